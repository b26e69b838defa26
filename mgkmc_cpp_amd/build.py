"""Build the device library and the native host driver in-tree.

    python -m mgkmc_cpp_amd.build

produces mgkmc_cpp_amd/lib/libmgkmc_gpu.so (HIP kernels + C ABI, gfx950) and
mgkmc_cpp_amd/lib/libmgkmc_host.so (drop-in KmcSimulation2D/3D over the C ABI).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["api.hip", "ls_kernels.hip", "kmc_kernels.hip", "gen_kernels.hip"]
ARCH = os.environ.get("MGK_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

# The KMC kernels evaluate barriers, rates and softening in the reference's
# operation order: no fused multiply-add contraction there.  The LS kernels
# fuse a multiply into an add only within one source expression: with the
# default cross-statement fusion the code generator picks which product of
# a + b*c-like chains to fuse per instantiation, so the single-context and
# batched instantiations of one pass (the same source) rounded differently.
EXTRA_FLAGS = {"kmc_kernels.hip": ["-ffp-contract=off"], "ls_kernels.hip": ["-ffp-contract=on"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r.stdout


def build(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        extra = EXTRA_FLAGS.get(src, [])
        cmds.append([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(len(cmds), max(1, (os.cpu_count() or 2)))) as ex:
        list(ex.map(_run, cmds))
    gpu_so = os.path.join(LIB, "libmgkmc_gpu.so")
    tmp = gpu_so + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, gpu_so)
    host_so = os.path.join(LIB, "libmgkmc_host.so")
    tmp = host_so + ".tmp"
    _run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-Wno-unused-function",
          "-fvisibility=hidden", "-fvisibility-inlines-hidden",
          os.path.join(HOST, "host_api.cpp"), "-o", tmp,
          "-L" + LIB, "-lmgkmc_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, host_so)
    callers = build_dropin_callers()
    if verbose:
        print("built", gpu_so, host_so, *callers)
    return gpu_so, host_so


DROPIN = os.path.join(os.path.dirname(HERE), "tests", "dropin")


def build_dropin_callers():
    """The reference-style caller of the drop-in C++ headers (tests/dropin),
    on the built-in tensor type and on a second layout-identical type, linked
    against the device library: test programs, built with the rest so that
    GPU runs only execute them."""
    src = os.path.join(DROPIN, "reference_style_caller.cpp")
    if not os.path.exists(src):
        return []
    out_dir = os.path.join(DROPIN, "_build")
    os.makedirs(out_dir, exist_ok=True)
    outs = []
    for name, extra in (("caller_builtin", []), ("caller_layout", ["-DCALLER_LAYOUT_TYPE"])):
        exe = os.path.join(out_dir, name)
        _run(["g++", "-O2", "-std=c++17", "-Wall"] + extra + ["-I" + HOST, "-I" + DROPIN, src, "-o", exe + ".tmp",
              "-L" + LIB, "-lmgkmc_gpu", "-Wl,-rpath,$ORIGIN/../../../mgkmc_cpp_amd/lib"])
        os.replace(exe + ".tmp", exe)
        outs.append(exe)
    return outs


if __name__ == "__main__":
    build(verbose=True)
