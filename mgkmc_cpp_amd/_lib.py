"""Loader of the in-tree native libraries.  There is no fallback: when the
HIP build is missing every entry point raises."""
import ctypes as C
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
GPU_LIB = os.path.join(LIB_DIR, "libmgkmc_gpu.so")
HOST_LIB = os.path.join(LIB_DIR, "libmgkmc_host.so")

_gpu = None
_host = None

P = C.c_void_p
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class MgkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mgkmc error {code}: {msg}")
        self.code = code


def _sig(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def gpu():
    """The device library (C ABI of include/mgkmc_gpu.h)."""
    global _gpu
    if _gpu is None:
        if not os.path.exists(GPU_LIB):
            raise ImportError(f"{GPU_LIB} is not built; run python -m mgkmc_cpp_amd.build")
        lib = C.CDLL(GPU_LIB)
        _sig(lib, "mgk_last_error", C.c_char_p, [])
        _sig(lib, "mgk_abi_version", C.c_int, [])
        _sig(lib, "mgk_create", C.c_int,
             [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, DP, DP, C.POINTER(P)])
        _sig(lib, "mgk_destroy", C.c_int, [P])
        _sig(lib, "mgk_set_precision", C.c_int, [P, C.c_int])
        _sig(lib, "mgk_ls_solve", C.c_int,
             [P, DP, DP, C.c_int, DP, DP, DP, DP, C.c_int, C.c_double, IP, DP])
        _sig(lib, "mgk_set_field", C.c_int, [P, C.c_int, DP])
        _sig(lib, "mgk_get_field", C.c_int, [P, C.c_int, DP])
        _sig(lib, "mgk_kmc_init", C.c_int, [P, C.POINTER(abi.KmcParams)])
        _sig(lib, "mgk_elastic_run", C.c_int, [P, DP, DP, C.c_int, C.c_double, IP, DP])
        _sig(lib, "mgk_update_rates", C.c_int, [P, C.c_double, DP, IP, IP])
        _sig(lib, "mgk_select", C.c_int, [P, C.c_double, IP, DP])
        _sig(lib, "mgk_select_exact", C.c_int, [P, C.c_double, IP, DP])
        _sig(lib, "mgk_state_size", C.c_longlong, [P])
        _sig(lib, "mgk_patch_init", C.c_int, [P, C.c_int])
        _sig(lib, "mgk_kmc_draw", C.c_int, [P, C.c_void_p, C.c_ulonglong, C.c_int])
        _sig(lib, "mgk_correlated_field", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                                    C.c_double, C.c_int, C.c_double, C.c_int, C.c_double,
                                                    C.c_ulonglong, DP, DP])
        _sig(lib, "mgk_correlated_field_noalloc", C.c_int,
             [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
              C.c_double, C.c_ulonglong, C.c_void_p, C.c_void_p, C.c_void_p])
        _sig(lib, "mgk_patch_unit", C.c_int, [P, DP])
        _sig(lib, "mgk_patch_apply", C.c_int, [P, C.c_int, DP, C.c_double])
        _sig(lib, "mgk_state_save", C.c_int, [P, C.c_void_p, C.c_longlong])
        _sig(lib, "mgk_state_load", C.c_int, [P, C.c_void_p, C.c_longlong])
        _sig(lib, "mgk_apply_flip", C.c_int, [P, C.c_int, C.c_double, DP, DP, C.c_int, C.c_int])
        _sig(lib, "mgk_find_unstable", C.c_int, [P, IP, DP, C.c_int, IP])
        _sig(lib, "mgk_select_flip", C.c_int, [P, C.c_double, C.c_double, DP, DP, C.c_int, C.c_int])
        _sig(lib, "mgk_last_selection", C.c_int, [P, IP, DP])
        _sig(lib, "mgk_kmc_event", C.c_int,
             [P, C.POINTER(abi.EventIn), DP, DP, C.POINTER(abi.EventOut)])
        _sig(lib, "mgk_batch_create", C.c_int, [C.POINTER(P), C.c_int, C.POINTER(P)])
        _sig(lib, "mgk_batch_destroy", C.c_int, [P])
        _sig(lib, "mgk_batch_size", C.c_int, [P])
        _sig(lib, "mgk_batch_set_graphs", C.c_int, [P, C.c_int])
        _sig(lib, "mgk_batch_kmc_event", C.c_int,
             [P, C.POINTER(abi.EventIn), IP, DP, DP, C.POINTER(abi.EventOut)])
        _sig(lib, "mgk_sync", C.c_int, [P])
        _sig(lib, "mgk_timer_start", C.c_int, [P])
        _sig(lib, "mgk_timer_stop", C.c_int, [P, DP])
        _sig(lib, "mgk_last_solve_stats", C.c_int, [P, IP, IP])
        _sig(lib, "mgk_last_solve_history", C.c_int, [P, DP, C.c_int, IP])
        _sig(lib, "mgk_device_bytes", C.c_longlong, [P])
        _sig(lib, "mgk_estimate_bytes", C.c_longlong, [C.c_int] * 6)
        _sig(lib, "mgk_time_ls_pass", C.c_int, [P, C.c_int, C.c_int, C.c_int, DP])
        _sig(lib, "mgk_cycle_supported", C.c_int, [P, IP])
        _sig(lib, "mgk_tuning", C.c_int, [P, IP, DP])
        _sig(lib, "mgk_set_tuning", C.c_int, [P, IP])
        _sig(lib, "mgk_cycle_timer", C.c_int, [P, C.c_int])
        _sig(lib, "mgk_cycle_timer_read", C.c_int, [P, DP, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)])
        _gpu = lib
    return _gpu


def host():
    """The native host driver (include/mgkmc_host.h)."""
    global _host
    if _host is None:
        gpu()
        if not os.path.exists(HOST_LIB):
            raise ImportError(f"{HOST_LIB} is not built; run python -m mgkmc_cpp_amd.build")
        lib = C.CDLL(HOST_LIB)
        _sig(lib, "mgk_host_last_error", C.c_char_p, [])
        _sig(lib, "mgk_sim_config_defaults", None, [C.POINTER(abi.SimConfig)])
        _sig(lib, "mgk_run_config_defaults", None, [C.POINTER(abi.RunConfig)])
        _sig(lib, "mgk_sim_create", C.c_int, [C.POINTER(abi.SimConfig), DP, DP, C.POINTER(P)])
        _sig(lib, "mgk_sim_destroy", C.c_int, [P])
        _sig(lib, "mgk_sim_run", C.c_int, [P, C.POINTER(abi.RunConfig), C.POINTER(abi.RunStats)])
        _sig(lib, "mgk_sim_event_steps", C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(abi.RunStats)])
        _sig(lib, "mgk_sim_set_checkpoint", C.c_int, [P, C.c_char_p, C.c_char_p])
        _sig(lib, "mgk_sim_save", C.c_int, [P, C.c_char_p])
        _sig(lib, "mgk_sim_set_output", C.c_int, [P, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int])
        _sig(lib, "mgk_sim_export_vtu", C.c_int, [P, C.c_char_p])
        _sig(lib, "mgk_vtu_write", C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, DP, DP, DP, DP,
                                             DP, DP])
        _sig(lib, "mgk_sim_load", C.c_int, [P, C.c_char_p])
        _sig(lib, "mgk_sim_elastic_run", C.c_int, [P])
        _sig(lib, "mgk_sim_get_field", C.c_int, [P, C.c_int, DP])
        _sig(lib, "mgk_sim_set_field", C.c_int, [P, C.c_int, DP])
        _sig(lib, "mgk_sim_get_macro", C.c_int, [P, DP, DP, DP])
        _sig(lib, "mgk_sim_set_macro", C.c_int, [P, DP, C.c_double])
        _sig(lib, "mgk_sim_num_events", C.c_longlong, [P])
        _sig(lib, "mgk_sim_get_events", C.c_int, [P, C.POINTER(abi.EventRecord), C.c_longlong])
        _sig(lib, "mgk_sim_device_ctx", C.c_int, [P, C.POINTER(P)])
        _sig(lib, "mgk_sim_batch_create", C.c_int, [C.POINTER(P), C.c_int, C.POINTER(P)])
        _sig(lib, "mgk_sim_batch_destroy", C.c_int, [P])
        _sig(lib, "mgk_sim_batch_event_steps", C.c_int,
             [P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_longlong), C.POINTER(abi.RunStats)])
        _sig(lib, "mgk_sim_batch_set_member", C.c_int, [P, C.c_int, P])
        _host = lib
    return _host


def check(rc, lib=None):
    if rc != abi.MGK_OK:
        lib = lib or gpu()
        msg = lib.mgk_last_error().decode() if hasattr(lib, "mgk_last_error") else ""
        raise MgkError(rc, msg)


def check_host(rc):
    if rc != abi.MGK_OK:
        h = host()
        msg = h.mgk_host_last_error().decode()
        g = gpu().mgk_last_error().decode()
        raise MgkError(rc, msg or g)
