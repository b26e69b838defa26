"""VTU output (vtu_writer.hpp:22-393) on the CPU: the library's writer
(mgk_vtu_write, the drop-in export_to_vtu / export_to_vtu_2d) against the
oracle's restatement of the reference writer, byte for byte, and that
restatement against the closed form of a uniform normal strain."""
import numpy as np
import pytest

from mgkmc_cpp_amd.solver import write_vtu
from oracle import vtu_ref


def _sym(rng, N, D, scale):
    a = scale * rng.standard_normal((N, D, D))
    return 0.5 * (a + a.transpose(0, 2, 1))


@pytest.mark.parametrize("shape", [(3, 2, 4), (1, 3, 2), (4, 3)])
@pytest.mark.parametrize("extras", [False, True])
def test_vtu_writer_matches_reference_layout(tmp_path, shape, extras):
    rng = np.random.default_rng(len(shape) * 10 + extras)
    D = len(shape)
    N = int(np.prod(shape))
    eps = _sym(rng, N, D, 1e-3)
    eps[0] = 0.0                                 # zeros and negative zeros print as the reference does
    eps[1, 0, 0] = -0.0
    sig = _sym(rng, N, D, 1e8)
    E = 70e9 * (1 + 0.1 * rng.standard_normal(N))
    nu = np.full(N, 0.3)
    ep = _sym(rng, N, D, 1e-2) if extras else None
    soft = np.abs(rng.standard_normal((N, 2))) if extras else None
    path = tmp_path / "out.vtu"
    write_vtu(path, shape, eps, sig, E, nu, pixel=0.5, eps_plastic=ep, soft=soft)
    got = path.read_text()
    if D == 2:
        want = vtu_ref.export_to_vtu_2d(shape[0], shape[1], eps, sig, E, nu, 0.5, ep, soft)
    else:
        want = vtu_ref.export_to_vtu(*shape, eps, sig, E, nu, 0.5, ep, soft)
    assert got == want


def test_reference_displacement_of_uniform_normal_strain():
    nx, ny, nz, pixel = 3, 4, 2, 0.5
    a, b, c = 1e-3, -2e-3, 5e-4
    eps = np.zeros((nx * ny * nz, 3, 3))
    eps[:, 0, 0], eps[:, 1, 1], eps[:, 2, 2] = a, b, c
    u = np.array(vtu_ref.reconstruct_displacement(nx, ny, nz, eps, pixel))
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    want = np.stack([a * pixel * i, b * pixel * (j - ny / 2), c * pixel * (k - nz / 2)], axis=-1).reshape(-1, 3)
    assert np.abs(u - want).max() < 1e-15
