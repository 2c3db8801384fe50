"""dmk_divergence_rms_noalloc, the divergence RMS that allocates no device
memory, against dmk_divergence_rms (a handle's buffers, host field) and the
oracle's utilities_divergenceRMS (spectral_utilities.f90:650-691).

Two calls in a row on one GPU, the second on a smaller grid, so the second
reads nothing the first left in the shared device array; then sizes beyond the
array's maximum are refused with DMK_ERR_ARG.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DMK_ERR_ARG = 1


def _noalloc(lib, cells, size, dc, field):
    import torch
    f = torch.from_numpy(np.ascontiguousarray(field)).to("cuda")
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    rc = lib.dmk_divergence_rms_noalloc((C.c_int32 * 3)(*cells), (C.c_double * 3)(*size), dc,
                                        C.c_void_p(f.data_ptr()), C.c_void_p(out.data_ptr()),
                                        C.c_void_p(stream.cuda_stream))
    stream.synchronize()
    return rc, float(out.cpu()[0])


def _reference(cells, size, dc, field, oracle):
    from damask_amd import _lib, config, workloads
    from damask_amd.grid import GridMechSpectralBasic
    grid, mat = workloads.polycrystal(cells, 3, 5, size=size)
    num = config.Numerics(divergence_correction=dc)
    g = GridMechSpectralBasic(grid, config.parse_material(mat), num)
    d = np.zeros(1)
    _lib.check(g.lib.dmk_divergence_rms(g.h, _lib.dptr(field), _lib.dptr(d)))
    g.close()
    o = oracle.Problem(grid, config.parse_material(mat), num)
    d_o = oracle.lib().orc_divergence_rms(o.grid, oracle.ptr(field))
    o.close()
    return d[0], d_o


def test_noalloc_divergence_rms_matches_handle_entry(gpu_lib, oracle):
    lib = gpu_lib
    rng = np.random.default_rng(11)
    cases = [([32, 24, 20], (1.0, 0.75, 1.25), 1),     # grid+size
             ([7, 6, 5], (0.5, 1.0, 0.8), 0)]          # smaller, odd, no correction
    for cells, size, dc in cases:
        N = int(np.prod(cells))
        field = 1e8 * rng.standard_normal(9 * N)
        print(f"noalloc divergence RMS: cells {cells}, size {size}, divergence_correction {dc}", flush=True)
        rc, r = _noalloc(lib, cells, size, dc, field)
        assert rc == 0, lib.dmk_last_error()
        d_h, d_o = _reference(cells, size, dc, field, oracle)
        assert abs(r - d_h) <= 1e-12 * d_h, (r, d_h)
        assert abs(r - d_o) <= 1e-12 * d_o, (r, d_o)


def test_noalloc_divergence_rms_refuses_oversized_grids(gpu_lib):
    lib = gpu_lib
    from damask_amd import _lib
    field = np.zeros(9)
    for cells in ([65, 64, 64], [2048, 1, 1]):
        assert _lib.DMK_NOALLOC_MAX_CELLS == 262144
        rc, _ = _noalloc(lib, cells, (1.0, 1.0, 1.0), 1, field)
        assert rc == DMK_ERR_ARG
        assert b"exceed" in lib.dmk_last_error()
