"""The allocation-free mixed-precision solve (mugpu_cg_solve_mixed_noalloc)
against mugpu_cg_solve_mixed on the same cell: identical iterations and
solution, on two calls in a row (the second on a smaller cell, so the fixed
work array holds the first call's values), and a cell larger than the work
array rejected with a clear error.  Called through the C-ABI."""
import ctypes

import numpy as np
import pytest

from muspectre_amd._device import DeviceBuffer

pytestmark = pytest.mark.gpu


def _cell(capi, res, seed):
    """FP32 small-strain two-phase LE1 cell with an evaluated tangent and the
    projected stress of a uniform shear as the right-hand side."""
    lib = capi.load()
    d = len(res)
    nc = d * d
    h = ctypes.c_void_p()
    r = (ctypes.c_int * d)(*res)
    L = (ctypes.c_double * d)(*([1.0] * d))
    even = int(any(n % 2 == 0 for n in res))
    capi.check(lib.mugpu_cell_create(d, r, L, capi.SMALL_STRAIN, capi.FP32, even, ctypes.byref(h)))
    rng = np.random.default_rng(seed)
    N = int(np.prod(res))
    hard = rng.random(N) < 0.3
    idx = np.arange(N, dtype=np.int64)
    for E, sel in ((100e9, hard), (10e9, ~hard)):
        m = ctypes.c_int()
        capi.check(lib.mugpu_cell_add_material(h, capi.LE1, E, 0.3, 0.0, 0.0, ctypes.byref(m)))
        pix = np.ascontiguousarray(idx[sel])
        capi.check(lib.mugpu_material_add_pixels(
            h, m.value, len(pix), pix.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), None))
    capi.check(lib.mugpu_cell_initialise(h))
    eps = np.zeros((d, d))
    eps[0, 1] = eps[1, 0] = 0.01
    capi.check(lib.mugpu_cell_set_uniform_strain(h, eps.ravel(order="F").ctypes.data_as(
        ctypes.POINTER(ctypes.c_double))))
    capi.check(lib.mugpu_cell_evaluate_stress_tangent(h))
    P = ctypes.c_void_p()
    capi.check(lib.mugpu_cell_stress(h, ctypes.byref(P)))
    rhs = DeviceBuffer(8 * N * nc)
    capi.check(lib.mugpu_scale_copy(-1.0, P, rhs.ptr, N * nc))
    capi.check(lib.mugpu_cell_apply_projection(h, rhs.ptr))
    return h, rhs, N * nc


def _solve(capi, fn, h, rhs, n):
    x = DeviceBuffer(8 * n)
    its, refs, rel = ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
    capi.check(fn(h, rhs.ptr, x.ptr, 1e-8, 2000, 1e-4, ctypes.byref(its), ctypes.byref(refs),
                  ctypes.byref(rel)))
    return x.download(n), its.value, refs.value, rel.value


def test_noalloc_equals_mixed_solve(capi):
    lib = capi.load()
    # the second cell is smaller than the first, so the work array still holds
    # the first solve's vectors when it starts
    for res, seed in (([48, 32, 64], 1), ([24, 16, 32], 2)):
        h, rhs, n = _cell(capi, res, seed)
        assert n <= capi.NOALLOC_MAX_VALUES
        try:
            ref = _solve(capi, lib.mugpu_cg_solve_mixed, h, rhs, n)
            got = _solve(capi, lib.mugpu_cg_solve_mixed_noalloc, h, rhs, n)
            print("noalloc {}: {} inner iterations, {} refinements, |r|/|b| = {:.2e}".format(
                "x".join(map(str, res)), got[1], got[2], got[3]), flush=True)
            assert got[1:] == ref[1:]
            np.testing.assert_array_equal(got[0], ref[0])
            assert ref[3] <= 1e-8
        finally:
            lib.mugpu_cell_destroy(h)


def test_noalloc_rejects_cells_beyond_the_work_array(capi):
    lib = capi.load()
    res = [64, 64, 128]  # 4.7M values > MUGPU_NOALLOC_MAX_VALUES
    h = ctypes.c_void_p()
    r = (ctypes.c_int * 3)(*res)
    L = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    capi.check(lib.mugpu_cell_create(3, r, L, capi.SMALL_STRAIN, capi.FP32, 1, ctypes.byref(h)))
    try:
        x = DeviceBuffer(8)
        its, refs, rel = ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
        rc = lib.mugpu_cg_solve_mixed_noalloc(h, x.ptr, x.ptr, 1e-8, 10, 1e-4, ctypes.byref(its),
                                              ctypes.byref(refs), ctypes.byref(rel))
        assert rc == capi.EARG
        assert "at most" in lib.mugpu_last_error().decode()
    finally:
        lib.mugpu_cell_destroy(h)
