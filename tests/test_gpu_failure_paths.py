"""Failure paths of the library's device allocations, forced from the host:
mugpu_debug_fail_allocations makes the library's own allocations report a
full device after a given count (no kernel is affected; every launch that
happens runs normally).  Called through the C-ABI."""
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
    capi.check(lib.mugpu_cell_create(d, r, L, capi.SMALL_STRAIN, capi.FP32, 1, ctypes.byref(h)))
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


def _mixed(capi, h, rhs, n):
    """mugpu_cg_solve_mixed: (return code, x, inner iterations, refinements)"""
    lib = capi.load()
    x = DeviceBuffer(8 * n)
    its, refs, rel = ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
    rc = lib.mugpu_cg_solve_mixed(h, rhs.ptr, x.ptr, 1e-8, 2000, 1e-4, ctypes.byref(its),
                                  ctypes.byref(refs), ctypes.byref(rel))
    return rc, (x.download(n) if rc == capi.OK else None), its.value, refs.value


def test_mixed_vectors_all_or_nothing(capi):
    """A cell's first mixed-precision solve allocates its five work vectors,
    all or nothing: with the third allocation refused the solve returns
    MUGPU_EMEMORY and keeps none of them, and the next solve on the same cell
    allocates them afresh and equals, bit for bit, the solve of an identical
    cell that never failed."""
    lib = capi.load()
    res = [16, 24, 32]
    h1, rhs1, n = _cell(capi, res, 3)
    h2, rhs2, _ = _cell(capi, res, 3)
    try:
        rc, ref, its, refs = _mixed(capi, h2, rhs2, n)
        assert rc == capi.OK
        capi.check(lib.mugpu_debug_fail_allocations(2))
        try:
            rc, _, _, _ = _mixed(capi, h1, rhs1, n)
            msg = lib.mugpu_last_error().decode()
        finally:
            capi.check(lib.mugpu_debug_fail_allocations(-1))
        assert rc == capi.EMEMORY
        assert "injected" in msg
        rc, got, its1, refs1 = _mixed(capi, h1, rhs1, n)
        assert rc == capi.OK
        assert (its1, refs1) == (its, refs)
        np.testing.assert_array_equal(got, ref)
    finally:
        lib.mugpu_cell_destroy(h1)
        lib.mugpu_cell_destroy(h2)


def test_cell_creation_and_initialise_fail_cleanly(capi):
    """A cell whose creation or initialisation meets a full device reports
    MUGPU_EMEMORY; with the device available again, the same steps succeed
    (initialise frees what a failed attempt left before allocating anew)."""
    lib = capi.load()
    r = (ctypes.c_int * 3)(16, 16, 16)
    L = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    h = ctypes.c_void_p()
    capi.check(lib.mugpu_debug_fail_allocations(1))
    try:
        rc = lib.mugpu_cell_create(3, r, L, capi.SMALL_STRAIN, capi.FP32, 1, ctypes.byref(h))
    finally:
        capi.check(lib.mugpu_debug_fail_allocations(-1))
    assert rc == capi.EMEMORY
    capi.check(lib.mugpu_cell_create(3, r, L, capi.SMALL_STRAIN, capi.FP32, 1, ctypes.byref(h)))
    try:
        m = ctypes.c_int()
        capi.check(lib.mugpu_cell_add_material(h, capi.LE1, 10e9, 0.3, 0.0, 0.0, ctypes.byref(m)))
        pix = np.arange(16 ** 3, dtype=np.int64)
        capi.check(lib.mugpu_material_add_pixels(
            h, m.value, len(pix), pix.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), None))
        capi.check(lib.mugpu_debug_fail_allocations(1))
        try:
            rc = lib.mugpu_cell_initialise(h)
        finally:
            capi.check(lib.mugpu_debug_fail_allocations(-1))
        assert rc == capi.EMEMORY
        capi.check(lib.mugpu_cell_initialise(h))
        capi.check(lib.mugpu_cell_evaluate_stress_tangent(h))
    finally:
        lib.mugpu_cell_destroy(h)


def _rve_request(res, seed):
    rng = np.random.default_rng(seed)
    N = int(np.prod(res))
    E = np.where(rng.random(N) < 0.4, 100e9, 10e9)
    nu = np.full(N, 0.3)
    d = len(res)
    a = rng.uniform(-0.01, 0.01, (d, d))
    return res, E, nu, 0.5 * (a + a.T)


def test_server_admission_fails_cleanly(capi):
    """Admitting the first request of a resolution sets up its class (the
    transform plan, the pixel list, the batch slots).  With each of those
    allocations refused in turn, progress reports MUGPU_EMEMORY and the
    request stays queued, admitted by nobody and holding nothing (its arena
    block, pinned words and event go back); once the device is available
    again the same server serves it, with the result of a server that never
    failed, bit for bit."""
    from muspectre_amd import serving
    lib = capi.load()
    res, E, nu, eps = _rve_request([32, 16, 48], 5)

    ref_srv = serving.RVEServer(hbm_budget=1 << 30, keep_fields=True)
    try:
        rid = ref_srv.submit(res, E, nu, eps)
        ref_srv.run()
        ref = ref_srv.result(rid, fields=True)
    finally:
        ref_srv.close()
    assert ref["ok"] and ref["converged"]

    srv = serving.RVEServer(hbm_budget=1 << 30, keep_fields=True)
    try:
        rid = srv.submit(res, E, nu, eps)
        refused = 0
        for after in range(64):
            capi.check(lib.mugpu_debug_fail_allocations(after))
            try:
                srv.progress()
            except capi.MugpuError as e:
                assert e.code == capi.EMEMORY and "injected" in str(e)
                st = srv.stats()
                assert st["admitted"] == 0 and st["peak_running"] == 0
                refused += 1
                continue
            finally:
                capi.check(lib.mugpu_debug_fail_allocations(-1))
            break
        else:
            pytest.fail("the class set-up never completed")
        assert refused >= 4  # plan tables and partials, pixel list, slots, control words
        srv.run()
        assert srv.poll() == [rid]
        got = srv.result(rid, fields=True)
        st = srv.stats()
        assert st["admitted"] == 1 and st["finished"] == 1 and st["failed"] == 0
    finally:
        srv.close()
    for k in ("status", "converged", "newton_iterations", "cg_iterations"):
        assert got[k] == ref[k]
    np.testing.assert_array_equal(got["strain"], ref["strain"])
    np.testing.assert_array_equal(got["stress"], ref["stress"])
    np.testing.assert_array_equal(got["mean_stress"], ref["mean_stress"])
