"""GPU parity of the device cell (materials, directional stiffness, CG) against
the oracle, called through the C-ABI.

Reference behaviour: CellBase::evaluate_stress_tangent and
evaluate_projected_directional_stiffness (src/cell/cell_base.cc:129-192),
SolverCG::solve (src/solver/solver_cg.cc:54-110).
"""
import ctypes

import numpy as np
import pytest

from muspectre_amd._device import DeviceBuffer

pytestmark = pytest.mark.gpu

DP = ctypes.POINTER(ctypes.c_double)


def _arr(a):
    return a.ctypes.data_as(DP)


class GpuCell:
    def __init__(self, capi, res, lengths, form, prec=0, allow_even=0):
        self.capi = capi
        self.lib = capi.load()
        self.res = list(res)
        self.d = len(res)
        self.nc = self.d ** 2
        self.N = int(np.prod(res))
        r = (ctypes.c_int * self.d)(*res)
        L = (ctypes.c_double * self.d)(*lengths)
        h = ctypes.c_void_p()
        capi.check(self.lib.mugpu_cell_create(self.d, r, L, form, prec, allow_even, ctypes.byref(h)))
        self.h = h

    def add_material(self, fam, E=0.0, nu=0.0, tau=0.0, H=0.0):
        m = ctypes.c_int()
        self.capi.check(self.lib.mugpu_cell_add_material(self.h, fam, E, nu, tau, H, ctypes.byref(m)))
        return m.value

    def add_pixels(self, m, pix, params=None):
        pix = np.ascontiguousarray(pix, dtype=np.int64)
        pp = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        self.capi.check(self.lib.mugpu_material_add_pixels(
            self.h, m, len(pix), pix.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            None if pp is None else _arr(pp)))

    def ptr(self, which):
        p = ctypes.c_void_p()
        fn = self.lib.mugpu_cell_strain if which == "F" else self.lib.mugpu_cell_stress
        self.capi.check(fn(self.h, ctypes.byref(p)))
        return p

    def set_F(self, F):
        F = np.ascontiguousarray(F, dtype=np.float64)
        self.capi.check(self.lib.mugpu_upload(self.ptr("F"), F.ctypes.data_as(ctypes.c_void_p), F.nbytes))

    def get(self, which):
        out = np.empty((self.N, self.nc))
        self.capi.check(self.lib.mugpu_download(out.ctypes.data_as(ctypes.c_void_p), self.ptr(which), out.nbytes))
        return out

    def __del__(self):
        self.lib.mugpu_cell_destroy(self.h)


def _two_phase(oc, gc, mask, fam_in, fam_out, Ein, Eout, nu, extra=None):
    mi_o = oc.add_material(fam_in, Ein, nu, *(extra or (0.0, 0.0)))
    mo_o = oc.add_material(fam_out, Eout, nu, *(extra or (0.0, 0.0)))
    mi_g = gc.add_material(fam_in, Ein, nu, *(extra or (0.0, 0.0)))
    mo_g = gc.add_material(fam_out, Eout, nu, *(extra or (0.0, 0.0)))
    idx = np.arange(mask.size)
    for i in idx:
        oc.add_pixel(mi_o if mask.flat[i] else mo_o, int(i))
    gc.add_pixels(mi_g, idx[mask.ravel()])
    gc.add_pixels(mo_g, idx[~mask.ravel()])


def _sphere_mask(res, r):
    g = np.meshgrid(*[np.arange(n) for n in res], indexing="ij")
    c = [n // 2 for n in res]
    return sum((gi - ci) ** 2 for gi, ci in zip(g, c)) < r * r


# (name, resolutions, lengths, formulation, family in, family out, E in,
#  E out[, allow_even]); power-of-two grids run the compile-time-length FFT
# passes and need the even-grid policy
CASES = [
    ("2d_small_le1", [51, 51], [7., 5.], 1, 1, 1, 10e9, 70e9),
    ("3d_small_le1", [9, 7, 11], [1., 1., 1.], 1, 1, 1, 100e9, 10e9),
    ("2d_finite_le1", [7, 9], [1., 1.], 0, 1, 1, 210e9, 70e9),
    ("3d_finite_le1", [5, 7, 5], [1., 1., 1.], 0, 1, 1, 210e9, 70e9),
    ("2d_small_le1_pow2", [64, 128], [1., 2.], 1, 1, 1, 10e9, 70e9, 1),
    ("2d_finite_le1_pow2", [32, 512], [1., 1.], 0, 1, 1, 210e9, 70e9, 1),
    ("3d_small_le1_pow2", [32, 16, 64], [1., 1., 1.], 1, 1, 1, 100e9, 10e9, 1),
    ("3d_finite_le1_pow2", [16, 32, 16], [1., 1.3, .7], 0, 1, 1, 210e9, 70e9, 1),
    ("3d_finite_le1_mixed", [16, 9, 32], [1., 1., 1.], 0, 1, 1, 210e9, 70e9, 1),
    ("3d_small_le1_radix3", [48, 16, 96], [1., 1., 1.], 1, 1, 1, 100e9, 10e9, 1),
    ("2d_small_le1_radix3", [96, 48], [1., 1.], 1, 1, 1, 100e9, 10e9, 1),
    # runtime-length last axis under 4096+ rows: several rows per workgroup
    ("3d_finite_le1_rows", [64, 72, 15], [1., 1., 1.], 0, 1, 1, 210e9, 70e9, 1),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_stress_tangent_and_directional_stiffness(capi, oracle, case):
    name, res, L, form, f_in, f_out, Ein, Eout = case[:8]
    even = case[8] if len(case) > 8 else 0
    d = len(res)
    nc = d * d
    oc = oracle.Cell(res, L, form, allow_even=bool(even))
    gc = GpuCell(capi, res, L, form, allow_even=even)
    _two_phase(oc, gc, _sphere_mask(res, min(res) / 3), f_in, f_out, Ein, Eout, .33)
    oc.initialise()
    capi.check(gc.lib.mugpu_cell_initialise(gc.h))
    rng = np.random.default_rng(11)
    F = 0.01 * rng.standard_normal((gc.N, nc))
    if form == 0:
        F += np.eye(d).ravel(order="F")
    oc.F[:] = F
    gc.set_F(F)
    oc.evaluate_stress_tangent()
    capi.check(gc.lib.mugpu_cell_evaluate_stress_tangent(gc.h))
    P = gc.get("P")
    assert np.abs(P - oc.P).max() <= 1e-12 * np.abs(oc.P).max()
    K = np.zeros((gc.N, nc * nc))
    capi.check(gc.lib.mugpu_cell_get_tangent(gc.h, _arr(K)))
    assert np.abs(K - oc.K).max() <= 1e-12 * np.abs(oc.K).max()
    dF = rng.standard_normal((gc.N, nc))
    ref = oc.directional_stiffness(dF)
    bi = DeviceBuffer.from_array(dF)
    bo = DeviceBuffer(dF.nbytes)
    capi.check(gc.lib.mugpu_cell_directional_stiffness(gc.h, bi.ptr, bo.ptr))
    out = bo.download(dF.shape)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


def _gpu_cg(gc, rhs, tol, maxiter):
    """Host loop of SolverCG::solve (src/solver/solver_cg.cc:54-110) over the
    device kernels."""
    lib, capi = gc.lib, gc.capi
    n = gc.N * gc.nc
    b = DeviceBuffer.from_array(rhs)
    x, r, p, Ap = (DeviceBuffer(8 * n) for _ in range(4))
    rdr = ctypes.c_double()
    bn2 = ctypes.c_double()
    capi.check(lib.mugpu_cg_start(gc.h, b.ptr, x.ptr, r.ptr, p.ptr, Ap.ptr, ctypes.byref(rdr), ctypes.byref(bn2)))
    tol2 = tol * tol * bn2.value
    rd = rdr.value
    it = 0
    while it < maxiter and (rd > tol2 or it == 0):
        nr = ctypes.c_double()
        capi.check(lib.mugpu_cg_step(gc.h, 1 if it > 0 else 0, x.ptr, r.ptr, p.ptr, Ap.ptr, ctypes.byref(nr)))
        rd = nr.value
        it += 1
    return x.download((gc.N, gc.nc)), it, rd < tol2


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_cg_solve_vs_oracle(capi, oracle, case):
    name, res, L, form, f_in, f_out, Ein, Eout = case[:8]
    even = case[8] if len(case) > 8 else 0
    d = len(res)
    nc = d * d
    oc = oracle.Cell(res, L, form, allow_even=bool(even))
    gc = GpuCell(capi, res, L, form, allow_even=even)
    _two_phase(oc, gc, _sphere_mask(res, min(res) / 3), f_in, f_out, Ein, Eout, .3)
    oc.initialise()
    capi.check(gc.lib.mugpu_cell_initialise(gc.h))
    F = np.zeros((gc.N, nc))
    if form == 0:
        F += np.eye(d).ravel(order="F")
    macro = np.zeros((d, d))
    macro[0, 1] = macro[1, 0] = 0.01
    F += macro.ravel(order="F")
    oc.F[:] = F
    gc.set_F(F)
    oc.evaluate_stress_tangent()
    capi.check(gc.lib.mugpu_cell_evaluate_stress_tangent(gc.h))
    rhs = -oc.apply_projection(oc.P.copy())
    start = oc.counter.value
    x_ref = oc.cg_solve(rhs, 1e-8, 400)
    it_ref = oc.counter.value - start
    x, it, conv = _gpu_cg(gc, rhs, 1e-8, 400)
    assert conv
    assert it == it_ref
    assert np.linalg.norm(x - x_ref) <= 1e-10 * np.linalg.norm(x_ref)


def _gpu_cg_chunked(gc, rhs, tol, maxiter, chunk):
    """The same loop with the iterations queued `chunk` at a time and the
    loop test applied on the device (mugpu_cg_steps)."""
    lib, capi = gc.lib, gc.capi
    n = gc.N * gc.nc
    b = DeviceBuffer.from_array(rhs)
    x, r, p, Ap = (DeviceBuffer(8 * n) for _ in range(4))
    rdr = ctypes.c_double()
    bn2 = ctypes.c_double()
    capi.check(lib.mugpu_cg_start(gc.h, b.ptr, x.ptr, r.ptr, p.ptr, Ap.ptr, ctypes.byref(rdr), ctypes.byref(bn2)))
    tol2 = tol * tol * bn2.value
    rd = rdr.value
    it = 0
    calls = 0
    while it < maxiter and (rd > tol2 or it == 0):
        nr = ctypes.c_double()
        done = ctypes.c_int()
        k = min(chunk, maxiter - it)
        capi.check(lib.mugpu_cg_steps(gc.h, 1 if it > 0 else 0, x.ptr, r.ptr, p.ptr, Ap.ptr, tol2, k,
                                      ctypes.byref(done), ctypes.byref(nr)))
        assert 1 <= done.value <= k
        rd = nr.value
        it += done.value
        calls += 1
    capi.check(lib.mugpu_cg_finish(gc.h))
    return x.download((gc.N, gc.nc)), r.download((gc.N, gc.nc)), it, rd, calls


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[5], CASES[6], CASES[7], CASES[9]],
                         ids=lambda c: c[0])
@pytest.mark.parametrize("chunk", [1, 3, 8])
def test_cg_steps_equal_step_loop(capi, case, chunk):
    """Iterations queued on the device with the loop test there give the
    host loop's iterates bit for bit, and stop where it stops: at
    convergence, and at maxiter in the middle of a chunk."""
    name, res, L, form, f_in, f_out, Ein, Eout = case[:8]
    even = case[8] if len(case) > 8 else 0
    d = len(res)
    gc = GpuCell(capi, res, L, form, allow_even=even)
    mask = _sphere_mask(res, min(res) / 3)
    mi = gc.add_material(f_in, Ein, .3)
    mo = gc.add_material(f_out, Eout, .3)
    idx = np.arange(mask.size)
    gc.add_pixels(mi, idx[mask.ravel()])
    gc.add_pixels(mo, idx[~mask.ravel()])
    capi.check(gc.lib.mugpu_cell_initialise(gc.h))
    F = np.zeros((gc.N, gc.nc))
    if form == 0:
        F += np.eye(d).ravel(order="F")
    F[:, 1] += 0.01
    F[:, d] += 0.01
    gc.set_F(F)
    capi.check(gc.lib.mugpu_cell_evaluate_stress_tangent(gc.h))
    rhs = np.random.default_rng(5).standard_normal((gc.N, gc.nc))
    for maxiter in (400, 7):
        x1, it1, conv1 = _gpu_cg(gc, rhs, 1e-8, maxiter)
        x2, _, it2, rd2, calls = _gpu_cg_chunked(gc, rhs, 1e-8, maxiter, chunk)
        assert it2 == it1
        assert calls <= -(-it1 // chunk) + 1
        np.testing.assert_array_equal(x2, x1)
    assert conv1 is False and it1 == 7


@pytest.mark.parametrize("case", [CASES[6], CASES[7], CASES[9], CASES[1]], ids=lambda c: c[0])
def test_cg_steps_pending_update_then_new_solve(capi, case):
    """A solve stopped after some chunks without mugpu_cg_finish: the next
    solve's mugpu_cg_start applies its pending x update with its own alpha
    (kept apart from the scalars the new solve overwrites), so its x is that
    of a host loop run for as many iterations, and the new solve is
    unaffected."""
    name, res, L, form, f_in, f_out, Ein, Eout = case[:8]
    even = case[8] if len(case) > 8 else 0
    gc = GpuCell(capi, res, L, form, allow_even=even)
    mask = _sphere_mask(res, min(res) / 3)
    mi = gc.add_material(f_in, Ein, .3)
    mo = gc.add_material(f_out, Eout, .3)
    idx = np.arange(mask.size)
    gc.add_pixels(mi, idx[mask.ravel()])
    gc.add_pixels(mo, idx[~mask.ravel()])
    capi.check(gc.lib.mugpu_cell_initialise(gc.h))
    F = np.zeros((gc.N, gc.nc))
    if form == 0:
        F += np.eye(len(res)).ravel(order="F")
    F[:, 1] += 0.01
    gc.set_F(F)
    capi.check(gc.lib.mugpu_cell_evaluate_stress_tangent(gc.h))
    lib = gc.lib
    n = gc.N * gc.nc
    rng = np.random.default_rng(11)
    rhs_a, rhs_b = rng.standard_normal((gc.N, gc.nc)), rng.standard_normal((gc.N, gc.nc))
    ref_a, it_a, _ = _gpu_cg(gc, rhs_a, 1e-30, 7)
    assert it_a == 7
    ref_b = _gpu_cg(gc, rhs_b, 1e-8, 200)[0]
    b = DeviceBuffer.from_array(rhs_a)
    x, r, p, Ap = (DeviceBuffer(8 * n) for _ in range(4))
    rdr, bn2 = ctypes.c_double(), ctypes.c_double()
    capi.check(lib.mugpu_cg_start(gc.h, b.ptr, x.ptr, r.ptr, p.ptr, Ap.ptr, ctypes.byref(rdr),
                                  ctypes.byref(bn2)))
    it = 0
    for k in (3, 4):
        nr, done = ctypes.c_double(), ctypes.c_int()
        capi.check(lib.mugpu_cg_steps(gc.h, 1 if it > 0 else 0, x.ptr, r.ptr, p.ptr, Ap.ptr, 0.0, k,
                                      ctypes.byref(done), ctypes.byref(nr)))
        assert done.value == k
        it += k
    # no mugpu_cg_finish: the next solve's start applies the pending update
    x_b, _, _, _, _ = _gpu_cg_chunked(gc, rhs_b, 1e-8, 200, 8)
    np.testing.assert_array_equal(x.download((gc.N, gc.nc)), ref_a)
    np.testing.assert_array_equal(x_b, ref_b)


def test_cg_steps_pending_update_and_free(capi):
    """Freeing the p of a pending x update (mugpu_free) applies it first;
    freeing its x drops it, so a block the field cache hands out again is
    never written by it."""
    name, res, L, form, f_in, f_out, Ein, Eout, even = CASES[6]
    gc = GpuCell(capi, res, L, form, allow_even=even)
    mask = _sphere_mask(res, min(res) / 3)
    mi = gc.add_material(f_in, Ein, .3)
    mo = gc.add_material(f_out, Eout, .3)
    idx = np.arange(mask.size)
    gc.add_pixels(mi, idx[mask.ravel()])
    gc.add_pixels(mo, idx[~mask.ravel()])
    capi.check(gc.lib.mugpu_cell_initialise(gc.h))
    F = np.zeros((gc.N, gc.nc))
    F[:, 1] += 0.01
    gc.set_F(F)
    capi.check(gc.lib.mugpu_cell_evaluate_stress_tangent(gc.h))
    lib = gc.lib
    n = gc.N * gc.nc
    rhs = np.random.default_rng(3).standard_normal((gc.N, gc.nc))
    ref, it, _ = _gpu_cg(gc, rhs, 1e-30, 5)
    assert it == 5

    def start():
        b = DeviceBuffer.from_array(rhs)
        x, r, p, Ap = (DeviceBuffer(8 * n) for _ in range(4))
        rdr, bn2 = ctypes.c_double(), ctypes.c_double()
        capi.check(lib.mugpu_cg_start(gc.h, b.ptr, x.ptr, r.ptr, p.ptr, Ap.ptr, ctypes.byref(rdr),
                                      ctypes.byref(bn2)))
        nr, done = ctypes.c_double(), ctypes.c_int()
        capi.check(lib.mugpu_cg_steps(gc.h, 0, x.ptr, r.ptr, p.ptr, Ap.ptr, 0.0, 5,
                                      ctypes.byref(done), ctypes.byref(nr)))
        assert done.value == 5
        return b, x, r, p, Ap

    b, x, r, p, Ap = start()
    p.free()  # applies the pending update
    np.testing.assert_array_equal(x.download((gc.N, gc.nc)), ref)
    b, x, r, p, Ap = start()
    x.free()  # drops it
    y = DeviceBuffer(8 * n)  # may reuse x's block
    marker = np.arange(n, dtype=np.float64)
    y.upload(marker)
    capi.check(lib.mugpu_cg_finish(gc.h))
    b2, x2, r2, p2, Ap2 = start()
    capi.check(lib.mugpu_cg_finish(gc.h))
    np.testing.assert_array_equal(y.download(n), marker)
    np.testing.assert_array_equal(x2.download((gc.N, gc.nc)), ref)


def _aniso_voigt(dim, seed, major_sym=True):
    vs = 3 if dim == 2 else 6
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((vs, vs))
    C = 50e9 * (A @ A.T / vs + np.eye(vs))
    if not major_sym:
        C += 5e9 * rng.standard_normal((vs, vs))
    return C


# (name, resolutions, formulation, major-symmetric C, allow_even)
GENERIC = [
    ("3d_small", [9, 7, 11], 1, True, 0),
    ("3d_finite", [5, 7, 5], 0, True, 0),
    ("3d_finite_unsym", [5, 7, 5], 0, False, 0),
    ("2d_small", [15, 13], 1, True, 0),
    ("2d_finite_unsym", [7, 9], 0, False, 0),
    ("3d_small_pow2", [32, 16, 64], 1, True, 1),
    ("3d_finite_pow2", [16, 32, 16], 0, True, 1),
]


@pytest.mark.parametrize("case", GENERIC, ids=lambda c: c[0])
def test_generic_material_vs_oracle(capi, oracle, case):
    """MaterialLinearElasticGeneric (material_linear_elastic_generic.hh)
    next to LE1: stress, tangent, directional stiffness and a CG solve
    against the oracle; small strain then needs the full (not isotropic)
    tangent representation, and a C without major symmetry the full matrix."""
    name, res, form, msym, even = case
    d = len(res)
    nc = d * d
    L = [1.0] * d
    Cv = _aniso_voigt(d, 7, msym)
    oc = oracle.Cell(res, L, form, allow_even=bool(even))
    gc = GpuCell(capi, res, L, form, allow_even=even)
    mask = _sphere_mask(res, min(res) / 3)
    idx = np.arange(mask.size)
    mg_o = oc.add_material_generic(Cv)
    ml_o = oc.add_material(1, 70e9, .3)
    cva = np.ascontiguousarray(Cv)
    mg = ctypes.c_int()
    capi.check(gc.lib.mugpu_cell_add_material_generic(gc.h, _arr(cva), ctypes.byref(mg)))
    ml = gc.add_material(1, 70e9, .3)
    for i in idx:
        oc.add_pixel(mg_o if mask.flat[i] else ml_o, int(i))
    gc.add_pixels(mg.value, idx[mask.ravel()])
    gc.add_pixels(ml, idx[~mask.ravel()])
    oc.initialise()
    capi.check(gc.lib.mugpu_cell_initialise(gc.h))
    rng = np.random.default_rng(3)
    F = 0.01 * rng.standard_normal((gc.N, nc))
    if form == 0:
        F += np.eye(d).ravel(order="F")
    oc.F[:] = F
    gc.set_F(F)
    oc.evaluate_stress_tangent()
    capi.check(gc.lib.mugpu_cell_evaluate_stress_tangent(gc.h))
    P = gc.get("P")
    assert np.abs(P - oc.P).max() <= 1e-12 * np.abs(oc.P).max()
    K = np.zeros((gc.N, nc * nc))
    capi.check(gc.lib.mugpu_cell_get_tangent(gc.h, _arr(K)))
    assert np.abs(K - oc.K).max() <= 1e-12 * np.abs(oc.K).max()
    dF = rng.standard_normal((gc.N, nc))
    ref = oc.directional_stiffness(dF)
    bi = DeviceBuffer.from_array(dF)
    bo = DeviceBuffer(dF.nbytes)
    capi.check(gc.lib.mugpu_cell_directional_stiffness(gc.h, bi.ptr, bo.ptr))
    assert np.abs(bo.download(dF.shape) - ref).max() <= 1e-12 * np.abs(ref).max()
    if msym:
        rhs = -oc.apply_projection(oc.P.copy())
        start = oc.counter.value
        x_ref = oc.cg_solve(rhs, 1e-8, 400)
        it_ref = oc.counter.value - start
        x, it, conv = _gpu_cg(gc, rhs, 1e-8, 400)
        assert conv and it == it_ref
        assert np.linalg.norm(x - x_ref) <= 1e-10 * np.linalg.norm(x_ref)


def test_generic_material_errors(capi):
    gc = GpuCell(capi, [5, 5], [1.0, 1.0], 1)
    m = ctypes.c_int()
    with pytest.raises(capi.MugpuError):
        capi.check(gc.lib.mugpu_cell_add_material(gc.h, 7, 1.0, .3, 0.0, 0.0, ctypes.byref(m)))


@pytest.mark.parametrize("case", [CASES[0], CASES[6], CASES[7]], ids=lambda c: c[0])
def test_cg_steps_graph_replay_equals_direct_launches(capi, case):
    """On small cells mugpu_cg_steps replays each chunk from a captured
    hipGraph.  Three solves in a row (the graphs replayed, with the solves'
    own start, finish and host synchronisations between the replays) give
    bit for bit the iterates of the same solves launched kernel by kernel,
    and the iteration counts of the host step loop."""
    name, res, L, form, f_in, f_out, Ein, Eout = case[:8]
    even = case[8] if len(case) > 8 else 0
    gc = GpuCell(capi, res, L, form, allow_even=even)
    mask = _sphere_mask(res, min(res) / 3)
    mi = gc.add_material(f_in, Ein, .3)
    mo = gc.add_material(f_out, Eout, .3)
    idx = np.arange(mask.size)
    gc.add_pixels(mi, idx[mask.ravel()])
    gc.add_pixels(mo, idx[~mask.ravel()])
    capi.check(gc.lib.mugpu_cell_initialise(gc.h))
    F = np.zeros((gc.N, gc.nc))
    if form == 0:
        F += np.eye(len(res)).ravel(order="F")
    F[:, 1] += 0.01
    gc.set_F(F)
    capi.check(gc.lib.mugpu_cell_evaluate_stress_tangent(gc.h))
    rhs = [np.random.default_rng(s).standard_normal((gc.N, gc.nc)) for s in (1, 2, 1)]
    try:
        capi.check(gc.lib.mugpu_cg_graphs(0))
        direct = [_gpu_cg_chunked(gc, b, 1e-8, 400, 8) for b in rhs]
        capi.check(gc.lib.mugpu_cg_graphs(1))
        replay = [_gpu_cg_chunked(gc, b, 1e-8, 400, 8) for b in rhs]
    finally:
        capi.check(gc.lib.mugpu_cg_graphs(1))
    for (xd, rd_, itd, _, _), (xg, rg, itg, _, _) in zip(direct, replay):
        assert itg == itd
        np.testing.assert_array_equal(xg, xd)
        np.testing.assert_array_equal(rg, rd_)
    ref_it = _gpu_cg(gc, rhs[0], 1e-8, 400)[1]
    assert replay[0][2] == ref_it and replay[2][2] == ref_it
    np.testing.assert_array_equal(replay[0][0], replay[2][0])
