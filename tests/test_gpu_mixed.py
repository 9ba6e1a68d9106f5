"""Mixed precision on the GPU: FP32 FFTs, projection, K:p and CG vectors with
FP64 dot products, FP64 alpha/beta, FP64 material evaluation and an FP64
residual-refinement loop, against the FP64 oracle.

Tolerance: the refined solve stops on the same FP64 residual rule as the
reference (|r| <= tol |b|, tol = 1e-8).  The homogenised stress must agree
with the FP64 solution to 1e-6 relative (the target for the precision form).
The strain field is held to 1e-5: two solves that each stop at a relative
residual of 1e-8 differ by up to the condition number times 1e-8 (measured:
~1e-6 at a stiffness contrast of 10).  Newton iteration counts must be equal.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-6
TOL_FIELD = 1e-5


@pytest.fixture(scope="module")
def µ():
    import muspectre_amd
    return muspectre_amd


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def _homog(stress, nc):
    return np.asarray(stress).reshape(-1, nc).mean(0)


def _disp_grad(grad, nc, finite):
    """Strain field compared without the identity of finite strain (F - I),
    so that the relative error measures the solution, not the offset."""
    g = np.asarray(grad).reshape(-1, nc).copy()
    if finite:
        d = int(round(np.sqrt(nc)))
        g -= np.eye(d).ravel()[None, :]
    return g


def test_c1_disc_fp32_refined(µ, oracle):
    """51x51 small strain, disc of radius 10 (bin/small_case.py:46-87)."""
    res, L = [51, 51], [7., 5.]
    x = np.arange(51)
    mask = np.hypot(x[:, None] - 25, x[None, :] - 25) < 10
    cell = µ.Cell(res, L, µ.Formulation.small_strain, precision="fp32")
    inc = µ.material.MaterialLinearElastic1_2d.make(cell, "inclusion", 10e9, .33)
    mat = µ.material.MaterialLinearElastic1_2d.make(cell, "matrix", 70e9, .33)
    idx = np.arange(mask.size)
    inc.add_pixels_by_index(idx[mask.ravel()])
    mat.add_pixels_by_index(idx[~mask.ravel()])
    load = np.array([[0, 0], [0, .03]])
    solver = µ.solvers.SolverCG(cell, 1e-8, 401, False)
    r = µ.solvers.newton_cg(cell, [load], solver, 1e-5, 1e-10, 0)[0]
    oc = oracle.Cell(res, L, oracle.SMALL_STRAIN)
    a = oc.add_material(oracle.LE1, 10e9, .33)
    b = oc.add_material(oracle.LE1, 70e9, .33)
    for i in range(mask.size):
        oc.add_pixel(a if mask.flat[i] else b, i)
    o = oc.newton_cg([load], 1e-8, 401, 1e-5, 1e-10)[0]
    assert r.success and r.nb_it == o["nb_it"]
    assert _rel(_homog(r.stress, 4), _homog(o["stress"], 4)) < TOL
    assert _rel(r.grad, o["grad"]) < TOL_FIELD
    assert solver.get_refinements() >= 2


@pytest.mark.parametrize("driver", ["newton_cg", "de_geus"])
@pytest.mark.parametrize("res", [[9, 11, 7], [16, 32, 16], [16, 9, 32], [32, 24, 15]],
                         ids=lambda r: "x".join(map(str, r)))
@pytest.mark.parametrize("form", ["small", "finite"])
def test_3d_two_phase_fp32_refined(µ, oracle, form, res, driver):
    """Runtime-length axes, compiled-length axes (the x update on the next
    first pass, the even policy) and grids that mix the two, under both
    drivers (de_geus takes its first right-hand side from -G:K:dF,
    solvers.cc:353-363)."""
    even = any(n % 2 == 0 for n in res)
    g = np.meshgrid(*[np.arange(n) for n in res], indexing="ij")
    mask = ((g[0] - 4) ** 2 + (g[1] - 5) ** 2 + (g[2] - 3) ** 2) < 9
    F_ = µ.Formulation.small_strain if form == "small" else µ.Formulation.finite_strain
    cell = µ.Cell(res, [1., 1.2, .8], F_, precision="fp32", allow_even=even)
    h = µ.material.MaterialLinearElastic1_3d.make(cell, "hard", 100e9, .3)
    s = µ.material.MaterialLinearElastic1_3d.make(cell, "soft", 10e9, .3)
    idx = np.arange(mask.size)
    h.add_pixels_by_index(idx[mask.ravel()])
    s.add_pixels_by_index(idx[~mask.ravel()])
    steps = [np.array([[.01, .002, 0], [.002, 0, 0], [0, 0, -.005]]) * k for k in (1, 2)]
    solver = µ.solvers.SolverCG(cell, 1e-8, 400, 0)
    rs = getattr(µ.solvers, driver)(cell, steps, solver, 1e-5, 1e-10, 0)
    oc = oracle.Cell(res, [1., 1.2, .8], oracle.SMALL_STRAIN if form == "small" else oracle.FINITE_STRAIN,
                     allow_even=even)
    a = oc.add_material(oracle.LE1, 100e9, .3)
    b = oc.add_material(oracle.LE1, 10e9, .3)
    for i in range(mask.size):
        oc.add_pixel(a if mask.flat[i] else b, i)
    ro = getattr(oc, driver)(steps, 1e-8, 400, 1e-5, 1e-10)
    assert len(rs) == len(ro) == 2
    for r, o in zip(rs, ro):
        assert r.success and r.nb_it == o["nb_it"]
        assert _rel(_homog(r.stress, 9), _homog(o["stress"], 9)) < TOL
        fin = form == "finite"
        assert _rel(_disp_grad(r.grad, 9, fin), _disp_grad(o["grad"], 9, fin)) < TOL_FIELD


def test_laminate_fp32_refined(µ):
    """Reuss laminate (test_solver_newton_cg.cc:101-323) with the refined solve."""
    cell = µ.Cell([3, 3], [1., 1.], µ.Formulation.small_strain, precision="fp32")
    hard = µ.material.MaterialLinearElastic1_2d.make(cell, "hard", 4., .33)
    soft = µ.material.MaterialLinearElastic1_2d.make(cell, "soft", 2., .33)
    for i in range(9):
        (hard if i // 3 < 1 else soft).add_pixel([i // 3, i % 3])
    solver = µ.solvers.SolverCG(cell, 1e-8, 20, False)
    r = µ.solvers.newton_cg(cell, np.array([[1., 0], [0, 0]]), solver, 1e-5, equil_tol=1e-10)
    factor = 0.5 / 3 + 1 - 1 / 3
    F = r.grad.reshape(-1, 4)
    for i in range(9):
        ref = (0.5 if i // 3 < 1 else 1.0) / factor
        assert abs(F[i, 0] - ref) < 1e-7 and np.abs(F[i, 1:]).max() < 1e-7


@pytest.mark.parametrize("res", [[9, 9, 9], [16, 16, 32]], ids=lambda r: "x".join(map(str, r)))
def test_hep1_consistent_fp32_refined(µ, oracle, res):
    """C3 first load step with the consistent tangent, FP32 operator; 9^3
    (runtime-length axes) and 16x16x32 (compiled-length axes: the
    benchmark's FP32 path)."""
    from test_gpu_hep1 import _grf_mask, E_SOFT, E_HARD, NU, TY_SOFT, TY_HARD, H
    even = any(n % 2 == 0 for n in res)
    mask = _grf_mask(res, 2.0, 1)
    cell = µ.Cell(res, [1., 1., 1.], µ.Formulation.finite_strain, precision="fp32",
                  allow_even=even)
    mk = µ.material.MaterialHyperElastoPlastic1_3d.make
    soft = mk(cell, "soft", E_SOFT, NU, TY_SOFT, H, consistent_tangent=True)
    hard = mk(cell, "hard", E_HARD, NU, TY_HARD, H, consistent_tangent=True)
    idx = np.arange(mask.size)
    hard.add_pixels_by_index(idx[mask.ravel()])
    soft.add_pixels_by_index(idx[~mask.ravel()])
    oc = oracle.Cell(res, [1., 1., 1.], oracle.FINITE_STRAIN, allow_even=even)
    a = oc.add_material(oracle.HEP1_CT, E_SOFT, NU, TY_SOFT, H)
    b = oc.add_material(oracle.HEP1_CT, E_HARD, NU, TY_HARD, H)
    for i in range(mask.size):
        oc.add_pixel(b if mask.flat[i] else a, i)
    load = [np.diag([0, 0, 0.005])]
    solver = µ.solvers.SolverCG(cell, 1e-8, 2000, 0)
    r = µ.solvers.newton_cg(cell, load, solver, 1e-5, 1e-10, 0)[0]
    o = oc.newton_cg(load, 1e-8, 2000, 1e-5, 1e-10)[0]
    assert r.success and r.nb_it == o["nb_it"]
    assert _rel(_homog(r.stress, 9), _homog(o["stress"], 9)) < TOL
    assert _rel(_disp_grad(r.grad, 9, True), _disp_grad(o["grad"], 9, True)) < TOL_FIELD


def test_inner_solves_stop_at_the_outer_tolerance(capi):
    """An inner FP32 solve stops at inner_tol (1e-4) relative or, earlier,
    once the outer test will pass with a factor of two to spare.  At tol 1e-8
    the two inner solves both run to 1e-4.  At tol 1e-6 the second one needs
    only about two of its four decades: the solve still meets its tolerance
    with three FP64 residuals, in clearly fewer inner iterations (were every
    inner solve run to 1e-4, the two solves would do the same work)."""
    import ctypes
    from muspectre_amd._device import DeviceBuffer
    lib = capi.load()
    res = [16, 24, 32]
    N = int(np.prod(res))
    n = N * 9
    h = ctypes.c_void_p()
    capi.check(lib.mugpu_cell_create(3, (ctypes.c_int * 3)(*res), (ctypes.c_double * 3)(1., 1., 1.),
                                     capi.SMALL_STRAIN, capi.FP32, 1, ctypes.byref(h)))
    try:
        hard = np.random.default_rng(3).random(N) < 0.3
        idx = np.arange(N, dtype=np.int64)
        for E, sel in ((100e9, hard), (10e9, ~hard)):
            m = ctypes.c_int()
            capi.check(lib.mugpu_cell_add_material(h, capi.LE1, E, 0.3, 0.0, 0.0, ctypes.byref(m)))
            pix = np.ascontiguousarray(idx[sel])
            capi.check(lib.mugpu_material_add_pixels(
                h, m.value, len(pix), pix.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), None))
        capi.check(lib.mugpu_cell_initialise(h))
        eps = np.zeros((3, 3))
        eps[0, 1] = eps[1, 0] = 0.01
        capi.check(lib.mugpu_cell_set_uniform_strain(h, eps.ravel(order="F").ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))
        capi.check(lib.mugpu_cell_evaluate_stress_tangent(h))
        P = ctypes.c_void_p()
        capi.check(lib.mugpu_cell_stress(h, ctypes.byref(P)))
        rhs = DeviceBuffer(8 * n)
        capi.check(lib.mugpu_scale_copy(-1.0, P, rhs.ptr, n))
        capi.check(lib.mugpu_cell_apply_projection(h, rhs.ptr))
        out = {}
        for tol in (1e-8, 1e-6):
            x = DeviceBuffer(8 * n)
            its, refs, rel = ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
            capi.check(lib.mugpu_cg_solve_mixed(h, rhs.ptr, x.ptr, tol, 2000, 1e-4, ctypes.byref(its),
                                                ctypes.byref(refs), ctypes.byref(rel)))
            out[tol] = (its.value, refs.value, rel.value)
    finally:
        lib.mugpu_cell_destroy(h)
    its8, refs8, rel8 = out[1e-8]
    its6, refs6, rel6 = out[1e-6]
    assert rel8 <= 1e-8 and rel6 <= 1e-6, out
    assert refs6 == 3 and refs8 >= 3, out
    assert its6 < 0.9 * its8, out
