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
