"""The reference's Python-level tests, run against this build's _muSpectre
(the HIP path), plus full newton_cg / de_geus parity against the oracle.

Sources of the checks: tests/python_binding_tests.py:47-194,
python_fft_tests.py:53-101, python_material_linear_elastic3_test.py:44-85,
python_material_linear_elastic4_test.py:44-86, test_solver_newton_cg.cc:
55-449 (reference repository).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def µ():
    import muspectre_amd
    return muspectre_amd


# ---- python_fft_tests.py:53-101 ------------------------------------------
@pytest.mark.parametrize("res", [[6, 4], [7, 5], [4, 6, 5], [5, 5, 5]])
def test_fft_forward_reverse(µ, res):
    dim = len(res)
    tol = 1e-14 * np.prod(res)
    engine = µ.fft.FFT(res, dim ** 2)
    engine.initialise()
    in_arr = np.random.random([*res, dim, dim])
    out_ref = np.fft.rfftn(in_arr, axes=tuple(range(dim)))
    out_msp = engine.fft(in_arr.reshape(-1, dim ** 2).T).T
    assert np.linalg.norm(out_ref - out_msp.reshape(out_ref.shape)) < tol
    complex_res = µ.get_hermitian_sizes(res)
    c = np.zeros([*complex_res, dim, dim], dtype=complex)
    c.real = np.random.random(c.shape)
    c.imag = np.random.random(c.shape)
    ref = np.fft.irfftn(c, s=res, axes=tuple(range(dim)))
    out = engine.ifft(c.reshape(-1, dim ** 2).T).T * engine.normalisation()
    assert np.linalg.norm(ref - out.reshape(ref.shape)) < tol


def test_fft_requires_initialise(µ):
    engine = µ.fft.FFT([5, 5], 4)
    with pytest.raises(RuntimeError):
        engine.fft(np.zeros((4, 25)))


# ---- python_binding_tests.py:47-91 ---------------------------------------
def test_cell_construction_and_add_pixel(µ):
    sys = µ.Cell([5, 7], [5.2, 8.3], µ.Formulation.small_strain)
    mat = µ.material.MaterialLinearElastic1_2d.make(sys, "material", 210e9, .33)
    mat.add_pixel([2, 1])
    assert mat.size() == 1
    assert len(sys) == 35
    pix = list(sys)
    assert pix[0] == [0, 0] and pix[1] == [0, 1] and pix[-1] == [4, 6]


def test_even_grid_rejected(µ):
    with pytest.raises(µ.ProjectionError):
        µ.Cell([4, 5], [1., 1.], µ.Formulation.small_strain)


def test_unassigned_pixel_raises(µ):
    sys = µ.Cell([3, 3], [1., 1.], µ.Formulation.small_strain)
    mat = µ.material.MaterialLinearElastic1_2d.make(sys, "m", 1., .3)
    mat.add_pixel([0, 0])
    with pytest.raises(RuntimeError):
        sys.initialise()


def test_duplicate_pixel_raises(µ):
    sys = µ.Cell([3, 3], [1., 1.], µ.Formulation.small_strain)
    a = µ.material.MaterialLinearElastic1_2d.make(sys, "a", 1., .3)
    b = µ.material.MaterialLinearElastic1_2d.make(sys, "b", 2., .3)
    for p in sys:
        a.add_pixel(p)
    b.add_pixel([1, 1])
    with pytest.raises(RuntimeError):
        sys.initialise()


def test_solver_finite_strain_de_geus(µ):
    sys = µ.Cell([3, 3], [3., 3.], µ.Formulation.finite_strain)
    hard = µ.material.MaterialLinearElastic1_2d.make(sys, "hard", 210e9, .33)
    soft = µ.material.MaterialLinearElastic1_2d.make(sys, "soft", 70e9, .33)
    for i, pixel in enumerate(sys):
        (hard if i < 3 else soft).add_pixel(pixel)
    sys.initialise()
    solver = µ.solvers.SolverCG(sys, 1e-6, 100, 0)
    r = µ.solvers.de_geus(sys, np.array([[0, .1], [0, 0]]), solver, 1e-6, 0)
    assert r.success and r.nb_it >= 1


# ---- python_binding_tests.py:105-194: eigenstrain equivalence -------------
def test_eigenstrain_solve(µ):
    cell1 = µ.Cell([3, 3], [3., 3.], µ.Formulation.small_strain)
    cell2 = µ.Cell([3, 3], [3., 3.], µ.Formulation.small_strain)
    mat1 = µ.material.MaterialLinearElastic1_2d.make(cell1, "simple", 210e9, .33)
    mat2 = µ.material.MaterialLinearElastic2_2d.make(cell2, "eigen", 210e9, .33)
    grad = np.array([[1.1, .2], [.3, 1.5]])
    gl_strain = -0.5 * (grad.T + grad - 2 * np.eye(2))
    grad = -gl_strain
    for pixel in cell1:
        mat1.add_pixel(pixel)
        mat2.add_pixel(pixel, gl_strain)
    cell1.initialise()
    cell2.initialise()
    tol = 1e-6

    def solve(cell, g):
        solver = µ.solvers.SolverCG(cell, tol, 2, 0)
        return µ.solvers.newton_cg(cell, g, solver, tol, tol, 0)

    r1 = solve(cell1, grad)
    r2 = solve(cell2, np.zeros_like(grad))
    P1, P2 = r1.stress, r2.stress
    assert np.linalg.norm(P1 - P2) / np.linalg.norm(.5 * (P1 + P2)) < tol
    eig = cell2.get_globalised_internal_real_array("Eigenstrain")
    assert eig.shape == (4, 9)
    np.testing.assert_allclose(eig[:, 0], gl_strain.ravel(order="F"))


# ---- python_material_linear_elastic{3,4}_test.py ---------------------------
@pytest.mark.parametrize("which,res,lengths", [("MaterialLinearElastic3_2d", [5, 5], [2.5, 3.1]),
                                               ("MaterialLinearElastic4_2d", [7, 7], [2.3, 3.9])])
def test_le34_homogeneous(µ, which, res, lengths):
    sys = µ.Cell(res, lengths, µ.Formulation.small_strain)
    mat = getattr(µ.material, which).make(sys, "material")
    Young, Poisson = 10., .3
    for pixel in sys:
        mat.add_pixel(pixel, Young, Poisson)
    sys.initialise()
    tol = 1e-6
    Del0 = np.array([[0, 0.025], [0.025, 0]])
    solver = µ.solvers.SolverCG(sys, tol, 100, False)
    r = µ.solvers.newton_cg(sys, Del0, solver, tol, tol, False)
    mu = Young / (2 * (1 + Poisson))
    stress = 2 * mu * Del0
    assert np.linalg.norm(r.stress.reshape(-1, 4) - stress.reshape(1, 4)) < 1e-8


# ---- test_solver_newton_cg.cc:206-323: laminate patch test (Reuss), with
# every solver of the reference's SolverList (:212-216) ---------------------
SOLVERS = ["SolverCG", "SolverCGEigen", "SolverGMRESEigen", "SolverDGMRESEigen",
           "SolverBiCGSTABEigen", "SolverMINRESEigen"]


@pytest.mark.parametrize("solver_name", SOLVERS)
@pytest.mark.parametrize("driver", ["newton_cg", "de_geus"])
def test_laminate_patch_exact(µ, driver, solver_name):
    sys = µ.Cell([3, 3], [1., 1.], µ.Formulation.small_strain)
    contrast, Young, Poisson = 2., 2., .33
    hard = µ.material.MaterialLinearElastic1_2d.make(sys, "hard", contrast * Young, Poisson)
    soft = µ.material.MaterialLinearElastic1_2d.make(sys, "soft", Young, Poisson)
    for pixel in sys:
        (hard if pixel[0] < 1 else soft).add_pixel(pixel)
    sys.initialise()
    factor = 1 / contrast * 1 / 3 + 1 - 1 / 3
    eps_soft = 1. / factor
    eps_hard = eps_soft / contrast
    for load, comps in [(np.array([[1., 0], [0, 0]]), [0]), (np.array([[0, 1.], [1., 0]]), [1, 2])]:
        solver = getattr(µ.solvers, solver_name)(sys, 1e-8, 20, False)
        r = getattr(µ.solvers, driver)(sys, load, solver, 1e-5, 1e-10, 0)
        F = r.grad.reshape(-1, 4)
        for i, pixel in enumerate(sys):
            ref = np.zeros(4)
            ref[comps] = eps_hard if pixel[0] < 1 else eps_soft
            assert np.linalg.norm(F[i] - ref) < 1e-12


# ---- full newton_cg parity against the oracle ------------------------------
def _disc(n, r):
    x = np.arange(n)
    return np.hypot(x[:, None] - n // 2, x[None, :] - n // 2) < r


def test_small_case_c1_parity(µ, oracle):
    """bin/small_case.py (config 1): 51x51, disc of radius 10, LE1, one step,
    equil_tol = 2 as the script passes it positionally."""
    res, lengths = [51, 51], [7., 5.]
    mask = _disc(51, 10)
    cell = µ.Cell(res, lengths, µ.Formulation.small_strain)
    inc = µ.material.MaterialLinearElastic1_2d.make(cell, "incl", 10e9, .33)
    mat = µ.material.MaterialLinearElastic1_2d.make(cell, "matrix", 70e9, .33)
    for i, pixel in enumerate(cell):
        (inc if mask.flat[i] else mat).add_pixel(pixel)
    solver = µ.solvers.SolverCG(cell, 1e-8, 401, 0)
    load = np.array([[0, 0], [0, .03]])
    r = µ.solvers.newton_cg(cell, load, solver, 1e-5, 2, 0)
    oc = oracle.Cell(res, lengths, oracle.SMALL_STRAIN)
    a = oc.add_material(oracle.LE1, 10e9, .33)
    b = oc.add_material(oracle.LE1, 70e9, .33)
    for i in range(51 * 51):
        oc.add_pixel(a if mask.flat[i] else b, i)
    ro = oc.newton_cg([load], 1e-8, 401, 1e-5, 2.0)[0]
    assert r.nb_it == ro["nb_it"] and r.nb_fev == ro["nb_fev"]
    assert r.message == ro["message"]
    F, P = r.grad.ravel(), r.stress.ravel()
    assert np.linalg.norm(F - ro["grad"].ravel()) <= 1e-10 * np.linalg.norm(ro["grad"])
    assert np.linalg.norm(P - ro["stress"].ravel()) <= 1e-10 * np.linalg.norm(ro["stress"])


@pytest.mark.parametrize("driver", ["newton_cg", "de_geus"])
@pytest.mark.parametrize("form", ["small", "finite"])
def test_3d_two_phase_multi_step_parity(µ, oracle, form, driver):
    """newton_cg (solvers.cc:44-217) and de_geus (solvers.cc:220-408) over
    two load steps of a heterogeneous 3D cell, against the oracle's same
    driver: identical Newton and cumulative CG counts, fields to 1e-10 (the
    oracle reproduces its own counts and fields to 3e-15 with 1, 2, 3 and 8
    threads on this cell)."""
    res = [9, 11, 7]
    g = np.meshgrid(*[np.arange(n) for n in res], indexing="ij")
    mask = ((g[0] - 4) ** 2 + (g[1] - 5) ** 2 + (g[2] - 3) ** 2) < 9
    F_ = µ.Formulation.small_strain if form == "small" else µ.Formulation.finite_strain
    cell = µ.Cell(res, [1., 1.2, .8], F_)
    h = µ.material.MaterialLinearElastic1_3d.make(cell, "hard", 100e9, .3)
    s = µ.material.MaterialLinearElastic1_3d.make(cell, "soft", 10e9, .3)
    idx = np.arange(mask.size)
    h.add_pixels_by_index(idx[mask.ravel()])
    s.add_pixels_by_index(idx[~mask.ravel()])
    steps = [np.array([[.01, .002, 0], [.002, 0, 0], [0, 0, -.005]]) * k for k in (1, 2)]
    solver = µ.solvers.SolverCG(cell, 1e-8, 400, 0)
    rs = getattr(µ.solvers, driver)(cell, steps, solver, 1e-5, 1e-10, 0)
    oc = oracle.Cell(res, [1., 1.2, .8], oracle.SMALL_STRAIN if form == "small" else oracle.FINITE_STRAIN)
    a = oc.add_material(oracle.LE1, 100e9, .3)
    b = oc.add_material(oracle.LE1, 10e9, .3)
    for i in range(mask.size):
        oc.add_pixel(a if mask.flat[i] else b, i)
    ro = getattr(oc, driver)(steps, 1e-8, 400, 1e-5, 1e-10)
    assert len(rs) == 2
    for r, o in zip(rs, ro):
        assert r.nb_it == o["nb_it"] and r.nb_fev == o["nb_fev"]
        assert np.linalg.norm(r.grad.ravel() - o["grad"].ravel()) <= 1e-10 * np.linalg.norm(o["grad"])
        assert np.linalg.norm(r.stress.ravel() - o["stress"].ravel()) <= 1e-10 * np.linalg.norm(o["stress"])


def test_de_geus_equals_newton_cg(µ):
    """test_solver_newton_cg.cc:55-99 on 5^3 (finite strain, one hard voxel)."""
    Young, Poisson = 1.0030648180242636, 0.29930675909878679
    out = []
    for driver in ("de_geus", "newton_cg"):
        sys = µ.Cell([5, 5, 5], [5., 5., 5.], µ.Formulation.finite_strain)
        hard = µ.material.MaterialLinearElastic1_3d.make(sys, "hard", 10 * Young, Poisson)
        soft = µ.material.MaterialLinearElastic1_3d.make(sys, "soft", Young, Poisson)
        for i, pixel in enumerate(sys):
            (hard if i == 0 else soft).add_pixel(pixel)
        sys.initialise()
        delF0 = np.zeros((3, 3))
        delF0[0, 1] = 1.
        solver = µ.solvers.SolverCG(sys, 1e-8, 125 * 9 * 10, 0)
        out.append(getattr(µ.solvers, driver)(sys, [delF0], solver, 1e-5, 1e-10, 0)[0].grad)
    assert np.abs(out[0] - out[1]).mean() <= 1e-8


def test_convergence_error_on_tiny_maxiter(µ):
    sys = µ.Cell([11, 11], [1., 1.], µ.Formulation.small_strain)
    a = µ.material.MaterialLinearElastic1_2d.make(sys, "a", 100., .3)
    b = µ.material.MaterialLinearElastic1_2d.make(sys, "b", 1., .3)
    for i, p in enumerate(sys):
        (a if (p[0] - 5) ** 2 + (p[1] - 5) ** 2 < 9 else b).add_pixel(p)
    solver = µ.solvers.SolverCG(sys, 1e-12, 2, 0)
    with pytest.raises(µ.ConvergenceError):
        µ.solvers.newton_cg(sys, np.array([[.01, 0], [0, 0]]), solver, 1e-5, 1e-10, 0)


def test_asymmetric_small_strain_load_raises(µ):
    sys = µ.Cell([3, 3], [1., 1.], µ.Formulation.small_strain)
    a = µ.material.MaterialLinearElastic1_2d.make(sys, "a", 1., .3)
    for p in sys:
        a.add_pixel(p)
    solver = µ.solvers.SolverCG(sys, 1e-8, 10, 0)
    with pytest.raises(µ.SolverError):
        µ.solvers.newton_cg(sys, np.array([[0, 1.], [0, 0]]), solver, 1e-5, 1e-10, 0)


def test_projection_api_vs_oracle(µ, oracle):
    res, L = [5, 7, 9], [1., 2., 3.]
    f = np.random.default_rng(1).standard_normal((9, int(np.prod(res))))
    for form, kind in [(µ.Formulation.small_strain, 0), (µ.Formulation.finite_strain, 1)]:
        proj = µ.fft.Projection(res, L, form)
        proj.initialise()
        out = proj.apply_projection(f)
        ref = oracle.apply_projection(kind, res, L, np.asfortranarray(f).T.reshape(-1, 9)).T
        assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("form", ["small_strain", "finite_strain"])
def test_generic_material_equals_le1(µ, form):
    """MaterialLinearElasticGeneric_3d.make(cell, name, C_voigt) with the
    Voigt matrix of an isotropic law solves exactly as LE1 with the same
    moduli (material_linear_elastic_generic.hh:164-176 vs
    material_linear_elastic1.hh); a wrong-sized matrix raises."""
    res = [7, 9, 5]
    lam, mu = 2e9, 1.5e9
    E, nu = mu * (3 * lam + 2 * mu) / (lam + mu), lam / (2 * (lam + mu))
    Cv = np.zeros((6, 6))
    Cv[:3, :3] = lam
    Cv[:3, :3] += 2 * mu * np.eye(3)
    Cv[3:, 3:] = mu * np.eye(3)
    F = getattr(µ.Formulation, form)
    out = []
    for generic in (True, False):
        cell = µ.Cell(res, [1.] * 3, F)
        if generic:
            soft = µ.material.MaterialLinearElasticGeneric_3d.make(cell, "soft", Cv)
        else:
            soft = µ.material.MaterialLinearElastic1_3d.make(cell, "soft", E, nu)
        hard = µ.material.MaterialLinearElastic1_3d.make(cell, "hard", 5 * E, nu)
        for i, pix in enumerate(cell):
            (hard if i % 5 == 0 else soft).add_pixel(pix)
        cell.initialise()
        solver = µ.solvers.SolverCG(cell, 1e-8, 200, False)
        g = np.zeros((3, 3))
        g[0, 1] = g[1, 0] = 0.01
        r = µ.solvers.newton_cg(cell, g, solver, 1e-6, 1e-10, 0)
        out.append((np.asarray(r.grad).copy(), np.asarray(r.stress).copy(), r.nb_fev))
    assert out[0][2] == out[1][2]
    assert np.abs(out[0][0] - out[1][0]).max() <= 1e-12 * np.abs(out[1][0]).max()
    assert np.abs(out[0][1] - out[1][1]).max() <= 1e-10 * np.abs(out[1][1]).max()
    cell = µ.Cell(res, [1.] * 3, F)
    with pytest.raises(RuntimeError):
        µ.material.MaterialLinearElasticGeneric_3d.make(cell, "bad", np.eye(3))


@pytest.mark.parametrize("solver_name", SOLVERS[1:])
def test_eigen_solvers_small_case(µ, solver_name):
    """bin/small_case.py:74-95 runs config 1 (51x51, disc of radius 10, LE1)
    with every Eigen solver; each reaches the SolverCG solution to within
    what two solves stopped at cg_tol = 1e-8 can differ by."""
    res, lengths = [51, 51], [7., 5.]
    mask = _disc(51, 10)
    out = {}
    for name in ("SolverCG", solver_name):
        cell = µ.Cell(res, lengths, µ.Formulation.small_strain)
        inc = µ.material.MaterialLinearElastic1_2d.make(cell, "incl", 10e9, .33)
        mat = µ.material.MaterialLinearElastic1_2d.make(cell, "matrix", 70e9, .33)
        for i, pixel in enumerate(cell):
            (inc if mask.flat[i] else mat).add_pixel(pixel)
        solver = getattr(µ.solvers, name)(cell, 1e-8, 401, False)
        assert solver.name() == {"SolverCG": "CG", "SolverCGEigen": "CG", "SolverGMRESEigen": "GMRES",
                                 "SolverDGMRESEigen": "DGMRES", "SolverBiCGSTABEigen": "BiCGSTAB",
                                 "SolverMINRESEigen": "MINRES"}[name]
        r = µ.solvers.newton_cg(cell, np.array([[0, 0], [0, .03]]), solver, 1e-5, 2, 0)
        assert r.success and solver.get_counter() > 0
        out[name] = np.asarray(r.grad).copy()
    ref = out["SolverCG"]
    assert np.linalg.norm(out[solver_name] - ref) <= 1e-6 * np.linalg.norm(ref)


@pytest.mark.parametrize("solver_name", SOLVERS[1:])
def test_eigen_solver_convergence_error(µ, solver_name):
    """SolverEigen::solve (solver_eigen.cc:69-77): too few iterations raise
    ConvergenceError naming the solver."""
    cell = µ.Cell([51, 51], [7., 5.], µ.Formulation.small_strain)
    mask = _disc(51, 10)
    inc = µ.material.MaterialLinearElastic1_2d.make(cell, "incl", 10e9, .33)
    mat = µ.material.MaterialLinearElastic1_2d.make(cell, "matrix", 70e9, .33)
    for i, pixel in enumerate(cell):
        (inc if mask.flat[i] else mat).add_pixel(pixel)
    solver = getattr(µ.solvers, solver_name)(cell, 1e-8, 2, False)
    with pytest.raises(µ.ConvergenceError, match="has not converged"):
        µ.solvers.newton_cg(cell, np.array([[0, 0], [0, .03]]), solver, 1e-5, 2, 0)


# ---- python_field_tests.py:40-113: a material's field collection ----------
def test_field_collection_eigenstrain(µ):
    res, lengths = [3, 3], [1.58, 5.87]
    cell = µ.Cell(res, lengths, µ.Formulation.finite_strain)
    mat = µ.material.MaterialLinearElastic2_2d.make(cell, "material", 210e9, .33)
    eigen_strain = np.array([[.01, .02], [.03, -.01]])
    for i, pixel in enumerate(cell):
        mat.add_pixel(pixel, i / cell.size * eigen_strain)
    cell.initialise()
    assert isinstance(mat, µ.material.Material_2d)
    collection = mat.collection
    field_name = collection.field_names[0]
    assert field_name == "Eigenstrain"
    for getter in (collection.get_complex_field, collection.get_int_field,
                   collection.get_uint_field):
        with pytest.raises(Exception):
            getter(field_name)
    field = collection.get_real_field(field_name)
    for i, row in enumerate(field.array.T):
        assert np.linalg.norm(i / cell.size * eigen_strain - row.reshape(eigen_strain.shape).T) == 0
    glob = cell.get_globalised_internal_real_array("Eigenstrain")
    assert np.linalg.norm(field.array - glob) == 0


def test_field_collection_hep1_state_vs_oracle(µ, oracle):
    """HEP1's state fields through the collection: the current and old
    sub-fields match the oracle's current and old values across a history
    cycle (statefield.hh:245-326), and the current one is what
    get_globalised_internal_real_array returns."""
    res = [5, 7, 3]
    N = int(np.prod(res))
    rng = np.random.default_rng(8)
    hard = rng.random(N) < 0.5
    cell = µ.Cell(res, [1.0] * 3, µ.Formulation.finite_strain)
    mk = µ.material.MaterialHyperElastoPlastic1_3d.make
    soft = mk(cell, "soft", 70e9, .33, 200e6, 1e9)
    hm = mk(cell, "hard", 210e9, .33, 800e6, 1e9)
    idx = np.arange(N)
    hm.add_pixels_by_index(idx[hard])
    soft.add_pixels_by_index(idx[~hard])
    oc = oracle.Cell(res, [1.0] * 3, oracle.FINITE_STRAIN)
    a = oc.add_material(oracle.HEP1, 70e9, .33, 200e6, 1e9)
    b = oc.add_material(oracle.HEP1, 210e9, .33, 800e6, 1e9)
    for i in range(N):
        oc.add_pixel(b if hard[i] else a, i)
    cell.initialise()
    oc.initialise()
    col = hm.collection
    assert col.statefield_names == sorted(["cumulated plastic flow εₚ", "Previous placement gradient Fᵗ",
                                           "Previous left Cauchy-Green deformation bₑᵗ"])
    assert len(col.field_names) == 6
    names = {"eps_p": "cumulated plastic flow εₚ", "F_prev": "Previous placement gradient Fᵗ",
             "be": "Previous left Cauchy-Green deformation bₑᵗ"}
    for step in range(3):
        F = np.eye(3).ravel(order="F") + 0.01 * (step + 1) * rng.standard_normal((N, 9))
        cell.evaluate_stress_tangent(F.T)
        oc.F[:] = F
        oc.evaluate_stress_tangent()
        for short, prefix in names.items():
            sf = col.get_real_statefield(prefix)
            cur = sf.get_current_field().array.T
            old = sf.get_old_field().array.T
            for which, val in ((0, cur), (1, old)):
                ref = oc.get_internal(short, which)[hard]
                assert np.abs(val - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (short, which)
            glob = cell.get_globalised_internal_real_array(prefix).T[hard]
            np.testing.assert_array_equal(glob, cur)
        cell.save_history_variables()
        oc.save_history()


def test_cell_get_projection(µ, oracle):
    """cell.get_projection() (bind_py_cell.cc:197-198): the projection of
    the cell's formulation; apply_projection equals cell.project and the
    operator equals the oracle's stored one."""
    for form, kind in ((µ.Formulation.small_strain, oracle.PROJ_SMALL),
                       (µ.Formulation.finite_strain, oracle.PROJ_FINITE_FAST)):
        res, lengths = [5, 7, 9], [1.0, 1.3, 0.7]
        cell = µ.Cell(res, lengths, form)
        m = µ.material.MaterialLinearElastic1_3d.make(cell, "m", 70e9, .3)
        for pixel in cell:
            m.add_pixel(pixel)
        cell.initialise()
        proj = cell.get_projection()
        assert proj.get_formulation() == form
        assert list(proj.get_domain_resolutions()) == res
        field = np.random.default_rng(2).standard_normal((9, int(np.prod(res))))
        np.testing.assert_array_equal(proj.apply_projection(field), cell.project(field))
        op = proj.get_operator()
        ref = oracle.projection_operator(kind, res, lengths)
        assert op.shape == ref.T.shape
        assert np.abs(op - ref.T).max() <= 1e-14
