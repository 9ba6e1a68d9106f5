"""CPU checks of the oracle's cell, SolverCG and Newton drivers.

Pins: the exact Reuss laminate solution (test_solver_newton_cg.cc:101-323),
homogeneous LE4 stress (python_material_linear_elastic4_test.py:82-86),
eigenstrain equivalence (python_binding_tests.py:151-194), de Geus vs
Newton-CG (test_solver_newton_cg.cc:55-99), a dense linear solve of the
same operator for SolverCG, and the reference's error cases.
"""
import numpy as np
import pytest


def _laminate(oracle):
    c = oracle.Cell([3, 3], [1., 1.], oracle.SMALL_STRAIN)
    hard = c.add_material(oracle.LE1, 4., .33)
    soft = c.add_material(oracle.LE1, 2., .33)
    for i in range(9):
        c.add_pixel(hard if i // 3 < 1 else soft, i)
    return c


@pytest.mark.parametrize("driver", ["newton_cg", "de_geus"])
def test_laminate_reuss_exact(oracle, driver):
    factor = 0.5 / 3 + 1 - 1 / 3
    es, eh = 1 / factor, 0.5 / factor
    for load, comps in [(np.array([[1., 0], [0, 0]]), [0]),
                        (np.array([[0, 1.], [1., 0]]), [1, 2])]:
        c = _laminate(oracle)
        r = getattr(c, driver)([load], 1e-8, 20, 1e-5, 1e-10)[0]
        for i in range(9):
            ref = np.zeros(4)
            ref[comps] = eh if i // 3 < 1 else es
            assert np.abs(r["grad"][i] - ref).max() < 1e-12


def test_le4_homogeneous(oracle):
    c = oracle.Cell([7, 7], [2.3, 3.9], oracle.SMALL_STRAIN)
    m = c.add_material(oracle.LE4)
    for i in range(49):
        c.add_pixel(m, i, [10., .3])
    load = np.array([[0, 0.025], [0.025, 0]])
    r = c.newton_cg([load], 1e-6, 100, 1e-6, 1e-6)[0]
    mu = 10. / (2 * 1.3)
    assert np.linalg.norm(r["stress"] - (2 * mu * load).ravel(order="F")[None]) < 1e-8


def test_eigenstrain_equivalence(oracle):
    grad = np.array([[1.1, .2], [.3, 1.5]])
    gl = -0.5 * (grad.T + grad - 2 * np.eye(2))
    c1 = oracle.Cell([3, 3], [3., 3.], oracle.SMALL_STRAIN)
    c2 = oracle.Cell([3, 3], [3., 3.], oracle.SMALL_STRAIN)
    m1 = c1.add_material(oracle.LE1, 210e9, .33)
    m2 = c2.add_material(oracle.LE2, 210e9, .33)
    for i in range(9):
        c1.add_pixel(m1, i)
        c2.add_pixel(m2, i, gl)
    P1 = c1.newton_cg([-gl], 1e-6, 2, 1e-6, 1e-6)[0]["stress"]
    P2 = c2.newton_cg([np.zeros((2, 2))], 1e-6, 2, 1e-6, 1e-6)[0]["stress"]
    assert np.linalg.norm(P1 - P2) / np.linalg.norm(.5 * (P1 + P2)) < 1e-6


def test_de_geus_equals_newton_cg_5cubed(oracle):
    Young, Poisson = 1.0030648180242636, 0.29930675909878679
    out = []
    for driver in ("de_geus", "newton_cg"):
        c = oracle.Cell([5, 5, 5], [5., 5., 5.], oracle.FINITE_STRAIN)
        h = c.add_material(oracle.LE1, 10 * Young, Poisson)
        s = c.add_material(oracle.LE1, Young, Poisson)
        for i in range(125):
            c.add_pixel(h if i == 0 else s, i)
        load = np.zeros((3, 3))
        load[0, 1] = 1.
        out.append(getattr(c, driver)([load], 1e-8, 125 * 90, 1e-5, 1e-10)[0]["grad"])
    assert np.abs(out[0] - out[1]).mean() < 1e-8


def test_cg_solution_equals_dense_solve(oracle):
    """SolverCG's x solves G K x = b on the compatible subspace: compare with a
    dense solve of the operator restricted to range(G), both assembled column
    by column (G K G is SPD there, so that solution is unique)."""
    res = [5, 3]
    c = oracle.Cell(res, [1., 1.], oracle.SMALL_STRAIN)
    a = c.add_material(oracle.LE1, 3., .3)
    b = c.add_material(oracle.LE1, 1., .3)
    for i in range(15):
        c.add_pixel(a if i in (0, 4, 7) else b, i)
    c.initialise()
    e = np.array([[0.01, 0.003], [0.003, -0.002]])
    c.F[:] = e.ravel(order="F")
    c.evaluate_stress_tangent()
    rhs = -c.apply_projection(c.P.copy())
    x = c.cg_solve(rhs, 1e-12, 200)
    n = 15 * 4
    A = np.zeros((n, n))
    G = np.zeros((n, n))
    for k in range(n):
        u = np.zeros(n)
        u[k] = 1.0
        A[:, k] = c.directional_stiffness(u.reshape(15, 4)).ravel()
        G[:, k] = c.apply_projection(u.reshape(15, 4)).ravel()
    w, V = np.linalg.eigh(0.5 * (G + G.T))
    Q = V[:, w > 0.5]                      # orthonormal basis of range(G)
    y = np.linalg.solve(Q.T @ A @ Q, Q.T @ rhs.ravel())
    xd = (Q @ y).reshape(15, 4)
    assert np.abs(A @ x.ravel() - rhs.ravel()).max() < 1e-10 * np.abs(rhs).max()
    assert np.abs(x - xd).max() < 1e-8 * np.abs(xd).max()


def test_counter_is_cumulative(oracle):
    c = _laminate(oracle)
    steps = [np.array([[1., 0], [0, 0]]) * k for k in (1, 2)]
    r = c.newton_cg(steps, 1e-8, 20, 1e-5, 1e-10)
    assert r[1]["nb_fev"] >= r[0]["nb_fev"]


def test_convergence_error(oracle):
    c = oracle.Cell([11, 11], [1., 1.], oracle.SMALL_STRAIN)
    a = c.add_material(oracle.LE1, 100., .3)
    b = c.add_material(oracle.LE1, 1., .3)
    for i in range(121):
        x, y = divmod(i, 11)
        c.add_pixel(a if (x - 5) ** 2 + (y - 5) ** 2 < 9 else b, i)
    with pytest.raises(oracle.OracleError) as e:
        c.newton_cg([np.array([[.01, 0], [0, 0]])], 1e-12, 2, 1e-5, 1e-10)
    assert e.value.kind == "ConvergenceError"


def test_asymmetric_small_strain_load(oracle):
    c = _laminate(oracle)
    with pytest.raises(oracle.OracleError) as e:
        c.newton_cg([np.array([[0, 1.], [0, 0]])], 1e-8, 10, 1e-5, 1e-10)
    assert e.value.kind == "SolverError"


def test_material_coverage(oracle):
    c = oracle.Cell([3, 3], [1., 1.], oracle.SMALL_STRAIN)
    m = c.add_material(oracle.LE1, 1., .3)
    c.add_pixel(m, 0)
    with pytest.raises(oracle.OracleError):
        c.initialise()
    c2 = oracle.Cell([3, 3], [1., 1.], oracle.SMALL_STRAIN)
    m2 = c2.add_material(oracle.LE1, 1., .3)
    for i in range(9):
        c2.add_pixel(m2, i)
    c2.add_pixel(m2, 4)
    with pytest.raises(oracle.OracleError):
        c2.initialise()


def test_hep1_as_coded_c3_load_raises(oracle):
    """The reference's HEP1 tangent makes CG diverge on the uniaxial C3 load:
    ConvergenceError is the oracle outcome (9^3, one load step)."""
    rng = np.random.default_rng(1)
    mask = rng.random(729) > 0.5
    c = oracle.Cell([9, 9, 9], [1., 1., 1.], oracle.FINITE_STRAIN)
    s = c.add_material(oracle.HEP1, 70e9, .33, 200e6, 1e9)
    h = c.add_material(oracle.HEP1, 210e9, .33, 800e6, 1e9)
    for i in range(729):
        c.add_pixel(h if mask[i] else s, i)
    with pytest.raises(oracle.OracleError) as e:
        c.newton_cg([np.diag([0, 0, 0.005])], 1e-8, 300, 1e-5, 1e-10)
    assert e.value.kind == "ConvergenceError"


def test_hep1_consistent_first_step_converges(oracle):
    rng = np.random.default_rng(1)
    mask = rng.random(729) > 0.5
    c = oracle.Cell([9, 9, 9], [1., 1., 1.], oracle.FINITE_STRAIN)
    s = c.add_material(oracle.HEP1_CT, 70e9, .33, 200e6, 1e9)
    h = c.add_material(oracle.HEP1_CT, 210e9, .33, 800e6, 1e9)
    for i in range(729):
        c.add_pixel(h if mask[i] else s, i)
    r = c.newton_cg([np.diag([0, 0, 0.005])], 1e-8, 2000, 1e-5, 1e-10)[0]
    assert r["success"] and r["nb_it"] <= 8
