"""CPU checks that pin the oracle's constitutive laws and tensor functions.

Pins: the reference's golden values for HEP1 (two states) and expm; what
HEP1's return mapping fixes at every step of a many-step path;
numpy.linalg.eigh for logm/expm on separated and repeated spectra; the
explicit Curnier forms of the LE1 finite-strain P and K used by the
reference (tests/test_goodies.hh:94-134); closed-form Hooke stresses
(test_material_linear_elastic3/4.cc); the elastic-modulus identities
(test_materials_toolbox.cc:211-268); central finite differences for every
tangent that is a derivative.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "reference_vectors.json")))


def _hep1_params():
    g = GOLDEN["hep1"]
    K, mu = g["K"], g["mu"]
    E = 9 * K * mu / (3 * K + mu)
    nu = (3 * K - 2 * mu) / (2 * (3 * K + mu))
    return E, nu, g["tau_y0"], g["H"]


def test_hep1_golden_two_states(oracle):
    E, nu, ty, H = _hep1_params()
    I = np.eye(3)
    g = GOLDEN["hep1"]
    F = np.eye(3)
    F[0, 1] = g["state1"]["F01"]
    tau, C, Fpc, bec, epc = oracle.hep1_eval(3, E, nu, ty, H, F, I, I, 0.0, 0.0)
    s = g["state1"]
    assert np.linalg.norm(tau - np.array(s["tau"])) < 1e-8
    assert np.linalg.norm(bec - np.array(s["be"])) < 1e-8
    assert abs(epc - s["eps_p"]) < 1e-8
    assert np.linalg.norm(C - np.array(s["C4"])) < 1e-8
    # second state: old values from the first evaluation, current buffer
    # holds the value of two cycles ago (statefield.hh:318-326): zero
    F2 = np.eye(3)
    F2[0, 1] = g["state2"]["F01"]
    tau2, C2, _, bec2, epc2 = oracle.hep1_eval(3, E, nu, ty, H, F2, Fpc, bec, epc, 0.0)
    s = g["state2"]
    assert np.linalg.norm(tau2 - np.array(s["tau"])) < 1e-8
    assert np.linalg.norm(bec2 - np.array(s["be"])) < 1e-8
    assert abs(epc2 - s["eps_p"]) < 1e-8
    assert np.linalg.norm(C2 - np.array(s["C4"])) < 1e-8


def test_hep1_eps_p_accumulates_on_current_buffer(oracle):
    """material_hyper_elasto_plastic1.hh:262: eps_p.current() += dgamma."""
    E, nu, ty, H = _hep1_params()
    F = np.eye(3)
    F[0, 1] = 0.2
    _, _, _, _, e1 = oracle.hep1_eval(3, E, nu, ty, H, F, np.eye(3), np.eye(3), 0.0, 0.0)
    _, _, _, _, e2 = oracle.hep1_eval(3, E, nu, ty, H, F, np.eye(3), np.eye(3), 0.0, e1)
    assert e1 > 0 and abs(e2 - 2 * e1) < 1e-15


def test_expm_golden(oracle):
    g = GOLDEN["expm"]
    A = np.array(g["input"])
    for fn in (oracle.expm, oracle.expm_eig):
        assert np.linalg.norm(fn(A) - np.array(g["expected"])) < 1e-8


def _spd(rng, spread):
    Q = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    return Q @ np.diag(1 + spread * rng.random(3)) @ Q.T


@pytest.mark.parametrize("seed", range(6))
def test_logm_expm_vs_eigh(oracle, seed):
    rng = np.random.default_rng(seed)
    A = _spd(rng, 0.5)
    w, V = np.linalg.eigh(A)
    assert np.abs(oracle.logm(A) - V @ np.diag(np.log(w)) @ V.T).max() < 1e-14
    S = 0.3 * (A - np.eye(3))
    w, V = np.linalg.eigh(S)
    assert np.abs(oracle.expm(S) - V @ np.diag(np.exp(w)) @ V.T).max() < 1e-14
    # the reference's literal Sylvester form agrees on separated spectra
    assert np.abs(oracle.logm_sylvester(A) - oracle.logm(A)).max() < 1e-13


@pytest.mark.parametrize("diag", [[1, 1, 1.010025], [0.99, 1.02, 1.02], [1.1, 1.1, 1.1]])
def test_logm_expm_repeated_eigenvalues(oracle, diag):
    rng = np.random.default_rng(1)
    Q = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    A = Q @ np.diag(diag) @ Q.T
    A = 0.5 * (A + A.T)
    ref = Q @ np.diag(np.log(diag)) @ Q.T
    assert np.abs(oracle.logm(A) - ref).max() < 1e-13
    assert np.abs(oracle.expm(oracle.logm(A)) - A).max() < 1e-13


def test_logm_2d(oracle):
    A = np.array([[1.2, 0.1], [0.1, 0.9]])
    w, V = np.linalg.eigh(A)
    assert np.abs(oracle.logm(A) - V @ np.diag(np.log(w)) @ V.T).max() < 1e-14


def test_jacobi_eigensystem(oracle):
    rng = np.random.default_rng(7)
    for spread in (0.5, 1e-9, 0.0):
        A = _spd(rng, spread)
        ev, V = oracle.jacobi_sym3(A)
        assert np.abs(V.T @ V - np.eye(3)).max() < 1e-14
        assert np.abs(V @ np.diag(ev) @ V.T - A).max() < 1e-14


def _lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def test_le1_finite_strain_vs_curnier(oracle):
    """P and K of LE1 in finite strain against the explicit forms of Curnier
    (tests/test_goodies.hh:94-134, test_material_linear_elastic1.cc:134-234)."""
    E, nu = 210e9, .33
    lam, mu = _lame(E, nu)
    F = np.eye(3) + 0.05 * np.random.default_rng(2).standard_normal((3, 3))
    P, K = oracle.le1_eval(3, 0, E, nu, F)
    I = np.eye(3)
    Fr = (F * F).sum()
    Pc = lam / 2 * (Fr - 3) * F + mu * F @ (F.T @ F - I)
    assert np.abs(P - Pc).max() < 1e-12 * np.abs(Pc).max()
    FtF, FFt = F.T @ F, F @ F.T
    Kc = np.zeros((9, 9))
    for i in range(3):
        for j in range(3):
            for m in range(3):
                for n in range(3):
                    Kc[i + 3 * m, j + 3 * n] = (lam * ((Fr - 3) / 2 * I[i, j] * I[m, n] + F[i, m] * F[j, n]) +
                                                mu * (I[i, j] * FtF[m, n] + FFt[i, j] * I[m, n] -
                                                      I[i, j] * I[m, n] + F[i, n] * F[j, m]))
    assert np.abs(K - Kc).max() < 1e-12 * np.abs(Kc).max()


def test_le1_small_strain_hooke(oracle):
    E, nu = 10., .3
    lam, mu = _lame(E, nu)
    eps = np.array([[0.01, 0.002, 0], [0.002, -0.004, 0.001], [0, 0.001, 0.003]])
    P, K = oracle.le1_eval(3, 1, E, nu, eps)
    assert np.abs(P - (lam * np.trace(eps) * np.eye(3) + 2 * mu * eps)).max() < 1e-15
    # the tangent is the Hooke tensor: contracting it with eps gives the stress
    assert np.abs(K @ eps.ravel(order="F") - P.ravel(order="F")).max() < 1e-15


def test_moduli_identities():
    """materials_toolbox.hh:488-525 conversions round-trip
    (test_materials_toolbox.cc:211-268)."""
    E, nu = 70e9, .33
    lam, mu = _lame(E, nu)
    K = E / (3 * (1 - 2 * nu))
    assert abs(9 * K * mu / (3 * K + mu) - E) < 1e-4
    assert abs((3 * K - 2 * mu) / (2 * (3 * K + mu)) - nu) < 1e-14
    assert abs(mu * (3 * lam + 2 * mu) / (lam + mu) - E) < 1e-4


def _fd_tangent(fun, F, h=1e-7):
    d = F.shape[0]
    K = np.zeros((d * d, d * d))
    for j in range(d):
        for n in range(d):
            dF = np.zeros_like(F)
            dF[j, n] = h
            K[:, j + d * n] = ((fun(F + dF) - fun(F - dF)) / (2 * h)).ravel(order="F")
    return K


def test_pk2_to_pk1_tangent_is_derivative(oracle):
    """materials_toolbox.hh:313-350: K = dP/dF for P = F S(E(F))."""
    E, nu = 210e9, .33
    F = np.eye(3) + 0.03 * np.random.default_rng(4).standard_normal((3, 3))
    _, K = oracle.le1_eval(3, 0, E, nu, F)
    Kfd = _fd_tangent(lambda G: oracle.le1_eval(3, 0, E, nu, G)[0], F)
    assert np.abs(K - Kfd).max() < 1e-6 * np.abs(K).max()


@pytest.mark.parametrize("seed", range(4))
def test_hep1_consistent_tangent_is_derivative(oracle, seed):
    rng = np.random.default_rng(seed)
    E, nu, ty, H = 70e9, .33, 200e6, 1e9
    Fp = np.eye(3) + 0.01 * rng.standard_normal((3, 3))
    be = np.eye(3) + 0.002 * rng.standard_normal((3, 3))
    be = 0.5 * (be + be.T)
    F = Fp + (0.003 if seed % 2 else 0.0005) * rng.standard_normal((3, 3))
    P, K = oracle.hep1_stress_and_tangent(3, E, nu, ty, H, F, Fp, be, 0.001)
    Kfd = _fd_tangent(lambda G: oracle.hep1_stress_and_tangent(3, E, nu, ty, H, G, Fp, be, 0.001,
                                                               False)[0], F)
    assert np.abs(K - Kfd).max() < 1e-7 * np.abs(Kfd).max()


def _spd_logm(A):
    w, V = np.linalg.eigh(0.5 * (A + A.T))
    return (V * np.log(w)) @ V.T


def _hep1_path(oracle, steps, amp, seed, E=70e9, nu=.33, ty=200e6, H=1e9):
    """HEP1 over a loading path, its state cycled as the reference's state
    fields cycle after each step (statefield.hh:318-326): F_prev and be take
    the last step's values as old; the current eps_p buffer holds the value
    of two steps back, which the step adds its increment to
    (material_hyper_elasto_plastic1.hh:262).  Yields (F, F_prev_old, be_old,
    eps_old, result of the step, eps_p increment) per step."""
    rng = np.random.default_rng(seed)
    Fprev, be = np.eye(3), np.eye(3)
    eps = [0.0, 0.0]  # values of two steps back and of the last step
    G = np.zeros((3, 3))
    for _ in range(steps):
        # a random walk of the distortion: loading and unloading, with and
        # without rotation
        G = G + amp * rng.standard_normal((3, 3))
        F = np.eye(3) + G
        out = oracle.hep1_eval(3, E, nu, ty, H, F, Fprev, be, eps[1], eps[0])
        dgamma = out[4] - eps[0]
        yield F, Fprev, be, eps[1], out, dgamma
        Fprev, be = out[2], out[3]
        eps = [eps[1], out[4]]


def test_hep1_return_mapping_invariants_over_a_path(oracle):
    """Sixteen steps of HEP1 with history, each checked against what the
    return mapping (material_hyper_elasto_plastic1.hh:227-269) fixes exactly
    for any state: F_prev becomes F; the flow direction is deviatoric, so
    det(be_new) = det(f be_old f^T) with f = F F_prev_old^-1; the stress is
    the Hencky stress of the new be, tau = lambda tr(e) I + 2 mu e with
    e = log(be_new) / 2, and commutes with be_new; the increment is not
    negative, and the equivalent stress sqrt(3/2 tau_d:tau_d) lies on the
    updated yield surface tau_y0 + H (eps_old + dgamma) after a plastic step
    and within the old one after an elastic step.  A wrong sign or factor in
    the return, in the yield function or in the Hooke law fails one of
    these; the path has both kinds of step."""
    E, nu, ty, H = 70e9, .33, 200e6, 1e9
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    plastic = elastic = 0
    for F, Fp, be, eps_old, out, dgamma in _hep1_path(oracle, 16, 0.002, 7, E, nu, ty, H):
        tau, _, Fpc, bec, _ = out
        assert np.array_equal(Fpc, F)
        f = F @ np.linalg.inv(Fp)
        bstar = f @ be @ f.T
        assert abs(np.linalg.det(bec) / np.linalg.det(bstar) - 1) < 1e-12
        e = 0.5 * _spd_logm(bec)
        hencky = lam * np.trace(e) * np.eye(3) + 2 * mu * e
        assert np.abs(tau - hencky).max() < 1e-9 * np.abs(tau).max()
        assert np.abs(tau @ bec - bec @ tau).max() < 1e-9 * np.abs(tau).max()
        td = tau - np.trace(tau) / 3 * np.eye(3)
        teq = np.sqrt(1.5 * (td * td.T).sum())
        assert dgamma >= 0.0
        if dgamma > 0:
            plastic += 1
            assert abs(teq - (ty + H * (eps_old + dgamma))) < 1e-9 * ty
        else:
            elastic += 1
            assert teq <= ty + H * eps_old
    assert plastic >= 3 and elastic >= 3, (plastic, elastic)


def test_hep1_elastic_path_be_is_left_cauchy_green(oracle):
    """Without plastic flow be is the left Cauchy-Green tensor F F^T at
    every step, whatever the path, and the stress is its Hencky stress; the
    relative deformation f = F F_prev^-1 taken in the other order,
    F_prev^-1 F, fails this from the second step on (the shears of the
    golden states commute, so they cannot tell the orders apart)."""
    E, nu, ty, H = 70e9, .33, 200e6, 1e9
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    for F, _, _, _, out, dgamma in _hep1_path(oracle, 8, 0.0002, 3, E, nu, ty, H):
        tau, _, _, bec, epc = out
        assert dgamma == 0.0 and epc == 0.0
        b = F @ F.T
        assert np.abs(bec - b).max() < 1e-13
        e = 0.5 * _spd_logm(b)
        assert np.abs(tau - (lam * np.trace(e) * np.eye(3) + 2 * mu * e)).max() < 1e-9 * np.abs(tau).max()


def test_hep1_coded_tangent_is_j_independent(oracle):
    """materials_toolbox.hh:394-406 as coded: K(i,m,j,n) does not depend on j."""
    E, nu, ty, H = 70e9, .33, 200e6, 1e9
    F = np.eye(3) + 0.004 * np.random.default_rng(0).standard_normal((3, 3))
    _, K = oracle.hep1_stress_and_tangent(3, E, nu, ty, H, F, np.eye(3), np.eye(3), 0.0, False)
    K4 = K.reshape(3, 3, 3, 3, order="F")  # (i, m, j, n)
    assert np.abs(K4 - K4[:, :, :1, :]).max() == 0.0


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("seed", range(3))
def test_kirchhoff_to_pk1_determined_by_identities(oracle, d, seed):
    """The coded Kirchhoff -> PK1 conversion (materials_toolbox.hh:356-410)
    on random inputs with no symmetry: P = tau F^-T, and the tangent
    K(i,m,j,n) = sum_r Finv(i,r) sum_s C(r,m,n,s), which is fixed uniquely by
    two identities that do not restate it: K does not depend on j, and
    contracting its first index with F from the left gives the sums of C
    over its last index, sum_i F(k,i) K(i,m,j,n) = sum_s C(k,m,n,s) (F is
    invertible, so no other K satisfies this).  A transposed inverse,
    Finv(r,i), or a swapped pair of C's indices fails it."""
    import ctypes
    rng = np.random.default_rng(100 + seed)
    F = np.eye(d) + 0.3 * rng.standard_normal((d, d))
    tau = rng.standard_normal((d, d))
    C = rng.standard_normal(d ** 4)
    Ff, tf = np.asfortranarray(F), np.asfortranarray(tau)
    P = np.zeros((d, d), order="F")
    K = np.zeros(d ** 4)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    oracle.lib().mo_pk1_from_kirchhoff(d, dp(Ff), dp(tf), dp(C), dp(P), dp(K))
    K4 = K.reshape(d, d, d, d, order="F")  # (i, m, j, n)
    C4 = C.reshape(d, d, d, d, order="F")  # (r, m, n, s)
    assert np.abs(P @ F.T - tau).max() <= 1e-13 * np.abs(tau).max()
    assert np.abs(K4 - K4[:, :, :1, :]).max() == 0.0
    S = C4.sum(axis=3)  # (k, m, n)
    FK = np.einsum("ki,imjn->kmjn", F, K4)
    for j in range(d):
        assert np.abs(FK[:, :, j, :] - S).max() <= 1e-12 * np.abs(S).max()


def _generic_cell(oracle, dim, form, Cv, F):
    """A 3^dim cell of one MaterialLinearElasticGeneric under a uniform F;
    returns P and K of pixel 0."""
    c = oracle.Cell([3] * dim, [1.0] * dim, form)
    m = c.add_material_generic(Cv)
    for p in range(3 ** dim):
        c.add_pixel(m, p)
    c.initialise()
    c.F[:] = np.asarray(F).ravel(order="F")
    c.evaluate_stress_tangent()
    nc = dim * dim
    # copies: the views die with the cell
    return (np.array(c.P[0]).reshape(dim, dim, order="F"),
            np.array(c.K[0]).reshape(nc, nc, order="F"))


def _voigt_iso(dim, lam, mu):
    """test_material_linear_elastic_generic.cc:63-72: lambda + 2 mu on the
    normal block diagonal, lambda off it, mu on the shear diagonal."""
    vs = 3 if dim == 2 else 6
    C = np.zeros((vs, vs))
    C[:dim, :dim] = lam
    C[:dim, :dim] += 2 * mu * np.eye(dim)
    C[dim:, dim:] += mu * np.eye(vs - dim)
    return C


@pytest.mark.parametrize("dim", [2, 3])
def test_generic_isotropic_voigt_is_hooke(oracle, dim):
    """Golden case of test_material_linear_elastic_generic.cc:52-92: the Voigt
    matrix of lambda = 2, mu = 1.5 gives Hooke's C_ijkl = lambda d_ij d_kl +
    mu (d_ik d_jl + d_il d_jk); and the generic law equals LE1 with the same
    moduli in both formulations."""
    lam, mu = 2.0, 1.5
    I = np.eye(dim)
    nc = dim * dim
    Ch = np.zeros((nc, nc))
    for i in range(dim):
        for j in range(dim):
            for k in range(dim):
                for l in range(dim):
                    Ch[i + dim * j, k + dim * l] = lam * I[i, j] * I[k, l] + mu * (I[i, k] * I[j, l] + I[i, l] * I[j, k])
    _, K = _generic_cell(oracle, dim, 1, _voigt_iso(dim, lam, mu), np.zeros((dim, dim)))
    assert np.abs(K - Ch).max() < 1e-14
    E, nu = mu * (3 * lam + 2 * mu) / (lam + mu), lam / (2 * (lam + mu))
    F = I + 0.04 * np.random.default_rng(9).standard_normal((dim, dim))
    # small strain: a symmetric strain (C : eps drops the skew part, which
    # LE1's lambda tr(eps) I + 2 mu eps keeps)
    eps = 0.5 * ((F - I) + (F - I).T)
    for form in (0, 1):
        G = F if form == 0 else eps
        P, K = _generic_cell(oracle, dim, form, _voigt_iso(dim, lam, mu), G)
        Pl, Kl = oracle.le1_eval(dim, form, E, nu, G)
        assert np.abs(P - Pl).max() < 1e-13 * np.abs(Pl).max()
        assert np.abs(K - Kl).max() < 1e-13 * np.abs(Kl).max()


def _voigt_strain(e):
    """[e11, e22, (e33,) 2e23, 2e13, 2e12] (engineering shears)"""
    if e.shape[0] == 2:
        return np.array([e[0, 0], e[1, 1], 2 * e[0, 1]])
    return np.array([e[0, 0], e[1, 1], e[2, 2], 2 * e[1, 2], 2 * e[0, 2], 2 * e[0, 1]])


def _voigt_stress(s):
    if s.shape[0] == 2:
        return np.array([s[0, 0], s[1, 1], s[0, 1]])
    return np.array([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]])


@pytest.mark.parametrize("dim", [2, 3])
def test_generic_anisotropic_voigt_relation(oracle, dim):
    """Small strain: sigma_voigt = C_voigt eps_voigt with engineering shear
    strains, the textbook meaning of a Voigt stiffness (fixes the component
    order xx, yy, zz, yz, xz, xy and the shear factors); the tangent
    contracted with eps gives sigma."""
    vs = 3 if dim == 2 else 6
    A = np.random.default_rng(dim).standard_normal((vs, vs))
    Cv = A @ A.T + vs * np.eye(vs)
    e = 0.01 * np.random.default_rng(10 + dim).standard_normal((dim, dim))
    e = 0.5 * (e + e.T)
    P, K = _generic_cell(oracle, dim, 1, Cv, e)
    assert np.abs(P - P.T).max() < 1e-15
    assert np.abs(_voigt_stress(P) - Cv @ _voigt_strain(e)).max() < 1e-14 * np.abs(Cv).max()
    assert np.abs(K @ e.ravel(order="F") - P.ravel(order="F")).max() < 1e-14 * np.abs(Cv).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_generic_finite_strain(oracle, dim):
    """Finite strain: S from the Voigt relation on the Green-Lagrange strain,
    P = F S, and K = dP/dF by finite differences."""
    vs = 3 if dim == 2 else 6
    A = np.random.default_rng(20 + dim).standard_normal((vs, vs))
    Cv = A @ A.T + vs * np.eye(vs)
    F = np.eye(dim) + 0.03 * np.random.default_rng(30 + dim).standard_normal((dim, dim))
    P, K = _generic_cell(oracle, dim, 0, Cv, F)
    E = 0.5 * (F.T @ F - np.eye(dim))
    Sv = Cv @ _voigt_strain(E)
    S = np.zeros((dim, dim))
    idx = [(0, 0), (1, 1), (0, 1)] if dim == 2 else [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for a, (i, j) in enumerate(idx):
        S[i, j] = S[j, i] = Sv[a]
    assert np.abs(P - F @ S).max() < 1e-13 * np.abs(P).max()
    Kfd = _fd_tangent(lambda G: _generic_cell(oracle, dim, 0, Cv, G)[0], F)
    assert np.abs(K - Kfd).max() < 1e-6 * np.abs(K).max()
