"""CPU checks that pin the oracle's FFT, frequencies and projections.

Pins: numpy.fft (a library routine) and a brute-force DFT for the
transforms; the reference's golden values for fft_freqs and the wave vectors
(tests/test_fft_utils.cc:43-83); for the projections, the defining
properties of an orthogonal projector onto compatible fields (idempotent,
self-adjoint, preserves gradients of periodic displacements, its output is a
gradient, zero mean) and the reference's even-grid rejection
(tests/test_projection_finite.cc:71-77) - none of which restates the
operator's formula.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "reference_vectors.json")))


def test_fft_freqs_golden(oracle):
    for case in GOLDEN["fft_freqs"].values():
        if not isinstance(case, dict):
            continue
        out = oracle.fft_freqs(case["n"], case["length"])
        assert np.abs(out - np.array(case["expected"])).sum() == 0.0


def test_wave_vectors_golden(oracle):
    wv = GOLDEN["wave_vectors"]
    f = [oracle.fft_freqs(n, L) for n, L in zip(wv["sizes"], wv["lengths"])]
    for case in wv["cases"]:
        xi = np.array([f[d][c] for d, c in enumerate(case["ccoord"])])
        assert np.linalg.norm(xi - np.array(case["xi"])) < 1e-14


@pytest.mark.parametrize("res", [[6, 4], [5, 5], [3, 5, 7], [8, 6, 4], [51, 51], [17, 4],
                                 [9, 11, 13], [2, 3, 2], [1, 10], [125, 3]])
def test_r2c_c2r_vs_numpy(oracle, res):
    d = len(res)
    rng = np.random.default_rng(len(res) * 100 + res[0])
    x = rng.random([*res, d, d])
    ref = np.fft.rfftn(x, axes=tuple(range(d)))
    out = oracle.fft_r2c(res, x.reshape(-1, d * d))
    assert np.linalg.norm(out.reshape(ref.shape) - ref) < 1e-14 * np.prod(res)
    herm = oracle.hermitian_sizes(res)
    c = rng.random([*herm, d, d]) + 1j * rng.random([*herm, d, d])
    ref2 = np.fft.irfftn(c, s=res, axes=tuple(range(d)))
    out2 = oracle.fft_c2r(res, c.reshape(-1, d * d)) / np.prod(res)
    assert np.linalg.norm(out2.reshape(ref2.shape) - ref2) < 1e-14 * np.prod(res)


@pytest.mark.parametrize("res", [[3, 5], [4, 3], [3, 2, 5]])
def test_r2c_brute_force_dft(oracle, res):
    d = len(res)
    nc = d * d
    x = np.random.default_rng(0).standard_normal((int(np.prod(res)), nc))
    herm = oracle.hermitian_sizes(res)
    grid = np.stack(np.meshgrid(*[np.arange(n) for n in res], indexing="ij"), -1).reshape(-1, d)
    kgrid = np.stack(np.meshgrid(*[np.arange(n) for n in herm], indexing="ij"), -1).reshape(-1, d)
    phase = np.exp(-2j * np.pi * sum(np.outer(kgrid[:, a], grid[:, a]) / res[a] for a in range(d)))
    ref = phase @ x
    out = oracle.fft_r2c(res, x)
    assert np.abs(out - ref).max() < 1e-12


def test_round_trip_and_dc_real(oracle):
    """test_fftw_engine.cc:91-135: ifft(fft(x))/N = x and the DC bin is real."""
    res = [7, 6, 5]
    x = np.random.default_rng(1).random((210, 9))
    s = oracle.fft_r2c(res, x)
    assert np.abs(s[0].imag).max() == 0.0
    back = oracle.fft_c2r(res, s) / 210
    assert np.abs(back - x).max() < 1e-14


# ---- projections -------------------------------------------------------------
CASES = [([5, 5], [5., 5.]), ([3, 5], [1.3, 2.1]), ([7, 7, 7], [7., 7., 7.]),
         ([3, 5, 7], [1., 2., 3.]), ([9, 7], [2., 1.])]


def _gradient_field(res, lengths, sym, seed):
    """Gradient of a random periodic displacement field, computed by finite
    sums of Fourier modes (exactly representable on the grid)."""
    d = len(res)
    rng = np.random.default_rng(seed)
    x = np.stack(np.meshgrid(*[np.arange(n) * L / n for n, L in zip(res, lengths)],
                             indexing="ij"), -1).reshape(-1, d)
    G = np.zeros((x.shape[0], d, d))
    for _ in range(4):
        k = np.array([rng.integers(-(n // 2), n // 2 + 1) for n in res])
        if not k.any():
            continue
        q = 2 * np.pi * k / np.array(lengths)
        a, ph = rng.standard_normal(d), rng.uniform(0, 2 * np.pi)
        arg = x @ q + ph
        # u = a sin(q.x + ph) -> grad u_i,j = a_i q_j cos(q.x + ph)
        G += np.cos(arg)[:, None, None] * np.outer(a, q)[None]
    if sym:
        G = 0.5 * (G + G.transpose(0, 2, 1))
    return G.transpose(0, 2, 1).reshape(-1, d * d)  # component i + d j of G(i,j)


@pytest.mark.parametrize("res,lengths", CASES)
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_projection_is_orthogonal_projector(oracle, res, lengths, kind):
    d = len(res)
    N = int(np.prod(res))
    rng = np.random.default_rng(3)
    x = rng.standard_normal((N, d * d))
    y = rng.standard_normal((N, d * d))
    Gx = oracle.apply_projection(kind, res, lengths, x)
    Gy = oracle.apply_projection(kind, res, lengths, y)
    GGx = oracle.apply_projection(kind, res, lengths, Gx)
    assert np.abs(GGx - Gx).max() < 1e-13 * np.abs(Gx).max()          # idempotent
    assert abs((Gx * y).sum() - (x * Gy).sum()) < 1e-12 * np.abs(x).sum()  # self-adjoint
    assert np.abs(Gx.mean(0)).max() < 1e-13                          # zero mean
    g = _gradient_field(res, lengths, kind == 0, 11)
    g -= g.mean(0)
    Gg = oracle.apply_projection(kind, res, lengths, g)
    assert np.abs(Gg - g).max() < 1e-12 * np.abs(g).max()              # keeps gradients


@pytest.mark.parametrize("res,lengths", CASES)
@pytest.mark.parametrize("kind", [0, 1])
def test_projection_output_is_compatible(oracle, res, lengths, kind):
    """Fourier components of the output are a (x) xi (finite strain) or
    sym(a (x) xi) (small strain): rank <= 2 in 3D, no part orthogonal to xi."""
    d = len(res)
    N = int(np.prod(res))
    x = np.random.default_rng(5).standard_normal((N, d * d))
    Gx = oracle.apply_projection(kind, res, lengths, x)
    s = oracle.fft_r2c(res, Gx)
    op = oracle.projection_operator(1, res, lengths)   # unit wave vectors
    for k in range(1, s.shape[0]):
        xi = op[k]
        if not xi.any():
            continue
        Fk = s[k].reshape(d, d, order="F")
        # a vector t orthogonal to xi must be annihilated on the right side
        t = np.linalg.svd(xi[None, :])[2][1:]  # basis of xi^perp
        if kind == 1:
            assert np.abs(Fk @ t.T).max() < 1e-10 * max(1, np.abs(Fk).max())
        else:
            assert np.abs(t @ Fk @ t.T).max() < 1e-10 * max(1, np.abs(Fk).max())
            assert np.abs(Fk - Fk.T).max() < 1e-10 * max(1, np.abs(Fk).max())


def test_finite_projections_agree(oracle):
    """ProjectionFiniteStrain (full G, outer_under(I, xi xi^T)) and the fast
    variant are the same operator."""
    res, L = [5, 7, 3], [1., 2., 3.]
    x = np.random.default_rng(2).standard_normal((105, 9))
    a = oracle.apply_projection(1, res, L, x)
    b = oracle.apply_projection(2, res, L, x)
    assert np.abs(a - b).max() < 1e-13 * np.abs(a).max()


def test_even_grid_rejected(oracle):
    with pytest.raises(oracle.OracleError) as e:
        oracle.apply_projection(1, [4, 5], [1., 1.], np.zeros((20, 4)))
    assert e.value.kind == "ProjectionError"


@pytest.mark.parametrize("res", [[8, 8], [6, 4, 6]])
def test_even_policy_is_projector(oracle, res):
    d = len(res)
    x = np.random.default_rng(4).standard_normal((int(np.prod(res)), d * d))
    for kind in (0, 1):
        Gx = oracle.apply_projection(kind, res, [1.] * d, x, allow_even=True)
        GGx = oracle.apply_projection(kind, res, [1.] * d, Gx, allow_even=True)
        assert np.abs(GGx - Gx).max() < 1e-13 * np.abs(Gx).max()
