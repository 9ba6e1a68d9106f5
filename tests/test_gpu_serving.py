"""The serving loop on the GPU.

A request is one small-strain newton_cg load step (src/solver/solvers.cc:
44-217) with the mixed-precision SolverCG on per-voxel isotropic elasticity
(MaterialLinearElastic4).  Checks:
  * a request served alone computes exactly what newton_cg computes on a Cell
    with the same material and precision (bit-identical fields and counts),
    which ties the server to the parity-tested cell path;
  * every request of a mix (several resolutions, more requests than slots,
    slots changing hands between batched launches) gives exactly its alone
    result;
  * a request whose CG fails (ConvergenceError) fails alone and leaves the
    others untouched;
  * admission keeps the device footprint within the budget;
  * replaying captured graphs and launching kernel by kernel agree.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def µ():
    import muspectre_amd
    return muspectre_amd


@pytest.fixture(scope="module")
def serving():
    from muspectre_amd import serving
    return serving


def _grf_phase(res, ell, seed):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(res)
    k2 = sum(np.meshgrid(*[np.fft.fftfreq(n) ** 2 for n in res], indexing="ij"))
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * np.exp(-0.5 * (2 * np.pi) ** 2 * k2 * ell ** 2)))
    return f > np.median(f)


def _request(res, seed):
    """C4-style request: thresholded GRF phases, contrast log-uniform in
    [1, 100], nu in [0.2, 0.4], one symmetric strain with entries in +-0.01."""
    rng = np.random.default_rng(1000 + seed)
    mask = _grf_phase(res, max(2.0, res[0] / 16), seed)
    contrast = np.exp(rng.uniform(0, np.log(100)))
    E = np.where(mask.ravel(), 10e9 * contrast, 10e9)
    nu = np.full(mask.size, rng.uniform(0.2, 0.4))
    d = len(res)
    a = rng.uniform(-0.01, 0.01, (d, d))
    return {"res": list(res), "E": E, "nu": nu, "strain": 0.5 * (a + a.T)}


def _serve(serving, reqs, **kw):
    kw.setdefault("hbm_budget", 2 << 30)
    srv = serving.RVEServer(keep_fields=True, **kw)
    ids = [srv.submit(q["res"], q["E"], q["nu"], q["strain"], cg_maxiter=q.get("maxiter", 0))
           for q in reqs]
    srv.run()
    done = set(srv.poll())
    assert done == set(ids)
    out = [srv.result(i, fields=True) for i in ids]
    st = srv.stats()
    srv.close()
    return out, st


def _alone(serving, reqs, **kw):
    return [_serve(serving, [q], **kw)[0][0] for q in reqs]


def _same(a, b):
    assert a["status"] == b["status"] and a["converged"] == b["converged"]
    assert a["newton_iterations"] == b["newton_iterations"]
    assert a["cg_iterations"] == b["cg_iterations"]
    np.testing.assert_array_equal(a["strain"], b["strain"])
    np.testing.assert_array_equal(a["stress"], b["stress"])
    np.testing.assert_array_equal(a["mean_stress"], b["mean_stress"])


@pytest.mark.parametrize("res", [[64, 64], [32, 32, 32], [48, 16, 32], [96, 64, 32], [512, 256]],
                         ids=lambda r: "x".join(map(str, r)))
def test_request_equals_cell_newton_cg(µ, serving, res):
    """A served request computes exactly what newton_cg computes on a Cell;
    96x64x32 and 512x256 are large enough for submit to split the
    per-voxel Lame parameters over three and two host threads."""
    q = _request(res, 3)
    out, _ = _serve(serving, [q])
    r = out[0]
    d = len(res)
    cell = µ.Cell(res, [1.0] * d, µ.Formulation.small_strain, precision="fp32", allow_even=True)
    mat = getattr(µ.material, "MaterialLinearElastic4_{}d".format(d)).make(cell, "rve")
    mat.add_pixels_by_index(np.arange(q["E"].size), np.stack([q["E"], q["nu"]], axis=1))
    solver = µ.solvers.SolverCG(cell, 1e-6, 2000, False)
    ref = µ.solvers.newton_cg(cell, [q["strain"]], solver, 1e-5, 1e-10, 0)[0]
    nc = d * d
    assert r["ok"] and r["converged"] == ref.success
    assert r["newton_iterations"] == ref.nb_it and r["cg_iterations"] == ref.nb_fev
    np.testing.assert_array_equal(r["strain"], np.asarray(ref.grad).reshape(-1, nc))
    np.testing.assert_array_equal(r["stress"], np.asarray(ref.stress).reshape(-1, nc))
    mean = np.asarray(ref.stress).reshape(-1, nc).mean(0)
    assert np.abs(r["mean_stress"] - mean).max() <= 1e-13 * np.abs(mean).max()


@pytest.mark.parametrize("res,lengths", [([32, 16, 48], [2.0, 1.0, 0.5]), ([64, 32], [7.0, 5.0])],
                         ids=["32x16x48", "64x32"])
def test_request_lengths_equal_cell(µ, serving, res, lengths):
    """A request with unequal cell lengths (the projection's wave vectors
    scale with 1/L) computes exactly what newton_cg computes on a Cell of the
    same lengths, and not what it computes on the unit cell."""
    q = _request(res, 11)
    d = len(res)
    srv = serving.RVEServer(hbm_budget=2 << 30, keep_fields=True)
    try:
        rid = srv.submit(q["res"], q["E"], q["nu"], q["strain"], lengths=lengths)
        unit = srv.submit(q["res"], q["E"], q["nu"], q["strain"])
        srv.run()
        r, u = srv.result(rid, fields=True), srv.result(unit, fields=True)
    finally:
        srv.close()
    cell = µ.Cell(res, lengths, µ.Formulation.small_strain, precision="fp32", allow_even=True)
    mat = getattr(µ.material, "MaterialLinearElastic4_{}d".format(d)).make(cell, "rve")
    mat.add_pixels_by_index(np.arange(q["E"].size), np.stack([q["E"], q["nu"]], axis=1))
    solver = µ.solvers.SolverCG(cell, 1e-6, 2000, False)
    ref = µ.solvers.newton_cg(cell, [q["strain"]], solver, 1e-5, 1e-10, 0)[0]
    nc = d * d
    assert r["ok"] and r["converged"] == ref.success
    assert r["newton_iterations"] == ref.nb_it and r["cg_iterations"] == ref.nb_fev
    np.testing.assert_array_equal(r["strain"], np.asarray(ref.grad).reshape(-1, nc))
    np.testing.assert_array_equal(r["stress"], np.asarray(ref.stress).reshape(-1, nc))
    assert np.abs(r["strain"] - u["strain"]).max() > 1e-3 * np.abs(u["strain"]).max()


MIX = [[64, 64], [32, 32, 32], [128, 128], [48, 48, 48], [32, 32, 32], [64, 64],
       [48, 48, 48], [32, 32, 32], [16, 16, 16], [32, 32, 32], [64, 64], [96, 32, 16]]


def test_alone_equals_mix(serving):
    reqs = [_request(r, i) for i, r in enumerate(MIX)]
    alone = _alone(serving, reqs)
    mixed, st = _serve(serving, reqs, max_batch=2, iters_per_launch=3)
    for a, b in zip(alone, mixed):
        assert a["ok"]
        _same(a, b)
    assert st["peak_running"] > 1 and st["batched_slots"] > st["batched_launches"]


def test_many_requests_of_one_resolution(serving):
    """70 requests of one small resolution admitted together: the material
    stage of the first pass runs as two launches (64 requests and 6), the
    inner iterations in batches of up to 64 slots, and every request still
    computes exactly its alone result."""
    reqs = [_request([16, 16, 16], 200 + i) for i in range(70)]
    alone = _alone(serving, reqs[:6]) + _alone(serving, reqs[64:])
    mixed, st = _serve(serving, reqs, max_batch=64, iters_per_launch=4)
    for a, b in zip(alone, mixed[:6] + mixed[64:]):
        assert a["ok"]
        _same(a, b)
    assert all(m["ok"] for m in mixed)
    assert st["peak_running"] == 70 and st["batched_slots"] > st["batched_launches"]


def test_failure_is_isolated(serving):
    reqs = [_request(r, 40 + i) for i, r in enumerate(MIX[:6])]
    reqs[2] = dict(reqs[2], maxiter=3)
    alone = _alone(serving, reqs)
    mixed, st = _serve(serving, reqs, max_batch=4, iters_per_launch=4)
    assert not alone[2]["ok"] and alone[2]["status"] == 1  # ConvergenceError
    for a, b in zip(alone, mixed):
        _same(a, b)
    assert st["failed"] == 1


def test_graphs_and_direct_launches_agree(serving):
    reqs = [_request(r, 70 + i) for i, r in enumerate(MIX[:5])]
    g, _ = _serve(serving, reqs, max_batch=4, iters_per_launch=5, use_graphs=True)
    d, _ = _serve(serving, reqs, max_batch=4, iters_per_launch=2, use_graphs=False)
    for a, b in zip(g, d):
        _same(a, b)


def test_admission_keeps_budget(serving):
    fp = serving.RVEServer.footprint([32, 32, 32])
    budget = 2 * fp + fp // 2
    reqs = [_request([32, 32, 32], 90 + i) for i in range(5)]
    out, st = _serve(serving, reqs, hbm_budget=budget)
    assert all(o["ok"] for o in out)
    assert st["peak_running"] == 2 and st["peak_bytes"] <= budget
    srv = serving.RVEServer(hbm_budget=budget)
    big = _request([64, 64, 64], 1)
    with pytest.raises(RuntimeError):
        srv.submit(big["res"], big["E"], big["nu"], big["strain"])
    with pytest.raises(RuntimeError):
        q = _request([32, 32, 32], 2)
        srv.submit(q["res"], q["E"], q["nu"], np.array([[0, 0.01, 0], [0, 0, 0], [0, 0, 0]]))
    with pytest.raises(RuntimeError):
        q = _request([16, 16, 16], 2)
        srv.submit([15, 16, 16], np.ones(15 * 256), np.full(15 * 256, .3), q["strain"])
    srv.close()


@pytest.mark.parametrize("res", [[32, 32, 32], [64, 64]], ids=lambda r: "x".join(map(str, r)))
def test_served_request_vs_oracle(serving, oracle, res):
    """A served request against the oracle's FP64 newton_cg on the same
    per-voxel LE4 cell and strain (the serving loop runs FP32 CG with FP64
    refinement): Newton count equal, homogenised stress to 1e-6, strain
    field to 1e-5, both solves to cg_tol 1e-8."""
    q = _request(res, 11)
    out, _ = _serve(serving, [q], cg_tol=1e-8, cg_maxiter=4000)
    r = out[0]
    d = len(res)
    nc = d * d
    oc = oracle.Cell(res, [1.0] * d, oracle.SMALL_STRAIN, allow_even=True)
    m = oc.add_material(oracle.LE4)
    for i in range(q["E"].size):
        oc.add_pixel(m, i, [q["E"][i], q["nu"][i]])
    o = oc.newton_cg([q["strain"]], 1e-8, 4000, 1e-5, 1e-10)[0]
    ref = o["stress"].mean(0)
    es = np.linalg.norm(r["mean_stress"].ravel() - ref) / np.linalg.norm(ref)
    eg = np.linalg.norm(r["strain"].reshape(-1, nc) - o["grad"]) / np.linalg.norm(o["grad"])
    print("served {}: stress {:.2e}, strain {:.2e}".format("x".join(map(str, res)), es, eg),
          flush=True)
    assert r["ok"] and r["newton_iterations"] == o["nb_it"]
    assert es <= 1e-6
    assert eg <= 1e-5
