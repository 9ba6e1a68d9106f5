"""MaterialLinearElastic2, 3 and 4 with heterogeneous per-pixel parameters on
the GPU against the oracle, through the Python API: 2D and 3D, small and
finite strain, odd grids (the reference's runtime-length FFT) and
compiled-length grids (the benchmark passes, even-grid policy).

  LE2  two materials (E = 100 GPa and 10 GPa, nu = .3), a random eigenstrain
       per pixel (material_linear_elastic2.hh:178-201)
  LE3  two materials, per-pixel (E, nu) -> per-pixel C
       (material_linear_elastic3.cc:57-68)
  LE4  two materials, per-pixel (E, nu) -> per-pixel (lambda, mu)
       (material_linear_elastic4.cc:62-73, .hh:189-210)

Per-pixel moduli: E lognormal about 50 GPa (sigma .3), nu uniform in
[.15, .45].  Per case: stress and tangent at a random strain (1e-12), then a
newton_cg load step: Newton and CG counts and the strain and stress fields
against the oracle's, to 1e-10 and the exact count wherever the oracle
reproduces itself that closely under another summation order
(tests/_parity.py), to its own rounding spread otherwise; a solve whose
right-hand side is the residual the previous one left at 1e-8 may take one
iteration more or less (its rhs is known only to that rounding).

The C5-shaped case (LE4, lognormal E with median 50 GPa and sigma .8, nu
uniform in [.15, .45], shear eps_xy = .01) runs at 25x27x15 here, as the
first Newton iteration: on a linear problem newton_cg's second iteration asks
CG to reduce a residual already at the operator's rounding level by a
further 1e-8, which the oracle itself fails at this grid (ConvergenceError
after ~4000 iterations; the reference does the same).  The benchmark's
512-point passes with this material are covered in test_gpu_bench_kernels.py
(16x16x512 and its permutations).
"""
import ctypes

import numpy as np
import pytest

from _parity import (CHAOTIC_COUNT, bound, cg_with_spread, chaotic_bound, check_count,
                     newton_with_spread, rounding_level_solves)

pytestmark = pytest.mark.gpu

SIGMA = 0.3

CASES = [
    ("2d_small_odd", [15, 13], 1),
    ("2d_finite_odd", [9, 11], 0),
    ("3d_small_odd", [9, 7, 11], 1),
    ("3d_finite_odd", [7, 5, 9], 0),
    ("2d_small_pow2", [64, 32], 1),
    ("3d_small_pow2", [16, 32, 16], 1),
    ("3d_finite_pow2", [32, 16, 16], 0),
]
FAMILIES = ["le2", "le3", "le4"]


@pytest.fixture(scope="module")
def µ():
    import muspectre_amd
    return muspectre_amd


def _cells(µ, oracle, res, form, fam, seed):
    d = len(res)
    N = int(np.prod(res))
    even = any(n % 2 == 0 for n in res)
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(n) for n in res], indexing="ij")
    mask = (sum((gi - n // 2) ** 2 for gi, n in zip(g, res)) < (min(res) / 3) ** 2).ravel()
    F_ = µ.Formulation.small_strain if form == 1 else µ.Formulation.finite_strain
    cell = µ.Cell(res, [1.0] * d, F_, allow_even=even)
    oc = oracle.Cell(res, [1.0] * d, form, allow_even=even)
    mk = getattr(µ.material, {"le2": "MaterialLinearElastic2", "le3": "MaterialLinearElastic3",
                              "le4": "MaterialLinearElastic4"}[fam] + "_{}d".format(d)).make
    ofam = {"le2": oracle.LE2, "le3": oracle.LE3, "le4": oracle.LE4}[fam]
    idx = np.arange(N, dtype=np.int64)
    if fam == "le2":
        a = 0.002 * rng.standard_normal((N, d, d))
        eig = 0.5 * (a + np.swapaxes(a, 1, 2))
        params = eig.reshape(N, d * d, order="C")  # symmetric: either order
        gm = [mk(cell, "hard", 100e9, .3), mk(cell, "soft", 10e9, .3)]
        om = [oc.add_material(ofam, 100e9, .3), oc.add_material(ofam, 10e9, .3)]
        for sel, m, o in ((mask, gm[0], om[0]), (~mask, gm[1], om[1])):
            m.add_pixels_by_index(idx[sel], params[sel])
            for i in idx[sel]:
                oc.add_pixel(o, int(i), eig[i])
    else:
        E = np.exp(np.log(50e9) + SIGMA * rng.standard_normal(N))
        nu = rng.uniform(0.15, 0.45, N)
        gm = [mk(cell, "a"), mk(cell, "b")]
        om = [oc.add_material(ofam), oc.add_material(ofam)]
        for sel, m, o in ((mask, gm[0], om[0]), (~mask, gm[1], om[1])):
            m.add_pixels_by_index(idx[sel], np.stack([E[sel], nu[sel]], axis=1))
            for i in idx[sel]:
                oc.add_pixel(o, int(i), [E[i], nu[i]])
    return cell, oc


@pytest.mark.parametrize("fam", FAMILIES)
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_le_family_vs_oracle(µ, oracle, case, fam):
    name, res, form = case
    d = len(res)
    nc = d * d
    N = int(np.prod(res))
    cell, oc = _cells(µ, oracle, res, form, fam, 17)
    cell.initialise()
    oc.initialise()
    rng = np.random.default_rng(4)
    F = 0.01 * rng.standard_normal((N, nc))
    if form == 0:
        F += np.eye(d).ravel(order="F")
    cell.evaluate_stress_tangent(F.T)
    oc.F[:] = F
    oc.evaluate_stress_tangent()
    P = np.asarray(cell.get_stress()).T
    K = np.asarray(cell.get_tangent()).T
    assert np.abs(P - oc.P).max() <= 1e-12 * np.abs(oc.P).max()
    assert np.abs(K - oc.K).max() <= 1e-12 * np.abs(oc.K).max()

    load = np.zeros((d, d))
    load[0, 1] = load[1, 0] = 0.01
    load[0, 0] = -0.004
    cell2, _ = _cells(µ, oracle, res, form, fam, 17)
    solver = µ.solvers.SolverCG(cell2, 1e-8, 2000, False)
    r = µ.solvers.newton_cg(cell2, [load], solver, 1e-5, 1e-10, 0)[0]
    ref, counts, sg, ss, records = newton_with_spread(
        oracle, lambda: _cells(µ, oracle, res, form, fam, 17)[1], [load], 1e-8, 2000, 1e-5, 1e-10)
    o = ref[0]
    g = np.asarray(r.grad).reshape(-1, nc)
    s = np.asarray(r.stress).reshape(-1, nc)
    eg = np.linalg.norm(g - o["grad"]) / np.linalg.norm(o["grad"])
    es = np.linalg.norm(s - o["stress"]) / np.linalg.norm(o["stress"])
    print(name, fam, "Newton", r.nb_it, "CG", r.nb_fev, "oracle", counts,
          "strain {:.1e} (spread {:.1e}) stress {:.1e} (spread {:.1e})".format(eg, sg, es, ss),
          flush=True)
    assert r.nb_it == o["nb_it"]
    check_count(int(r.nb_fev), counts, allowance=rounding_level_solves(records))
    assert eg <= bound(sg)
    assert es <= bound(ss)


def _c5_cells(capi, oracle, res, prec):
    from test_gpu_abi_cell import GpuCell
    N = int(np.prod(res))
    rng = np.random.default_rng(2026)
    E = np.exp(np.log(50e9) + 0.8 * rng.standard_normal(N))
    nu = rng.uniform(0.15, 0.45, N)
    gc = GpuCell(capi, res, [1.0] * 3, 1, prec=prec)
    m = gc.add_material(4)
    gc.add_pixels(m, np.arange(N), np.stack([E, nu], axis=1))
    oc = oracle.Cell(res, [1.0] * 3, oracle.SMALL_STRAIN)
    om = oc.add_material(oracle.LE4)
    for i in range(N):
        oc.add_pixel(om, i, [E[i], nu[i]])
    oc.initialise()
    capi.check(gc.lib.mugpu_cell_initialise(gc.h))
    e = np.zeros((3, 3))
    e[0, 1] = e[1, 0] = 0.01
    F0 = np.tile(e.ravel(order="F"), (N, 1))
    oc.F[:] = F0
    gc.set_F(F0)
    oc.evaluate_stress_tangent()
    capi.check(gc.lib.mugpu_cell_evaluate_stress_tangent(gc.h))
    return oc, gc, F0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_c5_shaped_le4_odd_grid(capi, oracle, prec):
    """C5's material and load on an odd grid (the reference's own grids are
    odd), the first Newton iteration: rhs = -G:P at the macroscopic shear,
    SolverCG to 1e-8.  FP64: the oracle's count and solution within its own
    rounding spread (tests/_parity.py); FP32 operator with FP64 refinement:
    homogenised stress at F0 + x to 1e-6, x to 1e-5."""
    from muspectre_amd._device import DeviceBuffer
    from test_gpu_abi_cell import _gpu_cg_chunked
    res = [25, 27, 15]
    oc, gc, F0 = _c5_cells(capi, oracle, res, 0 if prec == "fp64" else 1)
    rhs = -oc.apply_projection(oc.P.copy())
    x_ref, it_ref, counts, spread = cg_with_spread(oracle, oc, rhs, 1e-8, 4000)
    if prec == "fp64":
        x, _, it, _, _ = _gpu_cg_chunked(gc, rhs, 1e-8, 4000, 8)
    else:
        b = DeviceBuffer.from_array(rhs)
        xb = DeviceBuffer(rhs.nbytes)
        its, refs, rel = ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
        capi.check(gc.lib.mugpu_cg_solve_mixed(gc.h, b.ptr, xb.ptr, 1e-8, 8000, 1e-4,
                                               ctypes.byref(its), ctypes.byref(refs),
                                               ctypes.byref(rel)))
        x, it = xb.download(rhs.shape), its.value
    ex = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    oc.F[:] = F0 + x_ref
    oc.evaluate_stress_tangent()
    s_ref = oc.P.mean(0)
    oc.F[:] = F0 + x
    oc.evaluate_stress_tangent()
    es = np.linalg.norm(oc.P.mean(0) - s_ref) / np.linalg.norm(s_ref)
    print("c5-shaped 25x27x15 {}: CG {} (oracle {}), x {:.2e} (spread {:.2e}), <stress> {:.2e}"
          .format(prec, it, counts, ex, spread, es), flush=True)
    if prec == "fp64":
        check_count(it, counts, allowance=CHAOTIC_COUNT)
        assert ex <= chaotic_bound(spread)
    else:
        assert es <= 1e-6
        assert ex <= 1e-5
