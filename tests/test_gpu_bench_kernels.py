"""GPU parity of the kernel instantiations the benchmarks run, against the
oracle, called through the C-ABI.

The compile-time-length passes are instantiated per axis length, per
precision and per tangent representation (k_r2c_fast with the p update, the
x update of the previous iteration and K:p; k_strided_fast; k_axis0_fast with
the projection; k_c2r_fast with the p.Ap partials), and the tuned variant
differs from length to length.  The benchmarks run them at 128 (C2), 256 (C3)
and 512 (C5) on every axis; here each of those lengths sits on each axis in
turn of a 16 x 16 x 16 grid (the other axes at 16), with the tangent of each
benchmark:
  iso  small strain, MaterialLinearElastic4 with lognormal E and uniform nu
       per voxel (C5; C2's two-phase LE1 has the same (lambda, mu) tangent)
  sym  finite strain, HyperElastoPlastic1 with the consistent tangent (C3;
       the 45-entry upper triangle)
  hep  finite strain, HyperElastoPlastic1 as coded (the 27-entry
       j-independent tangent; CG diverges on it, so a fixed number of
       iterations is compared)

Tolerances: stress and tangent 1e-12, G:(K:dF) 1e-12 (max norm); SolverCG
(mugpu_cg_steps, iterations queued eight at a time with the x update carried
by the next r2c, as newton_cg runs them): the oracle's iteration count and x
to 1e-10 (2-norm), both relaxed only to the oracle's own rounding spread
where the oracle does not reproduce itself that closely (tests/_parity.py:
on C5's lognormal material, contrast ~10^3, the oracle's own count moves by
one or two and its solution by up to 3.7e-9 with the thread count, so there
the count is held to +-2 and x to 1e-8); the
mixed-precision solve (FP32 operator, FP64 refinement): x to 1e-5 and the
homogenised stress at F0 + x to 1e-6 of the oracle's.

Reference: CellBase::evaluate_stress_tangent and
evaluate_projected_directional_stiffness (src/cell/cell_base.cc:129-192),
SolverCG::solve (src/solver/solver_cg.cc:54-110).
"""
import ctypes

import numpy as np
import pytest

from _parity import (CHAOTIC_COUNT, bound, cg_with_spread, chaotic_bound, check_count,
                     perturbed)
from muspectre_amd._device import DeviceBuffer
from test_gpu_abi_cell import GpuCell, _arr

pytestmark = pytest.mark.gpu

E_SOFT, E_HARD, NU, TY_SOFT, TY_HARD, H = 70e9, 210e9, .33, 200e6, 800e6, 1e9


def _grids():
    out = []
    for n in (128, 256, 512):
        for ax in range(3):
            res = [16, 16, 16]
            res[ax] = n
            out.append(res)
    return out


GRIDS = _grids()
KINDS = ["iso", "sym", "hep"]


def _gid(res):
    return "x".join(map(str, res))


def _build(capi, oracle, res, kind, prec=0):
    """Oracle cell and GPU cell of one tangent kind, materials added, both
    initialised; returns (oc, gc, F0)."""
    N = int(np.prod(res))
    rng = np.random.default_rng(sum(res) + KINDS.index(kind))
    if kind == "iso":
        form = 1
        oc = oracle.Cell(res, [1.0] * 3, oracle.SMALL_STRAIN, allow_even=True)
        gc = GpuCell(capi, res, [1.0] * 3, form, prec=prec, allow_even=1)
        E = np.exp(np.log(50e9) + 0.8 * rng.standard_normal(N))
        nu = rng.uniform(0.15, 0.45, N)
        mo = oc.add_material(oracle.LE4)
        for i in range(N):
            oc.add_pixel(mo, i, [E[i], nu[i]])
        mg = gc.add_material(4)
        gc.add_pixels(mg, np.arange(N), np.stack([E, nu], axis=1))
        eps = np.zeros((3, 3))
        eps[0, 1] = eps[1, 0] = 0.01
        F0 = np.tile(eps.ravel(order="F"), (N, 1))
    else:
        form = 0
        fam = oracle.HEP1_CT if kind == "sym" else oracle.HEP1
        oc = oracle.Cell(res, [1.0] * 3, oracle.FINITE_STRAIN, allow_even=True)
        gc = GpuCell(capi, res, [1.0] * 3, form, prec=prec, allow_even=1)
        hard = rng.random(N) < 0.5
        a = oc.add_material(fam, E_SOFT, NU, TY_SOFT, H)
        b = oc.add_material(fam, E_HARD, NU, TY_HARD, H)
        for i in range(N):
            oc.add_pixel(b if hard[i] else a, i)
        ga = gc.add_material(fam, E_SOFT, NU, TY_SOFT, H)
        gb = gc.add_material(fam, E_HARD, NU, TY_HARD, H)
        idx = np.arange(N)
        gc.add_pixels(gb, idx[hard])
        gc.add_pixels(ga, idx[~hard])
        # the C3 load of the first step plus a small random field, so that
        # both phases and yielded voxels occur
        F0 = np.eye(3).ravel(order="F") + np.diag([0, 0, 0.005]).ravel(order="F") \
            + 0.002 * rng.standard_normal((N, 9))
    oc.initialise()
    capi.check(gc.lib.mugpu_cell_initialise(gc.h))
    return oc, gc, F0


def _evaluate(capi, oc, gc, F0):
    oc.F[:] = F0
    gc.set_F(F0)
    oc.evaluate_stress_tangent()
    capi.check(gc.lib.mugpu_cell_evaluate_stress_tangent(gc.h))


def _cg_chunked(gc, rhs, tol, maxiter, chunk=8):
    """SolverCG's loop as newton_cg runs it: mugpu_cg_steps, `chunk`
    iterations per call with the loop test on the device."""
    from test_gpu_abi_cell import _gpu_cg_chunked
    x, r, it, rd, calls = _gpu_cg_chunked(gc, rhs, tol, maxiter, chunk)
    return x, it


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("res", GRIDS, ids=_gid)
def test_operator_and_cg_fp64(capi, oracle, res, kind):
    oc, gc, F0 = _build(capi, oracle, res, kind)
    _evaluate(capi, oc, gc, F0)
    N, nc = gc.N, 9
    P = gc.get("P")
    assert np.abs(P - oc.P).max() <= 1e-12 * np.abs(oc.P).max()
    K = np.zeros((N, nc * nc))
    capi.check(gc.lib.mugpu_cell_get_tangent(gc.h, _arr(K)))
    assert np.abs(K - oc.K).max() <= 1e-12 * np.abs(oc.K).max()
    rng = np.random.default_rng(7)
    dF = rng.standard_normal((N, nc))
    ref = oc.directional_stiffness(dF)
    bi = DeviceBuffer.from_array(dF)
    bo = DeviceBuffer(dF.nbytes)
    capi.check(gc.lib.mugpu_cell_directional_stiffness(gc.h, bi.ptr, bo.ptr))
    assert np.abs(bo.download(dF.shape) - ref).max() <= 1e-12 * np.abs(ref).max()
    rhs = -oc.apply_projection(oc.P.copy())
    if kind == "hep":
        # the coded tangent is not symmetric: CG does not converge on it, so
        # a fixed number of iterations (r.r never below a zero threshold) is
        # compared
        iters = 6
        x_ref = oc.cg_fixed_iterations(rhs, iters)
        spread = max(np.linalg.norm(oc.cg_fixed_iterations(perturbed(rhs, s), iters) - x_ref)
                     for s in range(2)) / np.linalg.norm(x_ref)
        x, it = _cg_chunked(gc, rhs, 0.0, iters, chunk=4)
        err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
        print(_gid(res), kind, "{} CG iterations: x {:.2e} (oracle spread {:.2e})".format(
            iters, err, spread), flush=True)
        assert it == iters
        assert err <= bound(spread)
        return
    oracle.cg_records()
    x_ref, it_ref, counts, spread = cg_with_spread(oracle, oc, rhs, 1e-8, 2000)
    rec = oracle.cg_records()
    x, it = _cg_chunked(gc, rhs, 1e-8, 2000)
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    margin = "oracle stop: r.r/tol2 = {:.6f} at {} iterations, {:.6f} one before".format(
        rec[0][1] / rec[0][3], int(rec[0][0]), rec[0][2] / rec[0][3])
    print(_gid(res), kind, "CG", it, "oracle", counts, "x {:.2e} (oracle spread {:.2e})".format(
        err, spread), margin, flush=True)
    if kind == "iso":
        check_count(it, counts, margin, allowance=CHAOTIC_COUNT)
        assert err <= chaotic_bound(spread)
    else:
        check_count(it, counts, margin)
        assert err <= bound(spread)


def _homog_stress_at(oracle, oc, F0, x):
    oc.F[:] = F0 + x
    oc.evaluate_stress_tangent()
    return oc.P.mean(0)


@pytest.mark.parametrize("kind", ["iso", "sym"])
@pytest.mark.parametrize("res", GRIDS, ids=_gid)
def test_mixed_precision_solve(capi, oracle, res, kind):
    """The FP32 instantiations of the same passes: the refined solve on the
    oracle's right-hand side."""
    oc, gc, F0 = _build(capi, oracle, res, kind, prec=1)
    _evaluate(capi, oc, gc, F0)
    rhs = -oc.apply_projection(oc.P.copy())
    x_ref = oc.cg_solve(rhs, 1e-8, 2000)
    b = DeviceBuffer.from_array(rhs)
    xb = DeviceBuffer(rhs.nbytes)
    its, refs, rel = ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
    capi.check(gc.lib.mugpu_cg_solve_mixed(gc.h, b.ptr, xb.ptr, 1e-8, 4000, 1e-4, ctypes.byref(its),
                                           ctypes.byref(refs), ctypes.byref(rel)))
    x = xb.download(rhs.shape)
    ex = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    s_ref = _homog_stress_at(oracle, oc, F0, x_ref)
    s = _homog_stress_at(oracle, oc, F0, x)
    es = np.linalg.norm(s - s_ref) / np.linalg.norm(s_ref)
    print(_gid(res), kind, "mixed: {} inner, {} refinements, field {:.2e}, stress {:.2e}".format(
        its.value, refs.value, ex, es), flush=True)
    assert rel.value <= 1e-8
    assert ex <= 1e-5
    assert es <= 1e-6
