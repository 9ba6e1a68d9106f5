"""Every variant of every compile-time-length FFT pass, each forced in turn
with mugpu_debug_force_variant.  The autotuner picks among them by timing
alone, so any of them may end up in the tuned table; each must compute what
the oracle computes:

  * the passes of a CG iteration (last-axis r2c with the p and x updates and
    K:p, middle-axis c2c, first-axis c2c with the projection, last-axis c2r
    with the p.Ap partials) through newton_cg on two-phase LE1 cells whose
    axes take the compiled lengths 16 .. 512 in each role (first, middle,
    last), small and finite strain, against the oracle's newton_cg: in FP64
    the Newton count and, as every full-size parity test here holds them
    (tests/_parity.py), the CG count and fields at the oracle's own rounding
    floor (fields to 1e-10, the CG count equal to the oracle's or, for a
    solve whose right-hand side is a residual left at the tolerance, within
    one); FP32-refined,
    the homogenised stress to 1e-6 and the strain field to 1e-5 of the
    oracle, with at most three FP64 residuals per solve (a faulty FP32
    operator shows as more refinements before it shows in the result);
  * the engine transforms (r2c, c2r, the plain middle- and first-axis
    passes) against numpy rfftn / irfftn.
"""
import ctypes

import numpy as np
import pytest

from _parity import bound, check_count, newton_with_spread, rounding_level_solves
from muspectre_amd._device import DeviceBuffer

pytestmark = pytest.mark.gpu

# axis lengths in each role: first (projection pass), middle, last (r2c, c2r)
GRIDS = [[64, 32, 32], [256, 16, 128], [512, 16, 16], [16, 256, 16], [16, 16, 512]]
LENGTHS = [1.0, 1.3, 0.7]
# MUGPU_PASS_* and the number of variant codes of each
PASSES = {"r2c": (0, 4), "middle": (1, 2), "axis0": (2, 8), "c2r": (3, 4), "axis0_plain": (4, 2)}
FP64, FP32 = 0, 1
_REF = {}


@pytest.fixture(scope="module")
def µ():
    import muspectre_amd
    return muspectre_amd


def _exists(capi, res, prec, pas, v):
    """Whether variant v of the pass exists at these lengths (the autotuner's
    measurement entry point refuses the others with MUGPU_EARG)."""
    lib = capi.load()
    ms = ctypes.c_double()
    r = (ctypes.c_int * 3)(*res)
    rc = lib.mugpu_fft_tune_pass(3, r, prec, pas, v, 0, 1, ctypes.byref(ms))
    if rc == capi.EARG:
        return False
    capi.check(rc)
    return True


class _Forced:
    def __init__(self, capi, pas, prec, v):
        self.lib, self.capi, self.args = capi.load(), capi, (pas, prec, v)

    def __enter__(self):
        self.capi.check(self.lib.mugpu_debug_force_variant(*self.args))

    def __exit__(self, *exc):
        self.capi.check(self.lib.mugpu_debug_force_variant(self.args[0], self.args[1], -1))


def _mask(res):
    """two ellipsoidal inclusions, their axes a fixed fraction of the grid"""
    g = np.meshgrid(*[np.arange(n) / n for n in res], indexing="ij")
    return (((g[0] - .3) ** 2 + (g[1] - .3) ** 2 + (g[2] - .4) ** 2) < .05) | \
           (((g[0] - .75) ** 2 + (g[1] - .7) ** 2 + (g[2] - .75) ** 2) < .02)


STEPS = [np.array([[.01, .002, 0], [.002, 0, 0], [0, 0, -.005]])]


def _oracle(oracle, res, form):
    """The oracle's result, the CG counts and field spreads of its runs with
    other thread counts, and how many of its solves are rounding-level
    (tests/_parity.py)."""
    key = (tuple(res), form)
    if key not in _REF:
        def make():
            oc = oracle.Cell(res, LENGTHS, oracle.SMALL_STRAIN if form == "small" else oracle.FINITE_STRAIN,
                             allow_even=True)
            a = oc.add_material(oracle.LE1, 100e9, .3)
            b = oc.add_material(oracle.LE1, 10e9, .3)
            mask = _mask(res)
            for i in range(mask.size):
                oc.add_pixel(a if mask.flat[i] else b, i)
            return oc
        ref, counts, sg, ss, records = newton_with_spread(oracle, make, STEPS, 1e-8, 400, 1e-5, 1e-10)
        _REF[key] = (ref[0], counts, sg, ss, rounding_level_solves(records))
    return _REF[key]


def _gpu(µ, res, form, prec):
    F_ = µ.Formulation.small_strain if form == "small" else µ.Formulation.finite_strain
    cell = µ.Cell(res, LENGTHS, F_, precision="fp64" if prec == FP64 else "fp32", allow_even=True)
    h = µ.material.MaterialLinearElastic1_3d.make(cell, "hard", 100e9, .3)
    s = µ.material.MaterialLinearElastic1_3d.make(cell, "soft", 10e9, .3)
    idx = np.arange(int(np.prod(res)))
    mask = _mask(res)
    h.add_pixels_by_index(idx[mask.ravel()])
    s.add_pixels_by_index(idx[~mask.ravel()])
    solver = µ.solvers.SolverCG(cell, 1e-8, 400, 0)
    r = µ.solvers.newton_cg(cell, STEPS, solver, 1e-5, 1e-10, 0)[0]
    return r, solver


def _rel(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("form", ["small", "finite"])
@pytest.mark.parametrize("name", ["r2c", "middle", "axis0", "c2r"])
@pytest.mark.parametrize("res", GRIDS, ids=lambda r: "x".join(map(str, r)))
def test_cg_pass_variants_vs_oracle(µ, oracle, capi, res, name, form):
    pas, nv = PASSES[name]
    o, counts, sg, ss, allowance = _oracle(oracle, res, form)
    ran = []
    for prec in (FP64, FP32):
        for v in range(nv):
            if not _exists(capi, res, prec, pas, v):
                continue
            with _Forced(capi, pas, prec, v):
                r, solver = _gpu(µ, res, form, prec)
            case = "%s variant %d %s" % (name, v, "fp64" if prec == FP64 else "fp32")
            assert r.success, case
            assert r.nb_it == o["nb_it"], case
            if prec == FP64:
                check_count(int(r.nb_fev), counts, case, allowance=allowance)
                assert _rel(r.grad, o["grad"]) <= bound(sg), case
                assert _rel(r.stress, o["stress"]) <= bound(ss), case
            else:
                mean = lambda P: np.asarray(P).reshape(-1, 9).mean(0)
                assert _rel(mean(r.stress), mean(o["stress"])) <= 1e-6, case
                assert _rel(r.grad, o["grad"]) <= 1e-5, case
                assert solver.get_refinements() <= 3 * r.nb_it, case
            ran.append(case)
    assert any("fp64" in c for c in ran) and any("fp32" in c for c in ran), ran


def test_forced_variants_run(µ, capi):
    """The hook takes effect: the c2r variants change the threads of a
    workgroup and so the order in which each sums its p.Ap partial, and the
    FP64 fields of a solve differ in their last bits between them."""
    res = GRIDS[0]
    fields = set()
    for v in range(PASSES["c2r"][1]):
        if _exists(capi, res, FP64, PASSES["c2r"][0], v):
            with _Forced(capi, PASSES["c2r"][0], FP64, v):
                r, _ = _gpu(µ, res, "small", FP64)
            fields.add(np.asarray(r.grad).tobytes())
    assert len(fields) >= 2


@pytest.mark.parametrize("name", ["r2c", "middle", "axis0_plain", "c2r"])
@pytest.mark.parametrize("res", GRIDS, ids=lambda r: "x".join(map(str, r)))
def test_engine_pass_variants_vs_numpy(capi, res, name):
    pas, nv = PASSES[name]
    lib = capi.load()
    nc = 9
    rng = np.random.default_rng(pas)
    x = rng.random([*res, 3, 3])
    herm = res[:-1] + [res[-1] // 2 + 1]
    kp = int(np.prod(herm))
    c = rng.random([*herm, 3, 3]) + 1j * rng.random([*herm, 3, 3])
    ref = np.fft.rfftn(x, axes=(0, 1, 2)).reshape(kp, nc)
    ref2 = np.fft.irfftn(c, s=res, axes=(0, 1, 2)).reshape(-1, nc)
    ran = 0
    for prec in (FP64, FP32):
        dt = np.float64 if prec == FP64 else np.float32
        ct = np.complex128 if prec == FP64 else np.complex64
        for v in range(nv):
            if not _exists(capi, res, prec, pas, v):
                continue
            with _Forced(capi, pas, prec, v):
                h = ctypes.c_void_p()
                r = (ctypes.c_int * 3)(*res)
                capi.check(lib.mugpu_fft_create(3, r, nc, prec, ctypes.byref(h)))
            try:
                din = DeviceBuffer.from_array(x.reshape(-1, nc).astype(dt))
                dspec = DeviceBuffer(kp * nc * np.dtype(ct).itemsize)
                capi.check(lib.mugpu_fft_r2c(h, din.ptr, dspec.ptr))
                spec = dspec.download((kp, nc), ct)
                dspec.upload(c.reshape(kp, nc).astype(ct))
                dout = DeviceBuffer(int(np.prod(res)) * nc * np.dtype(dt).itemsize)
                capi.check(lib.mugpu_fft_c2r(h, dspec.ptr, dout.ptr))
                out = dout.download((int(np.prod(res)), nc), dt) * lib.mugpu_fft_normalisation(h)
            finally:
                lib.mugpu_fft_destroy(h)
            case = "%s variant %d %s" % (name, v, "fp64" if prec == FP64 else "fp32")
            if prec == FP64:
                assert np.linalg.norm(spec - ref) < 1e-14 * np.prod(res), case
                assert np.linalg.norm(out - ref2) < 1e-14 * np.prod(res), case
            else:
                assert np.abs(spec - ref).max() < 2e-6 * np.abs(ref).max(), case
                assert np.abs(out - ref2).max() < 2e-6 * np.abs(ref2).max(), case
            ran += 1
    assert ran >= 2
