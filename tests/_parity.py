"""Parity bounds of the GPU against the oracle that respect the reference
algorithm's own rounding floor.

A CG solve to a relative residual of 1e-8 fixes its solution only up to the
rounding of the arithmetic that produced it: on a badly conditioned cell
(C5's per-voxel lognormal moduli, sigma .8, contrast ~10^3) the oracle
itself, run on the same data with another thread count (another summation
order of every dot product), returns solutions that differ by ~2e-9 and
stops one or two iterations earlier or later (16x16x128: 112 or 113
iterations, r.r / tol2 one iteration before the stop between 1.05 and 1.64).
On the other benchmark materials it reproduces itself to ~1e-14 with
identical counts.  A faithful implementation cannot be held closer to the
oracle than the oracle is to itself.  So every comparison measures that
spread on the oracle first:

  * iteration counts: the GPU's count must equal the oracle's when the
    oracle's own count does not move under the perturbations, and lie within
    the range of the oracle's counts, widened by its width, otherwise;
  * fields: |x_gpu - x_oracle| <= max(1e-10, 10 * spread) * |x_oracle|, so
    the 1e-10 bound applies wherever the oracle reproduces itself to better
    than 1e-11, and a bound ten times the oracle's own spread elsewhere.
"""
import numpy as np

FIELD_TOL = 1e-10
SPREAD_FACTOR = 10.0


def perturbed(a, seed):
    """a with relative noise of 1e-16 (another rounding of the same data)."""
    return a * (1.0 + 1e-16 * np.random.default_rng(1000 + seed).standard_normal(a.shape))


def cg_with_spread(oracle, oc, rhs, tol, maxiter, threads=(1, 2, 3)):
    """The oracle's CG solve of rhs, and the same solve with other thread
    counts (another summation order of every dot product, at every
    iteration).  Returns (x, iterations, all iteration counts, spread)."""
    default = int(oracle.lib().mo_get_threads())
    c0 = oc.counter.value
    x = oc.cg_solve(rhs, tol, maxiter)
    its = oc.counter.value - c0
    counts = [its]
    spread = 0.0
    try:
        for th in threads:
            oracle.set_threads(th)
            c0 = oc.counter.value
            xp = oc.cg_solve(rhs, tol, maxiter)
            counts.append(oc.counter.value - c0)
            spread = max(spread, float(np.linalg.norm(xp - x) / np.linalg.norm(x)))
    finally:
        oracle.set_threads(default)
    return x, its, counts, spread


def newton_with_spread(oracle, make_cell, load_steps, cg_tol, maxiter, newton_tol, equil_tol,
                       threads=(1, 3)):
    """The oracle's newton_cg on a fresh cell from make_cell(), and again on
    fresh cells with other thread counts (another summation order of the
    dot products).  Returns (results of the first run, all CG counts of the
    last step, spread of grad and of stress over the runs, the stopping
    records of the first run's CG solves)."""
    default = int(oracle.lib().mo_get_threads())
    runs = []
    records = None
    try:
        for th in (default,) + tuple(threads):
            oracle.set_threads(th)
            oracle.cg_records()
            runs.append(make_cell().newton_cg(load_steps, cg_tol, maxiter, newton_tol, equil_tol))
            if records is None:
                records = oracle.cg_records()
    finally:
        oracle.set_threads(default)
    ref = runs[0]
    counts = [int(r[-1]["nb_fev"]) for r in runs]
    sg = ss = 0.0
    for r in runs[1:]:
        for a, b in zip(r, ref):
            sg = max(sg, float(np.linalg.norm(a["grad"] - b["grad"]) / np.linalg.norm(b["grad"])))
            ss = max(ss, float(np.linalg.norm(a["stress"] - b["stress"]) / np.linalg.norm(b["stress"])))
    return ref, counts, sg, ss, records


def bound(spread):
    return max(FIELD_TOL, SPREAD_FACTOR * spread)


def rounding_level_solves(records, rel=1e-6):
    """CG solves of a Newton loop whose right-hand side is the residual that
    the previous solve left at its tolerance (|rhs| below rel times the first
    solve's): that rhs, -G:P at the previous iterate, is known only to the
    rounding of G:P at full scale, ~1e-16 of |P| and so ~1e-8 of itself, and
    its CG count may differ by one between two faithful implementations (the
    fields are not affected: they match to ~1e-15).  records: the oracle's
    stopping records (tol2 = tol^2 |rhs|^2)."""
    if len(records) == 0:
        return 0
    first = records[0][3]
    return int(sum(1 for q in records[1:] if q[3] < rel * rel * first))


def check_count(it, counts, msg="", allowance=0):
    """Equal to the oracle's count when the oracle's runs agree on it;
    otherwise within their range widened by its own width (the runs are a
    small sample of the rounding variation); `allowance` more for solves
    known to be rounding-dominated (rounding_level_solves, the C5 material)."""
    lo, hi = min(counts), max(counts)
    w = hi - lo + allowance
    assert lo - w <= it <= hi + w, "GPU {} iterations, oracle {} under its own rounding {}".format(
        it, counts, msg)


# C5's material (per-voxel lognormal moduli, sigma .8, contrast ~10^3): the
# last CG iterations are dominated by rounding.  Measured on the oracle at
# the benchmark's axis lengths with 1 to 8 threads: counts 150-153 and
# solutions that differ by up to 3.7e-9 (256x16x16, 16x256x16), 112-113 and
# 1.9e-9 (16x16x128); a few sampled runs can miss that range, so a solve on
# this material is held to its count +-2 and to 1e-8 in the field
CHAOTIC_COUNT = 2
CHAOTIC_FIELD_TOL = 1e-8


def chaotic_bound(spread):
    return max(CHAOTIC_FIELD_TOL, SPREAD_FACTOR * spread)
