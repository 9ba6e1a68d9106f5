"""Writes the full-size solve fixtures of the benchmark workloads from the
oracle (the CPU restatement of the reference's FFTW-double Newton-CG path).

  c3_256_newton1.npz  C3 at 256^3: finite strain, two HyperElastoPlastic1
                      phases (consistent tangent) on the Gaussian-random-field
                      mask of the benchmark, first load step F_zz = 1.005, the
                      first Newton iteration (stress and tangent at F0 = I +
                      H_1, rhs = -G:P, SolverCG to 1e-8, F1 = F0 + dF), then
                      the history cycle and the stress at F1
  c2_128_shear.npz    C2 at 128^3: small strain LE1, 40 periodic spheres, the
                      shear load case eps_12 = eps_21 = 1, newton_cg
  c5_128_newton1.npz  C5's material at 128^3: LE4 with lognormal E and
                      uniform nu per voxel, shear eps_xy = .01, the first
                      Newton iteration as for C3
  c2_125_shear.npz    C2 at 125^3 = 5^3, the odd grid next to the benchmark's
                      (spheres of the benchmark's generator at this size)
  c5_125_newton1.npz  C5's material and load at 125^3 (odd grid)
  c3_243_newton1.npz  C3 at 243^3 = 3^5, the odd grid next to the benchmark's
                      (the reference's projection, no even-grid policy; the
                      runtime-length FFT passes on the GPU), mask of the
                      benchmark's generator at this size, the first Newton
                      iteration as for C3

Each file holds the inputs that are not regenerated bit for bit elsewhere
(the C3 phase mask), the iteration counts, the stopping-test records of
every CG solve, homogenised stresses, field norms, and the strain and
stress of 1024 sampled voxels; the first-Newton-iteration fixtures also
hold the oracle's own rounding spread (the same solve with 3 and 5 threads,
i.e. another summation order of every dot product: the iteration counts
and the largest |dx|/|x|).  Run from the repository root:
    python tests/golden/make_solve_fixtures.py [c3] [c2] [c5] [c3_243] [c2_125] [c5_125]
(needs ~32 GB of memory for c3 and c3_243; OMP threads from MO_THREADS,
default all).
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from bench import grf_mask, spheres_mask  # noqa: E402  (the benchmark's generators)
from oracle import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
NSAMPLE = 1024


def samples(N, seed):
    return np.sort(np.random.default_rng(seed).choice(N, NSAMPLE, replace=False))


def spread_solve(oc, rhs, x_ref, tol, maxiter, threads=(3, 5)):
    """The oracle's own rounding spread of a CG solve: the solve repeated
    with other thread counts (another summation order of every dot product,
    at every iteration); returns (iteration counts, max |x' - x| / |x|)."""
    default = int(O.lib().mo_get_threads())
    counts, spread = [], 0.0
    try:
        for th in threads:
            O.set_threads(th)
            c0 = oc.counter.value
            xp = oc.cg_solve(rhs, tol, maxiter)
            counts.append(oc.counter.value - c0)
            spread = max(spread, float(np.linalg.norm(xp - x_ref) / np.linalg.norm(x_ref)))
    finally:
        O.set_threads(default)
    return counts, spread


def first_newton(name, oc, F0, N, seed, extra):
    """The first Newton iteration of newton_cg (src/solver/solvers.cc:141-
    187): stress and tangent at F0, rhs = -G:P, SolverCG to 1e-8, F1 = F0 +
    dF; then the history cycle that ends the load step (:213) and the stress
    at F1."""
    oc.F[:] = F0
    t = time.time()
    oc.evaluate_stress_tangent()
    P0 = oc.P.copy()
    rhs = -oc.apply_projection(P0)
    O.cg_records()
    c0 = oc.counter.value
    dF = oc.cg_solve(rhs, 1e-8, 4000)
    its = oc.counter.value - c0
    rec = O.cg_records()
    its_p, spread = spread_solve(oc, rhs, dF, 1e-8, 4000)
    print("{}: {} CG iterations (other thread counts: {}, spread {:.2e}), {:.0f} s".format(
        name, its, its_p, spread, time.time() - t), flush=True)
    F1 = oc.F + dF
    oc.save_history()
    oc.F[:] = F1
    oc.evaluate_stress_tangent()
    P1 = oc.P
    idx = samples(N, seed)
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"), cg_iterations=its, cg_records=rec,
        spread_iterations=its_p, spread=spread, rhs_norm=np.linalg.norm(rhs),
        dF_norm=np.linalg.norm(dF), P0_norm=np.linalg.norm(P0), P0_mean=P0.mean(0),
        P1_mean=P1.mean(0), P1_norm=np.linalg.norm(P1), sample_index=idx, dF_sample=dF[idx],
        P0_sample=P0[idx], P1_sample=P1[idx], **extra)


def c3(n=256):
    res = [n] * 3
    N = n ** 3
    # the benchmark's mask (correlation length 16 voxels at 256^3)
    mask = grf_mask(n, 16.0 * n / 256.0, 2026)
    oc = O.Cell(res, [1.0] * 3, O.FINITE_STRAIN, allow_even=n % 2 == 0)
    soft = oc.add_material(O.HEP1_CT, 70e9, .33, 200e6, 1e9)
    hard = oc.add_material(O.HEP1_CT, 210e9, .33, 800e6, 1e9)
    flat = mask.ravel()
    for i in range(N):
        oc.add_pixel(hard if flat[i] else soft, i)
    oc.initialise()
    H1 = np.diag([0.0, 0.0, 0.005])
    F0 = (np.eye(3) + H1).ravel(order="F")
    first_newton("c3_{}_newton1".format(n), oc, F0, N, 1,
                 {"mask_bits": np.packbits(flat), "res": np.array(res), "load": H1})


def c2(n=128):
    mask = spheres_mask(n, 40, 2026).ravel()
    res = [n] * 3
    N = n ** 3
    oc = O.Cell(res, [1.0] * 3, O.SMALL_STRAIN, allow_even=n % 2 == 0)
    inc = oc.add_material(O.LE1, 100e9, .3)
    mat = oc.add_material(O.LE1, 10e9, .3)
    for i in range(N):
        oc.add_pixel(inc if mask[i] else mat, i)
    e = np.zeros((3, 3))
    e[0, 1] = e[1, 0] = 1.0
    O.cg_records()
    t = time.time()
    r = oc.newton_cg([e], 1e-8, 4000, 1e-5, 1e-10)[0]
    rec = O.cg_records()
    name = "c2_{}_shear".format(n)
    print("{}: {} Newton, {} CG iterations, {:.0f} s".format(name, r["nb_it"], r["nb_fev"],
                                                             time.time() - t), flush=True)
    idx = samples(N, 2)
    g, s = r["grad"], r["stress"]
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"), res=np.array(res), load=e, nb_it=r["nb_it"],
        nb_fev=r["nb_fev"], success=r["success"], cg_records=rec, grad_norm=np.linalg.norm(g),
        stress_norm=np.linalg.norm(s), grad_mean=g.mean(0), stress_mean=s.mean(0),
        sample_index=idx, grad_sample=g[idx], stress_sample=s[idx],
        inclusion_fraction=mask.mean())


def c5(n=128):
    """C5's material and load at n^3, first Newton iteration (on a linear
    problem the second Newton iteration asks CG to reduce a residual that
    is already at the operator's rounding level by a further 1e-8)."""
    N = n ** 3
    # the benchmark's generator at 128^3 (bench.py Workload c5)
    rng = np.random.default_rng(2026)
    E = np.exp(np.log(50e9) + 0.8 * rng.standard_normal(N))
    nu = rng.uniform(0.15, 0.45, N)
    oc = O.Cell([n] * 3, [1.0] * 3, O.SMALL_STRAIN, allow_even=n % 2 == 0)
    m = oc.add_material(O.LE4)
    for i in range(N):
        oc.add_pixel(m, i, [E[i], nu[i]])
    oc.initialise()
    e = np.zeros((3, 3))
    e[0, 1] = e[1, 0] = 0.01
    first_newton("c5_{}_newton1".format(n), oc, np.tile(e.ravel(order="F"), (N, 1)), N, 3,
                 {"res": np.array([n] * 3), "load": e})


if __name__ == "__main__":
    O.set_threads(int(os.environ.get("MO_THREADS", os.cpu_count() or 1)))
    which = sys.argv[1:] or ["c2", "c5", "c3"]
    for w in which:
        {"c3": c3, "c2": c2, "c5": c5, "c3_243": lambda: c3(243), "c2_125": lambda: c2(125),
         "c5_125": lambda: c5(125)}[w]()
