"""C3 scaled down on the oracle (a study, not a test: run by hand).

The C3 ramp (finite strain, two HyperElastoPlastic1 phases with the
consistent tangent, Gaussian-random-field mask with correlation length
16 n / 256, load steps H_k = diag(0, 0, .005 k)) at n^3 voxels: the Newton
loop of newton_cg (src/solver/solvers.cc:141-214) with SolverCG(1e-8), and at
every Newton iterate the extreme Ritz values of the projected operator
A = G:(K:.) on range(G) from 60 Lanczos steps with full
reorthogonalisation, i.e. whether CG meets an indefinite operator there.

    python tests/study_c3_convergence.py 16 32 48
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
from bench import grf_mask  # noqa: E402  (the benchmark's generator)
from oracle import oracle as O  # noqa: E402


def make(n):
    mask = grf_mask(n, 16.0 * n / 256, 2026)
    oc = O.Cell([n] * 3, [1.] * 3, O.FINITE_STRAIN, allow_even=True)
    soft = oc.add_material(O.HEP1_CT, 70e9, .33, 200e6, 1e9)
    hard = oc.add_material(O.HEP1_CT, 210e9, .33, 800e6, 1e9)
    for i, m in enumerate(mask.ravel()):
        oc.add_pixel(hard if m else soft, i)
    oc.initialise()
    return oc


def ritz(oc, k=60, seed=0):
    """Smallest two and largest Ritz values of A on range(G)."""
    rng = np.random.default_rng(seed)
    v = oc.apply_projection(rng.standard_normal((oc.N, 9))).ravel()
    v /= np.linalg.norm(v)
    V, a, b = [v], [], []
    for j in range(k):
        w = oc.directional_stiffness(V[-1].reshape(oc.N, 9)).ravel()
        al = w @ V[-1]
        a.append(al)
        w = w - al * V[-1] - (b[-1] * V[-2] if j > 0 else 0)
        for q in V:
            w -= (w @ q) * q
        be = np.linalg.norm(w)
        if be < 1e-14 * abs(al):
            break
        b.append(be)
        V.append(w / be)
    T = np.diag(a) + np.diag(b[:len(a) - 1], 1) + np.diag(b[:len(a) - 1], -1)
    ev = np.linalg.eigvalsh(T)
    return ev[0], ev[1], ev[-1]


def run(n, steps=10, maxcg=4000):
    oc = make(n)
    F = np.tile(np.eye(3).ravel(order="F"), (oc.N, 1))
    print("n = {}: {} voxels".format(n, oc.N), flush=True)
    for k in range(1, steps + 1):
        F += np.diag([0, 0, .005]).ravel(order="F")
        for it in range(50):
            oc.F[:] = F
            oc.evaluate_stress_tangent()
            rhs = -oc.apply_projection(oc.P.copy())
            nplast = int((oc.get_internal("eps_p") > 0).sum())
            lo, lo2, hi = ritz(oc)
            msg = ("  step {} newton {}: |G:P| {:.3e}, plastic voxels {}, Ritz min {:.3e} {:.3e} "
                   "max {:.3e}".format(k, it, np.linalg.norm(rhs), nplast, lo, lo2, hi))
            try:
                c0 = oc.counter.value
                x = oc.cg_solve(rhs, 1e-8, maxcg)
            except O.OracleError as e:
                print(msg + ", CG FAILED: " + str(e)[:120], flush=True)
                return
            F = F + x
            incr = np.linalg.norm(x) / np.linalg.norm(F)
            print(msg + ", CG {} its, |dF|/|F| {:.3e}".format(oc.counter.value - c0, incr), flush=True)
            if incr <= 1e-5:
                break
        else:
            print("  step {}: no convergence in 50 Newton iterations".format(k), flush=True)
            return
        oc.save_history()
    print("  all {} steps converged".format(steps), flush=True)


if __name__ == "__main__":
    O.set_threads(int(os.environ.get("MO_THREADS", os.cpu_count() or 1)))
    for n in [int(a) for a in sys.argv[1:]] or [16]:
        t = time.time()
        run(n)
        print("  ({:.0f} s)".format(time.time() - t), flush=True)
