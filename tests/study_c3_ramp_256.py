"""The C3 load ramp at full size on the GPU (a study, not a test: run by
hand on a GPU host): newton_cg over the ten load steps with verbose Newton
output, reporting where it stops (the ConvergenceError of SolverCG) and how
many voxels have yielded.

    python tests/study_c3_ramp_256.py [n]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
import muspectre_amd as mu  # noqa: E402


class _Args:
    workload = "c3"
    precision = "fp64"
    size = int(sys.argv[1]) if len(sys.argv) > 1 else 0


wl = bench.Workload(_Args())
cell = wl.make_cell(mu)
solver = mu.solvers.SolverCG(cell, 1e-8, 4000, False)
loads = [np.diag([0, 0, 0.005 * k]) for k in range(1, 11)]
t = time.time()
try:
    for i, r in enumerate(mu.solvers.newton_cg(cell, loads, solver, 1e-5, 1e-10, 1)):
        print("step", i + 1, r.success, r.nb_it, r.nb_fev, flush=True)
except Exception as e:  # the reference's ConvergenceError
    print("EXCEPTION after %.1f s: %s" % (time.time() - t, str(e)[:300]), flush=True)
eps = np.asarray(cell.get_globalised_internal_real_array("eps_p")).ravel()
print("plastic voxels (current eps_p > 0):", int((eps > 0).sum()), "of", eps.size, flush=True)
