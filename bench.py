#!/usr/bin/env python3
"""Benchmark of the Newton-CG solve on one MI355X.

Metric: voxel-CG-iters/s = N_voxels x (CG iterations counted by SolverCG)
/ wall time of the newton_cg call (cell construction and initialise are
outside the timed region).  One step = one complete newton_cg solve of the
workload from a freshly initialised cell.

After the warmup steps, one untimed step runs with HIP events around every
kernel launch and gives the per-kernel breakdown ("kernels") and the
dominant kernel; the timed steps then put events around the launches of
that kernel only ("roofline", measured live over the timed steps), so that
the markers between the other kernels do not slow the measured solve.

Workloads (synthetic inputs generated from fixed seeds):
  c3  256^3 finite strain, two HyperElastoPlastic1 phases from a Gaussian
      random field (correlation length 16 voxels) thresholded at its median;
      soft E=70 GPa nu=.33 tau_y0=200 MPa H=1 GPa, hard E=210 GPa
      tau_y0=800 MPa; first load step of the uniaxial ramp, F_zz = 1.005;
      consistent tangent (the coded HEP1 tangent makes CG diverge);
      the first three Newton iterations (return mapping, residual projection
      and the CG solve to tol 1e-8, three times): no load step of the ramp
      converges at 256^3, and these are the iterations before Newton's
      increment starts to grow
  c2  128^3 small strain, 40 periodic spheres (r ~ U[4,10]) of E=100 GPa in
      E=10 GPa, nu=.3, six unit load cases, CG tol 1e-8
  c5  512^3 small strain LE4, E lognormal (median 50 GPa, sigma_log .8),
      nu ~ U[.15,.45], shear eps_xy = .01, CG tol 1e-8
  c1  51x51 small strain, disc of radius 10 (bin/small_case.py)
  c4  serving stream: RVE requests, 3D with p = .7 (32^3, 48^3, 64^3, 96^3)
      else 2D (64^2 .. 512^2), thresholded Gaussian-random-field phases
      seeded per request, stiffness contrast log-uniform in [1, 100], nu in
      [.2, .4], one symmetric strain with entries in +-.01, small strain,
      FP32 + refined CG to 1e-6.  Phase 1 queues every request at t = 0
      (saturated throughput, RVE solves/s); phase 2 replays a fresh stream
      with Poisson arrivals at --load times that throughput (p50/p99
      latency from each request's arrival time).  Each phase's per-voxel
      E and nu arrays are formed in host memory before its timed region.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=["c3", "c2", "c5", "c1", "c4"])
    ap.add_argument("--requests", type=int, default=1000,
                    help="c4: requests per phase (the full stream is 10000)")
    ap.add_argument("--load", type=float, default=0.7,
                    help="c4: Poisson arrival rate as a fraction of the measured capacity")
    ap.add_argument("--size", type=int, default=0, help="override the grid size")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--kernel", default="auto",
                    choices=["auto", "r2c", "middle", "axis0", "c2r", "material", "update"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="c4: launch the batched iterations directly instead of replaying "
                         "captured graphs (the same kernels; for kernel tracers)")
    return ap.parse_args()


# ---------------------------------------------------------------- workloads
def grf_mask(n, ell, seed):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n)
    f = np.fft.rfftn(w)
    kz = np.fft.rfftfreq(n)
    K2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
    f *= np.exp(-0.5 * (2 * np.pi) ** 2 * K2 * ell ** 2)
    g = np.fft.irfftn(f, s=(n, n, n), axes=(0, 1, 2))
    return g > np.median(g)


def spheres_mask(n, count, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(n)
    mask = np.zeros((n, n, n), dtype=bool)
    for _ in range(count):
        c = rng.uniform(0, n, 3)
        r = rng.uniform(4, 10)
        d = [np.minimum(np.abs(x - ci), n - np.abs(x - ci)) for ci in c]
        mask |= (d[0][:, None, None] ** 2 + d[1][None, :, None] ** 2 +
                 d[2][None, None, :] ** 2) < r * r
    return mask


class Workload:
    def __init__(self, args):
        self.name = args.workload
        self.prec = args.precision
        w = self.name
        if w == "c3":
            self.n = args.size or 256
            self.res = [self.n] * 3
            self.mask = grf_mask(self.n, 16.0 * self.n / 256.0, 2026)
            self.form = "finite"
            self.loads = [[np.diag([0.0, 0.0, 0.005])]]
            self.newton_tol, self.cg_tol, self.maxiter = 1e-5, 1e-8, 4000
            self.equil_tol = 1e-10
            # No load step of this ramp converges at 256^3: in load step 1 the
            # Newton increment |dF|/|DF| falls for three iterations (2.2e-3,
            # 1.1e-3, 5.0e-4), then grows (7.2e-4, 2.5e-3) as voxels keep
            # crossing the yield surface (the return mapping is not smooth
            # there and newton_cg has no line search), and the CG of the sixth
            # iteration fails (|r|/|b| = .27 after 5582 iterations) on a
            # tangent that the divergence has made indefinite.  The timed
            # prefix is the three Newton iterations in which the increment
            # falls: three material evaluations and three CG solves.
            self.newton_maxiter = 3
            self.prefix_note = ("load step 1, the first 3 Newton iterations (the increment "
                                "falls 2.2e-3, 1.1e-3, 5.0e-4; from the 4th it grows, the 6th "
                                "CG fails): no load step of the ramp converges at 256^3")
            # 9*(6*8 + 9*8) + 360: the tangent is major-symmetric, stored as
            # its 45-entry upper triangle
            self.T = 360
        elif w == "c2":
            self.n = args.size or 128
            self.res = [self.n] * 3
            self.mask = spheres_mask(self.n, 40, 2026)
            # 40 spheres of radius U[4, 10] fill ~3 % of 128^3 (the configuration's
            # "~30 %" does not follow from its own numbers); the actual fraction
            # is reported
            self.inclusion_fraction = float(self.mask.mean())
            self.form = "small"
            eye = np.eye(3)
            cases = []
            for (i, j) in [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]:
                e = np.zeros((3, 3))
                e[i, j] = e[j, i] = 1.0
                cases.append([e])
            self.loads = cases
            self.newton_tol, self.cg_tol, self.maxiter = 1e-5, 1e-8, 4000
            self.equil_tol = 1e-10
            self.T = 16  # (lambda, mu) per voxel
        elif w == "c5":
            self.n = args.size or 512
            self.res = [self.n] * 3
            rng = np.random.default_rng(2026)
            N = self.n ** 3
            self.E = np.exp(np.log(50e9) + 0.8 * rng.standard_normal(N))
            self.nu = rng.uniform(0.15, 0.45, N)
            self.form = "small"
            e = np.zeros((3, 3))
            e[0, 1] = e[1, 0] = 0.01
            self.loads = [[e]]
            self.newton_tol, self.cg_tol, self.maxiter = 1e-5, 1e-8, 4000
            self.equil_tol = 1e-10
            self.T = 16
        else:
            self.n = args.size or 51
            self.res = [self.n, self.n]
            x = np.arange(self.n)
            c = self.n // 2
            self.mask = np.hypot(x[:, None] - c, x[None, :] - c) < 10 * self.n / 51
            self.form = "small"
            self.loads = [[np.array([[0, 0], [0, .03]])]]
            self.newton_tol, self.cg_tol, self.maxiter = 1e-5, 1e-8, 401
            self.equil_tol = 2.0
            self.T = 16
        self.N = int(np.prod(self.res))
        # algorithmic bytes per voxel per CG iteration: nc (6 s_f + 9 s_v) + T,
        # six sweeps of operator-precision fields (K:p input, two spectrum
        # round trips, Ap) and nine of CG vectors, plus the tangent; the FP32
        # form keeps operator fields, CG vectors and the tangent copy in FP32
        nc = len(self.res) ** 2
        s = 8 if self.prec == "fp64" else 4
        self.T_prec = self.T * s // 8
        self.B_alg = nc * (6 * s + 9 * s) + self.T_prec
        if not hasattr(self, "newton_maxiter"):
            self.newton_maxiter = 0
        self.converged = None

    def make_cell(self, mu):
        dim = len(self.res)
        form = mu.Formulation.finite_strain if self.form == "finite" else mu.Formulation.small_strain
        lengths = [1.0] * dim if self.name != "c1" else [7.0, 5.0]
        cell = mu.Cell(self.res, lengths, form, precision=self.prec,
                       allow_even=any(r % 2 == 0 for r in self.res))
        idx = np.arange(self.N, dtype=np.int64)
        m3 = mu.material
        if self.name == "c3":
            soft = m3.MaterialHyperElastoPlastic1_3d.make(cell, "soft", 70e9, .33, 200e6, 1e9,
                                                          consistent_tangent=True)
            hard = m3.MaterialHyperElastoPlastic1_3d.make(cell, "hard", 210e9, .33, 800e6, 1e9,
                                                          consistent_tangent=True)
            hard.add_pixels_by_index(idx[self.mask.ravel()])
            soft.add_pixels_by_index(idx[~self.mask.ravel()])
        elif self.name == "c2":
            inc = m3.MaterialLinearElastic1_3d.make(cell, "inclusion", 100e9, .3)
            mat = m3.MaterialLinearElastic1_3d.make(cell, "matrix", 10e9, .3)
            inc.add_pixels_by_index(idx[self.mask.ravel()])
            mat.add_pixels_by_index(idx[~self.mask.ravel()])
        elif self.name == "c5":
            m = m3.MaterialLinearElastic4_3d.make(cell, "lognormal")
            m.add_pixels_by_index(idx, np.stack([self.E, self.nu], axis=1))
        else:
            inc = m3.MaterialLinearElastic1_2d.make(cell, "inclusion", 10e9, .33)
            mat = m3.MaterialLinearElastic1_2d.make(cell, "matrix", 70e9, .33)
            inc.add_pixels_by_index(idx[self.mask.ravel()])
            mat.add_pixels_by_index(idx[~self.mask.ravel()])
        cell.initialise()
        return cell

    def solve(self, mu, cell):
        """All load cases of one step; returns (newton iterations, CG iterations)."""
        its = cg = 0
        for steps in self.loads:
            solver = mu.solvers.SolverCG(cell, self.cg_tol, self.maxiter, False)
            res = mu.solvers.newton_cg(cell, list(steps), solver, self.newton_tol,
                                       self.equil_tol, 0, newton_maxiter=self.newton_maxiter)
            self.converged = all(r.success for r in res)
            self.last = res
            its += sum(r.nb_it for r in res)
            cg += res[-1].nb_fev
        return its, cg


# -------------------------------------------------------- kernel accounting
K_IDS = {"r2c": 0, "middle": 1, "axis0": 2, "c2r": 3, "material": 4, "update": 5}
FAST = (16, 32, 48, 64, 96, 128, 256, 512)


def kernel_name(wl, k):
    """Symbol of the kernel behind a timing class for this grid: axis
    lengths 16..512 run the compile-time-length passes."""
    last, first = wl.res[-1], wl.res[0]
    mid = wl.res[1] if len(wl.res) == 3 else None
    return {0: "k_r2c_fast" if last in FAST else "k_last_r2c",
            1: "k_strided_fast (middle axis)" if mid in FAST else "k_strided (middle axis)",
            2: "k_axis0_fast" if first in FAST else "k_axis0_project",
            3: "k_c2r_fast" if last in FAST else "k_last_c2r",
            4: "material", 5: "k_cg_update"}[k]


def kernel_bytes(wl, kclass, prec):
    """Algorithmic HBM bytes of one launch: each input read once and each
    output written once.  For the spectrum passes this is a read and a write
    of the Hermitian half spectrum, Kpts x nc complex values."""
    dim = len(wl.res)
    nc = dim * dim
    s = 8 if prec == "fp64" else 4
    herm = list(wl.res[:-1]) + [wl.res[-1] // 2 + 1]
    kpts = int(np.prod(herm))
    if kclass in (1, 2):
        return 2 * kpts * nc * 2 * s
    # with compile-time-length passes on the last axis the x update of a CG
    # iteration rides on the next iteration's r2c (x += alpha p_old where p
    # is read anyway), so the r2c moves x twice more and the update kernel
    # reads neither x nor p
    fused = wl.res[-1] in FAST
    if kclass == 0:
        # r and p read, p written (p = -r + beta p), [x read and written],
        # tangent read, half spectrum written
        return (5 if fused else 3) * wl.N * nc * s + wl.N * wl.T_prec + kpts * nc * 2 * s
    if kclass == 3:
        # half spectrum read, Ap written, p read for the dot product
        return kpts * nc * 2 * s + 2 * wl.N * nc * s
    if kclass == 5:
        # [x, p read, x written,] r and Ap read, r written
        return (3 if fused else 6) * wl.N * nc * s
    return None


# ------------------------------------------------------------- CPU baseline
def host_info():
    """CPU model and core counts of the host this runs on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def _cpu_sample_cell(wl):
    """The oracle cell of the CPU sample: the workload's material and load at
    a reduced grid (at least 64^3 in 3D, so that fields and the stored tangent,
    170 MB and more, do not sit in the CPU's caches), strain, stress and
    tangent evaluated."""
    from oracle import oracle as O
    if wl.name in ("c3", "c2", "c5"):
        n = 64
        res = [n] * 3
    else:
        n = wl.n
        res = [n, n]
    N = int(np.prod(res))
    dim = len(res)
    if wl.name == "c3":
        mask = grf_mask(n, 16.0 * n / 256.0, 2026)
        c = O.Cell(res, [1.0] * 3, O.FINITE_STRAIN, allow_even=True)
        a = c.add_material(O.HEP1_CT, 70e9, .33, 200e6, 1e9)
        b = c.add_material(O.HEP1_CT, 210e9, .33, 800e6, 1e9)
        for i in range(N):
            c.add_pixel(b if mask.flat[i] else a, i)
        c.initialise()
        c.F[:] = np.eye(3).ravel(order="F") + np.diag([0, 0, .005]).ravel(order="F")
    elif wl.name == "c5":
        rng = np.random.default_rng(2026)
        E = np.exp(np.log(50e9) + 0.8 * rng.standard_normal(N))
        nu = rng.uniform(0.15, 0.45, N)
        c = O.Cell(res, [1.0] * 3, O.SMALL_STRAIN, allow_even=True)
        m = c.add_material(O.LE4)
        for i in range(N):
            c.add_pixel(m, i, [E[i], nu[i]])
        c.initialise()
        e = np.zeros((3, 3))
        e[0, 1] = e[1, 0] = 0.01
        c.F[:] = e.ravel(order="F")
    else:
        mask = spheres_mask(n, 40, 2026) if dim == 3 else wl.mask
        c = O.Cell(res, [1.0] * dim, O.SMALL_STRAIN, allow_even=n % 2 == 0)
        a = c.add_material(O.LE1, 100e9 if dim == 3 else 10e9, .3 if dim == 3 else .33)
        b = c.add_material(O.LE1, 10e9 if dim == 3 else 70e9, .3 if dim == 3 else .33)
        for i in range(N):
            c.add_pixel(a if mask.flat[i] else b, i)
        c.initialise()
        e = np.zeros((dim, dim))
        e[0, 1] = e[1, 0] = 0.01
        c.F[:] = e.ravel(order="F")
    c.evaluate_stress_tangent()
    return c, res, N


def _cpu_rate(O, c, rhs, N, threads, budget_s):
    O.set_threads(threads)
    c.cg_fixed_iterations(rhs, 1)  # warm-up
    iters = 0
    t0 = time.perf_counter()
    while True:
        c.cg_fixed_iterations(rhs, 1)
        iters += 1
        el = time.perf_counter() - t0
        if el > budget_s or iters >= 400:
            break
    return N * iters / el, iters, el


def cpu_baseline(wl, budget_s):
    """The oracle (CPU restatement of the reference's FFTW-double path) timed
    on this host over a bounded sample of the workload: one thread (the
    reference is single-threaded) and the host's CPU share (OMP_NUM_THREADS,
    else every core), each for about budget_s seconds of CG iterations."""
    from oracle import oracle as O
    c, res, N = _cpu_sample_cell(wl)
    rhs = -c.apply_projection(c.P.copy())
    info = host_info()
    share = int(info["omp_num_threads"] or info["affinity"] or 1)
    share = max(1, min(share, info["affinity"] or share))
    r1, i1, e1 = _cpu_rate(O, c, rhs, N, 1, budget_s)
    rn, i_n, e_n = _cpu_rate(O, c, rhs, N, share, budget_s)
    O.set_threads(1)
    dim = len(res)
    sample = ("{} {} voxels, CG iterations of the oracle (FP64, stored {}x{} tangent and "
              "projection operator as in the reference): {} iterations in {:.1f} s on {} threads"
              .format(wl.name, "x".join(map(str, res)), dim * dim, dim * dim, i_n, e_n, share))
    return {"value": round(rn, 1), "unit": "voxel-CG-iters/s", "cores": share, "kind": "port",
            "sample": sample,
            "single_thread": {"value": round(r1, 1), "cores": 1,
                              "sample": "{} iterations in {:.1f} s".format(i1, e1)},
            "host": info}


def copy_bandwidth(capi, nbytes=4 << 30, reps=5):
    """GB/s (read + write) of device copies: the library's copy kernel, a
    streaming copy with four 16-byte loads in flight per lane (plain and
    non-temporal) and hipMemcpyAsync, each timed with HIP events."""
    lib = capi.load()
    out = (ctypes.c_double * 6)()
    capi.check(lib.mugpu_copy_bandwidth(nbytes, reps, out))
    names = ["copy_kernel", "streaming", "streaming_nontemporal", "hipMemcpy", "read_only",
             "write_only"]
    return {n: round(v, 1) for n, v in zip(names, out)}


def line_bandwidth(capi, wl, reps=3):
    """GB/s (read + write) of the first-axis pass's HBM pattern alone, for
    this workload's spectrum (n rows of plane k-points, nc components) with
    tiles of 1, 2 and 4 k-points: the bandwidth that access pattern allows,
    before any arithmetic."""
    lib = capi.load()
    n = wl.res[0]
    if n % 16:
        return None
    dim = len(wl.res)
    nc = dim * dim
    herm = list(wl.res[:-1]) + [wl.res[-1] // 2 + 1]
    plane = int(np.prod(herm)) // n
    prec = capi.FP64 if wl.prec == "fp64" else capi.FP32
    out = {}
    for kp in (1, 2, 4):
        if kp * nc * (n // 16) > 1024:
            continue
        g = ctypes.c_double()
        capi.check(lib.mugpu_line_bandwidth(prec, n, plane, nc, kp, reps, ctypes.byref(g)))
        out["kpts_per_tile_{}".format(kp)] = round(g.value, 1)
    return out


def kernel_read_write_bytes(wl, kclass, prec):
    """(read, written) algorithmic bytes of one launch (their sum is
    kernel_bytes)."""
    dim = len(wl.res)
    nc = dim * dim
    s = 8 if prec == "fp64" else 4
    herm = list(wl.res[:-1]) + [wl.res[-1] // 2 + 1]
    spec = int(np.prod(herm)) * nc * 2 * s
    field = wl.N * nc * s
    fused = wl.res[-1] in FAST
    if kclass in (1, 2):
        return spec, spec
    if kclass == 0:
        return (3 if fused else 2) * field + wl.N * wl.T_prec, (2 if fused else 1) * field + spec
    if kclass == 3:
        return spec + field, field
    if kclass == 5:
        return (2 if fused else 4) * field, (1 if fused else 2) * field
    return None


def achievable_time(rw, copy_gbs):
    """Seconds one launch needs at the measured streams: its reads at the
    read-only rate and its writes at the write-only rate."""
    r, w = rw
    return r / (copy_gbs["read_only"] * 1e9) + w / (copy_gbs["write_only"] * 1e9)


def traffic_from_profiles(wl, kname):
    """HBM bytes per launch measured with rocprofv3 PMC counters, if a
    summary for this workload and kernel is committed under profiles/."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if (len(parts) >= 4 and parts[0] == wl.name + ":" + str(wl.n) and parts[1] == wl.prec
                and parts[2] == kname.split()[0]):
            try:
                return float(parts[3])
            except ValueError:
                return None
    return None


def homogenised_stress_at(mu, wl, grad):
    """<P> evaluated at a strain field on a fresh FP64 cell (initial history),
    so that two solutions are compared on the same footing: newton_cg reports
    the stress of the last evaluated strain, which precedes the last update."""
    nc = len(wl.res) ** 2
    prec = wl.prec
    wl.prec = "fp64"
    cell = wl.make_cell(mu)
    wl.prec = prec
    cell.evaluate_stress_tangent(np.asfortranarray(np.asarray(grad).reshape(-1, nc).T))
    out = np.asarray(cell.get_stress()).reshape(nc, -1).mean(1)
    del cell
    return out


def mixed_precision_error(mu, wl):
    """Untimed: the workload again in plain FP64 on the GPU (itself held to
    1e-10 of the FP64 oracle by the parity tests), then the relative error of
    the mixed-precision result in homogenised stress and in the strain field
    (F - I for finite strain)."""
    nc = len(wl.res) ** 2
    g32 = [np.asarray(r.grad).reshape(-1, nc) for r in wl.last]
    wl.prec = "fp64"
    cell = wl.make_cell(mu)
    wl.solve(mu, cell)
    del cell
    g64 = [np.asarray(r.grad).reshape(-1, nc) for r in wl.last]
    wl.prec = "fp32"
    eye = np.eye(len(wl.res)).ravel() if wl.form == "finite" else 0.0
    es = ef = 0.0
    for a, b in zip(g32, g64):
        s32 = homogenised_stress_at(mu, wl, a)
        s64 = homogenised_stress_at(mu, wl, b)
        es = max(es, float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64)))
        ef = max(ef, float(np.linalg.norm(a - b) / np.linalg.norm(b - eye)))
    return {"homogenised_stress_rel": es, "strain_field_rel": ef,
            "against": "plain FP64 GPU solve of the same workload",
            "strain_measure": "F - I" if wl.form == "finite" else "eps"}


# ------------------------------------------------------------ C4 serving
C4_3D = [32, 48, 64, 96]
C4_2D = [64, 128, 256, 512]


def c4_request(seed):
    """One request of the serving stream, from its own seed."""
    rng = np.random.default_rng(seed)
    if rng.random() < 0.7:
        res = [int(rng.choice(C4_3D))] * 3
    else:
        res = [int(rng.choice(C4_2D))] * 2
    d = len(res)
    w = rng.standard_normal(res).astype(np.float32)
    k2 = sum(np.meshgrid(*[np.fft.fftfreq(n).astype(np.float32) ** 2 for n in res],
                         indexing="ij", sparse=True))
    ell = 4.0 * res[0] / 32.0
    ax = tuple(range(d))
    f = np.fft.irfftn(np.fft.rfftn(w, axes=ax) *
                      np.exp(-0.5 * (2 * np.pi) ** 2 * k2[..., :res[-1] // 2 + 1] * ell ** 2),
                      s=res, axes=ax)
    mask = (f > np.median(f)).ravel()
    contrast = np.exp(rng.uniform(0.0, np.log(100.0)))
    nu = rng.uniform(0.2, 0.4)
    a = rng.uniform(-0.01, 0.01, (d, d))
    # phases kept as a mask; the per-voxel moduli are formed at submission
    return res, np.packbits(mask), 10e9 * contrast, 10e9, nu, 0.5 * (a + a.T)


def c4_arrays(q):
    res, bits, e_in, e_mat, nu, eps = q
    N = int(np.prod(res))
    mask = np.unpackbits(bits, count=N).astype(bool)
    return res, np.where(mask, e_in, e_mat), np.full(N, nu), eps


def c4_stream(n, seed0):
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=8) as ex:
        return list(ex.map(c4_request, range(seed0, seed0 + n), chunksize=16))


def c4_inputs(reqs):
    """The requests' per-voxel E and nu as host arrays: the inputs a client
    hands the server, formed before a timed phase starts."""
    return [c4_arrays(q) for q in reqs]


def c4_run(srv, reqs, arrivals):
    """Submit each request (res, E, nu, strain: host arrays formed before the
    call) at its arrival time (seconds after the start) and turn the loop
    until all have finished; returns (wall seconds, latencies, results)."""
    n = len(reqs)
    order = np.argsort(arrivals, kind="stable")
    t_done = np.zeros(n)
    idmap = {}
    nxt = 0
    finished = 0
    t0 = time.perf_counter()
    last_print = t0
    while finished < n:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[order[nxt]] <= now:
            i = order[nxt]
            res, E, nu, eps = reqs[i]
            idmap[srv.submit(res, E, nu, eps)] = i
            nxt += 1
        srv.progress(wait=False)
        for rid in srv.poll():
            t_done[idmap[rid]] = time.perf_counter() - t0
            finished += 1
        if time.perf_counter() - last_print > 20:
            last_print = time.perf_counter()
            print("bench: c4 {}/{} finished".format(finished, n), file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    lat = t_done - np.asarray(arrivals)
    results = {i: srv.result(rid) for rid, i in idmap.items()}
    for rid in idmap:
        srv.release(rid)
    return wall, lat, results


def c4_main(args):
    from muspectre_amd import serving
    n = args.requests
    print("bench: c4 generating {} + {} requests".format(n, n), file=sys.stderr, flush=True)
    sat_reqs = c4_stream(n, 1)
    poi_reqs = c4_stream(n, 100001)
    srv = serving.RVEServer(max_batch=32, iters_per_launch=8, cg_tol=1e-6, cg_maxiter=4000,
                            use_graphs=not args.no_graphs)
    # warm-up: one request of each resolution (plans, graphs)
    warm = {}
    for q in sat_reqs + poi_reqs:
        warm.setdefault(tuple(q[0]), q)
    c4_run(srv, c4_inputs(warm.values()), np.zeros(len(warm)))
    print("bench: c4 forming the saturated phase's inputs", file=sys.stderr, flush=True)
    inputs = c4_inputs(sat_reqs)
    wall, lat, res = c4_run(srv, inputs, np.zeros(n))
    del inputs
    throughput = n / wall
    voxel_iters = sum(int(np.prod(sat_reqs[i][0])) * r["cg_iterations"] for i, r in res.items())
    failed = sum(1 for r in res.values() if not r["ok"])
    print("bench: c4 saturated: {} requests in {:.2f} s, {:.1f} solves/s".format(n, wall, throughput),
          file=sys.stderr, flush=True)
    rate = args.load * throughput
    arr = np.cumsum(np.random.default_rng(7).exponential(1.0 / rate, n))
    print("bench: c4 forming the Poisson phase's inputs", file=sys.stderr, flush=True)
    inputs = c4_inputs(poi_reqs)
    wall2, lat2, res2 = c4_run(srv, inputs, arr)
    del inputs
    st = srv.stats()
    srv.close()
    p50, p99 = np.percentile(lat2, [50, 99])
    # latency by resolution: the small cells wait in the same queue as the
    # large ones, so their tail shows how the stream is shared
    by_res = {}
    for i, q in enumerate(poi_reqs):
        by_res.setdefault("x".join(map(str, q[0])), []).append(lat2[i])
    per_res = {k: {"requests": len(v),
                   "p50_ms": round(1e3 * float(np.percentile(v, 50)), 2),
                   "p99_ms": round(1e3 * float(np.percentile(v, 99)), 2)}
               for k, v in sorted(by_res.items(), key=lambda kv: (len(kv[0].split("x")),
                                                                   int(kv[0].split("x")[0])))}
    out = {
        "metric": "RVE solves/s",
        "value": round(throughput, 2),
        "unit": "RVE solves/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(1e3 * wall, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-operator/f64-refined",
        "data": "synthetic",
        "config": {"workload": "c4", "requests": n, "mix": "3D p=.7 {32,48,64,96}^3, 2D {64..512}^2",
                   "cg_tol": 1e-6, "formulation": "small", "max_batch": 32, "iters_per_launch": 8,
                   "graphs": not args.no_graphs,
                   "inputs": "per-voxel E and nu host arrays, formed before each timed phase"},
        "saturated": {"requests": n, "wall_s": round(wall, 3), "solves_per_s": round(throughput, 2),
                      "voxel_cg_iters_per_s": round(voxel_iters / wall, 1), "failed": failed},
        "poisson": {"rate_per_s": round(rate, 2), "load": args.load, "requests": n,
                    "wall_s": round(wall2, 3), "solves_per_s": round(n / wall2, 2),
                    "p50_ms": round(1e3 * p50, 2), "p99_ms": round(1e3 * p99, 2),
                    "per_resolution": per_res,
                    "failed": sum(1 for r in res2.values() if not r["ok"])},
        "server": st,
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU: every box here has one GPU and this build "
              "has no multi-GPU path (--gpus {} requested)".format(args.gpus), file=sys.stderr)
        sys.exit(2)
    if args.workload == "c4":
        return c4_main(args)
    import muspectre_amd as mu
    from muspectre_amd import _capi
    lib = _capi.load()
    _capi.check(lib.mugpu_init(0))
    wl = Workload(args)
    print("bench: workload {} {} {} ({} voxels), precision {}".format(
        wl.name, "x".join(map(str, wl.res)), wl.form, wl.N, wl.prec), file=sys.stderr, flush=True)
    for w in range(args.warmup):
        cell = wl.make_cell(mu)
        its, cg = wl.solve(mu, cell)
        print("bench: warmup {}: {} Newton, {} CG iterations".format(w, its, cg),
              file=sys.stderr, flush=True)
        del cell

    def timed_step(kernel_class):
        """One step: fresh cell (untimed), then newton_cg timed end to end with
        HIP events around every launch of kernel_class (MUGPU_K_ALL: all)."""
        t0 = time.perf_counter()
        cell = wl.make_cell(mu)
        _capi.check(lib.mugpu_synchronize())
        setup = time.perf_counter() - t0
        if kernel_class >= 0:
            _capi.check(lib.mugpu_timing_start(kernel_class))
        _capi.check(lib.mugpu_synchronize())
        t0 = time.perf_counter()
        its, cg = wl.solve(mu, cell)
        _capi.check(lib.mugpu_synchronize())
        dt = time.perf_counter() - t0
        launches = ctypes.c_int64()
        kms = ctypes.c_double()
        st = {k: (0, 0.0) for k in K_IDS.values()}
        if kernel_class >= 0:
            _capi.check(lib.mugpu_timing_stop(ctypes.byref(launches), ctypes.byref(kms)))
            for k in K_IDS.values():
                _capi.check(lib.mugpu_timing_class(k, ctypes.byref(launches), ctypes.byref(kms)))
                st[k] = (launches.value, kms.value)
        del cell
        return dt, setup, its, cg, st

    # per-kernel breakdown: one untimed step with events around every launch
    # (the events add a marker between every two kernels, so the timed steps
    # below time only the roofline kernel)
    bdt, _, _, _, bstats = timed_step(_capi.K_ALL)
    print("bench: breakdown step: {:.3f} s with events around every launch".format(bdt),
          file=sys.stderr, flush=True)
    kernels = {}
    for name, k in K_IDS.items():
        n, ms = bstats[k]
        if not n:
            continue
        avg = ms / n / 1e3
        kb = kernel_bytes(wl, k, wl.prec)
        e = {"kernel": kernel_name(wl, k), "launches": n, "avg_launch_us": round(avg * 1e6, 2),
             "share_of_step": round(ms / 1e3 / bdt, 4)}
        if kb:
            e["bytes_per_launch"] = kb
            e["achieved_GBs"] = round(kb / avg / 1e9, 1)
            e["frac"] = round(kb / avg / 1e9 / HBM_PEAK_GBS, 4)
        kernels[name] = e
    kchoice = args.kernel
    if kchoice == "auto":
        # the dominant kernel: largest share of the step among the HBM passes
        kchoice = max((n for n in kernels if "frac" in kernels[n]),
                      key=lambda n: kernels[n]["share_of_step"])
    kclass = K_IDS[kchoice]
    # On cells of at most 2^20 values the CG chunks are replayed from captured
    # hipGraphs, which per-launch events would switch off: the timed steps
    # then run without events and the roofline kernel is timed in the
    # breakdown step instead
    nc = len(wl.res) ** 2
    graph_regime = wl.N * nc <= (1 << 20)
    timed_class = -1 if graph_regime else kclass

    total = 0.0
    setup = 0.0
    tot_its = tot_cg = 0
    klaunch, kms_tot = 0, 0.0
    for s_ in range(args.steps):
        dt, st_setup, its, cg, st = timed_step(timed_class)
        total += dt
        setup += st_setup
        tot_its += its
        tot_cg += cg
        if not graph_regime:
            klaunch += st[kclass][0]
            kms_tot += st[kclass][1]
    if graph_regime:
        klaunch, kms_tot = bstats[kclass]
        print("bench: step {}: {:.3f} s, {} Newton, {} CG iterations".format(s_, dt, its, cg),
              file=sys.stderr, flush=True)
    value = wl.N * tot_cg / total
    roof = None
    kb = kernel_bytes(wl, kclass, wl.prec)
    if klaunch and kb:
        # live over the timed steps: HIP events on the library stream around
        # every launch of the roofline kernel
        avg = kms_tot / klaunch / 1e3
        achieved = kb / avg / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "timed_in": ("the breakdown step (the timed steps replay captured graphs)"
                             if graph_regime else "the timed steps"),
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(wl, kernel_name(wl, kclass)),
                "traffic_source": "committed rocprofv3 PMC figure (profiles/hbm_traffic.txt: "
                                  "FETCH_SIZE + WRITE_SIZE per launch of this kernel, "
                                  "calibrated on a copy of known size), not measured in this run",
                "kernel": kernel_name(wl, kclass), "launches": klaunch,
                "avg_launch_us": round(avg * 1e6, 2), "bytes_per_launch": kb}
    fp32_error = None
    if wl.prec == "fp32":
        fp32_error = mixed_precision_error(mu, wl)
        print("bench: fp32 error vs fp64: stress {:.3e}, strain {:.3e}".format(
            fp32_error["homogenised_stress_rel"], fp32_error["strain_field_rel"]),
            file=sys.stderr, flush=True)
    copy_gbs = copy_bandwidth(_capi)
    achievable = max(copy_gbs[k] for k in ("copy_kernel", "streaming", "streaming_nontemporal",
                                            "hipMemcpy"))
    if roof is not None:
        roof["frac_of_best_copy"] = round(roof["achieved"] / achievable, 4)
        # the kernel's own mix of reads and writes at the measured read-only and
        # write-only rates: the time it would take at those streams' speeds
        rw = kernel_read_write_bytes(wl, kclass, wl.prec)
        if rw:
            t_ach = achievable_time(rw, copy_gbs)
            roof["frac_of_stream_mix"] = round(t_ach / (roof["avg_launch_us"] * 1e-6), 4)
    for name, e in kernels.items():
        rw = kernel_read_write_bytes(wl, K_IDS[name], wl.prec)
        if rw:
            e["frac_of_stream_mix"] = round(achievable_time(rw, copy_gbs) /
                                            (e["avg_launch_us"] * 1e-6), 4)
    line_gbs = line_bandwidth(_capi, wl)
    if line_gbs and "axis0" in kernels and kernels["axis0"].get("achieved_GBs"):
        kernels["axis0"]["line_pattern_GBs"] = line_gbs
        kernels["axis0"]["frac_of_line_pattern"] = round(
            kernels["axis0"]["achieved_GBs"] / max(line_gbs.values()), 4)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, args.cpu_seconds)
    out = {
        "metric": "voxel-CG-iters/s",
        "value": round(value, 1),
        "unit": "voxel-CG-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if wl.prec == "fp64" else "f32-operator/f64-refined",
        "data": "synthetic",
        "config": {"workload": wl.name, "grid": "x".join(map(str, wl.res)), "formulation": wl.form,
                   "newton_iterations": tot_its, "cg_iterations": tot_cg,
                   "load_cases": len(wl.loads), "precision": wl.prec,
                   "newton_tol": wl.newton_tol, "cg_tol": wl.cg_tol,
                   "newton_maxiter": wl.newton_maxiter or None, "converged": wl.converged,
                   **({"timed_prefix": wl.prefix_note} if hasattr(wl, "prefix_note") else {}),
                   "setup_s_per_step": round(setup / args.steps, 2),
                   **({"inclusion_fraction": round(wl.inclusion_fraction, 4)}
                      if hasattr(wl, "inclusion_fraction") else {})},
        "roofline": roof,
        "whole_iteration_roofline": {
            "bytes_per_voxel_iter": wl.B_alg,
            "achieved": round(wl.B_alg * value / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(wl.B_alg * value / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels": kernels,
        "kernels_note": "untimed breakdown step with HIP events around every launch",
        "copy_bandwidth_GBs": copy_gbs,
        "best_copy_GBs": achievable,
        "fp32_error": fp32_error,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
