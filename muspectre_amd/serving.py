"""Serving loop for many independent RVE solves on one GPU.

Thin Python front end of the native server in libmugpu.so
(mugpu_server_* in include/mugpu.h): requests are queued with submit(),
the loop is turned with progress(), finished requests are collected with
poll() and read with result().  Scheduling, admission by device footprint,
batching and graph replay all happen in the native code.

Each request is one small-strain load step of newton_cg
(src/solver/solvers.cc:44-217) on a cell with per-voxel isotropic linear
elasticity (MaterialLinearElastic4, material_linear_elastic4.hh:189-210),
solved by the mixed-precision SolverCG (FP32 inner CG with FP64 dots,
FP64 residual refinement).  A request computes exactly what the same solve
computes alone, whatever else is in flight.
"""
import ctypes

import numpy as np

from . import _capi

_STATS = ("submitted", "admitted", "finished", "failed", "batched_launches",
          "batched_slots", "peak_bytes", "budget_bytes", "peak_running")


class RVEServer:
    """Continuous-batching RVE solver.

    hbm_budget: device bytes for admitted requests (0: 3/4 of free memory);
    max_batch: slots per resolution; iters_per_launch: inner CG iterations
    per batched launch; cg_tol, cg_maxiter, inner_tol: SolverCG parameters;
    newton_tol, equil_tol: newton_cg tolerances; keep_fields: keep strain and
    stress fields of finished requests for result(fields=True).
    """

    def __init__(self, hbm_budget=0, max_batch=32, iters_per_launch=8, use_graphs=True,
                 keep_fields=False, cg_tol=1e-6, cg_maxiter=2000, inner_tol=1e-4,
                 newton_tol=1e-5, equil_tol=1e-10):
        self._lib = _capi.load()
        cfg = _capi.ServerConfig(hbm_budget=int(hbm_budget), max_batch=int(max_batch),
                                 iters_per_launch=int(iters_per_launch),
                                 use_graphs=1 if use_graphs else 0,
                                 keep_fields=1 if keep_fields else 0, cg_tol=cg_tol,
                                 cg_maxiter=int(cg_maxiter), inner_tol=inner_tol,
                                 newton_tol=newton_tol, equil_tol=equil_tol)
        h = ctypes.c_void_p()
        _capi.check(self._lib.mugpu_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._keep = bool(keep_fields)
        self._shape = {}

    def close(self):
        if self._h:
            self._lib.mugpu_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def footprint(res):
        """Device bytes one request of this resolution occupies while it runs."""
        lib = _capi.load()
        r = (ctypes.c_int * len(res))(*res)
        b = ctypes.c_size_t()
        _capi.check(lib.mugpu_server_footprint(len(res), r, ctypes.byref(b)))
        return b.value

    def submit(self, res, young, poisson, strain, lengths=None, cg_maxiter=0):
        """Queue one solve; young and poisson are per-voxel arrays (row-major
        pixels), strain the symmetric macroscopic strain (dim x dim)."""
        dim = len(res)
        N = int(np.prod(res))
        E = np.ascontiguousarray(np.broadcast_to(np.asarray(young, dtype=np.float64).ravel(), (N,)))
        nu = np.ascontiguousarray(np.broadcast_to(np.asarray(poisson, dtype=np.float64).ravel(), (N,)))
        eps = np.asfortranarray(np.asarray(strain, dtype=np.float64).reshape(dim, dim))
        eps_flat = np.ascontiguousarray(eps.ravel(order="F"))
        r = (ctypes.c_int * dim)(*res)
        L = None if lengths is None else (ctypes.c_double * dim)(*lengths)
        rid = ctypes.c_int64()
        dp = ctypes.POINTER(ctypes.c_double)
        _capi.check(self._lib.mugpu_server_submit(
            self._h, dim, r, L, E.ctypes.data_as(dp), nu.ctypes.data_as(dp),
            eps_flat.ctypes.data_as(dp), int(cg_maxiter), ctypes.byref(rid)))
        self._shape[rid.value] = (N, dim * dim)
        return rid.value

    def progress(self, wait=False):
        """One turn of the loop; returns the number of unfinished requests."""
        n = ctypes.c_int()
        _capi.check(self._lib.mugpu_server_progress(self._h, 1 if wait else 0, ctypes.byref(n)))
        return n.value

    def poll(self, capacity=4096):
        """Ids of the requests finished since the last call."""
        ids = (ctypes.c_int64 * capacity)()
        n = ctypes.c_int()
        _capi.check(self._lib.mugpu_server_poll(self._h, ids, capacity, ctypes.byref(n)))
        return list(ids[:n.value])

    def result(self, rid, fields=False):
        N, nc = self._shape[rid]
        status, conv, nit = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        cg = ctypes.c_int64()
        mean = np.zeros(nc)
        dp = ctypes.POINTER(ctypes.c_double)
        F = P = None
        if fields:
            if not self._keep:
                raise ValueError("the server was created without keep_fields")
            F = np.zeros((N, nc))
            P = np.zeros((N, nc))
        _capi.check(self._lib.mugpu_server_result(
            self._h, rid, ctypes.byref(status), ctypes.byref(conv), ctypes.byref(nit),
            ctypes.byref(cg), mean.ctypes.data_as(dp),
            None if F is None else F.ctypes.data_as(dp),
            None if P is None else P.ctypes.data_as(dp)))
        out = {"status": status.value, "ok": status.value == _capi.OK,
               "converged": bool(conv.value), "newton_iterations": nit.value,
               "cg_iterations": cg.value, "mean_stress": mean}
        if fields:
            out["strain"] = F
            out["stress"] = P
        return out

    def release(self, rid):
        _capi.check(self._lib.mugpu_server_release(self._h, rid))
        self._shape.pop(rid, None)

    def stats(self):
        out = (ctypes.c_int64 * len(_STATS))()
        _capi.check(self._lib.mugpu_server_stats(self._h, out, len(_STATS)))
        return dict(zip(_STATS, list(out)))

    def run(self, wait=True):
        """Turn the loop until every submitted request has finished."""
        while self.progress(wait=wait):
            pass
