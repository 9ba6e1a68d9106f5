// Serving loop: many independent RVE solves of uneven resolution on one GPU.
//
// A request is one small-strain load step of newton_cg (src/solver/
// solvers.cc:44-217: check_load_steps, F = 0 + dE, then Newton iterations
// of stress/tangent evaluation, rhs = -G:P, SolverCG, F += dF and the
// incremental-norm / equilibrium tests) with the mixed-precision SolverCG of
// this library (FP64 residual refinement around an FP32 inner CG).  The
// control flow of both is replayed step for step as an event-driven state
// machine, so the host never blocks on one request while others could run.
//
// Each resolution class has two streams of its own.  Device work of a request
// runs on its class's request stream, except its inner CG iterations: those
// of all running requests of one resolution share batched launches on the
// class's batch stream (one workgroup grid per launch, one RVE per
// blockIdx.z; slots packed at the front).  A batch chunk is K iterations,
// replayed from a captured hipGraph (one graph per power-of-two slot
// count).  Every slot carries its own scalars and stopping test, evaluated
// on the device, so a slot that finishes is skipped by every later launch
// and stops exactly where the same solve stops alone; new requests take
// free slots between chunks (continuous batching).  The material stage
// (stress, tangent) of the requests of one resolution that reach it in the
// same pass of the loop runs as one launch too (one request per blockIdx.y).
//
// Memory: one device arena of the HBM budget; a request is admitted (first
// come, first served) when its whole footprint fits, and gives it back when
// its result has been copied out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <thread>
#include <vector>

#include "mugpu_internal.hpp"

namespace mugpu {
double lame_lambda_h(double E, double nu) { return E * nu / ((1 + nu) * (1 - 2 * nu)); }
double lame_mu_h(double E, double nu) { return E / (2 * (1 + nu)); }

// lm[2v], lm[2v+1] = (lambda, mu) of voxel v.  Submission runs on the thread
// that turns the serving loop, so a large request is split over up to eight
// host threads, which also share the first-touch page faults of lm (left
// uninitialised for that reason); every voxel's arithmetic is the same in
// any split.  A thread that cannot be started leaves its range to this one.
static void fill_lame(const double* E, const double* nu, int64_t N, double* lm) {
  auto part = [=](int64_t v0, int64_t v1) {
    for (int64_t v = v0; v < v1; ++v) {
      lm[2 * v] = lame_lambda_h(E[v], nu[v]);
      lm[2 * v + 1] = lame_mu_h(E[v], nu[v]);
    }
  };
  const int T = (int)std::min<int64_t>(8, std::max<int64_t>(1, N / 65536));
  std::vector<std::thread> pool;
  int started = 1;
  for (int t = 1; t < T; ++t) {
    try {
      pool.emplace_back(part, N * t / T, N * (t + 1) / T);
    } catch (...) {
      break;
    }
    ++started;
  }
  part(0, N / T);
  for (int t = started; t < T; ++t) part(N * t / T, N * (t + 1) / T);
  for (auto& th : pool) th.join();
}
}  // namespace mugpu

using namespace mugpu;

namespace {

constexpr size_t kAlign = 256;
constexpr int kMaxSlots = 64;
constexpr int kHostDoubles = 32;  // pinned host words per request
// host word layout: [0, SC_COUNT) scalar mirror, [16, 25) column sums,
// [31] staging of the inner threshold
constexpr int kHostColsum = 16, kHostTol2 = 31;

size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// per-component sums of an AoS field over its pixels (deterministic)
constexpr int kColT = 256, kColBlocks = 256;
__global__ __launch_bounds__(kColT) void k_colsum(const double* __restrict__ f, int64_t npix,
                                                  int nc, double* __restrict__ part) {
  __shared__ double red[kColT];
  double acc[9];
  for (int c = 0; c < 9; ++c) acc[c] = 0.0;
  for (int64_t v = (int64_t)blockIdx.x * kColT + threadIdx.x; v < npix;
       v += (int64_t)gridDim.x * kColT)
    for (int c = 0; c < nc; ++c) acc[c] += f[v * nc + c];
  for (int c = 0; c < nc; ++c) {
    red[threadIdx.x] = acc[c];
    __syncthreads();
    for (int s = kColT / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * nc + c] = red[0];
    __syncthreads();
  }
}
__global__ void k_colsum_final(const double* __restrict__ part, int nblocks, int nc,
                               double* __restrict__ out) {
  const int c = threadIdx.x;
  if (c >= nc) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += part[b * nc + c];
  out[c] = s;
}

// ---- device arena ---------------------------------------------------------------
class Arena {
 public:
  int init(size_t bytes) {
    if (hipMalloc((void**)&base_, bytes) != hipSuccess) {
      set_error("device allocation of the serving arena (%zu bytes) failed", bytes);
      return MUGPU_EMEMORY;
    }
    size_ = bytes;
    free_[0] = bytes;
    return MUGPU_OK;
  }
  ~Arena() {
    if (base_) (void)hipFree(base_);
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t start = it->first, rem = it->second - bytes;
      free_.erase(it);
      if (rem) free_[start + bytes] = rem;
      *off = start;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(bytes);
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
    used_ -= bytes;
  }
  char* base() const { return base_; }
  size_t size() const { return size_; }
  size_t peak() const { return peak_; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

// ---- per-request buffer layout -----------------------------------------------------
struct Layout {
  size_t F, P, rhs, x, r64, params, tan64, tan32, x32, r32, p32, Ap32, spec, scal, part, zpart,
      colpart, total;
};

Layout layout_of(int dim, const int* res) {
  int64_t N = 1, kp = 1;
  for (int d = 0; d < dim; ++d) N *= res[d];
  for (int d = 0; d < dim; ++d) kp *= (d == dim - 1 ? res[d] / 2 + 1 : res[d]);
  const int nc = dim * dim;
  const int64_t nd = N * nc;
  const int64_t rows = N / res[dim - 1];
  // c2r workgroups of the FP32 operator (one p.Ap partial each), counted for
  // the variant with the most (half the rows), so that the buffer holds the
  // partials whichever variant the class's plan runs
  const int rb = fast_rows_per_block(res[dim - 1], dim, 2);
  const int64_t nz = (rows + rb - 1) / rb;
  Layout l{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += align_up(bytes);
    return at;
  };
  l.F = take(8 * nd);
  l.P = take(8 * nd);
  l.rhs = take(8 * nd);
  l.x = take(8 * nd);
  l.r64 = take(8 * nd);
  l.params = take(8 * 2 * N);
  l.tan64 = take(8 * tangent_values(TAN_ISO, dim, N));
  l.tan32 = take(4 * tangent_values(TAN_ISO, dim, N));
  l.x32 = take(4 * nd);
  l.r32 = take(4 * nd);
  l.p32 = take(4 * nd);
  l.Ap32 = take(4 * nd);
  l.spec = take(16 * kp * nc);
  l.scal = take(8 * SC_COUNT);
  l.part = take(8 * kRedBlocks);
  l.zpart = take(8 * nz);
  l.colpart = take(8 * kColBlocks * nc);
  l.total = o;
  return l;
}

enum Stage {
  ST_QUEUED,
  ST_EVAL_QUEUED,  // waiting for its class's batched material stage
  ST_WAIT_EVAL,    // stress, tangent, rhs = -G:P and |rhs|^2 queued
  ST_WAIT_BNORM,   // |rhs|^2 of the solve
  ST_WAIT_RESID,   // FP64 residual |b - A x|^2
  ST_WAIT_RDR,     // inner CG start vectors and r.r
  ST_BATCH_QUEUE,  // waiting for a batch slot
  ST_BATCH,        // inner CG iterations in a batch
  ST_WAIT_NORMS,   // |dF|^2 and |F|^2 after the update
  ST_WAIT_RESULT,  // column sums (and fields) copied out
  ST_DONE
};

struct ResClass;

struct Request {
  int64_t id = 0;
  int dim = 0, nc = 0, res[3] = {1, 1, 1};
  double L[3] = {1, 1, 1};
  int64_t N = 0, nd = 0;
  // per-voxel (lambda, mu), interleaved; uninitialised until submit fills
  // it, freed once the first stage has completed
  std::unique_ptr<double[]> lam_mu;
  double strain[9] = {0};
  int maxiter = 0;
  Layout lay{};
  size_t off = 0;
  ResClass* cls = nullptr;
  double *F = nullptr, *P = nullptr, *rhs = nullptr, *x = nullptr, *r64 = nullptr,
         *params = nullptr, *tan64 = nullptr, *scal = nullptr, *part = nullptr,
         *zpart = nullptr, *colpart = nullptr;
  float *tan32 = nullptr, *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *Ap32 = nullptr;
  void* spec = nullptr;
  double* h = nullptr;  // pinned host words
  hipEvent_t ev = nullptr;
  Stage st = ST_QUEUED;
  // newton_cg state (solvers.cc:148-163)
  double incr_norm = 0, grad_norm = 1, stress_norm = 0;
  bool has_converged = false;
  int newt_iter = 0;
  int64_t counter = 0;  // SolverCG counter (cumulative over solves)
  // mixed-precision solve state
  int64_t its = 0;
  int refs = 0;
  double tol2 = 0, rr = 0, s = 1, bnorm2 = 0;
  int inner_maxit = 0;
  int inner_iter = 0;  // host mirror of the slot's iteration count
  // result
  int status = MUGPU_OK;
  bool converged = false;
  double mean_stress[9] = {0};
  std::vector<double> hF, hP;

  bool test(double newton_tol, double equil_tol) {
    const bool incr_test = incr_norm / grad_norm <= newton_tol;
    const bool stress_test = stress_norm < equil_tol;
    has_converged = incr_test || stress_test;
    return has_converged;
  }
};

// One resolution class: its plan, its batch slots and its two streams (the
// batched inner iterations on sb, the requests' own stages on sr), so that
// the work of one resolution never queues behind another's.  A request's
// stages and its class's chunks hand over to each other only after the host
// has seen the previous one complete (hipEventQuery), which orders them
// across the two streams.  It owns what it holds: a class that failed half
// way through its set-up frees what it had taken; the server destroys its
// classes only after the device is idle.
struct ResClass {
  ResClass() = default;
  ResClass(const ResClass&) = delete;
  ResClass& operator=(const ResClass&) = delete;
  ~ResClass() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (ev) (void)hipEventDestroy(ev);
    if (sb) (void)hipStreamDestroy(sb);
    if (sr) (void)hipStreamDestroy(sr);
    if (d_slots) (void)hipFree(d_slots);
    if (d_ctl) (void)hipFree(d_ctl);
    if (h_slots) (void)hipHostFree(h_slots);
    if (h_ctl_up) (void)hipHostFree(h_ctl_up);
    if (h_ctl_down) (void)hipHostFree(h_ctl_down);
    plan_free(plan);
  }
  int dim = 0, res[3] = {1, 1, 1};
  hipStream_t sb = nullptr, sr = nullptr;
  FftPlan plan;
  int64_t N = 0;
  BatchSlot* d_slots = nullptr;
  BatchSlot* h_slots = nullptr;  // pinned staging
  int* d_ctl = nullptr;
  int* h_ctl_up = nullptr;    // pinned
  int* h_ctl_down = nullptr;  // pinned
  std::vector<Request*> slot_req;
  std::deque<Request*> waiting;
  std::vector<Request*> eval_ready;  // at the material stage, not yet launched
  int nlaunch = 0;  // slots in the chunk in flight
  bool in_flight = false;
  hipEvent_t ev = nullptr;
  std::map<int, hipGraphExec_t> graphs;
};

// pinned host words for requests, handed out from a growing pool
class PinnedPool {
 public:
  ~PinnedPool() {
    for (double* b : blocks_) (void)hipHostFree(b);
  }
  double* get() {
    if (free_.empty()) {
      double* b = nullptr;
      if (hipHostMalloc((void**)&b, sizeof(double) * kHostDoubles * 256, hipHostMallocDefault) !=
          hipSuccess)
        return nullptr;
      blocks_.push_back(b);
      for (int i = 0; i < 256; ++i) free_.push_back(b + i * kHostDoubles);
    }
    double* p = free_.back();
    free_.pop_back();
    return p;
  }
  void put(double* p) { free_.push_back(p); }

 private:
  std::vector<double*> blocks_, free_;
};

class EventPool {
 public:
  ~EventPool() {
    for (hipEvent_t e : all_) (void)hipEventDestroy(e);
  }
  hipEvent_t get() {
    if (free_.empty()) {
      hipEvent_t e;
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
      all_.push_back(e);
      return e;
    }
    hipEvent_t e = free_.back();
    free_.pop_back();
    return e;
  }
  void put(hipEvent_t e) { free_.push_back(e); }

 private:
  std::vector<hipEvent_t> all_, free_;
};

}  // namespace

struct mugpu_server_s {
  mugpu_server_config cfg{};
  Arena arena;
  PinnedPool pinned;
  EventPool events;
  std::map<std::vector<int>, std::unique_ptr<ResClass>> classes;
  std::deque<Request*> queue;      // submitted, not admitted
  std::vector<Request*> running;   // admitted, not done
  std::map<int64_t, std::unique_ptr<Request>> all;
  std::vector<int64_t> done_ids;
  int64_t next_id = 1;
  int64_t stats[9] = {0};

  ~mugpu_server_s();
  int class_of(Request* r);
  int admit();
  int issue(Request* r);  // queue the next device stage of r (r->st tells which)
  int advance(Request* r);
  int start_eval(Request* r);
  int flush_evals(ResClass* c);
  int project_rhs(Request* r);
  int start_resid(Request* r);
  int finish(Request* r);
  int after_inner(Request* r, bool ran);
  int residual_known(Request* r);
  int launch_chunk(ResClass* c);
  int process_chunk(ResClass* c);
  int graph_for(ResClass* c, int cap, hipGraphExec_t* out);
  int record(Request* r) {
    hipStream_t sr = r->cls->sr;
    MUGPU_HIP(hipMemcpyAsync(r->h, r->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, sr));
    MUGPU_HIP(hipEventRecord(r->ev, sr));
    return MUGPU_OK;
  }
};

mugpu_server_s::~mugpu_server_s() {
  (void)hipDeviceSynchronize();
  classes.clear();
}

int mugpu_server_s::class_of(Request* r) {
  std::vector<int> key(r->res, r->res + r->dim);
  auto it = classes.find(key);
  if (it != classes.end()) {
    r->cls = it->second.get();
    return MUGPU_OK;
  }
  auto c = std::make_unique<ResClass>();
  c->dim = r->dim;
  for (int d = 0; d < 3; ++d) c->res[d] = r->res[d];
  c->N = r->N;
  MUGPU_TRY(plan_init(c->plan, r->dim, r->res, r->nc, MUGPU_FP32));
  const int cap = cfg.max_batch;
  MUGPU_TRY(dev_alloc(&c->d_slots, sizeof(BatchSlot) * cap, "the batch slots"));
  MUGPU_TRY(dev_alloc(&c->d_ctl, sizeof(int) * CTL_COUNT * cap, "the batch control words"));
  MUGPU_HIP(hipHostMalloc((void**)&c->h_slots, sizeof(BatchSlot) * cap, hipHostMallocDefault));
  MUGPU_HIP(hipHostMalloc((void**)&c->h_ctl_up, sizeof(int) * CTL_COUNT * cap, hipHostMallocDefault));
  MUGPU_HIP(hipHostMalloc((void**)&c->h_ctl_down, sizeof(int) * CTL_COUNT * cap, hipHostMallocDefault));
  // every slot's control words exist from the start (inactive), so a launch
  // over more slots than are in use reads valid flags
  std::memset(c->h_slots, 0, sizeof(BatchSlot) * cap);
  std::memset(c->h_ctl_up, 0, sizeof(int) * CTL_COUNT * cap);
  for (int i = 0; i < cap; ++i) c->h_slots[i].ctl = c->d_ctl + CTL_COUNT * i;
  MUGPU_HIP(hipMemcpy(c->d_slots, c->h_slots, sizeof(BatchSlot) * cap, hipMemcpyHostToDevice));
  MUGPU_HIP(hipMemcpy(c->d_ctl, c->h_ctl_up, sizeof(int) * CTL_COUNT * cap, hipMemcpyHostToDevice));
  MUGPU_HIP(hipEventCreateWithFlags(&c->ev, hipEventDisableTiming));
  MUGPU_HIP(hipStreamCreateWithFlags(&c->sb, hipStreamNonBlocking));
  MUGPU_HIP(hipStreamCreateWithFlags(&c->sr, hipStreamNonBlocking));
  r->cls = c.get();
  classes.emplace(key, std::move(c));
  return MUGPU_OK;
}

// first come, first served while the head request's footprint fits
int mugpu_server_s::admit() {
  while (!queue.empty()) {
    Request* r = queue.front();
    size_t off = 0;
    if (!arena.alloc(r->lay.total, &off)) break;
    // its pinned words, its event and its resolution class; when one of them
    // cannot be had, the request stays at the head of the queue holding
    // nothing, and the error is returned
    double* h = pinned.get();
    hipEvent_t ev = h ? events.get() : nullptr;
    int rc = MUGPU_OK;
    if (!h || !ev) {
      set_error("pinned host memory or events exhausted");
      rc = MUGPU_EMEMORY;
    } else {
      rc = class_of(r);
    }
    if (rc != MUGPU_OK) {
      if (h) pinned.put(h);
      if (ev) events.put(ev);
      arena.release(off, r->lay.total);
      return rc;
    }
    queue.pop_front();
    r->h = h;
    r->ev = ev;
    r->off = off;
    char* b = arena.base() + off;
    r->F = (double*)(b + r->lay.F);
    r->P = (double*)(b + r->lay.P);
    r->rhs = (double*)(b + r->lay.rhs);
    r->x = (double*)(b + r->lay.x);
    r->r64 = (double*)(b + r->lay.r64);
    r->params = (double*)(b + r->lay.params);
    r->tan64 = (double*)(b + r->lay.tan64);
    r->tan32 = (float*)(b + r->lay.tan32);
    r->x32 = (float*)(b + r->lay.x32);
    r->r32 = (float*)(b + r->lay.r32);
    r->p32 = (float*)(b + r->lay.p32);
    r->Ap32 = (float*)(b + r->lay.Ap32);
    r->spec = (void*)(b + r->lay.spec);
    r->scal = (double*)(b + r->lay.scal);
    r->part = (double*)(b + r->lay.part);
    r->zpart = (double*)(b + r->lay.zpart);
    r->colpart = (double*)(b + r->lay.colpart);
    hipStream_t sr = r->cls->sr;
    running.push_back(r);
    ++stats[1];
    stats[8] = std::max<int64_t>(stats[8], (int64_t)running.size());
    // cell set-up: per-voxel (lambda, mu), SC block, F = 0 + dE
    // (solvers.cc:96-145: set_uniform_strain(0), add_uniform_strain(dE))
    set_current_stream(sr);
    MUGPU_HIP(hipMemcpyAsync(r->params, r->lam_mu.get(), sizeof(double) * 2 * r->N,
                             hipMemcpyHostToDevice, sr));
    MUGPU_HIP(hipMemsetAsync(r->scal, 0, sizeof(double) * SC_COUNT, sr));
    const double zero[9] = {0};
    MUGPU_TRY(fill_uniform(r->F, zero, r->nc, r->N, 0));
    MUGPU_TRY(fill_uniform(r->F, r->strain, r->nc, r->N, 1));
    // the host copy of the parameters is kept until the first stage's event
    // (recorded after the transfer on the same stream) has completed
    r->incr_norm = 2 * cfg.newton_tol;
    r->grad_norm = 1.0;
    r->stress_norm = 2 * cfg.equil_tol;
    r->newt_iter = 0;
    MUGPU_TRY(start_eval(r));
  }
  return MUGPU_OK;
}

// one Newton iteration begins (solvers.cc:165-176): the request waits for
// its class's batched material stage, which flush_evals launches once per
// progress pass for every request that reached it
int mugpu_server_s::start_eval(Request* r) {
  const int newton_bound = r->maxiter;
  if (r->newt_iter >= newton_bound || r->has_converged) return finish(r);
  r->st = ST_EVAL_QUEUED;
  r->cls->eval_ready.push_back(r);
  return MUGPU_OK;
}

// stress and tangent of every request of c at the material stage, in
// launches of up to kServeMatSlots requests (the same per-voxel arithmetic
// as eval_material followed by the FP32 tangent copy and rhs = -P), then each
// request's rhs = -G:P and |rhs|^2.  A request leaves eval_ready once its
// stage is queued; after an error the rest stay there.
int mugpu_server_s::flush_evals(ResClass* c) {
  if (c->eval_ready.empty()) return MUGPU_OK;
  set_current_stream(c->sr);
  const size_t total = c->eval_ready.size();
  for (size_t i0 = 0; i0 < total; i0 += kServeMatSlots) {
    const int n = (int)std::min<size_t>(kServeMatSlots, total - i0);
    ServeMatSlot sl[kServeMatSlots];
    for (int i = 0; i < n; ++i) {
      const Request* r = c->eval_ready[i0 + i];
      sl[i] = ServeMatSlot{r->F, r->params, r->P, r->rhs, r->tan64, r->tan32};
    }
    MUGPU_TRY(serve_material_batch(c->dim, sl, n, c->N));
  }
  size_t done = 0;
  int rc = MUGPU_OK;
  for (; done < total && rc == MUGPU_OK; ++done) rc = project_rhs(c->eval_ready[done]);
  if (rc != MUGPU_OK) --done;
  c->eval_ready.erase(c->eval_ready.begin(), c->eval_ready.begin() + (std::ptrdiff_t)done);
  return rc;
}

// rhs = -G:P in FP64 (P negated by the material stage) and |rhs|^2
int mugpu_server_s::project_rhs(Request* r) {
  set_current_stream(r->cls->sr);
  OpSpec op;
  op.in_mode = IN_PLAIN;
  op.in = r->rhs;
  op.out_mode = OUT_PLAIN;
  op.out = r->rhs;
  ProjParams pj;
  pj.kind = MUGPU_PROJ_SMALL_STRAIN;
  pj.allow_even = 1;
  for (int d = 0; d < 3; ++d) pj.L[d] = r->L[d];
  pj.norm = 1.0 / (double)r->N;
  MUGPU_TRY(run_operator(r->cls->plan, MUGPU_FP64, op, pj, r->spec));
  MUGPU_TRY(dot_device(MUGPU_FP64, r->rhs, r->rhs, r->nd, r->part, r->scal, FIN_STORE_TMP0));
  r->st = ST_WAIT_EVAL;
  return record(r);
}

// r64 = b - A x in FP64, |r64|^2 (the refinement step of the mixed solve)
int mugpu_server_s::start_resid(Request* r) {
  hipStream_t sr = r->cls->sr;
  set_current_stream(sr);
  OpSpec op;
  op.in_mode = IN_TAN;
  op.in = r->x;
  op.tan_kind = TAN_ISO;
  op.tangent = r->tan64;
  op.out_mode = OUT_PLAIN;
  op.out = r->r64;
  ProjParams pj;
  pj.kind = MUGPU_PROJ_SMALL_STRAIN;
  pj.allow_even = 1;
  for (int d = 0; d < 3; ++d) pj.L[d] = r->L[d];
  pj.norm = 1.0 / (double)r->N;
  MUGPU_TRY(run_operator(r->cls->plan, MUGPU_FP64, op, pj, r->spec));
  MUGPU_TRY(axpby(MUGPU_FP64, 1.0, r->rhs, -1.0, r->r64, r->nd));
  MUGPU_TRY(dot_device(MUGPU_FP64, r->r64, r->r64, r->nd, r->part, r->scal, FIN_STORE_TMP1));
  r->st = ST_WAIT_RESID;
  return record(r);
}

// x += |r64| d after an inner solve; when it ran in the batch, the x update
// of its last iteration (alpha and p of that iteration, left pending by the
// batched r2c) is folded in, as in the single-cell mixed solve
int mugpu_server_s::after_inner(Request* r, bool ran) {
  hipStream_t sr = r->cls->sr;
  set_current_stream(sr);
  MUGPU_TRY(add_scaled_convert(r->x, r->x32, 1.0 / r->s, r->nd, ran ? r->p32 : nullptr,
                               ran ? r->scal : nullptr));
  return start_resid(r);
}

int mugpu_server_s::finish(Request* r) {
  hipStream_t sr = r->cls->sr;
  set_current_stream(sr);
  r->converged = r->test(cfg.newton_tol, cfg.equil_tol);
  const int g = (int)std::min<int64_t>(kColBlocks, (r->N + kColT - 1) / kColT);
  hipLaunchKernelGGL(k_colsum, dim3(g), dim3(kColT), 0, sr, r->P, r->N, r->nc, r->colpart);
  hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(64), 0, sr, r->colpart, g, r->nc, r->part);
  MUGPU_HIP(hipGetLastError());
  MUGPU_HIP(hipMemcpyAsync(r->h + kHostColsum, r->part, sizeof(double) * r->nc,
                           hipMemcpyDeviceToHost, sr));
  if (cfg.keep_fields) {
    r->hF.resize((size_t)r->nd);
    r->hP.resize((size_t)r->nd);
    MUGPU_HIP(hipMemcpyAsync(r->hF.data(), r->F, sizeof(double) * r->nd, hipMemcpyDeviceToHost, sr));
    MUGPU_HIP(hipMemcpyAsync(r->hP.data(), r->P, sizeof(double) * r->nd, hipMemcpyDeviceToHost, sr));
  }
  MUGPU_HIP(hipEventRecord(r->ev, sr));
  r->st = ST_WAIT_RESULT;
  return MUGPU_OK;
}

// the refinement step with the residual |r64|^2 = r->rr known (the loop of
// the mixed-precision solve: stop, or start the next inner CG)
int mugpu_server_s::residual_known(Request* r) {
  hipStream_t sr = r->cls->sr;
  ++r->refs;
  if (!(r->rr > r->tol2) || r->its >= r->maxiter || !(r->rr == r->rr)) {
    r->counter += r->its;
    if (r->rr > r->tol2 || !(r->rr == r->rr)) {
      // ConvergenceError: the request fails alone
      r->status = MUGPU_ECONVERGENCE;
      ++stats[3];
      return finish(r);
    }
    // F += dF, |dF|, |F| (solvers.cc:185-200)
    set_current_stream(sr);
    MUGPU_TRY(axpby(MUGPU_FP64, 1.0, r->x, 1.0, r->F, r->nd));
    MUGPU_TRY(dot_device(MUGPU_FP64, r->x, r->x, r->nd, r->part, r->scal, FIN_STORE_TMP0));
    MUGPU_TRY(dot_device(MUGPU_FP64, r->F, r->F, r->nd, r->part, r->scal, FIN_STORE_TMP1));
    r->st = ST_WAIT_NORMS;
    return record(r);
  }
  r->s = 1.0 / std::sqrt(r->rr);
  set_current_stream(sr);
  MUGPU_HIP(hipMemsetAsync(r->x32, 0, sizeof(float) * r->nd, sr));
  MUGPU_TRY(convert(MUGPU_FP64, r->r64, MUGPU_FP32, r->r32, r->nd, -r->s));
  MUGPU_TRY(convert(MUGPU_FP64, r->r64, MUGPU_FP32, r->p32, r->nd, r->s));
  MUGPU_TRY(dot_device(MUGPU_FP32, r->r32, r->r32, r->nd, r->part, r->scal, FIN_STORE_RDR));
  r->st = ST_WAIT_RDR;
  return record(r);
}

// the device work of r's current stage has completed: take the next step
int mugpu_server_s::advance(Request* r) {
  hipStream_t sr = r->cls->sr;
  const double* h = r->h;
  const double tol = cfg.cg_tol;
  switch (r->st) {
    case ST_WAIT_EVAL: {
      r->lam_mu.reset();
      r->stress_norm = std::sqrt(h[SC_TMP0]);
      if (r->test(cfg.newton_tol, cfg.equil_tol)) return finish(r);
      // SolverCG::solve, mixed precision: x = 0, |b|^2, then refinement
      set_current_stream(sr);
      MUGPU_HIP(hipMemsetAsync(r->x, 0, sizeof(double) * r->nd, sr));
      MUGPU_TRY(dot_device(MUGPU_FP64, r->rhs, r->rhs, r->nd, r->part, r->scal, FIN_STORE_TMP1));
      r->st = ST_WAIT_BNORM;
      return record(r);
    }
    case ST_WAIT_BNORM:
      r->bnorm2 = h[SC_TMP1];
      r->its = 0;
      r->refs = 0;
      r->tol2 = tol * tol * r->bnorm2;
      // first residual at x = 0: A x = 0 exactly, so r64 = b and
      // |r64|^2 = |b|^2 bit for bit (no operator application)
      set_current_stream(sr);
      MUGPU_TRY(convert(MUGPU_FP64, r->rhs, MUGPU_FP64, r->r64, r->nd, 1.0));
      r->rr = r->bnorm2;
      return residual_known(r);
    case ST_WAIT_RESID:
      r->rr = h[SC_TMP1];
      return residual_known(r);
    case ST_WAIT_RDR: {
      const double rdr = h[SC_RDR];
      const double tol2i = inner_threshold2(cfg.inner_tol, r->tol2, r->rr, rdr);
      if (r->its >= r->maxiter) return after_inner(r, false);  // no inner iteration
      r->h[kHostTol2] = tol2i;
      r->inner_maxit = (int)(r->maxiter - r->its);
      r->st = ST_BATCH_QUEUE;
      r->cls->waiting.push_back(r);
      return MUGPU_OK;
    }
    case ST_WAIT_NORMS:
      r->incr_norm = std::sqrt(h[SC_TMP0]);
      r->grad_norm = std::sqrt(h[SC_TMP1]);
      r->test(cfg.newton_tol, cfg.equil_tol);
      ++r->newt_iter;
      return start_eval(r);
    case ST_WAIT_RESULT:
      for (int c = 0; c < r->nc; ++c) r->mean_stress[c] = r->h[kHostColsum + c] / (double)r->N;
      arena.release(r->off, r->lay.total);
      pinned.put(r->h);
      events.put(r->ev);
      r->h = nullptr;
      r->ev = nullptr;
      r->st = ST_DONE;
      ++stats[2];
      done_ids.push_back(r->id);
      return MUGPU_OK;
    default:
      set_error("request %lld in an unexpected stage", (long long)r->id);
      return MUGPU_ERUNTIME;
  }
}

int mugpu_server_s::graph_for(ResClass* c, int cap, hipGraphExec_t* out) {
  hipStream_t sb = c->sb;
  auto it = c->graphs.find(cap);
  if (it != c->graphs.end()) {
    *out = it->second;
    return MUGPU_OK;
  }
  const int K = cfg.iters_per_launch;
  hipGraph_t g = nullptr;
  set_current_stream(sb);
  MUGPU_HIP(hipStreamBeginCapture(sb, hipStreamCaptureModeThreadLocal));
  set_capturing(true);
  int rc = MUGPU_OK;
  for (int k = 0; k < K && rc == MUGPU_OK; ++k) rc = batch_cg_iteration(c->plan, c->d_slots, cap);
  set_capturing(false);
  const hipError_t e = hipStreamEndCapture(sb, &g);
  if (rc != MUGPU_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  MUGPU_HIP(e);
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MUGPU_HIP(ei);
  c->graphs[cap] = ex;
  *out = ex;
  return MUGPU_OK;
}

// pack the running slots, fill free ones from the queue, launch K iterations
int mugpu_server_s::launch_chunk(ResClass* c) {
  hipStream_t sb = c->sb;
  std::vector<Request*> act(c->slot_req.begin(), c->slot_req.end());
  while ((int)act.size() < cfg.max_batch && !c->waiting.empty()) {
    Request* r = c->waiting.front();
    c->waiting.pop_front();
    // a joining slot starts its inner CG at iteration 0 with its threshold
    r->st = ST_BATCH;
    r->inner_iter = 0;
    MUGPU_HIP(hipMemcpyAsync(r->scal + SC_TOL2, r->h + kHostTol2, sizeof(double),
                             hipMemcpyHostToDevice, sb));
    act.push_back(r);
  }
  const int n = (int)act.size();
  if (n == 0) return MUGPU_OK;
  int cap = 1;
  while (cap < n) cap *= 2;
  c->slot_req = act;
  for (int i = 0; i < n; ++i) {
    Request* r = act[i];
    BatchSlot& sl = c->h_slots[i];
    sl.tan = r->tan32;
    sl.x = r->x32;
    sl.r = r->r32;
    sl.p = r->p32;
    sl.Ap = r->Ap32;
    sl.spec = (float2*)r->spec;
    sl.scal = r->scal;
    sl.zpart = r->zpart;
    sl.upart = r->part;
    sl.ctl = c->d_ctl + CTL_COUNT * i;
    sl.tan_kind = TAN_ISO;
    sl.kind = MUGPU_PROJ_SMALL_STRAIN;
    sl.allow_even = 1;
    for (int d = 0; d < 3; ++d) sl.L[d] = r->L[d];
    sl.norm = 1.0 / (double)r->N;
  }
  // control words of the launched slots: packed requests carry their own
  // iteration counts, slots beyond them are inactive
  for (int i = 0; i < cap; ++i) {
    int* ctl = c->h_ctl_up + CTL_COUNT * i;
    ctl[CTL_ACTIVE] = i < n ? 1 : 0;
    ctl[CTL_ITER] = i < n ? act[i]->inner_iter : 0;
    ctl[CTL_MAXIT] = i < n ? act[i]->inner_maxit : 0;
  }
  MUGPU_HIP(hipMemcpyAsync(c->d_slots, c->h_slots, sizeof(BatchSlot) * n, hipMemcpyHostToDevice, sb));
  MUGPU_HIP(hipMemcpyAsync(c->d_ctl, c->h_ctl_up, sizeof(int) * CTL_COUNT * cap,
                           hipMemcpyHostToDevice, sb));
  if (cfg.use_graphs) {
    hipGraphExec_t ex = nullptr;
    MUGPU_TRY(graph_for(c, cap, &ex));
    MUGPU_HIP(hipGraphLaunch(ex, sb));
  } else {
    set_current_stream(sb);
    for (int k = 0; k < cfg.iters_per_launch; ++k)
      MUGPU_TRY(batch_cg_iteration(c->plan, c->d_slots, cap));
  }
  MUGPU_HIP(hipMemcpyAsync(c->h_ctl_down, c->d_ctl, sizeof(int) * CTL_COUNT * n,
                           hipMemcpyDeviceToHost, sb));
  MUGPU_HIP(hipEventRecord(c->ev, sb));
  c->nlaunch = n;
  c->in_flight = true;
  ++stats[4];
  stats[5] += n;
  return MUGPU_OK;
}

int mugpu_server_s::process_chunk(ResClass* c) {
  c->in_flight = false;
  std::vector<Request*> keep;
  for (int i = 0; i < c->nlaunch; ++i) {
    Request* r = c->slot_req[i];
    const int* ctl = c->h_ctl_down + CTL_COUNT * i;
    r->inner_iter = ctl[CTL_ITER];
    if (ctl[CTL_ACTIVE]) {
      keep.push_back(r);
      continue;
    }
    // inner CG done: its iterations join the solve's count
    r->its += r->inner_iter;
    MUGPU_TRY(after_inner(r, true));
  }
  c->slot_req = keep;
  return MUGPU_OK;
}

// ---- C-ABI ------------------------------------------------------------------------
extern "C" int mugpu_server_footprint(int dim, const int* res, size_t* bytes) {
  if ((dim != 2 && dim != 3) || !res || !bytes) return MUGPU_EARG;
  for (int d = 0; d < dim; ++d)
    if (!fast_size(res[d])) {
      set_error("serving needs axis lengths in {16, 32, 48, 64, 96, 128, 256, 512}, got %d", res[d]);
      return MUGPU_EARG;
    }
  *bytes = layout_of(dim, res).total;
  return MUGPU_OK;
}

extern "C" int mugpu_server_create(const mugpu_server_config* cfg, mugpu_server* out) {
  if (!cfg || !out) return MUGPU_EARG;
  *out = nullptr;
  MUGPU_TRY(mugpu_init(0));
  if (cfg->max_batch < 1 || cfg->max_batch > kMaxSlots || cfg->iters_per_launch < 1 ||
      cfg->cg_maxiter < 1) {
    set_error("server config: max_batch in 1..%d, iters_per_launch >= 1, cg_maxiter >= 1", kMaxSlots);
    return MUGPU_EARG;
  }
  auto s = std::make_unique<mugpu_server_s>();
  s->cfg = *cfg;
  // field blocks cached by mugpu_free go back to the device first: the
  // arena is sized from (and taken out of) free memory
  MUGPU_TRY(mugpu_release_cached());
  size_t budget = cfg->hbm_budget;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    MUGPU_HIP(hipMemGetInfo(&fr, &tot));
    budget = fr / 4 * 3;
  }
  MUGPU_TRY(s->arena.init(budget));
  s->stats[7] = (int64_t)budget;
  *out = s.release();
  return MUGPU_OK;
}

extern "C" int mugpu_server_destroy(mugpu_server s) {
  if (!s) return MUGPU_OK;
  set_current_stream(nullptr);
  delete s;
  return MUGPU_OK;
}

extern "C" int mugpu_server_submit(mugpu_server s, int dim, const int* res, const double* lengths,
                                   const double* young, const double* poisson, const double* strain,
                                   int cg_maxiter, int64_t* id) {
  if (!s || !res || !young || !poisson || !strain || !id) return MUGPU_EARG;
  size_t bytes = 0;
  MUGPU_TRY(mugpu_server_footprint(dim, res, &bytes));
  if (bytes > s->arena.size()) {
    set_error("a request of %zu bytes exceeds the serving budget of %zu bytes", bytes, s->arena.size());
    return MUGPU_EMEMORY;
  }
  // check_symmetry of the small-strain load (solver_common.cc:39-43)
  double nrm = 0, asym = 0;
  for (int j = 0; j < dim; ++j)
    for (int i = 0; i < dim; ++i) {
      nrm += strain[i + dim * j] * strain[i + dim * j];
      const double a = strain[i + dim * j] - strain[j + dim * i];
      asym += a * a;
    }
  if (std::sqrt(asym) > 1e-8 * std::sqrt(nrm)) {
    set_error("all macroscopic strain tensors must be symmetric");
    return MUGPU_ESOLVER;
  }
  auto r = std::make_unique<Request>();
  r->id = s->next_id++;
  r->dim = dim;
  r->nc = dim * dim;
  r->N = 1;
  for (int d = 0; d < dim; ++d) {
    r->res[d] = res[d];
    r->L[d] = lengths ? lengths[d] : 1.0;
    r->N *= res[d];
  }
  r->nd = r->N * r->nc;
  r->lam_mu.reset(new (std::nothrow) double[(size_t)(2 * r->N)]);
  if (!r->lam_mu) {
    set_error("host allocation of a request's material parameters (%lld voxels) failed",
              (long long)r->N);
    return MUGPU_EMEMORY;
  }
  fill_lame(young, poisson, r->N, r->lam_mu.get());
  for (int c = 0; c < r->nc; ++c) r->strain[c] = strain[c];
  r->maxiter = cg_maxiter > 0 ? cg_maxiter : s->cfg.cg_maxiter;
  r->lay = layout_of(dim, res);
  *id = r->id;
  s->queue.push_back(r.get());
  s->all.emplace(r->id, std::move(r));
  ++s->stats[0];
  return MUGPU_OK;
}

extern "C" int mugpu_server_progress(mugpu_server s, int wait, int* unfinished) {
  if (!s) return MUGPU_EARG;
  // the stages queue their work on the streams of their resolution class; on
  // every return, an error's included, the library's own stream is current
  // again, so that later calls outside the server do not land on them
  struct StreamReset {
    ~StreamReset() { set_current_stream(nullptr); }
  } reset;
  MUGPU_TRY(s->admit());
  bool moved = false;
  // requests whose stage completed
  for (size_t i = 0; i < s->running.size(); ++i) {
    Request* r = s->running[i];
    if (r->st == ST_BATCH_QUEUE || r->st == ST_BATCH || r->st == ST_DONE ||
        r->st == ST_EVAL_QUEUED)
      continue;
    const hipError_t q = hipEventQuery(r->ev);
    if (q == hipErrorNotReady) continue;
    MUGPU_HIP(q);
    MUGPU_TRY(s->advance(r));
    moved = true;
  }
  // batches whose chunk completed, then new chunks
  for (auto& kv : s->classes) {
    ResClass* c = kv.second.get();
    if (c->in_flight) {
      const hipError_t q = hipEventQuery(c->ev);
      if (q == hipErrorNotReady) continue;
      MUGPU_HIP(q);
      MUGPU_TRY(s->process_chunk(c));
      moved = true;
    }
    if (!c->in_flight && (!c->slot_req.empty() || !c->waiting.empty())) {
      MUGPU_TRY(s->launch_chunk(c));
      moved = true;
    }
  }
  s->running.erase(std::remove_if(s->running.begin(), s->running.end(),
                                  [](Request* r) { return r->st == ST_DONE; }),
                   s->running.end());
  MUGPU_TRY(s->admit());
  // the material stages reached in this pass, batched per resolution
  for (auto& kv : s->classes) {
    if (kv.second->eval_ready.empty()) continue;
    MUGPU_TRY(s->flush_evals(kv.second.get()));
    moved = true;
  }
  if (wait && !moved) {
    // block on the oldest outstanding device work
    for (Request* r : s->running)
      if (r->st != ST_BATCH_QUEUE && r->st != ST_BATCH && r->ev) {
        MUGPU_HIP(hipEventSynchronize(r->ev));
        break;
      }
  }
  if (unfinished) *unfinished = (int)(s->queue.size() + s->running.size());
  return MUGPU_OK;
}

extern "C" int mugpu_server_poll(mugpu_server s, int64_t* ids, int capacity, int* count) {
  if (!s || !count || capacity < 0 || (capacity > 0 && !ids)) return MUGPU_EARG;
  const int n = std::min<int>(capacity, (int)s->done_ids.size());
  for (int i = 0; i < n; ++i) ids[i] = s->done_ids[i];
  s->done_ids.erase(s->done_ids.begin(), s->done_ids.begin() + n);
  *count = n;
  return MUGPU_OK;
}

extern "C" int mugpu_server_result(mugpu_server s, int64_t id, int* status, int* converged,
                                   int* newton_iterations, int64_t* cg_iterations,
                                   double* mean_stress, double* strain, double* stress) {
  if (!s) return MUGPU_EARG;
  auto it = s->all.find(id);
  if (it == s->all.end() || it->second->st != ST_DONE) {
    set_error("request %lld is unknown or not finished", (long long)id);
    return MUGPU_EARG;
  }
  const Request* r = it->second.get();
  if (status) *status = r->status;
  if (converged) *converged = r->converged ? 1 : 0;
  if (newton_iterations) *newton_iterations = r->newt_iter;
  if (cg_iterations) *cg_iterations = r->counter;
  if (mean_stress)
    for (int c = 0; c < r->nc; ++c) mean_stress[c] = r->mean_stress[c];
  if (strain && !r->hF.empty()) std::memcpy(strain, r->hF.data(), sizeof(double) * r->hF.size());
  if (stress && !r->hP.empty()) std::memcpy(stress, r->hP.data(), sizeof(double) * r->hP.size());
  return MUGPU_OK;
}

extern "C" int mugpu_server_release(mugpu_server s, int64_t id) {
  if (!s) return MUGPU_EARG;
  auto it = s->all.find(id);
  if (it == s->all.end() || it->second->st != ST_DONE) return MUGPU_EARG;
  s->all.erase(it);
  return MUGPU_OK;
}

extern "C" int mugpu_server_stats(mugpu_server s, int64_t* out, int n) {
  if (!s || !out) return MUGPU_EARG;
  s->stats[6] = (int64_t)s->arena.peak();
  for (int i = 0; i < n && i < 9; ++i) out[i] = s->stats[i];
  return MUGPU_OK;
}
