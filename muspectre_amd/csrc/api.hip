// C-ABI of libmugpu.so (declared in include/mugpu.h).  Handles wrap device
// buffers; every function reports errors through return codes and
// mugpu_last_error().  Work runs on one library stream.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <atomic>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "mugpu_internal.hpp"

namespace mugpu {

namespace {
// the message of the calling thread's last failed call
thread_local char g_err[2048] = "";
hipStream_t g_stream = nullptr;
int g_device = -1;
int g_maxlds = 65536;
double* g_host_scal = nullptr;  // pinned
int* g_host_ctl = nullptr;       // pinned, CTL_COUNT words after the scalars
double* g_dev_scratch = nullptr;  // partials + scalars for mugpu_dot
}  // namespace

// Work memory of mugpu_cg_solve_mixed_noalloc: a fixed array in device
// memory, one copy per GPU the library runs on, held until the process ends.
// Host code never uses the array's name as an address (on the host it is
// only a stand-in for the symbol): mugpu_init looks up its device address
// once per device (hipGetSymbolAddress) and every use goes through that.
static __device__ __align__(256) unsigned char scratch[MUGPU_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  return MUGPU_EHIP;
}

// the stream the calling thread's launches go to (null: the library's
// stream) and whether that thread is capturing a graph
thread_local hipStream_t g_cur_stream = nullptr;
hipStream_t stream() { return g_cur_stream ? g_cur_stream : g_stream; }
void set_current_stream(hipStream_t s) { g_cur_stream = s; }
thread_local bool g_capturing = false;
void set_capturing(bool on) { g_capturing = on; }
int max_lds_per_block() { return g_maxlds; }

namespace {
// graph replay of mugpu_cg_steps chunks on cells of at most this many values
constexpr int64_t kCgGraphMaxValues = int64_t(1) << 20;
constexpr size_t kCgGraphsPerCell = 16;
bool g_cg_graphs = true;
int g_timing_class = -1;
std::vector<hipEvent_t> g_ev_start, g_ev_stop;
int64_t g_timing_count = 0;
constexpr int64_t kMaxTimed = 16384;
}  // namespace

std::vector<int> g_ev_class;
int64_t g_stat_launches[MUGPU_K_ALL] = {0};
double g_stat_ms[MUGPU_K_ALL] = {0};

void timing_begin(int kclass) {
  if ((kclass != g_timing_class && g_timing_class != MUGPU_K_ALL) || g_timing_count >= kMaxTimed ||
      g_capturing)
    return;
  if ((int64_t)g_ev_start.size() <= g_timing_count) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    g_ev_start.push_back(a);
    g_ev_stop.push_back(b);
    g_ev_class.push_back(-1);
  }
  g_ev_class[g_timing_count] = kclass;
  (void)hipEventRecord(g_ev_start[g_timing_count], stream());
}

void timing_end(int kclass) {
  if ((kclass != g_timing_class && g_timing_class != MUGPU_K_ALL) || g_timing_count >= kMaxTimed ||
      g_capturing)
    return;
  if ((int64_t)g_ev_stop.size() <= g_timing_count) return;
  (void)hipEventRecord(g_ev_stop[g_timing_count], stream());
  ++g_timing_count;
}

int64_t timing_mark() { return g_timing_count; }

void timing_void(int64_t from) {
  for (int64_t i = from < 0 ? 0 : from; i < g_timing_count; ++i) g_ev_class[i] = -2;
}

int copy_device(void* dst, const void* src, size_t bytes);
int copy_device_wide(void* dst, const void* src, size_t bytes, bool nontemporal, int grid);
int read_device_wide(const void* src, size_t bytes, void* sink, int grid);
int fill_device_wide(void* dst, size_t bytes, int grid);

int run_tangent_product(int dim, int prec, int kind, const void* tan, const void* in, void* out,
                        int64_t npix);

}  // namespace mugpu

using namespace mugpu;

struct mugpu_fft_s {
  FftPlan plan;
};

struct mugpu_proj_s {
  mugpu_fft engine = nullptr;
  ProjParams pj;
  void* spectrum = nullptr;
};

struct mugpu_cell_s {
  int dim = 0, form = 0, prec = 0, allow_even = 0, nc = 0, initialised = 0;
  int res[3] = {1, 1, 1};
  double L[3] = {1, 1, 1};
  int64_t N = 0;
  FftPlan plan;
  ProjParams pj;
  double* F = nullptr;
  double* P = nullptr;
  int tan_kind = TAN_FULL;
  double* tan64 = nullptr;
  float* tan32 = nullptr;
  void* spectrum = nullptr;
  double* tmp = nullptr;
  double* scal = nullptr;
  double* partials = nullptr;
  std::vector<MaterialDev> mats;
  float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *Ap32 = nullptr;
  double* r64 = nullptr;
  int* ctl = nullptr;  // CTL_* words of the device-resident inner CG loop
  // mugpu_cg_steps left the x update of its last iteration to be applied
  // (to x_pend with p_pend) by the next call or by mugpu_cg_finish
  bool x_pending = false;
  void *x_pend = nullptr, *p_pend = nullptr;
  // captured chunks of mugpu_cg_steps (small cells), by vectors, chunk
  // length and form of the first iteration
  struct CgGraph {
    void *x, *r, *p, *Ap;
    int steps, first_update_p, first_x;
    hipGraphExec_t exec;
  };
  std::vector<CgGraph> graphs;
};

// Cells with a pending x update.  mugpu_free consults them: freeing the p of
// a pending update applies it first (p is still valid in stream order),
// freeing its x drops it, so a later call never writes to a block that the
// field cache has handed to someone else.
namespace {
std::mutex g_pend_mu;
std::vector<mugpu_cell> g_pending_cells;

void pending_set(mugpu_cell c, void* x, void* p) {
  std::lock_guard<std::mutex> lk(g_pend_mu);
  c->x_pending = true;
  c->x_pend = x;
  c->p_pend = p;
  for (mugpu_cell q : g_pending_cells)
    if (q == c) return;
  g_pending_cells.push_back(c);
}

void pending_clear(mugpu_cell c) {
  std::lock_guard<std::mutex> lk(g_pend_mu);
  c->x_pending = false;
  for (size_t i = 0; i < g_pending_cells.size(); ++i)
    if (g_pending_cells[i] == c) {
      g_pending_cells.erase(g_pending_cells.begin() + (std::ptrdiff_t)i);
      break;
    }
}
}  // namespace

int flush_x(mugpu_cell c);

namespace {

int ensure_init() {
  if (g_device < 0) return mugpu_init(0);
  return MUGPU_OK;
}

int sync_scalars(const double* d_scal) {
  MUGPU_HIP(hipMemcpyAsync(g_host_scal, d_scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost,
                           g_stream));
  MUGPU_HIP(hipStreamSynchronize(g_stream));
  return MUGPU_OK;
}

// scalars and the control words of a device-resident CG loop
int sync_ctl(const int* d_ctl, const double* d_scal) {
  MUGPU_HIP(hipMemcpyAsync(g_host_ctl, d_ctl, sizeof(int) * CTL_COUNT, hipMemcpyDeviceToHost,
                           g_stream));
  return sync_scalars(d_scal);
}

int check_odd(int dim, const int* res, int allow_even) {
  if (allow_even) return MUGPU_OK;
  for (int d = 0; d < dim; ++d)
    if (res[d] % 2 == 0) {
      set_error("Only an odd number of gridpoints in each direction is supported");
      return MUGPU_EPROJECTION;
    }
  return MUGPU_OK;
}

double lame_lambda(double E, double nu) { return E * nu / ((1 + nu) * (1 - 2 * nu)); }
double lame_mu(double E, double nu) { return E / (2 * (1 + nu)); }
double bulk_modulus(double E, double nu) { return E / (3 * (1 - 2 * nu)); }

template <typename P>
void free_dev(P*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// host evaluation of the stored operator for get_operator (debug view)
void host_operator(int kind, int dim, const int* res, const double* L, int allow_even,
                   double* out) {
  int h[3] = {1, 1, 1};
  for (int d = 0; d < dim; ++d) h[d] = res[d];
  h[dim - 1] = res[dim - 1] / 2 + 1;
  int64_t kp = 1;
  for (int d = 0; d < dim; ++d) kp *= h[d];
  const int nc = dim * dim;
  const int nop = kind == MUGPU_PROJ_FINITE_STRAIN_FAST ? dim : nc * nc;
  for (int64_t k = 0; k < kp; ++k) {
    int c[3];
    int64_t r = k;
    for (int d = dim - 1; d >= 0; --d) {
      c[d] = (int)(r % h[d]);
      r /= h[d];
    }
    double xi[3], n2 = 0;
    bool zero = k == 0, nyq = false;
    for (int d = 0; d < dim; ++d) {
      const int npos = (res[d] - 1) / 2 + 1;
      xi[d] = (double)(c[d] < npos ? c[d] : c[d] - res[d]) / L[d];
      n2 += xi[d] * xi[d];
      if (allow_even && res[d] % 2 == 0 && 2 * c[d] == res[d]) nyq = true;
    }
    double* o = out + k * nop;
    if (zero || nyq) {
      for (int q = 0; q < nop; ++q) o[q] = 0.0;
      continue;
    }
    const double nn = std::sqrt(n2);
    for (int d = 0; d < dim; ++d) xi[d] /= nn;
    if (kind == MUGPU_PROJ_FINITE_STRAIN_FAST) {
      for (int d = 0; d < dim; ++d) o[d] = xi[d];
      continue;
    }
    for (int i = 0; i < dim; ++i)
      for (int j = 0; j < dim; ++j)
        for (int l = 0; l < dim; ++l)
          for (int m = 0; m < dim; ++m) {
            double g;
            if (kind == MUGPU_PROJ_SMALL_STRAIN)
              g = 0.5 * (xi[i] * (j == l) * xi[m] + xi[i] * (j == m) * xi[l] +
                         xi[j] * (i == l) * xi[m] + xi[j] * (i == m) * xi[l]) -
                  xi[i] * xi[j] * xi[l] * xi[m];
            else
              g = (i == l) * xi[j] * xi[m];
            o[(i + dim * j) + nc * (l + dim * m)] = g;
          }
  }
}

int cell_operator64(mugpu_cell c, const void* in, void* out) {
  OpSpec op;
  op.in_mode = IN_TAN;
  op.in = in;
  op.tan_kind = c->tan_kind;
  op.tangent = c->tan64;
  op.out_mode = OUT_PLAIN;
  op.out = out;
  return run_operator(c->plan, MUGPU_FP64, op, c->pj, c->spectrum);
}


// work vectors of the mixed-precision solve: FP32 CG vectors, the FP64
// residual and the control words of the device-resident inner loop
struct MixedVecs {
  float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *Ap32 = nullptr;
  double* r64 = nullptr;
  int* ctl = nullptr;
};

// one iteration of the FP32 inner CG (solver_cg.cc:62-96): A p fused with
// the p update (all but the first iteration), p.Ap, the x and r updates and
// r.r, gated by the control words.  fuse: the previous iteration's x update
// rides on this iteration's r2c (x += alpha p_old on the way in) instead of
// the CG update kernel, which then reads neither x nor p
int inner_iteration(mugpu_cell c, const MixedVecs& v, bool first, bool fuse, int64_t n) {
  OpSpec op;
  op.in_mode = first ? IN_TAN : IN_TAN_PUPDATE;
  op.xvec = fuse && !first ? v.x32 : nullptr;
  op.in = v.p32;
  op.rvec = v.r32;
  op.pout = v.p32;
  op.scal = c->scal;
  op.tan_kind = c->tan_kind;
  op.tangent = c->tan32;
  op.out_mode = OUT_DOT;
  op.out = v.Ap32;
  op.dotv = v.p32;
  MUGPU_TRY(run_operator(c->plan, MUGPU_FP32, op, c->pj, c->spectrum));
  MUGPU_TRY(finalize((const double*)c->plan.partials, c->plan.nzb[1], c->scal, FIN_ALPHA, v.ctl));
  return cg_update(MUGPU_FP32, fuse ? nullptr : v.x32, v.r32, v.p32, v.Ap32, c->scal, n,
                   c->partials, v.ctl);
}

// the mixed vectors of a cell, allocated together: after a failed attempt
// none is kept, so a retry allocates them all again
int ensure_mixed(mugpu_cell c, int64_t n) {
  if (c->x32 && c->r32 && c->p32 && c->Ap32 && c->r64) return MUGPU_OK;
  auto drop = [&]() {
    free_dev(c->x32);
    free_dev(c->r32);
    free_dev(c->p32);
    free_dev(c->Ap32);
    free_dev(c->r64);
  };
  drop();
  int rc = dev_alloc(&c->x32, sizeof(float) * n, "the mixed-precision CG vectors");
  if (!rc) rc = dev_alloc(&c->r32, sizeof(float) * n, "the mixed-precision CG vectors");
  if (!rc) rc = dev_alloc(&c->p32, sizeof(float) * n, "the mixed-precision CG vectors");
  if (!rc) rc = dev_alloc(&c->Ap32, sizeof(float) * n, "the mixed-precision CG vectors");
  if (!rc) rc = dev_alloc(&c->r64, sizeof(double) * n, "the mixed-precision CG vectors");
  if (rc) drop();
  return rc;
}

}  // namespace

// ============================== runtime ======================================
extern "C" int mugpu_init(int device) {
  if (g_device >= 0) return MUGPU_OK;
  int count = 0;
  MUGPU_HIP(hipGetDeviceCount(&count));
  if (count <= 0) {
    set_error("no HIP device");
    return MUGPU_EHIP;
  }
  if (device < 0 || device >= count) {
    set_error("device %d out of range (%d devices)", device, count);
    return MUGPU_EARG;
  }
  MUGPU_HIP(hipSetDevice(device));
  MUGPU_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  MUGPU_HIP(hipDeviceGetAttribute(&g_maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  MUGPU_HIP(hipHostMalloc((void**)&g_host_scal, sizeof(double) * SC_COUNT + sizeof(int) * CTL_COUNT,
                          hipHostMallocDefault));
  g_host_ctl = reinterpret_cast<int*>(g_host_scal + SC_COUNT);
  MUGPU_HIP(hipMalloc((void**)&g_dev_scratch, sizeof(double) * (kRedBlocks + SC_COUNT)));
  if (device < kMaxDevices && !g_scratch_ptr[device]) {
    void* ptr = nullptr;
    MUGPU_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = ptr;
  }
  g_device = device;
  return MUGPU_OK;
}

extern "C" const char* mugpu_last_error(void) { return g_err; }

extern "C" int mugpu_device_info(char* name, int name_len, int* cu_count, size_t* hbm_bytes,
                                 int* lds_per_block) {
  MUGPU_TRY(ensure_init());
  hipDeviceProp_t prop;
  MUGPU_HIP(hipGetDeviceProperties(&prop, g_device));
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
  if (lds_per_block) *lds_per_block = g_maxlds;
  return MUGPU_OK;
}

// Field memory is cached: mugpu_free keeps a block for the next mugpu_alloc
// of the same (rounded) size, so repeated solves (a new solver, work and
// result fields per call) do not pay hipMalloc, which maps the pages, and
// hipFree, which also waits for the device, every time.  Every user of these
// blocks works on the library stream or copies synchronously, so stream
// order makes reuse safe.  The cache holds at most a quarter of the device
// memory and is emptied when an allocation fails (then retried) and by
// mugpu_release_cached.
namespace {
std::mutex g_alloc_mu;
std::multimap<size_t, void*> g_cached;      // size -> free block
std::unordered_map<void*, size_t> g_blocks;  // live block -> size
size_t g_cached_bytes = 0, g_cache_cap = 0;

size_t round_block(size_t bytes) {
  constexpr size_t kSmall = 512, kLarge = size_t(2) << 20;
  if (bytes == 0) bytes = 16;
  const size_t g = bytes >= (size_t(1) << 20) ? kLarge : kSmall;
  return (bytes + g - 1) / g * g;
}

void release_cached_locked() {
  for (auto& kv : g_cached) (void)hipFree(kv.second);
  g_cached.clear();
  g_cached_bytes = 0;
}
}  // namespace

// allocations dev_alloc still grants before it reports the device full
// (-1: no limit), set by mugpu_debug_fail_allocations for the tests of the
// failure paths
static std::atomic<int> g_fail_alloc_after{-1};

int mugpu::dev_alloc(void** p, size_t bytes, const char* what) {
  *p = nullptr;
  for (int left = g_fail_alloc_after.load(); left >= 0;) {
    if (left == 0) {
      set_error("device allocation of %s (%zu bytes) refused: a failure injected by "
                "mugpu_debug_fail_allocations", what, bytes);
      return MUGPU_EMEMORY;
    }
    if (g_fail_alloc_after.compare_exchange_weak(left, left - 1)) break;
  }
  hipError_t e = hipMalloc(p, bytes ? bytes : 16);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    if (!g_cached.empty()) {
      (void)hipStreamSynchronize(g_stream);
      release_cached_locked();
      e = hipMalloc(p, bytes ? bytes : 16);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
      }
    }
  }
  if (e != hipSuccess) {
    set_error("device allocation of %s (%zu bytes) failed: %s", what, bytes, hipGetErrorString(e));
    return MUGPU_EMEMORY;
  }
  return MUGPU_OK;
}

extern "C" int mugpu_debug_force_variant(int pass, int precision, int variant) {
  if (pass < 0 || pass >= FV_COUNT || (precision != MUGPU_FP64 && precision != MUGPU_FP32) ||
      variant > 7) {
    set_error("mugpu_debug_force_variant: pass %d, precision %d, variant %d out of range", pass,
              precision, variant);
    return MUGPU_EARG;
  }
  force_variant(pass, precision, variant < 0 ? -1 : variant);
  return MUGPU_OK;
}

extern "C" int mugpu_debug_fail_allocations(int after) {
  g_fail_alloc_after.store(after < 0 ? -1 : after);
  return MUGPU_OK;
}

extern "C" int mugpu_alloc(size_t bytes, void** d_ptr) {
  MUGPU_TRY(ensure_init());
  *d_ptr = nullptr;
  const size_t sz = round_block(bytes);
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  if (g_cache_cap == 0) {
    size_t fr = 0, tot = 0;
    MUGPU_HIP(hipMemGetInfo(&fr, &tot));
    g_cache_cap = tot / 4;
  }
  auto it = g_cached.find(sz);
  if (it != g_cached.end()) {
    *d_ptr = it->second;
    g_cached.erase(it);
    g_cached_bytes -= sz;
    g_blocks[*d_ptr] = sz;
    return MUGPU_OK;
  }
  hipError_t e = hipMalloc(d_ptr, sz);
  if (e != hipSuccess && !g_cached.empty()) {
    (void)hipGetLastError();
    release_cached_locked();
    e = hipMalloc(d_ptr, sz);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *d_ptr = nullptr;
    set_error("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    return MUGPU_EMEMORY;
  }
  g_blocks[*d_ptr] = sz;
  return MUGPU_OK;
}

extern "C" int mugpu_free(void* d_ptr) {
  if (!d_ptr) return MUGPU_OK;
  {
    // pending x updates that involve this block (see pending_set)
    std::vector<mugpu_cell> apply;
    {
      std::lock_guard<std::mutex> lk(g_pend_mu);
      for (size_t i = 0; i < g_pending_cells.size();) {
        mugpu_cell c = g_pending_cells[i];
        if (c->x_pend == d_ptr) {
          c->x_pending = false;
          g_pending_cells.erase(g_pending_cells.begin() + (std::ptrdiff_t)i);
          continue;
        }
        if (c->p_pend == d_ptr) apply.push_back(c);
        ++i;
      }
    }
    for (mugpu_cell c : apply) MUGPU_TRY(flush_x(c));
  }
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  auto it = g_blocks.find(d_ptr);
  if (it == g_blocks.end()) {
    set_error("mugpu_free: %p was not allocated by mugpu_alloc (or was freed already)", d_ptr);
    return MUGPU_EARG;
  }
  const size_t sz = it->second;
  g_blocks.erase(it);
  if (g_cached_bytes + sz <= g_cache_cap) {
    g_cached.emplace(sz, d_ptr);
    g_cached_bytes += sz;
    return MUGPU_OK;
  }
  MUGPU_HIP(hipFree(d_ptr));
  return MUGPU_OK;
}

extern "C" int mugpu_release_cached(void) {
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  if (g_cached.empty()) return MUGPU_OK;
  MUGPU_HIP(hipStreamSynchronize(g_stream));
  release_cached_locked();
  return MUGPU_OK;
}

extern "C" int mugpu_upload(void* d_dst, const void* h_src, size_t bytes) {
  MUGPU_TRY(ensure_init());
  MUGPU_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, g_stream));
  MUGPU_HIP(hipStreamSynchronize(g_stream));
  return MUGPU_OK;
}

extern "C" int mugpu_download(void* h_dst, const void* d_src, size_t bytes) {
  MUGPU_TRY(ensure_init());
  MUGPU_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, g_stream));
  MUGPU_HIP(hipStreamSynchronize(g_stream));
  return MUGPU_OK;
}

extern "C" int mugpu_memset_zero(void* d_ptr, size_t bytes) {
  MUGPU_TRY(ensure_init());
  MUGPU_HIP(hipMemsetAsync(d_ptr, 0, bytes, g_stream));
  return MUGPU_OK;
}

extern "C" int mugpu_synchronize(void) {
  MUGPU_TRY(ensure_init());
  MUGPU_HIP(hipStreamSynchronize(g_stream));
  return MUGPU_OK;
}

extern "C" void* mugpu_stream(void) { return (void*)g_stream; }

extern "C" int mugpu_cg_graphs(int enable) {
  g_cg_graphs = enable != 0;
  return MUGPU_OK;
}

extern "C" int mugpu_timing_start(int kernel_class) {
  MUGPU_TRY(ensure_init());
  g_timing_class = kernel_class;
  g_timing_count = 0;
  return MUGPU_OK;
}

extern "C" int mugpu_timing_stop(int64_t* launches, double* total_ms) {
  MUGPU_TRY(ensure_init());
  MUGPU_HIP(hipStreamSynchronize(g_stream));
  double tot = 0.0;
  for (int k = 0; k < MUGPU_K_ALL; ++k) {
    g_stat_launches[k] = 0;
    g_stat_ms[k] = 0.0;
  }
  int64_t counted = 0;
  for (int64_t i = 0; i < g_timing_count; ++i) {
    const int k = g_ev_class[i];
    if (k == -2) continue;  // a launch that returned at once (device CG loop stopped)
    float ms = 0.f;
    MUGPU_HIP(hipEventElapsedTime(&ms, g_ev_start[i], g_ev_stop[i]));
    tot += ms;
    ++counted;
    if (k >= 0 && k < MUGPU_K_ALL) {
      ++g_stat_launches[k];
      g_stat_ms[k] += ms;
    }
  }
  if (launches) *launches = counted;
  if (total_ms) *total_ms = tot;
  g_timing_class = -1;
  g_timing_count = 0;
  return MUGPU_OK;
}

extern "C" int mugpu_timing_class(int kernel_class, int64_t* launches, double* total_ms) {
  if (kernel_class < 0 || kernel_class >= MUGPU_K_ALL) {
    set_error("kernel class %d out of range", kernel_class);
    return MUGPU_EARG;
  }
  if (launches) *launches = g_stat_launches[kernel_class];
  if (total_ms) *total_ms = g_stat_ms[kernel_class];
  return MUGPU_OK;
}

// time one compile-time-length pass in one variant on scratch buffers of
// the grid res (the autotuner's measurement)
extern "C" int mugpu_fft_tune_pass(int dim, const int* res, int precision, int pass, int variant,
                                   int iso, int reps, double* avg_ms) {
  MUGPU_TRY(ensure_init());
  if (!res || !avg_ms || reps < 1 || pass < 0 || pass >= FV_COUNT || dim < 2 || dim > 3)
    return MUGPU_EARG;
  const int axis = (pass == FV_R2C || pass == FV_C2R) ? dim - 1 : pass == FV_STRIDED1 ? 1 : 0;
  if ((pass == FV_STRIDED1 && dim != 3) || !fast_size(res[axis]) ||
      !fast_variant_valid(pass, res[axis], dim, precision, variant)) {
    set_error("no variant %d of pass %d for an axis of %d points", variant, pass, res[axis]);
    return MUGPU_EARG;
  }
  FftPlan plan;
  const int nc = dim * dim;
  int rc = plan_init(plan, dim, res, nc, precision);
  if (rc) {
    plan_free(plan);
    return rc;
  }
  const int pr = precision == MUGPU_FP64 ? 0 : 1;
  for (int i = 0; i < 2; ++i) plan.fvar[pr][pass][i] = (unsigned char)variant;
  // (the plan's partials buffer holds the c2r partials of every variant)
  const size_t s = precision == MUGPU_FP64 ? 8 : 4;
  const int tk = iso ? TAN_ISO : (dim == 3 ? TAN_SYM : TAN_FULL);
  const size_t nd = (size_t)plan.N * nc;
  void *spec = nullptr, *a = nullptr, *b = nullptr, *c = nullptr, *tan = nullptr;
  double* scal = nullptr;
  bool ok = hipMalloc(&spec, 16 * (size_t)plan.kpts * nc) == hipSuccess &&
            hipMalloc(&a, s * nd) == hipSuccess && hipMalloc(&b, s * nd) == hipSuccess &&
            hipMalloc(&c, s * nd) == hipSuccess &&
            hipMalloc(&tan, s * (size_t)tangent_values(tk, dim, plan.N)) == hipSuccess &&
            hipMalloc((void**)&scal, sizeof(double) * SC_COUNT) == hipSuccess;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ok) {
    ok = hipMemsetAsync(spec, 0, 16 * (size_t)plan.kpts * nc, g_stream) == hipSuccess &&
         hipMemsetAsync(a, 0, s * nd, g_stream) == hipSuccess &&
         hipMemsetAsync(b, 0, s * nd, g_stream) == hipSuccess &&
         hipMemsetAsync(c, 0, s * nd, g_stream) == hipSuccess &&
         hipMemsetAsync(tan, 0, s * (size_t)tangent_values(tk, dim, plan.N), g_stream) == hipSuccess &&
         hipMemsetAsync(scal, 0, sizeof(double) * SC_COUNT, g_stream) == hipSuccess &&
         hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
  }
  if (!ok) {
    set_error("scratch allocation for the pass timing failed");
    rc = MUGPU_EMEMORY;
  }
  auto run = [&]() -> int {
    switch (pass) {
      case FV_R2C: {
        // the CG iteration's form: p update, the previous x update, K:p
        OpSpec op;
        op.in_mode = IN_TAN_PUPDATE;
        op.in = a;
        op.rvec = b;
        op.pout = a;
        op.xvec = c;
        op.scal = scal;
        op.tan_kind = tk;
        op.tangent = tan;
        return fast_r2c(plan, precision, op, spec);
      }
      case FV_C2R: {
        OpSpec op;
        op.out_mode = OUT_DOT;
        op.out = c;
        op.dotv = a;
        return fast_c2r(plan, precision, op, spec);
      }
      case FV_AXIS0: {
        ProjParams pj;
        pj.kind = MUGPU_PROJ_SMALL_STRAIN;
        pj.allow_even = 1;
        pj.norm = 1.0 / (double)plan.N;
        return fast_axis0_project(plan, precision, pj, spec);
      }
      case FV_STRIDED1: return fast_strided(plan, precision, 1, spec, -1);
      default: return fast_strided(plan, precision, 0, spec, -1);
    }
  };
  // each launch is timed on its own after a 1 GiB copy that evicts the
  // last-level cache (256 MB), so that a grid whose fields fit there is timed
  // as it runs inside a CG iteration, not from the cache of the launch before
  const size_t flush = (size_t)1 << 29;
  void* fl = nullptr;
  if (rc == MUGPU_OK && hipMalloc(&fl, 2 * flush) != hipSuccess) {
    set_error("scratch allocation for the pass timing failed");
    rc = MUGPU_EMEMORY;
  }
  if (rc == MUGPU_OK) rc = run();
  double total = 0.0;
  for (int i = 0; i < reps && rc == MUGPU_OK; ++i) {
    rc = copy_device((char*)fl + flush, fl, flush);
    if (rc == MUGPU_OK && hipEventRecord(e0, g_stream) != hipSuccess) rc = MUGPU_EHIP;
    if (rc == MUGPU_OK) rc = run();
    if (rc == MUGPU_OK && (hipEventRecord(e1, g_stream) != hipSuccess ||
                           hipEventSynchronize(e1) != hipSuccess))
      rc = MUGPU_EHIP;
    if (rc == MUGPU_OK) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      total += ms;
    }
  }
  if (rc == MUGPU_OK) *avg_ms = total / reps;
  if (fl) (void)hipFree(fl);
  (void)hipStreamSynchronize(g_stream);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* q : {spec, a, b, c, tan, (void*)scal})
    if (q) (void)hipFree(q);
  plan_free(plan);
  return rc;
}

extern "C" int mugpu_copy(void* d_dst, const void* d_src, size_t bytes) {
  MUGPU_TRY(ensure_init());
  if (bytes % 16) {
    set_error("mugpu_copy needs a multiple of 16 bytes");
    return MUGPU_EARG;
  }
  return copy_device(d_dst, d_src, bytes);
}

extern "C" int mugpu_copy_bandwidth(size_t bytes, int reps, double* gbs) {
  MUGPU_TRY(ensure_init());
  if (!gbs || reps < 1 || bytes < 16 || bytes % 16) {
    set_error("mugpu_copy_bandwidth needs reps >= 1 and a multiple of 16 bytes");
    return MUGPU_EARG;
  }
  int cus = 0;
  MUGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g_device));
  void *a = nullptr, *b = nullptr;
  MUGPU_TRY(dev_alloc(&a, bytes, "the copy source"));
  int rc = dev_alloc(&b, bytes, "the copy destination");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = MUGPU_EHIP;
  if (!rc && hipMemsetAsync(a, 1, bytes, g_stream) != hipSuccess) rc = MUGPU_EHIP;
  for (int v = 0; v < 6 && !rc; ++v) {
    auto run = [&]() -> int {
      switch (v) {
        case 0: return copy_device(b, a, bytes);
        case 1: return copy_device_wide(b, a, bytes, false, cus * 8);
        case 2: return copy_device_wide(b, a, bytes, true, cus * 8);
        case 3:
          return hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, g_stream) == hipSuccess
                     ? MUGPU_OK
                     : MUGPU_EHIP;
        case 4: return read_device_wide(a, bytes, b, cus * 8);
        default: return fill_device_wide(b, bytes, cus * 8);
      }
    };
    rc = run();  // warm-up
    if (!rc && hipEventRecord(e0, g_stream) != hipSuccess) rc = MUGPU_EHIP;
    for (int i = 0; i < reps && !rc; ++i) rc = run();
    if (!rc && (hipEventRecord(e1, g_stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
      rc = MUGPU_EHIP;
    float ms = 0.f;
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = MUGPU_EHIP;
    // bytes moved: read and written for the copies, read or written once
    // for the read-only and write-only streams
    if (!rc) gbs[v] = (v < 4 ? 2.0 : 1.0) * (double)bytes * reps / (ms * 1e-3) / 1e9;
  }
  if (rc == MUGPU_EHIP) set_error("HIP error in the bandwidth measurement");
  (void)hipStreamSynchronize(g_stream);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  free_dev(a);
  free_dev(b);
  return rc;
}

extern "C" int mugpu_line_bandwidth(int precision, int n, int64_t plane, int nb_components,
                                    int kpts_per_tile, int reps, double* gbs) {
  MUGPU_TRY(ensure_init());
  if (!gbs || reps < 1 || n < 16 || n % 16 || plane < 1 || nb_components < 1 ||
      kpts_per_tile < 1 || (int64_t)kpts_per_tile * nb_components * (n / 16) > 1024 ||
      (precision != MUGPU_FP64 && precision != MUGPU_FP32)) {
    set_error("mugpu_line_bandwidth needs n a multiple of 16, reps >= 1 and at most 1024 threads "
              "(kpts_per_tile * nb_components * n / 16)");
    return MUGPU_EARG;
  }
  const size_t bytes = (size_t)n * plane * nb_components * (precision == MUGPU_FP64 ? 16 : 8);
  void* a = nullptr;
  MUGPU_TRY(dev_alloc(&a, bytes, "the line-pattern buffer"));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MUGPU_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = MUGPU_EHIP;
  // 0x3f bytes: finite values in both precisions
  if (!rc && hipMemsetAsync(a, 0x3f, bytes, g_stream) != hipSuccess) rc = MUGPU_EHIP;
  if (!rc) rc = line_pattern(a, precision, n, plane, nb_components, kpts_per_tile);  // warm-up
  if (!rc && hipEventRecord(e0, g_stream) != hipSuccess) rc = MUGPU_EHIP;
  for (int i = 0; i < reps && !rc; ++i) rc = line_pattern(a, precision, n, plane, nb_components, kpts_per_tile);
  if (!rc && (hipEventRecord(e1, g_stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
    rc = MUGPU_EHIP;
  float ms = 0.f;
  if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = MUGPU_EHIP;
  if (!rc) *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
  if (rc == MUGPU_EHIP) set_error("HIP error in the line-pattern measurement");
  (void)hipStreamSynchronize(g_stream);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  free_dev(a);
  return rc;
}

// ============================== FFT engine ====================================
extern "C" int mugpu_fft_create(int dim, const int* res, int nb_components, int precision,
                                mugpu_fft* out) {
  MUGPU_TRY(ensure_init());
  *out = nullptr;
  mugpu_fft f = new (std::nothrow) mugpu_fft_s;
  if (!f) return MUGPU_EMEMORY;
  int rc = plan_init(f->plan, dim, res, nb_components, precision);
  if (rc) {
    plan_free(f->plan);
    delete f;
    return rc;
  }
  *out = f;
  return MUGPU_OK;
}

extern "C" int mugpu_fft_destroy(mugpu_fft plan) {
  if (!plan) return MUGPU_OK;
  plan_free(plan->plan);
  delete plan;
  return MUGPU_OK;
}

extern "C" int mugpu_fft_r2c(mugpu_fft plan, const void* d_real, void* d_spectrum) {
  if (!plan) return MUGPU_EARG;
  return run_r2c(plan->plan, plan->plan.prec, d_real, d_spectrum);
}

extern "C" int mugpu_fft_c2r(mugpu_fft plan, void* d_spectrum, void* d_real) {
  if (!plan) return MUGPU_EARG;
  return run_c2r(plan->plan, plan->plan.prec, d_spectrum, d_real);
}

extern "C" double mugpu_fft_normalisation(mugpu_fft plan) {
  return plan ? 1.0 / (double)plan->plan.N : 0.0;
}

extern "C" int64_t mugpu_fft_nb_fourier_points(mugpu_fft plan) {
  return plan ? plan->plan.kpts : 0;
}

// ============================== projection ====================================
extern "C" int mugpu_projection_create(mugpu_fft engine, int kind, const double* lengths,
                                       int allow_even, mugpu_proj* out) {
  *out = nullptr;
  if (!engine) return MUGPU_EARG;
  if (kind < MUGPU_PROJ_SMALL_STRAIN || kind > MUGPU_PROJ_FINITE_STRAIN) {
    set_error("unknown projection kind %d", kind);
    return MUGPU_EARG;
  }
  MUGPU_TRY(check_odd(engine->plan.dim, engine->plan.res, allow_even));
  mugpu_proj p = new (std::nothrow) mugpu_proj_s;
  if (!p) return MUGPU_EMEMORY;
  p->engine = engine;
  p->pj.kind = kind;
  p->pj.allow_even = allow_even;
  for (int d = 0; d < engine->plan.dim; ++d) p->pj.L[d] = lengths[d];
  p->pj.norm = 1.0 / (double)engine->plan.N;
  const size_t bytes = (size_t)engine->plan.kpts * engine->plan.nc * 2 *
                       (engine->plan.prec == MUGPU_FP64 ? sizeof(double) : sizeof(float));
  const int rc = dev_alloc(&p->spectrum, bytes, "the projection work space");
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return MUGPU_OK;
}

extern "C" int mugpu_projection_destroy(mugpu_proj proj) {
  if (!proj) return MUGPU_OK;
  free_dev(proj->spectrum);
  mugpu_fft_destroy(proj->engine);
  delete proj;
  return MUGPU_OK;
}

extern "C" int mugpu_projection_apply(mugpu_proj proj, void* d_field) {
  if (!proj) return MUGPU_EARG;
  OpSpec op;
  op.in_mode = IN_PLAIN;
  op.in = d_field;
  op.out_mode = OUT_PLAIN;
  op.out = d_field;
  return run_operator(proj->engine->plan, proj->engine->plan.prec, op, proj->pj, proj->spectrum);
}

extern "C" int mugpu_projection_get_operator(mugpu_proj proj, double* h_out) {
  if (!proj) return MUGPU_EARG;
  const FftPlan& p = proj->engine->plan;
  host_operator(proj->pj.kind, p.dim, p.res, proj->pj.L, proj->pj.allow_even, h_out);
  return MUGPU_OK;
}

// ============================== cell ==========================================
extern "C" int mugpu_cell_create(int dim, const int* res, const double* lengths, int formulation,
                                 int precision, int allow_even, mugpu_cell* out) {
  MUGPU_TRY(ensure_init());
  *out = nullptr;
  if (formulation != MUGPU_FINITE_STRAIN && formulation != MUGPU_SMALL_STRAIN) {
    set_error("unknow formulation");
    return MUGPU_ERUNTIME;
  }
  MUGPU_TRY(check_odd(dim, res, allow_even));
  mugpu_cell c = new (std::nothrow) mugpu_cell_s;
  if (!c) return MUGPU_EMEMORY;
  c->dim = dim;
  c->form = formulation;
  c->prec = precision;
  c->allow_even = allow_even;
  c->nc = dim * dim;
  int rc = plan_init(c->plan, dim, res, c->nc, MUGPU_FP64);
  if (rc) {
    plan_free(c->plan);
    delete c;
    return rc;
  }
  c->N = c->plan.N;
  for (int d = 0; d < dim; ++d) {
    c->res[d] = res[d];
    c->L[d] = lengths[d];
    c->pj.L[d] = lengths[d];
  }
  c->pj.kind = formulation == MUGPU_SMALL_STRAIN ? MUGPU_PROJ_SMALL_STRAIN
                                                 : MUGPU_PROJ_FINITE_STRAIN_FAST;
  c->pj.allow_even = allow_even;
  c->pj.norm = 1.0 / (double)c->N;
  const size_t fb = sizeof(double) * (size_t)c->N * c->nc;
  const size_t sb = sizeof(double) * 2 * (size_t)c->plan.kpts * c->nc;
  rc = dev_alloc(&c->F, fb, "the strain field");
  if (!rc) rc = dev_alloc(&c->P, fb, "the stress field");
  if (!rc) rc = dev_alloc(&c->tmp, fb, "the cell's work field");
  if (!rc) rc = dev_alloc(&c->spectrum, sb, "the cell's spectrum");
  if (!rc) rc = dev_alloc(&c->scal, sizeof(double) * SC_COUNT, "the CG scalars");
  if (!rc) rc = dev_alloc(&c->partials, sizeof(double) * kRedBlocks, "the partial sums");
  if (!rc) rc = dev_alloc(&c->ctl, sizeof(int) * CTL_COUNT, "the CG control words");
  if (rc) {
    const std::string msg = g_err;
    mugpu_cell_destroy(c);
    set_error("%s", msg.c_str());
    return rc;
  }
  (void)hipMemsetAsync(c->F, 0, fb, g_stream);
  (void)hipMemsetAsync(c->P, 0, fb, g_stream);
  (void)hipMemsetAsync(c->scal, 0, sizeof(double) * SC_COUNT, g_stream);
  *out = c;
  return MUGPU_OK;
}

extern "C" int mugpu_cell_destroy(mugpu_cell c) {
  if (!c) return MUGPU_OK;
  pending_clear(c);  // a pending x update is dropped with the cell
  (void)hipStreamSynchronize(g_stream);
  for (auto& m : c->mats) {
    free_dev(m.d_pix);
    free_dev(m.d_params);
    free_dev(m.d_C);
    for (int b = 0; b < 2; ++b) {
      free_dev(m.Fp[b]);
      free_dev(m.be[b]);
      free_dev(m.ep[b]);
    }
  }
  free_dev(c->F);
  free_dev(c->P);
  free_dev(c->tan64);
  free_dev(c->tan32);
  free_dev(c->spectrum);
  free_dev(c->tmp);
  free_dev(c->scal);
  free_dev(c->partials);
  free_dev(c->x32);
  free_dev(c->r32);
  free_dev(c->p32);
  free_dev(c->Ap32);
  free_dev(c->r64);
  free_dev(c->ctl);
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  plan_free(c->plan);
  delete c;
  return MUGPU_OK;
}

extern "C" int64_t mugpu_cell_nb_pixels(mugpu_cell c) { return c ? c->N : 0; }
extern "C" int mugpu_cell_nb_components(mugpu_cell c) { return c ? c->nc : 0; }

extern "C" int mugpu_cell_add_material(mugpu_cell c, int family, double young, double poisson,
                                       double tau_y0, double H, int* mat) {
  if (!c) return MUGPU_EARG;
  if (c->initialised) {
    set_error("materials must be added before the cell is initialised");
    return MUGPU_ERUNTIME;
  }
  if (family == MUGPU_LEG) {
    set_error("MaterialLinearElasticGeneric takes a stiffness matrix: use mugpu_cell_add_material_generic");
    return MUGPU_EARG;
  }
  if (family < MUGPU_LE1 || family > MUGPU_HEP1_CT) {
    set_error("unknown material family %d", family);
    return MUGPU_EARG;
  }
  if ((family == MUGPU_HEP1 || family == MUGPU_HEP1_CT) && c->form != MUGPU_FINITE_STRAIN) {
    set_error("MaterialHyperElastoPlastic1 needs the finite-strain formulation");
    return MUGPU_ERUNTIME;
  }
  MaterialDev m;
  m.family = family;
  m.young = young;
  m.poisson = poisson;
  m.tau_y0 = tau_y0;
  m.H = H;
  m.lambda = lame_lambda(young, poisson);
  m.mu = lame_mu(young, poisson);
  m.bulk = bulk_modulus(young, poisson);
  m.ppn = family == MUGPU_LE2 ? c->nc : (family == MUGPU_LE3 || family == MUGPU_LE4) ? 2 : 0;
  c->mats.push_back(std::move(m));
  *mat = (int)c->mats.size() - 1;
  return MUGPU_OK;
}

extern "C" int mugpu_cell_add_material_generic(mugpu_cell c, const double* C_voigt, int* mat) {
  if (!c || !C_voigt || !mat) return MUGPU_EARG;
  if (c->initialised) {
    set_error("materials must be added before the cell is initialised");
    return MUGPU_ERUNTIME;
  }
  // VoigtConversion<dim>::sym_mat (src/common/voigt_conversion.hh:118-125)
  static const int sym2[2][2] = {{0, 2}, {2, 1}};
  static const int sym3[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
  const int d = c->dim, nc = c->nc, vs = d == 2 ? 3 : 6;
  auto sym = [&](int i, int j) { return d == 2 ? sym2[i][j] : sym3[i][j]; };
  MaterialDev m;
  m.family = MUGPU_LEG;
  m.h_C.assign((size_t)nc * nc, 0.0);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j)
      for (int k = 0; k < d; ++k)
        for (int l = 0; l < d; ++l)
          m.h_C[(i + d * j) + nc * (k + d * l)] = C_voigt[sym(i, j) * vs + sym(k, l)];
  for (int r = 0; r < nc; ++r)
    for (int q = 0; q < nc; ++q)
      if (m.h_C[r + nc * q] != m.h_C[q + nc * r]) m.major_sym = false;
  c->mats.push_back(std::move(m));
  *mat = (int)c->mats.size() - 1;
  return MUGPU_OK;
}

extern "C" int mugpu_material_add_pixels(mugpu_cell c, int mat, int64_t count,
                                         const int64_t* pixels, const double* params) {
  if (!c || mat < 0 || mat >= (int)c->mats.size()) return MUGPU_EARG;
  if (c->initialised) {
    set_error("pixels must be added before the cell is initialised");
    return MUGPU_ERUNTIME;
  }
  MaterialDev& m = c->mats[mat];
  if (m.ppn && !params) {
    set_error(m.family == MUGPU_LE2 ? "this material needs pixels with and eigenstrain"
                                    : "this material needs pixels with Youngs modulus and Poisson ratio.");
    return MUGPU_ERUNTIME;
  }
  for (int64_t i = 0; i < count; ++i) {
    if (pixels[i] < 0 || pixels[i] >= c->N) {
      set_error("pixel index %lld outside the grid", (long long)pixels[i]);
      return MUGPU_EARG;
    }
  }
  for (int64_t i = 0; i < count; ++i) {
    m.h_pix.push_back(pixels[i]);
    if (m.family == MUGPU_LE2) {
      for (int k = 0; k < c->nc; ++k) m.h_params.push_back(params[i * c->nc + k]);
    } else if (m.family == MUGPU_LE3 || m.family == MUGPU_LE4) {
      const double E = params[2 * i], nu = params[2 * i + 1];
      m.h_params.push_back(lame_lambda(E, nu));
      m.h_params.push_back(lame_mu(E, nu));
    }
  }
  return MUGPU_OK;
}

extern "C" int mugpu_cell_initialise(mugpu_cell c) {
  if (!c) return MUGPU_EARG;
  if (c->initialised) return MUGPU_OK;
  // CellBase::check_material_coverage (src/cell/cell_base.cc:486-526)
  std::vector<int> assign((size_t)c->N, -1);
  for (size_t mi = 0; mi < c->mats.size(); ++mi)
    for (int64_t p : c->mats[mi].h_pix) {
      if (assign[(size_t)p] >= 0) {
        set_error("Pixel %lld is already assigned to material %d and cannot be reassigned to material %d",
                  (long long)p, assign[(size_t)p], (int)mi);
        return MUGPU_ERUNTIME;
      }
      assign[(size_t)p] = (int)mi;
    }
  for (int64_t i = 0; i < c->N; ++i)
    if (assign[(size_t)i] < 0) {
      set_error("The following pixels have were not assigned a material: %lld, ... and that cannot be handled",
                (long long)i);
      return MUGPU_ERUNTIME;
    }
  // tangent representation
  // small strain: isotropic Hooke per voxel -> (lambda, mu), unless a
  // generic material brings a general C; finite strain: the coded HEP1
  // tangent alone -> its dim^3 form; laws whose tangent is major-symmetric
  // (LE1-4 and a major-symmetric generic C pushed forward from PK2, the
  // consistent HEP1 tangent; 3D) -> upper triangle; otherwise the full matrix
  bool all_hep = !c->mats.empty(), all_sym = c->dim == 3 && !c->mats.empty(), any_gen = false;
  for (auto& m : c->mats) {
    all_hep = all_hep && m.family == MUGPU_HEP1;
    all_sym = all_sym && m.family != MUGPU_HEP1 && (m.family != MUGPU_LEG || m.major_sym);
    any_gen = any_gen || m.family == MUGPU_LEG;
  }
  c->tan_kind = c->form == MUGPU_SMALL_STRAIN && !any_gen ? TAN_ISO
                : all_hep                                 ? TAN_HEP
                : all_sym                                 ? TAN_SYM
                                                          : TAN_FULL;
  const size_t tn = (size_t)tangent_values(c->tan_kind, c->dim, c->N);
  // an initialise that failed part way is retried from scratch: what the
  // failed call allocated is released first
  free_dev(c->tan64);
  free_dev(c->tan32);
  MUGPU_TRY(dev_alloc(&c->tan64, tn * sizeof(double), "the tangent"));
  MUGPU_HIP(hipMemsetAsync(c->tan64, 0, tn * sizeof(double), g_stream));
  if (c->prec == MUGPU_FP32) MUGPU_TRY(dev_alloc(&c->tan32, tn * sizeof(float), "the FP32 tangent"));
  const int nc = c->nc;
  for (auto& m : c->mats) {
    free_dev(m.d_pix);
    free_dev(m.d_params);
    free_dev(m.d_C);
    for (int b = 0; b < 2; ++b) {
      free_dev(m.Fp[b]);
      free_dev(m.be[b]);
      free_dev(m.ep[b]);
    }
  }
  for (auto& m : c->mats) {
    const int64_t n = (int64_t)m.h_pix.size();
    std::vector<int> pix32(m.h_pix.begin(), m.h_pix.end());
    MUGPU_TRY(dev_alloc(&m.d_pix, sizeof(int) * (size_t)(n > 0 ? n : 1), "a material's pixel list"));
    if (n) MUGPU_HIP(hipMemcpy(m.d_pix, pix32.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    if (m.family == MUGPU_LEG) {
      MUGPU_TRY(dev_alloc(&m.d_C, sizeof(double) * m.h_C.size(), "a material's stiffness"));
      MUGPU_HIP(hipMemcpy(m.d_C, m.h_C.data(), sizeof(double) * m.h_C.size(), hipMemcpyHostToDevice));
    }
    if (m.ppn) {
      MUGPU_TRY(dev_alloc(&m.d_params, sizeof(double) * (size_t)(n > 0 ? n * m.ppn : 1),
                          "a material's per-pixel parameters"));
      if (n) MUGPU_HIP(hipMemcpy(m.d_params, m.h_params.data(), sizeof(double) * n * m.ppn,
                                 hipMemcpyHostToDevice));
    }
    if (m.family == MUGPU_HEP1 || m.family == MUGPU_HEP1_CT) {
      // MaterialHyperElastoPlastic1::initialise (.cc:76-84): current = I,
      // then cycle; both buffers start at zero
      const size_t sz = (size_t)(n > 0 ? n : 1);
      for (int b = 0; b < 2; ++b) {
        MUGPU_TRY(dev_alloc(&m.Fp[b], sizeof(double) * sz * nc, "the HEP1 state"));
        MUGPU_TRY(dev_alloc(&m.be[b], sizeof(double) * sz * nc, "the HEP1 state"));
        MUGPU_TRY(dev_alloc(&m.ep[b], sizeof(double) * sz, "the HEP1 state"));
        MUGPU_HIP(hipMemsetAsync(m.Fp[b], 0, sizeof(double) * sz * nc, g_stream));
        MUGPU_HIP(hipMemsetAsync(m.be[b], 0, sizeof(double) * sz * nc, g_stream));
        MUGPU_HIP(hipMemsetAsync(m.ep[b], 0, sizeof(double) * sz, g_stream));
      }
      std::vector<double> eye(nc, 0.0);
      for (int i = 0; i < c->dim; ++i) eye[i * (c->dim + 1)] = 1.0;
      MUGPU_TRY(fill_uniform(m.Fp[0], eye.data(), nc, n, 0));
      MUGPU_TRY(fill_uniform(m.be[0], eye.data(), nc, n, 0));
      m.cur = 1;
    }
  }
  MUGPU_HIP(hipStreamSynchronize(g_stream));
  c->initialised = 1;
  return MUGPU_OK;
}

extern "C" int mugpu_cell_is_initialised(mugpu_cell c) { return c ? c->initialised : 0; }

extern "C" int mugpu_cell_strain(mugpu_cell c, void** d_F) {
  if (!c) return MUGPU_EARG;
  *d_F = c->F;
  return MUGPU_OK;
}

extern "C" int mugpu_cell_stress(mugpu_cell c, void** d_P) {
  if (!c) return MUGPU_EARG;
  *d_P = c->P;
  return MUGPU_OK;
}

extern "C" int mugpu_cell_set_uniform_strain(mugpu_cell c, const double* h_strain) {
  if (!c) return MUGPU_EARG;
  return fill_uniform(c->F, h_strain, c->nc, c->N, 0);
}

extern "C" int mugpu_cell_add_uniform_strain(mugpu_cell c, const double* h_increment) {
  if (!c) return MUGPU_EARG;
  return fill_uniform(c->F, h_increment, c->nc, c->N, 1);
}

static int evaluate(mugpu_cell c, int need_tangent) {
  if (!c) return MUGPU_EARG;
  MUGPU_TRY(mugpu_cell_initialise(c));
  for (auto& m : c->mats)
    MUGPU_TRY(eval_material(m, c->dim, c->form, c->F, c->P, c->tan_kind, c->tan64, need_tangent));
  if (need_tangent && c->prec == MUGPU_FP32) {
    const int64_t tn = tangent_values(c->tan_kind, c->dim, c->N);
    MUGPU_TRY(convert(MUGPU_FP64, c->tan64, MUGPU_FP32, c->tan32, tn, 1.0));
  }
  return MUGPU_OK;
}

extern "C" int mugpu_cell_evaluate_stress_tangent(mugpu_cell c) { return evaluate(c, 1); }
extern "C" int mugpu_cell_evaluate_stress(mugpu_cell c) { return evaluate(c, 0); }

extern "C" int mugpu_cell_get_tangent(mugpu_cell c, double* h_K) {
  if (!c || !c->tan64) {
    set_error("K does not exist");
    return MUGPU_ERUNTIME;
  }
  const int d = c->dim, nc = c->nc, st = tangent_stride(c->tan_kind, d);
  std::vector<double> t((size_t)tangent_values(c->tan_kind, d, c->N));
  MUGPU_TRY(mugpu_download(t.data(), c->tan64, t.size() * sizeof(double)));
  for (int64_t v = 0; v < c->N; ++v) {
    double* K = h_K + v * nc * nc;
    auto s = [&](int q) { return t[(size_t)tan_off(v, q, st)]; };
    if (c->tan_kind == TAN_FULL) {
      for (int q = 0; q < nc * nc; ++q) K[q] = s(q);
    } else if (c->tan_kind == TAN_SYM) {
      for (int col = 0; col < nc; ++col)
        for (int row = 0; row < nc; ++row)
          K[row + nc * col] = row <= col ? s(sym_index(row, col, nc)) : s(sym_index(col, row, nc));
    } else if (c->tan_kind == TAN_ISO) {
      for (int l = 0; l < d; ++l)
        for (int k = 0; k < d; ++k)
          for (int j = 0; j < d; ++j)
            for (int i = 0; i < d; ++i) {
              const double itrac = (i == j && k == l) ? 1.0 : 0.0;
              const double isym = 0.5 * (((i == k && j == l) ? 1.0 : 0.0) + ((i == l && j == k) ? 1.0 : 0.0));
              K[(i + d * j) + nc * (k + d * l)] = s(0) * itrac + 2 * s(1) * isym;
            }
    } else {
      for (int n = 0; n < d; ++n)
        for (int j = 0; j < d; ++j)
          for (int m = 0; m < d; ++m)
            for (int i = 0; i < d; ++i) K[(i + d * m) + nc * (j + d * n)] = s(i + d * m + d * d * n);
    }
  }
  return MUGPU_OK;
}

extern "C" int mugpu_cell_directional_stiffness(mugpu_cell c, const void* d_in, void* d_out) {
  if (!c) return MUGPU_EARG;
  if (!c->tan64) {
    set_error("currently only implemented for cases where a stiffness matrix exists");
    return MUGPU_ERUNTIME;
  }
  return cell_operator64(c, d_in, d_out);
}

extern "C" int mugpu_cell_tangent_product(mugpu_cell c, const void* d_in, void* d_out) {
  if (!c) return MUGPU_EARG;
  if (!c->tan64) {
    set_error("currently only implemented for cases where a stiffness matrix exists");
    return MUGPU_ERUNTIME;
  }
  return run_tangent_product(c->dim, MUGPU_FP64, c->tan_kind, c->tan64, d_in, d_out, c->N);
}

extern "C" int mugpu_cell_apply_projection(mugpu_cell c, void* d_vec) {
  if (!c) return MUGPU_EARG;
  MUGPU_TRY(mugpu_cell_initialise(c));
  OpSpec op;
  op.in_mode = IN_PLAIN;
  op.in = d_vec;
  op.out_mode = OUT_PLAIN;
  op.out = d_vec;
  return run_operator(c->plan, MUGPU_FP64, op, c->pj, c->spectrum);
}

extern "C" int mugpu_cell_save_history(mugpu_cell c) {
  if (!c) return MUGPU_EARG;
  for (auto& m : c->mats)
    if (m.family == MUGPU_HEP1 || m.family == MUGPU_HEP1_CT) m.cur ^= 1;
  return MUGPU_OK;
}

extern "C" int mugpu_cell_get_internal(mugpu_cell c, const char* name, double* h_out,
                                       int* nb_components) {
  if (!c || !name) return MUGPU_EARG;
  // short names, or the reference's field names (material_linear_elastic2.cc:59,
  // material_hyper_elasto_plastic1.cc:44-52)
  std::string nm(name);
  if (nm == "Eigenstrain") nm = "eigenstrain";
  if (nm == "cumulated plastic flow \xce\xb5\xe2\x82\x9a") nm = "eps_p";
  if (nm == "Previous placement gradient F\xe1\xb5\x97") nm = "F_prev";
  if (nm == "Previous left Cauchy-Green deformation b\xe2\x82\x91\xe1\xb5\x97") nm = "be";
  int found = 0, ncomp = 0;
  for (auto& m : c->mats) {
    const double* src = nullptr;
    int k = 0;
    std::vector<double> host;
    if ((m.family == MUGPU_HEP1 || m.family == MUGPU_HEP1_CT) && m.Fp[0]) {
      if (nm == "eps_p") { src = m.ep[m.cur]; k = 1; }
      else if (nm == "F_prev") { src = m.Fp[m.cur]; k = c->nc; }
      else if (nm == "be") { src = m.be[m.cur]; k = c->nc; }
      if (src) {
        host.resize(m.h_pix.size() * k);
        MUGPU_TRY(mugpu_download(host.data(), src, host.size() * sizeof(double)));
      }
    } else if (m.family == MUGPU_LE2 && nm == "eigenstrain") {
      host = m.h_params;
      k = c->nc;
      src = host.data();
    }
    if (!k) continue;
    if (!found) {
      ncomp = k;
      std::memset(h_out, 0, sizeof(double) * (size_t)c->N * k);
    } else if (k != ncomp) {
      set_error("The fields named '%s' do not have the same number of components in every material", name);
      return MUGPU_ERUNTIME;
    }
    found = 1;
    for (size_t q = 0; q < m.h_pix.size(); ++q)
      std::memcpy(h_out + m.h_pix[q] * k, host.data() + q * k, sizeof(double) * k);
  }
  if (!found) {
    set_error("The field named '%s' does not exist in any of the materials and can therefore not be globalised!", name);
    return MUGPU_ERUNTIME;
  }
  *nb_components = ncomp;
  return MUGPU_OK;
}

extern "C" int mugpu_material_get_state(mugpu_cell c, int mat, const char* name, int sub_field,
                                        double* h_out, int* nb_components,
                                        int* current_sub_field) {
  if (!c || !name || mat < 0 || mat >= (int)c->mats.size()) return MUGPU_EARG;
  const MaterialDev& m = c->mats[(size_t)mat];
  if (m.family != MUGPU_HEP1 && m.family != MUGPU_HEP1_CT) {
    set_error("material %d holds no state fields", mat);
    return MUGPU_ERUNTIME;
  }
  if (!m.Fp[0]) {
    set_error("the state fields exist once the cell is initialised");
    return MUGPU_ERUNTIME;
  }
  if (sub_field < 0 || sub_field > 1) {
    set_error("sub_field index %d: a state field with one old value has sub-fields 0 and 1",
              sub_field);
    return MUGPU_EARG;
  }
  // short names or the reference's prefixes (material_hyper_elasto_plastic1.cc:46-53)
  const std::string nm(name);
  const double* src = nullptr;
  int k = 0;
  if (nm == "eps_p" || nm == "cumulated plastic flow \xce\xb5\xe2\x82\x9a") {
    src = m.ep[sub_field];
    k = 1;
  } else if (nm == "F_prev" || nm == "Previous placement gradient F\xe1\xb5\x97") {
    src = m.Fp[sub_field];
    k = c->nc;
  } else if (nm == "be" || nm == "Previous left Cauchy-Green deformation b\xe2\x82\x91\xe1\xb5\x97") {
    src = m.be[sub_field];
    k = c->nc;
  } else {
    set_error("no state field '%s' in this material", name);
    return MUGPU_ERUNTIME;
  }
  if (nb_components) *nb_components = k;
  if (current_sub_field) *current_sub_field = m.cur;
  if (h_out && !m.h_pix.empty())
    MUGPU_TRY(mugpu_download(h_out, src, sizeof(double) * m.h_pix.size() * (size_t)k));
  return MUGPU_OK;
}

extern "C" int mugpu_cell_projection_operator(mugpu_cell c, double* h_out) {
  if (!c || !h_out) return MUGPU_EARG;
  host_operator(c->pj.kind, c->dim, c->res, c->pj.L, c->pj.allow_even, h_out);
  return MUGPU_OK;
}

// ============================== vector kernels ================================
extern "C" int mugpu_dot(const void* d_a, const void* d_b, int64_t n, double* h_out) {
  MUGPU_TRY(ensure_init());
  double* part = g_dev_scratch;
  double* scal = g_dev_scratch + kRedBlocks;
  MUGPU_TRY(dot_device(MUGPU_FP64, d_a, d_b, n, part, scal, FIN_STORE_TMP0));
  MUGPU_TRY(sync_scalars(scal));
  *h_out = g_host_scal[SC_TMP0];
  return MUGPU_OK;
}

extern "C" int mugpu_axpby(double a, const void* d_x, double b, void* d_y, int64_t n) {
  MUGPU_TRY(ensure_init());
  return axpby(MUGPU_FP64, a, d_x, b, d_y, n);
}

extern "C" int mugpu_scale_copy(double a, const void* d_x, void* d_y, int64_t n) {
  MUGPU_TRY(ensure_init());
  return convert(MUGPU_FP64, d_x, MUGPU_FP64, d_y, n, a);
}

// ============================== SolverCG ======================================
extern "C" int mugpu_cg_start(mugpu_cell c, const void* d_rhs, void* d_x, void* d_r, void* d_p,
                              void* d_Ap, double* rdr, double* rhs_norm2) {
  if (!c) return MUGPU_EARG;
  if (!c->tan64) {
    set_error("currently only implemented for cases where a stiffness matrix exists");
    return MUGPU_ERUNTIME;
  }
  const int64_t n = c->N * c->nc;
  MUGPU_TRY(flush_x(c));
  MUGPU_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, g_stream));
  // r = A x - b (solver_cg.cc:62-63) at x = 0: the operator of a zero field
  // is zero exactly (FFT, projection and K : p of zeros), so r = 0 - b
  // without applying it
  MUGPU_TRY(cg_init_vectors(MUGPU_FP64, nullptr, d_rhs, d_r, d_p, n));
  MUGPU_TRY(dot_device(MUGPU_FP64, d_r, d_r, n, c->partials, c->scal, FIN_STORE_RDR));
  MUGPU_TRY(dot_device(MUGPU_FP64, d_rhs, d_rhs, n, c->partials, c->scal, FIN_STORE_TMP0));
  MUGPU_TRY(sync_scalars(c->scal));
  *rdr = g_host_scal[SC_RDR];
  *rhs_norm2 = g_host_scal[SC_TMP0];
  return MUGPU_OK;
}

// the x update left by mugpu_cg_steps: x += alpha p (alpha of its last iteration)
int flush_x(mugpu_cell c) {
  if (!c->x_pending) return MUGPU_OK;
  pending_clear(c);
  return axpy_alpha(MUGPU_FP64, c->x_pend, c->p_pend, c->scal, c->N * c->nc, SC_XALPHA);
}

extern "C" int mugpu_cg_finish(mugpu_cell c) {
  if (!c) return MUGPU_EARG;
  return flush_x(c);
}

extern "C" int mugpu_cg_step(mugpu_cell c, int update_p, void* d_x, void* d_r, void* d_p,
                             void* d_Ap, double* new_rdr) {
  if (!c) return MUGPU_EARG;
  MUGPU_TRY(flush_x(c));
  const int64_t n = c->N * c->nc;
  OpSpec op;
  op.in_mode = update_p ? IN_TAN_PUPDATE : IN_TAN;
  op.in = d_p;
  op.rvec = d_r;
  op.pout = d_p;
  op.scal = c->scal;
  op.tan_kind = c->tan_kind;
  op.tangent = c->tan64;
  op.out_mode = OUT_DOT;
  op.out = d_Ap;
  op.dotv = d_p;
  MUGPU_TRY(run_operator(c->plan, MUGPU_FP64, op, c->pj, c->spectrum));
  MUGPU_TRY(finalize((const double*)c->plan.partials, c->plan.nzb[0], c->scal, FIN_ALPHA));
  MUGPU_TRY(cg_update(MUGPU_FP64, d_x, d_r, d_p, d_Ap, c->scal, n, c->partials));
  MUGPU_TRY(sync_scalars(c->scal));
  *new_rdr = g_host_scal[SC_NEWRDR];
  return MUGPU_OK;
}

extern "C" int mugpu_cg_steps(mugpu_cell c, int update_p, void* d_x, void* d_r, void* d_p,
                              void* d_Ap, double tol2, int max_steps, int* steps,
                              double* new_rdr) {
  if (!c || max_steps < 1 || !steps || !new_rdr) return MUGPU_EARG;
  const int64_t n = c->N * c->nc;
  MUGPU_TRY(set_ctl(c->ctl, max_steps, c->scal, tol2));
  OpSpec op;
  op.in = d_p;
  op.rvec = d_r;
  op.pout = d_p;
  op.scal = c->scal;
  op.tan_kind = c->tan_kind;
  op.tangent = c->tan64;
  op.out_mode = OUT_DOT;
  op.out = d_Ap;
  op.dotv = d_p;
  // every iteration is queued; once the loop test fails on the device the
  // remaining kernels return at once
  int rc = MUGPU_OK;
  std::vector<int64_t> marks;
  // the x update of iteration i rides on the r2c of iteration i + 1 (x +=
  // alpha p_old on the way in) where the last axis has compiled passes; the
  // last one is left pending for the next call or mugpu_cg_finish
  const bool fuse = c->plan.fastax[c->dim - 1];
  const bool pend_here = c->x_pending && c->x_pend == d_x && c->p_pend == d_p && update_p;
  if (c->x_pending && !pend_here) MUGPU_TRY(flush_x(c));
  // the first r2c applies the pending update with scal[SC_ALPHA]
  if (pend_here) MUGPU_TRY(scal_copy(c->scal, SC_XALPHA, SC_ALPHA));
  pending_clear(c);
  c->plan.active = c->ctl;
  auto queue = [&](int i) -> int {
    op.in_mode = (i > 0 || update_p) ? IN_TAN_PUPDATE : IN_TAN;
    op.xvec = fuse && (i > 0 || pend_here) ? d_x : nullptr;
    int e = run_operator(c->plan, MUGPU_FP64, op, c->pj, c->spectrum);
    if (e == MUGPU_OK)
      e = finalize((const double*)c->plan.partials, c->plan.nzb[0], c->scal, FIN_ALPHA, c->ctl);
    if (e == MUGPU_OK)
      e = cg_update(MUGPU_FP64, fuse ? nullptr : d_x, d_r, d_p, d_Ap, c->scal, n, c->partials,
                    c->ctl);
    return e;
  };
  // On small cells every kernel of a chunk takes a few microseconds, and the
  // host's launches, not the device, set the pace: the chunk is captured once
  // into a hipGraph (per vectors, chunk length and first-iteration form; every
  // other pointer the kernels take belongs to the cell) and replayed.  Not
  // while kernels are timed one by one, which needs events between them.
  const bool graph = g_cg_graphs && n <= kCgGraphMaxValues && g_timing_class < 0;
  if (graph) {
    hipGraphExec_t ex = nullptr;
    const int fu = update_p ? 1 : 0, fx = pend_here ? 1 : 0;
    for (auto& g : c->graphs)
      if (g.x == d_x && g.r == d_r && g.p == d_p && g.Ap == d_Ap && g.steps == max_steps &&
          g.first_update_p == fu && g.first_x == fx)
        ex = g.exec;
    if (!ex) {
      hipGraph_t gr = nullptr;
      MUGPU_HIP(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
      set_capturing(true);
      for (int i = 0; i < max_steps && rc == MUGPU_OK; ++i) rc = queue(i);
      set_capturing(false);
      const hipError_t e = hipStreamEndCapture(g_stream, &gr);
      if (rc == MUGPU_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
      if (rc == MUGPU_OK) {
        const hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        if (ei != hipSuccess) rc = hip_fail(ei, "hipGraphInstantiate", __FILE__, __LINE__);
      }
      if (gr) (void)hipGraphDestroy(gr);
      if (rc == MUGPU_OK) {
        if (c->graphs.size() >= kCgGraphsPerCell) {
          (void)hipGraphExecDestroy(c->graphs.front().exec);
          c->graphs.erase(c->graphs.begin());
        }
        c->graphs.push_back({d_x, d_r, d_p, d_Ap, max_steps, fu, fx, ex});
      }
    }
    if (rc == MUGPU_OK && hipGraphLaunch(ex, g_stream) != hipSuccess)
      rc = hip_fail(hipGetLastError(), "hipGraphLaunch", __FILE__, __LINE__);
  } else {
    for (int i = 0; i < max_steps && rc == MUGPU_OK; ++i) {
      marks.push_back(timing_mark());
      rc = queue(i);
    }
  }
  c->plan.active = nullptr;
  MUGPU_TRY(rc);
  if (fuse) {
    // alpha of the last iteration run (launches after the stop leave it)
    MUGPU_TRY(scal_copy(c->scal, SC_ALPHA, SC_XALPHA));
    pending_set(c, d_x, d_p);
  }
  MUGPU_TRY(sync_ctl(c->ctl, c->scal));
  const double r = g_host_scal[SC_NEWRDR];
  // a NaN r.r stops the device loop without counting its iteration; the
  // host loop counts it
  *steps = g_host_ctl[CTL_ITER] + (r == r ? 0 : 1);
  *new_rdr = r;
  // launches after the loop test failed did no work: left out of the timing
  if (*steps < (int)marks.size()) timing_void(marks[*steps]);
  return MUGPU_OK;
}

// ==================== mixed precision CG with FP64 refinement ===================
// Body of the mixed-precision solve (mugpu_cg_solve_mixed) over the work
// vectors v; allocates nothing.
int solve_mixed(mugpu_cell c, const MixedVecs& v, const void* d_rhs, void* d_x, double tol,
                int maxiter, double inner_tol, int64_t* iterations, int* refinements,
                double* rel_residual) {
  const int64_t n = c->N * c->nc;
  MUGPU_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, g_stream));
  MUGPU_TRY(dot_device(MUGPU_FP64, d_rhs, d_rhs, n, c->partials, c->scal, FIN_STORE_TMP1));
  MUGPU_TRY(sync_scalars(c->scal));
  const double bnorm2 = g_host_scal[SC_TMP1];
  int64_t its = 0;
  int refs = 0;
  double rr = bnorm2;
  const double tol2 = tol * tol * bnorm2;
  bool first = true;
  while (true) {
    // true residual with the FP64 operator: r64 = b - A x; at the start
    // x = 0, so A x = 0 exactly and the residual is b itself (read in place
    // of r64 below), |b|^2 bit for bit
    const double* res64 = first ? (const double*)d_rhs : v.r64;
    if (first) {
      rr = bnorm2;
      first = false;
    } else {
      MUGPU_TRY(cell_operator64(c, d_x, v.r64));
      MUGPU_TRY(residual_dot((const double*)d_rhs, v.r64, n, c->partials, c->scal,
                             FIN_STORE_TMP1));
      MUGPU_TRY(sync_scalars(c->scal));
      rr = g_host_scal[SC_TMP1];
    }
    ++refs;
    if (!(rr > tol2) || its >= maxiter || !(rr == rr)) break;
    // inner CG in FP32 on A d = r64 / |r64|
    const double s = 1.0 / std::sqrt(rr);
    // r32 = -r64 s, p32 = -r32, x32 = 0, r32.r32: one pass over r64
    MUGPU_TRY(mixed_start(res64, s, v.r32, v.p32, v.x32, n, c->partials, c->scal,
                          FIN_STORE_RDR));
    MUGPU_TRY(sync_scalars(c->scal));
    const double rdr0 = g_host_scal[SC_RDR];
    const double tol2i = inner_threshold2(inner_tol, tol2, rr, rdr0);
    // SolverCG's loop, for (i = 0; its < maxiter && (r.r > tol2i || i == 0);
    // ++i, ++its), run on the device: iterations are queued kChunk at a time
    // ahead of the host, the FIN_BETA finalize applies the loop test and
    // clears ctl[CTL_ACTIVE], after which the remaining kernels of the chunk
    // return at once; the host reads the counters once per chunk
    constexpr int kChunk = 8;
    const int maxit = (int)(maxiter - its);
    MUGPU_TRY(set_ctl(v.ctl, maxit, c->scal, tol2i));
    int launched = 0;
    int rc = MUGPU_OK;
    // timing marks of the current chunk's iterations: the loop stops inside
    // the last chunk queued, so only its marks are needed
    int64_t marks[kChunk];
    int chunk0 = 0;
    // x updates fused into the next r2c (compile-time-length last axis)
    const bool fuse = c->plan.fastax[c->dim - 1];
    c->plan.active = v.ctl;
    while (rc == MUGPU_OK) {
      chunk0 = launched;
      for (int j = 0; j < kChunk && launched < maxit && rc == MUGPU_OK; ++j, ++launched) {
        marks[j] = timing_mark();
        rc = inner_iteration(c, v, launched == 0, fuse, n);
      }
      if (rc != MUGPU_OK) break;
      rc = sync_ctl(v.ctl, c->scal);
      if (rc != MUGPU_OK || !g_host_ctl[CTL_ACTIVE] || launched >= maxit) break;
    }
    c->plan.active = nullptr;
    MUGPU_TRY(rc);
    // launches after the loop test failed did no work: left out of the timing
    // (the iteration with a NaN r.r, not counted, did)
    {
      const int ran = g_host_ctl[CTL_ITER] + (g_host_scal[SC_NEWRDR] == g_host_scal[SC_NEWRDR] ? 0 : 1);
      if (ran >= chunk0 && ran < launched) timing_void(marks[ran - chunk0]);
    }
    its += g_host_ctl[CTL_ITER];
    // x += |r64| d, with the last iteration's x update (alpha and p of the
    // last iteration run) folded in when it is pending
    MUGPU_TRY(add_scaled_convert(d_x, v.x32, 1.0 / s, n, fuse ? v.p32 : nullptr, c->scal));
  }

  if (iterations) *iterations = its;
  if (refinements) *refinements = refs;
  if (rel_residual) *rel_residual = bnorm2 > 0 ? std::sqrt(rr / bnorm2) : 0.0;
  if (rr > tol2 || !(rr == rr)) {
    set_error("Conjugate gradient has not converged. After %lld steps, the mixed-precision solver FAILED with |r|/|b| = %g, cg_tol = %g",
              (long long)its, std::sqrt(rr / bnorm2), tol);
    return MUGPU_ECONVERGENCE;
  }
  return MUGPU_OK;
}

int check_mixed_cell(mugpu_cell c) {
  if (!c) return MUGPU_EARG;
  if (c->prec != MUGPU_FP32 || !c->tan32) {
    set_error("the mixed-precision solve needs a cell created with MUGPU_FP32 and an evaluated tangent");
    return MUGPU_ERUNTIME;
  }
  return MUGPU_OK;
}

// ==================== mixed precision CG with FP64 refinement ===================
extern "C" int mugpu_cg_solve_mixed(mugpu_cell c, const void* d_rhs, void* d_x, double tol,
                                    int maxiter, double inner_tol, int64_t* iterations,
                                    int* refinements, double* rel_residual) {
  MUGPU_TRY(check_mixed_cell(c));
  const int64_t n = c->N * c->nc;
  MUGPU_TRY(ensure_mixed(c, n));
  MixedVecs v;
  v.x32 = c->x32;
  v.r32 = c->r32;
  v.p32 = c->p32;
  v.Ap32 = c->Ap32;
  v.r64 = c->r64;
  v.ctl = c->ctl;
  return solve_mixed(c, v, d_rhs, d_x, tol, maxiter, inner_tol, iterations, refinements,
                     rel_residual);
}

// ============ mixed-precision CG without device allocation ============
// The work vectors of solve_mixed carved out of the per-device `scratch`
// array, 256-byte aligned each: control words, x32, r32, p32, Ap32 (FP32,
// n values each) and r64 (FP64, n values).
static size_t noalloc_bytes(int64_t n) {
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  return up(sizeof(int) * CTL_COUNT) + 4 * up(sizeof(float) * (size_t)n) + up(sizeof(double) * (size_t)n);
}

extern "C" int mugpu_cg_solve_mixed_noalloc(mugpu_cell c, const void* d_rhs, void* d_x, double tol,
                                            int maxiter, double inner_tol, int64_t* iterations,
                                            int* refinements, double* rel_residual) {
  if (!c) return MUGPU_EARG;
  const int64_t n = c->N * c->nc;
  if (n > MUGPU_NOALLOC_MAX_VALUES || noalloc_bytes(n) > MUGPU_NOALLOC_SCRATCH_BYTES) {
    set_error("mugpu_cg_solve_mixed_noalloc: the cell has %lld values (pixels x components); the "
              "allocation-free solve holds at most %lld in its fixed work array of %lld bytes",
              (long long)n, (long long)MUGPU_NOALLOC_MAX_VALUES,
              (long long)MUGPU_NOALLOC_SCRATCH_BYTES);
    return MUGPU_EARG;
  }
  // the library is initialised (the cell exists); the work array's address
  // was looked up for this device by mugpu_init
  if (g_device < 0 || g_device >= kMaxDevices || !g_scratch_ptr[g_device]) {
    set_error("mugpu_cg_solve_mixed_noalloc: the work array of device %d has no address", g_device);
    return MUGPU_ERUNTIME;
  }
  MUGPU_TRY(check_mixed_cell(c));
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[g_device]);
  MixedVecs v;
  v.ctl = reinterpret_cast<int*>(base);
  base += up(sizeof(int) * CTL_COUNT);
  float** f32[4] = {&v.x32, &v.r32, &v.p32, &v.Ap32};
  for (float** f : f32) {
    *f = reinterpret_cast<float*>(base);
    base += up(sizeof(float) * (size_t)n);
  }
  v.r64 = reinterpret_cast<double*>(base);
  // every vector is written before it is read (x32, r32, p32 by the inner
  // start, Ap32 by the operator, r64 by the residual, ctl by set_ctl), so
  // what an earlier call left in the array does not matter
  return solve_mixed(c, v, d_rhs, d_x, tol, maxiter, inner_tol, iterations, refinements,
                     rel_residual);
}
