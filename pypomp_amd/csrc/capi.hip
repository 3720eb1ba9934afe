// extern "C" entry points of libmipomp.so (declared in include/mipomp.h):
// argument checks, workspace carving, launch sequencing.  No C++ exception
// crosses the boundary; errors become status codes plus a per-thread message.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mipomp.h"
#include "engine.hpp"
#include "mech.hpp"

namespace mp {
hipError_t launch_bits(const uint32_t* keys, int64_t nkeys, int64_t n, void* out, int mode, hipStream_t s);
hipError_t launch_resample_rows(const float* logw, const uint32_t* keys, int64_t R, int J, int32_t* anc,
                                uint64_t* C, hipStream_t s);
hipError_t launch_batched(int algo, const BatchArgs& a, int block, int grid, size_t lds, hipStream_t st);
hipError_t launch_sched(int algo, const uint32_t* keys, const SeriesSet& ser, const int32_t* unit,
                        int64_t sched_len, uint32_t* sched, int R, hipStream_t st);
hipError_t plain_kernel(int which, const PlainArgs& a, int ntiles, int threads, hipStream_t s);
hipError_t plain_prepare(const ModelDesc& md, const float* theta, int64_t stride, int R, int flags, char* out,
                         int64_t pstride, hipStream_t s);
hipError_t plain_reset(const PlainArgs& a, hipStream_t s);
size_t batched_small_lds(int J, int block);
hipError_t launch_order(const SeriesSet& ser, const int32_t* unit, int R, int32_t* order, int32_t* queue,
                        hipStream_t st);
}  // namespace mp

namespace mp {
hipError_t launch_mop_grad(const GradArgs& a, int block, int grid, hipStream_t st);
int64_t mop_grad_scratch_bytes(int S, int nz, int J, int T, int K, bool lds_work);
int mop_grad_noise_words(const ModelDesc& md);
bool mop_grad_supported(const ModelDesc& md);
hipError_t launch_simulate(const SimArgs& a, hipStream_t st);
size_t mop_grad_lds_bytes(int S, int J, int block, bool lds_work);
hipError_t launch_mop_hess(const HessArgs& a, int grid, hipStream_t st);
hipError_t launch_valu_probe(int kind, int64_t iters, int blocks, float* out, hipStream_t s);
bool mop_hess_supported(const ModelDesc& md);
}  // namespace mp

using namespace mp;

// the timing events the next mp_launch attaches to its dispatch (engine.hpp)
thread_local LaunchTiming mp::g_launch_timing;

// Scratch of the no-allocation entry point (mp_pfilter_noalloc): one copy per
// GPU, held from the library's load until the process ends.  In host code
// `scratch` names a host-side stand-in, not device memory: the device address
// comes from hipGetSymbolAddress, once per device, and only that address is
// handed to kernels.
static __device__ __align__(256) char scratch[MP_NOALLOC_SCRATCH_BYTES];

static_assert(sizeof(mp_model_desc) == sizeof(ModelDesc), "model desc layout");
static_assert(sizeof(mp_series_desc) == sizeof(SeriesSet), "series desc layout");

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(MP_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

#define MP_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

// kParamSlot (engine.hpp): bytes per filter of prepared model parameters

// Launch-shape parameters (set by the autotuner through mp_tune_set).
// Defaults from a sweep on LG 2-D, J = 1e6 (2.86e10 particle-steps/s).
// Plain-form launch shapes; 0 = automatic (plain_shape below).
int g_plain_scan_rows = 0;     // particles per scan thread (1..16)
int g_plain_prop_rows = 0;     // particle rows per propagate block (1..16)
int g_plain_scan_threads = 0;  // scan workgroup size (256, 512, 1024)
int g_plain_prop_threads = 0;  // propagate workgroup size (256, 512, 1024)
// scan tiles whose ends every propagate block forms in LDS (above it, or at 0,
// a k_plain_ends launch per step writes them to global memory)
int g_plain_lds_tiles = kMaxLdsTiles;
// propagate blocks take the chunks of the scan tiles on their own XCD
int g_plain_xcd_map = 1;
// Replicate filters the plain form runs at once, each with its own buffers on
// its own stream (1..4): the latency-bound scan of one overlaps the VALU-bound
// propagate of another.
constexpr int kMaxLanes = 4;
int g_plain_lanes = 4;
int g_batched_block = 0;    // batched workgroup size: 0 = automatic, 256, 512, 768
int g_grad_block = 0;       // MOP-adjoint workgroup size: 0 = automatic, 256, 512

// Optional per-kernel timing (mp_profile_*): the events are attached to the
// kernel's own dispatch packet (hipExtLaunchKernel through mp_launch), so they
// time that kernel alone.  Kinds: 0 plain propagate, 1 plain scan, 2 batched
// filter, 3 key schedule, 4 VALU calibration probe.
struct ProfRec {
  int kind;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t take_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Set while a no-allocation entry point runs: its launches are not timed,
// because timing would create HIP events.
thread_local bool g_no_alloc = false;

constexpr int kMaxDevices = 64;

// Runs `launch` with the timing events handed to its kernel launch when
// profiling is on; a launch that does not take them is not timed.
template <class F>
hipError_t timed(int kind, hipStream_t, F&& launch) {
  if (!g_prof_on || g_no_alloc) return launch();
  std::lock_guard<std::mutex> lk(g_prof_mu);
  ProfRec r{kind, take_event(), take_event()};
  if (!r.a || !r.b) return launch();
  g_launch_timing = LaunchTiming{r.a, r.b};
  const hipError_t e = launch();
  const bool used = g_launch_timing.start == nullptr;
  g_launch_timing = LaunchTiming{};
  if (e == hipSuccess && used) {
    g_prof.push_back(r);
  } else {
    g_event_pool.push_back(r.a);
    g_event_pool.push_back(r.b);
  }
  return e;
}

int max_lds_bytes() {
  static int v = -1;
  if (v < 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return 65536;
    v = (int)p.sharedMemPerBlock;
  }
  return v;
}

int num_cus() {
  static int v = -1;
  if (v < 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      n = 256;
    v = n;
  }
  return v;
}

int check_model(const mp_model_desc* m) {
  if (!m) return fail(MP_ERR_ARG, "model descriptor is NULL");
  switch (m->kind) {
    case MP_MODEL_LG:
      if (m->dx < 1 || m->dx > kLgMaxDim || m->dy < 1 || m->dy > kLgMaxDim)
        return fail(MP_ERR_ARG, "LG dimensions must be 1..%d (dx=%d, dy=%d)", kLgMaxDim, m->dx, m->dy);
      if (m->S != m->dx) return fail(MP_ERR_ARG, "LG: S must equal dx");
      {
        // A (dx x dx), C (dy x dx), the Q and R triangles, X0 and, with a
        // separate initial covariance, its triangle (linear_gaussian.py:53-92)
        const int dx = m->dx, dy = m->dy, tq = dx * (dx + 1) / 2;
        const int d = dx * dx + dy * dx + tq + dy * (dy + 1) / 2 + dx + (m->lg_sep_init ? tq : 0);
        if (m->d != d) return fail(MP_ERR_ARG, "LG(%d, %d) has %d parameters, the descriptor says %d", dx, dy, d, m->d);
      }
      break;
    case MP_MODEL_SV:
      if (m->S != 2 || m->d != 6 || m->dy != 1) return fail(MP_ERR_ARG, "SV model expects S=2, d=6, dy=1");
      break;
    case MP_MODEL_SEIR:
      if (m->S != 5 || m->d != 10 || m->dy != 1) return fail(MP_ERR_ARG, "SEIR model expects S=5, d=10, dy=1");
      break;
    case MP_MODEL_SPX:
      if (m->S != 2 || m->d != 6 || m->dy != 1) return fail(MP_ERR_ARG, "SPX model expects S=2, d=6, dy=1");
      break;
    default:
      return fail(MP_ERR_ARG, "unknown model kind %d", m->kind);
  }
  if (m->vectorized && m->kind != MP_MODEL_SPX)
    return fail(MP_ERR_ARG, "model kind %d has no @vectorized rproc", m->kind);
  return MP_OK;
}

// covariates a model reads per grid row
int model_covariates(const mp_model_desc* m) { return m->kind == MP_MODEL_SPX ? SPX::DC : 0; }

int check_series(const mp_series_desc* s) {
  if (!s) return fail(MP_ERR_ARG, "series descriptor is NULL");
  if (s->U < 1 || s->Tmax < 1) return fail(MP_ERR_ARG, "series needs U >= 1 and Tmax >= 1");
  if (!s->ys || !s->times || !s->nstep || !s->tidx0 || !s->dt_ext || !s->obs_off || !s->dt_off)
    return fail(MP_ERR_ARG, "series descriptor has a NULL array");
  if (s->dc < 0) return fail(MP_ERR_ARG, "negative covariate count (dc = %d)", s->dc);
  if (s->dc > 0 && (!s->covars || !s->cv_off)) return fail(MP_ERR_ARG, "dc = %d but covars or cv_off is NULL", s->dc);
  if (s->max_nstep < 0) return fail(MP_ERR_ARG, "negative max_nstep");
  return MP_OK;
}

// Batched workgroup size.  Filters of >= 1536 particles take 768 threads (12
// waves: a CU's full share at the kernels' ~168 VGPRs), which measured best
// for both one filter per CU (SEIR, J = 20000, 256 filters: 1.7x over 256
// threads) and several (SEIR panel, J = 2000, 800 filters: 1.2x).  Smaller
// filters aim at ~12 waves per CU from several workgroups.  mp_tune_set
// overrides the choice.
int batched_block(int R, int J) {
  if (g_batched_block) return g_batched_block;
  if (J >= 1536) return 768;
  const int per_cu = (R + num_cus() - 1) / num_cus();
  int nt = per_cu >= 3 ? 256 : per_cu == 2 ? 512 : 768;
  while (nt > 256 && nt / 2 >= J) nt -= 256;  // at least two particles per thread
  return nt;
}

struct BatchedPlan {
  int W, block;
  int64_t sched_len, sched_bytes, order_bytes;
  int64_t per_filter;  // global scratch per filter (0 when LDS-resident)
  bool lds_resident;
  size_t lds;
  size_t total;
};

BatchedPlan plan_batched(int algo, const mp_model_desc* m, const mp_series_desc* s, int R, int J) {
  BatchedPlan p{};
  const SeriesSet& ser = *reinterpret_cast<const SeriesSet*>(s);
  // a perturbed filter also carries d parameters and one ancestry word
  p.W = m->S + (algo == ALGO_PER ? m->d + 1 : 0);
  p.sched_len = sched_stride(algo, ser);
  p.sched_bytes = align256((int64_t)R * p.sched_len * 8);
  p.order_bytes = align256((int64_t)R * 4) + 256;  // filter order + queue counter
  const FilterScratch L = filter_scratch_layout(p.W, J);
  // LG of dimension 5..8 has 256-thread workgroups only (a smaller build).
  // 1024 threads only where the kernel fits 128 VGPRs: SEIR with its
  // particle-independent tables (pfilter and MOP, not IF2's per-particle
  // parameters); automatic for the SEIR pfilter of >= 1536 particles in place
  // of 768 (4.43e9 -> 4.76e9 particle-steps/s on 256 filters of 20000).
  p.block = (m->kind == MP_MODEL_LG && (m->dx > 4 || m->dy > 4)) ? 256 : batched_block(R, J);
  const bool seir_tab = m->kind == MP_MODEL_SEIR && algo != ALGO_PER;
  if (!g_batched_block && p.block == 768 && seir_tab && algo == ALGO_PF && J >= 1536) p.block = 1024;
  if (p.block == 1024 && !seir_tab) p.block = 768;
  const size_t small = batched_small_lds(J, p.block);
  // Keep filters LDS-resident when two workgroups still fit per CU.
  p.lds_resident = (int64_t)small + L.total <= (int64_t)max_lds_bytes() / 2;
  p.per_filter = p.lds_resident ? 0 : L.total;
  p.lds = small + (p.lds_resident ? (size_t)L.total : 0);
  p.total = (size_t)p.sched_bytes + (size_t)p.order_bytes + (size_t)R * (size_t)p.per_filter + 256;
  return p;
}

struct PlainPlan {
  int ntiles, rows, tile_len, nprop, rows_prop, prop_len, scan_threads, prop_threads, xcd_q, win;
  int ndiag, lanes;
  bool global_ends;
  // off_x0 .. off_diag: lane 0's buffers; lane l's sit lane_stride * l further
  int64_t off_sched, off_params, off_x0, off_x1, off_lw, off_c, off_agg, off_ends, off_pmax, off_pmin, off_mm,
      off_steps, off_st, off_diag, lane_stride, total;
};

// Lanes of a plain-form call of R filters: the setting, at most R, and one in
// the no-allocation entry point (the other lanes' streams are created on first
// use).
int plain_lanes(int R) { return g_no_alloc ? 1 : std::min(g_plain_lanes, R); }

PlainPlan plan_plain(const mp_model_desc* m, const mp_series_desc* s, int R, int J) {
  PlainPlan p{};
  p.lanes = plain_lanes(R);
  // Automatic shapes, measured on config 2 (J = 1e6): alone or in two lanes,
  // 4096-particle scan tiles of 1024 threads x 4 and propagate blocks of
  // 512 x 2; with three or four lanes in flight, smaller workgroups that fit
  // beside the other lanes' kernels: scan 256 x 16, propagate 256 x 4
  // (4 lanes: 5.6e10 -> 6.7e10 particle-steps/s; 1 lane: 3.4e10 -> 3.1e10).
  const bool wide = p.lanes >= 3;
  p.scan_threads = g_plain_scan_threads ? g_plain_scan_threads : (wide ? 256 : 1024);
  p.prop_threads = g_plain_prop_threads ? g_plain_prop_threads : (wide ? 256 : 512);
  p.rows = g_plain_scan_rows ? g_plain_scan_rows : (wide ? 16 : 4);
  // Past kMaxLdsTiles tiles the scan threads take more particles each (up to
  // 16), so that propagate blocks can still form the tile ends in LDS.
  const int lds_cap = std::min(g_plain_lds_tiles, kMaxLdsTiles);
  while (p.rows < 16 && (int64_t)J > (int64_t)lds_cap * p.rows * p.scan_threads && lds_cap > 0) p.rows *= 2;
  if (p.rows > 16) p.rows = 16;
  const int tile = p.rows * p.scan_threads;
  p.tile_len = tile;
  p.ntiles = (J + tile - 1) / tile;
  p.global_ends = p.ntiles > lds_cap;
  p.rows_prop = g_plain_prop_rows ? g_plain_prop_rows : (wide ? 4 : 2);
  p.nprop = (J + p.rows_prop * p.prop_threads - 1) / (p.rows_prop * p.prop_threads);
  // Whole rounds of resident workgroups (8 waves per SIMD: 2048 threads per
  // CU): past half a round the block count is rounded up to a multiple of a
  // round and the particles split evenly, so no CU runs one workgroup more
  // than another.  A trailing block may be left without particles.
  const int round = num_cus() * (2048 / p.prop_threads);
  if (p.nprop > round / 2) p.nprop = (p.nprop + round - 1) / round * round;
  p.prop_len = (int)(((int64_t)J + p.nprop - 1) / p.nprop);
  // XCD-local chunks.  Workgroups are dispatched to the XCDs round-robin by
  // block index, and scan tile k runs on XCD k % 8.  When a scan tile holds a
  // whole number q of full propagate blocks, the propagate blocks take
  // block-sized chunks in an order that puts the q chunks of tile k on XCD
  // k % 8 as well, so the scan reads the log-weights and the next step reads
  // prefix sums and states from the L2 that holds them.  The block count is
  // then a multiple of 8 q (trailing blocks may be empty).
  p.xcd_q = 0;
  const int cap = p.rows_prop * p.prop_threads;
  if (g_plain_xcd_map && tile % cap == 0) {
    const int q = tile / cap, G = kXcds * q;
    const int n = (J + cap - 1) / cap;
    if (n >= G) {
      p.xcd_q = q;
      p.nprop = (n + G - 1) / G * G;
      p.prop_len = cap;
    }
  }
  // LDS ancestor window: twice a block's particles, 2048 to 4096 entries (16
  // to 32 KiB), so that a block whose comb spans up to that many parents
  // searches in LDS (a wider span searches global memory)
  p.win = std::min(4096, std::max(2048, 2 * p.rows_prop * p.prop_threads));
  const SeriesSet& ser = *reinterpret_cast<const SeriesSet*>(s);
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    int64_t r = o;
    o += align256(bytes);
    return r;
  };
  p.off_sched = take((int64_t)R * sched_stride(ALGO_PF, ser) * 8);  // all filters' key schedules
  p.off_params = take((int64_t)R * kParamSlot);                      // prepared parameters
  const int64_t lane0 = o;
  p.off_x0 = take((int64_t)m->S * J * 4);
  p.off_x1 = take((int64_t)m->S * J * 4);
  p.off_lw = take((int64_t)J * 4);
  p.off_c = take((int64_t)J * 8);
  p.off_agg = take((int64_t)p.ntiles * 8);
  p.off_ends = p.global_ends ? take((int64_t)p.ntiles * 8) : -1;
  p.off_pmax = take((int64_t)p.nprop * 4);
  p.off_pmin = take((int64_t)p.nprop * 4);
  p.off_mm = take(4 * sizeof(float));
  p.off_steps = take((int64_t)std::max(1, ser.Tmax) * (int64_t)sizeof(PlainStep));
  p.off_st = take(sizeof(PlainState));
  // step diagnostics: one block per 1024 particles, at most 1024 blocks
  p.ndiag = (int)std::min<int64_t>(1024, ((int64_t)J + 1023) / 1024);
  p.off_diag = take((int64_t)p.ndiag * kDiagWords * 8);
  p.lane_stride = o - lane0;
  p.total = lane0 + p.lanes * p.lane_stride + 256;
  return p;
}

char* align_ptr(void* p) { return reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(p) + 255) & ~uintptr_t(255)); }

int run_batched(int algo, const mp_model_desc* m, const mp_series_desc* s, BatchArgs a, void* ws, size_t ws_bytes,
                hipStream_t st) {
  const BatchedPlan p = plan_batched(algo, m, s, a.R, a.J);
  if (ws_bytes < p.total) return fail(MP_ERR_WORKSPACE, "workspace %zu < required %zu bytes", ws_bytes, p.total);
  char* base = align_ptr(ws);
  uint32_t* sched = reinterpret_cast<uint32_t*>(base);
  a.sched = sched;
  a.sched_len = p.sched_len;
  int32_t* order = reinterpret_cast<int32_t*>(base + p.sched_bytes);
  a.order = order;
  a.queue = reinterpret_cast<int32_t*>(base + p.sched_bytes + p.order_bytes - 256);
  a.ws = base + p.sched_bytes + p.order_bytes;
  a.ws_stride = p.per_filter;
  a.lds_resident = p.lds_resident;
  MP_HIP(timed(3, st, [&] { return launch_order(a.ser, a.unit, a.R, order, a.queue, st); }));
  MP_HIP(timed(3, st, [&] { return launch_sched(algo, a.keys, a.ser, a.unit, p.sched_len, sched, a.R, st); }));
  const int per_cu = p.lds_resident ? 2 : 1;
  int grid = a.R < num_cus() * per_cu * 4 ? a.R : num_cus() * per_cu * 4;
  MP_HIP(timed(2, st, [&] { return launch_batched(algo, a, p.block, grid, p.lds, st); }));
  return MP_OK;
}

// The streams of the plain form's lanes 1.. (lane 0 is the caller's stream)
// and the events that fork them from the caller's stream and join them back:
// one set per host thread and device, default priority, created on first use
// and kept until the process ends.
struct LaneSync {
  hipStream_t s[kMaxLanes] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kMaxLanes] = {};
};

int lane_sync(int lanes, LaneSync** out) {
  static thread_local LaneSync tl[kMaxDevices];
  int dev = -1;
  MP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(MP_ERR_ARG, "device index %d out of range", dev);
  LaneSync& L = tl[dev];
  if (!L.fork) MP_HIP(hipEventCreateWithFlags(&L.fork, hipEventDisableTiming));
  for (int l = 1; l < lanes; ++l) {
    if (!L.s[l]) MP_HIP(hipStreamCreateWithFlags(&L.s[l], hipStreamNonBlocking));
    if (!L.join[l]) MP_HIP(hipEventCreateWithFlags(&L.join[l], hipEventDisableTiming));
  }
  *out = &L;
  return MP_OK;
}

int run_plain_pf(const mp_model_desc* m, const mp_series_desc* s, const float* theta, int64_t theta_stride,
                 const uint32_t* keys, int R, int J, float logthresh, int flags, double* loglik, float* cll,
                 float* ess, float* fmean, float* pmean, void* ws, size_t ws_bytes, hipStream_t st) {
  if (s->U != 1) return fail(MP_ERR_ARG, "the plain form runs single-series filters (U = 1)");
  const PlainPlan p = plan_plain(m, s, R, J);
  if (ws_bytes < (size_t)p.total) return fail(MP_ERR_WORKSPACE, "workspace %zu < required %lld bytes", ws_bytes,
                                              (long long)p.total);
  char* b = align_ptr(ws);
  PlainArgs a{};
  a.md = *reinterpret_cast<const ModelDesc*>(m);
  a.ser = *reinterpret_cast<const SeriesSet*>(s);
  a.unit = 0;
  a.J = J;
  a.T = s->Tmax;
  a.flags = flags;
  a.logthresh = logthresh;
  a.ntiles = p.ntiles;
  a.tile_len = p.tile_len;
  a.tile_shift = __builtin_ctz((unsigned)p.tile_len);
  a.rows_prop = p.rows_prop;
  a.prop_len = p.prop_len;
  a.nprop = p.nprop;
  a.xcd_q = p.xcd_q;
  a.win = p.win;
  a.logJ = d_log((float)J);
  a.log1J = d_log(1.0f / (float)J);
  a.sh = shift_for(J);
  a.rows = p.rows;
  a.ndiag = p.ndiag;
  const bool diag = (flags & (MP_ESS | MP_FILTER_MEAN | MP_PREDICTION_MEAN)) != 0;
  const int64_t slen = sched_stride(ALGO_PF, a.ser);
  uint32_t* sched_all = reinterpret_cast<uint32_t*>(b + p.off_sched);
  MP_HIP(timed(3, st, [&] { return launch_sched(ALGO_PF, keys, a.ser, nullptr, slen, sched_all, R, st); }));
  char* params = b + p.off_params;
  MP_HIP(plain_prepare(a.md, theta, theta_stride, R, flags, params, kParamSlot, st));
  // each lane's buffers
  PlainArgs A[kMaxLanes] = {a, a, a, a};
  for (int l = 0; l < p.lanes; ++l) {
    char* lb = b + l * p.lane_stride;
    PlainArgs& x = A[l];
    x.X0 = reinterpret_cast<float*>(lb + p.off_x0);
    x.X1 = reinterpret_cast<float*>(lb + p.off_x1);
    x.lw = reinterpret_cast<float*>(lb + p.off_lw);
    x.C = reinterpret_cast<uint64_t*>(lb + p.off_c);
    x.agg = reinterpret_cast<uint64_t*>(lb + p.off_agg);
    x.ends = p.global_ends ? reinterpret_cast<uint64_t*>(lb + p.off_ends) : nullptr;
    x.pmax = reinterpret_cast<float*>(lb + p.off_pmax);
    x.pmin = reinterpret_cast<float*>(lb + p.off_pmin);
    x.mm = reinterpret_cast<float*>(lb + p.off_mm);
    x.steps = reinterpret_cast<PlainStep*>(lb + p.off_steps);
    x.st = reinterpret_cast<PlainState*>(lb + p.off_st);
    x.diag_part = reinterpret_cast<double*>(lb + p.off_diag);
  }
  hipStream_t ls[kMaxLanes] = {st, st, st, st};
  LaneSync* L = nullptr;
  if (p.lanes > 1) {
    int rc = lane_sync(p.lanes, &L);
    if (rc) return rc;
    MP_HIP(hipEventRecord(L->fork, st));
    for (int l = 1; l < p.lanes; ++l) {
      ls[l] = L->s[l];
      MP_HIP(hipStreamWaitEvent(ls[l], L->fork, 0));
    }
  }
  // Filters f0 .. f0 + nl - 1 run together, filter f0 + l in lane l; each
  // observation step is the propagate and scan launches of every lane.
  auto body = [&]() -> int {
    for (int f0 = 0; f0 < R; f0 += p.lanes) {
      const int nl = std::min(p.lanes, R - f0);
      for (int l = 0; l < nl; ++l) {
        const int f = f0 + l;
        PlainArgs& x = A[l];
        x.f = f;
        x.theta = theta + f * theta_stride;
        x.fkey = keys + 2 * f;
        x.loglik_out = loglik + f;
        x.cll = cll ? cll + (int64_t)f * s->Tmax : nullptr;
        x.ess = (flags & MP_ESS) ? ess + (int64_t)f * s->Tmax : nullptr;
        x.fmean = (flags & MP_FILTER_MEAN) ? fmean + (int64_t)f * s->Tmax * m->S : nullptr;
        x.pmean = (flags & MP_PREDICTION_MEAN) ? pmean + (int64_t)f * s->Tmax * m->S : nullptr;
        x.sched = sched_all + 2 * (int64_t)f * slen;
        x.params = params + (int64_t)f * kParamSlot;
        MP_HIP(plain_reset(x, ls[l]));
      }
      for (int t = 0; t < a.T; ++t) {
        for (int l = 0; l < nl; ++l) {
          PlainArgs& x = A[l];
          const hipStream_t q = ls[l];
          x.t = t;
          MP_HIP(timed(0, q, [&] { return plain_kernel(0, x, p.ntiles, p.prop_threads, q); }));
          MP_HIP(timed(1, q, [&] { return plain_kernel(1, x, p.ntiles, p.scan_threads, q); }));
          if (diag) MP_HIP(plain_kernel(2, x, p.ntiles, 256, q));
        }
      }
      for (int l = 0; l < nl; ++l) MP_HIP(plain_kernel(3, A[l], p.ntiles, 64, ls[l]));
    }
    return MP_OK;
  };
  const int rc = body();
  // join even after a failed launch, so the caller's stream still orders
  // after whatever the other lanes have queued
  for (int l = 1; l < p.lanes; ++l) {
    const hipError_t e1 = hipEventRecord(L->join[l], ls[l]);
    const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(st, L->join[l], 0) : e1;
    if (rc == MP_OK && e2 != hipSuccess) return hip_fail(e2, "joining a plain-form lane");
  }
  return rc;
}

BatchArgs base_args(const mp_model_desc* m, const mp_series_desc* s, const uint32_t* keys, const int32_t* unit,
                    int R, int J) {
  BatchArgs a{};
  a.md = *reinterpret_cast<const ModelDesc*>(m);
  a.ser = *reinterpret_cast<const SeriesSet*>(s);
  a.unit = unit;
  a.keys = keys;
  a.R = R;
  a.J = J;
  return a;
}

float log_thresh(float thresh) { return thresh > 0.f ? d_log(thresh) : -__builtin_inff(); }

struct GradPlan {
  int64_t sched_len, sched_bytes, per_filter;
  int block;
  bool lds_work;
  size_t total;
};

GradPlan plan_grad(const mp_model_desc* m, const mp_series_desc* s, int R, int J, int K) {
  GradPlan p{};
  const SeriesSet& ser = *reinterpret_cast<const SeriesSet*>(s);
  p.sched_len = sched_stride(ALGO_MOP, ser);
  p.sched_bytes = align256((int64_t)R * p.sched_len * 8);
  // 512 threads at <= 128 VGPRs (two filters per CU at 4 waves per SIMD)
  // measured 1.23x over 256 threads on SV, J = 1000, 512 filters
  p.block = g_grad_block ? g_grad_block : (J >= 512 ? 512 : 256);
  // per-step arrays in LDS when two workgroups still fit per CU
  p.lds_work = mop_grad_lds_bytes(m->S, J, p.block, true) <= (size_t)max_lds_bytes() / 2;
  p.per_filter = align256(mop_grad_scratch_bytes(
      m->S, mop_grad_noise_words(*reinterpret_cast<const ModelDesc*>(m)), J, s->Tmax, K, p.lds_work));
  p.total = (size_t)p.sched_bytes + (size_t)R * (size_t)p.per_filter + 256;
  return p;
}

int run_mop_grad(const mp_model_desc* m, const mp_series_desc* s, const float* theta, int64_t theta_stride,
                 const uint32_t* keys, const int32_t* unit, int R, int J, float alpha, int K, double* loglik,
                 float* grad, void* ws, size_t ws_bytes, hipStream_t st) {
  const GradPlan p = plan_grad(m, s, R, J, K);
  if (ws_bytes < p.total) return fail(MP_ERR_WORKSPACE, "workspace %zu < required %zu bytes", ws_bytes, p.total);
  char* base = align_ptr(ws);
  GradArgs a{};
  a.md = *reinterpret_cast<const ModelDesc*>(m);
  a.ser = *reinterpret_cast<const SeriesSet*>(s);
  a.unit = unit;
  a.theta = theta;
  a.theta_stride = theta_stride;
  a.keys = keys;
  a.sched = reinterpret_cast<uint32_t*>(base);
  a.sched_len = p.sched_len;
  a.R = R;
  a.J = J;
  a.K = K;
  a.alpha = alpha;
  a.loglik = loglik;
  a.grad = grad;
  a.ws = base + p.sched_bytes;
  a.ws_stride = p.per_filter;
  a.lds_work = p.lds_work;
  MP_HIP(timed(3, st, [&] { return launch_sched(ALGO_MOP, keys, a.ser, unit, p.sched_len,
                                                const_cast<uint32_t*>(a.sched), R, st); }));
  const int grid = R < num_cus() * 8 ? R : num_cus() * 8;
  hipError_t e = timed(2, st, [&] { return launch_mop_grad(a, p.block, grid, st); });
  if (e == hipErrorNotSupported) return fail(MP_ERR_ARG, "model kind %d has no adjoint (not differentiable)", m->kind);
  MP_HIP(e);
  return MP_OK;
}

// Device address of `scratch` on the current device, looked up on the first
// call made on that device and kept per device.
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void** out) {
  int dev = -1;
  MP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(MP_ERR_ARG, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(MP_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MP_OK;
}

// MOP Hessian: key schedules of the R rows, then one scratch slice per
// resident workgroup (the grid strides over the (row, pair) tasks).
struct HessPlan {
  int64_t sched_len, sched_bytes, per_block;
  int grid;
  size_t total;
};

HessPlan plan_hess(const mp_model_desc* m, const mp_series_desc* s, int R, int J) {
  HessPlan p{};
  const SeriesSet& ser = *reinterpret_cast<const SeriesSet*>(s);
  p.sched_len = sched_stride(ALGO_MOP, ser);
  p.sched_bytes = align256((int64_t)R * p.sched_len * 8);
  p.per_block = hess_scratch_layout(m->S, J).total;
  const int64_t tasks = (int64_t)R * (m->d * (m->d + 1) / 2);
  const int64_t cap = (int64_t)num_cus() * 4;
  p.grid = (int)(tasks < cap ? tasks : cap);
  p.total = (size_t)p.sched_bytes + (size_t)p.grid * (size_t)p.per_block + 256;
  return p;
}

// the model's from_est run on the host CPU (detmath is host-and-device code)
template <class M>
void host_from_est_rows(const ModelDesc& md, float* theta, int64_t n) {
  for (int64_t r = 0; r < n; ++r) {
    float th[M::MAXP];
    float* row = theta + r * md.d;
    for (int p = 0; p < md.d; ++p) th[p] = row[p];
    M::from_est(md, th);
    for (int p = 0; p < md.d; ++p) row[p] = th[p];
  }
}

}  // namespace

extern "C" {

const char* mp_last_error(void) { return g_err.c_str(); }
int32_t mp_abi_version(void) { return 4; }

int32_t mp_host_threefry2x32(const uint32_t* h_keys, const uint32_t* h_ctr, uint32_t* h_out, int64_t n) {
  if (n < 0 || (n > 0 && (!h_keys || !h_ctr || !h_out))) return fail(MP_ERR_ARG, "bad threefry arguments");
  for (int64_t i = 0; i < n; ++i) {
    uint32_t x0 = h_ctr[2 * i], x1 = h_ctr[2 * i + 1];
    threefry2x32(h_keys[2 * i], h_keys[2 * i + 1], x0, x1);
    h_out[2 * i] = x0;
    h_out[2 * i + 1] = x1;
  }
  return MP_OK;
}

int32_t mp_host_split(const uint32_t* h_key, int64_t n, uint32_t* h_out) {
  if (n < 0 || !h_key || (n > 0 && !h_out)) return fail(MP_ERR_ARG, "bad split arguments");
  const Key k{h_key[0], h_key[1]};
  for (int64_t i = 0; i < n; ++i) {
    Key s = split_at(k, (uint64_t)i);
    h_out[2 * i] = s.k0;
    h_out[2 * i + 1] = s.k1;
  }
  return MP_OK;
}

static int32_t bits_common(const uint32_t* keys, int64_t nkeys, int64_t n, void* out, int mode, void* stream) {
  if (nkeys < 0 || n < 0) return fail(MP_ERR_ARG, "negative size");
  if (nkeys == 0 || n == 0) return MP_OK;
  if (!keys || !out) return fail(MP_ERR_ARG, "NULL pointer");
  MP_HIP(launch_bits(keys, nkeys, n, out, mode, (hipStream_t)stream));
  return MP_OK;
}

int32_t mp_random_bits(const uint32_t* keys, int64_t nkeys, int64_t n, uint32_t* out, void* stream) {
  return bits_common(keys, nkeys, n, out, 0, stream);
}
int32_t mp_uniform(const uint32_t* keys, int64_t nkeys, int64_t n, float* out, void* stream) {
  return bits_common(keys, nkeys, n, out, 1, stream);
}
int32_t mp_normal(const uint32_t* keys, int64_t nkeys, int64_t n, float* out, void* stream) {
  return bits_common(keys, nkeys, n, out, 2, stream);
}
int32_t mp_split(const uint32_t* keys, int64_t nkeys, int64_t n, uint32_t* out, void* stream) {
  return bits_common(keys, nkeys, n, out, 3, stream);
}

size_t mp_resample_workspace_size(int64_t R, int64_t J) {
  if (R <= 0 || J <= 0) return 256;
  return (size_t)(R * J * 8) + 256;
}

int32_t mp_resample_systematic(const float* logw, const uint32_t* keys, int64_t R, int64_t J, int32_t* ancestors,
                               void* workspace, size_t ws_bytes, void* stream) {
  if (R < 0 || J < 1 || J > (int64_t)1 << 30) return fail(MP_ERR_ARG, "need R >= 0 and 1 <= J <= 2^30");
  if (R == 0) return MP_OK;
  if (!logw || !keys || !ancestors || !workspace) return fail(MP_ERR_ARG, "NULL pointer");
  if (J > 65536) return fail(MP_ERR_ARG, "mp_resample_systematic handles rows of at most 65536 weights");
  if (ws_bytes < mp_resample_workspace_size(R, J)) return fail(MP_ERR_WORKSPACE, "workspace too small");
  MP_HIP(launch_resample_rows(logw, keys, R, (int)J, ancestors, reinterpret_cast<uint64_t*>(align_ptr(workspace)),
                              (hipStream_t)stream));
  return MP_OK;
}

size_t mp_filter_workspace_size(int32_t algo, const mp_model_desc* h_model, const mp_series_desc* h_series,
                                int32_t R, int32_t J, int32_t form) {
  if (!h_model || !h_series || R < 1 || J < 1) return 0;
  // device queries are cached here, before any stream capture of the launches
  (void)num_cus();
  (void)max_lds_bytes();
  if (form == MP_FORM_PLAIN) return (size_t)plan_plain(h_model, h_series, R, J).total;
  return plan_batched(algo, h_model, h_series, R, J).total;
}

static int check_common(const mp_model_desc* m, const mp_series_desc* s, int R, int J, const void* keys) {
  int rc = check_model(m);
  if (rc) return rc;
  rc = check_series(s);
  if (rc) return rc;
  if (s->dc < model_covariates(m))
    return fail(MP_ERR_ARG, "model kind %d reads %d covariates per row, the series has %d", m->kind,
                model_covariates(m), s->dc);
  if (R < 1) return fail(MP_ERR_ARG, "need at least one filter (R = %d)", R);
  if (J < 1) return fail(MP_ERR_ARG, "J should be greater than 0 (J = %d)", J);
  if (!keys) return fail(MP_ERR_ARG, "keys is NULL");
  return MP_OK;
}

int32_t mp_pfilter(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
                   int64_t theta_stride, const uint32_t* keys, const int32_t* unit, int32_t R, int32_t J, float thresh,
                   int32_t flags, int32_t form, double* loglik, float* cll, float* ess, float* fmean, float* pmean,
                   void* workspace, size_t ws_bytes, void* stream) {
  int rc = check_common(h_model, h_series, R, J, keys);
  if (rc) return rc;
  if (!theta || !loglik || !workspace) return fail(MP_ERR_ARG, "NULL theta/loglik/workspace");
  if ((flags & MP_CLL) && !cll) return fail(MP_ERR_ARG, "MP_CLL set but cll is NULL");
  if ((flags & MP_ESS) && !ess) return fail(MP_ERR_ARG, "MP_ESS set but ess is NULL");
  if ((flags & MP_FILTER_MEAN) && !fmean) return fail(MP_ERR_ARG, "MP_FILTER_MEAN set but fmean is NULL");
  if ((flags & MP_PREDICTION_MEAN) && !pmean) return fail(MP_ERR_ARG, "MP_PREDICTION_MEAN set but pmean is NULL");
  hipStream_t st = (hipStream_t)stream;
  if (form == MP_FORM_PLAIN) {
    if (unit) return fail(MP_ERR_ARG, "the plain form runs single-series filters");
    return run_plain_pf(h_model, h_series, theta, theta_stride, keys, R, J, log_thresh(thresh), flags, loglik,
                        (flags & MP_CLL) ? cll : nullptr, ess, fmean, pmean, workspace, ws_bytes, st);
  }
  BatchArgs a = base_args(h_model, h_series, keys, unit, R, J);
  a.theta = theta;
  a.theta_stride = theta_stride;
  a.flags = flags;
  a.logthresh = log_thresh(thresh);
  a.loglik = loglik;
  a.cll = cll;
  a.ess = ess;
  a.fmean = fmean;
  a.pmean = pmean;
  return run_batched(ALGO_PF, h_model, h_series, a, workspace, ws_bytes, st);
}

int32_t mp_pfilter_noalloc(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
                           int64_t theta_stride, const uint32_t* keys, const int32_t* unit, int32_t R, int32_t J,
                           float thresh, int32_t flags, int32_t form, double* loglik, float* cll, float* ess,
                           float* fmean, float* pmean, void* stream) {
  int rc = check_common(h_model, h_series, R, J, keys);
  if (rc) return rc;
  // set before the size query: the plan depends on it (one plain-form lane)
  struct NoAllocScope {
    NoAllocScope() { g_no_alloc = true; }
    ~NoAllocScope() { g_no_alloc = false; }
  } no_alloc_scope;
  const size_t need = mp_filter_workspace_size(MP_ALGO_PF, h_model, h_series, R, J, form);
  if (need == 0) return fail(MP_ERR_ARG, "workspace size query failed");
  if (need > (size_t)MP_NOALLOC_SCRATCH_BYTES)
    return fail(MP_ERR_ARG,
                "mp_pfilter_noalloc: this call needs %zu bytes of scratch, more than the %zu the "
                "no-allocation entry point holds; use mp_pfilter with a workspace",
                need, (size_t)MP_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  rc = scratch_address(&ws);
  if (rc) return rc;
  return mp_pfilter(h_model, h_series, theta, theta_stride, keys, unit, R, J, thresh, flags, form, loglik, cll, ess,
                    fmean, pmean, ws, (size_t)MP_NOALLOC_SCRATCH_BYTES, stream);
}

int32_t mp_mop(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta, int64_t theta_stride,
               const uint32_t* keys, const int32_t* unit, int32_t R, int32_t J, float alpha, int32_t form,
               double* loglik, void* workspace, size_t ws_bytes, void* stream) {
  int rc = check_common(h_model, h_series, R, J, keys);
  if (rc) return rc;
  if (!theta || !loglik || !workspace) return fail(MP_ERR_ARG, "NULL theta/loglik/workspace");
  hipStream_t st = (hipStream_t)stream;
  if (form == MP_FORM_PLAIN) {
    // one filter per launch of the fused kernel
    for (int f = 0; f < R; ++f) {
      BatchArgs a = base_args(h_model, h_series, keys + 2 * f, unit ? unit + f : nullptr, 1, J);
      a.theta = theta + f * theta_stride;
      a.alpha = alpha;
      a.loglik = loglik + f;
      rc = run_batched(ALGO_MOP, h_model, h_series, a, workspace, ws_bytes, st);
      if (rc) return rc;
    }
    return MP_OK;
  }
  BatchArgs a = base_args(h_model, h_series, keys, unit, R, J);
  a.theta = theta;
  a.theta_stride = theta_stride;
  a.alpha = alpha;
  a.loglik = loglik;
  return run_batched(ALGO_MOP, h_model, h_series, a, workspace, ws_bytes, st);
}

size_t mp_mop_grad_workspace_size(const mp_model_desc* h_model, const mp_series_desc* h_series, int32_t R,
                                  int32_t J, int32_t K) {
  if (!h_model || !h_series || R < 1 || J < 1 || K < 1) return 0;
  return plan_grad(h_model, h_series, R, J, K).total;
}

int32_t mp_mop_grad(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
                    int64_t theta_stride, const uint32_t* keys, const int32_t* unit, int32_t R, int32_t J,
                    float alpha, int32_t K, int32_t form, double* loglik, float* grad, void* workspace,
                    size_t ws_bytes, void* stream) {
  int rc = check_common(h_model, h_series, R, J, keys);
  if (rc) return rc;
  if (!theta || !loglik || !grad || !workspace) return fail(MP_ERR_ARG, "NULL theta/loglik/grad/workspace");
  if (K < 1) return fail(MP_ERR_ARG, "checkpoint interval K must be >= 1 (K = %d)", K);
  if (h_series->U != 1 && !unit) return fail(MP_ERR_ARG, "a series of %d units needs a unit per filter", h_series->U);
  if (!mop_grad_supported(*reinterpret_cast<const ModelDesc*>(h_model)))
    return fail(MP_ERR_ARG, "model kind %d (dx %d, dy %d) has no adjoint", h_model->kind, h_model->dx, h_model->dy);
  if (h_series->max_euler != h_series->Tmax || h_series->max_nstep != 1)
    return fail(MP_ERR_ARG, "mp_mop_grad needs exactly one Euler step per observation interval "
                "(total %lld over %d intervals, at most %d per interval)", (long long)h_series->max_euler,
                h_series->Tmax, h_series->max_nstep);
  hipStream_t st = (hipStream_t)stream;
  if (form == MP_FORM_PLAIN) {
    for (int f = 0; f < R; ++f) {
      rc = run_mop_grad(h_model, h_series, theta + f * theta_stride, theta_stride, keys + 2 * f,
                        unit ? unit + f : nullptr, 1, J, alpha, K, loglik + f, grad + (int64_t)f * h_model->d,
                        workspace, ws_bytes, st);
      if (rc) return rc;
    }
    return MP_OK;
  }
  return run_mop_grad(h_model, h_series, theta, theta_stride, keys, unit, R, J, alpha, K, loglik, grad, workspace,
                      ws_bytes, st);
}

size_t mp_mop_hess_workspace_size(const mp_model_desc* h_model, const mp_series_desc* h_series, int32_t R,
                                  int32_t J) {
  if (!h_model || !h_series || R < 1 || J < 1) return 0;
  return plan_hess(h_model, h_series, R, J).total;
}

int32_t mp_mop_hess(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
                    int64_t theta_stride, const uint32_t* keys, int32_t R, int32_t J, float alpha, double* loglik,
                    float* grad, float* hess, void* workspace, size_t ws_bytes, void* stream) {
  int rc = check_common(h_model, h_series, R, J, keys);
  if (rc) return rc;
  if (!theta || !loglik || !grad || !hess || !workspace) return fail(MP_ERR_ARG, "NULL theta/loglik/grad/hess/workspace");
  if (h_series->U != 1) return fail(MP_ERR_ARG, "mp_mop_hess runs single-series filters (U = 1)");
  if (!mop_hess_supported(*reinterpret_cast<const ModelDesc*>(h_model)))
    return fail(MP_ERR_ARG, "model kind %d (dx %d, dy %d) has no Hessian", h_model->kind, h_model->dx,
                h_model->dy);
  const HessPlan p = plan_hess(h_model, h_series, R, J);
  if (ws_bytes < p.total) return fail(MP_ERR_WORKSPACE, "workspace %zu < required %zu bytes", ws_bytes, p.total);
  hipStream_t st = (hipStream_t)stream;
  char* base = align_ptr(workspace);
  HessArgs a{};
  a.md = *reinterpret_cast<const ModelDesc*>(h_model);
  a.ser = *reinterpret_cast<const SeriesSet*>(h_series);
  a.theta = theta;
  a.theta_stride = theta_stride;
  a.keys = keys;
  a.sched = reinterpret_cast<uint32_t*>(base);
  a.sched_len = p.sched_len;
  a.R = R;
  a.J = J;
  a.alpha = alpha;
  a.loglik = loglik;
  a.grad = grad;
  a.hess = hess;
  a.ws = base + p.sched_bytes;
  a.ws_stride = p.per_block;
  MP_HIP(timed(3, st, [&] { return launch_sched(ALGO_MOP, keys, a.ser, nullptr, p.sched_len,
                                                const_cast<uint32_t*>(a.sched), R, st); }));
  MP_HIP(timed(2, st, [&] { return launch_mop_hess(a, p.grid, st); }));
  return MP_OK;
}

int32_t mp_host_from_est(const mp_model_desc* h_model, float* h_theta, int64_t n) {
  int rc = check_model(h_model);
  if (rc) return rc;
  if (n < 0 || (n > 0 && !h_theta)) return fail(MP_ERR_ARG, "bad theta rows (n = %lld)", (long long)n);
  const ModelDesc md = *reinterpret_cast<const ModelDesc*>(h_model);
  if (md.identity_trans) return MP_OK;
  switch (md.kind) {
    case MODEL_LG: host_from_est_rows<LG8>(md, h_theta, n); break;
    case MODEL_SV: host_from_est_rows<SV>(md, h_theta, n); break;
    case MODEL_SEIR: host_from_est_rows<SEIR>(md, h_theta, n); break;
    case MODEL_SPX: host_from_est_rows<SPX>(md, h_theta, n); break;
  }
  return MP_OK;
}

int32_t mp_simulate(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
                    int64_t theta_stride, const uint32_t* keys, const float* times, int32_t R, int32_t nsim,
                    int32_t flags, float* X, float* Y, void* stream) {
  int rc = check_common(h_model, h_series, R, 1, keys);
  if (rc) return rc;
  if (!theta || !X || !Y) return fail(MP_ERR_ARG, "NULL theta/X/Y");
  if (nsim < 1) return fail(MP_ERR_ARG, "nsim must be >= 1 (nsim = %d)", nsim);
  if (h_series->U != 1) return fail(MP_ERR_ARG, "mp_simulate runs single-series models (U = 1)");
  if (h_model->dy > kSimMaxDy) return fail(MP_ERR_ARG, "dy = %d exceeds %d", h_model->dy, kSimMaxDy);
  SimArgs a{};
  a.md = *reinterpret_cast<const ModelDesc*>(h_model);
  a.ser = *reinterpret_cast<const SeriesSet*>(h_series);
  a.theta_stride = theta_stride;
  a.times = times ? times : a.ser.times;
  a.nsim = nsim;
  a.flags = flags;
  const int64_t S = h_model->S, T = h_series->Tmax, dy = h_model->dy;
  hipStream_t st = (hipStream_t)stream;
  // grid.y carries the rows: at most 65535 per launch
  for (int32_t r0 = 0; r0 < R; r0 += 65535) {
    a.R = R - r0 < 65535 ? R - r0 : 65535;
    a.theta = theta + r0 * theta_stride;
    a.keys = keys + 2 * (int64_t)r0;
    a.X = X + (int64_t)r0 * (T + 1) * S * nsim;
    a.Y = Y + (int64_t)r0 * T * dy * nsim;
    MP_HIP(launch_simulate(a, st));
  }
  return MP_OK;
}

int32_t mp_perfilter(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* thetas_in,
                     const uint32_t* keys, const int32_t* unit, int32_t R, int32_t J, float thresh, const float* sigmas,
                     const float* sigmas_init, float cool, const float* cooling, int32_t m, float* thetas_out,
                     float* theta_mean, double* loglik, int32_t* ancestry, void* workspace, size_t ws_bytes,
                     void* stream) {
  int rc = check_common(h_model, h_series, R, J, keys);
  if (rc) return rc;
  if (!thetas_in || !thetas_out || !theta_mean || !loglik || !sigmas || !sigmas_init || !workspace)
    return fail(MP_ERR_ARG, "NULL argument to mp_perfilter");
  if (thetas_in == thetas_out) return fail(MP_ERR_ARG, "thetas_in and thetas_out must not alias");
  BatchArgs a = base_args(h_model, h_series, keys, unit, R, J);
  a.theta = thetas_in;
  a.logthresh = log_thresh(thresh);
  a.sigmas = sigmas;
  a.sigmas_init = sigmas_init;
  a.cool_factor = cool;
  a.cool_table = cooling;
  a.mif_m = m;
  a.theta_out = thetas_out;
  a.theta_mean = theta_mean;
  a.loglik = loglik;
  a.ancestry = ancestry;
  return run_batched(ALGO_PER, h_model, h_series, a, workspace, ws_bytes, (hipStream_t)stream);
}

int32_t mp_tune_set(int32_t param, int32_t value) {
  switch (param) {
    case MP_TUNE_PLAIN_SCAN_ROWS:
      if (value < 0 || value > 16 || (value & (value - 1)))
        return fail(MP_ERR_ARG, "scan rows must be 0 (automatic), 1, 2, 4, 8 or 16");
      g_plain_scan_rows = value;
      return MP_OK;
    case MP_TUNE_PLAIN_PROP_ROWS:
      if (value < 0 || value > 16) return fail(MP_ERR_ARG, "propagate rows must be 0 (automatic) or 1..16");
      g_plain_prop_rows = value;
      return MP_OK;
    case MP_TUNE_BATCHED_BLOCK:
      if (value != 0 && value != 256 && value != 512 && value != 768 && value != 1024)
        return fail(MP_ERR_ARG, "batched workgroup size must be 0 (automatic), 256, 512, 768 or 1024");
      g_batched_block = value;
      return MP_OK;
    case MP_TUNE_PLAIN_SCAN_THREADS:
      if (value != 0 && value != 256 && value != 512 && value != 1024)
        return fail(MP_ERR_ARG, "plain scan workgroup size must be 0 (automatic), 256, 512 or 1024");
      g_plain_scan_threads = value;
      return MP_OK;
    case MP_TUNE_PLAIN_PROP_THREADS:
      if (value != 0 && value != 256 && value != 512 && value != 1024)
        return fail(MP_ERR_ARG, "plain workgroup size must be 0 (automatic), 256, 512 or 1024");
      g_plain_prop_threads = value;
      return MP_OK;
    case MP_TUNE_PLAIN_LDS_TILES:
      if (value < 0 || value > kMaxLdsTiles)
        return fail(MP_ERR_ARG, "LDS tile limit must be 0..%d", kMaxLdsTiles);
      g_plain_lds_tiles = value;
      return MP_OK;
    case MP_TUNE_PLAIN_XCD_MAP:
      if (value != 0 && value != 1) return fail(MP_ERR_ARG, "XCD map must be 0 or 1");
      g_plain_xcd_map = value;
      return MP_OK;
    case MP_TUNE_PLAIN_LANES:
      if (value < 1 || value > kMaxLanes) return fail(MP_ERR_ARG, "plain-form lanes must be 1..%d", kMaxLanes);
      g_plain_lanes = value;
      return MP_OK;
    case MP_TUNE_GRAD_BLOCK:
      if (value != 0 && value != 256 && value != 512)
        return fail(MP_ERR_ARG, "adjoint workgroup size must be 0 (automatic), 256 or 512");
      g_grad_block = value;
      return MP_OK;
  }
  return fail(MP_ERR_ARG, "unknown tuning parameter %d", param);
}

int32_t mp_tune_get(int32_t param) {
  switch (param) {
    case MP_TUNE_PLAIN_SCAN_ROWS: return g_plain_scan_rows;
    case MP_TUNE_PLAIN_PROP_ROWS: return g_plain_prop_rows;
    case MP_TUNE_BATCHED_BLOCK: return g_batched_block;
    case MP_TUNE_GRAD_BLOCK: return g_grad_block;
    case MP_TUNE_PLAIN_SCAN_THREADS: return g_plain_scan_threads;
    case MP_TUNE_PLAIN_PROP_THREADS: return g_plain_prop_threads;
    case MP_TUNE_PLAIN_LDS_TILES: return g_plain_lds_tiles;
    case MP_TUNE_PLAIN_XCD_MAP: return g_plain_xcd_map;
    case MP_TUNE_PLAIN_LANES: return g_plain_lanes;
  }
  return -1;
}

int32_t mp_valu_probe(int32_t kind, int64_t iters, int32_t blocks, float* out, void* stream) {
  if (kind < 0 || kind > 4) return fail(MP_ERR_ARG, "probe kind must be 0..4");
  if (iters < 1 || blocks < 1 || !out) return fail(MP_ERR_ARG, "need iters >= 1, blocks >= 1 and an output");
  hipStream_t st = (hipStream_t)stream;
  MP_HIP(timed(4, st, [&] { return launch_valu_probe(kind, iters, blocks, out, st); }));
  return MP_OK;
}

int32_t mp_profile_enable(int32_t on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = on != 0;
  return MP_OK;
}

int32_t mp_profile_read(double* h_ms, int64_t* h_count, int32_t nkinds) {
  if (!h_ms || !h_count || nkinds < 1) return fail(MP_ERR_ARG, "bad profile buffers");
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (int k = 0; k < nkinds; ++k) {
    h_ms[k] = 0.0;
    h_count[k] = 0;
  }
  for (const ProfRec& r : g_prof) {
    MP_HIP(hipEventSynchronize(r.b));
    float ms = 0.f;
    MP_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    if (r.kind < nkinds) {
      h_ms[r.kind] += ms;
      h_count[r.kind] += 1;
    }
    g_event_pool.push_back(r.a);
    g_event_pool.push_back(r.b);
  }
  g_prof.clear();
  return MP_OK;
}

}  // extern "C"
