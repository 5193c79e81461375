// liblbm: C ABI of the coupled LBM step on MI355X (declarations in include/lbm.h).
//
// Host side of the library: device state, the per-pass entry points (the
// reference's Sim passes one by one), the fused coupled step, field transfer and
// the host RigidSphere functions.  Kernels live in kernels_*.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "lbm.h"
#include "device_common.h"
#include "kernels_pass.h"
#include "kernels_fused.h"
#include "kernels_march.h"

using namespace lbm;

// Per-launch HIP event timing (lbm_set_timing / lbm_kernel_time).
struct Timer {
  bool on = false;
  std::map<std::string, std::pair<double, int64_t>> total;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
};

// Executable graphs of captured launch sequences, keyed by the bytes of every
// launch argument and grid size they were captured with: an entry is replayed
// only for launches identical to the captured ones.  Least recently used
// entries are dropped beyond kMax.
struct GraphCache {
  static constexpr int kMax = 8;
  struct Entry {
    std::vector<char> key;
    hipGraphExec_t exec;
    uint64_t used;
  };
  std::vector<Entry> entries;
  uint64_t clock = 0;
  // Steps per graph (even; 0: no graphs, every kernel launched on its own).
  int steps = 2;
  int64_t captures = 0, replays = 0;
  ~GraphCache() { clear(); }
  void clear() {
    for (auto &e : entries) hipGraphExecDestroy(e.exec);
    entries.clear();
  }
  hipGraphExec_t find(const std::vector<char> &key) {
    for (auto &e : entries)
      if (e.key == key) {
        e.used = ++clock;
        return e.exec;
      }
    return nullptr;
  }
  void put(std::vector<char> key, hipGraphExec_t exec) {
    if ((int)entries.size() >= kMax) {
      size_t lru = 0;
      for (size_t j = 1; j < entries.size(); j++)
        if (entries[j].used < entries[lru].used) lru = j;
      hipGraphExecDestroy(entries[lru].exec);
      entries.erase(entries.begin() + lru);
    }
    entries.push_back({std::move(key), exec, ++clock});
  }
};

struct lbm_sim {
  int nx, ny, nz;
  Geom g;
  int device = 0;
  hipStream_t stream = nullptr;
  // Memory pool of the server that made the handle (null: plain allocations).
  hipMemPool_t pool = nullptr;
  lbm_params params{};
  // Population buffers (row blocks of the 27 f and 7 h directions, one per step
  // parity; hbuf[b] points into fbuf[b]): fbuf[fcur] is the reference's f,
  // fbuf[1 - fcur] its f2;
  // likewise for h.  phibuf[phicur] is phi, phibuf[1 - phicur] phiPrev.
  float *fbuf[2] = {nullptr, nullptr};
  float *hbuf[2] = {nullptr, nullptr};
  int fcur = 0, hcur = 0;
  float *phibuf[2] = {nullptr, nullptr};
  int phicur = 0;
  float *rho = nullptr, *ux = nullptr, *uy = nullptr, *uz = nullptr;
  uint8_t *tag = nullptr, *tagPrev = nullptr;
  float *usx = nullptr, *usy = nullptr, *usz = nullptr;
  float *gx = nullptr, *gy = nullptr, *gz = nullptr, *lap = nullptr;
  float *bfx = nullptr, *bfy = nullptr, *bfz = nullptr;
  bool bodyF_armed = false;
  // Gradients stored in gx..lap are those of the last computePhiGradients; after
  // a fused step they are recomputed from phiPrev when read (grads_stale).
  bool grads_stale = false;
  // Lists of the last retag.
  int32_t *fresh = nullptr, *dead = nullptr;
  int list_cap = 0;  // power of two >= n
  int *level_scratch = nullptr;
  Ctrl *ctrl = nullptr;       // device
  Ctrl *ctrl_host = nullptr;  // pinned mirror
  // Momentum-exchange link list (per-pass path), its sorted copy and the
  // sort's temporary storage.
  int64_t *link_key = nullptr;
  double *link_m = nullptr;
  int link_cap = 0;
  int64_t *link_key2 = nullptr;
  double *link_m2 = nullptr;
  int link_cap2 = 0;
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  // Pinned staging of sphere uploads that the per-pass step does not wait for,
  // and an event after the last of them (the staging is reused once it fired).
  lbm_sphere *sphere_pin = nullptr;
  hipEvent_t sphere_ev = nullptr;
  bool sphere_ev_armed = false;
  // Staging for host transfers and layout transposes.
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  uint8_t *mask_dev = nullptr;
  float *mus_dev = nullptr;
  // Fused step.
  double *partials = nullptr;  // per-block link momentum, 6 doubles per block
  int64_t partial_blocks = 0;
  bool need_full_retag = true;  // next coupled step must retag the whole grid
  // The last fused call ended after a step's body update, with that step's band
  // update deferred to the next call (only the serving loop ends calls so; phi,
  // the tags and the sphere are then the reference's, the other fields not).
  bool live = false;
  bool sphere_dev_valid = false;
  std::vector<lbm_link> links;  // kept by lbm_stream_hydro_macro_links
  // Compressed populations of the fused step (lbm_set_storage): a pair of
  // 16-bit f and h buffers; cf[ccur]/ch[ccur] hold the reference's f and h when
  // comp_auth is set, and the float32 buffers are then stale.
  int storage = LBM_STORAGE_F32;
  bool comp_auth = false;
  int ccur = 0;
  void *cf[2] = {nullptr, nullptr}, *ch[2] = {nullptr, nullptr};
  // Timing.
  Timer timer;
  GraphCache graphs;  // steady steps of the fused call (lbm_step)
  int lanes = 2;      // cells per lane of the 16-bit stream kernel (1 or 2)
  // Float32 fused steps run the row march beside the band box's cells (2,
  // where rows fit it), launch the bulk tiles and the band-row tiles of the
  // stream kernel separately (1) or as one launch (0).
  int stream_split = 2;
  int march_rows = 2;  // rows per workgroup of the row march (1 or 2)
  int march_nseg = 0;  // plane segments of the row march (0: chosen per grid)
  std::string err;
};

static std::string g_create_error;

// Device and pinned-host allocation and release.  Each holds g_alloc_mu, which a
// launch-graph capture also holds from its start to the graph's instantiation,
// as do lbm_create and lbm_destroy (streams created and destroyed): an
// allocation, a release or a stream's creation or destruction on another thread
// (any of which may synchronise the device) then never meets a stream that is
// being captured.
static std::recursive_mutex g_alloc_mu;
template <typename T>
static hipError_t dev_malloc(T **p, size_t bytes) {
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  return hipMalloc((void **)p, bytes);
}
static hipError_t dev_free(void *p) {
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  return hipFree(p);
}
template <typename T>
static hipError_t host_malloc(T **p, size_t bytes, unsigned flags) {
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  return hipHostMalloc((void **)p, bytes, flags);
}
static hipError_t host_free(void *p) {
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  return hipHostFree(p);
}

// Device memory of a handle.  A handle that a server made takes its buffers
// from the server's memory pool, ordered on the handle's stream, and gives them
// back the same way: hipFree waits for the device (about 2.6 ms a buffer while
// a server steps other scenes, 80 ms for a scene's thirty buffers), and memory
// freed into the pool is taken by the next scene without going back to the
// driver.  An allocation waits for the handle's stream, so the buffer is ready
// for any stream once it returns; a buffer is given back only when no queued
// work uses it.
template <typename T>
static hipError_t sim_malloc(lbm_sim *s, T **p, size_t bytes) {
  if (!s->pool) return dev_malloc(p, bytes);
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  hipError_t e = hipMallocFromPoolAsync((void **)p, bytes, s->pool, s->stream);
  if (e == hipSuccess && ((uintptr_t)*p & 255u) != 0) {
    // The kernels assume 256-byte aligned buffers, as hipMalloc gives them.
    (void)hipFreeAsync(*p, s->stream);
    *p = nullptr;
    e = hipErrorInvalidValue;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  return e;
}
static hipError_t sim_free(lbm_sim *s, void *p) {
  if (!p) return hipSuccess;
  if (!s->pool) return dev_free(p);
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  return hipFreeAsync(p, s->stream);
}

// Small pinned host blocks (a handle's control-block mirror and sphere
// staging) are kept for the next handle instead of released: hipHostFree waits
// for the device as hipFree does.  At most kPinCacheMax blocks are kept.
static std::mutex g_pin_mu;
static std::multimap<size_t, void *> g_pin_cache;
constexpr size_t kPinCacheMax = 4096;
template <typename T>
static hipError_t pin_small_take(T **p, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    auto it = g_pin_cache.find(bytes);
    if (it != g_pin_cache.end()) {
      *p = (T *)it->second;
      g_pin_cache.erase(it);
      std::memset((void *)*p, 0, bytes);
      return hipSuccess;
    }
  }
  return host_malloc(p, bytes, hipHostMallocDefault);
}
static void pin_small_give(void *p, size_t bytes) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    if (g_pin_cache.size() < kPinCacheMax) {
      g_pin_cache.emplace(bytes, p);
      return;
    }
  }
  host_free(p);
}

// ------------------------------------------------------------------ no-alloc scratch
// Device memory of lbm_step_noalloc: the fused step's link-momentum partials
// (6 doubles per 256-cell tile) in a file-scope device array instead of the
// handle's own buffer, so that entry point allocates nothing.  Every GPU the
// library runs on holds its own copy, at its own address, until the process
// ends.  On the host the array's name is a stand-in, not a device address: host
// code uses only the address hipGetSymbolAddress gives for each device, looked
// up once per device by lbm_create (with that device current).
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];
static bool g_scratch_done[kMaxDevices];

// Looks up the current device's copy of `scratch` once; a failed lookup leaves
// that device without one (lbm_step_noalloc then reports it).
static void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
  g_scratch_done[device] = true;
}

static void *scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

// ------------------------------------------------------------------ helpers

#define HIP_TRY(sim, expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      (sim)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                  \
      return LBM_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define CHECK_SIM(sim)                    \
  do {                                    \
    if (!(sim)) return LBM_ERR_ARG;       \
  } while (0)

static int fail(lbm_sim *s, int code, const std::string &msg) {
  s->err = msg;
  return code;
}

static Derived derive(const lbm_params &p) {
  // src/sim.js:327-332 and :410-416, :490-492, evaluated exactly as there.
  Derived d;
  d.invTau = 1 / p.tau;
  d.beta = 12 * p.sigma / p.W;
  d.kappa = 1.5 * p.sigma * p.W;
  d.dRho = p.rhoL - p.rhoG;
  d.A = (1 - 0.5 * d.invTau) * INV_CS2_27;
  d.rhoG = p.rhoG;
  d.gravity = p.gravity;
  d.invTauPhi = 1 / p.tauPhi;
  const double Mphi = CS2_7 * (p.tauPhi - 0.5);
  const double guoHalf = 1.0 - 0.5 * d.invTauPhi;
  d.sharpC = 4.0 * Mphi * guoHalf / p.W;
  d.sharpFac = INV_CS2_7;
  return d;
}

static dim3 cell_grid(const lbm_sim *s) {
  return dim3((s->nx + 63) / 64, (s->ny + 3) / 4, s->nz);
}
static const dim3 kCellBlock(64, 4, 1);
static const dim3 kPairBlock(64, 2, 1);  // a tile of 256 cells, two per lane

static float *F(lbm_sim *s) { return s->fbuf[s->fcur]; }
static float *F2(lbm_sim *s) { return s->fbuf[1 - s->fcur]; }
static float *H(lbm_sim *s) { return s->hbuf[s->hcur]; }
static float *H2(lbm_sim *s) { return s->hbuf[1 - s->hcur]; }
static float *PHI(lbm_sim *s) { return s->phibuf[s->phicur]; }
static float *PHIPREV(lbm_sim *s) { return s->phibuf[1 - s->phicur]; }

static int ensure_scratch(lbm_sim *s, size_t bytes) {
  if (s->scratch_bytes >= bytes) return LBM_OK;
  sim_free(s, s->scratch);
  s->scratch = nullptr;
  s->scratch_bytes = 0;
  HIP_TRY(s, sim_malloc(s, &s->scratch, bytes));
  s->scratch_bytes = bytes;
  return LBM_OK;
}

// Makes the float32 populations current again after fused steps with
// compressed storage: decodes the reference's f and h (h against the current
// phi) into fbuf[fcur] and hbuf[hcur], on stream st.
static int ensure_f32(lbm_sim *s, hipStream_t st) {
  if (!s->comp_auth) return LBM_OK;
  const unsigned blocks = (unsigned)((s->g.n + 255) / 256);
  if (s->storage == LBM_STORAGE_F16)
    k_decode<ST_F16><<<blocks, 256, 0, st>>>(s->g, (const _Float16 *)s->cf[s->ccur],
                                             (const _Float16 *)s->ch[s->ccur], PHI(s), F(s), H(s));
  else
    k_decode<ST_BF16><<<blocks, 256, 0, st>>>(s->g, (const uint16_t *)s->cf[s->ccur],
                                              (const uint16_t *)s->ch[s->ccur], PHI(s), F(s),
                                              H(s));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(s, LBM_ERR_HIP, std::string("k_decode: ") + hipGetErrorString(e));
  s->comp_auth = false;
  return LBM_OK;
}

// After lbm_step_live the state is the fused step's own (post-collision
// populations for the next step, rho and u stored only near the sphere): entry
// points that need the reference's state refuse it until an lbm_step call ends
// in that state.
#define CHECK_NOT_LIVE(sim)                                                              \
  do {                                                                                   \
    if ((sim)->live)                                                                     \
      return fail((sim), LBM_ERR_STATE,                                                  \
                  "state between live calls (lbm_step_live): end with lbm_step first"); \
  } while (0)

// CHECK_SIM for entry points that read or write the float32 populations or
// anything their compressed baseline depends on.
#define CHECK_SIM_F32(sim)                                  \
  do {                                                      \
    if (!(sim)) return LBM_ERR_ARG;                         \
    CHECK_NOT_LIVE(sim);                                    \
    const int rc_f32_ = ensure_f32((sim), (sim)->stream);   \
    if (rc_f32_) return rc_f32_;                            \
  } while (0)

static int ensure_links(lbm_sim *s, int64_t cap) {
  if (cap > INT32_MAX) return fail(s, LBM_ERR_CAPACITY, "link list capacity exceeds 2^31");
  if (s->link_cap >= cap) return LBM_OK;
  sim_free(s, s->link_key);
  sim_free(s, s->link_m);
  s->link_key = nullptr;
  s->link_m = nullptr;
  s->link_cap = 0;
  HIP_TRY(s, sim_malloc(s, &s->link_key, cap * sizeof(int64_t)));
  HIP_TRY(s, sim_malloc(s, &s->link_m, cap * sizeof(double)));
  s->link_cap = (int)cap;
  return LBM_OK;
}

static int sync(lbm_sim *s) {
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  return LBM_OK;
}

// Sorted-copy buffers and the radix sort's temporary storage for `cap` links.
static int ensure_link_sort(lbm_sim *s, int64_t cap, int end_bit) {
  if (s->link_cap2 < cap) {
    sim_free(s, s->link_key2);
    sim_free(s, s->link_m2);
    s->link_key2 = nullptr;
    s->link_m2 = nullptr;
    s->link_cap2 = 0;
    HIP_TRY(s, sim_malloc(s, &s->link_key2, cap * sizeof(int64_t)));
    HIP_TRY(s, sim_malloc(s, &s->link_m2, cap * sizeof(double)));
    s->link_cap2 = (int)cap;
  }
  size_t need = 0;
  HIP_TRY(s, rocprim::radix_sort_pairs(nullptr, need, (uint64_t *)s->link_key,
                                       (uint64_t *)s->link_key2, s->link_m, s->link_m2,
                                       (unsigned)cap, 0, end_bit, s->stream));
  if (need > s->sort_tmp_bytes) {
    sim_free(s, s->sort_tmp);
    s->sort_tmp = nullptr;
    s->sort_tmp_bytes = 0;
    HIP_TRY(s, sim_malloc(s, &s->sort_tmp, need));
    s->sort_tmp_bytes = need;
  }
  return LBM_OK;
}

// Copies *b to the device sphere without waiting: through the pinned staging,
// which the previous upload must have read first.
static int upload_sphere_async(lbm_sim *s, const lbm_sphere *b) {
  if (s->sphere_ev_armed) HIP_TRY(s, hipEventSynchronize(s->sphere_ev));
  *s->sphere_pin = *b;
  HIP_TRY(s, hipMemcpyAsync(&s->ctrl->sphere, s->sphere_pin, sizeof(lbm_sphere),
                            hipMemcpyHostToDevice, s->stream));
  HIP_TRY(s, hipEventRecord(s->sphere_ev, s->stream));
  s->sphere_ev_armed = true;
  return LBM_OK;
}

static int read_ctrl(lbm_sim *s) {
  HIP_TRY(s, hipMemcpyAsync(s->ctrl_host, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
  return sync(s);
}

static int check_launch(lbm_sim *s, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(s, LBM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return LBM_OK;
}

// Timing brackets around a launch (only when timing is on).  The events only
// carry timestamps, so they skip the system-scope fence a default event record
// performs (an L2 writeback and invalidate on every XCD, ~5 us per record on
// MI355X, which would otherwise sit between the timed kernels).
struct TimedLaunch {
  Timer &t;
  hipStream_t st;
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  TimedLaunch(Timer &t_, hipStream_t st_, const char *n) : t(t_), st(st_), name(n) {
    if (t.on) {
      hipEventCreateWithFlags(&a, hipEventDisableSystemFence);
      hipEventCreateWithFlags(&b, hipEventDisableSystemFence);
      hipEventRecord(a, st);
    }
  }
  ~TimedLaunch() {
    if (t.on) {
      hipEventRecord(b, st);
      t.pending.push_back({name, {a, b}});
    }
  }
};

static void collect_timing(Timer &t) {
  for (auto &p : t.pending) {
    float ms = 0;
    hipEventSynchronize(p.second.second);
    hipEventElapsedTime(&ms, p.second.first, p.second.second);
    auto &e = t.total[p.first];
    e.first += ms;
    e.second += 1;
    hipEventDestroy(p.second.first);
    hipEventDestroy(p.second.second);
  }
  t.pending.clear();
}

// ------------------------------------------------------------------ lifetime

// lbm_create, with the handle's device memory from `pool` when it is not null
// (the serving loop's scenes).
static int create_sim(int nx, int ny, int nz, hipMemPool_t pool, lbm_sim **out) {
  if (!out) return LBM_ERR_ARG;
  *out = nullptr;
  if (nx < 3 || ny < 3 || nz < 3) {
    g_create_error = "grid must be at least 3 cells on each axis";
    return LBM_ERR_ARG;
  }
  const int64_t n = (int64_t)nx * ny * nz;
  if (n > ((int64_t)1 << 30)) {
    g_create_error = "grid too large: at most 2^30 cells";
    return LBM_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    g_create_error = "no HIP device available";
    return LBM_ERR_NODEVICE;
  }
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  lbm_sim *s = new lbm_sim();
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  s->g.nx = nx;
  s->g.ny = ny;
  s->g.nz = nz;
  s->g.nxny = nx * ny;
  s->g.n = n;
  s->g.S = (n + 63) / 64 * 64;
  choose_layout(s->g);
  hipGetDevice(&s->device);
  scratch_setup(s->device);
  // src/sim.js:88-95 defaults.
  s->params = {0.6, 0.7, 0.012, 0.0010, 1.0, 0.05, 4.0, 0.02};
  int cap = 1;
  while (cap < n) cap <<= 1;
  s->list_cap = cap;
  const int64_t S = s->g.S;
  s->pool = pool;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
#define ALLOC(ptr, bytes)                                        \
  if (e == hipSuccess) {                                         \
    e = sim_malloc(s, (void **)&(ptr), (bytes));                  \
  }
  ALLOC(s->fbuf[0], S * kRowQ * sizeof(float));
  ALLOC(s->fbuf[1], S * kRowQ * sizeof(float));
  ALLOC(s->phibuf[0], S * sizeof(float));
  ALLOC(s->phibuf[1], S * sizeof(float));
  ALLOC(s->rho, S * sizeof(float));
  ALLOC(s->ux, S * sizeof(float));
  ALLOC(s->uy, S * sizeof(float));
  ALLOC(s->uz, S * sizeof(float));
  ALLOC(s->tag, S);
  ALLOC(s->tagPrev, S);
  ALLOC(s->usx, S * sizeof(float));
  ALLOC(s->usy, S * sizeof(float));
  ALLOC(s->usz, S * sizeof(float));
  ALLOC(s->gx, S * sizeof(float));
  ALLOC(s->gy, S * sizeof(float));
  ALLOC(s->gz, S * sizeof(float));
  ALLOC(s->lap, S * sizeof(float));
  ALLOC(s->bfx, S * sizeof(float));
  ALLOC(s->bfy, S * sizeof(float));
  ALLOC(s->bfz, S * sizeof(float));
  ALLOC(s->fresh, (size_t)cap * sizeof(int32_t));
  ALLOC(s->dead, (size_t)cap * sizeof(int32_t));
  ALLOC(s->level_scratch, (size_t)cap * sizeof(int));
  ALLOC(s->ctrl, sizeof(Ctrl));
#undef ALLOC
  if (e == hipSuccess)
    for (int b = 0; b < 2; b++) s->hbuf[b] = s->fbuf[b] + Q27 * s->g.pb;
  if (e == hipSuccess) e = pin_small_take(&s->ctrl_host, sizeof(Ctrl));
  if (e == hipSuccess) e = pin_small_take(&s->sphere_pin, sizeof(lbm_sphere));
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s->sphere_ev, hipEventDisableTiming);
  if (e == hipSuccess) {
    // Zero the whole state so reads before an init are defined.
    const int64_t fb = S * sizeof(float);
    for (int b = 0; b < 2; b++) {
      hipMemsetAsync(s->fbuf[b], 0, S * kRowQ * sizeof(float), s->stream);
      hipMemsetAsync(s->phibuf[b], 0, fb, s->stream);
    }
    float *arrs[] = {s->rho, s->ux, s->uy, s->uz, s->usx, s->usy, s->usz, s->gx, s->gy, s->gz,
                     s->lap, s->bfx, s->bfy, s->bfz};
    for (float *a : arrs) hipMemsetAsync(a, 0, fb, s->stream);
    hipMemsetAsync(s->tag, 0, S, s->stream);
    hipMemsetAsync(s->tagPrev, 0, S, s->stream);
    hipMemsetAsync(s->ctrl, 0, sizeof(Ctrl), s->stream);
    e = hipStreamSynchronize(s->stream);
  }
  if (e != hipSuccess) {
    g_create_error = std::string("allocation of device state failed: ") + hipGetErrorString(e);
    lbm_destroy(s);
    return e == hipErrorOutOfMemory ? LBM_ERR_OOM : LBM_ERR_HIP;
  }
  *out = s;
  return LBM_OK;
}

extern "C" int lbm_create(int nx, int ny, int nz, lbm_sim **out) {
  return create_sim(nx, ny, nz, nullptr, out);
}

extern "C" void lbm_destroy(lbm_sim *s) {
  if (!s) return;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  if (s->stream) hipStreamSynchronize(s->stream);
  collect_timing(s->timer);
  void *ptrs[] = {s->fbuf[0], s->fbuf[1], s->phibuf[0], s->phibuf[1],
                  s->rho, s->ux, s->uy, s->uz, s->tag, s->tagPrev, s->usx, s->usy, s->usz, s->gx,
                  s->gy, s->gz, s->lap, s->bfx, s->bfy, s->bfz, s->fresh, s->dead,
                  s->level_scratch, s->ctrl, s->link_key, s->link_m, s->link_key2, s->link_m2,
                  s->sort_tmp, s->scratch,
                  s->mask_dev, s->mus_dev, s->partials, s->cf[0], s->cf[1], s->ch[0],
                  s->ch[1]};
  for (void *p : ptrs) sim_free(s, p);
  // Pooled buffers go back to the pool once the stream reaches their frees.
  if (s->pool && s->stream) hipStreamSynchronize(s->stream);
  pin_small_give(s->ctrl_host, sizeof(Ctrl));
  pin_small_give(s->sphere_pin, sizeof(lbm_sphere));
  if (s->sphere_ev) hipEventDestroy(s->sphere_ev);
  s->graphs.clear();
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
}

extern "C" const char *lbm_last_error(const lbm_sim *s) {
  return s ? s->err.c_str() : g_create_error.c_str();
}

extern "C" int lbm_dims(const lbm_sim *s, int *nx, int *ny, int *nz) {
  CHECK_SIM(s);
  if (nx) *nx = s->nx;
  if (ny) *ny = s->ny;
  if (nz) *nz = s->nz;
  return LBM_OK;
}

extern "C" int lbm_set_params(lbm_sim *s, const lbm_params *p) {
  CHECK_SIM(s);
  if (!p) return fail(s, LBM_ERR_ARG, "params is NULL");
  s->params = *p;
  return LBM_OK;
}

extern "C" int lbm_get_params(const lbm_sim *s, lbm_params *p) {
  CHECK_SIM(s);
  if (!p) return LBM_ERR_ARG;
  *p = s->params;
  return LBM_OK;
}

extern "C" const char *lbm_version(void) { return "liblbm 0.1 gfx950"; }

extern "C" void *lbm_stream(lbm_sim *s) { return s ? (void *)s->stream : nullptr; }

extern "C" int lbm_synchronize(lbm_sim *s) {
  CHECK_SIM(s);
  return sync(s);
}

// ------------------------------------------------------------------ init

static int init_columns(lbm_sim *s, const double *yw, double waterline) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const double W = s->params.W;
  std::vector<double> col;
  int per_cell = yw != nullptr;
  if (!per_cell) {
    col.resize(ny);
    // src/sim.js:116
    for (int y = 0; y < ny; y++) col[y] = 0.5 * (1 + std::tanh(2 * (y - waterline) / W));
  } else {
    col.resize((size_t)s->g.n);
    for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          const double wl = yw[(size_t)z * nx + x];
          col[((size_t)z * ny + y) * nx + x] = 0.5 * (1 + std::tanh(2 * (y - wl) / W));
        }
  }
  const size_t bytes = col.size() * sizeof(double);
  int rc = ensure_scratch(s, bytes);
  if (rc) return rc;
  HIP_TRY(s, hipMemcpyAsync(s->scratch, col.data(), bytes, hipMemcpyHostToDevice, s->stream));
  k_init_cells<<<cell_grid(s), kCellBlock, 0, s->stream>>>(
      s->g, (const double *)s->scratch, per_cell, PHI(s), PHIPREV(s), s->rho, s->ux, s->uy, s->uz,
      s->tag, s->tagPrev, F(s), H(s));
  if ((rc = check_launch(s, "k_init_cells"))) return rc;
  rc = lbm_seed_hydrostatic(s);
  if (rc) return rc;
  s->need_full_retag = true;
  s->grads_stale = false;
  return sync(s);
}

extern "C" int lbm_init_flat(lbm_sim *s, double waterline_y) {
  if (s) s->live = false;  // initialisation replaces the whole state
  CHECK_SIM_F32(s);
  return init_columns(s, nullptr, waterline_y);
}

extern "C" int lbm_init_waterline_map(lbm_sim *s, const double *yw) {
  if (s) s->live = false;  // initialisation replaces the whole state
  CHECK_SIM_F32(s);
  if (!yw) return fail(s, LBM_ERR_ARG, "yw is NULL");
  return init_columns(s, yw, 0.0);
}

extern "C" int lbm_seed_hydrostatic(lbm_sim *s) {
  CHECK_SIM_F32(s);
  const lbm_params &p = s->params;
  dim3 grid((s->nx + 63) / 64, s->nz);
  k_seed_hydrostatic<<<grid, 64, 0, s->stream>>>(s->g, p.rhoG, p.rhoL - p.rhoG, p.gravity,
                                                 PHI(s), s->rho, F(s));
  return check_launch(s, "k_seed_hydrostatic");
}

// ------------------------------------------------------------------ passes

static int retag_full(lbm_sim *s, int mode) {
  HIP_TRY(s, hipMemsetAsync(s->ctrl, 0, 4 * sizeof(int), s->stream));  // list counters
  k_retag<<<cell_grid(s), kCellBlock, 0, s->stream>>>(
      s->g, mode, &s->ctrl->sphere, s->mask_dev, s->mus_dev, s->tag, s->tagPrev, PHI(s),
      PHIPREV(s), 1, s->usx, s->usy, s->usz, s->ctrl, s->fresh, s->dead, s->list_cap);
  int rc = check_launch(s, "k_retag");
  if (rc) return rc;
  k_sort_lists<<<2, 1024, 0, s->stream>>>(s->fresh, s->dead, s->ctrl, s->list_cap);
  return check_launch(s, "k_sort_lists");
}

static int upload_sphere(lbm_sim *s, const lbm_sphere *b) {
  HIP_TRY(s, hipMemcpyAsync(&s->ctrl->sphere, b, sizeof(lbm_sphere), hipMemcpyHostToDevice,
                            s->stream));
  // The copy reads *b asynchronously; wait so the caller may modify it.
  return sync(s);
}

extern "C" int lbm_retag_cells(lbm_sim *s, const lbm_sphere *solid) {
  CHECK_SIM_F32(s);
  int rc;
  if (solid && (rc = upload_sphere(s, solid))) return rc;
  rc = retag_full(s, solid ? 1 : 0);
  if (rc) return rc;
  s->need_full_retag = true;  // coupled path re-derives its band from scratch
  return LBM_OK;
}

extern "C" int lbm_retag_cells_mask(lbm_sim *s, const uint8_t *mask, const float *us) {
  CHECK_SIM_F32(s);
  if (!mask || !us) return fail(s, LBM_ERR_ARG, "mask/us is NULL");
  const int64_t n = s->g.n;
  if (!s->mask_dev) {
    HIP_TRY(s, sim_malloc(s, &s->mask_dev, n));
    HIP_TRY(s, sim_malloc(s, &s->mus_dev, n * 3 * sizeof(float)));
  }
  HIP_TRY(s, hipMemcpyAsync(s->mask_dev, mask, n, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(s, hipMemcpyAsync(s->mus_dev, us, n * 3 * sizeof(float), hipMemcpyHostToDevice, s->stream));
  int rc = retag_full(s, 2);
  if (rc) return rc;
  s->need_full_retag = true;
  return sync(s);
}

static int fresh_pass(lbm_sim *s, float *phi_fb, float *phi_w1, float *phi_w2, float *f, float *h) {
  const lbm_params &p = s->params;
  k_fresh<<<1, 1024, 0, s->stream>>>(s->g, p.rhoG, p.rhoL, s->fresh, s->ctrl, s->list_cap,
                                     s->level_scratch, s->tag, phi_fb, phi_w1, phi_w2, s->rho,
                                     s->ux, s->uy, s->uz, f, h);
  return check_launch(s, "k_fresh");
}

extern "C" int lbm_handle_fresh_cells(lbm_sim *s) {
  CHECK_SIM_F32(s);
  return fresh_pass(s, PHIPREV(s), PHI(s), PHIPREV(s), F(s), H(s));
}

extern "C" int lbm_compute_phi_gradients(lbm_sim *s) {
  CHECK_SIM_F32(s);
  k_gradients<<<cell_grid(s), kCellBlock, 0, s->stream>>>(s->g, PHI(s), s->gx, s->gy, s->gz,
                                                          s->lap);
  s->grads_stale = false;
  return check_launch(s, "k_gradients");
}

extern "C" int lbm_collide_phase(lbm_sim *s) {
  CHECK_SIM_F32(s);
  const Derived d = derive(s->params);
  k_collide_phase<<<cell_grid(s), kCellBlock, 0, s->stream>>>(s->g, d, H(s), H2(s), PHI(s), s->ux,
                                                              s->uy, s->uz, s->tag, s->gx, s->gy,
                                                              s->gz);
  return check_launch(s, "k_collide_phase");
}

extern "C" int lbm_stream_phase(lbm_sim *s) {
  CHECK_SIM_F32(s);
  k_stream_phase<<<cell_grid(s), kCellBlock, 0, s->stream>>>(s->g, H2(s), H(s), s->tag);
  return check_launch(s, "k_stream_phase");
}

extern "C" int lbm_collide_hydro(lbm_sim *s) {
  CHECK_SIM_F32(s);
  const Derived d = derive(s->params);
  k_collide_hydro<<<cell_grid(s), kCellBlock, 0, s->stream>>>(
      s->g, d, F(s), F2(s), s->rho, s->ux, s->uy, s->uz, PHI(s), s->tag, s->gx, s->gy, s->gz,
      s->lap, s->bfx, s->bfy, s->bfz, s->bodyF_armed ? 1 : 0);
  return check_launch(s, "k_collide_hydro");
}

// Runs streamHydroAndMacro; with emit, collects the link list (sorted by key) on
// the host.  Clears the external body force afterwards (src/sim.js:581).
static int stream_macro_pass(lbm_sim *s, bool emit, std::vector<int64_t> *keys,
                             std::vector<double> *ms) {
  const Derived d = derive(s->params);
  int rc;
  if (emit) {
    // The tags may have been written by the host since the last retag, so count
    // the solid cells now; they bound the number of links.
    HIP_TRY(s, hipMemsetAsync(&s->ctrl->n_solid, 0, sizeof(int), s->stream));
    k_count_solid<<<1024, 256, 0, s->stream>>>(s->g.n, s->tag, s->ctrl);
    if ((rc = check_launch(s, "k_count_solid"))) return rc;
    rc = read_ctrl(s);
    if (rc) return rc;
    const int64_t bound = std::max<int64_t>(1, (int64_t)s->ctrl_host->n_solid * 26);
    if ((rc = ensure_links(s, bound))) return rc;
    HIP_TRY(s, hipMemsetAsync(&s->ctrl->n_links, 0, sizeof(int), s->stream));
  }
  k_stream_hydro_macro<<<cell_grid(s), kCellBlock, 0, s->stream>>>(
      s->g, d, F2(s), F(s), H(s), s->rho, s->ux, s->uy, s->uz, PHI(s), s->tag, s->usx, s->usy,
      s->usz, s->gx, s->gy, s->gz, s->lap, s->bfx, s->bfy, s->bfz, s->bodyF_armed ? 1 : 0,
      emit ? 1 : 0, s->link_key, s->link_m, s->link_cap, s->ctrl);
  if ((rc = check_launch(s, "k_stream_hydro_macro"))) return rc;
  if (s->bodyF_armed) {
    const int64_t fb = s->g.S * sizeof(float);
    HIP_TRY(s, hipMemsetAsync(s->bfx, 0, fb, s->stream));
    HIP_TRY(s, hipMemsetAsync(s->bfy, 0, fb, s->stream));
    HIP_TRY(s, hipMemsetAsync(s->bfz, 0, fb, s->stream));
    s->bodyF_armed = false;
  }
  if (!emit) return LBM_OK;
  if ((rc = read_ctrl(s))) return rc;
  const int cnt = s->ctrl_host->n_links;
  if (cnt > s->link_cap || (s->ctrl_host->err & kErrListOverflow))
    return fail(s, LBM_ERR_CAPACITY, "momentum-exchange link list overflowed");
  keys->resize(cnt);
  ms->resize(cnt);
  if (cnt) {
    HIP_TRY(s, hipMemcpyAsync(keys->data(), s->link_key, cnt * sizeof(int64_t),
                              hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipMemcpyAsync(ms->data(), s->link_m, cnt * sizeof(double), hipMemcpyDeviceToHost,
                              s->stream));
    if ((rc = sync(s))) return rc;
  }
  // Links were appended in arbitrary order; the reference visits cells in index
  // order and directions k = 0..26 within a cell, which is ascending key order.
  std::vector<int> ord(cnt);
  for (int j = 0; j < cnt; j++) ord[j] = j;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return (*keys)[a] < (*keys)[b]; });
  std::vector<int64_t> k2(cnt);
  std::vector<double> m2(cnt);
  for (int j = 0; j < cnt; j++) {
    k2[j] = (*keys)[ord[j]];
    m2[j] = (*ms)[ord[j]];
  }
  keys->swap(k2);
  ms->swap(m2);
  return LBM_OK;
}

static void link_to_record(const lbm_sim *s, int64_t key, double m, lbm_link *L) {
  const int64_t i = key / Q27;
  const int k = (int)(key % Q27);
  const int x = (int)(i % s->nx), y = (int)((i / s->nx) % s->ny), z = (int)(i / ((int64_t)s->nx * s->ny));
  const double ekx = EX27[k], eky = EY27[k], ekz = EZ27[k];
  // src/sim.js:533-535
  L->x = x - EX27[k];
  L->y = y - EY27[k];
  L->z = z - EZ27[k];
  L->k = k;
  L->px = -m * ekx;
  L->py = -m * eky;
  L->pz = -m * ekz;
}

extern "C" int lbm_stream_hydro_macro(lbm_sim *s, lbm_sphere *solid) {
  CHECK_SIM_F32(s);
  std::vector<int64_t> keys;
  std::vector<double> ms;
  int rc = stream_macro_pass(s, solid != nullptr, &keys, &ms);
  if (rc) return rc;
  if (solid) {
    for (size_t j = 0; j < keys.size(); j++) {
      lbm_link L;
      link_to_record(s, keys[j], ms[j], &L);
      sphere_apply_impulse(*solid, L.x, L.y, L.z, L.px, L.py, L.pz);
    }
  }
  return sync(s);
}

extern "C" int lbm_stream_hydro_macro_links(lbm_sim *s, int *count) {
  CHECK_SIM_F32(s);
  std::vector<int64_t> keys;
  std::vector<double> ms;
  int rc = stream_macro_pass(s, true, &keys, &ms);
  if (rc) return rc;
  s->links.resize(keys.size());
  for (size_t j = 0; j < keys.size(); j++) link_to_record(s, keys[j], ms[j], &s->links[j]);
  if (count) *count = (int)keys.size();
  return LBM_OK;
}

extern "C" int lbm_get_links(lbm_sim *s, lbm_link *out, int cap) {
  CHECK_SIM(s);
  if (!out && cap > 0) return fail(s, LBM_ERR_ARG, "out is NULL");
  const int m = std::min<int>(cap, (int)s->links.size());
  for (int j = 0; j < m; j++) out[j] = s->links[j];
  return LBM_OK;
}

// After a retag from sphere b, the solid cells are interior cells whose centre
// lies within r of b's centre (sphere_contains): along each axis within
// [c - r - 0.5, c + r - 0.5], and no more than the unit cubes around those
// centres fit in a ball of radius r + sqrt(3)/2.  Both bounds get a margin
// for rounding (one cell per side, radius r + 1).
static int64_t sphere_cell_bound(const lbm_sim *s, const lbm_sphere &b) {
  const double c[3] = {b.cx, b.cy, b.cz};
  const int n[3] = {s->nx, s->ny, s->nz};
  int64_t cells = 1;
  for (int a = 0; a < 3; a++) {
    double lo = 1, hi = n[a] - 2;  // interior cells
    if (std::isfinite(c[a]) && std::isfinite(b.r)) {
      lo = std::max(lo, std::floor(c[a] - b.r - 0.5) - 1);
      hi = std::min(hi, std::ceil(c[a] + b.r - 0.5) + 1);
    }
    cells *= hi >= lo ? (int64_t)(hi - lo) + 1 : 0;
  }
  if (std::isfinite(b.r)) {
    const double R = std::fabs(b.r) + 1;
    const double ball = std::ceil(4.0 / 3.0 * 3.141592653589793 * R * R * R);
    if (ball < (double)cells) cells = (int64_t)ball;
  }
  return cells;
}

// Sim.step with the library's sphere: the step's momentum exchange goes to the
// device sphere in the reference's order (links sorted by cell and direction,
// then applied one after the other), so the host waits once per step, for the
// sphere coming back, instead of copying, sorting and replaying the link list.
// Lists of up to kLinkSortLds links are sorted and applied by one kernel; a
// longer one is sorted by rocprim's radix sort once its length is known (one
// more wait).  The sphere ends as the host replay leaves it, bit for bit.
static int sim_step_sphere(lbm_sim *s, lbm_sphere *solid) {
  int rc;
  if ((rc = upload_sphere_async(s, solid))) return rc;
  if ((rc = retag_full(s, 1))) return rc;
  s->need_full_retag = true;
  // Every link has a solid source, and a cell sources at most 26 links; a list
  // that still overflows sets kErrListOverflow and the step fails.
  const int64_t bound = std::max<int64_t>(1, sphere_cell_bound(s, *solid) * 26);
  if ((rc = ensure_links(s, bound))) return rc;
  if ((rc = lbm_handle_fresh_cells(s))) return rc;
  if ((rc = lbm_compute_phi_gradients(s))) return rc;
  if ((rc = lbm_collide_phase(s))) return rc;
  if ((rc = lbm_stream_phase(s))) return rc;
  if ((rc = lbm_collide_hydro(s))) return rc;
  const Derived d = derive(s->params);
  HIP_TRY(s, hipMemsetAsync(&s->ctrl->n_links, 0, sizeof(int), s->stream));
  k_stream_hydro_macro<<<cell_grid(s), kCellBlock, 0, s->stream>>>(
      s->g, d, F2(s), F(s), H(s), s->rho, s->ux, s->uy, s->uz, PHI(s), s->tag, s->usx, s->usy,
      s->usz, s->gx, s->gy, s->gz, s->lap, s->bfx, s->bfy, s->bfz, s->bodyF_armed ? 1 : 0, 1,
      s->link_key, s->link_m, s->link_cap, s->ctrl);
  if ((rc = check_launch(s, "k_stream_hydro_macro"))) return rc;
  if (s->bodyF_armed) {
    const int64_t fb = s->g.S * sizeof(float);
    HIP_TRY(s, hipMemsetAsync(s->bfx, 0, fb, s->stream));
    HIP_TRY(s, hipMemsetAsync(s->bfy, 0, fb, s->stream));
    HIP_TRY(s, hipMemsetAsync(s->bfz, 0, fb, s->stream));
    s->bodyF_armed = false;
  }
  k_sort_apply_links<<<1, kApplyChunk, 0, s->stream>>>(s->g, s->link_key, s->link_m, s->ctrl,
                                                       s->link_cap);
  if ((rc = check_launch(s, "k_sort_apply_links"))) return rc;
  if ((rc = read_ctrl(s))) return rc;
  if (s->ctrl_host->n_links > s->link_cap || (s->ctrl_host->err & kErrListOverflow))
    return fail(s, LBM_ERR_CAPACITY, "momentum-exchange link list overflowed");
  if (s->ctrl_host->err & kErrLinksUnsorted) {
    const int cnt = s->ctrl_host->n_links;
    // Keys (cell * 27 + k) are below 2^end_bit.
    int end_bit = 1;
    while (end_bit < 62 && ((int64_t)1 << end_bit) <= s->g.n * Q27) end_bit++;
    if ((rc = ensure_link_sort(s, cnt, end_bit))) return rc;
    size_t tmp = s->sort_tmp_bytes;
    HIP_TRY(s, rocprim::radix_sort_pairs(s->sort_tmp, tmp, (uint64_t *)s->link_key,
                                         (uint64_t *)s->link_key2, s->link_m, s->link_m2,
                                         (unsigned)cnt, 0, end_bit, s->stream));
    k_apply_links<<<1, kApplyChunk, 0, s->stream>>>(s->g, s->link_key2, s->link_m2, s->ctrl, cnt);
    if ((rc = check_launch(s, "k_apply_links"))) return rc;
    s->ctrl_host->err &= ~kErrLinksUnsorted;
    HIP_TRY(s, hipMemcpyAsync(&s->ctrl->err, &s->ctrl_host->err, sizeof(int),
                              hipMemcpyHostToDevice, s->stream));
    if ((rc = read_ctrl(s))) return rc;
  }
  *solid = s->ctrl_host->sphere;
  return LBM_OK;
}

extern "C" int lbm_sim_step(lbm_sim *s, lbm_sphere *solid) {
  CHECK_SIM_F32(s);
  if (solid) return sim_step_sphere(s, solid);
  int rc;
  if ((rc = lbm_retag_cells(s, solid))) return rc;
  if ((rc = lbm_handle_fresh_cells(s))) return rc;
  if ((rc = lbm_compute_phi_gradients(s))) return rc;
  if ((rc = lbm_collide_phase(s))) return rc;
  if ((rc = lbm_stream_phase(s))) return rc;
  if ((rc = lbm_collide_hydro(s))) return rc;
  return lbm_stream_hydro_macro(s, solid);
}

// RigidSphere.absorbDeadCells (src/solid.js:87-101): the dead cells' impulses
// applied to the sphere on the device in the list's index order (the host's
// function, FP64), and the sphere read back in one transfer.
extern "C" int lbm_absorb_dead_cells(lbm_sim *s, lbm_sphere *solid) {
  CHECK_SIM(s);
  CHECK_NOT_LIVE(s);
  if (!solid) return fail(s, LBM_ERR_ARG, "solid is NULL");
  int rc;
  if ((rc = upload_sphere_async(s, solid))) return rc;
  const lbm_params &p = s->params;
  k_absorb_dead<<<1, kApplyChunk, 0, s->stream>>>(s->g, s->dead, s->ctrl, s->list_cap, p.rhoG,
                                                  p.rhoL - p.rhoG, PHI(s), s->ux, s->uy, s->uz);
  if ((rc = check_launch(s, "k_absorb_dead"))) return rc;
  if ((rc = read_ctrl(s))) return rc;
  *solid = s->ctrl_host->sphere;
  return LBM_OK;
}

// ------------------------------------------------------------------ fields

extern "C" int64_t lbm_field_count(const lbm_sim *s, int field) {
  if (!s) return -1;
  switch (field) {
    case LBM_FIELD_F:
    case LBM_FIELD_F2: return s->g.n * 27;
    case LBM_FIELD_H:
    case LBM_FIELD_H2: return s->g.n * 7;
    default: return (field >= 0 && field < LBM_FIELD_COUNT) ? s->g.n : -1;
  }
}

static void *field_ptr(lbm_sim *s, int field, int *q, size_t *elem) {
  *q = 1;
  *elem = sizeof(float);
  switch (field) {
    case LBM_FIELD_F: *q = 27; return F(s);
    case LBM_FIELD_F2: *q = 27; return F2(s);
    case LBM_FIELD_H: *q = 7; return H(s);
    case LBM_FIELD_H2: *q = 7; return H2(s);
    case LBM_FIELD_RHO: return s->rho;
    case LBM_FIELD_UX: return s->ux;
    case LBM_FIELD_UY: return s->uy;
    case LBM_FIELD_UZ: return s->uz;
    case LBM_FIELD_PHI: return PHI(s);
    case LBM_FIELD_PHIPREV: return PHIPREV(s);
    case LBM_FIELD_TAG: *elem = 1; return s->tag;
    case LBM_FIELD_TAGPREV: *elem = 1; return s->tagPrev;
    case LBM_FIELD_USX: return s->usx;
    case LBM_FIELD_USY: return s->usy;
    case LBM_FIELD_USZ: return s->usz;
    case LBM_FIELD_GX: return s->gx;
    case LBM_FIELD_GY: return s->gy;
    case LBM_FIELD_GZ: return s->gz;
    case LBM_FIELD_LAP: return s->lap;
    case LBM_FIELD_BODYFX: return s->bfx;
    case LBM_FIELD_BODYFY: return s->bfy;
    case LBM_FIELD_BODYFZ: return s->bfz;
    default: return nullptr;
  }
}

extern "C" int lbm_download(lbm_sim *s, int field, void *host) {
  CHECK_SIM(s);
  // Between live calls phi, phiPrev, the tags, the surface velocities, the
  // gradients (recomputed from phiPrev) and the cleared body force are the
  // reference's; populations, rho and u are not.
  if (s->live && (field <= LBM_FIELD_UZ)) CHECK_NOT_LIVE(s);
  if (field == LBM_FIELD_F || field == LBM_FIELD_F2 || field == LBM_FIELD_H ||
      field == LBM_FIELD_H2) {
    const int rc = ensure_f32(s, s->stream);
    if (rc) return rc;
  }
  if (!host) return fail(s, LBM_ERR_ARG, "host pointer is NULL");
  if (field >= LBM_FIELD_GX && field <= LBM_FIELD_LAP && s->grads_stale) {
    // After a fused step the gradients of the last step are those of phiPrev
    // (phi after that step's fresh-cell update).
    k_gradients<<<cell_grid(s), kCellBlock, 0, s->stream>>>(s->g, PHIPREV(s), s->gx, s->gy,
                                                            s->gz, s->lap);
    int rc = check_launch(s, "k_gradients");
    if (rc) return rc;
    s->grads_stale = false;
  }
  int q;
  size_t elem;
  void *dp = field_ptr(s, field, &q, &elem);
  if (!dp) return fail(s, LBM_ERR_ARG, "bad field id");
  const int64_t n = s->g.n;
  if (q == 1) {
    HIP_TRY(s, hipMemcpyAsync(host, dp, n * elem, hipMemcpyDeviceToHost, s->stream));
  } else {
    int rc = ensure_scratch(s, n * q * sizeof(float));
    if (rc) return rc;
    k_rows_to_aos<<<(n + 255) / 256, 256, 0, s->stream>>>(s->g, q, (const float *)dp,
                                                          (float *)s->scratch);
    if ((rc = check_launch(s, "k_rows_to_aos"))) return rc;
    HIP_TRY(s, hipMemcpyAsync(host, s->scratch, n * q * sizeof(float), hipMemcpyDeviceToHost,
                              s->stream));
  }
  return sync(s);
}

extern "C" int lbm_upload(lbm_sim *s, int field, const void *host) {
  CHECK_SIM_F32(s);
  if (!host) return fail(s, LBM_ERR_ARG, "host pointer is NULL");
  int q;
  size_t elem;
  void *dp = field_ptr(s, field, &q, &elem);
  if (!dp) return fail(s, LBM_ERR_ARG, "bad field id");
  const int64_t n = s->g.n;
  int rc;
  if (q == 1) {
    HIP_TRY(s, hipMemcpyAsync(dp, host, n * elem, hipMemcpyHostToDevice, s->stream));
  } else {
    if ((rc = ensure_scratch(s, n * q * sizeof(float)))) return rc;
    HIP_TRY(s, hipMemcpyAsync(s->scratch, host, n * q * sizeof(float), hipMemcpyHostToDevice,
                              s->stream));
    k_aos_to_rows<<<(n + 255) / 256, 256, 0, s->stream>>>(s->g, q, (const float *)s->scratch,
                                                          (float *)dp);
    if ((rc = check_launch(s, "k_aos_to_rows"))) return rc;
  }
  if (field == LBM_FIELD_BODYFX || field == LBM_FIELD_BODYFY || field == LBM_FIELD_BODYFZ)
    s->bodyF_armed = true;
  if (field == LBM_FIELD_TAG || field == LBM_FIELD_TAGPREV || field == LBM_FIELD_PHI ||
      field == LBM_FIELD_PHIPREV)
    s->need_full_retag = true;
  if (field >= LBM_FIELD_GX && field <= LBM_FIELD_LAP) s->grads_stale = false;
  return sync(s);
}

extern "C" int lbm_set_storage(lbm_sim *s, int kind) {
  CHECK_SIM(s);
  CHECK_NOT_LIVE(s);
  if (kind != LBM_STORAGE_F32 && kind != LBM_STORAGE_F16 && kind != LBM_STORAGE_BF16)
    return fail(s, LBM_ERR_ARG, "unknown storage kind");
  if (kind == s->storage) return LBM_OK;
  int rc = ensure_f32(s, s->stream);
  if (rc) return rc;
  if (kind != LBM_STORAGE_F32 && !s->cf[0]) {
    const int64_t S = s->g.S;
    for (int b = 0; b < 2; b++) {
      HIP_TRY(s, sim_malloc(s, &s->cf[b], S * 27 * 2));
      HIP_TRY(s, sim_malloc(s, &s->ch[b], S * 7 * 2));
      // The fused step never writes the populations of cells that are not
      // fluid: clear them so that what reads back there is defined (a zero
      // deviation) rather than whatever the allocation held.
      HIP_TRY(s, hipMemsetAsync(s->cf[b], 0, S * 27 * 2, s->stream));
      HIP_TRY(s, hipMemsetAsync(s->ch[b], 0, S * 7 * 2, s->stream));
    }
  }
  s->storage = kind;
  return sync(s);
}

extern "C" int lbm_get_storage(const lbm_sim *s) { return s ? s->storage : LBM_ERR_ARG; }

extern "C" int lbm_mass(lbm_sim *s, double *phi_sum, double *rho_sum) {
  CHECK_SIM(s);
  CHECK_NOT_LIVE(s);
  const int64_t n = s->g.n;
  const int64_t blocks = (n + 255) / 256;
  int rc = ensure_scratch(s, blocks * 2 * sizeof(double));
  if (rc) return rc;
  k_mass_partials<<<(unsigned)blocks, 256, 0, s->stream>>>(n, s->g, s->tag, PHI(s), s->rho,
                                                           (double *)s->scratch);
  if ((rc = check_launch(s, "k_mass_partials"))) return rc;
  std::vector<double> part(2 * blocks);
  HIP_TRY(s, hipMemcpyAsync(part.data(), s->scratch, 2 * blocks * sizeof(double),
                            hipMemcpyDeviceToHost, s->stream));
  if ((rc = sync(s))) return rc;
  double a = 0, b = 0;
  for (int64_t k = 0; k < blocks; k++) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  if (phi_sum) *phi_sum = a;
  if (rho_sum) *rho_sum = b;
  return LBM_OK;
}

extern "C" int lbm_check_fields(lbm_sim *s, int *all_finite, double *max_u) {
  CHECK_SIM(s);
  CHECK_NOT_LIVE(s);
  const int64_t n = s->g.n;
  const int64_t blocks = (n + 255) / 256;
  int rc = ensure_scratch(s, blocks * 2 * sizeof(double));
  if (rc) return rc;
  k_check_partials<<<(unsigned)blocks, 256, 0, s->stream>>>(n, s->tag, PHI(s), s->rho, s->ux,
                                                            s->uy, s->uz, (double *)s->scratch);
  if ((rc = check_launch(s, "k_check_partials"))) return rc;
  std::vector<double> part(2 * blocks);
  HIP_TRY(s, hipMemcpyAsync(part.data(), s->scratch, 2 * blocks * sizeof(double),
                            hipMemcpyDeviceToHost, s->stream));
  if ((rc = sync(s))) return rc;
  double bad = 0, um = 0;
  for (int64_t k = 0; k < blocks; k++) {
    bad = std::max(bad, part[2 * k]);
    um = std::max(um, part[2 * k + 1]);
  }
  if (all_finite) *all_finite = bad == 0 ? 1 : 0;
  if (max_u) *max_u = um;
  return LBM_OK;
}

extern "C" int lbm_pack_volume(lbm_sim *s, uint8_t *out) {
  CHECK_SIM(s);
  if (!out) return fail(s, LBM_ERR_ARG, "out is NULL");
  const int64_t n = s->g.n;
  int rc = ensure_scratch(s, 2 * n);
  if (rc) return rc;
  k_pack<<<(n + 255) / 256, 256, 0, s->stream>>>(n, PHI(s), s->tag, (uint8_t *)s->scratch);
  if ((rc = check_launch(s, "k_pack"))) return rc;
  HIP_TRY(s, hipMemcpyAsync(out, s->scratch, 2 * n, hipMemcpyDeviceToHost, s->stream));
  return sync(s);
}

static int copy_list(lbm_sim *s, bool fresh, int32_t *out, int cap, int *count) {
  int rc = read_ctrl(s);
  if (rc) return rc;
  const int cnt = std::min(fresh ? s->ctrl_host->n_fresh : s->ctrl_host->n_dead, s->list_cap);
  if (count) *count = cnt;
  const int m = std::min(cnt, cap);
  if (m > 0 && out) {
    HIP_TRY(s, hipMemcpyAsync(out, fresh ? s->fresh : s->dead, m * sizeof(int32_t),
                              hipMemcpyDeviceToHost, s->stream));
    return sync(s);
  }
  return LBM_OK;
}

extern "C" int lbm_fresh_cells(lbm_sim *s, int32_t *out, int cap, int *count) {
  CHECK_SIM(s);
  return copy_list(s, true, out, cap, count);
}

extern "C" int lbm_dead_cells(lbm_sim *s, int32_t *out, int cap, int *count) {
  CHECK_SIM(s);
  return copy_list(s, false, out, cap, count);
}

// ------------------------------------------------------------------ sphere (host)

extern "C" void lbm_sphere_init(lbm_sphere *b, int nx, int ny, int nz, double cx, double cy,
                                double cz, double r, double density) {
  if (!b) return;
  std::memset(b, 0, sizeof *b);
  b->nx = nx;
  b->ny = ny;
  b->nz = nz;
  // src/solid.js:11-14, 24: `opts.x ?? default`.
  b->cx = std::isnan(cx) ? nx * 0.5 : cx;
  b->cy = std::isnan(cy) ? ny * 0.3 : cy;
  b->cz = std::isnan(cz) ? nz * 0.5 : cz;
  b->r = std::isnan(r) ? std::min(std::min(nx, ny), nz) * 0.10 : r;
  b->density = std::isnan(density) ? 1.4 : density;
  sphere_recompute_mass(*b);
}

extern "C" void lbm_sphere_recompute_mass(lbm_sphere *b) {
  if (b) sphere_recompute_mass(*b);
}

extern "C" int lbm_sphere_contains(const lbm_sphere *b, int x, int y, int z) {
  return b ? (sphere_contains(*b, x, y, z) ? 1 : 0) : 0;
}

extern "C" void lbm_sphere_surface_velocity(const lbm_sphere *b, int x, int y, int z,
                                            double out[3]) {
  if (b && out) sphere_surface_velocity(*b, x, y, z, out);
}

extern "C" void lbm_sphere_apply_impulse_at_cell(lbm_sphere *b, int x, int y, int z, double px,
                                                 double py, double pz) {
  if (b) sphere_apply_impulse(*b, x, y, z, px, py, pz);
}

extern "C" void lbm_sphere_step(lbm_sphere *b, double dt, double gravity) {
  if (b) sphere_step(*b, dt, gravity);
}

extern "C" int lbm_lattice_tables(int32_t ex27[27], int32_t ey27[27], int32_t ez27[27],
                                  float w27_[27], int32_t opp27_[27], int32_t ex7[7],
                                  int32_t ey7[7], int32_t ez7[7], float w7_[7], int32_t opp7_[7]) {
  for (int k = 0; k < 27; k++) {
    ex27[k] = EX27[k];
    ey27[k] = EY27[k];
    ez27[k] = EZ27[k];
    w27_[k] = w27(k);
    opp27_[k] = opp27(k);
  }
  for (int k = 0; k < 7; k++) {
    ex7[k] = EX7[k];
    ey7[k] = EY7[k];
    ez7[k] = EZ7[k];
    w7_[k] = w7(k);
    opp7_[k] = opp7(k);
  }
  return LBM_OK;
}

// ------------------------------------------------------------------ timing

extern "C" int lbm_set_timing(lbm_sim *s, int on) {
  CHECK_SIM(s);
  int rc = sync(s);
  if (rc) return rc;
  collect_timing(s->timer);
  s->timer.on = on != 0;
  s->timer.total.clear();
  return LBM_OK;
}

static void timer_read(Timer &t, const char *name, double *total_ms, int64_t *launches) {
  collect_timing(t);
  auto it = t.total.find(name);
  if (total_ms) *total_ms = it == t.total.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == t.total.end() ? 0 : it->second.second;
}

extern "C" int lbm_kernel_time(lbm_sim *s, const char *name, double *total_ms, int64_t *launches) {
  CHECK_SIM(s);
  if (!name) return LBM_ERR_ARG;
  int rc = sync(s);
  if (rc) return rc;
  timer_read(s->timer, name, total_ms, launches);
  return LBM_OK;
}

static int graph_steps_ok(int steps) { return steps == 0 || (steps >= 2 && steps <= 64 && !(steps & 1)); }

extern "C" int lbm_set_graph_steps(lbm_sim *s, int steps) {
  CHECK_SIM(s);
  if (!graph_steps_ok(steps)) return fail(s, LBM_ERR_ARG, "graph steps must be 0 or even in 2..64");
  int rc = sync(s);
  if (rc) return rc;
  s->graphs.clear();
  s->graphs.steps = steps;
  return LBM_OK;
}

extern "C" int lbm_set_stream_split(lbm_sim *s, int mode) {
  CHECK_SIM(s);
  if (mode < 0 || mode > 2) return fail(s, LBM_ERR_ARG, "stream split must be 0, 1 or 2");
  s->stream_split = mode;
  s->graphs.clear();
  return LBM_OK;
}

extern "C" int lbm_set_march_shape(lbm_sim *s, int rows, int segments) {
  CHECK_SIM(s);
  if (rows != 1 && rows != 2) return fail(s, LBM_ERR_ARG, "march rows must be 1 or 2");
  if (segments < 0) return fail(s, LBM_ERR_ARG, "march segments < 0");
  s->march_rows = rows;
  s->march_nseg = segments;
  s->graphs.clear();
  return LBM_OK;
}

extern "C" int lbm_set_stream_lanes(lbm_sim *s, int lanes) {
  CHECK_SIM(s);
  if (lanes != 1 && lanes != 2) return fail(s, LBM_ERR_ARG, "cells per lane must be 1 or 2");
  s->lanes = lanes;
  return LBM_OK;
}

extern "C" int lbm_graph_counts(const lbm_sim *s, int64_t *captures, int64_t *replays) {
  CHECK_SIM(s);
  if (captures) *captures = s->graphs.captures;
  if (replays) *replays = s->graphs.replays;
  return LBM_OK;
}

// ------------------------------------------------------------------ bandwidth probe

__global__ __launch_bounds__(256) void k_copy_probe(const float4 *__restrict__ a,
                                                    float4 *__restrict__ b, int64_t n4) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4) b[i] = a[i];
}

// The population pattern of the fused stream kernels with no arithmetic: every
// one of the 27 + 7 directions pulled from its shifted source (the own element
// where the source is outside the grid) and stored, non-temporally, in the
// kernels' tiles, order and layout: row blocks of 4-byte populations with one
// cell per lane, or planes of 16-bit populations with two cells per lane in
// aligned 4-byte pairs (an odd offset straddles two pairs).  As in the kernels,
// every pull is issued before the first store: buffer loads are not moved past
// buffer stores by the compiler, and a probe that stores each direction after
// its load runs the 16-bit pattern at 4.2 instead of 5.0 TB/s (256^3, MI355X).
__global__ __launch_bounds__(256) void k_pull_probe32(const float *__restrict__ src,
                                                      float *__restrict__ dst, Geom g,
                                                      TileMap tm) {
  int tile;
  if (!map_tile(tm.T, tm.T8, blockIdx.x, tile)) return;
  const int64_t i = ((int64_t)tile << kTileShift) + threadIdx.x;
  if (i >= g.n) return;
  int x, y, z;
  cell_xyz(g.nx, g.nxny, i, x, y, z);
  const RowAddr ra = row_addr(g, i, z * g.ny + y);
  const float *lo = pull_base(src, ra, g);
  float v[Q27 + Q7];
#pragma unroll
  for (int k = 0; k < Q27 + Q7; k++) {
    const int kk = k < Q27 ? k : k - Q27;
    const int ex = EX27[kk], ey = EY27[kk], ez = EZ27[kk];
    const bool in = x - ex >= 0 && x - ex < g.nx && y - ey >= 0 && y - ey < g.ny &&
                    z - ez >= 0 && z - ez < g.nz;
    // Direction k of the merged row block; the opposite of an h direction is
    // another h direction.
    const int kopp = k < Q27 ? opp27(k) : Q27 + opp7(kk);
    v[k] = pull_row(lo, ra.relb, g, k, kopp, ex, ey, ez, !in);
  }
#pragma unroll
  for (int k = 0; k < Q27 + Q7; k++) stn(dst + ra.base, ra.relb + (uint32_t)k * (uint32_t)g.pb * 4u, v[k]);
}
__global__ __launch_bounds__(128) void k_pull_probe16(const uint16_t *__restrict__ src,
                                                      uint16_t *__restrict__ dst, Geom g,
                                                      TileMap tm) {
  int tile;
  if (!map_tile(tm.T, tm.T8, blockIdx.x, tile)) return;
  const uint32_t i0 = (uint32_t)(((int64_t)tile << kTileShift) + 128 * threadIdx.y + 2 * threadIdx.x);
  const uint32_t S32 = (uint32_t)g.S;
  const __amdgpu_buffer_rsrc_t r = pop_rsrc(src, (uint32_t)((Q27 + Q7) * 2 * g.S));
  const __amdgpu_buffer_rsrc_t w = pop_rsrc(dst, (uint32_t)((Q27 + Q7) * 2 * g.S));
  uint32_t lo[Q27 + Q7], hi[Q27 + Q7];
#pragma unroll
  for (int k = 0; k < Q27 + Q7; k++) {
    const int kk = k < Q27 ? k : k - Q27;
    const uint32_t off =
        (uint32_t)((int64_t)EZ27[kk] * g.nxny + (int64_t)EY27[kk] * g.nx + EX27[kk]);
    const uint32_t e = (uint32_t)k * S32 + i0 - off;
    if (off & 1u) {
      lo[k] = ld_pair<true>(r, e - 1u);
      hi[k] = ld_pair<true>(r, e + 1u);
    } else {
      lo[k] = ld_pair<true>(r, e);
      hi[k] = 0u;
    }
  }
  if (i0 >= (uint32_t)g.n) return;
#pragma unroll
  for (int k = 0; k < Q27 + Q7; k++) {
    const int kk = k < Q27 ? k : k - Q27;
    const uint32_t off =
        (uint32_t)((int64_t)EZ27[kk] * g.nxny + (int64_t)EY27[kk] * g.nx + EX27[kk]);
    st_pair(w, (uint32_t)k * S32 + i0, (off & 1u) ? (lo[k] >> 16) | (hi[k] << 16) : lo[k]);
  }
}

extern "C" int lbm_probe_pull_bandwidth(int nx, int ny, int nz, int storage, int reps,
                                        double *gbps) {
  if (nx < 3 || ny < 3 || nz < 3 || reps < 1 || !gbps) return LBM_ERR_ARG;
  if (storage != LBM_STORAGE_F32 && storage != LBM_STORAGE_F16 && storage != LBM_STORAGE_BF16)
    return LBM_ERR_ARG;
  Geom g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.nxny = nx * ny;
  g.n = (int64_t)g.nxny * nz;
  g.S = (g.n + 63) / 64 * 64;
  choose_layout(g);  // as lbm_create
  const int elem = storage == LBM_STORAGE_F32 ? 4 : 2;
  const int64_t bytes = (int64_t)(Q27 + Q7) * g.S * elem;
  if (elem == 4 ? !f32_fused_fits(g) : bytes >= ((int64_t)1 << 32)) return LBM_ERR_ARG;
  const TileMap tm = make_tile_map(nx, ny, nz);
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LBM_ERR_HIP;
  float ms = 0;
  if (dev_malloc(&a, bytes) != hipSuccess || dev_malloc(&b, bytes) != hipSuccess) {
    rc = LBM_ERR_OOM;
    goto done;
  }
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  if (hipMemsetAsync(a, 0, bytes, st) != hipSuccess) goto done;
  for (int r = -1; r < reps; r++) {  // r = -1: warm-up
    if (r == 0) hipEventRecord(e0, st);
    if (elem == 4)
      k_pull_probe32<<<tm.T8, 256, 0, st>>>((const float *)a, (float *)b, g, tm);
    else
      k_pull_probe16<<<tm.T8, dim3(64, 2, 1), 0, st>>>((const uint16_t *)a, (uint16_t *)b, g, tm);
  }
  hipEventRecord(e1, st);
  if (hipEventSynchronize(e1) != hipSuccess) goto done;
  hipEventElapsedTime(&ms, e0, e1);
  *gbps = 2.0 * (Q27 + Q7) * elem * (double)g.n * reps / (ms * 1e-3) / 1e9;
  rc = LBM_OK;
done:
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  if (a) dev_free(a);
  if (b) dev_free(b);
  return rc;
}

extern "C" int lbm_probe_copy_bandwidth(int64_t bytes, int reps, double *gbps) {
  if (bytes < 16 || reps < 1 || !gbps) return LBM_ERR_ARG;
  const int64_t n4 = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LBM_ERR_HIP;
  float ms = 0;
  if (dev_malloc(&a, n4 * 16) != hipSuccess || dev_malloc(&b, n4 * 16) != hipSuccess) {
    rc = LBM_ERR_OOM;
    goto done;
  }
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  if (hipMemsetAsync(a, 0, n4 * 16, st) != hipSuccess) goto done;
  {
    const int64_t blocks = (n4 + 255) / 256;
    k_copy_probe<<<blocks, 256, 0, st>>>(a, b, n4);  // warm-up
    hipEventRecord(e0, st);
    for (int r = 0; r < reps; r++) k_copy_probe<<<blocks, 256, 0, st>>>(a, b, n4);
    hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) goto done;
    hipEventElapsedTime(&ms, e0, e1);
    *gbps = 2.0 * (double)n4 * 16 * reps / (ms * 1e-3) / 1e9;
    rc = LBM_OK;
  }
done:
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  if (a) dev_free(a);
  if (b) dev_free(b);
  return rc;
}

// ------------------------------------------------------------------ coupled step

#include "coupled.inc"

// ------------------------------------------------------------------ serving

#include "serve.inc"
