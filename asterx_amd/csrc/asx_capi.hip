// C ABI of the substep library (include/asterx.h): argument validation,
// parameter resolution and kernel launches.  Host code only.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "asx_kernels.hpp"

using namespace asx;

namespace {

// A roctx range over one entry point's host call: `rocprofv3 --marker-trace`
// shows it beside the kernels the call launched (a batch plan's graph
// replays and persistent stages get one each).
struct Range {
  explicit Range(const char *name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range &) = delete;
  Range &operator=(const Range &) = delete;
};

constexpr int kStateSlots[ASX_NUM_STATE] = {
    ASX_DENS, ASX_MOMX, ASX_MOMY, ASX_MOMZ, ASX_TAU, ASX_DENT,
    ASX_DYE,  ASX_AVEC_X, ASX_AVEC_Y, ASX_AVEC_Z, ASX_PSI};

int centering(int s) {
  if (s >= ASX_ALP && s <= ASX_KZZ) return 7;
  if (s >= ASX_RHO && s <= ASX_MASK_CC) return 0;
  if (s == ASX_MASK_VC) return 7;
  if (s == ASX_DBX_STAG) return 1;
  if (s == ASX_DBY_STAG) return 2;
  if (s == ASX_DBZ_STAG) return 4;
  if (s == ASX_AVEC_X) return 6;
  if (s == ASX_AVEC_Y) return 5;
  if (s == ASX_AVEC_Z) return 3;
  if (s == ASX_PSI) return 7;
  if (s >= ASX_FXDENS && s <= ASX_FXBZ) return 1;
  if (s >= ASX_FYDENS && s <= ASX_FYBZ) return 2;
  if (s >= ASX_FZDENS && s <= ASX_FZBZ) return 4;
  if (s == ASX_VTILDE_Y_XFACE || s == ASX_VTILDE_Z_XFACE) return 1;
  if (s == ASX_VTILDE_X_YFACE || s == ASX_VTILDE_Z_YFACE) return 2;
  if (s == ASX_VTILDE_X_ZFACE || s == ASX_VTILDE_Y_ZFACE) return 4;
  if (s == ASX_AMAX_XFACE || s == ASX_AMIN_XFACE) return 1;
  if (s == ASX_AMAX_YFACE || s == ASX_AMIN_YFACE) return 2;
  if (s == ASX_AMAX_ZFACE || s == ASX_AMIN_ZFACE) return 4;
  if (s >= ASX_FX && s <= ASX_G) return 7;
  if (s >= ASX_DENSRHS && s <= ASX_DYE_RHS) return 0;
  if (s == ASX_AVEC_X_RHS) return 6;
  if (s == ASX_AVEC_Y_RHS) return 5;
  if (s == ASX_AVEC_Z_RHS) return 3;
  if (s == ASX_PSI_RHS) return 7;
  if (s >= ASX_ETTT && s <= ASX_ETZZ) return 7;
  if (s == ASX_REGRID_ERROR) return 0;
  return -1;
}

void strides(const int32_t n[3], int32_t ng, int c, int64_t &sj, int64_t &sk) {
  const int64_t ex = n[0] + 2 * ng + (c & 1);
  const int64_t ey = n[1] + 2 * ng + ((c >> 1) & 1);
  sj = ex;
  sk = ex * ey;
}

int64_t npoints(const int32_t n[3], int32_t ng, int c) {
  int64_t p = 1;
  for (int d = 0; d < 3; ++d) p *= n[d] + 2 * ng + ((c >> d) & 1);
  return p;
}

asx_status check_geom(const asx_patch *pt) {
  if (!pt) return ASX_ERR_NULL;
  for (int d = 0; d < 3; ++d)
    if (pt->n[d] < 1 || !(pt->dx[d] > 0.0)) return ASX_ERR_SIZE;
  if (pt->ng < 1) return ASX_ERR_GHOSTS;
  // int32 offsets inside kernels
  if (npoints(pt->n, pt->ng, 7) >= (int64_t(1) << 31)) return ASX_ERR_SIZE;
  return ASX_OK;
}

asx_status need(const asx_patch *pt, std::initializer_list<int> slots) {
  for (int s : slots) {
    const asx_gf &g = pt->gf[s];
    if (!g.data) return ASX_ERR_NULL;
    int64_t sj, sk;
    strides(pt->n, pt->ng, centering(s), sj, sk);
    if (g.sj != sj || g.sk != sk) return ASX_ERR_LAYOUT;
  }
  return ASX_OK;
}
asx_status need_range(const asx_patch *pt, int a, int b) {
  for (int s = a; s <= b; ++s) {
    asx_status st = need(pt, {s});
    if (st) return st;
  }
  return ASX_OK;
}
asx_status opt(const asx_patch *pt, int s) {
  if (!pt->gf[s].data) return ASX_OK;
  return need(pt, {s});
}

KPatch kpatch(const asx_patch *pt) {
  KPatch P;
  for (int s = 0; s < ASX_NUM_GF; ++s) P.g[s] = pt->gf[s].data;
  for (int c = 0; c < 8; ++c) {
    int64_t sj, sk;
    strides(pt->n, pt->ng, c, sj, sk);
    P.sj[c] = (int)sj;
    P.sk[c] = (int)sk;
  }
  for (int d = 0; d < 3; ++d) {
    P.n[d] = pt->n[d];
    P.dx[d] = pt->dx[d];
    P.idx[d] = 1.0 / pt->dx[d];
    P.x0[d] = pt->x0[d];
  }
  P.ng = pt->ng;
  return P;
}

Box box(const asx_patch *pt, int c, const int halo[3]) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = pt->ng - halo[d];
    b.cnt[d] = pt->n[d] + ((c >> d) & 1) + 2 * halo[d];
  }
  return b;
}

RParams rparams(const asx_params *p) {
  RParams r;
  r.ppm_shock_detection = p->ppm_shock_detection;
  r.ppm_zone_flattening = p->ppm_zone_flattening;
  r.kgamma = p->poly_k * p->poly_gamma;
  r.ppm_eta1 = p->ppm_eta1;
  r.ppm_eta2 = p->ppm_eta2;
  r.ppm_eps = p->ppm_eps;
  r.ppm_eps_shock = p->ppm_eps_shock;
  r.ppm_small = p->ppm_small;
  r.ppm_omega1 = p->ppm_omega1;
  r.ppm_omega2 = p->ppm_omega2;
  r.weno_eps = p->weno_eps;
  r.mp5_alpha = p->mp5_alpha;
  return r;
}

// EOSX/src/setup_eos.cxx:284-305 + eos_3p_idealgas.hxx:117-135
IdealGas ideal_gas(const asx_params *p) {
  IdealGas e;
  e.gamma = p->gl_gamma;
  e.gm1 = p->gl_gamma - 1;
  e.inv_gamma = 1 / p->gl_gamma;
  e.igm1 = 1 / e.gm1;
  e.eps_min = p->eps_min;
  e.eps_max = p->eps_max;
  if (e.gamma > 2) {
    const double cap = 1 / (e.gamma * (e.gamma - 2));
    e.eps_max = (cap < e.eps_max) ? cap : e.eps_max;
  }
  e.rho_min = p->rho_min;
  e.rho_max = p->rho_max;
  e.ye_min = p->ye_min;
  e.ye_max = p->ye_max;
  e.temp_over_eps = e.gm1 * p->particle_mass;
  e.temp_min = e.temp_over_eps * e.eps_min;
  e.temp_max = e.temp_over_eps * e.eps_max;
  return e;
}

C2PArgs c2p_args(const asx_params *p) {
  C2PArgs a;
  a.prime = p->c2p_prime;
  a.second = p->c2p_second;
  a.max_iter = p->max_iter;
  a.excise = p->excise;
  a.use_mask = p->use_mask;
  a.entropy_fix = p->use_entropy_fix;
  a.use_z = p->use_z;
  a.use_temp = p->use_temperature;
  a.use_press_atmo = p->use_press_atmo;
  a.thermal_eos_atmo = p->thermal_eos_atmo;
  a.tol = p->c2p_tol;
  a.alp_thresh = p->alp_thresh;
  a.cons_error = p->cons_error_limit;
  a.vw_lim = p->vw_lim;
  a.w_lim = std::sqrt(1.0 + p->vw_lim * p->vw_lim);
  a.v_lim = p->vw_lim / a.w_lim;
  a.Bsq_lim = p->B_lim * p->B_lim;
  a.rho_BH = p->rho_BH;
  a.eps_BH = p->eps_BH;
  a.vwlim_BH = p->vwlim_BH;
  a.rho_abs_min = p->rho_abs_min;
  a.atmo_tol = p->atmo_tol;
  a.p_atmo = p->p_atmo;
  a.t_atmo = p->t_atmo;
  a.Ye_atmo = p->Ye_atmo;
  a.r_atmo = p->r_atmo;
  a.n_rho_atmo = p->n_rho_atmo;
  a.n_press_atmo = p->n_press_atmo;
  a.n_temp_atmo = p->n_temp_atmo;
  const double n = 1.0 / (p->poly_gamma - 1.0);
  a.poly_gamma = p->poly_gamma;
  a.poly_np1 = n + 1;
  a.poly_invn = 1.0 / n;
  a.poly_rho_p = std::pow(p->poly_k, -n);
  // c2p_1DPalenzuela.hxx:388-390
  a.minbits = int(std::fabs(std::log(p->c2p_tol)) / std::log(2.0));
  a.tol0 = std::ldexp(1.0, -a.minbits);
  return a;
}

asx_status check_table(const asx_params *p);
bool tabulated(const asx_params *p) { return p->evolution_eos == ASX_EOS_TABULATED; }

// A table (with its descriptor) or the ideal gas.
asx_status check_par(const asx_params *p) {
  if (!p) return ASX_ERR_NULL;
  if (tabulated(p)) {
    if (asx_status st = check_table(p)) return st;
  } else if (p->evolution_eos != ASX_EOS_IDEALGAS) {
    return ASX_ERR_UNSUPPORTED;
  }
  if (p->reconstruction_method == ASX_RECON_EPPM) return ASX_ERR_UNSUPPORTED;
  if (p->reconstruction_method < 0 || p->reconstruction_method > 6) return ASX_ERR_PARAM;
  if (p->loworder_method < 0 || p->loworder_method > 2) return ASX_ERR_PARAM;
  if (p->flux_type != ASX_FLUX_LXF && p->flux_type != ASX_FLUX_HLLE) return ASX_ERR_PARAM;
  if (p->recon_type < 0 || p->recon_type > 2) return ASX_ERR_PARAM;
  if (p->vector_potential_gauge != ASX_GAUGE_ALGEBRAIC &&
      p->vector_potential_gauge != ASX_GAUGE_GENERALIZED_LORENZ)
    return ASX_ERR_PARAM;
  return ASX_OK;
}

// A table described by asx_eos_table, as EOSX sets it up
// (eos_3p_tabulated3d.hxx:165-183): idx from the first spacing, the ranges
// from the axis ends, T inverted to 1e-14 in log T (:217).
asx_status check_table(const asx_params *p) {
  const asx_eos_table *t = p->eos_table;
  if (!t || !t->data || !t->axis[0] || !t->axis[1] || !t->axis[2]) return ASX_ERR_NULL;
  if (t->nvars != 4) return ASX_ERR_PARAM;
  for (int d = 0; d < 3; ++d)
    if (t->n[d] < 2 || !(t->x_step[d] > 0.0)) return ASX_ERR_PARAM;
  if (!(t->eps_max >= t->eps_min)) return ASX_ERR_PARAM;
  return ASX_OK;
}
TabEOS tab_eos(const asx_params *p) {
  const asx_eos_table *t = p->eos_table;
  TabEOS e;
  e.tab = t->data;
  for (int d = 0; d < 3; ++d) {
    e.ax[d] = t->axis[d];
    e.n[d] = t->n[d];
    e.idx[d] = 1. / t->x_step[d];
  }
  e.shift = t->energy_shift;
  e.rho_min = std::exp(t->x_lo[0]);
  e.rho_max = std::exp(t->x_hi[0]);
  e.temp_min = std::exp(t->x_lo[1]);
  e.temp_max = std::exp(t->x_hi[1]);
  e.ye_min = t->x_lo[2];
  e.ye_max = t->x_hi[2];
  e.eps_min = t->eps_min;
  e.eps_max = t->eps_max;
  return e;
}
// An ideal gas or a table (entry points with composition)
asx_status check_eos(const asx_params *p) {
  if (!p) return ASX_ERR_NULL;
  if (p->evolution_eos == ASX_EOS_TABULATED) return check_table(p);
  if (p->evolution_eos != ASX_EOS_IDEALGAS) return ASX_ERR_UNSUPPORTED;
  return ASX_OK;
}

asx_status check_c2p_par(const asx_params *p) {
  asx_status st = check_par(p);
  if (st) return st;
  // a table: Palenzuela on temperature without a pressure atmosphere (Noble
  // needs the ideal-fluid Gamma; eps_from_valid_rho_press_ye and p, eps
  // from (rho, kappa), which the Entropy solver needs, are refused)
  if (tabulated(p) && (p->c2p_prime != ASX_C2P_PALENZUELA || p->c2p_second == ASX_C2P_NOBLE ||
                       p->c2p_second == ASX_C2P_ENTROPY || p->use_entropy_fix ||
                       !p->use_temperature || p->use_press_atmo))
    return ASX_ERR_UNSUPPORTED;
  if (p->c2p_prime != ASX_C2P_NOBLE && p->c2p_prime != ASX_C2P_PALENZUELA &&
      p->c2p_prime != ASX_C2P_ENTROPY)
    return ASX_ERR_UNSUPPORTED;
  if (p->c2p_second != ASX_C2P_NOBLE && p->c2p_second != ASX_C2P_PALENZUELA &&
      p->c2p_second != ASX_C2P_ENTROPY && p->c2p_second != ASX_C2P_NONE)
    return ASX_ERR_UNSUPPORTED;
  if (p->max_iter < 1) return ASX_ERR_PARAM;
  if (!(p->c2p_tol > 0.0)) return ASX_ERR_PARAM;
  return ASX_OK;
}

// Faces with a table: reconstruction in T (a table has no eps(rho, p)) and
// the high-order methods the table kernels are built for.
asx_status check_flux_eos(const asx_params *p) {
  if (!tabulated(p)) return ASX_OK;
  if (!p->reconstruct_with_temperature || !flux_tab_method(p->reconstruction_method))
    return ASX_ERR_UNSUPPORTED;
  return ASX_OK;
}

int recon_ghosts(int m) {
  return m == ASX_RECON_GODUNOV ? 1
         : (m == ASX_RECON_MINMOD || m == ASX_RECON_MONOCENTRAL) ? 2
                                                                  : 3;
}

asx_status dev_status(hipError_t e) { return e == hipSuccess ? ASX_OK : ASX_ERR_DEVICE; }

// Per-device buffers: counters, and face scratch per (stream, shape).
struct Scratch {
  std::vector<double *> bufs;
  double *slot[ASX_NUM_GF] = {};
};
std::mutex g_mu;
std::map<int, unsigned long long *> g_cnt;
std::map<std::tuple<int, void *, int, int, int, int>, Scratch> g_scratch;

// ---------------------------------------------------------------------------
// Kernel timing (diagnostics): HIP events around launches, on the caller's
// stream, per kernel class; events come from a per-device pool.
// ---------------------------------------------------------------------------
struct TimedLaunch {
  int dev, cls;
  hipEvent_t a, b;
};
std::atomic<bool> g_timing{false};
std::mutex g_tmu;
std::map<int, std::vector<hipEvent_t>> g_pool;
std::vector<TimedLaunch> g_pending;
double g_tms[ASX_KC_COUNT];
int64_t g_tn[ASX_KC_COUNT];

hipEvent_t take_event(int dev) {
  std::vector<hipEvent_t> &v = g_pool[dev];
  if (!v.empty()) {
    hipEvent_t e = v.back();
    v.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Drain finished (blocking) launches into the accumulators.
asx_status drain_timing() {
  for (const TimedLaunch &t : g_pending) {
    if (hipEventSynchronize(t.b) != hipSuccess) return ASX_ERR_DEVICE;
    float ms = 0;
    if (hipEventElapsedTime(&ms, t.a, t.b) != hipSuccess) return ASX_ERR_DEVICE;
    g_tms[t.cls] += ms;
    g_tn[t.cls] += 1;
    g_pool[t.dev].push_back(t.a);
    g_pool[t.dev].push_back(t.b);
  }
  g_pending.clear();
  return ASX_OK;
}

// Set while this thread records a graph: launches are then not timed.
thread_local bool g_capturing = false;

struct KTimer {
  int dev = 0, cls;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  KTimer(int c, hipStream_t st) : cls(c), s(st) {
    if (g_capturing || !g_timing.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> l(g_tmu);
    (void)hipGetDevice(&dev);
    a = take_event(dev);
    b = take_event(dev);
    if (!a || !b || hipEventRecord(a, s) != hipSuccess) a = nullptr;
  }
  ~KTimer() {
    if (!a) return;
    std::lock_guard<std::mutex> l(g_tmu);
    if (hipEventRecord(b, s) == hipSuccess) g_pending.push_back({dev, cls, a, b});
  }
};
template <class F> hipError_t timed(int cls, hipStream_t s, F &&f) {
  KTimer t(cls, s);
  return f();
}

// Launch-shape variant per device and kernel class (default 0).
std::map<int, std::array<int, ASX_KC_COUNT>> g_var;

int variant_count(int kc) {
  if (kc == ASX_KC_FACE_X || kc == ASX_KC_FACE_Y || kc == ASX_KC_FACE_Z) return kFaceVariants;
  if (kc == ASX_KC_UPDATE) return kUpdateVariants;
  if (kc == ASX_KC_TMUNU) return kTmunuVariants;
  if (kc == ASX_KC_MLP) return 2;
  return 1;
}
int variant_of(int kc) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_var.find(dev);
  return it == g_var.end() ? 0 : it->second[kc];
}
void set_variant(int kc, int v) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_var.find(dev);
  if (it == g_var.end()) {
    std::array<int, ASX_KC_COUNT> z{};
    it = g_var.emplace(dev, z).first;
  }
  it->second[kc] = v;
}

unsigned long long *counter_buf(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cnt.find(dev);
  if (it != g_cnt.end()) return it->second;
  unsigned long long *p = nullptr;
  if (hipMalloc(&p, 80 * sizeof(unsigned long long)) != hipSuccess) return nullptr;
  g_cnt[dev] = p;
  return p;
}

// Queue of zones the primary con2prim solver failed, per (device, stream):
// one int per interior zone and a count, grown on demand.
struct C2PQueue {
  int *q = nullptr;
  unsigned *n = nullptr;
  int64_t cap = 0;
};
std::map<std::pair<int, void *>, C2PQueue> g_c2pq;

C2PQueue *c2p_queue(void *stream, int64_t nz) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_mu);
  C2PQueue &Q = g_c2pq[{dev, stream}];
  if (!Q.n && hipMalloc(&Q.n, sizeof(unsigned)) != hipSuccess) return nullptr;
  if (Q.cap < nz) {
    // the previous buffer may still be read by work queued on this stream
    if (Q.q && (hipStreamSynchronize((hipStream_t)stream) != hipSuccess ||
                hipFree(Q.q) != hipSuccess))
      return nullptr;
    Q.q = nullptr;
    Q.cap = 0;
    if (hipMalloc(&Q.q, (size_t)nz * sizeof(int)) != hipSuccess) return nullptr;
    Q.cap = nz;
  }
  return &Q;
}

// Device buffers of one con2prim call: the queue of zones the primary
// solver failed (one int per interior zone), its count, and the counters.
struct C2PScratch {
  int *queue;
  unsigned *qcount;
  unsigned long long *cnt; // 80 counters
};

// Fixed device scratch of asx_con2prim_group_noalloc: the counters (80
// words at byte 0), the queue count (byte 1024) and the queue (one int per
// interior zone from byte 2048), for patches of up to kFixedC2PZones
// interior zones.  The runtime places one copy per device when it loads the
// code object; it lives until the process ends.  In host code `scratch` is
// only a handle for hipGetSymbolAddress: every host use goes through the
// per-device address fixed_c2p_scratch() returns.
constexpr int64_t kFixedC2PZones = int64_t(1) << 24; // 256^3
constexpr size_t kFixedC2PQueueByte = 2048;
static __device__ __align__(256) unsigned long long
    scratch[kFixedC2PQueueByte / 8 + kFixedC2PZones / 2];
std::map<int, unsigned char *> g_fixed_c2p;

// The device address of `scratch` on the current device, looked up once per
// device (under g_mu) and kept.
asx_status fixed_c2p_scratch(C2PScratch *out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ASX_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(g_mu);
  unsigned char *&base = g_fixed_c2p[dev];
  if (!base) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      g_fixed_c2p.erase(dev);
      return ASX_ERR_DEVICE;
    }
    base = (unsigned char *)p;
  }
  out->cnt = (unsigned long long *)base;
  out->qcount = (unsigned *)(base + 1024);
  out->queue = (int *)(base + kFixedC2PQueueByte);
  return ASX_OK;
}

// Device scratch per (device, stream, tag), grown on demand.
enum { SCR_MLP_GUESS = 1, SCR_MLP_IMAGE = 2 };
struct DevScratch {
  void *p = nullptr;
  size_t bytes = 0;
};
std::map<std::tuple<int, void *, int>, DevScratch> g_dscratch;

void *dev_scratch(int tag, void *stream, size_t bytes) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_mu);
  DevScratch &S = g_dscratch[std::make_tuple(dev, stream, tag)];
  if (S.bytes < bytes) {
    // the previous buffer may still be read by work queued on this stream
    if (S.p && (hipStreamSynchronize((hipStream_t)stream) != hipSuccess ||
                hipFree(S.p) != hipSuccess))
      return nullptr;
    S.p = nullptr;
    S.bytes = 0;
    if (hipMalloc(&S.p, bytes) != hipSuccess) return nullptr;
    S.bytes = bytes;
  }
  return S.p;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// ---------------------------------------------------------------------------
// Debug poison (CarpetX::poison_undefined_values): outputs filled with
// ASX_POISON_BITS before an entry point runs, its declared region counted
// afterwards.
// ---------------------------------------------------------------------------
std::atomic<int> g_poison{0};
std::mutex g_pmu;
int32_t g_poison_slot = -1;
int64_t g_poison_points = 0;
constexpr int kPoisonBatch = 64;
std::map<int, unsigned long long *> g_pcnt;

unsigned long long *poison_buf(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_pcnt.find(dev);
  if (it != g_pcnt.end()) return it->second;
  unsigned long long *p = nullptr;
  if (hipMalloc(&p, kPoisonBatch * sizeof(unsigned long long)) != hipSuccess) return nullptr;
  g_pcnt[dev] = p;
  return p;
}

// The outputs of one call.  out(): a grid function the call only writes,
// filled on every point and checked on box w; scratch(): library memory
// filled only.
struct Poisoner {
  struct Out {
    double *g;
    int64_t n;
    int sj, sk, slot;
    Box w;
  };
  bool on = g_poison.load() != 0;
  std::vector<Out> outs;
  std::vector<std::pair<double *, int64_t>> extra;
  void out(const asx_patch *pt, int slot, const Box &w) {
    if (!on || !pt->gf[slot].data) return;
    const int c = centering(slot);
    int64_t sj, sk;
    strides(pt->n, pt->ng, c, sj, sk);
    outs.push_back({pt->gf[slot].data, npoints(pt->n, pt->ng, c), (int)sj, (int)sk, slot, w});
  }
  void out_range(const asx_patch *pt, int a, int b, const Box &w) {
    for (int q = a; q <= b; ++q) out(pt, q, w);
  }
  // array g laid out as slot `slot` of a patch of geometry pt
  void out_array(double *g, const asx_patch *pt, int slot, const Box &w) {
    if (!on || !g) return;
    const int c = centering(slot);
    int64_t sj, sk;
    strides(pt->n, pt->ng, c, sj, sk);
    outs.push_back({g, npoints(pt->n, pt->ng, c), (int)sj, (int)sk, slot, w});
  }
  void scratch(double *p, int64_t n) {
    if (on && p) extra.push_back({p, n});
  }
  asx_status fill(hipStream_t s) const {
    for (const Out &o : outs)
      if (launch_poison_fill(o.g, o.n, ASX_POISON_BITS, s) != hipSuccess) return ASX_ERR_DEVICE;
    for (const auto &e : extra)
      if (launch_poison_fill(e.first, e.second, ASX_POISON_BITS, s) != hipSuccess)
        return ASX_ERR_DEVICE;
    return ASX_OK;
  }
  // st: the call's own status; the check runs only after a successful call
  asx_status check(hipStream_t s, asx_status st) const {
    if (!on || st != ASX_OK || outs.empty()) return st;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ASX_ERR_DEVICE;
    unsigned long long *cnt = poison_buf(dev);
    if (!cnt) return ASX_ERR_DEVICE;
    for (size_t a = 0; a < outs.size(); a += kPoisonBatch) {
      const int m = (int)std::min<size_t>(kPoisonBatch, outs.size() - a);
      hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(kPoisonBatch), 0, s, cnt, m);
      for (int q = 0; q < m; ++q) {
        const Out &o = outs[a + q];
        if (launch_poison_count(o.g, o.sj, o.sk, o.w, ASX_POISON_BITS, cnt + q, s) != hipSuccess)
          return ASX_ERR_DEVICE;
      }
      unsigned long long h[kPoisonBatch];
      if (hipMemcpyAsync(h, cnt, m * sizeof(unsigned long long), hipMemcpyDeviceToHost, s) !=
              hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        return ASX_ERR_DEVICE;
      for (int q = 0; q < m; ++q)
        if (h[q]) {
          std::lock_guard<std::mutex> lk(g_pmu);
          g_poison_slot = outs[a + q].slot;
          g_poison_points = (int64_t)h[q];
          return ASX_ERR_POISON;
        }
    }
    return ASX_OK;
  }
};

// Scratch slots the fused substep computes: fluxes, upwind-CT face data,
// aux, and dB_stag (written only by asx_substep_bc with use_uct; the faces
// take their normal field from A).
bool scratch_slot(int s) {
  return (s >= ASX_FXDENS && s <= ASX_AMIN_ZFACE) || (s >= ASX_FX && s <= ASX_G) ||
         (s >= ASX_DBX_STAG && s <= ASX_DBZ_STAG);
}

Scratch *scratch_for(const asx_patch *pt, void *stream) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  const auto key = std::make_tuple(dev, stream, pt->n[0], pt->n[1], pt->n[2], pt->ng);
  auto it = g_scratch.find(key);
  if (it != g_scratch.end()) return &it->second;
  Scratch s;
  for (int q = 0; q < ASX_NUM_GF; ++q) {
    if (!scratch_slot(q)) continue;
    double *p = nullptr;
    const size_t bytes = (size_t)npoints(pt->n, pt->ng, centering(q)) * sizeof(double);
    if (hipMalloc(&p, bytes) != hipSuccess) {
      for (double *b : s.bufs) (void)hipFree(b);
      return nullptr;
    }
    s.bufs.push_back(p);
    s.slot[q] = p;
  }
  return &(g_scratch[key] = s);
}

} // namespace

static_assert(ASX_MLP_NPARAM == kMlpNParam && ASX_MLP_NBLOB == kMlpNBlob, "MLP blob");
static_assert(ASX_MLP_YE_NPARAM == kMlpYeNParam && ASX_MLP_YE_NBLOB == kMlpYeNBlob, "MLP blob");

extern "C" {

const char *asx_version(void) { return "asterx-mi355x 0.1 (gfx950, FP64)"; }

int asx_gf_centering(int slot) { return centering(slot); }

void asx_gf_strides(const int32_t n[3], int32_t ng, int slot, int64_t *sj, int64_t *sk) {
  strides(n, ng, centering(slot), *sj, *sk);
}

void asx_params_default(asx_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->reconstruction_method = ASX_RECON_GODUNOV;
  p->loworder_method = ASX_RECON_GODUNOV;
  p->reconstruct_with_temperature = 0;
  p->ppm_shock_detection = 1;
  p->ppm_zone_flattening = 0;
  p->ppm_eta1 = 20.;
  p->ppm_eta2 = 0.05;
  p->ppm_eps = 0.33;
  p->ppm_eps_shock = 0.01;
  p->ppm_small = 1.e-7;
  p->ppm_omega1 = 0.75;
  p->ppm_omega2 = 10.;
  p->enhanced_ppm_C2 = 1.25;
  p->weno_eps = 1e-26;
  p->mp5_alpha = 4.0;
  p->vector_potential_gauge = ASX_GAUGE_GENERALIZED_LORENZ;
  p->use_uct = 0;
  p->use_entropy_fix = 0;
  p->flux_type = ASX_FLUX_LXF;
  p->recon_type = ASX_RECVAR_V;
  p->local_spatial_order = 2;
  p->excise = 0;
  p->use_mask = 0;
  p->lorenz_damp_fac = 0.0;
  p->evolution_eos = ASX_EOS_IDEALGAS;
  p->poly_gamma = 2.0;
  p->poly_k = 100.0;
  p->gl_gamma = 2.0;
  p->rho_max = 1.0;
  p->rho_min = 1.0e-40;
  p->eps_max = 2.0;
  p->eps_min = 0.0;
  p->ye_max = 1.0;
  p->ye_min = 0.0;
  p->particle_mass = 1.0;
  p->c2p_prime = ASX_C2P_NOBLE;
  p->c2p_second = ASX_C2P_PALENZUELA;
  p->max_iter = 30;
  p->use_press_atmo = 0;
  p->thermal_eos_atmo = 1;
  p->Ye_lenient = 0;
  p->use_z = 0;
  p->use_temperature = 0;
  p->rho_abs_min = 1e-13;
  p->atmo_tol = 1e-3;
  p->eps_atmo = 1e-10;
  p->p_atmo = 1e-10;
  p->t_atmo = 1e-2;
  p->Ye_atmo = 0.5;
  p->B_lim = 100.0;
  p->vw_lim = 10.0;
  p->c2p_tol = 1e-10;
  p->r_atmo = 1.0e100;
  p->n_rho_atmo = 0.0;
  p->n_press_atmo = 0.0;
  p->n_temp_atmo = 0.0;
  p->alp_thresh = 0.0;
  p->rho_BH = 1.0e100;
  p->eps_BH = 1.0e100;
  p->vwlim_BH = 1.0e100;
  p->cons_error_limit = -1;
  p->tmunu_interp_order = 4;
  p->update_tmunu = 1;
  p->local_estimate_error = 1;
  p->regrid_method = ASX_REGRID_FIRST_GRAD;
  p->epsilon_err = 0.01;
}

asx_status asx_fluxes(const asx_patch *pt, const asx_params *par, int32_t face_halo,
                      asx_counters *cnt, void *stream) {
  const Range range("asx_fluxes");
  asx_status st = check_geom(pt);
  if (st) return st;
  if ((st = check_par(par))) return st;
  if ((st = check_flux_eos(par))) return st;
  if (face_halo < 0 || face_halo > 1) return ASX_ERR_PARAM;
  if (pt->ng < recon_ghosts(par->reconstruction_method) ||
      pt->ng < recon_ghosts(par->loworder_method) || (face_halo > 0 && pt->ng < 3))
    return ASX_ERR_GHOSTS;
  if ((st = need_range(pt, ASX_ALP, ASX_GZZ))) return st;
  if ((st = need(pt, {ASX_RHO, ASX_VELX, ASX_VELY, ASX_VELZ, ASX_PRESS, ASX_ENTROPY,
                      ASX_YE, ASX_BVECX, ASX_BVECY, ASX_BVECZ, ASX_DBX_STAG,
                      ASX_DBY_STAG, ASX_DBZ_STAG})))
    return st;
  if (par->reconstruct_with_temperature && (st = need(pt, {ASX_TEMPERATURE}))) return st;
  if (par->recon_type == ASX_RECVAR_Z && (st = need_range(pt, ASX_ZVEC_X, ASX_ZVEC_Z)))
    return st;
  if (par->recon_type == ASX_RECVAR_S && (st = need_range(pt, ASX_SVEC_X, ASX_SVEC_Z)))
    return st;
  if ((st = need_range(pt, ASX_FXDENS, ASX_FZBZ))) return st;
  const bool uct = pt->gf[ASX_AMAX_XFACE].data != nullptr;
  if (uct && (st = need_range(pt, ASX_VTILDE_Y_XFACE, ASX_AMIN_ZFACE))) return st;
  if (par->use_uct && !uct) return ASX_ERR_NULL;
  const bool aux = pt->gf[ASX_FX].data != nullptr;
  if (aux && ((st = need_range(pt, ASX_FX, ASX_G)) ||
              (st = need(pt, {ASX_AVEC_X, ASX_AVEC_Y, ASX_AVEC_Z, ASX_PSI}))))
    return st;
  hipStream_t s = (hipStream_t)stream;
  const KPatch P = kpatch(pt);
  unsigned long long *cd = nullptr;
  if (cnt) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    cd = counter_buf(dev);
    if (!cd) return ASX_ERR_DEVICE;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(128), 0, s, cd, 80);
  }
  Poisoner pz;
  for (int d = 0; d < 3; ++d) {
    int halo[3] = {face_halo, face_halo, face_halo};
    halo[d] = 0;
    const Box b = box(pt, 1 << d, halo);
    pz.out_range(pt, ASX_FXDENS + 10 * d, ASX_FXDENS + 10 * d + 9, b);
    if (uct) {
      pz.out_range(pt, ASX_VTILDE_Y_XFACE + 2 * d, ASX_VTILDE_Y_XFACE + 2 * d + 1, b);
      pz.out(pt, ASX_AMAX_XFACE + 2 * d, b);
      pz.out(pt, ASX_AMIN_XFACE + 2 * d, b);
    }
  }
  if (aux) {
    const int h = 2 * face_halo;
    const int ha[3] = {h, h, h};
    pz.out_range(pt, ASX_FX, ASX_G, box(pt, 7, ha));
  }
  if ((st = pz.fill(s))) return st;
  FluxArgs A;
  A.rp = rparams(par);
  A.eos = ideal_gas(par);
  const bool tab = tabulated(par);
  if (tab) A.tab = tab_eos(par);
  A.lo_method = par->loworder_method;
  A.flux_type = par->flux_type;
  A.rec_var = par->recon_type;
  A.with_T = par->reconstruct_with_temperature;
  A.write_uct = uct;
  A.bstag_from_A = 0; // AsterX_Fluxes reads dB_stag (schedule.ccl:173-194)
  A.cnt_dev = cd;
  A.with_aux = 0;
  for (int d = 0; d < 3; ++d) {
    int halo[3] = {face_halo, face_halo, face_halo};
    halo[d] = 0;
    const Box b = box(pt, 1 << d, halo);
    for (int q = 0; q < 3; ++q) {
      A.lo[q] = b.lo[q];
      A.cnt[q] = b.cnt[q];
    }
    if (d == 2 && aux) { // CalcAuxForAvecPsi rides on the z-face launch
      const int h = 2 * face_halo;
      const int ha[3] = {h, h, h};
      A.with_aux = 1;
      A.aux = box(pt, 7, ha);
    }
    if ((st = dev_status(timed(ASX_KC_FACE_X + d, s, [&] { return launch_flux(d, P, A, par->reconstruction_method, variant_of(ASX_KC_FACE_X + d), s, tab); })))) return st;
  }
  if (cnt) {
    unsigned long long hbuf[2];
    if (hipMemcpyAsync(hbuf, cd, sizeof(hbuf), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return ASX_ERR_DEVICE;
    cnt->nan_faces += (int64_t)hbuf[0];
    cnt->lo_faces += (int64_t)hbuf[1];
  }
  return pz.check(s, ASX_OK);
}

static RhsArgs rhs_args(const asx_params *par) {
  RhsArgs A;
  A.rp = rparams(par);
  A.recon = par->reconstruction_method;
  A.gauge = par->vector_potential_gauge;
  A.use_uct = par->use_uct;
  A.fd_order = par->local_spatial_order;
  A.rec_var = par->recon_type;
  A.lorenz_damp_fac = par->lorenz_damp_fac;
  return A;
}

asx_status asx_source_terms(const asx_patch *pt, const asx_params *par, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if (!par) return ASX_ERR_NULL;
  if (par->local_spatial_order != 2 && par->local_spatial_order != 4) return ASX_ERR_PARAM;
  if (par->local_spatial_order == 4 && pt->ng < 2) return ASX_ERR_GHOSTS;
  if ((st = need_range(pt, ASX_ALP, ASX_KZZ))) return st;
  if ((st = need(pt, {ASX_RHO, ASX_EPS, ASX_PRESS, ASX_BVECX, ASX_BVECY, ASX_BVECZ})))
    return st;
  if ((st = (par->recon_type == ASX_RECVAR_V) ? need(pt, {ASX_VELX, ASX_VELY, ASX_VELZ})
                                              : need_range(pt, ASX_ZVEC_X, ASX_ZVEC_Z)))
    return st;
  if ((st = need_range(pt, ASX_DENSRHS, ASX_DYE_RHS))) return st;
  const int halo[3] = {0, 0, 0};
  Poisoner pz;
  pz.out_range(pt, ASX_DENSRHS, ASX_DYE_RHS, box(pt, 0, halo));
  if ((st = pz.fill((hipStream_t)stream))) return st;
  st = dev_status(timed(ASX_KC_RHS, (hipStream_t)stream, [&] { return launch_sources(kpatch(pt), box(pt, 0, halo), rhs_args(par),
                                   (hipStream_t)stream); }));
  return pz.check((hipStream_t)stream, st);
}

asx_status asx_tmunu(const asx_patch *pt, const asx_params *par, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if (!par) return ASX_ERR_NULL;
  if (par->tmunu_interp_order != 2 && par->tmunu_interp_order != 4) return ASX_ERR_PARAM;
  if (!par->update_tmunu) return ASX_OK;
  if (pt->ng < par->tmunu_interp_order / 2) return ASX_ERR_GHOSTS;
  const int use_z = par->recon_type != ASX_RECVAR_V;
  if ((st = need_range(pt, ASX_ALP, ASX_GZZ))) return st;
  if ((st = need(pt, {ASX_RHO, ASX_EPS, ASX_PRESS, ASX_BVECX, ASX_BVECY, ASX_BVECZ})))
    return st;
  if ((st = use_z ? need_range(pt, ASX_ZVEC_X, ASX_ZVEC_Z)
                  : need(pt, {ASX_VELX, ASX_VELY, ASX_VELZ})))
    return st;
  if ((st = need_range(pt, ASX_ETTT, ASX_ETZZ))) return st;
  // the kernel addresses every grid function with 32-bit byte offsets
  if (npoints(pt->n, pt->ng, 7) * 8 >= (int64_t(1) << 32)) return ASX_ERR_SIZE;
  const int v = variant_of(ASX_KC_TMUNU);
  return dev_status(timed(ASX_KC_TMUNU, (hipStream_t)stream, [&] {
    return launch_tmunu(kpatch(pt), par->tmunu_interp_order, use_z, v, (hipStream_t)stream);
  }));
}

asx_status asx_estimate_error(const asx_patch *pt, const asx_params *par, int32_t nvars,
                              const int32_t *slots, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if (!par) return ASX_ERR_NULL;
  if (par->regrid_method < ASX_REGRID_FIRST_DERIV || par->regrid_method > ASX_REGRID_FIRST_GRAD)
    return ASX_ERR_PARAM;
  if (!par->local_estimate_error) return ASX_OK;
  if (nvars < 0 || nvars > kMaxRegridVars) return ASX_ERR_SIZE;
  if (nvars > 0 && !slots) return ASX_ERR_NULL;
  RegridArgs A{};
  A.nv = nvars;
  A.method = par->regrid_method;
  A.eps = par->epsilon_err;
  for (int q = 0; q < nvars; ++q) {
    if (slots[q] < 0 || slots[q] >= ASX_NUM_GF || centering(slots[q]) < 0) return ASX_ERR_PARAM;
    if ((st = need(pt, {slots[q]}))) return st;
    A.slot[q] = slots[q];
    A.cent[q] = centering(slots[q]);
  }
  if ((st = need(pt, {ASX_REGRID_ERROR}))) return st;
  const int halo[3] = {0, 0, 0};
  Poisoner pz;
  pz.out(pt, ASX_REGRID_ERROR, box(pt, 0, halo));
  if ((st = pz.fill((hipStream_t)stream))) return st;
  st = dev_status(timed(ASX_KC_REGRID, (hipStream_t)stream, [&] {
    return launch_regrid(kpatch(pt), box(pt, 0, halo), A, (hipStream_t)stream);
  }));
  return pz.check((hipStream_t)stream, st);
}

asx_status asx_rhs(const asx_patch *pt, const asx_params *par, void *stream) {
  const Range range("asx_rhs");
  asx_status st = check_geom(pt);
  if (st) return st;
  if ((st = check_par(par))) return st;
  if ((st = need_range(pt, ASX_FXDENS, ASX_FZBZ))) return st;
  if ((st = need_range(pt, ASX_DENSRHS, ASX_PSI_RHS))) return st;
  if ((st = need(pt, {ASX_ALP, ASX_BETAX, ASX_BETAY, ASX_BETAZ, ASX_PSI}))) return st;
  if (par->vector_potential_gauge == ASX_GAUGE_GENERALIZED_LORENZ &&
      (st = need_range(pt, ASX_FX, ASX_G)))
    return st;
  if (par->use_uct) {
    if ((st = need_range(pt, ASX_VTILDE_Y_XFACE, ASX_AMIN_ZFACE))) return st;
    if ((st = need(pt, {ASX_DBX_STAG, ASX_DBY_STAG, ASX_DBZ_STAG, ASX_PRESS, ASX_VELX,
                        ASX_VELY, ASX_VELZ})))
      return st;
    if (pt->ng < recon_ghosts(par->reconstruction_method)) return ASX_ERR_GHOSTS;
  }
  if (pt->ng < 2) return ASX_ERR_GHOSTS;
  // the hydro right-hand sides are added to, not written: not poisoned
  Poisoner pz;
  const int h0[3] = {0, 0, 0};
  for (int q = ASX_AVEC_X_RHS; q <= ASX_PSI_RHS; ++q) pz.out(pt, q, box(pt, centering(q), h0));
  if ((st = pz.fill((hipStream_t)stream))) return st;
  st = dev_status(timed(ASX_KC_RHS, (hipStream_t)stream, [&] { return launch_rhs(kpatch(pt), rhs_args(par), (hipStream_t)stream); }));
  return pz.check((hipStream_t)stream, st);
}

asx_status asx_bstag_from_A(const asx_patch *pt, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if ((st = need(pt, {ASX_AVEC_X, ASX_AVEC_Y, ASX_AVEC_Z, ASX_DBX_STAG, ASX_DBY_STAG,
                      ASX_DBZ_STAG})))
    return st;
  const KPatch P = kpatch(pt);
  const int halo[3] = {0, 0, 0};
  Poisoner pz;
  for (int d = 0; d < 3; ++d) pz.out(pt, ASX_DBX_STAG + d, box(pt, 1 << d, halo));
  if ((st = pz.fill((hipStream_t)stream))) return st;
  for (int d = 0; d < 3; ++d)
    if ((st = dev_status(timed(ASX_KC_BSTAG, (hipStream_t)stream, [&] { return launch_bstag(d, P, box(pt, 1 << d, halo), pt->gf[ASX_DBX_STAG + d].data,
                                      (hipStream_t)stream); }))))
      return st;
  return pz.check((hipStream_t)stream, ASX_OK);
}

asx_status asx_dB_from_bstag(const asx_patch *pt, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if ((st = need(pt, {ASX_DBX_STAG, ASX_DBY_STAG, ASX_DBZ_STAG, ASX_DBX, ASX_DBY, ASX_DBZ})))
    return st;
  const int halo[3] = {0, 0, 0};
  Poisoner pz;
  pz.out_range(pt, ASX_DBX, ASX_DBZ, box(pt, 0, halo));
  if ((st = pz.fill((hipStream_t)stream))) return st;
  st = dev_status(timed(ASX_KC_OTHER, (hipStream_t)stream, [&] { return launch_dB(kpatch(pt), box(pt, 0, halo), (hipStream_t)stream); }));
  return pz.check((hipStream_t)stream, st);
}

asx_status asx_B_from_dB(const asx_patch *pt, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if ((st = need_range(pt, ASX_GXX, ASX_GZZ))) return st;
  if ((st = need(pt, {ASX_DBX, ASX_DBY, ASX_DBZ, ASX_BVECX, ASX_BVECY, ASX_BVECZ}))) return st;
  const int halo[3] = {0, 0, 0};
  Poisoner pz;
  pz.out_range(pt, ASX_BVECX, ASX_BVECZ, box(pt, 0, halo));
  if ((st = pz.fill((hipStream_t)stream))) return st;
  st = dev_status(timed(ASX_KC_OTHER, (hipStream_t)stream, [&] { return launch_B_from_dB(kpatch(pt), box(pt, 0, halo), (hipStream_t)stream); }));
  return pz.check((hipStream_t)stream, st);
}

asx_status asx_prim2con(const asx_patch *pt, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if ((st = need_range(pt, ASX_GXX, ASX_GZZ))) return st;
  if ((st = need(pt, {ASX_RHO, ASX_VELX, ASX_VELY, ASX_VELZ, ASX_EPS, ASX_PRESS,
                      ASX_ENTROPY, ASX_YE, ASX_BVECX, ASX_BVECY, ASX_BVECZ})))
    return st;
  if ((st = need_range(pt, ASX_DENS, ASX_DBZ))) return st;
  const int halo[3] = {0, 0, 0};
  Poisoner pz;
  pz.out_range(pt, ASX_DENS, ASX_DBZ, box(pt, 0, halo));
  if ((st = pz.fill((hipStream_t)stream))) return st;
  st = dev_status(timed(ASX_KC_OTHER, (hipStream_t)stream, [&] { return launch_prim2con(kpatch(pt), box(pt, 0, halo), (hipStream_t)stream); }));
  return pz.check((hipStream_t)stream, st);
}

static asx_status c2p_common(const asx_patch *pt, const asx_params *par, bool from_A,
                             asx_counters *cnt, void *stream,
                             const C2PScratch *fixed = nullptr) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if ((st = check_c2p_par(par))) return st;
  if ((st = need(pt, {ASX_ALP, ASX_GXX, ASX_GXY, ASX_GXZ, ASX_GYY, ASX_GYZ, ASX_GZZ})))
    return st;
  if ((st = need_range(pt, ASX_RHO, ASX_DBZ))) return st;
  if ((st = need_range(pt, ASX_SAVED_RHO, ASX_CON2PRIM_FLAG))) return st;
  if ((st = opt(pt, ASX_MASK_CC))) return st;
  if (par->use_mask && (st = need(pt, {ASX_MASK_VC}))) return st;
  if (from_A && (st = need(pt, {ASX_AVEC_X, ASX_AVEC_Y, ASX_AVEC_Z}))) return st;
  hipStream_t s = (hipStream_t)stream;
  const KPatch P = kpatch(pt);
  // With caller-fixed buffers the launches are not timed: the timer's event
  // pool grows by hipEventCreate.
  auto run = [&](int cls, auto &&f) { return fixed ? f() : timed(cls, s, f); };
  if (from_A && pt->gf[ASX_DBX_STAG].data) {
    if ((st = need(pt, {ASX_DBY_STAG, ASX_DBZ_STAG}))) return st;
    const int halo[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d)
      if ((st = dev_status(run(ASX_KC_BSTAG, [&] { return launch_bstag(d, P, box(pt, 1 << d, halo),
                                        pt->gf[ASX_DBX_STAG + d].data, s); }))))
        return st;
  }
  unsigned long long *cd = nullptr;
  if (cnt) {
    if (fixed) {
      cd = fixed->cnt;
    } else {
      int dev = 0;
      (void)hipGetDevice(&dev);
      cd = counter_buf(dev);
      if (!cd) return ASX_ERR_DEVICE;
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(128), 0, s, cd, 80);
  }
  int *queue;
  unsigned *qcount;
  if (fixed) {
    queue = fixed->queue;
    qcount = fixed->qcount;
  } else {
    C2PQueue *Q = c2p_queue(stream, (int64_t)pt->n[0] * pt->n[1] * pt->n[2]);
    if (!Q) return ASX_ERR_DEVICE;
    queue = Q->q;
    qcount = Q->n;
  }
  const C2PArgs ca = c2p_args(par);
  unsigned long long *hist = cd ? cd + 2 : nullptr;
  const bool tab = tabulated(par);
  if ((st = dev_status(run(ASX_KC_C2P, [&] {
         return tab ? launch_con2prim_main(P, tab_eos(par), ca, from_A, hist, queue, qcount, s)
                    : launch_con2prim_main(P, ideal_gas(par), ca, from_A, hist, queue, qcount, s);
       }))))
    return st;
  if ((st = dev_status(run(ASX_KC_C2P_REST, [&] {
         return tab ? launch_con2prim_rest(P, tab_eos(par), ca, from_A, hist, queue, qcount, s)
                    : launch_con2prim_rest(P, ideal_gas(par), ca, from_A, hist, queue, qcount, s);
       }))))
    return st;
  if (cnt) {
    unsigned long long h[71];
    if (hipMemcpyAsync(h, cd + 2, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return ASX_ERR_DEVICE;
    for (int q = 0; q < 7; ++q) cnt->c2p_flags[q] += (int64_t)h[q];
    for (int q = 0; q < 64; ++q) cnt->c2p_iters[q] += (int64_t)h[7 + q];
  }
  return ASX_OK;
}

asx_status asx_con2prim(const asx_patch *pt, const asx_params *par, asx_counters *cnt,
                        void *stream) {
  const Range range("asx_con2prim");
  return c2p_common(pt, par, false, cnt, stream);
}

asx_status asx_con2prim_group(const asx_patch *pt, const asx_params *par,
                              asx_counters *cnt, void *stream) {
  const Range range("asx_con2prim_group");
  return c2p_common(pt, par, true, cnt, stream);
}

asx_status asx_con2prim_group_noalloc(const asx_patch *pt, const asx_params *par,
                                      asx_counters *cnt, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if ((int64_t)pt->n[0] * pt->n[1] * pt->n[2] > kFixedC2PZones) return ASX_ERR_SIZE;
  C2PScratch f;
  if ((st = fixed_c2p_scratch(&f))) return st;
  return c2p_common(pt, par, true, cnt, stream, &f);
}

asx_status asx_state_lincomb(const asx_patch *pt, const asx_state *dst, double scale,
                             int32_t nsrc, const double *fac, const asx_state *const *src,
                             void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if (!dst || nsrc < 0 || nsrc > 8 || (nsrc > 0 && (!fac || !src))) return ASX_ERR_PARAM;
  for (int q = 0; q < ASX_NUM_STATE; ++q) {
    const int c = centering(kStateSlots[q]);
    int64_t sj, sk;
    strides(pt->n, pt->ng, c, sj, sk);
    if (!dst->gf[q].data) return ASX_ERR_NULL;
    if (dst->gf[q].sj != sj || dst->gf[q].sk != sk) return ASX_ERR_LAYOUT;
    double *srcp[8];
    for (int k = 0; k < nsrc; ++k) {
      if (!src[k] || !src[k]->gf[q].data) return ASX_ERR_NULL;
      if (src[k]->gf[q].sj != sj || src[k]->gf[q].sk != sk) return ASX_ERR_LAYOUT;
      srcp[k] = src[k]->gf[q].data;
    }
    if ((st = dev_status(timed(ASX_KC_OTHER, (hipStream_t)stream, [&] { return launch_lincomb(dst->gf[q].data, npoints(pt->n, pt->ng, c), scale,
                                        nsrc, fac, srcp, (hipStream_t)stream); }))))
      return st;
  }
  return ASX_OK;
}

asx_status asx_bc_fill(const asx_patch *pt, int32_t slot, int32_t kind, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if (slot < 0 || slot >= ASX_NUM_GF) return ASX_ERR_PARAM;
  if (kind < ASX_BC_LINEAR || kind > ASX_BC_COPY) return ASX_ERR_PARAM;
  if ((st = need(pt, {slot}))) return st;
  for (int d = 0; d < 3; ++d) {
    const int c = (centering(slot) >> d) & 1;
    if (kind == ASX_BC_LINEAR && pt->n[d] + c < 2) return ASX_ERR_SIZE;
    if (kind == ASX_BC_PERIODIC && pt->n[d] < pt->ng) return ASX_ERR_SIZE;
  }
  return dev_status(timed(ASX_KC_BC, (hipStream_t)stream, [&] { return launch_bc(kpatch(pt), slot, kind, (hipStream_t)stream); }));
}

asx_status asx_timing_enable(int32_t on) {
  std::lock_guard<std::mutex> l(g_tmu);
  asx_status st = drain_timing();
  for (int q = 0; q < ASX_KC_COUNT; ++q) {
    g_tms[q] = 0;
    g_tn[q] = 0;
  }
  g_timing.store(on != 0);
  return st;
}

asx_status asx_timing_read(int32_t kclass, int64_t *launches, double *ms) {
  if (kclass < 0 || kclass >= ASX_KC_COUNT) return ASX_ERR_PARAM;
  if (!launches || !ms) return ASX_ERR_NULL;
  std::lock_guard<std::mutex> l(g_tmu);
  asx_status st = drain_timing();
  *launches = g_tn[kclass];
  *ms = g_tms[kclass];
  return st;
}

asx_status asx_release_scratch(void *stream) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ASX_ERR_DEVICE;
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return ASX_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(g_mu);
  bool ok = true;
  for (auto it = g_scratch.begin(); it != g_scratch.end();) {
    if (std::get<0>(it->first) == dev && std::get<1>(it->first) == stream) {
      for (double *b : it->second.bufs) ok &= hipFree(b) == hipSuccess;
      it = g_scratch.erase(it);
    } else {
      ++it;
    }
  }
  auto q = g_c2pq.find({dev, stream});
  if (q != g_c2pq.end()) {
    if (q->second.q) ok &= hipFree(q->second.q) == hipSuccess;
    if (q->second.n) ok &= hipFree(q->second.n) == hipSuccess;
    g_c2pq.erase(q);
  }
  return ok ? ASX_OK : ASX_ERR_DEVICE;
}

asx_status asx_set_poison(int32_t on) {
  g_poison.store(on ? 1 : 0);
  std::lock_guard<std::mutex> lk(g_pmu);
  g_poison_slot = -1;
  g_poison_points = 0;
  return ASX_OK;
}

asx_status asx_poison_report(int32_t *slot, int64_t *points) {
  if (!slot || !points) return ASX_ERR_NULL;
  std::lock_guard<std::mutex> lk(g_pmu);
  *slot = g_poison_slot;
  *points = g_poison_points;
  return ASX_OK;
}

asx_status asx_poison_check(const asx_patch *pt, int32_t slot, const int32_t lo[3],
                            const int32_t cnt[3], int64_t *points, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if (!lo || !cnt || !points) return ASX_ERR_NULL;
  if (slot < 0 || slot >= ASX_NUM_GF || centering(slot) < 0) return ASX_ERR_PARAM;
  if ((st = need(pt, {slot}))) return st;
  const int c = centering(slot);
  Box B;
  for (int d = 0; d < 3; ++d) {
    const int E = pt->n[d] + ((c >> d) & 1) + 2 * pt->ng;
    if (lo[d] < 0 || cnt[d] < 0 || lo[d] + cnt[d] > E) return ASX_ERR_SIZE;
    B.lo[d] = lo[d];
    B.cnt[d] = cnt[d];
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ASX_ERR_DEVICE;
  unsigned long long *cd = poison_buf(dev);
  if (!cd) return ASX_ERR_DEVICE;
  hipStream_t s = (hipStream_t)stream;
  int64_t sj, sk;
  strides(pt->n, pt->ng, c, sj, sk);
  unsigned long long h = 0;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, cd, 1);
  if (launch_poison_count(pt->gf[slot].data, (int)sj, (int)sk, B, ASX_POISON_BITS, cd, s) !=
          hipSuccess ||
      hipMemcpyAsync(&h, cd, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ASX_ERR_DEVICE;
  *points = (int64_t)h;
  return ASX_OK;
}

asx_status asx_bc_fill_many(const asx_patch *pt, int32_t nslot, const int32_t *slots,
                            int32_t kind, void *stream) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if (nslot < 0 || nslot > kMaxBC) return ASX_ERR_PARAM;
  if (nslot > 0 && !slots) return ASX_ERR_NULL;
  if (kind < ASX_BC_LINEAR || kind > ASX_BC_COPY) return ASX_ERR_PARAM;
  BCList L;
  L.nf = nslot;
  for (int q = 0; q < nslot; ++q) {
    const int slot = slots[q];
    if (slot < 0 || slot >= ASX_NUM_GF) return ASX_ERR_PARAM;
    if ((st = need(pt, {slot}))) return st;
    const int c = centering(slot);
    for (int d = 0; d < 3; ++d) {
      const int cd = (c >> d) & 1;
      if (kind == ASX_BC_LINEAR && pt->n[d] + cd < 2) return ASX_ERR_SIZE;
      if (kind == ASX_BC_PERIODIC && pt->n[d] < pt->ng) return ASX_ERR_SIZE;
    }
    L.g[q] = pt->gf[slot].data;
    L.cent[q] = c;
  }
  return dev_status(timed(ASX_KC_BC, (hipStream_t)stream, [&] { return launch_bc_many(kpatch(pt), L, kind, (hipStream_t)stream); }));
}

}  // extern "C"

// Fine patch f one level finer than c (refinement 2): off[d] = fine index of
// f's first interior point counted from twice c's first interior point.
static asx_status amr_geometry(const asx_patch *c, const asx_patch *f, int off[3]) {
  for (int d = 0; d < 3; ++d) {
    if (std::fabs(2.0 * f->dx[d] - c->dx[d]) > 1e-12 * c->dx[d]) return ASX_ERR_LAYOUT;
    const double r = (f->x0[d] - c->x0[d]) / f->dx[d];
    const double ri = std::nearbyint(r);
    if (!(std::fabs(r - ri) <= 1e-6) || std::fabs(ri) > 1e9) return ASX_ERR_LAYOUT;
    off[d] = (int)ri;
    if ((off[d] & 1) || (f->n[d] & 1)) return ASX_ERR_LAYOUT;
  }
  return ASX_OK;
}

static asx_status amr_slots(const asx_patch *a, const asx_patch *b, int32_t nslot,
                            const int32_t *slots) {
  if (nslot < 0 || nslot > kMaxAmrSlots) return ASX_ERR_PARAM;
  if (nslot > 0 && !slots) return ASX_ERR_NULL;
  for (int q = 0; q < nslot; ++q) {
    if (slots[q] < 0 || slots[q] >= ASX_NUM_GF || centering(slots[q]) < 0) return ASX_ERR_PARAM;
    asx_status st = need(a, {slots[q]});
    if (st || (st = need(b, {slots[q]}))) return st;
  }
  return ASX_OK;
}

static void amr_strides(const asx_patch *pt, int sj[8], int sk[8]) {
  for (int c = 0; c < 8; ++c) {
    int64_t a, b;
    strides(pt->n, pt->ng, c, a, b);
    sj[c] = (int)a;
    sk[c] = (int)b;
  }
}

extern "C" {

asx_status asx_prolongate(const asx_patch *coarse, const asx_patch *fine, int32_t nslot,
                          const int32_t *slots, int32_t order, int32_t region, void *stream) {
  const Range range("asx_prolongate");
  asx_status st = check_geom(coarse);
  if (st || (st = check_geom(fine))) return st;
  if (region < 0 || region > 1) return ASX_ERR_PARAM;
  if (order != 1 && order != 3 && order != 5) return ASX_ERR_UNSUPPORTED;
  if ((st = amr_slots(coarse, fine, nslot, slots))) return st;
  ProlongArgs A{};
  if ((st = amr_geometry(coarse, fine, A.off))) return st;
  A.nslot = nslot;
  A.order = order;
  A.region = region;
  A.fng = fine->ng;
  A.cng = coarse->ng;
  for (int d = 0; d < 3; ++d) A.fn[d] = fine->n[d];
  amr_strides(fine, A.fsj, A.fsk);
  amr_strides(coarse, A.csj, A.csk);
  int64_t most = 0;
  for (int q = 0; q < nslot; ++q) {
    const int c = centering(slots[q]);
    A.src[q] = coarse->gf[slots[q]].data;
    A.dst[q] = fine->gf[slots[q]].data;
    A.cent[q] = c;
    // every coarse point a stencil reads must exist (every fine index along
    // each axis is reached, by the ghost slabs of the other axes too)
    int E3[3];
    for (int d = 0; d < 3; ++d) {
      const int vc = (c >> d) & 1;
      const int E = fine->n[d] + vc + 2 * fine->ng, Ec = coarse->n[d] + vc + 2 * coarse->ng;
      for (int i = 0; i < E; ++i) {
        int s0, n;
        axis_span(A.off[d] + i - fine->ng, vc, order, s0, n);
        if (s0 + coarse->ng < 0 || s0 + coarse->ng + n > Ec) return ASX_ERR_GHOSTS;
      }
      E3[d] = E;
    }
    most = std::max(most, BlockShell(A.off, E3, fine->ng).total(region == 1));
  }
  return dev_status(timed(ASX_KC_AMR, (hipStream_t)stream, [&] {
    return launch_prolong(A, most, (hipStream_t)stream);
  }));
}

asx_status asx_restrict(const asx_patch *fine, const asx_patch *coarse, int32_t nslot,
                        const int32_t *slots, void *stream) {
  const Range range("asx_restrict");
  asx_status st = check_geom(coarse);
  if (st || (st = check_geom(fine))) return st;
  if ((st = amr_slots(coarse, fine, nslot, slots))) return st;
  RestrictArgs A{};
  if ((st = amr_geometry(coarse, fine, A.off))) return st;
  A.nslot = nslot;
  A.fng = fine->ng;
  A.cng = coarse->ng;
  amr_strides(fine, A.fsj, A.fsk);
  amr_strides(coarse, A.csj, A.csk);
  int64_t most = 0;
  for (int q = 0; q < nslot; ++q) {
    const int c = centering(slots[q]);
    A.src[q] = fine->gf[slots[q]].data;
    A.dst[q] = coarse->gf[slots[q]].data;
    A.cent[q] = c;
    // coarse points (counted from the coarse interior origin) whose fine
    // points are the fine patch's own: vertex axis 2I - off in [0, n_f],
    // cell axis both children in [0, n_f - 1]; then inside the coarse interior
    Box &B = A.box[q];
    for (int d = 0; d < 3; ++d) {
      const int vc = (c >> d) & 1;
      const int lo = std::max(A.off[d] / 2, 0);
      const int hi = std::min(vc ? (A.off[d] + fine->n[d]) / 2 : A.off[d] / 2 + fine->n[d] / 2 - 1,
                              coarse->n[d] + vc - 1);
      B.lo[d] = lo + coarse->ng;
      B.cnt[d] = std::max(0, hi - lo + 1);
    }
    most = std::max(most, B.htotal());
  }
  return dev_status(timed(ASX_KC_AMR, (hipStream_t)stream, [&] {
    return launch_restrict(A, most, (hipStream_t)stream);
  }));
}

asx_status asx_regrid_boxes(const asx_patch *pt, double threshold, int32_t n_error_buf,
                            const int32_t blocking_factor[3], const int32_t max_grid_size[3],
                            double grid_eff, int32_t max_boxes, asx_box *boxes,
                            int32_t *nboxes, void *stream) {
  const Range range("asx_regrid_boxes");
  asx_status st = check_geom(pt);
  if (st) return st;
  if (!blocking_factor || !max_grid_size || !nboxes || (max_boxes > 0 && !boxes))
    return ASX_ERR_NULL;
  if (n_error_buf < 0 || !(grid_eff > 0.0 && grid_eff <= 1.0) || max_boxes < 0)
    return ASX_ERR_PARAM;
  if ((st = need(pt, {ASX_REGRID_ERROR}))) return st;
  TagArgs A{};
  int unit[3], per[3];
  for (int d = 0; d < 3; ++d) {
    const int bf = blocking_factor[d];
    if (bf < 1 || (bf & (bf - 1)) || max_grid_size[d] % bf) return ASX_ERR_PARAM;
    A.blk[d] = std::max(1, bf / 2);
    if (pt->n[d] % A.blk[d]) return ASX_ERR_PARAM;
    A.nb[d] = pt->n[d] / A.blk[d];
    A.n[d] = pt->n[d];
    unit[d] = 2 * A.blk[d];             // fine cells per block
    per[d] = max_grid_size[d] / unit[d]; // blocks per piece at most
    if (per[d] < 1) return ASX_ERR_PARAM;
  }
  int64_t sj, sk;
  strides(pt->n, pt->ng, 0, sj, sk);
  A.err = pt->gf[ASX_REGRID_ERROR].data;
  A.sj = (int)sj;
  A.sk = (int)sk;
  A.ng = pt->ng;
  A.buf = n_error_buf;
  A.thr = threshold;
  const int64_t nb = (int64_t)A.nb[0] * A.nb[1] * A.nb[2];
  hipStream_t s = (hipStream_t)stream;
  // the block tags live on the caller's stream only (stream-ordered
  // allocation: no device-wide synchronisation)
  uint8_t *dtags = nullptr;
  if (hipMallocAsync((void **)&dtags, nb, s) != hipSuccess) return ASX_ERR_DEVICE;
  A.tags = dtags;
  std::vector<uint8_t> tags(nb);
  hipError_t e = timed(ASX_KC_AMR, s, [&] { return launch_regrid_tags(A, s); });
  if (e == hipSuccess) e = hipMemcpyAsync(tags.data(), dtags, nb, hipMemcpyDeviceToHost, s);
  const hipError_t ef = hipFreeAsync(dtags, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess || ef != hipSuccess) return ASX_ERR_DEVICE;
  // clusters in blocks, each cut into pieces of at most per[d] blocks
  // (equal to within one block, the larger ones first), in fine cells
  int64_t found = 0;
  for (const ClusterBox &b : cluster_tags(tags.data(), A.nb, grid_eff)) {
    int np[3], len[3];
    for (int d = 0; d < 3; ++d) {
      len[d] = b.hi[d] - b.lo[d] + 1;
      np[d] = (len[d] + per[d] - 1) / per[d];
    }
    for (int pz = 0; pz < np[2]; ++pz)
      for (int py = 0; py < np[1]; ++py)
        for (int px = 0; px < np[0]; ++px) {
          const int p[3] = {px, py, pz};
          if (found < max_boxes) {
            asx_box &o = boxes[found];
            for (int d = 0; d < 3; ++d) {
              const int base = len[d] / np[d], extra = len[d] % np[d];
              const int a = b.lo[d] + p[d] * base + std::min(p[d], extra);
              const int w = base + (p[d] < extra ? 1 : 0);
              o.lo[d] = a * unit[d];
              o.hi[d] = (a + w) * unit[d] - 1;
            }
          }
          ++found;
        }
  }
  *nboxes = (int32_t)std::min<int64_t>(found, INT32_MAX);
  return found > max_boxes ? ASX_ERR_SIZE : ASX_OK;
}

// The fused substep.  bc_kind < 0: the face data of the one-cell transverse
// halo are computed from the ghost cells (what a neighbour patch's SYNC
// delivers there).  bc_kind >= 0: face data, aux and (with use_uct) the
// upwind-CT face data and dB_stag on the interior only, their ghost points
// filled by that boundary kind -- AsterX_Fluxes followed by the SYNC of
// AsterX/schedule.ccl:191-193 on a single patch.
static asx_status substep_impl(const asx_patch *stage, const asx_state *u0,
                               const asx_state *out, double c0, double c1, double c2,
                               const asx_params *par, int bc_kind, void *stream) {
  asx_status st = check_geom(stage);
  if (st) return st;
  if ((st = check_par(par))) return st;
  if ((st = check_flux_eos(par))) return st;
  if (!out || (c0 != 0.0 && !u0)) return ASX_ERR_NULL;
  const bool bc = bc_kind >= 0;
  if (bc && bc_kind > ASX_BC_COPY) return ASX_ERR_PARAM;
  // Upwind CT reconstructs face data three rows into the transverse ghost
  // region, which the halo computation cannot reach with 3 ghost cells; it
  // needs those faces from a SYNC (bc_kind >= 0).
  if (par->use_uct && !bc) return ASX_ERR_UNSUPPORTED;
  // the halo computation reconstructs one cell into the ghosts (3 more
  // cells of stencil); with SYNCed faces the stencils and the aux of the
  // Psi equation (+-2 vertices) need 2
  if (stage->ng < (bc ? 2 : 3) || stage->ng < recon_ghosts(par->reconstruction_method) ||
      stage->ng < recon_ghosts(par->loworder_method))
    return ASX_ERR_GHOSTS;
  if (bc)
    for (int d = 0; d < 3; ++d) {
      if (bc_kind == ASX_BC_LINEAR && stage->n[d] < 2) return ASX_ERR_SIZE;
      if (bc_kind == ASX_BC_PERIODIC && stage->n[d] < stage->ng) return ASX_ERR_SIZE;
    }
  if ((st = need_range(stage, ASX_ALP, ASX_KZZ))) return st;
  if ((st = need(stage, {ASX_RHO, ASX_VELX, ASX_VELY, ASX_VELZ, ASX_EPS, ASX_PRESS,
                         ASX_ENTROPY, ASX_YE, ASX_BVECX, ASX_BVECY, ASX_BVECZ})))
    return st;
  if (par->reconstruct_with_temperature && (st = need(stage, {ASX_TEMPERATURE}))) return st;
  if (par->recon_type == ASX_RECVAR_Z && (st = need_range(stage, ASX_ZVEC_X, ASX_ZVEC_Z)))
    return st;
  if (par->recon_type == ASX_RECVAR_S && (st = need_range(stage, ASX_SVEC_X, ASX_SVEC_Z)))
    return st;
  for (int q = 0; q < ASX_NUM_STATE; ++q) {
    if ((st = need(stage, {kStateSlots[q]}))) return st;
    const int c = centering(kStateSlots[q]);
    int64_t sj, sk;
    strides(stage->n, stage->ng, c, sj, sk);
    if (!out->gf[q].data) return ASX_ERR_NULL;
    if (out->gf[q].sj != sj || out->gf[q].sk != sk) return ASX_ERR_LAYOUT;
    if (c0 != 0.0) {
      if (!u0->gf[q].data) return ASX_ERR_NULL;
      if (u0->gf[q].sj != sj || u0->gf[q].sk != sk) return ASX_ERR_LAYOUT;
    }
  }
  hipStream_t s = (hipStream_t)stream;
  Scratch *sc = scratch_for(stage, stream);
  if (!sc) return ASX_ERR_DEVICE;
  Poisoner pz;
  if (pz.on) {
    // the out state (unless it shares an array with an input) and the face scratch
    bool alias = false;
    for (int q = 0; q < ASX_NUM_STATE; ++q) {
      for (int r = 0; r < ASX_NUM_GF; ++r) alias |= out->gf[q].data == stage->gf[r].data;
      if (c0 != 0.0)
        for (int r = 0; r < ASX_NUM_STATE; ++r) alias |= out->gf[q].data == u0->gf[r].data;
    }
    const int h0[3] = {0, 0, 0};
    if (!alias)
      for (int q = 0; q < ASX_NUM_STATE; ++q)
        pz.out_array(out->gf[q].data, stage, kStateSlots[q],
                     box(stage, centering(kStateSlots[q]), h0));
    for (int q = 0; q < ASX_NUM_GF; ++q)
      if (sc->slot[q]) pz.scratch(sc->slot[q], npoints(stage->n, stage->ng, centering(q)));
    if ((st = pz.fill(s))) return st;
  }
  KPatch P = kpatch(stage);
  for (int q = 0; q < ASX_NUM_GF; ++q)
    if (sc->slot[q]) P.g[q] = sc->slot[q];
  FluxArgs A;
  A.rp = rparams(par);
  A.eos = ideal_gas(par);
  const bool tab = tabulated(par);
  if (tab) A.tab = tab_eos(par);
  A.lo_method = par->loworder_method;
  A.flux_type = par->flux_type;
  A.rec_var = par->recon_type;
  A.with_T = par->reconstruct_with_temperature;
  A.write_uct = par->use_uct;
  // the faces take their normal field as the curl of the stage's A (what
  // ComputedBstagFromA would store on the interior faces)
  A.bstag_from_A = 1;
  A.cnt_dev = nullptr;
  A.with_aux = 0;
  const int fh = bc ? 0 : 1;
  for (int d = 0; d < 3; ++d) {
    int halo[3] = {fh, fh, fh};
    halo[d] = 0;
    const Box b = box(stage, 1 << d, halo);
    for (int q = 0; q < 3; ++q) {
      A.lo[q] = b.lo[q];
      A.cnt[q] = b.cnt[q];
    }
    if (d == 2) { // the vertex aux ride on the z-face launch
      const int h = 2 * fh;
      const int ha[3] = {h, h, h};
      A.with_aux = 1;
      A.aux = box(stage, 7, ha);
    }
    if ((st = dev_status(timed(ASX_KC_FACE_X + d, s, [&] { return launch_flux(d, P, A, par->reconstruction_method, variant_of(ASX_KC_FACE_X + d), s, tab); })))) return st;
  }
  if (bc) {
    BCList L;
    L.nf = 0;
    auto add = [&](int a, int b) {
      for (int q = a; q <= b; ++q) {
        L.g[L.nf] = P.g[q];
        L.cent[L.nf] = centering(q);
        ++L.nf;
      }
    };
    add(ASX_FXDENS, ASX_FZBZ);
    add(ASX_FX, ASX_G);
    if (par->use_uct) {
      // dB_stag on the interior faces (AsterX_ComputedBstagFromA), SYNCed
      // with the face data the upwind CT reconstructs
      const int halo[3] = {0, 0, 0};
      for (int d = 0; d < 3; ++d)
        if ((st = dev_status(timed(ASX_KC_BSTAG, s, [&] { return launch_bstag(d, P, box(stage, 1 << d, halo), P.g[ASX_DBX_STAG + d], s); }))))
          return st;
      add(ASX_VTILDE_Y_XFACE, ASX_AMIN_ZFACE);
      add(ASX_DBX_STAG, ASX_DBZ_STAG);
    }
    if ((st = dev_status(timed(ASX_KC_BC, s, [&] { return launch_bc_many(P, L, bc_kind, s); }))))
      return st;
  }
  StatePtrs pu0, pout;
  for (int q = 0; q < ASX_NUM_STATE; ++q) {
    pu0.g[q] = (c0 != 0.0) ? u0->gf[q].data : nullptr;
    pout.g[q] = out->gf[q].data;
  }
  st = dev_status(timed(ASX_KC_UPDATE, s, [&] { return launch_update(P, rhs_args(par), pu0, pout, c0, c1, c2, variant_of(ASX_KC_UPDATE), s); }));
  return pz.check(s, st);
}

asx_status asx_substep(const asx_patch *stage, const asx_state *u0, const asx_state *out,
                       double c0, double c1, double c2, const asx_params *par,
                       void *stream) {
  const Range range("asx_substep");
  return substep_impl(stage, u0, out, c0, c1, c2, par, -1, stream);
}

asx_status asx_substep_bc(const asx_patch *stage, const asx_state *u0, const asx_state *out,
                          double c0, double c1, double c2, const asx_params *par,
                          int32_t bc_kind, void *stream) {
  if (bc_kind < 0) return ASX_ERR_PARAM;
  const Range range("asx_substep_bc");
  return substep_impl(stage, u0, out, c0, c1, c2, par, bc_kind, stream);
}

asx_status asx_c2p_batch(int64_t nz, const double *cons, const double *g,
                         const double *seed, double *prim, double *flag, int32_t *iters,
                         const asx_params *par, void *stream) {
  asx_status st = check_c2p_par(par);
  if (st) return st;
  if (tabulated(par)) return ASX_ERR_UNSUPPORTED; // no Ye in this layout: asx_c2p_batch_ye
  if (nz < 0) return ASX_ERR_SIZE;
  if (nz > 0 && (!cons || !g || !seed || !prim || !flag || !iters)) return ASX_ERR_NULL;
  return dev_status(timed(ASX_KC_C2P, (hipStream_t)stream, [&] { return launch_c2p_batch(nz, cons, g, seed, prim, flag, iters, ideal_gas(par),
                                     c2p_args(par), (hipStream_t)stream); }));
}

int32_t asx_variant_count(int32_t kclass) {
  if (kclass < 0 || kclass >= ASX_KC_COUNT) return 0;
  return variant_count(kclass);
}

asx_status asx_set_variant(int32_t kclass, int32_t v) {
  if (kclass < 0 || kclass >= ASX_KC_COUNT) return ASX_ERR_PARAM;
  if (v < 0 || v >= variant_count(kclass)) return ASX_ERR_PARAM;
  set_variant(kclass, v);
  return ASX_OK;
}

asx_status asx_get_variant(int32_t kclass, int32_t *v) {
  if (kclass < 0 || kclass >= ASX_KC_COUNT) return ASX_ERR_PARAM;
  if (!v) return ASX_ERR_NULL;
  *v = variant_of(kclass);
  return ASX_OK;
}

asx_status asx_autotune_substep(const asx_patch *stage, const asx_state *u0,
                                const asx_state *out, double c0, double c1, double c2,
                                const asx_params *par, int32_t reps, void *stream) {
  if (reps < 1) return ASX_ERR_PARAM;
  const int tuned[4] = {ASX_KC_FACE_X, ASX_KC_FACE_Y, ASX_KC_FACE_Z, ASX_KC_UPDATE};
  int nv = 1;
  for (int kc : tuned) nv = variant_count(kc) > nv ? variant_count(kc) : nv;
  // keep the caller's timing state: drain it, measure, put it back
  bool was_on;
  double tms[ASX_KC_COUNT];
  int64_t tn[ASX_KC_COUNT];
  {
    std::lock_guard<std::mutex> l(g_tmu);
    asx_status st = drain_timing();
    if (st) return st;
    was_on = g_timing.load();
    for (int q = 0; q < ASX_KC_COUNT; ++q) {
      tms[q] = g_tms[q];
      tn[q] = g_tn[q];
    }
  }
  double best[4];
  int arg[4];
  for (int q = 0; q < 4; ++q) {
    best[q] = 1e300;
    arg[q] = variant_of(tuned[q]);
  }
  asx_status st = ASX_OK;
  for (int v = 0; v < nv && !st; ++v) {
    for (int kc : tuned) set_variant(kc, v < variant_count(kc) ? v : 0);
    // one untimed call (first launch of a variant loads its code object)
    if ((st = asx_substep(stage, u0, out, c0, c1, c2, par, stream))) break;
    if ((st = asx_timing_enable(1))) break;
    for (int r = 0; r < reps && !st; ++r)
      st = asx_substep(stage, u0, out, c0, c1, c2, par, stream);
    for (int q = 0; q < 4 && !st; ++q) {
      if (v >= variant_count(tuned[q])) continue;
      int64_t n = 0;
      double ms = 0;
      if ((st = asx_timing_read(tuned[q], &n, &ms))) break;
      if (n > 0 && ms / n < best[q]) {
        best[q] = ms / n;
        arg[q] = v;
      }
    }
  }
  for (int q = 0; q < 4; ++q) set_variant(tuned[q], arg[q]);
  {
    std::lock_guard<std::mutex> l(g_tmu);
    asx_status st2 = drain_timing();
    for (int q = 0; q < ASX_KC_COUNT; ++q) {
      g_tms[q] = tms[q];
      g_tn[q] = tn[q];
    }
    g_timing.store(was_on);
    if (!st) st = st2;
  }
  return st;
}

// ---------------------------------------------------------------------------
// Independent zones and the learned con2prim guess
// ---------------------------------------------------------------------------
asx_status asx_zone_sample(int64_t nz, uint64_t seed, double gamma, double seed_rel,
                           double *prim, double *g, double *seed_out, double *target,
                           void *stream) {
  if (nz < 0) return ASX_ERR_SIZE;
  if (nz > 0 && (!prim || !g)) return ASX_ERR_NULL;
  if (!(gamma > 1.0) || !(seed_rel >= 0.0)) return ASX_ERR_PARAM;
  return dev_status(timed(ASX_KC_OTHER, (hipStream_t)stream, [&] {
    return launch_zone_sample(nz, seed, gamma, seed_rel, prim, g, seed_out, target,
                              (hipStream_t)stream);
  }));
}

asx_status asx_prim2con_batch(int64_t nz, const double *prim, const double *g, double *cons,
                              double *cold_seed, const asx_params *par, void *stream) {
  asx_status st = check_par(par);
  if (st) return st;
  if (tabulated(par)) return ASX_ERR_UNSUPPORTED; // asx_prim2con_batch_ye
  if (nz < 0) return ASX_ERR_SIZE;
  if (nz > 0 && (!prim || !g || !cons)) return ASX_ERR_NULL;
  const IdealGas eos = ideal_gas(par);
  return dev_status(timed(ASX_KC_OTHER, (hipStream_t)stream, [&] {
    return launch_prim2con_batch(nz, prim, g, cons, cold_seed, eos, par->Ye_atmo,
                                 (hipStream_t)stream);
  }));
}

asx_status asx_mlp_features(int64_t nz, const double *cons, const double *g, float *x,
                            void *stream) {
  if (nz < 0) return ASX_ERR_SIZE;
  if (nz > 0 && (!cons || !g || !x)) return ASX_ERR_NULL;
  return dev_status(timed(ASX_KC_OTHER, (hipStream_t)stream, [&] {
    return launch_mlp_features<kMlpIdeal>(nz, cons, g, x, (hipStream_t)stream);
  }));
}

asx_status asx_mlp_forward(int64_t nz, const double *cons, const double *g, const float *blob,
                           double *out, void *stream) {
  if (nz < 0) return ASX_ERR_SIZE;
  if (nz > 0 && (!cons || !g || !blob || !out)) return ASX_ERR_NULL;
  if (nz == 0) return ASX_OK;
  void *img = dev_scratch(SCR_MLP_IMAGE, stream, kNetImageBytes);
  if (!img) return ASX_ERR_DEVICE;
  return dev_status(timed(ASX_KC_MLP, (hipStream_t)stream, [&] {
    return launch_mlp_forward<kMlpIdeal>(nz, cons, g, blob, img, out, (hipStream_t)stream);
  }));
}

asx_status asx_mlp_train_step(int64_t nz, const double *cons, const double *g,
                              const double *target, float *blob, float *m, float *v, float lr,
                              int64_t step, float *partial, float *loss, float *grad,
                              void *stream) {
  if (nz < 1 || nz > (int64_t)1 << 30) return ASX_ERR_SIZE;
  if (!cons || !g || !target || !blob || !partial) return ASX_ERR_NULL;
  if (lr != 0.0f && (!m || !v || step < 1)) return ASX_ERR_PARAM;
  hipStream_t s = (hipStream_t)stream;
  const int nw = (int)((nz + 63) / 64);
  asx_status st = dev_status(timed(ASX_KC_MLP, s, [&] {
    return launch_mlp_grad<kMlpIdeal>(nz, cons, g, target, blob, partial, s);
  }));
  if (st) return st;
  return dev_status(timed(ASX_KC_MLP, s, [&] {
    return launch_mlp_adam<kMlpIdeal>(nw, partial, blob, m, v, lr, 0.9f, 0.999f, 1e-8f,
                           step < 1 ? 1 : step, loss, grad, s);
  }));
}

asx_status asx_c2p_batch_mlp(int64_t nz, const double *cons, const double *g,
                             const double *seed, const float *blob, int32_t kguess,
                             double *prim, double *flag, int32_t *iters, int32_t *fallback,
                             const asx_params *par, void *stream) {
  asx_status st = check_c2p_par(par);
  if (st) return st;
  if (par->c2p_prime != ASX_C2P_NOBLE || tabulated(par)) return ASX_ERR_UNSUPPORTED;
  if (nz < 0) return ASX_ERR_SIZE;
  if (kguess < 1) return ASX_ERR_PARAM;
  if (nz >= (int64_t(1) << 31)) return ASX_ERR_SIZE;
  if (nz > 0 && (!cons || !g || !seed || !blob || !prim || !flag || !iters))
    return ASX_ERR_NULL;
  if (nz == 0) return ASX_OK;
  // variant 0: the network inside the primary-solver kernel; variant 1: a
  // forward-pass kernel writing the guesses to library scratch first
  const int v = variant_of(ASX_KC_MLP);
  double *guess = nullptr;
  if (v == 1) {
    guess = (double *)dev_scratch(SCR_MLP_GUESS, stream, (size_t)nz * 2 * sizeof(double));
    if (!guess) return ASX_ERR_DEVICE;
  }
  void *img = dev_scratch(SCR_MLP_IMAGE, stream, kNetImageBytes);
  C2PQueue *Q = c2p_queue(stream, nz);
  if (!img || !Q) return ASX_ERR_DEVICE;
  return dev_status(timed(ASX_KC_C2P, (hipStream_t)stream, [&] {
    return launch_c2p_batch_mlp(nz, cons, g, seed, blob, kguess, prim, flag, iters, fallback,
                                ideal_gas(par), c2p_args(par), guess, img, Q->q, Q->n, v,
                                (hipStream_t)stream);
  }));
}

} // extern "C"

// ---------------------------------------------------------------------------
// Independent zones with composition (any 3-parameter EOS)
// ---------------------------------------------------------------------------
namespace {
// con2prim parameters with a table: Palenzuela on temperature, no
// pressure atmosphere (the paths the reference's table supports)
asx_status check_c2p_eos(const asx_params *p) {
  asx_status st = check_eos(p);
  if (st) return st;
  if (p->max_iter < 1 || !(p->c2p_tol > 0.0)) return ASX_ERR_PARAM;
  if (p->evolution_eos == ASX_EOS_TABULATED) {
    if (p->c2p_prime != ASX_C2P_PALENZUELA || p->use_entropy_fix) return ASX_ERR_UNSUPPORTED;
    if (p->c2p_second != ASX_C2P_PALENZUELA && p->c2p_second != ASX_C2P_NONE)
      return ASX_ERR_UNSUPPORTED;
    if (!p->use_temperature || p->use_press_atmo) return ASX_ERR_UNSUPPORTED;
    return ASX_OK;
  }
  if (p->c2p_prime != ASX_C2P_NOBLE && p->c2p_prime != ASX_C2P_PALENZUELA &&
      p->c2p_prime != ASX_C2P_ENTROPY)
    return ASX_ERR_UNSUPPORTED;
  if (p->c2p_second != ASX_C2P_NOBLE && p->c2p_second != ASX_C2P_PALENZUELA &&
      p->c2p_second != ASX_C2P_ENTROPY && p->c2p_second != ASX_C2P_NONE)
    return ASX_ERR_UNSUPPORTED;
  return ASX_OK;
}
} // namespace

extern "C" {

asx_status asx_eos_from_rho_temp(int64_t n, const double *rho, const double *temp,
                                 const double *ye, double *out, const asx_params *par,
                                 void *stream) {
  asx_status st = check_eos(par);
  if (st) return st;
  if (n < 0) return ASX_ERR_SIZE;
  if (n > 0 && (!rho || !temp || !ye || !out)) return ASX_ERR_NULL;
  hipStream_t s = (hipStream_t)stream;
  return dev_status(timed(ASX_KC_OTHER, s, [&] {
    return par->evolution_eos == ASX_EOS_TABULATED
               ? launch_eos_points(0, n, rho, temp, ye, out, tab_eos(par), s)
               : launch_eos_points(0, n, rho, temp, ye, out, ideal_gas(par), s);
  }));
}

asx_status asx_eos_from_rho_eps(int64_t n, const double *rho, const double *eps,
                                const double *ye, double *out, const asx_params *par,
                                void *stream) {
  asx_status st = check_eos(par);
  if (st) return st;
  if (n < 0) return ASX_ERR_SIZE;
  if (n > 0 && (!rho || !eps || !ye || !out)) return ASX_ERR_NULL;
  hipStream_t s = (hipStream_t)stream;
  return dev_status(timed(ASX_KC_OTHER, s, [&] {
    return par->evolution_eos == ASX_EOS_TABULATED
               ? launch_eos_points(1, n, rho, eps, ye, out, tab_eos(par), s)
               : launch_eos_points(1, n, rho, eps, ye, out, ideal_gas(par), s);
  }));
}

asx_status asx_prim2con_batch_ye(int64_t nz, const double *prim, const double *g, double *cons,
                                 double *thermo, const asx_params *par, void *stream) {
  asx_status st = check_eos(par);
  if (st) return st;
  if (nz < 0) return ASX_ERR_SIZE;
  if (nz > 0 && (!prim || !g || !cons)) return ASX_ERR_NULL;
  hipStream_t s = (hipStream_t)stream;
  return dev_status(timed(ASX_KC_OTHER, s, [&] {
    return par->evolution_eos == ASX_EOS_TABULATED
               ? launch_prim2con_ye(nz, prim, g, cons, thermo, tab_eos(par), s)
               : launch_prim2con_ye(nz, prim, g, cons, thermo, ideal_gas(par), s);
  }));
}

asx_status asx_c2p_batch_ye(int64_t nz, const double *cons, const double *g,
                            const double *seed, double *prim, double *flag, int32_t *iters,
                            const asx_params *par, void *stream) {
  asx_status st = check_c2p_eos(par);
  if (st) return st;
  if (nz < 0 || nz >= (int64_t(1) << 31)) return ASX_ERR_SIZE;
  if (nz > 0 && (!cons || !g || !seed || !prim || !flag || !iters)) return ASX_ERR_NULL;
  if (nz == 0) return ASX_OK;
  C2PQueue *Q = c2p_queue(stream, nz);
  if (!Q) return ASX_ERR_DEVICE;
  hipStream_t s = (hipStream_t)stream;
  const C2PArgs ca = c2p_args(par);
  return dev_status(timed(ASX_KC_C2P, s, [&] {
    return par->evolution_eos == ASX_EOS_TABULATED
               ? launch_c2p_ye(nz, cons, g, seed, prim, flag, iters, tab_eos(par), ca, Q->q,
                               Q->n, s)
               : launch_c2p_ye(nz, cons, g, seed, prim, flag, iters, ideal_gas(par), ca, Q->q,
                               Q->n, s);
  }));
}

asx_status asx_zone_sample_ye(int64_t nz, uint64_t seed, const double *box, double seed_rel,
                              double *prim, double *g, double *seed_out, double *target,
                              const asx_params *par, void *stream) {
  asx_status st = check_eos(par);
  if (st) return st;
  if (nz < 0) return ASX_ERR_SIZE;
  if (!box || (nz > 0 && (!prim || !g))) return ASX_ERR_NULL;
  Box7 b;
  for (int i = 0; i < 7; ++i) b.v[i] = box[i];
  if (!(b.v[0] > 0.0 && b.v[1] >= b.v[0] && b.v[2] > 0.0 && b.v[3] >= b.v[2] &&
        b.v[5] >= b.v[4] && b.v[6] >= 1.0 && std::isfinite(b.v[1]) && std::isfinite(b.v[3]) &&
        std::isfinite(b.v[6]) && seed_rel >= 0.0))
    return ASX_ERR_PARAM;
  hipStream_t s = (hipStream_t)stream;
  return dev_status(timed(ASX_KC_OTHER, s, [&] {
    return tabulated(par) ? launch_zone_sample_ye(nz, seed, b, seed_rel, prim, g, seed_out,
                                                  target, tab_eos(par), s)
                          : launch_zone_sample_ye(nz, seed, b, seed_rel, prim, g, seed_out,
                                                  target, ideal_gas(par), s);
  }));
}

asx_status asx_mlp_ye_features(int64_t nz, const double *cons, const double *g, float *x,
                               void *stream) {
  if (nz < 0) return ASX_ERR_SIZE;
  if (nz > 0 && (!cons || !g || !x)) return ASX_ERR_NULL;
  return dev_status(timed(ASX_KC_OTHER, (hipStream_t)stream, [&] {
    return launch_mlp_features<kMlpTab>(nz, cons, g, x, (hipStream_t)stream);
  }));
}

asx_status asx_mlp_ye_forward(int64_t nz, const double *cons, const double *g,
                              const float *blob, double *out, void *stream) {
  if (nz < 0) return ASX_ERR_SIZE;
  if (nz > 0 && (!cons || !g || !blob || !out)) return ASX_ERR_NULL;
  if (nz == 0) return ASX_OK;
  void *img = dev_scratch(SCR_MLP_IMAGE, stream, kNetImageBytes);
  if (!img) return ASX_ERR_DEVICE;
  return dev_status(timed(ASX_KC_MLP, (hipStream_t)stream, [&] {
    return launch_mlp_forward<kMlpTab>(nz, cons, g, blob, img, out, (hipStream_t)stream);
  }));
}

asx_status asx_mlp_ye_train_step(int64_t nz, const double *cons, const double *g,
                                 const double *target, float *blob, float *m, float *v,
                                 float lr, int64_t step, float *partial, float *loss,
                                 float *grad, void *stream) {
  if (nz < 1 || nz > (int64_t)1 << 30) return ASX_ERR_SIZE;
  if (!cons || !g || !target || !blob || !partial) return ASX_ERR_NULL;
  if (lr != 0.0f && (!m || !v || step < 1)) return ASX_ERR_PARAM;
  hipStream_t s = (hipStream_t)stream;
  const int nw = (int)((nz + 63) / 64);
  asx_status st = dev_status(timed(ASX_KC_MLP, s, [&] {
    return launch_mlp_grad<kMlpTab>(nz, cons, g, target, blob, partial, s);
  }));
  if (st) return st;
  return dev_status(timed(ASX_KC_MLP, s, [&] {
    return launch_mlp_adam<kMlpTab>(nw, partial, blob, m, v, lr, 0.9f, 0.999f, 1e-8f,
                                    step < 1 ? 1 : step, loss, grad, s);
  }));
}

asx_status asx_c2p_batch_ye_mlp(int64_t nz, const double *cons, const double *g,
                                const double *seed, const float *blob, double bracket,
                                double *prim, double *flag, int32_t *iters,
                                const asx_params *par, void *stream) {
  asx_status st = check_c2p_eos(par);
  if (st) return st;
  if (!tabulated(par)) return ASX_ERR_UNSUPPORTED; // the ideal gas: asx_c2p_batch_mlp
  if (!(bracket > 0.0 && bracket < 1.0)) return ASX_ERR_PARAM;
  if (nz < 0 || nz >= (int64_t(1) << 31)) return ASX_ERR_SIZE;
  if (nz > 0 && (!cons || !g || !seed || !blob || !prim || !flag || !iters)) return ASX_ERR_NULL;
  if (nz == 0) return ASX_OK;
  double *guess = (double *)dev_scratch(SCR_MLP_GUESS, stream, (size_t)nz * 2 * sizeof(double));
  void *img = dev_scratch(SCR_MLP_IMAGE, stream, kNetImageBytes);
  C2PQueue *Q = c2p_queue(stream, nz);
  if (!guess || !img || !Q) return ASX_ERR_DEVICE;
  hipStream_t s = (hipStream_t)stream;
  return dev_status(timed(ASX_KC_C2P, s, [&] {
    return launch_c2p_ye_mlp(nz, cons, g, seed, blob, bracket, prim, flag, iters, tab_eos(par),
                             c2p_args(par), guess, img, Q->q, Q->n, s);
  }));
}

} // extern "C"

// ---------------------------------------------------------------------------
// Batched patches
// ---------------------------------------------------------------------------
namespace {

// the primitives SYNCed after con2prim (AsterX/schedule.ccl:78, with Bvec)
constexpr int kNumSyncPrims = 12;
constexpr int kSyncPrims[kNumSyncPrims] = {ASX_RHO,   ASX_VELX,        ASX_VELY,    ASX_VELZ,
                                           ASX_EPS,   ASX_PRESS,       ASX_TEMPERATURE,
                                           ASX_ENTROPY, ASX_YE,        ASX_BVECX,
                                           ASX_BVECY, ASX_BVECZ};
// SSPRK3 stages: out = c0 u0 + c1 u + c2 dt rhs(u)
constexpr double kRK3[3][3] = {{0.0, 1.0, 1.0}, {0.75, 0.25, 0.25},
                               {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
constexpr int kStageIn[3] = {0, 1, 2}, kStageOut[3] = {1, 2, 0};

// Launch kinds of a plan, each with its per-patch boxes and workgroup list.
enum BatchKind {
  BK_FACE = 0,   // + d: faces, tiles of the face variant
  BK_AUX = 3,    // vertices, two layers into the ghosts
  BK_UPDATE = 4, // interior vertices, tiles of the update variant
  BK_BC = 5,     // + d: ghost layers normal to d
  BK_UBSTAG = 8, // + d: interior faces of the caller's dB_stag
  BK_C2P = 11,   // interior cells
  BK_COUNT = 12
};

struct BatchLaunch {
  BatchBox *box = nullptr;
  BatchRef *map = nullptr;
  int nmap = 0;
};

} // namespace

struct asx_batch_s {
  int dev = 0;
  int np = 0;
  asx_params par;
  double dt = 0;
  int bc_kind = 0;
  bool user_bstag = false;
  std::vector<asx_patch> pt; // caller's descriptors (state of buffer 0)
  std::vector<void *> allocs;
  int64_t bytes = 0;
  int64_t nzones = 0;
  KPatch *stab[3] = {}; // stage views: scratch slots substituted, state of buffer k
  KPatch *utab[3] = {}; // caller views, state of buffer k
  StatePtrs *st[3] = {};
  int *zbase = nullptr;
  int *queue = nullptr;
  unsigned *qcount = nullptr;
  std::map<std::pair<int, int>, BatchLaunch> launches; // (kind, variant)
  int var[4] = {0, 0, 0, 0}; // launch shapes: faces x, y, z, update
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_var[4] = {-1, -1, -1, -1};
  PersistArgs *pargs = nullptr; // one per SSPRK3 stage, device memory
  int pgrid = 0;                // blocks of a persistent stage launch
};

namespace {

hipError_t plan_alloc(asx_batch *b, size_t bytes, void **p) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e == hipSuccess) {
    b->allocs.push_back(*p);
    b->bytes += (int64_t)bytes;
  }
  return e;
}

template <class T>
asx_status plan_upload(asx_batch *b, const std::vector<T> &v, T **dst) {
  void *p = nullptr;
  if (plan_alloc(b, v.size() * sizeof(T), &p) != hipSuccess) return ASX_ERR_DEVICE;
  if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) !=
                        hipSuccess)
    return ASX_ERR_DEVICE;
  *dst = (T *)p;
  return ASX_OK;
}

// Boxes and workgroup list of launch kind k (variant v for faces / update).
asx_status plan_launch(asx_batch *b, int k, int v, BatchLaunch **out) {
  const auto key = std::make_pair(k, v);
  auto it = b->launches.find(key);
  if (it != b->launches.end()) {
    *out = &it->second;
    return ASX_OK;
  }
  std::vector<BatchBox> boxes(b->np);
  std::vector<BatchRef> refs;
  for (int p = 0; p < b->np; ++p) {
    const asx_patch &pt = b->pt[p];
    BatchBox &bb = boxes[p];
    int tile[3] = {0, 0, 0}; // 0: 256-point chunks of the flattened box
    if (k >= BK_FACE && k < BK_FACE + 3) {
      const int d = k - BK_FACE;
      int halo[3] = {1, 1, 1};
      halo[d] = 0;
      const Box bx = box(&pt, 1 << d, halo);
      for (int q = 0; q < 3; ++q) {
        bb.lo[q] = bx.lo[q];
        bb.cnt[q] = bx.cnt[q];
      }
      face_tile_extent(d, v, tile);
      tile[d] -= 1; // tiles along d overlap by one cell
    } else if (k == BK_AUX) {
      const int halo[3] = {2, 2, 2};
      const Box bx = box(&pt, 7, halo);
      for (int q = 0; q < 3; ++q) {
        bb.lo[q] = bx.lo[q];
        bb.cnt[q] = bx.cnt[q];
      }
    } else if (k == BK_UPDATE) {
      const int halo[3] = {0, 0, 0};
      const Box bx = box(&pt, 7, halo);
      for (int q = 0; q < 3; ++q) {
        bb.lo[q] = bx.lo[q];
        bb.cnt[q] = bx.cnt[q];
      }
      update_tile_extent(v, tile);
    } else if (k >= BK_BC && k < BK_BC + 3) {
      const int d = k - BK_BC; // bc_point's pass normal to d, widest centering
      for (int q = 0; q < 3; ++q) {
        bb.lo[q] = 0;
        bb.cnt[q] = q > d ? pt.n[q] + 1 : pt.n[q] + 2 * pt.ng + 1;
      }
      bb.cnt[d] = 2 * pt.ng;
    } else if (k >= BK_UBSTAG && k < BK_UBSTAG + 3) {
      const int d = k - BK_UBSTAG;
      const int halo[3] = {0, 0, 0};
      const Box bx = box(&pt, 1 << d, halo);
      for (int q = 0; q < 3; ++q) {
        bb.lo[q] = bx.lo[q];
        bb.cnt[q] = bx.cnt[q];
      }
    } else { // BK_C2P
      for (int q = 0; q < 3; ++q) {
        bb.lo[q] = pt.ng;
        bb.cnt[q] = pt.n[q];
      }
    }
    int64_t chunks;
    if (tile[0] > 0) {
      for (int q = 0; q < 3; ++q) bb.nt[q] = (bb.cnt[q] + tile[q] - 1) / tile[q];
      chunks = (int64_t)bb.nt[0] * bb.nt[1] * bb.nt[2];
    } else {
      const int64_t tot = (int64_t)bb.cnt[0] * bb.cnt[1] * bb.cnt[2];
      chunks = (tot + 255) / 256;
      bb.nt[0] = (int)chunks;
      bb.nt[1] = bb.nt[2] = 1;
    }
    for (int64_t c = 0; c < chunks; ++c) refs.push_back(BatchRef{p, (int)c});
  }
  if (refs.size() >= (size_t(1) << 31)) return ASX_ERR_SIZE;
  BatchLaunch L;
  L.nmap = (int)refs.size();
  asx_status st;
  if ((st = plan_upload(b, boxes, &L.box)) || (st = plan_upload(b, refs, &L.map))) return st;
  *out = &(b->launches[key] = L);
  return ASX_OK;
}

FluxArgs plan_flux_args(const asx_params *par) {
  FluxArgs A;
  A.rp = rparams(par);
  A.eos = ideal_gas(par);
  A.lo_method = par->loworder_method;
  A.flux_type = par->flux_type;
  A.rec_var = par->recon_type;
  A.with_T = par->reconstruct_with_temperature;
  A.write_uct = 0;
  A.bstag_from_A = 1; // the faces take the curl of the stage's A
  A.cnt_dev = nullptr;
  for (int q = 0; q < 3; ++q) A.lo[q] = A.cnt[q] = 0;
  A.with_aux = 0; // the batch launches aux over its own (patch, chunk) list
  A.aux = Box{};
  return A;
}

// Stage-s stencil work: faces (normal field from A), aux, update into
// buffer (1,2,0)[s].
asx_status plan_stencil(asx_batch *b, int s, hipStream_t str) {
  const int bi = kStageIn[s], bo = kStageOut[s];
  const KPatch *T = b->stab[bi];
  BatchLaunch *L;
  asx_status st;
  const FluxArgs A = plan_flux_args(&b->par);
  for (int d = 0; d < 3; ++d) {
    const int v = b->var[d];
    if ((st = plan_launch(b, BK_FACE + d, v, &L))) return st;
    if ((st = dev_status(timed(ASX_KC_FACE_X + d, str, [&] {
           return launch_face_batch(d, T, A, L->box, L->map, L->nmap,
                                    b->par.reconstruction_method, v, str);
         }))))
      return st;
  }
  if ((st = plan_launch(b, BK_AUX, 0, &L))) return st;
  if ((st = dev_status(timed(ASX_KC_AUX, str, [&] {
         return launch_aux_batch(T, L->box, L->map, L->nmap, str);
       }))))
    return st;
  const int v = b->var[3];
  if ((st = plan_launch(b, BK_UPDATE, v, &L))) return st;
  const RhsArgs R = rhs_args(&b->par);
  return dev_status(timed(ASX_KC_UPDATE, str, [&] {
    return launch_update_batch(T, b->st[0], b->st[bo], R, kRK3[s][0], kRK3[s][1],
                               kRK3[s][2] * b->dt, L->box, L->map, L->nmap, v, str);
  }));
}

asx_status plan_bc(asx_batch *b, const KPatch *T, const int *slots, int n, hipStream_t str) {
  BCSlots S;
  S.nf = n;
  for (int q = 0; q < n; ++q) {
    S.slot[q] = slots[q];
    S.cent[q] = centering(slots[q]);
  }
  for (int d = 0; d < 3; ++d) {
    BatchLaunch *L;
    asx_status st;
    if ((st = plan_launch(b, BK_BC + d, 0, &L))) return st;
    if ((st = dev_status(timed(ASX_KC_BC, str, [&] {
           return launch_bc_batch(d, T, S, L->map, L->nmap, b->bc_kind, str);
         }))))
      return st;
  }
  return ASX_OK;
}

asx_status plan_stage(asx_batch *b, int s, hipStream_t str) {
  asx_status st;
  if ((st = plan_stencil(b, s, str))) return st;
  const KPatch *U = b->utab[kStageOut[s]];
  if ((st = plan_bc(b, U, kStateSlots, ASX_NUM_STATE, str))) return st;
  BatchLaunch *L;
  if (b->user_bstag)
    for (int d = 0; d < 3; ++d) {
      if ((st = plan_launch(b, BK_UBSTAG + d, 0, &L))) return st;
      if ((st = dev_status(timed(ASX_KC_BSTAG, str, [&] {
             return launch_bstag_batch(d, U, L->box, L->map, L->nmap, str);
           }))))
        return st;
    }
  if ((st = plan_launch(b, BK_C2P, 0, &L))) return st;
  const IdealGas eos = ideal_gas(&b->par);
  const C2PArgs ca = c2p_args(&b->par);
  if ((st = dev_status(timed(ASX_KC_C2P, str, [&] {
         return launch_con2prim_main_batch(U, b->zbase, b->np, eos, ca, L->map, L->nmap,
                                           b->queue, b->qcount, str);
       }))))
    return st;
  if ((st = dev_status(timed(ASX_KC_C2P_REST, str, [&] {
         return launch_con2prim_rest_batch(U, b->zbase, b->np, eos, ca, b->nzones, b->queue,
                                           b->qcount, str);
       }))))
    return st;
  return plan_bc(b, U, kSyncPrims, kNumSyncPrims, str);
}

// Build every launch list the stages use with the current variants (no
// allocation may happen while a graph is recorded).
asx_status plan_prepare(asx_batch *b) {
  BatchLaunch *L;
  asx_status st;
  for (int k = 0; k < BK_COUNT; ++k) {
    int v = 0;
    if (k >= BK_FACE && k < BK_FACE + 3) v = b->var[k - BK_FACE];
    if (k == BK_UPDATE) v = b->var[3];
    if ((st = plan_launch(b, k, v, &L))) return st;
  }
  return ASX_OK;
}

void plan_drop_graph(asx_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

asx_status plan_capture(asx_batch *b) {
  const int *var = b->var;
  if (b->exec && std::equal(var, var + 4, b->graph_var)) return ASX_OK;
  plan_drop_graph(b);
  asx_status st = plan_prepare(b);
  if (st) return st;
  if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return ASX_ERR_DEVICE;
  g_capturing = true;
  for (int s = 0; s < 3 && !st; ++s) st = plan_stage(b, s, b->cap);
  g_capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->cap, &g);
  if (st || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return st ? st : ASX_ERR_DEVICE;
  }
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    b->exec = nullptr;
    plan_drop_graph(b);
    return ASX_ERR_DEVICE;
  }
  std::copy(var, var + 4, b->graph_var);
  return ASX_OK;
}

// The three stages' argument blocks of the persistent form (built once):
// launch lists of the persistent kernel's fixed shapes (face variant 0 of
// every direction, update tiles of kPersistUpdateTile = variant 1).
asx_status plan_persist_prepare(asx_batch *b) {
  if (b->pargs) return ASX_OK;
  const asx_params &par = b->par;
  const RhsArgs R = rhs_args(&par);
  if (!persist_supported(par.reconstruction_method, R.fd_order) ||
      par.c2p_prime != ASX_C2P_NOBLE || b->user_bstag)
    return ASX_ERR_UNSUPPORTED;
  int ex[3];
  update_tile_extent(1, ex);
  if (ex[0] != kPersistUpdateTile[0] || ex[1] != kPersistUpdateTile[1] ||
      ex[2] != kPersistUpdateTile[2])
    return ASX_ERR_DEVICE;
  BatchLaunch *L;
  asx_status st;
  auto list = [&](int k, int v, BatchList &out) {
    if ((st = plan_launch(b, k, v, &L))) return false;
    out = BatchList{L->box, L->map, L->nmap};
    return true;
  };
  PersistArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int d = 0; d < 3; ++d)
    if (!list(BK_FACE + d, 0, a.face[d]) || !list(BK_BC + d, 0, a.bc[d])) return st;
  if (!list(BK_AUX, 0, a.aux) || !list(BK_UPDATE, 1, a.update) || !list(BK_C2P, 0, a.c2p))
    return st;
  a.A = plan_flux_args(&par);
  a.R = R;
  a.eos = ideal_gas(&par);
  a.ca = c2p_args(&par);
  a.state.nf = ASX_NUM_STATE;
  for (int q = 0; q < ASX_NUM_STATE; ++q) {
    a.state.slot[q] = kStateSlots[q];
    a.state.cent[q] = centering(kStateSlots[q]);
  }
  a.prims.nf = kNumSyncPrims;
  for (int q = 0; q < kNumSyncPrims; ++q) {
    a.prims.slot[q] = kSyncPrims[q];
    a.prims.cent[q] = centering(kSyncPrims[q]);
  }
  a.bc_kind = b->bc_kind;
  a.zbase = b->zbase;
  a.np = b->np;
  a.queue = b->queue;
  a.qcount = b->qcount;
  std::vector<PersistArgs> v(3, a);
  for (int s = 0; s < 3; ++s) {
    const int bi = kStageIn[s], bo = kStageOut[s];
    v[s].T = b->stab[bi];
    v[s].U = b->utab[bo];
    v[s].u0 = b->st[0];
    v[s].out = b->st[bo];
    v[s].c0 = kRK3[s][0];
    v[s].c1 = kRK3[s][1];
    v[s].c2 = kRK3[s][2] * b->dt;
  }
  const int grid = persist_grid(par.reconstruction_method, R.fd_order);
  if (grid < 8) return ASX_ERR_DEVICE;
  if ((st = plan_upload(b, v, &b->pargs))) return st;
  b->pgrid = grid;
  return ASX_OK;
}

// a plan runs on the device it was created on
asx_status plan_device(const asx_batch *b) {
  int d = -1;
  if (hipGetDevice(&d) != hipSuccess) return ASX_ERR_DEVICE;
  return d == b->dev ? ASX_OK : ASX_ERR_PARAM;
}

void plan_free(asx_batch *b) {
  plan_drop_graph(b);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  for (void *p : b->allocs) (void)hipFree(p);
  delete b;
}

// asx_substep + asx_con2prim_group requirements of one patch
asx_status plan_check_patch(const asx_patch *pt, const asx_params *par) {
  asx_status st = check_geom(pt);
  if (st) return st;
  if (pt->ng < 3 || pt->ng < recon_ghosts(par->reconstruction_method)) return ASX_ERR_GHOSTS;
  if ((st = need_range(pt, ASX_ALP, ASX_KZZ))) return st;
  if ((st = need_range(pt, ASX_RHO, ASX_DBZ))) return st;
  if ((st = need_range(pt, ASX_SAVED_RHO, ASX_CON2PRIM_FLAG))) return st;
  if ((st = opt(pt, ASX_MASK_CC))) return st;
  if (par->use_mask && (st = need(pt, {ASX_MASK_VC}))) return st;
  if ((st = need(pt, {ASX_AVEC_X, ASX_AVEC_Y, ASX_AVEC_Z, ASX_PSI}))) return st;
  return ASX_OK;
}

} // namespace

extern "C" {

asx_status asx_batch_create(int32_t npatch, const asx_patch *patches, const asx_state *bufs,
                            const asx_params *par, double dt, int32_t bc_kind,
                            asx_batch **plan) {
  if (!plan) return ASX_ERR_NULL;
  *plan = nullptr;
  if (!patches || !bufs || !par) return ASX_ERR_NULL;
  if (npatch < 1) return ASX_ERR_SIZE;
  asx_status st;
  if ((st = check_c2p_par(par))) return st;
  if (par->use_uct || tabulated(par)) return ASX_ERR_UNSUPPORTED;
  if (bc_kind < ASX_BC_LINEAR || bc_kind > ASX_BC_COPY) return ASX_ERR_PARAM;
  if (!std::isfinite(dt)) return ASX_ERR_PARAM;
  bool ubst = patches[0].gf[ASX_DBX_STAG].data != nullptr;
  int64_t zones = 0;
  for (int p = 0; p < npatch; ++p) {
    const asx_patch *pt = &patches[p];
    if ((st = plan_check_patch(pt, par))) return st;
    if (par->reconstruct_with_temperature && (st = need(pt, {ASX_TEMPERATURE}))) return st;
    if ((pt->gf[ASX_DBX_STAG].data != nullptr) != ubst) return ASX_ERR_NULL;
    if (ubst && (st = need_range(pt, ASX_DBX_STAG, ASX_DBZ_STAG))) return st;
    for (int d = 0; d < 3; ++d) {
      if (bc_kind == ASX_BC_LINEAR && pt->n[d] < 2) return ASX_ERR_SIZE;
      if (bc_kind == ASX_BC_PERIODIC && pt->n[d] < pt->ng) return ASX_ERR_SIZE;
    }
    for (int k = 0; k < 2; ++k) {
      const asx_state &S = bufs[2 * p + k];
      for (int q = 0; q < ASX_NUM_STATE; ++q) {
        int64_t sj, sk;
        strides(pt->n, pt->ng, centering(kStateSlots[q]), sj, sk);
        if (!S.gf[q].data) return ASX_ERR_NULL;
        if (S.gf[q].sj != sj || S.gf[q].sk != sk) return ASX_ERR_LAYOUT;
      }
    }
    zones += (int64_t)pt->n[0] * pt->n[1] * pt->n[2];
  }
  if (zones >= (int64_t(1) << 31)) return ASX_ERR_SIZE;
  asx_batch *b = new asx_batch;
  b->np = npatch;
  b->par = *par;
  b->dt = dt;
  b->bc_kind = bc_kind;
  b->user_bstag = ubst;
  b->nzones = zones;
  b->pt.assign(patches, patches + npatch);
  for (int q = 0; q < 3; ++q) b->var[q] = variant_of(ASX_KC_FACE_X + q);
  b->var[3] = variant_of(ASX_KC_UPDATE);
  st = ASX_OK;
  auto fail = [&](asx_status e) {
    plan_free(b);
    return e;
  };
  if (hipGetDevice(&b->dev) != hipSuccess ||
      hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess)
    return fail(ASX_ERR_DEVICE);
  // one scratch pool: face fluxes and aux of every patch (the faces take
  // their normal field from A, so no dB_stag scratch)
  std::vector<int> sslots;
  for (int q = ASX_FXDENS; q <= ASX_FZBZ; ++q) sslots.push_back(q);
  for (int q = ASX_FX; q <= ASX_G; ++q) sslots.push_back(q);
  size_t pool = 0;
  for (int p = 0; p < npatch; ++p)
    for (int q : sslots)
      pool += ((size_t)npoints(patches[p].n, patches[p].ng, centering(q)) * 8 + 255) & ~size_t(255);
  void *base = nullptr;
  if (plan_alloc(b, pool, &base) != hipSuccess) return fail(ASX_ERR_DEVICE);
  std::vector<KPatch> stab[3], utab[3];
  std::vector<StatePtrs> sp[3];
  char *cur = (char *)base;
  for (int p = 0; p < npatch; ++p) {
    KPatch U = kpatch(&patches[p]);
    KPatch S = U;
    for (int q : sslots) {
      S.g[q] = (double *)cur;
      cur += ((size_t)npoints(patches[p].n, patches[p].ng, centering(q)) * 8 + 255) &
             ~size_t(255);
    }
    for (int k = 0; k < 3; ++k) {
      StatePtrs P;
      for (int q = 0; q < ASX_NUM_STATE; ++q)
        P.g[q] = k == 0 ? patches[p].gf[kStateSlots[q]].data : bufs[2 * p + k - 1].gf[q].data;
      KPatch Sk = S, Uk = U;
      for (int q = 0; q < ASX_NUM_STATE; ++q) Sk.g[kStateSlots[q]] = Uk.g[kStateSlots[q]] = P.g[q];
      stab[k].push_back(Sk);
      utab[k].push_back(Uk);
      sp[k].push_back(P);
    }
  }
  std::vector<int> zb(npatch + 1, 0);
  for (int p = 0; p < npatch; ++p)
    zb[p + 1] = zb[p] + patches[p].n[0] * patches[p].n[1] * patches[p].n[2];
  for (int k = 0; k < 3 && !st; ++k)
    if (!(st = plan_upload(b, stab[k], &b->stab[k])) &&
        !(st = plan_upload(b, utab[k], &b->utab[k])))
      st = plan_upload(b, sp[k], &b->st[k]);
  if (!st) st = plan_upload(b, zb, &b->zbase);
  if (st) return fail(st);
  void *q = nullptr, *qc = nullptr;
  if (plan_alloc(b, (size_t)zones * sizeof(int), &q) != hipSuccess ||
      plan_alloc(b, sizeof(unsigned), &qc) != hipSuccess)
    return fail(ASX_ERR_DEVICE);
  b->queue = (int *)q;
  b->qcount = (unsigned *)qc;
  if ((st = plan_prepare(b))) return fail(st);
  *plan = b;
  return ASX_OK;
}

asx_status asx_batch_stage(asx_batch *plan, int32_t stage, void *stream) {
  if (!plan) return ASX_ERR_NULL;
  if (stage < 0 || stage > 2) return ASX_ERR_PARAM;
  if (asx_status st = plan_device(plan)) return st;
  return plan_stage(plan, stage, (hipStream_t)stream);
}

asx_status asx_batch_step(asx_batch *plan, int32_t nsteps, int32_t use_graph, void *stream) {
  if (!plan) return ASX_ERR_NULL;
  if (nsteps < 0) return ASX_ERR_PARAM;
  hipStream_t s = (hipStream_t)stream;
  asx_status st = plan_device(plan);
  if (st) return st;
  if (use_graph < 0 || use_graph > 2) return ASX_ERR_PARAM;
  if (!use_graph) {
    for (int n = 0; n < nsteps && !st; ++n)
      for (int k = 0; k < 3 && !st; ++k) st = plan_stage(plan, k, s);
    return st;
  }
  if (use_graph == 2) {
    if ((st = plan_persist_prepare(plan))) return st;
    const int m = plan->par.reconstruction_method, o = rhs_args(&plan->par).fd_order;
    for (int n = 0; n < nsteps && !st; ++n)
      for (int k = 0; k < 3 && !st; ++k)
        st = dev_status(timed(ASX_KC_PERSIST, s, [&] {
          const Range range("asx_batch_step persistent stage");
          return launch_stage_persist(plan->pargs + k, m, o, plan->pgrid, s);
        }));
    return st;
  }
  if ((st = plan_capture(plan))) return st;
  for (int n = 0; n < nsteps && !st; ++n)
    st = dev_status(timed(ASX_KC_GRAPH, s, [&] {
      const Range range("asx_batch_step graph replay");
      return hipGraphLaunch(plan->exec, s);
    }));
  return st;
}

asx_status asx_batch_autotune(asx_batch *plan, int32_t reps, void *stream) {
  if (!plan) return ASX_ERR_NULL;
  if (reps < 1) return ASX_ERR_PARAM;
  if (asx_status st = plan_device(plan)) return st;
  hipStream_t s = (hipStream_t)stream;
  const int tuned[4] = {ASX_KC_FACE_X, ASX_KC_FACE_Y, ASX_KC_FACE_Z, ASX_KC_UPDATE};
  int nv = 1;
  for (int kc : tuned) nv = variant_count(kc) > nv ? variant_count(kc) : nv;
  bool was_on;
  double tms[ASX_KC_COUNT];
  int64_t tn[ASX_KC_COUNT];
  {
    std::lock_guard<std::mutex> l(g_tmu);
    asx_status st = drain_timing();
    if (st) return st;
    was_on = g_timing.load();
    for (int q = 0; q < ASX_KC_COUNT; ++q) {
      tms[q] = g_tms[q];
      tn[q] = g_tn[q];
    }
  }
  double best[4];
  int arg[4];
  for (int q = 0; q < 4; ++q) {
    best[q] = 1e300;
    arg[q] = plan->var[q];
  }
  asx_status st = ASX_OK;
  for (int v = 0; v < nv && !st; ++v) {
    for (int q = 0; q < 4; ++q) plan->var[q] = v < variant_count(tuned[q]) ? v : 0;
    // one untimed pass (first launch of a variant loads its code object)
    if ((st = plan_stencil(plan, 0, s))) break;
    if ((st = asx_timing_enable(1))) break;
    for (int r = 0; r < reps && !st; ++r) st = plan_stencil(plan, 0, s);
    for (int q = 0; q < 4 && !st; ++q) {
      if (v >= variant_count(tuned[q])) continue;
      int64_t n = 0;
      double ms = 0;
      if ((st = asx_timing_read(tuned[q], &n, &ms))) break;
      if (n > 0 && ms / n < best[q]) {
        best[q] = ms / n;
        arg[q] = v;
      }
    }
  }
  for (int q = 0; q < 4; ++q) plan->var[q] = arg[q];
  {
    std::lock_guard<std::mutex> l(g_tmu);
    asx_status st2 = drain_timing();
    for (int q = 0; q < ASX_KC_COUNT; ++q) {
      g_tms[q] = tms[q];
      g_tn[q] = tn[q];
    }
    g_timing.store(was_on);
    if (!st) st = st2;
  }
  if (!st) st = plan_prepare(plan);
  return st;
}

asx_status asx_batch_get_variant(const asx_batch *plan, int32_t kclass, int32_t *v) {
  if (!plan || !v) return ASX_ERR_NULL;
  if (kclass >= ASX_KC_FACE_X && kclass <= ASX_KC_FACE_Z) *v = plan->var[kclass - ASX_KC_FACE_X];
  else if (kclass == ASX_KC_UPDATE) *v = plan->var[3];
  else if (kclass >= 0 && kclass < ASX_KC_COUNT) *v = 0;
  else return ASX_ERR_PARAM;
  return ASX_OK;
}

asx_status asx_batch_info(const asx_batch *plan, int64_t *zones, int64_t *device_bytes) {
  if (!plan) return ASX_ERR_NULL;
  if (zones) *zones = plan->nzones;
  if (device_bytes) *device_bytes = plan->bytes;
  return ASX_OK;
}

asx_status asx_batch_destroy(asx_batch *plan) {
  if (!plan) return ASX_OK;
  int cur = 0;
  const int dev = plan->dev;
  (void)hipGetDevice(&cur);
  if (cur != dev) (void)hipSetDevice(dev);
  const hipError_t e = hipDeviceSynchronize();
  plan_free(plan);
  if (cur != dev) (void)hipSetDevice(cur);
  return e == hipSuccess ? ASX_OK : ASX_ERR_DEVICE;
}

} // extern "C"
