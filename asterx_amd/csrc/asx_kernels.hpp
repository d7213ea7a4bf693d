// Kernel argument structs and launcher declarations shared by the .hip
// translation units and the C ABI layer.
#pragma once

#include <vector>

#include "asx_c2p.hpp"
#include "asx_device.hpp"

namespace asx {

// A box of points [lo, lo+cnt) flattened x-fastest.
struct Box {
  int lo[3], cnt[3];
  ASX_DEV int64_t total() const { return (int64_t)cnt[0] * cnt[1] * cnt[2]; }
  ASX_DEV void ijk(int64_t t, int &i, int &j, int &k) const {
    i = lo[0] + (int)(t % cnt[0]);
    j = lo[1] + (int)((t / cnt[0]) % cnt[1]);
    k = lo[2] + (int)(t / ((int64_t)cnt[0] * cnt[1]));
  }
  __host__ int64_t htotal() const { return (int64_t)cnt[0] * cnt[1] * cnt[2]; }
  __host__ dim3 grid() const { return dim3((unsigned)((htotal() + 255) / 256)); }
};

// Point of this work-item in tile b of a grid of ntx x nty tiles of
// TX x TY x TZ points over the box (lo, cnt); false past the box.
template <int TX, int TY, int TZ>
ASX_DEV bool tile_point(const int lo[3], const int cnt[3], int b, int ntx, int nty, int &i,
                        int &j, int &k) {
  const int bx = b % ntx, by = (b / ntx) % nty, bz = b / (ntx * nty);
  const int t = (int)threadIdx.x;
  const int x = bx * TX + t % TX, y = by * TY + (t / TX) % TY, z = bz * TZ + t / (TX * TY);
  i = lo[0] + x;
  j = lo[1] + y;
  k = lo[2] + z;
  return x < cnt[0] && y < cnt[1] && z < cnt[2];
}

// A box covered by workgroup tiles of TX x TY x TZ points (one work-item
// each), tiles ordered x, y, z and spread over the XCDs in contiguous runs.
// Deeper tiles (TZ > 1) read each shared vertex / face plane once per tile
// instead of once per plane.
template <int TX, int TY, int TZ = 1> struct TileGrid {
  static constexpr int kThreads = TX * TY * TZ;
  int ntx, nty, ntz;
  __host__ static TileGrid make(const Box &B) {
    return TileGrid{(B.cnt[0] + TX - 1) / TX, (B.cnt[1] + TY - 1) / TY,
                    (B.cnt[2] + TZ - 1) / TZ};
  }
  __host__ dim3 grid() const { return dim3((unsigned)(ntx * nty * ntz)); }
  // point of this work-item; false past the box
  ASX_DEV bool point(const Box &B, int &i, int &j, int &k) const {
    const int b = xcd_tile((int)blockIdx.x, ntx * nty * ntz);
    return tile_point<TX, TY, TZ>(B.lo, B.cnt, b, ntx, nty, i, j, k);
  }
};

// The same tiles visited in bands of YS tile rows: x fastest, then the YS
// rows of the band, then z, then the next band.  Tiles sharing a z plane of
// vertices or z faces are then YS * ntx launches apart instead of a whole
// x-y plane of tiles, close enough to meet in L2.
template <int TX, int TY, int TZ, int YS> struct TileGridBand {
  static constexpr int kThreads = TX * TY * TZ;
  int ntx, nty, ntz;
  __host__ static TileGridBand make(const Box &B) {
    return TileGridBand{(B.cnt[0] + TX - 1) / TX, (B.cnt[1] + TY - 1) / TY,
                        (B.cnt[2] + TZ - 1) / TZ};
  }
  __host__ dim3 grid() const { return dim3((unsigned)(ntx * nty * ntz)); }
  ASX_DEV bool point(const Box &B, int &i, int &j, int &k) const {
    const int b = xcd_tile((int)blockIdx.x, ntx * nty * ntz);
    const int full = nty / YS, blk = YS * ntz * ntx;
    int bx, by, bz;
    if (b < full * blk) {
      const int yb = b / blk, r = b % blk;
      bz = r / (YS * ntx);
      const int r2 = r % (YS * ntx);
      by = yb * YS + r2 / ntx;
      bx = r2 % ntx;
    } else {
      const int rem = nty - full * YS, b2 = b - full * blk;
      bz = b2 / (rem * ntx);
      const int r2 = b2 % (rem * ntx);
      by = full * YS + r2 / ntx;
      bx = r2 % ntx;
    }
    return tile_point<TX, TY, TZ>(B.lo, B.cnt, bx + ntx * (by + nty * bz), ntx, nty, i, j, k);
  }
};

// Launch-shape variants of the tunable kernels (selected per call; see
// asx_set_variant / asx_autotune_substep in the C ABI).
constexpr int kUpdateVariants = 7;
constexpr int kFaceVariants = 8; // 5-7 stage the tile's vertex metric in LDS
constexpr int kTmunuVariants = 9;

// Batched launches over many patches: the workgroups of one launch are
// listed as (patch, chunk) pairs, chunk = tile index or 256-point block of
// that patch's box; the per-patch data come from device tables.
struct BatchRef {
  int pid, chunk;
};
struct BatchBox {
  int lo[3], cnt[3], nt[3]; // box of points and its tile grid
};

struct FluxArgs {
  RParams rp;
  IdealGas eos;
  TabEOS tab; // used by the table instantiations (evolution_eos = Tabulated3d)
  int lo_method, flux_type, rec_var, with_T, write_uct;
  int bstag_from_A; // normal face field as the curl of A at the face (no dB_stag read)
  int lo[3], cnt[3];
  unsigned long long *cnt_dev; // [nan_faces, lo_faces] or null
  int with_aux; // the face launch also computes the vertex auxiliaries of box aux
  Box aux;
};

struct RhsArgs {
  RParams rp;
  int recon, gauge, use_uct, fd_order, rec_var;
  double lorenz_damp_fac;
};

// Evolved state pointers (asx_state order) for the fused update.
struct StatePtrs {
  double *g[ASX_NUM_STATE];
};

// tab: the EOS is A.tab (PPM, WENO-Z and MP5 only; reconstruction in T)
hipError_t launch_flux(int d, const KPatch &P, const FluxArgs &A, int method, int variant,
                       hipStream_t s, bool tab = false);
bool flux_tab_method(int method);
hipError_t launch_bstag(int d, const KPatch &P, const Box &B, double *out,
                        hipStream_t s);
hipError_t launch_dB(const KPatch &P, const Box &B, hipStream_t s);
hipError_t launch_B_from_dB(const KPatch &P, const Box &B, hipStream_t s);
hipError_t launch_prim2con(const KPatch &P, const Box &B, hipStream_t s);

// eT_{mu nu} += T_{mu nu} at the interior vertices (order 2 or 4; use_z:
// velocities from zvec instead of vel; variant: tile shape and wave budget,
// 0 .. kTmunuVariants-1).  The cell grid functions must span < 4 GiB each.
hipError_t launch_tmunu(const KPatch &P, int order, int use_z, int variant, hipStream_t s);
// regrid_error on the cells of B: max over nv grid functions (slot, centering)
constexpr int kMaxRegridVars = 9;
struct RegridArgs {
  int nv, method;
  double eps;
  int slot[kMaxRegridVars], cent[kMaxRegridVars];
};
hipError_t launch_regrid(const KPatch &P, const Box &B, const RegridArgs &A, hipStream_t s);

// Refinement by 2 between a coarse and a fine patch (asx_amr.hip).  off[d]:
// fine index of the fine patch's first interior point counted from twice
// the coarse patch's first interior point; strides per centering code.
constexpr int kMaxAmrSlots = 64;
__host__ __device__ inline int floor2(int v) { return (v - (v & 1)) / 2; }
// Coarse points of the prolongation stencil of fine index I along one axis
// (I counted from twice the coarse patch's first interior point): the first
// coarse index (counted from that point) and the number of points.  A
// vertex axis takes the coinciding point (even I) or the order + 1 points
// centred on the midpoint; a cell axis the order cells centred on the parent.
__host__ __device__ inline void axis_span(int I, bool vertex, int order, int &start, int &n) {
  if (vertex) {
    if ((I & 1) == 0) {
      start = I / 2;
      n = 1;
    } else {
      start = floor2(I - 1) - (order - 1) / 2;
      n = order + 1;
    }
  } else {
    start = floor2(I) - (order - 1) / 2;
    n = order;
  }
}
// The coarse blocks of a fine grid function (block a holds the fine points
// 2a and 2a+1 of each axis, fine indices counted from twice the coarse
// patch's first interior point) over its extent E[d] with fng ghosts and
// interior offset off[d]: nbk[d] blocks per axis, of which the first gl[d]
// and the last gh[d] hold ghost points.  decode() maps t to block indices
// over all blocks, or over the shell of blocks holding ghost points: the
// low and high z planes, then the y rows of the other planes, then the x
// ends of the remaining rows.
struct BlockShell {
  int nbk[3], gl[3], gh[3];
  __host__ __device__ BlockShell(const int off[3], const int E[3], int fng) {
    for (int d = 0; d < 3; ++d) {
      const int a0 = floor2(off[d] - fng);
      nbk[d] = floor2(off[d] - fng + E[d] - 1) - a0 + 1;
      gl[d] = off[d] / 2 - a0;
      gl[d] = gl[d] < 0 ? 0 : gl[d] > nbk[d] ? nbk[d] : gl[d];
      const int bh = floor2(E[d] - 2 * fng + off[d]) - a0; // first block with a high ghost
      gh[d] = nbk[d] - (bh < gl[d] ? gl[d] : bh > nbk[d] ? nbk[d] : bh);
    }
  }
  __host__ __device__ int64_t total(bool shell) const {
    const int64_t all = (int64_t)nbk[0] * nbk[1] * nbk[2];
    if (!shell) return all;
    return all - (int64_t)(nbk[0] - gl[0] - gh[0]) * (nbk[1] - gl[1] - gh[1]) *
                     (nbk[2] - gl[2] - gh[2]);
  }
  __host__ __device__ void decode(int64_t t, bool shell, int b[3]) const {
    if (!shell) {
      b[0] = (int)(t % nbk[0]);
      b[1] = (int)((t / nbk[0]) % nbk[1]);
      b[2] = (int)(t / ((int64_t)nbk[0] * nbk[1]));
      return;
    }
    const int64_t plane = (int64_t)nbk[0] * nbk[1];
    const int tz = gl[2] + gh[2], ty = gl[1] + gh[1], tx = gl[0] + gh[0];
    if (t < tz * plane) {
      const int p = (int)(t / plane), r = (int)(t % plane);
      b[2] = p < gl[2] ? p : nbk[2] - gh[2] + (p - gl[2]);
      b[1] = r / nbk[0];
      b[0] = r % nbk[0];
      return;
    }
    t -= tz * plane;
    const int mz = nbk[2] - tz;
    if (t < (int64_t)ty * nbk[0] * mz) {
      b[0] = (int)(t % nbk[0]);
      const int q = (int)(t / nbk[0]), row = q % ty;
      b[1] = row < gl[1] ? row : nbk[1] - gh[1] + (row - gl[1]);
      b[2] = gl[2] + q / ty;
      return;
    }
    t -= (int64_t)ty * nbk[0] * mz;
    const int my = nbk[1] - ty;
    const int col = (int)(t % tx), q = (int)(t / tx);
    b[0] = col < gl[0] ? col : nbk[0] - gh[0] + (col - gl[0]);
    b[1] = gl[1] + q % my;
    b[2] = gl[2] + q / my;
  }
};

struct ProlongArgs {
  const double *src[kMaxAmrSlots];
  double *dst[kMaxAmrSlots];
  int cent[kMaxAmrSlots];
  int nslot, order, region; // region 0: every fine point, 1: ghost points only
  int off[3], fn[3], fng, cng;
  int fsj[8], fsk[8], csj[8], csk[8];
};
struct RestrictArgs {
  const double *src[kMaxAmrSlots];
  double *dst[kMaxAmrSlots];
  int cent[kMaxAmrSlots];
  Box box[kMaxAmrSlots]; // coarse points written, per slot
  int nslot;
  int off[3], fng, cng;
  int fsj[8], fsk[8], csj[8], csk[8];
};
// max_blocks: the largest number of coarse blocks (2 x 2 x 2 fine points)
// of one slot; max_points: the largest coarse box of one slot
hipError_t launch_prolong(const ProlongArgs &A, int64_t max_blocks, hipStream_t s);
hipError_t launch_restrict(const RestrictArgs &A, int64_t max_points, hipStream_t s);

// Debug poison: fill n doubles with a bit pattern; add to *cnt the points of
// box B (strides sj, sk) holding it
hipError_t launch_poison_fill(double *p, int64_t n, uint64_t bits, hipStream_t s);
hipError_t launch_poison_count(const double *p, int sj, int sk, const Box &B, uint64_t bits,
                               unsigned long long *cnt, hipStream_t s);

// Regrid tags on blocks of blk[d] interior cells (cell layout strides sj, sk)
struct TagArgs {
  const double *err;
  uint8_t *tags;
  int sj, sk, ng, n[3], blk[3], nb[3], buf;
  double thr;
};
hipError_t launch_regrid_tags(const TagArgs &A, hipStream_t s);
// Berger-Rigoutsos clusters of the block tags (inclusive block indices)
struct ClusterBox {
  int lo[3], hi[3];
};
std::vector<ClusterBox> cluster_tags(const uint8_t *tags, const int nb[3], double grid_eff);
hipError_t launch_sources(const KPatch &P, const Box &B, const RhsArgs &A,
                          hipStream_t s);
hipError_t launch_rhs(const KPatch &P, const RhsArgs &A, hipStream_t s);
hipError_t launch_lincomb(double *dst, int64_t np, double scale, int nsrc,
                          const double *fac, double *const *src, hipStream_t s);
hipError_t launch_bc(const KPatch &P, int slot, int kind, hipStream_t s);
// Grid functions filled by one multi-field boundary launch.
constexpr int kMaxBC = 64;
struct BCList {
  double *g[kMaxBC];
  int cent[kMaxBC];
  int nf;
};
hipError_t launch_bc_many(const KPatch &P, const BCList &L, int kind, hipStream_t s);
hipError_t launch_update(const KPatch &P, const RhsArgs &A, const StatePtrs &u0,
                         const StatePtrs &out, double c0, double c1, double c2,
                         int variant, hipStream_t s);

// con2prim of the patch interior in two passes: every zone through the
// primary solver, the zones it failed queued (queue: one int per interior
// zone, qcount: one unsigned) and run through the full chain.
hipError_t launch_con2prim_main(const KPatch &P, const IdealGas &eos, const C2PArgs &a,
                                bool from_A, unsigned long long *hist, int *queue,
                                unsigned *qcount, hipStream_t s);
hipError_t launch_con2prim_rest(const KPatch &P, const IdealGas &eos, const C2PArgs &a,
                                bool from_A, unsigned long long *hist, const int *queue,
                                const unsigned *qcount, hipStream_t s);
// with a table: Palenzuela primary
hipError_t launch_con2prim_main(const KPatch &P, const TabEOS &eos, const C2PArgs &a,
                                bool from_A, unsigned long long *hist, int *queue,
                                unsigned *qcount, hipStream_t s);
hipError_t launch_con2prim_rest(const KPatch &P, const TabEOS &eos, const C2PArgs &a,
                                bool from_A, unsigned long long *hist, const int *queue,
                                const unsigned *qcount, hipStream_t s);

// ---------------------------------------------------------------------------
// Batched launches over many patches (device tables, one entry per patch;
// map = the (patch, chunk) list of the launch's workgroups)
// ---------------------------------------------------------------------------
// Tile extents (x, y, z work-items) of the launch-shape variants.
void face_tile_extent(int d, int variant, int ex[3]);
void update_tile_extent(int variant, int ex[3]);
// Slots of a multi-field boundary fill shared by every patch.
struct BCSlots {
  int slot[kMaxBC], cent[kMaxBC];
  int nf;
};

hipError_t launch_face_batch(int d, const KPatch *tab, const FluxArgs &A, const BatchBox *box,
                             const BatchRef *map, int nmap, int method, int variant,
                             hipStream_t s);
hipError_t launch_aux_batch(const KPatch *tab, const BatchBox *box, const BatchRef *map,
                            int nmap, hipStream_t s);
// face B from the vector potential into each patch's dB_stag slot d
hipError_t launch_bstag_batch(int d, const KPatch *tab, const BatchBox *box,
                              const BatchRef *map, int nmap, hipStream_t s);
hipError_t launch_update_batch(const KPatch *tab, const StatePtrs *u0, const StatePtrs *out,
                               const RhsArgs &A, double c0, double c1, double c2,
                               const BatchBox *box, const BatchRef *map, int nmap,
                               int variant, hipStream_t s);
// boundary fill along axis d; map chunks are 256-point blocks of the 2 ng
// ghost layers normal to d, sized for the widest centering
hipError_t launch_bc_batch(int d, const KPatch *tab, const BCSlots &L, const BatchRef *map,
                           int nmap, int kind, hipStream_t s);
// con2prim of the patch interiors; zbase[p] = first global zone of patch p
// (np + 1 entries), the queue holds global zone indices
hipError_t launch_con2prim_main_batch(const KPatch *tab, const int *zbase, int np,
                                      const IdealGas &eos, const C2PArgs &a,
                                      const BatchRef *map, int nmap, int *queue,
                                      unsigned *qcount, hipStream_t s);
hipError_t launch_con2prim_rest_batch(const KPatch *tab, const int *zbase, int np,
                                      const IdealGas &eos, const C2PArgs &a, int64_t nzones,
                                      const int *queue, const unsigned *qcount,
                                      hipStream_t s);
// A whole SSPRK3 stage of a batch as one persistent cooperative launch
// (asx_plan.hip): the launch lists of each phase and everything the phases
// read, in device memory.  Faces use tile variant 0 of every direction, the
// update 32x4x2 tiles (kPersistUpdateTile), con2prim Noble as the primary.
struct BatchList {
  const BatchBox *box;
  const BatchRef *map;
  int n;
};
struct PersistArgs {
  const KPatch *T;  // stage input views (faces, aux and update read them)
  const KPatch *U;  // views of the output buffer (fills and con2prim)
  const StatePtrs *u0, *out;
  FluxArgs A;
  RhsArgs R;
  double c0, c1, c2;
  IdealGas eos;
  C2PArgs ca;
  BatchList face[3], aux, update, bc[3], c2p;
  BCSlots state, prims; // grid functions filled after the update / con2prim
  int bc_kind;
  const int *zbase;
  int np;
  int *queue;
  unsigned *qcount;
};
constexpr int kPersistUpdateTile[3] = {32, 4, 2};
// the reconstruction methods and fd orders the persistent stage is built for
bool persist_supported(int method, int fd_order);
// blocks of one launch: every block resident at once (cooperative launch),
// a multiple of 8 (one run of the lists per XCD); 0 on error
int persist_grid(int method, int fd_order);
hipError_t launch_stage_persist(const PersistArgs *dev_args, int method, int fd_order,
                                int grid, hipStream_t s);

// ---------------------------------------------------------------------------
// Independent zones: sampler, batched prim2con, learned con2prim guess
// ---------------------------------------------------------------------------
// Network: 5 inputs -> 64 tanh -> 64 tanh -> 2 outputs.  One FP32 blob:
//   W1[64][5] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]   (4674 parameters)
//   in_mean[5] in_istd[5] out_mean[2] out_std[2]         (normalisation)
// Network layout of NIN inputs -> 64 tanh -> 64 tanh -> 2: offsets into
// the float32 blob (weights, biases, then the normalisation statistics).
// NIN = 5 for the ideal gas, 7 for a table (the 5 scale-free inputs,
// log D and Ye).
template <int NIN> struct MlpLayout {
  static constexpr int In = NIN, H = 64, Out = 2;
  static constexpr int W1 = 0, B1 = W1 + H * In, W2 = B1 + H, B2 = W2 + H * H, W3 = B2 + H,
                       B3 = W3 + Out * H, NParam = B3 + Out;
  static constexpr int InMean = NParam, InIstd = InMean + In, OutMean = InIstd + In,
                       OutStd = OutMean + Out, NBlob = OutStd + Out;
  static constexpr int Partial = NParam + 1; // per-wavefront gradient + loss
  static constexpr int NCons = NIN == 5 ? 8 : 9; // rows of the cons batch read
  // hidden activation: ReLU for the ideal-gas network (one VALU operation,
  // so the forward pass costs the solver little), tanh for the table's
  static constexpr bool Relu = NIN == 5;
};
constexpr int kMlpIdeal = 5, kMlpTab = 7;
constexpr int kMlpNParam = MlpLayout<kMlpIdeal>::NParam, kMlpNBlob = MlpLayout<kMlpIdeal>::NBlob;
constexpr int kMlpYeNParam = MlpLayout<kMlpTab>::NParam, kMlpYeNBlob = MlpLayout<kMlpTab>::NBlob;

hipError_t launch_zone_sample(int64_t nz, uint64_t seed, double gamma, double seed_rel,
                              double *prim, double *g, double *sd, double *target,
                              hipStream_t s);
hipError_t launch_prim2con_batch(int64_t nz, const double *prim, const double *g,
                                 double *cons, double *cold, const IdealGas &eos, double Ye,
                                 hipStream_t s);
// the network of NIN inputs (kMlpIdeal or kMlpTab) on cons of
// MlpLayout<NIN>::NCons rows
template <int NIN>
hipError_t launch_mlp_features(int64_t nz, const double *cons, const double *g, float *x,
                               hipStream_t s);
// img: device scratch of kNetImageBytes for the network's inference image
// (bf16 weights and fp32 biases, rebuilt from blob on every call)
constexpr size_t kNetImageBytes = 10512;
template <int NIN>
hipError_t launch_mlp_forward(int64_t nz, const double *cons, const double *g,
                              const float *blob, void *img, double *out, hipStream_t s);
template <int NIN>
hipError_t launch_mlp_grad(int64_t nz, const double *cons, const double *g,
                           const double *target, const float *blob, float *partial,
                           hipStream_t s);
template <int NIN>
hipError_t launch_mlp_adam(int nwaves, const float *partial, float *blob, float *m,
                           float *v, float lr, float b1, float b2, float eps, int64_t step,
                           float *loss, float *grad, hipStream_t s);
// the table network's guided con2prim on a table; guess: scratch of 2 nz
// doubles
hipError_t launch_c2p_ye_mlp(int64_t nz, const double *cons, const double *g,
                             const double *seed, const float *blob, double xrel, double *prim,
                             double *flag, int32_t *iters, const TabEOS &eos, const C2PArgs &a,
                             double *guess, void *img, int *queue, unsigned *qcount,
                             hipStream_t s);
// zones with composition: box = rho lo, hi, T lo, hi, Ye lo, hi, W max
struct Box7 {
  double v[7];
};
hipError_t launch_zone_sample_ye(int64_t nz, uint64_t seed, const Box7 &box, double seed_rel,
                                 double *prim, double *g, double *sd, double *target,
                                 const IdealGas &eos, hipStream_t s);
hipError_t launch_zone_sample_ye(int64_t nz, uint64_t seed, const Box7 &box, double seed_rel,
                                 double *prim, double *g, double *sd, double *target,
                                 const TabEOS &eos, hipStream_t s);
// guess: scratch of 2 nz doubles; queue: nz ints; qcount: one unsigned
hipError_t launch_c2p_batch_mlp(int64_t nz, const double *cons, const double *g,
                                const double *seed, const float *blob, int kguess,
                                double *prim, double *flag, int32_t *iters,
                                int32_t *fallback, const IdealGas &eos, const C2PArgs &a,
                                double *guess, void *img, int *queue, unsigned *qcount,
                                int variant, hipStream_t s);

// Independent zones with composition, any 3-parameter EOS (asx_eos.hip).
// mode 0: (rho, T, Ye) -> out[4][n] = press, eps, c_s, entropy;
// mode 1: (rho, eps, Ye) -> out[3][n] = T, press, eps after clamping
hipError_t launch_eos_points(int mode, int64_t n, const double *rho, const double *x,
                             const double *ye, double *out, const IdealGas &eos,
                             hipStream_t s);
hipError_t launch_eos_points(int mode, int64_t n, const double *rho, const double *x,
                             const double *ye, double *out, const TabEOS &eos,
                             hipStream_t s);
hipError_t launch_prim2con_ye(int64_t nz, const double *prim, const double *g, double *cons,
                              double *thermo, const IdealGas &eos, hipStream_t s);
hipError_t launch_prim2con_ye(int64_t nz, const double *prim, const double *g, double *cons,
                              double *thermo, const TabEOS &eos, hipStream_t s);
// queue: nz ints, qcount: one unsigned (device)
hipError_t launch_c2p_ye(int64_t nz, const double *cons, const double *g, const double *seed,
                         double *prim, double *flag, int32_t *iters, const IdealGas &eos,
                         const C2PArgs &a, int *queue, unsigned *qcount, hipStream_t s);
hipError_t launch_c2p_ye(int64_t nz, const double *cons, const double *g, const double *seed,
                         double *prim, double *flag, int32_t *iters, const TabEOS &eos,
                         const C2PArgs &a, int *queue, unsigned *qcount, hipStream_t s);

hipError_t launch_c2p_batch(int64_t nz, const double *cons, const double *g,
                            const double *seed, double *prim, double *flag,
                            int32_t *iters, const IdealGas &eos,
                            const C2PArgs &a, hipStream_t s);

} // namespace asx
