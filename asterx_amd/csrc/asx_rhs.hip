// Source terms, RHS (flux divergence, CT EMF, gauge), RK combination,
// boundary fill and the fused update of the substep, for gfx950.
#include <type_traits>

#include "asx_update.hpp"

namespace asx {

template <int ORD>
__global__ __launch_bounds__(256) void k_sources(KPatch P, Box B, RhsArgs A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  double r[7];
  source_cell<ORD>(P, i, j, k, A, r);
  const int oc = off(P, 0, i, j, k);
#pragma unroll
  for (int q = 0; q < 7; ++q) P.g[kRhsSlot(q)][oc] = r[q];
}

__global__ __launch_bounds__(256) void k_rhs_hydro(KPatch P, Box B) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  const int oc = off(P, 0, i, j, k);
#pragma unroll
  for (int q = 0; q < 7; ++q) P.g[kRhsSlot(q)][oc] += flux_div(P, i, j, k, kFluxVar(q));
}

template <int I>
__global__ __launch_bounds__(256) void k_rhs_avec(KPatch P, Box B, RhsArgs A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  constexpr int CE = 7 ^ (1 << I);
  P.g[ASX_AVEC_X_RHS + I][off(P, CE, i, j, k)] = avec_rhs<I>(P, i, j, k, A);
}

__global__ __launch_bounds__(256) void k_rhs_psi(KPatch P, Box B, RhsArgs A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  P.g[ASX_PSI_RHS][off(P, 7, i, j, k)] = psi_rhs(P, i, j, k, A);
}

// dst = scale*dst + sum fac[k]*src[k] (scale == 0: dst not read)
struct LinArgs {
  double fac[8];
  const double *src[8];
  int nsrc;
  double scale;
};
__global__ __launch_bounds__(256) void k_lincomb(double *dst, int64_t np, LinArgs L) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double acc = (L.scale == 0.0) ? 0.0 : L.scale * dst[p];
  for (int q = 0; q < L.nsrc; ++q) acc += L.fac[q] * L.src[q][p];
  dst[p] = acc;
}

// One pass of the boundary fill along axis D over every (a, b) line.
template <int D>
__global__ __launch_bounds__(256) void k_bc(double *g, int ex, int ey, int ez, int lo,
                                            int hi, int n, int kind) {
  constexpr int A = (D + 1) % 3, Bx = (D + 2) % 3;
  const int e[3] = {ex, ey, ez};
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int na = e[A], nb = e[Bx];
  if (t >= (int64_t)na * nb) return;
  const int ia = (int)(t % na), ib = (int)(t / na);
  const int64_t sd = D == 0 ? 1 : (D == 1 ? (int64_t)ex : (int64_t)ex * ey);
  const int64_t sa = A == 0 ? 1 : (A == 1 ? (int64_t)ex : (int64_t)ex * ey);
  const int64_t sb = Bx == 0 ? 1 : (Bx == 1 ? (int64_t)ex : (int64_t)ex * ey);
  double *l = g + ia * sa + ib * sb;
  for (int q = lo - 1; q >= 0; --q) {
    if (kind == ASX_BC_LINEAR)
      l[q * sd] = l[lo * sd] + (lo - q) * (l[lo * sd] - l[(lo + 1) * sd]);
    else if (kind == ASX_BC_PERIODIC)
      l[q * sd] = l[(q + n) * sd];
    else
      l[q * sd] = l[lo * sd];
  }
  for (int q = hi + 1; q < e[D]; ++q) {
    if (kind == ASX_BC_LINEAR)
      l[q * sd] = l[hi * sd] + (q - hi) * (l[hi * sd] - l[(hi - 1) * sd]);
    else if (kind == ASX_BC_PERIODIC)
      l[q * sd] = l[(q - n) * sd];
    else
      l[q * sd] = l[hi * sd];
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_bc_many(BCList L, int n0, int n1, int n2, int ng,
                                                 int kind) {
  const int f = blockIdx.y;
  bc_point<D>(L.g[f], L.cent[f], n0, n1, n2, ng, kind, blockIdx.x * blockDim.x + threadIdx.x);
}

template <int D>
__global__ __launch_bounds__(256) void k_bc_batch(const KPatch *__restrict__ tab, BCSlots L,
                                                  const BatchRef *__restrict__ map, int nmap,
                                                  int kind) {
  const BatchRef r = map[xcd_tile((int)blockIdx.x, nmap)];
  const KPatch &P = tab[r.pid];
  const int f = blockIdx.y;
  bc_point<D>(P.g[L.slot[f]], L.cent[f], P.n[0], P.n[1], P.n[2], P.ng, kind,
              r.chunk * 256 + (int)threadIdx.x);
}

template <int ORD, class G_>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kUpdateWaves<ORD>)))
void k_update_all(KPatch P, Box B, G_ G, RhsArgs A, StatePtrs u0, StatePtrs out, Comb cb) {
  int i, j, k;
  if (!G.point(B, i, j, k)) return;
  update_point<ORD>(P, i, j, k, A, u0, out, cb);
}

template <int ORD, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kUpdateWaves<ORD>)))
void k_update_batch(const KPatch *__restrict__ tab, const StatePtrs *__restrict__ u0,
                    const StatePtrs *__restrict__ out, RhsArgs A, Comb cb,
                    const BatchBox *__restrict__ box, const BatchRef *__restrict__ map,
                    int nmap) {
  const BatchRef r = map[xcd_tile((int)blockIdx.x, nmap)];
  const BatchBox &bb = box[r.pid];
  int i, j, k;
  if (!tile_point<TX, TY, TZ>(bb.lo, bb.cnt, r.chunk, bb.nt[0], bb.nt[1], i, j, k)) return;
  update_point<ORD>(tab[r.pid], i, j, k, A, u0[r.pid], out[r.pid], cb);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static Box interior_box(const KPatch &P, int c, int halo = 0) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = P.ng - halo;
    b.cnt[d] = P.n[d] + ((c >> d) & 1) + 2 * halo;
  }
  return b;
}

hipError_t launch_sources(const KPatch &P, const Box &B, const RhsArgs &A, hipStream_t s) {
  if (A.fd_order == 4)
    hipLaunchKernelGGL(k_sources<4>, B.grid(), dim3(256), 0, s, P, B, A);
  else
    hipLaunchKernelGGL(k_sources<2>, B.grid(), dim3(256), 0, s, P, B, A);
  return hipGetLastError();
}

hipError_t launch_rhs(const KPatch &P, const RhsArgs &A, hipStream_t s) {
  Box b = interior_box(P, 0);
  hipLaunchKernelGGL(k_rhs_hydro, b.grid(), dim3(256), 0, s, P, b);
  b = interior_box(P, 6);
  hipLaunchKernelGGL(k_rhs_avec<0>, b.grid(), dim3(256), 0, s, P, b, A);
  b = interior_box(P, 5);
  hipLaunchKernelGGL(k_rhs_avec<1>, b.grid(), dim3(256), 0, s, P, b, A);
  b = interior_box(P, 3);
  hipLaunchKernelGGL(k_rhs_avec<2>, b.grid(), dim3(256), 0, s, P, b, A);
  b = interior_box(P, 7);
  hipLaunchKernelGGL(k_rhs_psi, b.grid(), dim3(256), 0, s, P, b, A);
  return hipGetLastError();
}

hipError_t launch_lincomb(double *dst, int64_t np, double scale, int nsrc,
                          const double *fac, double *const *src, hipStream_t s) {
  LinArgs L;
  L.nsrc = nsrc;
  L.scale = scale;
  for (int q = 0; q < 8; ++q) {
    L.fac[q] = q < nsrc ? fac[q] : 0.0;
    L.src[q] = q < nsrc ? src[q] : nullptr;
  }
  hipLaunchKernelGGL(k_lincomb, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, dst, np, L);
  return hipGetLastError();
}

hipError_t launch_bc(const KPatch &P, int slot, int kind, hipStream_t s) {
  const int c = asx_gf_centering(slot);
  int e[3];
  for (int d = 0; d < 3; ++d) e[d] = P.n[d] + 2 * P.ng + ((c >> d) & 1);
  for (int d = 0; d < 3; ++d) {
    const int lo = P.ng, hi = P.ng + P.n[d] - 1 + ((c >> d) & 1);
    const int64_t lines = (int64_t)e[(d + 1) % 3] * e[(d + 2) % 3];
    const dim3 grid((unsigned)((lines + 255) / 256));
    if (d == 0) hipLaunchKernelGGL(k_bc<0>, grid, dim3(256), 0, s, P.g[slot], e[0], e[1], e[2], lo, hi, P.n[0], kind);
    if (d == 1) hipLaunchKernelGGL(k_bc<1>, grid, dim3(256), 0, s, P.g[slot], e[0], e[1], e[2], lo, hi, P.n[1], kind);
    if (d == 2) hipLaunchKernelGGL(k_bc<2>, grid, dim3(256), 0, s, P.g[slot], e[0], e[1], e[2], lo, hi, P.n[2], kind);
  }
  return hipGetLastError();
}

hipError_t launch_bc_many(const KPatch &P, const BCList &L, int kind, hipStream_t s) {
  if (L.nf <= 0) return hipSuccess;
  auto grid = [&](int d) {
    return dim3((unsigned)((bc_points(P.n, P.ng, d) + 255) / 256), (unsigned)L.nf);
  };
  hipLaunchKernelGGL(k_bc_many<0>, grid(0), dim3(256), 0, s, L, P.n[0], P.n[1], P.n[2], P.ng, kind);
  hipLaunchKernelGGL(k_bc_many<1>, grid(1), dim3(256), 0, s, L, P.n[0], P.n[1], P.n[2], P.ng, kind);
  hipLaunchKernelGGL(k_bc_many<2>, grid(2), dim3(256), 0, s, L, P.n[0], P.n[1], P.n[2], P.ng, kind);
  return hipGetLastError();
}

template <int ORD, int TX, int TY, int TZ>
static void launch_update_t(const KPatch &P, const Box &b, const RhsArgs &A,
                            const StatePtrs &u0, const StatePtrs &out, const Comb &cb,
                            hipStream_t s) {
  using G = TileGrid<TX, TY, TZ>;
  static_assert(G::kThreads == 256, "update tile");
  const G g = G::make(b);
  hipLaunchKernelGGL((k_update_all<ORD, G>), g.grid(), dim3(256), 0, s, P, b, g, A, u0, out, cb);
}
template <int ORD, int TX, int TY, int TZ, int YS>
static void launch_update_band_t(const KPatch &P, const Box &b, const RhsArgs &A,
                                 const StatePtrs &u0, const StatePtrs &out, const Comb &cb,
                                 hipStream_t s) {
  using G = TileGridBand<TX, TY, TZ, YS>;
  static_assert(G::kThreads == 256, "update tile");
  const G g = G::make(b);
  hipLaunchKernelGGL((k_update_all<ORD, G>), g.grid(), dim3(256), 0, s, P, b, g, A, u0, out, cb);
}
template <int ORD>
static void launch_update_v(int v, const KPatch &P, const Box &b, const RhsArgs &A,
                            const StatePtrs &u0, const StatePtrs &out, const Comb &cb,
                            hipStream_t s) {
  switch (v) {
  case 1: launch_update_t<ORD, 32, 4, 2>(P, b, A, u0, out, cb, s); break;
  case 2: launch_update_t<ORD, 32, 2, 4>(P, b, A, u0, out, cb, s); break;
  case 3: launch_update_t<ORD, 16, 4, 4>(P, b, A, u0, out, cb, s); break;
  case 4: launch_update_t<ORD, 64, 2, 2>(P, b, A, u0, out, cb, s); break;
  case 5: launch_update_t<ORD, 64, 4, 1>(P, b, A, u0, out, cb, s); break;
  case 6: launch_update_band_t<ORD, 32, 4, 2, 2>(P, b, A, u0, out, cb, s); break;
  // default: 32x2x4 tiles in bands of 4 tile rows (1.465 ms on the 192^3
  // torus against 1.481 ms for the same tiles in plain x, y, z order)
  default: launch_update_band_t<ORD, 32, 2, 4, 4>(P, b, A, u0, out, cb, s); break;
  }
}

template <int TX, int TY, int TZ> static void ext3(int ex[3]) {
  ex[0] = TX;
  ex[1] = TY;
  ex[2] = TZ;
}
void update_tile_extent(int variant, int ex[3]) {
  switch (variant) {
  case 1: ext3<32, 4, 2>(ex); break;
  case 2: ext3<32, 2, 4>(ex); break;
  case 3: ext3<16, 4, 4>(ex); break;
  case 4: ext3<64, 2, 2>(ex); break;
  case 5: ext3<64, 4, 1>(ex); break;
  case 6: ext3<32, 4, 2>(ex); break;
  default: ext3<32, 2, 4>(ex); break;
  }
}

template <int ORD, int TX, int TY, int TZ>
static void launch_update_batch_t(const KPatch *tab, const StatePtrs *u0, const StatePtrs *out,
                                  const RhsArgs &A, const Comb &cb, const BatchBox *box,
                                  const BatchRef *map, int nmap, hipStream_t s) {
  static_assert(TX * TY * TZ == 256, "update tile");
  hipLaunchKernelGGL((k_update_batch<ORD, TX, TY, TZ>), dim3((unsigned)nmap), dim3(256), 0, s,
                     tab, u0, out, A, cb, box, map, nmap);
}
template <int ORD>
static void launch_update_batch_v(int v, const KPatch *tab, const StatePtrs *u0,
                                  const StatePtrs *out, const RhsArgs &A, const Comb &cb,
                                  const BatchBox *box, const BatchRef *map, int nmap,
                                  hipStream_t s) {
  switch (v) {
  case 1: launch_update_batch_t<ORD, 32, 4, 2>(tab, u0, out, A, cb, box, map, nmap, s); break;
  case 2: launch_update_batch_t<ORD, 32, 2, 4>(tab, u0, out, A, cb, box, map, nmap, s); break;
  case 3: launch_update_batch_t<ORD, 16, 4, 4>(tab, u0, out, A, cb, box, map, nmap, s); break;
  case 4: launch_update_batch_t<ORD, 64, 2, 2>(tab, u0, out, A, cb, box, map, nmap, s); break;
  case 5: launch_update_batch_t<ORD, 64, 4, 1>(tab, u0, out, A, cb, box, map, nmap, s); break;
  // the band orders of 0 and 6 concern one large box; a batch keeps their
  // tile shapes in its own (patch, tile) order
  case 6: launch_update_batch_t<ORD, 32, 4, 2>(tab, u0, out, A, cb, box, map, nmap, s); break;
  default: launch_update_batch_t<ORD, 32, 2, 4>(tab, u0, out, A, cb, box, map, nmap, s); break;
  }
}

hipError_t launch_update_batch(const KPatch *tab, const StatePtrs *u0, const StatePtrs *out,
                               const RhsArgs &A, double c0, double c1, double c2,
                               const BatchBox *box, const BatchRef *map, int nmap,
                               int variant, hipStream_t s) {
  if (nmap <= 0) return hipSuccess;
  const Comb cb{c0, c1, c2};
  if (A.fd_order == 4)
    launch_update_batch_v<4>(variant, tab, u0, out, A, cb, box, map, nmap, s);
  else
    launch_update_batch_v<2>(variant, tab, u0, out, A, cb, box, map, nmap, s);
  return hipGetLastError();
}

hipError_t launch_bc_batch(int d, const KPatch *tab, const BCSlots &L, const BatchRef *map,
                           int nmap, int kind, hipStream_t s) {
  if (nmap <= 0 || L.nf <= 0) return hipSuccess;
  const dim3 grid((unsigned)nmap, (unsigned)L.nf);
  if (d == 0) hipLaunchKernelGGL(k_bc_batch<0>, grid, dim3(256), 0, s, tab, L, map, nmap, kind);
  else if (d == 1) hipLaunchKernelGGL(k_bc_batch<1>, grid, dim3(256), 0, s, tab, L, map, nmap, kind);
  else hipLaunchKernelGGL(k_bc_batch<2>, grid, dim3(256), 0, s, tab, L, map, nmap, kind);
  return hipGetLastError();
}

hipError_t launch_update(const KPatch &P, const RhsArgs &A, const StatePtrs &u0,
                         const StatePtrs &out, double c0, double c1, double c2,
                         int variant, hipStream_t s) {
  const Comb cb{c0, c1, c2};
  const Box b = interior_box(P, 7);
  if (A.fd_order == 4)
    launch_update_v<4>(variant, P, b, A, u0, out, cb, s);
  else
    launch_update_v<2>(variant, P, b, A, u0, out, cb, s);
  return hipGetLastError();
}

} // namespace asx
