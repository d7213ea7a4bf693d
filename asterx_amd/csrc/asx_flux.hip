// Face, vertex and curl kernels of AsterX_Fluxes and the B-field group on
// gfx950.  Vertex and curl kernels: one work-item per point, x-fastest so a
// wavefront reads 64 consecutive doubles of every stencil row.  Face
// kernels: cell tiles with an LDS hand-over between neighbouring cells.
#include <algorithm>

#include "asx_face.hpp"

namespace asx {

template <int D, int M, int V, bool TAB = false>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(TAB ? ASX_FACE_WAVES_TAB : ASX_FACE_WAVES)))
void k_face(KPatch P, FluxArgs A, int ntx, int nty, int ntz) {
  __shared__ FaceLDS S;
  using MB = MetricBox<D, V>;
  __shared__ double MS[MB::ok ? MB::N : 1];
  // workgroup u takes face tile u and, when the launch carries the vertex
  // auxiliaries, their 256-vertex chunk u: the aux loads and stores then
  // overlap the face arithmetic instead of running as a launch of their own
  const int u = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (u < ntx * nty * ntz)
    face_tile<D, M, V, TAB>(P, A, A.lo, A.cnt, u, ntx, nty, S, MB::ok ? MS : nullptr);
  if (A.with_aux) {
    const int64_t t = (int64_t)u * 256 + threadIdx.x;
    if (t < A.aux.total()) {
      int i, j, k;
      A.aux.ijk(t, i, j, k);
      aux_point(P, i, j, k);
    }
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_bstag(KPatch P, Box B, double *out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  bstag_point<D>(P, i, j, k, out);
}

// AsterX_ComputedBFromdBstag (computeBfromA.cxx:71-84)
__global__ __launch_bounds__(256) void k_dB(KPatch P, Box B) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  const int oc = off(P, 0, i, j, k);
  const int o1 = off(P, 1, i, j, k), o2 = off(P, 2, i, j, k), o4 = off(P, 4, i, j, k);
  P.g[ASX_DBX][oc] = 0.5 * (P.g[ASX_DBX_STAG][o1] + P.g[ASX_DBX_STAG][o1 + 1]);
  P.g[ASX_DBY][oc] = 0.5 * (P.g[ASX_DBY_STAG][o2] + P.g[ASX_DBY_STAG][o2 + P.sj[2]]);
  P.g[ASX_DBZ][oc] = 0.5 * (P.g[ASX_DBZ_STAG][o4] + P.g[ASX_DBZ_STAG][o4 + P.sk[4]]);
}

// AsterX_ComputeBFromdB (computeBfromA.cxx:91-112)
__global__ __launch_bounds__(256) void k_B_from_dB(KPatch P, Box B) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  const int oc = off(P, 0, i, j, k), ov = off(P, 7, i, j, k);
  const int sj = P.sj[7], sk = P.sk[7];
  const Sym3 g{avg_v2c(P.g[ASX_GXX], ov, sj, sk), avg_v2c(P.g[ASX_GXY], ov, sj, sk),
               avg_v2c(P.g[ASX_GXZ], ov, sj, sk), avg_v2c(P.g[ASX_GYY], ov, sj, sk),
               avg_v2c(P.g[ASX_GYZ], ov, sj, sk), avg_v2c(P.g[ASX_GZZ], ov, sj, sk)};
  const double sq = sqrt(det3(g));
  P.g[ASX_BVECX][oc] = P.g[ASX_DBX][oc] / sq;
  P.g[ASX_BVECY][oc] = P.g[ASX_DBY][oc] / sq;
  P.g[ASX_BVECZ][oc] = P.g[ASX_DBZ][oc] / sq;
}

// AsterX_Prim2Con_Initial (AsterX/src/prim2con.cxx:12-52, prim2con.hxx:38-83)
__global__ __launch_bounds__(256) void k_prim2con(KPatch P, Box B) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  const int oc = off(P, 0, i, j, k), ov = off(P, 7, i, j, k);
  const int sj = P.sj[7], sk = P.sk[7];
  const Sym3 g{avg_v2c(P.g[ASX_GXX], ov, sj, sk), avg_v2c(P.g[ASX_GXY], ov, sj, sk),
               avg_v2c(P.g[ASX_GXZ], ov, sj, sk), avg_v2c(P.g[ASX_GYY], ov, sj, sk),
               avg_v2c(P.g[ASX_GYZ], ov, sj, sk), avg_v2c(P.g[ASX_GZZ], ov, sj, sk)};
  const double vel[3] = {P.g[ASX_VELX][oc], P.g[ASX_VELY][oc], P.g[ASX_VELZ][oc]};
  const double Bv[3] = {P.g[ASX_BVECX][oc], P.g[ASX_BVECY][oc], P.g[ASX_BVECZ][oc]};
  double vl[3];
  lower(g, vel, vl);
  const double w = 1.0 / sqrt(1.0 - dot3(vl, vel));
  Cons cv;
  cons_from_prim(g, P.g[ASX_RHO][oc], P.g[ASX_EPS][oc], P.g[ASX_PRESS][oc], vel, w,
                 Bv, P.g[ASX_YE][oc], P.g[ASX_ENTROPY][oc], cv);
  P.g[ASX_DENS][oc] = cv.dens;
  P.g[ASX_MOMX][oc] = cv.mom[0];
  P.g[ASX_MOMY][oc] = cv.mom[1];
  P.g[ASX_MOMZ][oc] = cv.mom[2];
  P.g[ASX_TAU][oc] = cv.tau;
  P.g[ASX_DYE][oc] = cv.DYe;
  P.g[ASX_DBX][oc] = cv.dB[0];
  P.g[ASX_DBY][oc] = cv.dB[1];
  P.g[ASX_DBZ][oc] = cv.dB[2];
  P.g[ASX_DENT][oc] = cv.DEnt;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int D, int V>
static hipError_t launch_flux_dv(const KPatch &P, const FluxArgs &A, int method,
                                 hipStream_t s, bool tab) {
  using T = FaceTile<D, V>;
  if ((int64_t)A.cnt[0] * A.cnt[1] * A.cnt[2] <= 0) return hipSuccess;
  const int ex[3] = {T::X, T::Y, T::Z};
  int nt[3];
  for (int q = 0; q < 3; ++q) {
    const int per = (q == D) ? ex[q] - 1 : ex[q];
    nt[q] = (A.cnt[q] + per - 1) / per;
  }
  int64_t nb = (int64_t)nt[0] * nt[1] * nt[2];
  if (A.with_aux) nb = std::max<int64_t>(nb, (A.aux.htotal() + 255) / 256);
  const dim3 grid((unsigned)nb), block(256);
  if (tab) {
    switch (method) {
    case ASX_RECON_PPM: hipLaunchKernelGGL((k_face<D, ASX_RECON_PPM, V, true>), grid, block, 0, s, P, A, nt[0], nt[1], nt[2]); break;
    case ASX_RECON_WENOZ: hipLaunchKernelGGL((k_face<D, ASX_RECON_WENOZ, V, true>), grid, block, 0, s, P, A, nt[0], nt[1], nt[2]); break;
    case ASX_RECON_MP5: hipLaunchKernelGGL((k_face<D, ASX_RECON_MP5, V, true>), grid, block, 0, s, P, A, nt[0], nt[1], nt[2]); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  switch (method) {
  case ASX_RECON_GODUNOV: hipLaunchKernelGGL((k_face<D, ASX_RECON_GODUNOV, V>), grid, block, 0, s, P, A, nt[0], nt[1], nt[2]); break;
  case ASX_RECON_MINMOD: hipLaunchKernelGGL((k_face<D, ASX_RECON_MINMOD, V>), grid, block, 0, s, P, A, nt[0], nt[1], nt[2]); break;
  case ASX_RECON_MONOCENTRAL: hipLaunchKernelGGL((k_face<D, ASX_RECON_MONOCENTRAL, V>), grid, block, 0, s, P, A, nt[0], nt[1], nt[2]); break;
  case ASX_RECON_PPM: hipLaunchKernelGGL((k_face<D, ASX_RECON_PPM, V>), grid, block, 0, s, P, A, nt[0], nt[1], nt[2]); break;
  case ASX_RECON_WENOZ: hipLaunchKernelGGL((k_face<D, ASX_RECON_WENOZ, V>), grid, block, 0, s, P, A, nt[0], nt[1], nt[2]); break;
  default: hipLaunchKernelGGL((k_face<D, ASX_RECON_MP5, V>), grid, block, 0, s, P, A, nt[0], nt[1], nt[2]); break;
  }
  return hipGetLastError();
}
template <int D>
static hipError_t launch_flux_d(const KPatch &P, const FluxArgs &A, int method, int variant,
                                hipStream_t s, bool tab) {
  switch (variant) {
  case 1: return launch_flux_dv<D, 1>(P, A, method, s, tab);
  case 2: return launch_flux_dv<D, 2>(P, A, method, s, tab);
  case 3: return launch_flux_dv<D, 3>(P, A, method, s, tab);
  case 4: return launch_flux_dv<D, 4>(P, A, method, s, tab);
  case 5: return launch_flux_dv<D, 5>(P, A, method, s, tab);
  case 6: return launch_flux_dv<D, 6>(P, A, method, s, tab);
  case 7: return launch_flux_dv<D, 7>(P, A, method, s, tab);
  default: return launch_flux_dv<D, 0>(P, A, method, s, tab);
  }
}

bool flux_tab_method(int method) {
  return method == ASX_RECON_PPM || method == ASX_RECON_WENOZ || method == ASX_RECON_MP5;
}

hipError_t launch_flux(int d, const KPatch &P, const FluxArgs &A, int method, int variant,
                       hipStream_t s, bool tab) {
  if (d == 0) return launch_flux_d<0>(P, A, method, variant, s, tab);
  if (d == 1) return launch_flux_d<1>(P, A, method, variant, s, tab);
  return launch_flux_d<2>(P, A, method, variant, s, tab);
}

hipError_t launch_bstag(int d, const KPatch &P, const Box &B, double *out,
                        hipStream_t s) {
  if (B.htotal() <= 0) return hipSuccess;
  if (d == 0) hipLaunchKernelGGL(k_bstag<0>, B.grid(), dim3(256), 0, s, P, B, out);
  else if (d == 1) hipLaunchKernelGGL(k_bstag<1>, B.grid(), dim3(256), 0, s, P, B, out);
  else hipLaunchKernelGGL(k_bstag<2>, B.grid(), dim3(256), 0, s, P, B, out);
  return hipGetLastError();
}

hipError_t launch_dB(const KPatch &P, const Box &B, hipStream_t s) {
  hipLaunchKernelGGL(k_dB, B.grid(), dim3(256), 0, s, P, B);
  return hipGetLastError();
}
hipError_t launch_B_from_dB(const KPatch &P, const Box &B, hipStream_t s) {
  hipLaunchKernelGGL(k_B_from_dB, B.grid(), dim3(256), 0, s, P, B);
  return hipGetLastError();
}
hipError_t launch_prim2con(const KPatch &P, const Box &B, hipStream_t s) {
  hipLaunchKernelGGL(k_prim2con, B.grid(), dim3(256), 0, s, P, B);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched forms: one launch over the listed (patch, chunk) workgroups of
// many patches, per-patch views and boxes from device tables.
// ---------------------------------------------------------------------------
template <int D, int M, int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ASX_FACE_WAVES)))
void k_face_batch(const KPatch *__restrict__ tab, FluxArgs A, const BatchBox *__restrict__ box,
                  const BatchRef *__restrict__ map, int nmap) {
  const BatchRef r = map[xcd_tile((int)blockIdx.x, nmap)];
  const BatchBox &bb = box[r.pid];
  __shared__ FaceLDS S;
  using MB = MetricBox<D, V>;
  __shared__ double MS[MB::ok ? MB::N : 1];
  face_tile<D, M, V>(tab[r.pid], A, bb.lo, bb.cnt, r.chunk, bb.nt[0], bb.nt[1], S,
                     MB::ok ? MS : nullptr);
}

__global__ __launch_bounds__(256) void k_aux_batch(const KPatch *__restrict__ tab,
                                                   const BatchBox *__restrict__ box,
                                                   const BatchRef *__restrict__ map, int nmap) {
  const BatchRef r = map[xcd_tile((int)blockIdx.x, nmap)];
  int i, j, k;
  if (batch_point(box[r.pid], r.chunk, i, j, k)) aux_point(tab[r.pid], i, j, k);
}

template <int D>
__global__ __launch_bounds__(256) void k_bstag_batch(const KPatch *__restrict__ tab,
                                                     const BatchBox *__restrict__ box,
                                                     const BatchRef *__restrict__ map, int nmap) {
  const BatchRef r = map[xcd_tile((int)blockIdx.x, nmap)];
  int i, j, k;
  if (batch_point(box[r.pid], r.chunk, i, j, k)) {
    const KPatch &P = tab[r.pid];
    bstag_point<D>(P, i, j, k, P.g[ASX_DBX_STAG + D]);
  }
}

template <int D, int V>
static void launch_face_batch_dv(const KPatch *tab, const FluxArgs &A, const BatchBox *box,
                                 const BatchRef *map, int nmap, int method, hipStream_t s) {
  const dim3 grid((unsigned)nmap), block(256);
  switch (method) {
  case ASX_RECON_GODUNOV: hipLaunchKernelGGL((k_face_batch<D, ASX_RECON_GODUNOV, V>), grid, block, 0, s, tab, A, box, map, nmap); break;
  case ASX_RECON_MINMOD: hipLaunchKernelGGL((k_face_batch<D, ASX_RECON_MINMOD, V>), grid, block, 0, s, tab, A, box, map, nmap); break;
  case ASX_RECON_MONOCENTRAL: hipLaunchKernelGGL((k_face_batch<D, ASX_RECON_MONOCENTRAL, V>), grid, block, 0, s, tab, A, box, map, nmap); break;
  case ASX_RECON_PPM: hipLaunchKernelGGL((k_face_batch<D, ASX_RECON_PPM, V>), grid, block, 0, s, tab, A, box, map, nmap); break;
  case ASX_RECON_WENOZ: hipLaunchKernelGGL((k_face_batch<D, ASX_RECON_WENOZ, V>), grid, block, 0, s, tab, A, box, map, nmap); break;
  default: hipLaunchKernelGGL((k_face_batch<D, ASX_RECON_MP5, V>), grid, block, 0, s, tab, A, box, map, nmap); break;
  }
}
template <int D>
static void launch_face_batch_d(const KPatch *tab, const FluxArgs &A, const BatchBox *box,
                                const BatchRef *map, int nmap, int method, int variant,
                                hipStream_t s) {
  switch (variant) {
  case 1: launch_face_batch_dv<D, 1>(tab, A, box, map, nmap, method, s); break;
  case 2: launch_face_batch_dv<D, 2>(tab, A, box, map, nmap, method, s); break;
  case 3: launch_face_batch_dv<D, 3>(tab, A, box, map, nmap, method, s); break;
  case 4: launch_face_batch_dv<D, 4>(tab, A, box, map, nmap, method, s); break;
  case 5: launch_face_batch_dv<D, 5>(tab, A, box, map, nmap, method, s); break;
  case 6: launch_face_batch_dv<D, 6>(tab, A, box, map, nmap, method, s); break;
  case 7: launch_face_batch_dv<D, 7>(tab, A, box, map, nmap, method, s); break;
  default: launch_face_batch_dv<D, 0>(tab, A, box, map, nmap, method, s); break;
  }
}

template <int D, int V> static void face_ext(int ex[3]) {
  using T = FaceTile<D, V>;
  ex[0] = T::X;
  ex[1] = T::Y;
  ex[2] = T::Z;
}
template <int D> static void face_ext_d(int variant, int ex[3]) {
  switch (variant) {
  case 1: face_ext<D, 1>(ex); break;
  case 2: face_ext<D, 2>(ex); break;
  case 3: face_ext<D, 3>(ex); break;
  case 4: face_ext<D, 4>(ex); break;
  case 5: face_ext<D, 5>(ex); break;
  case 6: face_ext<D, 6>(ex); break;
  case 7: face_ext<D, 7>(ex); break;
  default: face_ext<D, 0>(ex); break;
  }
}
void face_tile_extent(int d, int variant, int ex[3]) {
  if (d == 0) face_ext_d<0>(variant, ex);
  else if (d == 1) face_ext_d<1>(variant, ex);
  else face_ext_d<2>(variant, ex);
}

hipError_t launch_face_batch(int d, const KPatch *tab, const FluxArgs &A, const BatchBox *box,
                             const BatchRef *map, int nmap, int method, int variant,
                             hipStream_t s) {
  if (nmap <= 0) return hipSuccess;
  if (d == 0) launch_face_batch_d<0>(tab, A, box, map, nmap, method, variant, s);
  else if (d == 1) launch_face_batch_d<1>(tab, A, box, map, nmap, method, variant, s);
  else launch_face_batch_d<2>(tab, A, box, map, nmap, method, variant, s);
  return hipGetLastError();
}
hipError_t launch_aux_batch(const KPatch *tab, const BatchBox *box, const BatchRef *map,
                            int nmap, hipStream_t s) {
  if (nmap <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_aux_batch, dim3((unsigned)nmap), dim3(256), 0, s, tab, box, map, nmap);
  return hipGetLastError();
}
hipError_t launch_bstag_batch(int d, const KPatch *tab, const BatchBox *box,
                              const BatchRef *map, int nmap, hipStream_t s) {
  if (nmap <= 0) return hipSuccess;
  const dim3 grid((unsigned)nmap), block(256);
  if (d == 0) hipLaunchKernelGGL(k_bstag_batch<0>, grid, block, 0, s, tab, box, map, nmap);
  else if (d == 1) hipLaunchKernelGGL(k_bstag_batch<1>, grid, block, 0, s, tab, box, map, nmap);
  else hipLaunchKernelGGL(k_bstag_batch<2>, grid, block, 0, s, tab, box, map, nmap);
  return hipGetLastError();
}

} // namespace asx
