// Vertex-centred stress-energy tensor of the fluid and the magnetic field,
// added into eTtt..eTzz (AsterX/src/tmunu.cxx:19-176): the primitives are
// interpolated from cells to the vertex (calc_avg_c2v,
// AsterUtils/src/aster_interp.hxx:97-148), then T_{mu nu} is formed with the
// vertex metric.
//
// Work decomposition: a workgroup owns a tile of TX x TY vertex columns
// and a run of TZ vertices along z, and marches through the cell planes
// the run needs.  Each plane of the 9 interpolated variables is staged in
// LDS once per tile (coalesced rows, each cell fetched once), every
// work-item reduces its (ORD x ORD) window of the plane in x and y from LDS,
// and keeps the last ORD reduced planes of each variable in registers, so
// the z reduction costs nothing extra.  The global loads of the next plane
// are issued before the reductions of the current one.
#include "asx_kernels.hpp"

namespace asx {
namespace {

constexpr int kTmunuVars = 9;  // rho, eps, press, B^i, v^i (or z^i)

// Cell-to-vertex weights of calc_avg_c2v<ORD> along one axis, on the cells
// I - ORD/2 .. I + ORD/2 - 1: (-1, 9, 9, -1)/16 for order 4; order 2 is the
// mean of the 8 surrounding cells, 1/2 per axis.
template <int ORD> ASX_DEV constexpr double c2v_w(int a) {
  return ORD == 2 ? 0.5 : ((a == 0 || a == 3) ? -1.0 / 16.0 : 9.0 / 16.0);
}

// Buffer descriptor word 3 for raw (untyped) buffer loads on gfx9-family GPUs.
constexpr int kBufferFlags = 0x00020000;

// Tiles of TX x TY vertex columns (TX * TY = 256 work-items) and runs of
// TZ vertices along z; the staged plane has (TX + ORD - 1) x (TY + ORD - 1)
// cells per variable.
template <int ORD, int TX, int TY> struct TmTile {
  static_assert(TX * TY == 256, "one work-item per vertex column");
  static constexpr int H = ORD / 2;
  static constexpr int CX = TX + ORD - 1, CY = TY + ORD - 1, NC = CX * CY;
  static constexpr int PER = (NC + 255) / 256; // cells per work-item and variable
};

// Element at byte offset ob of a grid function: a scalar base plus one
// 32-bit vector offset shared by every grid function of the same centering.
ASX_DEV double &gf_at(double *g, unsigned ob) { return *(double *)((char *)g + ob); }

// T_{mu nu} at the vertex at byte offset ob from the interpolated primitives
// (AsterX/src/tmunu.cxx:56-158), added into the ten eT slots.
ASX_DEV void tmunu_point(const KPatch &P, bool use_z, unsigned ob,
                         const double q[kTmunuVars]) {
  const double rho = q[0], eps = q[1], press = q[2];
  const double B[3] = {q[3], q[4], q[5]};
  const Sym3 g{gf_at(P.g[ASX_GXX], ob), gf_at(P.g[ASX_GXY], ob), gf_at(P.g[ASX_GXZ], ob),
               gf_at(P.g[ASX_GYY], ob), gf_at(P.g[ASX_GYZ], ob), gf_at(P.g[ASX_GZZ], ob)};
  const double betau[3] = {gf_at(P.g[ASX_BETAX], ob), gf_at(P.g[ASX_BETAY], ob),
                           gf_at(P.g[ASX_BETAZ], ob)};
  const double alp = gf_at(P.g[ASX_ALP], ob);
  double betal[3];
  lower(g, betau, betal);
  const double beta2 = dot3(betal, betau);

  double vu[3], vl[3], w;
  if (!use_z) {
    vu[0] = q[6], vu[1] = q[7], vu[2] = q[8];
    lower(g, vu, vl);
    w = 1.0 / sqrt(1.0 - dot3(vu, vl));
  } else {
    const double zu[3] = {q[6], q[7], q[8]};
    double zl[3];
    lower(g, zu, zl);
    w = sqrt(1.0 + dot3(zu, zl));
#pragma unroll
    for (int d = 0; d < 3; ++d) vu[d] = zu[d] / w, vl[d] = zl[d] / w;
  }

  const double rhw2 = (rho * (1.0 + eps) + press) * w * w;
  const double utl = -alp + dot3(betal, vu);
  double uiu[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) uiu[d] = vu[d] - betau[d] / alp;
  const double btu = w * dot3(B, vl) / alp;
  double bsu[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) bsu[d] = (B[d] + alp * btu * w * uiu[d]) / w;
  const double g00 = -alp * alp + beta2;
  const double btl = btu * g00 + dot3(bsu, betal);
  double gb[3], bsl[3];
  lower(g, bsu, gb);
#pragma unroll
  for (int d = 0; d < 3; ++d) bsl[d] = btu * betal[d] + gb[d];
  const double b2 = btu * btl + dot3(bsu, bsl);
  const double w2 = w * w;

  gf_at(P.g[ASX_ETTT], ob) += rhw2 * utl * utl + press * g00 + (w2 * utl * utl + 0.5 * g00) * b2 -
                       btl * btl;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    gf_at(P.g[ASX_ETTX + d], ob) += rhw2 * utl * vl[d] + press * betal[d] +
                             (w2 * utl * vl[d] + 0.5 * betal[d]) * b2 - btl * bsl[d];
  int s = ASX_ETXX;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b, ++s)
      gf_at(P.g[s], ob) += rhw2 * vl[a] * vl[b] + press * g(a, b) +
                    (w2 * vl[a] * vl[b] + 0.5 * g(a, b)) * b2 - bsl[a] * bsl[b];
}

// Tile b of an ntx x nty x ntz grid.  YS = 0: x fastest, then y, then z.
// YS > 0: bands of YS tile rows (x, the band's rows, then z, then the next
// band), so the runs above and below a tile, which stage ORD - 1 of the same
// planes, are YS * ntx launches away rather than a whole x-y layer of tiles.
template <int YS> ASX_DEV void tile_of(int b, int ntx, int nty, int ntz, int &bx, int &by,
                                       int &bz) {
  if (YS == 0) {
    bx = b % ntx, by = (b / ntx) % nty, bz = b / (ntx * nty);
    return;
  }
  const int full = nty / YS, blk = YS * ntz * ntx;
  if (b < full * blk) {
    const int r = b % blk;
    bz = r / (YS * ntx);
    by = (b / blk) * YS + (r % (YS * ntx)) / ntx;
    bx = r % ntx;
  } else {
    const int rem = nty - full * YS, b2 = b - full * blk;
    bz = b2 / (rem * ntx);
    by = full * YS + (b2 % (rem * ntx)) / ntx;
    bx = b2 % ntx;
  }
}

template <int ORD, int TX, int TY, int TZ, int WAVES, int YS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_tmunu(
    KPatch P, int ntx, int nty, int nb, int use_z) {
  using T = TmTile<ORD, TX, TY>;
  constexpr int H = T::H, CX = T::CX, NC = T::NC, PER = T::PER;
  __shared__ double lds[kTmunuVars][NC];
  const int b = xcd_tile((int)blockIdx.x, nb);
  int bx, by, bz;
  tile_of<YS>(b, ntx, nty, nb / (ntx * nty), bx, by, bz);
  const int tid = (int)threadIdx.x, tx = tid % TX, ty = tid / TX;
  const int i0 = P.ng + bx * TX, j0 = P.ng + by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int kb = P.ng + bz * TZ;
  const int ke = min(kb + TZ, P.ng + P.n[2] + 1);
  const bool mine = i <= P.ng + P.n[0] && j <= P.ng + P.n[1];
  // last cell any interior vertex's stencil reads, per axis
  const int cxmax = P.ng + P.n[0] + H - 1, cymax = P.ng + P.n[1] + H - 1;
  const int sj = P.sj[0], sk = P.sk[0];
  const double *src[kTmunuVars] = {
      P.g[ASX_RHO],   P.g[ASX_EPS],   P.g[ASX_PRESS],
      P.g[ASX_BVECX], P.g[ASX_BVECY], P.g[ASX_BVECZ],
      P.g[use_z ? ASX_ZVEC_X : ASX_VELX], P.g[use_z ? ASX_ZVEC_Y : ASX_VELY],
      P.g[use_z ? ASX_ZVEC_Z : ASX_VELZ]};

  // this work-item's cells of a plane (LDS slot tid + 256 r): byte offset
  // in the plane; cells past the tile or past the last cell any interior
  // vertex reads are fetched from the plane's first cell instead (in bounds,
  // never used)
  unsigned lofs[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int c = tid + 256 * r;
    const int ci = i0 - H + c % CX, cj = j0 - H + c / CX;
    lofs[r] = (c < NC && ci <= cxmax && cj <= cymax) ? 8u * (unsigned)(ci + sj * cj) : 0u;
  }
  // buffer resources of the 9 source grid functions: one scalar descriptor
  // each, so a fetch is one vector offset plus the plane's scalar offset
  __amdgpu_buffer_rsrc_t rs[kTmunuVars];
#pragma unroll
  for (int q = 0; q < kTmunuVars; ++q)
    rs[q] = __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(src[q]), (short)0,
                                              (int)(8u * (unsigned)sk * (unsigned)P.n[2] +
                                                    16u * (unsigned)sk * (unsigned)P.ng),
                                              kBufferFlags);
  double nxt[kTmunuVars][PER];
  auto fetch = [&](int kc) {
    const int pofs = (int)(8u * (unsigned)sk * (unsigned)kc);
#pragma unroll
    for (int q = 0; q < kTmunuVars; ++q)
#pragma unroll
      for (int r = 0; r < PER; ++r)
        nxt[q][r] = __builtin_bit_cast(
            double, __builtin_amdgcn_raw_buffer_load_b64(rs[q], lofs[r], pofs, 0));
  };

  const int np = ke - kb + ORD - 1; // planes kb-H .. ke+H-2
  double win[kTmunuVars][ORD];
  fetch(kb - H);
#pragma unroll 1
  for (int p = 0; p < np; ++p) {
    __syncthreads(); // the previous plane's LDS reads are done
#pragma unroll
    for (int q = 0; q < kTmunuVars; ++q)
#pragma unroll
      for (int r = 0; r < PER; ++r)
        if (tid + 256 * r < NC) lds[q][tid + 256 * r] = nxt[q][r];
    __syncthreads();
    double v[kTmunuVars];
#pragma unroll
    for (int q = 0; q < kTmunuVars; ++q) {
      double s = 0.0;
#pragma unroll
      for (int bb = 0; bb < ORD; ++bb) {
        double r = 0.0;
#pragma unroll
        for (int a = 0; a < ORD; ++a) r += c2v_w<ORD>(a) * lds[q][(ty + bb) * CX + tx + a];
        s += c2v_w<ORD>(bb) * r;
      }
#pragma unroll
      for (int c = 0; c < ORD - 1; ++c) win[q][c] = win[q][c + 1];
      win[q][ORD - 1] = s;
      double z = 0.0;
#pragma unroll
      for (int c = 0; c < ORD; ++c) z += c2v_w<ORD>(c) * win[q][c];
      v[q] = z;
    }
    if (p >= ORD - 1 && mine) {
      unsigned ob = 8u * (unsigned)(i + P.sj[7] * j + P.sk[7] * (kb + p - (ORD - 1)));
      // opaque to loop strength reduction: every vertex access stays a
      // scalar base plus this one offset
      asm volatile("" : "+v"(ob));
      tmunu_point(P, use_z != 0, ob, v);
    }
    if (p + 1 < np) fetch(kb - H + p + 1);
  }
}

}  // namespace

// Grid: tiles of TX x TY vertex columns x runs of TZ vertices over the
// interior vertices.  The cell stencil of vertex (i, j, k) spans
// i-ORD/2 .. i+ORD/2-1 per axis, inside the patch for vertices ng..ng+n when
// ng >= ORD/2 (checked by the caller); cells past the last one any interior
// vertex needs are not read.
template <int ORD, int TX, int TY, int TZ, int WAVES, int YS = 0>
static hipError_t tmunu_t(const KPatch &P, int use_z, hipStream_t s) {
  const int ntx = (P.n[0] + 1 + TX - 1) / TX, nty = (P.n[1] + 1 + TY - 1) / TY;
  const int ntz = (P.n[2] + 1 + TZ - 1) / TZ;
  const int nb = ntx * nty * ntz;
  hipLaunchKernelGGL((k_tmunu<ORD, TX, TY, TZ, WAVES, YS>), dim3(nb), dim3(256), 0, s, P, ntx,
                     nty, nb, use_z);
  return hipGetLastError();
}

template <int ORD>
static hipError_t tmunu_v(const KPatch &P, int use_z, int variant, hipStream_t s) {
  // default: 32 x 8 columns at a 3-wave register budget (0.83-0.89 ms for
  // order 4 on 193^3 vertices; the other shapes 0.93-0.98 ms; the banded
  // orders 7 and 8 measure within the same run-to-run spread)
  switch (variant) {
  case 1: return tmunu_t<ORD, 32, 8, 16, 1>(P, use_z, s);
  case 2: return tmunu_t<ORD, 32, 8, 16, 2>(P, use_z, s);
  case 3: return tmunu_t<ORD, 64, 4, 16, 2>(P, use_z, s);
  case 4: return tmunu_t<ORD, 16, 16, 16, 2>(P, use_z, s);
  case 5: return tmunu_t<ORD, 32, 8, 32, 3>(P, use_z, s);
  case 6: return tmunu_t<ORD, 64, 4, 16, 3>(P, use_z, s);
  case 7: return tmunu_t<ORD, 32, 8, 16, 3, 2>(P, use_z, s);
  case 8: return tmunu_t<ORD, 32, 8, 16, 3, 4>(P, use_z, s);
  default: return tmunu_t<ORD, 32, 8, 16, 3>(P, use_z, s);
  }
}

hipError_t launch_tmunu(const KPatch &P, int order, int use_z, int variant, hipStream_t s) {
  return order == 4 ? tmunu_v<4>(P, use_z, variant, s) : tmunu_v<2>(P, use_z, variant, s);
}

}  // namespace asx
