// Vertex-centred stress-energy tensor of the fluid and the magnetic field,
// added into eTtt..eTzz (AsterX/src/tmunu.cxx:19-171): the primitives are
// interpolated from cells to the vertex (calc_avg_c2v,
// AsterUtils/src/aster_interp.hxx:97-148), then T_{mu nu} is formed with the
// vertex metric.
//
// Work decomposition: one work-item per vertex column (i, j) and a run of
// kTmunuTZ vertices along z.  The tensor-product stencil is separable: each
// cell plane is reduced in x and y once (16 loads for order 4, 4 for order
// 2) and kept in a register window of ORD planes, so every vertex after the
// first costs one new plane per variable instead of the full 64 / 8 points.
#include "asx_kernels.hpp"

namespace asx {
namespace {

constexpr int kTmunuTZ = 8;    // vertices per work-item along z
constexpr int kTmunuVars = 9;  // rho, eps, press, B^i, v^i (or z^i)

// Cell-to-vertex weights of calc_avg_c2v<ORD> along one axis, on the cells
// I - ORD/2 .. I + ORD/2 - 1: (-1, 9, 9, -1)/16 for order 4; order 2 is the
// mean of the 8 surrounding cells, 1/2 per axis.
template <int ORD> ASX_DEV constexpr double c2v_w(int a) {
  return ORD == 2 ? 0.5 : ((a == 0 || a == 3) ? -1.0 / 16.0 : 9.0 / 16.0);
}

// x-then-y reduction of one cell plane.  g points at the stencil's first
// cell of the plane relative to the work-item (uniform: the same address for
// the whole workgroup), ob is the work-item's own offset in bytes, so every
// load is a scalar base plus one shared 32-bit vector offset.
template <int ORD>
ASX_DEV double plane_xy(const double *__restrict__ g, unsigned ob, int sj) {
  double s = 0.0;
#pragma unroll
  for (int b = 0; b < ORD; ++b) {
    const double *row = g + b * sj;
    double r = 0.0;
#pragma unroll
    for (int a = 0; a < ORD; ++a)
      r += c2v_w<ORD>(a) * *(const double *)((const char *)(row + a) + ob);
    s += c2v_w<ORD>(b) * r;
  }
  return s;
}

// T_{mu nu} at vertex offset ov from the interpolated primitives
// (AsterX/src/tmunu.cxx:56-170), added into the ten eT slots.
ASX_DEV void tmunu_point(const KPatch &P, bool use_z, int ov, const double q[kTmunuVars]) {
  const double rho = q[0], eps = q[1], press = q[2];
  const double B[3] = {q[3], q[4], q[5]};
  const Sym3 g{P.g[ASX_GXX][ov], P.g[ASX_GXY][ov], P.g[ASX_GXZ][ov],
               P.g[ASX_GYY][ov], P.g[ASX_GYZ][ov], P.g[ASX_GZZ][ov]};
  const double betau[3] = {P.g[ASX_BETAX][ov], P.g[ASX_BETAY][ov], P.g[ASX_BETAZ][ov]};
  const double alp = P.g[ASX_ALP][ov];
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

  P.g[ASX_ETTT][ov] += rhw2 * utl * utl + press * g00 + (w2 * utl * utl + 0.5 * g00) * b2 -
                       btl * btl;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    P.g[ASX_ETTX + d][ov] += rhw2 * utl * vl[d] + press * betal[d] +
                             (w2 * utl * vl[d] + 0.5 * betal[d]) * b2 - btl * bsl[d];
  int s = ASX_ETXX;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b, ++s)
      P.g[s][ov] += rhw2 * vl[a] * vl[b] + press * g(a, b) +
                    (w2 * vl[a] * vl[b] + 0.5 * g(a, b)) * b2 - bsl[a] * bsl[b];
}

// Workgroup w covers 256 consecutive vertex columns of the flattened (i, j)
// plane and one run of kTmunuTZ vertices in z; tiles spread over the XCDs in
// contiguous runs so that neighbouring columns share an L2.
template <int ORD>
__global__ __launch_bounds__(256) void k_tmunu(KPatch P, int nbxy, int nb, int use_z) {
  constexpr int H = ORD / 2;
  const int b = xcd_tile((int)blockIdx.x, nb);
  const int bxy = b % nbxy, bz = b / nbxy;
  const int cx = P.n[0] + 1, cy = P.n[1] + 1, cz = P.n[2] + 1;
  const int t = bxy * 256 + (int)threadIdx.x;
  if (t >= cx * cy) return;
  const int i = P.ng + t % cx, j = P.ng + t / cx;
  const int kb = P.ng + bz * kTmunuTZ;
  const int ke = min(kb + kTmunuTZ, P.ng + cz);
  const int sj = P.sj[0], sk = P.sk[0];
  const unsigned o0 = 8u * (unsigned)((i - H) + sj * (j - H));  // bytes
  const double *src[kTmunuVars] = {
      P.g[ASX_RHO],   P.g[ASX_EPS],   P.g[ASX_PRESS],
      P.g[ASX_BVECX], P.g[ASX_BVECY], P.g[ASX_BVECZ],
      P.g[use_z ? ASX_ZVEC_X : ASX_VELX], P.g[use_z ? ASX_ZVEC_Y : ASX_VELY],
      P.g[use_z ? ASX_ZVEC_Z : ASX_VELZ]};

  // win[q][c]: plane k - H + c reduced in x and y
  double win[kTmunuVars][ORD];
#pragma unroll
  for (int p = 0; p < ORD - 1; ++p)
#pragma unroll
    for (int q = 0; q < kTmunuVars; ++q)
      win[q][p + 1] = plane_xy<ORD>(src[q] + sk * (kb - H + p), o0, sj);

  for (int k = kb; k < ke; ++k) {
    double v[kTmunuVars];
#pragma unroll
    for (int q = 0; q < kTmunuVars; ++q) {
#pragma unroll
      for (int c = 0; c < ORD - 1; ++c) win[q][c] = win[q][c + 1];
      win[q][ORD - 1] = plane_xy<ORD>(src[q] + sk * (k + H - 1), o0, sj);
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < ORD; ++c) s += c2v_w<ORD>(c) * win[q][c];
      v[q] = s;
    }
    tmunu_point(P, use_z != 0, i + P.sj[7] * j + P.sk[7] * k, v);
  }
}

}  // namespace

// Grid: ceil((n0+1)(n1+1)/256) column tiles x ceil((n2+1)/kTmunuTZ) z runs.
// The cell stencil of vertex (i, j, k) spans i-ORD/2 .. i+ORD/2-1 per axis,
// inside the patch for vertices ng..ng+n when ng >= ORD/2 (checked by the
// caller).
hipError_t launch_tmunu(const KPatch &P, int order, int use_z, hipStream_t s) {
  const int64_t cols = (int64_t)(P.n[0] + 1) * (P.n[1] + 1);
  const int nbxy = (int)((cols + 255) / 256);
  const int nbz = (P.n[2] + 1 + kTmunuTZ - 1) / kTmunuTZ;
  const int nb = nbxy * nbz;
  if (order == 4)
    hipLaunchKernelGGL(k_tmunu<4>, dim3(nb), dim3(256), 0, s, P, nbxy, nb, use_z);
  else
    hipLaunchKernelGGL(k_tmunu<2>, dim3(nb), dim3(256), 0, s, P, nbxy, nb, use_z);
  return hipGetLastError();
}

}  // namespace asx
