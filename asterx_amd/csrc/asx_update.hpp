// Device code of the RHS and the fused update shared by the kernels of
// asx_rhs.hip and the batch plan's persistent stage kernel (asx_plan.hip):
// source terms, flux divergence, CT EMF and gauge terms, the Psi equation,
// the boundary fill of one ghost point and the update of one point.
#pragma once
#include <type_traits>

#include "asx_kernels.hpp"

namespace asx {

// calc_fd_v2c<2|4> of a vertex field at cell o (AsterUtils aster_fd.hxx:49-111)
template <int DIR>
ASX_DEV double fd_v2c(const double *g, int o, int sj, int sk, int order, double idx) {
  constexpr int D1 = (DIR + 1) % 3, D2 = (DIR + 2) % 3;
  const int sd = DIR == 0 ? 1 : (DIR == 1 ? sj : sk);
  const int s1 = D1 == 0 ? 1 : (D1 == 1 ? sj : sk);
  const int s2 = D2 == 0 ? 1 : (D2 == 1 ? sj : sk);
  if (order == 2) {
    const double g1 = (g[o + sd] - g[o]) * idx;
    const double g2 = (g[o + s1 + sd] - g[o + s1]) * idx;
    const double g3 = (g[o + s2 + sd] - g[o + s2]) * idx;
    const double g4 = (g[o + s1 + s2 + sd] - g[o + s1 + s2]) * idx;
    return 0.25 * (g1 + g2 + g3 + g4);
  }
  auto d4 = [&](int b) {
    return ((1.0 / 24.0) * g[b - sd] - (27.0 / 24.0) * g[b] + (27.0 / 24.0) * g[b + sd] -
            (1.0 / 24.0) * g[b + 2 * sd]) *
           idx;
  };
  return 0.25 * (d4(o) + d4(o + s1) + d4(o + s2) + d4(o + s1 + s2));
}
ASX_DEV double fdv(int d, const double *g, int o, int sj, int sk, int order, double idx) {
  return d == 0 ? fd_v2c<0>(g, o, sj, sk, order, idx)
                : (d == 1 ? fd_v2c<1>(g, o, sj, sk, order, idx) : fd_v2c<2>(g, o, sj, sk, order, idx));
}

// The 8 vertex values around cell (vertex base o), bit d of the index set
// for the +1 vertex along axis d.
ASX_DEV void corners(const double *g, int o, int sj, int sk, double c[8]) {
  c[0] = g[o];
  c[1] = g[o + 1];
  c[2] = g[o + sj];
  c[3] = g[o + sj + 1];
  c[4] = g[o + sk];
  c[5] = g[o + sk + 1];
  c[6] = g[o + sk + sj];
  c[7] = g[o + sk + sj + 1];
}
// calc_avg_v2c from corners, same summation order as avg_v2c
ASX_DEV double corner_avg(const double c[8]) {
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) s += c[q];
  return s / 8.0;
}
// calc_fd_v2c<2> along d from corners (same expression as fd_v2c)
ASX_DEV double corner_d2(const double c[8], int d, double idx) {
  const int bd = 1 << d, b1 = 1 << ((d + 1) % 3), b2 = 1 << ((d + 2) % 3);
  const double g1 = (c[bd] - c[0]) * idx;
  const double g2 = (c[b1 | bd] - c[b1]) * idx;
  const double g3 = (c[b2 | bd] - c[b2]) * idx;
  const double g4 = (c[b1 | b2 | bd] - c[b1 | b2]) * idx;
  return 0.25 * (g1 + g2 + g3 + g4);
}
// Scheduling fence: no instruction moves across it, so the loads of the
// next field are not hoisted above the arithmetic of the previous one.  It
// bounds how many stencil loads a work-item keeps in flight (each in-flight
// double holds two VGPRs) and with it the register footprint.
ASX_DEV void load_fence() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
}

// Gradient of vertex field g at cell (vertex base o), order 2 or 4.
template <int ORD>
ASX_DEV void cell_grad(const KPatch &P, const double *g, int o, double out[3]) {
  const int sj = P.sj[7], sk = P.sk[7];
  load_fence();
  if (ORD == 2) {
    double c[8];
    corners(g, o, sj, sk, c);
#pragma unroll
    for (int d = 0; d < 3; ++d) out[d] = corner_d2(c, d, P.idx[d]);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) out[d] = fdv(d, g, o, sj, sk, 4, P.idx[d]);
  }
}
// Cell average of vertex field g, fenced like cell_grad.
ASX_DEV double cell_avg(const double *g, int o, int sj, int sk) {
  load_fence();
  return avg_v2c(g, o, sj, sk);
}

// SourceTerms (AsterX/src/source.cxx:32-167).  r[] = densrhs, momx, momy,
// momz, tau, DEnt, DYe source contributions.
//
// Two passes over the vertex metric: the cell averages (and d alpha) give
// T^{mu nu}; the momentum source is then linear in the metric derivatives,
//   S_d = sum_{a<=b} C_ab d_d g_ab + sum_c T^0_c d_d beta^c - T^00 alpha d_d alpha,
//   C_ab = w_ab (T^00 beta^a beta^b + T^ab) / 2 + (t0^a beta^b + t0^b beta^a) / (1 + [a==b]),
// (w_ab = 1 on, 2 off the diagonal), which is source.cxx:138-150 regrouped
// by derivative, so that each derivative is consumed as it is loaded.
template <int ORD>
ASX_DEV void source_cell(const KPatch &P, int i, int j, int k, const RhsArgs &A,
                         double r[7]) {
  ASX_CONTRACT
  const int oc = off(P, 0, i, j, k), ov = off(P, 7, i, j, k);
  const int sj = P.sj[7], sk = P.sk[7];
  double c[8];
  corners(P.g[ASX_ALP], ov, sj, sk, c);
  const double alp = corner_avg(c);
  double d_alp[3];
  if (ORD == 2) {
#pragma unroll
    for (int d = 0; d < 3; ++d) d_alp[d] = corner_d2(c, d, P.idx[d]);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) d_alp[d] = fdv(d, P.g[ASX_ALP], ov, sj, sk, 4, P.idx[d]);
  }
  const double beta[3] = {cell_avg(P.g[ASX_BETAX], ov, sj, sk),
                          cell_avg(P.g[ASX_BETAY], ov, sj, sk),
                          cell_avg(P.g[ASX_BETAZ], ov, sj, sk)};
  const Sym3 g{cell_avg(P.g[ASX_GXX], ov, sj, sk), cell_avg(P.g[ASX_GXY], ov, sj, sk),
               cell_avg(P.g[ASX_GXZ], ov, sj, sk), cell_avg(P.g[ASX_GYY], ov, sj, sk),
               cell_avg(P.g[ASX_GYZ], ov, sj, sk), cell_avg(P.g[ASX_GZZ], ov, sj, sk)};
  const double detg = det3(g);
  const double sqrt_detg = sqrt(detg);
  const Sym3 ug = inv3(g, detg);
  double v_up[3], v_low[3], w;
  if (A.rec_var == ASX_RECVAR_V) {
    v_up[0] = P.g[ASX_VELX][oc];
    v_up[1] = P.g[ASX_VELY][oc];
    v_up[2] = P.g[ASX_VELZ][oc];
    lower(g, v_up, v_low);
    w = 1.0 / sqrt(1.0 - dot3(v_low, v_up));
  } else {
    const double z_up[3] = {P.g[ASX_ZVEC_X][oc], P.g[ASX_ZVEC_Y][oc], P.g[ASX_ZVEC_Z][oc]};
    double z_low[3];
    lower(g, z_up, z_low);
    w = sqrt(1.0 + dot3(z_low, z_up));
#pragma unroll
    for (int q = 0; q < 3; ++q) v_up[q] = z_up[q] / w;
#pragma unroll
    for (int q = 0; q < 3; ++q) v_low[q] = z_low[q] / w;
  }
  const double rho = P.g[ASX_RHO][oc], eps = P.g[ASX_EPS][oc], press = P.g[ASX_PRESS][oc];
  const double rhW2 = (rho * (1.0 + eps) + press) * w * w;
  const double ialp = 1.0 / alp, ialp2 = ialp * ialp, iw = 1.0 / w;
  double vs[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) vs[q] = v_up[q] - beta[q] * ialp;
  const double B_up[3] = {P.g[ASX_BVECX][oc], P.g[ASX_BVECY][oc], P.g[ASX_BVECZ][oc]};
  double B_low[3];
  lower(g, B_up, B_low);
  const double bs0 = w * dot3(B_up, v_low) * ialp;
  double bs[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) bs[q] = B_up[q] * iw + alp * bs0 * vs[q];
  const double bs2 = (dot3(B_up, B_low) + pow2(alp * bs0)) * (iw * iw);
  double gb[3], gbs[3], bs_low[3];
  lower(g, beta, gb);
  lower(g, bs, gbs);
#pragma unroll
  for (int q = 0; q < 3; ++q) bs_low[q] = gb[q] * bs0 + gbs[q];
  const double Eh = rhW2 + bs2 * pow2(w);
  const double Ph = press + 0.5 * bs2;
  const double t00 = (Eh - Ph) * ialp2 - pow2(bs0);
  double t0[3], t0low[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) t0[q] = Eh * vs[q] * ialp + Ph * beta[q] * ialp2 - bs0 * bs[q];
#pragma unroll
  for (int q = 0; q < 3; ++q) t0low[q] = Eh * v_low[q] * ialp - bs_low[q] * bs0;
  auto t = [&](int a, int b) {
    return Eh * vs[a] * vs[b] + Ph * (ug(a, b) - beta[a] * beta[b] * ialp2) - bs[a] * bs[b];
  };
  // tau source (source.cxx:152-158), linear in K:
  //   tau_S = -T^00 beta.d(alpha) - t0.d(alpha) + sum_{a<=b} K_ab D_ab,
  //   D_ab = w_ab (T^00 beta^a beta^b + T^ab) + 2 (t0^a beta^b + t0^b beta^a) / (1 + [a==b])
  double tau_source = -t00 * dot3(beta, d_alp) - dot3(t0, d_alp);
  {
    double dk[6];
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int y = x; y < 3; ++y) {
        const double wab = x == y ? 1.0 : 2.0;
        const double cross = x == y ? 2.0 * t0[x] * beta[x]
                                    : 2.0 * (t0[x] * beta[y] + t0[y] * beta[x]);
        dk[x == 0 ? y : (x == 1 ? 2 + y : 5)] =
            wab * (t00 * beta[x] * beta[y] + t(x, y)) + cross;
      }
#pragma unroll 1
    for (int f = 0; f < 6; ++f) {
      double c = dk[0];
#pragma unroll
      for (int q = 1; q < 6; ++q) c = (f == q) ? dk[q] : c;
      tau_source += c * cell_avg(P.g[ASX_KXX + f], ov, sj, sk);
    }
  }
  double ms[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) ms[d] = -(t00 * alp) * d_alp[d];
  // the derivative pass re-reads the metric (cache hits) rather than keeping
  // 80 corner values live across the stress-energy computation
  // coefficient of each derivative: beta^x, beta^y, beta^z, then g_ab
  double cf[9];
#pragma unroll
  for (int q = 0; q < 3; ++q) cf[q] = t0low[q];
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int y = x; y < 3; ++y) {
      const double wab = x == y ? 1.0 : 2.0;
      const double cross = x == y ? t0[x] * beta[x] : t0[x] * beta[y] + t0[y] * beta[x];
      cf[3 + (x == 0 ? y : (x == 1 ? 2 + y : 5))] =
          0.5 * wab * (t00 * beta[x] * beta[y] + t(x, y)) + cross;
    }
  // one field per iteration (not unrolled), so only that field's corner
  // values are live
#pragma unroll 1
  for (int f = 0; f < 9; ++f) {
    const int slot = f < 3 ? ASX_BETAX + f : ASX_GXX + (f - 3);
    double dq[3];
    cell_grad<ORD>(P, P.g[slot], ov, dq);
    double c = cf[0];
#pragma unroll
    for (int q = 1; q < 9; ++q) c = (f == q) ? cf[q] : c;
#pragma unroll
    for (int d = 0; d < 3; ++d) ms[d] += c * dq[d];
  }
  r[0] = 0.0;
  r[1] = alp * sqrt_detg * ms[0];
  r[2] = alp * sqrt_detg * ms[1];
  r[3] = alp * sqrt_detg * ms[2];
  r[4] = alp * sqrt_detg * tau_source;
  r[5] = 0.0;
  r[6] = 0.0;
}

// rhs slots in the order source_cell returns them
ASX_DEV constexpr int kRhsSlot(int q) {
  constexpr int t[7] = {ASX_DENSRHS, ASX_MOMXRHS, ASX_MOMYRHS, ASX_MOMZRHS,
                                ASX_TAURHS,  ASX_DENTRHS, ASX_DYE_RHS};
  return t[q];
}
// flux variable (offset in a flux group) of each of those
ASX_DEV constexpr int kFluxVar(int q) {
  constexpr int t[7] = {0, 1, 2, 3, 4, 6, 5};
  return t[q];
}

// -(idx . dF) of one flux variable at cell o (AsterX/src/rhs.cxx:107-114)
ASX_DEV double flux_div(const KPatch &P, int i, int j, int k, int var) {
  ASX_CONTRACT
  const double *idx = P.idx;
  const int o1 = off(P, 1, i, j, k), o2 = off(P, 2, i, j, k), o4 = off(P, 4, i, j, k);
  const double *Fx = P.g[ASX_FXDENS + var], *Fy = P.g[ASX_FYDENS + var],
               *Fz = P.g[ASX_FZDENS + var];
  const double df[3] = {Fx[o1 + 1] - Fx[o1], Fy[o2 + P.sj[2]] - Fy[o2],
                        Fz[o4 + P.sk[4]] - Fz[o4]};
  return -dot3(idx, df);
}

// UCT face data of AsterX/src/rhs.cxx:87-105 by edge direction
ASX_DEV constexpr int kDBone(int q) {
  constexpr int t[3] = {ASX_DBY_STAG, ASX_DBZ_STAG, ASX_DBX_STAG};
  return t[q];
}
ASX_DEV constexpr int kDBtwo(int q) {
  constexpr int t[3] = {ASX_DBZ_STAG, ASX_DBX_STAG, ASX_DBY_STAG};
  return t[q];
}
ASX_DEV constexpr int kVTone(int q) {
  constexpr int t[3] = {ASX_VTILDE_Z_YFACE, ASX_VTILDE_X_ZFACE, ASX_VTILDE_Y_XFACE};
  return t[q];
}
ASX_DEV constexpr int kVTtwo(int q) {
  constexpr int t[3] = {ASX_VTILDE_Y_ZFACE, ASX_VTILDE_Z_XFACE, ASX_VTILDE_X_YFACE};
  return t[q];
}
ASX_DEV constexpr int kAmaxOne(int q) {
  constexpr int t[3] = {ASX_AMAX_ZFACE, ASX_AMAX_XFACE, ASX_AMAX_YFACE};
  return t[q];
}
ASX_DEV constexpr int kAmaxTwo(int q) {
  constexpr int t[3] = {ASX_AMAX_YFACE, ASX_AMAX_ZFACE, ASX_AMAX_XFACE};
  return t[q];
}
ASX_DEV constexpr int kAminOne(int q) {
  constexpr int t[3] = {ASX_AMIN_ZFACE, ASX_AMIN_XFACE, ASX_AMIN_YFACE};
  return t[q];
}
ASX_DEV constexpr int kAminTwo(int q) {
  constexpr int t[3] = {ASX_AMIN_YFACE, ASX_AMIN_ZFACE, ASX_AMIN_XFACE};
  return t[q];
}

// Reconstruct a face field along `dir` at the edge index (as the reference
// does, with the edge's I used in the face array's layout).
template <int M>
ASX_DEV void recon_at(const KPatch &P, int slot, int dir, int i, int j, int k,
                      const RParams &rp, double &o0, double &o1) {
  const int c = cent(slot);
  const int o = off(P, c, i, j, k), st = step(P, c, dir);
  double s[6], ps[6], vs[6];
  constexpr int m0 = stencil_lo<M>();
#pragma unroll
  for (int m = 0; m < 6; ++m)
    if (m >= m0 && m <= 5 - m0) s[m] = P.g[slot][o + (m - 3) * st];
  if (M == ASX_RECON_PPM) {
    const int oc = off(P, 0, i, j, k), sc = step(P, 0, dir);
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      ps[m] = P.g[ASX_PRESS][oc + (m - 3) * sc];
      vs[m] = P.g[ASX_VELX + dir][oc + (m - 3) * sc];
    }
  }
  recon<M>(s, ps, vs, false, rp, o0, o1);
}
ASX_DEV void recon_rt(const KPatch &P, int m, int slot, int dir, int i, int j, int k,
                      const RParams &rp, double &o0, double &o1) {
  switch (m) {
  case ASX_RECON_GODUNOV: recon_at<ASX_RECON_GODUNOV>(P, slot, dir, i, j, k, rp, o0, o1); break;
  case ASX_RECON_MINMOD: recon_at<ASX_RECON_MINMOD>(P, slot, dir, i, j, k, rp, o0, o1); break;
  case ASX_RECON_MONOCENTRAL: recon_at<ASX_RECON_MONOCENTRAL>(P, slot, dir, i, j, k, rp, o0, o1); break;
  case ASX_RECON_PPM: recon_at<ASX_RECON_PPM>(P, slot, dir, i, j, k, rp, o0, o1); break;
  case ASX_RECON_WENOZ: recon_at<ASX_RECON_WENOZ>(P, slot, dir, i, j, k, rp, o0, o1); break;
  default: recon_at<ASX_RECON_MP5>(P, slot, dir, i, j, k, rp, o0, o1); break;
  }
}

// A_rhs on edges along I (AsterX/src/rhs.cxx:116-186)
template <int I>
ASX_DEV double avec_rhs(const KPatch &P, int i, int j, int k, const RhsArgs &A) {
  ASX_CONTRACT
  constexpr int J = (I + 1) % 3, K = (I + 2) % 3;
  constexpr int CE = 7 ^ (1 << I); // edge centering
  double E;
  if (A.use_uct) {
    double dB1[2], dB2[2], vt1[2], vt2[2];
    recon_rt(P, A.recon, kDBone(I), K, i, j, k, A.rp, dB1[0], dB1[1]);
    recon_rt(P, A.recon, kDBtwo(I), J, i, j, k, A.rp, dB2[0], dB2[1]);
    recon_rt(P, A.recon, kVTone(I), K, i, j, k, A.rp, vt1[0], vt1[1]);
    recon_rt(P, A.recon, kVTtwo(I), J, i, j, k, A.rp, vt2[0], vt2[1]);
    const double ax1 = at(P, kAmaxOne(I), i, j, k), an1 = at(P, kAminOne(I), i, j, k);
    const double ax2 = at(P, kAmaxTwo(I), i, j, k), an2 = at(P, kAminTwo(I), i, j, k);
    const double den1 = ax1 + an1;
    E = (ax1 * vt1[0] * dB1[0] + an1 * vt1[1] * dB1[1] - ax1 * an1 * (dB1[1] - dB1[0])) / den1;
    const double den2 = ax2 + an2;
    E -= (ax2 * vt2[0] * dB2[0] + an2 * vt2[1] * dB2[1] - ax2 * an2 * (dB2[1] - dB2[0])) / den2;
  } else {
    // flux of B_J in direction K (faces normal to K) and of B_K along J
    constexpr int SJK = ASX_FXDENS + 10 * K + 7 + J;
    constexpr int SKJ = ASX_FXDENS + 10 * J + 7 + K;
    const int oK = off(P, 1 << K, i, j, k), oJ = off(P, 1 << J, i, j, k);
    const double *Fjk = P.g[SJK], *Fkj = P.g[SKJ];
    E = 0.25 * ((Fjk[oK] + Fjk[oK - step(P, 1 << K, J)]) -
                (Fkj[oJ] + Fkj[oJ - step(P, 1 << J, K)]));
  }
  if (A.gauge == ASX_GAUGE_ALGEBRAIC) return -E;
  const int ov = off(P, 7, i, j, k);
  const double *G = P.g[ASX_G];
  (void)CE;
  return -E - (G[ov + step(P, 7, I)] - G[ov]) * P.idx[I];
}

// Psi_rhs at vertices (AsterX/src/rhs.cxx:227-252, aster_fd.hxx:20-36)
ASX_DEV double psi_rhs(const KPatch &P, int i, int j, int k, const RhsArgs &A) {
  ASX_CONTRACT
  const int ov = off(P, 7, i, j, k);
  if (A.gauge == ASX_GAUGE_ALGEBRAIC) return 0.0;
  double dF = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double hidx = 0.5 * P.idx[d]; // == 0.5 / dx exactly
    const int st = step(P, 7, d);
    const double *F = P.g[ASX_FX + d], *Fb = P.g[ASX_FBETAX + d];
    const double c2c = hidx * (F[ov + st] - F[ov - st]);
    const int sign = P.g[ASX_BETAX + d][ov] < 0 ? -1 : 1;
    const double one =
        -sign * (Fb[ov + 2 * sign * st] - 4.0 * Fb[ov + sign * st] + 3.0 * Fb[ov]) * hidx;
    dF += c2c - one;
  }
  return -dF - A.lorenz_damp_fac * P.g[ASX_ALP][ov] * P.g[ASX_PSI][ov];
}

// The same fill for up to kMaxBC grid functions of any centering in one
// launch per axis (blockIdx.y = grid function): one work-item per ghost point,
// point t of the 2 ng ghost layers normal to D (x fastest, so that the layers
// normal to y and z are written in full rows).  Each ghost value depends
// only on interior points along D, so the points of one axis are
// independent; the axes run in order x, y, z as in k_bc.
template <int D>
ASX_DEV void bc_point(double *g, int c, int n0, int n1, int n2, int ng, int kind, int t) {
  const int n[3] = {n0, n1, n2};
  const int e[3] = {n0 + 2 * ng + (c & 1), n1 + 2 * ng + ((c >> 1) & 1),
                    n2 + 2 * ng + ((c >> 2) & 1)};
  // The axes run in the order x, y, z, and each pass reads only interior
  // rows / planes along its own axis, so the x pass covers the interior rows
  // in y and z and the y pass the interior planes in z: the ghost rows and
  // planes of a later axis are written by that axis's pass.  (Covering them
  // too, as a single pass per axis over the full extent would, only writes
  // values the later pass overwrites.)
  int w[3] = {e[0], e[1], e[2]}, o[3] = {0, 0, 0};
#pragma unroll
  for (int a = D + 1; a < 3; ++a) {
    w[a] = e[a] - 2 * ng;
    o[a] = ng;
  }
  w[D] = 2 * ng;
  if (t >= w[0] * w[1] * w[2]) return;
  int x[3] = {t % w[0] + o[0], (t / w[0]) % w[1] + o[1], t / (w[0] * w[1]) + o[2]};
  const int lo = ng, hi = ng + n[D] - 1 + ((c >> D) & 1);
  const int q = x[D] < ng ? x[D] : hi + 1 + (x[D] - ng);
  const int64_t ex = e[0], exy = (int64_t)e[0] * e[1];
  const int64_t sd = D == 0 ? 1 : (D == 1 ? ex : exy);
  x[D] = 0;
  double *l = g + x[0] + x[1] * ex + x[2] * exy;
  double v;
  if (q < lo) {
    if (kind == ASX_BC_LINEAR)
      v = l[lo * sd] + (lo - q) * (l[lo * sd] - l[(lo + 1) * sd]);
    else if (kind == ASX_BC_PERIODIC)
      v = l[(q + n[D]) * sd];
    else
      v = l[lo * sd];
  } else {
    if (kind == ASX_BC_LINEAR)
      v = l[hi * sd] + (q - hi) * (l[hi * sd] - l[(hi - 1) * sd]);
    else if (kind == ASX_BC_PERIODIC)
      v = l[(q - n[D]) * sd];
    else
      v = l[hi * sd];
  }
  l[q * sd] = v;
}
// points of bc_point's pass normal to d for the widest centering (vertex on
// every axis): the full extent on the axes before d, the interior after it
__host__ inline int bc_points(const int n[3], int ng, int d) {
  int w[3] = {n[0] + 2 * ng + 1, n[1] + 2 * ng + 1, n[2] + 2 * ng + 1};
  for (int a = d + 1; a < 3; ++a) w[a] = n[a] + 1;
  w[d] = 2 * ng;
  return w[0] * w[1] * w[2];
}

// ---------------------------------------------------------------------------
// Fused update of the substep: out = c0*u0 + c1*u + c2*rhs on interior
// points, rhs = sources + flux divergence (cells), CT EMF + gauge (edges),
// Psi gauge (vertices).  Face fluxes and Aux come from scratch in P.
// ---------------------------------------------------------------------------
struct Comb {
  double c0, c1, c2;
  ASX_DEV double operator()(double u0, double u, double r) const {
  ASX_CONTRACT
    const double acc = (c0 != 0.0) ? c0 * u0 + c1 * u : c1 * u;
    return acc + c2 * r;
  }
};
ASX_DEV constexpr int kStateSlot(int q) {
  constexpr int t[ASX_NUM_STATE] = {ASX_DENS, ASX_MOMX, ASX_MOMY, ASX_MOMZ,
                                              ASX_TAU,  ASX_DENT, ASX_DYE,  ASX_AVEC_X,
                                              ASX_AVEC_Y, ASX_AVEC_Z, ASX_PSI};
  return t[q];
}
// state index of each rhs component returned by source_cell
ASX_DEV constexpr int kSrcState(int q) {
  constexpr int t[7] = {0, 1, 2, 3, 4, 5, 6};
  return t[q];
}

// One work-item per point (i, j, k) of the box [ng, ng+n] on every axis:
// the interior cell (sources + flux divergence), the three interior edges
// and the interior vertex that share the index, each combined with the
// previous stages as out = c0*u0 + c1*u + c2*rhs.
// waves per SIMD the register budget is sized for (order-4 derivatives
// need 32 more doubles live than order 2)
template <int ORD> constexpr int kUpdateWaves = ORD == 2 ? 4 : 3;
template <int ORD>
ASX_DEV void update_point(const KPatch &P, int i, int j, int k, const RhsArgs &A,
                          const StatePtrs &u0, const StatePtrs &out, const Comb &cb) {
  ASX_CONTRACT
  const bool cx = i < P.ng + P.n[0], cy = j < P.ng + P.n[1], cz = k < P.ng + P.n[2];
  if (cx && cy && cz) {
    double r[7];
    source_cell<ORD>(P, i, j, k, A, r);
    const int oc = off(P, 0, i, j, k);
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const double rhs = r[q] + flux_div(P, i, j, k, kFluxVar(q));
      const int s = kSrcState(q);
      const double u = P.g[kStateSlot(s)][oc];
      const double z = (cb.c0 != 0.0) ? u0.g[s][oc] : 0.0;
      out.g[s][oc] = cb(z, u, rhs);
    }
  }
  auto edge = [&](auto I_) {
    constexpr int I = decltype(I_)::value;
    constexpr int CE = 7 ^ (1 << I);
    const int o = off(P, CE, i, j, k);
    const double rhs = avec_rhs<I>(P, i, j, k, A);
    const double u = P.g[ASX_AVEC_X + I][o];
    const double z = (cb.c0 != 0.0) ? u0.g[7 + I][o] : 0.0;
    out.g[7 + I][o] = cb(z, u, rhs);
  };
  if (cx) edge(std::integral_constant<int, 0>{});
  if (cy) edge(std::integral_constant<int, 1>{});
  if (cz) edge(std::integral_constant<int, 2>{});
  {
    const int o = off(P, 7, i, j, k);
    const double rhs = psi_rhs(P, i, j, k, A);
    const double z = (cb.c0 != 0.0) ? u0.g[10][o] : 0.0;
    out.g[10][o] = cb(z, P.g[ASX_PSI][o], rhs);
  }
}

} // namespace asx
