// Source terms, RHS (flux divergence, CT EMF, gauge), RK combination,
// boundary fill and the fused update of the substep, for gfx950.
#include "asx_kernels.hpp"

namespace asx {

// calc_fd_v2c<2|4> of a vertex field at cell o (AsterUtils aster_fd.hxx:49-111)
template <int DIR>
ASX_DEV double fd_v2c(const double *g, int o, int sj, int sk, int order, double dx) {
  constexpr int D1 = (DIR + 1) % 3, D2 = (DIR + 2) % 3;
  const int sd = DIR == 0 ? 1 : (DIR == 1 ? sj : sk);
  const int s1 = D1 == 0 ? 1 : (D1 == 1 ? sj : sk);
  const int s2 = D2 == 0 ? 1 : (D2 == 1 ? sj : sk);
  if (order == 2) {
    const double g1 = (g[o + sd] - g[o]) / dx;
    const double g2 = (g[o + s1 + sd] - g[o + s1]) / dx;
    const double g3 = (g[o + s2 + sd] - g[o + s2]) / dx;
    const double g4 = (g[o + s1 + s2 + sd] - g[o + s1 + s2]) / dx;
    return 0.25 * (g1 + g2 + g3 + g4);
  }
  auto d4 = [&](int b) {
    return ((1.0 / 24.0) * g[b - sd] - (27.0 / 24.0) * g[b] + (27.0 / 24.0) * g[b + sd] -
            (1.0 / 24.0) * g[b + 2 * sd]) /
           dx;
  };
  return 0.25 * (d4(o) + d4(o + s1) + d4(o + s2) + d4(o + s1 + s2));
}
ASX_DEV double fdv(int d, const double *g, int o, int sj, int sk, int order, double dx) {
  return d == 0 ? fd_v2c<0>(g, o, sj, sk, order, dx)
                : (d == 1 ? fd_v2c<1>(g, o, sj, sk, order, dx) : fd_v2c<2>(g, o, sj, sk, order, dx));
}

// SourceTerms (AsterX/src/source.cxx:32-167).  r[] = densrhs, momx, momy,
// momz, tau, DEnt, DYe source contributions.
ASX_DEV void source_cell(const KPatch &P, int i, int j, int k, const RhsArgs &A,
                         double r[7]) {
  const int oc = off(P, 0, i, j, k), ov = off(P, 7, i, j, k);
  const int sj = P.sj[7], sk = P.sk[7];
  const double alp = avg_v2c(P.g[ASX_ALP], ov, sj, sk);
  const double beta[3] = {avg_v2c(P.g[ASX_BETAX], ov, sj, sk),
                          avg_v2c(P.g[ASX_BETAY], ov, sj, sk),
                          avg_v2c(P.g[ASX_BETAZ], ov, sj, sk)};
  const Sym3 g{avg_v2c(P.g[ASX_GXX], ov, sj, sk), avg_v2c(P.g[ASX_GXY], ov, sj, sk),
               avg_v2c(P.g[ASX_GXZ], ov, sj, sk), avg_v2c(P.g[ASX_GYY], ov, sj, sk),
               avg_v2c(P.g[ASX_GYZ], ov, sj, sk), avg_v2c(P.g[ASX_GZZ], ov, sj, sk)};
  const Sym3 kk{avg_v2c(P.g[ASX_KXX], ov, sj, sk), avg_v2c(P.g[ASX_KXY], ov, sj, sk),
                avg_v2c(P.g[ASX_KXZ], ov, sj, sk), avg_v2c(P.g[ASX_KYY], ov, sj, sk),
                avg_v2c(P.g[ASX_KYZ], ov, sj, sk), avg_v2c(P.g[ASX_KZZ], ov, sj, sk)};
  const double detg = det3(g);
  const double sqrt_detg = sqrt(detg);
  const Sym3 ug = inv3(g, detg);
  const int ord = A.fd_order;
  double d_alp[3], d_beta[3][3];
  Sym3 d_g[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double dx = P.dx[d];
    d_alp[d] = fdv(d, P.g[ASX_ALP], ov, sj, sk, ord, dx);
#pragma unroll
    for (int c = 0; c < 3; ++c) d_beta[d][c] = fdv(d, P.g[ASX_BETAX + c], ov, sj, sk, ord, dx);
    d_g[d] = Sym3{fdv(d, P.g[ASX_GXX], ov, sj, sk, ord, dx), fdv(d, P.g[ASX_GXY], ov, sj, sk, ord, dx),
                  fdv(d, P.g[ASX_GXZ], ov, sj, sk, ord, dx), fdv(d, P.g[ASX_GYY], ov, sj, sk, ord, dx),
                  fdv(d, P.g[ASX_GYZ], ov, sj, sk, ord, dx), fdv(d, P.g[ASX_GZZ], ov, sj, sk, ord, dx)};
  }
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
    for (int c = 0; c < 3; ++c) v_up[c] = z_up[c] / w;
#pragma unroll
    for (int c = 0; c < 3; ++c) v_low[c] = z_low[c] / w;
  }
  const double rho = P.g[ASX_RHO][oc], eps = P.g[ASX_EPS][oc], press = P.g[ASX_PRESS][oc];
  const double rhW2 = (rho * (1.0 + eps) + press) * w * w;
  double vs[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) vs[c] = v_up[c] - beta[c] / alp;
  const double B_up[3] = {P.g[ASX_BVECX][oc], P.g[ASX_BVECY][oc], P.g[ASX_BVECZ][oc]};
  double B_low[3];
  lower(g, B_up, B_low);
  const double bs0 = w * dot3(B_up, v_low) / alp;
  double bs[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) bs[c] = B_up[c] / w + alp * bs0 * vs[c];
  const double bs2 = (dot3(B_up, B_low) + pow2(alp * bs0)) / pow2(w);
  double gb[3], gbs[3], bs_low[3];
  lower(g, beta, gb);
  lower(g, bs, gbs);
#pragma unroll
  for (int c = 0; c < 3; ++c) bs_low[c] = gb[c] * bs0 + gbs[c];
  const double Eh = rhW2 + bs2 * pow2(w);
  const double Ph = press + 0.5 * bs2;
  const double t00 = (Eh - Ph) / pow2(alp) - pow2(bs0);
  double t0[3], t0low[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) t0[c] = Eh * vs[c] / alp + Ph * beta[c] / pow2(alp) - bs0 * bs[c];
#pragma unroll
  for (int c = 0; c < 3; ++c) t0low[c] = Eh * v_low[c] / alp - bs_low[c] * bs0;
  auto t = [&](int a, int b) {
    return Eh * vs[a] * vs[b] + Ph * (ug(a, b) - beta[a] * beta[b] / pow2(alp)) - bs[a] * bs[b];
  };
  // upper-triangle contraction with off-diagonal weight 2 (sum_symm)
  auto symm = [&](auto f) {
    double s = 0.0;
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int y = x; y < 3; ++y) s += (x == y ? 1 : 2) * f(x, y);
    return s;
  };
  double shiftk[3];
  lower(kk, beta, shiftk);
  const double shiftshiftk = dot3(shiftk, beta);
  double tmp[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) tmp[c] = -d_alp[c] + 2.0 * shiftk[c];
  const double tau_source = t00 * (shiftshiftk - dot3(beta, d_alp)) + dot3(t0, tmp) +
                            symm([&](int a, int b) { return kk(a, b) * t(a, b); });
  double ms[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const Sym3 &dg = d_g[d];
    double s2 = 0.0;
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int y = 0; y < 3; ++y) s2 += dg(x, y) * t0[x] * beta[y];
    ms[d] = t00 * (0.5 * symm([&](int a, int b) { return dg(a, b) * beta[a] * beta[b]; }) -
                   alp * d_alp[d]) +
            s2 + dot3(t0low, d_beta[d]) +
            0.5 * symm([&](int a, int b) { return t(a, b) * dg(a, b); });
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
__constant__ int kRhsSlot[7] = {ASX_DENSRHS, ASX_MOMXRHS, ASX_MOMYRHS, ASX_MOMZRHS,
                                ASX_TAURHS,  ASX_DENTRHS, ASX_DYE_RHS};
// flux variable (offset in a flux group) of each of those
__constant__ int kFluxVar[7] = {0, 1, 2, 3, 4, 6, 5};

__global__ __launch_bounds__(256) void k_sources(KPatch P, Box B, RhsArgs A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  double r[7];
  source_cell(P, i, j, k, A, r);
  const int oc = off(P, 0, i, j, k);
#pragma unroll
  for (int q = 0; q < 7; ++q) P.g[kRhsSlot[q]][oc] = r[q];
}

// -(idx . dF) of one flux variable at cell o (AsterX/src/rhs.cxx:107-114)
ASX_DEV double flux_div(const KPatch &P, int i, int j, int k, int var) {
  const double idx[3] = {1 / P.dx[0], 1 / P.dx[1], 1 / P.dx[2]};
  const int o1 = off(P, 1, i, j, k), o2 = off(P, 2, i, j, k), o4 = off(P, 4, i, j, k);
  const double *Fx = P.g[ASX_FXDENS + var], *Fy = P.g[ASX_FYDENS + var],
               *Fz = P.g[ASX_FZDENS + var];
  const double df[3] = {Fx[o1 + 1] - Fx[o1], Fy[o2 + P.sj[2]] - Fy[o2],
                        Fz[o4 + P.sk[4]] - Fz[o4]};
  return -dot3(idx, df);
}

__global__ __launch_bounds__(256) void k_rhs_hydro(KPatch P, Box B) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  const int oc = off(P, 0, i, j, k);
#pragma unroll
  for (int q = 0; q < 7; ++q) P.g[kRhsSlot[q]][oc] += flux_div(P, i, j, k, kFluxVar[q]);
}

// UCT face data of AsterX/src/rhs.cxx:87-105 by edge direction
__constant__ int kDBone[3] = {ASX_DBY_STAG, ASX_DBZ_STAG, ASX_DBX_STAG};
__constant__ int kDBtwo[3] = {ASX_DBZ_STAG, ASX_DBX_STAG, ASX_DBY_STAG};
__constant__ int kVTone[3] = {ASX_VTILDE_Z_YFACE, ASX_VTILDE_X_ZFACE, ASX_VTILDE_Y_XFACE};
__constant__ int kVTtwo[3] = {ASX_VTILDE_Y_ZFACE, ASX_VTILDE_Z_XFACE, ASX_VTILDE_X_YFACE};
__constant__ int kAmaxOne[3] = {ASX_AMAX_ZFACE, ASX_AMAX_XFACE, ASX_AMAX_YFACE};
__constant__ int kAmaxTwo[3] = {ASX_AMAX_YFACE, ASX_AMAX_ZFACE, ASX_AMAX_XFACE};
__constant__ int kAminOne[3] = {ASX_AMIN_ZFACE, ASX_AMIN_XFACE, ASX_AMIN_YFACE};
__constant__ int kAminTwo[3] = {ASX_AMIN_YFACE, ASX_AMIN_ZFACE, ASX_AMIN_XFACE};

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
  constexpr int J = (I + 1) % 3, K = (I + 2) % 3;
  constexpr int CE = 7 ^ (1 << I); // edge centering
  double E;
  if (A.use_uct) {
    double dB1[2], dB2[2], vt1[2], vt2[2];
    recon_rt(P, A.recon, kDBone[I], K, i, j, k, A.rp, dB1[0], dB1[1]);
    recon_rt(P, A.recon, kDBtwo[I], J, i, j, k, A.rp, dB2[0], dB2[1]);
    recon_rt(P, A.recon, kVTone[I], K, i, j, k, A.rp, vt1[0], vt1[1]);
    recon_rt(P, A.recon, kVTtwo[I], J, i, j, k, A.rp, vt2[0], vt2[1]);
    const double ax1 = at(P, kAmaxOne[I], i, j, k), an1 = at(P, kAminOne[I], i, j, k);
    const double ax2 = at(P, kAmaxTwo[I], i, j, k), an2 = at(P, kAminTwo[I], i, j, k);
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
  return -E - (G[ov + step(P, 7, I)] - G[ov]) / P.dx[I];
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

// Psi_rhs at vertices (AsterX/src/rhs.cxx:227-252, aster_fd.hxx:20-36)
ASX_DEV double psi_rhs(const KPatch &P, int i, int j, int k, const RhsArgs &A) {
  const int ov = off(P, 7, i, j, k);
  if (A.gauge == ASX_GAUGE_ALGEBRAIC) return 0.0;
  double dF = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double dx = P.dx[d];
    const int st = step(P, 7, d);
    const double *F = P.g[ASX_FX + d], *Fb = P.g[ASX_FBETAX + d];
    const double c2c = (0.5 / dx) * (F[ov + st] - F[ov - st]);
    const int sign = P.g[ASX_BETAX + d][ov] < 0 ? -1 : 1;
    const double one =
        -sign * (Fb[ov + 2 * sign * st] - 4.0 * Fb[ov + sign * st] + 3.0 * Fb[ov]) * (0.5 / dx);
    dF += c2c - one;
  }
  return -dF - A.lorenz_damp_fac * P.g[ASX_ALP][ov] * P.g[ASX_PSI][ov];
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

// ---------------------------------------------------------------------------
// Fused update of the substep: out = c0*u0 + c1*u + c2*rhs on interior
// points, rhs = sources + flux divergence (cells), CT EMF + gauge (edges),
// Psi gauge (vertices).  Face fluxes and Aux come from scratch in P.
// ---------------------------------------------------------------------------
struct Comb {
  double c0, c1, c2;
  ASX_DEV double operator()(double u0, double u, double r) const {
    const double acc = (c0 != 0.0) ? c0 * u0 + c1 * u : c1 * u;
    return acc + c2 * r;
  }
};
__constant__ int kStateSlot[ASX_NUM_STATE] = {ASX_DENS, ASX_MOMX, ASX_MOMY, ASX_MOMZ,
                                              ASX_TAU,  ASX_DENT, ASX_DYE,  ASX_AVEC_X,
                                              ASX_AVEC_Y, ASX_AVEC_Z, ASX_PSI};
// state index of each rhs component returned by source_cell
__constant__ int kSrcState[7] = {0, 1, 2, 3, 4, 5, 6};

__global__ __launch_bounds__(256) void k_update_cells(KPatch P, Box B, RhsArgs A,
                                                      StatePtrs u0, StatePtrs out,
                                                      Comb cb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  double r[7];
  source_cell(P, i, j, k, A, r);
  const int oc = off(P, 0, i, j, k);
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const double rhs = r[q] + flux_div(P, i, j, k, kFluxVar[q]);
    const int s = kSrcState[q];
    const double u = P.g[kStateSlot[s]][oc];
    const double z = (cb.c0 != 0.0) ? u0.g[s][oc] : 0.0;
    out.g[s][oc] = cb(z, u, rhs);
  }
}
template <int I>
__global__ __launch_bounds__(256) void k_update_avec(KPatch P, Box B, RhsArgs A,
                                                     StatePtrs u0, StatePtrs out,
                                                     Comb cb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  constexpr int CE = 7 ^ (1 << I);
  const int o = off(P, CE, i, j, k);
  const double rhs = avec_rhs<I>(P, i, j, k, A);
  const double u = P.g[ASX_AVEC_X + I][o];
  const double z = (cb.c0 != 0.0) ? u0.g[7 + I][o] : 0.0;
  out.g[7 + I][o] = cb(z, u, rhs);
}
__global__ __launch_bounds__(256) void k_update_psi(KPatch P, Box B, RhsArgs A,
                                                    StatePtrs u0, StatePtrs out,
                                                    Comb cb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  const int o = off(P, 7, i, j, k);
  const double rhs = psi_rhs(P, i, j, k, A);
  const double z = (cb.c0 != 0.0) ? u0.g[10][o] : 0.0;
  out.g[10][o] = cb(z, P.g[ASX_PSI][o], rhs);
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
  hipLaunchKernelGGL(k_sources, B.grid(), dim3(256), 0, s, P, B, A);
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

hipError_t launch_update(const KPatch &P, const RhsArgs &A, const StatePtrs &u0,
                         const StatePtrs &out, double c0, double c1, double c2,
                         hipStream_t s) {
  const Comb cb{c0, c1, c2};
  Box b = interior_box(P, 0);
  hipLaunchKernelGGL(k_update_cells, b.grid(), dim3(256), 0, s, P, b, A, u0, out, cb);
  b = interior_box(P, 6);
  hipLaunchKernelGGL(k_update_avec<0>, b.grid(), dim3(256), 0, s, P, b, A, u0, out, cb);
  b = interior_box(P, 5);
  hipLaunchKernelGGL(k_update_avec<1>, b.grid(), dim3(256), 0, s, P, b, A, u0, out, cb);
  b = interior_box(P, 3);
  hipLaunchKernelGGL(k_update_avec<2>, b.grid(), dim3(256), 0, s, P, b, A, u0, out, cb);
  b = interior_box(P, 7);
  hipLaunchKernelGGL(k_update_psi, b.grid(), dim3(256), 0, s, P, b, A, u0, out, cb);
  return hipGetLastError();
}

} // namespace asx
