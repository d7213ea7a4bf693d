// Face, vertex and curl kernels of AsterX_Fluxes and the B-field group on
// gfx950: one work-item per face / vertex, x-fastest index order so a
// wavefront reads 64 consecutive doubles of every stencil row.
#include "asx_kernels.hpp"

namespace asx {

// ---------------------------------------------------------------------------
// CalcFlux<D> (AsterX/src/fluxes.cxx:216-795) for faces normal to D.
// ---------------------------------------------------------------------------
template <int D, int M>
__global__ __launch_bounds__(256) void k_flux(KPatch P, FluxArgs A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nf = (int64_t)A.cnt[0] * A.cnt[1] * A.cnt[2];
  if (t >= nf) return;
  const int i = A.lo[0] + (int)(t % A.cnt[0]);
  const int j = A.lo[1] + (int)((t / A.cnt[0]) % A.cnt[1]);
  const int k = A.lo[2] + (int)(t / ((int64_t)A.cnt[0] * A.cnt[1]));
  constexpr int CF = 1 << D;      // face centering
  constexpr int D1 = (D + 1) % 3, D2 = (D + 2) % 3;
  const int oc = off(P, 0, i, j, k); // cell I (right of the face)
  const int sc = step(P, 0, D);
  const int ofc = off(P, CF, i, j, k);
  const int ov = off(P, 7, i, j, k);
  const int svj = P.sj[7], svk = P.sk[7];
  const RParams &rp = A.rp;
  const IdealGas &eos = A.eos;

  // metric at the face
  const double alp = avg_v2f<D>(P.g[ASX_ALP], ov, svj, svk);
  double betas[3];
  betas[0] = avg_v2f<D>(P.g[ASX_BETAX], ov, svj, svk);
  betas[1] = avg_v2f<D>(P.g[ASX_BETAY], ov, svj, svk);
  betas[2] = avg_v2f<D>(P.g[ASX_BETAZ], ov, svj, svk);
  Sym3 g;
  g.xx = avg_v2f<D>(P.g[ASX_GXX], ov, svj, svk);
  g.xy = avg_v2f<D>(P.g[ASX_GXY], ov, svj, svk);
  g.xz = avg_v2f<D>(P.g[ASX_GXZ], ov, svj, svk);
  g.yy = avg_v2f<D>(P.g[ASX_GYY], ov, svj, svk);
  g.yz = avg_v2f<D>(P.g[ASX_GYZ], ov, svj, svk);
  g.zz = avg_v2f<D>(P.g[ASX_GZZ], ov, svj, svk);
  const double detg = det3(g);
  const double sqrtg = sqrt(detg);

  constexpr int m0 = stencil_lo<M>();
  const int mlo = stencil_lo_rt(A.lo_method);
  auto load = [&](int slot, int from, double s[6]) {
    const double *p = P.g[slot];
#pragma unroll
    for (int m = 0; m < 6; ++m)
      if (m >= from && m <= 5 - (from == 2 ? 2 : (from == 1 ? 1 : 0)))
        s[m] = p[oc + (m - 3) * sc];
  };
  double ps[6] = {0, 0, 0, 0, 0, 0}, vs[6] = {0, 0, 0, 0, 0, 0};
  if (M == ASX_RECON_PPM) {
    load(ASX_PRESS, 0, ps);
    load(ASX_VELX + D, 0, vs);
  }
  // high-order face states of one variable
  auto rc = [&](int slot, bool is_rho, double &o0, double &o1) {
    double s[6];
    load(slot, m0, s);
    recon<M>(s, ps, vs, is_rho, rp, o0, o1);
  };
  auto rc_lo = [&](int slot, double &o0, double &o1) {
    double s[6];
    load(slot, mlo, s);
    recon_lo(A.lo_method, s, o0, o1);
  };

  double rho0, rho1, ent0, ent1, ye0, ye1, th0, th1;
  rc(ASX_RHO, true, rho0, rho1);
  rc(ASX_ENTROPY, false, ent0, ent1);
  rc(ASX_YE, false, ye0, ye1);
  const int thermo = A.with_T ? ASX_TEMPERATURE : ASX_PRESS;
  rc(thermo, false, th0, th1);
  const bool lo0 = (rho0 <= 0.0) || (ent0 <= 0.0) || (ye0 <= 0.0) || (th0 <= 0.0);
  const bool lo1 = (rho1 <= 0.0) || (ent1 <= 0.0) || (ye1 <= 0.0) || (th1 <= 0.0);
  if (lo0 || lo1) {
    double a, b;
    rc_lo(ASX_RHO, a, b);
    if (lo0) rho0 = a;
    if (lo1) rho1 = b;
    rc_lo(ASX_ENTROPY, a, b);
    if (lo0) ent0 = a;
    if (lo1) ent1 = b;
    rc_lo(ASX_YE, a, b);
    if (lo0) ye0 = a;
    if (lo1) ye1 = b;
    rc_lo(thermo, a, b);
    if (lo0) th0 = a;
    if (lo1) th1 = b;
  }
  double rho[2] = {rho0, rho1}, ent[2] = {ent0, ent1}, ye[2] = {ye0, ye1};
  double eps[2], press[2];
  if (A.with_T) {
    eps[0] = eos.eps_rt(th0);
    press[0] = eos.press_rt(rho0, th0);
    eps[1] = eos.eps_rt(th1);
    press[1] = eos.press_rt(rho1, th1);
  } else {
    press[0] = th0;
    eps[0] = eos.eps_rp(rho0, th0);
    press[1] = th1;
    eps[1] = eos.eps_rp(rho1, th1);
  }
  double rhoh[2];
  rhoh[0] = rho[0] + rho[0] * eps[0] + press[0];
  rhoh[1] = rho[1] + rho[1] * eps[1] + press[1];

  // magnetic field: normal from the staggered field, transverse recon
  double Bs[3][2];
  {
    const double val = P.g[ASX_DBX_STAG + D][ofc] / sqrtg;
    Bs[D][0] = val;
    Bs[D][1] = val;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int d = q == 0 ? D1 : D2;
      rc(ASX_BVECX + d, false, Bs[d][0], Bs[d][1]);
      if (lo0 || lo1) {
        double a, b;
        rc_lo(ASX_BVECX + d, a, b);
        if (lo0) Bs[d][0] = a;
        if (lo1) Bs[d][1] = b;
      }
    }
  }

  double vel[3][2], vlow[3][2], W[2];
  auto recon_vec = [&](int s0, double o[3][2]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rc(s0 + c, false, o[c][0], o[c][1]);
      if (lo0 || lo1) {
        double a, b;
        rc_lo(s0 + c, a, b);
        if (lo0) o[c][0] = a;
        if (lo1) o[c][1] = b;
      }
    }
  };
  auto lower2 = [&](const double a[3][2], double o[3][2]) {
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const double v[3] = {a[0][f], a[1][f], a[2][f]};
      double l[3];
      lower(g, v, l);
      o[0][f] = l[0];
      o[1][f] = l[1];
      o[2][f] = l[2];
    }
  };
  auto dot2 = [](const double a[3][2], const double b[3][2], int f) {
    return 0.0 + a[0][f] * b[0][f] + a[1][f] * b[1][f] + a[2][f] * b[2][f];
  };
  if (A.rec_var == ASX_RECVAR_V) {
    recon_vec(ASX_VELX, vel);
    lower2(vel, vlow);
    W[0] = 1 / sqrt(1 - dot2(vlow, vel, 0));
    W[1] = 1 / sqrt(1 - dot2(vlow, vel, 1));
  } else if (A.rec_var == ASX_RECVAR_Z) {
    double z[3][2], zl[3][2];
    recon_vec(ASX_ZVEC_X, z);
    lower2(z, zl);
    W[0] = sqrt(1 + dot2(zl, z, 0));
    W[1] = sqrt(1 + dot2(zl, z, 1));
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        vel[c][f] = z[c][f] / W[f];
        vlow[c][f] = zl[c][f] / W[f];
      }
  } else {
    double sv[3][2], sl[3][2];
    recon_vec(ASX_SVEC_X, sv);
    lower2(sv, sl);
#pragma unroll
    for (int f = 0; f < 2; ++f)
      W[f] = sqrt(0.5 + sqrt(0.25 + dot2(sl, sv, f) / rhoh[f] / rhoh[f]));
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        vel[c][f] = sv[c][f] / W[f] / W[f] / rhoh[f];
        vlow[c][f] = sl[c][f] / W[f] / W[f] / rhoh[f];
      }
  }

  double vt[3][2];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int f = 0; f < 2; ++f) vt[c][f] = alp * vel[c][f] - betas[c];
  double ab0[2], Bl[3][2], B2[2], bl[3][2], bsq[2];
#pragma unroll
  for (int f = 0; f < 2; ++f) ab0[f] = W[f] * dot2(Bs, vlow, f);
  lower2(Bs, Bl);
#pragma unroll
  for (int f = 0; f < 2; ++f) B2[f] = dot2(Bs, Bl, f);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int f = 0; f < 2; ++f) bl[c][f] = Bl[c][f] / W[f] + ab0[f] * vlow[c][f];
#pragma unroll
  for (int f = 0; f < 2; ++f) bsq[f] = (B2[f] + pow2(ab0[f])) / pow2(W[f]);

  double cs2[2], h[2], dens[2], dent[2], dhW[2], dhWb[2], ppm[2], tau[2];
  double mom[3][2], Bt[3][2];
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    const double cs = eos.csnd(eps[f]);
    cs2[f] = cs * cs;
    h[f] = 1 + eps[f] + press[f] / rho[f];
    dens[f] = sqrtg * rho[f] * W[f];
    dent[f] = sqrtg * rho[f] * W[f] * ent[f];
    dhW[f] = dens[f] * h[f] * W[f];
    dhWb[f] = dhW[f] + sqrtg * (ab0[f] * ab0[f] + B2[f]);
    ppm[f] = press[f] + 0.5 * bsq[f];
    tau[f] = dhW[f] - dens[f] + sqrtg * (B2[f] - ppm[f]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      mom[c][f] = dhWb[f] * vlow[c][f] - sqrtg * ab0[f] * bl[c][f];
      Bt[c][f] = sqrtg * Bs[c][f];
    }
  const double alp_sqrtg = alp * sqrtg;

  // characteristic speeds (u = (gamma^-1)^{DD})
  const Sym3 ginv = inv3(g, detg);
  const double u = D == 0 ? ginv.xx : (D == 1 ? ginv.yy : ginv.zz);
  double l0p, l0m, l1p, l1m;
  eig_side(alp, betas[D], u, vel[D][0], rho[0], cs2[0], W[0], h[0], bsq[0], l0p, l0m);
  eig_side(alp, betas[D], u, vel[D][1], rho[1], cs2[1], W[1], h[1], bsq[1], l1p, l1m);
  double cmax, cmin;
  char_extrema(l0p, l0m, l1p, l1m, cmax, cmin);
  const double cabs = smax(smax(smax(smax(0.0, fabs(l0p)), fabs(l0m)), fabs(l1p)),
                           fabs(l1m));
  const bool use_hlle = A.flux_type == ASX_FLUX_HLLE;
  auto num = [&](double U0, double U1, double F0, double F1) {
    return use_hlle ? hlle(cmax, cmin, U0, U1, F0, F1)
                    : laxf(cabs, U0, U1, F0, F1);
  };

  double BoW[2];
  BoW[0] = Bs[D][0] / W[0];
  BoW[1] = Bs[D][1] / W[1];
  double *const *F = P.g + ASX_FXDENS + 10 * D;
  F[0][ofc] = num(dens[0], dens[1], dens[0] * vt[D][0], dens[1] * vt[D][1]);
  F[6][ofc] = num(dent[0], dent[1], dent[0] * vt[D][0], dent[1] * vt[D][1]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double unit = (c == D) ? 1.0 : 0.0;
    double fm[2];
#pragma unroll
    for (int f = 0; f < 2; ++f)
      fm[f] = mom[c][f] * vt[D][f] + alp_sqrtg * (ppm[f] * unit - bl[c][f] * BoW[f]);
    F[1 + c][ofc] = num(mom[c][0], mom[c][1], fm[0], fm[1]);
  }
  {
    double ft[2];
#pragma unroll
    for (int f = 0; f < 2; ++f)
      ft[f] = tau[f] * vt[D][f] + alp_sqrtg * (ppm[f] * vel[D][f] - ab0[f] * BoW[f]);
    F[4][ofc] = num(tau[0], tau[1], ft[0], ft[1]);
  }
  {
    const double dye0 = dens[0] * ye[0], dye1 = dens[1] * ye[1];
    F[5][ofc] = num(dye0, dye1, dye0 * vt[D][0], dye1 * vt[D][1]);
  }
  // induction flux: E = Btilde x vtilde, flux(B)_c = (unit x E)_c
  double E[3][2];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int cj = (c + 1) % 3, ck = (c + 2) % 3;
#pragma unroll
    for (int f = 0; f < 2; ++f) E[c][f] = Bt[cj][f] * vt[ck][f] - Bt[ck][f] * vt[cj][f];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int cj = (c + 1) % 3, ck = (c + 2) % 3;
    const double uj = (cj == D) ? 1.0 : 0.0, uk = (ck == D) ? 1.0 : 0.0;
    double fb[2];
#pragma unroll
    for (int f = 0; f < 2; ++f) fb[f] = uj * E[ck][f] - uk * E[cj][f];
    F[7 + c][ofc] = (D != c) * num(Bt[c][0], Bt[c][1], fb[0], fb[1]);
  }

  if (A.write_uct) { // fluxes.cxx:776-793
    const double amx = cmax, amn = -1 * cmin;
    P.g[ASX_AMAX_XFACE + 2 * D][ofc] = amx;
    P.g[ASX_AMIN_XFACE + 2 * D][ofc] = amn;
    // vtilde^(D+1), vtilde^(D+2); a face stores its two transverse
    // components in increasing order (AsterX/interface.ccl:137-150)
    constexpr int s1 = ASX_VTILDE_Y_XFACE + 2 * D + (D1 < D2 ? 0 : 1);
    constexpr int s2 = ASX_VTILDE_Y_XFACE + 2 * D + (D1 < D2 ? 1 : 0);
    P.g[s1][ofc] = (amx * vt[D1][0] + amn * vt[D1][1]) / (amx + amn);
    P.g[s2][ofc] = (amx * vt[D2][0] + amn * vt[D2][1]) / (amx + amn);
  }
  if (A.cnt_dev) {
    const bool lo = lo0 || lo1;
    const bool bad = !(dens[0] == dens[0]) || !(dens[1] == dens[1]) ||
                     !(tau[0] == tau[0]) || !(tau[1] == tau[1]) ||
                     rho[0] < 0.0 || rho[1] < 0.0 || press[0] < 0.0 || press[1] < 0.0;
    const uint64_t mb = __ballot(bad), ml = __ballot(lo);
    if ((threadIdx.x & 63) == 0) {
      if (mb) atomicAdd((unsigned long long *)&A.cnt_dev[0], (unsigned long long)__popcll(mb));
      if (ml) atomicAdd((unsigned long long *)&A.cnt_dev[1], (unsigned long long)__popcll(ml));
    }
  }
}

// CalcAuxForAvecPsi (AsterX/src/fluxes.cxx:803-825)
__global__ __launch_bounds__(256) void k_aux(KPatch P, Box B) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  const int ov = off(P, 7, i, j, k);
  double A[3];
  {
    const int o = off(P, 6, i, j, k);
    A[0] = (0.0 + P.g[ASX_AVEC_X][o] + P.g[ASX_AVEC_X][o - 1]) / 2.0;
  }
  {
    const int o = off(P, 5, i, j, k);
    A[1] = (0.0 + P.g[ASX_AVEC_Y][o] + P.g[ASX_AVEC_Y][o - P.sj[5]]) / 2.0;
  }
  {
    const int o = off(P, 3, i, j, k);
    A[2] = (0.0 + P.g[ASX_AVEC_Z][o] + P.g[ASX_AVEC_Z][o - P.sk[3]]) / 2.0;
  }
  const Sym3 g{P.g[ASX_GXX][ov], P.g[ASX_GXY][ov], P.g[ASX_GXZ][ov],
               P.g[ASX_GYY][ov], P.g[ASX_GYZ][ov], P.g[ASX_GZZ][ov]};
  const double beta[3] = {P.g[ASX_BETAX][ov], P.g[ASX_BETAY][ov], P.g[ASX_BETAZ][ov]};
  const double detg = det3(g);
  const double sqrtg = sqrt(detg);
  const Sym3 ug = inv3(g, detg);
  double Aup[3];
  lower(ug, A, Aup);
  const double alp = P.g[ASX_ALP][ov], psi = P.g[ASX_PSI][ov];
  P.g[ASX_FX][ov] = alp * sqrtg * Aup[0];
  P.g[ASX_FY][ov] = alp * sqrtg * Aup[1];
  P.g[ASX_FZ][ov] = alp * sqrtg * Aup[2];
  P.g[ASX_FBETAX][ov] = beta[0] * psi;
  P.g[ASX_FBETAY][ov] = beta[1] * psi;
  P.g[ASX_FBETAZ][ov] = beta[2] * psi;
  P.g[ASX_G][ov] = alp * psi / sqrtg - dot3(beta, A);
}

// ComputeStaggeredB<D> (AsterX/src/computeBfromA.cxx:32-55)
template <int D>
__global__ __launch_bounds__(256) void k_bstag(KPatch P, Box B, double *out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  const double idx[3] = {1 / P.dx[0], 1 / P.dx[1], 1 / P.dx[2]};
  const double *Ax = P.g[ASX_AVEC_X], *Ay = P.g[ASX_AVEC_Y], *Az = P.g[ASX_AVEC_Z];
  double r;
  if (D == 0) {
    const int oz = off(P, 3, i, j, k), oy = off(P, 5, i, j, k);
    r = idx[1] * (Az[oz + P.sj[3]] - Az[oz]) - idx[2] * (Ay[oy + P.sk[5]] - Ay[oy]);
  } else if (D == 1) {
    const int ox = off(P, 6, i, j, k), oz = off(P, 3, i, j, k);
    r = idx[2] * (Ax[ox + P.sk[6]] - Ax[ox]) - idx[0] * (Az[oz + 1] - Az[oz]);
  } else {
    const int oy = off(P, 5, i, j, k), ox = off(P, 6, i, j, k);
    r = idx[0] * (Ay[oy + 1] - Ay[oy]) - idx[1] * (Ax[ox + P.sj[6]] - Ax[ox]);
  }
  out[off(P, 1 << D, i, j, k)] = r;
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
template <int D>
static hipError_t launch_flux_d(const KPatch &P, const FluxArgs &A, int method,
                                hipStream_t s) {
  const int64_t nf = (int64_t)A.cnt[0] * A.cnt[1] * A.cnt[2];
  if (nf <= 0) return hipSuccess;
  const dim3 grid((unsigned)((nf + 255) / 256)), block(256);
  switch (method) {
  case ASX_RECON_GODUNOV: hipLaunchKernelGGL((k_flux<D, ASX_RECON_GODUNOV>), grid, block, 0, s, P, A); break;
  case ASX_RECON_MINMOD: hipLaunchKernelGGL((k_flux<D, ASX_RECON_MINMOD>), grid, block, 0, s, P, A); break;
  case ASX_RECON_MONOCENTRAL: hipLaunchKernelGGL((k_flux<D, ASX_RECON_MONOCENTRAL>), grid, block, 0, s, P, A); break;
  case ASX_RECON_PPM: hipLaunchKernelGGL((k_flux<D, ASX_RECON_PPM>), grid, block, 0, s, P, A); break;
  case ASX_RECON_WENOZ: hipLaunchKernelGGL((k_flux<D, ASX_RECON_WENOZ>), grid, block, 0, s, P, A); break;
  default: hipLaunchKernelGGL((k_flux<D, ASX_RECON_MP5>), grid, block, 0, s, P, A); break;
  }
  return hipGetLastError();
}

hipError_t launch_flux(int d, const KPatch &P, const FluxArgs &A, int method,
                       hipStream_t s) {
  if (d == 0) return launch_flux_d<0>(P, A, method, s);
  if (d == 1) return launch_flux_d<1>(P, A, method, s);
  return launch_flux_d<2>(P, A, method, s);
}

hipError_t launch_aux(const KPatch &P, const Box &B, hipStream_t s) {
  if (B.htotal() <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_aux, B.grid(), dim3(256), 0, s, P, B);
  return hipGetLastError();
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

} // namespace asx
