// Device code of AsterX_Fluxes shared by the face, vertex and curl kernels
// (asx_flux.hip) and the batch plan's persistent stage kernel
// (asx_plan.hip): reconstruction, the numerical flux of one face, the face
// tile with its LDS hand-over, the vertex auxiliaries and the staggered
// field of one point.
#pragma once
#include "asx_kernels.hpp"

namespace asx {

// ---------------------------------------------------------------------------
// CalcFlux<D> (AsterX/src/fluxes.cxx:216-795) for faces normal to D.
//
// Work decomposition: one work-item per cell of a tile that is one cell
// longer along D than the faces it produces.  Each work-item reconstructs
// its cell once, both sides (ReconX's per-cell parabola / WENO weights are
// shared by the two faces of the cell), keeps the low side in registers as
// the right state of its own face and hands the high side to the next cell
// along D through LDS, where it is the left state of that face.  Tiles
// overlap by one cell along D.
// ---------------------------------------------------------------------------

// The reconstructed primitives of one side of a cell: rho, entropy, Ye,
// pressure (or temperature), the two transverse B components (D+1, D+2)
// and the velocity-like vector (v, z or s).
struct Side {
  double rho, ent, ye, th, b1, b2, v[3];
};
constexpr int kSideVars = 9;

// Points of a cell's stencil along D that method M reads: s[m] holds the
// value at oc + (m - 2) sc.
template <int M> struct StencilSpan {
  static constexpr int m0 = M == ASX_RECON_GODUNOV ? 2
                            : (M == ASX_RECON_MINMOD || M == ASX_RECON_MONOCENTRAL) ? 1
                                                                                   : 0;
};
template <int M> ASX_DEV void cell_stencil(const double *p, int oc, int sc, double s[5]) {
  constexpr int m0 = StencilSpan<M>::m0;
#pragma unroll
  for (int m = m0; m <= 4 - m0; ++m) s[m] = p[oc + (m - 2) * sc];
}
// Face values (lo = the cell's face at -D, up = at +D) of one grid function
// from its 5-cell stencil along D.  Equal, value for value, to the two
// sides that ReconX's reconstruct() yields for the faces of this cell.
template <int M>
ASX_DEV void cell_recon_s(const double s[5], const double ps[5], const double vs[5],
                          bool is_rho, const RParams &rp, double &lo, double &up) {
  if (M == ASX_RECON_GODUNOV) {
    lo = s[2];
    up = s[2];
  } else if (M == ASX_RECON_MINMOD) {
    const double sl = minmod(s[3] - s[2], s[2] - s[1]);
    up = s[2] + sl / 2;
    lo = s[2] - sl / 2;
  } else if (M == ASX_RECON_MONOCENTRAL) {
    const double sl = monocentral(s[3] - s[2], s[2] - s[1]);
    up = s[2] + sl / 2;
    lo = s[2] - sl / 2;
  } else if (M == ASX_RECON_PPM) {
    ppm_cell(s[0], s[1], s[2], s[3], s[4], ps[0], ps[1], ps[3], ps[4], vs[1], vs[3],
             is_rho, rp, lo, up);
  } else if (M == ASX_RECON_WENOZ) {
    wenoz_cell(s[0], s[1], s[2], s[3], s[4], rp.weno_eps, lo, up);
  } else {
    up = mp5_side(s[0], s[1], s[2], s[3], s[4], rp.mp5_alpha);
    lo = mp5_side(s[4], s[3], s[2], s[1], s[0], rp.mp5_alpha);
  }
}
template <int M>
ASX_DEV void cell_recon(const double *p, int oc, int sc, const double ps[5],
                        const double vs[5], bool is_rho, const RParams &rp,
                        double &lo, double &up) {
  double s[5];
  cell_stencil<M>(p, oc, sc, s);
  cell_recon_s<M>(s, ps, vs, is_rho, rp, lo, up);
}
// The low-order method m (Godunov, minmod or monocentral) on a stencil that
// holds at least s[1..3].
ASX_DEV void cell_recon_lo_s(int m, const double s[5], double &lo, double &up) {
  const double z[5] = {0, 0, 0, 0, 0};
  const RParams rp{};
  if (m == ASX_RECON_MINMOD)
    cell_recon_s<ASX_RECON_MINMOD>(s, z, z, false, rp, lo, up);
  else if (m == ASX_RECON_MONOCENTRAL)
    cell_recon_s<ASX_RECON_MONOCENTRAL>(s, z, z, false, rp, lo, up);
  else
    cell_recon_s<ASX_RECON_GODUNOV>(s, z, z, false, rp, lo, up);
}
ASX_DEV void cell_recon_lo(int m, const double *p, int oc, int sc, double &lo,
                           double &up) {
  double s[5];
  cell_stencil<ASX_RECON_MINMOD>(p, oc, sc, s);
  cell_recon_lo_s(m, s, lo, up);
}

// Both sides of cell oc with the low-order fallback of fluxes.cxx:258-330:
// a side whose rho, entropy, Ye or pressure/temperature is not positive is
// reconstructed again, every variable, with the low-order method.
struct NoHook {
  ASX_DEV void operator()() const {}
};
// `mid` runs once every stencil load has been issued, before any is used
// (the face tile writes its staged metric to LDS there).
template <int D, int M, class Mid = NoHook>
ASX_DEV void recon_cell(const KPatch &P, const FluxArgs &A, int oc, Side &lo, Side &up,
                        bool &flo, bool &fup, Mid &&mid = NoHook{}) {
  constexpr int D1 = (D + 1) % 3, D2 = (D + 2) % 3;
  const int sc = step(P, 0, D);
  double ps[5] = {0, 0, 0, 0, 0}, vs[5] = {0, 0, 0, 0, 0};
  if (M == ASX_RECON_PPM) {
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      ps[m] = P.g[ASX_PRESS][oc + (m - 2) * sc];
      vs[m] = P.g[ASX_VELX + D][oc + (m - 2) * sc];
    }
  }
  const int thermo = A.with_T ? ASX_TEMPERATURE : ASX_PRESS;
  const int v0 = A.rec_var == ASX_RECVAR_V   ? ASX_VELX
                 : A.rec_var == ASX_RECVAR_Z ? ASX_ZVEC_X
                                             : ASX_SVEC_X;
  // every stencil is loaded before any is used: the fallback branches below
  // would otherwise hold each variable's loads back until the previous
  // variable is done.  The stencils cover s[1..3] (what the low-order
  // fallback reads) for every method but Godunov.
  const int slot[kSideVars] = {ASX_RHO, ASX_ENTROPY, ASX_YE, thermo, ASX_BVECX + D1,
                               ASX_BVECX + D2, v0, v0 + 1, v0 + 2};
  double st[kSideVars][5];
#pragma unroll
  for (int q = 0; q < kSideVars; ++q) cell_stencil<M>(P.g[slot[q]], oc, sc, st[q]);
  mid();
  double *const lv[kSideVars] = {&lo.rho, &lo.ent, &lo.ye, &lo.th, &lo.b1,
                                 &lo.b2, &lo.v[0], &lo.v[1], &lo.v[2]};
  double *const uv[kSideVars] = {&up.rho, &up.ent, &up.ye, &up.th, &up.b1,
                                 &up.b2, &up.v[0], &up.v[1], &up.v[2]};
#pragma unroll
  for (int q = 0; q < 4; ++q) cell_recon_s<M>(st[q], ps, vs, q == 0, A.rp, *lv[q], *uv[q]);
  flo = (lo.rho <= 0.0) || (lo.ent <= 0.0) || (lo.ye <= 0.0) || (lo.th <= 0.0);
  fup = (up.rho <= 0.0) || (up.ent <= 0.0) || (up.ye <= 0.0) || (up.th <= 0.0);
  const bool any = flo || fup;
  // a side whose rho, entropy, Ye or pressure/temperature is not positive
  // takes the low-order method for every variable
  auto fix = [&](int q) {
    if (any) {
      double a, b;
      if (M == ASX_RECON_GODUNOV) cell_recon_lo(A.lo_method, P.g[slot[q]], oc, sc, a, b);
      else cell_recon_lo_s(A.lo_method, st[q], a, b);
      if (flo) *lv[q] = a;
      if (fup) *uv[q] = b;
    }
  };
#pragma unroll
  for (int q = 0; q < 4; ++q) fix(q);
#pragma unroll
  for (int q = 4; q < kSideVars; ++q) {
    cell_recon_s<M>(st[q], ps, vs, false, A.rp, *lv[q], *uv[q]);
    fix(q);
  }
}

// Numerical flux of every conserved variable at face (i, j, k) from the
// left (0) and right (1) states.
// dB_stag on the D-face at (i, j, k): the curl of the edge vector
// potential with first differences (ComputeStaggeredB<D>,
// AsterX/src/computeBfromA.cxx:32-55).  One expression for the curl kernel
// and for the face kernels that take it inline, so both give the same value.
template <int D>
ASX_DEV double curl_of(const double *idx, double b0, double b1, double b2, double b3) {
  constexpr int J = (D + 1) % 3, K = (D + 2) % 3;
  // (A_K at +J minus at 0) / dx_J - (A_J at +K minus at 0) / dx_K
  return idx[J] * (b1 - b0) - idx[K] * (b3 - b2);
}
template <int D> ASX_DEV double bstag_value(const KPatch &P, int i, int j, int k) {
  constexpr int J = (D + 1) % 3, K = (D + 2) % 3;
  constexpr int CK = 7 ^ (1 << K), CJ = 7 ^ (1 << J); // centerings of A_K, A_J
  const double *AK = P.g[ASX_AVEC_X + K], *AJ = P.g[ASX_AVEC_X + J];
  const int ok = off(P, CK, i, j, k), oj = off(P, CJ, i, j, k);
  return curl_of<D>(P.idx, AK[ok], AK[ok + step(P, CK, J)], AJ[oj], AJ[oj + step(P, CJ, K)]);
}
// The normal field on the D-face at (i, j, k) as loaded, without a branch
// (so that its loads issue with the reconstruction's): the curl's four
// points of A when the field is taken from A, else the stored field (read
// four times).  Every point exists wherever cell (i, j, k) does.
template <int D> struct FaceBn {
  double b[4];
  ASX_DEV FaceBn(const KPatch &P, const FluxArgs &A, int i, int j, int k) {
    constexpr int J = (D + 1) % 3, K = (D + 2) % 3;
    constexpr int CK = 7 ^ (1 << K), CJ = 7 ^ (1 << J);
    const bool c = A.bstag_from_A;
    const double *AK = c ? P.g[ASX_AVEC_X + K] : P.g[ASX_DBX_STAG + D];
    const double *AJ = c ? P.g[ASX_AVEC_X + J] : AK;
    const int ok = c ? off(P, CK, i, j, k) : off(P, 1 << D, i, j, k);
    const int oj = c ? off(P, CJ, i, j, k) : ok;
    const int sJ = c ? step(P, CK, J) : 0, sK = c ? step(P, CJ, K) : 0;
    b[0] = AK[ok], b[1] = AK[ok + sJ], b[2] = AJ[oj], b[3] = AJ[oj + sK];
  }
  ASX_DEV double value(const KPatch &P, const FluxArgs &A) const {
    return A.bstag_from_A ? curl_of<D>(P.idx, b[0], b[1], b[2], b[3]) : b[0];
  }
};

// The metric at face (i, j, k) (v2f averages of the vertex lapse, shift and
// 3-metric) and the staggered normal field there.  The face tile loads the
// metric before it reconstructs and the normal field with the
// reconstruction's stencils (FaceBn).
struct FaceGeom {
  double alp, beta[3];
  Sym3 g;
  double bn;
};
template <int D>
ASX_DEV void face_geom(const KPatch &P, const FluxArgs &A, int i, int j, int k, FaceGeom &G) {
  const int ov = off(P, 7, i, j, k);
  const int svj = P.sj[7], svk = P.sk[7];
  G.alp = avg_v2f<D>(P.g[ASX_ALP], ov, svj, svk);
  G.beta[0] = avg_v2f<D>(P.g[ASX_BETAX], ov, svj, svk);
  G.beta[1] = avg_v2f<D>(P.g[ASX_BETAY], ov, svj, svk);
  G.beta[2] = avg_v2f<D>(P.g[ASX_BETAZ], ov, svj, svk);
  G.g.xx = avg_v2f<D>(P.g[ASX_GXX], ov, svj, svk);
  G.g.xy = avg_v2f<D>(P.g[ASX_GXY], ov, svj, svk);
  G.g.xz = avg_v2f<D>(P.g[ASX_GXZ], ov, svj, svk);
  G.g.yy = avg_v2f<D>(P.g[ASX_GYY], ov, svj, svk);
  G.g.yz = avg_v2f<D>(P.g[ASX_GYZ], ov, svj, svk);
  G.g.zz = avg_v2f<D>(P.g[ASX_GZZ], ov, svj, svk);
  G.bn = 0.0; // set by the face tile from FaceBn
}

template <int D, bool TAB = false>
ASX_DEV void face_flux(const KPatch &P, const FluxArgs &A, int i, int j, int k,
                       const FaceGeom &G, const Side &L, const Side &R, bool lo0, bool lo1) {
  ASX_CONTRACT
  constexpr int CF = 1 << D;      // face centering
  constexpr int D1 = (D + 1) % 3, D2 = (D + 2) % 3;
  const int ofc = off(P, CF, i, j, k);
  const IdealGas &eos = A.eos;

  // metric at the face
  const double alp = G.alp;
  const double betas[3] = {G.beta[0], G.beta[1], G.beta[2]};
  const Sym3 g = G.g;
  const double detg = det3(g);
  const double sqrtg = sqrt(detg);
  // one reciprocal for 1/sqrt(gamma) and 1/alpha (both O(1) on any slice)
  const double ras = 1.0 / (alp * sqrtg);
  const double isqrtg = alp * ras;
  const double ialp = sqrtg * ras;

  const double rho[2] = {L.rho, R.rho}, ent[2] = {L.ent, R.ent}, ye[2] = {L.ye, R.ye};
  const double th[2] = {L.th, R.th};
  const double irho[2] = {1.0 / rho[0], 1.0 / rho[1]};
  double eps[2], press[2], cs2t[2];
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    if constexpr (TAB) {
      // fluxes.cxx:298-304 and the sound speed of :555-559 from one lookup
      // at the reconstructed (rho, T, Ye); the reference's csnd from
      // (rho, eps) inverts eps back to that T within 1e-14 in log T
      A.tab.state_rt(rho[f], th[f], ye[f], eps[f], press[f], cs2t[f]);
    } else if (A.with_T) {
      eps[f] = eos.eps_rt(th[f]);
      press[f] = eos.press_rt(rho[f], th[f]);
    } else {
      press[f] = th[f];
      eps[f] = th[f] * irho[f] * eos.igm1;
    }
  }
  double rhoh[2];
  rhoh[0] = rho[0] + rho[0] * eps[0] + press[0];
  rhoh[1] = rho[1] + rho[1] * eps[1] + press[1];

  // magnetic field: normal from the staggered field, transverse recon
  double Bs[3][2];
  {
    const double val = G.bn * isqrtg;
    Bs[D][0] = val;
    Bs[D][1] = val;
    Bs[D1][0] = L.b1;
    Bs[D1][1] = R.b1;
    Bs[D2][0] = L.b2;
    Bs[D2][1] = R.b2;
  }

  double vel[3][2], vlow[3][2], W[2], iW[2];
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
  double rv[3][2];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    rv[c][0] = L.v[c];
    rv[c][1] = R.v[c];
  }
  if (A.rec_var == ASX_RECVAR_V) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      vel[c][0] = rv[c][0];
      vel[c][1] = rv[c][1];
    }
    lower2(vel, vlow);
    iW[0] = sqrt(1 - dot2(vlow, vel, 0));
    iW[1] = sqrt(1 - dot2(vlow, vel, 1));
    const double rw = 1 / (iW[0] * iW[1]); // 1/W lies in (0, 1]
    W[0] = iW[1] * rw;
    W[1] = iW[0] * rw;
  } else if (A.rec_var == ASX_RECVAR_Z) {
    double zl[3][2];
    lower2(rv, zl);
    W[0] = sqrt(1 + dot2(zl, rv, 0));
    W[1] = sqrt(1 + dot2(zl, rv, 1));
    const double rw = 1 / (W[0] * W[1]);
    iW[0] = W[1] * rw;
    iW[1] = W[0] * rw;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        vel[c][f] = rv[c][f] * iW[f];
        vlow[c][f] = zl[c][f] * iW[f];
      }
  } else {
    double sl[3][2];
    lower2(rv, sl);
    double fac[2];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const double irh = 1.0 / rhoh[f];
      W[f] = sqrt(0.5 + sqrt(0.25 + dot2(sl, rv, f) * irh * irh));
      iW[f] = 1 / W[f];
      fac[f] = iW[f] * iW[f] * irh;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        vel[c][f] = rv[c][f] * fac[f];
        vlow[c][f] = sl[c][f] * fac[f];
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
    for (int f = 0; f < 2; ++f) bl[c][f] = Bl[c][f] * iW[f] + ab0[f] * vlow[c][f];
#pragma unroll
  for (int f = 0; f < 2; ++f) bsq[f] = (B2[f] + pow2(ab0[f])) * pow2(iW[f]);

  double cs2[2], h[2], dens[2], dent[2], dhW[2], dhWb[2], ppm[2], tau[2];
  double mom[3][2], Bt[3][2];
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    if constexpr (TAB) {
      cs2[f] = cs2t[f];
    } else {
      const double cs = eos.csnd(eps[f]);
      cs2[f] = cs * cs;
    }
    h[f] = 1 + eps[f] + press[f] * irho[f];
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

  // characteristic speeds (u = (gamma^-1)^{DD}, the cofactor over det
  // gamma = sqrt(gamma)^2, by the reciprocal already at hand)
  const double u = (D == 0   ? g.yy * g.zz - pow2(g.yz)
                    : D == 1 ? g.xx * g.zz - pow2(g.xz)
                             : g.xx * g.yy - pow2(g.xy)) *
                   (isqrtg * isqrtg);
  double l0p, l0m, l1p, l1m;
  eig_sides(alp, ialp, betas[D], u, vel[D], rho, cs2, W, h, bsq, l0p, l0m, l1p, l1m);
  double cmax, cmin;
  char_extrema(l0p, l0m, l1p, l1m, cmax, cmin);
  const double cabs = smax(smax(smax(smax(0.0, fabs(l0p)), fabs(l0m)), fabs(l1p)),
                           fabs(l1m));
  const bool use_hlle = A.flux_type == ASX_FLUX_HLLE;
  const double idc = 1.0 / (cmax - cmin);
  auto num = [&](double U0, double U1, double F0, double F1) {
    return use_hlle ? hlle(cmax, cmin, idc, U0, U1, F0, F1)
                    : laxf(cabs, U0, U1, F0, F1);
  };

  double BoW[2];
  BoW[0] = Bs[D][0] * iW[0];
  BoW[1] = Bs[D][1] * iW[1];
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
    const double ia = idc; // amx + amn is cmax - cmin
    P.g[s1][ofc] = (amx * vt[D1][0] + amn * vt[D1][1]) * ia;
    P.g[s2][ofc] = (amx * vt[D2][0] + amn * vt[D2][1]) * ia;
  }
  if (A.cnt_dev) {
    const bool lo = lo0 || lo1;
    const bool bad = !(dens[0] == dens[0]) || !(dens[1] == dens[1]) ||
                     !(tau[0] == tau[0]) || !(tau[1] == tau[1]) ||
                     rho[0] < 0.0 || rho[1] < 0.0 || press[0] < 0.0 || press[1] < 0.0;
    const uint64_t mb = __ballot(bad), ml = __ballot(lo);
    if ((threadIdx.x & 63) == __builtin_ctzll(__ballot(1))) {
      if (mb) atomicAdd((unsigned long long *)&A.cnt_dev[0], (unsigned long long)__popcll(mb));
      if (ml) atomicAdd((unsigned long long *)&A.cnt_dev[1], (unsigned long long)__popcll(ml));
    }
  }
}

// Face tile shapes (x, y, z work-items; 256 per block), variant V of the
// faces normal to D.  The extent along D is one cell more than the faces the
// tile produces.
template <int X_, int Y_, int Z_> struct Tile3 {
  static constexpr int X = X_, Y = Y_, Z = Z_;
};
template <int D, int V> struct FaceTileT;
template <> struct FaceTileT<0, 0> { using T = Tile3<64, 4, 1>; };
template <> struct FaceTileT<0, 1> { using T = Tile3<64, 2, 2>; };
template <> struct FaceTileT<0, 2> { using T = Tile3<32, 4, 2>; };
template <> struct FaceTileT<1, 0> { using T = Tile3<16, 16, 1>; };
template <> struct FaceTileT<1, 1> { using T = Tile3<32, 8, 1>; };
template <> struct FaceTileT<1, 2> { using T = Tile3<16, 8, 2>; };
template <> struct FaceTileT<2, 0> { using T = Tile3<16, 1, 16>; };
template <> struct FaceTileT<2, 1> { using T = Tile3<32, 1, 8>; };
template <> struct FaceTileT<2, 2> { using T = Tile3<16, 2, 8>; };
// narrower rows: fewer idle work-items on extents such as n + 2
template <> struct FaceTileT<0, 3> { using T = Tile3<16, 4, 4>; };
template <> struct FaceTileT<0, 4> { using T = Tile3<16, 8, 2>; };
template <> struct FaceTileT<1, 3> { using T = Tile3<8, 16, 2>; };
template <> struct FaceTileT<1, 4> { using T = Tile3<8, 32, 1>; };
template <> struct FaceTileT<2, 3> { using T = Tile3<8, 2, 16>; };
template <> struct FaceTileT<2, 4> { using T = Tile3<8, 4, 8>; };
// variants 5-7: the tile's vertex metric staged in LDS (MetricBox)
template <> struct FaceTileT<0, 5> { using T = Tile3<16, 4, 4>; };
template <> struct FaceTileT<0, 6> { using T = Tile3<16, 8, 2>; };
template <> struct FaceTileT<0, 7> { using T = Tile3<8, 8, 4>; };
template <> struct FaceTileT<1, 5> { using T = Tile3<16, 8, 2>; };
template <> struct FaceTileT<1, 6> { using T = Tile3<8, 16, 2>; };
template <> struct FaceTileT<1, 7> { using T = Tile3<8, 8, 4>; };
template <> struct FaceTileT<2, 5> { using T = Tile3<16, 2, 8>; };
template <> struct FaceTileT<2, 6> { using T = Tile3<8, 2, 16>; };
template <> struct FaceTileT<2, 7> { using T = Tile3<8, 4, 8>; };
constexpr int kFaceStagedVariant = 5; // variants from here on stage the metric
template <int D, int V> using FaceTile = typename FaceTileT<D, V>::T;

// LDS of a face tile: the high side of every cell, handed to the next cell
// along D, and its low-order flag.
struct FaceLDS {
  double sh[kSideVars][256];
  int shf[256];
};

// The vertex metric a face tile reads, staged in LDS: the ten fields (lapse,
// shift, 3-metric) on the box of vertices the tile's faces average over --
// along D the low-face vertex of every cell but the first, across it the
// tile's cells plus one.  Loaded cooperatively (x fastest, about 16 loads
// per work-item instead of 40) at the start of the tile, written to LDS
// after the reconstruction and summed after the hand-over barrier, so the
// metric's latency overlaps the reconstruction's.  Only for tile shapes
// (variants kFaceStagedVariant and up), whose box leaves room for three
// workgroups per CU (<= 33 KB beside the hand-over); the others load the
// metric per face.
template <int D, int V> struct MetricBox {
  using T = FaceTile<D, V>;
  static constexpr int BX = D == 0 ? T::X - 1 : T::X + 1;
  static constexpr int BY = D == 1 ? T::Y - 1 : T::Y + 1;
  static constexpr int BZ = D == 2 ? T::Z - 1 : T::Z + 1;
  // rows padded to an even length when that still fits: each work-item then
  // loads an x-adjacent pair of vertices per field (one 16-byte load)
  static constexpr int BXE = BX + (BX & 1);
  static constexpr bool PAIR = 10 * BXE * BY * BZ * 8 <= 33 * 1024;
  static constexpr int BXS = PAIR ? BXE : BX; // LDS row stride
  static constexpr int VB = BXS * BY * BZ, N = 10 * VB;
  static constexpr int NV = PAIR ? (VB / 2 + 255) / 256 : (VB + 255) / 256;
  static constexpr bool ok = V >= kFaceStagedVariant;
  static_assert(!ok || N * 8 <= 33 * 1024, "staged metric box too large for 3 workgroups per CU");
};

// One tile of a face box (lo, cnt), tile index b of a grid of ntx x nty x ntz
// tiles (ordered x, y, z).  Every work-item of the workgroup calls it; the
// caller owns the LDS (one FaceLDS serves tiles of every direction).
template <int D, int M, int V, bool TAB = false>
ASX_DEV void face_tile(const KPatch &P, const FluxArgs &A, const int lo3[3], const int cnt[3],
                       int b, int ntx, int nty, FaceLDS &S, double *MS = nullptr) {
  // MS: LDS for MetricBox<D, V>::N doubles, or null (metric loaded per face)
  using T = FaceTile<D, V>;
  static_assert(T::X * T::Y * T::Z == 256, "face tile");
  constexpr int TD = D == 0 ? T::X : (D == 1 ? T::Y : T::Z);
  constexpr int SD = D == 0 ? 1 : (D == 1 ? T::X : T::X * T::Y); // tile stride along D
  double (*sh)[256] = S.sh;
  int *shf = S.shf;
  const int bx = b % ntx, by = (b / ntx) % nty, bz = b / (ntx * nty);
  const int t = threadIdx.x;
  const int tx = t % T::X, ty = (t / T::X) % T::Y, tz = t / (T::X * T::Y);
  // cell (and face) coordinates relative to the face box; along D the tile
  // starts one cell below its first face
  const int fx = bx * (D == 0 ? TD - 1 : T::X) + tx - (D == 0);
  const int fy = by * (D == 1 ? TD - 1 : T::Y) + ty - (D == 1);
  const int fz = bz * (D == 2 ? TD - 1 : T::Z) + tz - (D == 2);
  const int td = D == 0 ? tx : (D == 1 ? ty : tz);
  const int fd = D == 0 ? fx : (D == 1 ? fy : fz);
  const bool need = fx < cnt[0] && fy < cnt[1] && fz < cnt[2];
  const int i = lo3[0] + fx, j = lo3[1] + fy, k = lo3[2] + fz;
  const bool flux = need && td >= 1 && fd >= 0;
  using MB = MetricBox<D, V>;
  const bool staged = MB::ok && MS != nullptr;
  // staged: this work-item's share of the tile's vertex metric (vertices
  // t and t + 256 of the box, every field), issued before the
  // reconstruction's loads; indices clamped to the box and the arrays, so no
  // branch (elements no face reads may hold any in-range value)
  constexpr int NV = MB::ok ? MB::NV : 1;
  constexpr int NE = MB::ok && MB::PAIR ? 2 : 1; // vertices per load
  double mv[10][NV][NE];
  int me[NV];
  bool hi[NV]; // a pair starting at the last vertex: loaded one vertex lower
  if (staged) {
    const int org[3] = {lo3[0] + bx * (D == 0 ? TD - 1 : T::X),
                        lo3[1] + by * (D == 1 ? TD - 1 : T::Y),
                        lo3[2] + bz * (D == 2 ? TD - 1 : T::Z)};
    const int ex = P.n[0] + 2 * P.ng, ey = P.n[1] + 2 * P.ng, ez = P.n[2] + 2 * P.ng;
    int mo[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      // first LDS element of this load (x fastest; pairs start at even x)
      const int e = NE * min(q * 256 + t, MB::VB / NE - 1);
      const int x = org[0] + e % MB::BXS;
      const int vy = min(org[1] + (e / MB::BXS) % MB::BY, ey);
      const int vz = min(org[2] + e / (MB::BXS * MB::BY), ez);
      const int vx = min(x, ex - (NE - 1));
      hi[q] = NE == 2 && x > vx;
      me[q] = e;
      mo[q] = vx + vy * P.sj[7] + vz * P.sk[7];
    }
#pragma unroll
    for (int f = 0; f < 10; ++f)
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const double *g = P.g[ASX_ALP + f] + mo[q];
        // an x-adjacent pair is one 16-byte load
#pragma unroll
        for (int h = 0; h < NE; ++h) mv[f][q][h] = g[h];
      }
  }
  FaceGeom G;
  if (flux && !staged) face_geom<D>(P, A, i, j, k, G);
  Side lo, up;
  bool flo = false, fup = false;
  double bn = 0.0;
  // the staged values go to LDS once the stencils' loads are out (their
  // registers are then free for the reconstruction)
  auto put = [&] {
    if (staged) {
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int f = 0; f < 10; ++f)
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          // a pair moved down one vertex: its second value is the
          // requested first one (the second slot is then past the array
          // and read by no face)
          MS[f * MB::VB + me[q]] = hi[q] ? mv[f][q][NE - 1] : mv[f][q][0];
          if (NE == 2) MS[f * MB::VB + me[q] + 1] = mv[f][q][NE - 1];
        }
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  if (need) {
    const FaceBn<D> fb(P, A, i, j, k);
    recon_cell<D, M>(P, A, off(P, 0, i, j, k), lo, up, flo, fup, put);
    bn = fb.value(P, A);
    sh[0][t] = up.rho;
    sh[1][t] = up.ent;
    sh[2][t] = up.ye;
    sh[3][t] = up.th;
    sh[4][t] = up.b1;
    sh[5][t] = up.b2;
    sh[6][t] = up.v[0];
    sh[7][t] = up.v[1];
    sh[8][t] = up.v[2];
    shf[t] = fup;
  } else {
    put();
  }
  __syncthreads();
  if (flux) {
    if (staged) {
      // avg_v2f of each field from LDS, in its order of summation
      constexpr int J = (D + 1) % 3, K = (D + 2) % 3;
      constexpr int s1 = J == 0 ? 1 : (J == 1 ? MB::BXS : MB::BXS * MB::BY);
      constexpr int s2 = K == 0 ? 1 : (K == 1 ? MB::BXS : MB::BXS * MB::BY);
      const int c0 = ((tz - (D == 2)) * MB::BY + (ty - (D == 1))) * MB::BXS + (tx - (D == 0));
      double v[10];
#pragma unroll
      for (int f = 0; f < 10; ++f) {
        const double *m = MS + f * MB::VB + c0;
        double s = 0.0;
        s += m[0];
        s += m[s1];
        s += m[s2];
        s += m[s1 + s2];
        v[f] = s / 4.0;
      }
      G.alp = v[0];
      G.beta[0] = v[1], G.beta[1] = v[2], G.beta[2] = v[3];
      G.g = Sym3{v[4], v[5], v[6], v[7], v[8], v[9]};
    }
    const int tl = t - SD;
    Side L;
    L.rho = sh[0][tl];
    L.ent = sh[1][tl];
    L.ye = sh[2][tl];
    L.th = sh[3][tl];
    L.b1 = sh[4][tl];
    L.b2 = sh[5][tl];
    L.v[0] = sh[6][tl];
    L.v[1] = sh[7][tl];
    L.v[2] = sh[8][tl];
    G.bn = bn;
    face_flux<D, TAB>(P, A, i, j, k, G, L, lo, shf[tl] != 0, flo);
  }
}

#ifndef ASX_FACE_WAVES
#define ASX_FACE_WAVES 3 // waves per SIMD the register budget is sized for
#endif
// the table's face kernels keep the lookup's 8 corners x 3 variables and
// the temperature bracket live beside the reconstruction: a 256-register
// budget (2 waves per SIMD) instead of spilling ~300 B per lane at 3
#define ASX_FACE_WAVES_TAB 2
// CalcAuxForAvecPsi (AsterX/src/fluxes.cxx:803-825) at vertex (i, j, k)
ASX_DEV void aux_point(const KPatch &P, int i, int j, int k) {
  ASX_CONTRACT
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
ASX_DEV void bstag_point(const KPatch &P, int i, int j, int k, double *out) {
  out[off(P, 1 << D, i, j, k)] = bstag_value<D>(P, i, j, k);
}
ASX_DEV bool batch_point(const BatchBox &bb, int chunk, int &i, int &j, int &k) {
  const int t = chunk * 256 + (int)threadIdx.x;
  const int nxy = bb.cnt[0] * bb.cnt[1];
  if (t >= nxy * bb.cnt[2]) return false;
  i = bb.lo[0] + t % bb.cnt[0];
  j = bb.lo[1] + (t / bb.cnt[0]) % bb.cnt[1];
  k = bb.lo[2] + t / nxy;
  return true;
}


} // namespace asx
