// Device-side building blocks of the GRMHD substep for gfx950.
//
// Per-point physics of the reference (ReconX, AsterX fluxes/eigenvalues,
// EOSX ideal gas, AsterUtils contractions) as __device__ functions.  The
// expressions keep the reference's evaluation order term for term and the
// library is compiled with -ffp-contract=off, so results agree bit for bit
// with a strict-IEEE CPU evaluation of the same formulas wherever only
// +,-,*,/,sqrt are involved (all correctly rounded on CDNA4).  Deliberate
// exceptions, which agree to rounding level: finite differences multiply by
// a host-computed 1/dx, and the source terms divide by alpha and W once.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "asterx.h"

namespace asx {

#define ASX_DEV __device__ __forceinline__
// The library builds with -ffp-contract=off so that con2prim and the
// reconstruction (whose limiter branches compare computed values) evaluate
// their expressions exactly as written.  The smooth arithmetic after the
// reconstruction -- Riemann fluxes, characteristic speeds, source terms,
// flux divergence, gauge terms, RK combination -- opts in to fused
// multiply-add with ASX_CONTRACT at the top of a function body: it changes
// results by rounding only (no branch depends on it), well inside the
// 1e-12 parity bar, and removes about a third of the FP64 instructions.
#define ASX_CONTRACT _Pragma("clang fp contract(fast)")

// ---------------------------------------------------------------------------
// Patch view passed to kernels by value
// ---------------------------------------------------------------------------
struct KPatch {
  double *g[ASX_NUM_GF];
  int sj[8], sk[8]; // dense strides per centering code (bit d: vertex on d)
  int n[3], ng;
  double dx[3], x0[3];
  double idx[3]; // 1/dx
};

// Centering code of each slot (include/asterx.h enum order).
ASX_DEV constexpr int cent(int s) {
  return (s <= ASX_KZZ)             ? 7
         : (s <= ASX_MASK_CC)       ? 0
         : (s == ASX_MASK_VC)       ? 7
         : (s == ASX_DBX_STAG)      ? 1
         : (s == ASX_DBY_STAG)      ? 2
         : (s == ASX_DBZ_STAG)      ? 4
         : (s == ASX_AVEC_X)        ? 6
         : (s == ASX_AVEC_Y)        ? 5
         : (s == ASX_AVEC_Z)        ? 3
         : (s == ASX_PSI)           ? 7
         : (s <= ASX_FXBZ)          ? 1
         : (s <= ASX_FYBZ)          ? 2
         : (s <= ASX_FZBZ)          ? 4
         : (s <= ASX_VTILDE_Z_XFACE) ? 1
         : (s <= ASX_VTILDE_Z_YFACE) ? 2
         : (s <= ASX_VTILDE_Y_ZFACE) ? 4
         : (s <= ASX_AMIN_XFACE)    ? 1
         : (s <= ASX_AMIN_YFACE)    ? 2
         : (s <= ASX_AMIN_ZFACE)    ? 4
         : (s <= ASX_G)             ? 7
         : (s <= ASX_DYE_RHS)       ? 0
         : (s == ASX_AVEC_X_RHS)    ? 6
         : (s == ASX_AVEC_Y_RHS)    ? 5
         : (s == ASX_AVEC_Z_RHS)    ? 3
         : (s == ASX_REGRID_ERROR)  ? 0
                                    : 7;
}

ASX_DEV int off(const KPatch &p, int c, int i, int j, int k) {
  return i + p.sj[c] * j + p.sk[c] * k;
}
template <int S> ASX_DEV double &at(const KPatch &p, int i, int j, int k) {
  return p.g[S][off(p, cent(S), i, j, k)];
}
ASX_DEV double &at(const KPatch &p, int s, int i, int j, int k) {
  return p.g[s][off(p, cent(s), i, j, k)];
}
// Linear offset of one step along axis d for centering c.
ASX_DEV int step(const KPatch &p, int c, int d) {
  return d == 0 ? 1 : (d == 1 ? p.sj[c] : p.sk[c]);
}

// ---------------------------------------------------------------------------
// Scalar helpers (std::min/max semantics: first operand kept on ties/NaN)
// ---------------------------------------------------------------------------
ASX_DEV double smin(double a, double b) { return (b < a) ? b : a; }
ASX_DEV double smax(double a, double b) { return (a < b) ? b : a; }
ASX_DEV double pow2(double x) { return x * x; }
ASX_DEV int sgn(double v) { return (0.0 < v) - (v < 0.0); }

// Workgroups are handed to the 8 XCDs of the chip round-robin (w % 8), and
// each XCD has its own L2.  Map workgroup w to a tile index such that every
// XCD receives one contiguous run of tiles, so neighbouring tiles (sharing
// halo cells and metric vertices) hit the same L2.  A bijection of [0, nb).
ASX_DEV int xcd_tile(int w, int nb) {
  constexpr int NX = 8;
  const int x = w % NX, l = w / NX, per = (nb + NX - 1) / NX, rem = nb % NX;
  return (rem == 0 || x < rem) ? x * per + l : rem * per + (x - rem) * (per - 1) + l;
}

struct Sym3 {
  double xx, xy, xz, yy, yz, zz;
  ASX_DEV double operator()(int i, int j) const {
    const int a = i < j ? i : j, b = i < j ? j : i;
    return a == 0 ? (b == 0 ? xx : (b == 1 ? xy : xz))
                  : (a == 1 ? (b == 1 ? yy : yz) : zz);
  }
};
// 3x3 symmetric determinant and inverse (cofactors over det)
ASX_DEV double det3(const Sym3 &g) {
  return 2 * g.xy * g.xz * g.yz + g.xx * g.yy * g.zz - g.yy * pow2(g.xz) -
         g.xx * pow2(g.yz) - g.zz * pow2(g.xy);
}
ASX_DEV Sym3 inv3(const Sym3 &g, double d) {
  return Sym3{(g.yy * g.zz - pow2(g.yz)) / d, (g.xz * g.yz - g.xy * g.zz) / d,
              (g.xy * g.yz - g.xz * g.yy) / d, (g.xx * g.zz - pow2(g.xz)) / d,
              (g.xy * g.xz - g.xx * g.yz) / d, (g.xx * g.yy - pow2(g.xy)) / d};
}
ASX_DEV void lower(const Sym3 &g, const double v[3], double o[3]) {
  o[0] = 0.0 + g.xx * v[0] + g.xy * v[1] + g.xz * v[2];
  o[1] = 0.0 + g.xy * v[0] + g.yy * v[1] + g.yz * v[2];
  o[2] = 0.0 + g.xz * v[0] + g.yz * v[1] + g.zz * v[2];
}
ASX_DEV double dot3(const double a[3], const double b[3]) {
  return 0.0 + a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// ---------------------------------------------------------------------------
// Reconstruction (ReconX)
// ---------------------------------------------------------------------------
struct RParams {
  int ppm_shock_detection, ppm_zone_flattening;
  double kgamma; // poly_k * poly_gamma (PPM contact test uses the ID polytrope)
  double ppm_eta1, ppm_eta2, ppm_eps, ppm_eps_shock, ppm_small, ppm_omega1,
      ppm_omega2, weno_eps, mp5_alpha;
};

// minmod, ReconX/src/minmod.hxx:14-26 (signbit test: -0 differs from +0)
ASX_DEV double minmod(double x, double y) {
  if (__builtin_signbit(x) != __builtin_signbit(y)) return 0.0;
  if (fabs(x) < fabs(y)) return x;
  return y;
}
ASX_DEV double monocentral(double x, double y) {
  if (sgn(x) != sgn(y)) return 0.0;
  return sgn(x) * smin(2 * fabs(x), smin(2 * fabs(y), fabs(x + y) / 2));
}

// PPM parabola of the centre cell from 5 values (ReconX/src/ppm.hxx:23-162)
ASX_DEV void ppm_cell(double gImm, double gIm, double gI, double gIp,
                      double gIpp, double pImm, double pIm, double pIp,
                      double pIpp, double vIm, double vIp, bool is_rho,
                      const RParams &rp, double &lo, double &up) {
  const double delta_Im = 0.5 * (gI - gImm);
  const double diff_I = gIp - gIm;
  const double delta_I = 0.5 * diff_I;
  const double delta_Ip = 0.5 * (gIpp - gI);
  const double f_Im_Imm = 2 * fabs(gIm - gImm);
  const double f_I_Im = 2 * fabs(gI - gIm);
  const double f_Ip_I = 2 * fabs(gIp - gI);
  const double f_Ipp_Ip = 2 * fabs(gIpp - gIp);
  const bool ss_Im = ((gI - gIm) * (gIm - gImm) > 0);
  const bool ss_I = ((gIp - gI) * (gI - gIm) > 0);
  const bool ss_Ip = ((gIpp - gIp) * (gIp - gI) > 0);
  const double dm_Im =
      ss_Im ? sgn(delta_Im) * smin(fabs(delta_Im), smin(f_Im_Imm, f_I_Im)) : 0;
  const double dm_I =
      ss_I ? sgn(delta_I) * smin(fabs(delta_I), smin(f_I_Im, f_Ip_I)) : 0;
  const double dm_Ip =
      ss_Ip ? sgn(delta_Ip) * smin(fabs(delta_Ip), smin(f_Ip_I, f_Ipp_Ip)) : 0;
  lo = 0.5 * (gIm + gI) + (1. / 6.) * (dm_Im - dm_I);
  up = 0.5 * (gI + gIp) + (1. / 6.) * (dm_I - dm_Ip);
  const double diff_press_I = pIp - pIm;
  const double min_press_I = smin(pIm, pIp);
  if (rp.ppm_shock_detection && is_rho) {
    const bool contact = (rp.kgamma * fabs(diff_I) * min_press_I >=
                          fabs(diff_press_I) * smin(gIp, gIm));
    if (contact) {
      const double d2_Im = gI - 2 * gIm + gImm;
      const double d2_Ip = gIpp - 2 * gIp + gI;
      const bool cond2 =
          (fabs(diff_I) - rp.ppm_eps_shock * smin(fabs(gIp), fabs(gIm))) > 0.;
      const double eta_t = ((d2_Im * d2_Ip < 0) && cond2)
                               ? ((-1. / 6.) * (d2_Ip - d2_Im) / diff_I)
                               : 0;
      const double eta = smax(0., smin(rp.ppm_eta1 * (eta_t - rp.ppm_eta2), 1.));
      lo = (1 - eta) * lo + eta * (gIm + 0.5 * dm_Im);
      up = (1 - eta) * up + eta * (gIp - 0.5 * dm_Ip);
    }
  }
  if (rp.ppm_zone_flattening) {
    const double w =
        ((fabs(diff_press_I) > rp.ppm_eps * min_press_I) && (vIm > vIp)) ? 1. : 0;
    const double ft =
        (fabs(pIpp - pImm) < rp.ppm_small)
            ? 1.0
            : smax(0., 1. - w * smax(0., rp.ppm_omega2 *
                                             ((diff_press_I / (pIpp - pImm)) -
                                              rp.ppm_omega1)));
    const double omf = (1 - ft) * gI;
    lo = ft * lo + omf;
    up = ft * up + omf;
  }
  if ((up - gI) * (gI - lo) <= 0) {
    lo = gI;
    up = gI;
  } else {
    const double d = up - lo;
    const double d2 = d * d;
    const double g6 = 6 * d * (gI - 0.5 * (lo + up));
    if (g6 > d2) {
      lo = 3 * gI - 2 * up;
    } else if (g6 < -d2) {
      up = 3 * gI - 2 * lo;
    }
  }
}

// WENO-Z face values of the centre cell (ReconX/src/wenoz.hxx:23-99).
// Same smoothness indicators, weights and stencils; the nonlinear weights
// a_k = 1 + tau5/(b_k + eps) are carried over the common denominator
// prod_k (b_k + eps), and the two face values share one reciprocal, so a
// cell costs one division instead of the reference's fourteen (identical up
// to rounding).  The (b_k + eps) are first
// rescaled by a power of two (exact) so the triple products cannot overflow.
ASX_DEV void wenoz_cell(double gImm, double gIm, double gI, double gIp,
                        double gIpp, double weno_eps, double &m, double &p) {
  const double b0 = (13.0 / 12.0) * pow2(gImm - 2.0 * gIm + gI) +
                    (1.0 / 4.0) * pow2(gImm - 4.0 * gIm + 3.0 * gI);
  const double b1 = (13.0 / 12.0) * pow2(gIm - 2.0 * gI + gIp) +
                    (1.0 / 4.0) * pow2(gIm - gIp);
  const double b2 = (13.0 / 12.0) * pow2(gI - 2.0 * gIp + gIpp) +
                    (1.0 / 4.0) * pow2(3.0 * gI - 4.0 * gIp + gIpp);
  double s0 = b0 + weno_eps, s1 = b1 + weno_eps, s2 = b2 + weno_eps;
  double tau5 = fabs(b0 - b2);
  const int e = -ilogb(smax(s0, smax(s1, s2)));
  s0 = ldexp(s0, e);
  s1 = ldexp(s1, e);
  s2 = ldexp(s2, e);
  tau5 = ldexp(tau5, e);
  // a_k * s0 s1 s2
  const double n0 = (s0 + tau5) * (s1 * s2);
  const double n1 = (s1 + tau5) * (s0 * s2);
  const double n2 = (s2 + tau5) * (s0 * s1);
  const double w0 = 3.0 / 10.0, w1 = 3.0 / 5.0, w2 = 1.0 / 10.0;
  const double am0 = w0 * n0, am1 = w1 * n1, am2 = w2 * n2;
  const double ap0 = w2 * n0, ap1 = w1 * n1, ap2 = w0 * n2;
  const double nm = am2 * (2.0 * gIpp - 7.0 * gIp + 11.0 * gI) +
                    am1 * (-1.0 * gIp + 5.0 * gI + 2.0 * gIm) +
                    am0 * (2.0 * gI + 5.0 * gIm - 1.0 * gImm);
  const double np = ap0 * (2.0 * gImm - 7.0 * gIm + 11.0 * gI) +
                    ap1 * (-1.0 * gIm + 5.0 * gI + 2.0 * gIp) +
                    ap2 * (2.0 * gI + 5.0 * gIp - 1.0 * gIpp);
  const double dm = 6.0 * (am0 + am1 + am2), dp = 6.0 * (ap0 + ap1 + ap2);
  // one reciprocal for both sides: after the rescaling the weights' sums
  // lie within ~1e-130 .. 1e2, so their product neither overflows nor
  // underflows
  const double r = 1.0 / (dm * dp);
  m = nm * (dp * r);
  p = np * (dm * r);
}

// MP5 value on the gIp side of cell gI (ReconX/src/mp5.hxx:26-77)
ASX_DEV double mp5_side(double gImm, double gIm, double gI, double gIp,
                        double gIpp, double alpha) {
  const double ul =
      (2 * gImm - 13 * gIm + 47 * gI + 27 * gIp - 3 * gIpp) * (1.0 / 60.0);
  const double deltam = gI - gIm;
  const double ump = gI + minmod(gIp - gI, alpha * deltam);
  if ((ul - gI) * (ul - ump) <= 1.0e-10) return ul;
  const double dm = gImm + gI - 2 * gIm;
  const double d = gIm + gIp - 2 * gI;
  const double dp = gI + gIpp - 2 * gIp;
  const double dmp = minmod(minmod(4 * d - dp, 4 * dp - d), minmod(d, dp));
  const double dmm = minmod(minmod(4 * dm - d, 4 * d - dm), minmod(dm, d));
  const double uul = gI + alpha * deltam;
  const double umd = (gI + gIp) / 2.0 - dmp / 2.0;
  const double ulc = gI + deltam / 2.0 + 4.0 / 3.0 * dmm;
  const double umin = smax(smin(gI, smin(gIp, umd)), smin(gI, smin(uul, ulc)));
  const double umax = smin(smax(gI, smax(gIp, umd)), smax(gI, smax(uul, ulc)));
  return ul + minmod(umin - ul, umax - ul);
}

// Face states from a 6-point stencil s[0..5] = cells I-3 .. I+2 along the
// face normal (ReconX/src/reconstruct.hxx:36-103): out0 from cell I-1,
// out1 from cell I.
template <int M>
ASX_DEV void recon(const double s[6], const double ps[6], const double vs[6],
                   bool is_rho, const RParams &rp, double &o0, double &o1) {
  if (M == ASX_RECON_GODUNOV) {
    o0 = s[2];
    o1 = s[3];
  } else if (M == ASX_RECON_MINMOD) {
    o0 = s[2] + minmod(s[3] - s[2], s[2] - s[1]) / 2;
    o1 = s[3] - minmod(s[4] - s[3], s[3] - s[2]) / 2;
  } else if (M == ASX_RECON_MONOCENTRAL) {
    o0 = s[2] + monocentral(s[3] - s[2], s[2] - s[1]) / 2;
    o1 = s[3] - monocentral(s[4] - s[3], s[3] - s[2]) / 2;
  } else if (M == ASX_RECON_PPM) {
    double a0, a1, b0, b1;
    ppm_cell(s[0], s[1], s[2], s[3], s[4], ps[0], ps[1], ps[3], ps[4], vs[1],
             vs[3], is_rho, rp, a0, a1);
    ppm_cell(s[1], s[2], s[3], s[4], s[5], ps[1], ps[2], ps[4], ps[5], vs[2],
             vs[4], is_rho, rp, b0, b1);
    o0 = a1;
    o1 = b0;
  } else if (M == ASX_RECON_WENOZ) {
    double a0, a1, b0, b1;
    wenoz_cell(s[0], s[1], s[2], s[3], s[4], rp.weno_eps, a0, a1);
    wenoz_cell(s[1], s[2], s[3], s[4], s[5], rp.weno_eps, b0, b1);
    o0 = a1;
    o1 = b0;
  } else { // MP5, right state from the mirrored stencil
    o0 = mp5_side(s[0], s[1], s[2], s[3], s[4], rp.mp5_alpha);
    o1 = mp5_side(s[5], s[4], s[3], s[2], s[1], rp.mp5_alpha);
  }
}
// Lowest stencil point m (of I-3+m) each method reads.
template <int M> ASX_DEV constexpr int stencil_lo() {
  return M == ASX_RECON_GODUNOV ? 2
         : (M == ASX_RECON_MINMOD || M == ASX_RECON_MONOCENTRAL) ? 1
                                                                  : 0;
}
ASX_DEV int stencil_lo_rt(int m) {
  return m == ASX_RECON_GODUNOV ? 2
         : (m == ASX_RECON_MINMOD || m == ASX_RECON_MONOCENTRAL) ? 1
                                                                  : 0;
}
// Runtime dispatch for the low-order fallback (Godunov/minmod/monocentral).
ASX_DEV void recon_lo(int m, const double s[6], double &o0, double &o1) {
  const RParams rp{};
  const double z[6] = {0, 0, 0, 0, 0, 0};
  if (m == ASX_RECON_MINMOD)
    recon<ASX_RECON_MINMOD>(s, z, z, false, rp, o0, o1);
  else if (m == ASX_RECON_MONOCENTRAL)
    recon<ASX_RECON_MONOCENTRAL>(s, z, z, false, rp, o0, o1);
  else
    recon<ASX_RECON_GODUNOV>(s, z, z, false, rp, o0, o1);
}

// ---------------------------------------------------------------------------
// Riemann fluxes and characteristic speeds (AsterX)
// ---------------------------------------------------------------------------
// Characteristic speeds (lambda+, lambda-) of the left (0) and right (1)
// states (AsterX/src/eigenvalues.hxx:19-79, the roots of a2 l^2 + a1 l + a0
// over alpha), one reciprocal serving both 1/(2 a2 alpha) (a2 > 0)
ASX_DEV void eig_sides(double alp, double ialp, double beta, double u, const double vel[2],
                       const double rho[2], const double cs2[2], const double w[2],
                       const double h[2], const double bsq[2], double &l0p, double &l0m,
                       double &l1p, double &l1m) {
  ASX_CONTRACT
  double a1[2], a2[2], sd[2];
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    const double a0 = (bsq[f] + cs2[f] * h[f] * rho[f]) * (pow2(beta) - pow2(alp) * u) -
                      (-1 + cs2[f]) * h[f] * rho[f] * pow2(beta - alp * vel[f]) * pow2(w[f]);
    a1[f] = 2 * beta * (bsq[f] + cs2[f] * h[f] * rho[f]) -
            2 * (-1 + cs2[f]) * h[f] * rho[f] * (beta - alp * vel[f]) * pow2(w[f]);
    a2[f] = bsq[f] + h[f] * rho[f] * (cs2[f] + pow2(w[f]) - cs2[f] * pow2(w[f]));
    double det = pow2(a1[f]) - 4 * a2[f] * a0;
    if (det < 0) det = 0;
    sd[f] = sqrt(det);
  }
  const double r = ialp / (4 * a2[0] * a2[1]); // 1 / (2 a2 alpha) = 2 a2' r
  const double f0 = 2 * a2[1] * r, f1 = 2 * a2[0] * r;
  l0p = (-a1[0] + sd[0]) * f0;
  l0m = (-a1[0] - sd[0]) * f0;
  l1p = (-a1[1] + sd[1]) * f1;
  l1m = (-a1[1] - sd[1]) * f1;
}
// max/min over {0, l0+, l0+, l0-, l0-, l1+, l1+, l1-, l1-} (fluxes.hxx:35-42)
ASX_DEV void char_extrema(double l0p, double l0m, double l1p, double l1m,
                          double &cmax, double &cmin) {
  cmax = smax(smax(smax(smax(0.0, l0p), l0m), l1p), l1m);
  cmin = smin(smin(smin(smin(0.0, l0p), l0m), l1p), l1m);
}
// HLLE with the common 1/(cmax - cmin) precomputed (AsterX/src/fluxes.hxx:31-50)
ASX_DEV double hlle(double cmax, double cmin, double idc, double U0, double U1,
                    double F0, double F1) {
  ASX_CONTRACT
  return (cmax * F0 - cmin * F1 + cmax * cmin * (U1 - U0)) * idc;
}
ASX_DEV double laxf(double cabs, double U0, double U1, double F0, double F1) {
  ASX_CONTRACT
  return 0.5 * ((F0 + F1) - cabs * (U1 - U0));
}

// ---------------------------------------------------------------------------
// Ideal-gas EOS (EOSX/src/eos_3p_idealgas/eos_3p_idealgas.hxx)
// ---------------------------------------------------------------------------
struct IdealGas {
  double gamma, gm1, inv_gamma, temp_over_eps;
  double eps_min, eps_max, rho_min, rho_max, ye_min, ye_max, temp_min,
      temp_max;
  double igm1; // 1/(gamma-1)
  ASX_DEV double press_re(double rho, double eps) const { return gm1 * rho * eps; }
  ASX_DEV double eps_rp(double rho, double press) const {
    return press / (rho * gm1);
  }
  ASX_DEV double csnd(double eps) const {
    return sqrt(gm1 * eps / (eps + inv_gamma));
  }
  ASX_DEV double temp_re(double eps) const { return temp_over_eps * eps; }
  ASX_DEV double eps_rt(double temp) const { return temp / temp_over_eps; }
  ASX_DEV double press_rt(double rho, double temp) const {
    return press_re(rho, eps_rt(temp));
  }
  ASX_DEV double kappa(double rho, double eps) const {
    return (gamma - 1.0) * eps * pow(rho, 1.0 - gamma);
  }
  // The 3-parameter interface shared with TabEOS (EOSX eos_3p.hxx): Ye is
  // not used by the ideal gas, and eps is never changed.
  ASX_DEV double press_re(double rho, double &eps, double, int * = nullptr) const {
    return press_re(rho, eps);
  }
  ASX_DEV double temp_re(double, double &eps, double) const { return temp_re(eps); }
  ASX_DEV double kappa_re(double rho, double &eps, double) const { return kappa(rho, eps); }
  ASX_DEV double csnd_re(double, double &eps, double) const { return csnd(eps); }
  // p and eps from (rho, kappa) (eos_3p_idealgas.hxx:201-215):
  // p = kappa rho^gamma, eps = kappa rho^(gamma-1) / (gamma-1)
  ASX_DEV double press_rk(double rho, double kap, double) const { return kap * pow(rho, gamma); }
  ASX_DEV double eps_rk(double rho, double kap, double) const {
    return kap * pow(rho, gamma - 1.0) / (gamma - 1.0);
  }
  // press_re, temp_re and kappa_re at one (rho, eps, Ye)
  ASX_DEV void state_re(double rho, double &eps, double ye, double &press, double &temp,
                        double &kap, int * = nullptr) const {
    press = press_re(rho, eps, ye);
    temp = temp_re(rho, eps, ye);
    kap = kappa_re(rho, eps, ye);
  }
  ASX_DEV double eps_rt(double, double temp, double) const { return eps_rt(temp); }
  ASX_DEV double press_rt(double rho, double temp, double) const { return press_rt(rho, temp); }
  ASX_DEV double eps_rp(double rho, double press, double) const { return eps_rp(rho, press); }
  ASX_DEV double csnd_rt(double, double temp, double) const { return csnd(eps_rt(temp)); }
  // eos_3p_idealgas.hxx: the entropy of (rho, T) is kappa(rho, eps(T))
  ASX_DEV double entropy_rt(double rho, double temp, double) const {
    return kappa(rho, eps_rt(temp));
  }
};

// v2f metric average at a face normal to D (AsterUtils aster_interp.hxx:237-249)
template <int D>
ASX_DEV double avg_v2f(const double *g, int o, int sj, int sk) {
  constexpr int J = (D + 1) % 3, K = (D + 2) % 3;
  const int s1 = J == 0 ? 1 : (J == 1 ? sj : sk);
  const int s2 = K == 0 ? 1 : (K == 1 ? sj : sk);
  double s = 0.0;
  s += g[o];
  s += g[o + s1];
  s += g[o + s2];
  s += g[o + s1 + s2];
  return s / 4.0;
}
// v2c average (aster_interp.hxx:191-205)
ASX_DEV double avg_v2c(const double *g, int o, int sj, int sk) {
  double s = 0.0;
  s += g[o];
  s += g[o + 1];
  s += g[o + sj];
  s += g[o + sj + 1];
  s += g[o + sk];
  s += g[o + sk + 1];
  s += g[o + sk + sj];
  s += g[o + sk + sj + 1];
  return s / 8.0;
}

} // namespace asx
