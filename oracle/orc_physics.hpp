// ORACLE -- test infrastructure only.  CPU restatement of the reference's
// per-point physics for the GRMHD substep (AsterX, ReconX, Con2PrimFactory,
// EOSX, AsterUtils).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline may load it; the product path never does.
//
// Every function cites the reference file:line it restates.  Operation order
// follows the reference expression by expression (evaluation left to right,
// no contraction: built with -ffp-contract=off) so that masks and flags can
// be compared bit for bit with the GPU path.
//
// Parity status per function (see the header of orc_grid.cpp for the grid
// routines):
//   pinned by reference golden TSV runs (AsterX/test/*): recon PPM, WENO-Z,
//     MP5, minmod, monocentral, Godunov; HLLE, LxF; eigenvalues; flux
//     assembly; Noble 2D; Palenzuela 1D (to solver tolerance); ideal gas.
//   pinned by closed forms / invariants in tests/test_oracle_*.py: the same,
//     plus det/inv (exact on diagonal metrics), c2p round trip.
//   parity unpinned at the bit level: calc_det / calc_inv / sum_symm
//     summation order (CarpetX Arith, not vendored), Algo::brent iterate
//     sequence (CarpetX Algo, not vendored; restated from Brent 1973).
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

namespace orc {

// std::min / std::max semantics (b < a ? b : a), also used for the
// initializer-list forms (first extreme wins).
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }
inline double pow2(double x) { return x * x; }
// ReconX/src/reconx_utils.hxx:30-33
inline int sgn(double v) { return (0.0 < v) - (v < 0.0); }

struct Sym3 {
  double xx, xy, xz, yy, yz, zz;
  double operator()(int i, int j) const {
    const int a = i < j ? i : j, b = i < j ? j : i;
    if (a == 0) return b == 0 ? xx : (b == 1 ? xy : xz);
    if (a == 1) return b == 1 ? yy : yz;
    return zz;
  }
};

// CarpetX Arith calc_det (not vendored; parity unpinned at the bit level).
inline double calc_det(const Sym3 &g) {
  return 2 * g.xy * g.xz * g.yz + g.xx * g.yy * g.zz - g.yy * pow2(g.xz) -
         g.xx * pow2(g.yz) - g.zz * pow2(g.xy);
}
// CarpetX Arith calc_inv(g, detg) (not vendored; parity unpinned).
inline Sym3 calc_inv(const Sym3 &g, double detg) {
  return Sym3{(g.yy * g.zz - pow2(g.yz)) / detg,
              (g.xz * g.yz - g.xy * g.zz) / detg,
              (g.xy * g.yz - g.xz * g.yy) / detg,
              (g.xx * g.zz - pow2(g.xz)) / detg,
              (g.xy * g.xz - g.xx * g.yz) / detg,
              (g.xx * g.yy - pow2(g.xy)) / detg};
}
// AsterUtils/src/aster_utils.hxx:27-33: (g.v)_i = sum_j g_ij v_j
inline void lower(const Sym3 &g, const double v[3], double out[3]) {
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int j = 0; j < 3; ++j) s += g(i, j) * v[j];
    out[i] = s;
  }
}
// AsterUtils/src/aster_utils.hxx:35-39
inline double dot3(const double a[3], const double b[3]) {
  double s = 0.0;
  for (int i = 0; i < 3; ++i) s += a[i] * b[i];
  return s;
}
// sum_symm: CarpetX Arith (not vendored): upper triangle, off-diagonal x2.
template <typename F> inline double sum_symm(const F &f) {
  double s = 0.0;
  for (int x = 0; x < 3; ++x)
    for (int y = x; y < 3; ++y) s += (x == y ? 1 : 2) * f(x, y);
  return s;
}
// 2-D sum<D>: CarpetX Arith (not vendored): x outer, y inner.
template <typename F> inline double sum2(const F &f) {
  double s = 0.0;
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 3; ++y) s += f(x, y);
  return s;
}

// ---------------------------------------------------------------------------
// Reconstruction (ReconX)
// ---------------------------------------------------------------------------
struct ReconParams {
  bool ppm_shock_detection, ppm_zone_flattening;
  double poly_k, poly_gamma, ppm_eta1, ppm_eta2, ppm_eps, ppm_eps_shock,
      ppm_small, ppm_omega1, ppm_omega2, weno_eps, mp5_alpha;
};

// ReconX/src/minmod.hxx:14-26
inline double minmod(double x, double y) {
  if (std::signbit(x) != std::signbit(y)) return 0.0;
  if (std::fabs(x) < std::fabs(y)) return x;
  return y;
}
// ReconX/src/minmod.hxx:28-44 -> {right state of Im, left state of Ip}
inline void minmod_rc(double qmm, double qm, double qp, double qpp,
                      double out[2]) {
  const double sp = qpp - qp, sc = qp - qm, sm = qm - qmm;
  out[0] = qm + minmod(sc, sm) / 2;
  out[1] = qp - minmod(sp, sc) / 2;
}
// ReconX/src/monocentral.hxx:17-27
inline double monocentral(double x, double y) {
  if (sgn(x) != sgn(y)) return 0.0;
  return sgn(x) * smin(2 * std::fabs(x),
                       smin(2 * std::fabs(y), std::fabs(x + y) / 2));
}
// ReconX/src/monocentral.hxx:29-45
inline void monocentral_rc(double qmm, double qm, double qp, double qpp,
                           double out[2]) {
  double sp = qp - qm, sm = qm - qmm;
  out[0] = qm + monocentral(sp, sm) / 2;
  sp = qpp - qp;
  sm = qp - qm;
  out[1] = qp - monocentral(sp, sm) / 2;
}

// ReconX/src/ppm.hxx:23-162: parabola of the centre cell I from a 5-cell
// stencil; returns {lower, upper} interface states of cell I.
inline void ppm_cell(double gImm, double gIm, double gI, double gIp,
                     double gIpp, double pImm, double pIm, double pIp,
                     double pIpp, double vIm, double vIp, bool is_rho,
                     const ReconParams &rp, double rc[2]) {
  const double diff_Im = gI - gImm;
  const double diff_I = gIp - gIm;
  const double diff_Ip = gIpp - gI;
  const double delta_Im = 0.5 * diff_Im;
  const double delta_I = 0.5 * diff_I;
  const double delta_Ip = 0.5 * diff_Ip;
  const double f_Im_Imm = 2 * std::fabs(gIm - gImm);
  const double f_I_Im = 2 * std::fabs(gI - gIm);
  const double f_Ip_I = 2 * std::fabs(gIp - gI);
  const double f_Ipp_Ip = 2 * std::fabs(gIpp - gIp);
  const bool ss_Im = ((gI - gIm) * (gIm - gImm) > 0);
  const bool ss_I = ((gIp - gI) * (gI - gIm) > 0);
  const bool ss_Ip = ((gIpp - gIp) * (gIp - gI) > 0);
  const double dm_Im =
      ss_Im ? sgn(delta_Im) * smin(std::fabs(delta_Im), smin(f_Im_Imm, f_I_Im))
            : 0;
  const double dm_I =
      ss_I ? sgn(delta_I) * smin(std::fabs(delta_I), smin(f_I_Im, f_Ip_I)) : 0;
  const double dm_Ip =
      ss_Ip ? sgn(delta_Ip) * smin(std::fabs(delta_Ip), smin(f_Ip_I, f_Ipp_Ip))
            : 0;
  const double g_Im_I = 0.5 * (gIm + gI) + (1. / 6.) * (dm_Im - dm_I);
  const double g_I_Ip = 0.5 * (gI + gIp) + (1. / 6.) * (dm_I - dm_Ip);
  double lo = g_Im_I, up = g_I_Ip;
  const double diff_press_I = pIp - pIm;
  const double min_press_I = smin(pIm, pIp);
  // ppm.hxx:94-114: contact steepening, rho only, uses poly_k*poly_gamma
  if (rp.ppm_shock_detection && is_rho) {
    const double k_gamma = rp.poly_k * rp.poly_gamma;
    const bool contact = (k_gamma * std::fabs(diff_I) * min_press_I >=
                          std::fabs(diff_press_I) * smin(gIp, gIm));
    if (contact) {
      const double d2_Im = gI - 2 * gIm + gImm;
      const double d2_Ip = gIpp - 2 * gIp + gI;
      const double d2p = d2_Im * d2_Ip;
      const bool cond2 =
          (std::fabs(diff_I) - rp.ppm_eps_shock * smin(std::fabs(gIp),
                                                       std::fabs(gIm))) > 0.;
      const double eta_t =
          ((d2p < 0) && cond2) ? ((-1. / 6.) * (d2_Ip - d2_Im) / diff_I) : 0;
      const double eta = smax(0., smin(rp.ppm_eta1 * (eta_t - rp.ppm_eta2), 1.));
      lo = (1 - eta) * lo + eta * (gIm + 0.5 * dm_Im);
      up = (1 - eta) * up + eta * (gIp - 0.5 * dm_Ip);
    }
  }
  // ppm.hxx:121-138: zone flattening (off by default)
  if (rp.ppm_zone_flattening) {
    const double w = ((std::fabs(diff_press_I) > rp.ppm_eps * min_press_I) &&
                      (vIm > vIp))
                         ? 1.
                         : 0;
    const double ft =
        (std::fabs(pIpp - pImm) < rp.ppm_small)
            ? 1.0
            : smax(0., 1. - w * smax(0., rp.ppm_omega2 *
                                             ((diff_press_I / (pIpp - pImm)) -
                                              rp.ppm_omega1)));
    const double omf = (1 - ft) * gI;
    lo = ft * lo + omf;
    up = ft * up + omf;
  }
  // ppm.hxx:141-157: monotonisation
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
  rc[0] = lo;
  rc[1] = up;
}

// ReconX/src/wenoz.hxx:23-99: {minus-side, plus-side} values of centre cell.
inline void wenoz_cell(double gImm, double gIm, double gI, double gIp,
                       double gIpp, double weno_eps, double out[2]) {
  const double b0 = (13.0 / 12.0) * pow2(gImm - 2.0 * gIm + gI) +
                    (1.0 / 4.0) * pow2(gImm - 4.0 * gIm + 3.0 * gI);
  const double b1 = (13.0 / 12.0) * pow2(gIm - 2.0 * gI + gIp) +
                    (1.0 / 4.0) * pow2(gIm - gIp);
  const double b2 = (13.0 / 12.0) * pow2(gI - 2.0 * gIp + gIpp) +
                    (1.0 / 4.0) * pow2(3.0 * gI - 4.0 * gIp + gIpp);
  const double tau5 = std::fabs(b0 - b2);
  const double a0 = 1.0 + tau5 / (b0 + weno_eps);
  const double a1 = 1.0 + tau5 / (b1 + weno_eps);
  const double a2 = 1.0 + tau5 / (b2 + weno_eps);
  const double w0 = 3.0 / 10.0, w1 = 3.0 / 5.0, w2 = 1.0 / 10.0;
  const double am0 = w0 * a0, am1 = w1 * a1, am2 = w2 * a2; // minus side
  const double ap0 = w2 * a0, ap1 = w1 * a1, ap2 = w0 * a2; // plus side
  const double dm = am0 + am1 + am2, dp = ap0 + ap1 + ap2;
  const double om0 = am0 / dm, om1 = am1 / dm, om2 = am2 / dm;
  const double op0 = ap0 / dp, op1 = ap1 / dp, op2 = ap2 / dp;
  out[0] = (om2 / 6.0) * (2.0 * gIpp - 7.0 * gIp + 11.0 * gI) +
           (om1 / 6.0) * (-1.0 * gIp + 5.0 * gI + 2.0 * gIm) +
           (om0 / 6.0) * (2.0 * gI + 5.0 * gIm - 1.0 * gImm);
  out[1] = (op0 / 6.0) * (2.0 * gImm - 7.0 * gIm + 11.0 * gI) +
           (op1 / 6.0) * (-1.0 * gIm + 5.0 * gI + 2.0 * gIp) +
           (op2 / 6.0) * (2.0 * gI + 5.0 * gIp - 1.0 * gIpp);
}

// ReconX/src/mp5.hxx:26-77: value at the face on the gIp side of cell gI.
inline double mp5_side(double gImm, double gIm, double gI, double gIp,
                       double gIpp, double alpha) {
  const double ul =
      (2 * gImm - 13 * gIm + 47 * gI + 27 * gIp - 3 * gIpp) / 60.0;
  const double deltam = gI - gIm;
  const double deltap = gIp - gI;
  const double ump = gI + minmod(deltap, alpha * deltam);
  if ((ul - gI) * (ul - ump) <= 1.0e-10) return ul; // absolute threshold
  const double dm = gImm + gI - 2 * gIm;
  const double d = gIm + gIp - 2 * gI;
  const double dp = gI + gIpp - 2 * gIp;
  const double dmp = minmod(minmod(4 * d - dp, 4 * dp - d), minmod(d, dp));
  const double dmm = minmod(minmod(4 * dm - d, 4 * d - dm), minmod(dm, d));
  const double uul = gI + alpha * deltam;
  const double uav = (gI + gIp) / 2.0;
  const double umd = uav - dmp / 2.0;
  const double ulc = gI + deltam / 2.0 + 4.0 / 3.0 * dmm;
  const double umin =
      smax(smin(gI, smin(gIp, umd)), smin(gI, smin(uul, ulc)));
  const double umax =
      smin(smax(gI, smax(gIp, umd)), smax(gI, smax(uul, ulc)));
  return ul + minmod(umin - ul, umax - ul); // median, mp5.hxx:14-17
}

enum { R_GODUNOV = 0, R_MINMOD = 1, R_MONOCENTRAL = 2, R_PPM = 3,
       R_EPPM = 4, R_WENOZ = 5, R_MP5 = 6 };

// ReconX/src/reconstruct.hxx:36-103 on a 6-point stencil
// s[0..5] = var(Immm..Ippp), p[0..5] = press(...), v[0..5] = vel_dir(...).
// Returns {left state (from cell Im), right state (from cell Ip)}.
inline void reconstruct(int method, const double s[6], const double p[6],
                        const double v[6], bool is_rho, const ReconParams &rp,
                        double out[2]) {
  switch (method) {
  case R_GODUNOV:
    out[0] = s[2];
    out[1] = s[3];
    return;
  case R_MINMOD:
    minmod_rc(s[1], s[2], s[3], s[4], out);
    return;
  case R_MONOCENTRAL:
    monocentral_rc(s[1], s[2], s[3], s[4], out);
    return;
  case R_PPM: { // ppm.hxx:165-180
    double a[2], b[2];
    ppm_cell(s[0], s[1], s[2], s[3], s[4], p[0], p[1], p[3], p[4], v[1], v[3],
             is_rho, rp, a);
    ppm_cell(s[1], s[2], s[3], s[4], s[5], p[1], p[2], p[4], p[5], v[2], v[4],
             is_rho, rp, b);
    out[0] = a[1];
    out[1] = b[0];
    return;
  }
  case R_WENOZ: { // wenoz.hxx:106-113
    double a[2], b[2];
    wenoz_cell(s[0], s[1], s[2], s[3], s[4], rp.weno_eps, a);
    wenoz_cell(s[1], s[2], s[3], s[4], s[5], rp.weno_eps, b);
    out[0] = a[1];
    out[1] = b[0];
    return;
  }
  case R_MP5: // mp5.hxx:198-211 (right state uses the mirrored stencil)
    out[0] = mp5_side(s[0], s[1], s[2], s[3], s[4], rp.mp5_alpha);
    out[1] = mp5_side(s[5], s[4], s[3], s[2], s[1], rp.mp5_alpha);
    return;
  default:
    out[0] = out[1] = std::numeric_limits<double>::quiet_NaN();
  }
}

// ---------------------------------------------------------------------------
// Riemann solvers and characteristic speeds (AsterX)
// ---------------------------------------------------------------------------
// AsterX/src/eigenvalues.hxx:19-79; lam[side][4]
inline void eigenvalues(double alp, double beta, double u, const double vel[2],
                        const double rho[2], const double cs2[2],
                        const double w[2], const double h[2],
                        const double bsq[2], double lam[2][4]) {
  for (int f = 0; f < 2; ++f) {
    const double a0 =
        (bsq[f] + cs2[f] * h[f] * rho[f]) * (pow2(beta) - pow2(alp) * u) -
        (-1 + cs2[f]) * h[f] * rho[f] * pow2(beta - alp * vel[f]) * pow2(w[f]);
    const double a1 =
        2 * beta * (bsq[f] + cs2[f] * h[f] * rho[f]) -
        2 * (-1 + cs2[f]) * h[f] * rho[f] * (beta - alp * vel[f]) * pow2(w[f]);
    const double a2 =
        bsq[f] + h[f] * rho[f] * (cs2[f] + pow2(w[f]) - cs2[f] * pow2(w[f]));
    double det = pow2(a1) - 4 * a2 * a0;
    if (det < 0) det = 0;
    const double lp = ((-a1 + std::sqrt(det)) / (2 * a2)) / alp;
    const double lm = ((-a1 - std::sqrt(det)) / (2 * a2)) / alp;
    lam[f][0] = lp;
    lam[f][1] = lp;
    lam[f][2] = lm;
    lam[f][3] = lm;
  }
}
// AsterX/src/fluxes.hxx:32-49
inline double hlle(const double lam[2][4], const double U[2],
                   const double F[2]) {
  double cmax = 0, cmin = 0;
  for (int f = 0; f < 2; ++f)
    for (int k = 0; k < 4; ++k) {
      cmax = smax(cmax, lam[f][k]);
      cmin = smin(cmin, lam[f][k]);
    }
  const double cpm = cmax - cmin;
  return (cmax * F[0] - cmin * F[1] + cmax * cmin * (U[1] - U[0])) / cpm;
}
// AsterX/src/fluxes.hxx:20-29
inline double laxf(const double lam[2][4], const double U[2],
                   const double F[2]) {
  double cmax = 0;
  for (int f = 0; f < 2; ++f)
    for (int k = 0; k < 4; ++k) cmax = smax(cmax, std::fabs(lam[f][k]));
  return 0.5 * ((F[0] + F[1]) - cmax * (U[1] - U[0]));
}

// ---------------------------------------------------------------------------
// EOS (EOSX)
// ---------------------------------------------------------------------------
// EOSX/src/eos_3p_idealgas/eos_3p_idealgas.hxx:117-237 with the ranges of
// EOSX/src/setup_eos.cxx:284-305
struct IdealGas {
  double gamma, gm1, inv_gamma, temp_over_eps;
  double eps_min, eps_max, rho_min, rho_max, ye_min, ye_max, temp_min,
      temp_max;
  void init(double gamma_, double umass, double epsmin, double epsmax,
            double rhomin, double rhomax, double yemin, double yemax) {
    gamma = gamma_;
    gm1 = gamma_ - 1;
    inv_gamma = 1 / gamma_;
    eps_min = epsmin;
    eps_max = epsmax;
    if (gamma > 2) eps_max = smin(eps_max, 1 / (gamma * (gamma - 2)));
    rho_min = rhomin;
    rho_max = rhomax;
    ye_min = yemin;
    ye_max = yemax;
    temp_over_eps = gm1 * umass;
    temp_min = temp_over_eps * eps_min;
    temp_max = temp_over_eps * eps_max;
  }
  double press_from_rho_eps(double rho, double eps) const {
    return gm1 * rho * eps;
  }
  double eps_from_rho_press(double rho, double press) const {
    return press / (rho * gm1);
  }
  double csnd_from_rho_eps(double, double eps) const {
    return std::sqrt(gm1 * eps / (eps + inv_gamma));
  }
  double temp_from_rho_eps(double, double eps) const {
    return temp_over_eps * eps;
  }
  double eps_from_rho_temp(double, double temp) const {
    return temp / temp_over_eps;
  }
  double press_from_rho_temp(double rho, double temp) const {
    return press_from_rho_eps(rho, eps_from_rho_temp(rho, temp));
  }
  // kappa = (gamma-1) eps rho^(1-gamma)  (eos_3p_idealgas.hxx:225-229)
  double kappa_from_rho_eps(double rho, double eps) const {
    return (gamma - 1.0) * eps * std::pow(rho, 1.0 - gamma);
  }  // The 3-parameter interface (EOSX eos_3p.hxx) shared with the tabulated
  // EOS: Ye unused, eps never changed by the ideal gas.
  double press_from_rho_eps_ye(double rho, double &eps, double) const {
    return press_from_rho_eps(rho, eps);
  }
  double temp_from_rho_eps_ye(double rho, double &eps, double) const {
    return temp_from_rho_eps(rho, eps);
  }
  double kappa_from_rho_eps_ye(double rho, double &eps, double) const {
    return kappa_from_rho_eps(rho, eps);
  }
  // eos_3p_idealgas.hxx:201-215: p = kappa rho^gamma,
  // eps = kappa rho^(gamma-1) / (gamma-1)
  double press_from_rho_kappa_ye(double rho, double kappa, double) const {
    return kappa * std::pow(rho, gamma);
  }
  double eps_from_rho_kappa_ye(double rho, double kappa, double) const {
    return kappa * std::pow(rho, gamma - 1.0) / (gamma - 1.0);
  }
  double csnd_from_rho_eps_ye(double rho, double &eps, double) const {
    return csnd_from_rho_eps(rho, eps);
  }
  double eps_from_rho_temp_ye(double rho, double temp, double) const {
    return eps_from_rho_temp(rho, temp);
  }
  double press_from_rho_temp_ye(double rho, double temp, double) const {
    return press_from_rho_temp(rho, temp);
  }
  double eps_from_rho_press_ye(double rho, double press, double) const {
    return eps_from_rho_press(rho, press);
  }
  double csnd_from_rho_temp_ye(double rho, double temp, double) const {
    return csnd_from_rho_eps(rho, eps_from_rho_temp(rho, temp));
  }
  double entropy_from_rho_temp_ye(double rho, double temp, double) const {
    return kappa_from_rho_eps(rho, eps_from_rho_temp(rho, temp));
  }

};

// EOSX/src/eos_1p_polytropic/eos_1p_polytropic.hxx:57-89 (atmosphere with
// thermal_eos_atmo = no)
struct Polytrope {
  double poly_k, poly_gamma, n, np1, invn, rho_p;
  void init(double gamma, double k) {
    poly_k = k;
    poly_gamma = gamma;
    n = 1.0 / (poly_gamma - 1.0);
    np1 = n + 1;
    invn = 1.0 / n;
    rho_p = std::pow(poly_k, -n);
  }
  double gm1_from_rho(double rho) const {
    return np1 * std::pow(rho / rho_p, invn);
  }
  double sed_from_gm1(double gm1) const { return gm1 / poly_gamma; }
};

} // namespace orc
