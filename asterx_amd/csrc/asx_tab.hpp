// Tabulated 3-parameter EOS for gfx950 (EOSX eos_3p_tabulated3d): a table
// of log p, log(eps + shift), c_s^2 and entropy over axes log rho, log T,
// Ye (code units; axes uniform), the variable index fastest, and the EOSX interface
// on it (EOSX/src/eos_3p_tabulated3d/eos_3p_tabulated3d.hxx:186-357):
// trilinear interpolation in the order of eos_linear_interp_ND.hxx (rho,
// then T, then Ye) and the same clamping of the inputs, including the
// change of eps the reference makes through its reference argument.  The
// temperature from (rho, eps, Ye) is the root the reference's Brent
// iteration (EOSX/src/utils/eos_brent.hxx, tolerance 1e-14 in log T)
// approaches, solved exactly on the piecewise-linear interpolant.
#pragma once

#include "asx_device.hpp"

namespace asx {

struct TabEOS {
  enum { PRESS = 0, EPS = 1, CS2 = 2, ENT = 3, NV = 4 };
  const double *tab;    // [v + NV * (ir + n0 * (it + n1 * iy))]
  const double *ax[3];  // axis values: log rho, log T, Ye
  int n[3];
  double idx[3];        // 1 / (ax[d][1] - ax[d][0])
  double shift;         // energy shift
  double rho_min, rho_max, temp_min, temp_max, ye_min, ye_max; // rgrho, rgtemp, rgye
  double eps_min, eps_max;                                     // rgeps

  // nearest lower node, uniform axis (eos_linear_interp_ND.hxx:57-73)
  ASX_DEV int find(int d, double x) const {
    const double xl = x - ax[d][0];
    if (!(xl > 0.0)) return 0;
    const double f = xl * idx[d];
    const int last = n[d] - 2;
    return f >= (double)last ? last : (int)f;
  }
  ASX_DEV static double lin(double A, double B, double xA, double xB, double x) {
    const double l = (x - xA) / (xB - xA);
    return (B - A) * l + A;
  }
  // one variable at (log rho, log T, Ye)
  ASX_DEV double interp(int v, double lr, double lt, double ye) const {
    const int i0 = find(0, lr), i1 = find(1, lt), i2 = find(2, ye);
    const int s1 = NV * n[0], s2 = NV * n[0] * n[1];
    const double *p = tab + v + NV * i0 + s1 * i1 + s2 * i2;
    const double r0 = ax[0][i0], r1 = ax[0][i0 + 1];
    const double t0 = ax[1][i1], t1 = ax[1][i1 + 1];
    const double y0 = ax[2][i2], y1 = ax[2][i2 + 1];
    const double a00 = lin(p[0], p[NV], r0, r1, lr);
    const double a10 = lin(p[s1], p[s1 + NV], r0, r1, lr);
    const double a01 = lin(p[s2], p[s2 + NV], r0, r1, lr);
    const double a11 = lin(p[s2 + s1], p[s2 + s1 + NV], r0, r1, lr);
    const double b0 = lin(a00, a10, t0, t1, lt);
    const double b1 = lin(a01, a11, t0, t1, lt);
    return lin(b0, b1, y0, y1, ye);
  }
  // std::fmin / std::fmax as the reference (a NaN argument is dropped)
  // p, eps and c_s^2 at one point (indices found once)
  ASX_DEV void interp3(double lr, double lt, double ye, double out[3]) const {
    const int i0 = find(0, lr), i1 = find(1, lt), i2 = find(2, ye);
    const int s1 = NV * n[0], s2 = NV * n[0] * n[1];
    const double *p = tab + NV * i0 + s1 * i1 + s2 * i2;
    const double r0 = ax[0][i0], r1 = ax[0][i0 + 1];
    const double t0 = ax[1][i1], t1 = ax[1][i1 + 1];
    const double y0 = ax[2][i2], y1 = ax[2][i2 + 1];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const double a00 = lin(p[v], p[v + NV], r0, r1, lr);
      const double a10 = lin(p[v + s1], p[v + s1 + NV], r0, r1, lr);
      const double a01 = lin(p[v + s2], p[v + s2 + NV], r0, r1, lr);
      const double a11 = lin(p[v + s2 + s1], p[v + s2 + s1 + NV], r0, r1, lr);
      const double b0 = lin(a00, a10, t0, t1, lt);
      const double b1 = lin(a01, a11, t0, t1, lt);
      out[v] = lin(b0, b1, y0, y1, ye);
    }
  }
  // eps_rt, press_rt and csnd_rt^2 at one (rho, T, Ye) from one lookup; the
  // pressure is looked up again at the unbounded (rho, T) when either lies
  // outside the table (press_rt takes the logs of the inputs)
  ASX_DEV void state_rt(double rho, double temp, double ye, double &eps, double &press,
                        double &cs2) const {
    const double rc = crho(rho), tc = ctemp(temp);
    double v[3];
    interp3(log(rc), log(tc), ye, v);
    eps = exp(v[EPS]) - shift;
    press = (rc == rho && tc == temp) ? exp(v[PRESS]) : press_rt(rho, temp, ye);
    const double cs = sqrt(v[CS2]);
    cs2 = cs * cs;
  }
  ASX_DEV double crho(double rho) const { return fmin(fmax(rho, rho_min), rho_max); }
  ASX_DEV double ctemp(double t) const { return fmin(fmax(t, temp_min), temp_max); }
  ASX_DEV double xmin1() const { return ax[1][0]; }
  ASX_DEV double xmax1() const { return ax[1][n[1] - 1]; }

  // logtemp_from_valid_rho_eps_ye (eos_3p_tabulated3d.hxx:186-218).  hint
  // (optional) is the T cell of the previous solve for the same zone: when
  // that cell, away from the table's ends, still brackets the value, the
  // root is solved there with two node values instead of the end values and
  // the bisection.  For a table whose eps grows with T the bracketing cell is
  // unique, so the result is the same bits either way.
  ASX_DEV double logtemp_re(double rho, double &eps, double ye, int *hint = nullptr) const {
    const double r = crho(rho);
    eps = fmax(eps + shift, eps_min + shift) - shift;
    const double lrho = log(r);
    const double leps = log(eps + shift);
    if (hint && *hint >= 0 && *hint <= n[1] - 3) {
      const int i0 = find(0, lrho), i2 = find(2, ye), h = *hint;
      const int s1 = NV * n[0], s2 = NV * n[0] * n[1];
      const double *p = tab + EPS + NV * i0 + s2 * i2;
      const double r0 = ax[0][i0], r1 = ax[0][i0 + 1];
      const double y0 = ax[2][i2], y1 = ax[2][i2 + 1];
      const double eh = enode(p + s1 * h, s2, r0, r1, y0, y1, lrho, ye);
      const double eh1 = enode(p + s1 * (h + 1), s2, r0, r1, y0, y1, lrho, ye);
      // at h = 0 the bisection takes the end value, equal to E_0: strictly
      // above it, no clamp applies
      if (eh <= leps && leps < eh1 && (h > 0 || leps > eh)) {
        const double t0 = ax[1][h], t1 = ax[1][h + 1];
        return t0 + (leps - eh) * (t1 - t0) / (eh1 - eh);
      }
    }
    const double vmin = interp(EPS, lrho, xmin1(), ye);
    const double vmax = interp(EPS, lrho, xmax1(), ye);
    if (leps <= vmin) {
      eps = exp(vmin) - shift;
      if (hint) *hint = -1;
      return xmin1();
    }
    if (leps >= vmax) {
      eps = exp(vmax) - shift;
      if (hint) *hint = -1;
      return xmax1();
    }
    return logtemp_nodes(lrho, leps, ye, vmin, vmax, hint);
  }
  // E_j at the node row q of the T axis: log(eps + shift) interpolated in
  // log rho, then in Ye (the reference's order)
  ASX_DEV static double enode(const double *q, int s2, double r0, double r1, double y0,
                              double y1, double lrho, double ye) {
    const double a0 = lin(q[0], q[NV], r0, r1, lrho);
    const double a1 = lin(q[s2], q[s2 + NV], r0, r1, lrho);
    return lin(a0, a1, y0, y1, ye);
  }
  // The root in log T of interp(EPS, lrho, lt, ye) = leps for
  // vmin < leps < vmax (the interpolant's values at the ends of the T
  // axis).  At fixed (lrho, ye) the interpolant is linear in lt inside each
  // cell [t_j, t_j+1], between the node values E_j (the bilinear rho, Ye
  // interpolation at T node j, in the reference's order), so the root is
  // found exactly: a bisection over the node index keeps
  // E_lo <= leps < E_hi -- a sign change, hence a root, in between -- down
  // to adjacent nodes, then that cell's linear equation is solved.  Brent's
  // iteration on the same function stops within 2 eps |lt| + 1e-14 of this
  // root, after about twice as many dependent steps of twice the reads.
  ASX_DEV double logtemp_nodes(double lrho, double leps, double ye, double vmin,
                               double vmax, int *hint = nullptr) const {
    const int i0 = find(0, lrho), i2 = find(2, ye);
    const int s1 = NV * n[0], s2 = NV * n[0] * n[1];
    const double *p = tab + EPS + NV * i0 + s2 * i2;
    const double r0 = ax[0][i0], r1 = ax[0][i0 + 1];
    const double y0 = ax[2][i2], y1 = ax[2][i2 + 1];
    int lo = 0, hi = n[1] - 1;
    double elo = vmin, ehi = vmax;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      const double em = enode(p + s1 * mid, s2, r0, r1, y0, y1, lrho, ye);
      if (em <= leps) {
        lo = mid;
        elo = em;
      } else {
        hi = mid;
        ehi = em;
      }
    }
    if (hint) *hint = lo;
    const double t0 = ax[1][lo], t1 = ax[1][hi];
    return t0 + (leps - elo) * (t1 - t0) / (ehi - elo);
  }
  // press_from_valid_rho_temp_ye: the bounded rho, T are not used there
  ASX_DEV double press_rt(double rho, double temp, double ye) const {
    return exp(interp(PRESS, log(rho), log(temp), ye));
  }
  ASX_DEV double press_re(double rho, double &eps, double ye, int *hint = nullptr) const {
    const double lr = log(crho(rho));
    const double lt = logtemp_re(rho, eps, ye, hint);
    return exp(interp(PRESS, lr, lt, ye));
  }
  ASX_DEV double eps_rt(double rho, double temp, double ye) const {
    return exp(interp(EPS, log(crho(rho)), log(ctemp(temp)), ye)) - shift;
  }
  ASX_DEV double csnd_rt(double rho, double temp, double ye) const {
    return sqrt(interp(CS2, log(crho(rho)), log(ctemp(temp)), ye));
  }
  ASX_DEV double csnd_re(double rho, double &eps, double ye) const {
    const double lr = log(crho(rho));
    const double lt = logtemp_re(rho, eps, ye);
    return sqrt(interp(CS2, lr, lt, ye));
  }
  ASX_DEV double temp_re(double rho, double &eps, double ye) const {
    return exp(logtemp_re(rho, eps, ye));
  }
  // press_re, temp_re and kappa_re at one (rho, eps, Ye) from one
  // temperature solve.  Called in turn, the reference's functions solve
  // three times, the second and third time at the eps the first clamped;
  // where that clamp applied, the later solves may land an ulp away from
  // the table edge, which is the only difference
  ASX_DEV void state_re(double rho, double &eps, double ye, double &press, double &temp,
                        double &kap, int *hint = nullptr) const {
    const double lt = logtemp_re(rho, eps, ye, hint);
    press = exp(interp(PRESS, log(crho(rho)), lt, ye));
    temp = exp(lt);
    kap = entropy_rt(rho, temp, ye);
  }
  // entropy_from_valid_rho_temp_ye (:320-329)
  ASX_DEV double entropy_rt(double rho, double temp, double ye) const {
    return interp(ENT, log(crho(rho)), log(ctemp(temp)), ye);
  }
  // kappa_from_valid_rho_eps_ye (:331-336): the entropy at T(rho, eps, Ye)
  ASX_DEV double kappa_re(double rho, double &eps, double ye) const {
    return entropy_rt(rho, temp_re(rho, eps, ye), ye);
  }
  // eps_from_valid_rho_press_ye is refused by the reference for tables;
  // the callers that reach it (pressure atmosphere, use_temperature = no)
  // are refused at the C ABI
  ASX_DEV double eps_rp(double, double, double) const { return __builtin_nan(""); }
  // p and eps from (rho, kappa): refused by the reference for tables
  // (eos_3p_tabulated3d.hxx:315-333), and at the C ABI here
  ASX_DEV double press_rk(double, double, double) const { return __builtin_nan(""); }
  ASX_DEV double eps_rk(double, double, double) const { return __builtin_nan(""); }
};

} // namespace asx
