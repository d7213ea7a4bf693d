// Tabulated 3-parameter EOS for gfx950 (EOSX eos_3p_tabulated3d): a table
// of log p, log(eps + shift), c_s^2 and entropy over axes log rho, log T,
// Ye (code units; axes uniform), the variable index fastest, and the EOSX interface
// on it (EOSX/src/eos_3p_tabulated3d/eos_3p_tabulated3d.hxx:186-357):
// trilinear interpolation in the order of eos_linear_interp_ND.hxx (rho,
// then T, then Ye), temperature from (rho, eps, Ye) by Brent's ZERO, and
// the same clamping of the inputs, including the change of eps the
// reference makes through its reference argument.
#pragma once

#include "asx_device.hpp"

namespace asx {

// Brent's ZERO (Brent 1973, "Algorithms for Minimization without
// Derivatives", ch. 4): the root of f in [a, b] (f(a) f(b) <= 0) to
// 2 eps |x| + t, returning the last iterate -- the form of
// EOSX/src/utils/eos_brent.hxx:8-166.  The loop is capped at kBrentCap
// iterations: on a valid bracket ZERO needs far fewer, and the cap only ends
// a NaN residual, which would otherwise step by t across the whole bracket.
constexpr int kBrentCap = 1000;
template <typename F>
ASX_DEV double zero_brent(double a, double b, double t, const F &f) {
  double sa = a, sb = b;
  double fa = f(sa), fb = f(sb);
  double c = sa, fc = fa, e = sb - sa, d = e;
  const double macheps = 2.220446049250313e-16;
  for (int it = 0; it < kBrentCap; ++it) {
    if (fabs(fc) < fabs(fb)) {
      sa = sb; sb = c; c = sa;
      fa = fb; fb = fc; fc = fa;
    }
    const double tol = 2.0 * macheps * fabs(sb) + t;
    const double m = 0.5 * (c - sb);
    if (fabs(m) <= tol || fb == 0.0) break;
    if (fabs(e) < tol || fabs(fa) <= fabs(fb)) {
      e = m;
      d = e;
    } else {
      double s = fb / fa, p, q;
      if (sa == c) {
        p = 2.0 * m * s;
        q = 1.0 - s;
      } else {
        q = fa / fc;
        const double r = fb / fc;
        p = s * (2.0 * m * q * (q - r) - (sb - sa) * (r - 1.0));
        q = (q - 1.0) * (r - 1.0) * (s - 1.0);
      }
      if (0.0 < p) q = -q;
      else p = -p;
      s = e;
      e = d;
      if (2.0 * p < 3.0 * m * q - fabs(tol * q) && p < fabs(0.5 * s * q)) {
        d = p / q;
      } else {
        e = m;
        d = e;
      }
    }
    sa = sb;
    fa = fb;
    if (tol < fabs(d)) sb = sb + d;
    else if (0.0 < m) sb = sb + tol;
    else sb = sb - tol;
    fb = f(sb);
    if ((0.0 < fb && 0.0 < fc) || (fb <= 0.0 && fc <= 0.0)) {
      c = sa;
      fc = fa;
      e = sb - sa;
      d = e;
    }
  }
  return sb;
}

struct TabEOS {
  enum { PRESS = 0, EPS = 1, CS2 = 2, ENT = 3, NV = 4 };
  const double *tab;    // [v + NV * (ir + n0 * (it + n1 * iy))]
  const double *ax[3];  // axis values: log rho, log T, Ye
  int n[3];
  double idx[3];        // 1 / (ax[d][1] - ax[d][0])
  double shift;         // energy shift
  double rho_min, rho_max, temp_min, temp_max, ye_min, ye_max; // rgrho, rgtemp, rgye
  double eps_min, eps_max;                                     // rgeps
  double tol;           // temperature inversion tolerance in log T

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

  // logtemp_from_valid_rho_eps_ye (eos_3p_tabulated3d.hxx:186-218)
  ASX_DEV double logtemp_re(double rho, double &eps, double ye) const {
    const double r = crho(rho);
    eps = fmax(eps + shift, eps_min + shift) - shift;
    const double lrho = log(r);
    const double leps = log(eps + shift);
    const double vmin = interp(EPS, lrho, xmin1(), ye);
    const double vmax = interp(EPS, lrho, xmax1(), ye);
    if (leps <= vmin) {
      eps = exp(vmin) - shift;
      return xmin1();
    }
    if (leps >= vmax) {
      eps = exp(vmax) - shift;
      return xmax1();
    }
    auto f = [&](double lt) { return leps - interp(EPS, lrho, lt, ye); };
    return zero_brent(xmin1(), xmax1(), tol, f);
  }
  // press_from_valid_rho_temp_ye: the bounded rho, T are not used there
  ASX_DEV double press_rt(double rho, double temp, double ye) const {
    return exp(interp(PRESS, log(rho), log(temp), ye));
  }
  ASX_DEV double press_re(double rho, double &eps, double ye) const {
    const double lr = log(crho(rho));
    const double lt = logtemp_re(rho, eps, ye);
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
};

} // namespace asx
