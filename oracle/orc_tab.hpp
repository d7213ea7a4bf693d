// ORACLE -- test infrastructure only (see orc_physics.hpp).
// CPU restatement of the tabulated 3-parameter EOS of EOSX:
//   eos_3p_tabulated3d.hxx:186-357 (the interface on the table),
//   utils/eos_linear_interp_ND.hxx:26-231 (uniform-axis linear
//   interpolation: lower node by truncation of (x - x0)/dx, clamped to
//   [0, n-2]; interpolation in rho, then T, then Ye),
//   utils/eos_brent.hxx:8-166 (zero_brent, Brent's ZERO).
// The table holds 4 variables, the variable index fastest: log p,
// log(eps + shift), c_s^2 and the entropy (the reference's table holds 19
// and reads these 4 on this path).
#pragma once

#include <cmath>
#include <limits>

namespace orc {

// zero_brent (EOSX/src/utils/eos_brent.hxx:8-166), with the loop capped
// at 1000 iterations (the reference's for (;;) never ends on a NaN
// residual; a valid bracket converges long before the cap)
template <typename F> double zero_brent(double a, double b, double t, F &f) {
  double sa = a, sb = b;
  double fa = f(sa), fb = f(sb);
  double c = sa, fc = fa, e = sb - sa, d = e;
  const double macheps = std::numeric_limits<double>::epsilon();
  for (int it = 0; it < 1000; ++it) {
    if (std::fabs(fc) < std::fabs(fb)) {
      sa = sb;
      sb = c;
      c = sa;
      fa = fb;
      fb = fc;
      fc = fa;
    }
    const double tol = 2.0 * macheps * std::fabs(sb) + t;
    const double m = 0.5 * (c - sb);
    if (std::fabs(m) <= tol || fb == 0.0) break;
    if (std::fabs(e) < tol || std::fabs(fa) <= std::fabs(fb)) {
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
      if (2.0 * p < 3.0 * m * q - std::fabs(tol * q) && p < std::fabs(0.5 * s * q)) {
        d = p / q;
      } else {
        e = m;
        d = e;
      }
    }
    sa = sb;
    fa = fb;
    if (tol < std::fabs(d)) sb = sb + d;
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
  const double *y;       // table, variable fastest
  const double *x[3];    // axes: log rho, log T, Ye
  int n[3];
  double idx[3];
  double shift;
  // rgrho, rgtemp, rgye (exp of the log axes), rgeps
  double rho_min, rho_max, temp_min, temp_max, ye_min, ye_max, eps_min, eps_max;
  double tol = 1.e-14; // eos_3p_tabulated3d.hxx:217

  // eos_linear_interp_ND.hxx:57-73 (uniform)
  int find_index(int d, double xin) const {
    const double xL = xin - x[d][0];
    if (!(xL > 0)) return 0; // a NaN input takes the first cell
    const double f = xL * idx[d];
    if (f >= (double)(n[d] - 2)) return n[d] - 2;
    return (int)f;
  }
  static double linterp1D(double A, double B, double xA, double xB, double xv) {
    const double lambda = (xv - xA) / (xB - xA);
    return (B - A) * lambda + A;
  }
  // eos_linear_interp_ND.hxx:119-160: Ye outermost, rho innermost
  double interpolate(int var, double lr, double lt, double ye) const {
    const int i0 = find_index(0, lr), i1 = find_index(1, lt), i2 = find_index(2, ye);
    auto F = [&](int a, int b, int c) {
      return y[var + NV * ((i0 + a) + n[0] * ((i1 + b) + n[1] * (i2 + c)))];
    };
    auto in_rho = [&](int b, int c) {
      return linterp1D(F(0, b, c), F(1, b, c), x[0][i0], x[0][i0 + 1], lr);
    };
    auto in_temp = [&](int c) {
      const double A = in_rho(0, c), B = in_rho(1, c);
      return linterp1D(A, B, x[1][i1], x[1][i1 + 1], lt);
    };
    const double A = in_temp(0), B = in_temp(1);
    return linterp1D(A, B, x[2][i2], x[2][i2 + 1], ye);
  }
  double xmin1() const { return x[1][0]; }
  double xmax1() const { return x[1][n[1] - 1]; }
  double crho(double rho) const { return std::fmin(std::fmax(rho, rho_min), rho_max); }
  double ctemp(double t) const { return std::fmin(std::fmax(t, temp_min), temp_max); }

  // eos_3p_tabulated3d.hxx:186-218
  double logtemp_from_rho_eps_ye(double rho, double &eps, double ye) const {
    const double r = crho(rho);
    eps = std::fmax(eps + shift, eps_min + shift) - shift;
    const double lrho = std::log(r);
    const double leps = std::log(eps + shift);
    const double vmin = interpolate(EPS, lrho, xmin1(), ye);
    const double vmax = interpolate(EPS, lrho, xmax1(), ye);
    if (leps <= vmin) {
      eps = std::exp(vmin) - shift;
      return xmin1();
    }
    if (leps >= vmax) {
      eps = std::exp(vmax) - shift;
      return xmax1();
    }
    auto func = [&](double &lt) { return leps - interpolate(EPS, lrho, lt, ye); };
    return zero_brent(xmin1(), xmax1(), tol, func);
  }
  // :220-229 (the bounded rho and T are computed but the logs take the inputs)
  double press_from_rho_temp_ye(double rho, double temp, double ye) const {
    return std::exp(interpolate(PRESS, std::log(rho), std::log(temp), ye));
  }
  // :231-238
  double press_from_rho_eps_ye(double rho, double &eps, double ye) const {
    const double lr = std::log(crho(rho));
    const double lt = logtemp_from_rho_eps_ye(rho, eps, ye);
    return std::exp(interpolate(PRESS, lr, lt, ye));
  }
  // :240-247
  double eps_from_rho_temp_ye(double rho, double temp, double ye) const {
    return std::exp(interpolate(EPS, std::log(crho(rho)), std::log(ctemp(temp)), ye)) - shift;
  }
  // :261-269
  double csnd_from_rho_temp_ye(double rho, double temp, double ye) const {
    return std::sqrt(interpolate(CS2, std::log(crho(rho)), std::log(ctemp(temp)), ye));
  }
  // :271-279
  double csnd_from_rho_eps_ye(double rho, double &eps, double ye) const {
    const double lr = std::log(crho(rho));
    const double lt = logtemp_from_rho_eps_ye(rho, eps, ye);
    return std::sqrt(interpolate(CS2, lr, lt, ye));
  }
  // :281-286
  double temp_from_rho_eps_ye(double rho, double &eps, double ye) const {
    return std::exp(logtemp_from_rho_eps_ye(rho, eps, ye));
  }
  // :320-329
  double entropy_from_rho_temp_ye(double rho, double temp, double ye) const {
    return interpolate(ENT, std::log(crho(rho)), std::log(ctemp(temp)), ye);
  }
  // :331-336
  double kappa_from_rho_eps_ye(double rho, double &eps, double ye) const {
    return entropy_from_rho_temp_ye(rho, temp_from_rho_eps_ye(rho, eps, ye), ye);
  }
  // :249-258: refused by the reference for tables
  double eps_from_rho_press_ye(double, double, double) const {
    return std::numeric_limits<double>::quiet_NaN();
  }
};

} // namespace orc
