// ORACLE -- test infrastructure only (see orc_physics.hpp).
// Flop-counting build of the oracle's con2prim: orc_physics.hpp and
// orc_c2p.hpp compiled with every `double` replaced by `real`, a type that
// counts its arithmetic (the Makefile writes the substituted copies under
// _flops/).  Convention: +, -, *, / and sqrt count 1 each; fabs, copysign,
// min/max, negation and comparisons count 0; pow, log, exp are counted on
// their own (the ideal-gas Noble path has none per iteration).  The counts are
// the algorithmic FLOPs of the reference's solver as restated here (the
// device evaluates the same expressions, with its divisions and square roots
// expanded into several instructions each).
#include <cmath>
#include <cstdint>
#include <limits>

namespace orcflops {
uint64_t g_arith = 0, g_trans = 0;
}

struct real {
  double v;
  real() = default;
  real(double x) : v(x) {}
  explicit operator int() const { return (int)v; }
  explicit operator double() const { return v; }
  real &operator+=(real b) { ++orcflops::g_arith; v += b.v; return *this; }
  real &operator-=(real b) { ++orcflops::g_arith; v -= b.v; return *this; }
  real &operator*=(real b) { ++orcflops::g_arith; v *= b.v; return *this; }
  real &operator/=(real b) { ++orcflops::g_arith; v /= b.v; return *this; }
};
inline real operator+(real a, real b) { ++orcflops::g_arith; return a.v + b.v; }
inline real operator-(real a, real b) { ++orcflops::g_arith; return a.v - b.v; }
inline real operator*(real a, real b) { ++orcflops::g_arith; return a.v * b.v; }
inline real operator/(real a, real b) { ++orcflops::g_arith; return a.v / b.v; }
inline real operator-(real a) { return -a.v; }
inline real operator+(real a) { return a; }
inline bool operator<(real a, real b) { return a.v < b.v; }
inline bool operator>(real a, real b) { return a.v > b.v; }
inline bool operator<=(real a, real b) { return a.v <= b.v; }
inline bool operator>=(real a, real b) { return a.v >= b.v; }
inline bool operator==(real a, real b) { return a.v == b.v; }
inline bool operator!=(real a, real b) { return a.v != b.v; }

namespace std {
inline real sqrt(real a) { ++orcflops::g_arith; return std::sqrt(a.v); }
inline real fabs(real a) { return std::fabs(a.v); }
inline real abs(real a) { return std::fabs(a.v); }
inline real fmax(real a, real b) { return std::fmax(a.v, b.v); }
inline real fmin(real a, real b) { return std::fmin(a.v, b.v); }
inline real pow(real a, real b) { ++orcflops::g_trans; return std::pow(a.v, b.v); }
inline real log(real a) { ++orcflops::g_trans; return std::log(a.v); }
inline real exp(real a) { ++orcflops::g_trans; return std::exp(a.v); }
inline real ldexp(real a, int e) { return std::ldexp(a.v, e); }
inline real copysign(real a, real b) { return std::copysign(a.v, b.v); }
inline bool isfinite(real a) { return std::isfinite(a.v); }
inline bool isnan(real a) { return std::isnan(a.v); }
inline bool signbit(real a) { return std::signbit(a.v); }
template <> class numeric_limits<real> {
public:
  static real quiet_NaN() { return std::numeric_limits<double>::quiet_NaN(); }
  static real infinity() { return std::numeric_limits<double>::infinity(); }
  static real epsilon() { return std::numeric_limits<double>::epsilon(); }
  static real max() { return std::numeric_limits<double>::max(); }
  static real min() { return std::numeric_limits<double>::min(); }
};
} // namespace std

#include "asterx.h"
#include "_flops/orc_c2p.hpp"

using namespace orc;

extern "C" {

// The con2prim of asx_c2p_batch (ideal gas; same inputs and driver as
// orc_c2p_batch) on nz zones, one at a time, recording per zone the
// arithmetic FLOPs, the pow/log/exp calls, the primary solver's iterations
// and the flag.
int c2p_flops(int64_t nz, const double *cons, const double *g, const double *seed,
              const asx_params *par, double *flops, double *trans, int32_t *iters,
              double *flag) {
  IdealGas eos;
  eos.init(par->gl_gamma, par->particle_mass, par->eps_min, par->eps_max, par->rho_min,
           par->rho_max, par->ye_min, par->ye_max);
  Polytrope poly;
  poly.init(par->poly_gamma, par->poly_k);
  C2PDriver d;
  d.prime = par->c2p_prime;
  d.second = par->c2p_second;
  d.excise = par->excise != 0;
  d.entropy_fix = par->use_entropy_fix != 0;
  d.ap = AtmoParams{par->rho_abs_min, par->atmo_tol, par->p_atmo, par->t_atmo,
                    par->Ye_atmo, par->r_atmo, par->n_rho_atmo, par->n_press_atmo,
                    par->n_temp_atmo, par->thermal_eos_atmo != 0, par->use_press_atmo != 0};
  d.cp.max_iter = par->max_iter;
  d.cp.tol = par->c2p_tol;
  d.cp.alp_thresh = par->alp_thresh;
  d.cp.cons_error = par->cons_error_limit;
  d.cp.vw_lim = par->vw_lim;
  d.cp.w_lim = std::sqrt(1.0 + par->vw_lim * par->vw_lim);
  d.cp.v_lim = par->vw_lim / d.cp.w_lim.v;
  d.cp.Bsq_lim = par->B_lim * par->B_lim;
  d.cp.rho_BH = par->rho_BH;
  d.cp.eps_BH = par->eps_BH;
  d.cp.vwlim_BH = par->vwlim_BH;
  d.cp.use_z = par->use_z != 0;
  d.cp.use_temp = par->use_temperature != 0;
  d.cp.use_press_atmo = par->use_press_atmo != 0;
  for (int64_t z = 0; z < nz; ++z) {
    C2PCellIn in;
    in.r = 0.0;
    in.alp = 1.0;
    in.glo = Sym3{g[0 * nz + z], g[1 * nz + z], g[2 * nz + z],
                  g[3 * nz + z], g[4 * nz + z], g[5 * nz + z]};
    in.mask = 1.0;
    in.cv.dens = cons[0 * nz + z];
    for (int i = 0; i < 3; ++i) in.cv.mom[i] = cons[(1 + i) * nz + z];
    in.cv.tau = cons[4 * nz + z];
    for (int i = 0; i < 3; ++i) in.cv.dB[i] = cons[(5 + i) * nz + z];
    in.cv.DYe = in.cv.dens.v * par->Ye_atmo;
    in.cv.DEnt = 0.0;
    in.saved_rho = seed[0 * nz + z];
    in.saved_eps = seed[1 * nz + z];
    in.saved_Ye = par->Ye_atmo;
    for (int i = 0; i < 3; ++i) in.saved_vel[i] = seed[(2 + i) * nz + z];
    in.temperature = 0.0;
    C2PCellOut out;
    orcflops::g_arith = 0;
    orcflops::g_trans = 0;
    con2prim_cell(eos, poly, d, in, out);
    flops[z] = (double)orcflops::g_arith;
    trans[z] = (double)orcflops::g_trans;
    iters[z] = out.iters;
    flag[z] = out.flag.v;
  }
  return ASX_OK;
}

} // extern "C"
