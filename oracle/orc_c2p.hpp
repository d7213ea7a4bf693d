// ORACLE -- test infrastructure only (see orc_physics.hpp).
// CPU restatement of Con2PrimFactory (c2p.hxx, c2p_2DNoble.hxx,
// c2p_1DPalenzuela.hxx, atmo.hxx, prims.hxx, cons.hxx, c2p_report.hxx) and of
// the per-cell driver AsterX/src/con2prim.cxx:73-454, ideal-gas EOS.
//
// Algo::brent (CarpetX Algo/roots.hxx) is not in the reference tree; it is
// restated as Brent's ZERO (Brent 1973), the form the reference ships in
// EOSX/src/utils/eos_brent.hxx, returning the final bracket, with the
// termination rule the reference documents at c2p_1DPalenzuela.hxx:380-393.
// Its iterate sequence is parity unpinned; Palenzuela results are pinned to
// solver tolerance by the round-trip and TSV tests.
#pragma once

#include "orc_physics.hpp"

#include <cstdio>
#include <type_traits>
#include <utility>

namespace orc {

// Con2PrimFactory/src/c2p_report.hxx:17-37
enum C2PStatus {
  ST_SUCCESS = 0, ST_INVALID_DETG, ST_NEG_BSQR, ST_NANS_IN_CONS, ST_RANGE_RHO,
  ST_RANGE_EPS, ST_SPEED_LIMIT, ST_RANGE_YE, ST_B_LIMIT, ST_ROOT_FAIL_CONV,
  ST_ROOT_FAIL_BRACKET, ST_PREP_ROOT_FAIL_CONV, ST_PREP_ROOT_FAIL_BRACKET,
  ST_NOT_SET
};
struct Report {
  int status = ST_NOT_SET;
  bool adjust_cons = false, set_atmo = false;
  int iters = 0;
  bool failed() const { return status != ST_SUCCESS; }
};

// Con2PrimFactory/src/prims.hxx:14-94
struct Prim {
  double rho, eps, Ye, press, temperature, entropy;
  double vel[3];
  double w_lor;
  double B[3];
};
// Con2PrimFactory/src/cons.hxx:15-113
struct Cons {
  double dens, mom[3], tau, DYe, DEnt, dB[3];
  void scale(double s) { // *= s on every component (re-densitize)
    dens *= s; tau *= s;
    for (int i = 0; i < 3; ++i) mom[i] *= s;
    for (int i = 0; i < 3; ++i) dB[i] *= s;
    DYe *= s; DEnt *= s;
  }
  void divide(double s) { // /= s on every component (undensitize)
    dens /= s; tau /= s;
    for (int i = 0; i < 3; ++i) mom[i] /= s;
    for (int i = 0; i < 3; ++i) dB[i] /= s;
    DYe /= s; DEnt /= s;
  }
  void set_nan() {
    const double n = std::numeric_limits<double>::quiet_NaN();
    dens = tau = DYe = DEnt = n;
    for (int i = 0; i < 3; ++i) mom[i] = dB[i] = n;
  }
  // cons.hxx:49-91 (uses the stored Lorentz factor)
  void from_prim(const Prim &pv, const Sym3 &g) {
    const double sqrt_detg = std::sqrt(calc_det(g));
    double v_low[3], B_low[3];
    lower(g, pv.vel, v_low);
    const double w = pv.w_lor;
    lower(g, pv.B, B_low);
    const double bst = w * dot3(pv.B, v_low);
    double b_low[3];
    for (int i = 0; i < 3; ++i) b_low[i] = B_low[i] / w + bst * v_low[i];
    const double bs2 = (dot3(pv.B, B_low) + bst * bst) / (w * w);
    dens = sqrt_detg * pv.rho * w;
    const double hW2 = w * w * (pv.rho * (1.0 + pv.eps) + pv.press + bs2);
    for (int i = 0; i < 3; ++i)
      mom[i] = sqrt_detg * (hW2 * v_low[i] - bst * b_low[i]);
    tau = sqrt_detg * (w * w * (pv.rho * (1.0 + pv.eps) + pv.press + bs2) -
                       (pv.press + 0.5 * bs2) - bst * bst) -
          dens;
    for (int i = 0; i < 3; ++i) dB[i] = sqrt_detg * pv.B[i];
    DYe = dens * pv.Ye;
    DEnt = dens * pv.entropy;
  }
};
inline void prim_set_nan(Prim &pv) {
  const double n = std::numeric_limits<double>::quiet_NaN();
  pv.rho = pv.eps = pv.Ye = pv.press = pv.temperature = pv.entropy = n;
  pv.vel[0] = pv.vel[1] = pv.vel[2] = pv.w_lor = n;
  pv.B[0] = pv.B[1] = pv.B[2] = n;
}

// Con2PrimFactory/src/atmo.hxx:16-83
struct Atmo {
  double rho, eps, ye, press, temp, entropy, rho_cut;
  void set(Prim &pv) const {
    pv.rho = rho; pv.eps = eps; pv.Ye = ye; pv.press = press;
    pv.temperature = temp; pv.entropy = entropy;
    pv.vel[0] = pv.vel[1] = pv.vel[2] = 0.0;
    pv.w_lor = 1.0;
  }
  void set(Prim &pv, Cons &cv, const Sym3 &g) const {
    set(pv);
    const double sqrt_detg = std::sqrt(calc_det(g));
    cv.dens = sqrt_detg * rho;
    cv.mom[0] = cv.mom[1] = cv.mom[2] = 0.0;
    cv.DYe = cv.dens * ye;
    cv.DEnt = cv.dens * entropy;
    double B_low[3];
    lower(g, pv.B, B_low);
    const double Bsq = dot3(pv.B, B_low);
    cv.tau = cv.dens * eps + 0.5 * sqrt_detg * Bsq;
  }
};

struct C2PParams {
  int max_iter;
  double tol, alp_thresh, cons_error, vw_lim, w_lim, v_lim, Bsq_lim, rho_BH,
      eps_BH, vwlim_BH;
  bool use_z, use_temp, use_press_atmo;
};

// Con2PrimFactory/src/c2p.hxx:86-216
template <class E>
inline void prims_floors(const E &eos, const Atmo &atmo,
                         const C2PParams &cp, Prim &pv, const Cons &cv,
                         const Sym3 &glo, Report &rep) {
  bool recomp = false;
  if (pv.Ye < eos.ye_min) { pv.Ye = eos.ye_min; rep.adjust_cons = true; recomp = true; }
  if (pv.Ye > eos.ye_max) { pv.Ye = eos.ye_max; rep.adjust_cons = true; recomp = true; }
  double v_low[3];
  lower(glo, pv.vel, v_low);
  const double vsq = dot3(v_low, pv.vel);
  const double sol_v = std::sqrt(vsq);
  if (sol_v > cp.v_lim) {
    pv.rho = cv.dens / cp.w_lim;
    const double s = cp.v_lim / sol_v;
    for (int i = 0; i < 3; ++i) pv.vel[i] *= s;
    pv.w_lor = cp.w_lim;
    rep.adjust_cons = true;
    if (cp.use_temp) {
      recomp = true;
    } else {
      recomp = false;
      pv.eps = eos.eps_from_rho_press_ye(pv.rho, pv.press, pv.Ye);
      pv.temperature = eos.temp_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
      pv.entropy = eos.kappa_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
    }
  }
  if (pv.rho > eos.rho_max) {
    pv.rho = eos.rho_max;
    rep.adjust_cons = true;
    if (cp.use_temp) {
      recomp = true;
    } else {
      recomp = false;
      pv.eps = eos.eps_from_rho_press_ye(pv.rho, pv.press, pv.Ye);
      pv.temperature = eos.temp_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
      pv.entropy = eos.kappa_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
    }
  }
  if (pv.temperature > eos.temp_max) {
    pv.temperature = eos.temp_max;
    recomp = true;
    rep.adjust_cons = true;
  }
  if (cp.use_press_atmo) {
    if (pv.press < atmo.press) {
      pv.press = atmo.press;
      pv.eps = eos.eps_from_rho_press_ye(pv.rho, pv.press, pv.Ye);
      pv.temperature = eos.temp_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
      pv.entropy = eos.kappa_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
      recomp = false;
      rep.adjust_cons = true;
    }
  } else {
    if (pv.temperature < atmo.temp) {
      pv.temperature = atmo.temp;
      recomp = true;
      rep.adjust_cons = true;
    }
  }
  if (recomp) {
    pv.eps = eos.eps_from_rho_temp_ye(pv.rho, pv.temperature, pv.Ye);
    pv.press = eos.press_from_rho_temp_ye(pv.rho, pv.temperature, pv.Ye);
    pv.entropy = eos.kappa_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
  }
}

// Con2PrimFactory/src/c2p.hxx:218-328
template <bool limiting, class E>
inline void bh_interior(const E &eos, const Atmo &atmo,
                        const C2PParams &cp, Prim &pv, Cons &cv,
                        const Sym3 &glo) {
  const double wlim_BH = std::sqrt(1.0 + cp.vwlim_BH * cp.vwlim_BH);
  const double vlim_BH = cp.vwlim_BH / wlim_BH;
  bool recomp = false;
  if (limiting) {
    if (pv.rho > cp.rho_BH) { pv.rho = cp.rho_BH; recomp = true; }
    if (pv.eps > cp.eps_BH) { pv.eps = cp.eps_BH; recomp = true; }
    const double sol_v = std::sqrt((pv.w_lor * pv.w_lor - 1.0)) / pv.w_lor;
    if (sol_v > vlim_BH) {
      const double s = vlim_BH / sol_v;
      for (int i = 0; i < 3; ++i) pv.vel[i] *= s;
      pv.w_lor = wlim_BH;
      recomp = true;
    }
    if (recomp) {
      pv.temperature = eos.temp_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
      pv.press = eos.press_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
      pv.entropy = eos.kappa_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
      cv.from_prim(pv, glo);
    }
  } else {
    pv.rho = cp.rho_BH;
    pv.eps = cp.eps_BH;
    pv.Ye = atmo.ye;
    pv.temperature = eos.temp_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
    pv.press = eos.press_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
    pv.entropy = eos.kappa_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
    const double detg = calc_det(glo);
    const Sym3 gup = calc_inv(glo, detg);
    const double Z = (pv.rho * (1.0 + pv.eps) + pv.press) * wlim_BH * wlim_BH;
    double B_low[3];
    lower(glo, pv.B, B_low);
    const double Bsq = dot3(B_low, pv.B);
    const double sq = std::sqrt(detg);
    double mom_low[3], mom_up[3];
    for (int i = 0; i < 3; ++i) mom_low[i] = cv.mom[i] / sq;
    lower(gup, mom_low, mom_up);
    const double Ssq_old = dot3(mom_low, mom_up);
    const double S_old = std::sqrt(Ssq_old) + 1e-50;
    const double BiSi_old = dot3(mom_low, pv.B);
    const double BiEsi = BiSi_old / S_old;
    const double Ssq_new = ((Z + Bsq) * (Z + Bsq) * vlim_BH * vlim_BH) /
                           (1.0 + BiEsi * BiEsi * (2.0 * Z + Bsq) / (Z * Z));
    const double S_new = std::sqrt(Ssq_new);
    for (int i = 0; i < 3; ++i) mom_low[i] *= S_new / S_old;
    for (int i = 0; i < 3; ++i) mom_up[i] *= S_new / S_old;
    for (int i = 0; i < 3; ++i) {
      pv.vel[i] = mom_up[i] / (Z + Bsq);
      pv.vel[i] += BiEsi * S_new * pv.B[i] / (Z * (Z + Bsq));
    }
    pv.w_lor = wlim_BH;
    cv.from_prim(pv, glo);
  }
}

// Con2PrimFactory/src/c2p.hxx:330-369
inline void cons_floors(Cons &cv, const Sym3 &glo) {
  const double detg = calc_det(glo);
  const double sqrt_detg = std::sqrt(detg);
  const Sym3 gup = calc_inv(glo, detg);
  double B_low[3];
  lower(glo, cv.dB, B_low);
  const double BsqL = dot3(B_low, cv.dB);
  const double tau_lim = 0.5 * BsqL / sqrt_detg;
  if (cv.tau <= tau_lim) cv.tau = tau_lim;
  double mom_up[3];
  lower(gup, cv.mom, mom_up);
  const double mom2L = dot3(cv.mom, mom_up);
  const double slim = cv.dens + cv.tau;
  const double slim2 = slim * slim;
  if (mom2L > slim2) {
    const double s = std::sqrt(slim2 / mom2L);
    for (int i = 0; i < 3; ++i) cv.mom[i] = cv.mom[i] * s;
  }
}

// Velocity from (Z, B, S) shared by Noble's WZ2Prim and Palenzuela's
// xPalenzuelaToPrim (c2p_2DNoble.hxx:238-256, c2p_1DPalenzuela.hxx:226-244,
// and the z-vector forms :204-236 / :192-225).
inline void vel_from_Z(double Z, double W, double Bsq, double BiSi,
                       const Cons &cv, const Sym3 &gup, const Sym3 &glo,
                       bool use_z, Prim &pv) {
  const double *m = cv.mom;
  const double *B = cv.dB;
  if (use_z) {
    double zx = W * (gup.xx * m[0] + gup.xy * m[1] + gup.xz * m[2]) / (Z + Bsq);
    zx += W * BiSi * B[0] / (Z * (Z + Bsq));
    double zy = W * (gup.xy * m[0] + gup.yy * m[1] + gup.yz * m[2]) / (Z + Bsq);
    zy += W * BiSi * B[1] / (Z * (Z + Bsq));
    double zz = W * (gup.xz * m[0] + gup.yz * m[1] + gup.zz * m[2]) / (Z + Bsq);
    zz += W * BiSi * B[2] / (Z * (Z + Bsq));
    const double zxd = glo.xx * zx + glo.xy * zy + glo.xz * zz;
    const double zyd = glo.xy * zx + glo.yy * zy + glo.yz * zz;
    const double zzd = glo.xz * zx + glo.yz * zy + glo.zz * zz;
    const double Zsq = zx * zxd + zy * zyd + zz * zzd;
    const double lor = std::sqrt(1.0 + Zsq);
    pv.vel[0] = zx / lor;
    pv.vel[1] = zy / lor;
    pv.vel[2] = zz / lor;
    pv.w_lor = lor;
  } else {
    pv.vel[0] = (gup.xx * m[0] + gup.xy * m[1] + gup.xz * m[2]) / (Z + Bsq);
    pv.vel[0] += BiSi * B[0] / (Z * (Z + Bsq));
    pv.vel[1] = (gup.xy * m[0] + gup.yy * m[1] + gup.yz * m[2]) / (Z + Bsq);
    pv.vel[1] += BiSi * B[1] / (Z * (Z + Bsq));
    pv.vel[2] = (gup.xz * m[0] + gup.yz * m[1] + gup.zz * m[2]) / (Z + Bsq);
    pv.vel[2] += BiSi * B[2] / (Z * (Z + Bsq));
    pv.w_lor = W;
  }
}

// Con2PrimFactory/src/c2p_2DNoble.hxx:275-594 (ideal gas; Gamma = eos.gamma,
// Zmin = rho_min as in the constructor :100-101).
inline void noble_solve(const IdealGas &eos, const Atmo &atmo,
                        const C2PParams &cp, Prim &pv, Prim &seed, Cons &cv,
                        const Sym3 &glo, Report &rep) {
  rep.iters = 0;
  rep.adjust_cons = false;
  rep.set_atmo = false;
  rep.status = ST_SUCCESS;
  const double G = eos.gamma;
  const double Zmin = eos.rho_min;
  const double detg = calc_det(glo);
  const double sqrt_detg = std::sqrt(detg);
  const bool minor1 = glo.xx > 0.0;
  const bool minor2 = glo.xx * glo.yy - glo.xy * glo.xy > 0.0;
  const bool minor3 = detg > 0.0;
  if (!(minor1 && minor2 && minor3)) {
    rep.status = ST_INVALID_DETG; rep.adjust_cons = true;
    prim_set_nan(pv); cv.set_nan();
    return;
  }
  const Sym3 gup = calc_inv(glo, detg);
  cv.divide(sqrt_detg);
  for (int i = 0; i < 3; ++i) seed.B[i] = cv.dB[i];
  double mom_up[3], B_low[3];
  lower(gup, cv.mom, mom_up);
  const double Ssq = dot3(cv.mom, mom_up);
  lower(glo, seed.B, B_low);
  const double Bsq = dot3(B_low, seed.B);
  const double BiSi = dot3(cv.mom, seed.B);
  double vsq_seed;
  if (cp.use_z) { // c2p_2DNoble.hxx:145-163
    double z_up[3], z_low[3];
    for (int i = 0; i < 3; ++i) z_up[i] = seed.vel[i] * seed.w_lor;
    lower(glo, z_up, z_low);
    const double zsq = dot3(z_low, z_up);
    const double w = std::sqrt(1. + zsq);
    vsq_seed = smin(zsq / w / w, 1. - 1.e-15);
    seed.w_lor = w;
  } else { // c2p_2DNoble.hxx:126-141
    double v_low[3];
    lower(glo, seed.vel, v_low);
    const double vsq = dot3(v_low, seed.vel);
    seed.w_lor = 1. / std::sqrt(1. - vsq);
    vsq_seed = vsq;
  }
  if ((!std::isfinite(cv.dens)) || (!std::isfinite(Ssq)) ||
      (!std::isfinite(Bsq)) || (!std::isfinite(BiSi)) ||
      (!std::isfinite(cv.DYe)) || (!std::isfinite(cv.DEnt))) {
    rep.status = ST_NANS_IN_CONS; rep.adjust_cons = true;
    prim_set_nan(pv); cv.set_nan();
    return;
  }
  if (Bsq > cp.Bsq_lim) {
    rep.status = ST_B_LIMIT; rep.adjust_cons = true;
    prim_set_nan(pv); cv.set_nan();
    return;
  }
  seed.rho = cv.dens / seed.w_lor;
  const double eps_last = smax(seed.eps, atmo.eps);
  seed.press = eos.press_from_rho_eps(seed.rho, eps_last);
  const double Z_seed =
      (seed.rho + eps_last * seed.rho + seed.press) * seed.w_lor * seed.w_lor;
  double x0 = std::fabs(Z_seed), x1 = vsq_seed;
  double xo0 = x0;
  const double dv = (1. - 1.e-10);
  if (x1 < 0.0) {
    x1 = 0.0;
  } else if (x1 >= 1.0) {
    x1 = dv;
  }
  if (x0 <= 0.0) {
    x0 = std::fabs(x0) + Zmin;
  } else if (x0 > 1e20) {
    x0 = xo0;
  }
  double errx = 1., df = 1., f = 1.;
  int k;
  for (k = 1; k <= cp.max_iter; k++) {
    const double Z = x0;
    const double invZ = 1.0 / Z;
    const double Vsq = x1;
    const double Sdotn = -(cv.tau + cv.dens);
    const double p_tmp =
        ((Z * (1.0 - Vsq) - cv.dens * std::sqrt(1.0 - Vsq)) * (G - 1.0) / (G));
    const double dPdvsq =
        ((-Z + cv.dens / (2.0 * std::sqrt(1.0 - Vsq))) * (G - 1.0) / G);
    const double dPdZ = ((1.0 - Vsq) * (G - 1.0) / G);
    const double j00 = -2 * (Vsq + BiSi * BiSi * invZ * invZ * invZ) * (Bsq + Z);
    const double j01 = -(Bsq + Z) * (Bsq + Z);
    const double j10 = -1.0 + dPdZ - BiSi * BiSi * invZ * invZ * invZ;
    const double j11 = -0.5 * Bsq + dPdvsq;
    const double r0 = Ssq - Vsq * (Bsq + Z) * (Bsq + Z) +
                      BiSi * BiSi * invZ * invZ * (Bsq + Z + Z);
    const double r1 = -Sdotn - 0.5 * Bsq * (1.0 + Vsq) +
                      0.5 * BiSi * BiSi * invZ * invZ - Z + p_tmp;
    const double detjac =
        (Bsq + Z) * (j10 * (Bsq + Z) +
                     (Bsq - 2.0 * dPdvsq) * (BiSi * BiSi * invZ * invZ + Vsq * Z) *
                         invZ);
    const double detjac_inv = 1.0 / detjac;
    const double dx0 = -(j11 * r0 - j01 * r1) * detjac_inv;
    const double dx1 = -(-j10 * r0 + j00 * r1) * detjac_inv;
    df = -r0 * r0 - r1 * r1;
    f = -0.5 * (df);
    errx = 0.;
    xo0 = x0;
    x0 += dx0;
    x1 += dx1;
    if (x1 < 0.0) {
      x1 = 0.0;
    } else if (x1 >= 1.0) {
      x1 = dv;
    }
    if (x0 <= 0.0) {
      x0 = std::fabs(x0) + Zmin;
    } else if (x0 > 1e20) {
      x0 = xo0;
    }
    errx = (x0 == 0.) ? std::fabs(dx0) : std::fabs(dx0 / x0);
    if (std::fabs(errx) <= cp.tol) break;
  }
  rep.iters = k;
  if (std::fabs(errx) > cp.tol) {
    rep.status = ST_ROOT_FAIL_CONV; rep.adjust_cons = true;
    cv.scale(sqrt_detg);
    return;
  }
  if ((!std::isfinite(f)) || (!std::isfinite(df))) {
    rep.status = ST_ROOT_FAIL_BRACKET; rep.adjust_cons = true;
  }
  // WZ2Prim, c2p_2DNoble.hxx:192-273
  const double Z_Sol = x0, vsq_Sol = x1;
  const double W_Sol = 1.0 / std::sqrt(1.0 - vsq_Sol);
  pv.rho = cv.dens / W_Sol;
  vel_from_Z(Z_Sol, W_Sol, Bsq, BiSi, cv, gup, glo, cp.use_z, pv);
  pv.eps = (Z_Sol / pv.w_lor / pv.w_lor / pv.rho - 1.0) / G;
  pv.Ye = cv.DYe / cv.dens;
  pv.press = eos.press_from_rho_eps(pv.rho, pv.eps);
  pv.temperature = eos.temp_from_rho_eps(pv.rho, pv.eps);
  pv.entropy = eos.kappa_from_rho_eps(pv.rho, pv.eps);
  for (int i = 0; i < 3; ++i) pv.B[i] = cv.dB[i];
  if (pv.rho <= 0.0) {
    rep.status = ST_RANGE_RHO; rep.set_atmo = false; rep.adjust_cons = true;
    cv.scale(sqrt_detg);
    return;
  }
  if (pv.eps <= 0.0) {
    rep.status = ST_RANGE_EPS; rep.set_atmo = false; rep.adjust_cons = true;
    cv.scale(sqrt_detg);
    return;
  }
  cv.DEnt = cv.dens * pv.entropy;
  if (pv.rho < atmo.rho_cut) {
    rep.status = ST_SUCCESS; rep.set_atmo = true; rep.adjust_cons = true;
    atmo.set(pv, cv, glo);
    return;
  }
  prims_floors(eos, atmo, cp, pv, cv, glo, rep);
  if (rep.adjust_cons) {
    cv.from_prim(pv, glo);
  } else {
    cv.scale(sqrt_detg);
  }
}

// Brent's bracketing root finder in the form of Brent's ZERO (Brent 1973),
// the structure the reference ships in EOSX/src/utils/eos_brent.hxx:7-165,
// with the stopping rule of the Palenzuela call site
// (c2p_1DPalenzuela.hxx:380-393): stop when the bracket [b, c] satisfies
// |b - c| <= 2^-minbits min(|b|, |c|), when f(b) == 0, or after maxiters
// evaluations.  Steps are at least half that width, so the bracket
// collapses once b has converged.  Returns {b, c}; iters = evaluations.
template <typename F>
inline std::pair<double, double> brent(const F &fn, double a, double b,
                                       int minbits, int maxiters, int &iters) {
  const double rel = std::ldexp(1.0, -minbits);
  double sa = a, sb = b;
  double fa = fn(sa), fb = fn(sb);
  iters = 0;
  if (fa == 0.0) return {sa, sa};
  if (fb == 0.0) return {sb, sb};
  if (fa * fb > 0.0) return {sa, sb}; // not bracketed: interval unchanged
  double c = sa, fc = fa, e = sb - sa, d = e;
  for (;;) {
    if (std::fabs(fc) < std::fabs(fb)) {
      sa = sb; sb = c; c = sa;
      fa = fb; fb = fc; fc = fa;
    }
    const double tol = 0.5 * rel * smin(std::fabs(sb), std::fabs(c));
    const double m = 0.5 * (c - sb);
    if (std::fabs(m) <= tol || fb == 0.0 || iters >= maxiters) break;
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
    fb = fn(sb);
    ++iters;
    if ((0.0 < fb && 0.0 < fc) || (fb <= 0.0 && fc <= 0.0)) {
      c = sa;
      fc = fa;
      e = sb - sa;
      d = e;
    }
  }
  if (fb == 0.0) return {sb, sb};
  return {sb, c};
}

// Con2PrimFactory/src/c2p_1DPalenzuela.hxx:260-304
template <class E>
inline double palenzuela_f(const E &eos, const Atmo &atmo,
                           const C2PParams &cp, double Ssq, double Bsq,
                           double BiSi, double x, const Cons &cv) {
  const double q = cv.tau / cv.dens;
  const double r = Ssq / pow2(cv.dens);
  const double s = Bsq / cv.dens;
  const double t = BiSi / std::pow(cv.dens, 3. / 2.);
  double Wm2 = 1.0 - (x * x * r + (2.0 * x + s) * t * t) /
                         (x * x * (x + s) * (x + s));
  Wm2 = std::fmin(std::fmax(Wm2, 1e-10), 1);
  const double W = std::pow(Wm2, -0.5);
  const double rho = cv.dens / W;
  double eps = W - 1.0 + (1.0 - W * W) * x / W +
               W * (q - s + t * t / (2 * x * x) + s / (2 * W * W));
  if (cp.use_temp) {
    eps = smax(eps, atmo.eps);
    eps = smin(eps, eos.eps_max);
  }
  const double P = eos.press_from_rho_eps_ye(rho, eps, cv.DYe / cv.dens);
  return (x - (1.0 + eps + P / rho) * W);
}

// Con2PrimFactory/src/c2p_1DPalenzuela.hxx:306-556
template <class E>
inline void palenzuela_solve(const E &eos, const Atmo &atmo,
                             const C2PParams &cp, Prim &pv, Cons &cv,
                             const Sym3 &glo, Report &rep) {
  rep.iters = 0;
  rep.adjust_cons = false;
  rep.set_atmo = false;
  rep.status = ST_SUCCESS;
  const double detg = calc_det(glo);
  const double sqrt_detg = std::sqrt(detg);
  const bool minor1 = glo.xx > 0.0;
  const bool minor2 = glo.xx * glo.yy - glo.xy * glo.xy > 0.0;
  const bool minor3 = detg > 0.0;
  if (!(minor1 && minor2 && minor3)) {
    rep.status = ST_INVALID_DETG; rep.adjust_cons = true;
    prim_set_nan(pv); cv.set_nan();
    return;
  }
  const Sym3 gup = calc_inv(glo, detg);
  cv.divide(sqrt_detg);
  const double *m = cv.mom;
  double Ssq = m[0] * (gup.xx * m[0] + gup.xy * m[1] + gup.xz * m[2]);
  Ssq += m[1] * (gup.xy * m[0] + gup.yy * m[1] + gup.yz * m[2]);
  Ssq += m[2] * (gup.xz * m[0] + gup.yz * m[1] + gup.zz * m[2]);
  double B_low[3];
  lower(glo, cv.dB, B_low);
  const double Bsq = dot3(B_low, cv.dB);
  const double BiSi = cv.dB[0] * m[0] + cv.dB[1] * m[1] + cv.dB[2] * m[2];
  if ((!std::isfinite(cv.dens)) || (!std::isfinite(Ssq)) ||
      (!std::isfinite(Bsq)) || (!std::isfinite(BiSi)) ||
      (!std::isfinite(cv.DYe)) || (!std::isfinite(cv.DEnt))) {
    rep.status = ST_NANS_IN_CONS; rep.adjust_cons = true;
    prim_set_nan(pv); cv.set_nan();
    return;
  }
  if (Bsq > cp.Bsq_lim) {
    rep.status = ST_B_LIMIT; rep.adjust_cons = true;
    prim_set_nan(pv); cv.set_nan();
    return;
  }
  const double log2 = std::log(2.0);
  const int minbits = int(std::fabs(std::log(cp.tol)) / log2);
  const double tol0 = std::ldexp(double(1.0), -minbits);
  const double q = cv.tau / cv.dens;
  const double s = Bsq / cv.dens;
  double a = smax(1e-10, 1.0 + q - s);
  double b = 2.0 + 2.0 * q - s;
  auto fn = [&](double x) {
    return palenzuela_f(eos, atmo, cp, Ssq, Bsq, BiSi, x, cv);
  };
  if (fn(a) * fn(b) > 0) { // dominant energy check, widen once
    b = 3.0 + 3.0 * q - 1.5 * s;
  }
  const auto res = brent(fn, a, b, minbits, cp.max_iter, rep.iters);
  const double x_sol = 0.5 * (res.first + res.second); // bracket midpoint
  // xPalenzuelaToPrim, c2p_1DPalenzuela.hxx:142-258
  {
    const double r = Ssq / pow2(cv.dens);
    const double t = BiSi / std::pow(cv.dens, 3. / 2.);
    double Wm2 = 1.0 - (x_sol * x_sol * r + (2.0 * x_sol + s) * t * t) /
                           (x_sol * x_sol * (x_sol + s) * (x_sol + s));
    Wm2 = std::fmin(std::fmax(Wm2, 1e-10), 1);
    const double W = std::pow(Wm2, -0.5);
    pv.rho = cv.dens / W;
    pv.eps = W - 1.0 + (1.0 - W * W) * x_sol / W +
             W * (q - s + t * t / (2 * x_sol * x_sol) + s / (2.0 * W * W));
    if (cp.use_temp) {
      pv.eps = smax(pv.eps, atmo.eps);
      pv.eps = smin(pv.eps, eos.eps_max);
    }
    const double Z_Sol = x_sol * pv.rho * W;
    vel_from_Z(Z_Sol, W, Bsq, BiSi, cv, gup, glo, cp.use_z, pv);
    pv.Ye = cv.DYe / cv.dens;
    pv.press = eos.press_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
    pv.temperature = eos.temp_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
    pv.entropy = eos.kappa_from_rho_eps_ye(pv.rho, pv.eps, pv.Ye);
    for (int i = 0; i < 3; ++i) pv.B[i] = cv.dB[i];
  }
  if (pv.rho <= 0.0) {
    rep.status = ST_RANGE_RHO; rep.adjust_cons = true;
    cv.scale(sqrt_detg);
    return;
  }
  if (pv.eps <= 0.0) {
    rep.status = ST_RANGE_EPS; rep.adjust_cons = true;
    cv.scale(sqrt_detg);
    return;
  }
  if (std::fabs(res.first - res.second) >
      tol0 * smin(std::fabs(res.first), std::fabs(res.second))) {
    Cons ck;
    ck.from_prim(pv, glo);
    ck.divide(sqrt_detg);
    const double small = 1e-50;
    double e = pow2((ck.dens - cv.dens) / (cv.dens + small));
    e = smax(e, pow2((ck.mom[0] - cv.mom[0]) / (cv.mom[0] + small)));
    e = smax(e, pow2((ck.mom[1] - cv.mom[1]) / (cv.mom[1] + small)));
    e = smax(e, pow2((ck.mom[2] - cv.mom[2]) / (cv.mom[2] + small)));
    e = smax(e, pow2((ck.tau - cv.tau) / (cv.tau + small)));
    const double max_error = std::sqrt(e);
    if (max_error > cp.cons_error) {
      rep.status = ST_ROOT_FAIL_CONV; rep.adjust_cons = true;
      cv.scale(sqrt_detg);
      return;
    }
  }
  cv.DEnt = cv.dens * pv.entropy;
  if (pv.rho < atmo.rho_cut) {
    rep.status = ST_SUCCESS; rep.set_atmo = true; rep.adjust_cons = true;
    atmo.set(pv, cv, glo);
    return;
  }
  prims_floors(eos, atmo, cp, pv, cv, glo, rep);
  if (rep.adjust_cons) {
    cv.from_prim(pv, glo);
  } else {
    cv.scale(sqrt_detg);
  }
}

// Con2PrimFactory/src/c2p_1DEntropy.hxx:229-255 (Appendix A.4 of
// arXiv:1112.0568): x - rho(x), with h from the conserved entropy
inline double entropy_f(const IdealGas &eos, double Ssq, double Bsq, double BiSi, double x,
                        const Cons &cv) {
  const double ent = cv.DEnt / cv.dens;
  const double ye = cv.DYe / cv.dens;
  const double press = eos.press_from_rho_kappa_ye(x, ent, ye);
  const double eps = eos.eps_from_rho_kappa_ye(x, ent, ye);
  const double lor = cv.dens / x;
  const double h = (1.0 + eps + press / x);
  const double W = x * h * lor * lor;
  const double Sfsq = (W * W * Ssq + BiSi * BiSi * (Bsq + 2.0 * W)) / ((W + Bsq) * (W + Bsq));
  const double rho = cv.dens / std::sqrt(1.0 + Sfsq / (cv.dens * cv.dens * h * h));
  return x - rho;
}

// Con2PrimFactory/src/c2p_1DEntropy.hxx:257-466 (ideal gas: the tabulated
// EOS of the reference refuses p and eps from (rho, kappa))
inline void entropy_solve(const IdealGas &eos, const Atmo &atmo, const C2PParams &cp,
                          Prim &pv, Cons &cv, const Sym3 &glo, Report &rep) {
  rep.iters = 0;
  rep.adjust_cons = false;
  rep.set_atmo = false;
  rep.status = ST_SUCCESS;
  const double detg = calc_det(glo);
  const double sqrt_detg = std::sqrt(detg);
  const bool minor1 = glo.xx > 0.0;
  const bool minor2 = glo.xx * glo.yy - glo.xy * glo.xy > 0.0;
  const bool minor3 = detg > 0.0;
  if (!(minor1 && minor2 && minor3)) {
    rep.status = ST_INVALID_DETG; rep.adjust_cons = true;
    prim_set_nan(pv); cv.set_nan();
    return;
  }
  const Sym3 gup = calc_inv(glo, detg);
  cv.divide(sqrt_detg);
  const double *m = cv.mom;
  double Ssq = m[0] * (gup.xx * m[0] + gup.xy * m[1] + gup.xz * m[2]);
  Ssq += m[1] * (gup.xy * m[0] + gup.yy * m[1] + gup.yz * m[2]);
  Ssq += m[2] * (gup.xz * m[0] + gup.yz * m[1] + gup.zz * m[2]);
  double B_low[3];
  lower(glo, cv.dB, B_low);
  const double Bsq = dot3(B_low, cv.dB);
  const double BiSi = cv.dB[0] * m[0] + cv.dB[1] * m[1] + cv.dB[2] * m[2];
  if ((!std::isfinite(cv.dens)) || (!std::isfinite(Ssq)) || (!std::isfinite(Bsq)) ||
      (!std::isfinite(BiSi)) || (!std::isfinite(cv.DYe)) || (!std::isfinite(cv.DEnt))) {
    rep.status = ST_NANS_IN_CONS; rep.adjust_cons = true;
    prim_set_nan(pv); cv.set_nan();
    return;
  }
  if (Bsq > cp.Bsq_lim) {
    rep.status = ST_B_LIMIT; rep.adjust_cons = true;
    prim_set_nan(pv); cv.set_nan();
    return;
  }
  // (A59): the root lies in [D / sqrt(1 + S^2/D^2), D]
  const double a = cv.dens / std::sqrt(1.0 + Ssq / (cv.dens * cv.dens));
  const double b = cv.dens;
  auto fn = [&](double x) { return entropy_f(eos, Ssq, Bsq, BiSi, x, cv); };
  const double log2 = std::log(2.0);
  const int minbits = int(std::fabs(std::log(cp.tol)) / log2);
  const double tol0 = std::ldexp(double(1.0), -minbits);
  const auto res = brent(fn, a, b, minbits, cp.max_iter, rep.iters);
  const double x_sol = 0.5 * (res.first + res.second);
  // xEntropyToPrim, c2p_1DEntropy.hxx:142-226
  pv.rho = x_sol;
  pv.entropy = cv.DEnt / cv.dens;
  pv.Ye = cv.DYe / cv.dens;
  pv.w_lor = cv.dens / x_sol;
  pv.press = eos.press_from_rho_kappa_ye(x_sol, pv.entropy, pv.Ye);
  pv.eps = eos.eps_from_rho_kappa_ye(x_sol, pv.entropy, pv.Ye);
  pv.temperature = eos.temp_from_rho_eps_ye(x_sol, pv.eps, pv.Ye);
  const double Z_Sol = (x_sol + pv.eps * x_sol + pv.press) * pv.w_lor * pv.w_lor;
  vel_from_Z(Z_Sol, pv.w_lor, Bsq, BiSi, cv, gup, glo, cp.use_z, pv);
  for (int i = 0; i < 3; ++i) pv.B[i] = cv.dB[i];
  if (pv.rho <= 0.0) {
    rep.status = ST_RANGE_RHO; rep.adjust_cons = true;
    cv.scale(sqrt_detg);
    return;
  }
  if (std::fabs(res.first - res.second) >
      tol0 * smin(std::fabs(res.first), std::fabs(res.second))) {
    // tau is not checked: the inversion uses the conserved entropy
    Cons ck;
    ck.from_prim(pv, glo);
    ck.divide(sqrt_detg);
    const double small = 1e-50;
    double e = pow2((ck.dens - cv.dens) / (cv.dens + small));
    e = smax(e, pow2((ck.mom[0] - cv.mom[0]) / (cv.mom[0] + small)));
    e = smax(e, pow2((ck.mom[1] - cv.mom[1]) / (cv.mom[1] + small)));
    e = smax(e, pow2((ck.mom[2] - cv.mom[2]) / (cv.mom[2] + small)));
    if (std::sqrt(e) >= cp.cons_error) {
      rep.status = ST_ROOT_FAIL_CONV; rep.adjust_cons = true;
      cv.scale(sqrt_detg);
      return;
    }
  }
  // tau from the primitives (:427-436)
  double v_low[3];
  lower(glo, pv.vel, v_low);
  const double bst = pv.w_lor * dot3(pv.B, v_low);
  const double bs2 = (Bsq + bst * bst) / (pv.w_lor * pv.w_lor);
  cv.tau = (pv.w_lor * pv.w_lor * (pv.rho * (1.0 + pv.eps) + pv.press + bs2) -
            (pv.press + 0.5 * bs2) - bst * bst) -
           cv.dens;
  if (pv.rho < atmo.rho_cut) {
    rep.status = ST_SUCCESS; rep.set_atmo = true; rep.adjust_cons = true;
    atmo.set(pv, cv, glo);
    return;
  }
  prims_floors(eos, atmo, cp, pv, cv, glo, rep);
  if (rep.adjust_cons) {
    cv.from_prim(pv, glo);
  } else {
    cv.scale(sqrt_detg);
  }
}

// Atmosphere of AsterX/src/con2prim.cxx:82-133 at radius r.
struct AtmoParams {
  double rho_abs_min, atmo_tol, p_atmo, t_atmo, Ye_atmo, r_atmo, n_rho_atmo,
      n_press_atmo, n_temp_atmo;
  bool thermal_eos_atmo, use_press_atmo;
};
template <class E>
inline Atmo make_atmo(const E &eos, const Polytrope &poly,
                      const AtmoParams &ap, double r) {
  double rho_atm = (r > ap.r_atmo)
                       ? (ap.rho_abs_min * std::pow((ap.r_atmo / r), ap.n_rho_atmo))
                       : ap.rho_abs_min;
  rho_atm = smax(eos.rho_min, rho_atm);
  double press_atm = 0.0, eps_atm = 0.0, temp_atm = 0.0;
  if (ap.thermal_eos_atmo) {
    if (ap.use_press_atmo) {
      press_atm = (r > ap.r_atmo)
                      ? (ap.p_atmo * std::pow(ap.r_atmo / r, ap.n_press_atmo))
                      : ap.p_atmo;
      press_atm = smax(eos.press_from_rho_temp_ye(rho_atm, eos.temp_min, ap.Ye_atmo), press_atm);
      eps_atm = eos.eps_from_rho_press_ye(rho_atm, press_atm, ap.Ye_atmo);
      temp_atm = eos.temp_from_rho_eps_ye(rho_atm, eps_atm, ap.Ye_atmo);
    } else {
      temp_atm = (r > ap.r_atmo)
                     ? (ap.t_atmo * std::pow(ap.r_atmo / r, ap.n_temp_atmo))
                     : ap.t_atmo;
      temp_atm = smax(eos.temp_min, temp_atm);
      press_atm = eos.press_from_rho_temp_ye(rho_atm, temp_atm, ap.Ye_atmo);
      eps_atm = eos.eps_from_rho_temp_ye(rho_atm, temp_atm, ap.Ye_atmo);
    }
  } else {
    const double gm1 = poly.gm1_from_rho(rho_atm);
    eps_atm = poly.sed_from_gm1(gm1);
    eps_atm = smin(smax(eos.eps_min, eps_atm), eos.eps_max);
    press_atm = eos.press_from_rho_eps_ye(rho_atm, eps_atm, ap.Ye_atmo);
  }
  const double entropy_atm = eos.kappa_from_rho_eps_ye(rho_atm, eps_atm, ap.Ye_atmo);
  const double rho_cut = rho_atm * (1 + ap.atmo_tol);
  return Atmo{rho_atm, eps_atm, ap.Ye_atmo, press_atm, temp_atm, entropy_atm,
              rho_cut};
}

// One cell of AsterX_Con2Prim_typeEoS (AsterX/src/con2prim.cxx:81-453).
struct C2PCellIn {
  double r;          // radial distance of the cell centre
  double alp;        // v2c-averaged lapse
  Sym3 glo;          // v2c-averaged metric
  double mask;       // v2c-averaged aster_mask_vc (1 when use_mask = no)
  Cons cv;           // densitized conserved variables incl. dB
  double saved_rho, saved_eps, saved_Ye, saved_vel[3], temperature;
};
struct C2PCellOut {
  Prim pv;
  Cons cv;
  double flag, mask_cc;
  int iters;
};
struct C2PDriver {
  int prime, second; // ASX_C2P_*
  bool excise;
  bool entropy_fix; // use_entropy_fix: Entropy after both solvers failed
  AtmoParams ap;
  C2PParams cp;
};

template <class E>
inline void con2prim_cell(const E &eos, const Polytrope &poly,
                          const C2PDriver &dp, const C2PCellIn &in,
                          C2PCellOut &out) {
  const Atmo atmo = make_atmo(eos, poly, dp.ap, in.r);
  const Sym3 &glo = in.glo;
  double mask_local = in.mask;
  const double detg = calc_det(glo);
  const double sqrt_detg = std::sqrt(detg);
  double v_up[3] = {in.saved_vel[0], in.saved_vel[1], in.saved_vel[2]};
  double v_low[3];
  lower(glo, v_up, v_low);
  double zsq = 0.0;
  const double vsq = dot3(v_low, v_up);
  if (vsq >= 1.0) {
    const double wlim = std::sqrt(1.0 + dp.cp.vw_lim * dp.cp.vw_lim);
    const double vlim = dp.cp.vw_lim / wlim;
    const double s = vlim / std::sqrt(vsq);
    for (int i = 0; i < 3; ++i) v_up[i] *= s;
    for (int i = 0; i < 3; ++i) v_low[i] *= s;
    zsq = dp.cp.vw_lim * dp.cp.vw_lim;
  } else {
    zsq = vsq / (1.0 - vsq);
  }
  double wlor = std::sqrt(1.0 + zsq);
  Cons cv = in.cv;
  double Bup[3];
  for (int i = 0; i < 3; ++i) Bup[i] = cv.dB[i] / sqrt_detg;
  Prim pv;
  prim_set_nan(pv); // the reference leaves it uninitialised
  Prim seed;
  seed.rho = in.saved_rho;
  seed.eps = in.saved_eps;
  seed.Ye = in.saved_Ye;
  {
    // con2prim.cxx:199-209 passes the saved eps grid function by reference
    // to both calls (a table may clamp it for the second); the seed keeps
    // the value read first
    double e = in.saved_eps;
    seed.press = eos.press_from_rho_eps_ye(in.saved_rho, e, in.saved_Ye);
    seed.entropy = eos.kappa_from_rho_eps_ye(in.saved_rho, e, in.saved_Ye);
  }
  seed.temperature = in.temperature;
  for (int i = 0; i < 3; ++i) seed.vel[i] = v_up[i];
  seed.w_lor = wlor;
  for (int i = 0; i < 3; ++i) seed.B[i] = Bup[i];

  bool flag_local = true;
  int code = 0; // C2P_INIT
  bool call_c2p = true;
  if (cv.dens <= sqrt_detg * atmo.rho_cut) {
    for (int i = 0; i < 3; ++i) pv.B[i] = Bup[i];
    atmo.set(pv, cv, glo);
    atmo.set(seed);
    code = 4; // C2P_ATMO
    call_c2p = false;
  }
  if (in.alp < dp.cp.alp_thresh) mask_local = 0.0;
  out.mask_cc = mask_local;
  if (dp.excise) {
    if (mask_local != 1.0) {
      bh_interior<false>(eos, atmo, dp.cp, seed, cv, glo);
      pv = seed;
      call_c2p = false;
    }
  }
  Report rep1, rep2, rep3;
  cons_floors(cv, glo);
  out.iters = 0;
  if (call_c2p) {
    code = 1; // C2P_PRIME
    if (dp.prime == 1) {
      if constexpr (std::is_same<E, IdealGas>::value)
        noble_solve(eos, atmo, dp.cp, pv, seed, cv, glo, rep1);
    } else if (dp.prime == 2) {
      palenzuela_solve(eos, atmo, dp.cp, pv, cv, glo, rep1);
    } else if (dp.prime == 3) {
      if constexpr (std::is_same<E, IdealGas>::value)
        entropy_solve(eos, atmo, dp.cp, pv, cv, glo, rep1);
    }
    out.iters = rep1.iters;
    if (rep1.failed()) {
      code = 2; // C2P_SECOND
      if (dp.second == 1) {
        if constexpr (std::is_same<E, IdealGas>::value)
          noble_solve(eos, atmo, dp.cp, pv, seed, cv, glo, rep2);
      } else if (dp.second == 2) {
        palenzuela_solve(eos, atmo, dp.cp, pv, cv, glo, rep2);
      } else if (dp.second == 3) {
        if constexpr (std::is_same<E, IdealGas>::value)
          entropy_solve(eos, atmo, dp.cp, pv, cv, glo, rep2);
      }
    }
    bool failed = rep1.failed() && rep2.failed();
    if (failed && dp.entropy_fix) { // con2prim.cxx:310-313
      code = 3; // C2P_ENTROPY
      if constexpr (std::is_same<E, IdealGas>::value)
        entropy_solve(eos, atmo, dp.cp, pv, cv, glo, rep3);
      failed = rep3.failed();
    }
    if (failed) {
      flag_local = false;
      code = 6; // C2P_FAIL
      if (mask_local != 1.0) {
        bh_interior<false>(eos, atmo, dp.cp, seed, cv, glo);
        pv = seed;
      } else {
        for (int i = 0; i < 3; ++i) cv.dB[i] = sqrt_detg * Bup[i];
        for (int i = 0; i < 3; ++i) pv.B[i] = Bup[i];
        atmo.set(pv, cv, glo);
      }
    }
    if ((mask_local != 1.0) && flag_local) {
      bh_interior<true>(eos, atmo, dp.cp, pv, cv, glo);
    }
  }
  out.flag = code;
  out.pv = pv;
  out.cv = cv;
}

} // namespace orc
