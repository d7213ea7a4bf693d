// Device con2prim for a 3-parameter EOS (ideal gas or table): the per-cell chain of
// AsterX/src/con2prim.cxx:81-453 with Con2PrimFactory's Noble 2D
// (c2p_2DNoble.hxx:275-594) and Palenzuela 1D (c2p_1DPalenzuela.hxx:306-556)
// solvers, atmosphere (atmo.hxx), floors (c2p.hxx:86-369).
//
// Palenzuela's bracketing root finder is Brent's ZERO (Brent 1973) returning
// the final bracket, stopping at |b-c| <= 2^-minbits min(|b|,|c|), as the
// reference's call site documents (c2p_1DPalenzuela.hxx:380-393).
#pragma once

#include <type_traits>

#include "asx_device.hpp"
#include "asx_tab.hpp"

namespace asx {

// Waves per SIMD the primary-solver kernels are budgeted for: on a table
// the nested temperature solve would otherwise take every register (one
// wave per SIMD, with the latency of the table reads exposed)
template <class E> constexpr int c2p_waves() { return std::is_same<E, TabEOS>::value ? 2 : 1; }


enum C2PStatus {
  ST_SUCCESS = 0, ST_INVALID_DETG, ST_NEG_BSQR, ST_NANS_IN_CONS, ST_RANGE_RHO,
  ST_RANGE_EPS, ST_SPEED_LIMIT, ST_RANGE_YE, ST_B_LIMIT, ST_ROOT_FAIL_CONV,
  ST_ROOT_FAIL_BRACKET, ST_PREP_ROOT_FAIL_CONV, ST_PREP_ROOT_FAIL_BRACKET,
  ST_NOT_SET,
  ST_GUESS_MISS // a guided primary whose guessed bracket holds no root (internal)
};
struct Report {
  int status;
  bool adjust_cons, set_atmo;
  int iters;
  ASX_DEV bool failed() const { return status != ST_SUCCESS; }
};

struct Prim {
  double rho, eps, Ye, press, temperature, entropy;
  double vel[3];
  double w_lor;
  double B[3];
};
struct Cons {
  double dens, mom[3], tau, DYe, DEnt, dB[3];
  ASX_DEV void scale(double s) {
    dens *= s; tau *= s;
    mom[0] *= s; mom[1] *= s; mom[2] *= s;
    dB[0] *= s; dB[1] *= s; dB[2] *= s;
    DYe *= s; DEnt *= s;
  }
  ASX_DEV void divide(double s) {
    dens /= s; tau /= s;
    mom[0] /= s; mom[1] /= s; mom[2] /= s;
    dB[0] /= s; dB[1] /= s; dB[2] /= s;
    DYe /= s; DEnt /= s;
  }
  ASX_DEV void set_nan() {
    const double n = __builtin_nan("");
    dens = tau = DYe = DEnt = n;
    mom[0] = mom[1] = mom[2] = dB[0] = dB[1] = dB[2] = n;
  }
};
ASX_DEV void prim_nan(Prim &pv) {
  const double n = __builtin_nan("");
  pv.rho = pv.eps = pv.Ye = pv.press = pv.temperature = pv.entropy = n;
  pv.vel[0] = pv.vel[1] = pv.vel[2] = pv.w_lor = n;
  pv.B[0] = pv.B[1] = pv.B[2] = n;
}

// cons_vars::from_prim (Con2PrimFactory/src/cons.hxx:49-91), stored W
ASX_DEV void cons_from_prim(const Sym3 &g, double rho, double eps, double press,
                            const double vel[3], double w, const double Bv[3],
                            double Ye, double entropy, Cons &cv) {
  const double sqrt_detg = sqrt(det3(g));
  double v_low[3], B_low[3];
  lower(g, vel, v_low);
  lower(g, Bv, B_low);
  const double bst = w * dot3(Bv, v_low);
  double b_low[3];
  for (int i = 0; i < 3; ++i) b_low[i] = B_low[i] / w + bst * v_low[i];
  const double bs2 = (dot3(Bv, B_low) + bst * bst) / (w * w);
  cv.dens = sqrt_detg * rho * w;
  const double hW2 = w * w * (rho * (1.0 + eps) + press + bs2);
  for (int i = 0; i < 3; ++i) cv.mom[i] = sqrt_detg * (hW2 * v_low[i] - bst * b_low[i]);
  cv.tau = sqrt_detg * (w * w * (rho * (1.0 + eps) + press + bs2) -
                        (press + 0.5 * bs2) - bst * bst) -
           cv.dens;
  for (int i = 0; i < 3; ++i) cv.dB[i] = sqrt_detg * Bv[i];
  cv.DYe = cv.dens * Ye;
  cv.DEnt = cv.dens * entropy;
}
ASX_DEV void cons_from_prim(const Sym3 &g, const Prim &pv, Cons &cv) {
  cons_from_prim(g, pv.rho, pv.eps, pv.press, pv.vel, pv.w_lor, pv.B, pv.Ye,
                 pv.entropy, cv);
}

struct Atmo {
  double rho, eps, ye, press, temp, entropy, rho_cut;
  ASX_DEV void set(Prim &pv) const {
    pv.rho = rho; pv.eps = eps; pv.Ye = ye; pv.press = press;
    pv.temperature = temp; pv.entropy = entropy;
    pv.vel[0] = pv.vel[1] = pv.vel[2] = 0.0;
    pv.w_lor = 1.0;
  }
  ASX_DEV void set(Prim &pv, Cons &cv, const Sym3 &g) const {
    set(pv);
    const double sqrt_detg = sqrt(det3(g));
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

// Solver and atmosphere parameters, resolved once on the host.
struct C2PArgs {
  int prime, second, max_iter, excise, use_mask;
  int entropy_fix; // use_entropy_fix: the Entropy solver after both failed
  int use_z, use_temp, use_press_atmo, thermal_eos_atmo;
  double tol, alp_thresh, cons_error, vw_lim, w_lim, v_lim, Bsq_lim, rho_BH,
      eps_BH, vwlim_BH;
  double rho_abs_min, atmo_tol, p_atmo, t_atmo, Ye_atmo, r_atmo, n_rho_atmo,
      n_press_atmo, n_temp_atmo;
  double poly_gamma, poly_np1, poly_invn, poly_rho_p; // ID polytrope
  int minbits;   // Palenzuela: int(|ln tol| / ln 2), computed on the host
  double tol0;   // 2^-minbits
};

template <class E> ASX_DEV Atmo make_atmo(const E &eos, const C2PArgs &a, double r) {
  double rho_atm = (r > a.r_atmo) ? (a.rho_abs_min * pow((a.r_atmo / r), a.n_rho_atmo))
                                  : a.rho_abs_min;
  rho_atm = smax(eos.rho_min, rho_atm);
  double press_atm = 0.0, eps_atm = 0.0, temp_atm = 0.0;
  if (a.thermal_eos_atmo) {
    if (a.use_press_atmo) {
      press_atm = (r > a.r_atmo) ? (a.p_atmo * pow(a.r_atmo / r, a.n_press_atmo)) : a.p_atmo;
      press_atm = smax(eos.press_rt(rho_atm, eos.temp_min, a.Ye_atmo), press_atm);
      eps_atm = eos.eps_rp(rho_atm, press_atm, a.Ye_atmo);
      temp_atm = eos.temp_re(rho_atm, eps_atm, a.Ye_atmo);
    } else {
      temp_atm = (r > a.r_atmo) ? (a.t_atmo * pow(a.r_atmo / r, a.n_temp_atmo)) : a.t_atmo;
      temp_atm = smax(eos.temp_min, temp_atm);
      press_atm = eos.press_rt(rho_atm, temp_atm, a.Ye_atmo);
      eps_atm = eos.eps_rt(rho_atm, temp_atm, a.Ye_atmo);
    }
  } else {
    const double gm1 = a.poly_np1 * pow(rho_atm / a.poly_rho_p, a.poly_invn);
    eps_atm = gm1 / a.poly_gamma;
    eps_atm = smin(smax(eos.eps_min, eps_atm), eos.eps_max);
    press_atm = eos.press_re(rho_atm, eps_atm, a.Ye_atmo);
  }
  const double entropy_atm = eos.kappa_re(rho_atm, eps_atm, a.Ye_atmo);
  return Atmo{rho_atm, eps_atm, a.Ye_atmo, press_atm, temp_atm, entropy_atm,
              rho_atm * (1 + a.atmo_tol)};
}

// c2p::prims_floors_and_ceilings (Con2PrimFactory/src/c2p.hxx:86-216)
template <class E>
ASX_DEV void prims_floors(const E &eos, const Atmo &atmo, const C2PArgs &a,
                          Prim &pv, const Cons &cv, const Sym3 &glo, Report &rep) {
  bool recomp = false;
  if (pv.Ye < eos.ye_min) { pv.Ye = eos.ye_min; rep.adjust_cons = true; recomp = true; }
  if (pv.Ye > eos.ye_max) { pv.Ye = eos.ye_max; rep.adjust_cons = true; recomp = true; }
  double v_low[3];
  lower(glo, pv.vel, v_low);
  const double sol_v = sqrt(dot3(v_low, pv.vel));
  if (sol_v > a.v_lim) {
    pv.rho = cv.dens / a.w_lim;
    const double s = a.v_lim / sol_v;
    pv.vel[0] *= s; pv.vel[1] *= s; pv.vel[2] *= s;
    pv.w_lor = a.w_lim;
    rep.adjust_cons = true;
    if (a.use_temp) {
      recomp = true;
    } else {
      recomp = false;
      pv.eps = eos.eps_rp(pv.rho, pv.press, pv.Ye);
      pv.temperature = eos.temp_re(pv.rho, pv.eps, pv.Ye);
      pv.entropy = eos.kappa_re(pv.rho, pv.eps, pv.Ye);
    }
  }
  if (pv.rho > eos.rho_max) {
    pv.rho = eos.rho_max;
    rep.adjust_cons = true;
    if (a.use_temp) {
      recomp = true;
    } else {
      recomp = false;
      pv.eps = eos.eps_rp(pv.rho, pv.press, pv.Ye);
      pv.temperature = eos.temp_re(pv.rho, pv.eps, pv.Ye);
      pv.entropy = eos.kappa_re(pv.rho, pv.eps, pv.Ye);
    }
  }
  if (pv.temperature > eos.temp_max) {
    pv.temperature = eos.temp_max;
    recomp = true;
    rep.adjust_cons = true;
  }
  if (a.use_press_atmo) {
    if (pv.press < atmo.press) {
      pv.press = atmo.press;
      pv.eps = eos.eps_rp(pv.rho, pv.press, pv.Ye);
      pv.temperature = eos.temp_re(pv.rho, pv.eps, pv.Ye);
      pv.entropy = eos.kappa_re(pv.rho, pv.eps, pv.Ye);
      recomp = false;
      rep.adjust_cons = true;
    }
  } else if (pv.temperature < atmo.temp) {
    pv.temperature = atmo.temp;
    recomp = true;
    rep.adjust_cons = true;
  }
  if (recomp) {
    pv.eps = eos.eps_rt(pv.rho, pv.temperature, pv.Ye);
    pv.press = eos.press_rt(pv.rho, pv.temperature, pv.Ye);
    pv.entropy = eos.kappa_re(pv.rho, pv.eps, pv.Ye);
  }
}

// c2p::bh_interior (Con2PrimFactory/src/c2p.hxx:218-328)
template <bool limiting, class E>
ASX_DEV void bh_interior(const E &eos, const Atmo &atmo, const C2PArgs &a,
                         Prim &pv, Cons &cv, const Sym3 &glo) {
  const double wlim_BH = sqrt(1.0 + a.vwlim_BH * a.vwlim_BH);
  const double vlim_BH = a.vwlim_BH / wlim_BH;
  if (limiting) {
    bool recomp = false;
    if (pv.rho > a.rho_BH) { pv.rho = a.rho_BH; recomp = true; }
    if (pv.eps > a.eps_BH) { pv.eps = a.eps_BH; recomp = true; }
    const double sol_v = sqrt((pv.w_lor * pv.w_lor - 1.0)) / pv.w_lor;
    if (sol_v > vlim_BH) {
      const double s = vlim_BH / sol_v;
      pv.vel[0] *= s; pv.vel[1] *= s; pv.vel[2] *= s;
      pv.w_lor = wlim_BH;
      recomp = true;
    }
    if (recomp) {
      pv.temperature = eos.temp_re(pv.rho, pv.eps, pv.Ye);
      pv.press = eos.press_re(pv.rho, pv.eps, pv.Ye);
      pv.entropy = eos.kappa_re(pv.rho, pv.eps, pv.Ye);
      cons_from_prim(glo, pv, cv);
    }
  } else {
    pv.rho = a.rho_BH;
    pv.eps = a.eps_BH;
    pv.Ye = atmo.ye;
    pv.temperature = eos.temp_re(pv.rho, pv.eps, pv.Ye);
    pv.press = eos.press_re(pv.rho, pv.eps, pv.Ye);
    pv.entropy = eos.kappa_re(pv.rho, pv.eps, pv.Ye);
    const double detg = det3(glo);
    const Sym3 gup = inv3(glo, detg);
    const double Z = (pv.rho * (1.0 + pv.eps) + pv.press) * wlim_BH * wlim_BH;
    double B_low[3];
    lower(glo, pv.B, B_low);
    const double Bsq = dot3(B_low, pv.B);
    const double sq = sqrt(detg);
    double mom_low[3], mom_up[3];
    for (int i = 0; i < 3; ++i) mom_low[i] = cv.mom[i] / sq;
    lower(gup, mom_low, mom_up);
    const double S_old = sqrt(dot3(mom_low, mom_up)) + 1e-50;
    const double BiEsi = dot3(mom_low, pv.B) / S_old;
    const double Ssq_new = ((Z + Bsq) * (Z + Bsq) * vlim_BH * vlim_BH) /
                           (1.0 + BiEsi * BiEsi * (2.0 * Z + Bsq) / (Z * Z));
    const double S_new = sqrt(Ssq_new);
    for (int i = 0; i < 3; ++i) mom_low[i] *= S_new / S_old;
    for (int i = 0; i < 3; ++i) mom_up[i] *= S_new / S_old;
    for (int i = 0; i < 3; ++i) {
      pv.vel[i] = mom_up[i] / (Z + Bsq);
      pv.vel[i] += BiEsi * S_new * pv.B[i] / (Z * (Z + Bsq));
    }
    pv.w_lor = wlim_BH;
    cons_from_prim(glo, pv, cv);
  }
}

// c2p::cons_floors_and_ceilings (Con2PrimFactory/src/c2p.hxx:330-369)
ASX_DEV void cons_floors(Cons &cv, const Sym3 &glo) {
  const double detg = det3(glo);
  const double sqrt_detg = sqrt(detg);
  const Sym3 gup = inv3(glo, detg);
  double B_low[3];
  lower(glo, cv.dB, B_low);
  const double tau_lim = 0.5 * dot3(B_low, cv.dB) / sqrt_detg;
  if (cv.tau <= tau_lim) cv.tau = tau_lim;
  double mom_up[3];
  lower(gup, cv.mom, mom_up);
  const double mom2L = dot3(cv.mom, mom_up);
  const double slim = cv.dens + cv.tau;
  const double slim2 = slim * slim;
  if (mom2L > slim2) {
    const double s = sqrt(slim2 / mom2L);
    cv.mom[0] = cv.mom[0] * s;
    cv.mom[1] = cv.mom[1] * s;
    cv.mom[2] = cv.mom[2] * s;
  }
}

ASX_DEV void vel_from_Z(double Z, double W, double Bsq, double BiSi, const Cons &cv,
                        const Sym3 &gup, const Sym3 &glo, bool use_z, Prim &pv) {
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
    const double lor = sqrt(1.0 + (zx * zxd + zy * zyd + zz * zzd));
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

// Noble 2D Newton on (Z, v^2), c2p_2DNoble.hxx:275-594.  z0/v0 (if >= 0)
// replace the seed-derived initial guess (used by the learned guess).
// SEEDLESS: the guess is given (zguess >= 0) and the seed is not read at
// all (the saved primitives need not be loaded); same iterates otherwise.
template <bool SEEDLESS = false>
ASX_DEV void noble_solve(const IdealGas &eos, const Atmo &atmo, const C2PArgs &a,
                         Prim &pv, Prim &seed, Cons &cv, const Sym3 &glo,
                         Report &rep, double zguess = -1.0, double vguess = -1.0) {
  rep.iters = 0;
  rep.adjust_cons = false;
  rep.set_atmo = false;
  rep.status = ST_SUCCESS;
  const double G = eos.gamma;
  const double Zmin = eos.rho_min;
  const double detg = det3(glo);
  const double sqrt_detg = sqrt(detg);
  if (!((glo.xx > 0.0) && (glo.xx * glo.yy - glo.xy * glo.xy > 0.0) && (detg > 0.0))) {
    rep.status = ST_INVALID_DETG; rep.adjust_cons = true;
    prim_nan(pv); cv.set_nan();
    return;
  }
  const Sym3 gup = inv3(glo, detg);
  cv.divide(sqrt_detg);
  seed.B[0] = cv.dB[0]; seed.B[1] = cv.dB[1]; seed.B[2] = cv.dB[2];
  double mom_up[3], B_low[3];
  lower(gup, cv.mom, mom_up);
  const double Ssq = dot3(cv.mom, mom_up);
  lower(glo, seed.B, B_low);
  const double Bsq = dot3(B_low, seed.B);
  const double BiSi = dot3(cv.mom, seed.B);
  double vsq_seed = 0.0;
  if (SEEDLESS) {
  } else if (a.use_z) {
    double z_up[3], z_low[3];
    for (int i = 0; i < 3; ++i) z_up[i] = seed.vel[i] * seed.w_lor;
    lower(glo, z_up, z_low);
    const double zsq = dot3(z_low, z_up);
    const double w = sqrt(1. + zsq);
    vsq_seed = smin(zsq / w / w, 1. - 1.e-15);
    seed.w_lor = w;
  } else {
    double v_low[3];
    lower(glo, seed.vel, v_low);
    const double vsq = dot3(v_low, seed.vel);
    seed.w_lor = 1. / sqrt(1. - vsq);
    vsq_seed = vsq;
  }
  if (!isfinite(cv.dens) || !isfinite(Ssq) || !isfinite(Bsq) || !isfinite(BiSi) ||
      !isfinite(cv.DYe) || !isfinite(cv.DEnt)) {
    rep.status = ST_NANS_IN_CONS; rep.adjust_cons = true;
    prim_nan(pv); cv.set_nan();
    return;
  }
  if (Bsq > a.Bsq_lim) {
    rep.status = ST_B_LIMIT; rep.adjust_cons = true;
    prim_nan(pv); cv.set_nan();
    return;
  }
  double x0 = zguess, x1 = vguess;
  if (!SEEDLESS) {
    seed.rho = cv.dens / seed.w_lor;
    const double eps_last = smax(seed.eps, atmo.eps);
    seed.press = eos.press_re(seed.rho, eps_last);
    const double Z_seed =
        (seed.rho + eps_last * seed.rho + seed.press) * seed.w_lor * seed.w_lor;
    x0 = fabs(Z_seed);
    x1 = vsq_seed;
  }
  if (zguess >= 0.0) {
    x0 = zguess;
    x1 = vguess;
  }
  double xo0 = x0;
  const double dv = (1. - 1.e-10);
  if (x1 < 0.0) x1 = 0.0;
  else if (x1 >= 1.0) x1 = dv;
  if (x0 <= 0.0) x0 = fabs(x0) + Zmin;
  else if (x0 > 1e20) x0 = xo0;
  double errx = 1., df = 1., f = 1.;
  int k;
  for (k = 1; k <= a.max_iter; k++) {
    const double Z = x0;
    const double invZ = 1.0 / Z;
    const double Vsq = x1;
    const double Sdotn = -(cv.tau + cv.dens);
    const double sq1v = sqrt(1.0 - Vsq);
    const double p_tmp = ((Z * (1.0 - Vsq) - cv.dens * sq1v) * (G - 1.0) / (G));
    const double dPdvsq = ((-Z + cv.dens / (2.0 * sq1v)) * (G - 1.0) / G);
    const double dPdZ = ((1.0 - Vsq) * (G - 1.0) / G);
    const double j00 = -2 * (Vsq + BiSi * BiSi * invZ * invZ * invZ) * (Bsq + Z);
    const double j01 = -(Bsq + Z) * (Bsq + Z);
    const double j10 = -1.0 + dPdZ - BiSi * BiSi * invZ * invZ * invZ;
    const double j11 = -0.5 * Bsq + dPdvsq;
    const double r0 = Ssq - Vsq * (Bsq + Z) * (Bsq + Z) +
                      BiSi * BiSi * invZ * invZ * (Bsq + Z + Z);
    const double r1 = -Sdotn - 0.5 * Bsq * (1.0 + Vsq) + 0.5 * BiSi * BiSi * invZ * invZ -
                      Z + p_tmp;
    const double detjac =
        (Bsq + Z) * (j10 * (Bsq + Z) +
                     (Bsq - 2.0 * dPdvsq) * (BiSi * BiSi * invZ * invZ + Vsq * Z) * invZ);
    const double detjac_inv = 1.0 / detjac;
    const double dx0 = -(j11 * r0 - j01 * r1) * detjac_inv;
    const double dx1 = -(-j10 * r0 + j00 * r1) * detjac_inv;
    df = -r0 * r0 - r1 * r1;
    f = -0.5 * (df);
    xo0 = x0;
    x0 += dx0;
    x1 += dx1;
    if (x1 < 0.0) x1 = 0.0;
    else if (x1 >= 1.0) x1 = dv;
    if (x0 <= 0.0) x0 = fabs(x0) + Zmin;
    else if (x0 > 1e20) x0 = xo0;
    errx = (x0 == 0.) ? fabs(dx0) : fabs(dx0 / x0);
    if (fabs(errx) <= a.tol) break;
  }
  rep.iters = k;
  if (fabs(errx) > a.tol) {
    rep.status = ST_ROOT_FAIL_CONV; rep.adjust_cons = true;
    cv.scale(sqrt_detg);
    return;
  }
  if (!isfinite(f) || !isfinite(df)) {
    rep.status = ST_ROOT_FAIL_BRACKET; rep.adjust_cons = true;
  }
  const double Z_Sol = x0, vsq_Sol = x1;
  const double W_Sol = 1.0 / sqrt(1.0 - vsq_Sol);
  pv.rho = cv.dens / W_Sol;
  vel_from_Z(Z_Sol, W_Sol, Bsq, BiSi, cv, gup, glo, a.use_z, pv);
  pv.eps = (Z_Sol / pv.w_lor / pv.w_lor / pv.rho - 1.0) / G;
  pv.Ye = cv.DYe / cv.dens;
  pv.press = eos.press_re(pv.rho, pv.eps);
  pv.temperature = eos.temp_re(pv.eps);
  pv.entropy = eos.kappa(pv.rho, pv.eps);
  pv.B[0] = cv.dB[0]; pv.B[1] = cv.dB[1]; pv.B[2] = cv.dB[2];
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
  cv.DEnt = cv.dens * pv.entropy;
  if (pv.rho < atmo.rho_cut) {
    rep.status = ST_SUCCESS; rep.set_atmo = true; rep.adjust_cons = true;
    atmo.set(pv, cv, glo);
    return;
  }
  prims_floors(eos, atmo, a, pv, cv, glo, rep);
  if (rep.adjust_cons) cons_from_prim(glo, pv, cv);
  else cv.scale(sqrt_detg);
}

// Palenzuela residual f(x), c2p_1DPalenzuela.hxx:260-304
template <class E> struct PalF {
  const E &eos;
  const Atmo &atmo;
  const C2PArgs &a;
  double q, r, s, t, dens, ye;
  // the table's temperature cell of the previous evaluation (successive
  // Brent iterates of one zone move T little; the solve is exact either way)
  mutable int thint = -1;
  ASX_DEV double operator()(double x) const {
    double Wm2 = 1.0 - (x * x * r + (2.0 * x + s) * t * t) / (x * x * (x + s) * (x + s));
    Wm2 = fmin(fmax(Wm2, 1e-10), 1);
    const double W = pow(Wm2, -0.5);
    const double rho = dens / W;
    double eps = W - 1.0 + (1.0 - W * W) * x / W +
                 W * (q - s + t * t / (2 * x * x) + s / (2 * W * W));
    if (a.use_temp) {
      eps = smax(eps, atmo.eps);
      eps = smin(eps, eos.eps_max);
    }
    const double P = eos.press_re(rho, eps, ye, &thint);
    return (x - (1.0 + eps + P / rho) * W);
  }
};

// Brent's ZERO (Brent 1973; the form of EOSX/src/utils/eos_brent.hxx)
// returning the final bracket (b, c): stops when |b - c| <= rel*min(|b|,|c|),
// f(b) == 0, or after maxiters evaluations; steps are at least half that
// width so the bracket collapses once b has converged.  fa0, fb0 are f(a),
// f(b), which the caller has evaluated to choose the bracket (the residual
// is deterministic, so this equals evaluating them again).
template <typename F>
ASX_DEV void brent(const F &fn, double a, double b, double fa0, double fb0, double rel,
                   int maxiters, double &ra, double &rb, int &iters) {
  double sa = a, sb = b;
  double fa = fa0, fb = fb0;
  iters = 0;
  if (fa == 0.0) { ra = rb = sa; return; }
  if (fb == 0.0) { ra = rb = sb; return; }
  if (fa * fb > 0.0) { ra = sa; rb = sb; return; }
  double c = sa, fc = fa, e = sb - sa, d = e;
  for (;;) {
    if (fabs(fc) < fabs(fb)) {
      sa = sb; sb = c; c = sa;
      fa = fb; fb = fc; fc = fa;
    }
    const double tol = 0.5 * rel * smin(fabs(sb), fabs(c));
    const double m = 0.5 * (c - sb);
    if (fabs(m) <= tol || fb == 0.0 || iters >= maxiters) break;
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
    fb = fn(sb);
    ++iters;
    if ((0.0 < fb && 0.0 < fc) || (fb <= 0.0 && fc <= 0.0)) {
      c = sa;
      fc = fa;
      e = sb - sa;
      d = e;
    }
  }
  if (fb == 0.0) { ra = rb = sb; return; }
  ra = sb;
  rb = c;
}

// Palenzuela 1D, c2p_1DPalenzuela.hxx:306-556
template <class E>
ASX_DEV void palenzuela_solve(const E &eos, const Atmo &atmo, const C2PArgs &a,
                              Prim &pv, Cons &cv, const Sym3 &glo, Report &rep,
                              double xguess = -1.0, double xrel = 0.0,
                              bool miss_fails = false) {
  rep.iters = 0;
  rep.adjust_cons = false;
  rep.set_atmo = false;
  rep.status = ST_SUCCESS;
  const double detg = det3(glo);
  const double sqrt_detg = sqrt(detg);
  if (!((glo.xx > 0.0) && (glo.xx * glo.yy - glo.xy * glo.xy > 0.0) && (detg > 0.0))) {
    rep.status = ST_INVALID_DETG; rep.adjust_cons = true;
    prim_nan(pv); cv.set_nan();
    return;
  }
  const Sym3 gup = inv3(glo, detg);
  cv.divide(sqrt_detg);
  const double *m = cv.mom;
  double Ssq = m[0] * (gup.xx * m[0] + gup.xy * m[1] + gup.xz * m[2]);
  Ssq += m[1] * (gup.xy * m[0] + gup.yy * m[1] + gup.yz * m[2]);
  Ssq += m[2] * (gup.xz * m[0] + gup.yz * m[1] + gup.zz * m[2]);
  double B_low[3];
  lower(glo, cv.dB, B_low);
  const double Bsq = dot3(B_low, cv.dB);
  const double BiSi = cv.dB[0] * m[0] + cv.dB[1] * m[1] + cv.dB[2] * m[2];
  if (!isfinite(cv.dens) || !isfinite(Ssq) || !isfinite(Bsq) || !isfinite(BiSi) ||
      !isfinite(cv.DYe) || !isfinite(cv.DEnt)) {
    rep.status = ST_NANS_IN_CONS; rep.adjust_cons = true;
    prim_nan(pv); cv.set_nan();
    return;
  }
  if (Bsq > a.Bsq_lim) {
    rep.status = ST_B_LIMIT; rep.adjust_cons = true;
    prim_nan(pv); cv.set_nan();
    return;
  }
  const double q = cv.tau / cv.dens;
  const double s = Bsq / cv.dens;
  const double r = Ssq / pow2(cv.dens);
  const double t = BiSi / pow(cv.dens, 3. / 2.);
  const PalF<E> fn{eos, atmo, a, q, r, s, t, cv.dens, cv.DYe / cv.dens};
  const double xa = smax(1e-10, 1.0 + q - s);
  double xb = 2.0 + 2.0 * q - s;
  double lo = xa, hi = xb, flo = 0.0, fhi = 0.0;
  bool hinted = false;
  if (xguess > 0.0) {
    // a bracket of relative half-width xrel around the guess of x = h W,
    // inside the widest bracket the reference may take, used when the
    // residual changes sign across it (the root it holds is the root)
    const double xw = 3.0 + 3.0 * q - 1.5 * s;
    const double gl = smax(xa, xguess * (1.0 - xrel)), gh = smin(xw, xguess * (1.0 + xrel));
    if (gl < gh) {
      flo = fn(gl);
      fhi = fn(gh);
      if (flo * fhi <= 0.0) {
        lo = gl;
        hi = gh;
        hinted = true;
      }
    }
  }
  if (xguess > 0.0 && !hinted && miss_fails) {
    // the zone goes to the full chain (a wavefront then keeps to the narrow
    // brackets instead of waiting on a lane that searches the wide one)
    rep.status = ST_GUESS_MISS;
    return;
  }
  if (!hinted) {
    flo = fn(xa);
    fhi = fn(xb);
    if (flo * fhi > 0) {
      hi = 3.0 + 3.0 * q - 1.5 * s;
      fhi = fn(hi);
    }
  }
  double ra, rb;
  brent(fn, lo, hi, flo, fhi, a.tol0, a.max_iter, ra, rb, rep.iters);
  const double x = 0.5 * (ra + rb);
  {
    double Wm2 = 1.0 - (x * x * r + (2.0 * x + s) * t * t) / (x * x * (x + s) * (x + s));
    Wm2 = fmin(fmax(Wm2, 1e-10), 1);
    const double W = pow(Wm2, -0.5);
    pv.rho = cv.dens / W;
    pv.eps = W - 1.0 + (1.0 - W * W) * x / W +
             W * (q - s + t * t / (2 * x * x) + s / (2.0 * W * W));
    if (a.use_temp) {
      pv.eps = smax(pv.eps, atmo.eps);
      pv.eps = smin(pv.eps, eos.eps_max);
    }
    vel_from_Z(x * pv.rho * W, W, Bsq, BiSi, cv, gup, glo, a.use_z, pv);
    pv.Ye = cv.DYe / cv.dens;
    eos.state_re(pv.rho, pv.eps, pv.Ye, pv.press, pv.temperature, pv.entropy, &fn.thint);
    pv.B[0] = cv.dB[0]; pv.B[1] = cv.dB[1]; pv.B[2] = cv.dB[2];
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
  if (fabs(ra - rb) > a.tol0 * smin(fabs(ra), fabs(rb))) {
    Cons ck;
    cons_from_prim(glo, pv, ck);
    ck.divide(sqrt_detg);
    const double small = 1e-50;
    double e = pow2((ck.dens - cv.dens) / (cv.dens + small));
    e = smax(e, pow2((ck.mom[0] - cv.mom[0]) / (cv.mom[0] + small)));
    e = smax(e, pow2((ck.mom[1] - cv.mom[1]) / (cv.mom[1] + small)));
    e = smax(e, pow2((ck.mom[2] - cv.mom[2]) / (cv.mom[2] + small)));
    e = smax(e, pow2((ck.tau - cv.tau) / (cv.tau + small)));
    if (sqrt(e) > a.cons_error) {
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
  prims_floors(eos, atmo, a, pv, cv, glo, rep);
  if (rep.adjust_cons) cons_from_prim(glo, pv, cv);
  else cv.scale(sqrt_detg);
}

// Con2PrimFactory/src/c2p_1DEntropy.hxx:229-255 (Appendix A.4 of
// arXiv:1112.0568): x - rho(x), h from the conserved entropy
template <class E> struct EntF {
  const E &eos;
  double Ssq, Bsq, BiSi, dens, ent, ye;
  ASX_DEV double operator()(double x) const {
    const double press = eos.press_rk(x, ent, ye);
    const double eps = eos.eps_rk(x, ent, ye);
    const double lor = dens / x;
    const double h = (1.0 + eps + press / x);
    const double W = x * h * lor * lor;
    const double Sfsq = (W * W * Ssq + BiSi * BiSi * (Bsq + 2.0 * W)) / ((W + Bsq) * (W + Bsq));
    const double rho = dens / sqrt(1.0 + Sfsq / (dens * dens * h * h));
    return x - rho;
  }
};

// Entropy 1D, c2p_1DEntropy.hxx:257-466 (the ideal gas; refused for tables)
template <class E>
ASX_DEV void entropy_solve(const E &eos, const Atmo &atmo, const C2PArgs &a, Prim &pv,
                           Cons &cv, const Sym3 &glo, Report &rep) {
  rep.iters = 0;
  rep.adjust_cons = false;
  rep.set_atmo = false;
  rep.status = ST_SUCCESS;
  const double detg = det3(glo);
  const double sqrt_detg = sqrt(detg);
  if (!((glo.xx > 0.0) && (glo.xx * glo.yy - glo.xy * glo.xy > 0.0) && (detg > 0.0))) {
    rep.status = ST_INVALID_DETG; rep.adjust_cons = true;
    prim_nan(pv); cv.set_nan();
    return;
  }
  const Sym3 gup = inv3(glo, detg);
  cv.divide(sqrt_detg);
  const double *m = cv.mom;
  double Ssq = m[0] * (gup.xx * m[0] + gup.xy * m[1] + gup.xz * m[2]);
  Ssq += m[1] * (gup.xy * m[0] + gup.yy * m[1] + gup.yz * m[2]);
  Ssq += m[2] * (gup.xz * m[0] + gup.yz * m[1] + gup.zz * m[2]);
  double B_low[3];
  lower(glo, cv.dB, B_low);
  const double Bsq = dot3(B_low, cv.dB);
  const double BiSi = cv.dB[0] * m[0] + cv.dB[1] * m[1] + cv.dB[2] * m[2];
  if (!isfinite(cv.dens) || !isfinite(Ssq) || !isfinite(Bsq) || !isfinite(BiSi) ||
      !isfinite(cv.DYe) || !isfinite(cv.DEnt)) {
    rep.status = ST_NANS_IN_CONS; rep.adjust_cons = true;
    prim_nan(pv); cv.set_nan();
    return;
  }
  if (Bsq > a.Bsq_lim) {
    rep.status = ST_B_LIMIT; rep.adjust_cons = true;
    prim_nan(pv); cv.set_nan();
    return;
  }
  // (A59): the root lies in [D / sqrt(1 + S^2/D^2), D]
  const EntF<E> fn{eos, Ssq, Bsq, BiSi, cv.dens, cv.DEnt / cv.dens, cv.DYe / cv.dens};
  const double xa = cv.dens / sqrt(1.0 + Ssq / (cv.dens * cv.dens));
  const double xb = cv.dens;
  double ra, rb;
  brent(fn, xa, xb, fn(xa), fn(xb), a.tol0, a.max_iter, ra, rb, rep.iters);
  const double x = 0.5 * (ra + rb);
  // xEntropyToPrim, c2p_1DEntropy.hxx:142-226
  pv.rho = x;
  pv.entropy = cv.DEnt / cv.dens;
  pv.Ye = cv.DYe / cv.dens;
  pv.w_lor = cv.dens / x;
  pv.press = eos.press_rk(x, pv.entropy, pv.Ye);
  pv.eps = eos.eps_rk(x, pv.entropy, pv.Ye);
  pv.temperature = eos.temp_re(x, pv.eps, pv.Ye);
  const double Z_Sol = (x + pv.eps * x + pv.press) * pv.w_lor * pv.w_lor;
  vel_from_Z(Z_Sol, pv.w_lor, Bsq, BiSi, cv, gup, glo, a.use_z, pv);
  pv.B[0] = cv.dB[0]; pv.B[1] = cv.dB[1]; pv.B[2] = cv.dB[2];
  if (pv.rho <= 0.0) {
    rep.status = ST_RANGE_RHO; rep.adjust_cons = true;
    cv.scale(sqrt_detg);
    return;
  }
  if (fabs(ra - rb) > a.tol0 * smin(fabs(ra), fabs(rb))) {
    // tau is not checked: the inversion uses the conserved entropy
    Cons ck;
    cons_from_prim(glo, pv, ck);
    ck.divide(sqrt_detg);
    const double small = 1e-50;
    double e = pow2((ck.dens - cv.dens) / (cv.dens + small));
    e = smax(e, pow2((ck.mom[0] - cv.mom[0]) / (cv.mom[0] + small)));
    e = smax(e, pow2((ck.mom[1] - cv.mom[1]) / (cv.mom[1] + small)));
    e = smax(e, pow2((ck.mom[2] - cv.mom[2]) / (cv.mom[2] + small)));
    if (sqrt(e) >= a.cons_error) {
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
  prims_floors(eos, atmo, a, pv, cv, glo, rep);
  if (rep.adjust_cons) cons_from_prim(glo, pv, cv);
  else cv.scale(sqrt_detg);
}

// Inputs of one cell of AsterX_Con2Prim_typeEoS.
struct C2PIn {
  double r, alp, mask;
  Sym3 glo;
  Cons cv;
  double saved_rho, saved_eps, saved_Ye, saved_vel[3], temperature;
};
struct C2POut {
  Prim pv;
  Cons cv;
  double flag, mask_cc;
  int iters;
};

// AsterX/src/con2prim.cxx:81-416.
// PRIME < 0: the full chain, primary and backup solver chosen at run time
// from a.prime / a.second (the reference's behaviour); returns true.
// PRIME = ASX_C2P_NOBLE or ASX_C2P_PALENZUELA: only the primary solver,
// compiled for that method (a.prime must equal it); returns false, with
// `out` not set, when it fails: the zone then needs the full chain, which
// repeats the same primary solve before the backup.
// zguess, vguess (optional, > 0) guide the primary solver: Noble starts
// from Z = zguess, v^2 = vguess; Palenzuela brackets x = h W within the
// relative half-width vguess of zguess when the residual changes sign
// there; otherwise the primary-only form fails (the zone is queued for the
// full chain) and the full chain takes the reference's bracket.
// SEEDLESS (Noble primary only, zguess >= 0, no excision or mask): the
// saved primitives are neither read nor derived; the zone is the same
// otherwise.
template <int PRIME, class E, bool SEEDLESS = false>
ASX_DEV bool con2prim_zone(const E &eos, const C2PArgs &a, const C2PIn &in,
                           C2POut &out, double zguess = -1.0, double vguess = -1.0) {
  static_assert(!SEEDLESS || PRIME == ASX_C2P_NOBLE, "seedless: the Noble primary only");
  const Atmo atmo = make_atmo(eos, a, in.r);
  const Sym3 &glo = in.glo;
  double mask_local = in.mask;
  const double sqrt_detg = sqrt(det3(glo));
  Cons cv = in.cv;
  double Bup[3];
  for (int i = 0; i < 3; ++i) Bup[i] = cv.dB[i] / sqrt_detg;
  Prim pv;
  prim_nan(pv);
  Prim seed;
  if (SEEDLESS) {
    prim_nan(seed);
  } else {
    double v_up[3] = {in.saved_vel[0], in.saved_vel[1], in.saved_vel[2]};
    double v_low[3];
    lower(glo, v_up, v_low);
    double zsq;
    const double vsq = dot3(v_low, v_up);
    if (vsq >= 1.0) {
      const double wlim = sqrt(1.0 + a.vw_lim * a.vw_lim);
      const double vlim = a.vw_lim / wlim;
      const double s = vlim / sqrt(vsq);
      v_up[0] *= s; v_up[1] *= s; v_up[2] *= s;
      zsq = a.vw_lim * a.vw_lim;
    } else {
      zsq = vsq / (1.0 - vsq);
    }
    seed.rho = in.saved_rho;
    seed.eps = in.saved_eps;
    seed.Ye = in.saved_Ye;
    {
      // p and the entropy of the saved state; the EOS takes eps by reference
      // (a table may clamp it), the seed keeps the value as loaded
      double e = in.saved_eps, temp;
      eos.state_re(in.saved_rho, e, in.saved_Ye, seed.press, temp, seed.entropy);
    }
    seed.temperature = in.temperature;
    seed.vel[0] = v_up[0]; seed.vel[1] = v_up[1]; seed.vel[2] = v_up[2];
    seed.w_lor = sqrt(1.0 + zsq);
  }
  seed.B[0] = Bup[0]; seed.B[1] = Bup[1]; seed.B[2] = Bup[2];
  bool flag_local = true;
  int code = ASX_C2P_FLAG_INIT;
  bool call_c2p = true;
  if (cv.dens <= sqrt_detg * atmo.rho_cut) {
    pv.B[0] = Bup[0]; pv.B[1] = Bup[1]; pv.B[2] = Bup[2];
    atmo.set(pv, cv, glo);
    atmo.set(seed);
    code = ASX_C2P_FLAG_ATMO;
    call_c2p = false;
  }
  if (in.alp < a.alp_thresh) mask_local = 0.0;
  // the BH-interior paths start from the seed: the full chain takes them
  if (SEEDLESS && mask_local != 1.0) return false;
  out.mask_cc = mask_local;
  if (a.excise && mask_local != 1.0) {
    bh_interior<false>(eos, atmo, a, seed, cv, glo);
    pv = seed;
    call_c2p = false;
  }
  Report rep1{ST_NOT_SET, false, false, 0}, rep2{ST_NOT_SET, false, false, 0};
  cons_floors(cv, glo);
  out.iters = 0;
  if (call_c2p) {
    code = ASX_C2P_FLAG_PRIME;
    const int prime = PRIME < 0 ? a.prime : PRIME;
    if constexpr (std::is_same<E, IdealGas>::value) {
      if (prime == ASX_C2P_NOBLE)
        noble_solve<SEEDLESS>(eos, atmo, a, pv, seed, cv, glo, rep1, zguess, vguess);
      else if (prime == ASX_C2P_ENTROPY) entropy_solve(eos, atmo, a, pv, cv, glo, rep1);
      else palenzuela_solve(eos, atmo, a, pv, cv, glo, rep1, zguess, vguess, PRIME >= 0);
    } else {
      (void)prime; // Noble needs the ideal-fluid Gamma; refused for tables at the C ABI
      palenzuela_solve(eos, atmo, a, pv, cv, glo, rep1, zguess, vguess, PRIME >= 0);
    }
    out.iters = rep1.iters;
    if (PRIME >= 0 && rep1.failed()) return false;
    if (PRIME < 0 && rep1.failed()) {
      code = ASX_C2P_FLAG_SECOND;
      if constexpr (std::is_same<E, IdealGas>::value) {
        if (a.second == ASX_C2P_NOBLE) noble_solve(eos, atmo, a, pv, seed, cv, glo, rep2);
        else if (a.second == ASX_C2P_PALENZUELA) palenzuela_solve(eos, atmo, a, pv, cv, glo, rep2);
        else if (a.second == ASX_C2P_ENTROPY) entropy_solve(eos, atmo, a, pv, cv, glo, rep2);
      } else {
        if (a.second == ASX_C2P_PALENZUELA) palenzuela_solve(eos, atmo, a, pv, cv, glo, rep2);
      }
    }
    bool failed = rep1.failed() && rep2.failed();
    if constexpr (std::is_same<E, IdealGas>::value) {
      if (failed && a.entropy_fix) { // con2prim.cxx:310-313
        code = ASX_C2P_FLAG_ENTROPY;
        Report rep3{ST_NOT_SET, false, false, 0};
        entropy_solve(eos, atmo, a, pv, cv, glo, rep3);
        failed = rep3.failed();
      }
    }
    if (failed) {
      flag_local = false;
      code = ASX_C2P_FLAG_FAIL;
      if (mask_local != 1.0) {
        bh_interior<false>(eos, atmo, a, seed, cv, glo);
        pv = seed;
      } else {
        for (int i = 0; i < 3; ++i) cv.dB[i] = sqrt_detg * Bup[i];
        pv.B[0] = Bup[0]; pv.B[1] = Bup[1]; pv.B[2] = Bup[2];
        atmo.set(pv, cv, glo);
      }
    }
    if ((mask_local != 1.0) && flag_local) bh_interior<true>(eos, atmo, a, pv, cv, glo);
  }
  out.flag = code;
  out.pv = pv;
  out.cv = cv;
  return true;
}
template <class E>
ASX_DEV void con2prim_cell(const E &eos, const C2PArgs &a, const C2PIn &in, C2POut &out,
                           double zguess = -1.0, double vguess = -1.0) {
  (void)con2prim_zone<-1>(eos, a, in, out, zguess, vguess);
}

// Append `value` of every work-item with `fail` set to the queue: a
// wavefront ballot, one atomic per wavefront, lanes in order.
ASX_DEV void queue_push(bool fail, int value, int *queue, unsigned *qcount) {
  const uint64_t m = __ballot(fail);
  if (m) {
    const int lane = threadIdx.x & 63;
    const int leader = __builtin_ctzll(m);
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(qcount, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (fail) queue[base + __popcll(m & ((1ull << lane) - 1))] = value;
  }
}

// Independent zone z of a batch (layout of asx_c2p_batch: cons[8][nz],
// g[6][nz], seed[5][nz]); no grid position, lapse 1, no mask.
// SEED = false: the saved primitives are not read (left NaN), for the
// seedless primary solve from a given guess.
template <bool SEED = true>
ASX_DEV void batch_zone_in(int64_t z, int64_t nz, const double *cons, const double *g,
                           const double *seed, const C2PArgs &a, C2PIn &in) {
  in.r = 0.0;
  in.alp = 1.0;
  in.mask = 1.0;
  in.glo = Sym3{g[z], g[nz + z], g[2 * nz + z], g[3 * nz + z], g[4 * nz + z], g[5 * nz + z]};
  in.cv.dens = cons[z];
  in.cv.mom[0] = cons[nz + z];
  in.cv.mom[1] = cons[2 * nz + z];
  in.cv.mom[2] = cons[3 * nz + z];
  in.cv.tau = cons[4 * nz + z];
  in.cv.dB[0] = cons[5 * nz + z];
  in.cv.dB[1] = cons[6 * nz + z];
  in.cv.dB[2] = cons[7 * nz + z];
  in.cv.DYe = in.cv.dens * a.Ye_atmo;
  in.cv.DEnt = 0.0;
  in.saved_Ye = a.Ye_atmo;
  if (SEED) {
    in.saved_rho = seed[z];
    in.saved_eps = seed[nz + z];
    in.saved_vel[0] = seed[2 * nz + z];
    in.saved_vel[1] = seed[3 * nz + z];
    in.saved_vel[2] = seed[4 * nz + z];
  } else {
    const double nan = __builtin_nan("");
    in.saved_rho = in.saved_eps = nan;
    in.saved_vel[0] = in.saved_vel[1] = in.saved_vel[2] = nan;
  }
  in.temperature = 0.0;
}
// prim[8][nz] = rho, eps, press, vel^i, w_lor, temperature; flag[nz]
ASX_DEV void batch_zone_out(int64_t z, int64_t nz, const C2POut &o, double *prim,
                            double *flag) {
  prim[z] = o.pv.rho;
  prim[nz + z] = o.pv.eps;
  prim[2 * nz + z] = o.pv.press;
  prim[3 * nz + z] = o.pv.vel[0];
  prim[4 * nz + z] = o.pv.vel[1];
  prim[5 * nz + z] = o.pv.vel[2];
  prim[6 * nz + z] = o.pv.w_lor;
  prim[7 * nz + z] = o.pv.temperature;
  flag[z] = o.flag;
}

} // namespace asx
