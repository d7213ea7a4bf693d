// Con2prim's zone I/O on a patch, shared by the con2prim kernels
// (asx_c2p.hip) and the batch plan's persistent stage kernel
// (asx_plan.hip): dB from A, the inputs and results of one cell, and the
// zone numbering of a batch of patches.
#pragma once
#include "asx_kernels.hpp"

namespace asx {

// Densitized cell B from the vector potential: curl at the six faces, then
// the face -> cell average (computeBfromA.cxx:32-55 and :71-84).
ASX_DEV void dB_from_A(const KPatch &P, int i, int j, int k, double dB[3]) {
  const double *idx = P.idx;
  const double *Ax = P.g[ASX_AVEC_X], *Ay = P.g[ASX_AVEC_Y], *Az = P.g[ASX_AVEC_Z];
  auto bx = [&](int ii, int jj, int kk) {
    const int oz = off(P, 3, ii, jj, kk), oy = off(P, 5, ii, jj, kk);
    return idx[1] * (Az[oz + P.sj[3]] - Az[oz]) - idx[2] * (Ay[oy + P.sk[5]] - Ay[oy]);
  };
  auto by = [&](int ii, int jj, int kk) {
    const int ox = off(P, 6, ii, jj, kk), oz = off(P, 3, ii, jj, kk);
    return idx[2] * (Ax[ox + P.sk[6]] - Ax[ox]) - idx[0] * (Az[oz + 1] - Az[oz]);
  };
  auto bz = [&](int ii, int jj, int kk) {
    const int oy = off(P, 5, ii, jj, kk), ox = off(P, 6, ii, jj, kk);
    return idx[0] * (Ay[oy + 1] - Ay[oy]) - idx[1] * (Ax[ox + P.sj[6]] - Ax[ox]);
  };
  dB[0] = 0.5 * (bx(i, j, k) + bx(i + 1, j, k));
  dB[1] = 0.5 * (by(i, j, k) + by(i, j + 1, k));
  dB[2] = 0.5 * (bz(i, j, k) + bz(i, j, k + 1));
}

// Inputs of the zone at cell (i, j, k): cell-averaged metric, conserved
// variables (dB from A or from the dB grid functions) and the seed.
template <bool FROM_A>
ASX_DEV void load_zone(const KPatch &P, int i, int j, int k, const C2PArgs &a, C2PIn &in) {
  const int oc = off(P, 0, i, j, k), ov = off(P, 7, i, j, k);
  const int sj = P.sj[7], sk = P.sk[7];
  const double x = P.x0[0] + (i - P.ng + 0.5) * P.dx[0];
  const double y = P.x0[1] + (j - P.ng + 0.5) * P.dx[1];
  const double z = P.x0[2] + (k - P.ng + 0.5) * P.dx[2];
  in.r = sqrt(x * x + y * y + z * z);
  in.alp = avg_v2c(P.g[ASX_ALP], ov, sj, sk);
  in.glo = Sym3{avg_v2c(P.g[ASX_GXX], ov, sj, sk), avg_v2c(P.g[ASX_GXY], ov, sj, sk),
                avg_v2c(P.g[ASX_GXZ], ov, sj, sk), avg_v2c(P.g[ASX_GYY], ov, sj, sk),
                avg_v2c(P.g[ASX_GYZ], ov, sj, sk), avg_v2c(P.g[ASX_GZZ], ov, sj, sk)};
  in.mask = a.use_mask ? avg_v2c(P.g[ASX_MASK_VC], ov, sj, sk) : 1.0;
  in.cv.dens = P.g[ASX_DENS][oc];
  in.cv.mom[0] = P.g[ASX_MOMX][oc];
  in.cv.mom[1] = P.g[ASX_MOMY][oc];
  in.cv.mom[2] = P.g[ASX_MOMZ][oc];
  in.cv.tau = P.g[ASX_TAU][oc];
  in.cv.DYe = P.g[ASX_DYE][oc];
  in.cv.DEnt = P.g[ASX_DENT][oc];
  if (FROM_A) {
    dB_from_A(P, i, j, k, in.cv.dB);
  } else {
    in.cv.dB[0] = P.g[ASX_DBX][oc];
    in.cv.dB[1] = P.g[ASX_DBY][oc];
    in.cv.dB[2] = P.g[ASX_DBZ][oc];
  }
  in.saved_rho = P.g[ASX_SAVED_RHO][oc];
  in.saved_eps = P.g[ASX_SAVED_EPS][oc];
  in.saved_Ye = P.g[ASX_SAVED_YE][oc];
  in.saved_vel[0] = P.g[ASX_SAVED_VELX][oc];
  in.saved_vel[1] = P.g[ASX_SAVED_VELY][oc];
  in.saved_vel[2] = P.g[ASX_SAVED_VELZ][oc];
  in.temperature = P.g[ASX_TEMPERATURE][oc];
}

// Results of the zone (con2prim.cxx:418-453).
ASX_DEV void store_zone(const KPatch &P, int oc, const C2POut &o) {
  const Prim &pv = o.pv;
  const double w = pv.w_lor;
  P.g[ASX_CON2PRIM_FLAG][oc] = o.flag;
  if (P.g[ASX_MASK_CC]) P.g[ASX_MASK_CC][oc] = o.mask_cc;
  P.g[ASX_RHO][oc] = pv.rho;
  P.g[ASX_EPS][oc] = pv.eps;
  P.g[ASX_YE][oc] = pv.Ye;
  P.g[ASX_PRESS][oc] = pv.press;
  P.g[ASX_TEMPERATURE][oc] = pv.temperature;
  P.g[ASX_ENTROPY][oc] = pv.entropy;
  P.g[ASX_VELX][oc] = pv.vel[0];
  P.g[ASX_VELY][oc] = pv.vel[1];
  P.g[ASX_VELZ][oc] = pv.vel[2];
  P.g[ASX_BVECX][oc] = pv.B[0];
  P.g[ASX_BVECY][oc] = pv.B[1];
  P.g[ASX_BVECZ][oc] = pv.B[2];
  P.g[ASX_ZVEC_X][oc] = w * pv.vel[0];
  P.g[ASX_ZVEC_Y][oc] = w * pv.vel[1];
  P.g[ASX_ZVEC_Z][oc] = w * pv.vel[2];
  const double hw = (pv.rho + pv.rho * pv.eps + pv.press) * w * w;
  P.g[ASX_SVEC_X][oc] = hw * pv.vel[0];
  P.g[ASX_SVEC_Y][oc] = hw * pv.vel[1];
  P.g[ASX_SVEC_Z][oc] = hw * pv.vel[2];
  const Cons &cv = o.cv;
  P.g[ASX_DENS][oc] = cv.dens;
  P.g[ASX_MOMX][oc] = cv.mom[0];
  P.g[ASX_MOMY][oc] = cv.mom[1];
  P.g[ASX_MOMZ][oc] = cv.mom[2];
  P.g[ASX_TAU][oc] = cv.tau;
  P.g[ASX_DYE][oc] = cv.DYe;
  P.g[ASX_DENT][oc] = cv.DEnt;
  P.g[ASX_DBX][oc] = cv.dB[0];
  P.g[ASX_DBY][oc] = cv.dB[1];
  P.g[ASX_DBZ][oc] = cv.dB[2];
  P.g[ASX_SAVED_RHO][oc] = pv.rho;
  P.g[ASX_SAVED_VELX][oc] = pv.vel[0];
  P.g[ASX_SAVED_VELY][oc] = pv.vel[1];
  P.g[ASX_SAVED_VELZ][oc] = pv.vel[2];
  P.g[ASX_SAVED_EPS][oc] = pv.eps;
  P.g[ASX_SAVED_YE][oc] = pv.Ye;
}

// Batched forms over many patches.  Zone t of patch p (x-fastest over the
// interior) has the global index zbase[p] + t.
ASX_DEV void interior_ijk(const KPatch &P, int t, int &i, int &j, int &k) {
  i = P.ng + t % P.n[0];
  j = P.ng + (t / P.n[0]) % P.n[1];
  k = P.ng + t / (P.n[0] * P.n[1]);
}
// patch of global zone z: zbase[p] <= z < zbase[p + 1]
ASX_DEV int patch_of(const int *zbase, int np, int z) {
  int lo = 0, hi = np;
  while (hi - lo > 1) {
    const int m = (lo + hi) >> 1;
    if (zbase[m] <= z) lo = m;
    else hi = m;
  }
  return lo;
}

} // namespace asx
