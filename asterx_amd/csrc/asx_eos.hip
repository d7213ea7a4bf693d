// Independent zones with composition for any 3-parameter EOS (ideal gas or
// table): the EOSX interface on points, prim -> cons, and con2prim in two
// passes (primary solver with a wavefront-ballot queue of the zones it
// failed, then the reference's full chain on the queue).
#include "asx_kernels.hpp"
#include "asx_rand.hpp"

namespace asx {

// EOSX on points: from (rho, T, Ye) -> press, eps, c_s, entropy
// (press_from_valid_rho_temp_ye, eps_from_..., csnd_from_..., entropy_from_...)
template <class E>
__global__ __launch_bounds__(256) void k_eos_rt(int64_t n, const double *rho,
                                                const double *temp, const double *ye,
                                                double *out, E eos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = eos.press_rt(rho[i], temp[i], ye[i]);
  out[n + i] = eos.eps_rt(rho[i], temp[i], ye[i]);
  out[2 * n + i] = eos.csnd_rt(rho[i], temp[i], ye[i]);
  out[3 * n + i] = eos.entropy_rt(rho[i], temp[i], ye[i]);
}
// from (rho, eps, Ye) -> T, press, eps after the EOS's clamping
template <class E>
__global__ __launch_bounds__(256) void k_eos_re(int64_t n, const double *rho,
                                                const double *eps, const double *ye,
                                                double *out, E eos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double e = eps[i];
  out[i] = eos.temp_re(rho[i], e, ye[i]);
  out[n + i] = eos.press_re(rho[i], e, ye[i]);
  out[2 * n + i] = e;
}

// prim[9][nz] = rho, T, Ye, v^i, B^i -> cons[9][nz] (dens, S_i, tau, dB^i,
// DYe); thermo[2][nz] (optional) = eps, press.  Con2PrimFactory cons.hxx:49-91
template <class E>
__global__ __launch_bounds__(256) void k_prim2con_ye(int64_t nz, const double *prim,
                                                     const double *g, double *cons,
                                                     double *thermo, E eos) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  const Sym3 gl{g[z], g[nz + z], g[2 * nz + z], g[3 * nz + z], g[4 * nz + z], g[5 * nz + z]};
  const double rho = prim[z], temp = prim[nz + z], ye = prim[2 * nz + z];
  const double vel[3] = {prim[3 * nz + z], prim[4 * nz + z], prim[5 * nz + z]};
  const double B[3] = {prim[6 * nz + z], prim[7 * nz + z], prim[8 * nz + z]};
  const double eps = eos.eps_rt(rho, temp, ye);
  const double press = eos.press_rt(rho, temp, ye);
  double vl[3];
  lower(gl, vel, vl);
  const double w = 1.0 / sqrt(1.0 - dot3(vl, vel));
  Cons cv;
  cons_from_prim(gl, rho, eps, press, vel, w, B, ye, 0.0, cv);
  cons[z] = cv.dens;
  cons[nz + z] = cv.mom[0];
  cons[2 * nz + z] = cv.mom[1];
  cons[3 * nz + z] = cv.mom[2];
  cons[4 * nz + z] = cv.tau;
  cons[5 * nz + z] = cv.dB[0];
  cons[6 * nz + z] = cv.dB[1];
  cons[7 * nz + z] = cv.dB[2];
  cons[8 * nz + z] = cv.DYe;
  if (thermo) {
    thermo[z] = eps;
    thermo[nz + z] = press;
  }
}

// Zone z: cons[9][nz] (DYe last), seed[7][nz] = rho, eps, v^i, Ye, T
ASX_DEV void zone_in_ye(int64_t z, int64_t nz, const double *cons, const double *g,
                        const double *seed, C2PIn &in) {
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
  in.cv.DYe = cons[8 * nz + z];
  in.cv.DEnt = 0.0;
  in.saved_rho = seed[z];
  in.saved_eps = seed[nz + z];
  in.saved_vel[0] = seed[2 * nz + z];
  in.saved_vel[1] = seed[3 * nz + z];
  in.saved_vel[2] = seed[4 * nz + z];
  in.saved_Ye = seed[5 * nz + z];
  in.temperature = seed[6 * nz + z];
}
// prim[9][nz]: the asx_c2p_batch rows, then Ye
ASX_DEV void zone_out_ye(int64_t z, int64_t nz, const C2POut &o, double *prim, double *flag,
                         int32_t *iters, int spent) {
  batch_zone_out(z, nz, o, prim, flag);
  prim[8 * nz + z] = o.pv.Ye;
  iters[z] = spent + o.iters;
}

// guess (optional, Palenzuela primary): guess[2][nz] = log W, log x from the
// network; the primary brackets x = h W within the relative half-width xrel
// of exp(guess[1]) where the residual changes sign there
template <int PRIME, class E>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(c2p_waves<E>())))
void k_c2p_ye_main(int64_t nz, const double *cons, const double *g, const double *seed,
                   double *prim, double *flag, int32_t *iters, E eos, C2PArgs a, int *queue,
                   unsigned *qcount, const double *guess, double xrel) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool fail = false;
  if (z < nz) {
    C2PIn in;
    zone_in_ye(z, nz, cons, g, seed, in);
    double xg = -1.0;
    if constexpr (PRIME == ASX_C2P_PALENZUELA)
      if (guess) xg = exp(guess[nz + z]);
    C2POut o;
    if (con2prim_zone<PRIME>(eos, a, in, o, xg, xrel)) zone_out_ye(z, nz, o, prim, flag, iters, 0);
    else fail = true;
  }
  queue_push(fail, (int)z, queue, qcount);
}

template <class E>
__global__ __launch_bounds__(256) void k_c2p_ye_rest(int64_t nz, const double *cons,
                                                     const double *g, const double *seed,
                                                     double *prim, double *flag,
                                                     int32_t *iters, E eos, C2PArgs a,
                                                     const int *queue,
                                                     const unsigned *qcount) {
  const unsigned nq = *qcount;
  for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += gridDim.x * blockDim.x) {
    const int64_t z = queue[q];
    C2PIn in;
    zone_in_ye(z, nz, cons, g, seed, in);
    C2POut o;
    con2prim_cell(eos, a, in, o);
    zone_out_ye(z, nz, o, prim, flag, iters, 0);
  }
}

__global__ void k_zero_qcount(unsigned *p) { *p = 0u; }

static dim3 g256(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

template <class E>
hipError_t launch_eos_points_t(int mode, int64_t n, const double *rho, const double *x,
                               const double *ye, double *out, const E &eos, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (mode == 0) hipLaunchKernelGGL(k_eos_rt<E>, g256(n), dim3(256), 0, s, n, rho, x, ye, out, eos);
  else hipLaunchKernelGGL(k_eos_re<E>, g256(n), dim3(256), 0, s, n, rho, x, ye, out, eos);
  return hipGetLastError();
}
hipError_t launch_eos_points(int mode, int64_t n, const double *rho, const double *x,
                             const double *ye, double *out, const IdealGas &eos,
                             hipStream_t s) {
  return launch_eos_points_t(mode, n, rho, x, ye, out, eos, s);
}
hipError_t launch_eos_points(int mode, int64_t n, const double *rho, const double *x,
                             const double *ye, double *out, const TabEOS &eos,
                             hipStream_t s) {
  return launch_eos_points_t(mode, n, rho, x, ye, out, eos, s);
}

template <class E>
static hipError_t prim2con_ye_t(int64_t nz, const double *prim, const double *g,
                                double *cons, double *thermo, const E &eos, hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prim2con_ye<E>, g256(nz), dim3(256), 0, s, nz, prim, g, cons, thermo,
                     eos);
  return hipGetLastError();
}
hipError_t launch_prim2con_ye(int64_t nz, const double *prim, const double *g, double *cons,
                              double *thermo, const IdealGas &eos, hipStream_t s) {
  return prim2con_ye_t(nz, prim, g, cons, thermo, eos, s);
}
hipError_t launch_prim2con_ye(int64_t nz, const double *prim, const double *g, double *cons,
                              double *thermo, const TabEOS &eos, hipStream_t s) {
  return prim2con_ye_t(nz, prim, g, cons, thermo, eos, s);
}

template <int PRIME, class E>
static hipError_t c2p_ye_t(int64_t nz, const double *cons, const double *g,
                           const double *seed, double *prim, double *flag, int32_t *iters,
                           const E &eos, const C2PArgs &a, int *queue, unsigned *qcount,
                           hipStream_t s, const double *guess = nullptr, double xrel = 0.0) {
  hipLaunchKernelGGL(k_zero_qcount, dim3(1), dim3(1), 0, s, qcount);
  hipLaunchKernelGGL((k_c2p_ye_main<PRIME, E>), g256(nz), dim3(256), 0, s, nz, cons, g, seed,
                     prim, flag, iters, eos, a, queue, qcount, guess, xrel);
  const unsigned g2 = (unsigned)((nz + 255) / 256 < 2048 ? (nz + 255) / 256 : 2048);
  hipLaunchKernelGGL(k_c2p_ye_rest<E>, dim3(g2), dim3(256), 0, s, nz, cons, g, seed, prim,
                     flag, iters, eos, a, queue, qcount);
  return hipGetLastError();
}
hipError_t launch_c2p_ye(int64_t nz, const double *cons, const double *g, const double *seed,
                         double *prim, double *flag, int32_t *iters, const IdealGas &eos,
                         const C2PArgs &a, int *queue, unsigned *qcount, hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  if (a.prime == ASX_C2P_NOBLE)
    return c2p_ye_t<ASX_C2P_NOBLE>(nz, cons, g, seed, prim, flag, iters, eos, a, queue, qcount, s);
  if (a.prime == ASX_C2P_ENTROPY)
    return c2p_ye_t<ASX_C2P_ENTROPY>(nz, cons, g, seed, prim, flag, iters, eos, a, queue, qcount,
                                     s);
  return c2p_ye_t<ASX_C2P_PALENZUELA>(nz, cons, g, seed, prim, flag, iters, eos, a, queue,
                                      qcount, s);
}
hipError_t launch_c2p_ye(int64_t nz, const double *cons, const double *g, const double *seed,
                         double *prim, double *flag, int32_t *iters, const TabEOS &eos,
                         const C2PArgs &a, int *queue, unsigned *qcount, hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  return c2p_ye_t<ASX_C2P_PALENZUELA>(nz, cons, g, seed, prim, flag, iters, eos, a, queue,
                                      qcount, s);
}

// con2prim on a table from the table network's guess of x = h W: the
// forward pass writes guess[2][nz], then the two passes above with the
// guided Palenzuela primary
hipError_t launch_c2p_ye_mlp(int64_t nz, const double *cons, const double *g,
                             const double *seed, const float *blob, double xrel, double *prim,
                             double *flag, int32_t *iters, const TabEOS &eos, const C2PArgs &a,
                             double *guess, void *img, int *queue, unsigned *qcount,
                             hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  hipError_t e = launch_mlp_forward<kMlpTab>(nz, cons, g, blob, img, guess, s);
  if (e != hipSuccess) return e;
  return c2p_ye_t<ASX_C2P_PALENZUELA>(nz, cons, g, seed, prim, flag, iters, eos, a, queue,
                                      qcount, s, guess, xrel);
}

// Zones with composition for a 3-parameter EOS, drawn on the device
// (draws as asx_zone_sample): rho log-uniform in [box[0], box[1]], T
// log-uniform in [box[2], box[3]], Ye uniform in [box[4], box[5]], W
// uniform in [1, box[6]] along a random direction, b^2/rho log-uniform in
// [1e-4, 1e2] along a random direction, a symmetric 3-metric within 10% of
// the identity.  prim[9] = rho, T, Ye, v^i, B^i (asx_prim2con_batch_ye);
// sd[7] (optional) = rho, eps, v^i, Ye, T with rho, eps, v^i and T times
// (1 + s u); target[2] (optional) = log W, log(h W) with h from the EOS.
template <class E>
__global__ __launch_bounds__(256) void k_zone_sample_ye(int64_t nz, uint64_t seed, Box7 box,
                                                        double seed_rel, double *prim,
                                                        double *g, double *sd, double *target,
                                                        E eos) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  const double rho = loguni(seed, z, 0, box.v[0], box.v[1]);
  const double temp = loguni(seed, z, 1, box.v[2], box.v[3]);
  const double ye = box.v[4] + (box.v[5] - box.v[4]) * urand(seed, z, 2);
  const double W = 1.0 + (box.v[6] - 1.0) * urand(seed, z, 3);
  double M[6];
  for (int q = 0; q < 6; ++q) M[q] = (2.0 * urand(seed, z, 4 + q) - 1.0) * (0.1 / 3.0);
  const Sym3 gl{1.0 + M[0], M[1], M[2], 1.0 + M[3], M[4], 1.0 + M[5]};
  double nv[3], nb[3];
  normal3(seed, z, 10, nv);
  unit_in(gl, nv);
  normal3(seed, z, 14, nb);
  unit_in(gl, nb);
  const double vabs = sqrt(1.0 - 1.0 / (W * W));
  const double vel[3] = {nv[0] * vabs, nv[1] * vabs, nv[2] * vabs};
  const double b2 = rho * loguni(seed, z, 18, 1e-4, 1e2);
  double vl[3];
  lower(gl, vel, vl);
  const double ndv = dot3(nb, vl);
  const double beta = sqrt(b2 / (1.0 / (W * W) + ndv * ndv));
  prim[z] = rho;
  prim[nz + z] = temp;
  prim[2 * nz + z] = ye;
  for (int i = 0; i < 3; ++i) prim[(3 + i) * nz + z] = vel[i];
  for (int i = 0; i < 3; ++i) prim[(6 + i) * nz + z] = nb[i] * beta;
  g[z] = gl.xx;
  g[nz + z] = gl.xy;
  g[2 * nz + z] = gl.xz;
  g[3 * nz + z] = gl.yy;
  g[4 * nz + z] = gl.yz;
  g[5 * nz + z] = gl.zz;
  const double eps = eos.eps_rt(rho, temp, ye);
  if (sd) {
    sd[z] = rho * (1.0 + seed_rel * (2.0 * urand(seed, z, 19) - 1.0));
    sd[nz + z] = eps * (1.0 + seed_rel * (2.0 * urand(seed, z, 20) - 1.0));
    for (int i = 0; i < 3; ++i)
      sd[(2 + i) * nz + z] = vel[i] * (1.0 + seed_rel * (2.0 * urand(seed, z, 21 + i) - 1.0));
    sd[5 * nz + z] = ye;
    sd[6 * nz + z] = temp * (1.0 + seed_rel * (2.0 * urand(seed, z, 24) - 1.0));
  }
  if (target) {
    const double press = eos.press_rt(rho, temp, ye);
    target[z] = log(W);
    target[nz + z] = log((1.0 + eps + press / rho) * W);
  }
}
template <class E>
static hipError_t zone_sample_ye_t(int64_t nz, uint64_t seed, const Box7 &box, double seed_rel,
                                   double *prim, double *g, double *sd, double *target,
                                   const E &eos, hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zone_sample_ye<E>, g256(nz), dim3(256), 0, s, nz, seed, box, seed_rel,
                     prim, g, sd, target, eos);
  return hipGetLastError();
}
hipError_t launch_zone_sample_ye(int64_t nz, uint64_t seed, const Box7 &box, double seed_rel,
                                 double *prim, double *g, double *sd, double *target,
                                 const IdealGas &eos, hipStream_t s) {
  return zone_sample_ye_t(nz, seed, box, seed_rel, prim, g, sd, target, eos, s);
}
hipError_t launch_zone_sample_ye(int64_t nz, uint64_t seed, const Box7 &box, double seed_rel,
                                 double *prim, double *g, double *sd, double *target,
                                 const TabEOS &eos, hipStream_t s) {
  return zone_sample_ye_t(nz, seed, box, seed_rel, prim, g, sd, target, eos, s);
}

} // namespace asx
