// Independent-zone kernels for gfx950: the sampler of the learned-guess
// workload, batched prim2con, and the learned con2prim initial guess --
// inference, training (forward + backward + Adam) and con2prim from the
// guess with the plain guess as fallback.
//
// The ideal-gas network is 5 -> 64 ReLU -> 64 ReLU -> 2, the table's
// 7 -> 64 tanh -> 64 tanh -> 2 (MlpLayout::Relu).  Inference runs all three
// layers on the matrix cores (v_mfma_f32_16x16x16_bf16: bf16 operands, fp32
// accumulation; net_wave), inside the con2prim kernel or as a kernel of its
// own.  Training runs the 64x64 layer and both 64x64 products of its
// backward pass on the matrix cores and the first and last layers (5 and 2
// wide) on the vector ALU with bf16-rounded operands and fp32
// accumulation.  Weights are kept in fp32 (the blob) and rounded to bf16
// where they enter a product.
#include "asx_kernels.hpp"
#include "asx_rand.hpp"

namespace asx {

// One zone of the learned-guess workload: rho log-uniform in [1e-10, 1e-2],
// p/rho log-uniform in [1e-6, 1e2], W uniform in [1, 10] along a random
// direction, b^2/rho log-uniform in [1e-4, 1e2] along a random direction, a
// symmetric 3-metric within 10% of the identity.  prim[8] = rho, eps, v^i,
// B^i; sd[5] (optional) = the true rho, eps, v^i each times (1 + s u), u
// uniform in [-1, 1]; target[2] (optional) = log W, log(Z / D) with
// Z = rho h W^2 and D = rho W.
__global__ __launch_bounds__(256) void k_zone_sample(int64_t nz, uint64_t seed, double gamma,
                                                     double seed_rel, double *prim, double *g,
                                                     double *sd, double *target) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  const double rho = loguni(seed, z, 0, 1e-10, 1e-2);
  const double press = rho * loguni(seed, z, 1, 1e-6, 1e2);
  const double eps = press / ((gamma - 1.0) * rho);
  const double W = 1.0 + 9.0 * urand(seed, z, 2);
  double M[6];
  for (int q = 0; q < 6; ++q) M[q] = (2.0 * urand(seed, z, 3 + q) - 1.0) * (0.1 / 3.0);
  const Sym3 gl{1.0 + M[0], M[1], M[2], 1.0 + M[3], M[4], 1.0 + M[5]};
  double nv[3], nb[3];
  normal3(seed, z, 9, nv);
  unit_in(gl, nv);
  normal3(seed, z, 13, nb);
  unit_in(gl, nb);
  const double vabs = sqrt(1.0 - 1.0 / (W * W));
  const double vel[3] = {nv[0] * vabs, nv[1] * vabs, nv[2] * vabs};
  const double b2 = rho * loguni(seed, z, 17, 1e-4, 1e2);
  double vl[3];
  lower(gl, vel, vl);
  const double ndv = dot3(nb, vl);
  const double beta = sqrt(b2 / (1.0 / (W * W) + ndv * ndv));
  prim[z] = rho;
  prim[nz + z] = eps;
  for (int i = 0; i < 3; ++i) prim[(2 + i) * nz + z] = vel[i];
  for (int i = 0; i < 3; ++i) prim[(5 + i) * nz + z] = nb[i] * beta;
  g[z] = gl.xx;
  g[nz + z] = gl.xy;
  g[2 * nz + z] = gl.xz;
  g[3 * nz + z] = gl.yy;
  g[4 * nz + z] = gl.yz;
  g[5 * nz + z] = gl.zz;
  if (sd) {
    sd[z] = rho * (1.0 + seed_rel * (2.0 * urand(seed, z, 18) - 1.0));
    sd[nz + z] = eps * (1.0 + seed_rel * (2.0 * urand(seed, z, 19) - 1.0));
    for (int i = 0; i < 3; ++i)
      sd[(2 + i) * nz + z] = vel[i] * (1.0 + seed_rel * (2.0 * urand(seed, z, 20 + i) - 1.0));
  }
  if (target) {
    target[z] = log(W);
    target[nz + z] = log((1.0 + eps + press / rho) * W);
  }
}

// Batched Valencia prim -> cons (Con2PrimFactory/src/cons.hxx:49-91 with W
// from v); cons[8] = dens, S_i, tau, dB^i (densitized).  cold (optional)
// receives the cold-start seed rho = D, eps = tau/D, v = 0.
__global__ __launch_bounds__(256) void k_prim2con_batch(int64_t nz, const double *prim,
                                                        const double *g, double *cons,
                                                        double *cold, IdealGas eos,
                                                        double Ye) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  const Sym3 gl{g[z], g[nz + z], g[2 * nz + z], g[3 * nz + z], g[4 * nz + z], g[5 * nz + z]};
  const double rho = prim[z], eps = prim[nz + z];
  const double vel[3] = {prim[2 * nz + z], prim[3 * nz + z], prim[4 * nz + z]};
  const double Bv[3] = {prim[5 * nz + z], prim[6 * nz + z], prim[7 * nz + z]};
  double vl[3];
  lower(gl, vel, vl);
  const double w = 1.0 / sqrt(1.0 - dot3(vl, vel));
  Cons cv;
  cons_from_prim(gl, rho, eps, eos.press_re(rho, eps), vel, w, Bv, Ye, 0.0, cv);
  cons[z] = cv.dens;
  for (int i = 0; i < 3; ++i) cons[(1 + i) * nz + z] = cv.mom[i];
  cons[4 * nz + z] = cv.tau;
  for (int i = 0; i < 3; ++i) cons[(5 + i) * nz + z] = cv.dB[i];
  if (cold) {
    cold[z] = cv.dens / sqrt(det3(gl));
    cold[nz + z] = cv.tau / cv.dens;
    cold[2 * nz + z] = 0.0;
    cold[3 * nz + z] = 0.0;
    cold[4 * nz + z] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Network
// ---------------------------------------------------------------------------
// Inputs from the densitized conserved variables: log D, log S^2, log tau,
// log B^2 (undensitized, S^2 = gamma^ij S_i S_j, B^2 = gamma_ij B^i B^j) in
// units of E = D + tau (so scale-free: the ideal-gas con2prim is invariant
// under D, S, tau -> l D, l S, l tau and B -> sqrt(l) B), and the cosine of
// S and B.  tau is floored at 1e-15 D and the ratios S^2/E^2, B^2/E at
// 1e-30; the ratios are formed in double precision and their logarithms
// taken in single precision (the network's own precision).
ASX_DEV void mlp_features(const double c[8], const Sym3 &gl, float x[5]) {
  const double detg = det3(gl), sq = sqrt(detg);
  const Sym3 gu = inv3(gl, detg);
  const double is = 1.0 / sq;
  const double D = c[0] * is, tau = c[4] * is;
  const double S[3] = {c[1] * is, c[2] * is, c[3] * is};
  const double B[3] = {c[5] * is, c[6] * is, c[7] * is};
  double Su[3], Bl[3];
  lower(gu, S, Su);
  lower(gl, B, Bl);
  const double S2 = dot3(S, Su), B2 = dot3(B, Bl), SB = dot3(S, B);
  const double taup = tau > 1e-15 * D ? tau : 1e-15 * D;
  const double iE = 1.0 / (D + taup);
  const double rS = S2 * iE * iE, rB = B2 * iE;
  x[0] = __logf((float)(D * iE));
  x[1] = __logf((float)(rS > 1e-30 ? rS : 1e-30));
  x[2] = __logf((float)(taup * iE));
  x[3] = __logf((float)(rB > 1e-30 ? rB : 1e-30));
  const double den = sqrt(S2 * B2);
  x[4] = den > 0.0 ? (float)(SB / den) : 0.0f;
}
// The two further inputs on a table, which sets a density scale of its
// own: log D (undensitized, floored at 1e-37) and Ye = DYe / D (c[8]).
ASX_DEV void mlp_features_tab(const double c[9], const Sym3 &gl, float x[2]) {
  const double D = c[0] / sqrt(det3(gl));
  x[0] = __logf((float)(D > 1e-37 ? D : 1e-37));
  x[1] = (float)(c[8] / c[0]);
}

// tanh from the hardware exp2 and reciprocal, 1 - 2 / (2^(2 x / ln 2) + 1):
// within about 1e-7 of tanh, saturating to +-1 (the network's activation in
// both inference and training)
ASX_DEV float fast_tanh(float x) {
  const float t = __builtin_amdgcn_exp2f(2.8853900817779268f * x);
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(t + 1.0f);
}

// The hidden activation of the network of NIN inputs and its derivative
// expressed through the activation's value h.
template <int NIN> ASX_DEV float act(float z) {
  if constexpr (MlpLayout<NIN>::Relu) return z > 0.0f ? z : 0.0f;
  else return fast_tanh(z);
}
template <int NIN> ASX_DEV float dact(float h) {
  if constexpr (MlpLayout<NIN>::Relu) return h > 0.0f ? 1.0f : 0.0f;
  else return 1.0f - h * h;
}

// bf16 round-to-nearest-even and back
ASX_DEV uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
ASX_DEV float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
ASX_DEV float bfr(float f) { return bf2f(f2bf(f)); }

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// C[16x16] += A[16x16] B[16x16] on one wavefront.  Fragments, for lane l:
//   A: A[l%16][4(l/16) .. 4(l/16)+3]      B: B[4(l/16) .. 4(l/16)+3][l%16]
//   C: C[4(l/16) .. 4(l/16)+3][l%16]
ASX_DEV f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// 64x64 bf16 tiles in LDS, row-major with a padded row of kTS elements
// (136 bytes: fragment reads stay 8-byte aligned and spread over banks)
constexpr int kTS = 68;
ASX_DEV bf16x4 frag(const uint16_t *t, int row, int col) {
  return *reinterpret_cast<const bf16x4 *>(t + row * kTS + col);
}

// Raw inputs of the network of NIN inputs from cons (MlpLayout<NIN>::NCons
// values) of one zone.
template <int NIN> ASX_DEV void net_features(const double *c, const Sym3 &gl, float *x) {
  mlp_features(c, gl, x);
  if constexpr (NIN == kMlpTab) mlp_features_tab(c, gl, x + 5);
}

// Normalised inputs of zone z (zeros for lanes past the batch).
template <int NIN>
ASX_DEV void zone_inputs(bool live, int64_t z, int64_t nz, const double *cons,
                         const double *g, const float *blob, float xn[NIN]) {
  using L = MlpLayout<NIN>;
  float x[NIN] = {};
  if (live) {
    double c[L::NCons];
    for (int q = 0; q < L::NCons; ++q) c[q] = cons[q * nz + z];
    const Sym3 gl{g[z], g[nz + z], g[2 * nz + z], g[3 * nz + z], g[4 * nz + z], g[5 * nz + z]};
    net_features<NIN>(c, gl, x);
  }
  for (int i = 0; i < NIN; ++i)
    xn[i] = live ? (x[i] - blob[L::InMean + i]) * blob[L::InIstd + i] : 0.0f;
}

// Stage W2 (bf16, [j][k]) of the blob into LDS; every work-item of the block.
template <int NIN> ASX_DEV void stage_w2(const float *blob, uint16_t *w2s, uint16_t *w2t) {
  using L = MlpLayout<NIN>;
  for (int q = threadIdx.x; q < L::H * L::H; q += blockDim.x) {
    const int j = q / L::H, k = q % L::H;
    const uint16_t b = f2bf(blob[L::W2 + q]);
    w2s[j * kTS + k] = b;
    if (w2t) w2t[k * kTS + j] = b;
  }
}

// Layer 1 for this lane's zone: h1 = tanh(W1 x + b1), written as bf16 to
// the zone's row of h1 (and column of h1t when given).
template <int NIN>
ASX_DEV void layer1(const float *blob, const float xb[NIN], uint16_t *h1, uint16_t *h1t, int l) {
  using L = MlpLayout<NIN>;
  for (int j = 0; j < L::H; ++j) {
    float zj = blob[L::B1 + j];
    for (int i = 0; i < L::In; ++i) zj += bfr(blob[L::W1 + j * L::In + i]) * xb[i];
    const uint16_t hb = f2bf(act<NIN>(zj));
    h1[l * kTS + j] = hb;
    if (h1t) h1t[j * kTS + l] = hb;
  }
}

// Layer 2 products Z2[z][j] = sum_k H1[z][k] W2[j][k] for the wavefront's
// 64 zones: 4 zone tiles x 4 output tiles x 4 k-steps of 16.
ASX_DEV void layer2(const uint16_t *h1, const uint16_t *w2s, f32x4 acc[4][4], int l) {
  constexpr int kH = 64;
  const int r16 = l & 15, k4 = 4 * (l >> 4);
  for (int rt = 0; rt < 4; ++rt)
    for (int ob = 0; ob < 4; ++ob) acc[rt][ob] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    bf16x4 b[4];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) b[ob] = frag(w2s, ob * 16 + r16, kb * 16 + k4);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const bf16x4 a = frag(h1, rt * 16 + r16, kb * 16 + k4);
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) acc[rt][ob] = mfma16(a, b[ob], acc[rt][ob]);
    }
  }
}

template <int NIN>
__global__ __launch_bounds__(256) void k_mlp_features(int64_t nz, const double *cons,
                                                      const double *g, float *xo) {
  using L = MlpLayout<NIN>;
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  double c[L::NCons];
  float x[NIN];
  for (int q = 0; q < L::NCons; ++q) c[q] = cons[q * nz + z];
  const Sym3 gl{g[z], g[nz + z], g[2 * nz + z], g[3 * nz + z], g[4 * nz + z], g[5 * nz + z]};
  net_features<NIN>(c, gl, x);
  for (int i = 0; i < NIN; ++i) xo[i * nz + z] = x[i];
}

// ---------------------------------------------------------------------------
// Training: gradient of the mean squared error of the normalised outputs
// over the batch, one wavefront (64 zones) per workgroup; every workgroup
// writes its partial gradient (and loss) to partial[blockIdx.x].  bf16
// rounding is taken as the identity in the derivatives; the products that
// run on the matrix cores use the bf16-rounded operands, as in the forward
// pass.
// ---------------------------------------------------------------------------
template <int NIN>
__global__ __launch_bounds__(64) void k_mlp_grad(int64_t nz, const double *cons, const double *g,
                                                 const double *target, const float *blob,
                                                 float *partial) {
  using L = MlpLayout<NIN>;
  __shared__ uint16_t w2s[L::H * kTS], w2t[L::H * kTS];
  __shared__ uint16_t h1[L::H * kTS], h1t[L::H * kTS];
  __shared__ uint16_t a2[L::H * kTS], dz2t[L::H * kTS];
  __shared__ float xs[64][L::In];
  __shared__ float dys[64][L::Out];
  __shared__ float red[4][L::H][L::In + 1];
  __shared__ float lsum[64];
  const int l = threadIdx.x;
  const int r16 = l & 15, grp = l >> 4, k4 = 4 * grp;
  stage_w2<NIN>(blob, w2s, w2t);
  const int64_t z = (int64_t)blockIdx.x * 64 + l;
  const bool live = z < nz;
  float xn[NIN];
  zone_inputs<NIN>(live, z, nz, cons, g, blob, xn);
  float xb[NIN];
  for (int i = 0; i < L::In; ++i) {
    xb[i] = bfr(xn[i]);
    xs[l][i] = xb[i];
  }
  float tn[2] = {0.f, 0.f};
  if (live)
    for (int o = 0; o < L::Out; ++o)
      tn[o] = ((float)target[o * nz + z] - blob[L::OutMean + o]) / blob[L::OutStd + o];
  __syncthreads();
  layer1<NIN>(blob, xb, h1, h1t, l);
  __syncthreads();
  f32x4 acc[4][4];
  layer2(h1, w2s, acc, l);
  // h2 = tanh(z2 + b2), in place of acc, and as bf16 into a2
  for (int rt = 0; rt < 4; ++rt)
    for (int ob = 0; ob < 4; ++ob)
      for (int r = 0; r < 4; ++r) {
        const int zr = rt * 16 + k4 + r, o = ob * 16 + r16;
        const float h = act<NIN>(acc[rt][ob][r] + blob[L::B2 + o]);
        acc[rt][ob][r] = h;
        a2[zr * kTS + o] = f2bf(h);
      }
  __syncthreads();
  // layer 3 and the loss
  const float inv_n = 1.0f / (float)nz;
  float loss = 0.f;
  for (int o = 0; o < L::Out; ++o) {
    float s = blob[L::B3 + o];
    for (int k = 0; k < L::H; ++k) s += bfr(blob[L::W3 + o * L::H + k]) * bf2f(a2[l * kTS + k]);
    const float e = live ? s - tn[o] : 0.f;
    loss += e * e * inv_n;
    dys[l][o] = 2.f * e * inv_n;
  }
  lsum[l] = loss;
  __syncthreads();
  float *out = partial + (int64_t)blockIdx.x * L::Partial;
  // dW3[o][k] (lane = k), db3, loss
  for (int o = 0; o < L::Out; ++o) {
    float s = 0.f;
    for (int q = 0; q < 64; ++q) s += dys[q][o] * bf2f(a2[q * kTS + l]);
    out[L::W3 + o * L::H + l] = s;
  }
  if (l < L::Out) {
    float s = 0.f;
    for (int q = 0; q < 64; ++q) s += dys[q][l];
    out[L::B3 + l] = s;
  }
  if (l == 0) {
    float s = 0.f;
    for (int q = 0; q < 64; ++q) s += lsum[q];
    out[L::NParam] = s;
  }
  // dz2 = (W3^T dy) act'(h2), in place of acc; db2 partials per lane group
  float db2p[4] = {0.f, 0.f, 0.f, 0.f};
  for (int rt = 0; rt < 4; ++rt)
    for (int ob = 0; ob < 4; ++ob)
      for (int r = 0; r < 4; ++r) {
        const int zr = rt * 16 + k4 + r, j = ob * 16 + r16;
        float dh = 0.f;
        for (int o = 0; o < L::Out; ++o) dh += dys[zr][o] * bfr(blob[L::W3 + o * L::H + j]);
        const float h = acc[rt][ob][r];
        const float dz = dh * dact<NIN>(h);
        acc[rt][ob][r] = dz;
        db2p[ob] += dz;
      }
  __syncthreads(); // a2 (h2) no longer read: reuse it for dz2
  for (int ob = 0; ob < 4; ++ob) red[grp][ob * 16 + r16][0] = db2p[ob];
  for (int rt = 0; rt < 4; ++rt)
    for (int ob = 0; ob < 4; ++ob)
      for (int r = 0; r < 4; ++r) {
        const int zr = rt * 16 + k4 + r, j = ob * 16 + r16;
        const uint16_t b = f2bf(acc[rt][ob][r]);
        a2[zr * kTS + j] = b;
        dz2t[j * kTS + zr] = b;
      }
  __syncthreads();
  out[L::B2 + l] = red[0][l][0] + red[1][l][0] + red[2][l][0] + red[3][l][0];
  // dW2[j][k] = sum_z dz2[z][j] H1[z][k]: A = dz2^T (rows j), B = H1 (cols k)
  {
    f32x4 gw[4][4];
    for (int jt = 0; jt < 4; ++jt)
      for (int kt = 0; kt < 4; ++kt) gw[jt][kt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kz = 0; kz < 4; ++kz) {
      bf16x4 b[4];
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) b[kt] = frag(h1t, kt * 16 + r16, kz * 16 + k4);
#pragma unroll
      for (int jt = 0; jt < 4; ++jt) {
        const bf16x4 a = frag(dz2t, jt * 16 + r16, kz * 16 + k4);
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) gw[jt][kt] = mfma16(a, b[kt], gw[jt][kt]);
      }
    }
    for (int jt = 0; jt < 4; ++jt)
      for (int kt = 0; kt < 4; ++kt)
        for (int r = 0; r < 4; ++r)
          out[L::W2 + (jt * 16 + k4 + r) * L::H + kt * 16 + r16] = gw[jt][kt][r];
  }
  // dH1[z][k] = sum_j dz2[z][j] W2[j][k]: A = dz2 (rows z), B = W2 (cols k)
  f32x4 dh1[4][4];
  for (int rt = 0; rt < 4; ++rt)
    for (int kt = 0; kt < 4; ++kt) dh1[rt][kt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kj = 0; kj < 4; ++kj) {
    bf16x4 b[4];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) b[kt] = frag(w2t, kt * 16 + r16, kj * 16 + k4);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const bf16x4 a = frag(a2, rt * 16 + r16, kj * 16 + k4);
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) dh1[rt][kt] = mfma16(a, b[kt], dh1[rt][kt]);
    }
  }
  // dz1 = dH1 act'(h1) with h1 recomputed in fp32; dW1 and db1 partials
  // per (lane group, column), reduced over the four groups
  __syncthreads();
  for (int kt = 0; kt < 4; ++kt) {
    const int k = kt * 16 + r16;
    float sd = 0.f, sw[L::In] = {};
    float wk[L::In];
    for (int i = 0; i < L::In; ++i) wk[i] = bfr(blob[L::W1 + k * L::In + i]);
    for (int rt = 0; rt < 4; ++rt)
      for (int r = 0; r < 4; ++r) {
        const int zr = rt * 16 + k4 + r;
        float zk = blob[L::B1 + k];
        for (int i = 0; i < L::In; ++i) zk += wk[i] * xs[zr][i];
        const float h = act<NIN>(zk);
        const float dz = dh1[rt][kt][r] * dact<NIN>(h);
        sd += dz;
        for (int i = 0; i < L::In; ++i) sw[i] += dz * xs[zr][i];
      }
    for (int i = 0; i < L::In; ++i) red[grp][k][i] = sw[i];
    red[grp][k][L::In] = sd;
  }
  __syncthreads();
  for (int i = 0; i < L::In; ++i)
    out[L::W1 + l * L::In + i] = red[0][l][i] + red[1][l][i] + red[2][l][i] + red[3][l][i];
  out[L::B1 + l] = red[0][l][L::In] + red[1][l][L::In] + red[2][l][L::In] + red[3][l][L::In];
}

// Sum the per-workgroup partials in a fixed order (deterministic) and take
// one Adam step (Kingma & Ba 2015) on the fp32 parameters.
template <int NIN>
__global__ __launch_bounds__(256) void k_mlp_adam(int nw, const float *partial, float *blob,
                                                  float *m, float *v, float lr, float b1,
                                                  float b2, float eps, float c1, float c2,
                                                  float *loss, float *grad) {
  using L = MlpLayout<NIN>;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > L::NParam) return;
  float s = 0.f;
  for (int w = 0; w < nw; ++w) s += partial[(int64_t)w * L::Partial + p];
  if (p == L::NParam) {
    if (loss) *loss = s;
    return;
  }
  if (grad) grad[p] = s;
  if (lr == 0.f) return;
  const float mp = b1 * m[p] + (1.f - b1) * s;
  const float vp = b2 * v[p] + (1.f - b2) * s * s;
  m[p] = mp;
  v[p] = vp;
  blob[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// ---------------------------------------------------------------------------
// Con2prim from the network's guess, after the forward pass has written
// guess[2][nz] (log W, log(Z/D)), in two passes:
//   1. Noble from (Z, v^2) of the guess for at most kguess iterations;
//      converged zones are stored, the others queued by a wavefront ballot
//      (the iterations they spent are kept in iters);
//   2. the queued zones run the full chain from the plain seed, as
//      asx_c2p_batch would (fallback[z] = 1), iters counting both.
// The primary-only pass carries neither the backup solver nor the network,
// so it runs at the occupancy of the plain solver.
// ---------------------------------------------------------------------------
ASX_DEV void guess_zv_of(double lw, double lzd, const C2PIn &in, double &Zg, double &vsq) {
  const double W = smax(1.0, exp(lw));
  vsq = 1.0 - 1.0 / (W * W);
  const double D = in.cv.dens / sqrt(det3(in.glo));
  Zg = D * exp(lzd);
  if (!(Zg > 0.0) || !isfinite(Zg) || !isfinite(vsq)) Zg = -1.0; // plain seed instead
}
ASX_DEV void guess_zv(const double *guess, int64_t z, int64_t nz, const C2PIn &in, double &Zg,
                      double &vsq) {
  guess_zv_of(guess[z], guess[nz + z], in, Zg, vsq);
}

__global__ __launch_bounds__(256) void k_c2p_guess_main(int64_t nz, const double *cons,
                                                        const double *g, const double *seed,
                                                        const double *guess, int kguess,
                                                        double *prim, double *flag,
                                                        int32_t *iters, int32_t *fallback,
                                                        IdealGas eos, C2PArgs a, int *queue,
                                                        unsigned *qcount) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool fail = false;
  if (z < nz) {
    C2PIn in;
    batch_zone_in(z, nz, cons, g, seed, a, in);
    double Zg, vsq;
    guess_zv(guess, z, nz, in, Zg, vsq);
    C2PArgs ak = a;
    ak.max_iter = kguess;
    C2POut o;
    if (con2prim_zone<ASX_C2P_NOBLE>(eos, ak, in, o, Zg, vsq)) {
      batch_zone_out(z, nz, o, prim, flag);
      if (fallback) fallback[z] = 0;
    } else {
      fail = true;
    }
    iters[z] = o.iters;
  }
  queue_push(fail, (int)z, queue, qcount);
}

__global__ __launch_bounds__(256) void k_c2p_guess_rest(int64_t nz, const double *cons,
                                                        const double *g, const double *seed,
                                                        double *prim, double *flag,
                                                        int32_t *iters, int32_t *fallback,
                                                        IdealGas eos, C2PArgs a,
                                                        const int *queue,
                                                        const unsigned *qcount) {
  const unsigned nq = *qcount;
  for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += gridDim.x * blockDim.x) {
    const int64_t z = queue[q];
    C2PIn in;
    batch_zone_in(z, nz, cons, g, seed, a, in);
    C2POut o;
    con2prim_cell(eos, a, in, o);
    batch_zone_out(z, nz, o, prim, flag);
    iters[z] += o.iters;
    if (fallback) fallback[z] = 1;
  }
}

__global__ void k_zero_count(unsigned *p) { *p = 0u; }

// ---------------------------------------------------------------------------
// The network on the matrix cores, all three layers, for the 64 zones of a
// wavefront (lane = zone).  Every layer is evaluated as Out^T = W In^T: in
// the 16x16x16 product lane l holds rows 4(l/16)..+3 of column l%16 of both
// the accumulator and the B operand, so one layer's activations are, lane
// for lane, the next layer's B operand and never leave registers.  Layer 1
// pads its K (the inputs) to 16, layer 3 its M (the outputs) to 16; biases
// initialise the fp32 accumulators.  Only the inputs (zone rows to B
// columns) and the outputs (back to the zone's lane) pass through the
// wavefront's own LDS rows.
// ---------------------------------------------------------------------------
// Orders a wavefront's LDS writes before its later LDS reads of other
// lanes' data (the rows used here are private to one wavefront, so no
// workgroup barrier is needed).
ASX_DEV void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Two floats to bf16 with the hardware conversion (v_cvt_pk_bf16_f32,
// round to nearest even: the rounding of f2bf for every non-NaN value).
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
ASX_DEV uint32_t pack_bf16(float a, float b) {
  const bf16x2_t h = __builtin_convertvector(f32x2_t{a, b}, bf16x2_t);
  return __builtin_bit_cast(uint32_t, h);
}
template <int NIN> ASX_DEV bf16x4 act_bf16x4(const f32x4 &v) {
  const uint32_t lo = pack_bf16(act<NIN>(v[0]), act<NIN>(v[1]));
  const uint32_t hi = pack_bf16(act<NIN>(v[2]), act<NIN>(v[3]));
  return bf16x4{(short)(lo & 0xffffu), (short)(lo >> 16), (short)(hi & 0xffffu),
                (short)(hi >> 16)};
}

// The network's inference weights as the matrix cores read them: W2 as bf16
// rows padded to kTS, W1 as bf16 rows of 8 inputs (zero padded), W3 as two
// bf16 rows, the biases in fp32.  Built once per call from the fp32 blob
// (k_net_image) and copied into LDS by every workgroup with 16-byte loads.
struct NetImage {
  uint16_t w2[64 * kTS];
  uint16_t w1[64 * 8];
  uint16_t w3[2 * 64];
  float b1[64], b2[64], b3[4];
};
static_assert(sizeof(NetImage) % 16 == 0, "NetImage is copied as uint4");
static_assert(sizeof(NetImage) == kNetImageBytes, "scratch size of the image");
template <int NIN> __global__ __launch_bounds__(256) void k_net_image(const float *blob,
                                                                      NetImage *img) {
  using L = MlpLayout<NIN>;
  static_assert(L::H == 64 && L::Out == 2, "layer shapes");
  const int t = threadIdx.x;
  for (int q = t; q < 64 * kTS; q += blockDim.x) {
    const int j = q / kTS, k = q % kTS;
    img->w2[q] = k < 64 ? f2bf(blob[L::W2 + j * 64 + k]) : (uint16_t)0;
  }
  for (int q = t; q < 64 * 8; q += blockDim.x) {
    const int j = q / 8, i = q % 8;
    img->w1[q] = i < NIN ? f2bf(blob[L::W1 + j * L::In + i]) : (uint16_t)0;
  }
  for (int q = t; q < 2 * 64; q += blockDim.x) img->w3[q] = f2bf(blob[L::W3 + q]);
  if (t < 64) {
    img->b1[t] = blob[L::B1 + t];
    img->b2[t] = blob[L::B2 + t];
  }
  if (t < 4) img->b3[t] = t < L::Out ? blob[L::B3 + t] : 0.0f;
}
// Every work-item of the workgroup copies its share; the caller then
// synchronises the workgroup.
ASX_DEV void load_image(const NetImage *img, NetImage &S) {
  const uint4 *src = reinterpret_cast<const uint4 *>(img);
  uint4 *dst = reinterpret_cast<uint4 *>(&S);
  for (int q = threadIdx.x; q < (int)(sizeof(NetImage) / 16); q += blockDim.x) dst[q] = src[q];
}

// Normalised outputs y[2] of this lane's zone from its normalised inputs
// xn.  S: the weights in LDS (load_image); xrow: the wavefront's 64 input
// rows; yrow: its 64 output pairs.  Every lane of the wavefront calls it.  Layer 3 consumes each 16-row block of layer 2's
// activations as it is produced, so only layer 1's activations (32 VGPRs)
// are held whole.
template <int NIN>
ASX_DEV void net_wave(const NetImage &S, uint4 *xrow, float (*yrow)[2], const float xn[NIN],
                      float y[2]) {
  using L = MlpLayout<NIN>;
  static_assert(NIN <= 8 && L::Out <= 4, "inputs padded to 8 of K = 16; outputs in rows 0..3");
  const int l = threadIdx.x & 63, r16 = l & 15, q4 = l >> 4, k4 = 4 * q4;
  const bf16x4 zero4{0, 0, 0, 0};
  uint32_t pk[4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
    pk[p] = pack_bf16(2 * p < NIN ? xn[2 * p] : 0.0f, 2 * p + 1 < NIN ? xn[2 * p + 1] : 0.0f);
  xrow[l] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
  wave_lds_sync();
  // layer 1: A = rows of W1 (inputs k4..k4+3, zero beyond 8), B = X^T,
  // rows k4..k4+3 (inputs) of column zt*16 + l%16 (zone)
  bf16x4 hb[4][4];
  {
    bf16x4 xb[4];
#pragma unroll
    for (int zt = 0; zt < 4; ++zt) {
      const uint4 v = xrow[zt * 16 + r16];
      const uint32_t u0 = q4 == 0 ? v.x : q4 == 1 ? v.z : 0u;
      const uint32_t u1 = q4 == 0 ? v.y : q4 == 1 ? v.w : 0u;
      xb[zt] = bf16x4{(short)(u0 & 0xffffu), (short)(u0 >> 16), (short)(u1 & 0xffffu),
                      (short)(u1 >> 16)};
    }
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) {
      const bf16x4 aw = q4 < 2 ? *reinterpret_cast<const bf16x4 *>(S.w1 + (jt * 16 + r16) * 8 + k4)
                               : zero4;
      const f32x4 b = *reinterpret_cast<const f32x4 *>(S.b1 + jt * 16 + k4);
#pragma unroll
      for (int zt = 0; zt < 4; ++zt) hb[jt][zt] = act_bf16x4<NIN>(mfma16(aw, xb[zt], b));
    }
  }
  // layer 2 (A = rows of W2 from LDS, B = layer 1's activations), each
  // 16-row block fed at once to layer 3 (A = W3 rows o = l%16 < 2, B = that
  // block); layer 3's rows o = k4 + r: outputs 0 and 1 sit in lanes 0..15
  f32x4 yv[4];
  {
    const f32x4 b3 = q4 == 0 ? *reinterpret_cast<const f32x4 *>(S.b3) : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int zt = 0; zt < 4; ++zt) yv[zt] = b3;
  }
#pragma unroll
  for (int jt = 0; jt < 4; ++jt) {
    const f32x4 b = *reinterpret_cast<const f32x4 *>(S.b2 + jt * 16 + k4);
    f32x4 acc[4] = {b, b, b, b};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      const bf16x4 aw = frag(S.w2, jt * 16 + r16, kb * 16 + k4);
#pragma unroll
      for (int zt = 0; zt < 4; ++zt) acc[zt] = mfma16(aw, hb[kb][zt], acc[zt]);
    }
    const bf16x4 a3 = r16 < L::Out
                          ? *reinterpret_cast<const bf16x4 *>(S.w3 + r16 * 64 + jt * 16 + k4)
                          : zero4;
#pragma unroll
    for (int zt = 0; zt < 4; ++zt) yv[zt] = mfma16(a3, act_bf16x4<NIN>(acc[zt]), yv[zt]);
  }
#pragma unroll
  for (int zt = 0; zt < 4; ++zt)
    if (q4 == 0) {
      yrow[zt * 16 + r16][0] = yv[zt][0];
      yrow[zt * 16 + r16][1] = yv[zt][1];
    }
  wave_lds_sync();
  y[0] = yrow[l][0];
  y[1] = yrow[l][1];
}

// Normalised inputs of zone z from its conserved variables and metric.
template <int NIN>
ASX_DEV void zone_inputs_of(bool live, const double *c, const Sym3 &gl, const float *blob,
                            float xn[NIN]) {
  using L = MlpLayout<NIN>;
  float x[NIN] = {};
  if (live) net_features<NIN>(c, gl, x);
#pragma unroll
  for (int i = 0; i < NIN; ++i)
    xn[i] = live ? (x[i] - blob[L::InMean + i]) * blob[L::InIstd + i] : 0.0f;
}

// ---------------------------------------------------------------------------
// Con2prim from the network's guess in one kernel: the wavefront evaluates
// the network for its 64 zones (net_wave), then each lane runs Noble from
// its zone's guess for at most kguess iterations, as k_c2p_guess_main,
// queueing failures for k_c2p_guess_rest.  The guided solve reads no seed
// (con2prim_zone's SEEDLESS form).  One chunk of 256 zones per workgroup: a
// loop over chunks made the compiler keep ~100 more VGPRs live (occupancy
// 1); without it the kernel runs at 4 waves per SIMD (120 VGPRs, no spills).
// ---------------------------------------------------------------------------
template <int NIN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_c2p_guess_fused(
    int64_t nz, const double *cons, const double *g, const double *seed, const float *blob,
    const NetImage *img, int kguess, double *prim, double *flag, int32_t *iters, int32_t *fallback, IdealGas eos,
    C2PArgs a, int *queue, unsigned *qcount) {
  using L = MlpLayout<NIN>;
  __shared__ NetImage S;
  __shared__ uint4 xs[4][64];
  __shared__ float ys[4][64][2];
  load_image(img, S);
  __syncthreads();
  const int wv = threadIdx.x >> 6;
  {
    const int64_t base = (int64_t)blockIdx.x * blockDim.x;
    const int64_t z = base + threadIdx.x;
    const bool live = z < nz;
    float xn[NIN], y[2];
    {
      double c[L::NCons];
      Sym3 gl{1.0, 0.0, 0.0, 1.0, 0.0, 1.0};
      if (live) {
#pragma unroll
        for (int q = 0; q < L::NCons; ++q) c[q] = cons[q * nz + z];
        gl = Sym3{g[z], g[nz + z], g[2 * nz + z], g[3 * nz + z], g[4 * nz + z], g[5 * nz + z]};
      }
      zone_inputs_of<NIN>(live, c, gl, blob, xn);
    }
    net_wave<NIN>(S, xs[wv], ys[wv], xn, y);
    bool fail = false;
    if (live) {
      const double lw = (double)(y[0] * blob[L::OutStd] + blob[L::OutMean]);
      const double lzd = (double)(y[1] * blob[L::OutStd + 1] + blob[L::OutMean + 1]);
      // the guided solve never reads the seed; a guess that is not usable
      // (non-finite) sends the zone to the full chain from the seed
      C2PIn in;
      batch_zone_in<false>(z, nz, cons, g, seed, a, in);
      double Zg, vsq;
      guess_zv_of(lw, lzd, in, Zg, vsq);
      C2PArgs ak = a;
      ak.max_iter = kguess;
      C2POut o;
      o.iters = 0;
      if (Zg >= 0.0 && con2prim_zone<ASX_C2P_NOBLE, IdealGas, true>(eos, ak, in, o, Zg, vsq)) {
        batch_zone_out(z, nz, o, prim, flag);
        if (fallback) fallback[z] = 0;
      } else {
        fail = true;
      }
      iters[z] = o.iters;
    }
    queue_push(fail, (int)z, queue, qcount);
  }
}

// Network outputs in target units (log W, log(Z/D)) for every zone, all
// three layers on the matrix cores (net_wave), one chunk of 256 zones per
// workgroup.
template <int NIN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_mlp_forward(int64_t nz, const double *cons,
                                                     const double *g, const float *blob,
                                                     const NetImage *img, double *out) {
  using L = MlpLayout<NIN>;
  __shared__ NetImage S;
  __shared__ uint4 xs[4][64];
  __shared__ float ys[4][64][2];
  load_image(img, S);
  __syncthreads();
  const int wv = threadIdx.x >> 6;
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = z < nz;
  float xn[NIN], y[2];
  zone_inputs<NIN>(live, z, nz, cons, g, blob, xn);
  net_wave<NIN>(S, xs[wv], ys[wv], xn, y);
  if (live) {
    out[z] = (double)(y[0] * blob[L::OutStd] + blob[L::OutMean]);
    out[nz + z] = (double)(y[1] * blob[L::OutStd + 1] + blob[L::OutMean + 1]);
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static dim3 grid256(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_zone_sample(int64_t nz, uint64_t seed, double gamma, double seed_rel,
                              double *prim, double *g, double *sd, double *target,
                              hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zone_sample, grid256(nz), dim3(256), 0, s, nz, seed, gamma, seed_rel,
                     prim, g, sd, target);
  return hipGetLastError();
}
hipError_t launch_prim2con_batch(int64_t nz, const double *prim, const double *g,
                                 double *cons, double *cold, const IdealGas &eos, double Ye,
                                 hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prim2con_batch, grid256(nz), dim3(256), 0, s, nz, prim, g, cons, cold,
                     eos, Ye);
  return hipGetLastError();
}
template <int NIN>
hipError_t launch_mlp_features(int64_t nz, const double *cons, const double *g, float *x,
                               hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_features<NIN>, grid256(nz), dim3(256), 0, s, nz, cons, g, x);
  return hipGetLastError();
}
template <int NIN>
hipError_t launch_mlp_forward(int64_t nz, const double *cons, const double *g,
                              const float *blob, void *img, double *out, hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  NetImage *im = static_cast<NetImage *>(img);
  hipLaunchKernelGGL(k_net_image<NIN>, dim3(1), dim3(256), 0, s, blob, im);
  hipLaunchKernelGGL(k_mlp_forward<NIN>, grid256(nz), dim3(256), 0, s, nz, cons, g, blob, im,
                     out);
  return hipGetLastError();
}
template <int NIN>
hipError_t launch_mlp_grad(int64_t nz, const double *cons, const double *g,
                           const double *target, const float *blob, float *partial,
                           hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_grad<NIN>, dim3((unsigned)((nz + 63) / 64)), dim3(64), 0, s, nz,
                     cons, g, target, blob, partial);
  return hipGetLastError();
}
template <int NIN>
hipError_t launch_mlp_adam(int nwaves, const float *partial, float *blob, float *m,
                           float *v, float lr, float b1, float b2, float eps, int64_t step,
                           float *loss, float *grad, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam<NIN>, grid256(MlpLayout<NIN>::NParam + 1), dim3(256), 0, s,
                     nwaves, partial, blob, m, v, lr, b1, b2, eps, c1, c2, loss, grad);
  return hipGetLastError();
}
#define ASX_MLP_INSTANTIATE(N)                                                                \
  template hipError_t launch_mlp_features<N>(int64_t, const double *, const double *, float *, \
                                             hipStream_t);                                     \
  template hipError_t launch_mlp_forward<N>(int64_t, const double *, const double *,          \
                                            const float *, void *, double *, hipStream_t);     \
  template hipError_t launch_mlp_grad<N>(int64_t, const double *, const double *,             \
                                         const double *, const float *, float *, hipStream_t); \
  template hipError_t launch_mlp_adam<N>(int, const float *, float *, float *, float *, float, \
                                         float, float, float, int64_t, float *, float *,       \
                                         hipStream_t);
ASX_MLP_INSTANTIATE(kMlpIdeal)
ASX_MLP_INSTANTIATE(kMlpTab)
#undef ASX_MLP_INSTANTIATE
hipError_t launch_c2p_batch_mlp(int64_t nz, const double *cons, const double *g,
                                const double *seed, const float *blob, int kguess,
                                double *prim, double *flag, int32_t *iters,
                                int32_t *fallback, const IdealGas &eos, const C2PArgs &a,
                                double *guess, void *img, int *queue, unsigned *qcount,
                                int variant, hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  if (variant == 1) {
    // forward pass into guess[2][nz], then the primary solve from it
    hipError_t e = launch_mlp_forward<kMlpIdeal>(nz, cons, g, blob, img, guess, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, s, qcount);
    hipLaunchKernelGGL(k_c2p_guess_main, grid256(nz), dim3(256), 0, s, nz, cons, g, seed,
                       guess, kguess, prim, flag, iters, fallback, eos, a, queue, qcount);
  } else {
    NetImage *im = static_cast<NetImage *>(img);
    hipLaunchKernelGGL(k_net_image<kMlpIdeal>, dim3(1), dim3(256), 0, s, blob, im);
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, s, qcount);
    hipLaunchKernelGGL(k_c2p_guess_fused<kMlpIdeal>, grid256(nz), dim3(256), 0, s, nz, cons, g,
                       seed, blob, im, kguess, prim, flag, iters, fallback, eos, a, queue,
                       qcount);
  }
  const unsigned g2 = (unsigned)((nz + 255) / 256 < 2048 ? (nz + 255) / 256 : 2048);
  hipLaunchKernelGGL(k_c2p_guess_rest, dim3(g2), dim3(256), 0, s, nz, cons, g, seed, prim,
                     flag, iters, fallback, eos, a, queue, qcount);
  return hipGetLastError();
}

} // namespace asx
