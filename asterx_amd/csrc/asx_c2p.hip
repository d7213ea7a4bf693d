// Con2prim kernels for gfx950: AsterX_Con2Prim on a patch (optionally fused
// with ComputedBstagFromA + ComputedBFromdBstag) and Con2PrimFactory solves
// on independent zones.  One work-item per cell; flag and iteration
// histograms are reduced in LDS per block before one atomic per bin.
#include "asx_zone.hpp"

namespace asx {

struct HistLDS {
  unsigned int flags[8];
  unsigned int iters[64];
};

ASX_DEV void hist_init(HistLDS &h) {
  for (int q = threadIdx.x; q < 72; q += blockDim.x) {
    if (q < 8) h.flags[q] = 0;
    else h.iters[q - 8] = 0;
  }
}
ASX_DEV void hist_flush(HistLDS &h, unsigned long long *g) {
  for (int q = threadIdx.x; q < 71; q += blockDim.x) {
    const unsigned v = q < 7 ? h.flags[q] : h.iters[q - 7];
    if (v) atomicAdd(&g[q], (unsigned long long)v);
  }
}

// Pass 1: every zone through the primary solver (compiled for PRIME only,
// so the backup solvers cost no registers here).  Zones whose primary solve
// fails store nothing; a wavefront ballot packs their indices into `queue`
// with one atomic per wavefront.
template <bool FROM_A, int PRIME, class E>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(c2p_waves<E>())))
void k_c2p_main(KPatch P, Box B, E eos, C2PArgs a,
                                                  unsigned long long *hist, int *queue,
                                                  unsigned *qcount) {
  __shared__ HistLDS h;
  if (hist) {
    hist_init(h);
    __syncthreads();
  }
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < B.total();
  bool fail = false;
  if (live) {
    int i, j, k;
    B.ijk(t, i, j, k);
    C2PIn in;
    load_zone<FROM_A>(P, i, j, k, a, in);
    C2POut o;
    if (con2prim_zone<PRIME>(eos, a, in, o)) {
      store_zone(P, off(P, 0, i, j, k), o);
      if (hist) {
        atomicAdd(&h.flags[(int)o.flag], 1u);
        atomicAdd(&h.iters[o.iters < 63 ? o.iters : 63], 1u);
      }
    } else {
      fail = true;
    }
  }
  queue_push(fail, (int)t, queue, qcount);
  if (hist) {
    __syncthreads();
    hist_flush(h, hist);
  }
}

// Pass 2: the queued zones through the reference's full chain (primary
// again, then the backup), grid-stride over the device-side count.
template <bool FROM_A, class E>
__global__ __launch_bounds__(256) void k_c2p_rest(KPatch P, Box B, E eos, C2PArgs a,
                                                  unsigned long long *hist, const int *queue,
                                                  const unsigned *qcount) {
  __shared__ HistLDS h;
  if (hist) {
    hist_init(h);
    __syncthreads();
  }
  const unsigned nq = *qcount;
  for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += gridDim.x * blockDim.x) {
    int i, j, k;
    B.ijk(queue[q], i, j, k);
    C2PIn in;
    load_zone<FROM_A>(P, i, j, k, a, in);
    C2POut o;
    con2prim_cell(eos, a, in, o);
    store_zone(P, off(P, 0, i, j, k), o);
    if (hist) {
      atomicAdd(&h.flags[(int)o.flag], 1u);
      atomicAdd(&h.iters[o.iters < 63 ? o.iters : 63], 1u);
    }
  }
  if (hist) {
    __syncthreads();
    hist_flush(h, hist);
  }
}

__global__ void k_zero_u32(unsigned *p) { *p = 0u; }

template <int PRIME>
__global__ __launch_bounds__(256) void k_c2p_main_batch(const KPatch *__restrict__ tab,
                                                        const int *__restrict__ zbase,
                                                        IdealGas eos, C2PArgs a,
                                                        const BatchRef *__restrict__ map,
                                                        int nmap, int *queue,
                                                        unsigned *qcount) {
  const BatchRef r = map[xcd_tile((int)blockIdx.x, nmap)];
  const KPatch &P = tab[r.pid];
  const int t = r.chunk * 256 + (int)threadIdx.x;
  bool fail = false;
  if (t < P.n[0] * P.n[1] * P.n[2]) {
    int i, j, k;
    interior_ijk(P, t, i, j, k);
    C2PIn in;
    load_zone<true>(P, i, j, k, a, in);
    C2POut o;
    if (con2prim_zone<PRIME>(eos, a, in, o)) store_zone(P, off(P, 0, i, j, k), o);
    else fail = true;
  }
  queue_push(fail, zbase[r.pid] + t, queue, qcount);
}

__global__ __launch_bounds__(256) void k_c2p_rest_batch(const KPatch *__restrict__ tab,
                                                        const int *__restrict__ zbase, int np,
                                                        IdealGas eos, C2PArgs a,
                                                        const int *queue,
                                                        const unsigned *qcount) {
  const unsigned nq = *qcount;
  for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += gridDim.x * blockDim.x) {
    const int z = queue[q];
    const int p = patch_of(zbase, np, z);
    const KPatch &P = tab[p];
    int i, j, k;
    interior_ijk(P, z - zbase[p], i, j, k);
    C2PIn in;
    load_zone<true>(P, i, j, k, a, in);
    C2POut o;
    con2prim_cell(eos, a, in, o);
    store_zone(P, off(P, 0, i, j, k), o);
  }
}

// Independent zones (layout of asx_c2p_batch in include/asterx.h).
__global__ __launch_bounds__(256) void k_c2p_batch(int64_t nz, const double *cons,
                                                   const double *g, const double *seed,
                                                   double *prim, double *flag,
                                                   int32_t *iters, IdealGas eos,
                                                   C2PArgs a) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  C2PIn in;
  batch_zone_in(z, nz, cons, g, seed, a, in);
  C2POut o;
  con2prim_cell(eos, a, in, o);
  batch_zone_out(z, nz, o, prim, flag);
  iters[z] = o.iters;
}

static Box c2p_box(const KPatch &P) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = P.ng;
    b.cnt[d] = P.n[d];
  }
  return b;
}

template <bool FROM_A, int PRIME, class E>
static void c2p_main_t(const KPatch &P, const Box &b, const E &eos, const C2PArgs &a,
                       unsigned long long *hist, int *queue, unsigned *qcount, hipStream_t s) {
  hipLaunchKernelGGL((k_c2p_main<FROM_A, PRIME, E>), b.grid(), dim3(256), 0, s, P, b, eos, a,
                     hist, queue, qcount);
}
template <class E>
static void c2p_rest_t(const KPatch &P, const Box &b, const E &eos, const C2PArgs &a,
                       bool from_A, unsigned long long *hist, const int *queue,
                       const unsigned *qcount, hipStream_t s) {
  const int64_t nz = b.htotal();
  // enough work-items for a few percent of the zones; more stride
  const unsigned g2 = (unsigned)((nz + 255) / 256 < 2048 ? (nz + 255) / 256 : 2048);
  if (from_A)
    hipLaunchKernelGGL((k_c2p_rest<true, E>), dim3(g2), dim3(256), 0, s, P, b, eos, a, hist, queue, qcount);
  else
    hipLaunchKernelGGL((k_c2p_rest<false, E>), dim3(g2), dim3(256), 0, s, P, b, eos, a, hist, queue, qcount);
}

hipError_t launch_con2prim_main(const KPatch &P, const IdealGas &eos, const C2PArgs &a,
                                bool from_A, unsigned long long *hist, int *queue,
                                unsigned *qcount, hipStream_t s) {
  const Box b = c2p_box(P);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, qcount);
  if (b.htotal() <= 0) return hipGetLastError();
  const int pr = a.prime;
  if (from_A) {
    if (pr == ASX_C2P_NOBLE) c2p_main_t<true, ASX_C2P_NOBLE>(P, b, eos, a, hist, queue, qcount, s);
    else if (pr == ASX_C2P_ENTROPY) c2p_main_t<true, ASX_C2P_ENTROPY>(P, b, eos, a, hist, queue, qcount, s);
    else c2p_main_t<true, ASX_C2P_PALENZUELA>(P, b, eos, a, hist, queue, qcount, s);
  } else {
    if (pr == ASX_C2P_NOBLE) c2p_main_t<false, ASX_C2P_NOBLE>(P, b, eos, a, hist, queue, qcount, s);
    else if (pr == ASX_C2P_ENTROPY) c2p_main_t<false, ASX_C2P_ENTROPY>(P, b, eos, a, hist, queue, qcount, s);
    else c2p_main_t<false, ASX_C2P_PALENZUELA>(P, b, eos, a, hist, queue, qcount, s);
  }
  return hipGetLastError();
}
hipError_t launch_con2prim_main(const KPatch &P, const TabEOS &eos, const C2PArgs &a,
                                bool from_A, unsigned long long *hist, int *queue,
                                unsigned *qcount, hipStream_t s) {
  const Box b = c2p_box(P);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, qcount);
  if (b.htotal() <= 0) return hipGetLastError();
  if (from_A) c2p_main_t<true, ASX_C2P_PALENZUELA>(P, b, eos, a, hist, queue, qcount, s);
  else c2p_main_t<false, ASX_C2P_PALENZUELA>(P, b, eos, a, hist, queue, qcount, s);
  return hipGetLastError();
}

hipError_t launch_con2prim_rest(const KPatch &P, const IdealGas &eos, const C2PArgs &a,
                                bool from_A, unsigned long long *hist, const int *queue,
                                const unsigned *qcount, hipStream_t s) {
  const Box b = c2p_box(P);
  if (b.htotal() <= 0) return hipSuccess;
  c2p_rest_t(P, b, eos, a, from_A, hist, queue, qcount, s);
  return hipGetLastError();
}
hipError_t launch_con2prim_rest(const KPatch &P, const TabEOS &eos, const C2PArgs &a,
                                bool from_A, unsigned long long *hist, const int *queue,
                                const unsigned *qcount, hipStream_t s) {
  const Box b = c2p_box(P);
  if (b.htotal() <= 0) return hipSuccess;
  c2p_rest_t(P, b, eos, a, from_A, hist, queue, qcount, s);
  return hipGetLastError();
}

hipError_t launch_con2prim_main_batch(const KPatch *tab, const int *zbase, int np,
                                      const IdealGas &eos, const C2PArgs &a,
                                      const BatchRef *map, int nmap, int *queue,
                                      unsigned *qcount, hipStream_t s) {
  (void)np;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, qcount);
  if (nmap <= 0) return hipGetLastError();
  const dim3 grid((unsigned)nmap), block(256);
  if (a.prime == ASX_C2P_NOBLE)
    hipLaunchKernelGGL(k_c2p_main_batch<ASX_C2P_NOBLE>, grid, block, 0, s, tab, zbase, eos, a, map, nmap, queue, qcount);
  else if (a.prime == ASX_C2P_ENTROPY)
    hipLaunchKernelGGL(k_c2p_main_batch<ASX_C2P_ENTROPY>, grid, block, 0, s, tab, zbase, eos, a, map, nmap, queue, qcount);
  else
    hipLaunchKernelGGL(k_c2p_main_batch<ASX_C2P_PALENZUELA>, grid, block, 0, s, tab, zbase, eos, a, map, nmap, queue, qcount);
  return hipGetLastError();
}

hipError_t launch_con2prim_rest_batch(const KPatch *tab, const int *zbase, int np,
                                      const IdealGas &eos, const C2PArgs &a, int64_t nzones,
                                      const int *queue, const unsigned *qcount,
                                      hipStream_t s) {
  if (nzones <= 0 || np <= 0) return hipSuccess;
  const unsigned g2 = (unsigned)((nzones + 255) / 256 < 2048 ? (nzones + 255) / 256 : 2048);
  hipLaunchKernelGGL(k_c2p_rest_batch, dim3(g2), dim3(256), 0, s, tab, zbase, np, eos, a, queue, qcount);
  return hipGetLastError();
}

hipError_t launch_c2p_batch(int64_t nz, const double *cons, const double *g,
                            const double *seed, double *prim, double *flag,
                            int32_t *iters, const IdealGas &eos, const C2PArgs &a,
                            hipStream_t s) {
  if (nz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_c2p_batch, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s, nz,
                     cons, g, seed, prim, flag, iters, eos, a);
  return hipGetLastError();
}

} // namespace asx
