// A batch plan's SSPRK3 stage (AsterX/schedule.ccl:119-232 over every local
// box) as ONE persistent cooperative launch for gfx950.  Every workgroup
// stays resident and walks the phases of the stage in order, separated by
// grid-wide barriers:
//   1. faces normal to x, y, z and the vertex auxiliaries of every patch
//      (independent of one another; the con2prim queue count is cleared)
//   2. the fused RHS + RK combination of every patch
//   3. boundary fill of the 11 evolved fields, axis x, then y, then z
//   4. con2prim, Noble primary, failures queued by a wavefront ballot
//   5. the queued zones through the full chain
//   6. boundary fill of the 12 synced primitives, axis x, then y, then z
// The work of each phase is the plan's (patch, chunk) list; XCD x (the
// workgroups with blockIdx.x % 8 == x) takes one contiguous run of it, so
// neighbouring tiles of a patch share one L2.  The bodies are the device
// functions of the single-launch kernels, so the results equal the plan's
// launch-per-phase form bit for bit.  No early return anywhere: every
// workgroup reaches every barrier.
#include <hip/hip_cooperative_groups.h>

#include "asx_face.hpp"
#include "asx_update.hpp"
#include "asx_zone.hpp"

namespace asx {

namespace cg = cooperative_groups;

// this workgroup's items of a list of n: a contiguous run per XCD
struct XcdRun {
  int first, end, step;
};
ASX_DEV XcdRun xcd_run(int n) {
  constexpr int NX = 8;
  const int x = (int)blockIdx.x % NX, l = (int)blockIdx.x / NX, nb = (int)gridDim.x / NX;
  const int lo = (int)((int64_t)n * x / NX), hi = (int)((int64_t)n * (x + 1) / NX);
  return XcdRun{lo + l, hi, nb};
}

// Barrier between phases: every workgroup's writes reach device scope
// before any workgroup reads them in the next phase.
ASX_DEV void phase_barrier(cg::grid_group &grid) {
  __threadfence();
  grid.sync();
}

// Item bodies (inlined: as calls, each would save and restore most of the
// register file around every item).
#define ASX_ITEM ASX_DEV

template <int M>
ASX_ITEM void face_item(const PersistArgs &S, int d, int q, FaceLDS &L) {
  const BatchRef ref = S.face[d].map[q];
  const BatchBox &bb = S.face[d].box[ref.pid];
  const KPatch &P = S.T[ref.pid];
  if (d == 0) face_tile<0, M, 0>(P, S.A, bb.lo, bb.cnt, ref.chunk, bb.nt[0], bb.nt[1], L);
  else if (d == 1) face_tile<1, M, 0>(P, S.A, bb.lo, bb.cnt, ref.chunk, bb.nt[0], bb.nt[1], L);
  else face_tile<2, M, 0>(P, S.A, bb.lo, bb.cnt, ref.chunk, bb.nt[0], bb.nt[1], L);
}

ASX_ITEM void aux_item(const PersistArgs &S, int q) {
  const BatchRef ref = S.aux.map[q];
  int i, j, k;
  if (batch_point(S.aux.box[ref.pid], ref.chunk, i, j, k)) aux_point(S.T[ref.pid], i, j, k);
}

template <int ORD> ASX_ITEM void update_item(const PersistArgs &S, int q) {
  const BatchRef ref = S.update.map[q];
  const BatchBox &bb = S.update.box[ref.pid];
  int i, j, k;
  if (tile_point<kPersistUpdateTile[0], kPersistUpdateTile[1], kPersistUpdateTile[2]>(
          bb.lo, bb.cnt, ref.chunk, bb.nt[0], bb.nt[1], i, j, k))
    update_point<ORD>(S.T[ref.pid], i, j, k, S.R, S.u0[ref.pid], S.out[ref.pid],
                      Comb{S.c0, S.c1, S.c2});
}

template <int D> ASX_ITEM void fill_item(const PersistArgs &S, const BCSlots &F, int q) {
  const BatchList &L = S.bc[D];
  const int f = q / L.n;
  const BatchRef ref = L.map[q - f * L.n];
  const KPatch &P = S.U[ref.pid];
  bc_point<D>(P.g[F.slot[f]], F.cent[f], P.n[0], P.n[1], P.n[2], P.ng, S.bc_kind,
              ref.chunk * 256 + (int)threadIdx.x);
}

// con2prim's primary solve on one chunk of 256 interior zones; failures
// queued (every lane of the wavefront takes part in the ballot)
ASX_ITEM void c2p_item(const PersistArgs &S, int q) {
  const BatchRef ref = S.c2p.map[q];
  const KPatch &P = S.U[ref.pid];
  const int t = ref.chunk * 256 + (int)threadIdx.x;
  bool fail = false;
  if (t < P.n[0] * P.n[1] * P.n[2]) {
    int i, j, k;
    interior_ijk(P, t, i, j, k);
    C2PIn in;
    load_zone<true>(P, i, j, k, S.ca, in);
    C2POut o;
    if (con2prim_zone<ASX_C2P_NOBLE>(S.eos, S.ca, in, o)) store_zone(P, off(P, 0, i, j, k), o);
    else fail = true;
  }
  queue_push(fail, S.zbase[ref.pid] + t, S.queue, S.qcount);
}

// one queued zone through the full chain
ASX_ITEM void c2p_rest_item(const PersistArgs &S, int z) {
  const int p = patch_of(S.zbase, S.np, z);
  const KPatch &P = S.U[p];
  int i, j, k;
  interior_ijk(P, z - S.zbase[p], i, j, k);
  C2PIn in;
  load_zone<true>(P, i, j, k, S.ca, in);
  C2POut o;
  con2prim_cell(S.eos, S.ca, in, o);
  store_zone(P, off(P, 0, i, j, k), o);
}

template <int D>
ASX_DEV void fill_axis(const PersistArgs &S, const BCSlots &F) {
  const XcdRun r = xcd_run(S.bc[D].n * F.nf);
  for (int it = r.first; it < r.end; it += r.step) fill_item<D>(S, F, it);
}

// Register budget: 2 waves per SIMD (256 VGPRs), what the con2prim solve
// uses in its own kernel; the faces, which run at 3 in their own kernels,
// run at 2 here.  With every phase in one body the allocator still spills
// ~200 B per lane (the single-launch kernels spill nothing): one of the
// costs of the single launch.
template <int M, int ORD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_stage_persist(const PersistArgs *__restrict__ SA) {
  cg::grid_group grid = cg::this_grid();
  __shared__ FaceLDS L;
  const PersistArgs &S = *SA;
  // 1. faces and aux
  if (blockIdx.x == 0 && threadIdx.x == 0) *S.qcount = 0u;
  {
    const int n0 = S.face[0].n, n1 = S.face[1].n, n2 = S.face[2].n;
    const XcdRun r = xcd_run(n0 + n1 + n2 + S.aux.n);
    for (int it = r.first; it < r.end; it += r.step) {
      if (it < n0) face_item<M>(S, 0, it, L);
      else if (it < n0 + n1) face_item<M>(S, 1, it - n0, L);
      else if (it < n0 + n1 + n2) face_item<M>(S, 2, it - n0 - n1, L);
      else aux_item(S, it - n0 - n1 - n2);
    }
  }
  phase_barrier(grid);
  // 2. RHS + RK combination
  {
    const XcdRun r = xcd_run(S.update.n);
    for (int it = r.first; it < r.end; it += r.step) update_item<ORD>(S, it);
  }
  phase_barrier(grid);
  // 3. fill of the evolved fields
  fill_axis<0>(S, S.state);
  phase_barrier(grid);
  fill_axis<1>(S, S.state);
  phase_barrier(grid);
  fill_axis<2>(S, S.state);
  phase_barrier(grid);
  // 4. con2prim, primary solver
  {
    const XcdRun r = xcd_run(S.c2p.n);
    for (int it = r.first; it < r.end; it += r.step) c2p_item(S, it);
  }
  phase_barrier(grid);
  // 5. queued zones, full chain
  {
    const unsigned nq = __hip_atomic_load(S.qcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned nthreads = gridDim.x * blockDim.x;
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += nthreads)
      c2p_rest_item(S, S.queue[q]);
  }
  phase_barrier(grid);
  // 6. fill of the synced primitives
  fill_axis<0>(S, S.prims);
  phase_barrier(grid);
  fill_axis<1>(S, S.prims);
  phase_barrier(grid);
  fill_axis<2>(S, S.prims);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int M, int ORD> static const void *persist_kernel() {
  return reinterpret_cast<const void *>(&k_stage_persist<M, ORD>);
}
static const void *persist_kernel(int method, int fd_order) {
  const bool o4 = fd_order == 4;
  switch (method) {
  case ASX_RECON_PPM: return o4 ? persist_kernel<ASX_RECON_PPM, 4>() : persist_kernel<ASX_RECON_PPM, 2>();
  case ASX_RECON_WENOZ: return o4 ? persist_kernel<ASX_RECON_WENOZ, 4>() : persist_kernel<ASX_RECON_WENOZ, 2>();
  case ASX_RECON_MP5: return o4 ? persist_kernel<ASX_RECON_MP5, 4>() : persist_kernel<ASX_RECON_MP5, 2>();
  default: return nullptr;
  }
}

bool persist_supported(int method, int fd_order) {
  return (fd_order == 2 || fd_order == 4) && persist_kernel(method, fd_order) != nullptr;
}

int persist_grid(int method, int fd_order) {
  const void *k = persist_kernel(method, fd_order);
  if (!k) return 0;
  int dev = 0, per_cu = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess ||
      per_cu < 1)
    return 0;
  const int g = per_cu * prop.multiProcessorCount;
  return g - g % 8;
}

hipError_t launch_stage_persist(const PersistArgs *dev_args, int method, int fd_order,
                                int grid, hipStream_t s) {
  const void *k = persist_kernel(method, fd_order);
  if (!k || grid < 8 || grid % 8) return hipErrorInvalidValue;
  void *args[] = {(void *)&dev_args};
  return hipLaunchCooperativeKernel(k, dim3((unsigned)grid), dim3(256), args, 0, s);
}

} // namespace asx
