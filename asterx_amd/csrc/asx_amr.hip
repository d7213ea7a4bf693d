// Refinement by a factor 2 between patches of adjacent levels: the "ddf"
// prolongation and the restriction CarpetX applies to AsterX's grid
// functions (CarpetX::prolongation_type = "ddf", AsterX/test/*.par), and
// the box table of the next level from regrid_error (tagging on the device,
// Berger-Rigoutsos clustering on the host).
//
// Prolongation: one work-item per coarse block of 2 x 2 x 2 fine points
// (the children of one coarse index per axis; for the ghost region only the
// shell of blocks holding ghost points is launched), templated on order and
// centering so that the stencils and weights are constants.  Per axis the
// children's stencils lie in P or P+1 coarse points (vertex axis: the
// coinciding point and the centred Lagrange midpoint weights; cell axis:
// the conservative child weights).  The 3-D value is the tensor product,
// summed x innermost, then y, then z; each partial sum starts at 0 and adds
// weight * value in stencil order, so the oracle's numpy restatement gives
// the same bits.  Each coarse row is read once per block; neighbouring
// blocks share their rows in L1/L2.
#include <algorithm>
#include <vector>

#include "asx_kernels.hpp"

namespace asx {

// Prolongation weights along one axis of vertex flag VC for order P.  A
// work-item takes a pair of fine points (children h = 0, 1 of coarse index
// a) whose stencils lie in the NU coarse points a - (P-1)/2 .. : along a
// vertex axis child 0 sits on point a (one term, weight 1) and child 1 is
// the midpoint of a and a+1 (the P+1 centred Lagrange weights); along a
// cell axis both children use the P cells centred on a with the
// conservative weights of the lower half (child 0) or their mirror image.
template <int P, int VC> struct AxisW {
  static constexpr int NU = VC ? P + 1 : P;
  static constexpr bool uses(int h, int m) { return !(VC && h == 0) || m == (P - 1) / 2; }
  static constexpr double poly(int m) {
    // centred Lagrange interpolant at the midpoint of P+1 unit-spaced points
    return P == 1 ? 0.5
           : P == 3 ? (m == 0 || m == 3 ? -1.0 / 16 : 9.0 / 16)
                    : (m == 0 || m == 5 ? 3.0 / 256 : m == 1 || m == 4 ? -25.0 / 256 : 75.0 / 128);
  }
  static constexpr double cons(int m) {
    // lower child's average of the degree P-1 polynomial with the P cell
    // averages centred on the parent
    return P == 1 ? 1.0
           : P == 3 ? (m == 0 ? 1.0 / 8 : m == 1 ? 1.0 : -1.0 / 8)
                    : (m == 0 ? -3.0 / 128 : m == 1 ? 11.0 / 64 : m == 2 ? 1.0
                       : m == 3 ? -11.0 / 64 : 3.0 / 128);
  }
  static constexpr double w(int h, int m) {
    return VC ? (h == 0 ? 1.0 : poly(m)) : cons(h ? P - 1 - m : m);
  }
};

// The 2 x 2 x 2 fine points of coarse block t of slot s: each fine point's
// value is the tensor-product sum, x innermost, then y, then z, every
// partial sum from 0 in increasing stencil order; the x sums of a coarse
// row serve all four (y, z) children and the y sums both z children, which
// leaves every fine point's sequence of operations unchanged.
template <int P, int C>
ASX_DEV void prolong_block(const ProlongArgs &A, int s, int64_t t) {
  using WX = AxisW<P, C & 1>;
  using WY = AxisW<P, (C >> 1) & 1>;
  using WZ = AxisW<P, (C >> 2) & 1>;
  int E[3], a[3];
  for (int d = 0; d < 3; ++d) E[d] = A.fn[d] + ((C >> d) & 1) + 2 * A.fng;
  const BlockShell sh(A.off, E, A.fng);
  if (t >= sh.total(A.region == 1)) return;
  sh.decode(t, A.region == 1, a);
  for (int d = 0; d < 3; ++d) a[d] += floor2(A.off[d] - A.fng);
  int ib[3];
  unsigned in[3] = {0, 0, 0}, own[3] = {0, 0, 0}; // bit h: child h inside / own point
  for (int d = 0; d < 3; ++d) {
    ib[d] = 2 * a[d] - A.off[d] + A.fng; // fine local index of child 0
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = ib[d] + h;
      in[d] |= (unsigned)(i >= 0 && i < E[d]) << h;
      own[d] |= (unsigned)(i >= A.fng && i < E[d] - A.fng) << h;
    }
  }
  unsigned wm = 0; // bit hx + 2 hy + 4 hz: fine point written
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int hx = q & 1, hy = (q >> 1) & 1, hz = q >> 2;
    const bool inside = ((in[0] >> hx) & (in[1] >> hy) & (in[2] >> hz) & 1u) != 0;
    const bool mine = ((own[0] >> hx) & (own[1] >> hy) & (own[2] >> hz) & 1u) != 0;
    if (inside && (A.region == 0 || !mine)) wm |= 1u << q;
  }
  if (!wm) return;
  // Along a vertex axis whose child 1 lies outside the patch only child 0's
  // coarse point is read: every load belongs to the stencil of a fine point
  // inside the patch, which the host checked against the coarse extent.
  constexpr int MID = (P - 1) / 2;
  const bool fx = !(C & 1) || (in[0] & 2u), fy = !((C >> 1) & 1) || (in[1] & 2u),
             fz = !((C >> 2) & 1) || (in[2] & 2u);
  const int csj = A.csj[C], csk = A.csk[C];
  const double *__restrict__ u = A.src[s] + (int64_t)(a[2] - (P - 1) / 2 + A.cng) * csk +
                                 (int64_t)(a[1] - (P - 1) / 2 + A.cng) * csj +
                                 (a[0] - (P - 1) / 2 + A.cng);
  double az[2][2][2] = {{{0.0, 0.0}, {0.0, 0.0}}, {{0.0, 0.0}, {0.0, 0.0}}};
  // order 5 keeps one coarse plane of loads in flight (all six would
  // spill at the 4-wave register budget)
#pragma unroll(P >= 5 ? 1 : WZ::NU)
  for (int cz = 0; cz < WZ::NU; ++cz) {
    if (!fz && cz != MID) continue;
    double ay[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int cy = 0; cy < WY::NU; ++cy) {
      if (!fy && cy != MID) continue;
      const double *row = u + (int64_t)cz * csk + (int64_t)cy * csj;
      double r[WX::NU];
#pragma unroll
      for (int m = 0; m < WX::NU; ++m) r[m] = 0.0;
      if (fx) {
#pragma unroll
        for (int m = 0; m < WX::NU; ++m) r[m] = row[m];
      } else {
        r[MID] = row[MID];
      }
      double ax[2];
#pragma unroll
      for (int hx = 0; hx < 2; ++hx) {
        ax[hx] = 0.0;
#pragma unroll
        for (int m = 0; m < WX::NU; ++m)
          if (WX::uses(hx, m)) ax[hx] = ax[hx] + WX::w(hx, m) * r[m];
      }
#pragma unroll
      for (int hy = 0; hy < 2; ++hy)
        if (WY::uses(hy, cy))
#pragma unroll
          for (int hx = 0; hx < 2; ++hx) ay[hy][hx] = ay[hy][hx] + WY::w(hy, cy) * ax[hx];
    }
#pragma unroll
    for (int hz = 0; hz < 2; ++hz)
      if (WZ::uses(hz, cz))
#pragma unroll
        for (int hy = 0; hy < 2; ++hy)
#pragma unroll
          for (int hx = 0; hx < 2; ++hx)
            az[hz][hy][hx] = az[hz][hy][hx] + WZ::w(hz, cz) * ay[hy][hx];
  }
  double *__restrict__ f = A.dst[s];
  const int fsj = A.fsj[C], fsk = A.fsk[C];
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if ((wm >> q) & 1u) {
      const int hx = q & 1, hy = (q >> 1) & 1, hz = q >> 2;
      f[(ib[0] + hx) + (int64_t)fsj * (ib[1] + hy) + (int64_t)fsk * (ib[2] + hz)] =
          az[hz][hy][hx];
    }
}

template <int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_prolong(ProlongArgs A) {
  const int s = (int)blockIdx.y;
  // contiguous runs of blocks per XCD: neighbouring blocks share most of
  // their coarse rows, which then meet in one L2
  const int64_t t =
      (int64_t)xcd_tile((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
  switch (A.cent[s]) {
  case 0: prolong_block<P, 0>(A, s, t); break;
  case 1: prolong_block<P, 1>(A, s, t); break;
  case 2: prolong_block<P, 2>(A, s, t); break;
  case 3: prolong_block<P, 3>(A, s, t); break;
  case 4: prolong_block<P, 4>(A, s, t); break;
  case 5: prolong_block<P, 5>(A, s, t); break;
  case 6: prolong_block<P, 6>(A, s, t); break;
  default: prolong_block<P, 7>(A, s, t); break;
  }
}

// Coarse points [lo, lo+cnt) per axis of slot s's centering
__global__ __launch_bounds__(256) void k_restrict(RestrictArgs A) {
  const int s = (int)blockIdx.y;
  const int c = A.cent[s];
  const Box &B = A.box[s];
  const int64_t t =
      (int64_t)xcd_tile((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int ic[3];
  B.ijk(t, ic[0], ic[1], ic[2]);
  int st[3], n[3];
  for (int d = 0; d < 3; ++d) {
    // fine index of the first fine point: 2 (ic - cng) - off, local + fng
    st[d] = 2 * (ic[d] - A.cng) - A.off[d] + A.fng;
    n[d] = ((c >> d) & 1) ? 1 : 2;
  }
  const double *__restrict__ u = A.src[s];
  const int fsj = A.fsj[c], fsk = A.fsk[c];
  double az = 0.0;
  for (int cz = 0; cz < n[2]; ++cz) {
    double ay = 0.0;
    for (int cy = 0; cy < n[1]; ++cy) {
      const double *row = u + (int64_t)(st[2] + cz) * fsk + (int64_t)(st[1] + cy) * fsj;
      double ax = 0.0;
      for (int cx = 0; cx < n[0]; ++cx) ax = ax + (n[0] == 1 ? 1.0 : 0.5) * row[st[0] + cx];
      ay = ay + (n[1] == 1 ? 1.0 : 0.5) * ax;
    }
    az = az + (n[2] == 1 ? 1.0 : 0.5) * ay;
  }
  A.dst[s][ic[0] + (int64_t)A.csj[c] * ic[1] + (int64_t)A.csk[c] * ic[2]] = az;
}

// One work-item per block of blk[0] x blk[1] x blk[2] interior cells: the
// block is tagged when a cell within buf cells of it (inside the interior)
// has regrid_error > thr.
__global__ __launch_bounds__(256) void k_regrid_tags(TagArgs A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)A.nb[0] * A.nb[1] * A.nb[2]) return;
  const int b[3] = {(int)(t % A.nb[0]), (int)((t / A.nb[0]) % A.nb[1]),
                    (int)(t / ((int64_t)A.nb[0] * A.nb[1]))};
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = max(0, b[d] * A.blk[d] - A.buf);
    hi[d] = min(A.n[d] - 1, (b[d] + 1) * A.blk[d] - 1 + A.buf);
  }
  bool tag = false;
  for (int k = lo[2]; k <= hi[2] && !tag; ++k)
    for (int j = lo[1]; j <= hi[1] && !tag; ++j) {
      const double *row = A.err + (int64_t)(k + A.ng) * A.sk + (int64_t)(j + A.ng) * A.sj + A.ng;
      for (int i = lo[0]; i <= hi[0]; ++i)
        if (row[i] > A.thr) {
          tag = true;
          break;
        }
    }
  A.tags[t] = tag ? 1 : 0;
}

hipError_t launch_prolong(const ProlongArgs &A, int64_t max_blocks, hipStream_t s) {
  if (max_blocks > 0 && A.nslot > 0) {
    const dim3 g((unsigned)((max_blocks + 255) / 256), (unsigned)A.nslot);
    if (A.order == 1) hipLaunchKernelGGL(k_prolong<1>, g, dim3(256), 0, s, A);
    else if (A.order == 3) hipLaunchKernelGGL(k_prolong<3>, g, dim3(256), 0, s, A);
    else hipLaunchKernelGGL(k_prolong<5>, g, dim3(256), 0, s, A);
  }
  return hipGetLastError();
}

hipError_t launch_restrict(const RestrictArgs &A, int64_t max_points, hipStream_t s) {
  if (max_points > 0 && A.nslot > 0)
    hipLaunchKernelGGL(k_restrict, dim3((unsigned)((max_points + 255) / 256), (unsigned)A.nslot),
                       dim3(256), 0, s, A);
  return hipGetLastError();
}

hipError_t launch_regrid_tags(const TagArgs &A, hipStream_t s) {
  const int64_t nb = (int64_t)A.nb[0] * A.nb[1] * A.nb[2];
  if (nb > 0) hipLaunchKernelGGL(k_regrid_tags, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, A);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Berger-Rigoutsos clustering of the block tags (host).  Boxes inclusive, in
// block indices.
// ---------------------------------------------------------------------------
namespace {

struct Clusterer {
  const uint8_t *t;
  int nb[3];
  double eff;
  std::vector<ClusterBox> out;

  uint8_t tag(int i, int j, int k) const { return t[i + (int64_t)nb[0] * (j + (int64_t)nb[1] * k)]; }

  void run(ClusterBox b) {
    // tag signatures of b along each axis
    std::vector<int64_t> sig[3];
    for (int d = 0; d < 3; ++d) sig[d].assign(b.hi[d] - b.lo[d] + 1, 0);
    int64_t total = 0;
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
      for (int j = b.lo[1]; j <= b.hi[1]; ++j)
        for (int i = b.lo[0]; i <= b.hi[0]; ++i)
          if (tag(i, j, k)) {
            ++sig[0][i - b.lo[0]];
            ++sig[1][j - b.lo[1]];
            ++sig[2][k - b.lo[2]];
            ++total;
          }
    if (total == 0) return;
    // shrink to the tags' bounding box (the signatures are unchanged)
    for (int d = 0; d < 3; ++d) {
      int a = 0, e = (int)sig[d].size() - 1;
      while (sig[d][a] == 0) ++a;
      while (sig[d][e] == 0) --e;
      sig[d] = std::vector<int64_t>(sig[d].begin() + a, sig[d].begin() + e + 1);
      b.hi[d] = b.lo[d] + e;
      b.lo[d] += a;
    }
    const int64_t vol = (int64_t)sig[0].size() * sig[1].size() * sig[2].size();
    if (vol == 1 || (double)total >= eff * (double)vol) {
      out.push_back(b);
      return;
    }
    int cd = -1, cut = 0; // cut: first plane (relative) of the upper part
    // 1. an empty plane, the one nearest the middle (relative to the length)
    double best = 2.0;
    for (int d = 0; d < 3; ++d) {
      const int L = (int)sig[d].size();
      for (int x = 1; x < L - 1; ++x)
        if (sig[d][x] == 0) {
          const double dist = std::abs(2.0 * x - (L - 1)) / L;
          if (dist < best) {
            best = dist;
            cd = d;
            cut = x;
          }
        }
    }
    // 2. the strongest sign change of the signature's second difference
    if (cd < 0) {
      int64_t strength = 0;
      double bdist = 2.0;
      for (int d = 0; d < 3; ++d) {
        const int L = (int)sig[d].size();
        for (int x = 1; x + 2 < L; ++x) {
          const int64_t l0 = sig[d][x + 1] - 2 * sig[d][x] + sig[d][x - 1];
          const int64_t l1 = sig[d][x + 2] - 2 * sig[d][x + 1] + sig[d][x];
          if ((l0 < 0 && l1 > 0) || (l0 > 0 && l1 < 0)) {
            const int64_t s = l1 > l0 ? l1 - l0 : l0 - l1;
            const double dist = std::abs(2.0 * (x + 1) - L) / L;
            if (s > strength || (s == strength && dist < bdist)) {
              strength = s;
              bdist = dist;
              cd = d;
              cut = x + 1;
            }
          }
        }
      }
    }
    // 3. halve the longest axis
    if (cd < 0) {
      cd = 0;
      for (int d = 1; d < 3; ++d)
        if (sig[d].size() > sig[cd].size()) cd = d;
      cut = (int)sig[cd].size() / 2;
    }
    ClusterBox lo = b, up = b;
    lo.hi[cd] = b.lo[cd] + cut - 1;
    up.lo[cd] = b.lo[cd] + cut;
    run(lo);
    run(up);
  }
};

}  // namespace

std::vector<ClusterBox> cluster_tags(const uint8_t *tags, const int nb[3], double grid_eff) {
  Clusterer C{tags, {nb[0], nb[1], nb[2]}, grid_eff, {}};
  C.run(ClusterBox{{0, 0, 0}, {nb[0] - 1, nb[1] - 1, nb[2] - 1}});
  return C.out;
}

}  // namespace asx
