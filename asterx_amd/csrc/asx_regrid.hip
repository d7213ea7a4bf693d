// Regrid error estimate on the interior cells (AsterX/src/estimate_error.cxx:
// 72-142 with the estimators of estimate_error.hxx:14-100): the maximum over
// up to 9 grid functions of a first- or second-difference indicator of each
// at the cell's index.  One work-item per cell; the 7-point reads of
// neighbouring work-items share cache lines.
#include "asx_kernels.hpp"

namespace asx {
namespace {

// std::max of the reference: keeps m unless m < x (NaN never replaces m)
ASX_DEV double amax(double m, double x) { return m < x ? x : m; }

// One grid function's indicator at offset o of its own layout; st[d] is the
// offset of one step along axis d.
ASX_DEV double estimate(const double *__restrict__ g, int o, const int st[3], int method,
                        double eps) {
  const double v0 = g[o];
  double err = 0.0, ep = 0.0, em = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double vm = g[o - st[d]], vp = g[o + st[d]];
    if (method == ASX_REGRID_FIRST_DERIV) {
      err = amax(amax(err, fabs(v0 - vm)), fabs(vp - v0));
    } else if (method == ASX_REGRID_SECOND_DERIV) {
      err = amax(err, fabs(vp + vm - 2 * v0));
    } else if (method == ASX_REGRID_SECOND_DERIV_NORM) {
      // Mignone et al. 2011 eq. 50
      const double dp = vp - v0, dm = v0 - vm;
      const double sigma = fabs(vp) + 2 * fabs(v0) + fabs(vm);
      const double num = fabs(dp - dm) * fabs(dp - dm);
      const double den = (fabs(dp) + fabs(dm) + eps * sigma) * (fabs(dp) + fabs(dm) + eps * sigma);
      err = amax(err, sqrt(num / den));
    } else {
      ep += (vp - v0) * (vp - v0);
      em += (v0 - vm) * (v0 - vm);
    }
  }
  return method == ASX_REGRID_FIRST_GRAD ? sqrt(amax(ep, em)) : err;
}

__global__ __launch_bounds__(256) void k_regrid(KPatch P, Box B, RegridArgs A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.total()) return;
  int i, j, k;
  B.ijk(t, i, j, k);
  double e = 0.0;
  for (int q = 0; q < A.nv; ++q) {
    const int c = A.cent[q];
    const int st[3] = {1, P.sj[c], P.sk[c]};
    e = amax(e, estimate(P.g[A.slot[q]], off(P, c, i, j, k), st, A.method, A.eps));
  }
  P.g[ASX_REGRID_ERROR][off(P, 0, i, j, k)] = e;
}

}  // namespace

hipError_t launch_regrid(const KPatch &P, const Box &B, const RegridArgs &A, hipStream_t s) {
  if (B.htotal() > 0) hipLaunchKernelGGL(k_regrid, B.grid(), dim3(256), 0, s, P, B, A);
  return hipGetLastError();
}

}  // namespace asx
