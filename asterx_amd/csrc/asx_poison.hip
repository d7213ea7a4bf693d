// Debug poison (CarpetX::poison_undefined_values): fill grid functions with
// a NaN bit pattern before an entry point writes them, and count the points
// of a box that still hold it afterwards.  Plain streaming kernels; the
// count is exact (bitwise comparison, so NaNs computed from poisoned inputs,
// which the hardware quiets, are not mistaken for unwritten points).
#include <algorithm>

#include "asx_kernels.hpp"

namespace asx {
namespace {

__global__ __launch_bounds__(256) void k_poison_fill(uint64_t *__restrict__ p, int64_t n,
                                                     uint64_t v) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}

__global__ __launch_bounds__(256) void k_poison_count(const uint64_t *__restrict__ p, int sj,
                                                      int sk, Box B, uint64_t v,
                                                      unsigned long long *cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool hit = false;
  if (t < B.total()) {
    int i, j, k;
    B.ijk(t, i, j, k);
    hit = p[i + (int64_t)sj * j + (int64_t)sk * k] == v;
  }
  const uint64_t m = __ballot(hit);
  if (m && (threadIdx.x & 63) == __builtin_ctzll(m))
    atomicAdd(cnt, (unsigned long long)__popcll(m));
}

}  // namespace

hipError_t launch_poison_fill(double *p, int64_t n, uint64_t bits, hipStream_t s) {
  if (n > 0) {
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_poison_fill, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<uint64_t *>(p), n, bits);
  }
  return hipGetLastError();
}

hipError_t launch_poison_count(const double *p, int sj, int sk, const Box &B, uint64_t bits,
                               unsigned long long *cnt, hipStream_t s) {
  if (B.htotal() > 0)
    hipLaunchKernelGGL(k_poison_count, B.grid(), dim3(256), 0, s,
                       reinterpret_cast<const uint64_t *>(p), sj, sk, B, bits, cnt);
  return hipGetLastError();
}

}  // namespace asx
