// Per-column step error of the fixed-dt integration loop:
//   delta_c = max_{z,i} delta[z,i] / max(|sol[z,i]|, 1e-30)
// (integrate.py:30-33).  One workgroup per column; a max reduction, so the
// result does not depend on the order of the lanes.
#include "vk_common.h"

__global__ __launch_bounds__(VK_NT) void k_delta_max(const VkCol* __restrict__ cols, int ncols,
                                                     const double* __restrict__ sol,
                                                     const double* __restrict__ delta,
                                                     double* __restrict__ out, int64_t stride) {
  __shared__ double red[VK_NT / 64];
  const int c = blockIdx.x;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  const int64_t n = (int64_t)col.nz * col.ni;
  double m = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += VK_NT) {
    const double s = fabs(sol[col.off_s + e]);
    const double r = delta[col.off_s + e] / (s > 1e-30 ? s : 1e-30);
    m = r > m || r != r ? r : m;  // NaN propagates, as jnp.max does
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(m, off, 64);
    m = o > m || o != o ? o : m;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < VK_NT / 64; ++w) r = red[w] > r || red[w] != red[w] ? red[w] : r;
    out[(int64_t)c * stride] = r;
  }
}

int vk_launch_delta_max(const VkCol* cols, int ncols, const double* sol, const double* delta,
                        double* out, int64_t stride, hipStream_t st) {
  hipLaunchKernelGGL(k_delta_max, dim3(ncols), dim3(VK_NT), 0, st, cols, ncols, sol, delta, out,
                     stride);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}
