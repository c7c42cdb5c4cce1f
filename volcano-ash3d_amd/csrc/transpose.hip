// Batched x<->y plane transpose through LDS tiles.  The x-direction line
// solves run on a transposed copy of the state (x lines become y lines, so
// one lane per line reads and writes coalesced rows), then the result is
// transposed back; both transposes read and write 32x32 tiles with
// coalesced 256-byte row segments.
#include "common.hpp"

namespace ash3d {

namespace {

constexpr int kT = 32;

// dst[p*dps + b + B*a] = src[p*sps + a + A*b] for a < A, b < B, plane p.
__global__ __launch_bounds__(256) void k_transpose(const double *__restrict__ src, double *__restrict__ dst,
                                                   int A, int B, long long sps, long long dps,
                                                   const int *skip_if_zero) {
  __shared__ double tile[kT][kT + 1];
  if (skip_if_zero && *skip_if_zero == 0) return;  // uniform
  const long long p = blockIdx.z;
  const int a0 = blockIdx.x * kT, b0 = blockIdx.y * kT;
  const double *s = src + p * sps;
  double *o = dst + p * dps;
  for (int r = threadIdx.y; r < kT; r += 8) {
    const int a = a0 + threadIdx.x, b = b0 + r;
    if (a < A && b < B) tile[r][threadIdx.x] = s[(long long)a + (long long)A * b];
  }
  __syncthreads();
  for (int r = threadIdx.y; r < kT; r += 8) {
    const int b = b0 + threadIdx.x, a = a0 + r;
    if (a < A && b < B) o[(long long)b + (long long)B * a] = tile[threadIdx.x][r];
  }
}

}  // namespace

hipError_t launch_transpose_xy(hipStream_t s, const double *src, double *dst, int A, int B, long long nplanes,
                               const int *skip_if_zero) {
  const long long ps = (long long)A * B;
  dim3 grid((unsigned)((A + kT - 1) / kT), (unsigned)((B + kT - 1) / kT), (unsigned)nplanes);
  hipLaunchKernelGGL(k_transpose, grid, dim3(kT, 8), 0, s, src, dst, A, B, ps, ps, skip_if_zero);
  return hipGetLastError();
}

}  // namespace ash3d
