// Per-bin mass budget on the device: the sums of Calc_AshVol_Aloft,
// Calc_AshVol_Deposit and Calc_AshVol_Outflow (src/Output_Vars.f90:1255-1293,
// :1375-1404, :1419-1459) before their conversion to volume, which
// TimeStepTotals (src/TimeStepTotals.f90:17-88) evaluates every step:
//   aloft(n)   = sum concen(1:nx,1:ny,1:nz,n,ts0) * kappa(1:nx,1:ny,1:nz)
//   deposit(n) = sum DepositGranularity(1:nx,1:ny,n) * kappa(1:nx,1:ny,0)
//   outflow(n) = sum over the yz1, yz2, xz1, xz2 and xy2 faces of the
//                outflow times the kappa of the ghost cell it belongs to.
// The reduction is deterministic: a fixed number of partial sums per bin,
// each over a fixed set of cells in a fixed order, then a fixed-order sum
// of the partials.
#include "common.hpp"

namespace ash3d {

namespace {

constexpr int kThreads = 256;
constexpr int kParts = 128;

__device__ inline double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

struct TotArgs {
  Dims d;
  const double *c0, *kap, *dep;
  const double *yz1, *yz2, *xz1, *xz2, *xy2;
  double *partial;  // [ns][kParts][3]
};

__global__ __launch_bounds__(kThreads) void k_mass_partials(TotArgs A) {
  __shared__ double sh[kThreads];
  const Dims d = A.d;
  const int n = blockIdx.y, p = blockIdx.x;
  const long long stride = (long long)kParts * kThreads;
  const long long t0 = (long long)p * kThreads + threadIdx.x;
  const double *c = A.c0 + d.ncell * (long long)n;
  // aloft
  double a = 0.0;
  const long long ncells = (long long)d.nx * d.ny * d.nz;
  for (long long t = t0; t < ncells; t += stride) {
    const int i = (int)(t % d.nx) + 1;
    const int j = (int)((t / d.nx) % d.ny) + 1;
    const int k = (int)(t / ((long long)d.nx * d.ny)) + 1;
    const long long o = d.p3(i, j, k);
    a += c[o] * A.kap[o];
  }
  // deposit
  double b = 0.0;
  const long long nxy = (long long)d.nx * d.ny;
  for (long long t = t0; t < nxy; t += stride) {
    const int i = (int)(t % d.nx) + 1, j = (int)(t / d.nx) + 1;
    b += A.dep[(long long)(i - 1) + (long long)d.nx * ((long long)(j - 1) + (long long)d.ny * n)] *
         A.kap[d.p3(i, j, 0)];
  }
  // outflow through five faces
  double o = 0.0;
  const long long nyz = (long long)d.ny * d.nz, nxz = (long long)d.nx * d.nz;
  const long long byz = (long long)(d.ny + 4) * (d.nz + 4) * n;
  const long long bxz = (long long)(d.nx + 4) * (d.nz + 4) * n;
  const long long bxy = (long long)(d.nx + 4) * (d.ny + 4) * n;
  for (long long t = t0; t < nyz; t += stride) {
    const int j = (int)(t % d.ny) + 1, k = (int)(t / d.ny) + 1;
    const long long f = (long long)(j + 1) + (long long)(d.ny + 4) * (k + 1) + byz;
    o += A.yz1[f] * A.kap[d.p3(0, j, k)];
    o += A.yz2[f] * A.kap[d.p3(d.nx + 1, j, k)];
  }
  for (long long t = t0; t < nxz; t += stride) {
    const int i = (int)(t % d.nx) + 1, k = (int)(t / d.nx) + 1;
    const long long f = (long long)(i + 1) + (long long)(d.nx + 4) * (k + 1) + bxz;
    o += A.xz1[f] * A.kap[d.p3(i, 0, k)];
    o += A.xz2[f] * A.kap[d.p3(i, d.ny + 1, k)];
  }
  for (long long t = t0; t < nxy; t += stride) {
    const int i = (int)(t % d.nx) + 1, j = (int)(t / d.nx) + 1;
    const long long f = (long long)(i + 1) + (long long)(d.nx + 4) * (j + 1) + bxy;
    o += A.xy2[f] * A.kap[d.p3(i, j, d.nz + 1)];
  }
  a = block_sum(a, sh);
  b = block_sum(b, sh);
  o = block_sum(o, sh);
  if (threadIdx.x == 0) {
    double *w = A.partial + ((long long)n * kParts + p) * 3;
    w[0] = a;
    w[1] = b;
    w[2] = o;
  }
}

__global__ void k_mass_final(const double *partial, int ns, double *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * ns) return;
  const int n = t / 3, q = t % 3;
  double s = 0.0;
  for (int p = 0; p < kParts; ++p) s += partial[((long long)n * kParts + p) * 3 + q];
  out[(long long)q * ns + n] = s;
}

}  // namespace

long long mass_partials_size(const Dims &d) { return (long long)d.ns * kParts * 3; }

hipError_t launch_mass_totals(hipStream_t s, const Dims &d, const double *c0, const double *kap,
                              const double *dep, const double *yz1, const double *yz2, const double *xz1,
                              const double *xz2, const double *xy2, double *partial, double *out) {
  TotArgs A{d, c0, kap, dep, yz1, yz2, xz1, xz2, xy2, partial};
  hipLaunchKernelGGL(k_mass_partials, dim3(kParts, d.ns), dim3(kThreads), 0, s, A);
  hipLaunchKernelGGL(k_mass_final, dim3((3 * d.ns + 255) / 256), dim3(256), 0, s, partial, d.ns, out);
  return hipGetLastError();
}

}  // namespace ash3d
