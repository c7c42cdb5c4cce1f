// Per-step driver work that the reference does on the host between the
// hot-path calls, kept on the device so that the state never leaves HBM:
//   * wind time interpolation between two met steps
//     (src/MesoInterpolater.f90:203-230),
//   * the CFL and diffusion maxima of Adjust_DT (src/Adjust_dt.f90:96-285),
//   * the slab maxima of get_minmax_index (src/Calc_Mesh.f90:459-511).
// Maxima and minima are exact in any order, so the reductions (per-block
// partials, then a fixed-order host pass) reproduce the reference's loops
// bit for bit.
#include "common.hpp"

namespace ash3d {

namespace {

constexpr double MPS_2_KMPHR = 3.6;  // src/Ash3d_VariableModules.f90:286

// v_pd(1:nx,1:ny,1:nz) = (real(last) + real(next - last) * frac) * 3.6,
// the difference in FP32 as the _sp arrays give it.
__global__ void k_met_interp(Dims d, const float *__restrict__ lx, const float *__restrict__ ly,
                             const float *__restrict__ lz, const float *__restrict__ nx_,
                             const float *__restrict__ ny_, const float *__restrict__ nz_, double frac,
                             double *vx, double *vy, double *vz) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n3 = (long long)d.nx * d.ny * d.nz;
  if (m >= n3) return;
  const int i = (int)(m % d.nx) + 1, j = (int)((m / d.nx) % d.ny) + 1, k = (int)(m / ((long long)d.nx * d.ny)) + 1;
  const long long o = d.p3(i, j, k);
  const float dx = nx_[m] - lx[m], dy = ny_[m] - ly[m], dz = nz_[m] - lz[m];
  vx[o] = ((double)lx[m] + (double)dx * frac) * MPS_2_KMPHR;
  vy[o] = ((double)ly[m] + (double)dy * frac) * MPS_2_KMPHR;
  vz[o] = ((double)lz[m] + (double)dz * frac) * MPS_2_KMPHR;
}

__global__ void k_met_interp_vf(Dims d, const float *__restrict__ lf, const float *__restrict__ nf, double frac,
                                double *vf) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n3 = (long long)d.nx * d.ny * d.nz;
  if (m >= n3 * d.ns) return;
  const long long c = m % n3;
  const int n = (int)(m / n3);
  const int i = (int)(c % d.nx) + 1, j = (int)((c / d.nx) % d.ny) + 1, k = (int)(c / ((long long)d.nx * d.ny)) + 1;
  const float dv = nf[m] - lf[m];
  vf[d.p3(i, j, k) + d.ncell * (long long)n] = ((double)lf[m] + (double)dv * frac) * MPS_2_KMPHR;
}

constexpr int kRed = 256;

__device__ inline double block_max(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kRed / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Per-block partials of Adjust_DT over the interior cells:
// [0] vxmax_dx (lon-lat) or max|vx| (projected), [1] the same for y,
// [2] vzmax_dz, [3] min of the diffusion times (negated, so that all four
// reduce with max).  met != 0 reads the FP32 next met step (in m/s).
struct DtArgs {
  Dims d;
  int latlon, met, diffusion;
  const double *vx, *vy, *vz, *vf, *kap, *sx, *sy, *sz, *kx, *ky, *kz, *dz;
  const float *mx, *my, *mz, *mf;
  double *partial;  // [blocks][4]
};

__global__ __launch_bounds__(kRed) void k_dt_partials(DtArgs A) {
  __shared__ double sh[kRed];
  const Dims d = A.d;
  const long long n3 = (long long)d.nx * d.ny * d.nz;
  const long long kxs = d.nx + 2, kxys = (long long)(d.nx + 2) * (d.ny + 2);
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = -INFINITY;
  for (long long m = (long long)blockIdx.x * kRed + threadIdx.x; m < n3; m += (long long)gridDim.x * kRed) {
    const int i = (int)(m % d.nx) + 1, j = (int)((m / d.nx) % d.ny) + 1, k = (int)(m / ((long long)d.nx * d.ny)) + 1;
    const long long o = d.p3(i, j, k);
    const double kap = A.kap[o];
    if (A.met) {
      if (A.latlon) {
        double ms = fmin(A.sx[o], A.sx[o + 1]);
        r0 = fmax(r0, (double)fabsf(A.mx[m]) * MPS_2_KMPHR * ms / kap);
        ms = fmin(A.sy[o], A.sy[o + d.sy]);
        r1 = fmax(r1, (double)fabsf(A.my[m]) * MPS_2_KMPHR * ms / kap);
      } else {
        r0 = fmax(r0, (double)fabsf(A.mx[m]));
        r1 = fmax(r1, (double)fabsf(A.my[m]));
      }
      float vfm = 0.0f;
      for (int n = 0; n < d.ns; ++n) vfm = fmaxf(vfm, fabsf(A.mf[m + n3 * n]));
      r2 = fmax(r2, (double)(fabsf(A.mz[m]) + vfm) * MPS_2_KMPHR / A.dz[k - 1]);
    } else {
      if (A.latlon) {
        double ms = fmin(A.sx[o], A.sx[o + 1]);
        r0 = fmax(r0, fabs(A.vx[o]) * ms / kap);
        ms = fmin(A.sy[o], A.sy[o + d.sy]);
        r1 = fmax(r1, fabs(A.vy[o]) * ms / kap);
      } else {
        r0 = fmax(r0, fabs(A.vx[o]));
        r1 = fmax(r1, fabs(A.vy[o]));
      }
      double vfm = 0.0;
      for (int n = 0; n < d.ns; ++n) vfm = fmax(vfm, fabs(A.vf[o + d.ncell * n]));
      r2 = fmax(r2, (fabs(A.vz[o]) + vfm) / A.dz[k - 1]);
    }
    if (A.diffusion) {
      const long long ok = (long long)i + kxs * j + kxys * k;
      double rr = kap / fmin(A.sx[o], A.sx[o + 1]);
      const double t1 = (rr * rr) / A.kx[ok];
      rr = kap / fmin(A.sy[o], A.sy[o + d.sy]);
      const double t2 = (rr * rr) / A.ky[ok];
      rr = kap / fmin(A.sz[o], A.sz[o + d.sz]);
      const double t3 = (rr * rr) / A.kz[ok];
      r3 = fmax(r3, -(t1 + t2 + t3));
    }
  }
  r0 = block_max(r0, sh);
  r1 = block_max(r1, sh);
  r2 = block_max(r2, sh);
  r3 = block_max(r3, sh);
  if (threadIdx.x == 0) {
    double *p = A.partial + 4 * (long long)blockIdx.x;
    p[0] = r0; p[1] = r1; p[2] = r2; p[3] = r3;
  }
}

// Slab maxima of get_minmax_index: px[i] = max over (j,k,n), py[j] over
// (i,k,n), pz[k] over (i,j,n) of concen(1:nx,1:ny,1:nz,1:ns,ts0), clamped
// at zero (the reference compares them with a positive threshold) and kept
// as the bit patterns of non-negative doubles, whose unsigned order is their
// numeric order.  One workgroup per (k, bin) plane.
__global__ __launch_bounds__(kRed) void k_slab_max(Dims d, const double *__restrict__ c0,
                                                   unsigned long long *px, unsigned long long *py,
                                                   unsigned long long *pz) {
  extern __shared__ unsigned long long ly[];  // ny entries, then one for the plane
  const int k = (int)(blockIdx.x % (unsigned)d.nz) + 1;
  const int n = (int)(blockIdx.x / (unsigned)d.nz);
  for (int j = threadIdx.x; j <= d.ny; j += kRed) ly[j] = 0ull;
  __syncthreads();
  const double *c = c0 + d.ncell * (long long)n;
  const int lane = threadIdx.x & 63;
  for (int i0 = 0; i0 < d.nx; i0 += kRed) {
    const int i = i0 + threadIdx.x + 1;
    double mi = 0.0;
    for (int j = 1; j <= d.ny; ++j) {
      const double v = (i <= d.nx) ? fmax(c[d.p3(i, j, k)], 0.0) : 0.0;
      mi = fmax(mi, v);
      // max over the row segment of this wave
      double w = v;
      for (int s = 32; s > 0; s >>= 1) w = fmax(w, __shfl_xor(w, s, 64));
      if (lane == 0) atomicMax(&ly[j - 1], (unsigned long long)__double_as_longlong(w));
    }
    if (i <= d.nx) atomicMax(&px[i - 1], (unsigned long long)__double_as_longlong(mi));
    atomicMax(&ly[d.ny], (unsigned long long)__double_as_longlong(mi));
  }
  __syncthreads();
  for (int j = threadIdx.x; j < d.ny; j += kRed) atomicMax(&py[j], ly[j]);
  if (threadIdx.x == 0) atomicMax(&pz[k - 1], ly[d.ny]);
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0ull;
}

}  // namespace

hipError_t launch_met_interp(hipStream_t s, const Dims &d, const float *const last[4], const float *const next[4],
                             double frac, bool with_vf, double *vx, double *vy, double *vz, double *vf) {
  const long long n3 = (long long)d.nx * d.ny * d.nz;
  hipLaunchKernelGGL(k_met_interp, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, d, last[0], last[1], last[2],
                     next[0], next[1], next[2], frac, vx, vy, vz);
  if (with_vf)
    hipLaunchKernelGGL(k_met_interp_vf, dim3((unsigned)((n3 * d.ns + 255) / 256)), dim3(256), 0, s, d, last[3],
                       next[3], frac, vf);
  return hipGetLastError();
}

int dt_partial_blocks(const Dims &d) {
  const long long n3 = (long long)d.nx * d.ny * d.nz;
  const long long b = (n3 + kRed - 1) / kRed;
  return (int)(b < 2048 ? b : 2048);
}

hipError_t launch_dt_partials(hipStream_t s, const Dims &d, int latlon, int met, int diffusion, const double *vx,
                              const double *vy, const double *vz, const double *vf, const double *kap,
                              const double *sx, const double *sy, const double *sz, const double *kx,
                              const double *ky, const double *kz, const double *dz, const float *const met_next[4],
                              double *partial) {
  DtArgs A{d, latlon, met, diffusion, vx, vy, vz, vf, kap, sx, sy, sz, kx, ky, kz, dz,
           met_next[0], met_next[1], met_next[2], met_next[3], partial};
  hipLaunchKernelGGL(k_dt_partials, dim3(dt_partial_blocks(d)), dim3(kRed), 0, s, A);
  return hipGetLastError();
}

hipError_t launch_slab_max(hipStream_t s, const Dims &d, const double *c0, unsigned long long *prof) {
  const int np = d.nx + d.ny + d.nz;
  hipLaunchKernelGGL(k_zero_u64, dim3((np + 255) / 256), dim3(256), 0, s, prof, np);
  hipLaunchKernelGGL(k_slab_max, dim3((unsigned)(d.nz * d.ns)), dim3(kRed), sizeof(unsigned long long) * (d.ny + 1), s,
                     d, c0, prof, prof + d.nx, prof + d.nx + d.ny);
  return hipGetLastError();
}

}  // namespace ash3d
