// advect_x (src/AdvectionHorz_DCU.f90:68-305) with one workgroup per x row.
//
// x rows are contiguous in memory, so lanes map to cells of the row: every
// load and store of concen is a coalesced row segment.  The row of one bin
// is staged through LDS and the update is built in three LDS phases (dq;
// fss, fs_l, fs_r; the cell update), each neighbour value read from LDS.
// The workgroup loops over all grain bins, so the bin-independent terms of
// the row (interface speeds usig, dt/kappa, the divergence corrections, the
// topographic factors) are computed once per row.  All global reads of a
// row happen before the barrier that precedes its in-place writes, and no
// other workgroup touches the row.
//
// Arithmetic: the reference expressions in the reference order, no
// contraction -- bitwise equal to the streaming kernel and the oracle.
#include "common.hpp"
#include "dcu.hpp"

namespace ash3d {

namespace {

template <int CPT, bool TOPO>
__global__ __launch_bounds__(256) void k_advect_x_rows(AdvArgs A) {
  extern __shared__ double smem[];
  const Dims d = A.d;
  const LineGeom g = A.g;
  const int nx = d.nx;
  const int W = nx + 4;  // cells -1..nx+2 at LDS index c+1
  double *s_q = smem;
  double *s_dq = smem + W;
  double *s_fss = smem + 2 * W;
  double *s_fsl = smem + 3 * W;
  double *s_tf = smem + 4 * W;
  double *s_D = smem + 5 * W;

  const int j = (int)(blockIdx.x % (unsigned)d.ny) + 1;
  const int k = (int)(blockIdx.x / (unsigned)d.ny) + 1;
  const int rmin = g.rmin, rmax = g.rmax;
  const bool row_on = j >= g.a0 && j <= g.a1 && k >= g.b0 && k <= g.b1;
  const long long base = d.p3(-1, j, k);  // cell -1 of the row
  const int t = threadIdx.x, BT = blockDim.x;
  const double *c0 = A.c0;
  double *co = A.cout;  // == c0 for the in-place sweep

  int cc[CPT];
  bool valid[CPT];
#pragma unroll
  for (int m = 0; m < CPT; ++m) {
    cc[m] = -1 + t + BT * m;
    valid[m] = cc[m] <= nx + 2;
  }

  bool any_on = false;
  if (row_on)
    for (int n = 0; n < d.ns; ++n) any_on = any_on || A.aloft.on(n);

  // ---- bin-independent terms of the row
  double u[CPT], dtv[CPT], divp[CPT], divm[CPT], tf[CPT], D[CPT], tfm1[CPT], Dm1[CPT];
  if (any_on) {
    double vel[CPT], sig[CPT], kap[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int c = cc[m];
      vel[m] = 0.0; sig[m] = 0.0; kap[m] = 1.0; D[m] = 0.0;
      if (valid[m] && c >= rmin - 2 && c <= rmax + 2) {
        const long long o = base + (c + 1);
        vel[m] = A.vel[o];
        sig[m] = A.sig[o];
        kap[m] = A.kap[o];
        if (TOPO && c >= rmin - 1 && c <= rmax + 1)
          D[m] = A.deld[(long long)(c + 1) + (long long)(nx + 4) * (j + 1)];
      }
      if (valid[m]) {
        s_fss[c + 1] = vel[m];  // scratch: vel
        s_fsl[c + 1] = sig[m];  // scratch: sigma
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int c = cc[m], p = c + 1;
      u[m] = 0.0; dtv[m] = 0.0; tf[m] = 0.0;
      if (!valid[m]) continue;
      const bool uin = c >= rmin - 1 && c <= rmax + 1;
      const double velm1 = p >= 1 ? s_fss[p - 1] : 0.0;
      const double sigp1 = p + 1 < W ? s_fsl[p + 1] : 0.0;
      // usig (:159-162), dt/kappa (:155-156)
      u[m] = uin ? 0.5 * (velm1 + vel[m]) * sig[m] : 0.0;
      dtv[m] = uin ? A.dt / kap[m] : 0.0;
      // topographic factor 0.5*vel_I*(sig_I+sig_{I+1}) (:252-254)
      if (TOPO) tf[m] = 0.5 * vel[m] * (sig[m] + sigp1);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      if (!valid[m]) continue;
      const int p = cc[m] + 1;
      s_q[p] = u[m];  // scratch: usig
      if (TOPO) { s_tf[p] = tf[m]; s_D[p] = D[m]; }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      divp[m] = divm[m] = tfm1[m] = Dm1[m] = 0.0;
      if (!valid[m]) continue;
      const int p = cc[m] + 1;
      const double um1 = p >= 1 ? s_q[p - 1] : 0.0;
      const double up1 = p + 1 < W ? s_q[p + 1] : 0.0;
      // divergence corrections (:245-246)
      divp[m] = (fmax(0.0, u[m]) - fmax(0.0, um1));
      divm[m] = (fmin(0.0, up1) - fmin(0.0, u[m]));
      if (TOPO && p >= 1) { tfm1[m] = s_tf[p - 1]; Dm1[m] = s_D[p - 1]; }
    }
    __syncthreads();
  }

  const int lim = A.limiter;
  for (int n = 0; n < d.ns; ++n) {
    const long long bofs = d.ncell * (long long)n;
    const bool on = row_on && A.aloft.on(n);
    if (!on) {
      // zero fill of ts1 then copy (:122, :302-303)
#pragma unroll
      for (int m = 0; m < CPT; ++m)
        if (cc[m] >= 1 && cc[m] <= nx) co[base + (cc[m] + 1) + bofs] = 0.0;
      continue;
    }
    double q[CPT], qm1[CPT], dq[CPT], fss[CPT], fsr[CPT];
    // phase 1: gather the row (:143)
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int c = cc[m];
      q[m] = 0.0;
      if (valid[m] && c >= rmin - 2 && c <= rmax + 2) q[m] = c0[base + (c + 1) + bofs];
      if (valid[m]) s_q[c + 1] = q[m];
    }
    __syncthreads();
    // phase 2: dq (:168-169)
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int p = cc[m] + 1;
      qm1[m] = (valid[m] && p >= 1) ? s_q[p - 1] : 0.0;
      dq[m] = q[m] - qm1[m];
      if (valid[m]) s_dq[p] = dq[m];
    }
    __syncthreads();
    // phase 3: limited flux and fluctuations at interface I = c (:178-259)
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int c = cc[m], p = c + 1;
      fss[m] = 0.0; fsr[m] = 0.0;
      double fsl = 0.0;
      if (valid[m]) {
        if (lim != 0 && c >= rmin && c <= rmax + 1)
          fss[m] = limited_flux(lim, s_dq[p - 1], dq[m], s_dq[p + 1], u[m], dtv[m]);
        if (c >= rmin - 1 && c <= rmax + 1) {
          fsr[m] = fmax(0.0, u[m]) * dq[m];
          fsl = fmin(0.0, u[m]) * dq[m];
          fsr[m] = fsr[m] + q[m] * divm[m];
          fsl = fsl + qm1[m] * divp[m];
          if (TOPO) {
            fsr[m] = fsr[m] + tf[m] * q[m] * D[m];
            fsl = fsl - tfm1[m] * qm1[m] * Dm1[m];
          }
        }
        s_fss[p] = fss[m];
        s_fsl[p] = fsl;
      }
    }
    __syncthreads();
    // phase 4: cell update (:264-281), interior write and outflow (:285-294)
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int c = cc[m], p = c + 1;
      if (!valid[m]) continue;
      if (c >= rmin - 1 && c <= rmax + 1) {
        const double upd = -dtv[m] * (s_fsl[p + 1] + fsr[m] + s_fss[p + 1] - fss[m]);
        if (c >= rmin && c <= rmax) co[base + p + bofs] = q[m] + upd;
        const long long of = (long long)(j + 1) + (long long)(d.ny + 4) * ((long long)(k + 1) + (long long)(d.nz + 4) * n);
        if (c == 0 && rmin == 1 && !A.periodic) A.of1[of] = A.of1[of] + upd;
        if (c == nx + 1 && rmax == nx && !A.periodic) A.of2[of] = A.of2[of] + upd;
      }
      if (c >= 1 && c <= nx && (c < rmin || c > rmax)) co[base + p + bofs] = 0.0;
    }
  }
}

template <int CPT>
hipError_t launch_cpt(hipStream_t s, const AdvArgs &a, int bt) {
  const size_t shmem = sizeof(double) * 6 * (size_t)(a.d.nx + 4);
  dim3 grid((unsigned)((long long)a.d.ny * a.d.nz));
  const void *kern = a.deld != nullptr ? (const void *)k_advect_x_rows<CPT, true>
                                       : (const void *)k_advect_x_rows<CPT, false>;
  if (shmem > 65536) {
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess) return e;
  }
  if (a.deld != nullptr)
    hipLaunchKernelGGL((k_advect_x_rows<CPT, true>), grid, dim3(bt), shmem, s, a);
  else
    hipLaunchKernelGGL((k_advect_x_rows<CPT, false>), grid, dim3(bt), shmem, s, a);
  return hipGetLastError();
}

}  // namespace

// Rows of up to 8*256-4 cells; returns hipErrorInvalidValue beyond that so
// the caller falls back to the streaming kernel.
hipError_t launch_advect_x_rows(hipStream_t s, const AdvArgs &a) {
  const int cells = a.d.nx + 4;
  int cpt = (cells + 255) / 256;
  if (cpt > 8) return hipErrorInvalidValue;
  if (cpt == 3) cpt = 4;
  if (cpt > 4) cpt = 8;
  int bt = (cells + cpt - 1) / cpt;
  bt = ((bt + 63) / 64) * 64;
  switch (cpt) {
    case 1: return launch_cpt<1>(s, a, bt);
    case 2: return launch_cpt<2>(s, a, bt);
    case 4: return launch_cpt<4>(s, a, bt);
    default: return launch_cpt<8>(s, a, bt);
  }
}

}  // namespace ash3d
