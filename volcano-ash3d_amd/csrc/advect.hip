// Limited donor-cell-upwind sweeps: advect_x, advect_y
// (src/AdvectionHorz_DCU.f90:68-305, :322-556) and advect_z / advect_sigma
// with settling and deposit (src/AdvectionVert_DCU.f90:100-323, :341-564).
//
// One thread owns one line (a row along the sweep direction) for NB grain
// bins and streams along it once: cell r is loaded, and the update of cell
// r-2 is formed from a register window, so concen is read once and written
// once, in place in ts0.  Bin-independent work per cell (dt/kappa, the
// interface speed of x/y sweeps) is done once for the NB bins.  Lanes map to
// the transverse index that is contiguous in memory (i for y and z lines),
// so every load and store of a wavefront is one coalesced 512-byte segment.
//
// The arithmetic reproduces the reference expression by expression (no
// contraction: the library is built with -ffp-contract=off), so results are
// bitwise those of the Fortran routines.
#include "common.hpp"

namespace ash3d {

namespace {

constexpr double EPS_THRESH = 1.0e-10;  // src/Ash3d_VariableModules.f90:273
constexpr int kBlock = 256;

// Limited Delta Q, src/AdvectionHorz_DCU.f90:186-226 (limiter is uniform
// across the launch, so the switch does not diverge).
__device__ inline double limited_dq(int lim, double dq, double dqu, double theta) {
  switch (lim) {
    case 1: return dq;                                      // Lax-Wendroff
    case 2: return dqu;                                     // Beam-Warming
    case 3: return dq * (0.5 * (1.0 + theta));              // Fromm
    case 4: return dq * fmax(0.0, fmin(1.0, theta));        // minmod
    case 5: return dq * fmax(fmax(0.0, fmin(1.0, 2.0 * theta)), fmin(2.0, theta));  // superbee
    case 6: return dq * fmax(0.0, fmin(fmin((1.0 + theta) / 2.0, 2.0), 2.0 * theta));  // MC
    default: return 0.0;
  }
}

// fss_I, the limited second-order flux correction at interface I
// (:179-231): dq0 = dq_I, dqm = dq_{I-1}, dqp = dq_{I+1}, u = usig_I,
// dtv = dt/kappa of cell I.
__device__ inline double limited_flux(int lim, double dqm, double dq0, double dqp, double u,
                                      double dtv) {
  if (fabs(dq0) <= EPS_THRESH) return 0.0;
  double dqu = 0.0, theta = 0.0;
  if (lim != 1) {
    dqu = (u > 0.0) ? dqm : dqp;
    if (fabs(dqu) > 3.0 * fabs(dq0) || fabs(dq0) < EPS_THRESH)
      theta = copysign(3.0, dqu * dq0);
    else
      theta = dqu / dq0;
  }
  const double ldq = limited_dq(lim, dq0, dqu, theta);
  const double aus = fabs(u);
  return 0.5 * aus * (1.0 - dtv * aus) * ldq;
}

template <int DIR, int NB, bool TOPO>
__global__ __launch_bounds__(kBlock) void k_advect(AdvArgs A) {
  const Dims d = A.d;
  const LineGeom g = A.g;
  const long long gid = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= (long long)g.Na * g.Nb) return;
  const int a = (int)(gid % g.Na) + 1;
  const int b = (int)(gid / g.Na) + 1;
  const int n0 = (int)blockIdx.y * NB;
  const int N = g.N, rmin = g.rmin, rmax = g.rmax;
  const long long sL = line_stride(d, DIR);
  const long long base = line_base(d, DIR, a, b);
  const bool line_on = a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  double *c = A.c0;

  bool exists[NB], on[NB];
  bool any_on = false;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int n = n0 + u;
    exists[u] = n < d.ns;
    on[u] = exists[u] && line_on && A.aloft.on(n);
    any_on = any_on || on[u];
  }
  // outflow entry of this line: outflow(a,b,n) with the transverse extents
  const long long ofs0 = (long long)(a + 1) + (long long)(g.Na + 4) * (long long)(b + 1);
  const long long ofs_bin = (long long)(g.Na + 4) * (g.Nb + 4);

  if (any_on) {
    const int lo = rmin - 2, hi = rmax + 2;  // cells gathered by the reference
    const int ncells = rmax - rmin + 1;
    const double dt = A.dt;
    const int lim = A.limiter;
    // register window (index 1 = r-1, 2 = r-2)
    double q0[NB], q1[NB], q2[NB], dq0[NB], dq1[NB], dq2[NB], fsr2[NB], fss2[NB], tp[NB];
    double vb1[NB], ub0[NB], ub1[NB], ub2[NB];  // z: per-bin velocity and usig
    double vel1 = 0.0, sig1 = 0.0, u0 = 0.0, u1 = 0.0, u2 = 0.0, dtv1 = 0.0, dtv2 = 0.0, D1 = 0.0;
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      q0[u] = q1[u] = q2[u] = dq0[u] = dq1[u] = dq2[u] = fsr2[u] = fss2[u] = tp[u] = 0.0;
      vb1[u] = ub0[u] = ub1[u] = ub2[u] = 0.0;
    }
    for (int r = -1; r <= N + 3; ++r) {
      const bool ld = (r >= lo && r <= hi);
      const long long cell = base + sL * (long long)(r + 1);
      // ---- gather cell r (:143-146 / :171-176)
      double vel0 = 0.0, sig0 = 0.0, kap0 = 1.0, vzh0 = 0.0, D0 = 0.0;
      if (ld) {
        sig0 = A.sig[cell];
        kap0 = A.kap[cell];
        if (DIR == 2) vzh0 = A.vel[cell] + A.vh[cell];
        else vel0 = A.vel[cell];
        // DelDxonD_cc(r,j) for x lines, DelDyonD_cc(i,r) for y lines (:163-166, :414-417)
        if (TOPO && r >= rmin - 1 && r <= rmax + 1)
          D0 = (DIR == 0) ? A.deld[(long long)(r + 1) + (long long)(d.nx + 4) * (a + 1)]
                          : A.deld[(long long)(a + 1) + (long long)(d.nx + 4) * (r + 1)];
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        q0[u] = 0.0;
        if (ld && exists[u]) q0[u] = c[cell + d.ncell * (long long)(n0 + u)];
      }
      const bool uin = (r >= rmin - 1 && r <= rmax + 1);
      // dt/kappa (:155-156) -- bin independent
      const double dtv0 = uin ? dt / kap0 : 0.0;
      // usig (:159-162)
      if (DIR != 2) u0 = uin ? 0.5 * (vel1 + vel0) * sig0 : 0.0;
      double vb0[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        vb0[u] = 0.0;
        if (DIR == 2) {
          if (ld && exists[u]) vb0[u] = vzh0 + A.vf[cell + d.ncell * (long long)(n0 + u)];
          ub0[u] = uin ? 0.5 * (vb1[u] + vb0[u]) * sig0 : 0.0;
        }
        dq0[u] = q0[u] - q1[u];  // (:168-169), used only for r >= rmin-1
      }
      // ---- interface I = r-1
      const int I = r - 1;
      const bool fss_in = (lim != 0) && I >= rmin && I <= rmax + 1;
      const bool fs_in = I >= rmin - 1 && I <= rmax + 1;
      // ---- cell c = r-2
      const int cc = r - 2;
      const bool upd_in = cc >= rmin - 1 && cc <= rmax + 1;
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const double uI = (DIR == 2) ? ub1[u] : u1;
        const double uIm = (DIR == 2) ? ub2[u] : u2;
        const double uIp = (DIR == 2) ? ub0[u] : u0;
        double fssI = 0.0, fsrI = 0.0, fslI = 0.0, TI = 0.0;
        if (fss_in) fssI = limited_flux(lim, dq2[u], dq1[u], dq0[u], uI, dtv1);
        if (fs_in) {
          // (:241-249)
          fsrI = fmax(0.0, uI) * dq1[u];
          fslI = fmin(0.0, uI) * dq1[u];
          const double divu_p = (fmax(0.0, uI) - fmax(0.0, uIm));
          const double divu_m = (fmin(0.0, uIp) - fmin(0.0, uI));
          fsrI = fsrI + q1[u] * divu_m;
          fslI = fslI + q2[u] * divu_p;
          if (TOPO) {
            // topographic term (:252-257): T_I on the right fluctuation of
            // interface I, T_{I-1} on the left one
            TI = 0.5 * vel1 * (sig1 + sig0) * q1[u] * D1;
            fsrI = fsrI + TI;
            fslI = fslI - tp[u];
          }
        }
        if (upd_in && on[u]) {
          // (:264-281)
          const double upd = -dtv2 * (fslI + fsr2[u] + fssI - fss2[u]);
          const long long cellc = base + sL * (long long)(cc + 1) + d.ncell * (long long)(n0 + u);
          if (cc >= rmin && cc <= rmax) c[cellc] = q2[u] + upd;
          const long long of = ofs0 + ofs_bin * (long long)(n0 + u);
          if (DIR == 0) {
            // (:289-294) x outflow, skipped when periodic
            if (cc == 0 && rmin == 1 && !A.periodic) A.of1[of] = A.of1[of] + upd;
            if (cc == N + 1 && rmax == N && !A.periodic) A.of2[of] = A.of2[of] + upd;
          } else if (DIR == 1) {
            // (:540-545)
            if (cc == 0 && rmin == 1) A.of1[of] = A.of1[of] + upd;
            if (cc == N + 1 && rmax == N) A.of2[of] = A.of2[of] + upd;
          } else {
            // deposit and top outflow (src/AdvectionVert_DCU.f90:306-311);
            // the top uses update_cc(ncells+1)
            if (cc == 0 && rmin == 1) A.of1[of] = A.of1[of] + upd;
            if (cc == ncells + 1 && rmax == N) A.of2[of] = A.of2[of] + upd;
          }
        }
        // shift the per-bin window
        fsr2[u] = fsrI;
        fss2[u] = fssI;
        tp[u] = TI;
        q2[u] = q1[u];
        q1[u] = q0[u];
        dq2[u] = dq1[u];
        dq1[u] = dq0[u];
        if (DIR == 2) {
          ub2[u] = ub1[u];
          ub1[u] = ub0[u];
          vb1[u] = vb0[u];
        }
      }
      // shift the bin-independent window
      u2 = u1;
      u1 = u0;
      dtv2 = dtv1;
      dtv1 = dtv0;
      vel1 = vel0;
      sig1 = sig0;
      D1 = D0;
    }
  }
  // cells the reference leaves at zero: inactive bins and lines, and cells
  // outside [rmin, rmax] (zero fill of ts1 then copy, :122, :302-303)
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    if (!exists[u]) continue;
    const long long bofs = d.ncell * (long long)(n0 + u);
    if (!on[u]) {
      for (int r = 1; r <= N; ++r) c[base + sL * (long long)(r + 1) + bofs] = 0.0;
    } else {
      for (int r = 1; r < rmin; ++r) c[base + sL * (long long)(r + 1) + bofs] = 0.0;
      for (int r = rmax + 1; r <= N; ++r) c[base + sL * (long long)(r + 1) + bofs] = 0.0;
    }
    if (DIR == 2) {
      // DepositGranularity = outflow_xy1 (src/AdvectionVert_DCU.f90:321)
      const long long of = ofs0 + ofs_bin * (long long)(n0 + u);
      A.dep[(long long)(a - 1) + (long long)d.nx * ((long long)(b - 1) + (long long)d.ny * (n0 + u))] =
          A.of1[of];
    }
  }
}

template <int DIR, int NB>
hipError_t launch_dir(hipStream_t s, const AdvArgs &a) {
  const long long lines = (long long)a.g.Na * a.g.Nb;
  dim3 grid((unsigned)((lines + kBlock - 1) / kBlock), (unsigned)((a.d.ns + NB - 1) / NB));
  if (DIR != 2 && a.deld != nullptr)
    hipLaunchKernelGGL((k_advect<DIR, NB, true>), grid, dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_advect<DIR, NB, false>), grid, dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <int NB>
hipError_t launch_nb(hipStream_t s, const AdvArgs &a) {
  switch (a.g.dir) {
    case 0: return launch_dir<0, NB>(s, a);
    case 1: return launch_dir<1, NB>(s, a);
    default: return launch_dir<2, NB>(s, a);
  }
}

}  // namespace

int pick_bins_per_thread(int ns) {
  if (ns >= 8) return 4;
  if (ns >= 4) return 4;
  if (ns >= 2) return 2;
  return 1;
}

hipError_t launch_advect(hipStream_t s, const AdvArgs &a, int nb) {
  switch (nb) {
    case 1: return launch_nb<1>(s, a);
    case 2: return launch_nb<2>(s, a);
    case 4: return launch_nb<4>(s, a);
    default: return launch_nb<8>(s, a);
  }
}

}  // namespace ash3d
