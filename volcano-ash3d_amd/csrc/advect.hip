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
#include <algorithm>

#include "common.hpp"
#include "dcu.hpp"

namespace ash3d {

namespace {

constexpr int kBlock = 256;

// Block = 64 lanes (consecutive transverse index a, i.e. coalesced for y
// and z lines) x W segments of the same lines.  Segment w emits the cell
// updates of cc in [lo, hi] of the line (cc = 0 and N+1 are the ghost
// updates that feed the outflow/deposit); it streams cells lo-2 .. hi+2.
// The two cells below lo and above hi are also written by the neighbouring
// segments, so they are read into registers before a workgroup barrier and
// never re-read from memory.
// GM: the bin groups of a tile in one workgroup, whole lines per wave (no
// halos to exchange); otherwise line segments per wave.
template <int DIR, int NB, bool TOPO, bool GM>
__device__ __forceinline__ void advect_body(const AdvArgs &A, const long long vb, const long long vn) {
  constexpr int PF = (NB >= 2) ? 2 : 4;  // cells in flight
  const Dims d = A.d;
  const LineGeom g = A.g;
  const int ntile = (g.Na + 63) / 64;
  // A.gy > 0: the waves of a workgroup take gy bin groups of the same lines
  // (whole lines, so the per-cell fields are read by all of them at once)
  const int G = (d.ns + NB - 1) / NB;
  constexpr bool gmode = GM;
  const TileGroup tg = xcd_tile_group(vb, gmode ? (G + A.gy - 1) / A.gy : G, vn);
  if (tg.tile >= (long long)ntile * g.Nb) return;  // padding workgroup (uniform)
  const int a = (int)(tg.tile % ntile) * 64 + (int)threadIdx.x + 1;
  const int b = (int)(tg.tile / ntile) + 1;
  const bool lane_ok = a <= g.Na;
  const int W = gmode ? 1 : blockDim.y, w = gmode ? 0 : threadIdx.y;
  const int n0 = (gmode ? tg.group * A.gy + (int)threadIdx.y : tg.group) * NB;
  const int N = g.N, rmin = g.rmin, rmax = g.rmax;
  const int seg = (N + 2 + W - 1) / W;
  const int lo = w * seg, hi = min(N + 1, lo + seg - 1);
  const long long sL = line_stride(d, DIR);
  const long long base = lane_ok ? line_base(d, DIR, a, b) : 0;
  const bool line_on = lane_ok && a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  const double *c = A.c0;
  double *co = A.cout;  // == c0 for the in-place sweep
  double *junk = A.junk + (threadIdx.x + 64 * threadIdx.y + 256 * (vb % (kJunk / 256)));

  bool exists[NB], on[NB];
  long long bofs[NB];  // bin offset (an existing bin for the others)
  bool any_on = false;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int n = n0 + u;
    exists[u] = lane_ok && n < d.ns;
    on[u] = exists[u] && line_on && A.aloft.on(n);
    any_on = any_on || on[u];
    bofs[u] = d.ncell * (long long)min(n, d.ns - 1);
  }
  // outflow entry of this line: outflow(a,b,n) with the transverse extents
  const long long ofs0 = (long long)(a + 1) + (long long)(g.Na + 4) * (long long)(b + 1);
  const long long ofs_bin = (long long)(g.Na + 4) * (g.Nb + 4);
  const int ldlo = rmin - 2, ldhi = rmax + 2;  // cells gathered by the reference

  // halo cells shared with the neighbouring segments (four named arrays so
  // that they stay in registers)
  // (whole lines: the ring reads those cells itself, nobody writes them).
  // The halo values wait in LDS (4*NB per thread, element-major), read back
  // in the four iterations that need them.
  extern __shared__ double s_halo[];
  const int tid = (int)threadIdx.x + 64 * (int)threadIdx.y, nthr = 64 * W;
  if constexpr (!GM) {
    auto halo = [&](int r, int u) {
      const bool qin = !A.zg || (r >= 1 && r <= N);
      return (any_on && exists[u] && qin && r >= ldlo && r <= ldhi) ? c[base + sL * (long long)(r + 1) + bofs[u]]
                                                                      : 0.0;
    };
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      s_halo[(4 * u + 0) * nthr + tid] = halo(lo - 2, u);
      s_halo[(4 * u + 1) * nthr + tid] = halo(lo - 1, u);
      s_halo[(4 * u + 2) * nthr + tid] = halo(hi + 1, u);
      s_halo[(4 * u + 3) * nthr + tid] = halo(hi + 2, u);
    }
    if (W > 1) __syncthreads();
  }

  // outflow increments of the line ends, applied after the sweep
  double upd_lo[NB], upd_hi[NB];
  bool has_lo = false, has_hi = false;
#pragma unroll
  for (int u = 0; u < NB; ++u) upd_lo[u] = upd_hi[u] = 0.0;

  if (any_on) {
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
    // ---- gather of cell r (:143-146 / :171-176): a ring of PF cells in
    // flight, raw loads from the cell clamped into the gathered range (the
    // values outside it are replaced when the cell is consumed)
    double rs_sig[PF], rs_kap[PF], rs_vel[PF], rs_vh[PF], rs_D[PF], rs_q[PF][NB], rs_vf[PF][NB];
    auto fetch = [&](int slot, int r) {
      const int rr = min(max(r, ldlo), ldhi);
      const long long cell = base + sL * (long long)(rr + 1);
      rs_sig[slot] = A.sig[cell];
      rs_kap[slot] = A.kap[cell];
      rs_vel[slot] = A.vel[cell];
      if (DIR == 2) rs_vh[slot] = A.vh[cell];
      // DelDxonD_cc(r,j) for x lines, DelDyonD_cc(i,r) for y lines (:163-166, :414-417)
      if (TOPO) {
        const int rd = min(max(r, rmin - 1), rmax + 1);
        rs_D[slot] = (DIR == 0) ? A.deld[(long long)(rd + 1) + (long long)(d.nx + 4) * (a + 1)]
                                : A.deld[(long long)(a + 1) + (long long)(d.nx + 4) * (rd + 1)];
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        rs_q[slot][u] = c[cell + bofs[u]];
        if (DIR == 2) rs_vf[slot][u] = A.vf[cell + bofs[u]];
      }
    };
#pragma unroll
    for (int p = 0; p < PF; ++p) fetch(p, lo - 2 + p);
    // whole turns of the ring: cells past hi+2 update nothing (their
    // stores go to junk and no outflow is recorded)
    for (int r0 = lo - 2; r0 <= hi + 2; r0 += PF) {
#pragma unroll
      for (int p = 0; p < PF; ++p) {
        const int r = r0 + p;
        const bool ld = (r >= ldlo && r <= ldhi);
        const double sig0 = ld ? rs_sig[p] : 0.0, kap0 = ld ? rs_kap[p] : 1.0;
        const double vel0 = (DIR == 2 || !ld) ? 0.0 : rs_vel[p];
        const double vzh0 = (DIR == 2 && ld) ? rs_vel[p] + rs_vh[p] : 0.0;
        const double D0 = (TOPO && ld && r >= rmin - 1 && r <= rmax + 1) ? rs_D[p] : 0.0;
        double vfc[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          // the halo cells come from the registers filled before the barrier
          // (A.zg: ghost cells read as Set_BC(1)'s zeros)
          double v = (ld && exists[u] && (!A.zg || (r >= 1 && r <= N))) ? rs_q[p][u] : 0.0;
          if constexpr (!GM) {
            // r, lo and hi are uniform across the wave: a scalar branch
            if (r == lo - 2 || r == lo - 1 || r == hi + 1 || r == hi + 2) {
              const int k = (r == lo - 2) ? 0 : (r == lo - 1) ? 1 : (r == hi + 1) ? 2 : 3;
              v = s_halo[(4 * u + k) * nthr + tid];
            }
          }
          q0[u] = v;
          vfc[u] = (DIR == 2 && ld && exists[u]) ? rs_vf[p][u] : 0.0;
        }
        fetch(p, r + PF);
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
            if (ld && exists[u]) vb0[u] = vzh0 + vfc[u];
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
        const bool upd_in = cc >= lo && cc <= hi && cc >= rmin - 1 && cc <= rmax + 1;
        const bool st_in = upd_in && cc >= rmin && cc <= rmax;
        const bool at_lo = upd_in && cc == 0 && rmin == 1 && (DIR != 0 || !A.periodic);
        const bool at_hi = upd_in && rmax == N && (DIR != 0 || !A.periodic) &&
                           cc == ((DIR == 2) ? ncells + 1 : N + 1);
        has_lo = has_lo || at_lo;
        has_hi = has_hi || at_hi;
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
          // (:264-281); every cell issues one store (masked-off ones to junk)
          const double upd = -dtv2 * (fslI + fsr2[u] + fssI - fss2[u]);
          *((st_in && on[u]) ? co + base + sL * (long long)(cc + 1) + bofs[u] : junk) = q2[u] + upd;
          // x outflow skipped when periodic (:289-294); y (:540-545); deposit
          // and top outflow of z, the top from update_cc(ncells+1)
          // (src/AdvectionVert_DCU.f90:306-311)
          if (at_lo) upd_lo[u] = upd;
          if (at_hi) upd_hi[u] = upd;
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
  }
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const long long of = ofs0 + ofs_bin * (long long)(n0 + u);
    if (has_lo && on[u]) A.of1[of] = A.of1[of] + upd_lo[u];
    if (has_hi && on[u]) A.of2[of] = A.of2[of] + upd_hi[u];
  }
  // cells the reference leaves at zero: inactive bins and lines, and cells
  // outside [rmin, rmax] (zero fill of ts1 then copy, :122, :302-303)
  const int zlo = max(lo, 1), zhi = min(hi, N);
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    if (!exists[u]) continue;
    for (int r = zlo; r <= zhi; ++r)
      if (!on[u] || r < rmin || r > rmax) co[base + sL * (long long)(r + 1) + bofs[u]] = 0.0;
    if (DIR == 2 && w == 0) {
      // DepositGranularity = outflow_xy1 (src/AdvectionVert_DCU.f90:321);
      // segment 0 owns the bottom ghost update
      const long long of = ofs0 + ofs_bin * (long long)(n0 + u);
      A.dep[(long long)(a - 1) + (long long)d.nx * ((long long)(b - 1) + (long long)d.ny * (n0 + u))] =
          A.of1[of];
    }
  }
}

template <int DIR, int NB, bool TOPO, bool GM>
__global__ __launch_bounds__(256) void k_advect(AdvArgs A) {
  advect_body<DIR, NB, TOPO, GM>(A, blockIdx.x, gridDim.x);
}
// the same sweep over the members of an ensemble (one launch)
template <int DIR, int NB, bool TOPO, bool GM>
__global__ __launch_bounds__(256) void k_advect_batch(
    const ASH3D_CONSTANT AdvArgs *__restrict__ args, const ASH3D_CONSTANT int *__restrict__ member,
    const ASH3D_CONSTANT long long *__restrict__ first) {
  const int m = member[blockIdx.x];
  const AdvArgs &A = *(const AdvArgs *)(args + m);
  advect_body<DIR, NB, TOPO, GM>(A, (long long)blockIdx.x - first[m], first[m + 1] - first[m]);
}

template <int DIR, int NB>
hipError_t launch_dir(hipStream_t s, const AdvArgs &a_in, LaunchSpec *sp = nullptr) {
  AdvArgs a = a_in;
  const int ntile = (a.g.Na + 63) / 64;
  const long long groups = (a.d.ns + NB - 1) / NB;
  dim3 grid, block;
  if (a.gy > 0) {
    // bin groups along y: gy of them per workgroup (at most 4 waves: the
    // kernel's launch bounds)
    a.gy = (int)std::min<long long>(std::min(a.gy, 4), groups);
    grid = dim3((unsigned)(padded_tiles((long long)ntile * a.g.Nb) * ((groups + a.gy - 1) / a.gy)));
    block = dim3(64, a.gy);
  } else {
    // segments per line: enough waves to fill the chip, segments >= 16
    // cells, at most 4 (one wave per SIMD of the CU holding the workgroup)
    const long long lanes = (long long)ntile * 64 * a.g.Nb * groups;
    int W = (int)((1LL << 21) / (lanes > 0 ? lanes : 1));
    W = W < 1 ? 1 : (W > 4 ? 4 : W);
    while (W > 1 && (a.g.N + 2) / W < 16) --W;
    grid = dim3((unsigned)(padded_tiles((long long)ntile * a.g.Nb) * groups));
    block = dim3(64, W);
  }
  const bool topo = DIR != 2 && a.deld != nullptr;
  const size_t lds = a.gy > 0 ? 0 : sizeof(double) * 4 * NB * (size_t)block.x * block.y;
  if (a.gy > 0)
    return issue(s, sp, topo ? (const void *)k_advect<DIR, NB, true, true> : (const void *)k_advect<DIR, NB, false, true>,
                 topo ? (const void *)k_advect_batch<DIR, NB, true, true>
                      : (const void *)k_advect_batch<DIR, NB, false, true>,
                 grid, block, lds, a);
  return issue(s, sp, topo ? (const void *)k_advect<DIR, NB, true, false> : (const void *)k_advect<DIR, NB, false, false>,
               topo ? (const void *)k_advect_batch<DIR, NB, true, false>
                    : (const void *)k_advect_batch<DIR, NB, false, false>,
               grid, block, lds, a);
}

template <int NB>
hipError_t launch_nb(hipStream_t s, const AdvArgs &a, LaunchSpec *sp = nullptr) {
  switch (a.g.dir) {
    case 0: return launch_dir<0, NB>(s, a, sp);
    case 1: return launch_dir<1, NB>(s, a, sp);
    default: return launch_dir<2, NB>(s, a, sp);
  }
}

}  // namespace

// Default bins per thread (kind 0 advection, 1 diffusion; dir 0..2):
// more bins share the per-cell terms, fewer keep more waves resident.
int pick_bins_per_thread(int ns, int kind, int dir) {
  // measured on the 360x240x50x15 forecast grid (bench autotuner): y
  // advection 3, z advection 2, x line solves (row tiles) 2, y/z line
  // solves 8
  if (kind == 0 && dir == 1 && ns >= 3) return 3;
  const int cap = (kind == 0) ? 2 : (dir == 0 ? 2 : 8);
  int nb = 1;
  while (nb * 2 <= cap && nb * 2 <= ns) nb *= 2;
  return nb;
}

static hipError_t advect_any(hipStream_t s, const AdvArgs &a, int nb, LaunchSpec *sp) {
  if (a.g.dir == 0) {
    hipError_t e = sp ? spec_advect_x_rows(a, sp) : launch_advect_x_rows(s, a);
    if (e != hipErrorInvalidValue) return e;
    (void)hipGetLastError();  // row too long for one workgroup: stream it
  }
  switch (nb) {
    case 1: return launch_nb<1>(s, a, sp);
    case 2: return launch_nb<2>(s, a, sp);
    case 3: return launch_nb<3>(s, a, sp);
    case 4: return launch_nb<4>(s, a, sp);
    default: return launch_nb<8>(s, a, sp);
  }
}

hipError_t launch_advect(hipStream_t s, const AdvArgs &a, int nb) { return advect_any(s, a, nb, nullptr); }

hipError_t spec_advect(const AdvArgs &a, int nb, LaunchSpec *sp) { return advect_any(nullptr, a, nb, sp); }

}  // namespace ash3d
