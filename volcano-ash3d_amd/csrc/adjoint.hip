// Discrete adjoint of the transport step.
//
// Advection.  For fixed limiter branches a DCU sweep is linear in q: the
// limited difference is ldq = alpha*dq_I + beta*dqu_I, with (alpha, beta)
// fixed by the branch phi(theta) takes (superbee: (0,0), (0,2), (1,0),
// (0,1), (2,0); the clamped theta = +-3 gives (phi(+-3), 0)), and every
// other term of the update (src/AdvectionHorz_DCU.f90:238-281) is linear.
// The adjoint kernels recompute the branches from the pre-sweep state
// exactly as the forward sweep does and apply the transpose of that linear
// map to lambda, the adjoint of the post-sweep interior.  With
// mu_c = -dt/kappa_c * lambda_c on the updated cells, the update pairs
// fs_l(I) with mu_{I-1}, fs_r(I) with mu_I and fss(I) with mu_{I-1} - mu_I;
// interface I contributes to the inputs q_{I-2} .. q_{I+1}.  Ghost inputs
// are zero (Set_BC(1)) except the periodic x copies, whose adjoint is
// folded back onto cells 1, 2, nx-1, nx.  Interior cells outside the
// active ranges leave the sweep as zero, so only their stencil terms
// remain.
//
// Diffusion.  A theta-scheme line maps q to x = A^{-1} B q
// (src/Diffusion.F90:807-890); its adjoint is lambda' = B^T A^{-T} lambda:
// a tridiagonal solve with the transposed matrix (the elimination with
// partial pivoting of dgtsv) and the transposed right-hand-side stencil.
// The row coefficients are rebuilt on both passes from the row factors L, R
// the forward sweeps cache.  The explicit scheme
// (src/Diffusion.F90:335-377) is a three-point stencil whose Neumann (or
// periodic x) ghosts fold back onto the end cells.
#include <algorithm>

#include "common.hpp"
#include "dcu.hpp"
#include "lanes.hpp"

namespace ash3d {

namespace {

constexpr int kBlock = 256;

// (alpha, beta) of ldq = alpha*dq + beta*dqu for the branch the forward
// limiter (dcu.hpp limited_flux) takes; (0,0) where it skips the interface.
__device__ inline void limiter_coeffs(int lim, double dqm, double dq0, double dqp, double u, double &alpha,
                                      double &beta) {
  alpha = 0.0;
  beta = 0.0;
  if (lim == 0 || fabs(dq0) <= EPS_THRESH) return;
  if (lim == 1) { alpha = 1.0; return; }
  const double dqu = (u > 0.0) ? dqm : dqp;
  const bool clamped = fabs(dqu) > 3.0 * fabs(dq0) || fabs(dq0) < EPS_THRESH;
  const double theta = clamped ? copysign(3.0, dqu * dq0) : dqu / dq0;
  // ldq = dq*phi(theta); on a branch phi = p0 + p1*theta, so
  // ldq = p0*dq + p1*dqu, or dq*phi(theta) with theta constant when clamped
  double p0 = 0.0, p1 = 0.0;
  switch (lim) {
    case 2: beta = 1.0; return;  // Beam-Warming: ldq = dqu
    case 3: p0 = 0.5; p1 = 0.5; break;  // Fromm
    case 4:  // minmod
      if (theta <= 0.0) {} else if (theta < 1.0) p1 = 1.0; else p0 = 1.0;
      break;
    case 5:  // superbee
      if (theta <= 0.0) {}
      else if (theta <= 0.5) p1 = 2.0;
      else if (theta <= 1.0) p0 = 1.0;
      else if (theta <= 2.0) p1 = 1.0;
      else p0 = 2.0;
      break;
    case 6: {  // MC
      const double a = (1.0 + theta) / 2.0, b2 = 2.0 * theta;
      if (theta <= 0.0) {}
      else if (b2 <= a && b2 <= 2.0) p1 = 2.0;
      else if (a <= 2.0) { p0 = 0.5; p1 = 0.5; }
      else p0 = 2.0;
      break;
    }
    default: return;
  }
  if (clamped) alpha = p0 + p1 * theta;
  else { alpha = p0; beta = p1; }
}

// Contributions of interface I to the adjoints of q_{I-2}, q_{I-1}, q_I,
// q_{I+1} (e0..e3).  cl = mu_{I-1}, cr = mu_I; g = 0.5|u|(1 - dtv_I|u|);
// tfD and tfDm1 the topographic factors tf*D of cells I and I-1.
__device__ inline void interface_adjoint(int lim, double qm2, double qm1, double q0, double qp1, double u,
                                         double divp, double divm, double g, double tfD, double tfDm1,
                                         double cl, double cr, double &e0, double &e1, double &e2,
                                         double &e3) {
  e0 = 0.0;
  e3 = 0.0;
  // fs_l(I) = min(0,u)(q_I - q_{I-1}) + q_{I-1} divp - tf_{I-1} q_{I-1} D_{I-1}
  // fs_r(I) = max(0,u)(q_I - q_{I-1}) + q_I divm + tf_I q_I D_I
  e2 = fmin(0.0, u) * cl + (fmax(0.0, u) + divm + tfD) * cr;
  e1 = (divp - fmin(0.0, u) - tfDm1) * cl - fmax(0.0, u) * cr;
  double al, be;
  limiter_coeffs(lim, qm1 - qm2, q0 - qm1, qp1 - q0, u, al, be);
  const double w = g * (cl - cr);
  e2 += al * w;
  e1 -= al * w;
  if (be != 0.0) {
    if (u > 0.0) { e1 += be * w; e0 -= be * w; }
    else { e3 += be * w; e2 -= be * w; }
  }
}

// ---- y and z: one lane per line, streaming along it -------------------
template <int DIR, bool TOPO>
__global__ __launch_bounds__(kBlock) void k_adj_advect_lines(AdvArgs A, const double *__restrict__ qpre,
                                                             double *lam, int nb) {
  const Dims d = A.d;
  const LineGeom g = A.g;
  const TileGroup tg = xcd_tile_group(blockIdx.x, (d.ns + nb - 1) / nb);
  const long long gid = tg.tile * kBlock + threadIdx.x;
  if (gid >= (long long)g.Na * g.Nb) return;
  const int a = (int)(gid % g.Na) + 1;
  const int b = (int)(gid / g.Na) + 1;
  const int N = g.N, rmin = g.rmin, rmax = g.rmax;
  const long long sL = line_stride(d, DIR), base = line_base(d, DIR, a, b);
  const bool line_on = a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  const int lim = A.limiter;
  const double dt = A.dt;
  auto cell = [&](int r) { return base + sL * (long long)(r + 1); };
  auto ld_ok = [&](int r) { return r >= rmin - 2 && r <= rmax + 2; };
  auto uin = [&](int r) { return r >= rmin - 1 && r <= rmax + 1; };

  for (int n = tg.group * nb; n < d.ns && n < (tg.group + 1) * nb; ++n) {
    const long long bofs = d.ncell * (long long)n;
    if (!(line_on && A.aloft.on(n))) {
      for (int r = 1; r <= N; ++r) lam[cell(r) + bofs] = 0.0;
      continue;
    }
    auto qv = [&](int r) { return ld_ok(r) ? qpre[cell(r) + bofs] : 0.0; };
    auto velv = [&](int r) -> double {
      if (!ld_ok(r)) return 0.0;
      if (DIR == 2) return A.vel[cell(r)] + A.vh[cell(r)] + A.vf[cell(r) + bofs];
      return A.vel[cell(r)];
    };
    auto sigv = [&](int r) { return ld_ok(r) ? A.sig[cell(r)] : 0.0; };
    auto uv = [&](int I) { return uin(I) ? 0.5 * (velv(I - 1) + velv(I)) * sigv(I) : 0.0; };
    auto tfDv = [&](int r) -> double {
      if (!TOPO || !uin(r)) return 0.0;
      const double D = (DIR == 0) ? A.deld[(long long)(r + 1) + (long long)(d.nx + 4) * (a + 1)]
                                  : A.deld[(long long)(a + 1) + (long long)(d.nx + 4) * (r + 1)];
      return 0.5 * velv(r) * (sigv(r) + sigv(r + 1)) * D;
    };
    auto muv = [&](int c) {
      return (c >= rmin && c <= rmax) ? -(dt / A.kap[cell(c)]) * lam[cell(c) + bofs] : 0.0;
    };
    double ghost[4] = {0.0, 0.0, 0.0, 0.0};  // adjoint reaching cells -1, 0, N+1, N+2
    // final adjoint of input cell r; lam at r is read here for the last time
    auto emit = [&](int r, double v) {
      if (r >= 1 && r <= N) {
        const double direct = (r >= rmin && r <= rmax) ? lam[cell(r) + bofs] : 0.0;
        lam[cell(r) + bofs] = direct + v;
      } else if (r == -1) ghost[0] += v;
      else if (r == 0) ghost[1] += v;
      else if (r == N + 1) ghost[2] += v;
      else if (r == N + 2) ghost[3] += v;
    };
    for (int r = 1; r < rmin - 2 && r <= N; ++r) emit(r, 0.0);
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;  // pending adjoints of I-2, I-1, I
    double q_m2 = qv(rmin - 2), q_m1 = qv(rmin - 1), q_0 = qv(rmin), q_p1 = qv(rmin + 1);
    double u_m1 = uv(rmin - 1), u_0 = uv(rmin), u_p1 = uv(rmin + 1);
    double mu_m1 = 0.0, mu_0 = muv(rmin);
    double tfD_m1 = tfDv(rmin - 1), tfD_0 = tfDv(rmin);
    for (int I = rmin; I <= rmax + 1; ++I) {
      const double divp = fmax(0.0, u_0) - fmax(0.0, u_m1);
      const double divm = fmin(0.0, u_p1) - fmin(0.0, u_0);
      const double aus = fabs(u_0);
      const double gI = 0.5 * aus * (1.0 - (dt / A.kap[cell(I)]) * aus);
      double e0, e1, e2, e3;
      interface_adjoint(lim, q_m2, q_m1, q_0, q_p1, u_0, divp, divm, gI, tfD_0, tfD_m1, mu_m1, mu_0, e0, e1,
                        e2, e3);
      emit(I - 2, acc0 + e0);
      acc0 = acc1 + e1;
      acc1 = acc2 + e2;
      acc2 = e3;
      q_m2 = q_m1; q_m1 = q_0; q_0 = q_p1; q_p1 = qv(I + 2);
      u_m1 = u_0; u_0 = u_p1; u_p1 = uv(I + 2);
      mu_m1 = mu_0; mu_0 = muv(I + 1);
      tfD_m1 = tfD_0; tfD_0 = tfDv(I + 1);
    }
    emit(rmax, acc0);
    emit(rmax + 1, acc1);
    emit(rmax + 2, acc2);
    for (int r = rmax + 3; r <= N; ++r) emit(r, 0.0);
    if (A.periodic) {
      // Set_BC(1) copies: cell -1 <- N-1, 0 <- N, N+1 <- 1, N+2 <- 2
      lam[cell(N - 1) + bofs] += ghost[0];
      lam[cell(N) + bofs] += ghost[1];
      lam[cell(1) + bofs] += ghost[2];
      lam[cell(2) + bofs] += ghost[3];
    }
  }
}

// ---- y and z: segmented lines, NB bins per lane -----------------------
// The layout of the forward streamed sweep: a workgroup is 64 lanes
// (consecutive transverse index a, coalesced) x W segments of the same
// lines; segment w emits the adjoint of cells [lo, hi] (a partition of
// 1..N) and streams cells lo-3 .. hi+3, since cell c collects interfaces
// c-1 .. c+2 and interface I reads cells I-2 .. I+1.  lambda is updated in
// place: the halo cells lo-2, lo-1, hi+1, hi+2 (emitted by the
// neighbouring segments) are read into registers before a barrier.  The
// pending sums of each cell are formed in the same order as the one-lane
// kernel above (e3, e2, e1, e0, then the direct term).
// GM: whole lines per wave (no halos: the ring reads those cells itself
// and nobody writes them).
template <int DIR, int NB, bool TOPO, bool GM>
__global__ __launch_bounds__(256) void k_adj_advect_seg(AdvArgs A, const double *__restrict__ qpre, double *lam) {
  constexpr int PF = 2;  // cells in flight
  const Dims d = A.d;
  const LineGeom g = A.g;
  const int ntile = (g.Na + 63) / 64;
  // A.gy > 0: the waves of a workgroup take gy bin groups of the same whole
  // lines (as the forward sweep), else W segments of the lines
  const int G = (d.ns + NB - 1) / NB;
  constexpr bool gmode = GM;
  const TileGroup tg = xcd_tile_group(blockIdx.x, gmode ? (G + A.gy - 1) / A.gy : G);
  if (tg.tile >= (long long)ntile * g.Nb) return;  // padding workgroup (uniform)
  const int a = (int)(tg.tile % ntile) * 64 + (int)threadIdx.x + 1;
  const int b = (int)(tg.tile / ntile) + 1;
  const bool lane_ok = a <= g.Na;
  const int W = gmode ? 1 : blockDim.y, w = gmode ? 0 : threadIdx.y;
  const int n0 = (gmode ? tg.group * A.gy + (int)threadIdx.y : tg.group) * NB;
  const int N = g.N, rmin = g.rmin, rmax = g.rmax;
  const int seg = (N + W - 1) / W;
  const int lo = 1 + w * seg, hi = min(N, lo + seg - 1);
  const long long sL = line_stride(d, DIR);
  const long long base = lane_ok ? line_base(d, DIR, a, b) : 0;
  const bool line_on = lane_ok && a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  double *junk = A.junk + (threadIdx.x + 64 * threadIdx.y + 256 * (blockIdx.x % (kJunk / 256)));

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
  const int ldlo = rmin - 2, ldhi = rmax + 2;  // cells the forward gathers
  auto act = [&](int r) { return r >= rmin && r <= rmax; };
  // lambda of the halo cells (zero where mu and the direct term vanish)
  // (held in LDS, 4*NB per thread element-major, until the four
  // iterations that read them)
  extern __shared__ double s_halo[];
  const int tid = (int)threadIdx.x + 64 * (int)threadIdx.y, nthr = 64 * W;
  if constexpr (!GM) {
    auto halo = [&](int r, int u) {
      return (any_on && exists[u] && act(r)) ? lam[base + sL * (long long)(r + 1) + bofs[u]] : 0.0;
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

  if (any_on && lo <= hi) {
    const double dt = A.dt;
    const int lim = A.limiter;
    // ring of PF cells in flight: raw loads from the cell clamped into the
    // gathered range (values outside it are replaced when consumed)
    double rs_vel[PF], rs_vh[PF], rs_sig[PF], rs_kap[PF], rs_D[PF], rs_q[PF][NB], rs_l[PF][NB], rs_vf[PF][NB];
    auto fetch = [&](int slot, int r) {
      const int rr = min(max(r, ldlo), ldhi);
      const long long cell = base + sL * (long long)(rr + 1);
      rs_vel[slot] = A.vel[cell];
      if (DIR == 2) rs_vh[slot] = A.vh[cell];
      rs_sig[slot] = A.sig[cell];
      rs_kap[slot] = A.kap[cell];
      if (TOPO) {
        const int rd = min(max(r, rmin - 1), rmax + 1);
        rs_D[slot] = A.deld[(long long)(a + 1) + (long long)(d.nx + 4) * (rd + 1)];
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        rs_q[slot][u] = qpre[cell + bofs[u]];
        rs_l[slot][u] = lam[cell + bofs[u]];
        if (DIR == 2) rs_vf[slot][u] = A.vf[cell + bofs[u]];
      }
    };
    // windows (suffix 1 = r-1, 2 = r-2, 3 = r-3)
    double q1[NB], q2[NB], q3[NB], mu1[NB], mu2[NB], dr1[NB], dr2[NB], dr3[NB];
    double acc0[NB], acc1[NB], acc2[NB];
    double vb1[NB], ub1[NB], ub2[NB];  // z: per-bin velocity and usig
    double vel1 = 0.0, sig1 = 0.0, u1 = 0.0, u2 = 0.0, dtk1 = 0.0, D1 = 0.0, tf2 = 0.0;
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      q1[u] = q2[u] = q3[u] = mu1[u] = mu2[u] = dr1[u] = dr2[u] = dr3[u] = 0.0;
      acc0[u] = acc1[u] = acc2[u] = 0.0;
      vb1[u] = ub1[u] = ub2[u] = 0.0;
    }
#pragma unroll
    for (int p = 0; p < PF; ++p) fetch(p, lo - 3 + p);
    for (int r0 = lo - 3; r0 <= hi + 3; r0 += PF) {
#pragma unroll
      for (int p = 0; p < PF; ++p) {
        const int r = r0 + p;
        const bool ld = r >= ldlo && r <= ldhi;
        const bool uin = r >= rmin - 1 && r <= rmax + 1;
        const double sig0 = ld ? rs_sig[p] : 0.0;
        const double vel0 = ld ? rs_vel[p] : 0.0;
        const double vzh0 = (DIR == 2 && ld) ? rs_vel[p] + rs_vh[p] : 0.0;
        const double dtk0 = dt / rs_kap[p];
        const double D0 = (TOPO && uin) ? rs_D[p] : 0.0;
        double q0[NB], l0[NB], vf0[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          q0[u] = (ld && exists[u]) ? rs_q[p][u] : 0.0;
          double v = rs_l[p][u];
          if constexpr (!GM) {
            // r, lo and hi are uniform across the wave: a scalar branch
            if (r == lo - 2 || r == lo - 1 || r == hi + 1 || r == hi + 2) {
              const int k = (r == lo - 2) ? 0 : (r == lo - 1) ? 1 : (r == hi + 1) ? 2 : 3;
              v = s_halo[(4 * u + k) * nthr + tid];
            }
          }
          l0[u] = v;
          vf0[u] = (DIR == 2 && ld && exists[u]) ? rs_vf[p][u] : 0.0;
        }
        fetch(p, r + PF);
        // usig of interface r (:159-162), tf*D of cell r-1 (:252-257)
        const double u0 = (DIR != 2 && uin) ? 0.5 * (vel1 + vel0) * sig0 : 0.0;
        const bool uin1 = r - 1 >= rmin - 1 && r - 1 <= rmax + 1;
        const double tf1 = (TOPO && uin1) ? 0.5 * vel1 * (sig1 + sig0) * D1 : 0.0;
        // interface I = r-1
        const int I = r - 1;
        const bool ifc = I >= rmin && I <= rmax + 1;
        const int cc = r - 3;  // the cell whose sum is complete
        const bool st = cc >= lo && cc <= hi;
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          double vb0 = 0.0, ub0 = 0.0;
          if (DIR == 2) {
            if (ld && exists[u]) vb0 = vzh0 + vf0[u];
            ub0 = uin ? 0.5 * (vb1[u] + vb0) * sig0 : 0.0;
          }
          const double uI = (DIR == 2) ? ub1[u] : u1;    // usig_I
          const double uIm = (DIR == 2) ? ub2[u] : u2;   // usig_{I-1}
          const double uIp = (DIR == 2) ? ub0 : u0;      // usig_{I+1}
          double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0;
          if (ifc) {
            const double divp = fmax(0.0, uI) - fmax(0.0, uIm);
            const double divm = fmin(0.0, uIp) - fmin(0.0, uI);
            const double aus = fabs(uI);
            const double gI = 0.5 * aus * (1.0 - dtk1 * aus);
            interface_adjoint(lim, q3[u], q2[u], q1[u], q0[u], uI, divp, divm, gI, tf1, tf2, mu2[u], mu1[u], e0,
                              e1, e2, e3);
          }
          const double v = acc0[u] + e0;
          *((st && on[u]) ? lam + base + sL * (long long)(cc + 1) + bofs[u] : junk) = dr3[u] + v;
          acc0[u] = acc1[u] + e1;
          acc1[u] = acc2[u] + e2;
          acc2[u] = e3;
          // mu_r = -dt/kappa_r lambda_r and the direct term of cell r
          const double mu0 = act(r) ? -dtk0 * l0[u] : 0.0;
          const double dr0 = act(r) ? l0[u] : 0.0;
          q3[u] = q2[u]; q2[u] = q1[u]; q1[u] = q0[u];
          mu2[u] = mu1[u]; mu1[u] = mu0;
          dr3[u] = dr2[u]; dr2[u] = dr1[u]; dr1[u] = dr0;
          if (DIR == 2) {
            ub2[u] = ub1[u];
            ub1[u] = ub0;
            vb1[u] = vb0;
          }
        }
        u2 = u1;
        u1 = u0;
        tf2 = tf1;
        vel1 = vel0;
        sig1 = sig0;
        dtk1 = dtk0;
        D1 = D0;
      }
    }
  }
  // inactive bins and lines leave the sweep as zero
  if (lo <= hi) {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      if (!exists[u] || on[u]) continue;
      for (int r = lo; r <= hi; ++r) lam[base + sL * (long long)(r + 1) + bofs[u]] = 0.0;
    }
  }
}

// ---- x: one workgroup per row, gather through LDS ---------------------
template <int CPT, bool TOPO>
__global__ __launch_bounds__(256) void k_adj_advect_x_rows(AdvArgs A, const double *__restrict__ qpre,
                                                           double *lam) {
  extern __shared__ double smem[];
  const Dims d = A.d;
  const LineGeom g = A.g;
  const int nx = d.nx;
  const int W = nx + 4;  // cells -1..nx+2 at LDS index c+1
  double *s_q = smem, *s_mu = smem + W;
  double *s_e0 = smem + 2 * W, *s_e1 = smem + 3 * W, *s_e2 = smem + 4 * W, *s_e3 = smem + 5 * W;

  const int j = (int)(blockIdx.x % (unsigned)d.ny) + 1;
  const int k = (int)(blockIdx.x / (unsigned)d.ny) + 1;
  const int rmin = g.rmin, rmax = g.rmax;
  const bool row_on = j >= g.a0 && j <= g.a1 && k >= g.b0 && k <= g.b1;
  const long long base = d.p3(-1, j, k);
  const int t = threadIdx.x, BT = blockDim.x;
  const int lim = A.limiter;

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

  // bin-independent terms at interface / cell c (as the forward row kernel)
  double u[CPT], dtv[CPT], divp[CPT], divm[CPT], gI[CPT], tfD[CPT], tfDm1[CPT];
  if (any_on) {
    double vel[CPT], sig[CPT], kap[CPT], D[CPT];
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
      if (valid[m]) { s_q[c + 1] = vel[m]; s_mu[c + 1] = sig[m]; }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int c = cc[m], p = c + 1;
      u[m] = 0.0; dtv[m] = 0.0; tfD[m] = 0.0; gI[m] = 0.0;
      if (!valid[m]) continue;
      const bool uin = c >= rmin - 1 && c <= rmax + 1;
      const double velm1 = p >= 1 ? s_q[p - 1] : 0.0;
      const double sigp1 = p + 1 < W ? s_mu[p + 1] : 0.0;
      u[m] = uin ? 0.5 * (velm1 + vel[m]) * sig[m] : 0.0;
      dtv[m] = uin ? A.dt / kap[m] : 0.0;
      if (TOPO) tfD[m] = 0.5 * vel[m] * (sig[m] + sigp1) * D[m];
      const double aus = fabs(u[m]);
      gI[m] = 0.5 * aus * (1.0 - dtv[m] * aus);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < CPT; ++m)
      if (valid[m]) { s_q[cc[m] + 1] = u[m]; s_mu[cc[m] + 1] = tfD[m]; }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      divp[m] = divm[m] = tfDm1[m] = 0.0;
      if (!valid[m]) continue;
      const int p = cc[m] + 1;
      const double um1 = p >= 1 ? s_q[p - 1] : 0.0;
      const double up1 = p + 1 < W ? s_q[p + 1] : 0.0;
      divp[m] = fmax(0.0, u[m]) - fmax(0.0, um1);
      divm[m] = fmin(0.0, up1) - fmin(0.0, u[m]);
      if (TOPO && p >= 1) tfDm1[m] = s_mu[p - 1];
    }
    __syncthreads();
  }

  for (int n = 0; n < d.ns; ++n) {
    const long long bofs = d.ncell * (long long)n;
    if (!(row_on && A.aloft.on(n))) {
#pragma unroll
      for (int m = 0; m < CPT; ++m)
        if (cc[m] >= 1 && cc[m] <= nx) lam[base + (cc[m] + 1) + bofs] = 0.0;
      continue;
    }
    double lm[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int c = cc[m];
      lm[m] = 0.0;
      if (!valid[m]) continue;
      const bool act = c >= rmin && c <= rmax;
      lm[m] = act ? lam[base + (c + 1) + bofs] : 0.0;
      s_q[c + 1] = (c >= rmin - 2 && c <= rmax + 2) ? qpre[base + (c + 1) + bofs] : 0.0;
      s_mu[c + 1] = act ? -dtv[m] * lm[m] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int c = cc[m], p = c + 1;
      if (!valid[m]) continue;
      double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0;
      if (c >= rmin && c <= rmax + 1)
        interface_adjoint(lim, s_q[p - 2], s_q[p - 1], s_q[p], s_q[p + 1], u[m], divp[m], divm[m], gI[m], tfD[m],
                          tfDm1[m], s_mu[p - 1], s_mu[p], e0, e1, e2, e3);
      s_e0[p] = e0; s_e1[p] = e1; s_e2[p] = e2; s_e3[p] = e3;
    }
    __syncthreads();
    double v[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int p = cc[m] + 1;
      v[m] = 0.0;
      if (!valid[m]) continue;
      // interfaces r+2, r+1, r, r-1 reach cell r as their I-2, I-1, I, I+1
      v[m] = (p + 2 < W ? s_e0[p + 2] : 0.0) + (p + 1 < W ? s_e1[p + 1] : 0.0) + s_e2[p] +
             (p >= 1 ? s_e3[p - 1] : 0.0);
    }
    if (A.periodic) {
      __syncthreads();  // s_q is free: hold the ghost adjoints there
#pragma unroll
      for (int m = 0; m < CPT; ++m) {
        const int c = cc[m];
        if (valid[m] && (c <= 0 || c >= nx + 1)) s_q[c + 1] = v[m];
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < CPT; ++m) {
        const int c = cc[m];
        if (c == 1) v[m] += s_q[nx + 2];  // from cell nx+1
        if (c == 2) v[m] += s_q[nx + 3];  // from cell nx+2
        if (c == nx - 1) v[m] += s_q[0];  // from cell -1
        if (c == nx) v[m] += s_q[1];      // from cell 0
      }
    }
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int c = cc[m];
      if (c >= 1 && c <= nx) lam[base + (c + 1) + bofs] = lm[m] + v[m];
    }
    __syncthreads();  // before the next bin overwrites the LDS rows
  }
}

// ---- x rows, one wavefront per row ------------------------------------
// The row in registers in the pair layout of lanes.hpp, as the forward
// wavefront kernel: the pre-sweep row and lambda of a bin are loaded two
// bins ahead; no LDS and no barriers.  Every load and store of the bin loop
// is issued on every iteration (clamped addresses, junk stores).  The
// terms, their order and the periodic fold are those of
// k_adj_advect_x_rows.
template <int NC, bool TOPO>
__global__ __launch_bounds__(256) void k_adj_advect_x_wave(AdvArgs A, const double *__restrict__ qpre,
                                                           double *lam) {
  const Dims d = A.d;
  const LineGeom g = A.g;
  const int nx = d.nx, W = nx + 4;
  const int l = (int)(threadIdx.x & 63);
  const long long row = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= (long long)d.ny * d.nz) return;  // whole wavefront
  const int j = (int)(row % d.ny) + 1;
  const int k = (int)(row / d.ny) + 1;
  const int rmin = g.rmin, rmax = g.rmax;
  const bool row_on = j >= g.a0 && j <= g.a1 && k >= g.b0 && k <= g.b1;
  const long long base = d.p3(-1, j, k);
  const int lim = A.limiter;
  double *junk = A.junk + 2 * (long long)(threadIdx.x + blockDim.x * (blockIdx.x % (kJunk / 512)));
  const bool lane0 = l == 0, lane63 = l == 63;

  int cc[NC][2];
  bool valid[NC][2], ld[NC][2], uin[NC][2], act[NC][2];
  int pp[NC];
  bool store[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    const int p0 = 128 * m + 2 * l;
    pp[m] = min(p0, W - 2);
    store[m] = p0 >= 2 && p0 + 1 <= nx + 1;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int c = p0 + e - 1;
      cc[m][e] = c;
      valid[m][e] = p0 + e < W;
      ld[m][e] = valid[m][e] && c >= rmin - 2 && c <= rmax + 2;
      uin[m][e] = valid[m][e] && c >= rmin - 1 && c <= rmax + 1;
      act[m][e] = valid[m][e] && c >= rmin && c <= rmax;
    }
  }
  if (!row_on) {
    for (int n = 0; n < d.ns; ++n) {
      const long long bofs = d.ncell * (long long)n;
#pragma unroll
      for (int m = 0; m < NC; ++m)
        if (store[m]) *(double2 *)(lam + base + pp[m] + bofs) = make_double2(0.0, 0.0);
    }
    return;
  }

  // bin-independent terms at interface / cell c (as the forward row kernel);
  // the divergence corrections and g_I are formed again in each bin from
  // u and dt/kappa (fewer registers, the same values)
  double u[NC][2], dtv[NC][2], tfD[NC][2], tfDm1[NC][2];
  {
    double vel[NC][2], sig[NC][2], kap[NC][2], D[NC][2];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const long long o = base + pp[m];
      const double2 v2 = *(const double2 *)(A.vel + o);
      const double2 s2 = *(const double2 *)(A.sig + o);
      const double2 k2 = *(const double2 *)(A.kap + o);
      double2 d2 = make_double2(0.0, 0.0);
      if (TOPO) d2 = *(const double2 *)(A.deld + (long long)pp[m] + (long long)(nx + 4) * (j + 1));
      const double vv[2] = {v2.x, v2.y}, ss[2] = {s2.x, s2.y}, kk[2] = {k2.x, k2.y}, dd[2] = {d2.x, d2.y};
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        vel[m][e] = ld[m][e] ? vv[e] : 0.0;
        sig[m][e] = ld[m][e] ? ss[e] : 0.0;
        kap[m][e] = ld[m][e] ? kk[e] : 1.0;
        D[m][e] = (TOPO && uin[m][e]) ? dd[e] : 0.0;
      }
    }
    double velm1[NC][2], sigp1[NC][2];
    row_prev<NC>(vel, velm1, lane0);
    if (TOPO) row_next<NC>(sig, sigp1, lane63);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        u[m][e] = uin[m][e] ? 0.5 * (velm1[m][e] + vel[m][e]) * sig[m][e] : 0.0;
        dtv[m][e] = uin[m][e] ? A.dt / kap[m][e] : 0.0;
        tfD[m][e] = (TOPO && valid[m][e]) ? 0.5 * vel[m][e] * (sig[m][e] + sigp1[m][e]) * D[m][e] : 0.0;
      }
    if (TOPO) row_prev<NC>(tfD, tfDm1, lane0);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e)
        if (!TOPO) tfDm1[m][e] = 0.0;
  }

  // periodic fold: cells nx+1, nx+2 are element 0, 1 of lane l_hi, chunk m_hi
  const int m_hi = (nx + 2) >> 7, l_hi = ((nx + 2) & 127) >> 1;
  double sa[NC][2], sq[NC][2], sb[NC][2], sr[NC][2];  // lambda and q of bins n, n+1
  auto fetch = [&](double (&sl)[NC][2], double (&sqq)[NC][2], int n) {
    const int nn = (n < d.ns && A.aloft.on(n)) ? n : max(0, min(n - 2, d.ns - 1));
    const long long bofs = d.ncell * (long long)nn;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const double2 l2 = *(const double2 *)(lam + base + pp[m] + bofs);
      const double2 q2 = *(const double2 *)(qpre + base + pp[m] + bofs);
      sl[m][0] = l2.x; sl[m][1] = l2.y;
      sqq[m][0] = q2.x; sqq[m][1] = q2.y;
    }
  };
  auto bin = [&](double (&sl)[NC][2], double (&sqq)[NC][2], int n) {
    const bool on = A.aloft.on(n);
    double lm[NC][2], q[NC][2], mu[NC][2];
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        lm[m][e] = act[m][e] ? sl[m][e] : 0.0;
        q[m][e] = ld[m][e] ? sqq[m][e] : 0.0;
      }
    fetch(sl, sqq, n + 2);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e) mu[m][e] = act[m][e] ? -dtv[m][e] * lm[m][e] : 0.0;
    double um1[NC][2], up1[NC][2];
    row_prev<NC>(u, um1, lane0);
    row_next<NC>(u, up1, lane63);
    double qm2[NC][2], qm1[NC][2], qp1[NC][2], mum1[NC][2];
    row_prev2<NC>(q, qm2, lane0);
    row_prev<NC>(q, qm1, lane0);
    row_next<NC>(q, qp1, lane63);
    row_prev<NC>(mu, mum1, lane0);
    double e0[NC][2], e1[NC][2], e2[NC][2], e3[NC][2];
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int c = cc[m][e];
        e0[m][e] = e1[m][e] = e2[m][e] = e3[m][e] = 0.0;
        if (valid[m][e] && c >= rmin && c <= rmax + 1) {
          const double divp = fmax(0.0, u[m][e]) - fmax(0.0, um1[m][e]);
          const double divm = fmin(0.0, up1[m][e]) - fmin(0.0, u[m][e]);
          const double aus = fabs(u[m][e]);
          const double gI = 0.5 * aus * (1.0 - dtv[m][e] * aus);
          interface_adjoint(lim, qm2[m][e], qm1[m][e], q[m][e], qp1[m][e], u[m][e], divp, divm, gI, tfD[m][e],
                            tfDm1[m][e], mum1[m][e], mu[m][e], e0[m][e], e1[m][e], e2[m][e], e3[m][e]);
        }
      }
    // interfaces r+2, r+1, r, r-1 reach cell r as their I-2, I-1, I, I+1
    double e0n[NC][2], e1n[NC][2], e3p[NC][2], v[NC][2];
    row_next2<NC>(e0, e0n, lane63);
    row_next<NC>(e1, e1n, lane63);
    row_prev<NC>(e3, e3p, lane0);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e) v[m][e] = e0n[m][e] + e1n[m][e] + e2[m][e] + e3p[m][e];
    if (A.periodic) {
      double vh0 = v[0][0], vh1 = v[0][1];
#pragma unroll
      for (int m = 0; m < NC; ++m)
        if (m == m_hi) { vh0 = v[m][0]; vh1 = v[m][1]; }
      const double ga = lane_value(vh0, l_hi), gb = lane_value(vh1, l_hi);  // cells nx+1, nx+2
      const double gc = lane_value(v[0][0], 0), gd = lane_value(v[0][1], 0);  // cells -1, 0
#pragma unroll
      for (int m = 0; m < NC; ++m)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int c = cc[m][e];
          if (c == 1) v[m][e] += ga;
          if (c == 2) v[m][e] += gb;
          if (c == nx - 1) v[m][e] += gc;
          if (c == nx) v[m][e] += gd;
        }
    }
    const long long bofs = d.ncell * (long long)n;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      double *dst = store[m] ? lam + base + pp[m] + bofs : junk;
      *(double2 *)dst = on ? make_double2(lm[m][0] + v[m][0], lm[m][1] + v[m][1]) : make_double2(0.0, 0.0);
    }
  };
  fetch(sa, sq, 0);
  fetch(sb, sr, 1);
  int n = 0;
  for (; n + 1 < d.ns; n += 2) {
    bin(sa, sq, n);
    bin(sb, sr, n + 1);
  }
  if (n < d.ns) bin(sa, sq, n);
}

// ---- x rows longer than one wavefront's registers (nx even) -----------
// WV waves per row in one workgroup, wave w owning the row positions
// [S, S + L) with S = w L (L even) and holding the extended range
// [S - 4, S + L + 4) in the pair layout of k_adj_advect_x_wave (lane l of
// chunk m: positions S - 4 + 128 m + 2 l + e).  The adjoint of a cell
// gathers the terms of the four interfaces around it, which read q three
// cells and lambda two cells beyond it, so every cell a wave stores has
// what it needs in its own registers and the arithmetic is the wavefront
// kernel's.  The pre-sweep q and the per-cell fields of the extended
// range come from memory (they are read only); lambda, which the
// neighbouring waves update in place, is loaded at the wave's own
// positions only, and the two halo pairs are the neighbours' first and
// last pairs, exchanged through LDS (two slots by bin parity, one barrier
// per bin).  On a periodic row the ghost adjoints of cells -1, 0 (wave 0)
// and nx+1, nx+2 (the last wave) fold back across the row through LDS
// after a second barrier.
template <int NC, bool TOPO>
__global__ __launch_bounds__(512) void k_adj_advect_x_seg(AdvArgs A, const double *__restrict__ qpre, double *lam,
                                                          int L) {
  constexpr int CS = 128;
  const Dims d = A.d;
  const LineGeom g = A.g;
  const int nx = d.nx, W = nx + 4;
  const int l = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6), WV = (int)(blockDim.x >> 6);
  const long long row = blockIdx.x;
  const int j = (int)(row % d.ny) + 1;
  const int k = (int)(row / d.ny) + 1;
  const int rmin = g.rmin, rmax = g.rmax;
  const bool row_on = j >= g.a0 && j <= g.a1 && k >= g.b0 && k <= g.b1;
  const long long base = d.p3(-1, j, k);
  const int lim = A.limiter;
  double *junk = A.junk + 2 * (long long)(threadIdx.x + blockDim.x * (blockIdx.x % (kJunk / 1024)));
  const bool lane0 = l == 0, lane63 = l == 63;
  const int S = w * L, Send = min(S + L, W), X = S - 4;
  const bool has_lo = w > 0, has_hi = w + 1 < WV;
  // halo pairs: positions S-2, S-1 (offset 2) and S+L, S+L+1 (offset L+4);
  // own edge pairs: S, S+1 (offset 4) and S+L-2, S+L-1 (offset L+2)
  const int m_hh = (L + 4) / CS, l_hh = ((L + 4) % CS) / 2;
  const int m_le = (L + 2) / CS, l_le = ((L + 2) % CS) / 2;

  int cc[NC][2];
  bool valid[NC][2], ld[NC][2], uin[NC][2], act[NC][2];
  int pf[NC], pl[NC];  // field / q and lambda offsets of the pair (clamped)
  bool own[NC], store[NC], halo_lo[NC], halo_hi[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    const int p0 = X + CS * m + 2 * l;
    own[m] = p0 >= S && p0 < Send;
    pf[m] = min(max(p0, 0), W - 2);
    pl[m] = own[m] ? p0 : S;
    store[m] = own[m] && p0 >= 2 && p0 + 1 <= nx + 1;
    halo_lo[m] = has_lo && m == 0 && l == 1;
    halo_hi[m] = has_hi && m == m_hh && l == l_hh;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int c = p0 + e - 1;
      cc[m][e] = c;
      valid[m][e] = p0 + e >= 0 && p0 + e < W && p0 + e < Send + 4;
      ld[m][e] = valid[m][e] && c >= rmin - 2 && c <= rmax + 2;
      uin[m][e] = valid[m][e] && c >= rmin - 1 && c <= rmax + 1;
      act[m][e] = valid[m][e] && c >= rmin && c <= rmax;
    }
  }
  if (!row_on) {
    for (int n = 0; n < d.ns; ++n) {
      const long long bofs = d.ncell * (long long)n;
#pragma unroll
      for (int m = 0; m < NC; ++m)
        if (store[m]) *(double2 *)(lam + base + pl[m] + bofs) = make_double2(0.0, 0.0);
    }
    return;  // the whole workgroup: row_on is the row's
  }

  // bin-independent terms, as the wavefront kernel
  double u[NC][2], dtv[NC][2], tfD[NC][2], tfDm1[NC][2];
  {
    double vel[NC][2], sig[NC][2], kap[NC][2], D[NC][2];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const long long o = base + pf[m];
      const double2 v2 = *(const double2 *)(A.vel + o);
      const double2 s2 = *(const double2 *)(A.sig + o);
      const double2 k2 = *(const double2 *)(A.kap + o);
      double2 d2 = make_double2(0.0, 0.0);
      if (TOPO) d2 = *(const double2 *)(A.deld + (long long)pf[m] + (long long)(nx + 4) * (j + 1));
      const double vv[2] = {v2.x, v2.y}, ss[2] = {s2.x, s2.y}, kk[2] = {k2.x, k2.y}, dd[2] = {d2.x, d2.y};
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        vel[m][e] = ld[m][e] ? vv[e] : 0.0;
        sig[m][e] = ld[m][e] ? ss[e] : 0.0;
        kap[m][e] = ld[m][e] ? kk[e] : 1.0;
        D[m][e] = (TOPO && uin[m][e]) ? dd[e] : 0.0;
      }
    }
    double velm1[NC][2], sigp1[NC][2];
    row_prev<NC>(vel, velm1, lane0);
    if (TOPO) row_next<NC>(sig, sigp1, lane63);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        u[m][e] = uin[m][e] ? 0.5 * (velm1[m][e] + vel[m][e]) * sig[m][e] : 0.0;
        dtv[m][e] = uin[m][e] ? A.dt / kap[m][e] : 0.0;
        tfD[m][e] = (TOPO && valid[m][e]) ? 0.5 * vel[m][e] * (sig[m][e] + sigp1[m][e]) * D[m][e] : 0.0;
      }
    if (TOPO) row_prev<NC>(tfD, tfDm1, lane0);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e)
        if (!TOPO) tfDm1[m][e] = 0.0;
  }

  // edge pairs s_edge[((slot * 8 + wave) * 2 + first/last) * 2 + e]; the
  // periodic fold's ghost adjoints s_fold = {v(-1), v(0), v(nx+1), v(nx+2)}
  __shared__ double s_edge[2 * 8 * 4];
  __shared__ double s_fold[4];
  // cells nx+1, nx+2 (positions nx+2, nx+3) in the last wave
  const int o_g = nx + 2 - ((WV - 1) * L - 4);
  const int m_g = o_g / CS, l_g = (o_g % CS) / 2;
  const bool first_w = w == 0, last_w = w == WV - 1;
  auto sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  };

  double sa[NC][2], sq[NC][2], sb[NC][2], sr[NC][2];  // lambda and q of bins n, n+1
  auto fetch = [&](double (&sl)[NC][2], double (&sqq)[NC][2], int n) {
    const int nn = (n < d.ns && A.aloft.on(n)) ? n : max(0, min(n - 2, d.ns - 1));
    const long long bofs = d.ncell * (long long)nn;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const double2 l2 = *(const double2 *)(lam + base + pl[m] + bofs);
      const double2 q2 = *(const double2 *)(qpre + base + pf[m] + bofs);
      sl[m][0] = l2.x; sl[m][1] = l2.y;
      sqq[m][0] = q2.x; sqq[m][1] = q2.y;
    }
  };
  auto bin = [&](double (&sl)[NC][2], double (&sqq)[NC][2], int n) {
    const bool on = A.aloft.on(n);
    double *edge = s_edge + (n & 1) * 32;
    if (l == 2) {
      edge[(w * 2 + 0) * 2 + 0] = sl[0][0];
      edge[(w * 2 + 0) * 2 + 1] = sl[0][1];
    }
#pragma unroll
    for (int m = 0; m < NC; ++m)
      if (m == m_le && l == l_le) {
        edge[(w * 2 + 1) * 2 + 0] = sl[m][0];
        edge[(w * 2 + 1) * 2 + 1] = sl[m][1];
      }
    sync();
    double hlo[2] = {0.0, 0.0}, hhi[2] = {0.0, 0.0};
    if (has_lo) {
      hlo[0] = edge[((w - 1) * 2 + 1) * 2 + 0];
      hlo[1] = edge[((w - 1) * 2 + 1) * 2 + 1];
    }
    if (has_hi) {
      hhi[0] = edge[((w + 1) * 2 + 0) * 2 + 0];
      hhi[1] = edge[((w + 1) * 2 + 0) * 2 + 1];
    }
    double lm[NC][2], q[NC][2], mu[NC][2];
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const double lv = halo_lo[m] ? hlo[e] : (halo_hi[m] ? hhi[e] : (own[m] ? sl[m][e] : 0.0));
        lm[m][e] = act[m][e] ? lv : 0.0;
        q[m][e] = ld[m][e] ? sqq[m][e] : 0.0;
      }
    fetch(sl, sqq, n + 2);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e) mu[m][e] = act[m][e] ? -dtv[m][e] * lm[m][e] : 0.0;
    double um1[NC][2], up1[NC][2];
    row_prev<NC>(u, um1, lane0);
    row_next<NC>(u, up1, lane63);
    double qm2[NC][2], qm1[NC][2], qp1[NC][2], mum1[NC][2];
    row_prev2<NC>(q, qm2, lane0);
    row_prev<NC>(q, qm1, lane0);
    row_next<NC>(q, qp1, lane63);
    row_prev<NC>(mu, mum1, lane0);
    double e0[NC][2], e1[NC][2], e2[NC][2], e3[NC][2];
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int c = cc[m][e];
        e0[m][e] = e1[m][e] = e2[m][e] = e3[m][e] = 0.0;
        if (valid[m][e] && c >= rmin && c <= rmax + 1) {
          const double divp = fmax(0.0, u[m][e]) - fmax(0.0, um1[m][e]);
          const double divm = fmin(0.0, up1[m][e]) - fmin(0.0, u[m][e]);
          const double aus = fabs(u[m][e]);
          const double gI = 0.5 * aus * (1.0 - dtv[m][e] * aus);
          interface_adjoint(lim, qm2[m][e], qm1[m][e], q[m][e], qp1[m][e], u[m][e], divp, divm, gI, tfD[m][e],
                            tfDm1[m][e], mum1[m][e], mu[m][e], e0[m][e], e1[m][e], e2[m][e], e3[m][e]);
        }
      }
    // interfaces r+2, r+1, r, r-1 reach cell r as their I-2, I-1, I, I+1
    double e0n[NC][2], e1n[NC][2], e3p[NC][2], v[NC][2];
    row_next2<NC>(e0, e0n, lane63);
    row_next<NC>(e1, e1n, lane63);
    row_prev<NC>(e3, e3p, lane0);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < 2; ++e) v[m][e] = e0n[m][e] + e1n[m][e] + e2[m][e] + e3p[m][e];
    if (A.periodic) {
      // cells -1, 0: element 0, 1 of lane 2, chunk 0 of wave 0
      if (first_w && l == 2) {
        s_fold[0] = v[0][0];
        s_fold[1] = v[0][1];
      }
#pragma unroll
      for (int m = 0; m < NC; ++m)
        if (last_w && m == m_g && l == l_g) {
          s_fold[2] = v[m][0];
          s_fold[3] = v[m][1];
        }
      sync();
      const double gc = s_fold[0], gd = s_fold[1], ga = s_fold[2], gb = s_fold[3];
#pragma unroll
      for (int m = 0; m < NC; ++m)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int c = cc[m][e];
          if (c == 1) v[m][e] += ga;
          if (c == 2) v[m][e] += gb;
          if (c == nx - 1) v[m][e] += gc;
          if (c == nx) v[m][e] += gd;
        }
    }
    const long long bofs = d.ncell * (long long)n;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      double *dst = store[m] ? lam + base + pl[m] + bofs : junk;
      *(double2 *)dst = on ? make_double2(lm[m][0] + v[m][0], lm[m][1] + v[m][1]) : make_double2(0.0, 0.0);
    }
  };
  fetch(sa, sq, 0);
  fetch(sb, sr, 1);
  int n = 0;
  for (; n + 1 < d.ns; n += 2) {
    bin(sa, sq, n);
    bin(sb, sr, n + 1);
  }
  if (n < d.ns) bin(sa, sq, n);
}

// ---- theta-scheme lines: lambda' = B^T A^{-T} lambda -------------------
struct RowC {
  double a, b, c;     // B row: a q_{l-1} + b q_l + c q_{l+1}
  double dl, dg, du;  // A row: dl x_{l-1} + dg x_l + du x_{l+1}
};

template <int DIR, int NB>
__global__ __launch_bounds__(kBlock) void k_adj_cn(DiffArgs A, double *lam) {
  // the partitioned adjoint solves ran (launched before this kernel) when
  // every line of the sweep is a diagonally dominant M-matrix (uniform exit)
  if (A.sel == 2 && *A.ndom == 0) return;
  const Dims d = A.d;
  const LineGeom g = A.g;
  const TileGroup tg = xcd_tile_group(blockIdx.x, (d.ns + NB - 1) / NB);
  const long long gid = tg.tile * kBlock + threadIdx.x;
  if (gid >= (long long)g.Na * g.Nb) return;
  if (A.jline) {  // a line of the Jacobi sweeps' adjoint, which ran before this kernel
    const int sw = A.jline[gid];
    if (sw >= 1 && sw <= A.jcap) return;
  }
  const int a = (int)(gid % g.Na) + 1;
  const int b = (int)(gid / g.Na) + 1;
  const int n0 = tg.group * NB;
  const int N = g.N, rmin = g.rmin, rmax = g.rmax;
  const int n = rmax - rmin + 1;
  const long long sL = line_stride(d, DIR), base = line_base(d, DIR, a, b);
  const bool line_on = a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  double *uD = A.fac;  // shared by the bin groups (identical values)
  double *uDU = uD + d.ncell, *uDL = uD + 2 * d.ncell;
  double *sb = A.scratch_b;
  auto cell = [&](int r) { return base + sL * (long long)(r + 1); };
  auto row = [&](int l) { return cell(rmin - 1 + l); };
  bool on[NB], any_on = false;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    on[u] = (n0 + u) < d.ns && line_on && A.aloft.on(n0 + u);
    any_on = any_on || on[u];
  }
  if (any_on) {
    const double th = A.theta;
    // the row coefficients (:807-853) from the cached row factors L, R
    auto rowc = [&](int l) {
      RowC r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      if (l < 1 || l > n) return r;
      const long long o = cell(rmin - 1 + l);
      const double L = A.lr[o];
      const double R = A.lr[o + d.ncell];
      const double C = L + R;
      r.a = (1.0 - th) * L;
      r.b = 1.0 - (1.0 - th) * C;
      r.c = (1.0 - th) * R;
      r.dl = -(th * L);
      r.dg = (l == 1) ? 1.0 + th * C - th * L : (l == n ? 1.0 + th * C - th * R : 1.0 + th * C);
      r.du = -(th * R);
      return r;
    };
    // A^T y = lambda.  Row l of A^T: sub-diagonal du_{l-1}, diagonal dg_l,
    // super-diagonal dl_{l+1}.  Streamed elimination with partial pivoting.
    RowC rprev = rowc(1), rcur = rowc(2);
    double Dc = rprev.dg, DUc = rcur.dl, Bc[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) Bc[u] = on[u] ? lam[row(1) + d.ncell * (long long)(n0 + u)] : 0.0;
    for (int l = 2; l <= n; ++l) {
      const int i = l - 1;
      const RowC rnext = rowc(l + 1);
      const double DLn = rprev.du;
      double Dn = rcur.dg;
      double DUn = (l < n) ? rnext.dl : 0.0;
      double Bn[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) Bn[u] = on[u] ? lam[row(l) + d.ncell * (long long)(n0 + u)] : 0.0;
      double DLo = 0.0;
      if (fabs(Dc) >= fabs(DLn)) {
        const double fact = DLn / Dc;
        Dn = Dn - fact * DUc;
#pragma unroll
        for (int u = 0; u < NB; ++u) Bn[u] = Bn[u] - fact * Bc[u];
      } else {
        const double fact = Dc / DLn;
        const double temp = Dn;
        Dn = DUc - fact * temp;
        if (i <= n - 2) {
          DLo = DUn;
          DUn = -fact * DLo;
        }
        DUc = temp;
        Dc = DLn;
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const double tb = Bc[u];
          Bc[u] = Bn[u];
          Bn[u] = tb - fact * Bn[u];
        }
      }
      const long long o = row(i);
      uD[o] = Dc; uDU[o] = DUc; uDL[o] = DLo;
#pragma unroll
      for (int u = 0; u < NB; ++u)
        if (on[u]) sb[o + d.ncell * (long long)(n0 + u)] = Bc[u];
      Dc = Dn; DUc = DUn;
#pragma unroll
      for (int u = 0; u < NB; ++u) Bc[u] = Bn[u];
      rprev = rcur;
      rcur = rnext;
    }
    // back substitution downward; row i+1 of lambda' = B^T y is final once
    // y_i is known:
    //   lambda'_l = c_{l-1} y_{l-1} + (b_l + [l=1] a_l + [l=n] c_l) y_l + a_{l+1} y_{l+1}
    double y1[NB], y2[NB];  // y_{i+1}, y_{i+2}
#pragma unroll
    for (int u = 0; u < NB; ++u) { y1[u] = Bc[u] / Dc; y2[u] = 0.0; }
    RowC rhi{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // row i+2
    RowC rmid = rowc(n);                      // row i+1
    for (int i = n - 1; i >= 0; --i) {
      double yi[NB];
      RowC rlo{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      if (i >= 1) {
        const long long o = row(i);
        const double ud = uD[o], udu = uDU[o], udl = uDL[o];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const double Bi = on[u] ? sb[o + d.ncell * (long long)(n0 + u)] : 0.0;
          yi[u] = (i == n - 1) ? (Bi - udu * y1[u]) / ud : (Bi - udu * y1[u] - udl * y2[u]) / ud;
        }
        rlo = rowc(i);
      } else {
#pragma unroll
        for (int u = 0; u < NB; ++u) yi[u] = 0.0;
      }
      const int l = i + 1;
      const double bd = rmid.b + (l == 1 ? rmid.a : 0.0) + (l == n ? rmid.c : 0.0);
      const long long ol = row(l);
#pragma unroll
      for (int u = 0; u < NB; ++u)
        if (on[u]) lam[ol + d.ncell * (long long)(n0 + u)] = rlo.c * yi[u] + bd * y1[u] + rhi.a * y2[u];
#pragma unroll
      for (int u = 0; u < NB; ++u) { y2[u] = y1[u]; y1[u] = yi[u]; }
      rhi = rmid;
      rmid = rlo;
    }
  }
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    if (n0 + u >= d.ns) break;
    const long long bofs = d.ncell * (long long)(n0 + u);
    for (int r = 1; r <= N; ++r)
      if (!on[u] || r < rmin || r > rmax) lam[cell(r) + bofs] = 0.0;
  }
}

// ---- explicit diffusion lines -----------------------------------------
// q'_c = q_c + Ep_c (q_{c+1} - q_c) - Em_c (q_c - q_{c-1}), c in [rmin,rmax],
// Ep_c = dt kds_{c+1} sig_{c+1}/kappa_c, Em_c = dt kds_c sig_c/kappa_c.
template <int DIR>
__global__ __launch_bounds__(kBlock) void k_adj_diff_explicit(DiffArgs A, double *lam, int nb, int periodic) {
  const Dims d = A.d;
  const LineGeom g = A.g;
  const TileGroup tg = xcd_tile_group(blockIdx.x, (d.ns + nb - 1) / nb);
  const long long gid = tg.tile * kBlock + threadIdx.x;
  if (gid >= (long long)g.Na * g.Nb) return;
  const int a = (int)(gid % g.Na) + 1;
  const int b = (int)(gid / g.Na) + 1;
  const int N = g.N, rmin = g.rmin, rmax = g.rmax;
  const long long sL = line_stride(d, DIR), base = line_base(d, DIR, a, b);
  const long long kS = kline_stride(d, DIR), kB = kline_base(d, DIR, a, b);
  const bool line_on = a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  const double dt = A.dt;
  auto cell = [&](int r) { return base + sL * (long long)(r + 1); };
  auto kds_at = [&](int I) {
    const double ds = 2.0 * A.sig[cell(I)] / (A.kap[cell(I - 1)] + A.kap[cell(I)]);
    return 0.5 * (A.kk[kB + kS * (long long)(I - 1)] + A.kk[kB + kS * (long long)I]) * ds;
  };
  for (int n = tg.group * nb; n < d.ns && n < (tg.group + 1) * nb; ++n) {
    const long long bofs = d.ncell * (long long)n;
    if (!(line_on && A.aloft.on(n))) {
      for (int r = 1; r <= N; ++r) lam[cell(r) + bofs] = 0.0;
      continue;
    }
    auto act = [&](int c) { return c >= rmin && c <= rmax; };
    auto lv = [&](int c) { return act(c) ? lam[cell(c) + bofs] : 0.0; };
    auto epv = [&](int c) { return act(c) ? dt * kds_at(c + 1) * (A.sig[cell(c + 1)] / A.kap[cell(c)]) : 0.0; };
    auto emv = [&](int c) { return act(c) ? dt * kds_at(c) * (A.sig[cell(c)] / A.kap[cell(c)]) : 0.0; };
    for (int r = 1; r < rmin - 1; ++r) lam[cell(r) + bofs] = 0.0;
    // sliding over r: lambda and Ep at r-1, lambda, Ep, Em at r, lambda and Em at r+1
    double l_m1 = lv(rmin - 2), l_0 = lv(rmin - 1), l_p1 = lv(rmin);
    double ep_m1 = epv(rmin - 2), ep_0 = epv(rmin - 1);
    double em_0 = emv(rmin - 1), em_p1 = emv(rmin);
    double glo = 0.0, ghi = 0.0;  // adjoints of ghost cells 0 and N+1
    for (int r = rmin - 1; r <= rmax + 1; ++r) {
      const double v = (1.0 - ep_0 - em_0) * l_0 + ep_m1 * l_m1 + em_p1 * l_p1;
      if (r == 0) glo = v;
      else if (r == N + 1) ghi = v;
      else lam[cell(r) + bofs] = v;  // r = rmin-1, rmax+1 (inactive) keep only the stencil terms
      l_m1 = l_0; l_0 = l_p1; l_p1 = lv(r + 2);
      ep_m1 = ep_0; ep_0 = epv(r + 1);
      em_0 = em_p1; em_p1 = emv(r + 2);
    }
    for (int r = rmax + 2; r <= N; ++r) lam[cell(r) + bofs] = 0.0;
    // Neumann ghosts: 0 <- cell 1, N+1 <- cell N; periodic x: 0 <- N, N+1 <- 1
    if (rmin == 1) lam[cell(periodic ? N : 1) + bofs] += glo;
    if (rmax == N) lam[cell(periodic ? 1 : N) + bofs] += ghi;
  }
}

// L(i,j) = sum_k (sum_n c(i,j,k,n)) * dz(k) / 1e3 (src/Output_Vars.f90:1061-1066)
__global__ void k_column_load(Dims d, const double *c0, const double *dz, double *out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)d.nx * d.ny) return;
  const int i = (int)(t % d.nx) + 1, j = (int)(t / d.nx) + 1;
  double L = 0.0;
  for (int k = 1; k <= d.nz; ++k) {
    double s = 0.0;
    for (int n = 0; n < d.ns; ++n) s += c0[d.p3(i, j, k) + d.ncell * (long long)n];
    L = L + s * dz[k - 1] / 1.0e3;
  }
  out[t] = L;
}

// lambda(i,j,k,n) += w(i,j) * dz(k) / 1e3
__global__ void k_adj_column(Dims d, double *lam, const double *dz, const double *w) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)d.nx * d.ny * d.nz) return;
  const int i = (int)(t % d.nx) + 1, j = (int)((t / d.nx) % d.ny) + 1;
  const int k = (int)(t / ((long long)d.nx * d.ny)) + 1;
  const double wk = w[(long long)(i - 1) + (long long)d.nx * (j - 1)] * dz[k - 1] / 1.0e3;
  for (int n = 0; n < d.ns; ++n) {
    const long long o = d.p3(i, j, k) + d.ncell * (long long)n;
    lam[o] = lam[o] + wk;
  }
}

// slot(k,n) += dt * lambda(iv,jv,k,n), k = 1..nz+1
__global__ void k_adj_source(Dims d, const double *lam, int iv, int jv, double dt, double *slot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (d.nz + 1) * d.ns) return;
  const int k = t % (d.nz + 1) + 1, n = t / (d.nz + 1);
  slot[t] = slot[t] + dt * lam[d.p3(iv, jv, k) + d.ncell * (long long)n];
}

template <int CPT>
hipError_t adj_x_rows_cpt(hipStream_t s, const AdvArgs &a, const double *qpre, double *lam, int bt) {
  const size_t shmem = sizeof(double) * 6 * (size_t)(a.d.nx + 4);
  dim3 grid((unsigned)((long long)a.d.ny * a.d.nz));
  const void *kern = a.deld != nullptr ? (const void *)k_adj_advect_x_rows<CPT, true>
                                       : (const void *)k_adj_advect_x_rows<CPT, false>;
  if (shmem > 65536) {
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess) return e;
  }
  if (a.deld != nullptr)
    hipLaunchKernelGGL((k_adj_advect_x_rows<CPT, true>), grid, dim3(bt), shmem, s, a, qpre, lam);
  else
    hipLaunchKernelGGL((k_adj_advect_x_rows<CPT, false>), grid, dim3(bt), shmem, s, a, qpre, lam);
  return hipGetLastError();
}

template <int NC>
hipError_t adj_x_wave(hipStream_t s, const AdvArgs &a, const double *qpre, double *lam) {
  const long long rows = (long long)a.d.ny * a.d.nz;
  dim3 grid((unsigned)((rows + 3) / 4));
  if (a.deld != nullptr)
    hipLaunchKernelGGL((k_adj_advect_x_wave<NC, true>), grid, dim3(256), 0, s, a, qpre, lam);
  else
    hipLaunchKernelGGL((k_adj_advect_x_wave<NC, false>), grid, dim3(256), 0, s, a, qpre, lam);
  return hipGetLastError();
}

// Segmented rows: WV = ceil(W / (128 NC - 8)) waves, L the even segment
// length; hipErrorInvalidValue (nothing launched) for rows that need more
// than 8 segments.
template <int NC>
hipError_t adj_x_seg(hipStream_t s, const AdvArgs &a, const double *qpre, double *lam) {
  const int W = a.d.nx + 4;
  const int cap = 128 * NC - 8;
  const int WV = (W + cap - 1) / cap;
  if (WV > 8) return hipErrorInvalidValue;
  int L = (W + WV - 1) / WV;
  L += L & 1;
  // the last segment holds cells nx-1 .. nx+2 (the periodic fold's targets
  // and sources) and every segment its first and last pair
  if (L > cap || L < 4 || (WV - 1) * L > W - 4) return hipErrorInvalidValue;
  dim3 grid((unsigned)((long long)a.d.ny * a.d.nz));
  if (a.deld != nullptr)
    hipLaunchKernelGGL((k_adj_advect_x_seg<NC, true>), grid, dim3(64 * WV), 0, s, a, qpre, lam, L);
  else
    hipLaunchKernelGGL((k_adj_advect_x_seg<NC, false>), grid, dim3(64 * WV), 0, s, a, qpre, lam, L);
  return hipGetLastError();
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

template <int DIR>
hipError_t adj_lines(hipStream_t s, const AdvArgs &a, const double *qpre, double *lam, int nb) {
  const long long lines = (long long)a.g.Na * a.g.Nb;
  dim3 grid((unsigned)(padded_tiles((lines + kBlock - 1) / kBlock) * ((a.d.ns + nb - 1) / nb)));
  if (DIR != 2 && a.deld != nullptr)
    hipLaunchKernelGGL((k_adj_advect_lines<DIR, true>), grid, dim3(kBlock), 0, s, a, qpre, lam, nb);
  else
    hipLaunchKernelGGL((k_adj_advect_lines<DIR, false>), grid, dim3(kBlock), 0, s, a, qpre, lam, nb);
  return hipGetLastError();
}

template <int DIR, int NB>
hipError_t adj_seg_dir(hipStream_t s, const AdvArgs &a_in, const double *qpre, double *lam) {
  AdvArgs a = a_in;
  const int ntile = (a.g.Na + 63) / 64;
  const long long groups = (a.d.ns + NB - 1) / NB;
  dim3 grid, block;
  if (a.gy > 0) {
    // bin groups along y, at most 4 per workgroup (the launch bounds)
    a.gy = (int)std::min<long long>(std::min(a.gy, 4), groups);
    grid = dim3((unsigned)(padded_tiles((long long)ntile * a.g.Nb) * ((groups + a.gy - 1) / a.gy)));
    block = dim3(64, a.gy);
  } else {
    // segments per line as for the forward sweep: enough waves to fill the
    // chip, segments of at least 16 cells, at most 4
    const long long lanes = (long long)ntile * 64 * a.g.Nb * groups;
    int W = (int)((1LL << 21) / (lanes > 0 ? lanes : 1));
    W = W < 1 ? 1 : (W > 4 ? 4 : W);
    while (W > 1 && a.g.N / W < 16) --W;
    grid = dim3((unsigned)(padded_tiles((long long)ntile * a.g.Nb) * groups));
    block = dim3(64, W);
  }
  const bool topo = DIR == 1 && a.deld != nullptr;
  if (a.gy > 0) {
    if (topo) hipLaunchKernelGGL((k_adj_advect_seg<DIR, NB, true, true>), grid, block, 0, s, a, qpre, lam);
    else hipLaunchKernelGGL((k_adj_advect_seg<DIR, NB, false, true>), grid, block, 0, s, a, qpre, lam);
  } else {
    const size_t halo = sizeof(double) * 4 * NB * (size_t)block.x * block.y;
    if (topo) hipLaunchKernelGGL((k_adj_advect_seg<DIR, NB, true, false>), grid, block, halo, s, a, qpre, lam);
    else hipLaunchKernelGGL((k_adj_advect_seg<DIR, NB, false, false>), grid, block, halo, s, a, qpre, lam);
  }
  return hipGetLastError();
}

template <int DIR>
hipError_t adj_seg(hipStream_t s, const AdvArgs &a, const double *qpre, double *lam, int nb) {
  switch (nb) {
    case 1: return adj_seg_dir<DIR, 1>(s, a, qpre, lam);
    case 2: return adj_seg_dir<DIR, 2>(s, a, qpre, lam);
    case 3: return adj_seg_dir<DIR, 3>(s, a, qpre, lam);
    default: return adj_seg_dir<DIR, 4>(s, a, qpre, lam);
  }
}

template <int DIR, int NB>
hipError_t adj_cn_dir(hipStream_t s, const DiffArgs &a, double *lam) {
  const long long lines = (long long)a.g.Na * a.g.Nb;
  dim3 grid((unsigned)(padded_tiles((lines + kBlock - 1) / kBlock) * ((a.d.ns + NB - 1) / NB)));
  hipLaunchKernelGGL((k_adj_cn<DIR, NB>), grid, dim3(kBlock), 0, s, a, lam);
  return hipGetLastError();
}

template <int DIR>
hipError_t adj_cn_nb(hipStream_t s, const DiffArgs &a, double *lam, int nb) {
  switch (nb) {
    case 1: return adj_cn_dir<DIR, 1>(s, a, lam);
    case 2: return adj_cn_dir<DIR, 2>(s, a, lam);
    case 4: return adj_cn_dir<DIR, 4>(s, a, lam);
    default: return adj_cn_dir<DIR, 8>(s, a, lam);
  }
}

}  // namespace

// x rows: even rows in wavefront registers (one wavefront per row up to
// 380 cells, 252 with topography; beyond, up to 8 segments of 248 cells),
// other rows through LDS (rows of up to 8*256-4 cells), longer
// rows streamed per lane; y and z lines in segments; a.periodic selects
// the periodic x fold.
hipError_t launch_adj_advect(hipStream_t s, const AdvArgs &a, const double *qpre, double *lam, int nb) {
  if (a.g.dir == 0) {
    const int cells = a.d.nx + 4;
    // one wavefront per row under the forward's conditions, longer rows in
    // segments of one wavefront each
    const bool pairs = a.xrows == 1 && a.d.nx % 2 == 0 && aligned16(lam) && aligned16(qpre) && aligned16(a.vel) &&
                       aligned16(a.sig) && aligned16(a.kap) && aligned16(a.junk) &&
                       (a.deld == nullptr || aligned16(a.deld));
    if (pairs && cells <= (a.deld ? 256 : 384)) {
      switch ((cells + 127) / 128) {
        case 1: return adj_x_wave<1>(s, a, qpre, lam);
        case 2: return adj_x_wave<2>(s, a, qpre, lam);
        default: return adj_x_wave<3>(s, a, qpre, lam);
      }
    }
    if (pairs) {
      // two chunks per wavefront: 188-200 VGPRs (three spill)
      hipError_t e = adj_x_seg<2>(s, a, qpre, lam);
      if (e != hipErrorInvalidValue) return e;
    }
    int cpt = (cells + 255) / 256;
    if (cpt > 8) return adj_lines<0>(s, a, qpre, lam, nb);
    if (cpt == 3) cpt = 4;
    if (cpt > 4) cpt = 8;
    int bt = (cells + cpt - 1) / cpt;
    bt = ((bt + 63) / 64) * 64;
    switch (cpt) {
      case 1: return adj_x_rows_cpt<1>(s, a, qpre, lam, bt);
      case 2: return adj_x_rows_cpt<2>(s, a, qpre, lam, bt);
      case 4: return adj_x_rows_cpt<4>(s, a, qpre, lam, bt);
      default: return adj_x_rows_cpt<8>(s, a, qpre, lam, bt);
    }
  }
  if (a.g.dir == 1) return adj_seg<1>(s, a, qpre, lam, nb);
  return adj_seg<2>(s, a, qpre, lam, nb);
}

// a.g.dir is 1 or 2 (x lines run on the transposed grid as dir 1)
hipError_t launch_adj_cn(hipStream_t s, const DiffArgs &a, double *lam, int nb) {
  if (a.g.dir == 2) return adj_cn_nb<2>(s, a, lam, nb);
  return adj_cn_nb<1>(s, a, lam, nb);
}

hipError_t launch_adj_diff_explicit(hipStream_t s, const DiffArgs &a, double *lam, int nb, int periodic) {
  const long long lines = (long long)a.g.Na * a.g.Nb;
  dim3 grid((unsigned)(padded_tiles((lines + kBlock - 1) / kBlock) * ((a.d.ns + nb - 1) / nb)));
  if (a.g.dir == 2)
    hipLaunchKernelGGL((k_adj_diff_explicit<2>), grid, dim3(kBlock), 0, s, a, lam, nb, periodic);
  else
    hipLaunchKernelGGL((k_adj_diff_explicit<1>), grid, dim3(kBlock), 0, s, a, lam, nb, periodic);
  return hipGetLastError();
}

hipError_t launch_column_load(hipStream_t s, const Dims &d, const double *c0, const double *dz, double *out) {
  const long long t = (long long)d.nx * d.ny;
  hipLaunchKernelGGL(k_column_load, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, d, c0, dz, out);
  return hipGetLastError();
}

hipError_t launch_adj_column(hipStream_t s, const Dims &d, double *lam, const double *dz, const double *w) {
  const long long t = (long long)d.nx * d.ny * d.nz;
  hipLaunchKernelGGL(k_adj_column, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, d, lam, dz, w);
  return hipGetLastError();
}

hipError_t launch_adj_source(hipStream_t s, const Dims &d, const double *lam, int iv, int jv, double dt,
                             double *slot) {
  const int t = (d.nz + 1) * d.ns;
  hipLaunchKernelGGL(k_adj_source, dim3((t + 255) / 256), dim3(256), 0, s, d, lam, iv, jv, dt, slot);
  return hipGetLastError();
}

}  // namespace ash3d
