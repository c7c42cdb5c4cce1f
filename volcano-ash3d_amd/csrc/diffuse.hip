// Diffusion sweeps: explicit diff_{x,y,z} (src/Diffusion.F90:266-650) and
// the theta-scheme diffCN_{x,y,z} (src/Diffusion.F90:671-1428) with the
// tridiagonal solve of LAPACK dgtsv (Gaussian elimination with partial
// pivoting), which the reference calls once per line and bin.
//
// One thread owns one line for NB bins.  The matrix of a line does not
// depend on the bin, so its coefficients (three divisions per cell) and its
// LU factorisation are formed once per line and applied to the NB
// right-hand sides; the elimination is streamed along the line in exactly
// dgtsv's order, the U factors go to a per-line scratch and the eliminated
// right-hand sides to the ts1 slot (which the reference zero-fills anyway),
// and the back substitution streams the line in reverse writing ts0.
// No contraction (-ffp-contract=off): bitwise the reference arithmetic.
#include "common.hpp"

namespace ash3d {

namespace {

constexpr int kBlock = 256;

// Row factors of the theta-scheme lines, one thread per cell (lanes along
// the contiguous transverse index).  For the row of cell c (:807-816):
//   kds_I = (0.5*(K_{I-1}+K_I)) * (sig_I / (0.5*(kap_{I-1}+kap_I)))
//   L = dt*kds_c*sig_c/kap_c,  R = dt*kds_{c+1}*sig_{c+1}/kap_c
// They depend on dt, the geometry and K only, so the context keeps them
// between steps; the line solves then read two numbers per row instead of
// evaluating three divisions in every bin group.
template <int DIR>
__global__ __launch_bounds__(kBlock) void k_line_lr(DiffArgs A, double *lr) {
  const Dims d = A.d;
  const LineGeom g = A.g;
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long long)g.Na * g.N * g.Nb) return;
  const int a = (int)(t % g.Na) + 1;
  const long long rest = t / g.Na;
  const int r = (int)(rest % g.N) + 1;
  const int b = (int)(rest / g.N) + 1;
  const long long sL = line_stride(d, DIR), base = line_base(d, DIR, a, b);
  const long long kS = kline_stride(d, DIR), kB = kline_base(d, DIR, a, b);
  auto kapat = [&](int q) { return A.kap[base + sL * (long long)(q + 1)]; };
  auto Kat = [&](int q) { return A.kk[kB + kS * (long long)q]; };
  auto sigat = [&](int q) { return A.sig[base + sL * (long long)(q + 1)]; };
  auto kds_at = [&](int I) {
    const double vavg = 0.5 * (kapat(I - 1) + kapat(I));
    const double kavg = 0.5 * (Kat(I - 1) + Kat(I));
    const double ds = sigat(I) / vavg;
    return kavg * ds;
  };
  const double dt = A.dt, kap_c = kapat(r);
  const long long o = base + sL * (long long)(r + 1);
  lr[o] = dt * kds_at(r) * sigat(r) / kap_c;
  lr[o + d.ncell] = dt * kds_at(r + 1) * sigat(r + 1) / kap_c;
}

template <int DIR, int NB>
__global__ __launch_bounds__(kBlock) void k_diff_explicit(DiffArgs A) {
  const Dims d = A.d;
  const LineGeom g = A.g;
  const TileGroup tg = xcd_tile_group(blockIdx.x, (d.ns + NB - 1) / NB);
  const long long gid = tg.tile * kBlock + threadIdx.x;
  if (gid >= (long long)g.Na * g.Nb) return;
  const int a = (int)(gid % g.Na) + 1;
  const int b = (int)(gid / g.Na) + 1;
  const int n0 = tg.group * NB;
  const int N = g.N, rmin = g.rmin, rmax = g.rmax;
  const long long sL = line_stride(d, DIR), base = line_base(d, DIR, a, b);
  const long long kS = kline_stride(d, DIR), kB = kline_base(d, DIR, a, b);
  const bool line_on = a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  double *c = A.c0;
  bool exists[NB], on[NB];
  bool any_on = false;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    exists[u] = (n0 + u) < d.ns;
    on[u] = exists[u] && line_on && A.aloft.on(n0 + u);
    any_on = any_on || on[u];
  }
  if (any_on) {
    const double dt = A.dt;
    // window: q_{r-1}, kap_{r-1}, K_{r-1}; interface r-1 quantities
    double qm[NB], dqm[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) qm[u] = dqm[u] = 0.0;
    double kapm = 1.0, Km = 0.0, sigm = 0.0, Am = 0.0;
    for (int r = rmin - 1; r <= rmax + 1; ++r) {
      const long long cell = base + sL * (long long)(r + 1);
      const double kap0 = A.kap[cell];
      const double K0 = A.kk[kB + kS * (long long)r];
      double q0[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) q0[u] = exists[u] ? c[cell + d.ncell * (long long)(n0 + u)] : 0.0;
      if (r >= rmin) {
        // interface I = r (:357-359)
        const double sig0 = A.sig[cell];
        const double ds = 2.0 * sig0 / (kapm + kap0);
        const double kds = 0.5 * (Km + K0) * ds;
        const double A0 = dt * kds;
        const int cc = r - 1;
        if (cc >= rmin) {
          // cell cc = r-1 (:367-372): interfaces cc (=r-1) and cc+1 (=r)
          const double sR = sig0 / kapm;  // sig_{c+1}/kap_c
          const double sLf = sigm / kapm; // sig_c/kap_c
#pragma unroll
          for (int u = 0; u < NB; ++u) {
            const double dq0 = q0[u] - qm[u];
            if (on[u]) {
              const double LF = A0 * dq0 * sR;
              const double RF = -(Am * dqm[u] * sLf);
              c[base + sL * (long long)(cc + 1) + d.ncell * (long long)(n0 + u)] = qm[u] + (LF + RF);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) dqm[u] = q0[u] - qm[u];
        Am = A0;
        sigm = sig0;
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) qm[u] = q0[u];
      kapm = kap0;
      Km = K0;
    }
  }
  // zero fill of ts1 then copy (:314, :385-386)
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
  }
}

// theta-scheme line solve.  Row l (cell c = rmin-1+l) of the system
// (:807-853):
//   L = dt*kds_c*sig_c/kap_c, R = dt*kds_{c+1}*sig_{c+1}/kap_c, C = L+R
//   first row: D = (1+th*C) - th*L, DU = -(th*R)
//   interior:  DL = -(th*L), D = 1+th*C, DU = -(th*R)
//   last row:  DL = -(th*L), D = (1+th*C) - th*R
//   B = ((1-th)*L*q_left + (1-(1-th)*C)*q_c) + (1-th)*R*q_right
// with q_left = q_c on the first row and q_right = q_c on the last.
template <int DIR, int NB>
__global__ __launch_bounds__(kBlock) void k_diff_cn(DiffArgs A) {
  constexpr int PF = 4;  // rows in flight per stream
  const Dims d = A.d;
  const LineGeom g = A.g;
  const TileGroup tg = xcd_tile_group(blockIdx.x, (d.ns + NB - 1) / NB);
  const long long gid = tg.tile * kBlock + threadIdx.x;
  if (gid >= (long long)g.Na * g.Nb) return;
  const int a = (int)(gid % g.Na) + 1;
  const int b = (int)(gid / g.Na) + 1;
  const int n0 = tg.group * NB;
  const int N = g.N, rmin = g.rmin, rmax = g.rmax;
  const int n = rmax - rmin + 1;  // nlineq
  const long long sL = line_stride(d, DIR), base = line_base(d, DIR, a, b);
  const long long kS = kline_stride(d, DIR), kB = kline_base(d, DIR, a, b);
  const bool line_on = a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  double *c = A.c0;
  double *sb = A.scratch_b;
  // per-line U factors: D, DU, DL (2nd super-diagonal).  They depend on the
  // line only, so every bin group writes the same values to the same slab
  // (identical bits whatever the interleaving) and the groups of a tile,
  // which run together on one XCD, share it through L2.
  double *fD = A.fac;
  double *fDU = fD + d.ncell;
  double *fDL = fDU + d.ncell;
  bool exists[NB], on[NB];
  bool any_on = false;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    exists[u] = (n0 + u) < d.ns;
    on[u] = exists[u] && line_on && A.aloft.on(n0 + u);
    any_on = any_on || on[u];
  }
  if (any_on) {
    const double th = A.theta;
    auto qat = [&](int u, int r) { return c[base + sL * (long long)(r + 1) + d.ncell * (long long)(n0 + u)]; };
    const double *Lf = A.lr, *Rf = A.lr + d.ncell;
    double q_l[NB], q_c[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      q_c[u] = exists[u] ? qat(u, rmin) : 0.0;
      q_l[u] = q_c[u];  // BC_left_t0 = q(rmin) (:822)
    }
    // elimination state: current row i (modified D, DU, B) and the
    // original sub-diagonal entry that row i+1 carries
    double Dc = 0.0, DUc = 0.0, Bc[NB];
    int singular = 0;
    // Row l needs its factors L, R (cell c = rmin-1+l) and the q of its
    // right neighbour (cell c+1, while it is active).  A ring of PF rows is
    // kept in flight: slot p holds row l0+p of the current block of PF rows
    // and is refilled with row l+PF when row l consumes it.
    double rL[PF], rR[PF], rq[PF][NB];
    auto fetch_row_in = [&](int slot, int l) {
      const long long o = base + sL * (long long)(rmin + l);  // cell rmin-1+l
      rL[slot] = Lf[o];
      rR[slot] = Rf[o];
#pragma unroll
      for (int u = 0; u < NB; ++u) rq[slot][u] = (l < n && exists[u]) ? qat(u, rmin + l) : 0.0;
    };
#pragma unroll
    for (int p = 0; p < PF; ++p)
      if (1 + p <= n) fetch_row_in(p, 1 + p);
    for (int l0 = 1; l0 <= n; l0 += PF) {
#pragma unroll
    for (int p = 0; p < PF; ++p) {
      const int l = l0 + p;
      if (l > n) break;
      const double LeftFac = rL[p], RightFac = rR[p];
      double q_next[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) q_next[u] = rq[p][u];
      if (l + PF <= n) fetch_row_in(p, l + PF);
      const double CenterFac = LeftFac + RightFac;
      double Dn, DUn = 0.0, DLn = 0.0;
      if (l == 1) {
        Dn = 1.0 + (th)*CenterFac - (th)*LeftFac;
        DUn = -((th)*RightFac);
      } else if (l < n) {
        DLn = -((th)*LeftFac);
        Dn = 1.0 + (th)*CenterFac;
        DUn = -((th)*RightFac);
      } else {
        DLn = -((th)*LeftFac);
        Dn = 1.0 + (th)*CenterFac - (th)*RightFac;
      }
      double Bn[NB], q_r[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        q_r[u] = (l < n) ? q_next[u] : q_c[u];  // BC_right_t0 (:845)
        Bn[u] = (1.0 - th) * LeftFac * q_l[u] + (1.0 - (1.0 - th) * CenterFac) * q_c[u] +
                (1.0 - th) * RightFac * q_r[u];
      }
      if (l == 1) {
        Dc = Dn;
        DUc = DUn;
#pragma unroll
        for (int u = 0; u < NB; ++u) Bc[u] = Bn[u];
      } else {
        // dgtsv elimination step i = l-1 with row i+1 = l (NRHS = 1 path)
        const int i = l - 1;
        const long long fo = base + sL * (long long)(rmin - 1 + i + 1);  // cell of row i
        double DLo;  // 2nd super-diagonal of U in row i
        if (fabs(Dc) >= fabs(DLn)) {
          if (Dc == 0.0) { if (!singular) singular = i; }
          const double fact = DLn / Dc;
          Dn = Dn - fact * DUc;
#pragma unroll
          for (int u = 0; u < NB; ++u) Bn[u] = Bn[u] - fact * Bc[u];
          DLo = (i <= n - 2) ? 0.0 : DLn;
        } else {
          const double fact = Dc / DLn;
          const double newDi = DLn;
          const double temp = Dn;
          Dn = DUc - fact * temp;
          if (i <= n - 2) {
            DLo = DUn;
            DUn = -fact * DLo;
          } else {
            DLo = DLn;
          }
          DUc = temp;
          Dc = newDi;
#pragma unroll
          for (int u = 0; u < NB; ++u) {
            const double tb = Bc[u];
            Bc[u] = Bn[u];
            Bn[u] = tb - fact * Bn[u];
          }
        }
        // row i is final: store its U entries and eliminated RHS
        fD[fo] = Dc;
        fDU[fo] = DUc;
        fDL[fo] = DLo;
#pragma unroll
        for (int u = 0; u < NB; ++u)
          if (on[u]) sb[fo + d.ncell * (long long)(n0 + u)] = Bc[u];
        Dc = Dn;
        DUc = DUn;
#pragma unroll
        for (int u = 0; u < NB; ++u) Bc[u] = Bn[u];
      }
      // advance the window
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        q_l[u] = q_c[u];
        q_c[u] = q_r[u];
      }
    }
    }
    if (Dc == 0.0 && !singular) singular = n;
    if (singular) atomicCAS(A.flag, 0, singular);
    // ---- back substitution (dgtsv, NRHS <= 2 path), row n first
    double x1[NB], x2[NB];
    {
      const double Dn_ = Dc;
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        x1[u] = Bc[u] / Dn_;
        x2[u] = 0.0;
        if (on[u]) c[base + sL * (long long)(rmax + 1) + d.ncell * (long long)(n0 + u)] = x1[u];
      }
    }
    // rows' factors and eliminated right-hand sides, PF rows in flight
    // (slot p holds row i0-p of the current block)
    double bD[PF], bDU[PF], bDL[PF], bB[PF][NB];
    auto fetch_row = [&](int slot, int i) {
      const long long fo = base + sL * (long long)(rmin - 1 + i + 1);
      bD[slot] = fD[fo];
      bDU[slot] = fDU[fo];
      bDL[slot] = fDL[fo];
#pragma unroll
      for (int u = 0; u < NB; ++u) bB[slot][u] = on[u] ? sb[fo + d.ncell * (long long)(n0 + u)] : 0.0;
    };
#pragma unroll
    for (int p = 0; p < PF; ++p)
      if (n - 1 - p >= 1) fetch_row(p, n - 1 - p);
    for (int i0 = n - 1; i0 >= 1; i0 -= PF) {
#pragma unroll
    for (int p = 0; p < PF; ++p) {
      const int i = i0 - p;
      if (i < 1) break;
      const long long fo = base + sL * (long long)(rmin - 1 + i + 1);
      const double Di = bD[p], DUi = bDU[p], DLi = bDL[p];
      double Bcur[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) Bcur[u] = bB[p][u];
      if (i - PF >= 1) fetch_row(p, i - PF);
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const double Bi = Bcur[u];
        double xi;
        if (i == n - 1) xi = (Bi - DUi * x1[u]) / Di;
        else xi = (Bi - DUi * x1[u] - DLi * x2[u]) / Di;
        x2[u] = x1[u];
        x1[u] = xi;
        if (on[u]) c[fo + d.ncell * (long long)(n0 + u)] = xi;
      }
    }
    }
  }
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
  }
}

// The x-line solve without transposing the state.  Lines run along x,
// which is contiguous in memory, so a lane per line would read a different
// row per lane.  One workgroup of 64 lanes owns 64 lines (consecutive j at
// one k); the state is staged through LDS in tiles of TX cells of all 64
// rows (each tile row is a coalesced 8*TX-byte segment), in both
// directions: q tiles are loaded ahead of the elimination, solution tiles
// are stored behind the back substitution.  Geometry, K, the U factors and
// the eliminated right-hand sides use the x<->y transposed layout (cached
// transposed kappa, sigma_nx, kx; scratch), where consecutive lanes are
// consecutive addresses.  The arithmetic is k_diff_cn's, row for row.
struct XState {
  Dims d;     // original dims
  double *c;  // concen ts0, original layout
};

template <int NB, int TX>
__global__ __launch_bounds__(64) void k_diff_cn_xrows(DiffArgs A, XState X) {
  constexpr int PF = 4;  // row factors in flight; TX cells per LDS tile
  __shared__ double s_t[NB][64][TX + 1];
  __shared__ long long s_row[64];          // original offset of cell -1 of each lane's line (-1: none)
  __shared__ unsigned char s_on[64];       // line_on of each lane
  const Dims d = A.d;  // transposed dims
  const Dims o = X.d;
  const LineGeom g = A.g;
  const int G = (d.ns + NB - 1) / NB;
  const TileGroup tg = xcd_tile_group(blockIdx.x, G);
  const long long lines = (long long)g.Na * g.Nb;
  if (tg.tile * 64 >= lines) return;  // padding workgroup (uniform)
  const int t = threadIdx.x;
  const long long gid = tg.tile * 64 + t;
  const bool lane_ok = gid < lines;
  const int a = lane_ok ? (int)(gid % g.Na) + 1 : 1;  // j
  const int b = lane_ok ? (int)(gid / g.Na) + 1 : 1;  // k
  const int n0 = tg.group * NB;
  const int N = g.N, rmin = g.rmin, rmax = g.rmax;
  const int n = rmax - rmin + 1;
  const long long sL = line_stride(d, 1), base = line_base(d, 1, a, b);
  const long long kS = kline_stride(d, 1), kB = kline_base(d, 1, a, b);
  const bool line_on = lane_ok && a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  const long long obase = o.p3(-1, a, b);
  double *c = X.c;
  double *sb = A.scratch_b;
  double *fD = A.fac;
  double *fDU = fD + d.ncell;
  double *fDL = fDU + d.ncell;
  bool exists[NB], bin_on[NB], on[NB];
  bool any_on = false, any_bin = false;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    exists[u] = (n0 + u) < d.ns;
    bin_on[u] = exists[u] && A.aloft.on(n0 + u);
    on[u] = bin_on[u] && line_on;
    any_on = any_on || on[u];
    any_bin = any_bin || bin_on[u];
  }
  s_row[t] = lane_ok ? obase : -1;
  s_on[t] = line_on ? 1 : 0;
  __syncthreads();
  // the block works if one of its lines is active for one of its bins
  const bool block_on = __syncthreads_or(any_on ? 1 : 0) != 0;

  // q of cells t0 .. t0+TX-1 (active cells only) of all 64 lines
  auto load_tile = [&](int t0) {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      if (!bin_on[u]) continue;  // uniform
#pragma unroll 4
      for (int m = 0; m < TX; ++m) {
        const int f = m * 64 + t, row = f / TX, col = f % TX, cell = t0 + col;
        const long long rb = s_row[row];
        s_t[u][row][col] = (rb >= 0 && cell >= rmin && cell <= rmax)
                               ? c[rb + (cell + 1) + o.ncell * (long long)(n0 + u)]
                               : 0.0;
      }
    }
  };
  // store the solution of cells lo .. lo+TX-1 (active cells of active lines)
  auto store_tile = [&](int lo) {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      if (!bin_on[u]) continue;  // uniform
#pragma unroll 4
      for (int m = 0; m < TX; ++m) {
        const int f = m * 64 + t, row = f / TX, col = f % TX, cell = lo + col;
        const long long rb = s_row[row];
        if (rb >= 0 && s_on[row] && cell >= rmin && cell <= rmax)
          c[rb + (cell + 1) + o.ncell * (long long)(n0 + u)] = s_t[u][row][col];
      }
    }
  };

  if (block_on && n > 1) {
    const double th = A.theta;
    const double *Lf = A.lr, *Rf = A.lr + d.ncell;
    load_tile(rmin);
    __syncthreads();
    int tlo = rmin;  // first cell of the tile in LDS
    double q_l[NB], q_c[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      q_c[u] = s_t[u][t][0];
      q_l[u] = q_c[u];  // BC_left_t0 = q(rmin) (:822)
    }
    double Dc = 0.0, DUc = 0.0, Bc[NB];
    int singular = 0;
    double rL[PF], rR[PF];
    auto fetch_lr = [&](int slot, int l) {
      const long long o = base + sL * (long long)(rmin + l);  // cell rmin-1+l
      rL[slot] = Lf[o];
      rR[slot] = Rf[o];
    };
#pragma unroll
    for (int p = 0; p < PF; ++p)
      if (1 + p <= n) fetch_lr(p, 1 + p);
    for (int l0 = 1; l0 <= n; l0 += PF) {
#pragma unroll
      for (int p = 0; p < PF; ++p) {
        const int l = l0 + p;
        if (l > n) break;
        const double LeftFac = rL[p], RightFac = rR[p];
        if (l + PF <= n) fetch_lr(p, l + PF);
        double q_next[NB];
        if (l < n) {
          if (rmin + l >= tlo + TX) {  // next tile (uniform: l and n are)
            __syncthreads();
            tlo += TX;
            load_tile(tlo);
            __syncthreads();
          }
#pragma unroll
          for (int u = 0; u < NB; ++u) q_next[u] = s_t[u][t][rmin + l - tlo];
        } else {
#pragma unroll
          for (int u = 0; u < NB; ++u) q_next[u] = 0.0;
        }
        const double CenterFac = LeftFac + RightFac;
        double Dn, DUn = 0.0, DLn = 0.0;
        if (l == 1) {
          Dn = 1.0 + (th)*CenterFac - (th)*LeftFac;
          DUn = -((th)*RightFac);
        } else if (l < n) {
          DLn = -((th)*LeftFac);
          Dn = 1.0 + (th)*CenterFac;
          DUn = -((th)*RightFac);
        } else {
          DLn = -((th)*LeftFac);
          Dn = 1.0 + (th)*CenterFac - (th)*RightFac;
        }
        double Bn[NB], q_r[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          q_r[u] = (l < n) ? q_next[u] : q_c[u];  // BC_right_t0 (:845)
          Bn[u] = (1.0 - th) * LeftFac * q_l[u] + (1.0 - (1.0 - th) * CenterFac) * q_c[u] +
                  (1.0 - th) * RightFac * q_r[u];
        }
        if (l == 1) {
          Dc = Dn;
          DUc = DUn;
#pragma unroll
          for (int u = 0; u < NB; ++u) Bc[u] = Bn[u];
        } else {
          const int i = l - 1;
          const long long fo = base + sL * (long long)(rmin - 1 + i + 1);
          double DLo;
          if (fabs(Dc) >= fabs(DLn)) {
            if (Dc == 0.0) { if (!singular) singular = i; }
            const double fact = DLn / Dc;
            Dn = Dn - fact * DUc;
#pragma unroll
            for (int u = 0; u < NB; ++u) Bn[u] = Bn[u] - fact * Bc[u];
            DLo = (i <= n - 2) ? 0.0 : DLn;
          } else {
            const double fact = Dc / DLn;
            const double newDi = DLn;
            const double temp = Dn;
            Dn = DUc - fact * temp;
            if (i <= n - 2) {
              DLo = DUn;
              DUn = -fact * DLo;
            } else {
              DLo = DLn;
            }
            DUc = temp;
            Dc = newDi;
#pragma unroll
            for (int u = 0; u < NB; ++u) {
              const double tb = Bc[u];
              Bc[u] = Bn[u];
              Bn[u] = tb - fact * Bn[u];
            }
          }
          if (any_on) {
            fD[fo] = Dc;
            fDU[fo] = DUc;
            fDL[fo] = DLo;
          }
#pragma unroll
          for (int u = 0; u < NB; ++u)
            if (on[u]) sb[fo + d.ncell * (long long)(n0 + u)] = Bc[u];
          Dc = Dn;
          DUc = DUn;
#pragma unroll
          for (int u = 0; u < NB; ++u) Bc[u] = Bn[u];
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          q_l[u] = q_c[u];
          q_c[u] = q_r[u];
        }
      }
    }
    if (Dc == 0.0 && !singular) singular = n;
    if (singular && any_on) atomicCAS(A.flag, 0, singular);
    // ---- back substitution; solutions collect in an LDS tile covering
    // cells [thi-TX+1, thi], stored when the sweep leaves it
    __syncthreads();  // the q tile is no longer read
    int thi = rmax;
    double x1[NB], x2[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      x1[u] = Bc[u] / Dc;
      x2[u] = 0.0;
      s_t[u][t][TX - 1] = x1[u];  // cell rmax at column TX-1 of tile [rmax-TX+1, rmax]
    }
    double bD[PF], bDU[PF], bDL[PF], bB[PF][NB];
    auto fetch_row = [&](int slot, int i) {
      const long long fo = base + sL * (long long)(rmin - 1 + i + 1);
      bD[slot] = fD[fo];
      bDU[slot] = fDU[fo];
      bDL[slot] = fDL[fo];
#pragma unroll
      for (int u = 0; u < NB; ++u) bB[slot][u] = on[u] ? sb[fo + d.ncell * (long long)(n0 + u)] : 0.0;
    };
#pragma unroll
    for (int p = 0; p < PF; ++p)
      if (n - 1 - p >= 1) fetch_row(p, n - 1 - p);
    for (int i0 = n - 1; i0 >= 1; i0 -= PF) {
#pragma unroll
      for (int p = 0; p < PF; ++p) {
        const int i = i0 - p;
        if (i < 1) break;
        const int cell = rmin - 1 + i;
        if (cell < thi - TX + 1) {  // left the tile (uniform)
          __syncthreads();
          store_tile(thi - TX + 1);
          __syncthreads();
          thi -= TX;
        }
        const double Di = bD[p], DUi = bDU[p], DLi = bDL[p];
        double Bcur[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) Bcur[u] = bB[p][u];
        if (i - PF >= 1) fetch_row(p, i - PF);
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const double Bi = Bcur[u];
          double xi;
          if (i == n - 1) xi = (Bi - DUi * x1[u]) / Di;
          else xi = (Bi - DUi * x1[u] - DLi * x2[u]) / Di;
          x2[u] = x1[u];
          x1[u] = xi;
          s_t[u][t][cell - (thi - TX + 1)] = xi;
        }
      }
    }
    __syncthreads();
    store_tile(thi - TX + 1);
  }
  // cells the reference leaves at zero (inactive bins and lines, cells
  // outside [rmin, rmax]); a one-cell line copies nothing (:770, :909)
  if (n > 1 && lane_ok) {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      if (!exists[u]) continue;
      const long long bofs = o.ncell * (long long)(n0 + u);
      if (!on[u]) {
        for (int r = 1; r <= N; ++r) c[obase + (r + 1) + bofs] = 0.0;
      } else {
        for (int r = 1; r < rmin; ++r) c[obase + (r + 1) + bofs] = 0.0;
        for (int r = rmax + 1; r <= N; ++r) c[obase + (r + 1) + bofs] = 0.0;
      }
    }
  }
}

template <int NB, int TX>
hipError_t launch_cn_xrows_nb(hipStream_t s, const DiffArgs &a, const XState &x) {
  const long long lines = (long long)a.g.Na * a.g.Nb;
  dim3 grid((unsigned)(padded_tiles((lines + 63) / 64) * ((a.d.ns + NB - 1) / NB)));
  hipLaunchKernelGGL((k_diff_cn_xrows<NB, TX>), grid, dim3(64), 0, s, a, x);
  return hipGetLastError();
}

template <int NB>
hipError_t launch_cn_xrows_tx(hipStream_t s, const DiffArgs &a, const XState &x, int tx) {
  // LDS per workgroup NB*64*(TX+1) doubles: at most 34 KiB for NB <= 2 and
  // TX = 32, 35 KiB for NB = 4 and TX = 16, 37 KiB for NB = 8 and TX = 8
  if (tx <= 8 || NB >= 8) return launch_cn_xrows_nb<NB, 8>(s, a, x);
  if (tx <= 16 || NB >= 4) return launch_cn_xrows_nb<NB, 16>(s, a, x);
  return launch_cn_xrows_nb<NB, 32>(s, a, x);
}

template <template <int, int> class K, int NB>
struct dummy {};

template <int NB>
hipError_t launch_explicit_nb(hipStream_t s, const DiffArgs &a) {
  const long long lines = (long long)a.g.Na * a.g.Nb;
  dim3 grid((unsigned)(padded_tiles((lines + kBlock - 1) / kBlock) * ((a.d.ns + NB - 1) / NB)));
  switch (a.g.dir) {
    case 0: hipLaunchKernelGGL((k_diff_explicit<0, NB>), grid, dim3(kBlock), 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_diff_explicit<1, NB>), grid, dim3(kBlock), 0, s, a); break;
    default: hipLaunchKernelGGL((k_diff_explicit<2, NB>), grid, dim3(kBlock), 0, s, a); break;
  }
  return hipGetLastError();
}

template <int NB>
hipError_t launch_cn_nb(hipStream_t s, const DiffArgs &a) {
  const long long lines = (long long)a.g.Na * a.g.Nb;
  dim3 grid((unsigned)(padded_tiles((lines + kBlock - 1) / kBlock) * ((a.d.ns + NB - 1) / NB)));
  switch (a.g.dir) {
    case 0: hipLaunchKernelGGL((k_diff_cn<0, NB>), grid, dim3(kBlock), 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_diff_cn<1, NB>), grid, dim3(kBlock), 0, s, a); break;
    default: hipLaunchKernelGGL((k_diff_cn<2, NB>), grid, dim3(kBlock), 0, s, a); break;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_diff_explicit(hipStream_t s, const DiffArgs &a, int nb) {
  switch (nb) {
    case 1: return launch_explicit_nb<1>(s, a);
    case 2: return launch_explicit_nb<2>(s, a);
    case 4: return launch_explicit_nb<4>(s, a);
    default: return launch_explicit_nb<8>(s, a);
  }
}

hipError_t launch_line_lr(hipStream_t s, const DiffArgs &a, double *lr) {
  const long long cells = (long long)a.g.Na * a.g.N * a.g.Nb;
  dim3 grid((unsigned)((cells + kBlock - 1) / kBlock));
  switch (a.g.dir) {
    case 0: hipLaunchKernelGGL((k_line_lr<0>), grid, dim3(kBlock), 0, s, a, lr); break;
    case 1: hipLaunchKernelGGL((k_line_lr<1>), grid, dim3(kBlock), 0, s, a, lr); break;
    default: hipLaunchKernelGGL((k_line_lr<2>), grid, dim3(kBlock), 0, s, a, lr); break;
  }
  return hipGetLastError();
}

hipError_t launch_diff_cn_xrows(hipStream_t s, const DiffArgs &a, const Dims &orig, double *c0, int nb, int tx) {
  const XState x{orig, c0};
  switch (nb) {
    case 1: return launch_cn_xrows_tx<1>(s, a, x, tx);
    case 2: return launch_cn_xrows_tx<2>(s, a, x, tx);
    case 4: return launch_cn_xrows_tx<4>(s, a, x, tx);
    default: return launch_cn_xrows_tx<8>(s, a, x, tx);
  }
}

hipError_t launch_diff_cn(hipStream_t s, const DiffArgs &a, int nb) {
  switch (nb) {
    case 1: return launch_cn_nb<1>(s, a);
    case 2: return launch_cn_nb<2>(s, a);
    case 4: return launch_cn_nb<4>(s, a);
    default: return launch_cn_nb<8>(s, a);
  }
}

}  // namespace ash3d
