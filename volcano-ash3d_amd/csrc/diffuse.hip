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
#include <float.h>

#include "common.hpp"
#include "lanes.hpp"

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

// LU factorisation of the theta-scheme lines: dgtsv's elimination
// (partial pivoting) of the matrix alone, one thread per line, in dgtsv's
// order and arithmetic.  The matrix depends on dt, theta, the active range
// [rmin, rmax] and the row factors only, so the context keeps the factors
// between steps and the line solves apply them to each bin's right-hand
// side.  Per row i (cell rmin-1+i), at that cell's offset: fact_i and the
// interchange flag piv_i of elimination step i (rows i and i+1), and the U
// entries D_i, DU_i, DL_i (DL_i: second super-diagonal); D_n on the last
// row.  sing[line] is dgtsv's INFO for the line (first zero pivot row).
// *ndom counts the lines that are not diagonally dominant M-matrix rows
// (some L or R negative or not finite, or theta < 0): the partitioned
// solves, which eliminate without pivoting, are used only when it is zero.
template <int DIR>
__global__ __launch_bounds__(kBlock) void k_line_lu(DiffArgs A, double *lu, uint8_t *piv, int *sing, int *ndom,
                                                   unsigned long long *rho, uint8_t *jl, unsigned long long *rho_j,
                                                   int *njl) {
  const Dims d = A.d;
  const LineGeom g = A.g;
  const long long gid = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= (long long)g.Na * g.Nb) return;
  const int a = (int)(gid % g.Na) + 1;
  const int b = (int)(gid / g.Na) + 1;
  const int rmin = g.rmin, n = g.rmax - g.rmin + 1;
  const long long sL = line_stride(d, DIR), base = line_base(d, DIR, a, b);
  const double th = A.theta;
  const double *Lf = A.lr, *Rf = A.lr + d.ncell;
  double *fF = lu, *fD = lu + d.ncell, *fDU = fD + d.ncell, *fDL = fDU + d.ncell;
  int singular = 0;
  bool dom = true;
  double rho_line = 0.0;  // largest (|DL| + |DU|) / |D| of the line's rows
  double Dc = 0.0, DUc = 0.0;
  // row factors PF rows ahead
  constexpr int PF = 8;
  double rL[PF], rR[PF];
#pragma unroll
  for (int p = 0; p < PF; ++p) {
    const long long o = base + sL * (long long)(rmin + min(1 + p, n));
    rL[p] = Lf[o];
    rR[p] = Rf[o];
  }
  for (int l0 = 1; l0 <= n; l0 += PF) {
#pragma unroll
    for (int p = 0; p < PF; ++p) {
      const int l = l0 + p;
      if (l > n) break;
      const double LeftFac = rL[p], RightFac = rR[p];
      {
        const long long o = base + sL * (long long)(rmin + min(l + PF, n));
        rL[p] = Lf[o];
        rR[p] = Rf[o];
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
      // L, R >= 0 and finite, theta >= 0: every row has D >= 1 + |DL| + |DU|
      // in exact arithmetic (false for NaN too)
      dom = dom && LeftFac >= 0.0 && RightFac >= 0.0 && LeftFac <= DBL_MAX && RightFac <= DBL_MAX && th >= 0.0;
      rho_line = fmax(rho_line, (fabs(DLn) + fabs(DUn)) / fabs(Dn));
      if (l == 1) {
        Dc = Dn;
        DUc = DUn;
        continue;
      }
      const int i = l - 1;
      const long long fo = base + sL * (long long)(rmin + i);  // cell of row i
      double DLo, fact;
      uint8_t p_i;
      if (fabs(Dc) >= fabs(DLn)) {
        if (Dc == 0.0) { if (!singular) singular = i; }
        fact = DLn / Dc;
        Dn = Dn - fact * DUc;
        DLo = (i <= n - 2) ? 0.0 : DLn;
        p_i = 0;
      } else {
        fact = Dc / DLn;
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
        p_i = 1;
      }
      fF[fo] = fact;
      piv[fo] = p_i;
      fD[fo] = Dc;
      fDU[fo] = DUc;
      fDL[fo] = DLo;
      Dc = Dn;
      DUc = DUn;
    }
  }
  if (Dc == 0.0 && !singular) singular = n;
  fD[base + sL * (long long)(rmin + n)] = Dc;
  sing[gid] = singular;
  if (!dom) atomicAdd(ndom, 1);
  // the sweep's largest row ratio (non-negative doubles order as their bit
  // patterns); lines that are not dominant count as 1
  const double r = (dom && rho_line < 1.0) ? rho_line : 1.0;
  if (rho) atomicMax(rho, (unsigned long long)__double_as_longlong(r));
  // the Jacobi sweeps the line needs (0: more than kJacobiMaxSweepsX); the
  // number of lines that need at most that many and the largest ratio
  // among them
  if (jl) {
    const int sw = dom ? jacobi_sweeps_needed(r, kJacobiMaxSweepsX) : 0;
    jl[gid] = (uint8_t)sw;
    if (sw > 0) {
      atomicMax(rho_j, (unsigned long long)__double_as_longlong(r));
      atomicAdd(njl, 1);
    }
  }
}

__global__ void k_zero_i32(int *p) { *p = 0; }
__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

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
// with q_left = q_c on the first row and q_right = q_c on the last.  The
// matrix part is factored once by k_line_lu; here each bin's right-hand
// side goes through dgtsv's elimination step i, which with rows (i, i+1)
// holding (Bc, Bn) is
//   no interchange: row i keeps Bc,  row i+1 gets Bn - fact_i*Bc
//   interchange:    row i takes Bn,  row i+1 gets Bc - fact_i*Bn
// i.e. row i keeps y and row i+1 gets x - fact_i*y with (x, y) = (Bn, Bc)
// or (Bc, Bn), then the back substitution.  Every row issues the same
// memory operations (masked-off lanes store to A.junk), so the loads kept
// in flight are waited for exactly.
template <int DIR, int NB>
__global__ __launch_bounds__(kBlock) void k_diff_cn(DiffArgs A) {
  constexpr int PF = 4;  // rows in flight per stream
  if (A.sel == 2 && *A.ndom == 0) return;  // the partitioned solves ran (uniform)
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
  const bool line_on = a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  double *c = A.c0;
  double *sb = A.scratch_b;
  double *junk = A.junk + (threadIdx.x + kBlock * (blockIdx.x % (kJunk / kBlock)));
  const double *fF = A.lu, *fD = A.lu + d.ncell, *fDU = fD + d.ncell, *fDL = fDU + d.ncell;
  const uint8_t *fP = A.piv;
  bool exists[NB], on[NB];
  bool any_on = false;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    exists[u] = (n0 + u) < d.ns;
    on[u] = exists[u] && line_on && A.aloft.on(n0 + u);
    any_on = any_on || on[u];
  }
  if (any_on) {
    const int singular = A.sing[gid];
    if (singular) atomicCAS(A.flag, 0, singular);
    const double th = A.theta;
    auto qofs = [&](int u, int r) { return base + sL * (long long)(r + 1) + d.ncell * (long long)(n0 + u); };
    const double *Lf = A.lr, *Rf = A.lr + d.ncell;
    double q_l[NB], q_c[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      q_c[u] = exists[u] ? c[qofs(u, rmin)] : 0.0;
      q_l[u] = q_c[u];  // BC_left_t0 = q(rmin) (:822)
    }
    double Bc[NB];  // right-hand side of the row being eliminated
    // Row l needs its factors L, R (cell c = rmin-1+l), fact and piv of
    // elimination step l-1 (cell c-1) and the q of its right neighbour
    // (cell c+1, while it is active).  A ring of PF rows is kept in
    // flight: slot p holds row l0+p of the current block of PF rows and is
    // refilled with row l+PF (the last row again past the end).
    double rL[PF], rR[PF], rF[PF], rq[PF][NB];
    uint8_t rP[PF];
    auto fetch_row_in = [&](int slot, int l) {
      const long long o = base + sL * (long long)(rmin + l);  // cell rmin-1+l
      rL[slot] = Lf[o];
      rR[slot] = Rf[o];
      rF[slot] = fF[o - sL];
      rP[slot] = fP[o - sL];
#pragma unroll
      for (int u = 0; u < NB; ++u) rq[slot][u] = c[(l < n && exists[u]) ? qofs(u, rmin + l) : 0];
    };
#pragma unroll
    for (int p = 0; p < PF; ++p) fetch_row_in(p, min(1 + p, n));
    for (int l0 = 1; l0 <= n; l0 += PF) {
#pragma unroll
    for (int p = 0; p < PF; ++p) {
      const int l = l0 + p;
      if (l > n) break;
      const double LeftFac = rL[p], RightFac = rR[p], fact = rF[p];
      const bool pv = rP[p] != 0;
      double q_next[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) q_next[u] = rq[p][u];
      fetch_row_in(p, min(l + PF, n));
      const double CenterFac = LeftFac + RightFac;
      double Bn[NB], q_r[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        q_r[u] = (l < n) ? q_next[u] : q_c[u];  // BC_right_t0 (:845)
        Bn[u] = (1.0 - th) * LeftFac * q_l[u] + (1.0 - (1.0 - th) * CenterFac) * q_c[u] +
                (1.0 - th) * RightFac * q_r[u];
      }
      if (l == 1) {
#pragma unroll
        for (int u = 0; u < NB; ++u) Bc[u] = Bn[u];
      } else {
        // elimination step i = l-1: row i is final
        const long long fo = base + sL * (long long)(rmin + l - 1);  // cell of row i
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const double y = pv ? Bn[u] : Bc[u];
          const double x = pv ? Bc[u] : Bn[u];
          *(on[u] ? sb + fo + d.ncell * (long long)(n0 + u) : junk) = y;
          Bc[u] = x - fact * y;
        }
      }
      // advance the window
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        q_l[u] = q_c[u];
        q_c[u] = q_r[u];
      }
    }
    }
    // ---- back substitution (dgtsv, NRHS <= 2 path), row n first
    double x1[NB], x2[NB];
    {
      const double Dn_ = fD[base + sL * (long long)(rmin + n)];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        x1[u] = Bc[u] / Dn_;
        x2[u] = 0.0;
        *(on[u] ? c + qofs(u, rmax) : junk) = x1[u];
      }
    }
    // rows' factors and eliminated right-hand sides, PF rows in flight
    // (slot p holds row i0-p of the current block; row 1 again past it)
    double bD[PF], bDU[PF], bDL[PF], bB[PF][NB];
    auto fetch_row = [&](int slot, int i) {
      const long long fo = base + sL * (long long)(rmin + i);
      bD[slot] = fD[fo];
      bDU[slot] = fDU[fo];
      bDL[slot] = fDL[fo];
#pragma unroll
      for (int u = 0; u < NB; ++u) bB[slot][u] = sb[on[u] ? fo + d.ncell * (long long)(n0 + u) : 0];
    };
#pragma unroll
    for (int p = 0; p < PF; ++p) fetch_row(p, max(n - 1 - p, 1));
    for (int i0 = n - 1; i0 >= 1; i0 -= PF) {
#pragma unroll
    for (int p = 0; p < PF; ++p) {
      const int i = i0 - p;
      if (i < 1) break;
      const long long fo = base + sL * (long long)(rmin + i);
      const double Di = bD[p], DUi = bDU[p], DLi = bDL[p];
      double Bcur[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) Bcur[u] = bB[p][u];
      fetch_row(p, max(i - PF, 1));
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const double Bi = Bcur[u];
        double xi;
        if (i == n - 1) xi = (Bi - DUi * x1[u]) / Di;
        else xi = (Bi - DUi * x1[u] - DLi * x2[u]) / Di;
        x2[u] = x1[u];
        x1[u] = xi;
        *(on[u] ? c + fo + d.ncell * (long long)(n0 + u) : junk) = xi;
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

// theta-scheme line solve by Jacobi sweeps, for sweeps whose rows are
// strongly diagonally dominant.  Row t of a line (the matrix of k_line_lu:
// DL = -th L, D, DU = -th R) gives x_t = y_t + a_t x_{t-1} + b_t x_{t+1}
// with y = B / D, a = th L / D, b = th R / D (a_1 = b_n = 0).  The sweeps
// x(0) = y, x(m) = y + a x(m-1)[t-1] + b x(m-1)[t+1] leave, after K of
// them, an error of at most rho^(K+1) of the line's largest |x| (rho =
// max (|DL| + |DU|) / |D| = max (a + b), the max norm of the iteration
// matrix); the host picks K.  The K sweeps run in one pass along the line,
// sweep m lagging m rows behind the rows entering: row s enters (its
// right-hand side B from q at rows s-1..s+1, as the dgtsv-order kernel
// forms it), sweep m updates row s-m from sweep m-1's values at rows s-m-1
// and s-m+1, and sweep K's row s-K is stored, in place (every q it
// overwrites has been read).  One thread per line and NB bins, lanes along
// the contiguous transverse index: q is read and written once per
// bin-cell, the row factors L, R once per bin group.
template <int DIR, int NB, int K, bool ADJ>
__global__ __launch_bounds__(kBlock) void k_cn_jacobi(DiffArgs A) {
  constexpr int PF = 3;  // rows loaded ahead (2: 7.12 ms, 3: 6.95, 4: 7.03 on the global y lines)
  constexpr int NY = K + 1 > 3 ? K + 1 : 3;  // y of rows s .. s-NY+1 (sweep 0 and the rows sweeps use)
  // windows of sweeps 1 .. K-1 (the adjoint: 1 .. K, its last sweep feeds B)
  constexpr int NW = ADJ ? (K > 0 ? K : 1) : (K > 1 ? K - 1 : 1);
  constexpr int KD = ADJ ? K + 2 : 1;  // the adjoint's rows s .. s-K-1 of L, R and kappa
  constexpr int LAG = ADJ ? K + 1 : K;  // rows between a row's entry and its store
  const Dims d = A.d;
  const LineGeom g = A.g;
  const TileGroup tg = xcd_tile_group(blockIdx.x, (d.ns + NB - 1) / NB);
  const long long gid = tg.tile * kBlock + threadIdx.x;
  if (gid >= (long long)g.Na * g.Nb) return;
  const int a = (int)(gid % g.Na) + 1;
  const int b = (int)(gid / g.Na) + 1;
  const int n0 = tg.group * NB;
  const int N = g.N, rmin = g.rmin, rmax = g.rmax;
  const int n = rmax - rmin + 1;
  const long long sL = line_stride(d, DIR), base = line_base(d, DIR, a, b);
  // the state: cell r of the line at qb + qs*(r+1), bin stride qn
  const long long qs = sL, qb = base, qn = d.ncell;
  const bool line_on = a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  double *c = A.c0;
  bool exists[NB], on[NB];
  long long bofs[NB];  // bin offset (an existing bin's for the others)
  bool any_on = false;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    exists[u] = (n0 + u) < d.ns;
    on[u] = exists[u] && line_on && A.aloft.on(n0 + u);
    any_on = any_on || on[u];
    bofs[u] = qn * (long long)min(n0 + u, d.ns - 1);
  }
  if (any_on) {
    const double th = A.theta, th1 = 1.0 - th;
    const double *Lf = A.lr, *Rf = A.lr + d.ncell;
    auto lofs = [&](int l) { return base + sL * (long long)(rmin + l); };  // row l: cell rmin-1+l
    auto ofs = [&](int l) { return qb + qs * (long long)(rmin + l); };
    // rows ahead: slot p holds L, R (the adjoint: and kappa) of row s and q
    // of row s+1
    double rL[PF], rR[PF], rK[PF], rq[PF][NB];
    auto fetch = [&](int slot, int l) {
      const long long o = lofs(min(l, n));
      rL[slot] = Lf[o];
      rR[slot] = Rf[o];
      if (ADJ) rK[slot] = A.kap[o];
      const long long oq = ofs(min(l + 1, n));
#pragma unroll
      for (int u = 0; u < NB; ++u) rq[slot][u] = c[oq + bofs[u]];
    };
    double qc[NB], qp[NB];  // q of rows s and s-1 as row s enters
#pragma unroll
    for (int u = 0; u < NB; ++u) qc[u] = qp[u] = c[ofs(1) + bofs[u]];
#pragma unroll
    for (int p = 0; p < PF; ++p) fetch(p, 1 + p);
    double y[NY][NB], ca[K + 1], cb[K + 1];  // index j: row s-j
    double xw[NW][3][NB];                    // sweep m: rows s-m, s-m-1, s-m-2
    double dL[KD], dR[KD], dK[KD];           // the adjoint's row factors, index j: row s-j
#pragma unroll
    for (int j = 0; j < NY; ++j)
#pragma unroll
      for (int u = 0; u < NB; ++u) y[j][u] = 0.0;
#pragma unroll
    for (int j = 0; j <= K; ++j) ca[j] = cb[j] = 0.0;
#pragma unroll
    for (int j = 0; j < KD; ++j) dL[j] = dR[j] = dK[j] = 0.0;
#pragma unroll
    for (int m = 0; m < NW; ++m)
#pragma unroll
      for (int e = 0; e < 3; ++e)
#pragma unroll
        for (int u = 0; u < NB; ++u) xw[m][e][u] = 0.0;
    for (int s0 = 1; s0 <= n + LAG; s0 += PF) {
#pragma unroll
      for (int p = 0; p < PF; ++p) {
        const int s = s0 + p;
        if (s > n + LAG) break;
        const double L = rL[p], R = rR[p], kap = ADJ ? rK[p] : 0.0;
        double qn[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) qn[u] = rq[p][u];
        fetch(p, s + PF);
#pragma unroll
        for (int j = NY - 1; j > 0; --j)
#pragma unroll
          for (int u = 0; u < NB; ++u) y[j][u] = y[j - 1][u];
#pragma unroll
        for (int j = K; j > 0; --j) {
          ca[j] = ca[j - 1];
          cb[j] = cb[j - 1];
        }
        if (ADJ) {
#pragma unroll
          for (int j = KD - 1; j > 0; --j) {
            dL[j] = dL[j - 1];
            dR[j] = dR[j - 1];
            dK[j] = dK[j - 1];
          }
        }
        if (s <= n) {
          const double C = L + R;
          double D = 1.0 + th * C;
          if (s == 1) D = D - th * L;
          if (s == n) D = D - th * R;
          const double rD = 1.0 / D;
          ca[0] = (s > 1) ? (th * L) * rD : 0.0;
          cb[0] = (s < n) ? (th * R) * rD : 0.0;
          if (ADJ) {
            // the adjoint solves A u = lambda / kappa (then lambda' = kappa B u)
            dL[0] = L;
            dR[0] = R;
            dK[0] = kap;
#pragma unroll
            for (int u = 0; u < NB; ++u) y[0][u] = (qc[u] / kap) * rD;
          } else {
#pragma unroll
            for (int u = 0; u < NB; ++u) {
              const double ql = (s == 1) ? qc[u] : qp[u];
              const double qr = (s == n) ? qc[u] : qn[u];
              y[0][u] = ((th1 * L * ql + (1.0 - th1 * C) * qc[u]) + th1 * R * qr) * rD;
            }
          }
        } else {
          ca[0] = cb[0] = 0.0;
          if (ADJ) dL[0] = dR[0] = dK[0] = 0.0;
#pragma unroll
          for (int u = 0; u < NB; ++u) y[0][u] = 0.0;
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          qp[u] = qc[u];
          qc[u] = qn[u];
        }
        if constexpr (K == 0 && !ADJ) {
          if (s <= n) {
#pragma unroll
            for (int u = 0; u < NB; ++u)
              if (on[u]) c[ofs(s) + bofs[u]] = y[0][u];
          }
        } else {
          // sweep m at row t = s-m from sweep m-1 at rows t-1 and t+1
          // (sweep 0: y at rows s-m-1, s-m+1; later sweeps: their windows)
#pragma unroll
          for (int m = 1; m <= K; ++m) {
            const int t = s - m;
            const bool tin = t >= 1 && t <= n;
            double xm[NB];
#pragma unroll
            for (int u = 0; u < NB; ++u) {
              const double lo = (m == 1) ? y[2][u] : xw[m > 1 ? m - 2 : 0][2][u];
              const double hi = (m == 1) ? y[0][u] : xw[m > 1 ? m - 2 : 0][0][u];
              xm[u] = tin ? (y[m][u] + ca[m] * lo) + cb[m] * hi : 0.0;
            }
            if (ADJ || m < K) {
              const int w = m - 1 < NW ? m - 1 : 0;
#pragma unroll
              for (int u = 0; u < NB; ++u) {
                xw[w][2][u] = xw[w][1][u];
                xw[w][1][u] = xw[w][0][u];
                xw[w][0][u] = xm[u];
              }
            } else if (tin) {
#pragma unroll
              for (int u = 0; u < NB; ++u)
                if (on[u]) c[ofs(t) + bofs[u]] = xm[u];
            }
          }
          if constexpr (ADJ) {
            // lambda'_t = kappa_t (B u)_t at t = s-K-1, B's line ends
            // u_0 = u_1 and u_{n+1} = u_n (the partitioned adjoint's B)
            const int t = s - K - 1;
            if (t >= 1 && t <= n) {
              const double Lt = dL[KD - 1], Rt = dR[KD - 1], kt = dK[KD - 1], Ct = Lt + Rt;
#pragma unroll
              for (int u = 0; u < NB; ++u) {
                const double up = K == 0 ? y[0][u] : xw[NW - 1][0][u];
                const double uc = K == 0 ? y[1][u] : xw[NW - 1][1][u];
                const double um = K == 0 ? y[2][u] : xw[NW - 1][2][u];
                const double xm_ = (t == 1) ? uc : um;
                const double xp_ = (t == n) ? uc : up;
                const double v = (th1 * Lt) * xm_ + (1.0 - th1 * Ct) * uc + (th1 * Rt) * xp_;
                if (on[u]) c[ofs(t) + bofs[u]] = kt * v;
              }
            }
          }
        }
      }
    }
  }
  // cells the reference leaves at zero (zero fill of ts1, then copy; the
  // adjoint of those writes zeroes lambda there)
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    if (!exists[u]) continue;
    const long long bo = qn * (long long)(n0 + u);
    if (!on[u]) {
      for (int r = 1; r <= N; ++r) c[qb + qs * (long long)(r + 1) + bo] = 0.0;
    } else {
      for (int r = 1; r < rmin; ++r) c[qb + qs * (long long)(r + 1) + bo] = 0.0;
      for (int r = rmax + 1; r <= N; ++r) c[qb + qs * (long long)(r + 1) + bo] = 0.0;
    }
  }
}

// x lines (contiguous in memory) by Jacobi sweeps, each line with the
// sweeps it needs (A.jline, up to A.jcap; the partitioned solves take the
// others).  One workgroup of WV waves per line; wave w owns rows
// 64 M w .. 64 M (w+1) - 1 (M = 6), and lane l of it the M consecutive rows
// 64 M w + M l + m (m < M), so that a sweep finds every neighbour in the
// lane's own registers but those of its first and last row (the
// neighbouring lanes' by whole-wave DPP rotations; lanes 0 and 63 the
// neighbouring waves' through LDS, one barrier per sweep).  Loads and
// stores stay coalesced: a bin's q is read row block by row block (lane l
// of block m: row 64 M w + 64 m + l) and turned into the lane's rows
// through LDS, the solution the same way back.  The row factors L, R are
// formed from kappa, sigma_nx and kx as k_line_lr forms them (same
// arithmetic, so the same values as the factor cache) once per line, with
// the Jacobi coefficients; then every aloft bin of the line in turn: its
// right-hand side as k_cn_jacobi forms it (the next bin's q already in
// flight), the sweeps (fused multiply-adds, alternating between two sets
// of registers), one store.  A is the x<->y transposed description the
// other x-line solves use (lines a = j, b = k, rows along x); X the
// original layout.
struct XJac {
  Dims o;                        // original dims
  double *c;                     // concen ts0
  const double *kap, *sig, *kk;  // kappa, sigma_nx (with ghosts) and kx
};

template <int WV, int M, bool ADJ>
__global__ __launch_bounds__(64 * WV) void k_cn_jacobi_x(DiffArgs A, XJac X) {
  constexpr int P = M + 1;  // LDS pitch of a lane's rows (bank spread)
  __shared__ double s_ex[2][WV][2];       // per parity: each wave's first and last value
  __shared__ double s_q[WV][64 * P];      // q, block order in, lane order out
  __shared__ double s_x[WV][64 * P];      // the solution, lane order in, block order out
  const LineGeom g = A.g;
  const long long line = blockIdx.x;
  const int sweeps = A.jline[line];
  if (sweeps < 1 || sweeps > A.jcap) return;  // a line of the partitioned solves (uniform)
  const int a = (int)(line % g.Na) + 1;   // j
  const int b = (int)(line / g.Na) + 1;   // k
  const Dims o = X.o;
  const int rmin = g.rmin, rmax = g.rmax, n = rmax - rmin + 1, N = g.N;
  const bool line_on = a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const bool lane0 = l == 0, lane63 = l == 63;
  const long long ob = o.p3(-1, a, b);          // cell r at ob + r + 1
  const long long kb = kline_base(o, 0, a, b);  // K of cell r at kb + r
  double *c = X.c;
  const double th = A.theta, th1 = 1.0 - th, dt = A.dt;
  const int w0 = w * 64 * M;
  auto own = [&](int m) { return w0 + M * l + m; };   // the lane's rows
  auto blk = [&](int m) { return w0 + 64 * m + l; };  // row of block m
  // LDS slot of row w0 + t: lane t / M, element t % M
  auto slot = [&](int t) { return (t / M) * P + t % M; };
  // per row: L, R, 1/D and the Jacobi coefficients (zero on padding rows);
  // the adjoint also kappa
  double rL[M], rR[M], rD[M], ca[M], cb[M], kc[ADJ ? M : 1];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = own(m);
    const bool real = i < n;
    const int r = rmin + (real ? i : 0);
    auto kapat = [&](int q) { return X.kap[ob + q + 1]; };
    auto Kat = [&](int q) { return X.kk[kb + q]; };
    auto sigat = [&](int q) { return X.sig[ob + q + 1]; };
    auto kds_at = [&](int I) {
      const double vavg = 0.5 * (kapat(I - 1) + kapat(I));
      const double kavg = 0.5 * (Kat(I - 1) + Kat(I));
      const double ds = sigat(I) / vavg;
      return kavg * ds;
    };
    const double kap_c = kapat(r);
    const double L = dt * kds_at(r) * sigat(r) / kap_c;
    const double R = dt * kds_at(r + 1) * sigat(r + 1) / kap_c;
    const double C = L + R;
    double D = 1.0 + th * C;
    if (i == 0) D = D - th * L;
    if (i == n - 1) D = D - th * R;
    const double rd = 1.0 / D;
    rL[m] = real ? L : 0.0;
    rR[m] = real ? R : 0.0;
    rD[m] = real ? rd : 0.0;
    ca[m] = (real && i > 0) ? (th * L) * rd : 0.0;
    cb[m] = (real && i < n - 1) ? (th * R) * rd : 0.0;
    if (ADJ) kc[ADJ ? m : 0] = kap_c;
  }
  // exchange slots: each wave writes its first and last value, then reads
  // the neighbouring waves' (zero beyond the line)
  int parity = 0;
  auto put_ends = [&](double first, double last) {
    if (WV > 1) {
      if (lane0) s_ex[parity][w][0] = first;
      if (lane63) s_ex[parity][w][1] = last;
    }
  };
  auto get_ends = [&](double &before, double &after) {
    before = (WV > 1 && w > 0) ? s_ex[parity][w - 1][1] : 0.0;
    after = (WV > 1 && w + 1 < WV) ? s_ex[parity][w + 1][0] : 0.0;
    parity ^= 1;
  };
  // the values at the lane's first row - 1 and last row + 1 (all lanes)
  auto ends = [&](const double (&v)[M], double before, double after, double &lo, double &hi) {
    const double p = dpp_f64<kWaveRor1>(v[M - 1]);  // lane l-1's last row
    const double nx = dpp_f64<kWaveRol1>(v[0]);     // lane l+1's first row
    lo = lane0 ? before : p;
    hi = lane63 ? after : nx;
  };
  const long long ncell = o.ncell;
  const int ns = o.ns;
  auto on_bin = [&](int u) { return line_on && A.aloft.on(u); };
  auto next_on = [&](int u) {
    while (u < ns && !on_bin(u)) ++u;
    return u;
  };
  double qb[M];  // q of the bin in flight, block order
  auto load = [&](int u) {
    const long long bo = ncell * (long long)u;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int i = blk(m);
      qb[m] = c[ob + rmin + (i < n ? i : 0) + 1 + bo];
    }
  };
  double xa[M], xb[M];  // sweeps alternate between these
  int u = next_on(0);
  int u_prev = ns;      // the bin whose solution waits in s_x
  if (u < ns) load(u);
  auto store_prev = [&]() {
    // the previous bin's solution: s_x in lane order, stored by blocks
    const long long bo = ncell * (long long)u_prev;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int i = blk(m);
      if (i < n) c[ob + rmin + i + 1 + bo] = s_x[w][slot(64 * m + l)];
    }
  };
  while (u < ns) {
    // q into the lane's rows (and the wave ends of q), with the previous
    // bin's solution out of s_x: one barrier
#pragma unroll
    for (int m = 0; m < M; ++m) s_q[w][slot(64 * m + l)] = qb[m];
    put_ends(qb[0], qb[M - 1]);  // block 0 lane 0 / block M-1 lane 63: the wave's first / last row
    __syncthreads();
    if (u_prev < ns) store_prev();
    double q[M];
#pragma unroll
    for (int m = 0; m < M; ++m) q[m] = s_q[w][l * P + m];
    double before, after, qlo, qhi;
    get_ends(before, after);
    const int un = next_on(u + 1);
    if (un < ns) load(un);
    ends(q, before, after, qlo, qhi);
    // right-hand side (the reference's q_0 = q_1 and q_{n+1} = q_n); the
    // adjoint solves A u = lambda / kappa
    double y[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int i = own(m);
      double ql = m > 0 ? q[m > 0 ? m - 1 : 0] : qlo;
      double qr = m + 1 < M ? q[m + 1 < M ? m + 1 : m] : qhi;
      if (i == 0) ql = q[m];
      if (i == n - 1) qr = q[m];
      const double L = rL[m], R = rR[m], C = L + R;
      if (ADJ)
        y[m] = (q[m] / kc[ADJ ? m : 0]) * rD[m];
      else
        y[m] = ((th1 * L * ql + (1.0 - th1 * C) * q[m]) + th1 * R * qr) * rD[m];
    }
    // sweep: dst = y + ca * src[row - 1] + cb * src[row + 1]
    auto sweep = [&](const double (&src)[M], double (&dst)[M]) {
      put_ends(src[0], src[M - 1]);
      if (WV > 1) __syncthreads();
      double bf, af, lo, hi;
      get_ends(bf, af);
      ends(src, bf, af, lo, hi);
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const double xl = m > 0 ? src[m > 0 ? m - 1 : 0] : lo;
        const double xr = m + 1 < M ? src[m + 1 < M ? m + 1 : m] : hi;
        dst[m] = fma(cb[m], xr, fma(ca[m], xl, y[m]));
      }
    };
    // sweep 1: x = y; then pairs of sweeps y -> xa -> xb -> xa ...
    int done = 1;
    bool in_a = false;  // the current iterate: y (false, before any sweep) or xa / xb
#pragma unroll
    for (int m = 0; m < M; ++m) xb[m] = y[m];
    for (; done + 2 <= sweeps; done += 2) {
      sweep(xb, xa);
      sweep(xa, xb);
    }
    if (done < sweeps) {
      sweep(xb, xa);
      in_a = true;
    }
    if constexpr (ADJ) {
      // lambda' = kappa B u, B's line ends u_0 = u_1 and u_{n+1} = u_n
      double uu[M];
#pragma unroll
      for (int m = 0; m < M; ++m) uu[m] = in_a ? xa[m] : xb[m];
      put_ends(uu[0], uu[M - 1]);
      if (WV > 1) __syncthreads();
      double bf, af, lo, hi;
      get_ends(bf, af);
      ends(uu, bf, af, lo, hi);
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int i = own(m);
        double um = m > 0 ? uu[m > 0 ? m - 1 : 0] : lo;
        double up = m + 1 < M ? uu[m + 1 < M ? m + 1 : m] : hi;
        if (i == 0) um = uu[m];
        if (i == n - 1) up = uu[m];
        const double L = rL[m], R = rR[m];
        const double v = (th1 * L) * um + (1.0 - th1 * (L + R)) * uu[m] + (th1 * R) * up;
        s_x[w][l * P + m] = kc[ADJ ? m : 0] * v;
      }
    } else {
      // the solution into s_x in lane order; stored after the next barrier
#pragma unroll
      for (int m = 0; m < M; ++m) s_x[w][l * P + m] = in_a ? xa[m] : xb[m];
    }
    u_prev = u;
    u = un;
  }
  if (u_prev < ns) {
    __syncthreads();
    store_prev();
  }
  // cells the reference leaves at zero: outside [rmin, rmax], and every
  // cell of bins that are not aloft and of lines outside the bounds
  for (int v = 0; v < ns; ++v) {
    const bool on = on_bin(v);
    const long long bo = ncell * (long long)v;
    for (int r = 1 + (int)threadIdx.x; r <= N; r += 64 * WV)
      if (!on || r < rmin || r > rmax) c[ob + r + 1 + bo] = 0.0;
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
  constexpr int PF = 4;  // rows in flight; TX cells per LDS tile
  if (A.sel == 2 && *A.ndom == 0) return;  // the partitioned solves ran (uniform)
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
  const bool line_on = lane_ok && a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  const long long obase = o.p3(-1, a, b);
  double *c = X.c;
  double *sb = A.scratch_b;
  double *junk = A.junk + (t + 64 * (blockIdx.x % (kJunk / 64)));
  const double *fF = A.lu, *fD = A.lu + d.ncell, *fDU = fD + d.ncell, *fDL = fDU + d.ncell;
  const uint8_t *fP = A.piv;
  bool exists[NB], bin_on[NB], on[NB];
  bool any_on = false;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    exists[u] = (n0 + u) < d.ns;
    bin_on[u] = exists[u] && A.aloft.on(n0 + u);
    on[u] = bin_on[u] && line_on;
    any_on = any_on || on[u];
  }
  s_row[t] = lane_ok ? obase : -1;
  s_on[t] = line_on ? 1 : 0;
  __syncthreads();
  // the block works if one of its lines is active for one of its bins
  const bool block_on = __syncthreads_or(any_on ? 1 : 0) != 0;

  // q of cells t0 .. t0+TX-1 of all 64 lines, loaded into registers one
  // tile ahead of its use: all loads of a tile are issued together (cells
  // outside [rmin, rmax] and missing lines read element 0 instead, and are
  // never used), and land in LDS when the sweep enters the tile
  double pf[NB][TX];
  auto issue_tile = [&](int t0) {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      if (!bin_on[u]) continue;  // uniform
#pragma unroll
      for (int m = 0; m < TX; ++m) {
        const int f = m * 64 + t, row = f / TX, col = f % TX, cell = t0 + col;
        const long long rb = s_row[row];
        const bool ok = rb >= 0 && cell >= rmin && cell <= rmax;
        pf[u][m] = c[ok ? rb + (cell + 1) + o.ncell * (long long)(n0 + u) : 0];
      }
    }
  };
  auto commit_tile = [&]() {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      if (!bin_on[u]) continue;  // uniform
#pragma unroll
      for (int m = 0; m < TX; ++m) {
        const int f = m * 64 + t, row = f / TX, col = f % TX;
        s_t[u][row][col] = pf[u][m];
      }
    }
  };
  // store the solution of cells lo .. lo+TX-1 (active cells of active
  // lines; the other lanes store to junk)
  auto store_tile = [&](int lo) {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      if (!bin_on[u]) continue;  // uniform
#pragma unroll
      for (int m = 0; m < TX; ++m) {
        const int f = m * 64 + t, row = f / TX, col = f % TX, cell = lo + col;
        const long long rb = s_row[row];
        const bool ok = rb >= 0 && s_on[row] && cell >= rmin && cell <= rmax;
        *(ok ? c + rb + (cell + 1) + o.ncell * (long long)(n0 + u) : junk) = s_t[u][row][col];
      }
    }
  };

  if (block_on && n > 1) {
    if (any_on) {
      const int singular = A.sing[gid];
      if (singular) atomicCAS(A.flag, 0, singular);
    }
    const double th = A.theta;
    const double *Lf = A.lr, *Rf = A.lr + d.ncell;
    issue_tile(rmin);
    commit_tile();
    __syncthreads();
    int tlo = rmin;  // first cell of the tile in LDS
    if (rmin + TX <= rmax) issue_tile(rmin + TX);
    double q_l[NB], q_c[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      q_c[u] = s_t[u][t][0];
      q_l[u] = q_c[u];  // BC_left_t0 = q(rmin) (:822)
    }
    double Bc[NB];
    // ring of row factors (L, R of row l; fact, piv of step l-1), PF rows
    // ahead, the last row again past the end
    double rL[PF], rR[PF], rF[PF];
    uint8_t rP[PF];
    auto fetch_lr = [&](int slot, int l) {
      const long long oo = base + sL * (long long)(rmin + l);  // cell rmin-1+l
      rL[slot] = Lf[oo];
      rR[slot] = Rf[oo];
      rF[slot] = fF[oo - sL];
      rP[slot] = fP[oo - sL];
    };
#pragma unroll
    for (int p = 0; p < PF; ++p) fetch_lr(p, min(1 + p, n));
    for (int l0 = 1; l0 <= n; l0 += PF) {
#pragma unroll
      for (int p = 0; p < PF; ++p) {
        const int l = l0 + p;
        if (l > n) break;
        const double LeftFac = rL[p], RightFac = rR[p], fact = rF[p];
        const bool pv = rP[p] != 0;
        fetch_lr(p, min(l + PF, n));
        double q_next[NB];
        if (l < n) {
          if (rmin + l >= tlo + TX) {  // next tile (uniform: l and n are)
            __syncthreads();
            tlo += TX;
            commit_tile();
            __syncthreads();
            if (tlo + TX <= rmax) issue_tile(tlo + TX);
          }
#pragma unroll
          for (int u = 0; u < NB; ++u) q_next[u] = s_t[u][t][rmin + l - tlo];
        } else {
#pragma unroll
          for (int u = 0; u < NB; ++u) q_next[u] = 0.0;
        }
        const double CenterFac = LeftFac + RightFac;
        double Bn[NB], q_r[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          q_r[u] = (l < n) ? q_next[u] : q_c[u];  // BC_right_t0 (:845)
          Bn[u] = (1.0 - th) * LeftFac * q_l[u] + (1.0 - (1.0 - th) * CenterFac) * q_c[u] +
                  (1.0 - th) * RightFac * q_r[u];
        }
        if (l == 1) {
#pragma unroll
          for (int u = 0; u < NB; ++u) Bc[u] = Bn[u];
        } else {
          // elimination step i = l-1 (see k_diff_cn): row i is final
          const long long fo = base + sL * (long long)(rmin + l - 1);
#pragma unroll
          for (int u = 0; u < NB; ++u) {
            const double y = pv ? Bn[u] : Bc[u];
            const double x = pv ? Bc[u] : Bn[u];
            *(on[u] ? sb + fo + d.ncell * (long long)(n0 + u) : junk) = y;
            Bc[u] = x - fact * y;
          }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          q_l[u] = q_c[u];
          q_c[u] = q_r[u];
        }
      }
    }
    // ---- back substitution; solutions collect in an LDS tile covering
    // cells [thi-TX+1, thi], stored when the sweep leaves it
    __syncthreads();  // the q tile is no longer read
    int thi = rmax;
    double x1[NB], x2[NB];
    {
      const double Dn_ = fD[base + sL * (long long)(rmin + n)];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        x1[u] = Bc[u] / Dn_;
        x2[u] = 0.0;
        s_t[u][t][TX - 1] = x1[u];  // cell rmax at column TX-1 of tile [rmax-TX+1, rmax]
      }
    }
    double bD[PF], bDU[PF], bDL[PF], bB[PF][NB];
    auto fetch_row = [&](int slot, int i) {
      const long long fo = base + sL * (long long)(rmin + i);
      bD[slot] = fD[fo];
      bDU[slot] = fDU[fo];
      bDL[slot] = fDL[fo];
#pragma unroll
      for (int u = 0; u < NB; ++u) bB[slot][u] = sb[on[u] ? fo + d.ncell * (long long)(n0 + u) : 0];
    };
#pragma unroll
    for (int p = 0; p < PF; ++p) fetch_row(p, max(n - 1 - p, 1));
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
        fetch_row(p, max(i - PF, 1));
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
  if constexpr (NB >= 8) {
    return launch_cn_xrows_nb<NB, 8>(s, a, x);
  } else if constexpr (NB >= 4) {
    return tx <= 8 ? launch_cn_xrows_nb<NB, 8>(s, a, x) : launch_cn_xrows_nb<NB, 16>(s, a, x);
  } else {
    if (tx <= 8) return launch_cn_xrows_nb<NB, 8>(s, a, x);
    return tx <= 16 ? launch_cn_xrows_nb<NB, 16>(s, a, x) : launch_cn_xrows_nb<NB, 32>(s, a, x);
  }
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

hipError_t launch_line_lu(hipStream_t s, const DiffArgs &a, double *lu, uint8_t *piv, int *sing, int *ndom,
                          unsigned long long *rho, uint8_t *jl, unsigned long long *rho_j, int *njl) {
  if ((jl == nullptr) != (rho_j == nullptr) || (jl == nullptr) != (njl == nullptr)) return hipErrorInvalidValue;
  const long long lines = (long long)a.g.Na * a.g.Nb;
  dim3 grid((unsigned)((lines + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, s, ndom);
  if (rho) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, rho);
  if (rho_j) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, rho_j);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, s, njl);
  }
  switch (a.g.dir) {
    case 0:
      hipLaunchKernelGGL((k_line_lu<0>), grid, dim3(kBlock), 0, s, a, lu, piv, sing, ndom, rho, jl, rho_j, njl);
      break;
    case 1:
      hipLaunchKernelGGL((k_line_lu<1>), grid, dim3(kBlock), 0, s, a, lu, piv, sing, ndom, rho, jl, rho_j, njl);
      break;
    default:
      hipLaunchKernelGGL((k_line_lu<2>), grid, dim3(kBlock), 0, s, a, lu, piv, sing, ndom, rho, jl, rho_j, njl);
      break;
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

namespace {

template <int DIR, int NB, bool ADJ>
hipError_t launch_jacobi_k(hipStream_t s, const DiffArgs &a, int k, dim3 grid) {
  switch (k) {
    case 0: hipLaunchKernelGGL((k_cn_jacobi<DIR, NB, 0, ADJ>), grid, dim3(kBlock), 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_cn_jacobi<DIR, NB, 1, ADJ>), grid, dim3(kBlock), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_cn_jacobi<DIR, NB, 2, ADJ>), grid, dim3(kBlock), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_cn_jacobi<DIR, NB, 3, ADJ>), grid, dim3(kBlock), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_cn_jacobi<DIR, NB, 4, ADJ>), grid, dim3(kBlock), 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_cn_jacobi<DIR, NB, 5, ADJ>), grid, dim3(kBlock), 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_cn_jacobi<DIR, NB, 6, ADJ>), grid, dim3(kBlock), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int NB, bool ADJ>
hipError_t launch_jacobi_nb(hipStream_t s, const DiffArgs &a, int k) {
  const long long lines = (long long)a.g.Na * a.g.Nb;
  dim3 grid((unsigned)(padded_tiles((lines + kBlock - 1) / kBlock) * ((a.d.ns + NB - 1) / NB)));
  return a.g.dir == 1 ? launch_jacobi_k<1, NB, ADJ>(s, a, k, grid) : launch_jacobi_k<2, NB, ADJ>(s, a, k, grid);
}

template <int WV, int M>
hipError_t launch_jacobi_x_wv(hipStream_t s, const DiffArgs &a, const XJac &x, bool adjoint) {
  const dim3 grid((unsigned)((long long)a.g.Na * a.g.Nb)), block(64 * WV);
  if (adjoint)
    hipLaunchKernelGGL((k_cn_jacobi_x<WV, M, true>), grid, block, 0, s, a, x);
  else
    hipLaunchKernelGGL((k_cn_jacobi_x<WV, M, false>), grid, block, 0, s, a, x);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_diff_cn_jacobi(hipStream_t s, const DiffArgs &a, int nb, int k, bool adjoint) {
  if (a.g.dir != 1 && a.g.dir != 2) return hipErrorInvalidValue;
  if (k < 0 || k > kJacobiMaxIters) return hipErrorInvalidValue;
  if (adjoint) {
    // (the adjoint keeps more rows in flight: at most 2 bins per thread)
    return nb == 1 ? launch_jacobi_nb<1, true>(s, a, k) : launch_jacobi_nb<2, true>(s, a, k);
  }
  switch (nb) {
    case 1: return launch_jacobi_nb<1, false>(s, a, k);
    case 2: return launch_jacobi_nb<2, false>(s, a, k);
    default: return launch_jacobi_nb<4, false>(s, a, k);
  }
}

hipError_t launch_diff_cn_jacobi_x(hipStream_t s, const DiffArgs &a, const Dims &orig, double *c_orig,
                                   const double *kap, const double *sig, const double *kx, bool adjoint) {
  if (a.g.dir != 1 || !a.jline) return hipErrorInvalidValue;
  const int n = a.g.rmax - a.g.rmin + 1;
  const XJac x{orig, c_orig, kap, sig, kx};
  // rows per lane (global grid, 1440-cell lines: M = 3 6.67 ms, 4 7.02,
  // 6 6.63, 8 7.16)
  constexpr int M = 6;
  switch ((n + 64 * M - 1) / (64 * M)) {
    case 1: return launch_jacobi_x_wv<1, M>(s, a, x, adjoint);
    case 2: return launch_jacobi_x_wv<2, M>(s, a, x, adjoint);
    case 3: return launch_jacobi_x_wv<3, M>(s, a, x, adjoint);
    case 4: return launch_jacobi_x_wv<4, M>(s, a, x, adjoint);
    case 5: return launch_jacobi_x_wv<5, M>(s, a, x, adjoint);
    case 6: return launch_jacobi_x_wv<6, M>(s, a, x, adjoint);
    case 7: return launch_jacobi_x_wv<7, M>(s, a, x, adjoint);
    case 8: return launch_jacobi_x_wv<8, M>(s, a, x, adjoint);
    default: return hipErrorInvalidValue;
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
