// advect_x (src/AdvectionHorz_DCU.f90:68-305) over whole x rows, in two
// forms that give the same bits:
//
// k_advect_x_wave: one wavefront per row, the row held in registers as
// 16-byte cell pairs with the neighbours moved by DPP wave rotations (the
// default where nx is even and the row fits three chunks of 128 cells);
//
// k_advect_x_rows: one workgroup per row.  x rows are contiguous in
// memory, so lanes map to cells of the row: every load and store of
// concen is a coalesced row segment.  The row of one bin is staged through
// LDS and the update is built in three LDS phases (dq; fss, fs_l, fs_r;
// the cell update), each neighbour value read from LDS.
//
// Both loop over all grain bins, so the bin-independent terms of the row
// (interface speeds usig, dt/kappa, the divergence corrections, the
// topographic factors) are computed once per row.  All global reads of a
// row and bin happen before its in-place writes, and no other wavefront or
// workgroup touches the row.
//
// Arithmetic: the reference expressions in the reference order, no
// contraction -- bitwise equal to the streaming kernel and the oracle.
#include <cstring>

#include "common.hpp"
#include "dcu.hpp"
#include "lanes.hpp"

namespace ash3d {

namespace {

template <int CPT, bool TOPO>
__device__ __forceinline__ void advect_x_rows_body(const AdvArgs &A, const long long vb) {
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

  const int j = (int)(vb % d.ny) + 1;
  const int k = (int)(vb / d.ny) + 1;
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
      if (valid[m] && c >= rmin - 2 && c <= rmax + 2 && (!A.zg || (c >= 1 && c <= nx)))
        q[m] = c0[base + (c + 1) + bofs];
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


// ---------------------------------------------------------------------
// One wavefront per x row, the row in registers.  Lane l of chunk m holds
// E cells c = -1 + 64 E m + E l + e (e < E): E = 2, the cell pair loaded
// and stored as one 16-byte access (nx even: every pair is two ghost or
// two interior cells); E = 1 for odd nx.  The neighbours c-1 and c+1 come
// from the same lane (the other element of a pair) or from the
// next/previous lane by a whole-wave DPP rotation, with the rotated value
// of the neighbouring chunk at lanes 0 and 63 and zero beyond the row
// (what the LDS version reads there).  The per-row terms wait in a
// wave-private LDS region; no barriers: the row of the next bin is loaded
// while the current one is updated.  Every load and store of the bin loop is issued on every
// iteration (clamped addresses, masked-off pairs stored to junk), so the
// wait before a row is used counts exactly the memory operations issued
// after its loads.  The outflow increments of the row ends are gathered in
// lane n for bin n and added after the loop.
//
// The arithmetic is that of k_advect_x_rows expression by expression.

template <int NC, int E, bool TOPO>
__device__ __forceinline__ void advect_x_wave_body(const AdvArgs &A, const long long vb) {
  constexpr int CS = 64 * E;  // cells per chunk
  // a lane's E consecutive cells (one 16-byte access for a pair)
  auto ld_cells = [](const double *src, double (&out)[E]) {
    if constexpr (E == 2) {
      const double2 v = *(const double2 *)src;
      out[0] = v.x;
      out[1] = v.y;
    } else {
      out[0] = *src;
    }
  };
  auto st_cells = [](double *dst, const double (&v)[E]) {
    if constexpr (E == 2) *(double2 *)dst = make_double2(v[0], v[1]);
    else *dst = v[0];
  };
  const Dims d = A.d;
  const LineGeom g = A.g;
  const int nx = d.nx, W = nx + 4;
  const int l = (int)(threadIdx.x & 63);
  const long long row = vb * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= (long long)d.ny * d.nz) return;  // whole wavefront
  const int j = (int)(row % d.ny) + 1;
  const int k = (int)(row / d.ny) + 1;
  const int rmin = g.rmin, rmax = g.rmax;
  const bool row_on = j >= g.a0 && j <= g.a1 && k >= g.b0 && k <= g.b1;
  const long long base = d.p3(-1, j, k);  // cell -1 of the row
  const double *c0 = A.c0;
  double *co = A.cout;  // == c0 for the in-place sweep
  double *junk = A.junk + 2 * (long long)(threadIdx.x + blockDim.x * (vb % (kJunk / 512)));
  const bool lane0 = l == 0, lane63 = l == 63;

  // per element: cell, row position, masks (bin-independent)
  int cc[NC][E];
  bool valid[NC][E], ld[NC][E], uin[NC][E], lq[NC][E];
  int pp[NC];       // row position of the pair, clamped into the row
  bool store[NC];   // an interior pair (both cells in 1..nx)
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    const int p0 = CS * m + E * l;
    pp[m] = min(p0, W - E);
    store[m] = p0 >= 2 && p0 + E - 1 <= nx + 1;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int c = p0 + e - 1;
      cc[m][e] = c;
      valid[m][e] = p0 + e < W;
      ld[m][e] = valid[m][e] && c >= rmin - 2 && c <= rmax + 2;
      uin[m][e] = valid[m][e] && c >= rmin - 1 && c <= rmax + 1;
      // concentrations gathered (A.zg: ghosts read as Set_BC(1)'s zeros)
      lq[m][e] = ld[m][e] && (!A.zg || (c >= 1 && c <= nx));
    }
  }

  if (!row_on) {
    for (int n = 0; n < d.ns; ++n) {
      const long long bofs = d.ncell * (long long)n;
#pragma unroll
      for (int m = 0; m < NC; ++m)
        if (store[m]) {
          const double z[E] = {};
          st_cells(co + base + pp[m] + bofs, z);
        }
    }
    return;
  }

  // ---- bin-independent terms of the row
  double u[NC][E], dtv[NC][E], divp[NC][E], divm[NC][E];
  // the four per-element row terms wait in LDS for the bin loop (wave-private:
  // 4 * 2*NC values per lane, element-major), which frees their registers
  extern __shared__ double s_row[];
  double *srow = s_row + (long long)(threadIdx.x >> 6) * (4 * E * NC * 64) + l;
  auto rt = [&](int which, int m, int e) -> double & { return srow[((which * NC + m) * E + e) * 64]; };
  double tf[NC][E], D[NC][E], tfm1[NC][E], Dm1[NC][E];
  {
    double vel[NC][E], sig[NC][E], kap[NC][E];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const long long o = base + pp[m];
      double vv[E], ss[E], kk[E], dd[E] = {};
      ld_cells(A.vel + o, vv);
      ld_cells(A.sig + o, ss);
      ld_cells(A.kap + o, kk);
      if (TOPO) ld_cells(A.deld + (long long)pp[m] + (long long)(nx + 4) * (j + 1), dd);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        vel[m][e] = ld[m][e] ? vv[e] : 0.0;
        sig[m][e] = ld[m][e] ? ss[e] : 0.0;
        kap[m][e] = ld[m][e] ? kk[e] : 1.0;
        D[m][e] = (TOPO && uin[m][e]) ? dd[e] : 0.0;
      }
    }
    double velm1[NC][E], sigp1[NC][E];
    row_prev<NC>(vel, velm1, lane0);
    if (TOPO) row_next<NC>(sig, sigp1, lane63);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        // usig (:159-162), dt/kappa (:155-156)
        u[m][e] = uin[m][e] ? 0.5 * (velm1[m][e] + vel[m][e]) * sig[m][e] : 0.0;
        dtv[m][e] = uin[m][e] ? A.dt / kap[m][e] : 0.0;
        // topographic factor 0.5*vel_I*(sig_I+sig_{I+1}) (:252-254)
        tf[m][e] = (TOPO && valid[m][e]) ? 0.5 * vel[m][e] * (sig[m][e] + sigp1[m][e]) : 0.0;
      }
    double um1[NC][E], up1[NC][E];
    row_prev<NC>(u, um1, lane0);
    row_next<NC>(u, up1, lane63);
    if (TOPO) {
      row_prev<NC>(tf, tfm1, lane0);
      row_prev<NC>(D, Dm1, lane0);
    }
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        // divergence corrections (:245-246)
        divp[m][e] = (fmax(0.0, u[m][e]) - fmax(0.0, um1[m][e]));
        divm[m][e] = (fmin(0.0, up1[m][e]) - fmin(0.0, u[m][e]));
        rt(0, m, e) = u[m][e];
        rt(1, m, e) = dtv[m][e];
        rt(2, m, e) = divp[m][e];
        rt(3, m, e) = divm[m][e];
        if (!TOPO) tfm1[m][e] = Dm1[m][e] = 0.0;
      }
  }

  // outflow increments: lane n keeps those of bin n (at most 64 bins)
  const bool of_lo = rmin == 1 && !A.periodic;
  const bool of_hi = rmax == nx && !A.periodic;
  const int m_hi = (nx + 2) / CS, l_hi = ((nx + 2) % CS) / E;  // cell nx+1: element 0 (nx even for pairs)
  double acc_lo = 0.0, acc_hi = 0.0;

  const int lim = A.limiter;
  // two bins in flight: the rows of bins n and n+1 are in slots A and B;
  // slot A is refilled with bin n+2 as soon as bin n is gathered
  double qa[NC][E], qb[NC][E];
  auto fetch = [&](double (&qs)[NC][E], int n) {
    // a bin past the end or one that is off: a cached re-read of bin n-2
    const int nn = (n < d.ns && A.aloft.on(n)) ? n : max(0, min(n - 2, d.ns - 1));
    const long long bofs = d.ncell * (long long)nn;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      ld_cells(c0 + base + pp[m] + bofs, qs[m]);
    }
  };
  auto bin = [&](double (&qs)[NC][E], int n) {
    const bool on = A.aloft.on(n);
    double q[NC][E];
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) q[m][e] = lq[m][e] ? qs[m][e] : 0.0;  // gather (:143)
    fetch(qs, n + 2);
    // dq (:168-169)
    double qm1[NC][E], dq[NC][E];
    row_prev<NC>(q, qm1, lane0);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) dq[m][e] = valid[m][e] ? q[m][e] - qm1[m][e] : 0.0;
    double dqm[NC][E], dqp[NC][E];
    row_prev<NC>(dq, dqm, lane0);
    row_next<NC>(dq, dqp, lane63);
    // limited flux and fluctuations at interface I = c (:178-259)
    double fss[NC][E], fsr[NC][E], fsl[NC][E];
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int c = cc[m][e];
        fss[m][e] = 0.0; fsr[m][e] = 0.0; fsl[m][e] = 0.0;
        if (lim != 0 && valid[m][e] && c >= rmin && c <= rmax + 1)
          fss[m][e] = limited_flux(lim, dqm[m][e], dq[m][e], dqp[m][e], rt(0, m, e), rt(1, m, e));
        if (uin[m][e]) {
          const double uu = rt(0, m, e);
          double r = fmax(0.0, uu) * dq[m][e];
          double t = fmin(0.0, uu) * dq[m][e];
          r = r + q[m][e] * rt(3, m, e);
          t = t + qm1[m][e] * rt(2, m, e);
          if (TOPO) {
            r = r + tf[m][e] * q[m][e] * D[m][e];
            t = t - tfm1[m][e] * qm1[m][e] * Dm1[m][e];
          }
          fsr[m][e] = r;
          fsl[m][e] = t;
        }
      }
    double fsln[NC][E], fssn[NC][E];
    row_next<NC>(fsl, fsln, lane63);
    row_next<NC>(fss, fssn, lane63);
    // cell update (:264-281), interior write and outflow (:285-294)
    double upd[NC][E], res[NC][E];
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int c = cc[m][e];
        upd[m][e] = uin[m][e] ? -rt(1, m, e) * (fsln[m][e] + fsr[m][e] + fssn[m][e] - fss[m][e]) : 0.0;
        res[m][e] = (on && c >= rmin && c <= rmax) ? q[m][e] + upd[m][e] : 0.0;
      }
    const long long bofs = d.ncell * (long long)n;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      double *dst = store[m] ? co + base + pp[m] + bofs : junk;
      st_cells(dst, res[m]);
    }
    if (on) {
      // cell 0 (row position 1) is element 1 % E of lane 1 / E, chunk 0; cell
      // nx+1 element 0 of lane l_hi, chunk m_hi
      double uh = upd[0][0];
#pragma unroll
      for (int m = 0; m < NC; ++m)
        if (m == m_hi) uh = upd[m][0];
      const double lo = lane_value(upd[0][1 % E], 1 / E);
      const double hi = lane_value(uh, l_hi);
      if (l == n) { acc_lo = lo; acc_hi = hi; }
    }
  };
  fetch(qa, 0);
  fetch(qb, 1);
  int n = 0;
  for (; n + 1 < d.ns; n += 2) {
    bin(qa, n);
    bin(qb, n + 1);
  }
  if (n < d.ns) bin(qa, n);
  // outflow (:285-294): bin n from lane n
  if (l < d.ns && A.aloft.on(l)) {
    const long long of = (long long)(j + 1) + (long long)(d.ny + 4) * ((long long)(k + 1) + (long long)(d.nz + 4) * l);
    if (of_lo) A.of1[of] = A.of1[of] + acc_lo;
    if (of_hi) A.of2[of] = A.of2[of] + acc_hi;
  }
}


// ---------------------------------------------------------------------
// Rows longer than one wavefront's registers hold (nx even): WV waves per
// row in one workgroup, wave w owning the row positions [S, S + L) with
// S = w L (L even).  Each wave holds the extended range [S - 2, S + L + 2)
// in the layout of k_advect_x_wave (NC chunks of 64 cell pairs, lane l of
// chunk m: positions S - 2 + 128 m + 2 l + e), so every cell it stores has
// its +-2 stencil in its own registers and the arithmetic is that kernel's.
// The per-cell fields of the extended range are read from memory; the
// concentrations of the two halo pairs are not, because the neighbouring
// wave updates them in place: every wave writes the first and the last
// pair of its own range to LDS (two slots, by bin parity) and reads its
// halos from there after a barrier, one barrier per bin.  Global loads and
// stores of concentrations touch only the wave's own range.
struct SegArgs {
  AdvArgs A;
  int L;  // even segment length
};

template <int NC, bool TOPO>
__device__ __forceinline__ void advect_x_seg_body(const AdvArgs &A, const int L, const long long vb) {
  constexpr int E = 2, CS = 128;
  auto ld_cells = [](const double *src, double (&out)[E]) {
    const double2 v = *(const double2 *)src;
    out[0] = v.x;
    out[1] = v.y;
  };
  auto st_cells = [](double *dst, const double (&v)[E]) { *(double2 *)dst = make_double2(v[0], v[1]); };
  const Dims d = A.d;
  const LineGeom g = A.g;
  const int nx = d.nx, W = nx + 4;
  const int l = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6), WV = (int)(blockDim.x >> 6);
  const long long row = vb;
  const int j = (int)(row % d.ny) + 1;
  const int k = (int)(row / d.ny) + 1;
  const int rmin = g.rmin, rmax = g.rmax;
  const bool row_on = j >= g.a0 && j <= g.a1 && k >= g.b0 && k <= g.b1;
  const long long base = d.p3(-1, j, k);  // cell -1 of the row
  const double *c0 = A.c0;
  double *co = A.cout;  // == c0 for the in-place sweep
  double *junk = A.junk + 2 * (long long)(threadIdx.x + blockDim.x * (vb % (kJunk / 1024)));
  const bool lane0 = l == 0, lane63 = l == 63;
  const int S = w * L, Send = min(S + L, W), X = S - 2;
  // the halo pairs (positions S-2, S-1 and S+L, S+L+1) and the own edge
  // pairs published to the neighbours (S, S+1 and S+L-2, S+L-1)
  const bool has_lo = w > 0, has_hi = w + 1 < WV;
  const int m_hh = (L + 2) / CS, l_hh = ((L + 2) % CS) / 2;  // upper halo pair
  const int m_le = L / CS, l_le = (L % CS) / 2;              // own last pair

  int cc[NC][E];
  bool valid[NC][E], ld[NC][E], uin[NC][E], lq[NC][E];
  int pf[NC], pq[NC];  // field and concentration offsets of the pair (clamped)
  bool store[NC], halo_lo[NC], halo_hi[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    const int p0 = X + CS * m + 2 * l;
    const bool own = p0 >= S && p0 < Send;
    pf[m] = min(max(p0, 0), W - 2);
    pq[m] = own ? p0 : S;
    store[m] = own && p0 >= 2 && p0 + 1 <= nx + 1;
    halo_lo[m] = has_lo && m == 0 && l == 0;
    halo_hi[m] = has_hi && m == m_hh && l == l_hh;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int c = p0 + e - 1;
      cc[m][e] = c;
      valid[m][e] = p0 + e >= 0 && p0 + e < W && p0 + e < Send + 2;
      ld[m][e] = valid[m][e] && c >= rmin - 2 && c <= rmax + 2;
      uin[m][e] = valid[m][e] && c >= rmin - 1 && c <= rmax + 1;
      lq[m][e] = ld[m][e] && (!A.zg || (c >= 1 && c <= nx));
    }
  }

  if (!row_on) {
    for (int n = 0; n < d.ns; ++n) {
      const long long bofs = d.ncell * (long long)n;
#pragma unroll
      for (int m = 0; m < NC; ++m)
        if (store[m]) {
          const double z[E] = {};
          st_cells(co + base + pq[m] + bofs, z);
        }
    }
    return;
  }

  extern __shared__ double s_row[];
  // wave-private row terms (4 * 2*NC values per lane), then the edge pairs:
  // s_edge[((slot * WV + wave) * 2 + first/last) * 2 + e]
  double *srow = s_row + (long long)w * (4 * E * NC * 64) + l;
  double *s_edge = s_row + (long long)WV * (4 * E * NC * 64);
  auto rt = [&](int which, int m, int e) -> double & { return srow[((which * NC + m) * E + e) * 64]; };
  double tf[NC][E], D[NC][E], tfm1[NC][E], Dm1[NC][E];
  {
    double u[NC][E], dtv[NC][E];
    double vel[NC][E], sig[NC][E], kap[NC][E];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const long long o = base + pf[m];
      double vv[E], ss[E], kk[E], dd[E] = {};
      ld_cells(A.vel + o, vv);
      ld_cells(A.sig + o, ss);
      ld_cells(A.kap + o, kk);
      if (TOPO) ld_cells(A.deld + (long long)pf[m] + (long long)(nx + 4) * (j + 1), dd);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        vel[m][e] = ld[m][e] ? vv[e] : 0.0;
        sig[m][e] = ld[m][e] ? ss[e] : 0.0;
        kap[m][e] = ld[m][e] ? kk[e] : 1.0;
        D[m][e] = (TOPO && uin[m][e]) ? dd[e] : 0.0;
      }
    }
    double velm1[NC][E], sigp1[NC][E];
    row_prev<NC>(vel, velm1, lane0);
    if (TOPO) row_next<NC>(sig, sigp1, lane63);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        // usig (:159-162), dt/kappa (:155-156)
        u[m][e] = uin[m][e] ? 0.5 * (velm1[m][e] + vel[m][e]) * sig[m][e] : 0.0;
        dtv[m][e] = uin[m][e] ? A.dt / kap[m][e] : 0.0;
        // topographic factor 0.5*vel_I*(sig_I+sig_{I+1}) (:252-254)
        tf[m][e] = (TOPO && valid[m][e]) ? 0.5 * vel[m][e] * (sig[m][e] + sigp1[m][e]) : 0.0;
      }
    double um1[NC][E], up1[NC][E];
    row_prev<NC>(u, um1, lane0);
    row_next<NC>(u, up1, lane63);
    if (TOPO) {
      row_prev<NC>(tf, tfm1, lane0);
      row_prev<NC>(D, Dm1, lane0);
    }
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        // divergence corrections (:245-246)
        rt(0, m, e) = u[m][e];
        rt(1, m, e) = dtv[m][e];
        rt(2, m, e) = (fmax(0.0, u[m][e]) - fmax(0.0, um1[m][e]));
        rt(3, m, e) = (fmin(0.0, up1[m][e]) - fmin(0.0, u[m][e]));
        if (!TOPO) tfm1[m][e] = Dm1[m][e] = 0.0;
      }
  }

  // outflow increments (wave 0 owns cell 0, the last wave cell nx+1): lane
  // n keeps those of bin n (at most 64 bins)
  const bool of_lo = rmin == 1 && !A.periodic && w == 0;
  const bool of_hi = rmax == nx && !A.periodic && w == WV - 1;
  const int o_hi = nx + 2 - X;  // extended offset of cell nx+1 (element 0) in the last wave
  const int m_hi = o_hi / CS, l_hi = (o_hi % CS) / E;
  double acc_lo = 0.0, acc_hi = 0.0;

  const int lim = A.limiter;
  double qa[NC][E], qb[NC][E];
  auto fetch = [&](double (&qs)[NC][E], int n) {
    const int nn = (n < d.ns && A.aloft.on(n)) ? n : max(0, min(n - 2, d.ns - 1));
    const long long bofs = d.ncell * (long long)nn;
#pragma unroll
    for (int m = 0; m < NC; ++m) ld_cells(c0 + base + pq[m] + bofs, qs[m]);
  };
  auto bin = [&](double (&qs)[NC][E], int n) {
    const bool on = A.aloft.on(n);
    double *edge = s_edge + (long long)(n & 1) * WV * 4;
    // publish the own first and last pairs, then take the halos
    if (l == 1) {
      edge[(w * 2 + 0) * 2 + 0] = qs[0][0];
      edge[(w * 2 + 0) * 2 + 1] = qs[0][1];
    }
#pragma unroll
    for (int m = 0; m < NC; ++m)
      if (m == m_le && l == l_le) {
        edge[(w * 2 + 1) * 2 + 0] = qs[m][0];
        edge[(w * 2 + 1) * 2 + 1] = qs[m][1];
      }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    double hlo[E] = {}, hhi[E] = {};
    if (has_lo) {
      hlo[0] = edge[((w - 1) * 2 + 1) * 2 + 0];
      hlo[1] = edge[((w - 1) * 2 + 1) * 2 + 1];
    }
    if (has_hi) {
      hhi[0] = edge[((w + 1) * 2 + 0) * 2 + 0];
      hhi[1] = edge[((w + 1) * 2 + 0) * 2 + 1];
    }
    double q[NC][E];
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const double v = halo_lo[m] ? hlo[e] : (halo_hi[m] ? hhi[e] : qs[m][e]);
        q[m][e] = lq[m][e] ? v : 0.0;  // gather (:143)
      }
    fetch(qs, n + 2);
    // dq (:168-169)
    double qm1[NC][E], dq[NC][E];
    row_prev<NC>(q, qm1, lane0);
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) dq[m][e] = valid[m][e] ? q[m][e] - qm1[m][e] : 0.0;
    double dqm[NC][E], dqp[NC][E];
    row_prev<NC>(dq, dqm, lane0);
    row_next<NC>(dq, dqp, lane63);
    // limited flux and fluctuations at interface I = c (:178-259)
    double fss[NC][E], fsr[NC][E], fsl[NC][E];
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int c = cc[m][e];
        fss[m][e] = 0.0; fsr[m][e] = 0.0; fsl[m][e] = 0.0;
        if (lim != 0 && valid[m][e] && c >= rmin && c <= rmax + 1)
          fss[m][e] = limited_flux(lim, dqm[m][e], dq[m][e], dqp[m][e], rt(0, m, e), rt(1, m, e));
        if (uin[m][e]) {
          const double uu = rt(0, m, e);
          double r = fmax(0.0, uu) * dq[m][e];
          double t = fmin(0.0, uu) * dq[m][e];
          r = r + q[m][e] * rt(3, m, e);
          t = t + qm1[m][e] * rt(2, m, e);
          if (TOPO) {
            r = r + tf[m][e] * q[m][e] * D[m][e];
            t = t - tfm1[m][e] * qm1[m][e] * Dm1[m][e];
          }
          fsr[m][e] = r;
          fsl[m][e] = t;
        }
      }
    double fsln[NC][E], fssn[NC][E];
    row_next<NC>(fsl, fsln, lane63);
    row_next<NC>(fss, fssn, lane63);
    // cell update (:264-281), interior write and outflow (:285-294)
    double upd[NC][E], res[NC][E];
#pragma unroll
    for (int m = 0; m < NC; ++m)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int c = cc[m][e];
        upd[m][e] = uin[m][e] ? -rt(1, m, e) * (fsln[m][e] + fsr[m][e] + fssn[m][e] - fss[m][e]) : 0.0;
        res[m][e] = (on && c >= rmin && c <= rmax) ? q[m][e] + upd[m][e] : 0.0;
      }
    const long long bofs = d.ncell * (long long)n;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      double *dst = store[m] ? co + base + pq[m] + bofs : junk;
      st_cells(dst, res[m]);
    }
    if (on) {
      // cell 0 (position 1): element 1 of lane 1, chunk 0 of wave 0; cell
      // nx+1: element 0 of lane l_hi, chunk m_hi of the last wave
      double uh = upd[0][0];
#pragma unroll
      for (int m = 0; m < NC; ++m)
        if (m == m_hi) uh = upd[m][0];
      const double lo = lane_value(upd[0][1], 1);
      const double hi = lane_value(uh, l_hi < 64 ? l_hi : 0);
      if (l == n) { acc_lo = lo; acc_hi = hi; }
    }
  };
  fetch(qa, 0);
  fetch(qb, 1);
  int n = 0;
  for (; n + 1 < d.ns; n += 2) {
    bin(qa, n);
    bin(qb, n + 1);
  }
  if (n < d.ns) bin(qa, n);
  // outflow (:285-294): bin n from lane n
  if (l < d.ns && A.aloft.on(l)) {
    const long long of = (long long)(j + 1) + (long long)(d.ny + 4) * ((long long)(k + 1) + (long long)(d.nz + 4) * l);
    if (of_lo) A.of1[of] = A.of1[of] + acc_lo;
    if (of_hi) A.of2[of] = A.of2[of] + acc_hi;
  }
}

// one kernel per body: a launch for one member and a batched launch over
// the members of an ensemble (workgroup index within the member's launch)
template <int CPT, bool TOPO>
__global__ __launch_bounds__(256) void k_advect_x_rows(AdvArgs A) {
  advect_x_rows_body<CPT, TOPO>(A, blockIdx.x);
}
template <int CPT, bool TOPO>
__global__ __launch_bounds__(256) void k_advect_x_rows_batch(
    const ASH3D_CONSTANT AdvArgs *__restrict__ args, const ASH3D_CONSTANT int *__restrict__ member,
    const ASH3D_CONSTANT long long *__restrict__ first) {
  const int m = member[blockIdx.x];
  const AdvArgs &A = *(const AdvArgs *)(args + m);
  advect_x_rows_body<CPT, TOPO>(A, (long long)blockIdx.x - first[m]);
}
template <int NC, int E, bool TOPO>
__global__ __launch_bounds__(256) void k_advect_x_wave(AdvArgs A) {
  advect_x_wave_body<NC, E, TOPO>(A, blockIdx.x);
}
template <int NC, int E, bool TOPO>
__global__ __launch_bounds__(256) void k_advect_x_wave_batch(
    const ASH3D_CONSTANT AdvArgs *__restrict__ args, const ASH3D_CONSTANT int *__restrict__ member,
    const ASH3D_CONSTANT long long *__restrict__ first) {
  const int m = member[blockIdx.x];
  const AdvArgs &A = *(const AdvArgs *)(args + m);
  advect_x_wave_body<NC, E, TOPO>(A, (long long)blockIdx.x - first[m]);
}
template <int NC, bool TOPO>
__global__ __launch_bounds__(512) void k_advect_x_seg(SegArgs S) {
  advect_x_seg_body<NC, TOPO>(S.A, S.L, blockIdx.x);
}
template <int NC, bool TOPO>
__global__ __launch_bounds__(512) void k_advect_x_seg_batch(
    const ASH3D_CONSTANT SegArgs *__restrict__ args, const ASH3D_CONSTANT int *__restrict__ member,
    const ASH3D_CONSTANT long long *__restrict__ first) {
  const int m = member[blockIdx.x];
  const SegArgs &S = *(const SegArgs *)(args + m);
  advect_x_seg_body<NC, TOPO>(S.A, S.L, (long long)blockIdx.x - first[m]);
}

template <int CPT>
hipError_t launch_cpt(hipStream_t s, const AdvArgs &a, int bt, LaunchSpec *sp) {
  const size_t shmem = sizeof(double) * 6 * (size_t)(a.d.nx + 4);
  dim3 grid((unsigned)((long long)a.d.ny * a.d.nz));
  const bool topo = a.deld != nullptr;
  return issue(s, sp, topo ? (const void *)k_advect_x_rows<CPT, true> : (const void *)k_advect_x_rows<CPT, false>,
               topo ? (const void *)k_advect_x_rows_batch<CPT, true> : (const void *)k_advect_x_rows_batch<CPT, false>,
               grid, dim3(bt), shmem, a);
}

template <int NC, int E>
hipError_t launch_wave(hipStream_t s, const AdvArgs &a, LaunchSpec *sp) {
  const long long rows = (long long)a.d.ny * a.d.nz;
  dim3 grid((unsigned)((rows + 3) / 4));
  const size_t lds = sizeof(double) * 4 * 4 * E * NC * 64;  // 4 waves x 4 row terms x E*NC per lane
  const bool topo = a.deld != nullptr;
  return issue(s, sp,
               topo ? (const void *)k_advect_x_wave<NC, E, true> : (const void *)k_advect_x_wave<NC, E, false>,
               topo ? (const void *)k_advect_x_wave_batch<NC, E, true>
                    : (const void *)k_advect_x_wave_batch<NC, E, false>,
               grid, dim3(256), lds, a);
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// Segmented rows: WV = ceil(W / (128 NC - 4)) waves, L the even segment
// length; the dynamic LDS holds each wave's row terms and the edge pairs.
template <int NC>
hipError_t launch_seg(hipStream_t s, const AdvArgs &a, LaunchSpec *sp) {
  const int W = a.d.nx + 4;
  const int cap = 128 * NC - 4;
  const int WV = (W + cap - 1) / cap;
  if (WV > 8) return hipErrorInvalidValue;
  int L = (W + WV - 1) / WV;
  L += L & 1;
  if (L > cap || (WV - 1) * L >= W) return hipErrorInvalidValue;
  const size_t lds = sizeof(double) * ((size_t)WV * 4 * 2 * NC * 64 + 2 * (size_t)WV * 4);
  dim3 grid((unsigned)((long long)a.d.ny * a.d.nz));
  const bool topo = a.deld != nullptr;
  return issue(s, sp, topo ? (const void *)k_advect_x_seg<NC, true> : (const void *)k_advect_x_seg<NC, false>,
               topo ? (const void *)k_advect_x_seg_batch<NC, true> : (const void *)k_advect_x_seg_batch<NC, false>,
               grid, dim3(64 * WV), lds, SegArgs{a, L});
}

}  // namespace

// Rows of up to 8*256-4 cells; returns hipErrorInvalidValue beyond that so
// the caller falls back to the streaming kernel.  One wavefront per row
// (odd nx: single cells per lane, 6 chunks of 64 at most) where the row
// fits 3 chunks of 128 cells in pairs (nx even, nx <= 380;
// with topography 2 chunks, nx <= 252: longer rows leave one wave per SIMD
// and spill, and run slower than the workgroup kernel), the outflow
// increments fit one lane per bin (ns <= 64) and the arrays allow 16-byte
// accesses.
static hipError_t advect_x_rows_any(hipStream_t s, const AdvArgs &a, LaunchSpec *sp) {
  const int cells = a.d.nx + 4;
  const bool common = a.xrows == 1 && cells <= (a.deld ? 256 : 384) && a.d.ns <= 64;
  if (common && a.d.nx % 2 == 0 && aligned16(a.c0) && aligned16(a.cout) && aligned16(a.vel) &&
      aligned16(a.sig) && aligned16(a.kap) && aligned16(a.junk) && (a.deld == nullptr || aligned16(a.deld))) {
    switch ((cells + 127) / 128) {  // cell pairs
      case 1: return launch_wave<1, 2>(s, a, sp);
      case 2: return launch_wave<2, 2>(s, a, sp);
      default: return launch_wave<3, 2>(s, a, sp);
    }
  }
  // longer rows (nx even): several waves per row, a segment each
  if (a.xrows == 1 && a.d.ns <= 64 && a.d.nx % 2 == 0 && !common && aligned16(a.c0) && aligned16(a.cout) &&
      aligned16(a.vel) && aligned16(a.sig) && aligned16(a.kap) && aligned16(a.junk) &&
      (a.deld == nullptr || aligned16(a.deld))) {
    hipError_t e = a.deld != nullptr ? launch_seg<2>(s, a, sp) : launch_seg<3>(s, a, sp);
    if (e != hipErrorInvalidValue) return e;
    (void)hipGetLastError();
  }
  if (common) {
    switch ((cells + 63) / 64) {  // single cells (odd nx)
      case 1: return launch_wave<1, 1>(s, a, sp);
      case 2: return launch_wave<2, 1>(s, a, sp);
      case 3: return launch_wave<3, 1>(s, a, sp);
      case 4: return launch_wave<4, 1>(s, a, sp);
      case 5: return launch_wave<5, 1>(s, a, sp);
      default: return launch_wave<6, 1>(s, a, sp);
    }
  }
  int cpt = (cells + 255) / 256;
  if (cpt > 8) return hipErrorInvalidValue;
  if (cpt == 3) cpt = 4;
  if (cpt > 4) cpt = 8;
  int bt = (cells + cpt - 1) / cpt;
  bt = ((bt + 63) / 64) * 64;
  switch (cpt) {
    case 1: return launch_cpt<1>(s, a, bt, sp);
    case 2: return launch_cpt<2>(s, a, bt, sp);
    case 4: return launch_cpt<4>(s, a, bt, sp);
    default: return launch_cpt<8>(s, a, bt, sp);
  }
}

hipError_t launch_advect_x_rows(hipStream_t s, const AdvArgs &a) { return advect_x_rows_any(s, a, nullptr); }

hipError_t spec_advect_x_rows(const AdvArgs &a, LaunchSpec *sp) { return advect_x_rows_any(nullptr, a, sp); }

}  // namespace ash3d
