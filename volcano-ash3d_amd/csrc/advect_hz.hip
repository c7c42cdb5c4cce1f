// The second horizontal advection sweep and the vertical sweep of a step in
// one pass (AdvectHorz's second direction, then AdvectVert:
// src/AdvectionHorz_DCU.f90:68-305 or :322-556, then
// src/AdvectionVert_DCU.f90:100-323), for the full-grid step whose sweeps
// read concentration ghosts as Set_BC(1)'s zeros (the step with
// diffusion, whose Set_BC(2) and Neumann passes rewrite every ghost
// afterwards).
//
// Each wave owns a set of columns -- 64 columns x TY rows for y lines,
// 64 cells of one row for x lines -- for NB grain bins and streams them
// upward through the levels.  At level k a lane loads the horizontal
// sweep's input along its line -- the first sweep's output, which that
// sweep wrote out of place to the ts1 slot -- with the two-cell halo
// along the sweep direction (rows j-2..j+TY+1 for y lines; columns
// i-2..i+2 for x lines, wrapped on a periodic grid; the neighbouring
// waves' cells come through L1), evaluates the horizontal update of its
// cells, and feeds the result into a register window that runs the
// vertical sweep of the column exactly as the streamed z kernel does (cell
// k-2 is updated once level k is known).  The updated column goes to ts0.
// So the pair costs one read of the first sweep's output and one write of
// the step's, instead of a read and a write for each sweep.  No LDS and
// no barriers: every wave is independent.
//
// Every update is the reference expression evaluated in the same order on
// the same inputs as the split sweeps (built with -ffp-contract=off), so
// the result is bitwise theirs.  Outflow increments of the horizontal line
// ends are formed by the tiles that own those ends; the deposit and top
// outflow by every column.
#include <algorithm>

#include "common.hpp"
#include "dcu.hpp"

namespace ash3d {

namespace {

// Updates of the R consecutive cells c0..c0+R-1 of a line of N interior
// cells (full active range) from q at c0-2..c0+R+1 (masked as the split
// sweep gathers them), vel at c0-1..c0+R, sigma and dt/kappa at c0..c0+R
// and DelD at c0..c0+R-1.  The interface terms (usig, dt/kappa, the limited
// flux) are formed once and shared by the two cells of each interface; the
// expressions and their order are those of the streamed sweep (advect.hip).
// out = q + update; upd = the update.
template <int R, bool TOPO>
__device__ inline void dcu_segment(int c0, int N, int lim, const double (&q)[R + 4], const double (&vel)[R + 2],
                                   const double (&sig)[R + 1], const double (&dtvc)[R + 1], const double (&D)[R],
                                   double (&out)[R], double (&upd)[R]) {
  double u[R + 1], dtv[R + 1], fss[R + 1], dq[R + 3];
#pragma unroll
  for (int e = 0; e < R + 3; ++e) dq[e] = q[e + 1] - q[e];  // dq at c0-1+e
#pragma unroll
  for (int m = 0; m <= R; ++m) {
    const int I = c0 + m;
    const bool uin = I >= 0 && I <= N + 1;
    u[m] = uin ? 0.5 * (vel[m] + vel[m + 1]) * sig[m] : 0.0;
    dtv[m] = uin ? dtvc[m] : 0.0;
  }
#pragma unroll
  for (int m = 0; m <= R; ++m) {
    const int I = c0 + m;
    fss[m] = (lim != 0 && I >= 1 && I <= N + 1) ? limited_flux(lim, dq[m], dq[m + 1], dq[m + 2], u[m], dtv[m]) : 0.0;
  }
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int c = c0 + t;
    double fsr = 0.0, fsl = 0.0, T = 0.0;
    if (c >= 0 && c <= N + 1) {
      fsr = fmax(0.0, u[t]) * dq[t + 1];
      const double divu_m = (fmin(0.0, u[t + 1]) - fmin(0.0, u[t]));
      fsr = fsr + q[t + 2] * divu_m;
      if (TOPO) {
        T = 0.5 * vel[t + 1] * (sig[t] + sig[t + 1]) * q[t + 2] * D[t];
        fsr = fsr + T;
      }
    }
    if (c + 1 >= 0 && c + 1 <= N + 1) {
      fsl = fmin(0.0, u[t + 1]) * dq[t + 2];
      const double divu_p = (fmax(0.0, u[t + 1]) - fmax(0.0, u[t]));
      fsl = fsl + q[t + 2] * divu_p;
      if (TOPO) fsl = fsl - T;
    }
    upd[t] = -dtv[t] * (fsl + fsr + fss[t + 1] - fss[t]);
    out[t] = q[t + 2] + upd[t];
  }
}

// One column of the streamed vertical sweep (advect.hip, DIR = 2) for one
// bin: cell r enters with its value q0 (zero outside the gathered range
// [-1, nz+2] and for the ghost levels, which read as Set_BC(1)'s zeros),
// sigma, vz + vh and vf; the update of cell r-2 is formed.  dt/kappa of
// the column's cells r-1 and r-2 (bin-independent) is kept by the caller,
// once for all bins of the column (ZDtv).  dq1 is q1 - q2, recomputed
// (the same operation on the same values as the dq0 of the step before).
struct ZDtv {
  double d1 = 0.0, d2 = 0.0;  // dt / kappa of cells r-1, r-2
  // dt / kappa of cell r as the sweep gathers it (zero outside [0, nz+1])
  __device__ static inline double gather(int r, int nz, double dtvc) { return (r >= 0 && r <= nz + 1) ? dtvc : 0.0; }
  __device__ inline void shift(double d0) {
    d2 = d1;
    d1 = d0;
  }
};
struct ZWin {
  double q1 = 0.0, q2 = 0.0, dq2 = 0.0, fsr2 = 0.0, fss2 = 0.0;
  double vb1 = 0.0, ub1 = 0.0, ub2 = 0.0;
  // returns the new value of cell r-2 (meaningful for r-2 in [1, nz]); the
  // update itself in *edge when r-2 is 0 or nz+1 (the bottom and top
  // outflow increments)
  __device__ inline double step(int r, int nz, int lim, double q0, double sig0, const ZDtv &dt, double vzh0,
                                double vfc, double *edge) {
    const bool ld = r >= -1 && r <= nz + 2;
    if (!ld) {
      q0 = 0.0;
      sig0 = 0.0;
      vzh0 = 0.0;
      vfc = 0.0;
    }
    const bool uin = r >= 0 && r <= nz + 1;
    const double dtv1 = dt.d1, dtv2 = dt.d2;
    const double vb0 = ld ? vzh0 + vfc : 0.0;
    const double ub0 = uin ? 0.5 * (vb1 + vb0) * sig0 : 0.0;
    const double dq0 = q0 - q1;
    const double dq1 = q1 - q2;
    const int I = r - 1, cc = r - 2;
    double fssI = 0.0, fsrI = 0.0, fslI = 0.0;
    if (lim != 0 && I >= 1 && I <= nz + 1) fssI = limited_flux(lim, dq2, dq1, dq0, ub1, dtv1);
    if (I >= 0 && I <= nz + 1) {
      fsrI = fmax(0.0, ub1) * dq1;
      fslI = fmin(0.0, ub1) * dq1;
      const double divu_p = (fmax(0.0, ub1) - fmax(0.0, ub2));
      const double divu_m = (fmin(0.0, ub0) - fmin(0.0, ub1));
      fsrI = fsrI + q1 * divu_m;
      fslI = fslI + q2 * divu_p;
    }
    double out = 0.0;
    if (cc >= 0 && cc <= nz + 1) {
      const double upd = -dtv2 * (fslI + fsr2 + fssI - fss2);
      out = q2 + upd;
      if (cc == 0 || cc == nz + 1) *edge = upd;
    }
    fsr2 = fsrI;
    fss2 = fssI;
    q2 = q1;
    q1 = q0;
    dq2 = dq1;
    ub2 = ub1;
    ub1 = ub0;
    vb1 = vb0;
    return out;
  }
};

// Tiles and waves.  y lines second (DIR2 = 1): a workgroup of 4 waves owns
// 64 columns (lane l: column i0+l) x 4*TY rows, wave w the rows
// j0 + w*TY .. + TY-1, each lane its TY cells of every level: the y
// stencil of those cells is in the lane's own registers (rows j-2..j+2
// loaded per level; the rows of the neighbouring waves come through L1).
// x lines second (DIR2 = 0): a workgroup of 4 waves owns 256 columns of
// one row, wave w the columns i0 + 64w .., lane l one cell, its x
// neighbours loaded from the row (through L1), wrapped on a periodic
// grid.  Each wave takes
// NB consecutive bins, so that the bin-independent fields of a level are
// loaded once for them.  Every wave streams its cells up the levels on
// its own (no barriers); the vertical sweep of each owned column and bin
// runs in a ZWin.
constexpr int kHzWaves = 4;

template <int DIR2, bool TOPO, int TY, int NB, bool WL>
__device__ __forceinline__ void advect_hz_body(const AdvArgs &H, const AdvArgs &Z, const long long vb,
                                               const long long vn) {
  const Dims d = Z.d;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  constexpr int R = DIR2 == 1 ? TY : 1;  // rows per lane
  constexpr int COLS = (WL ? 64 * kHzWaves : 64);           // columns of a workgroup's tile
  constexpr int ROWS = WL ? R : (DIR2 == 1 ? kHzWaves * TY : kHzWaves);  // rows of a workgroup's tile
  const int tiles_x = (nx + COLS - 1) / COLS, tiles_y = (ny + ROWS - 1) / ROWS;
  const int G = (d.ns + NB - 1) / NB;
  const TileGroup tg = xcd_tile_group(vb, G, vn);
  if (tg.tile >= (long long)tiles_x * tiles_y) return;  // padding workgroup (uniform)
  const int n0 = tg.group * NB;
  const int l = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  // WL: the waves side by side along x (the rows a workgroup reads are
  // contiguous over 64 * waves cells); else stacked along y
  const int i0w = (int)(tg.tile % tiles_x) * COLS + 1 + (WL ? 64 * w : 0);  // first column of this wave
  const int i = i0w + l;
  const int jw = (int)(tg.tile / tiles_x) * ROWS + 1 + (WL ? 0 : w * R);  // first row of this wave
  if (i0w > nx) return;  // the whole wave is past the grid (no barriers follow)
  const int ic = min(i, nx);
  const bool lane_in = i <= nx;
  const int lim = Z.limiter;
  const double *dtvp = Z.dtv;  // dt / kappa per cell
  const bool xper = DIR2 == 0 && H.periodic;
  const int sy = (int)d.sy, sz = (int)d.sz;
  const long long ofs_xy = (long long)(nx + 4) * (ny + 4);
  auto p3 = [&](int a, int b, int c) { return (a + 1) + sy * (b + 1) + sz * (c + 1); };
  if (jw > ny) return;  // the whole wave is past the grid (no barriers follow)

  bool own[R];
  int jr[R];  // rows of the owned cells (clamped into the grid for the others' reads)
#pragma unroll
  for (int t = 0; t < R; ++t) {
    own[t] = lane_in && jw + t <= ny;
    jr[t] = min(jw + t, ny);
  }
  // the bins' slices (uniform bases; cell offsets within a bin fit 32 bits,
  // the launcher checks ncell < 2^31); a bin past ns repeats the last one
  // and stores nothing
  const double *src[NB];
  double *dst[NB];
  const double *vfb[NB];
  bool ex[NB], on[NB];
  bool any_on = false;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int n = min(n0 + b, d.ns - 1);
    ex[b] = n0 + b < d.ns;
    on[b] = ex[b] && Z.aloft.on(n);
    any_on = any_on || on[b];
    const long long bofs = d.ncell * (long long)n;
    src[b] = H.c0 + bofs;  // the first sweep's output (ts1)
    dst[b] = Z.cout + bofs;  // ts0
    vfb[b] = Z.vf + bofs;
  }
  // bins that are not aloft: both sweeps leave every interior cell zero
  // (DepositGranularity is still copied from the unchanged outflow)
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (!ex[b] || on[b]) continue;
    const int n = n0 + b;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      if (!own[t]) continue;
      for (int k = 1; k <= nz; ++k) dst[b][p3(i, jr[t], k)] = 0.0;
      const long long of = (long long)(i + 1) + (long long)(nx + 4) * (jr[t] + 1) + ofs_xy * n;
      Z.dep[(long long)(i - 1) + (long long)nx * ((long long)(jr[t] - 1) + (long long)ny * n)] = Z.of1[of];
    }
  }
  if (!any_on) return;

  // topographic factor DelD at the owned cells (2-D, bin- and level-independent)
  double Dw[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    Dw[t] = 0.0;
    if (TOPO) Dw[t] = H.deld[(long long)(ic + 1) + (long long)(nx + 4) * (jr[t] + 1)];
  }

  ZWin zw[NB][R];
  ZDtv zd[R];
  // the bottom and top outflow increments of an owned column of an aloft
  // bin, added as the sweep forms them (src/AdvectionVert_DCU.f90:306-311);
  // the deposit is the bottom outflow (:321)
  auto edge = [&](int b, int t, int cc, double upd) {
    if (!(on[b] && own[t])) return;
    const int n = n0 + b;
    const long long of = (long long)(i + 1) + (long long)(nx + 4) * (jr[t] + 1) + ofs_xy * n;
    if (cc == 0) {
      Z.of1[of] = Z.of1[of] + upd;
      Z.dep[(long long)(i - 1) + (long long)nx * ((long long)(jr[t] - 1) + (long long)ny * n)] = Z.of1[of];
    } else {
      Z.of2[of] = Z.of2[of] + upd;
    }
  };
  // the vertical sweep steps of the ghost levels and past the top
  auto zghost = [&](int r) {
    const int rc = min(max(r, -1), nz + 2);  // clamped into the column
    const int cc = r - 2;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const int c = p3(ic, jr[t], rc);
      const double zs = Z.sig[c], dv = dtvp[c], vzh = Z.vel[c] + Z.vh[c];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        double e = 0.0;
        const double v = zw[b][t].step(r, nz, lim, 0.0, zs, zd[t], vzh, vfb[b][c], &e);
        if (on[b] && own[t] && cc >= 1 && cc <= nz) dst[b][p3(ic, jr[t], cc)] = v;
        if (cc == 0 || cc == nz + 1) edge(b, t, cc, e);
      }
      zd[t].shift(ZDtv::gather(r, nz, dv));
    }
  };
  zghost(-2);
  zghost(-1);
  zghost(0);

  const int hN = DIR2 == 1 ? ny : nx;
  constexpr int QN = R + 4, VN = R + 2, SN = R + 1;
  for (int k = 1; k <= nz; ++k) {
    // the inputs of level k: along the horizontal lines q at the cells
    // c-2..c+2 of the owned cells, vel c-1..c+1, sigma and dt/kappa c..c+1
    double q[NB][QN], vel[VN], sig[SN], dv[SN];
    if constexpr (DIR2 == 1) {
      const int c0 = p3(ic, 0, k) - sy;  // + sy * (r + 1) is row r of column ic
#pragma unroll
      for (int t = 0; t < QN; ++t) {
        const int r = jw - 2 + t;
        const bool v = lane_in && r >= 1 && r <= ny;
        const int rc = min(r, ny + 2);  // rows -1..ny+2 exist: every lane loads
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const double x = src[b][c0 + sy * (rc + 1)];
          q[b][t] = v ? x : 0.0;
        }
      }
#pragma unroll
      for (int t = 0; t < VN; ++t) vel[t] = H.vel[c0 + sy * (min(jw - 1 + t, ny + 2) + 1)];
#pragma unroll
      for (int t = 0; t < SN; ++t) {
        const int c = c0 + sy * (min(jw + t, ny + 2) + 1);
        sig[t] = H.sig[c];
        dv[t] = dtvp[c];
      }
    } else {
      const int c0 = p3(0, jr[0], k) - 1;  // + ci + 1 is cell ci of the row
#pragma unroll
      for (int t = 0; t < QN; ++t) {
        int ci = ic - 2 + t;
        if (xper) ci = ci < 1 ? ci + nx : (ci > nx ? ci - nx : ci);
        const bool v = lane_in && ci >= 1 && ci <= nx;  // ci in -1..nx+2: every lane loads
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const double x = src[b][c0 + ci + 1];
          q[b][t] = v ? x : 0.0;
        }
      }
#pragma unroll
      for (int t = 0; t < VN; ++t) vel[t] = H.vel[c0 + ic + t];
#pragma unroll
      for (int t = 0; t < SN; ++t) {
        sig[t] = H.sig[c0 + ic + 1 + t];
        dv[t] = dtvp[c0 + ic + 1 + t];
      }
    }
    double hv[NB][R];  // horizontal update of the owned cells at level k
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      double up[R];
      dcu_segment<R, TOPO>(DIR2 == 1 ? jw : ic, hN, lim, q[b], vel, sig, dv, Dw, hv[b], up);
    }
    // outflow of the line ends (non-periodic): the ghost updates at 0 and
    // N+1, formed by the owners of cells 1 and N
    if (!xper) {
#pragma unroll
      for (int t = 0; t < R; ++t) {
        const int hc = DIR2 == 1 ? jr[t] : ic;
        const bool lo = own[t] && hc == 1, hi = own[t] && hc == hN;
        if (!lo && !hi) continue;
        auto hcell = [&](int r) { return DIR2 == 1 ? p3(ic, r, k) : p3(r, jr[t], k); };
        double vg[3], sg[2], kg[2], Dg[1] = {0.0};
        const int cg = lo ? 0 : hN + 1;
#pragma unroll
        for (int e = 0; e < 3; ++e) vg[e] = H.vel[hcell(cg - 1 + e)];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          sg[e] = H.sig[hcell(cg + e)];
          kg[e] = dtvp[hcell(cg + e)];
        }
        if (TOPO)
          Dg[0] = DIR2 == 1 ? H.deld[(long long)(ic + 1) + (long long)(nx + 4) * (cg + 1)]
                            : H.deld[(long long)(cg + 1) + (long long)(nx + 4) * (jr[t] + 1)];
        const int a = DIR2 == 1 ? i : jr[t];
        const long long na = DIR2 == 1 ? (long long)(nx + 4) : (long long)(ny + 4);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (!on[b]) continue;
          double qg[5], og[1], u[1];
#pragma unroll
          for (int e = 0; e < 5; ++e) {
            const int r = cg - 2 + e;
            qg[e] = (r >= 1 && r <= hN) ? src[b][hcell(r)] : 0.0;
          }
          dcu_segment<1, TOPO>(cg, hN, lim, qg, vg, sg, kg, Dg, og, u);
          const long long of = (long long)(a + 1) + na * (long long)(k + 1) + na * (nz + 4) * (long long)(n0 + b);
          double *o = lo ? H.of1 : H.of2;
          o[of] = o[of] + u[0];
        }
      }
    }
    // the vertical sweep: level k enters, cell k-2 is updated (its dt/kappa
    // is the horizontal one of the same cell)
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const int c = p3(ic, jr[t], k);
      const double zs = Z.sig[c], vzh = Z.vel[c] + Z.vh[c];
      const double dvz = DIR2 == 1 ? dv[t] : dv[0];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        double e = 0.0;
        const double v = zw[b][t].step(k, nz, lim, hv[b][t], zs, zd[t], vzh, vfb[b][c], &e);
        if (on[b] && own[t] && k - 2 >= 1) dst[b][c - 2 * sz] = v;
        if (k == 2) edge(b, t, 0, e);
      }
      zd[t].shift(ZDtv::gather(k, nz, dvz));
    }
  }
  zghost(nz + 1);
  zghost(nz + 2);
  zghost(nz + 3);
}

struct HzArgs {
  AdvArgs h, z;
};

#ifndef ASH3D_HZ_TY
#define ASH3D_HZ_TY 2
#endif
#ifndef ASH3D_HZ_NB
#define ASH3D_HZ_NB 2
#endif
#ifndef ASH3D_HZ_NBX
#define ASH3D_HZ_NBX 3
#endif
constexpr int kHzTYy = ASH3D_HZ_TY;  // rows per lane, y lines second
// bins per wave: y lines second, x lines second (measured on 1440x720x80x20:
// two bins with two rows per lane for y+z, three bins for x+z)
template <int DIR2>
constexpr int kHzNB = DIR2 == 1 ? ASH3D_HZ_NB : ASH3D_HZ_NBX;
// waves of a workgroup: stacked along y for y lines (the halo rows of a
// wave are the neighbouring waves' rows, read through the same L1;
// side by side: 17.7 against 15.7 ms), side by side along x for x lines
// (a workgroup reads 256 contiguous cells of each row; 14.3 against 14.4)
template <int DIR2>
constexpr bool kHzWL = DIR2 == 0;
template <int DIR2, bool TOPO>
__global__ __launch_bounds__(64 * kHzWaves) void k_advect_hz(HzArgs A) {
  advect_hz_body<DIR2, TOPO, kHzTYy, kHzNB<DIR2>, kHzWL<DIR2>>(A.h, A.z, blockIdx.x, gridDim.x);
}
template <int DIR2, bool TOPO>
__global__ __launch_bounds__(64 * kHzWaves) void k_advect_hz_batch(
    const ASH3D_CONSTANT HzArgs *__restrict__ args, const ASH3D_CONSTANT int *__restrict__ member,
    const ASH3D_CONSTANT long long *__restrict__ first) {
  const int m = member[blockIdx.x];
  const HzArgs &A = *(const HzArgs *)(args + m);
  advect_hz_body<DIR2, TOPO, kHzTYy, kHzNB<DIR2>, kHzWL<DIR2>>(A.h, A.z, (long long)blockIdx.x - first[m],
                                                               first[m + 1] - first[m]);
}

template <int DIR2>
hipError_t hz_dir(hipStream_t s, const AdvArgs &h, const AdvArgs &z, LaunchSpec *sp) {
  const int r = DIR2 == 1 ? kHzTYy : 1;
  const int cols = kHzWL<DIR2> ? 64 * kHzWaves : 64, rows = kHzWL<DIR2> ? r : kHzWaves * r;
  const long long tiles = (long long)((z.d.nx + cols - 1) / cols) * ((z.d.ny + rows - 1) / rows);
  const dim3 grid((unsigned)(padded_tiles(tiles) * ((z.d.ns + kHzNB<DIR2> - 1) / kHzNB<DIR2>))), block(64 * kHzWaves);
  const bool topo = h.deld != nullptr;
  return issue(s, sp, topo ? (const void *)k_advect_hz<DIR2, true> : (const void *)k_advect_hz<DIR2, false>,
               topo ? (const void *)k_advect_hz_batch<DIR2, true> : (const void *)k_advect_hz_batch<DIR2, false>,
               grid, block, 0, HzArgs{h, z});
}

hipError_t hz_any(hipStream_t s, const AdvArgs &h, const AdvArgs &z, LaunchSpec *sp) {
  if (h.g.dir != 0 && h.g.dir != 1) return hipErrorInvalidValue;
  if (z.g.dir != 2 || z.limiter != h.limiter || z.dt != h.dt || z.d.ns != h.d.ns) return hipErrorInvalidValue;
  if (z.d.ncell >= (1LL << 31)) return hipErrorInvalidValue;  // 32-bit cell offsets within a bin
  if (!z.dtv) return hipErrorInvalidValue;
  return h.g.dir == 1 ? hz_dir<1>(s, h, z, sp) : hz_dir<0>(s, h, z, sp);
}

__global__ __launch_bounds__(256) void k_cell_dtv(const double *kap, double *out, long long n, double dt) {
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long long)gridDim.x * 256)
    out[c] = dt / kap[c];
}

}  // namespace

hipError_t launch_cell_dtv(hipStream_t s, const double *kap, double *out, long long n, double dt) {
  const long long b = std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_cell_dtv, dim3((unsigned)std::max<long long>(b, 1)), dim3(256), 0, s, kap, out, n, dt);
  return hipGetLastError();
}

hipError_t launch_advect_hz(hipStream_t s, const AdvArgs &h, const AdvArgs &z) { return hz_any(s, h, z, nullptr); }

hipError_t spec_advect_hz(const AdvArgs &h, const AdvArgs &z, LaunchSpec *sp) { return hz_any(nullptr, h, z, sp); }

}  // namespace ash3d
