// Partitioned theta-scheme line solves (the fast alternative to the
// dgtsv-order solves in diffuse.hip).  Same systems as diffCN_{x,y,z}
// (src/Diffusion.F90:671-1428): per line the tridiagonal matrix of rows
// l = 1..n (cells rmin..rmax)
//   a_l = -(th*L_l) (l > 1),  c_l = -(th*R_l) (l < n),
//   b_l = 1 + th*C_l, minus th*L_1 on the first row and th*R_n on the last,
//   d_l = (1-th)*L_l*q_{l-1} + (1-(1-th)*C_l)*q_l + (1-th)*R_l*q_{l+1}
// (q_0 = q_1, q_{n+1} = q_n).  These matrices are strictly row diagonally
// dominant (|b| >= 1 + |a| + |c| for L, R >= 0), so elimination needs no
// pivoting and may be reordered: the result agrees with dgtsv's to
// rounding (about 1e-14 relative), not bit for bit.
//
// Each line is cut into LP segments of M rows, one lane per segment (rows
// past n are identity rows).  A lane eliminates its segment so that every
// row reads  ah_k*x_first + x_k + ch_k*x_last = dh_k  (the "modified
// Thomas" step of Laszlo, Giles & Appleyard, ACM TOMS 42(4), 2016); the
// first and last rows of all segments form a tridiagonal system of 2*LP
// unknowns, solved by parallel cyclic reduction across the LP lanes
// (Hockney's PCR, log2(2*LP) levels of lane shuffles); the interior rows
// then follow from their segment's two end values.  Everything that
// depends on the matrix only (segment factors, PCR multipliers) is formed
// once per line and kept in registers while the workgroup runs the bins
// of its lines one after another; per bin a line costs one read of q and
// one write of the result.  A bin's tile of lines is staged through LDS
// (coalesced both ways) while the next bin's tile is already in flight.
// Sweeps with any line that is not a diagonally dominant M-matrix (negative
// diffusivities, non-finite factors) are left to the dgtsv-order kernels.
//
// Lines longer than one wave's 64 x M rows (WV > 1) are split into WV
// parts, one wave each, and solved by the SPIKE algorithm (Polizzi &
// Sameh, Parallel Computing 32, 2006) on top of the scheme above: every
// part drops its coupling to the neighbouring parts (the sub-diagonal
// entry alpha of its first row, the super-diagonal entry gamma of its
// last row) and solves its own system A_w y = d with the partitioned
// solve; the spikes v = A_w^{-1} e_first and w = A_w^{-1} e_last depend on
// the matrix only and are formed once per line.  Then
//   x = y - alpha * x_last(previous part) * v - gamma * x_first(next part) * w,
// and the 2 WV end values x_first, x_last of the parts solve a small
// system R z = (y_first, y_last, ...) whose inverse is formed once per
// line; per bin the parts exchange their y end values through LDS (one
// barrier) and each lane corrects its rows.
#include <type_traits>

#include "common.hpp"
#include "lanes.hpp"

namespace ash3d {

namespace {

// Offsets of cell r of line (a, b): st0 + sta*a + stb*b + str*r in the
// state (bin stride ncell) and lr0 + lra*a + lrb*b + lrr*r in the row
// factor cache (L at lr[o], R at lr[o + lr_ncell]).
struct PartArgs {
  long long st0, sta, stb, str;
  long long lr0, lra, lrb, lrr, lr_ncell;
  const double *kap;  // cell volumes at the state's cell offsets (adjoint)
  int bins_wg;        // bins per workgroup
};

// Workgroup b runs on XCD b mod 8.  Each XCD takes a contiguous range of
// the (tile, bin group) items, in order, so that neighbouring tiles (whose
// rows can share cache lines) go through the same L2.
struct Item {
  long long tile;
  int group;
};
__device__ inline Item xcd_block_item(long long b, long long nwg, int G) {
  const long long per = nwg / kXcds;  // nwg is a multiple of kXcds
  const long long it = (b % kXcds) * per + b / kXcds;
  return Item{it / G, (int)(it % G)};
}

template <int LP>
struct Lev {
  static constexpr int value = (LP == 16) ? 5 : (LP == 32) ? 6 : 7;  // log2(2*LP)
};

struct CnPartArgs {
  DiffArgs A;
  PartArgs Q;
};

template <int LP, int M, int T, bool CONTIG, int PD, bool ADJ, int WV>
__device__ __forceinline__ void cn_part_body(const DiffArgs &A, const PartArgs &Q, const long long vb,
                                             const long long vn) {
  constexpr int LPL = LP * WV;  // lanes per line
  constexpr int NT = T * LPL;
  constexpr int NPAD = LPL * M;
  constexpr int PITCH = NPAD + 1;
  constexpr int LEV = Lev<LP>::value;
  constexpr int NR = 2 * WV;  // unknowns of the parts' reduced system
  static_assert(PD == 1 || PD == 2, "one or two bins in flight");
  static_assert(WV == 1 || LP == 64, "several parts per line: whole waves");
  extern __shared__ double s_tile[];  // T lines x PITCH
  // a sweep with lines that are not diagonally dominant M-matrix rows runs
  // the dgtsv-order kernel instead (launched after this one; uniform exit)
  if (*A.ndom != 0) return;
  const LineGeom g = A.g;
  const int ns = A.d.ns;
  const long long ncell = A.d.ncell;
  const int tiles_a = (g.Na + T - 1) / T;
  const int G = (ns + Q.bins_wg - 1) / Q.bins_wg;
  const Item tg = xcd_block_item(vb, vn, G);
  if (tg.tile >= (long long)tiles_a * g.Nb) return;  // padding workgroup (uniform)
  const int b = (int)(tg.tile / tiles_a) + 1;
  const int a_first = (int)(tg.tile % tiles_a) * T + 1;
  // lines of the tile this kernel solves (bit l: line a_first + l): those
  // of the grid, but not those the Jacobi sweeps take (1 <= A.jline <=
  // A.jcap); a tile with none of them exits (uniform)
  static_assert(T <= 32, "one bit per line of the tile");
  unsigned own_mask = 0;
  for (int l_ = 0; l_ < T; ++l_) {
    const int al_ = a_first + l_;
    const int sw = (A.jline && al_ <= g.Na) ? A.jline[(long long)(al_ - 1) + (long long)g.Na * (b - 1)] : 0;
    if (al_ <= g.Na && !(sw >= 1 && sw <= A.jcap)) own_mask |= 1u << l_;
  }
  if (own_mask == 0) return;
  auto owned = [&](int al_) { return ((own_mask >> (al_ - a_first)) & 1u) != 0; };
  const int tid = threadIdx.x;
  const int ln = tid / LPL;  // line of this lane within the tile
  const int p = tid % LPL;   // segment of this lane
  const int pw = p % LP;     // segment within the part (the PCR group)
  const int wv = p / LP;     // part of the line (wave)
  const int a = a_first + ln;
  const bool has_line = owned(a);
  const int ac = has_line ? a : g.Na;  // an existing line for masked lanes' reads
  const int rmin = g.rmin, rmax = g.rmax, n = rmax - rmin + 1, N = g.N;
  const bool line_on = has_line && a >= g.a0 && a <= g.a1 && b >= g.b0 && b <= g.b1;
  const int u0 = tg.group * Q.bins_wg;
  const int u1 = min(ns, u0 + Q.bins_wg);
  bool any_bin = false;
  for (int u = u0; u < u1; ++u) any_bin = any_bin || A.aloft.on(u);
  if (line_on && any_bin) {
    const int singular = A.sing[(long long)(a - 1) + (long long)g.Na * (b - 1)];
    if (singular) atomicCAS(A.flag, 0, singular);
  }
  const double th = A.theta;

  const double omt = 1.0 - th;
  // ---- per-line factors of this lane's segment (rows i0 = p*M + k)
  double rL[M], rR[M];          // row factors (zero on padding rows)
  double al[M], be[M];          // forward: dh_k = al_k*d_k - be_k*dh_{k-1}
  double ga[M];                 // backward: dh_k -= ga_k*dh_{k+1}
  double ah[M], ch[M];          // final rows: ah_k*x_first + x_k + ch_k*x_last = dh_k
  double rho;
  double cpl_a = 0.0, cpl_c = 0.0;  // dropped couplings (first lane / last lane of a part)
  // adjoint: kappa of each row and its inverse, in LDS after the tile
  // (row k of this lane at [k*NT + tid], its inverse at [(M+k)*NT + tid])
  double *s_kap = s_tile + T * PITCH;
  {
    double av[M], bv[M], cv[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int i0 = p * M + k;
      const bool real = i0 < n;
      const long long o = Q.lr0 + Q.lra * ac + Q.lrb * b + Q.lrr * (long long)(rmin + (real ? i0 : 0));
      const double L = A.lr[o], R = A.lr[o + Q.lr_ncell];
      const double C = L + R;
      const int l = i0 + 1;
      if (real) {
        av[k] = (l == 1) ? 0.0 : -(th * L);
        cv[k] = (l == n) ? 0.0 : -(th * R);
        double D = 1.0 + th * C;
        if (l == 1) D = D - th * L;
        if (l == n) D = D - th * R;
        bv[k] = D;
        rL[k] = L;
        rR[k] = R;
      } else {
        av[k] = 0.0; bv[k] = 1.0; cv[k] = 0.0;
        rL[k] = rR[k] = 0.0;
      }
      if (WV > 1) {
        // each part's own system: drop the couplings to the other parts
        if (k == 0 && pw == 0) { cpl_a = av[0]; av[0] = 0.0; }
        if (k == M - 1 && pw == LP - 1) { cpl_c = cv[M - 1]; cv[M - 1] = 0.0; }
      }
      if (ADJ) {
        const double kv = Q.kap[Q.st0 + Q.sta * ac + Q.stb * b + Q.str * (long long)(rmin + (real ? i0 : 0))];
        s_kap[k * NT + tid] = kv;
        s_kap[(M + k) * NT + tid] = 1.0 / kv;
      }
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
      if (k < 2) {
        const double inv = 1.0 / bv[k];
        al[k] = inv;
        be[k] = 0.0;
        ah[k] = av[k] * inv;
        ch[k] = cv[k] * inv;
      } else {
        const int km = k > 0 ? k - 1 : 0;
        const double r = 1.0 / (bv[k] - av[k] * ch[km]);
        al[k] = r;
        be[k] = r * av[k];
        ah[k] = -be[k] * ah[km];
        ch[k] = r * cv[k];
      }
      ga[k] = ch[k];
    }
#pragma unroll
    for (int k = M - 3; k >= 1; --k) {
      ah[k] = ah[k] - ga[k] * ah[k + 1];
      ch[k] = -ga[k] * ch[k + 1];
    }
    rho = 1.0 / (1.0 - ga[0] * ah[1]);
    ah[0] = rho * ah[0];
    ch[0] = -rho * ga[0] * ch[1];
  }

  // ---- PCR multipliers of the reduced system: lane p holds equations
  // 2p (first row: ah_0, 1, ch_0) and 2p+1 (last row: ah_{M-1}, 1, ch_{M-1})
  double k1[LEV][2], k2[LEV][2], invB[2];
  {
    double RA[2] = {ah[0], ah[M - 1]}, RB[2] = {1.0, 1.0}, RC[2] = {ch[0], ch[M - 1]};
#pragma unroll
    for (int lev = 0; lev < LEV; ++lev) {
      double mA[2], mB[2], mC[2], pA[2], pB[2], pC[2];
      if (lev == 0) {
        // equation 2p: neighbours 2p-1 (lane p-1, eq 1) and 2p+1 (own eq 1);
        // equation 2p+1: 2p (own eq 0) and 2p+2 (lane p+1, eq 0)
        const double uA = __shfl_up(RA[1], 1, LP), uB = __shfl_up(RB[1], 1, LP), uC = __shfl_up(RC[1], 1, LP);
        const double dA = __shfl_down(RA[0], 1, LP), dB = __shfl_down(RB[0], 1, LP), dC = __shfl_down(RC[0], 1, LP);
        const bool lo = pw >= 1, hi = pw + 1 < LP;
        mA[0] = lo ? uA : 0.0; mB[0] = lo ? uB : 1.0; mC[0] = lo ? uC : 0.0;
        pA[0] = RA[1]; pB[0] = RB[1]; pC[0] = RC[1];
        mA[1] = RA[0]; mB[1] = RB[0]; mC[1] = RC[0];
        pA[1] = hi ? dA : 0.0; pB[1] = hi ? dB : 1.0; pC[1] = hi ? dC : 0.0;
      } else {
        const int h = 1 << (lev - 1);
        const bool lo = pw >= h, hi = pw + h < LP;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const double uA = __shfl_up(RA[j], h, LP), uB = __shfl_up(RB[j], h, LP), uC = __shfl_up(RC[j], h, LP);
          const double dA = __shfl_down(RA[j], h, LP), dB = __shfl_down(RB[j], h, LP), dC = __shfl_down(RC[j], h, LP);
          mA[j] = lo ? uA : 0.0; mB[j] = lo ? uB : 1.0; mC[j] = lo ? uC : 0.0;
          pA[j] = hi ? dA : 0.0; pB[j] = hi ? dB : 1.0; pC[j] = hi ? dC : 0.0;
        }
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        k1[lev][j] = RA[j] / mB[j];
        k2[lev][j] = RC[j] / pB[j];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const double nB = RB[j] - k1[lev][j] * mC[j] - k2[lev][j] * pA[j];
        RA[j] = -k1[lev][j] * mA[j];
        RC[j] = -k2[lev][j] * pC[j];
        RB[j] = nB;
      }
    }
    invB[0] = 1.0 / RB[0];
    invB[1] = 1.0 / RB[1];
  }

  // ---- tile staging: elements f of the tile's T lines x n active cells
  constexpr int PF = M;  // ceil(T * NPAD / NT)
  // element f: (line l_, cell index i_) with consecutive f along the
  // direction that is contiguous in memory
  auto elem_l = [&](int f) { return CONTIG ? f / n : f % T; };
  auto elem_i = [&](int f) { return CONTIG ? f % n : f / T; };
  const int nel = T * n;
  double pf[PD][PF];  // bins in flight, one register slot each
  auto issue = [&](auto slot_tag, int u) {
    constexpr int sl = decltype(slot_tag)::value;
#pragma unroll
    for (int m = 0; m < PF; ++m) {
      const int f = tid + NT * m;
      const int l_ = elem_l(f), i_ = elem_i(f);
      const bool ok = f < nel && a_first + l_ <= g.Na;
      const long long o = Q.st0 + Q.sta * (a_first + l_) + Q.stb * b + Q.str * (long long)(rmin + i_) +
                          ncell * (long long)u;
      pf[sl][m] = A.c0[ok ? o : 0];
    }
  };
  auto commit = [&](auto slot_tag) {
    constexpr int sl = decltype(slot_tag)::value;
#pragma unroll
    for (int m = 0; m < PF; ++m) {
      const int f = tid + NT * m;
      if (f < nel) s_tile[elem_l(f) * PITCH + elem_i(f)] = pf[sl][m];
    }
  };
  auto next_on = [&](int u) {
    while (u < u1 && !A.aloft.on(u)) ++u;
    return u;
  };
  double *trow = s_tile + ln * PITCH;

  // ---- solve this lane's segment of its line from the tile in LDS and
  // leave the solution there
  // the solve of this part's system for the right-hand side d of the
  // lane's rows (segment elimination, PCR of the segment ends, interior)
  auto local = [&](const double (&d)[M], double (&xs)[M]) {
    double dh[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      dh[k] = al[k] * d[k];
      if (k >= 2) dh[k] = dh[k] - be[k] * dh[k > 0 ? k - 1 : 0];
    }
#pragma unroll
    for (int k = M - 3; k >= 1; --k) dh[k] = dh[k] - ga[k] * dh[k + 1];
    dh[0] = rho * (dh[0] - ga[0] * dh[1]);
    double D[2] = {dh[0], dh[M - 1]};
    auto level = [&](auto lev_tag) {
      constexpr int lev = decltype(lev_tag)::value;
      constexpr int h = (lev == 0) ? 1 : 1 << (lev - 1);
      double mD[2], pD[2];
      if constexpr (lev == 0) {
        double uD, dD, t;
        lane_updown<LP, 1>(D[1], uD, t);
        lane_updown<LP, 1>(D[0], t, dD);
        mD[0] = pw >= 1 ? uD : 0.0;
        pD[0] = D[1];
        mD[1] = D[0];
        pD[1] = pw + 1 < LP ? dD : 0.0;
      } else {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          double uD, dD;
          lane_updown<LP, h>(D[j], uD, dD);
          mD[j] = pw >= h ? uD : 0.0;
          pD[j] = pw + h < LP ? dD : 0.0;
        }
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) D[j] = D[j] - k1[lev][j] * mD[j] - k2[lev][j] * pD[j];
    };
    level(std::integral_constant<int, 0>{});
    level(std::integral_constant<int, 1>{});
    level(std::integral_constant<int, 2>{});
    level(std::integral_constant<int, 3>{});
    level(std::integral_constant<int, 4>{});
    if constexpr (LEV > 5) level(std::integral_constant<int, 5>{});
    if constexpr (LEV > 6) level(std::integral_constant<int, 6>{});
    const double xf = D[0] * invB[0], xl = D[1] * invB[1];
#pragma unroll
    for (int k = 0; k < M; ++k) xs[k] = (k == 0) ? xf : (k == M - 1) ? xl : dh[k] - ah[k] * xf - ch[k] * xl;
  };

  // ---- SPIKE set-up (several parts per line): spikes of the lane's rows,
  // the parts' couplings and spike ends in LDS, the inverse of the reduced
  // system formed by the line's first lane (its rows for x_last of the
  // previous part and x_first of the next one are read from LDS per bin)
  double vS[M], wS[M];
  double alpha_w = 0.0, gamma_w = 0.0;
  double *s_cpl = s_tile + T * PITCH + (ADJ ? (size_t)2 * M * NT : 0);  // [T][WV][6]
  double *s_Ri = s_cpl + T * WV * 6;                                     // [T][NR][NR]
  double *s_Rm = s_Ri + T * NR * NR;                                     // [T][NR][NR]
  double *s_y = s_Rm + T * NR * NR;                                      // [T][NR]
  if constexpr (WV > 1) {
    double e1[M], e2[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      e1[k] = (k == 0 && pw == 0) ? 1.0 : 0.0;
      e2[k] = (k == M - 1 && pw == LP - 1) ? 1.0 : 0.0;
    }
    local(e1, vS);
    local(e2, wS);
    alpha_w = lane_value(cpl_a, 0);
    gamma_w = lane_value(cpl_c, LP - 1);
    double *cp = s_cpl + (ln * WV + wv) * 6;
    if (pw == 0) { cp[0] = alpha_w; cp[1] = vS[0]; cp[2] = wS[0]; }
    if (pw == LP - 1) { cp[3] = gamma_w; cp[4] = vS[M - 1]; cp[5] = wS[M - 1]; }
    __syncthreads();
    // R z = y with z = (x_first(0), x_last(0), x_first(1), ...): row 2u
    // (first row of part u) and 2u+1 (its last row) couple x_last(u-1) =
    // z[2u-1] through alpha_u * spike v and x_first(u+1) = z[2u+2] through
    // gamma_u * spike w.  R^{-1} by Gauss-Jordan (R is I plus couplings
    // below one in size for these dominant matrices), one thread of the
    // line per entry of R and of R^{-1}, two barriers per column
    double *Ri = s_Ri + ln * NR * NR, *Rm = s_Rm + ln * NR * NR;
    const bool gj = p < NR * NR;
    const int r_ = p / NR, c_ = p % NR;
    double rm = 0.0, ri = 0.0;
    if (gj) {
      rm = ri = (r_ == c_) ? 1.0 : 0.0;
      const int u = r_ / 2;
      const double *cu = s_cpl + (ln * WV + u) * 6;
      const bool first = (r_ % 2) == 0;
      if (u > 0 && c_ == 2 * u - 1) rm = cu[0] * (first ? cu[1] : cu[4]);
      if (u + 1 < WV && c_ == 2 * u + 2) rm = cu[3] * (first ? cu[2] : cu[5]);
      Rm[p] = rm;
      Ri[p] = ri;
    }
    __syncthreads();
#pragma unroll 1
    for (int c0 = 0; c0 < NR; ++c0) {
      double piv = 1.0, f = 0.0, pm = 0.0, pi = 0.0;
      if (gj) {
        piv = Rm[c0 * NR + c0];
        f = Rm[r_ * NR + c0];
        pm = Rm[c0 * NR + c_];
        pi = Ri[c0 * NR + c_];
      }
      __syncthreads();
      if (gj) {
        const double nm = pm / piv, ni = pi / piv;
        if (r_ == c0) {
          rm = nm;
          ri = ni;
        } else {
          rm = rm - f * nm;
          ri = ri - f * ni;
        }
        Rm[p] = rm;
        Ri[p] = ri;
      }
      __syncthreads();
    }
  }

  auto solve = [&]() {
    double q[M + 2];
#pragma unroll
    for (int k = 0; k < M + 2; ++k) {
      const int i = p * M + k - 1;
      q[k] = trow[min(max(i, 0), n - 1)];
    }
    double d[M], xs[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      // forward: the explicit half-step B q; adjoint: K^{-1} lambda
      d[k] = ADJ ? q[k + 1] * s_kap[(M + k) * NT + tid]
                 : (omt * rL[k]) * q[k] + (1.0 - omt * (rL[k] + rR[k])) * q[k + 1] + (omt * rR[k]) * q[k + 2];
    }
    local(d, xs);
    double xlp = 0.0, xfn = 0.0;  // x_last of the previous part, x_first of the next
    if constexpr (WV > 1) {
      // every part's y end values, then the coupled correction
      double *yv = s_y + ln * NR;
      if (pw == 0) yv[2 * wv] = xs[0];
      if (pw == LP - 1) yv[2 * wv + 1] = xs[M - 1];
      __syncthreads();
      // rows 2 wv - 1 and 2 wv + 2 of R^{-1} (zero rows at the line's ends)
      const double *Ri = s_Ri + ln * NR * NR;
      const double *rL_ = Ri + (wv > 0 ? 2 * wv - 1 : 0) * NR, *rF_ = Ri + (wv + 1 < WV ? 2 * wv + 2 : 0) * NR;
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const double y = yv[c];
        xlp = xlp + rL_[c] * y;
        xfn = xfn + rF_[c] * y;
      }
      if (wv == 0) xlp = 0.0;
      if (wv + 1 == WV) xfn = 0.0;
#pragma unroll
      for (int k = 0; k < M; ++k) xs[k] = xs[k] - alpha_w * xlp * vS[k] - gamma_w * xfn * wS[k];
    }
    if constexpr (ADJ) {
      // lambda' = K B u with u = A^{-1} K^{-1} lambda: B's stencil needs the
      // last row of the segment before and the first row of the one after;
      // the line ends take u_0 = u_1 and u_{n+1} = u_n as B does
      double xprev, xnext, t;
      lane_updown<LP, 1>(xs[M - 1], xprev, t);
      lane_updown<LP, 1>(xs[0], t, xnext);
      // several parts: the first and last lanes of a part take the
      // neighbouring parts' end values of the solution
      if (WV > 1 && pw == 0) xprev = xlp;
      if (WV > 1 && pw == LP - 1) xnext = xfn;
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const int i = p * M + k;
        double xm = (k == 0) ? xprev : xs[k > 0 ? k - 1 : 0];
        double xp = (k == M - 1) ? xnext : xs[k < M - 1 ? k + 1 : 0];
        if (i == 0) xm = xs[k];
        if (i == n - 1) xp = xs[k];
        const double v = (omt * rL[k]) * xm + (1.0 - omt * (rL[k] + rR[k])) * xs[k] + (omt * rR[k]) * xp;
        if (i < n) trow[i] = s_kap[k * NT + tid] * v;
      }
    } else {
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const int i = p * M + k;
        if (i < n) trow[i] = xs[k];
      }
    }
  };

  // ---- stores.  Cells 1..N of the tile's lines hold, after the sweep, the
  // solution on active cells of active lines of aloft bins and zero
  // everywhere else (as the reference leaves them).  The zeros outside the
  // active cells, and whole bins that are not aloft, are written first;
  // the per-bin store of the active cells then has the fixed shape of the
  // tile load (PF stores per lane, masked-off ones to the junk buffer), so
  // that the compiler can count the memory operations in flight exactly
  // and wait for one bin's loads without waiting for the next one's.
  double *junk = A.junk + (tid + NT * (vb % (kJunk / NT)));
  auto zero_cells = [&](int w, int r_lo, int r_hi) {
    const int len = r_hi - r_lo + 1;
    if (len <= 0) return;
    for (int f = tid; f < T * len; f += NT) {
      const int l_ = CONTIG ? f / len : f % T;
      const int r_ = (CONTIG ? f % len : f / T) + r_lo;
      const int al_ = a_first + l_;
      if (owned(al_)) A.c0[Q.st0 + Q.sta * al_ + Q.stb * b + Q.str * (long long)r_ + ncell * (long long)w] = 0.0;
    }
  };
  for (int w = u0; w < u1; ++w) {
    if (!A.aloft.on(w)) {
      zero_cells(w, 1, N);
    } else {
      zero_cells(w, 1, rmin - 1);
      zero_cells(w, rmax + 1, N);
    }
  }
  auto store = [&](int w) {
    auto one = [&](int f, bool in) {
      const int l_ = elem_l(f), i_ = elem_i(f);
      const int al_ = a_first + l_;
      const bool ok = in && owned(al_);
      const bool lon = al_ >= g.a0 && al_ <= g.a1 && b >= g.b0 && b <= g.b1;
      const double v = (ok && lon) ? s_tile[l_ * PITCH + i_] : 0.0;
      const long long o = Q.st0 + Q.sta * al_ + Q.stb * b + Q.str * (long long)(rmin + i_) + ncell * (long long)w;
      *(ok ? A.c0 + o : junk) = v;
    };
    if constexpr (PD == 2) {
#pragma unroll
      for (int m = 0; m < PF; ++m) one(tid + NT * m, tid + NT * m < nel);
    } else {
      // one bin in flight: a plain loop (fewer registers)
      for (int f = tid; f < nel; f += NT) {
        const int l_ = elem_l(f), i_ = elem_i(f);
        const int al_ = a_first + l_;
        if (!owned(al_)) continue;
        const bool lon = al_ >= g.a0 && al_ <= g.a1 && b >= g.b0 && b <= g.b1;
        A.c0[Q.st0 + Q.sta * al_ + Q.stb * b + Q.str * (long long)(rmin + i_) + ncell * (long long)w] =
            lon ? s_tile[l_ * PITCH + i_] : 0.0;
      }
    }
  };

  if constexpr (PD == 1) {
    int u = next_on(u0);
    if (u < u1) issue(std::integral_constant<int, 0>{}, u);
    while (u < u1) {
      __syncthreads();  // the previous bin's tile has been stored
      commit(std::integral_constant<int, 0>{});
      __syncthreads();
      const int un = next_on(u + 1);
      if (un < u1) issue(std::integral_constant<int, 0>{}, un);
      solve();
      __syncthreads();
      store(u);
      u = un;
    }
  } else {
    // the aloft bins two at a time, slot 0 then slot 1, so that each slot
    // stays in its registers; a load past the last aloft bin re-reads the
    // last one (its tile was just read: a cache hit) instead of being
    // skipped, which would make the slot's contents depend on the path
    int nact = 0, last = u0;
    for (int w = u0; w < u1; ++w)
      if (A.aloft.on(w)) { ++nact; last = w; }
    auto nxt = [&](int u) {
      const int v = next_on(u);
      return v < u1 ? v : last;
    };
    int ua = next_on(u0);
    int ub = nxt(ua + 1);
    if (nact > 0) {
      issue(std::integral_constant<int, 0>{}, ua);
      issue(std::integral_constant<int, 1>{}, ub);
    }
    for (int k = 0; k < nact; k += 2) {
      __syncthreads();  // the previous bin's tile has been stored
      commit(std::integral_constant<int, 0>{});
      __syncthreads();
      const int uc = nxt(ub + 1);
      issue(std::integral_constant<int, 0>{}, uc);
      solve();
      __syncthreads();
      store(ua);
      ua = uc;
      if (k + 1 >= nact) break;
      __syncthreads();
      commit(std::integral_constant<int, 1>{});
      __syncthreads();
      const int ud = nxt(uc + 1);
      issue(std::integral_constant<int, 1>{}, ud);
      solve();
      __syncthreads();
      store(ub);
      ub = ud;
    }
  }
}

template <int LP, int M, int T, bool CONTIG, int PD, bool ADJ, int WV>
__global__ __launch_bounds__(T *LP *WV) void k_cn_part(CnPartArgs P) {
  cn_part_body<LP, M, T, CONTIG, PD, ADJ, WV>(P.A, P.Q, blockIdx.x, gridDim.x);
}
// the same solves over the members of an ensemble (one launch)
template <int LP, int M, int T, bool CONTIG, int PD, bool ADJ, int WV>
__global__ __launch_bounds__(T *LP *WV) void k_cn_part_batch(
    const ASH3D_CONSTANT CnPartArgs *__restrict__ args, const ASH3D_CONSTANT int *__restrict__ member,
    const ASH3D_CONSTANT long long *__restrict__ first) {
  const int m = member[blockIdx.x];
  const CnPartArgs &P = *(const CnPartArgs *)(args + m);
  cn_part_body<LP, M, T, CONTIG, PD, ADJ, WV>(P.A, P.Q, (long long)blockIdx.x - first[m],
                                               first[m + 1] - first[m]);
}

template <int LP, int M, int T, bool CONTIG, bool ADJ, int PD = (M <= 6 ? 2 : 1), int WV = 1>
hipError_t launch_part(hipStream_t s, const DiffArgs &a, const PartArgs &q, LaunchSpec *sp) {
  const long long tiles = (long long)((a.g.Na + T - 1) / T) * a.g.Nb;
  const int G = (a.d.ns + q.bins_wg - 1) / q.bins_wg;
  // the tile, for the adjoint each lane's row kappas and inverses, for
  // several parts per line the couplings, the reduced inverse and the end values
  const size_t lds = sizeof(double) * ((size_t)T * (LP * WV * M + 1) + (ADJ ? (size_t)2 * M * T * LP * WV : 0) +
                                       (size_t)T * WV * 6 + (size_t)T * 8 * WV * WV + (size_t)T * 2 * WV);
  // (beyond 64 KiB of dynamic LDS for the adjoint's 512-lane tiles)
  return issue(s, sp, (const void *)k_cn_part<LP, M, T, CONTIG, PD, ADJ, WV>,
               (const void *)k_cn_part_batch<LP, M, T, CONTIG, PD, ADJ, WV>, dim3((unsigned)padded_tiles(tiles * G)),
               dim3(T * LP * WV), lds, CnPartArgs{a, q});
}

// bins in flight of the multi-wave line solves
#ifndef ASH3D_PDL
#define ASH3D_PDL 2
#endif
constexpr int PDL = ASH3D_PDL;

template <bool ADJ>
hipError_t launch_cn_part_shapes(hipStream_t s, const DiffArgs &a, const PartArgs &q, bool contig, int n, int shape,
                                 LaunchSpec *sp) {
  // lanes per line LP, rows per lane M (LP*M >= n), lines per workgroup T;
  // shape 1: contiguous lines two per workgroup (fewer waves held at each
  // barrier), strided lines in wider tiles (longer coalesced row segments)
  {
    // lines over several waves (SPIKE coupling of the parts; the adjoint
    // too); 4 rows per lane, 256 per wave; contiguous lines one per
    // workgroup, strided lines in tiles of 2-4 along the contiguous index
    if (n > 512) {
      if (contig) {
        if (n <= 768) return launch_part<64, 4, 1, true, ADJ, PDL, 3>(s, a, q, sp);
        if (n <= 1024) return launch_part<64, 4, 1, true, ADJ, PDL, 4>(s, a, q, sp);
        if (n <= 1536) return launch_part<64, 4, 1, true, ADJ, PDL, 6>(s, a, q, sp);
        return launch_part<64, 4, 1, true, ADJ, PDL, 8>(s, a, q, sp);
      }
      if (n <= 768) return launch_part<64, 4, 2, false, ADJ, PDL, 3>(s, a, q, sp);
      if (n <= 1024) return launch_part<64, 4, 2, false, ADJ, PDL, 4>(s, a, q, sp);
      if (n <= 1536) return launch_part<64, 4, 1, false, ADJ, PDL, 6>(s, a, q, sp);
      return launch_part<64, 4, 1, false, ADJ, PDL, 8>(s, a, q, sp);
    }
  }
  if (contig) {
    if (shape == 1) {
      if (n <= 256) return launch_part<64, 4, 2, true, ADJ>(s, a, q, sp);
      if (n <= 384) return launch_part<64, 6, 2, true, ADJ>(s, a, q, sp);
      return launch_part<64, 8, 2, true, ADJ>(s, a, q, sp);
    }
    if (n <= 256) return launch_part<64, 4, 4, true, ADJ>(s, a, q, sp);
    if (n <= 384) return launch_part<64, 6, 4, true, ADJ>(s, a, q, sp);
    return launch_part<64, 8, 4, true, ADJ>(s, a, q, sp);
  }
  // lines of 65..96 cells in 16 lanes of 5 or 6 rows (less padding than
  // 32 lanes of 4: 80-cell columns 11.5 -> 6.5 ms a sweep on the
  // 1440x720x80x20 global grid)
  if (shape == 1) {
    if (n <= 64) return launch_part<16, 4, 32, false, ADJ>(s, a, q, sp);
    if (n <= 80) return launch_part<16, 5, 16, false, ADJ>(s, a, q, sp);
    if (n <= 96) return launch_part<16, 6, 16, false, ADJ>(s, a, q, sp);
    if (n <= 128) return launch_part<32, 4, 16, false, ADJ>(s, a, q, sp);
    if (n <= 256) return launch_part<32, 8, 16, false, ADJ>(s, a, q, sp);
    return launch_part<64, 8, 4, false, ADJ>(s, a, q, sp);
  }
  if (n <= 64) return launch_part<16, 4, 16, false, ADJ>(s, a, q, sp);
  if (n <= 80) return launch_part<16, 5, 32, false, ADJ>(s, a, q, sp);
  if (n <= 96) return launch_part<16, 6, 32, false, ADJ>(s, a, q, sp);
  if (n <= 128) return launch_part<32, 4, 16, false, ADJ>(s, a, q, sp);
  if (n <= 256) return launch_part<64, 4, 8, false, ADJ>(s, a, q, sp);
  return launch_part<64, 8, 4, false, ADJ>(s, a, q, sp);
}

}  // namespace

static hipError_t diff_cn_part_any(hipStream_t s, const DiffArgs &a, const PartLayout &lay, int bins_wg, int shape,
                                   bool adjoint, LaunchSpec *sp) {
  const int n = a.g.rmax - a.g.rmin + 1;
  PartArgs q;
  q.st0 = lay.st0; q.sta = lay.sta; q.stb = lay.stb; q.str = lay.str;
  q.lr0 = lay.lr0; q.lra = lay.lra; q.lrb = lay.lrb; q.lrr = lay.lrr; q.lr_ncell = lay.lr_ncell;
  q.kap = lay.kap;
  q.bins_wg = (bins_wg > 0 && bins_wg < a.d.ns) ? bins_wg : a.d.ns;
  if (n < 3 || n > 2048) return hipErrorInvalidValue;
  if (adjoint && !q.kap) return hipErrorInvalidValue;
  return adjoint ? launch_cn_part_shapes<true>(s, a, q, lay.str == 1, n, shape, sp)
                 : launch_cn_part_shapes<false>(s, a, q, lay.str == 1, n, shape, sp);
}

hipError_t launch_diff_cn_part(hipStream_t s, const DiffArgs &a, const PartLayout &lay, int bins_wg, int shape,
                               bool adjoint) {
  return diff_cn_part_any(s, a, lay, bins_wg, shape, adjoint, nullptr);
}

hipError_t spec_diff_cn_part(const DiffArgs &a, const PartLayout &lay, int bins_wg, int shape, LaunchSpec *sp) {
  return diff_cn_part_any(nullptr, a, lay, bins_wg, shape, false, sp);
}

}  // namespace ash3d
