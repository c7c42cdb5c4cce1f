// Lane moves inside groups of LP consecutive lanes of a wave64 (LP = 16,
// 32 or 64): DPP row shifts within 16-lane rows, the whole-wave DPP shifts
// by one, the gfx950 row swaps v_permlane16_swap (odd rows of the first
// operand <-> even rows of the second) and v_permlane32_swap (rows 2,3 of
// the first <-> rows 0,1 of the second) for shifts by whole rows, and the
// LDS crossbar (ds_bpermute) for the rest.
#pragma once

#include <hip/hip_runtime.h>

namespace ash3d {

// dpp_ctrl: 0x100+h row_shl:h, 0x110+h row_shr:h, 0x120+h row_ror:h,
// 0x130 wave_shl:1, 0x134 wave_rol:1, 0x138 wave_shr:1, 0x13C wave_ror:1;
// source lanes outside the row (or wave) read 0
template <int CTRL>
__device__ inline int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}

struct UpDown {
  int up, dn;
};

// up = v of lane l-H, dn = v of lane l+H, for the lanes where that lane is
// in the same group of LP; other lanes get unspecified values.
template <int LP, int H>
__device__ inline UpDown lane_updown(int v) {
  static_assert(LP == 16 || LP == 32 || LP == 64, "groups of 16, 32 or 64 lanes");
  static_assert(H >= 1 && H < LP && (H & (H - 1)) == 0, "a power of two below LP");
  if constexpr (LP == 16) {
    return {dpp_i32<0x110 + H>(v), dpp_i32<0x100 + H>(v)};
  } else if constexpr (H == 1) {
    return {dpp_i32<0x138>(v), dpp_i32<0x130>(v)};
  } else if constexpr (H == 32) {
    const auto t = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return {(int)t[0], (int)t[1]};  // rows 2,3 <- rows 0,1 ; rows 0,1 <- rows 2,3
  } else if constexpr (H == 16) {
    // the same position in the previous and the next row
    const auto s = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    int prev = (int)s[0];  // odd rows: the row before
    int next = (int)s[1];  // even rows: the row after
    if constexpr (LP == 64) {
      // row 2 <- row 1 and row 1 <- row 2 from the two swaps above
      const int row = (int)__lane_id() >> 4;
      const auto t = __builtin_amdgcn_permlane32_swap((unsigned)prev, (unsigned)next, false, false);
      prev = (row == 2) ? (int)t[0] : prev;
      next = (row == 1) ? (int)t[1] : next;
    }
    return {prev, next};
  } else {
    // shifts by 2..8 across row boundaries: the LDS crossbar (measured
    // faster than composing row shifts, rotations and swaps with selects)
    return {__shfl_up(v, H, LP), __shfl_down(v, H, LP)};
  }
}

template <int LP, int H>
__device__ inline void lane_updown(double v, double &up, double &dn) {
  const UpDown lo = lane_updown<LP, H>(__double2loint(v));
  const UpDown hi = lane_updown<LP, H>(__double2hiint(v));
  up = __hiloint2double(hi.up, lo.up);
  dn = __hiloint2double(hi.dn, lo.dn);
}


// ---- rows held by one wavefront ---------------------------------------
// A row of up to 128*NC cells as NC chunks: lane l of chunk m holds the
// cell pair p = 128 m + 2 l + e (e = 0, 1) in v[m][e].  The neighbours of
// a cell come from the other element of the pair or from the next or
// previous lane by a whole-wave DPP rotation; lanes 0 and 63 take the
// rotated value of the neighbouring chunk, and positions before the row or
// after the last chunk read zero.  Every lane executes these moves (they
// must not sit under divergent control flow).
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  return __hiloint2double(dpp_i32<CTRL>(__double2hiint(v)), dpp_i32<CTRL>(__double2loint(v)));
}
constexpr int kWaveRor1 = 0x13C;  // lane l <- lane l-1, lane 0 <- lane 63
constexpr int kWaveRol1 = 0x134;  // lane l <- lane l+1, lane 63 <- lane 0

// o = v at p-1
template <int NC>
__device__ inline void row_prev(const double (&v)[NC][2], double (&o)[NC][2], bool lane0) {
  double r[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) r[m] = dpp_f64<kWaveRor1>(v[m][1]);
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    o[m][0] = lane0 ? (m > 0 ? r[m > 0 ? m - 1 : 0] : 0.0) : r[m];
    o[m][1] = v[m][0];
  }
}
// o = v at p+1
template <int NC>
__device__ inline void row_next(const double (&v)[NC][2], double (&o)[NC][2], bool lane63) {
  double r[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) r[m] = dpp_f64<kWaveRol1>(v[m][0]);
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    o[m][0] = v[m][1];
    o[m][1] = lane63 ? (m + 1 < NC ? r[m + 1 < NC ? m + 1 : m] : 0.0) : r[m];
  }
}
// single cells per lane (p = 64 m + l): o = v at p-1 / p+1, the same
// rotations with the neighbouring chunk's value at lanes 0 and 63
template <int NC>
__device__ inline void row_prev(const double (&v)[NC][1], double (&o)[NC][1], bool lane0) {
  double r[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) r[m] = dpp_f64<kWaveRor1>(v[m][0]);
#pragma unroll
  for (int m = 0; m < NC; ++m) o[m][0] = lane0 ? (m > 0 ? r[m > 0 ? m - 1 : 0] : 0.0) : r[m];
}
template <int NC>
__device__ inline void row_next(const double (&v)[NC][1], double (&o)[NC][1], bool lane63) {
  double r[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) r[m] = dpp_f64<kWaveRol1>(v[m][0]);
#pragma unroll
  for (int m = 0; m < NC; ++m) o[m][0] = lane63 ? (m + 1 < NC ? r[m + 1 < NC ? m + 1 : m] : 0.0) : r[m];
}
// o = v at p-2 (the previous lane's same element)
template <int NC>
__device__ inline void row_prev2(const double (&v)[NC][2], double (&o)[NC][2], bool lane0) {
  double r[NC][2];
#pragma unroll
  for (int m = 0; m < NC; ++m)
#pragma unroll
    for (int e = 0; e < 2; ++e) r[m][e] = dpp_f64<kWaveRor1>(v[m][e]);
#pragma unroll
  for (int m = 0; m < NC; ++m)
#pragma unroll
    for (int e = 0; e < 2; ++e) o[m][e] = lane0 ? (m > 0 ? r[m > 0 ? m - 1 : 0][e] : 0.0) : r[m][e];
}
// o = v at p+2 (the next lane's same element)
template <int NC>
__device__ inline void row_next2(const double (&v)[NC][2], double (&o)[NC][2], bool lane63) {
  double r[NC][2];
#pragma unroll
  for (int m = 0; m < NC; ++m)
#pragma unroll
    for (int e = 0; e < 2; ++e) r[m][e] = dpp_f64<kWaveRol1>(v[m][e]);
#pragma unroll
  for (int m = 0; m < NC; ++m)
#pragma unroll
    for (int e = 0; e < 2; ++e)
      o[m][e] = lane63 ? (m + 1 < NC ? r[m + 1 < NC ? m + 1 : m][e] : 0.0) : r[m][e];
}
// the double v of lane `lane` (uniform) in every lane
__device__ inline double lane_value(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

}  // namespace ash3d
