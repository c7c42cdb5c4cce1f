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
// 0x130 wave_shl:1, 0x138 wave_shr:1; source lanes outside the row read 0
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

}  // namespace ash3d
