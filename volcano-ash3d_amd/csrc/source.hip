// Source term of the standard single-column sources (host code: a few
// hundred operations per eruptive pulse or time step, next to the device
// column insertion ash3d_tr_source_column_add):
//   Calc_Normalized_SourceCol   src/Source.f90:287-490
//   EruptivePulse_MassFluxRate  src/Source.f90:508-589
//   CheckEruptivePulses         src/Source.f90:613-685
//   TephraSourceNodes           src/Source.f90:708-783
// Each is evaluated expression by expression in the reference's order with
// the C library's exp, so the results are those of the Fortran routines.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "ash3d_tr.h"
#include "common.hpp"

namespace {

constexpr double KM3_2_M3 = 1.0e9;   // src/Ash3d_VariableModules.f90:284
constexpr double EPS_SMALL = 1.0e-6; // :271

int src_fail(const char *msg) { return ash3d::report_error(ASH3D_TR_EINVAL, msg); }

bool suzuki_like(int t) {
  return t == ASH3D_SRC_SUZUKI || t == ASH3D_SRC_UMBRELLA || t == ASH3D_SRC_UMBRELLA_AIR;
}
bool standard(int t) { return suzuki_like(t) || t == ASH3D_SRC_LINE || t == ASH3D_SRC_POINT; }

// z (km) to the s coordinate of ZScaling_ID (:323-338, :437-446)
double to_s(const ash3d_tr_source_desc *s, double z) {
  if (s->zscaling_id == 1) return z - s->zsurf_vent;
  if (s->zscaling_id == 2) return s->ztop * (z - s->zsurf_vent) / (s->ztop - s->zsurf_vent);
  return z;
}

}  // namespace

extern "C" {

int ash3d_tr_source_column(const ash3d_tr_source_desc *s, int32_t nz, double *norm, double *mass_flux_rate,
                           double *end_time) {
  if (!s || !norm || nz < 1 || s->neruptions < 1 || !s->start_time || !s->duration || !s->volume ||
      !s->plume_height || !s->s_lb || !s->s_cc || !s->ds)
    return src_fail("ash3d_tr_source_column: null or empty input");
  if (s->source_type == ASH3D_SRC_PROFILE && (!s->prof_dz || !s->prof_nz || !s->prof_volume || s->prof_ld < 1))
    return src_fail("ash3d_tr_source_column: profile source without its profile");
  const int ne = s->neruptions, t = s->source_type;
  // NormSourceColumn(i,k) at norm[i-1 + ne*(k-1)]
  for (long long c = 0; c < (long long)ne * nz; ++c) norm[c] = 0.0;
  // Convert z_volcano to s-coordinate (:322-338)
  double z_volcano = s->z_volcano;
  if (s->zscaling_id == 1 || s->zscaling_id == 2) z_volcano = fmax(z_volcano, s->zsurf_vent);
  const double s_volcano = to_s(s, z_volcano);
  for (int i = 1; i <= ne; ++i) {
    const double sPH = to_s(s, s->plume_height[i - 1]);
    // the cell containing the vent, the cell containing the plume top (:340-352)
    int kground = 1, kPlumeTop = 0;
    for (int k = 1; k <= nz; ++k) {
      if (s_volcano >= s->s_lb[k - 1] && s_volcano < s->s_lb[k]) kground = k;
      if (sPH > s->s_lb[k - 1] && sPH <= s->s_lb[k]) kPlumeTop = k;
    }
    const double sground = s->s_lb[kground - 1];
    const double sPHag = sPH - sground;
    double Suzuki_k = 0.0;
    if (suzuki_like(t)) Suzuki_k = s->suzuki_a / (1.0 - (s->suzuki_a + 1.0) * exp(-s->suzuki_a));
    for (int k = kground; k <= kPlumeTop; ++k) {
      const double s_cell_top = fmin(s->s_cc[k - 1] + 0.5 * s->ds[k - 1], sPH);
      const double s_cell_bot = s->s_cc[k - 1] - 0.5 * s->ds[k - 1];
      double &N = norm[(i - 1) + (long long)ne * (k - 1)];
      if (suzuki_like(t)) {
        // the Suzuki distribution integrated over the cell (:375-392)
        const double sclfac1 = (s_cell_top - sground) / sPHag;
        const double sclfac2 = (s_cell_bot - sground) / sPHag;
        const double cmpsclfac1 = 1.0 - sclfac1;
        const double cmpsclfac2 = 1.0 - sclfac2;
        const double fac1 = 1.0 / s->suzuki_a + cmpsclfac1;
        const double fac2 = 1.0 / s->suzuki_a + cmpsclfac2;
        N = Suzuki_k * (fac1 * exp(-1.0 * s->suzuki_a * cmpsclfac1) - fac2 * exp(-1.0 * s->suzuki_a * cmpsclfac2));
      } else if (t == ASH3D_SRC_LINE) {
        N = (s_cell_top - s_cell_bot) / sPHag;  // (:394-397)
      } else if (t == ASH3D_SRC_POINT) {
        N = (s_cell_top >= sPH && s_cell_bot < sPH) ? 1.0 : 0.0;  // (:398-406)
      } else if (t == ASH3D_SRC_PROFILE) {
        // overlap of each profile step with the cell (:407-474)
        for (int kk = 1; kk <= s->prof_nz[i - 1]; ++kk) {
          const double zbot_prof = s->prof_dz[i - 1] * (double)(kk - 1);
          const double ztop_prof = s->prof_dz[i - 1] * (double)kk;
          const double sbot_prof = to_s(s, zbot_prof), stop_prof = to_s(s, ztop_prof);
          double frac;
          if (stop_prof < sground) {
            continue;  // below the vent: disregarded
          } else if (sbot_prof <= s_cell_bot && stop_prof >= s_cell_bot && stop_prof <= s_cell_top) {
            frac = (stop_prof - s_cell_bot) / s->ds[k - 1];
          } else if (sbot_prof >= s_cell_bot && stop_prof < s_cell_top) {
            frac = (stop_prof - sbot_prof) / s->ds[k - 1];
          } else if (sbot_prof >= s_cell_bot && sbot_prof <= s_cell_top && stop_prof >= s_cell_top) {
            frac = (s_cell_top - sbot_prof) / s->ds[k - 1];
          } else if (sbot_prof < s_cell_bot && stop_prof > s_cell_top) {
            frac = 1.0;
          } else {
            frac = 0.0;
          }
          N = N + s->prof_volume[(i - 1) + (long long)s->prof_ld * (kk - 1)] * frac;
        }
      } else {
        N = 0.0;  // not a standard source (:475-480)
      }
    }
    double tot = 0.0;  // sum(NormSourceColumn(i,:)) (:485-486)
    for (int k = 1; k <= nz; ++k) tot += norm[(i - 1) + (long long)ne * (k - 1)];
    for (int k = 1; k <= nz; ++k) {
      double &N = norm[(i - 1) + (long long)ne * (k - 1)];
      N = N / tot;
    }
  }
  // EruptivePulse_MassFluxRate (:530-563)
  for (int i = 1; i <= ne; ++i) {
    double mfr = 0.0;
    if (standard(t)) {
      mfr = s->magma_density * s->volume[i - 1] * KM3_2_M3 / s->duration[i - 1];
    } else if (t == ASH3D_SRC_PROFILE) {
      for (int kk = 1; kk <= s->prof_nz[i - 1]; ++kk)
        mfr += s->magma_density * s->prof_volume[(i - 1) + (long long)s->prof_ld * (kk - 1)] * KM3_2_M3 /
               s->duration[i - 1];
    }
    if (mass_flux_rate) mass_flux_rate[i - 1] = mfr;
    if (end_time) end_time[i - 1] = s->start_time[i - 1] + s->duration[i - 1];
  }
  return 0;
}

int ash3d_tr_eruptive_pulses(const ash3d_tr_source_desc *s, const double *end_time, double time, double dt,
                             int32_t *ieruption, int32_t *jeruption, double *frac, int32_t *in_dt) {
  if (!s || !end_time || !ieruption || !jeruption || !frac || !in_dt)
    return src_fail("ash3d_tr_eruptive_pulses: null input");
  const int ne = s->neruptions, t = s->source_type;
  const double tstart = time, tend = time + dt;
  *in_dt = 0;
  for (int i = 0; i < ne; ++i) frac[i] = 0.0;
  if (!(standard(t) || t == ASH3D_SRC_PROFILE)) return 0;  // (:674-680)
  for (int i = *ieruption; i <= ne; ++i) {
    bool contributes = false;
    if (tstart >= s->start_time[i - 1] && tstart < end_time[i - 1]) {
      contributes = true;  // touches the start of dt (:641-645)
      *jeruption = i;
    } else if (tend > s->start_time[i - 1] && tend <= end_time[i - 1]) {
      // touches the end of dt: the reference leaves the loop here, before
      // this pulse's fraction is recorded (:646-653)
      *jeruption = i;
      break;
    } else if (tstart < s->start_time[i - 1] && tend > end_time[i - 1]) {
      contributes = true;  // wholly within dt (:654-664)
      *jeruption = i;
    }
    if (contributes) {
      *in_dt = 1;
      frac[i - 1] = (fmin(tend, end_time[i - 1]) - fmax(tstart, s->start_time[i - 1])) / dt;
    }
  }
  return 0;
}

int ash3d_tr_source_nodes(const ash3d_tr_source_desc *s, int32_t nz, int32_t ns, int32_t n_gs_max,
                          const double *norm, const double *mass_flux_rate, const double *frac,
                          const double *kappa_col, int32_t *ieruption, int32_t jeruption, double *flux) {
  if (!s || !norm || !mass_flux_rate || !frac || !kappa_col || !ieruption || !flux || !s->bin_mass)
    return src_fail("ash3d_tr_source_nodes: null input");
  if (n_gs_max < 0 || n_gs_max > ns) return src_fail("ash3d_tr_source_nodes: n_gs_max outside 0..ns");
  const int ne = s->neruptions, ie = *ieruption;
  // SourceNodeFlux(1:nz+1, 1:ns) at flux[k-1 + (nz+1)*(n-1)]
  for (long long c = 0; c < (long long)(nz + 1) * ns; ++c) flux[c] = 0.0;
  double tfr_stack[512];
  double *TFR = nz <= 512 ? tfr_stack : new double[nz];
  for (int k = 0; k < nz; ++k) TFR[k] = 0.0;
  double mfr_now = 0.0;
  for (int i = ie; i <= jeruption; ++i) {  // (:733-740); NormSourceColumn(ieruption,:) for every pulse
    for (int k = 1; k <= nz; ++k)
      TFR[k - 1] = TFR[k - 1] + mass_flux_rate[i - 1] * frac[i - 1] * norm[(ie - 1) + (long long)ne * (k - 1)];
    mfr_now = mfr_now + mass_flux_rate[i - 1] * frac[i - 1];
  }
  double sum = 0.0;
  for (int k = 1; k <= nz; ++k) {  // (:746-757)
    const double kap = kappa_col[k - 1];
    double sk = 0.0;
    for (int n = 1; n <= n_gs_max; ++n) {
      const double f = s->bin_mass[n - 1] * TFR[k - 1] / kap;
      flux[(k - 1) + (long long)(nz + 1) * (n - 1)] = f;
      sk += f * kap;
    }
    sum = sum + sk / mfr_now;
  }
  if (TFR != tfr_stack) delete[] TFR;
  if (fabs(sum - 1.0) > EPS_SMALL) {  // (:760-770)
    char msg[160];
    snprintf(msg, sizeof msg,
             "Source Node Flux does not agree with calculations: (Sum(SourceNodeFlux)/MassFluxRate)-1 = %.5e",
             sum - 1.0);
    return src_fail(msg);
  }
  *ieruption = jeruption;  // (:774)
  return 0;
}

}  // extern "C"
