// Source term of the standard single-column sources (host code: a few
// hundred operations per eruptive pulse or time step, next to the device
// column insertion ash3d_tr_source_column_add):
//   Calc_Normalized_SourceCol   src/Source.f90:287-490
//   EruptivePulse_MassFluxRate  src/Source.f90:508-589
//   CheckEruptivePulses         src/Source.f90:613-685
//   TephraSourceNodes           src/Source.f90:708-783
//   Source_Umbrella             src/Source_Umbrella.f90:85-560 (stencil,
//                               levels, radial winds, 3x3 patch flux, volume)
// Each is evaluated expression by expression in the reference's order with
// the C library's exp, so the results are those of the Fortran routines.
#include <math.h>

#include <algorithm>
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

// ---- umbrella sources ----------------------------------------------------

int ash3d_tr_umbrella_stencil(const double *kappa_pd, int32_t nx, int32_t ny, int32_t nz, int32_t ivent,
                              int32_t jvent, double *avg_stencil, double *scale_fac) {
  if (!kappa_pd || !avg_stencil || !scale_fac || nx < 1 || ny < 1 || nz < 1)
    return src_fail("ash3d_tr_umbrella_stencil: null or empty input");
  if (ivent < 1 || ivent > nx || jvent < 1 || jvent > ny)
    return src_fail("ash3d_tr_umbrella_stencil: vent outside the grid");
  const long long sy = nx + 4, sz = (long long)(nx + 4) * (ny + 4);
  auto kap = [&](int i, int j, int k) { return kappa_pd[(i + 1) + sy * (j + 1) + sz * (k + 1)]; };
  for (int k = 1; k <= nz; ++k) {  // (:126-137)
    double tot_vol = 0.0;
    for (int i = 1; i <= 3; ++i)
      for (int j = 1; j <= 3; ++j) {
        const double v = kap(ivent - 2 + i, jvent - 2 + j, k);
        avg_stencil[(i - 1) + 3 * (j - 1) + 9 * (k - 1)] = v;
        tot_vol = tot_vol + v;
      }
    scale_fac[k - 1] = kap(ivent, jvent, k) / tot_vol;
    for (int q = 0; q < 9; ++q) avg_stencil[q + 9 * (k - 1)] = avg_stencil[q + 9 * (k - 1)] / tot_vol;
  }
  return 0;
}

int ash3d_tr_umbrella_levels(const double *z_vec_init, int32_t nz, double plume_height, int32_t *ibase,
                             int32_t *itop) {
  if (!z_vec_init || !ibase || !itop || nz < 1) return src_fail("ash3d_tr_umbrella_levels: null or empty input");
  *itop = 0;
  *ibase = 0;
  for (int k = 1; k <= nz; ++k) {  // (:139-150)
    if (z_vec_init[k] >= plume_height && z_vec_init[k - 1] < plume_height) *itop = k;
    if (z_vec_init[k] >= 0.75 * plume_height && z_vec_init[k - 1] < 0.75 * plume_height) *ibase = k;
  }
  return 0;
}

int ash3d_tr_umbrella_winds(const ash3d_tr_umbrella_wind_params *p, int32_t nx, int32_t ny, const double *lon_cc,
                            const double *lat_cc, double *uvx, double *uvy) {
  constexpr double PI = 3.141592653589793;  // src/Ash3d_VariableModules.f90:275-288
  constexpr double DEG2RAD = 1.7453292519943295e-2, DEG2KMLAT = 111.0, DEG2KMLON = 111.321;
  constexpr double KM_2_M = 1.0e3, MPS_2_KMPHR = 3.6, HR_2_S = 3.6e+3;
  if (!p || !lon_cc || !lat_cc || !uvx || !uvy || nx < 1 || ny < 1)
    return src_fail("ash3d_tr_umbrella_winds: null or empty input");
  for (long long c = 0; c < (long long)nx * ny; ++c) uvx[c] = uvy[c] = 0.0;  // (:257-258)
  if (!p->is_latlon)
    return src_fail("ash3d_tr_umbrella_winds: umbrella_winds is not yet set up to handle projected coordinates");
  if (p->vel_model != 1 && p->vel_model != 2)
    return src_fail("ash3d_tr_umbrella_winds: Umb.Vel.Model unknown.  Should be 1 or 2.");
  double mfr = p->mass_flux_rate / HR_2_S;  // kg/s (:270)
  if (p->airborne_one_bin) mfr = 20.0 * mfr;  // (:274-276)
  const double C_Costa = (fabs(p->lat_volcano) < 23.0) ? 0.43e3 : 0.87e3;  // (:279-285)
  // Eq. 2 of Mastin and Van Eaton (2020), m3/s (:288-289)
  const double qnow = C_Costa * sqrt(p->k_entrainment) * pow(mfr, 3.0 / 4.0) / pow(p->n_bv, 5.0 / 4.0);
  const double etime_s = p->first_time ? fmin(p->simtime, p->end_time) * HR_2_S  // (:318-324)
                                       : fmax(p->time, p->dt) * HR_2_S;
  // Eq. 1 and its time derivative (:327-336)
  const double base = 3.0 * p->lambda * p->n_bv * qnow / (2.0 * PI);
  const double cloud_radius = pow(base, 1.0 / 3.0) * pow(etime_s, 2.0 / 3.0) / KM_2_M;
  const double edge_speed = (2.0 / 3.0) * pow(base, 1.0 / 3.0) * pow(etime_s, -1.0 / 3.0);
  const double node_w = (3.0 / 2.0) * p->de_km, node_h = (3.0 / 2.0) * p->dn_km;  // (:122-123)
  if (cloud_radius <= fmax(node_w, node_h)) return 0;  // (:338-340)
  const int ew_nodes = (int)(cloud_radius / p->de_km) + 1, ns_nodes = (int)(cloud_radius / p->dn_km) + 1;
  const int west = std::max(1, p->ivent - ew_nodes), east = std::min(nx, p->ivent + ew_nodes);
  const int south = std::max(1, p->jvent - ns_nodes), north = std::min(ny, p->jvent + ns_nodes);
  for (int ii = west; ii <= east; ++ii) {  // (:351-389)
    for (int jj = south; jj <= north; ++jj) {
      if (ii == p->ivent && jj == p->jvent) continue;
      const double avg_lat = (lat_cc[jj - 1] + p->lat_volcano) / 2.0;
      const double ns_km = (lat_cc[jj - 1] - p->lat_volcano) * DEG2KMLAT;
      const double ew_km = (lon_cc[ii - 1] - p->lon_volcano) * cos(avg_lat * DEG2RAD) * DEG2KMLON;
      const double radnow = sqrt(ns_km * ns_km + ew_km * ew_km);
      if (radnow < cloud_radius) {
        double w;
        if (p->vel_model == 1) {
          const double rexp = 3.0;  // the legacy exponent (:366-367)
          w = edge_speed * MPS_2_KMPHR * (3.0 / 4.0) * (cloud_radius / radnow) *
              (1.0 + (1.0 / 3.0) * pow(radnow / cloud_radius, rexp));
        } else {
          w = edge_speed * MPS_2_KMPHR * pow(cloud_radius / radnow, 0.5);
        }
        const double thetanow = atan2(ns_km, ew_km);
        uvx[(ii - 1) + (long long)nx * (jj - 1)] = w * cos(thetanow);
        uvy[(ii - 1) + (long long)nx * (jj - 1)] = w * sin(thetanow);
      }
    }
  }
  return 0;
}

int ash3d_tr_umbrella_nodes(const double *flux, int32_t nz, int32_t ns, int32_t n_gs_max, int32_t ibase,
                            int32_t itop, const double *scale_fac, double *flux_umb) {
  if (!flux || !scale_fac || !flux_umb || nz < 1 || ns < 1)
    return src_fail("ash3d_tr_umbrella_nodes: null or empty input");
  if (n_gs_max < 0 || n_gs_max > ns) return src_fail("ash3d_tr_umbrella_nodes: n_gs_max outside 0..ns");
  if (itop > nz) return src_fail("ash3d_tr_umbrella_nodes: itop above the grid");
  auto fu = [&](int i, int j, int k, int n) -> double & {
    return flux_umb[(i - 1) + 3 * (j - 1) + 9 * (long long)(k - 1) + 9LL * (nz + 1) * (n - 1)];
  };
  for (long long c = 0; c < 9LL * (nz + 1) * ns; ++c) flux_umb[c] = 0.0;
  for (int k = 1; k <= nz + 1; ++k) {  // (:429-444)
    for (int n = 1; n <= n_gs_max; ++n) {
      const double f = flux[(k - 1) + (long long)(nz + 1) * (n - 1)];
      if (k < ibase) {
        fu(2, 2, k, n) = f;  // below the cloud: the column
      } else if (k <= itop) {
        for (int i = 1; i <= 3; ++i)
          for (int j = 1; j <= 3; ++j) fu(i, j, k, n) = f * scale_fac[k - 1];
      }  // above the cloud: nothing
    }
  }
  return 0;
}

int ash3d_tr_umbrella_vol_inc(const double *flux, int32_t nz, int32_t ns, int32_t n_gs_max, int32_t ibase,
                              int32_t itop, const double *avg_stencil, const double *kappa_vent_col, double dt,
                              double magma_density, double *vol) {
  if (!flux || !avg_stencil || !kappa_vent_col || !vol || nz < 1 || n_gs_max < 0 || n_gs_max > ns)
    return src_fail("ash3d_tr_umbrella_vol_inc: bad input");
  double tmp = 0.0;
  auto f = [&](int k, int n) { return flux[(k - 1) + (long long)(nz + 1) * (n - 1)]; };
  for (int n = 1; n <= n_gs_max; ++n)  // (:489-499)
    for (int k = 1; k <= ibase - 1; ++k)
      tmp = tmp + dt * f(k, n) * kappa_vent_col[k - 1] / magma_density / KM3_2_M3;
  for (int i = 1; i <= 3; ++i)  // (:501-516), levels below 1 skipped
    for (int j = 1; j <= 3; ++j)
      for (int k = std::max(ibase, 1); k <= itop; ++k)
        for (int n = 1; n <= n_gs_max; ++n)
          tmp = tmp + dt * f(k, n) * avg_stencil[(i - 1) + 3 * (j - 1) + 9 * (k - 1)] * kappa_vent_col[k - 1] /
                          magma_density / KM3_2_M3;
  *vol = tmp;
  return 0;
}

}  // extern "C"
