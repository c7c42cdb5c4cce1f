// Device pieces of the limited donor-cell-upwind update shared by the
// sweep kernels (src/AdvectionHorz_DCU.f90:178-231).
#pragma once

#include <hip/hip_runtime.h>

namespace ash3d {

constexpr double EPS_THRESH = 1.0e-10;  // src/Ash3d_VariableModules.f90:273
// Limited Delta Q, src/AdvectionHorz_DCU.f90:186-226 (limiter is uniform
// across the launch, so the switch does not diverge).
__device__ inline double limited_dq(int lim, double dq, double dqu, double theta) {
  switch (lim) {
    case 1: return dq;                                      // Lax-Wendroff
    case 2: return dqu;                                     // Beam-Warming
    case 3: return dq * (0.5 * (1.0 + theta));              // Fromm
    case 4: return dq * fmax(0.0, fmin(1.0, theta));        // minmod
    case 5: return dq * fmax(fmax(0.0, fmin(1.0, 2.0 * theta)), fmin(2.0, theta));  // superbee
    case 6: return dq * fmax(0.0, fmin(fmin((1.0 + theta) / 2.0, 2.0), 2.0 * theta));  // MC
    default: return 0.0;
  }
}

// fss_I, the limited second-order flux correction at interface I
// (:179-231): dq0 = dq_I, dqm = dq_{I-1}, dqp = dq_{I+1}, u = usig_I,
// dtv = dt/kappa of cell I.
__device__ inline double limited_flux(int lim, double dqm, double dq0, double dqp, double u,
                                      double dtv) {
  if (fabs(dq0) <= EPS_THRESH) return 0.0;
  double dqu = 0.0, theta = 0.0;
  if (lim != 1) {
    dqu = (u > 0.0) ? dqm : dqp;
    if (fabs(dqu) > 3.0 * fabs(dq0) || fabs(dq0) < EPS_THRESH)
      theta = copysign(3.0, dqu * dq0);
    else
      theta = dqu / dq0;
  }
  const double ldq = limited_dq(lim, dq0, dqu, theta);
  const double aus = fabs(u);
  return 0.5 * aus * (1.0 - dtv * aus) * ldq;
}

}  // namespace ash3d
