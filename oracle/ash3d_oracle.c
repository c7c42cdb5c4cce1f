/*
 * ash3d_oracle.c -- CPU oracle (plain C restatement) of the Ash3d transport
 * step.  TEST INFRASTRUCTURE ONLY: see ash3d_oracle.h for who may use it.
 *
 * Every routine follows the cited reference lines, including the Fortran
 * left-to-right evaluation order of each expression, so that a build with
 * -ffp-contract=off reproduces the reference arithmetic operation for
 * operation.  Rows are processed with private scratch arrays exactly as the
 * reference does (gather, compute, scatter), and the advection row loops
 * carry the same OpenMP static schedule as the reference
 * (src/AdvectionHorz_DCU.f90:126-138, src/AdvectionVert_DCU.f90:155-166);
 * diffusion is serial, as in the reference.
 */
#include "ash3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/Ash3d_VariableModules.f90:273 */
#define EPS_THRESH 1.0e-10

/* ---- layout helpers (lower bounds -1 for padded arrays) ---------------- */
#define NXP(s) ((s)->nx + 4)
#define NYP(s) ((s)->ny + 4)
#define NZP(s) ((s)->nz + 4)
#define NCELL(s) ((size_t)NXP(s) * NYP(s) * NZP(s))

static inline size_t P3(const orc_state *s, int i, int j, int k) {
  return (size_t)(i + 1) + (size_t)NXP(s) * ((size_t)(j + 1) + (size_t)NYP(s) * (size_t)(k + 1));
}
/* concen_pd(i,j,k,n,t), n is 1-based, t in {0,1} */
static inline size_t PC(const orc_state *s, int i, int j, int k, int n, int t) {
  return P3(s, i, j, k) + NCELL(s) * ((size_t)(n - 1) + (size_t)s->ns * (size_t)t);
}
/* vf_pd(i,j,k,n) */
static inline size_t PF(const orc_state *s, int i, int j, int k, int n) {
  return P3(s, i, j, k) + NCELL(s) * (size_t)(n - 1);
}
/* kx(0:nx+1,0:ny+1,0:nz+1) */
static inline size_t PK(const orc_state *s, int i, int j, int k) {
  return (size_t)i + (size_t)(s->nx + 2) * ((size_t)j + (size_t)(s->ny + 2) * (size_t)k);
}

#define TS0 0
#define TS1 1

/* ---- Set_BC ------------------------------------------------------------ */
/* src/Set_BC.f90:15-126 */
int orc_set_bc(orc_state *s, int bc_code) {
  const int nx = s->nx, ny = s->ny, nz = s->nz, ns = s->ns;
  int i, j, k, n;
  if (bc_code == 1) {
    /* velocities: zero-order extrapolation, :44-70 */
    for (k = -1; k <= nz + 2; k++)
      for (j = -1; j <= ny + 2; j++) {
        double *vx = s->vx_pd;
        if (s->is_periodic) {
          vx[P3(s, -1, j, k)] = vx[P3(s, nx - 1, j, k)];
          vx[P3(s, 0, j, k)] = vx[P3(s, nx, j, k)];
          vx[P3(s, nx + 1, j, k)] = vx[P3(s, 1, j, k)];
          vx[P3(s, nx + 2, j, k)] = vx[P3(s, 2, j, k)];
        } else {
          vx[P3(s, -1, j, k)] = vx[P3(s, 1, j, k)];
          vx[P3(s, 0, j, k)] = vx[P3(s, 1, j, k)];
          vx[P3(s, nx + 1, j, k)] = vx[P3(s, nx, j, k)];
          vx[P3(s, nx + 2, j, k)] = vx[P3(s, nx, j, k)];
        }
      }
    for (k = -1; k <= nz + 2; k++)
      for (i = -1; i <= nx + 2; i++) {
        double *vy = s->vy_pd;
        vy[P3(s, i, -1, k)] = vy[P3(s, i, 1, k)];
        vy[P3(s, i, 0, k)] = vy[P3(s, i, 1, k)];
        vy[P3(s, i, ny + 1, k)] = vy[P3(s, i, ny, k)];
        vy[P3(s, i, ny + 2, k)] = vy[P3(s, i, ny, k)];
      }
    for (j = -1; j <= ny + 2; j++)
      for (i = -1; i <= nx + 2; i++) {
        double *vz = s->vz_pd;
        vz[P3(s, i, j, -1)] = vz[P3(s, i, j, 1)];
        vz[P3(s, i, j, 0)] = vz[P3(s, i, j, 1)];
        vz[P3(s, i, j, nz + 1)] = vz[P3(s, i, j, nz)];
        vz[P3(s, i, j, nz + 2)] = vz[P3(s, i, j, nz)];
      }
    for (n = 1; n <= ns; n++)
      for (j = -1; j <= ny + 2; j++)
        for (i = -1; i <= nx + 2; i++) {
          double *vf = s->vf_pd;
          vf[PF(s, i, j, -1, n)] = vf[PF(s, i, j, 1, n)];
          vf[PF(s, i, j, 0, n)] = vf[PF(s, i, j, 1, n)];
          vf[PF(s, i, j, nz + 1, n)] = vf[PF(s, i, j, nz, n)];
          vf[PF(s, i, j, nz + 2, n)] = vf[PF(s, i, j, nz, n)];
        }
    /* concentrations: zero all ghosts, :77-84 */
    for (n = 1; n <= ns; n++) {
      double *c = s->concen_pd;
      for (k = -1; k <= nz + 2; k++)
        for (j = -1; j <= ny + 2; j++) {
          c[PC(s, -1, j, k, n, TS0)] = 0.0;
          c[PC(s, 0, j, k, n, TS0)] = 0.0;
          c[PC(s, nx + 1, j, k, n, TS0)] = 0.0;
          c[PC(s, nx + 2, j, k, n, TS0)] = 0.0;
        }
      for (k = -1; k <= nz + 2; k++)
        for (i = -1; i <= nx + 2; i++) {
          c[PC(s, i, -1, k, n, TS0)] = 0.0;
          c[PC(s, i, 0, k, n, TS0)] = 0.0;
          c[PC(s, i, ny + 1, k, n, TS0)] = 0.0;
          c[PC(s, i, ny + 2, k, n, TS0)] = 0.0;
        }
      for (j = -1; j <= ny + 2; j++)
        for (i = -1; i <= nx + 2; i++) {
          c[PC(s, i, j, -1, n, TS0)] = 0.0;
          c[PC(s, i, j, 0, n, TS0)] = 0.0;
          c[PC(s, i, j, nz + 1, n, TS0)] = 0.0;
          c[PC(s, i, j, nz + 2, n, TS0)] = 0.0;
        }
      /* periodic x copy, :86-91 */
      if (s->is_periodic)
        for (k = -1; k <= nz + 2; k++)
          for (j = -1; j <= ny + 2; j++) {
            c[PC(s, -1, j, k, n, TS0)] = c[PC(s, nx - 1, j, k, n, TS0)];
            c[PC(s, 0, j, k, n, TS0)] = c[PC(s, nx, j, k, n, TS0)];
            c[PC(s, nx + 1, j, k, n, TS0)] = c[PC(s, 1, j, k, n, TS0)];
            c[PC(s, nx + 2, j, k, n, TS0)] = c[PC(s, 2, j, k, n, TS0)];
          }
    }
    return 0;
  } else if (bc_code == 2) {
    /* Neumann mirror x (with periodic override), then y, then z, :98-118 */
    double *c = s->concen_pd;
    for (n = 1; n <= ns; n++) {
      for (k = -1; k <= nz + 2; k++)
        for (j = -1; j <= ny + 2; j++) {
          c[PC(s, -1, j, k, n, TS0)] = c[PC(s, 1, j, k, n, TS0)];
          c[PC(s, 0, j, k, n, TS0)] = c[PC(s, 1, j, k, n, TS0)];
          c[PC(s, nx + 1, j, k, n, TS0)] = c[PC(s, nx, j, k, n, TS0)];
          c[PC(s, nx + 2, j, k, n, TS0)] = c[PC(s, nx, j, k, n, TS0)];
        }
      if (s->is_periodic)
        for (k = -1; k <= nz + 2; k++)
          for (j = -1; j <= ny + 2; j++) {
            c[PC(s, -1, j, k, n, TS0)] = c[PC(s, nx - 1, j, k, n, TS0)];
            c[PC(s, 0, j, k, n, TS0)] = c[PC(s, nx, j, k, n, TS0)];
            c[PC(s, nx + 1, j, k, n, TS0)] = c[PC(s, 1, j, k, n, TS0)];
            c[PC(s, nx + 2, j, k, n, TS0)] = c[PC(s, 2, j, k, n, TS0)];
          }
      for (k = -1; k <= nz + 2; k++)
        for (i = -1; i <= nx + 2; i++) {
          c[PC(s, i, -1, k, n, TS0)] = c[PC(s, i, 1, k, n, TS0)];
          c[PC(s, i, 0, k, n, TS0)] = c[PC(s, i, 1, k, n, TS0)];
          c[PC(s, i, ny + 1, k, n, TS0)] = c[PC(s, i, ny, k, n, TS0)];
          c[PC(s, i, ny + 2, k, n, TS0)] = c[PC(s, i, ny, k, n, TS0)];
        }
      for (j = -1; j <= ny + 2; j++)
        for (i = -1; i <= nx + 2; i++) {
          c[PC(s, i, j, -1, n, TS0)] = c[PC(s, i, j, 1, n, TS0)];
          c[PC(s, i, j, 0, n, TS0)] = c[PC(s, i, j, 1, n, TS0)];
          c[PC(s, i, j, nz + 1, n, TS0)] = c[PC(s, i, j, nz, n, TS0)];
          c[PC(s, i, j, nz + 2, n, TS0)] = c[PC(s, i, j, nz, n, TS0)];
        }
    }
    return 0;
  }
  /* src/Set_BC.f90:119-124: unknown code is fatal */
  return 1;
}

/* ---- the limiter -------------------------------------------------------- */
/* Limited Delta Q for one interface, src/AdvectionHorz_DCU.f90:186-226.
 * theta is only meaningful for limiters that need it. */
static inline double limited_dq(int lim, double dq, double dqu, double theta) {
  switch (lim) {
  case ORC_LIM_LAXWEN:
    return dq; /* :188 */
  case ORC_LIM_BW:
    return dqu; /* :209 */
  case ORC_LIM_FROMM:
    return dq * (0.5 * (1.0 + theta)); /* :213 */
  case ORC_LIM_MINMOD:
    return dq * fmax(0.0, fmin(1.0, theta)); /* :217 */
  case ORC_LIM_SUPERBEE:
    return dq * fmax(fmax(0.0, fmin(1.0, 2.0 * theta)), fmin(2.0, theta)); /* :221 */
  case ORC_LIM_MC:
    return dq * fmax(0.0, fmin(fmin((1.0 + theta) / 2.0, 2.0), 2.0 * theta)); /* :225 */
  default:
    return 0.0; /* no limiter macro: ldq_I keeps its initial 0 (:173) */
  }
}

double orc_limiter_phi(int lim, double theta) { return limited_dq(lim, 1.0, theta, theta); }

/* ---- one DCU row ------------------------------------------------------- */
/* One row of the limited donor-cell-upwind sweep, shared by x, y and z
 * (src/AdvectionHorz_DCU.f90:143-281; src/AdvectionVert_DCU.f90:171-298).
 * All arrays are indexed by the cell index r in [-1, N+2] at offset r+1.
 * deld is NULL unless the topographic term is active (ZScaling_ID == 2). */
typedef struct {
  double *dt_vol, *usig, *dq, *fss, *fsl, *fsr, *deldond;
} row_work;

static void row_work_alloc(row_work *w, int N) {
  size_t m = (size_t)(N + 4);
  double *p = (double *)calloc(7 * m, sizeof(double));
  w->dt_vol = p;
  w->usig = p + m;
  w->dq = p + 2 * m;
  w->fss = p + 3 * m;
  w->fsl = p + 4 * m;
  w->fsr = p + 5 * m;
  w->deldond = p + 6 * m;
}
static void row_work_free(row_work *w) { free(w->dt_vol); }

#define R(a, r) ((a)[(r) + 1])

static void dcu_row(int N, int rmin, int rmax, double dt, int lim, const double *q,
                    const double *vel, const double *sig, const double *kap,
                    const double *deld, double *upd, row_work *w) {
  int r, I, c;
  /* :155-156  dt_vol = dt / kappa, on rmin-2 .. rmax+2 */
  for (r = rmin - 2; r <= rmax + 2; r++) R(w->dt_vol, r) = dt / R(kap, r);
  /* :159-162  usig is zero except on interfaces rmin-1 .. rmax+1 */
  for (r = -1; r <= N + 2; r++) R(w->usig, r) = 0.0;
  for (I = rmin - 1; I <= rmax + 1; I++)
    R(w->usig, I) = 0.5 * (R(vel, I - 1) + R(vel, I)) * R(sig, I);
  /* :163-166  topographic term */
  for (r = -1; r <= N + 2; r++) R(w->deldond, r) = 0.0;
  if (deld)
    for (I = rmin - 1; I <= rmax + 1; I++) R(w->deldond, I) = R(deld, I);
  /* :168-169 */
  for (I = rmin - 1; I <= rmax + 2; I++) R(w->dq, I) = R(q, I) - R(q, I - 1);
  /* :173-176 */
  for (I = rmin - 1; I <= rmax + 2; I++) {
    R(w->fsl, I) = 0.0;
    R(w->fsr, I) = 0.0;
    R(w->fss, I) = 0.0;
  }
  /* :178-231  limited second-order correction */
  if (lim != ORC_LIM_NONE) {
    for (I = rmin; I <= rmax + 1; I++) {
      double dqI = R(w->dq, I), dqu = 0.0, theta = 0.0, ldq, aus;
      if (fabs(dqI) <= EPS_THRESH) continue; /* :183 */
      if (lim != ORC_LIM_LAXWEN) {
        /* :193-197 upwind interface */
        if (R(w->usig, I) > 0.0)
          dqu = R(w->dq, I - 1);
        else
          dqu = R(w->dq, I + 1);
        /* :200-205 */
        if (fabs(dqu) > 3.0 * fabs(dqI) || fabs(dqI) < EPS_THRESH)
          theta = copysign(3.0, dqu * dqI);
        else
          theta = dqu / dqI;
      }
      ldq = limited_dq(lim, dqI, dqu, theta);
      aus = fabs(R(w->usig, I)); /* :227 */
      R(w->fss, I) = 0.5 * aus * (1.0 - R(w->dt_vol, I) * aus) * ldq; /* :230 */
    }
  }
  /* :238-259  fluctuations with the divergence correction */
  for (I = rmin - 1; I <= rmax + 1; I++) {
    double u = R(w->usig, I), um = R(w->usig, I - 1), up = R(w->usig, I + 1);
    double divu_p, divu_m;
    R(w->fsr, I) = fmax(0.0, u) * R(w->dq, I);
    R(w->fsl, I) = fmin(0.0, u) * R(w->dq, I);
    divu_p = (fmax(0.0, u) - fmax(0.0, um));
    divu_m = (fmin(0.0, up) - fmin(0.0, u));
    R(w->fsr, I) = R(w->fsr, I) + R(q, I) * divu_m;
    R(w->fsl, I) = R(w->fsl, I) + R(q, I - 1) * divu_p;
    if (deld) {
      R(w->fsr, I) = R(w->fsr, I) +
                     0.5 * R(vel, I) * (R(sig, I) + R(sig, I + 1)) * R(q, I) * R(w->deldond, I);
      R(w->fsl, I) = R(w->fsl, I) -
                     0.5 * R(vel, I - 1) * (R(sig, I - 1) + R(sig, I)) * R(q, I - 1) *
                         R(w->deldond, I - 1);
    }
  }
  /* :264-281  cell update, one ghost each side for the outflow */
  for (c = rmin - 1; c <= rmax + 1; c++) {
    double LFluct_Rbound = R(w->fsl, c + 1);
    double RFluct_Lbound = R(w->fsr, c);
    double LimFlux_Rbound = R(w->fss, c + 1);
    double LimFlux_Lbound = R(w->fss, c);
    R(upd, c) = -R(w->dt_vol, c) * (LFluct_Rbound + RFluct_Lbound + LimFlux_Rbound - LimFlux_Lbound);
  }
}

static void zero_ts1(orc_state *s) {
  memset(s->concen_pd + NCELL(s) * (size_t)s->ns, 0, sizeof(double) * NCELL(s) * (size_t)s->ns);
}

static void copy_ts1_to_ts0_interior(orc_state *s) {
  int i, j, k, n;
  for (n = 1; n <= s->ns; n++)
    for (k = 1; k <= s->nz; k++)
      for (j = 1; j <= s->ny; j++)
        for (i = 1; i <= s->nx; i++)
          s->concen_pd[PC(s, i, j, k, n, TS0)] = s->concen_pd[PC(s, i, j, k, n, TS1)];
}

/* ---- advect_x : src/AdvectionHorz_DCU.f90:68-305 ----------------------- */
void orc_advect_x(orc_state *s) {
  const int nx = s->nx, rmin = s->imin, rmax = s->imax;
  const int use_topo = (s->zscaling_id == 2);
  int n;
  orc_set_bc(s, 1); /* :115 */
  zero_ts1(s);      /* :122 */
  for (n = 1; n <= s->ns; n++) {
    int k, j;
    if (!s->is_aloft[n - 1]) continue; /* :125 */
#pragma omp parallel private(j, k)
    {
      /* private row arrays, as the reference's PRIVATE list (:133-137) */
      row_work w;
      double *buf = (double *)malloc(sizeof(double) * 6 * (size_t)(nx + 4));
      double *q = buf, *vel = buf + (nx + 4), *sig = buf + 2 * (nx + 4);
      double *kap = buf + 3 * (nx + 4), *deld = buf + 4 * (nx + 4), *upd = buf + 5 * (nx + 4);
      row_work_alloc(&w, nx);
#pragma omp for collapse(2) schedule(static)
    for (k = s->kmin; k <= s->kmax; k++)
      for (j = s->jmin; j <= s->jmax; j++) {
        int r;
        /* :143-146 gather */
        for (r = rmin - 2; r <= rmax + 2; r++) {
          R(q, r) = s->concen_pd[PC(s, r, j, k, n, TS0)];
          R(vel, r) = s->vx_pd[P3(s, r, j, k)];
          R(sig, r) = s->sigma_nx_pd[P3(s, r, j, k)];
          R(kap, r) = s->kappa_pd[P3(s, r, j, k)];
        }
        if (use_topo)
          for (r = rmin - 1; r <= rmax + 1; r++)
            R(deld, r) = s->deldxond_cc[(size_t)(r + 1) + (size_t)(nx + 4) * (size_t)(j + 1)];
        dcu_row(nx, rmin, rmax, s->dt, s->limiter, q, vel, sig, kap, use_topo ? deld : NULL, upd, &w);
        /* :285-287 */
        for (r = rmin; r <= rmax; r++)
          s->concen_pd[PC(s, r, j, k, n, TS1)] = s->concen_pd[PC(s, r, j, k, n, TS0)] + R(upd, r);
        /* :289-294 side outflow, skipped when periodic */
        if (rmin == 1 && !s->is_periodic) {
          size_t o = (size_t)(j + 1) + (size_t)NYP(s) * ((size_t)(k + 1) + (size_t)NZP(s) * (size_t)(n - 1));
          s->outflow_yz1_pd[o] = s->outflow_yz1_pd[o] + R(upd, 0);
        }
        if (rmax == nx && !s->is_periodic) {
          size_t o = (size_t)(j + 1) + (size_t)NYP(s) * ((size_t)(k + 1) + (size_t)NZP(s) * (size_t)(n - 1));
          s->outflow_yz2_pd[o] = s->outflow_yz2_pd[o] + R(upd, nx + 1);
        }
      }
      row_work_free(&w);
      free(buf);
    }
  }
  copy_ts1_to_ts0_interior(s); /* :302-303 */
}

/* ---- advect_y : src/AdvectionHorz_DCU.f90:322-556 ---------------------- */
void orc_advect_y(orc_state *s) {
  const int ny = s->ny, rmin = s->jmin, rmax = s->jmax;
  const int use_topo = (s->zscaling_id == 2);
  int n;
  orc_set_bc(s, 1); /* :366 */
  zero_ts1(s);      /* :373 */
  for (n = 1; n <= s->ns; n++) {
    int k, i;
    if (!s->is_aloft[n - 1]) continue;
#pragma omp parallel private(i, k)
    {
      row_work w;
      double *buf = (double *)malloc(sizeof(double) * 6 * (size_t)(ny + 4));
      double *q = buf, *vel = buf + (ny + 4), *sig = buf + 2 * (ny + 4);
      double *kap = buf + 3 * (ny + 4), *deld = buf + 4 * (ny + 4), *upd = buf + 5 * (ny + 4);
      row_work_alloc(&w, ny);
#pragma omp for collapse(2) schedule(static)
    for (k = s->kmin; k <= s->kmax; k++)
      for (i = s->imin; i <= s->imax; i++) {
        int r;
        /* :394-397 */
        for (r = rmin - 2; r <= rmax + 2; r++) {
          R(q, r) = s->concen_pd[PC(s, i, r, k, n, TS0)];
          R(vel, r) = s->vy_pd[P3(s, i, r, k)];
          R(sig, r) = s->sigma_ny_pd[P3(s, i, r, k)];
          R(kap, r) = s->kappa_pd[P3(s, i, r, k)];
        }
        if (use_topo)
          for (r = rmin - 1; r <= rmax + 1; r++)
            R(deld, r) = s->deldyond_cc[(size_t)(i + 1) + (size_t)NXP(s) * (size_t)(r + 1)];
        dcu_row(ny, rmin, rmax, s->dt, s->limiter, q, vel, sig, kap, use_topo ? deld : NULL, upd, &w);
        /* :536-538 */
        for (r = rmin; r <= rmax; r++)
          s->concen_pd[PC(s, i, r, k, n, TS1)] = s->concen_pd[PC(s, i, r, k, n, TS0)] + R(upd, r);
        /* :540-545 no periodic test in y */
        if (rmin == 1) {
          size_t o = (size_t)(i + 1) + (size_t)NXP(s) * ((size_t)(k + 1) + (size_t)NZP(s) * (size_t)(n - 1));
          s->outflow_xz1_pd[o] = s->outflow_xz1_pd[o] + R(upd, 0);
        }
        if (rmax == ny) {
          size_t o = (size_t)(i + 1) + (size_t)NXP(s) * ((size_t)(k + 1) + (size_t)NZP(s) * (size_t)(n - 1));
          s->outflow_xz2_pd[o] = s->outflow_xz2_pd[o] + R(upd, ny + 1);
        }
      }
      row_work_free(&w);
      free(buf);
    }
  }
  copy_ts1_to_ts0_interior(s); /* :553-554 */
}

/* ---- advect_z (== advect_sigma) : src/AdvectionVert_DCU.f90:100-323 ----- */
void orc_advect_z(orc_state *s) {
  const int nz = s->nz, rmin = s->kmin, rmax = s->kmax;
  const int ncells = rmax - rmin + 1;
  int n, i, j;
  orc_set_bc(s, 1); /* :144 */
  zero_ts1(s);      /* :151 */
  for (n = 1; n <= s->ns; n++) {
    if (!s->is_aloft[n - 1]) continue; /* :154 */
#pragma omp parallel private(i, j)
    {
      row_work w;
      double *buf = (double *)malloc(sizeof(double) * 5 * (size_t)(nz + 4));
      double *q = buf, *vel = buf + (nz + 4), *sig = buf + 2 * (nz + 4);
      double *kap = buf + 3 * (nz + 4), *upd = buf + 4 * (nz + 4);
      row_work_alloc(&w, nz);
#pragma omp for collapse(2) schedule(static)
    for (j = s->jmin; j <= s->jmax; j++)
      for (i = s->imin; i <= s->imax; i++) {
        int r;
        /* :171-176, velocity = (vz + vh) + vf(n): settling */
        for (r = rmin - 2; r <= rmax + 2; r++) {
          R(q, r) = s->concen_pd[PC(s, i, j, r, n, TS0)];
          R(vel, r) = (s->vz_pd[P3(s, i, j, r)] + s->vh_pd[P3(s, i, j, r)] + s->vf_pd[PF(s, i, j, r, n)]);
          R(sig, r) = s->sigma_nz_pd[P3(s, i, j, r)];
          R(kap, r) = s->kappa_pd[P3(s, i, j, r)];
        }
        dcu_row(nz, rmin, rmax, s->dt, s->limiter, q, vel, sig, kap, NULL, upd, &w);
        /* :302-304 */
        for (r = rmin; r <= rmax; r++)
          s->concen_pd[PC(s, i, j, r, n, TS1)] = s->concen_pd[PC(s, i, j, r, n, TS0)] + R(upd, r);
        /* :306-311 deposit and top outflow (top uses update_cc(ncells+1)) */
        {
          size_t o = (size_t)(i + 1) + (size_t)NXP(s) * ((size_t)(j + 1) + (size_t)NYP(s) * (size_t)(n - 1));
          if (rmin == 1) s->outflow_xy1_pd[o] = s->outflow_xy1_pd[o] + R(upd, 0);
          if (rmax == nz) {
            /* update_cc is only defined on rmin-1..rmax+1; outside that range
             * the reference reads an unset private array element.  The
             * oracle uses 0 there (only reachable with kmin > (nz+3)/2). */
            int t = ncells + 1;
            double u = (t >= rmin - 1 && t <= rmax + 1) ? R(upd, t) : 0.0;
            s->outflow_xy2_pd[o] = s->outflow_xy2_pd[o] + u;
          }
        }
      }
      row_work_free(&w);
      free(buf);
    }
  }
  copy_ts1_to_ts0_interior(s); /* :318-319 */
  /* :321 */
  for (n = 1; n <= s->ns; n++)
    for (j = 1; j <= s->ny; j++)
      for (i = 1; i <= s->nx; i++)
        s->deposit_granularity[(size_t)(i - 1) + (size_t)s->nx * ((size_t)(j - 1) + (size_t)s->ny * (size_t)(n - 1))] =
            s->outflow_xy1_pd[(size_t)(i + 1) + (size_t)NXP(s) * ((size_t)(j + 1) + (size_t)NYP(s) * (size_t)(n - 1))];
}

/* src/AdvectionHorz.f90:47-85 with FAST_SUBGRID off */
void orc_advect_horz(orc_state *s, int itoggle) {
  s->imin = 1; s->imax = s->nx;
  s->jmin = 1; s->jmax = s->ny;
  s->kmin = 1; s->kmax = s->nz;
  if (itoggle % 2 == 0) {
    orc_advect_y(s);
    orc_advect_x(s);
  } else {
    orc_advect_x(s);
    orc_advect_y(s);
  }
}

/* src/AdvectionVert_DCU.f90:64-82: advect_sigma has the same body as advect_z */
void orc_advect_vert(orc_state *s) { orc_advect_z(s); }

/* ---- explicit diffusion : src/Diffusion.F90:266-650 -------------------- */
/* One line of the explicit flux form (:335-377).  q, kap indexed r+1 for
 * r in [0, N+1]; sig and K indexed by interface I in [1, N+1] at I, K by
 * cell r in [0, N+1] at r. */
static void diff_line(int N, int rmin, int rmax, double dt, const double *q, const double *kap,
                      const double *sig, const double *kk, double *upd) {
  /* locals of :285-297, zero-initialised as at :335-341 */
  double *ds = (double *)calloc((size_t)(N + 2), sizeof(double));
  double *kds = (double *)calloc((size_t)(N + 2), sizeof(double));
  double *dq = (double *)calloc((size_t)(N + 2), sizeof(double));
  int I, c;
  for (I = rmin; I <= rmax + 1; I++) dq[I] = q[I] - q[I - 1]; /* :347-348 */
  for (I = rmin; I <= rmax + 1; I++) {
    ds[I] = 2.0 * sig[I] / (kap[I - 1] + kap[I]);   /* :357 */
    kds[I] = 0.5 * (kk[I - 1] + kk[I]) * ds[I];     /* :359 */
  }
  for (c = rmin; c <= rmax; c++) {
    double LFluct_Rbound = dt * kds[c + 1] * dq[c + 1] * (sig[c + 1] / kap[c]); /* :367-368 */
    double RFluct_Lbound = -(dt * kds[c] * dq[c] * (sig[c] / kap[c]));          /* :369-370 */
    upd[c] = LFluct_Rbound + RFluct_Lbound;                                      /* :372 */
  }
  free(ds);
  free(kds);
  free(dq);
}

/* Neumann ghosts of one direction (dir 0:x, 1:y, 2:z), as each diffusion
 * routine sets them before its solve (:317-327, :457-460, :587-590,
 * :756-768, :1018-1021, :1277-1280).  x applies the periodic override. */
static void neumann_ghosts(orc_state *s, int dir) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  double *c = s->concen_pd;
  int i, j, k, n;
  for (n = 1; n <= s->ns; n++) {
    if (dir == 0) {
      for (k = -1; k <= nz + 2; k++)
        for (j = -1; j <= ny + 2; j++) {
          c[PC(s, -1, j, k, n, TS0)] = c[PC(s, 1, j, k, n, TS0)];
          c[PC(s, 0, j, k, n, TS0)] = c[PC(s, 1, j, k, n, TS0)];
          c[PC(s, nx + 1, j, k, n, TS0)] = c[PC(s, nx, j, k, n, TS0)];
          c[PC(s, nx + 2, j, k, n, TS0)] = c[PC(s, nx, j, k, n, TS0)];
        }
      if (s->is_periodic)
        for (k = -1; k <= nz + 2; k++)
          for (j = -1; j <= ny + 2; j++) {
            c[PC(s, -1, j, k, n, TS0)] = c[PC(s, nx - 1, j, k, n, TS0)];
            c[PC(s, 0, j, k, n, TS0)] = c[PC(s, nx, j, k, n, TS0)];
            c[PC(s, nx + 1, j, k, n, TS0)] = c[PC(s, 1, j, k, n, TS0)];
            c[PC(s, nx + 2, j, k, n, TS0)] = c[PC(s, 2, j, k, n, TS0)];
          }
    } else if (dir == 1) {
      for (k = -1; k <= nz + 2; k++)
        for (i = -1; i <= nx + 2; i++) {
          c[PC(s, i, -1, k, n, TS0)] = c[PC(s, i, 1, k, n, TS0)];
          c[PC(s, i, 0, k, n, TS0)] = c[PC(s, i, 1, k, n, TS0)];
          c[PC(s, i, ny + 1, k, n, TS0)] = c[PC(s, i, ny, k, n, TS0)];
          c[PC(s, i, ny + 2, k, n, TS0)] = c[PC(s, i, ny, k, n, TS0)];
        }
    } else {
      for (j = -1; j <= ny + 2; j++)
        for (i = -1; i <= nx + 2; i++) {
          c[PC(s, i, j, -1, n, TS0)] = c[PC(s, i, j, 1, n, TS0)];
          c[PC(s, i, j, 0, n, TS0)] = c[PC(s, i, j, 1, n, TS0)];
          c[PC(s, i, j, nz + 1, n, TS0)] = c[PC(s, i, j, nz, n, TS0)];
          c[PC(s, i, j, nz + 2, n, TS0)] = c[PC(s, i, j, nz, n, TS0)];
        }
    }
  }
}

/* Strided access to one line of a padded 3-D array along direction dir. */
typedef struct {
  size_t base;   /* offset of the first cell of the line */
  size_t stride; /* offset step per cell */
} line_t;

static line_t pad_line(const orc_state *s, int dir, int a, int b) {
  line_t L;
  /* base is the offset of cell -1; cell r sits at base + stride*(r+1) */
  if (dir == 0) { L.base = P3(s, -1, a, b); L.stride = 1; }
  else if (dir == 1) { L.base = P3(s, a, -1, b); L.stride = (size_t)NXP(s); }
  else { L.base = P3(s, a, b, -1); L.stride = (size_t)NXP(s) * NYP(s); }
  return L;
}
static line_t k_line(const orc_state *s, int dir, int a, int b) {
  /* K arrays start at 0: cell r sits at base + stride*r */
  line_t L;
  if (dir == 0) { L.base = PK(s, 0, a, b); L.stride = 1; }
  else if (dir == 1) { L.base = PK(s, a, 0, b); L.stride = (size_t)(s->nx + 2); }
  else { L.base = PK(s, a, b, 0); L.stride = (size_t)(s->nx + 2) * (s->ny + 2); }
  return L;
}

/* Range of line indices (rmin..rmax) and of the two transverse indices. */
static void dir_ranges(const orc_state *s, int dir, int *rmin, int *rmax, int *a0, int *a1, int *b0, int *b1) {
  if (dir == 0) { *rmin = s->imin; *rmax = s->imax; *a0 = s->jmin; *a1 = s->jmax; *b0 = s->kmin; *b1 = s->kmax; }
  else if (dir == 1) { *rmin = s->jmin; *rmax = s->jmax; *a0 = s->imin; *a1 = s->imax; *b0 = s->kmin; *b1 = s->kmax; }
  else { *rmin = s->kmin; *rmax = s->kmax; *a0 = s->imin; *a1 = s->imax; *b0 = s->jmin; *b1 = s->jmax; }
}
static int dir_len(const orc_state *s, int dir) { return dir == 0 ? s->nx : (dir == 1 ? s->ny : s->nz); }
static const double *dir_sigma(const orc_state *s, int dir) {
  return dir == 0 ? s->sigma_nx_pd : (dir == 1 ? s->sigma_ny_pd : s->sigma_nz_pd);
}
static const double *dir_k(const orc_state *s, int dir) { return dir == 0 ? s->kx : (dir == 1 ? s->ky : s->kz); }

static void diff_explicit(orc_state *s, int dir) {
  int rmin, rmax, a0, a1, b0, b1, n, a, b, r;
  const int N = dir_len(s, dir);
  const double *sigma = dir_sigma(s, dir), *kk = dir_k(s, dir);
  double *q = (double *)malloc(sizeof(double) * (size_t)(N + 2));
  double *kap = (double *)malloc(sizeof(double) * (size_t)(N + 2));
  double *sig = (double *)malloc(sizeof(double) * (size_t)(N + 2));
  double *kl = (double *)malloc(sizeof(double) * (size_t)(N + 2));
  double *upd = (double *)malloc(sizeof(double) * (size_t)(N + 2));
  dir_ranges(s, dir, &rmin, &rmax, &a0, &a1, &b0, &b1);
  zero_ts1(s);
  neumann_ghosts(s, dir);
  for (n = 1; n <= s->ns; n++) {
    if (!s->is_aloft[n - 1]) continue;
    for (b = b0; b <= b1; b++)
      for (a = a0; a <= a1; a++) {
        line_t L = pad_line(s, dir, a, b), LK = k_line(s, dir, a, b);
        size_t cofs = NCELL(s) * (size_t)(n - 1);
        for (r = 0; r <= N + 1; r++) { q[r] = 0.0; kap[r] = 0.0; sig[r] = 0.0; upd[r] = 0.0; }
        for (r = rmin - 1; r <= rmax + 1; r++) {
          kap[r] = s->kappa_pd[L.base + L.stride * (size_t)(r + 1)];
          q[r] = s->concen_pd[cofs + L.base + L.stride * (size_t)(r + 1)];
        }
        for (r = rmin; r <= rmax + 1; r++) sig[r] = sigma[L.base + L.stride * (size_t)(r + 1)];
        for (r = rmin - 1; r <= rmax + 1; r++) kl[r] = kk[LK.base + LK.stride * (size_t)r];
        diff_line(N, rmin, rmax, s->dt, q, kap, sig, kl, upd);
        for (r = rmin; r <= rmax; r++) {
          size_t o = L.base + L.stride * (size_t)(r + 1);
          s->concen_pd[NCELL(s) * ((size_t)(n - 1) + (size_t)s->ns) + o] = s->concen_pd[cofs + o] + upd[r];
        }
      }
  }
  copy_ts1_to_ts0_interior(s);
  free(q); free(kap); free(sig); free(kl); free(upd);
}

void orc_diff_x(orc_state *s) { diff_explicit(s, 0); }
void orc_diff_y(orc_state *s) { diff_explicit(s, 1); }
void orc_diff_z(orc_state *s) { diff_explicit(s, 2); }

/* ---- LAPACK dgtsv (restated) ------------------------------------------- */
/* Solves A X = B for a general tridiagonal A by Gaussian elimination with
 * partial pivoting (LAPACK dgtsv, reference implementation).  On exit dl
 * holds the second super-diagonal of U, d the diagonal of U and du the first
 * super-diagonal of U; b holds the solution.  info > 0: U(info,info) == 0. */
void orc_dgtsv(int n, int nrhs, double *dl, double *d, double *du, double *b, int ldb, int *info) {
  int i, j;
  double fact, temp;
#define B_(ii, jj) b[(size_t)((ii) - 1) + (size_t)ldb * (size_t)((jj) - 1)]
#define DL_(ii) dl[(ii) - 1]
#define D_(ii) d[(ii) - 1]
#define DU_(ii) du[(ii) - 1]
  *info = 0;
  if (n < 0) { *info = -1; return; }
  if (nrhs < 0) { *info = -2; return; }
  if (ldb < (n > 1 ? n : 1)) { *info = -7; return; }
  if (n == 0) return;
  if (nrhs == 1) {
    for (i = 1; i <= n - 2; i++) {
      if (fabs(D_(i)) >= fabs(DL_(i))) {
        /* no row interchange */
        if (D_(i) != 0.0) {
          fact = DL_(i) / D_(i);
          D_(i + 1) = D_(i + 1) - fact * DU_(i);
          B_(i + 1, 1) = B_(i + 1, 1) - fact * B_(i, 1);
        } else {
          *info = i;
          return;
        }
        DL_(i) = 0.0;
      } else {
        /* interchange rows i and i+1 */
        fact = D_(i) / DL_(i);
        D_(i) = DL_(i);
        temp = D_(i + 1);
        D_(i + 1) = DU_(i) - fact * temp;
        DL_(i) = DU_(i + 1);
        DU_(i + 1) = -fact * DL_(i);
        DU_(i) = temp;
        temp = B_(i, 1);
        B_(i, 1) = B_(i + 1, 1);
        B_(i + 1, 1) = temp - fact * B_(i + 1, 1);
      }
    }
    if (n > 1) {
      i = n - 1;
      if (fabs(D_(i)) >= fabs(DL_(i))) {
        if (D_(i) != 0.0) {
          fact = DL_(i) / D_(i);
          D_(i + 1) = D_(i + 1) - fact * DU_(i);
          B_(i + 1, 1) = B_(i + 1, 1) - fact * B_(i, 1);
        } else {
          *info = i;
          return;
        }
      } else {
        fact = D_(i) / DL_(i);
        D_(i) = DL_(i);
        temp = D_(i + 1);
        D_(i + 1) = DU_(i) - fact * temp;
        DU_(i) = temp;
        temp = B_(i, 1);
        B_(i, 1) = B_(i + 1, 1);
        B_(i + 1, 1) = temp - fact * B_(i + 1, 1);
      }
    }
    if (D_(n) == 0.0) { *info = n; return; }
  } else {
    for (i = 1; i <= n - 2; i++) {
      if (fabs(D_(i)) >= fabs(DL_(i))) {
        if (D_(i) != 0.0) {
          fact = DL_(i) / D_(i);
          D_(i + 1) = D_(i + 1) - fact * DU_(i);
          for (j = 1; j <= nrhs; j++) B_(i + 1, j) = B_(i + 1, j) - fact * B_(i, j);
        } else {
          *info = i;
          return;
        }
        DL_(i) = 0.0;
      } else {
        fact = D_(i) / DL_(i);
        D_(i) = DL_(i);
        temp = D_(i + 1);
        D_(i + 1) = DU_(i) - fact * temp;
        DL_(i) = DU_(i + 1);
        DU_(i + 1) = -fact * DL_(i);
        DU_(i) = temp;
        for (j = 1; j <= nrhs; j++) {
          temp = B_(i, j);
          B_(i, j) = B_(i + 1, j);
          B_(i + 1, j) = temp - fact * B_(i + 1, j);
        }
      }
    }
    if (n > 1) {
      i = n - 1;
      if (fabs(D_(i)) >= fabs(DL_(i))) {
        if (D_(i) != 0.0) {
          fact = DL_(i) / D_(i);
          D_(i + 1) = D_(i + 1) - fact * DU_(i);
          for (j = 1; j <= nrhs; j++) B_(i + 1, j) = B_(i + 1, j) - fact * B_(i, j);
        } else {
          *info = i;
          return;
        }
      } else {
        fact = D_(i) / DL_(i);
        D_(i) = DL_(i);
        temp = D_(i + 1);
        D_(i + 1) = DU_(i) - fact * temp;
        DU_(i) = temp;
        for (j = 1; j <= nrhs; j++) {
          temp = B_(i, j);
          B_(i, j) = B_(i + 1, j);
          B_(i + 1, j) = temp - fact * B_(i + 1, j);
        }
      }
    }
    if (D_(n) == 0.0) { *info = n; return; }
  }
  /* back substitution with U */
  for (j = 1; j <= nrhs; j++) {
    B_(n, j) = B_(n, j) / D_(n);
    if (n > 1) B_(n - 1, j) = (B_(n - 1, j) - DU_(n - 1) * B_(n, j)) / D_(n - 1);
    for (i = n - 2; i >= 1; i--)
      B_(i, j) = (B_(i, j) - DU_(i) * B_(i + 1, j) - DL_(i) * B_(i + 2, j)) / D_(i);
  }
#undef B_
#undef DL_
#undef D_
#undef DU_
}

/* ---- Crank-Nicolson : src/Diffusion.F90:671-911 (x), :932-1167 (y),
 *      :1188-1428 (z) ----------------------------------------------------- */
static void diff_cn(orc_state *s, int dir) {
  int rmin, rmax, a0, a1, b0, b1, n, a, b, r, l;
  const int N = dir_len(s, dir);
  const double *sigma = dir_sigma(s, dir), *kk = dir_k(s, dir);
  const double th = s->imp_fac, dt = s->dt;
  int ncells;
  dir_ranges(s, dir, &rmin, &rmax, &a0, &a1, &b0, &b1);
  ncells = rmax - rmin + 1;
  zero_ts1(s);            /* :753 */
  neumann_ghosts(s, dir); /* :756-768 */
  if (ncells <= 1) return; /* :770, :909 */
  {
    double *q = (double *)malloc(sizeof(double) * (size_t)(N + 2));
    double *kap = (double *)malloc(sizeof(double) * (size_t)(N + 2));
    double *sig = (double *)malloc(sizeof(double) * (size_t)(N + 2));
    double *kl = (double *)malloc(sizeof(double) * (size_t)(N + 2));
    double *vavg = (double *)malloc(sizeof(double) * (size_t)(N + 2));
    double *kavg = (double *)malloc(sizeof(double) * (size_t)(N + 2));
    double *ds = (double *)malloc(sizeof(double) * (size_t)(N + 2));
    double *kds = (double *)malloc(sizeof(double) * (size_t)(N + 2));
    double *DL = (double *)malloc(sizeof(double) * (size_t)(ncells));
    double *D = (double *)malloc(sizeof(double) * (size_t)(ncells));
    double *DU = (double *)malloc(sizeof(double) * (size_t)(ncells));
    double *B = (double *)malloc(sizeof(double) * (size_t)(ncells));
    for (n = 1; n <= s->ns; n++) {
      if (!s->is_aloft[n - 1]) continue;
      for (b = b0; b <= b1; b++)
        for (a = a0; a <= a1; a++) {
          line_t L = pad_line(s, dir, a, b), LK = k_line(s, dir, a, b);
          size_t cofs = NCELL(s) * (size_t)(n - 1);
          int info = 0;
          /* :791-801 */
          for (r = rmin - 1; r <= rmax + 1; r++) {
            kap[r] = s->kappa_pd[L.base + L.stride * (size_t)(r + 1)];
            q[r] = s->concen_pd[cofs + L.base + L.stride * (size_t)(r + 1)];
            kl[r] = kk[LK.base + LK.stride * (size_t)r];
          }
          for (r = rmin; r <= rmax + 1; r++) {
            sig[r] = sigma[L.base + L.stride * (size_t)(r + 1)];
            vavg[r] = 0.5 * (kap[r - 1] + kap[r]);
            kavg[r] = 0.5 * (kl[r - 1] + kl[r]);
            ds[r] = sig[r] / vavg[r];
            kds[r] = kavg[r] * ds[r];
          }
          /* :807-853 */
          for (l = 1; l <= ncells; l++) {
            const int I = rmin - 1 + l, c = rmin - 1 + l;
            double LeftFac = dt * kds[I] * sig[I] / kap[c];
            double RightFac = dt * kds[I + 1] * sig[I + 1] / kap[c];
            double CenterFac = LeftFac + RightFac;
            if (l == 1) {
              double BC_left_t0 = q[rmin];
              D[l - 1] = 1.0 + (th)*CenterFac - (th)*LeftFac;
              DU[l - 1] = -((th)*RightFac);
              B[l - 1] = (1.0 - th) * LeftFac * BC_left_t0 + (1.0 - (1.0 - th) * CenterFac) * q[c] +
                         (1.0 - th) * RightFac * q[c + 1];
            } else if (l < ncells) {
              DL[l - 2] = -((th)*LeftFac);
              D[l - 1] = 1.0 + (th)*CenterFac;
              DU[l - 1] = -((th)*RightFac);
              B[l - 1] = (1.0 - th) * LeftFac * q[c - 1] + (1.0 - (1.0 - th) * CenterFac) * q[c] +
                         (1.0 - th) * RightFac * q[c + 1];
            } else {
              double BC_right_t0 = q[rmin - 1 + ncells];
              DL[l - 2] = -((th)*LeftFac);
              D[l - 1] = 1.0 + (th)*CenterFac - (th)*RightFac;
              B[l - 1] = (1.0 - th) * LeftFac * q[c - 1] + (1.0 - (1.0 - th) * CenterFac) * q[c] +
                         (1.0 - th) * RightFac * BC_right_t0;
            }
          }
          /* :881-890 (info is ignored by the reference; the oracle records it) */
          orc_dgtsv(ncells, 1, DL, D, DU, B, ncells, &info);
          if (info != 0) s->last_info = info;
          for (r = rmin; r <= rmax; r++)
            s->concen_pd[NCELL(s) * ((size_t)(n - 1) + (size_t)s->ns) + L.base + L.stride * (size_t)(r + 1)] =
                B[r - rmin];
        }
    }
    free(q); free(kap); free(sig); free(kl); free(vavg); free(kavg); free(ds); free(kds);
    free(DL); free(D); free(DU); free(B);
  }
  copy_ts1_to_ts0_interior(s); /* :906-907 */
}

void orc_diffcn_x(orc_state *s) { diff_cn(s, 0); }
void orc_diffcn_y(orc_state *s) { diff_cn(s, 1); }
void orc_diffcn_z(orc_state *s) { diff_cn(s, 2); }

/* src/Diffusion.F90:151-216 with FAST_SUBGRID off */
void orc_diffuse_horz(orc_state *s, int itoggle) {
  s->imin = 1; s->imax = s->nx;
  s->jmin = 1; s->jmax = s->ny;
  s->kmin = 1; s->kmax = s->nz;
  if (s->use_cn) {
    if (itoggle % 2 == 0) { orc_diffcn_x(s); orc_diffcn_y(s); }
    else { orc_diffcn_y(s); orc_diffcn_x(s); }
  } else {
    if (itoggle % 2 == 0) { orc_diff_x(s); orc_diff_y(s); }
    else { orc_diff_y(s); orc_diff_x(s); }
  }
}

/* src/Diffusion.F90:232-247 */
void orc_diffuse_vert(orc_state *s) {
  if (s->use_cn) orc_diffcn_z(s);
  else orc_diff_z(s);
}

/* src/Ash3d.F90:376-378 */
void orc_step_glue(orc_state *s) {
  size_t m = NCELL(s) * (size_t)s->ns;
  memcpy(s->concen_pd, s->concen_pd + m, sizeof(double) * m);
  memset(s->concen_pd + m, 0, sizeof(double) * m);
}

/* src/Ash3d.F90:428-431: concen(iv,jv,1:nz+1,n,ts0) += dt*SourceNodeFlux(k,n).
 * flux is SourceNodeFlux(1:nz+1,1:ns) in kg/km3/h. */
void orc_source_column_add(orc_state *s, int iv, int jv, const double *flux) {
  int k, n;
  for (n = 1; n <= s->ns; n++)
    for (k = 1; k <= s->nz + 1; k++) {
      size_t o = PC(s, iv, jv, k, n, TS0);
      s->concen_pd[o] = s->concen_pd[o] + s->dt * flux[(size_t)(k - 1) + (size_t)(s->nz + 1) * (size_t)(n - 1)];
    }
}

/* src/Ash3d.F90:447-460: umbrella source.  On levels ibase..itop each
   bin's 3x3 patch around the vent takes AvgCon_Umbrella
   (src/Source_Umbrella.f90:539-560) of itself; then dt*flux_umb is added to
   the patch on levels 1..nz+1.  avg_stencil(3,3,nz), flux_umb(3,3,nz+1,ns),
   Fortran order; bins 1..n_gs_max. */
void orc_umbrella_source_add(orc_state *s, int iv, int jv, int ibase, int itop, int n_gs_max,
                             const double *avg_stencil, const double *flux_umb) {
  int i, j, k, n;
  for (k = ibase; k <= itop; k++)
    for (n = 1; n <= n_gs_max; n++) {
      double avg = 0.0;
      for (i = 1; i <= 3; i++)
        for (j = 1; j <= 3; j++)
          avg = avg + s->concen_pd[PC(s, iv - 2 + i, jv - 2 + j, k, n, TS0)] *
                          avg_stencil[(size_t)(i - 1) + 3 * (size_t)(j - 1) + 9 * (size_t)(k - 1)];
      for (i = 1; i <= 3; i++)
        for (j = 1; j <= 3; j++) s->concen_pd[PC(s, iv - 2 + i, jv - 2 + j, k, n, TS0)] = avg;
    }
  for (n = 1; n <= n_gs_max; n++)
    for (k = 1; k <= s->nz + 1; k++)
      for (j = 1; j <= 3; j++)
        for (i = 1; i <= 3; i++) {
          size_t o = PC(s, iv - 2 + i, jv - 2 + j, k, n, TS0);
          s->concen_pd[o] = s->concen_pd[o] + s->dt * flux_umb[(size_t)(i - 1) + 3 * (size_t)(j - 1) +
                                                               9 * (size_t)(k - 1) +
                                                               9 * (size_t)(s->nz + 1) * (size_t)(n - 1)];
        }
}

/* src/MesoInterpolater.f90:238-243: vx_pd, vy_pd += uvx_pd, uvy_pd on
   levels ibase..itop (one value per column: umbrella_winds sets every
   level alike, src/Source_Umbrella.f90:381-385) */
void orc_umbrella_winds_add(orc_state *s, int ibase, int itop, const double *uvx, const double *uvy) {
  int i, j, k;
  for (k = ibase; k <= itop; k++)
    for (j = 1; j <= s->ny; j++)
      for (i = 1; i <= s->nx; i++) {
        size_t o = P3(s, i, j, k), q = (size_t)(i - 1) + (size_t)s->nx * (size_t)(j - 1);
        s->vx_pd[o] = s->vx_pd[o] + uvx[q];
        s->vy_pd[o] = s->vy_pd[o] + uvy[q];
      }
}

/* src/Ash3d.F90:491-510: the hot-path call sequence of one time step */
int orc_transport_step(orc_state *s, int itime, int use_diffusion) {
  if (orc_set_bc(s, 1)) return 1;
  orc_advect_horz(s, itime);
  orc_advect_vert(s);
  if (use_diffusion) {
    orc_set_bc(s, 2);
    orc_diffuse_horz(s, itime);
    orc_diffuse_vert(s);
  }
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* threads of the OpenMP row loops from now on (the CPU baseline's
 * one-thread figure) */
void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---- wind time interpolation : src/MesoInterpolater.f90:203-230 -------- */
/* v_pd = (real(last,8) + real(next - last,8) * frac) * MPS_2_KMPHR; the
 * difference is taken in FP32 (both operands are _sp arrays). */
void orc_met_interp(orc_state *s, const orc_met_step *last, const orc_met_step *next, double frac,
                    int use_vf) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const double MPS_2_KMPHR = 3.6; /* src/Ash3d_VariableModules.f90:286 */
  int i, j, k, n;
  for (k = 1; k <= nz; k++)
    for (j = 1; j <= ny; j++)
      for (i = 1; i <= nx; i++) {
        size_t m = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
        size_t o = P3(s, i, j, k);
        float dvx = next->vx[m] - last->vx[m], dvy = next->vy[m] - last->vy[m], dvz = next->vz[m] - last->vz[m];
        s->vx_pd[o] = ((double)last->vx[m] + (double)dvx * frac) * MPS_2_KMPHR;
        s->vy_pd[o] = ((double)last->vy[m] + (double)dvy * frac) * MPS_2_KMPHR;
        s->vz_pd[o] = ((double)last->vz[m] + (double)dvz * frac) * MPS_2_KMPHR;
      }
  if (!use_vf) return;
  for (n = 1; n <= s->ns; n++)
    for (k = 1; k <= nz; k++)
      for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
          size_t m = (size_t)(i - 1) +
                     (size_t)nx * ((size_t)(j - 1) + (size_t)ny * ((size_t)(k - 1) + (size_t)nz * (size_t)(n - 1)));
          float dv = next->vf[m] - last->vf[m];
          s->vf_pd[PF(s, i, j, k, n)] = ((double)last->vf[m] + (double)dv * frac) * MPS_2_KMPHR;
        }
}

/* ---- Adjust_DT : src/Adjust_dt.f90:89-333 ------------------------------ */
double orc_adjust_dt(const orc_state *s, const orc_dt_params *p, const orc_met_step *next, double *diag) {
  const int nx = s->nx, ny = s->ny, nz = s->nz, ns = s->ns;
  const double MPS_2_KMPHR = 3.6;
  double vxmax_dx = 0.0, vymax_dy = 0.0, vzmax_dz = 0.0, time_advect, time_diffuse, dt_tmp, dt;
  int i, j, k, n;
  if (p->use_met_next) { /* :96-136 */
    if (!p->is_latlon) {
      float mx = 0.0f, my = 0.0f;
      for (size_t m = 0; m < (size_t)nx * ny * nz; m++) {
        if (fabsf(next->vx[m]) > mx) mx = fabsf(next->vx[m]);
        if (fabsf(next->vy[m]) > my) my = fabsf(next->vy[m]);
      }
      vxmax_dx = (double)mx * MPS_2_KMPHR / p->dx;
      vymax_dy = (double)my * MPS_2_KMPHR / p->dy;
    }
    for (i = 1; i <= nx; i++)
      for (j = 1; j <= ny; j++)
        for (k = 1; k <= nz; k++) {
          size_t m = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
          double t3;
          float vfm = 0.0f;
          if (p->is_latlon) {
            double ms = fmin(s->sigma_nx_pd[P3(s, i, j, k)], s->sigma_nx_pd[P3(s, i + 1, j, k)]);
            double t1 = (double)fabsf(next->vx[m]) * MPS_2_KMPHR * ms / s->kappa_pd[P3(s, i, j, k)];
            double t2;
            if (t1 > vxmax_dx) vxmax_dx = t1;
            ms = fmin(s->sigma_ny_pd[P3(s, i, j, k)], s->sigma_ny_pd[P3(s, i, j + 1, k)]);
            t2 = (double)fabsf(next->vy[m]) * MPS_2_KMPHR * ms / s->kappa_pd[P3(s, i, j, k)];
            if (t2 > vymax_dy) vymax_dy = t2;
          }
          for (n = 0; n < ns; n++) {
            float a = fabsf(next->vf[m + (size_t)nx * ny * nz * (size_t)n]);
            if (a > vfm) vfm = a;
          }
          t3 = (double)(fabsf(next->vz[m]) + vfm) * MPS_2_KMPHR / p->dz[k - 1];
          if (t3 > vzmax_dz) vzmax_dz = t3;
        }
  } else { /* :137-172 */
    if (!p->is_latlon) {
      double mx = 0.0, my = 0.0;
      for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
          for (i = 1; i <= nx; i++) {
            if (fabs(s->vx_pd[P3(s, i, j, k)]) > mx) mx = fabs(s->vx_pd[P3(s, i, j, k)]);
            if (fabs(s->vy_pd[P3(s, i, j, k)]) > my) my = fabs(s->vy_pd[P3(s, i, j, k)]);
          }
      vxmax_dx = mx / p->dx;
      vymax_dy = my / p->dy;
    }
    for (i = 1; i <= nx; i++)
      for (j = 1; j <= ny; j++)
        for (k = 1; k <= nz; k++) {
          double t3, vfm = 0.0;
          if (p->is_latlon) {
            double ms = fmin(s->sigma_nx_pd[P3(s, i, j, k)], s->sigma_nx_pd[P3(s, i + 1, j, k)]);
            double t1 = fabs(s->vx_pd[P3(s, i, j, k)]) * ms / s->kappa_pd[P3(s, i, j, k)], t2;
            if (t1 > vxmax_dx) vxmax_dx = t1;
            ms = fmin(s->sigma_ny_pd[P3(s, i, j, k)], s->sigma_ny_pd[P3(s, i, j + 1, k)]);
            t2 = fabs(s->vy_pd[P3(s, i, j, k)]) * ms / s->kappa_pd[P3(s, i, j, k)];
            if (t2 > vymax_dy) vymax_dy = t2;
          }
          for (n = 1; n <= ns; n++) {
            double a = fabs(s->vf_pd[PF(s, i, j, k, n)]);
            if (a > vfm) vfm = a;
          }
          t3 = (fabs(s->vz_pd[P3(s, i, j, k)]) + vfm) / p->dz[k - 1];
          if (t3 > vzmax_dz) vzmax_dz = t3;
        }
  }
  if (p->n_gs_aloft == 0) vzmax_dz = 0.0; /* :176 */
  time_advect = 1.0 / fmax(fmax(vxmax_dx, vymax_dy), vzmax_dz);
  time_diffuse = p->dt_max;
  if (p->use_diffusion) { /* :182-246 (computed on every call here) */
    for (i = 1; i <= nx; i++)
      for (j = 1; j <= ny; j++)
        for (k = 1; k <= nz; k++) {
          double kap = s->kappa_pd[P3(s, i, j, k)], ms, t1, t2, t3, r;
          ms = fmin(s->sigma_nx_pd[P3(s, i, j, k)], s->sigma_nx_pd[P3(s, i + 1, j, k)]);
          r = kap / ms;
          t1 = (r * r) / s->kx[PK(s, i, j, k)];
          ms = fmin(s->sigma_ny_pd[P3(s, i, j, k)], s->sigma_ny_pd[P3(s, i, j + 1, k)]);
          r = kap / ms;
          t2 = (r * r) / s->ky[PK(s, i, j, k)];
          ms = fmin(s->sigma_nz_pd[P3(s, i, j, k)], s->sigma_nz_pd[P3(s, i, j, k + 1)]);
          r = kap / ms;
          t3 = (r * r) / s->kz[PK(s, i, j, k)];
          if (time_diffuse > t1 + t2 + t3) time_diffuse = t1 + t2 + t3;
        }
    time_diffuse = time_diffuse * (p->use_cn ? p->imp_dt_fac : 0.5);
  }
  dt_tmp = fmin(time_advect, time_diffuse); /* :296-323 */
  if (dt_tmp < p->dt_min) dt = p->dt_min;
  else if (dt_tmp > p->dt_max) dt = p->dt_max;
  else dt = fmin(1.0, p->cfl) * dt_tmp;
  dt = p->dt_min * (double)(int)(dt / p->dt_min);
  if ((p->next_write_time - p->time) > p->eps_small && (p->next_write_time - p->time) < dt)
    dt = p->next_write_time - p->time;
  if (p->time + dt > p->simtime) dt = p->simtime - p->time;
  if (diag) {
    diag[0] = vxmax_dx; diag[1] = vymax_dy; diag[2] = vzmax_dz; diag[3] = time_diffuse;
  }
  return dt;
}

/* ---- get_minmax_index : src/Calc_Mesh.f90:432-542 ---------------------- */
static double slab_max(const orc_state *s, int dir, int idx) {
  double m = -INFINITY;
  int a, b, n;
  const int A = dir == 0 ? s->ny : s->nx, B = dir == 2 ? s->ny : s->nz;
  for (n = 1; n <= s->ns; n++)
    for (b = 1; b <= B; b++)
      for (a = 1; a <= A; a++) {
        int i = dir == 0 ? idx : a, j = dir == 0 ? a : (dir == 1 ? idx : b), k = dir == 2 ? idx : b;
        double v = s->concen_pd[PC(s, i, j, k, n, TS0)];
        if (v > m) m = v;
      }
  return m;
}

int orc_minmax_index(const orc_state *s, int ivent, int jvent, int kvent, double thresh, int32_t b[6]) {
  const int N[3] = {s->nx, s->ny, s->nz}, vent[3] = {ivent, jvent, kvent};
  int d, i;
  for (d = 0; d < 3; d++) {
    int lo = 1, hi = N[d];
    for (i = 2; i <= vent[d] - 1; i++) {
      if (slab_max(s, d, i) < thresh) lo = i;
      else break;
    }
    for (i = N[d]; i >= lo + 2; i--) {
      if (slab_max(s, d, i) < thresh) hi = i;
      else break;
    }
    if (d == 0 && s->is_periodic && (lo == 1 || hi == N[0])) { lo = 1; hi = N[0]; }
    b[2 * d] = lo;
    b[2 * d + 1] = hi;
  }
  if (b[5] < b[4]) return 1;
  for (d = 0; d < 3; d++)
    if (b[2 * d + 1] - b[2 * d] < 2) {
      b[2 * d + 1] = b[2 * d + 1] + 1 < N[d] ? b[2 * d + 1] + 1 : N[d];
      b[2 * d] = b[2 * d + 1] - 2;
    }
  return 0;
}

/* src/Ash3d.F90:491-510 built with FAST_SUBGRID: AdvectHorz takes the box of
 * get_minmax_index (src/AdvectionHorz.f90:60-66), AdvectVert keeps it, and
 * DiffuseHorz grows it by one cell (src/Diffusion.F90:168-175). */
int orc_transport_step_subgrid(orc_state *s, int itime, int use_diffusion, int ivent, int jvent, int kvent,
                               double thresh, int32_t bounds[6]) {
  int32_t b[6];
  if (orc_set_bc(s, 1)) return 1;
  if (orc_minmax_index(s, ivent, jvent, kvent, thresh, b)) return 2;
  if (bounds) memcpy(bounds, b, sizeof b);
  s->imin = b[0]; s->imax = b[1]; s->jmin = b[2]; s->jmax = b[3]; s->kmin = b[4]; s->kmax = b[5];
  if (itime % 2 == 0) { orc_advect_y(s); orc_advect_x(s); }
  else { orc_advect_x(s); orc_advect_y(s); }
  orc_advect_z(s);
  if (use_diffusion) {
    orc_set_bc(s, 2);
    s->imin = s->imin - 1 > 1 ? s->imin - 1 : 1;
    s->imax = s->imax + 1 < s->nx ? s->imax + 1 : s->nx;
    s->jmin = s->jmin - 1 > 1 ? s->jmin - 1 : 1;
    s->jmax = s->jmax + 1 < s->ny ? s->jmax + 1 : s->ny;
    s->kmin = s->kmin - 1 > 1 ? s->kmin - 1 : 1;
    s->kmax = s->kmax + 1 < s->nz ? s->kmax + 1 : s->nz;
    if (s->use_cn) {
      if (itime % 2 == 0) { orc_diffcn_x(s); orc_diffcn_y(s); } else { orc_diffcn_y(s); orc_diffcn_x(s); }
      orc_diffcn_z(s);
    } else {
      if (itime % 2 == 0) { orc_diff_x(s); orc_diff_y(s); } else { orc_diff_y(s); orc_diff_x(s); }
      orc_diff_z(s);
    }
  }
  return 0;
}

/* ConcentrationCalculator (src/Output_Vars.f90:1026-1137), on ts1, and
 * AshThicknessCalculator (:855-893), loop for loop. */
void orc_products(const orc_state *s, const double *dz, const double *zcc, const double prm[4], double *out,
                  int32_t *masks, double areas[2]) {
  const int nx = s->nx, ny = s->ny, nz = s->nz, ns = s->ns;
  const long long nxy = (long long)nx * ny;
  const double CLOUDCON_THRESH = prm[0], CLOUDLOAD_THRESH = prm[1], DEPO_THRESH = prm[2], DepositDensity = prm[3];
  const double KG_2_MG = 1.0e6, KM3_2_M3 = 1.0e9, KM2_2_M2 = 1.0e6, M_2_MM = 1.0e3;
  double *CloudLoad = out, *MaxConcentration = out + nxy, *MaxHeight = out + 2 * nxy, *MinHeight = out + 3 * nxy;
  double *DepositThickness = out + 4 * nxy;
  int32_t *Mask_Cloud = masks, *Mask_Deposit = masks + nxy;
  double *TotalConcentration = (double *)malloc(sizeof(double) * (size_t)nz);
  double CloudArea = 0.0, DepositAreaCovered = 0.0;
  long long t;
  int i, j, k, n;
  for (t = 0; t < nxy; t++) {
    CloudLoad[t] = 0.0;
    MaxConcentration[t] = 0.0;
    MaxHeight[t] = 0.0;
    MinHeight[t] = 100.0;
    Mask_Cloud[t] = 0;
  }
  if (ns > 0) {
    for (i = s->imin; i <= s->imax; i++) {
      for (j = s->jmin; j <= s->jmax; j++) {
        const long long c = (long long)(i - 1) + (long long)nx * (j - 1);
        const double CellArea = s->sigma_nz_pd[P3(s, i, j, 1)];
        double MaxTotalConcentration;
        for (k = 1; k <= nz; k++) TotalConcentration[k - 1] = 0.0;
        for (k = 1; k <= nz; k++) {
          double sum = 0.0;
          for (n = 1; n <= ns; n++) sum += s->concen_pd[PC(s, i, j, k, n, TS1)];
          CloudLoad[c] = CloudLoad[c] + sum * dz[k - 1] / 1.0e3;
          TotalConcentration[k - 1] = sum;
          if (TotalConcentration[k - 1] >= CLOUDCON_THRESH) MaxHeight[c] = zcc[k - 1] + 0.5 * dz[k - 1];
        }
        for (k = nz; k >= 1; k--)
          if (TotalConcentration[k - 1] <= CLOUDCON_THRESH) MinHeight[c] = zcc[k - 1] - 0.5 * dz[k - 1];
        MaxTotalConcentration = TotalConcentration[0];
        for (k = 2; k <= nz; k++)
          if (TotalConcentration[k - 1] > MaxTotalConcentration) MaxTotalConcentration = TotalConcentration[k - 1];
        Mask_Cloud[c] = CloudLoad[c] >= CLOUDLOAD_THRESH;
        if (Mask_Cloud[c]) {
          if (MinHeight[c] < 0.0 && MaxHeight[c] > 0.0) MinHeight[c] = 0.0;
          MinHeight[c] = MaxHeight[c] < MinHeight[c] ? MaxHeight[c] : MinHeight[c];
          CloudArea = CloudArea + CellArea;
          MaxConcentration[c] = MaxTotalConcentration * KG_2_MG / KM3_2_M3;
        } else {
          CloudLoad[c] = 0.0;
          MaxConcentration[c] = 0.0;
          MaxHeight[c] = 0.0;
          MinHeight[c] = 0.0;
        }
      }
    }
  }
  free(TotalConcentration);
  for (t = 0; t < nxy; t++) {
    Mask_Deposit[t] = 0;
    DepositThickness[t] = 0.0;
  }
  if (ns > 0) {
    for (i = 1; i <= nx; i++) {
      for (j = 1; j <= ny; j++) {
        const long long c = (long long)(i - 1) + (long long)nx * (j - 1);
        double sum = 0.0;
        for (n = 1; n <= ns; n++) sum += s->deposit_granularity[c + nxy * (n - 1)];
        DepositThickness[c] = sum * dz[0] / KM2_2_M2 / DepositDensity * M_2_MM;
        if (DepositThickness[c] > DEPO_THRESH) {
          Mask_Deposit[c] = 1;
          DepositAreaCovered = DepositAreaCovered + s->sigma_nz_pd[P3(s, i, j, 1)];
        }
      }
    }
  }
  areas[0] = CloudArea;
  areas[1] = DepositAreaCovered;
}
