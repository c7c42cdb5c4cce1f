/*
 * ash3d_oracle.h -- CPU oracle for the Ash3d transport step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its
 * C ABI, the Fortran shims or the Python host mirror) may link, load or call
 * this code.  It is used by tests/ (as the checker), by the smoke entry
 * point (as the checker) and by bench.py's cpu_baseline leg (as the timed
 * CPU port of the reference).
 *
 * The oracle is a plain-C restatement of the reference Fortran hot path:
 *   Set_BC                 src/Set_BC.f90:15-126
 *   advect_x / advect_y    src/AdvectionHorz_DCU.f90:68-305 / :322-556
 *   AdvectHorz             src/AdvectionHorz.f90:47-85 (FAST_SUBGRID off)
 *   advect_z / advect_sigma src/AdvectionVert_DCU.f90:100-323 / :341-564
 *   AdvectVert             src/AdvectionVert_DCU.f90:64-82
 *   diff_x/y/z             src/Diffusion.F90:266-388 / :406-518 / :536-650
 *   diffCN_x/y/z           src/Diffusion.F90:671-911 / :932-1167 / :1188-1428
 *   DiffuseHorz/Vert       src/Diffusion.F90:151-216 / :232-247 (FAST_SUBGRID off)
 *   dgtsv                  LAPACK dgtsv (third-party, not vendored by the
 *                          reference; restated from LAPACK's published
 *                          algorithm, Gaussian elimination with partial
 *                          pivoting)
 *   step glue              src/Ash3d.F90:376-378, :428-431, :491-510
 *   wind time interpolation src/MesoInterpolater.f90:203-230
 *   Adjust_DT              src/Adjust_dt.f90:89-333
 *   get_minmax_index       src/Calc_Mesh.f90:432-542 (and the FAST_SUBGRID step)
 *   ConcentrationCalculator src/Output_Vars.f90:1026-1137
 *   AshThicknessCalculator src/Output_Vars.f90:855-893
 *
 * Arrays use the reference's Fortran column-major layouts and lower bounds
 * (src/Ash3d_VariableModules.f90:918-946, src/Diffusion.F90:95-101):
 *   concen_pd(-1:nx+2,-1:ny+2,-1:nz+2,1:ns,0:1)
 *   vx/vy/vz/vh_pd, kappa_pd, sigma_n{x,y,z}_pd (-1:nx+2,-1:ny+2,-1:nz+2)
 *   vf_pd(-1:nx+2,-1:ny+2,-1:nz+2,1:ns)
 *   kx, ky, kz (0:nx+1,0:ny+1,0:nz+1)
 *   DelDxonD_cc, DelDyonD_cc (-1:nx+2,-1:ny+2)
 *   outflow_xz{1,2}_pd(-1:nx+2,-1:nz+2,1:ns)
 *   outflow_yz{1,2}_pd(-1:ny+2,-1:nz+2,1:ns)
 *   outflow_xy{1,2}_pd(-1:nx+2,-1:ny+2,1:ns)
 *   DepositGranularity(1:nx,1:ny,1:ns)
 */
#ifndef ASH3D_ORACLE_H
#define ASH3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Limiter ids (compile-time LIMITER in the reference makefile). */
enum {
  ORC_LIM_NONE = 0,
  ORC_LIM_LAXWEN = 1,
  ORC_LIM_BW = 2,
  ORC_LIM_FROMM = 3,
  ORC_LIM_MINMOD = 4,
  ORC_LIM_SUPERBEE = 5,
  ORC_LIM_MC = 6
};

typedef struct {
  int32_t nx, ny, nz, ns;
  int32_t is_periodic;   /* mesh::IsPeriodic */
  int32_t zscaling_id;   /* mesh::ZScaling_ID */
  int32_t limiter;       /* ORC_LIM_* */
  int32_t use_cn;        /* global_param::useCN (CRANKNIC) */
  int32_t imin, imax, jmin, jmax, kmin, kmax; /* solution::imin..kmax */
  double dt;             /* time_data::dt (hours); dt_ip == dt */
  double imp_fac;        /* Diffusion::Imp_fac */
  const int32_t *is_aloft; /* solution::IsAloft as 0/1, length ns */
  double *concen_pd;
  double *vx_pd, *vy_pd, *vz_pd, *vh_pd, *vf_pd;
  double *kappa_pd, *sigma_nx_pd, *sigma_ny_pd, *sigma_nz_pd;
  double *kx, *ky, *kz;
  double *deldxond_cc, *deldyond_cc;
  double *outflow_xz1_pd, *outflow_xz2_pd;
  double *outflow_yz1_pd, *outflow_yz2_pd;
  double *outflow_xy1_pd, *outflow_xy2_pd;
  double *deposit_granularity;
  int32_t last_info;     /* last nonzero dgtsv INFO seen (0 if none) */
} orc_state;

int orc_set_bc(orc_state *s, int bc_code);
void orc_advect_x(orc_state *s);
void orc_advect_y(orc_state *s);
void orc_advect_z(orc_state *s);
void orc_advect_horz(orc_state *s, int itoggle);
void orc_advect_vert(orc_state *s);
void orc_diff_x(orc_state *s);
void orc_diff_y(orc_state *s);
void orc_diff_z(orc_state *s);
void orc_diffcn_x(orc_state *s);
void orc_diffcn_y(orc_state *s);
void orc_diffcn_z(orc_state *s);
void orc_diffuse_horz(orc_state *s, int itoggle);
void orc_diffuse_vert(orc_state *s);
void orc_step_glue(orc_state *s);
void orc_source_column_add(orc_state *s, int ivent, int jvent,
                           const double *flux /* (1:nz+1,1:ns) */);
void orc_umbrella_source_add(orc_state *s, int iv, int jv, int ibase, int itop, int n_gs_max,
                             const double *avg_stencil, const double *flux_umb);
void orc_umbrella_winds_add(orc_state *s, int ibase, int itop, const double *uvx, const double *uvy);
int orc_transport_step(orc_state *s, int itime, int use_diffusion);

/* Met-step winds in m/s, FP32, unpadded (1:nx,1:ny,1:nz) and vf
 * (1:nx,1:ny,1:nz,1:ns): the wind_grid module's v*_meso_{last,next}_step_sp. */
typedef struct {
  const float *vx, *vy, *vz, *vf;
} orc_met_step;

/* vx/vy/vz_pd(1:nx,1:ny,1:nz) and, if use_vf, vf_pd(1:nx,1:ny,1:nz,:) from
 * the bracketing met steps at Interval_Frac = frac. */
void orc_met_interp(orc_state *s, const orc_met_step *last, const orc_met_step *next, double frac,
                    int use_vf);

typedef struct {
  int32_t is_latlon;      /* mesh::IsLatLon */
  int32_t use_met_next;   /* Adjust_DT(mesostep) present: the met-step (FAST_DT) branch */
  int32_t n_gs_aloft;     /* Tephra::n_gs_aloft */
  int32_t use_diffusion, use_cn;
  double dx, dy;          /* mesh::dx, dy (km) */
  const double *dz;       /* dz_vec_pd(1:nz) */
  double imp_dt_fac, cfl, dt_min, dt_max, eps_small;
  double time, next_write_time, simtime; /* hours */
} orc_dt_params;

/* Adjust_DT: returns dt (hours); diag[0..3] = vxmax_dx, vymax_dy, vzmax_dz,
 * time_diffuse. */
double orc_adjust_dt(const orc_state *s, const orc_dt_params *p, const orc_met_step *next, double *diag);

/* get_minmax_index with the vent (ivent,jvent,kvent) and CLOUDCON_GRID_THRESH
 * = thresh; returns 0, or 1 for kmax < kmin (the reference stops). */
int orc_minmax_index(const orc_state *s, int ivent, int jvent, int kvent, double thresh, int32_t bounds[6]);

/* ash3d step with FAST_SUBGRID (src/Ash3d.F90:491-510, src/AdvectionHorz.f90:60-66,
 * src/Diffusion.F90:168-175); returns 2 if get_minmax_index stops. */
int orc_transport_step_subgrid(orc_state *s, int itime, int use_diffusion, int ivent, int jvent, int kvent,
                               double thresh, int32_t bounds[6]);

/* ConcentrationCalculator and AshThicknessCalculator.  dz, zcc: dz_vec_pd,
 * z_cc_pd (1:nz); prm = {CLOUDCON_THRESH, CLOUDLOAD_THRESH, DEPO_THRESH,
 * DepositDensity}; columns imin..imax x jmin..jmax of the state.  out:
 * CloudLoad, MaxConcentration, MaxHeight, MinHeight, DepositThickness
 * (5 x nx*ny, x fastest); masks: Mask_Cloud, Mask_Deposit (2 x nx*ny);
 * areas: CloudArea, DepositAreaCovered. */
void orc_products(const orc_state *s, const double *dz, const double *zcc, const double prm[4], double *out,
                  int32_t *masks, double areas[2]);

void orc_dgtsv(int n, int nrhs, double *dl, double *d, double *du, double *b,
               int ldb, int *info);
double orc_limiter_phi(int limiter, double theta);
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
