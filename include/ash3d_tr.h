/*
 * ash3d_tr.h -- C ABI of the MI355X transport step for Ash3d.
 *
 * Drop-in boundary for the reference's hot path.  The reference routines
 * take no arguments and work on Fortran module arrays; a Fortran shim with
 * the reference's module and routine names passes c_loc() of those arrays
 * here once (ash3d_tr_create) and then calls the per-routine entry points.
 * All arrays keep the reference's Fortran column-major layouts and lower
 * bounds (src/Ash3d_VariableModules.f90:918-946):
 *
 *   concen_pd(-1:nx+2,-1:ny+2,-1:nz+2,1:ns,0:1)      real(8)
 *   vx_pd, vy_pd, vz_pd, vh_pd (-1:nx+2,-1:ny+2,-1:nz+2)
 *   vf_pd(-1:nx+2,-1:ny+2,-1:nz+2,1:ns)
 *   kappa_pd, sigma_nx_pd, sigma_ny_pd, sigma_nz_pd (-1:nx+2,-1:ny+2,-1:nz+2)
 *   kx, ky, kz (0:nx+1,0:ny+1,0:nz+1)                  src/Diffusion.F90:95-101
 *   DelDxonD_cc, DelDyonD_cc (-1:nx+2,-1:ny+2)         src/Topography.f90:412-413
 *   outflow_xz{1,2}_pd(-1:nx+2,-1:nz+2,1:ns)
 *   outflow_yz{1,2}_pd(-1:ny+2,-1:nz+2,1:ns)
 *   outflow_xy{1,2}_pd(-1:nx+2,-1:ny+2,1:ns)
 *   DepositGranularity(1:nx,1:ny,1:ns)
 *
 * Units are the reference's: km, hours, km/h, km2/h, kg/km3.  IsAloft is a
 * Fortran LOGICAL array, which is not interoperable; the shim passes it as
 * an int32 0/1 mask of length ns.  Bounds are imin,imax,jmin,jmax,kmin,kmax
 * (solution module, src/Ash3d_VariableModules.f90:885-887).
 *
 * Device residency: the context owns device mirrors of every array.  Between
 * ash3d_tr_sync_from_host and ash3d_tr_sync_to_host the device copy is the
 * authoritative one; the host must sync before it reads or writes any array
 * (outputs, mass totals, source insertion, wind updates).  Calls on one
 * context are asynchronous on the context's HIP stream and must not overlap
 * (the reference calls these from one thread); several contexts may share a
 * device.  Every entry point returns 0 on success, or an ASH3D_TR_E* code
 * with a message available from ash3d_tr_last_error().
 */
#ifndef ASH3D_TR_H
#define ASH3D_TR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASH3D_TR_ABI_VERSION 1

enum {
  ASH3D_TR_OK = 0,
  ASH3D_TR_EINVAL = 1,    /* bad argument, bounds, shape or flag */
  ASH3D_TR_EDEVICE = 2,   /* HIP runtime error */
  ASH3D_TR_ENOMEM = 3,    /* device allocation failed */
  ASH3D_TR_EBC = 4,       /* unknown Set_BC code (src/Set_BC.f90:119-124) */
  ASH3D_TR_ESINGULAR = 5, /* a tridiagonal pivot was exactly zero (dgtsv INFO>0,
                             ignored by the reference, reported here) */
  ASH3D_TR_ENOLIB = 6     /* no usable HIP device */
};

/* Limiters, chosen at compile time in the reference (src/makefile:126). */
enum {
  ASH3D_LIM_NONE = 0,
  ASH3D_LIM_LAXWEN = 1,
  ASH3D_LIM_BW = 2,
  ASH3D_LIM_FROMM = 3,
  ASH3D_LIM_MINMOD = 4,
  ASH3D_LIM_SUPERBEE = 5,
  ASH3D_LIM_MC = 6
};

/* Array groups for the sync masks. */
enum {
  ASH3D_SYNC_CONCEN = 1u << 0,   /* concen_pd, both time levels */
  ASH3D_SYNC_WINDS = 1u << 1,    /* vx_pd, vy_pd, vz_pd, vh_pd */
  ASH3D_SYNC_VF = 1u << 2,       /* vf_pd */
  ASH3D_SYNC_GEOMETRY = 1u << 3, /* kappa_pd, sigma_nx/ny/nz_pd */
  ASH3D_SYNC_DIFFK = 1u << 4,    /* kx, ky, kz */
  ASH3D_SYNC_TOPO = 1u << 5,     /* DelDxonD_cc, DelDyonD_cc */
  ASH3D_SYNC_OUTFLOW = 1u << 6,  /* the six outflow_*_pd arrays */
  ASH3D_SYNC_DEPOSIT = 1u << 7,  /* DepositGranularity */
  ASH3D_SYNC_ALL = 0xffu
};

typedef struct ash3d_tr_desc {
  int32_t abi_version; /* ASH3D_TR_ABI_VERSION */
  int32_t nx, ny, nz, ns; /* nxmax, nymax, nzmax, nsmax */
  int32_t is_periodic;    /* mesh::IsPeriodic */
  int32_t zscaling_id;    /* mesh::ZScaling_ID: 2 = advect_sigma + topo terms */
  int32_t limiter;        /* ASH3D_LIM_* (reference default: SUPERBEE) */
  int32_t device;         /* HIP device ordinal */
  /* Host arrays (may be NULL: the device copy then starts at zero, except
   * the geometry, which starts at 1.0 as in Allocate_mesh). */
  double *concen_pd;
  double *vx_pd, *vy_pd, *vz_pd, *vh_pd, *vf_pd;
  double *kappa_pd, *sigma_nx_pd, *sigma_ny_pd, *sigma_nz_pd;
  double *kx, *ky, *kz;
  double *deldxond_cc, *deldyond_cc;
  double *outflow_xz1_pd, *outflow_xz2_pd;
  double *outflow_yz1_pd, *outflow_yz2_pd;
  double *outflow_xy1_pd, *outflow_xy2_pd;
  double *deposit_granularity;
} ash3d_tr_desc;

typedef struct ash3d_tr_ctx ash3d_tr_ctx;

/* Allocate device mirrors and upload every non-NULL host array.
 * Replaces the allocation glue of src/Alloc_Arrays.f90:48-53 for the
 * arrays above (the host arrays stay owned by the caller). */
int ash3d_tr_create(const ash3d_tr_desc *desc, ash3d_tr_ctx **ctx);
int ash3d_tr_destroy(ash3d_tr_ctx *ctx);

/* Host <-> device copies of the array groups in mask (ASH3D_SYNC_*).
 * sync_to_host blocks until the data is on the host. */
int ash3d_tr_sync_from_host(ash3d_tr_ctx *ctx, uint32_t mask);
int ash3d_tr_sync_to_host(ash3d_tr_ctx *ctx, uint32_t mask);
/* Wait for all queued work of the context. */
int ash3d_tr_synchronize(ash3d_tr_ctx *ctx);

/* Set_BC(bc_code), src/Set_BC.f90:15-126. */
int ash3d_tr_set_bc(ash3d_tr_ctx *ctx, int32_t bc_code);

/* advect_x / advect_y, src/AdvectionHorz_DCU.f90:68-305 / :322-556, each
 * starting with Set_BC(1) as the reference does.  dt in hours. */
int ash3d_tr_advect_x(ash3d_tr_ctx *ctx, double dt, const int32_t *is_aloft,
                      const int32_t bounds[6]);
int ash3d_tr_advect_y(ash3d_tr_ctx *ctx, double dt, const int32_t *is_aloft,
                      const int32_t bounds[6]);
/* advect_z / advect_sigma incl. settling with vz+vh+vf(n), the deposit and
 * DepositGranularity, src/AdvectionVert_DCU.f90:100-323, :341-564. */
int ash3d_tr_advect_z(ash3d_tr_ctx *ctx, double dt, const int32_t *is_aloft,
                      const int32_t bounds[6]);

/* AdvectHorz(itoggle) with FAST_SUBGRID off, src/AdvectionHorz.f90:47-85:
 * full-grid bounds, even itoggle = y then x, odd = x then y. */
int ash3d_tr_advect_horz(ash3d_tr_ctx *ctx, int32_t itoggle, double dt,
                         const int32_t *is_aloft);
/* AdvectVert, src/AdvectionVert_DCU.f90:64-82 (full-grid bounds). */
int ash3d_tr_advect_vert(ash3d_tr_ctx *ctx, double dt, const int32_t *is_aloft);

/* One diffusion sweep along dir (0:x, 1:y, 2:z): use_cn != 0 gives
 * diffCN_{x,y,z} with theta = imp_fac (src/Diffusion.F90:671-911,
 * :932-1167, :1188-1428), use_cn == 0 gives the explicit diff_{x,y,z}
 * (src/Diffusion.F90:266-388, :406-518, :536-650). */
int ash3d_tr_diffuse_dir(ash3d_tr_ctx *ctx, int32_t dir, int32_t use_cn, double dt,
                         double imp_fac, const int32_t *is_aloft, const int32_t bounds[6]);
/* DiffuseHorz(itoggle) / DiffuseVert, src/Diffusion.F90:151-216 / :232-247,
 * with FAST_SUBGRID off (full-grid bounds). */
int ash3d_tr_diffuse_horz(ash3d_tr_ctx *ctx, int32_t itoggle, int32_t use_cn, double dt,
                          double imp_fac, const int32_t *is_aloft);
int ash3d_tr_diffuse_vert(ash3d_tr_ctx *ctx, int32_t use_cn, double dt, double imp_fac,
                          const int32_t *is_aloft);

/* The hot-path call sequence of one time step, src/Ash3d.F90:491-510:
 * Set_BC(1); AdvectHorz(itime); AdvectVert; and if use_diffusion:
 * Set_BC(2); DiffuseHorz(itime); DiffuseVert. */
int ash3d_tr_step(ash3d_tr_ctx *ctx, int32_t itime, double dt, int32_t use_diffusion,
                  int32_t use_cn, double imp_fac, const int32_t *is_aloft);
/* ash3d_tr_step whose advection sweeps run out of place through checkpoint
 * slots slot_b and slot_z (ash3d_tr_checkpoints), which keep the inputs of
 * the second horizontal and of the vertical sweep for
 * ash3d_tr_adj_step_kept.  The step's results are ash3d_tr_step's, bit for
 * bit.  No counterpart in the reference (it has no adjoint). */
int ash3d_tr_step_keep(ash3d_tr_ctx *ctx, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                       double imp_fac, const int32_t *is_aloft, int32_t slot_b, int32_t slot_z);

/* Step glue of the driver, src/Ash3d.F90:376-378:
 * concen_pd(:,:,:,:,ts0) = concen_pd(:,:,:,:,ts1); concen_pd(...,ts1) = 0. */
int ash3d_tr_step_glue(ash3d_tr_ctx *ctx);

/* Source insertion on the vent column, src/Ash3d.F90:428-431:
 * concen_pd(ivent,jvent,1:nz+1,1:ns,ts0) += dt * flux(1:nz+1,1:ns),
 * flux = SourceNodeFlux in kg/km3/h, a host array (nz+1, ns).  The array is
 * copied before the call returns; the upload and the update are
 * asynchronous (the host may run up to four insertions ahead). */
int ash3d_tr_source_column_add(ash3d_tr_ctx *ctx, int32_t ivent, int32_t jvent, double dt,
                               const double *flux);

/* Per-bin mass budget in kg, evaluated on the device from the current
 * state (deterministic FP64 reduction), as the sums of Calc_AshVol_Aloft,
 * Calc_AshVol_Deposit and Calc_AshVol_Outflow (src/Output_Vars.f90:1255-1293,
 * :1375-1404, :1419-1459) before their division by MagmaDensity*1e9:
 *   aloft[n]   = sum concen(1:nx,1:ny,1:nz,n,ts0)*kappa(1:nx,1:ny,1:nz)
 *   deposit[n] = sum DepositGranularity(:,:,n)*kappa(1:nx,1:ny,0)
 *   outflow[n] = yz1, yz2, xz1, xz2, xy2 outflows times their ghost kappa.
 * Each output is a host array of ns doubles (may be NULL).  Blocks until
 * the values are on the host. */
int ash3d_tr_mass_totals(ash3d_tr_ctx *ctx, double *aloft, double *deposit, double *outflow);

/* ash3d_tr_mass_totals without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same kernels and the same bits as ash3d_tr_mass_totals, with the
 * partial sums in a fixed device array of the library (one per device,
 * ASH3D_TR_NOALLOC_MAX_BINS * 128 * 3 doubles = 192 KiB, held until the
 * process ends) and the result written to out, a device array of 3*ns
 * doubles (aloft[0..ns), deposit[ns..2ns), outflow[2ns..3ns)), on the
 * context's stream without waiting.  Contexts with more than
 * ASH3D_TR_NOALLOC_MAX_BINS bins are rejected (ASH3D_TR_EINVAL).  Calls
 * share their device's array: calls that overlap on one device (contexts
 * on different streams) race; serialise them. */
#define ASH3D_TR_NOALLOC_MAX_BINS 64
int ash3d_tr_mass_totals_noalloc(ash3d_tr_ctx *ctx, double *out);

/* Launch tuning (no counterpart in the reference): the number of grain
 * bins one thread carries through a sweep (1, 2, 4 or 8; 3 also for the
 * advection sweeps).  More bins share
 * the per-cell terms (dt/kappa, interface speeds, line factorisations);
 * fewer keep more waves resident.  Results do not depend on it. */
enum {
  ASH3D_TUNE_NB_ADVECT_X = 0, /* used when a row is too long for one workgroup */
  ASH3D_TUNE_NB_ADVECT_Y = 1,
  ASH3D_TUNE_NB_ADVECT_Z = 2,
  ASH3D_TUNE_NB_DIFFUSE_X = 3,
  ASH3D_TUNE_NB_DIFFUSE_Y = 4,
  ASH3D_TUNE_NB_DIFFUSE_Z = 5,
  ASH3D_TUNE_X_TILE = 6, /* cells per LDS tile of the x-line solves: 8, 16 or 32 */
  ASH3D_TUNE_LINE_SOLVER = 7, /* ASH3D_LINE_SOLVER_* below */
  ASH3D_TUNE_PART_BINS = 8, /* bins per workgroup of the partitioned solves (0: all) */
  ASH3D_TUNE_PART_SHAPE_X = 9, /* tile shape of the partitioned solves (0 or 1; defaults 1, 1, 0), x lines */
  ASH3D_TUNE_PART_SHAPE_Y = 10,
  ASH3D_TUNE_PART_SHAPE_Z = 11,
  /* streamed y/z advection: 0 cuts the lines of a workgroup into segments;
   * 1..4 gives each wave of a workgroup one bin group of the same whole
   * lines, so the per-cell fields are read once for all of them */
  ASH3D_TUNE_ADV_GROUPS_Y = 12,
  ASH3D_TUNE_ADV_GROUPS_Z = 13,
  /* x advection rows: 1 (default) rows in wavefront registers: one
   * wavefront per row for nx <= 380 (nx <= 252 with ZScaling_ID 2; the
   * forward sweep at most 64 bins, the adjoint even nx only), longer even
   * rows in up to 8 segments of one wavefront each (380 cells a segment
   * in the forward sweep, 248 in the adjoint); other rows as with 0: one
   * workgroup per row staged through LDS (rows of up to 2044 cells,
   * longer ones streamed per lane) */
  ASH3D_TUNE_ADV_X_ROWS = 14,
  /* longest active line (3..2048 cells) that the partitioned solves take,
   * for x lines (contiguous in memory; default 2048) and for y and z lines
   * (strided; default 512); lines beyond 512 cells run over several waves
   * per line (SPIKE coupling of the waves' parts), longer lines in dgtsv
   * order */
  ASH3D_TUNE_PART_MAX_LINE = 15,
  ASH3D_TUNE_PART_MAX_LINE_STRIDED = 16,
  /* advection of the step with diffusion (ash3d_tr_step, graph and
   * ensemble steps): 1 the second horizontal sweep and the vertical sweep
   * in one pass (the first sweep writes the ts1 slot, the fused pass reads
   * it with the halo along its lines and writes ts0; bitwise the split
   * sweeps), 0 three separate in-place sweeps; default 1 on grids of at
   * least 2^24 interior cells, 0 below */
  ASH3D_TUNE_ADV_FUSED = 17,
  /* partitioned mode, eager steps: the most Jacobi sweeps a line may take
   * (0..12; 0 never; default 12).  y and z theta-scheme sweeps whose lines
   * are too long for the partitioned solves are solved by Jacobi sweeps
   * in one pass when at most min(this, 7) bring the error under 2^-52 of
   * the line's largest value (rows with rho = max (|DL| + |DU|) / |D|
   * need the least K + 1 with rho^(K+1) <= 2^-52 (1 - rho)); otherwise
   * dgtsv order.  x lines: see ASH3D_TUNE_JACOBI_X */
  ASH3D_TUNE_JACOBI_SWEEPS = 18,
  /* read only: the Jacobi sweeps the last y (z) theta-scheme sweep took,
   * 0 when it used another solver */
  ASH3D_TUNE_JACOBI_LAST_Y = 19,
  ASH3D_TUNE_JACOBI_LAST_Z = 20,
  /* partitioned mode, eager steps, x lines: each line whose rows' rho
   * needs at most ASH3D_TUNE_JACOBI_SWEEPS sweeps (as above: seven take
   * rho up to 5.80e-3, twelve up to 4.8e-2) takes that many Jacobi sweeps, the partitioned
   * solves the others; 1 (default) for lines of at least 256 cells, 2 at
   * any length, 0 never */
  ASH3D_TUNE_JACOBI_X = 21,
  /* read only: the most Jacobi sweeps a line of the last x theta-scheme
   * sweep may have taken (0: none) */
  ASH3D_TUNE_JACOBI_LAST_X = 22
};
/* Theta-scheme line solves.  EXACT runs dgtsv's elimination (partial
 * pivoting) in dgtsv's order: bitwise the reference.  PARTITIONED splits
 * every line over the lanes of a wave (segment elimination, then parallel
 * cyclic reduction of the segment ends) for lines of 3..512 active cells
 * (other lines use EXACT); the matrices are strictly diagonally dominant,
 * so the results agree with dgtsv's to rounding (about 1e-14 relative). */
enum { ASH3D_LINE_SOLVER_EXACT = 0, ASH3D_LINE_SOLVER_PARTITIONED = 1 };
int ash3d_tr_tune(ash3d_tr_ctx *ctx, int32_t param, int32_t value);
/* The current value of a launch setting (the library's default until
 * ash3d_tr_tune changes it). */
int ash3d_tr_tune_get(ash3d_tr_ctx *ctx, int32_t param, int32_t *value);

/* Kernel timing.  No counterpart in the reference, whose only timers are
 * wall-clock reads around the whole run (src/Ash3d.F90:176-177, :727-732).
 * ash3d_tr_profile(ctx,1) clears the records and brackets every later
 * kernel launch of the context with HIP events on its stream;
 * ash3d_tr_profile(ctx,0) stops recording.  ash3d_tr_profile_read sums the
 * recorded GPU time (ms) and launch count of one kernel kind. */
enum {
  ASH3D_PROF_ADVECT_X = 0,
  ASH3D_PROF_ADVECT_Y = 1,
  ASH3D_PROF_ADVECT_Z = 2,
  ASH3D_PROF_DIFFUSE_X = 3,
  ASH3D_PROF_DIFFUSE_Y = 4,
  ASH3D_PROF_DIFFUSE_Z = 5,
  ASH3D_PROF_BC = 6,   /* Set_BC and per-direction ghost kernels */
  ASH3D_PROF_GLUE = 7, /* step glue */
  ASH3D_PROF_TRANSPOSE = 8, /* x<->y transposes around the x-line solves */
  ASH3D_PROF_ADJ_ADVECT = 9,   /* adjoint advection sweeps */
  ASH3D_PROF_ADJ_DIFFUSE = 10, /* adjoint diffusion sweeps */
  ASH3D_PROF_COPY = 11,        /* state snapshots and checkpoints */
  ASH3D_PROF_ADVECT_HZ = 12,   /* fused second horizontal + vertical advection */
  ASH3D_PROF_NKINDS = 13
};
int ash3d_tr_profile(ash3d_tr_ctx *ctx, int32_t enable);
int ash3d_tr_profile_read(ash3d_tr_ctx *ctx, int32_t kind, double *total_ms, int64_t *launches);

/* ---- Captured steps -----------------------------------------------------
 * No counterpart in the reference.  graph_capture records
 * ash3d_tr_step_glue followed by ash3d_tr_step(itime, ...) as a HIP graph
 * for the parity of itime (the x/y order alternates); graph_launch replays
 * it for a later step of the same parity, with the arguments captured
 * (dt, flags, IsAloft, the full-grid bounds).  Capture after at least one
 * eager step (scratch allocated, ts1 mirrored), with kernel timing off;
 * a synchronize-free replay costs one launch instead of about twenty. */
int ash3d_tr_graph_capture(ash3d_tr_ctx *ctx, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                           double imp_fac, const int32_t *is_aloft);
int ash3d_tr_graph_launch(ash3d_tr_ctx *ctx, int32_t itime);

/* ---- Batched ensembles --------------------------------------------------
 * No counterpart in the reference, which runs one simulation per process
 * (an ensemble forecast runs the executable once per member).  An ensemble
 * steps the states of n contexts on one device -- wind-perturbed members
 * on grids of different sizes, each with its own dt -- as batched launches:
 * every sweep kind of the step (each advection direction, each Set_BC(2)
 * face or Neumann pass, each diffusion direction) is one launch over all
 * members for each kernel variant the members' grids select, the members'
 * argument blocks and workgroup ranges held in device memory.  One
 * ensemble step of parity itime % 2 does, for every member m, what
 * ash3d_tr_step_glue(m) followed by ash3d_tr_step(m, itime, dt[m], 1, 1,
 * imp_fac, is_aloft[m]) does, bit for bit -- the glue only for members
 * whose last state change was a step: a member glued by the caller (to
 * insert a source between glue and step, as the reference's time loop
 * does) is stepped from its state as it is.  With use_graph the step is
 * one HIP graph launch (captured at create, one per parity).
 *   Requirements, checked at create and at every step: members on one
 * device, with diffusion by the partitioned Crank-Nicolson line solves
 * (ASH3D_TUNE_LINE_SOLVER = ASH3D_LINE_SOLVER_PARTITIONED) on diagonally
 * dominant lines of 3 to the solver's maximum active cells, and the last
 * state change on each member a step or a step glue (source insertion may
 * follow the glue).
 * is_aloft is null (every bin of every member aloft) or n pointers, each
 * null or an ns-long mask.  n_streams (0: 4; at most 16) splits the
 * members into that many groups of about equal work, each with its own
 * stream, so that one group's launches fill the others' tails; the graph
 * holds the groups as parallel branches.  The members' tuning at create fixes the
 * kernels; the step orders itself after earlier calls on each member's
 * stream, and later calls on the members after it.  Members must outlive
 * the ensemble, and must not be stepped on their own between ensemble
 * steps with another dt (the step rebuilds line factors that changed, at a
 * cost).  ensemble_info gives the launches of one step, batched and as
 * the members' own steps would issue them. */
typedef struct ash3d_tr_ens ash3d_tr_ens;
int ash3d_tr_ensemble_create(ash3d_tr_ctx *const *members, int32_t n, const double *dt, double imp_fac,
                             const int32_t *const *is_aloft, int32_t n_streams, ash3d_tr_ens **out);
int ash3d_tr_ensemble_step(ash3d_tr_ens *ens, int32_t itime, int32_t use_graph);
int ash3d_tr_ensemble_synchronize(ash3d_tr_ens *ens);
int ash3d_tr_ensemble_info(ash3d_tr_ens *ens, int32_t itime, int32_t *batched_launches, int32_t *member_launches);
int ash3d_tr_ensemble_destroy(ash3d_tr_ens *ens);

/* ---- Checkpoints ------------------------------------------------------
 * No counterpart in the reference (which only writes output files).  Slots
 * of device memory, each one copy of concen_pd(:,:,:,:,ts0).  Save copies
 * ts0 into a slot; load copies a slot into ts0 and leaves ts1 equal to its
 * interior with zero ghosts, as after a sweep, so that a following
 * ash3d_tr_step_glue or ash3d_tr_step continues from the saved state. */
enum { ASH3D_CKPT_SAVE = 0, ASH3D_CKPT_LOAD = 1 };
int ash3d_tr_checkpoints(ash3d_tr_ctx *ctx, int32_t nslots);
int ash3d_tr_checkpoint(ash3d_tr_ctx *ctx, int32_t slot, int32_t op);

/* ---- Discrete adjoint --------------------------------------------------
 * No counterpart in the reference.  The adjoint variable lambda has the
 * layout of concen_pd(:,:,:,:,ts0) (ghost entries stay zero).  Each adjoint
 * entry point applies the transpose of the Jacobian of the matching forward
 * routine, linearised at the forward input that is in ts0 when it is
 * called: lambda <- J(ts0)^T lambda.  For the limited DCU sweeps the
 * limiter branches are those the forward sweep takes on that input, so the
 * forward map is F(q) = J(q) q and the adjoint is exact for it; the
 * diffusion sweeps are linear.  ts0 is read, not changed (except that the
 * advection entry points set its Set_BC(1) ghosts, as the forward does).
 * Outflow and deposit are not differentiated. */
/* Allocates the adjoint buffers, zeroes lambda and n_source_slots source
 * gradient slots of (nz+1)*ns values. */
int ash3d_tr_adj_init(ash3d_tr_ctx *ctx, int32_t n_source_slots);
/* Host copies of lambda (ns*(nx+4)*(ny+4)*(nz+4) values); set(NULL) zeroes. */
int ash3d_tr_adj_set(ash3d_tr_ctx *ctx, const double *lambda);
int ash3d_tr_adj_get(ash3d_tr_ctx *ctx, double *lambda);
/* Adjoints of advect_x/y/z (dir 0/1/2; forward: src/AdvectionHorz_DCU.f90:
 * 68-305, :322-556; src/AdvectionVert_DCU.f90:100-323). */
int ash3d_tr_adj_advect(ash3d_tr_ctx *ctx, int32_t dir, double dt, const int32_t *is_aloft,
                        const int32_t bounds[6]);
/* Adjoints of diff_x/y/z and diffCN_x/y/z (src/Diffusion.F90:266-650,
 * :671-1428).  Linear, so ts0 is not used. */
int ash3d_tr_adj_diffuse(ash3d_tr_ctx *ctx, int32_t dir, int32_t use_cn, double dt, double imp_fac,
                         const int32_t *is_aloft, const int32_t bounds[6]);
/* Adjoint of ash3d_tr_step with its input state in ts0: the forward
 * advection sweeps are recomputed from ts0 into private snapshots, then the
 * adjoint sweeps run in reverse order. */
int ash3d_tr_adj_step(ash3d_tr_ctx *ctx, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                      double imp_fac, const int32_t *is_aloft);
/* ash3d_tr_adj_step with the sweep inputs kept by ash3d_tr_step_keep in
 * slots slot_b, slot_z instead of recomputed (ts0 holds the step input) */
int ash3d_tr_adj_step_kept(ash3d_tr_ctx *ctx, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                           double imp_fac, const int32_t *is_aloft, int32_t slot_b, int32_t slot_z);
/* Source gradient: slot(k,n) += dt * lambda(ivent,jvent,k,n), k = 1..nz+1,
 * the adjoint of ash3d_tr_source_column_add; read returns all slots
 * (slot-major, then bin, then k). */
int ash3d_tr_adj_source(ash3d_tr_ctx *ctx, int32_t ivent, int32_t jvent, double dt, int32_t slot);
int ash3d_tr_adj_source_read(ash3d_tr_ctx *ctx, double *out);
/* Column load L(i,j) = sum_k sum_n concen(i,j,k,n,ts0) * dz(k) / 1e3
 * (t/km2, ConcentrationCalculator, src/Output_Vars.f90:1058-1064, without
 * its display threshold), out(nx,ny) x fastest; dz(1:nz) in km.  The
 * adjoint adds L^T w to lambda. */
int ash3d_tr_column_load(ash3d_tr_ctx *ctx, const double *dz, double *out);
int ash3d_tr_adj_column_load(ash3d_tr_ctx *ctx, const double *dz, const double *w);

/* ---- Source term of the standard single-column sources (host) ---------
 * The routines of src/Source.f90 that build what ash3d_tr_source_column_add
 * inserts each step; no device or context is needed.  Umbrella sources
 * (3x3 insertion, src/Source_Umbrella.f90) are not provided: their
 * normalized column is the Suzuki one, but their insertion is not.
 *   ash3d_tr_source_column    Calc_Normalized_SourceCol (:287-490) and
 *                             EruptivePulse_MassFluxRate (:508-589):
 *                             norm(i,k) at norm[i-1 + neruptions*(k-1)],
 *                             k = 1..nz; mass_flux_rate (kg/hr) and
 *                             end_time (h) per pulse (either may be NULL)
 *   ash3d_tr_eruptive_pulses  CheckEruptivePulses (:613-685) for the step
 *                             [time, time+dt]: the pulse indices
 *                             ieruption/jeruption (1-based, kept by the
 *                             caller between steps; both start at 1), the
 *                             fraction of dt of each pulse, and whether a
 *                             source is inserted in this step.  As in the
 *                             reference, a pulse that starts inside dt is
 *                             not counted in that step.
 *   ash3d_tr_source_nodes     TephraSourceNodes (:708-783): flux(k,n) =
 *                             SourceNodeFlux (kg/km3/hr), k = 1..nz+1 (row
 *                             nz+1 zero), n = 1..ns (bins past n_gs_max
 *                             zero), the layout of ash3d_tr_source_column_add;
 *                             kappa_col = kappa_pd(ivent,jvent,1:nz); checks
 *                             that the fluxes carry the mass flux rate
 *                             (EINVAL otherwise, as the reference stops) and
 *                             sets ieruption = jeruption. */
enum {
  ASH3D_SRC_SUZUKI = 0,
  ASH3D_SRC_UMBRELLA = 1,     /* normalized column only */
  ASH3D_SRC_UMBRELLA_AIR = 2, /* normalized column only */
  ASH3D_SRC_LINE = 3,
  ASH3D_SRC_POINT = 4,
  ASH3D_SRC_PROFILE = 5,
  ASH3D_SRC_OTHER = 6 /* zero column and mass flux rate */
};
typedef struct {
  int32_t source_type;        /* ASH3D_SRC_* (SourceType) */
  int32_t neruptions;
  const double *start_time;   /* e_StartTime(1:neruptions), h, on the driver's clock */
  const double *duration;     /* e_Duration, h */
  const double *volume;       /* e_Volume, km3 DRE */
  const double *plume_height; /* e_PlumeHeight, km */
  double suzuki_a;            /* Suzuki_A */
  double magma_density;       /* MagmaDensity, kg/m3 */
  double z_volcano;           /* vent elevation, km */
  int32_t zscaling_id;        /* ZScaling_ID: 0 s = z, 1 shifted, 2 scaled */
  double zsurf_vent;          /* Zsurf(ivent,jvent), km */
  double ztop;                /* Ztop, km */
  const double *s_lb;         /* s_lb_pd(1:nz+1), cell lower bounds in s */
  const double *s_cc;         /* s_cc_pd(1:nz) */
  const double *ds;           /* ds_vec_pd(1:nz) */
  const double *bin_mass;     /* Tephra_bin_mass(1:n_gs_max) */
  /* profile sources */
  const double *prof_dz;      /* e_prof_dz(1:neruptions), km */
  const int32_t *prof_nz;     /* e_prof_nzpoints(1:neruptions) */
  const double *prof_volume;  /* e_prof_Volume(i,kk) at prof_volume[i-1 + prof_ld*(kk-1)], km3 */
  int32_t prof_ld;            /* leading dimension of prof_volume (neruptions) */
} ash3d_tr_source_desc;
int ash3d_tr_source_column(const ash3d_tr_source_desc *src, int32_t nz, double *norm, double *mass_flux_rate,
                           double *end_time);
int ash3d_tr_eruptive_pulses(const ash3d_tr_source_desc *src, const double *end_time, double time, double dt,
                             int32_t *ieruption, int32_t *jeruption, double *frac, int32_t *in_dt);
int ash3d_tr_source_nodes(const ash3d_tr_source_desc *src, int32_t nz, int32_t ns, int32_t n_gs_max,
                          const double *norm, const double *mass_flux_rate, const double *frac,
                          const double *kappa_col, int32_t *ieruption, int32_t jeruption, double *flux);

/* ---- Umbrella sources (src/Source_Umbrella.f90) -------------------------
 * Host work (a few hundred operations per call):
 *   ash3d_tr_umbrella_stencil  the weights of Allocate_Source_Umbrella
 *                              (:85-153): avg_stencil[(i-1)+3(j-1)+9(k-1)] =
 *                              kappa(ivent-2+i, jvent-2+j, k) / (sum of the
 *                              3x3 kappa of level k), scale_fac[k-1] =
 *                              kappa(ivent, jvent, k) / that sum, k = 1..nz;
 *                              kappa_pd is the padded (nx+4, ny+4, nz+4) array.
 *   ash3d_tr_umbrella_levels   ibase/itop (:139-150): the levels k = 1..nz
 *                              with z(k-1) < 0.75 H <= z(k) and
 *                              z(k-1) < H <= z(k), z = z_vec_init(0:nz) (km),
 *                              H the first pulse's plume height; 0 when no
 *                              level brackets them.
 *   ash3d_tr_umbrella_winds    umbrella_winds (:206-394): the radial spreading
 *                              wind (km/hr) of the umbrella cloud at every
 *                              cell centre (nx*ny, x fastest), the same on
 *                              every level ibase..itop; zero at the vent, off
 *                              the cloud and while the cloud is no wider than
 *                              the 3x3 source patch.  Lon-lat grids only.
 *   ash3d_tr_umbrella_nodes    TephraSourceNodes_Umbrella (:409-448): the
 *                              3x3 patch flux flux_umb[(i-1)+3(j-1)+9(k-1)+
 *                              9(nz+1)n] from the column flux of
 *                              ash3d_tr_source_nodes: the column below ibase,
 *                              flux*scale_fac on all nine cells of levels
 *                              ibase..itop, nothing above.
 *   ash3d_tr_umbrella_vol_inc  SourceVolInc_Umbrella (:464-522), km3 DRE.
 * Device work:
 *   ash3d_tr_umbrella_source_add  src/Ash3d.F90:447-460: on levels
 *                              ibase..itop each bin's 3x3 patch around the
 *                              vent is replaced by its stencil-weighted
 *                              average (AvgCon_Umbrella, :539-560), then
 *                              dt*flux_umb is added to the patch on levels
 *                              1..nz+1; bins 1..n_gs_max.
 *   ash3d_tr_umbrella_winds_add   src/MesoInterpolater.f90:238-243: vx_pd and
 *                              vy_pd += the umbrella winds on levels
 *                              ibase..itop (call after the met interpolation). */
typedef struct {
  int32_t vel_model;          /* VelMod_umb: 1 (Eq. 3 with the legacy r^3 term) or 2 (Eq. 4) */
  double k_entrainment;       /* k_entrainment_umb (0.1) */
  double lambda;              /* lambda_umb, cloud shape factor (0.2) */
  double n_bv;                /* N_BV_umb, Brunt-Vaisala frequency, 1/s (0.02) */
  double mass_flux_rate;      /* MassFluxRate(1), kg/hr */
  int32_t airborne_one_bin;   /* SourceType 'umbrella_air' with n_gs_max == 1: flux x 20 */
  double lat_volcano, lon_volcano;  /* degrees */
  double time, dt;            /* hours since the eruption start; the step */
  int32_t first_time;         /* 1: use min(simtime, end_time) instead of max(time, dt) */
  double simtime, end_time;   /* Simtime_in_hours, e_EndTime(1) */
  double de_km, dn_km;        /* cell width and height, km */
  int32_t ivent, jvent;
  int32_t is_latlon;          /* projected grids are an error, as in the reference */
} ash3d_tr_umbrella_wind_params;
int ash3d_tr_umbrella_stencil(const double *kappa_pd, int32_t nx, int32_t ny, int32_t nz, int32_t ivent,
                              int32_t jvent, double *avg_stencil, double *scale_fac);
int ash3d_tr_umbrella_levels(const double *z_vec_init, int32_t nz, double plume_height, int32_t *ibase,
                             int32_t *itop);
int ash3d_tr_umbrella_winds(const ash3d_tr_umbrella_wind_params *p, int32_t nx, int32_t ny, const double *lon_cc,
                            const double *lat_cc, double *uvx, double *uvy);
int ash3d_tr_umbrella_nodes(const double *flux, int32_t nz, int32_t ns, int32_t n_gs_max, int32_t ibase,
                            int32_t itop, const double *scale_fac, double *flux_umb);
int ash3d_tr_umbrella_vol_inc(const double *flux, int32_t nz, int32_t ns, int32_t n_gs_max, int32_t ibase,
                              int32_t itop, const double *avg_stencil, const double *kappa_vent_col, double dt,
                              double magma_density, double *vol);
int ash3d_tr_umbrella_source_add(ash3d_tr_ctx *ctx, int32_t ivent, int32_t jvent, double dt, int32_t ibase,
                                 int32_t itop, int32_t n_gs_max, const double *avg_stencil,
                                 const double *flux_umb);
int ash3d_tr_umbrella_winds_add(ash3d_tr_ctx *ctx, int32_t ibase, int32_t itop, const double *uvx,
                                const double *uvy);

/* ---- Output products and their ESRI ASCII files -----------------------
 * ash3d_tr_products replaces ConcentrationCalculator
 * (src/Output_Vars.f90:1026-1137) and AshThicknessCalculator (:855-893) on
 * the device copy of concen_pd(...,ts1) and DepositGranularity.  Inputs:
 * dz_vec_pd(1:nz), z_cc_pd(1:nz) (km); bounds = imin,imax,jmin,jmax of the
 * columns ConcentrationCalculator visits (NULL: the whole grid); the
 * thresholds and the deposit density.  Outputs (host, nx*ny, x fastest; any
 * may be NULL): CloudLoad (t/km2), MaxConcentration (mg/m3), MaxHeight and
 * MinHeight (km), DepositThickness (mm), Mask_Cloud and Mask_Deposit (0/1),
 * and CloudArea / DepositAreaCovered (sums of sigma_nz_pd(i,j,1) over the
 * masked columns, in the reference's loop order, from the descriptor's
 * host geometry). */
typedef struct {
  double cloudcon_thresh;  /* CLOUDCON_THRESH (kg/km3) */
  double cloudload_thresh; /* CLOUDLOAD_THRESH (t/km2) */
  double depo_thresh;      /* DEPO_THRESH (mm) */
  double deposit_density;  /* DepositDensity (kg/m3) */
} ash3d_tr_product_params;
typedef struct {
  double *cloud_load, *max_concentration, *max_height, *min_height, *deposit_thickness;
  int32_t *mask_cloud, *mask_deposit;
  double cloud_area, deposit_area; /* set by the call */
} ash3d_tr_products_out;
int ash3d_tr_products(ash3d_tr_ctx *ctx, const double *dz, const double *zcc, const int32_t *bounds,
                      const ash3d_tr_product_params *params, ash3d_tr_products_out *out);
/* write_2D_ASCII (src/write_ASCII.f90:264-382): field(nx,ny) x fastest,
 * mask (0/1, NULL = all) replaces unmasked cells by fill_value (up to six
 * characters, e.g. "-9999."); is_latlon selects the header (lonLL, latLL,
 * de, dn in degrees, lonLL > 180 shifted by -360) or the projected one
 * (xLL, yLL, dx, dy in km, written in m).  Values are written as real(4)
 * in 10f18.6 records from the northern row down.  Host only: no device or
 * context is needed. */
int ash3d_tr_write_esri_ascii(const char *path, int32_t nx, int32_t ny, const double *field,
                              const int32_t *mask, const char *fill_value, int32_t is_latlon, double xll,
                              double yll, double dx, double dy);
/* The file name write_2D_ASCII gives a product: the trimmed root, then
 * "____final" or the output time as 00h.hhhrs / 0hh.hhhrs / hhh.hhhrs, then
 * ".dat" (roots containing "outvar" or "ArrivalTime" get no time).  Writes
 * at most outlen bytes including the terminating NUL. */
int ash3d_tr_esri_ascii_name(const char *root, double time_h, int32_t is_final, char *out, int32_t outlen);

/* ---- Driver work between the hot-path calls ---------------------------
 * Met steps: the two bracketing wind fields of the wind_grid module
 * (v*_meso_{last,next}_step_sp, m/s, FP32, unpadded (1:nx,1:ny,1:nz) and
 * vf (1:nx,1:ny,1:nz,1:ns)).  push makes the current "next" step the "last"
 * one and uploads a new "next" (vf may be NULL); interp evaluates
 * src/MesoInterpolater.f90:203-230 into vx/vy/vz_pd (and vf_pd if use_vf)
 * at Interval_Frac = frac. */
int ash3d_tr_met_push(ash3d_tr_ctx *ctx, const float *vx, const float *vy, const float *vz, const float *vf);
int ash3d_tr_met_interp(ash3d_tr_ctx *ctx, double frac, int32_t use_vf);

/* Adjust_DT, src/Adjust_dt.f90:89-333.  use_met_next selects the met-step
 * branch (Adjust_DT(mesostep=.true.), reading the "next" met step); else the
 * current vx/vy/vz/vf_pd.  The diffusion restriction is evaluated on every
 * call (the reference caches it after the first call unless the
 * diffusivity varies).  Returns dt (hours) and diag = vxmax_dx, vymax_dy,
 * vzmax_dz, time_diffuse (diag may be NULL). */
typedef struct ash3d_tr_dt_params {
  int32_t is_latlon, use_met_next, n_gs_aloft, use_diffusion, use_cn;
  double dx, dy;      /* km (projected grids) */
  const double *dz;   /* dz_vec_pd(1:nz), km, host */
  double imp_dt_fac, cfl, dt_min, dt_max, eps_small;
  double time, next_write_time, simtime; /* hours */
} ash3d_tr_dt_params;
int ash3d_tr_adjust_dt(ash3d_tr_ctx *ctx, const ash3d_tr_dt_params *p, double *dt, double *diag);

/* get_minmax_index (FAST_SUBGRID), src/Calc_Mesh.f90:432-542, on ts0 with
 * CLOUDCON_GRID_THRESH = thresh (1e-7 in the reference,
 * src/Output_Vars.f90:120).  bounds = imin,imax,jmin,jmax,kmin,kmax.
 * kmax < kmin is ASH3D_TR_EINVAL (the reference stops). */
int ash3d_tr_minmax_index(ash3d_tr_ctx *ctx, int32_t ivent, int32_t jvent, int32_t kvent, double thresh,
                          int32_t bounds[6]);
/* ash3d_tr_step built with FAST_SUBGRID: AdvectHorz takes the box of
 * get_minmax_index (src/AdvectionHorz.f90:60-66), AdvectVert the same box,
 * DiffuseHorz/DiffuseVert the box grown by one cell
 * (src/Diffusion.F90:168-175).  bounds returns the advection box. */
int ash3d_tr_step_subgrid(ash3d_tr_ctx *ctx, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                          double imp_fac, const int32_t *is_aloft, int32_t ivent, int32_t jvent, int32_t kvent,
                          double thresh, int32_t bounds[6]);

/* Message of the last failed call on this thread. */
const char *ash3d_tr_last_error(void);

/* Library build description (compile flags, target). */
const char *ash3d_tr_build_info(void);

#ifdef __cplusplus
}
#endif
#endif
