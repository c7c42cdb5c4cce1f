// C ABI of the gfx950 transport step (declared in include/ash3d_tr.h).
//
// A context owns device mirrors of the reference's module arrays, one HIP
// stream and the scratch the line solvers need.  Entry points enqueue
// kernels on the stream and return; sync_to_host / synchronize wait.
//
// ts1 bookkeeping.  Every sweep of the reference zero-fills all of
// concen_pd(...,ts1), writes the swept rows and copies the interior back to
// ts0 (e.g. src/AdvectionHorz_DCU.f90:122, :285-287, :302-303), so after any
// sweep ts1 equals ts0's interior with zero ghosts.  The kernels update ts0
// in place and the context records that fact (TS1_MIRROR) instead of writing
// ts1; ts1 is materialised only when the host asks for it.  The driver glue
// ts0 <- ts1, ts1 <- 0 (src/Ash3d.F90:376-378) then only has to zero ts0's
// ghost cells.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ash3d_tr.h"
#include "common.hpp"

using namespace ash3d;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(ASH3D_TR_EDEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                  __FILE__, __LINE__);                                               \
  } while (0)

enum { TS1_EXPLICIT = 0, TS1_ZERO = 1, TS1_MIRROR = 2 };

}  // namespace

// Scratch of the allocation-free entry points (ash3d_tr_mass_totals_noalloc):
// the per-workgroup partial sums of the mass totals, for at most
// ASH3D_TR_NOALLOC_MAX_BINS bins.  Each device holds its own copy of this
// array from the library's load until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in):
// ash3d_tr_create resolves the device address once per device with
// hipGetSymbolAddress and every use goes through that address.
static constexpr long long kNoallocScratch = (long long)ASH3D_TR_NOALLOC_MAX_BINS * 128 * 3;
static __device__ __align__(256) double scratch[kNoallocScratch];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device `dev`, resolved on
// the first call for that device.
int resolve_scratch(int dev, double **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(ASH3D_TR_EINVAL, "device ordinal %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ASH3D_TR_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

struct ash3d_tr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double *scratch = nullptr;  // this device's address of the allocation-free scratch array
  Dims d{};
  int periodic = 0, zscaling = 0, limiter = ASH3D_LIM_SUPERBEE;
  ash3d_tr_desc host{};
  // device arrays
  double *c = nullptr;  // ts0 then ts1, ns*ncell each
  double *vx = nullptr, *vy = nullptr, *vz = nullptr, *vh = nullptr, *vf = nullptr;
  double *kap = nullptr, *sx = nullptr, *sy = nullptr, *sz = nullptr;
  double *kx = nullptr, *ky = nullptr, *kz = nullptr;
  double *ddx = nullptr, *ddy = nullptr;
  double *ofxz1 = nullptr, *ofxz2 = nullptr, *ofyz1 = nullptr, *ofyz2 = nullptr;
  double *ofxy1 = nullptr, *ofxy2 = nullptr;
  double *dep = nullptr;
  double *fac = nullptr;  // line-solve U factors (3 * ncell)
  // theta-scheme row factors L, R per direction (2 * ncell each), valid for lr_dt
  double *lr[3] = {nullptr, nullptr, nullptr};
  double lr_dt[3] = {0.0, 0.0, 0.0};
  bool lr_ok[3] = {false, false, false};
  // LU factors of the theta-scheme lines per direction (one allocation:
  // fact, D, DU, DL (4 * ncell doubles), then ncell interchange flags, one
  // INFO per line, the count of lines that are not diagonally dominant,
  // the largest row ratio (|DL| + |DU|) / |D| as double bits, per line the
  // Jacobi sweeps it needs (0: more than kJacobiMaxSweepsX), the number of
  // lines that need at most that many and their largest ratio), valid for
  // the key (dt, theta, rmin, rmax)
  double *lu[3] = {nullptr, nullptr, nullptr};
  uint8_t *lu_piv[3] = {nullptr, nullptr, nullptr};
  int *lu_sing[3] = {nullptr, nullptr, nullptr};
  int *lu_ndom[3] = {nullptr, nullptr, nullptr};
  unsigned long long *lu_rho[3] = {nullptr, nullptr, nullptr};
  uint8_t *lu_jl[3] = {nullptr, nullptr, nullptr};
  unsigned long long *lu_rhoj[3] = {nullptr, nullptr, nullptr};
  int *lu_njl[3] = {nullptr, nullptr, nullptr};
  double lu_dt[3] = {0.0, 0.0, 0.0}, lu_th[3] = {0.0, 0.0, 0.0};
  int lu_rmin[3] = {0, 0, 0}, lu_rmax[3] = {0, 0, 0};
  bool lu_ok[3] = {false, false, false};
  long long lu_count = 0;  // factorisations launched (any direction)
  // host copy of each direction's non-dominant line count (pinned, read
  // back asynchronously after a factorisation): 0 unknown, 1 in flight
  // (ndom_ev), 2 known.  When known, a sweep launches only the solver it
  // needs instead of both with the choice made on the device.  The row
  // ratios and the flagged line count come back with it (njl_h, rho_h,
  // rhoj_h, in the same pinned block).
  int *ndom_h = nullptr, *njl_h = nullptr;
  unsigned long long *rho_h = nullptr, *rhoj_h = nullptr;
  hipEvent_t ndom_ev[3] = {};
  int ndom_state[3] = {0, 0, 0};
  bool capturing = false;
  // the velocity ghosts hold Set_BC(1)'s values for the current interior
  // velocities (cleared by every entry point that writes velocities)
  bool vel_bc_ok = false;
  // the glue's zeroing of the ts0 ghosts, not yet launched: the next
  // transport step's Set_BC(1) rewrites every concentration ghost from
  // interior cells, so it stands in for it; every other entry point but
  // the met/wind updates launches it first (the vent-column source too: it
  // adds into the ghost cell above the column, which a later reader must
  // see as zero plus the source)
  bool ghost_zero_pending = false;
  // umbrella source: device copies of the stencil + patch flux and of the
  // two wind planes, uploaded through one pinned staging buffer each
  double *umb_src = nullptr, *umb_wind = nullptr;
  double *umb_pin[2] = {nullptr, nullptr};
  size_t umb_pin_bytes[2] = {0, 0};
  hipEvent_t umb_ev[2] = {};
  bool umb_used[2] = {false, false};
  double *junk = nullptr;  // kJunk doubles that masked-off lanes store to
  // the sweep arguments the factors of each direction were last requested
  // with, and those of the sweeps in each captured graph (graph_launch
  // brings the caches back to them before a replay)
  DiffArgs lu_args[3] = {};
  bool lu_args_set[3] = {false, false, false};
  DiffArgs graph_lu[2][3] = {};
  bool graph_lu_on[2][3] = {};
  bool graph_tgeom[2] = {false, false};  // the graph reads the transposed geometry
  long long fac_groups = 0;
  double *flux = nullptr;  // source column staging (nz+1)*ns
  // pinned host ring for the source columns: the caller's buffer is copied
  // into a slot, so the upload is asynchronous and the host may run ahead
  static constexpr int kFluxRing = 4;
  double *flux_pinned[kFluxRing] = {};
  hipEvent_t flux_ev[kFluxRing] = {};
  bool flux_used[kFluxRing] = {};
  int flux_slot = 0;
  int *flag = nullptr;     // singular-pivot flag
  double *tot = nullptr;   // mass-total partial sums and results
  // x-line solves run on an x<->y transposed copy: state, and the cached
  // transposed kappa, sigma_nx and kx (rebuilt after a geometry/K upload)
  double *tq = nullptr, *tkap = nullptr, *tsx = nullptr, *tkx = nullptr;
  bool tgeom_valid = false;
  // adjoint: lambda, snapshots of the recomputed sweep inputs, line-solve
  // scratch, an outflow sink for recomputed sweeps, source-gradient slots
  double *lam = nullptr, *snap_b = nullptr, *snap_z = nullptr, *adj_sb = nullptr, *sink = nullptr;
  double *src_grad = nullptr;
  int n_src_slots = 0;
  double *col = nullptr;  // column-load staging: nz dz values, then nx*ny loads / weights
  double *prod = nullptr; // output products: 5 planes, 2 mask planes, dz and z_cc
  std::vector<double *> ckpt;  // checkpoint slots
  // met steps: two slots of vx, vy, vz (n3) and vf (n3*ns) in FP32
  float *met[2][4] = {};
  bool met_vf[2] = {false, false};
  int met_next = 1, met_count = 0;
  double *dtp = nullptr;                 // Adjust_DT partials, then dz
  // captured glue + step, one graph per step parity
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  unsigned long long *prof = nullptr;    // get_minmax_index slab maxima
  int ts1_state = TS1_EXPLICIT;
  int nb_adv[3] = {4, 4, 4}, nb_diff[3] = {4, 4, 4};  // bins per thread, by direction
  int x_tile = 16;  // cells per LDS tile of the x-line solves
  int line_solver = ASH3D_LINE_SOLVER_EXACT;  // theta-scheme line solves
  int part_bins = 0;  // bins per workgroup of the partitioned solves (0: all)
  // tile shape of the partitioned solves, by direction (measured on the
  // 360x240x50 forecast grid: 2-line x workgroups, 16-line y tiles)
  int part_shape[3] = {1, 1, 0};
  // longest active line the partitioned solves take (longer lines: dgtsv
  // order); lines beyond 512 cells span several waves.  Strided lines
  // (y, z) default to 512: their multi-wave tiles hold only 1-2 lines, too
  // few for coalesced rows (720-cell y lines: 17.2 ms against 13.6 ms in
  // dgtsv order on the 1440x720x80x20 global grid)
  int part_max_line = 2048, part_max_line_strided = 512;
  // streamed advection: bin groups per workgroup (0: line segments); the
  // z default is the autotuner's choice on the 360x240x50x15 forecast grid
  int adv_gy[3] = {0, 0, 4};
  int adv_xrows = 1;
  int adv_fused = 1;  // second horizontal + vertical advection in one pass (step with diffusion)
  // most Jacobi sweeps a line may take (0: never; y and z lines at most
  // kJacobiMaxIters + 1 in any case), and the sweeps the last sweep of
  // each direction took (0: another solver)
  int jacobi_sweeps = kJacobiMaxSweepsX;
  int jacobi_last[3] = {0, 0, 0};
  // x lines: 1 the lines that need at most jacobi_sweeps sweeps take them
  // and the partitioned solves the rest, for lines of at least 256 cells
  // (shorter ones leave most of the kernel's lanes idle); 2 at any
  // length; 0 never
  int jacobi_x = 1;
  // dt / kappa per cell for the fused pass, valid for dtv_dt
  double *dtvc = nullptr;
  double dtv_dt = 0.0;
  bool dtv_ok = false;
  double graph_dtv[2] = {-1.0, -1.0};  // dt of each captured graph's fused pass (< 0: none)
  // kernel timing (ash3d_tr_profile): event pairs per kernel kind
  bool prof_on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, size_t>> ev_rec;  // (kind, pool index)

  long long n_conc() const { return d.ncell * d.ns; }
  long long n_k() const { return (long long)(d.nx + 2) * (d.ny + 2) * (d.nz + 2); }
  long long n_2d() const { return (long long)(d.nx + 4) * (d.ny + 4); }
  long long n_xz() const { return (long long)(d.nx + 4) * (d.nz + 4) * d.ns; }
  long long n_yz() const { return (long long)(d.ny + 4) * (d.nz + 4) * d.ns; }
  long long n_xy() const { return (long long)(d.nx + 4) * (d.ny + 4) * d.ns; }
  long long n_dep() const { return (long long)d.nx * d.ny * d.ns; }
  double *ts0() const { return c; }
  double *ts1() const { return c + n_conc(); }
};

namespace {

int dev_alloc(double **p, long long n) {
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void **)p, sizeof(double) * (size_t)n);
  if (e != hipSuccess) return fail(ASH3D_TR_ENOMEM, "hipMalloc of %lld doubles: %s", n, hipGetErrorString(e));
  return 0;
}

int upload(ash3d_tr_ctx *x, double *dst, const double *src, long long n, double fill) {
  if (src) {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, x->stream));
  } else {
    HIPCHK(launch_fill(x->stream, dst, n, fill));
  }
  return 0;
}

int download(ash3d_tr_ctx *x, double *dst, const double *src, long long n) {
  if (dst) HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, x->stream));
  return 0;
}

// Entry points that may leave the glue's ghost zeroing pending.
int check_ctx_keep(ash3d_tr_ctx *x) {
  if (!x) return fail(ASH3D_TR_EINVAL, "null context");
  HIPCHK(hipSetDevice(x->device));
  return 0;
}

int check_ctx(ash3d_tr_ctx *x) {
  int rc = check_ctx_keep(x);
  if (rc) return rc;
  if (x->ghost_zero_pending) {
    x->ghost_zero_pending = false;
    HIPCHK(launch_zero_ghosts(x->stream, x->d, x->ts0()));
  }
  return 0;
}

int make_mask(const ash3d_tr_ctx *x, const int32_t *is_aloft, AloftMask *m) {
  std::memset(m, 0, sizeof *m);
  for (int n = 0; n < x->d.ns; ++n)
    if (!is_aloft || is_aloft[n]) m->w[n >> 5] |= (1u << (n & 31));
  return 0;
}

int check_bounds(const ash3d_tr_ctx *x, const int32_t *b) {
  const Dims &d = x->d;
  if (!b) return fail(ASH3D_TR_EINVAL, "null bounds");
  if (b[0] < 1 || b[1] > d.nx || b[2] < 1 || b[3] > d.ny || b[4] < 1 || b[5] > d.nz || b[0] > b[1] ||
      b[2] > b[3] || b[4] > b[5])
    return fail(ASH3D_TR_EINVAL, "bounds (%d:%d,%d:%d,%d:%d) outside the grid %dx%dx%d", b[0], b[1], b[2],
                b[3], b[4], b[5], d.nx, d.ny, d.nz);
  return 0;
}

void full_bounds(const ash3d_tr_ctx *x, int32_t *b) {
  b[0] = 1; b[1] = x->d.nx; b[2] = 1; b[3] = x->d.ny; b[4] = 1; b[5] = x->d.nz;
}

// Event-bracketed launches for ash3d_tr_profile.
struct ProfScope {
  ash3d_tr_ctx *x;
  size_t idx;
  bool on;
  ProfScope(ash3d_tr_ctx *x_, int kind) : x(x_), idx(0), on(x_->prof_on) {
    if (!on) return;
    if (x->ev_used == x->ev_pool.size()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { on = false; return; }
      x->ev_pool.emplace_back(a, b);
    }
    idx = x->ev_used++;
    x->ev_rec.emplace_back(kind, idx);
    hipEventRecord(x->ev_pool[idx].first, x->stream);
  }
  ~ProfScope() {
    if (on) hipEventRecord(x->ev_pool[idx].second, x->stream);
  }
};

// Set_BC(1) on a state buffer (ts0 unless given); conc = false: the
// velocity ghosts only.
int set_bc1(ash3d_tr_ctx *x, bool velocities, double *state = nullptr, bool conc = true) {
  ProfScope ps(x, ASH3D_PROF_BC);
  // Set_BC(1) on velocities is a projection (every ghost a copy of an
  // interior or wrapped cell that the pass does not write), so it is
  // skipped while the ghosts already hold its values; a captured graph
  // always contains it
  if (velocities && (x->capturing || !x->vel_bc_ok))
    HIPCHK(launch_set_bc1_vel(x->stream, x->d, x->periodic, x->vx, x->vy, x->vz, x->vf));
  if (velocities && !x->capturing) x->vel_bc_ok = true;
  if (conc) HIPCHK(launch_set_bc1_conc(x->stream, x->d, x->periodic, state ? state : x->ts0()));
  return 0;
}

// Argument block of one advection sweep over state c0.
AdvArgs advect_args(ash3d_tr_ctx *x, int dir, double dt, const int32_t *is_aloft, const int32_t *bounds,
                    double *c0) {
  AdvArgs a{};
  a.d = x->d;
  a.g = make_line_geom(x->d, dir, bounds);
  a.c0 = c0;
  a.cout = c0;
  a.dt = dt;
  a.periodic = x->periodic;
  a.limiter = x->limiter;
  make_mask(x, is_aloft, &a.aloft);
  a.kap = x->kap;
  a.junk = x->junk;
  a.gy = x->adv_gy[dir];
  a.xrows = x->adv_xrows;
  if (dir == 0) {
    a.vel = x->vx; a.sig = x->sx; a.of1 = x->ofyz1; a.of2 = x->ofyz2;
    a.deld = (x->zscaling == 2) ? x->ddx : nullptr;
  } else if (dir == 1) {
    a.vel = x->vy; a.sig = x->sy; a.of1 = x->ofxz1; a.of2 = x->ofxz2;
    a.deld = (x->zscaling == 2) ? x->ddy : nullptr;
  } else {
    a.vel = x->vz; a.vh = x->vh; a.vf = x->vf; a.sig = x->sz; a.of1 = x->ofxy1; a.of2 = x->ofxy2;
    a.dep = x->dep;
  }
  return a;
}

// One advection sweep on ts0, or on another state buffer whose outflow and
// deposit go to the sink (recomputation for the adjoint; to_sink = false
// keeps the real outflow and deposit arrays).  With out, the sweep reads
// state and writes its result to out (every interior cell of the sweep's
// lines, then the ghost cells copied from state): out ends as state would
// after the in-place sweep, and state is left as it was.
// The fused advection of the step with diffusion: the first sweep (dir)
// from ts0 into the ts1 slot, its ghosts read as Set_BC(1)'s zeros (x on a
// periodic grid: the copies Set_BC(2) leaves), and the arguments of the
// pass that runs the second horizontal sweep (dir2) and the vertical
// sweep from ts1 into ts0
AdvArgs fused_first_args(ash3d_tr_ctx *x, int dir, double dt, const int32_t *is_aloft, const int32_t *bounds) {
  AdvArgs a = advect_args(x, dir, dt, is_aloft, bounds, x->ts0());
  a.zg = (x->periodic && dir == 0) ? 0 : 1;
  a.cout = x->ts1();
  return a;
}
void fused_hz_args(ash3d_tr_ctx *x, int dir2, double dt, const int32_t *is_aloft, const int32_t *bounds, AdvArgs *h,
                   AdvArgs *z) {
  *h = advect_args(x, dir2, dt, is_aloft, bounds, x->ts1());
  *z = advect_args(x, 2, dt, is_aloft, bounds, x->ts1());
  z->cout = x->ts0();
  h->zg = z->zg = 1;
  h->dtv = z->dtv = x->dtvc;
}
// dt / kappa of every cell for the fused pass, recomputed when dt changes
// (a geometry upload invalidates it)
int ensure_dtv(ash3d_tr_ctx *x, double dt) {
  int rc = 0;
  if (!x->dtvc && (rc = dev_alloc(&x->dtvc, x->d.ncell))) return rc;
  if (!x->dtv_ok || x->dtv_dt != dt) {
    HIPCHK(launch_cell_dtv(x->stream, x->kap, x->dtvc, x->d.ncell, dt));
    x->dtv_ok = true;
    x->dtv_dt = dt;
  }
  return 0;
}

int sweep_advect(ash3d_tr_ctx *x, int dir, double dt, const int32_t *is_aloft, const int32_t *bounds,
                 bool with_bc, bool bc_vel, double *state = nullptr, double *out = nullptr, bool to_sink = true,
                 bool zero_ghosts = false) {
  int rc = check_bounds(x, bounds);
  if (rc) return rc;
  if (!(dt >= 0.0)) return fail(ASH3D_TR_EINVAL, "dt must be >= 0 (got %g)", dt);
  if (with_bc && (rc = set_bc1(x, bc_vel, state))) return rc;
  AdvArgs a = advect_args(x, dir, dt, is_aloft, bounds, state ? state : x->ts0());
  // (periodic x: the x ghosts are copies, never zeros)
  a.zg = (zero_ghosts && !(x->periodic && dir == 0)) ? 1 : 0;
  if (state && out) a.cout = out;
  if (state && to_sink) {
    const long long face = std::max(x->n_yz(), std::max(x->n_xz(), x->n_xy()));
    a.of1 = x->sink;
    a.of2 = x->sink + face;
    a.dep = x->sink + 2 * face;
  }
  {
    ProfScope ps(x, ASH3D_PROF_ADVECT_X + dir);
    HIPCHK(launch_advect(x->stream, a, x->nb_adv[dir]));
  }
  if (state && out) {
    ProfScope ps(x, ASH3D_PROF_COPY);
    HIPCHK(launch_copy_ghosts(x->stream, x->d, state, out));
  }
  if (!state) x->ts1_state = TS1_MIRROR;
  return 0;
}

// Transposed kappa, sigma_nx and kx for the x-line solves.
int ensure_tgeom(ash3d_tr_ctx *x) {
  int rc = 0;
  const Dims &d = x->d;
  if (x->tgeom_valid) return 0;
  if (!x->tkap && ((rc = dev_alloc(&x->tkap, d.ncell)) || (rc = dev_alloc(&x->tsx, d.ncell)) ||
                   (rc = dev_alloc(&x->tkx, x->n_k()))))
    return rc;
  ProfScope ps(x, ASH3D_PROF_TRANSPOSE);
  HIPCHK(launch_transpose_xy(x->stream, x->kap, x->tkap, d.nx + 4, d.ny + 4, d.nz + 4));
  HIPCHK(launch_transpose_xy(x->stream, x->sx, x->tsx, d.nx + 4, d.ny + 4, d.nz + 4));
  HIPCHK(launch_transpose_xy(x->stream, x->kx, x->tkx, d.nx + 2, d.ny + 2, d.nz + 2));
  x->tgeom_valid = true;
  return 0;
}

// Row factors L, R of the theta-scheme lines of one direction, recomputed
// when dt changes (the geometry and K uploads invalidate them)
int ensure_lr(ash3d_tr_ctx *x, int dir, DiffArgs &a) {
  int rc = 0;
  if (!x->lr[dir] && (rc = dev_alloc(&x->lr[dir], 2 * x->d.ncell))) return rc;
  if (!x->lr_ok[dir] || x->lr_dt[dir] != a.dt) {
    HIPCHK(launch_line_lr(x->stream, a, x->lr[dir]));
    x->lr_ok[dir] = true;
    x->lr_dt[dir] = a.dt;
    x->lu_ok[dir] = false;
  }
  a.lr = x->lr[dir];
  return 0;
}

// LU factors of the theta-scheme lines of one direction (after ensure_lr),
// refactored when dt, theta or the active range along the lines changes
int ensure_lu(ash3d_tr_ctx *x, int dir, DiffArgs &a) {
  int rc = 0;
  const long long nc = x->d.ncell, lines = (long long)a.g.Na * a.g.Nb;
  if (!x->lu[dir]) {
    const long long tail = 4 * nc + (nc + 7) / 8 + (lines + 1) / 2;
    if ((rc = dev_alloc(&x->lu[dir], tail + 4 + (lines + 7) / 8))) return rc;
    x->lu_piv[dir] = (uint8_t *)(x->lu[dir] + 4 * nc);
    x->lu_sing[dir] = (int *)(x->lu[dir] + 4 * nc + (nc + 7) / 8);
    x->lu_ndom[dir] = (int *)(x->lu[dir] + tail);
    x->lu_rho[dir] = (unsigned long long *)(x->lu[dir] + tail + 1);
    x->lu_rhoj[dir] = (unsigned long long *)(x->lu[dir] + tail + 2);
    x->lu_njl[dir] = (int *)(x->lu[dir] + tail + 3);
    x->lu_jl[dir] = (uint8_t *)(x->lu[dir] + tail + 4);
  }
  if (!x->junk && (rc = dev_alloc(&x->junk, kJunk))) return rc;
  if (!x->lu_ok[dir] || x->lu_dt[dir] != a.dt || x->lu_th[dir] != a.theta || x->lu_rmin[dir] != a.g.rmin ||
      x->lu_rmax[dir] != a.g.rmax) {
    HIPCHK(launch_line_lu(x->stream, a, x->lu[dir], x->lu_piv[dir], x->lu_sing[dir], x->lu_ndom[dir],
                          x->lu_rho[dir], x->lu_jl[dir], x->lu_rhoj[dir], x->lu_njl[dir]));
    ++x->lu_count;
    x->ndom_state[dir] = 0;
    x->lu_ok[dir] = true;
    x->lu_dt[dir] = a.dt;
    x->lu_th[dir] = a.theta;
    x->lu_rmin[dir] = a.g.rmin;
    x->lu_rmax[dir] = a.g.rmax;
  }
  a.lu = x->lu[dir];
  a.piv = x->lu_piv[dir];
  a.sing = x->lu_sing[dir];
  a.ndom = x->lu_ndom[dir];
  a.sel = 0;
  a.junk = x->junk;
  x->lu_args[dir] = a;
  x->lu_args_set[dir] = true;
  return 0;
}

// Line-solve U factors (one set per line, shared by the bin groups)
int ensure_fac(ash3d_tr_ctx *x, int nb) {
  (void)nb;
  if (x->fac) return 0;
  int rc = dev_alloc(&x->fac, 3 * x->d.ncell);
  if (rc) return rc;
  x->fac_groups = 1;
  return 0;
}

// Partitioned line solves of one direction when selected and the active
// line length allows (3..512 cells); returns 1 when they were launched, and
// then sets a.sel = 2 so that the dgtsv-order kernel the caller launches
// next runs only if the sweep has lines that are not diagonally dominant
// (the partitioned kernel then returns at once).  a holds the sweep's
// geometry (for x: the x<->y transposed one, which the row factor cache
// uses) and the original state in a.c0.
// The non-dominant line count of direction dir's current factors as the
// host knows it, or -1 while unknown (the first sweeps after a
// factorisation, and always inside a graph capture, whose sweeps keep the
// device-side choice so that a replay after a refactorisation is right).
int host_ndom(ash3d_tr_ctx *x, int dir) {
  if (x->capturing || !x->lu_ok[dir]) return -1;
  if (x->ndom_state[dir] == 0) {
    if (!x->ndom_h) {
      if (hipHostMalloc((void **)&x->ndom_h, 128, hipHostMallocDefault) != hipSuccess) {
        x->ndom_h = nullptr;
        return -1;
      }
      x->njl_h = x->ndom_h + 4;
      x->rho_h = (unsigned long long *)(x->ndom_h + 8);
      x->rhoj_h = x->rho_h + 3;
    }
    if (!x->ndom_ev[dir] && hipEventCreateWithFlags(&x->ndom_ev[dir], hipEventDisableTiming) != hipSuccess) {
      x->ndom_ev[dir] = nullptr;
      return -1;
    }
    if (hipMemcpyAsync(x->ndom_h + dir, x->lu_ndom[dir], sizeof(int), hipMemcpyDeviceToHost, x->stream) !=
            hipSuccess ||
        hipMemcpyAsync(x->rho_h + dir, x->lu_rho[dir], sizeof(unsigned long long), hipMemcpyDeviceToHost,
                       x->stream) != hipSuccess ||
        hipMemcpyAsync(x->rhoj_h + dir, x->lu_rhoj[dir], sizeof(unsigned long long), hipMemcpyDeviceToHost,
                       x->stream) != hipSuccess ||
        hipMemcpyAsync(x->njl_h + dir, x->lu_njl[dir], sizeof(int), hipMemcpyDeviceToHost, x->stream) !=
            hipSuccess ||
        hipEventRecord(x->ndom_ev[dir], x->stream) != hipSuccess)
      return -1;
    x->ndom_state[dir] = 1;
    return -1;
  }
  if (x->ndom_state[dir] == 1) {
    if (hipEventQuery(x->ndom_ev[dir]) != hipSuccess) return -1;
    x->ndom_state[dir] = 2;
  }
  return x->ndom_h[dir];
}

// Jacobi sweeps for a theta-scheme sweep of y or z lines that would
// otherwise be solved in dgtsv order (partitioned mode, eager steps): the
// K of launch_diff_cn_jacobi (K + 1 sweeps from zero), or -1.  K + 1
// sweeps leave an error of at most rho^(K+1) of the line's largest
// solution value (rho: the largest row ratio (|DL| + |DU|) / |D| of the
// sweep, the max norm of the Jacobi iteration matrix); K is the least
// with rho^(K+1) <= 2^-52 (1 - rho), the rounding of the sweeps growing
// at most by 1 / (1 - rho).  -1 while the factors' ratio is not known on
// the host (the first sweep after a factorisation) and in graph captures.
int jacobi_k(double rho, int max_sweeps) {
  // y and z lines: at most kJacobiMaxIters + 1 sweeps (the kernels' windows)
  const int sw = jacobi_sweeps_needed(rho, kJacobiMaxIters + 1);
  return (sw == 0 || sw > max_sweeps) ? -1 : sw - 1;
}

int jacobi_iters(ash3d_tr_ctx *x, int dir) {
  if (dir == 0 || x->line_solver != ASH3D_LINE_SOLVER_PARTITIONED || x->jacobi_sweeps <= 0) return -1;
  if (host_ndom(x, dir) != 0) return -1;
  double rho;
  std::memcpy(&rho, x->rho_h + dir, sizeof rho);
  return jacobi_k(rho, x->jacobi_sweeps);
}

// x lines of n active cells: whether the lines that need at most
// jacobi_sweeps Jacobi sweeps take them (each line its own number, from the
// factorisation), the partitioned solves taking the others.  Returns the
// most sweeps a line takes (at most the sweeps of the largest ratio among
// the lines that need at most kJacobiMaxSweepsX, capped by jacobi_sweeps), or 0 (no
// Jacobi sweeps: the x-line solves as before).  Only where the partitioned
// solves certainly run for the sweep (every line dominant, n within their
// range).
int jacobi_x_sweeps(ash3d_tr_ctx *x, int n) {
  if (x->line_solver != ASH3D_LINE_SOLVER_PARTITIONED || x->jacobi_sweeps <= 0 || x->jacobi_x == 0) return 0;
  if (n < 3 || n > x->part_max_line || (x->jacobi_x == 1 && n < 256)) return 0;
  if (host_ndom(x, 0) != 0 || x->njl_h[0] == 0) return 0;
  double rho;
  std::memcpy(&rho, x->rhoj_h, sizeof rho);
  const int sw = jacobi_sweeps_needed(rho, kJacobiMaxSweepsX);
  return sw == 0 ? 0 : std::min(sw, x->jacobi_sweeps);
}

// Cell offsets of the partitioned solves of direction dir: the state in
// its own layout, the row factor cache in the sweep's (for x lines the
// x<->y transposed one, a.d)
PartLayout part_layout(const ash3d_tr_ctx *x, int dir, const DiffArgs &a) {
  const Dims &o = x->d;
  PartLayout lay;
  lay.st0 = 1 + o.sy + o.sz;
  if (dir == 0) {
    lay.sta = o.sy; lay.stb = o.sz; lay.str = 1;
    const Dims &t = a.d;  // transposed: cell (j, r, k) at (j+1) + t.sy*(r+1) + t.sz*(k+1)
    lay.lr0 = 1 + t.sy + t.sz; lay.lra = 1; lay.lrb = t.sz; lay.lrr = t.sy;
  } else {
    lay.sta = 1;
    lay.stb = (dir == 1) ? o.sz : o.sy;
    lay.str = (dir == 1) ? o.sy : o.sz;
    lay.lr0 = lay.st0; lay.lra = lay.sta; lay.lrb = lay.stb; lay.lrr = lay.str;
  }
  lay.lr_ncell = a.d.ncell;
  lay.kap = x->kap;
  return lay;
}

// Launches the partitioned solves of a sweep when they apply.  Returns 0
// when they were not launched (the caller runs the dgtsv-order kernel as
// usual), 1 when launched with the choice left to the device (the caller
// launches the dgtsv-order kernel too, a.sel = 2), 2 when launched for a
// sweep known to be dominant (the caller launches nothing else).
int try_cn_partitioned(ash3d_tr_ctx *x, int dir, DiffArgs &a, int *rc, bool adjoint = false) {
  *rc = 0;
  const int n = a.g.rmax - a.g.rmin + 1;
  a.sel = 0;
  // (x lines are contiguous in memory; y and z lines strided; the adjoint's
  // partitioned solves take the lines the forward's took, so that the
  // sweeps the forward ran as Jacobi sweeps get their transpose)
  const int nmax = dir == 0 ? x->part_max_line : x->part_max_line_strided;
  if (x->line_solver != ASH3D_LINE_SOLVER_PARTITIONED || n < 3 || n > nmax) return 0;
  const int hn = host_ndom(x, dir);
  if (hn > 0) return 0;
  const PartLayout lay = part_layout(x, dir, a);
  a.sel = 1;
  hipError_t e = launch_diff_cn_part(x->stream, a, lay, x->part_bins, x->part_shape[dir], adjoint);
  if (e != hipSuccess) *rc = fail(ASH3D_TR_EDEVICE, "partitioned line solve: %s", hipGetErrorString(e));
  a.sel = 2;
  return hn == 0 ? 2 : 1;
}

// Arguments of a diffusion sweep of direction dir over the state ts0
DiffArgs diff_args(ash3d_tr_ctx *x, int dir, double dt, double imp_fac, const int32_t *is_aloft,
                   const int32_t *bounds) {
  DiffArgs a{};
  a.d = x->d;
  a.g = make_line_geom(x->d, dir, bounds);
  a.c0 = x->ts0();
  a.scratch_b = x->ts1();
  a.dt = dt;
  a.theta = imp_fac;
  a.flag = x->flag;
  make_mask(x, is_aloft, &a.aloft);
  a.kap = x->kap;
  a.kk = dir == 0 ? x->kx : (dir == 1 ? x->ky : x->kz);
  a.sig = dir == 0 ? x->sx : (dir == 1 ? x->sy : x->sz);
  return a;
}

// x lines as the y lines of the x<->y transposed grid: dims, line
// geometry, kappa, sigma and K of the transposed copies
void transpose_x_args(const ash3d_tr_ctx *x, const int32_t *bounds, DiffArgs &a) {
  const Dims &d = x->d;
  const int32_t tb[6] = {bounds[2], bounds[3], bounds[0], bounds[1], bounds[4], bounds[5]};
  a.d = make_dims(d.ny, d.nx, d.nz, d.ns);
  a.g = make_line_geom(a.d, 1, tb);
  a.kap = x->tkap;
  a.sig = x->tsx;
  a.kk = x->tkx;
}

// neumann = false: the caller has just run Set_BC(2), whose ghosts are the
// ones the routine's own Neumann pass would write (every ghost equals the
// clamped, x-wrapped interior cell, and no interior cell has changed since)
int sweep_diffuse(ash3d_tr_ctx *x, int dir, int use_cn, double dt, double imp_fac, const int32_t *is_aloft,
                  const int32_t *bounds, bool neumann = true) {
  int rc = check_bounds(x, bounds);
  if (rc) return rc;
  if (dir < 0 || dir > 2) return fail(ASH3D_TR_EINVAL, "dir must be 0, 1 or 2 (got %d)", dir);
  // the routine's own Neumann ghosts of its direction (incl. periodic x)
  if (neumann) {
    ProfScope ps(x, ASH3D_PROF_BC);
    HIPCHK(launch_neumann_dir(x->stream, x->d, dir, x->periodic, x->ts0()));
  }
  DiffArgs a = diff_args(x, dir, dt, imp_fac, is_aloft, bounds);
  if (use_cn && a.g.rmax - a.g.rmin + 1 <= 1) {
    // ncells <= 1: the reference zero-fills ts1 and copies nothing (:770, :909)
    x->ts1_state = TS1_ZERO;
    return 0;
  }
  if (dir == 0 && use_cn) {
    // x lines: geometry, K, factors and scratch in the x<->y transposed
    // layout, the state read and written in place through LDS row tiles
    const Dims &d = x->d;
    if ((rc = ensure_tgeom(x))) return rc;
    transpose_x_args(x, bounds, a);
    if ((rc = ensure_lr(x, 0, a)) || (rc = ensure_lu(x, 0, a))) return rc;
    {
      ProfScope ps(x, ASH3D_PROF_DIFFUSE_X);  // one record for the kernels
      const int sw = jacobi_x_sweeps(x, a.g.rmax - a.g.rmin + 1);
      x->jacobi_last[0] = sw;
      if (sw > 0) {
        // the lines that need at most jacobi_sweeps sweeps, then the
        // partitioned solves of the others (which certainly run: every
        // line dominant, n in range)
        a.jline = x->lu_jl[0];
        a.jcap = x->jacobi_sweeps;
        HIPCHK(launch_diff_cn_jacobi_x(x->stream, a, d, x->ts0(), x->kap, x->sx, x->kx));
      }
      const int tp = try_cn_partitioned(x, 0, a, &rc);
      if (rc) return rc;
      if (sw > 0 && tp != 2) return fail(ASH3D_TR_EDEVICE, "x-line Jacobi sweeps without the partitioned solves");
      if (tp != 2) HIPCHK(launch_diff_cn_xrows(x->stream, a, d, x->ts0(), x->nb_diff[0], x->x_tile));
    }
    x->ts1_state = TS1_MIRROR;
    return 0;
  }
  const bool transposed = (dir == 0);
  if (transposed) {
    // explicit x lines as the y lines of the transposed grid (x and y swapped)
    const Dims &d = x->d;
    if ((rc = ensure_tgeom(x))) return rc;
    if (!x->tq && (rc = dev_alloc(&x->tq, x->n_conc()))) return rc;
    {
      ProfScope ps(x, ASH3D_PROF_TRANSPOSE);
      HIPCHK(launch_transpose_xy(x->stream, x->ts0(), x->tq, d.nx + 4, d.ny + 4, (long long)(d.nz + 4) * d.ns));
    }
    const int32_t tb[6] = {bounds[2], bounds[3], bounds[0], bounds[1], bounds[4], bounds[5]};
    a.d = make_dims(d.ny, d.nx, d.nz, d.ns);
    a.g = make_line_geom(a.d, 1, tb);
    a.c0 = x->tq;
    a.kap = x->tkap;
    a.sig = x->tsx;
    a.kk = x->tkx;
  }
  if (use_cn) {
    if ((rc = ensure_lr(x, dir, a)) || (rc = ensure_lu(x, dir, a))) return rc;
    ProfScope ps(x, ASH3D_PROF_DIFFUSE_X + dir);  // one record for both kernels
    const int tp = try_cn_partitioned(x, dir, a, &rc);
    if (rc) return rc;
    const int k = tp == 0 ? jacobi_iters(x, dir) : -1;
    x->jacobi_last[dir] = k + 1;
    if (k >= 0) HIPCHK(launch_diff_cn_jacobi(x->stream, a, x->nb_diff[dir], k));
    else if (tp != 2) HIPCHK(launch_diff_cn(x->stream, a, x->nb_diff[dir]));
  } else {
    ProfScope ps(x, ASH3D_PROF_DIFFUSE_X + dir);
    HIPCHK(launch_diff_explicit(x->stream, a, x->nb_diff[dir]));
  }
  if (transposed) {
    ProfScope ps(x, ASH3D_PROF_TRANSPOSE);
    HIPCHK(launch_transpose_xy(x->stream, x->tq, x->ts0(), x->d.ny + 4, x->d.nx + 4,
                               (long long)(x->d.nz + 4) * x->d.ns));
  }
  x->ts1_state = TS1_MIRROR;
  return 0;
}

int need_adjoint(ash3d_tr_ctx *x) {
  if (!x->lam) return fail(ASH3D_TR_EINVAL, "adjoint buffers not allocated (call ash3d_tr_adj_init)");
  return 0;
}

int copy_state(ash3d_tr_ctx *x, double *dst, const double *src) {
  ProfScope ps(x, ASH3D_PROF_COPY);
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)x->n_conc(), hipMemcpyDeviceToDevice, x->stream));
  return 0;
}

// lambda <- J_adv(qpre)^T lambda for one direction; qpre carries the
// Set_BC(1) ghosts the forward sweep saw.
int adj_advect_sweep(ash3d_tr_ctx *x, int dir, double dt, const int32_t *is_aloft, const int32_t *bounds,
                     double *qpre) {
  AdvArgs a = advect_args(x, dir, dt, is_aloft, bounds, qpre);
  a.periodic = (dir == 0) ? x->periodic : 0;
  ProfScope ps(x, ASH3D_PROF_ADJ_ADVECT);
  HIPCHK(launch_adj_advect(x->stream, a, qpre, x->lam, x->nb_adv[dir]));
  return 0;
}

// lambda <- (A^{-1} B)^T lambda, or the explicit stencil's transpose
int adj_diffuse_sweep(ash3d_tr_ctx *x, int dir, int use_cn, double dt, double imp_fac, const int32_t *is_aloft,
                      const int32_t *bounds) {
  int rc = check_bounds(x, bounds);
  if (rc) return rc;
  if (dir < 0 || dir > 2) return fail(ASH3D_TR_EINVAL, "dir must be 0, 1 or 2 (got %d)", dir);
  DiffArgs a{};
  a.d = x->d;
  a.g = make_line_geom(x->d, dir, bounds);
  a.dt = dt;
  a.theta = imp_fac;
  a.flag = x->flag;
  make_mask(x, is_aloft, &a.aloft);
  a.kap = x->kap;
  a.kk = dir == 0 ? x->kx : (dir == 1 ? x->ky : x->kz);
  a.sig = dir == 0 ? x->sx : (dir == 1 ? x->sy : x->sz);
  if (use_cn && a.g.rmax - a.g.rmin + 1 <= 1) return 0;  // the forward leaves ts0 as it is
  double *lam = x->lam;
  const bool transposed = (dir == 0);
  if (transposed) {
    // x lines as the y lines of the x<->y transposed grid
    const Dims &d = x->d;
    if ((rc = ensure_tgeom(x))) return rc;
    if (!x->tq && (rc = dev_alloc(&x->tq, x->n_conc()))) return rc;
    const int32_t tb[6] = {bounds[2], bounds[3], bounds[0], bounds[1], bounds[4], bounds[5]};
    a.d = make_dims(d.ny, d.nx, d.nz, d.ns);
    a.g = make_line_geom(a.d, 1, tb);
    a.kap = x->tkap;
    a.sig = x->tsx;
    a.kk = x->tkx;
  }
  // partitioned adjoint solves in place on lambda (x lines read along the
  // rows of the original layout); the dgtsv-order kernels below return at
  // once on the device when those ran
  const int *skip = nullptr;
  if (use_cn && x->line_solver == ASH3D_LINE_SOLVER_PARTITIONED) {
    if ((rc = ensure_lr(x, dir, a)) || (rc = ensure_lu(x, dir, a))) return rc;
    a.c0 = x->lam;
    ProfScope ps(x, ASH3D_PROF_ADJ_DIFFUSE);
    const int n = a.g.rmax - a.g.rmin + 1;
    if (dir == 0) {
      // the transpose of the forward's Jacobi sweeps for its x lines
      // (lambda in place, original layout); the transposed dgtsv-order
      // solves below skip those lines
      const int sw = jacobi_x_sweeps(x, n);
      x->jacobi_last[0] = sw;
      if (sw > 0) {
        a.jline = x->lu_jl[0];
        a.jcap = x->jacobi_sweeps;
        HIPCHK(launch_diff_cn_jacobi_x(x->stream, a, x->d, x->lam, x->kap, x->sx, x->kx, true));
      }
    }
    const int tp = try_cn_partitioned(x, dir, a, &rc, true);
    if (rc) return rc;
    if (tp == 2) return 0;  // known dominant: nothing else to run
    if (tp == 1) skip = a.ndom;
    if (dir != 0 && tp == 0) {
      // the transpose of the forward's Jacobi sweeps for its y or z lines
      // (where the forward took them: lines outside the partitioned range)
      const int k = (n < 3 || n > x->part_max_line_strided) ? jacobi_iters(x, dir) : -1;
      x->jacobi_last[dir] = k + 1;
      if (k >= 0) {
        HIPCHK(launch_diff_cn_jacobi(x->stream, a, x->nb_diff[dir], k, true));
        return 0;
      }
    }
  }
  if (transposed) {
    const Dims &d = x->d;
    ProfScope ps(x, ASH3D_PROF_TRANSPOSE);
    HIPCHK(launch_transpose_xy(x->stream, x->lam, x->tq, d.nx + 4, d.ny + 4, (long long)(d.nz + 4) * d.ns, skip));
    lam = x->tq;
  }
  {
    ProfScope ps(x, ASH3D_PROF_ADJ_DIFFUSE);
    if (use_cn) {
      if ((rc = ensure_fac(x, x->nb_diff[dir])) || (rc = ensure_lr(x, dir, a))) return rc;
      a.fac = x->fac;
      a.scratch_b = x->adj_sb;
      HIPCHK(launch_adj_cn(x->stream, a, lam, x->nb_diff[dir]));
    } else {
      HIPCHK(launch_adj_diff_explicit(x->stream, a, lam, x->nb_diff[dir], (dir == 0) ? x->periodic : 0));
    }
  }
  if (transposed) {
    ProfScope ps(x, ASH3D_PROF_TRANSPOSE);
    HIPCHK(launch_transpose_xy(x->stream, x->tq, x->lam, x->d.ny + 4, x->d.nx + 4,
                               (long long)(x->d.nz + 4) * x->d.ns, skip));
  }
  return 0;
}

int check_flag(ash3d_tr_ctx *x) {
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, x->flag, sizeof(int), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (h) {
    int z = 0;
    HIPCHK(hipMemcpy(x->flag, &z, sizeof(int), hipMemcpyHostToDevice));
    return fail(ASH3D_TR_ESINGULAR, "tridiagonal solve hit a zero pivot (dgtsv INFO=%d)", h);
  }
  return 0;
}

}  // namespace

extern "C" {

const char *ash3d_tr_last_error(void) { return g_err.c_str(); }

const char *ash3d_tr_build_info(void) {
  return "ash3d_tr: HIP, gfx950, FP64, -O3 -ffp-contract=off, ABI " "1";
}

int ash3d_tr_create(const ash3d_tr_desc *desc, ash3d_tr_ctx **out) {
  if (!desc || !out) return fail(ASH3D_TR_EINVAL, "null descriptor or output pointer");
  *out = nullptr;
  if (desc->abi_version != ASH3D_TR_ABI_VERSION)
    return fail(ASH3D_TR_EINVAL, "ABI version %d, library has %d", desc->abi_version, ASH3D_TR_ABI_VERSION);
  if (desc->nx < 2 || desc->ny < 1 || desc->nz < 1 || desc->ns < 1)
    return fail(ASH3D_TR_EINVAL, "bad grid %dx%dx%d with %d bins", desc->nx, desc->ny, desc->nz, desc->ns);
  if (desc->ns > kMaxBins) return fail(ASH3D_TR_EINVAL, "at most %d grain bins (got %d)", kMaxBins, desc->ns);
  if (desc->limiter < 0 || desc->limiter > 6) return fail(ASH3D_TR_EINVAL, "unknown limiter %d", desc->limiter);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(ASH3D_TR_ENOLIB, "no HIP device");
  if (desc->device < 0 || desc->device >= ndev) return fail(ASH3D_TR_EINVAL, "device %d of %d", desc->device, ndev);
  auto *x = new ash3d_tr_ctx();
  x->device = desc->device;
  x->host = *desc;
  x->d = make_dims(desc->nx, desc->ny, desc->nz, desc->ns);
  // the fused advection pass pays off on large grids (measured: 9 % less
  // advection time on 1440x720x80, 5 % more on the 120-400 x 80-300 x
  // 40-60 ensemble grids)
  x->adv_fused = (long long)desc->nx * desc->ny * desc->nz >= (1LL << 24) ? 1 : 0;
  x->periodic = desc->is_periodic ? 1 : 0;
  x->zscaling = desc->zscaling_id;
  x->limiter = desc->limiter;
  for (int k = 0; k < 3; ++k) {
    x->nb_adv[k] = pick_bins_per_thread(desc->ns, 0, k);
    x->nb_diff[k] = pick_bins_per_thread(desc->ns, 1, k);
  }
  int rc = 0;
  if (hipSetDevice(x->device) != hipSuccess || hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess) {
    delete x;
    return fail(ASH3D_TR_EDEVICE, "cannot create a stream on device %d", desc->device);
  }
  if ((rc = resolve_scratch(x->device, &x->scratch))) {
    std::string msg = g_err;
    ash3d_tr_destroy(x);
    g_err = msg;
    return rc;
  }
  const Dims &d = x->d;
  if ((rc = dev_alloc(&x->c, 2 * x->n_conc())) || (rc = dev_alloc(&x->vx, d.ncell)) ||
      (rc = dev_alloc(&x->vy, d.ncell)) || (rc = dev_alloc(&x->vz, d.ncell)) || (rc = dev_alloc(&x->vh, d.ncell)) ||
      (rc = dev_alloc(&x->vf, x->n_conc())) || (rc = dev_alloc(&x->kap, d.ncell)) ||
      (rc = dev_alloc(&x->sx, d.ncell)) || (rc = dev_alloc(&x->sy, d.ncell)) || (rc = dev_alloc(&x->sz, d.ncell)) ||
      (rc = dev_alloc(&x->kx, x->n_k())) || (rc = dev_alloc(&x->ky, x->n_k())) || (rc = dev_alloc(&x->kz, x->n_k())) ||
      (rc = dev_alloc(&x->ddx, x->n_2d())) || (rc = dev_alloc(&x->ddy, x->n_2d())) ||
      (rc = dev_alloc(&x->ofxz1, x->n_xz())) || (rc = dev_alloc(&x->ofxz2, x->n_xz())) ||
      (rc = dev_alloc(&x->ofyz1, x->n_yz())) || (rc = dev_alloc(&x->ofyz2, x->n_yz())) ||
      (rc = dev_alloc(&x->ofxy1, x->n_xy())) || (rc = dev_alloc(&x->ofxy2, x->n_xy())) ||
      (rc = dev_alloc(&x->dep, x->n_dep())) || (rc = dev_alloc(&x->flux, (long long)(d.nz + 1) * d.ns))) {
    std::string msg = g_err;
    ash3d_tr_destroy(x);
    g_err = msg;
    return rc;
  }
  if (hipMalloc((void **)&x->flag, sizeof(int)) != hipSuccess || hipMemset(x->flag, 0, sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&x->junk, sizeof(double) * kJunk) != hipSuccess) {
    ash3d_tr_destroy(x);
    return fail(ASH3D_TR_ENOMEM, "flag allocation failed");
  }
  rc = ash3d_tr_sync_from_host(x, ASH3D_SYNC_ALL);
  if (rc == 0) {
    hipError_t e = hipStreamSynchronize(x->stream);
    if (e != hipSuccess) rc = fail(ASH3D_TR_EDEVICE, "upload: %s", hipGetErrorString(e));
  }
  if (rc) {
    std::string msg = g_err;
    ash3d_tr_destroy(x);
    g_err = msg;
    return rc;
  }
  *out = x;
  return 0;
}

int ash3d_tr_destroy(ash3d_tr_ctx *x) {
  if (!x) return 0;
  hipSetDevice(x->device);
  if (x->stream) hipStreamSynchronize(x->stream);
  double *ptrs[] = {x->c, x->vx, x->vy, x->vz, x->vh, x->vf, x->kap, x->sx, x->sy, x->sz, x->kx, x->ky, x->kz,
                    x->ddx, x->ddy, x->ofxz1, x->ofxz2, x->ofyz1, x->ofyz2, x->ofxy1, x->ofxy2, x->dep,
                    x->fac, x->flux, x->tot, x->tq, x->tkap, x->tsx, x->tkx, x->lam, x->snap_b,
                    x->snap_z, x->adj_sb, x->sink, x->src_grad, x->col, x->lr[0], x->lr[1], x->lr[2],
                    x->lu[0], x->lu[1], x->lu[2], x->junk, x->prod, x->dtvc};
  for (double *p : ptrs)
    if (p) hipFree(p);
  for (double *p : x->ckpt)
    if (p) hipFree(p);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 4; ++b)
      if (x->met[a][b]) hipFree(x->met[a][b]);
  if (x->dtp) hipFree(x->dtp);
  for (int k = 0; k < 2; ++k)
    if (x->graph[k]) hipGraphExecDestroy(x->graph[k]);
  if (x->prof) hipFree(x->prof);
  for (int k = 0; k < ash3d_tr_ctx::kFluxRing; ++k) {
    if (x->flux_pinned[k]) hipHostFree(x->flux_pinned[k]);
    if (x->flux_ev[k]) hipEventDestroy(x->flux_ev[k]);
  }
  if (x->flag) hipFree(x->flag);
  if (x->ndom_h) hipHostFree(x->ndom_h);
  if (x->umb_src) hipFree(x->umb_src);
  if (x->umb_wind) hipFree(x->umb_wind);
  for (int k = 0; k < 2; ++k) {
    if (x->umb_pin[k]) hipHostFree(x->umb_pin[k]);
    if (x->umb_ev[k]) hipEventDestroy(x->umb_ev[k]);
  }
  for (int k = 0; k < 3; ++k)
    if (x->ndom_ev[k]) hipEventDestroy(x->ndom_ev[k]);
  for (auto &e : x->ev_pool) {
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  if (x->stream) hipStreamDestroy(x->stream);
  delete x;
  return 0;
}

int ash3d_tr_sync_from_host(ash3d_tr_ctx *x, uint32_t mask) {
  int rc = check_ctx(x);
  if (rc) return rc;
  const ash3d_tr_desc &h = x->host;
  const Dims &d = x->d;
  if (mask & ASH3D_SYNC_CONCEN) {
    if ((rc = upload(x, x->c, h.concen_pd, 2 * x->n_conc(), 0.0))) return rc;
    x->ts1_state = TS1_EXPLICIT;
  }
  if (mask & (ASH3D_SYNC_WINDS | ASH3D_SYNC_VF)) x->vel_bc_ok = false;
  if (mask & ASH3D_SYNC_WINDS) {
    if ((rc = upload(x, x->vx, h.vx_pd, d.ncell, 0.0)) || (rc = upload(x, x->vy, h.vy_pd, d.ncell, 0.0)) ||
        (rc = upload(x, x->vz, h.vz_pd, d.ncell, 0.0)) || (rc = upload(x, x->vh, h.vh_pd, d.ncell, 0.0)))
      return rc;
  }
  if (mask & ASH3D_SYNC_VF)
    if ((rc = upload(x, x->vf, h.vf_pd, x->n_conc(), 0.0))) return rc;
  if (mask & ASH3D_SYNC_GEOMETRY) x->dtv_ok = false;
  if (mask & (ASH3D_SYNC_GEOMETRY | ASH3D_SYNC_DIFFK)) {
    x->tgeom_valid = false;
    x->lr_ok[0] = x->lr_ok[1] = x->lr_ok[2] = false;
  }
  if (mask & ASH3D_SYNC_GEOMETRY) {
    if ((rc = upload(x, x->kap, h.kappa_pd, d.ncell, 1.0)) || (rc = upload(x, x->sx, h.sigma_nx_pd, d.ncell, 1.0)) ||
        (rc = upload(x, x->sy, h.sigma_ny_pd, d.ncell, 1.0)) || (rc = upload(x, x->sz, h.sigma_nz_pd, d.ncell, 1.0)))
      return rc;
  }
  if (mask & ASH3D_SYNC_DIFFK) {
    if ((rc = upload(x, x->kx, h.kx, x->n_k(), 0.0)) || (rc = upload(x, x->ky, h.ky, x->n_k(), 0.0)) ||
        (rc = upload(x, x->kz, h.kz, x->n_k(), 0.0)))
      return rc;
  }
  if (mask & ASH3D_SYNC_TOPO) {
    if ((rc = upload(x, x->ddx, h.deldxond_cc, x->n_2d(), 0.0)) || (rc = upload(x, x->ddy, h.deldyond_cc, x->n_2d(), 0.0)))
      return rc;
  }
  if (mask & ASH3D_SYNC_OUTFLOW) {
    if ((rc = upload(x, x->ofxz1, h.outflow_xz1_pd, x->n_xz(), 0.0)) ||
        (rc = upload(x, x->ofxz2, h.outflow_xz2_pd, x->n_xz(), 0.0)) ||
        (rc = upload(x, x->ofyz1, h.outflow_yz1_pd, x->n_yz(), 0.0)) ||
        (rc = upload(x, x->ofyz2, h.outflow_yz2_pd, x->n_yz(), 0.0)) ||
        (rc = upload(x, x->ofxy1, h.outflow_xy1_pd, x->n_xy(), 0.0)) ||
        (rc = upload(x, x->ofxy2, h.outflow_xy2_pd, x->n_xy(), 0.0)))
      return rc;
  }
  if (mask & ASH3D_SYNC_DEPOSIT)
    if ((rc = upload(x, x->dep, h.deposit_granularity, x->n_dep(), 0.0))) return rc;
  // pageable host memory: make the copies complete before the caller may
  // reuse its buffers
  HIPCHK(hipStreamSynchronize(x->stream));
  return 0;
}

int ash3d_tr_sync_to_host(ash3d_tr_ctx *x, uint32_t mask) {
  int rc = check_ctx(x);
  if (rc) return rc;
  const ash3d_tr_desc &h = x->host;
  const Dims &d = x->d;
  if (mask & ASH3D_SYNC_CONCEN) {
    // materialise ts1; the recorded state stays valid (ts1 now holds
    // exactly what the record says), so captured steps may still follow
    if (x->ts1_state == TS1_MIRROR) {
      HIPCHK(launch_mirror_interior(x->stream, d, x->ts0(), x->ts1()));
    } else if (x->ts1_state == TS1_ZERO) {
      HIPCHK(launch_fill(x->stream, x->ts1(), x->n_conc(), 0.0));
    }
    if ((rc = download(x, h.concen_pd, x->c, 2 * x->n_conc()))) return rc;
  }
  if (mask & ASH3D_SYNC_WINDS) {
    if ((rc = download(x, h.vx_pd, x->vx, d.ncell)) || (rc = download(x, h.vy_pd, x->vy, d.ncell)) ||
        (rc = download(x, h.vz_pd, x->vz, d.ncell)) || (rc = download(x, h.vh_pd, x->vh, d.ncell)))
      return rc;
  }
  if (mask & ASH3D_SYNC_VF)
    if ((rc = download(x, h.vf_pd, x->vf, x->n_conc()))) return rc;
  if (mask & ASH3D_SYNC_GEOMETRY) {
    if ((rc = download(x, h.kappa_pd, x->kap, d.ncell)) || (rc = download(x, h.sigma_nx_pd, x->sx, d.ncell)) ||
        (rc = download(x, h.sigma_ny_pd, x->sy, d.ncell)) || (rc = download(x, h.sigma_nz_pd, x->sz, d.ncell)))
      return rc;
  }
  if (mask & ASH3D_SYNC_DIFFK) {
    if ((rc = download(x, h.kx, x->kx, x->n_k())) || (rc = download(x, h.ky, x->ky, x->n_k())) ||
        (rc = download(x, h.kz, x->kz, x->n_k())))
      return rc;
  }
  if (mask & ASH3D_SYNC_TOPO) {
    if ((rc = download(x, h.deldxond_cc, x->ddx, x->n_2d())) || (rc = download(x, h.deldyond_cc, x->ddy, x->n_2d())))
      return rc;
  }
  if (mask & ASH3D_SYNC_OUTFLOW) {
    if ((rc = download(x, h.outflow_xz1_pd, x->ofxz1, x->n_xz())) ||
        (rc = download(x, h.outflow_xz2_pd, x->ofxz2, x->n_xz())) ||
        (rc = download(x, h.outflow_yz1_pd, x->ofyz1, x->n_yz())) ||
        (rc = download(x, h.outflow_yz2_pd, x->ofyz2, x->n_yz())) ||
        (rc = download(x, h.outflow_xy1_pd, x->ofxy1, x->n_xy())) ||
        (rc = download(x, h.outflow_xy2_pd, x->ofxy2, x->n_xy())))
      return rc;
  }
  if (mask & ASH3D_SYNC_DEPOSIT)
    if ((rc = download(x, h.deposit_granularity, x->dep, x->n_dep()))) return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  return check_flag(x);
}

int ash3d_tr_synchronize(ash3d_tr_ctx *x) {
  int rc = check_ctx(x);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  return check_flag(x);
}

int ash3d_tr_set_bc(ash3d_tr_ctx *x, int32_t bc_code) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (bc_code == 1) return set_bc1(x, true);
  if (bc_code == 2) {
    HIPCHK(launch_set_bc2(x->stream, x->d, x->periodic, x->ts0()));
    return 0;
  }
  return fail(ASH3D_TR_EBC, "ERROR BC code not recognized (%d)", bc_code);
}

int ash3d_tr_advect_x(ash3d_tr_ctx *x, double dt, const int32_t *is_aloft, const int32_t bounds[6]) {
  int rc = check_ctx(x);
  return rc ? rc : sweep_advect(x, 0, dt, is_aloft, bounds, true, true);
}

int ash3d_tr_advect_y(ash3d_tr_ctx *x, double dt, const int32_t *is_aloft, const int32_t bounds[6]) {
  int rc = check_ctx(x);
  return rc ? rc : sweep_advect(x, 1, dt, is_aloft, bounds, true, true);
}

int ash3d_tr_advect_z(ash3d_tr_ctx *x, double dt, const int32_t *is_aloft, const int32_t bounds[6]) {
  int rc = check_ctx(x);
  return rc ? rc : sweep_advect(x, 2, dt, is_aloft, bounds, true, true);
}

int ash3d_tr_advect_horz(ash3d_tr_ctx *x, int32_t itoggle, double dt, const int32_t *is_aloft) {
  int rc = check_ctx(x);
  if (rc) return rc;
  int32_t b[6];
  full_bounds(x, b);
  const int first = (itoggle % 2 == 0) ? 1 : 0;
  if ((rc = sweep_advect(x, first, dt, is_aloft, b, true, true))) return rc;
  return sweep_advect(x, 1 - first, dt, is_aloft, b, true, false);
}

int ash3d_tr_advect_vert(ash3d_tr_ctx *x, double dt, const int32_t *is_aloft) {
  int rc = check_ctx(x);
  if (rc) return rc;
  int32_t b[6];
  full_bounds(x, b);
  return sweep_advect(x, 2, dt, is_aloft, b, true, true);
}

int ash3d_tr_diffuse_dir(ash3d_tr_ctx *x, int32_t dir, int32_t use_cn, double dt, double imp_fac,
                         const int32_t *is_aloft, const int32_t bounds[6]) {
  int rc = check_ctx(x);
  return rc ? rc : sweep_diffuse(x, dir, use_cn, dt, imp_fac, is_aloft, bounds);
}

int ash3d_tr_diffuse_horz(ash3d_tr_ctx *x, int32_t itoggle, int32_t use_cn, double dt, double imp_fac,
                          const int32_t *is_aloft) {
  int rc = check_ctx(x);
  if (rc) return rc;
  int32_t b[6];
  full_bounds(x, b);
  const int first = (itoggle % 2 == 0) ? 0 : 1;  // even: x then y (:187-199)
  if ((rc = sweep_diffuse(x, first, use_cn, dt, imp_fac, is_aloft, b))) return rc;
  return sweep_diffuse(x, 1 - first, use_cn, dt, imp_fac, is_aloft, b);
}

int ash3d_tr_diffuse_vert(ash3d_tr_ctx *x, int32_t use_cn, double dt, double imp_fac, const int32_t *is_aloft) {
  int rc = check_ctx(x);
  if (rc) return rc;
  int32_t b[6];
  full_bounds(x, b);
  return sweep_diffuse(x, 2, use_cn, dt, imp_fac, is_aloft, b);
}

// DiffuseHorz / DiffuseVert of one step on ts0 (src/Ash3d.F90:496-508)
static int step_diffusion(ash3d_tr_ctx *x, int32_t itime, double dt, int32_t use_cn, double imp_fac,
                          const int32_t *is_aloft, const int32_t *b) {
  int rc;
  const int dfirst = (itime % 2 == 0) ? 0 : 1;
  {
    // Set_BC(2) where it is read or survives: the first direction's ghosts
    // of the interior rows (the later Neumann passes rewrite the rest)
    ProfScope ps(x, ASH3D_PROF_BC);
    HIPCHK(launch_set_bc2_face(x->stream, x->d, x->periodic, dfirst, x->ts0()));
  }
  if ((rc = sweep_diffuse(x, dfirst, use_cn, dt, imp_fac, is_aloft, b, false))) return rc;
  if ((rc = sweep_diffuse(x, 1 - dfirst, use_cn, dt, imp_fac, is_aloft, b))) return rc;
  return sweep_diffuse(x, 2, use_cn, dt, imp_fac, is_aloft, b);
}

int ash3d_tr_step(ash3d_tr_ctx *x, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                  double imp_fac, const int32_t *is_aloft) {
  int rc = check_ctx_keep(x);
  if (rc) return rc;
  int32_t b[6];
  full_bounds(x, b);
  // Set_BC(1) (src/Ash3d.F90:491), then AdvectHorz / AdvectVert.  Their own
  // Set_BC(1) calls repeat velocity ghosts that cannot have changed, and
  // re-zero concentration ghosts that the sweeps never write; only the
  // periodic x copies change, so only a periodic grid repeats the pass.
  // With diffusion the concentration part of the pass is not launched:
  // the sweeps read the ghost cells along their lines as the zeros it
  // would write, and the diffusion's Set_BC(2) and Neumann passes rewrite
  // every ghost before anything else reads it.
  // With periodic x the x ghosts the x sweep reads are the periodic copies
  // of the state it starts from: one pass over the x ghosts of the
  // interior rows right before it (the Set_BC(2) rule there is the same
  // copy) stands in for the full passes.
  const bool zg = use_diffusion != 0;
  if ((rc = set_bc1(x, true, nullptr, !zg))) return rc;
  const int first = (itime % 2 == 0) ? 1 : 0;
  const bool again = x->periodic != 0 && !zg;
  const int order[3] = {first, 1 - first, 2};
  if (zg && x->adv_fused) {
    // the first sweep into the ts1 slot, then the second horizontal and
    // the vertical sweep in one pass from ts1 into ts0
    if ((rc = ensure_dtv(x, dt))) return rc;
    if (x->capturing) x->graph_dtv[itime % 2] = dt;
    if (x->periodic && first == 0) {
      ProfScope ps(x, ASH3D_PROF_BC);
      HIPCHK(launch_set_bc2_face(x->stream, x->d, 1, 0, x->ts0()));
    }
    {
      AdvArgs a = fused_first_args(x, first, dt, is_aloft, b);
      ProfScope ps(x, ASH3D_PROF_ADVECT_X + first);
      HIPCHK(launch_advect(x->stream, a, x->nb_adv[first]));
    }
    {
      AdvArgs h, z;
      fused_hz_args(x, 1 - first, dt, is_aloft, b, &h, &z);
      ProfScope ps(x, ASH3D_PROF_ADVECT_HZ);
      HIPCHK(launch_advect_hz(x->stream, h, z));
    }
    x->ts1_state = TS1_MIRROR;
  } else {
    for (int q = 0; q < 3; ++q) {
      const int dir = order[q];
      if (zg && x->periodic && dir == 0) {
        ProfScope ps(x, ASH3D_PROF_BC);
        HIPCHK(launch_set_bc2_face(x->stream, x->d, 1, 0, x->ts0()));
      }
      if ((rc = sweep_advect(x, dir, dt, is_aloft, b, q > 0 && again, false, nullptr, nullptr, true, zg))) return rc;
    }
  }
  if (use_diffusion && (rc = step_diffusion(x, itime, dt, use_cn, imp_fac, is_aloft, b))) return rc;
  // every ts0 ghost has been rewritten (Set_BC(1), or the diffusion's
  // Set_BC(2) and Neumann passes); a step that failed part-way leaves the
  // glue's zeroing pending for the next entry point
  x->ghost_zero_pending = false;
  return 0;
}

int ash3d_tr_step_keep(ash3d_tr_ctx *x, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                       double imp_fac, const int32_t *is_aloft, int32_t slot_b, int32_t slot_z) {
  int rc = check_ctx_keep(x);
  if (rc) return rc;
  const int ns = (int)x->ckpt.size();
  if (slot_b < 0 || slot_b >= ns || slot_z < 0 || slot_z >= ns || slot_b == slot_z)
    return fail(ASH3D_TR_EINVAL, "snapshot slots %d, %d of %d (two different slots)", slot_b, slot_z, ns);
  int32_t b[6];
  full_bounds(x, b);
  // ash3d_tr_step's advection, each sweep out of place: ts0 -> slot_b ->
  // slot_z -> ts0 (the same kernels and Set_BC(1) passes, so the same bits)
  if ((rc = set_bc1(x, true))) return rc;
  x->ghost_zero_pending = false;  // Set_BC(1) above rewrote every ts0 ghost
  const int first = (itime % 2 == 0) ? 1 : 0;
  const bool again = x->periodic != 0;
  double *sb = x->ckpt[slot_b], *sz = x->ckpt[slot_z];
  if ((rc = sweep_advect(x, first, dt, is_aloft, b, false, false, x->ts0(), sb, false))) return rc;
  if ((rc = sweep_advect(x, 1 - first, dt, is_aloft, b, again, false, sb, sz, false))) return rc;
  if ((rc = sweep_advect(x, 2, dt, is_aloft, b, again, false, sz, x->ts0(), false))) return rc;
  x->ts1_state = TS1_MIRROR;
  if (use_diffusion) return step_diffusion(x, itime, dt, use_cn, imp_fac, is_aloft, b);
  return 0;
}

int ash3d_tr_step_glue(ash3d_tr_ctx *x) {
  int rc = check_ctx(x);
  if (rc) return rc;
  ProfScope ps(x, ASH3D_PROF_GLUE);
  if (x->ts1_state == TS1_MIRROR) {
    x->ghost_zero_pending = true;  // launched by the next entry point unless a step comes first
  } else if (x->ts1_state == TS1_ZERO) {
    HIPCHK(launch_fill(x->stream, x->ts0(), x->n_conc(), 0.0));
  } else {
    HIPCHK(hipMemcpyAsync(x->ts0(), x->ts1(), sizeof(double) * (size_t)x->n_conc(), hipMemcpyDeviceToDevice,
                          x->stream));
  }
  x->ts1_state = TS1_ZERO;
  return 0;
}

int ash3d_tr_source_column_add(ash3d_tr_ctx *x, int32_t ivent, int32_t jvent, double dt, const double *flux) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (!flux) return fail(ASH3D_TR_EINVAL, "null flux");
  if (ivent < 1 || ivent > x->d.nx || jvent < 1 || jvent > x->d.ny)
    return fail(ASH3D_TR_EINVAL, "vent (%d,%d) outside the grid", ivent, jvent);
  const size_t bytes = sizeof(double) * (size_t)(x->d.nz + 1) * x->d.ns;
  const int k = x->flux_slot;
  if (!x->flux_pinned[k]) {
    HIPCHK(hipHostMalloc((void **)&x->flux_pinned[k], bytes, hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&x->flux_ev[k], hipEventDisableTiming));
  }
  if (x->flux_used[k]) HIPCHK(hipEventSynchronize(x->flux_ev[k]));  // its last upload has been read
  std::memcpy(x->flux_pinned[k], flux, bytes);
  HIPCHK(hipMemcpyAsync(x->flux, x->flux_pinned[k], bytes, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipEventRecord(x->flux_ev[k], x->stream));
  x->flux_used[k] = true;
  x->flux_slot = (k + 1) % ash3d_tr_ctx::kFluxRing;
  HIPCHK(launch_source_add(x->stream, x->d, x->ts0(), ivent, jvent, dt, x->flux));
  return 0;
}

// Copy n doubles from host memory to dev on the context's stream through
// pinned staging buffer `which` (waits for that buffer's previous upload).
static int upload_staged(ash3d_tr_ctx *x, int which, double *dev, const double *const *parts, const size_t *counts,
                         int nparts) {
  size_t n = 0;
  for (int q = 0; q < nparts; ++q) n += counts[q];
  const size_t bytes = sizeof(double) * n;
  if (x->umb_pin_bytes[which] < bytes) {
    if (x->umb_used[which]) HIPCHK(hipEventSynchronize(x->umb_ev[which]));
    if (x->umb_pin[which]) HIPCHK(hipHostFree(x->umb_pin[which]));
    x->umb_pin[which] = nullptr;
    HIPCHK(hipHostMalloc((void **)&x->umb_pin[which], bytes, hipHostMallocDefault));
    x->umb_pin_bytes[which] = bytes;
    x->umb_used[which] = false;
  }
  if (!x->umb_ev[which]) HIPCHK(hipEventCreateWithFlags(&x->umb_ev[which], hipEventDisableTiming));
  if (x->umb_used[which]) HIPCHK(hipEventSynchronize(x->umb_ev[which]));
  size_t at = 0;
  for (int q = 0; q < nparts; ++q) {
    std::memcpy(x->umb_pin[which] + at, parts[q], sizeof(double) * counts[q]);
    at += counts[q];
  }
  HIPCHK(hipMemcpyAsync(dev, x->umb_pin[which], bytes, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipEventRecord(x->umb_ev[which], x->stream));
  x->umb_used[which] = true;
  return 0;
}

int ash3d_tr_umbrella_source_add(ash3d_tr_ctx *x, int32_t ivent, int32_t jvent, double dt, int32_t ibase,
                                 int32_t itop, int32_t n_gs_max, const double *avg_stencil,
                                 const double *flux_umb) {
  int rc = check_ctx(x);
  if (rc) return rc;
  const Dims &d = x->d;
  if (!avg_stencil || !flux_umb) return fail(ASH3D_TR_EINVAL, "null stencil or flux");
  if (ivent < 1 || ivent > d.nx || jvent < 1 || jvent > d.ny)
    return fail(ASH3D_TR_EINVAL, "vent (%d,%d) outside the grid", ivent, jvent);
  if (itop > d.nz || (itop >= ibase && ibase < 1))
    return fail(ASH3D_TR_EINVAL, "umbrella levels %d..%d outside 1..nz", ibase, itop);
  if (n_gs_max < 0 || n_gs_max > d.ns) return fail(ASH3D_TR_EINVAL, "n_gs_max %d outside 0..ns", n_gs_max);
  const size_t nst = 9 * (size_t)d.nz, nfu = 9 * (size_t)(d.nz + 1) * d.ns;
  if (!x->umb_src && (rc = dev_alloc(&x->umb_src, (long long)(nst + nfu)))) return rc;
  const double *parts[2] = {avg_stencil, flux_umb};
  const size_t counts[2] = {nst, nfu};
  if ((rc = upload_staged(x, 0, x->umb_src, parts, counts, 2))) return rc;
  HIPCHK(launch_umbrella_source(x->stream, d, x->ts0(), ivent, jvent, dt, ibase, itop, n_gs_max, x->umb_src,
                                x->umb_src + nst));
  return 0;
}

int ash3d_tr_umbrella_winds_add(ash3d_tr_ctx *x, int32_t ibase, int32_t itop, const double *uvx,
                                const double *uvy) {
  int rc = check_ctx_keep(x);
  if (rc) return rc;
  const Dims &d = x->d;
  if (!uvx || !uvy) return fail(ASH3D_TR_EINVAL, "null umbrella wind");
  if (itop < ibase) return 0;  // no level brackets the cloud
  if (ibase < 1 || itop > d.nz) return fail(ASH3D_TR_EINVAL, "umbrella levels %d..%d outside 1..nz", ibase, itop);
  const size_t plane = (size_t)d.nx * d.ny;
  if (!x->umb_wind && (rc = dev_alloc(&x->umb_wind, 2 * (long long)plane))) return rc;
  const double *parts[2] = {uvx, uvy};
  const size_t counts[2] = {plane, plane};
  if ((rc = upload_staged(x, 1, x->umb_wind, parts, counts, 2))) return rc;
  x->vel_bc_ok = false;
  HIPCHK(launch_umbrella_winds(x->stream, d, x->vx, x->vy, ibase, itop, x->umb_wind, x->umb_wind + plane));
  return 0;
}

int ash3d_tr_tune(ash3d_tr_ctx *x, int32_t param, int32_t value) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (param == ASH3D_TUNE_LINE_SOLVER) {
    if (value != ASH3D_LINE_SOLVER_EXACT && value != ASH3D_LINE_SOLVER_PARTITIONED)
      return fail(ASH3D_TR_EINVAL, "line solver must be %d (dgtsv order) or %d (partitioned) (got %d)",
                  ASH3D_LINE_SOLVER_EXACT, ASH3D_LINE_SOLVER_PARTITIONED, value);
    x->line_solver = value;
    return 0;
  }
  if (param == ASH3D_TUNE_ADV_GROUPS_Y || param == ASH3D_TUNE_ADV_GROUPS_Z) {
    if (value < 0 || value > 4)
      return fail(ASH3D_TR_EINVAL, "bin groups per workgroup must be 0..4 (got %d)", value);
    x->adv_gy[param == ASH3D_TUNE_ADV_GROUPS_Y ? 1 : 2] = value;
    return 0;
  }
  if (param == ASH3D_TUNE_ADV_FUSED) {
    if (value != 0 && value != 1) return fail(ASH3D_TR_EINVAL, "fused advection must be 0 or 1 (got %d)", value);
    x->adv_fused = value;
    return 0;
  }
  if (param == ASH3D_TUNE_JACOBI_SWEEPS) {
    if (value < 0 || value > kJacobiMaxSweepsX)
      return fail(ASH3D_TR_EINVAL, "Jacobi sweeps must be 0..%d (got %d)", kJacobiMaxSweepsX, value);
    x->jacobi_sweeps = value;
    return 0;
  }
  if (param == ASH3D_TUNE_JACOBI_LAST_Y || param == ASH3D_TUNE_JACOBI_LAST_Z || param == ASH3D_TUNE_JACOBI_LAST_X)
    return fail(ASH3D_TR_EINVAL, "tuning parameter %d can only be read", param);
  if (param == ASH3D_TUNE_JACOBI_X) {
    if (value < 0 || value > 2) return fail(ASH3D_TR_EINVAL, "x-line Jacobi sweeps must be 0, 1 or 2 (got %d)", value);
    x->jacobi_x = value;
    return 0;
  }
  if (param == ASH3D_TUNE_ADV_X_ROWS) {
    if (value != 0 && value != 1)
      return fail(ASH3D_TR_EINVAL, "x rows kernel must be 0 (workgroup rows) or 1 (wavefront rows) (got %d)", value);
    x->adv_xrows = value;
    return 0;
  }
  if (param >= ASH3D_TUNE_PART_SHAPE_X && param <= ASH3D_TUNE_PART_SHAPE_Z) {
    if (value != 0 && value != 1) return fail(ASH3D_TR_EINVAL, "tile shape must be 0 or 1 (got %d)", value);
    x->part_shape[param - ASH3D_TUNE_PART_SHAPE_X] = value;
    return 0;
  }
  if (param == ASH3D_TUNE_PART_MAX_LINE || param == ASH3D_TUNE_PART_MAX_LINE_STRIDED) {
    if (value < 3 || value > 2048) return fail(ASH3D_TR_EINVAL, "longest partitioned line must be 3..2048 (got %d)", value);
    (param == ASH3D_TUNE_PART_MAX_LINE ? x->part_max_line : x->part_max_line_strided) = value;
    return 0;
  }
  if (param == ASH3D_TUNE_PART_BINS) {
    if (value < 0) return fail(ASH3D_TR_EINVAL, "bins per workgroup must be >= 0 (got %d)", value);
    x->part_bins = value;
    return 0;
  }
  if (param == ASH3D_TUNE_X_TILE) {
    if (value != 8 && value != 16 && value != 32)
      return fail(ASH3D_TR_EINVAL, "x-line tile must be 8, 16 or 32 cells (got %d)", value);
    x->x_tile = value;
    return 0;
  }
  const bool adv = param >= ASH3D_TUNE_NB_ADVECT_X && param <= ASH3D_TUNE_NB_ADVECT_Z;
  if (value != 1 && value != 2 && !(adv && value == 3) && value != 4 && value != 8)
    return fail(ASH3D_TR_EINVAL, "bins per thread must be 1, 2, 4 or 8, or 3 for the advection (got %d)", value);
  if (adv) {
    x->nb_adv[param - ASH3D_TUNE_NB_ADVECT_X] = value;
  } else if (param >= ASH3D_TUNE_NB_DIFFUSE_X && param <= ASH3D_TUNE_NB_DIFFUSE_Z) {
    HIPCHK(hipStreamSynchronize(x->stream));  // the CN scratch is sized by it
    x->nb_diff[param - ASH3D_TUNE_NB_DIFFUSE_X] = value;
  } else {
    return fail(ASH3D_TR_EINVAL, "unknown tuning parameter %d", param);
  }
  return 0;
}

int ash3d_tr_tune_get(ash3d_tr_ctx *x, int32_t param, int32_t *value) {
  int rc = check_ctx_keep(x);
  if (rc) return rc;
  if (!value) return fail(ASH3D_TR_EINVAL, "null output");
  switch (param) {
    case ASH3D_TUNE_NB_ADVECT_X: case ASH3D_TUNE_NB_ADVECT_Y: case ASH3D_TUNE_NB_ADVECT_Z:
      *value = x->nb_adv[param - ASH3D_TUNE_NB_ADVECT_X]; return 0;
    case ASH3D_TUNE_NB_DIFFUSE_X: case ASH3D_TUNE_NB_DIFFUSE_Y: case ASH3D_TUNE_NB_DIFFUSE_Z:
      *value = x->nb_diff[param - ASH3D_TUNE_NB_DIFFUSE_X]; return 0;
    case ASH3D_TUNE_X_TILE: *value = x->x_tile; return 0;
    case ASH3D_TUNE_LINE_SOLVER: *value = x->line_solver; return 0;
    case ASH3D_TUNE_PART_BINS: *value = x->part_bins; return 0;
    case ASH3D_TUNE_PART_SHAPE_X: case ASH3D_TUNE_PART_SHAPE_Y: case ASH3D_TUNE_PART_SHAPE_Z:
      *value = x->part_shape[param - ASH3D_TUNE_PART_SHAPE_X]; return 0;
    case ASH3D_TUNE_ADV_GROUPS_Y: *value = x->adv_gy[1]; return 0;
    case ASH3D_TUNE_ADV_GROUPS_Z: *value = x->adv_gy[2]; return 0;
    case ASH3D_TUNE_ADV_X_ROWS: *value = x->adv_xrows; return 0;
    case ASH3D_TUNE_PART_MAX_LINE: *value = x->part_max_line; return 0;
    case ASH3D_TUNE_PART_MAX_LINE_STRIDED: *value = x->part_max_line_strided; return 0;
    case ASH3D_TUNE_ADV_FUSED: *value = x->adv_fused; return 0;
    case ASH3D_TUNE_JACOBI_SWEEPS: *value = x->jacobi_sweeps; return 0;
    case ASH3D_TUNE_JACOBI_LAST_Y: *value = x->jacobi_last[1]; return 0;
    case ASH3D_TUNE_JACOBI_LAST_Z: *value = x->jacobi_last[2]; return 0;
    case ASH3D_TUNE_JACOBI_X: *value = x->jacobi_x; return 0;
    case ASH3D_TUNE_JACOBI_LAST_X: *value = x->jacobi_last[0]; return 0;
    default: return fail(ASH3D_TR_EINVAL, "unknown tuning parameter %d", param);
  }
}

int ash3d_tr_profile(ash3d_tr_ctx *x, int32_t enable) {
  int rc = check_ctx(x);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  x->prof_on = enable != 0;
  if (enable) {
    x->ev_used = 0;
    x->ev_rec.clear();
  }
  return 0;
}

int ash3d_tr_profile_read(ash3d_tr_ctx *x, int32_t kind, double *total_ms, int64_t *launches) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (!total_ms || !launches) return fail(ASH3D_TR_EINVAL, "null output");
  HIPCHK(hipStreamSynchronize(x->stream));
  double t = 0.0;
  int64_t n = 0;
  for (auto &r : x->ev_rec) {
    if (r.first != kind) continue;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, x->ev_pool[r.second].first, x->ev_pool[r.second].second));
    t += ms;
    ++n;
  }
  *total_ms = t;
  *launches = n;
  return 0;
}

int ash3d_tr_mass_totals(ash3d_tr_ctx *x, double *aloft, double *deposit, double *outflow) {
  int rc = check_ctx(x);
  if (rc) return rc;
  const Dims &d = x->d;
  const long long np = mass_partials_size(d);
  if (!x->tot && (rc = dev_alloc(&x->tot, np + 3LL * d.ns))) return rc;
  double *out = x->tot + np;
  HIPCHK(launch_mass_totals(x->stream, d, x->ts0(), x->kap, x->dep, x->ofyz1, x->ofyz2, x->ofxz1, x->ofxz2,
                            x->ofxy2, x->tot, out));
  std::vector<double> h(3 * (size_t)d.ns);
  HIPCHK(hipMemcpyAsync(h.data(), out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  for (int n = 0; n < d.ns; ++n) {
    if (aloft) aloft[n] = h[n];
    if (deposit) deposit[n] = h[d.ns + n];
    if (outflow) outflow[n] = h[2 * d.ns + n];
  }
  return 0;
}

int ash3d_tr_mass_totals_noalloc(ash3d_tr_ctx *x, double *out) {
  int rc = check_ctx(x);
  if (rc) return rc;
  const Dims &d = x->d;
  if (!out) return fail(ASH3D_TR_EINVAL, "null output");
  if (d.ns > ASH3D_TR_NOALLOC_MAX_BINS)
    return fail(ASH3D_TR_EINVAL, "ash3d_tr_mass_totals_noalloc handles at most %d bins (the context has %d)",
                ASH3D_TR_NOALLOC_MAX_BINS, d.ns);
  if (mass_partials_size(d) > kNoallocScratch)
    return fail(ASH3D_TR_EINVAL, "mass-total partials of %lld doubles exceed the scratch of %lld",
                mass_partials_size(d), kNoallocScratch);
  if (!x->scratch) return fail(ASH3D_TR_EDEVICE, "allocation-free scratch not resolved on device %d", x->device);
  // every partial is written by k_mass_partials before k_mass_final reads
  // it on the same stream, so nothing needs clearing
  HIPCHK(launch_mass_totals(x->stream, d, x->ts0(), x->kap, x->dep, x->ofyz1, x->ofyz2, x->ofxz1, x->ofxz2,
                            x->ofxy2, x->scratch, out));
  return 0;
}

int ash3d_tr_checkpoints(ash3d_tr_ctx *x, int32_t nslots) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (nslots < 0) return fail(ASH3D_TR_EINVAL, "negative checkpoint count %d", nslots);
  HIPCHK(hipStreamSynchronize(x->stream));
  while ((int)x->ckpt.size() > nslots) {
    HIPCHK(hipFree(x->ckpt.back()));
    x->ckpt.pop_back();
  }
  while ((int)x->ckpt.size() < nslots) {
    double *p = nullptr;
    if ((rc = dev_alloc(&p, x->n_conc()))) return rc;
    x->ckpt.push_back(p);
  }
  return 0;
}

int ash3d_tr_checkpoint(ash3d_tr_ctx *x, int32_t slot, int32_t op) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (slot < 0 || slot >= (int)x->ckpt.size())
    return fail(ASH3D_TR_EINVAL, "checkpoint slot %d of %d", slot, (int)x->ckpt.size());
  if (op == ASH3D_CKPT_SAVE) return copy_state(x, x->ckpt[slot], x->ts0());
  if (op == ASH3D_CKPT_LOAD) {
    if ((rc = copy_state(x, x->ts0(), x->ckpt[slot]))) return rc;
    x->ts1_state = TS1_MIRROR;
    return 0;
  }
  return fail(ASH3D_TR_EINVAL, "unknown checkpoint operation %d", op);
}

int ash3d_tr_adj_init(ash3d_tr_ctx *x, int32_t n_source_slots) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (n_source_slots < 0) return fail(ASH3D_TR_EINVAL, "negative source slot count %d", n_source_slots);
  const Dims &d = x->d;
  const long long face = std::max(x->n_yz(), std::max(x->n_xz(), x->n_xy()));
  if (!x->lam) {
    if ((rc = dev_alloc(&x->lam, x->n_conc())) || (rc = dev_alloc(&x->snap_b, x->n_conc())) ||
        (rc = dev_alloc(&x->snap_z, x->n_conc())) || (rc = dev_alloc(&x->adj_sb, x->n_conc())) ||
        (rc = dev_alloc(&x->sink, 2 * face + x->n_dep())))
      return rc;
    HIPCHK(launch_fill(x->stream, x->sink, 2 * face + x->n_dep(), 0.0));
  }
  const long long per = (long long)(d.nz + 1) * d.ns;
  if (n_source_slots > x->n_src_slots || !x->src_grad) {
    HIPCHK(hipStreamSynchronize(x->stream));
    if (x->src_grad) HIPCHK(hipFree(x->src_grad));
    x->src_grad = nullptr;
    x->n_src_slots = 0;
    if ((rc = dev_alloc(&x->src_grad, per * std::max(1, (int)n_source_slots)))) return rc;
  }
  x->n_src_slots = n_source_slots;
  HIPCHK(launch_fill(x->stream, x->lam, x->n_conc(), 0.0));
  HIPCHK(launch_fill(x->stream, x->src_grad, per * std::max(1, (int)n_source_slots), 0.0));
  return 0;
}

int ash3d_tr_adj_set(ash3d_tr_ctx *x, const double *lambda) {
  int rc = check_ctx(x);
  if (rc || (rc = need_adjoint(x))) return rc;
  if (!lambda) {
    HIPCHK(launch_fill(x->stream, x->lam, x->n_conc(), 0.0));
    return 0;
  }
  HIPCHK(hipMemcpyAsync(x->lam, lambda, sizeof(double) * (size_t)x->n_conc(), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(launch_zero_ghosts(x->stream, x->d, x->lam));
  return 0;
}

int ash3d_tr_adj_get(ash3d_tr_ctx *x, double *lambda) {
  int rc = check_ctx(x);
  if (rc || (rc = need_adjoint(x))) return rc;
  if (!lambda) return fail(ASH3D_TR_EINVAL, "null output");
  HIPCHK(hipMemcpyAsync(lambda, x->lam, sizeof(double) * (size_t)x->n_conc(), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return 0;
}

int ash3d_tr_adj_advect(ash3d_tr_ctx *x, int32_t dir, double dt, const int32_t *is_aloft, const int32_t bounds[6]) {
  int rc = check_ctx(x);
  if (rc || (rc = need_adjoint(x)) || (rc = check_bounds(x, bounds))) return rc;
  if (dir < 0 || dir > 2) return fail(ASH3D_TR_EINVAL, "dir must be 0, 1 or 2 (got %d)", dir);
  if (!(dt >= 0.0)) return fail(ASH3D_TR_EINVAL, "dt must be >= 0 (got %g)", dt);
  if ((rc = set_bc1(x, true))) return rc;
  return adj_advect_sweep(x, dir, dt, is_aloft, bounds, x->ts0());
}

int ash3d_tr_adj_diffuse(ash3d_tr_ctx *x, int32_t dir, int32_t use_cn, double dt, double imp_fac,
                         const int32_t *is_aloft, const int32_t bounds[6]) {
  int rc = check_ctx(x);
  if (rc || (rc = need_adjoint(x))) return rc;
  return adj_diffuse_sweep(x, dir, use_cn, dt, imp_fac, is_aloft, bounds);
}

static int adj_step_impl(ash3d_tr_ctx *x, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                         double imp_fac, const int32_t *is_aloft, double *kept_b, double *kept_z);

int ash3d_tr_adj_step(ash3d_tr_ctx *x, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                      double imp_fac, const int32_t *is_aloft) {
  return adj_step_impl(x, itime, dt, use_diffusion, use_cn, imp_fac, is_aloft, nullptr, nullptr);
}

int ash3d_tr_adj_step_kept(ash3d_tr_ctx *x, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                           double imp_fac, const int32_t *is_aloft, int32_t slot_b, int32_t slot_z) {
  int rc = check_ctx(x);
  if (rc) return rc;
  const int ns = (int)x->ckpt.size();
  if (slot_b < 0 || slot_b >= ns || slot_z < 0 || slot_z >= ns || slot_b == slot_z)
    return fail(ASH3D_TR_EINVAL, "snapshot slots %d, %d of %d (two different slots)", slot_b, slot_z, ns);
  return adj_step_impl(x, itime, dt, use_diffusion, use_cn, imp_fac, is_aloft, x->ckpt[slot_b], x->ckpt[slot_z]);
}

static int adj_step_impl(ash3d_tr_ctx *x, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                         double imp_fac, const int32_t *is_aloft, double *kept_b, double *kept_z) {
  int rc = check_ctx(x);
  if (rc || (rc = need_adjoint(x))) return rc;
  if (!(dt >= 0.0)) return fail(ASH3D_TR_EINVAL, "dt must be >= 0 (got %g)", dt);
  int32_t b[6];
  full_bounds(x, b);
  if (use_diffusion && use_cn && x->d.nz <= 1) {
    // diffCN_z with one layer zero-fills ts1 and copies nothing: the step's
    // output (ts1) is zero whatever its input
    HIPCHK(launch_fill(x->stream, x->lam, x->n_conc(), 0.0));
    return 0;
  }
  // the inputs of the second horizontal sweep and of the vertical sweep:
  // kept by ash3d_tr_step_keep, or recomputed from the step input (ts0) as
  // ash3d_tr_step computes them
  if ((rc = set_bc1(x, true))) return rc;
  const int first = (itime % 2 == 0) ? 1 : 0;
  double *qb = kept_b ? kept_b : x->snap_b, *qz = kept_z ? kept_z : x->snap_z;
  if (!kept_b) {
    if ((rc = sweep_advect(x, first, dt, is_aloft, b, false, false, x->ts0(), x->snap_b))) return rc;
    if (x->periodic && (rc = set_bc1(x, false, x->snap_b))) return rc;
    if ((rc = sweep_advect(x, 1 - first, dt, is_aloft, b, false, false, x->snap_b, x->snap_z))) return rc;
    if (x->periodic && (rc = set_bc1(x, false, x->snap_z))) return rc;
  }
  // adjoint sweeps in reverse order (Set_BC(2) ghosts are not read by the
  // diffusion sweeps, which set their own)
  if (use_diffusion) {
    const int dfirst = (itime % 2 == 0) ? 0 : 1;
    if ((rc = adj_diffuse_sweep(x, 2, use_cn, dt, imp_fac, is_aloft, b))) return rc;
    if ((rc = adj_diffuse_sweep(x, 1 - dfirst, use_cn, dt, imp_fac, is_aloft, b))) return rc;
    if ((rc = adj_diffuse_sweep(x, dfirst, use_cn, dt, imp_fac, is_aloft, b))) return rc;
  }
  if ((rc = adj_advect_sweep(x, 2, dt, is_aloft, b, qz))) return rc;
  if ((rc = adj_advect_sweep(x, 1 - first, dt, is_aloft, b, qb))) return rc;
  return adj_advect_sweep(x, first, dt, is_aloft, b, x->ts0());
}

int ash3d_tr_adj_source(ash3d_tr_ctx *x, int32_t ivent, int32_t jvent, double dt, int32_t slot) {
  int rc = check_ctx(x);
  if (rc || (rc = need_adjoint(x))) return rc;
  if (ivent < 1 || ivent > x->d.nx || jvent < 1 || jvent > x->d.ny)
    return fail(ASH3D_TR_EINVAL, "vent (%d,%d) outside the grid", ivent, jvent);
  if (slot < 0 || slot >= x->n_src_slots)
    return fail(ASH3D_TR_EINVAL, "source slot %d of %d", slot, x->n_src_slots);
  HIPCHK(launch_adj_source(x->stream, x->d, x->lam, ivent, jvent, dt,
                           x->src_grad + (long long)slot * (x->d.nz + 1) * x->d.ns));
  return 0;
}

int ash3d_tr_adj_source_read(ash3d_tr_ctx *x, double *out) {
  int rc = check_ctx(x);
  if (rc || (rc = need_adjoint(x))) return rc;
  if (!out) return fail(ASH3D_TR_EINVAL, "null output");
  const size_t n = (size_t)x->n_src_slots * (size_t)(x->d.nz + 1) * x->d.ns;
  if (n) HIPCHK(hipMemcpyAsync(out, x->src_grad, sizeof(double) * n, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return 0;
}

static int ensure_col(ash3d_tr_ctx *x) {
  if (x->col) return 0;
  return dev_alloc(&x->col, (long long)x->d.nz + (long long)x->d.nx * x->d.ny);
}

int ash3d_tr_column_load(ash3d_tr_ctx *x, const double *dz, double *out) {
  int rc = check_ctx(x);
  if (rc || (rc = ensure_col(x))) return rc;
  if (!dz || !out) return fail(ASH3D_TR_EINVAL, "null dz or output");
  const Dims &d = x->d;
  const size_t nxy = (size_t)d.nx * d.ny;
  if (x->ts1_state == TS1_ZERO) {  // ConcentrationCalculator reads ts1
    std::memset(out, 0, sizeof(double) * nxy);
    return 0;
  }
  const double *src = (x->ts1_state == TS1_EXPLICIT) ? x->ts1() : x->ts0();
  HIPCHK(hipMemcpyAsync(x->col, dz, sizeof(double) * (size_t)d.nz, hipMemcpyHostToDevice, x->stream));
  HIPCHK(launch_column_load(x->stream, d, src, x->col, x->col + d.nz));
  HIPCHK(hipMemcpyAsync(out, x->col + d.nz, sizeof(double) * nxy, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return 0;
}

int ash3d_tr_products(ash3d_tr_ctx *x, const double *dz, const double *zcc, const int32_t *bounds,
                      const ash3d_tr_product_params *prm, ash3d_tr_products_out *out) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (!dz || !zcc || !prm || !out) return fail(ASH3D_TR_EINVAL, "null dz, z_cc, params or output");
  const Dims &d = x->d;
  const long long nxy = (long long)d.nx * d.ny;
  int32_t b[4] = {1, d.nx, 1, d.ny};
  if (bounds) {
    for (int k = 0; k < 4; ++k) b[k] = bounds[k];
    if (b[0] < 1 || b[1] > d.nx || b[2] < 1 || b[3] > d.ny)
      return fail(ASH3D_TR_EINVAL, "product bounds outside the grid");
  }
  if (!x->prod && (rc = dev_alloc(&x->prod, 6 * nxy + 2 * (long long)d.nz))) return rc;
  double *dzd = x->prod + 6 * nxy, *zcd = dzd + d.nz;
  HIPCHK(hipStreamSynchronize(x->stream));  // the staging buffer may still be read
  HIPCHK(hipMemcpyAsync(dzd, dz, sizeof(double) * (size_t)d.nz, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(zcd, zcc, sizeof(double) * (size_t)d.nz, hipMemcpyHostToDevice, x->stream));
  ProdArgs p{};
  p.d = d;
  // ConcentrationCalculator reads ts1
  p.c = (x->ts1_state == TS1_ZERO) ? nullptr : (x->ts1_state == TS1_EXPLICIT) ? x->ts1() : x->ts0();
  p.dep = x->dep;
  p.dz = dzd;
  p.zcc = zcd;
  p.cloudcon = prm->cloudcon_thresh;
  p.cloudload = prm->cloudload_thresh;
  p.density = prm->deposit_density;
  p.depo = prm->depo_thresh;
  p.i0 = b[0]; p.i1 = b[1]; p.j0 = b[2]; p.j1 = b[3];
  p.out = x->prod;
  HIPCHK(launch_products(x->stream, p));
  double *planes[5] = {out->cloud_load, out->max_concentration, out->max_height, out->min_height,
                       out->deposit_thickness};
  for (int k = 0; k < 5; ++k)
    if (planes[k])
      HIPCHK(hipMemcpyAsync(planes[k], x->prod + k * nxy, sizeof(double) * (size_t)nxy, hipMemcpyDeviceToHost,
                            x->stream));
  std::vector<int32_t> masks((size_t)(2 * nxy));
  HIPCHK(hipMemcpyAsync(masks.data(), x->prod + 5 * nxy, sizeof(int32_t) * (size_t)(2 * nxy),
                        hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (out->mask_cloud) std::memcpy(out->mask_cloud, masks.data(), sizeof(int32_t) * (size_t)nxy);
  if (out->mask_deposit) std::memcpy(out->mask_deposit, masks.data() + nxy, sizeof(int32_t) * (size_t)nxy);
  // CloudArea (i outer, j inner over the box) and DepositAreaCovered (whole
  // grid), sigma_nz_pd(i,j,1) from the host geometry
  const double *sz = x->host.sigma_nz_pd;
  double ca = 0.0, da = 0.0;
  for (int i = 1; i <= d.nx; ++i)
    for (int j = 1; j <= d.ny; ++j) {
      const long long t = (long long)(i - 1) + (long long)d.nx * (j - 1);
      const double area = sz ? sz[d.p3(i, j, 1)] : 0.0;
      if (i >= b[0] && i <= b[1] && j >= b[2] && j <= b[3] && masks[(size_t)t]) ca = ca + area;
      if (masks[(size_t)(nxy + t)]) da = da + area;
    }
  out->cloud_area = ca;
  out->deposit_area = da;
  return 0;
}

int ash3d_tr_write_esri_ascii(const char *path, int32_t nx, int32_t ny, const double *field, const int32_t *mask,
                              const char *fill_value, int32_t is_latlon, double xll, double yll, double dx,
                              double dy) {
  if (!path || !field || !fill_value) return fail(ASH3D_TR_EINVAL, "null path, field or fill value");
  if (nx < 1 || ny < 1) return fail(ASH3D_TR_EINVAL, "bad field size %d x %d", nx, ny);
  const std::string text = esri_ascii_text(nx, ny, field, mask, fill_value, is_latlon, xll, yll, dx, dy);
  FILE *f = fopen(path, "wb");
  if (!f) return fail(ASH3D_TR_EINVAL, "cannot open %s for writing", path);
  const size_t w = fwrite(text.data(), 1, text.size(), f);
  const int c = fclose(f);
  if (w != text.size() || c != 0) return fail(ASH3D_TR_EINVAL, "writing %s failed", path);
  return 0;
}

int ash3d_tr_esri_ascii_name(const char *root, double time_h, int32_t is_final, char *out, int32_t outlen) {
  if (!root || !out || outlen < 1) return fail(ASH3D_TR_EINVAL, "null root or output");
  const std::string n = esri_ascii_name(root, time_h, is_final);
  if ((int32_t)n.size() + 1 > outlen) return fail(ASH3D_TR_EINVAL, "name needs %d bytes", (int)n.size() + 1);
  std::memcpy(out, n.c_str(), n.size() + 1);
  return 0;
}

int ash3d_tr_adj_column_load(ash3d_tr_ctx *x, const double *dz, const double *w) {
  int rc = check_ctx(x);
  if (rc || (rc = need_adjoint(x)) || (rc = ensure_col(x))) return rc;
  if (!dz || !w) return fail(ASH3D_TR_EINVAL, "null dz or weights");
  const Dims &d = x->d;
  HIPCHK(hipStreamSynchronize(x->stream));  // the staging buffer may still be read
  HIPCHK(hipMemcpyAsync(x->col, dz, sizeof(double) * (size_t)d.nz, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(x->col + d.nz, w, sizeof(double) * (size_t)d.nx * d.ny, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(launch_adj_column(x->stream, d, x->lam, x->col, x->col + d.nz));
  return 0;
}

int ash3d_tr_graph_capture(ash3d_tr_ctx *x, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                           double imp_fac, const int32_t *is_aloft) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (x->ts1_state != TS1_MIRROR)
    return fail(ASH3D_TR_EINVAL, "capture after an eager step (the glue must be the ghost-zeroing kernel)");
  if (x->prof_on) return fail(ASH3D_TR_EINVAL, "kernel timing is on; switch it off before capturing");
  if (use_diffusion && use_cn && !(x->lu[0] && x->lu[1] && x->lu[2] && x->junk))
    return fail(ASH3D_TR_EINVAL, "capture after an eager step (line-solve factors not allocated)");
  if (use_diffusion && !x->tgeom_valid)
    return fail(ASH3D_TR_EINVAL, "capture after an eager step (transposed geometry not built)");
  HIPCHK(hipStreamSynchronize(x->stream));
  const int slot = itime % 2;
  if (x->graph[slot]) {
    HIPCHK(hipGraphExecDestroy(x->graph[slot]));
    x->graph[slot] = nullptr;
  }
  for (int k = 0; k < 3; ++k) x->lu_args_set[k] = false;
  x->graph_dtv[slot] = -1.0;
  if (use_diffusion && x->adv_fused && (rc = ensure_dtv(x, dt))) return rc;  // (outside the capture)
  HIPCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
  x->capturing = true;
  int rc2 = ash3d_tr_step_glue(x);
  if (!rc2) rc2 = ash3d_tr_step(x, itime, dt, use_diffusion, use_cn, imp_fac, is_aloft);
  x->capturing = false;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(x->stream, &g);
  if (rc2) {
    if (g) hipGraphDestroy(g);
    return rc2;
  }
  if (e != hipSuccess) return fail(ASH3D_TR_EDEVICE, "stream capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&x->graph[slot], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return fail(ASH3D_TR_EDEVICE, "graph instantiation: %s", hipGetErrorString(e));
  for (int k = 0; k < 3; ++k) {
    x->graph_lu_on[slot][k] = x->lu_args_set[k];
    x->graph_lu[slot][k] = x->lu_args[k];
  }
  x->graph_tgeom[slot] = use_diffusion != 0;
  return 0;
}

int ash3d_tr_graph_launch(ash3d_tr_ctx *x, int32_t itime) {
  int rc = check_ctx(x);
  if (rc) return rc;
  const int slot = itime % 2;
  if (!x->graph[slot]) return fail(ASH3D_TR_EINVAL, "no graph captured for step parity %d", slot);
  if (x->ts1_state != TS1_MIRROR) return fail(ASH3D_TR_EINVAL, "the state changed since the capture (sync or load)");
  // the graph's sweeps read cached geometry and line factors: rebuild
  // (eagerly, ahead of the graph on the stream) what a geometry/K upload
  // or an eager step with another dt or range replaced since the capture
  if (x->graph_tgeom[slot] && (rc = ensure_tgeom(x))) return rc;
  if (x->graph_dtv[slot] >= 0.0 && (rc = ensure_dtv(x, x->graph_dtv[slot]))) return rc;
  for (int k = 0; k < 3; ++k) {
    if (!x->graph_lu_on[slot][k]) continue;
    DiffArgs a = x->graph_lu[slot][k];
    if ((rc = ensure_lr(x, k, a)) || (rc = ensure_lu(x, k, a))) return rc;
  }
  HIPCHK(hipGraphLaunch(x->graph[slot], x->stream));
  x->ts1_state = TS1_MIRROR;
  return 0;
}

int ash3d_tr_met_push(ash3d_tr_ctx *x, const float *vx, const float *vy, const float *vz, const float *vf) {
  int rc = check_ctx_keep(x);
  if (rc) return rc;
  if (!vx || !vy || !vz) return fail(ASH3D_TR_EINVAL, "null met-step wind");
  const Dims &d = x->d;
  const size_t n3 = (size_t)d.nx * d.ny * d.nz;
  if (!x->met[0][0]) {
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 4; ++b) {
        const size_t n = (b == 3) ? n3 * d.ns : n3;
        hipError_t e = hipMalloc((void **)&x->met[a][b], sizeof(float) * n);
        if (e != hipSuccess) return fail(ASH3D_TR_ENOMEM, "met-step allocation: %s", hipGetErrorString(e));
      }
  }
  const int slot = 1 - x->met_next;  // the old "last": overwritten by the new "next"
  const float *src[4] = {vx, vy, vz, vf};
  for (int b = 0; b < 4; ++b) {
    if (!src[b]) continue;
    const size_t n = (b == 3) ? n3 * d.ns : n3;
    HIPCHK(hipMemcpyAsync(x->met[slot][b], src[b], sizeof(float) * n, hipMemcpyHostToDevice, x->stream));
  }
  HIPCHK(hipStreamSynchronize(x->stream));  // pageable sources
  x->met_vf[slot] = vf != nullptr;
  x->met_next = slot;
  ++x->met_count;
  return 0;
}

int ash3d_tr_met_interp(ash3d_tr_ctx *x, double frac, int32_t use_vf) {
  int rc = check_ctx_keep(x);
  if (rc) return rc;
  if (x->met_count < 2) return fail(ASH3D_TR_EINVAL, "two met steps are needed (have %d)", x->met_count);
  const int nxt = x->met_next, lst = 1 - nxt;
  if (use_vf && !(x->met_vf[0] && x->met_vf[1])) return fail(ASH3D_TR_EINVAL, "fall velocities were not pushed");
  const float *last[4] = {x->met[lst][0], x->met[lst][1], x->met[lst][2], x->met[lst][3]};
  const float *next[4] = {x->met[nxt][0], x->met[nxt][1], x->met[nxt][2], x->met[nxt][3]};
  x->vel_bc_ok = false;
  HIPCHK(launch_met_interp(x->stream, x->d, last, next, frac, use_vf != 0, x->vx, x->vy, x->vz, x->vf));
  return 0;
}

int ash3d_tr_adjust_dt(ash3d_tr_ctx *x, const ash3d_tr_dt_params *p, double *dt_out, double *diag) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (!p || !dt_out || !p->dz) return fail(ASH3D_TR_EINVAL, "null parameters, dz or output");
  if (p->use_met_next && x->met_count < 1) return fail(ASH3D_TR_EINVAL, "no met step pushed");
  if (p->use_met_next && !x->met_vf[x->met_next]) return fail(ASH3D_TR_EINVAL, "the next met step has no vf");
  const Dims &d = x->d;
  const int nb = dt_partial_blocks(d);
  if (!x->dtp && (rc = dev_alloc(&x->dtp, 4LL * 2048 + d.nz))) return rc;
  double *dzd = x->dtp + 4LL * 2048;
  HIPCHK(hipMemcpyAsync(dzd, p->dz, sizeof(double) * (size_t)d.nz, hipMemcpyHostToDevice, x->stream));
  const int nxt = x->met_next;
  const float *mn[4] = {x->met[nxt][0], x->met[nxt][1], x->met[nxt][2], x->met[nxt][3]};
  HIPCHK(launch_dt_partials(x->stream, d, p->is_latlon, p->use_met_next, p->use_diffusion, x->vx, x->vy, x->vz, x->vf,
                            x->kap, x->sx, x->sy, x->sz, x->kx, x->ky, x->kz, dzd, mn, x->dtp));
  std::vector<double> h(4 * (size_t)nb);
  HIPCHK(hipMemcpyAsync(h.data(), x->dtp, sizeof(double) * h.size(), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = -INFINITY;
  for (int b = 0; b < nb; ++b) {
    r0 = std::fmax(r0, h[4 * b]);
    r1 = std::fmax(r1, h[4 * b + 1]);
    r2 = std::fmax(r2, h[4 * b + 2]);
    r3 = std::fmax(r3, h[4 * b + 3]);
  }
  const double MPS_2_KMPHR = 3.6;
  double vxmax_dx = r0, vymax_dy = r1, vzmax_dz = r2;
  if (!p->is_latlon) {  // src/Adjust_dt.f90:100-106, :141-146
    vxmax_dx = p->use_met_next ? r0 * MPS_2_KMPHR / p->dx : r0 / p->dx;
    vymax_dy = p->use_met_next ? r1 * MPS_2_KMPHR / p->dy : r1 / p->dy;
  }
  if (p->n_gs_aloft == 0) vzmax_dz = 0.0;  // :176
  const double time_advect = 1.0 / std::fmax(std::fmax(vxmax_dx, vymax_dy), vzmax_dz);
  double time_diffuse = p->dt_max;
  if (p->use_diffusion) {  // :182-246
    if (time_diffuse > -r3) time_diffuse = -r3;
    time_diffuse = time_diffuse * (p->use_cn ? p->imp_dt_fac : 0.5);
  }
  const double dt_tmp = std::fmin(time_advect, time_diffuse);  // :296-333
  double dt;
  if (dt_tmp < p->dt_min) dt = p->dt_min;
  else if (dt_tmp > p->dt_max) dt = p->dt_max;
  else dt = std::fmin(1.0, p->cfl) * dt_tmp;
  dt = p->dt_min * (double)(int)(dt / p->dt_min);
  if ((p->next_write_time - p->time) > p->eps_small && (p->next_write_time - p->time) < dt)
    dt = p->next_write_time - p->time;
  if (p->time + dt > p->simtime) dt = p->simtime - p->time;
  *dt_out = dt;
  if (diag) {
    diag[0] = vxmax_dx; diag[1] = vymax_dy; diag[2] = vzmax_dz; diag[3] = time_diffuse;
  }
  return 0;
}

static int minmax_index(ash3d_tr_ctx *x, int32_t ivent, int32_t jvent, int32_t kvent, double thresh, int32_t *b) {
  const Dims &d = x->d;
  const int np = d.nx + d.ny + d.nz;
  int rc = 0;
  if (!x->prof) {
    hipError_t e = hipMalloc((void **)&x->prof, sizeof(unsigned long long) * (size_t)np);
    if (e != hipSuccess) return fail(ASH3D_TR_ENOMEM, "profile allocation: %s", hipGetErrorString(e));
  }
  HIPCHK(launch_slab_max(x->stream, d, x->ts0(), x->prof));
  std::vector<double> pm(np);
  HIPCHK(hipMemcpyAsync(pm.data(), x->prof, sizeof(double) * (size_t)np, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  // src/Calc_Mesh.f90:459-541, scanning the slab maxima
  const int N[3] = {d.nx, d.ny, d.nz}, vent[3] = {ivent, jvent, kvent};
  const double *slab[3] = {pm.data(), pm.data() + d.nx, pm.data() + d.nx + d.ny};
  for (int q = 0; q < 3; ++q) {
    int lo = 1, hi = N[q];
    for (int i = 2; i <= vent[q] - 1; ++i) {
      if (slab[q][i - 1] < thresh) lo = i;
      else break;
    }
    for (int i = N[q]; i >= lo + 2; --i) {
      if (slab[q][i - 1] < thresh) hi = i;
      else break;
    }
    if (q == 0 && x->periodic && (lo == 1 || hi == N[0])) { lo = 1; hi = N[0]; }
    b[2 * q] = lo;
    b[2 * q + 1] = hi;
  }
  if (b[5] < b[4]) return fail(ASH3D_TR_EINVAL, "ERROR: kmax<kmin in get_minmax_index");
  for (int q = 0; q < 3; ++q)
    if (b[2 * q + 1] - b[2 * q] < 2) {
      b[2 * q + 1] = std::min(b[2 * q + 1] + 1, N[q]);
      b[2 * q] = b[2 * q + 1] - 2;
    }
  (void)rc;
  return 0;
}

int ash3d_tr_minmax_index(ash3d_tr_ctx *x, int32_t ivent, int32_t jvent, int32_t kvent, double thresh,
                          int32_t bounds[6]) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (!bounds) return fail(ASH3D_TR_EINVAL, "null bounds");
  return minmax_index(x, ivent, jvent, kvent, thresh, bounds);
}

int ash3d_tr_step_subgrid(ash3d_tr_ctx *x, int32_t itime, double dt, int32_t use_diffusion, int32_t use_cn,
                          double imp_fac, const int32_t *is_aloft, int32_t ivent, int32_t jvent, int32_t kvent,
                          double thresh, int32_t bounds[6]) {
  int rc = check_ctx(x);
  if (rc) return rc;
  int32_t b[6];
  if ((rc = set_bc1(x, true))) return rc;  // src/Ash3d.F90:491
  if ((rc = minmax_index(x, ivent, jvent, kvent, thresh, b))) return rc;
  if (bounds)
    for (int q = 0; q < 6; ++q) bounds[q] = b[q];
  const int first = (itime % 2 == 0) ? 1 : 0;
  const bool again = x->periodic != 0;
  if ((rc = sweep_advect(x, first, dt, is_aloft, b, false, false))) return rc;
  if ((rc = sweep_advect(x, 1 - first, dt, is_aloft, b, again, false))) return rc;
  if ((rc = sweep_advect(x, 2, dt, is_aloft, b, again, false))) return rc;
  if (use_diffusion) {
    const int dfirst = (itime % 2 == 0) ? 0 : 1;
    {
      // as in step_diffusion: Set_BC(2) on the first direction's ghosts of
      // the interior rows
      ProfScope ps(x, ASH3D_PROF_BC);
      HIPCHK(launch_set_bc2_face(x->stream, x->d, x->periodic, dfirst, x->ts0()));
    }
    const int32_t g[6] = {std::max(1, b[0] - 1), std::min(x->d.nx, b[1] + 1), std::max(1, b[2] - 1),
                          std::min(x->d.ny, b[3] + 1), std::max(1, b[4] - 1), std::min(x->d.nz, b[5] + 1)};
    if ((rc = sweep_diffuse(x, dfirst, use_cn, dt, imp_fac, is_aloft, g, false))) return rc;
    if ((rc = sweep_diffuse(x, 1 - dfirst, use_cn, dt, imp_fac, is_aloft, g))) return rc;
    if ((rc = sweep_diffuse(x, 2, use_cn, dt, imp_fac, is_aloft, g))) return rc;
  }
  return 0;
}

}  // extern "C"

// ---- batched ensembles ------------------------------------------------------

// One batched launch of an ensemble step: the members (n of them) whose
// launch in this slot of the step selects the same kernel, block and LDS;
// their argument blocks, the member of each workgroup and the members'
// first workgroups at byte offsets args_off, member_off and first_off of
// the ensemble's device buffer.
struct EnsGroup {
  LaunchSpec spec;  // the first member's
  int sg = 0;       // stream group it runs on
  size_t args_off = 0, member_off = 0, first_off = 0;
  int n = 0;
  long long total = 0;  // workgroups along x, all members
  unsigned gy = 1;      // grid.y: the largest of the members'
};

struct ash3d_tr_ens {
  int device = 0;
  // stream groups: the members are split into groups of about equal work,
  // each group's batched launches run on its own stream so that one
  // group's launches fill the others' tails; streams[0] orders the step
  // after the members' earlier work and the members' later work after it
  std::vector<hipStream_t> streams;
  std::vector<int> group_of;  // per member
  std::vector<hipEvent_t> fork, join;  // per stream group
  std::vector<ash3d_tr_ctx *> m;
  std::vector<bool> fused;  // the member's planned advection: the fused pass
  std::vector<double> dt;
  std::vector<std::vector<int32_t>> aloft;  // per member (empty: every bin aloft)
  double imp_fac = 0.5;
  std::vector<EnsGroup> plan[2];  // by step parity
  int member_launches[2] = {0, 0};
  unsigned char *dev = nullptr;  // argument blocks and offsets of both plans
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  hipEvent_t done = nullptr;
  std::vector<hipEvent_t> ready;  // per member
};

namespace ash3d {

hipError_t launch_batch(hipStream_t s, const LaunchSpec &sp, const void *dev_args, const int *dev_member,
                        const long long *dev_first, long long total, unsigned grid_y) {
  if (total <= 0) return hipSuccess;
  if (total > 0x7fffffffLL || grid_y == 0 || grid_y > 65535) return hipErrorInvalidValue;
  void *kargs[] = {(void *)&dev_args, (void *)&dev_member, (void *)&dev_first};
  hipError_t e = hipLaunchKernel(sp.batch, dim3((unsigned)total, grid_y), sp.block, kargs, sp.lds, s);
  return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace ash3d

namespace {

// slots of one glue + step: [x ghosts before x, periodic only] around the
// horizontal advection pair, advection z, Set_BC(2) face, diffusion of the
// first direction, Neumann + diffusion of the second, Neumann + diffusion z
constexpr int kEnsSlots = 10;

const int32_t *ens_aloft(const ash3d_tr_ens *e, size_t k) {
  return e->aloft[k].empty() ? nullptr : e->aloft[k].data();
}

// Brings member x's caches for a step with dt, imp_fac, is_aloft up to date
// on its stream (velocity ghosts, transposed geometry, row and LU factors)
// and checks that the step takes only launches the ensemble batches.
// fused: the member's launches use the fused advection pass (its dt/kappa
// cache is kept current); refactored (optional): set when a line
// factorisation had to be launched (the caller checks dominance again).
int ens_prepare(ash3d_tr_ctx *x, double dt, double imp_fac, const int32_t *is_aloft, bool fused,
                bool *refactored = nullptr) {
  int rc = check_ctx_keep(x);
  if (rc) return rc;
  const long long lu_before = x->lu_count;
  // after a step (ts1 mirrored) the ensemble step's glue only zeroes ghost
  // cells the step rewrites; after a glue (ts1 zero) the member steps from
  // ts0 as it is
  if (x->ts1_state != TS1_MIRROR && x->ts1_state != TS1_ZERO)
    return fail(ASH3D_TR_EINVAL, "ensemble member: the last call must be a step or a step glue");
  if (x->line_solver != ASH3D_LINE_SOLVER_PARTITIONED)
    return fail(ASH3D_TR_EINVAL, "ensemble member: the line solver must be the partitioned one");
  if (!(dt >= 0.0)) return fail(ASH3D_TR_EINVAL, "dt must be >= 0 (got %g)", dt);
  if ((rc = set_bc1(x, true, nullptr, false))) return rc;
  if ((rc = ensure_tgeom(x))) return rc;
  if (fused && (rc = ensure_dtv(x, dt))) return rc;
  int32_t b[6];
  full_bounds(x, b);
  for (int dir = 0; dir < 3; ++dir) {
    DiffArgs a = diff_args(x, dir, dt, imp_fac, is_aloft, b);
    if (dir == 0) transpose_x_args(x, b, a);
    const int n = a.g.rmax - a.g.rmin + 1;
    const int nmax = dir == 0 ? x->part_max_line : x->part_max_line_strided;
    if (n < 3 || n > std::min(nmax, 2048))
      return fail(ASH3D_TR_EINVAL, "ensemble member: %d active cells along dir %d (partitioned solves take 3..%d)",
                  n, dir, std::min(nmax, 2048));
    if ((rc = ensure_lr(x, dir, a)) || (rc = ensure_lu(x, dir, a))) return rc;
  }
  if (refactored) *refactored = x->lu_count != lu_before;
  return 0;
}

// Every line of member x's three directions is a diagonally dominant
// M-matrix row (the partitioned kernels exit at once otherwise); waits for
// the member's stream.
int ens_check_dominant(ash3d_tr_ctx *x, size_t k) {
  HIPCHK(hipStreamSynchronize(x->stream));
  for (int dir = 0; dir < 3; ++dir) {
    int h = 0;
    HIPCHK(hipMemcpy(&h, x->lu_ndom[dir], sizeof(int), hipMemcpyDeviceToHost));
    if (h != 0)
      return fail(ASH3D_TR_EINVAL, "ensemble member %zu: %d lines along dir %d are not diagonally dominant", k, h, dir);
  }
  return 0;
}

// The launches of member x's glue + step of parity itime % 2, one per slot
// (args_size 0 where the member has none), after ens_prepare.
int ens_member_specs(ash3d_tr_ctx *x, int itime, double dt, double imp_fac, const int32_t *is_aloft,
                     LaunchSpec *slots) {
  int rc = 0;
  for (int q = 0; q < kEnsSlots; ++q) slots[q] = LaunchSpec{};
  int32_t b[6];
  full_bounds(x, b);
  int q = 0;
  const int first = (itime % 2 == 0) ? 1 : 0;
  const int order[3] = {first, 1 - first, 2};
  if (x->adv_fused) {
    // the first sweep into ts1 and the fused pass, in the slots of the
    // first and second sweeps of the split advection
    if (first == 0 && x->periodic) spec_set_bc2_face(x->d, 1, 0, x->ts0(), &slots[0]);
    hipError_t e = spec_advect(fused_first_args(x, first, dt, is_aloft, b), x->nb_adv[first], &slots[first == 0 ? 1 : 0]);
    if (e != hipSuccess) return fail(ASH3D_TR_EDEVICE, "ensemble advection dir %d: %s", first, hipGetErrorString(e));
    AdvArgs h, z;
    fused_hz_args(x, 1 - first, dt, is_aloft, b, &h, &z);
    if ((e = spec_advect_hz(h, z, &slots[2])) != hipSuccess)
      return fail(ASH3D_TR_EDEVICE, "ensemble fused advection: %s", hipGetErrorString(e));
    q = 4;
  } else {
    for (int i = 0; i < 3; ++i) {
      const int dir = order[i];
      if (dir == 0) {
        // (periodic x: the x ghosts the sweep reads are the periodic copies)
        if (x->periodic) spec_set_bc2_face(x->d, 1, 0, x->ts0(), &slots[q]);
        ++q;
      }
      AdvArgs a = advect_args(x, dir, dt, is_aloft, b, x->ts0());
      a.zg = (x->periodic && dir == 0) ? 0 : 1;
      hipError_t e = spec_advect(a, x->nb_adv[dir], &slots[q++]);
      if (e != hipSuccess) return fail(ASH3D_TR_EDEVICE, "ensemble advection dir %d: %s", dir, hipGetErrorString(e));
    }
  }
  const int dfirst = (itime % 2 == 0) ? 0 : 1;
  for (int i = 0; i < 3; ++i) {
    const int dir = i == 0 ? dfirst : (i == 1 ? 1 - dfirst : 2);
    if (i == 0) spec_set_bc2_face(x->d, x->periodic, dfirst, x->ts0(), &slots[q++]);
    else spec_neumann_dir(x->d, dir, x->periodic, x->ts0(), &slots[q++]);
    DiffArgs a = diff_args(x, dir, dt, imp_fac, is_aloft, b);
    if (dir == 0) transpose_x_args(x, b, a);
    if ((rc = ensure_lr(x, dir, a)) || (rc = ensure_lu(x, dir, a))) return rc;
    a.sel = 1;
    hipError_t e = spec_diff_cn_part(a, part_layout(x, dir, a), x->part_bins, x->part_shape[dir], &slots[q++]);
    if (e != hipSuccess) return fail(ASH3D_TR_EDEVICE, "ensemble diffusion dir %d: %s", dir, hipGetErrorString(e));
  }
  return q == kEnsSlots ? 0 : fail(ASH3D_TR_EINVAL, "ensemble step slots");
}

// The batched launches of one step on the stream groups, forked from and
// joined back into streams[0] (captured into one graph the same way).
int ens_run_plan(ash3d_tr_ens *e, int p) {
  const size_t ng = e->streams.size();
  HIPCHK(hipEventRecord(e->fork[0], e->streams[0]));
  for (size_t s = 1; s < ng; ++s) HIPCHK(hipStreamWaitEvent(e->streams[s], e->fork[0], 0));
  for (const EnsGroup &g : e->plan[p])
    HIPCHK(launch_batch(e->streams[g.sg], g.spec, e->dev + g.args_off, (const int *)(e->dev + g.member_off),
                        (const long long *)(e->dev + g.first_off), g.total, g.gy));
  for (size_t s = 1; s < ng; ++s) {
    HIPCHK(hipEventRecord(e->join[s], e->streams[s]));
    HIPCHK(hipStreamWaitEvent(e->streams[0], e->join[s], 0));
  }
  return 0;
}

int ens_destroy(ash3d_tr_ens *e) {
  if (!e) return 0;
  hipSetDevice(e->device);
  for (hipStream_t st : e->streams)
    if (st) hipStreamSynchronize(st);
  for (int k = 0; k < 2; ++k)
    if (e->graph[k]) hipGraphExecDestroy(e->graph[k]);
  if (e->dev) hipFree(e->dev);
  if (e->done) hipEventDestroy(e->done);
  for (hipEvent_t v : e->ready)
    if (v) hipEventDestroy(v);
  for (hipEvent_t v : e->fork)
    if (v) hipEventDestroy(v);
  for (hipEvent_t v : e->join)
    if (v) hipEventDestroy(v);
  for (hipStream_t st : e->streams)
    if (st) hipStreamDestroy(st);
  delete e;
  return 0;
}

int ens_build(ash3d_tr_ens *e) {
  int rc = 0;
  const size_t n = e->m.size();
  e->fused.resize(n);
  for (size_t k = 0; k < n; ++k) {
    e->fused[k] = e->m[k]->adv_fused != 0;
    if ((rc = ens_prepare(e->m[k], e->dt[k], e->imp_fac, ens_aloft(e, k), e->fused[k]))) return rc;
  }
  for (size_t k = 0; k < n; ++k)
    if ((rc = ens_check_dominant(e->m[k], k))) return rc;
  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  // stream groups of about equal work (bin-cells), the largest members
  // first; members are laid out largest first inside each launch too
  const int ng = (int)e->streams.size();
  std::vector<size_t> order(n);
  for (size_t k = 0; k < n; ++k) order[k] = k;
  auto work = [&](size_t k) { return e->m[k]->d.ncell * (long long)e->m[k]->d.ns; };
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work(a) > work(b); });
  {
    std::vector<long long> load(ng, 0);
    e->group_of.assign(n, 0);
    for (size_t k : order) {
      const int gsel = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      e->group_of[k] = gsel;
      load[gsel] += work(k);
    }
  }
  std::vector<unsigned char> host;
  std::vector<LaunchSpec> specs(n * kEnsSlots);
  for (int p = 0; p < 2; ++p) {
    const int itime = p == 0 ? 2 : 1;
    for (size_t k = 0; k < n; ++k)
      if ((rc = ens_member_specs(e->m[k], itime, e->dt[k], e->imp_fac, ens_aloft(e, k), &specs[k * kEnsSlots])))
        return rc;
    e->plan[p].clear();
    e->member_launches[p] = 0;
    for (int q = 0; q < kEnsSlots; ++q) {
      // member indices, by stream group, kernel, block and LDS
      std::vector<std::vector<size_t>> groups;
      std::vector<int> group_sg;
      for (size_t k : order) {
        const LaunchSpec &sp = specs[k * kEnsSlots + q];
        if (sp.args_size == 0) continue;
        ++e->member_launches[p];
        if (sp.grid.z != 1 || sp.grid.x == 0) return fail(ASH3D_TR_EINVAL, "ensemble launch shape");
        size_t gi = 0;
        for (; gi < groups.size(); ++gi) {
          const LaunchSpec &o = specs[groups[gi][0] * kEnsSlots + q];
          if (group_sg[gi] == e->group_of[k] && o.batch == sp.batch && o.block.x == sp.block.x &&
              o.block.y == sp.block.y && o.block.z == sp.block.z && o.lds == sp.lds && o.args_size == sp.args_size)
            break;
        }
        if (gi == groups.size()) {
          groups.emplace_back();
          group_sg.push_back(e->group_of[k]);
        }
        groups[gi].push_back(k);
      }
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        const auto &mem = groups[gi];
        EnsGroup g;
        g.sg = group_sg[gi];
        g.spec = specs[mem[0] * kEnsSlots + q];
        g.n = (int)mem.size();
        long long tot = 0;
        for (size_t k : mem) {
          tot += specs[k * kEnsSlots + q].grid.x;
          g.gy = std::max(g.gy, specs[k * kEnsSlots + q].grid.y);
        }
        if (tot > 0x7fffffffLL) return fail(ASH3D_TR_EINVAL, "ensemble launch of %lld workgroups", tot);
        g.total = tot;
        g.args_off = align(host.size());
        g.first_off = align(g.args_off + (size_t)g.n * g.spec.args_size);
        g.member_off = align(g.first_off + sizeof(long long) * (size_t)(g.n + 1));
        host.resize(align(g.member_off + sizeof(int) * (size_t)tot), 0);
        long long f = 0;
        for (int i = 0; i <= g.n; ++i) {
          std::memcpy(&host[g.first_off + sizeof(long long) * i], &f, sizeof f);
          if (i == g.n) break;
          const LaunchSpec &sp = specs[mem[i] * kEnsSlots + q];
          std::memcpy(&host[g.args_off + (size_t)i * sp.args_size], sp.args, sp.args_size);
          for (long long w = 0; w < (long long)sp.grid.x; ++w)
            std::memcpy(&host[g.member_off + sizeof(int) * (size_t)(f + w)], &i, sizeof i);
          f += sp.grid.x;
        }
        e->plan[p].push_back(g);
      }
    }
  }
  HIPCHK(hipMalloc((void **)&e->dev, host.size()));
  HIPCHK(hipMemcpy(e->dev, host.data(), host.size(), hipMemcpyHostToDevice));
  for (int p = 0; p < 2; ++p)
    for (const EnsGroup &g : e->plan[p])
      if (g.spec.lds > 65536)
        HIPCHK(hipFuncSetAttribute(g.spec.batch, hipFuncAttributeMaxDynamicSharedMemorySize, (int)g.spec.lds));
  // one graph per parity (the stream groups as parallel branches)
  for (int p = 0; p < 2; ++p) {
    HIPCHK(hipStreamBeginCapture(e->streams[0], hipStreamCaptureModeThreadLocal));
    int rc2 = ens_run_plan(e, p);
    hipGraph_t gr = nullptr;
    hipError_t er = hipStreamEndCapture(e->streams[0], &gr);
    if (rc2) {
      if (gr) hipGraphDestroy(gr);
      return rc2;
    }
    if (er != hipSuccess) return fail(ASH3D_TR_EDEVICE, "ensemble stream capture: %s", hipGetErrorString(er));
    er = hipGraphInstantiate(&e->graph[p], gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    if (er != hipSuccess) return fail(ASH3D_TR_EDEVICE, "ensemble graph instantiation: %s", hipGetErrorString(er));
  }
  return 0;
}

}  // namespace

extern "C" {

int ash3d_tr_ensemble_create(ash3d_tr_ctx *const *members, int32_t n, const double *dt, double imp_fac,
                             const int32_t *const *is_aloft, int32_t n_streams, ash3d_tr_ens **out) {
  if (!out) return fail(ASH3D_TR_EINVAL, "null output pointer");
  *out = nullptr;
  if (!members || !dt || n < 1) return fail(ASH3D_TR_EINVAL, "ensemble: members and dt of n >= 1 members");
  for (int k = 0; k < n; ++k) {
    if (!members[k]) return fail(ASH3D_TR_EINVAL, "ensemble: null member %d", k);
    if (members[k]->device != members[0]->device) return fail(ASH3D_TR_EINVAL, "ensemble: members on two devices");
    for (int j = 0; j < k; ++j)
      if (members[j] == members[k]) return fail(ASH3D_TR_EINVAL, "ensemble: member %d repeated", k);
  }
  ash3d_tr_ens *e = new ash3d_tr_ens();
  e->device = members[0]->device;
  e->m.assign(members, members + n);
  e->dt.assign(dt, dt + n);
  e->imp_fac = imp_fac;
  e->aloft.resize(n);
  for (int k = 0; k < n; ++k)
    if (is_aloft && is_aloft[k]) e->aloft[k].assign(is_aloft[k], is_aloft[k] + members[k]->d.ns);
  int rc = 0;
  if (n_streams < 0 || n_streams > 16) {
    ens_destroy(e);
    return fail(ASH3D_TR_EINVAL, "ensemble: 0..16 stream groups (got %d)", n_streams);
  }
  const int ng = std::min<int>(n, n_streams > 0 ? n_streams : 4);
  e->streams.assign(ng, nullptr);
  e->fork.assign(ng, nullptr);
  e->join.assign(ng, nullptr);
  bool ok = hipSetDevice(e->device) == hipSuccess &&
            hipEventCreateWithFlags(&e->done, hipEventDisableTiming) == hipSuccess;
  for (int g = 0; ok && g < ng; ++g)
    ok = hipStreamCreateWithFlags(&e->streams[g], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&e->fork[g], hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&e->join[g], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    ens_destroy(e);
    return fail(ASH3D_TR_EDEVICE, "ensemble stream / event creation");
  }
  e->ready.assign(n, nullptr);
  for (int k = 0; k < n; ++k)
    if (hipEventCreateWithFlags(&e->ready[k], hipEventDisableTiming) != hipSuccess) {
      ens_destroy(e);
      return fail(ASH3D_TR_EDEVICE, "ensemble event creation");
    }
  if ((rc = ens_build(e))) {
    std::string msg = g_err;
    ens_destroy(e);
    g_err = msg;
    return rc;
  }
  *out = e;
  return 0;
}

int ash3d_tr_ensemble_step(ash3d_tr_ens *e, int32_t itime, int32_t use_graph) {
  if (!e) return fail(ASH3D_TR_EINVAL, "null ensemble");
  int rc = 0;
  const int p = ((itime % 2) + 2) % 2;
  for (size_t k = 0; k < e->m.size(); ++k) {
    ash3d_tr_ctx *x = e->m[k];
    bool refactored = false;
    if ((rc = ens_prepare(x, e->dt[k], e->imp_fac, ens_aloft(e, k), e->fused[k], &refactored))) return rc;
    if (refactored && (rc = ens_check_dominant(x, k))) return rc;
    HIPCHK(hipEventRecord(e->ready[k], x->stream));
    HIPCHK(hipStreamWaitEvent(e->streams[0], e->ready[k], 0));
  }
  HIPCHK(hipSetDevice(e->device));
  if (use_graph) HIPCHK(hipGraphLaunch(e->graph[p], e->streams[0]));
  else if ((rc = ens_run_plan(e, p))) return rc;
  HIPCHK(hipEventRecord(e->done, e->streams[0]));
  for (ash3d_tr_ctx *x : e->m) {
    HIPCHK(hipStreamWaitEvent(x->stream, e->done, 0));
    // as after ash3d_tr_step_glue + ash3d_tr_step
    x->ts1_state = TS1_MIRROR;
    x->ghost_zero_pending = false;
  }
  return 0;
}

int ash3d_tr_ensemble_synchronize(ash3d_tr_ens *e) {
  if (!e) return fail(ASH3D_TR_EINVAL, "null ensemble");
  HIPCHK(hipSetDevice(e->device));
  for (hipStream_t st : e->streams) HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int ash3d_tr_ensemble_info(ash3d_tr_ens *e, int32_t itime, int32_t *batched_launches, int32_t *member_launches) {
  if (!e) return fail(ASH3D_TR_EINVAL, "null ensemble");
  const int p = ((itime % 2) + 2) % 2;
  if (batched_launches) *batched_launches = (int32_t)e->plan[p].size();
  if (member_launches) *member_launches = e->member_launches[p];
  return 0;
}

int ash3d_tr_ensemble_destroy(ash3d_tr_ens *e) { return ens_destroy(e); }

}  // extern "C"

namespace ash3d {
int report_error(int code, const char *msg) { return fail(code, "%s", msg); }
}  // namespace ash3d
