// Shared definitions of the gfx950 transport-step library: padded-grid
// indexing in the reference's Fortran layout, kernel argument blocks and the
// internal launchers the C ABI (ash3d_tr.hip) drives.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <stdint.h>

#include <string>

namespace ash3d {

// concen_pd(-1:nx+2,-1:ny+2,-1:nz+2,1:ns,0:1) and the other padded arrays
// (src/Ash3d_VariableModules.f90:918-946): x fastest, lower bound -1.
struct Dims {
  int nx, ny, nz, ns;
  long long sy;     // nx+4
  long long sz;     // (nx+4)*(ny+4)
  long long ncell;  // (nx+4)*(ny+4)*(nz+4)
  __host__ __device__ inline long long p3(int i, int j, int k) const {
    return (long long)(i + 1) + sy * (long long)(j + 1) + sz * (long long)(k + 1);
  }
};

inline Dims make_dims(int nx, int ny, int nz, int ns) {
  Dims d;
  d.nx = nx; d.ny = ny; d.nz = nz; d.ns = ns;
  d.sy = nx + 4;
  d.sz = (long long)(nx + 4) * (ny + 4);
  d.ncell = d.sz * (nz + 4);
  return d;
}

// IsAloft as a bit mask passed by value (no per-call upload).
constexpr int kMaxBins = 1024;
struct AloftMask {
  uint32_t w[kMaxBins / 32];
  __host__ __device__ inline bool on(int n) const { return (w[n >> 5] >> (n & 31)) & 1u; }
};

// Geometry of one sweep direction: lines of length N along dir, transverse
// interior extents Na (fastest, mapped to lanes) and Nb.
struct LineGeom {
  int dir;          // 0:x 1:y 2:z
  int N, Na, Nb;
  int rmin, rmax;   // active range along the line
  int a0, a1, b0, b1;  // active transverse ranges
};

__host__ __device__ inline LineGeom make_line_geom(const Dims &d, int dir, const int32_t *bnd) {
  LineGeom g;
  g.dir = dir;
  if (dir == 0) {
    g.N = d.nx; g.Na = d.ny; g.Nb = d.nz;
    g.rmin = bnd[0]; g.rmax = bnd[1]; g.a0 = bnd[2]; g.a1 = bnd[3]; g.b0 = bnd[4]; g.b1 = bnd[5];
  } else if (dir == 1) {
    g.N = d.ny; g.Na = d.nx; g.Nb = d.nz;
    g.rmin = bnd[2]; g.rmax = bnd[3]; g.a0 = bnd[0]; g.a1 = bnd[1]; g.b0 = bnd[4]; g.b1 = bnd[5];
  } else {
    g.N = d.nz; g.Na = d.nx; g.Nb = d.ny;
    g.rmin = bnd[4]; g.rmax = bnd[5]; g.a0 = bnd[0]; g.a1 = bnd[1]; g.b0 = bnd[2]; g.b1 = bnd[3];
  }
  return g;
}

// Offset of cell -1 of the line through transverse (a,b) and its stride.
__host__ __device__ inline long long line_base(const Dims &d, int dir, int a, int b) {
  return dir == 0 ? d.p3(-1, a, b) : (dir == 1 ? d.p3(a, -1, b) : d.p3(a, b, -1));
}
__host__ __device__ inline long long line_stride(const Dims &d, int dir) {
  return dir == 0 ? 1LL : (dir == 1 ? d.sy : d.sz);
}
// Same for the diffusivity arrays kx/ky/kz(0:nx+1,0:ny+1,0:nz+1): offset of cell 0.
__host__ __device__ inline long long kline_base(const Dims &d, int dir, int a, int b) {
  const long long kx = d.nx + 2, kxy = (long long)(d.nx + 2) * (d.ny + 2);
  return dir == 0 ? (kx * a + kxy * b) : (dir == 1 ? (a + kxy * b) : (a + kx * b));
}
__host__ __device__ inline long long kline_stride(const Dims &d, int dir) {
  return dir == 0 ? 1LL : (dir == 1 ? (long long)(d.nx + 2) : (long long)(d.nx + 2) * (d.ny + 2));
}

// XCD-aware order of a flattened (tile, bin group) grid.  Workgroups are
// dispatched round-robin over the XCDs (workgroup b to XCD b mod 8), each
// XCD with its own L2.  XCD x takes the x-th eighth of the (tile, group)
// items in order: the m-th workgroup dispatched to it takes item
// x * (items / 8) + m, i.e. tile item / G and bin group item mod G.  So the
// G bin groups of a tile run back to back on one XCD and read the tile's
// per-cell fields (winds, geometry, K, line factors) through the same L2,
// and neighbouring tiles, whose rows share the cache lines at their edges,
// go through that L2 too.  A launch has padded_tiles(T) * G workgroups
// (a multiple of 8); those with tile >= T return at once.
constexpr int kXcds = 8;
__host__ __device__ inline long long padded_tiles(long long T) { return (T + kXcds - 1) / kXcds * kXcds; }
struct TileGroup {
  long long tile;
  int group;
};
__device__ inline TileGroup xcd_tile_group(long long b, int G, long long nblocks) {
  const long long per = nblocks / kXcds;
  const long long it = (b % kXcds) * per + b / kXcds;
  return TileGroup{it / G, (int)(it % G)};
}
__device__ inline TileGroup xcd_tile_group(long long b, int G) { return xcd_tile_group(b, G, gridDim.x); }

// Batched launches (ensembles): one launch covers the workgroups of several
// members (contexts).  A batched kernel takes the members' argument blocks
// (args), the member of each workgroup (member[blockIdx.x]) and each
// member's first workgroup (first, n + 1 entries): member m owns
// workgroups first[m] .. first[m+1]-1 and reads args[m].  The three are
// kernel parameters in the constant address space (ASH3D_CONSTANT), so
// that the argument blocks are read with scalar loads and the pointers in
// them are known to address global memory: the kernels issue global, not
// flat, loads and stores, as with arguments passed by value.  Every member's
// workgroup count is a multiple of kXcds where its kernel maps workgroups
// to XCDs, so a workgroup sits on the XCD it would have in the member's
// own launch.  A kernel's body takes the workgroup index and count within
// its member's launch.

#define ASH3D_CONSTANT __attribute__((address_space(4)))

// Argument block of the ghost-cell passes that a step launches (Set_BC(2)
// on one face, the Neumann ghosts of one direction).
struct GhostArgs {
  Dims d;
  int periodic, dir;
  double *c0;
};

constexpr int kJunk = 4096;

struct AdvArgs {
  Dims d;
  LineGeom g;
  double *c0;          // concen ts0 (read)
  double *cout;        // where the interior cells of the sweep's lines are written (c0: in place)
  const double *vel;   // vx_pd / vy_pd / vz_pd
  const double *vh;    // vh_pd (z)
  const double *vf;    // vf_pd (z)
  const double *sig;   // sigma_n{x,y,z}_pd
  const double *kap;   // kappa_pd
  const double *deld;  // DelD{x,y}onD_cc (x,y with ZScaling_ID == 2)
  double *of1, *of2;   // outflow on the low / high side of the line
  double *dep;         // DepositGranularity (z)
  double *junk;        // kJunk doubles that masked-off lanes store to
  double dt;
  int periodic;
  int limiter;
  int gy;              // streamed sweeps: bin groups per workgroup along y (0: line segments along y)
  int xrows;           // x rows: 1 one wavefront per row where it applies, 0 one workgroup per row
  int zg;              // 1: the concentration ghosts along the lines read as zero (the non-periodic
                       // Set_BC(1) values) instead of being loaded
  const double *dtv;   // dt / kappa per cell (the fused pass; null elsewhere)
  AloftMask aloft;
};

struct DiffArgs {
  Dims d;
  LineGeom g;
  double *c0;          // concen ts0 (in/out)
  double *scratch_b;   // per-bin scratch, same layout as concen ts0 (the ts1 slot)
  double *fac;         // per-line U factors, 3 * ncell, shared by the bin groups
  const double *kk;    // kx / ky / kz
  const double *sig;
  const double *kap;
  // theta-scheme row factors of each cell c (src/Diffusion.F90:807-816):
  // lr[c] = L = dt*kds_c*sig_c/kap_c, lr[c + ncell] = R = dt*kds_{c+1}*sig_{c+1}/kap_c
  const double *lr;
  // cached LU factors of the lines (bin-independent; see k_line_lu): per
  // row i, lu[c] = fact_i, lu[c+ncell] = D_i, lu[c+2*ncell] = DU_i,
  // lu[c+3*ncell] = DL_i, piv[c] = row interchange at elimination step i;
  // sing[line] = first zero pivot row of the line (0: none)
  const double *lu;
  const uint8_t *piv;
  const int *sing;
  // number of lines of the sweep whose matrix is not diagonally dominant
  // (from k_line_lu); sel = 1: the kernel runs only when it is zero (the
  // partitioned solves), sel = 2: only when it is not (their dgtsv-order
  // alternative), sel = 0: always
  const int *ndom;
  int sel;
  double *junk;        // kJunk doubles that masked-off lanes store to
  // per line of the sweep (k_line_lu's line index): the Jacobi sweeps the
  // line needs (0: more than kJacobiMaxSweepsX); lines with 1 <= jline <= jcap are
  // solved by the Jacobi sweeps and skipped by the partitioned solves.
  // null: none
  const uint8_t *jline;
  int jcap;
  double dt, theta;
  int *flag;           // set to the first singular pivot row (dgtsv INFO)
  AloftMask aloft;
};

// A kernel launch described instead of issued: the kernel for one member
// (single), the same kernel over a Batch of members (batch), the launch
// shape and dynamic LDS of one member, and its argument block (args_size
// bytes of args; 0 when the sweep needs no launch).  Launches with the
// same batch kernel, block and LDS can run as one batched launch.
struct LaunchSpec {
  const void *single = nullptr, *batch = nullptr;
  dim3 grid, block;
  size_t lds = 0;
  alignas(16) unsigned char args[1024];
  size_t args_size = 0;
};
// Fill sp (kernels, shape, LDS, argument block) or, with sp null, launch
// the single kernel.
template <class Args>
hipError_t issue(hipStream_t s, LaunchSpec *sp, const void *single, const void *batch, dim3 grid, dim3 block,
                 size_t lds, const Args &args) {
  if (sp) {
    static_assert(sizeof(Args) <= sizeof(sp->args), "argument block");
    sp->single = single;
    sp->batch = batch;
    sp->grid = grid;
    sp->block = block;
    sp->lds = lds;
    std::memcpy(sp->args, &args, sizeof args);
    sp->args_size = sizeof args;
    return hipSuccess;
  }
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(single, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  void *kargs[] = {(void *)&args};
  hipError_t e = hipLaunchKernel(single, grid, block, kargs, lds, s);
  return e != hipSuccess ? e : hipGetLastError();
}

// launchers (defined next to their kernels)
hipError_t launch_set_bc1_vel(hipStream_t s, const Dims &d, int periodic, double *vx, double *vy,
                              double *vz, double *vf);
hipError_t launch_set_bc1_conc(hipStream_t s, const Dims &d, int periodic, double *c0);
hipError_t launch_set_bc2(hipStream_t s, const Dims &d, int periodic, double *c0);
// Set_BC(2) on the dir (0: x, 1: y) ghosts of the interior rows only
hipError_t launch_set_bc2_face(hipStream_t s, const Dims &d, int periodic, int dir, double *c0);
hipError_t launch_neumann_dir(hipStream_t s, const Dims &d, int dir, int periodic, double *c0);
hipError_t launch_zero_ghosts(hipStream_t s, const Dims &d, double *c0);
// dst's ghost cells (every bin) = src's
hipError_t launch_copy_ghosts(hipStream_t s, const Dims &d, const double *src, double *dst);
hipError_t launch_fill(hipStream_t s, double *p, long long n, double v);
hipError_t launch_mirror_interior(hipStream_t s, const Dims &d, const double *c0, double *c1);
// umbrella source: 3x3 patch averaging (levels ibase..itop) and insertion
// (levels 1..nz+1) for bins < ngs; umbrella winds added to vx, vy
hipError_t launch_umbrella_source(hipStream_t s, const Dims &d, double *c0, int iv, int jv, double dt, int ibase,
                                  int itop, int ngs, const double *st, const double *fu);
hipError_t launch_umbrella_winds(hipStream_t s, const Dims &d, double *vx, double *vy, int ibase, int itop,
                                 const double *uvx, const double *uvy);
hipError_t launch_source_add(hipStream_t s, const Dims &d, double *c0, int iv, int jv, double dt,
                             const double *flux);
hipError_t launch_advect(hipStream_t s, const AdvArgs &a, int nb);

// the launch launch_advect(s, a, nb) would issue
hipError_t spec_advect(const AdvArgs &a, int nb, LaunchSpec *sp);
// Set_BC(2) on one face (launch_set_bc2_face), the Neumann ghosts of one
// direction (launch_neumann_dir)
void spec_set_bc2_face(const Dims &d, int periodic, int dir, double *c0, LaunchSpec *sp);
void spec_neumann_dir(const Dims &d, int dir, int periodic, double *c0, LaunchSpec *sp);
// one batched launch: members' argument blocks (device, each sp.args_size
// bytes), the member of each of the total workgroups and the members'
// first workgroups (device), and the launch shape (block, LDS) of the
// members with grid.y
hipError_t launch_batch(hipStream_t s, const LaunchSpec &sp, const void *dev_args, const int *dev_member,
                        const long long *dev_first, long long total, unsigned grid_y);
// the second horizontal sweep (h: lines along x or y, input h.c0) and the
// vertical sweep (z: output z.cout) in one pass, for the full-grid step
// that reads concentration ghosts as zeros (advect_hz.hip)
hipError_t launch_advect_hz(hipStream_t s, const AdvArgs &h, const AdvArgs &z);
// out[c] = dt / kap[c] for the n cells (the fused pass's cached dt/kappa)
hipError_t launch_cell_dtv(hipStream_t s, const double *kap, double *out, long long n, double dt);
hipError_t spec_advect_hz(const AdvArgs &h, const AdvArgs &z, LaunchSpec *sp);
hipError_t launch_advect_x_rows(hipStream_t s, const AdvArgs &a);
// the launch launch_advect_x_rows would issue (hipErrorInvalidValue: the
// row needs the streamed kernel)
hipError_t spec_advect_x_rows(const AdvArgs &a, LaunchSpec *sp);
hipError_t launch_diff_explicit(hipStream_t s, const DiffArgs &a, int nb);
hipError_t launch_diff_cn(hipStream_t s, const DiffArgs &a, int nb);
// the row factors L, R of every cell of the sweep's lines (bin-independent)
hipError_t launch_line_lr(hipStream_t s, const DiffArgs &a, double *lr);
// partitioned line solves (linesolve.hip): offsets of cell r of line (a, b)
// are st0 + sta*a + stb*b + str*r in the state and lr0 + lra*a + lrb*b +
// lrr*r in the row factor cache (R at + lr_ncell); hipErrorInvalidValue
// when the active line length is outside [3, 512]
struct PartLayout {
  long long st0, sta, stb, str;
  long long lr0, lra, lrb, lrr, lr_ncell;
  const double *kap = nullptr;  // cell volumes in the state's layout (adjoint only)
};
// adjoint: lambda <- K B A^{-1} K^{-1} lambda (= B^T A^{-T} lambda, since
// kappa_l R_l = kappa_{l+1} L_{l+1} makes K A and K B symmetric, K = diag kappa)
hipError_t launch_diff_cn_part(hipStream_t s, const DiffArgs &a, const PartLayout &lay, int bins_wg, int shape,
                               bool adjoint = false);
// the launch of the forward partitioned solves (launch_diff_cn_part)
hipError_t spec_diff_cn_part(const DiffArgs &a, const PartLayout &lay, int bins_wg, int shape, LaunchSpec *sp);
// the LU factors of every line of the sweep (reads a.lr)
// (*rho: the largest row ratio (|DL| + |DU|) / |D| over the sweep's lines as
// double bits, 1.0 when a line is not diagonally dominant; may be null.
// jl[line]: the Jacobi sweeps the line needs (jacobi_sweeps_needed of its
// ratio; 0: more than kJacobiMaxSweepsX), *rho_j the largest ratio of the
// lines that need at most that many and *njl their number; all three null
// or all given)
hipError_t launch_line_lu(hipStream_t s, const DiffArgs &a, double *lu, uint8_t *piv, int *sing, int *ndom,
                          unsigned long long *rho, uint8_t *jl, unsigned long long *rho_j, int *njl);
// theta-scheme solves of y or z lines by K + 1 Jacobi sweeps (0 <= k <=
// kJacobiMaxIters), in one pass, for sweeps whose rows are all diagonally
// dominant (nb: bins per thread, 1, 2 or 4)
constexpr int kJacobiMaxIters = 6;
// x lines take up to this many sweeps (a runtime count per line)
constexpr int kJacobiMaxSweepsX = 12;
// The least number of sweeps s <= max_sweeps with rho^s <= 2^-52 (1 - rho)
// -- after s sweeps from zero the error is at most rho^s of the line's
// largest |x|, and the rounding grows at most by 1 / (1 - rho) -- or 0
// when more are needed (rho not in [0, 1): 0).  Seven sweeps take rho up
// to 5.7998e-3, twelve up to 4.8e-2.
__host__ __device__ inline int jacobi_sweeps_needed(double rho, int max_sweeps) {
  if (!(rho >= 0.0 && rho < 1.0)) return 0;
  const double bound = 0x1p-52 * (1.0 - rho);
  double e = rho;
  for (int s = 1; s <= max_sweeps; ++s) {
    if (e <= bound) return s;
    e *= rho;
  }
  return 0;
}
// adjoint: lambda <- K B p(A) K^-1 lambda in place of the solve's K B A^-1 K^-1
// (p(A): the same sweeps; the transpose of the forward map, since
// p(A)^T = p(A^T) = K p(A) K^-1 for K A symmetric)
hipError_t launch_diff_cn_jacobi(hipStream_t s, const DiffArgs &a, int nb, int k, bool adjoint = false);
// the same for x lines, each with the sweeps it needs: the lines with
// 1 <= a.jline[line] <= a.jcap (the others are the partitioned solves');
// a describes the x<->y transposed grid as for the other x-line solves;
// the state c_orig, kappa, sigma_nx and kx in the original layout (orig);
// lines of up to 2048 cells
hipError_t launch_diff_cn_jacobi_x(hipStream_t s, const DiffArgs &a, const Dims &orig, double *c_orig,
                                   const double *kap, const double *sig, const double *kx, bool adjoint = false);
// x-line solves on the original state layout (geometry, K and scratch in
// the x<->y transposed layout described by a)
hipError_t launch_diff_cn_xrows(hipStream_t s, const DiffArgs &a, const Dims &orig, double *c0, int nb, int tx);
int pick_bins_per_thread(int ns, int kind, int dir);
long long mass_partials_size(const Dims &d);
// skip_if_zero: when given, the launch does nothing if *skip_if_zero == 0
hipError_t launch_transpose_xy(hipStream_t s, const double *src, double *dst, int A, int B, long long nplanes,
                               const int *skip_if_zero = nullptr);
hipError_t launch_mass_totals(hipStream_t s, const Dims &d, const double *c0, const double *kap,
                              const double *dep, const double *yz1, const double *yz2, const double *xz1,
                              const double *xz2, const double *xy2, double *partial, double *out);
// output products (products.hip): c = the ts1 view of the state (null:
// all zero), dep = DepositGranularity, dz/zcc = dz_vec_pd, z_cc_pd (nz);
// out = 5 planes of nx*ny doubles then 2 planes of int32 masks
struct ProdArgs {
  Dims d;
  const double *c, *dep, *dz, *zcc;
  double cloudcon, cloudload, density, depo;
  int i0, i1, j0, j1;
  double *out;
};
hipError_t launch_products(hipStream_t s, const ProdArgs &p);
std::string esri_ascii_text(int nx, int ny, const double *field, const int32_t *mask, const char *fill,
                            int latlon, double xll, double yll, double dx, double dy);
std::string esri_ascii_name(const char *root, double time_h, int is_final);
// record msg as the library's last error (ash3d_tr_last_error); returns code
int report_error(int code, const char *msg);
// driver work on the device (meteo.hip)
hipError_t launch_met_interp(hipStream_t s, const Dims &d, const float *const last[4], const float *const next[4],
                             double frac, bool with_vf, double *vx, double *vy, double *vz, double *vf);
int dt_partial_blocks(const Dims &d);
hipError_t launch_dt_partials(hipStream_t s, const Dims &d, int latlon, int met, int diffusion, const double *vx,
                              const double *vy, const double *vz, const double *vf, const double *kap,
                              const double *sx, const double *sy, const double *sz, const double *kx,
                              const double *ky, const double *kz, const double *dz, const float *const met_next[4],
                              double *partial);
hipError_t launch_slab_max(hipStream_t s, const Dims &d, const double *c0, unsigned long long *prof);
// adjoint (adjoint.hip)
hipError_t launch_adj_advect(hipStream_t s, const AdvArgs &a, const double *qpre, double *lam, int nb);
hipError_t launch_adj_cn(hipStream_t s, const DiffArgs &a, double *lam, int nb);
hipError_t launch_adj_diff_explicit(hipStream_t s, const DiffArgs &a, double *lam, int nb, int periodic);
hipError_t launch_column_load(hipStream_t s, const Dims &d, const double *c0, const double *dz, double *out);
hipError_t launch_adj_column(hipStream_t s, const Dims &d, double *lam, const double *dz, const double *w);
hipError_t launch_adj_source(hipStream_t s, const Dims &d, const double *lam, int iv, int jv, double dt,
                             double *slot);

}  // namespace ash3d
