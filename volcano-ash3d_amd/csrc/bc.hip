// Ghost-cell kernels: Set_BC (src/Set_BC.f90:15-126), the per-direction
// Neumann ghosts the diffusion routines set (src/Diffusion.F90:317-327,
// :457-460, :587-590, :756-768, :1018-1021, :1277-1280), the driver's step
// glue (src/Ash3d.F90:376-378) and the vent-column source add (:428-431).
// All of these touch O(surface) cells; every thread writes one cell and
// reads only interior cells, so no ordering between threads is needed.
#include <algorithm>

#include "common.hpp"

namespace ash3d {

namespace {

constexpr int kBlock = 256;

// The t-th ghost cell of plane k.  In the four ghost planes every cell is a
// ghost; in an interior plane the ghosts are the ring of 4 rows of nx+4
// cells (j = -1,0,ny+1,ny+2) and 4 columns of ny cells (i = -1,0,nx+1,nx+2).
__device__ inline bool ghost_cell(const Dims &d, int k, long long t, int &i, int &j) {
  const int W = d.nx + 4;
  const bool gplane = (k < 1 || k > d.nz);
  if (gplane) {
    if (t >= (long long)W * (d.ny + 4)) return false;
    i = (int)(t % W) - 1;
    j = (int)(t / W) - 1;
    return true;
  }
  const long long rowcells = 4LL * W;
  if (t < rowcells) {
    const int r = (int)(t / W);
    i = (int)(t % W) - 1;
    j = (r < 2) ? (r - 1) : (d.ny + (r - 1));  // -1, 0, ny+1, ny+2
    return true;
  }
  t -= rowcells;
  if (t >= 4LL * d.ny) return false;
  // the four ghost columns of a row in adjacent lanes: i = -1, 0 and
  // nx+1, nx+2 are two 16-byte pieces of the row
  const int col = (int)(t & 3);
  j = (int)(t >> 2) + 1;
  i = (col < 2) ? (col - 1) : (d.nx + (col - 1));
  return true;
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// x index of the Neumann / periodic source of ghost column i (interior i maps to itself)
__device__ inline int xmap(int i, int nx, int periodic) {
  if (i >= 1 && i <= nx) return i;
  if (periodic) return i < 1 ? i + nx : i - nx;  // -1->nx-1, 0->nx, nx+1->1, nx+2->2
  return i < 1 ? 1 : nx;
}

enum { GH_ZERO = 0, GH_BC1 = 1, GH_BC2 = 2, GH_COPY = 3 };

// Every ghost cell of every bin in one launch (bin blockIdx.y): per bin
// the four ghost planes k = -1, 0, nz+1, nz+2, then the rings of the
// interior planes k = 1..nz.  mode GH_COPY: src's value; GH_ZERO: 0; GH_BC1: Set_BC(1)
// concentration rule (:77-91); GH_BC2: Set_BC(2) (:98-118) -- the
// sequential x (periodic override), y, z mirrors leave every ghost equal to
// c(xmap(i), clamp(j), clamp(k)).
__global__ __launch_bounds__(kBlock) void k_ghosts_all(Dims d, int periodic, int mode, double *c0,
                                                      const double *src) {
  const int n = (int)blockIdx.y;
  long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long plane = (long long)(d.nx + 4) * (d.ny + 4);
  const long long ring = 4LL * (d.nx + 4) + 4LL * d.ny;
  int k, i, j;
  if (t < 4 * plane) {
    const int p = (int)(t / plane);
    k = (p < 2) ? p - 1 : d.nz + p - 1;  // -1, 0, nz+1, nz+2
    t -= p * plane;
  } else {
    t -= 4 * plane;
    if (t >= (long long)d.nz * ring) return;
    k = 1 + (int)(t / ring);
    t %= ring;
  }
  if (!ghost_cell(d, k, t, i, j)) return;
  double *c = c0 + d.ncell * n;
  double v = 0.0;
  if (mode == GH_BC1) {
    if (periodic && j >= 1 && j <= d.ny && k >= 1 && k <= d.nz && (i < 1 || i > d.nx))
      v = c[d.p3(xmap(i, d.nx, 1), j, k)];
  } else if (mode == GH_BC2) {
    v = c[d.p3(xmap(i, d.nx, periodic), clampi(j, 1, d.ny), clampi(k, 1, d.nz))];
  } else if (mode == GH_COPY) {
    v = src[d.p3(i, j, k) + d.ncell * n];
  }
  c[d.p3(i, j, k)] = v;
}

hipError_t launch_ghosts(hipStream_t s, const Dims &d, int periodic, int mode, double *c0,
                         const double *src = nullptr) {
  const long long plane = (long long)(d.nx + 4) * (d.ny + 4);
  const long long ring = 4LL * (d.nx + 4) + 4LL * d.ny;
  const long long per_bin = 4 * plane + (long long)d.nz * ring;
  dim3 grid((unsigned)((per_bin + kBlock - 1) / kBlock), d.ns);
  hipLaunchKernelGGL(k_ghosts_all, grid, dim3(kBlock), 0, s, d, periodic, mode, c0, src);
  return hipGetLastError();
}

// Set_BC(1) velocities (:44-70): zero-order extrapolation into the ghost
// layers of vx (x, periodic wrap), vy (y), vz and every vf bin (z).
__device__ inline void vel_x(Dims d, int periodic, double *vx, long long t) {
  const long long nl = (long long)(d.ny + 4) * (d.nz + 4);
  if (t >= nl) return;
  const int j = (int)(t % (d.ny + 4)) - 1, k = (int)(t / (d.ny + 4)) - 1;
  const int nx = d.nx;
  if (periodic) {
    vx[d.p3(-1, j, k)] = vx[d.p3(nx - 1, j, k)];
    vx[d.p3(0, j, k)] = vx[d.p3(nx, j, k)];
    vx[d.p3(nx + 1, j, k)] = vx[d.p3(1, j, k)];
    vx[d.p3(nx + 2, j, k)] = vx[d.p3(2, j, k)];
  } else {
    vx[d.p3(-1, j, k)] = vx[d.p3(1, j, k)];
    vx[d.p3(0, j, k)] = vx[d.p3(1, j, k)];
    vx[d.p3(nx + 1, j, k)] = vx[d.p3(nx, j, k)];
    vx[d.p3(nx + 2, j, k)] = vx[d.p3(nx, j, k)];
  }
}

__device__ inline void vel_y(Dims d, double *vy, long long t) {
  const long long nl = (long long)(d.nx + 4) * (d.nz + 4);
  if (t >= nl) return;
  const int i = (int)(t % (d.nx + 4)) - 1, k = (int)(t / (d.nx + 4)) - 1;
  const int ny = d.ny;
  vy[d.p3(i, -1, k)] = vy[d.p3(i, 1, k)];
  vy[d.p3(i, 0, k)] = vy[d.p3(i, 1, k)];
  vy[d.p3(i, ny + 1, k)] = vy[d.p3(i, ny, k)];
  vy[d.p3(i, ny + 2, k)] = vy[d.p3(i, ny, k)];
}

// vz or one bin of vf
__device__ inline void vel_z(Dims d, double *v, long long t) {
  const long long nl = (long long)(d.nx + 4) * (d.ny + 4);
  if (t >= nl) return;
  const int i = (int)(t % (d.nx + 4)) - 1, j = (int)(t / (d.nx + 4)) - 1;
  const int nz = d.nz;
  v[d.p3(i, j, -1)] = v[d.p3(i, j, 1)];
  v[d.p3(i, j, 0)] = v[d.p3(i, j, 1)];
  v[d.p3(i, j, nz + 1)] = v[d.p3(i, j, nz)];
  v[d.p3(i, j, nz + 2)] = v[d.p3(i, j, nz)];
}

// All velocity ghosts in one launch: blockIdx.y 0 = vx (x lines), 1 = vy
// (y lines), 2 = vz, 3 + n = vf of bin n (z lines).
__global__ __launch_bounds__(kBlock) void k_vel_all(Dims d, int periodic, double *vx, double *vy, double *vz,
                                                   double *vf) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const int y = (int)blockIdx.y;
  if (y == 0) vel_x(d, periodic, vx, t);
  else if (y == 1) vel_y(d, vy, t);
  else vel_z(d, y == 2 ? vz : vf + d.ncell * (long long)(y - 3), t);
}

// Neumann ghosts along one direction for all bins of ts0, all transverse
// indices including ghosts; x applies the periodic override.  y and z: a
// thread per line writes its four ghosts (lanes along the contiguous x);
// x: a thread per ghost, the four of a row in adjacent lanes.
__device__ __forceinline__ void neumann_body(const GhostArgs &G, const long long t, const int n) {
  const Dims d = G.d;
  const int dir = G.dir, periodic = G.periodic;
  double *c0 = G.c0;
  int A, B, N;
  if (dir == 0) { A = d.ny + 4; B = d.nz + 4; N = d.nx; }
  else if (dir == 1) { A = d.nx + 4; B = d.nz + 4; N = d.ny; }
  else { A = d.nx + 4; B = d.ny + 4; N = d.nz; }
  const int per = (dir == 0) ? 4 : 1;  // threads per line
  if (t >= (long long)A * B * per) return;
  const long long line = (dir == 0) ? (t >> 2) : t;
  const int u0 = (dir == 0) ? (int)(t & 3) : 0, u1 = (dir == 0) ? u0 + 1 : 4;
  const int a = (int)(line % A) - 1, b = (int)(line / A) - 1;
  double *c = c0 + d.ncell * n;
  const long long base = line_base(d, dir, a, b), st = line_stride(d, dir);
  const int p = (dir == 0) ? periodic : 0;
  const int g[4] = {-1, 0, N + 1, N + 2};
  for (int u = u0; u < u1; ++u) {
    int src;
    if (p) src = g[u] < 1 ? g[u] + N : g[u] - N;
    else src = g[u] < 1 ? 1 : N;
    c[base + st * (g[u] + 1)] = c[base + st * (src + 1)];
  }
}

// The part of Set_BC(2) (rule GH_BC2) that the diffusion of a step reads
// or leaves in the final state when the sweeps start along dir (0: x, 1:
// y): the dir ghosts of the interior rows.  Every other ghost the full
// pass writes is rewritten, before anything reads it, by the Neumann
// passes of the later sweeps (which cover all transverse indices,
// ghosts included: the second horizontal direction's over every ghost
// plane index, the z pass over every ghost plane).  dir 0: the four x
// ghosts of a row in adjacent lanes; dir 1: lanes along x.
__device__ __forceinline__ void bc2_face_body(const GhostArgs &G, const long long t, const int n) {
  const Dims d = G.d;
  const int dir = G.dir, periodic = G.periodic;
  double *c = G.c0 + d.ncell * n;
  if (dir == 0) {
    if (t >= 4LL * d.ny * d.nz) return;
    const int col = (int)(t & 3);
    const long long row = t >> 2;
    const int j = (int)(row % d.ny) + 1, k = (int)(row / d.ny) + 1;
    const int i = (col < 2) ? (col - 1) : (d.nx + (col - 1));
    c[d.p3(i, j, k)] = c[d.p3(xmap(i, d.nx, periodic), j, k)];
  } else {
    if (t >= 4LL * d.nx * d.nz) return;
    const int i = (int)(t % d.nx) + 1;
    const long long rk = t / d.nx;
    const int r = (int)(rk & 3), k = (int)(rk >> 2) + 1;
    const int j = (r < 2) ? (r - 1) : (d.ny + (r - 1));
    c[d.p3(i, j, k)] = c[d.p3(i, clampi(j, 1, d.ny), k)];
  }
}

// one launch per member (bin blockIdx.y) and one over the members of an
// ensemble (grid.y = the largest bin count)
__global__ __launch_bounds__(kBlock) void k_neumann(GhostArgs G) {
  neumann_body(G, (long long)blockIdx.x * kBlock + threadIdx.x, (int)blockIdx.y);
}
__global__ __launch_bounds__(kBlock) void k_neumann_batch(
    const ASH3D_CONSTANT GhostArgs *__restrict__ args, const ASH3D_CONSTANT int *__restrict__ member,
    const ASH3D_CONSTANT long long *__restrict__ first) {
  const int m = member[blockIdx.x];
  const GhostArgs &G = *(const GhostArgs *)(args + m);
  if ((int)blockIdx.y >= G.d.ns) return;
  neumann_body(G, ((long long)blockIdx.x - first[m]) * kBlock + threadIdx.x, (int)blockIdx.y);
}
__global__ __launch_bounds__(kBlock) void k_bc2_face(GhostArgs G) {
  bc2_face_body(G, (long long)blockIdx.x * kBlock + threadIdx.x, (int)blockIdx.y);
}
__global__ __launch_bounds__(kBlock) void k_bc2_face_batch(
    const ASH3D_CONSTANT GhostArgs *__restrict__ args, const ASH3D_CONSTANT int *__restrict__ member,
    const ASH3D_CONSTANT long long *__restrict__ first) {
  const int m = member[blockIdx.x];
  const GhostArgs &G = *(const GhostArgs *)(args + m);
  if ((int)blockIdx.y >= G.d.ns) return;
  bc2_face_body(G, ((long long)blockIdx.x - first[m]) * kBlock + threadIdx.x, (int)blockIdx.y);
}

__global__ __launch_bounds__(kBlock) void k_fill(double *p, long long n, double v) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long q = t; q < n; q += stride) p[q] = v;
}

// ts1 = interior(ts0), zero ghosts: the content the reference leaves in ts1
// after any sweep (zero fill, interior rows written, copied to ts0).
__global__ __launch_bounds__(kBlock) void k_mirror(Dims d, const double *c0, double *c1) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long tot = d.ncell * d.ns;
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long q = t; q < tot; q += stride) {
    const long long cell = q % d.ncell;
    const int i = (int)(cell % d.sy) - 1;
    const int j = (int)((cell / d.sy) % (d.ny + 4)) - 1;
    const int k = (int)(cell / d.sz) - 1;
    const bool in = i >= 1 && i <= d.nx && j >= 1 && j <= d.ny && k >= 1 && k <= d.nz;
    c1[q] = in ? c0[q] : 0.0;
  }
}

// concen(iv,jv,k,n,ts0) += dt*flux(k,n), k = 1..nz+1 (src/Ash3d.F90:428-431)
__global__ __launch_bounds__(kBlock) void k_source(Dims d, double *c0, int iv, int jv, double dt,
                                                  const double *flux) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long tot = (long long)(d.nz + 1) * d.ns;
  if (t >= tot) return;
  const int k = (int)(t % (d.nz + 1)) + 1, n = (int)(t / (d.nz + 1));
  const long long o = d.p3(iv, jv, k) + d.ncell * n;
  c0[o] = c0[o] + dt * flux[t];
}

// Umbrella source (src/Ash3d.F90:447-460), one thread per (level k =
// 1..nz+1, bin n < n_gs_max): on levels ibase..itop the 3x3 patch is set to
// its stencil-weighted average (AvgCon_Umbrella, src/Source_Umbrella.f90:
// 539-560, summed i-outer, j-inner), then dt*flux_umb is added to it.
__global__ __launch_bounds__(kBlock) void k_umbrella_source(Dims d, double *c0, int iv, int jv, double dt,
                                                           int ibase, int itop, int ngs, const double *st,
                                                           const double *fu) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long long)(d.nz + 1) * ngs) return;
  const int k = (int)(t % (d.nz + 1)) + 1, n = (int)(t / (d.nz + 1));
  double *c = c0 + d.ncell * n;
  if (k >= ibase && k <= itop) {
    double a = 0.0;
    for (int i = 1; i <= 3; ++i)
      for (int j = 1; j <= 3; ++j)
        a = a + c[d.p3(iv - 2 + i, jv - 2 + j, k)] * st[(i - 1) + 3 * (j - 1) + 9 * (k - 1)];
    for (int i = 1; i <= 3; ++i)
      for (int j = 1; j <= 3; ++j) c[d.p3(iv - 2 + i, jv - 2 + j, k)] = a;
  }
  const double *f = fu + 9 * (long long)(k - 1) + 9LL * (d.nz + 1) * n;
  for (int i = 1; i <= 3; ++i)
    for (int j = 1; j <= 3; ++j) {
      const long long o = d.p3(iv - 2 + i, jv - 2 + j, k);
      c[o] = c[o] + dt * f[(i - 1) + 3 * (j - 1)];
    }
}

// vx_pd, vy_pd += the umbrella winds (one value per column) on levels
// ibase..itop (src/MesoInterpolater.f90:238-243)
__global__ __launch_bounds__(kBlock) void k_umbrella_winds(Dims d, double *vx, double *vy, int ibase, int itop,
                                                          const double *uvx, const double *uvy) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long plane = (long long)d.nx * d.ny;
  if (t >= plane * (itop - ibase + 1)) return;
  const int i = (int)(t % d.nx) + 1, j = (int)((t / d.nx) % d.ny) + 1, k = ibase + (int)(t / plane);
  const long long o = d.p3(i, j, k), q = (long long)(i - 1) + (long long)d.nx * (j - 1);
  vx[o] = vx[o] + uvx[q];
  vy[o] = vy[o] + uvy[q];
}

unsigned blocks_for(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b > 65535LL * 16) b = 65535LL * 16;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

hipError_t launch_set_bc1_vel(hipStream_t s, const Dims &d, int periodic, double *vx, double *vy,
                              double *vz, double *vf) {
  const long long nx_l = (long long)(d.ny + 4) * (d.nz + 4);
  const long long ny_l = (long long)(d.nx + 4) * (d.nz + 4);
  const long long nz_l = (long long)(d.nx + 4) * (d.ny + 4);
  const long long nl = std::max(nx_l, std::max(ny_l, nz_l));
  hipLaunchKernelGGL(k_vel_all, dim3(blocks_for(nl), 3 + d.ns), dim3(kBlock), 0, s, d, periodic, vx, vy, vz, vf);
  return hipGetLastError();
}

hipError_t launch_set_bc1_conc(hipStream_t s, const Dims &d, int periodic, double *c0) {
  return launch_ghosts(s, d, periodic, GH_BC1, c0);
}

hipError_t launch_set_bc2(hipStream_t s, const Dims &d, int periodic, double *c0) {
  return launch_ghosts(s, d, periodic, GH_BC2, c0);
}

void spec_set_bc2_face(const Dims &d, int periodic, int dir, double *c0, LaunchSpec *sp) {
  const long long n = 4LL * (dir == 0 ? d.ny : d.nx) * d.nz;
  (void)issue(nullptr, sp, (const void *)k_bc2_face, (const void *)k_bc2_face_batch, dim3(blocks_for(n), d.ns),
              dim3(kBlock), 0, GhostArgs{d, periodic, dir, c0});
}

hipError_t launch_set_bc2_face(hipStream_t s, const Dims &d, int periodic, int dir, double *c0) {
  const long long n = 4LL * (dir == 0 ? d.ny : d.nx) * d.nz;
  hipLaunchKernelGGL(k_bc2_face, dim3(blocks_for(n), d.ns), dim3(kBlock), 0, s, GhostArgs{d, periodic, dir, c0});
  return hipGetLastError();
}

hipError_t launch_zero_ghosts(hipStream_t s, const Dims &d, double *c0) {
  return launch_ghosts(s, d, 0, GH_ZERO, c0);
}

hipError_t launch_copy_ghosts(hipStream_t s, const Dims &d, const double *src, double *dst) {
  return launch_ghosts(s, d, 0, GH_COPY, dst, src);
}

static long long neumann_threads(const Dims &d, int dir) {
  if (dir == 0) return 4LL * (d.ny + 4) * (d.nz + 4);  // a thread per x ghost
  if (dir == 1) return (long long)(d.nx + 4) * (d.nz + 4);
  return (long long)(d.nx + 4) * (d.ny + 4);
}

void spec_neumann_dir(const Dims &d, int dir, int periodic, double *c0, LaunchSpec *sp) {
  (void)issue(nullptr, sp, (const void *)k_neumann, (const void *)k_neumann_batch,
              dim3(blocks_for(neumann_threads(d, dir)), d.ns), dim3(kBlock), 0, GhostArgs{d, periodic, dir, c0});
}

hipError_t launch_neumann_dir(hipStream_t s, const Dims &d, int dir, int periodic, double *c0) {
  hipLaunchKernelGGL(k_neumann, dim3(blocks_for(neumann_threads(d, dir)), d.ns), dim3(kBlock), 0, s,
                     GhostArgs{d, periodic, dir, c0});
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, double *p, long long n, double v) {
  if (n <= 0) return hipSuccess;
  unsigned b = blocks_for(n);
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_fill, dim3(b), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_mirror_interior(hipStream_t s, const Dims &d, const double *c0, double *c1) {
  unsigned b = blocks_for(d.ncell * d.ns);
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_mirror, dim3(b), dim3(kBlock), 0, s, d, c0, c1);
  return hipGetLastError();
}

hipError_t launch_umbrella_source(hipStream_t s, const Dims &d, double *c0, int iv, int jv, double dt, int ibase,
                                  int itop, int ngs, const double *st, const double *fu) {
  const long long tot = (long long)(d.nz + 1) * ngs;
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_umbrella_source, dim3(blocks_for(tot)), dim3(kBlock), 0, s, d, c0, iv, jv, dt, ibase, itop,
                     ngs, st, fu);
  return hipGetLastError();
}

hipError_t launch_umbrella_winds(hipStream_t s, const Dims &d, double *vx, double *vy, int ibase, int itop,
                                 const double *uvx, const double *uvy) {
  const long long tot = (long long)d.nx * d.ny * (itop - ibase + 1);
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_umbrella_winds, dim3(blocks_for(tot)), dim3(kBlock), 0, s, d, vx, vy, ibase, itop, uvx, uvy);
  return hipGetLastError();
}

hipError_t launch_source_add(hipStream_t s, const Dims &d, double *c0, int iv, int jv, double dt,
                             const double *flux) {
  const long long tot = (long long)(d.nz + 1) * d.ns;
  hipLaunchKernelGGL(k_source, dim3(blocks_for(tot)), dim3(kBlock), 0, s, d, c0, iv, jv, dt, flux);
  return hipGetLastError();
}

}  // namespace ash3d
