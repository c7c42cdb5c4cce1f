// Output products of the transport state and their ESRI ASCII files:
//   ConcentrationCalculator  src/Output_Vars.f90:1026-1137 (cloud load,
//                            peak concentration, cloud top and bottom,
//                            cloud mask and area)
//   AshThicknessCalculator   src/Output_Vars.f90:855-893 (deposit thickness,
//                            deposit mask and area)
//   write_2D_ASCII           src/write_ASCII.f90:264-382 (header, 10f18.6
//                            rows from the northern row down, real(4) values)
// The products are one pass over the state on the device, one thread per
// column (lanes along x, so every load of a wavefront is one coalesced
// segment), in the reference's summation order.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "common.hpp"

namespace ash3d {

namespace {

constexpr double KG_2_MG = 1.0e6;    // src/Ash3d_VariableModules.f90:285
constexpr double KM3_2_M3 = 1.0e9;   // :284
constexpr double KM2_2_M2 = 1.0e6;   // :283
constexpr double M_2_MM = 1.0e3;     // :281

// out: 5 planes of nx*ny doubles (cloud load, max concentration, max
// height, min height, deposit thickness) then 2 planes of int32 masks.
__global__ void k_products(ProdArgs P) {
  const Dims d = P.d;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nxy = (long long)d.nx * d.ny;
  if (t >= nxy) return;
  const int i = (int)(t % d.nx) + 1, j = (int)(t / d.nx) + 1;
  double *load = P.out, *maxc = P.out + nxy, *maxh = maxc + nxy, *minh = maxh + nxy, *thick = minh + nxy;
  int *mcloud = (int *)(P.out + 5 * nxy), *mdep = mcloud + nxy;
  // ConcentrationCalculator, columns of the box imin..imax x jmin..jmax
  double L = 0.0, H = 0.0, B = 100.0, C = 0.0;
  int mask = 0;
  if (i >= P.i0 && i <= P.i1 && j >= P.j0 && j <= P.j1) {
    double tcmax = 0.0;
    bool first_le = false;
    for (int k = 1; k <= d.nz; ++k) {
      double tc = 0.0;  // sum(concen_pd(i,j,k,1:n_gs_max,ts1)), bins in order
      if (P.c != nullptr)
        for (int n = 0; n < d.ns; ++n) tc += P.c[d.p3(i, j, k) + d.ncell * (long long)n];
      L = L + tc * P.dz[k - 1] / 1.0e3;
      if (tc >= P.cloudcon) H = P.zcc[k - 1] + 0.5 * P.dz[k - 1];
      // the top-down scan keeps the lowest level at or below the threshold
      if (tc <= P.cloudcon && !first_le) {
        B = P.zcc[k - 1] - 0.5 * P.dz[k - 1];
        first_le = true;
      }
      tcmax = (k == 1) ? tc : fmax(tcmax, tc);  // maxval(TotalConcentration(:))
    }
    mask = (L >= P.cloudload) ? 1 : 0;
    if (mask) {
      if (B < 0.0 && H > 0.0) B = 0.0;
      B = fmin(H, B);
      C = tcmax * KG_2_MG / KM3_2_M3;
    } else {
      L = 0.0;
      C = 0.0;
      H = 0.0;
      B = 0.0;
    }
  }
  load[t] = L;
  maxc[t] = C;
  maxh[t] = H;
  minh[t] = B;
  mcloud[t] = mask;
  // AshThicknessCalculator, every column
  double s = 0.0;  // sum(DepositGranularity(i,j,1:n_gs_max))
  for (int n = 0; n < d.ns; ++n) s += P.dep[(long long)(i - 1) + (long long)d.nx * ((long long)(j - 1) + (long long)d.ny * n)];
  const double th = s * P.dz[0] / KM2_2_M2 / P.density * M_2_MM;
  thick[t] = th;
  mdep[t] = (th > P.depo) ? 1 : 0;
}

// Fortran Fw.d of a real(4) value: the exact decimal expansion rounded to d
// places, right-justified in w columns, asterisks when it does not fit.
void put_f(std::string &o, double v, int w, int dd) {
  char buf[64];
  int len = snprintf(buf, sizeof buf, "%*.*f", w, dd, v);
  if (len < 0 || len > w) o.append((size_t)w, '*');
  else o.append(buf, (size_t)len);
}

void put_i(std::string &o, int v, int w) {
  char buf[32];
  int len = snprintf(buf, sizeof buf, "%*d", w, v);
  if (len < 0 || len > w) o.append((size_t)w, '*');
  else o.append(buf, (size_t)len);
}

}  // namespace

hipError_t launch_products(hipStream_t s, const ProdArgs &p) {
  const long long nxy = (long long)p.d.nx * p.d.ny;
  hipLaunchKernelGGL(k_products, dim3((unsigned)((nxy + 255) / 256)), dim3(256), 0, s, p);
  return hipGetLastError();
}

// write_2D_ASCII (src/write_ASCII.f90:334-373) into a string
std::string esri_ascii_text(int nx, int ny, const double *field, const int32_t *mask, const char *fill,
                            int latlon, double xll, double yll, double dx, double dy) {
  std::string o;
  o.reserve((size_t)nx * ny * 18 + (size_t)ny * (nx / 10 + 4) + 256);
  const double KM_2_M = 1.0e3;  // src/Ash3d_VariableModules.f90:280
  o += "NCOLS ";
  put_i(o, nx, 5);
  o += "\nNROWS ";
  put_i(o, ny, 5);
  o += "\nXLLCORNER ";
  // lon-lat: GIS prefers -180..180 (:343-346); projected: km to m (:350-352)
  put_f(o, latlon ? (xll > 180.0 ? xll - 360.0 : xll) : xll * KM_2_M, 15, 3);
  o += "\nYLLCORNER ";
  put_f(o, latlon ? yll : yll * KM_2_M, 15, 3);
  o += "\nCELLSIZE ";
  put_f(o, latlon ? dx : dx * KM_2_M, 15, 3);
  put_f(o, latlon ? dy : dy * KM_2_M, 15, 3);
  char fv[7] = "      ";  // Fill_Value, character(len=6), a6
  for (int c = 0; c < 6 && fill[c]; ++c) fv[c] = fill[c];
  o += "\nNODATA_VALUE ";
  o += fv;
  o += "\n";
  const float fvalue = (float)strtod(fill, nullptr);  // read(Fill_Value,*) FValue, real(op=4)
  for (int j = ny; j >= 1; --j) {
    for (int i0 = 1; i0 <= nx; i0 += 10) {  // format 10f18.6: records of ten
      for (int i = i0; i <= nx && i < i0 + 10; ++i) {
        const long long t = (long long)(i - 1) + (long long)nx * (j - 1);
        const float v = (mask == nullptr || mask[t]) ? (float)field[t] : fvalue;  // real(..., kind=op)
        put_f(o, (double)v, 18, 6);
      }
      o += "\n";
    }
    o += "  \n";  // write(fid,*)" ": list-directed blank, then the string
  }
  return o;
}

// the output file name of write_2D_ASCII (:316-331)
std::string esri_ascii_name(const char *root, double time_h, int is_final) {
  std::string r(root);
  const size_t a = r.find_first_not_of(' ');
  r = (a == std::string::npos) ? std::string() : r.substr(a);
  const size_t b = r.find_last_not_of(' ');
  r = (b == std::string::npos) ? std::string() : r.substr(0, b + 1);
  if (r.find("outvar") != std::string::npos || r.find("ArrivalTime") != std::string::npos) return r + ".dat";
  std::string cio;
  if (is_final) {
    cio = "____final";
  } else {
    std::string tbuf;
    if (time_h < 10.0) {
      cio = "00";
      put_f(tbuf, time_h, 4, 2);
    } else if (time_h < 100.0) {
      cio = "0";
      put_f(tbuf, time_h, 5, 2);
    } else {
      put_f(tbuf, time_h, 6, 2);
    }
    cio += tbuf + "hrs";
  }
  return r + cio + ".dat";
}

}  // namespace ash3d
