// gx_host.hpp -- host-side helpers shared by the single-chart and the batch
// entry points: error text, argument checks, chart geometry and state
// transfers in the precision of the state.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/gx_fluid.h"
#include "gx_kernels.hpp"

namespace gx {

// Message of the last failed call, per host thread (gx_last_error): a fixed
// buffer, so recording an error allocates nothing.
constexpr size_t ERROR_CAP = 512;
extern thread_local char g_last_error[ERROR_CAP];

inline int fail(int code, const char *msg) noexcept {
  size_t i = 0;
  for (; msg && msg[i] && i + 1 < ERROR_CAP; i++) g_last_error[i] = msg[i];
  g_last_error[i] = '\0';
  return code;
}

inline int fail(int code, const std::string &msg) noexcept { return fail(code, msg.c_str()); }

// Every extern "C" body that returns a status runs between GX_API_BEGIN and
// GX_API_END: a C++ exception (a host allocation in std::vector or
// std::string, or anything else) becomes a status code and never crosses
// the C-ABI into the caller.
#define GX_API_BEGIN try {
#define GX_API_END(who)                                                      \
  }                                                                          \
  catch (const std::bad_alloc &) {                                           \
    return gx::fail(GX_E_NOMEM, who ": host allocation failed");             \
  }                                                                          \
  catch (...) {                                                              \
    return gx::fail(GX_E_STATE, who ": internal error (C++ exception)");     \
  }

#define HIP_TRY(expr)                                                       \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(GX_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

inline int check_params(const gx_params &p, const char *who) {
  if (p.Precision != GX_FP64 && p.Precision != GX_FP32)
    return fail(GX_E_ARG, std::string(who) + ": Precision must be GX_FP64 or GX_FP32");
  if (p.RiemannSolverType != GX_HLL && p.RiemannSolverType != GX_HLLC)
    return fail(GX_E_ARG, std::string(who) + ": RiemannSolverType must be GX_HLL or GX_HLLC");
  if (p.ReconstructionType != GX_LINEAR && p.ReconstructionType != GX_PARABOLIC)
    return fail(GX_E_ARG,
                std::string(who) + ": ReconstructionType must be GX_LINEAR or GX_PARABOLIC");
  return GX_OK;
}

inline int check_grid(const gx_grid &g, const char *who) {
  const std::string w(who);
  for (int d = 0; d < 3; d++) {
    if (g.nCells[d] < 2) return fail(GX_E_ARG, w + ": nCells must be >= 2 in every dimension");
    const int bi = g.BoundaryCondition[2 * d], bo = g.BoundaryCondition[2 * d + 1];
    if (bi < 0 || bi > 2 || bo < 0 || bo > 2)
      return fail(GX_E_ARG, w + ": unknown boundary condition");
    if ((bi == GX_BC_PERIODIC) != (bo == GX_BC_PERIODIC))
      return fail(GX_E_ARG, w + ": PERIODIC must hold on both faces of a dimension");
    if (!(g.MaxCoordinate[d] > g.MinCoordinate[d]))
      return fail(GX_E_ARG, w + ": MaxCoordinate must exceed MinCoordinate");
  }
  long long cells = 1;
  for (int d = 0; d < 3; d++) cells *= g.nCells[d];
  // byte offsets within one field stay below 2^32
  if (cells >= (1ll << 29)) return fail(GX_E_ARG, w + ": at most 2^29 - 1 cells per chart");
  return GX_OK;
}

// Per-index 1-D geometry of one dimension (GeoArray order), padded index
// p = c + 2 for c in [-2, n+2): the same construction as the reference chart
// (Chart_Template.f90:463-593), including single-precision ghost widths
// (real(KDL) at :477-478).
inline void geometry_1d(int n, double lo, double hi, std::vector<double> out[N_GEO],
                        std::vector<double> *center = nullptr) {
  const int g = 2;
  std::vector<double> e(n + 2 * g + 1);
  auto E = [&](int i) -> double & { return e[i + g - 1]; };
  E(1) = lo;
  const double width = (hi - lo) / n;
  for (int ic = 2; ic <= n + 1; ic++) E(ic) = E(ic - 1) + width;
  for (int ic = 1; ic <= g; ic++) {
    const float w_ig = (float)(E(ic + 1) - E(ic));
    const float w_og = (float)(E(n - ic + 2) - E(n - ic + 1));
    E(1 - ic) = E(2 - ic) - (double)w_ig;
    E(n + 1 + ic) = E(n + ic) + (double)w_og;
  }
  const int m = n + 2 * g;
  std::vector<double> c(m);
  for (int a = 0; a < N_GEO; a++) out[a].assign(m, 0.0);
  for (int p = 0; p < m; p++) {
    const int ic = p - g + 1;
    c[p] = 0.5 * (E(ic) + E(ic + 1));
    out[G_WL][p] = c[p] - E(ic);
    out[G_WR][p] = E(ic + 1) - c[p];
  }
  if (center) *center = c;
  for (int p = 0; p < m; p++) {
    if (p >= 1 && p + 1 < m) {
      const double dxl = c[p] - c[p - 1];
      const double dxr = c[p + 1] - c[p];
      out[G_DXL][p] = dxl;
      out[G_DXR][p] = dxr;
      out[G_DXL2][p] = dxl * dxl;
      out[G_DXR2][p] = dxr * dxr;
      out[G_DEN][p] = dxl * dxr * (dxl + dxr);
    } else {
      out[G_DXL][p] = out[G_DXR][p] = out[G_DXL2][p] = out[G_DXR2][p] = 1.0;
      out[G_DEN][p] = 1.0;
    }
    out[G_WIDTH][p] = out[G_WL][p] + out[G_WR][p];
    out[G_RDXL][p] = 1.0 / out[G_DXL][p];
    out[G_RDXR][p] = 1.0 / out[G_DXR][p];
    out[G_RDEN][p] = 1.0 / out[G_DEN][p];
    out[G_CENTER][p] = c[p];
  }
}

// Geometry of one chart: the three 1-D geometries, stride gs = max(n) + 4.
struct ChartGeometry {
  int n[3] = {0, 0, 0};
  int gs = 0;
  std::vector<double> a[3][N_GEO];
  std::vector<double> center[3];  // Center (Chart_Template.f90:827-841), padded
  // every centre spacing, slope denominator and cell width is a power of two
  // (then so is every cell volume, and every division by them is exact as a
  // multiplication by the reciprocal: the stage kernels' P2 variant)
  bool pow2 = false;
};

// Builds the geometry and checks the divisors of the exactly rounded
// reciprocal division: centre spacings, slope denominators and cell volumes
// must lie in [2^-500, 2^122].
inline int make_geometry(const gx_grid &grid, ChartGeometry &cg, const char *who) {
  cg.gs = 0;
  for (int d = 0; d < 3; d++) {
    cg.n[d] = grid.nCells[d];
    geometry_1d(cg.n[d], grid.MinCoordinate[d], grid.MaxCoordinate[d], cg.a[d],
                &cg.center[d]);
    cg.gs = cg.n[d] + 4 > cg.gs ? cg.n[d] + 4 : cg.gs;
  }
  const double lo = 0x1p-500, hi = 0x1p122;
  double vmin = 1.0, vmax = 1.0;
  bool ok = true;
  auto is_pow2 = [](double x) {
    int e;
    return x > 0.0 && frexp(x, &e) == 0.5;
  };
  cg.pow2 = true;
  for (int d = 0; d < 3; d++) {
    const std::vector<double> *g = cg.a[d];
    double wmin = HUGE_VAL, wmax = 0.0;
    for (int q = 0; q < cg.n[d] + 4; q++) {
      for (int k : {G_DXL, G_DXR, G_DEN}) {
        ok = ok && g[k][q] >= lo && g[k][q] <= hi;
        cg.pow2 = cg.pow2 && is_pow2(g[k][q]);
      }
      if (q >= 2 && q < cg.n[d] + 2) {
        wmin = fmin(wmin, g[G_WIDTH][q]);
        wmax = fmax(wmax, g[G_WIDTH][q]);
        cg.pow2 = cg.pow2 && is_pow2(g[G_WIDTH][q]);
      }
    }
    vmin *= wmin;
    vmax *= wmax;
  }
  if (!ok || !(vmin >= lo) || !(vmax <= hi))
    return fail(GX_E_ARG,
                std::string(who) + ": cell spacings and volumes must lie in [2^-500, 2^122]");
  return GX_OK;
}

// The device image of a geometry: array g of dimension d at (d*N_GEO + g)*gs,
// in double, or in float with the reciprocals of the rounded values.
inline std::vector<char> geometry_image(const ChartGeometry &cg, bool fp32) {
  const size_t len = 3 * (size_t)N_GEO * cg.gs;
  std::vector<double> host(len, 1.0);
  for (int d = 0; d < 3; d++)
    for (int a = 0; a < N_GEO; a++)
      memcpy(&host[((size_t)d * N_GEO + a) * cg.gs], cg.a[d][a].data(),
             sizeof(double) * (cg.n[d] + 4));
  std::vector<char> out;
  if (!fp32) {
    out.resize(len * sizeof(double));
    memcpy(out.data(), host.data(), out.size());
    return out;
  }
  std::vector<float> h32(len);
  for (size_t i = 0; i < len; i++) h32[i] = (float)host[i];
  const int recip[3][2] = {{G_RDXL, G_DXL}, {G_RDXR, G_DXR}, {G_RDEN, G_DEN}};
  for (int d = 0; d < 3; d++)
    for (int r = 0; r < 3; r++)
      for (int q = 0; q < cg.gs; q++)
        h32[((size_t)d * N_GEO + recip[r][0]) * cg.gs + q] =
            1.0f / h32[((size_t)d * N_GEO + recip[r][1]) * cg.gs + q];
  out.resize(len * sizeof(float));
  memcpy(out.data(), h32.data(), out.size());
  return out;
}

inline int grid_blocks(long long ncell) {
  long long b = (ncell + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

// n host doubles into a device state buffer of the given precision, and
// back.  Both synchronise the stream.
inline int put_values(hipStream_t s, bool fp32, void *dst, const double *src, size_t n) {
  if (!fp32) {
    HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, s));
  } else {
    std::vector<float> t(n);
    for (size_t i = 0; i < n; i++) t[i] = (float)src[i];
    HIP_TRY(hipMemcpyAsync(dst, t.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  return GX_OK;
}

inline int get_values(hipStream_t s, bool fp32, double *dst, const void *src, size_t n) {
  if (!fp32) {
    HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  } else {
    std::vector<float> t(n);
    HIP_TRY(hipMemcpyAsync(t.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; i++) dst[i] = (double)t[i];
  }
  return GX_OK;
}

// Runs f(R()) with R = float when fp32, else double.
template <typename F>
inline int by_precision(bool fp32, F &&f) {
  return fp32 ? f(float()) : f(double());
}

}  // namespace gx
