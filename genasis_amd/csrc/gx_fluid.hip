// gx_fluid.hip -- MI355X (gfx950) finite-volume step for the GenASiS
// polytropic ideal-gas fluid, behind the C-ABI of include/gx_fluid.h: one
// chart per context.
//
// One RK2 step = two launches of the fused stage kernel (reconstruction,
// fluxes and divergence of all three dimensions, plus the RK combination;
// blocks appended to it record the boundary fluence), one launch of the
// shock-flag kernel for the next step, and one small launch that adds the
// step's surface integral to the boundary tally.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/gx_fluid.h"
#include "gx_host.hpp"
#include "gx_kernels.hpp"
#include "gx_tiles.hpp"
#include "gx_tally.hpp"
#include "gx_context.hpp"

namespace gx {

// Self-test of div_rcp against the IEEE division on pseudo-random operands:
// |a| in [2^-60, 2^67], b in [2^-32, 2^31], random signs and mantissas; one
// pair in 64 each has a numerator of +-0, +-inf, NaN, a normal below the fast
// range (2^-950), a subnormal or a large value (2^600).  NaN results match
// when both are NaN.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void division_selftest_kernel(
    long long n, unsigned long long seed, unsigned long long *mismatches) {
  unsigned long long bad = 0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256) {
    const unsigned long long r1 = mix64(seed ^ (2ull * i));
    const unsigned long long r2 = mix64(seed ^ (2ull * i + 1));
    const long long ea = (long long)(r1 >> 57) - 60;  // [-60, 67]
    const long long eb = (long long)(r2 >> 58) - 32;  // [-32, 31]
    const double ma = 1.0 + (double)(r1 & ((1ull << 52) - 1)) * 0x1p-52;
    const double mb = 1.0 + (double)(r2 & ((1ull << 52) - 1)) * 0x1p-52;
    double a = ldexp(ma, (int)ea), b = ldexp(mb, (int)eb);
    switch ((int)((r2 >> 52) & 63)) {
      case 0: a = 0.0; break;
      case 1: a = ldexp(ma, 900); break;
      case 2: a = ldexp(ma, -950); break;
      case 3: a = ldexp(ma, -1060); break;
      case 4: a = ldexp(ma, 600); break;
      case 5: a = __longlong_as_double(0x7ff8000000000001ll); break;
      default: break;
    }
    if (r1 & (1ull << 52)) a = -a;
    const double y = 1.0 / b;
    const double q1 = div_rcp(a, b, y);
    const double q2 = a / b;
    const bool both_nan = q1 != q1 && q2 != q2;
    // the exact domain: zeros, and |a| >= 2^-969 with a normal quotient;
    // elsewhere one ulp is allowed
    const bool exact = a == 0.0 || (fabs(a) >= 0x1p-969 && fabs(q2) >= 0x1p-1022);
    const long long d = __double_as_longlong(q1) - __double_as_longlong(q2);
    if (!both_nan && (exact ? d != 0 : (d > 1 || d < -1))) bad++;
  }
  if (bad) atomicAdd(mismatches, bad);
}

}  // namespace gx

// ---------------------------------------------------------------------------
// Scratch of the allocation-free time-step candidate
// (gx_compute_time_step_noalloc): the CFL accumulator in the first 256 bytes,
// then a chart geometry image of 3 x N_GEO rows of NOALLOC_GS values (only
// the WidthLeft + WidthRight rows are written).  1.18 MB per GPU, held until
// the process ends; every call on a GPU shares it.
// ---------------------------------------------------------------------------
namespace gx {
constexpr int NOALLOC_MAX_EDGE = 4096;
constexpr int NOALLOC_GS = NOALLOC_MAX_EDGE + 4;
constexpr size_t NOALLOC_GEO_OFFSET = 32;  // doubles
constexpr size_t NOALLOC_DOUBLES = NOALLOC_GEO_OFFSET + 3 * (size_t)N_GEO * NOALLOC_GS;
}  // namespace gx

static __device__ __align__(256) double scratch[gx::NOALLOC_DOUBLES];

namespace gx {

// Clears the accumulator and writes WidthLeft + WidthRight of the proper
// cells of each dimension (one thread per dimension), by the construction
// of geometry_1d: edges by repeated addition of (Max - Min)/n, centre
// 0.5 (E(i) + E(i+1)), widths centre - E(i) and E(i+1) - centre.
template <typename R>
__global__ __launch_bounds__(64) void noalloc_setup_kernel(gx_grid grid, double *scr) {
  const int d = threadIdx.x;
  if (d == 0) *(unsigned long long *)scr = 0ull;
  if (d >= 3) return;
  R *geo = (R *)(scr + NOALLOC_GEO_OFFSET) + (size_t)(d * N_GEO + G_WIDTH) * NOALLOC_GS;
  const int n = grid.nCells[d];
  const double width = (grid.MaxCoordinate[d] - grid.MinCoordinate[d]) / n;
  double e0 = grid.MinCoordinate[d];
  for (int i = 0; i < n; i++) {
    const double e1 = e0 + width;
    const double c = 0.5 * (e0 + e1);
    geo[i + 2] = (R)((c - e0) + (e1 - c));
    e0 = e1;
  }
}

// ComputeTimeStep_C_ASC's tail on the device: CourantFactor * min(huge,
// 1/max(tiny, invdt)) from the accumulated maximum.
__global__ void noalloc_candidate_kernel(const double *scr, double courant, double *candidate) {
  const double invdt = __longlong_as_double(*(const long long *)scr);
  const double tsi = fmax(DBL_MIN, invdt);
  const double ts = fmin(DBL_MAX, 1.0 / tsi);
  *candidate = courant * ts;
}

}  // namespace gx

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

using namespace gx;

thread_local char gx::g_last_error[gx::ERROR_CAP];


// The stage kernel and the boundary-line kernel of the context's
// ReconstructionType.
template <typename R, int STAGE>
static void launch_stage(const gx_ctx *c, int nblk, hipStream_t st, const StageArgs<R> &a) {
  const bool para = c->params.ReconstructionType == GX_PARABOLIC;
  if (a.sources) {
    if (para)
      stage_kernel_parabolic<R, STAGE, true><<<nblk, NT, 0, st>>>(a);
    else
      stage_kernel<R, STAGE, true><<<nblk, NT, 0, st>>>(a);
  } else if (para) {
    stage_kernel_parabolic<R, STAGE><<<nblk, NT, 0, st>>>(a);
  } else {
    if constexpr (std::is_same<R, double>::value) {
      if (c->geom.pow2) {
        // every geometry divisor a power of two: the same bits, one
        // multiplication per division by the geometry
        if (stage_shape(c) == 1)
          stage_kernel<R, STAGE, false, false, Tile32x8, true><<<nblk, NT, 0, st>>>(a);
        else
          stage_kernel<R, STAGE, false, false, Tile16x16, true><<<nblk, NT, 0, st>>>(a);
        return;
      }
    }
    if (stage_shape(c) == 1)
      stage_kernel<R, STAGE, false, false, Tile32x8><<<nblk, NT, 0, st>>>(a);
    else
      stage_kernel<R, STAGE><<<nblk, NT, 0, st>>>(a);
  }
}

template <typename R, int STAGE>
static void launch_boundary(const gx_ctx *c, hipStream_t st, const StageArgs<R> &a,
                            const BoundaryArgs &B) {
  if (c->params.ReconstructionType == GX_PARABOLIC)
    boundary_flux_kernel<R, STAGE, RECON_PARABOLIC><<<c->n_bpartial, TALLY_BLOCK, 0, st>>>(a, B);
  else
    boundary_flux_kernel<R, STAGE, RECON_LINEAR><<<c->n_bpartial, TALLY_BLOCK, 0, st>>>(a, B);
}

extern "C" {

void gx_default_params(gx_params *p) {
  if (!p) return;
  p->AdiabaticIndex = 1.4;
  p->LimiterParameter = 1.4;
  p->ShockThreshold = 0.1;
  p->CourantFactor = 0.7;
  p->BaryonDensityMin = 0.0;
  p->BaryonMassReference = 1.0;
  p->UseLimiter = 1;
  p->RiemannSolverType = GX_HLLC;
  p->Precision = GX_FP64;
  p->ReconstructionType = GX_LINEAR;
}

const char *gx_last_error(void) { return g_last_error; }

int gx_self_test_division(long long n, unsigned long long seed,
                          long long *mismatches) {
  GX_API_BEGIN
  if (n < 0 || !mismatches) return fail(GX_E_ARG, "gx_self_test_division: bad argument");
  unsigned long long *d = nullptr;
  HIP_TRY(hipMalloc(&d, sizeof(unsigned long long)));
  HIP_TRY(hipMemset(d, 0, sizeof(unsigned long long)));
  division_selftest_kernel<<<4096, 256>>>(n, seed, d);
  hipError_t e = hipGetLastError();
  unsigned long long h = 0;
  if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess)
    return fail(GX_E_HIP, std::string("gx_self_test_division: ") + hipGetErrorString(e));
  *mismatches = (long long)h;
  return GX_OK;
  GX_API_END("gx_self_test_division")
}

int gx_destroy(gx_ctx *c) {
  GX_API_BEGIN
  if (!c) return GX_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  gx::release_adjoint(c);
  gx::release_bricks(c);
  for (int k = 0; k < 2; k++)
    for (int e = 0; e < 6; e++)
      if (c->ev[k][e]) (void)hipEventDestroy(c->ev[k][e]);
  if (c->ev_cfl) (void)hipEventDestroy(c->ev_cfl);
  if (c->invdt_host) (void)hipHostFree(c->invdt_host);
  (void)hipFree(c->fluence);
  (void)hipFree(c->tgeo);
  (void)hipFree(c->tally_local);
  (void)hipFree(c->bpartial);
  (void)hipFree(c->tpartial);
  (void)hipFree(c->geo_buf);
  (void)hipFree(c->U);
  (void)hipFree(c->Y);
  (void)hipFree(c->W);
  (void)hipFree(c->P0);
  (void)hipFree(c->flags);
  (void)hipFree(c->sources);
  (void)hipFree(c->invdt);
  (void)hipFree(c->partial);
  (void)hipFree(c->scratch);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GX_OK;
  GX_API_END("gx_destroy")
}

int gx_create(const gx_grid *grid, const gx_params *params, gx_ctx **out) {
  GX_API_BEGIN
  if (!grid || !out) return fail(GX_E_ARG, "gx_create: null argument");
  *out = nullptr;
  gx_params p;
  if (params)
    p = *params;
  else
    gx_default_params(&p);
  int rc = check_params(p, "gx_create");
  if (rc) return rc;
  rc = check_grid(*grid, "gx_create");
  if (rc) return rc;
  gx_ctx *c = new (std::nothrow) gx_ctx();
  if (!c) return fail(GX_E_NOMEM, "gx_create: host allocation failed");
  c->grid = *grid;
  c->params = p;
  c->fp32 = p.Precision == GX_FP32;
  c->esz = c->fp32 ? sizeof(float) : sizeof(double);
  c->ncell = 1;
  for (int d = 0; d < 3; d++) {
    c->n[d] = grid->nCells[d];
    c->ncell *= c->n[d];
  }
  c->tz = stage_tz(((c->n[0] + TX - 1) / TX) * ((c->n[1] + TY - 1) / TY), c->n[2]);
#define TRY_CLEAN(expr)                                                   \
  do {                                                                    \
    hipError_t e_ = (expr);                                               \
    if (e_ != hipSuccess) {                                               \
      std::string m_ = std::string(#expr ": ") + hipGetErrorString(e_);   \
      gx_destroy(c);                                                      \
      return fail(GX_E_HIP, m_);                                          \
    }                                                                     \
  } while (0)
  rc = make_geometry(*grid, c->geom, "gx_create");
  if (rc) {
    gx_destroy(c);
    return rc;
  }
  c->gs = c->geom.gs;
  TRY_CLEAN(hipStreamCreate(&c->stream));
  {
    const std::vector<char> img = geometry_image(c->geom, c->fp32);
    TRY_CLEAN(hipMalloc(&c->geo_buf, img.size()));
    TRY_CLEAN(hipMemcpy(c->geo_buf, img.data(), img.size(), hipMemcpyHostToDevice));
  }
  const size_t bytes = c->esz * (size_t)c->ncell;
  TRY_CLEAN(hipMalloc(&c->U, 5 * bytes));
  TRY_CLEAN(hipMalloc(&c->Y, 5 * bytes));
  TRY_CLEAN(hipMalloc(&c->W, 5 * bytes));
  TRY_CLEAN(hipMalloc(&c->P0, 5 * bytes));
  TRY_CLEAN(hipMalloc(&c->flags, (size_t)c->ncell));
  TRY_CLEAN(hipMemset(c->flags, 0, (size_t)c->ncell));
  TRY_CLEAN(hipMalloc(&c->invdt, sizeof(unsigned long long)));
  TRY_CLEAN(hipHostMalloc((void **)&c->invdt_host, sizeof(unsigned long long)));
  TRY_CLEAN(hipEventCreateWithFlags(&c->ev_cfl, hipEventDisableTiming));
  c->n_partial = grid_blocks(c->ncell);
  if (c->n_partial > 1024) c->n_partial = 1024;
  TRY_CLEAN(hipMalloc(&c->partial, sizeof(double) * 5 * c->n_partial));
  {
    const size_t nf = (size_t)fluence_count(c->n);
    TRY_CLEAN(hipMalloc(&c->fluence, sizeof(double) * nf));
    TRY_CLEAN(hipMemset(c->fluence, 0, sizeof(double) * nf));
    std::vector<double> tg((size_t)3 * N_TG * c->gs, 0.0);
    for (int d = 0; d < 3; d++)
      for (int p = 0; p < c->n[d] + 4; p++) {
        const double ctr = c->geom.center[d][p], wl = c->geom.a[d][G_WL][p];
        tg[(size_t)(3 * d + TG_CENTER) * c->gs + p] = ctr;
        tg[(size_t)(3 * d + TG_WIDTH) * c->gs + p] = wl + c->geom.a[d][G_WR][p];
        tg[(size_t)(3 * d + TG_FACE) * c->gs + p] = ctr - wl;
      }
    TRY_CLEAN(hipMalloc(&c->tgeo, sizeof(double) * tg.size()));
    TRY_CLEAN(hipMemcpy(c->tgeo, tg.data(), sizeof(double) * tg.size(), hipMemcpyHostToDevice));
    TRY_CLEAN(hipMalloc(&c->tally_local, sizeof(double) * TALLY_N));
    TRY_CLEAN(hipMemset(c->tally_local, 0, sizeof(double) * TALLY_N));
    c->n_bpartial = (int)((boundary_lines(c->n) + TALLY_BLOCK - 1) / TALLY_BLOCK);
    TRY_CLEAN(hipMalloc(&c->bpartial, sizeof(double) * NB_SUM * c->n_bpartial));
    TRY_CLEAN(hipMalloc(&c->tpartial, sizeof(double) * NI_SUM * c->n_partial));
  }
#undef TRY_CLEAN
  *out = c;
  return GX_OK;
  GX_API_END("gx_create")
}

int gx_set_tuning(gx_ctx *c, int zChunk) {
  GX_API_BEGIN
  if (!c || zChunk < 1) return fail(GX_E_ARG, "gx_set_tuning: bad argument");
  c->tz = std::min(zChunk, c->n[2]);  // a longer chunk is the whole chart
  return GX_OK;
  GX_API_END("gx_set_tuning")
}

int gx_set_tile(gx_ctx *c, int tileX, int tileY) {
  GX_API_BEGIN
  if (!c) return fail(GX_E_ARG, "gx_set_tile: null context");
  int shape = -1;
  for (int k = 0; k < N_TILE_SHAPES; k++)
    if (tileX == tile_x(k) && tileY == tile_y(k)) shape = k;
  if (shape < 0) return fail(GX_E_ARG, "gx_set_tile: the stage tile is 16 x 16 or 32 x 8");
  if (shape != 0 && c->params.ReconstructionType != GX_LINEAR)
    return fail(GX_E_ARG, "gx_set_tile: PARABOLIC steps take the 16 x 16 tile");
  c->tile = shape;
  return GX_OK;
  GX_API_END("gx_set_tile")
}

int gx_get_tile(gx_ctx *c, int *tileX, int *tileY) {
  GX_API_BEGIN
  if (!c || !tileX || !tileY) return fail(GX_E_ARG, "gx_get_tile: null argument");
  *tileX = tile_x(c->tile);
  *tileY = tile_y(c->tile);
  return GX_OK;
  GX_API_END("gx_get_tile")
}

int gx_set_record_sources(gx_ctx *c, int on) {
  GX_API_BEGIN
  if (!c) return fail(GX_E_ARG, "gx_set_record_sources: null context");
  if (c->bricks) return fail(GX_E_STATE, "gx_set_record_sources: not available on a chart split into bricks");
  if (on) {
    const size_t bytes = 5 * c->esz * (size_t)c->ncell;
    if (!c->sources) HIP_TRY(hipMalloc(&c->sources, bytes));
    HIP_TRY(hipMemsetAsync(c->sources, 0, bytes, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  c->record_sources = on != 0;
  return GX_OK;
  GX_API_END("gx_set_record_sources")
}

int gx_get_sources(gx_ctx *c, double *sources) {
  GX_API_BEGIN
  if (!c || !sources) return fail(GX_E_ARG, "gx_get_sources: null argument");
  if (c->bricks) return fail(GX_E_STATE, "gx_get_sources: not available on a chart split into bricks");
  if (!c->sources)
    return fail(GX_E_STATE, "gx_get_sources: recording was never enabled (gx_set_record_sources)");
  return get_values(c->stream, c->fp32, sources, c->sources, 5 * (size_t)c->ncell);
  GX_API_END("gx_get_sources")
}

int gx_get_geometry_pow2(gx_ctx *c, int *pow2) {
  GX_API_BEGIN
  if (!c || !pow2) return fail(GX_E_ARG, "gx_get_geometry_pow2: null argument");
  *pow2 = c->geom.pow2 ? 1 : 0;
  return GX_OK;
  GX_API_END("gx_get_geometry_pow2")
}

int gx_get_tuning(gx_ctx *c, int *zChunk) {
  GX_API_BEGIN
  if (!c || !zChunk) return fail(GX_E_ARG, "gx_get_tuning: null argument");
  *zChunk = c->tz;
  return GX_OK;
  GX_API_END("gx_get_tuning")
}

void *gx_stream(gx_ctx *c) { return c ? (void *)c->stream : nullptr; }

int gx_set_profiling(gx_ctx *c, int on) {
  GX_API_BEGIN
  if (!c) return fail(GX_E_ARG, "gx_set_profiling: null context");
  if (c->bricks) return fail(GX_E_STATE, "gx_set_profiling: not available on a chart split into bricks");
  if (on && !c->ev[0][0])
    for (int k = 0; k < 2; k++)
      for (int e = 0; e < 6; e++) HIP_TRY(hipEventCreate(&c->ev[k][e]));
  for (int k = 0; k < 2; k++) {  // the steps recorded so far are dropped
    if (c->ev_used[k]) HIP_TRY(hipEventSynchronize(c->ev[k][5]));
    c->ev_used[k] = false;
  }
  c->profiling = on != 0;
  for (int k = 0; k < 4; k++) {
    c->kernel_ms[k] = 0.0;
    c->kernel_n[k] = 0;
  }
  return GX_OK;
  GX_API_END("gx_set_profiling")
}

int gx_kernel_times(gx_ctx *c, double ms[4], long long count[4]) {
  GX_API_BEGIN
  if (!c || !ms || !count) return fail(GX_E_ARG, "gx_kernel_times: null argument");
  for (int k = 0; k < 2; k++) {
    const int rc = gx::collect_times(c, k);
    if (rc) return rc;
  }
  for (int k = 0; k < 4; k++) {
    ms[k] = c->kernel_ms[k];
    count[k] = c->kernel_n[k];
  }
  return GX_OK;
  GX_API_END("gx_kernel_times")
}

// A new state starts new tallies: fluence, boundary tallies, total and
// change cleared (the reference's tallies start at zero with the Current).
static int reset_tallies(gx_ctx *c) {
  HIP_TRY(hipMemsetAsync(c->fluence, 0, sizeof(double) * (size_t)fluence_count(c->n),
                         c->stream));
  HIP_TRY(hipMemsetAsync(c->tally_local, 0, sizeof(double) * TALLY_N, c->stream));
  for (int t = 0; t < TALLY_N; t++)
    c->tally_global[t] = c->tally_total[t] = c->tally_change[t] = 0.0;
  return GX_OK;
}

int gx_set_state(gx_ctx *c, const double *primitive, const double *conserved) {
  GX_API_BEGIN
  if (!c || !primitive || !conserved) return fail(GX_E_ARG, "gx_set_state: null argument");
  const size_t n = 5 * (size_t)c->ncell;
  int rc;
  if (c->bricks) {
    rc = gx::bricks_set_state(c, primitive, conserved);
    if (rc) return rc;
  } else {
    rc = put_values(c->stream, c->fp32, c->P0, primitive, n);
    if (rc) return rc;
    rc = put_values(c->stream, c->fp32, c->U, conserved, n);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(c->flags, 0, (size_t)c->ncell, c->stream));
  }
  rc = reset_tallies(c);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->uploaded = true;
  c->prims_valid = true;
  c->invdt_valid = false;
  c->cfl_pending = false;
  return GX_OK;
  GX_API_END("gx_set_state")
}

int gx_set_primitive(gx_ctx *c, const double *primitive) {
  GX_API_BEGIN
  if (!c || !primitive) return fail(GX_E_ARG, "gx_set_primitive: null argument");
  if (c->bricks) {
    int rb = gx::bricks_set_state(c, primitive, nullptr);
    if (rb) return rb;
    rb = reset_tallies(c);
    if (rb) return rb;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->uploaded = true;
    c->prims_valid = true;
    c->invdt_valid = false;
    c->cfl_pending = false;
    return GX_OK;
  }
  const int rc = put_values(c->stream, c->fp32, c->P0, primitive, 5 * (size_t)c->ncell);
  if (rc) return rc;
  by_precision(c->fp32, [&](auto tag) {
    using R = decltype(tag);
    from_primitive_kernel<R><<<grid_blocks(c->ncell), 256, 0, c->stream>>>(
        c->ncell, c->ncell, c->params.BaryonMassReference, (R *)c->P0, (R *)c->U);
    return 0;
  });
  HIP_TRY(hipGetLastError());
  const int rt = reset_tallies(c);
  if (rt) return rt;
  HIP_TRY(hipMemsetAsync(c->flags, 0, (size_t)c->ncell, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->uploaded = true;
  c->prims_valid = true;
  c->invdt_valid = false;
  c->cfl_pending = false;
  return GX_OK;
  GX_API_END("gx_set_primitive")
}

static const int PRIM_FIELDS[5] = {GX_COMOVING_BARYON_DENSITY, GX_VELOCITY_U_1,
                                   GX_VELOCITY_U_2, GX_VELOCITY_U_3,
                                   GX_INTERNAL_ENERGY};
int gx_upload(gx_ctx *c, const double *value, long long ld, const int *map) {
  GX_API_BEGIN
  if (!c || !value || !map) return fail(GX_E_ARG, "gx_upload: null argument");
  const int m0 = c->n[0] + 4, m1 = c->n[1] + 4, m2 = c->n[2] + 4;
  if (ld < (long long)m0 * m1 * m2) return fail(GX_E_ARG, "gx_upload: ldValue < nValues");
  for (int f = 0; f < 5; f++)
    if (map[PRIM_FIELDS[f]] <= 0)
      return fail(GX_E_ARG, "gx_upload: the primitive fields N, V_1..3, E are required");
  // PrepareEvolution: the Current's proper-cell primitives; the conserved
  // fields follow from them on the device (ComputeFromInitial).
  std::vector<double> prim(5 * (size_t)c->ncell);
  for (int f = 0; f < 5; f++) {
    const double *vp = value + (size_t)(map[PRIM_FIELDS[f]] - 1) * ld;
    for (int z = 0; z < c->n[2]; z++)
      for (int y = 0; y < c->n[1]; y++)
        for (int x = 0; x < c->n[0]; x++) {
          const size_t src = (size_t)(x + 2) + (size_t)m0 * ((y + 2) + (size_t)m1 * (z + 2));
          const size_t dst = (size_t)f * c->ncell + x + (size_t)c->n[0] * (y + (size_t)c->n[1] * z);
          prim[dst] = vp[src];
        }
  }
  return gx_set_primitive(c, prim.data());
  GX_API_END("gx_upload")
}

int gx_get_conserved(gx_ctx *c, double *conserved) {
  GX_API_BEGIN
  if (!c || !conserved) return fail(GX_E_ARG, "gx_get_conserved: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_get_conserved: no state uploaded");
  if (c->bricks) return gx::bricks_get(c, 0, conserved);
  return get_values(c->stream, c->fp32, conserved, c->U, 5 * (size_t)c->ncell);
  GX_API_END("gx_get_conserved")
}

// The 23 Current fields of the proper cells in chart order into host
// (GX_N_FIELDS x ncell doubles).
static int current_to_host(gx_ctx *c, double *host);

static int run_current(gx_ctx *c) {
  if (!c->scratch)
    HIP_TRY(hipMalloc(&c->scratch, sizeof(double) * GX_N_FIELDS * (size_t)c->ncell));
  // c_v = k/(mu(gamma-1)) with the default gamma (Fluid_P_I__Form.f90:206-223)
  const double c_v = 1.0 / (1.0 * (1.4 - 1.0));
  by_precision(c->fp32, [&](auto tag) {
    using R = decltype(tag);
    StageArgs<R> a = base_args<R>(c);
    a.aux = (const R *)c->U;
    current_kernel<R><<<grid_blocks(c->ncell), 256, 0, c->stream>>>(
        a, c_v, c->prims_valid ? (const R *)c->P0 : nullptr, c->scratch);
    return 0;
  });
  HIP_TRY(hipGetLastError());
  return GX_OK;
}

static int current_to_host(gx_ctx *c, double *host) {
  if (c->bricks) return gx::bricks_current(c, host);
  int rc = run_current(c);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(host, c->scratch, sizeof(double) * GX_N_FIELDS * (size_t)c->ncell,
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
}

int gx_get_primitive(gx_ctx *c, double *primitive) {
  GX_API_BEGIN
  if (!c || !primitive) return fail(GX_E_ARG, "gx_get_primitive: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_get_primitive: no state uploaded");
  const size_t bytes = sizeof(double) * (size_t)c->ncell;
  if (c->bricks) {
    if (c->prims_valid) return gx::bricks_get(c, 1, primitive);
    std::vector<double> host(GX_N_FIELDS * (size_t)c->ncell);
    const int rc = current_to_host(c, host.data());
    if (rc) return rc;
    for (int f = 0; f < 5; f++)
      memcpy(primitive + (size_t)f * c->ncell, host.data() + (size_t)PRIM_FIELDS[f] * c->ncell,
             bytes);
    return GX_OK;
  }
  if (c->prims_valid) {
    return get_values(c->stream, c->fp32, primitive, c->P0, 5 * (size_t)c->ncell);
  } else {
    int rc = run_current(c);
    if (rc) return rc;
    for (int f = 0; f < 5; f++)
      HIP_TRY(hipMemcpyAsync(primitive + (size_t)f * c->ncell,
                             c->scratch + (size_t)PRIM_FIELDS[f] * c->ncell, bytes,
                             hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
  GX_API_END("gx_get_primitive")
}

int gx_download(gx_ctx *c, double *value, long long ld, const int *map) {
  GX_API_BEGIN
  if (!c || !value || !map) return fail(GX_E_ARG, "gx_download: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_download: no state uploaded");
  const int m0 = c->n[0] + 4, m1 = c->n[1] + 4, m2 = c->n[2] + 4;
  if (ld < (long long)m0 * m1 * m2) return fail(GX_E_ARG, "gx_download: ldValue < nValues");
  std::vector<double> host(GX_N_FIELDS * (size_t)c->ncell);
  int rc = current_to_host(c, host.data());
  if (rc) return rc;
  const int g = 2;
  const int n0 = c->n[0], n1 = c->n[1], n2 = c->n[2];
  for (int f = 0; f < GX_N_FIELDS; f++) {
    if (map[f] <= 0) continue;
    double *col = value + (size_t)(map[f] - 1) * ld;
    const double *src = host.data() + (size_t)f * c->ncell;
    for (int pz = 0; pz < m2; pz++)
      for (int py = 0; py < m1; py++)
        for (int px = 0; px < m0; px++) {
          int q[3] = {px - g, py - g, pz - g};
          bool ok = true;
          for (int d = 0; d < 3; d++) {
            if (q[d] < 0 || q[d] >= c->n[d]) {
              if (c->grid.BoundaryCondition[2 * d] != GX_BC_PERIODIC) ok = false;
              q[d] = (q[d] + c->n[d]) % c->n[d];
            }
          }
          if (!ok) continue;
          col[(size_t)px + (size_t)m0 * (py + (size_t)m1 * pz)] =
              src[(size_t)q[0] + (size_t)n0 * (q[1] + (size_t)n1 * q[2])];
        }
    (void)n2;
  }
  return GX_OK;
  GX_API_END("gx_download")
}

int gx_get_flags(gx_ctx *c, unsigned char *flags) {
  GX_API_BEGIN
  if (!c || !flags) return fail(GX_E_ARG, "gx_get_flags: null argument");
  if (c->bricks) return gx::bricks_get_flags(c, flags);
  HIP_TRY(hipMemcpyAsync(flags, c->flags, (size_t)c->ncell, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
  GX_API_END("gx_get_flags")
}

int gx_set_flags(gx_ctx *c, const unsigned char *flags) {
  GX_API_BEGIN
  if (!c || !flags) return fail(GX_E_ARG, "gx_set_flags: null argument");
  if (c->bricks) return gx::bricks_set_flags(c, flags);
  HIP_TRY(hipMemcpyAsync(c->flags, flags, (size_t)c->ncell, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
  GX_API_END("gx_set_flags")
}

int gx_upload_features(gx_ctx *c, const double *df, long long ld) {
  GX_API_BEGIN
  if (!c || !df) return fail(GX_E_ARG, "gx_upload_features: null argument");
  const int m0 = c->n[0] + 4, m1 = c->n[1] + 4;
  std::vector<unsigned char> b((size_t)c->ncell, 0);
  for (int d = 0; d < 3; d++)
    for (int z = 0; z < c->n[2]; z++)
      for (int y = 0; y < c->n[1]; y++)
        for (int x = 0; x < c->n[0]; x++) {
          const size_t src = (size_t)d * ld + (size_t)(x + 2) + (size_t)m0 * ((y + 2) + (size_t)m1 * (z + 2));
          if (df[src] > 0.0)
            b[x + (size_t)c->n[0] * (y + (size_t)c->n[1] * z)] |= (unsigned char)(1u << d);
        }
  return gx_set_flags(c, b.data());
  GX_API_END("gx_upload_features")
}

int gx_download_features(gx_ctx *c, double *df, long long ld) {
  GX_API_BEGIN
  if (!c || !df) return fail(GX_E_ARG, "gx_download_features: null argument");
  const int m0 = c->n[0] + 4, m1 = c->n[1] + 4;
  std::vector<unsigned char> b((size_t)c->ncell);
  int rc = gx_get_flags(c, b.data());
  if (rc) return rc;
  for (int d = 0; d < 3; d++)
    for (int z = 0; z < c->n[2]; z++)
      for (int y = 0; y < c->n[1]; y++)
        for (int x = 0; x < c->n[0]; x++) {
          const size_t dst = (size_t)d * ld + (size_t)(x + 2) + (size_t)m0 * ((y + 2) + (size_t)m1 * (z + 2));
          df[dst] = ((b[x + (size_t)c->n[0] * (y + (size_t)c->n[1] * z)] >> d) & 1) ? 1.0 : 0.0;
        }
  return GX_OK;
  GX_API_END("gx_download_features")
}

int gx_compute_time_step(gx_ctx *c, double *candidate) {
  GX_API_BEGIN
  if (!c || !candidate) return fail(GX_E_ARG, "gx_compute_time_step: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_compute_time_step: no state uploaded");
  if (!c->invdt_valid && c->bricks) {
    const int rb = gx::bricks_cfl(c);
    if (rb) return rb;
  } else if (!c->invdt_valid) {
    HIP_TRY(hipMemsetAsync(c->invdt, 0, sizeof(unsigned long long), c->stream));
    by_precision(c->fp32, [&](auto tag) {
      using R = decltype(tag);
      StageArgs<R> a = base_args<R>(c);
      a.src_is_prim = c->prims_valid ? 1 : 0;
      a.src = (const R *)(c->prims_valid ? c->P0 : c->U);
      cfl_kernel<R><<<grid_blocks(c->ncell), 256, 0, c->stream>>>(a);
      return 0;
    });
    HIP_TRY(hipGetLastError());
  }
  unsigned long long bits = 0;
  if (c->invdt_valid && c->cfl_pending) {
    HIP_TRY(hipEventSynchronize(c->ev_cfl));
    bits = *c->invdt_host;
  } else {
    HIP_TRY(hipMemcpyAsync(&bits, c->invdt, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  double invdt;
  memcpy(&invdt, &bits, sizeof(invdt));
  c->invdt_valid = true;
  // ComputeTimeStepKernel_CSL tail and ComputeTimeStep_C_ASC
  const double tsi = fmax(DBL_MIN, invdt);
  const double ts = fmin(DBL_MAX, 1.0 / tsi);
  *candidate = c->params.CourantFactor * ts;
  return GX_OK;
  GX_API_END("gx_compute_time_step")
}

// The address of this GPU's scratch (hipGetSymbolAddress once per device,
// with the device current); the array's name is not a device address on the
// host.
static int noalloc_scratch(double **out) {
  static std::mutex mu;
  static double *ptr[64] = {};
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return fail(GX_E_HIP, "gx_compute_time_step_noalloc: no current HIP device");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GX_E_HIP, "gx_compute_time_step_noalloc: hipGetSymbolAddress of the "
                            "scratch failed");
    ptr[dev] = (double *)p;
  }
  *out = ptr[dev];
  return GX_OK;
}

int gx_compute_time_step_noalloc(const gx_grid *grid, const gx_params *params,
                                 const void *conserved, double *candidate, void *stream) {
  GX_API_BEGIN
  if (!grid || !params || !conserved || !candidate)
    return fail(GX_E_ARG, "gx_compute_time_step_noalloc: null argument");
  if (params->Precision != GX_FP64 && params->Precision != GX_FP32)
    return fail(GX_E_ARG, "gx_compute_time_step_noalloc: Precision must be GX_FP64 or GX_FP32");
  long long ncell = 1;
  for (int d = 0; d < 3; d++) {
    if (grid->nCells[d] < 1 || grid->nCells[d] > NOALLOC_MAX_EDGE)
      return fail(GX_E_ARG, "gx_compute_time_step_noalloc: nCells must lie in [1, 4096] "
                            "in every dimension (the size of its static scratch)");
    if (!(grid->MaxCoordinate[d] > grid->MinCoordinate[d]))
      return fail(GX_E_ARG, "gx_compute_time_step_noalloc: MaxCoordinate must exceed "
                            "MinCoordinate");
    ncell *= grid->nCells[d];
  }
  if (ncell >= (1ll << 29))
    return fail(GX_E_ARG, "gx_compute_time_step_noalloc: at most 2^29 - 1 cells");
  double *scr = nullptr;
  int rc = noalloc_scratch(&scr);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const bool fp32 = params->Precision == GX_FP32;
  by_precision(fp32, [&](auto tag) {
    using R = decltype(tag);
    StageArgs<R> a;
    memset(&a, 0, sizeof(a));
    for (int d = 0; d < 3; d++) a.n[d] = grid->nCells[d];
    for (int i = 0; i < 6; i++) a.bc[i] = grid->BoundaryCondition[i];
    a.gs = NOALLOC_GS;
    a.fs = ncell;
    a.ph.gamma = params->AdiabaticIndex;
    a.ph.mass = params->BaryonMassReference;
    a.ph.nmin = params->BaryonDensityMin;
    a.src = (const R *)conserved;
    a.src_is_prim = 0;
    a.geo = (const R *)(scr + NOALLOC_GEO_OFFSET);
    a.invdt_bits = (unsigned long long *)scr;
    noalloc_setup_kernel<R><<<1, 64, 0, st>>>(*grid, scr);
    cfl_kernel<R><<<grid_blocks(ncell), 256, 0, st>>>(a);
    return 0;
  });
  noalloc_candidate_kernel<<<1, 1, 0, st>>>(scr, params->CourantFactor, candidate);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char msg[256];
    snprintf(msg, sizeof(msg), "gx_compute_time_step_noalloc: launch failed: %s",
             hipGetErrorString(e));
    return fail(GX_E_HIP, msg);
  }
  return GX_OK;
  GX_API_END("gx_compute_time_step_noalloc")
}

int gx_step_rk2(gx_ctx *c, double time, double dt) {
  GX_API_BEGIN
  (void)time;
  if (!c) return fail(GX_E_ARG, "gx_step_rk2: null context");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_step_rk2: no state uploaded");
  if (c->bricks) return gx::bricks_step(c, dt);
  return gx::step_rk2(c, dt, nullptr);
  GX_API_END("gx_step_rk2")
}

}  // extern "C"

namespace gx {

int step_rk2(gx_ctx *c, double dt, const StepRecord *rec) {
  if (rec && c->fp32) return fail(GX_E_ARG, "step_rk2: a recorded step is FP64");
  const int shape = stage_shape(c);
  const int nbx = (c->n[0] + tile_x(shape) - 1) / tile_x(shape),
            nby = (c->n[1] + tile_y(shape) - 1) / tile_y(shape),
            nbz = (c->n[2] + c->tz - 1) / c->tz;
  const int fbx = (c->n[0] + TX - 1) / TX, fby = (c->n[1] + TY - 1) / TY;  // flag tiles
  const int nblk = nbx * nby * nbz;
  BoundaryArgs B;
  B.bf = c->fluence;
  B.fw = 0.5 * dt;  // Weight_RK * TimeStep, B = (1/2, 1/2)
  B.tg = c->tgeo;
  B.partial = c->bpartial;
  hipStream_t st = c->stream;
  const bool prof = c->profiling;
  const int slot = c->prof_step & 1;
  hipEvent_t *ev = c->ev[slot];
  if (prof) {
    const int rp = collect_times(c, slot);  // the step before last, long done
    if (rp) return rp;
    HIP_TRY(hipEventRecord(ev[0], st));
  }
  const int rc = by_precision(c->fp32, [&](auto tag) {
    using R = decltype(tag);
    // Stage 1: Current = uploaded primitives or the Solution.
    StageArgs<R> a1 = base_args<R>(c);
    a1.dt = dt;
    a1.src_is_prim = c->prims_valid ? 1 : 0;
    const R *Uin = rec ? (const R *)rec->Uin : (const R *)c->U;
    R *Y = rec ? (R *)rec->Y : (R *)c->Y, *W = rec ? (R *)rec->W : (R *)c->W;
    R *Uout = rec ? (R *)rec->Uout : (R *)c->U;
    a1.src = c->prims_valid ? (const R *)c->P0 : Uin;
    a1.aux = Uin;
    a1.out1 = Y;
    a1.out2 = W;
    a1.nbx = nbx;
    a1.nby = nby;
    a1.nblk = nblk;
    if (c->record_sources) {
      // Weight_RK * Increment / TimeStep with RN(1/dt) (div_rcp's range)
      a1.sources = (R *)c->sources;
      a1.rdt = (dt >= 0x1p-500 && dt <= 0x1p122) ? 1.0 / dt : 0.0;
    }
    launch_stage<R, 1>(c, nblk, st, a1);
    HIP_TRY(hipGetLastError());
    if (prof) HIP_TRY(hipEventRecord(ev[1], st));
    // Stage-1 boundary fluxes from the stage-1 input (stage 2 overwrites the
    // Solution).
    launch_boundary<R, 1>(c, st, a1, B);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(c->invdt, 0, sizeof(unsigned long long), st));
    if (prof) HIP_TRY(hipEventRecord(ev[2], st));
    // Stage 2: Current = Y.
    StageArgs<R> a2 = base_args<R>(c);
    a2.dt = dt;
    a2.src = Y;
    a2.aux = W;
    a2.out1 = Uout;
    a2.out2 = rec ? (R *)rec->Z : nullptr;
    a2.nbx = nbx;
    a2.nby = nby;
    a2.nblk = nblk;
    a2.sources = a1.sources;
    a2.rdt = a1.rdt;
    launch_stage<R, 2>(c, nblk, st, a2);
    HIP_TRY(hipGetLastError());
    if (prof) HIP_TRY(hipEventRecord(ev[3], st));
    // the next candidate's CFL maximum, for the host as soon as it is formed
    HIP_TRY(hipMemcpyAsync(c->invdt_host, c->invdt, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(c->ev_cfl, st));
    // Stage-2 boundary fluxes (input Y, the step's flags) and the step's
    // surface integral into the boundary tally.
    launch_boundary<R, 2>(c, st, a2, B);
    HIP_TRY(hipGetLastError());
    boundary_tally_finish<<<1, TALLY_BLOCK, 0, st>>>(c->bpartial, c->n_bpartial,
                                                     c->tally_local);
    HIP_TRY(hipGetLastError());
    if (prof) HIP_TRY(hipEventRecord(ev[4], st));
    // Shock flags of the next step from the new state.
    StageArgs<R> a3 = base_args<R>(c);
    a3.src = Uout;
    a3.tz = flag_tz(fbx * fby, c->n[2]);
    a3.nbx = fbx;
    a3.nby = fby;
    a3.nblk = fbx * fby * ((c->n[2] + a3.tz - 1) / a3.tz);
    shock_flags_kernel<R><<<a3.nblk, NT, 0, st>>>(a3, c->flags);
    HIP_TRY(hipGetLastError());
    return GX_OK;
  });
  if (rc) return rc;
  if (prof) {
    HIP_TRY(hipEventRecord(ev[5], st));
    c->ev_used[slot] = true;
    c->prof_step++;
  }
  c->prims_valid = false;
  c->invdt_valid = true;
  c->cfl_pending = true;
  return GX_OK;
}

// Adds the kernel times of the step recorded in event set k (waiting for it
// if it is still running) and frees the set.  The boundary intervals hold
// the invdt reset (a few-byte fill) and the 8-byte CFL copy.
int collect_times(gx_ctx *c, int k) {
  if (!c->ev_used[k]) return GX_OK;
  hipEvent_t *ev = c->ev[k];
  HIP_TRY(hipEventSynchronize(ev[5]));
  float t[5];
  for (int e = 0; e < 5; e++) HIP_TRY(hipEventElapsedTime(&t[e], ev[e], ev[e + 1]));
  c->kernel_ms[0] += t[0];
  c->kernel_ms[1] += t[2];
  c->kernel_ms[2] += t[4];
  c->kernel_ms[3] += t[1] + t[3];
  for (int j = 0; j < 4; j++) c->kernel_n[j] += 1;
  c->ev_used[k] = false;
  return GX_OK;
}

}  // namespace gx

extern "C" {


int gx_sync(gx_ctx *c) {
  GX_API_BEGIN
  if (!c) return fail(GX_E_ARG, "gx_sync: null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
  GX_API_END("gx_sync")
}

int gx_get_boundary_fluence(gx_ctx *c, double *fluence) {
  GX_API_BEGIN
  if (!c || !fluence) return fail(GX_E_ARG, "gx_get_boundary_fluence: null argument");
  const size_t n = (size_t)fluence_count(c->n);
  HIP_TRY(hipMemcpyAsync(fluence, c->fluence, sizeof(double) * n, hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
  GX_API_END("gx_get_boundary_fluence")
}

int gx_compute_tally(gx_ctx *c, int computeChange, double *interior, double *boundary,
                     double *total, double *change) {
  GX_API_BEGIN
  if (!c) return fail(GX_E_ARG, "gx_compute_tally: null context");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_compute_tally: no state uploaded");
  if (c->bricks) {
    const int rb = gx::bricks_interior_partials(c);
    if (rb) return rb;
  } else {
    by_precision(c->fp32, [&](auto tag) {
      using R = decltype(tag);
      StageArgs<R> a = base_args<R>(c);
      a.aux = (const R *)c->U;
      interior_tally_kernel<R><<<c->n_partial, TALLY_BLOCK, 0, c->stream>>>(
          a, c->prims_valid ? (const R *)c->P0 : nullptr, c->tgeo, c->tpartial);
      return 0;
    });
    HIP_TRY(hipGetLastError());
  }
  std::vector<double> h((size_t)NI_SUM * c->n_partial);
  double local[TALLY_N];
  HIP_TRY(hipMemcpyAsync(h.data(), c->tpartial, sizeof(double) * h.size(),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(local, c->tally_local, sizeof(local), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipMemsetAsync(c->tally_local, 0, sizeof(local), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  // Current_ASC % ComputeTallyTemplate (FLD/Current_ASC__Template.f90:239-349),
  // one chart and one boundary, no reduction across ranks.
  double in[TALLY_N] = {};
  for (int k = 0; k < NI_SUM; k++) {
    double s = 0.0;
    for (int b = 0; b < c->n_partial; b++) s += h[(size_t)b * NI_SUM + k];
    in[interior_slot(k)] = s;
  }
  for (int t = 0; t < TALLY_N; t++) {
    const double old_total = c->tally_total[t];
    c->tally_global[t] = c->tally_global[t] + local[t];
    c->tally_total[t] = in[t] + c->tally_global[t];
    if (computeChange) c->tally_change[t] = c->tally_change[t] + (c->tally_total[t] - old_total);
  }
  if (interior) memcpy(interior, in, sizeof(in));
  if (boundary) memcpy(boundary, c->tally_global, sizeof(in));
  if (total) memcpy(total, c->tally_total, sizeof(in));
  if (change) memcpy(change, c->tally_change, sizeof(in));
  return GX_OK;
  GX_API_END("gx_compute_tally")
}

int gx_totals(gx_ctx *c, double out[5]) {
  GX_API_BEGIN
  if (!c || !out) return fail(GX_E_ARG, "gx_totals: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_totals: no state uploaded");
  if (c->bricks) return gx::bricks_totals(c, out);
  by_precision(c->fp32, [&](auto tag) {
    using R = decltype(tag);
    StageArgs<R> a = base_args<R>(c);
    a.aux = (const R *)c->U;
    totals_kernel<R><<<c->n_partial, 256, 0, c->stream>>>(a, c->partial);
    return 0;
  });
  HIP_TRY(hipGetLastError());
  std::vector<double> h(5 * (size_t)c->n_partial);
  HIP_TRY(hipMemcpyAsync(h.data(), c->partial, sizeof(double) * h.size(),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int f = 0; f < 5; f++) {
    double s = 0.0;
    for (int b = 0; b < c->n_partial; b++) s += h[(size_t)b * 5 + f];
    out[f] = s;
  }
  return GX_OK;
  GX_API_END("gx_totals")
}

}  // extern "C"
