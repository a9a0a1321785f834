// gx_batch.hip -- many independent charts of uneven sizes advanced together
// (the gx_batch_* entry points of include/gx_fluid.h).
//
// The states of all problems live in one allocation per buffer (problem p's
// field f at base + 5*off_p + f*ncell_p).  Every RK stage of every problem is
// one persistent launch: the resident blocks take tiles of all problems from
// a per-XCD atomic counter (tiles of one XCD are a contiguous run of the tile
// list), and each tile runs the same stage tile as the single-chart kernel,
// so every problem gets the bits the single-chart path gives it.  The time
// step of each problem is computed on the device from its own CFL maximum
// (Integrator_C_PS % ComputeCycle_ASC: ramp and finish-time clip), so a call
// of nSteps cycles is a fixed kernel sequence captured in one HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/gx_fluid.h"
#include "gx_host.hpp"
#include "gx_kernels.hpp"
#include "gx_tiles.hpp"

namespace gx {

// Tile counters per launch, NXCD each: stage 1 of each tile shape, stage 2
// of each tile shape, the flags.
constexpr int N_COUNTERS = 2 * N_TILE_SHAPES + 1;
__host__ __device__ constexpr int counter_set(int stage, int shape) {
  return stage == 3 ? 2 * N_TILE_SHAPES : (stage - 1) * N_TILE_SHAPES + shape;
}

// One problem of a batch, as the device reads it.
struct BatchProblem {
  int n[3];
  int bc[6];
  int nbx, nby, tz;    // stage tiles; tz = z-chunk of the problem's size class
  int shape;           // stage-tile shape of the size class (tile_x, tile_y)
  int fnbx, fnby;      // shock-flag tiles (16 x 16)
  int tile_start;      // first stage tile of the problem in the tile list
  int flag_tile_start; // first shock-flag tile
  int gs;              // geometry stride
  long long off;       // first cell of the problem in the state buffers
  long long ncell;
  long long geo_off;   // first value of the problem's geometry
};

template <typename R>
struct BatchArgs {
  const BatchProblem *prob;
  const int *tile_prob;   // problem of each tile of this launch
  int ntile;
  Physics ph;
  const R *src;           // stage input (stage 1: U, or P0 where src_prim)
  const R *prim;          // uploaded primitives
  const R *aux;
  R *out1, *out2;
  const R *geo;
  uint8_t *flags;
  unsigned long long *invdt;  // per problem
  const double *dt;           // per problem
  const int *active;          // per problem: stepped in this cycle
  const int *src_prim;        // per problem: stage 1 reads the uploaded primitives
  int *counters;              // NXCD tile counters of this launch
};

// Tile i of this block's XCD run, from the XCD's counter; the value is
// broadcast through LDS and is uniform over the block.
__device__ __forceinline__ int next_tile(int *counter, int *sNext) {
  __syncthreads();  // the previous tile's LDS reads (and sNext) are done
  if (threadIdx.x == 0) *sNext = atomicAdd(counter, 1);
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*sNext);
}

// Stage arguments of problem p (dt is left 0: the device reads it).
template <typename R>
__host__ __device__ inline StageArgs<R> problem_args(const BatchArgs<R> &B,
                                                     const BatchProblem &P, int p,
                                                     bool use_prim) {
  StageArgs<R> a;
  for (int d = 0; d < 3; d++) a.n[d] = P.n[d];
  for (int i = 0; i < 6; i++) a.bc[i] = P.bc[i];
  a.tz = P.tz;
  a.src_is_prim = use_prim ? 1 : 0;
  a.gs = P.gs;
  a.nbx = P.nbx;
  a.nby = P.nby;
  a.nblk = 0;
  a.fs = P.ncell;
  a.dt = 0.0;
  a.ph = B.ph;
  const long long o = 5 * P.off;
  a.src = (use_prim ? B.prim : B.src) + o;
  a.aux = B.aux ? B.aux + o : nullptr;
  a.out1 = B.out1 ? B.out1 + o : nullptr;
  a.out2 = B.out2 ? B.out2 + o : nullptr;  // stage 2 writes the pre-fix state only if set
  a.geo = B.geo + P.geo_off;
  a.flags = B.flags + P.off;
  a.invdt_bits = B.invdt + p;
  a.sources = nullptr;  // batches record no divergence sources
  a.rdt = 0.0;
  return a;
}

// One launch per stage-tile shape in use: its tile list holds the tiles of
// the size classes of that shape.
template <typename R, int STAGE, typename T>
__global__ __launch_bounds__(NT, StageOccupancy<R>::waves) void batch_stage_kernel(
    BatchArgs<R> B) {
  __shared__ StageShared<R, T> L;
  __shared__ int sNext;
  const int xcd = blockIdx.x % NXCD;
  const int q = B.ntile / NXCD, r = B.ntile % NXCD;
  const int begin = xcd * q + min(xcd, r), count = q + (xcd < r ? 1 : 0);
  for (;;) {
    const int i = next_tile(B.counters + xcd, &sNext);
    if (i >= count) break;
    const int t = begin + i;
    const int p = B.tile_prob[t];
    if (!B.active[p]) continue;
    const BatchProblem &P = B.prob[p];
    StageArgs<R> a = problem_args(B, P, p, STAGE == 1 && B.src_prim[p]);
    a.dt = B.dt[p];
    const int lt = t - P.tile_start;
    stage_tile<R, STAGE, RECON_LINEAR, false, false, T>(a, (lt % P.nbx) * T::X,
                                                        ((lt / P.nbx) % P.nby) * T::Y,
                                                        (lt / (P.nbx * P.nby)) * P.tz, L);
  }
}

template <typename R>
__global__ __launch_bounds__(NT) void batch_flags_kernel(BatchArgs<R> B) {
  __shared__ FlagShared<R> L;
  __shared__ int sNext;
  const int xcd = blockIdx.x % NXCD;
  const int q = B.ntile / NXCD, r = B.ntile % NXCD;
  const int begin = xcd * q + min(xcd, r), count = q + (xcd < r ? 1 : 0);
  for (;;) {
    const int i = next_tile(B.counters + xcd, &sNext);
    if (i >= count) break;
    const int t = begin + i;
    const int p = B.tile_prob[t];
    if (!B.active[p]) continue;
    const BatchProblem &P = B.prob[p];
    StageArgs<R> a = problem_args(B, P, p, false);
    a.tz = FLAG_TZ;
    const int lt = t - P.flag_tile_start;
    shock_flags_tile(a, (lt % P.fnbx) * TX, ((lt / P.fnbx) % P.fnby) * TY,
                     (lt / (P.fnbx * P.fnby)) * FLAG_TZ, B.flags + P.off, L);
  }
}

// Integrator state of every problem on the device.
struct BatchClock {
  double *time;
  double *dt;
  int *cycle;
  int *active;
  int *src_prim;      // stage 1 of this cycle reads the uploaded primitives
  int *prim_pending;  // set by an upload, moved to src_prim by the next cycle
  unsigned long long *invdt;
  int *counters;      // N_COUNTERS * NXCD: stage 1 and 2 per shape, flags
  int ncounters;
};

// Start of one cycle of Integrator_C_PS % ComputeCycle_ASC for every problem
// (Integrator_Template.f90:842-939 with Integrator_C_PS__Form.f90:328-423):
// the candidate CourantFactor * min(huge, 1/max(tiny, invdt)), the ramp
// min((iCycle+1)/nRampCycles, 1), the clip onto FinishTime (the checkpoint
// of a one-interval run), then dt = (Time + dt) - Time.  A problem at or past
// FinishTime, or at FinishCycle, is not stepped.  Also clears the CFL
// maxima that stage 2 accumulates and the tile counters.
__global__ __launch_bounds__(256) void batch_clock_kernel(int nprob, BatchClock c,
                                                          double courant, double finish,
                                                          int finish_cycle, int n_ramp) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid < c.ncounters) c.counters[tid] = 0;
  for (int p = tid; p < nprob; p += gridDim.x * blockDim.x) {
    c.src_prim[p] = c.prim_pending[p];
    c.prim_pending[p] = 0;
    const double t = c.time[p];
    const int cyc = c.cycle[p];
    if (!(t < finish) || cyc >= finish_cycle) {
      c.active[p] = 0;  // dt keeps the last step taken
      continue;
    }
    const double invdt = __longlong_as_double((long long)c.invdt[p]);
    const double candidate = courant * fmin(DBL_MAX, 1.0 / fmax(DBL_MIN, invdt));
    double ts = fmin(DBL_MAX, candidate);
    const double ramp = fmin((double)(cyc + 1) / (double)n_ramp, 1.0);
    if (ramp < 1.0) ts = ramp * ts;
    if (t + ts > finish) ts = finish - t;
    const double t_new = t + ts;
    const double dt = t_new - t;
    c.dt[p] = dt;
    c.time[p] = t + dt;
    c.cycle[p] = cyc + 1;
    c.active[p] = 1;
    c.invdt[p] = 0ull;
  }
}

}  // namespace gx

using namespace gx;

struct gx_batch {
  int np = 0;
  gx_params params;
  bool fp32 = false;
  size_t esz = sizeof(double);
  hipStream_t stream = nullptr;
  std::vector<gx_grid> grids;
  std::vector<BatchProblem> prob;  // host copy of the descriptors
  std::vector<int> size_class;     // per problem
  std::map<int, int> class_tz;     // z-chunk per size class
  std::map<int, int> class_shape;  // stage-tile shape per size class
  long long total_cells = 0;
  // device buffers
  void *U = nullptr, *Y = nullptr, *W = nullptr, *P0 = nullptr, *geo = nullptr;
  uint8_t *flags = nullptr;
  BatchProblem *prob_d = nullptr;
  // stage tiles of each shape (problem of each tile), and the flag tiles
  int *tile_prob_d[N_TILE_SHAPES] = {}, *flag_tile_prob_d = nullptr;
  int ntile[N_TILE_SHAPES] = {}, tile_cap[N_TILE_SHAPES] = {};
  int nflag_tile = 0, flag_tile_cap = 0;
  double *time = nullptr, *dt = nullptr;
  int *ints = nullptr;  // cycle, active, src_prim, prim_pending (np each), counters
  unsigned long long *invdt = nullptr;
  // host bookkeeping
  std::vector<char> uploaded, cfl_valid;
  double finish_time = 1.0;
  int finish_cycle = 2147483647;
  int n_ramp = 100;
  int use_graph = 1;
  int stage_blocks[N_TILE_SHAPES] = {}, flag_blocks = 0;  // persistent grid sizes
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;

  int *cycle() const { return ints; }
  int *active() const { return ints + np; }
  int *src_prim() const { return ints + 2 * np; }
  int *prim_pending() const { return ints + 3 * np; }
  int *counters() const { return ints + 4 * np; }
};

static void drop_graph(gx_batch *b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_steps = 0;
}

// Size class of a problem: half-octaves of its largest edge.
static int size_class_of(const gx_grid &g) {
  const int m = std::max(g.nCells[0], std::max(g.nCells[1], g.nCells[2]));
  return (int)std::floor(2.0 * std::log2((double)m));
}

// Rebuilds the stage and flag tile lists from the per-class z-chunks; a
// problem is listed when keep[p] (all when keep is empty).
static int build_tiles(gx_batch *b, const std::vector<char> &keep) {
  std::vector<int> tiles[N_TILE_SHAPES], ftiles;
  for (int p = 0; p < b->np; p++) {
    BatchProblem &P = b->prob[p];
    auto it = b->class_tz.find(b->size_class[p]);
    P.tz = std::max(1, std::min(it == b->class_tz.end() ? 16 : it->second, P.n[2]));
    auto is = b->class_shape.find(b->size_class[p]);
    P.shape = is == b->class_shape.end() ? 0 : is->second;
    P.nbx = (P.n[0] + tile_x(P.shape) - 1) / tile_x(P.shape);
    P.nby = (P.n[1] + tile_y(P.shape) - 1) / tile_y(P.shape);
    P.fnbx = (P.n[0] + TX - 1) / TX;
    P.fnby = (P.n[1] + TY - 1) / TY;
    const int nbz = (P.n[2] + P.tz - 1) / P.tz;
    const int nfz = (P.n[2] + FLAG_TZ - 1) / FLAG_TZ;
    std::vector<int> &tl = tiles[P.shape];
    P.tile_start = (int)tl.size();
    P.flag_tile_start = (int)ftiles.size();
    if (!keep.empty() && !keep[p]) continue;
    tl.insert(tl.end(), (size_t)P.nbx * P.nby * nbz, p);
    ftiles.insert(ftiles.end(), (size_t)P.fnbx * P.fnby * nfz, p);
  }
  for (int sh = 0; sh < N_TILE_SHAPES; sh++) {
    if ((int)tiles[sh].size() > b->tile_cap[sh]) {
      (void)hipFree(b->tile_prob_d[sh]);
      b->tile_prob_d[sh] = nullptr;
      HIP_TRY(hipMalloc(&b->tile_prob_d[sh], sizeof(int) * tiles[sh].size()));
      b->tile_cap[sh] = (int)tiles[sh].size();
    }
    b->ntile[sh] = (int)tiles[sh].size();
    if (!tiles[sh].empty())
      HIP_TRY(hipMemcpy(b->tile_prob_d[sh], tiles[sh].data(), sizeof(int) * tiles[sh].size(),
                        hipMemcpyHostToDevice));
  }
  if ((int)ftiles.size() > b->flag_tile_cap) {
    (void)hipFree(b->flag_tile_prob_d);
    b->flag_tile_prob_d = nullptr;
    HIP_TRY(hipMalloc(&b->flag_tile_prob_d, sizeof(int) * ftiles.size()));
    b->flag_tile_cap = (int)ftiles.size();
  }
  b->nflag_tile = (int)ftiles.size();
  if (!ftiles.empty())
    HIP_TRY(hipMemcpy(b->flag_tile_prob_d, ftiles.data(), sizeof(int) * ftiles.size(),
                      hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->prob_d, b->prob.data(), sizeof(BatchProblem) * b->np,
                    hipMemcpyHostToDevice));
  drop_graph(b);
  return GX_OK;
}

template <typename R>
static BatchArgs<R> batch_args(const gx_batch *b, int stage, int shape = 0) {
  BatchArgs<R> B;
  memset(&B, 0, sizeof(B));
  B.prob = b->prob_d;
  B.ph.gamma = b->params.AdiabaticIndex;
  B.ph.theta = b->params.LimiterParameter;
  B.ph.mass = b->params.BaryonMassReference;
  B.ph.nmin = b->params.BaryonDensityMin;
  B.ph.st = b->params.ShockThreshold;
  B.ph.use_limiter = b->params.UseLimiter;
  B.ph.hllc = b->params.RiemannSolverType == GX_HLLC;
  B.prim = (const R *)b->P0;
  B.geo = (const R *)b->geo;
  B.flags = b->flags;
  B.invdt = b->invdt;
  B.dt = b->dt;
  B.active = b->active();
  B.src_prim = b->src_prim();
  B.counters = b->counters() + counter_set(stage, shape) * NXCD;
  if (stage == 1) {
    B.tile_prob = b->tile_prob_d[shape];
    B.ntile = b->ntile[shape];
    B.src = (const R *)b->U;
    B.aux = (const R *)b->U;
    B.out1 = (R *)b->Y;
    B.out2 = (R *)b->W;
  } else if (stage == 2) {
    B.tile_prob = b->tile_prob_d[shape];
    B.ntile = b->ntile[shape];
    B.src = (const R *)b->Y;
    B.aux = (const R *)b->W;
    B.out1 = (R *)b->U;
  } else {
    B.tile_prob = b->flag_tile_prob_d;
    B.ntile = b->nflag_tile;
    B.src = (const R *)b->U;
  }
  return B;
}

// The kernels of one cycle, in stream order.
static int enqueue_cycle(gx_batch *b) {
  BatchClock clk;
  clk.time = b->time;
  clk.dt = b->dt;
  clk.cycle = b->cycle();
  clk.active = b->active();
  clk.src_prim = b->src_prim();
  clk.prim_pending = b->prim_pending();
  clk.invdt = b->invdt;
  clk.counters = b->counters();
  clk.ncounters = N_COUNTERS * NXCD;
  const int cb = std::max(1, (std::max(b->np, clk.ncounters) + 255) / 256);
  batch_clock_kernel<<<cb, 256, 0, b->stream>>>(b->np, clk, b->params.CourantFactor,
                                                 b->finish_time, b->finish_cycle, b->n_ramp);
  HIP_TRY(hipGetLastError());
  return by_precision(b->fp32, [&](auto tag) {
    using R = decltype(tag);
    for (int stage = 1; stage <= 2; stage++) {
      for (int sh = 0; sh < N_TILE_SHAPES; sh++) {
        if (b->ntile[sh] == 0) continue;
        const BatchArgs<R> B = batch_args<R>(b, stage, sh);
        if (stage == 1 && sh == 0)
          batch_stage_kernel<R, 1, Tile16x16><<<b->stage_blocks[0], NT, 0, b->stream>>>(B);
        else if (stage == 1)
          batch_stage_kernel<R, 1, Tile32x8><<<b->stage_blocks[1], NT, 0, b->stream>>>(B);
        else if (sh == 0)
          batch_stage_kernel<R, 2, Tile16x16><<<b->stage_blocks[0], NT, 0, b->stream>>>(B);
        else
          batch_stage_kernel<R, 2, Tile32x8><<<b->stage_blocks[1], NT, 0, b->stream>>>(B);
        HIP_TRY(hipGetLastError());
      }
    }
    batch_flags_kernel<R><<<b->flag_blocks, NT, 0, b->stream>>>(batch_args<R>(b, 3));
    HIP_TRY(hipGetLastError());
    return GX_OK;
  });
}

// CFL maxima of problems whose state changed on the host side.
static int refresh_cfl(gx_batch *b) {
  for (int p = 0; p < b->np; p++) {
    if (b->cfl_valid[p]) continue;
    const BatchProblem &P = b->prob[p];
    HIP_TRY(hipMemsetAsync(b->invdt + p, 0, sizeof(unsigned long long), b->stream));
    const int rc = by_precision(b->fp32, [&](auto tag) {
      using R = decltype(tag);
      BatchArgs<R> B = batch_args<R>(b, 1);
      StageArgs<R> a = problem_args(B, P, p, true);
      cfl_kernel<R><<<grid_blocks(P.ncell), 256, 0, b->stream>>>(a);
      HIP_TRY(hipGetLastError());
      return GX_OK;
    });
    if (rc) return rc;
    b->cfl_valid[p] = 1;
  }
  return GX_OK;
}

static int check_problem(gx_batch *b, int p, const char *who) {
  if (!b) return fail(GX_E_ARG, std::string(who) + ": null batch");
  if (p < 0 || p >= b->np) return fail(GX_E_ARG, std::string(who) + ": problem index out of range");
  return GX_OK;
}

extern "C" {

int gx_batch_destroy(gx_batch *b) {
  GX_API_BEGIN
  if (!b) return GX_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (void *p : {(void *)b->U, (void *)b->Y, (void *)b->W, (void *)b->P0, b->geo,
                  (void *)b->flags, (void *)b->prob_d, (void *)b->tile_prob_d[0],
                  (void *)b->tile_prob_d[1], (void *)b->flag_tile_prob_d, (void *)b->time, (void *)b->dt,
                  (void *)b->ints, (void *)b->invdt})
    (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return GX_OK;
  GX_API_END("gx_batch_destroy")
}

int gx_batch_create(int nProblems, const gx_grid *grids, const gx_params *params,
                    gx_batch **out) {
  GX_API_BEGIN
  if (!out || !grids || nProblems < 1)
    return fail(GX_E_ARG, "gx_batch_create: bad argument");
  *out = nullptr;
  gx_params p;
  if (params)
    p = *params;
  else
    gx_default_params(&p);
  int rc = check_params(p, "gx_batch_create");
  if (rc) return rc;
  if (p.ReconstructionType != GX_LINEAR)
    return fail(GX_E_ARG, "gx_batch_create: batches take ReconstructionType GX_LINEAR");
  std::vector<ChartGeometry> geoms(nProblems);
  for (int i = 0; i < nProblems; i++) {
    rc = check_grid(grids[i], "gx_batch_create");
    if (rc) return rc;
    rc = make_geometry(grids[i], geoms[i], "gx_batch_create");
    if (rc) return rc;
  }
  gx_batch *b = new (std::nothrow) gx_batch();
  if (!b) return fail(GX_E_NOMEM, "gx_batch_create: host allocation failed");
  b->np = nProblems;
  b->params = p;
  b->fp32 = p.Precision == GX_FP32;
  b->esz = b->fp32 ? sizeof(float) : sizeof(double);
  b->grids.assign(grids, grids + nProblems);
  b->prob.resize(nProblems);
  b->size_class.resize(nProblems);
  b->uploaded.assign(nProblems, 0);
  b->cfl_valid.assign(nProblems, 0);
#define TRY_CLEAN(expr)                                                   \
  do {                                                                    \
    hipError_t e_ = (expr);                                               \
    if (e_ != hipSuccess) {                                               \
      std::string m_ = std::string(#expr ": ") + hipGetErrorString(e_);   \
      gx_batch_destroy(b);                                                \
      return fail(GX_E_HIP, m_);                                          \
    }                                                                     \
  } while (0)
  TRY_CLEAN(hipStreamCreate(&b->stream));
  std::vector<char> geo_img;
  long long off = 0, goff = 0;
  for (int i = 0; i < nProblems; i++) {
    BatchProblem &P = b->prob[i];
    memset(&P, 0, sizeof(P));
    P.ncell = 1;
    for (int d = 0; d < 3; d++) {
      P.n[d] = grids[i].nCells[d];
      P.ncell *= P.n[d];
    }
    for (int k = 0; k < 6; k++) P.bc[k] = grids[i].BoundaryCondition[k];
    P.gs = geoms[i].gs;
    P.off = off;
    P.geo_off = goff;
    off += P.ncell;
    const std::vector<char> img = geometry_image(geoms[i], b->fp32);
    geo_img.insert(geo_img.end(), img.begin(), img.end());
    goff += (long long)(img.size() / b->esz);
    b->size_class[i] = size_class_of(grids[i]);
    b->class_tz.emplace(b->size_class[i], 16);
    b->class_shape.emplace(b->size_class[i], 0);
  }
  b->total_cells = off;
  const size_t bytes = 5 * b->esz * (size_t)off;
  TRY_CLEAN(hipMalloc(&b->U, bytes));
  TRY_CLEAN(hipMalloc(&b->Y, bytes));
  TRY_CLEAN(hipMalloc(&b->W, bytes));
  TRY_CLEAN(hipMalloc(&b->P0, bytes));
  TRY_CLEAN(hipMalloc(&b->geo, geo_img.size()));
  TRY_CLEAN(hipMemcpy(b->geo, geo_img.data(), geo_img.size(), hipMemcpyHostToDevice));
  TRY_CLEAN(hipMalloc(&b->flags, (size_t)off));
  TRY_CLEAN(hipMemset(b->flags, 0, (size_t)off));
  TRY_CLEAN(hipMalloc(&b->prob_d, sizeof(BatchProblem) * nProblems));
  TRY_CLEAN(hipMalloc(&b->time, sizeof(double) * nProblems));
  TRY_CLEAN(hipMalloc(&b->dt, sizeof(double) * nProblems));
  TRY_CLEAN(hipMemset(b->time, 0, sizeof(double) * nProblems));
  TRY_CLEAN(hipMemset(b->dt, 0, sizeof(double) * nProblems));
  const size_t nints = 4 * (size_t)nProblems + N_COUNTERS * NXCD;
  TRY_CLEAN(hipMalloc(&b->ints, sizeof(int) * nints));
  TRY_CLEAN(hipMemset(b->ints, 0, sizeof(int) * nints));
  TRY_CLEAN(hipMalloc(&b->invdt, sizeof(unsigned long long) * nProblems));
  TRY_CLEAN(hipMemset(b->invdt, 0, sizeof(unsigned long long) * nProblems));
  // Persistent grids: every block resident at once (per shape, the smaller
  // occupancy of its two stage kernels).
  int dev = 0, ncu = 0, per_cu_f = 0;
  TRY_CLEAN(hipGetDevice(&dev));
  TRY_CLEAN(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const void *kern[N_TILE_SHAPES][2];
  if (b->fp32) {
    kern[0][0] = (const void *)batch_stage_kernel<float, 1, Tile16x16>;
    kern[0][1] = (const void *)batch_stage_kernel<float, 2, Tile16x16>;
    kern[1][0] = (const void *)batch_stage_kernel<float, 1, Tile32x8>;
    kern[1][1] = (const void *)batch_stage_kernel<float, 2, Tile32x8>;
    TRY_CLEAN(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu_f, (const void *)batch_flags_kernel<float>, NT, 0));
  } else {
    kern[0][0] = (const void *)batch_stage_kernel<double, 1, Tile16x16>;
    kern[0][1] = (const void *)batch_stage_kernel<double, 2, Tile16x16>;
    kern[1][0] = (const void *)batch_stage_kernel<double, 1, Tile32x8>;
    kern[1][1] = (const void *)batch_stage_kernel<double, 2, Tile32x8>;
    TRY_CLEAN(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu_f, (const void *)batch_flags_kernel<double>, NT, 0));
  }
  for (int sh = 0; sh < N_TILE_SHAPES; sh++) {
    int o1 = 0, o2 = 0;
    TRY_CLEAN(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o1, kern[sh][0], NT, 0));
    TRY_CLEAN(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o2, kern[sh][1], NT, 0));
    b->stage_blocks[sh] = std::max(NXCD, ncu * std::max(1, std::min(o1, o2)));
  }
  b->flag_blocks = std::max(NXCD, ncu * std::max(1, per_cu_f));
#undef TRY_CLEAN
  rc = build_tiles(b, {});
  if (rc) {
    gx_batch_destroy(b);
    return rc;
  }
  *out = b;
  return GX_OK;
  GX_API_END("gx_batch_create")
}

int gx_batch_set_primitive(gx_batch *b, int iProblem, const double *primitive) {
  GX_API_BEGIN
  int rc = check_problem(b, iProblem, "gx_batch_set_primitive");
  if (rc) return rc;
  if (!primitive) return fail(GX_E_ARG, "gx_batch_set_primitive: null argument");
  const BatchProblem &P = b->prob[iProblem];
  const size_t o = 5 * (size_t)P.off * b->esz;
  rc = put_values(b->stream, b->fp32, (char *)b->P0 + o, primitive, 5 * (size_t)P.ncell);
  if (rc) return rc;
  by_precision(b->fp32, [&](auto tag) {
    using R = decltype(tag);
    from_primitive_kernel<R><<<grid_blocks(P.ncell), 256, 0, b->stream>>>(
        P.ncell, P.ncell, b->params.BaryonMassReference, (R *)((char *)b->P0 + o),
        (R *)((char *)b->U + o));
    return 0;
  });
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemsetAsync(b->flags + P.off, 0, (size_t)P.ncell, b->stream));
  const double zero = 0.0;
  const int one = 1, none = 0;
  HIP_TRY(hipMemcpyAsync(b->time + iProblem, &zero, sizeof(double), hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipMemcpyAsync(b->cycle() + iProblem, &none, sizeof(int), hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipMemcpyAsync(b->prim_pending() + iProblem, &one, sizeof(int),
                         hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->uploaded[iProblem] = 1;
  b->cfl_valid[iProblem] = 0;
  return GX_OK;
  GX_API_END("gx_batch_set_primitive")
}

int gx_batch_set_integrator(gx_batch *b, double FinishTime, int FinishCycle, int nRampCycles) {
  GX_API_BEGIN
  if (!b || nRampCycles < 1 || FinishCycle < 0)
    return fail(GX_E_ARG, "gx_batch_set_integrator: bad argument");
  b->finish_time = FinishTime;
  b->finish_cycle = FinishCycle;
  b->n_ramp = nRampCycles;
  drop_graph(b);
  return GX_OK;
  GX_API_END("gx_batch_set_integrator")
}

int gx_batch_step(gx_batch *b, int nSteps) {
  GX_API_BEGIN
  if (!b || nSteps < 0) return fail(GX_E_ARG, "gx_batch_step: bad argument");
  for (int p = 0; p < b->np; p++)
    if (!b->uploaded[p]) return fail(GX_E_STATE, "gx_batch_step: a problem has no state");
  int rc = refresh_cfl(b);
  if (rc) return rc;
  if (nSteps == 0) return GX_OK;
  if (!b->use_graph) {
    for (int s = 0; s < nSteps; s++) {
      rc = enqueue_cycle(b);
      if (rc) return rc;
    }
    return GX_OK;
  }
  if (!b->graph || b->graph_steps != nSteps) {
    drop_graph(b);
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < nSteps && rc == GX_OK; s++) rc = enqueue_cycle(b);
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess)
      return fail(GX_E_HIP, std::string("gx_batch_step: capture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      b->graph = nullptr;
      return fail(GX_E_HIP, std::string("gx_batch_step: instantiate: ") + hipGetErrorString(ei));
    }
    b->graph_steps = nSteps;
  }
  HIP_TRY(hipGraphLaunch(b->graph, b->stream));
  return GX_OK;
  GX_API_END("gx_batch_step")
}

int gx_batch_sync(gx_batch *b) {
  GX_API_BEGIN
  if (!b) return fail(GX_E_ARG, "gx_batch_sync: null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return GX_OK;
  GX_API_END("gx_batch_sync")
}

int gx_batch_cfl(gx_batch *b, double *candidates) {
  GX_API_BEGIN
  if (!b || !candidates) return fail(GX_E_ARG, "gx_batch_cfl: null argument");
  for (int p = 0; p < b->np; p++)
    if (!b->uploaded[p]) return fail(GX_E_STATE, "gx_batch_cfl: a problem has no state");
  int rc = refresh_cfl(b);
  if (rc) return rc;
  std::vector<unsigned long long> bits(b->np);
  HIP_TRY(hipMemcpyAsync(bits.data(), b->invdt, sizeof(unsigned long long) * b->np,
                         hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (int p = 0; p < b->np; p++) {
    double invdt;
    memcpy(&invdt, &bits[p], sizeof(invdt));
    candidates[p] = b->params.CourantFactor * fmin(DBL_MAX, 1.0 / fmax(DBL_MIN, invdt));
  }
  return GX_OK;
  GX_API_END("gx_batch_cfl")
}

int gx_batch_times(gx_batch *b, double *time, double *lastTimeStep, int *cycle) {
  GX_API_BEGIN
  if (!b || !time) return fail(GX_E_ARG, "gx_batch_times: null argument");
  HIP_TRY(hipMemcpyAsync(time, b->time, sizeof(double) * b->np, hipMemcpyDeviceToHost,
                         b->stream));
  if (lastTimeStep)
    HIP_TRY(hipMemcpyAsync(lastTimeStep, b->dt, sizeof(double) * b->np,
                           hipMemcpyDeviceToHost, b->stream));
  if (cycle)
    HIP_TRY(hipMemcpyAsync(cycle, b->cycle(), sizeof(int) * b->np, hipMemcpyDeviceToHost,
                           b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return GX_OK;
  GX_API_END("gx_batch_times")
}

int gx_batch_get_conserved(gx_batch *b, int iProblem, double *conserved) {
  GX_API_BEGIN
  int rc = check_problem(b, iProblem, "gx_batch_get_conserved");
  if (rc) return rc;
  if (!conserved) return fail(GX_E_ARG, "gx_batch_get_conserved: null argument");
  const BatchProblem &P = b->prob[iProblem];
  return get_values(b->stream, b->fp32, conserved,
                    (const char *)b->U + 5 * (size_t)P.off * b->esz, 5 * (size_t)P.ncell);
  GX_API_END("gx_batch_get_conserved")
}

int gx_batch_get_flags(gx_batch *b, int iProblem, unsigned char *flags) {
  GX_API_BEGIN
  int rc = check_problem(b, iProblem, "gx_batch_get_flags");
  if (rc) return rc;
  if (!flags) return fail(GX_E_ARG, "gx_batch_get_flags: null argument");
  const BatchProblem &P = b->prob[iProblem];
  HIP_TRY(hipMemcpyAsync(flags, b->flags + P.off, (size_t)P.ncell, hipMemcpyDeviceToHost,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return GX_OK;
  GX_API_END("gx_batch_get_flags")
}

int gx_batch_set_tuning(gx_batch *b, int nClasses, const int *sizeClass, const int *zChunk,
                        int useGraph) {
  GX_API_BEGIN
  if (!b || nClasses < 0 || (nClasses > 0 && (!sizeClass || !zChunk)))
    return fail(GX_E_ARG, "gx_batch_set_tuning: bad argument");
  for (int i = 0; i < nClasses; i++) {
    if (zChunk[i] < 1) return fail(GX_E_ARG, "gx_batch_set_tuning: zChunk must be >= 1");
    b->class_tz[sizeClass[i]] = zChunk[i];
  }
  b->use_graph = useGraph != 0;
  return build_tiles(b, {});
  GX_API_END("gx_batch_set_tuning")
}

int gx_batch_set_tiles(gx_batch *b, int nClasses, const int *sizeClass, const int *tileX,
                       const int *tileY, const int *zChunk, int useGraph) {
  GX_API_BEGIN
  if (!b || nClasses < 0 || (nClasses > 0 && (!sizeClass || !tileX || !tileY || !zChunk)))
    return fail(GX_E_ARG, "gx_batch_set_tiles: bad argument");
  std::vector<int> shapes(nClasses, -1);
  for (int i = 0; i < nClasses; i++) {
    if (zChunk[i] < 1) return fail(GX_E_ARG, "gx_batch_set_tiles: zChunk must be >= 1");
    for (int k = 0; k < N_TILE_SHAPES; k++)
      if (tileX[i] == tile_x(k) && tileY[i] == tile_y(k)) shapes[i] = k;
    if (shapes[i] < 0) return fail(GX_E_ARG, "gx_batch_set_tiles: the stage tile is 16 x 16 or 32 x 8");
  }
  for (int i = 0; i < nClasses; i++) {
    b->class_tz[sizeClass[i]] = zChunk[i];
    b->class_shape[sizeClass[i]] = shapes[i];
  }
  b->use_graph = useGraph != 0;
  return build_tiles(b, {});
  GX_API_END("gx_batch_set_tiles")
}

int gx_batch_get_tiles(gx_batch *b, int capacity, int *sizeClass, int *tileX, int *tileY,
                       int *zChunk, int *nClasses) {
  GX_API_BEGIN
  if (!b || !nClasses) return fail(GX_E_ARG, "gx_batch_get_tiles: null argument");
  int i = 0;
  for (const auto &kv : b->class_tz) {
    if (i < capacity && sizeClass && tileX && tileY && zChunk) {
      const int sh = b->class_shape[kv.first];
      sizeClass[i] = kv.first;
      tileX[i] = tile_x(sh);
      tileY[i] = tile_y(sh);
      zChunk[i] = kv.second;
    }
    i++;
  }
  *nClasses = i;
  return GX_OK;
  GX_API_END("gx_batch_get_tiles")
}

int gx_batch_get_tuning(gx_batch *b, int capacity, int *sizeClass, int *zChunk,
                        int *nClasses) {
  GX_API_BEGIN
  if (!b || !nClasses) return fail(GX_E_ARG, "gx_batch_get_tuning: null argument");
  int i = 0;
  for (const auto &kv : b->class_tz) {
    if (i < capacity && sizeClass && zChunk) {
      sizeClass[i] = kv.first;
      zChunk[i] = kv.second;
    }
    i++;
  }
  *nClasses = i;
  return GX_OK;
  GX_API_END("gx_batch_get_tuning")
}

int gx_batch_autotune(gx_batch *b, int nRepeat) {
  GX_API_BEGIN
  if (!b || nRepeat < 1) return fail(GX_E_ARG, "gx_batch_autotune: bad argument");
  for (int p = 0; p < b->np; p++)
    if (!b->uploaded[p]) return fail(GX_E_STATE, "gx_batch_autotune: a problem has no state");
  // Stage 1 writes only Y and W, so timing it leaves the state untouched.
  // Every problem is made active with its current dt; the flags stay.  On
  // any failure the z-chunks in force before the call and the full tile
  // list are restored before returning.
  const std::map<int, int> before = b->class_tz, before_shape = b->class_shape;
  std::vector<int> ones(b->np, 1);
  HIP_TRY(hipMemcpy(b->active(), ones.data(), sizeof(int) * b->np, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(b->src_prim(), 0, sizeof(int) * b->np));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GX_OK;
  auto hip_ok = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && rc == GX_OK)
      rc = fail(GX_E_HIP, std::string("gx_batch_autotune: ") + what + ": " + hipGetErrorString(e));
    return rc == GX_OK;
  };
  hip_ok(hipEventCreate(&e0), "hipEventCreate");
  hip_ok(hipEventCreate(&e1), "hipEventCreate");
  std::map<int, int> best, best_shape;
  for (const auto &kv : before) {
    if (rc) break;
    const int cls = kv.first;
    std::vector<char> keep(b->np, 0);
    int nmax = 1;
    for (int p = 0; p < b->np; p++)
      if (b->size_class[p] == cls) {
        keep[p] = 1;
        nmax = std::max(nmax, b->prob[p].n[2]);
      }
    float best_ms = 1e30f;
    int best_tz = kv.second, best_sh = before_shape.at(cls);
    for (int trial = 0; trial < 5 * N_TILE_SHAPES; trial++) {
      const int sh = trial / 5, tz = 4 << (trial % 5);  // z-chunks 4 .. 64
      if (rc) break;
      if (tz > nmax && tz != 4) continue;
      b->class_tz[cls] = tz;
      b->class_shape[cls] = sh;
      rc = build_tiles(b, keep);
      float ms_min = 1e30f;
      for (int r = 0; r < nRepeat + 1 && rc == GX_OK; r++) {
        const int cset = counter_set(1, sh);
        if (!hip_ok(hipMemsetAsync(b->counters() + cset * NXCD, 0, sizeof(int) * NXCD,
                                   b->stream),
                    "hipMemsetAsync") ||
            !hip_ok(hipEventRecord(e0, b->stream), "hipEventRecord"))
          break;
        rc = by_precision(b->fp32, [&](auto tag) {
          using R = decltype(tag);
          if (sh == 1)
            batch_stage_kernel<R, 1, Tile32x8><<<b->stage_blocks[1], NT, 0, b->stream>>>(
                batch_args<R>(b, 1, 1));
          else
            batch_stage_kernel<R, 1, Tile16x16><<<b->stage_blocks[0], NT, 0, b->stream>>>(
                batch_args<R>(b, 1, 0));
          HIP_TRY(hipGetLastError());
          return GX_OK;
        });
        float ms = 0.0f;
        if (rc || !hip_ok(hipEventRecord(e1, b->stream), "hipEventRecord") ||
            !hip_ok(hipEventSynchronize(e1), "hipEventSynchronize") ||
            !hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"))
          break;
        if (r > 0) ms_min = std::min(ms_min, ms);  // the first launch warms up
      }
      if (rc == GX_OK && ms_min < best_ms) {
        best_ms = ms_min;
        best_tz = tz;
        best_sh = sh;
      }
    }
    best[cls] = best_tz;
    best_shape[cls] = best_sh;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  b->class_tz = before;
  b->class_shape = before_shape;
  if (rc == GX_OK) {
    for (const auto &kv : best) b->class_tz[kv.first] = kv.second;
    for (const auto &kv : best_shape) b->class_shape[kv.first] = kv.second;
  }
  const int rc2 = build_tiles(b, {});
  return rc ? rc : rc2;
  GX_API_END("gx_batch_autotune")
}

long long gx_batch_cells(gx_batch *b) { return b ? b->total_cells : 0; }

void *gx_batch_stream(gx_batch *b) { return b ? (void *)b->stream : nullptr; }

}  // extern "C"
