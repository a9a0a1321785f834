// gx_bricks.hip -- a chart split into bricks in one GPU's HBM: the single-GPU
// part of the reference's brick decomposition (Chart_SLD with nBricks,
// Modules/Mathematics/Manifolds/Atlases/Charts/SingleLevelCharts/
// Chart_SLD__Form.f90:131-201, 324-554, 557-1007; SetBrick and the brick's
// geometry, ChartBasics/Chart_Template.f90:383-455, Chart_SLD__Form.f90:
// 295-317).
//
// Every brick holds its proper cells and, in each five-field buffer, a halo
// block of two ghost layers for every face it shares with another brick
// (edges and corners included, so the shock tests near a brick edge see the
// diagonal brick).  After each RK stage one halo kernel copies into every
// halo the neighbours' values of the next stage's input -- what
// ExchangeGhostData does between ranks -- and the bricks' stage kernels read
// their ghosts there.  Flags are cleared only at the faces of the chart
// (iaBrick == 1 or nBricks, FluidFeatures_P__Form.f90:333-338), and a second
// small exchange gives each brick the flags of the faces just beyond its
// outer faces.  The geometry of a brick is the slice of the chart's
// (Chart_SLD__Form.f90:312-317), and the CFL maximum is one accumulator for
// all bricks (the MPI MIN of the candidates equals the candidate of the
// maximum, the candidate being a non-increasing function of it), so a
// bricked chart steps bit for bit as the one-brick chart does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/gx_fluid.h"
#include "gx_context.hpp"
#include "gx_host.hpp"
#include "gx_kernels.hpp"
#include "gx_tally.hpp"
#include "gx_tiles.hpp"

namespace gx {

// One brick, as the halo kernels read it.
struct BrickDesc {
  int n[3];         // cells of the brick
  int org[3];       // first chart cell of the brick
  int bc[6];        // face codes: a physical BC, PERIODIC (one brick along a
                    // periodic dimension) or BC_NEIGHBOR
  int hoff[6];      // halo block of face b in each field slab (-1: none)
  int hlen[6];      // values of that block
  int fhoff[3];     // flags beyond the outer face of dimension d (-1: none)
  int chart_face;   // bit b: face b is a face of the chart
  int gs;           // geometry stride of the brick
  long long ncell;  // proper cells
  long long fs;     // field stride: cells + halo values
  long long base;   // first value of the brick in a five-field buffer
  long long fbase;  // first flag of the brick
  long long geo_off;  // geometry image of the brick (values of the state type)
  long long tg_off;   // FP64 tally geometry of the brick
  int part_off;     // first per-block partial of the interior sums
  int npart;        // blocks of the interior sums
  int bpart_off;    // first per-block partial of the boundary kernel
  int nbpart;       // blocks of the boundary kernel
};

// The brick decomposition of a context (gx_ctx::bricks).
struct BrickSet {
  int nb[3] = {1, 1, 1};
  int nbc[3] = {1, 1, 1};  // cells per brick
  int count = 0;
  std::vector<BrickDesc> desc;
  BrickDesc *desc_d = nullptr;
  long long total_fs = 0, total_flags = 0;
  long long nhalo = 0, nfhalo = 0;
  long long *hstart_d = nullptr;  // 6 count + 1 prefix sums of the halo blocks
  long long *fstart_d = nullptr;  // 3 count + 1 prefix sums of the flag halos
  std::vector<long long> fstart;  // host copy
  void *U = nullptr, *Y = nullptr, *W = nullptr, *P0 = nullptr, *geo = nullptr;
  uint8_t *flags = nullptr;
  double *tgeo = nullptr;
  double *cur = nullptr;  // download staging: GX_N_FIELDS x the largest brick
  int periodic = 0;       // bit d: the chart is periodic along d
  // Stage and shock-flag launches over groups of bricks: the z-chunk of the
  // shock-flag tiles, chosen for all bricks together.
  int ftz = 0;  // 0: not chosen yet
};

// The chart-level view the halo kernels need.
struct ChartView {
  int N[3];      // chart cells
  int nb[3];     // bricks per dimension
  int nbc[3];    // cells per brick
  int periodic;  // bit d: periodic along d
};

__device__ inline int find_segment(const long long *start, int nseg, long long h) {
  int lo = 0, hi = nseg;  // start[lo] <= h < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= h)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// The brick owning chart cell g (wrapped into the chart when periodic), and
// its local index; false for a cell beyond a non-periodic chart face.
__device__ inline bool owner_of(const ChartView &cv, const BrickDesc *D, int g[3], int *sb,
                                unsigned *li) {
#pragma unroll
  for (int e = 0; e < 3; e++) {
    if (g[e] < 0 || g[e] >= cv.N[e]) {
      if (!((cv.periodic >> e) & 1)) return false;
      g[e] = (g[e] + cv.N[e]) % cv.N[e];
    }
  }
  const int b0 = g[0] / cv.nbc[0], b1 = g[1] / cv.nbc[1], b2 = g[2] / cv.nbc[2];
  *sb = b0 + cv.nb[0] * (b1 + cv.nb[1] * b2);
  const BrickDesc &S = D[*sb];
  *li = lin(S.n, g[0] - S.org[0], g[1] - S.org[1], g[2] - S.org[2]);
  return true;
}

// Halo exchange of one five-field buffer: every halo value takes its
// neighbour brick's value of the same chart cell; halo values beyond a
// non-periodic chart face (which no kernel reads) are set to 0.
template <typename R>
__global__ __launch_bounds__(256) void brick_halo_kernel(const BrickDesc *D, int count,
                                                         const long long *hstart, long long nhalo,
                                                         ChartView cv, R *buf) {
  for (long long h = blockIdx.x * 256ll + threadIdx.x; h < nhalo;
       h += (long long)gridDim.x * 256) {
    const int q = find_segment(hstart, 6 * count, h);
    const int b = q / 6, face = q % 6;
    const BrickDesc &B = D[b];
    const long long r = h - hstart[q];
    const int d = face >> 1;
    const int j = d == 0 ? 1 : 0, k = d == 2 ? 1 : 2;
    const int layer = (int)(r & 1);
    const long long t = r >> 1;
    const int cj = (int)(t % (B.n[j] + 4)) - 2, ck = (int)(t / (B.n[j] + 4)) - 2;
    int c[3];
    c[d] = (face & 1) ? B.n[d] + layer : -1 - layer;
    c[j] = cj;
    c[k] = ck;
    int g[3] = {B.org[0] + c[0], B.org[1] + c[1], B.org[2] + c[2]};
    int sb;
    unsigned li;
    const bool ok = owner_of(cv, D, g, &sb, &li);
    R *dst = buf + B.base + B.hoff[face] + r;
#pragma unroll
    for (int f = 0; f < 5; f++)
      dst[f * B.fs] = ok ? buf[D[sb].base + f * D[sb].fs + li] : R(0);
  }
}

// Flag halo exchange: the flags of the cells just beyond the outer face of
// dimension d (inside the chart) are the neighbour's flags of its first
// layer, DiffusiveFlux_d only -- the flag of the face between the bricks.
__global__ __launch_bounds__(256) void brick_flag_halo_kernel(const BrickDesc *D, int count,
                                                              const long long *fstart,
                                                              long long nfhalo, ChartView cv,
                                                              uint8_t *flags) {
  for (long long h = blockIdx.x * 256ll + threadIdx.x; h < nfhalo;
       h += (long long)gridDim.x * 256) {
    const int q = find_segment(fstart, 3 * count, h);
    const int b = q / 3, d = q % 3;
    const BrickDesc &B = D[b];
    const long long r = h - fstart[q];
    const int j = d == 0 ? 1 : 0, k = d == 2 ? 1 : 2;
    int c[3];
    c[d] = B.n[d];
    c[j] = (int)(r % B.n[j]);
    c[k] = (int)(r / B.n[j]);
    int g[3] = {B.org[0] + c[0], B.org[1] + c[1], B.org[2] + c[2]};
    int sb;
    unsigned li;
    uint8_t v = 0;
    if (owner_of(cv, D, g, &sb, &li)) v = flags[D[sb].fbase + li] & (uint8_t)(1u << d);
    flags[B.fbase + B.fhoff[d] + r] = v;
  }
}

// Which faces of a brick are faces of the chart, and where its fluence goes.
struct BrickFaces {
  int N[3];    // chart cells (the fluence layout)
  int org[3];  // first chart cell of the brick
  int rec;     // bit b: record the fluence of face b (a face of the chart)
};

// The fluxes of the chart faces among a line's two ends, from the line's
// four-cell stencils with the brick's ghost rule, bit for bit the stage
// kernel's; recorded at the chart's slab index of the line.
template <typename R, int STAGE, int D>
__device__ inline void brick_line(const StageArgs<R> &a, const BoundaryArgs &B,
                                  const BrickFaces &bf, long long idx, double I[NB_SUM]) {
  int u, v;
  line_coords<D>(a.n, idx, &u, &v);
  const int j = D == 0 ? 1 : 0, k = D == 2 ? 1 : 2;
  const long long gidx = (long long)(u + bf.org[j]) + (long long)bf.N[j] * (v + bf.org[k]);
#pragma unroll
  for (int side = 0; side < 2; side++) {
    if (!((bf.rec >> (2 * D + side)) & 1)) continue;
    const int cl = side ? a.n[D] - 1 : -1;  // the cell left of the face
    R Vs[4][5];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int cd = cl - 1 + q;
      const int x = D == 0 ? cd : u, y = D == 0 ? u : (D == 1 ? cd : v), z = D == 2 ? cd : v;
      load_prim<true>(a, x, y, z, Vs[q]);
    }
    const bool df = side ? false : ((ld(a.flags, line_cell<D>(a.n, u, v, 0)) >> D) & 1u);
    R F[5];
    stencil_flux<R, D, RECON_LINEAR>(a, cl, Vs[0], Vs[1], Vs[2], Vs[3], df, F);
    double val[5];
    record_face<STAGE>(B, bf.N, 2 * D + side, gidx, F, val);
    if (STAGE == 2) add_integrands<D>(B, a.gs, side, u, v, side ? a.n[D] : 0, val, I);
  }
}

template <typename R, int STAGE>
__global__ __launch_bounds__(TALLY_BLOCK) void brick_boundary_kernel(StageArgs<R> a,
                                                                     BoundaryArgs B,
                                                                     BrickFaces bf) {
  const int n[3] = {a.n[0], a.n[1], a.n[2]};
  const long long e = blockIdx.x * (long long)TALLY_BLOCK + threadIdx.x;
  double I[NB_SUM];
#pragma unroll
  for (int i = 0; i < NB_SUM; i++) I[i] = 0.0;
  if (e < boundary_lines(n)) {
    int d;
    long long idx;
    boundary_line_of(n, e, &d, &idx);
    if (d == 0)
      brick_line<R, STAGE, 0>(a, B, bf, idx, I);
    else if (d == 1)
      brick_line<R, STAGE, 1>(a, B, bf, idx, I);
    else
      brick_line<R, STAGE, 2>(a, B, bf, idx, I);
  }
  if (STAGE == 2) {
    __shared__ double red[NB_SUM][TALLY_BLOCK];
#pragma unroll
    for (int i = 0; i < NB_SUM; i++) red[i][threadIdx.x] = I[i];
    __syncthreads();
    for (int o = TALLY_BLOCK / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o)
#pragma unroll
        for (int i = 0; i < NB_SUM; i++) red[i][threadIdx.x] += red[i][threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x < NB_SUM) B.partial[blockIdx.x * NB_SUM + threadIdx.x] = red[threadIdx.x][0];
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static ChartView chart_view(const gx_ctx *c) {
  const BrickSet &S = *c->bricks;
  ChartView cv;
  for (int d = 0; d < 3; d++) {
    cv.N[d] = c->n[d];
    cv.nb[d] = S.nb[d];
    cv.nbc[d] = S.nbc[d];
  }
  cv.periodic = S.periodic;
  return cv;
}

// Up to BRICKS_PER_LAUNCH bricks of one stage (or of the shock flags) in one
// launch: their arguments travel by value in the kernel-argument segment, a
// brick's tiles follow the previous brick's (tstart), and the launch's tiles
// are dealt to the XCDs as the single-chart kernel deals one chart's, so
// small bricks fill the chip together.  A block reads its brick's arguments
// in place (a reference into the argument segment, loaded as scalars), as
// the single-chart kernel reads its own.
constexpr int BRICKS_PER_LAUNCH = 8;

template <typename R> struct BrickLaunch {
  StageArgs<R> a[BRICKS_PER_LAUNCH];
  int tstart[BRICKS_PER_LAUNCH + 1];
  int count;  // bricks in this launch
  int ntile;  // tiles in this launch (tstart[count])
};
// kernel arguments are limited to 4 KB
static_assert(sizeof(BrickLaunch<double>) <= 4096 && sizeof(BrickLaunch<float>) <= 4096,
              "BrickLaunch must fit the kernel-argument segment");

// The brick of this launch whose run of tiles holds tile t (uniform).
template <typename R>
__device__ __forceinline__ int launch_brick(const BrickLaunch<R> &L, int t) {
  int b = 0;
#pragma unroll
  for (int i = 1; i < BRICKS_PER_LAUNCH; i++) b += (i < L.count && t >= L.tstart[i]) ? 1 : 0;
  return b;
}

// P2: every divisor of the chart's geometry is a power of two (gdiv); the
// bricks' geometries are slices of the chart's.
template <typename R, int STAGE, typename T, bool P2 = false>
__global__ __launch_bounds__(NT, StageOccupancy<R>::waves) void bricks_stage_kernel(
    BrickLaunch<R> L) {
  __shared__ StageShared<R, T> S;
  const int t = tile_of_block(blockIdx.x, L.ntile);
  const int b = launch_brick(L, t);
  const StageArgs<R> &a = L.a[b];
  const int lt = t - L.tstart[b];
  stage_tile<R, STAGE, RECON_LINEAR, false, true, T, P2>(a, (lt % a.nbx) * T::X,
                                                     ((lt / a.nbx) % a.nby) * T::Y,
                                                     (lt / (a.nbx * a.nby)) * a.tz, S);
}

// The shock flags of the launch's bricks (16 x 16 tiles of a.tz planes),
// written to each brick's flag array (a.flags).
template <typename R>
__global__ __launch_bounds__(NT) void bricks_flags_kernel(BrickLaunch<R> L) {
  __shared__ FlagShared<R> S;
  const int t = tile_of_block(blockIdx.x, L.ntile);
  const int b = launch_brick(L, t);
  const StageArgs<R> &a = L.a[b];
  const int fnbx = (a.n[0] + TX - 1) / TX, fnby = (a.n[1] + TY - 1) / TY;
  const int lt = t - L.tstart[b];
  shock_flags_tile<R, true>(a, (lt % fnbx) * TX, ((lt / fnbx) % fnby) * TY,
                            (lt / (fnbx * fnby)) * a.tz, const_cast<uint8_t *>(a.flags), S);
}

template <typename R>
static StageArgs<R> brick_args(const gx_ctx *c, int b) {
  const BrickSet &S = *c->bricks;
  const BrickDesc &B = S.desc[b];
  StageArgs<R> a = base_args<R>(c);
  for (int d = 0; d < 3; d++) a.n[d] = B.n[d];
  for (int i = 0; i < 6; i++) {
    a.bc[i] = B.bc[i];
    a.hoff[i] = B.hoff[i];
  }
  for (int d = 0; d < 3; d++) a.fhoff[d] = B.fhoff[d];
  a.chart_face = B.chart_face;
  a.gs = B.gs;
  a.fs = B.fs;
  a.geo = (const R *)S.geo + B.geo_off;
  a.flags = S.flags + B.fbase;
  const int shape = stage_shape(c);
  a.nbx = (B.n[0] + tile_x(shape) - 1) / tile_x(shape);
  a.nby = (B.n[1] + tile_y(shape) - 1) / tile_y(shape);
  a.nblk = a.nbx * a.nby * ((B.n[2] + c->tz - 1) / c->tz);
  return a;
}

// Launches of BrickLaunch groups: set(a, b) completes brick b's arguments
// (buffers, dt, ...), tiles(a) gives its tile count.
template <typename R, typename Set, typename Tiles, typename Launch>
static int launch_groups(gx_ctx *c, Set set, Tiles tiles, Launch launch) {
  BrickSet &S = *c->bricks;
  for (int g0 = 0; g0 < S.count; g0 += BRICKS_PER_LAUNCH) {
    BrickLaunch<R> L;
    memset(&L, 0, sizeof(L));
    L.count = std::min(BRICKS_PER_LAUNCH, S.count - g0);
    long long sum = 0;
    for (int i = 0; i < L.count; i++) {
      L.a[i] = brick_args<R>(c, g0 + i);
      set(L.a[i], g0 + i);
      L.tstart[i] = (int)sum;
      sum += tiles(L.a[i]);
    }
    if (sum >= (1ll << 31))
      return fail(GX_E_ARG, "gx_step_rk2: too many tiles in one launch of bricks");
    for (int i = L.count; i <= BRICKS_PER_LAUNCH; i++) L.tstart[i] = (int)sum;
    L.ntile = (int)sum;
    if (L.ntile > 0) {
      launch(L);
      HIP_TRY(hipGetLastError());
    }
  }
  return GX_OK;
}

// One RK stage of every brick: set(a, b) gives the stage's buffers and dt.
template <typename R, int STAGE, typename Set>
static int launch_bricks_stage(gx_ctx *c, Set set) {
  const int shape = stage_shape(c);
  auto tiles = [](const StageArgs<R> &a) { return (long long)a.nblk; };
  return launch_groups<R>(c, set, tiles, [&](const BrickLaunch<R> &L) {
    if constexpr (std::is_same<R, double>::value) {
      if (c->geom.pow2) {
        if (shape == 1)
          bricks_stage_kernel<R, STAGE, Tile32x8, true><<<L.ntile, NT, 0, c->stream>>>(L);
        else
          bricks_stage_kernel<R, STAGE, Tile16x16, true><<<L.ntile, NT, 0, c->stream>>>(L);
        return;
      }
    }
    if (shape == 1)
      bricks_stage_kernel<R, STAGE, Tile32x8><<<L.ntile, NT, 0, c->stream>>>(L);
    else
      bricks_stage_kernel<R, STAGE, Tile16x16><<<L.ntile, NT, 0, c->stream>>>(L);
  });
}

// The shock flags of every brick from src (16 x 16 tiles; the z-chunk is
// chosen for all bricks together).
template <typename R>
static int launch_bricks_flags(gx_ctx *c, const R *src) {
  BrickSet &S = *c->bricks;
  if (S.ftz == 0) {
    long long fxy = 0;
    for (const BrickDesc &B : S.desc)
      fxy += (long long)((B.n[0] + TX - 1) / TX) * ((B.n[1] + TY - 1) / TY);
    // every brick has the chart's cells divided by nBricks: one z extent
    S.ftz = flag_tz((int)std::min(fxy, 1ll << 30), S.desc[0].n[2]);
  }
  auto set = [&](StageArgs<R> &a, int b) {
    const BrickDesc &B = S.desc[b];
    a.src = src + B.base;
    a.tz = S.ftz;
  };
  auto tiles = [](const StageArgs<R> &a) {
    return (long long)((a.n[0] + TX - 1) / TX) * ((a.n[1] + TY - 1) / TY) *
           ((a.n[2] + a.tz - 1) / a.tz);
  };
  return launch_groups<R>(c, set, tiles, [&](const BrickLaunch<R> &L) {
    bricks_flags_kernel<R><<<L.ntile, NT, 0, c->stream>>>(L);
  });
}

template <typename R>
static int exchange_halo(gx_ctx *c, void *buf) {
  BrickSet &S = *c->bricks;
  if (S.nhalo == 0) return GX_OK;
  brick_halo_kernel<R><<<grid_blocks(S.nhalo), 256, 0, c->stream>>>(
      S.desc_d, S.count, S.hstart_d, S.nhalo, chart_view(c), (R *)buf);
  HIP_TRY(hipGetLastError());
  return GX_OK;
}

static int exchange_flag_halo(gx_ctx *c) {
  BrickSet &S = *c->bricks;
  if (S.nfhalo == 0) return GX_OK;
  brick_flag_halo_kernel<<<grid_blocks(S.nfhalo), 256, 0, c->stream>>>(
      S.desc_d, S.count, S.fstart_d, S.nfhalo, chart_view(c), S.flags);
  HIP_TRY(hipGetLastError());
  return GX_OK;
}

// nf host fields of the chart's proper cells (x fastest) into a buffer of
// bricks (field f of brick b at base + f * fs), and back.
static int bricks_put(gx_ctx *c, void *dev, const double *host, int nf) {
  BrickSet &S = *c->bricks;
  const size_t esz = c->esz;
  for (const BrickDesc &B : S.desc) {
    std::vector<char> t((size_t)nf * B.ncell * esz);
    for (int f = 0; f < nf; f++)
      for (int z = 0; z < B.n[2]; z++)
        for (int y = 0; y < B.n[1]; y++)
          for (int x = 0; x < B.n[0]; x++) {
            const size_t src = (size_t)f * c->ncell + (size_t)(x + B.org[0]) +
                               (size_t)c->n[0] * ((y + B.org[1]) + (size_t)c->n[1] * (z + B.org[2]));
            const size_t dst = (size_t)f * B.ncell + x + (size_t)B.n[0] * (y + (size_t)B.n[1] * z);
            if (c->fp32)
              ((float *)t.data())[dst] = (float)host[src];
            else
              ((double *)t.data())[dst] = host[src];
          }
    for (int f = 0; f < nf; f++)
      HIP_TRY(hipMemcpyAsync((char *)dev + (size_t)(B.base + f * B.fs) * esz,
                             t.data() + (size_t)f * B.ncell * esz, (size_t)B.ncell * esz,
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return GX_OK;
}

static int bricks_fetch(gx_ctx *c, double *host, const void *dev, int nf) {
  BrickSet &S = *c->bricks;
  const size_t esz = c->esz;
  for (const BrickDesc &B : S.desc) {
    std::vector<char> t((size_t)nf * B.ncell * esz);
    for (int f = 0; f < nf; f++)
      HIP_TRY(hipMemcpyAsync(t.data() + (size_t)f * B.ncell * esz,
                             (const char *)dev + (size_t)(B.base + f * B.fs) * esz,
                             (size_t)B.ncell * esz, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int f = 0; f < nf; f++)
      for (int z = 0; z < B.n[2]; z++)
        for (int y = 0; y < B.n[1]; y++)
          for (int x = 0; x < B.n[0]; x++) {
            const size_t dst = (size_t)f * c->ncell + (size_t)(x + B.org[0]) +
                               (size_t)c->n[0] * ((y + B.org[1]) + (size_t)c->n[1] * (z + B.org[2]));
            const size_t src = (size_t)f * B.ncell + x + (size_t)B.n[0] * (y + (size_t)B.n[1] * z);
            host[dst] = c->fp32 ? (double)((const float *)t.data())[src]
                                : ((const double *)t.data())[src];
          }
  }
  return GX_OK;
}

void release_bricks(gx_ctx *c) {
  if (!c || !c->bricks) return;
  BrickSet *S = c->bricks;
  for (void *p : {(void *)S->desc_d, (void *)S->hstart_d, (void *)S->fstart_d, S->U, S->Y, S->W,
                  S->P0, S->geo, (void *)S->flags, (void *)S->tgeo, (void *)S->cur})
    (void)hipFree(p);
  delete S;
  c->bricks = nullptr;
}

// Lays out the bricks of c (c->n, c->grid and c->geom set) and allocates
// their buffers.
int make_bricks(gx_ctx *c, const int nBricks[3]) {
  BrickSet *S = new BrickSet();
  c->bricks = S;
  S->count = nBricks[0] * nBricks[1] * nBricks[2];
  for (int d = 0; d < 3; d++) {
    S->nb[d] = nBricks[d];
    S->nbc[d] = c->n[d] / nBricks[d];
    if (c->grid.BoundaryCondition[2 * d] == GX_BC_PERIODIC) S->periodic |= 1 << d;
  }
  S->desc.resize(S->count);
  std::vector<long long> hstart(6 * (size_t)S->count + 1, 0), fstart(3 * (size_t)S->count + 1, 0);
  long long base = 0, fbase = 0, geo_off = 0, tg_off = 0;
  int part = 0, bpart = 0;
  std::vector<char> geo_img;
  std::vector<double> tg_img;
  for (int b2 = 0; b2 < nBricks[2]; b2++)
    for (int b1 = 0; b1 < nBricks[1]; b1++)
      for (int b0 = 0; b0 < nBricks[0]; b0++) {
        const int b = b0 + nBricks[0] * (b1 + nBricks[1] * b2);
        const int ia[3] = {b0, b1, b2};
        BrickDesc &B = S->desc[b];
        memset(&B, 0, sizeof(B));
        B.ncell = 1;
        for (int d = 0; d < 3; d++) {
          B.n[d] = S->nbc[d];
          B.org[d] = ia[d] * S->nbc[d];
          B.ncell *= B.n[d];
        }
        long long halo = 0;
        for (int face = 0; face < 6; face++) {
          const int d = face >> 1, side = face & 1;
          const bool at_chart = side ? ia[d] == nBricks[d] - 1 : ia[d] == 0;
          const int cbc = c->grid.BoundaryCondition[face];
          if (at_chart) B.chart_face |= 1 << face;
          if (!at_chart || (cbc == GX_BC_PERIODIC && nBricks[d] > 1))
            B.bc[face] = BC_NEIGHBOR;
          else
            B.bc[face] = cbc;
          B.hoff[face] = -1;
          if (B.bc[face] == BC_NEIGHBOR) {
            const int j = d == 0 ? 1 : 0, k = d == 2 ? 1 : 2;
            B.hlen[face] = 2 * (B.n[j] + 4) * (B.n[k] + 4);
            B.hoff[face] = (int)(B.ncell + halo);
            halo += B.hlen[face];
          }
          hstart[6 * (size_t)b + face + 1] = B.hlen[face];
        }
        long long fh = 0;
        for (int d = 0; d < 3; d++) {
          B.fhoff[d] = -1;
          long long len = 0;
          if (!(B.chart_face & (2 << (2 * d)))) {  // outer face inside the chart
            const int j = d == 0 ? 1 : 0, k = d == 2 ? 1 : 2;
            len = (long long)B.n[j] * B.n[k];
            B.fhoff[d] = (int)(B.ncell + fh);
            fh += len;
          }
          fstart[3 * (size_t)b + d + 1] = len;
        }
        B.fs = B.ncell + halo;
        // The brick's geometry: the slice of the chart's padded arrays from
        // its first cell's lower ghosts on (Chart_SLD__Form.f90:312-317).
        ChartGeometry g;
        g.gs = 0;
        for (int d = 0; d < 3; d++) {
          g.n[d] = B.n[d];
          for (int a = 0; a < N_GEO; a++)
            g.a[d][a].assign(c->geom.a[d][a].begin() + B.org[d],
                             c->geom.a[d][a].begin() + B.org[d] + B.n[d] + 4);
          g.center[d].assign(c->geom.center[d].begin() + B.org[d],
                             c->geom.center[d].begin() + B.org[d] + B.n[d] + 4);
          g.gs = std::max(g.gs, B.n[d] + 4);
        }
        B.gs = g.gs;
        const std::vector<char> img = geometry_image(g, c->fp32);
        B.geo_off = geo_off;
        geo_img.insert(geo_img.end(), img.begin(), img.end());
        geo_off += (long long)(img.size() / c->esz);
        B.tg_off = tg_off;
        std::vector<double> tg((size_t)3 * N_TG * B.gs, 0.0);
        for (int d = 0; d < 3; d++)
          for (int p = 0; p < B.n[d] + 4; p++) {
            const double ctr = g.center[d][p], wl = g.a[d][G_WL][p];
            tg[(size_t)(3 * d + TG_CENTER) * B.gs + p] = ctr;
            tg[(size_t)(3 * d + TG_WIDTH) * B.gs + p] = wl + g.a[d][G_WR][p];
            tg[(size_t)(3 * d + TG_FACE) * B.gs + p] = ctr - wl;
          }
        tg_img.insert(tg_img.end(), tg.begin(), tg.end());
        tg_off += (long long)tg.size();
        B.npart = std::min(grid_blocks(B.ncell), 1024);
        B.part_off = part;
        part += B.npart;
        B.nbpart = (int)((boundary_lines(B.n) + TALLY_BLOCK - 1) / TALLY_BLOCK);
        B.bpart_off = bpart;
        bpart += B.nbpart;
      }
  for (int b = 0; b < S->count; b++) {
    BrickDesc &B = S->desc[b];
    B.base = base;
    B.fbase = fbase;
    base += 5 * B.fs;
    fbase += B.ncell;
    for (int d = 0; d < 3; d++)
      if (B.fhoff[d] >= 0) fbase += fstart[3 * (size_t)b + d + 1];
  }
  for (size_t i = 1; i < hstart.size(); i++) hstart[i] += hstart[i - 1];
  for (size_t i = 1; i < fstart.size(); i++) fstart[i] += fstart[i - 1];
  S->fstart = fstart;
  S->total_fs = base / 5;
  S->total_flags = fbase;
  S->nhalo = hstart.back();
  S->nfhalo = fstart.back();
  c->n_partial = part;
  c->n_bpartial = bpart;
  const size_t bytes = c->esz * (size_t)base;
  HIP_TRY(hipMalloc(&S->U, bytes));
  HIP_TRY(hipMalloc(&S->Y, bytes));
  HIP_TRY(hipMalloc(&S->W, bytes));
  HIP_TRY(hipMalloc(&S->P0, bytes));
  for (void *p : {S->U, S->Y, S->W, S->P0}) HIP_TRY(hipMemset(p, 0, bytes));
  HIP_TRY(hipMalloc(&S->flags, (size_t)fbase));
  HIP_TRY(hipMemset(S->flags, 0, (size_t)fbase));
  HIP_TRY(hipMalloc(&S->geo, geo_img.size()));
  HIP_TRY(hipMemcpy(S->geo, geo_img.data(), geo_img.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&S->tgeo, sizeof(double) * tg_img.size()));
  HIP_TRY(hipMemcpy(S->tgeo, tg_img.data(), sizeof(double) * tg_img.size(),
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&S->desc_d, sizeof(BrickDesc) * S->count));
  HIP_TRY(hipMemcpy(S->desc_d, S->desc.data(), sizeof(BrickDesc) * S->count,
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&S->hstart_d, sizeof(long long) * hstart.size()));
  HIP_TRY(hipMemcpy(S->hstart_d, hstart.data(), sizeof(long long) * hstart.size(),
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&S->fstart_d, sizeof(long long) * fstart.size()));
  HIP_TRY(hipMemcpy(S->fstart_d, fstart.data(), sizeof(long long) * fstart.size(),
                    hipMemcpyHostToDevice));
  return GX_OK;
}

int bricks_set_state(gx_ctx *c, const double *primitive, const double *conserved) {
  BrickSet &S = *c->bricks;
  int rc = bricks_put(c, S.P0, primitive, 5);
  if (rc) return rc;
  if (conserved) {
    rc = bricks_put(c, S.U, conserved, 5);
    if (rc) return rc;
  }
  return by_precision(c->fp32, [&](auto tag) {
    using R = decltype(tag);
    if (!conserved)
      for (const BrickDesc &B : S.desc)
        from_primitive_kernel<R><<<grid_blocks(B.ncell), 256, 0, c->stream>>>(
            B.ncell, B.fs, c->params.BaryonMassReference, (R *)S.P0 + B.base, (R *)S.U + B.base);
    HIP_TRY(hipGetLastError());
    // stage 1 of the next step reads the primitives' halos; the conserved
    // halos serve the flags and later steps
    int r = exchange_halo<R>(c, S.P0);
    if (r) return r;
    r = exchange_halo<R>(c, S.U);
    if (r) return r;
    HIP_TRY(hipMemsetAsync(S.flags, 0, (size_t)S.total_flags, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GX_OK;
  });
}

int bricks_get(gx_ctx *c, int which, double *out) {
  BrickSet &S = *c->bricks;
  return bricks_fetch(c, out, which == 0 ? S.U : S.P0, 5);
}

int bricks_current(gx_ctx *c, double *host) {
  BrickSet &S = *c->bricks;
  long long most = 0;
  for (const BrickDesc &B : S.desc) most = std::max(most, B.ncell);
  if (!S.cur) HIP_TRY(hipMalloc(&S.cur, sizeof(double) * GX_N_FIELDS * (size_t)most));
  // c_v = k/(mu(gamma-1)) with the default gamma (Fluid_P_I__Form.f90:206-223)
  const double c_v = 1.0 / (1.0 * (1.4 - 1.0));
  std::vector<double> t(GX_N_FIELDS * (size_t)most);
  for (int b = 0; b < S.count; b++) {
    const BrickDesc &B = S.desc[b];
    by_precision(c->fp32, [&](auto tag) {
      using R = decltype(tag);
      StageArgs<R> a = brick_args<R>(c, b);
      a.aux = (const R *)S.U + B.base;
      current_kernel<R><<<grid_blocks(B.ncell), 256, 0, c->stream>>>(
          a, c_v, c->prims_valid ? (const R *)S.P0 + B.base : nullptr, S.cur);
      return 0;
    });
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(t.data(), S.cur, sizeof(double) * GX_N_FIELDS * B.ncell,
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int f = 0; f < GX_N_FIELDS; f++)
      for (int z = 0; z < B.n[2]; z++)
        for (int y = 0; y < B.n[1]; y++)
          for (int x = 0; x < B.n[0]; x++)
            host[(size_t)f * c->ncell + (size_t)(x + B.org[0]) +
                 (size_t)c->n[0] * ((y + B.org[1]) + (size_t)c->n[1] * (z + B.org[2]))] =
                t[(size_t)f * B.ncell + x + (size_t)B.n[0] * (y + (size_t)B.n[1] * z)];
  }
  return GX_OK;
}

int bricks_get_flags(gx_ctx *c, unsigned char *flags) {
  BrickSet &S = *c->bricks;
  std::vector<unsigned char> t((size_t)S.total_flags);
  HIP_TRY(hipMemcpyAsync(t.data(), S.flags, t.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (const BrickDesc &B : S.desc)
    for (int z = 0; z < B.n[2]; z++)
      for (int y = 0; y < B.n[1]; y++)
        for (int x = 0; x < B.n[0]; x++)
          flags[(size_t)(x + B.org[0]) +
                (size_t)c->n[0] * ((y + B.org[1]) + (size_t)c->n[1] * (z + B.org[2]))] =
              t[B.fbase + x + (size_t)B.n[0] * (y + (size_t)B.n[1] * z)];
  return GX_OK;
}

int bricks_set_flags(gx_ctx *c, const unsigned char *flags) {
  BrickSet &S = *c->bricks;
  std::vector<unsigned char> t((size_t)S.total_flags, 0);
  for (const BrickDesc &B : S.desc)
    for (int z = 0; z < B.n[2]; z++)
      for (int y = 0; y < B.n[1]; y++)
        for (int x = 0; x < B.n[0]; x++)
          t[B.fbase + x + (size_t)B.n[0] * (y + (size_t)B.n[1] * z)] =
              flags[(size_t)(x + B.org[0]) +
                    (size_t)c->n[0] * ((y + B.org[1]) + (size_t)c->n[1] * (z + B.org[2]))];
  HIP_TRY(hipMemcpyAsync(S.flags, t.data(), t.size(), hipMemcpyHostToDevice, c->stream));
  int rc = exchange_flag_halo(c);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
}

// CFL maximum of the current state over every brick into c->invdt.
int bricks_cfl(gx_ctx *c) {
  BrickSet &S = *c->bricks;
  HIP_TRY(hipMemsetAsync(c->invdt, 0, sizeof(unsigned long long), c->stream));
  for (int b = 0; b < S.count; b++) {
    const BrickDesc &B = S.desc[b];
    by_precision(c->fp32, [&](auto tag) {
      using R = decltype(tag);
      StageArgs<R> a = brick_args<R>(c, b);
      a.src_is_prim = c->prims_valid ? 1 : 0;
      a.src = (const R *)(c->prims_valid ? S.P0 : S.U) + B.base;
      cfl_kernel<R><<<grid_blocks(B.ncell), 256, 0, c->stream>>>(a);
      return 0;
    });
    HIP_TRY(hipGetLastError());
  }
  return GX_OK;
}

int bricks_step(gx_ctx *c, double dt) {
  BrickSet &S = *c->bricks;
  hipStream_t st = c->stream;
  BoundaryArgs BA;
  BA.bf = c->fluence;
  BA.fw = 0.5 * dt;  // Weight_RK * TimeStep, B = (1/2, 1/2)
  const int rc = by_precision(c->fp32, [&](auto tag) {
    using R = decltype(tag);
    R *U = (R *)S.U, *Y = (R *)S.Y, *W = (R *)S.W, *P0 = (R *)S.P0;
    auto faces = [&](const BrickDesc &B) {
      BrickFaces bf;
      for (int d = 0; d < 3; d++) {
        bf.N[d] = c->n[d];
        bf.org[d] = B.org[d];
      }
      bf.rec = B.chart_face;
      return bf;
    };
    // Stage 1: Current = uploaded primitives or the Solution.
    const int prim = c->prims_valid ? 1 : 0;
    const R *src1 = prim ? P0 : U;
    int r = launch_bricks_stage<R, 1>(c, [&](StageArgs<R> &a, int b) {
      const BrickDesc &B = S.desc[b];
      a.dt = dt;
      a.src_is_prim = prim;
      a.src = src1 + B.base;
      a.aux = U + B.base;
      a.out1 = Y + B.base;
      a.out2 = W + B.base;
    });
    if (r) return r;
    for (int b = 0; b < S.count; b++) {
      const BrickDesc &B = S.desc[b];
      if (!B.chart_face) continue;
      StageArgs<R> a = brick_args<R>(c, b);
      a.src_is_prim = prim;
      a.src = src1 + B.base;
      BoundaryArgs Bb = BA;
      Bb.tg = S.tgeo + B.tg_off;
      Bb.partial = c->bpartial + (size_t)B.bpart_off * NB_SUM;
      brick_boundary_kernel<R, 1><<<B.nbpart, TALLY_BLOCK, 0, st>>>(a, Bb, faces(B));
      HIP_TRY(hipGetLastError());
    }
    r = exchange_halo<R>(c, Y);
    if (r) return r;
    HIP_TRY(hipMemsetAsync(c->invdt, 0, sizeof(unsigned long long), st));
    HIP_TRY(hipMemsetAsync(c->bpartial, 0, sizeof(double) * NB_SUM * c->n_bpartial, st));
    // Stage 2: Current = Y.
    r = launch_bricks_stage<R, 2>(c, [&](StageArgs<R> &a, int b) {
      const BrickDesc &B = S.desc[b];
      a.dt = dt;
      a.src = Y + B.base;
      a.aux = W + B.base;
      a.out1 = U + B.base;
    });
    if (r) return r;
    for (int b = 0; b < S.count; b++) {
      const BrickDesc &B = S.desc[b];
      if (!B.chart_face) continue;
      StageArgs<R> a = brick_args<R>(c, b);
      a.src = Y + B.base;
      BoundaryArgs Bb = BA;
      Bb.tg = S.tgeo + B.tg_off;
      Bb.partial = c->bpartial + (size_t)B.bpart_off * NB_SUM;
      brick_boundary_kernel<R, 2><<<B.nbpart, TALLY_BLOCK, 0, st>>>(a, Bb, faces(B));
      HIP_TRY(hipGetLastError());
    }
    boundary_tally_finish<<<1, TALLY_BLOCK, 0, st>>>(c->bpartial, c->n_bpartial,
                                                     c->tally_local);
    HIP_TRY(hipGetLastError());
    r = exchange_halo<R>(c, U);
    if (r) return r;
    // Shock flags of the next step from the new state, then the flags of
    // the faces beyond each brick.
    r = launch_bricks_flags<R>(c, (const R *)U);
    if (r) return r;
    return exchange_flag_halo(c);
  });
  if (rc) return rc;
  c->prims_valid = false;
  c->invdt_valid = true;
  c->cfl_pending = false;
  return GX_OK;
}

int bricks_totals(gx_ctx *c, double out[5]) {
  BrickSet &S = *c->bricks;
  for (int b = 0; b < S.count; b++) {
    const BrickDesc &B = S.desc[b];
    by_precision(c->fp32, [&](auto tag) {
      using R = decltype(tag);
      StageArgs<R> a = brick_args<R>(c, b);
      a.aux = (const R *)S.U + B.base;
      totals_kernel<R><<<B.npart, 256, 0, c->stream>>>(a, c->partial + 5 * (size_t)B.part_off);
      return 0;
    });
    HIP_TRY(hipGetLastError());
  }
  std::vector<double> h(5 * (size_t)c->n_partial);
  HIP_TRY(hipMemcpyAsync(h.data(), c->partial, sizeof(double) * h.size(),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int f = 0; f < 5; f++) {
    double s = 0.0;
    for (int p = 0; p < c->n_partial; p++) s += h[(size_t)p * 5 + f];
    out[f] = s;
  }
  return GX_OK;
}

int bricks_interior_partials(gx_ctx *c) {
  BrickSet &S = *c->bricks;
  for (int b = 0; b < S.count; b++) {
    const BrickDesc &B = S.desc[b];
    by_precision(c->fp32, [&](auto tag) {
      using R = decltype(tag);
      StageArgs<R> a = brick_args<R>(c, b);
      a.aux = (const R *)S.U + B.base;
      interior_tally_kernel<R><<<B.npart, TALLY_BLOCK, 0, c->stream>>>(
          a, c->prims_valid ? (const R *)S.P0 + B.base : nullptr, S.tgeo + B.tg_off,
          c->tpartial + (size_t)NI_SUM * B.part_off);
      return 0;
    });
    HIP_TRY(hipGetLastError());
  }
  return GX_OK;
}

}  // namespace gx

using namespace gx;

extern "C" {

int gx_create_bricks(const gx_grid *grid, const gx_params *params, const int nBricks[3],
                     gx_ctx **out) {
  GX_API_BEGIN
  if (!grid || !nBricks || !out) return fail(GX_E_ARG, "gx_create_bricks: null argument");
  *out = nullptr;
  gx_params p;
  if (params)
    p = *params;
  else
    gx_default_params(&p);
  int rc = check_params(p, "gx_create_bricks");
  if (rc) return rc;
  if (p.ReconstructionType != GX_LINEAR)
    return fail(GX_E_ARG, "gx_create_bricks: bricked charts take ReconstructionType GX_LINEAR");
  rc = check_grid(*grid, "gx_create_bricks");
  if (rc) return rc;
  for (int d = 0; d < 3; d++) {
    if (nBricks[d] < 1 || grid->nCells[d] % nBricks[d] != 0)
      return fail(GX_E_ARG, "gx_create_bricks: nBricks must divide nCells evenly in every "
                            "dimension (Chart_Template % SetBrick)");
    if (grid->nCells[d] / nBricks[d] < 2)
      return fail(GX_E_ARG, "gx_create_bricks: a brick needs at least 2 cells per dimension");
  }
  if (nBricks[0] * nBricks[1] * nBricks[2] > 4096)
    return fail(GX_E_ARG, "gx_create_bricks: at most 4096 bricks");
  gx_ctx *c = new (std::nothrow) gx_ctx();
  if (!c) return fail(GX_E_NOMEM, "gx_create_bricks: host allocation failed");
  c->grid = *grid;
  c->params = p;
  c->fp32 = p.Precision == GX_FP32;
  c->esz = c->fp32 ? sizeof(float) : sizeof(double);
  c->ncell = 1;
  for (int d = 0; d < 3; d++) {
    c->n[d] = grid->nCells[d];
    c->ncell *= c->n[d];
  }
  rc = make_geometry(*grid, c->geom, "gx_create_bricks");
  if (rc) {
    delete c;
    return rc;
  }
  c->gs = c->geom.gs;
  auto fail_clean = [&](int code) {
    gx_destroy(c);
    return code;
  };
  if (hipStreamCreate(&c->stream) != hipSuccess)
    return fail_clean(fail(GX_E_HIP, "gx_create_bricks: hipStreamCreate failed"));
  rc = make_bricks(c, nBricks);
  if (rc) return fail_clean(rc);
  const BrickDesc &B0 = c->bricks->desc[0];
  const int tiles = ((B0.n[0] + TX - 1) / TX) * ((B0.n[1] + TY - 1) / TY);
  // the default z-chunk of the whole chart's launches, shared by its bricks
  c->tz = stage_tz(tiles * c->bricks->count, B0.n[2]);
  const size_t nf = (size_t)fluence_count(c->n);
#define TRY_CLEAN(expr)                                                   \
  do {                                                                    \
    hipError_t e_ = (expr);                                               \
    if (e_ != hipSuccess)                                                 \
      return fail_clean(fail(GX_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_))); \
  } while (0)
  TRY_CLEAN(hipMalloc(&c->invdt, sizeof(unsigned long long)));
  TRY_CLEAN(hipMalloc(&c->partial, sizeof(double) * 5 * c->n_partial));
  TRY_CLEAN(hipMalloc(&c->tpartial, sizeof(double) * NI_SUM * c->n_partial));
  TRY_CLEAN(hipMalloc(&c->bpartial, sizeof(double) * NB_SUM * c->n_bpartial));
  TRY_CLEAN(hipMalloc(&c->fluence, sizeof(double) * nf));
  TRY_CLEAN(hipMemset(c->fluence, 0, sizeof(double) * nf));
  TRY_CLEAN(hipMalloc(&c->tally_local, sizeof(double) * TALLY_N));
  TRY_CLEAN(hipMemset(c->tally_local, 0, sizeof(double) * TALLY_N));
#undef TRY_CLEAN
  *out = c;
  return GX_OK;
  GX_API_END("gx_create_bricks")
}

int gx_get_bricks(gx_ctx *c, int nBricks[3]) {
  GX_API_BEGIN
  if (!c || !nBricks) return fail(GX_E_ARG, "gx_get_bricks: null argument");
  for (int d = 0; d < 3; d++) nBricks[d] = c->bricks ? c->bricks->nb[d] : 1;
  return GX_OK;
  GX_API_END("gx_get_bricks")
}

}  // extern "C"
