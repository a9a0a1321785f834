// gx_tiles.hpp -- device side of the Fluid_P_I finite-volume step: launch
// geometry, ghost rule, the fused stage tile, the shock-flag tile and the
// small whole-chart kernels, templated on the real type of the state.
//
// Device layout: the proper cells only (no ghost storage), five conserved
// fields D, S_1, S_2, S_3, G as separate slabs of n0*n1*n2 values, x fastest.
// Ghost values are produced on load by index mapping: periodic wrap, or the
// mirror copy of OUTFLOW / REFLECTING (CH/SingleLevelCharts/
// Chart_SL__Template.f90:486-538, Chart_SL__Kernel.f90:12-113).  This equals
// the reference's ghost handling for every value its sweeps read: the swept
// dimension's ghosts over the transverse interior.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/gx_fluid.h"
#include "gx_kernels.hpp"

namespace gx {

constexpr int TX = 16;           // tile width in x (cells): the default stage tile
constexpr int TY = 16;           // tile height in y; also the shock-flag tile
constexpr int NT = TX * TY;      // threads per block
constexpr int WAVE = 64;
constexpr int NXCD = 8;          // accelerator complex dies, one L2 each

// Shape of a stage tile: X x Y cells, one thread each (NT threads); the
// plane tile with two ghost layers is HX x HY, the x-slope and x-flux rows
// are SXW and FXW long.
template <int X_, int Y_>
struct TileShape {
  static constexpr int X = X_, Y = Y_;
  static constexpr int HX = X + 4, HY = Y + 4, SXW = X + 2, FXW = X + 1;
  static_assert(X * Y == NT, "a stage tile has one cell per thread");
  static_assert(X <= WAVE && Y <= WAVE, "wave 0 takes the right column, wave 1 the top row");
};
using Tile16x16 = TileShape<16, 16>;
using Tile32x8 = TileShape<32, 8>;
// Tile shapes by number (gx_set_tile, gx_batch_set_tiles): 0 = 16 x 16,
// 1 = 32 x 8.
constexpr int N_TILE_SHAPES = 2;
__host__ __device__ constexpr int tile_x(int shape) { return shape == 1 ? 32 : 16; }
__host__ __device__ constexpr int tile_y(int shape) { return shape == 1 ? 8 : 16; }

__host__ __device__ inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Field f of a five-field state lives at base + f * fs (fs = cells).  The
// geometry is one buffer: array g of dimension d at geo + (d*N_GEO + g)*gs.
// Few pointers keep the kernel's scalar registers free.  R is the real type
// of the state and the sweep arithmetic (double, or float for the FP32
// variant); physics constants stay double and are converted on use.
template <typename R>
struct StageArgs {
  int n[3];
  int bc[6];
  int tz;               // planes per block (z-chunk)
  int src_is_prim;      // stage input holds primitives (first stage after upload)
  int gs;               // geometry stride (max n + 4)
  int nbx, nby, nblk;   // stage-kernel tiles in x, in y, in total
  long long fs;         // field stride (cells)
  double dt;
  Physics ph;
  const R *src;         // stage input: conserved (or primitive) state
  const R *aux;         // stage 1: U (Solution), the same buffer as src unless src
                        // holds primitives; stage 2: W = U + K1/2
  R *out1;              // stage 1: Y = U + K1; stage 2: U_new
  R *out2;              // stage 1: W; stage 2: the state before the fix, if set
  const R *geo;
  const uint8_t *flags; // DiffusiveFlux bits of the step
  unsigned long long *invdt_bits;  // stage 2: max CFL term (as FP64 bits)
  R *sources;           // Current % Sources of the divergence, if recorded
  double rdt;           // RN(1/dt) for the recording, 0: divide
  // Bricks only.  hoff[b]: first value of the halo block of face b (bc
  // BC_NEIGHBOR) in every field slab; fhoff[d]: first flag of the ghost layer
  // beyond the outer face of dimension d when that face is inside the chart
  // (-1: none, the flags there are 0); chart_face: bit b set when face b is
  // a face of the chart (ClearBoundary clears there, periodic or not).
  int hoff[6];
  int fhoff[3];
  int chart_face;
};

// Waves per SIMD the stage kernel is compiled for: the FP32 tile uses half
// the registers and LDS of the FP64 one.  The single-chart LINEAR stage fits
// the 128 registers of four FP32 waves without spilling; the other variants
// are bounded at three (about 168 registers) and take four where they fit.
template <typename R, bool LEAN = false> struct StageOccupancy {
  static constexpr int waves = 2;
};
template <bool LEAN> struct StageOccupancy<float, LEAN> {
  static constexpr int waves = LEAN ? 4 : 3;
};

// Loads and stores at a 32-bit element offset from a base pointer: the byte
// offset stays 32-bit (charts hold < 2^29 cells, checked at creation), so the
// address is a scalar base plus one 32-bit VGPR.
template <typename T>
__device__ __forceinline__ T ld(const T *base, unsigned idx) {
  return *(const T *)((const char *)base + (size_t)(idx * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st(T *base, unsigned idx, T v) {
  *(T *)((char *)base + (size_t)(idx * (unsigned)sizeof(T))) = v;
}

template <typename R>
__device__ __forceinline__ R geo(const StageArgs<R> &a, int d, int g, int p) {
  return ld(a.geo + (d * N_GEO + g) * a.gs, (unsigned)p);
}

__device__ __forceinline__ unsigned lin(const int n[3], int x, int y, int z) {
  return (unsigned)x + (unsigned)n[0] * ((unsigned)y + (unsigned)n[1] * (unsigned)z);
}

// Ghost index rule of one dimension: periodic wrap, or the mirror copy of
// OUTFLOW / REFLECTING; *flip is set for a reflecting face.
__device__ __forceinline__ int map_coord(int c, int n, int bc_in, int bc_out,
                                         bool *flip) {
  if (c < 0) {
    if (bc_in == BC_PERIODIC) return c + n;
    *flip = bc_in == BC_REFLECTING;
    return -1 - c;
  }
  if (c >= n) {
    if (bc_out == BC_PERIODIC) return c - n;
    *flip = bc_out == BC_REFLECTING;
    return 2 * n - 1 - c;
  }
  return c;
}

// Halo value of the ghost cell c of brick face b = 2d + s (BC_NEIGHBOR): a
// block of two layers (layer 0 next to the brick) by the two transverse
// extents with two ghost cells on each side (edges included), the layer
// fastest, then the lower, then the higher transverse dimension.
// v[i] for a run-time i, as a chain of selects over constant indices: the
// arguments of a kernel that builds them in registers (one launch over many
// bricks) then stay in registers instead of going to scratch memory.
template <int N>
__device__ __forceinline__ int pick(const int (&v)[N], int i) {
  int r = v[0];
#pragma unroll
  for (int k = 1; k < N; k++) r = i == k ? v[k] : r;
  return r;
}

template <typename A>
__device__ __forceinline__ unsigned halo_index(const A &a, int b, int cx, int cy, int cz) {
  const int d = b >> 1;
  const int cd = d == 0 ? cx : (d == 1 ? cy : cz);
  const int cj = d == 0 ? cy : cx, ck = d == 2 ? cy : cz;
  const int nj = d == 0 ? a.n[1] : a.n[0];
  const int layer = (b & 1) ? cd - pick(a.n, d) : -1 - cd;
  return (unsigned)pick(a.hoff, b) + (unsigned)layer +
         2u * ((unsigned)(cj + 2) + (unsigned)(nj + 4) * (unsigned)(ck + 2));
}

// Linear index of cell (x, y, z) with at most one coordinate outside the
// chart by at most two cells, after the ghost rule of that face; *flip = the
// dimension whose velocity a reflecting face negates, or -1.  A brick reads
// the ghosts of its BC_NEIGHBOR faces from their halos.
template <bool BRICK = false, typename A>
__device__ __forceinline__ unsigned cell_index(const A &a, int x, int y, int z,
                                               int *flip) {
  *flip = -1;
  if ((unsigned)x < (unsigned)a.n[0] && (unsigned)y < (unsigned)a.n[1] &&
      (unsigned)z < (unsigned)a.n[2])
    return lin(a.n, x, y, z);
  if constexpr (BRICK) {
    const int b = x < 0 ? 0 : (x >= a.n[0] ? 1 : (y < 0 ? 2 : (y >= a.n[1] ? 3 : (z < 0 ? 4 : 5))));
    if (pick(a.bc, b) == BC_NEIGHBOR) return halo_index(a, b, x, y, z);
  }
  bool fx = false, fy = false, fz = false;
  const int cx = map_coord(x, a.n[0], a.bc[0], a.bc[1], &fx);
  const int cy = map_coord(y, a.n[1], a.bc[2], a.bc[3], &fy);
  const int cz = map_coord(z, a.n[2], a.bc[4], a.bc[5], &fz);
  *flip = fx ? 0 : (fy ? 1 : (fz ? 2 : -1));
  return lin(a.n, cx, cy, cz);
}

// Raw stage input (conserved, or primitive after an upload) of one cell.
template <bool BRICK = false, typename R>
__device__ __forceinline__ void load_raw(const StageArgs<R> &a, int x, int y, int z,
                                         R raw[5], int *flip) {
  const unsigned i = cell_index<BRICK>(a, x, y, z, flip);
#pragma unroll
  for (int f = 0; f < 5; f++) raw[f] = ld(a.src + f * a.fs, i);
}

// -v when f, as a sign-bit operation: a select on an array element is turned
// into an indexed (scratch) store by the compiler.
__device__ __forceinline__ double negate_if(double v, bool f) {
  return __longlong_as_double(__double_as_longlong(v) ^
                              (f ? (long long)0x8000000000000000ull : 0ll));
}
__device__ __forceinline__ float negate_if(float v, bool f) {
  return __int_as_float(__float_as_int(v) ^ (f ? (int)0x80000000u : 0));
}

// Primitives of a raw value, with the reflecting sign flip.
template <typename R>
__device__ __forceinline__ void raw_to_prim(const StageArgs<R> &a, const R raw[5],
                                            int flip, R V[5]) {
  if (a.src_is_prim) {
#pragma unroll
    for (int f = 0; f < 5; f++) V[f] = raw[f];
  } else {
    prim_from_cons(raw, R(a.ph.mass), a.ph.nmin, V);
  }
  V[1] = negate_if(V[1], flip == 0);
  V[2] = negate_if(V[2], flip == 1);
  V[3] = negate_if(V[3], flip == 2);
}

template <bool BRICK = false, typename R>
__device__ __forceinline__ void load_prim(const StageArgs<R> &a, int x, int y,
                                          int z, R V[5]) {
  R raw[5];
  int flip;
  load_raw<BRICK>(a, x, y, z, raw, &flip);
  raw_to_prim(a, raw, flip, V);
}

template <bool P2 = false, typename R>
__device__ __forceinline__ R slope_at(const StageArgs<R> &a, int d, int c, R vm,
                                      R v0, R vp) {
  const R *g = a.geo + d * N_GEO * a.gs + (c + 2);
  const int s = a.gs;
  return slope<P2>(vm, v0, vp, g[G_DXL * s], g[G_DXR * s], g[G_DXL2 * s], g[G_DXR2 * s],
               g[G_DEN * s], g[G_RDXL * s], g[G_RDXR * s], g[G_RDEN * s],
               R(a.ph.theta), a.ph.use_limiter);
}

// slope_at with the left quotient carried in *q (slope_carry).
template <bool P2 = false, typename R>
__device__ __forceinline__ R slope_carry_at(const StageArgs<R> &a, int d, int c, R vm,
                                            R v0, R vp, R *q) {
  const R *g = a.geo + d * N_GEO * a.gs + (c + 2);
  const int s = a.gs;
  return slope_carry<P2>(vm, v0, vp, g[G_DXR * s], g[G_DXL2 * s], g[G_DXR2 * s], g[G_DEN * s],
                     g[G_RDXR * s], g[G_RDEN * s], R(a.ph.theta), a.ph.use_limiter, q);
}

// Left quotient |theta dVL / dxl| of cell c (slope_carry's *q).
template <bool P2 = false, typename R>
__device__ __forceinline__ R left_quotient_at(const StageArgs<R> &a, int d, int c, R vm,
                                              R v0) {
  const R *g = a.geo + d * N_GEO * a.gs + (c + 2);
  const int s = a.gs;
  return fabs(gdiv<P2>(R(a.ph.theta) * (v0 - vm), g[G_DXL * s], g[G_RDXL * s]));
}

// Parabolic face values of cell c of dimension d from the global geometry.
template <typename R>
__device__ __forceinline__ void parabolic_at(const StageArgs<R> &a, int d, int c, R vm,
                                             R v0, R vp, R *vI, R *vO) {
  const R *g = a.geo + d * N_GEO * a.gs + (c + 2);
  const int s = a.gs;
  parabolic_faces(vm, v0, vp, g[G_CENTER * s - 1], g[G_CENTER * s], g[G_CENTER * s + 1],
                  g[G_WL * s], g[G_WR * s], vI, vO);
}

// XCD-aware tile order.  The dispatcher deals consecutive blocks to the XCDs
// in turn; block b is given tile t so that every XCD owns one contiguous run
// of tiles (x fastest, then y, then z-chunk).  Tiles that share halo cache
// lines then run on the same L2 at about the same time.  A bijection of
// [0, nblk): XCD j holds q + (j < r) blocks and tiles from j*q + min(j, r).
__device__ __forceinline__ int tile_of_block(int b, int nblk) {
  const int q = nblk / NXCD, r = nblk % NXCD;
  const int xcd = b % NXCD, slot = b / NXCD;
  return xcd * q + min(xcd, r) + slot;
}

// Halo cell t of the plane tile: 4 per row for x (two each side), 4 per
// column for y.
__device__ __forceinline__ void halo_cell(int t, int x0, int y0, int tw, int th,
                                          int *hx, int *hy) {
  if (t < 4 * th) {
    const int j = t >> 2, s = t & 3;
    *hy = y0 + j;
    *hx = s < 2 ? x0 - 2 + s : x0 + tw + (s - 2);
  } else {
    const int u = t - 4 * th;
    const int i = u >> 2, s = u & 3;
    *hx = x0 + i;
    *hy = s < 2 ? y0 - 2 + s : y0 + th + (s - 2);
  }
}

// Geometry of the tile in LDS: array g of dimension d (0 = x, 1 = y) at the
// local index c - origin + 2 for c in [origin-2, origin+X+2) (the x extent
// bounds both dimensions: X >= Y).
template <typename R, typename T = Tile16x16> using TileGeo = R[2][N_GEO][T::HX];

template <bool P2 = false, typename R, typename G>
__device__ __forceinline__ R slope_lds(const G &sG, int d, int l, R vm,
                                       R v0, R vp, R theta, int use_limiter) {
  return slope<P2>(vm, v0, vp, sG[d][G_DXL][l], sG[d][G_DXR][l], sG[d][G_DXL2][l],
               sG[d][G_DXR2][l], sG[d][G_DEN][l], sG[d][G_RDXL][l],
               sG[d][G_RDXR][l], sG[d][G_RDEN][l], theta, use_limiter);
}

// Parabolic face values of the cell at local geometry index l.
template <typename R, typename G>
__device__ __forceinline__ void parabolic_lds(const G &sG, int d, int l, R vm,
                                              R v0, R vp, R *vI, R *vO) {
  parabolic_faces(vm, v0, vp, sG[d][G_CENTER][l - 1], sG[d][G_CENTER][l],
                  sG[d][G_CENTER][l + 1], sG[d][G_WL][l], sG[d][G_WR][l], vI, vO);
}

// Face flux from given face states into slot of sF.
template <typename R, typename SF>
__device__ __forceinline__ void face_to_lds(const Physics &ph, const R WL[5], const R WR[5],
                                            int d, bool df, SF sF, int slot) {
  R F[5];
  face_flux(WL, WR, d, df, ph, F);
#pragma unroll
  for (int f = 0; f < 5; f++) sF[f][slot] = F[f];
}

// x-face i (between cells x0+i-1 and x0+i) of row j into sFx.
template <typename T, typename R, typename SP, typename SS, typename SF>
__device__ __forceinline__ void x_face(const Physics &ph, SP sP, SS sSx, SF sFx,
                                       const TileGeo<R, T> &sG, int i, int j, bool df) {
  constexpr int HX = T::HX, SXW = T::SXW, FXW = T::FXW;
  const int bl = (j + 2) * HX + (i + 1);
  const R wr = sG[0][G_WR][i + 1], wl = sG[0][G_WL][i + 2];
  R WL[5], WR[5], F[5];
#pragma unroll
  for (int f = 0; f < 5; f++) {
    WL[f] = sP[f][bl] + wr * sSx[f][j * SXW + i];
    WR[f] = sP[f][bl + 1] - wl * sSx[f][j * SXW + i + 1];
  }
  face_flux(WL, WR, 0, df, ph, F);
#pragma unroll
  for (int f = 0; f < 5; f++) sFx[f][j * FXW + i] = F[f];
}

// y-face j (between rows y0+j-1 and y0+j) of column i into sFy.
template <typename T, typename R, typename SP, typename SS, typename SF>
__device__ __forceinline__ void y_face(const Physics &ph, SP sP, SS sSy, SF sFy,
                                       const TileGeo<R, T> &sG, int i, int j, bool df) {
  constexpr int HX = T::HX, TX = T::X;
  const int bl = (j + 1) * HX + (i + 2);
  const R wr = sG[1][G_WR][j + 1], wl = sG[1][G_WL][j + 2];
  R WL[5], WR[5], F[5];
#pragma unroll
  for (int f = 0; f < 5; f++) {
    WL[f] = sP[f][bl] + wr * sSy[f][j * TX + i];
    WR[f] = sP[f][bl + HX] - wl * sSy[f][(j + 1) * TX + i];
  }
  face_flux(WL, WR, 1, df, ph, F);
#pragma unroll
  for (int f = 0; f < 5; f++) sFy[f][j * TX + i] = F[f];
}

// Fused stage: K = dt * (-div F) over the three dimensions, then the RK
// combination.  A block owns a TX x TY column tile and marches along z over
// a.tz planes; x/y stencils come from the plane tile in LDS, the z stencil
// from registers.  Every thread computes the x- and y-slopes and the left and
// bottom faces of its own cell; the tile-edge slopes and faces are extra
// tasks.  Global loads of plane k+1 (halo, z-stencil) are issued during plane
// k, and the RK input of plane k at its start, so that their latency overlaps
// the slope and flux work.  Stage 2 applies the conserved fix and the CFL term
// in FP64 whatever R is.
// LDS of one stage tile.
template <typename R, typename T = Tile16x16> struct StageShared {
  R sP[5][T::HY * T::HX];
  R sSx[5][T::Y * T::SXW];       // x-slopes of cells x0-1 .. x0+tw
  R sSy[5][(T::Y + 2) * T::X];   // y-slopes of cells y0-1 .. y0+th
  R sFx[5][T::Y * T::FXW];       // x-face fluxes x0 .. x0+tw
  R sFy[5][(T::Y + 1) * T::X];   // y-face fluxes y0 .. y0+th
  double sRed[NT / WAVE];
  TileGeo<R, T> sG;
};

// PARABOLIC only: the inner-face values of the cells just beyond the tile's
// right column and top row (the slopes' LDS then holds outer-face values).
template <typename R, typename T = Tile16x16> struct EdgeShared {
  R sEx[5][T::Y];
  R sEy[5][T::X];
};

// REC: record the divergence sources (a.sources).  BRICK: a brick of a
// chart split into bricks (ghosts of BC_NEIGHBOR faces from the halos, and
// the flags of the faces beyond the brick from the flag halo).  T: the
// tile's shape.  P2: every divisor of the chart's geometry is a power of two
// (gdiv).
template <typename R, int STAGE, int RECON = RECON_LINEAR, bool REC = false,
          bool BRICK = false, typename T = Tile16x16, bool P2 = false>
__device__ __forceinline__ void stage_tile(const StageArgs<R> &a, int x0, int y0, int z0,
                                           StageShared<R, T> &L,
                                           EdgeShared<R, T> *E = nullptr) {
  constexpr bool PARA = RECON == RECON_PARABOLIC;
  constexpr int TX = T::X, TY = T::Y, HX = T::HX, SXW = T::SXW, FXW = T::FXW;
  auto &sP = L.sP;
  auto &sSx = L.sSx;
  auto &sSy = L.sSy;
  auto &sFx = L.sFx;
  auto &sFy = L.sFy;
  auto &sRed = L.sRed;
  auto &sG = L.sG;
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const int tw = min(TX, nx - x0), th = min(TY, ny - y0);
  const int zend = min(z0 + a.tz, nz);
  const int tid = threadIdx.x;
  // Tile geometry: x for cells x0-2 .. x0+tw+1, y for y0-2 .. y0+th+1.
  for (int t = tid; t < 2 * N_GEO * HX; t += NT) {
    const int d = t / (N_GEO * HX), g = (t / HX) % N_GEO, l = t % HX;
    const int o = d == 0 ? x0 : y0, w = d == 0 ? tw : th;
    sG[d][g][l] = l < w + 4 ? geo(a, d, g, o + l) : R(0);
  }
  const int tx = tid % TX, ty = tid / TX;
  const bool active = tx < tw && ty < th;
  const int x = x0 + tx, y = y0 + ty;
  const int ctr = (ty + 2) * HX + tx + 2;  // own cell in the plane tile
  const Physics ph = a.ph;
  const R theta = R(ph.theta);
  const int nhalo = 4 * th + 4 * tw;
  const int nxs = 2 * th, nys = 2 * tw;  // extra slope tasks

  R Vk[5], Vk1[5], sk[5], Flo[5];
  // The FP64 stage 1 carries the z-slope's right quotient to the next plane
  // (−1.9 %; stage 2, with its CFL and fix epilogue, gained nothing, and
  // FP32 keeps its 128 registers instead)
  constexpr bool ZCARRY = sizeof(R) == sizeof(double) && STAGE == 1;
  R qz[5];
  // Stage 1 takes the RK input U of its own cell from the z-stencil values
  // it loaded two planes earlier (raw values of planes k and k+1 carried in
  // registers) instead of loading U again, which by then has left L2 (40 B
  // per cell of HBM reads).  After an upload the stencil holds primitives and
  // U is loaded.
  // (FP64 only: FP32 keeps the 128 registers of four waves per SIMD)
  constexpr bool UCARRY = STAGE == 1 && sizeof(R) == sizeof(double);
  R rk[5], rk1[5];
  R zr[5], hr[5];  // prefetched raw values: z-stencil k+2, halo of k
  int zflip = -1, hflip = -1, hpos = 0;
  unsigned flk = 0;     // flag byte of the own cell at plane k (then k+1)
  double tmax = 0.0;

  if (active) {
    R Vm2[5], Vm1[5], sm1[5], WL[5], WR[5];
    load_prim<BRICK>(a, x, y, z0 - 2, Vm2);
    load_prim<BRICK>(a, x, y, z0 - 1, Vm1);
    if constexpr (UCARRY) {
      int fl;
      load_raw<BRICK>(a, x, y, z0, rk, &fl);
      raw_to_prim(a, rk, fl, Vk);
      load_raw<BRICK>(a, x, y, z0 + 1, rk1, &fl);
      raw_to_prim(a, rk1, fl, Vk1);
    } else {
      load_prim<BRICK>(a, x, y, z0, Vk);
      load_prim<BRICK>(a, x, y, z0 + 1, Vk1);
    }
    load_raw<BRICK>(a, x, y, z0 + 2, zr, &zflip);
    flk = ld(a.flags, lin(a.n, x, y, z0));
    if constexpr (PARA) {
      // sk holds the outer-face value of plane k
#pragma unroll
      for (int f = 0; f < 5; f++) {
        R iM1, iK;
        parabolic_at(a, 2, z0 - 1, Vm2[f], Vm1[f], Vk[f], &iM1, &sm1[f]);
        parabolic_at(a, 2, z0, Vm1[f], Vk[f], Vk1[f], &iK, &sk[f]);
        WL[f] = sm1[f];
        WR[f] = iK;
      }
    } else {
#pragma unroll
      for (int f = 0; f < 5; f++) {
        if constexpr (ZCARRY) {
          qz[f] = left_quotient_at<P2>(a, 2, z0 - 1, Vm2[f], Vm1[f]);
          sm1[f] = slope_carry_at<P2>(a, 2, z0 - 1, Vm2[f], Vm1[f], Vk[f], &qz[f]);
          sk[f] = slope_carry_at<P2>(a, 2, z0, Vm1[f], Vk[f], Vk1[f], &qz[f]);
        } else {
          sm1[f] = slope_at<P2>(a, 2, z0 - 1, Vm2[f], Vm1[f], Vk[f]);
          sk[f] = slope_at<P2>(a, 2, z0, Vm1[f], Vk[f], Vk1[f]);
        }
        WL[f] = Vm1[f] + geo(a, 2, G_WR, z0 - 1 + 2) * sm1[f];
        WR[f] = Vk[f] - geo(a, 2, G_WL, z0 + 2) * sk[f];
      }
    }
    face_flux(WL, WR, 2, (flk >> 2) & 1, ph, Flo);
  }
  // Halo cells of the plane tile.  FP64: waves 2 and 3 convert the next
  // plane's halo while waves 0 and 1 compute the tile-edge faces, so the
  // plane's load phase only stores them.
  // (the shape's halo must fit waves 2 and 3)
  constexpr bool LATE = sizeof(R) == sizeof(double) && 4 * (TX + TY) <= 2 * WAVE;
  const int wave = __builtin_amdgcn_readfirstlane(tid / WAVE);
  const int hid = LATE ? tid - 2 * WAVE : tid;
  const bool htask = hid >= 0 && hid < nhalo;
  R hv[5];
  if (htask) {
    int hx, hy;
    halo_cell(hid, x0, y0, tw, th, &hx, &hy);
    hpos = (hy - y0 + 2) * HX + (hx - x0 + 2);
    load_raw<BRICK>(a, hx, hy, z0, hr, &hflip);
    if constexpr (LATE) raw_to_prim(a, hr, hflip, hv);
  }

  for (int k = z0; k < zend; ++k) {
    const bool more = k + 1 < zend;
    R ar[5], zn[5];
    R Vk2[5];  // primitives of the own column at k+2 (FP64: converted early)
    int znflip = -1;
    unsigned fln = 0, flx = 0;  // own flags at k+1; flag of this thread's edge face
    if (active) {
#pragma unroll
      for (int f = 0; f < 5; f++) sP[f][ctr] = Vk[f];
    }
    if (htask) {
      if constexpr (!LATE) raw_to_prim(a, hr, hflip, hv);
#pragma unroll
      for (int f = 0; f < 5; f++) sP[f][hpos] = hv[f];
      if (more) {
        int hx, hy;
        halo_cell(hid, x0, y0, tw, th, &hx, &hy);
        load_raw<BRICK>(a, hx, hy, k + 1, hr, &hflip);
      }
    }
    if (active) {
      const unsigned ic = lin(a.n, x, y, k);
      if (UCARRY && !a.src_is_prim) {
#pragma unroll
        for (int f = 0; f < 5; f++) ar[f] = rk[f];
      } else {
#pragma unroll
        for (int f = 0; f < 5; f++) ar[f] = ld(a.aux + f * a.fs, ic);
      }
      if (more) load_raw<BRICK>(a, x, y, k + 3, zn, &znflip);
      if (k + 1 < nz)
        fln = ld(a.flags, lin(a.n, x, y, k + 1));
      else if (BRICK && a.fhoff[2] >= 0)
        fln = ld(a.flags, (unsigned)a.fhoff[2] + (unsigned)x + (unsigned)nx * (unsigned)y);
    }
    if (tid < th) {
      if (x0 + tw < nx)
        flx = ld(a.flags, lin(a.n, x0 + tw, y0 + tid, k)) & 1u;
      else if (BRICK && a.fhoff[0] >= 0)
        flx = ld(a.flags, (unsigned)a.fhoff[0] + (unsigned)(y0 + tid) + (unsigned)ny * (unsigned)k) &
              1u;
    } else if (tid >= WAVE && tid - WAVE < tw) {
      if (y0 + th < ny)
        flx = (ld(a.flags, lin(a.n, x0 + tid - WAVE, y0 + th, k)) >> 1) & 1u;
      else if (BRICK && a.fhoff[1] >= 0)
        flx = (ld(a.flags, (unsigned)a.fhoff[1] + (unsigned)(x0 + tid - WAVE) +
                               (unsigned)nx * (unsigned)k) >> 1) & 1u;
    }
    __syncthreads();

    // Slopes: own cell (x and y), then the tile-edge cells.  PARABOLIC: the
    // outer-face values go where the slopes go, the own inner-face values
    // stay in registers, the edge cells' inner-face values go to E.
    R ix[5], iy[5];
    if (active) {
      const int bx = ctr;
      if constexpr (PARA) {
#pragma unroll
        for (int f = 0; f < 5; f++)
          parabolic_lds(sG, 0, tx + 2, sP[f][bx - 1], sP[f][bx], sP[f][bx + 1], &ix[f],
                        &sSx[f][ty * SXW + tx + 1]);
#pragma unroll
        for (int f = 0; f < 5; f++)
          parabolic_lds(sG, 1, ty + 2, sP[f][bx - HX], sP[f][bx], sP[f][bx + HX], &iy[f],
                        &sSy[f][(ty + 1) * TX + tx]);
      } else {
#pragma unroll
        for (int f = 0; f < 5; f++)
          sSx[f][ty * SXW + tx + 1] = slope_lds<P2>(sG, 0, tx + 2, sP[f][bx - 1], sP[f][bx],
                                                sP[f][bx + 1], theta, ph.use_limiter);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int f = 0; f < 5; f++)
          sSy[f][(ty + 1) * TX + tx] = slope_lds<P2>(sG, 1, ty + 2, sP[f][bx - HX], sP[f][bx],
                                                 sP[f][bx + HX], theta, ph.use_limiter);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // FP64: every wave takes the edge cells, wave w field w (wave 3 also
    // field 4), so no wave adds more than two fields per edge task to its
    // own work.  FP32 has wave 0 take all five fields: with four blocks per
    // CU the others fill its barrier waits, and sharing measured slower.
    // (with 2 (TX + TY) <= 64 edge cells, one wave-pass per field)
  constexpr bool SPREAD = sizeof(R) == sizeof(double) && 2 * (TX + TY) <= WAVE;
    const int ut = SPREAD ? tid % WAVE : tid;
    if (ut < nxs + nys) {
      int base, stride, l, out, gd;
      bool isx = ut < nxs;
      if (isx) {
        const int j = ut >> 1, side = ut & 1;
        const int i = side ? tw + 1 : 0;  // slope index of cell x0-1+i
        base = (j + 2) * HX + (i + 1);
        stride = 1;
        l = i + 1;  // local geometry index of cell x0-1+i
        gd = 0;
        out = j * SXW + i;
      } else {
        const int u = ut - nxs;
        const int i = u >> 1, side = u & 1;
        const int j = side ? th + 1 : 0;  // slope index of cell y0-1+j
        base = (j + 1) * HX + (i + 2);
        stride = HX;
        l = j + 1;
        gd = 1;
        out = j * TX + i;
      }
      if constexpr (PARA) {
        // left/bottom edge cell: its outer-face value; right/top: its inner
        const bool hi = isx ? (ut & 1) : ((ut - nxs) & 1);
        const int e = isx ? (ut >> 1) : ((ut - nxs) >> 1);
        const int f0 = SPREAD ? wave : 0;
        const int fs = SPREAD ? (wave == 3 ? 1 : 5) : 1;
        for (int f = f0; f < 5; f += fs) {
          R vI, vO;
          parabolic_lds(sG, gd, l, sP[f][base - stride], sP[f][base], sP[f][base + stride],
                        &vI, &vO);
          if (!hi) {
            if (isx)
              sSx[f][out] = vO;
            else
              sSy[f][out] = vO;
          } else {
            if (isx)
              E->sEx[f][e] = vI;
            else
              E->sEy[f][e] = vI;
          }
        }
      } else if constexpr (SPREAD) {
        for (int f = wave; f < 5; f += wave == 3 ? 1 : 5) {
          const R sv = slope_lds<P2>(sG, gd, l, sP[f][base - stride], sP[f][base],
                                 sP[f][base + stride], theta, ph.use_limiter);
          if (isx)
            sSx[f][out] = sv;
          else
            sSy[f][out] = sv;
        }
      } else {
#pragma unroll
        for (int f = 0; f < 5; f++) {
          const R sv = slope_lds<P2>(sG, gd, l, sP[f][base - stride], sP[f][base],
                                 sP[f][base + stride], theta, ph.use_limiter);
          if (isx)
            sSx[f][out] = sv;
          else
            sSy[f][out] = sv;
        }
      }
    }
    // FP64: waves 0 and 1, whose slope work ends a field early, convert the
    // own column's next z-stencil value here; waves 2 and 3 do it beside the
    // halo conversion in the face step
    if constexpr (LATE) {
      if (active && wave < 2) raw_to_prim(a, zr, zflip, Vk2);
    }
    __syncthreads();

    // Faces: left and bottom faces of the own cell, then the right column of
    // x-faces (wave 0) and the top row of y-faces (wave 1).
    if constexpr (PARA) {
      R WL[5], WR[5];
      if (active) {
#pragma unroll
        for (int f = 0; f < 5; f++) {
          WL[f] = sSx[f][ty * SXW + tx];
          WR[f] = ix[f];
        }
        face_to_lds(ph, WL, WR, 0, flk & 1u, sFx, ty * FXW + tx);
#pragma unroll
        for (int f = 0; f < 5; f++) {
          WL[f] = sSy[f][ty * TX + tx];
          WR[f] = iy[f];
        }
        face_to_lds(ph, WL, WR, 1, (flk >> 1) & 1u, sFy, ty * TX + tx);
      }
      if (tid < th) {
#pragma unroll
        for (int f = 0; f < 5; f++) {
          WL[f] = sSx[f][tid * SXW + tw];
          WR[f] = E->sEx[f][tid];
        }
        face_to_lds(ph, WL, WR, 0, flx, sFx, tid * FXW + tw);
      }
      if (tid >= WAVE && tid - WAVE < tw) {
        const int i = tid - WAVE;
#pragma unroll
        for (int f = 0; f < 5; f++) {
          WL[f] = sSy[f][th * TX + i];
          WR[f] = E->sEy[f][i];
        }
        face_to_lds(ph, WL, WR, 1, flx, sFy, th * TX + i);
      }
    } else {
      if (active) {
        x_face<T>(ph, sP, sSx, sFx, sG, tx, ty, flk & 1u);
        __builtin_amdgcn_sched_barrier(0);
        y_face<T>(ph, sP, sSy, sFy, sG, tx, ty, (flk >> 1) & 1u);
        __builtin_amdgcn_sched_barrier(0);
      }
      if (tid < th) x_face<T>(ph, sP, sSx, sFx, sG, tw, tid, flx);
      __builtin_amdgcn_sched_barrier(0);
      if (tid >= WAVE && tid - WAVE < tw) y_face<T>(ph, sP, sSy, sFy, sG, tid - WAVE, th, flx);
    }
    if constexpr (LATE) {
      if (htask && more) raw_to_prim(a, hr, hflip, hv);  // the next plane's halo
      if (active && wave >= 2) raw_to_prim(a, zr, zflip, Vk2);
    }
    __syncthreads();

    if (active) {
      // z stencil of plane k: slope at k+1, flux through the face k+1/2.
      R sk1[5], Fhi[5];
      {
        R WL[5], WR[5];
        if constexpr (!LATE) raw_to_prim(a, zr, zflip, Vk2);
        if constexpr (PARA) {
#pragma unroll
          for (int f = 0; f < 5; f++) {
            R iK1;
            parabolic_at(a, 2, k + 1, sP[f][ctr], Vk1[f], Vk2[f], &iK1, &sk1[f]);
            WL[f] = sk[f];
            WR[f] = iK1;
          }
        } else {
#pragma unroll
          for (int f = 0; f < 5; f++) {
            const R vk = sP[f][ctr];
            if constexpr (ZCARRY)
              sk1[f] = slope_carry_at<P2>(a, 2, k + 1, vk, Vk1[f], Vk2[f], &qz[f]);
            else
              sk1[f] = slope_at<P2>(a, 2, k + 1, vk, Vk1[f], Vk2[f]);
            WL[f] = vk + geo(a, 2, G_WR, k + 2) * sk[f];
            WR[f] = Vk1[f] - geo(a, 2, G_WL, k + 1 + 2) * sk1[f];
          }
        }
        face_flux(WL, WR, 2, (fln >> 2) & 1u, ph, Fhi);
      }
      // ComputeIncrement_CSL_Kernel (EV/Increments/IncrementDivergence_FV__
      // Kernel.f90:672-752), dimensions in order 1, 2, 3, onto K = 0; the
      // division by the volume uses its reciprocal (FP64: rounded exactly).
      const R dX = sG[0][G_WIDTH][tx + 2], dY = sG[1][G_WIDTH][ty + 2],
              dZ = geo(a, 2, G_WIDTH, k + 2);
      const R Vol = dX * dY * dZ;
      R rVol;
      if constexpr (P2) {
        // Vol = 2^(e-1) exactly (frexp gives the mantissa 1/2): 1/Vol = 2^(1-e)
        int e;
        (void)frexp(Vol, &e);
        rVol = ldexp(R(1), 1 - e);
      } else {
        rVol = Arith<R>::recip(Vol);
      }
      const R A1 = dY * dZ, A2 = dZ * dX, A3 = dX * dY;
      const R dt = R(a.dt);
      const unsigned ic = lin(a.n, x, y, k);
      R Kc[5];
#pragma unroll
      for (int f = 0; f < 5; f++) {
        const R fxl = sFx[f][ty * FXW + tx], fxh = sFx[f][ty * FXW + tx + 1];
        const R fyl = sFy[f][ty * TX + tx], fyh = sFy[f][(ty + 1) * TX + tx];
        R K = R(0);
        K = K + gdiv<P2>(dt * (A1 * fxl - A1 * fxh), Vol, rVol);
        K = K + gdiv<P2>(dt * (A2 * fyl - A2 * fyh), Vol, rVol);
        K = K + gdiv<P2>(dt * (A3 * Flo[f] - A3 * Fhi[f]), Vol, rVol);
        Kc[f] = K;
      }
      if constexpr (REC) {
        // RecordDivergenceKernel (EV/Steps/Step_RK_C_ASC__Kernel.f90:12-46):
        // Sources = Sources + Weight_RK * Increment / TimeStep, Weight_RK =
        // 1/2, Sources cleared in stage 1 (Step_RK_C_ASC__Template.f90:
        // 1572-1589)
        const R hdt = R(a.dt), rdt = R(a.rdt);
#pragma unroll
        for (int f = 0; f < 5; f++) {
          R *sp = a.sources + f * a.fs;
          const R q = rdt != R(0) ? Arith<R>::rdiv(R(0.5) * Kc[f], hdt, rdt)
                                  : R(0.5) * Kc[f] / hdt;
          st(sp, ic, (STAGE == 1 ? R(0) : ld(sp, ic)) + q);
        }
      }
      if (STAGE == 1) {
        // Y = Solution + 1*K1; Solution + 1/2 K1 kept for the final combination
        // (Step_RK__Template.f90:240-278, MultiplyAdd_Command.f90:56-94).
#pragma unroll
        for (int f = 0; f < 5; f++) {
          st(a.out1 + f * a.fs, ic, ar[f] + R(1) * Kc[f]);
          st(a.out2 + f * a.fs, ic, ar[f] + R(0.5) * Kc[f]);
        }
      } else {
        // Solution = (Solution + 1/2 K1) + 1/2 K2, then StoreSolution's
        // conserved fix and the CFL term of the next step in FP64.
        double U[5], V[5];
#pragma unroll
        for (int f = 0; f < 5; f++) U[f] = (double)(ar[f] + R(0.5) * Kc[f]);
        if (a.out2 != nullptr) {
#pragma unroll
          for (int f = 0; f < 5; f++) st(a.out2 + f * a.fs, ic, (R)U[f]);
        }
        fix_conserved(U, ph.mass, ph.nmin, V);
#pragma unroll
        for (int f = 0; f < 5; f++) st(a.out1 + f * a.fs, ic, (R)U[f]);
        const double tc = cfl_term(V, ph.gamma, ph.mass, (double)dX, (double)dY,
                                   (double)dZ);
        tmax = tc > tmax ? tc : tmax;
      }
#pragma unroll
      for (int f = 0; f < 5; f++) {
        Flo[f] = Fhi[f];
        sk[f] = sk1[f];
        Vk[f] = Vk1[f];
        Vk1[f] = Vk2[f];
        if constexpr (UCARRY) {
          rk[f] = rk1[f];
          rk1[f] = zr[f];
        }
        zr[f] = zn[f];
      }
      zflip = znflip;
      flk = fln;
    }
    // No barrier here: the next plane writes only the own centre and the
    // halo cells, which nobody reads after the face barrier, and its first
    // barrier orders the slope and flux buffers.
  }

  if (STAGE == 2) {
    // Block max of the non-negative CFL terms; IEEE order of non-negative
    // doubles is the order of their bit patterns.
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
      const double v = __shfl_xor(tmax, o, WAVE);
      tmax = v > tmax ? v : tmax;
    }
    if ((tid % WAVE) == 0) sRed[tid / WAVE] = tmax;
    __syncthreads();
    if (tid == 0) {
      double m = 0.0;
      for (int w = 0; w < NT / WAVE; w++) m = sRed[w] > m ? sRed[w] : m;
      atomicMax(a.invdt_bits, (unsigned long long)__double_as_longlong(m));
    }
  }
}

template <typename R, int STAGE, bool REC = false, bool BRICK = false,
          typename T = Tile16x16, bool P2 = false>
__global__ __launch_bounds__(NT, (StageOccupancy<R, !BRICK>::waves)) void stage_kernel(
    StageArgs<R> a) {
  __shared__ StageShared<R, T> L;
  const int tile = tile_of_block(blockIdx.x, a.nblk);
  stage_tile<R, STAGE, RECON_LINEAR, REC, BRICK, T, P2>(a, (tile % a.nbx) * T::X,
                                                    ((tile / a.nbx) % a.nby) * T::Y,
                                                    (tile / (a.nbx * a.nby)) * a.tz, L);
}

// The same stage with PARABOLIC reconstruction.
template <typename R, int STAGE, bool REC = false>
__global__ __launch_bounds__(NT, StageOccupancy<R>::waves) void stage_kernel_parabolic(
    StageArgs<R> a) {
  __shared__ StageShared<R> L;
  __shared__ EdgeShared<R> E;
  const int tile = tile_of_block(blockIdx.x, a.nblk);
  stage_tile<R, STAGE, RECON_PARABOLIC, REC>(a, (tile % a.nbx) * TX,
                                             ((tile / a.nbx) % a.nby) * TY,
                                             (tile / (a.nbx * a.nby)) * a.tz, L, &E);
}

// Max CFL term over the current state (the first step after an upload, or
// whenever the stage-2 value is not available), in FP64: the stored values
// (of either precision) are widened and converted to primitives in double,
// as the stage-2 epilogue does with the values it stores.
template <typename R>
__global__ __launch_bounds__(256) void cfl_kernel(StageArgs<R> a) {
  __shared__ double sRed[256 / WAVE];
  const long long ntot = (long long)a.n[0] * a.n[1] * a.n[2];
  double tmax = 0.0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < ntot;
       i += (long long)gridDim.x * 256) {
    const int x = (int)(i % a.n[0]);
    const int y = (int)((i / a.n[0]) % a.n[1]);
    const int z = (int)(i / ((long long)a.n[0] * a.n[1]));
    R raw[5];
    int flip;
    load_raw(a, x, y, z, raw, &flip);  // a proper cell: no ghost rule
    double U[5], V[5];
#pragma unroll
    for (int f = 0; f < 5; f++) U[f] = (double)raw[f];
    if (a.src_is_prim) {
#pragma unroll
      for (int f = 0; f < 5; f++) V[f] = U[f];
    } else {
      prim_from_cons<double>(U, a.ph.mass, a.ph.nmin, V);
    }
    const double tc = cfl_term(V, a.ph.gamma, a.ph.mass, (double)geo(a, 0, G_WIDTH, x + 2),
                               (double)geo(a, 1, G_WIDTH, y + 2),
                               (double)geo(a, 2, G_WIDTH, z + 2));
    tmax = tc > tmax ? tc : tmax;
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    const double v = __shfl_xor(tmax, o, WAVE);
    tmax = v > tmax ? v : tmax;
  }
  if ((threadIdx.x % WAVE) == 0) sRed[threadIdx.x / WAVE] = tmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < 256 / WAVE; w++) m = sRed[w] > m ? sRed[w] : m;
    atomicMax(a.invdt_bits, (unsigned long long)__double_as_longlong(m));
  }
}

// Shock flags of the next step from the stored state, in one pass per (x, y)
// tile marching along z.  DetectShocks_CSL_Kernel (F/FluidFeatures_P__Kernel.
// f90:12-186) tests each cell against its lower neighbour in every dimension
// i -- |dP|/max(min(P), sqrt(tiny)) > ShockThreshold and dV_i <= 0 -- and a
// shock marks the DiffusiveFlux flags of the transverse faces on both sides
// of the two cells it separates (:137-147); ClearBoundary_CSL_Kernel
// (:189-313) then clears the first and last two layers of every dimension.
// So flags live on cells 1 .. n-2 and read shock tests of chart cells only:
// flag_d(c) = OR over i != d of test_i at c, c + e_i, c - e_d, c + e_i - e_d.
// Per plane k: P and V of the tile with two cells of halo below and one
// above go to LDS (ping-pong with plane k-1 for the z test); the tests of
// the tile with a one-cell halo go to a ring of three planes; the flags of
// plane k-1 follow from the ring.
constexpr int QW = TX + 3, QH = TY + 3;  // primitives of x0-2 .. x0+TX, y0-2 .. y0+TY
constexpr int BW = TX + 2, BH = TY + 2;  // tests of x0-1 .. x0+TX, y0-1 .. y0+TY
constexpr int FLAG_TZ = 16;              // planes per block (at 256^3: 4096 blocks,
                                         // 2.7 rounds of 1536; 32 left a third of a
                                         // round, 8 or 64 were slower)

// Default planes per stage block (z-chunk): 32, halved (down to 4) while a
// launch would hold fewer than 512 blocks -- two per CU of 256, what the FP64
// stage kernel keeps resident -- so small charts still fill the chip.
inline int stage_tz(int ntile_xy, int nz) {
  int tz = 32;
  while (tz > 4 && (long long)ntile_xy * ((nz + tz - 1) / tz) < 512) tz /= 2;
  return tz;
}

// Planes per single-chart shock-flag block: FLAG_TZ, halved (down to 4) while
// the launch would hold fewer than 1536 blocks -- six per CU, what its LDS
// allows -- so small charts still fill the chip.
inline int flag_tz(int ntile_xy, int nz) {
  int tz = FLAG_TZ;
  while (tz > 4 && (long long)ntile_xy * ((nz + tz - 1) / tz) < 1536) tz /= 2;
  return tz;
}

template <typename R> struct FlagShared {
  R sQ[2][4][QH * QW];     // P, V_1, V_2, V_3 of planes k and k-1
  uint8_t sB[3][BH * BW];  // tests of plane p at (p + 3) % 3
};

// Bricks: whether the data of cell (x, y, z) is at hand -- a proper cell, or
// a ghost (up to two cells out, edges and corners included) of BC_NEIGHBOR
// faces only -- and its index (the halo of the first such face, whose block
// covers the other ghost coordinates).  Ghosts of the other faces (outflow,
// reflecting, or periodic within the brick) are not: every flag they could
// set is cleared at that face of the chart.
template <typename R>
__device__ __forceinline__ bool brick_cell(const StageArgs<R> &a, int x, int y, int z,
                                           unsigned *idx) {
  const int c[3] = {x, y, z};
  int face = -1;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (c[d] < 0) {
      if (c[d] < -2 || a.bc[2 * d] != BC_NEIGHBOR) return false;
      if (face < 0) face = 2 * d;
    } else if (c[d] >= a.n[d]) {
      if (c[d] > a.n[d] + 1 || a.bc[2 * d + 1] != BC_NEIGHBOR) return false;
      if (face < 0) face = 2 * d + 1;
    }
  }
  *idx = face < 0 ? lin(a.n, x, y, z) : halo_index(a, face, x, y, z);
  return true;
}

// BRICK: the tests also run where the lower neighbour is a halo cell, and
// ClearBoundary clears only at the faces of the chart (FluidFeatures_P__
// Form.f90:333-338: iaBrick == 1 or nBricks, periodic or not), so the flags
// of a brick's first and last layers survive at faces between bricks.
template <typename R, bool BRICK = false>
__device__ __forceinline__ void shock_flags_tile(const StageArgs<R> &a, int x0, int y0,
                                                 int z0, uint8_t *flags, FlagShared<R> &L) {
  auto &sQ = L.sQ;
  auto &sB = L.sB;
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const int zend = min(z0 + a.tz, nz);
  const int tid = threadIdx.x;
  const R gm1 = R(a.ph.gamma) - R(1), st = R(a.ph.st);
  const R mass = R(a.ph.mass);
  const double nmin = a.ph.nmin;
  const R sqrt_tiny = Arith<R>::sqrt_tiny;
  auto avail = [&](int x, int y, int z) {
    unsigned i;
    return brick_cell(a, x, y, z, &i);
  };

  // Each thread converts the plane-tile cells t = tid + j NT (j < QJ); their
  // raw values for plane k + 1 are loaded while plane k's tests run.
  constexpr int QJ = (QH * QW + NT - 1) / NT;
  R Un[QJ][5];
  bool Uav[QJ];
  auto load_plane = [&](int k) {
#pragma unroll
    for (int j = 0; j < QJ; j++) {
      const int t = tid + j * NT;
      const int x = x0 - 2 + t % QW, y = y0 - 2 + t / QW;
      unsigned i = 0u;
      bool in;
      if constexpr (BRICK) {
        in = t < QH * QW && brick_cell(a, x, y, k, &i);
      } else {
        in = t < QH * QW && k >= 0 && k < nz && (unsigned)x < (unsigned)nx &&
             (unsigned)y < (unsigned)ny;
        i = in ? lin(a.n, x, y, k) : 0u;
      }
      Uav[j] = in;
#pragma unroll
      for (int f = 0; f < 5; f++) Un[j][f] = in ? ld(a.src + f * a.fs, i) : R(0);
    }
  };
  load_plane(z0 - 2);
  for (int k = z0 - 2; k <= zend; ++k) {
    R(*Q)[QH * QW] = sQ[(k + 2) & 1];
    R(*Qm)[QH * QW] = sQ[(k + 3) & 1];
    const bool plane = BRICK || (k >= 0 && k < nz);
    if (plane) {
#pragma unroll
      for (int j = 0; j < QJ; j++) {
        const int t = tid + j * NT;
        if (t < QH * QW) {
          R V[5] = {R(0), R(0), R(0), R(0), R(0)};
          if (Uav[j]) prim_from_cons(Un[j], mass, nmin, V);
          Q[0][t] = gm1 * V[4];
          Q[1][t] = V[1];
          Q[2][t] = V[2];
          Q[3][t] = V[3];
        }
      }
    }
    if (k < zend) load_plane(k + 1);
    __syncthreads();
    if (plane && k >= z0 - 1) {
      uint8_t *B = sB[(k + 3) % 3];
      for (int t = tid; t < BH * BW; t += NT) {
        const int bx = t % BW, by = t / BW;
        const int x = x0 - 1 + bx, y = y0 - 1 + by;
        const int q = (by + 1) * QW + (bx + 1);
        uint8_t b = 0;
        const bool here = BRICK ? avail(x, y, k)
                                : ((unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny);
        if (here) {
          const R P = Q[0][q];
          // lower neighbour in x, y (same plane) and z (plane k-1)
          const int cs[3] = {x, y, k};
          const R Pm[3] = {Q[0][q - 1], Q[0][q - QW], Qm[0][q]};
          const R dV[3] = {Q[1][q] - Q[1][q - 1], Q[2][q] - Q[2][q - QW],
                           Q[3][q] - Qm[3][q]};
#pragma unroll
          for (int d = 0; d < 3; d++) {
            if constexpr (BRICK) {
              if (!avail(x - (d == 0), y - (d == 1), k - (d == 2))) continue;
            } else {
              if (cs[d] == 0) continue;
            }
            const R dP = fabs(P - Pm[d]);
            const R P_Min = fmax(fmin(P, Pm[d]), sqrt_tiny);
            const R dLnP = Arith<R>::div(dP, P_Min);
            if (dLnP > st && dV[d] <= R(0)) b |= (uint8_t)(1u << d);
          }
        }
        B[t] = b;
      }
    }
    __syncthreads();
    const int kf = k - 1;  // flags of plane k - 1
    if (kf >= z0 && kf < zend) {
      const int tx = tid % TX, ty = tid / TX;
      const int x = x0 + tx, y = y0 + ty;
      if (x < nx && y < ny) {
        uint8_t out = 0;
        // the layers ClearBoundary leaves: all but the first and last at a
        // face of the chart (the ghost layers hold no flags here)
        const int cm = BRICK ? a.chart_face : 63;
        const int lx = cm & 1 ? 1 : 0, hx = cm & 2 ? nx - 2 : nx - 1;
        const int ly = cm & 4 ? 1 : 0, hy = cm & 8 ? ny - 2 : ny - 1;
        const int lz = cm & 16 ? 1 : 0, hz = cm & 32 ? nz - 2 : nz - 1;
        if (x >= lx && x <= hx && y >= ly && y <= hy && kf >= lz && kf <= hz) {
          const uint8_t *Bm = sB[(kf + 2) % 3], *B0 = sB[(kf + 3) % 3],
                        *Bp = sB[(kf + 4) % 3];
          const int i = (ty + 1) * BW + (tx + 1);
          const unsigned fx = ((B0[i] | B0[i + BW] | B0[i - 1] | B0[i + BW - 1]) & 2u) |
                              ((B0[i] | Bp[i] | B0[i - 1] | Bp[i - 1]) & 4u);
          const unsigned fy = ((B0[i] | B0[i + 1] | B0[i - BW] | B0[i + 1 - BW]) & 1u) |
                              ((B0[i] | Bp[i] | B0[i - BW] | Bp[i - BW]) & 4u);
          const unsigned fz = ((B0[i] | B0[i + 1] | Bm[i] | Bm[i + 1]) & 1u) |
                              ((B0[i] | B0[i + BW] | Bm[i] | Bm[i + BW]) & 2u);
          out = (uint8_t)((fx ? 1u : 0u) | (fy ? 2u : 0u) | (fz ? 4u : 0u));
        }
        flags[lin(a.n, x, y, kf)] = out;
      }
    }
  }
}

template <typename R, bool BRICK = false>
__global__ __launch_bounds__(NT) void shock_flags_kernel(StageArgs<R> a, uint8_t *flags) {
  __shared__ FlagShared<R> L;
  const int tile = tile_of_block(blockIdx.x, a.nblk);
  shock_flags_tile<R, BRICK>(a, (tile % a.nbx) * TX, ((tile / a.nbx) % a.nby) * TY,
                             (tile / (a.nbx * a.nby)) * a.tz, flags, L);
}

// The 23 Current fields for download, proper cells, in FP64: those of
// ComputeFromConservedCommon (F/Fluid_P_I__Form.f90:604-722) on the stored
// conserved state, or, while the Current still holds uploaded primitives
// (prim != null), those of ComputeFromPrimitiveCommon (:509-601) with the
// conserved fields the upload derived from them.
template <typename R>
__global__ __launch_bounds__(256) void current_kernel(StageArgs<R> a, double c_v,
                                                      const R *prim, double *out) {
  const long long ntot = (long long)a.n[0] * a.n[1] * a.n[2];
  const double sqrt_huge = 1.3407807929942596e+154;  // sqrt(DBL_MAX)
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < ntot;
       i += (long long)gridDim.x * 256) {
    double U[5], V[5];
#pragma unroll
    for (int f = 0; f < 5; f++) U[f] = (double)a.aux[f * a.fs + i];
    if (prim) {
#pragma unroll
      for (int f = 0; f < 5; f++) V[f] = (double)prim[f * a.fs + i];
    } else {
      fix_conserved(U, a.ph.mass, a.ph.nmin, V);
    }
    double *o = out;
    const double g = a.ph.gamma, M = a.ph.mass;
    const double N = V[0], E = V[4];
    const double P = (g - 1.0) * E;
    double T, SB, CS;
    if (N > 0.0 && P > 0.0) {
      T = E / (c_v * N);
      SB = c_v * log(P / 1.0 * pow(1.0 / N, g));
      CS = sqrt(g * P / (M * N));
    } else {
      T = 0.0;
      SB = -sqrt_huge;
      CS = 0.0;
    }
    double MN = 0.0;
    if (CS > 0.0) MN = sqrt((V[1] * V[1] + V[2] * V[2]) + V[3] * V[3]) / CS;
    const double vals[GX_N_FIELDS] = {
        V[1] + CS, V[2] + CS, V[3] + CS, V[1] - CS, V[2] - CS, V[3] - CS,
        N,         U[0],      M,         V[1],      V[2],      V[3],
        U[1],      U[2],      U[3],      E,         U[4],      P,
        T,         SB,        SB * N,    CS,        MN};
#pragma unroll
    for (int f = 0; f < GX_N_FIELDS; f++) o[(long long)f * ntot + i] = vals[f];
  }
}

// ComputeFromInitial = ComputeFromPrimitiveCommon on the Current
// (F/Fluid_P_I__Form.f90:509-601): the density floor N<0 -> 0 is written back
// to the primitive field, then D, S, G.
template <typename R>
__global__ __launch_bounds__(256) void from_primitive_kernel(long long ntot, long long fs,
                                                             double mass_d, R *P, R *U) {
  const R mass = R(mass_d);
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < ntot;
       i += (long long)gridDim.x * 256) {
    R N = P[i];
    if (N < R(0)) {
      N = R(0);
      P[i] = N;
    }
    const R V1 = P[fs + i], V2 = P[2 * fs + i], V3 = P[3 * fs + i];
    const R S1 = mass * N * V1, S2 = mass * N * V2, S3 = mass * N * V3;
    U[i] = N;
    U[fs + i] = S1;
    U[2 * fs + i] = S2;
    U[3 * fs + i] = S3;
    U[4 * fs + i] = P[4 * fs + i] + R(0.5) * ((S1 * V1 + S2 * V2) + S3 * V3);
  }
}

// Per-block partial sums of U * volume over proper cells, in FP64.
template <typename R>
__global__ __launch_bounds__(256) void totals_kernel(StageArgs<R> a,
                                                     double *partial) {
  __shared__ double s[5][256];
  const long long ntot = (long long)a.n[0] * a.n[1] * a.n[2];
  double acc[5] = {0, 0, 0, 0, 0};
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < ntot;
       i += (long long)gridDim.x * 256) {
    const int x = (int)(i % a.n[0]);
    const int y = (int)((i / a.n[0]) % a.n[1]);
    const int z = (int)(i / ((long long)a.n[0] * a.n[1]));
    const double vol = (double)geo(a, 0, G_WIDTH, x + 2) *
                       (double)geo(a, 1, G_WIDTH, y + 2) *
                       (double)geo(a, 2, G_WIDTH, z + 2);
#pragma unroll
    for (int f = 0; f < 5; f++) acc[f] += (double)a.aux[f * a.fs + i] * vol;
  }
#pragma unroll
  for (int f = 0; f < 5; f++) s[f][threadIdx.x] = acc[f];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int f = 0; f < 5; f++) s[f][threadIdx.x] += s[f][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 5) partial[blockIdx.x * 5 + threadIdx.x] = s[threadIdx.x][0];
}

}  // namespace gx
