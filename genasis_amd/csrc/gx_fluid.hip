// gx_fluid.hip -- MI355X (gfx950) finite-volume step for the GenASiS
// polytropic ideal-gas fluid, behind the C-ABI of include/gx_fluid.h.
//
// Device layout: the proper cells only (no ghost storage), five conserved
// fields D, S_1, S_2, S_3, G as separate FP64 slabs of n0*n1*n2 values, x
// fastest.  Ghost values are produced on load by index mapping: periodic wrap,
// or the mirror copy of OUTFLOW / REFLECTING (CH/SingleLevelCharts/
// Chart_SL__Template.f90:486-538, Chart_SL__Kernel.f90:12-113).  This equals
// the reference's ghost handling for every value its sweeps read: the swept
// dimension's ghosts over the transverse interior.
//
// One RK2 step = two launches of the fused stage kernel (reconstruction,
// fluxes and divergence of all three dimensions, plus the RK combination)
// and two small launches that build the shock flags of the next step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/gx_fluid.h"
#include "gx_kernels.hpp"

namespace gx {

constexpr int TX = 32;           // tile width in x (cells)
constexpr int TY = 8;            // tile height in y
constexpr int NT = TX * TY;      // threads per block
constexpr int HX = TX + 4;       // plane tile with two ghost layers
constexpr int HY = TY + 4;
constexpr int SXW = TX + 2;      // x-slope row length
constexpr int FXW = TX + 1;      // x-flux row length
constexpr int WAVE = 64;
constexpr int NXCD = 8;          // accelerator complex dies, one L2 each

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
  const R *aux;         // stage 1: U (Solution); stage 2: W = U + K1/2
  R *out1;              // stage 1: Y = U + K1; stage 2: U_new
  R *out2;              // stage 1: W
  const R *geo;
  const uint8_t *flags; // DiffusiveFlux bits of the step
  unsigned long long *invdt_bits;  // stage 2: max CFL term (as FP64 bits)
};

// Waves per SIMD the stage kernel is compiled for: the FP32 tile uses half
// the registers and LDS of the FP64 one.
template <typename R> struct StageOccupancy { static constexpr int waves = 2; };
template <> struct StageOccupancy<float> { static constexpr int waves = 3; };

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

// Linear index of cell (x, y, z) with at most one coordinate outside the
// chart by at most two cells, after the ghost rule of that face; *flip = the
// dimension whose velocity a reflecting face negates, or -1.
template <typename A>
__device__ __forceinline__ unsigned cell_index(const A &a, int x, int y, int z,
                                               int *flip) {
  *flip = -1;
  if ((unsigned)x < (unsigned)a.n[0] && (unsigned)y < (unsigned)a.n[1] &&
      (unsigned)z < (unsigned)a.n[2])
    return lin(a.n, x, y, z);
  bool fx = false, fy = false, fz = false;
  const int cx = map_coord(x, a.n[0], a.bc[0], a.bc[1], &fx);
  const int cy = map_coord(y, a.n[1], a.bc[2], a.bc[3], &fy);
  const int cz = map_coord(z, a.n[2], a.bc[4], a.bc[5], &fz);
  *flip = fx ? 0 : (fy ? 1 : (fz ? 2 : -1));
  return lin(a.n, cx, cy, cz);
}

// Raw stage input (conserved, or primitive after an upload) of one cell.
template <typename R>
__device__ __forceinline__ void load_raw(const StageArgs<R> &a, int x, int y, int z,
                                         R raw[5], int *flip) {
  const unsigned i = cell_index(a, x, y, z, flip);
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
    prim_from_cons(raw, R(a.ph.mass), R(a.ph.nmin), V);
  }
  V[1] = negate_if(V[1], flip == 0);
  V[2] = negate_if(V[2], flip == 1);
  V[3] = negate_if(V[3], flip == 2);
}

template <typename R>
__device__ __forceinline__ void load_prim(const StageArgs<R> &a, int x, int y,
                                          int z, R V[5]) {
  R raw[5];
  int flip;
  load_raw(a, x, y, z, raw, &flip);
  raw_to_prim(a, raw, flip, V);
}

template <typename R>
__device__ __forceinline__ R slope_at(const StageArgs<R> &a, int d, int c, R vm,
                                      R v0, R vp) {
  const R *g = a.geo + d * N_GEO * a.gs + (c + 2);
  const int s = a.gs;
  return slope(vm, v0, vp, g[G_DXL * s], g[G_DXR * s], g[G_DXL2 * s], g[G_DXR2 * s],
               g[G_DEN * s], g[G_RDXL * s], g[G_RDXR * s], g[G_RDEN * s],
               R(a.ph.theta), a.ph.use_limiter);
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
// local index c - origin + 2 for c in [origin-2, origin+TX+2).
template <typename R> using TileGeo = R[2][N_GEO][HX];

template <typename R>
__device__ __forceinline__ R slope_lds(const TileGeo<R> &sG, int d, int l, R vm,
                                       R v0, R vp, R theta, int use_limiter) {
  return slope(vm, v0, vp, sG[d][G_DXL][l], sG[d][G_DXR][l], sG[d][G_DXL2][l],
               sG[d][G_DXR2][l], sG[d][G_DEN][l], sG[d][G_RDXL][l],
               sG[d][G_RDXR][l], sG[d][G_RDEN][l], theta, use_limiter);
}

// x-face i (between cells x0+i-1 and x0+i) of row j into sFx.
template <typename R, typename SP, typename SS, typename SF>
__device__ __forceinline__ void x_face(const Physics &ph, SP sP, SS sSx, SF sFx,
                                       const TileGeo<R> &sG, int i, int j, bool df) {
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
template <typename R, typename SP, typename SS, typename SF>
__device__ __forceinline__ void y_face(const Physics &ph, SP sP, SS sSy, SF sFy,
                                       const TileGeo<R> &sG, int i, int j, bool df) {
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
template <typename R, int STAGE>
__global__ __launch_bounds__(NT, StageOccupancy<R>::waves) void stage_kernel(
    StageArgs<R> a) {
  __shared__ R sP[5][HY * HX];
  __shared__ R sSx[5][TY * SXW];       // x-slopes of cells x0-1 .. x0+tw
  __shared__ R sSy[5][(TY + 2) * TX];  // y-slopes of cells y0-1 .. y0+th
  __shared__ R sFx[5][TY * FXW];       // x-face fluxes x0 .. x0+tw
  __shared__ R sFy[5][(TY + 1) * TX];  // y-face fluxes y0 .. y0+th
  __shared__ double sRed[NT / WAVE];
  __shared__ TileGeo<R> sG;

  const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const int tile = tile_of_block(blockIdx.x, a.nblk);
  const int x0 = (tile % a.nbx) * TX, y0 = ((tile / a.nbx) % a.nby) * TY,
            z0 = (tile / (a.nbx * a.nby)) * a.tz;
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
  R zr[5], hr[5];  // prefetched raw values: z-stencil k+2, halo of k
  int zflip = -1, hflip = -1, hpos = 0;
  unsigned flk = 0;     // flag byte of the own cell at plane k (then k+1)
  double tmax = 0.0;

  if (active) {
    R Vm2[5], Vm1[5], sm1[5], WL[5], WR[5];
    load_prim(a, x, y, z0 - 2, Vm2);
    load_prim(a, x, y, z0 - 1, Vm1);
    load_prim(a, x, y, z0, Vk);
    load_prim(a, x, y, z0 + 1, Vk1);
    load_raw(a, x, y, z0 + 2, zr, &zflip);
    flk = ld(a.flags, lin(a.n, x, y, z0));
#pragma unroll
    for (int f = 0; f < 5; f++) {
      sm1[f] = slope_at(a, 2, z0 - 1, Vm2[f], Vm1[f], Vk[f]);
      sk[f] = slope_at(a, 2, z0, Vm1[f], Vk[f], Vk1[f]);
      WL[f] = Vm1[f] + geo(a, 2, G_WR, z0 - 1 + 2) * sm1[f];
      WR[f] = Vk[f] - geo(a, 2, G_WL, z0 + 2) * sk[f];
    }
    face_flux(WL, WR, 2, (flk >> 2) & 1, ph, Flo);
  }
  if (tid < nhalo) {
    int hx, hy;
    halo_cell(tid, x0, y0, tw, th, &hx, &hy);
    hpos = (hy - y0 + 2) * HX + (hx - x0 + 2);
    load_raw(a, hx, hy, z0, hr, &hflip);
  }

  for (int k = z0; k < zend; ++k) {
    const bool more = k + 1 < zend;
    R ar[5], zn[5];
    int znflip = -1;
    unsigned fln = 0, flx = 0;  // own flags at k+1; flag of this thread's edge face
    if (active) {
#pragma unroll
      for (int f = 0; f < 5; f++) sP[f][ctr] = Vk[f];
    }
    if (tid < nhalo) {
      R V[5];
      raw_to_prim(a, hr, hflip, V);
#pragma unroll
      for (int f = 0; f < 5; f++) sP[f][hpos] = V[f];
      if (more) {
        int hx, hy;
        halo_cell(tid, x0, y0, tw, th, &hx, &hy);
        load_raw(a, hx, hy, k + 1, hr, &hflip);
      }
    }
    if (active) {
      const unsigned ic = lin(a.n, x, y, k);
#pragma unroll
      for (int f = 0; f < 5; f++) ar[f] = ld(a.aux + f * a.fs, ic);
      if (more) load_raw(a, x, y, k + 3, zn, &znflip);
      if (k + 1 < nz) fln = ld(a.flags, lin(a.n, x, y, k + 1));
    }
    if (tid < th) {
      if (x0 + tw < nx) flx = ld(a.flags, lin(a.n, x0 + tw, y0 + tid, k)) & 1u;
    } else if (tid >= WAVE && tid - WAVE < tw) {
      if (y0 + th < ny) flx = (ld(a.flags, lin(a.n, x0 + tid - WAVE, y0 + th, k)) >> 1) & 1u;
    }
    __syncthreads();

    // Slopes: own cell (x and y), then the tile-edge cells.
    if (active) {
      const int bx = ctr;
#pragma unroll
      for (int f = 0; f < 5; f++)
        sSx[f][ty * SXW + tx + 1] = slope_lds(sG, 0, tx + 2, sP[f][bx - 1], sP[f][bx],
                                              sP[f][bx + 1], theta, ph.use_limiter);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int f = 0; f < 5; f++)
        sSy[f][(ty + 1) * TX + tx] = slope_lds(sG, 1, ty + 2, sP[f][bx - HX], sP[f][bx],
                                               sP[f][bx + HX], theta, ph.use_limiter);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (tid < nxs + nys) {
      int base, stride, l, out, gd;
      bool isx = tid < nxs;
      if (isx) {
        const int j = tid >> 1, side = tid & 1;
        const int i = side ? tw + 1 : 0;  // slope index of cell x0-1+i
        base = (j + 2) * HX + (i + 1);
        stride = 1;
        l = i + 1;  // local geometry index of cell x0-1+i
        gd = 0;
        out = j * SXW + i;
      } else {
        const int u = tid - nxs;
        const int i = u >> 1, side = u & 1;
        const int j = side ? th + 1 : 0;  // slope index of cell y0-1+j
        base = (j + 1) * HX + (i + 2);
        stride = HX;
        l = j + 1;
        gd = 1;
        out = j * TX + i;
      }
#pragma unroll
      for (int f = 0; f < 5; f++) {
        const R sv = slope_lds(sG, gd, l, sP[f][base - stride], sP[f][base],
                               sP[f][base + stride], theta, ph.use_limiter);
        if (isx)
          sSx[f][out] = sv;
        else
          sSy[f][out] = sv;
      }
    }
    __syncthreads();

    // Faces: left and bottom faces of the own cell, then the right column of
    // x-faces (wave 0) and the top row of y-faces (wave 1).
    if (active) {
      x_face(ph, sP, sSx, sFx, sG, tx, ty, flk & 1u);
      __builtin_amdgcn_sched_barrier(0);
      y_face(ph, sP, sSy, sFy, sG, tx, ty, (flk >> 1) & 1u);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (tid < th) x_face(ph, sP, sSx, sFx, sG, tw, tid, flx);
    __builtin_amdgcn_sched_barrier(0);
    if (tid >= WAVE && tid - WAVE < tw) y_face(ph, sP, sSy, sFy, sG, tid - WAVE, th, flx);
    __syncthreads();

    if (active) {
      // z stencil of plane k: slope at k+1, flux through the face k+1/2.
      R Vk2[5], sk1[5], Fhi[5];
      {
        R WL[5], WR[5];
        raw_to_prim(a, zr, zflip, Vk2);
#pragma unroll
        for (int f = 0; f < 5; f++) {
          const R vk = sP[f][ctr];
          sk1[f] = slope_at(a, 2, k + 1, vk, Vk1[f], Vk2[f]);
          WL[f] = vk + geo(a, 2, G_WR, k + 2) * sk[f];
          WR[f] = Vk1[f] - geo(a, 2, G_WL, k + 1 + 2) * sk1[f];
        }
        face_flux(WL, WR, 2, (fln >> 2) & 1u, ph, Fhi);
      }
      // ComputeIncrement_CSL_Kernel (EV/Increments/IncrementDivergence_FV__
      // Kernel.f90:672-752), dimensions in order 1, 2, 3, onto K = 0; the
      // division by the volume uses its reciprocal (FP64: rounded exactly).
      const R dX = sG[0][G_WIDTH][tx + 2], dY = sG[1][G_WIDTH][ty + 2],
              dZ = geo(a, 2, G_WIDTH, k + 2);
      const R Vol = dX * dY * dZ;
      const R rVol = Arith<R>::recip(Vol);
      const R A1 = dY * dZ, A2 = dZ * dX, A3 = dX * dY;
      const R dt = R(a.dt);
      const unsigned ic = lin(a.n, x, y, k);
      R Kc[5];
#pragma unroll
      for (int f = 0; f < 5; f++) {
        const R fxl = sFx[f][ty * FXW + tx], fxh = sFx[f][ty * FXW + tx + 1];
        const R fyl = sFy[f][ty * TX + tx], fyh = sFy[f][(ty + 1) * TX + tx];
        R K = R(0);
        K = K + Arith<R>::rdiv(dt * (A1 * fxl - A1 * fxh), Vol, rVol);
        K = K + Arith<R>::rdiv(dt * (A2 * fyl - A2 * fyh), Vol, rVol);
        K = K + Arith<R>::rdiv(dt * (A3 * Flo[f] - A3 * Fhi[f]), Vol, rVol);
        Kc[f] = K;
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

// Max CFL term over the current state (the first step after an upload, or
// whenever the stage-2 value is not available), in FP64.
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
    R Vr[5];
    load_prim(a, x, y, z, Vr);
    double V[5];
#pragma unroll
    for (int f = 0; f < 5; f++) V[f] = (double)Vr[f];
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
constexpr int FLAG_TZ = 32;              // planes per block

template <typename R>
__global__ __launch_bounds__(NT) void shock_flags_kernel(StageArgs<R> a, uint8_t *flags) {
  __shared__ R sQ[2][4][QH * QW];     // P, V_1, V_2, V_3 of planes k and k-1
  __shared__ uint8_t sB[3][BH * BW];  // tests of plane p at (p + 3) % 3
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const int tile = tile_of_block(blockIdx.x, a.nblk);
  const int x0 = (tile % a.nbx) * TX, y0 = ((tile / a.nbx) % a.nby) * TY,
            z0 = (tile / (a.nbx * a.nby)) * a.tz;
  const int zend = min(z0 + a.tz, nz);
  const int tid = threadIdx.x;
  const R gm1 = R(a.ph.gamma) - R(1), st = R(a.ph.st);
  const R mass = R(a.ph.mass), nmin = R(a.ph.nmin);
  const R sqrt_tiny = Arith<R>::sqrt_tiny;

  for (int k = z0 - 2; k <= zend; ++k) {
    R(*Q)[QH * QW] = sQ[(k + 2) & 1];
    R(*Qm)[QH * QW] = sQ[(k + 3) & 1];
    const bool plane = k >= 0 && k < nz;
    if (plane) {
      for (int t = tid; t < QH * QW; t += NT) {
        const int x = x0 - 2 + t % QW, y = y0 - 2 + t / QW;
        R V[5] = {R(0), R(0), R(0), R(0), R(0)};
        if ((unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny) {
          const unsigned i = lin(a.n, x, y, k);
          R U[5];
#pragma unroll
          for (int f = 0; f < 5; f++) U[f] = ld(a.src + f * a.fs, i);
          prim_from_cons(U, mass, nmin, V);
        }
        Q[0][t] = gm1 * V[4];
        Q[1][t] = V[1];
        Q[2][t] = V[2];
        Q[3][t] = V[3];
      }
    }
    __syncthreads();
    if (plane && k >= z0 - 1) {
      uint8_t *B = sB[(k + 3) % 3];
      for (int t = tid; t < BH * BW; t += NT) {
        const int bx = t % BW, by = t / BW;
        const int x = x0 - 1 + bx, y = y0 - 1 + by;
        const int q = (by + 1) * QW + (bx + 1);
        uint8_t b = 0;
        if ((unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny) {
          const R P = Q[0][q];
          // lower neighbour in x, y (same plane) and z (plane k-1)
          const int cs[3] = {x, y, k};
          const R Pm[3] = {Q[0][q - 1], Q[0][q - QW], Qm[0][q]};
          const R dV[3] = {Q[1][q] - Q[1][q - 1], Q[2][q] - Q[2][q - QW],
                           Q[3][q] - Qm[3][q]};
#pragma unroll
          for (int d = 0; d < 3; d++) {
            if (cs[d] == 0) continue;
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
        if (x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2 && kf >= 1 && kf <= nz - 2) {
          const uint8_t *Bm = sB[(kf + 2) % 3], *B0 = sB[(kf + 3) % 3],
                        *Bp = sB[(kf + 4) % 3];
          const int i = (ty + 1) * BW + (tx + 1);
          const unsigned fx = (B0[i] | B0[i + BW] | B0[i - 1] | B0[i + BW - 1]) & 2u |
                              (B0[i] | Bp[i] | B0[i - 1] | Bp[i - 1]) & 4u;
          const unsigned fy = (B0[i] | B0[i + 1] | B0[i - BW] | B0[i + 1 - BW]) & 1u |
                              (B0[i] | Bp[i] | B0[i - BW] | Bp[i - BW]) & 4u;
          const unsigned fz = (B0[i] | B0[i + 1] | Bm[i] | Bm[i + 1]) & 1u |
                              (B0[i] | B0[i + BW] | Bm[i] | Bm[i + BW]) & 2u;
          out = (uint8_t)((fx ? 1u : 0u) | (fy ? 2u : 0u) | (fz ? 4u : 0u));
        }
        flags[lin(a.n, x, y, kf)] = out;
      }
    }
  }
}

// Current fields of ComputeFromConservedCommon for download
// (F/Fluid_P_I__Form.f90:604-722), proper cells, on the stored state, in FP64.
template <typename R>
__global__ __launch_bounds__(256) void current_kernel(StageArgs<R> a, double c_v,
                                                      double *out) {
  const long long ntot = (long long)a.n[0] * a.n[1] * a.n[2];
  const double sqrt_huge = 1.3407807929942596e+154;  // sqrt(DBL_MAX)
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < ntot;
       i += (long long)gridDim.x * 256) {
    double U[5], V[5];
#pragma unroll
    for (int f = 0; f < 5; f++) U[f] = (double)a.aux[f * a.fs + i];
    fix_conserved(U, a.ph.mass, a.ph.nmin, V);
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
__global__ __launch_bounds__(256) void from_primitive_kernel(long long ntot,
                                                             double mass_d, R *P,
                                                             R *U) {
  const R mass = R(mass_d);
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < ntot;
       i += (long long)gridDim.x * 256) {
    R N = P[i];
    if (N < R(0)) {
      N = R(0);
      P[i] = N;
    }
    const R V1 = P[ntot + i], V2 = P[2 * ntot + i], V3 = P[3 * ntot + i];
    const R S1 = mass * N * V1, S2 = mass * N * V2, S3 = mass * N * V3;
    U[i] = N;
    U[ntot + i] = S1;
    U[2 * ntot + i] = S2;
    U[3 * ntot + i] = S3;
    U[4 * ntot + i] = P[4 * ntot + i] + R(0.5) * ((S1 * V1 + S2 * V2) + S3 * V3);
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
    if (__double_as_longlong(q1) != __double_as_longlong(q2) && !both_nan) bad++;
  }
  if (bad) atomicAdd(mismatches, bad);
}

}  // namespace gx

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

using namespace gx;

static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                       \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(GX_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

struct gx_ctx {
  gx_grid grid;
  gx_params params;
  int n[3];
  long long ncell;
  hipStream_t stream = nullptr;
  int tz = 16;
  bool fp32 = false;          // state and sweeps in float (Precision = GX_FP32)
  size_t esz = sizeof(double);  // bytes per state value
  // device buffers, of the state's real type
  void *geo_buf = nullptr;    // 3 dimensions x N_GEO arrays x gs
  int gs = 0;                 // geometry stride: max(n) + 4
  // five-field states, field f at base + f * ncell
  void *U = nullptr, *Y = nullptr, *W = nullptr, *P0 = nullptr;
  uint8_t *flags = nullptr;
  unsigned long long *invdt = nullptr;
  double *partial = nullptr;
  int n_partial = 0;
  double *scratch = nullptr;  // download staging: GX_N_FIELDS x ncell
  bool uploaded = false;
  bool prims_valid = false;   // next stage 1 reads P0
  bool invdt_valid = false;   // invdt holds the max for the current state
  // per-kernel timing (gx_set_profiling): stage 1, stage 2, shock flags, (unused)
  bool profiling = false;
  hipEvent_t ev[5] = {};
  double kernel_ms[4] = {0, 0, 0, 0};
  long long kernel_n[4] = {0, 0, 0, 0};
  std::vector<double> geo_host[3][N_GEO];
};

static void geometry_1d(int n, double lo, double hi, std::vector<double> out[N_GEO]) {
  // Same construction as the reference chart (Chart_Template.f90:463-593),
  // including single-precision ghost widths (real(KDL) at :477-478).
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
    out[0][p] = c[p] - E(ic);
    out[1][p] = E(ic + 1) - c[p];
  }
  for (int p = 0; p < m; p++) {
    if (p >= 1 && p + 1 < m) {
      const double dxl = c[p] - c[p - 1];
      const double dxr = c[p + 1] - c[p];
      out[2][p] = dxl;
      out[3][p] = dxr;
      out[4][p] = dxl * dxl;
      out[5][p] = dxr * dxr;
      out[6][p] = dxl * dxr * (dxl + dxr);
    } else {
      out[2][p] = out[3][p] = out[4][p] = out[5][p] = 1.0;
      out[6][p] = 1.0;
    }
    out[7][p] = out[0][p] + out[1][p];
    out[8][p] = 1.0 / out[2][p];
    out[9][p] = 1.0 / out[3][p];
    out[10][p] = 1.0 / out[6][p];
  }
}

template <typename R>
static StageArgs<R> base_args(const gx_ctx *c) {
  StageArgs<R> a;
  memset(&a, 0, sizeof(a));
  for (int d = 0; d < 3; d++) a.n[d] = c->n[d];
  for (int i = 0; i < 6; i++) a.bc[i] = c->grid.BoundaryCondition[i];
  a.tz = c->tz;
  a.ph.gamma = c->params.AdiabaticIndex;
  a.ph.theta = c->params.LimiterParameter;
  a.ph.mass = c->params.BaryonMassReference;
  a.ph.nmin = c->params.BaryonDensityMin;
  a.ph.st = c->params.ShockThreshold;
  a.ph.use_limiter = c->params.UseLimiter;
  a.ph.hllc = c->params.RiemannSolverType == GX_HLLC;
  a.gs = c->gs;
  a.fs = c->ncell;
  a.geo = (const R *)c->geo_buf;
  a.flags = c->flags;
  a.invdt_bits = c->invdt;
  return a;
}

static int grid_blocks(long long ncell) {
  long long b = (ncell + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

// Runs f(R()) with R the context's real type.
template <typename F>
static int by_precision(const gx_ctx *c, F &&f) {
  return c->fp32 ? f(float()) : f(double());
}

// n host doubles into a device state buffer of the context's real type, and
// back.  Both synchronise the stream.
static int put_values(gx_ctx *c, void *dst, const double *src, size_t n) {
  if (!c->fp32) {
    HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  } else {
    std::vector<float> t(n);
    for (size_t i = 0; i < n; i++) t[i] = (float)src[i];
    HIP_TRY(hipMemcpyAsync(dst, t.data(), n * sizeof(float), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
}

static int get_values(gx_ctx *c, double *dst, const void *src, size_t n) {
  if (!c->fp32) {
    HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> t(n);
    HIP_TRY(hipMemcpyAsync(t.data(), src, n * sizeof(float), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++) dst[i] = (double)t[i];
  }
  return GX_OK;
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
}

const char *gx_last_error(void) { return g_last_error.c_str(); }

int gx_self_test_division(long long n, unsigned long long seed,
                          long long *mismatches) {
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
}

int gx_destroy(gx_ctx *c) {
  if (!c) return GX_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int e = 0; e < 5; e++)
    if (c->ev[e]) (void)hipEventDestroy(c->ev[e]);
  (void)hipFree(c->geo_buf);
  (void)hipFree(c->U);
  (void)hipFree(c->Y);
  (void)hipFree(c->W);
  (void)hipFree(c->P0);
  (void)hipFree(c->flags);
  (void)hipFree(c->invdt);
  (void)hipFree(c->partial);
  (void)hipFree(c->scratch);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GX_OK;
}

int gx_create(const gx_grid *grid, const gx_params *params, gx_ctx **out) {
  if (!grid || !out) return fail(GX_E_ARG, "gx_create: null argument");
  *out = nullptr;
  gx_params p;
  if (params)
    p = *params;
  else
    gx_default_params(&p);
  if (p.Precision != GX_FP64 && p.Precision != GX_FP32)
    return fail(GX_E_ARG, "gx_create: Precision must be GX_FP64 or GX_FP32");
  if (p.RiemannSolverType != GX_HLL && p.RiemannSolverType != GX_HLLC)
    return fail(GX_E_ARG, "gx_create: RiemannSolverType must be GX_HLL or GX_HLLC");
  for (int d = 0; d < 3; d++) {
    if (grid->nCells[d] < 2)
      return fail(GX_E_ARG, "gx_create: nCells must be >= 2 in every dimension");
    const int bi = grid->BoundaryCondition[2 * d], bo = grid->BoundaryCondition[2 * d + 1];
    if (bi < 0 || bi > 2 || bo < 0 || bo > 2)
      return fail(GX_E_ARG, "gx_create: unknown boundary condition");
    if ((bi == GX_BC_PERIODIC) != (bo == GX_BC_PERIODIC))
      return fail(GX_E_ARG, "gx_create: PERIODIC must hold on both faces of a dimension");
    if (!(grid->MaxCoordinate[d] > grid->MinCoordinate[d]))
      return fail(GX_E_ARG, "gx_create: MaxCoordinate must exceed MinCoordinate");
  }
  {
    long long cells = 1;
    for (int d = 0; d < 3; d++) cells *= grid->nCells[d];
    // byte offsets within one field stay below 2^32
    if (cells >= (1ll << 29))
      return fail(GX_E_ARG, "gx_create: at most 2^29 - 1 cells per chart");
  }
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
#define TRY_CLEAN(expr)                                                   \
  do {                                                                    \
    hipError_t e_ = (expr);                                               \
    if (e_ != hipSuccess) {                                               \
      std::string m_ = std::string(#expr ": ") + hipGetErrorString(e_);   \
      gx_destroy(c);                                                      \
      return fail(GX_E_HIP, m_);                                          \
    }                                                                     \
  } while (0)
  c->gs = 0;
  for (int d = 0; d < 3; d++) {
    geometry_1d(c->n[d], grid->MinCoordinate[d], grid->MaxCoordinate[d], c->geo_host[d]);
    c->gs = c->n[d] + 4 > c->gs ? c->n[d] + 4 : c->gs;
  }
  {
    // Divisors of the exactly rounded reciprocal division: centre spacings,
    // slope denominators and cell volumes must lie in [2^-500, 2^122].
    const double lo = 0x1p-500, hi = 0x1p122;
    double vmin = 1.0, vmax = 1.0;
    bool ok = true;
    for (int d = 0; d < 3; d++) {
      const std::vector<double> *g = c->geo_host[d];
      double wmin = HUGE_VAL, wmax = 0.0;
      for (int q = 0; q < c->n[d] + 4; q++) {
        for (int k : {G_DXL, G_DXR, G_DEN})
          ok = ok && g[k][q] >= lo && g[k][q] <= hi;
        if (q >= 2 && q < c->n[d] + 2) {
          wmin = fmin(wmin, g[G_WIDTH][q]);
          wmax = fmax(wmax, g[G_WIDTH][q]);
        }
      }
      vmin *= wmin;
      vmax *= wmax;
    }
    if (!ok || !(vmin >= lo) || !(vmax <= hi)) {
      gx_destroy(c);
      return fail(GX_E_ARG, "gx_create: cell spacings and volumes must lie in [2^-500, 2^122]");
    }
  }
  TRY_CLEAN(hipStreamCreate(&c->stream));
  const size_t geo_len = 3 * (size_t)N_GEO * c->gs;
  TRY_CLEAN(hipMalloc(&c->geo_buf, geo_len * c->esz));
  {
    std::vector<double> host(geo_len, 1.0);
    for (int d = 0; d < 3; d++)
      for (int a = 0; a < N_GEO; a++)
        memcpy(&host[((size_t)d * N_GEO + a) * c->gs], c->geo_host[d][a].data(),
               sizeof(double) * (c->n[d] + 4));
    if (!c->fp32) {
      TRY_CLEAN(hipMemcpy(c->geo_buf, host.data(), geo_len * sizeof(double),
                          hipMemcpyHostToDevice));
    } else {
      // Rounded geometry, with the reciprocals of the rounded values.
      std::vector<float> h32(geo_len);
      for (size_t i = 0; i < geo_len; i++) h32[i] = (float)host[i];
      const int recip[3][2] = {{G_RDXL, G_DXL}, {G_RDXR, G_DXR}, {G_RDEN, G_DEN}};
      for (int d = 0; d < 3; d++)
        for (int r = 0; r < 3; r++)
          for (int q = 0; q < c->gs; q++)
            h32[((size_t)d * N_GEO + recip[r][0]) * c->gs + q] =
                1.0f / h32[((size_t)d * N_GEO + recip[r][1]) * c->gs + q];
      TRY_CLEAN(hipMemcpy(c->geo_buf, h32.data(), geo_len * sizeof(float),
                          hipMemcpyHostToDevice));
    }
  }
  const size_t bytes = c->esz * (size_t)c->ncell;
  TRY_CLEAN(hipMalloc(&c->U, 5 * bytes));
  TRY_CLEAN(hipMalloc(&c->Y, 5 * bytes));
  TRY_CLEAN(hipMalloc(&c->W, 5 * bytes));
  TRY_CLEAN(hipMalloc(&c->P0, 5 * bytes));
  TRY_CLEAN(hipMalloc(&c->flags, (size_t)c->ncell));
  TRY_CLEAN(hipMemset(c->flags, 0, (size_t)c->ncell));
  TRY_CLEAN(hipMalloc(&c->invdt, sizeof(unsigned long long)));
  c->n_partial = grid_blocks(c->ncell);
  if (c->n_partial > 1024) c->n_partial = 1024;
  TRY_CLEAN(hipMalloc(&c->partial, sizeof(double) * 5 * c->n_partial));
#undef TRY_CLEAN
  *out = c;
  return GX_OK;
}

int gx_set_tuning(gx_ctx *c, int zChunk) {
  if (!c || zChunk < 1) return fail(GX_E_ARG, "gx_set_tuning: bad argument");
  c->tz = zChunk;
  return GX_OK;
}

void *gx_stream(gx_ctx *c) { return c ? (void *)c->stream : nullptr; }

int gx_set_profiling(gx_ctx *c, int on) {
  if (!c) return fail(GX_E_ARG, "gx_set_profiling: null context");
  if (on && !c->ev[0])
    for (int e = 0; e < 5; e++) HIP_TRY(hipEventCreate(&c->ev[e]));
  c->profiling = on != 0;
  for (int k = 0; k < 4; k++) {
    c->kernel_ms[k] = 0.0;
    c->kernel_n[k] = 0;
  }
  return GX_OK;
}

int gx_kernel_times(gx_ctx *c, double ms[4], long long count[4]) {
  if (!c || !ms || !count) return fail(GX_E_ARG, "gx_kernel_times: null argument");
  for (int k = 0; k < 4; k++) {
    ms[k] = c->kernel_ms[k];
    count[k] = c->kernel_n[k];
  }
  return GX_OK;
}

int gx_set_state(gx_ctx *c, const double *primitive, const double *conserved) {
  if (!c || !primitive || !conserved) return fail(GX_E_ARG, "gx_set_state: null argument");
  const size_t n = 5 * (size_t)c->ncell;
  int rc = put_values(c, c->P0, primitive, n);
  if (rc) return rc;
  rc = put_values(c, c->U, conserved, n);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(c->flags, 0, (size_t)c->ncell, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->uploaded = true;
  c->prims_valid = true;
  c->invdt_valid = false;
  return GX_OK;
}

int gx_set_primitive(gx_ctx *c, const double *primitive) {
  if (!c || !primitive) return fail(GX_E_ARG, "gx_set_primitive: null argument");
  const int rc = put_values(c, c->P0, primitive, 5 * (size_t)c->ncell);
  if (rc) return rc;
  by_precision(c, [&](auto tag) {
    using R = decltype(tag);
    from_primitive_kernel<R><<<grid_blocks(c->ncell), 256, 0, c->stream>>>(
        c->ncell, c->params.BaryonMassReference, (R *)c->P0, (R *)c->U);
    return 0;
  });
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemsetAsync(c->flags, 0, (size_t)c->ncell, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->uploaded = true;
  c->prims_valid = true;
  c->invdt_valid = false;
  return GX_OK;
}

static const int PRIM_FIELDS[5] = {GX_COMOVING_BARYON_DENSITY, GX_VELOCITY_U_1,
                                   GX_VELOCITY_U_2, GX_VELOCITY_U_3,
                                   GX_INTERNAL_ENERGY};
static const int CONS_FIELDS[5] = {GX_CONSERVED_BARYON_DENSITY,
                                   GX_MOMENTUM_DENSITY_D_1,
                                   GX_MOMENTUM_DENSITY_D_2,
                                   GX_MOMENTUM_DENSITY_D_3,
                                   GX_CONSERVED_ENERGY};

int gx_upload(gx_ctx *c, const double *value, long long ld, const int *map) {
  if (!c || !value || !map) return fail(GX_E_ARG, "gx_upload: null argument");
  const int m0 = c->n[0] + 4, m1 = c->n[1] + 4, m2 = c->n[2] + 4;
  if (ld < (long long)m0 * m1 * m2) return fail(GX_E_ARG, "gx_upload: ldValue < nValues");
  for (int f = 0; f < 5; f++)
    if (map[PRIM_FIELDS[f]] <= 0 || map[CONS_FIELDS[f]] <= 0)
      return fail(GX_E_ARG, "gx_upload: primitive and conserved fields are required");
  std::vector<double> prim(5 * (size_t)c->ncell), cons(5 * (size_t)c->ncell);
  for (int f = 0; f < 5; f++) {
    const double *vp = value + (size_t)(map[PRIM_FIELDS[f]] - 1) * ld;
    const double *vc = value + (size_t)(map[CONS_FIELDS[f]] - 1) * ld;
    for (int z = 0; z < c->n[2]; z++)
      for (int y = 0; y < c->n[1]; y++)
        for (int x = 0; x < c->n[0]; x++) {
          const size_t src = (size_t)(x + 2) + (size_t)m0 * ((y + 2) + (size_t)m1 * (z + 2));
          const size_t dst = (size_t)f * c->ncell + x + (size_t)c->n[0] * (y + (size_t)c->n[1] * z);
          prim[dst] = vp[src];
          cons[dst] = vc[src];
        }
  }
  return gx_set_state(c, prim.data(), cons.data());
}

int gx_get_conserved(gx_ctx *c, double *conserved) {
  if (!c || !conserved) return fail(GX_E_ARG, "gx_get_conserved: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_get_conserved: no state uploaded");
  return get_values(c, conserved, c->U, 5 * (size_t)c->ncell);
}

static int run_current(gx_ctx *c) {
  if (!c->scratch)
    HIP_TRY(hipMalloc(&c->scratch, sizeof(double) * GX_N_FIELDS * (size_t)c->ncell));
  // c_v = k/(mu(gamma-1)) with the default gamma (Fluid_P_I__Form.f90:206-223)
  const double c_v = 1.0 / (1.0 * (1.4 - 1.0));
  by_precision(c, [&](auto tag) {
    using R = decltype(tag);
    StageArgs<R> a = base_args<R>(c);
    a.aux = (const R *)c->U;
    current_kernel<R><<<grid_blocks(c->ncell), 256, 0, c->stream>>>(a, c_v, c->scratch);
    return 0;
  });
  HIP_TRY(hipGetLastError());
  return GX_OK;
}

int gx_get_primitive(gx_ctx *c, double *primitive) {
  if (!c || !primitive) return fail(GX_E_ARG, "gx_get_primitive: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_get_primitive: no state uploaded");
  const size_t bytes = sizeof(double) * (size_t)c->ncell;
  if (c->prims_valid) {
    return get_values(c, primitive, c->P0, 5 * (size_t)c->ncell);
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
}

int gx_download(gx_ctx *c, double *value, long long ld, const int *map) {
  if (!c || !value || !map) return fail(GX_E_ARG, "gx_download: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_download: no state uploaded");
  const int m0 = c->n[0] + 4, m1 = c->n[1] + 4, m2 = c->n[2] + 4;
  if (ld < (long long)m0 * m1 * m2) return fail(GX_E_ARG, "gx_download: ldValue < nValues");
  int rc = run_current(c);
  if (rc) return rc;
  std::vector<double> host(GX_N_FIELDS * (size_t)c->ncell);
  HIP_TRY(hipMemcpyAsync(host.data(), c->scratch, sizeof(double) * host.size(),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->prims_valid) {
    // The Current still holds the uploaded primitives.
    std::vector<double> p(5 * (size_t)c->ncell);
    rc = get_values(c, p.data(), c->P0, p.size());
    if (rc) return rc;
    for (int f = 0; f < 5; f++)
      memcpy(host.data() + (size_t)PRIM_FIELDS[f] * c->ncell,
             p.data() + (size_t)f * c->ncell, sizeof(double) * c->ncell);
  }
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
}

int gx_get_flags(gx_ctx *c, unsigned char *flags) {
  if (!c || !flags) return fail(GX_E_ARG, "gx_get_flags: null argument");
  HIP_TRY(hipMemcpyAsync(flags, c->flags, (size_t)c->ncell, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
}

int gx_set_flags(gx_ctx *c, const unsigned char *flags) {
  if (!c || !flags) return fail(GX_E_ARG, "gx_set_flags: null argument");
  HIP_TRY(hipMemcpyAsync(c->flags, flags, (size_t)c->ncell, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
}

int gx_upload_features(gx_ctx *c, const double *df, long long ld) {
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
}

int gx_download_features(gx_ctx *c, double *df, long long ld) {
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
}

int gx_compute_time_step(gx_ctx *c, double *candidate) {
  if (!c || !candidate) return fail(GX_E_ARG, "gx_compute_time_step: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_compute_time_step: no state uploaded");
  if (!c->invdt_valid) {
    HIP_TRY(hipMemsetAsync(c->invdt, 0, sizeof(unsigned long long), c->stream));
    by_precision(c, [&](auto tag) {
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
  HIP_TRY(hipMemcpyAsync(&bits, c->invdt, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  double invdt;
  memcpy(&invdt, &bits, sizeof(invdt));
  c->invdt_valid = true;
  // ComputeTimeStepKernel_CSL tail and ComputeTimeStep_C_ASC
  const double tsi = fmax(DBL_MIN, invdt);
  const double ts = fmin(DBL_MAX, 1.0 / tsi);
  *candidate = c->params.CourantFactor * ts;
  return GX_OK;
}

int gx_step_rk2(gx_ctx *c, double time, double dt) {
  (void)time;
  if (!c) return fail(GX_E_ARG, "gx_step_rk2: null context");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_step_rk2: no state uploaded");
  const dim3 block(NT);
  const int nbx = (c->n[0] + TX - 1) / TX, nby = (c->n[1] + TY - 1) / TY,
            nbz = (c->n[2] + c->tz - 1) / c->tz;
  const int nblk = nbx * nby * nbz;
  const dim3 grid(nblk);
  if (c->profiling) HIP_TRY(hipEventRecord(c->ev[0], c->stream));
  const int rc = by_precision(c, [&](auto tag) {
    using R = decltype(tag);
    // Stage 1: Current = uploaded primitives or the Solution.
    StageArgs<R> a1 = base_args<R>(c);
    a1.dt = dt;
    a1.src_is_prim = c->prims_valid ? 1 : 0;
    a1.src = (const R *)(c->prims_valid ? c->P0 : c->U);
    a1.aux = (const R *)c->U;
    a1.out1 = (R *)c->Y;
    a1.out2 = (R *)c->W;
    a1.nbx = nbx;
    a1.nby = nby;
    a1.nblk = nblk;
    stage_kernel<R, 1><<<grid, block, 0, c->stream>>>(a1);
    HIP_TRY(hipGetLastError());
    // Stage 2: Current = Y.
    HIP_TRY(hipMemsetAsync(c->invdt, 0, sizeof(unsigned long long), c->stream));
    if (c->profiling) HIP_TRY(hipEventRecord(c->ev[1], c->stream));
    StageArgs<R> a2 = base_args<R>(c);
    a2.dt = dt;
    a2.src = (const R *)c->Y;
    a2.aux = (const R *)c->W;
    a2.out1 = (R *)c->U;
    a2.nbx = nbx;
    a2.nby = nby;
    a2.nblk = nblk;
    stage_kernel<R, 2><<<grid, block, 0, c->stream>>>(a2);
    HIP_TRY(hipGetLastError());
    if (c->profiling) HIP_TRY(hipEventRecord(c->ev[2], c->stream));
    // Shock flags of the next step from the new state.
    StageArgs<R> a3 = base_args<R>(c);
    a3.src = (const R *)c->U;
    a3.tz = FLAG_TZ;
    a3.nbx = nbx;
    a3.nby = nby;
    a3.nblk = nbx * nby * ((c->n[2] + FLAG_TZ - 1) / FLAG_TZ);
    shock_flags_kernel<R><<<a3.nblk, NT, 0, c->stream>>>(a3, c->flags);
    HIP_TRY(hipGetLastError());
    return GX_OK;
  });
  if (rc) return rc;
  if (c->profiling) {
    // The stage-1 interval ends after the invdt reset; the reset is a
    // few-byte fill and is counted with stage 1.
    HIP_TRY(hipEventRecord(c->ev[3], c->stream));
    HIP_TRY(hipEventSynchronize(c->ev[3]));
    for (int k = 0; k < 3; k++) {
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]));
      c->kernel_ms[k] += ms;
      c->kernel_n[k] += 1;
    }
  }
  c->prims_valid = false;
  c->invdt_valid = true;
  return GX_OK;
}

int gx_sync(gx_ctx *c) {
  if (!c) return fail(GX_E_ARG, "gx_sync: null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
}

int gx_totals(gx_ctx *c, double out[5]) {
  if (!c || !out) return fail(GX_E_ARG, "gx_totals: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_totals: no state uploaded");
  by_precision(c, [&](auto tag) {
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
}

}  // extern "C"
