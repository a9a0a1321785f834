// gx_adjoint.hip -- discrete adjoint of the FP64 step of one chart
// (gx_adjoint_* of include/gx_fluid.h).
//
// The forward run records the time steps, the shock flags of every step and
// the state every `every` steps (a segment), and keeps the last two
// segments' per-step states.  The reverse pass walks the segments backwards,
// each segment's states (the step inputs U, the stage values Y and
// W = U + K1/2, and Z = W + K2/2) recomputed from its checkpoint with the
// forward stage kernels on a second stream while the reverse pass of the
// segment after it runs, and applies per step the transposed step:
//
//   U' = fix(Z), Z = W + K2/2, K2 = dt R(Y), Y = U + K1, W = U + K1/2,
//   K1 = dt R(U or the uploaded primitives), dt = dt(CFL maximum of U),
//
// where R(S) = -div F / Vol of the stage input S.  The transpose of R is a
// primitives pass onto the padded layout (n + 4 per dimension, ghosts by the
// ghost rule), one line pass per dimension (slopes, the face flux's reverse
// sweep and the limiter's reverse along lines, neighbours by DPP lane
// shifts), a fold of the ghost bars onto the cells they copy, and the
// conserved -> primitive reverse.  The fold of a stage, with the RK and
// conserved-fix reverses that follow it, is done by the x pass of the stage
// transposed next, before its sweep, so that its memory traffic runs beside
// the sweep's arithmetic.  The recorded shock flags are replayed as
// constants; dt and gamma bars are reduced per block, then in two levels, in
// a fixed order.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/gx_fluid.h"
#include "gx_adjoint.hpp"
#include "gx_context.hpp"
#include "gx_host.hpp"
#include "gx_kernels.hpp"
#include "gx_tiles.hpp"

namespace gx {

constexpr int AB = 256;  // threads per block of the adjoint kernels

// The reverse sweeps' work arrays hold the five values of a cell together,
// so that a y or z line pass, whose lanes walk a line across rows or planes,
// touches one or two cache lines per cell and array instead of five: the
// primitives and their bars on the padded layout in 40-byte records (two
// 16-byte and one 8-byte access, 8-byte aligned), the stage bars K1b, K2b in
// 48-byte records whose sixth slot carries the cell's shock-flag byte.
constexpr int AOS = 6;   // stage bars
constexpr int PAD5 = 5;  // padded primitives and bars

struct __attribute__((aligned(8))) Pair8 {  // two doubles, 8-byte aligned
  double x, y;
};

__device__ __forceinline__ void loadp(const double *base, long long pos, double v[5]) {
  const double *q = base + PAD5 * pos;
  const Pair8 a = reinterpret_cast<const Pair8 *>(q)[0];
  const Pair8 b = reinterpret_cast<const Pair8 *>(q + 2)[0];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
  v[4] = q[4];
}

__device__ __forceinline__ void storep(double *base, long long pos, const double v[5]) {
  double *q = base + PAD5 * pos;
  reinterpret_cast<Pair8 *>(q)[0] = Pair8{v[0], v[1]};
  reinterpret_cast<Pair8 *>(q + 2)[0] = Pair8{v[2], v[3]};
  q[4] = v[4];
}

__device__ __forceinline__ void load5(const double *base, long long cell, double v[5]) {
  const double2 *p = reinterpret_cast<const double2 *>(base + AOS * cell);
  const double2 a = p[0], b = p[1], c = p[2];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
  v[4] = c.x;
}

// A stage bar K of a cell with the cell's shock-flag byte in the sixth slot,
// so that the y and z passes, whose lanes read cells on 64 rows or planes,
// take the flags with K's third 16-byte access.
__device__ __forceinline__ void store5_flags(double *base, long long cell, const double v[5],
                                             unsigned flags) {
  double2 *p = reinterpret_cast<double2 *>(base + AOS * cell);
  p[0] = make_double2(v[0], v[1]);
  p[1] = make_double2(v[2], v[3]);
  p[2] = make_double2(v[4], __longlong_as_double((long long)flags));
}

__device__ __forceinline__ unsigned load5_flags(const double *base, long long cell,
                                                double v[5]) {
  const double2 *p = reinterpret_cast<const double2 *>(base + AOS * cell);
  const double2 a = p[0], b = p[1], c = p[2];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
  v[4] = c.x;
  return (unsigned)__double_as_longlong(c.y);
}

struct PadGeom {
  int n[3];
  int m[3];
  long long np;
  long long s[3];  // padded strides
};

__host__ __device__ inline long long plin(const PadGeom &g, int x, int y, int z) {
  return (long long)(x + 2) + g.m[0] * ((long long)(y + 2) + (long long)g.m[1] * (z + 2));
}

// The states of one segment, per step (every steps): the input U^n (slot 0
// unused: the segment's checkpoint is its first input), the stage values Y
// and W, and the state before the fix Z.
struct SegmentBuffers {
  double *U = nullptr, *Y = nullptr, *W = nullptr, *Z = nullptr;
  int held = -1;                 // first step of the segment held, or -1
  hipEvent_t ready = nullptr;    // recorded when its recomputation is done
};

struct AdjointState {
  int nsteps = 0, every = 0;
  std::vector<double> dt, ts, invdt, ramp;
  std::vector<double *> ckpt;
  // Segment s lives in sb[s % 2]: the reverse pass of one segment runs on the
  // context's stream while the next (earlier) one is recomputed into the
  // other set on `side`.
  SegmentBuffers sb[2];
  hipStream_t side = nullptr;
  hipEvent_t mark = nullptr;
  uint8_t *stepF = nullptr;    // the shock flags each step used, nsteps ncell
  double *trash = nullptr;
  double *Vp = nullptr;               // primitives of a stage input, PAD5 np
  // bars of stage 2 and stage 1 (PAD5 np), one array per line pass: the x
  // pass of one stage folds in the other's three
  double *Vb[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  double *Ub = nullptr, *P0b = nullptr;    // 5 ncell, field slabs
  double *K1b = nullptr, *K2b = nullptr;   // AOS ncell, with the flag byte
  bool seed_ready = false;     // Ub holds dJ/dU_final (gx_adjoint_kinetic_energy)
  bool bar_ready = false;      // P0b holds the last reverse pass's result
  double *partials = nullptr;  // 2 per line block, both stages, three dimensions
  int part_stride = 0;         // blocks of the largest line pass
  double *partials2 = nullptr; // 2 per block of the first reduction level
  double *gb = nullptr;        // gamma bar
  int *argmax = nullptr;
  double *band = nullptr;      // per y-row weight of the observable, n1
  double *lossp = nullptr;     // per-block loss partials
  PadGeom pg;
  bool forward_done = false;
};

// ---- kernels ---------------------------------------------------------------

__global__ __launch_bounds__(AB) void adj_fill_kernel(double *p, long long n, double v) {
  for (long long i = blockIdx.x * (long long)AB + threadIdx.x; i < n;
       i += (long long)gridDim.x * AB)
    p[i] = v;
}

__device__ inline double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = AB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ inline double cell_width(const StageArgs<double> &a, int e, int c) {
  return geo(a, e, G_WIDTH, c + 2);
}

// ---- fused transposed residual, one wave per line segment -------------------
//
// A wave holds 64 consecutive cells c = c0 - 3 + lane of a line along D
// (cells -2 .. n+1 exist; others are zero).  Neighbour values come by lane
// shuffles: the slope of each cell, the face (c-1 | c) of each lane with its
// flux's reverse sweep, the slope bar s_b(c) = wr plus - wl minus and the
// limiter's reverse, and the cell bar: own face-state bars plus the slope
// partials of cells c-1, c, c+1.  Lanes 3 .. 60 (58 cells) are complete.  It
// is the slope, face and gather passes of the reverse in one step, with the
// face bars never leaving registers.  A face counts into the dt and gamma
// bars in the segment that owns its cell c.
constexpr int LINE_OUT = 58;  // complete cells per wave

// a/b exactly as the forward forms it (div_rcp, y = RN(1/b), or a * y for a
// power-of-two geometry: gdiv), so that the limiter picks the forward's
// branch.
template <bool P2>
__device__ __forceinline__ double quot_rev(double a, double b, double y) {
  return gdiv<P2>(a, b, y);
}

// The limited slope of one cell and its reverse multipliers: s = slope(...)
// as the forward computes it (the quotients div_rcp gives, which also pick
// the limiter's branch), and kl, kr with
// ds = kl d(dVL) + kr d(dVR), dVL = v0 - vm, dVR = vp - v0.  Branch-free:
// every case is formed and selected, so lanes on different limiter branches
// do not serialise.  Multipliers use the cell's geometry factors m (a
// gradient needs no exactly rounded quotient), formed once per cell for the
// five fields: cl = dxr^2/den, cr = dxl^2/den, tl = theta/dxl,
// tr = theta/dxr.  Rules at ties as in the header of gx_adjoint.hpp: the
// centred slope when it equals the smaller one-sided one; where exactly one
// difference is zero, the mean of the two sides (tl/2 or tr/2).
struct SlopeFactors {
  double cl, cr, tl, tr;
};

template <bool P2>
__device__ __forceinline__ double slope_fwd_rev(double vm, double v0, double vp, double dxl,
                                                double dxr, double dxl2, double dxr2,
                                                double den, double rdxl, double rdxr,
                                                double rden, double theta, int use_limiter,
                                                const SlopeFactors &m, double &kl,
                                                double &kr) {
  const double dVL = v0 - vm, dVR = vp - v0;
  const double c = quot_rev<P2>(dxr2 * dVL + dxl2 * dVR, den, rden);
  if (!use_limiter) {  // uniform over the launch
    kl = m.cl;
    kr = m.cr;
    return c;
  }
  const double ql = quot_rev<P2>(theta * dVL, dxl, rdxl);
  const double qr = quot_rev<P2>(theta * dVR, dxr, rdxr);
  const double tl = fabs(ql), tr = fabs(qr), ac = fabs(c);
  // sf = sign(1/2, dVL) + sign(1/2, dVR), in {-1, 0, 1}: products with it
  // are exact
  const double sf = copysign(0.5, dVL) + copysign(0.5, dVR);
  const double s = sf * fmin(fmin(tl, tr), ac);
  const double m1 = tl <= tr ? tl : tr;
  const bool zL = dVL == 0.0, zR = dVR == 0.0;
  const bool cen = ac <= m1, lef = !cen && tl <= tr;
  // d|q|/dq = sign(q): the branch's factor carries the sign of its quotient
  double l = cen ? copysign(m.cl, c) : (lef ? copysign(m.tl, ql) : 0.0);
  double r = cen ? copysign(m.cr, c) : (lef ? 0.0 : copysign(m.tr, qr));
  l *= sf;
  r *= sf;
  if (zL != zR) {  // exactly one difference is zero: one-sided, mean of the sides
    l = zL ? 0.5 * m.tl : 0.0;
    r = zL ? 0.0 : 0.5 * m.tr;
  }
  kl = l;
  kr = r;
  return s;
}

// Neighbour lanes by DPP wavefront shifts (wave_shr:1, wave_shl:1), a VALU
// move instead of an LDS permute: lane_up gives lane i the value of lane
// i - 1, lane_down that of lane i + 1; the end lanes, which have no source
// lane, read 0 (bound_ctrl), so no destination has to be cleared first; they
// are never used (every use is guarded by the lane range).
__device__ __forceinline__ double lane_up(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_down(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

// get_K(K) gives the lane's stage bar K of its cell (0 off the interior); it
// is called after the slopes, so the loads behind it overlap their
// arithmetic.
template <int D, bool P2, typename GetK>
__device__ __forceinline__ void line_vjp(const StageArgs<double> &a, int n, int c, int lane,
                                         bool own, double A, double dt, const double V[5],
                                         GetK get_K, bool df, double volc,
                                         double volm, double Vb[5], double &dtb,
                                         double &gbs, double (*kst)[AB]) {
  double Vm[5], Vq[5], s[5], sm[5], Kbm[5], Kb[5];
#pragma unroll
  for (int f = 0; f < 5; f++) {
    Vm[f] = lane_up(V[f]);
    Vq[f] = lane_down(V[f]);
  }
  const bool has_slope = c >= -1 && c <= n && lane >= 1 && lane <= 62;
  double kl[5], kr[5];  // the limiter's reverse multipliers of cell c
  {
    // positions -2 .. n+1 of the geometry (n + 4 entries, finite and
    // positive); a lane without a slope reads a clamped entry and its slope
    // is cleared below
    const int q = min(max(c + 2, 0), n + 3);
    const double dxl = geo(a, D, G_DXL, q), dxr = geo(a, D, G_DXR, q);
    const double dxl2 = geo(a, D, G_DXL2, q), dxr2 = geo(a, D, G_DXR2, q);
    const double den = geo(a, D, G_DEN, q), rden = geo(a, D, G_RDEN, q);
    const double rdxl = geo(a, D, G_RDXL, q), rdxr = geo(a, D, G_RDXR, q);
    const SlopeFactors m = {dxr2 * rden, dxl2 * rden, a.ph.theta * rdxl, a.ph.theta * rdxr};
#pragma unroll
    for (int f = 0; f < 5; f++) {
      s[f] = slope_fwd_rev<P2>(Vm[f], V[f], Vq[f], dxl, dxr, dxl2, dxr2, den, rdxl, rdxr, rden,
                           a.ph.theta, a.ph.use_limiter, m, kl[f], kr[f]);
      if (!has_slope) s[f] = kl[f] = kr[f] = 0.0;
    }
  }
  // the multipliers wait in LDS while the face flux's reverse runs (ten
  // registers pairs fewer at its peak)
#pragma unroll
  for (int f = 0; f < 5; f++) {
    kst[f][threadIdx.x] = kl[f];
    kst[5 + f][threadIdx.x] = kr[f];
  }
  __asm__ volatile("" ::: "memory");
  get_K(Kb);
#pragma unroll
  for (int f = 0; f < 5; f++) {
    sm[f] = lane_up(s[f]);
    Kbm[f] = lane_up(Kb[f]);
  }
  // face (c-1 | c)
  double WLb[5] = {0, 0, 0, 0, 0}, WRb[5] = {0, 0, 0, 0, 0};
  if (c >= 0 && c <= n && lane >= 2 && lane <= 62) {
    const double wr = geo(a, D, G_WR, c + 1), wl = geo(a, D, G_WL, c + 2);
    double WL[5], WR[5], Gf[5];
#pragma unroll
    for (int f = 0; f < 5; f++) {  // the face states in the forward's order
      WL[f] = Vm[f] + wr * sm[f];
      WR[f] = V[f] - wl * s[f];
    }
    {
      GX_REV_FMA
      const double rc = c < n ? A * adj::rcp_rev(volc) : 0.0;
      const double rm = c >= 1 ? A * adj::rcp_rev(volm) : 0.0;
#pragma unroll
      for (int f = 0; f < 5; f++) Gf[f] = rc * Kb[f] - rm * Kbm[f];
    }
#pragma unroll
    for (int f = 0; f < 5; f++) kst[10 + f][threadIdx.x] = Gf[f];
    __asm__ volatile("" ::: "memory");
    double Fb[5], F[5], gl = 0.0;
#pragma unroll
    for (int f = 0; f < 5; f++) Fb[f] = dt * Gf[f];
    adj::face_flux_vjp(WL, WR, D, df && c < n, a.ph.gamma, a.ph.mass, a.ph.hllc != 0, Fb, WLb,
                       WRb, gl, F);
    __asm__ volatile("" ::: "memory");
    if (own) {
      GX_REV_FMA
#pragma unroll
      for (int f = 0; f < 5; f++) dtb += F[f] * kst[10 + f][threadIdx.x];
      gbs += gl;
    }
  }
  // slope bar of cell c and the limiter's reverse onto c-1, c, c+1
  double dm[5], d0[5], dp[5], up[5];
#pragma unroll
  for (int f = 0; f < 5; f++) {
    up[f] = lane_down(WLb[f]);  // bar of c as the left state of face c+1
    dm[f] = d0[f] = dp[f] = 0.0;
  }
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int f = 0; f < 5; f++) {
    kl[f] = kst[f][threadIdx.x];
    kr[f] = kst[5 + f][threadIdx.x];
  }
  if (has_slope && lane >= 2 && lane <= 61) {
    GX_REV_FMA
    const int q = c + 2;
    const double wr = geo(a, D, G_WR, q), wl = geo(a, D, G_WL, q);
#pragma unroll
    for (int f = 0; f < 5; f++) {
      double sb = 0.0;
      if (c <= n - 1) sb += wr * up[f];
      if (c >= 0) sb -= wl * WRb[f];
      const double dVLb = sb * kl[f], dVRb = sb * kr[f];
      dm[f] = -dVLb;
      d0[f] = dVLb - dVRb;
      dp[f] = dVRb;
    }
  }
#pragma unroll
  for (int f = 0; f < 5; f++) {
    const double from_left = lane_up(dp[f]);    // cell c-1's partial on c
    const double from_right = lane_down(dm[f]); // cell c+1's partial on c
    double acc = 0.0;
    if (c >= -1 && c <= n - 1) acc += up[f];
    if (c >= 0 && c <= n) acc += WRb[f];
    Vb[f] = ((acc + from_left) + d0[f]) + from_right;
  }
}

// Bars of a stage on the padded layout, one array per line pass: pass d
// writes the interior cells and the ghost positions along d.
struct StageBars {
  double *d[3];
};

// The fold of a finished stage: the bar of an interior cell on the padded
// layout (the sum of the three passes' bars) plus the bars of the ghost
// values copied from it (normal velocity negated across a REFLECTING face).
__device__ __forceinline__ void gather_bars(const StageArgs<double> &a, const PadGeom &g,
                                            const StageBars &Vb, const int c3[3], long long p,
                                            double V[5]) {
  {
    double vy[5], vz[5];
    loadp(Vb.d[0], p, V);
    loadp(Vb.d[1], p, vy);
    loadp(Vb.d[2], p, vz);
#pragma unroll
    for (int f = 0; f < 5; f++) V[f] = (V[f] + vy[f]) + vz[f];
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int c = c3[d], n = g.n[d];
    const int bi = a.bc[2 * d], bo = a.bc[2 * d + 1];
    int gh[2], fl[2], ng = 0;
    if (bi == BC_PERIODIC) {
      if (c >= n - 2) { gh[ng] = c - n; fl[ng++] = 0; }
    } else if (c <= 1) {
      gh[ng] = -1 - c;
      fl[ng++] = bi == BC_REFLECTING;
    }
    if (bo == BC_PERIODIC) {
      if (c <= 1) { gh[ng] = c + n; fl[ng++] = 0; }
    } else if (c >= n - 2) {
      gh[ng] = 2 * n - 1 - c;
      fl[ng++] = bo == BC_REFLECTING;
    }
    for (int k = 0; k < ng; k++) {
      const long long pg = p + (long long)(gh[k] - c) * g.s[d];
      double gv[5];
      loadp(Vb.d[d], pg, gv);
#pragma unroll
      for (int f = 0; f < 5; f++) V[f] += (fl[k] && f == 1 + d) ? -gv[f] : gv[f];
    }
  }
}

// What the x pass of a stage's transpose folds in before its sweep, per
// interior cell of its lanes (all lanes form the cell's bar K of their
// stage; the lanes that own the cell store what they change):
//   FOLD_SEED (the first reverse step): K2b = fix^T(Ub)/2 at the step's Z;
//   FOLD_STAGE1 (the stage-1 transpose of the step after is finished):
//     K2b as above of Ub + prim^T(gathered bars) at that step's input;
//   FOLD_STAGE2 (this step's stage-2 transpose is finished): with
//     y = prim^T(gathered bars) at Y, Ub = y + 2 K2b (Y = U + K1,
//     W = U + K1/2, Z = W + K2/2: Yb = y, Wb = 2 K2b) and K1b = y + K2b.
enum { FOLD_SEED = 0, FOLD_STAGE1 = 1, FOLD_STAGE2 = 2 };

struct FoldIn {
  StageBars Vb;       // the finished stage's bars on the padded layout (PAD5)
  const double *src;  // that stage's input (conserved field slabs)
  double *out;        // Ub (field slabs)
  const double *K2b;  // FOLD_STAGE2: the stage-2 bar (AOS)
  const double *Z;    // FOLD_SEED, FOLD_STAGE1: the state before the fix (field slabs)
  double *Kout;       // the bar of the x pass's own stage: K1b or K2b (AOS)
};

// The loads of a cell's fold (issued before the sweep's slopes) and its
// arithmetic and stores (after them).
struct FoldVals {
  double Vg[5], U[5], o[5], w[5];  // w: Z (FOLD_SEED, FOLD_STAGE1) or K2b
};

template <int MODE>
__device__ __forceinline__ void fold_load(const StageArgs<double> &a, const PadGeom &g,
                                          const FoldIn &fi, const int c3[3], long long p,
                                          unsigned ic, FoldVals &v) {
  const long long fs = (long long)g.n[0] * g.n[1] * g.n[2];
  if (MODE != FOLD_SEED) {
    gather_bars(a, g, fi.Vb, c3, p, v.Vg);
#pragma unroll
    for (int f = 0; f < 5; f++) v.U[f] = fi.src[f * fs + ic];
  }
  if (MODE == FOLD_STAGE2) {
    load5(fi.K2b, ic, v.w);
  } else {
#pragma unroll
    for (int f = 0; f < 5; f++) {
      v.o[f] = fi.out[f * fs + ic];
      v.w[f] = fi.Z[f * fs + ic];
    }
  }
}

template <int MODE>
__device__ __forceinline__ void fold_finish(const StageArgs<double> &a, const PadGeom &g,
                                            const FoldIn &fi, unsigned ic, bool own,
                                            unsigned flags, FoldVals &v, double K[5]) {
  const long long fs = (long long)g.n[0] * g.n[1] * g.n[2];
  if (MODE != FOLD_SEED) {
    double yv[5] = {0, 0, 0, 0, 0};
    adj::prim_vjp(v.U, a.ph.mass, a.ph.nmin, v.Vg, yv);
    if (MODE == FOLD_STAGE2) {
#pragma unroll
      for (int f = 0; f < 5; f++) {
        v.o[f] = yv[f] + 2.0 * v.w[f];
        K[f] = yv[f] + v.w[f];
      }
      if (own) {
#pragma unroll
        for (int f = 0; f < 5; f++) fi.out[f * fs + ic] = v.o[f];
        store5_flags(fi.Kout, ic, K, flags);
      }
      return;
    }
    // (Ub itself is not stored: the stage-2 fold of this step replaces it)
#pragma unroll
    for (int f = 0; f < 5; f++) v.o[f] += yv[f];
  }
  adj::fix_vjp(v.w, a.ph.mass, a.ph.nmin, v.o, K);
#pragma unroll
  for (int f = 0; f < 5; f++) K[f] *= 0.5;
  if (own) store5_flags(fi.Kout, ic, K, flags);
}

// D = 0: the rows (y, z) of the padded layout, cells -2 .. n+1 each, are
// consecutive in memory; waves take 58 positions each of that sequence
// (lanes along x, coalesced), so a wave may hold the end of one row and the
// start of the next.  Every use of a neighbour lane is guarded by the own
// row's cell range, and the lanes of another row contribute zeros there.
// The lanes fold in the finished stage: fold_load issues its loads before
// the slopes and fold_finish (through get_K) uses them after, so that the
// fold's memory traffic overlaps the sweep's arithmetic.
template <int MODE, bool P2>
__global__ __launch_bounds__(AB) void adj_line_x_kernel(StageArgs<double> a, PadGeom g,
                                                        const double *Vp, FoldIn fi,
                                                        double dt, double *Vb,
                                                        double *partial) {
  __shared__ double sh[AB];
  __shared__ double kst[15][AB];
  const int n = g.n[0];
  const int L = n + 4;
  const long long T = (long long)L * g.n[1] * g.n[2];
  const int lane = threadIdx.x % WAVE;
  const long long w = blockIdx.x * (long long)(AB / WAVE) + threadIdx.x / WAVE;
  const long long p = w * LINE_OUT - 3 + lane;
  double dtb = 0.0, gbs = 0.0;
  if (w * LINE_OUT < T) {  // uniform over the wave
    const bool valid = p >= 0 && p < T;
    const long long r = valid ? p / L : 0;
    const int c = valid ? (int)(p - r * L) - 2 : -100;  // -100: in no cell range
    const int y = (int)(r % g.n[1]), z = (int)(r / g.n[1]);
    const long long pl = plin(g, 0, y, z);
    double V[5];
    bool df = false;
    double volc = 0.0, volm = 0.0;
    const double wy = cell_width(a, 1, y), wz = cell_width(a, 2, z);
    const double A = wy * wz;
    loadp(Vp, pl + (valid ? c : 0), V);
#pragma unroll
    for (int f = 0; f < 5; f++) V[f] = valid ? V[f] : 0.0;
    const bool own = valid && lane >= 3 && lane <= 60;
    const bool interior = c >= 0 && c < n;
    const unsigned ic = interior ? lin(a.n, c, y, z) : 0u;
    FoldVals fv;
    unsigned fl = 0;
    if (interior) {
      const int c3[3] = {c, y, z};
      fold_load<MODE>(a, g, fi, c3, pl + c, ic, fv);
      fl = a.flags[ic];
      df = fl & 1;
      volc = cell_width(a, 0, c) * wy * wz;
    }
    if (c >= 1 && c <= n) volm = cell_width(a, 0, c - 1) * wy * wz;
    double out[5];
    auto get_K = [&](double K[5]) {
#pragma unroll
      for (int f = 0; f < 5; f++) K[f] = 0.0;
      if (interior) fold_finish<MODE>(a, g, fi, ic, own, fl, fv, K);
    };
    line_vjp<0, P2>(a, n, c, lane, own, A, dt, V, get_K, df, volc, volm, out, dtb, gbs, kst);
    if (own) storep(Vb, pl + c, out);  // this pass's own array
  }
  const double s1 = block_sum(dtb, sh);
  const double s2 = block_sum(gbs, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s1;
    partial[2 * blockIdx.x + 1] = s2;
  }
}

// D = 1, 2 without staging: wave (x, segment) holds 64 consecutive positions
// of the sequence of lines along D (cells -2 .. n+1 of line o, then of line
// o+1, ...) in column x; the four waves of a block take four consecutive x
// at the same positions, so each strided load of a wave shares its cache
// lines with the other three.  The pass writes its bars (interior cells and
// the ghost positions along D) to an array of its own, which the fold adds
// to the other passes', so it reads no earlier bars.
constexpr int DXW = AB / WAVE;  // x columns per block
template <int D, bool P2>
__global__ __launch_bounds__(AB) void adj_line_t_kernel(StageArgs<double> a, PadGeom g,
                                                        const double *Vp, const double *Kb,
                                                        double dt, double *Vb,
                                                        double *partial) {
  constexpr int O = D == 1 ? 2 : 1;  // the other transverse dimension
  __shared__ double sh[AB];
  __shared__ double kst[15][AB];
  const int n = g.n[D];
  const int L = n + 4;
  const int T = L * g.n[O];
  const int nxb = (g.n[0] + DXW - 1) / DXW;
  const int b = tile_of_block(blockIdx.x, gridDim.x);
  const int xb = b % nxb, sgi = b / nxb;
  const int lane = threadIdx.x % WAVE;
  const int x = xb * DXW + threadIdx.x / WAVE;
  const int p = sgi * LINE_OUT - 3 + lane;
  double dtb = 0.0, gbs = 0.0;
  if (x < g.n[0]) {  // uniform over the wave
    const bool valid = p >= 0 && p < T;
    const int o = valid ? p / L : 0;
    const int c = valid ? p - o * L - 2 : -100;  // -100: in no cell range
    int q[3];
    q[0] = x;
    q[D] = valid ? c : 0;
    q[O] = o;
    const long long pv = plin(g, q[0], q[1], q[2]);
    double V[5], K[5];
    bool df = false;
    const bool interior = c >= 0 && c < n;
    const bool own = valid && lane >= 3 && lane <= 60;
    // pv is in the padded array for every lane (q[D] = 0 off the sequence)
    loadp(Vp, pv, V);
#pragma unroll
    for (int f = 0; f < 5; f++) {
      V[f] = valid ? V[f] : 0.0;
      K[f] = 0.0;
    }
    if (interior) {
      const unsigned ik = lin(a.n, q[0], q[1], q[2]);
      df = (load5_flags(Kb, ik, K) >> D) & 1;  // the flags the x pass stored
    }
    const double wo = cell_width(a, O, o);
    const double wx = cell_width(a, 0, x);
    const double A = D == 1 ? wo * wx : wx * wo;  // D=1: wz*wx; D=2: wx*wy
    auto vol_of = [&](int cc) {
      const double wd = cell_width(a, D, cc);
      return D == 1 ? (wx * wd) * wo : (wx * wo) * wd;
    };
    const double volc = (c >= 0 && c < n) ? vol_of(c) : 0.0;
    const double volm = (c >= 1 && c <= n) ? vol_of(c - 1) : 0.0;
    double out[5];
    auto get_K = [&](double Kc[5]) {
#pragma unroll
      for (int f = 0; f < 5; f++) Kc[f] = K[f];
    };
    line_vjp<D, P2>(a, n, c, lane, own, A, dt, V, get_K, df, volc, volm, out, dtb, gbs, kst);
    if (own) storep(Vb, pv, out);  // this pass's own array
  }
  const double s1 = block_sum(dtb, sh);
  const double s2 = block_sum(gbs, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s1;
    partial[2 * blockIdx.x + 1] = s2;
  }
}

// The fold of the last stage transpose (stage 1 of the first step): out +=
// the gathered bars of each cell, through the conserved -> primitive
// reverse unless the stage input is the uploaded primitives.
__global__ __launch_bounds__(AB) void adj_fold_kernel(StageArgs<double> a, PadGeom g,
                                                      StageBars Vb, double *out) {
  const long long ncell = (long long)g.n[0] * g.n[1] * g.n[2];
  for (long long t = blockIdx.x * (long long)AB + threadIdx.x; t < ncell;
       t += (long long)gridDim.x * AB) {
    const int c3[3] = {(int)(t % g.n[0]), (int)((t / g.n[0]) % g.n[1]),
                       (int)(t / ((long long)g.n[0] * g.n[1]))};
    double V[5], y[5] = {0, 0, 0, 0, 0};
    gather_bars(a, g, Vb, c3, plin(g, c3[0], c3[1], c3[2]), V);
    if (a.src_is_prim) {
#pragma unroll
      for (int f = 0; f < 5; f++) y[f] = V[f];
    } else {
      double U[5];
#pragma unroll
      for (int f = 0; f < 5; f++) U[f] = a.src[f * a.fs + t];
      adj::prim_vjp(U, a.ph.mass, a.ph.nmin, V, y);
    }
#pragma unroll
    for (int f = 0; f < 5; f++) out[f * ncell + t] += y[f];
  }
}

// First level of the dt and gamma bar sums: block b adds the pairs of its
// contiguous chunk of the per-block partials (fixed order).
constexpr int PART_BLOCKS = 128;

__global__ __launch_bounds__(AB) void adj_partsum_kernel(const double *partial, int nparts,
                                                         double *out) {
  __shared__ double sh[AB];
  const int chunk = (nparts + PART_BLOCKS - 1) / PART_BLOCKS;
  const int i0 = blockIdx.x * chunk, i1 = min(nparts, i0 + chunk);
  double sd = 0.0, sg = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += AB) {
    sd += partial[2 * i];
    sg += partial[2 * i + 1];
  }
  const double a = block_sum(sd, sh);
  const double b = block_sum(sg, sh);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = a;
    out[2 * blockIdx.x + 1] = b;
  }
}

// The dt bar of one step: sum of the per-block partials (fixed order), the
// CFL maximum's reverse at its cell (dt = ts = ramp * C / invdt), added to
// out through the conserved -> primitive reverse; gamma bars into *gbar.
// The next x pass folds the stage-1 bars into out after it, so the cell's
// conserved-fix reverse sees this term too.  One block: the dt kernel of the
// first step, and for the other steps the last block of the next stage
// transpose's primitives pass, where its latency hides behind that pass.
struct DtArgs {
  StageArgs<double> a;  // the step's stage-1 input
  const double *partial;
  int nparts;
  double ts, invdt;
  int *idx;
  double *out, *gbar;
};

__device__ __forceinline__ void dt_work(const DtArgs &d, double *sh) {
  const StageArgs<double> &a = d.a;
  double sd = 0.0, sg = 0.0;
  for (int i = threadIdx.x; i < d.nparts; i += AB) {
    sd += d.partial[2 * i];
    sg += d.partial[2 * i + 1];
  }
  const double dtb = block_sum(sd, sh);
  const double gsum = block_sum(sg, sh);
  if (threadIdx.x != 0) return;
  double gb = gsum;
  const long long ncell = (long long)a.n[0] * a.n[1] * a.n[2];
  const int t = d.idx[0];
  if (d.invdt > DBL_MIN && d.ts > 0.0 && !(t >= 0 && t < ncell)) {
    d.idx[1] += 1;  // no cell reproduced the maximum: gx_adjoint_backward fails
  } else if (d.invdt > DBL_MIN && d.ts > 0.0) {
    const double invb = dtb * (-d.ts / d.invdt);
    const int x = t % a.n[0], y = (t / a.n[0]) % a.n[1], z = t / (a.n[0] * a.n[1]);
    double V[5];
    load_prim(a, x, y, z, V);
    const double w[3] = {geo(a, 0, G_WIDTH, x + 2), geo(a, 1, G_WIDTH, y + 2),
                         geo(a, 2, G_WIDTH, z + 2)};
    double Vb[5] = {0, 0, 0, 0, 0};
    adj::cfl_term_vjp(V, a.ph.gamma, a.ph.mass, w, invb, Vb, gb);
    if (a.src_is_prim) {
      for (int f = 0; f < 5; f++) d.out[f * ncell + t] += Vb[f];
    } else {
      double U[5], Ub[5] = {0, 0, 0, 0, 0};
      for (int f = 0; f < 5; f++) U[f] = a.src[f * a.fs + t];
      adj::prim_vjp(U, a.ph.mass, a.ph.nmin, Vb, Ub);
      for (int f = 0; f < 5; f++) d.out[f * ncell + t] += Ub[f];
    }
  }
  *d.gbar += gb;
  d.idx[0] = 0x7fffffff;  // ready for the next step's search
}

__global__ __launch_bounds__(AB) void adj_dt_kernel(DtArgs d) {
  __shared__ double sh[AB];
  dt_work(d, sh);
}

// Primitives of the stage input on the padded layout, ghosts by the ghost
// rule (cells with two or more coordinates outside are unused and 0).  With
// idx given (stage 1), it also finds the first cell (x fastest) whose CFL
// term has the bits of the step's maximum (atomicMin; *idx starts at
// INT_MAX): the term is formed from the same primitives as the forward's.
// With with_dt, its last block does the dt work of the step reversed before.
__global__ __launch_bounds__(AB) void adj_prims_kernel(StageArgs<double> a, PadGeom g,
                                                       double *Vp, unsigned long long bits,
                                                       int *idx, DtArgs dtw, int with_dt) {
  const int nb = gridDim.x - with_dt;
  if ((int)blockIdx.x >= nb) {
    __shared__ double sh[AB];
    dt_work(dtw, sh);
    return;
  }
  for (long long t = blockIdx.x * (long long)AB + threadIdx.x; t < g.np;
       t += (long long)nb * AB) {
    const int x = (int)(t % g.m[0]) - 2;
    const int y = (int)((t / g.m[0]) % g.m[1]) - 2;
    const int z = (int)(t / ((long long)g.m[0] * g.m[1])) - 2;
    const int out = (x < 0 || x >= g.n[0]) + (y < 0 || y >= g.n[1]) + (z < 0 || z >= g.n[2]);
    double V[5] = {0, 0, 0, 0, 0};
    if (out <= 1) load_prim(a, x, y, z, V);
    storep(Vp, t, V);
    if (idx && out == 0) {
      const double tc = cfl_term(V, a.ph.gamma, a.ph.mass, geo(a, 0, G_WIDTH, x + 2),
                                 geo(a, 1, G_WIDTH, y + 2), geo(a, 2, G_WIDTH, z + 2));
      if ((unsigned long long)__double_as_longlong(tc) == bits)
        atomicMin(idx, (int)lin(a.n, x, y, z));
    }
  }
}

// ComputeFromInitial's reverse: D = N, S_k = (M N) V_k, G = E + S.V/2 with
// N >= 0 (P holds the floored primitives); Pb += (dU/dP)^T Ub.
__global__ __launch_bounds__(AB) void adj_from_prim_kernel(long long ncell, double mass,
                                                           const double *P, const double *Ub,
                                                           double *Pb) {
  for (long long t = blockIdx.x * (long long)AB + threadIdx.x; t < ncell;
       t += (long long)gridDim.x * AB) {
    const double N = P[t], V[3] = {P[ncell + t], P[2 * ncell + t], P[3 * ncell + t]};
    double ub[5];
    for (int f = 0; f < 5; f++) ub[f] = Ub[f * ncell + t];
    double S[3], Sb[3], Vb[3];
    for (int k = 0; k < 3; k++) S[k] = mass * N * V[k];
    // G = E + 1/2 ((S1 V1 + S2 V2) + S3 V3)
    const double Eb = ub[4];
    for (int k = 0; k < 3; k++) {
      Sb[k] = ub[1 + k] + 0.5 * ub[4] * V[k];
      Vb[k] = 0.5 * ub[4] * S[k];
    }
    double Nb = ub[0];
    for (int k = 0; k < 3; k++) {
      Nb += Sb[k] * mass * V[k];
      Vb[k] += Sb[k] * mass * N;
    }
    Pb[t] += Nb;
    for (int k = 0; k < 3; k++) Pb[(1 + k) * ncell + t] += Vb[k];
    Pb[4 * ncell + t] += Eb;
  }
}

// The observable sum_c band(y_c) * 1/2 N V_k^2 dV on the state U (per-block
// partials) and its bar with respect to U.
__global__ __launch_bounds__(AB) void adj_loss_kernel(StageArgs<double> a, int comp,
                                                      const double *band, double *loss_part,
                                                      double *Ub) {
  __shared__ double sh[AB];
  const long long ncell = (long long)a.n[0] * a.n[1] * a.n[2];
  double acc = 0.0;
  for (long long t = blockIdx.x * (long long)AB + threadIdx.x; t < ncell;
       t += (long long)gridDim.x * AB) {
    const int x = (int)(t % a.n[0]), y = (int)((t / a.n[0]) % a.n[1]),
              z = (int)(t / ((long long)a.n[0] * a.n[1]));
    double U[5], V[5];
    for (int f = 0; f < 5; f++) U[f] = a.src[f * a.fs + t];
    prim_from_cons(U, a.ph.mass, a.ph.nmin, V);
    const double vol = geo(a, 0, G_WIDTH, x + 2) * geo(a, 1, G_WIDTH, y + 2) *
                       geo(a, 2, G_WIDTH, z + 2);
    const double w = band[y];
    const double v = V[1 + comp];
    acc += w * (0.5 * V[0] * v * v) * vol;
    double Vb[5] = {0, 0, 0, 0, 0}, ub[5] = {0, 0, 0, 0, 0};
    Vb[0] = w * (0.5 * v * v) * vol;
    Vb[1 + comp] = w * (V[0] * v) * vol;
    adj::prim_vjp(U, a.ph.mass, a.ph.nmin, Vb, ub);
    for (int f = 0; f < 5; f++) Ub[f * ncell + t] = ub[f];
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) loss_part[blockIdx.x] = s;
}

}  // namespace gx

using namespace gx;

namespace gx {

void release_adjoint(gx_ctx *c) {
  AdjointState *A = c->adj;
  if (!A) return;
  for (double *p : A->ckpt) (void)hipFree(p);
  if (A->side) (void)hipStreamSynchronize(A->side);
  for (SegmentBuffers &b : A->sb) {
    for (void *p : {(void *)b.U, (void *)b.Y, (void *)b.W, (void *)b.Z}) (void)hipFree(p);
    if (b.ready) (void)hipEventDestroy(b.ready);
  }
  if (A->mark) (void)hipEventDestroy(A->mark);
  if (A->side) (void)hipStreamDestroy(A->side);
  for (int s = 0; s < 2; s++)
    for (int d = 0; d < 3; d++) (void)hipFree(A->Vb[s][d]);
  for (void *p : {(void *)A->stepF, (void *)A->trash, (void *)A->Vp,
                  (void *)A->Ub, (void *)A->K1b, (void *)A->K2b,
                  (void *)A->P0b, (void *)A->partials, (void *)A->partials2, (void *)A->gb, (void *)A->argmax,
                  (void *)A->band, (void *)A->lossp})
    (void)hipFree(p);
  delete A;
  c->adj = nullptr;
}

}  // namespace gx

static int blocks_for(long long n) {
  long long b = (n + AB - 1) / AB;
  if (b > 65536) b = 65536;
  return (int)std::max(1ll, b);
}

static int fill(gx_ctx *c, double *p, long long n, double v) {
  adj_fill_kernel<<<blocks_for(n), AB, 0, c->stream>>>(p, n, v);
  HIP_TRY(hipGetLastError());
  return GX_OK;
}

// Forward pieces, as gx_step_rk2 launches them.
static void stage_geometry(const gx_ctx *c, StageArgs<double> &a) {
  a.nbx = (c->n[0] + TX - 1) / TX;
  a.nby = (c->n[1] + TY - 1) / TY;
  a.nblk = a.nbx * a.nby * ((c->n[2] + c->tz - 1) / c->tz);
}

// One stage-1 launch: out1 = aux + K(src), out2 = aux + K(src)/2.
static int launch_stage1(gx_ctx *c, hipStream_t s, const double *src, bool prim,
                         const double *aux, double *out1, double *out2, double dt,
                         const uint8_t *flags) {
  StageArgs<double> a = base_args<double>(c);
  a.dt = dt;
  a.src_is_prim = prim ? 1 : 0;
  a.src = src;
  a.aux = aux;
  a.out1 = out1;
  a.out2 = out2;
  a.flags = flags;
  stage_geometry(c, a);
  if (c->geom.pow2)
    stage_kernel<double, 1, false, false, Tile16x16, true><<<a.nblk, NT, 0, s>>>(a);
  else
    stage_kernel<double, 1><<<a.nblk, NT, 0, s>>>(a);
  HIP_TRY(hipGetLastError());
  return GX_OK;
}

static int launch_stage2(gx_ctx *c, hipStream_t s, const double *Y, const double *W,
                         double *out, double *Z, double dt, const uint8_t *flags) {
  StageArgs<double> a = base_args<double>(c);
  a.dt = dt;
  a.src = Y;
  a.aux = W;
  a.out1 = out;
  a.out2 = Z;
  a.flags = flags;
  stage_geometry(c, a);
  HIP_TRY(hipMemsetAsync(c->invdt, 0, sizeof(unsigned long long), s));
  if (c->geom.pow2)
    stage_kernel<double, 2, false, false, Tile16x16, true><<<a.nblk, NT, 0, s>>>(a);
  else
    stage_kernel<double, 2><<<a.nblk, NT, 0, s>>>(a);
  HIP_TRY(hipGetLastError());
  return GX_OK;
}

// Blocks of the line kernel of dimension d.
static int line_blocks(const PadGeom &g, int d) {
  const long long other = d == 0 ? (long long)g.n[1] * g.n[2] : g.n[d == 1 ? 2 : 1];
  const long long nseg = ((g.n[d] + 4) * other + LINE_OUT - 1) / LINE_OUT;
  if (d == 0) return (int)((nseg + AB / WAVE - 1) / (AB / WAVE));
  const long long nxb = (g.n[0] + DXW - 1) / DXW;
  return (int)(nxb * nseg);
}

// Transposed residual of one stage: the primitives pass of its input src
// (with the CFL argmax search when idx is given, and the dt work of the step
// reversed before when dtw is given), the x pass folding in the
// finished stage (fi, MODE) and forming this stage's bar Kout, then the y
// and z passes with that bar; pass d's bars land in Vb.d[d], the dt and
// gamma bar partials at part (3 passes).
template <int MODE>
static int stage_vjp(gx_ctx *c, const double *src, bool prim, const FoldIn &fi, double dt,
                     const uint8_t *flags, const StageBars &Vb, double *part,
                     unsigned long long cfl_bits = 0, int *idx = nullptr,
                     const DtArgs *dtw = nullptr) {
  AdjointState *A = c->adj;
  const PadGeom &g = A->pg;
  StageArgs<double> a = base_args<double>(c);
  a.src = src;
  a.src_is_prim = prim ? 1 : 0;
  a.flags = flags;
  const int wdt = dtw ? 1 : 0;
  adj_prims_kernel<<<blocks_for(g.np) + wdt, AB, 0, c->stream>>>(a, g, A->Vp, cfl_bits, idx,
                                                                 dtw ? *dtw : DtArgs{}, wdt);
  HIP_TRY(hipGetLastError());
  // cells with two coordinates outside are never written nor read
  auto passes = [&](auto p2) {
    constexpr bool P2 = decltype(p2)::value;
    adj_line_x_kernel<MODE, P2><<<line_blocks(g, 0), AB, 0, c->stream>>>(a, g, A->Vp, fi, dt,
                                                                         Vb.d[0], part);
    adj_line_t_kernel<1, P2><<<line_blocks(g, 1), AB, 0, c->stream>>>(
        a, g, A->Vp, fi.Kout, dt, Vb.d[1], part + 2 * (size_t)A->part_stride);
    adj_line_t_kernel<2, P2><<<line_blocks(g, 2), AB, 0, c->stream>>>(
        a, g, A->Vp, fi.Kout, dt, Vb.d[2], part + 4 * (size_t)A->part_stride);
  };
  if (c->geom.pow2)
    passes(std::true_type{});
  else
    passes(std::false_type{});
  HIP_TRY(hipGetLastError());
  return GX_OK;
}

// Recomputes segment si into its buffers on the side stream, after the work
// the context's stream holds so far (the reverse pass that last used them).
static int recompute_segment(gx_ctx *c, int si) {
  AdjointState *A = c->adj;
  SegmentBuffers &b = A->sb[si & 1];
  const int every = A->every, seg0 = si * every, len = std::min(every, A->nsteps - seg0);
  const long long n5 = 5 * c->ncell;
  const double *P0 = (const double *)c->P0;
  HIP_TRY(hipEventRecord(A->mark, c->stream));
  HIP_TRY(hipStreamWaitEvent(A->side, A->mark, 0));
  const double *ck = A->ckpt[si];
  for (int j = 0; j < len; j++) {
    const int n = seg0 + j;
    const double *Un = j == 0 ? ck : b.U + (size_t)j * n5;
    double *Un1 = j + 1 < len ? b.U + (size_t)(j + 1) * n5 : A->trash;
    const uint8_t *Fn = A->stepF + (size_t)n * c->ncell;
    double *Yn = b.Y + (size_t)j * n5, *Wn = b.W + (size_t)j * n5;
    int rc = launch_stage1(c, A->side, n == 0 ? P0 : Un, n == 0, Un, Yn, Wn, A->dt[n], Fn);
    if (rc) return rc;
    rc = launch_stage2(c, A->side, Yn, Wn, Un1, b.Z + (size_t)j * n5, A->dt[n], Fn);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(b.ready, A->side));
  b.held = seg0;
  return GX_OK;
}

static int ensure_state(gx_ctx *c, int nsteps, int every) {
  if (c->fp32) return fail(GX_E_ARG, "gx_adjoint: only Precision = GX_FP64");
  if (c->adj && c->adj->nsteps == nsteps && c->adj->every == every) {
    // a run of the same shape reuses the buffers (the tails of the partials
    // stay 0; argmax is set up again in case a reverse pass stopped early)
    AdjointState *A = c->adj;
    const int init[2] = {0x7fffffff, 0};
    HIP_TRY(hipMemcpy(A->argmax, init, sizeof(init), hipMemcpyHostToDevice));
    A->dt.clear();
    A->ts.clear();
    A->invdt.clear();
    A->ramp.clear();
    for (SegmentBuffers &b : A->sb) b.held = -1;
    A->seed_ready = A->bar_ready = A->forward_done = false;
    HIP_TRY(hipStreamSynchronize(A->side));
    return GX_OK;
  }
  release_adjoint(c);
  AdjointState *A = new (std::nothrow) AdjointState();
  if (!A) return fail(GX_E_NOMEM, "gx_adjoint: host allocation failed");
  c->adj = A;
  A->nsteps = nsteps;
  A->every = every;
  PadGeom &g = A->pg;
  g.np = 1;
  for (int d = 0; d < 3; d++) {
    g.n[d] = c->n[d];
    g.m[d] = c->n[d] + 4;
    g.np *= g.m[d];
  }
  g.s[0] = 1;
  g.s[1] = g.m[0];
  g.s[2] = (long long)g.m[0] * g.m[1];
  const size_t s5 = 5 * sizeof(double) * (size_t)c->ncell;
  const int nck = (nsteps + every - 1) / every + 1;
  A->ckpt.assign(nck, nullptr);
  for (int i = 0; i < nck; i++) HIP_TRY(hipMalloc(&A->ckpt[i], s5));
  for (SegmentBuffers &b : A->sb) {
    for (double **p : {&b.U, &b.Y, &b.W, &b.Z}) HIP_TRY(hipMalloc(p, s5 * every));
    HIP_TRY(hipEventCreateWithFlags(&b.ready, hipEventDisableTiming));
  }
  HIP_TRY(hipEventCreateWithFlags(&A->mark, hipEventDisableTiming));
  HIP_TRY(hipStreamCreate(&A->side));
  HIP_TRY(hipMalloc(&A->stepF, (size_t)c->ncell * nsteps));
  for (double **p : {&A->trash, &A->Ub, &A->P0b}) HIP_TRY(hipMalloc(p, s5));
  // cell-interleaved work arrays: stage bars in AOS records per cell,
  // padded primitives and bars in PAD5 records per position
  for (double **p : {&A->K1b, &A->K2b})
    HIP_TRY(hipMalloc(p, AOS * sizeof(double) * (size_t)c->ncell));
  for (double **p : {&A->Vp, &A->Vb[0][0], &A->Vb[0][1], &A->Vb[0][2], &A->Vb[1][0],
                     &A->Vb[1][1], &A->Vb[1][2]})
    HIP_TRY(hipMalloc(p, PAD5 * sizeof(double) * (size_t)g.np));
  A->part_stride = 0;
  for (int d = 0; d < 3; d++) A->part_stride = std::max(A->part_stride, line_blocks(g, d));
  // every line block writes its pair; the tails of the shorter passes stay 0
  HIP_TRY(hipMalloc(&A->partials, sizeof(double) * 2 * 6 * (size_t)A->part_stride));
  HIP_TRY(hipMemset(A->partials, 0, sizeof(double) * 2 * 6 * (size_t)A->part_stride));
  HIP_TRY(hipMalloc(&A->partials2, sizeof(double) * 2 * PART_BLOCKS));
  HIP_TRY(hipMalloc(&A->gb, sizeof(double)));
  // argmax[0]: cell of the step's CFL maximum (the dt kernel restores
  // INT_MAX after each use); argmax[1]: steps whose maximum no cell matched
  HIP_TRY(hipMalloc(&A->argmax, 2 * sizeof(int)));
  {
    const int init[2] = {0x7fffffff, 0};
    HIP_TRY(hipMemcpy(A->argmax, init, sizeof(init), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMalloc(&A->band, sizeof(double) * c->n[1]));
  HIP_TRY(hipMalloc(&A->lossp, sizeof(double) * blocks_for(c->ncell)));
  return GX_OK;
}

extern "C" {

int gx_adjoint_forward(gx_ctx *c, int nSteps, int checkpointEvery, int nRampCycles) {
  GX_API_BEGIN
  if (c && c->bricks)
    return fail(GX_E_STATE, "gx_adjoint_forward: not available on a chart split into bricks");
  if (!c || nSteps < 1 || checkpointEvery < 1 || nRampCycles < 1)
    return fail(GX_E_ARG, "gx_adjoint_forward: bad argument");
  if (!c->uploaded || !c->prims_valid)
    return fail(GX_E_STATE, "gx_adjoint_forward: needs freshly uploaded primitives");
  if (c->params.ReconstructionType != GX_LINEAR)
    return fail(GX_E_ARG, "gx_adjoint_forward: the adjoint differentiates ReconstructionType GX_LINEAR");
  int rc = ensure_state(c, nSteps, checkpointEvery);
  if (rc) return rc;
  AdjointState *A = c->adj;
  const size_t s5 = 5 * sizeof(double) * (size_t)c->ncell;
  const long long n5 = 5 * c->ncell;
  // the last two segments, whose states the reverse pass needs first, are
  // kept in the two sets of segment buffers as they are run, not recomputed
  const int nseg = (nSteps + checkpointEvery - 1) / checkpointEvery;
  const int keep0 = std::max(0, nseg - 2) * checkpointEvery;
  double time = 0.0;
  for (int n = 0; n < nSteps; n++) {
    if (n % checkpointEvery == 0)
      HIP_TRY(hipMemcpyAsync(A->ckpt[n / checkpointEvery], c->U, s5, hipMemcpyDeviceToDevice,
                             c->stream));
    // the flags this step uses (those of the upload, then of the last step)
    HIP_TRY(hipMemcpyAsync(A->stepF + (size_t)n * c->ncell, c->flags, (size_t)c->ncell,
                           hipMemcpyDeviceToDevice, c->stream));
    double cand;
    rc = gx_compute_time_step(c, &cand);
    if (rc) return rc;
    unsigned long long bits;
    HIP_TRY(hipMemcpy(&bits, c->invdt, sizeof(bits), hipMemcpyDeviceToHost));
    double invdt;
    memcpy(&invdt, &bits, sizeof(invdt));
    // Integrator_Template % ComputeTimeStep and ComputeNewTime
    double ts = fmin(DBL_MAX, cand);
    const double ramp = fmin((double)(n + 1) / (double)nRampCycles, 1.0);
    if (ramp < 1.0) ts = ramp * ts;
    const double dt = (time + ts) - time;
    A->dt.push_back(dt);
    A->ts.push_back(ts);
    A->invdt.push_back(invdt);
    A->ramp.push_back(ramp);
    if (n < keep0) {
      rc = gx_step_rk2(c, time, dt);
    } else {
      const int si = n / checkpointEvery, j = n - si * checkpointEvery;
      SegmentBuffers &b = A->sb[si & 1];
      StepRecord rec;  // the first input is the segment's checkpoint
      rec.Uin = j == 0 ? A->ckpt[si] : b.U + (size_t)j * n5;
      rec.Y = b.Y + (size_t)j * n5;
      rec.W = b.W + (size_t)j * n5;
      rec.Z = b.Z + (size_t)j * n5;
      // the segment's last step writes the next checkpoint's state (or the
      // final state) in the context
      rec.Uout = j + 1 < checkpointEvery && n + 1 < nSteps ? b.U + (size_t)(j + 1) * n5
                                                            : (double *)c->U;
      rc = step_rk2(c, dt, &rec);
    }
    if (rc) return rc;
    time = time + dt;
  }
  for (SegmentBuffers &b : A->sb) {
    HIP_TRY(hipEventRecord(b.ready, c->stream));
    b.held = -1;
  }
  for (int si = keep0 / checkpointEvery; si < nseg; si++) A->sb[si & 1].held = si * checkpointEvery;
  HIP_TRY(hipStreamSynchronize(c->stream));
  A->forward_done = true;
  return GX_OK;
  GX_API_END("gx_adjoint_forward")
}

int gx_adjoint_backward(gx_ctx *c, const double *Ubar, double *P0bar, double *gammaBar) {
  GX_API_BEGIN
  if (!c || !gammaBar) return fail(GX_E_ARG, "gx_adjoint_backward: null argument");
  AdjointState *A = c->adj;
  if (!A || !A->forward_done)
    return fail(GX_E_STATE, "gx_adjoint_backward: no forward run recorded");
  if (!Ubar && !A->seed_ready)
    return fail(GX_E_STATE,
                "gx_adjoint_backward: Ubar is NULL and no device seed is held "
                "(gx_adjoint_kinetic_energy with Ubar = NULL leaves one)");
  const long long n5 = 5 * c->ncell;
  const size_t s5 = sizeof(double) * (size_t)n5;
  if (Ubar) HIP_TRY(hipMemcpyAsync(A->Ub, Ubar, s5, hipMemcpyHostToDevice, c->stream));
  A->seed_ready = false;  // the reverse pass overwrites Ub
  A->bar_ready = false;
  int rc = fill(c, A->P0b, n5, 0.0);
  if (rc) return rc;
  rc = fill(c, A->gb, 1, 0.0);
  if (rc) return rc;
  const double *P0 = (const double *)c->P0;
  const int every = A->every, N = A->nsteps;
  const double *Unext = nullptr;  // the stage-1 input of the step reversed last
  DtArgs dt_next = {};             // and its dt work, done by the next primitives pass
  for (int si = (N - 1) / every; si >= 0; si--) {
    const int seg0 = si * every, len = std::min(every, N - seg0);
    // Segment si's states (recomputed from its checkpoint with the recorded
    // flags, unless its buffers hold them), then the recomputation of
    // segment si - 1 on the side stream, beside this segment's reverse pass.
    SegmentBuffers &b = A->sb[si & 1];
    if (b.held != seg0) {
      rc = recompute_segment(c, si);
      if (rc) return rc;
    }
    HIP_TRY(hipStreamWaitEvent(c->stream, b.ready, 0));
    if (si >= 1 && A->sb[(si - 1) & 1].held != seg0 - every) {
      rc = recompute_segment(c, si - 1);
      if (rc) return rc;
    }
    const double *ck = A->ckpt[si];
    for (int j = len - 1; j >= 0; j--) {
      const int n = seg0 + j;
      const double *Un = j == 0 ? ck : b.U + (size_t)j * n5;
      const double *Yn = b.Y + (size_t)j * n5;
      const double *Zn = b.Z + (size_t)j * n5;
      const uint8_t *Fn = A->stepF + (size_t)n * c->ncell;
      const bool first = n == 0;
      const double dt = A->dt[n];
      // stage 2's transpose; its x pass first folds in stage 1 of step n + 1
      // (or takes the seed) and forms K2b = fix^T(Ub)/2
      FoldIn f2 = {};
      f2.out = A->Ub;
      f2.Z = Zn;
      f2.Kout = A->K2b;
      double *part2 = A->partials + 6 * (size_t)A->part_stride;
      const StageBars bars2 = {{A->Vb[0][0], A->Vb[0][1], A->Vb[0][2]}};
      const StageBars bars1 = {{A->Vb[1][0], A->Vb[1][1], A->Vb[1][2]}};
      if (Unext) {
        f2.Vb = bars1;
        f2.src = Unext;
        rc = stage_vjp<FOLD_STAGE1>(c, Yn, false, f2, dt, Fn, bars2, part2, 0, nullptr,
                                    &dt_next);
      } else {
        rc = stage_vjp<FOLD_SEED>(c, Yn, false, f2, dt, Fn, bars2, part2);
      }
      if (rc) return rc;
      // stage 1's transpose; its x pass folds in stage 2 (Ub = Yb + Wb,
      // K1b = Yb + Wb/2), its primitives pass finds the CFL maximum's cell
      FoldIn f1 = {};
      f1.Vb = bars2;
      f1.src = Yn;
      f1.out = A->Ub;
      f1.K2b = A->K2b;
      f1.Kout = A->K1b;
      unsigned long long bits;
      memcpy(&bits, &A->invdt[n], sizeof(bits));
      rc = stage_vjp<FOLD_STAGE2>(c, first ? P0 : Un, first, f1, dt, Fn, bars1,
                                  A->partials, bits, A->argmax);
      if (rc) return rc;
      StageArgs<double> a = base_args<double>(c);
      a.src = first ? P0 : Un;
      a.src_is_prim = first ? 1 : 0;
      double *out = first ? A->P0b : A->Ub;
      if (first) {  // the last stage transpose: its own fold
        adj_fold_kernel<<<blocks_for(c->ncell), AB, 0, c->stream>>>(a, A->pg, bars1, out);
        HIP_TRY(hipGetLastError());
      }
      // dt = ts(invdt(U^n)): the partials of both stages summed now, the
      // rest by the next primitives pass (or here, on the first step)
      adj_partsum_kernel<<<PART_BLOCKS, AB, 0, c->stream>>>(A->partials, 6 * A->part_stride,
                                                            A->partials2);
      HIP_TRY(hipGetLastError());
      dt_next = DtArgs{a, A->partials2, PART_BLOCKS, A->ts[n], A->invdt[n], A->argmax, out,
                       A->gb};
      if (first) {
        adj_dt_kernel<<<1, AB, 0, c->stream>>>(dt_next);
        HIP_TRY(hipGetLastError());
      }
      Unext = Un;
    }
  }
  // U^0 = ComputeFromInitial(P0)
  adj_from_prim_kernel<<<blocks_for(c->ncell), AB, 0, c->stream>>>(
      c->ncell, c->params.BaryonMassReference, P0, A->Ub, A->P0b);
  HIP_TRY(hipGetLastError());
  if (P0bar) HIP_TRY(hipMemcpyAsync(P0bar, A->P0b, s5, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(gammaBar, A->gb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  int missed = 0;
  HIP_TRY(hipMemcpyAsync(&missed, A->argmax + 1, sizeof(int), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipMemsetAsync(A->argmax + 1, 0, sizeof(int), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->invdt_valid = false;  // the recomputation reused the CFL accumulator
  c->cfl_pending = false;
  if (missed)
    return fail(GX_E_STATE,
                "gx_adjoint_backward: the recomputed CFL terms did not reproduce the "
                "recorded maximum in " + std::to_string(missed) +
                " step(s); the time-step part of the gradient is missing");
  A->bar_ready = true;
  return GX_OK;
  GX_API_END("gx_adjoint_backward")
}

int gx_adjoint_get_bar(gx_ctx *c, double *P0bar) {
  GX_API_BEGIN
  if (!c || !P0bar) return fail(GX_E_ARG, "gx_adjoint_get_bar: null argument");
  AdjointState *A = c->adj;
  if (!A || !A->bar_ready)
    return fail(GX_E_STATE, "gx_adjoint_get_bar: no reverse pass has completed");
  HIP_TRY(hipMemcpyAsync(P0bar, A->P0b, 5 * sizeof(double) * (size_t)c->ncell,
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GX_OK;
  GX_API_END("gx_adjoint_get_bar")
}

int gx_adjoint_kinetic_energy(gx_ctx *c, int component, double lo, double hi,
                              double *loss, double *Ubar) {
  GX_API_BEGIN
  if (!c || !loss || component < 0 || component > 2)
    return fail(GX_E_ARG, "gx_adjoint_kinetic_energy: bad argument");
  if (!c->adj) return fail(GX_E_STATE, "gx_adjoint_kinetic_energy: no forward run recorded");
  AdjointState *A = c->adj;
  std::vector<double> band(c->n[1]);
  {
    // cell centres of dimension 1 (Chart_Template.f90:827-841)
    const int n = c->n[1];
    const double y0 = c->grid.MinCoordinate[1], w = (c->grid.MaxCoordinate[1] - y0) / n;
    double e = y0;
    for (int j = 0; j < n; j++) {
      const double e1 = e + w;
      const double yc = 0.5 * (e + e1);
      band[j] = (yc > lo && yc < hi) ? 1.0 : 0.0;
      e = e1;
    }
  }
  HIP_TRY(hipMemcpyAsync(A->band, band.data(), sizeof(double) * band.size(),
                         hipMemcpyHostToDevice, c->stream));
  StageArgs<double> a = base_args<double>(c);
  a.src = (const double *)c->U;
  const int nb = blocks_for(c->ncell);
  adj_loss_kernel<<<nb, AB, 0, c->stream>>>(a, component, A->band, A->lossp, A->Ub);
  HIP_TRY(hipGetLastError());
  std::vector<double> part(nb);
  HIP_TRY(hipMemcpyAsync(part.data(), A->lossp, sizeof(double) * nb, hipMemcpyDeviceToHost,
                         c->stream));
  if (Ubar)
    HIP_TRY(hipMemcpyAsync(Ubar, A->Ub, 5 * sizeof(double) * (size_t)c->ncell,
                           hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  A->seed_ready = true;  // the seed stays on the device for gx_adjoint_backward
  double s = 0.0;
  for (double v : part) s += v;
  *loss = s;
  return GX_OK;
  GX_API_END("gx_adjoint_kinetic_energy")
}

// Host evaluations of the device code's reverse sweeps, for checking.
void gx_adjoint_face_vjp(const double WL[5], const double WR[5], int d, int df,
                         double gamma, double mass, int hllc, const double Fbar[5],
                         double WLbar[5], double WRbar[5], double *gammaBar, double F[5]) {
  for (int f = 0; f < 5; f++) WLbar[f] = WRbar[f] = 0.0;
  double gb = 0.0;
  adj::face_flux_vjp(WL, WR, d, df != 0, gamma, mass, hllc != 0, Fbar, WLbar, WRbar, gb, F);
  *gammaBar = gb;
}

void gx_adjoint_slope_vjp(double vm, double v0, double vp, double dxl, double dxr,
                          double theta, int useLimiter, double sbar, double out[3]) {
  double a = 0.0, b = 0.0, cc = 0.0;
  adj::slope_vjp(vm, v0, vp, dxl, dxr, dxl * dxl, dxr * dxr, dxl * dxr * (dxl + dxr), theta,
                 useLimiter, sbar, a, b, cc);
  out[0] = a;
  out[1] = b;
  out[2] = cc;
}

}  // extern "C"
