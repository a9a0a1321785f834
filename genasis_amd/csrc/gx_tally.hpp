// gx_tally.hpp -- boundary fluence and the Tally_F_P conservation tallies of
// one chart.
//
// Boundary fluence: Step_RK_C_ASC clears BoundaryFluence_CSL before each step
// (EV/Steps/Step_RK_C_ASC__Template.f90:526-529) and every stage adds
// Weight_RK * dt * F_I of the chart's boundary faces to it
// (EV/Increments/IncrementDivergence_FV__Form.f90:989-991, 1002-1069;
// IncrementDivergence_FV__Kernel.f90:755-800).  Here one small kernel per
// stage recomputes the flux of each boundary face from the stage input with
// the stage kernel's own device functions (same loads, ghost rule, slopes,
// reconstruction and Riemann flux), so its values are the stage kernel's bit
// for bit and the fused stage kernel is left as it is.
//
// After the step the surface integral of the boundary integrands over the
// fluence is added to the chart's boundary tally (AccumulateBoundaryTally,
// Step_RK_C_ASC__Template.f90:533-537; FLD/Current_ASC__Template.f90:203-219;
// FLD/Tally_C__Form.f90:215-279; OP/Calculus/Integrals/SurfaceIntegral_Form.
// f90:25-177): the stage-2 kernel writes per-block partial sums and one block
// adds them, in a fixed order, to the device accumulator.  The interior tally
// is a volume integral over the proper cells (VolumeIntegral_Form.f90:25-117).
// All sums are in FP64, with a fixed reduction order (run-to-run identical).
#pragma once

#include "gx_tiles.hpp"

namespace gx {

// Tally_F_P % Value (Tally_F_D__Form.f90:83-88, Tally_F_P__Form.f90:63-64):
// BaryonNumber, Momentum_1..3, KineticEnergy, AngularMomentum_1..3,
// GravitationalEnergy, TotalEnergy, FluidEnergy, InternalEnergy.  The flat
// geometry selects all but the gravitational and total energies
// (Tally_F_P__Form.f90:105-109).
constexpr int TALLY_N = 12;
enum {
  T_BN = 0, T_M1, T_M2, T_M3, T_KE, T_AM1, T_AM2, T_AM3, T_GE, T_TE, T_FE, T_IE
};
// Boundary integrands with a value: BN, M1..3, AM1..3, FE (the kinetic and
// internal energies have none, Tally_F_P__Form.f90:190-239).
constexpr int NB_SUM = 8;
__host__ __device__ constexpr int boundary_slot(int i) {
  return i < 4 ? i : (i < 7 ? i + 1 : T_FE);
}
// Interior integrands: all selected ones.
constexpr int NI_SUM = 10;
__host__ __device__ constexpr int interior_slot(int i) { return i < 8 ? i : i + 2; }

constexpr int TALLY_BLOCK = 256;

// Boundary fluence layout: face b = 2*d (inner) or 2*d + 1 (outer) of
// dimension d holds the five conserved fields D, S_1, S_2, S_3, G, each a
// slab of the transverse proper cells with x fastest (d = 0: y + n1*z;
// d = 1: x + n0*z; d = 2: x + n0*y).  Faces follow one another, b = 0 .. 5.
__host__ __device__ inline long long fluence_slab(const int n[3], int d) {
  return d == 0 ? (long long)n[1] * n[2]
                : (d == 1 ? (long long)n[0] * n[2] : (long long)n[0] * n[1]);
}
__host__ __device__ inline long long fluence_face(const int n[3], int b) {
  long long o = 0;
  for (int e = 0; e < b / 2; e++) o += 10 * fluence_slab(n, e);
  return o + (b & 1) * 5 * fluence_slab(n, b / 2);
}
__host__ __device__ inline long long fluence_count(const int n[3]) {
  return fluence_face(n, 6);
}

// FP64 geometry of the tallies: array k of dimension d at tg + (3*d + k)*gs,
// k = 0 Center, 1 Width (WidthLeft + WidthRight), 2 Center - WidthLeft (the
// inner face position, Tally_C_Form % ComputeFacePositions), padded index.
enum { TG_CENTER = 0, TG_WIDTH, TG_FACE, N_TG };

struct BoundaryArgs {
  double *bf;           // fluence slabs
  double fw;            // Weight_RK * dt of the stage
  const double *tg;     // tally geometry
  double *partial;      // stage 2: NB_SUM per block
};

// Boundary lines: one per slab cell of each dimension, each holding the
// inner and the outer face of that line.
__host__ __device__ inline long long boundary_lines(const int n[3]) {
  return fluence_slab(n, 0) + fluence_slab(n, 1) + fluence_slab(n, 2);
}

// Line e of [0, boundary_lines): dimension d and slab index.
__device__ inline void boundary_line_of(const int n[3], long long e, int *d, long long *idx) {
  const long long s0 = fluence_slab(n, 0), s1 = fluence_slab(n, 1);
  if (e < s0) {
    *d = 0;
    *idx = e;
  } else if (e < s0 + s1) {
    *d = 1;
    *idx = e - s0;
  } else {
    *d = 2;
    *idx = e - s0 - s1;
  }
}

// Transverse coordinates of slab index idx of dimension D (x-fastest order).
template <int D>
__device__ inline void line_coords(const int n[3], long long idx, int *u, int *v) {
  const int nj = D == 0 ? n[1] : n[0];
  *u = (int)(idx % nj);
  *v = (int)(idx / nj);
}

template <int D>
__device__ inline unsigned line_cell(const int n[3], int u, int v, int c) {
  return D == 0 ? lin(n, c, u, v) : (D == 1 ? lin(n, u, c, v) : lin(n, u, v, c));
}

// HLL(C) flux of the face between cells with primitives Vm1 | V0, from the
// four-cell stencil, as the stage kernel computes it (slopes of the two
// cells beside the face and the linear reconstruction, or the two cells'
// parabolic face values; then the face flux).
template <typename R, int D, int RECON>
__device__ inline void stencil_flux(const StageArgs<R> &a, int cl, const R Vm2[5],
                                    const R Vm1[5], const R V0[5], const R Vp1[5], bool df,
                                    R F[5]) {
  const int cr = cl + 1;
  R WL[5], WR[5];
  if constexpr (RECON == RECON_PARABOLIC) {
#pragma unroll
    for (int f = 0; f < 5; f++) {
      R iL, oR;
      parabolic_at(a, D, cl, Vm2[f], Vm1[f], V0[f], &iL, &WL[f]);
      parabolic_at(a, D, cr, Vm1[f], V0[f], Vp1[f], &WR[f], &oR);
    }
  } else {
    const R wr = geo(a, D, G_WR, cl + 2), wl = geo(a, D, G_WL, cr + 2);
#pragma unroll
    for (int f = 0; f < 5; f++) {
      const R sl = slope_at(a, D, cl, Vm2[f], Vm1[f], V0[f]);
      const R sr = slope_at(a, D, cr, Vm1[f], V0[f], Vp1[f]);
      WL[f] = Vm1[f] + wr * sl;
      WR[f] = V0[f] - wl * sr;
    }
  }
  face_flux(WL, WR, D, df, a.ph, F);
}

// Primitives of a raw value with the velocity component D negated when flip.
template <typename R, int D>
__device__ inline void prim_of(const StageArgs<R> &a, const R raw[5], bool flip, R V[5]) {
  raw_to_prim(a, raw, flip ? D : -1, V);
}

// Fluxes through the inner and the outer face of one line of dimension D.
// A line's two faces read only its first two and last two cells (the ghost
// rule maps the ghosts onto them), so each raw value is loaded once.  The
// inner face takes the DiffusiveFlux flag of the line's first cell; the
// outer face, between the last cell and a ghost, takes none.
template <typename R, int D, int RECON>
__device__ inline void line_fluxes(const StageArgs<R> &a, long long idx, R Fin[5],
                                   R Fout[5]) {
  int u, v;
  line_coords<D>(a.n, idx, &u, &v);
  const int nd = a.n[D];
  const unsigned i0 = line_cell<D>(a.n, u, v, 0), i1 = line_cell<D>(a.n, u, v, 1),
                 j1 = line_cell<D>(a.n, u, v, nd - 1), j2 = line_cell<D>(a.n, u, v, nd - 2);
  R r0[5], r1[5], q1[5], q2[5];  // cells 0, 1, n-1, n-2
#pragma unroll
  for (int f = 0; f < 5; f++) {
    r0[f] = ld(a.src + f * a.fs, i0);
    r1[f] = ld(a.src + f * a.fs, i1);
    q1[f] = ld(a.src + f * a.fs, j1);
    q2[f] = ld(a.src + f * a.fs, j2);
  }
  const bool df = (ld(a.flags, i0) >> D) & 1u;
  const int bin = a.bc[2 * D], bout = a.bc[2 * D + 1];
  R P0[5], P1[5], Q1[5], Q2[5], G1[5], G2[5];
  prim_of<R, D>(a, r0, false, P0);
  prim_of<R, D>(a, r1, false, P1);
  prim_of<R, D>(a, q1, false, Q1);
  prim_of<R, D>(a, q2, false, Q2);
  // inner ghosts -1, -2: periodic images n-1, n-2, or mirrors of 0, 1
  if (bin == BC_PERIODIC) {
    prim_of<R, D>(a, q1, false, G1);
    prim_of<R, D>(a, q2, false, G2);
  } else {
    prim_of<R, D>(a, r0, bin == BC_REFLECTING, G1);
    prim_of<R, D>(a, r1, bin == BC_REFLECTING, G2);
  }
  stencil_flux<R, D, RECON>(a, -1, G2, G1, P0, P1, df, Fin);
  // outer ghosts n, n+1: periodic images 0, 1, or mirrors of n-1, n-2
  if (bout == BC_PERIODIC) {
    prim_of<R, D>(a, r0, false, G1);
    prim_of<R, D>(a, r1, false, G2);
  } else {
    prim_of<R, D>(a, q1, bout == BC_REFLECTING, G1);
    prim_of<R, D>(a, q2, bout == BC_REFLECTING, G2);
  }
  stencil_flux<R, D, RECON>(a, nd - 1, Q2, Q1, G1, G2, false, Fout);
}

// BF = BF + fw * F for the five fields of one face value (stage 1 adds to
// the cleared fluence), returning the new values.
template <int STAGE, typename R>
__device__ inline void record_face(const BoundaryArgs &B, const int n[3], int b,
                                   long long idx, const R F[5], double v[5]) {
  const long long slab = fluence_slab(n, b / 2);
  double *p = B.bf + fluence_face(n, b) + idx;
#pragma unroll
  for (int f = 0; f < 5; f++) {
    const double x = B.fw * (double)F[f];
    v[f] = STAGE == 1 ? 0.0 + x : p[f * slab] + x;
    p[f * slab] = v[f];
  }
}

// Integrand * Direction * area of one face value (Tally_F_D__Form.f90:
// 387-480, 1007-1034; Tally_F_P__Form.f90:190-239; SurfaceIntegral_Form.f90:
// 150-177): positions are the transverse centres and, along d, the face; the
// area is that of SetFiniteVolumeRectangularKernel.  Added to I.
template <int D>
__device__ inline void add_integrands(const BoundaryArgs &B, int gs, int side, int u, int v,
                                      int cd, const double F[5], double I[NB_SUM]) {
  const int cx = D == 0 ? cd : u, cy = D == 0 ? u : (D == 1 ? cd : v), cz = D == 2 ? cd : v;
  const double *tg = B.tg;
  const double X0 = tg[(D == 0 ? TG_FACE : TG_CENTER) * gs + cx + 2];
  const double X1 = tg[(3 + (D == 1 ? TG_FACE : TG_CENTER)) * gs + cy + 2];
  const double X2 = tg[(6 + (D == 2 ? TG_FACE : TG_CENTER)) * gs + cz + 2];
  const double w0 = tg[TG_WIDTH * gs + cx + 2], w1 = tg[(3 + TG_WIDTH) * gs + cy + 2],
               w2 = tg[(6 + TG_WIDTH) * gs + cz + 2];
  const double dA = D == 0 ? w1 * w2 : (D == 1 ? w2 * w0 : w0 * w1);
  const double dir = side ? 1.0 : -1.0;
  const double S1 = F[1], S2 = F[2], S3 = F[3];
  const double J[NB_SUM] = {F[0], S1, S2, S3, X1 * S3 - X2 * S2,
                            X2 * S1 - X0 * S3, X0 * S2 - X1 * S1, F[4]};
#pragma unroll
  for (int i = 0; i < NB_SUM; i++) I[i] += dir * J[i] * dA;
}

template <typename R, int STAGE, int D, int RECON>
__device__ inline void boundary_line(const StageArgs<R> &a, const BoundaryArgs &B,
                                     const int n[3], long long idx, double I[NB_SUM]) {
  R Fin[5], Fout[5];
  line_fluxes<R, D, RECON>(a, idx, Fin, Fout);
  double vin[5], vout[5];
  record_face<STAGE>(B, n, 2 * D, idx, Fin, vin);
  record_face<STAGE>(B, n, 2 * D + 1, idx, Fout, vout);
  if (STAGE == 2) {
    int u, v;
    line_coords<D>(n, idx, &u, &v);
    add_integrands<D>(B, a.gs, 0, u, v, 0, vin, I);
    add_integrands<D>(B, a.gs, 1, u, v, n[D], vout, I);
  }
}

// One stage of fluence recording, one boundary line (two face values) per
// thread; stage 2 also sums the boundary integrands of the finished fluence
// into per-block partials.
template <typename R, int STAGE, int RECON>
__global__ __launch_bounds__(TALLY_BLOCK) void boundary_flux_kernel(StageArgs<R> a,
                                                                    BoundaryArgs B) {
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
      boundary_line<R, STAGE, 0, RECON>(a, B, n, idx, I);
    else if (d == 1)
      boundary_line<R, STAGE, 1, RECON>(a, B, n, idx, I);
    else
      boundary_line<R, STAGE, 2, RECON>(a, B, n, idx, I);
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

// TallyBoundaryLocal += the sum of the nblk block partials: thread t sums
// blocks t, t + 256, ... in turn (loads issued in groups of 8), then a tree
// over the threads; the order is fixed.
static __global__ __launch_bounds__(TALLY_BLOCK) void boundary_tally_finish(
    const double *partial, int nblk, double *local) {
  __shared__ double s[NB_SUM][TALLY_BLOCK];
  const int t = threadIdx.x;
  double acc[NB_SUM];
#pragma unroll
  for (int i = 0; i < NB_SUM; i++) acc[i] = 0.0;
  for (int b = t; b < nblk; b += TALLY_BLOCK) {
    double v[NB_SUM];
#pragma unroll
    for (int i = 0; i < NB_SUM; i++) v[i] = partial[b * NB_SUM + i];
#pragma unroll
    for (int i = 0; i < NB_SUM; i++) acc[i] += v[i];
  }
#pragma unroll
  for (int i = 0; i < NB_SUM; i++) s[i][t] = acc[i];
  __syncthreads();
  for (int o = TALLY_BLOCK / 2; o > 0; o >>= 1) {
    if (t < o)
#pragma unroll
      for (int i = 0; i < NB_SUM; i++) s[i][t] += s[i][t + o];
    __syncthreads();
  }
  if (t < NB_SUM) local[boundary_slot(t)] += s[t][0];
}

// Interior tally integrands of the Current (Tally_F_D__Form.f90:248-384,
// 737-782; Tally_F_P__Form.f90:147-187) times the cell volume, per-block
// partials of NI_SUM values.  The Current is the conserved state with its
// fix applied and the primitives that follow (as downloaded), or, right
// after an upload, the uploaded primitives.
template <typename R>
__global__ __launch_bounds__(TALLY_BLOCK) void interior_tally_kernel(StageArgs<R> a,
                                                                     const R *prim,
                                                                     const double *tg,
                                                                     double *partial) {
  __shared__ double s[NI_SUM][TALLY_BLOCK];
  const long long ntot = (long long)a.n[0] * a.n[1] * a.n[2];
  double acc[NI_SUM];
#pragma unroll
  for (int i = 0; i < NI_SUM; i++) acc[i] = 0.0;
  for (long long i = blockIdx.x * (long long)TALLY_BLOCK + threadIdx.x; i < ntot;
       i += (long long)gridDim.x * TALLY_BLOCK) {
    const int c[3] = {(int)(i % a.n[0]), (int)((i / a.n[0]) % a.n[1]),
                      (int)(i / ((long long)a.n[0] * a.n[1]))};
    double U[5], V[5], X[3], w[3];
#pragma unroll
    for (int f = 0; f < 5; f++) U[f] = (double)a.aux[f * a.fs + i];
    fix_conserved(U, a.ph.mass, a.ph.nmin, V);
    if (prim) {
#pragma unroll
      for (int f = 0; f < 5; f++) V[f] = (double)prim[f * a.fs + i];
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      X[q] = tg[(3 * q + TG_CENTER) * a.gs + c[q] + 2];
      w[q] = tg[(3 * q + TG_WIDTH) * a.gs + c[q] + 2];
    }
    const double dV = w[0] * w[1] * w[2];
    const double S1 = U[1], S2 = U[2], S3 = U[3];
    const double I[NI_SUM] = {U[0], S1, S2, S3,
                              0.5 * ((S1 * V[1] + S2 * V[2]) + S3 * V[3]),
                              X[1] * S3 - X[2] * S2, X[2] * S1 - X[0] * S3,
                              X[0] * S2 - X[1] * S1, U[4], V[4]};
#pragma unroll
    for (int k = 0; k < NI_SUM; k++) acc[k] += I[k] * dV;
  }
#pragma unroll
  for (int k = 0; k < NI_SUM; k++) s[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int o = TALLY_BLOCK / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
#pragma unroll
      for (int k = 0; k < NI_SUM; k++) s[k][threadIdx.x] += s[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < NI_SUM) partial[blockIdx.x * NI_SUM + threadIdx.x] = s[threadIdx.x][0];
}

}  // namespace gx
