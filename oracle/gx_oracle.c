/*
 * gx_oracle.c -- CPU restatement of the GenASiS Fluid_P_I RK2 step.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker of the HIP path and the CPU
 * baseline timed by bench.py.  The product never links or calls this file.
 *
 * Every pass below restates one reference kernel over the reference's storage
 * layout (one ghost-padded array per field, x fastest, 2 ghost layers) and
 * keeps the reference's expression order, so that with -ffp-contract=off the
 * arithmetic is the reference's.  Paths are relative to the GenASiS tree:
 *   F/   Modules/Physics/StressEnergies/Fluids/
 *   FLD/ Modules/Mathematics/Solvers/DifferentialEquations/Fields/
 *   EV/  Modules/Mathematics/Solvers/DifferentialEquations/Evolutions/
 *   CH/  Modules/Mathematics/Manifolds/Atlases/Charts/
 *   OP/  Modules/Mathematics/Operations/
 *
 * Conventions restated from the reference (the Fortran intrinsics):
 *   sign(0.5, x) -> copysign(0.5, x); max/min -> fmax/fmin; x**2 -> x*x;
 *   tiny(0.0_KDR) -> DBL_MIN; huge(1.0_KDR) -> DBL_MAX.
 * Fortran evaluates a*b*c as (a*b)*c and a+b-c as (a+b)-c; the parentheses
 * below spell that out.  Metric factors of the GALILEAN rectangular chart are
 * exactly 1.0 and multiplications by them are exact, so they are omitted.
 */
#include "gx_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NG 2 /* nGhostLayers (CH/ChartBasics/ChartHeader_SL__Form.f90:362) */

/* Conserved order of the step (Solution, Y, K): D, S_1, S_2, S_3, G, DS
 * (F/Fluid_D__Form.f90:273-274, F/Fluid_P__Template.f90:365-366). */
static const int IA_CONSERVED[6] = {GXO_D, GXO_S_1, GXO_S_2, GXO_S_3, GXO_G,
                                    GXO_DS};
/* Reconstructed = primitive: N, V_1, V_2, V_3, E (F/Fluid_ASC__Form.f90:156-165,
 * FLD/Current_Template.f90:309-312). */
static const int IA_PRIMITIVE[5] = {GXO_N, GXO_V_1, GXO_V_2, GXO_V_3, GXO_E};

struct gxo_ctx {
  int n[3], m[3];
  long nv;
  int bc[3][2];
  double gamma, theta, st, courant, n_min, c_v, n0, p0, mass;
  int use_limiter, riemann;
  int parabolic; /* ReconstructionType = 'PARABOLIC' */
  double *edge[3], *center[3], *wl[3], *wr[3];
  double *cur[GXO_N_FIELDS];
  double *df[3];
  double *sol[6], *y[6], *k1[6], *k2[6];
  double *sources[6]; /* Current % Sources, the divergence part (RecordDivergence) */
  /* IncrementDivergence scratch (EV/Increments/StorageDivergence_Form.f90) */
  double *il[GXO_N_FIELDS], *ir[GXO_N_FIELDS];
  double *icl[GXO_N_FIELDS], *icr[GXO_N_FIELDS];
  double *fil[6], *fir[6], *fi[6];
  double *ap, *am, *ac;
  double *dxi, *dvi[5], *grad[5], *xc;
  /* BoundaryFluence_CSL of the step: face 2*d (inner) and 2*d+1 (outer) of
   * dimension d, conserved fields D, S_1, S_2, S_3, G; slab of the transverse
   * proper cells, x fastest (size bf_n[d]). */
  double *bf[6][5];
  long bf_n[3];
  double weight_rk; /* Weight_RK of the stage being computed */
  /* Tally_F_P values (12, Tally_F_D__Form.f90:83-88, Tally_F_P__Form.f90:
   * 63-64): TallyBoundaryLocal, TallyBoundaryGlobal, TallyTotal, TallyChange */
  double tally_local[GXO_TALLY_N], tally_global[GXO_TALLY_N];
  double tally_total[GXO_TALLY_N], tally_change[GXO_TALLY_N];
  /* Proper cells that ComputeFromConserved repaired since the last
   * gxo_fallback_counts reset, per call site (0: the stage-2 input Y,
   * 1: StoreSolution and gxo_set_conserved) and branch (0: D <= N_Min,
   * vacuum; 1: E < 0). */
  long n_repaired[2][2];
};

static inline long IDX(const gxo_ctx *c, int p0, int p1, int p2) {
  return (long)p0 + (long)c->m[0] * ((long)p1 + (long)c->m[1] * (long)p2);
}

static double *alloc0(long n) {
  return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
}

/* ---- geometry: CH/ChartBasics/Chart_Template.f90:463-593, 699-723, 827-841,
 * 985-987 (SetGeometryCell, SetEdgeEqual, SetCenterRectangular,
 * SetWidthLeftRight). */
static void set_geometry(gxo_ctx *c, int d, double lo, double hi) {
  int n = c->n[d];
  /* Edge(1-nGL : n+nGL+1) stored at offset NG-1 so that Edge(1) -> e[NG]. */
  double *e = c->edge[d];
#define EDGE(i) e[(i) + NG - 1]
  EDGE(1) = lo;
  double width = (hi - lo) / n;
  for (int ic = 2; ic <= n + 1; ic++) EDGE(ic) = EDGE(ic - 1) + width;
  /* Ghost edges: Width_IG / Width_OG are declared real(KDL) in the reference
   * (Chart_Template.f90:477-478), KDL = kind(.true.) = 4, so each ghost width
   * is rounded to single precision before it is used. */
  for (int ic = 1; ic <= NG; ic++) {
    float w_ig = (float)(EDGE(ic + 1) - EDGE(ic));
    float w_og = (float)(EDGE(n - ic + 2) - EDGE(n - ic + 1));
    EDGE(1 - ic) = EDGE(2 - ic) - (double)w_ig;
    EDGE(n + 1 + ic) = EDGE(n + ic) + (double)w_og;
  }
  for (int p = 0; p < c->m[d]; p++) {
    int ic = p - NG + 1; /* 1-based cell index */
    c->center[d][p] = 0.5 * (EDGE(ic) + EDGE(ic + 1));
    c->wl[d][p] = c->center[d][p] - EDGE(ic);
    c->wr[d][p] = EDGE(ic + 1) - c->center[d][p];
  }
#undef EDGE
}

gxo_ctx *gxo_create(const int n[3], const double lo[3], const double hi[3],
                    const int bc[6], double gamma, double theta,
                    int use_limiter, int riemann, double shock_threshold,
                    double courant, double n_min, int n_threads) {
  gxo_ctx *c = (gxo_ctx *)calloc(1, sizeof(gxo_ctx));
  if (!c) return NULL;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
  c->nv = 1;
  for (int d = 0; d < 3; d++) {
    c->n[d] = n[d];
    c->m[d] = n[d] + 2 * NG;
    c->nv *= c->m[d];
    c->bc[d][0] = bc[2 * d];
    c->bc[d][1] = bc[2 * d + 1];
  }
  c->gamma = gamma;
  c->theta = theta;
  c->use_limiter = use_limiter;
  c->riemann = riemann;
  c->st = shock_threshold;
  c->courant = courant;
  c->n_min = n_min;
  /* F/Fluid_P_I__Form.f90:206-223: k = mu = 1, c_v = k/(mu(gamma-1)) with the
   * default gamma 1.4 (SetAdiabaticIndex later changes gamma only). */
  c->c_v = 1.0 / (1.0 * (1.4 - 1.0));
  c->n0 = 1.0;
  c->p0 = 1.0;
  c->mass = 1.0; /* BaryonMassReference */
  for (int d = 0; d < 3; d++) {
    c->edge[d] = alloc0(c->m[d] + 1);
    c->center[d] = alloc0(c->m[d]);
    c->wl[d] = alloc0(c->m[d]);
    c->wr[d] = alloc0(c->m[d]);
    set_geometry(c, d, lo[d], hi[d]);
    c->df[d] = alloc0(c->nv);
  }
  for (int f = 0; f < GXO_N_FIELDS; f++) {
    c->cur[f] = alloc0(c->nv);
    c->il[f] = alloc0(c->nv);
    c->ir[f] = alloc0(c->nv);
    c->icl[f] = alloc0(c->nv);
    c->icr[f] = alloc0(c->nv);
  }
  for (int f = 0; f < 6; f++) {
    c->sol[f] = alloc0(c->nv);
    c->y[f] = alloc0(c->nv);
    c->k1[f] = alloc0(c->nv);
    c->k2[f] = alloc0(c->nv);
    c->sources[f] = alloc0(c->nv);
    c->fil[f] = alloc0(c->nv);
    c->fir[f] = alloc0(c->nv);
    c->fi[f] = alloc0(c->nv);
  }
  c->ap = alloc0(c->nv);
  c->am = alloc0(c->nv);
  c->ac = alloc0(c->nv);
  c->dxi = alloc0(c->nv);
  c->xc = alloc0(c->nv);
  for (int f = 0; f < 5; f++) {
    c->dvi[f] = alloc0(c->nv);
    c->grad[f] = alloc0(c->nv);
  }
  for (int d = 0; d < 3; d++) {
    c->bf_n[d] = 1;
    for (int e = 0; e < 3; e++)
      if (e != d) c->bf_n[d] *= c->n[e];
    for (int s = 0; s < 2; s++)
      for (int f = 0; f < 5; f++) c->bf[2 * d + s][f] = alloc0(c->bf_n[d]);
  }
  return c;
}

void gxo_set_reconstruction(gxo_ctx *c, int parabolic) { c->parabolic = parabolic != 0; }

void gxo_destroy(gxo_ctx *c) {
  if (!c) return;
  for (int d = 0; d < 3; d++) {
    free(c->edge[d]); free(c->center[d]); free(c->wl[d]); free(c->wr[d]);
    free(c->df[d]);
  }
  for (int f = 0; f < GXO_N_FIELDS; f++) {
    free(c->cur[f]); free(c->il[f]); free(c->ir[f]); free(c->icl[f]);
    free(c->icr[f]);
  }
  for (int f = 0; f < 6; f++) {
    free(c->sol[f]); free(c->y[f]); free(c->k1[f]); free(c->k2[f]); free(c->sources[f]);
    free(c->fil[f]); free(c->fir[f]); free(c->fi[f]);
  }
  free(c->ap); free(c->am); free(c->ac); free(c->dxi); free(c->xc);
  for (int f = 0; f < 5; f++) { free(c->dvi[f]); free(c->grad[f]); }
  for (int b = 0; b < 6; b++)
    for (int f = 0; f < 5; f++) free(c->bf[b][f]);
  free(c);
}

void gxo_extents(const gxo_ctx *c, int m[3], long *n_values) {
  for (int d = 0; d < 3; d++) m[d] = c->m[d];
  *n_values = c->nv;
}

void gxo_geometry(const gxo_ctx *c, int d, double *center, double *wl,
                  double *wr) {
  memcpy(center, c->center[d], sizeof(double) * c->m[d]);
  memcpy(wl, c->wl[d], sizeof(double) * c->m[d]);
  memcpy(wr, c->wr[d], sizeof(double) * c->m[d]);
}

/* Padded 3-D loop over [lo, hi) per dimension. */
#define LOOP3(c, lo, hi)                                   \
  _Pragma("omp parallel for collapse(3) schedule(static)") \
  for (int p2 = (lo)[2]; p2 < (hi)[2]; p2++)               \
    for (int p1 = (lo)[1]; p1 < (hi)[1]; p1++)             \
      for (int p0 = (lo)[0]; p0 < (hi)[0]; p0++)

/* ---- EOS: F/Fluid_P_I__Kernel.f90:197-271 (Apply_EOS_I_E_Kernel) -------- */
static inline void eos_e(const gxo_ctx *c, double **F, long iv) {
  const double sqrt_huge = sqrt(DBL_MAX);
  double gamma = c->gamma;
  double N = F[GXO_N][iv], E = F[GXO_E][iv], M = F[GXO_M][iv];
  double P = (gamma - 1.0) * E;
  F[GXO_P][iv] = P;
  if (N > 0.0 && P > 0.0) {
    F[GXO_T][iv] = E / (c->c_v * N);
    F[GXO_SB][iv] = c->c_v * log(P / c->p0 * pow(c->n0 / N, gamma));
    F[GXO_CS][iv] = sqrt(gamma * P / (M * N));
  } else {
    F[GXO_T][iv] = 0.0;
    F[GXO_SB][iv] = -sqrt_huge;
    F[GXO_CS][iv] = 0.0;
  }
}

/* F/Fluid_P__Kernel.f90 Compute_FE_P_G_Kernel (MN and fast eigenspeeds). */
static inline void fe_p_g(double **F, long iv) {
  double V1 = F[GXO_V_1][iv], V2 = F[GXO_V_2][iv], V3 = F[GXO_V_3][iv];
  double CS = F[GXO_CS][iv];
  if (CS > 0.0)
    F[GXO_MN][iv] = sqrt((V1 * V1 + V2 * V2) + V3 * V3) / CS;
  else
    F[GXO_MN][iv] = 0.0;
  F[GXO_FEP_1][iv] = V1 + CS;
  F[GXO_FEP_2][iv] = V2 + CS; /* sqrt(M_UU_22) = 1 */
  F[GXO_FEP_3][iv] = V3 + CS;
  F[GXO_FEM_1][iv] = V1 - CS;
  F[GXO_FEM_2][iv] = V2 - CS;
  F[GXO_FEM_3][iv] = V3 - CS;
}

/* Fluid_P_I % ComputeFromPrimitiveCommon (F/Fluid_P_I__Form.f90:509-601):
 * Compute_M, Apply_EOS_I_E, Compute_D_S_G (F/Fluid_D__Kernel.f90:60-100),
 * Compute_G_G, Compute_DS_G, Compute_FE_P_G (F/Fluid_P__Kernel.f90). */
static inline void from_primitive_value(const gxo_ctx *c, double **F, long iv) {
  F[GXO_M][iv] = c->mass;
  eos_e(c, F, iv);
  double N = F[GXO_N][iv];
  if (N < 0.0) { N = 0.0; F[GXO_N][iv] = 0.0; }
  double M = F[GXO_M][iv];
  double V1 = F[GXO_V_1][iv], V2 = F[GXO_V_2][iv], V3 = F[GXO_V_3][iv];
  F[GXO_D][iv] = N;
  double S1 = M * N * V1, S2 = M * N * V2, S3 = M * N * V3;
  F[GXO_S_1][iv] = S1;
  F[GXO_S_2][iv] = S2;
  F[GXO_S_3][iv] = S3;
  F[GXO_G][iv] = F[GXO_E][iv] + 0.5 * ((S1 * V1 + S2 * V2) + S3 * V3);
  F[GXO_DS][iv] = F[GXO_SB][iv] * N;
  fe_p_g(F, iv);
}

/* Fluid_P_I % ComputeFromConservedCommon (F/Fluid_P_I__Form.f90:604-722) with
 * UseEntropy = .false.: Compute_N_V_E_G_Kernel (F/Fluid_P__Kernel.f90:94-192),
 * Apply_EOS_I_E, Compute_DS_G, Compute_FE_P_G. */
static inline int from_conserved_value(const gxo_ctx *c, double **F, long iv) {
  int repaired = 0;
  F[GXO_M][iv] = c->mass;
  double M = F[GXO_M][iv];
  double D = F[GXO_D][iv];
  if (D > c->n_min) {
    F[GXO_N][iv] = D;
    F[GXO_V_1][iv] = F[GXO_S_1][iv] / (M * D);
    F[GXO_V_2][iv] = F[GXO_S_2][iv] / (M * D);
    F[GXO_V_3][iv] = F[GXO_S_3][iv] / (M * D);
    F[GXO_E][iv] = F[GXO_G][iv] -
                   0.5 * ((F[GXO_S_1][iv] * F[GXO_V_1][iv] +
                           F[GXO_S_2][iv] * F[GXO_V_2][iv]) +
                          F[GXO_S_3][iv] * F[GXO_V_3][iv]);
  } else {
    repaired |= 1;
    F[GXO_N][iv] = c->n_min;
    F[GXO_V_1][iv] = 0.0; F[GXO_V_2][iv] = 0.0; F[GXO_V_3][iv] = 0.0;
    F[GXO_E][iv] = 0.0;
    F[GXO_D][iv] = c->n_min;
    F[GXO_S_1][iv] = 0.0; F[GXO_S_2][iv] = 0.0; F[GXO_S_3][iv] = 0.0;
    F[GXO_G][iv] = 0.0;
  }
  if (F[GXO_E][iv] < 0.0) {
    repaired |= 2;
    F[GXO_E][iv] = 0.0;
    F[GXO_G][iv] = F[GXO_E][iv] +
                   0.5 * ((F[GXO_S_1][iv] * F[GXO_V_1][iv] +
                           F[GXO_S_2][iv] * F[GXO_V_2][iv]) +
                          F[GXO_S_3][iv] * F[GXO_V_3][iv]);
  }
  eos_e(c, F, iv);
  F[GXO_DS][iv] = F[GXO_SB][iv] * F[GXO_N][iv];
  fe_p_g(F, iv);
  return repaired;
}

static void from_primitive_all(const gxo_ctx *c, double **F) {
  long nv = c->nv;
#pragma omp parallel for schedule(static)
  for (long iv = 0; iv < nv; iv++) from_primitive_value(c, F, iv);
}

static inline int is_proper(const gxo_ctx *c, long iv) {
  const int p0 = (int)(iv % c->m[0]);
  const int p1 = (int)((iv / c->m[0]) % c->m[1]);
  const int p2 = (int)(iv / ((long)c->m[0] * c->m[1]));
  return p0 >= NG && p0 < c->m[0] - NG && p1 >= NG && p1 < c->m[1] - NG &&
         p2 >= NG && p2 < c->m[2] - NG;
}

static void from_conserved_all(gxo_ctx *c, double **F, int site) {
  long nv = c->nv, nvac = 0, neng = 0;
#pragma omp parallel for schedule(static) reduction(+ : nvac, neng)
  for (long iv = 0; iv < nv; iv++) {
    const int r = from_conserved_value(c, F, iv);
    if (r && is_proper(c, iv)) {
      nvac += r & 1;
      neng += (r >> 1) & 1;
    }
  }
  c->n_repaired[site][0] += nvac;
  c->n_repaired[site][1] += neng;
}

/* ---- Chart_SLD % ExchangeGhostData on one rank (CH/SingleLevelCharts/
 * Chart_SLD__Form.f90:131-201, 324-554): every ghost value whose out-of-range
 * coordinates all lie in periodic dimensions takes the value of its periodic
 * image; ghosts across non-periodic physical faces are left alone. */
static void exchange_ghosts(const gxo_ctx *c, double **A, int nA) {
  int lo[3] = {0, 0, 0};
  const int *hi = c->m;
  LOOP3(c, lo, hi) {
    int p[3] = {p0, p1, p2}, q[3];
    int ghost = 0, ok = 1;
    for (int d = 0; d < 3; d++) {
      int ci = p[d] - NG;
      q[d] = p[d];
      if (ci < 0 || ci >= c->n[d]) {
        ghost = 1;
        if (c->bc[d][0] != GXO_BC_PERIODIC) ok = 0;
        ci = ((ci % c->n[d]) + c->n[d]) % c->n[d];
        q[d] = ci + NG;
      }
    }
    if (ghost && ok) {
      long dst = IDX(c, p[0], p[1], p[2]), src = IDX(c, q[0], q[1], q[2]);
      for (int a = 0; a < nA; a++) A[a][dst] = A[a][src];
    }
  }
}

/* ---- Atlas_SC % ApplyBoundaryConditions for dimension d (IC/Atlas_SC__
 * Template.f90:244-324, 360-406) -> Chart_SL % CopyBoundary / ReverseBoundary
 * (CH/SingleLevelCharts/Chart_SL__Template.f90:486-538, Chart_SL__Kernel.f90:
 * 12-113): mirror copy ghost g <- interior g over the transverse interior
 * range, on the reconstructed (primitive) fields; REFLECTING also negates the
 * velocity component along d.  PERIODIC is skipped. */
static void apply_bc(const gxo_ctx *c, int d) {
  for (int side = 0; side < 2; side++) {
    int bc = c->bc[d][side];
    if (bc == GXO_BC_PERIODIC) continue;
    int jd = (d + 1) % 3, kd = (jd + 1) % 3;
    int n = c->n[d];
    for (int f = 0; f < 5; f++) {
      double *V = c->cur[IA_PRIMITIVE[f]];
      int reverse = (bc == GXO_BC_REFLECTING) && (IA_PRIMITIVE[f] == GXO_V_1 + d);
#pragma omp parallel for collapse(2) schedule(static)
      for (int kk = 0; kk < c->n[kd]; kk++)
        for (int jj = 0; jj < c->n[jd]; jj++)
          for (int ig = 1; ig <= NG; ig++) {
            int pe[3], pi[3];
            pe[jd] = pi[jd] = jj + NG;
            pe[kd] = pi[kd] = kk + NG;
            if (side == 0) { /* inner: ghost (1-ig) <- cell ig (1-based) */
              pi[d] = NG + ig - 1;
              pe[d] = NG - ig;
            } else { /* outer: ghost (n+ig) <- cell (n+1-ig) */
              pi[d] = NG + n - ig;
              pe[d] = NG + n - 1 + ig;
            }
            long e = IDX(c, pe[0], pe[1], pe[2]);
            V[e] = V[IDX(c, pi[0], pi[1], pi[2])];
            if (reverse) V[e] = -V[e];
          }
    }
  }
}

/* ---- OP/Calculus/Derivatives/Gradient_Kernel.f90:166-190 -------------- */
double gxo_limited_slope(double dV_L, double dV_R, double dX_L, double dX_R,
                         double theta) {
  return (copysign(0.5, dV_L) + copysign(0.5, dV_R)) *
         fmin(fmin(fabs(theta * dV_L / dX_L), fabs(theta * dV_R / dX_R)),
              fabs((dX_R * dX_R * dV_L + dX_L * dX_L * dV_R) /
                   (dX_L * dX_R * (dX_L + dX_R))));
}

static inline double unlimited_slope(double dV_L, double dV_R, double dX_L,
                                     double dX_R) {
  return (dX_R * dX_R * dV_L + dX_L * dX_L * dV_R) /
         (dX_L * dX_R * (dX_L + dX_R));
}

/* Difference_Kernel (OP/Calculus/Derivatives/Difference_Kernel.f90:12-81):
 * dV_I(p) = V(p) - V(p - e_d), p_d in [1, m_d), transverse interior. */
static void difference(const gxo_ctx *c, const double *V, double *dV, int d) {
  int lo[3], hi[3];
  for (int e = 0; e < 3; e++) { lo[e] = NG; hi[e] = c->m[e] - NG; }
  lo[d] = NG - 1;
  hi[d] = c->m[d];
  long s = (d == 0) ? 1 : (d == 1) ? c->m[0] : (long)c->m[0] * c->m[1];
  LOOP3(c, lo, hi) {
    long iv = IDX(c, p0, p1, p2);
    dV[iv] = V[iv] - V[iv - s];
  }
}

/* Gradient_Form % ComputeChart_SL_G (OP/Calculus/Derivatives/Gradient_Form.
 * f90:120-203): coordinate difference, variable difference, limited slope for
 * p_d in [NG-1, m_d-NG], transverse interior. */
static void gradient(gxo_ctx *c, int d) {
  int lo[3], hi[3];
  for (int e = 0; e < 3; e++) { lo[e] = NG; hi[e] = c->m[e] - NG; }
  /* Center_U(d) as a 3-D field (Chart_SL__Template.f90:199-227) */
  {
    int alo[3] = {0, 0, 0};
    const int *ahi = c->m;
    LOOP3(c, alo, ahi) {
      int p[3] = {p0, p1, p2};
      c->xc[IDX(c, p0, p1, p2)] = c->center[d][p[d]];
    }
  }
  difference(c, c->xc, c->dxi, d);
  for (int f = 0; f < 5; f++) difference(c, c->cur[IA_PRIMITIVE[f]], c->dvi[f], d);
  lo[d] = NG - 1;
  hi[d] = c->m[d] - NG + 1;
  long s = (d == 0) ? 1 : (d == 1) ? c->m[0] : (long)c->m[0] * c->m[1];
  double theta = c->theta;
  for (int f = 0; f < 5; f++) {
    const double *dv = c->dvi[f];
    double *g = c->grad[f];
    const double *dx = c->dxi;
    int lim = c->use_limiter;
    LOOP3(c, lo, hi) {
      long iv = IDX(c, p0, p1, p2);
      double dV_L = dv[iv], dV_R = dv[iv + s], dX_L = dx[iv], dX_R = dx[iv + s];
      g[iv] = lim ? gxo_limited_slope(dV_L, dV_R, dX_L, dX_R, theta)
                  : unlimited_slope(dV_L, dV_R, dX_L, dX_R);
    }
  }
}

/* ComputeReconstructionParabolic_CSL_Kernel (EV/Increments/IncrementDivergence_
 * FV__Kernel.f90:296-596) for one cell, with XA = X and X2A = X**2 as
 * ComputeReconstructionParabolic_CSL passes them (IncrementDivergence_FV__
 * Form.f90:852-931): the cell's inner and outer face values vI, vO from the
 * values and centres of the cell and its two neighbours.  A local extremum of
 * the three values gives the constant vC; otherwise the parabola through the
 * three points, replaced on the side of an interior vertex by the parabola
 * flat at the neighbour; then the parabola through (xI, vI), (xC, vC),
 * (xO, vO), likewise replaced by one flat at xI or xO.  Fortran's a*b*c is
 * (a*b)*c, a+b+c is (a+b)+c, x**2 is x*x; -2*a*b/D is -((2*a*b)/D), which
 * rounds as ((-2*a)*b)/D. */
void gxo_parabolic_face_values(double vM, double vC, double vP, double xM,
                               double xC, double xP, double dX_L, double dX_R,
                               double out[2]) {
  const double sqrt_tiny = DBL_MIN; /* SqrtTiny = tiny ( 0.0_KDR ) */
  const double xAM = xM, xAC = xC, xAP = xP;
  const double x2AM = xM * xM, x2AC = xC * xC, x2AP = xP * xP;
  const double xI = xC - dX_L, xO = xC + dX_R;
  double vI, vO, D, c0, c1, c2, c2s, xE;
  if ((vC - vM) * (vP - vC) <= 0.0) {
    vI = vC;
    vO = vC;
  } else {
    /* first parabola */
    D = ((x2AM - x2AP) * xAC + (x2AP - x2AC) * xAM) + (x2AC - x2AM) * xAP;
    c0 = ((vP * (x2AM * xAC - x2AC * xAM) + vM * (x2AC * xAP - x2AP * xAC)) +
          vC * (x2AP * xAM - x2AM * xAP)) / D;
    c1 = ((vP * (x2AC - x2AM) + vC * (x2AM - x2AP)) + vM * (x2AP - x2AC)) / D;
    c2 = ((vP * (xAM - xAC) + vM * (xAC - xAP)) + vC * (xAP - xAM)) / D;
    c2s = copysign(fmax(fabs(c2), sqrt_tiny), c2);
    xE = -c1 / (2.0 * c2s);
    vI = (c0 + c1 * xI) + c2 * (xI * xI);
    vO = (c0 + c1 * xO) + c2 * (xO * xO);
    if (xE > xM && xE <= xC) {
      D = (x2AC - x2AM) + (2 * (xAM - xAC)) * xAM;
      c0 = ((vM * x2AC - vC * x2AM) + (2 * (vC * xAM - vM * xAC)) * xAM) / D;
      c1 = ((-2 * (vC - vM)) * xAM) / D;
      c2 = (vC - vM) / D;
      vI = (c0 + c1 * xI) + c2 * (xI * xI);
    } else if (xE > xC && xE < xP) {
      D = (x2AC - x2AP) + (2 * (xAP - xAC)) * xAP;
      c0 = ((vP * x2AC - vC * x2AP) + (2 * (vC * xAP - vP * xAC)) * xAP) / D;
      c1 = ((-2 * (vC - vP)) * xAP) / D;
      c2 = (vC - vP) / D;
      vO = (c0 + c1 * xO) + c2 * (xO * xO);
    }
    /* second parabola */
    D = (xI - xO) * ((x2AC + xI * xO) - xAC * (xI + xO));
    c0 = (((vO * xI) * (x2AC - xAC * xI) + ((vC * xI) * xO) * (xI - xO)) +
          (vI * xO) * (xAC * xO - x2AC)) / D;
    c1 = ((vO * (xI * xI - x2AC) + vI * (x2AC - xO * xO)) + vC * (xO * xO - xI * xI)) / D;
    c2 = ((vO * (xAC - xI) + vC * (xI - xO)) + vI * (xO - xAC)) / D;
    c2s = copysign(fmax(fabs(c2), sqrt_tiny), c2);
    xE = -c1 / (2.0 * c2s);
    vI = (c0 + c1 * xI) + c2 * (xI * xI);
    vO = (c0 + c1 * xO) + c2 * (xO * xO);
    if (xE > xI && xE <= xC) {
      D = (x2AC - xO * xO) + (2 * xI) * (xO - xAC);
      c0 = ((vO * x2AC - vC * (xO * xO)) + (2 * xI) * (vC * xO - vO * xAC)) / D;
      c1 = ((-2 * (vC - vO)) * xI) / D;
      c2 = (vC - vO) / D;
      vI = (c0 + c1 * xI) + c2 * (xI * xI);
    } else if (xE > xC && xE < xO) {
      D = (x2AC - xI * xI) + (2 * xO) * (xI - xAC);
      c0 = ((vI * x2AC - vC * (xI * xI)) + (2 * xO) * (vC * xI - vI * xAC)) / D;
      c1 = ((-2 * (vC - vI)) * xO) / D;
      c2 = (vC - vI) / D;
      vO = (c0 + c1 * xO) + c2 * (xO * xO);
    }
  }
  out[0] = vI;
  out[1] = vO;
}

/* The parabolic reconstruction of dimension d: cells p_d in [NG-1, m_d-NG]
 * (the proper cells and one ghost on each side), transverse interior;
 * V_IR(p) = vI, V_IL(p+e_d) = vO. */
static void reconstruct_parabolic(gxo_ctx *c, int d) {
  int lo[3], hi[3];
  for (int e = 0; e < 3; e++) { lo[e] = NG; hi[e] = c->m[e] - NG; }
  lo[d] = NG - 1;
  hi[d] = c->m[d] - NG + 1;
  long s = (d == 0) ? 1 : (d == 1) ? c->m[0] : (long)c->m[0] * c->m[1];
  const double *X = c->center[d], *WL = c->wl[d], *WR = c->wr[d];
  for (int f = 0; f < 5; f++) {
    const double *V = c->cur[IA_PRIMITIVE[f]];
    double *VL = c->il[IA_PRIMITIVE[f]], *VR = c->ir[IA_PRIMITIVE[f]];
    LOOP3(c, lo, hi) {
      int p[3] = {p0, p1, p2};
      long iv = IDX(c, p0, p1, p2);
      const int q = p[d];
      double out[2];
      gxo_parabolic_face_values(V[iv - s], V[iv], V[iv + s], X[q - 1], X[q], X[q + 1],
                                WL[q], WR[q], out);
      VR[iv] = out[0];
      VL[iv + s] = out[1];
    }
  }
}

/* ComputeReconstructionLinear_CSL_Kernel (EV/Increments/IncrementDivergence_
 * FV__Kernel.f90:11-94): V_IL(p) = V(p-e) + dX_R(p-e)*dVdX(p-e),
 * V_IR(p) = V(p) - dX_L(p)*dVdX(p); p_d in [NG, m_d-NG], transverse interior. */
static void reconstruct(gxo_ctx *c, int d) {
  int lo[3], hi[3];
  for (int e = 0; e < 3; e++) { lo[e] = NG; hi[e] = c->m[e] - NG; }
  hi[d] = c->m[d] - NG + 1;
  long s = (d == 0) ? 1 : (d == 1) ? c->m[0] : (long)c->m[0] * c->m[1];
  const double *WL = c->wl[d], *WR = c->wr[d];
  for (int f = 0; f < 5; f++) {
    const double *V = c->cur[IA_PRIMITIVE[f]];
    const double *g = c->grad[f];
    double *VL = c->il[IA_PRIMITIVE[f]], *VR = c->ir[IA_PRIMITIVE[f]];
    LOOP3(c, lo, hi) {
      int p[3] = {p0, p1, p2};
      long iv = IDX(c, p0, p1, p2);
      VL[iv] = V[iv - s] + WR[p[d] - 1] * g[iv - s];
      VR[iv] = V[iv] - WL[p[d]] * g[iv];
    }
  }
}

/* Face range of a sweep in dimension d. */
static void face_range(const gxo_ctx *c, int d, int lo[3], int hi[3]) {
  for (int e = 0; e < 3; e++) { lo[e] = NG; hi[e] = c->m[e] - NG; }
  hi[d] = c->m[d] - NG + 1;
}

/* ComputeRawFluxes: Fluid_D ComputeRawFluxes_G_Kernel (F/Fluid_D__Kernel.f90:
 * 220-262) then ComputeRawFluxesTemplate_P_Kernel (F/Fluid_P__Kernel.f90:
 * 671-709).  Output order D, S_1, S_2, S_3, G, DS. */
static inline void raw_flux(double **F, long iv, int d, double **RF) {
  double Vd = F[GXO_V_1 + d][iv];
  RF[0][iv] = F[GXO_D][iv] * Vd;
  RF[1][iv] = F[GXO_S_1][iv] * Vd;
  RF[2][iv] = F[GXO_S_2][iv] * Vd;
  RF[3][iv] = F[GXO_S_3][iv] * Vd;
  RF[1 + d][iv] = RF[1 + d][iv] + F[GXO_P][iv];
  RF[4][iv] = (F[GXO_G][iv] + F[GXO_P][iv]) * Vd;
  RF[5][iv] = F[GXO_DS][iv] * Vd;
}

/* Fluid_P_Template % ComputeFluxes (F/Fluid_P__Template.f90:400-531) with
 * CurrentTemplate % ComputeFluxes_HLL (FLD/Current_Template.f90:609-675,
 * FLD/Current_Kernel.f90:50-137) and, for HLLC, ComputeCenterSpeedKernel,
 * ComputeCenterStatesKernel and ComputeFluxes_HLLC_Kernel
 * (F/Fluid_P__Kernel.f90:319-668). */
static void fluxes(gxo_ctx *c, int d) {
  int lo[3], hi[3];
  face_range(c, d, lo, hi);
  const double tiny = DBL_MIN, sqrt_tiny = sqrt(DBL_MIN);
  double **IL = c->il, **IR = c->ir, **ICL = c->icl, **ICR = c->icr;
  double **FIL = c->fil, **FIR = c->fir, **FI = c->fi;
  double *AP = c->ap, *AM = c->am, *AC = c->ac;
  const double *DFd = c->df[d];
  int hllc = (c->riemann == GXO_HLLC);
  LOOP3(c, lo, hi) {
    long iv = IDX(c, p0, p1, p2);
    raw_flux(IL, iv, d, FIL);
    raw_flux(IR, iv, d, FIR);
    /* ComputeSolverSpeeds_HLL_Kernel */
    AP[iv] = fmax(fmax(0.0, +IL[GXO_FEP_1 + d][iv]), +IR[GXO_FEP_1 + d][iv]);
    AM[iv] = fmax(fmax(0.0, -IL[GXO_FEM_1 + d][iv]), -IR[GXO_FEM_1 + d][iv]);
    /* ComputeFluxes_HLL_Kernel with DF = 1 (SetDiffusionFactorUnity) */
    for (int f = 0; f < 6; f++) {
      double ap = AP[iv], am = AM[iv], df = 1.0;
      FI[f][iv] = (ap * FIL[f][iv] + am * FIR[f][iv] -
                   df * ap * am * (IR[IA_CONSERVED[f]][iv] - IL[IA_CONSERVED[f]][iv])) /
                  fmax(ap + am, tiny);
    }
    if (!hllc) continue;
    /* ComputeCenterSpeedKernel, M_UU = 1 */
    {
      double ap = AP[iv], am = AM[iv];
      double MD_Numerator = ap * IR[GXO_M][iv] * IR[GXO_D][iv] +
                            am * IL[GXO_M][iv] * IL[GXO_D][iv] - FIR[0][iv] +
                            FIL[0][iv];
      double S_Numerator = ap * IR[GXO_S_1 + d][iv] + am * IL[GXO_S_1 + d][iv] -
                           FIR[1 + d][iv] + FIL[1 + d][iv];
      double MD_Numerator_Inv =
          fmax(MD_Numerator, 0.0) / fmax(MD_Numerator * MD_Numerator, tiny);
      AC[iv] = 1.0 * S_Numerator * MD_Numerator_Inv;
    }
    /* ComputeCenterStatesKernel */
    {
      double ap = AP[iv], am = AM[iv], ac = AC[iv];
      ICL[GXO_V_1][iv] = IL[GXO_V_1][iv];
      ICR[GXO_V_1][iv] = IR[GXO_V_1][iv];
      ICL[GXO_V_2][iv] = IL[GXO_V_2][iv];
      ICR[GXO_V_2][iv] = IR[GXO_V_2][iv];
      ICL[GXO_V_3][iv] = IL[GXO_V_3][iv];
      ICR[GXO_V_3][iv] = IR[GXO_V_3][iv];
      ICL[GXO_V_1 + d][iv] = ac;
      ICR[GXO_V_1 + d][iv] = ac;
      double AM_VL = am + IL[GXO_V_1 + d][iv];
      double AM_AC = am + ac;
      double AM_AC_Inv = 1.0 / fmax(fabs(AM_AC), sqrt_tiny);
      double AP_VR = ap - IR[GXO_V_1 + d][iv];
      double AP_AC = ap - ac;
      double AP_AC_Inv = 1.0 / fmax(fabs(AP_AC), sqrt_tiny);
      ICL[GXO_M][iv] = IL[GXO_M][iv];
      ICR[GXO_M][iv] = IR[GXO_M][iv];
      ICL[GXO_D][iv] = IL[GXO_D][iv] * AM_VL * AM_AC_Inv;
      ICR[GXO_D][iv] = IR[GXO_D][iv] * AP_VR * AP_AC_Inv;
      for (int k = 0; k < 3; k++) {
        ICL[GXO_S_1 + k][iv] = ICL[GXO_M][iv] * ICL[GXO_D][iv] * ICL[GXO_V_1 + k][iv];
        ICR[GXO_S_1 + k][iv] = ICR[GXO_M][iv] * ICR[GXO_D][iv] * ICR[GXO_V_1 + k][iv];
      }
      ICL[GXO_P][iv] = IL[GXO_P][iv] + IL[GXO_S_1 + d][iv] * AM_VL -
                       ICL[GXO_S_1 + d][iv] * AM_AC;
      ICR[GXO_P][iv] = IR[GXO_P][iv] - IR[GXO_S_1 + d][iv] * AP_VR +
                       ICR[GXO_S_1 + d][iv] * AP_AC;
      ICL[GXO_G][iv] = (IL[GXO_G][iv] * AM_VL + IL[GXO_V_1 + d][iv] * IL[GXO_P][iv] -
                        ac * ICL[GXO_P][iv]) * AM_AC_Inv;
      ICR[GXO_G][iv] = (IR[GXO_G][iv] * AP_VR - IR[GXO_V_1 + d][iv] * IR[GXO_P][iv] +
                        ac * ICR[GXO_P][iv]) * AP_AC_Inv;
      ICL[GXO_DS][iv] = IL[GXO_DS][iv] * AM_VL * AM_AC_Inv;
      ICR[GXO_DS][iv] = IR[GXO_DS][iv] * AP_VR * AP_AC_Inv;
    }
    /* Overwrite F_IL and F_IR with the raw fluxes of the center states */
    raw_flux(ICL, iv, d, FIL);
    raw_flux(ICR, iv, d, FIR);
    /* ComputeFluxes_HLLC_Kernel */
    for (int f = 0; f < 6; f++) {
      if (DFd[iv] > 0.0) continue;
      if (AP[iv] != 0.0 && AM[iv] != 0.0) {
        if (AC[iv] >= 0.0)
          FI[f][iv] = FIL[f][iv];
        else
          FI[f][iv] = FIR[f][iv];
      }
    }
  }
}

/* ComputeIncrement_CSL_Kernel (EV/Increments/IncrementDivergence_FV__Kernel.
 * f90:672-752) over proper cells, with the rectangular volume and inner-face
 * areas of SetFiniteVolumeRectangularKernel (CH/ChartBasics/GeometryFlat_
 * Kernel.f90:12-47). */
static void increment(gxo_ctx *c, double **K, double dt, int d) {
  int lo[3], hi[3];
  for (int e = 0; e < 3; e++) { lo[e] = NG; hi[e] = c->m[e] - NG; }
  long s = (d == 0) ? 1 : (d == 1) ? c->m[0] : (long)c->m[0] * c->m[1];
  for (int f = 0; f < 6; f++) {
    double *dU = K[f];
    const double *F = c->fi[f];
    LOOP3(c, lo, hi) {
      int p[3] = {p0, p1, p2};
      long iv = IDX(c, p0, p1, p2);
      double dX = c->wl[0][p0] + c->wr[0][p0];
      double dY = c->wl[1][p1] + c->wr[1][p1];
      double dZ = c->wl[2][p2] + c->wr[2][p2];
      double V = dX * dY * dZ;
      double A;
      if (d == 0) A = dY * dZ;
      else if (d == 1) A = dZ * dX;
      else A = dX * dY;
      (void)p;
      dU[iv] = dU[iv] + dt * (A * F[iv] - A * F[iv + s]) / V;
    }
  }
}

/* RecordBoundaryFluence_CSL (EV/Increments/IncrementDivergence_FV__Form.f90:
 * 1002-1069) with its kernel (IncrementDivergence_FV__Kernel.f90:755-800):
 * BF = BF + (Weight_RK*dT) * F_I over the transverse proper cells, F_I taken
 * on the inner face of the first proper layer (inner boundary) and of the
 * first outer ghost layer (outer boundary) of dimension d. */
static void record_fluence(gxo_ctx *c, double dt, int d) {
  const double factor = c->weight_rk * dt;
  int nb[3];
  for (int e = 0; e < 3; e++) nb[e] = (e == d) ? 1 : c->n[e];
  for (int side = 0; side < 2; side++) {
    int ob[3] = {NG, NG, NG};
    if (side) ob[d] += c->n[d];
    for (int f = 0; f < 5; f++) {
      double *B = c->bf[2 * d + side][f];
      const double *F = c->fi[f];
      long i = 0;
      for (int k = 0; k < nb[2]; k++)
        for (int j = 0; j < nb[1]; j++)
          for (int l = 0; l < nb[0]; l++, i++)
            B[i] = B[i] + factor * F[IDX(c, ob[0] + l, ob[1] + j, ob[2] + k)];
    }
  }
}

/* IncrementDivergence_FV_Form % Compute (EV/Increments/IncrementDivergence_FV__
 * Form.f90:344-412, 582-686, 689-738, 934-999). */
static void increment_divergence(gxo_ctx *c, double **K, double dt) {
  for (int d = 0; d < 3; d++) {
    apply_bc(c, d);
    if (c->parabolic) {
      reconstruct_parabolic(c, d);
    } else {
      gradient(c, d);
      reconstruct(c, d);
    }
    int lo[3], hi[3];
    face_range(c, d, lo, hi);
    LOOP3(c, lo, hi) {
      long iv = IDX(c, p0, p1, p2);
      from_primitive_value(c, c->il, iv);
      from_primitive_value(c, c->ir, iv);
    }
    fluxes(c, d);
    increment(c, K, dt, d);
    record_fluence(c, dt, d);
  }
}

/* The Tally_F_P integrands of one value, in Value order (GXO_TALLY_*).
 * Interior (Tally_F_D__Form.f90:248-384, 737-782; Tally_F_P__Form.f90:
 * 147-187): D, S, KE = 0.5 (S.V), L = X x S, G, E; boundary (Tally_F_D__Form.
 * f90:387-480, 1007-1034; Tally_F_P__Form.f90:190-239): the fluences of D, S
 * and G with L = X x S from the momentum fluences; KE and E have no boundary
 * integrand.  ComputeDensity_AM_Rectangular (X_J, X_K, S_J, S_K) is
 * X_J*S_K - X_K*S_J with (J, K) = (2, 3), (3, 1), (1, 2). */
static void tally_integrands(const double X[3], double D, const double S[3],
                             double G, const double *V, double E,
                             double out[GXO_TALLY_N]) {
  for (int t = 0; t < GXO_TALLY_N; t++) out[t] = 0.0;
  out[GXO_TALLY_BARYON_NUMBER] = D;
  if (V)
    out[GXO_TALLY_KINETIC_ENERGY] =
        0.5 * ((S[0] * V[0] + S[1] * V[1]) + S[2] * V[2]);
  for (int e = 0; e < 3; e++) out[GXO_TALLY_MOMENTUM_1 + e] = S[e];
  out[GXO_TALLY_ANGULAR_MOMENTUM_1] = X[1] * S[2] - X[2] * S[1];
  out[GXO_TALLY_ANGULAR_MOMENTUM_2] = X[2] * S[0] - X[0] * S[2];
  out[GXO_TALLY_ANGULAR_MOMENTUM_3] = X[0] * S[1] - X[1] * S[0];
  out[GXO_TALLY_FLUID_ENERGY] = G;
  if (V) out[GXO_TALLY_INTERNAL_ENERGY] = E;
}

/* SurfaceIntegral_Form % Compute_CSL (OP/Calculus/Integrals/SurfaceIntegral_
 * Form.f90:25-146, 150-177) of the boundary integrands over the step's
 * fluence: I = I + Direction * dIdA * dA, Direction -1 on inner faces and +1
 * on outer ones, dA the inner-face area of the slab's cell (GeometryFlat_
 * Kernel.f90:12-47) and X the face positions of Tally_C_Form %
 * ComputeFacePositions (transverse centres; Center - WidthLeft along d). */
static void surface_integral(const gxo_ctx *c, double out[GXO_TALLY_N]) {
  for (int t = 0; t < GXO_TALLY_N; t++) out[t] = 0.0;
  for (int d = 0; d < 3; d++) {
    int nb[3];
    for (int e = 0; e < 3; e++) nb[e] = (e == d) ? 1 : c->n[e];
    for (int side = 0; side < 2; side++) {
      const double dir = side ? 1.0 : -1.0;
      int ob[3] = {NG, NG, NG};
      if (side) ob[d] += c->n[d];
      double *const *B = c->bf[2 * d + side];
      long i = 0;
      for (int k = 0; k < nb[2]; k++)
        for (int j = 0; j < nb[1]; j++)
          for (int l = 0; l < nb[0]; l++, i++) {
            const int p[3] = {ob[0] + l, ob[1] + j, ob[2] + k};
            double X[3], w[3], v[GXO_TALLY_N];
            for (int e = 0; e < 3; e++) {
              X[e] = c->center[e][p[e]];
              w[e] = c->wl[e][p[e]] + c->wr[e][p[e]];
            }
            X[d] = X[d] - c->wl[d][p[d]];
            const double dA = d == 0 ? w[1] * w[2] : d == 1 ? w[2] * w[0] : w[0] * w[1];
            const double S[3] = {B[1][i], B[2][i], B[3][i]};
            tally_integrands(X, B[0][i], S, B[4][i], NULL, 0.0, v);
            for (int t = 0; t < GXO_TALLY_N; t++) out[t] = out[t] + dir * v[t] * dA;
          }
    }
  }
}

/* VolumeIntegral_Form % Compute_CSL (OP/Calculus/Integrals/VolumeIntegral_
 * Form.f90:25-117) of the interior integrands over the proper cells of the
 * Current: I = I + dIdV * dV, dV = (dX*dY)*dZ. */
static void volume_integral(const gxo_ctx *c, double out[GXO_TALLY_N]) {
  for (int t = 0; t < GXO_TALLY_N; t++) out[t] = 0.0;
  double *const *F = c->cur;
  for (int p2 = NG; p2 < c->m[2] - NG; p2++)
    for (int p1 = NG; p1 < c->m[1] - NG; p1++)
      for (int p0 = NG; p0 < c->m[0] - NG; p0++) {
        const long iv = IDX(c, p0, p1, p2);
        const int p[3] = {p0, p1, p2};
        double X[3], w[3], v[GXO_TALLY_N];
        for (int e = 0; e < 3; e++) {
          X[e] = c->center[e][p[e]];
          w[e] = c->wl[e][p[e]] + c->wr[e][p[e]];
        }
        const double dV = w[0] * w[1] * w[2];
        const double S[3] = {F[GXO_S_1][iv], F[GXO_S_2][iv], F[GXO_S_3][iv]};
        const double V[3] = {F[GXO_V_1][iv], F[GXO_V_2][iv], F[GXO_V_3][iv]};
        tally_integrands(X, F[GXO_D][iv], S, F[GXO_G][iv], V, F[GXO_E][iv], v);
        for (int t = 0; t < GXO_TALLY_N; t++) out[t] = out[t] + v[t] * dV;
      }
}

static void reset_tallies(gxo_ctx *c) {
  for (int t = 0; t < GXO_TALLY_N; t++)
    c->tally_local[t] = c->tally_global[t] = c->tally_total[t] = c->tally_change[t] = 0.0;
  for (int b = 0; b < 6; b++)
    for (int f = 0; f < 5; f++) memset(c->bf[b][f], 0, sizeof(double) * c->bf_n[b / 2]);
}

static void clear6(const gxo_ctx *c, double **A) {
  for (int f = 0; f < 6; f++) memset(A[f], 0, sizeof(double) * c->nv);
}

static void copy_conserved_to(const gxo_ctx *c, double **dst) {
  for (int f = 0; f < 6; f++)
    memcpy(dst[f], c->cur[IA_CONSERVED[f]], sizeof(double) * c->nv);
}

static void copy_conserved_from(gxo_ctx *c, double **src) {
  for (int f = 0; f < 6; f++)
    memcpy(c->cur[IA_CONSERVED[f]], src[f], sizeof(double) * c->nv);
}

/* MultiplyAdd in place, A = A + C*B (OP/Algebra/MultiplyAdd_Command.f90:56-94) */
static void multiply_add(const gxo_ctx *c, double **A, double **B, double C) {
  long nv = c->nv;
  for (int f = 0; f < 6; f++) {
    double *a = A[f];
    const double *b = B[f];
#pragma omp parallel for schedule(static)
    for (long iv = 0; iv < nv; iv++) a[iv] = a[iv] + C * b[iv];
  }
}

/* ---- FluidFeatures_P % Detect (F/FluidFeatures_P__Form.f90:145-171,
 * 267-344; F/FluidFeatures_P__Kernel.f90:12-313).  Only DiffusiveFlux_I is
 * kept: Shock and Shock_I feed nothing on the IDEAL path. */
static void detect_features(gxo_ctx *c) {
  for (int d = 0; d < 3; d++) memset(c->df[d], 0, sizeof(double) * c->nv);
  const double sqrt_tiny = sqrt(DBL_MIN);
  const double *P = c->cur[GXO_P];
  long st[3] = {1, c->m[0], (long)c->m[0] * c->m[1]};
  for (int id = 0; id < 3; id++) {
    int jd = (id + 1) % 3, kd = (jd + 1) % 3;
    const double *Vd = c->cur[GXO_V_1 + id];
    double *DFj = c->df[jd], *DFk = c->df[kd];
    int lo[3], hi[3];
    for (int e = 0; e < 3; e++) { lo[e] = NG - 1; hi[e] = c->m[e] - NG + 1; }
    /* serial scan: the flag writes are idempotent stores of 1.0 */
    for (int p2 = lo[2]; p2 < hi[2]; p2++)
      for (int p1 = lo[1]; p1 < hi[1]; p1++)
        for (int p0 = lo[0]; p0 < hi[0]; p0++) {
          long iv = IDX(c, p0, p1, p2);
          long ivi = iv - st[id];
          double dP = fabs(P[iv] - P[ivi]);
          double P_Min = fmax(fmin(P[iv], P[ivi]), sqrt_tiny);
          double dLnP = dP / P_Min;
          double dV_iD = Vd[iv] - Vd[ivi];
          if (dLnP > c->st && dV_iD <= 0.0) {
            DFj[iv] = 1.0;
            DFj[ivi] = 1.0;
            DFj[iv + st[jd]] = 1.0;
            DFj[ivi + st[jd]] = 1.0;
            DFk[iv] = 1.0;
            DFk[ivi] = 1.0;
            DFk[iv + st[kd]] = 1.0;
            DFk[ivi + st[kd]] = 1.0;
          }
        }
  }
  /* ClearBoundary_CSL_Kernel on one brick: inner and outer faces of every
   * dimension, whether periodic or not. */
  for (int id = 0; id < 3; id++) {
    int lo[3], hi[3];
    /* inner: p_id in {NG-1, NG}, others [NG-1, m-NG] */
    for (int e = 0; e < 3; e++) { lo[e] = NG - 1; hi[e] = c->m[e] - NG + 1; }
    hi[id] = lo[id] + 2;
    for (int p2 = lo[2]; p2 < hi[2]; p2++)
      for (int p1 = lo[1]; p1 < hi[1]; p1++)
        for (int p0 = lo[0]; p0 < hi[0]; p0++) {
          long iv = IDX(c, p0, p1, p2);
          for (int e = 0; e < 3; e++) c->df[e][iv] = 0.0;
        }
    /* outer: p_id in {m-NG-1, m-NG}, others [NG-2, m-NG] */
    for (int e = 0; e < 3; e++) { lo[e] = NG - 2; hi[e] = c->m[e] - NG + 1; }
    lo[id] = c->m[id] - NG - 1;
    for (int p2 = lo[2]; p2 < hi[2]; p2++)
      for (int p1 = lo[1]; p1 < hi[1]; p1++)
        for (int p0 = lo[0]; p0 < hi[0]; p0++) {
          long iv = IDX(c, p0, p1, p2);
          for (int e = 0; e < 3; e++) c->df[e][iv] = 0.0;
        }
  }
}

/* ---- public entry points ---------------------------------------------- */

void gxo_set_primitive(gxo_ctx *c, const double *N, const double *V1,
                       const double *V2, const double *V3, const double *E) {
  size_t b = sizeof(double) * c->nv;
  memcpy(c->cur[GXO_N], N, b);
  memcpy(c->cur[GXO_V_1], V1, b);
  memcpy(c->cur[GXO_V_2], V2, b);
  memcpy(c->cur[GXO_V_3], V3, b);
  memcpy(c->cur[GXO_E], E, b);
  /* PrepareEvolution (EV/Integrators/Integrator_C_PS__Form.f90:175-224) */
  double *prim[5] = {c->cur[GXO_N], c->cur[GXO_V_1], c->cur[GXO_V_2],
                     c->cur[GXO_V_3], c->cur[GXO_E]};
  exchange_ghosts(c, prim, 5);
  from_primitive_all(c, c->cur);
  for (int d = 0; d < 3; d++) memset(c->df[d], 0, b);
  reset_tallies(c);
}

void gxo_set_conserved(gxo_ctx *c, const double *D, const double *S1,
                       const double *S2, const double *S3, const double *G) {
  size_t b = sizeof(double) * c->nv;
  memcpy(c->cur[GXO_D], D, b);
  memcpy(c->cur[GXO_S_1], S1, b);
  memcpy(c->cur[GXO_S_2], S2, b);
  memcpy(c->cur[GXO_S_3], S3, b);
  memcpy(c->cur[GXO_G], G, b);
  from_conserved_all(c, c->cur, 1);
}

/* Integrator_C_PS % ComputeTimeStep_C_ASC (EV/Integrators/Integrator_C_PS__
 * Form.f90:374-423) and ComputeTimeStepKernel_CSL (…Integrator_C_PS__Kernel.
 * f90:12-144), nDimensions = 3, Crsn = 1. */
double gxo_time_step_candidate(gxo_ctx *c, double *invdt_out) {
  double tsi = -DBL_MAX;
  int lo[3], hi[3];
  for (int e = 0; e < 3; e++) { lo[e] = NG; hi[e] = c->m[e] - NG; }
  double *const *F = c->cur;
#pragma omp parallel for collapse(3) schedule(static) reduction(max : tsi)
  for (int p2 = lo[2]; p2 < hi[2]; p2++)
    for (int p1 = lo[1]; p1 < hi[1]; p1++)
      for (int p0 = lo[0]; p0 < hi[0]; p0++) {
        long iv = IDX(c, p0, p1, p2);
        double v = fmax(F[GXO_FEP_1][iv], -F[GXO_FEM_1][iv]) /
                       (c->wl[0][p0] + c->wr[0][p0]) +
                   fmax(F[GXO_FEP_2][iv], -F[GXO_FEM_2][iv]) /
                       (c->wl[1][p1] + c->wr[1][p1]) +
                   fmax(F[GXO_FEP_3][iv], -F[GXO_FEM_3][iv]) /
                       (c->wl[2][p2] + c->wr[2][p2]);
        tsi = fmax(tsi, v);
      }
  if (invdt_out) *invdt_out = tsi;
  tsi = fmax(DBL_MIN, tsi);
  double ts = fmin(DBL_MAX, 1.0 / tsi);
  return c->courant * ts;
}

/* ApplyDivergence_C's recording (EV/Steps/Step_RK_C_ASC__Template.f90:
 * 1572-1589) with RecordDivergenceKernel (EV/Steps/Step_RK_C_ASC__Kernel.f90:
 * 12-46): Sources is cleared in stage 1, then Sources = Sources + Weight_RK *
 * Increment / TimeStep over all values, after the increment and before the
 * ghost exchange of K. */
static void record_divergence(gxo_ctx *c, double **K, double dt, int stage) {
  for (int f = 0; f < 6; f++) {
    double *s = c->sources[f];
    if (stage == 1) memset(s, 0, sizeof(double) * c->nv);
    for (long v = 0; v < c->nv; v++) s[v] = s[v] + c->weight_rk * K[f][v] / dt;
  }
}

void gxo_get_sources(const gxo_ctx *c, int f, double *out) {
  memcpy(out, c->sources[f], sizeof(double) * c->nv);
}

/* Step_RK_C_ASC % Compute -> Step_RK % ComputeTemplate (EV/Steps/Step_RK__
 * Template.f90:207-288; EV/Steps/Step_RK_C_ASC__Template.f90:504-542,
 * 692-816, 833-914, 1112-1281) with the Heun tableau of Step_RK2_C_ASC
 * (EV/Steps/Step_RK2_C_ASC__Form.f90:51-57): A21 = 1, B = (1/2, 1/2). */
void gxo_step(gxo_ctx *c, double dt) {
  /* ClearBoundaryFluence (EV/Steps/Step_RK_C_ASC__Template.f90:526-529) */
  for (int b = 0; b < 6; b++)
    for (int f = 0; f < 5; f++) memset(c->bf[b][f], 0, sizeof(double) * c->bf_n[b / 2]);
  /* LoadSolution */
  copy_conserved_to(c, c->sol);
  /* Stage 1: Y = Solution; K1 = L(Current)*dt; exchange K1 */
  for (int f = 0; f < 6; f++) memcpy(c->y[f], c->sol[f], sizeof(double) * c->nv);
  clear6(c, c->k1);
  c->weight_rk = 0.5; /* B(1) */
  increment_divergence(c, c->k1, dt);
  record_divergence(c, c->k1, dt, 1);
  exchange_ghosts(c, c->k1, 6);
  /* Stage 2: Y = Solution + 1*K1; Current <- Y; K2 = L(Current)*dt */
  for (int f = 0; f < 6; f++) memcpy(c->y[f], c->sol[f], sizeof(double) * c->nv);
  multiply_add(c, c->y, c->k1, 1.0);
  copy_conserved_from(c, c->y);
  from_conserved_all(c, c->cur, 0);
  clear6(c, c->k2);
  c->weight_rk = 0.5; /* B(2) */
  increment_divergence(c, c->k2, dt);
  record_divergence(c, c->k2, dt, 2);
  exchange_ghosts(c, c->k2, 6);
  /* IncrementSolution, K1 first */
  multiply_add(c, c->sol, c->k1, 0.5);
  multiply_add(c, c->sol, c->k2, 0.5);
  /* StoreSolution with DetectFeatures = .true. */
  copy_conserved_from(c, c->sol);
  from_conserved_all(c, c->cur, 1);
  detect_features(c);
  /* AccumulateBoundaryTally (Step_RK_C_ASC__Template.f90:533-537;
   * FLD/Current_ASC__Template.f90:203-219; FLD/Tally_C__Form.f90:215-279):
   * TallyBoundaryLocal = TallyBoundaryLocal + SurfaceIntegral */
  double si[GXO_TALLY_N];
  surface_integral(c, si);
  for (int t = 0; t < GXO_TALLY_N; t++) c->tally_local[t] = c->tally_local[t] + si[t];
}

void gxo_boundary_fluence(const gxo_ctx *c, int b, int f, double *out) {
  memcpy(out, c->bf[b][f], sizeof(double) * c->bf_n[b / 2]);
}

void gxo_set_boundary_fluence(gxo_ctx *c, int b, int f, const double *in) {
  memcpy(c->bf[b][f], in, sizeof(double) * c->bf_n[b / 2]);
}

void gxo_surface_integral(const gxo_ctx *c, double *out) { surface_integral(c, out); }

/* Current_ASC % ComputeTallyTemplate (FLD/Current_ASC__Template.f90:239-349),
 * one chart, one boundary (AtlasHeader_Form.f90:164), no reduction. */
void gxo_compute_tally(gxo_ctx *c, int compute_change, double *interior,
                       double *boundary, double *total, double *change) {
  double old_total[GXO_TALLY_N], in[GXO_TALLY_N];
  memcpy(old_total, c->tally_total, sizeof(old_total));
  volume_integral(c, in);
  for (int t = 0; t < GXO_TALLY_N; t++) {
    c->tally_global[t] = c->tally_global[t] + c->tally_local[t];
    c->tally_local[t] = 0.0;
    c->tally_total[t] = in[t];
    c->tally_total[t] = c->tally_total[t] + c->tally_global[t];
    if (compute_change)
      c->tally_change[t] = c->tally_change[t] + (c->tally_total[t] - old_total[t]);
  }
  if (interior) memcpy(interior, in, sizeof(in));
  if (boundary) memcpy(boundary, c->tally_global, sizeof(in));
  if (total) memcpy(total, c->tally_total, sizeof(in));
  if (change) memcpy(change, c->tally_change, sizeof(in));
}

void gxo_get_field(const gxo_ctx *c, int f, double *out) {
  memcpy(out, c->cur[f], sizeof(double) * c->nv);
}

void gxo_fallback_counts(gxo_ctx *c, long out[4], int reset) {
  for (int k = 0; k < 4; k++) out[k] = c->n_repaired[k / 2][k % 2];
  if (reset) memset(c->n_repaired, 0, sizeof(c->n_repaired));
}

void gxo_get_flags(const gxo_ctx *c, int d, double *out) {
  memcpy(out, c->df[d], sizeof(double) * c->nv);
}

void gxo_set_flags(gxo_ctx *c, int d, const double *in) {
  memcpy(c->df[d], in, sizeof(double) * c->nv);
}

/* One face, in the order of the passes above: a 1x1x1 chart whose first
 * face along d carries the given reconstructed states. */
void gxo_face_flux(const double prim_l[5], const double prim_r[5], int d,
                   double gamma, int riemann, double df, double f_out[5]) {
  int n[3] = {1, 1, 1};
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  int bc[6] = {0, 0, 0, 0, 0, 0};
  gxo_ctx *c = gxo_create(n, lo, hi, bc, gamma, 1.4, 1, riemann, 0.1, 0.7,
                          0.0, 0);
  if (!c) return;
  long iv = IDX(c, NG, NG, NG);
  for (int f = 0; f < 5; f++) {
    c->il[IA_PRIMITIVE[f]][iv] = prim_l[f];
    c->ir[IA_PRIMITIVE[f]][iv] = prim_r[f];
  }
  from_primitive_value(c, c->il, iv);
  from_primitive_value(c, c->ir, iv);
  c->df[d][iv] = df;
  /* The face range along d also holds the (zero) face p_d = NG+1; it only
   * reads its own zero states. */
  fluxes(c, d);
  for (int f = 0; f < 5; f++) f_out[f] = c->fi[f][iv];
  gxo_destroy(c);
}
