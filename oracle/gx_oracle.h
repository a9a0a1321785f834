/*
 * gx_oracle.h -- CPU restatement of the GenASiS Fluid_P_I finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path in genasis_amd/ and the timed CPU baseline of bench.py.  Nothing in the
 * product (genasis_amd/, include/) links, loads or calls it.
 *
 * It restates, pass by pass and in the reference's own expression order, the
 * unfused kernels of the reference's RK2 step on the reference's ghost-padded
 * storage layout (Value(nValues, nFields), 2 ghost layers, x fastest).
 */
#ifndef GX_ORACLE_H
#define GX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary condition codes (IC/Atlas_SC__Template.f90:244-324). */
#define GXO_BC_PERIODIC   0
#define GXO_BC_OUTFLOW    1
#define GXO_BC_REFLECTING 2

/* Riemann solver codes (F/Fluid_P__Template.f90:435-518). */
#define GXO_HLL  0
#define GXO_HLLC 1

/* Fluid_P_I field map, 0-based (Appendix of the field list in
 * FLD/Current_Template.f90:752-753, F/Fluid_D__Form.f90:650-654,
 * F/Fluid_P__Template.f90:718-725). */
enum {
  GXO_FEP_1 = 0, GXO_FEP_2, GXO_FEP_3,
  GXO_FEM_1, GXO_FEM_2, GXO_FEM_3,
  GXO_N, GXO_D, GXO_M,
  GXO_V_1, GXO_V_2, GXO_V_3,
  GXO_S_1, GXO_S_2, GXO_S_3,
  GXO_E, GXO_G, GXO_P, GXO_T, GXO_SB, GXO_DS, GXO_CS, GXO_MN,
  GXO_N_FIELDS
};

/* Tally_F_P % Value, 0-based (Tally_F_D__Form.f90:83-88, Tally_F_P__Form.
 * f90:63-64); gravitational and total energy are not selected in flat
 * geometry (Tally_F_P__Form.f90:105-109) and stay 0. */
enum {
  GXO_TALLY_BARYON_NUMBER = 0,
  GXO_TALLY_MOMENTUM_1, GXO_TALLY_MOMENTUM_2, GXO_TALLY_MOMENTUM_3,
  GXO_TALLY_KINETIC_ENERGY,
  GXO_TALLY_ANGULAR_MOMENTUM_1, GXO_TALLY_ANGULAR_MOMENTUM_2,
  GXO_TALLY_ANGULAR_MOMENTUM_3,
  GXO_TALLY_GRAVITATIONAL_ENERGY, GXO_TALLY_TOTAL_ENERGY,
  GXO_TALLY_FLUID_ENERGY, GXO_TALLY_INTERNAL_ENERGY,
  GXO_TALLY_N
};

typedef struct gxo_ctx gxo_ctx;

/* nCells[3] proper cells, MinCoordinate/MaxCoordinate, bc[2*d+0] inner and
 * bc[2*d+1] outer face of dimension d.  nThreads <= 0 keeps the OpenMP
 * default. */
gxo_ctx *gxo_create(const int n[3], const double lo[3], const double hi[3],
                    const int bc[6], double gamma, double theta,
                    int use_limiter, int riemann, double shock_threshold,
                    double courant, double n_min, int n_threads);
void gxo_destroy(gxo_ctx *c);

/* ReconstructionType: 0 'LINEAR' (default), 1 'PARABOLIC'
 * (EV/Increments/IncrementDivergence_FV__Form.f90:239-241, 637-655). */
void gxo_set_reconstruction(gxo_ctx *c, int parabolic);

/* Padded extents and value count of one field. */
void gxo_extents(const gxo_ctx *c, int m[3], long *n_values);

/* 1-D geometry of dimension d over the padded index range (n+4 values):
 * Center, WidthLeft, WidthRight (CH/ChartBasics/Chart_Template.f90:463-593). */
void gxo_geometry(const gxo_ctx *c, int d, double *center, double *wl,
                  double *wr);

/* SetProblem + PrepareEvolution: primitives on every value (ghosts too), then
 * ExchangeGhostData(Current) and ComputeFromPrimitive.  Features are cleared. */
void gxo_set_primitive(gxo_ctx *c, const double *N, const double *V1,
                       const double *V2, const double *V3, const double *E);

/* Overwrite conserved D, S1..3, G on every value, ComputeFromConserved
 * (no feature detection). */
void gxo_set_conserved(gxo_ctx *c, const double *D, const double *S1,
                       const double *S2, const double *S3, const double *G);

/* Integrator_C_PS % ComputeTimeStep_C_ASC: CourantFactor * min(huge,
 * 1/max(tiny, invdt)).  invdt_out (may be NULL) receives the max before the
 * guard. */
double gxo_time_step_candidate(gxo_ctx *c, double *invdt_out);

/* Step_RK_C_ASC % Compute(Time, TimeStep): one Heun RK2 step. */
void gxo_step(gxo_ctx *c, double dt);

/* Copy Current field f (padded layout) out. */
void gxo_get_field(const gxo_ctx *c, int f, double *out);

/* Current % Sources of the divergence (RecordDivergence) after the last
 * step, conserved field f (D, S_1..3, G, DS order), all nv values. */
void gxo_get_sources(const gxo_ctx *c, int f, double *out);

/* Proper cells that ComputeFromConserved (F/Fluid_P__Kernel.f90:94-192)
 * repaired since the last reset: out[0], out[1] in the stage-2 input Y,
 * out[2], out[3] in StoreSolution (and gxo_set_conserved); out[0], out[2]
 * with D <= N_Min (set to vacuum), out[1], out[3] with E < 0 (E = 0, G
 * rebuilt).  reset != 0 clears the counts. */
void gxo_fallback_counts(gxo_ctx *c, long out[4], int reset);

/* Copy DiffusiveFlux_I_d (d = 0,1,2) out, padded layout. */
void gxo_get_flags(const gxo_ctx *c, int d, double *out);

/* Overwrite DiffusiveFlux_I_d (d = 0,1,2), padded layout. */
void gxo_set_flags(gxo_ctx *c, int d, const double *in);

/* BoundaryFluence_CSL of the last step (EV/Increments/IncrementDivergence_FV__
 * Form.f90:1002-1069): face b = 2*d (inner) or 2*d+1 (outer), field f of D,
 * S_1, S_2, S_3, G; the slab of transverse proper cells, x fastest. */
void gxo_boundary_fluence(const gxo_ctx *c, int b, int f, double *out);

/* Current_ASC % ComputeTallyTemplate (FLD/Current_ASC__Template.f90:239-349)
 * for one chart: interior = VolumeIntegral of the Tally_F_P integrands, the
 * accumulated boundary tally, total = interior + boundary, change += total -
 * old total when compute_change.  Any output may be NULL. */
void gxo_compute_tally(gxo_ctx *c, int compute_change, double *interior,
                       double *boundary, double *total, double *change);

/* ---- single-kernel entry points used to pin the restatement ------------ */

/* Overwrite one fluence slab, and the SurfaceIntegral of the boundary
 * integrands over the fluence held (what a step adds to the boundary tally). */
void gxo_set_boundary_fluence(gxo_ctx *c, int b, int f, const double *in);
void gxo_surface_integral(const gxo_ctx *c, double *out);

/* Gradient_Kernel limited slope for one stencil (OP/Calculus/Derivatives/
 * Gradient_Kernel.f90:166-190). */
double gxo_limited_slope(double dV_L, double dV_R, double dX_L, double dX_R,
                         double theta);

/* Parabolic face values of one cell (EV/Increments/IncrementDivergence_FV__
 * Kernel.f90:296-596): out[0] = value at the inner face xC - dX_L, out[1] at
 * the outer face xC + dX_R, from the values vM, vC, vP at the centres xM, xC,
 * xP. */
void gxo_parabolic_face_values(double vM, double vC, double vP, double xM,
                               double xC, double xP, double dX_L, double dX_R,
                               double out[2]);

/* HLL(+HLLC) flux of one face from reconstructed primitive states
 * (N, V1, V2, V3, E) on the left and right, sweep dimension d (0..2),
 * shock flag df.  f_out[5] = fluxes of D, S1, S2, S3, G. */
void gxo_face_flux(const double prim_l[5], const double prim_r[5], int d,
                   double gamma, int riemann, double df, double f_out[5]);

#ifdef __cplusplus
}
#endif

#endif
