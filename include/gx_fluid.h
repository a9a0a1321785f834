/*
 * gx_fluid.h -- C-ABI of the MI355X finite-volume step for the GenASiS
 * polytropic ideal-gas fluid (Fluid_P_I): PLM/HLLC RK2 with CFL time step,
 * periodic / outflow / reflecting ghost fill and shock flags.
 *
 * Plain C types only.  Every call returns 0 on success and a negative GX_E_*
 * code on failure; gx_last_error() then holds the message (per thread).  One
 * context owns all device memory of one problem and one HIP stream; host
 * buffers passed in are borrowed for the duration of the call.  A context is
 * not re-entrant: call it from one host thread at a time.
 *
 * Reference interfaces each entry point replaces (GenASiS tree, one rank):
 *   gx_create          FluidBox_Form % Initialize with Fluid_P_I, Chart_SLD and
 *                      Step_RK2_C_ASC (Modules/Physics/Universes/
 *                      FluidBox_Form.f90:37-92, 205-241)
 *   gx_upload          Integrator_C_PS % PrepareEvolution: ExchangeGhostData +
 *                      ComputeFromInitial + UpdateDevice (Modules/Mathematics/
 *                      Solvers/DifferentialEquations/Evolutions/Integrators/
 *                      Integrator_C_PS__Form.f90:175-224)
 *   gx_compute_time_step  Integrator_C_PS % ComputeTimeStep_C_ASC
 *                      (Integrator_C_PS__Form.f90:374-423,
 *                      Integrator_C_PS__Kernel.f90:12-144)
 *   gx_step_rk2        Step_RK_C_ASC_Template % Compute (…/Evolutions/Steps/
 *                      Step_RK_C_ASC__Template.f90:504-542) with the Heun
 *                      tableau of Step_RK2_C_ASC__Form.f90:51-57
 *   gx_download        Integrator_C_PS % UpdateHost (Integrator_C_PS__Form.
 *                      f90:280-299): the Current in the reference layout
 *   gx_totals          VolumeIntegral of the conserved fields (Modules/
 *                      Mathematics/Operations/Calculus/Integrals/
 *                      VolumeIntegral_Form.f90:63), one rank
 *   gx_get_boundary_fluence  Step_RK_C_ASC % BoundaryFluence_CSL after a step
 *                      (…/Evolutions/Steps/Step_RK_C_ASC__Template.f90:
 *                      526-537, 1370-1422; …/Increments/IncrementDivergence_
 *                      FV__Form.f90:1002-1069)
 *   gx_set_record_sources / gx_get_sources  ApplyDivergence_C's
 *                      RecordDivergenceKernel into Current % Sources
 *                      (…/Evolutions/Steps/Step_RK_C_ASC__Template.f90:
 *                      1572-1589, Step_RK_C_ASC__Kernel.f90:12-46)
 *   gx_write_checkpoint / gx_read_checkpoint  Integrator_Template % Write at a
 *                      checkpoint and the RestartFrom path of PrepareInitial
 *                      (…/Evolutions/Integrators/Integrator_Template.f90:
 *                      476-502, 966-996)
 *   gx_compute_tally   Current_ASC % ComputeTally with the Tally_F_P tallies
 *                      (Modules/Mathematics/Solvers/DifferentialEquations/
 *                      Fields/Current_ASC__Template.f90:222-349, Tally_C__
 *                      Form.f90:170-279; Modules/Physics/StressEnergies/
 *                      Fluids/Tally_F_P__Form.f90, Tally_F_D__Form.f90), one
 *                      chart, one rank
 */
#ifndef GX_FLUID_H
#define GX_FLUID_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GX_OK 0
#define GX_E_ARG -1     /* invalid argument */
#define GX_E_HIP -2     /* HIP runtime error */
#define GX_E_NOMEM -3   /* device or host allocation failed */
#define GX_E_STATE -4   /* call out of order (e.g. step before upload) */

/* boundary conditions (Atlas_SC__Template.f90:244-324) */
#define GX_BC_PERIODIC 0
#define GX_BC_OUTFLOW 1
#define GX_BC_REFLECTING 2

/* RiemannSolverType (Fluid_ASC__Form.f90:128-136) */
#define GX_HLL 0
#define GX_HLLC 1

/* ReconstructionType (IncrementDivergence_FV__Form.f90:239-241, 637-655) */
#define GX_LINEAR 0
#define GX_PARABOLIC 1

/* Precision of the state and sweep arithmetic */
#define GX_FP64 0
#define GX_FP32 1

/* Fluid_P_I fields, in the order of the reference's field map
 * (FLD/Current_Template.f90:752-753, F/Fluid_D__Form.f90:650-654,
 * F/Fluid_P__Template.f90:718-725). */
enum {
  GX_FAST_EIGENSPEED_PLUS_1 = 0,
  GX_FAST_EIGENSPEED_PLUS_2,
  GX_FAST_EIGENSPEED_PLUS_3,
  GX_FAST_EIGENSPEED_MINUS_1,
  GX_FAST_EIGENSPEED_MINUS_2,
  GX_FAST_EIGENSPEED_MINUS_3,
  GX_COMOVING_BARYON_DENSITY,
  GX_CONSERVED_BARYON_DENSITY,
  GX_BARYON_MASS,
  GX_VELOCITY_U_1,
  GX_VELOCITY_U_2,
  GX_VELOCITY_U_3,
  GX_MOMENTUM_DENSITY_D_1,
  GX_MOMENTUM_DENSITY_D_2,
  GX_MOMENTUM_DENSITY_D_3,
  GX_INTERNAL_ENERGY,
  GX_CONSERVED_ENERGY,
  GX_PRESSURE,
  GX_TEMPERATURE,
  GX_ENTROPY_PER_BARYON,
  GX_CONSERVED_ENTROPY,
  GX_SOUND_SPEED,
  GX_MACH_NUMBER,
  GX_N_FIELDS
};

/* Chart: proper cells, coordinate box and one boundary condition per face
 * (BoundaryCondition[2*d] inner, [2*d+1] outer).  nGhostLayers is 2. */
typedef struct {
  int nCells[3];
  double MinCoordinate[3];
  double MaxCoordinate[3];
  int BoundaryCondition[6];
} gx_grid;

/* Parameters, named as the reference's GetParameter keys. */
typedef struct {
  double AdiabaticIndex;      /* 1.4  (Fluid_P_I__Form.f90:210) */
  double LimiterParameter;    /* 1.4  (Fluid_ASC__Form.f90:173-177) */
  double ShockThreshold;      /* 0.1  (FluidFeatures_ASC__Form.f90:61) */
  double CourantFactor;       /* 0.7  (Integrator_C_PS__Form.f90:136) */
  double BaryonDensityMin;    /* 0.0  (Fluid_D__Form.f90:35) */
  double BaryonMassReference; /* 1.0 */
  int UseLimiter;             /* 1 */
  int RiemannSolverType;      /* GX_HLLC */
  int Precision;              /* GX_FP64 */
  int ReconstructionType;     /* GX_LINEAR (IncrementDivergence_FV__Form.f90:239) */
} gx_params;

typedef struct gx_ctx gx_ctx;

/* Fills p with the reference defaults above. */
void gx_default_params(gx_params *p);

int gx_create(const gx_grid *grid, const gx_params *params, gx_ctx **ctx);
int gx_destroy(gx_ctx *ctx);

/* A chart split into nBricks[0] x nBricks[1] x nBricks[2] bricks in one
 * GPU's memory: Chart_SLD with the nBricks parameter (Modules/Mathematics/
 * Manifolds/Atlases/Charts/ChartBasics/Chart_Template.f90:383-455; ghost
 * exchange Charts/SingleLevelCharts/Chart_SLD__Form.f90:131-201, 324-554,
 * 557-1007).  nBricks must divide nCells in every dimension (SetBrick), with
 * at least 2 cells per brick and dimension.  Each brick keeps its own state
 * with a two-layer halo for every face it shares with another brick; every
 * RK stage exchanges the halos, the CFL maximum is taken over all bricks,
 * and shock flags are cleared only at the faces of the chart (FluidFeatures_
 * P__Form.f90:333-338), so the chart steps bit for bit as the one-brick
 * chart of gx_create does.  The context takes every call of a gx_create
 * context on the whole chart's proper cells, except the source recording,
 * profiling, checkpoints and the adjoint (GX_E_STATE); ReconstructionType
 * must be GX_LINEAR.  The one-rank, many-GPU layout of the same bricks is
 * described in the design notes; here all bricks share one GPU. */
int gx_create_bricks(const gx_grid *grid, const gx_params *params, const int nBricks[3],
                     gx_ctx **ctx);
/* nBricks of a context (1, 1, 1 for gx_create). */
int gx_get_bricks(gx_ctx *ctx, int nBricks[3]);

/* Upload the Current from the reference layout Value(nValues, nVariables),
 * column-major: field f of value iv at value[(fieldMap[f]-1)*ldValue + iv],
 * nValues = prod(nCells+4) with x fastest.  fieldMap has GX_N_FIELDS entries,
 * 1-based column numbers; an entry <= 0 marks an absent field.  The five
 * primitive fields (N, V_1..3, E) must be present; the conserved fields and
 * the other Current fields follow from them on the device, as
 * PrepareEvolution's ComputeFromInitial does (the columns the host holds for
 * them are not read).  The first step after an upload reconstructs from the
 * uploaded primitives, as the reference's first stage does.  Shock flags,
 * boundary fluence and tallies are cleared. */
int gx_upload(gx_ctx *ctx, const double *value, long long ldValue,
              const int *fieldMap);

/* Write the Current (all present fields of fieldMap) for proper cells, and
 * for ghost cells of periodic dimensions (their periodic image).  Ghost cells
 * across non-periodic faces are not written. */
int gx_download(gx_ctx *ctx, double *value, long long ldValue,
                const int *fieldMap);

/* DiffusiveFlux_I_1..3 of FluidFeatures_P (F/FluidFeatures_Template.f90:156)
 * in the same padded layout, three consecutive columns of ldValue values. */
int gx_upload_features(gx_ctx *ctx, const double *df, long long ldValue);
int gx_download_features(gx_ctx *ctx, double *df, long long ldValue);

/* ComputeTimeStep_C_ASC: *candidate = CourantFactor * min(huge,
 * 1/max(tiny, invdt)), invdt = max over proper cells of
 * sum_d max(FEP_d, -FEM_d)/(WidthLeft_d + WidthRight_d).  Synchronises. */
int gx_compute_time_step(gx_ctx *ctx, double *candidate);

/* gx_compute_time_step without a context and without allocating.  It has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).
 * conserved: device pointer to the five proper-cell fields D, S_1..3, G
 * (x fastest), double for Precision GX_FP64, float for GX_FP32; candidate:
 * device pointer to one double that receives CourantFactor * min(huge,
 * 1/max(tiny, invdt)) as gx_compute_time_step computes it; stream: a
 * hipStream_t (NULL: the null stream).  Stream-ordered and asynchronous:
 * nothing is allocated and nothing synchronises, so the call can be captured
 * in a HIP graph.  Its scratch (the CFL maximum and the chart's cell widths,
 * 1.18 MB) is a static device array per GPU, held until the process ends:
 * at most 4096 cells per dimension and 2^29 - 1 cells in all (GX_E_ARG
 * beyond).  Every call on a GPU shares that array: calls that overlap on
 * one GPU (other streams, other host threads) race on it. */
int gx_compute_time_step_noalloc(const gx_grid *grid, const gx_params *params,
                                 const void *conserved, double *candidate, void *stream);

/* One Heun RK2 step of size dt (stream-ordered, asynchronous).  time is the
 * reference's Time argument and does not enter the arithmetic. */
int gx_step_rk2(gx_ctx *ctx, double time, double dt);

/* Block until all work of the context is done. */
int gx_sync(gx_ctx *ctx);

/* Sum over proper cells of D, S_1, S_2, S_3, G times the cell volume, in
 * FP64.  Synchronises. */
int gx_totals(gx_ctx *ctx, double out[5]);

/* Conservation tallies (Tally_F_P % Value, Tally_F_D__Form.f90:83-88,
 * Tally_F_P__Form.f90:63-64), 0-based.  Flat geometry selects all but the
 * gravitational and total energies, which stay 0. */
enum {
  GX_TALLY_BARYON_NUMBER = 0,
  GX_TALLY_MOMENTUM_1,
  GX_TALLY_MOMENTUM_2,
  GX_TALLY_MOMENTUM_3,
  GX_TALLY_KINETIC_ENERGY,
  GX_TALLY_ANGULAR_MOMENTUM_1,
  GX_TALLY_ANGULAR_MOMENTUM_2,
  GX_TALLY_ANGULAR_MOMENTUM_3,
  GX_TALLY_GRAVITATIONAL_ENERGY,
  GX_TALLY_TOTAL_ENERGY,
  GX_TALLY_FLUID_ENERGY,
  GX_TALLY_INTERNAL_ENERGY,
  GX_TALLY_N
};

/* Boundary fluence of the last gx_step_rk2: sum over its two stages of
 * (1/2) dt times the flux through each boundary face of the chart, for D,
 * S_1, S_2, S_3, G (the reference's conserved entropy DS is recomputed from
 * the primitives and has no tally; its fluence is not kept).  Layout: face
 * b = 2*d (inner) or 2*d+1 (outer) of dimension d, then field, then the slab
 * of transverse proper cells with x fastest (d = 0: y + n1*z; d = 1:
 * x + n0*z; d = 2: x + n0*y); faces in the order b = 0..5, so face b starts
 * at 10 * sum_{e < d} slab_e + 5 * (b & 1) * slab_d.  fluence holds
 * 10 * (n1*n2 + n0*n2 + n0*n1) doubles.  Zero before the first step.
 * Synchronises. */
int gx_get_boundary_fluence(gx_ctx *ctx, double *fluence);

/* ComputeTallyTemplate: interior = VolumeIntegral of the tally integrands of
 * the Current (D, S, 0.5 S.V, X x S, G, E); the boundary tally accumulated
 * by every step since the upload (the SurfaceIntegral of the boundary
 * integrands over each step's fluence, outward positive) is added to the
 * global boundary tally; total = interior + boundary; change += total -
 * previous total when computeChange.  An upload starts all tallies at zero,
 * so the first call (computeChange = 0) sets total = interior.  Sums are in
 * FP64 with a fixed order.  Any output may be NULL (GX_TALLY_N values each).
 * Steps taken by gx_adjoint_forward record the boundary fluence and add to
 * the boundary tally as gx_step_rk2's do.  Synchronises. */
int gx_compute_tally(gx_ctx *ctx, int computeChange, double *interior,
                     double *boundary, double *total, double *change);

/* Checkpoint files.  gx_write_checkpoint stores the host integrator's Time,
 * cycle and iCheckpoint with the Current's primitives (FP64, proper cells)
 * and the device state (conserved and uploaded fields in the context's
 * precision, shock flags, CFL maximum, boundary fluence, tallies); written to
 * "<path>.tmp", then renamed.  gx_read_checkpoint checks that the file's
 * chart and parameters are the context's and restores:
 *   GX_RESTART_PRIMITIVE  the reference's RestartFrom (Integrator_Template.
 *                         f90:476-502, 966-996; PrepareEvolution): the
 *                         primitives, with the conserved fields recomputed
 *                         from them (as gx_set_primitive), flags and tallies
 *                         cleared;
 *   GX_RESTART_EXACT      the whole device state: the run continues bit for
 *                         bit as if it had not stopped (same precision).
 * time, cycle, iCheckpoint receive the stored values (each may be NULL). */
#define GX_RESTART_PRIMITIVE 0
#define GX_RESTART_EXACT 1
int gx_write_checkpoint(gx_ctx *ctx, const char *path, double time, long long cycle,
                        int iCheckpoint);
int gx_read_checkpoint(gx_ctx *ctx, const char *path, int mode, double *time,
                       long long *cycle, int *iCheckpoint);

/* Proper-cell views without the padded layout: five contiguous fields of
 * nCells[0]*nCells[1]*nCells[2] doubles each, x fastest. */
int gx_set_state(gx_ctx *ctx, const double *primitive, const double *conserved);
/* Primitives only; the conserved fields follow on the device as
 * ComputeFromInitial does (F/Fluid_P_I__Form.f90:509-601). */
int gx_set_primitive(gx_ctx *ctx, const double *primitive);
int gx_get_conserved(gx_ctx *ctx, double *conserved);
int gx_get_primitive(gx_ctx *ctx, double *primitive);
/* Packed shock flags: one byte per proper cell, bit d = DiffusiveFlux_I_(d+1). */
int gx_get_flags(gx_ctx *ctx, unsigned char *flags);
int gx_set_flags(gx_ctx *ctx, const unsigned char *flags);

/* Launch-shape control for measurement: z-chunk length of the fused stage
 * kernel (planes per thread block), >= 1; a value above nCells[2] is taken
 * as nCells[2].  The default, set by gx_create, is 32 planes, halved (down
 * to 4) while a launch would hold fewer than 512 blocks of 16 x 16 cells;
 * gx_get_tuning reads the current value. */
int gx_set_tuning(gx_ctx *ctx, int zChunk);
int gx_get_tuning(gx_ctx *ctx, int *zChunk);

/* Stage-tile shape of the fused stage kernels: tileX x tileY cells in x and y
 * per thread block, 16 x 16 (the default) or 32 x 8.  The results do not
 * depend on it (every cell takes the same operations); it changes the halo
 * share and the partial tiles of a chart.  PARABOLIC steps and steps that
 * record the divergence sources use 16 x 16. */
int gx_set_tile(gx_ctx *ctx, int tileX, int tileY);
int gx_get_tile(gx_ctx *ctx, int *tileX, int *tileY);

/* No counterpart in the reference: *pow2 = 1 when every centre spacing, slope
 * denominator and cell width of the chart is a power of two (a uniform grid
 * of 2^k cells per unit length, as the unit cube with a power-of-two cell
 * count).  Every division by the geometry is then exact as a multiplication
 * by the reciprocal, and the FP64 LINEAR stage kernels (single charts,
 * bricks, the adjoint's line passes) take that form: the same bits, fewer
 * instructions.  Steps that record the divergence sources, PARABOLIC steps
 * and batches keep the corrected quotient.  Otherwise 0. */
int gx_get_geometry_pow2(gx_ctx *ctx, int *pow2);

/* Current % Sources of the divergence (ApplyDivergence_C's recording,
 * Step_RK_C_ASC__Template.f90:1548-1595 with RecordDivergenceKernel,
 * Step_RK_C_ASC__Kernel.f90:12-46): with recording on, every gx_step_rk2
 * sets Sources = sum over its stages of Weight_RK * Increment / TimeStep for
 * D, S_1..3 and G (the conserved entropy is not evolved).  Off by default
 * (a diagnostic; it adds a read and a write of five fields per stage);
 * turning it on clears Sources.  gx_get_sources copies the five proper-cell
 * fields (FP64 values of the context's precision).  Both synchronise. */
int gx_set_record_sources(gx_ctx *ctx, int on);
int gx_get_sources(gx_ctx *ctx, double *sources);

/* The HIP stream of the context (a hipStream_t), for event timing. */
void *gx_stream(gx_ctx *ctx);

/* Per-kernel timing with HIP events on the context's stream: when on, each
 * gx_step_rk2 waits for its own kernels and adds their durations to
 * ms[0..3] (stage 1, stage 2, shock flags, and the boundary-fluence and
 * boundary-tally kernels of both stages); count[k] steps.
 * Switching it on (or off) clears the sums. */
int gx_set_profiling(gx_ctx *ctx, int on);
int gx_kernel_times(gx_ctx *ctx, double ms[4], long long count[4]);

/* Device self-test of the exactly rounded reciprocal division used for the
 * divisions by geometry constants and cell volumes: n pseudo-random operand
 * pairs (zeros, tiny, subnormal and NaN numerators among them),
 * *mismatches = number of quotients whose bits differ from the IEEE division
 * where the division is exact (a zero numerator, or |a| >= 2^-969 with a
 * normal quotient) or that differ by more than one ulp elsewhere.
 * Synchronises. */
int gx_self_test_division(long long n, unsigned long long seed,
                          long long *mismatches);

const char *gx_last_error(void);

/* ---------------------------------------------------------------------------
 * Batches: many independent charts (of any sizes) advanced together, each with
 * its own integrator state.  One batch call does for every problem what a
 * loop over separate FluidBox programs does for each (FluidBox_Form %
 * Initialize, Modules/Physics/Universes/FluidBox_Form.f90:37-92; Integrator_
 * C_PS % ComputeCycle_ASC, Integrator_C_PS__Form.f90:328-371), and gives
 * each problem the bits a gx_ctx of the same chart gives it.  Every RK stage
 * of all problems is one persistent launch; gx_batch_step captures its
 * cycles in one HIP graph.  Parameters are shared by the problems.
 * ------------------------------------------------------------------------- */
typedef struct gx_batch gx_batch;

int gx_batch_create(int nProblems, const gx_grid *grids, const gx_params *params,
                    gx_batch **batch);
int gx_batch_destroy(gx_batch *batch);

/* Primitives (N, V_1..3, E: five contiguous proper-cell fields) of problem
 * iProblem; the conserved fields follow on the device as ComputeFromInitial
 * does.  Resets the problem's Time and cycle count to 0. */
int gx_batch_set_primitive(gx_batch *batch, int iProblem, const double *primitive);

/* Integrator parameters of every problem (IntegratorHeader_Form.f90:95-131):
 * FinishTime (default 1), FinishCycle (default huge), nRampCycles (100). */
int gx_batch_set_integrator(gx_batch *batch, double FinishTime, int FinishCycle,
                            int nRampCycles);

/* nSteps cycles of Integrator_C_PS % ComputeCycle_ASC for every problem:
 * dt from the problem's own CFL candidate (ComputeTimeStep_C_ASC), times the
 * ramp min((iCycle+1)/nRampCycles, 1), clipped onto FinishTime; a problem at
 * FinishTime or FinishCycle is not stepped.  Stream-ordered, asynchronous. */
int gx_batch_step(gx_batch *batch, int nSteps);
int gx_batch_sync(gx_batch *batch);

/* ComputeTimeStep_C_ASC of every problem: candidates[nProblems].
 * Synchronises. */
int gx_batch_cfl(gx_batch *batch, double *candidates);

/* Time, last time step and cycle count of every problem (lastTimeStep and
 * cycle may be NULL).  Synchronises. */
int gx_batch_times(gx_batch *batch, double *time, double *lastTimeStep, int *cycle);

/* Proper-cell views of problem iProblem, as gx_get_conserved / gx_get_flags. */
int gx_batch_get_conserved(gx_batch *batch, int iProblem, double *conserved);
int gx_batch_get_flags(gx_batch *batch, int iProblem, unsigned char *flags);

/* Launch shape: the z-chunk of each size class (a size class is
 * floor(2*log2(largest edge)), half-octaves), and whether gx_batch_step
 * captures a HIP graph (useGraph) or launches its kernels one by one. */
int gx_batch_set_tuning(gx_batch *batch, int nClasses, const int *sizeClass,
                        const int *zChunk, int useGraph);
/* Up to capacity (class, z-chunk) pairs; *nClasses = number of classes. */
int gx_batch_get_tuning(gx_batch *batch, int capacity, int *sizeClass, int *zChunk,
                        int *nClasses);
/* Stage-tile shape (tileX x tileY: 16 x 16 or 32 x 8, as gx_set_tile) and
 * z-chunk of each size class; useGraph as gx_batch_set_tuning.  The stage
 * launches of a cycle run once per shape in use, each over the tiles of its
 * size classes. */
int gx_batch_set_tiles(gx_batch *batch, int nClasses, const int *sizeClass,
                       const int *tileX, const int *tileY, const int *zChunk, int useGraph);
int gx_batch_get_tiles(gx_batch *batch, int capacity, int *sizeClass, int *tileX,
                       int *tileY, int *zChunk, int *nClasses);
/* Chooses the stage-tile shape and the z-chunk of every size class by timing
 * the stage-1 launch over the problems of that class for both shapes and
 * z-chunks 4 .. 64 (best of nRepeat).  Needs every problem's state; leaves
 * the states unchanged. */
int gx_batch_autotune(gx_batch *batch, int nRepeat);

/* Total proper cells of the batch, and its HIP stream (a hipStream_t). */
long long gx_batch_cells(gx_batch *batch);
void *gx_batch_stream(gx_batch *batch);

/* ---------------------------------------------------------------------------
 * Discrete adjoint (FP64 contexts).  The reference has no adjoint; these
 * differentiate the step that gx_step_rk2 takes (Step_RK_C_ASC_Template %
 * Compute, Step_RK_C_ASC__Template.f90:504-542) driven by the integrator's
 * time steps (Integrator_Template.f90:896-939, nRampCycles ramp), time steps
 * included (they follow the CFL maximum of the state).  Shock flags and the
 * HLLC/HLL choice are replayed as constants; at ties of the limiter the
 * centred slope is taken, and where exactly one of its two differences is
 * zero the derivative is the mean of the two sides.
 * ------------------------------------------------------------------------- */

/* nSteps cycles from freshly uploaded primitives (gx_set_primitive /
 * gx_upload), recording the time steps and the state every checkpointEvery
 * steps.  Leaves the final state in the context. */
int gx_adjoint_forward(gx_ctx *ctx, int nSteps, int checkpointEvery, int nRampCycles);

/* Reverse pass of the recorded run: given Ubar = dJ/dU_final (five proper-cell
 * fields D, S_1..3, G), returns P0bar = dJ/d(N, V_1..3, E) of the uploaded
 * primitives and *gammaBar = dJ/dAdiabaticIndex.  Recomputes each segment
 * from its checkpoint, replaying the shock flags the forward run recorded.
 * Ubar = NULL takes the seed that gx_adjoint_kinetic_energy left on the
 * device (GX_E_STATE if there is none); P0bar = NULL leaves the result on
 * the device for gx_adjoint_get_bar.  Synchronises. */
int gx_adjoint_backward(gx_ctx *ctx, const double *Ubar, double *P0bar, double *gammaBar);

/* Copies the result of the last completed gx_adjoint_backward (P0bar, five
 * proper-cell fields) to the host.  Synchronises. */
int gx_adjoint_get_bar(gx_ctx *ctx, double *P0bar);

/* The observable J = sum over cells with lo < y_c < hi of 1/2 N V_comp^2 dV
 * (comp 0..2) on the current state, and Ubar = dJ/dU (copied to the host
 * unless Ubar is NULL; it stays on the device as the seed of the next
 * gx_adjoint_backward either way).  Synchronises. */
int gx_adjoint_kinetic_energy(gx_ctx *ctx, int component, double lo, double hi,
                              double *loss, double *Ubar);

/* Host evaluations of the device code's reverse sweeps, for checking: one
 * face flux (returns F and adds nothing: the bars are overwritten) and one
 * limited slope (out = bars of vm, v0, vp for sbar). */
void gx_adjoint_face_vjp(const double WL[5], const double WR[5], int d, int df,
                         double gamma, double mass, int hllc, const double Fbar[5],
                         double WLbar[5], double WRbar[5], double *gammaBar, double F[5]);
void gx_adjoint_slope_vjp(double vm, double v0, double vp, double dxl, double dxr,
                          double theta, int useLimiter, double sbar, double out[3]);

#ifdef __cplusplus
}
#endif

#endif
