// gx_context.hpp -- the single-chart context (gx_ctx of include/gx_fluid.h)
// shared by the step and the adjoint translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/gx_fluid.h"
#include "gx_host.hpp"
#include "gx_tally.hpp"
#include "gx_tiles.hpp"

namespace gx {
struct AdjointState;
// Frees the adjoint state of a context (defined with the adjoint).
void release_adjoint(gx_ctx *c);

// A chart split into bricks (gx_create_bricks, defined in gx_bricks.hip):
// the entry points below stand in for the single-brick code paths.
struct BrickSet;
void release_bricks(gx_ctx *c);
// Primitives (and conserved fields, or NULL: from the primitives) of the
// chart's proper cells into the bricks, with their halos and cleared flags.
int bricks_set_state(gx_ctx *c, const double *primitive, const double *conserved);
// Five proper-cell fields in chart order: which = 0 the conserved state,
// 1 the uploaded primitives.
int bricks_get(gx_ctx *c, int which, double *out);
// The 23 Current fields of the proper cells in chart order (host).
int bricks_current(gx_ctx *c, double *host);
int bricks_get_flags(gx_ctx *c, unsigned char *flags);
int bricks_set_flags(gx_ctx *c, const unsigned char *flags);
// CFL maximum of the current state into c->invdt (stream-ordered).
int bricks_cfl(gx_ctx *c);
int bricks_step(gx_ctx *c, double dt);
int bricks_totals(gx_ctx *c, double out[5]);
// Interior tally partials of every brick into c->tpartial (stream-ordered).
int bricks_interior_partials(gx_ctx *c);
}  // namespace gx

using gx::ChartGeometry;
using gx::StageArgs;

struct gx_ctx {
  gx_grid grid;
  gx_params params;
  int n[3];
  long long ncell;
  hipStream_t stream = nullptr;
  int tz = 16;                // z-chunk of the stage kernels (stage_tz at creation)
  int tile = 0;               // stage-tile shape (gx_set_tile): 0 = 16 x 16, 1 = 32 x 8
  bool fp32 = false;          // state and sweeps in float (Precision = GX_FP32)
  size_t esz = sizeof(double);  // bytes per state value
  // device buffers, of the state's real type
  void *geo_buf = nullptr;    // 3 dimensions x N_GEO arrays x gs
  int gs = 0;                 // geometry stride: max(n) + 4
  // five-field states, field f at base + f * ncell
  void *U = nullptr, *Y = nullptr, *W = nullptr, *P0 = nullptr;
  uint8_t *flags = nullptr;
  unsigned long long *invdt = nullptr;
  void *sources = nullptr;        // Current % Sources of the divergence, 5 fields
  bool record_sources = false;    // gx_set_record_sources
  double *partial = nullptr;
  int n_partial = 0;
  double *scratch = nullptr;  // download staging: GX_N_FIELDS x ncell
  bool uploaded = false;
  bool prims_valid = false;   // next stage 1 reads P0
  bool invdt_valid = false;   // invdt holds the max for the current state
  // boundary fluence of the last step and the conservation tallies
  double *fluence = nullptr;       // fluence_count(n) values
  double *tgeo = nullptr;          // FP64 tally geometry: N_TG x 3 dimensions x gs
  double *tally_local = nullptr;   // device TallyBoundaryLocal (TALLY_N)
  double *bpartial = nullptr;      // boundary partials (n_bpartial x NB_SUM)
  int n_bpartial = 0;              // blocks of the boundary kernels: one line each
  double *tpartial = nullptr;      // interior partials (n_partial x NI_SUM)
  double tally_global[gx::TALLY_N] = {};  // TallyBoundaryGlobal
  double tally_total[gx::TALLY_N] = {};   // TallyTotal
  double tally_change[gx::TALLY_N] = {};  // TallyChange
  // per-kernel timing (gx_set_profiling): stage 1, stage 2, shock flags,
  // boundary fluence and tally.  Alternate steps record into the two sets of
  // events, and a set is read when it comes round again (or by
  // gx_kernel_times), so a profiled step does not wait for its own kernels.
  bool profiling = false;
  hipEvent_t ev[2][6] = {};
  bool ev_used[2] = {false, false};
  int prof_step = 0;
  double kernel_ms[4] = {0, 0, 0, 0};
  long long kernel_n[4] = {0, 0, 0, 0};
  // The CFL maximum a step's stage 2 left, copied to pinned host memory right
  // after stage 2 (ev_cfl): gx_compute_time_step waits for that copy alone,
  // not for the boundary and shock-flag kernels queued behind it, so the
  // next step's launches overlap them.
  unsigned long long *invdt_host = nullptr;
  hipEvent_t ev_cfl = nullptr;
  bool cfl_pending = false;  // invdt_host holds the current invdt (after ev_cfl)
  ChartGeometry geom;
  gx::AdjointState *adj = nullptr;  // discrete adjoint (gx_adjoint_*), on demand
  gx::BrickSet *bricks = nullptr;   // a chart split into bricks (gx_create_bricks)
};

namespace gx {
// The stage-tile shape of the context's stage kernels: the chosen one for
// LINEAR steps without source recording, 16 x 16 otherwise.
inline int stage_shape(const gx_ctx *c) {
  return c->params.ReconstructionType == GX_LINEAR && !c->record_sources ? c->tile : 0;
}

// Where one RK2 step reads and writes when a caller records it (the discrete
// adjoint's forward run): the step input Uin (which it leaves intact), the
// stage values Y and W, the new state Uout and the state before the
// conserved fix Z.  Without a record the step works in the context's U, Y, W.
struct StepRecord {
  const double *Uin;
  double *Y, *W, *Uout, *Z;
};
// gx_step_rk2 with an optional record (FP64 contexts only when rec is set).
int step_rk2(gx_ctx *c, double dt, const StepRecord *rec);
// Kernel times of the profiled step recorded in event set k, into the totals.
int collect_times(gx_ctx *c, int k);
}  // namespace gx

template <typename R>
inline StageArgs<R> base_args(const gx_ctx *c) {
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

