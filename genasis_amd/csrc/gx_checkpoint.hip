// gx_checkpoint.hip -- checkpoint files of one context (gx_write_checkpoint,
// gx_read_checkpoint of include/gx_fluid.h).
//
// The reference writes the Current to Silo at each checkpoint and restarts
// with RestartFrom = k: Integrator_Template % PrepareInitial / ResetInitial
// read checkpoint k's Time, cycle number and fields, set iCheckpoint = k,
// and PrepareEvolution recomputes the conserved fields from the primitives
// (ComputeFromInitial) before evolving on (Modules/Mathematics/Solvers/
// DifferentialEquations/Evolutions/Integrators/Integrator_Template.f90:
// 476-502, 966-996; Integrator_C_PS__Form.f90:175-224).  A checkpoint file
// here holds that (Time, cycle, iCheckpoint, the Current's primitives in
// FP64) and, beside it, the device state itself (conserved and uploaded
// fields in the context's precision, shock flags, CFL maximum, fluence and
// tallies), so that a restart can either repeat the reference's
// (GX_RESTART_PRIMITIVE) or continue bit for bit (GX_RESTART_EXACT).
//
// Layout (little-endian, 600-byte header with the fields in the order of
// Header below and no implicit padding), then the primitives N, V_1,
// V_2, V_3, E of the proper cells (5 x ncell doubles, x fastest), then U
// and P0 (5 x ncell values each, 8 or 4 bytes by precision), the flags
// (ncell bytes) and the fluence (nfluence doubles).  Files are written to
// "<path>.tmp" and renamed.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gx_fluid.h"
#include "gx_host.hpp"
#include "gx_tally.hpp"
#include "gx_context.hpp"

using namespace gx;

namespace {

constexpr char MAGIC[8] = {'G', 'X', 'C', 'K', 'P', 'T', '0', '1'};
constexpr int32_t VERSION = 1;

struct Header {
  char magic[8];
  int32_t version;
  int32_t precision;
  int32_t nCells[3];
  int32_t bc[6];
  int32_t useLimiter;
  int32_t riemann;
  int32_t iCheckpoint;
  int32_t primsValid;
  int32_t invdtValid;
  int32_t reconstruction;
  int32_t pad;
  double lo[3], hi[3];
  double gamma, theta, shockThreshold, courant, nMin, mass;
  double time;
  int64_t cycle;
  uint64_t invdtBits;
  int64_t ncell;
  int64_t nfluence;
  double tallyLocal[TALLY_N], tallyGlobal[TALLY_N], tallyTotal[TALLY_N], tallyChange[TALLY_N];
};
static_assert(sizeof(Header) == 600, "header layout: no implicit padding");

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

bool put(FILE *f, const void *p, size_t n) { return fwrite(p, 1, n, f) == n; }
bool get(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n; }

int io_fail(const std::string &who, const std::string &path, const char *what) {
  return fail(GX_E_ARG, who + ": " + what + " (" + path + ")");
}

// The chart and parameters of the file must be the context's.
int check_match(const gx_ctx *c, const Header &h, const std::string &path) {
  const char *who = "gx_read_checkpoint";
  for (int d = 0; d < 3; d++)
    if (h.nCells[d] != c->grid.nCells[d] || h.lo[d] != c->grid.MinCoordinate[d] ||
        h.hi[d] != c->grid.MaxCoordinate[d])
      return io_fail(who, path, "the file's chart (nCells, coordinates) differs from the context's");
  for (int i = 0; i < 6; i++)
    if (h.bc[i] != c->grid.BoundaryCondition[i])
      return io_fail(who, path, "the file's boundary conditions differ from the context's");
  const gx_params &p = c->params;
  if (h.gamma != p.AdiabaticIndex || h.theta != p.LimiterParameter ||
      h.shockThreshold != p.ShockThreshold || h.courant != p.CourantFactor ||
      h.nMin != p.BaryonDensityMin || h.mass != p.BaryonMassReference ||
      h.useLimiter != p.UseLimiter || h.riemann != p.RiemannSolverType ||
      h.reconstruction != p.ReconstructionType)
    return io_fail(who, path, "the file's fluid parameters differ from the context's");
  if (h.ncell != c->ncell || h.nfluence != fluence_count(c->n))
    return io_fail(who, path, "the file's sizes do not match the chart");
  return GX_OK;
}

}  // namespace

extern "C" {

int gx_write_checkpoint(gx_ctx *c, const char *path, double time, long long cycle,
                        int iCheckpoint) {
  GX_API_BEGIN
  if (c && c->bricks)
    return fail(GX_E_STATE, "gx_write_checkpoint: not available on a chart split into bricks");
  const char *who = "gx_write_checkpoint";
  if (!c || !path) return fail(GX_E_ARG, "gx_write_checkpoint: null argument");
  if (!c->uploaded) return fail(GX_E_STATE, "gx_write_checkpoint: no state uploaded");
  const size_t n5 = 5 * (size_t)c->ncell, nf = (size_t)fluence_count(c->n);
  Header h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, MAGIC, 8);
  h.version = VERSION;
  h.precision = c->params.Precision;
  for (int d = 0; d < 3; d++) {
    h.nCells[d] = c->grid.nCells[d];
    h.lo[d] = c->grid.MinCoordinate[d];
    h.hi[d] = c->grid.MaxCoordinate[d];
  }
  for (int i = 0; i < 6; i++) h.bc[i] = c->grid.BoundaryCondition[i];
  h.useLimiter = c->params.UseLimiter;
  h.riemann = c->params.RiemannSolverType;
  h.reconstruction = c->params.ReconstructionType;
  h.gamma = c->params.AdiabaticIndex;
  h.theta = c->params.LimiterParameter;
  h.shockThreshold = c->params.ShockThreshold;
  h.courant = c->params.CourantFactor;
  h.nMin = c->params.BaryonDensityMin;
  h.mass = c->params.BaryonMassReference;
  h.time = time;
  h.cycle = cycle;
  h.iCheckpoint = iCheckpoint;
  h.primsValid = c->prims_valid ? 1 : 0;
  h.invdtValid = c->invdt_valid ? 1 : 0;
  h.ncell = c->ncell;
  h.nfluence = (int64_t)nf;
  memcpy(h.tallyGlobal, c->tally_global, sizeof(h.tallyGlobal));
  memcpy(h.tallyTotal, c->tally_total, sizeof(h.tallyTotal));
  memcpy(h.tallyChange, c->tally_change, sizeof(h.tallyChange));
  // The Current's primitives (as gx_get_primitive gives them).
  std::vector<double> prim(n5);
  int rc = gx_get_primitive(c, prim.data());
  if (rc) return rc;
  std::vector<char> U(n5 * c->esz), P(n5 * c->esz);
  std::vector<unsigned char> flags((size_t)c->ncell);
  std::vector<double> fl(nf);
  HIP_TRY(hipMemcpyAsync(U.data(), c->U, U.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(P.data(), c->P0, P.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(flags.data(), c->flags, flags.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(fl.data(), c->fluence, sizeof(double) * nf, hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipMemcpyAsync(h.tallyLocal, c->tally_local, sizeof(h.tallyLocal),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&h.invdtBits, c->invdt, sizeof(h.invdtBits), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const std::string tmp = std::string(path) + ".tmp";
  {
    File f;
    f.f = fopen(tmp.c_str(), "wb");
    if (!f.f) return io_fail(who, tmp, "cannot open for writing");
    const bool ok = put(f.f, &h, sizeof(h)) && put(f.f, prim.data(), sizeof(double) * n5) &&
                    put(f.f, U.data(), U.size()) && put(f.f, P.data(), P.size()) &&
                    put(f.f, flags.data(), flags.size()) &&
                    put(f.f, fl.data(), sizeof(double) * nf);
    const bool closed = fclose(f.f) == 0;
    f.f = nullptr;
    if (!ok || !closed) {
      remove(tmp.c_str());
      return io_fail(who, tmp, "write failed");
    }
  }
  if (rename(tmp.c_str(), path) != 0) {
    remove(tmp.c_str());
    return io_fail(who, path, "cannot rename the written file into place");
  }
  return GX_OK;
  GX_API_END("gx_write_checkpoint")
}

int gx_read_checkpoint(gx_ctx *c, const char *path, int mode, double *time,
                       long long *cycle, int *iCheckpoint) {
  GX_API_BEGIN
  if (c && c->bricks)
    return fail(GX_E_STATE, "gx_read_checkpoint: not available on a chart split into bricks");
  const char *who = "gx_read_checkpoint";
  if (!c || !path) return fail(GX_E_ARG, "gx_read_checkpoint: null argument");
  if (mode != GX_RESTART_PRIMITIVE && mode != GX_RESTART_EXACT)
    return fail(GX_E_ARG, "gx_read_checkpoint: mode must be GX_RESTART_PRIMITIVE or GX_RESTART_EXACT");
  File f;
  f.f = fopen(path, "rb");
  if (!f.f) return io_fail(who, path, "cannot open for reading");
  Header h;
  if (!get(f.f, &h, sizeof(h))) return io_fail(who, path, "file too short for a header");
  if (memcmp(h.magic, MAGIC, 8) != 0 || h.version != VERSION)
    return io_fail(who, path, "not a checkpoint file of this format version");
  int rc = check_match(c, h, path);
  if (rc) return rc;
  const size_t n5 = 5 * (size_t)c->ncell, nf = (size_t)h.nfluence;
  std::vector<double> prim(n5);
  if (!get(f.f, prim.data(), sizeof(double) * n5)) return io_fail(who, path, "file truncated");
  if (mode == GX_RESTART_PRIMITIVE) {
    // The reference's restart: primitives read, conserved recomputed,
    // features and tallies start over.
    rc = gx_set_primitive(c, prim.data());
    if (rc) return rc;
  } else {
    if (h.precision != c->params.Precision)
      return io_fail(who, path, "an exact restart needs the precision the file was written in");
    const size_t esz = c->esz;
    std::vector<char> U(n5 * esz), P(n5 * esz);
    std::vector<unsigned char> flags((size_t)c->ncell);
    std::vector<double> fl(nf);
    if (!get(f.f, U.data(), U.size()) || !get(f.f, P.data(), P.size()) ||
        !get(f.f, flags.data(), flags.size()) || !get(f.f, fl.data(), sizeof(double) * nf))
      return io_fail(who, path, "file truncated");
    HIP_TRY(hipMemcpyAsync(c->U, U.data(), U.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->P0, P.data(), P.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->flags, flags.data(), flags.size(), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(c->fluence, fl.data(), sizeof(double) * nf, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(c->tally_local, h.tallyLocal, sizeof(h.tallyLocal),
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->invdt, &h.invdtBits, sizeof(h.invdtBits), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    memcpy(c->tally_global, h.tallyGlobal, sizeof(h.tallyGlobal));
    memcpy(c->tally_total, h.tallyTotal, sizeof(h.tallyTotal));
    memcpy(c->tally_change, h.tallyChange, sizeof(h.tallyChange));
    c->uploaded = true;
    c->prims_valid = h.primsValid != 0;
    c->invdt_valid = h.invdtValid != 0;
    c->cfl_pending = false;
  }
  if (time) *time = h.time;
  if (cycle) *cycle = h.cycle;
  if (iCheckpoint) *iCheckpoint = h.iCheckpoint;
  return GX_OK;
  GX_API_END("gx_read_checkpoint")
}

}  // extern "C"
