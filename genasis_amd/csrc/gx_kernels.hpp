// gx_kernels.hpp -- device arithmetic of the Fluid_P_I finite-volume step.
//
// Every function keeps the reference's expression order (Fortran evaluates
// a*b*c as (a*b)*c and a+b-c as (a+b)-c); the library is compiled with
// -ffp-contract=off so no multiply-add is fused, and FP64 division and square
// root are the IEEE correctly rounded operations.  Metric factors of the
// GALILEAN rectangular chart are exactly 1 and are omitted.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace gx {

constexpr int BC_PERIODIC = 0;
constexpr int BC_OUTFLOW = 1;
constexpr int BC_REFLECTING = 2;
// A brick face shared with another brick of the chart (or with the chart's
// other end across a periodic face split into bricks): its ghost values are
// the halo the brick exchange fills (Chart_SLD % ExchangeGhostData).
constexpr int BC_NEIGHBOR = 3;

// Per-index 1-D geometry of one dimension, padded index p = c + 2 for the
// interior-based cell index c in [-2, n+2).
enum GeoArray {
  G_WL = 0,  // WidthLeft (Chart_Template.f90:985-987)
  G_WR,      // WidthRight
  G_DXL,     // Center(c) - Center(c-1) (Difference_Kernel)
  G_DXR,     // Center(c+1) - Center(c)
  G_DXL2,    // dxl*dxl
  G_DXR2,    // dxr*dxr
  G_DEN,     // (dxl*dxr)*(dxl+dxr)
  G_WIDTH,   // WidthLeft + WidthRight (GeometryFlat_Kernel.f90:21-23)
  G_RDXL,    // RN(1/dxl)
  G_RDXR,    // RN(1/dxr)
  G_RDEN,    // RN(1/den)
  G_CENTER,  // Center (Chart_Template.f90:827-841), for PARABOLIC
  N_GEO
};

// a/b rounded to nearest for b > 0, given y = RN(1/b): q0 = RN(a y) is
// corrected twice with the exact-remainder fma; after the first correction q
// is within one ulp of a/b, so the remainder of the second is exact and its
// correction rounds to RN(a/b) (Markstein, IBM J. Res. Dev. 34 (1990)
// 111-119).  Callers keep b in [2^-500, 2^122].  The result is RN(a/b) bit
// for bit when a = +-0 (the sign is put back: the corrections of -0 give +0)
// and when |a| >= 2^-969 with a normal quotient (|a/b| >= 2^-1022), where
// every remainder stays exact; below that (operands near 1e-292, quotients
// in the subnormal range) it is within one ulp.  There is no branch to an
// IEEE division: the per-call test and branch cost 13 % of the step, and no
// state of the tested workloads reaches those magnitudes.  For a = +-inf or
// an overflowing quotient the result is NaN where the IEEE division gives
// +-inf: a state that has already overflowed.
__device__ __forceinline__ double div_rcp(double a, double b, double y) {
  double q = a * y;
  double r = fma(-b, q, a);
  q = fma(r, y, q);
  r = fma(-b, q, a);
  return copysign(fma(r, y, q), a);
}

struct Physics {
  double gamma, theta, mass, nmin, st;
  int use_limiter, hllc;
};

// Arithmetic of one precision.  FP64 keeps the reference's IEEE operations
// (divisions by precomputed reciprocals are corrected to the exact quotient).
// FP32 multiplies by the hardware reciprocal (v_rcp_f32, 1 ulp) and uses the
// hardware square root (v_sqrt_f32, 1 ulp): a quotient is within 2 ulp.
template <typename R> struct Arith;
template <> struct Arith<double> {
  static constexpr double tiny = DBL_MIN;
  static constexpr double sqrt_tiny = 1.4916681462400413e-154;  // sqrt(DBL_MIN)
  __device__ static __forceinline__ double rdiv(double a, double b, double y) {
    return div_rcp(a, b, y);
  }
  __device__ static __forceinline__ double recip(double b) { return 1.0 / b; }
  __device__ static __forceinline__ double div(double a, double b) { return a / b; }
  __device__ static __forceinline__ double root(double a) { return sqrt(a); }
};
template <> struct Arith<float> {
  static constexpr float tiny = FLT_MIN;
  static constexpr float sqrt_tiny = 1.0842021724855044e-19f;  // sqrt(FLT_MIN)
  __device__ static __forceinline__ float rdiv(float a, float, float y) { return a * y; }
  __device__ static __forceinline__ float recip(float b) { return __builtin_amdgcn_rcpf(b); }
  __device__ static __forceinline__ float div(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
  }
  __device__ static __forceinline__ float root(float a) { return __builtin_amdgcn_sqrtf(a); }
};

// a / b for a divisor b of the chart geometry (a spacing, a limiter
// denominator or a cell volume) with y = RN(1/b).  P2: every such divisor of
// the chart is a normal power of two, so y = 1/b exactly and a * y is the
// correctly rounded quotient for every a (a / 2^k and a * 2^-k are the same
// real number, rounded once): one multiplication instead of the corrected
// quotient, bit for bit the same value.
template <bool P2, typename R>
__device__ __forceinline__ R gdiv(R a, R b, R y) {
  if constexpr (P2)
    return a * y;
  else
    return Arith<R>::rdiv(a, b, y);
}

// Fluid_P_I % ComputeFromConservedCommon, primitive part (F/Fluid_P__Kernel.
// f90:94-192): N, V, E from D, S, G with the density floor and E >= 0.  The
// floor N_Min is compared in FP64 whatever the state's precision.
template <typename R>
__device__ __forceinline__ void prim_from_cons(const R U[5], R mass, double nmin, R V[5]);

// S/(M D) of a live cell, as div_rcp forms it for M D in [2^-100, 2^100] and
// by IEEE division outside (a branch that every lane of a wave skips unless
// one of them holds such a density).
__device__ __forceinline__ void velocity_of(const double U[5], double MD, bool live,
                                            double v[3]) {
  const double y = 1.0 / MD;
  v[0] = div_rcp(U[1], MD, y);
  v[1] = div_rcp(U[2], MD, y);
  v[2] = div_rcp(U[3], MD, y);
  if (live && !(MD > 0x1p-100 && MD < 0x1p100)) {
    v[0] = U[1] / MD;
    v[1] = U[2] / MD;
    v[2] = U[3] / MD;
  }
}

// Both branches of the floor are formed and the live one selected (a vacuum
// cell's quotients are discarded), so the common path has no branch.
template <>
__device__ __forceinline__ void prim_from_cons<double>(const double U[5], double mass,
                                                       double nmin, double V[5]) {
  const double D = U[0];
  const bool live = D > nmin;
  double v[3];
  velocity_of(U, mass * D, live, v);
  const double E = U[4] - 0.5 * ((U[1] * v[0] + U[2] * v[1]) + U[3] * v[2]);
  V[0] = live ? D : nmin;
  V[1] = live ? v[0] : 0.0;
  V[2] = live ? v[1] : 0.0;
  V[3] = live ? v[2] : 0.0;
  const double e = live ? E : 0.0;
  V[4] = e < 0.0 ? 0.0 : e;
}

// FP32 state: the positivity decisions are FP64 ones.  D > N_Min is compared
// in double; E = G - S.V/2 is formed in float, and where it lies within the
// float evaluation's error bound of zero -- 2^-18 (|G| + S.V/2), some 60
// times the bound of the three roundings of S/(M D) and the five of the sum
// -- its sign is taken in double from 2 G (M D) - |S|^2 (the product of two
// floats is exact in double, so only the last three roundings remain).
// Outside that band the float sign is the exact sign.  A cell whose E the
// double test finds negative is repaired (E = 0); otherwise E is the float
// value, clamped at 0.
template <>
__device__ __forceinline__ void prim_from_cons<float>(const float U[5], float mass,
                                                      double nmin, float V[5]) {
  const float D = U[0];
  if ((double)D > nmin) {
    const float y = __builtin_amdgcn_rcpf(mass * D);
    V[0] = D;
    V[1] = U[1] * y;
    V[2] = U[2] * y;
    V[3] = U[3] * y;
    const float ke = 0.5f * ((U[1] * V[1] + U[2] * V[2]) + U[3] * V[3]);
    float E = U[4] - ke;
    if (!(fabsf(E) > 0x1p-18f * (fabsf(U[4]) + ke))) {
      const double MD = (double)mass * (double)D;
      const double s2 =
          ((double)U[1] * (double)U[1] + (double)U[2] * (double)U[2]) + (double)U[3] * (double)U[3];
      E = 2.0 * (double)U[4] * MD - s2 < 0.0 ? 0.0f : fmaxf(E, 0.0f);
    }
    V[4] = E < 0.0f ? 0.0f : E;
  } else {
    V[0] = (float)nmin;
    V[1] = 0.0f;
    V[2] = 0.0f;
    V[3] = 0.0f;
    V[4] = 0.0f;
  }
}

// Gradient_Kernel limited slope (OP/Calculus/Derivatives/Gradient_Kernel.f90:
// 166-190), or the unlimited centred slope when UseLimiter is false.
// The three divisions use the precomputed reciprocals of the geometry.
template <bool P2 = false, typename R>
__device__ __forceinline__ R slope(R vm, R v0, R vp, R dxl, R dxr, R dxl2, R dxr2,
                                   R den, R rdxl, R rdxr, R rden, R theta,
                                   int use_limiter) {
  const R dVL = v0 - vm;
  const R dVR = vp - v0;
  const R central = gdiv<P2>(dxr2 * dVL + dxl2 * dVR, den, rden);
  if (!use_limiter) return central;
  return (copysign(R(0.5), dVL) + copysign(R(0.5), dVR)) *
         fmin(fmin(fabs(gdiv<P2>(theta * dVL, dxl, rdxl)),
                   fabs(gdiv<P2>(theta * dVR, dxr, rdxr))),
              fabs(central));
}

// The same slope with the left quotient |theta dVL / dxl| given (*q) and the
// right one returned in *q.  The right quotient of a cell is the left one of
// its upper neighbour bit for bit (dV and the centre spacing are the same
// differences of the same values), so a march along one dimension computes
// each once.
template <bool P2 = false, typename R>
__device__ __forceinline__ R slope_carry(R vm, R v0, R vp, R dxr, R dxl2, R dxr2, R den,
                                         R rdxr, R rden, R theta, int use_limiter, R *q) {
  const R dVL = v0 - vm;
  const R dVR = vp - v0;
  const R central = gdiv<P2>(dxr2 * dVL + dxl2 * dVR, den, rden);
  if (!use_limiter) return central;
  const R qL = *q;
  const R qR = fabs(gdiv<P2>(theta * dVR, dxr, rdxr));
  *q = qR;
  return (copysign(R(0.5), dVL) + copysign(R(0.5), dVR)) * fmin(fmin(qL, qR), fabs(central));
}

// ReconstructionType (IncrementDivergence_FV__Form.f90:239-241, 637-655).
enum Recon { RECON_LINEAR = 0, RECON_PARABOLIC = 1 };

// ComputeReconstructionParabolic_CSL_Kernel (EV/Increments/IncrementDivergence
// _FV__Kernel.f90:296-596) for one cell, with XA = X and X2A = X**2 as
// ComputeReconstructionParabolic_CSL passes them (IncrementDivergence_FV__
// Form.f90:852-931): the values at the cell's inner face xC - dXL (*vI) and
// outer face xC + dXR (*vO) from the values and centres of the cell and its
// neighbours.  A local extremum of the three values gives vC on both faces;
// otherwise the parabola through the three points, replaced on the side of
// a vertex inside the stencil by the parabola flat at the neighbour; then
// the parabola through (xI, vI), (xC, vC), (xO, vO), likewise replaced by one
// flat at xI or xO.  The reference's expression order: a*b*c = (a*b)*c,
// a+b+c = (a+b)+c, x**2 = x*x; -2*a*b/D rounds as ((-2*a)*b)/D.
//
// FP64 keeps the reference's absolute coordinates (bit parity).  FP32 takes
// coordinates relative to the cell centre: the parabolas and the vertex
// tests are invariant under the shift, and the reference's determinants,
// differences of O(x^3) terms that are O(dx^3), would otherwise lose about
// log10(x^2 / dx^3) digits, more than single precision holds.
template <typename R>
__device__ __forceinline__ void parabolic_faces(R vM, R vC, R vP, R xM, R xC, R xP, R dXL,
                                                R dXR, R *pI, R *pO) {
  if constexpr (sizeof(R) == sizeof(float)) {
    xM = xM - xC;
    xP = xP - xC;
    xC = R(0);
  }
  const R tiny = Arith<R>::tiny;  // SqrtTiny = tiny ( 0.0_KDR )
  const R xAM = xM, xAC = xC, xAP = xP;
  const R x2AM = xM * xM, x2AC = xC * xC, x2AP = xP * xP;
  const R xI = xC - dXL, xO = xC + dXR;
  R vI = vC, vO = vC;
  if (!((vC - vM) * (vP - vC) <= R(0))) {
    R D = ((x2AM - x2AP) * xAC + (x2AP - x2AC) * xAM) + (x2AC - x2AM) * xAP;
    R c0 = Arith<R>::div((vP * (x2AM * xAC - x2AC * xAM) + vM * (x2AC * xAP - x2AP * xAC)) +
                             vC * (x2AP * xAM - x2AM * xAP),
                         D);
    R c1 = Arith<R>::div((vP * (x2AC - x2AM) + vC * (x2AM - x2AP)) + vM * (x2AP - x2AC), D);
    R c2 = Arith<R>::div((vP * (xAM - xAC) + vM * (xAC - xAP)) + vC * (xAP - xAM), D);
    R xE = -Arith<R>::div(c1, R(2) * copysign(fmax(fabs(c2), tiny), c2));
    vI = (c0 + c1 * xI) + c2 * (xI * xI);
    vO = (c0 + c1 * xO) + c2 * (xO * xO);
    if (xE > xM && xE <= xC) {
      D = (x2AC - x2AM) + (R(2) * (xAM - xAC)) * xAM;
      c0 = Arith<R>::div((vM * x2AC - vC * x2AM) + (R(2) * (vC * xAM - vM * xAC)) * xAM, D);
      c1 = Arith<R>::div((R(-2) * (vC - vM)) * xAM, D);
      c2 = Arith<R>::div(vC - vM, D);
      vI = (c0 + c1 * xI) + c2 * (xI * xI);
    } else if (xE > xC && xE < xP) {
      D = (x2AC - x2AP) + (R(2) * (xAP - xAC)) * xAP;
      c0 = Arith<R>::div((vP * x2AC - vC * x2AP) + (R(2) * (vC * xAP - vP * xAC)) * xAP, D);
      c1 = Arith<R>::div((R(-2) * (vC - vP)) * xAP, D);
      c2 = Arith<R>::div(vC - vP, D);
      vO = (c0 + c1 * xO) + c2 * (xO * xO);
    }
    D = (xI - xO) * ((x2AC + xI * xO) - xAC * (xI + xO));
    c0 = Arith<R>::div(((vO * xI) * (x2AC - xAC * xI) + ((vC * xI) * xO) * (xI - xO)) +
                           (vI * xO) * (xAC * xO - x2AC),
                       D);
    c1 = Arith<R>::div((vO * (xI * xI - x2AC) + vI * (x2AC - xO * xO)) + vC * (xO * xO - xI * xI),
                       D);
    c2 = Arith<R>::div((vO * (xAC - xI) + vC * (xI - xO)) + vI * (xO - xAC), D);
    xE = -Arith<R>::div(c1, R(2) * copysign(fmax(fabs(c2), tiny), c2));
    vI = (c0 + c1 * xI) + c2 * (xI * xI);
    vO = (c0 + c1 * xO) + c2 * (xO * xO);
    if (xE > xI && xE <= xC) {
      D = (x2AC - xO * xO) + (R(2) * xI) * (xO - xAC);
      c0 = Arith<R>::div((vO * x2AC - vC * (xO * xO)) + (R(2) * xI) * (vC * xO - vO * xAC), D);
      c1 = Arith<R>::div((R(-2) * (vC - vO)) * xI, D);
      c2 = Arith<R>::div(vC - vO, D);
      vI = (c0 + c1 * xI) + c2 * (xI * xI);
    } else if (xE > xC && xE < xO) {
      D = (x2AC - xI * xI) + (R(2) * xO) * (xI - xAC);
      c0 = Arith<R>::div((vI * x2AC - vC * (xI * xI)) + (R(2) * xO) * (vC * xI - vI * xAC), D);
      c1 = Arith<R>::div((R(-2) * (vC - vI)) * xO, D);
      c2 = Arith<R>::div(vC - vI, D);
      vO = (c0 + c1 * xO) + c2 * (xO * xO);
    }
  }
  *pI = vI;
  *pO = vO;
}

// One reconstructed face state after ComputeFromPrimitiveCommon
// (F/Fluid_P_I__Form.f90:509-601): EOS (Fluid_P_I__Kernel.f90:197-271) on the
// unclipped density, then the N<0 floor (Fluid_D__Kernel.f90:87-88), D, S, G.
template <typename R> struct FaceState {
  R N, V1, V2, V3, E, P, CS, D, S1, S2, S3, G;
};

template <typename R>
__device__ __forceinline__ FaceState<R> face_state(const R W[5], R gamma, R mass) {
  FaceState<R> s;
  s.N = W[0];
  s.V1 = W[1];
  s.V2 = W[2];
  s.V3 = W[3];
  s.E = W[4];
  s.P = (gamma - R(1)) * s.E;
  // the EOS guard N > 0 and P > 0 selects between the formed sound speed
  // and 0 (no branch: the expensive side is the common one)
  const bool ok = s.N > R(0) && s.P > R(0);
  const R q = Arith<R>::div(gamma * s.P, mass * s.N);
  const R cs = Arith<R>::root(ok ? q : R(1));
  s.CS = ok ? cs : R(0);
  if (s.N < R(0)) s.N = R(0);
  s.D = s.N;
  s.S1 = mass * s.N * s.V1;
  s.S2 = mass * s.N * s.V2;
  s.S3 = mass * s.N * s.V3;
  s.G = s.E + R(0.5) * ((s.S1 * s.V1 + s.S2 * s.V2) + s.S3 * s.V3);
  return s;
}

template <typename R>
__device__ __forceinline__ R sel3(int d, R a, R b, R c) {
  return d == 0 ? a : (d == 1 ? b : c);
}

// Raw flux in dimension d (F/Fluid_D__Kernel.f90:220-262 then
// F/Fluid_P__Kernel.f90:671-709): D V, S_i V, S_d V + P, (G + P) V.
template <typename R>
__device__ __forceinline__ void raw_flux(R D, R S1, R S2, R S3, R G, R P, R Vd,
                                         int d, R F[5]) {
  F[0] = D * Vd;
  F[1] = S1 * Vd;
  F[2] = S2 * Vd;
  F[3] = S3 * Vd;
  if (d == 0) F[1] = F[1] + P;
  if (d == 1) F[2] = F[2] + P;
  if (d == 2) F[3] = F[3] + P;
  F[4] = (G + P) * Vd;
}

// Face flux of the sweep in dimension d from reconstructed primitive states.
// HLL (FLD/Current_Kernel.f90:50-137, DiffusionFactor 1); for HLLC the flux
// of the star state on the side of the contact speed replaces it when the
// face carries no shock flag and both solver speeds are non-zero
// (F/Fluid_P__Kernel.f90:319-668).  Only the flux the reference keeps is
// evaluated: the HLL flux is skipped where HLLC overwrites it, and only the
// star state that is selected is built.
template <typename R>
__device__ __forceinline__ void face_flux(const R WL[5], const R WR[5], int d, bool df,
                                          const Physics &ph, R F[5]) {
  const R tiny = Arith<R>::tiny;
  const R sqrt_tiny = Arith<R>::sqrt_tiny;
  const R gamma = R(ph.gamma), mass = R(ph.mass);
  const FaceState<R> L = face_state(WL, gamma, mass);
  const FaceState<R> Rt = face_state(WR, gamma, mass);
  const R vdL = sel3(d, L.V1, L.V2, L.V3);
  const R vdR = sel3(d, Rt.V1, Rt.V2, Rt.V3);
  const R AP = fmax(fmax(R(0), vdL + L.CS), vdR + Rt.CS);
  const R AM = fmax(fmax(R(0), -(vdL - L.CS)), -(vdR - Rt.CS));
  R FL[5], FR[5];
  raw_flux(L.D, L.S1, L.S2, L.S3, L.G, L.P, vdL, d, FL);
  raw_flux(Rt.D, Rt.S1, Rt.S2, Rt.S3, Rt.G, Rt.P, vdR, d, FR);
  if (ph.hllc && !df && AP != R(0) && AM != R(0)) {
    const R SdL = sel3(d, L.S1, L.S2, L.S3);
    const R SdR = sel3(d, Rt.S1, Rt.S2, Rt.S3);
    const R MD = AP * mass * Rt.D + AM * mass * L.D - FR[0] + FL[0];
    const R FSdL = sel3(d, FL[1], FL[2], FL[3]);
    const R FSdR = sel3(d, FR[1], FR[2], FR[3]);
    const R SN = AP * SdR + AM * SdL - FSdR + FSdL;
    const R MDinv = Arith<R>::div(fmax(MD, R(0)), fmax(MD * MD, tiny));
    const R AC = SN * MDinv;  // M_UU_11 = 1: 1 * SN is SN exactly
    // The star state on the side of the contact speed, written once for both
    // sides: side X = L with A = AM and s = +1 when AC >= 0, else X = R with
    // A = AP and s = -1.  A + s v is AM + vL or AP - vR, and (P + s a) - s b
    // is (P + a) - b or (P - a) + b, bit for bit (s x is exact), so a wave
    // whose faces take both sides runs one star state, not two.
    const bool lft = AC >= R(0);
    const R s = lft ? R(1) : R(-1);
    const R A = lft ? AM : AP;
    const R v = lft ? vdL : vdR;
    const R XD = lft ? L.D : Rt.D, XP = lft ? L.P : Rt.P, XG = lft ? L.G : Rt.G;
    const R XSd = lft ? SdL : SdR;
    const R AV = A + s * v;
    const R AAC = A + s * AC;
    const R inv = Arith<R>::div(R(1), fmax(fabs(AAC), sqrt_tiny));
    const R Ds = XD * AV * inv;
    const R V1 = d == 0 ? AC : (lft ? L.V1 : Rt.V1);
    const R V2 = d == 1 ? AC : (lft ? L.V2 : Rt.V2);
    const R V3 = d == 2 ? AC : (lft ? L.V3 : Rt.V3);
    const R S1 = mass * Ds * V1;
    const R S2 = mass * Ds * V2;
    const R S3 = mass * Ds * V3;
    const R Sds = sel3(d, S1, S2, S3);
    const R Ps = (XP + s * (XSd * AV)) - s * (Sds * AAC);
    const R Gs = ((XG * AV + s * (v * XP)) - s * (AC * Ps)) * inv;
    raw_flux(Ds, S1, S2, S3, Gs, Ps, AC, d, F);
    return;
  }
  const R UL[5] = {L.D, L.S1, L.S2, L.S3, L.G};
  const R UR[5] = {Rt.D, Rt.S1, Rt.S2, Rt.S3, Rt.G};
  const R den = fmax(AP + AM, tiny);
  const R APAM = AP * AM;  // DiffusionFactor 1: (1 * AP) * AM is AP * AM exactly
#pragma unroll
  for (int f = 0; f < 5; f++)
    F[f] = Arith<R>::div(AP * FL[f] + AM * FR[f] - APAM * (UR[f] - UL[f]), den);
}

// Conserved fix of ComputeFromConserved on the new state (FP64), returning
// the primitive velocity and internal energy it implies (F/Fluid_P__Kernel.
// f90:94-192).
__device__ __forceinline__ void fix_conserved(double U[5], double mass,
                                              double nmin, double V[5]) {
  const double D = U[0];
  const bool live = D > nmin;
  double v[3];
  velocity_of(U, mass * D, live, v);
  const double E = U[4] - 0.5 * ((U[1] * v[0] + U[2] * v[1]) + U[3] * v[2]);
  V[0] = live ? D : nmin;
  V[1] = live ? v[0] : 0.0;
  V[2] = live ? v[1] : 0.0;
  V[3] = live ? v[2] : 0.0;
  V[4] = live ? E : 0.0;
  U[0] = live ? U[0] : nmin;
  U[1] = live ? U[1] : 0.0;
  U[2] = live ? U[2] : 0.0;
  U[3] = live ? U[3] : 0.0;
  U[4] = live ? U[4] : 0.0;
  // E < 0: E = 0 and G = 0 + S.V/2 (the repaired value)
  const bool neg = V[4] < 0.0;
  const double G = 0.0 + 0.5 * ((U[1] * V[1] + U[2] * V[2]) + U[3] * V[3]);
  V[4] = neg ? 0.0 : V[4];
  U[4] = neg ? G : U[4];
}

// Per-cell CFL term (Integrator_C_PS__Kernel.f90:12-144, nDimensions 3,
// Crsn = 1) from primitives, with the EOS sound speed, in FP64.
__device__ __forceinline__ double cfl_term(const double V[5], double gamma,
                                          double mass, double w1, double w2,
                                          double w3) {
  const double P = (gamma - 1.0) * V[4];
  const bool ok = V[0] > 0.0 && P > 0.0;
  const double q = gamma * P / (mass * V[0]);
  const double r = sqrt(ok ? q : 1.0);  // both formed, the EOS guard selects
  const double CS = ok ? r : 0.0;
  const double t1 = fmax(V[1] + CS, -(V[1] - CS)) / w1;
  const double t2 = fmax(V[2] + CS, -(V[2] - CS)) / w2;
  const double t3 = fmax(V[3] + CS, -(V[3] - CS)) / w3;
  return t1 + t2 + t3;
}

}  // namespace gx
