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
  N_GEO
};

// a/b rounded to nearest, given y = RN(1/b): q0 = RN(a y) is corrected twice
// with the exact-remainder fma; after the first correction q is within one
// ulp of a/b, so the remainder of the second is exact and its correction
// rounds to RN(a/b) (Markstein, IBM J. Res. Dev. 34 (1990) 111-119).  The
// quotient, a and the remainders must be normal: callers keep b in
// [2^-500, 2^500] and route |a| < 2^-900 to the IEEE division.
__device__ __forceinline__ double div_rcp(double a, double b, double y) {
  if (fabs(a) < 0x1p-900) return a / b;
  double q = a * y;
  double r = fma(-b, q, a);
  q = fma(r, y, q);
  r = fma(-b, q, a);
  return fma(r, y, q);
}

struct Physics {
  double gamma, theta, mass, nmin, st;
  int use_limiter, hllc;
};

// Fluid_P_I % ComputeFromConservedCommon, primitive part (F/Fluid_P__Kernel.
// f90:94-192): N, V, E from D, S, G with the density floor and E >= 0.
__device__ __forceinline__ void prim_from_cons(const double U[5], double mass,
                                               double nmin, double V[5]) {
  const double D = U[0];
  if (D > nmin) {
    const double MD = mass * D;
    V[0] = D;
    if (MD > 0x1p-500 && MD < 0x1p500) {
      const double y = 1.0 / MD;
      V[1] = div_rcp(U[1], MD, y);
      V[2] = div_rcp(U[2], MD, y);
      V[3] = div_rcp(U[3], MD, y);
    } else {
      V[1] = U[1] / MD;
      V[2] = U[2] / MD;
      V[3] = U[3] / MD;
    }
    V[4] = U[4] - 0.5 * ((U[1] * V[1] + U[2] * V[2]) + U[3] * V[3]);
  } else {
    V[0] = nmin;
    V[1] = 0.0;
    V[2] = 0.0;
    V[3] = 0.0;
    V[4] = 0.0;
  }
  if (V[4] < 0.0) V[4] = 0.0;
}

// Gradient_Kernel limited slope (OP/Calculus/Derivatives/Gradient_Kernel.f90:
// 166-190), or the unlimited centred slope when UseLimiter is false.
// The three divisions use the precomputed reciprocals of the geometry.
__device__ __forceinline__ double slope(double vm, double v0, double vp,
                                        double dxl, double dxr, double dxl2,
                                        double dxr2, double den, double rdxl,
                                        double rdxr, double rden, double theta,
                                        int use_limiter) {
  const double dVL = v0 - vm;
  const double dVR = vp - v0;
  const double central = div_rcp(dxr2 * dVL + dxl2 * dVR, den, rden);
  if (!use_limiter) return central;
  return (copysign(0.5, dVL) + copysign(0.5, dVR)) *
         fmin(fmin(fabs(div_rcp(theta * dVL, dxl, rdxl)),
                   fabs(div_rcp(theta * dVR, dxr, rdxr))),
              fabs(central));
}

// One reconstructed face state after ComputeFromPrimitiveCommon
// (F/Fluid_P_I__Form.f90:509-601): EOS (Fluid_P_I__Kernel.f90:197-271) on the
// unclipped density, then the N<0 floor (Fluid_D__Kernel.f90:87-88), D, S, G.
struct FaceState {
  double N, V1, V2, V3, E, P, CS, D, S1, S2, S3, G;
};

__device__ __forceinline__ FaceState face_state(const double W[5],
                                                const Physics &ph) {
  FaceState s;
  s.N = W[0];
  s.V1 = W[1];
  s.V2 = W[2];
  s.V3 = W[3];
  s.E = W[4];
  s.P = (ph.gamma - 1.0) * s.E;
  if (s.N > 0.0 && s.P > 0.0)
    s.CS = sqrt(ph.gamma * s.P / (ph.mass * s.N));
  else
    s.CS = 0.0;
  if (s.N < 0.0) s.N = 0.0;
  s.D = s.N;
  s.S1 = ph.mass * s.N * s.V1;
  s.S2 = ph.mass * s.N * s.V2;
  s.S3 = ph.mass * s.N * s.V3;
  s.G = s.E + 0.5 * ((s.S1 * s.V1 + s.S2 * s.V2) + s.S3 * s.V3);
  return s;
}

__device__ __forceinline__ double sel3(int d, double a, double b, double c) {
  return d == 0 ? a : (d == 1 ? b : c);
}

// Raw flux in dimension d (F/Fluid_D__Kernel.f90:220-262 then
// F/Fluid_P__Kernel.f90:671-709): D V, S_i V, S_d V + P, (G + P) V.
__device__ __forceinline__ void raw_flux(double D, double S1, double S2,
                                         double S3, double G, double P,
                                         double Vd, int d, double F[5]) {
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
__device__ __forceinline__ void face_flux(const double WL[5],
                                          const double WR[5], int d, bool df,
                                          const Physics &ph, double F[5]) {
  const double tiny = DBL_MIN;
  const double sqrt_tiny = 1.4916681462400413e-154;  // sqrt(DBL_MIN), exact
  const FaceState L = face_state(WL, ph);
  const FaceState R = face_state(WR, ph);
  const double vdL = sel3(d, L.V1, L.V2, L.V3);
  const double vdR = sel3(d, R.V1, R.V2, R.V3);
  const double AP = fmax(fmax(0.0, vdL + L.CS), vdR + R.CS);
  const double AM = fmax(fmax(0.0, -(vdL - L.CS)), -(vdR - R.CS));
  double FL[5], FR[5];
  raw_flux(L.D, L.S1, L.S2, L.S3, L.G, L.P, vdL, d, FL);
  raw_flux(R.D, R.S1, R.S2, R.S3, R.G, R.P, vdR, d, FR);
  if (ph.hllc && !df && AP != 0.0 && AM != 0.0) {
    const double SdL = sel3(d, L.S1, L.S2, L.S3);
    const double SdR = sel3(d, R.S1, R.S2, R.S3);
    const double MD = AP * ph.mass * R.D + AM * ph.mass * L.D - FR[0] + FL[0];
    const double FSdL = sel3(d, FL[1], FL[2], FL[3]);
    const double FSdR = sel3(d, FR[1], FR[2], FR[3]);
    const double SN = AP * SdR + AM * SdL - FSdR + FSdL;
    const double MDinv = fmax(MD, 0.0) / fmax(MD * MD, tiny);
    const double AC = 1.0 * SN * MDinv;
    if (AC >= 0.0) {
      const double AM_VL = AM + vdL;
      const double AM_AC = AM + AC;
      const double inv = 1.0 / fmax(fabs(AM_AC), sqrt_tiny);
      const double Ds = L.D * AM_VL * inv;
      const double V1 = d == 0 ? AC : L.V1;
      const double V2 = d == 1 ? AC : L.V2;
      const double V3 = d == 2 ? AC : L.V3;
      const double S1 = ph.mass * Ds * V1;
      const double S2 = ph.mass * Ds * V2;
      const double S3 = ph.mass * Ds * V3;
      const double Sds = sel3(d, S1, S2, S3);
      const double Ps = L.P + SdL * AM_VL - Sds * AM_AC;
      const double Gs = (L.G * AM_VL + vdL * L.P - AC * Ps) * inv;
      raw_flux(Ds, S1, S2, S3, Gs, Ps, AC, d, F);
    } else {
      const double AP_VR = AP - vdR;
      const double AP_AC = AP - AC;
      const double inv = 1.0 / fmax(fabs(AP_AC), sqrt_tiny);
      const double Ds = R.D * AP_VR * inv;
      const double V1 = d == 0 ? AC : R.V1;
      const double V2 = d == 1 ? AC : R.V2;
      const double V3 = d == 2 ? AC : R.V3;
      const double S1 = ph.mass * Ds * V1;
      const double S2 = ph.mass * Ds * V2;
      const double S3 = ph.mass * Ds * V3;
      const double Sds = sel3(d, S1, S2, S3);
      const double Ps = R.P - SdR * AP_VR + Sds * AP_AC;
      const double Gs = (R.G * AP_VR - vdR * R.P + AC * Ps) * inv;
      raw_flux(Ds, S1, S2, S3, Gs, Ps, AC, d, F);
    }
    return;
  }
  const double UL[5] = {L.D, L.S1, L.S2, L.S3, L.G};
  const double UR[5] = {R.D, R.S1, R.S2, R.S3, R.G};
  const double den = fmax(AP + AM, tiny);
#pragma unroll
  for (int f = 0; f < 5; f++)
    F[f] = (AP * FL[f] + AM * FR[f] - 1.0 * AP * AM * (UR[f] - UL[f])) / den;
}

// Conserved fix of ComputeFromConserved on the new state, returning the
// primitive velocity and internal energy it implies (F/Fluid_P__Kernel.f90:
// 94-192).
__device__ __forceinline__ void fix_conserved(double U[5], double mass,
                                              double nmin, double V[5]) {
  const double D = U[0];
  if (D > nmin) {
    V[0] = D;
    V[1] = U[1] / (mass * D);
    V[2] = U[2] / (mass * D);
    V[3] = U[3] / (mass * D);
    V[4] = U[4] - 0.5 * ((U[1] * V[1] + U[2] * V[2]) + U[3] * V[3]);
  } else {
    V[0] = nmin;
    V[1] = 0.0;
    V[2] = 0.0;
    V[3] = 0.0;
    V[4] = 0.0;
    U[0] = nmin;
    U[1] = 0.0;
    U[2] = 0.0;
    U[3] = 0.0;
    U[4] = 0.0;
  }
  if (V[4] < 0.0) {
    V[4] = 0.0;
    U[4] = V[4] + 0.5 * ((U[1] * V[1] + U[2] * V[2]) + U[3] * V[3]);
  }
}

// Per-cell CFL term (Integrator_C_PS__Kernel.f90:12-144, nDimensions 3,
// Crsn = 1) from primitives, with the EOS sound speed.
__device__ __forceinline__ double cfl_term(const double V[5], double gamma,
                                          double mass, double w1, double w2,
                                          double w3) {
  const double P = (gamma - 1.0) * V[4];
  double CS = 0.0;
  if (V[0] > 0.0 && P > 0.0) CS = sqrt(gamma * P / (mass * V[0]));
  const double t1 = fmax(V[1] + CS, -(V[1] - CS)) / w1;
  const double t2 = fmax(V[2] + CS, -(V[2] - CS)) / w2;
  const double t3 = fmax(V[3] + CS, -(V[3] - CS)) / w3;
  return t1 + t2 + t3;
}

}  // namespace gx
