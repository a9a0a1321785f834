// gx_adjoint.hpp -- reverse-mode derivatives (vector-Jacobian products) of
// the pieces of the FP64 Fluid_P_I step, host- and device-callable.
//
// Each function recomputes the forward values of its piece in the forward
// kernels' expression order, so it takes the branch the forward took, then
// runs the reverse sweep.  Rules at ties (where the forward value is the
// same either way): fmax/fmin take the first operand; the limiter takes the
// centred slope when it equals the smaller one-sided slope; |x| has
// derivative copysign(1, x); where exactly one of the two differences of a
// limited slope is zero the slope is one-sided linear in it and its
// derivative is the mean of the two sides, theta/(2 dx).  Shock flags and the
// HLLC / HLL choice are replayed as constants.
#pragma once

#include <cfloat>
#include <cmath>

#ifdef __HIPCC__
#define GX_HD __host__ __device__ inline
#else
#define GX_HD inline
#endif

namespace gx {
namespace adj {

constexpr double kTiny = DBL_MIN;
constexpr double kSqrtTiny = 1.4916681462400413e-154;

GX_HD double sgn1(double x) { return std::signbit(x) ? -1.0 : 1.0; }

// On the device the face flux's reverse lets the compiler fuse a * b + c
// into one FMA: its values are gradients, or recomputed forward values whose
// rounding picks no branch the forward replays exactly (those already come
// from the refined hardware reciprocal and square root below).  The limited
// slope's quotients and the face states are formed outside these functions,
// in the forward's order.  The host keeps separate roundings, so its
// evaluation repeats the forward bit for bit.
#ifdef __HIP_DEVICE_COMPILE__
#define GX_REV_FMA _Pragma("clang fp contract(fast)")
#else
#define GX_REV_FMA
#endif

// 1/x for the reverse sweeps' multipliers (x positive and normal), where a
// gradient needs no correctly rounded quotient: on the device the hardware
// reciprocal refined by two Newton steps (within an ulp or two), on the host
// the division.
GX_HD double rcp_rev(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  return fma(y, fma(-x, y, 1.0), y);
#else
  return 1.0 / x;
#endif
}

// The forward values a face flux's reverse recomputes (sound speed, contact
// speed, star and HLL denominators): on the host the forward's IEEE division
// and square root (so the host evaluation repeats the forward bit for bit);
// on the device a times rb = rcp_rev(b), and a square root from the hardware
// reciprocal square root refined by Goldschmidt steps, within an ulp or two,
// which moves a gradient by rounding only (identical inputs still give
// identical values, so the ties of identical states stay ties).
GX_HD double div_rev(double a, double b, double rb) {
#ifdef __HIP_DEVICE_COMPILE__
  (void)b;
  return a * rb;
#else
  (void)rb;
  return a / b;
#endif
}

// sqrt(x) for x > 0 and normal; half_rcp receives 1/(2 sqrt(x)) for the
// reverse sweep (on the device the Goldschmidt iterate, one refinement of
// the hardware estimate).
GX_HD double sqrt_rev(double x, double &half_rcp) {
#ifdef __HIP_DEVICE_COMPILE__
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  half_rcp = h;
  return fma(fma(-g, g, x), h, g);
#else
  const double c = sqrt(x);
  half_rcp = 0.5 / c;
  return c;
#endif
}

// ---- limited slope (Gradient_Kernel.f90:166-190) --------------------------
// s = (sign(1/2, dVL) + sign(1/2, dVR)) * min(|theta dVL/dxl|, |theta dVR/dxr|,
// |(dxr^2 dVL + dxl^2 dVR)/den|); adds sb * ds/d(vm, v0, vp) to the bars.
GX_HD void slope_vjp(double vm, double v0, double vp, double dxl, double dxr,
                     double dxl2, double dxr2, double den, double theta,
                     int use_limiter, double sb, double &vmb, double &v0b,
                     double &vpb) {
  const double dVL = v0 - vm, dVR = vp - v0;
  double dVLb = 0.0, dVRb = 0.0;
  if (!use_limiter) {
    dVLb = sb * dxr2 / den;
    dVRb = sb * dxl2 / den;
  } else {
    const bool zL = dVL == 0.0, zR = dVR == 0.0;
    if (zL && !zR) {
      dVLb = sb * theta / (2.0 * dxl);
    } else if (zR && !zL) {
      dVRb = sb * theta / (2.0 * dxr);
    } else {
      const double c = (dxr2 * dVL + dxl2 * dVR) / den;
      const double ql = theta * dVL / dxl, qr = theta * dVR / dxr;
      const double tl = fabs(ql), tr = fabs(qr);
      const double m1 = tl <= tr ? tl : tr;
      const double factor = (std::signbit(dVL) ? -0.5 : 0.5) + (std::signbit(dVR) ? -0.5 : 0.5);
      const double mb = sb * factor;
      if (fabs(c) <= m1) {
        const double cb = mb * sgn1(c);
        dVLb = cb * dxr2 / den;
        dVRb = cb * dxl2 / den;
      } else if (tl <= tr) {
        dVLb = mb * sgn1(ql) * theta / dxl;
      } else {
        dVRb = mb * sgn1(qr) * theta / dxr;
      }
    }
  }
  vmb -= dVLb;
  v0b += dVLb - dVRb;
  vpb += dVRb;
}

// ---- face state after ComputeFromPrimitiveCommon --------------------------
struct FState {
  double N, V[3], E, P, q, CS, Nc, D, S[3], G;
  double rMN, hCS;  // 1/(M N) and 1/(2 CS), where ok
  bool ok;
};

struct FBar {
  double N = 0, V[3] = {0, 0, 0}, E = 0, P = 0, CS = 0, D = 0, S[3] = {0, 0, 0}, G = 0;
};

GX_HD void fstate(const double W[5], double gamma, double mass, FState &s) {
  GX_REV_FMA
  s.N = W[0];
  s.V[0] = W[1];
  s.V[1] = W[2];
  s.V[2] = W[3];
  s.E = W[4];
  s.P = (gamma - 1.0) * s.E;
  s.ok = s.N > 0.0 && s.P > 0.0;
  const double MN = s.ok ? mass * s.N : 1.0;
  s.rMN = rcp_rev(MN);
  s.q = s.ok ? div_rev(gamma * s.P, MN, s.rMN) : 0.0;
  s.CS = s.ok ? sqrt_rev(s.ok ? s.q : 1.0, s.hCS) : 0.0;
  s.Nc = s.N < 0.0 ? 0.0 : s.N;
  s.D = s.Nc;
  for (int k = 0; k < 3; k++) s.S[k] = mass * s.Nc * s.V[k];
  s.G = s.E + 0.5 * ((s.S[0] * s.V[0] + s.S[1] * s.V[1]) + s.S[2] * s.V[2]);
}

// Bars of the face state -> bars of its primitive W (added) and gamma.
GX_HD void fstate_vjp(const FState &s, FBar b, double gamma, double mass, double Wb[5],
                      double &gb) {
  GX_REV_FMA
  b.E += b.G;
  for (int k = 0; k < 3; k++) {
    b.S[k] += 0.5 * b.G * s.V[k];
    b.V[k] += 0.5 * b.G * s.S[k];
  }
  double Ncb = b.D;
  for (int k = 0; k < 3; k++) {
    Ncb += b.S[k] * mass * s.V[k];
    b.V[k] += b.S[k] * mass * s.Nc;
  }
  double Nb = s.N < 0.0 ? 0.0 : Ncb;
  if (s.ok && s.CS > 0.0) {
    // q = gamma P / (M N), CS = sqrt(q); dq/dN = -q/N = -q M/(M N)
    const double qb = b.CS * s.hCS;
    gb += (qb * s.P) * s.rMN;
    b.P += (qb * gamma) * s.rMN;
    Nb -= (qb * s.q) * (mass * s.rMN);
  }
  b.E += b.P * (gamma - 1.0);
  gb += b.P * s.E;
  Wb[0] += Nb;
  for (int k = 0; k < 3; k++) Wb[1 + k] += b.V[k];
  Wb[4] += b.E;
}

GX_HD void raw(double D, const double S[3], double G, double P, double vd, int d,
               double F[5]) {
  GX_REV_FMA
  F[0] = D * vd;
  F[1] = S[0] * vd;
  F[2] = S[1] * vd;
  F[3] = S[2] * vd;
  F[1 + d] = F[1 + d] + P;
  F[4] = (G + P) * vd;
}

GX_HD void raw_vjp(double D, const double S[3], double G, double P, double vd, int d,
                   const double Fb[5], double &Db, double Sb[3], double &Gb, double &Pb,
                   double &vdb) {
  GX_REV_FMA
  Db += Fb[0] * vd;
  vdb += Fb[0] * D;
  for (int k = 0; k < 3; k++) {
    Sb[k] += Fb[1 + k] * vd;
    vdb += Fb[1 + k] * S[k];
  }
  Pb += Fb[1 + d];
  Gb += Fb[4] * vd;
  Pb += Fb[4] * vd;
  vdb += Fb[4] * (G + P);
}

// ---- face flux (HLL / HLLC) -------------------------------------------------
// F (returned) and the bars of WL, WR (added) and gamma for a given Fb.
GX_HD void face_flux_vjp(const double WL[5], const double WR[5], int d, bool df,
                         double gamma, double mass, bool hllc, const double Fb[5],
                         double WLb[5], double WRb[5], double &gb, double F[5]) {
  GX_REV_FMA
  FState L, R;
  fstate(WL, gamma, mass, L);
  fstate(WR, gamma, mass, R);
  FBar Lb, Rb;
  const double vdL = L.V[d], vdR = R.V[d];
  double vdLb = 0.0, vdRb = 0.0, APb = 0.0, AMb = 0.0;
  const double aL = vdL + L.CS, aR = vdR + R.CS;
  const bool pL0 = 0.0 >= aL;
  const double m1p = pL0 ? 0.0 : aL;
  const bool pAR = !(m1p >= aR);
  const double AP = pAR ? aR : m1p;
  const double bL = -(vdL - L.CS), bR = -(vdR - R.CS);
  const bool mL0 = 0.0 >= bL;
  const double m1m = mL0 ? 0.0 : bL;
  const bool mBR = !(m1m >= bR);
  const double AM = mBR ? bR : m1m;
  double FL[5], FR[5];
  raw(L.D, L.S, L.G, L.P, vdL, d, FL);
  raw(R.D, R.S, R.G, R.P, vdR, d, FR);
  double FLb[5] = {0, 0, 0, 0, 0}, FRb[5] = {0, 0, 0, 0, 0};
  const bool star = hllc && !df && AP != 0.0 && AM != 0.0;
  if (star) {
    const double SdL = L.S[d], SdR = R.S[d];
    const double MD = AP * mass * R.D + AM * mass * L.D - FR[0] + FL[0];
    const double SN = AP * SdR + AM * SdL - FR[1 + d] + FL[1 + d];
    const double nn = MD >= 0.0 ? MD : 0.0;
    const double dd = MD * MD >= kTiny ? MD * MD : kTiny;
    const double rdd = rcp_rev(dd);
    const double MDinv = div_rev(nn, dd, rdd);
    const double AC = 1.0 * SN * MDinv;
    double ACb = 0.0;
    // The star state on the side of the contact speed, written once for
    // both sides: side X = L with A = AM and s = +1 when AC >= 0, else X = R
    // with A = AP and s = -1.  A + s v is AM + vL or AP - vR, and
    // (P + s a) - s b is (P + a) - b or (P - a) + b, bit for bit.
    {
      const bool lft = AC >= 0.0;
      const double s = lft ? 1.0 : -1.0;
      const double A = lft ? AM : AP;
      const double v = lft ? vdL : vdR;
      const double XD = lft ? L.D : R.D, XP = lft ? L.P : R.P, XG = lft ? L.G : R.G;
      const double XSd = lft ? SdL : SdR;
      double XV[3];
      for (int k = 0; k < 3; k++) XV[k] = lft ? L.V[k] : R.V[k];
      const double AV = A + s * v, AAC = A + s * AC;
      const double aa = fabs(AAC);
      const double den = aa >= kSqrtTiny ? aa : kSqrtTiny;
      const double inv = rcp_rev(den);
      const double Ds = XD * AV * inv;
      double Vs[3], Ss[3];
      for (int k = 0; k < 3; k++) {
        Vs[k] = k == d ? AC : XV[k];
        Ss[k] = mass * Ds * Vs[k];
      }
      const double Ps = (XP + s * (XSd * AV)) - s * (Ss[d] * AAC);
      const double t = (XG * AV + s * (v * XP)) - s * (AC * Ps);
      const double Gs = t * inv;
      raw(Ds, Ss, Gs, Ps, AC, d, F);
      double Dsb = 0.0, Ssb[3] = {0, 0, 0}, Gsb = 0.0, Psb = 0.0;
      raw_vjp(Ds, Ss, Gs, Ps, AC, d, Fb, Dsb, Ssb, Gsb, Psb, ACb);
      FBar Xb;
      double AVb = 0.0, AACb = 0.0, vb = 0.0;
      const double tb = Gsb * inv;
      double invb = Gsb * t;
      Xb.G += tb * AV;
      AVb += tb * XG;
      vb += s * (tb * XP);
      Xb.P += s * (tb * v);
      ACb -= s * (tb * Ps);
      Psb -= s * (tb * AC);
      Xb.P += Psb;
      Xb.S[d] += s * (Psb * AV);
      AVb += s * (Psb * XSd);
      Ssb[d] -= s * (Psb * AAC);
      AACb -= s * (Psb * Ss[d]);
      for (int k = 0; k < 3; k++) {
        Dsb += Ssb[k] * mass * Vs[k];
        const double vk = Ssb[k] * mass * Ds;
        if (k == d)
          ACb += vk;
        else
          Xb.V[k] += vk;
      }
      Xb.D += Dsb * AV * inv;
      AVb += Dsb * XD * inv;
      invb += Dsb * XD * AV;
      if (aa >= kSqrtTiny) AACb += -invb * inv * inv * sgn1(AAC);
      ACb += s * AACb;
      vb += s * AVb;
      const double Ab = AACb + AVb;
      // scatter onto the side's bars
      auto onL = [lft](double x) { return lft ? x : 0.0; };
      auto onR = [lft](double x) { return lft ? 0.0 : x; };
      AMb += onL(Ab);
      APb += onR(Ab);
      vdLb += onL(vb);
      vdRb += onR(vb);
      Lb.D += onL(Xb.D);
      Rb.D += onR(Xb.D);
      Lb.P += onL(Xb.P);
      Rb.P += onR(Xb.P);
      Lb.G += onL(Xb.G);
      Rb.G += onR(Xb.G);
      for (int k = 0; k < 3; k++) {
        Lb.S[k] += onL(Xb.S[k]);
        Rb.S[k] += onR(Xb.S[k]);
        Lb.V[k] += onL(Xb.V[k]);
        Rb.V[k] += onR(Xb.V[k]);
      }
    }
    // AC = (1*SN)*MDinv, MDinv = max(MD, 0)/max(MD^2, tiny)
    const double SNb = ACb * MDinv;
    const double MDinvb = ACb * SN;
    const double nnb = MDinvb * rdd;
    const double ddb = (-MDinvb * MDinv) * rdd;
    double MDb = 0.0;
    if (MD >= 0.0) MDb += nnb;
    if (MD * MD >= kTiny) MDb += ddb * 2.0 * MD;
    APb += SNb * SdR;
    Rb.S[d] += SNb * AP;
    AMb += SNb * SdL;
    Lb.S[d] += SNb * AM;
    FRb[1 + d] -= SNb;
    FLb[1 + d] += SNb;
    APb += MDb * mass * R.D;
    Rb.D += MDb * AP * mass;
    AMb += MDb * mass * L.D;
    Lb.D += MDb * AM * mass;
    FRb[0] -= MDb;
    FLb[0] += MDb;
  } else {
    const double UL[5] = {L.D, L.S[0], L.S[1], L.S[2], L.G};
    const double UR[5] = {R.D, R.S[0], R.S[1], R.S[2], R.G};
    const double sum = AP + AM;
    const double den = sum >= kTiny ? sum : kTiny;
    double denb = 0.0, ULb[5], URb[5];
    const double rden = rcp_rev(den);
    for (int f = 0; f < 5; f++) {
      const double dU = UR[f] - UL[f];
      F[f] = div_rev(AP * FL[f] + AM * FR[f] - 1.0 * AP * AM * dU, den, rden);
      const double numb = Fb[f] * rden;
      denb -= (Fb[f] * F[f]) * rden;
      APb += numb * (FL[f] - AM * dU);
      AMb += numb * (FR[f] - AP * dU);
      FLb[f] += numb * AP;
      FRb[f] += numb * AM;
      URb[f] = -numb * AP * AM;
      ULb[f] = numb * AP * AM;
    }
    if (sum >= kTiny) {
      APb += denb;
      AMb += denb;
    }
    Lb.D += ULb[0];
    Rb.D += URb[0];
    for (int k = 0; k < 3; k++) {
      Lb.S[k] += ULb[1 + k];
      Rb.S[k] += URb[1 + k];
    }
    Lb.G += ULb[4];
    Rb.G += URb[4];
  }
  raw_vjp(L.D, L.S, L.G, L.P, vdL, d, FLb, Lb.D, Lb.S, Lb.G, Lb.P, vdLb);
  raw_vjp(R.D, R.S, R.G, R.P, vdR, d, FRb, Rb.D, Rb.S, Rb.G, Rb.P, vdRb);
  // AP = max(max(0, aL), aR), AM = max(max(0, bL), bR)
  double aLb = 0.0, aRb = 0.0, bLb = 0.0, bRb = 0.0;
  if (pAR)
    aRb = APb;
  else if (!pL0)
    aLb = APb;
  if (mBR)
    bRb = AMb;
  else if (!mL0)
    bLb = AMb;
  vdLb += aLb - bLb;
  Lb.CS += aLb + bLb;
  vdRb += aRb - bRb;
  Rb.CS += aRb + bRb;
  Lb.V[d] += vdLb;
  Rb.V[d] += vdRb;
  fstate_vjp(L, Lb, gamma, mass, WLb, gb);
  fstate_vjp(R, Rb, gamma, mass, WRb, gb);
}

// ---- conserved -> primitive (Fluid_P__Kernel.f90:94-192) -----------------
// V = (N, V_1..3, E) from U with the density floor and E >= 0; Ub += Vb dV/dU.
GX_HD void prim_vjp(const double U[5], double mass, double nmin, const double Vb[5],
                    double Ub[5]) {
  const double D = U[0];
  if (!(D > nmin)) return;
  const double MD = mass * D;
  const double V1 = U[1] / MD, V2 = U[2] / MD, V3 = U[3] / MD;
  const double E = U[4] - 0.5 * ((U[1] * V1 + U[2] * V2) + U[3] * V3);
  const double Eb = E < 0.0 ? 0.0 : Vb[4];
  const double V[3] = {V1, V2, V3};
  Ub[4] += Eb;
  double MDb = 0.0;
  for (int k = 0; k < 3; k++) {
    Ub[1 + k] -= 0.5 * Eb * V[k];
    const double vb = Vb[1 + k] - 0.5 * Eb * U[1 + k];
    Ub[1 + k] += vb / MD;
    MDb -= vb * V[k] / MD;
  }
  Ub[0] += Vb[0] + MDb * mass;
}

// ---- StoreSolution's conserved fix (Fluid_P__Kernel.f90:94-192) ----------
// U' = fix(Z); Zb = (dU'/dZ)^T Ub.
GX_HD void fix_vjp(const double Z[5], double mass, double nmin, const double Ub[5],
                   double Zb[5]) {
  for (int f = 0; f < 5; f++) Zb[f] = 0.0;
  const double D = Z[0];
  if (!(D > nmin)) return;
  const double MD = mass * D;
  const double V[3] = {Z[1] / MD, Z[2] / MD, Z[3] / MD};
  const double E = Z[4] - 0.5 * ((Z[1] * V[0] + Z[2] * V[1]) + Z[3] * V[2]);
  if (!(E < 0.0)) {
    for (int f = 0; f < 5; f++) Zb[f] = Ub[f];
    return;
  }
  // G' = 0 + 1/2 sum_k Z_k Z_k / (M D)
  const double Gp = 0.0 + 0.5 * ((Z[1] * V[0] + Z[2] * V[1]) + Z[3] * V[2]);
  Zb[0] = Ub[0] - Ub[4] * Gp / D;
  for (int k = 0; k < 3; k++) Zb[1 + k] = Ub[1 + k] + Ub[4] * V[k];
  Zb[4] = 0.0;
}

// ---- CFL term (Integrator_C_PS__Kernel.f90:12-144) ------------------------
// term = sum_d max(V_d + CS, -(V_d - CS)) / w_d; Vb, gb += tb dterm/d(V, gamma).
GX_HD void cfl_term_vjp(const double V[5], double gamma, double mass, const double w[3],
                        double tb, double Vb[5], double &gb) {
  const double P = (gamma - 1.0) * V[4];
  const bool ok = V[0] > 0.0 && P > 0.0;
  const double q = ok ? gamma * P / (mass * V[0]) : 0.0;
  const double CS = ok ? sqrt(q) : 0.0;
  double CSb = 0.0;
  for (int d = 0; d < 3; d++) {
    const double a = V[1 + d] + CS, b = -(V[1 + d] - CS);
    const double xb = tb / w[d];
    if (a >= b) {
      Vb[1 + d] += xb;
    } else {
      Vb[1 + d] -= xb;
    }
    CSb += xb;
  }
  if (ok && CS > 0.0) {
    const double qb = CSb * 0.5 / CS;
    gb += qb * P / (mass * V[0]);
    const double Pb = qb * gamma / (mass * V[0]);
    Vb[0] -= qb * q / V[0];
    Vb[4] += Pb * (gamma - 1.0);
    gb += Pb * V[4];
  }
}

}  // namespace adj
}  // namespace gx
