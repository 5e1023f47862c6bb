// mcf_prec.h -- point routines of the sweep, generic over the compute type R.
//
// R = double is the strict path: every expression keeps the reference's
// evaluation order and IEEE division / square root, so it is bit-identical
// to the FP64 routines of mcf_device.h.
// R = float is the mixed path: reconstruction, Riemann solver and
// characteristic (Alfven) states in FP32 with the hardware reciprocal and
// square root (v_rcp_f32 / v_sqrt_f32, 1 ulp).  Differences of neighbouring
// state values are taken in FP64 before rounding to FP32, so a smooth
// profile keeps its slope to FP32 relative accuracy.
#pragma once
#include "mcf_device.h"

namespace mcf {

template <class R> struct Ar;
template <> struct Ar<double> {
  static constexpr bool strict = true;
  __device__ static __forceinline__ double div(double a, double b) { return a / b; }
  __device__ static __forceinline__ double rsqrt_(double a) { return 1.0 / sqrt(a); }
  __device__ static __forceinline__ double sqrt_(double a) { return sqrt(a); }
};
template <> struct Ar<float> {
  static constexpr bool strict = false;
  __device__ static __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  __device__ static __forceinline__ float sqrt_(float a) { return __builtin_amdgcn_sqrtf(a); }
};

template <class R> __device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }
template <class R> __device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }

// Limiter switch of VLIMIT (M3:1721): minmod when T = 1.27*U5/(kb*rho) >= 3,
// i.e. when sign(0.5, T-3) is +0.5.
__device__ __forceinline__ bool use_minmod(double u5, double rho) {
  return !signbit(tswitch(u5, rho) - 3.0);
}
template <class R> __device__ __forceinline__ bool use_minmod_r(double u5, double rho) {
  if constexpr (Ar<R>::strict) {
    return use_minmod(u5, rho);
  } else {
    float t = Ar<float>::div(1.27f * (float)u5, (float)KB * (float)rho);
    return !signbit(t - 3.0f);
  }
}

// VLIMIT slope (M3:1716-1724) with only the selected candidate evaluated:
// van Albada  g  = max(0,(2d+d- + eps)/(d+^2 + d-^2 + eps)) (q+ - q-)/(dx0 + dx-/2 + dx+/2)
// minmod      gm = minmod(2d+/(dx0+dx+), 2d-/(dx0+dx-)).
// The reference blends g*(0.5-s) + gm*(0.5+s) with s = +-0.5, which equals
// the selected candidate for every finite input.  Both candidates are two
// divisions; the operands of the two are selected first, so the code has
// no branch (the loads of many slopes can be issued together) and each
// candidate keeps its own operation order.
template <class R>
__device__ __forceinline__ R vslope(double qm, double q0, double qp, R dxm, R dx0, R dxp, bool mm) {
  if constexpr (Ar<R>::strict) {
    const double delp = qp - q0;
    const double delm = q0 - qm;
    const double tp = 2.0 * delp;
    const double n1 = mm ? tp : tp * delm + EPSL;
    const double d1 = mm ? dx0 + dxp : delp * delp + delm * delm + EPSL;
    const double r1 = n1 / d1;
    const double n2 = mm ? 2.0 * delm : dmax(0.0, r1) * (qp - qm);
    const double d2 = mm ? dx0 + dxm : dx0 + 0.5 * dxm + 0.5 * dxp;
    const double r2 = n2 / d2;
    const double gm = (0.5 + copysign(0.5, r1 * r2)) * copysign(1.0, r1) * dmin(fabs(r1), fabs(r2));
    return mm ? gm : r2;
  } else {
    using A = Ar<float>;
    const float delp = (float)(qp - q0);
    const float delm = (float)(q0 - qm);
    const float eps = (float)EPSL;
    const float tp = 2.0f * delp;
    const float n1 = mm ? tp : tp * delm + eps;
    const float d1 = mm ? dx0 + dxp : delp * delp + delm * delm + eps;
    const float r1 = A::div(n1, d1);
    const float n2 = mm ? 2.0f * delm : rmax(0.0f, r1) * (float)(qp - qm);
    const float d2 = mm ? dx0 + dxm : dx0 + 0.5f * dxm + 0.5f * dxp;
    const float r2 = A::div(n2, d2);
    const float m = rmin(fabsf(r1), fabsf(r2));
    // minmod keeps the smaller slope when the two have the same sign; the
    // signs are compared directly because the FP32 product of two slopes of
    // a tiny field (a species near its 1e-20 floor: r ~ 1e-22) underflows
    // to zero and would zero the slope where the FP64 product does not
    const bool same = (r1 > 0.0f && r2 > 0.0f) || (r1 < 0.0f && r2 < 0.0f);
    const float gm = same ? copysignf(m, r1) : 0.0f;
    return mm ? gm : r2;
  }
}

// RIEMAN (M3:1474-1551) in R.
template <class R>
__device__ __forceinline__ void rieman_r(R D1, R V1, R P1, R B1sq, R D2, R V2, R P2, R B2sq, R gam, R &Pa,
                                         R &Va) {
  if constexpr (Ar<R>::strict) {
    rieman(D1, V1, P1, B1sq, D2, V2, P2, B2sq, gam, Pa, Va);
  } else {
    using A = Ar<float>;
    float gammi1 = gam - 1.0f, gammi2 = gam - 2.0f, gammi3 = gam - 3.0f, gampl1 = gam + 1.0f, gampl3 = gam + 3.0f;
    float B1 = 0.5f * B1sq, B2 = 0.5f * B2sq;
    P1 = P1 + B1;
    P2 = P2 + B2;
    float GP1 = gammi1 * P1, GP2 = gammi1 * P2;
    float GP1B1 = gammi3 * P1 - gammi2 * B1 * 2.0f;
    float GP2B2 = gammi3 * P2 - gammi2 * B2 * 2.0f;
    float pa = (P1 + P2) * 0.5f, q1 = 0.0f, q2 = 0.0f;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      float pp1 = gampl3 * pa + GP1B1;
      float s1 = pp1 + A::sqrt_(pp1 * pp1 - 8.0f * (pa - P1) * (gampl1 * pa + GP1));
      q1 = A::sqrt_(s1 * D1) * 0.5f;
      float pp2 = gampl3 * pa + GP2B2;
      float s2 = pp2 + A::sqrt_(pp2 * pp2 - 8.0f * (pa - P2) * (gampl1 * pa + GP2));
      q2 = A::sqrt_(s2 * D2) * 0.5f;
      float old = pa;
      pa = A::div(q2 * P1 + q1 * P2 + q2 * q1 * (V1 - V2), q2 + q1);
      if (pa <= 0.0f) pa = 0.1f * old;
    }
    Va = A::div(q1 * V1 + q2 * V2 + P1 - P2, q2 + q1);
    Pa = pa;
  }
}

// CHAREQ (M3:2101-2111) in R.
template <class R>
__device__ __forceinline__ void chareq_r(R vyp, R Byp, R rl, R vym, R Bym, R rr, R &Bay, R &Vay) {
  if constexpr (Ar<R>::strict) {
    chareq(vyp, Byp, rl, vym, Bym, rr, Bay, Vay);
  } else {
    using A = Ar<float>;
    float sp = A::sqrt_(rl), sm = A::sqrt_(rr);
    float inv = __builtin_amdgcn_rcpf(sp + sm);
    Vay = (sp * vyp + sm * vym + Bym - Byp) * inv;
    Bay = (sp * sm * (vym - vyp) + sp * Bym + sm * Byp) * inv;
  }
}

} // namespace mcf
