// mcf_device.h -- FP64 point routines of the MCF3DSG* step for gfx950.
//
// Each routine restates the reference routine named beside it
// (M3 = 3DMHD-CfCA/MCF3DSGincfca3.f90) with the reference's evaluation
// order; the translation unit is built with -ffp-contract=off so no
// multiply-add is fused and IEEE division / square root are used, which
// keeps every cell bit-identical to the reference wherever no
// transcendental function is involved.
#pragma once
#include <hip/hip_runtime.h>

namespace mcf {

constexpr double KB = 8.63359;      // M3:11
constexpr double KCOND = 1.6384e-2; // M3:11
constexpr double EPSL = 1.0e-10;    // VLIMIT eps (M3:1697)

// dmax1 / dmin1 as flang lowers them (compare + select)
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

// (0.5-sign(0.5,-Q))*Q + (0.5+sign(0.5,-Q))*U : positivity fallback (M3:1389)
__device__ __forceinline__ double posfix(double Q, double U) {
  return (0.5 - copysign(0.5, -Q)) * Q + (0.5 + copysign(0.5, -Q)) * U;
}

// The two slope candidates of VLIMIT (M3:1716-1724) at one position:
// van Albada g and minmod gm.  The limiter switch then takes
// g*(0.5-s) + gm*(0.5+s) with s = sign(0.5, T-3).
struct Slope2 {
  double g, gm;
};
__device__ __forceinline__ Slope2 vlim_parts(double qm, double q0, double qp, double dxm, double dx0,
                                             double dxp) {
  double delp = qp - q0;
  double delm = q0 - qm;
  double flmt = (2.0 * delp * delm + EPSL) / (delp * delp + delm * delm + EPSL);
  flmt = dmax(0.0, flmt);
  Slope2 r;
  r.g = flmt * (qp - qm) / (dx0 + 0.5 * dxm + 0.5 * dxp);
  delp = 2.0 * delp / (dx0 + dxp);
  delm = 2.0 * delm / (dx0 + dxm);
  r.gm = (0.5 + copysign(0.5, delp * delm)) * copysign(1.0, delp) * dmin(fabs(delp), fabs(delm));
  return r;
}
__device__ __forceinline__ double vlim_pick(Slope2 s, double T) {
  return s.g * (0.5 - copysign(0.5, T - 3.0)) + s.gm * (0.5 + copysign(0.5, T - 3.0));
}
// T = 1.27*U5/(kb*rho): the limiter switch variable (M3:1721)
__device__ __forceinline__ double tswitch(double u5, double rho) { return 1.27 * u5 / (KB * rho); }

// RIEMAN (M3:1474-1551): four unrolled two-shock iterations with the
// tangential magnetic pressure added to P.
__device__ __forceinline__ void rieman(double D1, double V1, double P1, double B1sq, double D2, double V2,
                                       double P2, double B2sq, double gam, double &Pa, double &Va) {
  double gammi1 = gam - 1.0, gammi2 = gam - 2.0, gammi3 = gam - 3.0, gampl1 = gam + 1.0, gampl3 = gam + 3.0;
  double B1 = 0.5 * B1sq, B2 = 0.5 * B2sq;
  P1 = P1 + B1;
  P2 = P2 + B2;
  double GP1 = gammi1 * P1, GP2 = gammi1 * P2;
  double GP1B1 = gammi3 * P1 - gammi2 * B1 * 2.0;
  double GP2B2 = gammi3 * P2 - gammi2 * B2 * 2.0;
  double pa = (P1 + P2) * 0.5, q1 = 0.0, q2 = 0.0;
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    double pp1 = gampl3 * pa + GP1B1;
    double s1 = pp1 + sqrt(pp1 * pp1 - 8.0 * (pa - P1) * (gampl1 * pa + GP1));
    q1 = sqrt(s1 * D1) * 0.5;
    double pp2 = gampl3 * pa + GP2B2;
    double s2 = pp2 + sqrt(pp2 * pp2 - 8.0 * (pa - P2) * (gampl1 * pa + GP2));
    q2 = sqrt(s2 * D2) * 0.5;
    double old = pa;
    pa = q2 * P1 + q1 * P2 + q2 * q1 * (V1 - V2);
    pa = pa / (q2 + q1);
    if (pa <= 0.0) pa = 0.1 * old;
  }
  double va = q1 * V1 + q2 * V2 + P1 - P2;
  Va = va / (q2 + q1);
  Pa = pa;
}

// CHAREQ (M3:2101-2111)
__device__ __forceinline__ void chareq(double vyp, double Byp, double rl, double vym, double Bym, double rr,
                                       double &Bay, double &Vay) {
  double sp = sqrt(rl), sm = sqrt(rr);
  Vay = (sp * vyp + sm * vym + Bym - Byp) / (sp + sm);
  Bay = (sp * sm * (vym - vyp) + sp * Bym + sm * Byp) / (sp + sm);
}

// Koyama-Inutsuka net cooling (1DMHDKIcool/1DMHDwCC3KIcool.f90:1165-1171)
// with N = rho/1.27 (3DMHD-feedback/MCF3DSGincfca3-feedback.f90:3586).
__device__ __forceinline__ double fcool_ki(double rho, double T) {
  double N = rho / 1.27;
  double Laml = 1.0e7 * exp(-1.184e5 / (T * 1.0e3 + 1.0e3));
  Laml = dmin(Laml, 5.0e3);
  double Lamc = 1.4e-2 * sqrt(T * 1.0e3) * exp(-9.2e1 / (T * 1.0e3));
  return (N * N * (Laml + Lamc) - N) * 3.94656e1;
}

// IMC (M3:2898-2921): backward-Euler P + gammi1*Lambda(P)*dt = P_old by
// bisection in log10 P on [1e-10, 1e10], at most 30 steps.
__device__ __forceinline__ double imc(double P, double n, double dt, double rho, double gammi1) {
  double Pu = 1.0e10, Pd = 1.0e-10, Pm = 1.0e1, Pmold = -1.0;
  double nkbi = 1.0 / (KB * n);
  for (int it = 0; it < 30; ++it) {
    double T = Pm * nkbi;
    double CooL = fcool_ki(rho, T);
    double fev = Pm + CooL * dt * gammi1 - P;
    double iud = copysign(1.0, fev);
    Pd = dmax(-iud * Pm, Pd);
    Pu = dmin(2.0 * Pm / (iud + 1.0 + 1.0e-10), Pu);
    Pm = pow(1.0e1, 0.5 * (log10(Pu) + log10(Pd)));
    if (fabs(Pm - Pmold) < 1.0e-5) break;
    Pmold = Pm;
  }
  return Pm;
}

} // namespace mcf
