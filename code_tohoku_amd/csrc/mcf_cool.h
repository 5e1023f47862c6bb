// mcf_cool.h -- the backward-Euler KI cooling solve with the learned
// surrogate: the MLP (mcf_mlp.h) predicts T_end, two Newton steps on the
// backward-Euler equation refine it, and the residual of that equation
// decides per cell whether the result stands or the reference's IMC
// bisection (3DMHD-CfCA/MCF3DSGincfca3.f90:2898-2921) runs instead.
#pragma once
#include "mcf_device.h"
#include "mcf_mlp.h"

namespace mcf {

constexpr double SUR_RESID_TOL = 1.0e-4; // |P + gammi1*Lambda*dt - P_old| / P_old

// Lambda(N, T) of fcool_ki and dLambda/dT (T in units of 10^3 K).
__device__ __forceinline__ double fcool_ki_d(double rho, double T, double &dLdT) {
  double N = rho / 1.27;
  double TK = T * 1.0e3;
  double a = TK + 1.0e3;
  double Laml = 1.0e7 * exp(-1.184e5 / a);
  double dLaml = Laml * 1.184e5 * 1.0e3 / (a * a);
  if (Laml > 5.0e3) { Laml = 5.0e3; dLaml = 0.0; }
  double Lamc = 1.4e-2 * sqrt(TK) * exp(-9.2e1 / TK);
  double dLamc = Lamc * (0.5 / T + 9.2e1 / (TK * T));
  dLdT = N * N * (dLaml + dLamc) * 3.94656e1;
  return (N * N * (Laml + Lamc) - N) * 3.94656e1;
}

// Surrogate features of one backward-Euler solve P + gammi1*Lambda(P)*dt = P0.
__device__ __forceinline__ void sur_features(double P0, double n, double dt, double rho, double gammi1, float &u0,
                                             float &u1, float &u2) {
  double T0 = P0 / (KB * n);
  double L0 = fcool_ki(rho, T0);
  double r = dt * gammi1 * fabs(L0) / P0;
  u0 = ((float)log10(rho / 1.27) - MLP_LN_C) / MLP_LN_S;
  u1 = ((float)log10(T0 * 1.0e3) - MLP_LT_C) / MLP_LT_S;
  u2 = (float)log10(r) / MLP_LR_S;
  if (!(u2 > -10.0f)) u2 = -10.0f; // Lambda = 0: no cooling
}

// Refine the surrogate's prediction y (scaled log10 T_end/T0) by two Newton
// steps and test the residual.  Returns true and P_end when accepted.
__device__ __forceinline__ bool sur_refine(double P0, double n, double dt, double rho, double gammi1, float y,
                                           double &Pend) {
  const double kbn = KB * n;
  const double T0 = P0 / kbn;
  const double a = gammi1 * dt / kbn;
  double T = T0 * exp10((double)y * (double)MLP_Y_S);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    double dL;
    double L = fcool_ki_d(rho, T, dL);
    double g = T + a * L - T0;
    double gp = 1.0 + a * dL;
    double Tn = T - g / gp;
    if (!(gp > 0.0) || !(Tn > 0.0)) Tn = 0.5 * T;
    T = Tn;
  }
  double res = fabs(T + a * fcool_ki(rho, T) - T0) / T0;
  Pend = kbn * T;
  return res <= SUR_RESID_TOL && Pend >= 1.0e-10 && Pend <= 1.0e10;
}

} // namespace mcf
