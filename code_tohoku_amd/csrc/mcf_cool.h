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

// Surrogate features of one backward-Euler solve P + gammi1*Lambda(P)*dt = P0
// (FP32 is enough for network inputs).
__device__ __forceinline__ void sur_features(double P0, double n, double dt, double rho, double gammi1, float &u0,
                                             float &u1, float &u2) {
  const float T0 = (float)(P0 / (KB * n));
  const float N = (float)rho * (1.0f / 1.27f);
  const float TK = T0 * 1.0e3f;
  const float laml = fminf(1.0e7f * __expf(-1.184e5f / (TK + 1.0e3f)), 5.0e3f);
  const float lamc = 1.4e-2f * sqrtf(TK) * __expf(-9.2e1f / TK);
  const float L0 = (N * N * (laml + lamc) - N) * 3.94656e1f;
  const float r = (float)dt * (float)gammi1 * fabsf(L0) / (float)P0;
  u0 = (log10f(N) - MLP_LN_C) / MLP_LN_S;
  u1 = (log10f(TK) - MLP_LT_C) / MLP_LT_S;
  u2 = log10f(r) / MLP_LR_S;
  if (!(u2 > -10.0f)) u2 = -10.0f; // Lambda = 0: no cooling
}

// Refine the surrogate's prediction y (scaled log10 T_end/T0) by Newton
// steps on g(u) = T + a*Lambda(T) - T0 in u = ln T (at most SUR_NEWTON,
// each limited to a factor e, stopping once |du| < 1e-12), then test the
// relative residual of the backward-Euler equation.  Returns true and
// P_end when accepted.
constexpr int SUR_NEWTON = 4;
__device__ __forceinline__ bool sur_refine(double P0, double n, double dt, double rho, double gammi1, float y,
                                           double &Pend) {
  const double kbn = KB * n;
  const double T0 = P0 / kbn;
  const double a = gammi1 * dt / kbn;
  double u = log(T0) + (double)y * (double)MLP_Y_S * 2.302585092994046;
  for (int it = 0; it < SUR_NEWTON; ++it) {
    const double T = exp(u);
    double dL;
    const double L = fcool_ki_d(rho, T, dL);
    const double g = T + a * L - T0;
    const double gp = T * (1.0 + a * dL);
    double du = (gp > 0.0) ? -g / gp : 0.0;
    du = fmin(fmax(du, -1.0), 1.0);
    u += du;
    if (fabs(du) < 1.0e-12) break;
  }
  const double T = exp(u);
  const double res = fabs(T + a * fcool_ki(rho, T) - T0) / T0;
  Pend = kbn * T;
  return res <= SUR_RESID_TOL && Pend >= 1.0e-10 && Pend <= 1.0e10;
}

} // namespace mcf
