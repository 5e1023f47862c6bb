// mcf_chem.h -- the reference's chemistry-network cooling, CII level
// populations, escape probabilities, rate coefficients and network step
// (M3 = 3DMHD-CfCA/MCF3DSGincfca3.f90), FP64, for gfx950.
//
// Species arrays are in restart order: 0 H, 1 p, 2 H2, 3 He, 4 He+, 5 C,
// 6 CO, 7 C+.  Column densities (SHIELD) are ten arrays col[2q + dir] with
// q = 0 Ntot, 1 NH2, 2 NnC, 3 NCO, 4 tCII and dir 0 = seen from the left x
// end, 1 = from the right one.  Every expression keeps the reference's
// evaluation order (the translation units are built with
// -ffp-contract=off); exp, log, pow and sqrt are the device library's, and
// powers T**a of the temperature are exp(a*log(T)) with log(T) formed once.
#pragma once
#include "mcf_device.h"

namespace mcf {

// chmvar (M3:27-28)
constexpr double CH_G0 = 1.0, CH_XO = 3.2e-4, CH_DV = 2.0, CH_TGR = 5.0e-3;
enum { CQ_NTOT = 0, CQ_NH2, CQ_NNC, CQ_NCO, CQ_TCII };

struct ChemCols {
  double *c[10]; // c[2q + dir]
};

// fesc (M3:2883-2888)
__device__ __forceinline__ double ch_fesc(double tau) {
  return (0.5 + copysign(0.5, 0.1 - 1.0e-16 - tau)) * (0.5 - 0.585 * tau + 0.4563 * (tau * tau)) +
         (0.5 + copysign(0.5, 11.9025 - (tau - 3.55) * (tau - 3.55))) * (1.0 - exp(-2.34 * tau)) / (4.68 * tau + 1.0e-8) +
         (0.5 + copysign(0.5, tau - 7.0)) / (4.0 * tau * sqrt(log(dmax(tau * 0.56419, 4.0))) + 1.0e-8);
}

// LEVC2 (M3:2890-2895); t007 = T**0.07
__device__ __forceinline__ void ch_levc2(double T, double t007, double b21, double ndH, double ndH2, double nde,
                                         double ndCp, double &n1, double &n2) {
  const double c21 = 8.854e-8 * nde / sqrt(T) + 9.399e-10 * t007 * (ndH + ndH2);
  const double c12 = 2.0 * c21 * exp(-0.092 / T);
  const double A21 = 2.4e-6 * b21;
  n2 = ndCp * c12 / (c12 + c21 + A21);
  n1 = ndCp * (c21 + A21) / (c12 + c21 + A21);
}

// The temperature-independent parts of Fcool (M3:2825-2881) at one cell,
// formed once so that IMC's bisection re-evaluates only what depends on T.
// Each stored prefix is the left-associated leading factor of its term, so
// the products are the reference's.
struct ChemFc {
  double ndH, ndH2, ndCp, nde, ATN1, b21;
  double laml_h, laml_p; // ndH*nde*1.44049e9, ndp*nde*1.48e6
  double lamo, feso;     // (max(ndtot*xo-ndCO,0)/xo)*(ndH+0.5ndH2)*1.23916e1, fesO1+fesO2
  double lamd;           // nde*ndtot*6.06236
  double nsum;           // ndH+ndp+ndH2+ndHe+ndHep (ncr)
  double tco1, tco2;     // 1.33194e1*NCO(1|2)
  double lcor, lcoh, lcoh2; // ndCO*1.91505e10, ndH*ndCO*7.96086e4, ndH2*ndCO*3.60834e4
  double gampe;          // ndtot*2.56526e3*G0*ATN1
  double gamcr, gampd;
};

__device__ __forceinline__ ChemFc ch_fc_setup(const double *const *sp, const ChemCols &C, int c) {
  ChemFc f;
  const double ndH = sp[0][c], ndp = sp[1][c], ndH2 = sp[2][c], ndHe = sp[3][c], ndHep = sp[4][c];
  const double ndCO = sp[6][c], ndCp = sp[7][c];
  const double nde = ndp + ndHep + ndCp;
  const double ndtot = ndp + ndH + 2.0 * ndH2 + ndHe + ndHep;
  const double Nt1 = C.c[0][c], Nt2 = C.c[1][c];
  const double Av1 = 1.63542e-3 * Nt1, x1 = 6.1714e3 * C.c[2][c];
  const double Av2 = 1.63542e-3 * Nt2, x2 = 6.1714e3 * C.c[3][c];
  f.ATN1 = (exp(-2.5 * Av1) + exp(-2.5 * Av2)) * 0.5;
  const double ATN2 = (exp(-3.77358 * Av1) + exp(-3.77358 * Av2)) * 0.5;
  f.ndH = ndH; f.ndH2 = ndH2; f.ndCp = ndCp; f.nde = nde;
  f.laml_h = ndH * nde * 1.44049e9;
  f.laml_p = ndp * nde * 1.48e6;
  f.b21 = ch_fesc(C.c[8][c]) + ch_fesc(C.c[9][c]);
  const double tO1 = 3.40057e-6 * Nt1 / CH_XO, tO2 = 3.40057e-6 * Nt2 / CH_XO;
  f.feso = ch_fesc(tO1) + ch_fesc(tO2);
  f.lamo = (dmax(ndtot * CH_XO - ndCO, 0.0) / CH_XO) * (ndH + 0.5 * ndH2) * 1.23916e1;
  f.lamd = nde * ndtot * 6.06236;
  f.nsum = ndH + ndp + ndH2 + ndHe + ndHep;
  f.tco1 = 1.33194e1 * C.c[6][c];
  f.tco2 = 1.33194e1 * C.c[7][c];
  f.lcor = ndCO * 1.91505e10;
  f.lcoh = ndH * ndCO * 7.96086e4;
  f.lcoh2 = ndH2 * ndCO * 3.60834e4;
  f.gampe = ndtot * 2.56526e3 * CH_G0 * f.ATN1;
  f.gamcr = (ndH + ndHe + ndH2) * 1.89435;
  const double SHLD1 = 0.965 / ((1.0 + x1 / CH_DV) * (1.0 + x1 / CH_DV)) + 0.035 / sqrt(1.0 + x1) * exp(-8.5e-4 * sqrt(1.0 + x1));
  const double SHLD2 = 0.965 / ((1.0 + x2 / CH_DV) * (1.0 + x2 / CH_DV)) + 0.035 / sqrt(1.0 + x2) * exp(-8.5e-4 * sqrt(1.0 + x2));
  f.gampd = ndH2 * 4.16362e4 * CH_G0 * ATN2 * (SHLD1 + SHLD2) * 0.5;
  return f;
}

// Fcool (M3:2825-2881) at temperature T (units of 10^3 K).  Its nine powers
// x**a are formed as exp(a*log(x)) from one log(T) and one log(pha), which
// differs from pow(x, a) by about (|a log x| + 2) units of the last place
// (|a log x| stays below 20 over 10 K - 10^5 K) and costs a third as much.
__device__ __forceinline__ double ch_fcool(const ChemFc &f, double T) {
  const double pha = CH_G0 * sqrt(1.0e3 * T) * f.ATN1 / f.nde;
  const double lT = log(T), lpha = log(pha);
  double Laml = f.laml_h * exp(-1.184e2 / T);
  Laml = Laml + f.laml_p * exp(-8.0e1 / T);
  double n1, n2;
  ch_levc2(T, exp(0.07 * lT), f.b21, f.ndH, f.ndH2, f.nde, f.ndCp, n1, n2);
  const double Lamc = 6.0157e7 * n2 * f.b21;
  double Lamo = f.lamo * exp(0.4 * lT) * exp(-0.228 / T);
  Lamo = Lamo * f.feso;
  const double Lamd = f.lamd * exp(0.94 * lT) * exp(0.462628 / exp(6.8e-2 * lT) * lpha);
  const double ncr = 3.3e6 * exp(0.75 * lT) / f.nsum;
  const double tau1 = f.tco1 / (T * CH_DV), tau2 = f.tco2 / (T * CH_DV);
  const double ct1 = tau1 * sqrt(6.283 * log(2.13 + (tau1 * 0.36788) * (tau1 * 0.36788)));
  const double ct2 = tau2 * sqrt(6.283 * log(2.13 + (tau2 * 0.36788) * (tau2 * 0.36788)));
  const double ym1 = log(1.0 + ct1 / (1.0 + 1.0e1 * ncr)), ym2 = log(1.0 + ct2 / (1.0 + 1.0e1 * ncr));
  const double fes = (2.0 + ym1 + 0.6 * (ym1 * ym1)) / (1.0 + ct1 + ncr + 1.5 * sqrt(ncr)) +
                     (2.0 + ym2 + 0.6 * (ym2 * ym2)) / (1.0 + ct2 + ncr + 1.5 * sqrt(ncr));
  const double LCOr = f.lcor * (T * T) * fes;
  const double e308 = exp(-3.08 / T);
  const double LCOH = f.lcoh * sqrt(T) * exp(-exp(3.43 * (0.69314718055994531 - lT))) * e308;
  const double LCOH2 = f.lcoh2 * T * exp(-exp(0.333 * (5.7493929859082531 - lT))) * e308;
  const double Gampe = f.gampe * (7.382e-3 * exp(0.7 * lT) / (1.0 + 2.0e-4 * pha) + 4.9e-2 / (1.0 + 4.0e-3 * exp(0.73 * lpha)));
  return Laml + Lamc + Lamo + Lamd + LCOr + LCOH + LCOH2 - Gampe - f.gamcr - f.gampd;
}

// IMC (M3:2898-2921) with the chemistry Fcool
__device__ __forceinline__ double ch_imc(const ChemFc &f, double P, double n, double dt, double gammi1) {
  double Pu = 1.0e10, Pd = 1.0e-10, Pm = 1.0e1, Pmold = -1.0;
  const double nkbi = 1.0 / (KB * n);
  for (int it = 0; it < 30; ++it) {
    const double T = Pm * nkbi;
    const double CooL = ch_fcool(f, T);
    const double fev = Pm + CooL * dt * gammi1 - P;
    const double iud = copysign(1.0, fev);
    Pd = dmax(-iud * Pm, Pd);
    Pu = dmin(2.0 * Pm / (iud + 1.0 + 1.0e-10), Pu);
    Pm = pow(1.0e1, 0.5 * (log10(Pu) + log10(Pd)));
    if (fabs(Pm - Pmold) < 1.0e-5) break;
    Pmold = Pm;
  }
  return Pm;
}

// Omexp (M3:3136-3141): 1 - exp(-eps), its Taylor series below 1e-4
__device__ __forceinline__ double ch_omexp(double eps) {
  return (0.5 + copysign(0.5, eps - 1.0e-4)) * (1.0 - exp(-eps)) +
         (0.5 - copysign(0.5, eps - 1.0e-4)) *
             (eps - 0.5 * (eps * eps) + 0.16666666666666667 * (eps * eps * eps) - 4.16666666666666667e-2 * (eps * eps * eps * eps));
}

// One cell's chemistry network over dt (SOURCE, M3:2576-2774) with the
// rate coefficients of RATES (M3:3040-3133).  Writes the eight species
// (with their floors) and the density formed from them.
__device__ __forceinline__ void ch_network(double *const *sp, double *rho, const ChemCols &C, int c, double u5,
                                           double dt, const double *spmin) {
  double ndpold = sp[1][c], ndHold = sp[0][c], ndH2old = sp[2][c], ndHeold = sp[3][c];
  double ndHepold = sp[4][c], ndCold = sp[5][c], ndCpold = sp[7][c], ndCOold = sp[6][c];
  double ndH, ndp, ndH2, ndHe, ndHep, ndC, ndCp, ndCO;
  const double ndtot = ndpold + ndHold + 2.0 * ndH2old + ndHeold + ndHepold;
  double nde = ndpold + ndHepold + ndCpold;
  const double T = u5 / KB / (ndpold + ndHold + ndH2old + ndHeold + ndHepold);
  // RATES
  const double NT1 = C.c[0][c], NT2 = C.c[1][c], NH21 = C.c[2][c], NH22 = C.c[3][c];
  const double NnC1 = C.c[4][c], NnC2 = C.c[5][c], NC1 = C.c[6][c], NC2 = C.c[7][c];
  const double Av1 = 1.63542e-3 * NT1, x1 = 6.1714e3 * NH21;
  const double Av2 = 1.63542e-3 * NT2, x2 = 6.1714e3 * NH22;
  const double ATN2 = (exp(-3.77358 * Av1) + exp(-3.77358 * Av2)) * 0.5;
  const double ATN3 = (exp(-2.3585 * Av1) + exp(-2.3585 * Av2)) * 0.5;
  double ATN4 = dmin(1.0, pow(3.0857e1 * NH21, -3.0e-2) * exp(-1.54285e-3 * NH21)) +
                dmin(1.0, pow(3.0857e1 * NH22, -3.0e-2) * exp(-1.54285e-3 * NH22));
  ATN4 = ATN4 * 0.5;
  const double p06 = (double)0.6f; // a default-real (REAL*4) exponent in the reference
  double ATN5 = (0.5 - copysign(0.5, NC1 - 1.40995e-3)) * exp(-1.0 * pow(1.02857e3 * NC1, p06)) +
                (0.5 + copysign(0.5, NC1 - 1.40995e-3)) *
                    dmin(pow(3.75439e3 * NC1 + 1.0e-100, -0.75), pow(7.3469e1 * NC1 + 1.0e-100, -1.3));
  ATN5 = ATN5 + (0.5 - copysign(0.5, NC2 - 1.40995e-3)) * exp(-1.0 * pow(1.02857e3 * NC2, p06)) +
         (0.5 + copysign(0.5, NC2 - 1.40995e-3)) *
             dmin(pow(3.75439e3 * NC2 + 1.0e-100, -0.75), pow(7.3469e1 * NC2 + 1.0e-100, -1.3));
  ATN5 = ATN5 * 0.5;
  const double zeta = 9.4671e-4;
  const double lT = log(T); // T**a below is exp(a*log(T)), as in Fcool
  double kHrec = (0.5 + copysign(0.5, 15.78 - T)) * 0.45 * log(1.578e2 / T) +
                 (0.5 - copysign(0.5, 15.78 - T)) * 0.4 * sqrt(1.578e2 / T);
  kHrec = 2.05572e1 * exp(-0.5 * lT) * kHrec;
  const double kHerec = 6.37623e1 * exp(-0.672 * lT);
  const double kH2 = 3.4569e-3 * sqrt(T) / (1.0 + 1.26491 * sqrt(T + CH_TGR) + 2.0 * T + 8.0 * (T * T));
  const double SHLD1 = 0.965 / ((1.0 + x1 / CH_DV) * (1.0 + x1 / CH_DV)) + 0.035 / sqrt(1.0 + x1) * exp(-8.5e-4 * sqrt(1.0 + x1));
  const double SHLD2 = 0.965 / ((1.0 + x2 / CH_DV) * (1.0 + x2 / CH_DV)) + 0.035 / sqrt(1.0 + x2) * exp(-8.5e-4 * sqrt(1.0 + x2));
  const double kH2ph = 1.04138e3 * CH_G0 * ATN2 * (SHLD1 + SHLD2) * 0.5;
  const double kH2dH = 1.07294e5 * exp(-4.39e1 / T);
  const double kH2de = 1.133e6 * exp(0.35 * lT) * exp(-1.02e2 / T);
  const double kCO = 1.57785e-2 / (1.0 + CH_G0 * ATN3 / (ndH2old * CH_XO));
  const double kCOph = 3.155e3 * CH_G0 * ATN3 * ATN4 * ATN5;
  const double SHLC1 = exp(-2.6 * Av1 - 3.0857e1 * NnC1 - 4.33743413e-4 * NH21) / (1.0 + 4.33743413e-4 * NH21);
  const double SHLC2 = exp(-2.6 * Av2 - 3.0857e1 * NnC2 - 4.33743413e-4 * NH22) / (1.0 + 4.33743413e-4 * NH22);
  const double kCi = 3.97 * zeta + 6.62697e3 * CH_G0 * (SHLC1 + SHLC2) * 0.5;
  const double kCrec = 2.69267e1 * exp(-0.62 * lT) * (1.0 + 1.0e-4 * sqrt(nde) / (T * T));
  const double kCOde = 1.4e4 * exp(-0.5 * lT) * exp(-1.14e2 / T);
  const double kCOdH = 1.91858e3 * exp(-0.5 * lT) * exp(-7.77e1 / T);
  const double kHie = 5.76e4 * sqrt(T) * exp(-1.58e2 / T);
  const double kHeie = 2.76e4 * exp(0.43 * lT) * exp(-2.85e2 / T);
  const double kCie = 1.74e5 * exp(0.4 * lT) * exp(-1.31e2 / T);
  const double kHiH = 9.79 * sqrt(T) * exp(-1.58e2 / T);
  const double kHeiH = 4.57 * exp(0.43 * lT) * exp(-2.85e2 / T);
  const double kCiH = 3.01e1 * exp(0.4 * lT) * exp(-1.31e2 / T);
  const double kH2dHep = 1.1676 * exp(-0.035 / T);
  const double kCOdHep = 5.0491e4;
  double temp1, temp2, temp3, omeps, eps;
  // H recombination & ionization by CR
  temp1 = kHrec * nde;
  temp2 = exp(-dt * (zeta + temp1));
  temp3 = 1.0 / (zeta + temp1);
  omeps = ch_omexp(dt * (zeta + temp1));
  ndH = (temp1 * ndHold + zeta * ndHold * temp2 + temp1 * ndpold * omeps) * temp3;
  ndp = (zeta * ndpold + temp1 * ndpold * temp2 + zeta * ndHold * omeps) * temp3;
  ndHold = ndH; ndpold = ndp; nde = ndp + ndHepold + ndCpold;
  // He recombination & ionization by CR
  temp1 = kHerec * nde;
  temp2 = exp(-dt * (zeta + temp1));
  temp3 = 1.0 / (zeta + temp1);
  omeps = ch_omexp(dt * (zeta + temp1));
  ndHe = (temp1 * ndHeold + zeta * ndHeold * temp2 + temp1 * ndHepold * omeps) * temp3;
  ndHep = (zeta * ndHepold + temp1 * ndHepold * temp2 + zeta * ndHeold * omeps) * temp3;
  ndHeold = ndHe; ndHepold = ndHep; nde = ndp + ndHep + ndCpold;
  // H2 formation & dissociation by UV & electron collision
  temp1 = kH2ph + kH2de * nde;
  temp2 = temp1 + 2.0 * kH2 * ndtot;
  temp3 = exp(-dt * temp2);
  omeps = ch_omexp(dt * temp2);
  ndH = (ndHold + 2.0 * ndH2old * omeps) * temp1 + 2.0 * kH2 * ndHold * ndtot * temp3;
  ndH = ndH / temp2;
  ndH2 = (2.0 * ndH2old + ndHold * omeps) * kH2 * ndtot + temp1 * ndH2old * temp3;
  ndH2 = ndH2 / temp2;
  ndHold = ndH; ndH2old = ndH2;
  // H2 dissociation by H collision
  temp1 = ndHold + 2.0 * ndH2old;
  temp2 = exp(-dt * temp1 * kH2dH);
  temp3 = 1.0 / (ndHold + 2.0 * ndH2old * temp2);
  ndH = ndHold * temp1 * temp3;
  ndH2 = ndH2old * temp1 * temp2 * temp3;
  ndHold = ndH; ndH2old = ndH2;
  // CO formation
  temp1 = exp(-dt * kCO * ndtot);
  eps = dt * kCO * ndtot;
  omeps = ch_omexp(eps);
  ndCO = ndCOold + omeps * ndCpold;
  ndCp = temp1 * ndCpold;
  ndCOold = ndCO; ndCpold = ndCp; nde = ndp + ndHep + ndCp;
  // CO dissociation by UV & electron collision & H collision
  temp1 = exp(-dt * (kCOph + kCOde * nde + kCOdH * ndHold));
  eps = dt * (kCOph + kCOde * nde + kCOdH * ndHold);
  omeps = ch_omexp(eps);
  ndC = ndCold + omeps * ndCOold;
  ndCO = temp1 * ndCOold;
  ndCold = ndC; ndCOold = ndCO;
  // C recombination & ionization by CR & UV
  temp1 = kCrec * nde;
  temp2 = exp(-dt * (kCi + temp1));
  temp3 = 1.0 / (kCi + temp1);
  omeps = ch_omexp(dt * (kCi + temp1));
  ndC = (temp1 * ndCold + kCi * ndCold * temp2 + temp1 * ndCpold * omeps) * temp3;
  ndCp = (kCi * ndCpold + temp1 * ndCpold * temp2 + kCi * ndCold * omeps) * temp3;
  ndCold = ndC; ndCpold = ndCp; nde = ndp + ndHep + ndCp;
  // H, He, C ionization by e collision
  ndH = ndHold * exp(-dt * kHie * nde);
  ndHe = ndHeold * exp(-dt * kHeie * nde);
  ndC = ndCold * exp(-dt * kCie * nde);
  omeps = ch_omexp(dt * kHie * nde); ndp = omeps * ndHold + ndpold;
  omeps = ch_omexp(dt * kHeie * nde); ndHep = omeps * ndHeold + ndHepold;
  omeps = ch_omexp(dt * kCie * nde); ndCp = omeps * ndCold + ndCpold;
  ndHold = ndH; ndHeold = ndHe; ndCold = ndC;
  ndpold = ndp; ndHepold = ndHep; ndCpold = ndCp;
  nde = ndp + ndHep + ndCp;
  // H, He, C ionization by H or H2 or p collision
  temp1 = 1.0 / (1.0 + kHiH * dt * ndHold);
  ndH = ndHold * temp1;
  ndp = ndpold + temp1 * kHiH * dt * (ndHold * ndHold);
  temp1 = ndH + ndp + ndH2old;
  ndHe = ndHeold * exp(-dt * kHeiH * temp1);
  ndC = ndCold * exp(-dt * kCiH * temp1);
  omeps = ch_omexp(dt * kHeiH * temp1); ndHep = omeps * ndHeold + ndHepold;
  omeps = ch_omexp(dt * kCiH * temp1); ndCp = omeps * ndCold + ndCpold;
  ndHold = ndH; ndHeold = ndHe; ndCold = ndC;
  ndpold = ndp; ndHepold = ndHep; ndCpold = ndCp;
  // H2 dissociation by He+ recombination
  temp1 = ndHepold - ndH2old;
  if (temp1 >= 0.0) {
    temp3 = exp(-dt * kH2dHep * temp1);
    temp2 = 1.0 / (ndHepold - ndH2old * temp3);
    ndHep = ndHepold * temp1 * temp2;
    ndH2 = ndH2old * temp1 * temp2 * temp3;
    temp3 = ndHepold * (1.0 - temp1 * temp2);
  } else {
    temp3 = exp(dt * kH2dHep * temp1);
    temp2 = 1.0 / (ndHepold * temp3 - ndH2old);
    ndHep = ndHepold * temp1 * temp2 * temp3;
    ndH2 = ndH2old * temp1 * temp2;
    temp3 = ndHepold * (1.0 - temp1 * temp2 * temp3);
  }
  ndp = ndpold + temp3;
  ndH = ndHold + temp3;
  ndHe = ndHeold + temp3;
  ndHepold = ndHep; ndHeold = ndHe;
  // CO dissociation by He+ recombination
  temp1 = ndHepold - ndCOold;
  if (temp1 >= 0.0) {
    temp3 = exp(-dt * kCOdHep * temp1);
    temp2 = 1.0 / (ndHepold - ndCOold * temp3);
    ndHep = ndHepold * temp1 * temp2;
    ndCO = ndCOold * temp1 * temp2 * temp3;
    temp3 = ndHepold * (1.0 - temp1 * temp2);
  } else {
    temp3 = exp(dt * kCOdHep * temp1);
    temp2 = 1.0 / (ndHepold * temp3 - ndCOold);
    ndHep = ndHepold * temp1 * temp2 * temp3;
    ndCO = ndCOold * temp1 * temp2;
    temp3 = ndHepold * (1.0 - temp1 * temp2 * temp3);
  }
  ndCp = ndCpold + temp3;
  ndHe = ndHeold + temp3;
  ndH = dmax(spmin[0], ndH);
  ndp = dmax(spmin[1], ndp);
  ndH2 = dmax(spmin[2], ndH2);
  ndHe = dmax(spmin[3], ndHe);
  ndHep = dmax(spmin[4], ndHep);
  ndC = dmax(spmin[5], ndC);
  ndCO = dmax(spmin[6], ndCO);
  ndCp = dmax(spmin[7], ndCp);
  sp[0][c] = ndH; sp[1][c] = ndp; sp[2][c] = ndH2; sp[3][c] = ndHe;
  sp[4][c] = ndHep; sp[5][c] = ndC; sp[6][c] = ndCO; sp[7][c] = ndCp;
  rho[c] = 1.0 * ndp + 1.0 * ndH + 2.0 * ndH2 + 4.0 * ndHe + 4.0 * ndHep;
}

} // namespace mcf
