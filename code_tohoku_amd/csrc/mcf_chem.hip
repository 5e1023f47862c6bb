// mcf_chem.hip -- SHIELD column densities and the chemistry network of
// SOURCE (M3 = 3DMHD-CfCA/MCF3DSGincfca3.f90) for gfx950, FP64.
#include "mcf_chem.h"
#include "mcf_kernels.h"

namespace mcf {

// Sum over the 64 lanes by butterfly: every lane ends with the same value
// (each pairwise sum a+b is formed as b+a by the partner lane, and FP
// addition is commutative).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}

// Inclusive prefix sums over the lanes (Hillis-Steele): upward (lane 0
// first) and downward (lane 63 first).
__device__ __forceinline__ double wave_scan_up(double v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double u = __shfl_up(v, o, 64);
    if (lane >= o) v = v + u;
  }
  return v;
}
__device__ __forceinline__ double wave_scan_down(double v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double u = __shfl_down(v, o, 64);
    if (lane + o < 64) v = v + u;
  }
  return v;
}

// SHIELD (M3:2924-3036), one wave per (j,k) row of x.  Per x block b of the
// reproduced MPI split (cx cells, half width dxh = dx(first cell)/2) the
// reference accumulates, for Ntot (from ndtot), NH2, NnC, NCO and tCII (from
// the CII optical-depth rate temp), a running column t with each cell's
// value taken half before and half after it is stored (t += dxh*v; store;
// t += dxh*v), from the left end of the block and from the right one; the
// stored columns then get the block totals of the blocks beyond, nearest
// first.  Here each 64-cell chunk of a block is one prefix scan over the
// lanes (the sums differ from the sequential ones only in rounding):
//   pass 1: temp of every cell (from the previous tCII) into S.temp, and the
//           block totals into LDS;
//   pass 2: the columns from the left, plus the totals of the blocks left of b;
//   pass 3: the columns from the right, plus the totals of the blocks right of b.
// Lane l always handles cell 64*chunk + l of a block, so every lane reads
// back only the temp values it wrote itself; the right-hand columns come
// from downward scans over the chunks taken last to first.
template <int DIR>
__device__ __forceinline__ void shield_pass(const ShieldArgs &S, const Geo &g, int j, int k, int lane,
                                            const double (*tot)[5]) {
  const int nsx = S.nsx, cx = S.cx, nch = (cx + 63) / 64;
  for (int b = 0; b < nsx; ++b) {
    const int ib = b * cx + 1;
    const double dxh = 0.5 * S.gdx0[ib + 1];
    double carry[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int ch = 0; ch < nch; ++ch) {
      const int m = 64 * (DIR == 0 ? ch : nch - 1 - ch) + lane;
      const bool on = m < cx;
      const int c = ix3(g, ib + (on ? m : 0), j, k);
      double a[5];
      a[0] = on ? dxh * (S.sp[1][c] + S.sp[0][c] + 2.0 * S.sp[2][c] + S.sp[3][c] + S.sp[4][c]) : 0.0;
      a[1] = on ? dxh * S.sp[2][c] : 0.0;
      a[2] = on ? dxh * S.sp[5][c] : 0.0;
      a[3] = on ? dxh * S.sp[6][c] : 0.0;
      a[4] = on ? dxh * S.temp[c] : 0.0;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        double incl, ex, chunk;
        if (DIR == 0) {
          incl = wave_scan_up(a[q] + a[q], lane);
          ex = __shfl_up(incl, 1, 64);
          if (lane == 0) ex = 0.0;
          chunk = __shfl(incl, 63, 64);
        } else {
          incl = wave_scan_down(a[q] + a[q], lane);
          ex = __shfl_down(incl, 1, 64);
          if (lane == 63) ex = 0.0;
          chunk = __shfl(incl, 0, 64);
        }
        double v = (carry[q] + ex) + a[q];
        if (DIR == 0) {
          for (int L = 1; b - L >= 0; ++L) v = v + tot[b - L][q];
        } else {
          for (int L = 1; b + L < nsx; ++L) v = v + tot[b + L][q];
        }
        if (on) S.col[2 * q + DIR][c] = v;
        carry[q] = carry[q] + chunk;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_shield(ShieldArgs S, Geo g) {
  __shared__ double tots[4][SHIELD_MAXB][5];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int R = g.n[1] * g.n[2];
  const int r = blockIdx.x * 4 + w;
  if (r >= R) return; // whole waves only, and no barrier below
  const int j = 1 + r % g.n[1], k = 1 + r / g.n[1];
  const int nsx = S.nsx, cx = S.cx;
  double(*tot)[5] = tots[w];
  // pass 1
  for (int b = 0; b < nsx; ++b) {
    const int ib = b * cx + 1;
    const double dxh = 0.5 * S.gdx0[ib + 1];
    double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int off = 0; off < cx; off += 64) {
      const int m = off + lane;
      if (m < cx) {
        const int c = ix3(g, ib + m, j, k);
        const double ndH = S.sp[0][c], ndp = S.sp[1][c], ndH2 = S.sp[2][c], ndHe = S.sp[3][c], ndHep = S.sp[4][c];
        const double ndCp = S.sp[7][c];
        const double b21 = ch_fesc(S.col[8][c]) + ch_fesc(S.col[9][c]);
        const double T = S.p[c] / (KB * (ndp + ndH + ndH2 + ndHe + ndHep));
        double n1, n2;
        ch_levc2(T, exp(0.07 * log(T)), b21, ndH, ndH2, ndp + ndHep + ndCp, ndCp, n1, n2);
        const double tmp = 1.15563e1 * (n1 - n2) * (1.0 / CH_DV);
        S.temp[c] = tmp;
        const double a0 = dxh * (ndp + ndH + 2.0 * ndH2 + ndHe + ndHep);
        const double a1 = dxh * ndH2, a2 = dxh * S.sp[5][c], a3 = dxh * S.sp[6][c], a4 = dxh * tmp;
        t[0] = t[0] + (a0 + a0);
        t[1] = t[1] + (a1 + a1);
        t[2] = t[2] + (a2 + a2);
        t[3] = t[3] + (a3 + a3);
        t[4] = t[4] + (a4 + a4);
      }
    }
    // every lane holds the same sum and writes it, so each lane later reads
    // a value it wrote itself
#pragma unroll
    for (int q = 0; q < 5; ++q) tot[b][q] = wave_sum(t[q]);
  }
  shield_pass<0>(S, g, j, k, lane, tot);
  shield_pass<1>(S, g, j, k, lane, tot);
}

// Chemistry network (M3:2576-2774) over the interior cells
__global__ __launch_bounds__(256) void k_chem(ChemArgs A, Geo g, double dt) {
  const long n = (long)g.n[0] * g.n[1] * g.n[2];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = 1 + (int)(t % g.n[0]);
  const int j = 1 + (int)((t / g.n[0]) % g.n[1]);
  const int k = 1 + (int)(t / ((long)g.n[0] * g.n[1]));
  const int c = ix3(g, i, j, k);
  ch_network(A.sp, A.rho, A.cols, c, A.p[c], dt, A.spmin);
}

} // namespace mcf
