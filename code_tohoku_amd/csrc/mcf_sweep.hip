// mcf_sweep.hip -- one MHD sweep (M3:1264-1324) as three fused stage
// kernels: RHS (M3:1330-1467), MOC (M3:1939-2098), REMAP + primitive
// recovery (M3:1556-1686, 1291-1321).
//
// RHS and MOC run in three phases per tile, with LDS between them:
//   1. every cell of the tile and its two halo cells along the sweep axis
//      computes its limited slopes once (RHS also its left/right
//      reconstructed states);
//   2. every interface of the tile solves its Riemann / characteristic
//      problem from the two cells' LDS entries;
//   3. every cell is updated from its two interfaces.
// REMAP evaluates the upwind slope per interface (one cell per interface)
// and then updates the cells.  No per-interface array goes to HBM except
// the interface velocity and Lagrangian position that later stages read.
//
// Tiles.  x sweep: 8 lines of 64 consecutive x cells (one wave per line,
// lanes along x); wave 8 takes the halo cells and the leftmost interface of
// the 8 lines.  y/z sweeps: 64 consecutive x cells (lanes) at 8 positions
// along the sweep axis (one wave per position); wave 8 takes the position
// before the tile (cell and interface), wave 9 the position after it
// (cell only).  All loads stay coalesced along x.
//
// Along the periodic y and z axes RHS and MOC instead run as one column
// march (k_rhsmoc_col): one thread per x column carries the stencils and
// the Lagrangian state of RHS in registers into MOC, so only MOC's results
// go to HBM.  The tiled RHS and MOC kernels remain for the x sweep and for
// axes shorter than 8 cells.
//
// Outputs go to the other buffer of each written field (the host swaps the
// pointers), so a tile never reads cells another tile has already updated.
//
// The kernels are generic over the compute type R (mcf_prec.h):
// R = double reproduces the reference bit for bit; R = float is the mixed
// path (FP32 reconstruction, Riemann and characteristic states; state
// updates, positions and primitive recovery in FP64).
#include "mcf_kernels.h"
#include "mcf_prec.h"

namespace mcf {

namespace {

__device__ __forceinline__ int ix3s(const Geo &g, int i, int j, int k) {
  return (i + 1) + (j + 1) * g.st[1] + (k + 1) * g.st[2];
}

__device__ __forceinline__ double gam_of_s(const double *const *sp, int c) {
  double s4 = sp[0][c] + sp[1][c] + sp[3][c] + sp[4][c];
  double g = 3.0 * s4 + 5.0 * sp[2][c];
  return (5.0 * s4 + 7.0 * sp[2][c]) / g;
}
__device__ __forceinline__ double gammi1_of_s(const double *const *sp, int c) {
  double s4 = sp[0][c] + sp[1][c] + sp[3][c] + sp[4][c];
  double g = 3.0 * s4 + 5.0 * sp[2][c];
  return (2.0 * s4 + 2.0 * sp[2][c]) / g;
}
template <class R> __device__ __forceinline__ R gam_of_r(const double *const *sp, int c) {
  if constexpr (Ar<R>::strict) {
    return gam_of_s(sp, c);
  } else {
    float s4 = (float)(sp[0][c] + sp[1][c] + sp[3][c] + sp[4][c]);
    float h2 = (float)sp[2][c];
    return Ar<float>::div(5.0f * s4 + 7.0f * h2, 3.0f * s4 + 5.0f * h2);
  }
}

constexpr int XL = 66; // x sweep: cell slots per line (64 + 2 halo)
constexpr int XF = 65; // x sweep: interface slots per line

// Thread roles in a tile.  The phase-1 cell of a thread is also the left
// cell of its interface (phase 2) and, for the main waves, the cell it
// updates (phase 3).
struct Roles {
  int i, j, k, q;    // phase-1 cell and its position along the sweep axis
  int cslot;         // LDS slot of that cell
  int rslot_c;       // LDS slot of the cell to its right
  int fslot;         // LDS slot of the interface right of the cell
  int lslot, rslot;  // interface slots left/right of an updated cell
  bool c_ok;         // phase 1: cell exists (position 0..N+1)
  bool f_ok;         // phase 2: interface exists (position 0..N)
  bool own;          // interface is written to HBM by this tile
  bool u_ok;         // phase 3: cell 1..N updated by this thread
};

template <int D>
__device__ __forceinline__ Roles roles(const Geo &g) {
  Roles r;
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  r.lslot = r.rslot = 0;
  r.u_ok = false;
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  if (D == 0) {
    const int s0 = 1 + 64 * bx;
    const int jr = 1 + 8 * by;
    r.k = 1 + bz;
    if (w < 8) {
      r.i = s0 + l; r.j = jr + w;
      r.cslot = w * XL + l + 1;
      r.fslot = w * XF + l + 1;
      r.lslot = w * XF + l; r.rslot = w * XF + l + 1;
      bool row = r.j <= ny;
      r.c_ok = row && r.i <= nx + 1;
      r.f_ok = row && r.i <= nx;
      r.u_ok = r.f_ok;
      r.own = true;
    } else if (l < 8) {
      r.i = s0 - 1; r.j = jr + l;
      r.cslot = l * XL;
      r.fslot = l * XF;
      r.c_ok = r.f_ok = r.j <= ny;
      r.own = (s0 == 1);
    } else {
      const int ll = l & 7;
      r.i = s0 + 64; r.j = jr + ll;
      r.cslot = ll * XL + 65;
      r.fslot = 0;
      r.c_ok = l < 16 && r.j <= ny && r.i <= nx + 1;
      r.f_ok = false;
      r.own = false;
    }
    r.rslot_c = r.cslot + 1;
    r.q = r.i;
  } else {
    const int i0 = 1 + 64 * bx;
    const int s0 = 1 + 8 * by;
    const int tr = 1 + bz;
    const int n = (D == 1) ? ny : nz;
    const int ntr = (D == 1) ? nz : ny;
    const bool col = (i0 + l <= nx) && (tr <= ntr);
    int p;
    if (w < 8) {
      p = s0 + w;
      r.cslot = (w + 1) * 64 + l;
      r.fslot = (w + 1) * 64 + l;
      r.lslot = w * 64 + l; r.rslot = (w + 1) * 64 + l;
      r.c_ok = col && p <= n + 1;
      r.f_ok = col && p <= n;
      r.u_ok = r.f_ok;
      r.own = true;
    } else if (w == 8) {
      p = s0 - 1;
      r.cslot = l;
      r.fslot = l;
      r.c_ok = r.f_ok = col;
      r.own = (s0 == 1);
    } else {
      p = s0 + 8;
      r.cslot = 9 * 64 + l;
      r.fslot = 0;
      r.c_ok = col && p <= n + 1;
      r.f_ok = false;
      r.own = false;
    }
    r.rslot_c = r.cslot + 64;
    r.i = i0 + l;
    if (D == 1) { r.j = p; r.k = tr; } else { r.k = p; r.j = tr; }
    r.q = p;
  }
  return r;
}

template <int D> constexpr int ncell_slots() { return D == 0 ? 8 * XL : 10 * 64; }
template <int D> constexpr int nface_slots() { return D == 0 ? 8 * XF : 9 * 64; }

// Reconstructed states q +- depend*g with the positivity fallback (M3:1389).
template <class R> __device__ __forceinline__ R left_state_pos(double q, R depend, R gr) {
  return (R)posfix(q + (double)(depend * gr), q);
}
template <class R> __device__ __forceinline__ R right_state_pos(double q, R depend, R gr) {
  return (R)posfix(q - (double)(depend * gr), q);
}

} // namespace

// ---------------------------------------------------------------------------
// RHS
// ---------------------------------------------------------------------------
// Occupancy targets (waves per SIMD) of RHS and MOC: 5 lets two tiles share
// a CU; the FP64 versions hold more live values and stay at 4, where they do
// not spill.
template <class R> constexpr int stage_waves() { return Ar<R>::strict ? 4 : 5; }

// Every global load of the three phases is issued before the first barrier
// (the cell's stencil, its species and those of the right neighbour, the
// remaining velocity and field components), so one memory latency covers
// the whole kernel.
template <int D, class R>
__global__ __launch_bounds__(D == 0 ? 576 : 640, stage_waves<R>()) void k_rhs(SweepIO A, Geo g, double dt) {
  constexpr int NC = ncell_slots<D>(), NF = nface_slots<D>();
  constexpr int VN = D + 1, VT1 = (D + 1) % 3 + 1, VT2 = (D + 2) % 3 + 1;
  constexpr int BT1 = (D + 1) % 3, BT2 = (D + 2) % 3;
  __shared__ R sQ[8][NC];       // QL(rho, v_n, p, |B_t|^2) and QR(...) per cell
  __shared__ R sPa[NF], sVa[NF];
  const Roles t = roles<D>(g);
  const int s = g.st[D];
  const int q = t.q;
  const int c = ix3s(g, t.i, t.j, t.k);
  const double *dxe = A.dxe, *xe = A.xe;
  const double *U1 = A.U[0], *UVN = A.U[VN], *U5 = A.U[4];
  const double *B1 = A.bcc[BT1], *B2 = A.bcc[BT2];
  const double *const *sp = A.sp;
  double rho = 0, vn = 0, p = 0, b1 = 0, b2 = 0;
  double r_m = 0, r_p = 0, v_m = 0, v_p = 0, p_m = 0, p_p = 0, a_m = 0, a_p = 0, c_m = 0, c_p = 0;
  double dxm_d = 1, dx0_d = 1, dxp_d = 1;
  double spc[8], spn[5];
#pragma unroll
  for (int f = 0; f < 8; ++f) spc[f] = 0.0;
#pragma unroll
  for (int f = 0; f < 5; ++f) spn[f] = 0.0;
  double vt1 = 0, vt2 = 0, bnc = 0, xq0 = 0, xq1 = 0;
  if (t.c_ok) {
    rho = U1[c]; vn = UVN[c]; p = U5[c]; b1 = B1[c]; b2 = B2[c];
    r_m = U1[c - s]; r_p = U1[c + s]; v_m = UVN[c - s]; v_p = UVN[c + s];
    p_m = U5[c - s]; p_p = U5[c + s]; a_m = B1[c - s]; a_p = B1[c + s]; c_m = B2[c - s]; c_p = B2[c + s];
    dxm_d = dxe[q]; dx0_d = dxe[q + 1]; dxp_d = dxe[q + 2];
#pragma unroll
    for (int f = 0; f < 5; ++f) spc[f] = sp[f][c];
  }
  if (t.f_ok) {
#pragma unroll
    for (int f = 0; f < 5; ++f) spn[f] = sp[f][c + s];
    xq1 = xe[q + 1];
  }
  if (t.u_ok) {
#pragma unroll
    for (int f = 5; f < 8; ++f) spc[f] = sp[f][c];
    vt1 = A.U[VT1][c]; vt2 = A.U[VT2][c]; bnc = A.bcc[D][c];
    xq0 = xe[q];
  }
  if (t.c_ok) {
    const R dxm = (R)dxm_d, dx0 = (R)dx0_d, dxp = (R)dxp_d;
    const bool mm = use_minmod_r<R>(p, rho);
    const R g1 = vslope<R>(r_m, rho, r_p, dxm, dx0, dxp, mm);
    const R gv = vslope<R>(v_m, vn, v_p, dxm, dx0, dxp, mm);
    const R g5 = vslope<R>(p_m, p, p_p, dxm, dx0, dxp, mm);
    const R gA = vslope<R>(a_m, b1, a_p, dxm, dx0, dxp, mm);
    const R gB = vslope<R>(c_m, b2, c_p, dxm, dx0, dxp, mm);
    // fast magnetosonic speed and the traced half-width (M3:1380-1388)
    R gamma;
    if constexpr (Ar<R>::strict) {
      double s4 = spc[0] + spc[1] + spc[3] + spc[4];
      double gg = 3.0 * s4 + 5.0 * spc[2];
      gamma = (5.0 * s4 + 7.0 * spc[2]) / gg;
    } else {
      float s4 = (float)(spc[0] + spc[1] + spc[3] + spc[4]);
      float h2 = (float)spc[2];
      gamma = Ar<float>::div(5.0f * s4 + 7.0f * h2, 3.0f * s4 + 5.0f * h2);
    }
    const R cm = Ar<R>::sqrt_(Ar<R>::div(gamma * (R)p + (R)(b1 * b1) + (R)(b2 * b2), (R)rho));
    R depend = (R)0.5 * (dx0 - cm * (R)dt) * (R)A.facdep;
    depend = rmax((R)0.0, depend);
    sQ[0][t.cslot] = left_state_pos<R>(rho, depend, g1);
    sQ[1][t.cslot] = (R)(vn + (double)(depend * gv));
    sQ[2][t.cslot] = left_state_pos<R>(p, depend, g5);
    R a1 = (R)(b1 + (double)(depend * gA)), a2 = (R)(b2 + (double)(depend * gB));
    sQ[3][t.cslot] = a1 * a1 + a2 * a2;
    sQ[4][t.cslot] = right_state_pos<R>(rho, depend, g1);
    sQ[5][t.cslot] = (R)(vn - (double)(depend * gv));
    sQ[6][t.cslot] = right_state_pos<R>(p, depend, g5);
    a1 = (R)(b1 - (double)(depend * gA));
    a2 = (R)(b2 - (double)(depend * gB));
    sQ[7][t.cslot] = a1 * a1 + a2 * a2;
  }
  __syncthreads();
  if (t.f_ok) {
    // interface gamma from the mean species (M3:1440-1446)
    R gm;
    if constexpr (Ar<R>::strict) {
      double ndHm = 0.5 * (spc[0] + spn[0]);
      double ndpm = 0.5 * (spc[1] + spn[1]);
      double ndHem = 0.5 * (spc[3] + spn[3]);
      double ndHepm = 0.5 * (spc[4] + spn[4]);
      double ndH2m = 0.5 * (spc[2] + spn[2]);
      gm = (5.0 * (ndHm + ndpm + ndHem + ndHepm) + 7.0 * ndH2m) / (3.0 * (ndHm + ndpm + ndHem + ndHepm) + 5.0 * ndH2m);
    } else {
      float s4 = (float)(spc[0] + spn[0] + spc[1] + spn[1] + spc[3] + spn[3] + spc[4] + spn[4]);
      float h2 = (float)(spc[2] + spn[2]);
      gm = Ar<float>::div(5.0f * s4 + 7.0f * h2, 3.0f * s4 + 5.0f * h2);
    }
    const int a = t.cslot, b = t.rslot_c;
    R Pa, Va;
    rieman_r<R>(sQ[0][a], sQ[1][a], sQ[2][a], sQ[3][a], sQ[4][b], sQ[5][b], sQ[6][b], sQ[7][b], gm, Pa, Va);
    sPa[t.fslot] = Pa;
    sVa[t.fslot] = Va;
    if (t.own) {
      A.va[c] = (double)Va;
      A.xlag[c] = xq1 + (double)Va * dt;
    }
  }
  __syncthreads();
  if (!t.u_ok) return;
  // Lagrangian update of the cell (M3:1448-1466)
  double vv[3], bb[3];
  vv[D] = vn; vv[(D + 1) % 3] = vt1; vv[(D + 2) % 3] = vt2;
  bb[D] = bnc; bb[BT1] = b1; bb[BT2] = b2;
  const double vx = vv[0], vy = vv[1], vz = vv[2];
  const double bx = bb[0], by = bb[1], bz = bb[2];
  const double dxq = dx0_d;
  const double dm = dxq * rho;
  const double pa = (double)sPa[t.rslot], pam = (double)sPa[t.lslot];
  const double va = (double)sVa[t.rslot], vam = (double)sVa[t.lslot];
  const double xl = xq1 + va * dt, xlm = xq0 + vam * dt;
  const double dxl = xl - xlm;
  if constexpr (Ar<R>::strict) {
    double s4 = spc[0] + spc[1] + spc[3] + spc[4];
    double gg = 3.0 * s4 + 5.0 * spc[2];
    double gi = (2.0 * s4 + 2.0 * spc[2]) / gg;
    double e = p / (rho * gi) + 0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz) / rho;
    A.Uo[0][c] = dm / dxl;
    A.Uo[VN][c] = vn - dt / dm * (pa - pam);
    A.u5e[c] = e - dt / dm * (pa * va - pam * vam);
    A.blgo[0][c] = b1 * dxq / dxl;
    A.blgo[1][c] = b2 * dxq / dxl;
#pragma unroll
    for (int f = 0; f < 8; ++f) A.spo[f][c] = dxq * spc[f] / dxl;
  } else {
    double s4 = spc[0] + spc[1] + spc[3] + spc[4];
    double irho = 1.0 / rho;
    double e = p * (3.0 * s4 + 5.0 * spc[2]) / (2.0 * s4 + 2.0 * spc[2]) * irho +
               0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz) * irho;
    double idl = 1.0 / dxl, dtdm = dt / dm, cmp = dxq * idl;
    A.Uo[0][c] = dm * idl;
    A.Uo[VN][c] = vn - dtdm * (pa - pam);
    A.u5e[c] = e - dtdm * (pa * va - pam * vam);
    A.blgo[0][c] = b1 * cmp;
    A.blgo[1][c] = b2 * cmp;
#pragma unroll
    for (int f = 0; f < 8; ++f) A.spo[f][c] = spc[f] * cmp;
  }
  A.dxlagM[c] = dxl;
}

// ---------------------------------------------------------------------------
// RHS from register values (the column march of k_rhsmoc_col)
// ---------------------------------------------------------------------------
// Cell states of RHS phase 1 (M3:1380-1406) from the cell's stencil:
// q[0..7] = QL(rho, v_n, p, |B_t|^2), QR(...).
template <class R>
__device__ __forceinline__ void rhs_states(double rho, double vn, double p, double b1, double b2, double r_m,
                                           double r_p, double v_m, double v_p, double p_m, double p_p, double a_m,
                                           double a_p, double c_m, double c_p, const double *spc, double dxm_d,
                                           double dx0_d, double dxp_d, double dt, double facdep, R *q) {
  const R dxm = (R)dxm_d, dx0 = (R)dx0_d, dxp = (R)dxp_d;
  const bool mm = use_minmod_r<R>(p, rho);
  const R g1 = vslope<R>(r_m, rho, r_p, dxm, dx0, dxp, mm);
  const R gv = vslope<R>(v_m, vn, v_p, dxm, dx0, dxp, mm);
  const R g5 = vslope<R>(p_m, p, p_p, dxm, dx0, dxp, mm);
  const R gA = vslope<R>(a_m, b1, a_p, dxm, dx0, dxp, mm);
  const R gB = vslope<R>(c_m, b2, c_p, dxm, dx0, dxp, mm);
  R gamma;
  if constexpr (Ar<R>::strict) {
    double s4 = spc[0] + spc[1] + spc[3] + spc[4];
    double gg = 3.0 * s4 + 5.0 * spc[2];
    gamma = (5.0 * s4 + 7.0 * spc[2]) / gg;
  } else {
    float s4 = (float)(spc[0] + spc[1] + spc[3] + spc[4]);
    float h2 = (float)spc[2];
    gamma = Ar<float>::div(5.0f * s4 + 7.0f * h2, 3.0f * s4 + 5.0f * h2);
  }
  const R cm = Ar<R>::sqrt_(Ar<R>::div(gamma * (R)p + (R)(b1 * b1) + (R)(b2 * b2), (R)rho));
  R depend = (R)0.5 * (dx0 - cm * (R)dt) * (R)facdep;
  depend = rmax((R)0.0, depend);
  q[0] = left_state_pos<R>(rho, depend, g1);
  q[1] = (R)(vn + (double)(depend * gv));
  q[2] = left_state_pos<R>(p, depend, g5);
  R a1 = (R)(b1 + (double)(depend * gA)), a2 = (R)(b2 + (double)(depend * gB));
  q[3] = a1 * a1 + a2 * a2;
  q[4] = right_state_pos<R>(rho, depend, g1);
  q[5] = (R)(vn - (double)(depend * gv));
  q[6] = right_state_pos<R>(p, depend, g5);
  a1 = (R)(b1 - (double)(depend * gA));
  a2 = (R)(b2 - (double)(depend * gB));
  q[7] = a1 * a1 + a2 * a2;
}

// RIEMAN at the interface between a left cell (species sl, states ql) and a
// right cell (species sr, states qr) with the interface gamma of the mean
// species (M3:1440-1446).
template <class R>
__device__ __forceinline__ void rhs_iface(const double *sl, const double *sr, const R *ql, const R *qr, R &Pa,
                                          R &Va) {
  R gm;
  if constexpr (Ar<R>::strict) {
    double ndHm = 0.5 * (sl[0] + sr[0]);
    double ndpm = 0.5 * (sl[1] + sr[1]);
    double ndHem = 0.5 * (sl[3] + sr[3]);
    double ndHepm = 0.5 * (sl[4] + sr[4]);
    double ndH2m = 0.5 * (sl[2] + sr[2]);
    gm = (5.0 * (ndHm + ndpm + ndHem + ndHepm) + 7.0 * ndH2m) / (3.0 * (ndHm + ndpm + ndHem + ndHepm) + 5.0 * ndH2m);
  } else {
    float s4 = (float)(sl[0] + sr[0] + sl[1] + sr[1] + sl[3] + sr[3] + sl[4] + sr[4]);
    float h2 = (float)(sl[2] + sr[2]);
    gm = Ar<float>::div(5.0f * s4 + 7.0f * h2, 3.0f * s4 + 5.0f * h2);
  }
  rieman_r<R>(ql[0], ql[1], ql[2], ql[3], qr[4], qr[5], qr[6], qr[7], gm, Pa, Va);
}

// ---------------------------------------------------------------------------
// MOC
// ---------------------------------------------------------------------------
template <class R> struct MocV {
  R Bay, Vay, Baz, Vaz;
};

// Slopes of rho, v_t1, v_t2, Blg1, Blg2 at cell c with the limiter choice mm.
template <class R>
__device__ __forceinline__ void moc_slopes(const SweepIO &A, int c, int s, bool mm, R *gs) {
  const double *Q[5] = {A.U[0], A.U[A.VT1], A.U[A.VT2], A.blg[0], A.blg[1]};
  const R dxm = (R)A.dxlagM[c - s], dx0 = (R)A.dxlagM[c], dxp = (R)A.dxlagM[c + s];
#pragma unroll
  for (int v = 0; v < 5; ++v) gs[v] = vslope<R>(Q[v][c - s], Q[v][c], Q[v][c + s], dxm, dx0, dxp, mm);
}

// One characteristic interface (M3:2000-2051) from register values: v0 and
// v1 hold (rho, v_t1, v_t2, Blg1, Blg2, dx_lag) of the left and right cell,
// gi and gp their slopes, Bnap_d the normal field at the Lagrangian
// interface position.
template <class R>
__device__ __forceinline__ MocV<R> moc_iface_r(const double *v0, const double *v1, const R *gi, const R *gp,
                                               double Bnap_d, double facdep, double dt) {
  const double dl1 = v0[5], dl2 = v1[5];
  R deni;
  if constexpr (Ar<R>::strict) {
    deni = sqrt((dl1 + dl2) / (dl1 * v1[0] + dl2 * v0[0]));
  } else {
    deni = Ar<float>::sqrt_(Ar<float>::div((float)(dl1 + dl2), (float)(dl1 * v1[0] + dl2 * v0[0])));
  }
  const R Bnap = (R)Bnap_d;
  const R fac = (R)facdep;
  const R rdt = (R)dt;
  R depend1, depend2, ca;
  R QL11, QL12, QL13, QR11, QR12, QR13, QL21, QL22, QR21, QR22;
  if (Bnap > (R)0.0) {
    ca = fabs(Bnap * deni) * (R)0.5 * rdt;
    ca = rmin((R)dl1, ca);
    depend1 = ((R)0.5 * (R)dl1 - ca) * fac;
    ca = fabs(Bnap * deni) * (R)0.5 * rdt;
    ca = rmin((R)dl2, ca);
    depend2 = ((R)0.5 * (R)dl2 - ca) * fac;
    QL11 = (R)(v0[1] + (double)(depend1 * gi[1]));
    QR11 = (R)(v1[1] - (double)(depend2 * gp[1]));
    QL12 = (R)(v0[3] + (double)(depend1 * gi[3]));
    QR12 = (R)(v1[3] - (double)(depend2 * gp[3]));
    QL13 = left_state_pos<R>(v0[0], depend1, gi[0]);
    QR13 = right_state_pos<R>(v1[0], depend2, gp[0]);
    QL21 = (R)(v0[2] + (double)(depend1 * gi[2]));
    QR21 = (R)(v1[2] - (double)(depend2 * gp[2]));
    QL22 = (R)(v0[4] + (double)(depend1 * gi[4]));
    QR22 = (R)(v1[4] - (double)(depend2 * gp[4]));
  } else {
    ca = fabs(Bnap * deni) * (R)0.5 * rdt;
    ca = rmin((R)dl2, ca);
    depend1 = ((R)0.5 * (R)dl2 - ca) * fac;
    ca = fabs(Bnap * deni) * (R)0.5 * rdt;
    ca = rmin((R)dl1, ca);
    depend2 = ((R)0.5 * (R)dl1 - ca) * fac;
    QL11 = (R)(v1[1] - (double)(depend1 * gp[1]));
    QR11 = (R)(v0[1] + (double)(depend2 * gi[1]));
    QL12 = (R)(v1[3] - (double)(depend1 * gp[3]));
    QR12 = (R)(v0[3] + (double)(depend2 * gi[3]));
    QL13 = right_state_pos<R>(v1[0], depend1, gp[0]);
    QR13 = left_state_pos<R>(v0[0], depend2, gi[0]);
    QL21 = (R)(v1[2] - (double)(depend1 * gp[2]));
    QR21 = (R)(v0[2] + (double)(depend2 * gi[2]));
    QL22 = (R)(v1[4] - (double)(depend1 * gp[4]));
    QR22 = (R)(v0[4] + (double)(depend2 * gi[4]));
  }
  MocV<R> o;
  chareq_r<R>(QL11, QL12, QL13, QR11, QR12, QR13, o.Bay, o.Vay);
  chareq_r<R>(QL21, QL22, QL13, QR21, QR22, QR13, o.Baz, o.Vaz);
  return o;
}

// B_n at the Lagrangian position of interface q (M3:2008): linear
// interpolation of the face values f0, f1, f2 at faces q, q+1, q+2.
__device__ __forceinline__ double moc_bn(double f0, double f1, double f2, double xlag, const double *xe, int q) {
  return f1 + (xlag - xe[q + 1]) * (f2 - f0) / (xe[q + 2] - xe[q]);
}

// The interface between c and cp = c + s at position q, from the slopes gi
// (cell c) and gp (cell cp), with the cell values read from the arrays.
template <class R>
__device__ __forceinline__ MocV<R> moc_iface(const SweepIO &A, int c, int cp, int q, const R *gi, const R *gp,
                                             double dt) {
  const double *UBn = A.U[A.BN];
  const int cpp = cp + (cp - c);
  const double v0[6] = {A.U[0][c], A.U[A.VT1][c], A.U[A.VT2][c], A.blg[0][c], A.blg[1][c], A.dxlagM[c]};
  const double v1[6] = {A.U[0][cp], A.U[A.VT1][cp], A.U[A.VT2][cp], A.blg[0][cp], A.blg[1][cp], A.dxlagM[cp]};
  const double Bnap_d = moc_bn(UBn[c], UBn[cp], UBn[cpp], A.xlag[c], A.xe, q);
  return moc_iface_r<R>(v0, v1, gi, gp, Bnap_d, A.facdep, dt);
}

// The limiter switch of a cell in MOC uses the specific total energy e left
// by RHS for interior cells and the pre-sweep pressure for ghost cells.
// An interface at a block edge of the reproduced MPI run (position a
// multiple of the block length) is evaluated twice: in the left block's
// view the right cell is that block's ghost copy (pressure switch), in the
// right block's view the left cell is.  The ghost-view slopes differ from
// the interior ones only when the two switch values fall on opposite sides
// of 3; they are recomputed for that case.
template <int D, class R>
__global__ __launch_bounds__(D == 0 ? 576 : 640, stage_waves<R>()) void k_moc(SweepIO A, Geo g, double dt) {
  constexpr int NC = ncell_slots<D>(), NF = nface_slots<D>();
  constexpr int NSH = (5 * NC > 8 * NF) ? 5 * NC : 8 * NF;
  __shared__ R sh[NSH];
  R(*sG)[NC] = reinterpret_cast<R(*)[NC]>(sh); // phases 1-2: slopes per cell
  R(*sI)[NF] = reinterpret_cast<R(*)[NF]>(sh); // phases 2-3: left view 0..3, right view 4..7
  const Roles t = roles<D>(g);
  const int s = g.st[D];
  const int n = g.n[D];
  const double *U1 = A.U[0];
  if (t.c_ok) {
    const int q = t.q;
    const int c = ix3s(g, t.i, t.j, t.k);
    const bool inner = q >= 1 && q <= n;
    const bool mm = use_minmod_r<R>(inner ? A.u5e_in[c] : A.U[4][c], U1[c]);
    R gs[5];
    moc_slopes<R>(A, c, s, mm, gs);
#pragma unroll
    for (int v = 0; v < 5; ++v) sG[v][t.cslot] = gs[v];
  }
  __syncthreads();
  MocV<R> L{}, Rv{};
  if (t.f_ok) {
    const int q = t.q;
    const int c = ix3s(g, t.i, t.j, t.k), cp = c + s;
    const int a = t.cslot, b = t.rslot_c;
    R gi[5], gp[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) { gi[v] = sG[v][a]; gp[v] = sG[v][b]; }
    if ((q % A.nblk) != 0) {
      L = moc_iface<R>(A, c, cp, q, gi, gp, dt);
      Rv = L;
    } else {
      // left cell q (interior unless q = 0), right cell q+1 (interior unless q = n)
      const bool mm0p = use_minmod_r<R>(A.U[4][c], U1[c]);
      const bool mm0e = q >= 1 ? use_minmod_r<R>(A.u5e_in[c], U1[c]) : mm0p;
      const bool mm1p = use_minmod_r<R>(A.U[4][cp], U1[cp]);
      const bool mm1e = q + 1 <= n ? use_minmod_r<R>(A.u5e_in[cp], U1[cp]) : mm1p;
      R ga[5];
      // left-block view: cell q interior, cell q+1 its ghost copy
      if (mm1p != mm1e) {
        moc_slopes<R>(A, cp, s, mm1p, ga);
        L = moc_iface<R>(A, c, cp, q, gi, ga, dt);
      } else {
        L = moc_iface<R>(A, c, cp, q, gi, gp, dt);
      }
      // right-block view: cell q its ghost copy, cell q+1 interior
      if (mm0p != mm0e) {
        moc_slopes<R>(A, c, s, mm0p, ga);
        Rv = moc_iface<R>(A, c, cp, q, ga, gp, dt);
      } else if (mm1p != mm1e) {
        Rv = moc_iface<R>(A, c, cp, q, gi, gp, dt);
      } else {
        Rv = L;
      }
    }
  }
  __syncthreads();
  if (t.f_ok) {
    sI[0][t.fslot] = L.Bay; sI[1][t.fslot] = L.Vay; sI[2][t.fslot] = L.Baz; sI[3][t.fslot] = L.Vaz;
    sI[4][t.fslot] = Rv.Bay; sI[5][t.fslot] = Rv.Vay; sI[6][t.fslot] = Rv.Baz; sI[7][t.fslot] = Rv.Vaz;
  }
  __syncthreads();
  if (!t.u_ok) return;
  // cell update (M3:2053-2095): left view of the right interface, right
  // view of the left interface
  const int q = t.q;
  const int c = ix3s(g, t.i, t.j, t.k), cm = c - s, cp = c + s, cpp = cp + s;
  const double *UBn = A.U[A.BN];
  const double *xe = A.xe;
  const double Bnap = moc_bn(UBn[c], UBn[cp], UBn[cpp], A.xlag[c], xe, q);
  const double Bnam = moc_bn(UBn[cm], UBn[c], UBn[cp], A.xlag[cm], xe, q - 1);
  const double Vnap = A.va[c], Vnam = A.va[cm];
  const double dxl = A.dxlagM[c];
  const double rho = U1[c];
  const double BayL = (double)sI[0][t.rslot], VayL = (double)sI[1][t.rslot];
  const double BazL = (double)sI[2][t.rslot], VazL = (double)sI[3][t.rslot];
  const double BayR = (double)sI[4][t.lslot], VayR = (double)sI[5][t.lslot];
  const double BazR = (double)sI[6][t.lslot], VazR = (double)sI[7][t.lslot];
  if constexpr (Ar<R>::strict) {
    A.Uo[A.VN][c] = A.U[A.VN][c] * rho + dt / dxl * 0.5 * (Bnap * Bnap - Bnam * Bnam);
    A.Uo[A.VT1][c] = A.U[A.VT1][c] * rho + dt / dxl * (Bnap * BayL - Bnam * BayR);
    A.Uo[A.VT2][c] = A.U[A.VT2][c] * rho + dt / dxl * (Bnap * BazL - Bnam * BazR);
    A.Uo[4][c] = A.u5e_in[c] * rho + dt / dxl *
                                        ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                         (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
    A.blgo[0][c] = A.blg[0][c] + dt / dxl * (Bnap * VayL - Bnam * VayR);
    A.blgo[1][c] = A.blg[1][c] + dt / dxl * (Bnap * VazL - Bnam * VazR);
  } else {
    const double f = dt / dxl;
    A.Uo[A.VN][c] = A.U[A.VN][c] * rho + f * 0.5 * (Bnap * Bnap - Bnam * Bnam);
    A.Uo[A.VT1][c] = A.U[A.VT1][c] * rho + f * (Bnap * BayL - Bnam * BayR);
    A.Uo[A.VT2][c] = A.U[A.VT2][c] * rho + f * (Bnap * BazL - Bnam * BazR);
    A.Uo[4][c] = A.u5e_in[c] * rho + f * ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                          (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
    A.blgo[0][c] = A.blg[0][c] + f * (Bnap * VayL - Bnam * VayR);
    A.blgo[1][c] = A.blg[1][c] + f * (Bnap * VazL - Bnam * VazR);
  }
}

// ---------------------------------------------------------------------------
// RHS + MOC fused along a periodic axis (y or z)
// ---------------------------------------------------------------------------
// One thread per x column i of transverse row tr marches along the sweep
// axis over positions p0-4 .. p1+5 and produces the MOC results of cells
// p0..p1 (one segment), so the Lagrangian state of RHS never goes to HBM
// (its species are formed again by k_remap_col, which follows).
// Per iteration q: RHS states of cell q, the Riemann interface q-1|q and the
// Lagrangian update of cell k = q-1 (RHS result O(k)); then the MOC slope of
// cell k-1, the characteristic interface k-2|k-1 and the MOC update of cell
// k-2.
//
// Periodicity: every value at a position outside 1..n is the value of the
// periodic image (the ghost fills are copies), so loads and the positions
// used for the Lagrangian update of a cell go through the image position.
// This needs a grid spacing that is the same at every position of the axis
// (checked by the host); the absolute positions xe still enter per cell.
// What does not follow the image is the MOC limiter switch: it uses e from
// RHS for cells 1..n and the pre-sweep pressure for cells outside (ghost
// cells), and each interface at a block edge of the reproduced MPI run
// (k_moc) is evaluated in both blocks' views.
template <class R> struct OR {
  double v[6]; // rho, v_t1, v_t2, Blg1, Blg2, dx_lag (the MOC slope fields)
  double vn, e;
  bool mme, mmp; // limiter switch from e and from the pre-sweep pressure
};

template <int D, class R>
__global__ __launch_bounds__(256) void k_rhsmoc_col(SweepIO A, Geo g, double dt) {
  constexpr int VN = D + 1, VT1 = (D + 1) % 3 + 1, VT2 = (D + 2) % 3 + 1;
  constexpr int BT1 = (D + 1) % 3, BT2 = (D + 2) % 3;
  const int nx = g.n[0];
  const int n = g.n[D], ntr = g.n[D == 1 ? 2 : 1];
  // grid: segments fastest, so the segments of a column (which share halo
  // cells) run together on one XCD
  int bseg, bxc, btr;
  xcd_tile(bseg, bxc, btr);
  const int i = 1 + 64 * bxc + (int)(threadIdx.x & 63);
  const int tr = 1 + 4 * btr + (int)(threadIdx.x >> 6);
  const int p0 = 1 + FUSE_SEG * bseg;
  if (i > nx || tr > ntr || p0 > n) return;
  const int p1 = min(n, p0 + FUSE_SEG - 1);
  const int nblk = A.nblk;
  // image position of p (n >= 8, so one shift suffices for p0-6 .. p1+6)
  auto img = [&](int p) { return p < 1 ? p + n : (p > n ? p - n : p); };
  auto at = [&](int p) { return D == 1 ? ix3s(g, i, p, tr) : ix3s(g, i, tr, p); };
  auto cell = [&](int p) { return at(img(p)); };
  const double *U1 = A.U[0], *UVN = A.U[VN], *U5 = A.U[4];
  const double *B1 = A.bcc[BT1], *B2 = A.bcc[BT2];
  const double *UBn = A.U[5 + D];
  const double *const *sp = A.sp;
  const double *xe = A.xe;
  const double dxu = A.dxe[1];
  // RHS window: positions q-1, q, q+1 and the values of q+2 in flight
  double wr[4], wv[4], wp[4], wa[4], wc[4];
  double spw[3][5];
  R qprev[8];
  R pa_prev = 0, va_prev = 0;
  const int ps = p0 - 2, pe = p1 + 3;
  {
    const int c0 = cell(ps - 2), c1 = cell(ps - 1), c2 = cell(ps), c3 = cell(ps + 1);
    wr[0] = U1[c0]; wr[1] = U1[c1]; wr[2] = U1[c2]; wr[3] = U1[c3];
    wv[0] = UVN[c0]; wv[1] = UVN[c1]; wv[2] = UVN[c2]; wv[3] = UVN[c3];
    wp[0] = U5[c0]; wp[1] = U5[c1]; wp[2] = U5[c2]; wp[3] = U5[c3];
    wa[0] = B1[c0]; wa[1] = B1[c1]; wa[2] = B1[c2]; wa[3] = B1[c3];
    wc[0] = B2[c0]; wc[1] = B2[c1]; wc[2] = B2[c2]; wc[3] = B2[c3];
#pragma unroll
    for (int f = 0; f < 5; ++f) { spw[0][f] = sp[f][c1]; spw[1][f] = sp[f][c2]; }
    rhs_states<R>(wr[1], wv[1], wp[1], wa[1], wc[1], wr[0], wr[2], wv[0], wv[2], wp[0], wp[2], wa[0], wa[2], wc[0],
                  wc[2], spw[0], dxu, dxu, dxu, dt, A.facdep, qprev);
#pragma unroll
    for (int k = 0; k < 3; ++k) { wr[k] = wr[k + 1]; wv[k] = wv[k + 1]; wp[k] = wp[k + 1]; wa[k] = wa[k + 1]; wc[k] = wc[k + 1]; }
  }
  // inputs of the Lagrangian update of cell q-1 (loaded one iteration ahead)
  double vt1 = 0, vt2 = 0, bnc = 0, fb = 0;
  // MOC windows: RHS results of k-2, k-1, k; faces k-2, k-1, k; interface
  // velocities of k-2|k-1, k-1|k, k|k+1; slopes of k-2 and k-1 (regular and,
  // next to block edges, with the ghost switch)
  OR<R> o0{}, o1{}, o2{};
  double f0 = 0, f1 = 0, f2 = 0;
  double va0 = 0, va1 = 0, va2 = 0;
  R G0[5], G1[5], H0[5], H1[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) { G0[v] = G1[v] = H0[v] = H1[v] = (R)0; }
  MocV<R> rvp{};           // right view of interface J-1
  double bnp = 0, vap = 0; // B_n and V* of interface J-1
  for (int q = ps; q <= pe; ++q) {
    const int cf = cell(q + 2), cn = cell(q + 1), cq = cell(q);
    wr[3] = U1[cf]; wv[3] = UVN[cf]; wp[3] = U5[cf]; wa[3] = B1[cf]; wc[3] = B2[cf];
#pragma unroll
    for (int f = 0; f < 5; ++f) spw[2][f] = sp[f][cn];
    const double nvt1 = A.U[VT1][cq], nvt2 = A.U[VT2][cq], nbnc = A.bcc[D][cq], nfb = UBn[cq];
    R qcur[8];
    rhs_states<R>(wr[1], wv[1], wp[1], wa[1], wc[1], wr[0], wr[2], wv[0], wv[2], wp[0], wp[2], wa[0], wa[2], wc[0],
                  wc[2], spw[1], dxu, dxu, dxu, dt, A.facdep, qcur);
    R Pa, Va;
    rhs_iface<R>(spw[0], spw[1], qprev, qcur, Pa, Va);
    const int k = q - 1;
    if (k >= p0 - 1 && k <= p1 && k >= 0 && (k >= p0 || k == 0)) A.va[at(k)] = (double)Va;
    if (k >= ps) {
      // Lagrangian update of cell k (M3:1448-1466) at its image position
      const int kk = img(k);
      const double xq = xe[kk + 1], xqm = xe[kk];
      const double rho = wr[0], vn = wv[0], p = wp[0], b1 = wa[0], b2 = wc[0];
      double vv[3], bb[3];
      vv[D] = vn; vv[(D + 1) % 3] = vt1; vv[(D + 2) % 3] = vt2;
      bb[D] = bnc; bb[BT1] = b1; bb[BT2] = b2;
      const double vx = vv[0], vy = vv[1], vz = vv[2];
      const double bx = bb[0], by = bb[1], bz = bb[2];
      const double dm = dxu * rho;
      const double pa = (double)Pa, pam = (double)pa_prev, va = (double)Va, vam = (double)va_prev;
      const double xl = xq + va * dt, xlm = xqm + vam * dt;
      const double dxl = xl - xlm;
      const double *spl = spw[0];
      double rhoL, vnL, eL, bl1, bl2;
      if constexpr (Ar<R>::strict) {
        double s4 = spl[0] + spl[1] + spl[3] + spl[4];
        double gg = 3.0 * s4 + 5.0 * spl[2];
        double gi = (2.0 * s4 + 2.0 * spl[2]) / gg;
        double e = p / (rho * gi) + 0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz) / rho;
        rhoL = dm / dxl;
        vnL = vn - dt / dm * (pa - pam);
        eL = e - dt / dm * (pa * va - pam * vam);
        bl1 = b1 * dxu / dxl;
        bl2 = b2 * dxu / dxl;
      } else {
        double s4 = spl[0] + spl[1] + spl[3] + spl[4];
        double irho = 1.0 / rho;
        double e = p * (3.0 * s4 + 5.0 * spl[2]) / (2.0 * s4 + 2.0 * spl[2]) * irho +
                   0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz) * irho;
        double idl = 1.0 / dxl, dtdm = dt / dm, cmp = dxu * idl;
        rhoL = dm * idl;
        vnL = vn - dtdm * (pa - pam);
        eL = e - dtdm * (pa * va - pam * vam);
        bl1 = b1 * cmp;
        bl2 = b2 * cmp;
      }
      if (k >= p0 && k <= p1) {
        const int ck = at(k);
        A.Uo[0][ck] = rhoL;
        A.dxlagM[ck] = dxl;
      }
      o2.v[0] = rhoL; o2.v[1] = vt1; o2.v[2] = vt2; o2.v[3] = bl1; o2.v[4] = bl2; o2.v[5] = dxl;
      o2.vn = vnL; o2.e = eL;
      o2.mme = use_minmod_r<R>(eL, rhoL);
      o2.mmp = use_minmod_r<R>(p, rhoL);
      f2 = fb;
      va2 = va;
      if (k >= p0) {
        // MOC slope of cell k-1 (M3:2011-2020)
        const int km = k - 1;
        const bool inner = km >= 1 && km <= n;
        const bool mm = inner ? o1.mme : o1.mmp;
        const R dxm = (R)o0.v[5], dx0 = (R)o1.v[5], dxp = (R)o2.v[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) G1[v] = vslope<R>(o0.v[v], o1.v[v], o2.v[v], dxm, dx0, dxp, mm);
        if ((km % nblk == 0 || (km - 1) % nblk == 0) && o1.mmp != mm) {
#pragma unroll
          for (int v = 0; v < 5; ++v) H1[v] = vslope<R>(o0.v[v], o1.v[v], o2.v[v], dxm, dx0, dxp, o1.mmp);
        }
        if (k >= p0 + 1) {
          // interface J = k-2 between cells J and J+1 (local positions 0..n)
          const int J = k - 2;
          const double bnJ = moc_bn(f0, f1, f2, xe[J + 1] + va0 * dt, xe, J);
          MocV<R> L, Rv;
          if (J % nblk != 0) {
            L = moc_iface_r<R>(o0.v, o1.v, G0, G1, bnJ, A.facdep, dt);
            Rv = L;
          } else {
            const bool mm0p = o0.mmp, mm0e = J >= 1 ? o0.mme : mm0p;
            const bool mm1p = o1.mmp, mm1e = J + 1 <= n ? o1.mme : mm1p;
            L = moc_iface_r<R>(o0.v, o1.v, G0, mm1p != mm1e ? H1 : G1, bnJ, A.facdep, dt);
            if (mm0p != mm0e) Rv = moc_iface_r<R>(o0.v, o1.v, H0, G1, bnJ, A.facdep, dt);
            else if (mm1p != mm1e) Rv = moc_iface_r<R>(o0.v, o1.v, G0, G1, bnJ, A.facdep, dt);
            else Rv = L;
          }
          if (J >= p0) {
            // MOC update of cell J (M3:2053-2095)
            const int cJ = at(J);
            const double Bnap = bnJ, Bnam = bnp, Vnap = va0, Vnam = vap;
            const double dxl0 = o0.v[5], rho0 = o0.v[0];
            const double BayL = (double)L.Bay, VayL = (double)L.Vay, BazL = (double)L.Baz, VazL = (double)L.Vaz;
            const double BayR = (double)rvp.Bay, VayR = (double)rvp.Vay, BazR = (double)rvp.Baz, VazR = (double)rvp.Vaz;
            if constexpr (Ar<R>::strict) {
              A.Uo[VN][cJ] = o0.vn * rho0 + dt / dxl0 * 0.5 * (Bnap * Bnap - Bnam * Bnam);
              A.Uo[VT1][cJ] = o0.v[1] * rho0 + dt / dxl0 * (Bnap * BayL - Bnam * BayR);
              A.Uo[VT2][cJ] = o0.v[2] * rho0 + dt / dxl0 * (Bnap * BazL - Bnam * BazR);
              A.Uo[4][cJ] = o0.e * rho0 + dt / dxl0 *
                                              ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                               (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
              A.blgo[0][cJ] = o0.v[3] + dt / dxl0 * (Bnap * VayL - Bnam * VayR);
              A.blgo[1][cJ] = o0.v[4] + dt / dxl0 * (Bnap * VazL - Bnam * VazR);
            } else {
              const double fq = dt / dxl0;
              A.Uo[VN][cJ] = o0.vn * rho0 + fq * 0.5 * (Bnap * Bnap - Bnam * Bnam);
              A.Uo[VT1][cJ] = o0.v[1] * rho0 + fq * (Bnap * BayL - Bnam * BayR);
              A.Uo[VT2][cJ] = o0.v[2] * rho0 + fq * (Bnap * BazL - Bnam * BazR);
              A.Uo[4][cJ] = o0.e * rho0 + fq * ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                                (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
              A.blgo[0][cJ] = o0.v[3] + fq * (Bnap * VayL - Bnam * VayR);
              A.blgo[1][cJ] = o0.v[4] + fq * (Bnap * VazL - Bnam * VazR);
            }
          }
          rvp = Rv;
          bnp = bnJ;
          vap = va0;
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) { G0[v] = G1[v]; H0[v] = H1[v]; }
      }
      o0 = o1; o1 = o2;
      f0 = f1; f1 = f2;
      va0 = va1; va1 = va2;
    }
    // advance the RHS window by one position
#pragma unroll
    for (int j = 0; j < 3; ++j) { wr[j] = wr[j + 1]; wv[j] = wv[j + 1]; wp[j] = wp[j + 1]; wa[j] = wa[j + 1]; wc[j] = wc[j + 1]; }
#pragma unroll
    for (int f = 0; f < 5; ++f) { spw[0][f] = spw[1][f]; spw[1][f] = spw[2][f]; }
#pragma unroll
    for (int f = 0; f < 8; ++f) qprev[f] = qcur[f];
    pa_prev = Pa;
    va_prev = Va;
    vt1 = nvt1; vt2 = nvt2; bnc = nbnc; fb = nfb;
  }
}

// ---------------------------------------------------------------------------
// REMAP + primitives
// ---------------------------------------------------------------------------
// The 15 remapped quantities (rho, m(3), E, Blg(2), 8 species) go through LDS
// RG at a time, which keeps the tile's LDS small enough for several tiles
// per CU.  Each interface loads the four cells c-1..c+2 around it before it
// knows which of c, c+1 is upwind, so no load waits for the interface
// velocity.
constexpr int RG = 5;

__device__ __forceinline__ const double *remap_field(const SweepIO &A, int v) {
  return v < 5 ? A.U[v] : v < 7 ? A.blg[v - 5] : A.sp[v - 7];
}

template <int D, class R>
__global__ __launch_bounds__(576, 7) void k_remap(SweepIO A, Geo g, double dt) {
  constexpr int NF = nface_slots<D>();
  __shared__ double sF[RG][NF];
  const Roles t = roles<D>(g);
  const int s = g.st[D];
  const int c = ix3s(g, t.i, t.j, t.k);
  // upwind linear flux through the face swept by V*dt (M3:1600-1640)
  double vadt = 0.0, depend = 0.0, dxl = 0.0;
  bool up = true, mm = false;
  R dxm = (R)1, dx0 = (R)1, dxp = (R)1;
  if (t.f_ok) {
    vadt = A.va[c] * dt;
    up = !(vadt <= 0.0);
    const double w0 = A.dxlagM[c - s], w1 = A.dxlagM[c], w2 = A.dxlagM[c + s], w3 = A.dxlagM[c + 2 * s];
    const double u4 = up ? A.U[4][c] : A.U[4][c + s], u0 = up ? A.U[0][c] : A.U[0][c + s];
    dxl = w1;
    if (up) depend = 0.5 * (w1 - vadt) * A.facdep;
    else depend = 0.5 * (w2 + vadt) * A.facdep;
    dxm = (R)(up ? w0 : w1);
    dx0 = (R)(up ? w1 : w2);
    dxp = (R)(up ? w2 : w3);
    mm = use_minmod_r<R>(u4, u0);
  }
  const R rdep = (R)depend;
  const int q = t.q;
  const double dxE = t.u_ok ? A.xe[q + 1] - A.xe[q] : 1.0;
  const double bn = t.u_ok ? A.bcc[D][c] : 0.0;
  double Qn[15];
#pragma unroll
  for (int g0 = 0; g0 < 15; g0 += RG) {
    double Qc[RG];
    if (t.f_ok) {
#pragma unroll
      for (int k = 0; k < RG; ++k) {
        const double *Qv = remap_field(A, g0 + k);
        const double q0 = Qv[c - s], q1 = Qv[c], q2 = Qv[c + s], q3 = Qv[c + 2 * s];
        Qc[k] = q1;
        const R gr = vslope<R>(up ? q0 : q1, up ? q1 : q2, up ? q2 : q3, dxm, dx0, dxp, mm);
        const double qu = up ? q1 : q2;
        double face;
        if constexpr (Ar<R>::strict) face = up ? qu + depend * gr : qu - depend * gr;
        else face = up ? qu + (double)(rdep * gr) : qu - (double)(rdep * gr);
        sF[k][t.fslot] = -(vadt * face);
      }
    }
    __syncthreads();
    if (t.u_ok) {
      // remapped conserved state (M3:1645-1686)
#pragma unroll
      for (int k = 0; k < RG; ++k) {
        if constexpr (Ar<R>::strict) Qn[g0 + k] = (Qc[k] * dxl + sF[k][t.rslot] - sF[k][t.lslot]) / dxE;
        else Qn[g0 + k] = (Qc[k] * dxl + sF[k][t.rslot] - sF[k][t.lslot]) * (1.0 / dxE);
      }
    }
    __syncthreads();
  }
  if (!t.u_ok) return;
  // primitives (M3:1291-1321)
  double rho = dmax(1.27 * A.rmin, Qn[0]);
  double invd = 1.0 / rho;
  double vx = Qn[1] * invd, vy = Qn[2] * invd, vz = Qn[3] * invd;
  double s4 = Qn[7] + Qn[8] + Qn[10] + Qn[11];
  double gg = 3.0 * s4 + 5.0 * Qn[9];
  double gi = (2.0 * s4 + 2.0 * Qn[9]) / gg;
  double ekin = 0.5 * (vx * vx + vy * vy + vz * vz) * rho;
  double emag = 0.5 * (bn * bn + Qn[5] * Qn[5] + Qn[6] * Qn[6]);
  double p = (Qn[4] - ekin - emag) * gi;
  p = dmax(A.pmin, p);
  p = dmin(A.pmax, p);
  A.Uo[0][c] = rho;
  A.Uo[1][c] = vx;
  A.Uo[2][c] = vy;
  A.Uo[3][c] = vz;
  A.Uo[4][c] = p;
#pragma unroll
  for (int f = 0; f < 8; ++f) A.spo[f][c] = dmax(A.spmin[f], Qn[7 + f]);
}

// REMAP + primitives along a periodic axis (y or z) as a column march: one
// thread per x column, cells p0..p1 of a segment, faces p0-1..p1.  The 15
// remapped quantities of positions J-1..J+2 sit in registers; each
// iteration takes one new position (in FP64 loaded one iteration ahead), forms the
// upwind flux through face J and updates cell J.  Values outside 1..n are
// read at the periodic image (the ghost fills are copies), so no fills
// precede this kernel.  The Lagrangian species are formed here from the
// pre-sweep species and dx_lag, with RHS's expression for them, instead of
// being stored by RHS.
template <int D, class R>
__global__ __launch_bounds__(256) void k_remap_col(SweepIO A, Geo g, double dt) {
  const int nx = g.n[0];
  const int n = g.n[D], ntr = g.n[D == 1 ? 2 : 1];
  // grid: segments fastest, so the segments of a column (which share halo
  // cells) run together on one XCD
  int bseg, bxc, btr;
  xcd_tile(bseg, bxc, btr);
  const int i = 1 + 64 * bxc + (int)(threadIdx.x & 63);
  const int tr = 1 + 4 * btr + (int)(threadIdx.x >> 6);
  const int p0 = 1 + FUSE_SEG * bseg;
  if (i > nx || tr > ntr || p0 > n) return;
  const int p1 = min(n, p0 + FUSE_SEG - 1);
  auto img = [&](int p) { return p < 1 ? p + n : (p > n ? p - n : p); };
  auto at = [&](int p) { return D == 1 ? ix3s(g, i, p, tr) : ix3s(g, i, tr, p); };
  auto cell = [&](int p) { return at(img(p)); };
  const double dxu = A.dxe[1];
  const double *xe = A.xe;
  // raw values of one position: rho, m(3), E, Blg(2), species(8), dx_lag
  auto fetch = [&](int p, double *r) {
    const int c = cell(p);
#pragma unroll
    for (int v = 0; v < 5; ++v) r[v] = A.U[v][c];
    r[5] = A.blg[0][c];
    r[6] = A.blg[1][c];
#pragma unroll
    for (int f = 0; f < 8; ++f) r[7 + f] = A.sp[f][c];
    r[15] = A.dxlagM[c];
  };
  // window entry from raw values: the Lagrangian species (M3:1466)
  auto enter = [&](const double *r, double *q, double &w) {
#pragma unroll
    for (int v = 0; v < 7; ++v) q[v] = r[v];
    w = r[15];
    if constexpr (Ar<R>::strict) {
#pragma unroll
      for (int f = 0; f < 8; ++f) q[7 + f] = dxu * r[7 + f] / w;
    } else {
      const double cmp = dxu * (1.0 / w);
#pragma unroll
      for (int f = 0; f < 8; ++f) q[7 + f] = r[7 + f] * cmp;
    }
  };
  double Q0[15], Q1[15], Q2[15], Q3[15], W0, W1, W2, W3;
  {
    double r[16];
    fetch(p0 - 2, r); enter(r, Q0, W0);
    fetch(p0 - 1, r); enter(r, Q1, W1);
    fetch(p0, r); enter(r, Q2, W2);
  }
  // FP64: the next position's loads are issued one iteration ahead; the
  // FP32 instance keeps fewer registers live without it and runs faster
  double nr[16];
  double nva = 0.0;
  if constexpr (Ar<R>::strict) {
    fetch(p0 + 1, nr);
    nva = A.va[at(p0 - 1)];
  }
  double Fp[15];
#pragma unroll
  for (int v = 0; v < 15; ++v) Fp[v] = 0.0;
  for (int J = p0 - 1; J <= p1; ++J) {
    double vaJ;
    if constexpr (Ar<R>::strict) {
      enter(nr, Q3, W3);
      vaJ = nva;
      fetch(J + 3, nr);
      nva = A.va[at(min(J + 1, n))];
    } else {
      fetch(J + 2, nr);
      enter(nr, Q3, W3);
      vaJ = A.va[at(J)];
    }
    // upwind linear flux through face J (M3:1600-1640)
    const double vadt = vaJ * dt;
    const bool up = !(vadt <= 0.0);
    const double u4 = up ? Q1[4] : Q2[4], u0 = up ? Q1[0] : Q2[0];
    const double depend = up ? 0.5 * (W1 - vadt) * A.facdep : 0.5 * (W2 + vadt) * A.facdep;
    const R dxm = (R)(up ? W0 : W1), dx0 = (R)(up ? W1 : W2), dxp = (R)(up ? W2 : W3);
    const bool mm = use_minmod_r<R>(u4, u0);
    const R rdep = (R)depend;
    double F[15];
#pragma unroll
    for (int v = 0; v < 15; ++v) {
      const R gr = vslope<R>(up ? Q0[v] : Q1[v], up ? Q1[v] : Q2[v], up ? Q2[v] : Q3[v], dxm, dx0, dxp, mm);
      const double qu = up ? Q1[v] : Q2[v];
      double face;
      if constexpr (Ar<R>::strict) face = up ? qu + depend * gr : qu - depend * gr;
      else face = up ? qu + (double)(rdep * gr) : qu - (double)(rdep * gr);
      F[v] = -(vadt * face);
    }
    if (J >= p0) {
      // remapped conserved state (M3:1645-1686) and primitives (M3:1291-1321)
      const double dxE = xe[J + 1] - xe[J];
      double Qn[15];
#pragma unroll
      for (int v = 0; v < 15; ++v) {
        if constexpr (Ar<R>::strict) Qn[v] = (Q1[v] * W1 + F[v] - Fp[v]) / dxE;
        else Qn[v] = (Q1[v] * W1 + F[v] - Fp[v]) * (1.0 / dxE);
      }
      const int c = at(J);
      const double bn = A.bcc[D][c];
      double rho = dmax(1.27 * A.rmin, Qn[0]);
      double invd = 1.0 / rho;
      double vx = Qn[1] * invd, vy = Qn[2] * invd, vz = Qn[3] * invd;
      double s4 = Qn[7] + Qn[8] + Qn[10] + Qn[11];
      double gg = 3.0 * s4 + 5.0 * Qn[9];
      double gi = (2.0 * s4 + 2.0 * Qn[9]) / gg;
      double ekin = 0.5 * (vx * vx + vy * vy + vz * vz) * rho;
      double emag = 0.5 * (bn * bn + Qn[5] * Qn[5] + Qn[6] * Qn[6]);
      double p = (Qn[4] - ekin - emag) * gi;
      p = dmax(A.pmin, p);
      p = dmin(A.pmax, p);
      A.Uo[0][c] = rho;
      A.Uo[1][c] = vx;
      A.Uo[2][c] = vy;
      A.Uo[3][c] = vz;
      A.Uo[4][c] = p;
#pragma unroll
      for (int f = 0; f < 8; ++f) A.spo[f][c] = dmax(A.spmin[f], Qn[7 + f]);
    }
#pragma unroll
    for (int v = 0; v < 15; ++v) { Fp[v] = F[v]; Q0[v] = Q1[v]; Q1[v] = Q2[v]; Q2[v] = Q3[v]; }
    W0 = W1; W1 = W2; W2 = W3;
  }
}

// ---------------------------------------------------------------------------
// A whole y or z sweep in one column march
// ---------------------------------------------------------------------------
// The chain of k_rhsmoc_col (RHS states, Riemann interface, Lagrangian
// update, MOC slope, characteristic interface, MOC update) and, in the same
// thread, that of k_remap_col (upwind flux, remap, primitive recovery): the
// MOC result of each cell enters REMAP's register window instead of HBM, and
// the Lagrangian species are formed from the pre-sweep species as they enter
// it.  Only the pre-sweep state is read and only the new state is written.
//
// A segment updates cells p0..p1; REMAP needs the MOC results of p0-2..p1+2.
// The result of a halo cell must be the one its owning block of the
// reproduced MPI run computes, which for a position outside 1..n is that of
// the periodic image.  So every cell is evaluated in its own block's view,
// with that block's coordinates: its MOC limiter switch from e (RHS's
// specific energy; ghost cells of a view take the pre-sweep pressure
// instead), and an interface at a block edge (position a multiple of the
// block length; 0 and n are block edges) evaluated once in each neighbour's
// view -- the left view for the update of the left cell, the right view for
// the right cell.  The two views also differ in the position at which B_n is
// interpolated when the two cells' owners see the interface at different
// coordinates (interface 0 = interface n of the last block).  For the cells
// 1..n this is the evaluation of k_rhsmoc_col, which treats positions outside
// 1..n as ghosts of the adjacent block.
template <class R> struct MocOut {
  double v[7]; // rho, m_n, m_t1, m_t2, E, Blg1, Blg2 (global component order via SweepIO)
  double w;    // dx_lag
  double va;   // V* of the interface to the right
};

template <int D, class R>
__global__ __launch_bounds__(256) void k_sweep_col(SweepIO A, Geo g, double dt, int seg) {
  constexpr int VN = D + 1, VT1 = (D + 1) % 3 + 1, VT2 = (D + 2) % 3 + 1;
  constexpr int BT1 = (D + 1) % 3, BT2 = (D + 2) % 3;
  const int nx = g.n[0];
  const int n = g.n[D], ntr = g.n[D == 1 ? 2 : 1];
  int bseg, bxc, btr;
  xcd_tile(bseg, bxc, btr);
  const int i = 1 + 64 * bxc + (int)(threadIdx.x & 63);
  const int tr = 1 + 4 * btr + (int)(threadIdx.x >> 6);
  const int p0 = 1 + seg * bseg;
  const int p1 = min(n, p0 + seg - 1);
  const int m0 = p0 - 2; // first MOC result (m0 .. p1+2)
  const int nblk = A.nblk;
  // cell edges along the sweep axis in LDS: every iteration reads several
  // (the same for the whole block), which would otherwise wait on memory
  __shared__ double sxe[SWEEP_COL_MAXN + 4];
  for (int t = (int)threadIdx.x; t < n + 4; t += (int)blockDim.x) sxe[t] = A.xe[t];
  __syncthreads();
  if (i > nx || tr > ntr || p0 > n) return;
  // image position of p (n >= 8, so one shift suffices for p0-6 .. p1+7)
  auto img = [&](int p) { return p < 1 ? p + n : (p > n ? p - n : p); };
  auto at = [&](int p) { return D == 1 ? ix3s(g, i, p, tr) : ix3s(g, i, tr, p); };
  auto cell = [&](int p) { return at(img(p)); };
  const double *U1 = A.U[0], *UVN = A.U[VN], *U5 = A.U[4];
  const double *B1 = A.bcc[BT1], *B2 = A.bcc[BT2];
  const double *UBn = A.U[5 + D];
  const double *const *sp = A.sp;
  const double *xe = sxe;
  const double dxu = A.dxe[1];
  // ---- RHS window: positions q-1, q, q+1 and the values of q+2 in flight
  double wr[4], wv[4], wp[4], wa[4], wc[4];
  double spw[3][5];
  R qprev[8];
  R pa_prev = 0, va_prev = 0;
  const int ps = m0 - 2, pe = p1 + 5;
  {
    const int c0 = cell(ps - 2), c1 = cell(ps - 1), c2 = cell(ps), c3 = cell(ps + 1);
    wr[0] = U1[c0]; wr[1] = U1[c1]; wr[2] = U1[c2]; wr[3] = U1[c3];
    wv[0] = UVN[c0]; wv[1] = UVN[c1]; wv[2] = UVN[c2]; wv[3] = UVN[c3];
    wp[0] = U5[c0]; wp[1] = U5[c1]; wp[2] = U5[c2]; wp[3] = U5[c3];
    wa[0] = B1[c0]; wa[1] = B1[c1]; wa[2] = B1[c2]; wa[3] = B1[c3];
    wc[0] = B2[c0]; wc[1] = B2[c1]; wc[2] = B2[c2]; wc[3] = B2[c3];
#pragma unroll
    for (int f = 0; f < 5; ++f) { spw[0][f] = sp[f][c1]; spw[1][f] = sp[f][c2]; }
    rhs_states<R>(wr[1], wv[1], wp[1], wa[1], wc[1], wr[0], wr[2], wv[0], wv[2], wp[0], wp[2], wa[0], wa[2], wc[0],
                  wc[2], spw[0], dxu, dxu, dxu, dt, A.facdep, qprev);
#pragma unroll
    for (int k = 0; k < 3; ++k) { wr[k] = wr[k + 1]; wv[k] = wv[k + 1]; wp[k] = wp[k + 1]; wa[k] = wa[k + 1]; wc[k] = wc[k + 1]; }
  }
  double vt1 = 0, vt2 = 0, bnc = 0, fb = 0;
  // ---- MOC windows (as k_rhsmoc_col)
  OR<R> o0{}, o1{}, o2{};
  double f0 = 0, f1 = 0, f2 = 0;
  double va0 = 0, va1 = 0, va2 = 0;
  R G0[5], G1[5], H0[5], H1[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) { G0[v] = G1[v] = H0[v] = H1[v] = (R)0; }
  MocV<R> rvp{};            // right view of interface J-1
  double bnp = 0, vap = 0;  // B_n of interface J-1 in the right view, V* of J-1
  // ---- REMAP window: remapped quantities of cells Jf-1 .. Jf+2 (15 each),
  // their dx_lag, V* of face Jf, the flux of face Jf-1
  double Q0[15], Q1[15], Q2[15], Q3[15], W0 = 0, W1 = 0, W2 = 0, W3 = 0, V1 = 0, V2 = 0, V3 = 0;
  double Fp[15];
#pragma unroll
  for (int v = 0; v < 15; ++v) { Q0[v] = Q1[v] = Q2[v] = Q3[v] = 0.0; Fp[v] = 0.0; }
  // pre-sweep species of the cell whose MOC result enters REMAP's window in
  // this iteration (J = q-3) and B_n of the cell REMAP updates (q-5), each
  // loaded one iteration ahead
  double csp[8], cbn;
  {
    const int cs = cell(ps - 3), cb = cell(ps - 5);
#pragma unroll
    for (int f = 0; f < 8; ++f) csp[f] = sp[f][cs];
    cbn = A.bcc[D][cb];
  }
  for (int q = ps; q <= pe; ++q) {
    const int cf = cell(q + 2), cn = cell(q + 1), cq = cell(q);
    wr[3] = U1[cf]; wv[3] = UVN[cf]; wp[3] = U5[cf]; wa[3] = B1[cf]; wc[3] = B2[cf];
#pragma unroll
    for (int f = 0; f < 5; ++f) spw[2][f] = sp[f][cn];
    const double nvt1 = A.U[VT1][cq], nvt2 = A.U[VT2][cq], nbnc = A.bcc[D][cq], nfb = UBn[cq];
    double nsp[8];
    {
      const int cs = cell(q - 2);
#pragma unroll
      for (int f = 0; f < 8; ++f) nsp[f] = sp[f][cs];
    }
    const double nbn = A.bcc[D][cell(q - 4)];
    R qcur[8];
    rhs_states<R>(wr[1], wv[1], wp[1], wa[1], wc[1], wr[0], wr[2], wv[0], wv[2], wp[0], wp[2], wa[0], wa[2], wc[0],
                  wc[2], spw[1], dxu, dxu, dxu, dt, A.facdep, qcur);
    R Pa, Va;
    rhs_iface<R>(spw[0], spw[1], qprev, qcur, Pa, Va);
    const int k = q - 1;
    if (k >= ps) {
      // Lagrangian update of cell k (M3:1448-1466) at its image position
      const int kk = img(k);
      const double xq = xe[kk + 1], xqm = xe[kk];
      const double rho = wr[0], vn = wv[0], p = wp[0], b1 = wa[0], b2 = wc[0];
      double vv[3], bb[3];
      vv[D] = vn; vv[(D + 1) % 3] = vt1; vv[(D + 2) % 3] = vt2;
      bb[D] = bnc; bb[BT1] = b1; bb[BT2] = b2;
      const double vx = vv[0], vy = vv[1], vz = vv[2];
      const double bx = bb[0], by = bb[1], bz = bb[2];
      const double dm = dxu * rho;
      const double pa = (double)Pa, pam = (double)pa_prev, va = (double)Va, vam = (double)va_prev;
      const double xl = xq + va * dt, xlm = xqm + vam * dt;
      const double dxl = xl - xlm;
      const double *spl = spw[0];
      double rhoL, vnL, eL, bl1, bl2;
      if constexpr (Ar<R>::strict) {
        double s4 = spl[0] + spl[1] + spl[3] + spl[4];
        double gg = 3.0 * s4 + 5.0 * spl[2];
        double gi = (2.0 * s4 + 2.0 * spl[2]) / gg;
        double e = p / (rho * gi) + 0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz) / rho;
        rhoL = dm / dxl;
        vnL = vn - dt / dm * (pa - pam);
        eL = e - dt / dm * (pa * va - pam * vam);
        bl1 = b1 * dxu / dxl;
        bl2 = b2 * dxu / dxl;
      } else {
        double s4 = spl[0] + spl[1] + spl[3] + spl[4];
        double irho = 1.0 / rho;
        double e = p * (3.0 * s4 + 5.0 * spl[2]) / (2.0 * s4 + 2.0 * spl[2]) * irho +
                   0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz) * irho;
        double idl = 1.0 / dxl, dtdm = dt / dm, cmp = dxu * idl;
        rhoL = dm * idl;
        vnL = vn - dtdm * (pa - pam);
        eL = e - dtdm * (pa * va - pam * vam);
        bl1 = b1 * cmp;
        bl2 = b2 * cmp;
      }
      o2.v[0] = rhoL; o2.v[1] = vt1; o2.v[2] = vt2; o2.v[3] = bl1; o2.v[4] = bl2; o2.v[5] = dxl;
      o2.vn = vnL; o2.e = eL;
      o2.mme = use_minmod_r<R>(eL, rhoL);
      o2.mmp = use_minmod_r<R>(p, rhoL);
      f2 = fb;
      va2 = va;
      if (k >= m0) {
        // MOC slope of cell k-1 (M3:2011-2020) in its own block's view; next
        // to a block edge also with the ghost switch (the neighbour's view)
        const int km = k - 1;
        const bool mm = o1.mme;
        const R dxm = (R)o0.v[5], dx0 = (R)o1.v[5], dxp = (R)o2.v[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) G1[v] = vslope<R>(o0.v[v], o1.v[v], o2.v[v], dxm, dx0, dxp, mm);
        if ((km % nblk == 0 || (km - 1) % nblk == 0) && o1.mmp != mm) {
#pragma unroll
          for (int v = 0; v < 5; ++v) H1[v] = vslope<R>(o0.v[v], o1.v[v], o2.v[v], dxm, dx0, dxp, o1.mmp);
        }
        if (k >= m0 + 1) {
          // interface J = k-2 between cells J and J+1; the owners of J and
          // J+1 see it at positions img(J) and img(J+1)-1
          const int J = k - 2;
          const int gL = img(J), gR = img(J + 1) - 1;
          const double bnL = moc_bn(f0, f1, f2, xe[gL + 1] + va0 * dt, xe, gL);
          const double bnR = gR == gL ? bnL : moc_bn(f0, f1, f2, xe[gR + 1] + va0 * dt, xe, gR);
          MocV<R> L, Rv;
          if (J % nblk != 0) {
            L = moc_iface_r<R>(o0.v, o1.v, G0, G1, bnL, A.facdep, dt);
            Rv = L;
          } else {
            const bool g0 = o0.mmp != o0.mme, g1 = o1.mmp != o1.mme;
            L = moc_iface_r<R>(o0.v, o1.v, G0, g1 ? H1 : G1, bnL, A.facdep, dt);
            if (g0 || g1 || gR != gL) Rv = moc_iface_r<R>(o0.v, o1.v, g0 ? H0 : G0, G1, bnR, A.facdep, dt);
            else Rv = L;
          }
          if (J >= m0) {
            // MOC update of cell J (M3:2053-2095) into REMAP's window
            const double Bnap = bnL, Bnam = bnp, Vnap = va0, Vnam = vap;
            const double dxl0 = o0.v[5], rho0 = o0.v[0];
            const double BayL = (double)L.Bay, VayL = (double)L.Vay, BazL = (double)L.Baz, VazL = (double)L.Vaz;
            const double BayR = (double)rvp.Bay, VayR = (double)rvp.Vay, BazR = (double)rvp.Baz, VazR = (double)rvp.Vaz;
            if constexpr (Ar<R>::strict) {
              Q3[VN] = o0.vn * rho0 + dt / dxl0 * 0.5 * (Bnap * Bnap - Bnam * Bnam);
              Q3[VT1] = o0.v[1] * rho0 + dt / dxl0 * (Bnap * BayL - Bnam * BayR);
              Q3[VT2] = o0.v[2] * rho0 + dt / dxl0 * (Bnap * BazL - Bnam * BazR);
              Q3[4] = o0.e * rho0 + dt / dxl0 *
                                        ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                         (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
              Q3[5] = o0.v[3] + dt / dxl0 * (Bnap * VayL - Bnam * VayR);
              Q3[6] = o0.v[4] + dt / dxl0 * (Bnap * VazL - Bnam * VazR);
            } else {
              const double fq = dt / dxl0;
              Q3[VN] = o0.vn * rho0 + fq * 0.5 * (Bnap * Bnap - Bnam * Bnam);
              Q3[VT1] = o0.v[1] * rho0 + fq * (Bnap * BayL - Bnam * BayR);
              Q3[VT2] = o0.v[2] * rho0 + fq * (Bnap * BazL - Bnam * BazR);
              Q3[4] = o0.e * rho0 + fq * ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                          (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
              Q3[5] = o0.v[3] + fq * (Bnap * VayL - Bnam * VayR);
              Q3[6] = o0.v[4] + fq * (Bnap * VazL - Bnam * VazR);
            }
            Q3[0] = rho0;
            W3 = dxl0;
            V3 = Vnap;
            // Lagrangian species of cell J (M3:1466) from the pre-sweep ones
            if constexpr (Ar<R>::strict) {
#pragma unroll
              for (int f = 0; f < 8; ++f) Q3[7 + f] = dxu * csp[f] / dxl0;
            } else {
              const double cmp = dxu * (1.0 / dxl0);
#pragma unroll
              for (int f = 0; f < 8; ++f) Q3[7 + f] = csp[f] * cmp;
            }
            // REMAP of face Jf = J-2 (cells Jf-1 .. Jf+2 in Q0 .. Q3) and of
            // cell Jf (faces Jf-1 and Jf)
            const int Jf = J - 2;
            if (Jf >= p0 - 1) {
              const double vadt = V1 * dt;
              const bool up = !(vadt <= 0.0);
              const double u4 = up ? Q1[4] : Q2[4], u0 = up ? Q1[0] : Q2[0];
              const double depend = up ? 0.5 * (W1 - vadt) * A.facdep : 0.5 * (W2 + vadt) * A.facdep;
              const R dxm2 = (R)(up ? W0 : W1), dx02 = (R)(up ? W1 : W2), dxp2 = (R)(up ? W2 : W3);
              const bool mmr = use_minmod_r<R>(u4, u0);
              const R rdep = (R)depend;
              double F[15];
#pragma unroll
              for (int v = 0; v < 15; ++v) {
                const R gr = vslope<R>(up ? Q0[v] : Q1[v], up ? Q1[v] : Q2[v], up ? Q2[v] : Q3[v], dxm2, dx02, dxp2, mmr);
                const double qu = up ? Q1[v] : Q2[v];
                double face;
                if constexpr (Ar<R>::strict) face = up ? qu + depend * gr : qu - depend * gr;
                else face = up ? qu + (double)(rdep * gr) : qu - (double)(rdep * gr);
                F[v] = -(vadt * face);
              }
              if (Jf >= p0) {
                // remapped conserved state (M3:1645-1686), primitives (M3:1291-1321)
                const double dxE = xe[Jf + 1] - xe[Jf];
                double Qn[15];
#pragma unroll
                for (int v = 0; v < 15; ++v) {
                  if constexpr (Ar<R>::strict) Qn[v] = (Q1[v] * W1 + F[v] - Fp[v]) / dxE;
                  else Qn[v] = (Q1[v] * W1 + F[v] - Fp[v]) * (1.0 / dxE);
                }
                const int c = at(Jf);
                const double bn = cbn;
                double rho = dmax(1.27 * A.rmin, Qn[0]);
                double invd = 1.0 / rho;
                double vx = Qn[1] * invd, vy = Qn[2] * invd, vz = Qn[3] * invd;
                double s4 = Qn[7] + Qn[8] + Qn[10] + Qn[11];
                double gg = 3.0 * s4 + 5.0 * Qn[9];
                double gi = (2.0 * s4 + 2.0 * Qn[9]) / gg;
                double ekin = 0.5 * (vx * vx + vy * vy + vz * vz) * rho;
                double emag = 0.5 * (bn * bn + Qn[5] * Qn[5] + Qn[6] * Qn[6]);
                double pr = (Qn[4] - ekin - emag) * gi;
                pr = dmax(A.pmin, pr);
                pr = dmin(A.pmax, pr);
                A.Uo[0][c] = rho;
                A.Uo[1][c] = vx;
                A.Uo[2][c] = vy;
                A.Uo[3][c] = vz;
                A.Uo[4][c] = pr;
#pragma unroll
                for (int f = 0; f < 8; ++f) A.spo[f][c] = dmax(A.spmin[f], Qn[7 + f]);
              }
#pragma unroll
              for (int v = 0; v < 15; ++v) Fp[v] = F[v];
            }
#pragma unroll
            for (int v = 0; v < 15; ++v) { Q0[v] = Q1[v]; Q1[v] = Q2[v]; Q2[v] = Q3[v]; }
            W0 = W1; W1 = W2; W2 = W3;
            V1 = V2; V2 = V3;
          }
          rvp = Rv;
          bnp = bnR;
          vap = va0;
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) { G0[v] = G1[v]; H0[v] = H1[v]; }
      }
      o0 = o1; o1 = o2;
      f0 = f1; f1 = f2;
      va0 = va1; va1 = va2;
    }
    // advance the RHS window by one position
#pragma unroll
    for (int j = 0; j < 3; ++j) { wr[j] = wr[j + 1]; wv[j] = wv[j + 1]; wp[j] = wp[j + 1]; wa[j] = wa[j + 1]; wc[j] = wc[j + 1]; }
#pragma unroll
    for (int f = 0; f < 5; ++f) { spw[0][f] = spw[1][f]; spw[1][f] = spw[2][f]; }
#pragma unroll
    for (int f = 0; f < 8; ++f) qprev[f] = qcur[f];
    pa_prev = Pa;
    va_prev = Va;
    vt1 = nvt1; vt2 = nvt2; bnc = nbnc; fb = nfb;
#pragma unroll
    for (int f = 0; f < 8; ++f) csp[f] = nsp[f];
    cbn = nbn;
  }
}

// A whole y or z sweep as k_sweep_col, over two waves per row of 64
// columns: the producer wave runs the RHS and MOC chain and puts, per cell J,
// its MOC-updated state (rho, m_n, m_t1, m_t2, E, Blg1, Blg2), dx_lag and V*
// into an LDS slot; the consumer wave takes it one iteration later, forms
// the Lagrangian species and remaps.  Each wave holds only its own windows,
// so a SIMD runs a producer and a consumer wave side by side.  The arithmetic
// is k_sweep_col's, statement for statement.  Lanes past the grid's end in x
// or in the transverse axis march the last column and write nothing (every
// wave reaches each barrier).
template <int D, class R>
__global__ __launch_bounds__(512) void k_sweep_col2(SweepIO A, Geo g, double dt, int seg) {
  constexpr int VN = D + 1, VT1 = (D + 1) % 3 + 1, VT2 = (D + 2) % 3 + 1;
  constexpr int BT1 = (D + 1) % 3, BT2 = (D + 2) % 3;
  const int nx = g.n[0];
  const int n = g.n[D], ntr = g.n[D == 1 ? 2 : 1];
  int bseg, bxc, btr;
  xcd_tile(bseg, bxc, btr);
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6, r4 = wv & 3;
  const bool producer = wv < 4;
  const int i_ = 1 + 64 * bxc + lane, tr_ = 1 + 4 * btr + r4;
  const bool act = i_ <= nx && tr_ <= ntr;
  const int i = min(i_, nx), tr = min(tr_, ntr);
  const int p0 = 1 + seg * bseg;
  const int p1 = min(n, p0 + seg - 1);
  const int m0 = p0 - 2; // first MOC result (m0 .. p1+2)
  const int nblk = A.nblk;
  __shared__ double sxe[SWEEP_COL_MAXN + 4];
  // producer -> consumer: [slot][quantity][row][lane], quantities 0..6 the
  // MOC-updated state in REMAP's order, 7 dx_lag, 8 V*
  __shared__ double hand[2][9][4][64];
  for (int t = (int)threadIdx.x; t < n + 4; t += (int)blockDim.x) sxe[t] = A.xe[t];
  __syncthreads();
  // image position of p (n >= 8, so one shift suffices for p0-6 .. p1+7)
  auto img = [&](int p) { return p < 1 ? p + n : (p > n ? p - n : p); };
  auto at = [&](int p) { return D == 1 ? ix3s(g, i, p, tr) : ix3s(g, i, tr, p); };
  auto cell = [&](int p) { return at(img(p)); };
  const double *const *sp = A.sp;
  const double *xe = sxe;
  const double dxu = A.dxe[1];
  const int ps = m0 - 2, pe = p1 + 5;
  if (producer) {
    const double *U1 = A.U[0], *UVN = A.U[VN], *U5 = A.U[4];
    const double *B1 = A.bcc[BT1], *B2 = A.bcc[BT2];
    const double *UBn = A.U[5 + D];
    // ---- RHS window: positions q-1, q, q+1 and the values of q+2 in flight
    double wr[4], wv[4], wp[4], wa[4], wc[4];
    double spw[3][5];
    R qprev[8];
    R pa_prev = 0, va_prev = 0;
    {
      const int c0 = cell(ps - 2), c1 = cell(ps - 1), c2 = cell(ps), c3 = cell(ps + 1);
      wr[0] = U1[c0]; wr[1] = U1[c1]; wr[2] = U1[c2]; wr[3] = U1[c3];
      wv[0] = UVN[c0]; wv[1] = UVN[c1]; wv[2] = UVN[c2]; wv[3] = UVN[c3];
      wp[0] = U5[c0]; wp[1] = U5[c1]; wp[2] = U5[c2]; wp[3] = U5[c3];
      wa[0] = B1[c0]; wa[1] = B1[c1]; wa[2] = B1[c2]; wa[3] = B1[c3];
      wc[0] = B2[c0]; wc[1] = B2[c1]; wc[2] = B2[c2]; wc[3] = B2[c3];
#pragma unroll
      for (int f = 0; f < 5; ++f) { spw[0][f] = sp[f][c1]; spw[1][f] = sp[f][c2]; }
      rhs_states<R>(wr[1], wv[1], wp[1], wa[1], wc[1], wr[0], wr[2], wv[0], wv[2], wp[0], wp[2], wa[0], wa[2], wc[0],
                    wc[2], spw[0], dxu, dxu, dxu, dt, A.facdep, qprev);
#pragma unroll
      for (int k = 0; k < 3; ++k) { wr[k] = wr[k + 1]; wv[k] = wv[k + 1]; wp[k] = wp[k + 1]; wa[k] = wa[k + 1]; wc[k] = wc[k + 1]; }
    }
    double vt1 = 0, vt2 = 0, bnc = 0, fb = 0;
    // ---- MOC windows (as k_rhsmoc_col)
    OR<R> o0{}, o1{}, o2{};
    double f0 = 0, f1 = 0, f2 = 0;
    double va0 = 0, va1 = 0, va2 = 0;
    R G0[5], G1[5], H0[5], H1[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) { G0[v] = G1[v] = H0[v] = H1[v] = (R)0; }
    MocV<R> rvp{};            // right view of interface J-1
    double bnp = 0, vap = 0;  // B_n of interface J-1 in the right view, V* of J-1
    for (int q = ps; q <= pe + 1; ++q) {
      __syncthreads();
      if (q > pe) continue;
      const int cf = cell(q + 2), cn = cell(q + 1), cq = cell(q);
      wr[3] = U1[cf]; wv[3] = UVN[cf]; wp[3] = U5[cf]; wa[3] = B1[cf]; wc[3] = B2[cf];
#pragma unroll
      for (int f = 0; f < 5; ++f) spw[2][f] = sp[f][cn];
      const double nvt1 = A.U[VT1][cq], nvt2 = A.U[VT2][cq], nbnc = A.bcc[D][cq], nfb = UBn[cq];
      R qcur[8];
      rhs_states<R>(wr[1], wv[1], wp[1], wa[1], wc[1], wr[0], wr[2], wv[0], wv[2], wp[0], wp[2], wa[0], wa[2], wc[0],
                    wc[2], spw[1], dxu, dxu, dxu, dt, A.facdep, qcur);
      R Pa, Va;
      rhs_iface<R>(spw[0], spw[1], qprev, qcur, Pa, Va);
      const int k = q - 1;
      if (k >= ps) {
        // Lagrangian update of cell k (M3:1448-1466) at its image position
        const int kk = img(k);
        const double xq = xe[kk + 1], xqm = xe[kk];
        const double rho = wr[0], vn = wv[0], p = wp[0], b1 = wa[0], b2 = wc[0];
        double vv[3], bb[3];
        vv[D] = vn; vv[(D + 1) % 3] = vt1; vv[(D + 2) % 3] = vt2;
        bb[D] = bnc; bb[BT1] = b1; bb[BT2] = b2;
        const double vx = vv[0], vy = vv[1], vz = vv[2];
        const double bx = bb[0], by = bb[1], bz = bb[2];
        const double dm = dxu * rho;
        const double pa = (double)Pa, pam = (double)pa_prev, va = (double)Va, vam = (double)va_prev;
        const double xl = xq + va * dt, xlm = xqm + vam * dt;
        const double dxl = xl - xlm;
        const double *spl = spw[0];
        double rhoL, vnL, eL, bl1, bl2;
        if constexpr (Ar<R>::strict) {
          double s4 = spl[0] + spl[1] + spl[3] + spl[4];
          double gg = 3.0 * s4 + 5.0 * spl[2];
          double gi = (2.0 * s4 + 2.0 * spl[2]) / gg;
          double e = p / (rho * gi) + 0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz) / rho;
          rhoL = dm / dxl;
          vnL = vn - dt / dm * (pa - pam);
          eL = e - dt / dm * (pa * va - pam * vam);
          bl1 = b1 * dxu / dxl;
          bl2 = b2 * dxu / dxl;
        } else {
          double s4 = spl[0] + spl[1] + spl[3] + spl[4];
          double irho = 1.0 / rho;
          double e = p * (3.0 * s4 + 5.0 * spl[2]) / (2.0 * s4 + 2.0 * spl[2]) * irho +
                     0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz) * irho;
          double idl = 1.0 / dxl, dtdm = dt / dm, cmp = dxu * idl;
          rhoL = dm * idl;
          vnL = vn - dtdm * (pa - pam);
          eL = e - dtdm * (pa * va - pam * vam);
          bl1 = b1 * cmp;
          bl2 = b2 * cmp;
        }
        o2.v[0] = rhoL; o2.v[1] = vt1; o2.v[2] = vt2; o2.v[3] = bl1; o2.v[4] = bl2; o2.v[5] = dxl;
        o2.vn = vnL; o2.e = eL;
        o2.mme = use_minmod_r<R>(eL, rhoL);
        o2.mmp = use_minmod_r<R>(p, rhoL);
        f2 = fb;
        va2 = va;
        if (k >= m0) {
          // MOC slope of cell k-1 (M3:2011-2020) in its own block's view; next
          // to a block edge also with the ghost switch (the neighbour's view)
          const int km = k - 1;
          const bool mm = o1.mme;
          const R dxm = (R)o0.v[5], dx0 = (R)o1.v[5], dxp = (R)o2.v[5];
#pragma unroll
          for (int v = 0; v < 5; ++v) G1[v] = vslope<R>(o0.v[v], o1.v[v], o2.v[v], dxm, dx0, dxp, mm);
          if ((km % nblk == 0 || (km - 1) % nblk == 0) && o1.mmp != mm) {
#pragma unroll
            for (int v = 0; v < 5; ++v) H1[v] = vslope<R>(o0.v[v], o1.v[v], o2.v[v], dxm, dx0, dxp, o1.mmp);
          }
          if (k >= m0 + 1) {
            // interface J = k-2 between cells J and J+1; the owners of J and
            // J+1 see it at positions img(J) and img(J+1)-1
            const int J = k - 2;
            const int gL = img(J), gR = img(J + 1) - 1;
            const double bnL = moc_bn(f0, f1, f2, xe[gL + 1] + va0 * dt, xe, gL);
            const double bnR = gR == gL ? bnL : moc_bn(f0, f1, f2, xe[gR + 1] + va0 * dt, xe, gR);
            MocV<R> L, Rv;
            if (J % nblk != 0) {
              L = moc_iface_r<R>(o0.v, o1.v, G0, G1, bnL, A.facdep, dt);
              Rv = L;
            } else {
              const bool g0 = o0.mmp != o0.mme, g1 = o1.mmp != o1.mme;
              L = moc_iface_r<R>(o0.v, o1.v, G0, g1 ? H1 : G1, bnL, A.facdep, dt);
              if (g0 || g1 || gR != gL) Rv = moc_iface_r<R>(o0.v, o1.v, g0 ? H0 : G0, G1, bnR, A.facdep, dt);
              else Rv = L;
            }
            if (J >= m0) {
              // MOC update of cell J (M3:2053-2095), to the consumer wave
              const double Bnap = bnL, Bnam = bnp, Vnap = va0, Vnam = vap;
              const double dxl0 = o0.v[5], rho0 = o0.v[0];
              const double BayL = (double)L.Bay, VayL = (double)L.Vay, BazL = (double)L.Baz, VazL = (double)L.Vaz;
              const double BayR = (double)rvp.Bay, VayR = (double)rvp.Vay, BazR = (double)rvp.Baz, VazR = (double)rvp.Vaz;
              double Q3[7];
              if constexpr (Ar<R>::strict) {
                Q3[VN] = o0.vn * rho0 + dt / dxl0 * 0.5 * (Bnap * Bnap - Bnam * Bnam);
                Q3[VT1] = o0.v[1] * rho0 + dt / dxl0 * (Bnap * BayL - Bnam * BayR);
                Q3[VT2] = o0.v[2] * rho0 + dt / dxl0 * (Bnap * BazL - Bnam * BazR);
                Q3[4] = o0.e * rho0 + dt / dxl0 *
                                          ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                           (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
                Q3[5] = o0.v[3] + dt / dxl0 * (Bnap * VayL - Bnam * VayR);
                Q3[6] = o0.v[4] + dt / dxl0 * (Bnap * VazL - Bnam * VazR);
              } else {
                const double fq = dt / dxl0;
                Q3[VN] = o0.vn * rho0 + fq * 0.5 * (Bnap * Bnap - Bnam * Bnam);
                Q3[VT1] = o0.v[1] * rho0 + fq * (Bnap * BayL - Bnam * BayR);
                Q3[VT2] = o0.v[2] * rho0 + fq * (Bnap * BazL - Bnam * BazR);
                Q3[4] = o0.e * rho0 + fq * ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                            (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
                Q3[5] = o0.v[3] + fq * (Bnap * VayL - Bnam * VayR);
                Q3[6] = o0.v[4] + fq * (Bnap * VazL - Bnam * VazR);
              }
              Q3[0] = rho0;
#pragma unroll
              for (int v = 0; v < 7; ++v) hand[q & 1][v][r4][lane] = Q3[v];
              hand[q & 1][7][r4][lane] = dxl0;
              hand[q & 1][8][r4][lane] = Vnap;
            }
            rvp = Rv;
            bnp = bnR;
            vap = va0;
          }
#pragma unroll
          for (int v = 0; v < 5; ++v) { G0[v] = G1[v]; H0[v] = H1[v]; }
        }
        o0 = o1; o1 = o2;
        f0 = f1; f1 = f2;
        va0 = va1; va1 = va2;
      }
      // advance the RHS window by one position
#pragma unroll
      for (int j = 0; j < 3; ++j) { wr[j] = wr[j + 1]; wv[j] = wv[j + 1]; wp[j] = wp[j + 1]; wa[j] = wa[j + 1]; wc[j] = wc[j + 1]; }
#pragma unroll
      for (int f = 0; f < 5; ++f) { spw[0][f] = spw[1][f]; spw[1][f] = spw[2][f]; }
#pragma unroll
      for (int f = 0; f < 8; ++f) qprev[f] = qcur[f];
      pa_prev = Pa;
      va_prev = Va;
      vt1 = nvt1; vt2 = nvt2; bnc = nbnc; fb = nfb;
    }
  } else {
    // ---- REMAP window: remapped quantities of cells Jf-1 .. Jf+2 (15 each),
    // their dx_lag, V* of face Jf, the flux of face Jf-1
    double Q0[15], Q1[15], Q2[15], Q3[15], W0 = 0, W1 = 0, W2 = 0, W3 = 0, V1 = 0, V2 = 0, V3 = 0;
    double Fp[15];
#pragma unroll
    for (int v = 0; v < 15; ++v) { Q0[v] = Q1[v] = Q2[v] = Q3[v] = 0.0; Fp[v] = 0.0; }
    // pre-sweep species of the cell whose MOC result enters REMAP's window in
    // this iteration (J = q-3) and B_n of the cell REMAP updates (q-5), each
    // loaded one iteration ahead
    double csp[8], cbn;
    {
      const int cs = cell(ps - 3), cb = cell(ps - 5);
#pragma unroll
      for (int f = 0; f < 8; ++f) csp[f] = sp[f][cs];
      cbn = A.bcc[D][cb];
    }
    for (int t = ps; t <= pe + 1; ++t) {
      __syncthreads();
      // the producer's iteration whose cell J = q-3 the slot holds
      const int q = t - 1;
      if (q < ps) continue;
      double nsp[8];
      {
        const int cs = cell(q - 2);
#pragma unroll
        for (int f = 0; f < 8; ++f) nsp[f] = sp[f][cs];
      }
      const double nbn = A.bcc[D][cell(q - 4)];
      const int J = q - 3;
      if (J >= m0) {
#pragma unroll
        for (int v = 0; v < 7; ++v) Q3[v] = hand[q & 1][v][r4][lane];
        const double dxl0 = hand[q & 1][7][r4][lane];
        W3 = dxl0;
        V3 = hand[q & 1][8][r4][lane];
        // Lagrangian species of cell J (M3:1466) from the pre-sweep ones
        if constexpr (Ar<R>::strict) {
#pragma unroll
          for (int f = 0; f < 8; ++f) Q3[7 + f] = dxu * csp[f] / dxl0;
        } else {
          const double cmp = dxu * (1.0 / dxl0);
#pragma unroll
          for (int f = 0; f < 8; ++f) Q3[7 + f] = csp[f] * cmp;
        }
        // REMAP of face Jf = J-2 (cells Jf-1 .. Jf+2 in Q0 .. Q3) and of
        // cell Jf (faces Jf-1 and Jf)
        const int Jf = J - 2;
        if (Jf >= p0 - 1) {
          const double vadt = V1 * dt;
          const bool up = !(vadt <= 0.0);
          const double u4 = up ? Q1[4] : Q2[4], u0 = up ? Q1[0] : Q2[0];
          const double depend = up ? 0.5 * (W1 - vadt) * A.facdep : 0.5 * (W2 + vadt) * A.facdep;
          const R dxm2 = (R)(up ? W0 : W1), dx02 = (R)(up ? W1 : W2), dxp2 = (R)(up ? W2 : W3);
          const bool mmr = use_minmod_r<R>(u4, u0);
          const R rdep = (R)depend;
          double F[15];
#pragma unroll
          for (int v = 0; v < 15; ++v) {
            const R gr = vslope<R>(up ? Q0[v] : Q1[v], up ? Q1[v] : Q2[v], up ? Q2[v] : Q3[v], dxm2, dx02, dxp2, mmr);
            const double qu = up ? Q1[v] : Q2[v];
            double face;
            if constexpr (Ar<R>::strict) face = up ? qu + depend * gr : qu - depend * gr;
            else face = up ? qu + (double)(rdep * gr) : qu - (double)(rdep * gr);
            F[v] = -(vadt * face);
          }
          if (Jf >= p0 && act) {
            // remapped conserved state (M3:1645-1686), primitives (M3:1291-1321)
            const double dxE = xe[Jf + 1] - xe[Jf];
            double Qn[15];
#pragma unroll
            for (int v = 0; v < 15; ++v) {
              if constexpr (Ar<R>::strict) Qn[v] = (Q1[v] * W1 + F[v] - Fp[v]) / dxE;
              else Qn[v] = (Q1[v] * W1 + F[v] - Fp[v]) * (1.0 / dxE);
            }
            const int c = at(Jf);
            const double bn = cbn;
            double rho = dmax(1.27 * A.rmin, Qn[0]);
            double invd = 1.0 / rho;
            double vx = Qn[1] * invd, vy = Qn[2] * invd, vz = Qn[3] * invd;
            double s4 = Qn[7] + Qn[8] + Qn[10] + Qn[11];
            double gg = 3.0 * s4 + 5.0 * Qn[9];
            double gi = (2.0 * s4 + 2.0 * Qn[9]) / gg;
            double ekin = 0.5 * (vx * vx + vy * vy + vz * vz) * rho;
            double emag = 0.5 * (bn * bn + Qn[5] * Qn[5] + Qn[6] * Qn[6]);
            double pr = (Qn[4] - ekin - emag) * gi;
            pr = dmax(A.pmin, pr);
            pr = dmin(A.pmax, pr);
            A.Uo[0][c] = rho;
            A.Uo[1][c] = vx;
            A.Uo[2][c] = vy;
            A.Uo[3][c] = vz;
            A.Uo[4][c] = pr;
#pragma unroll
            for (int f = 0; f < 8; ++f) A.spo[f][c] = dmax(A.spmin[f], Qn[7 + f]);
          }
#pragma unroll
          for (int v = 0; v < 15; ++v) Fp[v] = F[v];
        }
#pragma unroll
        for (int v = 0; v < 15; ++v) { Q0[v] = Q1[v]; Q1[v] = Q2[v]; Q2[v] = Q3[v]; }
        W0 = W1; W1 = W2; W2 = W3;
        V1 = V2; V2 = V3;
      }
#pragma unroll
      for (int f = 0; f < 8; ++f) csp[f] = nsp[f];
      cbn = nbn;
    }
  }
}

// ---------------------------------------------------------------------------
// RHS + MOC fused along x
// ---------------------------------------------------------------------------
// A wave holds 64 consecutive positions p = c0-4 .. c0+59 of one x row (one
// per lane, loads coalesced) and produces the MOC results of the 56 cells
// c0 .. c0+55: neighbour values move between lanes by shuffles, so the RHS
// results never go to HBM.  Stencil chain per lane: RHS states of p (inputs
// p-1..p+1), Riemann interface p-1|p, Lagrangian update of p (interfaces
// p-1|p and p|p+1), MOC slope of p, characteristic interface p|p+1 and the
// MOC update of p; each step needs one more lane on each side, which the
// four halo lanes provide.
//
// Positions outside 1..n: on a periodic side every value is that of the
// periodic image (spacing uniform, checked by the host), with the image's
// positions x for its Lagrangian update; on an inflow side the RHS inputs
// are the filled ghost cells and the RHS results MOC reads there are the
// MOC stage's ghost rules (rho, v_t and B_n as filled; Blg and dx_lag the
// side's constants).  The MOC limiter switch and block-edge views are
// those of k_moc.  The Lagrangian species go to HBM for REMAP.
template <class R> __device__ __forceinline__ R sh_up(R v) { return __shfl_up(v, 1u, 64); }
template <class R> __device__ __forceinline__ R sh_dn(R v) { return __shfl_down(v, 1u, 64); }

template <class R>
__global__ __launch_bounds__(256) void k_rhsmoc_x(SweepIO A, Geo g, XSide S, double dt) {
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int lane = threadIdx.x & 63;
  const int n = g.n[0];
  const int j = 1 + 4 * by + (int)(threadIdx.x >> 6), k = 1 + bz;
  if (j > g.n[1]) return; // whole waves
  const int c0 = 1 + XCELLS * bx;
  const int p = c0 - XHALO + lane;
  // position whose values the lane holds, and the one of its geometry
  int m = p;
  if (p < 1) m = S.inflow[0] ? p : p + n;
  else if (p > n) m = S.inflow[1] ? p : p - n;
  m = min(max(m, -1), n + 2);
  const int mg = min(max(m, 0), n + 1);
  const int c = ix3s(g, m, j, k);
  const double *const *sp = A.sp;
  const double rho = A.U[0][c], vn = A.U[1][c], pp = A.U[4][c];
  const double b1 = A.bcc[1][c], b2 = A.bcc[2][c], bnc = A.bcc[0][c];
  const double vt1 = A.U[2][c], vt2 = A.U[3][c];
  double spc[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) spc[f] = sp[f][c];
  const double *dxe = A.dxe, *xe = A.xe;
  // RHS states of p (M3:1380-1406)
  R q[8];
  rhs_states<R>(rho, vn, pp, b1, b2, sh_up(rho), sh_dn(rho), sh_up(vn), sh_dn(vn), sh_up(pp), sh_dn(pp), sh_up(b1),
                sh_dn(b1), sh_up(b2), sh_dn(b2), spc, dxe[mg], dxe[mg + 1], dxe[mg + 2], dt, A.facdep, q);
  // Riemann interface p-1|p
  R ql[8];
#pragma unroll
  for (int v = 0; v < 4; ++v) ql[v] = sh_up(q[v]);
#pragma unroll
  for (int v = 4; v < 8; ++v) ql[v] = q[v];
  double spl[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) spl[f] = sh_up(spc[f]);
  R Pa, Va;
  rhs_iface<R>(spl, spc, ql, q, Pa, Va);
  const R Pr = sh_dn(Pa), Vr = sh_dn(Va);
  // Lagrangian update of p (M3:1448-1466); interfaces p-1|p (Pa, Va) and
  // p|p+1 (Pr, Vr)
  const double xq = xe[mg + 1], xqm = xe[mg], dxq = dxe[mg + 1];
  const double dm = dxq * rho;
  const double pa = (double)Pr, pam = (double)Pa, va = (double)Vr, vam = (double)Va;
  const double xl = xq + va * dt, xlm = xqm + vam * dt;
  double dxl = xl - xlm;
  double rhoL, vnL, eL, bl1, bl2, spL[8];
  {
    const double vx = vn, vy = vt1, vz = vt2, bx_ = bnc, by_ = b1, bz_ = b2;
    if constexpr (Ar<R>::strict) {
      double s4 = spc[0] + spc[1] + spc[3] + spc[4];
      double gg = 3.0 * s4 + 5.0 * spc[2];
      double gi = (2.0 * s4 + 2.0 * spc[2]) / gg;
      double e = pp / (rho * gi) + 0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx_ * bx_ + by_ * by_ + bz_ * bz_) / rho;
      rhoL = dm / dxl;
      vnL = vn - dt / dm * (pa - pam);
      eL = e - dt / dm * (pa * va - pam * vam);
      bl1 = b1 * dxq / dxl;
      bl2 = b2 * dxq / dxl;
#pragma unroll
      for (int f = 0; f < 8; ++f) spL[f] = dxq * spc[f] / dxl;
    } else {
      double s4 = spc[0] + spc[1] + spc[3] + spc[4];
      double irho = 1.0 / rho;
      double e = pp * (3.0 * s4 + 5.0 * spc[2]) / (2.0 * s4 + 2.0 * spc[2]) * irho +
                 0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx_ * bx_ + by_ * by_ + bz_ * bz_) * irho;
      double idl = 1.0 / dxl, dtdm = dt / dm, cmp = dxq * idl;
      rhoL = dm * idl;
      vnL = vn - dtdm * (pa - pam);
      eL = e - dtdm * (pa * va - pam * vam);
      bl1 = b1 * cmp;
      bl2 = b2 * cmp;
#pragma unroll
      for (int f = 0; f < 8; ++f) spL[f] = spc[f] * cmp;
    }
  }
  // inflow ghosts as the MOC stage's fills leave them
  const int side = p < 1 ? 0 : 1;
  if ((p < 1 || p > n) && S.inflow[side]) {
    rhoL = rho;
    bl1 = S.blg[side][0];
    bl2 = S.blg[side][1];
    dxl = S.dxl[side];
  }
  const bool own = lane >= XHALO && lane < XHALO + XCELLS && p <= n;
  if (own) {
    A.Uo[0][c] = rhoL;
#pragma unroll
    for (int f = 0; f < 8; ++f) A.spo[f][c] = spL[f];
    A.dxlagM[c] = dxl;
  }
  // interface p|p+1 velocity for REMAP (interfaces c0..c0+55, and 0)
  if (own || (p == 0 && c0 == 1)) A.va[ix3s(g, p, j, k)] = va;
  // MOC slope of p (M3:2011-2020)
  const bool mme = use_minmod_r<R>(eL, rhoL), mmp = use_minmod_r<R>(pp, rhoL);
  const bool inner = p >= 1 && p <= n;
  const bool mm = inner ? mme : mmp;
  double v0[6] = {rhoL, vt1, vt2, bl1, bl2, dxl};
  double vm[6], vp[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) { vm[v] = sh_up(v0[v]); vp[v] = sh_dn(v0[v]); }
  const R dxm_r = (R)vm[5], dx0_r = (R)dxl, dxp_r = (R)vp[5];
  R G[5], H[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) { G[v] = vslope<R>(vm[v], v0[v], vp[v], dxm_r, dx0_r, dxp_r, mm); H[v] = G[v]; }
  const int nblk = A.nblk;
  if ((p % nblk == 0 || (p - 1) % nblk == 0) && mmp != mm) {
#pragma unroll
    for (int v = 0; v < 5; ++v) H[v] = vslope<R>(vm[v], v0[v], vp[v], dxm_r, dx0_r, dxp_r, mmp);
  }
  // characteristic interface J = p between p and p+1 (local positions 0..n)
  R G1[5], H1[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) { G1[v] = sh_dn(G[v]); H1[v] = sh_dn(H[v]); }
  const int fl = (mme ? 1 : 0) | (mmp ? 2 : 0);
  const int fl1 = __shfl_down(fl, 1u, 64);
  const int J = min(max(p, 0), n + 1);
  const double fown = A.U[5][ix3s(g, min(max(p, -1), n + 2), j, k)];
  const double f1 = sh_dn(fown), f2 = sh_dn(f1);
  const double bnJ = moc_bn(fown, f1, f2, xe[J + 1] + va * dt, xe, J);
  MocV<R> L, Rv;
  if (J % nblk != 0) {
    L = moc_iface_r<R>(v0, vp, G, G1, bnJ, A.facdep, dt);
    Rv = L;
  } else {
    const bool mm0p = mmp, mm0e = J >= 1 ? mme : mm0p;
    const bool mm1p = (fl1 & 2) != 0, mm1e = J + 1 <= n ? (fl1 & 1) != 0 : mm1p;
    L = moc_iface_r<R>(v0, vp, G, mm1p != mm1e ? H1 : G1, bnJ, A.facdep, dt);
    if (mm0p != mm0e) Rv = moc_iface_r<R>(v0, vp, H, G1, bnJ, A.facdep, dt);
    else if (mm1p != mm1e) Rv = moc_iface_r<R>(v0, vp, G, G1, bnJ, A.facdep, dt);
    else Rv = L;
  }
  // MOC update of p (M3:2053-2095): left view of interface p, right view
  // of interface p-1
  MocV<R> rvp;
  rvp.Bay = sh_up(Rv.Bay); rvp.Vay = sh_up(Rv.Vay); rvp.Baz = sh_up(Rv.Baz); rvp.Vaz = sh_up(Rv.Vaz);
  const double bnm = sh_up(bnJ);
  if (!own) return;
  const double Bnap = bnJ, Bnam = bnm, Vnap = va, Vnam = vam;
  const double BayL = (double)L.Bay, VayL = (double)L.Vay, BazL = (double)L.Baz, VazL = (double)L.Vaz;
  const double BayR = (double)rvp.Bay, VayR = (double)rvp.Vay, BazR = (double)rvp.Baz, VazR = (double)rvp.Vaz;
  if constexpr (Ar<R>::strict) {
    A.Uo[1][c] = vnL * rhoL + dt / dxl * 0.5 * (Bnap * Bnap - Bnam * Bnam);
    A.Uo[2][c] = vt1 * rhoL + dt / dxl * (Bnap * BayL - Bnam * BayR);
    A.Uo[3][c] = vt2 * rhoL + dt / dxl * (Bnap * BazL - Bnam * BazR);
    A.Uo[4][c] = eL * rhoL + dt / dxl *
                                ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                 (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
    A.blgo[0][c] = bl1 + dt / dxl * (Bnap * VayL - Bnam * VayR);
    A.blgo[1][c] = bl2 + dt / dxl * (Bnap * VazL - Bnam * VazR);
  } else {
    const double fq = dt / dxl;
    A.Uo[1][c] = vnL * rhoL + fq * 0.5 * (Bnap * Bnap - Bnam * Bnam);
    A.Uo[2][c] = vt1 * rhoL + fq * (Bnap * BayL - Bnam * BayR);
    A.Uo[3][c] = vt2 * rhoL + fq * (Bnap * BazL - Bnam * BazR);
    A.Uo[4][c] = eL * rhoL + fq * ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                   (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
    A.blgo[0][c] = bl1 + fq * (Bnap * VayL - Bnam * VayR);
    A.blgo[1][c] = bl2 + fq * (Bnap * VazL - Bnam * VazR);
  }
}

// The whole x sweep in one kernel when both x sides are inflow: the
// k_rhsmoc_x chain, then REMAP from the MOC results of lanes p-2..p+2 held
// in registers (the inflow ghosts take the REMAP stage's fill rules), so no
// Lagrangian or MOC result goes to HBM and no fill runs inside the sweep.
// Halo 6: a wave updates the 52 cells c0..c0+51.
template <class R>
__global__ __launch_bounds__(256, Ar<R>::strict ? 3 : 4) void k_sweep_x(SweepIO A, Geo g, XSide S, double dt) {
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int lane = threadIdx.x & 63;
  const int n = g.n[0];
  const int j = 1 + 2 * by + (int)(threadIdx.x >> 6), k = 1 + bz;
  if (j > g.n[1]) return; // whole waves
  const int c0 = 1 + XSCELLS * bx;
  const int p = c0 - 6 + lane;
  // position whose values the lane holds, and the one of its geometry
  int m = p;
  if (p < 1) m = S.inflow[0] ? p : p + n;
  else if (p > n) m = S.inflow[1] ? p : p - n;
  m = min(max(m, -1), n + 2);
  const int mg = min(max(m, 0), n + 1);
  const int c = ix3s(g, m, j, k);
  const double *const *sp = A.sp;
  const double rho = A.U[0][c], vn = A.U[1][c], pp = A.U[4][c];
  const double b1 = A.bcc[1][c], b2 = A.bcc[2][c], bnc = A.bcc[0][c];
  const double vt1 = A.U[2][c], vt2 = A.U[3][c];
  double spc[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) spc[f] = sp[f][c];
  const double *dxe = A.dxe, *xe = A.xe;
  // RHS states of p (M3:1380-1406)
  R q[8];
  rhs_states<R>(rho, vn, pp, b1, b2, sh_up(rho), sh_dn(rho), sh_up(vn), sh_dn(vn), sh_up(pp), sh_dn(pp), sh_up(b1),
                sh_dn(b1), sh_up(b2), sh_dn(b2), spc, dxe[mg], dxe[mg + 1], dxe[mg + 2], dt, A.facdep, q);
  // Riemann interface p-1|p
  R ql[8];
#pragma unroll
  for (int v = 0; v < 4; ++v) ql[v] = sh_up(q[v]);
#pragma unroll
  for (int v = 4; v < 8; ++v) ql[v] = q[v];
  double spl[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) spl[f] = sh_up(spc[f]);
  R Pa, Va;
  rhs_iface<R>(spl, spc, ql, q, Pa, Va);
  const R Pr = sh_dn(Pa), Vr = sh_dn(Va);
  // Lagrangian update of p (M3:1448-1466); interfaces p-1|p (Pa, Va) and
  // p|p+1 (Pr, Vr)
  const double xq = xe[mg + 1], xqm = xe[mg], dxq = dxe[mg + 1];
  const double dm = dxq * rho;
  const double pa = (double)Pr, pam = (double)Pa, va = (double)Vr, vam = (double)Va;
  const double xl = xq + va * dt, xlm = xqm + vam * dt;
  double dxl = xl - xlm;
  double rhoL, vnL, eL, bl1, bl2, spL[8];
  {
    const double vx = vn, vy = vt1, vz = vt2, bx_ = bnc, by_ = b1, bz_ = b2;
    if constexpr (Ar<R>::strict) {
      double s4 = spc[0] + spc[1] + spc[3] + spc[4];
      double gg = 3.0 * s4 + 5.0 * spc[2];
      double gi = (2.0 * s4 + 2.0 * spc[2]) / gg;
      double e = pp / (rho * gi) + 0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx_ * bx_ + by_ * by_ + bz_ * bz_) / rho;
      rhoL = dm / dxl;
      vnL = vn - dt / dm * (pa - pam);
      eL = e - dt / dm * (pa * va - pam * vam);
      bl1 = b1 * dxq / dxl;
      bl2 = b2 * dxq / dxl;
#pragma unroll
      for (int f = 0; f < 8; ++f) spL[f] = dxq * spc[f] / dxl;
    } else {
      double s4 = spc[0] + spc[1] + spc[3] + spc[4];
      double irho = 1.0 / rho;
      double e = pp * (3.0 * s4 + 5.0 * spc[2]) / (2.0 * s4 + 2.0 * spc[2]) * irho +
                 0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx_ * bx_ + by_ * by_ + bz_ * bz_) * irho;
      double idl = 1.0 / dxl, dtdm = dt / dm, cmp = dxq * idl;
      rhoL = dm * idl;
      vnL = vn - dtdm * (pa - pam);
      eL = e - dtdm * (pa * va - pam * vam);
      bl1 = b1 * cmp;
      bl2 = b2 * cmp;
#pragma unroll
      for (int f = 0; f < 8; ++f) spL[f] = spc[f] * cmp;
    }
  }
  // inflow ghosts as the MOC stage's fills leave them
  const int side = p < 1 ? 0 : 1;
  if ((p < 1 || p > n) && S.inflow[side]) {
    rhoL = rho;
    bl1 = S.blg[side][0];
    bl2 = S.blg[side][1];
    dxl = S.dxl[side];
  }
  // MOC slope of p (M3:2011-2020)
  const bool mme = use_minmod_r<R>(eL, rhoL), mmp = use_minmod_r<R>(pp, rhoL);
  const bool inner = p >= 1 && p <= n;
  const bool mm = inner ? mme : mmp;
  double v0[6] = {rhoL, vt1, vt2, bl1, bl2, dxl};
  double vm[6], vp[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) { vm[v] = sh_up(v0[v]); vp[v] = sh_dn(v0[v]); }
  const R dxm_r = (R)vm[5], dx0_r = (R)dxl, dxp_r = (R)vp[5];
  R G[5], H[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) { G[v] = vslope<R>(vm[v], v0[v], vp[v], dxm_r, dx0_r, dxp_r, mm); H[v] = G[v]; }
  const int nblk = A.nblk;
  if ((p % nblk == 0 || (p - 1) % nblk == 0) && mmp != mm) {
#pragma unroll
    for (int v = 0; v < 5; ++v) H[v] = vslope<R>(vm[v], v0[v], vp[v], dxm_r, dx0_r, dxp_r, mmp);
  }
  // characteristic interface J = p between p and p+1 (local positions 0..n)
  R G1[5], H1[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) { G1[v] = sh_dn(G[v]); H1[v] = sh_dn(H[v]); }
  const int fl = (mme ? 1 : 0) | (mmp ? 2 : 0);
  const int fl1 = __shfl_down(fl, 1u, 64);
  const int J = min(max(p, 0), n + 1);
  const double fown = A.U[5][ix3s(g, min(max(p, -1), n + 2), j, k)];
  const double f1 = sh_dn(fown), f2 = sh_dn(f1);
  const double bnJ = moc_bn(fown, f1, f2, xe[J + 1] + va * dt, xe, J);
  MocV<R> L, Rv;
  if (J % nblk != 0) {
    L = moc_iface_r<R>(v0, vp, G, G1, bnJ, A.facdep, dt);
    Rv = L;
  } else {
    const bool mm0p = mmp, mm0e = J >= 1 ? mme : mm0p;
    const bool mm1p = (fl1 & 2) != 0, mm1e = J + 1 <= n ? (fl1 & 1) != 0 : mm1p;
    L = moc_iface_r<R>(v0, vp, G, mm1p != mm1e ? H1 : G1, bnJ, A.facdep, dt);
    if (mm0p != mm0e) Rv = moc_iface_r<R>(v0, vp, H, G1, bnJ, A.facdep, dt);
    else if (mm1p != mm1e) Rv = moc_iface_r<R>(v0, vp, G, G1, bnJ, A.facdep, dt);
    else Rv = L;
  }
  // MOC update of p (M3:2053-2095): left view of interface p, right view
  // of interface p-1
  MocV<R> rvp;
  rvp.Bay = sh_up(Rv.Bay); rvp.Vay = sh_up(Rv.Vay); rvp.Baz = sh_up(Rv.Baz); rvp.Vaz = sh_up(Rv.Vaz);
  const double bnm = sh_up(bnJ);
  const double Bnap = bnJ, Bnam = bnm, Vnap = va, Vnam = vam;
  const double BayL = (double)L.Bay, VayL = (double)L.Vay, BazL = (double)L.Baz, VazL = (double)L.Vaz;
  const double BayR = (double)rvp.Bay, VayR = (double)rvp.Vay, BazR = (double)rvp.Baz, VazR = (double)rvp.Vaz;
  double Q[15];
  if constexpr (Ar<R>::strict) {
    Q[1] = vnL * rhoL + dt / dxl * 0.5 * (Bnap * Bnap - Bnam * Bnam);
    Q[2] = vt1 * rhoL + dt / dxl * (Bnap * BayL - Bnam * BayR);
    Q[3] = vt2 * rhoL + dt / dxl * (Bnap * BazL - Bnam * BazR);
    Q[4] = eL * rhoL + dt / dxl *
                          ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                           (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
    Q[5] = bl1 + dt / dxl * (Bnap * VayL - Bnam * VayR);
    Q[6] = bl2 + dt / dxl * (Bnap * VazL - Bnam * VazR);
  } else {
    const double fq = dt / dxl;
    Q[1] = vnL * rhoL + fq * 0.5 * (Bnap * Bnap - Bnam * Bnam);
    Q[2] = vt1 * rhoL + fq * (Bnap * BayL - Bnam * BayR);
    Q[3] = vt2 * rhoL + fq * (Bnap * BazL - Bnam * BazR);
    Q[4] = eL * rhoL + fq * ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                             (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
    Q[5] = bl1 + fq * (Bnap * VayL - Bnam * VayR);
    Q[6] = bl2 + fq * (Bnap * VazL - Bnam * VazR);
  }
  Q[0] = rhoL;
#pragma unroll
  for (int f = 0; f < 8; ++f) Q[7 + f] = spL[f];
  // inflow ghosts as the REMAP stage's fills leave them (M3:818-884)
  if (p < 1 || p > n) {
    const double v = S.vin[side];
    Q[0] = rho;
    Q[1] = rho * v;
    Q[2] = S.mt[side][0];
    Q[3] = S.mt[side][1];
    Q[4] = S.ein[side] * 1.5 + 0.5 * (rho * (v * v) + bnc * bnc + bl1 * bl1 + bl2 * bl2);
    Q[5] = bl1;
    Q[6] = bl2;
#pragma unroll
    for (int f = 0; f < 8; ++f) Q[7 + f] = S.spin[f] * rho;
  }
  // REMAP of cell p (M3:1600-1686): upwind flux through face p|p+1 from
  // cells p-1..p+2, the flux of face p-1|p from the lane below
  const double W1 = dxl;
  const double W0 = sh_up(W1), W2 = sh_dn(W1), W3 = sh_dn(W2);
  const double vadt = va * dt;
  const bool up = !(vadt <= 0.0);
  const double u4n = sh_dn(Q[4]), u0n = sh_dn(Q[0]);
  const double u4 = up ? Q[4] : u4n, u0 = up ? Q[0] : u0n;
  const double depend = up ? 0.5 * (W1 - vadt) * A.facdep : 0.5 * (W2 + vadt) * A.facdep;
  const R rdxm = (R)(up ? W0 : W1), rdx0 = (R)(up ? W1 : W2), rdxp = (R)(up ? W2 : W3);
  const bool rmm = use_minmod_r<R>(u4, u0);
  const R rdep = (R)depend;
  const bool own = lane >= 6 && lane < 6 + XSCELLS && p <= n;
  const double dxE = own ? xe[p + 1] - xe[p] : 1.0;
  double Qn[15];
#pragma unroll
  for (int v = 0; v < 15; ++v) {
    const double q0 = sh_up(Q[v]), q1 = Q[v], q2 = sh_dn(Q[v]), q3 = sh_dn(q2);
    const R gr = vslope<R>(up ? q0 : q1, up ? q1 : q2, up ? q2 : q3, rdxm, rdx0, rdxp, rmm);
    const double qu = up ? q1 : q2;
    double face;
    if constexpr (Ar<R>::strict) face = up ? qu + depend * gr : qu - depend * gr;
    else face = up ? qu + (double)(rdep * gr) : qu - (double)(rdep * gr);
    const double F = -(vadt * face);
    const double Fl = sh_up(F);
    if constexpr (Ar<R>::strict) Qn[v] = (q1 * W1 + F - Fl) / dxE;
    else Qn[v] = (q1 * W1 + F - Fl) * (1.0 / dxE);
  }
  if (!own) return;
  // primitives (M3:1291-1321)
  double rr = dmax(1.27 * A.rmin, Qn[0]);
  double invd = 1.0 / rr;
  double vx = Qn[1] * invd, vy = Qn[2] * invd, vz = Qn[3] * invd;
  double s4 = Qn[7] + Qn[8] + Qn[10] + Qn[11];
  double gg = 3.0 * s4 + 5.0 * Qn[9];
  double gi = (2.0 * s4 + 2.0 * Qn[9]) / gg;
  double ekin = 0.5 * (vx * vx + vy * vy + vz * vz) * rr;
  double emag = 0.5 * (bnc * bnc + Qn[5] * Qn[5] + Qn[6] * Qn[6]);
  double pr = (Qn[4] - ekin - emag) * gi;
  pr = dmax(A.pmin, pr);
  pr = dmin(A.pmax, pr);
  A.Uo[0][c] = rr;
  A.Uo[1][c] = vx;
  A.Uo[2][c] = vy;
  A.Uo[3][c] = vz;
  A.Uo[4][c] = pr;
#pragma unroll
  for (int f = 0; f < 8; ++f) A.spo[f][c] = dmax(A.spmin[f], Qn[7 + f]);
}

// REMAP + primitives along x with the layout of k_rhsmoc_x: a wave holds
// positions c0-2 .. c0+61 of one row and updates cells c0 .. c0+59.  Lane p
// forms the upwind flux through face p|p+1 from cells p-1..p+2 (neighbours
// by shuffles) and updates cell p with the flux of face p-1|p from the lane
// below.  Ghost cells are read as the REMAP stage's fills leave them.
template <class R>
__global__ __launch_bounds__(256) void k_remap_x(SweepIO A, Geo g, double dt) {
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int lane = threadIdx.x & 63;
  const int n = g.n[0];
  const int j = 1 + 4 * by + (int)(threadIdx.x >> 6), k = 1 + bz;
  if (j > g.n[1]) return; // whole waves
  const int c0 = 1 + XRCELLS * bx;
  const int p = c0 - 2 + lane;
  const int c = ix3s(g, min(max(p, -1), n + 2), j, k);
  double Q[15];
#pragma unroll
  for (int v = 0; v < 15; ++v) Q[v] = remap_field(A, v)[c];
  const double W1 = A.dxlagM[c];
  const double W0 = sh_up(W1), W2 = sh_dn(W1), W3 = sh_dn(W2);
  // upwind linear flux through face p|p+1 (M3:1600-1640)
  const double vadt = A.va[ix3s(g, min(max(p, 0), n), j, k)] * dt;
  const bool up = !(vadt <= 0.0);
  const double u4n = sh_dn(Q[4]), u0n = sh_dn(Q[0]);
  const double u4 = up ? Q[4] : u4n, u0 = up ? Q[0] : u0n;
  const double depend = up ? 0.5 * (W1 - vadt) * A.facdep : 0.5 * (W2 + vadt) * A.facdep;
  const R dxm = (R)(up ? W0 : W1), dx0 = (R)(up ? W1 : W2), dxp = (R)(up ? W2 : W3);
  const bool mm = use_minmod_r<R>(u4, u0);
  const R rdep = (R)depend;
  const bool own = lane >= 2 && lane < 2 + XRCELLS && p <= n;
  const double dxE = own ? A.xe[p + 1] - A.xe[p] : 1.0;
  double Qn[15];
#pragma unroll
  for (int v = 0; v < 15; ++v) {
    const double q0 = sh_up(Q[v]), q1 = Q[v], q2 = sh_dn(Q[v]), q3 = sh_dn(q2);
    const R gr = vslope<R>(up ? q0 : q1, up ? q1 : q2, up ? q2 : q3, dxm, dx0, dxp, mm);
    const double qu = up ? q1 : q2;
    double face;
    if constexpr (Ar<R>::strict) face = up ? qu + depend * gr : qu - depend * gr;
    else face = up ? qu + (double)(rdep * gr) : qu - (double)(rdep * gr);
    const double F = -(vadt * face);
    const double Fl = sh_up(F);
    // remapped conserved state (M3:1645-1686)
    if constexpr (Ar<R>::strict) Qn[v] = (q1 * W1 + F - Fl) / dxE;
    else Qn[v] = (q1 * W1 + F - Fl) * (1.0 / dxE);
  }
  if (!own) return;
  // primitives (M3:1291-1321)
  const double bn = A.bcc[0][c];
  double rho = dmax(1.27 * A.rmin, Qn[0]);
  double invd = 1.0 / rho;
  double vx = Qn[1] * invd, vy = Qn[2] * invd, vz = Qn[3] * invd;
  double s4 = Qn[7] + Qn[8] + Qn[10] + Qn[11];
  double gg = 3.0 * s4 + 5.0 * Qn[9];
  double gi = (2.0 * s4 + 2.0 * Qn[9]) / gg;
  double ekin = 0.5 * (vx * vx + vy * vy + vz * vz) * rho;
  double emag = 0.5 * (bn * bn + Qn[5] * Qn[5] + Qn[6] * Qn[6]);
  double pr = (Qn[4] - ekin - emag) * gi;
  pr = dmax(A.pmin, pr);
  pr = dmin(A.pmax, pr);
  A.Uo[0][c] = rho;
  A.Uo[1][c] = vx;
  A.Uo[2][c] = vy;
  A.Uo[3][c] = vz;
  A.Uo[4][c] = pr;
#pragma unroll
  for (int f = 0; f < 8; ++f) A.spo[f][c] = dmax(A.spmin[f], Qn[7 + f]);
}

template __global__ void k_rhsmoc_col<1, double>(SweepIO, Geo, double);
template __global__ void k_rhsmoc_col<2, double>(SweepIO, Geo, double);
template __global__ void k_rhsmoc_col<1, float>(SweepIO, Geo, double);
template __global__ void k_rhsmoc_col<2, float>(SweepIO, Geo, double);
template __global__ void k_sweep_col<1, double>(SweepIO, Geo, double, int);
template __global__ void k_sweep_col<2, double>(SweepIO, Geo, double, int);
template __global__ void k_sweep_col2<1, double>(SweepIO, Geo, double, int);
template __global__ void k_sweep_col2<2, double>(SweepIO, Geo, double, int);
template __global__ void k_sweep_col2<1, float>(SweepIO, Geo, double, int);
template __global__ void k_sweep_col2<2, float>(SweepIO, Geo, double, int);
template __global__ void k_remap_col<1, double>(SweepIO, Geo, double);
template __global__ void k_remap_col<2, double>(SweepIO, Geo, double);
template __global__ void k_remap_col<1, float>(SweepIO, Geo, double);
template __global__ void k_remap_col<2, float>(SweepIO, Geo, double);
template __global__ void k_rhsmoc_x<double>(SweepIO, Geo, XSide, double);
template __global__ void k_rhsmoc_x<float>(SweepIO, Geo, XSide, double);
template __global__ void k_remap_x<double>(SweepIO, Geo, double);
template __global__ void k_remap_x<float>(SweepIO, Geo, double);
template __global__ void k_sweep_x<double>(SweepIO, Geo, XSide, double);
template __global__ void k_sweep_x<float>(SweepIO, Geo, XSide, double);

#define MCF_INST(D, R)                                                                                  \
  template __global__ void k_rhs<D, R>(SweepIO, Geo, double);                                          \
  template __global__ void k_moc<D, R>(SweepIO, Geo, double);                                          \
  template __global__ void k_remap<D, R>(SweepIO, Geo, double);
MCF_INST(0, double)
MCF_INST(1, double)
MCF_INST(2, double)
MCF_INST(0, float)
MCF_INST(1, float)
MCF_INST(2, float)
#undef MCF_INST

} // namespace mcf
