// mcf_sweep.hip -- one MHD sweep (M3:1264-1324) as three fused stage
// kernels: RHS (M3:1330-1467), MOC (M3:1939-2098), REMAP + primitive
// recovery (M3:1556-1686, 1291-1321).
//
// Each kernel computes the interface values of a tile once into LDS and
// then updates the tile's cells from LDS, so no per-interface array goes to
// HBM.  A block is 576 threads: waves 0..7 own a tile of 8 lines x 64 cells
// and the interface to the right of each cell; wave 8 computes the one
// interface left of each line.
//   x sweep: a line is 64 consecutive x cells (one wave), 8 lines = 8 y rows.
//   y/z sweep: a wave is 64 consecutive x cells at one position along the
//   sweep axis, 8 waves = 8 positions; loads stay coalesced along x.
// Stage outputs go to the other buffer of each written field (the host
// swaps the pointers), so a tile never reads cells another tile has
// already updated.
#include "mcf_kernels.h"

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

constexpr int LW = 65;        // LDS row pitch (x layout: 65 interfaces per line)
constexpr int NSLOT = 9 * 65; // LDS interface slots per value

// Thread roles of a 576-thread tile.
struct Tile {
  int i, j, k;   // cell of a main thread / left cell of the extra interface
  int q;         // position of that cell along the sweep axis
  int fslot;     // LDS slot of the interface computed by this thread
  int lslot, rslot;
  bool face_ok;  // this thread computes an interface (position 0..N)
  bool cell_ok;  // this thread updates a cell (main waves, inside 1..N)
  bool own;      // interface owned by this tile (not the left halo of a later tile)
};

template <int D>
__device__ __forceinline__ Tile tile_map(const Geo &g) {
  Tile t;
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  if (D == 0) {
    const int s0 = 1 + 64 * (int)blockIdx.x;
    const int jr = 1 + 8 * (int)blockIdx.y;
    const int kk = 1 + (int)blockIdx.z;
    if (w < 8) {
      t.i = s0 + l; t.j = jr + w; t.k = kk;
      t.fslot = w * LW + l + 1; t.lslot = w * LW + l; t.rslot = w * LW + l + 1;
      t.face_ok = t.i <= nx && t.j <= ny;
      t.cell_ok = t.face_ok;
      t.own = true;
    } else {
      t.i = s0 - 1; t.j = jr + l; t.k = kk;
      t.fslot = l * LW; t.lslot = t.rslot = 0;
      t.face_ok = l < 8 && t.j <= ny;
      t.cell_ok = false;
      t.own = (s0 - 1 == 0);
    }
    t.q = t.i;
  } else {
    const int i0 = 1 + 64 * (int)blockIdx.x;
    const int s0 = 1 + 8 * (int)blockIdx.y;
    const int tr = 1 + (int)blockIdx.z;
    const int n = (D == 1) ? ny : nz;
    const int ntr = (D == 1) ? nz : ny;
    int p;
    if (w < 8) {
      p = s0 + w;
      t.fslot = (w + 1) * 64 + l; t.lslot = w * 64 + l; t.rslot = (w + 1) * 64 + l;
      t.face_ok = p <= n && i0 + l <= nx && tr <= ntr;
      t.cell_ok = t.face_ok;
      t.own = true;
    } else {
      p = s0 - 1;
      t.fslot = l; t.lslot = t.rslot = 0;
      t.face_ok = i0 + l <= nx && tr <= ntr;
      t.cell_ok = false;
      t.own = (s0 - 1 == 0);
    }
    t.i = i0 + l;
    if (D == 1) { t.j = p; t.k = tr; } else { t.k = p; t.j = tr; }
    t.q = p;
  }
  return t;
}

} // namespace

// ---------------------------------------------------------------------------
// RHS
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(576) void k_rhs(SweepIO A, Geo g, double dt) {
  __shared__ double sPa[NSLOT], sVa[NSLOT];
  const Tile t = tile_map<D>(g);
  const int s = g.st[D];
  const double *dxe = A.dxe, *xe = A.xe;
  if (t.face_ok) {
    const int q = t.q;
    const int c = ix3s(g, t.i, t.j, t.k), cp = c + s;
    const double *U1 = A.U[0], *UVN = A.U[A.VN], *U5 = A.U[4];
    const double *B1 = A.bcc[A.BT1], *B2 = A.bcc[A.BT2];
    double dxm = dxe[q], dx0 = dxe[q + 1], dxp = dxe[q + 2], dxpp = dxe[q + 3];
    double T0 = tswitch(U5[c], U1[c]), T1 = tswitch(U5[cp], U1[cp]);
    double g1a = vlim_pick(vlim_parts(U1[c - s], U1[c], U1[cp], dxm, dx0, dxp), T0);
    double g1b = vlim_pick(vlim_parts(U1[c], U1[cp], U1[cp + s], dx0, dxp, dxpp), T1);
    double gva = vlim_pick(vlim_parts(UVN[c - s], UVN[c], UVN[cp], dxm, dx0, dxp), T0);
    double gvb = vlim_pick(vlim_parts(UVN[c], UVN[cp], UVN[cp + s], dx0, dxp, dxpp), T1);
    double g5a = vlim_pick(vlim_parts(U5[c - s], U5[c], U5[cp], dxm, dx0, dxp), T0);
    double g5b = vlim_pick(vlim_parts(U5[c], U5[cp], U5[cp + s], dx0, dxp, dxpp), T1);
    double gAa = vlim_pick(vlim_parts(B1[c - s], B1[c], B1[cp], dxm, dx0, dxp), T0);
    double gAb = vlim_pick(vlim_parts(B1[c], B1[cp], B1[cp + s], dx0, dxp, dxpp), T1);
    double gBa = vlim_pick(vlim_parts(B2[c - s], B2[c], B2[cp], dxm, dx0, dxp), T0);
    double gBb = vlim_pick(vlim_parts(B2[c], B2[cp], B2[cp + s], dx0, dxp, dxpp), T1);
    const double *const *sp = A.sp;
    double gamma = gam_of_s(sp, c);
    double cm = sqrt((gamma * U5[c] + B1[c] * B1[c] + B2[c] * B2[c]) / U1[c]);
    double depend1 = 0.5 * (dx0 - cm * dt) * A.facdep;
    depend1 = dmax(0.0, depend1);
    gamma = gam_of_s(sp, cp);
    cm = sqrt((gamma * U5[cp] + B1[cp] * B1[cp] + B2[cp] * B2[cp]) / U1[cp]);
    double depend2 = 0.5 * (dxp - cm * dt) * A.facdep;
    depend2 = dmax(0.0, depend2);
    double QL1 = posfix(U1[c] + depend1 * g1a, U1[c]);
    double QR1 = posfix(U1[cp] - depend2 * g1b, U1[cp]);
    double QL2 = UVN[c] + depend1 * gva;
    double QR2 = UVN[cp] - depend2 * gvb;
    double QL3 = posfix(U5[c] + depend1 * g5a, U5[c]);
    double QR3 = posfix(U5[cp] - depend2 * g5b, U5[cp]);
    double a1 = B1[c] + depend1 * gAa, a2 = B2[c] + depend1 * gBa;
    double QL4 = a1 * a1 + a2 * a2;
    double b1 = B1[cp] - depend2 * gAb, b2 = B2[cp] - depend2 * gBb;
    double QR4 = b1 * b1 + b2 * b2;
    double ndHm = 0.5 * (sp[0][c] + sp[0][cp]);
    double ndpm = 0.5 * (sp[1][c] + sp[1][cp]);
    double ndHem = 0.5 * (sp[3][c] + sp[3][cp]);
    double ndHepm = 0.5 * (sp[4][c] + sp[4][cp]);
    double ndH2m = 0.5 * (sp[2][c] + sp[2][cp]);
    double gm = (5.0 * (ndHm + ndpm + ndHem + ndHepm) + 7.0 * ndH2m) / (3.0 * (ndHm + ndpm + ndHem + ndHepm) + 5.0 * ndH2m);
    double Pa, Va;
    rieman(QL1, QL2, QL3, QL4, QR1, QR2, QR3, QR4, gm, Pa, Va);
    sPa[t.fslot] = Pa;
    sVa[t.fslot] = Va;
    if (t.own) {
      A.va[c] = Va;
      A.xlag[c] = xe[q + 1] + Va * dt;
    }
  }
  __syncthreads();
  if (!t.cell_ok) return;
  const int q = t.q;
  const int c = ix3s(g, t.i, t.j, t.k);
  double rho = A.U[0][c];
  double gi = gammi1_of_s(A.sp, c);
  double vx = A.U[1][c], vy = A.U[2][c], vz = A.U[3][c];
  double bx = A.bcc[0][c], by = A.bcc[1][c], bz = A.bcc[2][c];
  double e = A.U[4][c] / (rho * gi) + 0.5 * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz) / rho;
  double dxq = dxe[q + 1];
  double dm = dxq * rho;
  double pa = sPa[t.rslot], pam = sPa[t.lslot], va = sVa[t.rslot], vam = sVa[t.lslot];
  double xl = xe[q + 1] + va * dt, xlm = xe[q] + vam * dt;
  A.Uo[0][c] = dm / (xl - xlm);
  A.Uo[A.VN][c] = A.U[A.VN][c] - dt / dm * (pa - pam);
  A.u5e[c] = e - dt / dm * (pa * va - pam * vam);
  A.blgo[0][c] = A.bcc[A.BT1][c] * dxq / (xl - xlm);
  A.blgo[1][c] = A.bcc[A.BT2][c] * dxq / (xl - xlm);
#pragma unroll
  for (int f = 0; f < 8; ++f) A.spo[f][c] = dxq * A.sp[f][c] / (xl - xlm);
  A.dxlagM[c] = xl - xlm;
}

// ---------------------------------------------------------------------------
// MOC
// ---------------------------------------------------------------------------
struct MocV {
  double Bay, Vay, Baz, Vaz;
};

__device__ __forceinline__ MocV moc_iface2(const SweepIO &A, int c, int cp, int q, const double *dxl,
                                          const double *gi, const double *gp, double dt) {
  const double *U1 = A.U[0], *UT1 = A.U[A.VT1], *UT2 = A.U[A.VT2], *UBn = A.U[A.BN];
  const double *Bl1 = A.blg[0], *Bl2 = A.blg[1];
  const double fac = A.facdep;
  const double *xe = A.xe;
  int cpp = cp + (cp - c);
  double Bnap = UBn[cp] + (A.xlag[c] - xe[q + 1]) * (UBn[cpp] - UBn[c]) / (xe[q + 2] - xe[q]);
  double deni = sqrt((dxl[1] + dxl[2]) / (dxl[1] * U1[cp] + dxl[2] * U1[c]));
  double depend1, depend2, ca;
  double QL11, QL12, QL13, QR11, QR12, QR13, QL21, QL22, QR21, QR22;
  if (Bnap > 0.0) {
    ca = fabs(Bnap * deni) * 0.5 * dt;
    ca = dmin(dxl[1], ca);
    depend1 = (0.5 * dxl[1] - ca) * fac;
    ca = fabs(Bnap * deni) * 0.5 * dt;
    ca = dmin(dxl[2], ca);
    depend2 = (0.5 * dxl[2] - ca) * fac;
    QL11 = UT1[c] + depend1 * gi[1];
    QR11 = UT1[cp] - depend2 * gp[1];
    QL12 = Bl1[c] + depend1 * gi[3];
    QR12 = Bl1[cp] - depend2 * gp[3];
    QL13 = posfix(U1[c] + depend1 * gi[0], U1[c]);
    QR13 = posfix(U1[cp] - depend2 * gp[0], U1[cp]);
    QL21 = UT2[c] + depend1 * gi[2];
    QR21 = UT2[cp] - depend2 * gp[2];
    QL22 = Bl2[c] + depend1 * gi[4];
    QR22 = Bl2[cp] - depend2 * gp[4];
  } else {
    ca = fabs(Bnap * deni) * 0.5 * dt;
    ca = dmin(dxl[2], ca);
    depend1 = (0.5 * dxl[2] - ca) * fac;
    ca = fabs(Bnap * deni) * 0.5 * dt;
    ca = dmin(dxl[1], ca);
    depend2 = (0.5 * dxl[1] - ca) * fac;
    QL11 = UT1[cp] - depend1 * gp[1];
    QR11 = UT1[c] + depend2 * gi[1];
    QL12 = Bl1[cp] - depend1 * gp[3];
    QR12 = Bl1[c] + depend2 * gi[3];
    QL13 = posfix(U1[cp] - depend1 * gp[0], U1[cp]);
    QR13 = posfix(U1[c] + depend2 * gi[0], U1[c]);
    QL21 = UT2[cp] - depend1 * gp[2];
    QR21 = UT2[c] + depend2 * gi[2];
    QL22 = Bl2[cp] - depend1 * gp[4];
    QR22 = Bl2[c] + depend2 * gi[4];
  }
  MocV o;
  chareq(QL11, QL12, QL13, QR11, QR12, QR13, o.Bay, o.Vay);
  chareq(QL21, QL22, QL13, QR21, QR22, QR13, o.Baz, o.Vaz);
  return o;
}

// At a block edge of the reproduced MPI run the interface is evaluated twice
// (left-block view: the right cell's limiter switch reads the pre-sweep
// pressure of its ghost copy; right-block view: the left cell's does).
template <int D>
__global__ __launch_bounds__(576) void k_moc(SweepIO A, Geo g, double dt) {
  __shared__ double sL[4][NSLOT], sR[4][NSLOT];
  const Tile t = tile_map<D>(g);
  const int s = g.st[D];
  const int n = g.n[D];
  if (t.face_ok) {
    const int q = t.q;
    const int c = ix3s(g, t.i, t.j, t.k), cp = c + s;
    double dxl[4] = {A.dxlagM[c - s], A.dxlagM[c], A.dxlagM[cp], A.dxlagM[cp + s]};
    const double *U1 = A.U[0];
    const double *Q[5] = {A.U[0], A.U[A.VT1], A.U[A.VT2], A.blg[0], A.blg[1]};
    Slope2 sa[5], sb[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      sa[v] = vlim_parts(Q[v][c - s], Q[v][c], Q[v][cp], dxl[0], dxl[1], dxl[2]);
      sb[v] = vlim_parts(Q[v][c], Q[v][cp], Q[v][cp + s], dxl[1], dxl[2], dxl[3]);
    }
    double Tp0 = tswitch(A.U[4][c], U1[c]), Tp1 = tswitch(A.U[4][cp], U1[cp]);
    double Te0 = (q >= 1) ? tswitch(A.u5e_in[c], U1[c]) : Tp0;
    double Te1 = (q + 1 <= n) ? tswitch(A.u5e_in[cp], U1[cp]) : Tp1;
    bool edge = (q % A.nblk) == 0;
    double gi[5], gp[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      gi[v] = vlim_pick(sa[v], Te0);
      gp[v] = vlim_pick(sb[v], edge ? Tp1 : Te1);
    }
    MocV L = moc_iface2(A, c, cp, q, dxl, gi, gp, dt);
    MocV R = L;
    if (edge) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        gi[v] = vlim_pick(sa[v], Tp0);
        gp[v] = vlim_pick(sb[v], Te1);
      }
      R = moc_iface2(A, c, cp, q, dxl, gi, gp, dt);
    }
    sL[0][t.fslot] = L.Bay; sL[1][t.fslot] = L.Vay; sL[2][t.fslot] = L.Baz; sL[3][t.fslot] = L.Vaz;
    sR[0][t.fslot] = R.Bay; sR[1][t.fslot] = R.Vay; sR[2][t.fslot] = R.Baz; sR[3][t.fslot] = R.Vaz;
  }
  __syncthreads();
  if (!t.cell_ok) return;
  const int q = t.q;
  const int c = ix3s(g, t.i, t.j, t.k), cm = c - s, cp = c + s, cpp = cp + s;
  const double *UBn = A.U[A.BN];
  const double *xe = A.xe;
  double Bnap = UBn[cp] + (A.xlag[c] - xe[q + 1]) * (UBn[cpp] - UBn[c]) / (xe[q + 2] - xe[q]);
  double Bnam = UBn[c] + (A.xlag[cm] - xe[q]) * (UBn[cp] - UBn[cm]) / (xe[q + 1] - xe[q - 1]);
  double Vnap = A.va[c], Vnam = A.va[cm];
  double dxl = A.dxlagM[c];
  double rho = A.U[0][c];
  double BayL = sL[0][t.rslot], VayL = sL[1][t.rslot], BazL = sL[2][t.rslot], VazL = sL[3][t.rslot];
  double BayR = sR[0][t.lslot], VayR = sR[1][t.lslot], BazR = sR[2][t.lslot], VazR = sR[3][t.lslot];
  A.Uo[A.VN][c] = A.U[A.VN][c] * rho + dt / dxl * 0.5 * (Bnap * Bnap - Bnam * Bnam);
  A.Uo[A.VT1][c] = A.U[A.VT1][c] * rho + dt / dxl * (Bnap * BayL - Bnam * BayR);
  A.Uo[A.VT2][c] = A.U[A.VT2][c] * rho + dt / dxl * (Bnap * BazL - Bnam * BazR);
  A.Uo[4][c] = A.u5e_in[c] * rho + dt / dxl *
                                      ((0.5 * Bnap * Vnap + BayL * VayL + BazL * VazL) * Bnap -
                                       (0.5 * Bnam * Vnam + BayR * VayR + BazR * VazR) * Bnam);
  A.blgo[0][c] = A.blg[0][c] + dt / dxl * (Bnap * VayL - Bnam * VayR);
  A.blgo[1][c] = A.blg[1][c] + dt / dxl * (Bnap * VazL - Bnam * VazR);
}

// ---------------------------------------------------------------------------
// REMAP + primitives
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(576) void k_remap(SweepIO A, Geo g, double dt) {
  __shared__ double sF[15][NSLOT];
  const Tile t = tile_map<D>(g);
  const int s = g.st[D];
  if (t.face_ok) {
    const int c = ix3s(g, t.i, t.j, t.k);
    double vadt = A.va[c] * dt;
    int u;
    double depend, sgn;
    if (vadt <= 0.0) {
      u = c + s;
      depend = 0.5 * (A.dxlagM[u] + vadt) * A.facdep;
      sgn = -1.0;
    } else {
      u = c;
      depend = 0.5 * (A.dxlagM[u] - vadt) * A.facdep;
      sgn = 1.0;
    }
    double dxm = A.dxlagM[u - s], dx0 = A.dxlagM[u], dxp = A.dxlagM[u + s];
    double T = tswitch(A.U[4][u], A.U[0][u]);
#pragma unroll
    for (int v = 0; v < 15; ++v) {
      const double *Qv = v < 5 ? A.U[v] : v < 7 ? A.blg[v - 5] : A.sp[v - 7];
      double gr = vlim_pick(vlim_parts(Qv[u - s], Qv[u], Qv[u + s], dxm, dx0, dxp), T);
      double face = sgn > 0.0 ? Qv[u] + depend * gr : Qv[u] - depend * gr;
      sF[v][t.fslot] = -(vadt * face);
    }
  }
  __syncthreads();
  if (!t.cell_ok) return;
  const int q = t.q;
  const int c = ix3s(g, t.i, t.j, t.k);
  double dxl = A.dxlagM[c];
  double dxE = A.xe[q + 1] - A.xe[q];
  double Qn[15];
#pragma unroll
  for (int v = 0; v < 15; ++v) {
    const double *Qv = v < 5 ? A.U[v] : v < 7 ? A.blg[v - 5] : A.sp[v - 7];
    Qn[v] = (Qv[c] * dxl + sF[v][t.rslot] - sF[v][t.lslot]) / dxE;
  }
  double rho = dmax(1.27 * A.rmin, Qn[0]);
  double invd = 1.0 / rho;
  double vx = Qn[1] * invd, vy = Qn[2] * invd, vz = Qn[3] * invd;
  double s4 = Qn[7] + Qn[8] + Qn[10] + Qn[11];
  double gg = 3.0 * s4 + 5.0 * Qn[9];
  double gi = (2.0 * s4 + 2.0 * Qn[9]) / gg;
  double ekin = 0.5 * (vx * vx + vy * vy + vz * vz) * rho;
  double bn = A.bcc[D][c];
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

template __global__ void k_rhs<0>(SweepIO, Geo, double);
template __global__ void k_rhs<1>(SweepIO, Geo, double);
template __global__ void k_rhs<2>(SweepIO, Geo, double);
template __global__ void k_moc<0>(SweepIO, Geo, double);
template __global__ void k_moc<1>(SweepIO, Geo, double);
template __global__ void k_moc<2>(SweepIO, Geo, double);
template __global__ void k_remap<0>(SweepIO, Geo, double);
template __global__ void k_remap<1>(SweepIO, Geo, double);
template __global__ void k_remap<2>(SweepIO, Geo, double);

} // namespace mcf
