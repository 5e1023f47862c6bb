// mcf_kernels.hip -- gfx950 kernels of the per-timestep MHD update
// (FP64 path).  Layout: one FP64 array per field on the Fortran index box
// (-1:N+2)^3, x fastest; one thread per cell, face or edge of the stage being
// computed, consecutive threads along x so every load is coalesced.
//
// Stage structure follows EVOLVE (3DMHDCfCA/MCF3DSGincfca3.f90, "M3"):
// ghost fills (BC_MPI / BC_MPI_OT, M3:800-1254) are device kernels with the
// reference's x -> y -> z pass order and x-inflow rules; every stage kernel
// reads the arrays its reference routine reads, so the fills between stages
// reproduce the reference's halo contents exactly.
#include "mcf_kernels.h"
#include "mcf_prec.h"
#include "mcf_cool.h"

namespace mcf {

// ---------------------------------------------------------------------------
// index helpers
// ---------------------------------------------------------------------------
// thread -> (i,j,k) of a box, x fastest; false when out of range
__device__ __forceinline__ bool box_ijk(const Box &b, int &i, int &j, int &k) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long n01 = (long)b.cnt[0] * b.cnt[1];
  if (t >= n01 * b.cnt[2]) return false;
  k = b.lo[2] + (int)(t / n01);
  long r = t - (long)(k - b.lo[2]) * n01;
  j = b.lo[1] + (int)(r / b.cnt[0]);
  i = b.lo[0] + (int)(r - (long)(j - b.lo[1]) * b.cnt[0]);
  return true;
}

__device__ __forceinline__ double gam_of(const double *const *sp, int c) {
  double s4 = sp[0][c] + sp[1][c] + sp[3][c] + sp[4][c];
  double g = 3.0 * s4 + 5.0 * sp[2][c];
  return (5.0 * s4 + 7.0 * sp[2][c]) / g;
}
__device__ __forceinline__ double gammi1_of(const double *const *sp, int c) {
  double s4 = sp[0][c] + sp[1][c] + sp[3][c] + sp[4][c];
  double g = 3.0 * s4 + 5.0 * sp[2][c];
  return (2.0 * s4 + 2.0 * sp[2][c]) / g;
}

// ---------------------------------------------------------------------------
// ghost fills
// ---------------------------------------------------------------------------
// The periodic y pass and then the periodic z pass of a fill in one launch.
// Done in sequence, the passes leave at a ghost point (i, j, k) the value of
// (i, jw, kw), where jw = j -+ n1 if the point's field has a y pass and j is
// one of its nol ghost layers (else jw = j), and kw likewise along z; no
// such source point is written by either pass, so every point can copy its
// source directly.  Points: the y ghost layers over all (i, k), then the z
// ghost layers over all i and the j outside the y ghost layers.
__global__ void k_fill_yz(FillList L, Geo g, int nol) {
  const int m0 = g.n[0] + 4, m1 = g.n[1] + 4, m2 = g.n[2] + 4, n1 = g.n[1], n2 = g.n[2];
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ra = 2L * nol * m0 * m2, rb = 2L * nol * m0 * (m1 - 2 * nol);
  int i, j, k; // memory positions (index + 1)
  if (t < ra) {
    i = (int)(t % m0);
    long r = t / m0;
    k = (int)(r % m2);
    const int layer = (int)(r / m2);
    j = layer < nol ? 2 - nol + layer : n1 + 2 + (layer - nol);
  } else if ((t -= ra) < rb) {
    i = (int)(t % m0);
    long r = t / m0;
    const int q = (int)(r % (m1 - 2 * nol));
    const int layer = (int)(r / (m1 - 2 * nol));
    j = q < 2 - nol ? q : (q < 2 - nol + n1 ? q + nol : q + 2 * nol);
    k = layer < nol ? 2 - nol + layer : n2 + 2 + (layer - nol);
  } else {
    return;
  }
  // ghost layers within nol: memory positions 2-nol..1 and n+2..n+1+nol
  const bool yg = (j >= 2 - nol && j <= 1) || (j >= n1 + 2 && j <= n1 + 1 + nol);
  const bool zg = (k >= 2 - nol && k <= 1) || (k >= n2 + 2 && k <= n2 + 1 + nol);
  const int jw = j <= 1 ? j + n1 : j - n1, kw = k <= 1 ? k + n2 : k - n2;
  const int c = i + j * g.st[1] + k * g.st[2];
  for (int f = 0; f < L.nf; ++f) {
    const bool wy = yg && (L.dm[f] & 1), wz = zg && (L.dm[f] & 2);
    if (!wy && !wz) continue;
    L.a[f][c] = L.a[f][i + (wy ? jw : j) * g.st[1] + (wz ? kw : k) * g.st[2]];
  }
}

// x-ghost rules of BC_MPI / BC_MPI_OT, the periodic copy (XR_PERIODIC) or
// an inflow (BC=4) end's rules
// (M3:818-884, 948-1092).  One thread per ghost point and side, the ghost
// layers of a row on adjacent threads (adjacent in memory); the fields are
// processed in list order, as the reference's K loop does.  A thread reads
// only its own ghost point and interior points, so the order of the points
// does not matter.
__global__ void k_fill_xrule(XFill X, Geo g, int nol) {
  int ny4 = g.n[1] + 4, nz4 = g.n[2] + 4, nx = g.n[0];
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)ny4 * nz4 * (2 * nol);
  if (t >= tot) return;
  int layer = (int)(t % (2 * nol));
  long r = t / (2 * nol);
  int j = (int)(r % ny4) - 1;
  int k = (int)(r / ny4) - 1;
  int side = layer < nol ? 0 : 1;
  int i = side == 0 ? 1 - nol + layer : nx + 1 + (layer - nol);
  int c = ix3(g, i, j, k);
  bool inner = (j >= 1 && j <= g.n[1] && k >= 1 && k <= g.n[2]);
  for (int f = 0; f < X.nf; ++f) {
    double *a = X.a[f];
    switch (X.rule[f][side]) {
    case XR_PERIODIC: a[c] = a[ix3(g, side == 0 ? i + nx : i - nx, j, k)]; break;
    case XR_CONST: a[c] = X.val[f][side]; break;
    case XR_COPY: a[c] = a[ix3(g, X.src[f][side], j, k)]; break;
    case XR_FACE: // U(:,:,:,6): left copies face 1, right only N+2 copies face N+1
      if (side == 0) a[c] = a[ix3(g, 1, j, k)];
      else if (i >= nx + 2) a[c] = a[ix3(g, nx + 1, j, k)];
      break;
    case XR_MOM: a[c] = X.rho[c] * X.val[f][side]; break;
    case XR_ENERGY: {
      double v = X.val2[side];
      a[c] = X.val[f][side] * 1.5 + 0.5 * (X.rho[c] * (v * v) + X.bcc1[c] * X.bcc1[c] +
                                          X.blg1[c] * X.blg1[c] + X.blg2[c] * X.blg2[c]);
      break;
    }
    case XR_SPECIES:
      if (inner) a[c] = X.val[f][side] * X.rho[c];
      break;
    case XR_DTF: { // M3:818-826 (left), 854-861 (right), in the reference's operation order
      const DtfShift &D = X.dtf;
      const double dx0 = D.dx0[side], L = D.L[side];
      double xp;
      int nn;
      if (side == 0) {
        xp = D.xe[1] - 0.5 * dx0 + (double)i * dx0 - D.t * D.v[0];
        nn = (int)(fabs(xp) / L) + 1;
        xp = xp + L * (double)nn;
      } else {
        xp = 0.5 * (D.xe[i] + D.xe[i + 1]) - D.t * D.v[1];
        nn = (int)(fabs(xp) / L);
        xp = xp - L * (double)nn;
      }
      const int ii = (int)(xp / dx0) + 1;
      a[c] = (double)D.f[ix3(g, ii, j, k)];
      break;
    }
    }
  }
}

// ---------------------------------------------------------------------------
// Constrained transport (M3:2115-2445)
// ---------------------------------------------------------------------------
// CC(1): face velocities (dx-weighted) on the interior faces
__global__ void k_cc1(CtArgs A, Geo g, Box b) {
  int i, j, k;
  if (!box_ijk(b, i, j, k)) return;
  int c = ix3(g, i, j, k);
  int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  if (j <= ny && k <= nz) {
    double a = A.gdx[0][i], bb = A.gdx[0][i + 1]; // dx(i-1), dx(i)
    A.vfc[0][c] = (a * A.U[1][c] + bb * A.U[1][c - 1]) / (a + bb);
  }
  if (i <= nx && k <= nz) {
    double a = A.gdx[1][j], bb = A.gdx[1][j + 1];
    A.vfc[1][c] = (a * A.U[2][c] + bb * A.U[2][c - g.st[1]]) / (a + bb);
  }
  if (i <= nx && j <= ny) {
    double a = A.gdx[2][k], bb = A.gdx[2][k + 1];
    A.vfc[2][c] = (a * A.U[3][c] + bb * A.U[3][c - g.st[2]]) / (a + bb);
  }
}

// What CCT needs at one edge of orientation (d1, d2), corner c0 - s1/2 -
// s2/2: the face values of V_d1, B_d1 at c0 - s2 ([0]) and c0 ([1]) and of
// V_d2, B_d2 at c0 - s1 ([0]) and c0 ([1]); their transverse VLIMIT slopes
// dQ0/dQ2 (of V_d1, B_d1 along d2) and dQ1/dQ3 (of V_d2, B_d2 along d1) at
// the same cells; the widths dy = dx_d2 at p2 and p2+1 ... as dye[p2],
// dye[p2+1] and dx = dxe[p1], dxe[p1+1]; and the corner density dnc.
template <class R> struct EdgeIn {
  R v1[2], b1[2], v2[2], b2[2];
  R dq0[2], dq2[2], dq1[2], dq3[2];
  R dy[2], dx[2];
  double dnc;
};

// CCT (M3:2186-2445) from the edge's inputs: the four candidate pairs with
// their 2x2 corrections (M3:2278-2385), the flags and 1e13 weights
// (M3:2387-2415), the weighted averages (M3:2419-2433) and EMF =
// 1/4 (v1 B2 - v2 B1) sqrt(dnc).  Arithmetic in R (FP32 on the mixed path).
template <class R> __device__ __forceinline__ double cct_core(const EdgeIn<R> &E, double facdep, double dt_d) {
  using Q = Ar<R>;
  const R dnc = (R)E.dnc;
  const R dt = (R)dt_d;
  const R dsqinv = Q::div((R)1.0, Q::sqrt_(dnc));
  const R facm = (R)facdep * (R)0.5;
  R C1p[4], C1m[4], C2p[4], C2m[4];
#pragma unroll
  for (int cs = 0; cs < 4; ++cs) {
    // 0: LL (S1=+1,S2=+1)  1: RL (-1,+1)  2: LR (+1,-1)  3: RR (-1,-1)
    const R S1 = (cs == 0 || cs == 2) ? (R)1.0 : (R)-1.0;
    const R S2 = (cs == 0 || cs == 1) ? (R)1.0 : (R)-1.0;
    const int u1 = (cs == 0 || cs == 2) ? 0 : 1; // q1 = c0 - s2 or c0
    const int u2 = (cs == 0 || cs == 1) ? 0 : 1; // q2 = c0 - s1 or c0
    const R dyv = E.dy[u1];
    const R dxv = E.dx[u2];
    const R vx = E.v1[u1], bx = E.b1[u1], vy = E.v2[u2], by = E.b2[u2];
    R a1p = vx + bx * dsqinv;
    R da1p = (E.dq0[u1] + E.dq2[u1] * dsqinv) * facm;
    R a1m = vx - bx * dsqinv;
    R da1m = (E.dq0[u1] - E.dq2[u1] * dsqinv) * facm;
    R a2m = vy - by * dsqinv;
    R da2m = (E.dq1[u2] - E.dq3[u2] * dsqinv) * facm;
    R a2p = vy + by * dsqinv;
    R da2p = (E.dq1[u2] + E.dq3[u2] * dsqinv) * facm;
    R t;
    t = a1p + da1p * dyv * S1 - da1p * (a2m + da2m * dxv * S2) * dt;
    C1p[cs] = Q::div(t, (R)1.0 - da1p * da2m * dt * dt);
    t = a2m + da2m * dxv * S2 - da2m * (a1p + da1p * dyv * S1) * dt;
    C2m[cs] = Q::div(t, (R)1.0 - da1p * da2m * dt * dt);
    t = a1m + da1m * dyv * S1 - da1m * (a2p + da2p * dxv * S2) * dt;
    C1m[cs] = Q::div(t, (R)1.0 - da1m * da2p * dt * dt);
    t = a2p + da2p * dxv * S2 - da2p * (a1m + da1m * dyv * S1) * dt;
    C2p[cs] = Q::div(t, (R)1.0 - da1m * da2p * dt * dt);
  }
  const R z = (R)0.0;
  int f0[4], f1[4];
  f0[0] = (C1p[0] > z) && (C2m[0] > z);
  f1[0] = (C1m[0] > z) && (C2p[0] > z);
  f0[1] = (C1p[1] > z) && (C2m[1] < z);
  f1[1] = (C1m[1] > z) && (C2p[1] < z);
  f0[2] = (C1p[2] < z) && (C2m[2] > z);
  f1[2] = (C1m[2] < z) && (C2p[2] > z);
  f0[3] = (C1p[3] < z) && (C2m[3] < z);
  f1[3] = (C1m[3] < z) && (C2p[3] < z);
  R W0[4], W1[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    // the weight of a candidate pair whose two flags are both set is 1e13:
    // int(0.51 f0 + 0.51 f1) of flags in {0, 1} is 1 exactly when both are
    const int both = f0[m] & f1[m];
    W0[m] = (R)f0[m] + (R)(1.0e13 - 1.0) * (R)both;
    W1[m] = (R)f1[m] + (R)(1.0e13 - 1.0) * (R)both;
  }
  R c1p = z, c2m = z, c1m = z, c2p = z;
  if (f0[0] + f0[1] + f0[2] + f0[3] != 0) {
    c1p = W0[0] * fabs(C2m[0]) * C1p[0] + W0[1] * fabs(C2m[1]) * C1p[1] + W0[2] * fabs(C2m[2]) * C1p[2] +
          W0[3] * fabs(C2m[3]) * C1p[3];
    c1p = Q::div(c1p, W0[0] * fabs(C2m[0]) + W0[1] * fabs(C2m[1]) + W0[2] * fabs(C2m[2]) + W0[3] * fabs(C2m[3]));
    c2m = W0[0] * fabs(C1p[0]) * C2m[0] + W0[1] * fabs(C1p[1]) * C2m[1] + W0[2] * fabs(C1p[2]) * C2m[2] +
          W0[3] * fabs(C1p[3]) * C2m[3];
    c2m = Q::div(c2m, W0[0] * fabs(C1p[0]) + W0[1] * fabs(C1p[1]) + W0[2] * fabs(C1p[2]) + W0[3] * fabs(C1p[3]));
  }
  if (f1[0] + f1[1] + f1[2] + f1[3] != 0) {
    c1m = W1[0] * fabs(C2p[0]) * C1m[0] + W1[1] * fabs(C2p[1]) * C1m[1] + W1[2] * fabs(C2p[2]) * C1m[2] +
          W1[3] * fabs(C2p[3]) * C1m[3];
    c1m = Q::div(c1m, W1[0] * fabs(C2p[0]) + W1[1] * fabs(C2p[1]) + W1[2] * fabs(C2p[2]) + W1[3] * fabs(C2p[3]));
    c2p = W1[0] * fabs(C1m[0]) * C2p[0] + W1[1] * fabs(C1m[1]) * C2p[1] + W1[2] * fabs(C1m[2]) * C2p[2] +
          W1[3] * fabs(C1m[3]) * C2p[3];
    c2p = Q::div(c2p, W1[0] * fabs(C1m[0]) + W1[1] * fabs(C1m[1]) + W1[2] * fabs(C1m[2]) + W1[3] * fabs(C1m[3]));
  }
  R v1a = c1p + c1m, v2a = c2p + c2m, B1a = c1p - c1m, B2a = c2p - c2m;
  return (double)((R)0.25 * (v1a * B2a - v2a * B1a) * Q::sqrt_(dnc));
}

// CCT for the whole box (1..n+1)^3 as a march along z, every transverse
// slope formed once per cell.  A wave holds 64 consecutive cells of one
// (j, k) row, lane l at x = ib - 2 + l, and writes the EMFs of lanes 2..62
// (61 cells: the x slopes at i and i-1 need the values at i-2..i+1).  Per
// plane k each lane forms, with its own cell's limiter switch:
//   along x: V_y, B_y, V_z, B_z from its neighbours' values (lane shuffles);
//            the slopes at i-1 then come from the left lane;
//   along y: V_x, B_x, V_z, B_z at rows j and j-1 (rows j-2..j+1 loaded;
//            the slope at j-1 is also formed by the wave of row j-1);
//   along z: V_x, B_x, V_y, B_y at plane k (planes k-1, k+1 in registers),
//            the slopes at k-1 kept from the previous plane.
// The three edges of the cell then go through cct_core.  Against a kernel
// with one thread per cell that formed all 24 slopes its three edges use
// (each slope twice), the march took 1.00 / 1.77 ms instead of 1.12 / 1.93
// (mixed / FP64, 256^3) with bit-identical EMFs; sharing the row-(j-1)
// slopes through LDS instead of forming them in each wave needs a barrier
// per plane and took 1.29 / 1.94.
template <class R> __device__ __forceinline__ void ct_march(const CtArgs &A, const Geo &g, double dt_d) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int j = 1 + CTM_ROWS * (int)blockIdx.y + w;
  if (j > ny + 1) return; // whole waves; no barrier in this kernel
  const int ib = 1 + CTM_OUT * (int)blockIdx.x;
  const int i = ib - 2 + lane;
  const int ic = min(i, nx + 2); // lanes past the last ghost read the last ghost
  const bool outx = lane >= 2 && lane <= 62 && i <= nx + 1;
  const int k0 = 1 + CTM_ZSEG * (int)blockIdx.z;
  const int k1 = min(nz + 1, k0 + CTM_ZSEG - 1);
  const int sy = g.st[1], sz = g.st[2];
  const double *Vx = A.vfc[0], *Vy = A.vfc[1], *Vz = A.vfc[2];
  const double *Bx = A.U[5], *By = A.U[6], *Bz = A.U[7], *rh = A.U[0], *pp = A.U[4];
  // widths: x at the lane's cell (indices clamped to the dx array for the
  // halo lanes, whose slopes are not used), y at row j
  const double *gx = A.gdx[0], *gy = A.gdx[1], *gz = A.gdx[2];
  const int ix0 = max(0, min(ic, nx + 3)), ix1 = min(ic + 1, nx + 3), ix2 = min(ic + 2, nx + 3);
  const R xw0 = (R)gx[ix0], xw1 = (R)gx[ix1], xw2 = (R)gx[ix2];
  const R yw_m = (R)gy[j - 1], yw0 = (R)gy[j], yw1 = (R)gy[j + 1], yw2 = (R)gy[j + 2];
  // z window of the own column: V_x, B_x, V_y, B_y at planes k-1 and k,
  // their z slopes at k-1, and rho at k-1 in rows j and j-1
  int c = ix3(g, ic, j, k0 - 1);
  double wVx0 = Vx[c], wBx0 = Bx[c], wVy0 = Vy[c], wBy0 = By[c];
  double wVx1 = Vx[c + sz], wBx1 = Bx[c + sz], wVy1 = Vy[c + sz], wBy1 = By[c + sz];
  double rho_km = rh[c], rhoJ_km = rh[c - sy];
  R zVx0, zBx0, zVy0, zBy0;
  {
    const bool mm = use_minmod_r<R>(pp[c], rh[c]);
    const R z0 = (R)gz[k0 - 1], z1 = (R)gz[k0], z2 = (R)gz[k0 + 1];
    const int cm = c - sz;
    zVx0 = vslope<R>(Vx[cm], wVx0, wVx1, z0, z1, z2, mm);
    zBx0 = vslope<R>(Bx[cm], wBx0, wBx1, z0, z1, z2, mm);
    zVy0 = vslope<R>(Vy[cm], wVy0, wVy1, z0, z1, z2, mm);
    zBy0 = vslope<R>(By[cm], wBy0, wBy1, z0, z1, z2, mm);
  }
  for (int k = k0; k <= k1; ++k) {
    c += sz; // (ic, j, k)
    // loads: own column at k+1 (z window) and k; rows j-1, j-2, j+1 at k
    const double nVx = Vx[c + sz], nBx = Bx[c + sz], nVy = Vy[c + sz], nBy = By[c + sz];
    const double vz = Vz[c], bz = Bz[c], rho = rh[c], p = pp[c];
    const int cJ = c - sy, cJJ = c - 2 * sy, cP = c + sy;
    const double VxJ = Vx[cJ], BxJ = Bx[cJ], VzJ = Vz[cJ], BzJ = Bz[cJ], rhoJ = rh[cJ], pJ = pp[cJ];
    const double VxJJ = Vx[cJJ], BxJJ = Bx[cJJ], VzJJ = Vz[cJJ], BzJJ = Bz[cJJ];
    const double VxP = Vx[cP], BxP = Bx[cP], VzP = Vz[cP], BzP = Bz[cP];
    const bool mm0 = use_minmod_r<R>(p, rho);
    const bool mmJ = use_minmod_r<R>(pJ, rhoJ);
    const R zw0 = (R)gz[k], zw1 = (R)gz[k + 1], zw2 = (R)gz[k + 2];
    // along z at k
    const R zVx1 = vslope<R>(wVx0, wVx1, nVx, zw0, zw1, zw2, mm0);
    const R zBx1 = vslope<R>(wBx0, wBx1, nBx, zw0, zw1, zw2, mm0);
    const R zVy1 = vslope<R>(wVy0, wVy1, nVy, zw0, zw1, zw2, mm0);
    const R zBy1 = vslope<R>(wBy0, wBy1, nBy, zw0, zw1, zw2, mm0);
    // along y at j and j-1
    const R yVx1 = vslope<R>(VxJ, wVx1, VxP, yw0, yw1, yw2, mm0);
    const R yBx1 = vslope<R>(BxJ, wBx1, BxP, yw0, yw1, yw2, mm0);
    const R yVz1 = vslope<R>(VzJ, vz, VzP, yw0, yw1, yw2, mm0);
    const R yBz1 = vslope<R>(BzJ, bz, BzP, yw0, yw1, yw2, mm0);
    const R yVx0 = vslope<R>(VxJJ, VxJ, wVx1, yw_m, yw0, yw1, mmJ);
    const R yBx0 = vslope<R>(BxJJ, BxJ, wBx1, yw_m, yw0, yw1, mmJ);
    const R yVz0 = vslope<R>(VzJJ, VzJ, vz, yw_m, yw0, yw1, mmJ);
    const R yBz0 = vslope<R>(BzJJ, BzJ, bz, yw_m, yw0, yw1, mmJ);
    // along x at i (neighbour values from the adjacent lanes), then at i-1
    const double VyL = __shfl_up(wVy1, 1, 64), VyR = __shfl_down(wVy1, 1, 64);
    const double ByL = __shfl_up(wBy1, 1, 64), ByR = __shfl_down(wBy1, 1, 64);
    const double VzL = __shfl_up(vz, 1, 64), VzR = __shfl_down(vz, 1, 64);
    const double BzL = __shfl_up(bz, 1, 64), BzR = __shfl_down(bz, 1, 64);
    const R xVy1 = vslope<R>(VyL, wVy1, VyR, xw0, xw1, xw2, mm0);
    const R xBy1 = vslope<R>(ByL, wBy1, ByR, xw0, xw1, xw2, mm0);
    const R xVz1 = vslope<R>(VzL, vz, VzR, xw0, xw1, xw2, mm0);
    const R xBz1 = vslope<R>(BzL, bz, BzR, xw0, xw1, xw2, mm0);
    const R xVy0 = __shfl_up(xVy1, 1, 64), xBy0 = __shfl_up(xBy1, 1, 64);
    const R xVz0 = __shfl_up(xVz1, 1, 64), xBz0 = __shfl_up(xBz1, 1, 64);
    const double rhoL = __shfl_up(rho, 1, 64), rhoJL = __shfl_up(rhoJ, 1, 64), rhoKL = __shfl_up(rho_km, 1, 64);
    if (outx) {
      double e0 = 0.0, e1 = 0.0, e2 = 0.0;
      if (k <= nz) { // edge along z: (d1, d2) = (x, y)
        EdgeIn<R> E;
        E.v1[0] = (R)VxJ; E.b1[0] = (R)BxJ; E.v1[1] = (R)wVx1; E.b1[1] = (R)wBx1;
        E.v2[0] = (R)VyL; E.b2[0] = (R)ByL; E.v2[1] = (R)wVy1; E.b2[1] = (R)wBy1;
        E.dq0[0] = yVx0; E.dq2[0] = yBx0; E.dq0[1] = yVx1; E.dq2[1] = yBx1;
        E.dq1[0] = xVy0; E.dq3[0] = xBy0; E.dq1[1] = xVy1; E.dq3[1] = xBy1;
        E.dy[0] = yw0; E.dy[1] = yw1;
        E.dx[0] = xw0; E.dx[1] = xw1;
        E.dnc = 0.25 * (rho + rhoL + rhoJ + rhoJL);
        e0 = cct_core<R>(E, A.facdep, dt_d);
      }
      if (i <= nx) { // edge along x: (d1, d2) = (y, z)
        EdgeIn<R> E;
        E.v1[0] = (R)wVy0; E.b1[0] = (R)wBy0; E.v1[1] = (R)wVy1; E.b1[1] = (R)wBy1;
        E.v2[0] = (R)VzJ; E.b2[0] = (R)BzJ; E.v2[1] = (R)vz; E.b2[1] = (R)bz;
        E.dq0[0] = zVy0; E.dq2[0] = zBy0; E.dq0[1] = zVy1; E.dq2[1] = zBy1;
        E.dq1[0] = yVz0; E.dq3[0] = yBz0; E.dq1[1] = yVz1; E.dq3[1] = yBz1;
        E.dy[0] = zw0; E.dy[1] = zw1;
        E.dx[0] = yw0; E.dx[1] = yw1;
        E.dnc = 0.25 * (rho + rhoJ + rho_km + rhoJ_km);
        e1 = cct_core<R>(E, A.facdep, dt_d);
      }
      if (j <= ny) { // edge along y: (d1, d2) = (z, x)
        EdgeIn<R> E;
        E.v1[0] = (R)VzL; E.b1[0] = (R)BzL; E.v1[1] = (R)vz; E.b1[1] = (R)bz;
        E.v2[0] = (R)wVx0; E.b2[0] = (R)wBx0; E.v2[1] = (R)wVx1; E.b2[1] = (R)wBx1;
        E.dq0[0] = xVz0; E.dq2[0] = xBz0; E.dq0[1] = xVz1; E.dq2[1] = xBz1;
        E.dq1[0] = zVx0; E.dq3[0] = zBx0; E.dq1[1] = zVx1; E.dq3[1] = zBx1;
        E.dy[0] = xw0; E.dy[1] = xw1;
        E.dx[0] = zw0; E.dx[1] = zw1;
        E.dnc = 0.25 * (rho + rhoL + rho_km + rhoKL);
        e2 = cct_core<R>(E, A.facdep, dt_d);
      }
      if (k <= nz) A.emf[2][c] = e0;
      if (i <= nx) A.emf[0][c] = e1;
      if (j <= ny) A.emf[1][c] = e2;
    }
    // advance the z window
    wVx0 = wVx1; wBx0 = wBx1; wVy0 = wVy1; wBy0 = wBy1;
    wVx1 = nVx; wBx1 = nBx; wVy1 = nVy; wBy1 = nBy;
    zVx0 = zVx1; zBx0 = zBx1; zVy0 = zVy1; zBy0 = zBy1;
    rho_km = rho; rhoJ_km = rhoJ;
  }
}
template <class R>
__global__ __launch_bounds__(64 * CTM_ROWS) void k_ct_march(CtArgs A, Geo g, double dt_d) {
  ct_march<R>(A, g, dt_d);
}
template __global__ void k_ct_march<double>(CtArgs, Geo, double);
// The FP32 (mixed) march capped at 168 registers, three waves per SIMD
// instead of two (36 bytes of scratch spill): 0.99-1.02 against 1.03 ms at
// 256^3, the same EMFs.  The FP64 march would spill 348 bytes at that cap.
__global__ __launch_bounds__(64 * CTM_ROWS, 3) void k_ct_march_mixed(CtArgs A, Geo g, double dt_d) {
  ct_march<float>(A, g, dt_d);
}

// CC(3): face-B induction (M3:2159-2169).  A variant fused with CC(4),
// out of place over the whole box (each thread also forming the new faces
// at i+1, j+1, k+1 for its cell-centred field), took 0.57 ms against 0.40
// for the two kernels (256^3).
__global__ void k_cc3(CtArgs A, Geo g, Box b, double dt) {
  int i, j, k;
  if (!box_ijk(b, i, j, k)) return;
  int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  int c = ix3(g, i, j, k);
  int sx = 1, sy = g.st[1], sz = g.st[2];
  const double *Ex = A.emf[0], *Ey = A.emf[1], *Ez = A.emf[2];
  double dxi = A.gdx[0][i + 1], dyj = A.gdx[1][j + 1], dzk = A.gdx[2][k + 1];
  if (j <= ny && k <= nz)
    A.U[5][c] = A.U[5][c] + dt / dyj * (Ez[c + sy] - Ez[c]) - dt / dzk * (Ey[c + sz] - Ey[c]);
  if (i <= nx && k <= nz)
    A.U[6][c] = A.U[6][c] + dt / dzk * (Ex[c + sz] - Ex[c]) - dt / dxi * (Ez[c + sx] - Ez[c]);
  if (i <= nx && j <= ny)
    A.U[7][c] = A.U[7][c] + dt / dxi * (Ey[c + sx] - Ey[c]) - dt / dyj * (Ex[c + sy] - Ex[c]);
}

// CC(4): cell-centred field (M3:2171-2181)
__global__ void k_cc4(CtArgs A, Geo g, Box b) {
  int i, j, k;
  if (!box_ijk(b, i, j, k)) return;
  int c = ix3(g, i, j, k);
  A.bcc[0][c] = 0.5 * (A.U[5][c] + A.U[5][c + 1]);
  A.bcc[1][c] = 0.5 * (A.U[6][c] + A.U[6][c + g.st[1]]);
  A.bcc[2][c] = 0.5 * (A.U[7][c] + A.U[7][c + g.st[2]]);
}

// ---------------------------------------------------------------------------
// SOURCE (M3:2448-2572), KI cooling + saturated conduction
// ---------------------------------------------------------------------------
// T = P/(kb n) of a cell, from the pressure array the stage reads T from
__device__ __forceinline__ double src_T(const SrcArgs &A, int c) {
  const double *const *sp = A.sp;
  return A.Tp[c] / (KB * (sp[1][c] + sp[0][c] + sp[2][c] + sp[3][c] + sp[4][c]));
}

__device__ __forceinline__ double qflux(double Tc, double Tcs, double dxi, double dxp) {
  double rt = sqrt((dxi * Tcs + dxp * Tc) / (dxi + dxp));
  return KCOND * dmin(rt, 3.873) * (Tcs - Tc) * 2.0 / (dxi + dxp);
}

// One stage of SOURCE's thermal update (M3:2461-2572), out of place.
// Stage 0 (predictor, dt/2): T from the pressure P (= A.Pn), result to
// A.Pout.  Stage 1 (corrector, dt): T from the predictor (A.Tp), P_old =
// A.Pn, result written over A.Pn (= A.Pout; each cell reads only its own
// P_old there).  Both stages cover the interior; the corrector's first
// ghost layer takes the predictor's ghost values (k_shell_copy), as the
// reference leaves them.
// The backward-Euler cells (dt > 0.2*t_cool) are not solved here: they are
// queued with their conduction increment for k_src_stiff (the surrogate) or
// k_src_imc, so the bisection runs on dense waves.
//
// The cells are visited as a march along z: a block is SRC_ROWS x rows of 64
// cells (one wave per row) over SRC_ZSEG_F32 / _F64 planes.  Per plane the whole block
// forms the temperatures of its rows plus a one-cell frame in x and y
// (coalesced rows) into a 4-plane LDS ring, so each cell reads its six
// neighbours from LDS and every temperature is formed once per block (one
// barrier per plane).
// R = double: the reference's arithmetic, operation by operation.
// R = float (mixed precision): temperatures and their differences in FP64,
// the rates (cooling function, conduction fluxes, time-step tests) in FP32,
// the pressure accumulated in FP64.
template <class R> struct SrcW {
  double T;
  R gi, ns;
};
template <class R> __device__ __forceinline__ SrcW<R> src_w(const SrcArgs &A, int c) {
  const double *const *sp = A.sp;
  const double s0 = sp[0][c], s1 = sp[1][c], s2 = sp[2][c], s3 = sp[3][c], s4 = sp[4][c];
  SrcW<R> w;
  w.T = A.Tp[c] / (KB * (s1 + s0 + s2 + s3 + s4));
  const double q4 = s0 + s1 + s3 + s4;
  if constexpr (Ar<R>::strict) {
    w.gi = (2.0 * q4 + 2.0 * s2) / (3.0 * q4 + 5.0 * s2);
  } else {
    w.gi = Ar<float>::div((float)(2.0 * q4 + 2.0 * s2), (float)(3.0 * q4 + 5.0 * s2));
  }
  w.ns = (R)(s0 + s1 + s3 + s4 + s2);
  return w;
}

__device__ __forceinline__ float fcool_ki_f(float rho, float T) {
  const float N = rho * (1.0f / 1.27f);
  float Laml = 1.0e7f * __expf(Ar<float>::div(-1.184e5f, T * 1.0e3f + 1.0e3f));
  Laml = fminf(Laml, 5.0e3f);
  const float Lamc = 1.4e-2f * Ar<float>::sqrt_(T * 1.0e3f) * __expf(Ar<float>::div(-9.2e1f, T * 1.0e3f));
  return (N * N * (Laml + Lamc) - N) * 3.94656e1f;
}

// conduction flux of qflux with the weights and the temperature difference
// in FP64, the rest in FP32
__device__ __forceinline__ float qflux_f(double Tc, double Tcs, double dxi, double dxp) {
  const float rt = Ar<float>::sqrt_(Ar<float>::div((float)(dxi * Tcs + dxp * Tc), (float)(dxi + dxp)));
  return (float)KCOND * fminf(rt, 3.873f) * (float)(Tcs - Tc) * Ar<float>::div(2.0f, (float)(dxi + dxp));
}

template <class R, bool CHEM>
__global__ __launch_bounds__(64 * SRC_ROWS) void k_src_col(SrcArgs A, Geo g, Box b, double dt, int stage) {
  static_assert(!CHEM || Ar<R>::strict, "the chemistry Fcool runs in FP64");
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  // the queue counters alternate between two pairs: this stage counts in
  // A.qn (cleared by the previous stage) and clears A.qz for the next one,
  // whose consumers of it (k_src_stiff, k_src_imc) finished before this
  // launch
  if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x < 2) A.qz[threadIdx.x] = 0u;
  const int lane = threadIdx.x & 63;
  const int hi0 = b.lo[0] + b.cnt[0] - 1, hi1 = b.lo[1] + b.cnt[1] - 1, hi2 = b.lo[2] + b.cnt[2] - 1;
  const int i = b.lo[0] + 64 * bx + lane;
  const int j = b.lo[1] + SRC_ROWS * by + (int)(threadIdx.x >> 6);
  constexpr int ZSEG = Ar<R>::strict ? SRC_ZSEG_F64 : SRC_ZSEG_F32;
  const int k0 = b.lo[2] + ZSEG * bz;
  const int k1 = min(hi2, k0 + ZSEG - 1);
  const bool own = i <= hi0 && j <= hi1;
  const int ic = min(i, hi0 + 1), jc = min(j, hi1 + 1);
  const int w = (int)(threadIdx.x >> 6);
  const int i0 = b.lo[0] + 64 * bx, j0 = b.lo[1] + SRC_ROWS * by;
  // backward-Euler cells: counted in LDS and staged in the block's own range
  // of A.fqueue (free while this kernel runs), at qoff: the tiles numbered
  // z segment, then row group, then x tile, each range as long as the
  // block's own cells
  __shared__ unsigned s_qn, s_qbase;
  if (threadIdx.x == 0) s_qn = 0u;
  const long qoff = (long)(k0 - b.lo[2]) * b.cnt[0] * b.cnt[1] +
                    (long)(k1 - k0 + 1) * ((long)(j0 - b.lo[1]) * b.cnt[0] +
                                           (long)min(SRC_ROWS, b.cnt[1] - SRC_ROWS * by) * (i0 - b.lo[0]));
  // temperature tiles of planes k-1 .. k+2 (ring of 4): rows j0-1 ..
  // j0+SRC_ROWS by columns i0-1 .. i0+64, formed by all threads together
  // (coalesced rows, the frame included); gamma-1 and n of the block's own
  // cells beside them
  constexpr int TW = 66, TR = SRC_ROWS + 2, TN = TW * TR, NT = 64 * SRC_ROWS;
  __shared__ double sT[4][TR][TW];
  __shared__ R sG[4][SRC_ROWS][64], sN[4][SRC_ROWS][64];
  auto tile = [&](int q) {
    const int sl = q & 3;
#pragma unroll
    for (int t0 = 0; t0 < TN; t0 += NT) {
      const int t = t0 + (int)threadIdx.x;
      if (t < TN) {
        const int r = t / TW, cl = t - r * TW;
        const int x = min(i0 - 1 + cl, hi0 + 1), y = min(j0 - 1 + r, hi1 + 1);
        const SrcW<R> sw = src_w<R>(A, ix3(g, x, y, q));
        sT[sl][r][cl] = sw.T;
        if (r >= 1 && r <= SRC_ROWS && cl >= 1 && cl <= 64) {
          sG[sl][r - 1][cl - 1] = sw.gi;
          sN[sl][r - 1][cl - 1] = sw.ns;
        }
      }
    }
  };
  const double dxi = A.gdx[0][ic + 1], dxp = A.gdx[0][ic + 2], dxm = A.gdx[0][ic];
  const double dyj = A.gdx[1][jc + 1], dyp = A.gdx[1][jc + 2], dym = A.gdx[1][jc];
  tile(k0 - 1);
  tile(k0);
  tile(k0 + 1);
  for (int k = k0; k <= k1; ++k) {
    const int c = ix3(g, ic, jc, k);
    // the slot of plane k+2 was last read for plane k-2, before this barrier
    __syncthreads();
    if (k < k1) tile(k + 2);
    // planes k-1, k, k+1 were written at least one barrier ago
    const int s0 = k & 3, sm = (k - 1) & 3, sp = (k + 1) & 3;
    const double T = sT[s0][w + 1][lane + 1];
    const double Txm = sT[s0][w + 1][lane], Txp = sT[s0][w + 1][lane + 2];
    const double Tym = sT[s0][w][lane + 1], Typ = sT[s0][w + 2][lane + 1];
    const double Tzm = sT[sm][w + 1][lane + 1], Tzp = sT[sp][w + 1][lane + 1];
    SrcW<R> wc;
    wc.T = T;
    wc.gi = sG[s0][w][lane];
    wc.ns = sN[s0][w][lane];
    const bool valid = own;
    const double rho = A.rho[c];
    const double Pn = A.Pn[c];
    const double dzk = A.gdx[2][k + 1], dzp = A.gdx[2][k + 2], dzm = A.gdx[2][k];
    double u, dc;
    bool expl;
    if constexpr (Ar<R>::strict) {
      double CooL;
      if constexpr (CHEM) CooL = own ? ch_fcool(ch_fc_setup(A.sp, A.cols, c), T) : 0.0;
      else CooL = fcool_ki(rho, T);
      const double gi = wc.gi, nsum = wc.ns;
      expl = dt <= 0.2 * Pn / (gi * fabs(CooL));
      u = stage == 0 ? Pn - gi * CooL * dt * 0.5 : Pn - gi * CooL * dt;
      double tcd = gi * KCOND * sqrt(T);
      tcd = 0.49 * nsum * KB / tcd * (dxi * dxi);
      tcd = copysign(0.5, tcd - dt);
      const double qxp = qflux(T, Txp, dxi, dxp);
      const double qxm = qflux(Txm, T, dxm, dxi);
      const double qyp = qflux(T, Typ, dyj, dyp);
      const double qym = qflux(Tym, T, dym, dyj);
      const double qzp = qflux(T, Tzp, dzk, dzp);
      const double qzm = qflux(Tzm, T, dzm, dzk);
      const double div = (qxp - qxm) / dxi + (qyp - qym) / dyj + (qzp - qzm) / dzk;
      dc = stage == 0 ? (tcd + 0.5) * gi * dt * 0.5 * div : (tcd + 0.5) * gi * dt * div;
    } else {
      const float Tf = (float)T, dtf = (float)dt;
      const float CooL = fcool_ki_f((float)rho, Tf);
      const float gi = wc.gi, nsum = wc.ns;
      expl = dtf <= Ar<float>::div(0.2f * (float)Pn, gi * fabsf(CooL));
      const float cool = stage == 0 ? gi * CooL * dtf * 0.5f : gi * CooL * dtf;
      u = Pn - (double)cool;
      const float dxf = (float)dxi;
      float tcd = Ar<float>::div(0.49f * nsum * (float)KB, gi * (float)KCOND * Ar<float>::sqrt_(Tf)) * (dxf * dxf);
      tcd = copysignf(0.5f, tcd - dtf);
      const float qxp = qflux_f(T, Txp, dxi, dxp);
      const float qxm = qflux_f(Txm, T, dxm, dxi);
      const float qyp = qflux_f(T, Typ, dyj, dyp);
      const float qym = qflux_f(Tym, T, dym, dyj);
      const float qzp = qflux_f(T, Tzp, dzk, dzp);
      const float qzm = qflux_f(Tzm, T, dzm, dzk);
      const float div = (qxp - qxm) * __builtin_amdgcn_rcpf(dxf) + (qyp - qym) * __builtin_amdgcn_rcpf((float)dyj) +
                        (qzp - qzm) * __builtin_amdgcn_rcpf((float)dzk);
      dc = (double)(stage == 0 ? (tcd + 0.5f) * gi * dtf * 0.5f * div : (tcd + 0.5f) * gi * dtf * div);
    }
    const bool stiff = valid && !expl;
    const unsigned long long m = __ballot(stiff);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(&s_qn, (unsigned)__popcll(m));
      base = __shfl(base, leader, 64);
      if (stiff) {
        const unsigned pos = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
        A.fqueue[qoff + (long)pos] = c;
        A.qdc[c] = dc;
      }
    }
    if (valid && !stiff) {
      u = u + dc;
      u = dmax(u, A.pmin);
      u = dmin(u, A.pmax);
      A.Pout[c] = u;
    }
  }
  // the block's backward-Euler cells into the stage's queue: one range per
  // block, taken with one atomic
  __syncthreads();
  if (threadIdx.x == 0) s_qbase = s_qn ? atomicAdd(A.qn, s_qn) : 0u;
  __syncthreads();
  for (unsigned t = threadIdx.x; t < s_qn; t += blockDim.x) A.queue[s_qbase + t] = A.fqueue[qoff + (long)t];
}
template __global__ void k_src_col<double, false>(SrcArgs, Geo, Box, double, int);
template __global__ void k_src_col<float, false>(SrcArgs, Geo, Box, double, int);
template __global__ void k_src_col<double, true>(SrcArgs, Geo, Box, double, int);

// Copy of the first ghost layer around the interior, src -> dst: the faces
// k = 0 and n2+1 with their edges and corners, then j = 0 and n1+1 over
// k = 1..n2, then i = 0 and n0+1 over the interior (j, k).
__global__ void k_shell_copy(const double *__restrict__ src, double *__restrict__ dst, Geo g) {
  const long a0 = g.n[0] + 2, a1 = g.n[1] + 2, n1 = g.n[1], n2 = g.n[2];
  const long fz = a0 * a1, fy = a0 * n2, fx = n1 * n2;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (t < 2 * fz) {
    i = (int)(t % a0); j = (int)((t / a0) % a1); k = t < fz ? 0 : g.n[2] + 1;
  } else if ((t -= 2 * fz) < 2 * fy) {
    i = (int)(t % a0); k = 1 + (int)((t / a0) % n2); j = t < fy ? 0 : g.n[1] + 1;
  } else if ((t -= 2 * fy) < 2 * fx) {
    j = 1 + (int)(t % n1); k = 1 + (int)((t / n1) % n2); i = t < fx ? 0 : g.n[0] + 1;
  } else {
    return;
  }
  const int c = ix3(g, i, j, k);
  dst[c] = src[c];
}

// The queued backward-Euler cells of one SOURCE stage, 64 per wave: the
// surrogate predicts, Newton steps refine, the residual test accepts; the
// cell's conduction increment and the pressure clamps are then applied as
// in k_src_col.  Rejected cells go to a second queue for k_src_imc,
// so IMC runs on dense waves.  Waves stride over the queue.
__device__ __forceinline__ void src_finish(const SrcArgs &A, int c, double Pi) {
  double u = Pi + A.qdc[c];
  u = dmax(u, A.pmin);
  u = dmin(u, A.pmax);
  A.Pout[c] = u;
}

__global__ __launch_bounds__(256) void k_src_stiff(SrcArgs A, double dt, int stage) {
  const unsigned nq = *A.qn;
  const unsigned wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const unsigned nw = (gridDim.x * blockDim.x) >> 6;
  const int lane = threadIdx.x & 63;
  for (unsigned base = wid * 64u; base < nq; base += nw * 64u) {
    const unsigned e = base + (unsigned)lane;
    const bool ok = e < nq;
    const int c = ok ? A.queue[e] : A.queue[base];
    const double *const *sp = A.sp;
    const double rho = A.rho[c];
    const double gi = gammi1_of(sp, c);
    const double nsum = sp[0][c] + sp[1][c] + sp[3][c] + sp[4][c] + sp[2][c];
    const double P0 = A.Pn[c];
    const double dts = stage == 0 ? dt * 0.5 : dt;
    float u0, u1, u2;
    sur_features(P0, nsum, dts, rho, gi, u0, u1, u2);
    const float y = mlp_wave_forward(A.pk, A.mw, u0, u1, u2);
    double Pi;
    const bool fell = !sur_refine(P0, nsum, dts, rho, gi, y, Pi);
    const unsigned long long mf = __ballot(ok && fell);
    if (mf) {
      const int leader = __ffsll((long long)mf) - 1;
      unsigned fb = 0;
      if (lane == leader) fb = atomicAdd(A.qn + 1, (unsigned)__popcll(mf));
      fb = __shfl(fb, leader, 64);
      if (ok && fell) A.fqueue[fb + (unsigned)__popcll(mf & ((1ull << lane) - 1ull))] = c;
    }
    if (ok && !fell) src_finish(A, c, Pi);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&A.cnt[0], (unsigned long long)nq);
}

// IMC (M3:2898-2921) on a queue of backward-Euler cells: all of them on the
// plain path (which = 0), the ones the surrogate's residual test rejected
// (which = 1).
template <bool CHEM>
__global__ __launch_bounds__(256) void k_src_imc(SrcArgs A, double dt, int stage, int which) {
  const unsigned nq = A.qn[which];
  const int *q = which ? A.fqueue : A.queue;
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < nq; e += stride) {
    const int c = q[e];
    const double *const *sp = A.sp;
    const double rho = A.rho[c];
    const double gi = gammi1_of(sp, c);
    const double nsum = sp[0][c] + sp[1][c] + sp[3][c] + sp[4][c] + sp[2][c];
    const double dts = stage == 0 ? dt * 0.5 : dt;
    if constexpr (CHEM) src_finish(A, c, ch_imc(ch_fc_setup(sp, A.cols, c), A.Pn[c], nsum, dts, gi));
    else src_finish(A, c, imc(A.Pn[c], nsum, dts, rho, gi));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&A.cnt[which], (unsigned long long)nq);
}
template __global__ void k_src_imc<false>(SrcArgs, double, int, int);
template __global__ void k_src_imc<true>(SrcArgs, double, int, int);

__global__ void k_zero_u32x2(unsigned *p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0u;
}


// ---------------------------------------------------------------------------
// Couran + Stblty: min over interior cells
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// Two-pass min: each block strides over (j,k) rows and writes one pair of
// partial minima; k_dtfinal folds them.  min is exact, so the result does
// not depend on the order.
template <bool CHEM> __global__ void k_dtmin(DtArgs A, Geo g, double *part) {
  __shared__ double s0[16], s1[16];
  const int nx = g.n[0];
  const long rows = (long)g.n[1] * g.n[2];
  const double *const *sp = A.sp;
  double tc = A.tfinal, tl = A.tfinal;
  for (long r = blockIdx.x; r < rows; r += gridDim.x) {
    int j = 1 + (int)(r % g.n[1]), k = 1 + (int)(r / g.n[1]);
    double dyj = A.gdx[1][j + 1], dzk = A.gdx[2][k + 1];
    for (int i = 1 + threadIdx.x; i <= nx; i += blockDim.x) {
      int c = ix3(g, i, j, k);
      double rho = A.U[0][c], p = A.U[4][c];
      double bx = A.bcc[0][c], by = A.bcc[1][c], bz = A.bcc[2][c];
      double gamma = gam_of(sp, c);
      double c2 = (gamma * p + bx * bx + by * by + bz * bz) / rho;
      double dxi = A.gdx[0][i + 1];
      double cf = sqrt(c2);
      double t = dxi / (cf + fabs(A.U[1][c]));
      tc = dmin(tc, t);
      t = dyj / (cf + fabs(A.U[2][c]));
      tc = dmin(tc, t);
      t = dzk / (cf + fabs(A.U[3][c]));
      tc = dmin(tc, t);
      double Nn = sp[1][c] + sp[0][c] + sp[2][c] + sp[3][c] + sp[4][c];
      double Tn = p / (KB * Nn);
      double dl = dmin(dxi, dyj);
      dl = dmin(dl, dzk);
      double gi = gammi1_of(sp, c);
      double alpha = gi * KCOND * sqrt(dmin(Tn, 15.0));
      alpha = 0.4 * Nn * KB / alpha * (dl * dl);
      double CooL;
      if constexpr (CHEM) CooL = ch_fcool(ch_fc_setup(sp, A.cols, c), Tn);
      else CooL = fcool_ki(rho, Tn);
      double tauC = p / gi / fabs(CooL);
      tauC = dmax(0.2 * tauC, 2.5e-4);
      tl = dmin(tl, tauC);
      tl = dmin(tl, alpha);
    }
  }
  tc = wave_min(tc);
  tl = wave_min(tl);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { s0[wid] = tc; s1[wid] = tl; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int w = 1; w < nw; ++w) { tc = fmin(tc, s0[w]); tl = fmin(tl, s1[w]); }
    part[2 * blockIdx.x] = tc;
    part[2 * blockIdx.x + 1] = tl;
  }
}

template __global__ void k_dtmin<false>(DtArgs, Geo, double *);
template __global__ void k_dtmin<true>(DtArgs, Geo, double *);

__global__ void k_dtfinal(const double *part, int nb, double *out, double tfinal) {
  __shared__ double s0[16], s1[16];
  double tc = tfinal, tl = tfinal;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    tc = fmin(tc, part[2 * b]);
    tl = fmin(tl, part[2 * b + 1]);
  }
  tc = wave_min(tc);
  tl = wave_min(tl);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { s0[wid] = tc; s1[wid] = tl; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int w = 1; w < nw; ++w) { tc = fmin(tc, s0[w]); tl = fmin(tl, s1[w]); }
    out[0] = tc;
    out[1] = tl;
  }
}

// ---------------------------------------------------------------------------
// DISSIP (M3:3143-3173)
// ---------------------------------------------------------------------------
__global__ void k_dissip_div(DisArgs A, Geo g, int e0) {
  int ny = g.n[1], nz = g.n[2];
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long plane = 12L * (ny + 2) * (nz + 2);
  long tot = (2 - e0) * plane;
  if (t >= tot) return;
  int e = e0 + (int)(t / plane);
  long r = t % plane;
  int i = (int)(r % 12);
  int j = (int)((r / 12) % (ny + 2));
  int k = (int)(r / (12L * (ny + 2)));
  int x = i + (e == 0 ? 0 : g.n[0] - 10);
  int c = ix3(g, x, j, k);
  int sy = g.st[1], sz = g.st[2];
  A.divv[(long)e * plane + r] =
      (A.U[1][c + 1] - A.U[1][c - 1] + A.U[2][c + sy] - A.U[2][c - sy] + A.U[3][c + sz] - A.U[3][c - sz]) * 0.25;
}

__global__ void k_dissip_upd(DisArgs A, Geo g, int e0) {
  int ny = g.n[1], nz = g.n[2];
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long cnt = 10L * ny * nz;
  if (t >= (2 - e0) * cnt) return;
  int e = e0 + (int)(t / cnt);
  long r = t % cnt;
  int i = 1 + (int)(r % 10);
  int j = 1 + (int)((r / 10) % ny);
  int k = 1 + (int)(r / (10L * ny));
  long plane = 12L * (ny + 2) * (nz + 2);
  const double *dv = A.divv + e * plane;
  auto D = [&](int a, int bb, int cc) { return dv[a + 12L * (bb + (long)(ny + 2) * cc)]; };
  int c = ix3(g, i + (e == 0 ? 0 : g.n[0] - 10), j, k);
  A.U[1][c] = A.U[1][c] + 0.1 * (D(i + 1, j, k) - D(i - 1, j, k));
  A.U[2][c] = A.U[2][c] + 0.1 * (D(i, j + 1, k) - D(i, j - 1, k));
  A.U[3][c] = A.U[3][c] + 0.1 * (D(i, j, k + 1) - D(i, j, k - 1));
}

// Plain streaming copy (8-byte words, grid-stride): the bandwidth and
// hardware-counter reference for a kernel that moves a known byte count.
__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

} // namespace mcf
