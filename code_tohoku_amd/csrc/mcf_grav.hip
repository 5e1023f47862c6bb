// mcf_grav.hip -- self-gravity GRAVTY (3DMHD-CfCA/selfG_MPIfull8.f90, "SG8")
// for gfx950, FP64.
//
// Split levels (one block of nodes 0..n per reference rank): every operator
// is one launch over all ranks' blocks.  A halo exchange (BCsgr_MPI) is a
// single launch that pulls each ghost node's final value from where the
// reference's six sequential sendrecv phases (x low, x high, y low, y high,
// z low, z high; x ends closed, y and z periodic) would take it: walking the
// phases backwards from z, a ghost coordinate with its phase active moves to
// the neighbouring rank's opposite interior node.  The node it lands on is
// never written by any phase, so the pull reads pre-exchange values and
// one launch reproduces the reference's result exactly.
//
// Gauss-Seidel (relaxMPI, relax) runs over each block in lexicographic order
// in the reference.  Updating all nodes of one hyperplane i+j+k = t at once,
// planes in increasing t, gives the same values: a node's lower neighbours
// (and the wrapped-around upper neighbour of the whole-domain levels) lie on
// earlier planes, its upper neighbours on later ones, exactly as in the
// sequential order.  One workgroup sweeps one block, a barrier per plane.
//
// The whole-domain levels 1..NGcr (at most 17^3 nodes) run as one
// workgroup (k_grav_coarse) with the potential, right-hand side and
// residual of all levels in LDS (3 x 45 KB), the reference's level-stacked
// layout kept so that the residual's reads past a level's ends hit the same
// neighbouring values.
//
// PB's boundary potential: the x planes' y-z Fourier modes by a radix-2 LDS
// FFT of the plane packed as nn1/2 complex values per row, then the
// reference's real-FFT unpacking (rlft3, SG8:1466-1512) including its
// single-precision twiddles and the in-place double pass over the middle
// column; the Green's-function weights and the sum over planes per mode;
// the same unpacking and FFT backwards.
#include "mcf_grav.h"

namespace mcf {

__device__ __forceinline__ void g_rank(const GravGeo &G, int r, int c[3]) {
  c[0] = r % G.S[0];
  c[2] = r / (G.S[0] * G.S[1]);
  c[1] = r / G.S[0] - G.S[1] * c[2];
}
__device__ __forceinline__ int g_rid(const GravGeo &G, int a, int b, int c) { return a + G.S[0] * (b + G.S[1] * c); }
// Skewed block layout, hyperplane-major: node (i,j,k) at
// j + J (k + K (i+j+k)), J = n1+1, K = n2+1, so that one hyperplane
// i+j+k = t is one contiguous J*K slab, rows of j; the upper neighbours
// sit at +JK (x), +1+JK (y), +J+JK (z), the lower ones at the negatives.
// Slots with i outside 0..n0 are never used.
__device__ __forceinline__ long g_idx(const int *n, int i, int j, int k) {
  return j + (long)(n[1] + 1) * (k + (long)(n[2] + 1) * (i + j + k));
}
__device__ __forceinline__ long g_tid() { return blockIdx.x * (long)blockDim.x + threadIdx.x; }

// Node q (0 <= q < A B) of an A x B rectangle of nodes (a, b) in
// anti-diagonal order: a+b ascending, then b.  With a = i-1, b = j-1 of one
// k plane, neighbouring q are neighbouring slots of the skewed layout (j up,
// i down, i+j+k fixed).  The diagonals' lengths rise 1..m, stay at m, fall
// m-1..1 (m = min(A, B)); the rising and falling parts invert a triangular
// number, the float square root corrected to the exact integer.
__device__ __forceinline__ int g_tri(int q) { // largest s with s(s+1)/2 <= q
  int s = (int)((__builtin_amdgcn_sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
  while (s * (s + 1) / 2 > q) --s;
  while ((s + 1) * (s + 2) / 2 <= q) ++s;
  return s;
}
__device__ __forceinline__ void g_diag(int A, int B, int q, int &a, int &b) {
  const int m = min(A, B), M = max(A, B), T1 = m * (m + 1) / 2, F = (M - m) * m;
  int s, o;
  if (q < T1) {
    s = g_tri(q);
    o = q - s * (s + 1) / 2;
  } else if (q < T1 + F) {
    s = m + (q - T1) / m;
    o = (q - T1) % m;
  } else {
    const int r = A * B - 1 - q, u = g_tri(r);
    s = A + B - 2 - u;
    o = u - (r - u * (u + 1) / 2);
  }
  b = max(0, s - (A - 1)) + o;
  a = s - b;
}
// node q of a block's nodes 1..n0 x 1..n1 x 1..n2: k planes, each in
// anti-diagonal order
__device__ __forceinline__ void g_node(const int *n, long q, int &i, int &j, int &k) {
  const long AB = (long)n[0] * n[1];
  k = 1 + (int)(q / AB);
  int a, b;
  g_diag(n[0], n[1], (int)(q - (k - 1) * AB), a, b);
  i = 1 + a;
  j = 1 + b;
}

__global__ void k_grav_zero(double *a, long n) {
  for (long t = g_tid(); t < n; t += (long)gridDim.x * blockDim.x) a[t] = 0.0;
}

// mglin's right-hand side at the finest level (SG8:163-167): 4 pi G rho on
// every node 0..C+1 of every rank, the density ghosts included
__global__ void k_grav_rhs(GravGeo G, const double *rho, Geo g, double *rhs2) {
  const int L = G.NGL;
  const int *n = G.n[L];
  const long B = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1), t = g_tid();
  if (t >= B * G.npe) return;
  const int r = (int)(t / B);
  long q = t - r * B;
  const int i = (int)(q % (n[0] + 1));
  q /= n[0] + 1;
  const int j = (int)(q % (n[1] + 1)), k = (int)(q / (n[1] + 1));
  int c[3];
  g_rank(G, r, c);
  rhs2[G.off[L] + r * G.bsz[L] + g_idx(n, i, j, k)] =
      rho[ix3(g, c[0] * G.C[0] + i, c[1] * G.C[1] + j, c[2] * G.C[2] + k)] * GRAV_G4PI;
}

// rstrctMPI (SG8:385-456) into split level lc from lc+1, coarse nodes
// 1..nc of every rank: the first node of the first rank and the last node
// of the last rank along an axis are copied, the other inner nodes take
// 1/2 + 6 x 1/12; mode 1 zeroes the x-end nodes.  The exchange follows as
// its own launch.
__global__ void k_grav_rstrct(GravGeo G, double *uc2, const double *uf2, int lc, int mode) {
  const int *nc = G.n[lc], *nf = G.n[lc + 1];
  const long cnt = (long)nc[0] * nc[1] * nc[2], t = g_tid();
  if (t >= cnt * G.npe) return;
  const int r = (int)(t / cnt);
  int ic[3];
  g_node(nc, t - r * cnt, ic[0], ic[1], ic[2]);
  int c[3];
  g_rank(G, r, c);
  const double *uf = uf2 + G.off[lc + 1] + r * G.bsz[lc + 1];
  double *uc = uc2 + G.off[lc] + r * G.bsz[lc];
  bool copy = false, inner = true;
  for (int d = 0; d < 3; ++d) {
    const bool first = c[d] == 0, last = c[d] == G.S[d] - 1;
    if ((first && ic[d] == 1) || (last && ic[d] == nc[d])) copy = true;
    if (ic[d] < (first ? 2 : 1) || ic[d] > nc[d] - 1) inner = false;
  }
  const int fi = 2 * ic[0] - 1, fj = 2 * ic[1] - 1, fk = 2 * ic[2] - 1;
  const long p = g_idx(nf, fi, fj, fk), J = nf[1] + 1, JK = J * (nf[2] + 1);
  const long sx = JK, sy = 1 + JK, sz = J + JK;
  double v = 0.0;
  bool write = true;
  if (copy) {
    v = uf[p];
  } else if (inner) {
    const double w = 1.0 / 12.0;
    v = 0.5 * uf[p] + w * (uf[p + sx] + uf[p - sx] + uf[p + sy] + uf[p - sy] + uf[p + sz] + uf[p - sz]);
  } else {
    write = false;
  }
  if (mode == 1 && ((c[0] == 0 && ic[0] == 1) || (c[0] == G.S[0] - 1 && ic[0] == nc[0]))) {
    v = 0.0;
    write = true;
  }
  if (write) uc[g_idx(nc, ic[0], ic[1], ic[2])] = v;
}

// Where BCsgr_MPI's six phases take ghost node (i,j,k) of rank c from: walks
// the phases backwards (z, y, x); returns false when no phase moves it.
__device__ __forceinline__ bool g_resolve(const GravGeo &G, const int *n, int c[3], int &i, int &j, int &k, int flags) {
  const int i0 = i, j0 = j, k0 = k;
  if (k == 0 && (flags & 16)) { c[2] = (c[2] + G.S[2] - 1) % G.S[2]; k = n[2] - 1; }
  else if (k == n[2] && (flags & 32)) { c[2] = (c[2] + 1) % G.S[2]; k = 1; }
  if (j == 0 && (flags & 4)) { c[1] = (c[1] + G.S[1] - 1) % G.S[1]; j = n[1] - 1; }
  else if (j == n[1] && (flags & 8)) { c[1] = (c[1] + 1) % G.S[1]; j = 1; }
  if (i == 0 && (flags & 1) && c[0] > 0) { c[0] -= 1; i = n[0] - 1; }
  else if (i == n[0] && (flags & 2) && c[0] < G.S[0] - 1) { c[0] += 1; i = 1; }
  return i != i0 || j != j0 || k != k0;
}

// Ghost node q (0 <= q < grav_nghost(n)) of a block: faces counted once,
// z planes whole, y planes without their z rows, x planes without both.  A
// z plane is walked as (i+j, j) with j fastest, so that neighbouring threads
// touch neighbouring slots of the skewed layout (there and at the source);
// the slots of that walk with i outside 0..n0 are not nodes (false).
__device__ __forceinline__ bool g_ghost(const int *n, int q, int &i, int &j, int &k) {
  const int gz = 2 * (n[0] + n[1] + 1) * (n[1] + 1), gy = 2 * (n[0] + 1) * (n[2] - 1);
  if (q < gz) {
    const int h = (n[0] + n[1] + 1) * (n[1] + 1);
    const int s = q / h, m = q - s * h;
    j = m % (n[1] + 1);
    i = m / (n[1] + 1) - j;
    k = s ? n[2] : 0;
    return i >= 0 && i <= n[0];
  }
  if (q < gz + gy) {
    q -= gz;
    const int h = (n[0] + 1) * (n[2] - 1);
    const int s = q / h, m = q - s * h;
    i = m % (n[0] + 1);
    k = 1 + m / (n[0] + 1);
    j = s ? n[1] : 0;
    return true;
  }
  q -= gz + gy;
  const int h = (n[1] - 1) * (n[2] - 1);
  const int s = q / h, m = q - s * h;
  j = 1 + m % (n[1] - 1);
  k = 1 + m / (n[1] - 1);
  i = s ? n[0] : 0;
  return true;
}

// BCsgr_MPI (SG8:298-336) at split level l.  flags: bit 0 x low, 1 x high,
// 2 y low, 3 y high, 4 z low, 5 z high (the reference's lx1..lz2).  One
// thread per ghost node (a node with a coordinate 0 or n); faces counted
// once: z planes whole, y planes without their z rows, x planes without both.
__global__ void k_grav_exchange(GravGeo G, double *u2, int l, int flags) {
  const int *n = G.n[l];
  const long per = grav_nghost(n), t = g_tid();
  if (t >= per * G.npe) return;
  const int r = (int)(t / per);
  int i, j, k;
  if (!g_ghost(n, (int)(t - r * per), i, j, k)) return;
  int c[3];
  g_rank(G, r, c);
  int si = i, sj = j, sk = k;
  if (!g_resolve(G, n, c, si, sj, sk, flags)) return;
  const int rs = g_rid(G, c[0], c[1], c[2]);
  u2[G.off[l] + r * G.bsz[l] + g_idx(n, i, j, k)] = u2[G.off[l] + rs * G.bsz[l] + g_idx(n, si, sj, sk)];
}

// relaxMPI (SG8:740-870): one Gauss-Seidel sweep of nodes 1..n-1 of one
// rank's block per workgroup, as hyperplanes i+j+k = t; node 1 of the first
// x rank is the boundary, kept by the reference's arithmetic mask.
__global__ __launch_bounds__(GRAV_RELAX_THREADS) void k_grav_relax(GravGeo G, double *phi2, const double *rhs2, int l) {
  const int r = blockIdx.x;
  const int *n = G.n[l];
  int c[3];
  g_rank(G, r, c);
  double *u = phi2 + G.off[l] + r * G.bsz[l];
  const double *rhs = rhs2 + G.off[l] + r * G.bsz[l];
  const double w = 1.0 / 6.0;
  const double h = G.Lbox / (double)((n[0] - 1) * G.S[0]);
  const double h2 = h * h;
  const int li = c[0] == 0 ? 1 : 0;
  const int m0 = n[0] - 1, m1 = n[1] - 1, m2 = n[2] - 1;
  const long J = n[1] + 1, JK = J * (n[2] + 1), sx = JK, sy = 1 + JK, sz = J + JK;
  for (int t = 3; t <= m0 + m1 + m2; ++t) {
    const int klo = max(1, t - m0 - m1), khi = min(m2, t - 2);
    const int cnt = m1 * (khi - klo + 1);
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
      const int j = 1 + q % m1, k = klo + q / m1, i = t - j - k;
      if (i < 1 || i > m0) continue;
      const long p = g_idx(n, i, j, k);
      const int ifl = li * (1 / i);
      const double s = u[p + sx] + u[p - sx] + u[p + sy] + u[p - sy] + u[p + sz] + u[p - sz] - h2 * rhs[p];
      u[p] = w * s * (double)(1 - ifl) + (0.5 + copysign(0.5, (double)ifl - 0.5)) * u[p];
    }
    __syncthreads();
  }
}

// Barrier that orders LDS accesses only: the workgroup's threads exchange
// values through LDS, and global loads still in flight (the next plane's
// prefetch) need not complete first, as they would under __syncthreads.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The same sweep with each thread owning fixed (j,k) columns of the block
// (PP of them) and walking i = t-j-k along them: the node's lower x
// neighbour is the thread's own result of the previous plane (a register),
// its lower y and z neighbours are the neighbouring columns' results of the
// previous plane (LDS, two buffers by plane parity) or, on the j = 0 and
// k = 0 faces, fixed ghost values staged in LDS at the start.  Everything
// else a node reads is a value from before the sweep -- its upper
// neighbours, the right-hand side, and its own old value (the upper x
// neighbour loaded one plane earlier) -- and is loaded D planes ahead into a
// ring of registers, so that D planes of loads are in flight behind the
// plane being computed: a node on plane t+D+1 is written only after the
// barrier that ends plane t+D, by which time its old value has arrived.
// LMAX bounds (n1+1)(n2+1), (n0+1)(n2+1) and (n0+1)(n1+1).
//
// Sweeps in a row exchange halos through boundary planes instead of a
// separate launch: with Eout, every node the sweep computes on the six
// planes next to the block's faces (node 1 or n-1 along an axis) is also
// written there; with Ein, the ghost nodes are first set from the
// neighbours' planes of the previous sweep, resolved as BCsgr_MPI would
// resolve them (a source at a closed x end, never written, is read from the
// neighbour's block itself).  Plane layout per block: 6 planes of EP values
// (x planes indexed by (j,k), y by (i,k), z by (i,j)); the instances with
// PL false have no plane code.
__device__ __forceinline__ long g_plane_at(const int *n, long EP, int i, int j, int k) {
  if (i == 1) return 0 * EP + j + (long)(n[1] + 1) * k;
  if (i == n[0] - 1) return 1 * EP + j + (long)(n[1] + 1) * k;
  if (j == 1) return 2 * EP + i + (long)(n[0] + 1) * k;
  if (j == n[1] - 1) return 3 * EP + i + (long)(n[0] + 1) * k;
  if (k == 1) return 4 * EP + i + (long)(n[0] + 1) * j;
  return 5 * EP + i + (long)(n[0] + 1) * j;
}

template <int PP, int LMAX, int D, bool PL>
__global__ __launch_bounds__(GRAV_RELAX_THREADS) void k_grav_relax_col(GravGeo G, double *phi2, const double *rhs2,
                                                                       int l, const double *Ein, double *Eout, long EP) {
  // two plane buffers, then the fixed ghost faces j = 0 and k = 0
  __shared__ double lds[4 * LMAX];
  const int r = blockIdx.x;
  const int *n = G.n[l];
  int c[3];
  g_rank(G, r, c);
  double *u = phi2 + G.off[l] + r * G.bsz[l];
  const double *rhs = rhs2 + G.off[l] + r * G.bsz[l];
  if (PL && Ein) {
    const int ng = (int)grav_nghost(n);
    for (int q = threadIdx.x; q < ng; q += blockDim.x) {
      int i, j, k, cc[3] = {c[0], c[1], c[2]};
      if (!g_ghost(n, q, i, j, k)) continue;
      int si = i, sj = j, sk = k;
      if (!g_resolve(G, n, cc, si, sj, sk, 63)) continue;
      const int rs = g_rid(G, cc[0], cc[1], cc[2]);
      u[g_idx(n, i, j, k)] = (si == 0 || si == n[0]) ? phi2[G.off[l] + rs * G.bsz[l] + g_idx(n, si, sj, sk)]
                                                     : Ein[rs * 6 * EP + g_plane_at(n, EP, si, sj, sk)];
    }
    __syncthreads();
  }
  const double w = 1.0 / 6.0;
  const double h = G.Lbox / (double)((n[0] - 1) * G.S[0]);
  const double h2 = h * h;
  const int li = c[0] == 0 ? 1 : 0;
  const int m0 = n[0] - 1, m1 = n[1] - 1, m2 = n[2] - 1, LJ = n[1] + 1, T = m0 + m1 + m2, I1 = n[0] + 1;
  const long J = n[1] + 1, JK = J * (n[2] + 1), sx = JK, sy = 1 + JK, sz = J + JK;
  double *gyf = lds + 2 * LMAX, *gzf = lds + 3 * LMAX;
  for (int q = threadIdx.x; q < I1 * (n[2] + 1); q += blockDim.x) gyf[q] = u[g_idx(n, q % I1, 0, q / I1)];
  for (int q = threadIdx.x; q < I1 * (n[1] + 1); q += blockDim.x) gzf[q] = u[g_idx(n, q % I1, q / I1, 0)];
  __syncthreads();
  int pj[PP], pk[PP];
  bool on[PP];
  // per column: the lower x neighbour (the ghost u(0,j,k), then the last
  // result) and the old value of the node to compute (u(1,j,k) at first)
  double prev[PP], fon[PP];
  // ring slot s: the loads of the node on plane t = 3 + s (mod D)
  double rx[D][PP], ry[D][PP], rz[D][PP], rr[D][PP];
#pragma unroll
  for (int p = 0; p < PP; ++p) {
    const int q = threadIdx.x + p * blockDim.x;
    on[p] = q < m1 * m2;
    pj[p] = on[p] ? 1 + q % m1 : 1;
    pk[p] = on[p] ? 1 + q / m1 : 1;
    const long q1 = g_idx(n, 1, pj[p], pk[p]);
    prev[p] = u[q1 - sx];
    fon[p] = u[q1];
  }
  // 32-bit byte offsets from the block's (uniform) base: one register per address
  const unsigned bx = 8u * (unsigned)sx, by = 8u * (unsigned)sy, bz = 8u * (unsigned)sz;
  auto at = [](const double *base, unsigned off) { return *(const double *)((const char *)base + off); };
  auto fetch = [&](int s, int p, int i) {
    // with one column per thread only the columns whose node i exists load
    // (the sweep is bound by the instructions its waves issue more than by
    // the latency of these loads); with four, every column loads, a column
    // off its nodes at an unused slot of the plane (held to 3..T) on the
    // same cache lines as the others
    if (PP == 1 && (!on[p] || i < 1 || i > m0)) return;
    const int tq = min(i + pj[p] + pk[p], T);
    const unsigned q = 8u * (unsigned)(pj[p] + LJ * (pk[p] + (n[2] + 1) * tq));
    rx[s][p] = at(u, q + bx);
    ry[s][p] = at(u, q + by);
    rz[s][p] = at(u, q + bz);
    rr[s][p] = at(rhs, q);
  };
#pragma unroll
  for (int s = 0; s < D; ++s)
#pragma unroll
    for (int p = 0; p < PP; ++p) fetch(s, p, 3 + s - pj[p] - pk[p]);
  // whole groups of D planes: on the planes past T every column's node is
  // past its end (i > n0-1), so they only wait at the barrier, and the loop
  // body has one shape
  for (int tb = 3; tb <= T; tb += D) {
#pragma unroll
    for (int s = 0; s < D; ++s) {
      const int t = tb + s;
      double *Lc = lds + (t & 1) * LMAX;
      const double *Lp = lds + ((t + 1) & 1) * LMAX;
#pragma unroll
      for (int p = 0; p < PP; ++p) {
        const int i = t - pj[p] - pk[p];
        if (on[p] && i >= 1 && i <= m0) {
          const double ym = pj[p] == 1 ? gyf[i + I1 * pk[p]] : Lp[(pj[p] - 1) + LJ * pk[p]];
          const double zm = pk[p] == 1 ? gzf[i + I1 * pj[p]] : Lp[pj[p] + LJ * (pk[p] - 1)];
          const int ifl = li * (1 / i);
          const double sum = rx[s][p] + prev[p] + ry[s][p] + ym + rz[s][p] + zm - h2 * rr[s][p];
          const double v = w * sum * (double)(1 - ifl) + (0.5 + copysign(0.5, (double)ifl - 0.5)) * fon[p];
          u[g_idx(n, i, pj[p], pk[p])] = v;
          Lc[pj[p] + LJ * pk[p]] = v;
          prev[p] = v;
          if (PL && Eout) {
            double *E = Eout + r * 6 * EP;
            if (i == 1) E[0 * EP + pj[p] + LJ * pk[p]] = v;
            if (i == m0) E[1 * EP + pj[p] + LJ * pk[p]] = v;
            if (pj[p] == 1) E[2 * EP + i + (long)(n[0] + 1) * pk[p]] = v;
            if (pj[p] == m1) E[3 * EP + i + (long)(n[0] + 1) * pk[p]] = v;
            if (pk[p] == 1) E[4 * EP + i + (long)(n[0] + 1) * pj[p]] = v;
            if (pk[p] == m2) E[5 * EP + i + (long)(n[0] + 1) * pj[p]] = v;
          }
        }
        if (i >= 1) fon[p] = rx[s][p];
        fetch(s, p, i + D);
      }
      lds_barrier();
    }
  }
}
template __global__ void k_grav_relax_col<1, 1156, GRAV_RELAX_D1, true>(GravGeo, double *, const double *, int,
                                                                        const double *, double *, long);
template __global__ void k_grav_relax_col<1, 1156, GRAV_RELAX_D1, false>(GravGeo, double *, const double *, int,
                                                                         const double *, double *, long);
// (four columns per thread: the instance with the plane code, called
// without planes, measured faster than the one without: 605 against 660 us
// per sweep on 65^3-node blocks)
template __global__ void k_grav_relax_col<4, 4356, GRAV_RELAX_D4, true>(GravGeo, double *, const double *, int,
                                                                        const double *, double *, long);

// residMPI (SG8:918-952): nodes 1..n-1, then zero at the x-end nodes
// (nodes in g_node's order, so that neighbouring threads touch neighbouring
// addresses of the skewed layout)
__global__ void k_grav_resid(GravGeo G, double *res2, const double *phi2, const double *rhs2, int l) {
  const int *n = G.n[l];
  const long cnt = (long)n[0] * n[1] * n[2], t = g_tid();
  if (t >= cnt * G.npe) return;
  const int r = (int)(t / cnt);
  int i, j, k;
  g_node(n, t - r * cnt, i, j, k);
  int c[3];
  g_rank(G, r, c);
  const long base = G.off[l] + r * G.bsz[l], p = g_idx(n, i, j, k);
  if ((c[0] == 0 && i == 1) || (c[0] == G.S[0] - 1 && i == n[0])) {
    res2[base + p] = 0.0;
    return;
  }
  if (i > n[0] - 1 || j > n[1] - 1 || k > n[2] - 1) return;
  const double *u = phi2 + base;
  const double h = G.Lbox / (double)((n[0] - 1) * G.S[0]);
  const double h2i = 1.0 / (h * h);
  const long J = n[1] + 1, JK = J * (n[2] + 1), sx = JK, sy = 1 + JK, sz = J + JK;
  res2[base + p] =
      -h2i * (u[p + sx] + u[p - sx] + u[p + sy] + u[p - sy] + u[p + sz] + u[p - sz] - 6.0 * u[p]) + rhs2[base + p];
}

// Prolongation of interpMPI / interp (SG8:500-599): injection at odd
// nodes, then the mean of the two neighbours along x (odd y, z), along y
// (odd z) and along z, in that order; each node's value is formed directly
// from the coarse nodes with the same operations.
template <class F>
__device__ __forceinline__ double pro_w(const F &C, int i, int j, int k) { // j, k odd
  return (i & 1) ? C(i, j, k) : .5 * (C(i + 1, j, k) + C(i - 1, j, k));
}
template <class F>
__device__ __forceinline__ double pro_v(const F &C, int i, int j, int k) { // k odd
  return (j & 1) ? pro_w(C, i, j, k) : .5 * (pro_w(C, i, j + 1, k) + pro_w(C, i, j - 1, k));
}
template <class F>
__device__ __forceinline__ double pro_node(const F &C, int i, int j, int k) {
  return (k & 1) ? pro_v(C, i, j, k) : .5 * (pro_v(C, i, j, k + 1) + pro_v(C, i, j, k - 1));
}

// interpMPI into split level lf (nodes 1..nf of every rank); mode 1 puts
// the end-plane potential on the x-end nodes.  The exchange (low ghosts
// only) follows as its own launch.
__global__ void k_grav_prolong(GravGeo G, double *uf2, const double *uc2, int lf, int mode, const double *bphi2) {
  const int *nf = G.n[lf], *nc = G.n[lf - 1];
  const long cnt = (long)nf[0] * nf[1] * nf[2], t = g_tid();
  if (t >= cnt * G.npe) return;
  const int r = (int)(t / cnt);
  int i, j, k;
  g_node(nf, t - r * cnt, i, j, k);
  int c[3];
  g_rank(G, r, c);
  const double *uc = uc2 + G.off[lf - 1] + r * G.bsz[lf - 1];
  double v;
  const int side = (c[0] == 0 && i == 1) ? 0 : ((c[0] == G.S[0] - 1 && i == nf[0]) ? 1 : -1);
  if (mode == 1 && side >= 0) {
    v = bphi2[G.boff[lf] + (2L * r + side) * (long)(nf[1] + 1) * (nf[2] + 1) + j + (long)(nf[1] + 1) * k];
  } else {
    auto C = [&](int a, int b, int e) { return uc[g_idx(nc, (a + 1) / 2, (b + 1) / 2, (e + 1) / 2)]; };
    v = pro_node(C, i, j, k);
  }
  uf2[G.off[lf] + r * G.bsz[lf] + g_idx(nf, i, j, k)] = v;
}

// addintMPI's update (SG8:602-624): every node 0..n, uf += res
// (over the slots of the skewed layout, the unused ones skipped: a third of
// the slots are nodes)
__global__ void k_grav_add(GravGeo G, double *uf2, const double *res2, int l) {
  const int *n = G.n[l];
  const long n_ = G.bsz[l] * G.npe, t = g_tid();
  if (t >= n_) return;
  const int J = n[1] + 1, K = n[2] + 1;
  const int s = (int)(t % G.bsz[l]);
  const int j = s % J, k = (s / J) % K, i = s / (J * K) - j - k;
  if (i < 0 || i > n[0]) return;
  uf2[G.off[l] + t] = uf2[G.off[l] + t] + res2[G.off[l] + t];
}

// collect (SG8:339-366): the level-NGcr rank blocks into the whole-domain
// level NGcr (2^NGcr+1 nodes per axis, node 1 first, 1-based offset p1)
__global__ void k_grav_collect(GravGeo G, double *u1, const double *u2) {
  const int L = G.NGcr, N1 = (1 << L) + 1;
  const int *m = G.n[L];
  const long t = g_tid();
  if (t >= (long)N1 * N1 * N1) return;
  const int gi = 1 + (int)(t % N1), gj = 1 + (int)((t / N1) % N1), gk = 1 + (int)(t / ((long)N1 * N1));
  const int g3[3] = {gi, gj, gk};
  int c[3], lo[3];
  for (int d = 0; d < 3; ++d) {
    c[d] = min((g3[d] - 1) / (m[d] - 1), G.S[d] - 1);
    lo[d] = g3[d] - c[d] * (m[d] - 1);
  }
  const int r = g_rid(G, c[0], c[1], c[2]);
  u1[G.p1[L] + t] = u2[G.off[L] + r * G.bsz[L] + g_idx(m, lo[0], lo[1], lo[2])];
}

// divide (SG8:369-382): whole-domain level NGcr into the rank blocks, the
// node-0 planes of the first rank along each axis left as they are
__global__ void k_grav_divide(GravGeo G, double *u2, const double *u1) {
  const int L = G.NGcr, N1 = (1 << L) + 1;
  const int *m = G.n[L];
  const long B = (long)(m[0] + 1) * (m[1] + 1) * (m[2] + 1), t = g_tid();
  if (t >= B * G.npe) return;
  const int r = (int)(t / B);
  long q = t - r * B;
  int id[3];
  id[0] = (int)(q % (m[0] + 1));
  q /= m[0] + 1;
  id[1] = (int)(q % (m[1] + 1));
  id[2] = (int)(q / (m[1] + 1));
  int c[3];
  g_rank(G, r, c);
  int g3[3];
  for (int d = 0; d < 3; ++d) {
    if (c[d] == 0 && id[d] == 0) return;
    g3[d] = c[d] * (m[d] - 1) + id[d];
  }
  u2[G.off[L] + r * G.bsz[L] + g_idx(m, id[0], id[1], id[2])] =
      u1[G.p1[L] + (g3[0] - 1) + (long)N1 * ((g3[1] - 1) + (long)N1 * (g3[2] - 1))];
}

// ---------------------------------------------------------------------------
// whole-domain levels, one workgroup
// ---------------------------------------------------------------------------
struct Coarse {
  const GravGeo &G;
  const GravBuf &B;
  double *phi, *rhs, *res; // LDS, 1-based
};

__device__ __forceinline__ long c_at(long off, int n, int i, int j, int k) {
  return off + (i - 1) + (long)n * ((j - 1) + (long)n * (k - 1));
}

// rstrct (SG8:459-497): inner nodes weighted, the six faces copied, mode 1
// zeroes the x faces
__device__ void c_rstrct(double *uc, long offc, const double *uf, long offf, int nc, int mode) {
  const int nf = 2 * nc - 1;
  const long cnt = (long)nc * nc * nc;
  for (long q = threadIdx.x; q < cnt; q += blockDim.x) {
    const int ic = 1 + (int)(q % nc), jc = 1 + (int)((q / nc) % nc), kc = 1 + (int)(q / ((long)nc * nc));
    const bool face = ic == 1 || ic == nc || jc == 1 || jc == nc || kc == 1 || kc == nc;
    const long p = c_at(offf, nf, 2 * ic - 1, 2 * jc - 1, 2 * kc - 1);
    double v;
    if (face) {
      v = uf[p];
    } else {
      const double w = 1.0 / 12.0;
      v = 0.5 * uf[p] + w * (uf[p + 1] + uf[p - 1] + uf[p + nf] + uf[p - nf] + uf[p + (long)nf * nf] + uf[p - (long)nf * nf]);
    }
    if (mode == 1 && (ic == 1 || ic == nc)) v = 0.0;
    uc[c_at(offc, nc, ic, jc, kc)] = v;
  }
  __syncthreads();
}

// interp (SG8:557-599); mode 1 sets the x faces from the end-plane potential
__device__ void c_interp(const Coarse &S, double *uf, long offf, const double *uc, long offc, int nf, long np, int mode) {
  const int nc = nf / 2 + 1;
  const long cnt = (long)nf * nf * nf;
  for (long q = threadIdx.x; q < cnt; q += blockDim.x) {
    const int i = 1 + (int)(q % nf), j = 1 + (int)((q / nf) % nf), k = 1 + (int)(q / ((long)nf * nf));
    double v;
    if (mode == 1 && (i == 1 || i == nf)) {
      v = S.B.bphi1[(i == 1 ? 0 : S.G.needb1 + 1) + np + (j - 1) + (long)nf * (k - 1)];
    } else {
      auto C = [&](int a, int b, int e) { return uc[c_at(offc, nc, (a + 1) / 2, (b + 1) / 2, (e + 1) / 2)]; };
      v = pro_node(C, i, j, k);
    }
    uf[c_at(offf, nf, i, j, k)] = v;
  }
  __syncthreads();
}

// relax (SG8:873-915): x nodes 2..n-1, y and z wrapped with period n
// Hyperplanes i+j+k = t; each thread keeps one (j,k) column for the whole
// sweep (n^2 <= 289 columns, fewer than the workgroup's threads), with its
// wrapped neighbours' offsets formed once.
static_assert(((1 << GRAV_MAX_NGCR) + 1) * ((1 << GRAV_MAX_NGCR) + 1) <= GRAV_COARSE_THREADS,
              "one thread per column of the whole-domain levels");
__device__ void c_relax(const Coarse &S, double *u, long off, const double *rhs, int n) {
  const double w = 1.0 / 6.0;
  const double h = S.G.Lbox / (double)(n - 1);
  const double h2 = h * h;
  const int q = threadIdx.x;
  const bool on = q < n * n;
  const int j = 1 + q % n, k = 1 + q / n;
  const int jm = j == 1 ? n : j - 1, jp = j == n ? 1 : j + 1;
  const int km = k == 1 ? n : k - 1, kp = k == n ? 1 : k + 1;
  const long base = c_at(off, n, 0, j, k);
  const long oyp = (long)n * (jp - j), oym = (long)n * (jm - j), ozp = (long)n * n * (kp - k),
             ozm = (long)n * n * (km - k);
  for (int t = 4; t <= (n - 1) + 2 * n; ++t) {
    const int i = t - j - k;
    if (on && i >= 2 && i <= n - 1) {
      const long p = base + i;
      u[p] = w * (u[p + 1] + u[p - 1] + u[p + oyp] + u[p + oym] + u[p + ozp] + u[p + ozm] - h2 * rhs[p]);
    }
    __syncthreads();
  }
}

// resid (SG8:955-971): the y and z neighbours at flat offsets +-n, +-n^2,
// past the level's ends for its first and last rows and planes
__device__ void c_resid(const Coarse &S, double *res, const double *u, const double *rhs, long off, int n) {
  const double h = S.G.Lbox / (double)(n - 1);
  const double h2i = 1.0 / (h * h);
  const long cnt = (long)n * n * n, nn = (long)n * n;
  for (long q = threadIdx.x; q < cnt; q += blockDim.x) {
    const int i = 1 + (int)(q % n);
    const long p = off + q;
    if (i == 1 || i == n) {
      res[p] = 0.0;
    } else {
      res[p] = -h2i * (u[p + 1] + u[p - 1] + u[p + n] + u[p - n] + u[p + nn] + u[p - nn] - 6.0 * u[p]) + rhs[p];
    }
  }
  __syncthreads();
}

__device__ void c_fill0(double *u, long off, int n) {
  const long cnt = (long)n * n * n;
  for (long q = threadIdx.x; q < cnt; q += blockDim.x) u[off + q] = 0.0;
  __syncthreads();
}

__device__ void c_copy(double *dst, long off, const double *src, int n) {
  const long cnt = (long)n * n * n;
  for (long q = threadIdx.x; q < cnt; q += blockDim.x) dst[off + q] = src[off + q];
  __syncthreads();
}

__device__ __forceinline__ double w3(int j, int k, int jp, int kp) {
  return (j == jp && k == kp) ? 8.0 : ((j == jp || k == kp) ? 2.0 : 1.0);
}

// slvsml (SG8:651-679): exact 3^3 solve, zero x faces; b = 1 is slvsmlb
// (SG8:682-737) with the end-plane potential on the x faces
__device__ void c_slvsml(const Coarse &S, const double *rhs, int b) {
  const double h = 0.5 * S.G.Lbox;
  const long off = S.G.p1[1];
  double *u = S.phi;
  const int q = threadIdx.x;
  if (q < 9) {
    const int j = 1 + q % 3, k = 1 + q / 3;
    u[c_at(off, 3, 1, j, k)] = b ? S.B.bphi1[S.G.pb1[1] + (j - 1) + 3 * (k - 1)] : 0.0;
    u[c_at(off, 3, 3, j, k)] = b ? S.B.bphi1[S.G.needb1 + 1 + S.G.pb1[1] + (j - 1) + 3 * (k - 1)] : 0.0;
  }
  __syncthreads();
  if (q < 9) {
    const int j = 1 + q % 3, k = 1 + q / 3;
    double s = 0.0;
    if (b) {
      bool first = true;
      for (int side = 1; side <= 3; side += 2)
        for (int jp = 1; jp <= 3; ++jp)
          for (int kp = 1; kp <= 3; ++kp) {
            const double t = w3(j, k, jp, kp) * u[c_at(off, 3, side, jp, kp)];
            s = first ? t : s + t;
            first = false;
          }
      for (int jp = 1; jp <= 3; ++jp)
        for (int kp = 1; kp <= 3; ++kp) s = s - w3(j, k, jp, kp) * h * h * rhs[c_at(off, 3, 2, jp, kp)];
      s = 0.025 * s;
    } else {
      bool first = true;
      for (int jp = 1; jp <= 3; ++jp)
        for (int kp = 1; kp <= 3; ++kp) {
          const double t = w3(j, k, jp, kp) * rhs[c_at(off, 3, 2, jp, kp)];
          s = first ? -t : s - t;
          first = false;
        }
      s = 0.025 * h * h * s;
    }
    u[c_at(off, 3, 2, j, k)] = s;
  }
  __syncthreads();
}

// addint (SG8:627-648): res = prolongation of level l-1, phi(l) += res
__device__ void c_addint(const Coarse &S, int l) {
  const int nf = (1 << l) + 1;
  c_interp(S, S.res, S.G.p1[l], S.phi, S.G.p1[l - 1], nf, 0, 0);
  const long cnt = (long)nf * nf * nf, off = S.G.p1[l];
  for (long q = threadIdx.x; q < cnt; q += blockDim.x) S.phi[off + q] = S.phi[off + q] + S.res[off + q];
  __syncthreads();
}

// the downward leg from level top-1 (whose right-hand side is set) to 2,
// the exact 3^3 solve and the upward leg to top-1 (mglin SG8:200-270)
__device__ void c_vcycle_lower(const Coarse &S, int top) {
  const GravGeo &G = S.G;
  for (int jj = top - 1; jj >= 2; --jj) {
    const int n = (1 << jj) + 1;
    for (int p = 0; p < 5; ++p) c_relax(S, S.phi, G.p1[jj], S.rhs, n);
    c_resid(S, S.res, S.phi, S.rhs, G.p1[jj], n);
    c_rstrct(S.rhs, G.p1[jj - 1], S.res, G.p1[jj], (1 << (jj - 1)) + 1, 1);
    c_fill0(S.phi, G.p1[jj - 1], (1 << (jj - 1)) + 1);
  }
  c_slvsml(S, S.rhs, 0);
  for (int jj = 2; jj <= top - 1; ++jj) {
    c_addint(S, jj);
    for (int p = 0; p < 5; ++p) c_relax(S, S.phi, G.p1[jj], S.rhs, (1 << jj) + 1);
  }
}

// 2D restriction of an end plane (PB, SG8:1394-1460): 1/2 centre + 1/8
// per in-plane neighbour, a missing neighbour's weight moved to the centre
__device__ __forceinline__ double plane_restrict(const double *bf, long kf, long stride, int ic, int jc, int mcx, int mcy) {
  const double w = 0.125;
  const bool hl = ic > 1, hr = ic < mcx, hd = jc > 1, hu = jc < mcy;
  const double cw = (double)(4 + !hl + !hr + !hd + !hu) * w;
  const long nb[4] = {kf - 1, kf + 1, kf - stride, kf + stride};
  const bool has[4] = {hl, hr, hd, hu};
  double sum = 0.0;
  bool first = true;
  for (int qq = 0; qq < 4; ++qq)
    if (has[qq]) {
      sum = first ? bf[nb[qq]] : sum + bf[nb[qq]];
      first = false;
    }
  return cw * bf[kf] + w * sum;
}

__global__ __launch_bounds__(GRAV_COARSE_THREADS) void k_grav_coarse(GravGeo G, GravBuf B, int prog, int jst) {
  __shared__ double sphi[GRAV_NEED1_MAX + 1], srhs[GRAV_NEED1_MAX + 1], sres[GRAV_NEED1_MAX + 1];
  // all loads of the three arrays in flight at once (GRAV_COARSE_THREADS
  // threads, GRAV_NEED1_MAX + 1 entries each)
  constexpr int CU = (GRAV_NEED1_MAX + GRAV_COARSE_THREADS) / GRAV_COARSE_THREADS;
  const long need = G.need1;
  {
    double a[CU], b[CU], c[CU];
#pragma unroll
    for (int q = 0; q < CU; ++q) {
      const long p = threadIdx.x + (long)q * GRAV_COARSE_THREADS;
      if (p <= need) {
        a[q] = B.phi1[p];
        b[q] = B.rhs1[p];
        c[q] = B.res1[p];
      }
    }
#pragma unroll
    for (int q = 0; q < CU; ++q) {
      const long p = threadIdx.x + (long)q * GRAV_COARSE_THREADS;
      if (p <= need) {
        sphi[p] = a[q];
        srhs[p] = b[q];
        sres[p] = c[q];
      }
    }
  }
  __syncthreads();
  const Coarse S{G, B, sphi, srhs, sres};
  const int L = G.NGcr, N1 = (1 << L) + 1;
  if (prog == 0) {
    // crho1 down to level 1, the end-plane potential likewise, slvsmlb
    for (int l = L - 1; l >= 1; --l) c_rstrct(B.rho1, G.p1[l], B.rho1, G.p1[l + 1], (1 << l) + 1, 0);
    for (int lc = L - 1; lc >= 1; --lc) {
      const int ncu = (1 << lc) + 1, nf = 2 * ncu - 1;
      for (long q = threadIdx.x; q < 2L * ncu * ncu; q += blockDim.x) {
        const int side = (int)(q / ((long)ncu * ncu)), m = (int)(q % ((long)ncu * ncu));
        const int ic = 1 + m % ncu, jc = 1 + m / ncu;
        double *bl = B.bphi1 + side * (G.needb1 + 1);
        const long kf = (2 * ic - 2) + (long)nf * (2 * jc - 2), kc = (ic - 1) + (long)ncu * (jc - 1);
        bl[G.pb1[lc] + kc] = plane_restrict(bl + G.pb1[lc + 1], kf, nf, ic, jc, ncu, ncu);
      }
      __syncthreads();
    }
    c_slvsml(S, B.rho1, 1);
  } else if (prog == 1) {
    // FMG stage jst < NGcr: candidate solution, then two V-cycles
    const int nf = (1 << jst) + 1;
    c_interp(S, sphi, G.p1[jst], sphi, G.p1[jst - 1], nf, G.pb1[jst], 1);
    c_copy(srhs, G.p1[jst], B.rho1, nf);
    for (int cyc = 0; cyc < 2; ++cyc) {
      for (int p = 0; p < 5; ++p) c_relax(S, sphi, G.p1[jst], srhs, nf);
      c_resid(S, sres, sphi, srhs, G.p1[jst], nf);
      c_rstrct(srhs, G.p1[jst - 1], sres, G.p1[jst], (1 << (jst - 1)) + 1, 1);
      c_fill0(sphi, G.p1[jst - 1], (1 << (jst - 1)) + 1);
      c_vcycle_lower(S, jst);
      c_addint(S, jst);
      for (int p = 0; p < 5; ++p) c_relax(S, sphi, G.p1[jst], srhs, nf);
    }
  } else if (prog == 2) {
    // FMG stage NGcr, whole-domain part
    c_interp(S, sphi, G.p1[L], sphi, G.p1[L - 1], N1, G.pb1[L], 1);
    c_copy(srhs, G.p1[L], B.rho1, N1);
  } else {
    // below the coarsest split level: the collected residual restricted,
    // the lower levels, then the correction of level NGcr
    const int nc = (1 << (L - 1)) + 1;
    c_rstrct(srhs, G.p1[L - 1], sres, G.p1[L], nc, 1);
    c_fill0(sphi, G.p1[L - 1], nc);
    c_vcycle_lower(S, L);
    c_addint(S, L);
  }
  for (long p = threadIdx.x; p <= need; p += blockDim.x) {
    B.phi1[p] = sphi[p];
    B.rhs1[p] = srhs[p];
    B.res1[p] = sres[p];
  }
}

// ---------------------------------------------------------------------------
// PB: boundary potential on the x ends
// ---------------------------------------------------------------------------
// the real plane of global x index gx packed as nn1/2 complex values per
// row (y = 2k-1 real part, 2k imaginary part), rows along z
// Through an LDS tile of 64 planes x 64 rows at one z: the reads run along
// x (the planes, consecutive x within a block), the writes along the row
// pairs.  Grid (plane tiles, pair tiles of 32, nn2), 256 threads.
__global__ __launch_bounds__(256) void k_pb_pack(PbArgs P, const double *rho, Geo g, double2 *Z) {
  __shared__ double T[64][65];
  const int H = P.nn1 / 2, q = blockIdx.z;
  const int pl0 = 64 * blockIdx.x, pp0 = 32 * blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int pl = pl0 + lane;
  const int gx = pl < P.np ? P.gx[pl] : 0;
  for (int r = w; r < 64; r += 4) { // row j = 2 pp0 + 1 + r
    const int pp = pp0 + (r >> 1);
    T[r][lane] = (pl < P.np && pp < H) ? rho[ix3(g, gx, 2 * pp0 + 1 + r, q + 1)] : 0.0;
  }
  __syncthreads();
  const int e = threadIdx.x & 31, h = threadIdx.x >> 5; // pair pp0 + e, planes h, h + 8, ...
  for (int c = h; c < 64; c += 8) {
    const int p2 = pl0 + c, pp = pp0 + e;
    if (p2 < P.np && pp < H) Z[((long)p2 * P.nn2 + q) * H + pp] = make_double2(T[2 * e][c], T[2 * e + 1][c]);
  }
}

// radix-2 FFT of lines of L complex values in LDS: X_k = sum_j z_j
// exp(sign 2 pi i j k / L).  Line id -> (outer, inner); element e of the
// line at outer*outer_stride + inner*inner_stride + e*es.  A workgroup
// transforms 2^lcw lines of consecutive inner index (ninner a multiple of
// it, 2^lcw L <= 2 GRAV_FFT_MAXL): with inner_stride 1 and es > 1 (the
// columns of a plane) its loads and stores run over 2^lcw neighbouring
// columns at once.
__global__ __launch_bounds__(512) void k_pb_fft(double2 *Z, int L, long es, long inner_stride, long outer_stride,
                                                 int ninner, int sign, int lcw) {
  __shared__ double2 buf[2 * GRAV_FFT_MAXL];
  // the twiddles exp(sign 2 pi i k / L), k < L/2, once per workgroup: stage
  // len takes k = pos L/len, whose sincospi argument 2k/L is the same double
  // as 2 pos/len
  __shared__ double2 tw[GRAV_FFT_MAXL / 2];
  const int CW = 1 << lcw, H2 = L / 2;
  const long line0 = (long)blockIdx.x << lcw;
  double2 *z = Z + (line0 / ninner) * outer_stride + (line0 % ninner) * inner_stride;
  int lg = 0;
  while ((1 << lg) < L) ++lg;
  for (int idx = threadIdx.x; idx < CW * L; idx += blockDim.x) {
    const int e = idx >> lcw, c = idx & (CW - 1);
    const int rev = (int)(__brev((unsigned)e) >> (32 - lg));
    buf[c * L + rev] = z[c * inner_stride + e * es];
  }
  for (int k = threadIdx.x; k < H2; k += blockDim.x) {
    double sn, cs;
    sincospi(2.0 * (double)k / (double)L, &sn, &cs);
    tw[k] = make_double2(cs, sign > 0 ? sn : -sn);
  }
  __syncthreads();
  for (int len = 2; len <= L; len <<= 1) {
    const int half = len >> 1;
    for (int b = threadIdx.x; b < CW * H2; b += blockDim.x) {
      const int c = b / H2, bb = b - c * H2;
      const int grp = bb / half, pos = bb - grp * half;
      const int i0 = c * L + grp * len + pos, i1 = i0 + half;
      const double2 W = tw[pos * (L / len)];
      const double cs = W.x, sn = W.y;
      const double2 a = buf[i0], d = buf[i1];
      const double tr = cs * d.x - sn * d.y, ti = cs * d.y + sn * d.x;
      buf[i1] = make_double2(a.x - tr, a.y - ti);
      buf[i0] = make_double2(a.x + tr, a.y + ti);
    }
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < CW * L; idx += blockDim.x) {
    const int e = idx >> lcw, c = idx & (CW - 1);
    z[c * inner_stride + e * es] = buf[c * L + e];
  }
}

// one rlft3 unpacking step: h1 = c1 (a + b), h2 = c2 (a - b) with
// c1 = (0.5, 0), c2 = (0, c2i), b already conjugated
struct H12 { double h1r, h1i, h2r, h2i; };
__device__ __forceinline__ H12 rl_h(double ar, double ai, double br, double bi, double c2i) {
  const double sr = ar + br, si = ai + bi, dr = ar - br, di = ai - bi;
  return {0.5 * sr - 0.0 * si, 0.5 * si + 0.0 * sr, 0.0 * dr - c2i * di, 0.0 * di + c2i * dr};
}

// rlft3's unpacking loop (SG8:1484-1508) after the forward FFT (isign 1:
// S receives the Nyquist column) or before the inverse one (isign -1: S
// holds it).  One thread per independent piece: a pair (i2, mirror) of
// the first column, an (i1, i2) of columns 2..nn1/4 with its partner
// (nn1/2-i1+2, mirror), or a pair of the middle column, which the
// reference's in-place loop passes twice.
__global__ void k_pb_unpack(PbArgs P, double2 *Z, double2 *S, int nplanes, int isign) {
  const int H = P.nn1 / 2, nn2 = P.nn2, q4 = P.nn1 / 4;
  const int npair = nn2 / 2 + 1;
  const long per = 2L * npair + (long)(q4 - 1) * nn2, t = g_tid();
  if (t >= per * nplanes) return;
  const int pl = (int)(t / per);
  long m = t - pl * per;
  double2 *D = Z + (long)pl * nn2 * H;
  double2 *Sp = S + (long)pl * nn2;
  const double c2i = -0.5 * (double)isign;
  const double2 *wt = isign > 0 ? P.wf : P.wi;
  auto at = [&](int i1, int i2) -> double2 & { return D[(long)(i2 - 1) * H + (i1 - 1)]; };
  auto mir = [&](int i2) { return i2 == 1 ? 1 : nn2 - i2 + 2; };
  if (m < npair) {
    const int a = (int)m + 1, b = mir(a);
    if (isign > 0) {
      const double2 za = at(1, a), zb = at(1, b);
      H12 h = rl_h(za.x, za.y, zb.x, -zb.y, c2i);
      at(1, a) = make_double2(h.h1r + h.h2r, h.h1i + h.h2i);
      Sp[b - 1] = make_double2(h.h1r - h.h2r, -(h.h1i - h.h2i));
      if (b != a) {
        h = rl_h(zb.x, zb.y, za.x, -za.y, c2i);
        at(1, b) = make_double2(h.h1r + h.h2r, h.h1i + h.h2i);
        Sp[a - 1] = make_double2(h.h1r - h.h2r, -(h.h1i - h.h2i));
      }
    } else {
      const double2 da = at(1, a), db = at(1, b), sa = Sp[a - 1], sb = Sp[b - 1];
      H12 h = rl_h(da.x, da.y, sb.x, -sb.y, c2i);
      at(1, a) = make_double2(h.h1r + h.h2r, h.h1i + h.h2i);
      if (b != a) {
        h = rl_h(db.x, db.y, sa.x, -sa.y, c2i);
        at(1, b) = make_double2(h.h1r + h.h2r, h.h1i + h.h2i);
      }
    }
    return;
  }
  m -= npair;
  if (m < npair) {
    // middle column i1 = nn1/4 + 1: step a, then step b on the new values
    const int i1 = q4 + 1, a = (int)m + 1, b = mir(a);
    const double wr = wt[i1].x, wi = wt[i1].y;
    double2 da = at(i1, a), db = at(i1, b);
    H12 h = rl_h(da.x, da.y, db.x, -db.y, c2i);
    double whr = wr * h.h2r - wi * h.h2i, whi = wr * h.h2i + wi * h.h2r;
    const double2 na = make_double2(h.h1r + whr, h.h1i + whi);
    const double2 nb = make_double2(h.h1r - whr, -(h.h1i - whi));
    if (b == a) {
      at(i1, a) = nb;
      return;
    }
    h = rl_h(nb.x, nb.y, na.x, -na.y, c2i);
    whr = wr * h.h2r - wi * h.h2i;
    whi = wr * h.h2i + wi * h.h2r;
    at(i1, b) = make_double2(h.h1r + whr, h.h1i + whi);
    at(i1, a) = make_double2(h.h1r - whr, -(h.h1i - whi));
    return;
  }
  m -= npair;
  const int i1 = 2 + (int)(m / nn2), i2 = 1 + (int)(m % nn2);
  const int j1 = H - i1 + 2, j2 = mir(i2);
  const double wr = wt[i1].x, wi = wt[i1].y;
  const double2 da = at(i1, i2), db = at(j1, j2);
  const H12 h = rl_h(da.x, da.y, db.x, -db.y, c2i);
  const double whr = wr * h.h2r - wi * h.h2i, whi = wr * h.h2i + wi * h.h2r;
  at(i1, i2) = make_double2(h.h1r + whr, h.h1i + whi);
  at(j1, j2) = make_double2(h.h1r - whr, -(h.h1i - whi));
}

// PB's mode weights (SG8:1146-1229) summed over the planes: for mode
// (l, m) of the y (nn1) and z (nn2) transforms, 0.5 exp(-z kappa)/kappa
// facG with kappa^2 = 4 (sin^2(pi (l-1)/nn1)/dy^2 + sin^2(pi (m-1)/nn2)/dz^2)
// for the left (z = zp) and right (z = Lbox - zp) end; the mean mode takes
// -0.5 z facG; the Nyquist column of y (the speq values) takes
// kappa = sqrt(sqrt(kappa^2) + 1e-20) for m <= nn2/2+1, as the reference.
// One thread per mode (pp = 0..nn1/2, q = 0..nn2-1).
// One wave per mode: the lanes take the planes in turn and the four sums
// are reduced across the wave (the reference's sum runs over each rank's
// planes and then over the ranks in the MPI library's order, so no order
// is the reference's bit for bit).
__global__ void k_pb_weight(PbArgs P, const double2 *Z, const double2 *S, double2 *D, double2 *SD) {
  const int H = P.nn1 / 2, nn1 = P.nn1, nn2 = P.nn2;
  const long t = g_tid() >> 6;
  const int lane = threadIdx.x & 63;
  if (t >= (long)(H + 1) * nn2) return;
  const int pp = (int)(t % (H + 1)), q = (int)(t / (H + 1));
  const double pi = 3.14159265359;
  const int l = pp + 1, m2 = q + 1;
  auto k2 = [&](int li, int mi) {
    const double s1 = sin(pi * (double)(li - 1) / (double)nn1), s2 = sin(pi * (double)(mi - 1) / (double)nn2);
    return 4.0 * (s1 * s1 / (P.dxx * P.dxx) + s2 * s2 / (P.dyy * P.dyy));
  };
  const bool dc = pp == 0 && m2 == 1;
  double kap = 0.0;
  if (pp < H) {
    if (!dc) kap = sqrt(k2(l, m2 <= nn2 / 2 + 1 ? m2 : nn2 + 2 - m2));
  } else {
    kap = m2 <= nn2 / 2 + 1 ? sqrt(sqrt(k2(l, m2)) + 1.0e-20) : sqrt(k2(l, nn2 + 2 - m2));
  }
  double ar[2] = {0.0, 0.0}, ai[2] = {0.0, 0.0};
  for (int pl = lane; pl < P.np; pl += 64) {
    const double2 X = pp < H ? Z[((long)pl * nn2 + q) * H + pp] : S[(long)pl * nn2 + q];
    const double fg = P.facG[pl];
    const double z[2] = {P.zp[pl], P.zq[pl]};
    for (int s = 0; s < 2; ++s) {
      if (dc) {
        ar[s] = ar[s] + (0.0 - X.x * 0.5 * z[s] * fg);
        ai[s] = ai[s] + (0.0 - X.y * 0.5 * z[s] * fg);
      } else {
        const double e = exp(-z[s] * kap);
        ar[s] = ar[s] + X.x * 0.5 * e / kap * fg;
        ai[s] = ai[s] + X.y * 0.5 * e / kap * fg;
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      ar[s] += __shfl_xor(ar[s], o);
      ai[s] += __shfl_xor(ai[s], o);
    }
  if (lane != 0) return;
  for (int s = 0; s < 2; ++s) {
    if (pp < H) D[((long)s * nn2 + q) * H + pp] = make_double2(ar[s], ai[s]);
    else SD[(long)s * nn2 + q] = make_double2(ar[s], ai[s]);
  }
}

// the end-plane potential after the inverse transform, divided by
// 2 (nn1/2)^2, onto the finest split level of every rank (SG8:1259-1272):
// node (j, k) of the rank's plane, 0..C+1, the outer nodes wrapped
__global__ void k_pb_bphi(GravGeo G, PbArgs P, const double2 *D, double *bphi2) {
  const int L = G.NGL, ny = G.C[1] + 1, nz = G.C[2] + 1;
  const long per = (long)(ny + 1) * (nz + 1), t = g_tid();
  if (t >= per * 2 * G.npe) return;
  const int rs = (int)(t / per), r = rs / 2, s = rs % 2;
  const long nidx = t - rs * per;
  const int j = (int)(nidx % (ny + 1)), k = (int)(nidx / (ny + 1));
  int c[3];
  g_rank(G, r, c);
  int jb = c[1] * G.C[1] + j, kb = c[2] * G.C[2] + k;
  if (j == ny && c[1] == G.S[1] - 1) jb = 1;
  if (k == nz && c[2] == G.S[2] - 1) kb = 1;
  if (j == 0 && c[1] == 0) jb = P.nn1;
  if (k == 0 && c[2] == 0) kb = P.nn2;
  const int H = P.nn1 / 2;
  const double2 z = D[((long)s * P.nn2 + (kb - 1)) * H + (jb - 1) / 2];
  const double fac = (double)(2 * (P.nn1 / 2) * (P.nn1 / 2));
  bphi2[G.boff[L] + t] = ((jb & 1) ? z.x : z.y) / fac;
}

// the end planes restricted to split level lc (SG8:1288-1363)
__global__ void k_pb_restrict(GravGeo G, double *bphi2, int lc) {
  const int mcx = G.n[lc][1], mcy = G.n[lc][2];
  const long cper = (long)(mcx + 1) * (mcy + 1), fper = (long)(G.n[lc + 1][1] + 1) * (G.n[lc + 1][2] + 1);
  const long cnt = (long)mcx * mcy, t = g_tid();
  if (t >= cnt * 2 * G.npe) return;
  const int rs = (int)(t / cnt);
  const long m = t - rs * cnt;
  const int ic = 1 + (int)(m % mcx), jc = 1 + (int)(m / mcx);
  const double *bf = bphi2 + G.boff[lc + 1] + rs * fper;
  const long kf = (2 * ic - 1) + (long)(mcx * 2) * (2 * jc - 1), kc = ic + (long)(mcx + 1) * jc;
  bphi2[G.boff[lc] + rs * cper + kc] = plane_restrict(bf, kf, 2L * mcx, ic, jc, mcx, mcy);
}

// the level-NGcr end planes of the first / last x ranks gathered into the
// whole-domain level NGcr (SG8:1366-1390)
__global__ void k_pb_collect(GravGeo G, double *bphi1, const double *bphi2) {
  const int L = G.NGcr, N1 = (1 << L) + 1;
  const int *m = G.n[L];
  const long t = g_tid();
  if (t >= 2L * N1 * N1) return;
  const int s = (int)(t / ((long)N1 * N1));
  const long q = t - s * (long)N1 * N1;
  const int j = 1 + (int)(q % N1), k = 1 + (int)(q / N1);
  const int cy = min((j - 1) / (m[1] - 1), G.S[1] - 1), cz = min((k - 1) / (m[2] - 1), G.S[2] - 1);
  const int jj = j - cy * (m[1] - 1), kk = k - cz * (m[2] - 1);
  const int r = g_rid(G, s == 0 ? 0 : G.S[0] - 1, cy, cz);
  bphi1[s * (G.needb1 + 1) + G.pb1[L] + (j - 1) + (long)N1 * (k - 1)] =
      bphi2[G.boff[L] + (2L * r + s) * (long)(m[1] + 1) * (m[2] + 1) + jj + (long)(m[1] + 1) * kk];
}

// Phi at the cells: each rank's nodes 1..C of the finest level (SG8:282-291)
// (a rank block's cells in g_node's order: the reads of the skewed layout
// contiguous, the writes scattered)
__global__ void k_grav_phi(GravGeo G, const double *phi2, double *Phi, Geo g) {
  const long cnt = (long)G.C[0] * G.C[1] * G.C[2], t = g_tid();
  if (t >= cnt * G.npe) return;
  const int r = (int)(t / cnt), L = G.NGL;
  int i, j, k, c[3];
  g_node(G.C, t - r * cnt, i, j, k);
  g_rank(G, r, c);
  Phi[ix3(g, c[0] * G.C[0] + i, c[1] * G.C[1] + j, c[2] * G.C[2] + k)] = phi2[G.off[L] + r * G.bsz[L] + g_idx(G.n[L], i, j, k)];
}

// Phi's face ghosts (SG8:22-67): y and z periodic, the x ends copied from
// the first / last cell (zero gradient)
__global__ void k_grav_phi_ghosts(double *Phi, Geo g) {
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2];
  const long a0 = (long)N1 * N2, a1 = (long)N0 * N2, a2 = (long)N0 * N1, t = g_tid();
  if (t < a0) {
    const int j = 1 + (int)(t % N1), k = 1 + (int)(t / N1);
    const double l = Phi[ix3(g, 1, j, k)], r = Phi[ix3(g, N0, j, k)];
    Phi[ix3(g, 0, j, k)] = l;
    Phi[ix3(g, -1, j, k)] = l;
    Phi[ix3(g, N0 + 1, j, k)] = r;
    Phi[ix3(g, N0 + 2, j, k)] = r;
  } else if (t < a0 + a1) {
    const long u = t - a0;
    const int i = 1 + (int)(u % N0), k = 1 + (int)(u / N0);
    Phi[ix3(g, i, -1, k)] = Phi[ix3(g, i, N1 - 1, k)];
    Phi[ix3(g, i, 0, k)] = Phi[ix3(g, i, N1, k)];
    Phi[ix3(g, i, N1 + 1, k)] = Phi[ix3(g, i, 1, k)];
    Phi[ix3(g, i, N1 + 2, k)] = Phi[ix3(g, i, 2, k)];
  } else if (t < a0 + a1 + a2) {
    const long u = t - a0 - a1;
    const int i = 1 + (int)(u % N0), j = 1 + (int)(u / N0);
    Phi[ix3(g, i, j, -1)] = Phi[ix3(g, i, j, N2 - 1)];
    Phi[ix3(g, i, j, 0)] = Phi[ix3(g, i, j, N2)];
    Phi[ix3(g, i, j, N2 + 1)] = Phi[ix3(g, i, j, 1)];
    Phi[ix3(g, i, j, N2 + 2)] = Phi[ix3(g, i, j, 2)];
  }
}

// GRAVTY(dt,3) (SG8:82-89): v -= dt/2 * the fourth-order gradient of Phi,
// 1/(12 dx(0)) with dx(0) of the cell's x rank
__global__ void k_grav_kick(double *U2, double *U3, double *U4, const double *Phi, Geo g, const double *gdx0, int cx,
                            double dt) {
  const long N0 = g.n[0], N1 = g.n[1];
  const long t = g_tid();
  if (t >= N0 * N1 * g.n[2]) return;
  const int i = 1 + (int)(t % N0), j = 1 + (int)((t / N0) % N1), k = 1 + (int)(t / (N0 * N1));
  const int c = ix3(g, i, j, k), sx = 1, sy = g.st[1], sz = g.st[2];
  const double dxi = 1.0 / (12.0 * gdx0[((i - 1) / cx) * cx + 1]);
  const double *P = Phi;
  U2[c] = U2[c] - dt * (-P[c + 2 * sx] + 8.0 * P[c + sx] - 8.0 * P[c - sx] + P[c - 2 * sx]) * dxi * 0.5;
  U3[c] = U3[c] - dt * (-P[c + 2 * sy] + 8.0 * P[c + sy] - 8.0 * P[c - sy] + P[c - 2 * sy]) * dxi * 0.5;
  U4[c] = U4[c] - dt * (-P[c + 2 * sz] + 8.0 * P[c + sz] - 8.0 * P[c - sz] + P[c - 2 * sz]) * dxi * 0.5;
}

} // namespace mcf
