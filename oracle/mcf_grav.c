/*
 * mcf_grav.c -- CPU restatement (plain C, FP64) of the self-gravity solver
 * GRAVTY of 3DMHD-CfCA/selfG_MPIfull8.f90 ("SG8"), called by the MHD
 * restatement (mcf_oracle.c) when ifgrv = 2.  TEST INFRASTRUCTURE ONLY: the
 * checker of the HIP path; the product never links it.
 *
 * Poisson's equation lap(Phi) = 4 pi G rho is solved on nodes that coincide
 * with the cell centres: periodic in y and z, with the potential on the two
 * x ends taken from the free-space Green's function of the y-z Fourier
 * modes of every x plane (PB, SG8:1026-1463).  The solver is full multigrid
 * (mglin, SG8:151-295): two V-cycles per level with 5 + 5 Gauss-Seidel
 * sweeps (2 + 2 on the split levels once the finest level is reached).
 * Levels 1..NGcr are whole-domain grids of 2^l+1 nodes per axis; levels
 * NGcr..NGL are split over the NSPLTx*NSPLTy*NSPLTz ranks, each rank holding
 * nodes 0..n of its block, node 0 and node n being copies of the
 * neighbours' nodes n-1 and 1.  Each rank's Gauss-Seidel sweep runs over its
 * own block in lexicographic order with its ghost nodes fixed, so the
 * result depends on the block split; this restatement keeps one set of
 * arrays per rank and performs the exchanges (BCsgr_MPI), gathers (collect)
 * and scatters (divide) between them in the reference's order.
 *
 * Reproduced properties of the reference (each cited where it is used):
 *  - the whole-domain levels wrap y and z with period n, not n-1 (relax);
 *  - the whole-domain residual reads past the ends of its level's block in
 *    the level-stacked storage (resid), so that storage is kept as one array
 *    with the reference's offsets;
 *  - the real-FFT unpacking twiddles are rounded to single precision (CMPLX
 *    without a kind, rlft3);
 *  - the Nyquist column of the y transform gets kappa = sqrt(sqrt(k2)+1e-20)
 *    for z frequencies 0..nn2/2 (PB);
 *  - only x planes Nir+1+P*m (m = 0..loopbc = 2, P = NSPLTy*NSPLTz, Nir the
 *    rank's y-z index) of each x block enter the boundary potential.
 * Every array starts each solve at zero (mcf_grav.h).  The sum over ranks of
 * the boundary spectra (MPI_ALLREDUCE in the reference) is taken in rank
 * order; the reference's order belongs to the MPI library, so results agree
 * to rounding, not bit for bit.
 *
 * Built with -ffp-contract=off and the reference's evaluation order.
 */
#include "mcf_grav.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* M3:35: G = 1.11142d-4, G4pi = 12.56637d0*G */
#define G4PI (12.56637 * 1.11142e-4)
#define LOOPBC 2  /* M3:16 */
#define NCYCLE 2  /* mglin(Nmem1,Nmem2,2,5,5), SG8:18 */
#define NPRE 5
#define NPOST 5

struct mcfg {
  int N[3], S[3], C[3], npe, NGL, NGcr;
  long point1[18], point2[18], pointb1[18], pointb2[18];
  long Need1, Need2, Needb1, Needb2;
  int ncr[3];              /* per-rank nodes at level NGcr (ncxcr) */
  double Lbox;
  double *xi0, *dxi0, dy1, dz1;  /* x_i and dx_i of x (index -1..N+2), dy(1), dz(1) */
  /* whole-domain arrays, 1-based: a[p], p = 1..Need1 */
  double *cphi1, *crho1, *cres1, *crhs1;
  /* per-rank arrays, 1-based: a[r*(Need2+1) + p] */
  double *cphi2, *crho2, *cres2, *crhs2;
  double *bphi1;  /* [2][Needb1+1] */
  double *bphi2;  /* [npe][2][Needb2+1] */
};

static void die(const char *m) {
  fprintf(stderr, "mcf_grav: %s\n", m);
  abort();
}

static int ilog2(int v) { /* int(dlog(dble(v))/dlog(2.d0)+1.d-3) for powers of two */
  int l = 0;
  while ((1 << (l + 1)) <= v) l++;
  return l;
}

static inline void rank_coords(const mcfg *g, int r, int *ist, int *jst, int *kst) {
  *ist = r % g->S[0];
  *kst = r / (g->S[0] * g->S[1]);
  *jst = r / g->S[0] - g->S[1] * *kst;
}
static inline int rank_of(const mcfg *g, int ist, int jst, int kst) {
  return ist + g->S[0] * (jst + g->S[1] * kst);
}

/* per-rank level arrays: u(0:n0,0:n1,0:n2) starting at 1-based offset off */
static inline double *R(const mcfg *g, double *base, int r) { return base + (size_t)r * (size_t)(g->Need2 + 1); }
#define U2(a, off, n, i, j, k) (a)[(off) + (i) + (long)((n)[0] + 1) * ((j) + (long)((n)[1] + 1) * (k))]
/* whole-domain level arrays: u(1:n,1:n,1:n) at 1-based offset off; the flat
 * index may leave the level (resid), it must stay inside 1..Need1 */
static inline long F1(const mcfg *g, long off, int n, int i, int j, int k) {
  long p = off + (i - 1) + (long)n * ((j - 1) + (long)n * (k - 1));
  if (p < 1 || p > g->Need1) die("whole-domain index outside the level storage");
  return p;
}

/* ------------------------------------------------------------------ */
/* BCsgr_MPI (SG8:298-336): x ends are not exchanged (PROC_NULL at IST = 0
 * and NSPLTx-1); y and z are periodic.  Phases in the reference's order,
 * each a full plane (ghost rows of the other axes included). */
static void bcsgr(const mcfg *g, double *base, long off, const int n[3], int lx1, int lx2, int ly1, int ly2,
                  int lz1, int lz2) {
  const int flags[6] = {lx1, lx2, ly1, ly2, lz1, lz2};
  for (int ph = 0; ph < 6; ph++) {
    if (!flags[ph]) continue;
    const int ax = ph / 2, hi = ph % 2;
    for (int r = 0; r < g->npe; r++) {
      int c[3];
      rank_coords(g, r, &c[0], &c[1], &c[2]);
      int s[3] = {c[0], c[1], c[2]};
      if (ax == 0) {
        if (!hi && c[0] == 0) continue;
        if (hi && c[0] == g->S[0] - 1) continue;
        s[0] += hi ? 1 : -1;
      } else {
        s[ax] = (c[ax] + (hi ? 1 : g->S[ax] - 1)) % g->S[ax];
      }
      double *dst = R(g, base, r), *src = R(g, base, rank_of(g, s[0], s[1], s[2]));
      const int dplane = hi ? n[ax] : 0, splane = hi ? 1 : n[ax] - 1;
      const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
      for (int v = 0; v <= n[a2]; v++)
        for (int u = 0; u <= n[a1]; u++) {
          int id[3], is[3];
          id[ax] = dplane; is[ax] = splane;
          id[a1] = is[a1] = u;
          id[a2] = is[a2] = v;
          U2(dst, off, n, id[0], id[1], id[2]) = U2(src, off, n, is[0], is[1], is[2]);
        }
    }
  }
}

/* collect (SG8:339-366): rank blocks (0:n2) -> whole domain (1:n1) */
static void collect(const mcfg *g, double *u1, long off1, int n1, double *base2, long off2, const int n2[3]) {
  for (int r = 0; r < g->npe; r++) {
    int c[3];
    rank_coords(g, r, &c[0], &c[1], &c[2]);
    int ed[3];
    for (int d = 0; d < 3; d++) ed[d] = (c[d] == g->S[d] - 1) ? n2[d] : n2[d] - 1;
    const double *u2 = R(g, base2, r);
    for (int kk = 1; kk <= ed[2]; kk++)
      for (int jj = 1; jj <= ed[1]; jj++)
        for (int ii = 1; ii <= ed[0]; ii++)
          u1[F1(g, off1, n1, c[0] * (n2[0] - 1) + ii, c[1] * (n2[1] - 1) + jj, c[2] * (n2[2] - 1) + kk)] =
              U2(u2, off2, n2, ii, jj, kk);
  }
}

/* divide (SG8:369-382): whole domain -> rank blocks; the node-0 planes of
 * the first rank along each axis are not written */
static void divide(const mcfg *g, double *base2, long off2, const int n2[3], const double *u1, long off1, int n1) {
  for (int r = 0; r < g->npe; r++) {
    int c[3];
    rank_coords(g, r, &c[0], &c[1], &c[2]);
    int st[3];
    for (int d = 0; d < 3; d++) st[d] = c[d] == 0 ? 1 : 0;
    double *u2 = R(g, base2, r);
    for (int kk = st[2]; kk <= n2[2]; kk++)
      for (int jj = st[1]; jj <= n2[1]; jj++)
        for (int ii = st[0]; ii <= n2[0]; ii++)
          U2(u2, off2, n2, ii, jj, kk) =
              u1[F1(g, off1, n1, c[0] * (n2[0] - 1) + ii, c[1] * (n2[1] - 1) + jj, c[2] * (n2[2] - 1) + kk)];
  }
}

/* rstrctMPI (SG8:385-456): full weighting 1/2 + 6 x 1/12, the block's first
 * node along an axis copied (not weighted) on the first rank and its last
 * node on the last rank; mode 1 zeroes the x-end nodes. */
static void rstrct_mpi(const mcfg *g, double *bc, long offc, double *bf, long offf, const int nc[3], int mode) {
  const double w = 1.0 / 12.0;
  const int nf[3] = {2 * nc[0] - 1, 2 * nc[1] - 1, 2 * nc[2] - 1};
  for (int r = 0; r < g->npe; r++) {
    int c[3];
    rank_coords(g, r, &c[0], &c[1], &c[2]);
    double *uc = R(g, bc, r), *uf = R(g, bf, r);
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) { lo[d] = c[d] == 0 ? 2 : 1; hi[d] = nc[d] - 1; }
    for (int kc = lo[2]; kc <= hi[2]; kc++)
      for (int jc = lo[1]; jc <= hi[1]; jc++)
        for (int ic = lo[0]; ic <= hi[0]; ic++) {
          const int fi = 2 * ic - 1, fj = 2 * jc - 1, fk = 2 * kc - 1;
          U2(uc, offc, nc, ic, jc, kc) =
              0.5 * U2(uf, offf, nf, fi, fj, fk) +
              w * (U2(uf, offf, nf, fi + 1, fj, fk) + U2(uf, offf, nf, fi - 1, fj, fk) + U2(uf, offf, nf, fi, fj + 1, fk) +
                   U2(uf, offf, nf, fi, fj - 1, fk) + U2(uf, offf, nf, fi, fj, fk + 1) + U2(uf, offf, nf, fi, fj, fk - 1));
        }
    /* copies at the first / last node of the first / last rank (every copy
     * of a node takes the same fine value, so their order is immaterial) */
    for (int d = 0; d < 3; d++)
      for (int side = 0; side < 2; side++) {
        if (side == 0 && c[d] != 0) continue;
        if (side == 1 && c[d] != g->S[d] - 1) continue;
        const int a1 = (d + 1) % 3, a2 = (d + 2) % 3;
        for (int v = 1; v <= nc[a2]; v++)
          for (int u = 1; u <= nc[a1]; u++) {
            int ic[3], fi[3];
            ic[d] = side ? nc[d] : 1;
            fi[d] = side ? nf[d] : 1;
            ic[a1] = u; fi[a1] = 2 * u - 1;
            ic[a2] = v; fi[a2] = 2 * v - 1;
            U2(uc, offc, nc, ic[0], ic[1], ic[2]) = U2(uf, offf, nf, fi[0], fi[1], fi[2]);
          }
      }
    if (mode == 1) {
      for (int kc = 1; kc <= nc[2]; kc++)
        for (int jc = 1; jc <= nc[1]; jc++) {
          if (c[0] == 0) U2(uc, offc, nc, 1, jc, kc) = 0.0;
          if (c[0] == g->S[0] - 1) U2(uc, offc, nc, nc[0], jc, kc) = 0.0;
        }
    }
  }
  bcsgr(g, bc, offc, nc, 1, 1, 1, 1, 1, 1);
}

/* rstrct (SG8:459-497): whole-domain restriction, all six faces copied */
static void rstrct(const mcfg *g, double *uc, long offc, const double *uf, long offf, int nc, int mode) {
  const double w = 1.0 / 12.0;
  const int nf = 2 * nc - 1;
#define UF(i, j, k) uf[F1(g, offf, nf, i, j, k)]
#define UC(i, j, k) uc[F1(g, offc, nc, i, j, k)]
  for (int kc = 2; kc <= nc - 1; kc++)
    for (int jc = 2; jc <= nc - 1; jc++)
      for (int ic = 2; ic <= nc - 1; ic++) {
        const int fi = 2 * ic - 1, fj = 2 * jc - 1, fk = 2 * kc - 1;
        UC(ic, jc, kc) = 0.5 * UF(fi, fj, fk) + w * (UF(fi + 1, fj, fk) + UF(fi - 1, fj, fk) + UF(fi, fj + 1, fk) +
                                                     UF(fi, fj - 1, fk) + UF(fi, fj, fk + 1) + UF(fi, fj, fk - 1));
      }
  for (int kc = 1; kc <= nc; kc++)
    for (int jc = 1; jc <= nc; jc++) {
      UC(1, jc, kc) = UF(1, 2 * jc - 1, 2 * kc - 1);
      UC(nc, jc, kc) = UF(nf, 2 * jc - 1, 2 * kc - 1);
    }
  for (int jc = 1; jc <= nc; jc++)
    for (int ic = 1; ic <= nc; ic++) {
      UC(ic, jc, 1) = UF(2 * ic - 1, 2 * jc - 1, 1);
      UC(ic, jc, nc) = UF(2 * ic - 1, 2 * jc - 1, nf);
    }
  for (int kc = 1; kc <= nc; kc++)
    for (int ic = 1; ic <= nc; ic++) {
      UC(ic, 1, kc) = UF(2 * ic - 1, 1, 2 * kc - 1);
      UC(ic, nc, kc) = UF(2 * ic - 1, nf, 2 * kc - 1);
    }
  if (mode == 1)
    for (int kc = 1; kc <= nc; kc++)
      for (int jc = 1; jc <= nc; jc++) {
        UC(1, jc, kc) = 0.0;
        UC(nc, jc, kc) = 0.0;
      }
#undef UF
#undef UC
}

/* Prolongation of one block (0:n) from (0:n/2+1): injection, then the
 * midpoints along x (odd y, z), y (odd z) and z, each the mean of its two
 * neighbours (interpMPI / interp, SG8:500-599) */
static void prolong_block(double *uf, long offf, const int nf[3], const double *uc, long offc, const int nc[3]) {
  for (int kc = 1; kc <= nf[2] / 2 + 1; kc++)
    for (int jc = 1; jc <= nf[1] / 2 + 1; jc++)
      for (int ic = 1; ic <= nf[0] / 2 + 1; ic++)
        U2(uf, offf, nf, 2 * ic - 1, 2 * jc - 1, 2 * kc - 1) = U2(uc, offc, nc, ic, jc, kc);
  for (int kf = 1; kf <= nf[2]; kf += 2)
    for (int jf = 1; jf <= nf[1]; jf += 2)
      for (int fi = 2; fi <= nf[0]; fi += 2)
        U2(uf, offf, nf, fi, jf, kf) = .5 * (U2(uf, offf, nf, fi + 1, jf, kf) + U2(uf, offf, nf, fi - 1, jf, kf));
  for (int kf = 1; kf <= nf[2]; kf += 2)
    for (int jf = 2; jf <= nf[1]; jf += 2)
      for (int fi = 1; fi <= nf[0]; fi++)
        U2(uf, offf, nf, fi, jf, kf) = .5 * (U2(uf, offf, nf, fi, jf + 1, kf) + U2(uf, offf, nf, fi, jf - 1, kf));
  for (int kf = 2; kf <= nf[2]; kf += 2)
    for (int jf = 1; jf <= nf[1]; jf++)
      for (int fi = 1; fi <= nf[0]; fi++)
        U2(uf, offf, nf, fi, jf, kf) = .5 * (U2(uf, offf, nf, fi, jf, kf + 1) + U2(uf, offf, nf, fi, jf, kf - 1));
}

/* interpMPI (SG8:500-554); mode 1 sets the x-end nodes from the boundary
 * potential bphi2 at offset np (node (j,k) at np + j + (n1+1)*k) */
static void interp_mpi(const mcfg *g, double *bf, long offf, double *bc, long offc, const int nf[3], long np, int mode) {
  const int nc[3] = {nf[0] / 2 + 1, nf[1] / 2 + 1, nf[2] / 2 + 1};
  for (int r = 0; r < g->npe; r++) {
    int c[3];
    rank_coords(g, r, &c[0], &c[1], &c[2]);
    double *uf = R(g, bf, r);
    prolong_block(uf, offf, nf, R(g, bc, r), offc, nc);
    if (mode == 1) {
      const double *b = g->bphi2 + (size_t)r * 2 * (size_t)(g->Needb2 + 1);
      for (int k = 0; k <= nf[2]; k++)
        for (int j = 0; j <= nf[1]; j++) {
          const long nn = j + (long)(nf[1] + 1) * k;
          if (c[0] == 0) U2(uf, offf, nf, 1, j, k) = b[np + nn];
          if (c[0] == g->S[0] - 1) U2(uf, offf, nf, nf[0], j, k) = b[(g->Needb2 + 1) + np + nn];
        }
    }
  }
  bcsgr(g, bf, offf, nf, 1, 0, 1, 0, 1, 0);
}

/* interp (SG8:557-599), whole domain; mode 1 sets the x-end nodes from
 * bphi1 (node (j,k) at np + (j-1) + n*(k-1)) */
static void interp(const mcfg *g, double *uf, long offf, const double *uc, long offc, int nf, long np, int mode) {
  const int nc = nf / 2 + 1;
#define UF(i, j, k) uf[F1(g, offf, nf, i, j, k)]
#define UC(i, j, k) uc[F1(g, offc, nc, i, j, k)]
  for (int kc = 1; kc <= nc; kc++)
    for (int jc = 1; jc <= nc; jc++)
      for (int ic = 1; ic <= nc; ic++) UF(2 * ic - 1, 2 * jc - 1, 2 * kc - 1) = UC(ic, jc, kc);
  for (int kf = 1; kf <= nf; kf += 2)
    for (int jf = 1; jf <= nf; jf += 2)
      for (int fi = 2; fi <= nf; fi += 2) UF(fi, jf, kf) = .5 * (UF(fi + 1, jf, kf) + UF(fi - 1, jf, kf));
  for (int kf = 1; kf <= nf; kf += 2)
    for (int jf = 2; jf <= nf; jf += 2)
      for (int fi = 1; fi <= nf; fi++) UF(fi, jf, kf) = .5 * (UF(fi, jf + 1, kf) + UF(fi, jf - 1, kf));
  for (int kf = 2; kf <= nf; kf += 2)
    for (int jf = 1; jf <= nf; jf++)
      for (int fi = 1; fi <= nf; fi++) UF(fi, jf, kf) = .5 * (UF(fi, jf, kf + 1) + UF(fi, jf, kf - 1));
  if (mode == 1)
    for (int kf = 1; kf <= nf; kf++)
      for (int jf = 1; jf <= nf; jf++) {
        const long nn = (jf - 1) + (long)nf * (kf - 1);
        UF(1, jf, kf) = g->bphi1[np + nn];
        UF(nf, jf, kf) = g->bphi1[(g->Needb1 + 1) + np + nn];
      }
#undef UF
#undef UC
}

/* addintMPI (SG8:602-624): res = prolongation of uc (exchanged), uf += res
 * over every node 0..n */
static void addint_mpi(const mcfg *g, double *bf, long offf, double *bc, long offc, double *bres, long offr,
                       const int nf[3]) {
  interp_mpi(g, bres, offr, bc, offc, nf, 0, 0);
  for (int r = 0; r < g->npe; r++) {
    double *uf = R(g, bf, r), *res = R(g, bres, r);
    for (int k = 0; k <= nf[2]; k++)
      for (int j = 0; j <= nf[1]; j++)
        for (int i = 0; i <= nf[0]; i++) U2(uf, offf, nf, i, j, k) = U2(uf, offf, nf, i, j, k) + U2(res, offr, nf, i, j, k);
  }
}

/* addint (SG8:627-648) */
static void addint(const mcfg *g, double *uf, long offf, const double *uc, long offc, double *res, long offr, int nf) {
  interp(g, res, offr, uc, offc, nf, 0, 0);
  for (int k = 1; k <= nf; k++)
    for (int j = 1; j <= nf; j++)
      for (int i = 1; i <= nf; i++) uf[F1(g, offf, nf, i, j, k)] = uf[F1(g, offf, nf, i, j, k)] + res[F1(g, offr, nf, i, j, k)];
}

/* Exact solves on the 3^3 grid (slvsml SG8:651-679 with zero x ends,
 * slvsmlb SG8:682-737 with the boundary potential on them): u(2,j,k) from
 * the weights 8 (same j and k), 2 (one of them the same), 1 (neither), the
 * terms summed in the order (j',k') = (1,1),(1,2),(1,3),(2,1),...,(3,3). */
static inline double w3(int j, int k, int jp, int kp) { return (j == jp && k == kp) ? 8.0 : ((j == jp || k == kp) ? 2.0 : 1.0); }

static void slvsml(const mcfg *g, double *u, long off, const double *rhs, long offr) {
  const double h = 0.5 * g->Lbox;
  for (int k = 1; k <= 3; k++)
    for (int j = 1; j <= 3; j++) {
      u[F1(g, off, 3, 1, j, k)] = 0.0;
      u[F1(g, off, 3, 3, j, k)] = 0.0;
    }
  for (int k = 1; k <= 3; k++)
    for (int j = 1; j <= 3; j++) {
      double s = 0.0;
      int first = 1;
      for (int jp = 1; jp <= 3; jp++)
        for (int kp = 1; kp <= 3; kp++) {
          const double t = w3(j, k, jp, kp) * rhs[F1(g, offr, 3, 2, jp, kp)];
          s = first ? -t : s - t;
          first = 0;
        }
      u[F1(g, off, 3, 2, j, k)] = 0.025 * h * h * s;
    }
}

static void slvsmlb(const mcfg *g, double *u, long off, const double *rhs, long offr) {
  const double h = 0.5 * g->Lbox;
  for (int k = 1; k <= 3; k++)
    for (int j = 1; j <= 3; j++) {
      u[F1(g, off, 3, 1, j, k)] = g->bphi1[(j - 1) + 3 * (k - 1) + 1];
      u[F1(g, off, 3, 3, j, k)] = g->bphi1[(g->Needb1 + 1) + (j - 1) + 3 * (k - 1) + 1];
    }
  for (int k = 1; k <= 3; k++)
    for (int j = 1; j <= 3; j++) {
      double s = 0.0;
      int first = 1;
      for (int side = 1; side <= 3; side += 2)
        for (int jp = 1; jp <= 3; jp++)
          for (int kp = 1; kp <= 3; kp++) {
            const double t = w3(j, k, jp, kp) * u[F1(g, off, 3, side, jp, kp)];
            s = first ? t : s + t;
            first = 0;
          }
      for (int jp = 1; jp <= 3; jp++)
        for (int kp = 1; kp <= 3; kp++) s = s - w3(j, k, jp, kp) * h * h * rhs[F1(g, offr, 3, 2, jp, kp)];
      u[F1(g, off, 3, 2, j, k)] = 0.025 * s;
    }
}

/* relaxMPI (SG8:740-870): one lexicographic Gauss-Seidel sweep over each
 * rank's nodes 1..n-1 (node 1 of the first x rank is the boundary and is
 * kept), then the exchange */
static void relax_mpi(const mcfg *g, double *bu, long off, double *brhs, long offr, const int n[3]) {
  const double w = 1.0 / 6.0;
  const double h = g->Lbox / (double)((n[0] - 1) * g->S[0]);
  const double h2 = h * h;
  for (int r = 0; r < g->npe; r++) {
    int c[3];
    rank_coords(g, r, &c[0], &c[1], &c[2]);
    const int li = c[0] == 0 ? 1 : 0;
    double *u = R(g, bu, r);
    const double *rhs = R(g, brhs, r);
    for (int k = 1; k <= n[2] - 1; k++)
      for (int j = 1; j <= n[1] - 1; j++)
        for (int i = 1; i <= n[0] - 1; i++) {
          const int ifl = li * (1 / i);
          const double s = U2(u, off, n, i + 1, j, k) + U2(u, off, n, i - 1, j, k) + U2(u, off, n, i, j + 1, k) +
                           U2(u, off, n, i, j - 1, k) + U2(u, off, n, i, j, k + 1) + U2(u, off, n, i, j, k - 1) -
                           h2 * U2(rhs, offr, n, i, j, k);
          U2(u, off, n, i, j, k) =
              w * s * (double)(1 - ifl) + (0.5 + copysign(0.5, (double)ifl - 0.5)) * U2(u, off, n, i, j, k);
        }
  }
  bcsgr(g, bu, off, n, 1, 1, 1, 1, 1, 1);
}

/* relax (SG8:873-915): whole domain, x nodes 2..n-1, y and z wrapped with
 * period n (node 1's lower neighbour is node n) */
static void relax(const mcfg *g, double *u, long off, const double *rhs, long offr, int n) {
  const double w = 1.0 / 6.0;
  const double h = g->Lbox / (double)(n - 1);
  const double h2 = h * h;
  for (int k = 1; k <= n; k++) {
    const int km = k == 1 ? n : k - 1, kp = k == n ? 1 : k + 1;
    for (int j = 1; j <= n; j++) {
      const int jm = j == 1 ? n : j - 1, jp = j == n ? 1 : j + 1;
      for (int i = 2; i <= n - 1; i++)
        u[F1(g, off, n, i, j, k)] =
            w * (u[F1(g, off, n, i + 1, j, k)] + u[F1(g, off, n, i - 1, j, k)] + u[F1(g, off, n, i, jp, k)] +
                 u[F1(g, off, n, i, jm, k)] + u[F1(g, off, n, i, j, kp)] + u[F1(g, off, n, i, j, km)] -
                 h2 * rhs[F1(g, offr, n, i, j, k)]);
    }
  }
}

/* residMPI (SG8:918-952) */
static void resid_mpi(const mcfg *g, double *bres, long offr, double *bu, long off, double *brhs, long offh,
                      const int n[3]) {
  const double h = g->Lbox / (double)((n[0] - 1) * g->S[0]);
  const double h2i = 1.0 / (h * h);
  for (int r = 0; r < g->npe; r++) {
    int c[3];
    rank_coords(g, r, &c[0], &c[1], &c[2]);
    double *res = R(g, bres, r);
    const double *u = R(g, bu, r), *rhs = R(g, brhs, r);
    for (int k = 1; k <= n[2] - 1; k++)
      for (int j = 1; j <= n[1] - 1; j++)
        for (int i = 1; i <= n[0] - 1; i++)
          U2(res, offr, n, i, j, k) =
              -h2i * (U2(u, off, n, i + 1, j, k) + U2(u, off, n, i - 1, j, k) + U2(u, off, n, i, j + 1, k) +
                      U2(u, off, n, i, j - 1, k) + U2(u, off, n, i, j, k + 1) + U2(u, off, n, i, j, k - 1) -
                      6.0 * U2(u, off, n, i, j, k)) +
              U2(rhs, offh, n, i, j, k);
    for (int k = 1; k <= n[2]; k++)
      for (int j = 1; j <= n[1]; j++) {
        if (c[0] == 0) U2(res, offr, n, 1, j, k) = 0.0;
        if (c[0] == g->S[0] - 1) U2(res, offr, n, n[0], j, k) = 0.0;
      }
  }
  bcsgr(g, bres, offr, n, 1, 1, 1, 1, 1, 1);
}

/* resid (SG8:955-971): whole domain; y and z neighbours of the first and
 * last nodes are read at the flat offsets +-n, +-n^2, i.e. in the
 * neighbouring rows, planes or levels of the level-stacked storage */
static void resid(const mcfg *g, double *res, long offr, const double *u, long off, const double *rhs, long offh, int n) {
  const double h = g->Lbox / (double)(n - 1);
  const double h2i = 1.0 / (h * h);
  for (int k = 1; k <= n; k++)
    for (int j = 1; j <= n; j++)
      for (int i = 2; i <= n - 1; i++)
        res[F1(g, offr, n, i, j, k)] =
            -h2i * (u[F1(g, off, n, i + 1, j, k)] + u[F1(g, off, n, i - 1, j, k)] + u[F1(g, off, n, i, j + 1, k)] +
                    u[F1(g, off, n, i, j - 1, k)] + u[F1(g, off, n, i, j, k + 1)] + u[F1(g, off, n, i, j, k - 1)] -
                    6.0 * u[F1(g, off, n, i, j, k)]) +
            rhs[F1(g, offh, n, i, j, k)];
  for (int k = 1; k <= n; k++)
    for (int j = 1; j <= n; j++) {
      res[F1(g, offr, n, 1, j, k)] = 0.0;
      res[F1(g, offr, n, n, j, k)] = 0.0;
    }
}

static void copy1(const mcfg *g, double *a, long offa, const double *b, long offb, int n) {
  for (long t = 0; t < (long)n * n * n; t++) a[offa + t] = b[offb + t];
  (void)g;
}
static void copy_mpi(const mcfg *g, double *ba, long offa, double *bb, long offb, const int n[3]) {
  const long m = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
  for (int r = 0; r < g->npe; r++) {
    double *a = R(g, ba, r), *b = R(g, bb, r);
    for (long t = 0; t < m; t++) a[offa + t] = b[offb + t];
  }
}
static void fill0(double *a, long off, int n) {
  for (long t = 0; t < (long)n * n * n; t++) a[off + t] = 0.0;
}
static void fill0_mpi(const mcfg *g, double *ba, long off, const int n[3]) {
  const long m = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
  for (int r = 0; r < g->npe; r++) {
    double *a = R(g, ba, r);
    for (long t = 0; t < m; t++) a[off + t] = 0.0;
  }
}

/* ------------------------------------------------------------------ */
/* fourn (SG8:1514-1584): in-place n-dimensional complex FFT of Numerical
 * Recipes, data 1-based interleaved (re, im), e^{+2 pi i ...} for isign=1 */
static void fourn(double *data, const int *nn, int ndim, int isign) {
  long ntot = 1;
  for (int d = 0; d < ndim; d++) ntot *= nn[d];
  long nprev = 1;
  for (int d = 0; d < ndim; d++) {
    const long n = nn[d], nrem = ntot / (n * nprev);
    const long ip1 = 2 * nprev, ip2 = ip1 * n, ip3 = ip2 * nrem;
    long i2rev = 1;
    for (long i2 = 1; i2 <= ip2; i2 += ip1) {
      if (i2 < i2rev) {
        for (long i1 = i2; i1 <= i2 + ip1 - 2; i1 += 2)
          for (long i3 = i1; i3 <= ip3; i3 += ip2) {
            const long i3rev = i2rev + i3 - i2;
            double t = data[i3];
            data[i3] = data[i3rev];
            data[i3rev] = t;
            t = data[i3 + 1];
            data[i3 + 1] = data[i3rev + 1];
            data[i3rev + 1] = t;
          }
      }
      long ibit = ip2 / 2;
      while (ibit >= ip1 && i2rev > ibit) {
        i2rev -= ibit;
        ibit /= 2;
      }
      i2rev += ibit;
    }
    long ifp1 = ip1;
    while (ifp1 < ip2) {
      const long ifp2 = 2 * ifp1;
      const double theta = (double)isign * 6.28318530717959 / (double)(ifp2 / ip1);
      const double sh = sin(0.5 * theta);
      const double wpr = -(2.0 * (sh * sh)), wpi = sin(theta);
      double wr = 1.0, wi = 0.0;
      for (long i3 = 1; i3 <= ifp1; i3 += ip1) {
        for (long i1 = i3; i1 <= i3 + ip1 - 2; i1 += 2)
          for (long i2 = i1; i2 <= ip3; i2 += ifp2) {
            const long k1 = i2, k2 = k1 + ifp1;
            const double tempr = wr * data[k2] - wi * data[k2 + 1];
            const double tempi = wr * data[k2 + 1] + wi * data[k2];
            data[k2] = data[k1] - tempr;
            data[k2 + 1] = data[k1 + 1] - tempi;
            data[k1] = data[k1] + tempr;
            data[k1 + 1] = data[k1 + 1] + tempi;
          }
        const double wtemp = wr;
        wr = wr * wpr - wi * wpi + wr;
        wi = wi * wpr + wtemp * wpi + wi;
      }
      ifp1 = ifp2;
    }
    nprev *= n;
  }
}

/* rlft3 (SG8:1466-1512) for 2D: data nn1 x nn2 reals seen as nn1/2 x nn2
 * complex values; speq the nn2 values at the first axis' Nyquist
 * frequency.  The unpacking twiddle w = CMPLX(wr, wi) is single precision. */
void mcfg_rlft3(double *data0, double *speq0, int nn1, int nn2, int isign) {
  double *data = data0 - 1, *speq = speq0 - 1; /* 1-based reals */
  const int h1n = nn1 / 2;
#define DRE(i1, i2) data[2 * ((i1) - 1) + 1 + (long)nn1 * ((i2) - 1)]
#define DIM(i1, i2) data[2 * ((i1) - 1) + 2 + (long)nn1 * ((i2) - 1)]
#define SRE(i2) speq[2 * ((i2) - 1) + 1]
#define SIM(i2) speq[2 * ((i2) - 1) + 2]
  const double c2i = -0.5 * (double)isign; /* c1 = (0.5, 0), c2 = (0, -0.5*isign) */
  const double theta = 6.28318530717959 / (double)(isign * nn1);
  const double sh = sin(0.5 * theta);
  const double wpr = -(2.0 * (sh * sh)), wpi = sin(theta);
  const int nn[2] = {h1n, nn2};
  if (isign == 1) {
    fourn(data, nn, 2, isign);
    for (int i2 = 1; i2 <= nn2; i2++) {
      SRE(i2) = DRE(1, i2);
      SIM(i2) = DIM(1, i2);
    }
  }
  double wr = 1.0, wi = 0.0, w_r = 0.0, w_i = 0.0;
  for (int i1 = 1; i1 <= nn1 / 4 + 1; i1++) {
    const int j1 = h1n - i1 + 2;
    for (int i2 = 1; i2 <= nn2; i2++) {
      const int j2 = (i2 != 1) ? nn2 - i2 + 2 : 1;
      double ar, ai, br, bi; /* a = first operand, b = conjg(second) */
      if (i1 == 1) {
        ar = DRE(1, i2); ai = DIM(1, i2);
        br = SRE(j2); bi = -SIM(j2);
      } else {
        ar = DRE(i1, i2); ai = DIM(i1, i2);
        br = DRE(j1, j2); bi = -DIM(j1, j2);
      }
      /* h1 = c1*(a+b), h2 = c2*(a-b) with the full complex products */
      const double sr = ar + br, si = ai + bi, dr = ar - br, di = ai - bi;
      const double h1r = 0.5 * sr - 0.0 * si, h1i = 0.5 * si + 0.0 * sr;
      const double h2r = 0.0 * dr - c2i * di, h2i = 0.0 * di + c2i * dr;
      if (i1 == 1) {
        DRE(1, i2) = h1r + h2r;
        DIM(1, i2) = h1i + h2i;
        SRE(j2) = h1r - h2r;
        SIM(j2) = -(h1i - h2i);
      } else {
        const double whr = w_r * h2r - w_i * h2i, whi = w_r * h2i + w_i * h2r;
        DRE(i1, i2) = h1r + whr;
        DIM(i1, i2) = h1i + whi;
        DRE(j1, j2) = h1r - whr;
        DIM(j1, j2) = -(h1i - whi);
      }
    }
    const double wtemp = wr;
    wr = wr * wpr - wi * wpi + wr;
    wi = wi * wpr + wtemp * wpi + wi;
    w_r = (double)(float)wr;
    w_i = (double)(float)wi;
  }
  if (isign == -1) fourn(data, nn, 2, isign);
#undef DRE
#undef DIM
#undef SRE
#undef SIM
}

/* ------------------------------------------------------------------ */
/* PB (SG8:1026-1463): boundary potential on the two x ends at the finest
 * level, then restricted to every level. */
static void pb(mcfg *g, const double *rho, const long st[3]) {
  const int nn1 = g->N[1], nn2 = g->N[2], C0 = g->C[0], P = g->S[1] * g->S[2];
  const double pi = 3.14159265359; /* SG8:1034 */
  const double dxx = g->dy1, dyy = g->dz1;
  const long npl = (long)nn1 * nn2;
  double *plane = (double *)malloc(sizeof(double) * npl), *speq = (double *)malloc(sizeof(double) * 2 * nn2);
  double *d[2], *s[2], *acc[2], *sacc[2], *prev[2], *sprev[2];
  for (int l = 0; l < 2; l++) {
    d[l] = (double *)malloc(sizeof(double) * npl);
    s[l] = (double *)malloc(sizeof(double) * 2 * nn2);
    acc[l] = (double *)calloc(npl, sizeof(double));
    sacc[l] = (double *)calloc(2 * nn2, sizeof(double));
    prev[l] = (double *)malloc(sizeof(double) * npl);
    sprev[l] = (double *)malloc(sizeof(double) * 2 * nn2);
  }
#define RHO(gi, gj, gk) rho[(long)((gi) + 1) + (long)((gj) + 1) * st[1] + (long)((gk) + 1) * st[2]]
  for (int r = 0; r < g->npe; r++) {
    int c[3];
    rank_coords(g, r, &c[0], &c[1], &c[2]);
    const int Nir = c[1] + g->S[1] * c[2];
    const double dzz = g->dxi0[c[0] * C0 + 1 + 1]; /* dx(1) of the rank */
    const double facG = -G4PI * dzz;
    for (int l = 0; l < 2; l++) {
      memset(prev[l], 0, sizeof(double) * npl);
      memset(sprev[l], 0, sizeof(double) * 2 * nn2);
    }
    for (int nlp2 = 0; nlp2 <= LOOPBC; nlp2++) {
      const int klr = Nir + 1 + P * nlp2;
      for (int l = 0; l < 2; l++) {
        memset(d[l], 0, sizeof(double) * npl);
        memset(s[l], 0, sizeof(double) * 2 * nn2);
      }
      if (klr <= C0) {
        const int gx = c[0] * C0 + klr; /* global x index of the plane */
        for (int kz = 1; kz <= nn2; kz++)
          for (int jy = 1; jy <= nn1; jy++) plane[(jy - 1) + (long)nn1 * (kz - 1)] = RHO(gx, jy, kz);
        mcfg_rlft3(plane, speq, nn1, nn2, 1);
        const double zp[2] = {g->xi0[gx + 1] - 0.5 * dzz, 0.0};
        const double zpp[2] = {zp[0], g->Lbox - zp[0]};
#define PD(ri, m) plane[(ri) - 1 + (long)nn1 * ((m) - 1)]
#define DD(l, ri, m) d[l][(ri) - 1 + (long)nn1 * ((m) - 1)]
        double tre[2], tim[2];
        for (int l = 0; l < 2; l++) {
          tre[l] = DD(l, 1, 1) - PD(1, 1) * 0.5 * zpp[l] * facG;
          tim[l] = DD(l, 2, 1) - PD(2, 1) * 0.5 * zpp[l] * facG;
        }
        /* weight of mode (li, mi) at distance z: 0.5*exp(-z*kap)/kap*facG */
#define KAP2(li, mi) (4.0 * (sin(pi * (double)((li) - 1) / (double)nn1) * sin(pi * (double)((li) - 1) / (double)nn1) / (dxx * dxx) + \
                             sin(pi * (double)((mi) - 1) / (double)nn2) * sin(pi * (double)((mi) - 1) / (double)nn2) / (dyy * dyy)))
        for (int m = 2; m <= nn2 / 2 + 1; m++)
          for (int li = 1; li <= nn1 / 2; li++) {
            const double kap = sqrt(KAP2(li, m));
            for (int l = 0; l < 2; l++) {
              DD(l, 2 * li - 1, m) = DD(l, 2 * li - 1, m) + PD(2 * li - 1, m) * 0.5 * exp(-zpp[l] * kap) / kap * facG;
              DD(l, 2 * li, m) = DD(l, 2 * li, m) + PD(2 * li, m) * 0.5 * exp(-zpp[l] * kap) / kap * facG;
            }
          }
        for (int li = 2; li <= nn1 / 2; li++) {
          const double kap = sqrt(KAP2(li, 1));
          for (int l = 0; l < 2; l++) {
            DD(l, 2 * li - 1, 1) = DD(l, 2 * li - 1, 1) + PD(2 * li - 1, 1) * 0.5 * exp(-zpp[l] * kap) / kap * facG;
            DD(l, 2 * li, 1) = DD(l, 2 * li, 1) + PD(2 * li, 1) * 0.5 * exp(-zpp[l] * kap) / kap * facG;
          }
        }
        {
          const int li = nn1 / 2 + 1;
          for (int m = 1; m <= nn2 / 2 + 1; m++) {
            const double kap = sqrt(sqrt(KAP2(li, m)) + 1.0e-20);
            for (int l = 0; l < 2; l++) {
              s[l][2 * (m - 1)] = s[l][2 * (m - 1)] + speq[2 * (m - 1)] * 0.5 * exp(-zpp[l] * kap) / kap * facG;
              s[l][2 * (m - 1) + 1] = s[l][2 * (m - 1) + 1] + speq[2 * (m - 1) + 1] * 0.5 * exp(-zpp[l] * kap) / kap * facG;
            }
          }
        }
        for (int m2 = nn2 / 2 + 2; m2 <= nn2; m2++) {
          const int m = nn2 + 2 - m2;
          for (int li = 1; li <= nn1 / 2; li++) {
            const double kap = sqrt(KAP2(li, m));
            for (int l = 0; l < 2; l++) {
              DD(l, 2 * li - 1, m2) = DD(l, 2 * li - 1, m2) + PD(2 * li - 1, m2) * 0.5 * exp(-zpp[l] * kap) / kap * facG;
              DD(l, 2 * li, m2) = DD(l, 2 * li, m2) + PD(2 * li, m2) * 0.5 * exp(-zpp[l] * kap) / kap * facG;
            }
          }
        }
        {
          const int li = nn1 / 2 + 1;
          for (int m2 = nn2 / 2 + 2; m2 <= nn2; m2++) {
            const int m = nn2 + 2 - m2;
            const double kap = sqrt(KAP2(li, m));
            for (int l = 0; l < 2; l++) {
              s[l][2 * (m2 - 1)] = s[l][2 * (m2 - 1)] + speq[2 * (m2 - 1)] * 0.5 * exp(-zpp[l] * kap) / kap * facG;
              s[l][2 * (m2 - 1) + 1] =
                  s[l][2 * (m2 - 1) + 1] + speq[2 * (m2 - 1) + 1] * 0.5 * exp(-zpp[l] * kap) / kap * facG;
            }
          }
        }
        for (int l = 0; l < 2; l++) {
          DD(l, 1, 1) = tre[l];
          DD(l, 2, 1) = tim[l];
        }
#undef KAP2
#undef PD
#undef DD
      }
      for (int l = 0; l < 2; l++) {
        for (long t = 0; t < npl; t++) d[l][t] = d[l][t] + prev[l][t];
        for (int t = 0; t < 2 * nn2; t++) s[l][t] = s[l][t] + sprev[l][t];
        memcpy(prev[l], d[l], sizeof(double) * npl);
        memcpy(sprev[l], s[l], sizeof(double) * 2 * nn2);
      }
    }
    /* MPI_ALLREDUCE(SUM), taken in rank order */
    for (int l = 0; l < 2; l++) {
      if (r == 0) {
        memcpy(acc[l], d[l], sizeof(double) * npl);
        memcpy(sacc[l], s[l], sizeof(double) * 2 * nn2);
      } else {
        for (long t = 0; t < npl; t++) acc[l][t] = acc[l][t] + d[l][t];
        for (int t = 0; t < 2 * nn2; t++) sacc[l][t] = sacc[l][t] + s[l][t];
      }
    }
  }
#undef RHO
  const double fac = (double)(2 * (nn1 / 2) * (nn1 / 2));
  for (int l = 0; l < 2; l++) {
    mcfg_rlft3(acc[l], sacc[l], nn1, nn2, -1);
    for (long t = 0; t < npl; t++) acc[l][t] = acc[l][t] / fac;
  }
  /* finest level, every rank's patch (SG8:1259-1272) */
  const int ncy = g->C[1] + 1, ncz = g->C[2] + 1;
  for (int r = 0; r < g->npe; r++) {
    int c[3];
    rank_coords(g, r, &c[0], &c[1], &c[2]);
    double *b = g->bphi2 + (size_t)r * 2 * (size_t)(g->Needb2 + 1);
    for (int k = 0; k <= ncz; k++)
      for (int j = 0; j <= ncy; j++) {
        const long nidx = j + (long)(ncy + 1) * k;
        int jb = c[1] * g->C[1] + j, kb = c[2] * g->C[2] + k;
        if (j == ncy && c[1] == g->S[1] - 1) jb = 1;
        if (k == ncz && c[2] == g->S[2] - 1) kb = 1;
        if (j == 0 && c[1] == 0) jb = nn1;
        if (k == 0 && c[2] == 0) kb = nn2;
        for (int l = 0; l < 2; l++) b[l * (g->Needb2 + 1) + g->pointb2[g->NGL] + nidx] = acc[l][(jb - 1) + (long)nn1 * (kb - 1)];
      }
  }
  for (int l = 0; l < 2; l++) { free(d[l]); free(s[l]); free(acc[l]); free(sacc[l]); free(prev[l]); free(sprev[l]); }
  free(plane);
  free(speq);

  /* restriction of the end planes over the split levels (SG8:1288-1363):
   * 4/8 centre + 1/8 each in-plane neighbour, a missing neighbour's weight
   * moved to the centre at the first and last node of the block */
  const double w = 0.125;
  int nc[3] = {g->C[0] + 1, g->C[1] + 1, g->C[2] + 1};
  for (int lc = g->NGL - 1; lc >= g->NGcr - 1; lc--) {
    int nf[3] = {nc[0], nc[1], nc[2]};
    for (int d2 = 0; d2 < 3; d2++) nc[d2] = nc[d2] / 2 + 1;
    const int mfx = nf[1], mcx = nc[1], mcy = nc[2];
    for (int r = 0; r < g->npe; r++) {
      double *b = g->bphi2 + (size_t)r * 2 * (size_t)(g->Needb2 + 1);
      for (int l = 0; l < 2; l++) {
        double *bl = b + l * (g->Needb2 + 1);
        const long pf = g->pointb2[lc + 1], pc = g->pointb2[lc];
        for (int jc = 1; jc <= mcy; jc++)
          for (int ic = 1; ic <= mcx; ic++) {
            const int fi = 2 * ic - 1, fj = 2 * jc - 1;
            const long kf = fi + (long)(mcx * 2) * fj, kc = ic + (long)(mcx + 1) * jc;
            const int hl = ic > 1, hr = ic < mcx, hd = jc > 1, hu = jc < mcy;
            const double cw = (double)(4 + !hl + !hr + !hd + !hu) * w;
            double sum = 0.0;
            int first = 1;
            const long nb[4] = {kf - 1, kf + 1, kf - mfx - 1, kf + mfx + 1};
            const int has[4] = {hl, hr, hd, hu};
            for (int q = 0; q < 4; q++)
              if (has[q]) {
                sum = first ? bl[pf + nb[q]] : sum + bl[pf + nb[q]];
                first = 0;
              }
            bl[pc + kc] = cw * bl[pf + kf] + w * sum;
          }
      }
    }
  }
  /* gather the level-NGcr end planes into the whole-domain arrays
   * (SG8:1366-1390) */
  const int n1 = (1 << g->NGcr) + 1;
  for (int r = 0; r < g->npe; r++) {
    int c[3];
    rank_coords(g, r, &c[0], &c[1], &c[2]);
    const int *m = g->ncr;
    for (int l = 0; l < 2; l++) {
      if (l == 0 && c[0] != 0) continue;
      if (l == 1 && c[0] != g->S[0] - 1) continue;
      const double *bl = g->bphi2 + (size_t)r * 2 * (size_t)(g->Needb2 + 1) + l * (g->Needb2 + 1);
      const int edy = c[1] == g->S[1] - 1 ? m[1] : m[1] - 1, edz = c[2] == g->S[2] - 1 ? m[2] : m[2] - 1;
      for (int kk = 1; kk <= edz; kk++)
        for (int jj = 1; jj <= edy; jj++) {
          const int j = c[1] * (m[1] - 1) + jj, k = c[2] * (m[2] - 1) + kk;
          g->bphi1[l * (g->Needb1 + 1) + g->pointb1[g->NGcr] + (j - 1) + (long)n1 * (k - 1)] =
              bl[g->pointb2[g->NGcr] + jj + (long)(m[1] + 1) * kk];
        }
    }
  }
  /* whole-domain levels (SG8:1394-1460) */
  int ncu = n1;
  for (int lc = g->NGcr - 1; lc >= 1; lc--) {
    const int nf = ncu;
    ncu = ncu / 2 + 1;
    for (int l = 0; l < 2; l++) {
      double *bl = g->bphi1 + l * (g->Needb1 + 1);
      const long pf = g->pointb1[lc + 1], pc = g->pointb1[lc];
      for (int jc = 1; jc <= ncu; jc++)
        for (int ic = 1; ic <= ncu; ic++) {
          const int fi = 2 * ic - 1, fj = 2 * jc - 1;
          const long kf = fi - 1 + (long)(ncu * 2 - 1) * (fj - 1), kc = ic - 1 + (long)ncu * (jc - 1);
          const int hl = ic > 1, hr = ic < ncu, hd = jc > 1, hu = jc < ncu;
          const double cw = (double)(4 + !hl + !hr + !hd + !hu) * w;
          double sum = 0.0;
          int first = 1;
          const long nb[4] = {kf - 1, kf + 1, kf - nf, kf + nf};
          const int has[4] = {hl, hr, hd, hu};
          for (int q = 0; q < 4; q++)
            if (has[q]) {
              sum = first ? bl[pf + nb[q]] : sum + bl[pf + nb[q]];
              first = 0;
            }
          bl[pc + kc] = cw * bl[pf + kf] + w * sum;
        }
    }
  }
}

/* ------------------------------------------------------------------ */
static void mglin(mcfg *g, const double *rho, const long st[3]) {
  const int NGL = g->NGL, NGcr = g->NGcr;
  const long *P1 = g->point1, *P2 = g->point2;
  /* right-hand side at the finest level: 4 pi G rho on nodes 0..C+1 */
  {
    const int nfin[3] = {g->C[0] + 1, g->C[1] + 1, g->C[2] + 1};
    for (int r = 0; r < g->npe; r++) {
      int c[3];
      rank_coords(g, r, &c[0], &c[1], &c[2]);
      double *a = R(g, g->crhs2, r);
      for (int k = 0; k <= nfin[2]; k++)
        for (int j = 0; j <= nfin[1]; j++)
          for (int i = 0; i <= nfin[0]; i++) {
            const int gi = c[0] * g->C[0] + i, gj = c[1] * g->C[1] + j, gk = c[2] * g->C[2] + k;
            U2(a, P2[NGL], nfin, i, j, k) = rho[(long)(gi + 1) + (long)(gj + 1) * st[1] + (long)(gk + 1) * st[2]] * G4PI;
          }
    }
  }
  int n[3] = {(g->C[0] + 1) / 2 + 1, (g->C[1] + 1) / 2 + 1, (g->C[2] + 1) / 2 + 1};
  int ngrid = NGL - 1;
  rstrct_mpi(g, g->crho2, P2[ngrid], g->crhs2, P2[ngrid + 1], n, 0);
  while (ngrid != NGcr) {
    for (int d = 0; d < 3; d++) n[d] = n[d] / 2 + 1;
    ngrid--;
    rstrct_mpi(g, g->crho2, P2[ngrid], g->crho2, P2[ngrid + 1], n, 0);
  }
  const int *ncr = g->ncr;
  const int N1 = (1 << NGcr) + 1;
  collect(g, g->crho1, P1[NGcr], N1, g->crho2, P2[NGcr], ncr);
  int nu = N1;
  while (ngrid != 1) {
    nu = nu / 2 + 1;
    ngrid--;
    rstrct(g, g->crho1, P1[ngrid], g->crho1, P1[ngrid + 1], nu, 0);
  }
  slvsmlb(g, g->cphi1, P1[1], g->crho1, P1[1]);
  /* the level sizes as the reference tracks them: whole-domain nu, split n */
  int split = 0;
  for (int j = 2; j <= NGL; j++) {
    if (j <= NGcr) {
      nu = nu * 2 - 1;
      interp(g, g->cphi1, P1[j], g->cphi1, P1[j - 1], nu, g->pointb1[j], 1);
      copy1(g, g->crhs1, P1[j], g->crho1, P1[j], nu);
      if (j == NGcr) {
        for (int d = 0; d < 3; d++) n[d] = ncr[d];
        split = 1;
        divide(g, g->cphi2, P2[NGcr], ncr, g->cphi1, P1[NGcr], N1);
        copy_mpi(g, g->crhs2, P2[NGcr], g->crho2, P2[NGcr], ncr);
      }
    } else {
      for (int d = 0; d < 3; d++) n[d] = n[d] * 2 - 1;
      split = 1;
      interp_mpi(g, g->cphi2, P2[j], g->cphi2, P2[j - 1], n, g->pointb2[j], 1);
      if (j != NGL) copy_mpi(g, g->crhs2, P2[j], g->crho2, P2[j], n);
    }
    for (int cyc = 1; cyc <= NCYCLE; cyc++) {
      int nf[3], nfu = 0;
      if (split) for (int d = 0; d < 3; d++) nf[d] = n[d];
      else nfu = nu;
      for (int jj = j; jj >= 2; jj--) {
        if (jj < NGcr) {
          for (int p = 0; p < NPRE; p++) relax(g, g->cphi1, P1[jj], g->crhs1, P1[jj], nfu);
          resid(g, g->cres1, P1[jj], g->cphi1, P1[jj], g->crhs1, P1[jj], nfu);
          nfu = nfu / 2 + 1;
          rstrct(g, g->crhs1, P1[jj - 1], g->cres1, P1[jj], nfu, 1);
          fill0(g->cphi1, P1[jj - 1], nfu);
        } else {
          const int npre1 = j >= NGL ? 2 : NPRE;
          for (int p = 0; p < npre1; p++) relax_mpi(g, g->cphi2, P2[jj], g->crhs2, P2[jj], nf);
          resid_mpi(g, g->cres2, P2[jj], g->cphi2, P2[jj], g->crhs2, P2[jj], nf);
          if (jj == NGcr) {
            collect(g, g->cphi1, P1[NGcr], N1, g->cphi2, P2[NGcr], ncr);
            collect(g, g->cres1, P1[NGcr], N1, g->cres2, P2[NGcr], ncr);
            nfu = N1 / 2 + 1;
            rstrct(g, g->crhs1, P1[jj - 1], g->cres1, P1[jj], nfu, 1);
            fill0(g->cphi1, P1[jj - 1], nfu);
          } else {
            for (int d = 0; d < 3; d++) nf[d] = nf[d] / 2 + 1;
            rstrct_mpi(g, g->crhs2, P2[jj - 1], g->cres2, P2[jj], nf, 1);
            fill0_mpi(g, g->cphi2, P2[jj - 1], nf);
          }
        }
      }
      slvsml(g, g->cphi1, P1[1], g->crhs1, P1[1]);
      nfu = 3;
      for (int jj = 2; jj <= j; jj++) {
        if (jj <= NGcr) {
          nfu = 2 * nfu - 1;
          addint(g, g->cphi1, P1[jj], g->cphi1, P1[jj - 1], g->cres1, P1[jj], nfu);
          if (jj == NGcr) {
            for (int d = 0; d < 3; d++) nf[d] = ncr[d];
            divide(g, g->cphi2, P2[NGcr], ncr, g->cphi1, P1[NGcr], N1);
            for (int p = 0; p < NPOST; p++) relax_mpi(g, g->cphi2, P2[jj], g->crhs2, P2[jj], nf);
          } else {
            for (int p = 0; p < NPOST; p++) relax(g, g->cphi1, P1[jj], g->crhs1, P1[jj], nfu);
          }
        } else {
          for (int d = 0; d < 3; d++) nf[d] = 2 * nf[d] - 1;
          addint_mpi(g, g->cphi2, P2[jj], g->cphi2, P2[jj - 1], g->cres2, P2[jj], nf);
          const int npost1 = j >= NGL ? 2 : NPOST;
          for (int p = 0; p < npost1; p++) relax_mpi(g, g->cphi2, P2[jj], g->crhs2, P2[jj], nf);
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
mcfg *mcfg_create(const int n[3], const int s[3], double lbox, const double *const xi[3], const double *const dxi[3]) {
  for (int d = 0; d < 3; d++) {
    if (n[d] < 4 || (n[d] & (n[d] - 1)) || s[d] < 1 || n[d] % s[d] || (s[d] & (s[d] - 1))) return NULL;
    if (n[d] / s[d] < 2) return NULL;
  }
  mcfg *g = (mcfg *)calloc(1, sizeof(mcfg));
  for (int d = 0; d < 3; d++) { g->N[d] = n[d]; g->S[d] = s[d]; g->C[d] = n[d] / s[d]; }
  g->npe = s[0] * s[1] * s[2];
  int mn = n[0];
  for (int d = 1; d < 3; d++) if (n[d] < mn) mn = n[d];
  g->NGL = ilog2(mn);
  int ms = s[0];
  for (int d = 1; d < 3; d++) if (s[d] > ms) ms = s[d];
  g->NGcr = ilog2(ms) + 1;
  if (g->NGcr > g->NGL || g->NGL > 15) { free(g); return NULL; }
  /* pinter (SG8:102-146) */
  g->point1[1] = 1;
  {
    long m = 3;
    for (int nl = 1; nl <= g->NGcr; nl++) { g->point1[nl + 1] = g->point1[nl] + m * m * m; m = 2 * m - 1; }
  }
  g->Need1 = g->point1[g->NGcr + 1];
  {
    long m[3];
    for (int d = 0; d < 3; d++) m[d] = (1 << g->NGcr) / s[d] + 2;
    g->point2[g->NGcr - 1] = 1;
    for (int nl = g->NGcr - 1; nl <= g->NGL; nl++) {
      g->point2[nl + 1] = g->point2[nl] + m[0] * m[1] * m[2];
      for (int d = 0; d < 3; d++) m[d] = 2 * m[d] - 2;
    }
  }
  g->Need2 = g->point2[g->NGL + 1];
  /* PB's boundary pointers (SG8:1058-1080) */
  g->pointb1[1] = 1;
  {
    long m = 3;
    for (int nl = 1; nl <= g->NGcr; nl++) { g->pointb1[nl + 1] = g->pointb1[nl] + m * m; m = 2 * m - 1; }
  }
  g->Needb1 = g->pointb1[g->NGcr + 1];
  {
    long m[3];
    for (int d = 0; d < 3; d++) m[d] = (1 << g->NGcr) / s[d] + 2;
    g->pointb2[g->NGcr - 1] = 1;
    for (int nl = g->NGcr - 1; nl <= g->NGL; nl++) {
      long a = m[0] * m[1], b = m[1] * m[2], c = m[2] * m[0];
      long mx = a > b ? a : b;
      mx = mx > c ? mx : c;
      g->pointb2[nl + 1] = g->pointb2[nl] + mx;
      for (int d = 0; d < 3; d++) m[d] = 2 * m[d] - 2;
    }
  }
  g->Needb2 = g->pointb2[g->NGL + 1];
  for (int d = 0; d < 3; d++) {
    int m = g->C[d] + 1;
    for (int l = g->NGL; l > g->NGcr; l--) m = m / 2 + 1;
    g->ncr[d] = m;
  }
  g->Lbox = lbox;
  g->xi0 = (double *)malloc(sizeof(double) * (n[0] + 4));
  g->dxi0 = (double *)malloc(sizeof(double) * (n[0] + 4));
  memcpy(g->xi0, xi[0], sizeof(double) * (n[0] + 4));
  memcpy(g->dxi0, dxi[0], sizeof(double) * (n[0] + 4));
  g->dy1 = dxi[1][1 + 1];
  g->dz1 = dxi[2][1 + 1];
  double **one[] = {&g->cphi1, &g->crho1, &g->cres1, &g->crhs1};
  for (int t = 0; t < 4; t++) *one[t] = (double *)calloc(g->Need1 + 1, sizeof(double));
  double **two[] = {&g->cphi2, &g->crho2, &g->cres2, &g->crhs2};
  for (int t = 0; t < 4; t++) *two[t] = (double *)calloc((size_t)g->npe * (g->Need2 + 1), sizeof(double));
  g->bphi1 = (double *)calloc(2 * (g->Needb1 + 1), sizeof(double));
  g->bphi2 = (double *)calloc((size_t)g->npe * 2 * (g->Needb2 + 1), sizeof(double));
  return g;
}

void mcfg_destroy(mcfg *g) {
  if (!g) return;
  free(g->cphi1); free(g->crho1); free(g->cres1); free(g->crhs1);
  free(g->cphi2); free(g->crho2); free(g->cres2); free(g->crhs2);
  free(g->bphi1); free(g->bphi2); free(g->xi0); free(g->dxi0);
  free(g);
}

const double *mcfg_boundary(const mcfg *g, int rank, int side) {
  if (rank < 0 || rank >= g->npe || side < 0 || side > 1) return NULL;
  return g->bphi2 + (size_t)rank * 2 * (size_t)(g->Needb2 + 1) + side * (g->Needb2 + 1) + g->pointb2[g->NGL];
}

void mcfg_solve(mcfg *g, const double *rho, double *phi) {
  const long st[3] = {1, g->N[0] + 4, (long)(g->N[0] + 4) * (g->N[1] + 4)};
  double *all[] = {g->cphi1, g->crho1, g->cres1, g->crhs1};
  for (int t = 0; t < 4; t++) memset(all[t], 0, sizeof(double) * (g->Need1 + 1));
  double *all2[] = {g->cphi2, g->crho2, g->cres2, g->crhs2};
  for (int t = 0; t < 4; t++) memset(all2[t], 0, sizeof(double) * (size_t)g->npe * (g->Need2 + 1));
  memset(g->bphi1, 0, sizeof(double) * 2 * (g->Needb1 + 1));
  memset(g->bphi2, 0, sizeof(double) * (size_t)g->npe * 2 * (g->Needb2 + 1));
  pb(g, rho, st);
  mglin(g, rho, st);
  /* Phi at the cells: the owning rank's nodes 1..C (SG8:282-291) */
  const int nfin[3] = {g->C[0] + 1, g->C[1] + 1, g->C[2] + 1};
  const int Nx = g->N[0], Ny = g->N[1], Nz = g->N[2];
#define PHI(i, j, k) phi[(long)((i) + 1) + (long)((j) + 1) * st[1] + (long)((k) + 1) * st[2]]
  for (int r = 0; r < g->npe; r++) {
    int c[3];
    rank_coords(g, r, &c[0], &c[1], &c[2]);
    const double *a = R(g, g->cphi2, r);
    for (int k = 1; k <= g->C[2]; k++)
      for (int j = 1; j <= g->C[1]; j++)
        for (int i = 1; i <= g->C[0]; i++)
          PHI(c[0] * g->C[0] + i, c[1] * g->C[1] + j, c[2] * g->C[2] + k) = U2(a, g->point2[g->NGL], nfin, i, j, k);
  }
  /* face ghosts (SG8:22-67): periodic images in y and z, the x ends copied
   * from the first / last cell (grad = 0) */
  for (int k = 1; k <= Nz; k++)
    for (int j = 1; j <= Ny; j++) {
      PHI(0, j, k) = PHI(1, j, k);
      PHI(-1, j, k) = PHI(1, j, k);
      PHI(Nx + 1, j, k) = PHI(Nx, j, k);
      PHI(Nx + 2, j, k) = PHI(Nx, j, k);
    }
  for (int k = 1; k <= Nz; k++)
    for (int i = 1; i <= Nx; i++) {
      PHI(i, -1, k) = PHI(i, Ny - 1, k);
      PHI(i, 0, k) = PHI(i, Ny, k);
      PHI(i, Ny + 1, k) = PHI(i, 1, k);
      PHI(i, Ny + 2, k) = PHI(i, 2, k);
    }
  for (int j = 1; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      PHI(i, j, -1) = PHI(i, j, Nz - 1);
      PHI(i, j, 0) = PHI(i, j, Nz);
      PHI(i, j, Nz + 1) = PHI(i, j, 1);
      PHI(i, j, Nz + 2) = PHI(i, j, 2);
    }
#undef PHI
}
