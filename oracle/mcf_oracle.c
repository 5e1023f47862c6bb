/*
 * mcf_oracle.c -- CPU restatement (plain C, FP64) of one timestep of the
 * MCF3DSG* 3D MHD code.  TEST INFRASTRUCTURE ONLY: it is the checker for
 * the HIP path and the timed CPU baseline; the product never links it.
 *
 * Each routine follows the reference routine named in its comment
 * (M3 = 3DMHD-CfCA/MCF3DSGincfca3.f90).  Expressions keep the reference's
 * evaluation order (Fortran evaluates left to right within a precedence
 * level and flang does not reassociate), build with -ffp-contract=off so no
 * multiply-add is fused, and the intrinsics follow flang's lowering:
 * dmax1(a,b) = (a>b ? a : b), dmin1(a,b) = (a<b ? a : b), dsign = copysign,
 * x**2.d0 = x*x, 1.d1**y = pow(10,y), dexp/dsqrt/dlog10 = libm exp/sqrt/log10.
 *
 * Single block: every MPI halo exchange becomes an in-place ghost fill with
 * the reference's pass order (x over all j,k; then y over all i,k; then z
 * over whole planes) and its physical x-boundary rules (BC=4 inflow).
 *
 * Threads (OpenMP): the pencil and plane loops of every stage are split over
 * threads.  Each pencil, cell, face or edge is computed by one thread with
 * exactly the serial expressions (per-thread pencil work arrays), and the
 * dt minima are combined with the same comparisons, so the result does not
 * depend on the thread count (tests/test_oracle_golden.py checks 1 and 4).
 */
#include "mcf_oracle.h"
#include "mcf_grav.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KB 8.63359            /* M3:11 */
#define KCOND 1.6384e-2       /* M3:11 */
#define EPSL 1.0e-10          /* VLIMIT eps, M3:1697 */

#define DMAX(a, b) ((a) > (b) ? (a) : (b))
#define DMIN(a, b) ((a) < (b) ? (a) : (b))

enum { SH = 0, SP = 1, SH2 = 2, SHE = 3, SHEP = 4, SC = 5, SCO = 6, SCP = 7 };

struct mcfo {
  mcfo_params p;
  int n[3];
  long st[3];
  long ncell;
  int nmax;
  double *gx[3], *gdx[3]; /* offset by one: gx[d][i+1] = x(i) */
  double bbrv[11][3][3];  /* 1-based BBRV(k,mode,side) */
  double bbrv_cm[8];
  double dtf;
  double pmin, pmax, rmin, spmin[8];
  double time;
  int ifevo;
  double *f[MCFO_NFIELD + 1];
  double *blg[2], *vfc1, *xlag, *dxlagM;
  double *vfc[3], *dnc, *emf[3], *dq[4];
  double *Tn, *Pn, *Qx, *Qy, *Qz;
  double *ppre; /* U(:,:,:,5) right after the RHS ghost fill of a sweep */
  float *dtfa;  /* inhomogeneous inflow density field DTF (REAL*4), NULL = uniform */
  double *col[10]; /* Ntot, NH2, NnC, NCO, tCII seen from the left / right x end (col[2q+dir]) */
  int nsp[3];   /* emulated MPI block split */
  double *phi;  /* gravitational potential Phi (ifgrv = 2), same layout as U */
  mcfg *grav;   /* the multigrid solver of GRAVTY (mcf_grav.c) */
  double *pw;   /* pencil work pool, one slice of 64 arrays per thread */
  int nthr;     /* threads the pool has slices for */
};

static int nthreads_max(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
static int thread_id(void) {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

#define U(k) (o->f[(k)])
#define SPC(s) (o->f[9 + (s)])
#define BCC(c) (o->f[16 + (c)])

static inline long ix3(const mcfo *o, int i, int j, int k) {
  return (long)(i + 1) + (long)(j + 1) * o->st[1] + (long)(k + 1) * o->st[2];
}
static inline double xg(const mcfo *o, int d, int i) { return o->gx[d][i + 1]; }
static inline double dxg(const mcfo *o, int d, int i) { return o->gdx[d][i + 1]; }

void mcfo_default_params(mcfo_params *p) {
  memset(p, 0, sizeof(*p));
  p->np1x = p->np2x = p->np1y = p->np2y = p->np1z = p->np2z = 32;
  p->ql1x = p->ql2x = p->ql1y = p->ql2y = p->ql1z = p->ql2z = 10.0;
  p->vinitx1 = 9.778; p->vinitx2 = -9.778;
  p->binitx1 = p->binitx2 = 0.0; p->binity1 = p->binity2 = 2.2307;
  p->cfl = 0.49; p->facdep = 1.0;
  p->maxstp = 4; p->nitera = 5000; p->tfinal = 100.0;
  p->bcx1 = p->bcx2 = 4; p->bcy1 = p->bcy2 = p->bcz1 = p->bcz2 = 3;
  p->ifchem = 0; p->ifthrm = 2; p->ifrad = 0; p->ifgrv = 0;
  /* M3:155-157 */
  p->pinit_over_kb = 8.810807e3;
  p->Hini = 0.9219098; p->pini = 0.9503446e-2; p->H2ini = 0.9465513e-8;
  p->Heini = 0.9155226e-1; p->Hepini = 0.5655353e-3;
  p->Cini = 0.1565848e-8; p->COini = 0.2202631e-20; p->Cpini = 0.1433520e-3;
  p->nsplitx = p->nsplity = p->nsplitz = 2;
}

/* INITIA (M3:114-486): grid, floors, boundary states BBRV/BBRV_cm and DTF. */
static void setup_constants(mcfo *o) {
  const mcfo_params *p = &o->p;
  const double mH = 1.0, mHe = 4.0, mH2 = 2.0, xc = 1.4e-4;
  double Hini = p->Hini, pini = p->pini, H2ini = p->H2ini, Heini = p->Heini, Hepini = p->Hepini;
  double pinit1 = p->pinit_over_kb * KB * 1.0e-3;
  double dinit1 = mH * Hini + mH * pini + mH2 * H2ini + mHe * Heini + mHe * Hepini;
  double gamma = (5.0 * (Hini + pini + Heini + Hepini) + 7.0 * H2ini) /
                 (3.0 * (Hini + pini + Heini + Hepini) + 5.0 * H2ini);
  double gammi1 = gamma - 1.0;
  o->pmin = 1.829797 * 8.6336;
  o->pmax = 1.0e10;
  o->rmin = 0.1949628;
  o->spmin[SH] = o->rmin * 0.91; o->spmin[SP] = 1.0e-20; o->spmin[SH2] = 1.0e-20;
  o->spmin[SHE] = o->rmin * 0.09; o->spmin[SHEP] = 1.0e-20; o->spmin[SC] = 1.0e-20;
  o->spmin[SCO] = 1.0e-20; o->spmin[SCP] = 1.0e-20;
  double *cm = o->bbrv_cm;
  cm[0] = 0.91 / 1.27;
  cm[1] = 0.91 * pini / Hini / 1.27;
  cm[2] = 0.91 * H2ini / Hini / 1.27;
  cm[3] = 0.09 / 1.27;
  cm[4] = 0.09 * Hepini / Heini / 1.27;
  cm[7] = xc / 1.27;
  cm[5] = xc * p->Cini / p->Cpini / 1.27;
  cm[6] = xc * p->COini / p->Cpini / 1.27;
  double dBC = mH * cm[0] + mH * cm[1] + mH2 * cm[2] + mHe * cm[3] + mHe * cm[4];
  double (*B)[3][3] = o->bbrv;
  memset(o->bbrv, 0, sizeof(o->bbrv));
  B[1][1][1] = dBC; B[1][2][1] = dBC; B[1][1][2] = dBC; B[1][2][2] = dBC;
  B[2][1][1] = p->vinitx1; B[2][2][1] = dBC * p->vinitx1; B[2][1][2] = -p->vinitx1; B[2][2][2] = -dBC * p->vinitx1;
  B[3][1][1] = p->vinity1; B[3][2][1] = dBC * p->vinity1; B[3][1][2] = p->vinity1; B[3][2][2] = dBC * p->vinity1;
  B[4][1][1] = p->vinitz1; B[4][2][1] = dBC * p->vinitz1; B[4][1][2] = p->vinitz1; B[4][2][2] = dBC * p->vinitz1;
  double e5 = pinit1 / gammi1 + 0.5 * (dBC * (p->vinitx1 * p->vinitx1) + p->binitx1 * p->binitx1 +
                                       p->binity1 * p->binity1 + p->binitz1 * p->binitz1);
  B[5][1][1] = pinit1; B[5][2][1] = e5; B[5][1][2] = pinit1; B[5][2][2] = e5;
  B[6][1][1] = p->binitx1; B[6][2][1] = p->binitx1; B[6][1][2] = p->binitx2; B[6][2][2] = p->binitx2;
  B[7][1][1] = p->binity1; B[7][2][1] = p->binity1; B[7][1][2] = p->binity2; B[7][2][2] = p->binity2;
  B[8][1][1] = p->binitz1; B[8][2][1] = p->binitz1; B[8][1][2] = p->binitz2; B[8][2][2] = p->binitz2;
  /* DTF is REAL*4 (M3:22) set to dinit1 (M3:405); ghosts read dble(DTF). */
  o->dtf = (double)(float)dinit1;

  /* grid (M3:255-298), single block */
  int np1[3] = {p->np1x, p->np1y, p->np1z};
  double ql1[3] = {p->ql1x, p->ql1y, p->ql1z};
  for (int d = 0; d < 3; d++) {
    int n = o->n[d];
    for (int i = -1; i <= n + 2; i++) {
      double v;
      if (d == 0) v = (i <= p->np1x) ? p->ql1x / (double)p->np1x : p->ql2x / (double)p->np2x;
      else v = ql1[d] / (double)np1[d];
      o->gdx[d][i + 1] = v;
    }
    o->gx[d][0] = -o->gdx[d][1];
    for (int i = 0; i <= n + 2; i++) o->gx[d][i + 1] = o->gx[d][i] + o->gdx[d][i + 1];
  }
}

mcfo *mcfo_create(const mcfo_params *p) {
  mcfo *o = (mcfo *)calloc(1, sizeof(mcfo));
  o->p = *p;
  o->n[0] = p->np1x + p->np2x;
  o->n[1] = p->np1y + p->np2y;
  o->n[2] = p->np1z + p->np2z;
  o->st[0] = 1;
  o->st[1] = o->n[0] + 4;
  o->st[2] = (long)(o->n[0] + 4) * (o->n[1] + 4);
  o->ncell = o->st[2] * (o->n[2] + 4);
  o->nmax = o->n[0];
  if (o->n[1] > o->nmax) o->nmax = o->n[1];
  if (o->n[2] > o->nmax) o->nmax = o->n[2];
  for (int d = 0; d < 3; d++) {
    o->gx[d] = (double *)calloc(o->n[d] + 4, sizeof(double));
    o->gdx[d] = (double *)calloc(o->n[d] + 4, sizeof(double));
  }
  for (int k = 1; k <= MCFO_NFIELD; k++) o->f[k] = (double *)calloc(o->ncell, sizeof(double));
  double **tmp[] = {&o->blg[0], &o->blg[1], &o->vfc1, &o->xlag, &o->dxlagM, &o->vfc[0], &o->vfc[1],
                    &o->vfc[2], &o->dnc, &o->emf[0], &o->emf[1], &o->emf[2], &o->dq[0], &o->dq[1],
                    &o->dq[2], &o->dq[3], &o->Tn, &o->Pn, &o->Qx, &o->Qy, &o->Qz, &o->ppre};
  for (size_t t = 0; t < sizeof(tmp) / sizeof(tmp[0]); t++) *tmp[t] = (double *)calloc(o->ncell, sizeof(double));
  if (!p->kicool || p->ifchem == 2 || p->ifrad == 2)
    for (int q = 0; q < 10; q++) o->col[q] = (double *)calloc(o->ncell, sizeof(double));
  o->nthr = nthreads_max();
  o->pw = (double *)calloc((size_t)o->nthr * 64 * (o->nmax + 8), sizeof(double));
  int nspl[3] = {p->nsplitx, p->nsplity, p->nsplitz};
  for (int d = 0; d < 3; d++) {
    o->nsp[d] = nspl[d] < 1 ? 1 : nspl[d];
    if (o->n[d] % o->nsp[d]) o->nsp[d] = 1;
  }
  o->ifevo = 1;
  setup_constants(o);
  if (p->ifgrv == 2) {
    const double *xs[3] = {o->gx[0], o->gx[1], o->gx[2]}, *ds[3] = {o->gdx[0], o->gdx[1], o->gdx[2]};
    o->phi = (double *)calloc(o->ncell, sizeof(double));
    o->grav = mcfg_create(o->n, o->nsp, p->ql1x + p->ql2x, xs, ds);
    if (!o->grav) {
      mcfo_destroy(o);
      return NULL;
    }
  }
  return o;
}

/* DTF (M3:402-431): the REAL*4 field read from DTF/D{rank}.dat over the
 * global interior (x fastest), x ghosts left at dinit1, y and z ghosts the
 * periodic images the reference's halo exchange gives them. */
int mcfo_set_dtf(mcfo *o, const float *interior) {
  if (!interior) { free(o->dtfa); o->dtfa = NULL; return 0; }
  if (!o->dtfa) o->dtfa = (float *)malloc(sizeof(float) * (size_t)o->ncell);
  if (!o->dtfa) return -1;
  const int nx = o->n[0], ny = o->n[1], nz = o->n[2];
  for (long c = 0; c < o->ncell; c++) o->dtfa[c] = (float)o->dtf;
  for (int k = -1; k <= nz + 2; k++)
    for (int j = -1; j <= ny + 2; j++) {
      const int jw = j < 1 ? j + ny : (j > ny ? j - ny : j);
      const int kw = k < 1 ? k + nz : (k > nz ? k - nz : k);
      for (int i = 1; i <= nx; i++)
        o->dtfa[ix3(o, i, j, k)] = interior[(size_t)(i - 1) + (size_t)nx * ((size_t)(jw - 1) + (size_t)ny * (size_t)(kw - 1))];
    }
  return 0;
}

int mcfo_set_threads(mcfo *o, int nthreads) {
  if (nthreads < 1) nthreads = nthreads_max();
  double *pw = (double *)calloc((size_t)nthreads * 64 * (o->nmax + 8), sizeof(double));
  if (!pw) return -1;
  free(o->pw);
  o->pw = pw;
  o->nthr = nthreads;
  return nthreads;
}

void mcfo_destroy(mcfo *o) {
  if (!o) return;
  for (int d = 0; d < 3; d++) { free(o->gx[d]); free(o->gdx[d]); }
  for (int k = 1; k <= MCFO_NFIELD; k++) free(o->f[k]);
  double *tmp[] = {o->blg[0], o->blg[1], o->vfc1, o->xlag, o->dxlagM, o->vfc[0], o->vfc[1], o->vfc[2],
                   o->dnc, o->emf[0], o->emf[1], o->emf[2], o->dq[0], o->dq[1], o->dq[2], o->dq[3],
                   o->Tn, o->Pn, o->Qx, o->Qy, o->Qz, o->ppre};
  for (size_t t = 0; t < sizeof(tmp) / sizeof(tmp[0]); t++) free(tmp[t]);
  free(o->pw);
  free(o->dtfa);
  for (int q = 0; q < 10; q++) free(o->col[q]);
  free(o->phi);
  mcfg_destroy(o->grav);
  free(o);
}

double *mcfo_field(mcfo *o, int id) { return (id >= 1 && id <= MCFO_NFIELD) ? o->f[id] : NULL; }
double *mcfo_grid(mcfo *o, int axis, int spacing) { return spacing ? o->gdx[axis] : o->gx[axis]; }
void mcfo_get_time(const mcfo *o, double *time, int *ifevo) { *time = o->time; *ifevo = o->ifevo; }
void mcfo_set_time(mcfo *o, double time, int ifevo) { o->time = time; o->ifevo = ifevo; }
void mcfo_bbrv(const mcfo *o, double *bbrv, double *bbrv_cm, double *dtf) {
  for (int k = 1; k <= 10; k++)
    for (int m = 1; m <= 2; m++)
      for (int s = 1; s <= 2; s++) bbrv[(k - 1) + 10 * (m - 1) + 20 * (s - 1)] = o->bbrv[k][m][s];
  for (int s = 0; s < 8; s++) bbrv_cm[s] = o->bbrv_cm[s];
  *dtf = o->dtf;
}

/* ------------------------------------------------------------------ */
/* Ghost fills (BC_MPI M3:800-926, BC_MPI_OT M3:930-1254), single block */
/* ------------------------------------------------------------------ */

/* Periodic copy of the ghost slabs of array a along axis d (width nol),
 * over the full transverse index range -1..n+2. */
static void fill_periodic(mcfo *o, double *a, int d, int nol, int left, int t1lo, int t1hi, int t2lo,
                          int t2hi) {
  int n = o->n[d];
  long sd = o->st[d], sa = o->st[d == 0 ? 1 : 0], sb = o->st[d == 2 ? 1 : 2];
  for (int b = t2lo; b <= t2hi; b++)
    for (int a1 = t1lo; a1 <= t1hi; a1++) {
      long base = (long)(a1 + 1) * sa + (long)(b + 1) * sb + sd; /* position 0 along d */
      if (left)
        for (int g = 1 - nol; g <= 0; g++) a[base + g * sd] = a[base + (g + n) * sd];
      else
        for (int g = n + 1; g <= n + nol; g++) a[base + g * sd] = a[base + (g - n) * sd];
    }
}

static void trange(const mcfo *o, int d, int *t1hi, int *t2hi) {
  if (d == 0) { *t1hi = o->n[1] + 2; *t2hi = o->n[2] + 2; }
  if (d == 1) { *t1hi = o->n[0] + 2; *t2hi = o->n[2] + 2; }
  if (d == 2) { *t1hi = o->n[0] + 2; *t2hi = o->n[1] + 2; }
}

static void periodic(mcfo *o, double *a, int d, int nol, int left) {
  int t1hi = 0, t2hi = 0;
  trange(o, d, &t1hi, &t2hi);
  fill_periodic(o, a, d, nol, left, -1, t1hi, -1, t2hi);
}

/* Inflow ghost density read from the DTF field shifted by -v t and wrapped
 * periodically in x (left M3:818-826, right M3:854-861).  The block edge
 * quantities are those of the rank that owns the boundary: dx(0) of the
 * first or last x block, L = dx(0)*Ncellx*NSPLTx. */
static double dtf_ghost(const mcfo *o, int side, int IX, int JY, int KZ) {
  const int cx = o->n[0] / o->nsp[0], nsx = o->nsp[0];
  double xp, dx0, L;
  int nn;
  if (side == 1) {
    dx0 = dxg(o, 0, 0);
    L = dx0 * cx * nsx;
    xp = xg(o, 0, 0) - 0.5 * dx0 + IX * dx0 - o->time * o->bbrv[2][1][1];
    nn = (int)(fabs(xp) / L) + 1;
    xp = xp + L * nn;
  } else {
    dx0 = dxg(o, 0, (nsx - 1) * cx);
    L = dx0 * cx * nsx;
    xp = 0.5 * (xg(o, 0, IX - 1) + xg(o, 0, IX)) - o->time * o->bbrv[2][1][2];
    nn = (int)(fabs(xp) / L);
    xp = xp - L * nn;
  }
  const int II = (int)(xp / dx0) + 1;
  return (double)o->dtfa[ix3(o, II, JY, KZ)];
}

/* BC_MPI x-side rule for U(:,:,:,K) (M3:818-884). side 1 = left, 2 = right. */
static void xrule_U(mcfo *o, int K, int side, int nol, int mode) {
  int nx = o->n[0];
  int lo = (side == 1) ? 1 - nol : nx + 1, hi = (side == 1) ? 0 : nx + nol;
  if (K == 6 && side == 2) lo = nx + 2;
  for (int k = -1; k <= o->n[2] + 2; k++)
    for (int j = -1; j <= o->n[1] + 2; j++)
      for (int i = lo; i <= hi; i++) {
        long c = ix3(o, i, j, k);
        if (K == 1) {
          U(1)[c] = o->dtfa ? dtf_ghost(o, side, i, j, k) : o->dtf;
        } else if (K == 6) {
          U(6)[c] = U(6)[ix3(o, side == 1 ? 1 : nx + 1, j, k)];
        } else if (K == 5 && mode == 2) {
          double v = o->bbrv[2][1][side];
          U(5)[c] = o->bbrv[5][1][side] * 1.5 +
                    0.5 * (U(1)[c] * (v * v) + BCC(1)[c] * BCC(1)[c] + o->blg[0][c] * o->blg[0][c] +
                           o->blg[1][c] * o->blg[1][c]);
        } else if (K == 2 && mode == 2) {
          U(2)[c] = U(1)[c] * o->bbrv[2][1][side];
        } else {
          U(K)[c] = o->bbrv[K][mode][side];
        }
      }
}

/* Outflow (free) x boundary, BC = 1: the zero-gradient rule of the 1D code
 * (1DMHDKIcool/1DMHDwCC3KIcool.f90:352-372), Q(0) = Q(-1) = Q(1) and
 * Q(N+1) = Q(N+2) = Q(N), for every cell-centred array; the face field
 * U(:,:,:,6) keeps the divergence-free copy of the inflow rule (left ghosts
 * take face 1, only N+2 takes face N+1: face N+1 is a boundary face CT
 * updates).  The 3D reference has no BC = 1; this is the build's
 * extension of it. */
static void outflow_copy(mcfo *o, double *a, int side, int nol, int face) {
  int nx = o->n[0];
  int lo = (side == 1) ? 1 - nol : nx + 1, hi = (side == 1) ? 0 : nx + nol;
  if (face && side == 2) lo = nx + 2;
  int src = (side == 1) ? 1 : (face ? nx + 1 : nx);
  for (int k = -1; k <= o->n[2] + 2; k++)
    for (int j = -1; j <= o->n[1] + 2; j++)
      for (int i = lo; i <= hi; i++) a[ix3(o, i, j, k)] = a[ix3(o, src, j, k)];
}

/* BC_MPI(N_ol,mode) for the listed U components along the axes in dmask. */
static void fillU(mcfo *o, int dmask, int nol, int mode, const int *list, int nk) {
  if (dmask & 1) {
    for (int t = 0; t < nk; t++) {
      if (o->p.bcx1 == 4) xrule_U(o, list[t], 1, nol, mode);
      else if (o->p.bcx1 == 1) outflow_copy(o, U(list[t]), 1, nol, list[t] == 6);
      else periodic(o, U(list[t]), 0, nol, 1);
    }
    for (int t = 0; t < nk; t++) {
      if (o->p.bcx2 == 4) xrule_U(o, list[t], 2, nol, mode);
      else if (o->p.bcx2 == 1) outflow_copy(o, U(list[t]), 2, nol, list[t] == 6);
      else periodic(o, U(list[t]), 0, nol, 0);
    }
  }
  for (int d = 1; d <= 2; d++)
    if (dmask & (1 << d)) {
      for (int t = 0; t < nk; t++) periodic(o, U(list[t]), d, nol, 1);
      for (int t = 0; t < nk; t++) periodic(o, U(list[t]), d, nol, 0);
    }
}

/* Generic BC_MPI_OT x-rule: value = constant, or copy of interior index src. */
static void xrule_const(mcfo *o, double *a, int side, int nol, double v) {
  int nx = o->n[0];
  int lo = (side == 1) ? 1 - nol : nx + 1, hi = (side == 1) ? 0 : nx + nol;
  for (int k = -1; k <= o->n[2] + 2; k++)
    for (int j = -1; j <= o->n[1] + 2; j++)
      for (int i = lo; i <= hi; i++) a[ix3(o, i, j, k)] = v;
}
static void xrule_copy(mcfo *o, double *a, int side, int nol, int src) {
  int nx = o->n[0];
  int lo = (side == 1) ? 1 - nol : nx + 1, hi = (side == 1) ? 0 : nx + nol;
  for (int k = -1; k <= o->n[2] + 2; k++)
    for (int j = -1; j <= o->n[1] + 2; j++)
      for (int i = lo; i <= hi; i++) a[ix3(o, i, j, k)] = a[ix3(o, src, j, k)];
}

/* BC_MPI_OT pieces.  what: 1 = dxlagM, 2 = Bcc, 3 = Blg, 4 = Vfc, 5 = species. */
static void fill_ot(mcfo *o, int dmask, int nol, int what) {
  int nx = o->n[0];
  if ((dmask & 1) && (o->p.bcx1 == 1 || o->p.bcx2 == 1)) {
    /* outflow: every BC_MPI_OT array takes the zero-gradient copy */
    double *arr[8];
    int na = 0;
    if (what == 1) arr[na++] = o->dxlagM;
    if (what == 2) for (int K = 1; K <= 3; K++) arr[na++] = BCC(K);
    if (what == 3) { arr[na++] = o->blg[0]; arr[na++] = o->blg[1]; }
    if (what == 4) { arr[na++] = o->vfc[1]; arr[na++] = o->vfc[2]; }
    if (what == 5) for (int sp = 0; sp < 8; sp++) arr[na++] = SPC(sp);
    for (int t = 0; t < na; t++) {
      if (o->p.bcx1 == 1) outflow_copy(o, arr[t], 1, nol, 0); else periodic(o, arr[t], 0, nol, 1);
      if (o->p.bcx2 == 1) outflow_copy(o, arr[t], 2, nol, 0); else periodic(o, arr[t], 0, nol, 0);
    }
    dmask &= ~1;
  }
  if (dmask & 1) {
    int b1 = (o->p.bcx1 == 4), b2 = (o->p.bcx2 == 4);
    if (what == 1) {
      if (b1) xrule_const(o, o->dxlagM, 1, nol, dxg(o, 0, 0)); else periodic(o, o->dxlagM, 0, nol, 1);
      if (b2) xrule_const(o, o->dxlagM, 2, nol, dxg(o, 0, nx + 1)); else periodic(o, o->dxlagM, 0, nol, 0);
    } else if (what == 2) {
      for (int K = 1; K <= 3; K++) {
        if (b1) { if (K == 1) xrule_copy(o, BCC(1), 1, nol, 1); else xrule_const(o, BCC(K), 1, nol, o->bbrv[5 + K][1][1]); }
        else periodic(o, BCC(K), 0, nol, 1);
        if (b2) { if (K == 1) xrule_copy(o, BCC(1), 2, nol, nx); else xrule_const(o, BCC(K), 2, nol, o->bbrv[5 + K][1][2]); }
        else periodic(o, BCC(K), 0, nol, 0);
      }
    } else if (what == 3) {
      for (int K = 1; K <= 2; K++) {
        if (b1) xrule_const(o, o->blg[K - 1], 1, nol, o->bbrv[6 + K][1][1]); else periodic(o, o->blg[K - 1], 0, nol, 1);
        if (b2) xrule_const(o, o->blg[K - 1], 2, nol, o->bbrv[6 + K][1][2]); else periodic(o, o->blg[K - 1], 0, nol, 0);
      }
    } else if (what == 4) {
      for (int K = 2; K <= 3; K++) {
        if (b1) xrule_copy(o, o->vfc[K - 1], 1, nol, 1); else periodic(o, o->vfc[K - 1], 0, nol, 1);
        if (b2) xrule_copy(o, o->vfc[K - 1], 2, nol, nx); else periodic(o, o->vfc[K - 1], 0, nol, 0);
      }
    } else if (what == 5) {
      for (int s = 0; s < 8; s++)
        if (!b1) periodic(o, SPC(s), 0, nol, 1);
      if (b1)
        for (int k = 1; k <= o->n[2]; k++)
          for (int j = 1; j <= o->n[1]; j++)
            for (int i = 1 - nol; i <= 0; i++) {
              long c = ix3(o, i, j, k);
              for (int s = 0; s < 8; s++) SPC(s)[c] = o->bbrv_cm[s] * U(1)[c];
            }
      for (int s = 0; s < 8; s++)
        if (!b2) periodic(o, SPC(s), 0, nol, 0);
      if (b2)
        for (int k = 1; k <= o->n[2]; k++)
          for (int j = 1; j <= o->n[1]; j++)
            for (int i = nx + 1; i <= nx + nol; i++) {
              long c = ix3(o, i, j, k);
              for (int s = 0; s < 8; s++) SPC(s)[c] = o->bbrv_cm[s] * U(1)[c];
            }
    }
  }
  for (int d = 1; d <= 2; d++) {
    if (!(dmask & (1 << d))) continue;
    if (what == 1) { periodic(o, o->dxlagM, d, nol, 1); periodic(o, o->dxlagM, d, nol, 0); }
    if (what == 2) for (int K = 1; K <= 3; K++) { periodic(o, BCC(K), d, nol, 1); periodic(o, BCC(K), d, nol, 0); }
    if (what == 3) for (int K = 0; K < 2; K++) { periodic(o, o->blg[K], d, nol, 1); periodic(o, o->blg[K], d, nol, 0); }
    if (what == 4) {
      int c1 = (d == 1) ? 0 : 0, c2 = (d == 1) ? 2 : 1; /* y: Vfc 1,3 ; z: Vfc 1,2 */
      periodic(o, o->vfc[c1], d, nol, 1); periodic(o, o->vfc[c1], d, nol, 0);
      periodic(o, o->vfc[c2], d, nol, 1); periodic(o, o->vfc[c2], d, nol, 0);
    }
    if (what == 5) {
      for (int s = 0; s < 8; s++) periodic(o, SPC(s), d, nol, 1);
      for (int s = 0; s < 8; s++) periodic(o, SPC(s), d, nol, 0);
    }
  }
}

/* ------------------------------------------------------------------ */
/* Point routines                                                      */
/* ------------------------------------------------------------------ */

/* RIEMAN (M3:1474-1551): 4 unrolled two-shock iterations. */
static void rieman(double uL1, double uL2, double uL3, double uL4, double uR1, double uR2, double uR3,
                   double uR4, double gammi1, double gammi2, double gammi3, double gampl1, double gampl3,
                   double *Pai, double *Vai) {
  double D1 = uL1, V1 = uL2, P1 = uL3, B1 = 0.5 * uL4;
  double D2 = uR1, V2 = uR2, P2 = uR3, B2 = 0.5 * uR4;
  P1 = P1 + B1;
  P2 = P2 + B2;
  double GP1 = gammi1 * P1, GP2 = gammi1 * P2;
  double GP1B1 = gammi3 * P1 - gammi2 * B1 * 2.0;
  double GP2B2 = gammi3 * P2 - gammi2 * B2 * 2.0;
  double Pa = (P1 + P2) * 0.5, qM1 = 0.0, qM2 = 0.0;
  for (int it = 0; it < 4; it++) {
    double pp1 = gampl3 * Pa + GP1B1;
    double qM1sq = pp1 + sqrt(pp1 * pp1 - 8.0 * (Pa - P1) * (gampl1 * Pa + GP1));
    qM1 = sqrt(qM1sq * D1) * 0.5;
    double pp2 = gampl3 * Pa + GP2B2;
    double qM2sq = pp2 + sqrt(pp2 * pp2 - 8.0 * (Pa - P2) * (gampl1 * Pa + GP2));
    qM2 = sqrt(qM2sq * D2) * 0.5;
    double PaOLD = Pa;
    Pa = qM2 * P1 + qM1 * P2 + qM2 * qM1 * (V1 - V2);
    Pa = Pa / (qM2 + qM1);
    if (Pa <= 0.0) Pa = 0.1 * PaOLD;
  }
  double Va = qM1 * V1 + qM2 * V2 + P1 - P2;
  Va = Va / (qM2 + qM1);
  *Pai = Pa;
  *Vai = Va;
}

/* CHAREQ (M3:2101-2111) */
static void chareq(double QL1, double QL2, double QL3, double QR1, double QR2, double QR3, double *Bay,
                   double *Vay) {
  double vyp = QL1, vym = QR1, Byp = QL2, Bym = QR2;
  double sqdyp = sqrt(QL3), sqdym = sqrt(QR3);
  *Vay = (sqdyp * vyp + sqdym * vym + Bym - Byp) / (sqdyp + sqdym);
  *Bay = (sqdyp * sqdym * (vym - vyp) + sqdyp * Bym + sqdym * Byp) / (sqdyp + sqdym);
}

/* KI Fcool (K1:1146-1173) with N = rho/1.27 (feedback variant :3586). */
static double fcool_ki(double rho, double T) {
  double N = rho / 1.27;
  double Laml = 1.0e7 * exp(-1.184e5 / (T * 1.0e3 + 1.0e3));
  Laml = DMIN(Laml, 5.0e3);
  double Lamc = 1.4e-2 * sqrt(T * 1.0e3) * exp(-9.2e1 / (T * 1.0e3));
  return (N * N * (Laml + Lamc) - N) * 3.94656e1;
}

/* IMC (M3:2898-2921): backward Euler by log-space bisection. */
static double imc(double P, double n, double dt, double rho, double gammi1) {
  double Pu = 1.0e10, Pd = 1.0e-10, Pm = 1.0e1, Pmold = -1.0;
  double nkbi = 1.0 / (KB * n);
  for (int kkk = 1; kkk <= 30; kkk++) {
    double T = Pm * nkbi;
    double CooL = fcool_ki(rho, T);
    double fev = Pm + CooL * dt * gammi1 - P;
    double iud = copysign(1.0, fev);
    Pd = DMAX(-iud * Pm, Pd);
    Pu = DMIN(2.0 * Pm / (iud + 1.0 + 1.0e-10), Pu);
    Pm = pow(1.0e1, 0.5 * (log10(Pu) + log10(Pd)));
    if (fabs(Pm - Pmold) < 1.0e-5) break;
    Pmold = Pm;
  }
  return Pm;
}

/* ------------------------------------------------------------------ */
/* Chemistry form of the cooling and the network (ifchem, ifrad = 2)    */
/* ------------------------------------------------------------------ */
/* Constants of module chmvar (M3:27-28). */
#define CH_G0 1.0
#define CH_XO 3.2e-4
#define CH_DV 2.0
#define CH_TGR 5.0e-3

#define COL(q, d) (o->col[2 * (q) + (d)])
enum { C_NTOT = 0, C_NH2, C_NNC, C_NCO, C_TCII };

/* fesc (M3:2883-2888): escape probability of an optical depth */
static double fesc(double tau) {
  return (0.5 + copysign(0.5, 0.1 - 1.0e-16 - tau)) * (0.5 - 0.585 * tau + 0.4563 * (tau * tau)) +
         (0.5 + copysign(0.5, 11.9025 - (tau - 3.55) * (tau - 3.55))) * (1.0 - exp(-2.34 * tau)) / (4.68 * tau + 1.0e-8) +
         (0.5 + copysign(0.5, tau - 7.0)) / (4.0 * tau * sqrt(log(DMAX(tau * 0.56419, 4.0))) + 1.0e-8);
}

/* LEVC2 (M3:2890-2895): CII two-level populations */
static void levc2(double T, double b21, double ndH, double ndH2, double nde, double ndCp, double *n1, double *n2) {
  double c21 = 8.854e-8 * nde / sqrt(T) + 9.399e-10 * pow(T, 0.07) * (ndH + ndH2);
  double c12 = 2.0 * c21 * exp(-0.092 / T);
  double A21 = 2.4e-6 * b21;
  *n2 = ndCp * c12 / (c12 + c21 + A21);
  *n1 = ndCp * (c21 + A21) / (c12 + c21 + A21);
}

static inline double nde_of(const mcfo *o, long c) { return SPC(SP)[c] + SPC(SHEP)[c] + SPC(SCP)[c]; }
static inline double ndtot_of(const mcfo *o, long c) {
  return SPC(SP)[c] + SPC(SH)[c] + 2.0 * SPC(SH2)[c] + SPC(SHE)[c] + SPC(SHEP)[c];
}

/* Fcool, chemistry form (M3:2825-2881), at cell c with the electron and
 * total number densities nde, ndtot the reference holds for it. */
static double fcool_chem(const mcfo *o, long c, double T, double nde, double ndtot) {
  const double ndH = SPC(SH)[c], ndp = SPC(SP)[c], ndH2 = SPC(SH2)[c], ndHe = SPC(SHE)[c], ndHep = SPC(SHEP)[c];
  const double ndCO = SPC(SCO)[c], ndCp = SPC(SCP)[c];
  double Av1 = 1.63542e-3 * COL(C_NTOT, 0)[c], x1 = 6.1714e3 * COL(C_NH2, 0)[c];
  double Av2 = 1.63542e-3 * COL(C_NTOT, 1)[c], x2 = 6.1714e3 * COL(C_NH2, 1)[c];
  double ATN1 = (exp(-2.5 * Av1) + exp(-2.5 * Av2)) * 0.5;
  double ATN2 = (exp(-3.77358 * Av1) + exp(-3.77358 * Av2)) * 0.5;
  double pha = CH_G0 * sqrt(1.0e3 * T) * ATN1 / nde;
  double Laml = ndH * nde * 1.44049e9 * exp(-1.184e2 / T);
  Laml = Laml + ndp * nde * 1.48e6 * exp(-8.0e1 / T);
  double fesC1 = fesc(COL(C_TCII, 0)[c]), fesC2 = fesc(COL(C_TCII, 1)[c]);
  double b21 = fesC1 + fesC2, n1, n2;
  levc2(T, b21, ndH, ndH2, nde, ndCp, &n1, &n2);
  double Lamc = 6.0157e7 * n2 * b21;
  double tO1 = 3.40057e-6 * COL(C_NTOT, 0)[c] / CH_XO, tO2 = 3.40057e-6 * COL(C_NTOT, 1)[c] / CH_XO;
  double fesO1 = fesc(tO1), fesO2 = fesc(tO2);
  double Lamo = (DMAX(ndtot * CH_XO - ndCO, 0.0) / CH_XO) * (ndH + 0.5 * ndH2) * 1.23916e1 * pow(T, 0.4) * exp(-0.228 / T);
  Lamo = Lamo * (fesO1 + fesO2);
  double Lamd = nde * ndtot * 6.06236 * pow(T, 0.94) * pow(pha, 0.462628 / pow(T, 6.8e-2));
  double ncr = 3.3e6 * pow(T, 0.75) / (ndH + ndp + ndH2 + ndHe + ndHep);
  double tau1 = 1.33194e1 * COL(C_NCO, 0)[c] / (T * CH_DV), tau2 = 1.33194e1 * COL(C_NCO, 1)[c] / (T * CH_DV);
  double ct1 = tau1 * sqrt(6.283 * log(2.13 + (tau1 * 0.36788) * (tau1 * 0.36788)));
  double ct2 = tau2 * sqrt(6.283 * log(2.13 + (tau2 * 0.36788) * (tau2 * 0.36788)));
  double ym1 = log(1.0 + ct1 / (1.0 + 1.0e1 * ncr)), ym2 = log(1.0 + ct2 / (1.0 + 1.0e1 * ncr));
  double fes = (2.0 + ym1 + 0.6 * (ym1 * ym1)) / (1.0 + ct1 + ncr + 1.5 * sqrt(ncr)) +
               (2.0 + ym2 + 0.6 * (ym2 * ym2)) / (1.0 + ct2 + ncr + 1.5 * sqrt(ncr));
  double LCOr = ndCO * 1.91505e10 * (T * T) * fes;
  double LCOH = ndH * ndCO * 7.96086e4 * sqrt(T) * exp(-pow(2.0 / T, 3.43)) * exp(-3.08 / T);
  double LCOH2 = ndH2 * ndCO * 3.60834e4 * T * exp(-pow(3.14e2 / T, 0.333)) * exp(-3.08 / T);
  double Gampe = ndtot * 2.56526e3 * CH_G0 * ATN1 *
                 (7.382e-3 * pow(T, 0.7) / (1.0 + 2.0e-4 * pha) + 4.9e-2 / (1.0 + 4.0e-3 * pow(pha, 0.73)));
  double Gamcr = (ndH + ndHe + ndH2) * 1.89435;
  double SHLD1 = 0.965 / ((1.0 + x1 / CH_DV) * (1.0 + x1 / CH_DV)) + 0.035 / sqrt(1.0 + x1) * exp(-8.5e-4 * sqrt(1.0 + x1));
  double SHLD2 = 0.965 / ((1.0 + x2 / CH_DV) * (1.0 + x2 / CH_DV)) + 0.035 / sqrt(1.0 + x2) * exp(-8.5e-4 * sqrt(1.0 + x2));
  double Gampd = ndH2 * 4.16362e4 * CH_G0 * ATN2 * (SHLD1 + SHLD2) * 0.5;
  return Laml + Lamc + Lamo + Lamd + LCOr + LCOH + LCOH2 - Gampe - Gamcr - Gampd;
}

/* The Fcool the run uses at cell c (the KI form reads only the density). */
static double fcool_cell(const mcfo *o, long c, double T) {
  if (o->p.kicool) return fcool_ki(U(1)[c], T);
  return fcool_chem(o, c, T, nde_of(o, c), ndtot_of(o, c));
}

/* IMC (M3:2898-2921) with the run's Fcool at cell c. */
static double imc_cell(const mcfo *o, long c, double P, double n, double dt, double gammi1) {
  double Pu = 1.0e10, Pd = 1.0e-10, Pm = 1.0e1, Pmold = -1.0;
  double nkbi = 1.0 / (KB * n);
  for (int kkk = 1; kkk <= 30; kkk++) {
    double T = Pm * nkbi;
    double CooL = fcool_cell(o, c, T);
    double fev = Pm + CooL * dt * gammi1 - P;
    double iud = copysign(1.0, fev);
    Pd = DMAX(-iud * Pm, Pd);
    Pu = DMIN(2.0 * Pm / (iud + 1.0 + 1.0e-10), Pu);
    Pm = pow(1.0e1, 0.5 * (log10(Pu) + log10(Pd)));
    if (fabs(Pm - Pmold) < 1.0e-5) break;
    Pmold = Pm;
  }
  return Pm;
}

/* SHIELD (M3:2924-3036): column densities of the total gas, H2, C, CO and
 * the CII optical depth from each x end, per (j,k) row.  Each x block of
 * the reproduced MPI run scans its own cells (half a cell's column at the
 * cell, dx of the block's first cell), then adds the totals of the blocks
 * beyond it in order of distance. */
static void shield(mcfo *o) {
  const int nx = o->n[0], ny = o->n[1], nz = o->n[2], nsx = o->nsp[0], cx = nx / nsx;
  const double dvin = 1.0 / CH_DV;
  double *temp = (double *)malloc(sizeof(double) * (size_t)o->ncell);
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long c = ix3(o, i, j, k);
        double b21 = fesc(COL(C_TCII, 0)[c]) + fesc(COL(C_TCII, 1)[c]);
        double T = U(5)[c] / (KB * (SPC(SP)[c] + SPC(SH)[c] + SPC(SH2)[c] + SPC(SHE)[c] + SPC(SHEP)[c]));
        double n1, n2;
        levc2(T, b21, SPC(SH)[c], SPC(SH2)[c], nde_of(o, c), SPC(SCP)[c], &n1, &n2);
        temp[c] = 1.15563e1 * (n1 - n2) * dvin;
      }
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      double tot[nsx][5];
      double *q[5] = {NULL, SPC(SH2), SPC(SC), SPC(SCO), temp};
      for (int b = 0; b < nsx; b++) {
        const double dxh = 0.5 * dxg(o, 0, b * cx + 1);
        double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int il = 1; il <= cx; il++) {
          long c = ix3(o, b * cx + il, j, k);
          for (int m = 0; m < 5; m++) {
            double v = m == 0 ? ndtot_of(o, c) : q[m][c];
            t[m] = t[m] + dxh * v;
            COL(m, 0)[c] = t[m];
            t[m] = t[m] + dxh * v;
          }
        }
        for (int m = 0; m < 5; m++) t[m] = 0.0;
        for (int il = cx; il >= 1; il--) {
          long c = ix3(o, b * cx + il, j, k);
          for (int m = 0; m < 5; m++) {
            double v = m == 0 ? ndtot_of(o, c) : q[m][c];
            t[m] = t[m] + dxh * v;
            COL(m, 1)[c] = t[m];
            t[m] = t[m] + dxh * v;
          }
        }
        for (int m = 0; m < 5; m++) tot[b][m] = t[m];
      }
      for (int b = 0; b < nsx; b++)
        for (int il = 1; il <= cx; il++) {
          long c = ix3(o, b * cx + il, j, k);
          for (int m = 0; m < 5; m++) {
            for (int L = 1; b + L < nsx; L++) COL(m, 1)[c] = COL(m, 1)[c] + tot[b + L][m];
            for (int L = 1; b - L >= 0; L++) COL(m, 0)[c] = COL(m, 0)[c] + tot[b - L][m];
          }
        }
    }
  free(temp);
}

void mcfo_init_columns(mcfo *o) {
  if (!o->col[0]) return;
  for (int q = 0; q < 10; q++) memset(o->col[q], 0, sizeof(double) * (size_t)o->ncell);
  if (o->p.ifrad == 2)
    for (int l = 0; l < 20; l++) shield(o);
}

double *mcfo_column(mcfo *o, int q, int dir) {
  if (q < 0 || q > 4 || dir < 0 || dir > 1 || !o->col[0]) return NULL;
  return COL(q, dir);
}

double mcfo_fcool_cell(mcfo *o, int i, int j, int k, double T) {
  long c = ix3(o, i, j, k);
  return fcool_chem(o, c, T, nde_of(o, c), ndtot_of(o, c));
}

/* Omexp (M3:3136-3141): 1 - exp(-eps), by its Taylor series below 1e-4 */
static double omexp(double eps) {
  return (0.5 + copysign(0.5, eps - 1.0e-4)) * (1.0 - exp(-eps)) +
         (0.5 - copysign(0.5, eps - 1.0e-4)) *
             (eps - 0.5 * (eps * eps) + 0.16666666666666667 * (eps * eps * eps) - 4.16666666666666667e-2 * (eps * eps * eps * eps));
}

typedef struct {
  double zeta, kHrec, kHerec, kH2, kH2ph, kH2dH, kH2de, kCO, kCOph, kCi, kCrec, kCOde, kCOdH, kHie, kHeie, kCie, kHiH,
      kHeiH, kCiH, kCOdHep, kH2dHep;
} chem_rates;

/* RATES (M3:3040-3133) at cell c and temperature T, with the cell's
 * electron density nde. */
static chem_rates rates(const mcfo *o, long c, double T, double nde) {
  chem_rates r;
  double Av1 = 1.63542e-3 * COL(C_NTOT, 0)[c], x1 = 6.1714e3 * COL(C_NH2, 0)[c];
  double Av2 = 1.63542e-3 * COL(C_NTOT, 1)[c], x2 = 6.1714e3 * COL(C_NH2, 1)[c];
  double ATN2 = (exp(-3.77358 * Av1) + exp(-3.77358 * Av2)) * 0.5;
  double ATN3 = (exp(-2.3585 * Av1) + exp(-2.3585 * Av2)) * 0.5;
  double NH21 = COL(C_NH2, 0)[c], NH22 = COL(C_NH2, 1)[c];
  double ATN4 = DMIN(1.0, pow(3.0857e1 * NH21, -3.0e-2) * exp(-1.54285e-3 * NH21)) +
                DMIN(1.0, pow(3.0857e1 * NH22, -3.0e-2) * exp(-1.54285e-3 * NH22));
  ATN4 = ATN4 * 0.5;
  const double p06 = (double)0.6f; /* the exponent is a default-real (REAL*4) literal */
  double NC1 = COL(C_NCO, 0)[c], NC2 = COL(C_NCO, 1)[c];
  double ATN5 = (0.5 - copysign(0.5, NC1 - 1.40995e-3)) * exp(-1.0 * pow(1.02857e3 * NC1, p06)) +
                (0.5 + copysign(0.5, NC1 - 1.40995e-3)) *
                    DMIN(pow(3.75439e3 * NC1 + 1.0e-100, -0.75), pow(7.3469e1 * NC1 + 1.0e-100, -1.3));
  ATN5 = ATN5 + (0.5 - copysign(0.5, NC2 - 1.40995e-3)) * exp(-1.0 * pow(1.02857e3 * NC2, p06)) +
         (0.5 + copysign(0.5, NC2 - 1.40995e-3)) *
             DMIN(pow(3.75439e3 * NC2 + 1.0e-100, -0.75), pow(7.3469e1 * NC2 + 1.0e-100, -1.3));
  ATN5 = ATN5 * 0.5;
  r.zeta = 9.4671e-4;
  r.kHrec = (0.5 + copysign(0.5, 15.78 - T)) * 0.45 * log(1.578e2 / T) +
            (0.5 - copysign(0.5, 15.78 - T)) * 0.4 * sqrt(1.578e2 / T);
  r.kHrec = 2.05572e1 * pow(T, -0.5) * r.kHrec;
  r.kHerec = 6.37623e1 * pow(T, -0.672);
  r.kH2 = 3.4569e-3 * sqrt(T) / (1.0 + 1.26491 * sqrt(T + CH_TGR) + 2.0 * T + 8.0 * (T * T));
  double SHLD1 = 0.965 / ((1.0 + x1 / CH_DV) * (1.0 + x1 / CH_DV)) + 0.035 / sqrt(1.0 + x1) * exp(-8.5e-4 * sqrt(1.0 + x1));
  double SHLD2 = 0.965 / ((1.0 + x2 / CH_DV) * (1.0 + x2 / CH_DV)) + 0.035 / sqrt(1.0 + x2) * exp(-8.5e-4 * sqrt(1.0 + x2));
  r.kH2ph = 1.04138e3 * CH_G0 * ATN2 * (SHLD1 + SHLD2) * 0.5;
  r.kH2dH = 1.07294e5 * exp(-4.39e1 / T);
  r.kH2de = 1.133e6 * pow(T, 0.35) * exp(-1.02e2 / T);
  r.kCO = 1.57785e-2 / (1.0 + CH_G0 * ATN3 / (SPC(SH2)[c] * CH_XO));
  r.kCOph = 3.155e3 * CH_G0 * ATN3 * ATN4 * ATN5;
  double SHLC1 = exp(-2.6 * Av1 - 3.0857e1 * COL(C_NNC, 0)[c] - 4.33743413e-4 * NH21) / (1.0 + 4.33743413e-4 * NH21);
  double SHLC2 = exp(-2.6 * Av2 - 3.0857e1 * COL(C_NNC, 1)[c] - 4.33743413e-4 * NH22) / (1.0 + 4.33743413e-4 * NH22);
  r.kCi = 3.97 * r.zeta + 6.62697e3 * CH_G0 * (SHLC1 + SHLC2) * 0.5;
  r.kCrec = 2.69267e1 * pow(T, -0.62) * (1.0 + 1.0e-4 * sqrt(nde) / (T * T));
  r.kCOde = 1.4e4 * pow(T, -0.5) * exp(-1.14e2 / T);
  r.kCOdH = 1.91858e3 * pow(T, -0.5) * exp(-7.77e1 / T);
  r.kHie = 5.76e4 * sqrt(T) * exp(-1.58e2 / T);
  r.kHeie = 2.76e4 * pow(T, 0.43) * exp(-2.85e2 / T);
  r.kCie = 1.74e5 * pow(T, 0.4) * exp(-1.31e2 / T);
  r.kHiH = 9.79 * sqrt(T) * exp(-1.58e2 / T);
  r.kHeiH = 4.57 * pow(T, 0.43) * exp(-2.85e2 / T);
  r.kCiH = 3.01e1 * pow(T, 0.4) * exp(-1.31e2 / T);
  r.kH2dHep = 1.1676 * exp(-0.035 / T);
  r.kCOdHep = 5.0491e4;
  return r;
}

/* The chemistry network of SOURCE (M3:2576-2774) at cell c over dt. */
static void chem_cell(mcfo *o, long c, double dt) {
  const double mH = 1.0, mHe = 4.0, mH2 = 2.0;
  double ndpold = SPC(SP)[c], ndHold = SPC(SH)[c], ndH2old = SPC(SH2)[c], ndHeold = SPC(SHE)[c];
  double ndHepold = SPC(SHEP)[c], ndCold = SPC(SC)[c], ndCpold = SPC(SCP)[c], ndCOold = SPC(SCO)[c];
  double ndH, ndp, ndH2, ndHe, ndHep, ndC, ndCp, ndCO;
  const double ndtot = ndpold + ndHold + 2.0 * ndH2old + ndHeold + ndHepold;
  double nde = ndpold + ndHepold + ndCpold;
  double T = U(5)[c] / KB / (ndpold + ndHold + ndH2old + ndHeold + ndHepold);
  chem_rates r = rates(o, c, T, nde);
  double temp1, temp2, temp3, omeps, eps;
  /* H recombination & ionization by CR */
  temp1 = r.kHrec * nde;
  temp2 = exp(-dt * (r.zeta + temp1));
  temp3 = 1.0 / (r.zeta + temp1);
  omeps = omexp(dt * (r.zeta + temp1));
  ndH = (temp1 * ndHold + r.zeta * ndHold * temp2 + temp1 * ndpold * omeps) * temp3;
  ndp = (r.zeta * ndpold + temp1 * ndpold * temp2 + r.zeta * ndHold * omeps) * temp3;
  ndHold = ndH; ndpold = ndp; nde = ndp + ndHepold + ndCpold;
  /* He recombination & ionization by CR */
  temp1 = r.kHerec * nde;
  temp2 = exp(-dt * (r.zeta + temp1));
  temp3 = 1.0 / (r.zeta + temp1);
  omeps = omexp(dt * (r.zeta + temp1));
  ndHe = (temp1 * ndHeold + r.zeta * ndHeold * temp2 + temp1 * ndHepold * omeps) * temp3;
  ndHep = (r.zeta * ndHepold + temp1 * ndHepold * temp2 + r.zeta * ndHeold * omeps) * temp3;
  ndHeold = ndHe; ndHepold = ndHep; nde = ndp + ndHep + ndCpold;
  /* H2 formation & dissociation by UV & electron collision */
  temp1 = r.kH2ph + r.kH2de * nde;
  temp2 = temp1 + 2.0 * r.kH2 * ndtot;
  temp3 = exp(-dt * temp2);
  omeps = omexp(dt * temp2);
  ndH = (ndHold + 2.0 * ndH2old * omeps) * temp1 + 2.0 * r.kH2 * ndHold * ndtot * temp3;
  ndH = ndH / temp2;
  ndH2 = (2.0 * ndH2old + ndHold * omeps) * r.kH2 * ndtot + temp1 * ndH2old * temp3;
  ndH2 = ndH2 / temp2;
  ndHold = ndH; ndH2old = ndH2;
  /* H2 dissociation by H collision */
  temp1 = ndHold + 2.0 * ndH2old;
  temp2 = exp(-dt * temp1 * r.kH2dH);
  temp3 = 1.0 / (ndHold + 2.0 * ndH2old * temp2);
  ndH = ndHold * temp1 * temp3;
  ndH2 = ndH2old * temp1 * temp2 * temp3;
  ndHold = ndH; ndH2old = ndH2;
  /* CO formation */
  temp1 = exp(-dt * r.kCO * ndtot);
  eps = dt * r.kCO * ndtot;
  omeps = omexp(eps);
  ndCO = ndCOold + omeps * ndCpold;
  ndCp = temp1 * ndCpold;
  ndCOold = ndCO; ndCpold = ndCp; nde = ndp + ndHep + ndCp;
  /* CO dissociation by UV & electron collision & H collision */
  temp1 = exp(-dt * (r.kCOph + r.kCOde * nde + r.kCOdH * ndHold));
  eps = dt * (r.kCOph + r.kCOde * nde + r.kCOdH * ndHold);
  omeps = omexp(eps);
  ndC = ndCold + omeps * ndCOold;
  ndCO = temp1 * ndCOold;
  ndCold = ndC; ndCOold = ndCO;
  /* C recombination & ionization by CR & UV */
  temp1 = r.kCrec * nde;
  temp2 = exp(-dt * (r.kCi + temp1));
  temp3 = 1.0 / (r.kCi + temp1);
  omeps = omexp(dt * (r.kCi + temp1));
  ndC = (temp1 * ndCold + r.kCi * ndCold * temp2 + temp1 * ndCpold * omeps) * temp3;
  ndCp = (r.kCi * ndCpold + temp1 * ndCpold * temp2 + r.kCi * ndCold * omeps) * temp3;
  ndCold = ndC; ndCpold = ndCp; nde = ndp + ndHep + ndCp;
  /* H, He, C ionization by e collision */
  ndH = ndHold * exp(-dt * r.kHie * nde);
  ndHe = ndHeold * exp(-dt * r.kHeie * nde);
  ndC = ndCold * exp(-dt * r.kCie * nde);
  omeps = omexp(dt * r.kHie * nde); ndp = omeps * ndHold + ndpold;
  omeps = omexp(dt * r.kHeie * nde); ndHep = omeps * ndHeold + ndHepold;
  omeps = omexp(dt * r.kCie * nde); ndCp = omeps * ndCold + ndCpold;
  ndHold = ndH; ndHeold = ndHe; ndCold = ndC;
  ndpold = ndp; ndHepold = ndHep; ndCpold = ndCp;
  nde = ndp + ndHep + ndCp;
  /* H, He, C ionization by H or H2 or p collision */
  temp1 = 1.0 / (1.0 + r.kHiH * dt * ndHold);
  ndH = ndHold * temp1;
  ndp = ndpold + temp1 * r.kHiH * dt * (ndHold * ndHold);
  temp1 = ndH + ndp + ndH2old;
  ndHe = ndHeold * exp(-dt * r.kHeiH * temp1);
  ndC = ndCold * exp(-dt * r.kCiH * temp1);
  omeps = omexp(dt * r.kHeiH * temp1); ndHep = omeps * ndHeold + ndHepold;
  omeps = omexp(dt * r.kCiH * temp1); ndCp = omeps * ndCold + ndCpold;
  ndHold = ndH; ndHeold = ndHe; ndCold = ndC;
  ndpold = ndp; ndHepold = ndHep; ndCpold = ndCp;
  nde = ndp + ndHep + ndCp;
  /* H2 dissociation by Hep recombination */
  temp1 = ndHepold - ndH2old;
  if (temp1 >= 0.0) {
    temp3 = exp(-dt * r.kH2dHep * temp1);
    temp2 = 1.0 / (ndHepold - ndH2old * temp3);
    ndHep = ndHepold * temp1 * temp2;
    ndH2 = ndH2old * temp1 * temp2 * temp3;
    temp3 = ndHepold * (1.0 - temp1 * temp2);
    ndp = ndpold + temp3;
    ndH = ndHold + temp3;
    ndHe = ndHeold + temp3;
  } else {
    temp3 = exp(dt * r.kH2dHep * temp1);
    temp2 = 1.0 / (ndHepold * temp3 - ndH2old);
    ndHep = ndHepold * temp1 * temp2 * temp3;
    ndH2 = ndH2old * temp1 * temp2;
    temp3 = ndHepold * (1.0 - temp1 * temp2 * temp3);
    ndp = ndpold + temp3;
    ndH = ndHold + temp3;
    ndHe = ndHeold + temp3;
  }
  ndHold = ndH; ndpold = ndp; ndH2old = ndH2;
  ndHeold = ndHe; ndHepold = ndHep;
  /* CO dissociation by Hep recombination */
  temp1 = ndHepold - ndCOold;
  if (temp1 >= 0.0) {
    temp3 = exp(-dt * r.kCOdHep * temp1);
    temp2 = 1.0 / (ndHepold - ndCOold * temp3);
    ndHep = ndHepold * temp1 * temp2;
    ndCO = ndCOold * temp1 * temp2 * temp3;
    temp3 = ndHepold * (1.0 - temp1 * temp2);
    ndCp = ndCpold + temp3;
    ndHe = ndHeold + temp3;
  } else {
    temp3 = exp(dt * r.kCOdHep * temp1);
    temp2 = 1.0 / (ndHepold * temp3 - ndCOold);
    ndHep = ndHepold * temp1 * temp2 * temp3;
    ndCO = ndCOold * temp1 * temp2;
    temp3 = ndHepold * (1.0 - temp1 * temp2 * temp3);
    ndCp = ndCpold + temp3;
    ndHe = ndHeold + temp3;
  }
  SPC(SH)[c] = DMAX(o->spmin[SH], ndH);
  SPC(SP)[c] = DMAX(o->spmin[SP], ndp);
  SPC(SH2)[c] = DMAX(o->spmin[SH2], ndH2);
  SPC(SHE)[c] = DMAX(o->spmin[SHE], ndHe);
  SPC(SHEP)[c] = DMAX(o->spmin[SHEP], ndHep);
  SPC(SC)[c] = DMAX(o->spmin[SC], ndC);
  SPC(SCP)[c] = DMAX(o->spmin[SCP], ndCp);
  SPC(SCO)[c] = DMAX(o->spmin[SCO], ndCO);
  U(1)[c] = mH * SPC(SP)[c] + mH * SPC(SH)[c] + mH2 * SPC(SH2)[c] + mHe * SPC(SHE)[c] + mHe * SPC(SHEP)[c];
}

/* VLIMIT / VLIMIT_OT body (M3:1692-1906) for one pencil.  q, U5, U1 are
 * base pointers with stride s addressing position 0; dxx points at index 0
 * of an array valid on -1..n+2; grd receives positions 0..n+1. */
static void vlimit(const double *q, const double *U5, const double *U1, long base, long s, int n,
                   const double *dxx, double *grd) {
  for (int i = 0; i <= n + 1; i++) {
    long c = base + i * s, cp = c + s, cm = c - s;
    double delp = q[cp] - q[c];
    double delm = q[c] - q[cm];
    double flmt = (2.0 * delp * delm + EPSL) / (delp * delp + delm * delm + EPSL);
    flmt = DMAX(0.0, flmt);
    double g = flmt * (q[cp] - q[cm]) / (dxx[i] + 0.5 * dxx[i - 1] + 0.5 * dxx[i + 1]);
    double T = 1.27 * U5[c] / (KB * U1[c]);
    delp = 2.0 * delp / (dxx[i] + dxx[i + 1]);
    delm = 2.0 * delm / (dxx[i] + dxx[i - 1]);
    double ad = fabs(delp), bd = fabs(delm);
    double gmm = (0.5 + copysign(0.5, delp * delm)) * copysign(1.0, delp) * DMIN(ad, bd);
    grd[i] = g * (0.5 - copysign(0.5, T - 3.0)) + gmm * (0.5 + copysign(0.5, T - 3.0));
  }
}

static inline double posfix(double Q, double Uc) {
  return (0.5 - copysign(0.5, -Q)) * Q + (0.5 + copysign(0.5, -Q)) * Uc;
}

static inline double gam_of(const mcfo *o, long c) {
  double s4 = SPC(SH)[c] + SPC(SP)[c] + SPC(SHE)[c] + SPC(SHEP)[c];
  double g = 3.0 * s4 + 5.0 * SPC(SH2)[c];
  return (5.0 * s4 + 7.0 * SPC(SH2)[c]) / g;
}
static inline double gammi1_of(const mcfo *o, long c) {
  double s4 = SPC(SH)[c] + SPC(SP)[c] + SPC(SHE)[c] + SPC(SHEP)[c];
  double g = 3.0 * s4 + 5.0 * SPC(SH2)[c];
  return (2.0 * s4 + 2.0 * SPC(SH2)[c]) / g;
}

/* ------------------------------------------------------------------ */
/* Sweep: MHD -> RHS -> MOC -> REMAP                                   */
/* ------------------------------------------------------------------ */

typedef struct {
  int d, n, ncm, ncl, VN, VT1, VT2, BN, BT1, BT2;
  long s;
} sweepdef;

static sweepdef sweep_def(const mcfo *o, int d) {
  sweepdef w;
  w.d = d;
  w.s = o->st[d];
  w.n = o->n[d];
  if (d == 0) { w.ncm = o->n[1]; w.ncl = o->n[2]; w.VN = 2; w.VT1 = 3; w.VT2 = 4; w.BN = 6; w.BT1 = 2; w.BT2 = 3; }
  if (d == 1) { w.ncm = o->n[2]; w.ncl = o->n[0]; w.VN = 3; w.VT1 = 4; w.VT2 = 2; w.BN = 7; w.BT1 = 3; w.BT2 = 1; }
  if (d == 2) { w.ncm = o->n[0]; w.ncl = o->n[1]; w.VN = 4; w.VT1 = 2; w.VT2 = 3; w.BN = 8; w.BT1 = 1; w.BT2 = 2; }
  return w;
}

/* index of position 0 of pencil (Lnum, Mnum) (M3:1368-1374 mapping) */
static long pencil_base(const mcfo *o, int d, int L, int M) {
  if (d == 0) return ix3(o, 0, M, L);
  if (d == 1) return ix3(o, L, 0, M);
  return ix3(o, M, L, 0);
}

/* pencil work arrays of the calling thread: slot t gives an array valid on
 * -1..nmax+2, returned pointing at index 0 */
static double *pwk(mcfo *o, int t) {
  return o->pw + ((size_t)thread_id() * 64 + (size_t)t) * (o->nmax + 8) + 2;
}

static void rhs(mcfo *o, int d, double dt) {
  sweepdef w = sweep_def(o, d);
  int list[3] = {1, w.VN, 5};
  fillU(o, 1 << d, 2, 1, list, 3);
  fill_ot(o, 1 << d, 2, 5);
  memcpy(o->ppre, U(5), sizeof(double) * (size_t)o->ncell);
  const double *dxe = o->gdx[d] + 1;
  double *U1 = U(1), *UVN = U(w.VN), *U5 = U(5), *Bt1 = BCC(w.BT1), *Bt2 = BCC(w.BT2);
  double dtv = dt, fac = o->p.facdep;
#pragma omp parallel num_threads(o->nthr)
  {
  double *g1 = pwk(o, 0), *gVN = pwk(o, 1), *g5 = pwk(o, 2), *gB1 = pwk(o, 3), *gB2 = pwk(o, 4);
  double *Pa = pwk(o, 5), *Va = pwk(o, 6), *dm = pwk(o, 7);
#pragma omp for collapse(2) schedule(static)
  for (int L = 1; L <= w.ncl; L++)
    for (int M = 1; M <= w.ncm; M++) {
      long b = pencil_base(o, d, L, M), s = w.s;
      vlimit(U1, U5, U1, b, s, w.n, dxe, g1);
      vlimit(UVN, U5, U1, b, s, w.n, dxe, gVN);
      vlimit(U5, U5, U1, b, s, w.n, dxe, g5);
      vlimit(Bt1, U5, U1, b, s, w.n, dxe, gB1);
      vlimit(Bt2, U5, U1, b, s, w.n, dxe, gB2);
      for (int i = 0; i <= w.n; i++) {
        long c = b + i * s, cp = c + s;
        double gamma = gam_of(o, c);
        double cm = sqrt((gamma * U5[c] + Bt1[c] * Bt1[c] + Bt2[c] * Bt2[c]) / U1[c]);
        double depend1 = 0.5 * (dxe[i] - cm * dtv) * fac;
        depend1 = DMAX(0.0, depend1);
        gamma = gam_of(o, cp);
        cm = sqrt((gamma * U5[cp] + Bt1[cp] * Bt1[cp] + Bt2[cp] * Bt2[cp]) / U1[cp]);
        double depend2 = 0.5 * (dxe[i + 1] - cm * dtv) * fac;
        depend2 = DMAX(0.0, depend2);
        double QL1 = posfix(U1[c] + depend1 * g1[i], U1[c]);
        double QR1 = posfix(U1[cp] - depend2 * g1[i + 1], U1[cp]);
        double QL2 = UVN[c] + depend1 * gVN[i];
        double QR2 = UVN[cp] - depend2 * gVN[i + 1];
        double QL3 = posfix(U5[c] + depend1 * g5[i], U5[c]);
        double QR3 = posfix(U5[cp] - depend2 * g5[i + 1], U5[cp]);
        double a1 = Bt1[c] + depend1 * gB1[i], a2 = Bt2[c] + depend1 * gB2[i];
        double QL4 = a1 * a1 + a2 * a2;
        double b1 = Bt1[cp] - depend2 * gB1[i + 1], b2 = Bt2[cp] - depend2 * gB2[i + 1];
        double QR4 = b1 * b1 + b2 * b2;
        double ndHm = 0.5 * (SPC(SH)[c] + SPC(SH)[cp]);
        double ndpm = 0.5 * (SPC(SP)[c] + SPC(SP)[cp]);
        double ndHem = 0.5 * (SPC(SHE)[c] + SPC(SHE)[cp]);
        double ndHepm = 0.5 * (SPC(SHEP)[c] + SPC(SHEP)[cp]);
        double ndH2m = 0.5 * (SPC(SH2)[c] + SPC(SH2)[cp]);
        double gm = (5.0 * (ndHm + ndpm + ndHem + ndHepm) + 7.0 * ndH2m) /
                    (3.0 * (ndHm + ndpm + ndHem + ndHepm) + 5.0 * ndH2m);
        rieman(QL1, QL2, QL3, QL4, QR1, QR2, QR3, QR4, gm - 1.0, gm - 2.0, gm - 3.0, gm + 1.0, gm + 3.0,
               &Pa[i], &Va[i]);
      }
      for (int i = 0; i <= w.n; i++) {
        long c = b + i * s;
        dm[i] = dxe[i] * U1[c];
        o->xlag[c] = o->xlag[c] + Va[i] * dtv;
      }
      for (int i = 1; i <= w.n; i++) {
        long c = b + i * s, cmn = c - s;
        double gi = gammi1_of(o, c);
        U5[c] = U5[c] / (U1[c] * gi) + 0.5 * (U(2)[c] * U(2)[c] + U(3)[c] * U(3)[c] + U(4)[c] * U(4)[c]) +
                0.5 * (BCC(1)[c] * BCC(1)[c] + BCC(2)[c] * BCC(2)[c] + BCC(3)[c] * BCC(3)[c]) / U1[c];
        U1[c] = dm[i] / (o->xlag[c] - o->xlag[cmn]);
        UVN[c] = UVN[c] - dtv / dm[i] * (Pa[i] - Pa[i - 1]);
        U5[c] = U5[c] - dtv / dm[i] * (Pa[i] * Va[i] - Pa[i - 1] * Va[i - 1]);
        o->blg[0][c] = Bt1[c] * dxe[i] / (o->xlag[c] - o->xlag[cmn]);
        o->blg[1][c] = Bt2[c] * dxe[i] / (o->xlag[c] - o->xlag[cmn]);
        for (int sp = 0; sp < 8; sp++) SPC(sp)[c] = dxe[i] * SPC(sp)[c] / (o->xlag[c] - o->xlag[cmn]);
      }
      for (int i = 1; i <= w.n + 1; i++) o->vfc1[b + i * s] = Va[i - 1];
    }
  }
}

/* One MOC interface (M3:1997-2056) from the slopes at positions i and i+1
 * (5 quantities each: rho, vt1, vt2, Blg1, Blg2). */
static void moc_iface(const mcfo *o, const sweepdef *w, long b, int i, const double *dxl, const double *gi,
                      const double *gp, double dt, double *Bay, double *Vay, double *Baz, double *Vaz) {
  long s = w->s, c = b + i * s, cp = c + s, cpp = cp + s;
  const double *xe = o->gx[w->d] + 1;
  const double *U1 = o->f[1], *UT1 = o->f[w->VT1], *UT2 = o->f[w->VT2], *UBn = o->f[w->BN];
  const double *Bl1 = o->blg[0], *Bl2 = o->blg[1];
  double fac = o->p.facdep;
  double Bnap = UBn[cp] + (o->xlag[c] - xe[i]) * (UBn[cpp] - UBn[c]) / (xe[i + 1] - xe[i - 1]);
  double deni = sqrt((dxl[i] + dxl[i + 1]) / (dxl[i] * U1[cp] + dxl[i + 1] * U1[c]));
  double depend1, depend2, ca;
  double QL11, QL12, QL13, QR11, QR12, QR13, QL21, QL22, QR21, QR22;
  if (Bnap > 0.0) {
    ca = fabs(Bnap * deni) * 0.5 * dt;
    ca = DMIN(dxl[i], ca);
    depend1 = (0.5 * dxl[i] - ca) * fac;
    ca = fabs(Bnap * deni) * 0.5 * dt;
    ca = DMIN(dxl[i + 1], ca);
    depend2 = (0.5 * dxl[i + 1] - ca) * fac;
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
    ca = DMIN(dxl[i + 1], ca);
    depend1 = (0.5 * dxl[i + 1] - ca) * fac;
    ca = fabs(Bnap * deni) * 0.5 * dt;
    ca = DMIN(dxl[i], ca);
    depend2 = (0.5 * dxl[i] - ca) * fac;
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
  chareq(QL11, QL12, QL13, QR11, QR12, QR13, Bay, Vay);
  chareq(QL21, QL22, QL13, QR21, QR22, QR13, Baz, Vaz);
}

/* MOC (M3:1939-2098).  The reference evaluates this per MPI block: an
 * interface on a block edge is computed by both neighbouring blocks, each
 * reading the limiter switch T of the other block's cell from its ghost
 * copy of U(:,:,:,5) -- the pre-sweep pressure -- instead of the specific
 * energy the interior holds.  Both views are kept: the left view (owner =
 * the cell on the left) feeds that cell's update, the right view the other. */
static void moc(mcfo *o, int d, double dt) {
  sweepdef w = sweep_def(o, d);
  int list[4] = {1, w.VT1, w.VT2, w.BN};
  fillU(o, 1 << d, 2, 1, list, 4);
#pragma omp parallel for collapse(2) schedule(static) num_threads(o->nthr)
  for (int L = 1; L <= w.ncl; L++)
    for (int M = 1; M <= w.ncm; M++) {
      long b = pencil_base(o, d, L, M), s = w.s;
      for (int i = 1; i <= w.n; i++) o->dxlagM[b + i * s] = o->xlag[b + i * s] - o->xlag[b + (i - 1) * s];
    }
  fill_ot(o, 1 << d, 2, 1);
  fill_ot(o, 1 << d, 2, 3);
  const double *xe = o->gx[d] + 1;
  int nc = w.n / o->nsp[d];
  double *U1 = U(1), *U5 = U(5), *UT1 = U(w.VT1), *UT2 = U(w.VT2), *UVN = U(w.VN), *UBn = U(w.BN);
  double *Q[5] = {U1, UT1, UT2, o->blg[0], o->blg[1]};
  double *Bl1 = o->blg[0], *Bl2 = o->blg[1];
#pragma omp parallel num_threads(o->nthr)
  {
  double *dxl = pwk(o, 0);
  double *g[5], *h[5];
  for (int q = 0; q < 5; q++) { g[q] = pwk(o, 1 + q); h[q] = pwk(o, 6 + q); }
  double *VayL = pwk(o, 11), *BayL = pwk(o, 12), *VazL = pwk(o, 13), *BazL = pwk(o, 14);
  double *VayR = pwk(o, 15), *BayR = pwk(o, 16), *VazR = pwk(o, 17), *BazR = pwk(o, 18);
#pragma omp for collapse(2) schedule(static)
  for (int L = 1; L <= w.ncl; L++)
    for (int M = 1; M <= w.ncm; M++) {
      long b = pencil_base(o, d, L, M), s = w.s;
      for (int i = -1; i <= w.n + 2; i++) dxl[i] = o->dxlagM[b + i * s];
      for (int q = 0; q < 5; q++) {
        vlimit(Q[q], U5, U1, b, s, w.n, dxl, g[q]);
        vlimit(Q[q], o->ppre, U1, b, s, w.n, dxl, h[q]);
      }
      for (int i = 0; i <= w.n; i++) {
        double gi[5], gp[5];
        int edge = (i % nc) == 0;
        for (int q = 0; q < 5; q++) { gi[q] = g[q][i]; gp[q] = edge ? h[q][i + 1] : g[q][i + 1]; }
        moc_iface(o, &w, b, i, dxl, gi, gp, dt, &BayL[i], &VayL[i], &BazL[i], &VazL[i]);
        if (edge) {
          for (int q = 0; q < 5; q++) { gi[q] = h[q][i]; gp[q] = g[q][i + 1]; }
          moc_iface(o, &w, b, i, dxl, gi, gp, dt, &BayR[i], &VayR[i], &BazR[i], &VazR[i]);
        } else {
          BayR[i] = BayL[i]; VayR[i] = VayL[i]; BazR[i] = BazL[i]; VazR[i] = VazL[i];
        }
      }
      for (int i = 1; i <= w.n; i++) {
        long c = b + i * s, cm = c - s, cp = c + s, cpp = cp + s;
        double Bnap = UBn[cp] + (o->xlag[c] - xe[i]) * (UBn[cpp] - UBn[c]) / (xe[i + 1] - xe[i - 1]);
        double Bnam = UBn[c] + (o->xlag[cm] - xe[i - 1]) * (UBn[cp] - UBn[cm]) / (xe[i] - xe[i - 2]);
        double Vnap = o->vfc1[cp], Vnam = o->vfc1[c];
        UVN[c] = UVN[c] * U1[c] + dt / dxl[i] * 0.5 * (Bnap * Bnap - Bnam * Bnam);
        UT1[c] = UT1[c] * U1[c] + dt / dxl[i] * (Bnap * BayL[i] - Bnam * BayR[i - 1]);
        UT2[c] = UT2[c] * U1[c] + dt / dxl[i] * (Bnap * BazL[i] - Bnam * BazR[i - 1]);
        U5[c] = U5[c] * U1[c] + dt / dxl[i] *
                                    ((0.5 * Bnap * Vnap + BayL[i] * VayL[i] + BazL[i] * VazL[i]) * Bnap -
                                     (0.5 * Bnam * Vnam + BayR[i - 1] * VayR[i - 1] + BazR[i - 1] * VazR[i - 1]) * Bnam);
        Bl1[c] = Bl1[c] + dt / dxl[i] * (Bnap * VayL[i] - Bnam * VayR[i - 1]);
        Bl2[c] = Bl2[c] + dt / dxl[i] * (Bnap * VazL[i] - Bnam * VazR[i - 1]);
      }
    }
  }
}

static void remap(mcfo *o, int d, double dt) {
  sweepdef w = sweep_def(o, d);
  fill_ot(o, 1 << d, 2, 3);
  int list[5] = {1, 2, 3, 4, 5};
  fillU(o, 1 << d, 2, 2, list, 5);
  fill_ot(o, 1 << d, 2, 5);
  const double *xe = o->gx[d] + 1;
  double *Q[15];
  for (int q = 0; q < 5; q++) Q[q] = U(q + 1);
  Q[5] = o->blg[0];
  Q[6] = o->blg[1];
  for (int sp = 0; sp < 8; sp++) Q[7 + sp] = SPC(sp);
  double fac = o->p.facdep;
#pragma omp parallel num_threads(o->nthr)
  {
  double *dxl = pwk(o, 0);
  double *grd[15], *F[15];
  for (int q = 0; q < 15; q++) { grd[q] = pwk(o, 1 + q); F[q] = pwk(o, 16 + q); }
#pragma omp for collapse(2) schedule(static)
  for (int L = 1; L <= w.ncl; L++)
    for (int M = 1; M <= w.ncm; M++) {
      long b = pencil_base(o, d, L, M), s = w.s;
      for (int i = -1; i <= w.n + 2; i++) dxl[i] = o->dxlagM[b + i * s];
      for (int q = 0; q < 15; q++) vlimit(Q[q], U(5), U(1), b, s, w.n, dxl, grd[q]);
      for (int i = 0; i <= w.n; i++) {
        long c = b + i * s, cp = c + s;
        double vadt = o->vfc1[cp] * dt;
        if (vadt <= 0.0) {
          double depend = 0.5 * (dxl[i + 1] + vadt) * fac;
          for (int q = 0; q < 15; q++) F[q][i] = -(vadt * (Q[q][cp] - depend * grd[q][i + 1]));
        } else {
          double depend = 0.5 * (dxl[i] - vadt) * fac;
          for (int q = 0; q < 15; q++) F[q][i] = -(vadt * (Q[q][c] + depend * grd[q][i]));
        }
      }
      for (int i = 1; i <= w.n; i++) {
        long c = b + i * s;
        for (int q = 0; q < 15; q++) Q[q][c] = (Q[q][c] * dxl[i] + F[q][i] - F[q][i - 1]) / (xe[i] - xe[i - 1]);
      }
    }
  }
}

static void mhd_sweep(mcfo *o, int d, double dt) {
  /* MHD (M3:1264-1324) */
  int n = o->n[d];
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = -1; k <= o->n[2] + 2; k++)
    for (int j = -1; j <= o->n[1] + 2; j++)
      for (int i = -1; i <= o->n[0] + 2; i++) {
        int id = (d == 0) ? i : (d == 1) ? j : k;
        o->xlag[ix3(o, i, j, k)] = xg(o, d, id);
      }
  (void)n;
  rhs(o, d, dt);
  moc(o, d, dt);
  remap(o, d, dt);
  int iBcc = d + 1;
  double rfloor = 1.27 * o->rmin;
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= o->n[2]; k++)
    for (int j = 1; j <= o->n[1]; j++)
      for (int i = 1; i <= o->n[0]; i++) {
        long c = ix3(o, i, j, k);
        U(1)[c] = DMAX(rfloor, U(1)[c]);
        double invd = 1.0 / U(1)[c];
        U(2)[c] = U(2)[c] * invd;
        U(3)[c] = U(3)[c] * invd;
        U(4)[c] = U(4)[c] * invd;
        double gi = gammi1_of(o, c);
        double ekin = 0.5 * (U(2)[c] * U(2)[c] + U(3)[c] * U(3)[c] + U(4)[c] * U(4)[c]) * U(1)[c];
        double emag = 0.5 * (BCC(iBcc)[c] * BCC(iBcc)[c] + o->blg[0][c] * o->blg[0][c] + o->blg[1][c] * o->blg[1][c]);
        U(5)[c] = (U(5)[c] - ekin - emag) * gi;
        U(5)[c] = DMAX(o->pmin, U(5)[c]);
        U(5)[c] = DMIN(o->pmax, U(5)[c]);
        for (int sp = 0; sp < 8; sp++) SPC(sp)[c] = DMAX(o->spmin[sp], SPC(sp)[c]);
      }
}

void mcfo_sweep(mcfo *o, int dir, double dt) { mhd_sweep(o, dir, dt); }

/* ------------------------------------------------------------------ */
/* Constrained transport: CC(1), CC(2)+CCT x3, CC(3), CC(4)            */
/* ------------------------------------------------------------------ */

static void cc1(mcfo *o) {
  int lx[3] = {2, 7, 8}, ly[3] = {3, 6, 8}, lz[3] = {4, 6, 7};
  fillU(o, 1, 2, 1, lx, 3);
  fillU(o, 2, 2, 1, ly, 3);
  fillU(o, 4, 2, 1, lz, 3);
  int nx = o->n[0], ny = o->n[1], nz = o->n[2];
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx + 1; i++)
        o->vfc[0][ix3(o, i, j, k)] = (dxg(o, 0, i - 1) * U(2)[ix3(o, i, j, k)] + dxg(o, 0, i) * U(2)[ix3(o, i - 1, j, k)]) /
                                     (dxg(o, 0, i - 1) + dxg(o, 0, i));
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny + 1; j++)
      for (int i = 1; i <= nx; i++)
        o->vfc[1][ix3(o, i, j, k)] = (dxg(o, 1, j - 1) * U(3)[ix3(o, i, j, k)] + dxg(o, 1, j) * U(3)[ix3(o, i, j - 1, k)]) /
                                     (dxg(o, 1, j - 1) + dxg(o, 1, j));
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= nz + 1; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        o->vfc[2][ix3(o, i, j, k)] = (dxg(o, 2, k - 1) * U(4)[ix3(o, i, j, k)] + dxg(o, 2, k) * U(4)[ix3(o, i, j, k - 1)]) /
                                     (dxg(o, 2, k - 1) + dxg(o, 2, k));
  fill_ot(o, 7, 2, 4);
  int l5[1] = {5};
  fillU(o, 7, 2, 1, l5, 1);
}

/* local (i along d1, j along d2, k along d3) -> global index, for the
 * three edge orientations Ez (x,y,z), Ex (y,z,x), Ey (z,x,y) */
static inline long lix(const mcfo *o, int orient, int i, int j, int k) {
  if (orient == 0) return ix3(o, i, j, k);
  if (orient == 1) return ix3(o, k, i, j);
  return ix3(o, j, k, i);
}

/* one VLIMIT pencil along axis ax at arbitrary transverse position,
 * base = global index of position 0 */
static void vlim_axis(mcfo *o, const double *q, int ax, long base, double *grd) {
  vlimit(q, U(5), U(1), base, o->st[ax], o->n[ax], o->gdx[ax] + 1, grd);
}

static void cct(mcfo *o, int orient, double dt) {
  int d1 = orient == 0 ? 0 : orient == 1 ? 1 : 2;
  int d2 = (d1 + 1) % 3, d3 = (d1 + 2) % 3;
  int n1 = o->n[d1], n2 = o->n[d2], n3 = o->n[d3];
  double *Vx = o->vfc[d1], *Vy = o->vfc[d2], *Bx = U(6 + d1), *By = U(6 + d2);
  double *dQ1 = o->dq[0], *dQ2 = o->dq[1], *dQ3 = o->dq[2], *dQ4 = o->dq[3];
  double facm = o->p.facdep * 0.5;
#pragma omp parallel num_threads(o->nthr)
  {
  double *grd = pwk(o, 0), *grd2 = pwk(o, 1);
  /* slopes along d1 of Vfc_d2 and B_d2 (dQ 2, 4), at d2 = 1..n2+1, d3 = 1..n3 */
#pragma omp for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2 + 1; j++) {
      long base = lix(o, orient, 0, j, k);
      vlim_axis(o, Vy, d1, base, grd);
      vlim_axis(o, By, d1, base, grd2);
      for (int i = 0; i <= n1 + 1; i++) {
        long c = lix(o, orient, i, j, k);
        dQ2[c] = grd[i];
        dQ4[c] = grd2[i];
      }
    }
  /* slopes along d2 of Vfc_d1 and B_d1 (dQ 1, 3), at d1 = 1..n1+1, d3 = 1..n3 */
#pragma omp for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int i = 1; i <= n1 + 1; i++) {
      long base = lix(o, orient, i, 0, k);
      vlim_axis(o, Vx, d2, base, grd);
      vlim_axis(o, Bx, d2, base, grd2);
      for (int j = 0; j <= n2 + 1; j++) {
        long c = lix(o, orient, i, j, k);
        dQ1[c] = grd[j];
        dQ3[c] = grd2[j];
      }
    }
  }
  const double *dxe = o->gdx[d1] + 1, *dye = o->gdx[d2] + 1;
  double *emf = o->emf[(d1 + 2) % 3]; /* Ez for orient 0, Ex for 1, Ey for 2 */
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2 + 1; j++)
      for (int i = 1; i <= n1 + 1; i++) {
        long c0 = lix(o, orient, i, j, k);
        double dsqinv = 1.0 / sqrt(o->dnc[c0]);
        double C1p[4], C1m[4], C2p[4], C2m[4];
        for (int cs = 0; cs < 4; cs++) {
          /* cs: 0 LL (S1=1,S2=1), 1 RL (-1,1), 2 LR (1,-1), 3 RR (-1,-1) */
          double S1 = (cs == 0 || cs == 2) ? 1.0 : -1.0;
          double S2 = (cs == 0 || cs == 1) ? 1.0 : -1.0;
          long p1 = (cs == 0 || cs == 2) ? lix(o, orient, i, j - 1, k) : c0;
          long p2 = (cs == 0 || cs == 1) ? lix(o, orient, i - 1, j, k) : c0;
          double dyv = (cs == 0 || cs == 2) ? dye[j - 1] : dye[j];
          double dxv = (cs == 0 || cs == 1) ? dxe[i - 1] : dxe[i];
          double a1p = Vx[p1] + Bx[p1] * dsqinv;
          double da1p = (dQ1[p1] + dQ3[p1] * dsqinv) * facm;
          double a1m = Vx[p1] - Bx[p1] * dsqinv;
          double da1m = (dQ1[p1] - dQ3[p1] * dsqinv) * facm;
          double a2m = Vy[p2] - By[p2] * dsqinv;
          double da2m = (dQ2[p2] - dQ4[p2] * dsqinv) * facm;
          double a2p = Vy[p2] + By[p2] * dsqinv;
          double da2p = (dQ2[p2] + dQ4[p2] * dsqinv) * facm;
          double t;
          t = a1p + da1p * dyv * S1 - da1p * (a2m + da2m * dxv * S2) * dt;
          C1p[cs] = t / (1.0 - da1p * da2m * dt * dt);
          t = a2m + da2m * dxv * S2 - da2m * (a1p + da1p * dyv * S1) * dt;
          C2m[cs] = t / (1.0 - da1p * da2m * dt * dt);
          t = a1m + da1m * dyv * S1 - da1m * (a2p + da2p * dxv * S2) * dt;
          C1m[cs] = t / (1.0 - da1m * da2p * dt * dt);
          t = a2p + da2p * dxv * S2 - da2p * (a1m + da1m * dyv * S1) * dt;
          C2p[cs] = t / (1.0 - da1m * da2p * dt * dt);
        }
        int ifW[4][2];
        ifW[0][0] = (C1p[0] > 0.0) && (C2m[0] > 0.0);
        ifW[0][1] = (C1m[0] > 0.0) && (C2p[0] > 0.0);
        ifW[1][0] = (C1p[1] > 0.0) && (C2m[1] < 0.0);
        ifW[1][1] = (C1m[1] > 0.0) && (C2p[1] < 0.0);
        ifW[2][0] = (C1p[2] < 0.0) && (C2m[2] > 0.0);
        ifW[2][1] = (C1m[2] < 0.0) && (C2p[2] > 0.0);
        ifW[3][0] = (C1p[3] < 0.0) && (C2m[3] < 0.0);
        ifW[3][1] = (C1m[3] < 0.0) && (C2p[3] < 0.0);
        double W[4][2];
        for (int n = 0; n < 4; n++) {
          int both = (int)(0.51 * (double)ifW[n][0] + 0.51 * (double)ifW[n][1]);
          W[n][0] = (double)ifW[n][0] + (1.0e13 - 1.0) * (double)both;
          W[n][1] = (double)ifW[n][1] + (1.0e13 - 1.0) * (double)both;
        }
        double c1p = 0.0, c2m = 0.0, c1m = 0.0, c2p = 0.0;
        if (ifW[0][0] + ifW[1][0] + ifW[2][0] + ifW[3][0] != 0) {
          c1p = W[0][0] * fabs(C2m[0]) * C1p[0] + W[1][0] * fabs(C2m[1]) * C1p[1] + W[2][0] * fabs(C2m[2]) * C1p[2] +
                W[3][0] * fabs(C2m[3]) * C1p[3];
          c1p = c1p / (W[0][0] * fabs(C2m[0]) + W[1][0] * fabs(C2m[1]) + W[2][0] * fabs(C2m[2]) + W[3][0] * fabs(C2m[3]));
          c2m = W[0][0] * fabs(C1p[0]) * C2m[0] + W[1][0] * fabs(C1p[1]) * C2m[1] + W[2][0] * fabs(C1p[2]) * C2m[2] +
                W[3][0] * fabs(C1p[3]) * C2m[3];
          c2m = c2m / (W[0][0] * fabs(C1p[0]) + W[1][0] * fabs(C1p[1]) + W[2][0] * fabs(C1p[2]) + W[3][0] * fabs(C1p[3]));
        }
        if (ifW[0][1] + ifW[1][1] + ifW[2][1] + ifW[3][1] != 0) {
          c1m = W[0][1] * fabs(C2p[0]) * C1m[0] + W[1][1] * fabs(C2p[1]) * C1m[1] + W[2][1] * fabs(C2p[2]) * C1m[2] +
                W[3][1] * fabs(C2p[3]) * C1m[3];
          c1m = c1m / (W[0][1] * fabs(C2p[0]) + W[1][1] * fabs(C2p[1]) + W[2][1] * fabs(C2p[2]) + W[3][1] * fabs(C2p[3]));
          c2p = W[0][1] * fabs(C1m[0]) * C2p[0] + W[1][1] * fabs(C1m[1]) * C2p[1] + W[2][1] * fabs(C1m[2]) * C2p[2] +
                W[3][1] * fabs(C1m[3]) * C2p[3];
          c2p = c2p / (W[0][1] * fabs(C1m[0]) + W[1][1] * fabs(C1m[1]) + W[2][1] * fabs(C1m[2]) + W[3][1] * fabs(C1m[3]));
        }
        double v1a = c1p + c1m, v2a = c2p + c2m, B1a = c1p - c1m, B2a = c2p - c2m;
        emf[c0] = 0.25 * (v1a * B2a - v2a * B1a) * sqrt(o->dnc[c0]);
      }
}

static void cc2(mcfo *o, int orient) {
  int nx = o->n[0], ny = o->n[1], nz = o->n[2];
  double *r = U(1);
  if (orient == 0) {
    int l1[1] = {1};
    fillU(o, 7, 2, 1, l1, 1);
#pragma omp parallel for schedule(static) num_threads(o->nthr)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny + 1; j++)
        for (int i = 1; i <= nx + 1; i++)
          o->dnc[ix3(o, i, j, k)] = 0.25 * (r[ix3(o, i, j, k)] + r[ix3(o, i - 1, j, k)] + r[ix3(o, i, j - 1, k)] +
                                            r[ix3(o, i - 1, j - 1, k)]);
  } else if (orient == 1) {
#pragma omp parallel for schedule(static) num_threads(o->nthr)
    for (int k = 1; k <= nz + 1; k++)
      for (int j = 1; j <= ny + 1; j++)
        for (int i = 1; i <= nx; i++)
          o->dnc[ix3(o, i, j, k)] = 0.25 * (r[ix3(o, i, j, k)] + r[ix3(o, i, j - 1, k)] + r[ix3(o, i, j, k - 1)] +
                                            r[ix3(o, i, j - 1, k - 1)]);
  } else {
#pragma omp parallel for schedule(static) num_threads(o->nthr)
    for (int k = 1; k <= nz + 1; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx + 1; i++)
          o->dnc[ix3(o, i, j, k)] = 0.25 * (r[ix3(o, i, j, k)] + r[ix3(o, i - 1, j, k)] + r[ix3(o, i, j, k - 1)] +
                                            r[ix3(o, i - 1, j, k - 1)]);
  }
}

static void cc3(mcfo *o, double dt) {
  int nx = o->n[0], ny = o->n[1], nz = o->n[2];
  double *Ex = o->emf[0], *Ey = o->emf[1], *Ez = o->emf[2];
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx + 1; i++) {
        long c = ix3(o, i, j, k);
        U(6)[c] = U(6)[c] + dt / dxg(o, 1, j) * (Ez[ix3(o, i, j + 1, k)] - Ez[c]) -
                  dt / dxg(o, 2, k) * (Ey[ix3(o, i, j, k + 1)] - Ey[c]);
      }
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny + 1; j++)
      for (int i = 1; i <= nx; i++) {
        long c = ix3(o, i, j, k);
        U(7)[c] = U(7)[c] + dt / dxg(o, 2, k) * (Ex[ix3(o, i, j, k + 1)] - Ex[c]) -
                  dt / dxg(o, 0, i) * (Ez[ix3(o, i + 1, j, k)] - Ez[c]);
      }
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= nz + 1; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long c = ix3(o, i, j, k);
        U(8)[c] = U(8)[c] + dt / dxg(o, 0, i) * (Ey[ix3(o, i + 1, j, k)] - Ey[c]) -
                  dt / dxg(o, 1, j) * (Ex[ix3(o, i, j + 1, k)] - Ex[c]);
      }
}

void mcfo_cc4(mcfo *o) {
  int nx = o->n[0], ny = o->n[1], nz = o->n[2];
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long c = ix3(o, i, j, k);
        BCC(1)[c] = 0.5 * (U(6)[c] + U(6)[ix3(o, i + 1, j, k)]);
        BCC(2)[c] = 0.5 * (U(7)[c] + U(7)[ix3(o, i, j + 1, k)]);
        BCC(3)[c] = 0.5 * (U(8)[c] + U(8)[ix3(o, i, j, k + 1)]);
      }
  fill_ot(o, 7, 2, 2);
}

void mcfo_ct(mcfo *o, double dt) {
  cc1(o);
  for (int orient = 0; orient < 3; orient++) {
    cc2(o, orient);
    cct(o, orient, dt);
  }
  cc3(o, dt);
  mcfo_cc4(o);
}

/* ------------------------------------------------------------------ */
/* Source, time step, dissipation                                      */
/* ------------------------------------------------------------------ */

static void temp_and_flux(mcfo *o) {
  int nx = o->n[0], ny = o->n[1], nz = o->n[2];
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 0; k <= nz + 1; k++)
    for (int j = 0; j <= ny + 1; j++)
      for (int i = 0; i <= nx + 1; i++) {
        long c = ix3(o, i, j, k);
        o->Tn[c] = U(5)[c] / (KB * (SPC(SP)[c] + SPC(SH)[c] + SPC(SH2)[c] + SPC(SHE)[c] + SPC(SHEP)[c]));
      }
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        long c = ix3(o, i, j, k);
        long cx = ix3(o, i + 1, j, k), cy = ix3(o, i, j + 1, k), cz = ix3(o, i, j, k + 1);
        double dxi = dxg(o, 0, i), dxp = dxg(o, 0, i + 1);
        double dyi = dxg(o, 1, j), dyp = dxg(o, 1, j + 1);
        double dzi = dxg(o, 2, k), dzp = dxg(o, 2, k + 1);
        double rtTx = sqrt((dxi * o->Tn[cx] + dxp * o->Tn[c]) / (dxi + dxp));
        double rtTy = sqrt((dyi * o->Tn[cy] + dyp * o->Tn[c]) / (dyi + dyp));
        double rtTz = sqrt((dzi * o->Tn[cz] + dzp * o->Tn[c]) / (dzi + dzp));
        o->Qx[c] = KCOND * DMIN(rtTx, 3.873) * (o->Tn[cx] - o->Tn[c]) * 2.0 / (dxi + dxp);
        o->Qy[c] = KCOND * DMIN(rtTy, 3.873) * (o->Tn[cy] - o->Tn[c]) * 2.0 / (dyi + dyp);
        o->Qz[c] = KCOND * DMIN(rtTz, 3.873) * (o->Tn[cz] - o->Tn[c]) * 2.0 / (dzi + dzp);
      }
}

static void thermal(mcfo *o, double dt);

/* SOURCE (M3:2448-2780): SHIELD (ifrad = 2), the thermal stage (ifthrm =
 * 2), then the chemistry network (ifchem = 2).  nde and ndtot are formed
 * from the species wherever they are read; the reference's arrays of them
 * always equal that sum when read. */
void mcfo_source(mcfo *o, double dt) {
  if (o->p.ifrad == 2) shield(o);
  if (o->p.ifthrm == 2) thermal(o, dt);
  if (o->p.ifchem == 2) {
    const int nx = o->n[0], ny = o->n[1], nz = o->n[2];
#pragma omp parallel for schedule(static) num_threads(o->nthr)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) chem_cell(o, ix3(o, i, j, k), dt);
  }
}

static void thermal(mcfo *o, double dt) {
  int nx = o->n[0], ny = o->n[1], nz = o->n[2];
  int l15[2] = {1, 5}, l5[1] = {5};
  fillU(o, 7, 1, 1, l15, 2);
  fill_ot(o, 7, 1, 5);
  temp_and_flux(o);
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 0; k <= nz + 1; k++)
    for (int j = 0; j <= ny + 1; j++)
      for (int i = 0; i <= nx + 1; i++) o->Pn[ix3(o, i, j, k)] = U(5)[ix3(o, i, j, k)];
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      fillU(o, 7, 1, 1, l5, 1);
      temp_and_flux(o);
    }
#pragma omp parallel for schedule(static) num_threads(o->nthr)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          long c = ix3(o, i, j, k);
          double CooL = fcool_cell(o, c, o->Tn[c]);
          double gi = gammi1_of(o, c);
          double nsum = SPC(SH)[c] + SPC(SP)[c] + SPC(SHE)[c] + SPC(SHEP)[c] + SPC(SH2)[c];
          if (pass == 0) {
            if (dt <= 0.2 * o->Pn[c] / (gi * fabs(CooL)))
              U(5)[c] = U(5)[c] - gi * CooL * dt * 0.5;
            else
              U(5)[c] = imc_cell(o, c, U(5)[c], nsum, dt * 0.5, gi);
          } else {
            if (dt <= 0.2 * o->Pn[c] / (gi * fabs(CooL)))
              U(5)[c] = o->Pn[c] - gi * CooL * dt;
            else {
              o->Pn[c] = imc_cell(o, c, o->Pn[c], nsum, dt, gi);
              U(5)[c] = o->Pn[c];
            }
          }
          double tcd = gi * KCOND * sqrt(o->Tn[c]);
          double dxi = dxg(o, 0, i);
          tcd = 0.49 * nsum * KB / tcd * (dxi * dxi);
          tcd = copysign(0.5, tcd - dt);
          double div = (o->Qx[c] - o->Qx[ix3(o, i - 1, j, k)]) / dxi + (o->Qy[c] - o->Qy[ix3(o, i, j - 1, k)]) / dxg(o, 1, j) +
                       (o->Qz[c] - o->Qz[ix3(o, i, j, k - 1)]) / dxg(o, 2, k);
          if (pass == 0)
            U(5)[c] = U(5)[c] + (tcd + 0.5) * gi * dt * 0.5 * div;
          else
            U(5)[c] = U(5)[c] + (tcd + 0.5) * gi * dt * div;
          U(5)[c] = DMAX(U(5)[c], o->pmin);
          U(5)[c] = DMIN(U(5)[c], o->pmax);
        }
  }
}

/* The minima of Couran and Stblty are formed per z plane (in parallel) and
 * then over the planes in order; min is exact, so this equals the serial
 * running minimum for any finite values. */
double mcfo_couran(mcfo *o) {
  double *part = (double *)malloc(sizeof(double) * (size_t)o->n[2]);
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= o->n[2]; k++) {
    double tCFL = o->p.tfinal;
    for (int j = 1; j <= o->n[1]; j++)
      for (int i = 1; i <= o->n[0]; i++) {
        long c = ix3(o, i, j, k);
        double gamma = gam_of(o, c);
        double c2 = (gamma * U(5)[c] + BCC(1)[c] * BCC(1)[c] + BCC(2)[c] * BCC(2)[c] + BCC(3)[c] * BCC(3)[c]) / U(1)[c];
        double t;
        t = dxg(o, 0, i) / (sqrt(c2) + fabs(U(2)[c])); tCFL = DMIN(tCFL, t);
        t = dxg(o, 1, j) / (sqrt(c2) + fabs(U(3)[c])); tCFL = DMIN(tCFL, t);
        t = dxg(o, 2, k) / (sqrt(c2) + fabs(U(4)[c])); tCFL = DMIN(tCFL, t);
      }
    part[k - 1] = tCFL;
  }
  double tCFL = o->p.tfinal;
  for (int k = 0; k < o->n[2]; k++) tCFL = DMIN(tCFL, part[k]);
  free(part);
  return tCFL;
}

double mcfo_stblty(mcfo *o) {
  double *part = (double *)malloc(sizeof(double) * (size_t)o->n[2]);
#pragma omp parallel for schedule(static) num_threads(o->nthr)
  for (int k = 1; k <= o->n[2]; k++) {
    double tLMT = o->p.tfinal;
    for (int j = 1; j <= o->n[1]; j++)
      for (int i = 1; i <= o->n[0]; i++) {
        long c = ix3(o, i, j, k);
        double Nn = SPC(SP)[c] + SPC(SH)[c] + SPC(SH2)[c] + SPC(SHE)[c] + SPC(SHEP)[c];
        double Tn = U(5)[c] / (KB * Nn);
        double dl = DMIN(dxg(o, 0, i), dxg(o, 1, j));
        dl = DMIN(dl, dxg(o, 2, k));
        double gi = gammi1_of(o, c);
        double alpha = gi * KCOND * sqrt(DMIN(Tn, 15.0));
        alpha = 0.4 * Nn * KB / alpha * (dl * dl);
        double CooL = fcool_cell(o, c, Tn);
        double tauC = U(5)[c] / gi / fabs(CooL);
        tauC = DMAX(0.2 * tauC, 2.5e-4);
        tLMT = DMIN(tLMT, tauC);
        tLMT = DMIN(tLMT, alpha);
      }
    part[k - 1] = tLMT;
  }
  double tLMT = o->p.tfinal;
  for (int k = 0; k < o->n[2]; k++) tLMT = DMIN(tLMT, part[k]);
  free(part);
  return tLMT;
}

double mcfo_timestep(mcfo *o, double tsave, int *st) {
  double tfinal = o->p.tfinal;
  double dt = tfinal;
  double t = o->p.cfl * mcfo_couran(o);
  dt = DMIN(dt, t);
  double stt = dt;
  int tag = 1;
  double tl = mcfo_stblty(o);
  dt = DMIN(dt, tl);
  if (dt < stt) tag = 2;
  if (st) *st = tag;
  if (o->time + dt > tfinal) dt = tfinal - o->time;
  if (o->time + dt > tsave) dt = tsave - o->time;
  return dt;
}

void mcfo_dissip(mcfo *o) {
  int l[3] = {2, 3, 4};
  fillU(o, 7, 2, 1, l, 3);
  int nx = o->n[0], ny = o->n[1], nz = o->n[2];
  long plane = (long)(ny + 2) * (nz + 2);
  double *divv = (double *)calloc((size_t)2 * 12 * plane, sizeof(double));
#define DV(e, i, j, k) divv[(size_t)(e) * 12 * plane + (size_t)(i) + 12 * ((size_t)(j) + (size_t)(ny + 2) * (size_t)(k))]
  int offs[2] = {0, nx - 10};
  int e0 = (o->nsp[0] == 1) ? 1 : 0; /* one block in x: only the right end (M3:3156-3158) */
  for (int e = e0; e < 2; e++)
    for (int k = 0; k <= nz + 1; k++)
      for (int j = 0; j <= ny + 1; j++)
        for (int i = 0; i <= 11; i++) {
          int x = i + offs[e];
          DV(e, i, j, k) = (U(2)[ix3(o, x + 1, j, k)] - U(2)[ix3(o, x - 1, j, k)] + U(3)[ix3(o, x, j + 1, k)] -
                            U(3)[ix3(o, x, j - 1, k)] + U(4)[ix3(o, x, j, k + 1)] - U(4)[ix3(o, x, j, k - 1)]) *
                           0.25;
        }
  for (int e = e0; e < 2; e++)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= 10; i++) {
          long c = ix3(o, i + offs[e], j, k);
          U(2)[c] = U(2)[c] + 0.1 * (DV(e, i + 1, j, k) - DV(e, i - 1, j, k));
          U(3)[c] = U(3)[c] + 0.1 * (DV(e, i, j + 1, k) - DV(e, i, j - 1, k));
          U(4)[c] = U(4)[c] + 0.1 * (DV(e, i, j, k + 1) - DV(e, i, j, k - 1));
        }
#undef DV
  free(divv);
}

/* GRAVTY(dt,2) (SG8:15-79): the density ghosts as BC_MPI(1,1) and PB's
 * BC_MPI(2,1) fill them, then the multigrid solve into Phi. */
int mcfo_grav_solve(mcfo *o) {
  if (!o->grav) return -1;
  static const int l1[1] = {1};
  fillU(o, 7, 1, 1, l1, 1);
  fillU(o, 7, 2, 1, l1, 1);
  mcfg_solve(o->grav, U(1), o->phi);
  return 0;
}

/* GRAVTY(dt,3) (SG8:82-89): the velocities take dt/2 of the fourth-order
 * potential gradient, dxi = 1/(12 dx(0)) with dx(0) of the cell's rank. */
void mcfo_grav_kick(mcfo *o, double dt) {
  if (!o->phi) return;
  const int nx = o->n[0], cx = o->n[0] / o->nsp[0];
  const double *P = o->phi;
  const long sx = o->st[0], sy = o->st[1], sz = o->st[2];
  for (int k = 1; k <= o->n[2]; k++)
    for (int j = 1; j <= o->n[1]; j++)
      for (int i = 1; i <= nx; i++) {
        const double dxi = 1.0 / (12.0 * dxg(o, 0, ((i - 1) / cx) * cx));
        const long c = ix3(o, i, j, k);
        U(2)[c] = U(2)[c] - dt * (-P[c + 2 * sx] + 8.0 * P[c + sx] - 8.0 * P[c - sx] + P[c - 2 * sx]) * dxi * 0.5;
        U(3)[c] = U(3)[c] - dt * (-P[c + 2 * sy] + 8.0 * P[c + sy] - 8.0 * P[c - sy] + P[c - 2 * sy]) * dxi * 0.5;
        U(4)[c] = U(4)[c] - dt * (-P[c + 2 * sz] + 8.0 * P[c + sz] - 8.0 * P[c - sz] + P[c - 2 * sz]) * dxi * 0.5;
      }
}

double *mcfo_phi(mcfo *o) { return o->phi; }

/* Sweep order for ifEVO = 1..6 (M3:611-635) */
static const int sweep_order[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {1, 0, 2}, {2, 1, 0}};

void mcfo_step(mcfo *o, double dt) {
  if (o->p.ifgrv == 2) mcfo_grav_kick(o, dt); /* M3:607 */
  mcfo_source(o, 0.5 * dt);
  const int *ord = sweep_order[o->ifevo - 1];
  for (int t = 0; t < 3; t++) mhd_sweep(o, ord[t], dt);
  o->ifevo = (o->ifevo % 6) + 1;
  mcfo_ct(o, dt);
  mcfo_source(o, 0.5 * dt);
  if (o->p.ifgrv == 2) { /* M3:653 */
    mcfo_grav_solve(o);
    mcfo_grav_kick(o, dt);
  }
  mcfo_dissip(o);
  o->time = o->time + dt;
}

int mcfo_run(mcfo *o, int nsteps, double *dts, int *tags) {
  const double dtsave = 0.25;
  int itime = 1 + (int)((o->time + 1.0e-8) / dtsave);
  double tsave = dtsave * (double)itime;
  int done = 0;
  for (int s = 0; s < nsteps; s++) {
    if (o->time >= o->p.tfinal) break;
    if (o->time >= tsave) break;
    int st = 1;
    double dt = mcfo_timestep(o, tsave, &st);
    mcfo_step(o, dt);
    if (dts) dts[s] = dt;
    if (tags) tags[s] = st;
    done++;
  }
  return done;
}

/* ---------------- known-answer entry points ---------------- */
void mcfo_rieman(const double *in8, double gamma, double *pa, double *va) {
  rieman(in8[0], in8[1], in8[2], in8[3], in8[4], in8[5], in8[6], in8[7], gamma - 1.0, gamma - 2.0, gamma - 3.0,
         gamma + 1.0, gamma + 3.0, pa, va);
}
void mcfo_chareq(const double *in6, double *bay, double *vay) {
  chareq(in6[0], in6[1], in6[2], in6[3], in6[4], in6[5], bay, vay);
}
double mcfo_fcool_ki(double rho, double T) { return fcool_ki(rho, T); }
double mcfo_imc(double P, double n, double dt, double rho, double gammi1) { return imc(P, n, dt, rho, gammi1); }
void mcfo_vlimit1d(int n, const double *q, const double *T, const double *dxx, double *grd) {
  /* express T through U5 = T*kb/1.27 with U1 = 1 */
  double *u5 = (double *)malloc(sizeof(double) * (n + 4));
  double *u1 = (double *)malloc(sizeof(double) * (n + 4));
  for (int i = 0; i < n + 4; i++) { u5[i] = T[i] * KB / 1.27; u1[i] = 1.0; }
  vlimit(q, u5, u1, 1, 1, n, dxx + 1, grd);
  free(u5);
  free(u1);
}
