// mcf_host.hip -- host side of the C ABI (include/mcf_abi.h): the device
// context, the ghost-fill schedule of the reference's BC_MPI/BC_MPI_OT calls
// and the stage order of EVOLVE (3DMHD-CfCA/MCF3DSGincfca3.f90, "M3").
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mcf_abi.h"
#include "mcf_grav.h"
#include "mcf_kernels.h"
#include "mcf_mlp.h"

using namespace mcf;

static thread_local std::string g_err;

#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) {                                                                        \
      g_err = std::string(#x) + ": " + hipGetErrorString(e_);                                     \
      return MCF_E_HIP;                                                                            \
    }                                                                                              \
  } while (0)
#define CHK(x)                                                                                     \
  do {                                                                                             \
    int r_ = (x);                                                                                  \
    if (r_ != MCF_OK) return r_;                                                                   \
  } while (0)

enum { F_RHO = 0, F_VX, F_VY, F_VZ, F_P, F_BX, F_BY, F_BZ };

// Device scratch of mcf_cool_solve_noalloc: the fallback counter pair in the
// first 256 bytes, then the queue of cells the surrogate rejected.  Each GPU
// holds its own copy until the process ends.  Host code never passes this
// name to a kernel: the device address is taken once per device with
// hipGetSymbolAddress (noalloc_scratch) and only that address is used.
static constexpr size_t NOALLOC_HDR = 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_HDR + (size_t)MCF_NOALLOC_MAX_CELLS * sizeof(int)];
static constexpr int NOALLOC_MAX_DEV = 64;
static void *g_scratch_ptr[NOALLOC_MAX_DEV];
static std::mutex g_scratch_mu;

// The device address of `scratch` on device dev (current), looked up on the
// first call for that device.
static int noalloc_scratch(int dev, void **ptr) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEV) { g_err = "device ordinal out of range for the no-allocation scratch"; return MCF_E_ARG; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return MCF_E_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return MCF_OK;
}

struct mcf_ctx {
  mcf_config cfg;
  int dev = 0;
  hipStream_t stream = nullptr;
  Geo g{};
  size_t ncell = 0;
  double *pool = nullptr;
  double *U[8]{}, *sp[8]{}, *bcc[3]{};
  // second buffers of the fields the sweep stages write (swapped per stage)
  double *Ub[5]{}, *spb[8]{};
  // sweep scratch
  double *u5e = nullptr, *blg[2]{}, *blgb[2]{}, *xlag = nullptr, *dxlagM = nullptr, *va = nullptr;
  // CT scratch
  double *vfc[3]{}, *emf[3]{};
  double *srcq[3]{}; // SOURCE: conduction increments and the two stiff-cell queues
  double *divv = nullptr, *red = nullptr;
  double *gx[3]{}, *gdx[3]{};
  std::vector<double> hx[3], hdx[3];
  double bbrv[11][3][3]{}; // BBRV(k, mode, side), 1-based
  double bbrv_cm[8]{};
  double dtf = 0, pmin = 0, pmax = 0, rmin = 0, spmin[8]{};
  float *dtfd = nullptr; // DTF field on the device (mcf_set_inflow_density), null = uniform dtf
  // column densities Ntot, NH2, NnC, NCO, tCII from the left / right x end
  // (col[2q + dir]) and SHIELD's per-cell scratch; allocated when the
  // chemistry Fcool, the network or SHIELD is on
  double *col[10]{}, *ctemp = nullptr;
  bool chem = false;    // MCF_COOL_CHEM: the chemistry-network Fcool
  bool has_cols = false;
  int nsp[3]{1, 1, 1};
  int ncu = 256; // compute units of the device
  double time = 0.0;
  int ifevo = 1;
  double *hred = nullptr; // pinned 2 doubles
  // per-kernel event timing (mcf_profile)
  bool prof = false;
  bool mixed = false; // MCF_PREC_MIXED: FP32 reconstruction / Riemann / EMF
  // y/z sweeps run as column marches: needs a uniform spacing along the
  // axis (ghosts included) and n >= 8
  bool fuse[3]{false, false, false};
  // cells per segment of k_sweep_col: 256 (the whole 256^3 column) measured
  // 1.66 / 2.32 ms per y sweep (mixed / FP64) against 1.68 / 2.37 with 128,
  // 1.92 / 2.70 with 96 and 2.48 / 3.57 with 192 (an uneven split)
  int col_seg = 256;
  // cooling surrogate: FP32 weights, Adam moments, gradient (4 x MLP_NP),
  // packed BF16 operands, counters {backward-Euler cells, IMC fallbacks}
  float *mlp = nullptr;
  __bf16 *mlp_pk = nullptr;
  unsigned long long *cool_cnt = nullptr;
  bool mlp_ready = false;
  int src_pair = 0; // which of the two SOURCE queue-counter pairs the next stage counts in
  void *scratch = nullptr; // this device's address of the no-allocation scratch
  long adam_t = 0;
  struct Rec { const char *name; const char *stage; hipEvent_t a, b; double bytes; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> evpool;
  // stage of the step the launches belong to (mcf_profile_stages) and how
  // many times each stage ran while profiling
  const char *stage = "other";
  std::map<std::string, int> stage_runs;
  // self-gravity (ifgrv = 2): multigrid geometry and arrays, the potential
  // Phi (layout of U) and PB's Fourier work arrays
  bool grav = false;
  GravGeo gg{};
  GravBuf gb{};
  PbArgs pb{};
  double *gpool = nullptr; // one allocation for all of it
  long gpool_n = 0;        // doubles in gpool cleared at each solve
  double *phi = nullptr;
  double2 *pbZ = nullptr, *pbS = nullptr, *pbD = nullptr, *pbSD = nullptr;
};

// Marks the launches made while it lives as one run of a stage of the step.
struct StageScope {
  mcf_ctx *x;
  const char *prev;
  StageScope(mcf_ctx *c, const char *name) : x(c), prev(c->stage) {
    x->stage = name;
    if (x->prof && std::strcmp(prev, name) != 0) x->stage_runs[name] += 1;
  }
  ~StageScope() { x->stage = prev; }
};

static hipEvent_t prof_event(mcf_ctx *x) {
  if (!x->evpool.empty()) { hipEvent_t e = x->evpool.back(); x->evpool.pop_back(); return e; }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

extern "C" void mcf_default_config(mcf_config *c) {
  std::memset(c, 0, sizeof(*c));
  c->np1x = c->np2x = c->np1y = c->np2y = c->np1z = c->np2z = 64;
  c->ql1x = c->ql2x = c->ql1y = c->ql2y = c->ql1z = c->ql2z = 10.0;
  c->vinitx1 = 9.778; c->vinitx2 = -9.778;
  c->binity1 = c->binity2 = 2.2307;
  c->cfl = 0.49; c->facdep = 1.0;
  c->maxstp = 4; c->nitera = 5000; c->tfinal = 100.0;
  c->bcx1 = c->bcx2 = 4; c->bcy1 = c->bcy2 = c->bcz1 = c->bcz2 = 3;
  c->ifchem = 0; c->ifthrm = 2; c->ifrad = 0; c->ifgrv = 0;
  c->pinit_over_kb = 8.810807e3;
  c->Hini = 0.9219098; c->pini = 0.9503446e-2; c->H2ini = 0.9465513e-8;
  c->Heini = 0.9155226e-1; c->Hepini = 0.5655353e-3;
  c->Cini = 0.1565848e-8; c->COini = 0.2202631e-20; c->Cpini = 0.1433520e-3;
  c->nsplitx = c->nsplity = c->nsplitz = 2;
  c->precision = MCF_PREC_FP64;
  c->cooling = MCF_COOL_KI;
}

extern "C" const char *mcf_last_error(void) { return g_err.c_str(); }

// INITIA (M3:155-298): boundary states, floors, grid
static void setup_constants(mcf_ctx *x) {
  const mcf_config &p = x->cfg;
  const double mH = 1.0, mHe = 4.0, mH2 = 2.0, xc = 1.4e-4;
  double pinit1 = p.pinit_over_kb * KB * 1.0e-3;
  double dinit1 = mH * p.Hini + mH * p.pini + mH2 * p.H2ini + mHe * p.Heini + mHe * p.Hepini;
  double s4 = p.Hini + p.pini + p.Heini + p.Hepini;
  double gamma = (5.0 * s4 + 7.0 * p.H2ini) / (3.0 * s4 + 5.0 * p.H2ini);
  double gammi1 = gamma - 1.0;
  x->pmin = 1.829797 * 8.6336;
  x->pmax = 1.0e10;
  x->rmin = 0.1949628;
  double smin[8] = {x->rmin * 0.91, 1.0e-20, 1.0e-20, x->rmin * 0.09, 1.0e-20, 1.0e-20, 1.0e-20, 1.0e-20};
  std::memcpy(x->spmin, smin, sizeof(smin));
  double *cm = x->bbrv_cm;
  cm[0] = 0.91 / 1.27;
  cm[1] = 0.91 * p.pini / p.Hini / 1.27;
  cm[2] = 0.91 * p.H2ini / p.Hini / 1.27;
  cm[3] = 0.09 / 1.27;
  cm[4] = 0.09 * p.Hepini / p.Heini / 1.27;
  cm[7] = xc / 1.27;
  cm[5] = xc * p.Cini / p.Cpini / 1.27;
  cm[6] = xc * p.COini / p.Cpini / 1.27;
  double dBC = mH * cm[0] + mH * cm[1] + mH2 * cm[2] + mHe * cm[3] + mHe * cm[4];
  auto &B = x->bbrv;
  B[1][1][1] = B[1][2][1] = B[1][1][2] = B[1][2][2] = dBC;
  B[2][1][1] = p.vinitx1; B[2][2][1] = dBC * p.vinitx1; B[2][1][2] = -p.vinitx1; B[2][2][2] = -dBC * p.vinitx1;
  B[3][1][1] = p.vinity1; B[3][2][1] = dBC * p.vinity1; B[3][1][2] = p.vinity1; B[3][2][2] = dBC * p.vinity1;
  B[4][1][1] = p.vinitz1; B[4][2][1] = dBC * p.vinitz1; B[4][1][2] = p.vinitz1; B[4][2][2] = dBC * p.vinitz1;
  double e5 = pinit1 / gammi1 + 0.5 * (dBC * (p.vinitx1 * p.vinitx1) + p.binitx1 * p.binitx1 +
                                       p.binity1 * p.binity1 + p.binitz1 * p.binitz1);
  B[5][1][1] = pinit1; B[5][2][1] = e5; B[5][1][2] = pinit1; B[5][2][2] = e5;
  B[6][1][1] = B[6][2][1] = p.binitx1; B[6][1][2] = B[6][2][2] = p.binitx2;
  B[7][1][1] = B[7][2][1] = p.binity1; B[7][1][2] = B[7][2][2] = p.binity2;
  B[8][1][1] = B[8][2][1] = p.binitz1; B[8][1][2] = B[8][2][2] = p.binitz2;
  x->dtf = (double)(float)dinit1; // DTF is REAL*4 (M3:22), filled with dinit1
  int np1[3] = {p.np1x, p.np1y, p.np1z};
  double ql1[3] = {p.ql1x, p.ql1y, p.ql1z};
  for (int d = 0; d < 3; ++d) {
    int n = x->g.n[d];
    x->hx[d].assign(n + 4, 0.0);
    x->hdx[d].assign(n + 4, 0.0);
    for (int i = -1; i <= n + 2; ++i) {
      double v;
      if (d == 0) v = (i <= p.np1x) ? p.ql1x / (double)p.np1x : p.ql2x / (double)p.np2x;
      else v = ql1[d] / (double)np1[d];
      x->hdx[d][i + 1] = v;
    }
    x->hx[d][0] = -x->hdx[d][1];
    for (int i = 0; i <= n + 2; ++i) x->hx[d][i + 1] = x->hx[d][i] + x->hdx[d][i + 1];
  }
}

// ----------------------------------------------------------------------------
// launch helpers
// ----------------------------------------------------------------------------
static const int TPB = 256;
// blocks of the dt reduction's first pass (one partial min pair each)
static const int DT_BLOCKS = 8192;

static Box mkbox(int x0, int x1, int y0, int y1, int z0, int z1) {
  Box b;
  b.lo[0] = x0; b.cnt[0] = x1 - x0 + 1;
  b.lo[1] = y0; b.cnt[1] = y1 - y0 + 1;
  b.lo[2] = z0; b.cnt[2] = z1 - z0 + 1;
  return b;
}
// box with axis d over [a0,a1] and the other axes over [1,n]
static Box sweepbox(const mcf_ctx *x, int d, int a0, int a1) {
  int lo[3], hi[3];
  for (int e = 0; e < 3; ++e) { lo[e] = 1; hi[e] = x->g.n[e]; }
  lo[d] = a0; hi[d] = a1;
  return mkbox(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
}
static unsigned nblocks(long n) { return (unsigned)((n + TPB - 1) / TPB); }
static long boxsize(const Box &b) { return (long)b.cnt[0] * b.cnt[1] * b.cnt[2]; }

// Launch on an explicit grid; bytes = algorithmic bytes the launch moves,
// recorded under name for mcf_profile.
#define KLAUNCH_N(name, kern, grid, block, bytes, ...)                                              \
  do {                                                                                             \
    hipEvent_t ea_ = nullptr, eb_ = nullptr;                                                       \
    if (x->prof) { ea_ = prof_event(x); eb_ = prof_event(x); HIPCHK(hipEventRecord(ea_, x->stream)); } \
    hipLaunchKernelGGL(kern, (grid), (block), 0, x->stream, __VA_ARGS__);                          \
    HIPCHK(hipGetLastError());                                                                     \
    if (x->prof) {                                                                                 \
      HIPCHK(hipEventRecord(eb_, x->stream));                                                      \
      x->recs.push_back({name, x->stage, ea_, eb_, (double)(bytes)});                              \
    }                                                                                              \
  } while (0)
#define KLAUNCH_G(kern, grid, block, bytes, ...) KLAUNCH_N(#kern, kern, grid, block, bytes, __VA_ARGS__)

// Launch with nthreads threads; narr = arrays the kernel reads or writes
// per thread, for the algorithmic byte count of mcf_profile.
#define KLAUNCH(kern, nthreads, narr, ...)                                                          \
  do {                                                                                             \
    long nt_ = (nthreads);                                                                         \
    if (nt_ > 0)                                                                                   \
      KLAUNCH_G(kern, dim3(nblocks(nt_)), dim3(TPB), 8.0 * (double)nt_ * (double)(narr), __VA_ARGS__); \
  } while (0)

// KLAUNCH with blocks of 512 threads (the CT face and cell kernels: 3-8%
// faster than 256; the fills are slower with it)
#define KLAUNCH512(kern, nthreads, narr, ...)                                                       \
  do {                                                                                             \
    long nt_ = (nthreads);                                                                         \
    if (nt_ > 0)                                                                                   \
      KLAUNCH_G(kern, dim3((unsigned)((nt_ + 511) / 512)), dim3(512), 8.0 * (double)nt_ * (double)(narr), __VA_ARGS__); \
  } while (0)

// KLAUNCH of the instance kern<R>, recorded under kern's name
#define KLAUNCH_R(kern, R, nthreads, narr, ...)                                                     \
  do {                                                                                             \
    long nt_ = (nthreads);                                                                         \
    if (nt_ > 0)                                                                                   \
      KLAUNCH_N(#kern, (kern<R>), dim3(nblocks(nt_)), dim3(TPB), 8.0 * (double)nt_ * (double)(narr), __VA_ARGS__); \
  } while (0)

// y/z sweeps as two waves per row of columns (k_sweep_col2) on the mixed
// path: 1.39 / 1.45 ms per y / z sweep against 1.62 / 1.70 for
// k_sweep_col at 256^3 (250 registers, no spill, two waves per SIMD).  The
// FP64 instance spills 224 bytes per lane to scratch at two waves per SIMD
// (4.47 against 2.31 ms at 256^3), so FP64 takes it only for launches of
// fewer blocks than the device has CUs, where the second wave per column
// row is otherwise idle hardware (2.2x faster on a 72 x 44 x 136 grid).

// ----------------------------------------------------------------------------
// ghost fills
// ----------------------------------------------------------------------------
// The y and z periodic passes of one fill in one launch (k_fill_yz); dm[f]
// says which of the two passes field f takes (bit 0: y, bit 1: z).
static int fill_yz(mcf_ctx *x, double *const *a, const int *dm, int nf, int nol) {
  FillList L{};
  L.nf = nf;
  for (int f = 0; f < nf; ++f) { L.a[f] = a[f]; L.dm[f] = dm[f]; }
  const long m0 = x->g.n[0] + 4, m1 = x->g.n[1] + 4, m2 = x->g.n[2] + 4;
  const long nt = 2L * nol * m0 * m2 + 2L * nol * m0 * (m1 - 2 * nol);
  KLAUNCH(k_fill_yz, nt, 2 * nf, L, x->g, nol);
  return MCF_OK;
}

static bool x_inflow(const mcf_ctx *x) { return x->cfg.bcx1 == 4 || x->cfg.bcx2 == 4; }

static int launch_xfill(mcf_ctx *x, XFill &X, int nol) {
  X.rho = x->U[F_RHO];
  X.bcc1 = x->bcc[0];
  X.blg1 = x->blg[0];
  X.blg2 = x->blg[1];
  if (x->dtfd) {
    // the boundary ranks' dx(0) and L = dx(0)*Ncellx*NSPLTx (M3:821-822, 857-858)
    const int cx = x->g.n[0] / x->nsp[0], nsx = x->nsp[0];
    X.dtf.f = x->dtfd;
    X.dtf.xe = x->gx[0];
    X.dtf.t = x->time;
    for (int s = 0; s < 2; ++s) {
      const double dx0 = x->hdx[0][(s == 0 ? 0 : (nsx - 1) * cx) + 1];
      X.dtf.dx0[s] = dx0;
      X.dtf.L[s] = dx0 * (double)cx * (double)nsx;
      X.dtf.v[s] = x->bbrv[2][1][s + 1];
    }
  }
  long nt = (long)(x->g.n[1] + 4) * (x->g.n[2] + 4) * 2 * nol;
  KLAUNCH(k_fill_xrule, nt, 2 * X.nf, X, x->g, nol);
  return MCF_OK;
}

enum { OT_DXLAG = 1, OT_BCC, OT_BLG, OT_VFC, OT_SPECIES };

// the fields a BC_MPI_OT piece fills along axis d
static int ot_fields(mcf_ctx *x, int what, int d, double **a) {
  int nf = 0;
  if (what == OT_DXLAG) a[nf++] = x->dxlagM;
  if (what == OT_BCC) for (int c = 0; c < 3; ++c) a[nf++] = x->bcc[c];
  if (what == OT_BLG) { a[nf++] = x->blg[0]; a[nf++] = x->blg[1]; }
  if (what == OT_VFC) {
    int c1 = d == 0 ? 1 : 0, c2 = d == 2 ? 1 : 2; // x: Vfc 2,3; y: 1,3; z: 1,2
    a[nf++] = x->vfc[c1]; a[nf++] = x->vfc[c2];
  }
  if (what == OT_SPECIES) for (int s = 0; s < 8; ++s) a[nf++] = x->sp[s];
  return nf;
}

// Fills that run back to back, gathered into one x launch (the x rules or
// the periodic copy, fields in call order) and one y/z launch.  Equal to the
// calls in sequence when no x rule reads, at a y or z ghost row, a field an
// earlier call fills: the rules that read other fields are the species rule
// (interior rows only) and the mode-2 momentum and energy rules, which are
// never gathered with another call.
struct FillPlan {
  int nol;
  XFill X{};
  double *yz[16];
  int dm[16];
  int nyz = 0;
  bool over = false; // more than 16 fields in a pass
  explicit FillPlan(int n) : nol(n) {}
  int add_x() {
    if (X.nf == 16) { over = true; return 15; }
    return X.nf++;
  }
  void add_yz(double *p, int bits) {
    int q = 0;
    while (q < nyz && yz[q] != p) ++q;
    if (q == nyz) {
      if (nyz == 16) { over = true; return; }
      yz[nyz] = p; dm[nyz] = 0; ++nyz;
    }
    dm[q] |= bits;
  }
};

// BC_MPI(nol, mode) for the 1-based U components in list (M3:800-926)
static void planU(mcf_ctx *x, FillPlan &P, int dmask, int mode, const int *list, int nk) {
  int bc[2] = {x->cfg.bcx1, x->cfg.bcx2};
  for (int t = 0; t < nk; ++t) {
    const int K = list[t];
    double *a = x->U[K - 1];
    if (dmask & 1) {
      XFill &X = P.X;
      const int u = P.add_x();
      X.a[u] = a;
      for (int s = 0; s < 2; ++s) {
        int side = s + 1;
        if (bc[s] == 1) { // outflow: zero-gradient copy; face B_x as the inflow rule
          if (K == 6) X.rule[u][s] = XR_FACE;
          else { X.rule[u][s] = XR_COPY; X.src[u][s] = s == 0 ? 1 : x->g.n[0]; }
          continue;
        }
        if (bc[s] != 4) { X.rule[u][s] = XR_PERIODIC; continue; }
        if (K == 1) { X.rule[u][s] = x->dtfd ? XR_DTF : XR_CONST; X.val[u][s] = x->dtf; }
        else if (K == 6) X.rule[u][s] = XR_FACE;
        else if (K == 5 && mode == 2) {
          X.rule[u][s] = XR_ENERGY; X.val[u][s] = x->bbrv[5][1][side]; X.val2[s] = x->bbrv[2][1][side];
        } else if (K == 2 && mode == 2) { X.rule[u][s] = XR_MOM; X.val[u][s] = x->bbrv[2][1][side]; }
        else { X.rule[u][s] = XR_CONST; X.val[u][s] = x->bbrv[K][mode][side]; }
      }
    }
    if (dmask & 6) P.add_yz(a, (dmask >> 1) & 3);
  }
}

// BC_MPI_OT(nol, .) pieces (M3:930-1254)
static void planOT(mcf_ctx *x, FillPlan &P, int dmask, int what) {
  int nx = x->g.n[0];
  int bc[2] = {x->cfg.bcx1, x->cfg.bcx2};
  double *a[16];
  if (dmask & 1) {
    const int nf = ot_fields(x, what, 0, a);
    XFill &X = P.X;
    for (int t = 0; t < nf; ++t) {
      const int u = P.add_x();
      X.a[u] = a[t];
      for (int s = 0; s < 2; ++s) {
        int side = s + 1;
        if (bc[s] == 1) { X.rule[u][s] = XR_COPY; X.src[u][s] = s == 0 ? 1 : nx; continue; }
        if (bc[s] != 4) { X.rule[u][s] = XR_PERIODIC; continue; }
        switch (what) {
        case OT_DXLAG: X.rule[u][s] = XR_CONST; X.val[u][s] = x->hdx[0][s == 0 ? 1 : nx + 2]; break;
        case OT_BCC:
          if (t == 0) { X.rule[u][s] = XR_COPY; X.src[u][s] = s == 0 ? 1 : nx; }
          else { X.rule[u][s] = XR_CONST; X.val[u][s] = x->bbrv[6 + t][1][side]; }
          break;
        case OT_BLG: X.rule[u][s] = XR_CONST; X.val[u][s] = x->bbrv[7 + t][1][side]; break;
        case OT_VFC: X.rule[u][s] = XR_COPY; X.src[u][s] = s == 0 ? 1 : nx; break;
        case OT_SPECIES: X.rule[u][s] = XR_SPECIES; X.val[u][s] = x->bbrv_cm[t]; break;
        }
      }
    }
  }
  for (int d = 1; d <= 2; ++d) {
    if (!(dmask & (1 << d))) continue;
    const int nf = ot_fields(x, what, d, a);
    for (int t = 0; t < nf; ++t) P.add_yz(a[t], d);
  }
}

// the x launch, then the y/z launch
static int flush(mcf_ctx *x, FillPlan &P) {
  if (P.over) { g_err = "internal: a fill plan holds more than 16 fields"; return MCF_E_STATE; }
  if (P.X.nf > 0) CHK(launch_xfill(x, P.X, P.nol));
  if (P.nyz > 0) CHK(fill_yz(x, P.yz, P.dm, P.nyz, P.nol));
  return MCF_OK;
}

static int fillU(mcf_ctx *x, int dmask, int nol, int mode, const int *list, int nk) {
  FillPlan P(nol);
  planU(x, P, dmask, mode, list, nk);
  return flush(x, P);
}

static int fillOT(mcf_ctx *x, int dmask, int nol, int what) {
  FillPlan P(nol);
  planOT(x, P, dmask, what);
  return flush(x, P);
}

// ----------------------------------------------------------------------------
// stages
// ----------------------------------------------------------------------------
static SweepIO sweep_io(mcf_ctx *x, int d) {
  SweepIO A{};
  for (int f = 0; f < 8; ++f) { A.U[f] = x->U[f]; A.sp[f] = x->sp[f]; A.spo[f] = x->spb[f]; A.spmin[f] = x->spmin[f]; }
  for (int f = 0; f < 5; ++f) A.Uo[f] = x->Ub[f];
  for (int c = 0; c < 3; ++c) A.bcc[c] = x->bcc[c];
  A.blg[0] = x->blg[0]; A.blg[1] = x->blg[1]; A.blgo[0] = x->blgb[0]; A.blgo[1] = x->blgb[1];
  A.u5e_in = x->u5e; A.u5e = x->u5e;
  A.xlag = x->xlag; A.dxlagM = x->dxlagM; A.va = x->va;
  A.xe = x->gx[d]; A.dxe = x->gdx[d];
  A.d = d;
  static const int VN[3] = {1, 2, 3}, VT1[3] = {2, 3, 1}, VT2[3] = {3, 1, 2}, BN[3] = {5, 6, 7};
  static const int BT1[3] = {1, 2, 0}, BT2[3] = {2, 0, 1};
  A.VN = VN[d]; A.VT1 = VT1[d]; A.VT2 = VT2[d]; A.BN = BN[d]; A.BT1 = BT1[d]; A.BT2 = BT2[d];
  A.nblk = x->g.n[d] / x->nsp[d];
  A.facdep = x->cfg.facdep; A.rmin = x->rmin; A.pmin = x->pmin; A.pmax = x->pmax;
  return A;
}

static void swapU(mcf_ctx *x, int f) { double *t = x->U[f]; x->U[f] = x->Ub[f]; x->Ub[f] = t; }
static void swapSp(mcf_ctx *x) { for (int f = 0; f < 8; ++f) { double *t = x->sp[f]; x->sp[f] = x->spb[f]; x->spb[f] = t; } }
static void swapBlg(mcf_ctx *x) { for (int c = 0; c < 2; ++c) { double *t = x->blg[c]; x->blg[c] = x->blgb[c]; x->blgb[c] = t; } }

static dim3 sweep_grid(const mcf_ctx *x, int d) {
  int nx = x->g.n[0], ny = x->g.n[1], nz = x->g.n[2];
  unsigned gx = (unsigned)((nx + 63) / 64);
  if (d == 0) return dim3(gx, (unsigned)((ny + 7) / 8), (unsigned)nz);
  if (d == 1) return dim3(gx, (unsigned)((ny + 7) / 8), (unsigned)nz);
  return dim3(gx, (unsigned)((nz + 7) / 8), (unsigned)ny);
}

// Stage kernel launch: strict FP64 (R = double) or mixed (R = float).
// RHS and MOC use 9 waves on the x sweep and 10 on y/z (halo positions),
// REMAP 9 waves everywhere.
// Profile names carry the sweep axis as the template argument rocprofv3
// shows first ("k_rhs<1>" for k_rhs<1, double> and k_rhs<1, float>).
#define SWEEP_LAUNCH(kern, d, wide, narr, ...)                                                      \
  do {                                                                                             \
    dim3 gr_ = sweep_grid(x, d);                                                                   \
    const unsigned bt_ = (d != 0 && (wide)) ? 640u : 576u;                                         \
    const double by_ = 8.0 * (double)x->g.n[0] * x->g.n[1] * x->g.n[2] * (narr);                   \
    static const char *const nm_[3] = {#kern "<0>", #kern "<1>", #kern "<2>"};                     \
    if (x->mixed) {                                                                                \
      if (d == 0) KLAUNCH_N(nm_[0], (kern<0, float>), gr_, dim3(bt_), by_, __VA_ARGS__);                  \
      else if (d == 1) KLAUNCH_N(nm_[1], (kern<1, float>), gr_, dim3(bt_), by_, __VA_ARGS__);             \
      else KLAUNCH_N(nm_[2], (kern<2, float>), gr_, dim3(bt_), by_, __VA_ARGS__);                         \
    } else {                                                                                       \
      if (d == 0) KLAUNCH_N(nm_[0], (kern<0, double>), gr_, dim3(bt_), by_, __VA_ARGS__);                 \
      else if (d == 1) KLAUNCH_N(nm_[1], (kern<1, double>), gr_, dim3(bt_), by_, __VA_ARGS__);            \
      else KLAUNCH_N(nm_[2], (kern<2, double>), gr_, dim3(bt_), by_, __VA_ARGS__);                        \
    }                                                                                              \
  } while (0)

// MHD (M3:1264-1324): RHS -> MOC -> REMAP -> primitives
static int do_sweep(mcf_ctx *x, int d, double dt) {
  static const char *const names[3] = {"sweep_x", "sweep_y", "sweep_z"};
  StageScope scope(x, names[d]);
  SweepIO A = sweep_io(x, d);
  bool dxlag_fill = false;
  if (d == 0 && x->fuse[0]) {
    // x: RHS and MOC in one kernel, waves of 64 cells along x exchanging
    // neighbours by shuffles; the ghost cells both stages read are filled
    // first (rho, v(3), p, face B_x, species), as the two stages' fills do
    int l6[6] = {1, 2, 3, 4, 5, 6};
    {
      FillPlan P(2);
      planU(x, P, 1, 1, l6, 6);
      planOT(x, P, 1, OT_SPECIES);
      CHK(flush(x, P));
    }
    A = sweep_io(x, 0);
    XSide S{};
    const int bc[2] = {x->cfg.bcx1, x->cfg.bcx2};
    for (int s = 0; s < 2; ++s) {
      S.inflow[s] = bc[s] == 4;
      S.blg[s][0] = x->bbrv[7][1][s + 1];
      S.blg[s][1] = x->bbrv[8][1][s + 1];
      S.dxl[s] = x->hdx[0][s == 0 ? 1 : x->g.n[0] + 2];
      S.vin[s] = x->bbrv[2][1][s + 1];
      S.mt[s][0] = x->bbrv[3][2][s + 1];
      S.mt[s][1] = x->bbrv[4][2][s + 1];
      S.ein[s] = x->bbrv[5][1][s + 1];
    }
    for (int t = 0; t < 8; ++t) S.spin[t] = x->bbrv_cm[t];
    const int nx = x->g.n[0];
    if (S.inflow[0] && S.inflow[1]) {
      // inflow on both sides: the whole sweep in one kernel (the REMAP
      // stage's ghosts are the fill rules, applied in the kernel)
      // two rows (waves) per block: 2% faster than four, 6% faster than eight
      dim3 gs((unsigned)((nx + XSCELLS - 1) / XSCELLS), (unsigned)((x->g.n[1] + 1) / 2), (unsigned)x->g.n[2]);
      // reads rho, v(3), p, Bcc(3), face B_x, species(8); writes rho, v(3), p, species(8)
      const double bys = 8.0 * (double)nx * x->g.n[1] * x->g.n[2] * 30;
      if (x->mixed) KLAUNCH_N("k_sweep_x", k_sweep_x<float>, gs, dim3(128), bys, A, x->g, S, dt);
      else KLAUNCH_N("k_sweep_x", k_sweep_x<double>, gs, dim3(128), bys, A, x->g, S, dt);
      for (int f = 0; f < 5; ++f) swapU(x, f);
      swapSp(x);
      return MCF_OK;
    }
    dim3 gr((unsigned)((nx + XCELLS - 1) / XCELLS), (unsigned)((x->g.n[1] + 3) / 4), (unsigned)x->g.n[2]);
    const double by = 8.0 * (double)nx * x->g.n[1] * x->g.n[2] * 33;
    if (x->mixed) KLAUNCH_N("k_rhsmoc_x", k_rhsmoc_x<float>, gr, dim3(256), by, A, x->g, S, dt);
    else KLAUNCH_N("k_rhsmoc_x", k_rhsmoc_x<double>, gr, dim3(256), by, A, x->g, S, dt);
    for (int f = 0; f < 5; ++f) swapU(x, f);
    swapSp(x); swapBlg(x);
    dxlag_fill = true; // gathered with REMAP's fills below
  } else if (x->fuse[d] && x->g.n[d] <= SWEEP_COL_MAXN) {
    // y/z: the whole sweep in one column march over the periodic axis; it
    // reads only interior cells (periodic images), so no fills precede it.
    // One kernel instead of k_rhsmoc_col + k_remap_col (the MOC results no
    // longer go through HBM): 1.66 against 1.93 ms per sweep (mixed), 2.32
    // against 2.75 (FP64), at 1 wave per SIMD (354 / 450 registers); on the
    // mixed path split over two waves per column row (k_sweep_col2)
    const int n = x->g.n[d], ntr = x->g.n[d == 1 ? 2 : 1];
    const int seg = x->col_seg;
    dim3 gr((unsigned)((n + seg - 1) / seg), (unsigned)((x->g.n[0] + 63) / 64), (unsigned)((ntr + 3) / 4));
    // reads rho, v(3), p, Bcc(3), face B_n, species(8); writes rho, v(3), p, species(8)
    const double by = 8.0 * (double)x->g.n[0] * x->g.n[1] * x->g.n[2] * 30;
    if (x->mixed) {
      if (d == 1) KLAUNCH_N("k_sweep_col2<1>", (k_sweep_col2<1, float>), gr, dim3(512), by, A, x->g, dt, seg);
      else KLAUNCH_N("k_sweep_col2<2>", (k_sweep_col2<2, float>), gr, dim3(512), by, A, x->g, dt, seg);
    } else if ((long)gr.x * gr.y * gr.z < (long)x->ncu) {
      if (d == 1) KLAUNCH_N("k_sweep_col2<1>", (k_sweep_col2<1, double>), gr, dim3(512), by, A, x->g, dt, seg);
      else KLAUNCH_N("k_sweep_col2<2>", (k_sweep_col2<2, double>), gr, dim3(512), by, A, x->g, dt, seg);
    } else {
      if (d == 1) KLAUNCH_N("k_sweep_col<1>", (k_sweep_col<1, double>), gr, dim3(256), by, A, x->g, dt, seg);
      else KLAUNCH_N("k_sweep_col<2>", (k_sweep_col<2, double>), gr, dim3(256), by, A, x->g, dt, seg);
    }
    for (int f = 0; f < 5; ++f) swapU(x, f);
    swapSp(x);
    return MCF_OK;
  } else if (x->fuse[d]) {
    // y/z axes longer than k_sweep_col's LDS copy of the cell edges: RHS and
    // MOC in one column march over the periodic axis (it reads only interior
    // cells, periodic images, so no fills precede it), then REMAP
    const int n = x->g.n[d], ntr = x->g.n[d == 1 ? 2 : 1];
    dim3 gr((unsigned)((n + FUSE_SEG - 1) / FUSE_SEG), (unsigned)((x->g.n[0] + 63) / 64), (unsigned)((ntr + 3) / 4));
    // algorithmic bytes: 14 arrays read, 9 written (RHS+MOC); 18 read, 13
    // written (REMAP)
    const double cb = 8.0 * (double)x->g.n[0] * x->g.n[1] * x->g.n[2];
    const double by = cb * 23, byr = cb * 31;
    if (x->mixed) {
      if (d == 1) KLAUNCH_N("k_rhsmoc_col<1>", (k_rhsmoc_col<1, float>), gr, dim3(256), by, A, x->g, dt);
      else KLAUNCH_N("k_rhsmoc_col<2>", (k_rhsmoc_col<2, float>), gr, dim3(256), by, A, x->g, dt);
    } else {
      if (d == 1) KLAUNCH_N("k_rhsmoc_col<1>", (k_rhsmoc_col<1, double>), gr, dim3(256), by, A, x->g, dt);
      else KLAUNCH_N("k_rhsmoc_col<2>", (k_rhsmoc_col<2, double>), gr, dim3(256), by, A, x->g, dt);
    }
    for (int f = 0; f < 5; ++f) swapU(x, f);
    swapBlg(x);
    // REMAP as a column march over the same segments (periodic images, no
    // fills); it forms the Lagrangian species from the pre-sweep ones
    A = sweep_io(x, d);
    if (x->mixed) {
      if (d == 1) KLAUNCH_N("k_remap_col<1>", (k_remap_col<1, float>), gr, dim3(256), byr, A, x->g, dt);
      else KLAUNCH_N("k_remap_col<2>", (k_remap_col<2, float>), gr, dim3(256), byr, A, x->g, dt);
    } else {
      if (d == 1) KLAUNCH_N("k_remap_col<1>", (k_remap_col<1, double>), gr, dim3(256), byr, A, x->g, dt);
      else KLAUNCH_N("k_remap_col<2>", (k_remap_col<2, double>), gr, dim3(256), byr, A, x->g, dt);
    }
    for (int f = 0; f < 5; ++f) swapU(x, f);
    swapSp(x);
    return MCF_OK;
  } else {
    // RHS (fills M3:1346-1357): reads rho, v(3), p, Bcc(3), species; writes
    // rho, v_n, e, Blg(2), species, interface velocity / position, widths
    int lr[3] = {1, A.VN + 1, 5};
    {
      FillPlan P(2);
      planU(x, P, 1 << d, 1, lr, 3);
      planOT(x, P, 1 << d, OT_SPECIES);
      CHK(flush(x, P));
    }
    A = sweep_io(x, d);
    SWEEP_LAUNCH(k_rhs, d, true, 33, A, x->g, dt);
    swapU(x, 0); swapU(x, A.VN); swapSp(x); swapBlg(x);
    // MOC (fills M3:1950-1976): reads rho, v_t(2), Blg(2), p, e, widths, Bn;
    // writes momenta(3), E, Blg(2)
    int lm[4] = {1, A.VT1 + 1, A.VT2 + 1, A.BN + 1};
    {
      FillPlan P(2);
      planU(x, P, 1 << d, 1, lm, 4);
      planOT(x, P, 1 << d, OT_DXLAG);
      planOT(x, P, 1 << d, OT_BLG);
      CHK(flush(x, P));
    }
    A = sweep_io(x, d);
    SWEEP_LAUNCH(k_moc, d, true, 20, A, x->g, dt);
    swapU(x, A.VN); swapU(x, A.VT1); swapU(x, A.VT2); swapU(x, 4); swapBlg(x);
  }
  // REMAP (fills M3:1571-1594): reads rho, m(3), E, Blg(2), species,
  // widths, interface velocity, Bcc_n; writes rho, v(3), p, species
  int l5[5] = {1, 2, 3, 4, 5};
  {
    // mode 2's x rules read rho, Bcc_x and Blg on every row, so with x
    // inflow these fills run call by call; otherwise they are gathered
    FillPlan P(2);
    if (dxlag_fill) planOT(x, P, 1, OT_DXLAG);
    if (d == 0 && x_inflow(x)) {
      CHK(flush(x, P));
      CHK(fillOT(x, 1 << d, 2, OT_BLG));
      CHK(fillU(x, 1 << d, 2, 2, l5, 5));
      CHK(fillOT(x, 1 << d, 2, OT_SPECIES));
    } else {
      planOT(x, P, 1 << d, OT_BLG);
      planU(x, P, 1 << d, 2, l5, 5);
      planOT(x, P, 1 << d, OT_SPECIES);
      CHK(flush(x, P));
    }
  }
  A = sweep_io(x, d);
  if (d == 0 && x->fuse[0]) {
    dim3 gr((unsigned)((x->g.n[0] + XRCELLS - 1) / XRCELLS), (unsigned)((x->g.n[1] + 3) / 4), (unsigned)x->g.n[2]);
    const double by = 8.0 * (double)x->g.n[0] * x->g.n[1] * x->g.n[2] * 31;
    if (x->mixed) KLAUNCH_N("k_remap_x", k_remap_x<float>, gr, dim3(256), by, A, x->g, dt);
    else KLAUNCH_N("k_remap_x", k_remap_x<double>, gr, dim3(256), by, A, x->g, dt);
  } else {
    SWEEP_LAUNCH(k_remap, d, false, 31, A, x->g, dt);
  }
  for (int f = 0; f < 5; ++f) swapU(x, f);
  swapSp(x);
  return MCF_OK;
}

static CtArgs ct_args(mcf_ctx *x) {
  CtArgs A{};
  for (int f = 0; f < 8; ++f) A.U[f] = x->U[f];
  for (int c = 0; c < 3; ++c) { A.bcc[c] = x->bcc[c]; A.vfc[c] = x->vfc[c]; A.emf[c] = x->emf[c]; A.gdx[c] = x->gdx[c]; }
  A.facdep = x->cfg.facdep;
  return A;
}

static int do_cc4(mcf_ctx *x) {
  CtArgs A = ct_args(x);
  int nx = x->g.n[0], ny = x->g.n[1], nz = x->g.n[2];
  Box b = mkbox(1, nx, 1, ny, 1, nz);
  KLAUNCH512(k_cc4, boxsize(b), 6, A, x->g, b);
  return fillOT(x, 7, 2, OT_BCC);
}

// CT (M3:638-649)
static int do_ct(mcf_ctx *x, double dt) {
  StageScope scope(x, "ct");
  CtArgs A = ct_args(x);
  int n[3] = {x->g.n[0], x->g.n[1], x->g.n[2]};
  int lx[3] = {2, 7, 8}, ly[3] = {3, 6, 8}, lz[3] = {4, 6, 7}, l5[1] = {5}, l1[1] = {1};
  {
    FillPlan P(2);
    planU(x, P, 1, 1, lx, 3);
    planU(x, P, 2, 1, ly, 3);
    planU(x, P, 4, 1, lz, 3);
    CHK(flush(x, P));
  }
  Box b1 = mkbox(1, n[0] + 1, 1, n[1] + 1, 1, n[2] + 1);
  KLAUNCH512(k_cc1, boxsize(b1), 6, A, x->g, b1);
  {
    FillPlan P(2);
    planOT(x, P, 7, OT_VFC);
    planU(x, P, 7, 1, l5, 1);
    planU(x, P, 7, 1, l1, 1);
    CHK(flush(x, P));
  }
  // the three edge orientations in one launch over (1..n+1)^3: reads rho, p,
  // Vfc(3), face B(3), writes EMF(3)
  {
    // z march over 61-cell rows, CTM_ROWS rows and CTM_ZSEG planes per
    // block.  A variant with full 64-cell tiles (the tile's outer
    // neighbours loaded by the edge lanes, the 257th column per cell) was
    // slower: 1.08 / 2.11 ms against 1.00 / 1.77 (mixed / FP64), its extra
    // registers cost more than the idle lanes of the fifth tile.
    dim3 gr((unsigned)((n[0] + 1 + CTM_OUT - 1) / CTM_OUT), (unsigned)((n[1] + 1 + CTM_ROWS - 1) / CTM_ROWS),
            (unsigned)((n[2] + 1 + CTM_ZSEG - 1) / CTM_ZSEG));
    if (x->mixed) KLAUNCH_N("k_ct_march_mixed", k_ct_march_mixed, gr, dim3(64 * CTM_ROWS), 8.0 * boxsize(b1) * 11, A, x->g, dt);
    else KLAUNCH_N("k_ct_march", k_ct_march<double>, gr, dim3(64 * CTM_ROWS), 8.0 * boxsize(b1) * 11, A, x->g, dt);
  }
  KLAUNCH512(k_cc3, boxsize(b1), 9, A, x->g, b1, dt);
  return do_cc4(x);
}

static ChemCols chem_cols(const mcf_ctx *x) {
  ChemCols C{};
  for (int q = 0; q < 10; ++q) C.c[q] = x->col[q];
  return C;
}

// SHIELD (M3:2924-3036): one wave per (j,k) row
static int do_shield(mcf_ctx *x) {
  ShieldArgs S{};
  for (int s = 0; s < 8; ++s) S.sp[s] = x->sp[s];
  S.p = x->U[F_P];
  for (int q = 0; q < 10; ++q) S.col[q] = x->col[q];
  S.temp = x->ctemp;
  S.gdx0 = x->gdx[0];
  S.nsx = x->nsp[0];
  S.cx = x->g.n[0] / x->nsp[0];
  const long rows = (long)x->g.n[1] * x->g.n[2];
  const double cells = (double)rows * x->g.n[0];
  // reads 8 species, p, tCII(2) and the 8 scan inputs twice; writes temp
  // and 10 columns
  KLAUNCH_G(k_shield, dim3((unsigned)((rows + 3) / 4)), dim3(256), 8.0 * cells * 38, S, x->g);
  return MCF_OK;
}

static int do_thermal(mcf_ctx *x, double dt);

// SOURCE (M3:2448-2780): SHIELD (ifrad = 2), the thermal update (ifthrm =
// 2), the chemistry network (ifchem = 2)
static int do_source(mcf_ctx *x, double dt) {
  StageScope scope(x, "source");
  if (x->cfg.ifrad == 2) CHK(do_shield(x));
  if (x->cfg.ifthrm == 2) CHK(do_thermal(x, dt));
  if (x->cfg.ifchem == 2) {
    ChemArgs C{};
    for (int s = 0; s < 8; ++s) { C.sp[s] = x->sp[s]; C.spmin[s] = x->spmin[s]; }
    C.rho = x->U[F_RHO];
    C.p = x->U[F_P];
    C.cols = chem_cols(x);
    const long n = (long)x->g.n[0] * x->g.n[1] * x->g.n[2];
    KLAUNCH(k_chem, n, 26, C, x->g, dt);
  }
  return MCF_OK;
}

// The thermal part of SOURCE (M3:2477-2572): two-stage cooling +
// saturated conduction
static int do_thermal(mcf_ctx *x, double dt) {
  SrcArgs A{};
  A.rho = x->U[F_RHO];
  for (int s = 0; s < 8; ++s) A.sp[s] = x->sp[s];
  for (int d = 0; d < 3; ++d) A.gdx[d] = x->gdx[d];
  A.pmin = x->pmin; A.pmax = x->pmax;
  A.cnt = x->cool_cnt;
  A.queue = reinterpret_cast<int *>(x->srcq[1]);
  A.fqueue = reinterpret_cast<int *>(x->srcq[2]);
  A.qdc = x->srcq[0];
  A.cols = chem_cols(x);
  if (x->cfg.cooling == MCF_COOL_SURROGATE) {
    if (!x->mlp_ready) { g_err = "cooling surrogate selected but no weights loaded (mcf_mlp_set_weights / mcf_mlp_train)"; return MCF_E_STATE; }
    A.sur = 1;
    A.pk = x->mlp_pk;
    A.mw = x->mlp;
  }
  int nx = x->g.n[0], ny = x->g.n[1], nz = x->g.n[2];
  int l15[2] = {1, 5}, l5[1] = {5};
  {
    FillPlan P(1);
    planU(x, P, 7, 1, l15, 2);
    planOT(x, P, 7, OT_SPECIES);
    CHK(flush(x, P));
  }
  Box bc = mkbox(1, nx, 1, ny, 1, nz);
  double *P = x->U[F_P], *Q = x->Ub[F_P];
  for (int stage = 0; stage < 2; ++stage) {
    if (stage == 0) {
      // predictor: T and P_old from P, result into the second pressure buffer
      A.Tp = P; A.Pn = P; A.Pout = Q;
    } else {
      // ghosts of the predictor (the fill works on U[F_P])
      x->U[F_P] = Q; x->Ub[F_P] = P;
      CHK(fillU(x, 7, 1, 1, l5, 1));
      x->U[F_P] = P; x->Ub[F_P] = Q;
      // corrector: T from the predictor, result written over P; P's first
      // ghost layer takes the predictor's
      A.Tp = Q; A.Pn = P; A.Pout = P;
      const long a0 = nx + 2, a1 = ny + 2;
      KLAUNCH(k_shell_copy, 2 * (a0 * a1 + a0 * nz + (long)ny * nz), 2, Q, P, x->g);
    }
    const Box &b = bc;
    // queue counters: cool_cnt[2] and cool_cnt[3] (two u32 each), zero at
    // init; the stage counts in one pair and its k_src_col clears the other
    A.qn = reinterpret_cast<unsigned *>(x->cool_cnt + 2 + x->src_pair);
    A.qz = reinterpret_cast<unsigned *>(x->cool_cnt + 3 - x->src_pair);
    {
      // the chemistry Fcool always runs in FP64
      const bool f32 = x->mixed && !x->chem;
      const int zseg = f32 ? SRC_ZSEG_F32 : SRC_ZSEG_F64;
      dim3 gr((unsigned)((b.cnt[0] + 63) / 64), (unsigned)((b.cnt[1] + SRC_ROWS - 1) / SRC_ROWS), (unsigned)((b.cnt[2] + zseg - 1) / zseg));
      if (x->chem) KLAUNCH_N("k_src_col", (k_src_col<double, true>), gr, dim3(64 * SRC_ROWS), 8.0 * (double)boxsize(b) * 19, A, x->g, b, dt, stage);
      else if (f32) KLAUNCH_N("k_src_col", (k_src_col<float, false>), gr, dim3(64 * SRC_ROWS), 8.0 * (double)boxsize(b) * 9, A, x->g, b, dt, stage);
      else KLAUNCH_N("k_src_col", (k_src_col<double, false>), gr, dim3(64 * SRC_ROWS), 8.0 * (double)boxsize(b) * 9, A, x->g, b, dt, stage);
    }
    x->src_pair ^= 1;
    if (A.sur) {
      KLAUNCH_G(k_src_stiff, dim3(1024), dim3(256), 0, A, dt, stage);
      KLAUNCH_N("k_src_imc", k_src_imc<false>, dim3(256), dim3(256), 0, A, dt, stage, 1);
    } else if (x->chem) {
      KLAUNCH_N("k_src_imc", k_src_imc<true>, dim3(256), dim3(256), 0, A, dt, stage, 0);
    } else {
      KLAUNCH_N("k_src_imc", k_src_imc<false>, dim3(256), dim3(256), 0, A, dt, stage, 0);
    }
  }
  return MCF_OK;
}

static int do_dissip(mcf_ctx *x) {
  StageScope scope(x, "dissip");
  int l[3] = {2, 3, 4};
  CHK(fillU(x, 7, 2, 1, l, 3));
  DisArgs A{};
  for (int f = 0; f < 8; ++f) A.U[f] = x->U[f];
  A.divv = x->divv;
  int e0 = x->nsp[0] == 1 ? 1 : 0; // one x block: only the right end (M3:3156-3158)
  int ny = x->g.n[1], nz = x->g.n[2];
  KLAUNCH(k_dissip_div, (2 - e0) * 12L * (ny + 2) * (nz + 2), 4, A, x->g, e0);
  KLAUNCH(k_dissip_upd, (2 - e0) * 10L * ny * nz, 6, A, x->g, e0);
  return MCF_OK;
}

// ----------------------------------------------------------------------------
// self-gravity (GRAVTY, 3DMHD-CfCA/selfG_MPIfull8.f90)
// ----------------------------------------------------------------------------
static int ilog2i(int v) {
  int l = 0;
  while ((1 << (l + 1)) <= v) ++l;
  return l;
}

// The grid GRAVTY takes: a cube of 2^NGL cells per axis, power-of-two
// splits, NGcr = log2(max split) + 1 below NGL; here also NGcr <= 4 (the
// whole-domain levels live in one workgroup's LDS) and n <= 1024 (FFT
// lines in LDS).
static const char *grav_check(const int n[3], const int nsp[3]) {
  if (n[0] != n[1] || n[0] != n[2]) return "self-gravity needs the same cell count on every axis";
  if (n[0] < 8 || n[0] > 1024 || (n[0] & (n[0] - 1))) return "self-gravity needs 8..1024 cells per axis, a power of two";
  int ms = 1;
  for (int d = 0; d < 3; ++d) {
    if (nsp[d] & (nsp[d] - 1)) return "self-gravity needs power-of-two block splits";
    if (n[d] / nsp[d] < 2) return "self-gravity needs at least 2 cells per block";
    ms = std::max(ms, nsp[d]);
  }
  const int NGL = ilog2i(n[0]), NGcr = ilog2i(ms) + 1;
  if (NGcr >= NGL) return "self-gravity needs more cells per block (NGcr < NGL)";
  if (NGcr > GRAV_MAX_NGCR) return "self-gravity takes at most 8 blocks along an axis";
  return nullptr;
}

// GravGeo and the device arrays (pinter SG8:102-146, PB's pointers
// SG8:1058-1080), the x planes PB transforms and the rlft3 twiddles
static int grav_setup(mcf_ctx *x) {
  GravGeo &G = x->gg;
  const mcf_config &cfg = x->cfg;
  for (int d = 0; d < 3; ++d) { G.S[d] = x->nsp[d]; G.C[d] = x->g.n[d] / x->nsp[d]; }
  G.npe = G.S[0] * G.S[1] * G.S[2];
  G.NGL = ilog2i(x->g.n[0]);
  G.NGcr = ilog2i(std::max(G.S[0], std::max(G.S[1], G.S[2]))) + 1;
  G.Lbox = cfg.ql1x + cfg.ql2x;
  long tot2 = 0, totb = 0;
  for (int l = G.NGL; l >= G.NGcr; --l) {
    for (int d = 0; d < 3; ++d) G.n[l][d] = l == G.NGL ? G.C[d] + 1 : G.n[l + 1][d] / 2 + 1;
  }
  for (int l = G.NGcr; l <= G.NGL; ++l) {
    // skewed block layout (k_grav_relax): (n1+1) x (n0+n1+n2+1) x (n2+1) slots
    G.bsz[l] = (long)(G.n[l][1] + 1) * (G.n[l][0] + G.n[l][1] + G.n[l][2] + 1) * (G.n[l][2] + 1);
    G.off[l] = tot2;
    tot2 += G.bsz[l] * G.npe;
    G.boff[l] = totb;
    totb += 2L * G.npe * (G.n[l][1] + 1) * (G.n[l][2] + 1);
  }
  G.p1[1] = 1;
  for (int l = 1; l <= G.NGcr; ++l) { const long m = (1L << l) + 1; G.p1[l + 1] = G.p1[l] + m * m * m; }
  G.need1 = G.p1[G.NGcr + 1];
  G.pb1[1] = 1;
  for (int l = 1; l <= G.NGcr; ++l) { const long m = (1L << l) + 1; G.pb1[l + 1] = G.pb1[l] + m * m; }
  G.needb1 = G.pb1[G.NGcr + 1];
  if (G.need1 > GRAV_NEED1_MAX) { g_err = "self-gravity: whole-domain levels exceed the LDS arrays"; return MCF_E_ARG; }
  // PB's planes: x planes 1..min(C0, 3P) of every x block (SG8:1106-1125)
  const int P = G.S[1] * G.S[2], C0 = G.C[0];
  std::vector<int> gxs;
  std::vector<double> zp, zq, fg;
  for (int ist = 0; ist < G.S[0]; ++ist) {
    const double dzz = x->hdx[0][ist * C0 + 1 + 1];
    for (int klr = 1; klr <= std::min(C0, 3 * P); ++klr) {
      const int gx = ist * C0 + klr;
      const double z = x->hx[0][gx + 1] - 0.5 * dzz;
      gxs.push_back(gx);
      zp.push_back(z);
      zq.push_back(G.Lbox - z);
      fg.push_back(-GRAV_G4PI * dzz);
    }
  }
  const int npl = (int)gxs.size(), nn1 = x->g.n[1], nn2 = x->g.n[2], H = nn1 / 2, q4 = nn1 / 4;
  // rlft3's unpacking twiddles: the recurrence in double, each value
  // rounded to single precision by CMPLX (SG8:1476-1508); index i1 >= 2
  std::vector<double2> tw[2];
  for (int s = 0; s < 2; ++s) {
    const int isign = s == 0 ? 1 : -1;
    const double theta = 6.28318530717959 / (double)(isign * nn1);
    const double sh = std::sin(0.5 * theta);
    const double wpr = -(2.0 * (sh * sh)), wpi = std::sin(theta);
    double wr = 1.0, wi = 0.0;
    tw[s].assign(q4 + 2, make_double2(0.0, 0.0));
    for (int i1 = 1; i1 <= q4; ++i1) {
      const double wtemp = wr;
      wr = wr * wpr - wi * wpi + wr;
      wi = wi * wpr + wtemp * wpi + wi;
      tw[s][i1 + 1] = make_double2((double)(float)wr, (double)(float)wi);
    }
  }
  // every segment starts at an even double (16-byte alignment for double2)
  auto ev = [](long c) { return (c + 1) & ~1L; };
  const long n1 = ev(G.need1 + 1), nb1 = ev(2 * (G.needb1 + 1));
  const long nz = 4 * ev(tot2) + 4 * n1 + nb1 + ev(totb); // cleared at every solve
  const long nZ = 2L * npl * nn2 * H, nS = 2L * npl * nn2, nD = 2L * 2 * nn2 * H, nSD = 2L * 2 * nn2;
  const long ntab = 4L * (q4 + 2);
  // boundary planes of consecutive sweeps: 6 per rank block, two sets
  long nex = 0;
  for (int l = G.NGcr; l <= G.NGL; ++l) {
    const int *m = G.n[l];
    const long ep = std::max((long)(m[1] + 1) * (m[2] + 1), std::max((long)(m[0] + 1) * (m[2] + 1), (long)(m[0] + 1) * (m[1] + 1)));
    nex = std::max(nex, 6L * G.npe * ep);
  }
  const long total = nz + ev((long)x->ncell) + nZ + nS + nD + nSD + ntab + 4 * ev(npl) + 2 * ev(nex) + 16;
  HIPCHK(hipMalloc(&x->gpool, total * sizeof(double)));
  double *p = x->gpool;
  auto take = [&](long c) { double *r = p; p += ev(c); return r; };
  x->gb.phi2 = take(tot2);
  x->gb.rho2 = take(tot2);
  x->gb.res2 = take(tot2);
  x->gb.rhs2 = take(tot2);
  x->gb.phi1 = take(n1);
  x->gb.rho1 = take(n1);
  x->gb.res1 = take(n1);
  x->gb.rhs1 = take(n1);
  x->gb.bphi1 = take(nb1);
  x->gb.bphi2 = take(totb);
  x->gpool_n = nz;
  x->phi = take((long)x->ncell);
  x->pbZ = (double2 *)take(nZ);
  x->pbS = (double2 *)take(nS);
  x->pbD = (double2 *)take(nD);
  x->pbSD = (double2 *)take(nSD);
  double2 *wf = (double2 *)take(2L * (q4 + 2));
  double2 *wi = (double2 *)take(2L * (q4 + 2));
  double *dzp = take(npl);
  double *dzq = take(npl);
  double *dfg = take(npl);
  int *dgx = (int *)take(npl);
  x->gb.ex[0] = take(nex);
  x->gb.ex[1] = take(nex);
  if (p > x->gpool + total) { g_err = "self-gravity: pool layout"; return MCF_E_ARG; }
  HIPCHK(hipMemsetAsync(x->gpool, 0, total * sizeof(double), x->stream));
  HIPCHK(hipMemcpyAsync(wf, tw[0].data(), (q4 + 2) * sizeof(double2), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(wi, tw[1].data(), (q4 + 2) * sizeof(double2), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(dzp, zp.data(), npl * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(dzq, zq.data(), npl * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(dfg, fg.data(), npl * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(dgx, gxs.data(), npl * sizeof(int), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  PbArgs &B = x->pb;
  B.np = npl;
  B.nn1 = nn1;
  B.nn2 = nn2;
  B.gx = dgx;
  B.zp = dzp;
  B.zq = dzq;
  B.facG = dfg;
  B.dxx = x->hdx[1][1 + 1];
  B.dyy = x->hdx[2][1 + 1];
  B.wf = wf;
  B.wi = wi;
  x->grav = true;
  return MCF_OK;
}

static unsigned gblocks(long n) { return (unsigned)((n + TPB - 1) / TPB); }

// per-level names of the exchange launches (mcf_profile)
static const char *const k_exchange_names[GRAV_MAXL] = {
    "k_grav_exchange<0>", "k_grav_exchange<1>", "k_grav_exchange<2>", "k_grav_exchange<3>",
    "k_grav_exchange<4>", "k_grav_exchange<5>", "k_grav_exchange<6>", "k_grav_exchange<7>",
    "k_grav_exchange<8>", "k_grav_exchange<9>", "k_grav_exchange<10>", "k_grav_exchange<11>",
    "k_grav_exchange<12>", "k_grav_exchange<13>", "k_grav_exchange<14>", "k_grav_exchange<15>"};

// one exchange of split level l (flags: x low, x high, y low, y high, z low, z high)
static int grav_exchange(mcf_ctx *x, double *u2, int l, int flags) {
  const GravGeo &G = x->gg;
  const int *n = G.n[l];
  const long per = grav_nghost(n);
  KLAUNCH_N(k_exchange_names[l], k_grav_exchange, dim3(gblocks(per * G.npe)), dim3(TPB), 16.0 * per * G.npe, G, u2, l, flags);
  return MCF_OK;
}
static const int XALL = 63, XLOW = 1 | 4 | 16;

// per-level names of the relaxation launches (mcf_profile)
static const char *const k_relax_names[GRAV_MAXL] = {
    "k_grav_relax<0>", "k_grav_relax<1>", "k_grav_relax<2>", "k_grav_relax<3>", "k_grav_relax<4>", "k_grav_relax<5>",
    "k_grav_relax<6>", "k_grav_relax<7>", "k_grav_relax<8>", "k_grav_relax<9>", "k_grav_relax<10>", "k_grav_relax<11>",
    "k_grav_relax<12>", "k_grav_relax<13>", "k_grav_relax<14>", "k_grav_relax<15>"};

// Gauss-Seidel sweeps at split level l: the column kernel (k_grav_relax_col)
// when the block's (j,k) columns fit 1 or 4 per thread of 1024 and its
// y-z plane and ghost faces fit the kernel's LDS buffers, else the
// plane-by-plane kernel
static int grav_relax(mcf_ctx *x, int l, int times) {
  const GravGeo &G = x->gg;
  const int *n = G.n[l];
  const long cols = (long)(n[1] - 1) * (n[2] - 1);
  const long plane = std::max((long)(n[1] + 1) * (n[2] + 1),
                              (long)(n[0] + 1) * std::max(n[1] + 1, n[2] + 1)); // the LDS planes and ghost faces
  const double bytes = 8.0 * 8.0 * (double)(n[0] - 1) * cols * G.npe;
  const unsigned tb = (unsigned)std::min<long>(GRAV_RELAX_THREADS, std::max<long>(64, (cols + 63) / 64 * 64));
  const long ep = std::max(plane, std::max((long)(n[0] + 1) * (n[2] + 1), (long)(n[0] + 1) * (n[1] + 1)));
  const int kind = (cols <= GRAV_RELAX_THREADS && plane <= 1156) ? 1
                   : (cols <= 4 * GRAV_RELAX_THREADS && plane <= 4356) ? 4 : 0;
  if (kind == 0) {
    for (int t = 0; t < times; ++t) {
      KLAUNCH_N(k_relax_names[l], k_grav_relax, dim3(G.npe), dim3(GRAV_RELAX_THREADS), bytes, G, x->gb.phi2,
                x->gb.rhs2, l);
      CHK(grav_exchange(x, x->gb.phi2, l, XALL));
    }
    return MCF_OK;
  }
  // Blocks of at most 9 nodes per side pass the halo between the sweeps of
  // the sequence through boundary planes and exchange once after the last
  // one; larger blocks exchange after every sweep (256^3, 8x8x8: the planes
  // saved 9 us of launch per sweep at 5^3-node blocks but cost 130 us per
  // sweep at 33^3).
  const bool planes = std::max(n[0], std::max(n[1], n[2])) <= 9;
  for (int t = 0; t < times; ++t) {
    const double *ein = planes && t > 0 ? x->gb.ex[(t - 1) & 1] : nullptr;
    double *eout = planes && t < times - 1 ? x->gb.ex[t & 1] : nullptr;
    if (!planes && t > 0) CHK(grav_exchange(x, x->gb.phi2, l, XALL));
    if (kind == 1 && planes)
      KLAUNCH_N(k_relax_names[l], (k_grav_relax_col<1, 1156, GRAV_RELAX_D1, true>), dim3(G.npe), dim3(tb), bytes, G,
                x->gb.phi2, x->gb.rhs2, l, ein, eout, ep);
    else if (kind == 1)
      KLAUNCH_N(k_relax_names[l], (k_grav_relax_col<1, 1156, GRAV_RELAX_D1, false>), dim3(G.npe), dim3(tb), bytes, G,
                x->gb.phi2, x->gb.rhs2, l, nullptr, nullptr, ep);
    else
      KLAUNCH_N(k_relax_names[l], (k_grav_relax_col<4, 4356, GRAV_RELAX_D4, true>), dim3(G.npe),
                dim3(GRAV_RELAX_THREADS), bytes, G, x->gb.phi2, x->gb.rhs2, l, nullptr, nullptr, ep);
  }
  return grav_exchange(x, x->gb.phi2, l, XALL);
}

static int grav_zero(mcf_ctx *x, double *a, long n) {
  if (n > 0) KLAUNCH_N("k_grav_zero", k_grav_zero, dim3((unsigned)std::min<long>(gblocks(n), 4096)), dim3(TPB), 8.0 * n, a, n);
  return MCF_OK;
}

// GRAVTY(dt,2) (SG8:15-79): density ghosts (BC_MPI(1,1), PB's BC_MPI(2,1)),
// PB, mglin, Phi with its face ghosts
static int do_grav_solve(mcf_ctx *x) {
  StageScope scope(x, "grav_solve");
  const GravGeo &G = x->gg;
  const GravBuf &B = x->gb;
  const PbArgs &P = x->pb;
  static const int l1[1] = {F_RHO};
  CHK(fillU(x, 7, 1, 1, l1, 1));
  CHK(fillU(x, 7, 2, 1, l1, 1));
  // every array starts the solve at zero, as in the reference's fixture
  // build; the finest split level needs no clearing: what a solve reads
  // there before writing it (node 0 of the first x rank) is never written
  for (double *a : {B.phi2, B.rho2, B.res2, B.rhs2}) CHK(grav_zero(x, a, G.off[G.NGL]));
  CHK(grav_zero(x, B.phi1, x->gpool_n - (B.phi1 - x->gpool)));
  // PB
  const int nn1 = P.nn1, nn2 = P.nn2, H = nn1 / 2, NGL = G.NGL, NGcr = G.NGcr;
  (void)nn1;
  const long nZ = (long)P.np * nn2 * H;
  KLAUNCH_N("k_pb_pack", k_pb_pack, dim3((unsigned)((P.np + 63) / 64), (unsigned)((H + 31) / 32), (unsigned)nn2),
            dim3(256), 16.0 * nZ, P, x->U[F_RHO], x->g, x->pbZ);
  // rows: one line per workgroup; columns: 8 neighbouring columns per
  // workgroup where H allows, so that each load covers 128 bytes
  const int lcw = (H % 8 == 0 && 8 * nn2 <= 2 * GRAV_FFT_MAXL) ? 3 : 0;
  const int t1 = std::min(512, std::max(64, H / 2)), t2 = std::min(512, std::max(64, (nn2 << lcw) / 2));
  KLAUNCH_N("k_pb_fft", k_pb_fft, dim3((unsigned)((long)P.np * nn2)), dim3(t1), 32.0 * nZ, x->pbZ, H, 1L, (long)H,
            (long)nn2 * H, nn2, 1, 0);
  KLAUNCH_N("k_pb_fft", k_pb_fft, dim3((unsigned)(((long)P.np * H) >> lcw)), dim3(t2), 32.0 * nZ, x->pbZ, nn2, (long)H,
            1L, (long)nn2 * H, H, 1, lcw);
  const long nu = 2L * (nn2 / 2 + 1) + (long)(nn1 / 4 - 1) * nn2;
  KLAUNCH_N("k_pb_unpack", k_pb_unpack, dim3(gblocks(nu * P.np)), dim3(TPB), 32.0 * nZ, P, x->pbZ, x->pbS, P.np, 1);
  KLAUNCH_N("k_pb_weight", k_pb_weight, dim3(gblocks(64L * (H + 1) * nn2)), dim3(TPB), 16.0 * nZ, P, x->pbZ, x->pbS,
            x->pbD, x->pbSD);
  KLAUNCH_N("k_pb_unpack", k_pb_unpack, dim3(gblocks(nu * 2)), dim3(TPB), 64.0 * nn2 * H, P, x->pbD, x->pbSD, 2, -1);
  KLAUNCH_N("k_pb_fft", k_pb_fft, dim3((unsigned)(2L * nn2)), dim3(t1), 64.0 * nn2 * H, x->pbD, H, 1L, (long)H,
            (long)nn2 * H, nn2, -1, 0);
  KLAUNCH_N("k_pb_fft", k_pb_fft, dim3((unsigned)((2L * H) >> lcw)), dim3(t2), 64.0 * nn2 * H, x->pbD, nn2, (long)H,
            1L, (long)nn2 * H, H, -1, lcw);
  {
    const long nb = 2L * G.npe * (G.C[1] + 2) * (G.C[2] + 2);
    KLAUNCH_N("k_pb_bphi", k_pb_bphi, dim3(gblocks(nb)), dim3(TPB), 16.0 * nb, G, P, x->pbD, B.bphi2);
    for (int lc = NGL - 1; lc >= NGcr; --lc) {
      const long nr = 2L * G.npe * G.n[lc][1] * G.n[lc][2];
      KLAUNCH_N("k_pb_restrict", k_pb_restrict, dim3(gblocks(nr)), dim3(TPB), 48.0 * nr, G, B.bphi2, lc);
    }
    const long N1 = (1L << NGcr) + 1;
    KLAUNCH_N("k_pb_collect", k_pb_collect, dim3(gblocks(2 * N1 * N1)), dim3(TPB), 16.0 * N1 * N1, G, B.bphi1, B.bphi2);
  }
  // mglin (SG8:151-295)
  {
    const long nr = (long)(G.n[NGL][0] + 1) * (G.n[NGL][1] + 1) * (G.n[NGL][2] + 1) * G.npe;
    KLAUNCH_N("k_grav_rhs", k_grav_rhs, dim3(gblocks(nr)), dim3(TPB), 16.0 * nr, G, x->U[F_RHO], x->g, B.rhs2);
  }
  auto rstrct = [&](double *uc, const double *uf, int lc, int mode) -> int {
    const long nr = (long)G.n[lc][0] * G.n[lc][1] * G.n[lc][2] * G.npe;
    KLAUNCH_N("k_grav_rstrct", k_grav_rstrct, dim3(gblocks(nr)), dim3(TPB), 64.0 * nr, G, uc, uf, lc, mode);
    return grav_exchange(x, uc, lc, XALL);
  };
  auto collect = [&](double *u1, const double *u2) {
    const long N1 = (1L << NGcr) + 1;
    KLAUNCH_N("k_grav_collect", k_grav_collect, dim3(gblocks(N1 * N1 * N1)), dim3(TPB), 16.0 * N1 * N1 * N1, G, u1, u2);
    return MCF_OK;
  };
  auto coarse = [&](int prog, int j) {
    KLAUNCH_N("k_grav_coarse", k_grav_coarse, dim3(1), dim3(GRAV_COARSE_THREADS), 48.0 * G.need1, G, B, prog, j);
    return MCF_OK;
  };
  auto divide = [&]() {
    const long nr = (long)(G.n[NGcr][0] + 1) * (G.n[NGcr][1] + 1) * (G.n[NGcr][2] + 1) * G.npe;
    KLAUNCH_N("k_grav_divide", k_grav_divide, dim3(gblocks(nr)), dim3(TPB), 16.0 * nr, G, B.phi2, B.phi1);
    return MCF_OK;
  };
  auto copy_level = [&](double *dst, const double *src, int l) {
    const long nr = G.bsz[l] * G.npe;
    KLAUNCH_N("k_copy", k_copy, dim3(gblocks(nr)), dim3(TPB), 16.0 * nr, src + G.off[l], dst + G.off[l], nr);
    return MCF_OK;
  };
  auto prolong = [&](double *uf, const double *uc, int lf, int mode) -> int {
    const long nr = (long)G.n[lf][0] * G.n[lf][1] * G.n[lf][2] * G.npe;
    KLAUNCH_N("k_grav_prolong", k_grav_prolong, dim3(gblocks(nr)), dim3(TPB), 16.0 * nr, G, uf, uc, lf, mode, B.bphi2);
    return grav_exchange(x, uf, lf, XLOW);
  };
  CHK(rstrct(B.rho2, B.rhs2, NGL - 1, 0));
  for (int l = NGL - 2; l >= NGcr; --l) CHK(rstrct(B.rho2, B.rho2, l, 0));
  CHK(collect(B.rho1, B.rho2));
  CHK(coarse(0, 0));
  for (int j = 2; j <= NGL; ++j) {
    if (j < NGcr) {
      CHK(coarse(1, j));
      continue;
    }
    if (j == NGcr) {
      CHK(coarse(2, j));
      CHK(divide());
      CHK(copy_level(B.rhs2, B.rho2, NGcr));
    } else {
      CHK(prolong(B.phi2, B.phi2, j, 1));
      if (j != NGL) CHK(copy_level(B.rhs2, B.rho2, j));
    }
    for (int cyc = 0; cyc < 2; ++cyc) {
      for (int jj = j; jj >= NGcr; --jj) {
        CHK(grav_relax(x, jj, j >= NGL ? 2 : 5));
        {
          const int *m = G.n[jj];
          const long nr = (long)m[0] * m[1] * m[2] * G.npe;
          KLAUNCH_N("k_grav_resid", k_grav_resid, dim3(gblocks(nr)), dim3(TPB), 24.0 * nr, G, B.res2, B.phi2,
                    B.rhs2, jj);
          CHK(grav_exchange(x, B.res2, jj, XALL));
        }
        if (jj == NGcr) {
          CHK(collect(B.phi1, B.phi2));
          CHK(collect(B.res1, B.res2));
        } else {
          CHK(rstrct(B.rhs2, B.res2, jj - 1, 1));
          CHK(grav_zero(x, B.phi2 + G.off[jj - 1], G.bsz[jj - 1] * G.npe));
        }
      }
      CHK(coarse(3, j));
      CHK(divide());
      CHK(grav_relax(x, NGcr, 5));
      for (int jj = NGcr + 1; jj <= j; ++jj) {
        CHK(prolong(B.res2, B.phi2, jj, 0));
        const long nr = G.bsz[jj] * G.npe;
        KLAUNCH_N("k_grav_add", k_grav_add, dim3(gblocks(nr)), dim3(TPB), 24.0 * nr, G, B.phi2, B.res2, jj);
        CHK(grav_relax(x, jj, j >= NGL ? 2 : 5));
      }
    }
  }
  const long nc = (long)x->g.n[0] * x->g.n[1] * x->g.n[2];
  KLAUNCH_N("k_grav_phi", k_grav_phi, dim3(gblocks(nc)), dim3(TPB), 16.0 * nc, G, B.phi2, x->phi, x->g);
  const long ng = (long)x->g.n[1] * x->g.n[2] + (long)x->g.n[0] * x->g.n[2] + (long)x->g.n[0] * x->g.n[1];
  KLAUNCH_N("k_grav_phi_ghosts", k_grav_phi_ghosts, dim3(gblocks(ng)), dim3(TPB), 64.0 * ng, x->phi, x->g);
  return MCF_OK;
}

// GRAVTY(dt,3) (SG8:82-89)
static int do_grav_kick(mcf_ctx *x, double dt) {
  StageScope scope(x, "grav_kick");
  const long nc = (long)x->g.n[0] * x->g.n[1] * x->g.n[2];
  KLAUNCH_N("k_grav_kick", k_grav_kick, dim3(gblocks(nc)), dim3(TPB), 56.0 * nc, x->U[F_VX], x->U[F_VY], x->U[F_VZ],
            x->phi, x->g, x->gdx[0], x->gg.C[0], dt);
  return MCF_OK;
}

// ----------------------------------------------------------------------------
// C ABI
// ----------------------------------------------------------------------------
extern "C" int mcf_init(const mcf_config *cfg, mcf_ctx **out) {
  if (!cfg || !out) { g_err = "null argument"; return MCF_E_ARG; }
  *out = nullptr;
  int n[3] = {cfg->np1x + cfg->np2x, cfg->np1y + cfg->np2y, cfg->np1z + cfg->np2z};
  int nsp[3] = {cfg->nsplitx, cfg->nsplity, cfg->nsplitz};
  for (int d = 0; d < 3; ++d) {
    if (n[d] < 4) { g_err = "grid must have at least 4 cells per axis"; return MCF_E_ARG; }
    if (nsp[d] < 1 || n[d] % nsp[d]) { g_err = "nsplit must divide the cell count"; return MCF_E_ARG; }
  }
  if (cfg->np1x < 1 || cfg->np2x < 0 || cfg->np1y < 1 || cfg->np1z < 1) { g_err = "bad cell counts"; return MCF_E_ARG; }
  if ((cfg->bcx1 != 1 && cfg->bcx1 != 3 && cfg->bcx1 != 4) || (cfg->bcx2 != 1 && cfg->bcx2 != 3 && cfg->bcx2 != 4)) {
    g_err = "x boundaries must be 1 (outflow), 3 (periodic) or 4 (inflow)"; return MCF_E_ARG;
  }
  // an inflow end opposite a periodic end is no configuration of the
  // reference (INPUT3D.DAT: 4,4) and its parity is not established
  if (cfg->bcx1 != cfg->bcx2) {
    g_err = "x boundaries must be alike: both outflow (1), both periodic (3) or both inflow (4)"; return MCF_E_ARG;
  }
  // y and z are always periodic in the reference (BC_MPI, M3:888-926)
  if (cfg->bcy1 != 3 || cfg->bcy2 != 3 || cfg->bcz1 != 3 || cfg->bcz2 != 3) {
    g_err = "y and z boundaries must be 3 (periodic)"; return MCF_E_ARG;
  }
  // DISSIP works on 10-cell layers at the x ends (M3:3152-3170)
  if (n[0] / nsp[0] < 10) { g_err = "x block must hold at least 10 cells (DISSIP layers)"; return MCF_E_ARG; }
  if (cfg->precision != MCF_PREC_FP64 && cfg->precision != MCF_PREC_MIXED) { g_err = "unknown precision mode"; return MCF_E_ARG; }
  if (cfg->cooling != MCF_COOL_KI && cfg->cooling != MCF_COOL_SURROGATE && cfg->cooling != MCF_COOL_CHEM) {
    g_err = "unknown cooling mode"; return MCF_E_ARG;
  }
  if (cfg->ifgrv == 2) {
    const char *why = grav_check(n, nsp);
    if (why) { g_err = why; return MCF_E_ARG; }
  }
  if (cfg->ifrad == 2 && nsp[0] > SHIELD_MAXB) { g_err = "SHIELD takes at most 64 x blocks"; return MCF_E_ARG; }
  long long vol = (long long)(n[0] + 4) * (n[1] + 4) * (n[2] + 4);
  if (vol >= (1LL << 31)) { g_err = "grid too large for 32-bit indexing"; return MCF_E_ARG; }
  mcf_ctx *x = new mcf_ctx();
  x->cfg = *cfg;
  x->mixed = cfg->precision == MCF_PREC_MIXED;
  x->chem = cfg->cooling == MCF_COOL_CHEM;
  x->has_cols = x->chem || cfg->ifchem == 2 || cfg->ifrad == 2;
  for (int d = 0; d < 3; ++d) { x->g.n[d] = n[d]; x->nsp[d] = nsp[d]; }
  x->g.st[0] = 1;
  x->g.st[1] = n[0] + 4;
  x->g.st[2] = (n[0] + 4) * (n[1] + 4);
  x->ncell = (size_t)vol;
  setup_constants(x);
  hipError_t e = hipGetDevice(&x->dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&x->ncu, hipDeviceAttributeMultiprocessorCount, x->dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { g_err = std::string("HIP init: ") + hipGetErrorString(e); delete x; return MCF_E_HIP; }
  {
    const int r = noalloc_scratch(x->dev, &x->scratch);
    if (r != MCF_OK) { (void)hipStreamDestroy(x->stream); delete x; return r; }
  }
  const int nstate = 19, nbuf = 13, nsweep = 8, nscratch = 9;
  size_t nc = x->ncell;
  size_t ndivv = 2 * 12 * (size_t)(n[1] + 2) * (n[2] + 2);
  size_t ngrid = 0;
  for (int d = 0; d < 3; ++d) ngrid += 2 * (size_t)(n[d] + 4);
  const int nchem = x->has_cols ? 11 : 0;
  size_t total = (nstate + nbuf + nsweep + nscratch + nchem) * nc + ndivv + ngrid + 16 + 2 * DT_BLOCKS;
  e = hipMalloc(&x->pool, total * sizeof(double));
  if (e != hipSuccess) { g_err = std::string("hipMalloc: ") + hipGetErrorString(e); (void)hipStreamDestroy(x->stream); delete x; return MCF_E_NOMEM; }
  double *p = x->pool;
  for (int f = 0; f < 8; ++f) { x->U[f] = p; p += nc; }
  for (int f = 0; f < 8; ++f) { x->sp[f] = p; p += nc; }
  for (int f = 0; f < 3; ++f) { x->bcc[f] = p; p += nc; }
  for (int f = 0; f < 5; ++f) { x->Ub[f] = p; p += nc; }
  for (int f = 0; f < 8; ++f) { x->spb[f] = p; p += nc; }
  x->u5e = p; p += nc;
  x->blg[0] = p; p += nc; x->blg[1] = p; p += nc; x->blgb[0] = p; p += nc; x->blgb[1] = p; p += nc;
  x->xlag = p; p += nc; x->dxlagM = p; p += nc; x->va = p; p += nc;
  for (int c = 0; c < 3; ++c) { x->vfc[c] = p; p += nc; }
  for (int t = 0; t < 3; ++t) { x->srcq[t] = p; p += nc; }
  for (int c = 0; c < 3; ++c) { x->emf[c] = p; p += nc; }
  x->divv = p; p += ndivv;
  for (int d = 0; d < 3; ++d) { x->gx[d] = p; p += n[d] + 4; x->gdx[d] = p; p += n[d] + 4; }
  x->red = p; p += 16 + 2 * DT_BLOCKS;
  if (x->has_cols) {
    for (int q = 0; q < 10; ++q) { x->col[q] = p; p += nc; }
    x->ctemp = p; p += nc;
  }
  // from here on a failure frees everything allocated so far
#define INITCHK(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) {                                                                        \
      g_err = std::string(#call) + ": " + hipGetErrorString(e_);                                  \
      (void)mcf_finalize(x);                                                                       \
      return e_ == hipErrorOutOfMemory ? MCF_E_NOMEM : MCF_E_HIP;                                  \
    }                                                                                              \
  } while (0)
  INITCHK(hipMemsetAsync(x->pool, 0, total * sizeof(double), x->stream));
  for (int d = 0; d < 3; ++d) {
    bool uni = n[d] >= 8;
    for (int k = 0; k < n[d] + 4 && uni; ++k) uni = std::memcmp(&x->hdx[d][k], &x->hdx[d][0], sizeof(double)) == 0;
    // x with inflow on both sides needs no periodic images, so no uniform spacing
    if (d == 0 && n[0] >= 8 && cfg->bcx1 == 4 && cfg->bcx2 == 4) uni = true;
    // an outflow x sweep runs as the stage kernels with the copy fills between
    if (d == 0 && cfg->bcx1 == 1) uni = false;
    x->fuse[d] = uni;
  }

  for (int d = 0; d < 3; ++d) {
    INITCHK(hipMemcpyAsync(x->gx[d], x->hx[d].data(), (n[d] + 4) * sizeof(double), hipMemcpyHostToDevice, x->stream));
    INITCHK(hipMemcpyAsync(x->gdx[d], x->hdx[d].data(), (n[d] + 4) * sizeof(double), hipMemcpyHostToDevice, x->stream));
  }
  INITCHK(hipHostMalloc((void **)&x->hred, 2 * sizeof(double), hipHostMallocDefault));
  INITCHK(hipMalloc(&x->mlp, 4 * MLP_NP * sizeof(float)));
  INITCHK(hipMalloc(&x->mlp_pk, MLP_NPK * sizeof(__bf16)));
  INITCHK(hipMalloc(&x->cool_cnt, 4 * sizeof(unsigned long long)));
  INITCHK(hipMemsetAsync(x->mlp, 0, 4 * MLP_NP * sizeof(float), x->stream));
  INITCHK(hipMemsetAsync(x->cool_cnt, 0, 4 * sizeof(unsigned long long), x->stream));
  INITCHK(hipStreamSynchronize(x->stream));
#undef INITCHK
  if (cfg->ifgrv == 2) {
    const int r = grav_setup(x);
    if (r != MCF_OK) {
      const std::string why = g_err;
      (void)mcf_finalize(x);
      g_err = why;
      return r;
    }
  }
  *out = x;
  return MCF_OK;
}

extern "C" int mcf_finalize(mcf_ctx *x) {
  if (!x) return MCF_OK;
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  for (auto &r : x->recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : x->evpool) (void)hipEventDestroy(e);
  if (x->pool) (void)hipFree(x->pool);
  if (x->mlp) (void)hipFree(x->mlp);
  if (x->mlp_pk) (void)hipFree(x->mlp_pk);
  if (x->cool_cnt) (void)hipFree(x->cool_cnt);
  if (x->hred) (void)hipHostFree(x->hred);
  if (x->dtfd) (void)hipFree(x->dtfd);
  if (x->gpool) (void)hipFree(x->gpool);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
  return MCF_OK;
}

extern "C" int mcf_upload(mcf_ctx *x, const double *U8, const double *sp8) {
  if (!x || !U8 || !sp8) { g_err = "null argument"; return MCF_E_ARG; }
  size_t bytes = x->ncell * sizeof(double);
  for (int f = 0; f < 8; ++f) {
    HIPCHK(hipMemcpyAsync(x->U[f], U8 + f * x->ncell, bytes, hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(x->sp[f], sp8 + f * x->ncell, bytes, hipMemcpyHostToDevice, x->stream));
  }
  StageScope scope(x, "upload");
  CHK(do_cc4(x)); // CC(4,0) after the restart read (M3:470)
  // column densities and CII optical depths zeroed, then 20 SHIELD sweeps
  // with ifrad = 2 (M3:472-478)
  if (x->has_cols) {
    for (int q = 0; q < 10; ++q) HIPCHK(hipMemsetAsync(x->col[q], 0, bytes, x->stream));
    if (x->cfg.ifrad == 2)
      for (int l = 0; l < 20; ++l) CHK(do_shield(x));
  }
  // GRAVTY(0,2) after the restart read (M3:481-484)
  if (x->grav) CHK(do_grav_solve(x));
  HIPCHK(hipStreamSynchronize(x->stream));
  return MCF_OK;
}

extern "C" int mcf_download_columns(mcf_ctx *x, double *col10) {
  if (!x || !col10) { g_err = "null argument"; return MCF_E_ARG; }
  if (!x->has_cols) { g_err = "no column densities: the chemistry Fcool, network and SHIELD are all off"; return MCF_E_STATE; }
  for (int q = 0; q < 10; ++q)
    HIPCHK(hipMemcpyAsync(col10 + q * x->ncell, x->col[q], x->ncell * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return MCF_OK;
}

extern "C" int mcf_download(mcf_ctx *x, double *U8, double *sp8) {
  if (!x || !U8 || !sp8) { g_err = "null argument"; return MCF_E_ARG; }
  size_t bytes = x->ncell * sizeof(double);
  for (int f = 0; f < 8; ++f) {
    HIPCHK(hipMemcpyAsync(U8 + f * x->ncell, x->U[f], bytes, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(sp8 + f * x->ncell, x->sp[f], bytes, hipMemcpyDeviceToHost, x->stream));
  }
  HIPCHK(hipStreamSynchronize(x->stream));
  return MCF_OK;
}

extern "C" int mcf_download_bcc(mcf_ctx *x, double *bcc3) {
  if (!x || !bcc3) { g_err = "null argument"; return MCF_E_ARG; }
  for (int c = 0; c < 3; ++c)
    HIPCHK(hipMemcpyAsync(bcc3 + c * x->ncell, x->bcc[c], x->ncell * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return MCF_OK;
}

// DTF (M3:402-431): interior from the caller, x ghosts dinit1 as REAL*4,
// y and z ghosts the periodic images the reference's halo exchange gives.
extern "C" int mcf_set_inflow_density(mcf_ctx *x, const float *dtf) {
  if (!x) { g_err = "null argument"; return MCF_E_ARG; }
  if (!dtf) {
    if (x->dtfd) { HIPCHK(hipStreamSynchronize(x->stream)); HIPCHK(hipFree(x->dtfd)); }
    x->dtfd = nullptr;
    return MCF_OK;
  }
  const int nx = x->g.n[0], ny = x->g.n[1], nz = x->g.n[2];
  std::vector<float> h(x->ncell, (float)x->dtf);
  for (int k = -1; k <= nz + 2; ++k)
    for (int j = -1; j <= ny + 2; ++j) {
      const int jw = j < 1 ? j + ny : (j > ny ? j - ny : j);
      const int kw = k < 1 ? k + nz : (k > nz ? k - nz : k);
      for (int i = 1; i <= nx; ++i) {
        const float v = dtf[(size_t)(i - 1) + (size_t)nx * ((size_t)(jw - 1) + (size_t)ny * (size_t)(kw - 1))];
        if (!std::isfinite(v) || v <= 0.0f) { g_err = "inflow density must be finite and positive"; return MCF_E_ARG; }
        h[(size_t)(i + 1) + (size_t)(j + 1) * x->g.st[1] + (size_t)(k + 1) * x->g.st[2]] = v;
      }
    }
  if (!x->dtfd) {
    hipError_t e = hipMalloc(&x->dtfd, x->ncell * sizeof(float));
    if (e != hipSuccess) { x->dtfd = nullptr; g_err = std::string("hipMalloc: ") + hipGetErrorString(e); return MCF_E_NOMEM; }
  }
  HIPCHK(hipMemcpyAsync(x->dtfd, h.data(), x->ncell * sizeof(float), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return MCF_OK;
}

extern "C" int mcf_set_time(mcf_ctx *x, double time, int ifevo) {
  if (!x || ifevo < 1 || ifevo > 6) { g_err = "bad argument"; return MCF_E_ARG; }
  x->time = time;
  x->ifevo = ifevo;
  return MCF_OK;
}
extern "C" int mcf_get_time(mcf_ctx *x, double *time, int *ifevo) {
  if (!x) { g_err = "null argument"; return MCF_E_ARG; }
  if (time) *time = x->time;
  if (ifevo) *ifevo = x->ifevo;
  return MCF_OK;
}

extern "C" int mcf_timestep(mcf_ctx *x, double tsave, double *dt_out, int *tag) {
  if (!x || !dt_out) { g_err = "null argument"; return MCF_E_ARG; }
  StageScope scope(x, "dt");
  DtArgs A{};
  for (int f = 0; f < 8; ++f) { A.U[f] = x->U[f]; A.sp[f] = x->sp[f]; }
  for (int c = 0; c < 3; ++c) { A.bcc[c] = x->bcc[c]; A.gdx[c] = x->gdx[c]; }
  A.out = x->red;
  A.tfinal = x->cfg.tfinal;
  A.cols = chem_cols(x);
  long rows = (long)x->g.n[1] * x->g.n[2];
  unsigned nb = (unsigned)(rows < DT_BLOCKS ? rows : DT_BLOCKS);
  if (x->chem) KLAUNCH_N("k_dtmin", k_dtmin<true>, dim3(nb), dim3(256), 184.0 * (double)rows * x->g.n[0], A, x->g, x->red + 16);
  else KLAUNCH_N("k_dtmin", k_dtmin<false>, dim3(nb), dim3(256), 104.0 * (double)rows * x->g.n[0], A, x->g, x->red + 16);
  KLAUNCH_G(k_dtfinal, dim3(1), dim3(256), 16.0 * nb, x->red + 16, (int)nb, x->red, x->cfg.tfinal);
  HIPCHK(hipMemcpyAsync(x->hred, x->red, 2 * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  double tCFL = x->hred[0], tLMT = x->hred[1];
  double tfinal = x->cfg.tfinal;
  double dt = tfinal;
  double t = x->cfg.cfl * tCFL;
  dt = dt < t ? dt : t;
  double stt = dt;
  dt = dt < tLMT ? dt : tLMT;
  int st = dt < stt ? 2 : 1;
  if (x->time + dt > tfinal) dt = tfinal - x->time;
  if (x->time + dt > tsave) dt = tsave - x->time;
  if (!(std::isfinite(dt)) || tCFL < 0.0 || tLMT < 0.0) {
    g_err = "non-finite or negative time step (err at Couran/Stblty)";
    return MCF_E_NONFINITE;
  }
  // EVOLVE stops before asking for a step once time reaches tfinal or the
  // next save time (M3:567-568); a step of zero or negative length
  // means the caller did not
  if (dt <= 0.0) {
    g_err = "no time step left: time has reached tfinal or tsave";
    return MCF_E_STATE;
  }
  *dt_out = dt;
  if (tag) *tag = st;
  return MCF_OK;
}

extern "C" int mcf_source(mcf_ctx *x, double dt) {
  if (!x) { g_err = "null argument"; return MCF_E_ARG; }
  return do_source(x, dt);
}
extern "C" int mcf_sweep(mcf_ctx *x, int dir, double dt) {
  if (!x || dir < 0 || dir > 2) { g_err = "bad argument"; return MCF_E_ARG; }
  return do_sweep(x, dir, dt);
}
extern "C" int mcf_ct(mcf_ctx *x, double dt) {
  if (!x) { g_err = "null argument"; return MCF_E_ARG; }
  return do_ct(x, dt);
}
extern "C" int mcf_dissip(mcf_ctx *x) {
  if (!x) { g_err = "null argument"; return MCF_E_ARG; }
  return do_dissip(x);
}

// ifEVO sweep orders xyz, yzx, zxy, xzy, yxz, zyx (M3:611-635)
static const int k_order[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {1, 0, 2}, {2, 1, 0}};

extern "C" int mcf_gravty(mcf_ctx *x, double dt, int mode) {
  if (!x) { g_err = "null argument"; return MCF_E_ARG; }
  if (!x->grav) { g_err = "self-gravity is off (ifgrv != 2)"; return MCF_E_STATE; }
  if (mode == 1) return MCF_OK;
  if (mode == 2) return do_grav_solve(x);
  if (mode == 3) return do_grav_kick(x, dt);
  g_err = "GRAVTY mode must be 1, 2 or 3";
  return MCF_E_ARG;
}

extern "C" int mcf_download_phi(mcf_ctx *x, double *phi) {
  if (!x || !phi) { g_err = "null argument"; return MCF_E_ARG; }
  if (!x->grav) { g_err = "self-gravity is off (ifgrv != 2)"; return MCF_E_STATE; }
  HIPCHK(hipMemcpyAsync(phi, x->phi, x->ncell * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return MCF_OK;
}

extern "C" int mcf_step(mcf_ctx *x, double dt) {
  if (!x) { g_err = "null argument"; return MCF_E_ARG; }
  if (x->grav) CHK(do_grav_kick(x, dt)); // M3:607
  CHK(do_source(x, 0.5 * dt));
  const int *o = k_order[x->ifevo - 1];
  for (int t = 0; t < 3; ++t) CHK(do_sweep(x, o[t], dt));
  x->ifevo = x->ifevo % 6 + 1;
  CHK(do_ct(x, dt));
  CHK(do_source(x, 0.5 * dt));
  if (x->grav) { // M3:653
    CHK(do_grav_solve(x));
    CHK(do_grav_kick(x, dt));
  }
  CHK(do_dissip(x));
  x->time = x->time + dt;
  return MCF_OK;
}

extern "C" int mcf_run(mcf_ctx *x, int nsteps, double tsave, double *dts, int *ndone) {
  if (!x || nsteps < 0) { g_err = "bad argument"; return MCF_E_ARG; }
  int done = 0;
  for (int s = 0; s < nsteps; ++s) {
    if (x->time >= x->cfg.tfinal || x->time >= tsave) break;
    double dt;
    int tag;
    CHK(mcf_timestep(x, tsave, &dt, &tag));
    CHK(mcf_step(x, dt));
    if (dts) dts[s] = dt;
    ++done;
  }
  if (ndone) *ndone = done;
  return MCF_OK;
}

extern "C" int mcf_sync(mcf_ctx *x) {
  if (!x) { g_err = "null argument"; return MCF_E_ARG; }
  HIPCHK(hipStreamSynchronize(x->stream));
  return MCF_OK;
}

extern "C" void *mcf_stream(mcf_ctx *x) { return x ? (void *)x->stream : nullptr; }

extern "C" int mcf_copy_probe(mcf_ctx *x, long nwords, int reps, double *ms) {
  if (!x || nwords <= 0 || reps <= 0 || !ms) { g_err = "bad argument"; return MCF_E_ARG; }
  double *buf = nullptr;
  HIPCHK(hipMalloc(&buf, 2 * (size_t)nwords * sizeof(double)));
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipMemsetAsync(buf, 0, 2 * (size_t)nwords * sizeof(double), x->stream));
  unsigned nb = 256u * 8u * 32u; // 32 blocks of 256 per CU
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(TPB), 0, x->stream, buf, buf + nwords, nwords); // warm
  HIPCHK(hipEventRecord(a, x->stream));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(nb), dim3(TPB), 0, x->stream, buf, buf + nwords, nwords);
  HIPCHK(hipEventRecord(b, x->stream));
  HIPCHK(hipEventSynchronize(b));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, a, b));
  *ms = (double)t / reps;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  HIPCHK(hipFree(buf));
  return MCF_OK;
}

static void profile_clear(mcf_ctx *x) {
  for (auto &r : x->recs) {
    x->evpool.push_back(r.a);
    x->evpool.push_back(r.b);
  }
  x->recs.clear();
  x->stage_runs.clear();
}

extern "C" int mcf_profile(mcf_ctx *x, int enable) {
  if (!x) { g_err = "null argument"; return MCF_E_ARG; }
  if (enable && !x->prof) {
    HIPCHK(hipStreamSynchronize(x->stream));
    profile_clear(x);
  }
  x->prof = enable != 0;
  return MCF_OK;
}

static int profile_copy(const std::map<std::string, mcf_kstat> &acc, mcf_kstat *out, int maxn, int *n) {
  int i = 0;
  for (auto &kv : acc) {
    if (out && i < maxn) out[i] = kv.second;
    ++i;
  }
  *n = i;
  return MCF_OK;
}

extern "C" int mcf_profile_report(mcf_ctx *x, mcf_kstat *out, int maxn, int *n) {
  if (!x || !n) { g_err = "null argument"; return MCF_E_ARG; }
  HIPCHK(hipStreamSynchronize(x->stream));
  std::map<std::string, mcf_kstat> acc;
  for (auto &r : x->recs) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    mcf_kstat &k = acc[r.name];
    std::snprintf(k.name, sizeof(k.name), "%s", r.name);
    k.launches += 1;
    k.ms += ms;
    k.bytes += r.bytes;
  }
  return profile_copy(acc, out, maxn, n);
}

// Compulsory HBM bytes of one run of a stage: every persistent array the
// stage must read or write, once per interior cell (ghost re-reads and the
// stage's own temporaries excluded).  A sweep reads rho, v(3), p, the 8
// species, Bcc(3) and the normal face B and writes rho, v(3), p and the
// species (17 + 13 doubles); CT reads rho, v(3), p and face B(3) and writes
// face B(3) and Bcc(3) (8 + 6); a SOURCE call reads rho, p and the five
// active species and writes p (7 + 1); the dt reduction reads rho, v(3), p,
// Bcc(3) and the five active species (13); DISSIP reads and writes v(3) in
// the ten cells next to each x end.
static double stage_bytes(const mcf_ctx *x, const std::string &st) {
  const double cells = (double)x->g.n[0] * x->g.n[1] * x->g.n[2];
  if (st.rfind("sweep_", 0) == 0) return 8.0 * 30.0 * cells;
  if (st == "ct") return 8.0 * 14.0 * cells;
  if (st == "source") {
    // thermal stage 8 arrays (+ the 10 columns the chemistry Fcool reads),
    // SHIELD 11 read + 10 written, network 17 read + 9 written
    double a = 0.0;
    if (x->cfg.ifthrm == 2) a += x->chem ? 18.0 : 8.0;
    if (x->cfg.ifrad == 2) a += 21.0;
    if (x->cfg.ifchem == 2) a += 26.0;
    return 8.0 * a * cells;
  }
  if (st == "dt") return 8.0 * 13.0 * cells;
  if (st == "dissip") return 8.0 * 6.0 * 20.0 * x->g.n[1] * x->g.n[2];
  // GRAVTY(dt,2) reads rho and writes Phi; GRAVTY(dt,3) reads Phi and
  // updates three velocities
  if (st == "grav_solve") return 8.0 * 2.0 * cells;
  if (st == "grav_kick") return 8.0 * 7.0 * cells;
  return 0.0;
}

extern "C" int mcf_profile_stages(mcf_ctx *x, mcf_kstat *out, int maxn, int *n) {
  if (!x || !n) { g_err = "null argument"; return MCF_E_ARG; }
  HIPCHK(hipStreamSynchronize(x->stream));
  std::map<std::string, mcf_kstat> acc;
  for (auto &r : x->recs) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    mcf_kstat &k = acc[r.stage];
    std::snprintf(k.name, sizeof(k.name), "%s", r.stage);
    k.ms += ms;
  }
  for (auto &kv : acc) {
    auto it = x->stage_runs.find(kv.first);
    const int runs = it == x->stage_runs.end() ? 0 : it->second;
    kv.second.launches = runs;
    kv.second.bytes = stage_bytes(x, kv.first) * runs;
  }
  return profile_copy(acc, out, maxn, n);
}

// ----------------------------------------------------------------------------
// KI-cooling surrogate
// ----------------------------------------------------------------------------
static int mlp_pack(mcf_ctx *x) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_NPK + 255) / 256), dim3(256), 0, x->stream, x->mlp, x->mlp_pk);
  HIPCHK(hipGetLastError());
  return MCF_OK;
}

extern "C" int mcf_mlp_set_weights(mcf_ctx *x, const float *w) {
  if (!x || !w) { g_err = "null argument"; return MCF_E_ARG; }
  for (int e = 0; e < MLP_NP; ++e)
    if (!std::isfinite(w[e])) { g_err = "non-finite weight"; return MCF_E_ARG; }
  HIPCHK(hipMemcpyAsync(x->mlp, w, MLP_NP * sizeof(float), hipMemcpyHostToDevice, x->stream));
  hipLaunchKernelGGL(k_fill_f32, dim3((2 * MLP_NP + 255) / 256), dim3(256), 0, x->stream, x->mlp + MLP_NP, 2 * MLP_NP, 0.0f);
  HIPCHK(hipGetLastError());
  CHK(mlp_pack(x));
  HIPCHK(hipStreamSynchronize(x->stream));
  x->adam_t = 0;
  x->mlp_ready = true;
  return MCF_OK;
}

extern "C" int mcf_mlp_get_weights(mcf_ctx *x, float *w) {
  if (!x || !w) { g_err = "null argument"; return MCF_E_ARG; }
  if (!x->mlp_ready) { g_err = "no weights loaded"; return MCF_E_STATE; }
  HIPCHK(hipMemcpyAsync(w, x->mlp, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return MCF_OK;
}

// Gaussian initial weights, std 1/sqrt(fan_in), zero biases.
extern "C" int mcf_mlp_init(mcf_ctx *x, uint64_t seed) {
  if (!x) { g_err = "null argument"; return MCF_E_ARG; }
  std::vector<float> w(MLP_NP, 0.0f);
  uint64_t s = seed;
  auto next = [&s]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  auto gauss = [&]() { return std::sqrt(-2.0 * std::log(next())) * std::cos(6.283185307179586 * next()); };
  for (int e = 0; e < 192; ++e) w[MLP_OW1 + e] = (float)(gauss() / std::sqrt(3.0));
  for (int e = 0; e < 4096; ++e) w[MLP_OW2 + e] = (float)(gauss() / 8.0);
  for (int e = 0; e < 64; ++e) w[MLP_OW3 + e] = (float)(gauss() / 8.0);
  return mcf_mlp_set_weights(x, w.data());
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
};

extern "C" int mcf_cool_samples(mcf_ctx *x, int64_t n, uint64_t seed, float *X, float *Y) {
  if (!x || n <= 0 || !X || !Y) { g_err = "bad argument"; return MCF_E_ARG; }
  DevBuf b;
  HIPCHK(hipMalloc(&b.p, 4 * (size_t)n * sizeof(float)));
  float *dX = (float *)b.p, *dY = dX + 3 * n;
  hipLaunchKernelGGL(k_cool_samples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, x->stream, (long)n, (unsigned long long)seed, dX, dY);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(X, dX, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(Y, dY, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return MCF_OK;
}

extern "C" int mcf_mlp_forward(mcf_ctx *x, const float *X, int64_t n, float *Y) {
  if (!x || n <= 0 || !X || !Y) { g_err = "bad argument"; return MCF_E_ARG; }
  if (!x->mlp_ready) { g_err = "no weights loaded"; return MCF_E_STATE; }
  DevBuf b;
  HIPCHK(hipMalloc(&b.p, 4 * (size_t)n * sizeof(float)));
  float *dX = (float *)b.p, *dY = dX + 3 * n;
  HIPCHK(hipMemcpyAsync(dX, X, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, x->stream));
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, x->stream, x->mlp_pk, x->mlp, dX, (long)n, dY);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(Y, dY, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return MCF_OK;
}

static int mlp_grad_dev(mcf_ctx *x, const float *dX, const float *dY, long n, float *dg, float *dloss) {
  long nch = (n + 63) / 64;
  unsigned nb = (unsigned)(nch < 4096 ? nch : 4096);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(64), 0, x->stream, x->mlp_pk, x->mlp, dX, dY, n, 1.0f / (float)n, dg, dloss);
  HIPCHK(hipGetLastError());
  return MCF_OK;
}

extern "C" int mcf_mlp_grad(mcf_ctx *x, const float *X, const float *T, int64_t n, float *grad, float *loss) {
  if (!x || n <= 0 || !X || !T || !grad) { g_err = "bad argument"; return MCF_E_ARG; }
  if (!x->mlp_ready) { g_err = "no weights loaded"; return MCF_E_STATE; }
  DevBuf b;
  HIPCHK(hipMalloc(&b.p, (4 * (size_t)n + MLP_NP + 64) * sizeof(float)));
  float *dX = (float *)b.p, *dY = dX + 3 * n, *dg = dY + n, *dl = dg + MLP_NP;
  HIPCHK(hipMemcpyAsync(dX, X, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(dY, T, (size_t)n * sizeof(float), hipMemcpyHostToDevice, x->stream));
  hipLaunchKernelGGL(k_fill_f32, dim3((MLP_NP + 64 + 255) / 256), dim3(256), 0, x->stream, dg, MLP_NP + 64, 0.0f);
  CHK(mlp_grad_dev(x, dX, dY, (long)n, dg, dl));
  HIPCHK(hipMemcpyAsync(grad, dg, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, x->stream));
  float hl = 0.0f;
  HIPCHK(hipMemcpyAsync(&hl, dl, sizeof(float), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (loss) *loss = hl;
  return MCF_OK;
}

// Adam on the mean-squared error of IMC samples, cosine learning-rate decay.
extern "C" int mcf_mlp_train(mcf_ctx *x, const mcf_train_opts *o, float *epoch_loss, float *val_loss) {
  if (!x || !o || o->nsamples < o->batch || o->batch < 64 || o->batch % 64 || o->epochs <= 0 || o->nval < 64) {
    g_err = "bad training options (batch a multiple of 64, nsamples >= batch, nval >= 64)";
    return MCF_E_ARG;
  }
  if (!x->mlp_ready) CHK(mcf_mlp_init(x, o->seed ^ 0x2545F4914F6CDD1Dull));
  const long n = (long)o->nsamples, nv = (long)o->nval, B = o->batch;
  DevBuf b;
  HIPCHK(hipMalloc(&b.p, (4 * (size_t)(n + nv) + 2 * (size_t)o->epochs) * sizeof(float)));
  float *dX = (float *)b.p, *dY = dX + 3 * n, *vX = dY + n, *vY = vX + 3 * nv, *el = vY + nv, *vl = el + o->epochs;
  hipLaunchKernelGGL(k_cool_samples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, x->stream, n, (unsigned long long)o->seed, dX, dY);
  hipLaunchKernelGGL(k_cool_samples, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, x->stream, nv,
                     (unsigned long long)(o->seed ^ 0x6A09E667F3BCC909ull), vX, vY);
  hipLaunchKernelGGL(k_fill_f32, dim3((2 * o->epochs + 255) / 256), dim3(256), 0, x->stream, el, 2 * o->epochs, 0.0f);
  HIPCHK(hipGetLastError());
  float *w = x->mlp, *m = w + MLP_NP, *v = m + MLP_NP, *g = v + MLP_NP;
  const long steps = n / B, total = steps * o->epochs;
  const unsigned nbp = (MLP_NP + 255) / 256;
  for (int ep = 0; ep < o->epochs; ++ep) {
    for (long s = 0; s < steps; ++s) {
      hipLaunchKernelGGL(k_fill_f32, dim3(nbp), dim3(256), 0, x->stream, g, MLP_NP, 0.0f);
      hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)(B / 64)), dim3(64), 0, x->stream, x->mlp_pk, w, dX + 3 * s * B,
                         dY + s * B, (long)B, 1.0f / (float)B, g, el + ep);
      x->adam_t += 1;
      const double tt = (double)(ep * steps + s + 1) / (double)total;
      const float lr = (float)(o->lr_min + 0.5 * (o->lr - o->lr_min) * (1.0 + std::cos(3.141592653589793 * tt)));
      hipLaunchKernelGGL(k_mlp_adam, dim3(nbp), dim3(256), 0, x->stream, w, m, v, g, lr,
                         (float)std::pow(0.9, (double)x->adam_t), (float)std::pow(0.999, (double)x->adam_t));
      hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_NPK + 255) / 256), dim3(256), 0, x->stream, w, x->mlp_pk);
    }
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_fill_f32, dim3(nbp), dim3(256), 0, x->stream, g, MLP_NP, 0.0f);
    CHK(mlp_grad_dev(x, vX, vY, nv, g, vl + ep));
  }
  std::vector<float> hl(2 * o->epochs);
  HIPCHK(hipMemcpyAsync(hl.data(), el, 2 * o->epochs * sizeof(float), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  for (int ep = 0; ep < o->epochs; ++ep) {
    if (epoch_loss) epoch_loss[ep] = hl[ep] / (float)steps;
    if (val_loss) val_loss[ep] = hl[o->epochs + ep];
  }
  return MCF_OK;
}

extern "C" int mcf_cool_solve(mcf_ctx *x, const double *P0, const double *nn, const double *dt, const double *rho,
                              const double *gammi1, int64_t n, int mode, double *P, int32_t *flag, double *ms) {
  if (!x || n <= 0 || !P0 || !nn || !dt || !rho || !gammi1 || !P || (mode != 0 && mode != 1)) {
    g_err = "bad argument";
    return MCF_E_ARG;
  }
  if (mode == 1 && !x->mlp_ready) { g_err = "no weights loaded"; return MCF_E_STATE; }
  DevBuf b;
  const size_t nb = (size_t)n * sizeof(double);
  HIPCHK(hipMalloc(&b.p, 6 * nb + 2 * (size_t)n * sizeof(int) + 64));
  double *d0 = (double *)b.p, *d1 = d0 + n, *d2 = d1 + n, *d3 = d2 + n, *d4 = d3 + n, *dP = d4 + n;
  int *df = (int *)(dP + n), *fq = df + n;
  unsigned *nfb = (unsigned *)(fq + n);
  hipLaunchKernelGGL(k_zero_u32x2, dim3(1), dim3(64), 0, x->stream, nfb);
  HIPCHK(hipMemcpyAsync(d0, P0, nb, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(d1, nn, nb, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(d2, dt, nb, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(d3, rho, nb, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(d4, gammi1, nb, hipMemcpyHostToDevice, x->stream));
  hipEvent_t ea, eb;
  HIPCHK(hipEventCreate(&ea));
  HIPCHK(hipEventCreate(&eb));
  HIPCHK(hipEventRecord(ea, x->stream));
  hipLaunchKernelGGL(k_cool_solve, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, x->stream, d0, d1, d2, d3, d4,
                     (long)n, mode, x->mlp_pk, x->mlp, dP, df, nfb, fq);
  if (mode == 1)
    hipLaunchKernelGGL(k_cool_fallback, dim3(256), dim3(256), 0, x->stream, d0, d1, d2, d3, d4, nfb, fq, dP);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(eb, x->stream));
  HIPCHK(hipMemcpyAsync(P, dP, nb, hipMemcpyDeviceToHost, x->stream));
  if (flag) HIPCHK(hipMemcpyAsync(flag, df, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, ea, eb));
  if (ms) *ms = t;
  (void)hipEventDestroy(ea);
  (void)hipEventDestroy(eb);
  return MCF_OK;
}

// mcf_cool_solve's kernels on caller-owned device arrays, with the counter
// pair and the fallback queue in this device's scratch; no allocation, no
// synchronisation.  The counters are cleared by a kernel (k_zero_u32x2), so
// the call can be captured into a HIP graph.
extern "C" int mcf_cool_solve_noalloc(mcf_ctx *x, const double *P0, const double *nn, const double *dt,
                                      const double *rho, const double *gammi1, int64_t n, int mode, double *P,
                                      int32_t *flag) {
  if (!x || !P0 || !nn || !dt || !rho || !gammi1 || !P || (mode != 0 && mode != 1)) {
    g_err = "bad argument";
    return MCF_E_ARG;
  }
  if (n <= 0 || n > (int64_t)MCF_NOALLOC_MAX_CELLS) {
    g_err = "ncell must be in 1..MCF_NOALLOC_MAX_CELLS (2097152) for mcf_cool_solve_noalloc";
    return MCF_E_ARG;
  }
  if (mode == 1 && !x->mlp_ready) { g_err = "no weights loaded"; return MCF_E_STATE; }
  if (!x->scratch) { g_err = "no-allocation scratch not set up"; return MCF_E_STATE; }
  unsigned *nfb = reinterpret_cast<unsigned *>(x->scratch);
  int *fq = reinterpret_cast<int *>(static_cast<unsigned char *>(x->scratch) + NOALLOC_HDR);
  hipLaunchKernelGGL(k_zero_u32x2, dim3(1), dim3(64), 0, x->stream, nfb);
  hipLaunchKernelGGL(k_cool_solve, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, x->stream, P0, nn, dt, rho,
                     gammi1, (long)n, mode, x->mlp_pk, x->mlp, P, flag, nfb, fq);
  if (mode == 1)
    hipLaunchKernelGGL(k_cool_fallback, dim3(256), dim3(256), 0, x->stream, P0, nn, dt, rho, gammi1, nfb, fq, P);
  HIPCHK(hipGetLastError());
  return MCF_OK;
}

extern "C" int mcf_cool_stats(mcf_ctx *x, int64_t *counts, int reset) {
  if (!x || !counts) { g_err = "null argument"; return MCF_E_ARG; }
  unsigned long long h[2];
  HIPCHK(hipMemcpyAsync(h, x->cool_cnt, sizeof(h), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  counts[0] = (int64_t)h[0];
  counts[1] = (int64_t)h[1];
  if (reset) {
    HIPCHK(hipMemsetAsync(x->cool_cnt, 0, sizeof(h), x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
  }
  return MCF_OK;
}
