// mcf_kernels.h -- kernel argument blocks and declarations (FP64 path).
#pragma once
#include <hip/hip_runtime.h>

#include "mcf_chem.h"
#include "mcf_device.h"

namespace mcf {

// Index box (-1:n+2)^3, x fastest: st = {1, nx+4, (nx+4)*(ny+4)}
struct Geo {
  int n[3];
  int st[3];
};

// linear index of cell (i,j,k), i, j, k in -1..n+2
__device__ __forceinline__ int ix3(const Geo &g, int i, int j, int k) {
  return (i + 1) + (j + 1) * g.st[1] + (k + 1) * g.st[2];
}

// Range of (i,j,k) a kernel covers: lo .. lo+cnt-1 per axis
struct Box {
  int lo[3];
  int cnt[3];
};

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Renumber them so that every XCD takes one contiguous range of tiles,
// ordered with the first two grid axes fastest: tiles that share halo rows
// (neighbours along x for the x sweep, along the sweep axis for y/z) then
// run on the same XCD at about the same time.  The renumbering is a
// bijection: with nt = 8q + r tiles, XCD x < r takes q+1 of them, the
// others q.
__device__ __forceinline__ void xcd_tile(int &bx, int &by, int &bz) {
  const unsigned gx = gridDim.x, gy = gridDim.y, nt = gx * gy * gridDim.z;
  unsigned t = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned q = nt >> 3, r = nt & 7u, x = t & 7u, s = t >> 3;
  t = x < r ? x * (q + 1u) + s : r * (q + 1u) + (x - r) * q + s;
  bx = (int)(t % gx);
  by = (int)((t / gx) % gy);
  bz = (int)(t / (gx * gy));
}

struct FillList {
  double *a[16];
  int nf;
  int dm[16]; // k_fill_yz: the passes of each field (bit 0: y, bit 1: z)
};

enum XRule { XR_PERIODIC = 0, XR_CONST, XR_COPY, XR_FACE, XR_MOM, XR_ENERGY, XR_SPECIES, XR_DTF };

// Inflow density read from the DTF field shifted by -v t and wrapped in x
// (XR_DTF): the field (float, the (-1:N+2)^3 box), the simulation time, the
// inflow velocity, dx(0) and L = dx(0)*Ncellx*NSPLTx of the boundary rank on
// each side, and the global cell edges x(i) at xe[i+1].
struct DtfShift {
  const float *f;
  const double *xe;
  double t;
  double v[2], dx0[2], L[2];
};

// x-direction fill with the inflow (BC=4) rules of BC_MPI / BC_MPI_OT
struct XFill {
  double *a[16];
  int rule[16][2];   // per field and side (0 = left ghosts 1-nol..0, 1 = right N+1..N+nol)
  double val[16][2];
  int src[16][2];
  int nf;
  const double *rho, *bcc1, *blg1, *blg2;
  double val2[2];
  DtfShift dtf;
};

// Fused sweep stages (mcf_sweep.hip): inputs are the current buffers,
// outputs the other buffer of each field the stage writes.
struct SweepIO {
  const double *U[8];
  const double *sp[8];
  const double *bcc[3];
  double *Uo[8];
  double *spo[8];
  const double *blg[2];
  double *blgo[2];
  const double *u5e_in; // specific total energy from RHS (read by MOC)
  double *u5e;          // written by RHS
  double *xlag, *va, *dxlagM;
  const double *xe, *dxe;
  int d, VN, VT1, VT2, BN, BT1, BT2;
  int nblk;
  double facdep, rmin, pmin, pmax;
  double spmin[8];
};

// Ghost values of the x sweep's RHS results as MOC reads them, per side
// (0 = left, 1 = right): inflow sides (BC=4) take the fill rules of the MOC
// stage, periodic sides the periodic images.
struct XSide {
  int inflow[2];
  double blg[2][2]; // Blg(t) ghost value per side
  double dxl[2];    // dx_lag ghost value per side
  // inflow ghost rules of the REMAP stage (k_sweep_x): momentum_x = rho *
  // vin, momenta y, z constants, E = ein*1.5 + (rho vin^2 + Bcc_x^2 +
  // Blg^2)/2, species = spin * rho
  double vin[2], mt[2][2], ein[2], spin[8];
};

// R = double: strict FP64; R = float: mixed precision
template <int D, class R> __global__ void k_rhs(SweepIO A, Geo g, double dt);
template <int D, class R> __global__ void k_rhsmoc_col(SweepIO A, Geo g, double dt); // D = 1, 2
template <int D, class R> __global__ void k_remap_col(SweepIO A, Geo g, double dt); // D = 1, 2
template <int D, class R> __global__ void k_sweep_col(SweepIO A, Geo g, double dt, int seg); // D = 1, 2: whole sweep
template <int D, class R> __global__ void k_sweep_col2(SweepIO A, Geo g, double dt, int seg); // two waves per column row
constexpr int SWEEP_COL_MAXN = 2048; // longest axis k_sweep_col takes (cell edges in LDS)
template <class R> __global__ void k_rhsmoc_x(SweepIO A, Geo g, XSide S, double dt);
constexpr int XHALO = 4, XCELLS = 64 - 2 * XHALO; // k_rhsmoc_x: halo and cells per wave
template <class R> __global__ void k_remap_x(SweepIO A, Geo g, double dt);
constexpr int XRCELLS = 60; // k_remap_x: cells per wave (halo 2)
template <class R> __global__ void k_sweep_x(SweepIO A, Geo g, XSide S, double dt);
constexpr int XSCELLS = 52; // k_sweep_x: cells per wave (halo 6)
constexpr int FUSE_SEG = 128; // cells per column segment of k_rhsmoc_col
template <int D, class R> __global__ void k_moc(SweepIO A, Geo g, double dt);
template <int D, class R> __global__ void k_remap(SweepIO A, Geo g, double dt);

struct CtArgs {
  double *U[8];
  double *bcc[3];
  double *vfc[3];
  double *emf[3];
  const double *gdx[3]; // dx arrays, dx(i) at [i+1]
  double facdep;
};

struct SrcArgs {
  const double *Tp;   // pressure the stage computes T from
  double *Pout;       // result of the stage
  double *Pn;         // pressure before SOURCE (P_old of both stages)
  const double *rho;
  const double *sp[8];
  const double *gdx[3];
  double pmin, pmax;
  // cooling surrogate (sur != 0): packed MLP operands, FP32 weights and the
  // counters {backward-Euler cells, IMC fallbacks}
  // and the queue of backward-Euler cells with their conduction increments
  int sur;
  const __bf16 *pk;
  const float *mw;
  unsigned long long *cnt;
  unsigned *qn;   // {queued cells, rejected cells}
  unsigned *qz;   // the other stage's pair of counters, cleared by k_src_col
  int *queue, *fqueue;
  double *qdc;
  ChemCols cols; // column densities the chemistry Fcool reads
};

struct DtArgs {
  const double *U[8];
  const double *sp[8];
  const double *bcc[3];
  const double *gdx[3];
  double *out;
  double tfinal;
  ChemCols cols; // column densities the chemistry Fcool reads
};

// SHIELD (k_shield): species, pressure, the ten column arrays, a per-cell
// scratch for the CII optical-depth rate, dx (dx(i) at [i+1]) and the x
// block split of the reproduced MPI run
struct ShieldArgs {
  const double *sp[8];
  const double *p;
  double *col[10];
  double *temp;
  const double *gdx0;
  int nsx, cx;
};
constexpr int SHIELD_MAXB = 64; // x blocks k_shield takes

// chemistry network (k_chem)
struct ChemArgs {
  double *sp[8];
  double *rho;
  const double *p;
  ChemCols cols;
  double spmin[8];
};

struct DisArgs {
  double *U[8];
  double *divv;
};

__global__ void k_fill_yz(FillList L, Geo g, int nol);
__global__ void k_fill_xrule(XFill X, Geo g, int nol);
__global__ void k_cc1(CtArgs A, Geo g, Box b);
template <class R> __global__ void k_ct_march(CtArgs A, Geo g, double dt);
__global__ void k_ct_march_mixed(CtArgs A, Geo g, double dt); // the FP32 march at three waves per SIMD
// k_ct_march: EMF cells per wave along x, rows (waves) per block, planes
// per block (16, 32 and 64 planes measured 1.01 / 1.04 / 1.11 ms, mixed,
// 256^3; 2 and 8 rows no faster than 4)
constexpr int CTM_OUT = 61, CTM_ROWS = 4, CTM_ZSEG = 16;
__global__ void k_cc3(CtArgs A, Geo g, Box b, double dt);
__global__ void k_cc4(CtArgs A, Geo g, Box b);
// CHEM: the chemistry-network Fcool (R = double only) instead of the KI one
template <class R, bool CHEM = false> __global__ void k_src_col(SrcArgs A, Geo g, Box b, double dt, int stage);
// planes per block of k_src_col: FP32 rates (mixed) 64, FP64 32
constexpr int SRC_ZSEG_F32 = 64, SRC_ZSEG_F64 = 32;
constexpr int SRC_ROWS = 8;  // x rows (waves) per block of k_src_col
__global__ void k_shell_copy(const double *__restrict__ src, double *__restrict__ dst, Geo g);
__global__ void k_src_stiff(SrcArgs A, double dt, int stage);
template <bool CHEM = false> __global__ void k_src_imc(SrcArgs A, double dt, int stage, int which);
__global__ void k_zero_u32x2(unsigned *p);
template <bool CHEM = false> __global__ void k_dtmin(DtArgs A, Geo g, double *part);
__global__ void k_dtfinal(const double *part, int nb, double *out, double tfinal);
__global__ void k_dissip_div(DisArgs A, Geo g, int e0);
__global__ void k_dissip_upd(DisArgs A, Geo g, int e0);
__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, long n);
__global__ void k_shield(ShieldArgs S, Geo g);
__global__ void k_chem(ChemArgs A, Geo g, double dt);

// cooling surrogate (mcf_cool.hip)
__global__ void k_cool_samples(long n, unsigned long long seed, float *X, float *Y);
__global__ void k_mlp_pack(const float *w, __bf16 *pk);
__global__ void k_mlp_infer(const __bf16 *pk, const float *w, const float *X, long n, float *Y);
__global__ void k_mlp_grad(const __bf16 *pk, const float *w, const float *X, const float *Yt, long n, float inv_b,
                           float *grad, float *loss);
__global__ void k_mlp_adam(float *w, float *m, float *v, const float *g, float lr, float b1t, float b2t);
__global__ void k_fill_f32(float *a, int n, float v);
__global__ void k_cool_solve(const double *P0, const double *nn, const double *dt, const double *rho,
                             const double *gi, long n, int mode, const __bf16 *pk, const float *w, double *Pout,
                             int *flag, unsigned *nfb, int *fq);
__global__ void k_cool_fallback(const double *P0, const double *nn, const double *dt, const double *rho,
                                const double *gi, const unsigned *nfb, const int *fq, double *Pout);

} // namespace mcf
