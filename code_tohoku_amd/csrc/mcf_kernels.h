// mcf_kernels.h -- kernel argument blocks and declarations (FP64 path).
#pragma once
#include <hip/hip_runtime.h>

#include "mcf_device.h"

namespace mcf {

// Index box (-1:n+2)^3, x fastest: st = {1, nx+4, (nx+4)*(ny+4)}
struct Geo {
  int n[3];
  int st[3];
};

// Range of (i,j,k) a kernel covers: lo .. lo+cnt-1 per axis
struct Box {
  int lo[3];
  int cnt[3];
};

struct FillList {
  double *a[16];
  int nf;
};

enum XRule { XR_PERIODIC = 0, XR_CONST, XR_COPY, XR_FACE, XR_MOM, XR_ENERGY, XR_SPECIES };

// x-direction fill with the inflow (BC=4) rules of BC_MPI / BC_MPI_OT
struct XFill {
  double *a[16];
  int rule[16][2];   // per field and side (0 = left ghosts 1-nol..0, 1 = right N+1..N+nol)
  double val[16][2];
  int src[16][2];
  int nf;
  const double *rho, *bcc1, *blg1, *blg2;
  double val2[2];
};

// One MHD sweep along axis d (M3:1264-1324).  Field indices into U:
// 0 rho, 1..3 v (or momenta), 4 p (or E), 5..7 face B.
struct SweepArgs {
  double *U[8];
  double *sp[8];
  const double *bcc[3];
  double *u5e;       // specific total energy of interior cells (RHS -> MOC)
  double *blg[2];    // transverse B on the Lagrangian grid
  double *xlag;      // Lagrangian interface positions
  double *dxlagM;    // Lagrangian cell widths
  double *pa, *va;   // interface pressure / velocity (Va = Vfc of the sweep)
  double *mf[8];     // MOC interface values, left view then right view
  double *fl[15];    // REMAP fluxes
  const double *xe;  // x(i), index -1..N+2 at [i+1]
  const double *dxe; // dx(i)
  int d, VN, VT1, VT2, BN, BT1, BT2;
  int nblk;          // cells per MPI block of the reproduced run along d
  double facdep, rmin, pmin, pmax;
  double spmin[8];
};

struct CtArgs {
  double *U[8];
  double *bcc[3];
  double *vfc[3];
  double *dq[4];
  double *emf[3];
  const double *gdx[3]; // dx arrays, dx(i) at [i+1]
  double facdep;
};

struct SrcArgs {
  double *U5;
  const double *rho;
  const double *sp[8];
  double *Tn, *Pn;
  const double *gdx[3];
  double pmin, pmax;
};

struct DtArgs {
  const double *U[8];
  const double *sp[8];
  const double *bcc[3];
  const double *gdx[3];
  double *out;
  double tfinal;
};

struct DisArgs {
  double *U[8];
  double *divv;
};

__global__ void k_fill_periodic(FillList L, Geo g, int d, int nol);
__global__ void k_fill_xrule(XFill X, Geo g, int nol);
__global__ void k_rhs_face(SweepArgs A, Geo g, Box b, double dt);
__global__ void k_rhs_cell(SweepArgs A, Geo g, Box b, double dt);
__global__ void k_moc_face(SweepArgs A, Geo g, Box b, double dt);
__global__ void k_moc_cell(SweepArgs A, Geo g, Box b, double dt);
__global__ void k_remap_face(SweepArgs A, Geo g, Box b, double dt);
__global__ void k_remap_cell(SweepArgs A, Geo g, Box b, double dt);
__global__ void k_cc1(CtArgs A, Geo g, Box b);
__global__ void k_cct_slopes(CtArgs A, Geo g, Box b, int orient, int mode);
__global__ void k_cct_emf(CtArgs A, Geo g, Box b, int orient, double dt);
__global__ void k_cc3(CtArgs A, Geo g, Box b, double dt);
__global__ void k_cc4(CtArgs A, Geo g, Box b);
__global__ void k_src_temp(SrcArgs A, Geo g, Box b, int savep);
__global__ void k_src_cell(SrcArgs A, Geo g, Box b, double dt, int stage);
__global__ void k_dtmin(DtArgs A, Geo g, Box b);
__global__ void k_dtinit(double *out, double v);
__global__ void k_dissip_div(DisArgs A, Geo g, int e0);
__global__ void k_dissip_upd(DisArgs A, Geo g, int e0);
__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, long n);

} // namespace mcf
