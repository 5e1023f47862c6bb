// mcf_grav.h -- self-gravity GRAVTY (3DMHD-CfCA/selfG_MPIfull8.f90, "SG8")
// for gfx950: the multigrid Poisson solver with the free-space boundary
// potential on the x ends, and the velocity kick.
//
// The reference's result depends on its MPI block split (Gauss-Seidel per
// rank with fixed ghost nodes), so the solver keeps the reference's rank
// layout: on the split levels NGcr..NGL every rank's block of nodes
// 0..n (per axis) is a contiguous array, blocks of one level side by side;
// the whole-domain levels 1..NGcr (2^l+1 nodes per axis) are stacked in one
// flat array with the reference's offsets (point1, SG8:111-121), which
// matters because the whole-domain residual reads past the ends of its
// level (resid, SG8:955-971).
#pragma once
#include <hip/hip_runtime.h>

#include "mcf_kernels.h"

namespace mcf {

constexpr int GRAV_MAXL = 16;
// whole-domain levels held in LDS by k_grav_coarse: NGcr <= 4 (at most 8
// ranks along an axis), point1(5) = 5795 nodes per array
constexpr int GRAV_MAX_NGCR = 4;
constexpr int GRAV_NEED1_MAX = 5795;
constexpr int GRAV_COARSE_THREADS = 1024;
constexpr int GRAV_RELAX_THREADS = 1024;
// planes of loads in flight in k_grav_relax_col with 1 and 4 columns per thread
constexpr int GRAV_RELAX_D1 = 4, GRAV_RELAX_D4 = 1;
constexpr int GRAV_FFT_MAXL = 1024; // longest FFT line (nn1/2 and nn2)
// 4 pi G (M3:35: G = 1.11142d-4, G4pi = 12.56637d0*G)
constexpr double GRAV_G4PI = 12.56637 * 1.11142e-4;

struct GravGeo {
  int S[3], C[3], npe, NGL, NGcr;
  double Lbox;
  int n[GRAV_MAXL][3];      // split level l: nodes 0..n[l][d] per rank block
  long off[GRAV_MAXL];      // split level l: offset of rank 0's block (doubles)
  long bsz[GRAV_MAXL];      // split level l: block size (n0+1)(n1+1)(n2+1)
  long p1[GRAV_MAXL + 2];   // whole-domain level l at 1-based offset p1[l]
  long need1;               // p1[NGcr+1]
  long pb1[GRAV_MAXL + 2];  // end-plane potential of whole-domain level l (pointb1)
  long needb1;
  long boff[GRAV_MAXL];     // end planes of split level l: rank r, side s at boff[l] + (2r+s)*(n1+1)(n2+1)
};

struct GravBuf {
  double *phi2, *rho2, *res2, *rhs2; // split levels
  double *phi1, *rho1, *res1, *rhs1; // whole-domain levels, need1+1 entries (1-based)
  double *bphi1;                     // 2 x (needb1+1), 1-based
  double *bphi2;
  double *ex[2]; // boundary planes of consecutive sweeps (k_grav_relax_col), 6 per rank block
};

// threads of an exchange per rank block at a split level of nodes 0..n
// (k_grav_exchange): the two z planes walked as (i+j, j), the y planes
// without their z rows, the x planes without both
__host__ __device__ inline long grav_nghost(const int *n) {
  return 2L * (n[0] + n[1] + 1) * (n[1] + 1) + 2L * (n[0] + 1) * (n[2] - 1) + 2L * (n[1] - 1) * (n[2] - 1);
}

// PB (SG8:1026-1463): the x planes whose Fourier modes make the end-plane
// potential, with their distances from the two ends and 4 pi G dx factors
struct PbArgs {
  int np;          // planes
  int nn1, nn2;    // global y, z cells
  const int *gx;   // global x index of each plane
  const double *zp;   // distance of the plane from the left end
  const double *zq;   // ... from the right end (Lbox - zp)
  const double *facG; // -4 pi G dx(1) of the plane's rank
  double dxx, dyy;    // dy(1), dz(1)
  const double2 *wf;  // rlft3 unpacking twiddles, forward (index i1, single precision values)
  const double2 *wi;  // ... inverse
};

__global__ void k_grav_zero(double *a, long n);
__global__ void k_grav_rhs(GravGeo G, const double *rho, Geo g, double *rhs2);
__global__ void k_grav_rstrct(GravGeo G, double *uc2, const double *uf2, int lc, int mode);
__global__ void k_grav_exchange(GravGeo G, double *u2, int l, int flags);
__global__ void k_grav_relax(GravGeo G, double *phi2, const double *rhs2, int l);
template <int PP, int LMAX, int D, bool PL>
__global__ void k_grav_relax_col(GravGeo G, double *phi2, const double *rhs2, int l, const double *Ein, double *Eout,
                                 long EP);
__global__ void k_grav_resid(GravGeo G, double *res2, const double *phi2, const double *rhs2, int l);
__global__ void k_grav_prolong(GravGeo G, double *uf2, const double *uc2, int lf, int mode, const double *bphi2);
__global__ void k_grav_add(GravGeo G, double *uf2, const double *res2, int l);
__global__ void k_grav_collect(GravGeo G, double *u1, const double *u2);
__global__ void k_grav_divide(GravGeo G, double *u2, const double *u1);
// the whole-domain part of the solve as one workgroup (prog: 0 set-up,
// 1 FMG stage j < NGcr, 2 FMG stage NGcr before the split levels, 3 the
// bottom of a V-cycle below the coarsest split level)
__global__ void k_grav_coarse(GravGeo G, GravBuf B, int prog, int j);
__global__ void k_pb_pack(PbArgs P, const double *rho, Geo g, double2 *Z);
__global__ void k_pb_fft(double2 *Z, int L, long es, long inner_stride, long outer_stride, int ninner, int sign,
                         int lcw);
__global__ void k_pb_unpack(PbArgs P, double2 *Z, double2 *S, int nplanes, int isign);
__global__ void k_pb_weight(PbArgs P, const double2 *Z, const double2 *S, double2 *D, double2 *SD);
__global__ void k_pb_bphi(GravGeo G, PbArgs P, const double2 *D, double *bphi2);
__global__ void k_pb_restrict(GravGeo G, double *bphi2, int lc);
__global__ void k_pb_collect(GravGeo G, double *bphi1, const double *bphi2);
__global__ void k_grav_phi(GravGeo G, const double *phi2, double *Phi, Geo g);
__global__ void k_grav_phi_ghosts(double *Phi, Geo g);
__global__ void k_grav_kick(double *U2, double *U3, double *U4, const double *Phi, Geo g, const double *gdx0, int cx,
                            double dt);

} // namespace mcf
