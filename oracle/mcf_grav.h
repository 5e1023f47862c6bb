/*
 * mcf_grav.h -- CPU restatement of the self-gravity solver GRAVTY
 * (3DMHD-CfCA/selfG_MPIfull8.f90, below "SG8"), used ONLY as the checker
 * for the HIP path.  TEST INFRASTRUCTURE: nothing in the product links it.
 *
 * The solver reproduces the reference's MPI run with NSPLTx x NSPLTy x
 * NSPLTz ranks: every rank's multigrid arrays are kept separately and the
 * halo exchanges (BCsgr_MPI), gathers (collect) and scatters (divide) move
 * data between them in the reference's order.
 *
 * Determinism: mglin (SG8:151-295) reads parts of its automatic arrays
 * before writing them (the y/z ghost planes of the coarsest split level
 * that `divide` skips on the first ranks, SG8:369-382), and PB leaves the
 * index-0 entries of the coarse boundary arrays unset.  This restatement
 * starts every solve with all of them zero, as the fixture build of the
 * reference does (oracle/patch_ref.py, option "grav").
 */
#ifndef MCF_GRAV_H
#define MCF_GRAV_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct mcfg mcfg;

/* n: global cells per axis (equal powers of two), s: NSPLTx/y/z, lbox: the
 * x length ql1x+ql2x (M3:483), xi/dxi: the global cell edges x_i and widths
 * dx_i of each axis, index -1..n+2 (pointer to index -1, M3:255-281).
 * Returns NULL on an unsupported shape. */
mcfg *mcfg_create(const int n[3], const int s[3], double lbox, const double *const xi[3],
                  const double *const dxi[3]);
void mcfg_destroy(mcfg *g);
/* GRAVTY(dt,2) after the density ghosts are filled (SG8:15-79): rho and phi
 * are (n+4)^3 arrays, x fastest, index -1..n+2.  Writes phi at every
 * interior cell, at the face ghosts (periodic images in y and z, the
 * zero-gradient copies at the x ends) and at the planes the reference's
 * exchanges leave defined. */
void mcfg_solve(mcfg *g, const double *rho, double *phi);
/* Pieces exposed for known-answer tests. */
/* rlft3 (SG8:1466-1512): data nn1*nn2 reals (first index fastest), speq
 * 2*nn2 (complex), isign +1 forward / -1 inverse. */
void mcfg_rlft3(double *data, double *speq, int nn1, int nn2, int isign);
/* The boundary potential of PB at the finest level for rank r, side l
 * (0 = x left end, 1 = right end): (C_y+2) x (C_z+2) values, node j fastest. */
const double *mcfg_boundary(const mcfg *g, int rank, int side);

#ifdef __cplusplus
}
#endif
#endif
