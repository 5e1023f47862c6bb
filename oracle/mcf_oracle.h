/*
 * mcf_oracle.h -- CPU restatement (plain C, FP64) of one timestep of the
 * MCF3DSG* 3D MHD code, used ONLY as the checker for the HIP path.
 *
 * TEST INFRASTRUCTURE.  Nothing in the product links or calls this library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and only as the reference the HIP results are compared against (or as
 * the timed CPU baseline).
 *
 * Reference: 3DMHD-CfCA/MCF3DSGincfca3.f90 (below "M3"), with the
 * Koyama-Inutsuka cooling function of 1DMHDKIcool/1DMHDwCC3KIcool.f90:1146-1173
 * ("K1") evaluated with N = rho/1.27 as in
 * 3DMHD-feedback/MCF3DSGincfca3-feedback.f90:3586.
 *
 * The whole domain is one block: the MPI halo exchanges of BC_MPI /
 * BC_MPI_OT (M3:800-1254) become in-place ghost fills with the same
 * pass order (x, then y, then z) and the same physical x-boundary rules.
 * Arrays are (nx+4)*(ny+4)*(nz+4) doubles, x fastest, index -1..n+2 per axis,
 * exactly the Fortran U(-1:N+2,-1:N+2,-1:N+2) layout.
 */
#ifndef MCF_ORACLE_H
#define MCF_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* INPUT3D.DAT (M3:134-151) plus the WNM state hard-wired in INITIA (M3:155-157). */
typedef struct {
  int np1x, np2x, np1y, np2y, np1z, np2z;
  double ql1x, ql2x, ql1y, ql2y, ql1z, ql2z;
  double vinitx1, vinitx2, vinity1, vinity2, vinitz1, vinitz2;
  double binitx1, binitx2, binity1, binity2, binitz1, binitz2;
  double cfl, facdep;
  int maxstp, nitera;
  double tfinal;
  int bcx1, bcx2, bcy1, bcy2, bcz1, bcz2;
  int ifchem, ifthrm, ifrad, ifgrv;
  double pinit_over_kb;                 /* 8.810807d3 */
  double Hini, pini, H2ini, Heini, Hepini, Cini, COini, Cpini;
  /* MPI block split of the reference run being reproduced (NSPLTx/y/z,
   * M3:64-72).  The reference's MOC limiter reads U(:,:,:,5) in rank-local
   * ghost cells, which still hold the pre-sweep pressure while interior cells
   * hold the specific energy, so the result depends on the block edges;
   * DISSIP treats only the right x end when NSPLTx = 1 (M3:3156-3158). */
  int nsplitx, nsplity, nsplitz;
  /* cooling function: 1 = the Koyama-Inutsuka form (K1:1146-1173) the
   * KI-cooling configurations substitute; 0 = the reference's own
   * chemistry-network form (M3:2825-2881, with fesc and LEVC2), which reads
   * the column densities SHIELD maintains (ifrad = 2). */
  int kicool;
} mcfo_params;

/* Field ids for mcfo_field(): 1..8 = U(:,:,:,1..8); 9..16 = species in the
 * restart-file order ndH, ndp, ndH2, ndHe, ndHep, ndC, ndCO, ndCp;
 * 17..19 = Bcc(:,:,:,1..3). */
enum { MCFO_NFIELD = 19 };

typedef struct mcfo mcfo;

void    mcfo_default_params(mcfo_params *p);
mcfo   *mcfo_create(const mcfo_params *p);
void    mcfo_destroy(mcfo *o);
/* Threads this instance splits its loops over (default: OpenMP's maximum;
 * nthreads < 1 restores it).  Results do not depend on it.  Returns the
 * count, or -1 if the work arrays cannot be allocated. */
int     mcfo_set_threads(mcfo *o, int nthreads);
/* Inhomogeneous x-inflow density DTF (M3:402-431, 818-826, 854-861):
 * nx*ny*nz float32 values over the interior, x fastest (the reference's
 * DTF/D{rank}.dat files in global order).  NULL restores the uniform dinit1
 * of the shipped reference, which skips the read (M3:404).  Returns 0, or
 * -1 if the copy cannot be allocated. */
int     mcfo_set_dtf(mcfo *o, const float *interior);
double *mcfo_field(mcfo *o, int id);
double *mcfo_grid(mcfo *o, int axis, int spacing); /* x/dx arrays, index -1..n+2 */
void    mcfo_get_time(const mcfo *o, double *time, int *ifevo);
void    mcfo_set_time(mcfo *o, double time, int ifevo);
void    mcfo_bbrv(const mcfo *o, double *bbrv /* 10*2*2, Fortran order */, double *bbrv_cm /* 8 */, double *dtf);

/* INITIA's set-up of the radiation state after a restart read (M3:471-478):
 * column densities and CII optical depths zeroed, then 20 SHIELD sweeps when
 * ifrad = 2.  Call after the state is set. */
void    mcfo_init_columns(mcfo *o);
/* Column density array q (0 Ntot, 1 NH2, 2 NnC, 3 NCO, 4 tCII) seen from the
 * left (dir 0) or the right (dir 1) end of x; NULL when not carried. */
double *mcfo_column(mcfo *o, int q, int dir);
/* Chemistry-form Fcool at cell (i,j,k) and temperature T (units of 10^3 K). */
double  mcfo_fcool_cell(mcfo *o, int i, int j, int k, double T);

/* Self-gravity (ifgrv = 2; mcf_grav.c).  mcfo_grav_solve is GRAVTY(dt,2):
 * fills the density ghosts and solves for Phi (INITIA calls it once after
 * the state is read, M3:481-484); returns -1 without gravity.
 * mcfo_grav_kick is GRAVTY(dt,3).  mcfo_phi: Phi, the layout of U (NULL
 * without gravity).  mcfo_create returns NULL for a grid the solver does
 * not take (power-of-two cells per axis, power-of-two splits). */
int     mcfo_grav_solve(mcfo *o);
void    mcfo_grav_kick(mcfo *o, double dt);
double *mcfo_phi(mcfo *o);

/* Pieces of the step (same names as the reference routines). */
void   mcfo_cc4(mcfo *o);                           /* CC(4): Bcc from faces + halo   */
double mcfo_couran(mcfo *o);                        /* Couran  M3:1912-1930           */
double mcfo_stblty(mcfo *o);                        /* Stblty  M3:2784-2822           */
double mcfo_timestep(mcfo *o, double tsave, int *st);/* EVOLVE M3:567-604 dt logic     */
void   mcfo_source(mcfo *o, double dt);             /* SOURCE  M3:2448-2572 (ifthrm=2)*/
void   mcfo_sweep(mcfo *o, int dir, double dt);     /* MHD     M3:1264-1324           */
void   mcfo_ct(mcfo *o, double dt);                 /* CC/CCT  M3:638-649             */
void   mcfo_dissip(mcfo *o);                        /* DISSIP  M3:3143-3173           */
void   mcfo_step(mcfo *o, double dt);               /* one EVOLVE inner iteration body*/
/* EVOLVE inner loop (M3:562-657) for nitera steps, with the tsave/tfinal clips.
 * Writes the dt of each step into dts[] (may be NULL).  Returns steps done. */
int    mcfo_run(mcfo *o, int nsteps, double *dts, int *tags);

/* Known-answer entry points for single routines. */
void   mcfo_rieman(const double *in8, double gamma, double *pa, double *va);
void   mcfo_chareq(const double *in6, double *bay, double *vay);
double mcfo_fcool_ki(double rho, double T);
double mcfo_imc(double P, double n, double dt, double rho, double gammi1);
/* VLIMIT on one pencil of n+4 values q[-1..n+2] (pointer to q[-1]), with
 * T[-1..n+2] and dxx[-1..n+2]; writes grd[0..n+1] (pointer to grd[0]). */
void   mcfo_vlimit1d(int n, const double *q, const double *T, const double *dxx, double *grd);

#ifdef __cplusplus
}
#endif
#endif
