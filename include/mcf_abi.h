/*
 * mcf_abi.h -- C ABI of the MI355X (gfx950) per-timestep MHD update.
 *
 * The reference (3DMHD-CfCA/MCF3DSGincfca3.f90, "M3") has no plugin layer:
 * EVOLVE (M3:507-679) calls Fortran subroutines that share module-global
 * arrays.  Each entry point below replaces one of those calls and is bound
 * from Fortran with ISO_C_BINDING (code_tohoku_amd/fdriver/mcf_abi_mod.f90)
 * and from Python with ctypes (code_tohoku_amd/mcf.py).
 *
 * Conventions: every function returns 0 on success or a negative MCF_E*
 * code; the last error text is available from mcf_last_error().  Arrays
 * are FP64 in the Fortran layout A(-1:N+2,-1:N+2,-1:N+2), x fastest.  One
 * context owns all device memory of one domain on the current HIP device;
 * calls on one context must not overlap.  Only mcf_timestep, the download
 * calls and mcf_sync wait for the device.
 */
#ifndef MCF_ABI_H
#define MCF_ABI_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  MCF_OK = 0,
  MCF_E_ARG = -1,      /* invalid argument                       */
  MCF_E_HIP = -2,      /* HIP runtime error                      */
  MCF_E_NOMEM = -3,    /* device allocation failed                */
  MCF_E_NONFINITE = -4,/* dt not finite / negative (err at Couran)*/
  MCF_E_STATE = -5     /* call out of order                      */
};

enum { MCF_PREC_FP64 = 0, MCF_PREC_MIXED = 1 };
/* MCF_COOL_KI: KI cooling with the reference's explicit/IMC update.
 * MCF_COOL_SURROGATE: the same, with the backward-Euler (IMC) cells solved
 * by the learned surrogate (see the surrogate section below).
 * MCF_COOL_CHEM: the reference's own chemistry-network Fcool (M3:2825-2881:
 * Ly-alpha, CII with fesc/LEVC2, OI, dust recombination and CO line
 * cooling; photo-electric, cosmic-ray and H2 photo-destruction heating),
 * which reads the SHIELD column densities; always FP64, also under
 * MCF_PREC_MIXED. */
enum { MCF_COOL_KI = 1, MCF_COOL_SURROGATE = 2, MCF_COOL_CHEM = 3 };

/* INPUT3D.DAT (M3:134-151), the WNM state hard-wired by INITIA
 * (M3:155-157) and the options of this implementation. */
typedef struct mcf_config {
  int np1x, np2x, np1y, np2y, np1z, np2z;
  double ql1x, ql2x, ql1y, ql2y, ql1z, ql2z;
  double vinitx1, vinitx2, vinity1, vinity2, vinitz1, vinitz2;
  double binitx1, binitx2, binity1, binity2, binitz1, binitz2;
  double cfl, facdep;
  int maxstp, nitera;
  double tfinal;
  /* x: 1 outflow (zero-gradient copy, the 1D code's BC=1), 3 periodic,
   * 4 inflow; both x ends alike.  y, z: 3 (periodic) only. */
  int bcx1, bcx2, bcy1, bcy2, bcz1, bcz2;
  /* ifthrm = 2: cooling + conduction; ifchem = 2: the chemistry network;
   * ifrad = 2: SHIELD column densities; ifgrv = 2: self-gravity (GRAVTY,
   * see mcf_gravty), which needs a cube of 2^k <= 1024 cells per axis,
   * power-of-two splits with at most 8 blocks per axis and
   * log2(max split) + 1 < k */
  int ifchem, ifthrm, ifrad, ifgrv;
  double pinit_over_kb;                      /* 8.810807d3 */
  double Hini, pini, H2ini, Heini, Hepini, Cini, COini, Cpini;
  /* MPI block split NSPLTx/y/z of the reference run to reproduce (M3:64-72):
   * the reference's MOC limiter switch reads rank-local ghost cells, so its
   * results depend on the block edges.  1 = one block on that axis. */
  int nsplitx, nsplity, nsplitz;
  int precision;                             /* MCF_PREC_*  */
  int cooling;                               /* MCF_COOL_*  */
} mcf_config;

typedef struct mcf_ctx mcf_ctx;

/* Fill *cfg with the defaults of 3DMHD-CfCA/INPUT3D.DAT and M3:155-157. */
void mcf_default_config(mcf_config *cfg);

/* INITIA (M3:114-486) minus the file I/O: validates cfg, builds the grid,
 * floors and inflow states (BBRV, BBRV_cm, DTF), allocates device memory. */
int mcf_init(const mcf_config *cfg, mcf_ctx **ctx);
/* Frees the context (DEALLOCATE at M3:101-104). */
int mcf_finalize(mcf_ctx *ctx);

/* Restart read (M3:454-465) followed by CC(4) (M3:470): copies the 8 U
 * components (U(:,:,:,1:8), contiguous) and the 8 species arrays
 * (ndH, ndp, ndH2, ndHe, ndHep, ndC, ndCO, ndCp; contiguous, restart order)
 * to the device and rebuilds the cell-centred field Bcc.  Then, as INITIA
 * does (M3:472-478), zeroes the column densities and CII optical depths and
 * runs SHIELD 20 times when ifrad = 2. */
int mcf_upload(mcf_ctx *ctx, const double *U8, const double *species8);
/* The column densities of module chmvar (M3:29): Ntot, NH2, NnC, NCO,
 * tCII, each seen from the left x end then the right one (10 arrays,
 * contiguous, (-1:N+2)^3 layout; interior cells hold values).  The
 * reference keeps them in memory only; MCF_E_STATE when none is carried
 * (chemistry Fcool, network and SHIELD all off). */
int mcf_download_columns(mcf_ctx *ctx, double *col10);
/* SAVEU (M3:684-757) without the file: copies the state back. */
int mcf_download(mcf_ctx *ctx, double *U8, double *species8);
/* Cell-centred field Bcc(:,:,:,1:3) (contiguous). */
int mcf_download_bcc(mcf_ctx *ctx, double *bcc3);

/* Inhomogeneous x-inflow density: the REAL*4 field DTF that INITIA reads
 * from DTF/D{rank}.dat (M3:402-431), given as nx*ny*nz float32 values over
 * the interior, x fastest (the rank files in global order).  The inflow
 * ghost density of every fill is then DTF at the ghost's position shifted by
 * -v t and wrapped periodically in x (M3:818-826, 854-861) instead of the
 * uniform dinit1 of the shipped reference, which skips the read (M3:404).
 * NULL restores the uniform value.  Values must be finite and positive. */
int mcf_set_inflow_density(mcf_ctx *ctx, const float *dtf);

/* Simulation clock and sweep-order counter (time, ifEVO of EVOLVE). */
int mcf_set_time(mcf_ctx *ctx, double time, int ifevo);
int mcf_get_time(mcf_ctx *ctx, double *time, int *ifevo);

/* Couran (M3:1912-1930) + Stblty (M3:2784-2822) + the dt reduction and
 * clips of EVOLVE (M3:567-604): dt = min(tfinal, CFL*tCFL, tLMT), clipped
 * to tfinal and tsave.  *tag = 1 (CFL bound, 'FastSpd') or 2 ('Conduct'). */
int mcf_timestep(mcf_ctx *ctx, double tsave, double *dt, int *tag);
/* SOURCE(dt) (M3:2448-2780): SHIELD (ifrad = 2), the two-stage cooling +
 * saturated conduction (ifthrm = 2) with the configured Fcool, then the
 * chemistry network (ifchem = 2). */
int mcf_source(mcf_ctx *ctx, double dt);
/* MHD(x|y|z) (M3:1264-1324): RHS + MOC + REMAP sweep, dir 0/1/2. */
int mcf_sweep(mcf_ctx *ctx, int dir, double dt);
/* CC(1), CC(2)+CCT x3, CC(3), CC(4) (M3:638-649). */
int mcf_ct(mcf_ctx *ctx, double dt);
/* DISSIP (M3:3143-3173). */
int mcf_dissip(mcf_ctx *ctx);
/* One EVOLVE inner-loop body (M3:607-656): SOURCE(dt/2), three sweeps in
 * the ifEVO order, CT, SOURCE(dt/2), DISSIP, time += dt. */
int mcf_step(mcf_ctx *ctx, double dt);
/* mcf_timestep + mcf_step repeated nsteps times (the EVOLVE inner loop,
 * M3:562-657, with the tsave/tfinal stops).  dts may be NULL. */
int mcf_run(mcf_ctx *ctx, int nsteps, double tsave, double *dts, int *ndone);

/* GRAVTY(dt,mode) of 3DMHD-CfCA/selfG_MPIfull8.f90:1-100 (ifgrv = 2):
 * mode 2 solves Poisson's equation lap Phi = 4 pi G rho (multigrid, with
 * the free-space potential of the y-z Fourier modes on the x ends, on the
 * reference's block split) into Phi, after filling the density ghosts;
 * mode 3 kicks the velocities by -dt/2 grad Phi (fourth-order
 * differences); mode 1 (the reference's pointer set-up) does nothing.
 * mcf_upload runs mode 2 once, as INITIA does (M3:481-484), and mcf_step
 * calls mode 3 before the first SOURCE and modes 2, 3 after the second
 * (M3:607, 653).  MCF_E_STATE when self-gravity is off. */
int mcf_gravty(mcf_ctx *ctx, double dt, int mode);
/* The potential Phi (M3:37), (-1:N+2)^3 layout: interior cells and face
 * ghosts hold values. */
int mcf_download_phi(mcf_ctx *ctx, double *phi);

/* Per-kernel timing: while enabled, every kernel launch of the context is
 * bracketed by HIP events on the context's stream (enabling clears what was
 * recorded before).  mcf_profile_report waits for the device, then returns
 * one entry per kernel name with the number of launches, the summed event
 * time and the summed bytes of the arrays the kernel touches (each array it
 * reads or writes counted once per element of its launch).
 * mcf_profile_stages returns one entry per stage of the step ("dt",
 * "source", "sweep_x", "sweep_y", "sweep_z", "ct", "dissip", "upload"):
 * launches = runs of the stage, ms = summed event time of all its launches
 * (ghost fills included), bytes = the stage's compulsory HBM bytes (each
 * persistent array it must read or write, once per interior cell) times its
 * runs.  No reference counterpart (the reference only keeps the wall clock
 * of M3:556,663-666). */
typedef struct mcf_kstat {
  char name[48];
  int launches;
  double ms;
  double bytes;
} mcf_kstat;
int mcf_profile(mcf_ctx *ctx, int enable);
int mcf_profile_report(mcf_ctx *ctx, mcf_kstat *out, int maxn, int *n);
int mcf_profile_stages(mcf_ctx *ctx, mcf_kstat *out, int maxn, int *n);

/* Streams nwords doubles from one device buffer to another reps times with
 * a plain copy kernel (k_copy) and returns the mean time per copy: the
 * achievable-bandwidth and hardware-counter reference for the roofline.
 * No reference counterpart. */
int mcf_copy_probe(mcf_ctx *ctx, long nwords, int reps, double *ms);

/* ---------------------------------------------------------------------
 * KI-cooling surrogate.  No counterpart in the reference: it replaces the
 * IMC bisection (M3:2898-2921) of SOURCE's backward-Euler cells
 * (dt > 0.2 P/(gammi1 |Lambda|), M3:2480-2500) when cfg.cooling is
 * MCF_COOL_SURROGATE.  An MLP 3 -> 64 -> 64 -> 1 (SiLU; BF16 MFMA operands,
 * FP32 accumulation) maps (log10 N, log10 T, log10 dt/t_cool) to
 * log10 T_end/T; two Newton steps on the backward-Euler equation refine it
 * and a cell whose relative residual exceeds 1e-4 is solved by IMC.
 * Parameters: MCF_MLP_NPARAM floats, W1[64][3] b1[64] W2[64][64] b2[64]
 * W3[64] b3.  Features and target are scaled as in mcf_cool_samples.
 * --------------------------------------------------------------------- */
#define MCF_MLP_NPARAM 4481
typedef struct mcf_train_opts {
  int64_t nsamples;   /* training samples drawn from IMC */
  int64_t nval;       /* held-out samples */
  uint64_t seed;
  int batch;          /* multiple of 64 */
  int epochs;
  float lr, lr_min;   /* Adam, cosine decay from lr to lr_min */
} mcf_train_opts;
/* Gaussian initial weights (std 1/sqrt(fan_in)), zero biases. */
int mcf_mlp_init(mcf_ctx *ctx, uint64_t seed);
int mcf_mlp_set_weights(mcf_ctx *ctx, const float *w);
int mcf_mlp_get_weights(mcf_ctx *ctx, float *w);
/* Generates o->nsamples IMC samples on the device and trains the current
 * weights (initialised first if none are loaded).  epoch_loss / val_loss
 * (NULL or o->epochs floats) receive the mean training loss and the
 * held-out loss after each epoch. */
int mcf_mlp_train(mcf_ctx *ctx, const mcf_train_opts *o, float *epoch_loss, float *val_loss);
/* Forward pass: X[n][3] scaled features -> Y[n] (host arrays). */
int mcf_mlp_forward(mcf_ctx *ctx, const float *X, int64_t n, float *Y);
/* Mean-squared-error loss over (X, T) and its gradient with respect to the
 * MCF_MLP_NPARAM parameters (backward pass of the training kernels). */
int mcf_mlp_grad(mcf_ctx *ctx, const float *X, const float *T, int64_t n, float *grad, float *loss);
/* n training samples: log-uniform N in [0.1,1e4] cm^-3, T in [10,1e4] K,
 * dt/t_cool in [1e-3,1e3] with gammi1 = 2/3 and n = N.  X[i][0..2] =
 * ((log10 N - 1.5)/2.5, (log10 T[K] - 2.5)/1.5, log10(dt/t_cool)/3);
 * Y[i] = log10(P_IMC/P_old)/3. */
int mcf_cool_samples(mcf_ctx *ctx, int64_t n, uint64_t seed, float *X, float *Y);
/* Backward-Euler solves P + gammi1*Lambda(rho, P/(kb n))*dt = P0 for n
 * cells (host arrays): mode 0 IMC, mode 1 surrogate + Newton + residual
 * test with IMC fallback (flag[i] = 1 where IMC ran; flag may be NULL).
 * *ms = kernel time. */
int mcf_cool_solve(mcf_ctx *ctx, const double *P0, const double *n, const double *dt, const double *rho,
                   const double *gammi1, int64_t ncell, int mode, double *P, int32_t *flag, double *ms);
/* Allocation-free variant of mcf_cool_solve.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * All array arguments are DEVICE pointers: P0, n, dt, rho, gammi1 (ncell
 * doubles each) in, P (ncell doubles) and flag (ncell int32, or NULL) out.
 * The call only enqueues kernels on the context's stream (mcf_stream) and
 * returns; it calls no hipMalloc, hipFree or synchronisation.  The queue of
 * cells the surrogate rejects lives in a library-owned device array of
 * MCF_NOALLOC_MAX_CELLS entries (8 MiB + 256 B on each GPU the library
 * runs on, held until the process ends); ncell above that is refused with
 * MCF_E_ARG.  Every call on one GPU shares that array and the last call's
 * contents stay in it: calls that overlap on one GPU (two streams, two
 * contexts) race. */
#define MCF_NOALLOC_MAX_CELLS 2097152
int mcf_cool_solve_noalloc(mcf_ctx *ctx, const double *P0, const double *n, const double *dt, const double *rho,
                           const double *gammi1, int64_t ncell, int mode, double *P, int32_t *flag);
/* counts[0] = backward-Euler cells solved in SOURCE since the last reset
 * (both cooling modes), counts[1] = of those, the cells the surrogate
 * rejected and IMC solved (surrogate mode). */
int mcf_cool_stats(mcf_ctx *ctx, int64_t *counts, int reset);

/* Waits for all work of the context. */
int mcf_sync(mcf_ctx *ctx);
/* HIP stream the context launches on (hipStream_t), for event timing. */
void *mcf_stream(mcf_ctx *ctx);
const char *mcf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
