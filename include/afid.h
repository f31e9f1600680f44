/*
 * afid.h -- C ABI of the MI355X engine for AFiD's per-substep time advance.
 *
 * Plain C types only (pointers, sizes, doubles).  One opaque context owns all
 * device buffers, streams and FFT tables for one Rayleigh-Benard cell; host
 * arrays are borrowed only for the duration of a synchronous upload/download.
 *
 * Each entry point replaces one reference routine (AFiD 1.1, Fortran).  The
 * reference routines take no arguments and work on module globals (param.F90,
 * local_arrays); here the context plays the role of those modules.  The
 * Fortran binding (iso_c_binding) lives in afid_amd/fortran/afid_hip.F90.
 *
 * Return codes: 0 = AFID_OK; < 0 = runtime/HIP/argument error; > 0 =
 * numerical condition (see afid_strerror).  Calls on one context are not
 * re-entrant; they are stream-ordered on the context's stream, and every
 * call that returns data to the host synchronises that stream first.
 */
#ifndef AFID_H
#define AFID_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define AFID_OK 0
#define AFID_ERR_ARG (-1)        /* invalid argument or size                 */
#define AFID_ERR_HIP (-2)        /* HIP runtime error (see afid_last_hip_error) */
#define AFID_ERR_NOMEM (-3)      /* device allocation failed                 */
#define AFID_ERR_FFTSIZE (-4)    /* lateral size has a prime factor > 5 (errorcode 444) */
#define AFID_ERR_STATE (-5)      /* call not valid in the context's state    */
#define AFID_ERR_IO (-6)         /* continuation file missing, unreadable or of another
                                    shape (afid_io.h) */
#define AFID_ERR_GRID (-7)       /* continuation grid that cannot be interpolated from
                                    (ReadFlowField.F90:72-82, interp.F90:254-299) */
#define AFID_NUM_PIVOT 1         /* an implicit factorisation would need a row
                                    interchange (dgttrf pivoting) */
#define AFID_NUM_SINGULAR 2      /* singular pressure tridiagonal (zgttrf info>0),
                                    SolvePressureCorrection.F90:130-134 */

typedef struct afid_ctx afid_ctx;

/* bou.in subset, ReadInputFile.F90:16-36 (names as in param.F90:8-19). */
typedef struct afid_params {
  int nxm, nym, nzm;   /* cells: wall-normal x, periodic y, periodic z        */
  int nsst;            /* >1: RK3 (3 substeps), else AB2 (InitTimeMarchScheme) */
  int istr3;           /* x clustering: 0 uniform, 4 tanh, 6 clipped Chebyshev */
  double str3;         /* clustering strength                                  */
  double alx3;         /* cell height                                          */
  double ylen, zlen;   /* periodic extents                                     */
  double ray, pra;     /* Rayleigh, Prandtl                                    */
  int inslws, inslwn;  /* 1 no-slip, 0 stress-free: lower / upper plate        */
  double tempbp, temptp; /* plate temperatures, SetTempBCs.F90:15-20 (1, 0)   */
  int precision;       /* AFID_PREC_FP64 (the reference's) or AFID_PREC_MIXED  */
} afid_params;

/* Precision modes.  MIXED (no counterpart in the reference, which is FP64
 * throughout): the explicit-term histories rux..rutemp are stored in FP32 and
 * the y/z FFTs of the pressure solve run on an FP32 spectrum; the fields, the
 * wall-normal factorisations and solves, the pressure x-solve arithmetic and
 * all reductions stay FP64, and each step ends with one refinement sweep: the
 * divergence left after the last correction, computed in FP64, is solved for
 * and corrected once more.  The per-routine ImplicitAndUpdate entry points
 * return AFID_ERR_STATE in this mode. */
#define AFID_PREC_FP64 0
#define AFID_PREC_MIXED 1

/* Optional caller-computed 1-D metrics, Fortran 1-based arrays as declared in
 * param.F90:24-48 (pass the address of element 1).  Any NULL member is
 * computed by the library with CreateGrid.F90 / InitPressureSolver.F90. */
typedef struct afid_metrics {
  const double *xc, *xm, *g3rc, *g3rm, *udx3c, *udx3m;      /* length nx  */
  const double *ap3ck, *ac3ck, *am3ck, *ap3sk, *ac3sk, *am3sk;
  const double *ap3ssk, *ac3ssk, *am3ssk;                    /* length nx  */
  const double *amphk, *acphk, *apphk;                       /* length nxm */
  const double *ak1, *ak2;                                   /* nzm, nym   */
} afid_metrics;

/* Field identifiers for upload/download (local_arrays, param.F90:79-87). */
enum afid_field {
  AFID_VX = 0, AFID_VY, AFID_VZ, AFID_TEMP, AFID_PR,
  AFID_RUX, AFID_RUY, AFID_RUZ, AFID_RUTEMP,
  AFID_DPH,   /* dph(1:nxm,..): ExplicitTermsVY output, then pressure correction */
  AFID_QCAP, AFID_DQ, AFID_HRO, AFID_RHS,
  AFID_NFIELDS
};

/* ---- lifetime (InitVariables.F90, CreateGrid.F90, InitPressureSolver.F90,
 *      InitTimeMarchScheme.F90, SetTempBCs.F90) --------------------------- */
int afid_create(afid_ctx **ctx, const afid_params *p, const afid_metrics *m);
int afid_destroy(afid_ctx *ctx);
const char *afid_strerror(int code);
int afid_last_hip_error(afid_ctx *ctx);
/* Sizes: out[0..]= nx, nxm, nym, nzm, device x pitch, bytes of device memory held. */
int afid_get_dims(afid_ctx *ctx, int64_t *out6);
/* 1-D metric copies (1-based name as in param.F90), n values into out; the
 * scalars dx, dy, dz, ren, pec, al, ga, ro, dt, beta, time and the bou.in
 * values ylen, zlen, alx3, str3, istr3, ray, pra come back with n = 1. */
int afid_get_metric(afid_ctx *ctx, const char *name, double *out, int n);

/* ---- host <-> device (replaces the module arrays as the data exchange) ----
 * Host element (k,j,i), k in 1..n1 (n1 = nx, or nxm for AFID_DPH),
 * j in 1..nym, i in 1..nzm, is host[(k-1) + ld1*((j-j0) + ld2*(i-i0))]:
 * a reference halo field (1:nx,0:nym+1,0:nzm+1) is ld1=nx, ld2=nym+2,
 * j0=i0=0; a plain field (1:nx,1:nym,1:nzm) is ld1=nx, ld2=nym, j0=i0=1.
 * Downloads into a halo layout also fill the periodic halos (update_halo). */
int afid_upload(afid_ctx *ctx, int field, const double *host, int ld1, int ld2, int j0, int i0);
int afid_download(afid_ctx *ctx, int field, double *host, int ld1, int ld2, int j0, int i0);
/* Seeded initial state: vx=vy=vz=pr=0, temp(1)=tempbp, temp(nx)=temptp,
 * temp(k)=1-xc(k)+(2u-1)*amp, u from splitmix64(seed, linear index). */
int afid_init_linear_noise(afid_ctx *ctx, uint64_t seed, double amp);

/* ---- per-routine entry points (one reference subroutine each) ------------
 * Substep coefficients as TimeMarcher.F90:22-30 sets them: al, ga, ro, dt and
 * beta = dt/ren/2.  afid_set_substep(ctx, ns, dt) takes them from the scheme. */
int afid_set_coefficients(afid_ctx *ctx, double al, double ga, double ro, double dt, double beta);
int afid_set_substep(afid_ctx *ctx, int ns, double dt);
int afid_explicit_terms_vx(afid_ctx *ctx);        /* ExplicitTermsVX.F90:11-109   -> qcap */
int afid_explicit_terms_vy(afid_ctx *ctx);        /* ExplicitTermsVY.F90:11-105   -> dph  */
int afid_explicit_terms_vz(afid_ctx *ctx);        /* ExplicitTermsVZ.F90:11-109   -> dq   */
int afid_explicit_terms_temp(afid_ctx *ctx);      /* ExplicitTermsTemp.F90:11-100 -> hro  */
int afid_implicit_and_update_vx(afid_ctx *ctx);   /* ImplicitAndUpdateVX.F90:14-80 + SolveImpEqnUpdate_X.F90 */
int afid_implicit_and_update_vy(afid_ctx *ctx);   /* ImplicitAndUpdateVY.F90:14-78 + SolveImpEqnUpdate_YZ.F90 */
int afid_implicit_and_update_vz(afid_ctx *ctx);   /* ImplicitAndUpdateVZ.F90:14-75 + SolveImpEqnUpdate_YZ.F90 */
int afid_implicit_and_update_temp(afid_ctx *ctx); /* ImplicitAndUpdateTemp.F90:13-77 + SolveImpEqnUpdate_Temp.F90 */
int afid_calc_local_divergence(afid_ctx *ctx);    /* CalcLocalDivergence.F90:12-44 -> dph */
int afid_solve_pressure_correction(afid_ctx *ctx);/* SolvePressureCorrection.F90:12-154, dph in place */
int afid_correct_velocity(afid_ctx *ctx);         /* CorrectVelocity.F90:11-48 (uses dph) */
int afid_correct_pressure(afid_ctx *ctx);         /* CorrectPressure.F90:11-55 (uses dph) */
int afid_update_halo(afid_ctx *ctx, int field);   /* halo.F90 update_halo: no-op, the
                                                     device applies periodic wrap */
/* Stages of SolvePressureCorrection on the current dph, for inspection: runs
 * stage 1 (y r2c FFT, :98), 2 (+ z forward FFT, :102), 3 (+ x tridiagonal
 * solves incl. the 1/(nym nzm) scaling, :105-141) or 4 (+ z backward FFT,
 * :145) and copies the spectrum cz1(k, m, n) into out as interleaved complex,
 * Fortran order (nxm, nym/2+1, nzm).  dph itself is left unchanged.
 * AFID_ERR_STATE in the mixed-precision mode (FP32 spectrum). */
int afid_pressure_spectrum(afid_ctx *ctx, int stage, double *out);

/* ---- fused time advance (TimeMarcher.F90:13-78) -----------------------------
 * One substep / one full TimeMarcher call with fused kernels: explicit terms,
 * implicit solves and updates in one pass; divergence + y-FFT; z-FFT; x
 * tridiagonals; inverse FFTs; corrections.  qcap/dq/hro/rhs are not kept.
 * With nsst = 3, substep 3 does not store its explicit terms in rux..rutemp
 * (no later substep reads them: the next step's substep 1 has ro = 0), so
 * after a full RK3 step those arrays hold substep 2's; the per-routine
 * ImplicitAndUpdate* entries always store them. */
int afid_substep(afid_ctx *ctx, int ns, double dt);
int afid_time_marcher(afid_ctx *ctx, double dt);

/* main.F90:169-200 step body: CalcMaxCFL, dt selection (idtv=1: CFL-limited,
 * dtmin for ntime<=1, clipped at dtmax; idtv=0: fixed dt), TimeMarcher,
 * time += dt.  *errorcode receives 0, 165 (CFL too high) or 166 (dt<dtmin). */
int afid_step(afid_ctx *ctx, int ntime, int idtv, double dt, double dtmin, double dtmax,
              double limitcfl, double *dt_used, int *errorcode);
/* main.F90:167-200 for nsteps consecutive steps (ntime = ntime0, ntime0+1, ...)
 * without host round trips: CalcMaxCFL, dt selection, the substeps with their
 * factorisations and time += dt all run on the device from one captured graph
 * per step (the first call captures it).  A step that raises errorcode 165 or
 * 166 still completes, as in main.F90, and the remaining steps are skipped on
 * the device.  Returns the dt of the last step run, its errorcode and the
 * number of steps run.  AFID_NUM_PIVOT if a factorisation would need a row
 * interchange. */
int afid_steps(afid_ctx *ctx, int ntime0, int nsteps, int idtv, double dt, double dtmin, double dtmax,
               double limitcfl, double *dt_last, int *errorcode, int *steps_done);
/* CalcMaxCFL (CalcMaxCFL.F90:11-48, with its 1e-8 floor) of the fields the
 * last device step left, as the device step control computed it for the next
 * step when the context was created with AFID_CFL=fold in the environment
 * (FP64 only): the step graph then folds the reduction into its last pressure
 * correction instead of reducing the fields at the start of each step.
 * Equals afid_calc_max_cfl of the same fields.  AFID_ERR_STATE before any
 * device step and without the fold. */
int afid_steps_cfl(afid_ctx *ctx, double *cflm);

/* ---- concurrent cells (batched sweeps, serving) --------------------------------
 * Each context has its own stream and its own captured step graph, so the
 * steps of several cells overlap on the GPU.  No counterpart in the
 * reference, which runs one cell per job. */
/* afid_steps without the final wait: sets the device step state as afid_steps
 * does, queues nsteps steps and returns. */
int afid_steps_async(afid_ctx *ctx, int ntime0, int nsteps, int idtv, double dt, double dtmin, double dtmax,
                     double limitcfl);
/* Queues nsteps more steps that continue the device step state (time, ntime,
 * dt, halt). */
int afid_steps_more(afid_ctx *ctx, int nsteps);
/* *busy = 1 while queued work of this context has not finished. */
int afid_steps_query(afid_ctx *ctx, int *busy);
/* Waits for the queued steps; results as afid_steps (steps_done counts from
 * the last afid_steps_async). */
int afid_steps_wait(afid_ctx *ctx, double *dt_last, int *errorcode, int *steps_done);
/* Queues nsteps more steps on each of ncells contexts, step n of every cell
 * before step n+1 of any; each cell must have had afid_steps_async. */
int afid_batch_steps(afid_ctx *const *cells, int ncells, int nsteps);
/* Autotuner for the stencil pass: times z planes per block in
 * {8, 16, 32, 64} (not above nzm) on this context and keeps the fastest for
 * its later launches; *zm receives the choice.  It
 * writes the histories and scratch while timing and clears them afterwards:
 * call it before setting the initial state.  No counterpart in the
 * reference. */
int afid_tune(afid_ctx *ctx, int *zm);
/* Device bytes a context for p holds at most (fields, spectra, factor rows,
 * device step state, staging, statistics, nine slab planes); for admission
 * by memory budget. */
int afid_estimate_bytes(const afid_params *p, int64_t *bytes);
int afid_get_time(afid_ctx *ctx, double *time);
int afid_set_time(afid_ctx *ctx, double time);      /* time of a continuation (ReadFlowField.F90:46) */
/* The stale row rhs(1,j,i) the x-solves read (SolveImpEqnUpdate_X.F90:44):
 * plane[j + nym*i], j < nym, i < nzm.  Part of the state a restart needs to
 * continue bit for bit. */
int afid_get_stale(afid_ctx *ctx, double *plane);
int afid_set_stale(afid_ctx *ctx, const double *plane);

/* ---- diagnostics (FP64 device reductions) ----------------------------------- */
int afid_calc_max_cfl(afid_ctx *ctx, double *cflm);          /* CalcMaxCFL.F90:11-48 */
int afid_check_divergence(afid_ctx *ctx, double *qmax);      /* CheckDivergence.F90:11-48 */
int afid_calc_plate_nu(afid_ctx *ctx, double *nu_bot, double *nu_top); /* CalcPlateNu.F90:11-57 */
/* GlobalQuantities.F90:11-102: out[0]=Nu_vol, [1]=tempm, [2]=tempmax, [3]=tempmin,
 * [4..6]=vmax(1..3) (vz,vy,vx), [7]=vz_rms, [8]=vy_rms, [9]=vx_rms, [10]=vzvyvx_rms */
int afid_global_quantities(afid_ctx *ctx, double *out11);
/* CalcDissipationNu.F90:12-120: Nu_eps_u, Nu_eps_T and this call's per-plane
 * contributions disste/dissth (nxm values each; either may be NULL). */
int afid_calc_dissipation_nu(afid_ctx *ctx, double *nu_mu, double *nu_th, double *disste,
                             double *dissth);

/* ---- statistics and slabs (StatRoutines.F90, SlabDumpRoutines.F90) ---------- */
/* CalcStats (StatRoutines.F90:12-44): adds this sample's plane averages over
 * (j, i) of vx, vy, vz, temp, their squares and temp*vx into the context's
 * statistics (device reduction); nstatsamples += 1. */
int afid_calc_stats(afid_ctx *ctx);
/* The accumulated statistics, out[q*nxm + k-1] for q = vx_mean, vy_mean,
 * vz_mean, temp_mean, vx_rms, vy_rms, vz_rms, temp_rms, tempvx_mean, and the
 * dissipation profiles disste, dissth summed over the afid_calc_dissipation_nu
 * calls (CalcDissipationNu.F90:101-104): 11*nxm values; *nsamples receives
 * the number of afid_calc_stats calls. */
int afid_get_stats(afid_ctx *ctx, double *out, int *nsamples);
int afid_reset_stats(afid_ctx *ctx);
/* No counterpart in the reference.  afid_calc_stats for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): it allocates nothing and adds this
 * sample's plane averages (the nine quantities of afid_calc_stats, in the
 * order of afid_get_stats' first 9*nxm values) into acc, a device array of
 * 9*nxm doubles that the caller owns and clears.  The context's own
 * statistics and sample count are not touched.  The block partials use a
 * static device array of AFID_NOALLOC_SCRATCH_BYTES per GPU, shared by every
 * call on that GPU and holding the last call's contents: calls that overlap
 * on one GPU (contexts on different streams) race.  AFID_ERR_ARG when the
 * partials, 256*9*ldx doubles with ldx = 16*ceil((nxm+1)/16), exceed that
 * size (nxm >= 896).  Stream-ordered on the context's stream; returns
 * without synchronising. */
#define AFID_NOALLOC_SCRATCH_BYTES (16u << 20)
int afid_calc_stats_noalloc(afid_ctx *ctx, double *acc);
/* SlabDumper (SlabDumpRoutines.F90:20-31): the planes k = kslab[m] (1-based,
 * m < nslab) of (vx(k)+vx(k+1))/2, vy, vz and temp, as
 * out[((m*4 + q)*nzm + i-1)*nym + j-1] (4*nslab*nym*nzm values). */
int afid_get_slabs(afid_ctx *ctx, const int *kslab, int nslab, double *out);

/* ---- restart regridding (interp.F90, ReadFlowField.F90:56-158) -------------- */
/* interp (interp.F90:11-208): sets field (AFID_VX, AFID_VY, AFID_VZ or
 * AFID_TEMP; intvar 1..4) on this grid from the field of an old grid with nxo,
 * nyo, nzo points (continua_master.h5's nx, ny, nz) and stretching istro3,
 * stro3.  old is the continuation array, Fortran (nxo, nyo-1, nzo-1).  The old
 * window is ReadFlowField's at one rank (:91-99); the trilinear weights and
 * the corner values are those of interp.F90:128-207 as written.  Halo edges
 * and corners of the old array, which the reference leaves unset, read as 0.
 * AFID_ERR_GRID: more than twice the old resolution (ReadFlowField.F90:72-82),
 * a stretching other than 0 or 6 (gridnew handles no other), or a point whose
 * bracket falls outside the old window (the reference would read out of
 * bounds). */
int afid_interp_field(afid_ctx *ctx, int field, const double *old, int nxo, int nyo, int nzo, int istro3,
                      double stro3);

/* ---- profiling ----------------------------------------------------------------
 * When enabled, every kernel launch is bracketed by HIP events on the
 * context's stream; totals are per kernel name.  Device steps (afid_steps and
 * the asynchronous forms) then replay a second capture of the step graph whose
 * kernels are bracketed by event-record nodes, one step at a time, and wait
 * for each step to read its events: the kernels are those of the timed steps,
 * the host no longer runs ahead of the device. */
int afid_profile_enable(afid_ctx *ctx, int on);
/* Writes up to cap entries: names (NUL-separated into names_buf of size nbuf),
 * total milliseconds and launch counts.  Returns the number of kernels. */
int afid_profile_read(afid_ctx *ctx, char *names_buf, int nbuf, double *total_ms, int64_t *count,
                      int cap);
int afid_profile_reset(afid_ctx *ctx);
int afid_synchronize(afid_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
