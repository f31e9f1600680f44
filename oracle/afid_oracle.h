/*
 * afid_oracle.h -- CPU oracle for the AFiD per-substep time advance.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a line-faithful C restatement of the
 * reference Fortran (AFiD 1.1) used as the parity checker for the HIP
 * engine.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product library never links or calls it.
 *
 * Arrays keep the reference's Fortran layout and index bases:
 *   halo fields  vx,vy,vz,pr,temp : (1:nx , 0:nym+1, 0:nzm+1)   param.F90:82-86,
 *                                                            InitVariables.F90:95-99
 *   dphhalo                         : (1:nxm, 0:nym+1, 0:nzm+1)
 *   plain fields rhs,rux,ruy,ruz,rutemp,qcap,dq,hro : (1:nx,1:nym,1:nzm)
 *   dph                             : (1:nxm,1:nym,1:nzm)      InitVariables.F90:105-113
 *   1-D metric arrays are 1-based (element 0 unused).
 */
#ifndef AFID_ORACLE_H
#define AFID_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct afo afo_t;

/* bou.in subset (ReadInputFile.F90:16-36).  nsst>1 selects RK3, else AB2. */
afo_t *afo_create(int nxm, int nym, int nzm, int nsst, double alx3, int istr3,
                  double str3, double ylen, double zlen, double ray, double pra,
                  int inslws, int inslwn);
void afo_destroy(afo_t *o);

/* Named array lookup: returns base pointer (double* or int*) and element count.
 * is_int is set to 1 for the integer index arrays.  NULL if unknown name. */
void *afo_array(afo_t *o, const char *name, long *count, int *is_int);
/* Named scalar lookup (dx, dy, ren, pec, beta, al, ga, ro, dt, time, ...). */
int afo_get_scalar(afo_t *o, const char *name, double *out);
int afo_set_scalar(afo_t *o, const char *name, double val);

/* Seeded initial condition: T = 1 - xc(k) + U(-1e-3,1e-3) noise, velocities 0. */
void afo_init_linear_noise(afo_t *o, unsigned long long seed, double amp);
/* Deterministic sinusoidal IC of CreateInitialConditions.F90:20-54. */
void afo_init_reference_ic(afo_t *o);

/* Routines of the substep (names mirror the reference subroutines). */
void afo_explicit_terms_vx(afo_t *o);
void afo_explicit_terms_vy(afo_t *o);
void afo_explicit_terms_vz(afo_t *o);
void afo_explicit_terms_temp(afo_t *o);
void afo_implicit_and_update_vx(afo_t *o);
void afo_implicit_and_update_vy(afo_t *o);
void afo_implicit_and_update_vz(afo_t *o);
void afo_implicit_and_update_temp(afo_t *o);
void afo_calc_local_divergence(afo_t *o);
int  afo_solve_pressure_correction(afo_t *o);  /* returns zgttrf info (0 ok) */
void afo_copy_dphhalo(afo_t *o);
void afo_correct_velocity(afo_t *o);
void afo_correct_pressure(afo_t *o);
void afo_update_halo(afo_t *o, const char *name);
void afo_update_all_halos(afo_t *o);

/* One TimeMarcher call (all nsst substeps) at the given dt. */
int afo_time_marcher(afo_t *o, double dt);
/* One substep ns (1-based) at the given dt; beta = dt/ren/2. */
int afo_substep(afo_t *o, int ns, double dt);
/* main.F90:167-200 step body: CFL, dt selection, TimeMarcher, time += dt.
 * Returns errorcode (0, 165, 166) and the dt used. */
int afo_step(afo_t *o, int ntime, int idtv, double dtfixed, double dtmin,
             double dtmax, double limitcfl, double *dt_used);

/* Diagnostics (a26-a30). */
double afo_calc_max_cfl(afo_t *o);
double afo_check_divergence(afo_t *o);
void afo_calc_plate_nu(afo_t *o, double *nu_bot, double *nu_top);
/* out[0..8]: nu_vol, tempm, tempmax, tempmin, vmax1(vz), vmax2(vy), vmax3(vx),
 *            vz_rms, vy_rms; out[9]: vx_rms; out[10]: vzvyvx_rms */
void afo_global_quantities(afo_t *o, double *out);
/* nu_mu, nu_th and this call's per-plane profiles disste/dissth (nxm each). */
void afo_calc_dissipation_nu(afo_t *o, double *nu_mu, double *nu_th,
                             double *disste, double *dissth);

/* Standalone FFT (FFTW conventions: unnormalised, forward sign -1). */
void afo_fft_c2c(int n, const double *in, double *out, int sign);   /* interleaved complex */
void afo_fft_r2c(int n, const double *in, double *out);            /* n/2+1 complex out */
void afo_fft_c2r(int n, const double *in, double *out);            /* n/2+1 complex in */
/* LAPACK-order tridiagonal kernels (netlib dgttrf/dgttrs algorithm). */
int afo_dgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv);
void afo_dgttrs(int n, int nrhs, const double *dl, const double *d, const double *du,
                const double *du2, const int *ipiv, double *b, int ldb);

void afo_set_threads(int n);

/* Restart regridding: interp (interp.F90:11-208) of one field from a grid
 * (nxo, nyo, nzo, istro3, stro3) onto (nxm+1, nym+1, nzm+1, istr3, str3),
 * with the one-rank old window of ReadFlowField.F90:64-98.  old is the
 * continuation array, Fortran (nxo, nyo-1, nzo-1); out is Fortran
 * (nxm+1, nym, nzm).  intvar 1..4 = vx, vy, vz, temp.  Returns 0, or
 * 1 more than twice the old resolution (ReadFlowField.F90:72-82),
 * 2 a stretching other than 0 / 6 or a window narrower than 2,
 * 3 a read outside the old window (the reference would read out of bounds). */
int afo_interp(const double *old, int nxo, int nyo, int nzo, int istro3, double stro3,
               int nxm, int nym, int nzm, int istr3, double str3, double ylen, double zlen,
               int intvar, double *out);
/* gridnew (interp.F90:254-299): rc[0..n-1], rm[0..n-1].  Returns 0, or 2 for
 * a stretching other than 0 / 6. */
int afo_gridnew(int n, double rext, int istr, double str, double *rc, double *rm);

#ifdef __cplusplus
}
#endif
#endif
