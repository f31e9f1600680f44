/*
 * afid_oracle.c -- CPU oracle for the AFiD per-substep time advance.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline; never shipped
 * in the product path).  Every routine restates one reference Fortran
 * routine, cited by file:line, keeping its loop bounds, index arrays and the
 * order of floating-point operations.  Third-party pieces on the path are
 * restated from their published algorithms:
 *   - LAPACK dgttrf/dgttrs/zgttrf/zgttrs (netlib reference, partial pivoting
 *     on |d|>=|dl|, CABS1 for the complex variant),
 *   - FFTW3 transform conventions (unnormalised, forward sign -1, r2c keeps
 *     n/2+1 bins, c2r ignores the imaginary parts of the DC/Nyquist bins).
 * Parity pins: see tests/test_oracle_*.py (reference Fortran routines built
 * into oracle/_ref, scipy LAPACK, numpy.fft, closed forms and invariants).
 */
#include "afid_oracle.h"
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct afo {
  int nx, ny, nz, nxm, nym, nzm, nsst, istr3, inslws, inslwn;
  double alx3, str3, ylen, zlen, ray, pra;
  double ren, pec, pi;
  double dx, dy, dz, dxq, dyq, dzq;
  double al, ga, ro, beta, dt, time;
  double gam[4], rom[4], alm[4];
  /* 1-D, 1-based */
  double *xc, *xm, *yc, *ym, *zc, *zm, *g3rc, *g3rm, *udx3c, *udx3m;
  double *ap3ck, *ac3ck, *am3ck, *ap3sk, *ac3sk, *am3sk, *ap3ssk, *ac3ssk, *am3ssk;
  double *ak1, *ao, *ak2, *ap, *amphk, *acphk, *apphk;
  int *kmc, *kpc, *kmv, *kpv;
  /* 2-D (1:nym,1:nzm) plate temperatures */
  double *tempbp, *temptp;
  /* 3-D */
  double *vx, *vy, *vz, *pr, *temp, *dphhalo;
  double *rhs, *rux, *ruy, *ruz, *rutemp, *qcap, *dq, *hro, *dph;
};

/* ---- index helpers (Fortran column-major, reference bounds) ---------------- */
#define HX(a, k, j, i) (a)[((size_t)(i) * (o->nym + 2) + (size_t)(j)) * o->nx + (size_t)(k) - 1]
#define PX(a, k, j, i) (a)[((size_t)((i) - 1) * o->nym + (size_t)((j) - 1)) * o->nx + (size_t)(k) - 1]
#define HM(a, k, j, i) (a)[((size_t)(i) * (o->nym + 2) + (size_t)(j)) * o->nxm + (size_t)(k) - 1]
#define PM(a, k, j, i) (a)[((size_t)((i) - 1) * o->nym + (size_t)((j) - 1)) * o->nxm + (size_t)(k) - 1]
#define P2(a, j, i) (a)[(size_t)((i) - 1) * o->nym + (size_t)(j) - 1]

static double *dvec(int n) { return (double *)calloc((size_t)n + 2, sizeof(double)); }
static int *ivec(int n) { return (int *)calloc((size_t)n + 2, sizeof(int)); }
static double *d3(size_t n) { return (double *)calloc(n, sizeof(double)); }

void afo_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---- InitTimeMarchScheme.F90:16-49 ------------------------------------------ */
static void init_time_march(afo_t *o) {
  if (o->nsst > 1) {
    o->gam[1] = 8.0 / 15.0; o->gam[2] = 5.0 / 12.0; o->gam[3] = 3.0 / 4.0;
    o->rom[1] = 0.0;        o->rom[2] = -17.0 / 60.0; o->rom[3] = -5.0 / 12.0;
  } else {
    o->gam[1] = 1.5; o->gam[2] = 0.0; o->gam[3] = 0.0;
    o->rom[1] = -0.5; o->rom[2] = 0.0; o->rom[3] = 0.0;
  }
  for (int ns = 1; ns <= o->nsst; ns++) o->alm[ns] = o->gam[ns] + o->rom[ns];
}

/* ---- CreateGrid.F90:11-269 --------------------------------------------------- */
static void create_grid(afo_t *o) {
  const int nx = o->nx, nxm = o->nxm, nym = o->nym, nzm = o->nzm;
  for (int kc = 1; kc <= nxm; kc++) {                       /* :27-33 */
    o->kmv[kc] = kc - 1;
    o->kpv[kc] = kc + 1;
    if (kc == 1) o->kmv[kc] = kc;
    if (kc == nxm) o->kpv[kc] = kc;
  }
  for (int kc = 1; kc <= nxm; kc++) {                       /* :34-37 */
    o->kpc[kc] = o->kpv[kc] - kc;
    o->kmc[kc] = kc - o->kmv[kc];
  }
  for (int i = 1; i <= o->nz; i++) {                        /* :41-48 */
    double x1 = (double)(i - 1) / (double)nzm;
    o->zc[i] = o->zlen * x1;
  }
  for (int i = 1; i <= nzm; i++) o->zm[i] = (o->zc[i] + o->zc[i + 1]) * 0.5;
  for (int j = 1; j <= o->ny; j++) {
    double x2 = (double)(j - 1) / (double)nym;
    o->yc[j] = o->ylen * x2;
  }
  for (int j = 1; j <= nym; j++) o->ym[j] = (o->yc[j] + o->yc[j + 1]) * 0.5;

  double *etaz = dvec(nx + 500), *etazm = dvec(nx + 500);
  if (o->istr3 == 0) {                                      /* :59-66 */
    for (int kc = 1; kc <= nx; kc++) {
      double x3 = (double)(kc - 1) / (double)nxm;
      etaz[kc] = o->alx3 * x3;
      o->xc[kc] = etaz[kc];
    }
  }
  double tstr3 = tanh(o->str3);                             /* :70-80 */
  if (o->istr3 == 4) {
    o->xc[1] = 0.0;
    for (int kc = 2; kc <= nx; kc++) {
      double z2dp = (double)(2 * kc - nx - 1) / (double)nxm;
      o->xc[kc] = (1 + tanh(o->str3 * z2dp) / tstr3) * 0.5 * o->alx3;
    }
  }
  if (o->istr3 == 6) {                                      /* :84-119 */
    int nclip = (int)o->str3;
    int nxmo = nx + nclip + nclip;
    for (int kc = 1; kc <= nxmo; kc++)
      etazm[kc] = +cos(o->pi * ((double)kc - 0.5) / (double)nxmo);
    for (int kc = 1; kc <= nx; kc++) etaz[kc] = etazm[kc + nclip];
    double delet = etaz[1] - etaz[nx];
    for (int kc = 1; kc <= nx; kc++) etaz[kc] = etaz[kc] / (0.5 * delet);
    o->xc[1] = 0.0;
    for (int kc = 2; kc <= nxm; kc++) o->xc[kc] = o->alx3 * (1.0 - etaz[kc]) * 0.5;
    o->xc[nx] = o->alx3;
  }
  free(etaz);
  free(etazm);

  o->dx = (double)nxm / o->alx3;                            /* :129-138 */
  o->dy = (double)nym / o->ylen;
  o->dz = (double)nzm / o->zlen;
  o->dxq = o->dx * o->dx;
  o->dyq = o->dy * o->dy;
  o->dzq = o->dz * o->dz;

  for (int kc = 1; kc <= nxm; kc++) {                       /* :143-153 */
    o->xm[kc] = (o->xc[kc] + o->xc[kc + 1]) * 0.5;
    o->g3rm[kc] = (o->xc[kc + 1] - o->xc[kc]) * o->dx;
  }
  for (int kc = 2; kc <= nxm; kc++) o->g3rc[kc] = (o->xc[kc + 1] - o->xc[kc - 1]) * o->dx * 0.5;
  o->g3rc[1] = (o->xc[2] - o->xc[1]) * o->dx;
  o->g3rc[nx] = (o->xc[nx] - o->xc[nxm]) * o->dx;
  for (int kc = 1; kc <= nxm; kc++) {                       /* :155-159 */
    o->udx3m[kc] = o->dx / o->g3rm[kc];
    o->udx3c[kc] = o->dx / o->g3rc[kc];
  }
  o->udx3c[nx] = o->dx / o->g3rc[nx];

  o->am3ck[1] = 0.0; o->ap3ck[1] = 0.0; o->ac3ck[1] = 1.0; /* :186-202 */
  o->am3ck[nx] = 0.0; o->ap3ck[nx] = 0.0; o->ac3ck[nx] = 1.0;
  for (int kc = 2; kc <= nxm; kc++) {
    int km = kc - 1;
    double a33 = o->dxq / o->g3rc[kc];
    double a33p = 1.0 / o->g3rm[kc];
    double a33m = 1.0 / o->g3rm[km];
    o->ap3ck[kc] = a33 * a33p;
    o->am3ck[kc] = a33 * a33m;
    o->ac3ck[kc] = -(o->ap3ck[kc] + o->am3ck[kc]);
  }
  for (int kc = 2; kc <= nxm - 1; kc++) {                   /* :209-218 */
    int kp = kc + 1;
    double a33 = o->dxq / o->g3rm[kc];
    double a33p = +a33 / o->g3rc[kp];
    double a33m = +a33 / o->g3rc[kc];
    o->ap3sk[kc] = a33p;
    o->am3sk[kc] = a33m;
    o->ac3sk[kc] = -(o->ap3sk[kc] + o->am3sk[kc]);
  }
  {                                                         /* :222-231 lower wall */
    int kc = 1, kp = kc + 1;
    double a33 = o->dxq / o->g3rm[kc];
    double a33p = +a33 / o->g3rc[kp];
    double a33m = +a33 / o->g3rc[kc];
    o->ap3sk[kc] = a33p;
    o->am3sk[kc] = 0.0;
    o->ac3sk[kc] = -(a33p + o->inslws * a33m * 2.0);
  }
  {                                                         /* :233-242 upper wall */
    int kc = nxm, kp = kc + 1;
    double a33 = o->dxq / o->g3rm[kc];
    double a33p = +a33 / o->g3rc[kp];
    double a33m = +a33 / o->g3rc[kc];
    o->am3sk[kc] = a33m;
    o->ap3sk[kc] = 0.0;
    o->ac3sk[kc] = -(a33m + o->inslwn * a33p * 2.0);
  }
  o->am3ssk[1] = 0.0; o->ap3ssk[1] = 0.0; o->ac3ssk[1] = 1.0; /* :244-266 */
  for (int kc = 2; kc <= nxm; kc++) {
    int km = kc - 1;
    double a33 = o->dxq / o->g3rc[kc];
    double a33p = 1.0 / o->g3rm[kc];
    double a33m = 1.0 / o->g3rm[km];
    o->ap3ssk[kc] = a33 * a33p;
    o->am3ssk[kc] = a33 * a33m;
    o->ac3ssk[kc] = -(o->ap3ssk[kc] + o->am3ssk[kc]);
  }
  o->am3ssk[nx] = 0.0; o->ap3ssk[nx] = 0.0; o->ac3ssk[nx] = 1.0;
}

/* ---- InitPressureSolver.F90:12-66 (without the FFT plan set-up) ------------- */
static void init_pressure_solver(afo_t *o) {
  const int nym = o->nym, nzm = o->nzm;
  int nymh = nym / 2 + 1, nymp = nymh + 1, nzmh = nzm / 2 + 1, nzmp = nzmh + 1;
  for (int i = 1; i <= nzmh; i++) o->ao[i] = (i - 1) * 2.0 * o->pi;
  for (int i = nzmp; i <= nzm; i++) o->ao[i] = -(nzm - i + 1) * 2.0 * o->pi;
  for (int i = 1; i <= nzm; i++) {
    double f = (double)nzm / o->zlen;
    o->ak1[i] = 2.0 * (1.0 - cos(o->ao[i] / nzm)) * (f * f);
  }
  for (int j = 1; j <= nymh; j++) o->ap[j] = (j - 1) * 2.0 * o->pi;
  for (int j = nymp; j <= nym; j++) o->ap[j] = -(nym - j + 1) * 2.0 * o->pi;
  for (int j = 1; j <= nym; j++) {
    double f = (double)nym / o->ylen;
    o->ak2[j] = 2.0 * (1.0 - cos(o->ap[j] / nym)) * (f * f);
  }
  for (int kc = 1; kc <= o->nxm; kc++) {
    int kp = o->kpv[kc];
    double a33icc = o->kmc[kc] * o->dxq / o->g3rc[kc];
    double a33icp = o->kpc[kc] * o->dxq / o->g3rc[kp];
    double ugmmm = 1.0 / o->g3rm[kc];
    o->amphk[kc] = a33icc * ugmmm;
    o->apphk[kc] = a33icp * ugmmm;
    o->acphk[kc] = -(o->amphk[kc] + o->apphk[kc]);
  }
}

afo_t *afo_create(int nxm, int nym, int nzm, int nsst, double alx3, int istr3,
                  double str3, double ylen, double zlen, double ray, double pra,
                  int inslws, int inslwn) {
  afo_t *o = (afo_t *)calloc(1, sizeof(afo_t));
  o->nxm = nxm; o->nym = nym; o->nzm = nzm;
  o->nx = nxm + 1; o->ny = nym + 1; o->nz = nzm + 1;        /* ReadInputFile.F90:38-40 */
  o->nsst = nsst; o->alx3 = alx3; o->istr3 = istr3; o->str3 = str3;
  o->ylen = ylen; o->zlen = zlen; o->ray = ray; o->pra = pra;
  o->inslws = inslws; o->inslwn = inslwn;
  o->ren = sqrt(ray / pra);                                 /* :44-47 */
  o->pec = sqrt(pra * ray);
  o->pi = 2.0 * asin(1.0);
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  o->zc = dvec(nz); o->zm = dvec(nz); o->ak1 = dvec(nz); o->ao = dvec(nz);
  o->yc = dvec(ny); o->ym = dvec(ny); o->ak2 = dvec(ny); o->ap = dvec(ny);
  o->xc = dvec(nx); o->xm = dvec(nx); o->g3rc = dvec(nx); o->g3rm = dvec(nx);
  o->udx3c = dvec(nx); o->udx3m = dvec(nx);
  o->ap3ck = dvec(nx); o->ac3ck = dvec(nx); o->am3ck = dvec(nx);
  o->ap3sk = dvec(nx); o->ac3sk = dvec(nx); o->am3sk = dvec(nx);
  o->ap3ssk = dvec(nx); o->ac3ssk = dvec(nx); o->am3ssk = dvec(nx);
  o->amphk = dvec(nx); o->acphk = dvec(nx); o->apphk = dvec(nx);
  o->kmc = ivec(nx); o->kpc = ivec(nx); o->kmv = ivec(nx); o->kpv = ivec(nx);
  o->tempbp = d3((size_t)nym * nzm);
  o->temptp = d3((size_t)nym * nzm);
  size_t nh = (size_t)nx * (nym + 2) * (nzm + 2);
  size_t np = (size_t)nx * nym * nzm;
  o->vx = d3(nh); o->vy = d3(nh); o->vz = d3(nh); o->pr = d3(nh); o->temp = d3(nh);
  o->dphhalo = d3((size_t)nxm * (nym + 2) * (nzm + 2));
  o->rhs = d3(np); o->rux = d3(np); o->ruy = d3(np); o->ruz = d3(np); o->rutemp = d3(np);
  o->qcap = d3(np); o->dq = d3(np); o->hro = d3(np);
  o->dph = d3((size_t)nxm * nym * nzm);
  init_time_march(o);
  create_grid(o);
  init_pressure_solver(o);
  for (int i = 1; i <= nzm; i++)                            /* SetTempBCs.F90:15-20 */
    for (int j = 1; j <= nym; j++) {
      P2(o->temptp, j, i) = 0.0;
      P2(o->tempbp, j, i) = 1.0;
    }
  return o;
}

void afo_destroy(afo_t *o) {
  if (!o) return;
  void *p[] = {o->xc, o->xm, o->yc, o->ym, o->zc, o->zm, o->g3rc, o->g3rm, o->udx3c,
               o->udx3m, o->ap3ck, o->ac3ck, o->am3ck, o->ap3sk, o->ac3sk, o->am3sk,
               o->ap3ssk, o->ac3ssk, o->am3ssk, o->ak1, o->ao, o->ak2, o->ap, o->amphk,
               o->acphk, o->apphk, o->kmc, o->kpc, o->kmv, o->kpv, o->tempbp, o->temptp,
               o->vx, o->vy, o->vz, o->pr, o->temp, o->dphhalo, o->rhs, o->rux, o->ruy,
               o->ruz, o->rutemp, o->qcap, o->dq, o->hro, o->dph};
  for (size_t n = 0; n < sizeof(p) / sizeof(p[0]); n++) free(p[n]);
  free(o);
}

void *afo_array(afo_t *o, const char *name, long *count, int *is_int) {
  const long nx = o->nx, ny = o->ny, nz = o->nz, nxm = o->nxm, nym = o->nym, nzm = o->nzm;
  const long nh = nx * (nym + 2) * (nzm + 2), np = nx * nym * nzm;
  struct { const char *n; void *p; long c; int i; } t[] = {
      {"xc", o->xc, nx + 2, 0}, {"xm", o->xm, nx + 2, 0}, {"yc", o->yc, ny + 2, 0},
      {"ym", o->ym, ny + 2, 0}, {"zc", o->zc, nz + 2, 0}, {"zm", o->zm, nz + 2, 0},
      {"g3rc", o->g3rc, nx + 2, 0}, {"g3rm", o->g3rm, nx + 2, 0},
      {"udx3c", o->udx3c, nx + 2, 0}, {"udx3m", o->udx3m, nx + 2, 0},
      {"ap3ck", o->ap3ck, nx + 2, 0}, {"ac3ck", o->ac3ck, nx + 2, 0}, {"am3ck", o->am3ck, nx + 2, 0},
      {"ap3sk", o->ap3sk, nx + 2, 0}, {"ac3sk", o->ac3sk, nx + 2, 0}, {"am3sk", o->am3sk, nx + 2, 0},
      {"ap3ssk", o->ap3ssk, nx + 2, 0}, {"ac3ssk", o->ac3ssk, nx + 2, 0}, {"am3ssk", o->am3ssk, nx + 2, 0},
      {"ak1", o->ak1, nz + 2, 0}, {"ao", o->ao, nz + 2, 0}, {"ak2", o->ak2, ny + 2, 0},
      {"ap", o->ap, ny + 2, 0}, {"amphk", o->amphk, nx + 2, 0}, {"acphk", o->acphk, nx + 2, 0},
      {"apphk", o->apphk, nx + 2, 0}, {"kmc", o->kmc, nx + 2, 1}, {"kpc", o->kpc, nx + 2, 1},
      {"kmv", o->kmv, nx + 2, 1}, {"kpv", o->kpv, nx + 2, 1},
      {"tempbp", o->tempbp, nym * nzm, 0}, {"temptp", o->temptp, nym * nzm, 0},
      {"vx", o->vx, nh, 0}, {"vy", o->vy, nh, 0}, {"vz", o->vz, nh, 0}, {"pr", o->pr, nh, 0},
      {"temp", o->temp, nh, 0}, {"dphhalo", o->dphhalo, nxm * (nym + 2) * (nzm + 2), 0},
      {"rhs", o->rhs, np, 0}, {"rux", o->rux, np, 0}, {"ruy", o->ruy, np, 0}, {"ruz", o->ruz, np, 0},
      {"rutemp", o->rutemp, np, 0}, {"qcap", o->qcap, np, 0}, {"dq", o->dq, np, 0},
      {"hro", o->hro, np, 0}, {"dph", o->dph, nxm * nym * nzm, 0},
  };
  for (size_t n = 0; n < sizeof(t) / sizeof(t[0]); n++)
    if (!strcmp(t[n].n, name)) {
      if (count) *count = t[n].c;
      if (is_int) *is_int = t[n].i;
      return t[n].p;
    }
  return NULL;
}

static double *scalar_ptr(afo_t *o, const char *name) {
  struct { const char *n; double *p; } t[] = {
      {"dx", &o->dx}, {"dy", &o->dy}, {"dz", &o->dz}, {"dxq", &o->dxq}, {"dyq", &o->dyq},
      {"dzq", &o->dzq}, {"ren", &o->ren}, {"pec", &o->pec}, {"pi", &o->pi}, {"al", &o->al},
      {"ga", &o->ga}, {"ro", &o->ro}, {"beta", &o->beta}, {"dt", &o->dt}, {"time", &o->time},
      {"gam1", &o->gam[1]}, {"gam2", &o->gam[2]}, {"gam3", &o->gam[3]},
      {"rom1", &o->rom[1]}, {"rom2", &o->rom[2]}, {"rom3", &o->rom[3]},
      {"alm1", &o->alm[1]}, {"alm2", &o->alm[2]}, {"alm3", &o->alm[3]},
  };
  for (size_t n = 0; n < sizeof(t) / sizeof(t[0]); n++)
    if (!strcmp(t[n].n, name)) return t[n].p;
  return NULL;
}
int afo_get_scalar(afo_t *o, const char *name, double *out) {
  double *p = scalar_ptr(o, name);
  if (!p) return -1;
  *out = *p;
  return 0;
}
int afo_set_scalar(afo_t *o, const char *name, double val) {
  double *p = scalar_ptr(o, name);
  if (!p) return -1;
  *p = val;
  return 0;
}

/* ---- seeded initial condition (shared recipe with the HIP engine) ----------
 * noise(k,j,i) = (2u-1)*amp, u = (h>>11)*2^-53, h = splitmix64 finaliser of
 * seed*0x9E3779B97F4A7C15 + (lin+1)*0xD1B54A32D192ED03, lin the 0-based
 * Fortran linear index (k-1) + nx*((j-1) + nym*(i-1)) of a plain (nx,nym,nzm)
 * array.  temp(1)=tempbp, temp(nx)=temptp, temp(k)=1-xc(k)+noise (k=2..nxm). */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
void afo_init_linear_noise(afo_t *o, unsigned long long seed, double amp) {
  const int nx = o->nx, nym = o->nym, nzm = o->nzm;
  memset(o->vx, 0, sizeof(double) * (size_t)nx * (nym + 2) * (nzm + 2));
  memset(o->vy, 0, sizeof(double) * (size_t)nx * (nym + 2) * (nzm + 2));
  memset(o->vz, 0, sizeof(double) * (size_t)nx * (nym + 2) * (nzm + 2));
  memset(o->pr, 0, sizeof(double) * (size_t)nx * (nym + 2) * (nzm + 2));
  for (int i = 1; i <= nzm; i++)
    for (int j = 1; j <= nym; j++) {
      for (int k = 2; k <= o->nxm; k++) {
        uint64_t lin = (uint64_t)(k - 1) + (uint64_t)nx * ((uint64_t)(j - 1) + (uint64_t)nym * (uint64_t)(i - 1));
        uint64_t h = mix64((uint64_t)seed * 0x9E3779B97F4A7C15ULL + (lin + 1) * 0xD1B54A32D192ED03ULL);
        double u = (double)(h >> 11) * 0x1.0p-53;
        HX(o->temp, k, j, i) = 1.0 - o->xc[k] + (2.0 * u - 1.0) * amp;
      }
      HX(o->temp, 1, j, i) = P2(o->tempbp, j, i);
      HX(o->temp, nx, j, i) = P2(o->temptp, j, i);
    }
  afo_update_all_halos(o);
}

/* CreateInitialConditions.F90:20-54 */
void afo_init_reference_ic(afo_t *o) {
  const int nx = o->nx, nxm = o->nxm, nym = o->nym, nzm = o->nzm;
  double eps = 0.01;
  for (int i = 1; i <= nzm; i++)
    for (int j = 1; j <= nym; j++)
      for (int k = 1; k <= nxm; k++) {
        HX(o->vz, k, j, i) = 0.0;
        double yyy = o->xm[k], xxx = o->yc[j];
        HX(o->vy, k, j, i) = (2.0 * yyy - 6.0 * yyy * yyy + 4.0 * yyy * yyy * yyy) * sin(3 * xxx) * eps;
        yyy = o->xc[k];
        xxx = o->ym[j];
        HX(o->vx, k, j, i) = -(yyy * yyy) * ((1.0 - yyy) * (1.0 - yyy)) * cos(3.1 * xxx) * eps;
      }
  for (int i = 1; i <= nzm; i++)
    for (int j = 1; j <= nym; j++)
      for (int k = 2; k <= nxm; k++)
        HX(o->temp, k, j, i) = P2(o->tempbp, j, i) - (P2(o->tempbp, j, i) - P2(o->temptp, j, i)) * o->xc[k] / o->xc[nx];
  for (int i = 1; i <= nzm; i++)
    for (int j = 1; j <= nym; j++) {
      HX(o->temp, 1, j, i) = P2(o->tempbp, j, i);
      HX(o->temp, nx, j, i) = P2(o->temptp, j, i);
    }
  afo_update_all_halos(o);
}

/* ---- periodic halo update, one rank: halo_common.F90:67-70,105-108 ---------- */
static void halo_update(afo_t *o, double *a, int n1) {
  const int nym = o->nym, nzm = o->nzm;
  const size_t s1 = (size_t)n1, plane = s1 * (nym + 2);
  for (int i = 0; i <= nzm + 1; i++) {
    double *b = a + (size_t)i * plane;
    memcpy(b + s1 * (nym + 1), b + s1 * 1, s1 * sizeof(double));
    memcpy(b + 0, b + s1 * nym, s1 * sizeof(double));
  }
  memcpy(a + plane * (nzm + 1), a + plane * 1, plane * sizeof(double));
  memcpy(a + 0, a + plane * nzm, plane * sizeof(double));
}
void afo_update_halo(afo_t *o, const char *name) {
  if (!strcmp(name, "dphhalo")) { halo_update(o, o->dphhalo, o->nxm); return; }
  long c; int ii;
  double *p = (double *)afo_array(o, name, &c, &ii);
  if (p) halo_update(o, p, o->nx);
}
void afo_update_all_halos(afo_t *o) {
  halo_update(o, o->vx, o->nx);
  halo_update(o, o->vy, o->nx);
  halo_update(o, o->vz, o->nx);
  halo_update(o, o->temp, o->nx);
  halo_update(o, o->pr, o->nx);
}

/* ---- ExplicitTermsVX.F90:11-109 ----------------------------------------------- */
void afo_explicit_terms_vx(afo_t *o) {
  const double udy = o->dy * 0.25, udz = o->dz * 0.25;
  const double udyq = o->dyq / o->ren, udzq = o->dzq / o->ren;
  double *vx = o->vx, *vy = o->vy, *vz = o->vz, *temp = o->temp, *qcap = o->qcap;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++) {
    int imm = ic - 1, ipp = ic + 1;
    for (int jc = 1; jc <= o->nym; jc++) {
      int jmm = jc - 1, jpp = jc + 1;
      for (int kc = 2; kc <= o->nxm; kc++) {
        int km = kc - 1, kp = kc + 1;
        double hxz = (((HX(vz, kc, jc, ipp) + HX(vz, km, jc, ipp)) * (HX(vx, kc, jc, ipp) + HX(vx, kc, jc, ic))) -
                      ((HX(vz, kc, jc, ic) + HX(vz, km, jc, ic)) * (HX(vx, kc, jc, ic) + HX(vx, kc, jc, imm)))) * udz;
        double hxy = (((HX(vy, kc, jpp, ic) + HX(vy, km, jpp, ic)) * (HX(vx, kc, jpp, ic) + HX(vx, kc, jc, ic))) -
                      ((HX(vy, kc, jc, ic) + HX(vy, km, jc, ic)) * (HX(vx, kc, jc, ic) + HX(vx, kc, jmm, ic)))) * udy;
        double hxx = ((HX(vx, kp, jc, ic) + HX(vx, kc, jc, ic)) * (HX(vx, kp, jc, ic) + HX(vx, kc, jc, ic)) -
                      (HX(vx, kc, jc, ic) + HX(vx, km, jc, ic)) * (HX(vx, kc, jc, ic) + HX(vx, km, jc, ic))) *
                     o->udx3c[kc] * 0.25;
        double tempit = HX(temp, kc, jc, ic);
        double dzzvx = (HX(vx, kc, jc, imm) - 2.0 * HX(vx, kc, jc, ic) + HX(vx, kc, jc, ipp)) * udzq;
        double dyyvx = (HX(vx, kc, jmm, ic) - 2.0 * HX(vx, kc, jc, ic) + HX(vx, kc, jpp, ic)) * udyq;
        PX(qcap, kc, jc, ic) = -(hxx + hxy + hxz) + dyyvx + dzzvx + tempit;
      }
    }
  }
}

/* ---- ExplicitTermsVY.F90:11-105 ------------------------------------------------ */
void afo_explicit_terms_vy(afo_t *o) {
  const double udyq = o->dyq / o->ren, udzq = o->dzq / o->ren;
  const double udy = o->dy * 0.25, udz = o->dz * 0.25;
  double *vx = o->vx, *vy = o->vy, *vz = o->vz, *dph = o->dph;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++) {
    int imm = ic - 1, ipp = ic + 1;
    for (int jc = 1; jc <= o->nym; jc++) {
      int jmm = jc - 1, jpp = jc + 1;
      for (int kc = 1; kc <= o->nxm; kc++) {
        int kmm = o->kmv[kc], kpp = o->kpv[kc], kp = kc + 1;
        double hyx = ((HX(vx, kp, jc, ic) + HX(vx, kp, jmm, ic)) * (HX(vy, kpp, jc, ic) + HX(vy, kc, jc, ic)) -
                      (HX(vx, kc, jc, ic) + HX(vx, kc, jmm, ic)) * (HX(vy, kc, jc, ic) + HX(vy, kmm, jc, ic))) *
                     o->udx3m[kc] * 0.25;
        double hyy = ((HX(vy, kc, jpp, ic) + HX(vy, kc, jc, ic)) * (HX(vy, kc, jpp, ic) + HX(vy, kc, jc, ic)) -
                      (HX(vy, kc, jmm, ic) + HX(vy, kc, jc, ic)) * (HX(vy, kc, jmm, ic) + HX(vy, kc, jc, ic))) * udy;
        double hyz = ((HX(vy, kc, jc, ipp) + HX(vy, kc, jc, ic)) * (HX(vz, kc, jc, ipp) + HX(vz, kc, jmm, ipp)) -
                      (HX(vy, kc, jc, ic) + HX(vy, kc, jc, imm)) * (HX(vz, kc, jc, ic) + HX(vz, kc, jmm, ic))) * udz;
        double dyyvy = (HX(vy, kc, jpp, ic) - 2.0 * HX(vy, kc, jc, ic) + HX(vy, kc, jmm, ic)) * udyq;
        double dzzvy = (HX(vy, kc, jc, ipp) - 2.0 * HX(vy, kc, jc, ic) + HX(vy, kc, jc, imm)) * udzq;
        PM(dph, kc, jc, ic) = -(hyx + hyy + hyz) + dyyvy + dzzvy;
      }
    }
  }
}

/* ---- ExplicitTermsVZ.F90:11-109 ------------------------------------------------ */
void afo_explicit_terms_vz(afo_t *o) {
  const double udyq = o->dyq / o->ren, udzq = o->dzq / o->ren;
  const double udy = o->dy * 0.25, udz = o->dz * 0.25;
  double *vx = o->vx, *vy = o->vy, *vz = o->vz, *dq = o->dq;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++) {
    int imm = ic - 1, ipp = ic + 1;
    for (int jc = 1; jc <= o->nym; jc++) {
      int jmm = jc - 1, jpp = jc + 1;
      for (int kc = 1; kc <= o->nxm; kc++) {
        int kmm = o->kmv[kc], kpp = o->kpv[kc], kp = kc + 1;
        double hzz = ((HX(vz, kc, jc, ipp) + HX(vz, kc, jc, ic)) * (HX(vz, kc, jc, ipp) + HX(vz, kc, jc, ic)) -
                      (HX(vz, kc, jc, imm) + HX(vz, kc, jc, ic)) * (HX(vz, kc, jc, imm) + HX(vz, kc, jc, ic))) * udz;
        double hzy = ((HX(vy, kc, jpp, ic) + HX(vy, kc, jpp, imm)) * (HX(vz, kc, jpp, ic) + HX(vz, kc, jc, ic)) -
                      (HX(vy, kc, jc, ic) + HX(vy, kc, jc, imm)) * (HX(vz, kc, jc, ic) + HX(vz, kc, jmm, ic))) * udy;
        double hzx = ((HX(vx, kp, jc, ic) + HX(vx, kp, jc, imm)) * (HX(vz, kpp, jc, ic) + HX(vz, kc, jc, ic)) -
                      (HX(vx, kc, jc, ic) + HX(vx, kc, jc, imm)) * (HX(vz, kc, jc, ic) + HX(vz, kmm, jc, ic))) *
                     o->udx3m[kc] * 0.25;
        double dzzvz = (HX(vz, kc, jc, ipp) - 2.0 * HX(vz, kc, jc, ic) + HX(vz, kc, jc, imm)) * udzq;
        double dyyvz = (HX(vz, kc, jpp, ic) - 2.0 * HX(vz, kc, jc, ic) + HX(vz, kc, jmm, ic)) * udyq;
        PX(dq, kc, jc, ic) = -(hzx + hzy + hzz) + dyyvz + dzzvz;
      }
    }
  }
}

/* ---- ExplicitTermsTemp.F90:11-100 ---------------------------------------------- */
void afo_explicit_terms_temp(afo_t *o) {
  const double udz = o->dz * 0.25, udy = o->dy * 0.25;
  const double udzq = o->dzq / o->pec, udyq = o->dyq / o->pec;
  double *vx = o->vx, *vy = o->vy, *vz = o->vz, *temp = o->temp, *hro = o->hro;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++) {
    int im = ic - 1, ip = ic + 1;
    for (int jc = 1; jc <= o->nym; jc++) {
      int jm = jc - 1, jp = jc + 1;
      for (int kc = 2; kc <= o->nxm; kc++) {
        int km = kc - 1, kp = kc + 1;
        double htz = ((HX(vz, km, jc, ip) + HX(vz, kc, jc, ip)) * (HX(temp, kc, jc, ip) + HX(temp, kc, jc, ic)) -
                      (HX(vz, km, jc, ic) + HX(vz, kc, jc, ic)) * (HX(temp, kc, jc, ic) + HX(temp, kc, jc, im))) * udz;
        double hty = ((HX(vy, kc, jp, ic) + HX(vy, km, jp, ic)) * (HX(temp, kc, jp, ic) + HX(temp, kc, jc, ic)) -
                      (HX(vy, kc, jc, ic) + HX(vy, km, jc, ic)) * (HX(temp, kc, jc, ic) + HX(temp, kc, jm, ic))) * udy;
        double htx = ((HX(vx, kp, jc, ic) + HX(vx, kc, jc, ic)) * (HX(temp, kp, jc, ic) + HX(temp, kc, jc, ic)) -
                      (HX(vx, kc, jc, ic) + HX(vx, km, jc, ic)) * (HX(temp, kc, jc, ic) + HX(temp, km, jc, ic))) *
                     o->udx3c[kc] * 0.25;
        double dzzt = (HX(temp, kc, jc, ip) - 2.0 * HX(temp, kc, jc, ic) + HX(temp, kc, jc, im)) * udzq;
        double dyyt = (HX(temp, kc, jp, ic) - 2.0 * HX(temp, kc, jc, ic) + HX(temp, kc, jm, ic)) * udyq;
        PX(hro, kc, jc, ic) = -(htx + hty + htz) + dyyt + dzzt;
      }
    }
  }
}

/* ---- LAPACK restatements (netlib dgttrf / dgtts2 'N') --------------------------
 * Arrays are 0-based here: dl[0..n-2], d[0..n-1], du[0..n-2], du2[0..n-3]. */
int afo_dgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv) {
  for (int i = 0; i < n; i++) ipiv[i] = i;
  for (int i = 0; i < n - 2; i++) du2[i] = 0.0;
  for (int i = 0; i < n - 2; i++) {
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (d[i] != 0.0) {
        double fact = dl[i] / d[i];
        dl[i] = fact;
        d[i + 1] = d[i + 1] - fact * du[i];
      }
    } else {
      double fact = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = fact;
      double temp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = temp - fact * d[i + 1];
      du2[i] = du[i + 1];
      du[i + 1] = -fact * du[i + 1];
      ipiv[i] = i + 1;
    }
  }
  if (n > 1) {
    int i = n - 2;
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (d[i] != 0.0) {
        double fact = dl[i] / d[i];
        dl[i] = fact;
        d[i + 1] = d[i + 1] - fact * du[i];
      }
    } else {
      double fact = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = fact;
      double temp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = temp - fact * d[i + 1];
      ipiv[i] = i + 1;
    }
  }
  for (int i = 0; i < n; i++)
    if (d[i] == 0.0) return i + 1;
  return 0;
}

static void gtts2_one(int n, const double *dl, const double *d, const double *du,
                      const double *du2, const int *ipiv, double *b, ptrdiff_t inc) {
  for (int i = 0; i < n - 1; i++) {
    if (ipiv[i] == i) {
      b[(i + 1) * inc] = b[(i + 1) * inc] - dl[i] * b[i * inc];
    } else {
      double temp = b[i * inc];
      b[i * inc] = b[(i + 1) * inc];
      b[(i + 1) * inc] = temp - dl[i] * b[i * inc];
    }
  }
  b[(n - 1) * inc] = b[(n - 1) * inc] / d[n - 1];
  if (n > 1) b[(n - 2) * inc] = (b[(n - 2) * inc] - du[n - 2] * b[(n - 1) * inc]) / d[n - 2];
  for (int i = n - 3; i >= 0; i--)
    b[i * inc] = (b[i * inc] - du[i] * b[(i + 1) * inc] - du2[i] * b[(i + 2) * inc]) / d[i];
}

void afo_dgttrs(int n, int nrhs, const double *dl, const double *d, const double *du,
                const double *du2, const int *ipiv, double *b, int ldb) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < nrhs; j++) gtts2_one(n, dl, d, du, du2, ipiv, b + (size_t)j * ldb, 1);
}

/* ---- SolveImpEqnUpdate_X.F90:12-67 --------------------------------------------- */
static void solve_imp_eqn_update_x(afo_t *o) {
  const int nx = o->nx, nxm = o->nxm;
  double *amkl = dvec(nx), *apkl = dvec(nx), *ackl = dvec(nx);
  double *amkT = dvec(nx), *apkT = dvec(nx), *ackT = dvec(nx), *appk = dvec(nx);
  int *ipkv = ivec(nx);
  double betadx = o->beta * o->al;
  amkl[1] = 0.0; apkl[1] = 0.0; ackl[1] = 1.0;
  for (int kc = 2; kc <= nxm; kc++) {
    double ackl_b = 1.0 / (1.0 - o->ac3ck[kc] * betadx);
    amkl[kc] = -o->am3ssk[kc] * betadx * ackl_b;
    ackl[kc] = 1.0;
    apkl[kc] = -o->ap3ssk[kc] * betadx * ackl_b;
  }
  amkl[nx] = 0.0; apkl[nx] = 0.0; ackl[nx] = 1.0;
  for (int k = 0; k < nx - 1; k++) { amkT[k] = amkl[k + 2]; apkT[k] = apkl[k + 1]; }
  for (int k = 0; k < nx; k++) ackT[k] = ackl[k + 1];
  afo_dgttrf(nx, amkT, ackT, apkT, appk, ipkv);
  double *rhs = o->rhs, *vx = o->vx;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++)
    for (int jc = 1; jc <= o->nym; jc++) {
      for (int kc = 2; kc <= nxm; kc++) {
        double ackl_b = 1.0 / (1.0 - o->ac3ck[kc] * betadx);
        PX(rhs, kc, jc, ic) = PX(rhs, kc, jc, ic) * ackl_b;
      }
      gtts2_one(nx, amkT, ackT, apkT, appk, ipkv, &PX(rhs, 1, jc, ic), 1);
      for (int kc = 2; kc <= nxm; kc++) HX(vx, kc, jc, ic) = HX(vx, kc, jc, ic) + PX(rhs, kc, jc, ic);
    }
  free(amkl); free(apkl); free(ackl); free(amkT); free(apkT); free(ackT); free(appk); free(ipkv);
}

/* ---- SolveImpEqnUpdate_YZ.F90:12-61 (q = vy or vz, rhs shared scratch) --------- */
static void solve_imp_eqn_update_yz(afo_t *o, double *q) {
  const int nx = o->nx, nxm = o->nxm;
  double *amkl = dvec(nx), *apkl = dvec(nx), *ackl = dvec(nx);
  double *amkT = dvec(nx), *apkT = dvec(nx), *ackT = dvec(nx), *appk = dvec(nx);
  int *ipkv = ivec(nx);
  double betadx = o->beta * o->al;
  for (int kc = 1; kc <= nxm; kc++) {
    double ackl_b = 1.0 / (1.0 - o->ac3sk[kc] * betadx);
    amkl[kc] = -o->am3sk[kc] * betadx * ackl_b;
    ackl[kc] = 1.0;
    apkl[kc] = -o->ap3sk[kc] * betadx * ackl_b;
  }
  for (int k = 0; k < nxm - 1; k++) { amkT[k] = amkl[k + 2]; apkT[k] = apkl[k + 1]; }
  for (int k = 0; k < nxm; k++) ackT[k] = ackl[k + 1];
  afo_dgttrf(nxm, amkT, ackT, apkT, appk, ipkv);
  double *rhs = o->rhs;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++)
    for (int jc = 1; jc <= o->nym; jc++) {
      for (int kc = 1; kc <= nxm; kc++) {
        double ackl_b = 1.0 / (1.0 - o->ac3sk[kc] * betadx);
        PX(rhs, kc, jc, ic) = PX(rhs, kc, jc, ic) * ackl_b;
      }
      gtts2_one(nxm, amkT, ackT, apkT, appk, ipkv, &PX(rhs, 1, jc, ic), 1);
      for (int kc = 1; kc <= nxm; kc++) HX(q, kc, jc, ic) = HX(q, kc, jc, ic) + PX(rhs, kc, jc, ic);
    }
  free(amkl); free(apkl); free(ackl); free(amkT); free(apkT); free(ackT); free(appk); free(ipkv);
}

/* ---- SolveImpEqnUpdate_Temp.F90:11-70 ------------------------------------------ */
static void solve_imp_eqn_update_temp(afo_t *o) {
  const int nx = o->nx, nxm = o->nxm;
  double *amkl = dvec(nx), *apkl = dvec(nx), *ackl = dvec(nx);
  double *amkT = dvec(nx), *apkT = dvec(nx), *ackT = dvec(nx), *appk = dvec(nx);
  int *ipkv = ivec(nx);
  double betadx = 0.5 * o->al * o->dt / o->pec;
  amkl[1] = 0.0; apkl[1] = 0.0; ackl[1] = 1.0;
  for (int kc = 2; kc <= nxm; kc++) {
    double ackl_b = 1.0 / (1. - o->ac3ssk[kc] * betadx);
    amkl[kc] = -o->am3ssk[kc] * betadx * ackl_b;
    ackl[kc] = 1.0;
    apkl[kc] = -o->ap3ssk[kc] * betadx * ackl_b;
  }
  amkl[nx] = 0.0; apkl[nx] = 0.0; ackl[nx] = 1.0;
  for (int k = 0; k < nx - 1; k++) { amkT[k] = amkl[k + 2]; apkT[k] = apkl[k + 1]; }
  for (int k = 0; k < nx; k++) ackT[k] = ackl[k + 1];
  afo_dgttrf(nx, amkT, ackT, apkT, appk, ipkv);
  double *rhs = o->rhs, *temp = o->temp;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++)
    for (int jc = 1; jc <= o->nym; jc++) {
      for (int kc = 2; kc <= nxm; kc++) {
        double ackl_b = 1.0 / (1.0 - o->ac3ssk[kc] * betadx);
        PX(rhs, kc, jc, ic) = PX(rhs, kc, jc, ic) * ackl_b;
      }
      gtts2_one(nx, amkT, ackT, apkT, appk, ipkv, &PX(rhs, 1, jc, ic), 1);
      for (int kc = 2; kc <= nxm; kc++) HX(temp, kc, jc, ic) = HX(temp, kc, jc, ic) + PX(rhs, kc, jc, ic);
    }
  free(amkl); free(apkl); free(ackl); free(amkT); free(apkT); free(ackT); free(appk); free(ipkv);
}

/* ---- ImplicitAndUpdateVX.F90:14-80 ---------------------------------------------- */
void afo_implicit_and_update_vx(afo_t *o) {
  const double alre = o->al / o->ren;
  double *vx = o->vx, *pr = o->pr, *rhs = o->rhs, *rux = o->rux, *qcap = o->qcap;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++)
    for (int jc = 1; jc <= o->nym; jc++)
      for (int kc = 2; kc <= o->nxm; kc++) {
        int km = kc - 1, kp = kc + 1;
        double udx3 = o->al * o->udx3c[kc];
        double amm = o->am3ck[kc], acc = o->ac3ck[kc], app = o->ap3ck[kc];
        double dxxvx = HX(vx, kp, jc, ic) * app + HX(vx, kc, jc, ic) * acc + HX(vx, km, jc, ic) * amm;
        double dxp = (HX(pr, kc, jc, ic) - HX(pr, km, jc, ic)) * udx3;
        PX(rhs, kc, jc, ic) = (o->ga * PX(qcap, kc, jc, ic) + o->ro * PX(rux, kc, jc, ic) + alre * dxxvx - dxp) * o->dt;
        PX(rux, kc, jc, ic) = PX(qcap, kc, jc, ic);
      }
  solve_imp_eqn_update_x(o);
  for (int i = 0; i <= o->nzm + 1; i++)
    for (int j = 0; j <= o->nym + 1; j++) {
      HX(vx, 1, j, i) = 0.0;
      HX(vx, o->nx, j, i) = 0.0;
    }
}

/* ---- ImplicitAndUpdateVY.F90:14-78 ---------------------------------------------- */
void afo_implicit_and_update_vy(afo_t *o) {
  const double alre = o->al / o->ren, udy = o->dy * o->al;
  double *vy = o->vy, *pr = o->pr, *rhs = o->rhs, *ruy = o->ruy, *dph = o->dph;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++)
    for (int jc = 1; jc <= o->nym; jc++) {
      int jmm = jc - 1;
      for (int kc = 1; kc <= o->nxm; kc++) {
        int kmm = o->kmv[kc], kpp = o->kpv[kc];
        double amm = o->am3sk[kc], acc = o->ac3sk[kc], app = o->ap3sk[kc];
        double dxxvy = HX(vy, kpp, jc, ic) * app + HX(vy, kc, jc, ic) * acc + HX(vy, kmm, jc, ic) * amm;
        double dyp = (HX(pr, kc, jc, ic) - HX(pr, kc, jmm, ic)) * udy;
        PX(rhs, kc, jc, ic) = (o->ga * PM(dph, kc, jc, ic) + o->ro * PX(ruy, kc, jc, ic) + alre * dxxvy - dyp) * o->dt;
        PX(ruy, kc, jc, ic) = PM(dph, kc, jc, ic);
      }
    }
  solve_imp_eqn_update_yz(o, o->vy);
}

/* ---- ImplicitAndUpdateVZ.F90:14-75 ---------------------------------------------- */
void afo_implicit_and_update_vz(afo_t *o) {
  const double alre = o->al / o->ren;
  double *vz = o->vz, *pr = o->pr, *rhs = o->rhs, *ruz = o->ruz, *dq = o->dq;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++) {
    int imm = ic - 1;
    for (int jc = 1; jc <= o->nym; jc++)
      for (int kc = 1; kc <= o->nxm; kc++) {
        int kmm = o->kmv[kc], kpp = o->kpv[kc];
        double amm = o->am3sk[kc], acc = o->ac3sk[kc], app = o->ap3sk[kc];
        double dxxvz = HX(vz, kpp, jc, ic) * app + HX(vz, kc, jc, ic) * acc + HX(vz, kmm, jc, ic) * amm;
        double dzp = (HX(pr, kc, jc, ic) - HX(pr, kc, jc, imm)) * o->dz * o->al;
        PX(rhs, kc, jc, ic) = (o->ga * PX(dq, kc, jc, ic) + o->ro * PX(ruz, kc, jc, ic) + alre * dxxvz - dzp) * o->dt;
        PX(ruz, kc, jc, ic) = PX(dq, kc, jc, ic);
      }
  }
  solve_imp_eqn_update_yz(o, o->vz);
}

/* ---- ImplicitAndUpdateTemp.F90:13-77 -------------------------------------------- */
void afo_implicit_and_update_temp(afo_t *o) {
  const double alpec = o->al / o->pec;
  double *temp = o->temp, *rhs = o->rhs, *rutemp = o->rutemp, *hro = o->hro;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++)
    for (int jc = 1; jc <= o->nym; jc++)
      for (int kc = 2; kc <= o->nxm; kc++) {
        double dxxt = HX(temp, kc + 1, jc, ic) * o->ap3ck[kc] + HX(temp, kc, jc, ic) * o->ac3ck[kc] +
                      HX(temp, kc - 1, jc, ic) * o->am3ck[kc];
        PX(rhs, kc, jc, ic) = (o->ga * PX(hro, kc, jc, ic) + o->ro * PX(rutemp, kc, jc, ic) + alpec * dxxt) * o->dt;
        PX(rutemp, kc, jc, ic) = PX(hro, kc, jc, ic);
      }
  solve_imp_eqn_update_temp(o);
  for (int ic = 1; ic <= o->nzm; ic++)
    for (int jc = 1; jc <= o->nym; jc++) {
      HX(temp, 1, jc, ic) = P2(o->tempbp, jc, ic);
      HX(temp, o->nx, jc, ic) = P2(o->temptp, jc, ic);
    }
}

/* ---- CalcLocalDivergence.F90:12-44 ---------------------------------------------- */
void afo_calc_local_divergence(afo_t *o) {
  const double usdtal = 1.0 / (o->dt * o->al);
  double *vx = o->vx, *vy = o->vy, *vz = o->vz, *dph = o->dph;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++) {
    int ip = ic + 1;
    for (int jc = 1; jc <= o->nym; jc++) {
      int jp = jc + 1;
      for (int kc = 1; kc <= o->nxm; kc++) {
        int kp = kc + 1;
        double dqcap = (HX(vz, kc, jc, ip) - HX(vz, kc, jc, ic)) * o->dz +
                       (HX(vy, kc, jp, ic) - HX(vy, kc, jc, ic)) * o->dy +
                       (HX(vx, kp, jc, ic) - HX(vx, kc, jc, ic)) * o->udx3m[kc];
        PM(dph, kc, jc, ic) = dqcap * usdtal;
      }
    }
  }
}

/* ---- FFT with FFTW conventions (mixed radix, recursive decimation in time) ---- */
static void twiddle(int n, long m, int sign, double *c, double *s) {
  /* exp(sign*2*pi*i*m/n) evaluated in long double (x87 80-bit) and rounded once */
  m %= n;
  if (m < 0) m += n;
  long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
  *c = (double)cosl(a);
  *s = (double)(sign * sinl(a));
}

typedef struct { double re, im; } cplx;

static void fft_rec(int n, const cplx *in, ptrdiff_t istride, cplx *out, int sign, const cplx *w, int wn) {
  /* w: table of exp(sign*2*pi*i*t/wn), n divides wn */
  if (n == 1) { out[0] = in[0]; return; }
  int p;
  if (n % 4 == 0) p = 4;
  else if (n % 2 == 0) p = 2;
  else if (n % 3 == 0) p = 3;
  else if (n % 5 == 0) p = 5;
  else p = n; /* naive DFT for other prime factors */
  int m = n / p;
  /* sub-transforms of the decimated sequences x[r + p*t] */
  for (int r = 0; r < p; r++) fft_rec(m, in + r * istride, istride * p, out + r * m, sign, w, wn);
  cplx tmp_s[5], res_s[5];
  cplx *tmp = p <= 5 ? tmp_s : (cplx *)malloc(sizeof(cplx) * (size_t)p);
  cplx *res = p <= 5 ? res_s : (cplx *)malloc(sizeof(cplx) * (size_t)p);
  int stride_w = wn / n;
  for (int k = 0; k < m; k++) {
    for (int r = 0; r < p; r++) {
      cplx y = out[r * m + k];
      if (r == 0 || k == 0) { tmp[r] = y; continue; }
      cplx t = w[(size_t)((long)r * k * stride_w) % (size_t)wn];
      tmp[r].re = y.re * t.re - y.im * t.im;
      tmp[r].im = y.re * t.im + y.im * t.re;
    }
    for (int q = 0; q < p; q++) {
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < p; r++) {
        cplx t = w[(size_t)(((long)r * q) % p) * (size_t)(wn / p)];
        if (r == 0 || q == 0) { sr += tmp[r].re; si += tmp[r].im; continue; }
        sr += tmp[r].re * t.re - tmp[r].im * t.im;
        si += tmp[r].re * t.im + tmp[r].im * t.re;
      }
      res[q].re = sr;
      res[q].im = si;
    }
    for (int q = 0; q < p; q++) out[q * m + k] = res[q];
  }
  if (p > 5) { free(tmp); free(res); }
}

static cplx *twiddle_table(int n, int sign) {
  cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int t = 0; t < n; t++) twiddle(n, t, sign, &w[t].re, &w[t].im);
  return w;
}

void afo_fft_c2c(int n, const double *in, double *out, int sign) {
  cplx *w = twiddle_table(n, sign < 0 ? -1 : 1);
  cplx *o = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  fft_rec(n, (const cplx *)in, 1, o, sign, w, n);
  memcpy(out, o, sizeof(cplx) * (size_t)n);
  free(o);
  free(w);
}

void afo_fft_r2c(int n, const double *in, double *out) {
  cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n), *y = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int t = 0; t < n; t++) { x[t].re = in[t]; x[t].im = 0.0; }
  cplx *w = twiddle_table(n, -1);
  fft_rec(n, x, 1, y, -1, w, n);
  memcpy(out, y, sizeof(cplx) * (size_t)(n / 2 + 1));
  free(x); free(y); free(w);
}

void afo_fft_c2r(int n, const double *in, double *out) {
  const cplx *c = (const cplx *)in;
  cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n), *y = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  int h = n / 2;
  for (int m = 0; m <= h; m++) x[m] = c[m];
  x[0].im = 0.0;
  if (n % 2 == 0) x[h].im = 0.0;
  for (int m = h + 1; m < n; m++) { x[m].re = c[n - m].re; x[m].im = -c[n - m].im; }
  cplx *w = twiddle_table(n, +1);
  fft_rec(n, x, 1, y, +1, w, n);
  for (int t = 0; t < n; t++) out[t] = y[t].re;
  free(x); free(y); free(w);
}

/* Batched line transforms over strided data (used by the Poisson solve). */
static void fft_lines_c2c(int n, cplx *base, ptrdiff_t stride, long nlines,
                          ptrdiff_t (*line_off)(long, const void *), const void *ctx, int sign) {
  cplx *w = twiddle_table(n, sign);
#pragma omp parallel
  {
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n), *y = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
    for (long l = 0; l < nlines; l++) {
      cplx *p = base + line_off(l, ctx);
      for (int t = 0; t < n; t++) x[t] = p[t * stride];
      fft_rec(n, x, 1, y, sign, w, n);
      for (int t = 0; t < n; t++) p[t * stride] = y[t];
    }
    free(x); free(y);
  }
  free(w);
}

struct zline_ctx { long nxm, nymh; };
static ptrdiff_t zline_off(long l, const void *v) {
  const struct zline_ctx *c = (const struct zline_ctx *)v;
  long k = l % c->nxm, j = l / c->nxm;
  return (ptrdiff_t)(k + c->nxm * j);
}

/* ---- SolvePressureCorrection.F90:12-154 (one rank: transposes are identities) -- */
int afo_solve_pressure_correction(afo_t *o) {
  const int nxm = o->nxm, nym = o->nym, nzm = o->nzm, nymh = nym / 2 + 1;
  const size_t nsp = (size_t)nxm * nymh * nzm;
  /* ry1 = dph; cy1(nxm, nymh, nzm) = r2c along y (stride nxm) */
  cplx *cz1 = (cplx *)calloc(nsp, sizeof(cplx));
  double *dph = o->dph;
  {
    cplx *w = twiddle_table(nym, -1);
#pragma omp parallel
    {
      cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)nym), *y = (cplx *)malloc(sizeof(cplx) * (size_t)nym);
#pragma omp for schedule(static)
      for (long l = 0; l < (long)nxm * nzm; l++) {
        int k = (int)(l % nxm), i = (int)(l / nxm);
        for (int j = 0; j < nym; j++) { x[j].re = dph[(size_t)k + (size_t)nxm * ((size_t)j + (size_t)nym * i)]; x[j].im = 0.0; }
        fft_rec(nym, x, 1, y, -1, w, nym);
        for (int m = 0; m < nymh; m++) cz1[(size_t)k + (size_t)nxm * ((size_t)m + (size_t)nymh * i)] = y[m];
      }
      free(x); free(y);
    }
    free(w);
  }
  /* forward c2c along z, stride nxm*nymh */
  struct zline_ctx zc = {nxm, nymh};
  fft_lines_c2c(nzm, cz1, (ptrdiff_t)nxm * nymh, (long)nxm * nymh, zline_off, &zc, -1);
  /* cz1 = cz1 / (nzm*nym) */
  {
    double nrm = (double)(nzm * nym);
#pragma omp parallel for schedule(static)
    for (size_t t = 0; t < nsp; t++) { cz1[t].re = cz1[t].re / nrm; cz1[t].im = cz1[t].im / nrm; }
  }
  /* per-mode tridiagonal (:117-141), zgttrf/zgttrs restated in real arithmetic:
     every coefficient is real, so the complex factorisation equals the real one
     and the complex solve is the real solve applied to re and im parts. */
  int info_all = 0;
#pragma omp parallel
  {
    double *amph = dvec(nxm), *apph = dvec(nxm), *acphT = dvec(nxm), *appph = dvec(nxm);
    double *dl = dvec(nxm), *du = dvec(nxm);
    int *piv = ivec(nxm);
    double *br = dvec(nxm), *bi = dvec(nxm);
#pragma omp for collapse(2) schedule(static)
    for (int i = 1; i <= nzm; i++)
      for (int j = 1; j <= nymh; j++) {
        cplx *col = cz1 + (size_t)nxm * ((size_t)(j - 1) + (size_t)nymh * (i - 1));
        for (int k = 1; k <= nxm; k++) {
          double acphT_b = 1.0 / (o->acphk[k] - o->ak2[j] - o->ak1[i]);
          col[k - 1].re = col[k - 1].re * acphT_b;
          col[k - 1].im = col[k - 1].im * acphT_b;
          apph[k] = o->apphk[k] * acphT_b;
          amph[k] = o->amphk[k] * acphT_b;
          acphT[k] = 1.0;
        }
        for (int k = 0; k < nxm - 1; k++) { dl[k] = amph[k + 2]; du[k] = apph[k + 1]; }
        int info = afo_dgttrf(nxm, dl, acphT + 1, du, appph, piv);
        if (info > 0) {
#pragma omp critical
          info_all = info;
        }
        for (int k = 0; k < nxm; k++) { br[k] = col[k].re; bi[k] = col[k].im; }
        gtts2_one(nxm, dl, acphT + 1, du, appph, piv, br, 1);
        gtts2_one(nxm, dl, acphT + 1, du, appph, piv, bi, 1);
        for (int k = 0; k < nxm; k++) { col[k].re = br[k]; col[k].im = bi[k]; }
      }
    free(amph); free(apph); free(acphT); free(appph); free(dl); free(du); free(piv); free(br); free(bi);
  }
  /* backward c2c along z, then c2r along y back into dph */
  fft_lines_c2c(nzm, cz1, (ptrdiff_t)nxm * nymh, (long)nxm * nymh, zline_off, &zc, +1);
  {
    cplx *w = twiddle_table(nym, +1);
    int h = nym / 2;
#pragma omp parallel
    {
      cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)nym), *y = (cplx *)malloc(sizeof(cplx) * (size_t)nym);
#pragma omp for schedule(static)
      for (long l = 0; l < (long)nxm * nzm; l++) {
        int k = (int)(l % nxm), i = (int)(l / nxm);
        for (int m = 0; m <= h; m++) x[m] = cz1[(size_t)k + (size_t)nxm * ((size_t)m + (size_t)nymh * i)];
        x[0].im = 0.0;
        if (nym % 2 == 0) x[h].im = 0.0;
        for (int m = h + 1; m < nym; m++) { x[m].re = x[nym - m].re; x[m].im = -x[nym - m].im; }
        fft_rec(nym, x, 1, y, +1, w, nym);
        for (int j = 0; j < nym; j++) dph[(size_t)k + (size_t)nxm * ((size_t)j + (size_t)nym * i)] = y[j].re;
      }
      free(x); free(y);
    }
    free(w);
  }
  free(cz1);
  return info_all;
}

/* ---- TimeMarcher.F90:55-63 ---------------------------------------------------- */
void afo_copy_dphhalo(afo_t *o) {
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= o->nzm; i++)
    for (int j = 1; j <= o->nym; j++)
      for (int k = 1; k <= o->nxm; k++) HM(o->dphhalo, k, j, i) = PM(o->dph, k, j, i);
  halo_update(o, o->dphhalo, o->nxm);
}

/* ---- CorrectVelocity.F90:11-48 ------------------------------------------------- */
void afo_correct_velocity(afo_t *o) {
  const double udy = o->al * o->dt * o->dy, udz = o->al * o->dt * o->dz;
  double *vx = o->vx, *vy = o->vy, *vz = o->vz, *dh = o->dphhalo;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++) {
    int im = ic - 1;
    for (int jc = 1; jc <= o->nym; jc++) {
      int jm = jc - 1;
      for (int kc = 1; kc <= o->nxm; kc++) {
        int km = o->kmv[kc];
        double usukm = o->al * o->dt * o->udx3c[kc];
        double locdph = HM(dh, kc, jc, ic);
        HX(vx, kc, jc, ic) = HX(vx, kc, jc, ic) - (locdph - HM(dh, km, jc, ic)) * usukm;
        HX(vy, kc, jc, ic) = HX(vy, kc, jc, ic) - (locdph - HM(dh, kc, jm, ic)) * udy;
        HX(vz, kc, jc, ic) = HX(vz, kc, jc, ic) - (locdph - HM(dh, kc, jc, im)) * udz;
      }
    }
  }
}

/* ---- CorrectPressure.F90:11-55 ------------------------------------------------- */
void afo_correct_pressure(afo_t *o) {
  const double be = o->al * o->beta;
  double *pr = o->pr, *dh = o->dphhalo;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++) {
    int im = ic - 1, ip = ic + 1;
    for (int jc = 1; jc <= o->nym; jc++) {
      int jm = jc - 1, jp = jc + 1;
      for (int kc = 1; kc <= o->nxm; kc++) {
        int kp = o->kpv[kc], km = o->kmv[kc];
        double amm = o->amphk[kc], acc = o->acphk[kc], app = o->apphk[kc];
        HX(pr, kc, jc, ic) = HX(pr, kc, jc, ic) + HM(dh, kc, jc, ic) -
                             be * ((HM(dh, kc, jc, ip) - 2.0 * HM(dh, kc, jc, ic) + HM(dh, kc, jc, im)) * o->dzq +
                                   (HM(dh, kc, jp, ic) - 2.0 * HM(dh, kc, jc, ic) + HM(dh, kc, jm, ic)) * o->dyq +
                                   (HM(dh, kp, jc, ic) * app + HM(dh, kc, jc, ic) * acc + HM(dh, km, jc, ic) * amm));
      }
    }
  }
}

/* ---- TimeMarcher.F90:24-74 ----------------------------------------------------- */
int afo_substep(afo_t *o, int ns, double dt) {
  o->dt = dt;
  o->beta = dt / o->ren * 0.5;
  o->al = o->alm[ns];
  o->ga = o->gam[ns];
  o->ro = o->rom[ns];
  afo_explicit_terms_vx(o);
  afo_explicit_terms_vy(o);
  afo_explicit_terms_vz(o);
  afo_explicit_terms_temp(o);
  afo_implicit_and_update_vx(o);
  afo_implicit_and_update_vy(o);
  afo_implicit_and_update_vz(o);
  afo_implicit_and_update_temp(o);
  halo_update(o, o->vy, o->nx);
  halo_update(o, o->vz, o->nx);
  afo_calc_local_divergence(o);
  int info = afo_solve_pressure_correction(o);
  afo_copy_dphhalo(o);
  afo_correct_velocity(o);
  afo_correct_pressure(o);
  afo_update_all_halos(o);
  return info;
}

int afo_time_marcher(afo_t *o, double dt) {
  int info = 0;
  for (int ns = 1; ns <= o->nsst; ns++) {
    int r = afo_substep(o, ns, dt);
    if (r) info = r;
  }
  return info;
}

/* ---- main.F90:169-200 ---------------------------------------------------------- */
int afo_step(afo_t *o, int ntime, int idtv, double dtfixed, double dtmin, double dtmax,
             double limitcfl, double *dt_used) {
  int errorcode = 0;
  double instcfl = afo_calc_max_cfl(o);
  double dt = dtfixed;
  if (idtv) {
    if (ntime > 1) {
      if (instcfl < 1.0e-8) dt = dtmax;
      else dt = limitcfl / instcfl;
      if (dt > dtmax) dt = dtmax;
    } else {
      dt = dtmin;
    }
    if (dt < dtmin) errorcode = 166;
  } else {
    instcfl = instcfl * dt;
    if (instcfl > limitcfl) errorcode = 165;
  }
  afo_time_marcher(o, dt);
  o->time = o->time + dt;
  if (dt_used) *dt_used = dt;
  return errorcode;
}

/* ---- CalcMaxCFL.F90:11-48 -------------------------------------------------------- */
double afo_calc_max_cfl(afo_t *o) {
  double cflm = 0.00000001;
  double *vx = o->vx, *vy = o->vy, *vz = o->vz;
#pragma omp parallel for reduction(max : cflm) schedule(static)
  for (int i = 1; i <= o->nzm; i++) {
    int ip = i + 1;
    for (int j = 1; j <= o->nym; j++) {
      int jp = j + 1;
      for (int k = 1; k <= o->nxm; k++) {
        int kp = k + 1;
        double qcf = (fabs((HX(vz, k, j, i) + HX(vz, k, j, ip)) * 0.5 * o->dz) +
                      fabs((HX(vy, k, j, i) + HX(vy, k, jp, i)) * 0.5 * o->dy) +
                      fabs((HX(vx, k, j, i) + HX(vx, kp, j, i)) * 0.5 * o->udx3m[k]));
        cflm = fmax(cflm, qcf);
      }
    }
  }
  return cflm;
}

/* ---- CheckDivergence.F90:11-48 ---------------------------------------------------- */
double afo_check_divergence(afo_t *o) {
  double qmax = -HUGE_VAL;
  double *vx = o->vx, *vy = o->vy, *vz = o->vz;
#pragma omp parallel for reduction(max : qmax) schedule(static)
  for (int ic = 1; ic <= o->nzm; ic++) {
    int ip = ic + 1;
    for (int jc = 1; jc <= o->nym; jc++) {
      int jp = jc + 1;
      for (int kc = 1; kc <= o->nxm; kc++) {
        int kp = kc + 1;
        double dqcap = (HX(vz, kc, jc, ip) - HX(vz, kc, jc, ic)) * o->dz +
                       (HX(vy, kc, jp, ic) - HX(vy, kc, jc, ic)) * o->dy +
                       (HX(vx, kp, jc, ic) - HX(vx, kc, jc, ic)) * o->udx3m[kc];
        qmax = fmax(fabs(dqcap), qmax);
      }
    }
  }
  return qmax;
}

/* ---- CalcPlateNu.F90:11-57 -------------------------------------------------------- */
void afo_calc_plate_nu(afo_t *o, double *nu_bot, double *nu_top) {
  double nuslow = 0.0, nusupp = 0.0;
  double del = 1.0 / (o->xc[2] - o->xc[1]);
  double deln = 1.0 / (o->xc[o->nx] - o->xc[o->nxm]);
  double *temp = o->temp;
  for (int i = 1; i <= o->nzm; i++)
    for (int j = 1; j <= o->nym; j++) {
      nuslow = nuslow + (HX(temp, 1, j, i) - HX(temp, 2, j, i)) * del;
      nusupp = nusupp + (HX(temp, o->nxm, j, i) - HX(temp, o->nx, j, i)) * deln;
    }
  nuslow = nuslow / (o->nzm * o->nym);
  nusupp = nusupp / (o->nzm * o->nym);
  *nu_bot = nuslow;
  *nu_top = nusupp;
}

/* ---- GlobalQuantities.F90:11-102 --------------------------------------------------- */
void afo_global_quantities(afo_t *o, double *out) {
  double vmax1 = 0.0, vmax2 = 0.0, vmax3 = 0.0; /* vmax = 0 after the -huge init (:34) */
  double tempmax = -HUGE_VAL, tempmin = HUGE_VAL, tempm = 0.0, anusin = 0.0;
  double vx_rms_vol = 0.0, vy_rms_vol = 0.0, vz_rms_vol = 0.0, vzvyvx_rms_vol = 0.0;
  double vol = 1.0 / (o->alx3 * o->dx * (double)o->nzm * (double)o->nym);
  double *vx = o->vx, *vy = o->vy, *vz = o->vz, *temp = o->temp;
  for (int ic = 1; ic <= o->nzm; ic++)
    for (int jc = 1; jc <= o->nym; jc++)
      for (int kc = 1; kc <= o->nxm; kc++) {
        int kp = kc + 1;
        double fac2 = o->g3rc[kc];
        vmax1 = fmax(vmax1, fabs(HX(vz, kc, jc, ic)));
        vmax2 = fmax(vmax2, fabs(HX(vy, kc, jc, ic)));
        vmax3 = fmax(vmax3, fabs(HX(vx, kc, jc, ic)));
        tempmax = fmax(tempmax, HX(temp, kc, jc, ic));
        tempmin = fmin(tempmin, HX(temp, kc, jc, ic));
        double vxcen = (HX(vx, kc, jc, ic) + HX(vx, kp, jc, ic)) * 0.5;
        double tempcen = (HX(temp, kc, jc, ic) + HX(temp, kp, jc, ic)) * 0.5;
        anusin = anusin + tempcen * vxcen * fac2;
        tempm = tempm + tempcen * fac2;
        vx_rms_vol = vx_rms_vol + fac2 * (HX(vx, kc, jc, ic) * HX(vx, kc, jc, ic));
        vy_rms_vol = vy_rms_vol + fac2 * (HX(vy, kc, jc, ic) * HX(vy, kc, jc, ic));
        vz_rms_vol = vz_rms_vol + fac2 * (HX(vz, kc, jc, ic) * HX(vz, kc, jc, ic));
        vzvyvx_rms_vol = vzvyvx_rms_vol + fac2 * (HX(vz, kc, jc, ic) * HX(vz, kc, jc, ic) +
                                                  HX(vy, kc, jc, ic) * HX(vy, kc, jc, ic) +
                                                  HX(vx, kc, jc, ic) * HX(vx, kc, jc, ic));
      }
  anusin = 1.0 + sqrt(o->pra * o->ray) * anusin * vol;
  double rradpr = sqrt(o->ray / o->pra);
  out[0] = anusin;
  out[1] = tempm * vol;
  out[2] = tempmax;
  out[3] = tempmin;
  out[4] = vmax1;
  out[5] = vmax2;
  out[6] = vmax3;
  out[7] = sqrt(vz_rms_vol * vol) * rradpr;
  out[8] = sqrt(vy_rms_vol * vol) * rradpr;
  out[9] = sqrt(vx_rms_vol * vol) * rradpr;
  out[10] = sqrt(vzvyvx_rms_vol * vol) * rradpr;
}

/* ---- CalcDissipationNu.F90:12-120 -------------------------------------------------- */
void afo_calc_dissipation_nu(afo_t *o, double *nu_mu_out, double *nu_th_out, double *disste,
                             double *dissth) {
  const int nxm = o->nxm;
  double *dxm = dvec(nxm), *dxc = dvec(nxm), *lem = dvec(nxm), *lec = dvec(nxm);
  for (int k = 1; k <= nxm; k++) {
    int kp = k + 1;
    dxm[k] = 1.0 / (o->xm[kp] - o->xm[k]);
    dxc[k] = 1.0 / (o->xc[kp] - o->xc[k]);
    lem[k] = (o->xc[kp] - o->xc[k]) * o->dx;
    lec[k] = (o->xm[kp] - o->xm[k]) * o->dx;
  }
  lec[nxm] = (o->xm[nxm] - o->xm[nxm - 1]) * o->dx;
  dxm[nxm] = 1.0 / (o->xm[nxm] - o->xm[nxm - 1]);
  double nu_mu = 0.0, nu_th = 0.0;
  for (int k = 0; k < nxm; k++) { disste[k] = 0.0; dissth[k] = 0.0; }
  double *vx = o->vx, *vy = o->vy, *vz = o->vz, *temp = o->temp;
  for (int i = 1; i <= o->nzm; i++) {
    int ip = i + 1;
    for (int j = 1; j <= o->nym; j++) {
      int jp = j + 1;
      for (int k = 1; k <= nxm; k++) {
        int kp = k + 1;
        double hxx = (HX(vx, kp, j, i) - HX(vx, k, j, i)) * dxc[k];
        double hxy = (HX(vx, k, jp, i) - HX(vx, k, j, i)) * o->dy;
        double hxz = (HX(vx, k, j, ip) - HX(vx, k, j, i)) * o->dz;
        double hyx = (HX(vy, kp, j, i) - HX(vy, k, j, i)) * dxm[k];
        double hyy = (HX(vy, k, jp, i) - HX(vy, k, j, i)) * o->dy;
        double hyz = (HX(vy, k, j, ip) - HX(vy, k, j, i)) * o->dz;
        double hzx = (HX(vz, kp, j, i) - HX(vz, k, j, i)) * dxm[k];
        double hzy = (HX(vz, k, jp, i) - HX(vz, k, j, i)) * o->dy;
        double hzz = (HX(vz, k, j, ip) - HX(vz, k, j, i)) * o->dz;
        double dissipte = (hxx * hxx + hxy * hxy + hxz * hxz);
        double dissipte2 = (hyx * hyx + hyy * hyy + hyz * hyz) + (hzx * hzx + hzy * hzy + hzz * hzz);
        nu_mu = nu_mu + dissipte * lem[k] * o->pra + dissipte2 * lec[k] * o->pra;
        double tx = (HX(temp, kp, j, i) - HX(temp, k, j, i)) * dxc[k];
        double ty = (HX(temp, k, jp, i) - HX(temp, k, j, i)) * o->dy;
        double tz = (HX(temp, k, j, ip) - HX(temp, k, j, i)) * o->dz;
        double dissipth = tx * tx + ty * ty + tz * tz;
        nu_th = nu_th + dissipth * lem[k];
        disste[k - 1] = disste[k - 1] + (dissipte + dissipte2) / (o->ren * (double)o->nym * (double)o->nzm);
        dissth[k - 1] = dissth[k - 1] + dissipth / (o->pec * (double)o->nym * (double)o->nzm);
      }
    }
  }
  double volt = 1.0 / ((double)nxm * (double)o->nzm * (double)o->nym);
  *nu_mu_out = nu_mu * volt + 1;
  *nu_th_out = nu_th * volt;
  free(dxm); free(dxc); free(lem); free(lec);
}

/* ---- restart regridding -------------------------------------------------- */

/* gridnew (interp.F90:254-299), 0-based outputs.  Uniform faces for istr 0;
 * for istr 6 the clipped Chebyshev nodes, with nclip = int(str) nodes
 * dropped at each end of n + 2 nclip.  Centres are face midpoints and the
 * last one is extrapolated, rm(n) = 2 rc(n) - rm(n-1). */
int afo_gridnew(int n, double rext, int istr, double str, double *rc, double *rm) {
  if (istr == 0) {
    for (int i = 1; i <= n; i++) {
      double x3 = (double)(i - 1) / (double)(n - 1);
      rc[i - 1] = rext * x3;
    }
  } else if (istr == 6) {
    const double pi = 2.0 * asin(1.0);
    const int nclip = (int)str;
    if (nclip < 0 || nclip > 200) return 2;        /* etazm(1:n+400) */
    const int nxmo = n + nclip + nclip;
    double *etaz = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    for (int i = 1; i <= n; i++) {
      const int ii = i + nclip;
      etaz[i] = +cos(pi * ((double)ii - 0.5) / (double)nxmo);
    }
    const double delet = etaz[1] - etaz[n];
    for (int i = 1; i <= n; i++) etaz[i] = etaz[i] / (0.5 * delet);
    rc[0] = 0.0;
    for (int i = 2; i <= n - 1; i++) rc[i - 1] = rext * (1. - etaz[i]) * 0.5;
    rc[n - 1] = rext;
    free(etaz);
  } else {
    return 2;
  }
  for (int i = 1; i <= n - 1; i++) rm[i - 1] = (rc[i - 1] + rc[i]) * 0.5;
  rm[n - 1] = 2 * rc[n - 1] - rm[n - 2];
  return 0;
}

/* Old grid, 0..n+1 with the second-order end extrapolation
 * (interp.F90:47-62). */
static int old_axis(int n, double rext, int istr, double str, double *c, double *m) {
  int rc = afo_gridnew(n, rext, istr, str, c + 1, m + 1);
  if (rc) return rc;
  c[0] = 2 * c[1] - c[2];
  c[n + 1] = 2 * c[n] - c[n - 1];
  m[0] = 2 * m[1] - m[2];
  m[n + 1] = 2 * m[n] - m[n - 1];
  return 0;
}

/* First l in 1..n with old[l] >= x gives (l-1, l); (n-1, n) when none
 * (interp.F90:139-170). */
static void bracket(const double *old, int n, double x, int *lo, int *hi) {
  *lo = n - 1;
  *hi = n;
  for (int l = 1; l <= n; l++)
    if (old[l] >= x) {
      *lo = l - 1;
      *hi = l;
      return;
    }
}

int afo_interp(const double *old, int nxo, int nyo, int nzo, int istro3, double stro3,
               int nxm, int nym, int nzm, int istr3, double str3, double ylen, double zlen,
               int intvar, double *out) {
  const int nx = nxm + 1, ny = nym + 1, nz = nzm + 1;
  const int nyom = nyo - 1, nzom = nzo - 1;
  if (nz > nzo * 2 || ny > nyo * 2 || nx > nxo * 2) return 1;
  if (intvar < 1 || intvar > 4) return 2;
  /* window of ReadFlowField.F90:91-99 at one rank (xstart = 1, xend = n):
   * floor(real(1*nyom/nym)) is an integer division */
  int js = nyom / nym, je = nym * nyom / nym, is = nzom / nzm, ie = nzm * nzom / nzm;
  if (js < 1) js = 1;
  if (je > nyom) je = nyom;
  if (is < 1) is = 1;
  if (ie > nzom) ie = nzom;
  if (je - js < 1 || ie - is < 1 || nxo < 2) return 2;
  double *tco = dvec(nzo), *tmo = dvec(nzo), *rco = dvec(nyo), *rmo = dvec(nyo), *zzo = dvec(nxo),
         *zmo = dvec(nxo);
  double *tcn = dvec(nz), *tmn = dvec(nz), *rcn = dvec(ny), *rmn = dvec(ny), *zzn = dvec(nx), *zmn = dvec(nx);
  int rc = old_axis(nzo, zlen, 0, 1.0, tco, tmo);
  if (!rc) rc = old_axis(nyo, ylen, 0, 1.0, rco, rmo);
  if (!rc) rc = old_axis(nxo, 1.0, istro3, stro3, zzo, zmo);
  if (!rc) rc = afo_gridnew(nz, zlen, 0, 1.0, tcn + 1, tmn + 1);
  if (!rc) rc = afo_gridnew(ny, ylen, 0, 1.0, rcn + 1, rmn + 1);
  if (!rc) rc = afo_gridnew(nx, 1.0, istr3, str3, zzn + 1, zmn + 1);
  /* interp.F90:73-101: "x" is the third index (z), "z" the first (x) */
  const double *xo = intvar == 1 ? tco : tmo, *xn = intvar == 1 ? tcn : tmn;
  const double *yo = intvar == 2 ? rco : rmo, *yn = intvar == 2 ? rcn : rmn;
  const double *zo = intvar == 3 ? zzo : zmo, *zn = intvar == 3 ? zzn : zmn;
  /* padded old array arrold(0:nxo+1, js-1:je+1, is-1:ie+1), zero where the
   * reference leaves it unset (edges and corners of the halo) */
  const int n1 = nxo + 2, n2 = je - js + 3, n3 = ie - is + 3;
  double *a = rc ? NULL : (double *)calloc((size_t)n1 * n2 * n3, sizeof(double));
#define AO(k, j, i) a[(((size_t)((i) - (is - 1)) * n2 + ((j) - (js - 1))) * n1) + (k)]
  if (a) {
    for (int i = is; i <= ie; i++)
      for (int j = js; j <= je; j++)
        for (int k = 1; k <= nxo; k++) AO(k, j, i) = old[((size_t)(i - 1) * nyom + (j - 1)) * nxo + (k - 1)];
    /* interp.F90:103-124 */
    for (int i = is; i <= ie; i++)
      for (int j = js; j <= je; j++) {
        AO(0, j, i) = 2.0 * AO(1, j, i) - AO(2, j, i);
        AO(nxo + 1, j, i) = 2.0 * AO(nxo, j, i) - AO(nxo - 1, j, i);
      }
    for (int j = js; j <= je; j++)
      for (int k = 1; k <= nxo; k++) {
        AO(k, j, is - 1) = 2.0 * AO(k, j, is) - AO(k, j, is + 1);
        AO(k, j, ie + 1) = 2.0 * AO(k, j, ie) - AO(k, j, ie - 1);
      }
    for (int i = is; i <= ie; i++)
      for (int k = 1; k <= nxo; k++) {
        AO(k, js - 1, i) = 2.0 * AO(k, js, i) - AO(k, js + 1, i);
        AO(k, je + 1, i) = 2.0 * AO(k, je, i) - AO(k, je - 1, i);
      }
  }
  for (int i = 1; i <= nzm && !rc; i++) {
    int bifi, bici;
    const double bix = xn[i];
    bracket(xo, nzo, bix, &bifi, &bici);
    if (bifi < is - 1 || bici > ie + 1) rc = 3;
    for (int j = 1; j <= nym && !rc; j++) {
      int bifj, bicj;
      const double biy = yn[j];
      bracket(yo, nyo, biy, &bifj, &bicj);
      if (bifj < js - 1 || bicj > je + 1) {
        rc = 3;
        break;
      }
      for (int k = 1; k <= nx; k++) {
        int bifk, bick;
        const double biz = zn[k];
        bracket(zo, nxo, biz, &bifk, &bick);
        const double bix1 = xo[bifi], bix2 = xo[bici];
        const double biy1 = yo[bifj], biy2 = yo[bicj];
        const double biz1 = zo[bifk], biz2 = zo[bick];
        /* interp.F90:182-189, with its corner indices as written */
        const double aficjfk = AO(bifk, bicj, bifi);
        const double afifjfk = AO(bifk, bifj, bifi);
        const double acicjfk = AO(bifk, bicj, bici);
        const double acifjfk = AO(bifk, bifj, bici);
        const double aficjck = AO(bick, bicj, bifi);
        const double afifjck = AO(bifk, bifj, bifi);
        const double acicjck = AO(bifk, bicj, bici);
        const double acifjck = AO(bifk, bifj, bici);
        /* interptrilin (interp.F90:210-252) */
        const double dxdydz = 1.0 / ((bix2 - bix1) * (biy2 - biy1) * (biz2 - biz1));
        out[((size_t)(i - 1) * nym + (j - 1)) * nx + (k - 1)] =
            (aficjfk * (bix2 - bix) * (biy - biy1) * (biz2 - biz) +
             afifjfk * (bix2 - bix) * (biy2 - biy) * (biz2 - biz) +
             acicjfk * (bix - bix1) * (biy - biy1) * (biz2 - biz) +
             acifjfk * (bix - bix1) * (biy2 - biy) * (biz2 - biz) +
             aficjck * (bix2 - bix) * (biy - biy1) * (biz - biz1) +
             afifjck * (bix2 - bix) * (biy2 - biy) * (biz - biz1) +
             acicjck * (bix - bix1) * (biy - biy1) * (biz - biz1) +
             acifjck * (bix - bix1) * (biy2 - biy) * (biz - biz1)) *
            dxdydz;
      }
    }
  }
#undef AO
  free(a);
  free(tco); free(tmo); free(rco); free(rmo); free(zzo); free(zmo);
  free(tcn); free(tmn); free(rcn); free(rmn); free(zzn); free(zmn);
  return rc;
}
