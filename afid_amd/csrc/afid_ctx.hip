// afid_ctx.hip -- host side of libafid_hip: context lifetime, metric set-up,
// per-substep factorisations, FFT plans, and the C-ABI entry points of
// include/afid.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/afid.h"
#include "afid_internal.h"
#include "kernels.h"

using namespace afid;

namespace {

struct PendingTiming {
  std::string name;
  hipEvent_t a, b;
};

struct DevPlan {
  double2 *tw = nullptr;
  int *rev = nullptr;
  double2 *post = nullptr;
};

}  // namespace

struct afid_ctx {
  afid_params p{};
  Geom g{};
  int nymh = 0, ldm = 0;
  hipStream_t st = nullptr;
  int last_hip = 0;
  // host metrics, Fortran 1-based (index 0 unused), CreateGrid.F90 / InitPressureSolver.F90
  std::vector<double> xc, xm, g3rc, g3rm, udx3c, udx3m, ap3ck, ac3ck, am3ck, ap3sk, ac3sk, am3sk, ap3ssk,
      ac3ssk, am3ssk, amphk, acphk, apphk, ak1, ak2;
  double dx = 0, dy = 0, dz = 0, dxq = 0, dyq = 0, dzq = 0, ren = 0, pec = 0, pi = 0;
  double gam[4] = {0}, rom[4] = {0}, alm[4] = {0};
  double al = 0, ga = 0, ro = 0, dt = 0, beta = 0, time = 0;
  int singular_mode = 0;       // zgttrf info of the (0,0) pressure mode (>0: reference aborts)
  int dev = -1;                // HIP device the context was created on
  // device state
  double *f[AFID_NFIELDS] = {nullptr};
  double *bp[4] = {nullptr};   // substep update: right-hand sides b + A q_old of vx, vy, vz, temp;
                               // after the column solve, u* (the velocities before the correction)
  // device step control (afid_steps): step state, factorisation inputs/scratch,
  // factor rows per substep, and the captured graph of one step
  DevStep *dstep = nullptr;
  double *mc = nullptr, *fwork = nullptr;
  double *fac_dev[3] = {nullptr, nullptr, nullptr};
  StepConsts kconst{};
  hipGraphExec_t step_graph = nullptr;
  bool capturing = false;
  // statistics (StatRoutines.F90): device accumulators [STATS_N][nxm], their
  // block partials, the dissipation profiles (host) and the sample count
  double *stat_acc = nullptr, *stat_partial = nullptr;
  std::vector<double> disste_acc, dissth_acc;
  int nstatsamples = 0;
  int *d_kslab = nullptr;
  double *slab_buf = nullptr;
  int slab_n = 0;
  double *stale = nullptr;     // rhs(1,j,i)
  double *kc = nullptr;
  double *fac_x = nullptr;     // per-routine factors (SYS_N*FR_N*ldx), current coefficients
  double *fac_sub[4] = {nullptr};  // fused path, per substep ns = 1..3
  double fac_sub_dt[4] = {-1, -1, -1, -1};
  double fac_x_key[2] = {-1, -1};
  double *pin_fac = nullptr;   // pinned staging for factor uploads
  hipEvent_t fac_ev = nullptr;
  double2 *S = nullptr;
  double *piv = nullptr;       // pressure x-solve pivots at chunk starts (per mode)
  double *nat = nullptr;       // natural-order arrays: g3rc, xc, amphk, acphk, apphk, ak1, ak2, dxc, dxm, lem, lec
  double *partial = nullptr, *out = nullptr;
  double *stage = nullptr;
  size_t stage_bytes = 0;
  size_t bytes_held = 0;
  bool dev_steps_ready = false;   // device step state set and the step graph captured
  bool mixed = false;             // AFID_PREC_MIXED: FP32 histories and FFT spectrum
  int zm = 0;                     // z planes per stencil block chosen by afid_tune (0: default)
  bool cfl_fold = false;          // device steps take CalcMaxCFL from the last correction (FP64 only)
  DevPlan dpy, dpz, dpzq;
  FFTPlan pzq{};                  // Q = nzm/16-point plan of the matrix-core z transform
  bool zfft_mfma = false;         // AFID_ZFFT=mfma: z transforms with a 16-point DFT stage on MFMA
  RealPlan py{};
  FFTPlan pz{};
  // natural-order device array views
  const double *d_g3rc = nullptr, *d_xc = nullptr, *d_amphk = nullptr, *d_acphk = nullptr, *d_apphk = nullptr,
               *d_ak1 = nullptr, *d_ak2 = nullptr, *d_dxc = nullptr, *d_dxm = nullptr, *d_lem = nullptr,
               *d_lec = nullptr;
  // profiling
  bool prof = false;
  std::vector<PendingTiming> pending;
  // Device steps while profiling replay a second capture of the step graph
  // whose kernels are bracketed by event-record nodes; its event pairs are
  // re-recorded at every replay and read after each one.
  bool prof_capture = false;
  hipGraphExec_t prof_graph = nullptr;
  std::vector<PendingTiming> graph_timing;
  std::map<std::string, std::pair<double, int64_t>> prof_tot;
};

#define HIPCHK(ctx, call)                       \
  do {                                          \
    hipError_t e_ = (call);                     \
    if (e_ != hipSuccess) {                     \
      (ctx)->last_hip = (int)e_;                \
      return AFID_ERR_HIP;                      \
    }                                           \
  } while (0)

namespace {

// ---------------------------------------------------------------- metrics
// CreateGrid.F90:11-269 and InitPressureSolver.F90:12-66 (coefficients only).
void create_grid(afid_ctx *c) {
  const afid_params &p = c->p;
  const int nx = p.nxm + 1, nxm = p.nxm, nym = p.nym, nzm = p.nzm;
  auto vec = [&](int n) { return std::vector<double>(n + 2, 0.0); };
  c->xc = vec(nx); c->xm = vec(nx); c->g3rc = vec(nx); c->g3rm = vec(nx);
  c->udx3c = vec(nx); c->udx3m = vec(nx);
  c->ap3ck = vec(nx); c->ac3ck = vec(nx); c->am3ck = vec(nx);
  c->ap3sk = vec(nx); c->ac3sk = vec(nx); c->am3sk = vec(nx);
  c->ap3ssk = vec(nx); c->ac3ssk = vec(nx); c->am3ssk = vec(nx);
  c->amphk = vec(nx); c->acphk = vec(nx); c->apphk = vec(nx);
  c->ak1 = vec(nzm); c->ak2 = vec(nym);
  std::vector<int> kmv(nx + 2), kpv(nx + 2), kmc(nx + 2), kpc(nx + 2);
  for (int kc = 1; kc <= nxm; kc++) {
    kmv[kc] = kc - 1;
    kpv[kc] = kc + 1;
    if (kc == 1) kmv[kc] = kc;
    if (kc == nxm) kpv[kc] = kc;
    kpc[kc] = kpv[kc] - kc;
    kmc[kc] = kc - kmv[kc];
  }
  c->pi = 2.0 * std::asin(1.0);
  std::vector<double> &xc = c->xc;
  if (p.istr3 == 0) {
    for (int kc = 1; kc <= nx; kc++) xc[kc] = p.alx3 * ((double)(kc - 1) / (double)nxm);
  } else if (p.istr3 == 4) {
    double tstr3 = std::tanh(p.str3);
    xc[1] = 0.0;
    for (int kc = 2; kc <= nx; kc++) {
      double z2dp = (double)(2 * kc - nx - 1) / (double)nxm;
      xc[kc] = (1 + std::tanh(p.str3 * z2dp) / tstr3) * 0.5 * p.alx3;
    }
  } else if (p.istr3 == 6) {
    int nclip = (int)p.str3;
    int nxmo = nx + nclip + nclip;
    std::vector<double> etazm(nxmo + 2), etaz(nx + 2);
    for (int kc = 1; kc <= nxmo; kc++) etazm[kc] = +std::cos(c->pi * ((double)kc - 0.5) / (double)nxmo);
    for (int kc = 1; kc <= nx; kc++) etaz[kc] = etazm[kc + nclip];
    double delet = etaz[1] - etaz[nx];
    for (int kc = 1; kc <= nx; kc++) etaz[kc] = etaz[kc] / (0.5 * delet);
    xc[1] = 0.0;
    for (int kc = 2; kc <= nxm; kc++) xc[kc] = p.alx3 * (1.0 - etaz[kc]) * 0.5;
    xc[nx] = p.alx3;
  }
  c->dx = (double)nxm / p.alx3;
  c->dy = (double)nym / p.ylen;
  c->dz = (double)nzm / p.zlen;
  c->dxq = c->dx * c->dx;
  c->dyq = c->dy * c->dy;
  c->dzq = c->dz * c->dz;
  for (int kc = 1; kc <= nxm; kc++) {
    c->xm[kc] = (xc[kc] + xc[kc + 1]) * 0.5;
    c->g3rm[kc] = (xc[kc + 1] - xc[kc]) * c->dx;
  }
  for (int kc = 2; kc <= nxm; kc++) c->g3rc[kc] = (xc[kc + 1] - xc[kc - 1]) * c->dx * 0.5;
  c->g3rc[1] = (xc[2] - xc[1]) * c->dx;
  c->g3rc[nx] = (xc[nx] - xc[nxm]) * c->dx;
  for (int kc = 1; kc <= nxm; kc++) {
    c->udx3m[kc] = c->dx / c->g3rm[kc];
    c->udx3c[kc] = c->dx / c->g3rc[kc];
  }
  c->udx3c[nx] = c->dx / c->g3rc[nx];
  c->am3ck[1] = 0.0; c->ap3ck[1] = 0.0; c->ac3ck[1] = 1.0;
  c->am3ck[nx] = 0.0; c->ap3ck[nx] = 0.0; c->ac3ck[nx] = 1.0;
  for (int kc = 2; kc <= nxm; kc++) {
    double a33 = c->dxq / c->g3rc[kc];
    double a33p = 1.0 / c->g3rm[kc], a33m = 1.0 / c->g3rm[kc - 1];
    c->ap3ck[kc] = a33 * a33p;
    c->am3ck[kc] = a33 * a33m;
    c->ac3ck[kc] = -(c->ap3ck[kc] + c->am3ck[kc]);
  }
  for (int kc = 2; kc <= nxm - 1; kc++) {
    double a33 = c->dxq / c->g3rm[kc];
    double a33p = +a33 / c->g3rc[kc + 1], a33m = +a33 / c->g3rc[kc];
    c->ap3sk[kc] = a33p;
    c->am3sk[kc] = a33m;
    c->ac3sk[kc] = -(c->ap3sk[kc] + c->am3sk[kc]);
  }
  {
    int kc = 1;
    double a33 = c->dxq / c->g3rm[kc];
    double a33p = +a33 / c->g3rc[kc + 1], a33m = +a33 / c->g3rc[kc];
    c->ap3sk[kc] = a33p;
    c->am3sk[kc] = 0.0;
    c->ac3sk[kc] = -(a33p + p.inslws * a33m * 2.0);
  }
  {
    int kc = nxm;
    double a33 = c->dxq / c->g3rm[kc];
    double a33p = +a33 / c->g3rc[kc + 1], a33m = +a33 / c->g3rc[kc];
    c->am3sk[kc] = a33m;
    c->ap3sk[kc] = 0.0;
    c->ac3sk[kc] = -(a33m + p.inslwn * a33p * 2.0);
  }
  c->am3ssk[1] = 0.0; c->ap3ssk[1] = 0.0; c->ac3ssk[1] = 1.0;
  for (int kc = 2; kc <= nxm; kc++) {
    double a33 = c->dxq / c->g3rc[kc];
    double a33p = 1.0 / c->g3rm[kc], a33m = 1.0 / c->g3rm[kc - 1];
    c->ap3ssk[kc] = a33 * a33p;
    c->am3ssk[kc] = a33 * a33m;
    c->ac3ssk[kc] = -(c->ap3ssk[kc] + c->am3ssk[kc]);
  }
  c->am3ssk[nx] = 0.0; c->ap3ssk[nx] = 0.0; c->ac3ssk[nx] = 1.0;
  // InitPressureSolver.F90:22-59
  int nymh = nym / 2 + 1, nzmh = nzm / 2 + 1;
  for (int i = 1; i <= nzm; i++) {
    double ao = i <= nzmh ? (i - 1) * 2.0 * c->pi : -(nzm - i + 1) * 2.0 * c->pi;
    double f = (double)nzm / p.zlen;
    c->ak1[i] = 2.0 * (1.0 - std::cos(ao / nzm)) * (f * f);
  }
  for (int j = 1; j <= nym; j++) {
    double ap = j <= nymh ? (j - 1) * 2.0 * c->pi : -(nym - j + 1) * 2.0 * c->pi;
    double f = (double)nym / p.ylen;
    c->ak2[j] = 2.0 * (1.0 - std::cos(ap / nym)) * (f * f);
  }
  for (int kc = 1; kc <= nxm; kc++) {
    int kp = kpv[kc];
    double a33icc = kmc[kc] * c->dxq / c->g3rc[kc];
    double a33icp = kpc[kc] * c->dxq / c->g3rc[kp];
    double ugmmm = 1.0 / c->g3rm[kc];
    c->amphk[kc] = a33icc * ugmmm;
    c->apphk[kc] = a33icp * ugmmm;
    c->acphk[kc] = -(c->amphk[kc] + c->apphk[kc]);
  }
  // InitTimeMarchScheme.F90:16-49
  if (p.nsst > 1) {
    c->gam[1] = 8.0 / 15.0; c->gam[2] = 5.0 / 12.0; c->gam[3] = 3.0 / 4.0;
    c->rom[1] = 0.0; c->rom[2] = -17.0 / 60.0; c->rom[3] = -5.0 / 12.0;
  } else {
    c->gam[1] = 1.5; c->gam[2] = 0.0; c->gam[3] = 0.0;
    c->rom[1] = -0.5; c->rom[2] = 0.0; c->rom[3] = 0.0;
  }
  for (int ns = 1; ns <= 3; ns++) c->alm[ns] = c->gam[ns] + c->rom[ns];
  c->ren = std::sqrt(p.ray / p.pra);
  c->pec = std::sqrt(p.pra * p.ray);
}

bool fft_factors(int n, int *f, int *nf) {
  int k = 0;
  while (n % 4 == 0) { f[k++] = 4; n /= 4; }
  while (n % 2 == 0) { f[k++] = 2; n /= 2; }
  while (n % 3 == 0) { f[k++] = 3; n /= 3; }
  while (n % 5 == 0) { f[k++] = 5; n /= 5; }
  *nf = k;
  return n == 1 && k <= 16;
}

int make_plan(afid_ctx *c, int n, FFTPlan *pl, DevPlan *dp, int realN) {
  pl->n = n;
  if (!fft_factors(n, pl->f, &pl->nf)) return AFID_ERR_FFTSIZE;
  for (int s = 0, m = 1; s < pl->nf; m *= pl->f[s], s++)
    pl->mdiv[s] = m == 1 ? 0u : (unsigned)(((1ull << 32) + (uint64_t)m - 1) / (uint64_t)m);
  std::vector<double2> tw(n);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int t = 0; t < n; t++) {
    long double a = twopi * (long double)t / (long double)n;
    tw[t] = make_double2((double)cosl(a), (double)(-sinl(a)));
  }
  // Position q of the in-place decimation-in-time input holds element rev(q)
  // (mixed-radix digit reversal).  The kernels scatter element t to position
  // rev^-1(t), so the table uploaded is the inverse permutation (rev is its own
  // inverse only when all radices are equal).
  std::vector<int> rev(n);
  for (int q = 0; q < n; q++) {
    int t = q, r = 0, rem = n;
    for (int s = 0; s < pl->nf; s++) {
      int d = t % pl->f[s];
      t /= pl->f[s];
      rem /= pl->f[s];
      r += d * rem;
    }
    rev[r] = q;
  }
  HIPCHK(c, hipMalloc(&dp->tw, sizeof(double2) * n));
  HIPCHK(c, hipMalloc(&dp->rev, sizeof(int) * n));
  HIPCHK(c, hipMemcpy(dp->tw, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(dp->rev, rev.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  c->bytes_held += (sizeof(double2) + sizeof(int)) * n;
  pl->tw = dp->tw;
  pl->rev = dp->rev;
  if (realN > 0) {
    int h = realN / 2;
    std::vector<double2> post(h + 1);
    for (int m = 0; m <= h; m++) {
      long double a = twopi * (long double)m / (long double)realN;
      post[m] = make_double2((double)cosl(a), (double)(-sinl(a)));
    }
    HIPCHK(c, hipMalloc(&dp->post, sizeof(double2) * (h + 1)));
    HIPCHK(c, hipMemcpy(dp->post, post.data(), sizeof(double2) * (h + 1), hipMemcpyHostToDevice));
  }
  return AFID_OK;
}

// Interleave a natural-order array v(k), k = 0..n-1, into an ldx row.
void interleave(const Geom &g, const double *v, int n, double *row, double pad = 0.0) {
  for (int s = 0; s < g.ldx; s++) row[s] = pad;
  for (int k = 0; k < n; k++) row[pos_of_k(k, g.L)] = v[k];
}

// LU of one wall-normal system (unit diagonal rows, dl/du given per row in
// natural order), written as the FR_N interleaved rows.  Returns
// AFID_NUM_PIVOT when dgttrf would interchange rows.
int factor_system(const Geom &g, int n, const std::vector<double> &dl, const std::vector<double> &du,
                  const std::vector<double> &scale, double *rows) {
  std::vector<double> l(n, 0.0), u(n, 1.0), rinv(n), e(n, 0.0), G(n), H(n);
  int rc = AFID_OK;
  for (int k = 1; k < n; k++) {
    if (std::fabs(u[k - 1]) < std::fabs(dl[k])) rc = AFID_NUM_PIVOT;
    l[k] = dl[k] / u[k - 1];
    u[k] = 1.0 - l[k] * du[k - 1];
  }
  for (int k = 0; k < n; k++) rinv[k] = 1.0 / u[k];
  for (int k = 0; k + 1 < n; k++) e[k] = du[k] * rinv[k];
  for (int p = 0; p < P; p++) {
    int k0 = p * g.L, k1 = std::min(n, k0 + g.L);
    double gg = 1.0;
    for (int k = k0; k < k1; k++) { gg *= -l[k]; G[k] = gg; }
    double hh = 1.0;
    for (int k = k1 - 1; k >= k0; k--) { hh *= -e[k]; H[k] = hh; }
  }
  interleave(g, scale.data(), n, rows + FR_SCALE * g.ldx);
  interleave(g, l.data(), n, rows + FR_L * g.ldx);
  interleave(g, rinv.data(), n, rows + FR_RINV * g.ldx);
  interleave(g, e.data(), n, rows + FR_E * g.ldx);
  interleave(g, G.data(), n, rows + FR_G * g.ldx);
  interleave(g, H.data(), n, rows + FR_H * g.ldx);
  std::vector<double> dlk(n, 0.0), duk(n, 0.0);
  for (int k = 1; k < n; k++) dlk[k] = dl[k];
  for (int k = 0; k + 1 < n; k++) duk[k] = du[k];
  interleave(g, dlk.data(), n, rows + FR_DL * g.ldx);
  interleave(g, duk.data(), n, rows + FR_DU * g.ldx);
  return rc;
}

// The three wall-normal systems of one substep (SolveImpEqnUpdate_X.F90:27-44,
// _YZ.F90:25-38, _Temp.F90:31-47) for coefficients (al, dt).
int build_factors(afid_ctx *c, double al, double dt, double *rows) {
  const int nx = c->g.nx, nxm = c->g.nxm;
  const double beta = dt / c->ren * 0.5;
  int rc = AFID_OK;
  {  // X: n = nx, rows 1 and nx identity
    double betadx = beta * al;
    std::vector<double> dl(nx, 0.0), du(nx, 0.0), sc(nx, 0.0);
    for (int kc = 2; kc <= nxm; kc++) {
      double ackl_b = 1.0 / (1.0 - c->ac3ck[kc] * betadx);
      dl[kc - 1] = -c->am3ssk[kc] * betadx * ackl_b;
      du[kc - 1] = -c->ap3ssk[kc] * betadx * ackl_b;
      sc[kc - 1] = 1.0 / (1.0 - c->ac3ck[kc] * betadx);
    }
    int r = factor_system(c->g, nx, dl, du, sc, rows + (SYS_X * FR_N) * c->g.ldx);
    if (r) rc = r;
  }
  {  // YZ: n = nxm
    double betadx = beta * al;
    std::vector<double> dl(nxm, 0.0), du(nxm, 0.0), sc(nxm, 0.0);
    for (int kc = 1; kc <= nxm; kc++) {
      double ackl_b = 1.0 / (1.0 - c->ac3sk[kc] * betadx);
      dl[kc - 1] = kc > 1 ? -c->am3sk[kc] * betadx * ackl_b : 0.0;
      du[kc - 1] = kc < nxm ? -c->ap3sk[kc] * betadx * ackl_b : 0.0;
      sc[kc - 1] = 1.0 / (1.0 - c->ac3sk[kc] * betadx);
    }
    int r = factor_system(c->g, nxm, dl, du, sc, rows + (SYS_YZ * FR_N) * c->g.ldx);
    if (r) rc = r;
  }
  {  // Temp: n = nx, betadx = al*dt/(2 pec)
    double betadx = 0.5 * al * dt / c->pec;
    std::vector<double> dl(nx, 0.0), du(nx, 0.0), sc(nx, 0.0);
    for (int kc = 2; kc <= nxm; kc++) {
      double ackl_b = 1.0 / (1. - c->ac3ssk[kc] * betadx);
      dl[kc - 1] = -c->am3ssk[kc] * betadx * ackl_b;
      du[kc - 1] = -c->ap3ssk[kc] * betadx * ackl_b;
      sc[kc - 1] = 1.0 / (1.0 - c->ac3ssk[kc] * betadx);
    }
    int r = factor_system(c->g, nx, dl, du, sc, rows + (SYS_T * FR_N) * c->g.ldx);
    if (r) rc = r;
  }
  return rc;
}

// zgttrf of the (ky,kz)=(0,0) pressure system in netlib order (the complex
// factorisation of real coefficients equals the real one).  Only this mode can
// be singular: its rows sum to zero, and whether rounding leaves an exactly
// zero last pivot decides whether the reference aborts
// (SolvePressureCorrection.F90:128-134).  Data independent, so done once.
int zero_mode_info(const afid_ctx *c) {
  const int n = c->g.nxm;
  std::vector<double> dl(n), d(n, 1.0), du(n), amph(n + 2), apph(n + 2);
  for (int k = 1; k <= n; k++) {
    double b = 1.0 / (c->acphk[k] - c->ak2[1] - c->ak1[1]);
    apph[k] = c->apphk[k] * b;
    amph[k] = c->amphk[k] * b;
  }
  for (int k = 0; k < n - 1; k++) {
    dl[k] = amph[k + 2];
    du[k] = apph[k + 1];
  }
  for (int i = 0; i < n - 1; i++) {
    if (std::fabs(d[i]) >= std::fabs(dl[i])) {
      if (d[i] != 0.0) {
        double fact = dl[i] / d[i];
        d[i + 1] = d[i + 1] - fact * du[i];
      }
    } else {
      double fact = d[i] / dl[i];
      d[i] = dl[i];
      double temp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = temp - fact * d[i + 1];
      if (i < n - 2) du[i + 1] = -fact * du[i + 1];
    }
  }
  for (int i = 0; i < n; i++)
    if (d[i] == 0.0) return i + 1;
  return 0;
}

size_t fac_bytes(const afid_ctx *c) { return sizeof(double) * SYS_N * FR_N * (size_t)c->g.ldx; }

// Upload factors for (al, dt) into dst through the pinned staging buffer.
int upload_factors(afid_ctx *c, double al, double dt, double *dst) {
  HIPCHK(c, hipEventSynchronize(c->fac_ev));  // previous upload has left the staging buffer
  int rc = build_factors(c, al, dt, c->pin_fac);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(dst, c->pin_fac, fac_bytes(c), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipEventRecord(c->fac_ev, c->st));
  return AFID_OK;
}

StepScalars scalars(const afid_ctx *c) {
  StepScalars s;
  s.al = c->al; s.ga = c->ga; s.ro = c->ro; s.dt = c->dt; s.beta = c->beta;
  s.alre = c->al / c->ren;
  s.alpec = c->al / c->pec;
  s.udy4 = c->dy * 0.25;
  s.udz4 = c->dz * 0.25;
  s.udyq_re = c->dyq / c->ren;
  s.udzq_re = c->dzq / c->ren;
  s.udyq_pe = c->dyq / c->pec;
  s.udzq_pe = c->dzq / c->pec;
  s.dy = c->dy; s.dz = c->dz; s.dyq = c->dyq; s.dzq = c->dzq;
  s.udy_al = c->dy * c->al;
  s.tempbp = c->p.tempbp;
  s.temptp = c->p.temptp;
  s.use_ru = c->ro != 0.0;
  return s;
}

FieldPtrs fields(const afid_ctx *c) {
  FieldPtrs f;
  f.vx = c->f[AFID_VX]; f.vy = c->f[AFID_VY]; f.vz = c->f[AFID_VZ];
  f.temp = c->f[AFID_TEMP]; f.pr = c->f[AFID_PR];
  return f;
}

size_t field_bytes(const afid_ctx *c) { return sizeof(double) * (size_t)c->g.ncol * c->g.ldx; }

int alloc_field(afid_ctx *c, double **p) {
  if (*p) return AFID_OK;
  if (hipMalloc(p, field_bytes(c)) != hipSuccess) {
    *p = nullptr;
    return AFID_ERR_NOMEM;
  }
  c->bytes_held += field_bytes(c);
  HIPCHK(c, hipMemsetAsync(*p, 0, field_bytes(c), c->st));
  return AFID_OK;
}

// ---- profiling helpers
struct Timed {
  afid_ctx *c;
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  bool on;
  Timed(afid_ctx *cc, const char *n) : c(cc), name(n), on(cc->prof && (!cc->capturing || cc->prof_capture)) {
    if (on) {
      hipEventCreate(&a);
      hipEventCreate(&b);
      record(a);
    }
  }
  void record(hipEvent_t e) {
    if (c->capturing) hipEventRecordWithFlags(e, c->st, hipEventRecordExternal);   // an event-record node
    else hipEventRecord(e, c->st);
  }
  ~Timed() {
    if (on) {
      record(b);
      (c->capturing ? c->graph_timing : c->pending).push_back({name, a, b});
    }
  }
};

static void drop_prof_graph(afid_ctx *c) {
  if (c->prof_graph) hipGraphExecDestroy(c->prof_graph);
  c->prof_graph = nullptr;
  for (auto &pt : c->graph_timing) {
    hipEventDestroy(pt.a);
    hipEventDestroy(pt.b);
  }
  c->graph_timing.clear();
}

int kerr(afid_ctx *c, hipError_t e) {
  if (e != hipSuccess) {
    c->last_hip = (int)e;
    return AFID_ERR_HIP;
  }
  return AFID_OK;
}

// from_ustar: the divergence is taken of u* in the b buffers (substep path)
PoissonArgs poisson_args(afid_ctx *c, int from_velocity, int ns = 0, bool dev = false, bool from_ustar = false) {
  PoissonArgs a;
  a.ds = dev ? c->dstep : nullptr;
  a.ns = ns;
  a.g = c->g;
  a.nymh = c->nymh;
  a.ldm = c->ldm;
  a.S = c->S;
  a.piv = c->piv;
  a.kc = c->kc;
  a.amphk = c->d_amphk;
  a.acphk = c->d_acphk;
  a.apphk = c->d_apphk;
  a.ak1 = c->d_ak1;
  a.ak2 = c->d_ak2;
  a.f = fields(c);
  if (from_ustar) {
    a.f.vx = c->bp[0];
    a.f.vy = c->bp[1];
    a.f.vz = c->bp[2];
  }
  a.dph = c->f[AFID_DPH];
  a.udx3m = c->kc + KC_UDX3M * c->g.ldx;
  a.sc = scalars(c);
  a.py = c->py;
  a.pz = c->pz;
  a.pzq = c->pzq;
  a.zmf = c->zfft_mfma ? 1 : 0;
  a.from_velocity = from_velocity;
  a.spec32 = c->mixed;
  return a;
}

int run_poisson(afid_ctx *c, int from_velocity, int ns = 0, bool dev = false, bool from_ustar = false) {
  if (c->singular_mode) return AFID_NUM_SINGULAR;
  PoissonArgs a = poisson_args(c, from_velocity, ns, dev, from_ustar);
  int rc;
  {
    Timed t(c, from_velocity ? "div_yfft_r2c" : "yfft_r2c");
    if ((rc = kerr(c, launch_yfft_forward(a, c->st)))) return rc;
  }
  {
    Timed t(c, "zfft_fwd");
    if ((rc = kerr(c, launch_zfft(a, -1, c->st)))) return rc;
  }
  {
    Timed t(c, "xsolve_pressure");
    if ((rc = kerr(c, launch_xsolve_pressure(a, c->st)))) return rc;
  }
  {
    Timed t(c, "zfft_bwd");
    if ((rc = kerr(c, launch_zfft(a, +1, c->st)))) return rc;
  }
  {
    Timed t(c, "yfft_c2r");
    if ((rc = kerr(c, launch_yfft_inverse(a, c->st)))) return rc;
  }
  return AFID_OK;
}

ReduceArgs reduce_args(afid_ctx *c) {
  ReduceArgs a;
  a.g = c->g;
  a.f = fields(c);
  a.kc = c->kc;
  a.g3rc = c->d_g3rc;
  a.dxc = c->d_dxc; a.dxm = c->d_dxm; a.lem = c->d_lem; a.lec = c->d_lec;
  a.dy = c->dy; a.dz = c->dz; a.ren = c->ren; a.pec = c->pec; a.pra = c->p.pra;
  a.partial = c->partial;
  a.out = c->out;
  return a;
}

int read_out(afid_ctx *c, double *h, int n) {
  HIPCHK(c, hipMemcpyAsync(h, c->out, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return AFID_OK;
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

const char *afid_strerror(int code) {
  switch (code) {
    case AFID_OK: return "ok";
    case AFID_ERR_ARG: return "invalid argument";
    case AFID_ERR_HIP: return "HIP runtime error";
    case AFID_ERR_NOMEM: return "device allocation failed";
    case AFID_ERR_FFTSIZE: return "lateral size must factor into 2, 3 and 5 (errorcode 444)";
    case AFID_ERR_STATE: return "call not valid in this state";
    case AFID_NUM_PIVOT: return "implicit factorisation needs a row interchange";
    case AFID_NUM_SINGULAR: return "singular pressure tridiagonal";
    case AFID_ERR_IO: return "continuation file could not be read or written, or holds a dataset of another shape";
    case AFID_ERR_GRID:
      return "old grid cannot be interpolated from (more than twice its resolution, a stretching other than 0 "
             "or 6, or a point outside its window)";
    default: return "unknown error";
  }
}

int afid_last_hip_error(afid_ctx *ctx) { return ctx ? ctx->last_hip : 0; }

int afid_destroy(afid_ctx *c) {
  if (!c) return AFID_OK;
  if (c->st) hipStreamSynchronize(c->st);
  for (auto &pt : c->pending) {
    hipEventDestroy(pt.a);
    hipEventDestroy(pt.b);
  }
  for (int q = 0; q < AFID_NFIELDS; q++) hipFree(c->f[q]);
  for (int q = 0; q < 4; q++) hipFree(c->bp[q]);
  for (int q = 0; q < 3; q++) hipFree(c->fac_dev[q]);
  if (c->step_graph) hipGraphExecDestroy(c->step_graph);
  drop_prof_graph(c);
  hipFree(c->dstep);
  hipFree(c->mc);
  hipFree(c->fwork);
  hipFree(c->stat_acc);
  hipFree(c->stat_partial);
  hipFree(c->d_kslab);
  hipFree(c->slab_buf);
  for (int q = 0; q < 4; q++) hipFree(c->fac_sub[q]);
  void *ps[] = {c->stale, c->kc, c->fac_x, c->S, c->piv, c->nat, c->partial, c->out, c->stage,
                c->dpy.tw, c->dpy.rev, c->dpy.post, c->dpz.tw, c->dpz.rev, c->dpz.post,
                c->dpzq.tw, c->dpzq.rev, c->dpzq.post};
  for (void *q : ps) hipFree(q);
  if (c->pin_fac) hipHostFree(c->pin_fac);
  if (c->fac_ev) hipEventDestroy(c->fac_ev);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
  return AFID_OK;
}

static int scratch_setup(afid_ctx *c);   // the no-allocation scratch of this device

int afid_create(afid_ctx **out, const afid_params *p, const afid_metrics *m) {
  if (!out || !p) return AFID_ERR_ARG;
  *out = nullptr;
  if (p->nxm < 2 || p->nym < 2 || p->nzm < 2 || p->alx3 <= 0 || p->ylen <= 0 || p->zlen <= 0 ||
      p->ray <= 0 || p->pra <= 0 || (p->precision != AFID_PREC_FP64 && p->precision != AFID_PREC_MIXED))
    return AFID_ERR_ARG;
  afid_ctx *c = new afid_ctx();
  c->p = *p;
  Geom &g = c->g;
  g.nxm = p->nxm;
  g.nx = p->nxm + 1;
  g.nym = p->nym;
  g.nzm = p->nzm;
  g.L = (g.nx + P - 1) / P;
  g.ldx = P * g.L;
  g.ncol = (int64_t)g.nym * g.nzm;
  if (g.L > max_supported_L() || g.nxm > xsolve_max_rows()) {
    delete c;
    return AFID_ERR_ARG;
  }
  c->nymh = g.nym / 2 + 1;
  c->ldm = (c->nymh + 3) / 4 * 4;
  create_grid(c);
  if (m) {  // caller-supplied metrics override the computed ones (1-based arrays)
    const int nx = g.nx, nxm = g.nxm;
    auto take = [](std::vector<double> &v, const double *src, int n) {
      if (src)
        for (int k = 1; k <= n; k++) v[k] = src[k - 1];
    };
    take(c->xc, m->xc, nx); take(c->xm, m->xm, nx); take(c->g3rc, m->g3rc, nx); take(c->g3rm, m->g3rm, nx);
    take(c->udx3c, m->udx3c, nx); take(c->udx3m, m->udx3m, nx);
    take(c->ap3ck, m->ap3ck, nx); take(c->ac3ck, m->ac3ck, nx); take(c->am3ck, m->am3ck, nx);
    take(c->ap3sk, m->ap3sk, nx); take(c->ac3sk, m->ac3sk, nx); take(c->am3sk, m->am3sk, nx);
    take(c->ap3ssk, m->ap3ssk, nx); take(c->ac3ssk, m->ac3ssk, nx); take(c->am3ssk, m->am3ssk, nx);
    take(c->amphk, m->amphk, nxm); take(c->acphk, m->acphk, nxm); take(c->apphk, m->apphk, nxm);
    take(c->ak1, m->ak1, g.nzm); take(c->ak2, m->ak2, g.nym);
  }
  // decided on the coefficients the solve actually uses (caller metrics included)
  c->singular_mode = zero_mode_info(c);
  {
    // AFID_CFL=fold: device steps fold CalcMaxCFL into the last correction.
    // Measured at 384x768x768 that correction takes 5.8 ms against 3.8 in
    // place, more than the 1.55 ms separate reduction it saves, so by default
    // each step graph reduces the fields at its start.
    const char *f = std::getenv("AFID_CFL");
    c->cfl_fold = p->precision == AFID_PREC_FP64 && f && std::strcmp(f, "fold") == 0;
  }
  int rc = AFID_OK;
#define TRY(x)                  \
  do {                          \
    if ((rc = (x)) != AFID_OK) { \
      afid_destroy(c);          \
      return rc;                \
    }                           \
  } while (0)
#define TRYH(x) TRY(kerr(c, (x)))
  TRYH(hipGetDevice(&c->dev));
  TRY(scratch_setup(c));
  TRYH(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  TRYH(hipEventCreateWithFlags(&c->fac_ev, hipEventDisableTiming));
  TRYH(hipEventRecord(c->fac_ev, c->st));
  // FFT plans (checked before the large allocations)
  {
    int N = g.nym, half = (N % 2 == 0);
    c->py.N = N;
    c->py.half = half;
    c->py.H = half ? N / 2 : N;
    TRY(make_plan(c, c->py.H, &c->py.c, &c->dpy, N));
    c->py.post = c->dpy.post;
    TRY(make_plan(c, g.nzm, &c->pz, &c->dpz, 0));
    {   // AFID_ZFFT=mfma: the z transforms' 16-point stage on the matrix cores (FP64)
      const char *zf = std::getenv("AFID_ZFFT");
      if (zf && std::strcmp(zf, "mfma") == 0 && p->precision == AFID_PREC_FP64 && zfft_mfma_fits(g.nzm)) {
        TRY(make_plan(c, g.nzm / 16, &c->pzq, &c->dpzq, 0));
        c->zfft_mfma = true;
      }
    }
    if (c->py.H > 1280 || g.nzm > 2560) TRY(AFID_ERR_ARG);
    // the dynamic-LDS limits are per kernel and process-wide: only ever raised,
    // so that cells created earlier keep what they launch with
    static size_t fft_lds_set = 65536, col_lds_set = 65536;
    size_t lds = std::max(sizeof(double2) * 8 * (size_t)c->py.H, sizeof(double2) * 4 * (size_t)g.nzm);
    if (lds > fft_lds_set) {
      TRYH(fft_configure_lds(lds));
      fft_lds_set = lds;
    }
    size_t clds = std::max(sizeof(double) * 256 * (size_t)g.L, column_lds_need(g));
    if (clds > col_lds_set) {
      TRYH(column_configure_lds(clds));
      col_lds_set = clds;
    }
  }
  for (int q : {AFID_VX, AFID_VY, AFID_VZ, AFID_TEMP, AFID_PR, AFID_RUX, AFID_RUY, AFID_RUZ, AFID_RUTEMP, AFID_DPH})
    TRY(alloc_field(c, &c->f[q]));
  c->mixed = p->precision == AFID_PREC_MIXED;
  for (int q = 0; q < 4; q++) TRY(alloc_field(c, &c->bp[q]));
  TRYH(hipMalloc(&c->stale, sizeof(double) * g.ncol));
  c->bytes_held += sizeof(double) * g.ncol;
  TRYH(hipMemsetAsync(c->stale, 0, sizeof(double) * g.ncol, c->st));
  {
    size_t ns = (size_t)g.nxm * g.nzm * c->ldm;
    if (hipMalloc(&c->S, sizeof(double2) * ns) != hipSuccess) TRY(AFID_ERR_NOMEM);
    c->bytes_held += ns * sizeof(double2);
    int np, lm;
    xsolve_shape(g.nxm, &np, &lm);
    const size_t npiv = (size_t)c->nymh * g.nzm * np;
    if (hipMalloc(&c->piv, sizeof(double) * npiv) != hipSuccess) TRY(AFID_ERR_NOMEM);
    c->bytes_held += npiv * sizeof(double);
  }
  // interleaved coefficient rows
  {
    std::vector<double> rows((size_t)KC_N * g.ldx, 0.0), v(g.nx);
    auto put = [&](int row, const std::vector<double> &src, int n) {
      for (int k = 0; k < n; k++) v[k] = src[k + 1];
      interleave(g, v.data(), n, rows.data() + (size_t)row * g.ldx);
    };
    put(KC_UDX3C, c->udx3c, g.nx); put(KC_UDX3M, c->udx3m, g.nxm);
    put(KC_AM3CK, c->am3ck, g.nx); put(KC_AC3CK, c->ac3ck, g.nx); put(KC_AP3CK, c->ap3ck, g.nx);
    put(KC_AM3SK, c->am3sk, g.nx); put(KC_AC3SK, c->ac3sk, g.nx); put(KC_AP3SK, c->ap3sk, g.nx);
    put(KC_AMPHK, c->amphk, g.nxm); put(KC_ACPHK, c->acphk, g.nxm); put(KC_APPHK, c->apphk, g.nxm);
    TRYH(hipMalloc(&c->kc, sizeof(double) * rows.size()));
    c->bytes_held += sizeof(double) * rows.size();
    TRYH(hipMemcpy(c->kc, rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice));
  }
  // natural-order arrays
  {
    const int nx = g.nx, nxm = g.nxm;
    std::vector<double> nat;
    std::vector<size_t> off;
    auto add = [&](const double *src, int n) {
      off.push_back(nat.size());
      nat.insert(nat.end(), src, src + n);
    };
    add(c->g3rc.data() + 1, nx);
    add(c->xc.data() + 1, nx);
    add(c->amphk.data() + 1, nxm);
    add(c->acphk.data() + 1, nxm);
    add(c->apphk.data() + 1, nxm);
    add(c->ak1.data() + 1, g.nzm);
    add(c->ak2.data() + 1, g.nym);
    // CalcDissipationNu.F90:44-55 geometric factors
    std::vector<double> dxm(nxm), dxc(nxm), lem(nxm), lec(nxm);
    for (int k = 1; k <= nxm; k++) {
      int kp = k + 1;
      dxm[k - 1] = 1.0 / (c->xm[kp] - c->xm[k]);
      dxc[k - 1] = 1.0 / (c->xc[kp] - c->xc[k]);
      lem[k - 1] = (c->xc[kp] - c->xc[k]) * c->dx;
      lec[k - 1] = (c->xm[kp] - c->xm[k]) * c->dx;
    }
    lec[nxm - 1] = (c->xm[nxm] - c->xm[nxm - 1]) * c->dx;
    dxm[nxm - 1] = 1.0 / (c->xm[nxm] - c->xm[nxm - 1]);
    add(dxc.data(), nxm);
    add(dxm.data(), nxm);
    add(lem.data(), nxm);
    add(lec.data(), nxm);
    TRYH(hipMalloc(&c->nat, sizeof(double) * nat.size()));
    c->bytes_held += sizeof(double) * nat.size();
    TRYH(hipMemcpy(c->nat, nat.data(), sizeof(double) * nat.size(), hipMemcpyHostToDevice));
    c->d_g3rc = c->nat + off[0];
    c->d_xc = c->nat + off[1];
    c->d_amphk = c->nat + off[2];
    c->d_acphk = c->nat + off[3];
    c->d_apphk = c->nat + off[4];
    c->d_ak1 = c->nat + off[5];
    c->d_ak2 = c->nat + off[6];
    c->d_dxc = c->nat + off[7];
    c->d_dxm = c->nat + off[8];
    c->d_lem = c->nat + off[9];
    c->d_lec = c->nat + off[10];
  }
  // pressure x-solve pivots at the chunk starts: they depend on the grid and the
  // wavenumbers only, not on dt
  TRYH(launch_xsolve_pivots(poisson_args(c, 0), c->piv, c->st));
  TRYH(hipMalloc(&c->partial, sizeof(double) * std::max<int64_t>(reduce_partial_blocks() * 16, 2 * g.nxm + 8)));
  TRYH(hipMalloc(&c->out, sizeof(double) * (2 * g.nxm + 16)));
  c->bytes_held += sizeof(double) * (std::max<int64_t>(reduce_partial_blocks() * 16, 2 * g.nxm + 8) + 2 * g.nxm + 16);
  TRYH(hipHostMalloc(&c->pin_fac, fac_bytes(c)));
  TRYH(hipMalloc(&c->fac_x, fac_bytes(c)));
  for (int ns = 1; ns <= 3; ns++) TRYH(hipMalloc(&c->fac_sub[ns], fac_bytes(c)));
  c->bytes_held += 4 * fac_bytes(c);
  TRYH(hipStreamSynchronize(c->st));
#undef TRY
#undef TRYH
  *out = c;
  return AFID_OK;
}

int afid_get_dims(afid_ctx *c, int64_t *o) {
  if (!c || !o) return AFID_ERR_ARG;
  o[0] = c->g.nx; o[1] = c->g.nxm; o[2] = c->g.nym; o[3] = c->g.nzm; o[4] = c->g.ldx;
  o[5] = (int64_t)c->bytes_held;
  return AFID_OK;
}

int afid_get_metric(afid_ctx *c, const char *name, double *out, int n) {
  if (!c || !name || !out) return AFID_ERR_ARG;
  std::map<std::string, const std::vector<double> *> t = {
      {"xc", &c->xc}, {"xm", &c->xm}, {"g3rc", &c->g3rc}, {"g3rm", &c->g3rm}, {"udx3c", &c->udx3c},
      {"udx3m", &c->udx3m}, {"ap3ck", &c->ap3ck}, {"ac3ck", &c->ac3ck}, {"am3ck", &c->am3ck},
      {"ap3sk", &c->ap3sk}, {"ac3sk", &c->ac3sk}, {"am3sk", &c->am3sk}, {"ap3ssk", &c->ap3ssk},
      {"ac3ssk", &c->ac3ssk}, {"am3ssk", &c->am3ssk}, {"amphk", &c->amphk}, {"acphk", &c->acphk},
      {"apphk", &c->apphk}, {"ak1", &c->ak1}, {"ak2", &c->ak2}};
  auto it = t.find(name);
  if (it != t.end()) {
    const std::vector<double> &v = *it->second;
    if (n > (int)v.size() - 2) return AFID_ERR_ARG;
    for (int k = 0; k < n; k++) out[k] = v[k + 1];
    return AFID_OK;
  }
  std::map<std::string, double> s = {{"dx", c->dx},   {"dy", c->dy},   {"dz", c->dz},     {"ren", c->ren},
                                     {"pec", c->pec}, {"al", c->al},   {"ga", c->ga},     {"ro", c->ro},
                                     {"dt", c->dt},   {"beta", c->beta}, {"time", c->time},
                                     {"ylen", c->p.ylen}, {"zlen", c->p.zlen}, {"alx3", c->p.alx3},
                                     {"str3", c->p.str3}, {"istr3", (double)c->p.istr3},
                                     {"ray", c->p.ray}, {"pra", c->p.pra}};
  auto is = s.find(name);
  if (is != s.end() && n >= 1) {
    out[0] = is->second;
    return AFID_OK;
  }
  return AFID_ERR_ARG;
}

static int ensure_stage(afid_ctx *c, size_t bytes) {
  if (c->stage_bytes >= bytes) return AFID_OK;
  hipFree(c->stage);
  c->bytes_held -= c->stage_bytes;
  c->stage = nullptr;
  c->stage_bytes = 0;
  if (hipMalloc(&c->stage, bytes) != hipSuccess) return AFID_ERR_NOMEM;
  c->bytes_held += bytes;
  c->stage_bytes = bytes;
  return AFID_OK;
}

static int field_rows(afid_ctx *c, int field) { return field == AFID_DPH ? c->g.nxm : c->g.nx; }

static int host_extent(afid_ctx *c, int field, int ld1, int ld2, int j0, int i0, size_t *n) {
  if (j0 < 0 || j0 > 1 || i0 < 0 || i0 > 1) return AFID_ERR_ARG;
  const int n1 = field_rows(c, field);
  const int ny = c->g.nym + 2 * (1 - j0), nz = c->g.nzm + 2 * (1 - i0);
  if (ld1 < n1 || ld2 < ny) return AFID_ERR_ARG;
  *n = (size_t)ld1 * ld2 * nz;
  return AFID_OK;
}

static bool is_history(int field) {
  return field == AFID_RUX || field == AFID_RUY || field == AFID_RUZ || field == AFID_RUTEMP;
}

int afid_upload(afid_ctx *c, int field, const double *host, int ld1, int ld2, int j0, int i0) {
  if (!c || !host || field < 0 || field >= AFID_NFIELDS) return AFID_ERR_ARG;
  size_t n;
  int rc = host_extent(c, field, ld1, ld2, j0, i0, &n);
  if (rc) return rc;
  if (field == AFID_QCAP || field == AFID_DQ || field == AFID_HRO || field == AFID_RHS)
    if ((rc = alloc_field(c, &c->f[field]))) return rc;
  if ((rc = ensure_stage(c, n * sizeof(double)))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(c->stage, host, n * sizeof(double), hipMemcpyHostToDevice));
  if (c->mixed && is_history(field)) {   // FP64 layout into b' scratch, then FP32 into the history
    if ((rc = kerr(c, launch_gather_in(c->stage, c->bp[0], c->g, c->g.nx, ld1, ld2, j0, i0, c->st)))) return rc;
    if ((rc = kerr(c, launch_convert(c->bp[0], c->f[field], c->g.ncol * c->g.ldx, 1, c->st)))) return rc;
  } else if ((rc = kerr(c, launch_gather_in(c->stage, c->f[field], c->g, field_rows(c, field), ld1, ld2, j0, i0,
                                            c->st)))) {
    return rc;
  }
  if (field == AFID_RHS) {  // rhs(1,j,i) is the stale row shared by the x-solves
    // gather row k=0 into the stale plane: reuse scatter of a 1-row view on the host side
    std::vector<double> plane((size_t)c->g.ncol);
    for (int i = 0; i < c->g.nzm; i++)
      for (int j = 0; j < c->g.nym; j++)
        plane[(size_t)i * c->g.nym + j] = host[(size_t)ld1 * ((j + 1 - j0) + (size_t)ld2 * (i + 1 - i0))];
    HIPCHK(c, hipMemcpy(c->stale, plane.data(), sizeof(double) * plane.size(), hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  return AFID_OK;
}

int afid_download(afid_ctx *c, int field, double *host, int ld1, int ld2, int j0, int i0) {
  if (!c || !host || field < 0 || field >= AFID_NFIELDS) return AFID_ERR_ARG;
  if (!c->f[field]) return AFID_ERR_STATE;
  size_t n;
  int rc = host_extent(c, field, ld1, ld2, j0, i0, &n);
  if (rc) return rc;
  if ((rc = ensure_stage(c, n * sizeof(double)))) return rc;
  const int halo = (j0 == 0 && i0 == 0);
  HIPCHK(c, hipMemsetAsync(c->stage, 0, n * sizeof(double), c->st));
  const double *src = c->f[field];
  if (c->mixed && is_history(field)) {   // FP32 history widened into b' scratch first
    if ((rc = kerr(c, launch_convert(c->f[field], c->bp[0], c->g.ncol * c->g.ldx, 0, c->st)))) return rc;
    src = c->bp[0];
  }
  if ((rc = kerr(c, launch_scatter_out(src, c->stage, c->g, field_rows(c, field), ld1, ld2, j0, i0, halo, c->st))))
    return rc;
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(host, c->stage, n * sizeof(double), hipMemcpyDeviceToHost));
  if (field == AFID_RHS) {
    std::vector<double> plane((size_t)c->g.ncol);
    HIPCHK(c, hipMemcpy(plane.data(), c->stale, sizeof(double) * plane.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < c->g.nzm; i++)
      for (int j = 0; j < c->g.nym; j++)
        host[(size_t)ld1 * ((j + 1 - j0) + (size_t)ld2 * (i + 1 - i0))] = plane[(size_t)i * c->g.nym + j];
  }
  return AFID_OK;
}

int afid_init_linear_noise(afid_ctx *c, uint64_t seed, double amp) {
  if (!c) return AFID_ERR_ARG;
  double *z[4] = {c->f[AFID_VX], c->f[AFID_VY], c->f[AFID_VZ], c->f[AFID_PR]};
  int rc = kerr(c, launch_init_noise(c->f[AFID_TEMP], z, c->d_xc, c->g, seed, amp, c->p.tempbp, c->p.temptp, c->st));
  if (rc) return rc;
  for (int q : {AFID_RUX, AFID_RUY, AFID_RUZ, AFID_RUTEMP})
    HIPCHK(c, hipMemsetAsync(c->f[q], 0, field_bytes(c), c->st));
  HIPCHK(c, hipMemsetAsync(c->stale, 0, sizeof(double) * c->g.ncol, c->st));
  c->time = 0.0;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return AFID_OK;
}

int afid_set_coefficients(afid_ctx *c, double al, double ga, double ro, double dt, double beta) {
  if (!c || dt <= 0) return AFID_ERR_ARG;
  c->al = al; c->ga = ga; c->ro = ro; c->dt = dt; c->beta = beta;
  return AFID_OK;
}

int afid_set_substep(afid_ctx *c, int ns, double dt) {
  if (!c || ns < 1 || ns > 3 || dt <= 0) return AFID_ERR_ARG;
  return afid_set_coefficients(c, c->alm[ns], c->gam[ns], c->rom[ns], dt, dt / c->ren * 0.5);
}

static RoutineArgs routine_args(afid_ctx *c) {
  RoutineArgs a;
  a.g = c->g;
  a.f = fields(c);
  a.kc = c->kc;
  a.out = nullptr;
  a.expl = nullptr;
  a.ru = nullptr;
  a.rhs = nullptr;
  a.sc = scalars(c);
  return a;
}

static int explicit_terms(afid_ctx *c, int which, int outfield, const char *name) {
  if (!c) return AFID_ERR_ARG;
  int rc = alloc_field(c, &c->f[outfield]);
  if (rc) return rc;
  RoutineArgs a = routine_args(c);
  a.out = c->f[outfield];
  Timed t(c, name);
  return kerr(c, launch_explicit(a, which, c->st));
}
int afid_explicit_terms_vx(afid_ctx *c) { return explicit_terms(c, 0, AFID_QCAP, "explicit_vx"); }
int afid_explicit_terms_vy(afid_ctx *c) { return explicit_terms(c, 1, AFID_DPH, "explicit_vy"); }
int afid_explicit_terms_vz(afid_ctx *c) { return explicit_terms(c, 2, AFID_DQ, "explicit_vz"); }
int afid_explicit_terms_temp(afid_ctx *c) { return explicit_terms(c, 3, AFID_HRO, "explicit_temp"); }

static int implicit_update(afid_ctx *c, int which) {
  if (!c) return AFID_ERR_ARG;
  static const int expl_of[4] = {AFID_QCAP, AFID_DPH, AFID_DQ, AFID_HRO};
  static const int ru_of[4] = {AFID_RUX, AFID_RUY, AFID_RUZ, AFID_RUTEMP};
  static const int q_of[4] = {AFID_VX, AFID_VY, AFID_VZ, AFID_TEMP};
  static const char *names[4] = {"implicit_vx", "implicit_vy", "implicit_vz", "implicit_temp"};
  if (c->mixed) return AFID_ERR_STATE;   // these read and write the histories in FP64
  int rc;
  if (!c->f[expl_of[which]]) return AFID_ERR_STATE;
  if ((rc = alloc_field(c, &c->f[AFID_RHS]))) return rc;
  if (c->fac_x_key[0] != c->al || c->fac_x_key[1] != c->dt) {
    if ((rc = upload_factors(c, c->al, c->dt, c->fac_x))) return rc;
    c->fac_x_key[0] = c->al;
    c->fac_x_key[1] = c->dt;
  }
  RoutineArgs a = routine_args(c);
  a.expl = c->f[expl_of[which]];
  a.ru = c->f[ru_of[which]];
  a.rhs = c->f[AFID_RHS];
  Timed t(c, names[which]);
  if ((rc = kerr(c, launch_rhs(a, which, c->st)))) return rc;
  SolveArgs s;
  s.g = c->g;
  const int sys = which == 0 ? SYS_X : (which == 3 ? SYS_T : SYS_YZ);
  s.fac = c->fac_x + (size_t)sys * FR_N * c->g.ldx;
  s.rhs = c->f[AFID_RHS];
  s.q = c->f[q_of[which]];
  s.stale = c->stale;
  s.tempbp = c->p.tempbp;
  s.temptp = c->p.temptp;
  return kerr(c, launch_solve_update(s, which, c->st));
}
int afid_implicit_and_update_vx(afid_ctx *c) { return implicit_update(c, 0); }
int afid_implicit_and_update_vy(afid_ctx *c) { return implicit_update(c, 1); }
int afid_implicit_and_update_vz(afid_ctx *c) { return implicit_update(c, 2); }
int afid_implicit_and_update_temp(afid_ctx *c) { return implicit_update(c, 3); }

int afid_calc_local_divergence(afid_ctx *c) {
  if (!c) return AFID_ERR_ARG;
  RoutineArgs a = routine_args(c);
  a.out = c->f[AFID_DPH];
  Timed t(c, "local_divergence");
  return kerr(c, launch_divergence(a, c->st));
}

int afid_solve_pressure_correction(afid_ctx *c) {
  if (!c) return AFID_ERR_ARG;
  return run_poisson(c, 0);
}

int afid_pressure_spectrum(afid_ctx *c, int stage, double *out) {
  if (!c || !out || stage < 1 || stage > 4) return AFID_ERR_ARG;
  if (c->mixed) return AFID_ERR_STATE;   // the spectrum is FP32 in that mode
  PoissonArgs a = poisson_args(c, 0);
  int rc;
  if ((rc = kerr(c, launch_yfft_forward(a, c->st)))) return rc;
  if (stage >= 2 && (rc = kerr(c, launch_zfft(a, -1, c->st)))) return rc;
  if (stage >= 3) {
    if (c->singular_mode) return AFID_NUM_SINGULAR;
    if ((rc = kerr(c, launch_xsolve_pressure(a, c->st)))) return rc;
  }
  if (stage >= 4 && (rc = kerr(c, launch_zfft(a, +1, c->st)))) return rc;
  const size_t n = (size_t)c->g.nxm * c->g.nzm * c->ldm;
  std::vector<double2> h(n);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->S, sizeof(double2) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const int nxm = c->g.nxm, nymh = c->nymh, nzm = c->g.nzm;
  for (int k = 0; k < nxm; k++)
    for (int nn = 0; nn < nzm; nn++)
      for (int m = 0; m < nymh; m++) {
        const double2 v = h[((size_t)k * nzm + nn) * c->ldm + m];
        const size_t o = (size_t)k + (size_t)nxm * ((size_t)m + (size_t)nymh * nn);
        out[2 * o] = v.x;
        out[2 * o + 1] = v.y;
      }
  return AFID_OK;
}

// from_ustar: the velocities before the correction are u* in the b buffers
// (substep path); otherwise the fields are corrected in place.
static int correct(afid_ctx *c, int which, const char *name, int ns = 0, bool dev = false, bool from_ustar = false,
                   bool cfl_fold = false) {
  if (!c) return AFID_ERR_ARG;
  CorrectArgs a;
  a.ds = dev ? c->dstep : nullptr;
  a.ns = ns;
  a.g = c->g;
  a.dph = c->f[AFID_DPH];
  a.vx = c->f[AFID_VX]; a.vy = c->f[AFID_VY]; a.vz = c->f[AFID_VZ]; a.pr = c->f[AFID_PR];
  a.ux = from_ustar ? c->bp[0] : a.vx;
  a.uy = from_ustar ? c->bp[1] : a.vy;
  a.uz = from_ustar ? c->bp[2] : a.vz;
  a.kc = c->kc;
  a.sc = scalars(c);
  a.cflbits = cfl_fold ? &c->dstep->cflbits : nullptr;
  Timed t(c, name);
  return kerr(c, launch_correct(a, which, c->st));
}
int afid_correct_velocity(afid_ctx *c) { return correct(c, 1, "correct_velocity"); }
int afid_correct_pressure(afid_ctx *c) { return correct(c, 2, "correct_pressure"); }
int afid_update_halo(afid_ctx *c, int field) {
  if (!c || field < 0 || field >= AFID_NFIELDS) return AFID_ERR_ARG;
  return AFID_OK;
}

// One substep's kernels on the context's stream.  dev = false: coefficients
// and factor rows from the host (afid_set_substep, upload_factors); dev = true:
// both come from the device step state (afid_steps), nothing is copied.
static FusedArgs update_args(afid_ctx *c, int ns, bool dev) {
  FusedArgs a;
  a.g = c->g;
  a.f = fields(c);
  a.tn = c->f[AFID_TEMP];   // the column solve reads only b' and vz(k=1)
  a.rux = c->f[AFID_RUX]; a.ruy = c->f[AFID_RUY]; a.ruz = c->f[AFID_RUZ]; a.rut = c->f[AFID_RUTEMP];
  a.stale = c->stale;
  a.kc = c->kc;
  a.fac = dev ? c->fac_dev[ns - 1] : c->fac_sub[ns];
  a.bx = c->bp[0]; a.by = c->bp[1]; a.bz = c->bp[2]; a.bt = c->bp[3];
  a.ux = c->f[AFID_VX]; a.uy = c->f[AFID_VY]; a.uz = c->f[AFID_VZ];   // in place unless set apart
  a.sc = scalars(c);
  a.sc.use_ru = c->rom[ns] != 0.0;
  a.ds = dev ? c->dstep : nullptr;
  a.ns = ns;
  a.h32 = c->mixed;
  a.zm = c->zm;
  // The histories written here are read by the next substep only when its
  // ro is nonzero: never after RK3 substep 3 (ro(1) = 0, TimeMarcher.F90:22-30).
  a.keep_ru = c->rom[ns % c->p.nsst + 1] != 0.0;
  return a;
}

// cfl_fold: the correction also reduces CalcMaxCFL into the device step
// state for the next step (the last substep of a device step).  u* then has
// to stay apart from the corrected fields, in the b buffers.  Elsewhere the
// column solve writes u* into the fields and the correction works in place,
// which moves the same words through fewer distinct arrays and runs faster.
static int enqueue_substep(afid_ctx *c, int ns, bool dev, bool cfl_fold = false) {
  int rc;
  FusedArgs a = update_args(c, ns, dev);
  const bool apart = cfl_fold;
  if (apart) {
    a.ux = c->bp[0]; a.uy = c->bp[1]; a.uz = c->bp[2];
  }
  {
    Timed t(c, "stencil_rhs");
    if ((rc = kerr(c, launch_stencil_rhs(a, c->st)))) return rc;
  }
  {
    Timed t(c, "column_solve");
    if ((rc = kerr(c, launch_column_solve(a, c->st)))) return rc;
  }
  if ((rc = run_poisson(c, 1, ns, dev, apart))) return rc;
  return correct(c, 3, "correct_vel_pr", ns, dev, apart, cfl_fold);
}

int afid_substep(afid_ctx *c, int ns, double dt) {
  if (!c || ns < 1 || ns > c->p.nsst || ns > 3 || dt <= 0) return AFID_ERR_ARG;
  int rc = afid_set_substep(c, ns, dt);
  if (rc) return rc;
  if (c->fac_sub_dt[ns] != dt) {
    if ((rc = upload_factors(c, c->al, dt, c->fac_sub[ns]))) return rc;
    c->fac_sub_dt[ns] = dt;
  }
  return enqueue_substep(c, ns, false);
}

// Mixed precision: the divergence the FP32 pressure solve of the last
// substep left behind, computed in FP64, is solved for and corrected once more
// (velocity and pressure, with the last substep's scalars).
static int refine_projection(afid_ctx *c, int ns, bool dev) {
  int rc = run_poisson(c, 1, ns, dev);
  return rc ? rc : correct(c, 3, "refine_correct", ns, dev);
}

int afid_time_marcher(afid_ctx *c, double dt) {
  if (!c || dt <= 0) return AFID_ERR_ARG;
  // every substep's factorisation is checked before the first one runs, so a
  // refused step (AFID_NUM_PIVOT) leaves the fields as they were, as the
  // device step path does
  {
    std::vector<double> rows(fac_bytes(c) / sizeof(double));
    for (int q = 1; q <= c->p.nsst && q <= 3; q++) {
      int rc = build_factors(c, c->alm[q], dt, rows.data());
      if (rc) return rc;
    }
  }
  int ns = 1;
  for (; ns <= c->p.nsst && ns <= 3; ns++) {
    int rc = afid_substep(c, ns, dt);
    if (rc) return rc;
  }
  return c->mixed ? refine_projection(c, ns - 1, false) : AFID_OK;
}

int afid_calc_max_cfl(afid_ctx *c, double *cflm) {
  if (!c || !cflm) return AFID_ERR_ARG;
  int rc;
  {
    Timed t(c, "reduce_cfl");
    if ((rc = kerr(c, launch_reduce_cfl(reduce_args(c), c->st)))) return rc;
  }
  double v;
  if ((rc = read_out(c, &v, 1))) return rc;
  *cflm = std::max(0.00000001, v);
  return AFID_OK;
}

int afid_step(afid_ctx *c, int ntime, int idtv, double dtin, double dtmin, double dtmax, double limitcfl,
              double *dt_used, int *errorcode) {
  if (!c) return AFID_ERR_ARG;
  double instcfl;
  int rc = afid_calc_max_cfl(c, &instcfl);
  if (rc) return rc;
  int err = 0;
  double dt = dtin;
  if (idtv) {
    if (ntime > 1) {
      dt = instcfl < 1.0e-8 ? dtmax : limitcfl / instcfl;
      if (dt > dtmax) dt = dtmax;
    } else {
      dt = dtmin;
    }
    if (dt < dtmin) err = 166;
  } else {
    if (instcfl * dt > limitcfl) err = 165;
  }
  if ((rc = afid_time_marcher(c, dt))) return rc;
  c->time = c->time + dt;
  if (dt_used) *dt_used = dt;
  if (errorcode) *errorcode = err;
  return AFID_OK;
}

// ---- whole steps on the device ------------------------------------------------
static int ensure_dev_step(afid_ctx *c) {
  if (c->dstep) return AFID_OK;
  const Geom &g = c->g;
  HIPCHK(c, hipMalloc(&c->dstep, sizeof(DevStep)));
  HIPCHK(c, hipMalloc(&c->fwork, sizeof(double) * SYS_N * 6 * (size_t)g.nx));
  for (int q = 0; q < 3; q++) HIPCHK(c, hipMalloc(&c->fac_dev[q], sizeof(double) * SYS_N * FR_N * (size_t)g.ldx));
  // natural-order rows (index kc-1) of the metric arrays the factorisation uses
  const std::vector<double> *rows[7] = {&c->ac3ck, &c->am3ssk, &c->ap3ssk, &c->ac3sk,
                                        &c->am3sk, &c->ap3sk, &c->ac3ssk};
  std::vector<double> mc((size_t)7 * g.nx);
  for (int q = 0; q < 7; q++)
    for (int k = 0; k < g.nx; k++) mc[(size_t)q * g.nx + k] = (*rows[q])[k + 1];
  HIPCHK(c, hipMalloc(&c->mc, sizeof(double) * mc.size()));
  HIPCHK(c, hipMemcpy(c->mc, mc.data(), sizeof(double) * mc.size(), hipMemcpyHostToDevice));
  c->bytes_held += sizeof(DevStep) + sizeof(double) * (SYS_N * 6 * (size_t)g.nx + 3 * SYS_N * FR_N * (size_t)g.ldx +
                                                       mc.size());
  StepConsts &k = c->kconst;
  k.ren = c->ren; k.pec = c->pec; k.dy = c->dy; k.dz = c->dz; k.dyq = c->dyq; k.dzq = c->dzq;
  k.tempbp = c->p.tempbp; k.temptp = c->p.temptp;
  k.nsst = std::min(c->p.nsst, 3);
  for (int ns = 1; ns <= 3; ns++) {
    k.gam[ns - 1] = c->gam[ns];
    k.rom[ns - 1] = c->rom[ns];
    k.alm[ns - 1] = c->alm[ns];
  }
  return AFID_OK;
}

// One main.F90 step entirely on the stream: CalcMaxCFL, dt and substep
// scalars, factorisations, the substeps, time += dt.  CalcMaxCFL is a
// reduction over the fields at the start of the step; with cfl_fold it comes
// instead from the last correction of the previous step (dev_steps_prepare
// seeds it for the first), which the mixed mode's in-place refinement sweep
// rules out.
static int enqueue_step_device(afid_ctx *c) {
  int rc;
  if (!c->cfl_fold) {
    {
      Timed t(c, "reduce_cfl");
      if ((rc = kerr(c, launch_reduce_cfl(reduce_args(c), c->st)))) return rc;
    }
    if ((rc = kerr(c, launch_step_begin(c->dstep, c->out, c->kconst, c->st)))) return rc;
  } else if ((rc = kerr(c, launch_step_begin(c->dstep, nullptr, c->kconst, c->st)))) {
    return rc;
  }
  FactorArgs fa;
  fa.g = c->g;
  fa.mc = c->mc;
  for (int q = 0; q < 3; q++) fa.fac[q] = c->fac_dev[q];
  fa.work = c->fwork;
  fa.ds = c->dstep;
  fa.k = c->kconst;
  if ((rc = kerr(c, launch_factor(fa, c->st)))) return rc;
  for (int ns = 1; ns <= c->kconst.nsst; ns++)
    if ((rc = enqueue_substep(c, ns, true, c->cfl_fold && ns == c->kconst.nsst))) return rc;
  if (c->mixed && (rc = refine_projection(c, c->kconst.nsst, true))) return rc;
  return kerr(c, launch_step_end(c->dstep, c->st));
}

// One step captured into a graph; its kernels take no per-step host data.
// timed = true brackets each kernel by event-record nodes (profiling).
static int capture_step(afid_ctx *c, hipGraphExec_t *out, bool timed) {
  hipGraph_t graph = nullptr;
  HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  c->prof_capture = timed;
  int rc = enqueue_step_device(c);
  c->capturing = false;
  c->prof_capture = false;
  hipError_t e = hipStreamEndCapture(c->st, &graph);
  if (rc) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  HIPCHK(c, e);
  e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  HIPCHK(c, e);
  return AFID_OK;
}

// Device step state from the host values, and the step graph, captured once.
static int dev_steps_prepare(afid_ctx *c, int ntime0, int idtv, double dt, double dtmin, double dtmax,
                             double limitcfl) {
  int rc = ensure_dev_step(c);
  if (rc) return rc;
  DevStep h{};
  h.time = c->time; h.ntime = ntime0; h.idtv = idtv; h.dtfix = dt; h.dtmin = dtmin; h.dtmax = dtmax;
  h.limitcfl = limitcfl; h.dt = dt;
  if (c->cfl_fold) {   // CalcMaxCFL of the current fields for the first step
    if ((rc = kerr(c, launch_reduce_cfl(reduce_args(c), c->st)))) return rc;
    double v;
    if ((rc = read_out(c, &v, 1))) return rc;
    std::memcpy(&h.cflbits, &v, sizeof v);
  }
  HIPCHK(c, hipMemcpyAsync(c->dstep, &h, sizeof(DevStep), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (!c->step_graph && (rc = capture_step(c, &c->step_graph, false))) return rc;
  if (c->prof && !c->prof_graph && (rc = capture_step(c, &c->prof_graph, true))) {
    drop_prof_graph(c);
    return rc;
  }
  c->dev_steps_ready = true;
  return AFID_OK;
}

// Profiling replays the timed copy of the step graph one step at a time and
// adds each kernel's event-pair time to the profile, as host-driven steps do.
static int dev_steps_enqueue(afid_ctx *c, int nsteps) {
  if (c->prof && c->prof_graph) {
    for (int n = 0; n < nsteps; n++) {
      HIPCHK(c, hipGraphLaunch(c->prof_graph, c->st));
      HIPCHK(c, hipStreamSynchronize(c->st));
      for (auto &pt : c->graph_timing) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, pt.a, pt.b));
        auto &e = c->prof_tot[pt.name];
        e.first += ms;
        e.second += 1;
      }
    }
    return AFID_OK;
  }
  for (int n = 0; n < nsteps; n++) HIPCHK(c, hipGraphLaunch(c->step_graph, c->st));
  return AFID_OK;
}

static int dev_steps_finish(afid_ctx *c, double *dt_last, int *errorcode, int *steps_done) {
  DevStep h;
  HIPCHK(c, hipMemcpyAsync(&h, c->dstep, sizeof(DevStep), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->time = h.time;
  if (dt_last) *dt_last = h.dt;
  if (errorcode) *errorcode = h.errorcode;
  if (steps_done) *steps_done = h.steps_done;
  return h.pivot ? AFID_NUM_PIVOT : AFID_OK;
}

int afid_steps(afid_ctx *c, int ntime0, int nsteps, int idtv, double dt, double dtmin, double dtmax,
               double limitcfl, double *dt_last, int *errorcode, int *steps_done) {
  if (!c || nsteps < 0 || (!idtv && dt <= 0)) return AFID_ERR_ARG;
  if (c->singular_mode) return AFID_NUM_SINGULAR;
  int rc = dev_steps_prepare(c, ntime0, idtv, dt, dtmin, dtmax, limitcfl);
  if (!rc) rc = dev_steps_enqueue(c, nsteps);
  if (rc) return rc;
  return dev_steps_finish(c, dt_last, errorcode, steps_done);
}

int afid_steps_cfl(afid_ctx *c, double *cflm) {
  if (!c || !cflm) return AFID_ERR_ARG;
  if (!c->dev_steps_ready || !c->cfl_fold) return AFID_ERR_STATE;
  DevStep h;
  HIPCHK(c, hipMemcpyAsync(&h, c->dstep, sizeof(DevStep), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  double v;
  std::memcpy(&v, &h.cflbits, sizeof v);
  *cflm = std::max(0.00000001, v);
  return AFID_OK;
}

int afid_steps_async(afid_ctx *c, int ntime0, int nsteps, int idtv, double dt, double dtmin, double dtmax,
                     double limitcfl) {
  if (!c || nsteps < 0 || (!idtv && dt <= 0)) return AFID_ERR_ARG;
  if (c->singular_mode) return AFID_NUM_SINGULAR;
  int rc = dev_steps_prepare(c, ntime0, idtv, dt, dtmin, dtmax, limitcfl);
  return rc ? rc : dev_steps_enqueue(c, nsteps);
}

int afid_steps_more(afid_ctx *c, int nsteps) {
  if (!c || nsteps < 0) return AFID_ERR_ARG;
  if (!c->dev_steps_ready) return AFID_ERR_STATE;
  return dev_steps_enqueue(c, nsteps);
}

int afid_steps_query(afid_ctx *c, int *busy) {
  if (!c || !busy) return AFID_ERR_ARG;
  hipError_t e = hipStreamQuery(c->st);
  if (e == hipErrorNotReady) {
    *busy = 1;
    return AFID_OK;
  }
  *busy = 0;
  HIPCHK(c, e);
  return AFID_OK;
}

int afid_steps_wait(afid_ctx *c, double *dt_last, int *errorcode, int *steps_done) {
  if (!c) return AFID_ERR_ARG;
  if (!c->dev_steps_ready) return AFID_ERR_STATE;
  return dev_steps_finish(c, dt_last, errorcode, steps_done);
}

int afid_batch_steps(afid_ctx *const *cells, int ncells, int nsteps) {
  if (!cells || ncells < 0 || nsteps < 0) return AFID_ERR_ARG;
  for (int q = 0; q < ncells; q++) {
    if (!cells[q]) return AFID_ERR_ARG;
    if (!cells[q]->dev_steps_ready) return AFID_ERR_STATE;
  }
  // step n of every cell is queued before step n+1 of any, each cell on its own stream
  for (int n = 0; n < nsteps; n++)
    for (int q = 0; q < ncells; q++) {
      int rc = dev_steps_enqueue(cells[q], 1);
      if (rc) return rc;
    }
  return AFID_OK;
}

// Autotuning of the stencil pass's z planes per block: each candidate is
// timed with HIP events on this context's stream (one warm-up and three timed
// launches of the substep-2 form, which reads the histories); the fastest is
// kept.  The launches write the histories and the b' scratch, which are
// cleared afterwards, so the state is to be set after this call.
int afid_tune(afid_ctx *c, int *zm_out) {
  if (!c) return AFID_ERR_ARG;
  static const int cands[] = {8, 16, 32, 64};
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  FusedArgs a = update_args(c, 2, false);
  a.sc.use_ru = 1;
  int best = 0, rc = AFID_OK;
  float best_ms = 0.f;
  for (int zm : cands) {
    if (zm > c->g.nzm && zm != cands[0]) break;   // a larger block than the grid adds nothing
    a.zm = zm;
    auto launch = [&]() { return launch_stencil_rhs(a, c->st); };
    if ((rc = kerr(c, launch()))) break;
    if ((rc = kerr(c, hipEventRecord(e0, c->st)))) break;
    for (int r = 0; r < 3 && !rc; r++) rc = kerr(c, launch());
    if (rc || (rc = kerr(c, hipEventRecord(e1, c->st))) || (rc = kerr(c, hipEventSynchronize(e1)))) break;
    float ms = 0.f;
    if ((rc = kerr(c, hipEventElapsedTime(&ms, e0, e1)))) break;
    if (best == 0 || ms < best_ms) {
      best = zm;
      best_ms = ms;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  for (int q : {AFID_RUX, AFID_RUY, AFID_RUZ, AFID_RUTEMP})
    HIPCHK(c, hipMemsetAsync(c->f[q], 0, field_bytes(c), c->st));
  for (int q = 0; q < 4; q++) HIPCHK(c, hipMemsetAsync(c->bp[q], 0, field_bytes(c), c->st));
  HIPCHK(c, hipMemsetAsync(c->stale, 0, sizeof(double) * c->g.ncol, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->zm = best;
  if (c->step_graph) {   // a captured step holds the previous launch shape
    hipGraphExecDestroy(c->step_graph);
    c->step_graph = nullptr;
    c->dev_steps_ready = false;
  }
  drop_prof_graph(c);
  if (zm_out) *zm_out = best;
  return AFID_OK;
}

int afid_estimate_bytes(const afid_params *p, int64_t *bytes) {
  if (!p || !bytes || p->nxm < 2 || p->nym < 2 || p->nzm < 2) return AFID_ERR_ARG;
  const int64_t nx = p->nxm + 1, nxm = p->nxm, nym = p->nym, nzm = p->nzm;
  const int64_t L = (nx + P - 1) / P, ldx = P * L, ncol = nym * nzm;
  const int64_t ldm = ((nym / 2 + 1) + 3) / 4 * 4;
  const int64_t field = 8 * ncol * ldx, fac = 8 * (int64_t)SYS_N * FR_N * ldx;
  int64_t b = 14 * field;                                   // 10 fields + 4 b' (or 4 ping-pong) arrays
  b += 8 * ncol;                                            // stale row
  int np, lm;
  xsolve_shape((int)nxm, &np, &lm);
  b += nxm * nzm * ldm * 16 + 8 * (nym / 2 + 1) * nzm * np;  // spectrum, x-solve pivot table
  b += 8 * (int64_t)KC_N * ldx + 8 * (2 * nx + 7 * nxm + nym + nzm);   // coefficient rows
  b += 4 * fac + 3 * fac + 8 * (int64_t)SYS_N * 6 * nx + 8 * 7 * nx + (int64_t)sizeof(DevStep);
  b += 20 * (nym + nzm) + 16 * nym;                         // FFT tables
  b += 8 * (reduce_partial_blocks() * 16 + 4 * nxm + 32);   // reduction scratch
  b += 8 * nx * (nym + 2) * (nzm + 2);                      // staging of uploads and downloads
  Geom gs{};
  gs.ldx = (int)ldx;
  b += 8 * (STATS_N * nxm + (int64_t)stats_partial_doubles(gs));  // statistics
  b += (4 + 8 * 4 * ncol) * 9;                              // nine slab planes of four fields
  *bytes = b;
  return AFID_OK;
}

int afid_get_time(afid_ctx *c, double *t) {
  if (!c || !t) return AFID_ERR_ARG;
  *t = c->time;
  return AFID_OK;
}

int afid_set_time(afid_ctx *c, double t) {
  if (!c) return AFID_ERR_ARG;
  c->time = t;
  return AFID_OK;
}

int afid_get_stale(afid_ctx *c, double *plane) {
  if (!c || !plane) return AFID_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(plane, c->stale, sizeof(double) * c->g.ncol, hipMemcpyDeviceToHost));
  return AFID_OK;
}

int afid_set_stale(afid_ctx *c, const double *plane) {
  if (!c || !plane) return AFID_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(c->stale, plane, sizeof(double) * c->g.ncol, hipMemcpyHostToDevice));
  return AFID_OK;
}

int afid_check_divergence(afid_ctx *c, double *qmax) {
  if (!c || !qmax) return AFID_ERR_ARG;
  int rc = kerr(c, launch_reduce_divmax(reduce_args(c), c->st));
  if (rc) return rc;
  return read_out(c, qmax, 1);
}

int afid_calc_plate_nu(afid_ctx *c, double *nb, double *nt) {
  if (!c || !nb || !nt) return AFID_ERR_ARG;
  double del = 1.0 / (c->xc[2] - c->xc[1]);
  double deln = 1.0 / (c->xc[c->g.nx] - c->xc[c->g.nxm]);
  int rc = kerr(c, launch_reduce_platenu(reduce_args(c), del, deln, c->st));
  if (rc) return rc;
  double v[2];
  if ((rc = read_out(c, v, 2))) return rc;
  *nb = v[0] / (c->g.nzm * c->g.nym);
  *nt = v[1] / (c->g.nzm * c->g.nym);
  return AFID_OK;
}

int afid_global_quantities(afid_ctx *c, double *o) {
  if (!c || !o) return AFID_ERR_ARG;
  int rc = kerr(c, launch_reduce_global(reduce_args(c), c->st));
  if (rc) return rc;
  double v[11];
  if ((rc = read_out(c, v, 11))) return rc;
  // GlobalQuantities.F90:77-96 finalisation
  double vol = 1.0 / (c->p.alx3 * c->dx * (double)c->g.nzm * (double)c->g.nym);
  double rradpr = std::sqrt(c->p.ray / c->p.pra);
  o[0] = 1.0 + std::sqrt(c->p.pra * c->p.ray) * v[0] * vol;
  o[1] = v[1] * vol;
  o[2] = v[9];
  o[3] = v[10];
  o[4] = v[6];
  o[5] = v[7];
  o[6] = v[8];
  o[7] = std::sqrt(v[4] * vol) * rradpr;
  o[8] = std::sqrt(v[3] * vol) * rradpr;
  o[9] = std::sqrt(v[2] * vol) * rradpr;
  o[10] = std::sqrt(v[5] * vol) * rradpr;
  return AFID_OK;
}

int afid_calc_dissipation_nu(afid_ctx *c, double *nu_mu, double *nu_th, double *disste, double *dissth) {
  if (!c || !nu_mu || !nu_th) return AFID_ERR_ARG;
  int rc = kerr(c, launch_reduce_dissipation(reduce_args(c), c->st));
  if (rc) return rc;
  std::vector<double> v(2 + 2 * c->g.nxm);
  if ((rc = read_out(c, v.data(), (int)v.size()))) return rc;
  double volt = 1.0 / ((double)c->g.nxm * (double)c->g.nzm * (double)c->g.nym);
  *nu_mu = v[0] * volt + 1;
  *nu_th = v[1] * volt;
  if (disste) std::copy(v.begin() + 2, v.begin() + 2 + c->g.nxm, disste);
  if (dissth) std::copy(v.begin() + 2 + c->g.nxm, v.end(), dissth);
  // the profiles also accumulate into the statistics (CalcDissipationNu.F90:101-104)
  c->disste_acc.resize(c->g.nxm, 0.0);
  c->dissth_acc.resize(c->g.nxm, 0.0);
  for (int k = 0; k < c->g.nxm; k++) {
    c->disste_acc[k] += v[2 + k];
    c->dissth_acc[k] += v[2 + c->g.nxm + k];
  }
  return AFID_OK;
}

// ---- statistics and slabs -------------------------------------------------------
int afid_calc_stats(afid_ctx *c) {
  if (!c) return AFID_ERR_ARG;
  if (!c->stat_acc) {
    HIPCHK(c, hipMalloc(&c->stat_acc, sizeof(double) * STATS_N * c->g.nxm));
    HIPCHK(c, hipMemsetAsync(c->stat_acc, 0, sizeof(double) * STATS_N * c->g.nxm, c->st));
    HIPCHK(c, hipMalloc(&c->stat_partial, sizeof(double) * stats_partial_doubles(c->g)));
    c->bytes_held += sizeof(double) * (STATS_N * c->g.nxm + stats_partial_doubles(c->g));
  }
  ReduceArgs a = reduce_args(c);
  a.partial = c->stat_partial;
  int rc = kerr(c, launch_stats(a, c->stat_acc, c->st));
  if (rc) return rc;
  c->nstatsamples += 1;
  return AFID_OK;
}

// ---- statistics without allocation ----------------------------------------------
// The block partials of afid_calc_stats_noalloc live in one static device array
// per GPU.  Host code never uses the array's name as a device address: its
// address on each device comes from hipGetSymbolAddress, taken once per device
// (with that device current) by the first afid_create on it, and only that
// pointer is passed to kernels.
static __device__ __align__(256) double scratch[AFID_NOALLOC_SCRATCH_BYTES / sizeof(double)];
constexpr int SCRATCH_MAX_DEV = 64;
static double *scratch_dev[SCRATCH_MAX_DEV];
static std::mutex scratch_mu;

static int scratch_setup(afid_ctx *c) {
  if (c->dev < 0 || c->dev >= SCRATCH_MAX_DEV) return AFID_ERR_ARG;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_dev[c->dev]) return AFID_OK;
  void *ptr = nullptr;
  HIPCHK(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  scratch_dev[c->dev] = static_cast<double *>(ptr);
  return AFID_OK;
}

int afid_calc_stats_noalloc(afid_ctx *c, double *acc) {
  if (!c || !acc) return AFID_ERR_ARG;
  if (stats_partial_doubles(c->g) * sizeof(double) > AFID_NOALLOC_SCRATCH_BYTES) return AFID_ERR_ARG;
  double *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (c->dev >= 0 && c->dev < SCRATCH_MAX_DEV) ptr = scratch_dev[c->dev];
  }
  if (!ptr) return AFID_ERR_STATE;   // afid_create on this device did not set it up
  ReduceArgs a = reduce_args(c);
  a.partial = ptr;   // fully written by the partial pass before the finish pass reads it
  return kerr(c, launch_stats(a, acc, c->st));
}

int afid_get_stats(afid_ctx *c, double *out, int *nsamples) {
  if (!c || !out) return AFID_ERR_ARG;
  const int nxm = c->g.nxm;
  std::fill(out, out + (STATS_N + 2) * nxm, 0.0);
  if (c->stat_acc) {
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(out, c->stat_acc, sizeof(double) * STATS_N * nxm, hipMemcpyDeviceToHost));
  }
  for (int k = 0; k < nxm && k < (int)c->disste_acc.size(); k++) {
    out[STATS_N * nxm + k] = c->disste_acc[k];
    out[(STATS_N + 1) * nxm + k] = c->dissth_acc[k];
  }
  if (nsamples) *nsamples = c->nstatsamples;
  return AFID_OK;
}

int afid_reset_stats(afid_ctx *c) {
  if (!c) return AFID_ERR_ARG;
  if (c->stat_acc) HIPCHK(c, hipMemsetAsync(c->stat_acc, 0, sizeof(double) * STATS_N * c->g.nxm, c->st));
  std::fill(c->disste_acc.begin(), c->disste_acc.end(), 0.0);
  std::fill(c->dissth_acc.begin(), c->dissth_acc.end(), 0.0);
  c->nstatsamples = 0;
  return AFID_OK;
}

int afid_get_slabs(afid_ctx *c, const int *kslab, int nslab, double *out) {
  if (!c || !kslab || !out || nslab < 1 || nslab > 64) return AFID_ERR_ARG;
  for (int m = 0; m < nslab; m++)
    if (kslab[m] < 1 || kslab[m] > c->g.nxm) return AFID_ERR_ARG;
  if (nslab > c->slab_n) {
    hipFree(c->d_kslab);
    hipFree(c->slab_buf);
    c->bytes_held -= (sizeof(int) + sizeof(double) * 4 * (size_t)c->g.ncol) * c->slab_n;
    c->d_kslab = nullptr;
    c->slab_buf = nullptr;
    c->slab_n = 0;
    HIPCHK(c, hipMalloc(&c->d_kslab, sizeof(int) * nslab));
    HIPCHK(c, hipMalloc(&c->slab_buf, sizeof(double) * 4 * (size_t)nslab * c->g.ncol));
    c->slab_n = nslab;
    c->bytes_held += (sizeof(int) + sizeof(double) * 4 * (size_t)c->g.ncol) * nslab;
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(c->d_kslab, kslab, sizeof(int) * nslab, hipMemcpyHostToDevice));
  int rc = kerr(c, launch_slabs(fields(c), c->g, c->d_kslab, nslab, c->slab_buf, c->st));
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(out, c->slab_buf, sizeof(double) * 4 * (size_t)nslab * c->g.ncol, hipMemcpyDeviceToHost));
  return AFID_OK;
}

// ---- restart regridding ---------------------------------------------------
// Grid of interp.F90's gridnew (:254-299): n faces rc and centres rm of
// extent rext, uniform (istr 0) or clipped Chebyshev (istr 6, int(str)
// nodes dropped at each end); rm(n) is extrapolated.  Other stretchings
// are rejected: gridnew leaves rc undefined for them.
static bool regrid_axis(int n, double rext, int istr, double str, std::vector<double> &rc,
                        std::vector<double> &rm) {
  rc.assign(n + 2, 0.0);
  rm.assign(n + 2, 0.0);
  if (n < 2) return false;
  if (istr == 0) {
    for (int i = 1; i <= n; i++) rc[i] = rext * ((double)(i - 1) / (double)(n - 1));
  } else if (istr == 6) {
    const int nclip = (int)str;
    if (nclip < 0 || nclip > 200) return false;
    const double pi = 2.0 * std::asin(1.0);
    const int m = n + 2 * nclip;
    std::vector<double> eta(n + 1);
    for (int i = 1; i <= n; i++) eta[i] = std::cos(pi * ((double)(i + nclip) - 0.5) / (double)m);
    const double half = 0.5 * (eta[1] - eta[n]);
    for (int i = 1; i <= n; i++) eta[i] = eta[i] / half;
    rc[1] = 0.0;
    for (int i = 2; i < n; i++) rc[i] = rext * (1. - eta[i]) * 0.5;
    rc[n] = rext;
  } else {
    return false;
  }
  for (int i = 1; i < n; i++) rm[i] = (rc[i] + rc[i + 1]) * 0.5;
  rm[n] = 2 * rc[n] - rm[n - 1];
  return true;
}

// Old axes get the linear end extrapolation of interp.F90:47-62.
static void extend_axis(std::vector<double> &v, int n) {
  v[0] = 2 * v[1] - v[2];
  v[n + 1] = 2 * v[n] - v[n - 1];
}

int afid_interp_field(afid_ctx *c, int field, const double *old, int nxo, int nyo, int nzo, int istro3,
                      double stro3) {
  if (!c || !old) return AFID_ERR_ARG;
  int intvar;
  switch (field) {
    case AFID_VX: intvar = 1; break;
    case AFID_VY: intvar = 2; break;
    case AFID_VZ: intvar = 3; break;
    case AFID_TEMP: intvar = 4; break;
    default: return AFID_ERR_ARG;
  }
  const Geom &g = c->g;
  const int nx = g.nx, nym = g.nym, nzm = g.nzm, ny = nym + 1, nz = nzm + 1;
  if (nxo < 2 || nyo < 3 || nzo < 3 || nxo > 65536 || nyo > 65536 || nzo > 65536) return AFID_ERR_ARG;
  if (nz > nzo * 2 || ny > nyo * 2 || nx > nxo * 2) return AFID_ERR_GRID;   // ReadFlowField.F90:72-82
  const int nyom = nyo - 1, nzom = nzo - 1;
  // one-rank window of ReadFlowField.F90:91-99 (integer divisions)
  const int js = std::max(nyom / nym, 1), je = std::min(nym * nyom / nym, nyom);
  const int is = std::max(nzom / nzm, 1), ie = std::min(nzm * nzom / nzm, nzom);
  if (je - js < 1 || ie - is < 1) return AFID_ERR_GRID;
  std::vector<double> tco, tmo, rco, rmo, zzo, zmo, tcn, tmn, rcn, rmn, zzn, zmn;
  if (!regrid_axis(nzo, c->p.zlen, 0, 1.0, tco, tmo) || !regrid_axis(nyo, c->p.ylen, 0, 1.0, rco, rmo) ||
      !regrid_axis(nxo, 1.0, istro3, stro3, zzo, zmo) || !regrid_axis(nz, c->p.zlen, 0, 1.0, tcn, tmn) ||
      !regrid_axis(ny, c->p.ylen, 0, 1.0, rcn, rmn) || !regrid_axis(nx, 1.0, c->p.istr3, c->p.str3, zzn, zmn))
    return AFID_ERR_GRID;
  for (auto *v : {&tco, &tmo}) extend_axis(*v, nzo);
  for (auto *v : {&rco, &rmo}) extend_axis(*v, nyo);
  for (auto *v : {&zzo, &zmo}) extend_axis(*v, nxo);
  // interp.F90:73-101: its "x" is the third index (z), its "z" the first (x)
  const std::vector<double> &xo = intvar == 1 ? tco : tmo, &xn = intvar == 1 ? tcn : tmn;
  const std::vector<double> &yo = intvar == 2 ? rco : rmo, &yn = intvar == 2 ? rcn : rmn;
  const std::vector<double> &zo = intvar == 3 ? zzo : zmo, &zn = intvar == 3 ? zzn : zmn;
  const int ne = nx + nym + nzm;
  std::vector<int> lohi(2 * (size_t)ne);
  std::vector<double> crd(3 * (size_t)ne);
  // first l in 1..n with old(l) >= x brackets x as (l-1, l); (n-1, n) when none
  auto fill = [&](int e0, int cnt, const std::vector<double> &on, int n, const std::vector<double> &nw, int lo_min,
                  int hi_max) {
    for (int q = 0; q < cnt; q++) {
      const double x = nw[q + 1];
      int lo = n - 1, hi = n;
      for (int l = 1; l <= n; l++)
        if (on[l] >= x) {
          lo = l - 1;
          hi = l;
          break;
        }
      if (lo < lo_min || hi > hi_max) return false;   // outside the reference's old window
      const int e = e0 + q;
      lohi[2 * e] = lo;
      lohi[2 * e + 1] = hi;
      crd[3 * e] = x;
      crd[3 * e + 1] = on[lo];
      crd[3 * e + 2] = on[hi];
    }
    return true;
  };
  if (!fill(0, nx, zo, nxo, zn, 0, nxo + 1) || !fill(nx, nym, yo, nyo, yn, js - 1, je + 1) ||
      !fill(nx + nym, nzm, xo, nzo, xn, is - 1, ie + 1))
    return AFID_ERR_GRID;
  const size_t nold = (size_t)nxo * nyom * nzom;
  const size_t b_old = nold * sizeof(double), b_crd = crd.size() * sizeof(double);
  const size_t b_lohi = lohi.size() * sizeof(int);
  int rc = ensure_stage(c, b_old + b_crd + b_lohi);
  if (rc) return rc;
  char *st = (char *)c->stage;
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(st, old, b_old, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(st + b_old, crd.data(), b_crd, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(st + b_old + b_crd, lohi.data(), b_lohi, hipMemcpyHostToDevice));
  InterpArgs a;
  a.g = g;
  a.out = c->f[field];
  a.old = (const double *)st;
  a.nxo = nxo;
  a.nyom = nyom;
  a.js = js;
  a.je = je;
  a.is = is;
  a.ie = ie;
  a.crd = (const double *)(st + b_old);
  a.lohi = (const int *)(st + b_old + b_crd);
  if ((rc = kerr(c, launch_interp(a, c->st)))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return AFID_OK;
}

int afid_synchronize(afid_ctx *c) {
  if (!c) return AFID_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return AFID_OK;
}

int afid_profile_enable(afid_ctx *c, int on) {
  if (!c) return AFID_ERR_ARG;
  c->prof = on != 0;
  return AFID_OK;
}

static void collect(afid_ctx *c) {
  hipStreamSynchronize(c->st);
  for (auto &pt : c->pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, pt.a, pt.b);
    auto &e = c->prof_tot[pt.name];
    e.first += ms;
    e.second += 1;
    hipEventDestroy(pt.a);
    hipEventDestroy(pt.b);
  }
  c->pending.clear();
}

int afid_profile_read(afid_ctx *c, char *names, int nbuf, double *total_ms, int64_t *count, int cap) {
  if (!c) return AFID_ERR_ARG;
  collect(c);
  int n = 0, used = 0;
  for (auto &kv : c->prof_tot) {
    if (n >= cap) break;
    int len = (int)kv.first.size() + 1;
    if (names && used + len <= nbuf) {
      std::memcpy(names + used, kv.first.c_str(), len);
      used += len;
    }
    if (total_ms) total_ms[n] = kv.second.first;
    if (count) count[n] = kv.second.second;
    n++;
  }
  return n;
}

int afid_profile_reset(afid_ctx *c) {
  if (!c) return AFID_ERR_ARG;
  collect(c);
  c->prof_tot.clear();
  return AFID_OK;
}

}  // extern "C"
