// kernels.h -- argument blocks and launchers of the libafid_hip kernels.
#pragma once
#include "afid_internal.h"

namespace afid {

struct FusedArgs {
  Geom g;
  FieldPtrs f;                       // fields at the start of the substep
  double *tn;                        // updated temperature
  double *rux, *ruy, *ruz, *rut;     // explicit-term history, updated in place
  double *stale;                     // rhs(1,j,i) of the reference, nym*nzm
  const double *kc;                  // KC_N interleaved coefficient rows
  const double *fac;                 // SYS_N*FR_N interleaved factor rows
  double *bx, *by, *bz, *bt;         // right-hand sides b + A q_old
  double *ux, *uy, *uz;              // u* (velocities before the correction) from the column solve:
                                     // the fields themselves (in place) or bx, by, bz
  StepScalars sc;
  const DevStep *ds;                 // non-null: scalars of substep ns come from the device
  int ns;
  int h32;                           // histories stored in FP32 (mixed precision)
  int zm;                            // z planes per stencil block (0: the default)
  int keep_ru;                       // 0: no later substep reads this substep's histories; not stored
};

struct RoutineArgs {
  Geom g;
  FieldPtrs f;
  const double *kc;
  double *out;          // explicit term / divergence output
  const double *expl;   // explicit term input of the RHS loop
  double *ru;           // history array of the RHS loop
  double *rhs;          // rhs scratch (reference local_arrays rhs)
  StepScalars sc;
};

struct SolveArgs {
  Geom g;
  const double *fac;    // FR_N rows of this system
  double *rhs;
  double *q;            // field updated in place
  double *stale;
  double tempbp, temptp;
};

struct CorrectArgs {
  Geom g;
  const double *dph;
  const double *ux, *uy, *uz;   // velocities before the correction (may equal vx, vy, vz)
  double *vx, *vy, *vz, *pr;
  const double *kc;
  StepScalars sc;
  const DevStep *ds;
  int ns;
  unsigned long long *cflbits;  // non-null: CalcMaxCFL of the corrected velocities, atomic max of the bits
};

// One complex FFT length with its radix stages, twiddles exp(-2 pi i t/n)
// and the mixed-radix digit reversal of the in-place decimation-in-time order.
// mdiv[s] = ceil(2^32 / m_s) for the span m_s of stage s (0 when m_s = 1):
// floor(w / m_s) = umulhi(w, mdiv[s]) exactly for w, m_s < 2^16.
struct FFTPlan {
  int n, nf;
  int f[16];
  unsigned mdiv[16];
  const double2 *tw;
  const int *rev;
};

// Real transform of length N along a lateral direction, computed as a complex
// transform of H = N/2 points (N even) or N points (N odd) in LDS.
struct RealPlan {
  int N, H, half;
  FFTPlan c;
  const double2 *post;   // exp(-2 pi i m/N), m = 0..N/2
};

struct PoissonArgs {
  Geom g;
  int nymh, ldm;         // y wavenumbers kept (nym/2+1), padded row length
  double2 *S;            // spectral array S[k][n][m]
  const double *piv;     // x-solve pivots at chunk starts, filled once per grid (launch_xsolve_pivots)
  const double *kc;      // for KC_ACPHK/AMPHK/APPHK (read through kx arrays below)
  const double *amphk, *acphk, *apphk;  // natural k order, length nxm
  const double *ak1, *ak2;              // z (nzm) and y (nym) modified wavenumbers
  FieldPtrs f;           // velocities for the fused divergence
  double *dph;           // real field (interleaved layout)
  const double *udx3m;   // interleaved row KC_UDX3M
  StepScalars sc;
  RealPlan py;
  FFTPlan pz;
  FFTPlan pzq;           // z transform on the matrix cores (zmf): the Q = nzm/16-point plan
  int zmf;               // 1: z transforms with a 16-point DFT stage on the matrix cores
  int from_velocity;     // 1: y-FFT input is div(u)/(dt al) computed on the fly; 0: dph
  const DevStep *ds;
  int ns;
  int spec32;            // S holds FP32 complex values and the y/z FFTs run in FP32 (mixed precision)
};

hipError_t launch_stencil_rhs(const FusedArgs &a, hipStream_t st);
hipError_t launch_column_solve(const FusedArgs &a, hipStream_t st);
hipError_t launch_solve_update(const SolveArgs &a, int which, hipStream_t st);
hipError_t launch_explicit(const RoutineArgs &a, int which, hipStream_t st);
hipError_t launch_rhs(const RoutineArgs &a, int which, hipStream_t st);
hipError_t launch_divergence(const RoutineArgs &a, hipStream_t st);
hipError_t launch_correct(const CorrectArgs &a, int which, hipStream_t st);
int max_supported_L();
size_t column_lds_need(const Geom &g);   // largest dynamic LDS of the split-update kernels
hipError_t column_configure_lds(size_t bytes);

hipError_t launch_yfft_forward(const PoissonArgs &a, hipStream_t st);
hipError_t launch_zfft(const PoissonArgs &a, int sign, hipStream_t st);
bool zfft_mfma_fits(int nzm);   // nzm = 16 Q with Q in {16, 32, 48, 64} (radices 2..5)
hipError_t launch_xsolve_pressure(const PoissonArgs &a, hipStream_t st);
hipError_t launch_xsolve_pivots(const PoissonArgs &a, double *piv, hipStream_t st);
void xsolve_shape(int nxm, int *np, int *lm);   // lanes per mode, register rows per lane
int xsolve_max_rows();
hipError_t launch_yfft_inverse(const PoissonArgs &a, hipStream_t st);
int fft_max_len();
hipError_t fft_configure_lds(size_t bytes);

// Reductions: results land in dev_out (double); the caller copies them.
struct ReduceArgs {
  Geom g;
  FieldPtrs f;
  const double *kc;
  const double *g3rc;      // natural order, length nx
  const double *dxc, *dxm, *lem, *lec;  // natural order, length nxm (dissipation)
  double dy, dz, ren, pec, pra;
  double *partial;         // scratch, nblk * NQ
  double *out;
};
hipError_t launch_reduce_cfl(const ReduceArgs &a, hipStream_t st);        // out[0]
hipError_t launch_reduce_divmax(const ReduceArgs &a, hipStream_t st);     // out[0]
hipError_t launch_reduce_global(const ReduceArgs &a, hipStream_t st);     // out[0..10] raw sums/max
hipError_t launch_reduce_platenu(const ReduceArgs &a, double del, double deln, hipStream_t st);  // out[0..1]
hipError_t launch_reduce_dissipation(const ReduceArgs &a, hipStream_t st);  // out[0..1], out[2..] profiles
int reduce_partial_blocks();
constexpr int STATS_N = 9;   // vx, vy, vz, T means; vx^2, vy^2, vz^2, T^2; T*vx
hipError_t launch_stats(const ReduceArgs &a, double *acc, hipStream_t st);   // acc[STATS_N][nxm] +=
size_t stats_partial_doubles(const Geom &g);
// Restart regridding (interp.F90:11-208).  The old field as the
// continuation file holds it, old[((i-1)*nyom + j-1)*nxo + k-1], read
// through the padded view arrold(0:nxo+1, js-1:je+1, is-1:ie+1) of the
// reference; per new k, j, i the bracketing old indices (lo, hi) and the
// coordinates (x, x1, x2), laid out k (nx), then j (nym), then i (nzm).
struct InterpArgs {
  Geom g;
  double *out;
  const double *old;
  int nxo, nyom, js, je, is, ie;
  const int *lohi;      // 2 per entry
  const double *crd;    // 3 per entry
};
hipError_t launch_interp(const InterpArgs &a, hipStream_t st);
// n FP64 values to FP32 (to_f32) or back, element for element.
hipError_t launch_convert(const void *src, void *dst, int64_t n, int to_f32, hipStream_t st);
hipError_t launch_slabs(const FieldPtrs &f, const Geom &g, const int *kslab, int nslab, double *out,
                        hipStream_t st);

// Layout conversion host-array staging <-> interleaved device field.
hipError_t launch_gather_in(const double *stage, double *dst, const Geom &g, int n1, int ld1, int ld2,
                            int j0, int i0, hipStream_t st);
hipError_t launch_scatter_out(const double *src, double *stage, const Geom &g, int n1, int ld1, int ld2,
                              int j0, int i0, int halo, hipStream_t st);
hipError_t launch_init_noise(double *temp, double *zero_fields[4], const double *xc_nat, const Geom &g,
                             uint64_t seed, double amp, double tbp, double ttp, hipStream_t st);
hipError_t launch_fill(double *p, double v, int64_t n, hipStream_t st);

// Device step control (DevStep): dt selection and substep scalars from the CFL
// reduction result, the wall-normal factorisations of the step's substeps, and
// the end-of-step bookkeeping.
struct FactorArgs {
  Geom g;
  const double *mc;     // natural-order rows (index kc-1): ac3ck, am3ssk, ap3ssk, ac3sk, am3sk, ap3sk, ac3ssk
  double *fac[3];       // factor rows per substep (SYS_N*FR_N*ldx each)
  double *work;         // scratch, SYS_N * 6 * nx doubles
  DevStep *ds;
  StepConsts k;
};
// cfl_out: CalcMaxCFL from k_red_cfl; null: from ds->cflbits, which the last
// correction of the previous step (or the host, before the first) filled.
hipError_t launch_step_begin(DevStep *ds, const double *cfl_out, StepConsts k, hipStream_t st);
hipError_t launch_factor(const FactorArgs &a, hipStream_t st);
hipError_t launch_step_end(DevStep *ds, hipStream_t st);

}  // namespace afid
