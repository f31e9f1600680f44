// afid_internal.h -- device data layout and kernel argument blocks shared by
// the HIP translation units of libafid_hip.
//
// Field layout in HBM (all FP64, one contiguous block per field):
//   column (j,i), j in [0,nym), i in [0,nzm), starts at ((i*nym)+j)*ldx.
//   Inside a column the nx wall-normal points are chunk-interleaved: the
//   column is cut into P=16 chunks of L=ceil(nx/16) points, and point k is
//   stored at pos(k) = (k mod L)*16 + k div L, ldx = 16*L.  A wave owns four
//   columns with one lane per chunk, so lane p's t-th point of all four
//   columns sits in one 128-byte run: every load of the column kernels is
//   coalesced, and the wall-normal tridiagonal solves stay in registers.
//   Periodic y/z neighbours are reached by index wrap; there are no halos.
//
// Spectral array (pressure solve): S[k][n][m] complex128, k in [0,nxm),
// n in [0,nzm) (z wavenumber), m in [0,ldm) (y wavenumber, m<nym/2+1 used),
// m fastest.  The x-solve holds each (m,n) mode's line in the registers of
// a group of lanes, one chunk of consecutive k per lane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace afid {

constexpr int P = 16;        // chunk lanes per column
constexpr int COLS_PER_WG = 16;  // 256 threads = 4 waves x 4 columns

struct Geom {
  int nx, nxm, nym, nzm;
  int L, ldx;       // chunk length, column pitch (16*L)
  int64_t ncol;     // nym*nzm
};

__host__ __device__ inline int pos_of_k(int k, int L) { return (k % L) * P + k / L; }
__host__ __device__ inline int k_of_pos(int s, int L) { return (s % P) * L + s / P; }
__host__ __device__ inline int64_t col_off(int j, int i, const Geom &g) {
  return ((int64_t)i * g.nym + j) * g.ldx;
}
__host__ __device__ inline int wrapm(int a, int n) { return a < 0 ? a + n : a; }
__host__ __device__ inline int wrapp(int a, int n) { return a >= n ? a - n : a; }

// XCD-aware block order: the hardware deals blocks b, b+8, b+16 ... to one
// XCD, so each XCD is given one contiguous range of the logical block order
// (grid padded to a multiple of 8).  Returns -1 for padding blocks.
__device__ __forceinline__ int64_t xcd_block(int64_t nblk) {
  int64_t b = blockIdx.x;
  int64_t per = (gridDim.x + 7) / 8;
  int64_t nb = (b % 8) * per + b / 8;
  return nb < nblk ? nb : -1;
}

// Logical column order of the point kernels: strips STRIP_W columns wide in y,
// z next, y within the strip fastest.  A column's y and z neighbours are then
// at most STRIP_W columns away in that order and are read from the L2 of the
// XCD that streams the strip.
constexpr int STRIP_W = 16;
__device__ __forceinline__ void strip_col(int64_t c, const Geom &g, int &j, int &i) {
  const int64_t per_strip = (int64_t)STRIP_W * g.nzm;
  const int strip = (int)(c / per_strip);
  const int r = (int)(c - strip * per_strip);
  const int sw = min(STRIP_W, g.nym - strip * STRIP_W);
  i = r / sw;
  j = strip * STRIP_W + (r - i * sw);
}

// Per-k coefficient rows held on device in the interleaved layout: row q of
// the block kc is kc[q*ldx + pos(k)].
enum KCoef {
  KC_UDX3C = 0, KC_UDX3M, KC_AM3CK, KC_AC3CK, KC_AP3CK, KC_AM3SK, KC_AC3SK, KC_AP3SK,
  KC_AMPHK, KC_ACPHK, KC_APPHK, KC_KMVPOS, KC_KPVPOS, KC_N
};

// Factor rows of one wall-normal system (interleaved layout, length ldx):
// scale = row normalisation 1/(1 - ac*betadx) (SolveImpEqnUpdate_*),
// l = LU multipliers, rinv = 1/u, e = du/u, G/H = chunk prefix/suffix
// products of -l / -e used by the lane scan, dl/du = the normalised off-
// diagonals (unit diagonal) themselves.
enum FacRow { FR_SCALE = 0, FR_L, FR_RINV, FR_E, FR_G, FR_H, FR_DL, FR_DU, FR_N };
enum System { SYS_X = 0, SYS_YZ, SYS_T, SYS_N };

struct StepScalars {
  double al, ga, ro, dt, beta;
  double alre, alpec;            // al/ren, al/pec
  double udy4, udz4;             // dy/4, dz/4
  double udyq_re, udzq_re;       // dyq/ren, dzq/ren
  double udyq_pe, udzq_pe;       // dyq/pec, dzq/pec
  double dy, dz, dyq, dzq;
  double udy_al;                 // dy*al (ImplicitAndUpdateVY.F90:33)
  double tempbp, temptp;
  int use_ru;                    // ro != 0
};

// Step state kept on the device so that whole time steps run without a host
// round trip (main.F90:167-200): dt selection from the CFL reduction, the
// substep scalars, error codes and time.  `halted` is set at the end of a step
// that raised an error code; every field-writing kernel of later steps then
// returns at once, so a graph of many steps stops where main.F90 would.
struct DevStep {
  double dt, time, dtfix, dtmin, dtmax, limitcfl;
  int ntime, idtv, errorcode, halted, pivot, steps_done;
  StepScalars sc[3];   // scalars of substeps 1..3 of the current step
  unsigned long long cflbits;   // CalcMaxCFL for the next step (bits of a non-negative double)
};

// Values that one kernel of a step writes and a later kernel reads (the
// device step state, the stale row, reduction results read by one thread)
// are loaded with device-coherent vector loads (agent-scope atomic loads from
// L2).  A plain load at a uniform address compiles to a scalar load, which a
// later kernel of a replayed step graph can find served from the scalar cache
// with an earlier step's line.
__device__ __forceinline__ int coh_int(const int *p) {
  return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long coh_u64(const unsigned long long *p) {
  return __hip_atomic_load(const_cast<unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double coh_dbl(const double *p) {
  return __longlong_as_double((long long)coh_u64(reinterpret_cast<const unsigned long long *>(p)));
}
// The scalars are the same in every lane: each word is moved to scalar
// registers (readfirstlane) after its coherent load, so that they do not hold
// vector registers for the whole kernel (the stencil pass would drop from
// three to two waves per SIMD).  Call with every lane of the wave active.
__device__ __forceinline__ StepScalars coh_scalars(const StepScalars *p) {
  static_assert(sizeof(StepScalars) % 8 == 0, "StepScalars is read in 8-byte words");
  StepScalars s;
  const unsigned long long *src = reinterpret_cast<const unsigned long long *>(p);
  unsigned long long *dst = reinterpret_cast<unsigned long long *>(&s);
#pragma unroll
  for (int q = 0; q < (int)(sizeof(StepScalars) / 8); q++) {
    const unsigned long long v = coh_u64(src + q);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    dst[q] = ((unsigned long long)hi << 32) | lo;
  }
  return s;
}

// Host-set constants the device step control needs to rebuild StepScalars.
struct StepConsts {
  double ren, pec, dy, dz, dyq, dzq, tempbp, temptp;
  double gam[3], rom[3], alm[3];
  int nsst;
};

struct FieldPtrs {
  const double *vx, *vy, *vz, *temp, *pr;
};

}  // namespace afid
