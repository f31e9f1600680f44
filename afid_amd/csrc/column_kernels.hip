// column_kernels.hip -- explicit terms, implicit Crank-Nicolson solves in the
// wall-normal direction, corrections, on the chunk-interleaved column layout.
//
// Reference routines restated (AFiD 1.1):
//   ExplicitTermsVX.F90:42-104, ExplicitTermsVY.F90:40-100,
//   ExplicitTermsVZ.F90:41-104, ExplicitTermsTemp.F90:37-95,
//   ImplicitAndUpdateVX.F90:33-77, ImplicitAndUpdateVY.F90:36-74,
//   ImplicitAndUpdateVZ.F90:35-71, ImplicitAndUpdateTemp.F90:34-73,
//   SolveImpEqnUpdate_X.F90, _YZ.F90, _Temp.F90 (dgttrf/dgttrs replaced by a
//   pivot-free LU applied with a 16-lane affine scan),
//   CalcLocalDivergence.F90:27-41, CorrectVelocity.F90:28-45,
//   CorrectPressure.F90:32-52.
// The order of floating-point operations inside each stencil follows the
// reference expressions (parenthesisation kept), so per-routine results match
// the reference to rounding.
#include <algorithm>
#include <type_traits>

#include "afid_internal.h"
#include "kernels.h"

namespace afid {

struct Nb {
  int64_t C, CJM, CJP, CIM, CIP, CJMIP, CJPIM;
};

__device__ __forceinline__ Nb make_nb(int j, int i, const Geom &g) {
  int jm = wrapm(j - 1, g.nym), jp = wrapp(j + 1, g.nym);
  int im = wrapm(i - 1, g.nzm), ip = wrapp(i + 1, g.nzm);
  Nb n;
  n.C = col_off(j, i, g);
  n.CJM = col_off(jm, i, g);
  n.CJP = col_off(jp, i, g);
  n.CIM = col_off(j, im, g);
  n.CIP = col_off(j, ip, g);
  n.CJMIP = col_off(jm, ip, g);
  n.CJPIM = col_off(jp, im, g);
  return n;
}

// ---------------------------------------------------------------- explicit terms
// ExplicitTermsVX.F90:55-101 at interior point k (1..nx-2)
__device__ __forceinline__ double expl_vx(const FieldPtrs &f, const Nb &n, int s, int skm, int skp,
                                          double udx3c, const StepScalars &c) {
  const double *vx = f.vx, *vy = f.vy, *vz = f.vz;
  const double vxc = vx[n.C + s];
  double hxz = (((vz[n.CIP + s] + vz[n.CIP + skm]) * (vx[n.CIP + s] + vxc)) -
                ((vz[n.C + s] + vz[n.C + skm]) * (vxc + vx[n.CIM + s]))) * c.udz4;
  double hxy = (((vy[n.CJP + s] + vy[n.CJP + skm]) * (vx[n.CJP + s] + vxc)) -
                ((vy[n.C + s] + vy[n.C + skm]) * (vxc + vx[n.CJM + s]))) * c.udy4;
  double a = vx[n.C + skp] + vxc, b = vxc + vx[n.C + skm];
  double hxx = (a * a - b * b) * udx3c * 0.25;
  double dzzvx = (vx[n.CIM + s] - 2.0 * vxc + vx[n.CIP + s]) * c.udzq_re;
  double dyyvx = (vx[n.CJM + s] - 2.0 * vxc + vx[n.CJP + s]) * c.udyq_re;
  return -(hxx + hxy + hxz) + dyyvx + dzzvx + f.temp[n.C + s];
}

// ExplicitTermsVY.F90:52-98 at k (0..nxm-1); skmm/skpp are the clamped kmv/kpv
__device__ __forceinline__ double expl_vy(const FieldPtrs &f, const Nb &n, int s, int skp, int skmm,
                                          int skpp, double udx3m, const StepScalars &c) {
  const double *vx = f.vx, *vy = f.vy, *vz = f.vz;
  const double vyc = vy[n.C + s];
  double hyx = ((vx[n.C + skp] + vx[n.CJM + skp]) * (vy[n.C + skpp] + vyc) -
                (vx[n.C + s] + vx[n.CJM + s]) * (vyc + vy[n.C + skmm])) * udx3m * 0.25;
  double yp = vy[n.CJP + s] + vyc, ym = vy[n.CJM + s] + vyc;
  double hyy = (yp * yp - ym * ym) * c.udy4;
  double hyz = ((vy[n.CIP + s] + vyc) * (vz[n.CIP + s] + vz[n.CJMIP + s]) -
                (vyc + vy[n.CIM + s]) * (vz[n.C + s] + vz[n.CJM + s])) * c.udz4;
  double dyyvy = (vy[n.CJP + s] - 2.0 * vyc + vy[n.CJM + s]) * c.udyq_re;
  double dzzvy = (vy[n.CIP + s] - 2.0 * vyc + vy[n.CIM + s]) * c.udzq_re;
  return -(hyx + hyy + hyz) + dyyvy + dzzvy;
}

// ExplicitTermsVZ.F90:56-101 at k (0..nxm-1)
__device__ __forceinline__ double expl_vz(const FieldPtrs &f, const Nb &n, int s, int skp, int skmm,
                                          int skpp, double udx3m, const StepScalars &c) {
  const double *vx = f.vx, *vy = f.vy, *vz = f.vz;
  const double vzc = vz[n.C + s];
  double zp = vz[n.CIP + s] + vzc, zm = vz[n.CIM + s] + vzc;
  double hzz = (zp * zp - zm * zm) * c.udz4;
  double hzy = ((vy[n.CJP + s] + vy[n.CJPIM + s]) * (vz[n.CJP + s] + vzc) -
                (vy[n.C + s] + vy[n.CIM + s]) * (vzc + vz[n.CJM + s])) * c.udy4;
  double hzx = ((vx[n.C + skp] + vx[n.CIM + skp]) * (vz[n.C + skpp] + vzc) -
                (vx[n.C + s] + vx[n.CIM + s]) * (vzc + vz[n.C + skmm])) * udx3m * 0.25;
  double dzzvz = (vz[n.CIP + s] - 2.0 * vzc + vz[n.CIM + s]) * c.udzq_re;
  double dyyvz = (vz[n.CJP + s] - 2.0 * vzc + vz[n.CJM + s]) * c.udyq_re;
  return -(hzx + hzy + hzz) + dyyvz + dzzvz;
}

// ExplicitTermsTemp.F90:52-93 at interior k (1..nx-2)
__device__ __forceinline__ double expl_temp(const FieldPtrs &f, const Nb &n, int s, int skm, int skp,
                                            double udx3c, const StepScalars &c) {
  const double *vx = f.vx, *vy = f.vy, *vz = f.vz, *T = f.temp;
  const double tc = T[n.C + s];
  double htz = ((vz[n.CIP + skm] + vz[n.CIP + s]) * (T[n.CIP + s] + tc) -
                (vz[n.C + skm] + vz[n.C + s]) * (tc + T[n.CIM + s])) * c.udz4;
  double hty = ((vy[n.CJP + s] + vy[n.CJP + skm]) * (T[n.CJP + s] + tc) -
                (vy[n.C + s] + vy[n.C + skm]) * (tc + T[n.CJM + s])) * c.udy4;
  double htx = ((vx[n.C + skp] + vx[n.C + s]) * (T[n.C + skp] + tc) -
                (vx[n.C + s] + vx[n.C + skm]) * (tc + T[n.C + skm])) * udx3c * 0.25;
  double dzzt = (T[n.CIP + s] - 2.0 * tc + T[n.CIM + s]) * c.udzq_pe;
  double dyyt = (T[n.CJP + s] - 2.0 * tc + T[n.CJM + s]) * c.udyq_pe;
  return -(htx + hty + htz) + dyyt + dzzt;
}

// ---------------------------------------------------------------- RHS assembly
// ImplicitAndUpdateVX.F90:47-60 (unscaled rhs; the x-solve applies ackl_b)
__device__ __forceinline__ double rhs_vx(double q, double ru, const double *vx, const double *pr,
                                         const Nb &n, int s, int skm, int skp, const double *kc,
                                         int ldx, const StepScalars &c) {
  double dxxvx = vx[n.C + skp] * kc[KC_AP3CK * ldx + s] + vx[n.C + s] * kc[KC_AC3CK * ldx + s] +
                 vx[n.C + skm] * kc[KC_AM3CK * ldx + s];
  double udx3 = c.al * kc[KC_UDX3C * ldx + s];
  double dxp = (pr[n.C + s] - pr[n.C + skm]) * udx3;
  double r = c.use_ru ? c.ro * ru : 0.0;
  return (c.ga * q + r + c.alre * dxxvx - dxp) * c.dt;
}
// ImplicitAndUpdateVY.F90:52-68
__device__ __forceinline__ double rhs_vy(double q, double ru, const double *vy, const double *pr,
                                         const Nb &n, int s, int skmm, int skpp, const double *kc,
                                         int ldx, const StepScalars &c) {
  double dxxvy = vy[n.C + skpp] * kc[KC_AP3SK * ldx + s] + vy[n.C + s] * kc[KC_AC3SK * ldx + s] +
                 vy[n.C + skmm] * kc[KC_AM3SK * ldx + s];
  double dyp = (pr[n.C + s] - pr[n.CJM + s]) * c.udy_al;
  double r = c.use_ru ? c.ro * ru : 0.0;
  return (c.ga * q + r + c.alre * dxxvy - dyp) * c.dt;
}
// ImplicitAndUpdateVZ.F90:48-64
__device__ __forceinline__ double rhs_vz(double q, double ru, const double *vz, const double *pr,
                                         const Nb &n, int s, int skmm, int skpp, const double *kc,
                                         int ldx, const StepScalars &c) {
  double dxxvz = vz[n.C + skpp] * kc[KC_AP3SK * ldx + s] + vz[n.C + s] * kc[KC_AC3SK * ldx + s] +
                 vz[n.C + skmm] * kc[KC_AM3SK * ldx + s];
  double dzp = (pr[n.C + s] - pr[n.CIM + s]) * c.dz * c.al;
  double r = c.use_ru ? c.ro * ru : 0.0;
  return (c.ga * q + r + c.alre * dxxvz - dzp) * c.dt;
}
// ImplicitAndUpdateTemp.F90:41-54
__device__ __forceinline__ double rhs_temp(double q, double ru, const double *T, const Nb &n, int s,
                                           int skm, int skp, const double *kc, int ldx,
                                           const StepScalars &c) {
  double dxxt = T[n.C + skp] * kc[KC_AP3CK * ldx + s] + T[n.C + s] * kc[KC_AC3CK * ldx + s] +
                T[n.C + skm] * kc[KC_AM3CK * ldx + s];
  double r = c.use_ru ? c.ro * ru : 0.0;
  return (c.ga * q + r + c.alpec * dxxt) * c.dt;
}

// ---------------------------------------------------------------- lane scan solve
// Solves (L U) x = b for one column held as 16 chunks: lane p owns points
// k = p*L + t (t < L) of an n-point system; its values sit in LDS at
// bl[t*WG] (WG = block size, so the 256 lanes of a block hit distinct banks).
// Forward  y_k = b_k - l_k y_{k-1}, backward x_k = y_k rinv_k - e_k x_{k+1}:
// each is an affine recurrence, evaluated per chunk from a zero carry and
// fixed up with the true carry from a Kogge-Stone scan of the chunk maps
// (A = product of -l or -e over the chunk, B = local end value).  The forward
// fix-up is folded into the backward sweep; the backward fix-up is returned
// as cin so that the caller folds it into its write-out: x = bl[t] + H_t*cin.
constexpr int WG = 256;

__device__ __forceinline__ double scan_up(double A, double B, int p) {
#pragma unroll
  for (int off = 1; off < P; off <<= 1) {
    double A2 = __shfl_up(A, off, P), B2 = __shfl_up(B, off, P);
    if (p >= off) {
      B = A * B2 + B;
      A = A * A2;
    }
  }
  double cin = __shfl_up(B, 1, P);
  return p == 0 ? 0.0 : cin;
}
__device__ __forceinline__ double scan_down(double A, double B, int p) {
#pragma unroll
  for (int off = 1; off < P; off <<= 1) {
    double A2 = __shfl_down(A, off, P), B2 = __shfl_down(B, off, P);
    if (p + off < P) {
      B = A * B2 + B;
      A = A * A2;
    }
  }
  double cin = __shfl_down(B, 1, P);
  return p == P - 1 ? 0.0 : cin;
}

__device__ __forceinline__ double chunk_solve(double *bl, const double *fac, int ldx, int cnt, int p) {
  const double *fl = fac + FR_L * ldx, *fr = fac + FR_RINV * ldx, *fe = fac + FR_E * ldx;
  const double *fg = fac + FR_G * ldx, *fh = fac + FR_H * ldx;
  double acc = 0.0;
  for (int t = 0; t < cnt; t++) {
    acc = bl[t * WG] - fl[t * P + p] * acc;
    bl[t * WG] = acc;
  }
  double cin = scan_up(cnt > 0 ? fg[(cnt - 1) * P + p] : 1.0, cnt > 0 ? acc : 0.0, p);
  acc = 0.0;
  for (int t = cnt - 1; t >= 0; t--) {
    const int s = t * P + p;
    double y = bl[t * WG] + fg[s] * cin;
    acc = y * fr[s] - fe[s] * acc;
    bl[t * WG] = acc;
  }
  return scan_down(cnt > 0 ? fh[p] : 1.0, cnt > 0 ? acc : 0.0, p);
}

// ---------------------------------------------------------------- load/store helpers
// Byte-offset loads and stores against a uniform base pointer (32-bit offsets
// while a field fits in 4 GiB).
template <typename OffT>
__device__ __forceinline__ double ldf(const double *base, OffT boff) {
  return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + boff);
}

// Non-temporal stores (streamed past the caches' normal retention): the
// stencil pass's b' and history stores and the column solve's new fields,
// which no later access of the same pass reads (about 2-3 % of each pass's
// time at 384x768x768; the arrays are far larger than the caches).
template <typename OffT>
__device__ __forceinline__ void stf_nt(double *base, OffT boff, double v) {
  __builtin_nontemporal_store(v, reinterpret_cast<double *>(reinterpret_cast<char *>(base) + boff));
}
// History arrays (rux..rutemp): FP64, or FP32 in the mixed-precision mode (H32),
// where element e sits at byte 4e, half the FP64 byte offset.
// FP32 histories store normally: a tile row of 16 positions is half a 128-byte
// line, whose other half the neighbouring tile writes, and the halves merge in
// the L2 only when the stores are ordinary.
template <bool H32, typename OffT>
__device__ __forceinline__ void sth_nt(double *base, OffT boff, double v) {
  if (H32) *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + (boff >> 1)) = (float)v;
  else stf_nt(base, boff, v);
}
// FP64 history loads of the stencil pass (read once per substep, by that pass
// alone) stream past the caches too, leaving the L2 to the field lines that
// neighbouring tiles read again: 3 % less fetched, about 1 % faster at
// 384x768x768.  FP32 histories, whose tile rows are 64-byte half lines, load
// normally (streamed, that pass ran 4 % slower).
template <bool H32, typename OffT>
__device__ __forceinline__ double ldh_nt(const double *base, OffT boff) {
  if (H32) return (double)*reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + (boff >> 1));
  return __builtin_nontemporal_load(reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + boff));
}

// ---------------------------------------------------------------- substep update
// The explicit terms, right-hand sides and wall-normal solves of one substep
// in two passes.  Each system is solved for the new value directly:
// A q_new = b + A q_old, with A the unit-diagonal normalised matrix, which
// equals q_old + A^-1 b of SolveImpEqnUpdate_* and spares re-reading q_old.
// k_stencil_rhs: the explicit terms and right-hand sides of all four fields
// (ExplicitTerms*, ImplicitAndUpdate* RHS loops), the history update, and
// b' = ackl_b*rhs + A q_old, where A is the unit-diagonal normalised system of
// SolveImpEqnUpdate_*; identity rows store q_old (+ rhs(1) = the stale row for
// vx).  A block holds ST_GC = 16 neighbouring y columns, all at the same
// ST_SP = 16 stored positions (one 128-byte line of each column; a wave covers
// four columns), and marches through zplanes consecutive z planes: each thread
// keeps its position's values of planes i-1, i and i+1 in registers, so the z
// neighbours cost no loads.  Each thread publishes the plane values its y
// neighbours need in LDS (double-buffered by plane, one barrier per plane), so
// only the group's two edge columns read a y neighbour from memory.  A column
// holds ldx = 16 L stored positions, so its L tiles are all full.  Blocks are
// ordered position-range fastest, then y group, then z range (XCD-remapped),
// so the k+-1 lines a tile reads belong to tiles running on the same XCD at
// the same time.  A plane's loads issue in one batch before any value is
// used; the history rows outside the reference loops (k = 1 and nx for
// vx/temp, nx for vy/vz), which nothing reads, are kept at 0.  Histories that no later substep reads (keep_ru = 0: RK3
// substep 3, since the next step's substep 1 has ro = 0) are not stored.
// k_column_solve: a wave owns four columns (16 chunk lanes each) and solves the
// four systems in registers: A q_new = b', with the factor rows staged in LDS.
// The new velocities u* (before the pressure correction) go to (ux, uy, uz):
// the fields themselves, in place (each lane rewrites only positions of its
// own column, and reads vz at k = 1 before writing it), or, where the
// correction must see u* apart from the corrected fields, the b buffers over
// b' itself.  The new temperature is final and goes to its field.  The vz
// increment at k=1, q_new - q_old, becomes the new stale row and the
// temperature's rhs(1).
constexpr int ZM_DEFAULT = 32;   // z planes per block unless afid_tune chose another count
constexpr int SY_WG = 256;             // threads per block of the stencil and correction passes
constexpr int ST_SP = 16;              // stored positions per column tile (one 128-byte line)
constexpr int ST_GC = SY_WG / ST_SP;   // y columns per block, four per wave
// Coefficient rows staged in LDS by k_stencil_rhs: eight per-k metric rows,
// then scale, dl, du of the x, y/z and temperature systems.
enum StencilCoef {
  SC_UDX3C = 0, SC_UDX3M, SC_AMC, SC_ACC, SC_APC, SC_AMS, SC_ACS, SC_APS,
  SC_XSC, SC_XDL, SC_XDU, SC_YSC, SC_YDL, SC_YDU, SC_TSC, SC_TDL, SC_TDU, SC_N
};
__device__ __forceinline__ int kc_row_of(int r) {
  const int rows[8] = {KC_UDX3C, KC_UDX3M, KC_AM3CK, KC_AC3CK, KC_AP3CK, KC_AM3SK, KC_AC3SK, KC_AP3SK};
  return rows[r];
}
__device__ __forceinline__ int fac_row_of(int q) { return q == 0 ? FR_SCALE : (q == 1 ? FR_DL : FR_DU); }

// The centre-column values of one plane at a thread's position (k-1, k, k+1).
struct Plane {
  double vxm, vxc, vxp, vym, vyc, vyp, vzm, vzc, vzp, tm, tc, tp, prm, prc;
};
template <typename OffT>
__device__ __forceinline__ Plane load_plane(const double *__restrict__ vx, const double *__restrict__ vy,
                                            const double *__restrict__ vz, const double *__restrict__ T,
                                            const double *__restrict__ pr, OffT C, OffT smb, OffT sb, OffT spb) {
  Plane q;
  q.vxm = ldf(vx, C + smb); q.vxc = ldf(vx, C + sb); q.vxp = ldf(vx, C + spb);
  q.vym = ldf(vy, C + smb); q.vyc = ldf(vy, C + sb); q.vyp = ldf(vy, C + spb);
  q.vzm = ldf(vz, C + smb); q.vzc = ldf(vz, C + sb); q.vzp = ldf(vz, C + spb);
  q.tm = ldf(T, C + smb); q.tc = ldf(T, C + sb); q.tp = ldf(T, C + spb);
  q.prm = ldf(pr, C + smb); q.prc = ldf(pr, C + sb);
  return q;
}

// Plane values a wave publishes for its y neighbours (the k-1 value of vy and
// the k+1 value of vx are those at the neighbour's own position s).
enum { SV_VX = 0, SV_VY, SV_VYM, SV_VZ, SV_T, SV_VXP, SV_PR, SV_VZIP, SV_N };
static size_t stencil_lds(const Geom &) { return sizeof(double) * (SC_N * ST_SP + 2 * SV_N * SY_WG); }

// DEV: the substep scalars come from the device step state (a.ds) rather
// than from the argument block; two instantiations keep each one's scalar
// loads unconditional.
template <typename OffT, bool USE_RU, bool DEV, bool H32>
__global__ void __launch_bounds__(SY_WG)
    k_stencil_rhs(FusedArgs a, const double *__restrict__ vx, const double *__restrict__ vy,
                  const double *__restrict__ vz, const double *__restrict__ T, const double *__restrict__ pr,
                  double *__restrict__ rux, double *__restrict__ ruy, double *__restrict__ ruz,
                  double *__restrict__ rut, double *__restrict__ bx, double *__restrict__ by,
                  double *__restrict__ bz, double *__restrict__ bt, const double *__restrict__ stalep,
                  const double *__restrict__ kc, const double *__restrict__ facp) {
  extern __shared__ double sh[];
  double *cf = sh;                    // [SC_N][ST_SP] coefficient rows of the block's positions
  double *xv = sh + SC_N * ST_SP;     // [2][SV_N][SY_WG] published plane values
  const Geom g = a.g;
  const int L = g.L, ldx = g.ldx, nx = g.nx, nxm = g.nxm;
  const int nq = (ldx + ST_SP - 1) / ST_SP;
  const int ngr = (g.nym + ST_GC - 1) / ST_GC;
  const int zplanes = a.zm > 0 ? a.zm : ZM_DEFAULT;
  const int nib = (g.nzm + zplanes - 1) / zplanes;
  const int64_t tile = xcd_block((int64_t)nib * ngr * nq);
  if (tile < 0) return;
  const int q = (int)(tile % nq);
  const int64_t rest = tile / nq;
  const int gr = (int)(rest % ngr), ib = (int)(rest / ngr);
  if (DEV && coh_int(&a.ds->halted)) return;
  const StepScalars c = DEV ? coh_scalars(&a.ds->sc[a.ns - 1]) : a.sc;
  const bool keep_ru = a.keep_ru != 0;
  for (int idx = threadIdx.x; idx < SC_N * ST_SP; idx += SY_WG) {
    const int r = idx / ST_SP, sl = min(q * ST_SP + idx - r * ST_SP, ldx - 1);
    const int src = r < SC_XSC ? kc_row_of(r) * ldx : (((r - SC_XSC) / 3) * FR_N + fac_row_of((r - SC_XSC) % 3)) * ldx;
    cf[idx] = (r < SC_XSC ? kc : facp)[src + sl];
  }
  __syncthreads();
  const int lane = (int)threadIdx.x & (ST_SP - 1);        // position within the tile
  const int w = (int)threadIdx.x / ST_SP;                  // column within the group
  const int ncols = min(ST_GC, g.nym - gr * ST_GC);        // columns of this group
  const bool colok = w < ncols;
  const int jc = colok ? gr * ST_GC + w : g.nym - 1;       // idle threads read a valid column, store nothing
  const int sraw = q * ST_SP + lane;
  const int s = min(sraw, ldx - 1);
  const int k = k_of_pos(s, L);
  // Every stored position of a column is written, the padding (k >= nx) as
  // zeros: a 128-byte line written in part costs the memory system about as
  // much again as a whole one (1.6 ms of 8.6 per launch at 384x768x768).
  const bool stv = colok && sraw < ldx;
  const bool pad = k >= nx;
  const bool jp_in = w + 1 < ncols, jm_in = w > 0;        // neighbour column in this block
  const int jm = wrapm(jc - 1, g.nym), jp = wrapp(jc + 1, g.nym);
  const OffT cb = (OffT)ldx * sizeof(double);
  auto colb = [&](int jj, int ii) { return ((OffT)ii * (OffT)g.nym + (OffT)jj) * cb; };
  const OffT sb = (OffT)s * sizeof(double);
  const OffT smb = (OffT)(k >= 1 ? pos_of_k(k - 1, L) : s) * sizeof(double);
  const OffT spb = (OffT)(k + 1 < nx ? pos_of_k(k + 1, L) : s) * sizeof(double);
  const bool inx = k >= 1 && k <= nx - 2;
  const bool iny = k <= nxm - 1;
  auto cfv = [&](int r) { return cf[r * ST_SP + lane]; };
  const double udx3c = cfv(SC_UDX3C), udx3m = cfv(SC_UDX3M);
  const double amc = cfv(SC_AMC), acc = cfv(SC_ACC), apc = cfv(SC_APC);
  const double ams = cfv(SC_AMS), acs = cfv(SC_ACS), aps = cfv(SC_APS);
  const int i0 = ib * zplanes, i1 = min(i0 + zplanes, g.nzm);
  Plane cur = load_plane<OffT>(vx, vy, vz, T, pr, colb(jc, i0), smb, sb, spb);
  const int ib0 = wrapm(i0 - 1, g.nzm);
  const OffT CP = colb(jc, ib0);
  double vxim = ldf(vx, CP + sb), vximp = ldf(vx, CP + spb), vyim = ldf(vy, CP + sb);
  double vzim = ldf(vz, CP + sb), tim = ldf(T, CP + sb), prim = ldf(pr, CP + sb);
  double vyjpim = ldf(vy, colb(jp, ib0) + sb);
  const int tid = (int)threadIdx.x;
  int par = 0;
  for (int ic = i0; ic < i1; ic++) {
    const int ip = wrapp(ic + 1, g.nzm);
    const OffT C = colb(jc, ic), CJM = colb(jm, ic), CJP = colb(jp, ic);
    const Plane nxt = load_plane<OffT>(vx, vy, vz, T, pr, colb(jc, ip), smb, sb, spb);
    double vxjp = 0.0, vyjp = 0.0, vyjpm = 0.0, vzjp = 0.0, tjp = 0.0;
    double vxjm = 0.0, vxjmp = 0.0, vyjm = 0.0, vzjm = 0.0, tjm = 0.0, prjm = 0.0, vzjmip = 0.0;
    if (!jp_in) {
      vxjp = ldf(vx, CJP + sb); vyjp = ldf(vy, CJP + sb); vyjpm = ldf(vy, CJP + smb);
      vzjp = ldf(vz, CJP + sb); tjp = ldf(T, CJP + sb);
    }
    if (!jm_in) {
      vxjm = ldf(vx, CJM + sb); vxjmp = ldf(vx, CJM + spb); vyjm = ldf(vy, CJM + sb);
      vzjm = ldf(vz, CJM + sb); tjm = ldf(T, CJM + sb); prjm = ldf(pr, CJM + sb);
      vzjmip = ldf(vz, colb(jm, ip) + sb);
    }
    const double rxo = USE_RU ? ldh_nt<H32>(rux, C + sb) : 0.0, ryo = USE_RU ? ldh_nt<H32>(ruy, C + sb) : 0.0;
    const double rzo = USE_RU ? ldh_nt<H32>(ruz, C + sb) : 0.0, rto = USE_RU ? ldh_nt<H32>(rut, C + sb) : 0.0;
    const double st0 = q == 0 ? coh_dbl(stalep + (int64_t)ic * g.nym + jc) : 0.0;   // k = 0 lies in row 0
    const double stale0 = k == 0 ? st0 : 0.0;
    double *buf = xv + par * (SV_N * SY_WG);
    buf[SV_VX * SY_WG + tid] = cur.vxc;
    buf[SV_VY * SY_WG + tid] = cur.vyc;
    buf[SV_VYM * SY_WG + tid] = cur.vym;
    buf[SV_VZ * SY_WG + tid] = cur.vzc;
    buf[SV_T * SY_WG + tid] = cur.tc;
    buf[SV_VXP * SY_WG + tid] = cur.vxp;
    buf[SV_PR * SY_WG + tid] = cur.prc;
    buf[SV_VZIP * SY_WG + tid] = nxt.vzc;
    __syncthreads();
    if (jp_in) {
      const int o = tid + ST_SP;
      vxjp = buf[SV_VX * SY_WG + o]; vyjp = buf[SV_VY * SY_WG + o]; vyjpm = buf[SV_VYM * SY_WG + o];
      vzjp = buf[SV_VZ * SY_WG + o]; tjp = buf[SV_T * SY_WG + o];
    }
    if (jm_in) {
      const int o = tid - ST_SP;
      vxjm = buf[SV_VX * SY_WG + o]; vxjmp = buf[SV_VXP * SY_WG + o]; vyjm = buf[SV_VY * SY_WG + o];
      vzjm = buf[SV_VZ * SY_WG + o]; tjm = buf[SV_T * SY_WG + o]; prjm = buf[SV_PR * SY_WG + o];
      vzjmip = buf[SV_VZIP * SY_WG + o];
    }
    par ^= 1;
    const double vxm = cur.vxm, vxc = cur.vxc, vxp = cur.vxp, vym = cur.vym, vyc = cur.vyc, vyp = cur.vyp;
    const double vzm = cur.vzm, vzc = cur.vzc, vzp = cur.vzp, tm = cur.tm, tc = cur.tc, tp = cur.tp;
    const double prm = cur.prm, prc = cur.prc;
    const double vxip = nxt.vxc, vyip = nxt.vyc, vzip = nxt.vzc, vzipm = nxt.vzm, tip = nxt.tc;
    if (stv) {
      // The explicit terms in E: FP64, or FP32 in the mixed-precision mode (their
      // inputs rounded from the FP64 fields); the right-hand sides, b' and the
      // solves stay FP64.  With E = double the expressions are the FP64 ones.
      using E = typename std::conditional<H32, float, double>::type;
      const E Evxm = vxm, Evxc = vxc, Evxp = vxp, Evym = vym, Evyc = vyc, Evyp = vyp;
      const E Evzm = vzm, Evzc = vzc, Evzp = vzp, Etm = tm, Etc = tc, Etp = tp;
      const E Evxip = vxip, Evyip = vyip, Evzip = vzip, Evzipm = vzipm, Etip = tip;
      const E Evxim = vxim, Evximp = vximp, Evyim = vyim, Evzim = vzim, Etim = tim;
      const E Evxjp = vxjp, Evyjp = vyjp, Evyjpm = vyjpm, Evzjp = vzjp, Etjp = tjp, Evyjpim = vyjpim;
      const E Evxjm = vxjm, Evxjmp = vxjmp, Evyjm = vyjm, Evzjm = vzjm, Etjm = tjm, Evzjmip = vzjmip;
      const E Eudz4 = c.udz4, Eudy4 = c.udy4, Eudx3c = udx3c, Eudx3m = udx3m;
      const E Eudzq_re = c.udzq_re, Eudyq_re = c.udyq_re, Eudzq_pe = c.udzq_pe, Eudyq_pe = c.udyq_pe;
      // ---- vx (ExplicitTermsVX.F90:55-101, ImplicitAndUpdateVX.F90:47-60)
      {
        E hxz = (((Evzip + Evzipm) * (Evxip + Evxc)) - ((Evzc + Evzm) * (Evxc + Evxim))) * Eudz4;
        E hxy = (((Evyjp + Evyjpm) * (Evxjp + Evxc)) - ((Evyc + Evym) * (Evxc + Evxjm))) * Eudy4;
        E aa = Evxp + Evxc, bb = Evxc + Evxm;
        E hxx = (aa * aa - bb * bb) * Eudx3c * (E)0.25;
        E dzzvx = (Evxim - (E)2.0 * Evxc + Evxip) * Eudzq_re;
        E dyyvx = (Evxjm - (E)2.0 * Evxc + Evxjp) * Eudyq_re;
        double q = -(hxx + hxy + hxz) + dyyvx + dzzvx + Etc;
        double dxxvx = vxp * apc + vxc * acc + vxm * amc;
        double dxp = (prc - prm) * (c.al * udx3c);
        double r = USE_RU ? c.ro * rxo : 0.0;
        double rhs = (c.ga * q + r + c.alre * dxxvx - dxp) * c.dt;
        double vin = rhs * cfv(SC_XSC) + ((cfv(SC_XDL) * vxm + vxc) + cfv(SC_XDU) * vxp);
        double vbd = stale0 + vxc;
        if (keep_ru) sth_nt<H32>(rux, C + sb, pad ? 0.0 : (inx ? q : 0.0));
        stf_nt(bx, C + sb, pad ? 0.0 : (inx ? vin : vbd));
      }
      // ---- vy (ExplicitTermsVY.F90:52-98, ImplicitAndUpdateVY.F90:52-68); kmv/kpv clamps
      {
        const double vymm = k == 0 ? vyc : vym, vypp = k == nxm - 1 ? vyc : vyp;
        const E Evymm = vymm, Evypp = vypp;
        E hyx = ((Evxp + Evxjmp) * (Evypp + Evyc) - (Evxc + Evxjm) * (Evyc + Evymm)) * Eudx3m * (E)0.25;
        E yp = Evyjp + Evyc, ym = Evyjm + Evyc;
        E hyy = (yp * yp - ym * ym) * Eudy4;
        E hyz = ((Evyip + Evyc) * (Evzip + Evzjmip) - (Evyc + Evyim) * (Evzc + Evzjm)) * Eudz4;
        E dyyvy = (Evyjp - (E)2.0 * Evyc + Evyjm) * Eudyq_re;
        E dzzvy = (Evyip - (E)2.0 * Evyc + Evyim) * Eudzq_re;
        double h = -(hyx + hyy + hyz) + dyyvy + dzzvy;
        double dxxvy = vypp * aps + vyc * acs + vymm * ams;
        double dyp = (prc - prjm) * c.udy_al;
        double r = USE_RU ? c.ro * ryo : 0.0;
        double rhs = (c.ga * h + r + c.alre * dxxvy - dyp) * c.dt;
        double vin = rhs * cfv(SC_YSC) + ((cfv(SC_YDL) * vym + vyc) + cfv(SC_YDU) * vyp);
        if (keep_ru) sth_nt<H32>(ruy, C + sb, pad ? 0.0 : (iny ? h : 0.0));
        stf_nt(by, C + sb, pad ? 0.0 : vin);   // row nx-1 is outside the nxm system and never read
      }
      // ---- vz (ExplicitTermsVZ.F90:56-101, ImplicitAndUpdateVZ.F90:48-64)
      {
        const double vzmm = k == 0 ? vzc : vzm, vzpp = k == nxm - 1 ? vzc : vzp;
        const E Evzmm = vzmm, Evzpp = vzpp;
        E zp = Evzip + Evzc, zm = Evzim + Evzc;
        E hzz = (zp * zp - zm * zm) * Eudz4;
        E hzy = ((Evyjp + Evyjpim) * (Evzjp + Evzc) - (Evyc + Evyim) * (Evzc + Evzjm)) * Eudy4;
        E hzx = ((Evxp + Evximp) * (Evzpp + Evzc) - (Evxc + Evxim) * (Evzc + Evzmm)) * Eudx3m * (E)0.25;
        E dzzvz = (Evzip - (E)2.0 * Evzc + Evzim) * Eudzq_re;
        E dyyvz = (Evzjp - (E)2.0 * Evzc + Evzjm) * Eudyq_re;
        double h = -(hzx + hzy + hzz) + dyyvz + dzzvz;
        double dxxvz = vzpp * aps + vzc * acs + vzmm * ams;
        double dzp = (prc - prim) * c.dz * c.al;
        double r = USE_RU ? c.ro * rzo : 0.0;
        double rhs = (c.ga * h + r + c.alre * dxxvz - dzp) * c.dt;
        double vin = rhs * cfv(SC_YSC) + ((cfv(SC_YDL) * vzm + vzc) + cfv(SC_YDU) * vzp);
        if (keep_ru) sth_nt<H32>(ruz, C + sb, pad ? 0.0 : (iny ? h : 0.0));
        stf_nt(bz, C + sb, pad ? 0.0 : vin);
      }
      // ---- temperature (ExplicitTermsTemp.F90:52-93, ImplicitAndUpdateTemp.F90:41-54)
      {
        E htz = ((Evzipm + Evzip) * (Etip + Etc) - (Evzm + Evzc) * (Etc + Etim)) * Eudz4;
        E hty = ((Evyjp + Evyjpm) * (Etjp + Etc) - (Evyc + Evym) * (Etc + Etjm)) * Eudy4;
        E htx = ((Evxp + Evxc) * (Etp + Etc) - (Evxc + Evxm) * (Etc + Etm)) * Eudx3c * (E)0.25;
        E dzzt = (Etip - (E)2.0 * Etc + Etim) * Eudzq_pe;
        E dyyt = (Etjp - (E)2.0 * Etc + Etjm) * Eudyq_pe;
        double h = -(htx + hty + htz) + dyyt + dzzt;
        double dxxt = tp * apc + tc * acc + tm * amc;
        double r = USE_RU ? c.ro * rto : 0.0;
        double rhs = (c.ga * h + r + c.alpec * dxxt) * c.dt;
        double vin = rhs * cfv(SC_TSC) + ((cfv(SC_TDL) * tm + tc) + cfv(SC_TDU) * tp);
        if (keep_ru) sth_nt<H32>(rut, C + sb, pad ? 0.0 : (inx ? h : 0.0));
        stf_nt(bt, C + sb, pad ? 0.0 : (inx ? vin : tc));   // row 1 gets the vz increment added by the solve pass
      }
    }
    vxim = vxc; vximp = vxp; vyim = vyc; vzim = vzc; tim = tc; prim = prc;
    vyjpim = vyjp;
    cur = nxt;
  }
}

// Lane p's chunk of one column's b' into registers (zeros past cnt).  The
// loads are unconditional (slot index clamped into the column's L rows) so
// that all LM of them issue back to back; the select discards the extras.
template <int LM, typename OffT>
__device__ __forceinline__ void load_chunk(double (&b)[LM], const double *__restrict__ src, OffT C, int cnt,
                                           int L, int p) {
#pragma unroll
  for (int t = 0; t < LM; t++) {
    const double v = ldf(src, C + (OffT)(min(t, L - 1) * P + p) * sizeof(double));
    b[t] = t < cnt ? v : 0.0;
  }
}

// Solves the chunk held in b in place; returns cin, with the solution
// x = b + H cin.  `rows` holds the system's factor rows L, RINV, E, G, H
// (row r at rows + r*ldx), staged in LDS by the caller.
enum { RW_L = 0, RW_RINV, RW_E, RW_G, RW_H, RW_N };
template <int LM>
__device__ __forceinline__ double solve_chunk(double (&b)[LM], const double *rows, int ldx, int cnt, int p) {
  const double *fl = rows + RW_L * ldx, *fr = rows + RW_RINV * ldx, *fe = rows + RW_E * ldx;
  const double *fg = rows + RW_G * ldx, *fh = rows + RW_H * ldx;
  double acc = 0.0;
#pragma unroll
  for (int t = 0; t < LM; t++) {
    if (t < cnt) {
      acc = b[t] - fl[t * P + p] * acc;
      b[t] = acc;
    }
  }
  const double cin = scan_up(cnt > 0 ? fg[(cnt - 1) * P + p] : 1.0, cnt > 0 ? acc : 0.0, p);
  acc = 0.0;
#pragma unroll
  for (int t = LM - 1; t >= 0; t--) {
    if (t < cnt) {
      const int s = t * P + p;
      const double y = b[t] + fg[s] * cin;
      acc = y * fr[s] - fe[s] * acc;
      b[t] = acc;
    }
  }
  return scan_down(cnt > 0 ? fh[p] : 1.0, cnt > 0 ? acc : 0.0, p);
}

constexpr int SOLVE_WG = 512;                 // 8 waves, 32 columns per block
constexpr int SOLVE_COLS = SOLVE_WG / P;

// LDS bytes of k_column_solve: rows L..H of the three systems.
static size_t solve_lds(const Geom &g) { return sizeof(double) * SYS_N * RW_N * (size_t)g.ldx; }

template <int LM, typename OffT>
__global__ void __launch_bounds__(SOLVE_WG)
    k_column_solve(FusedArgs a, const double *vz, const double *bx, const double *by, const double *bz,
                   const double *__restrict__ bt, double *ox, double *oy, double *oz, double *__restrict__ tn,
                   double *__restrict__ stalep, const double *__restrict__ facp) {
  extern __shared__ double rows[];   // [SYS_N][RW_N][ldx]
  const Geom g = a.g;
  const int64_t blk = xcd_block((g.ncol + SOLVE_COLS - 1) / SOLVE_COLS);
  if (blk < 0 || (a.ds && coh_int(&a.ds->halted))) return;
  const int L = g.L, ldx = g.ldx, nx = g.nx, nxm = g.nxm;
  // the factor rows are the same for every column: stage them once per block
  for (int idx = threadIdx.x; idx < SYS_N * RW_N * ldx; idx += SOLVE_WG) {
    const int sr = idx / ldx, s = idx - sr * ldx;
    const int sys = sr / RW_N, r = sr - sys * RW_N;
    rows[idx] = facp[(sys * FR_N + FR_L + r) * ldx + s];
  }
  __syncthreads();
  const int p = threadIdx.x % P;
  const int64_t col = blk * SOLVE_COLS + threadIdx.x / P;
  const bool live = col < g.ncol;
  const int64_t colc = live ? col : 0;
  const OffT C = (OffT)colc * (OffT)g.ldx * sizeof(double);
  const int k0 = p * L;
  const int cntx = live ? max(0, min(L, nx - k0)) : 0;
  const int cntm = live ? max(0, min(L, nxm - k0)) : 0;
  const double *rX = rows + (SYS_X * RW_N) * ldx, *rY = rows + (SYS_YZ * RW_N) * ldx,
               *rT = rows + (SYS_T * RW_N) * ldx;
  // Row nx-1 of vy and vz is outside the nxm systems and keeps the field's
  // value; it is stored again with the rest, so that every 128-byte line of a
  // column is written whole (one broadcast load per column).
  const OffT Ctop = C + (OffT)pos_of_k(nx - 1, L) * sizeof(double);
  const double vytop = ldf(a.f.vy, Ctop), vztop = ldf(vz, Ctop);
  double b[LM];
  load_chunk<LM, OffT>(b, bx, C, cntx, L, p);
  {  // vx
    const double *fh = rX + RW_H * ldx;
    const double cin = solve_chunk<LM>(b, rX, ldx, cntx, p);
    if (live)
#pragma unroll
      for (int t = 0; t < LM; t++)
        if (t < L) {   // slots of k >= nx (t >= cntx) are written 0
          const int k = k0 + t, s = t * P + p;
          stf_nt(ox, C + (OffT)s * sizeof(double), (k >= 1 && k <= nx - 2) ? b[t] + fh[s] * cin : 0.0);
        }
  }
  load_chunk<LM, OffT>(b, by, C, cntm, L, p);
  {  // vy
    const double *fh = rY + RW_H * ldx;
    const double cin = solve_chunk<LM>(b, rY, ldx, cntm, p);
    if (live)
#pragma unroll
      for (int t = 0; t < LM; t++)
        if (t < L) {
          const int k = k0 + t, s = t * P + p;
          stf_nt(oy, C + (OffT)s * sizeof(double), k == nx - 1 ? vytop : (k < nxm ? b[t] + fh[s] * cin : 0.0));
        }
  }
  load_chunk<LM, OffT>(b, bz, C, cntm, L, p);
  const double vz0 = (live && p == 0) ? ldf(vz, C) : 0.0;   // position of k = 0 is slot 0
  double dvz0 = 0.0;
  {  // vz
    const double *fh = rY + RW_H * ldx;
    const double cin = solve_chunk<LM>(b, rY, ldx, cntm, p);
    dvz0 = b[0] + fh[p] * cin - vz0;   // used from lane p = 0 (k = 0) only
    if (live)
#pragma unroll
      for (int t = 0; t < LM; t++)
        if (t < L) {
          const int k = k0 + t, s = t * P + p;
          stf_nt(oz, C + (OffT)s * sizeof(double), k == nx - 1 ? vztop : (k < nxm ? b[t] + fh[s] * cin : 0.0));
        }
  }
  dvz0 = __shfl(dvz0, 0, P);
  load_chunk<LM, OffT>(b, bt, C, cntx, L, p);
  if (live && p == 0) {
    stalep[colc] = dvz0;
    b[0] += dvz0;   // temperature row 1: rhs(1) = vz increment (lane 0 holds k = 0)
  }
  {  // temperature
    const double *fh = rT + RW_H * ldx;
    const double cin = solve_chunk<LM>(b, rT, ldx, cntx, p);
    if (live)
#pragma unroll
      for (int t = 0; t < LM; t++)
        if (t < L) {
          const int k = k0 + t, s = t * P + p;
          double y = b[t] + fh[s] * cin;
          y = k == 0 ? a.sc.tempbp : y;
          y = k == nx - 1 ? a.sc.temptp : y;
          stf_nt(tn, C + (OffT)s * sizeof(double), k < nx ? y : 0.0);
        }
  }
}

// ---------------------------------------------------------------- per-routine kernels
// One thread per stored point; k recovered from the interleaved position.
struct PointIdx {
  int64_t col;
  int s, k;
  bool live;
};
__device__ __forceinline__ PointIdx point_idx(const Geom &g) {
  PointIdx q;
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  q.col = gid / g.ldx;
  q.s = (int)(gid % g.ldx);
  q.k = k_of_pos(q.s, g.L);
  q.live = q.col < g.ncol;
  return q;
}

__global__ void k_explicit(RoutineArgs a, int which) {
  const Geom g = a.g;
  PointIdx q = point_idx(g);
  if (!q.live) return;
  const int j = (int)(q.col % g.nym), i = (int)(q.col / g.nym);
  const Nb n = make_nb(j, i, g);
  const int k = q.k, s = q.s, L = g.L, ldx = g.ldx;
  const double *kc = a.kc;
  if (which == 0 || which == 3) {  // vx, temp: k = 1..nx-2
    if (k < 1 || k > g.nx - 2) return;
    int skm = pos_of_k(k - 1, L), skp = pos_of_k(k + 1, L);
    if (which == 0) a.out[n.C + s] = expl_vx(a.f, n, s, skm, skp, kc[KC_UDX3C * ldx + s], a.sc);
    else a.out[n.C + s] = expl_temp(a.f, n, s, skm, skp, kc[KC_UDX3C * ldx + s], a.sc);
  } else {  // vy, vz: k = 0..nxm-1
    if (k > g.nxm - 1) return;
    int skp = pos_of_k(k + 1, L);
    int skmm = k == 0 ? s : pos_of_k(k - 1, L);
    int skpp = k == g.nxm - 1 ? s : skp;
    if (which == 1) a.out[n.C + s] = expl_vy(a.f, n, s, skp, skmm, skpp, kc[KC_UDX3M * ldx + s], a.sc);
    else a.out[n.C + s] = expl_vz(a.f, n, s, skp, skmm, skpp, kc[KC_UDX3M * ldx + s], a.sc);
  }
}

// ImplicitAndUpdate* RHS loops: rhs (unscaled) and history ru = explicit term.
__global__ void k_rhs(RoutineArgs a, int which) {
  const Geom g = a.g;
  PointIdx q = point_idx(g);
  if (!q.live) return;
  const int j = (int)(q.col % g.nym), i = (int)(q.col / g.nym);
  const Nb n = make_nb(j, i, g);
  const int k = q.k, s = q.s, L = g.L, ldx = g.ldx;
  const double *kc = a.kc;
  const int64_t o = n.C + s;
  if (which == 0 || which == 3) {
    if (k < 1 || k > g.nx - 2) return;
    int skm = pos_of_k(k - 1, L), skp = pos_of_k(k + 1, L);
    double qv = a.expl[o];
    double ru = a.sc.use_ru ? a.ru[o] : 0.0;
    a.rhs[o] = which == 0 ? rhs_vx(qv, ru, a.f.vx, a.f.pr, n, s, skm, skp, kc, ldx, a.sc)
                          : rhs_temp(qv, ru, a.f.temp, n, s, skm, skp, kc, ldx, a.sc);
    a.ru[o] = qv;
  } else {
    if (k > g.nxm - 1) return;
    int skp = pos_of_k(k + 1, L);
    int skmm = k == 0 ? s : pos_of_k(k - 1, L);
    int skpp = k == g.nxm - 1 ? s : skp;
    double qv = a.expl[o];
    double ru = a.sc.use_ru ? a.ru[o] : 0.0;
    a.rhs[o] = which == 1 ? rhs_vy(qv, ru, a.f.vy, a.f.pr, n, s, skmm, skpp, kc, ldx, a.sc)
                          : rhs_vz(qv, ru, a.f.vz, a.f.pr, n, s, skmm, skpp, kc, ldx, a.sc);
    a.ru[o] = qv;
  }
}

// SolveImpEqnUpdate_{X,YZ,Temp}: scale rhs rows, solve in place (rhs receives
// the increments, as dgttrs overwrites B), add to the field, apply the
// ImplicitAndUpdate* boundary rows.  which: 0 vx, 1 vy, 2 vz, 3 temp.
__global__ void __launch_bounds__(WG) k_solve_update(SolveArgs a, int which) {
  extern __shared__ double lds[];
  const Geom g = a.g;
  const int64_t blk = xcd_block((g.ncol + COLS_PER_WG - 1) / COLS_PER_WG);
  if (blk < 0) return;
  const int p = threadIdx.x % P;
  const int64_t col = blk * COLS_PER_WG + threadIdx.x / P;
  const bool live = col < g.ncol;
  const int64_t colc = live ? col : 0;
  const int64_t C = colc * g.ldx;
  const int L = g.L, ldx = g.ldx, nx = g.nx, nxm = g.nxm;
  const bool yz = which == 1 || which == 2;
  const int nsys = yz ? nxm : nx;
  const int k0 = p * L;
  const int cnt = live ? max(0, min(L, nsys - k0)) : 0;
  const double *fac = a.fac;
  double *bl = lds + threadIdx.x;
  for (int t = 0; t < cnt; t++) {
    const int k = k0 + t, s = t * P + p;
    double v = 0.0;
    if (yz || (k >= 1 && k <= nx - 2)) v = a.rhs[C + s] * fac[FR_SCALE * ldx + s];
    else if (k == 0) v = a.stale[colc];
    else v = a.rhs[C + s];  // rhs(nx): identity row, never written by the reference routines
    bl[t * WG] = v;
  }
  const double cin = chunk_solve(bl, fac, ldx, cnt, p);
  const double *fh = fac + FR_H * ldx;
  double x0 = 0.0;
  for (int t = 0; t < cnt; t++) {
    const int k = k0 + t, s = t * P + p;
    const double x = bl[t * WG] + fh[s] * cin;
    if (k == 0) x0 = x;
    if (yz) {
      a.rhs[C + s] = x;
      a.q[C + s] = a.q[C + s] + x;
    } else {
      if (k <= nx - 2) a.rhs[C + s] = x;
      double v = a.q[C + s];
      if (k >= 1 && k <= nx - 2) v = v + x;
      if (which == 0 && (k == 0 || k == nx - 1)) v = 0.0;
      if (which == 3 && k == 0) v = a.tempbp;
      if (which == 3 && k == nx - 1) v = a.temptp;
      a.q[C + s] = v;
    }
  }
  if (yz && live && p == 0) a.stale[colc] = x0;
}

// CalcLocalDivergence.F90:27-41 -> dph (k = 0..nxm-1)
__global__ void k_divergence(RoutineArgs a) {
  const Geom g = a.g;
  PointIdx q = point_idx(g);
  if (!q.live || q.k > g.nxm - 1) return;
  const int j = (int)(q.col % g.nym), i = (int)(q.col / g.nym);
  const int64_t C = col_off(j, i, g), CJP = col_off(wrapp(j + 1, g.nym), i, g),
                CIP = col_off(j, wrapp(i + 1, g.nzm), g);
  const int s = q.s, skp = pos_of_k(q.k + 1, g.L);
  const double usdtal = 1.0 / (a.sc.dt * a.sc.al);
  double dqcap = (a.f.vz[CIP + s] - a.f.vz[C + s]) * a.sc.dz + (a.f.vy[CJP + s] - a.f.vy[C + s]) * a.sc.dy +
                 (a.f.vx[C + skp] - a.f.vx[C + s]) * a.kc[KC_UDX3M * g.ldx + s];
  a.out[C + s] = dqcap * usdtal;
}

// CorrectVelocity.F90:28-45 and CorrectPressure.F90:32-52 from dph (= dphhalo).
// which: 1 velocity, 2 pressure, 3 both (the substep path).  The velocities
// are read from (ux, uy, uz) and written to (vx, vy, vz): the u* buffers of
// the substep path, or the fields themselves (in place) for the per-routine
// entries and the refinement sweep.  The block shape is that of
// k_stencil_rhs: ST_GC neighbouring y columns at the same ST_SP stored
// positions, marching through zplanes z planes with dph of planes i-1, i, i+1
// in registers; each thread publishes its plane-i dph in LDS for its y
// neighbours (double-buffered, one barrier per plane), so only the group's
// edge columns load a y neighbour.
//
// CFL = true (the last correction of a device step, not in place) also
// reduces CalcMaxCFL.F90:26-41 over the corrected velocities into
// a.cflbits (the bits of a non-negative double, so an integer atomic max is
// the floating-point max).  The corrected value of each neighbour a point
// needs (vx at k+1, vy at j+1, vz at i+1) is recomputed from u* and dph by
// the same fused multiply-add that stores it, so the maximum equals that of
// k_red_cfl over the stored fields bit for bit.
constexpr int CZ_ROWS = 6;   // udx3c, amphk, acphk, apphk, udx3m, udx3c at k+1
__device__ __forceinline__ double corr(double u, double dd, double f) { return __builtin_fma(-dd, f, u); }
template <bool CFL>
__global__ void __launch_bounds__(SY_WG) k_correct(CorrectArgs a, int which, int zplanes) {
  extern __shared__ double sh[];
  double *cr = sh;                    // [CZ_ROWS][ST_SP]
  double *xv = sh + CZ_ROWS * ST_SP;  // [2][2][SY_WG] published dph (and u*_y) of plane i
  const Geom g = a.g;
  const int L = g.L, ldx = g.ldx, nxm = g.nxm;
  const int nq = (ldx + ST_SP - 1) / ST_SP;
  const int ngr = (g.nym + ST_GC - 1) / ST_GC;
  const int nib = (g.nzm + zplanes - 1) / zplanes;
  const int64_t tile = xcd_block((int64_t)nib * ngr * nq);
  if (tile < 0 || (a.ds && coh_int(&a.ds->halted))) return;
  const int q = (int)(tile % nq);
  const int64_t rest = tile / nq;
  const int gr = (int)(rest % ngr), ib = (int)(rest / ngr);
  const StepScalars c = a.ds ? coh_scalars(&a.ds->sc[a.ns - 1]) : a.sc;
  for (int idx = threadIdx.x; idx < CZ_ROWS * ST_SP; idx += SY_WG) {
    const int r = idx / ST_SP, sl = min(q * ST_SP + idx - r * ST_SP, ldx - 1);
    int sp = sl;
    if (r == 5) {   // the position of k+1 (k+1 <= nx-1 holds for every stored k < nx-1)
      const int kk = k_of_pos(sl, L);
      sp = kk + 1 < g.nx ? pos_of_k(kk + 1, L) : sl;
    }
    const int src = r == 0 || r == 5 ? KC_UDX3C
                    : (r == 1 ? KC_AMPHK : (r == 2 ? KC_ACPHK : (r == 3 ? KC_APPHK : KC_UDX3M)));
    cr[idx] = a.kc[src * ldx + sp];
  }
  __syncthreads();
  const int lane = (int)threadIdx.x & (ST_SP - 1);   // position within the tile
  const int w = (int)threadIdx.x / ST_SP;             // column within the group
  const int ncols = min(ST_GC, g.nym - gr * ST_GC);
  const bool colok = w < ncols;
  const int jc = colok ? gr * ST_GC + w : g.nym - 1;
  const int sraw = q * ST_SP + lane;
  const int s = min(sraw, ldx - 1);
  const int k = k_of_pos(s, L);
  const bool live = colok && sraw < ldx && k <= nxm - 1;
  // Positions k >= nxm (the wall row of vx, where u* is 0, row nx-1 of the
  // other fields, and the padding) keep what u* and pr hold there; they are
  // stored too, so that every 128-byte line of a column is written whole.
  const bool stv = colok && sraw < ldx;
  const bool jp_in = w + 1 < ncols, jm_in = w > 0;
  const int jm = wrapm(jc - 1, g.nym), jp = wrapp(jc + 1, g.nym);
  const int skm = (k == 0 || k > nxm - 1) ? s : pos_of_k(k - 1, L);
  const int skp = (k >= nxm - 1) ? s : pos_of_k(k + 1, L);
  const int sx1 = k + 1 < g.nx ? pos_of_k(k + 1, L) : s;   // vx at k+1 (the wall row for k = nxm-1)
  const double *__restrict__ dph = a.dph;
  const double *ux = a.ux, *uy = a.uy, *uz = a.uz;
  double *qx = a.vx, *qy = a.vy, *qz = a.vz, *__restrict__ qp = a.pr;
  const double usukm = c.al * c.dt * cr[lane];   // row r of cr at cr[r * ST_SP + lane]
  const double udy = c.al * c.dt * c.dy, udz = c.al * c.dt * c.dz;
  const double be = c.al * c.beta;
  const double amm = cr[ST_SP + lane], acc = cr[2 * ST_SP + lane], app = cr[3 * ST_SP + lane];
  const double u3m = cr[4 * ST_SP + lane], usukp = c.al * c.dt * cr[5 * ST_SP + lane];
  const int i0 = ib * zplanes, i1 = min(i0 + zplanes, g.nzm);
  double dim_ = dph[col_off(jc, wrapm(i0 - 1, g.nzm), g) + s];
  double d = dph[col_off(jc, i0, g) + s];
  double uzc = CFL ? uz[col_off(jc, i0, g) + s] : 0.0;   // u*_z of the current plane, carried
  const int tid = (int)threadIdx.x;
  int par = 0;
  double cflm = 0.0;
  for (int ic = i0; ic < i1; ic++) {
    const int64_t C = col_off(jc, ic, g);
    const int64_t CIP = col_off(jc, wrapp(ic + 1, g.nzm), g);
    const double dip = dph[CIP + s];
    const double dkm = dph[C + skm], dkp = dph[C + skp];
    double djm = 0.0, djp = 0.0;
    if (!jm_in) djm = dph[col_off(jm, ic, g) + s];
    if (!jp_in) djp = dph[col_off(jp, ic, g) + s];
    double vxo = 0.0, vyo = 0.0, vzo = 0.0, pro = 0.0;
    double uxk1 = 0.0, uyjp = 0.0, uzip = 0.0;
    if (which & 1) {   // read here only: streamed past the caches (2 % less fetched, 2 % faster)
      vxo = __builtin_nontemporal_load(&ux[C + s]);
      vyo = __builtin_nontemporal_load(&uy[C + s]);
      vzo = CFL ? uzc : __builtin_nontemporal_load(&uz[C + s]);
    }
    if (CFL) {
      uxk1 = ux[C + sx1];
      uzip = uz[CIP + s];
      if (!jp_in) uyjp = uy[col_off(jp, ic, g) + s];
    }
    if (which & 2) pro = __builtin_nontemporal_load(&qp[C + s]);
    double *buf = xv + par * (2 * SY_WG);
    buf[tid] = d;
    if (CFL) buf[SY_WG + tid] = vyo;
    __syncthreads();
    if (jm_in) djm = buf[tid - ST_SP];
    if (jp_in) djp = buf[tid + ST_SP];
    if (CFL && jp_in) uyjp = buf[SY_WG + tid + ST_SP];
    par ^= 1;
    if (stv) {
      if (which & 1) {
        const double nvx = corr(vxo, d - dkm, usukm);
        const double nvy = corr(vyo, d - djm, udy);
        const double nvz = corr(vzo, d - dim_, udz);
        qx[C + s] = live ? nvx : vxo;
        qy[C + s] = live ? nvy : vyo;
        qz[C + s] = live ? nvz : vzo;
        if (CFL && live) {
          const double nvxk1 = k + 1 <= nxm - 1 ? corr(uxk1, dkp - d, usukp) : uxk1;
          const double nvyjp = corr(uyjp, djp - d, udy);
          const double nvzip = corr(uzip, dip - d, udz);
          const double qcf = (fabs((nvz + nvzip) * 0.5 * c.dz) + fabs((nvy + nvyjp) * 0.5 * c.dy) +
                              fabs((nvx + nvxk1) * 0.5 * u3m));
          cflm = fmax(cflm, qcf);
        }
      }
      if (which & 2)
        qp[C + s] = live ? pro + d - be * ((dip - 2.0 * d + dim_) * c.dzq + (djp - 2.0 * d + djm) * c.dyq +
                                           (dkp * app + d * acc + dkm * amm))
                         : pro;
    }
    dim_ = d;
    d = dip;
    uzc = uzip;
  }
  if (CFL) {   // one atomic per block
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cflm = fmax(cflm, __shfl_xor(cflm, off, 64));
    __syncthreads();   // every wave is past its last read of xv
    if ((threadIdx.x & 63) == 0) xv[threadIdx.x >> 6] = cflm;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = xv[0];
      for (int q = 1; q < SY_WG / 64; q++) m = fmax(m, xv[q]);
      atomicMax(a.cflbits, (unsigned long long)__double_as_longlong(m));
    }
  }
}

// ---------------------------------------------------------------- launchers
static size_t column_lds(const Geom &g) { return sizeof(double) * WG * (size_t)g.L; }

#define AFID_LM_LIST(X) X(2) X(4) X(8) X(12) X(16) X(25) X(32) X(48) X(64) X(80)

static int pick_lm(int L) {
  const int lms[] = {2, 4, 8, 12, 16, 25, 32, 48, 64, 80};
  for (int v : lms)
    if (L <= v) return v;
  return -1;
}

hipError_t column_configure_lds(size_t bytes) {
  hipError_t e = hipSuccess;
#define SETB(V)                                                                                          \
  e = hipFuncSetAttribute((const void *)k_column_solve<V, uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                          (int)bytes);                                                                   \
  if (e != hipSuccess) return e;                                                                         \
  e = hipFuncSetAttribute((const void *)k_column_solve<V, uint64_t>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                          (int)bytes);                                                                   \
  if (e != hipSuccess) return e;
  AFID_LM_LIST(SETB)
#undef SETB
#define SETC1(O, R, D, H)                                                                                \
  e = hipFuncSetAttribute((const void *)k_stencil_rhs<O, R, D, H>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                          (int)bytes);                                                                   \
  if (e != hipSuccess) return e;
#define SETC(O, R) SETC1(O, R, true, false) SETC1(O, R, false, false) SETC1(O, R, true, true) SETC1(O, R, false, true)
  SETC(uint32_t, true) SETC(uint32_t, false) SETC(uint64_t, true) SETC(uint64_t, false)
#undef SETC
#undef SETC1
  return hipSuccess;
}

static dim3 column_grid(const Geom &g) {
  int64_t nblk = (g.ncol + COLS_PER_WG - 1) / COLS_PER_WG;
  return dim3((unsigned)(((nblk + 7) / 8) * 8));
}

hipError_t launch_stencil_rhs(const FusedArgs &a, hipStream_t st) {
  const int64_t n = a.g.ncol * a.g.ldx;
  const int zplanes = a.zm > 0 ? a.zm : ZM_DEFAULT;
  const int64_t ntile = (int64_t)((a.g.nzm + zplanes - 1) / zplanes) * ((a.g.nym + ST_GC - 1) / ST_GC) *
                        ((a.g.ldx + ST_SP - 1) / ST_SP);
  const dim3 grid((unsigned)(((ntile + 7) / 8) * 8));
  const size_t shm = stencil_lds(a.g);
  const bool small = (uint64_t)n * sizeof(double) < (1ull << 32);
  const bool ru = a.sc.use_ru != 0, dev = a.ds != nullptr, h32 = a.h32 != 0;
#define SR1(O, R, D, H)                                                                                  \
  if (std::is_same<O, uint32_t>::value == small && R == ru && D == dev && H == h32) {                  \
    hipLaunchKernelGGL((k_stencil_rhs<O, R, D, H>), grid, dim3(SY_WG), shm, st, a, a.f.vx, a.f.vy,       \
                       a.f.vz, a.f.temp, a.f.pr, a.rux, a.ruy, a.ruz, a.rut, a.bx, a.by, a.bz, a.bt, a.stale, \
                       a.kc, a.fac);                                                                   \
    return hipGetLastError();                                                                          \
  }
#define SR(O, R) SR1(O, R, true, false) SR1(O, R, false, false) SR1(O, R, true, true) SR1(O, R, false, true)
  SR(uint32_t, true) SR(uint32_t, false) SR(uint64_t, true) SR(uint64_t, false)
#undef SR
#undef SR1
  return hipErrorInvalidValue;
}

hipError_t launch_column_solve(const FusedArgs &a, hipStream_t st) {
  const bool small = (uint64_t)a.g.ncol * a.g.ldx * sizeof(double) < (1ull << 32);
  const int64_t nblk = (a.g.ncol + SOLVE_COLS - 1) / SOLVE_COLS;
  const dim3 grid((unsigned)(((nblk + 7) / 8) * 8));
  const size_t shm = solve_lds(a.g);
  switch (pick_lm(a.g.L)) {
#define CASE(V)                                                                                        \
  case V:                                                                                              \
    if (small)                                                                                         \
      hipLaunchKernelGGL((k_column_solve<V, uint32_t>), grid, dim3(SOLVE_WG), shm, st, a, a.f.vz, a.bx, \
                         a.by, a.bz, a.bt, a.ux, a.uy, a.uz, a.tn, a.stale, a.fac);                      \
    else                                                                                               \
      hipLaunchKernelGGL((k_column_solve<V, uint64_t>), grid, dim3(SOLVE_WG), shm, st, a, a.f.vz, a.bx, \
                         a.by, a.bz, a.bt, a.ux, a.uy, a.uz, a.tn, a.stale, a.fac);                      \
    return hipGetLastError();
    AFID_LM_LIST(CASE)
#undef CASE
    default: break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_solve_update(const SolveArgs &a, int which, hipStream_t st) {
  hipLaunchKernelGGL(k_solve_update, column_grid(a.g), dim3(WG), column_lds(a.g), st, a, which);
  return hipGetLastError();
}

static dim3 point_grid(const Geom &g) {
  int64_t n = g.ncol * g.ldx;
  return dim3((unsigned)((n + 255) / 256));
}

hipError_t launch_explicit(const RoutineArgs &a, int which, hipStream_t st) {
  hipLaunchKernelGGL(k_explicit, point_grid(a.g), dim3(256), 0, st, a, which);
  return hipGetLastError();
}
hipError_t launch_rhs(const RoutineArgs &a, int which, hipStream_t st) {
  hipLaunchKernelGGL(k_rhs, point_grid(a.g), dim3(256), 0, st, a, which);
  return hipGetLastError();
}
hipError_t launch_divergence(const RoutineArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_divergence, point_grid(a.g), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_correct(const CorrectArgs &a, int which, hipStream_t st) {
  const int zplanes = ZM_DEFAULT;
  const int64_t nt = (int64_t)((a.g.nzm + zplanes - 1) / zplanes) * ((a.g.nym + ST_GC - 1) / ST_GC) *
                     ((a.g.ldx + ST_SP - 1) / ST_SP);
  const size_t shm = sizeof(double) * (CZ_ROWS * ST_SP + 2 * 2 * SY_WG);
  const dim3 grid((unsigned)(((nt + 7) / 8) * 8));
  if (a.cflbits) {
    if (which != 3 || a.ux == a.vx) return hipErrorInvalidValue;   // the fold needs u* apart from the fields
    hipLaunchKernelGGL(k_correct<true>, grid, dim3(SY_WG), shm, st, a, which, zplanes);
  } else {
    hipLaunchKernelGGL(k_correct<false>, grid, dim3(SY_WG), shm, st, a, which, zplanes);
  }
  return hipGetLastError();
}

// LDS per block: WG*L doubles (per-routine solve), SC_N*ST_SP + 2*SV_N*SY_WG (stencil
// pass), SYS_N*RW_N*16*L (column solve); L <= 80 keeps all within gfx950's 160 KiB.
int max_supported_L() { return 80; }
size_t column_lds_need(const Geom &g) {
  size_t b = std::max(stencil_lds(g), solve_lds(g));
  return b;
}

}  // namespace afid
