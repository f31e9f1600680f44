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

// XCD-aware block order: blocks b, b+8, b+16 ... land on one XCD, so give
// each XCD one contiguous range of columns (neighbouring columns share L2).
__device__ __forceinline__ int64_t xcd_block(int64_t nblk) {
  int64_t b = blockIdx.x;
  int64_t per = (gridDim.x + 7) / 8;
  int64_t nb = (b % 8) * per + b / 8;
  return nb < nblk ? nb : -1;
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

// Positions of k-1 / k+1 for lane p, step t (valid only when those k exist).
__device__ __forceinline__ int lane_skm(int p, int t, int L) { return t > 0 ? (t - 1) * P + p : (L - 1) * P + p - 1; }
__device__ __forceinline__ int lane_skp(int p, int t, int L) { return t < L - 1 ? (t + 1) * P + p : p + 1; }

// ---------------------------------------------------------------- fused substep
// One pass per column over all four prognostic fields: explicit terms, RHS,
// x-implicit solve and update, with the new fields written to separate
// buffers (explicit terms of all fields read the old ones, TimeMarcher.F90:32-40).
// rhs(1,j,i) of the reference (stale plane) is read by the vx solve and
// replaced by the vz increment at k=1 (SolveImpEqnUpdate_YZ.F90:50), which
// the temperature solve of the same substep then reads.
__global__ void __launch_bounds__(WG) k_fused_update(FusedArgs a) {
  extern __shared__ double lds[];
  const Geom g = a.g;
  const int64_t blk = xcd_block((g.ncol + COLS_PER_WG - 1) / COLS_PER_WG);
  if (blk < 0) return;
  const int p = threadIdx.x % P;
  const int64_t col = blk * COLS_PER_WG + threadIdx.x / P;
  const bool live = col < g.ncol;
  const int64_t colc = live ? col : 0;
  const int j = (int)(colc % g.nym), i = (int)(colc / g.nym);
  const Nb n = make_nb(j, i, g);
  const int L = g.L, ldx = g.ldx, nx = g.nx, nxm = g.nxm;
  const StepScalars &c = a.sc;
  const double *kc = a.kc;
  double *bl = lds + threadIdx.x;
  const int k0 = p * L;
  const int cntx = live ? max(0, min(L, nx - k0)) : 0;    // points of an nx system
  const int cntm = live ? max(0, min(L, nxm - k0)) : 0;   // points of an nxm system

  // ---- vx: ImplicitAndUpdateVX + SolveImpEqnUpdate_X (n = nx, rows 1 and nx identity)
  {
    const double *fac = a.fac + (SYS_X * FR_N) * ldx;
    for (int t = 0; t < cntx; t++) {
      const int k = k0 + t, s = t * P + p;
      double v = 0.0;
      if (k >= 1 && k <= nx - 2) {
        const int skm = lane_skm(p, t, L), skp = lane_skp(p, t, L);
        double q = expl_vx(a.f, n, s, skm, skp, kc[KC_UDX3C * ldx + s], c);
        double ru = c.use_ru ? a.rux[n.C + s] : 0.0;
        v = rhs_vx(q, ru, a.f.vx, a.f.pr, n, s, skm, skp, kc, ldx, c) * fac[FR_SCALE * ldx + s];
        a.rux[n.C + s] = q;
      } else if (k == 0) {
        v = a.stale[colc];
      }
      bl[t * WG] = v;
    }
    const double cin = chunk_solve(bl, fac, ldx, cntx, p);
    const double *fh = fac + FR_H * ldx;
    for (int t = 0; t < cntx; t++) {
      const int k = k0 + t, s = t * P + p;
      const double x = bl[t * WG] + fh[s] * cin;
      a.vxn[n.C + s] = (k >= 1 && k <= nx - 2) ? a.f.vx[n.C + s] + x : 0.0;
    }
  }
  // ---- vy, vz: SolveImpEqnUpdate_YZ (n = nxm)
  double dvz0 = 0.0;
  for (int comp = 0; comp < 2; comp++) {
    const double *fac = a.fac + (SYS_YZ * FR_N) * ldx;
    const double *q0 = comp == 0 ? a.f.vy : a.f.vz;
    double *ru = comp == 0 ? a.ruy : a.ruz;
    double *qn = comp == 0 ? a.vyn : a.vzn;
    for (int t = 0; t < cntm; t++) {
      const int k = k0 + t, s = t * P + p;
      const int skp = lane_skp(p, t, L);
      const int skmm = k == 0 ? s : lane_skm(p, t, L);
      const int skpp = k == nxm - 1 ? s : skp;
      const double udx3m = kc[KC_UDX3M * ldx + s];
      const double r0 = c.use_ru ? ru[n.C + s] : 0.0;
      double h, v;
      if (comp == 0) {
        h = expl_vy(a.f, n, s, skp, skmm, skpp, udx3m, c);
        v = rhs_vy(h, r0, q0, a.f.pr, n, s, skmm, skpp, kc, ldx, c);
      } else {
        h = expl_vz(a.f, n, s, skp, skmm, skpp, udx3m, c);
        v = rhs_vz(h, r0, q0, a.f.pr, n, s, skmm, skpp, kc, ldx, c);
      }
      bl[t * WG] = v * fac[FR_SCALE * ldx + s];
      ru[n.C + s] = h;
    }
    const double cin = chunk_solve(bl, fac, ldx, cntm, p);
    const double *fh = fac + FR_H * ldx;
    for (int t = 0; t < cntx; t++) {
      const int k = k0 + t, s = t * P + p;
      if (k < nxm) {
        const double x = bl[t * WG] + fh[s] * cin;
        qn[n.C + s] = q0[n.C + s] + x;
        if (comp == 1 && k == 0) dvz0 = x;
      } else {
        qn[n.C + s] = 0.0;
      }
    }
  }
  dvz0 = __shfl(dvz0, 0, P);  // lane p=0 of the column holds k=0
  if (live && p == 0) a.stale[colc] = dvz0;
  // ---- temperature: SolveImpEqnUpdate_Temp (n = nx), rhs(1) = the vz increment at k=1
  {
    const double *fac = a.fac + (SYS_T * FR_N) * ldx;
    for (int t = 0; t < cntx; t++) {
      const int k = k0 + t, s = t * P + p;
      double v = 0.0;
      if (k >= 1 && k <= nx - 2) {
        const int skm = lane_skm(p, t, L), skp = lane_skp(p, t, L);
        double h = expl_temp(a.f, n, s, skm, skp, kc[KC_UDX3C * ldx + s], c);
        double ru = c.use_ru ? a.rut[n.C + s] : 0.0;
        v = rhs_temp(h, ru, a.f.temp, n, s, skm, skp, kc, ldx, c) * fac[FR_SCALE * ldx + s];
        a.rut[n.C + s] = h;
      } else if (k == 0) {
        v = dvz0;
      }
      bl[t * WG] = v;
    }
    const double cin = chunk_solve(bl, fac, ldx, cntx, p);
    const double *fh = fac + FR_H * ldx;
    for (int t = 0; t < cntx; t++) {
      const int k = k0 + t, s = t * P + p;
      double v = a.f.temp[n.C + s] + (bl[t * WG] + fh[s] * cin);
      if (k == 0) v = c.tempbp;
      if (k == nx - 1) v = c.temptp;
      a.tn[n.C + s] = v;
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
// which: 1 velocity, 2 pressure, 3 both (the fused path).
__global__ void k_correct(CorrectArgs a, int which) {
  const Geom g = a.g;
  PointIdx q = point_idx(g);
  if (!q.live || q.k > g.nxm - 1) return;
  const int j = (int)(q.col % g.nym), i = (int)(q.col / g.nym);
  const int jm = wrapm(j - 1, g.nym), jp = wrapp(j + 1, g.nym);
  const int im = wrapm(i - 1, g.nzm), ip = wrapp(i + 1, g.nzm);
  const int64_t C = col_off(j, i, g);
  const int s = q.s, k = q.k, ldx = g.ldx;
  const int skm = k == 0 ? s : pos_of_k(k - 1, g.L);
  const int skp = k == g.nxm - 1 ? s : pos_of_k(k + 1, g.L);
  const double *dph = a.dph;
  const double d = dph[C + s];
  const StepScalars &c = a.sc;
  if (which & 1) {
    double usukm = c.al * c.dt * a.kc[KC_UDX3C * ldx + s];
    double udy = c.al * c.dt * c.dy, udz = c.al * c.dt * c.dz;
    a.vx[C + s] = a.vx[C + s] - (d - dph[C + skm]) * usukm;
    a.vy[C + s] = a.vy[C + s] - (d - dph[col_off(jm, i, g) + s]) * udy;
    a.vz[C + s] = a.vz[C + s] - (d - dph[col_off(j, im, g) + s]) * udz;
  }
  if (which & 2) {
    double be = c.al * c.beta;
    double amm = a.kc[KC_AMPHK * ldx + s], acc = a.kc[KC_ACPHK * ldx + s], app = a.kc[KC_APPHK * ldx + s];
    a.pr[C + s] = a.pr[C + s] + d -
                  be * ((dph[col_off(j, ip, g) + s] - 2.0 * d + dph[col_off(j, im, g) + s]) * c.dzq +
                        (dph[col_off(jp, i, g) + s] - 2.0 * d + dph[col_off(jm, i, g) + s]) * c.dyq +
                        (dph[C + skp] * app + d * acc + dph[C + skm] * amm));
  }
}

// ---------------------------------------------------------------- launchers
static size_t column_lds(const Geom &g) { return sizeof(double) * WG * (size_t)g.L; }

hipError_t column_configure_lds(size_t bytes) {
  hipError_t e = hipFuncSetAttribute((const void *)k_fused_update, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void *)k_solve_update, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

static dim3 column_grid(const Geom &g) {
  int64_t nblk = (g.ncol + COLS_PER_WG - 1) / COLS_PER_WG;
  return dim3((unsigned)(((nblk + 7) / 8) * 8));
}

hipError_t launch_fused_update(const FusedArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_fused_update, column_grid(a.g), dim3(WG), column_lds(a.g), st, a);
  return hipGetLastError();
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
  hipLaunchKernelGGL(k_correct, point_grid(a.g), dim3(256), 0, st, a, which);
  return hipGetLastError();
}

// LDS holds WG*L doubles per block: L <= 80 keeps it within gfx950's 160 KiB.
int max_supported_L() { return 80; }

}  // namespace afid
