// misc_kernels.hip -- diagnostics reductions (CalcMaxCFL.F90, CheckDivergence.F90,
// CalcPlateNu.F90, GlobalQuantities.F90, CalcDissipationNu.F90), layout
// conversion between the reference's Fortran arrays and the interleaved
// device layout, and initial conditions.  All reductions are FP64: per-block
// partials in a fixed order, then one block combines them (deterministic).
#include "afid_internal.h"
#include "kernels.h"

namespace afid {

constexpr int RBLK = 1024;  // partial blocks
int reduce_partial_blocks() { return RBLK; }

template <typename Op>
__device__ double block_reduce(double v, Op op, double init) {
  __shared__ double sh[256];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = op(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  (void)init;
  return r;
}
struct OpMax { __device__ double operator()(double a, double b) const { return fmax(a, b); } };
struct OpMin { __device__ double operator()(double a, double b) const { return fmin(a, b); } };
struct OpSum { __device__ double operator()(double a, double b) const { return a + b; } };

// ops: 0 sum, 1 max, 2 min for each of nq quantities
__global__ void k_finalize(const double *partial, double *out, int nq, const int *ops_dummy, int op_mask_max,
                           int op_mask_min) {
  const int q = blockIdx.x;
  const bool ismax = (op_mask_max >> q) & 1, ismin = (op_mask_min >> q) & 1;
  double v = ismax ? -HUGE_VAL : (ismin ? HUGE_VAL : 0.0);
  for (int b = threadIdx.x; b < RBLK; b += blockDim.x) {
    double x = partial[(int64_t)b * nq + q];
    v = ismax ? fmax(v, x) : (ismin ? fmin(v, x) : v + x);
  }
  if (ismax) v = block_reduce(v, OpMax(), 0);
  else if (ismin) v = block_reduce(v, OpMin(), 0);
  else v = block_reduce(v, OpSum(), 0);
  if (threadIdx.x == 0) out[q] = v;
  (void)ops_dummy;
}

// Interior points (k < kmax) of all columns: each of the RBLK blocks walks a
// contiguous range of the strip column order (XCD-remapped, so y/z neighbour
// columns come from the same L2), its threads covering the stored positions.
template <typename F>
__device__ __forceinline__ void for_points(const Geom &g, int kmax, F &&f) {
  const int64_t cpb = (g.ncol + RBLK - 1) / RBLK;
  const int64_t b = xcd_block(RBLK);
  if (b < 0) return;
  const int64_t c1 = min((b + 1) * cpb, g.ncol);
  for (int64_t c = b * cpb; c < c1; c++) {
    int j, i;
    strip_col(c, g, j, i);
    const int64_t C = col_off(j, i, g);
    for (int s = threadIdx.x; s < g.ldx; s += blockDim.x) {
      const int k = k_of_pos(s, g.L);
      if (k < kmax) f(j, i, C, s, k);
    }
  }
}

// CalcMaxCFL (CalcMaxCFL.F90:26-41): four waves on four neighbouring y columns
// at the same 64 stored positions march through CFL_ZP z planes, with vz of
// plane i+1 carried in registers to the next plane and vy of column j+1 taken
// from the neighbouring wave through LDS (the group's last column loads it),
// so each field is read about once.  (The stencil pass's block of sixteen
// columns at one 128-byte line each ran slower here, 1.34-1.38 against
// 1.15 ms at 384x768x768: these blocks are persistent and walk their tiles
// one after another.)
// Each of the RBLK blocks walks a contiguous range of these tiles (XCD-
// remapped) and leaves one maximum; the terms are summed in the reference
// order, so the maximum equals that of the point-wise form bit for bit.
constexpr int CFL_COLS = 4, CFL_WG = CFL_COLS * 64, CFL_ZP = 32;
__global__ void __launch_bounds__(CFL_WG) k_red_cfl(ReduceArgs a) {
  __shared__ double xv[2][CFL_WG];
  const Geom g = a.g;
  const int L = g.L, ldx = g.ldx, nxm = g.nxm;
  const int nq = (ldx + 63) / 64, ngr = (g.nym + CFL_COLS - 1) / CFL_COLS;
  const int nib = (g.nzm + CFL_ZP - 1) / CFL_ZP;
  const int64_t ntile = (int64_t)nq * ngr * nib;
  const int64_t b = xcd_block(RBLK);   // the grid is RBLK blocks, a multiple of 8
  const int64_t per = (ntile + RBLK - 1) / RBLK;
  const int64_t t0 = b * per, t1 = min(t0 + per, ntile);
  const int lane = (int)threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  double m = 0.0;
  int par = 0;
  for (int64_t tile = t0; tile < t1; tile++) {
    const int q = (int)(tile % nq);
    const int64_t rest = tile / nq;
    const int gr = (int)(rest % ngr), ib = (int)(rest / ngr);
    const int ncols = min(CFL_COLS, g.nym - gr * CFL_COLS);
    const bool colok = w < ncols;
    const int jc = colok ? gr * CFL_COLS + w : g.nym - 1;
    const int sraw = q * 64 + lane, s = min(sraw, ldx - 1), k = k_of_pos(s, L);
    const bool live = colok && sraw < ldx && k < nxm;
    const int skp = pos_of_k(min(k + 1, g.nx - 1), L);
    const bool jp_in = w + 1 < ncols;
    const int jp = wrapp(jc + 1, g.nym);
    const double u3m = a.kc[KC_UDX3M * ldx + s];
    const int i0 = ib * CFL_ZP, i1 = min(i0 + CFL_ZP, g.nzm);
    double vzc = a.f.vz[col_off(jc, i0, g) + s];
    for (int ic = i0; ic < i1; ic++) {
      const int64_t C = col_off(jc, ic, g);
      const double vzip = a.f.vz[col_off(jc, wrapp(ic + 1, g.nzm), g) + s];
      const double vyc = a.f.vy[C + s];
      const double vxc = a.f.vx[C + s], vxk1 = a.f.vx[C + skp];
      double vyjp = jp_in ? 0.0 : a.f.vy[col_off(jp, ic, g) + s];
      xv[par][threadIdx.x] = vyc;
      __syncthreads();
      if (jp_in) vyjp = xv[par][threadIdx.x + 64];
      par ^= 1;
      const double qcf = (fabs((vzc + vzip) * 0.5 * a.dz) + fabs((vyc + vyjp) * 0.5 * a.dy) +
                          fabs((vxc + vxk1) * 0.5 * u3m));
      if (live) m = fmax(m, qcf);
      vzc = vzip;
    }
  }
  m = block_reduce(m, OpMax(), 0);
  if (threadIdx.x == 0) a.partial[blockIdx.x] = m;
}

__global__ void __launch_bounds__(256) k_red_div(ReduceArgs a) {
  const Geom g = a.g;
  double m = 0.0;
  for_points(g, g.nxm, [&](int j, int i, int64_t C, int s, int k) {
    const int64_t CJP = col_off(wrapp(j + 1, g.nym), i, g), CIP = col_off(j, wrapp(i + 1, g.nzm), g);
    const int skp = pos_of_k(k + 1, g.L);
    double dq = (a.f.vz[CIP + s] - a.f.vz[C + s]) * a.dz + (a.f.vy[CJP + s] - a.f.vy[C + s]) * a.dy +
                (a.f.vx[C + skp] - a.f.vx[C + s]) * a.kc[KC_UDX3M * g.ldx + s];
    m = fmax(m, fabs(dq));
  });
  m = block_reduce(m, OpMax(), 0);
  if (threadIdx.x == 0) a.partial[blockIdx.x] = m;
}

// GlobalQuantities.F90:38-56 sums/extrema: q0 anusin, q1 tempm, q2 vx2, q3 vy2,
// q4 vz2, q5 all2 (sums); q6..q8 vmax(vz,vy,vx), q9 tempmax (max); q10 tempmin (min)
__global__ void __launch_bounds__(256) k_red_global(ReduceArgs a) {
  const Geom g = a.g;
  double an = 0, tm = 0, x2 = 0, y2 = 0, z2 = 0, a2 = 0, mz = 0, my = 0, mx = 0, tmax = -HUGE_VAL,
         tmin = HUGE_VAL;
  for_points(g, g.nxm, [&](int, int, int64_t C, int s, int k) {
    const int skp = pos_of_k(k + 1, g.L);
    const double fac2 = a.g3rc[k];
    const double vx = a.f.vx[C + s], vy = a.f.vy[C + s], vz = a.f.vz[C + s], T = a.f.temp[C + s];
    mz = fmax(mz, fabs(vz));
    my = fmax(my, fabs(vy));
    mx = fmax(mx, fabs(vx));
    tmax = fmax(tmax, T);
    tmin = fmin(tmin, T);
    double vxcen = (vx + a.f.vx[C + skp]) * 0.5;
    double tcen = (T + a.f.temp[C + skp]) * 0.5;
    an += tcen * vxcen * fac2;
    tm += tcen * fac2;
    x2 += fac2 * (vx * vx);
    y2 += fac2 * (vy * vy);
    z2 += fac2 * (vz * vz);
    a2 += fac2 * (vz * vz + vy * vy + vx * vx);
  });
  double v[11] = {an, tm, x2, y2, z2, a2, mz, my, mx, tmax, tmin};
  for (int q = 0; q < 11; q++) {
    double r = q < 6 ? block_reduce(v[q], OpSum(), 0)
                     : (q < 10 ? block_reduce(v[q], OpMax(), 0) : block_reduce(v[q], OpMin(), 0));
    if (threadIdx.x == 0) a.partial[(int64_t)blockIdx.x * 11 + q] = r;
  }
}

// CalcPlateNu.F90:38-44: sums over columns
__global__ void k_red_platenu(ReduceArgs a, double del, double deln) {
  const Geom g = a.g;
  double lo = 0, up = 0;
  const int s0 = pos_of_k(0, g.L), s1 = pos_of_k(1, g.L), sm = pos_of_k(g.nxm - 1, g.L),
            sn = pos_of_k(g.nx - 1, g.L);
  for (int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; col < g.ncol;
       col += (int64_t)gridDim.x * blockDim.x) {
    const double *T = a.f.temp + col * g.ldx;
    lo += (T[s0] - T[s1]) * del;
    up += (T[sm] - T[sn]) * deln;
  }
  double r0 = block_reduce(lo, OpSum(), 0), r1 = block_reduce(up, OpSum(), 0);
  if (threadIdx.x == 0) {
    a.partial[(int64_t)blockIdx.x * 2 + 0] = r0;
    a.partial[(int64_t)blockIdx.x * 2 + 1] = r1;
  }
}

// CalcDissipationNu.F90:63-113: one block per plane k; out[2+k] disste, out[2+nxm+k] dissth,
// partial[k*2+{0,1}] = plane sums of nu_mu, nu_th contributions.
__global__ void k_red_diss(ReduceArgs a) {
  const Geom g = a.g;
  const int k = blockIdx.x;
  const int s = pos_of_k(k, g.L), sp = pos_of_k(k + 1, g.L);
  double mu = 0, th = 0, pe = 0, pt = 0;
  const double dxc = a.dxc[k], dxm = a.dxm[k], lem = a.lem[k], lec = a.lec[k];
  for (int64_t col = threadIdx.x; col < g.ncol; col += blockDim.x) {
    const int j = (int)(col % g.nym), i = (int)(col / g.nym);
    const int64_t C = col * g.ldx, CJP = col_off(wrapp(j + 1, g.nym), i, g),
                  CIP = col_off(j, wrapp(i + 1, g.nzm), g);
    const double *vx = a.f.vx, *vy = a.f.vy, *vz = a.f.vz, *T = a.f.temp;
    double hxx = (vx[C + sp] - vx[C + s]) * dxc;
    double hxy = (vx[CJP + s] - vx[C + s]) * a.dy;
    double hxz = (vx[CIP + s] - vx[C + s]) * a.dz;
    double hyx = (vy[C + sp] - vy[C + s]) * dxm;
    double hyy = (vy[CJP + s] - vy[C + s]) * a.dy;
    double hyz = (vy[CIP + s] - vy[C + s]) * a.dz;
    double hzx = (vz[C + sp] - vz[C + s]) * dxm;
    double hzy = (vz[CJP + s] - vz[C + s]) * a.dy;
    double hzz = (vz[CIP + s] - vz[C + s]) * a.dz;
    double d1 = (hxx * hxx + hxy * hxy + hxz * hxz);
    double d2 = (hyx * hyx + hyy * hyy + hyz * hyz) + (hzx * hzx + hzy * hzy + hzz * hzz);
    mu += d1 * lem * a.pra + d2 * lec * a.pra;
    double tx = (T[C + sp] - T[C + s]) * dxc;
    double ty = (T[CJP + s] - T[C + s]) * a.dy;
    double tz = (T[CIP + s] - T[C + s]) * a.dz;
    double dth = tx * tx + ty * ty + tz * tz;
    th += dth * lem;
    pe += (d1 + d2);
    pt += dth;
  }
  mu = block_reduce(mu, OpSum(), 0);
  th = block_reduce(th, OpSum(), 0);
  pe = block_reduce(pe, OpSum(), 0);
  pt = block_reduce(pt, OpSum(), 0);
  if (threadIdx.x == 0) {
    const double nn = (double)g.nym * (double)g.nzm;
    a.partial[k * 2 + 0] = mu;
    a.partial[k * 2 + 1] = th;
    a.out[2 + k] = pe / (a.ren * nn);
    a.out[2 + g.nxm + k] = pt / (a.pec * nn);
  }
}

// CalcStats (StatRoutines.F90:22-44): plane sums over (j, i) of vx, vy, vz, T,
// their squares and T*vx, per k.  Pass 1: block b sums its contiguous range of
// columns into partial[b][q][s] (thread = stored position s, so the loads are
// coalesced along the column); pass 2 adds the blocks' partials in a fixed
// order and accumulates sum * (1/nzm)(1/nym) into acc[q][k].
constexpr int SBLK = 256;
__global__ void __launch_bounds__(256) k_stats_partial(ReduceArgs a) {
  const Geom g = a.g;
  const int64_t cpb = (g.ncol + SBLK - 1) / SBLK;
  const int64_t c0 = blockIdx.x * cpb, c1 = min(c0 + cpb, g.ncol);
  for (int s = threadIdx.x; s < g.ldx; s += blockDim.x) {
    double q[STATS_N] = {0};
    for (int64_t c = c0; c < c1; c++) {
      const int64_t o = c * g.ldx + s;
      const double vx = a.f.vx[o], vy = a.f.vy[o], vz = a.f.vz[o], T = a.f.temp[o];
      q[0] += vx; q[1] += vy; q[2] += vz; q[3] += T;
      q[4] += vx * vx; q[5] += vy * vy; q[6] += vz * vz; q[7] += T * T; q[8] += T * vx;
    }
    for (int r = 0; r < STATS_N; r++) a.partial[((int64_t)blockIdx.x * STATS_N + r) * g.ldx + s] = q[r];
  }
}
__global__ void k_stats_finish(ReduceArgs a, double *acc) {
  const Geom g = a.g;
  const double w = (1.0 / g.nzm) * (1.0 / g.nym);
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < STATS_N * g.nxm; idx += gridDim.x * blockDim.x) {
    const int r = idx / g.nxm, k = idx - r * g.nxm, s = pos_of_k(k, g.L);
    double v = 0.0;
    for (int b = 0; b < SBLK; b++) v += a.partial[((int64_t)b * STATS_N + r) * g.ldx + s];
    acc[r * g.nxm + k] += v * w;
  }
}
hipError_t launch_stats(const ReduceArgs &a, double *acc, hipStream_t st) {
  hipLaunchKernelGGL(k_stats_partial, dim3(SBLK), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_stats_finish, dim3(16), dim3(256), 0, st, a, acc);
  return hipGetLastError();
}
size_t stats_partial_doubles(const Geom &g) { return (size_t)SBLK * STATS_N * g.ldx; }

// SlabDumper (SlabDumpRoutines.F90:20-31): planes k = kslab(m) of
// (vx(k)+vx(k+1))/2, vy, vz, temp for the nslab levels, as
// out[((m*4 + q) * nzm + i) * nym + j] (Fortran (j, i) order per plane).
__global__ void k_slabs(FieldPtrs f, Geom g, const int *kslab, int nslab, double *out) {
  const int64_t n = (int64_t)nslab * g.ncol;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
    const int m = (int)(idx / g.ncol);
    const int64_t c = idx - (int64_t)m * g.ncol;   // c = i * nym + j
    const int k = kslab[m] - 1;
    const int64_t o = c * g.ldx;
    const int s = pos_of_k(k, g.L), sp = pos_of_k(min(k + 1, g.nx - 1), g.L);
    double *dst = out + (int64_t)m * 4 * g.ncol + c;
    dst[0] = (f.vx[o + s] + f.vx[o + sp]) * 0.5;
    dst[g.ncol] = f.vy[o + s];
    dst[2 * g.ncol] = f.vz[o + s];
    dst[3 * g.ncol] = f.temp[o + s];
  }
}
hipError_t launch_slabs(const FieldPtrs &f, const Geom &g, const int *kslab, int nslab, double *out,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_slabs, dim3(4096), dim3(256), 0, st, f, g, kslab, nslab, out);
  return hipGetLastError();
}

// Restart regridding, interp + interptrilin (interp.F90:103-252), one
// thread per new point.  The padded old array of the reference is read
// through old_at: interior points from the file array, halo faces
// extrapolated as 2 a(1) - a(2) (interp.F90:103-124), and halo edges and
// corners, which the reference leaves unset, as 0 (a fresh allocation).
__device__ inline double old_raw(const InterpArgs &a, int k, int j, int i) {
  return a.old[((int64_t)(i - 1) * a.nyom + (j - 1)) * a.nxo + (k - 1)];
}
__device__ inline double old_at(const InterpArgs &a, int k, int j, int i) {
  const int hk = (k == 0) | (k == a.nxo + 1);
  const int hj = (j == a.js - 1) | (j == a.je + 1);
  const int hi = (i == a.is - 1) | (i == a.ie + 1);
  if (hk + hj + hi >= 2) return 0.0;
  if (hk) {
    const int k1 = k == 0 ? 1 : a.nxo, k2 = k == 0 ? 2 : a.nxo - 1;
    return 2.0 * old_raw(a, k1, j, i) - old_raw(a, k2, j, i);
  }
  if (hi) {
    const int i1 = i == a.is - 1 ? a.is : a.ie, i2 = i == a.is - 1 ? a.is + 1 : a.ie - 1;
    return 2.0 * old_raw(a, k, j, i1) - old_raw(a, k, j, i2);
  }
  if (hj) {
    const int j1 = j == a.js - 1 ? a.js : a.je, j2 = j == a.js - 1 ? a.js + 1 : a.je - 1;
    return 2.0 * old_raw(a, k, j1, i) - old_raw(a, k, j2, i);
  }
  return old_raw(a, k, j, i);
}
__global__ void __launch_bounds__(256) k_interp(InterpArgs a) {
#pragma clang fp contract(off)
  const Geom g = a.g;
  const int64_t n = g.ncol * (int64_t)g.nx;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx / g.nx;
    const int k = (int)(idx - c * g.nx);
    const int i = (int)(c / g.nym), j = (int)(c - (int64_t)i * g.nym);
    const int ek = k, ej = g.nx + j, ei = g.nx + g.nym + i;
    const int bifk = a.lohi[2 * ek], bick = a.lohi[2 * ek + 1];
    const int bifj = a.lohi[2 * ej], bicj = a.lohi[2 * ej + 1];
    const int bifi = a.lohi[2 * ei], bici = a.lohi[2 * ei + 1];
    const double biz = a.crd[3 * ek], biz1 = a.crd[3 * ek + 1], biz2 = a.crd[3 * ek + 2];
    const double biy = a.crd[3 * ej], biy1 = a.crd[3 * ej + 1], biy2 = a.crd[3 * ej + 2];
    const double bix = a.crd[3 * ei], bix1 = a.crd[3 * ei + 1], bix2 = a.crd[3 * ei + 2];
    // the eight corners as interp.F90:182-189 picks them: all but one of the
    // upper-x corners come from the lower x level
    const double aficjfk = old_at(a, bifk, bicj, bifi);
    const double afifjfk = old_at(a, bifk, bifj, bifi);
    const double acicjfk = old_at(a, bifk, bicj, bici);
    const double acifjfk = old_at(a, bifk, bifj, bici);
    const double aficjck = old_at(a, bick, bicj, bifi);
    const double afifjck = afifjfk;
    const double acicjck = acicjfk;
    const double acifjck = acifjfk;
    const double dxdydz = 1.0 / ((bix2 - bix1) * (biy2 - biy1) * (biz2 - biz1));
    const double ans = (aficjfk * (bix2 - bix) * (biy - biy1) * (biz2 - biz) +
                        afifjfk * (bix2 - bix) * (biy2 - biy) * (biz2 - biz) +
                        acicjfk * (bix - bix1) * (biy - biy1) * (biz2 - biz) +
                        acifjfk * (bix - bix1) * (biy2 - biy) * (biz2 - biz) +
                        aficjck * (bix2 - bix) * (biy - biy1) * (biz - biz1) +
                        afifjck * (bix2 - bix) * (biy2 - biy) * (biz - biz1) +
                        acicjck * (bix - bix1) * (biy - biy1) * (biz - biz1) +
                        acifjck * (bix - bix1) * (biy2 - biy) * (biz - biz1)) *
                       dxdydz;
    a.out[c * g.ldx + pos_of_k(k, g.L)] = ans;
  }
}
hipError_t launch_interp(const InterpArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_interp, dim3(8192), dim3(256), 0, st, a);
  return hipGetLastError();
}

__global__ void k_convert(const void *src, void *dst, int64_t n, int to_f32) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    if (to_f32) static_cast<float *>(dst)[e] = (float)static_cast<const double *>(src)[e];
    else static_cast<double *>(dst)[e] = (double)static_cast<const float *>(src)[e];
  }
}
hipError_t launch_convert(const void *src, void *dst, int64_t n, int to_f32, hipStream_t st) {
  hipLaunchKernelGGL(k_convert, dim3(4096), dim3(256), 0, st, src, dst, n, to_f32);
  return hipGetLastError();
}

static hipError_t finalize(const ReduceArgs &a, int nq, int maxmask, int minmask, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(nq), dim3(256), 0, st, a.partial, a.out, nq, (const int *)nullptr,
                     maxmask, minmask);
  return hipGetLastError();
}

hipError_t launch_reduce_cfl(const ReduceArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_red_cfl, dim3(RBLK), dim3(256), 0, st, a);
  return finalize(a, 1, 1, 0, st);
}
hipError_t launch_reduce_divmax(const ReduceArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_red_div, dim3(RBLK), dim3(256), 0, st, a);
  return finalize(a, 1, 1, 0, st);
}
hipError_t launch_reduce_global(const ReduceArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_red_global, dim3(RBLK), dim3(256), 0, st, a);
  return finalize(a, 11, (0xF << 6), (1 << 10), st);
}
hipError_t launch_reduce_platenu(const ReduceArgs &a, double del, double deln, hipStream_t st) {
  hipLaunchKernelGGL(k_red_platenu, dim3(RBLK), dim3(256), 0, st, a, del, deln);
  return finalize(a, 2, 0, 0, st);
}
__global__ void k_diss_sum(ReduceArgs a) {
  double mu = 0, th = 0;
  for (int k = threadIdx.x; k < a.g.nxm; k += blockDim.x) {
    mu += a.partial[k * 2];
    th += a.partial[k * 2 + 1];
  }
  mu = block_reduce(mu, OpSum(), 0);
  th = block_reduce(th, OpSum(), 0);
  if (threadIdx.x == 0) {
    a.out[0] = mu;
    a.out[1] = th;
  }
}
hipError_t launch_reduce_dissipation(const ReduceArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_red_diss, dim3(a.g.nxm), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_diss_sum, dim3(1), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------- layout conversion
// stage: raw copy of the host array; element (k,j,i) (0-based, k<n1) at
// stage[k + ld1*((j+1-j0) + ld2*(i+1-i0))]  (j0/i0 are the host lower bounds).
__global__ void k_gather_in(const double *stage, double *dst, Geom g, int n1, int ld1, int ld2, int j0,
                            int i0) {
  for (int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < g.ncol * g.ldx;
       gid += (int64_t)gridDim.x * blockDim.x) {
    const int64_t col = gid / g.ldx;
    const int s = (int)(gid % g.ldx), k = k_of_pos(s, g.L);
    if (k >= n1) continue;
    const int j = (int)(col % g.nym), i = (int)(col / g.nym);
    dst[gid] = stage[k + (int64_t)ld1 * ((j + 1 - j0) + (int64_t)ld2 * (i + 1 - i0))];
  }
}
// inverse; with halo != 0 the host layout's halo planes (j = j0-1.. etc.) are
// filled with periodic copies, as update_halo leaves them.
__global__ void k_scatter_out(const double *src, double *stage, Geom g, int n1, int ld1, int ld2, int j0,
                              int i0, int halo) {
  const int ny = halo ? g.nym + 2 : g.nym, nz = halo ? g.nzm + 2 : g.nzm;
  const int64_t total = (int64_t)n1 * ny * nz;
  for (int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(gid % n1);
    const int64_t r = gid / n1;
    const int jj = (int)(r % ny), ii = (int)(r / ny);
    int j = halo ? jj - 1 : jj, i = halo ? ii - 1 : ii;  // 0-based interior index (may be -1 or n)
    int jw = j < 0 ? j + g.nym : (j >= g.nym ? j - g.nym : j);
    int iw = i < 0 ? i + g.nzm : (i >= g.nzm ? i - g.nzm : i);
    stage[k + (int64_t)ld1 * ((j + 1 - j0) + (int64_t)ld2 * (i + 1 - i0))] =
        src[col_off(jw, iw, g) + pos_of_k(k, g.L)];
  }
}
hipError_t launch_gather_in(const double *stage, double *dst, const Geom &g, int n1, int ld1, int ld2, int j0,
                            int i0, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_in, dim3(4096), dim3(256), 0, st, stage, dst, g, n1, ld1, ld2, j0, i0);
  return hipGetLastError();
}
hipError_t launch_scatter_out(const double *src, double *stage, const Geom &g, int n1, int ld1, int ld2, int j0,
                              int i0, int halo, hipStream_t st) {
  hipLaunchKernelGGL(k_scatter_out, dim3(4096), dim3(256), 0, st, src, stage, g, n1, ld1, ld2, j0, i0, halo);
  return hipGetLastError();
}

// ---------------------------------------------------------------- initial condition
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
struct ZeroSet { double *p[4]; };
__global__ void k_init_noise(double *temp, ZeroSet zs, const double *xc, Geom g, uint64_t seed, double amp,
                             double tbp, double ttp) {
  for (int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < g.ncol * g.ldx;
       gid += (int64_t)gridDim.x * blockDim.x) {
    const int64_t col = gid / g.ldx;
    const int s = (int)(gid % g.ldx), k = k_of_pos(s, g.L);
    for (int q = 0; q < 4; q++) zs.p[q][gid] = 0.0;
    if (k >= g.nx) {
      temp[gid] = 0.0;
      continue;
    }
    const int j = (int)(col % g.nym), i = (int)(col / g.nym);
    double v;
    if (k == 0) v = tbp;
    else if (k == g.nx - 1) v = ttp;
    else {
      uint64_t lin = (uint64_t)k + (uint64_t)g.nx * ((uint64_t)j + (uint64_t)g.nym * (uint64_t)i);
      uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ULL + (lin + 1) * 0xD1B54A32D192ED03ULL);
      double u = (double)(h >> 11) * 0x1.0p-53;
      v = 1.0 - xc[k] + (2.0 * u - 1.0) * amp;
    }
    temp[gid] = v;
  }
}
hipError_t launch_init_noise(double *temp, double *zero_fields[4], const double *xc_nat, const Geom &g,
                             uint64_t seed, double amp, double tbp, double ttp, hipStream_t st) {
  ZeroSet zs;
  for (int q = 0; q < 4; q++) zs.p[q] = zero_fields[q];
  hipLaunchKernelGGL(k_init_noise, dim3(4096), dim3(256), 0, st, temp, zs, xc_nat, g, seed, amp, tbp, ttp);
  return hipGetLastError();
}

__global__ void k_fill(double *p, double v, int64_t n) {
  for (int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < n; gid += (int64_t)gridDim.x * blockDim.x)
    p[gid] = v;
}
hipError_t launch_fill(double *p, double v, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

}  // namespace afid
