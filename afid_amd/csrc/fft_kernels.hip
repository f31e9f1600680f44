// fft_kernels.hip -- pressure correction (SolvePressureCorrection.F90:12-154)
// as four kernels over the spectral array S[k][n][m]:
//   1. y real-to-complex FFT of dph (optionally computing dph as the local
//      divergence of u*, CalcLocalDivergence.F90, on the fly),
//   2. z complex FFT (forward, and later backward),
//   3. per-(m,n) tridiagonal solve in x (the zgttrf/zgttrs loop, :117-141,
//      including the 1/(nym*nzm) normalisation of :105), each mode's line
//      held in registers across a group of lanes,
//   4. y complex-to-real FFT back into dph.
// FFTs are in-LDS mixed-radix (4,2,3,5) decimation in time, unnormalised,
// forward sign -1, with FFTW's c2r convention (imaginary parts of the DC and
// Nyquist bins ignored).  A real line of even length N is transformed as a
// complex line of N/2 points plus a post-twiddle.
#include "afid_internal.h"
#include "kernels.h"

namespace afid {

// Complex arithmetic in FP64 (double2) and, for the mixed-precision mode, FP32
// (float2); cplx_t<R> names the complex type of real type R.
template <typename R> struct CplxOf;
template <> struct CplxOf<double> { using T = double2; };
template <> struct CplxOf<float> { using T = float2; };
template <typename R> using cplx_t = typename CplxOf<R>::T;

__device__ __forceinline__ double2 mkc(double x, double y) { return make_double2(x, y); }
__device__ __forceinline__ float2 mkc(float x, float y) { return make_float2(x, y); }
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  return mkc(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return mkc(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return mkc(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C conj2(C a) { return mkc(a.x, -a.y); }
// multiply by s*i (s = +1/-1)
template <typename C, typename R> __device__ __forceinline__ C mul_si(C a, R s) { return mkc(-s * a.y, s * a.x); }
// FP64 table value in the working precision
template <typename R> __device__ __forceinline__ cplx_t<R> cvt(double2 t) { return mkc((R)t.x, (R)t.y); }
__device__ __forceinline__ double2 widen(double2 v) { return v; }
__device__ __forceinline__ double2 widen(float2 v) { return make_double2(v.x, v.y); }
template <typename C> __device__ __forceinline__ C narrow(double2 v);
template <> __device__ __forceinline__ double2 narrow<double2>(double2 v) { return v; }
template <> __device__ __forceinline__ float2 narrow<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }

template <typename R>
__device__ __forceinline__ cplx_t<R> twid(const double2 *tw, int idx, R sign) {
  cplx_t<R> t = cvt<R>(tw[idx]);
  if (sign > 0) t.y = -t.y;
  return t;
}

// LDS element order within a line: for lines whose length is a multiple of 8,
// element e lives at e ^ ((e >> 3) & 7).  A double2 spans 4 of the 32 LDS
// banks, so 8 consecutive elements fill one bank row; the XOR spreads the
// strided butterfly accesses of the early radix stages over the rows.
__device__ __forceinline__ int swz(int e, bool on) { return on ? e ^ ((e >> 3) & 7) : e; }

// LDS writes of this wave are complete and visible to its own later reads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Global loads below are issued in batches of LB (the whole batch is in
// flight before the first value is used) so that a thread never waits for one
// load at a time.
constexpr int LB = 8;
constexpr int YILB = 6;  // the y-inverse batch: two equal batches of its 12 loads per thread
constexpr int ZLB = 12;  // the z transform's batch: 12 loads cover a 4 x 768 tile per 256 threads in one
constexpr int TB = 4;   // butterfly iterations per twiddle batch

// In-place DIT stages over `nlines` (1 or 2) lines of length pl.n stored
// contiguously in LDS from `lines`, input already in digit-reversed order,
// computed by ONE wavefront (lane = 0..63): the stages synchronise within the
// wave only, so the waves of a block progress independently.  sign = -1
// forward, +1 backward.  Stage with radix r and span m combines, for every
// block b and offset u < m, the r points b*m*r + u + v*m (v < r) after
// twiddling point v by W_{mr}^{uv}.  The twiddles of TB butterflies per lane
// are loaded together before any of them is computed.
template <typename R>
__device__ __forceinline__ void lds_fft_wave(cplx_t<R> *lines, int nlines, const FFTPlan &pl, R sign, int lane) {
  using C = cplx_t<R>;
  const int n = pl.n;
  const bool sw = (n & 7) == 0;
  const double2 *tw = pl.tw;
  int m = 1;
  for (int st = 0; st < pl.nf; st++) {
    const int r = pl.f[st];
    const int M = m * r;
    const int nb = n / r;
    const int tws = n / M;
    const unsigned mdiv = pl.mdiv[st];
    const int total = nlines * nb;
    for (int w0 = 0; w0 < total; w0 += 64 * TB) {
      C t1[TB], t2[TB], t3[TB], t4[TB];
      int lo[TB], e0[TB];
      bool ok[TB];
#pragma unroll
      for (int b = 0; b < TB; b++) {
        const int wid = w0 + b * 64 + lane;
        ok[b] = wid < total;
        const int wl = ok[b] ? wid : 0;
        const int q = wl >= nb ? 1 : 0;
        const int w = wl - q * nb;
        const int bb = m == 1 ? w : (int)__umulhi((unsigned)w, mdiv);
        const int u = w - bb * m;
        lo[b] = q * n;
        e0[b] = bb * M + u;
        const int t1i = u * tws;   // v*t1i < n for v < r; clamp the unused ones
        t1[b] = twid(tw, t1i, sign);
        t2[b] = twid(tw, min(2 * t1i, n - 1), sign);
        t3[b] = twid(tw, min(3 * t1i, n - 1), sign);
        t4[b] = twid(tw, min(4 * t1i, n - 1), sign);
      }
#pragma unroll
      for (int b = 0; b < TB; b++) {
        if (!ok[b]) continue;
        C *line = lines + lo[b];
#define X(v) line[swz(e0[b] + (v) * m, sw)]
        if (r == 2) {
          C x0 = X(0), x1 = cmul(X(1), t1[b]);
          X(0) = cadd(x0, x1);
          X(1) = csub(x0, x1);
        } else if (r == 4) {
          C x0 = X(0);
          C x1 = cmul(X(1), t1[b]);
          C x2 = cmul(X(2), t2[b]);
          C x3 = cmul(X(3), t3[b]);
          C a0 = cadd(x0, x2), a1 = csub(x0, x2);
          C b0 = cadd(x1, x3), b1 = mul_si(csub(x1, x3), sign);
          X(0) = cadd(a0, b0);
          X(1) = cadd(a1, b1);
          X(2) = csub(a0, b0);
          X(3) = csub(a1, b1);
        } else if (r == 3) {
          const R s3 = (R) 0.86602540378443864676372317075294;  // sin(2 pi/3)
          C x0 = X(0);
          C x1 = cmul(X(1), t1[b]);
          C x2 = cmul(X(2), t2[b]);
          C p1 = cadd(x1, x2);
          C p2 = mul_si(csub(x1, x2), sign * s3);
          C q0 = mkc(x0.x - (R)0.5 * p1.x, x0.y - (R)0.5 * p1.y);
          X(0) = cadd(x0, p1);
          X(1) = cadd(q0, p2);
          X(2) = csub(q0, p2);
        } else {  // r == 5
          const R c1 = (R) 0.30901699437494742410229341718282;   // cos(2 pi/5)
          const R c2 = (R) -0.80901699437494742410229341718282;  // cos(4 pi/5)
          const R s1 = (R) 0.95105651629515357211643933337938;   // sin(2 pi/5)
          const R s2 = (R) 0.58778525229247312916870595463907;   // sin(4 pi/5)
          C x0 = X(0);
          C x1 = cmul(X(1), t1[b]);
          C x2 = cmul(X(2), t2[b]);
          C x3 = cmul(X(3), t3[b]);
          C x4 = cmul(X(4), t4[b]);
          C p1 = cadd(x1, x4), p2 = cadd(x2, x3);
          C p3 = csub(x1, x4), p4 = csub(x2, x3);
          C a1 = mkc(x0.x + c1 * p1.x + c2 * p2.x, x0.y + c1 * p1.y + c2 * p2.y);
          C a2 = mkc(x0.x + c2 * p1.x + c1 * p2.x, x0.y + c2 * p1.y + c1 * p2.y);
          C b1 = mul_si(mkc(s1 * p3.x + s2 * p4.x, s1 * p3.y + s2 * p4.y), sign);
          C b2 = mul_si(mkc(s2 * p3.x - s1 * p4.x, s2 * p3.y - s1 * p4.y), sign);
          X(0) = cadd(x0, cadd(p1, p2));
          X(1) = cadd(a1, b1);
          X(2) = cadd(a2, b2);
          X(3) = csub(a2, b2);
          X(4) = csub(a1, b1);
        }
#undef X
      }
    }
    wave_lds_sync();
    m = M;
  }
}

// Block index of a 1-D launch of `total` logical tiles (grid padded to a
// multiple of 8), remapped so that each XCD (blocks are dealt to the 8 XCDs
// round-robin) works on one contiguous range of tiles: neighbouring tiles share
// cache lines and stay in one L2.  Returns -1 for padding blocks.
__device__ __forceinline__ int64_t fft_tile(int64_t total) {
  const int64_t b = blockIdx.x, per = gridDim.x / 8;
  const int64_t t = (b % 8) * per + b / 8;
  return t < total ? t : -1;
}

constexpr int TK = 8;   // real lines (x slots) per y-FFT block
constexpr int TM = 4;   // complex lines (y wavenumbers) per z-FFT block

// ---------------------------------------------------------------- y forward
// tile: (TK consecutive interleaved positions, z index i); wave w transforms
// lines 2w and 2w+1 of the tile.  Thread t loads slot q = t % TK of the rows
// j = t / TK + 32 it, so its k, neighbour slot and udx3m stay fixed.
// The same stages for a line length known at compile time.  The radices
// come in the order the host plan uses (4s, then 2s, 3s, 5s) and every
// butterfly performs the same operations on the same twiddles, so a line
// transformed here equals the runtime form bit for bit; with the radix, span
// and line count constant, the index arithmetic folds and only the twiddles a
// radix uses are loaded.
template <int N> struct CtPlan {
  int f[16] = {};
  int nf = 0;
  constexpr CtPlan() {
    int n = N;
    while (n % 4 == 0) { f[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { f[nf++] = 2; n /= 2; }
    while (n % 3 == 0) { f[nf++] = 3; n /= 3; }
    while (n % 5 == 0) { f[nf++] = 5; n /= 5; }
  }
};

template <typename R, int N, int NL, int ST, int M>
__device__ __forceinline__ void ct_fft_stage(cplx_t<R> *lines, const double2 *tw, R sign, int lane) {
  using C = cplx_t<R>;
  constexpr CtPlan<N> PL{};
  constexpr int r = PL.f[ST];
  constexpr int MR = M * r, nb = N / r, tws = N / MR, total = NL * nb;
  constexpr bool sw = (N & 7) == 0;
#pragma unroll
  for (int w0 = 0; w0 < total; w0 += 64 * TB) {
    C t1[TB], t2[TB], t3[TB], t4[TB];
    int lo[TB], e0[TB];
    bool ok[TB];
#pragma unroll
    for (int b = 0; b < TB; b++) {
      if (w0 + b * 64 >= total) break;
      const int wid = w0 + b * 64 + lane;
      ok[b] = wid < total;
      const int wl = ok[b] ? wid : 0;
      const int q = (NL > 1 && wl >= nb) ? 1 : 0;
      const int w = wl - q * nb;
      const int bb = w / M;
      const int u = w - bb * M;
      lo[b] = q * N;
      e0[b] = bb * MR + u;
      const int t1i = u * tws;   // v * t1i < N for v < r
      t1[b] = twid(tw, t1i, sign);
      if constexpr (r >= 3) t2[b] = twid(tw, 2 * t1i, sign);
      if constexpr (r >= 4) t3[b] = twid(tw, 3 * t1i, sign);
      if constexpr (r >= 5) t4[b] = twid(tw, 4 * t1i, sign);
    }
#pragma unroll
    for (int b = 0; b < TB; b++) {
      if (w0 + b * 64 >= total) break;
      if (!ok[b]) continue;
      C *line = lines + lo[b];
#define X(v) line[swz(e0[b] + (v) * M, sw)]
      if constexpr (r == 2) {
        C x0 = X(0), x1 = cmul(X(1), t1[b]);
        X(0) = cadd(x0, x1);
        X(1) = csub(x0, x1);
      } else if constexpr (r == 4) {
        C x0 = X(0);
        C x1 = cmul(X(1), t1[b]);
        C x2 = cmul(X(2), t2[b]);
        C x3 = cmul(X(3), t3[b]);
        C a0 = cadd(x0, x2), a1 = csub(x0, x2);
        C b0 = cadd(x1, x3), b1 = mul_si(csub(x1, x3), sign);
        X(0) = cadd(a0, b0);
        X(1) = cadd(a1, b1);
        X(2) = csub(a0, b0);
        X(3) = csub(a1, b1);
      } else if constexpr (r == 3) {
        const R s3 = (R) 0.86602540378443864676372317075294;  // sin(2 pi/3)
        C x0 = X(0);
        C x1 = cmul(X(1), t1[b]);
        C x2 = cmul(X(2), t2[b]);
        C p1 = cadd(x1, x2);
        C p2 = mul_si(csub(x1, x2), sign * s3);
        C q0 = mkc(x0.x - (R)0.5 * p1.x, x0.y - (R)0.5 * p1.y);
        X(0) = cadd(x0, p1);
        X(1) = cadd(q0, p2);
        X(2) = csub(q0, p2);
      } else {  // r == 5
        const R c1 = (R) 0.30901699437494742410229341718282;   // cos(2 pi/5)
        const R c2 = (R) -0.80901699437494742410229341718282;  // cos(4 pi/5)
        const R s1 = (R) 0.95105651629515357211643933337938;   // sin(2 pi/5)
        const R s2 = (R) 0.58778525229247312916870595463907;   // sin(4 pi/5)
        C x0 = X(0);
        C x1 = cmul(X(1), t1[b]);
        C x2 = cmul(X(2), t2[b]);
        C x3 = cmul(X(3), t3[b]);
        C x4 = cmul(X(4), t4[b]);
        C p1 = cadd(x1, x4), p2 = cadd(x2, x3);
        C p3 = csub(x1, x4), p4 = csub(x2, x3);
        C a1 = mkc(x0.x + c1 * p1.x + c2 * p2.x, x0.y + c1 * p1.y + c2 * p2.y);
        C a2 = mkc(x0.x + c2 * p1.x + c1 * p2.x, x0.y + c2 * p1.y + c1 * p2.y);
        C b1 = mul_si(mkc(s1 * p3.x + s2 * p4.x, s1 * p3.y + s2 * p4.y), sign);
        C b2 = mul_si(mkc(s2 * p3.x - s1 * p4.x, s2 * p3.y - s1 * p4.y), sign);
        X(0) = cadd(x0, cadd(p1, p2));
        X(1) = cadd(a1, b1);
        X(2) = cadd(a2, b2);
        X(3) = csub(a2, b2);
        X(4) = csub(a1, b1);
      }
#undef X
    }
  }
  wave_lds_sync();
  if constexpr (ST + 1 < PL.nf) ct_fft_stage<R, N, NL, ST + 1, MR>(lines, tw, sign, lane);
}

// Dispatch on the plan's length: compile-time stages for the lengths the
// configurations use, the runtime stages for any other.
template <typename R, int NL>
__device__ __forceinline__ void fft_wave(cplx_t<R> *lines, const FFTPlan &pl, R sign, int lane) {
  switch (pl.n) {
#define CT_CASE(NN) \
  case NN: ct_fft_stage<R, NN, NL, 0, 1>(lines, pl.tw, sign, lane); return;
    CT_CASE(768) CT_CASE(512) CT_CASE(384) CT_CASE(256) CT_CASE(192) CT_CASE(128) CT_CASE(96) CT_CASE(64)
    CT_CASE(48) CT_CASE(32)
#undef CT_CASE
    default: lds_fft_wave<R>(lines, NL, pl, sign, lane);
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_yfft_forward(PoissonArgs a) {
  using C = cplx_t<R>;
  extern __shared__ double2 lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  C *S = reinterpret_cast<C *>(a.S);
  const Geom g = a.g;
  const RealPlan &py = a.py;
  const int H = py.H, N = py.N;
  const int ntile = g.ldx / TK;
  const int64_t tile = fft_tile((int64_t)ntile * g.nzm);
  if (tile < 0) return;
  if (a.ds && coh_int(&a.ds->halted)) return;
  const StepScalars sc = a.ds ? coh_scalars(&a.ds->sc[a.ns - 1]) : a.sc;
  const int s0 = (int)(tile % ntile) * TK, i = (int)(tile / ntile);
  const bool hsw = (H & 7) == 0;
  const double usdtal = 1.0 / (sc.dt * sc.al);
  {
    const int q = threadIdx.x % TK, jr = threadIdx.x / TK;   // 32 rows per pass
    const int s = s0 + q, k = k_of_pos(s, g.L);
    const bool live = k < g.nxm;
    const int skp = live ? pos_of_k(k + 1, g.L) : s;
    const double u3 = a.udx3m[s];
    const int ip = wrapp(i + 1, g.nzm);
    constexpr int YB = 8;   // rows per batch: 6 field loads each, all issued before any is used
    for (int jb = jr; jb < N; jb += 32 * YB) {
      double z0[YB], z1[YB], y0[YB], y1[YB], x0[YB], x1[YB];
      int rv[YB];
#pragma unroll
      for (int b = 0; b < YB; b++) {
        const int jc = min(jb + 32 * b, N - 1);
        const int64_t C = col_off(jc, i, g);
        if (a.from_velocity) {
          const int64_t CJP = col_off(wrapp(jc + 1, g.nym), i, g), CIP = col_off(jc, ip, g);
          z1[b] = a.f.vz[CIP + s];
          z0[b] = a.f.vz[C + s];
          y1[b] = a.f.vy[CJP + s];
          y0[b] = a.f.vy[C + s];
          x1[b] = a.f.vx[C + skp];
          x0[b] = a.f.vx[C + s];
        } else {
          z0[b] = a.dph[C + s];
          z1[b] = y0[b] = y1[b] = x0[b] = x1[b] = 0.0;
        }
        rv[b] = py.half ? py.c.rev[jc >> 1] : py.c.rev[jc];
      }
#pragma unroll
      for (int b = 0; b < YB; b++) {
        const int j = jb + 32 * b;
        if (j >= N) continue;
        double x = z0[b];
        if (a.from_velocity)
          x = ((z1[b] - z0[b]) * sc.dz + (y1[b] - y0[b]) * sc.dy + (x1[b] - x0[b]) * u3) * usdtal;
        const R v = live ? (R)x : (R)0;   // divergence in FP64, spectrum in R
        C *e = &lds[q * H + swz(rv[b], hsw)];
        if (!py.half) *e = mkc(v, (R)0);
        else if (j & 1) e->y = v;
        else e->x = v;
      }
    }
  }
  __syncthreads();
  fft_wave<R, TK / 4>(lds + (threadIdx.x / 64) * (TK / 4) * H, py.c, (R)-1, threadIdx.x % 64);
  __syncthreads();
  // Post-processing: 32 lanes per real line q, wavenumbers m = lane + 32 t.
  // The padded wavenumbers m in [nymh, ldm) of each row store zeros, so that
  // the rows' 128-byte lines are written whole.
  static_assert(TK * 32 == 256, "one real line per 32 lanes");
  {
    const int nymh = a.nymh, ldm = a.ldm;
    const int q = (int)threadIdx.x >> 5, lm = (int)threadIdx.x & 31;
    const int k = k_of_pos(s0 + q, g.L);
    const bool rowok = k < g.nxm;
    C *Srow = S + ((int64_t)(rowok ? k : 0) * g.nzm + i) * ldm;
    for (int m0 = lm; m0 < ldm; m0 += 32 * LB) {
      C pw[LB];
#pragma unroll
      for (int b = 0; b < LB; b++) {
        const int m = min(m0 + 32 * b, nymh - 1);
        pw[b] = py.half ? cvt<R>(py.post[m]) : mkc((R)1, (R)0);
      }
#pragma unroll
      for (int b = 0; b < LB; b++) {
        const int m = m0 + 32 * b;
        if (m >= ldm) continue;
        C X = mkc((R)0, (R)0);
        if (m < nymh) {
          if (py.half) {
            C zm = lds[q * H + swz(m == H ? 0 : m, hsw)];
            C zc = conj2(lds[q * H + swz(m == 0 ? 0 : H - m, hsw)]);
            C E = mkc((R)0.5 * (zm.x + zc.x), (R)0.5 * (zm.y + zc.y));
            C D = csub(zm, zc);
            C O = mkc((R)0.5 * D.y, (R)-0.5 * D.x);  // D / (2i)
            X = cadd(E, cmul(pw[b], O));
          } else {
            X = lds[q * H + swz(m, hsw)];
          }
        }
        if (rowok) Srow[m] = X;
      }
    }
  }
}

// ---------------------------------------------------------------- z transforms
// tile: (TM wavenumbers m, k); wave w transforms line w
template <typename R>
__global__ void __launch_bounds__(256) k_zfft(PoissonArgs a, R sign) {
  using C = cplx_t<R>;
  extern __shared__ double2 lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  const int N = a.pz.n;
  const int ntile = (a.nymh + TM - 1) / TM;
  const int64_t tile = fft_tile((int64_t)ntile * a.g.nxm);
  if (tile < 0 || (a.ds && coh_int(&a.ds->halted))) return;
  const int m0 = (int)(tile % ntile) * TM, k = (int)(tile / ntile);
  const bool zsw = (N & 7) == 0;
  C *S = reinterpret_cast<C *>(a.S) + (int64_t)k * a.g.nzm * a.ldm;
  const int tot = TM * N;
  for (int base = threadIdx.x; base < tot; base += 256 * ZLB) {
    C v[ZLB];
    int rv[ZLB];
#pragma unroll
    for (int b = 0; b < ZLB; b++) {   // all loads of the batch first
      const int ic = min(base + 256 * b, tot - 1);
      const int q = ic % TM, nn = ic / TM, m = m0 + q;
      v[b] = S[(int64_t)nn * a.ldm + min(m, a.nymh - 1)];
      rv[b] = a.pz.rev[nn];
    }
#pragma unroll
    for (int b = 0; b < ZLB; b++) {
      const int idx = base + 256 * b;
      if (idx >= tot) continue;
      const int q = idx % TM;
      lds[q * N + swz(rv[b], zsw)] = m0 + q < a.nymh ? v[b] : mkc((R)0, (R)0);
    }
  }
  __syncthreads();
  fft_wave<R, 1>(lds + (threadIdx.x / 64) * N, a.pz, sign, threadIdx.x % 64);
  __syncthreads();
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int q = idx % TM, nn = idx / TM;
    const int m = m0 + q;   // m < ldm: the padded wavenumbers carry zeros, so rows are written whole
    S[(int64_t)nn * a.ldm + m] = lds[q * N + swz(nn, zsw)];
  }
}

// ---------------------------------------------------------------- z transform with a matrix-core stage
// N = 16 Q.  With n = Q a + b (a < 16, b < Q) and k = c + 16 d:
//   Y[c + 16 d] = sum_b W_Q^{bd} W_N^{bc} sum_a W_16^{ac} X[Q a + b],
// so the 16-point DFTs over a are one complex matrix product per tile of 16
// rows (q, b): Z = X F16, with X[(q,b)][a] and F16[a][c] = W_16^{ac}, done as
// four real products per K step of 4 on v_mfma_f64_16x16x4f64.  Lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16], and register r of D[(l / 16) + 4r]
// [l % 16] (checked on the device).  The twiddles W_N^{bc} follow, then the
// Q-point transforms over b for the 16 c of each line run in LDS with the
// radix stages of lds_fft_wave.  A wave holds the products of its TPW tiles
// in registers until every wave has read the input, then writes them in place.
typedef double dbl4_t __attribute__((ext_vector_type(4)));

// In-place DIT over nl lines of length pl.n stored contiguously from `lines`,
// input in digit-reversed order, by one wavefront (as lds_fft_wave, for any
// number of lines).
__device__ __forceinline__ void lds_fft_lines(double2 *lines, int nl, const FFTPlan &pl, double sign, int lane) {
  const int n = pl.n;
  const bool sw = (n & 7) == 0;
  const double2 *tw = pl.tw;
  int m = 1;
  for (int st = 0; st < pl.nf; st++) {
    const int r = pl.f[st];
    const int M = m * r;
    const int nb = n / r;
    const int tws = n / M;
    const unsigned mdiv = pl.mdiv[st];
    const int total = nl * nb;
    for (int wid = lane; wid < total; wid += 64) {
      const int q = wid / nb, w = wid - q * nb;
      const int bb = m == 1 ? w : (int)__umulhi((unsigned)w, mdiv);
      const int u = w - bb * m;
      double2 *line = lines + q * n;
      const int e0 = bb * M + u;
      const int t1i = u * tws;
      const double2 t1 = twid(tw, t1i, sign), t2 = twid(tw, min(2 * t1i, n - 1), sign);
      const double2 t3 = twid(tw, min(3 * t1i, n - 1), sign), t4 = twid(tw, min(4 * t1i, n - 1), sign);
#define X(v) line[swz(e0 + (v) * m, sw)]
      if (r == 2) {
        double2 x0 = X(0), x1 = cmul(X(1), t1);
        X(0) = cadd(x0, x1);
        X(1) = csub(x0, x1);
      } else if (r == 4) {
        double2 x0 = X(0), x1 = cmul(X(1), t1), x2 = cmul(X(2), t2), x3 = cmul(X(3), t3);
        double2 a0 = cadd(x0, x2), a1 = csub(x0, x2);
        double2 b0 = cadd(x1, x3), b1 = mul_si(csub(x1, x3), sign);
        X(0) = cadd(a0, b0);
        X(1) = cadd(a1, b1);
        X(2) = csub(a0, b0);
        X(3) = csub(a1, b1);
      } else if (r == 3) {
        const double s3 = 0.86602540378443864676372317075294;  // sin(2 pi/3)
        double2 x0 = X(0), x1 = cmul(X(1), t1), x2 = cmul(X(2), t2);
        double2 p1 = cadd(x1, x2), p2 = mul_si(csub(x1, x2), sign * s3);
        double2 q0 = mkc(x0.x - 0.5 * p1.x, x0.y - 0.5 * p1.y);
        X(0) = cadd(x0, p1);
        X(1) = cadd(q0, p2);
        X(2) = csub(q0, p2);
      } else {  // r == 5
        const double c1 = 0.30901699437494742410229341718282, c2 = -0.80901699437494742410229341718282;
        const double s1 = 0.95105651629515357211643933337938, s2 = 0.58778525229247312916870595463907;
        double2 x0 = X(0), x1 = cmul(X(1), t1), x2 = cmul(X(2), t2), x3 = cmul(X(3), t3), x4 = cmul(X(4), t4);
        double2 p1 = cadd(x1, x4), p2 = cadd(x2, x3), p3 = csub(x1, x4), p4 = csub(x2, x3);
        double2 a1 = mkc(x0.x + c1 * p1.x + c2 * p2.x, x0.y + c1 * p1.y + c2 * p2.y);
        double2 a2 = mkc(x0.x + c2 * p1.x + c1 * p2.x, x0.y + c2 * p1.y + c1 * p2.y);
        double2 b1 = mul_si(mkc(s1 * p3.x + s2 * p4.x, s1 * p3.y + s2 * p4.y), sign);
        double2 b2 = mul_si(mkc(s2 * p3.x - s1 * p4.x, s2 * p3.y - s1 * p4.y), sign);
        X(0) = cadd(x0, cadd(p1, p2));
        X(1) = cadd(a1, b1);
        X(2) = cadd(a2, b2);
        X(3) = csub(a2, b2);
        X(4) = csub(a1, b1);
      }
#undef X
    }
    wave_lds_sync();
    m = M;
  }
}

template <int TPW>
__global__ void __launch_bounds__(256) k_zfft_mfma(PoissonArgs a, double sign) {
  extern __shared__ double2 lds[];
  const int N = a.pz.n, Q = 16 * TPW;
  const int ntile = (a.nymh + TM - 1) / TM;
  const int64_t tile = fft_tile((int64_t)ntile * a.g.nxm);
  if (tile < 0 || (a.ds && coh_int(&a.ds->halted))) return;
  const int m0 = (int)(tile % ntile) * TM, k = (int)(tile / ntile);
  double2 *S = a.S + (int64_t)k * a.g.nzm * a.ldm;
  const int tot = TM * N;
  for (int base = threadIdx.x; base < tot; base += 256 * ZLB) {   // natural order: line q at lds[q N + n]
    double2 v[ZLB];
#pragma unroll
    for (int b = 0; b < ZLB; b++) {
      const int ic = min(base + 256 * b, tot - 1);
      const int q = ic % TM, nn = ic / TM;
      v[b] = S[(int64_t)nn * a.ldm + min(m0 + q, a.nymh - 1)];
    }
#pragma unroll
    for (int b = 0; b < ZLB; b++) {
      const int idx = base + 256 * b;
      if (idx >= tot) continue;
      const int q = idx % TM, nn = idx / TM;
      lds[q * N + nn] = m0 + q < a.nymh ? v[b] : mkc(0.0, 0.0);
    }
  }
  __syncthreads();
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const int i = lane & 15, kq = lane >> 4, cc = lane & 15;
  const double2 *tw = a.pz.tw;
  dbl4_t dre[TPW], dim[TPW];
#pragma unroll
  for (int t = 0; t < TPW; t++) {
    const int rho = (w * TPW + t) * 16 + i;   // row (q, b) of the product
    const int q = rho / Q, b = rho - q * Q;
    dre[t] = dbl4_t{0.0, 0.0, 0.0, 0.0};
    dim[t] = dbl4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int aa = 4 * kk + kq;
      const double2 x = lds[q * N + Q * aa + b];
      const double2 f = twid(tw, Q * ((aa * cc) & 15), sign);   // W_16^{a c} = W_N^{Q a c}
      dre[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, f.x, dre[t], 0, 0, 0);
      dre[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-x.y, f.y, dre[t], 0, 0, 0);
      dim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, f.y, dim[t], 0, 0, 0);
      dim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, f.x, dim[t], 0, 0, 0);
    }
  }
  __syncthreads();   // every wave has read its input rows
  const bool sw = (Q & 7) == 0;
#pragma unroll
  for (int t = 0; t < TPW; t++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int rho = (w * TPW + t) * 16 + 4 * r + kq;
      const int q = rho / Q, b = rho - q * Q;
      const double2 z = cmul(mkc(dre[t][r], dim[t][r]), twid(tw, b * cc, sign));
      lds[q * N + cc * Q + swz(a.pzq.rev[b], sw)] = z;   // line (q, c), digit-reversed for the DIT
    }
  }
  __syncthreads();
  lds_fft_lines(lds + w * N, 16, a.pzq, sign, lane);   // wave w: the 16 lines of q = w
  __syncthreads();
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int q = idx % TM, nn = idx / TM;
    const int c = nn & 15, d = nn >> 4;
    if (m0 + q < a.nymh) S[(int64_t)nn * a.ldm + m0 + q] = lds[q * N + c * Q + swz(d, sw)];
  }
}

bool zfft_mfma_fits(int nzm) { return nzm == 256 || nzm == 512 || nzm == 768 || nzm == 1024; }

// ---------------------------------------------------------------- x solve
// Row k of the (m,n) mode's system, normalised to a unit diagonal
// (SolvePressureCorrection.F90:117-141):
//   amph(k) x(k-1) + x(k) + apph(k) x(k+1) = dphc(k)/(nym nzm) * acphT_b(k),
//   acphT_b = 1/(acphk(k) - ak2(m) - ak1(n)), amph = amphk*acphT_b, apph = apphk*acphT_b.
// It is eliminated in Thomas order without pivoting: u(0) = 1,
// l(k) = amph(k)/u(k-1), u(k) = 1 - l(k) apph(k-1), y(k) = r(k) - l(k) y(k-1),
// z(k) = y(k)/u(k), x(k) = z(k) - (apph(k)/u(k)) x(k+1).  The (0,0) mode is
// singular (zero row sums); its last pivot is left out and x(nxm) := 0 fixes
// the additive constant that the reference leaves to rounding.
//
// The whole x line of a mode stays in registers: NP lanes own chunks of
// Lc = ceil(nxm/NP) consecutive rows each, so the spectrum is read once and
// written once.  The pivot recurrence u(k) is nonlinear, so it is not scanned:
// the pivot at the row before each chunk, u(k0-1), comes from a table filled
// once per grid by k_xsolve_pivots (the coefficients do not depend on dt), and
// each lane continues the same recurrence through its own rows.  The two
// affine recurrences (y forward, x backward) are run per chunk from a zero
// carry and joined by a scan of the chunk maps over the NP lanes; each
// lane's rows then take the true carry in a short fix-up pass.
template <int W>
__device__ __forceinline__ double2 scan_up_c(double A, double2 B, int p) {
#pragma unroll
  for (int off = 1; off < W; off <<= 1) {
    const double A2 = __shfl_up(A, off, W), Bx = __shfl_up(B.x, off, W), By = __shfl_up(B.y, off, W);
    if (p >= off) {
      B.x = A * Bx + B.x;
      B.y = A * By + B.y;
      A = A * A2;
    }
  }
  const double cx = __shfl_up(B.x, 1, W), cy = __shfl_up(B.y, 1, W);
  return p == 0 ? make_double2(0.0, 0.0) : make_double2(cx, cy);
}
template <int W>
__device__ __forceinline__ double2 scan_down_c(double A, double2 B, int p) {
#pragma unroll
  for (int off = 1; off < W; off <<= 1) {
    const double A2 = __shfl_down(A, off, W), Bx = __shfl_down(B.x, off, W), By = __shfl_down(B.y, off, W);
    if (p + off < W) {
      B.x = A * Bx + B.x;
      B.y = A * By + B.y;
      A = A * A2;
    }
  }
  const double cx = __shfl_down(B.x, 1, W), cy = __shfl_down(B.y, 1, W);
  return p == W - 1 ? make_double2(0.0, 0.0) : make_double2(cx, cy);
}

// Pivot table: piv[mode*np + p] = u(p*lc - 1) for the chunks p >= 1 that hold
// rows (mode = n*nymh + m); one thread per mode walks the recurrence.
__global__ void __launch_bounds__(256) k_xsolve_pivots(PoissonArgs a, double *piv, int np, int lc) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nymh = a.nymh, nzm = a.g.nzm, nxm = a.g.nxm;
  if (id >= (int64_t)nymh * nzm) return;
  const int n = (int)(id / nymh), m = (int)(id - (int64_t)n * nymh);
  const double lam = a.ak2[m] + 0.0, ak1n = a.ak1[n];
  double *row = piv + id * np;
  row[0] = 1.0;
  double u_prev = 1.0, du_prev = 0.0;
  for (int k = 0; k < nxm; k++) {
    const double b = 1.0 / (a.acphk[k] - lam - ak1n);
    const double dl = a.amphk[k] * b, du = a.apphk[k] * b;
    double u = 1.0;
    if (k > 0) {
      const double l = dl / u_prev;
      u = 1.0 - l * du_prev;
    }
    const int q = (k + 1) / lc;
    if ((k + 1) - q * lc == 0 && q < np) row[q] = u;
    u_prev = u;
    du_prev = du;
  }
}

template <int NP, int LM, typename R>
__global__ void __launch_bounds__(256) k_xsolve_pressure(PoissonArgs a) {
  using C = cplx_t<R>;
  constexpr int MPB = 256 / NP;   // modes per block, consecutive in m
  extern __shared__ double xco[];   // [3][nxm]: acphk, amphk, apphk
  const int nymh = a.nymh, nzm = a.g.nzm, nxm = a.g.nxm;
  const int nmb = (nymh + MPB - 1) / MPB;
  const int64_t tile = xcd_block((int64_t)nmb * nzm);
  if (tile < 0 || (a.ds && coh_int(&a.ds->halted))) return;
  for (int i = threadIdx.x; i < nxm; i += 256) {
    xco[i] = a.acphk[i];
    xco[nxm + i] = a.amphk[i];
    xco[2 * nxm + i] = a.apphk[i];
  }
  const int n = (int)(tile / nmb), mb = (int)(tile - (int64_t)n * nmb);
  const int p = (int)threadIdx.x % NP;
  const int m = mb * MPB + (int)threadIdx.x / NP;
  const bool live = m < nymh;
  const int mc = live ? m : nymh - 1;   // lanes past the last mode solve a valid one and store nothing
  const int lc = (nxm + NP - 1) / NP;
  const int k0 = p * lc;
  const int cnt = max(0, min(lc, nxm - k0));
  const int64_t stride = (int64_t)nzm * a.ldm;
  C *S = reinterpret_cast<C *>(a.S) + (int64_t)n * a.ldm + mc;
  C *Sp = reinterpret_cast<C *>(a.S) + (int64_t)n * a.ldm + min(m, a.ldm - 1);   // padded wavenumbers: zeros
  double2 y[LM];
#pragma unroll
  for (int t = 0; t < LM; t++) y[t] = widen(S[(int64_t)min(k0 + min(t, lc - 1), nxm - 1) * stride]);
  const double lam = a.ak2[mc] + 0.0, ak1n = a.ak1[n];
  const double nrm = (double)(nzm * a.g.nym);
  const bool zero_mode = (mc == 0 && n == 0);
  double u0 = 1.0, du0 = 0.0;   // u and apph of row k0-1
  if (p > 0 && cnt > 0) {
    u0 = a.piv[((int64_t)n * nymh + mc) * NP + p];
    const double bp = 1.0 / (a.acphk[k0 - 1] - lam - ak1n);
    du0 = a.apphk[k0 - 1] * bp;
  }
  __syncthreads();
  const double *ac = xco, *am = xco + nxm, *ap = xco + 2 * nxm;
  // pass 1: pivots u(k) of the chunk (kept in w) and the chunk's y from a zero carry
  double w[LM];
  double Ay = 1.0;
  double2 yl = make_double2(0.0, 0.0);
  {
    double u_prev = u0, du_prev = du0;
#pragma unroll
    for (int t = 0; t < LM; t++) {
      w[t] = 1.0;
      if (t < cnt) {
        const int k = k0 + t;
        const double b = 1.0 / (ac[k] - lam - ak1n);
        const double dl = am[k] * b, du = ap[k] * b;
        double2 r = y[t];
        r.x = r.x / nrm * b;
        r.y = r.y / nrm * b;
        double u = 1.0;
        if (k > 0) {
          const double l = dl / u_prev;
          u = 1.0 - l * du_prev;
          r.x = r.x - l * yl.x;
          r.y = r.y - l * yl.y;
          Ay *= -l;
        }
        yl = r;
        y[t] = r;
        w[t] = u;
        u_prev = u;
        du_prev = du;
      }
    }
  }
  // pass 2: y(k0-1) from the lanes before; z = y/u; w becomes apph/u
  {
    double2 d = scan_up_c<NP>(Ay, yl, p);
    double u_prev = u0, du_prev = du0;
#pragma unroll
    for (int t = 0; t < LM; t++) {
      if (t < cnt) {
        const int k = k0 + t;
        const double b = 1.0 / (ac[k] - lam - ak1n);
        const double dl = am[k] * b, du = ap[k] * b;
        const double u = w[t];
        if (k > 0) {
          const double l = dl / u_prev;
          d.x = -l * d.x;
          d.y = -l * d.y;
        }
        const double uinv = (zero_mode && k == nxm - 1) ? 0.0 : 1.0 / u;
        y[t].x = (y[t].x + d.x) * uinv;
        y[t].y = (y[t].y + d.y) * uinv;
        w[t] = du * uinv;
        u_prev = u;
        du_prev = du;
      }
    }
  }
  // pass 3: back substitution of the chunk from a zero carry
  double Ax = 1.0;
  double2 xl = make_double2(0.0, 0.0);
#pragma unroll
  for (int t = LM - 1; t >= 0; t--) {
    if (t < cnt) {
      xl.x = y[t].x - w[t] * xl.x;
      xl.y = y[t].y - w[t] * xl.y;
      y[t] = xl;
      Ax *= -w[t];
    }
  }
  // pass 4: x(k0+cnt) from the lanes after, and the stores
  double2 e = scan_down_c<NP>(Ax, xl, p);
#pragma unroll
  for (int t = LM - 1; t >= 0; t--) {
    if (t < cnt) {
      e.x = -w[t] * e.x;
      e.y = -w[t] * e.y;
      if (live) S[(int64_t)(k0 + t) * stride] = narrow<C>(make_double2(y[t].x + e.x, y[t].y + e.y));
      else if (m < a.ldm) Sp[(int64_t)(k0 + t) * stride] = narrow<C>(make_double2(0.0, 0.0));
    }
  }
}

// Lanes per mode and register rows per lane of the x solve for nxm rows.
void xsolve_shape(int nxm, int *np, int *lm) {
  if (nxm <= 32 * 4) { *np = 32; *lm = 4; }
  else if (nxm <= 32 * 8) { *np = 32; *lm = 8; }
  else if (nxm <= 32 * 12) { *np = 32; *lm = 12; }
  else if (nxm <= 64 * 12) { *np = 64; *lm = 12; }
  else { *np = 64; *lm = 20; }
}
int xsolve_max_rows() { return 64 * 20; }

// ---------------------------------------------------------------- y inverse
template <typename R>
__global__ void __launch_bounds__(256) k_yfft_inverse(PoissonArgs a) {
  using C = cplx_t<R>;
  extern __shared__ double2 lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  const Geom g = a.g;
  const RealPlan &py = a.py;
  const int H = py.H, N = py.N;
  const int ntile = g.ldx / TK;
  const int64_t tile = fft_tile((int64_t)ntile * g.nzm);
  if (tile < 0) return;
  if (a.ds && coh_int(&a.ds->halted)) return;
  const int s0 = (int)(tile % ntile) * TK, i = (int)(tile / ntile);
  const bool hsw = (H & 7) == 0;
  const int nymh = a.nymh;
  // Loads: 32 lanes per real line q, points t = lane + 32 u.
  static_assert(TK * 32 == 256, "one real line per 32 lanes");
  {
    const int q = (int)threadIdx.x >> 5, lt = (int)threadIdx.x & 31;
    const int kq = k_of_pos(s0 + q, g.L);
    const bool live = kq < g.nxm;
    const C *X = reinterpret_cast<const C *>(a.S) + ((int64_t)min(kq, g.nxm - 1) * g.nzm + i) * a.ldm;
    for (int t0 = lt; t0 < H; t0 += 32 * YILB) {
      C xa[YILB], xb[YILB], pw[YILB];
      int rv[YILB];
#pragma unroll
      for (int b = 0; b < YILB; b++) {   // all loads of the batch first
        const int t = min(t0 + 32 * b, H - 1);
        if (py.half) {
          xa[b] = X[t];
          xb[b] = X[H - t];
          pw[b] = cvt<R>(py.post[t]);
        } else {
          xa[b] = X[min(t, nymh - 1)];
          xb[b] = X[min(N - t, nymh - 1)];
          pw[b] = mkc((R)1, (R)0);
        }
        rv[b] = py.c.rev[t];
      }
#pragma unroll
      for (int b = 0; b < YILB; b++) {
        const int t = t0 + 32 * b;
        if (t >= H) continue;
        C z;
        if (py.half) {
          C xm = xa[b], xh = xb[b];
          if (t == 0) { xm.y = (R)0; xh.y = (R)0; }  // DC and Nyquist: imaginary parts ignored
          C xc = conj2(xh);
          C E = cadd(xm, xc);
          C O = cmul(csub(xm, xc), conj2(pw[b]));
          z = mkc(E.x - O.y, E.y + O.x);  // E + i O
        } else {
          if (t < nymh) {
            z = xa[b];
            if (t == 0) z.y = (R)0;
          } else {
            z = conj2(xb[b]);
          }
        }
        lds[q * H + swz(rv[b], hsw)] = live ? z : mkc((R)0, (R)0);
      }
    }
  }
  __syncthreads();
  fft_wave<R, TK / 4>(lds + (threadIdx.x / 64) * (TK / 4) * H, py.c, (R)1, threadIdx.x % 64);
  __syncthreads();
  for (int idx = threadIdx.x; idx < TK * N; idx += blockDim.x) {
    const int q = idx % TK, j = idx / TK;
    // positions k >= nxm (their lines were zero) store 0, so that every
    // 128-byte line of the column is written whole
    const int s = s0 + q;
    double v;
    if (py.half) {
      C z = lds[q * H + swz(j >> 1, hsw)];
      v = (j & 1) ? z.y : z.x;
    } else {
      v = lds[q * H + swz(j, hsw)].x;
    }
    a.dph[col_off(j, i, g) + s] = v;
  }
}

// ---------------------------------------------------------------- launchers
int fft_max_len() { return 4096; }

// Dynamic LDS above 64 KiB must be opted into per kernel (gfx950: 160 KiB per workgroup).
template <typename R>
static hipError_t fft_configure_lds_r(size_t bytes) {
  hipError_t e = hipFuncSetAttribute((const void *)k_yfft_forward<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void *)k_yfft_inverse<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void *)k_zfft<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}
hipError_t fft_configure_lds(size_t bytes) {
  hipError_t e = fft_configure_lds_r<double>(bytes);
  if (e != hipSuccess) return e;
  e = fft_configure_lds_r<float>(bytes);
  if (e != hipSuccess) return e;
  const void *mk[] = {(const void *)k_zfft_mfma<1>, (const void *)k_zfft_mfma<2>, (const void *)k_zfft_mfma<3>,
                      (const void *)k_zfft_mfma<4>};
  for (const void *f : mk)
    if ((e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)) != hipSuccess) return e;
  return hipSuccess;
}

static dim3 tile_grid(int64_t total) { return dim3((unsigned)(((total + 7) / 8) * 8)); }
static size_t cplx_bytes(const PoissonArgs &a) { return a.spec32 ? sizeof(float2) : sizeof(double2); }

hipError_t launch_yfft_forward(const PoissonArgs &a, hipStream_t st) {
  dim3 grid = tile_grid((int64_t)(a.g.ldx / TK) * a.g.nzm);
  size_t shm = cplx_bytes(a) * (size_t)TK * a.py.H;
  if (a.spec32) hipLaunchKernelGGL(k_yfft_forward<float>, grid, dim3(256), shm, st, a);
  else hipLaunchKernelGGL(k_yfft_forward<double>, grid, dim3(256), shm, st, a);
  return hipGetLastError();
}
hipError_t launch_zfft(const PoissonArgs &a, int sign, hipStream_t st) {
  dim3 grid = tile_grid((int64_t)((a.nymh + TM - 1) / TM) * a.g.nxm);
  size_t shm = cplx_bytes(a) * (size_t)TM * a.pz.n;
  if (a.zmf && !a.spec32) {
    switch (a.pz.n / 256) {
      case 1: hipLaunchKernelGGL(k_zfft_mfma<1>, grid, dim3(256), shm, st, a, (double)sign); break;
      case 2: hipLaunchKernelGGL(k_zfft_mfma<2>, grid, dim3(256), shm, st, a, (double)sign); break;
      case 3: hipLaunchKernelGGL(k_zfft_mfma<3>, grid, dim3(256), shm, st, a, (double)sign); break;
      default: hipLaunchKernelGGL(k_zfft_mfma<4>, grid, dim3(256), shm, st, a, (double)sign); break;
    }
    return hipGetLastError();
  }
  if (a.spec32) hipLaunchKernelGGL(k_zfft<float>, grid, dim3(256), shm, st, a, (float)sign);
  else hipLaunchKernelGGL(k_zfft<double>, grid, dim3(256), shm, st, a, (double)sign);
  return hipGetLastError();
}
hipError_t launch_xsolve_pivots(const PoissonArgs &a, double *piv, hipStream_t st) {
  int np, lm;
  xsolve_shape(a.g.nxm, &np, &lm);
  const int64_t n = (int64_t)a.nymh * a.g.nzm;
  hipLaunchKernelGGL(k_xsolve_pivots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, piv, np,
                     (a.g.nxm + np - 1) / np);
  return hipGetLastError();
}
template <int NP, int LM>
static void launch_xsolve_r(const PoissonArgs &a, hipStream_t st) {
  const int64_t nmb = (a.nymh + 256 / NP - 1) / (256 / NP);
  const dim3 grid((unsigned)(((nmb * a.g.nzm + 7) / 8) * 8));
  const size_t shm = sizeof(double) * 3 * a.g.nxm;
  if (a.spec32) hipLaunchKernelGGL((k_xsolve_pressure<NP, LM, float>), grid, dim3(256), shm, st, a);
  else hipLaunchKernelGGL((k_xsolve_pressure<NP, LM, double>), grid, dim3(256), shm, st, a);
}
hipError_t launch_xsolve_pressure(const PoissonArgs &a, hipStream_t st) {
  int np, lm;
  xsolve_shape(a.g.nxm, &np, &lm);
  if (np == 32 && lm == 4) launch_xsolve_r<32, 4>(a, st);
  else if (np == 32 && lm == 8) launch_xsolve_r<32, 8>(a, st);
  else if (np == 32) launch_xsolve_r<32, 12>(a, st);
  else if (lm == 12) launch_xsolve_r<64, 12>(a, st);
  else launch_xsolve_r<64, 20>(a, st);
  return hipGetLastError();
}
hipError_t launch_yfft_inverse(const PoissonArgs &a, hipStream_t st) {
  dim3 grid = tile_grid((int64_t)(a.g.ldx / TK) * a.g.nzm);
  size_t shm = cplx_bytes(a) * (size_t)TK * a.py.H;
  if (a.spec32) hipLaunchKernelGGL(k_yfft_inverse<float>, grid, dim3(256), shm, st, a);
  else hipLaunchKernelGGL(k_yfft_inverse<double>, grid, dim3(256), shm, st, a);
  return hipGetLastError();
}

}  // namespace afid
