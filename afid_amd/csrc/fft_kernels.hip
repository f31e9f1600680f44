// fft_kernels.hip -- pressure correction (SolvePressureCorrection.F90:12-154)
// as four kernels over the spectral array S[k][n][m]:
//   1. y real-to-complex FFT of dph (optionally computing dph as the local
//      divergence of u*, CalcLocalDivergence.F90, on the fly),
//   2. z complex FFT (forward, and later backward),
//   3. per-(m,n) tridiagonal solve in x (the zgttrf/zgttrs loop, :117-141,
//      including the 1/(nym*nzm) normalisation of :105),
//   4. y complex-to-real FFT back into dph.
// FFTs are in-LDS mixed-radix (4,2,3,5) decimation in time, unnormalised,
// forward sign -1, with FFTW's c2r convention (imaginary parts of the DC and
// Nyquist bins ignored).  A real line of even length N is transformed as a
// complex line of N/2 points plus a post-twiddle.
#include "afid_internal.h"
#include "kernels.h"

namespace afid {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }
// multiply by s*i (s = +1/-1)
__device__ __forceinline__ double2 mul_si(double2 a, double s) { return make_double2(-s * a.y, s * a.x); }

__device__ __forceinline__ double2 twid(const double2 *tw, int idx, double sign) {
  double2 t = tw[idx];
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

// In-place DIT stages over `nlines` lines of length pl.n stored contiguously
// in LDS from `lines`, input already in digit-reversed order, computed by ONE
// wavefront (lane = 0..63): the stages synchronise within the wave only, so
// the waves of a block progress independently.  sign = -1 forward, +1 backward.
// Stage with radix r and span m combines, for every block b and offset u < m,
// the r points b*m*r + u + v*m (v < r) after twiddling point v by W_{mr}^{uv}.
__device__ __forceinline__ void lds_fft_wave(double2 *lines, int nlines, const FFTPlan &pl, double sign,
                                             int lane) {
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
    for (int q = 0; q < nlines; q++) {
      double2 *line = lines + q * n;
      for (int w = lane; w < nb; w += 64) {
        const int bb = m == 1 ? w : (int)__umulhi((unsigned)w, mdiv);
        const int u = w - bb * m;
        const int e0 = bb * M + u;
        double2 *base = line;
        const int t1i = u * tws;
        // element v of the butterfly: base[X(v)], X(v) = swz(e0 + v*m)
#define X(v) swz(e0 + (v) * m, sw)
        if (r == 2) {
          double2 x0 = base[X(0)], x1 = cmul(base[X(1)], twid(tw, t1i, sign));
          base[X(0)] = cadd(x0, x1);
          base[X(1)] = csub(x0, x1);
        } else if (r == 4) {
          double2 x0 = base[X(0)];
          double2 x1 = cmul(base[X(1)], twid(tw, t1i, sign));
          double2 x2 = cmul(base[X(2)], twid(tw, 2 * t1i, sign));
          double2 x3 = cmul(base[X(3)], twid(tw, 3 * t1i, sign));
          double2 a0 = cadd(x0, x2), a1 = csub(x0, x2);
          double2 b0 = cadd(x1, x3), b1 = mul_si(csub(x1, x3), sign);
          base[X(0)] = cadd(a0, b0);
          base[X(1)] = cadd(a1, b1);
          base[X(2)] = csub(a0, b0);
          base[X(3)] = csub(a1, b1);
        } else if (r == 3) {
          const double s3 = 0.86602540378443864676372317075294;  // sin(2 pi/3)
          double2 x0 = base[X(0)];
          double2 x1 = cmul(base[X(1)], twid(tw, t1i, sign));
          double2 x2 = cmul(base[X(2)], twid(tw, 2 * t1i, sign));
          double2 t1 = cadd(x1, x2);
          double2 t2 = mul_si(csub(x1, x2), sign * s3);
          double2 a = make_double2(x0.x - 0.5 * t1.x, x0.y - 0.5 * t1.y);
          base[X(0)] = cadd(x0, t1);
          base[X(1)] = cadd(a, t2);
          base[X(2)] = csub(a, t2);
        } else {  // r == 5
          const double c1 = 0.30901699437494742410229341718282;   // cos(2 pi/5)
          const double c2 = -0.80901699437494742410229341718282;  // cos(4 pi/5)
          const double s1 = 0.95105651629515357211643933337938;   // sin(2 pi/5)
          const double s2 = 0.58778525229247312916870595463907;   // sin(4 pi/5)
          double2 x0 = base[X(0)];
          double2 x1 = cmul(base[X(1)], twid(tw, t1i, sign));
          double2 x2 = cmul(base[X(2)], twid(tw, 2 * t1i, sign));
          double2 x3 = cmul(base[X(3)], twid(tw, 3 * t1i, sign));
          double2 x4 = cmul(base[X(4)], twid(tw, 4 * t1i, sign));
          double2 t1 = cadd(x1, x4), t2 = cadd(x2, x3);
          double2 t3 = csub(x1, x4), t4 = csub(x2, x3);
          double2 a1 = make_double2(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
          double2 a2 = make_double2(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
          double2 b1 = mul_si(make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), sign);
          double2 b2 = mul_si(make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), sign);
          base[X(0)] = cadd(x0, cadd(t1, t2));
          base[X(1)] = cadd(a1, b1);
          base[X(2)] = cadd(a2, b2);
          base[X(3)] = csub(a2, b2);
          base[X(4)] = csub(a1, b1);
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
// lines 2w and 2w+1 of the tile
__global__ void __launch_bounds__(256) k_yfft_forward(PoissonArgs a) {
  extern __shared__ double2 lds[];
  const Geom g = a.g;
  const RealPlan &py = a.py;
  const int H = py.H, N = py.N;
  const int ntile = g.ldx / TK;
  const int64_t tile = fft_tile((int64_t)ntile * g.nzm);
  if (tile < 0) return;
  const int s0 = (int)(tile % ntile) * TK, i = (int)(tile / ntile);
  const bool hsw = (H & 7) == 0;
  const double usdtal = 1.0 / (a.sc.dt * a.sc.al);
  for (int idx = threadIdx.x; idx < TK * N; idx += blockDim.x) {
    const int q = idx % TK, j = idx / TK;
    const int s = s0 + q, k = k_of_pos(s, g.L);
    double v = 0.0;
    if (k < g.nxm) {
      const int64_t C = col_off(j, i, g);
      if (a.from_velocity) {
        const int64_t CJP = col_off(wrapp(j + 1, g.nym), i, g), CIP = col_off(j, wrapp(i + 1, g.nzm), g);
        const int skp = pos_of_k(k + 1, g.L);
        double dq = (a.f.vz[CIP + s] - a.f.vz[C + s]) * a.sc.dz + (a.f.vy[CJP + s] - a.f.vy[C + s]) * a.sc.dy +
                    (a.f.vx[C + skp] - a.f.vx[C + s]) * a.udx3m[s];
        v = dq * usdtal;
      } else {
        v = a.dph[C + s];
      }
    }
    if (py.half) {
      double2 *e = &lds[q * H + swz(py.c.rev[j >> 1], hsw)];
      if (j & 1) e->y = v;
      else e->x = v;
    } else {
      lds[q * H + swz(py.c.rev[j], hsw)] = make_double2(v, 0.0);
    }
  }
  __syncthreads();
  lds_fft_wave(lds + (threadIdx.x / 64) * (TK / 4) * H, TK / 4, py.c, -1.0, threadIdx.x % 64);
  __syncthreads();
  const int nymh = a.nymh;
  for (int idx = threadIdx.x; idx < TK * nymh; idx += blockDim.x) {
    const int q = idx / nymh, m = idx - q * nymh;
    const int k = k_of_pos(s0 + q, g.L);
    if (k >= g.nxm) continue;
    double2 X;
    if (py.half) {
      double2 zm = lds[q * H + swz(m == H ? 0 : m, hsw)];
      double2 zc = conj2(lds[q * H + swz(m == 0 ? 0 : H - m, hsw)]);
      double2 E = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      double2 D = csub(zm, zc);
      double2 O = make_double2(0.5 * D.y, -0.5 * D.x);  // D / (2i)
      X = cadd(E, cmul(py.post[m], O));
    } else {
      X = lds[q * H + swz(m, hsw)];
    }
    a.S[((int64_t)k * g.nzm + i) * a.ldm + m] = X;
  }
}

// ---------------------------------------------------------------- z transforms
// tile: (TM wavenumbers m, k); wave w transforms line w
__global__ void __launch_bounds__(256) k_zfft(PoissonArgs a, double sign) {
  extern __shared__ double2 lds[];
  const int N = a.pz.n;
  const int ntile = (a.nymh + TM - 1) / TM;
  const int64_t tile = fft_tile((int64_t)ntile * a.g.nxm);
  if (tile < 0) return;
  const int m0 = (int)(tile % ntile) * TM, k = (int)(tile / ntile);
  const bool zsw = (N & 7) == 0;
  double2 *S = a.S + (int64_t)k * a.g.nzm * a.ldm;
  for (int idx = threadIdx.x; idx < TM * N; idx += blockDim.x) {
    const int q = idx % TM, nn = idx / TM;
    const int m = m0 + q;
    double2 v = make_double2(0.0, 0.0);
    if (m < a.nymh) v = S[(int64_t)nn * a.ldm + m];
    lds[q * N + swz(a.pz.rev[nn], zsw)] = v;
  }
  __syncthreads();
  lds_fft_wave(lds + (threadIdx.x / 64) * N, 1, a.pz, sign, threadIdx.x % 64);
  __syncthreads();
  for (int idx = threadIdx.x; idx < TM * N; idx += blockDim.x) {
    const int q = idx % TM, nn = idx / TM;
    const int m = m0 + q;
    if (m < a.nymh) S[(int64_t)nn * a.ldm + m] = lds[q * N + swz(nn, zsw)];
  }
}

// ---------------------------------------------------------------- x solve
// One thread per (m,n).  Row k of the normalised system (diagonal 1):
//   amph(k) x(k-1) + x(k) + apph(k) x(k+1) = dphc(k)/(nym nzm) * acphT_b(k),
//   acphT_b = 1/(acphk(k) - ak2(m) - ak1(n)), amph = amphk*acphT_b, apph = apphk*acphT_b.
// The (0,0) mode is singular (zero row sums); its last pivot is left out and
// x(nxm) := 0 fixes the additive constant that the reference leaves to rounding.
__global__ void __launch_bounds__(256) k_xsolve_pressure(PoissonArgs a) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nymh = a.nymh, nzm = a.g.nzm, nxm = a.g.nxm;
  if (id >= (int64_t)nymh * nzm) return;
  const int n = (int)(id / nymh), m = (int)(id - (int64_t)n * nymh);
  const int64_t stride = (int64_t)nzm * a.ldm;
  double2 *S = a.S + (int64_t)n * a.ldm + m;
  double *cp = a.cp + (int64_t)n * a.ldm + m;
  const double lam = a.ak2[m] + 0.0;  // ak2 indexed by y wavenumber
  const double ak1n = a.ak1[n];
  const double nrm = (double)(nzm * a.g.nym);
  const bool zero_mode = (m == 0 && n == 0);
  double u_prev = 1.0, du_prev = 0.0;
  double2 y_prev = make_double2(0.0, 0.0);
  for (int k = 0; k < nxm; k++) {
    const double b = 1.0 / (a.acphk[k] - lam - ak1n);
    const double dl = a.amphk[k] * b, du = a.apphk[k] * b;
    double2 r = S[k * stride];
    r.x = r.x / nrm * b;
    r.y = r.y / nrm * b;
    double u = 1.0;
    if (k > 0) {
      const double l = dl / u_prev;
      u = 1.0 - l * du_prev;
      r.x = r.x - l * y_prev.x;
      r.y = r.y - l * y_prev.y;
    }
    y_prev = r;
    u_prev = u;
    du_prev = du;
    double ui = (zero_mode && k == nxm - 1) ? 0.0 : 1.0 / u;
    S[k * stride] = make_double2(r.x * ui, r.y * ui);
    cp[k * stride] = du * ui;
  }
  double2 x = S[(nxm - 1) * stride];
  for (int k = nxm - 2; k >= 0; k--) {
    double2 yk = S[k * stride];
    const double c = cp[k * stride];
    x = make_double2(yk.x - c * x.x, yk.y - c * x.y);
    S[k * stride] = x;
  }
}

// ---------------------------------------------------------------- y inverse
__global__ void __launch_bounds__(256) k_yfft_inverse(PoissonArgs a) {
  extern __shared__ double2 lds[];
  const Geom g = a.g;
  const RealPlan &py = a.py;
  const int H = py.H, N = py.N;
  const int ntile = g.ldx / TK;
  const int64_t tile = fft_tile((int64_t)ntile * g.nzm);
  if (tile < 0) return;
  const int s0 = (int)(tile % ntile) * TK, i = (int)(tile / ntile);
  const bool hsw = (H & 7) == 0;
  const int nymh = a.nymh;
  for (int idx = threadIdx.x; idx < TK * H; idx += blockDim.x) {
    const int q = idx / H, t = idx - q * H;
    const int k = k_of_pos(s0 + q, g.L);
    double2 z = make_double2(0.0, 0.0);
    if (k < g.nxm) {
      const double2 *X = a.S + ((int64_t)k * g.nzm + i) * a.ldm;
      if (py.half) {
        double2 xm = X[t], xh = X[H - t];
        if (t == 0) { xm.y = 0.0; xh.y = 0.0; }  // DC and Nyquist: imaginary parts ignored
        double2 xc = conj2(xh);
        double2 E = cadd(xm, xc);
        double2 O = cmul(csub(xm, xc), conj2(py.post[t]));
        z = make_double2(E.x - O.y, E.y + O.x);  // E + i O
      } else {
        if (t < nymh) {
          z = X[t];
          if (t == 0) z.y = 0.0;
        } else {
          z = conj2(X[N - t]);
        }
      }
    }
    lds[q * H + swz(py.c.rev[t], hsw)] = z;
  }
  __syncthreads();
  lds_fft_wave(lds + (threadIdx.x / 64) * (TK / 4) * H, TK / 4, py.c, +1.0, threadIdx.x % 64);
  __syncthreads();
  for (int idx = threadIdx.x; idx < TK * N; idx += blockDim.x) {
    const int q = idx % TK, j = idx / TK;
    const int s = s0 + q, k = k_of_pos(s, g.L);
    if (k >= g.nxm) continue;
    double v;
    if (py.half) {
      double2 z = lds[q * H + swz(j >> 1, hsw)];
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
hipError_t fft_configure_lds(size_t bytes) {
  hipError_t e = hipFuncSetAttribute((const void *)k_yfft_forward, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void *)k_yfft_inverse, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void *)k_zfft, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

static dim3 tile_grid(int64_t total) { return dim3((unsigned)(((total + 7) / 8) * 8)); }

hipError_t launch_yfft_forward(const PoissonArgs &a, hipStream_t st) {
  dim3 grid = tile_grid((int64_t)(a.g.ldx / TK) * a.g.nzm);
  size_t shm = sizeof(double2) * (size_t)TK * a.py.H;
  hipLaunchKernelGGL(k_yfft_forward, grid, dim3(256), shm, st, a);
  return hipGetLastError();
}
hipError_t launch_zfft(const PoissonArgs &a, int sign, hipStream_t st) {
  dim3 grid = tile_grid((int64_t)((a.nymh + TM - 1) / TM) * a.g.nxm);
  size_t shm = sizeof(double2) * (size_t)TM * a.pz.n;
  hipLaunchKernelGGL(k_zfft, grid, dim3(256), shm, st, a, (double)sign);
  return hipGetLastError();
}
hipError_t launch_xsolve_pressure(const PoissonArgs &a, hipStream_t st) {
  int64_t n = (int64_t)a.nymh * a.g.nzm;
  hipLaunchKernelGGL(k_xsolve_pressure, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_yfft_inverse(const PoissonArgs &a, hipStream_t st) {
  dim3 grid = tile_grid((int64_t)(a.g.ldx / TK) * a.g.nzm);
  size_t shm = sizeof(double2) * (size_t)TK * a.py.H;
  hipLaunchKernelGGL(k_yfft_inverse, grid, dim3(256), shm, st, a);
  return hipGetLastError();
}

}  // namespace afid
