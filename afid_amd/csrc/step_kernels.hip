// step_kernels.hip -- time-step control on the device (DevStep), so that whole
// steps of main.F90 run without a host round trip and can be replayed as one
// captured graph:
//   k_step_begin  CalcMaxCFL result (from k_red_cfl, or folded into the last
//                 correction of the previous step) -> dt and error code (main.F90:169-189),
//                 substep scalars al, ga, ro, beta (TimeMarcher.F90:22-30)
//   k_factor      the wall-normal systems of each substep for that dt
//                 (SolveImpEqnUpdate_X.F90:27-44, _YZ.F90:25-38, _Temp.F90:31-47)
//   k_step_end    time += dt, ntime += 1, halt after an error code
// A factorisation that would need a row interchange halts the cell in
// k_factor, before the substeps of that step run.
#include "afid_internal.h"
#include "kernels.h"

namespace afid {

__global__ void k_step_begin(DevStep *ds, const double *cfl_out, StepConsts k) {
  // every field is read with coherent loads: earlier kernels of the step
  // graph (and of the previous replay) wrote them
  if (threadIdx.x != 0 || coh_int(&ds->halted)) return;
  const double cfl = cfl_out ? coh_dbl(cfl_out) : __longlong_as_double((long long)coh_u64(&ds->cflbits));
  ds->cflbits = 0;   // the last correction of this step accumulates the next step's value
  const double instcfl = fmax(0.00000001, cfl);   // CalcMaxCFL floor
  const double dtmin = coh_dbl(&ds->dtmin), dtmax = coh_dbl(&ds->dtmax), limitcfl = coh_dbl(&ds->limitcfl);
  double dt = coh_dbl(&ds->dtfix);
  int err = 0;
  if (coh_int(&ds->idtv)) {
    if (coh_int(&ds->ntime) > 1) {
      dt = instcfl < 1.0e-8 ? dtmax : limitcfl / instcfl;
      if (dt > dtmax) dt = dtmax;
    } else {
      dt = dtmin;
    }
    if (dt < dtmin) err = 166;
  } else {
    if (instcfl * dt > limitcfl) err = 165;
  }
  ds->dt = dt;
  ds->errorcode = err;
  for (int ns = 0; ns < k.nsst; ns++) {
    StepScalars s;
    s.al = k.alm[ns]; s.ga = k.gam[ns]; s.ro = k.rom[ns]; s.dt = dt; s.beta = dt / k.ren * 0.5;
    s.alre = s.al / k.ren;
    s.alpec = s.al / k.pec;
    s.udy4 = k.dy * 0.25;
    s.udz4 = k.dz * 0.25;
    s.udyq_re = k.dyq / k.ren;
    s.udzq_re = k.dzq / k.ren;
    s.udyq_pe = k.dyq / k.pec;
    s.udzq_pe = k.dzq / k.pec;
    s.dy = k.dy; s.dz = k.dz; s.dyq = k.dyq; s.dzq = k.dzq;
    s.udy_al = k.dy * s.al;
    s.tempbp = k.tempbp;
    s.temptp = k.temptp;
    s.use_ru = s.ro != 0.0;
    ds->sc[ns] = s;
  }
}

__global__ void k_step_end(DevStep *ds) {
  if (threadIdx.x != 0 || coh_int(&ds->halted)) return;
  ds->time = coh_dbl(&ds->time) + coh_dbl(&ds->dt);
  ds->ntime = coh_int(&ds->ntime) + 1;
  ds->steps_done = coh_int(&ds->steps_done) + 1;
  if (coh_int(&ds->errorcode) != 0) ds->halted = 1;
}

// One block.  Per substep and system: coefficients (all threads), LU in
// netlib order (one thread per system), 1/u and e, chunk products G/H (one
// thread per system and chunk), then the interleaved factor rows.
enum { W_DL = 0, W_DU, W_SC, W_L, W_U, W_E, W_N };
__global__ void __launch_bounds__(256) k_factor(FactorArgs a) {
  if (coh_int(&a.ds->halted)) return;
  const Geom g = a.g;
  const int nx = g.nx, nxm = g.nxm, ldx = g.ldx, L = g.L;
  const double *ac3ck = a.mc, *am3ssk = a.mc + nx, *ap3ssk = a.mc + 2 * nx, *ac3sk = a.mc + 3 * nx;
  const double *am3sk = a.mc + 4 * nx, *ap3sk = a.mc + 5 * nx, *ac3ssk = a.mc + 6 * nx;
  const double dt = coh_dbl(&a.ds->dt);
  auto W = [&](int q, int r) { return a.work + ((size_t)q * W_N + r) * nx; };
  auto nsys = [&](int q) { return q == SYS_YZ ? nxm : nx; };
  for (int ns = 0; ns < a.k.nsst; ns++) {
    const double al = coh_dbl(&a.ds->sc[ns].al);
    const double beta = dt / a.k.ren * 0.5;
    for (int idx = threadIdx.x; idx < SYS_N * nx; idx += blockDim.x) {
      const int q = idx / nx, k = idx - q * nx, kc = k + 1;
      double dl = 0.0, du = 0.0, sc = 0.0;
      if (q == SYS_X && kc >= 2 && kc <= nxm) {
        const double betadx = beta * al;
        const double ackl_b = 1.0 / (1.0 - ac3ck[k] * betadx);
        dl = -am3ssk[k] * betadx * ackl_b;
        du = -ap3ssk[k] * betadx * ackl_b;
        sc = 1.0 / (1.0 - ac3ck[k] * betadx);
      } else if (q == SYS_YZ && kc <= nxm) {
        const double betadx = beta * al;
        const double ackl_b = 1.0 / (1.0 - ac3sk[k] * betadx);
        dl = kc > 1 ? -am3sk[k] * betadx * ackl_b : 0.0;
        du = kc < nxm ? -ap3sk[k] * betadx * ackl_b : 0.0;
        sc = 1.0 / (1.0 - ac3sk[k] * betadx);
      } else if (q == SYS_T && kc >= 2 && kc <= nxm) {
        const double betadx = 0.5 * al * dt / a.k.pec;
        const double ackl_b = 1.0 / (1. - ac3ssk[k] * betadx);
        dl = -am3ssk[k] * betadx * ackl_b;
        du = -ap3ssk[k] * betadx * ackl_b;
        sc = 1.0 / (1.0 - ac3ssk[k] * betadx);
      }
      W(q, W_DL)[k] = dl;
      W(q, W_DU)[k] = du;
      W(q, W_SC)[k] = sc;
    }
    __syncthreads();
    if (threadIdx.x < SYS_N) {
      // The LU recurrence, one thread per system.  Its operands are loaded in
      // batches of FB rows ahead of the serial chain, which then runs in
      // registers (one memory round trip per batch instead of per row).
      constexpr int FB = 32;
      const int q = threadIdx.x, n = nsys(q);
      const double *dl = W(q, W_DL), *du = W(q, W_DU);
      double *l = W(q, W_L), *u = W(q, W_U);
      l[0] = 0.0;
      u[0] = 1.0;
      double up = 1.0;   // u(k-1)
      bool piv = false;
      for (int k0 = 1; k0 < n; k0 += FB) {
        double d[FB], e[FB], lo[FB], uo[FB];
#pragma unroll
        for (int t = 0; t < FB; t++) {
          const int k = min(k0 + t, n - 1);
          d[t] = dl[k];
          e[t] = du[k - 1];
        }
#pragma unroll
        for (int t = 0; t < FB; t++) {
          if (k0 + t < n) {
            // dgttrf would interchange rows here: the pivot-free LU is refused,
            // and the cell halts before any substep of this step changes a field
            piv = piv || fabs(up) < fabs(d[t]);
            lo[t] = d[t] / up;
            up = 1.0 - lo[t] * e[t];
            uo[t] = up;
          }
        }
#pragma unroll
        for (int t = 0; t < FB; t++) {
          if (k0 + t < n) {
            l[k0 + t] = lo[t];
            u[k0 + t] = uo[t];
          }
        }
      }
      if (piv) {
        a.ds->pivot = 1;
        a.ds->halted = 1;
      }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < SYS_N * nx; idx += blockDim.x) {
      const int q = idx / nx, k = idx - q * nx, n = nsys(q);
      if (k >= n) continue;
      double *u = W(q, W_U);
      const double rinv = 1.0 / u[k];
      W(q, W_E)[k] = k + 1 < n ? W(q, W_DU)[k] * rinv : 0.0;
      u[k] = rinv;   // W_U now holds 1/u
    }
    __syncthreads();
    double *rows = a.fac[ns];
    for (int idx = threadIdx.x; idx < SYS_N * P; idx += blockDim.x) {
      const int q = idx / P, p = idx - q * P, n = nsys(q);
      double *fr = rows + (size_t)(q * FR_N) * ldx;
      const int k0 = p * L, k1 = min(n, k0 + L);
      double gg = 1.0;
      for (int k = k0; k < k1; k++) {
        gg *= -W(q, W_L)[k];
        fr[FR_G * ldx + pos_of_k(k, L)] = gg;
      }
      double hh = 1.0;
      for (int k = k1 - 1; k >= k0; k--) {
        hh *= -W(q, W_E)[k];
        fr[FR_H * ldx + pos_of_k(k, L)] = hh;
      }
    }
    for (int idx = threadIdx.x; idx < SYS_N * ldx; idx += blockDim.x) {
      const int q = idx / ldx, s = idx - q * ldx, k = k_of_pos(s, L), n = nsys(q);
      double *fr = rows + (size_t)(q * FR_N) * ldx;
      const bool in = k < n;
      fr[FR_SCALE * ldx + s] = in ? W(q, W_SC)[k] : 0.0;
      fr[FR_L * ldx + s] = in ? W(q, W_L)[k] : 0.0;
      fr[FR_RINV * ldx + s] = in ? W(q, W_U)[k] : 0.0;
      fr[FR_E * ldx + s] = in ? W(q, W_E)[k] : 0.0;
      fr[FR_DL * ldx + s] = in && k >= 1 ? W(q, W_DL)[k] : 0.0;
      fr[FR_DU * ldx + s] = in && k + 1 < n ? W(q, W_DU)[k] : 0.0;
      if (!in) {
        fr[FR_G * ldx + s] = 0.0;
        fr[FR_H * ldx + s] = 0.0;
      }
    }
    __syncthreads();
  }
}

hipError_t launch_step_begin(DevStep *ds, const double *cfl_out, StepConsts k, hipStream_t st) {
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, st, ds, cfl_out, k);
  return hipGetLastError();
}
hipError_t launch_factor(const FactorArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_factor, dim3(1), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_step_end(DevStep *ds, hipStream_t st) {
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, st, ds);
  return hipGetLastError();
}

}  // namespace afid
