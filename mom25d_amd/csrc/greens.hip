// Sommerfeld integrals of the layered-medium Green's-function remainders.
//
// For every table row (one rho of one frequency's table) the kernel forms
//     out[row, c] = sum_k J0(rho * kc_k) * Cc[k, c]   (lifted complex path)
//                 + sum_k J0(rho * kr_k) * Cr[k, c]   (real tail)
// where C[k, c] = w_k kr_k g~_c(kr_k) / (2 pi) is prepared on the host.  This
// is the (n_rho x n_kr) Bessel-matrix product of Engine/mom25d/greens.py:157-173
// (slab) and Engine/mom25d/layered.py:360-377 (stackups), batched over all
// frequencies of a sweep in one launch: one workgroup per table row.
//
// J0 of a complex argument is this file's own implementation:
//   |z| < 4   power series  sum (-z^2/4)^k / (k!)^2  (17 terms),
//   |z| < 20  Miller's backward recurrence J_{n-1} = (2n/z) J_n - J_{n+1}
//             from n = N >= |z| + 36, normalised with 1 = J0 + 2 sum J_2k,
//   else      Hankel's asymptotic expansion (DLMF 10.17.3) truncated at the
//             first term below 1e-17 (the smallest term is ~exp(-2|z|)).
// cos/sin(z - pi/4) are formed as (cos z +- sin z)/sqrt(2) so that no
// rounding of z - pi/4 enters at large |z|.

#include "common.h"

namespace mom {

__device__ inline z64 j0_series(z64 z) {
  z64 w = {-0.25 * (z.re * z.re - z.im * z.im), -0.5 * z.re * z.im};
  z64 term = {1.0, 0.0}, sum = {1.0, 0.0};
#pragma unroll
  for (int k = 1; k <= 17; ++k) {
    double inv = 1.0 / double(k * k);
    term = (term * w) * inv;
    sum += term;
  }
  return sum;
}

__device__ inline z64 j0_miller(z64 z, double r) {
  int N = 2 * (int)((r + 36.0) * 0.5);
  z64 inv2z = cdiv(z64{2.0, 0.0}, z);
  z64 jp1 = {0.0, 0.0};
  z64 j = {1e-30, 0.0};
  z64 norm = {2e-30, 0.0};  // N is even: J_N enters the normalisation
  for (int n = N; n >= 1; --n) {
    z64 jm1 = (double(n) * inv2z) * j - jp1;
    jp1 = j;
    j = jm1;
    if (((n - 1) & 1) == 0 && n - 1 > 0) norm += 2.0 * jm1;
  }
  norm += j;
  return cdiv(j, norm);
}

template <bool REAL>
__device__ inline z64 j0_asym(z64 z) {
  z64 iz = REAL ? z64{1.0 / z.re, 0.0} : cdiv(z64{1.0, 0.0}, z);
  z64 t = {1.0, 0.0};
  z64 P = {1.0, 0.0}, Q = {0.0, 0.0};
  double prev = 1.0;
  for (int k = 1; k <= 60; ++k) {
    double f = -double((2 * k - 1) * (2 * k - 1)) / (8.0 * k);
    t = (f * t) * iz;
    double mag = fabs(t.re) + fabs(t.im);
    if (mag > prev) break;  // asymptotic series starts to diverge
    prev = mag;
    // P = t0 - t2 + t4 - ..., Q = t1 - t3 + t5 - ...
    int m = k & 3;
    if (m == 1) Q += t;
    else if (m == 2) P -= t;
    else if (m == 3) Q -= t;
    else P += t;
    if (mag < 1e-17) break;
  }
  double sx, cx;
  sincos(z.re, &sx, &cx);
  const double rs2 = 0.70710678118654752440;
  double c4 = (cx + sx) * rs2;  // cos(x - pi/4)
  double s4 = (sx - cx) * rs2;  // sin(x - pi/4)
  z64 cosx, sinx;
  if (REAL) {
    cosx = {c4, 0.0};
    sinx = {s4, 0.0};
  } else {
    double e = exp(z.im), ei = 1.0 / e;
    double ch = 0.5 * (e + ei), sh = 0.5 * (e - ei);
    cosx = {c4 * ch, -s4 * sh};
    sinx = {s4 * ch, c4 * sh};
  }
  z64 u = (2.0 / M_PI) * iz;  // amplitude sqrt(2 / (pi z)), principal branch
  double m = sqrt(u.re * u.re + u.im * u.im);
  double sr = sqrt(0.5 * (m + u.re));
  z64 amp = {sr, (sr > 0.0) ? 0.5 * u.im / sr : sqrt(0.5 * (m - u.re))};
  return amp * (P * cosx - Q * sinx);
}

__device__ z64 bessel_j0(z64 z) {
  double r = sqrt(z.re * z.re + z.im * z.im);
  if (r < 4.0) return j0_series(z);
  if (r < 20.0) return j0_miller(z, r);
  return j0_asym<false>(z);
}

__device__ double bessel_j0_real(double x) {
  x = fabs(x);
  if (x < 4.0) return j0_series(z64{x, 0.0}).re;
  if (x < 20.0) return j0_miller(z64{x, 0.0}, x).re;
  return j0_asym<true>(z64{x, 0.0}).re;
}

__device__ inline double shfl_down_d(double v, int off) {
  return __shfl_down(v, off, 64);
}

template <int NC>
__global__ __launch_bounds__(256) void sommerfeld_kernel(const double* __restrict__ rho,
                                                         const int32_t* __restrict__ row_prob,
                                                         const int64_t* __restrict__ kc_off,
                                                         const int32_t* __restrict__ kc_len,
                                                         const z64* __restrict__ kc,
                                                         const z64* __restrict__ Cc,
                                                         const int64_t* __restrict__ kr_off,
                                                         const int32_t* __restrict__ kr_len,
                                                         const double* __restrict__ kr,
                                                         const z64* __restrict__ Cr,
                                                         z64* __restrict__ out) {
  const int row = blockIdx.x;
  const int tid = threadIdx.x;
  const double r = rho[row];
  const int p = row_prob[row];
  z64 acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = {0.0, 0.0};

  const int64_t c0 = kc_off[p];
  const int nc = kc_len[p];
  for (int i = tid; i < nc; i += 256) {
    z64 k = kc[c0 + i];
    z64 J = bessel_j0(z64{r * k.re, r * k.im});
    const z64* C = Cc + (c0 + i) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += J * C[c];
  }
  const int64_t r0 = kr_off[p];
  const int nr = kr_len[p];
  for (int i = tid; i < nr; i += 256) {
    double J = bessel_j0_real(r * kr[r0 + i]);
    const z64* C = Cr + (r0 + i) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += J * C[c];
  }

  __shared__ z64 part[4][NC];
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double re = acc[c].re, im = acc[c].im;
    for (int off = 32; off > 0; off >>= 1) {
      re += shfl_down_d(re, off);
      im += shfl_down_d(im, off);
    }
    if (lane == 0) part[wave][c] = {re, im};
  }
  __syncthreads();
  if (tid < NC) {
    z64 s = part[0][tid];
    s += part[1][tid];
    s += part[2][tid];
    s += part[3][tid];
    out[(int64_t)row * NC + tid] = s;
  }
}

// Element-wise J0 (complex and real), for checking the Bessel routines.
__global__ void j0_eval_kernel(int n, const z64* __restrict__ z, z64* __restrict__ out, int real_axis) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = real_axis ? z64{bessel_j0_real(z[i].re), 0.0} : bessel_j0(z[i]);
}

}  // namespace mom

using namespace mom;

extern "C" int mom_sommerfeld(int n_rows, const double* rho, const int32_t* row_prob, const int64_t* kc_off,
                              const int32_t* kc_len, const double* kc, const double* Cc, const int64_t* kr_off,
                              const int32_t* kr_len, const double* kr, const double* Cr, int ncol, double* out,
                              void* stream) {
  if (n_rows < 0) return fail("mom_sommerfeld: negative row count");
  if (n_rows == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  dim3 grid(n_rows), block(256);
#define MOM_SOMMERFELD_CASE(NC)                                                                            \
  case NC:                                                                                                 \
    hipLaunchKernelGGL(sommerfeld_kernel<NC>, grid, block, 0, s, rho, row_prob, kc_off, kc_len, (const z64*)kc, \
                       (const z64*)Cc, kr_off, kr_len, kr, (const z64*)Cr, (z64*)out);                      \
    break;
  switch (ncol) {
    MOM_SOMMERFELD_CASE(1)
    MOM_SOMMERFELD_CASE(2)
    MOM_SOMMERFELD_CASE(3)
    MOM_SOMMERFELD_CASE(4)
    MOM_SOMMERFELD_CASE(8)
    MOM_SOMMERFELD_CASE(9)
    MOM_SOMMERFELD_CASE(12)
    MOM_SOMMERFELD_CASE(18)
    MOM_SOMMERFELD_CASE(27)
    default:
      return fail("mom_sommerfeld: ncol must be k * levels^2 with k = 1, 2 or 3 and 1 to 3 levels");
  }
#undef MOM_SOMMERFELD_CASE
  return check_launch("sommerfeld_kernel");
}

extern "C" int mom_j0(int n, const double* z, double* out, int real_axis, void* stream) {
  if (n <= 0) return n == 0 ? 0 : fail("mom_j0: negative length");
  hipLaunchKernelGGL(j0_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, (const z64*)z,
                     (z64*)out, real_axis);
  return check_launch("j0_eval_kernel");
}
