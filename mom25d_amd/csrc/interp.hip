// Frequency interpolation of the impedance matrix (Engine/mom25d/freqinterp.py:1-209).
//
// Z is filled exactly at anchor frequencies; between them the smooth
// W = j w Z exp(+j k_e R_mn) is Lagrange-interpolated from the nearest
// anchors and Z = W exp(-j k_e R_mn) / (j w).  R_mn is the 3D distance of the
// basis-function centres (RWG edge midpoints on their level, vias at mid
// span) and is 0 on the via-via block, which gets no phase extraction
// (freqinterp.py:70-83).  Both steps stream whole matrices through HBM once:
// one thread per entry along a row, the centre of the row's basis function
// uniform per workgroup.
#include "common.h"

namespace mom {

// Distance of the centres of basis functions m and n (0 on the via-via block).
__device__ inline double centre_distance(const double* pos, int nb, int m, int n) {
  if (m >= nb && n >= nb) return 0.0;
  const double dx = pos[3 * m] - pos[3 * n];
  const double dy = pos[3 * m + 1] - pos[3 * n + 1];
  const double dz = pos[3 * m + 2] - pos[3 * n + 2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// W[m, n] = j w Z[m, n] exp(+j ke R_mn), stored in precision T.
template <typename T>
__global__ __launch_bounds__(256) void interp_anchor_kernel(int N, int nb, const z64* __restrict__ Z, int64_t ldz,
                                                            const double* __restrict__ pos, double omega, double ke,
                                                            cplx<T>* __restrict__ W, int64_t ldw) {
  const int m = blockIdx.y, n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  double s, c;
  sincos(ke * centre_distance(pos, nb, m, n), &s, &c);
  const z64 z = Z[(int64_t)m * ldz + n];
  const z64 w = {-omega * z.im, omega * z.re};  // j w z
  W[(int64_t)m * ldw + n] = {(T)(w.re * c - w.im * s), (T)(w.re * s + w.im * c)};
}

// Up to 8 anchors and their Lagrange weights, passed by value.
template <typename T>
struct InterpAnchors {
  const cplx<T>* W[8];
  double L[8];
  int K;
};

// Z[m, n] = (sum_k L_k W_k[m, n]) exp(-j ke R_mn) / (j w).
template <typename T>
__global__ __launch_bounds__(256) void interp_combine_kernel(int N, int nb, InterpAnchors<T> a, int64_t ldw,
                                                             const double* __restrict__ pos, double omega, double ke,
                                                             z64* __restrict__ Z, int64_t ldz) {
  const int m = blockIdx.y, n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  const int64_t q = (int64_t)m * ldw + n;
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (k < a.K) {
      const cplx<T> w = a.W[k][q];
      re += a.L[k] * (double)w.re;
      im += a.L[k] * (double)w.im;
    }
  double s, c;
  sincos(ke * centre_distance(pos, nb, m, n), &s, &c);
  const double pr = re * c + im * s, pi = im * c - re * s;  // times exp(-j ke R)
  Z[(int64_t)m * ldz + n] = {pi / omega, -pr / omega};    // divided by j w
}

}  // namespace mom

using namespace mom;

extern "C" int mom_interp_anchor(int prec_w, int N, int nb, const void* Z, int64_t ldz, const double* pos,
                                 double omega, double ke, void* W, int64_t ldw, void* stream) {
  if (N <= 0) return 0;
  if (!Z || !pos || !W) return fail("mom_interp_anchor: null device pointer");
  if (omega <= 0) return fail("mom_interp_anchor: omega must be positive");
  const dim3 grid((unsigned)((N + 255) / 256), (unsigned)N);
  hipStream_t s = (hipStream_t)stream;
  if (prec_w == 64)
    hipLaunchKernelGGL(interp_anchor_kernel<double>, grid, dim3(256), 0, s, N, nb, (const z64*)Z, ldz, pos, omega, ke,
                       (z64*)W, ldw);
  else if (prec_w == 32)
    hipLaunchKernelGGL(interp_anchor_kernel<float>, grid, dim3(256), 0, s, N, nb, (const z64*)Z, ldz, pos, omega, ke,
                       (c32*)W, ldw);
  else
    return fail("mom_interp_anchor: prec must be 64 or 32");
  return check_launch("interp_anchor_kernel");
}

extern "C" int mom_interp_combine(int prec_w, int N, int nb, int K, const void* const* W, const double* L,
                                  int64_t ldw, const double* pos, double omega, double ke, void* Z, int64_t ldz,
                                  void* stream) {
  if (N <= 0) return 0;
  if (K < 1 || K > 8) return fail("mom_interp_combine: 1 to 8 anchors");
  if (!W || !L || !pos || !Z) return fail("mom_interp_combine: null pointer");
  if (omega <= 0) return fail("mom_interp_combine: omega must be positive");
  const dim3 grid((unsigned)((N + 255) / 256), (unsigned)N);
  hipStream_t s = (hipStream_t)stream;
  if (prec_w == 64) {
    InterpAnchors<double> a{};
    for (int k = 0; k < K; ++k) {
      if (!W[k]) return fail("mom_interp_combine: null anchor");
      a.W[k] = (const z64*)W[k];
      a.L[k] = L[k];
    }
    a.K = K;
    hipLaunchKernelGGL(interp_combine_kernel<double>, grid, dim3(256), 0, s, N, nb, a, ldw, pos, omega, ke, (z64*)Z,
                       ldz);
  } else if (prec_w == 32) {
    InterpAnchors<float> a{};
    for (int k = 0; k < K; ++k) {
      if (!W[k]) return fail("mom_interp_combine: null anchor");
      a.W[k] = (const c32*)W[k];
      a.L[k] = L[k];
    }
    a.K = K;
    hipLaunchKernelGGL(interp_combine_kernel<float>, grid, dim3(256), 0, s, N, nb, a, ldw, pos, omega, ke, (z64*)Z,
                       ldz);
  } else {
    return fail("mom_interp_combine: prec must be 64 or 32");
  }
  return check_launch("interp_combine_kernel");
}
