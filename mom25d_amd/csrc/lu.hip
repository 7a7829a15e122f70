// Dense complex LU with partial pivoting, triangular solves and FP64
// iterative refinement, for complex128 (R = double) and complex64 (R = float)
// matrices stored row-major with interleaved (re, im).
//
// Replaces scipy.linalg.lu_factor / lu_solve (LAPACK zgetrf / zgetrs) of
// Engine/mom25d/solver.py:11-23.  Blocked right-looking factorisation with
// panel width jb:
//   1. panel  : GEPP of the (n - k0) x jb column panel by a grid whose
//               workgroups are all resident (rows split over workgroups, held
//               in LDS; one exchange per column: every workgroup publishes its
//               local pivot candidate row in its own slot, and once all slots
//               carry the column's generation every workgroup knows the winner);
//               pivots are chosen by |re| + |im|, first maximum wins, as
//               LAPACK's izamax does;
//   2. laswp  : the panel's row interchanges applied to the other columns;
//   3. trsm   : U12 = L11^-1 A12 (unit lower triangular);
//   4. update : A22 -= L21 U12, a complex GEMM on the matrix cores, written as
//               one real GEMM: with A' the interleaved (re, im) view of L21
//               (m x 2jb) and B'' the 2x2 real expansion of U12 (2jb x 2n),
//               [Re C, Im C] = A' B''.  MFMA 16x16x4 (f64 or f32) tiles.

#include "common.h"

#include <algorithm>
#include <map>
#include <type_traits>
#include <utility>
#include <mutex>
#include <vector>

namespace mom {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename R>
__device__ inline R cabs1(cplx<R> a) { return fabs(a.re) + fabs(a.im); }

template <typename T>
__device__ inline T load_coherent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename R>
__device__ inline cplx<R> load_coherent_c(const cplx<R>* p) {
  const R* q = reinterpret_cast<const R*>(p);
  return {load_coherent(q), load_coherent(q + 1)};
}
// complex64: one 8-byte device-coherent load
template <>
__device__ inline cplx<float> load_coherent_c(const cplx<float>* p) {
  const unsigned long long v = load_coherent(reinterpret_cast<const unsigned long long*>(p));
  return {__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32))};
}

// Cross-workgroup publication without cache maintenance: data is written
// and read with device-coherent (relaxed, agent-scope) accesses, and every
// published value is its own "ready" flag (a sentinel before it is written;
// see the panel below), so no ordering between stores is needed.  Agent-scope
// release / acquire would instead write back / invalidate the whole L2 of the
// XCD on every flag, which costs microseconds when other kernels keep the L2
// busy.
template <typename T>
__device__ inline void store_coherent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename R>
__device__ inline void store_coherent_c(cplx<R>* p, cplx<R> v) {
  R* q = reinterpret_cast<R*>(p);
  store_coherent(q, v.re);
  store_coherent(q + 1, v.im);
}

// Back-off between polls of a flag in memory (units of 64 clocks).
#ifndef MOM_POLL_SLEEP
#define MOM_POLL_SLEEP 1
#endif
__device__ inline void poll_pause() { __builtin_amdgcn_s_sleep(MOM_POLL_SLEEP); }

// Upper bound on the polls of one wait (each at least a memory round trip,
// so 2^21 polls take about a second or more): a wait that runs out ends the
// kernel with an error flag instead of holding the GPU, and the workgroups
// waiting on it run out in turn.  mom_set_spin_limit changes it (tests force
// the error path with a tiny limit).
static unsigned g_spin_limit = 1u << 21;

// ------------------------------------------------------------------- panel
// Values that are themselves the "ready" flags: the publication buffers of a
// panel (and the solutions of the triangular solves below) are filled with a
// sentinel NaN (a payload no arithmetic produces) before the launch, every
// column of a panel has buffers of its own, and a reader polls each value
// until it is no longer the sentinel.  A writer needs no completion wait and
// no separate flag, so a column costs one memory round trip less.
template <typename R>
struct Sentinel;
template <>
struct Sentinel<double> {
  static constexpr unsigned long long bits = 0x7FF4D0D0D0D0D0D0ull;
  static __device__ inline bool is(double v) { return (unsigned long long)__double_as_longlong(v) == bits; }
  static __device__ inline double value() { return __longlong_as_double((long long)bits); }
};
template <>
struct Sentinel<float> {
  static constexpr unsigned bits = 0x7FA0D0D0u;
  static __device__ inline bool is(float v) { return __float_as_uint(v) == bits; }
  static __device__ inline float value() { return __uint_as_float(bits); }
};
constexpr int32_t kIdxSentinel = -2;

template <typename R>
__device__ inline bool is_sentinel(cplx<R> v) {
  return Sentinel<R>::is(v.re) || Sentinel<R>::is(v.im);
}

// Per-(column, workgroup) publication slot: the candidate pivot's |re|+|im|
// and row (each checked against its own sentinel: a half-written slot reads
// as not ready).
struct alignas(16) PivotSlot {
  double val;
  int32_t idx;
  int32_t pad;
};

template <typename R>
struct PanelArgs {
  cplx<R>* A;
  int64_t lda;
  int n, k0, jb, rows_per;  // rows_per: rows owned by one workgroup
  uint32_t spin_limit;      // polls before a wait gives up
  int32_t* ipiv;
  int32_t* info;
  PivotSlot* slots;         // [jb][G]
  cplx<R>* slot_row;        // [jb][G][jb]: each workgroup's candidate row
  cplx<R>* diag_row;        // [jb][jb]: row c before the interchange
};

// The sentinel in a panel's publication buffers: nslot slots and nrow
// complex entries (candidate rows followed by the diagonal rows).
template <typename R>
__global__ void panel_reset_kernel(PivotSlot* slots, int64_t nslot, cplx<R>* rowbuf, int64_t nrow) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrow; i += stride)
    rowbuf[i] = {Sentinel<R>::value(), Sentinel<R>::value()};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslot; i += stride)
    slots[i] = {Sentinel<double>::value(), kIdxSentinel, 0};
}

// (value, index) maximum over the workgroup, ties to the smaller index.
template <typename R>
__device__ inline void block_argmax(R& v, int& i, R* sv, int* si) {
  for (int off = 32; off > 0; off >>= 1) {
    R v2 = __shfl_xor(v, off, 64);
    int i2 = __shfl_xor(i, off, 64);
    if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { sv[wave] = v; si[wave] = i; }
  __syncthreads();
  v = sv[0];
  i = si[0];
#pragma unroll
  for (int w = 1; w < 4; ++w)
    if (sv[w] > v || (sv[w] == v && si[w] < i)) { v = sv[w]; i = si[w]; }
  __syncthreads();
}

template <typename R>
__global__ __launch_bounds__(256) void panel_kernel(PanelArgs<R> a) {
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* rows = reinterpret_cast<cplx<R>*>(smem_raw);  // [rows_per][jb]
  __shared__ cplx<R> prow[128];
  __shared__ cplx<R> drow[128];
  __shared__ R red_v[4];
  __shared__ int red_i[4];
  __shared__ int s_win;
  __shared__ int s_abort;

  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  if (tid == 0) s_abort = 0;
  const int jb = a.jb, rp = a.rows_per;
  const int r0 = a.k0 + g * rp;                 // first owned row
  const int r1 = min(a.n, r0 + rp);             // one past the last owned row
  const int nown = max(0, r1 - r0);

  for (int q = tid; q < nown * jb; q += 256) {
    int lr = q / jb, c = q % jb;
    rows[q] = a.A[(int64_t)(r0 + lr) * a.lda + a.k0 + c];
  }
  __syncthreads();

  for (int j = 0; j < jb; ++j) {
    const int c = a.k0 + j;
    cplx<R>* const crow = a.slot_row + (int64_t)j * G * jb;  // this column's candidate rows
    cplx<R>* const drow_pub = a.diag_row + (int64_t)j * jb;
    // phase A: local candidate among owned rows >= c, published with its row
    R best = (R)-1;
    int bidx = 0x7fffffff;
    for (int lr = max(0, c - r0) + tid; lr < nown; lr += 256) {
      R v = cabs1(rows[lr * jb + j]);
      if (v > best) { best = v; bidx = r0 + lr; }
    }
    block_argmax(best, bidx, red_v, red_i);
    if (bidx != 0x7fffffff)
      for (int q = tid; q < jb; q += 256) store_coherent_c(crow + (int64_t)g * jb + q, rows[(bidx - r0) * jb + q]);
    if (c >= r0 && c < r1)
      for (int q = tid; q < jb; q += 256) store_coherent_c(drow_pub + q, rows[(c - r0) * jb + q]);
    if (tid == 0) {
      PivotSlot* sl = a.slots + (int64_t)j * G + g;
      store_coherent(&sl->val, (double)best);
      store_coherent(&sl->idx, bidx);
    }
    // phase B: wait for every workgroup's slot, global pivot
    R v = (R)-1;
    int i = 0x7fffffff;
    if (tid < G) {
      PivotSlot* sl = a.slots + (int64_t)j * G + tid;
      unsigned spins = 0;
      double sv;
      int si;
      for (;;) {
        sv = load_coherent(&sl->val);
        si = load_coherent(&sl->idx);
        if (!Sentinel<double>::is(sv) && si != kIdxSentinel) break;
        poll_pause();
        if (++spins > a.spin_limit) {
          s_abort = 1;
          break;
        }
      }
      v = (R)sv;
      i = si;
    }
    const int mine_i = i;
    block_argmax(v, i, red_v, red_i);
    const int piv = i;
    if (tid < G && mine_i == piv) s_win = tid;
    __syncthreads();
    if (s_abort) {  // a workgroup never published: give up (reported in info[1])
      // the remaining pivots of the panel become identities, so that the row
      // interchanges that follow stay within the matrix
      if (g == 0)
        for (int q = tid; q < jb - j; q += 256) a.ipiv[c + q] = c + q;
      if (tid == 0) a.info[1] = 1;
      return;
    }
    const int win = s_win;
    // the winner's row and row c, polled element by element (published with
    // the slots, so normally already there)
    for (int q = tid; q < 2 * jb; q += 256) {
      const cplx<R>* src = q < jb ? crow + (int64_t)win * jb + q : drow_pub + (q - jb);
      unsigned spins = 0;
      cplx<R> x;
      for (;;) {
        x = load_coherent_c(src);
        if (!is_sentinel(x)) break;
        poll_pause();
        if (++spins > a.spin_limit) {
          s_abort = 1;
          break;
        }
      }
      if (q < jb)
        prow[q] = x;
      else
        drow[q - jb] = x;
    }
    __syncthreads();
    if (s_abort) {  // as above: identity pivots for the rest of the panel
      if (g == 0)
        for (int q = tid; q < jb - j; q += 256) a.ipiv[c + q] = c + q;
      if (tid == 0) a.info[1] = 1;
      return;
    }
    const bool nonzero = (prow[j].re != (R)0) || (prow[j].im != (R)0);
    const cplx<R> inv = nonzero ? cdiv(cplx<R>{(R)1, (R)0}, prow[j]) : cplx<R>{(R)0, (R)0};
    if (tid == 0) {
      if (g == 0) a.ipiv[c] = piv;
      if (g == 0 && !nonzero && *a.info == 0) *a.info = c + 1;
    }
    if (nonzero) {
      if (c >= r0 && c < r1)
        for (int q = tid; q < jb; q += 256) rows[(c - r0) * jb + q] = prow[q];
      if (piv != c && piv >= r0 && piv < r1)
        for (int q = tid; q < jb; q += 256) rows[(piv - r0) * jb + q] = drow[q];
    }
    __syncthreads();
    // scale the pivot column and apply the rank-1 update, row by row
    const int lstart = max(0, c + 1 - r0);
    const int w = jb - j - 1;
    if (nonzero)
      for (int lr = lstart + tid; lr < nown; lr += 256) rows[lr * jb + j] = rows[lr * jb + j] * inv;
    __syncthreads();
    if (w > 0)
      for (int q = tid; q < (nown - lstart) * w; q += 256) {
        int lr = lstart + q / w, cc = j + 1 + q % w;
        rows[lr * jb + cc] -= rows[lr * jb + j] * prow[cc];
      }
    __syncthreads();
  }
  for (int q = tid; q < nown * jb; q += 256) {
    int lr = q / jb, c = q % jb;
    a.A[(int64_t)(r0 + lr) * a.lda + a.k0 + c] = rows[q];
  }
}

// The kernels of a GMRES step (preconditioner solves, products,
// orthogonalisation) form a latency-bound chain per lane, while other lanes'
// factorisations and fills run beside them on the same CUs: their waves
// issue first (wave priority 2 of 3; the factorisation's panels take 3).
__device__ __forceinline__ void gmres_step_priority() { __builtin_amdgcn_s_setprio(2); }

// ------------------------------------------------- unpivoted panel
template <typename R>
static constexpr int kNoPivPanelMax() {  // widest unpivoted panel (columns held by the register layout of panel_below_kernel)
  return sizeof(R) == sizeof(double) ? 64 : 128;
}

// The panel of the LU without row interchanges (mom_getrf_nopiv: the
// preconditioner factorisation of the mixed path, where GMRES absorbs the
// difference): no column needs a cross-workgroup pivot search, so a panel is
// two ordinary launches -- the b x b diagonal block factored by one workgroup,
// then the rows below it, L21 = A21 U11^-1, by independent workgroups of 16
// rows -- instead of one exchange per column.  The rows below keep their part
// of the matrix in registers (thread (c, q): column c of the rows i = q mod Q)
// and exchange only the column being finished through LDS (1 KB), so that
// these 256-thread workgroups find room on CUs that the concurrent
// trailing-update GEMM of the look-ahead still occupies.  (A register form of
// the diagonal block, two barriers per column, measured slower than the LDS
// form below: 280 against 250 us per 128-column block.)  An exactly zero
// pivot is replaced by 1 and reported in
// info[0] (first one, 1-based), which leaves finite factors that the caller
// discards.
template <typename R>
__device__ inline cplx<R> nonzero_pivot(cplx<R> p) {
  return (p.re == (R)0 && p.im == (R)0) ? cplx<R>{(R)1, (R)0} : p;
}

// The diagonal block: one workgroup of 1024 threads with the block in LDS
// (b x b, up to 128 KB, which also keeps the look-ahead's update GEMM off
// its CU); thread (c, q) updates column c of the rows i = q mod 8, 8 rows at
// a time (their loads issued together), one barrier per column.  Column j is
// scaled by 1 / u_jj in step j + 1, when no one reads it any more.
template <typename R>
__global__ __launch_bounds__(1024) void diag_lu_kernel(cplx<R>* A, int64_t lda, int k0, int b, int32_t* ipiv,
                                                       int32_t* info, int sym) {
  // the block in LDS; thread (c, q) updates column c of the rows i = q mod 8,
  // 8 rows at a time (their loads issued together).  sym: the block of a
  // complex-symmetric matrix whose updates reached only its lower triangle;
  // the upper one is taken from it.
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* T = reinterpret_cast<cplx<R>*>(smem_raw);  // [b][b]
  const int tid = threadIdx.x, c = tid % 128;
  const int q = __builtin_amdgcn_readfirstlane(tid / 128);  // uniform per wavefront
  for (int e = tid; e < b * b; e += 1024) {
    int i = e / b, j = e % b;
    if (sym && j > i) {
      const int t = i;
      i = j;
      j = t;
    }
    T[e] = A[(int64_t)(k0 + i) * lda + k0 + j];
  }
  for (int e = tid; e < b; e += 1024) ipiv[k0 + e] = k0 + e;
  __syncthreads();
  for (int j = 0; j < b; ++j) {
    const cplx<R> p = T[j * b + j];  // final since step j - 1
    if (j > 0) {  // finish column j - 1: its rows below the diagonal scaled, one row per thread
      const cplx<R> pp = T[(j - 1) * b + j - 1];
      if (j + tid < b) T[(j + tid) * b + j - 1] = T[(j + tid) * b + j - 1] * cdiv(cplx<R>{(R)1, (R)0},
                                                                                     nonzero_pivot(pp));
      if (tid == 0) T[(j - 1) * b + j - 1] = nonzero_pivot(pp);
    }
    if (tid == 0 && p.re == (R)0 && p.im == (R)0 && *info == 0) *info = k0 + j + 1;
    if (c > j && c < b) {
      const cplx<R> m = cdiv(T[j * b + c], nonzero_pivot(p));
      const int first = j + 1 + ((q - (j + 1)) & 7);  // first row > j of this class
      for (int i0 = first; i0 < b; i0 += 64) {
        cplx<R> a[8], t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int i = min(i0 + 8 * k, b - 1);
          a[k] = T[i * b + j];
          t[k] = T[i * b + c];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (i0 + 8 * k < b) T[(i0 + 8 * k) * b + c] = t[k] - a[k] * m;
      }
    }
    __syncthreads();
  }
  if (tid == 0) T[(b - 1) * b + b - 1] = nonzero_pivot(T[(b - 1) * b + b - 1]);
  __syncthreads();
  for (int e = tid; e < b * b; e += 1024) A[(int64_t)(k0 + e / b) * lda + k0 + e % b] = T[e];
}

// L21 = A21 U11^-1 for the rows [k0 + b, n) below a factored diagonal block:
// 256 threads per workgroup of 16 rows, thread (c, q) = (tid % NC, tid / NC)
// holds column c of the rows q, q + Q, ... in registers.  Row j of U11 is
// read from memory (L2: every workgroup reads the same block) two steps
// ahead of its use; the column being finished is exchanged through LDS.
template <typename R>
__global__ __launch_bounds__(256) void panel_below_kernel(cplx<R>* A, int64_t lda, int n, int k0, int b) {
  constexpr int NC = kNoPivPanelMax<R>(), Q = 256 / NC, RT = 16, HR = RT / Q;
  __shared__ cplx<R> colbuf[2][RT];
  const int tid = threadIdx.x, c = tid % NC;
  const int q = __builtin_amdgcn_readfirstlane(tid / NC);
  const bool mine = c < b;
  const int r0 = k0 + b + blockIdx.x * RT, nr = min(RT, n - r0);
  const cplx<R>* U = A + (int64_t)k0 * lda + k0;
  constexpr cplx<R> zero = {(R)0, (R)0};
  cplx<R> v[HR];
#pragma unroll
  for (int t = 0; t < HR; ++t) {
    const int i = q + Q * t;
    v[t] = zero;
    if (mine && i < nr) v[t] = A[(int64_t)(r0 + i) * lda + k0 + c];
  }
  // u_jc and u_jj of rows j and j + 1 (rows past b, and columns past b, read as 0)
  cplx<R> uc0 = zero, ud0 = U[0], uc1 = zero, ud1 = zero;
  if (mine) uc0 = U[c];
  if (b > 1) {
    ud1 = U[lda + 1];
    if (mine) uc1 = U[lda + c];
  }
  for (int j = 0; j < b; ++j) {
    const int buf = j & 1;
    cplx<R> uc2 = zero, ud2 = zero;
    if (j + 2 < b) {
      ud2 = U[(int64_t)(j + 2) * lda + j + 2];
      if (mine) uc2 = U[(int64_t)(j + 2) * lda + c];
    }
    if (c == j) {
#pragma unroll
      for (int t = 0; t < HR; ++t) colbuf[buf][q + Q * t] = v[t];
    }
    __syncthreads();
    const cplx<R> inv = cdiv(cplx<R>{(R)1, (R)0}, nonzero_pivot(ud0));
    if (mine && c > j) {
      const cplx<R> m = uc0 * inv;
#pragma unroll
      for (int t = 0; t < HR; ++t) v[t] -= colbuf[buf][q + Q * t] * m;
    } else if (c == j) {
#pragma unroll
      for (int t = 0; t < HR; ++t) v[t] = v[t] * inv;
    }
    uc0 = uc1;
    ud0 = ud1;
    uc1 = uc2;
    ud1 = ud2;
  }
#pragma unroll
  for (int t = 0; t < HR; ++t) {
    const int i = q + Q * t;
    if (mine && i < nr) A[(int64_t)(r0 + i) * lda + k0 + c] = v[t];
  }
}

// ----------------------------------------------------------------- laswp
// Row interchanges ipiv[k0 .. k0+b) applied, in order, to columns [c0, c1)
// of A.  The first wavefront folds the b interchanges into one permutation
// of the (at most 2b) rows they touch; every workgroup then moves its 16
// columns of those rows through LDS: all loads, then all stores.
template <typename R>
__global__ __launch_bounds__(256) void laswp_kernel(cplx<R>* A, int64_t lda, int n, int c0, int c1, int k0, int b,
                                                    const int32_t* ipiv) {
  __shared__ int rows[256];
  __shared__ int src[256];
  __shared__ int s_nrows;
  __shared__ cplx<R> buf[256][17];
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid < 64) {
    for (int i = lane; i < 2 * b; i += 64) {
      rows[i] = i < b ? k0 + i : -1;
      src[i] = i;
    }
    int nout = 0;
    for (int k = 0; k < b; ++k) {
      int p = ipiv[k0 + k];
      if (p < k0 + k || p >= n) p = k0 + k;  // a valid pivot row lies in [k0 + k, n)
      int pp;
      if (p < k0 + b) {
        pp = p - k0;
      } else {
        int found = -1;
        for (int base = 0; base < nout; base += 64) {
          const int j = base + lane;
          const bool hit = j < nout && rows[b + j] == p;
          const unsigned long long m = __ballot(hit);
          if (m) { found = base + __ffsll((long long)m) - 1; break; }
        }
        if (found < 0) {
          found = nout;
          if (lane == 0) rows[b + nout] = p;
          ++nout;
        }
        pp = b + found;
      }
      if (lane == 0 && pp != k) {
        int t = src[k];
        src[k] = src[pp];
        src[pp] = t;
      }
    }
    if (lane == 0) s_nrows = b + nout;
  }
  __syncthreads();
  const int nrows = s_nrows;
  const int tx = tid & 15, ty = tid >> 4;
  const int col = c0 + blockIdx.x * 16 + tx;
  if (col < c1)
    for (int i = ty; i < nrows; i += 16) buf[i][tx] = A[(int64_t)rows[src[i]] * lda + col];
  __syncthreads();
  if (col < c1)
    for (int i = ty; i < nrows; i += 16)
      if (src[i] != i) A[(int64_t)rows[i] * lda + col] = buf[i][tx];
}

// ------------------------------------------------------------------ trsm
// B (jb x ncols at A[k0, c0..]) <- L11^-1 B, L11 unit lower in A[k0.., k0..].
// One workgroup per 16 columns; B staged in LDS.
template <typename R>
__global__ __launch_bounds__(256) void trsm_lower_unit_kernel(cplx<R>* A, int64_t lda, int k0, int jb, int c0,
                                                              int c1) {
  __shared__ cplx<R> Bs[128][17];
  const int cb = c0 + blockIdx.x * 16;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int col = cb + tx;
  for (int r = ty; r < jb; r += 16) Bs[r][tx] = col < c1 ? A[(int64_t)(k0 + r) * lda + col] : cplx<R>{0, 0};
  __syncthreads();
  for (int i = 0; i < jb - 1; ++i) {
    const cplx<R> xi = Bs[i][tx];
    for (int r = i + 1 + ty; r < jb; r += 16) Bs[r][tx] -= A[(int64_t)(k0 + r) * lda + k0 + i] * xi;
    __syncthreads();
  }
  for (int r = ty; r < jb; r += 16)
    if (col < c1) A[(int64_t)(k0 + r) * lda + col] = Bs[r][tx];
}

// The block row U12 of a complex-symmetric matrix factored without
// interchanges (A = L D L^T, stored as L and U = D L^T): U[i][j] = U[i][i]
// L[j][i] for rows i in [k0, k0 + b) and columns j in [c0, c1), from the
// already final L21 column block -- instead of the triangular solve
// U12 = L11^-1 A12 of the general factorisation.  A 32 x 32 tile per
// workgroup, transposed through LDS so that both reads and writes run along
// rows.
template <typename R>
__global__ __launch_bounds__(256) void u12_from_l21_kernel(cplx<R>* A, int64_t lda, int k0, int b, int c0, int c1) {
  __shared__ cplx<R> t[32][33];
  const int i0 = k0 + blockIdx.y * 32, j0 = c0 + blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows of 32 per pass
  for (int r = ty; r < 32; r += 8) {  // t[jj][ii] = L[j0 + jj][i0 + ii]
    const int j = j0 + r, i = i0 + tx;
    t[r][tx] = (j < c1 && i < k0 + b) ? A[(int64_t)j * lda + i] : cplx<R>{0, 0};
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + r, j = j0 + tx;
    if (i < k0 + b && j < c1) A[(int64_t)i * lda + j] = A[(int64_t)i * lda + i] * t[tx][r];
  }
}

// ------------------------------------------------------------------ update
// C (m x n complex, at Cp, ldc) -= A (m x kk, at Ap, lda) * B (kk x n, at Bp, ldb)
// Real view: C' (m x 2n) -= A' (m x 2kk) B'' (2kk x 2n).  Workgroup tile: 64
// rows x 64 real columns; each of the 4 wavefronts owns 32 x 32 as 2 x 2
// MFMA 16x16 tiles; K' processed in chunks of 16 real (8 complex).
template <typename R>
struct mfma_traits;
// Operand layout of the 16x16x4 MFMAs: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator element v of lane l is D[row(l, v)][l % 16]
// with row = l / 16 + 4 v for f64 and 4 (l / 16) + v for f32.
template <>
struct mfma_traits<double> {
  typedef d4 acc_t;
  static __device__ inline acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct mfma_traits<float> {
  typedef f4 acc_t;
  static __device__ inline acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Workgroup tile: BM rows x BN real columns (BN/2 complex); the 4 wavefronts
// form a 2 x 2 grid, each owning (BM/2) x (BN/2) as MFMA 16x16 tiles.  K' is
// streamed in stages of 16 real columns through double-buffered LDS: the
// global loads of stage s+1 are in flight while stage s runs on the matrix
// cores.  The tile is chosen per precision at run time (g_tile below).
#ifndef MOM_GEMM_BK
#define MOM_GEMM_BK 16
#endif
constexpr int kGemmBK = MOM_GEMM_BK;

// ASSIGN: C = -A B (C is not read) instead of C -= A B.
// skip_above >= 0: C's column j lies on the global diagonal of row
// j + skip_above (the trailing block of a symmetric factorisation), and tiles
// whose rows all lie above the diagonal are not computed (they are never
// read again).
template <typename R, int BM, int BN, bool ASSIGN = false>
__global__ __launch_bounds__(256) void cgemm_sub_kernel(int m, int n, int kk, const cplx<R>* __restrict__ Ap,
                                                        int64_t lda, const cplx<R>* __restrict__ Bp, int64_t ldb,
                                                        cplx<R>* __restrict__ Cp, int64_t ldc, int skip_above = -1) {
  if (skip_above >= 0 && (int)(blockIdx.y * BM + BM - 1) < (int)(blockIdx.x * (BN / 2)) + skip_above) return;
  typedef typename mfma_traits<R>::acc_t acc_t;
  constexpr int BK = kGemmBK;
  constexpr int TM = BM / 32, TN = BN / 32;          // MFMA tiles per wavefront
  constexpr int AE = BM * BK / 256;                  // A' reals per thread per stage
  constexpr int BE = (BK / 2) * (BN / 2) / 256;      // B complex per thread per stage
  static_assert(AE >= 1 && BE >= 1, "tile too small for 256 threads");
  __shared__ R As[2][BM][BK + 1];
  __shared__ R Bs[2][BK][BN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row0 = blockIdx.y * BM;
  const int colr0 = blockIdx.x * BN;
  const int wr = (wave >> 1) * (BM / 2), wc = (wave & 1) * (BN / 2);
  acc_t acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  const R* Ar = reinterpret_cast<const R*>(Ap);
  const int K2 = 2 * kk;
  const int nst = (K2 + BK - 1) / BK;
  constexpr int APR = BK / AE;                        // threads per A' row
  const int a_row = tid / APR, a_col = (tid % APR) * AE;
  constexpr int BPR = (BN / 2) / BE;                  // threads per complex B row
  const int b_row = tid / BPR, b_col = (tid % BPR) * BE;
  R ra[AE];
  cplx<R> rb[BE];
  auto load = [&](int st) {
    const int k0 = st * BK;
    const int gr = row0 + a_row;
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      const int gk = k0 + a_col + e;
      ra[e] = (gr < m && gk < K2) ? Ar[(int64_t)gr * 2 * lda + gk] : (R)0;
    }
    const int gk = (k0 >> 1) + b_row;
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      const int gj = (colr0 >> 1) + b_col + e;
      rb[e] = (gk < kk && gj < n) ? Bp[(int64_t)gk * ldb + gj] : cplx<R>{0, 0};
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < AE; ++e) As[buf][a_row][a_col + e] = ra[e];
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      const int c = 2 * (b_col + e);
      Bs[buf][2 * b_row][c] = rb[e].re;
      Bs[buf][2 * b_row][c + 1] = rb[e].im;
      Bs[buf][2 * b_row + 1][c] = -rb[e].im;
      Bs[buf][2 * b_row + 1][c + 1] = rb[e].re;
    }
  };
  load(0);
  store(0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int cur = st & 1;
    if (st + 1 < nst) load(st + 1);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      R af[TM], bf[TN];
#pragma unroll
      for (int a = 0; a < TM; ++a) af[a] = As[cur][wr + 16 * a + (lane & 15)][ks + (lane >> 4)];
#pragma unroll
      for (int b = 0; b < TN; ++b) bf[b] = Bs[cur][ks + (lane >> 4)][wc + 16 * b + (lane & 15)];
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = mfma_traits<R>::mfma(af[a], bf[b], acc[a][b]);
    }
    if (st + 1 < nst) store(cur ^ 1);
    __syncthreads();
  }
  R* Cr = reinterpret_cast<R*>(Cp);
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = row0 + wr + 16 * a + mfma_traits<R>::row(lane, v);
        const int cr = colr0 + wc + 16 * b + (lane & 15);
        if (r < m && cr < 2 * n) {
          R* c = Cr + (int64_t)r * 2 * ldc + cr;
          *c = ASSIGN ? -acc[a][b][v] : *c - acc[a][b][v];
        }
      }
}

// Workgroup tiles of the trailing update, chosen per precision at run time
// (mom_set_gemm_tile; the autotuner sets them).  Defaults: the fastest on
// MI355X for the N ~ 12k factorisation.
static int g_tile[2][2] = {{64, 128}, {64, 128}};  // [complex128, complex64][BM, BN]

// flops of the tiles a launch computes (all, or with skip_above those on or
// below the diagonal), for the roofline report
template <int BM, int BN>
static double cgemm_flops(int m, int n, int k, int skip_above) {
  if (skip_above < 0) return 8.0 * m * (double)n * k;
  double f = 0;
  for (int ty = 0; ty * BM < m; ++ty)
    for (int tx = 0; tx * (BN / 2) < n; ++tx)
      if (ty * BM + BM - 1 >= tx * (BN / 2) + skip_above)
        f += 8.0 * std::min(BM, m - ty * BM) * (double)std::min(BN / 2, n - tx * (BN / 2)) * k;
  return f;
}

static thread_local double g_last_gemm_flops = 0;  // of this thread's last launch_cgemm (timing report)

template <typename R, int BM, int BN, bool ASSIGN = false>
static int launch_cgemm_tile(int m, int n, int k, const cplx<R>* A, int64_t lda, const cplx<R>* B, int64_t ldb,
                             cplx<R>* C, int64_t ldc, hipStream_t s, int skip_above) {
  if ((m + BM - 1) / BM > 65535) return fail("cgemm: too many row tiles");
  dim3 grid((2 * n + BN - 1) / BN, (m + BM - 1) / BM);
  hipLaunchKernelGGL((cgemm_sub_kernel<R, BM, BN, ASSIGN>), grid, dim3(256), 0, s, m, n, k, A, lda, B, ldb, C, ldc,
                     skip_above);
  g_last_gemm_flops = cgemm_flops<BM, BN>(m, n, k, skip_above);
  return check_launch("cgemm_sub_kernel");
}

template <typename R, bool ASSIGN = false>
static int launch_cgemm(int m, int n, int k, const cplx<R>* A, int64_t lda, const cplx<R>* B, int64_t ldb,
                        cplx<R>* C, int64_t ldc, hipStream_t s, int skip_above = -1) {
  const int* t = g_tile[sizeof(R) == sizeof(double) ? 0 : 1];
#define MOM_TILE(BM, BN)                                                                                      \
  if (t[0] == BM && t[1] == BN)                                                                               \
    return launch_cgemm_tile<R, BM, BN, ASSIGN>(m, n, k, A, lda, B, ldb, C, ldc, s, skip_above);
  MOM_TILE(64, 64)
  MOM_TILE(64, 128)
  MOM_TILE(128, 64)
  MOM_TILE(128, 128)
#undef MOM_TILE
  return fail("cgemm: unsupported tile");
}

// ------------------------------------------------------------ solve phase
// perm: result of applying the interchanges ipiv[0..n) to the identity
// (a single workgroup, in LDS when it fits).
__global__ void build_perm_kernel(int n, const int32_t* ipiv, int32_t* perm) {
  extern __shared__ int32_t sp[];
  const bool lds = n <= 16000;
  int32_t* P = lds ? sp : perm;
  for (int i = threadIdx.x; i < n; i += blockDim.x) P[i] = i;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < n; ++k) {
      int p = ipiv[k];
      if (p <= k || p >= n) continue;  // a valid interchange has k < p < n
      int t = P[k];
      P[k] = P[p];
      P[p] = t;
    }
  __syncthreads();
  if (lds)
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = P[i];
}

// Triangular solves as a chain of 64-row blocks: block k needs x_0 .. x_{k-1}
// (lower; the mirror for upper).  Workgroup k, given block k by ticket, first
// copies the triangle of its diagonal block into LDS (packed by columns, the
// reciprocal of U's diagonal in place of the diagonal), then catches up with
// the blocks before it in order, accumulating A[k, j] x_j in registers.  Once
// its last dependency is in, wavefront 0 runs the substitution on the
// diagonal block (row i in lane i, x_k broadcast from lane k) -- the
// backward-stable form of LAPACK's ztrsv -- and stores x_k.
//
// The solution array itself is the flag: it is filled with a sentinel NaN
// before the launch, and a reader polls each element it needs until it is no
// longer the sentinel.  A finished block is therefore visible after one
// memory round trip, with no completion wait and no counter.  Tickets are
// taken in launch order, so a workgroup only waits for workgroups that are
// already running.
constexpr int kTrsvBS = 64;
constexpr int kMaxRhs = 8;  // right-hand sides per triangular solve

template <bool LOWER>
__device__ __forceinline__ int packed_index(int i, int k) {
  // column k of the strict lower triangle holds rows k+1 .. 63; of the upper
  // triangle (with diagonal) rows 0 .. k
  return LOWER ? k * (2 * kTrsvBS - k - 1) / 2 + (i - k - 1) : k * (k + 1) / 2 + i;
}

// Value of lane k (uniform k) in every lane, through the scalar unit.
__device__ __forceinline__ double lane_value(double v, int k) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_value(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

template <typename R>
__device__ __forceinline__ cplx<R> lane_value(cplx<R> v, int k) {
  return {lane_value(v.re, k), lane_value(v.im, k)};
}

// a - x y in four fused multiply-adds
template <typename R>
__device__ __forceinline__ cplx<R> fms(cplx<R> a, cplx<R> x, cplx<R> y) {
  return {fma(-x.re, y.re, fma(x.im, y.im, a.re)), fma(-x.re, y.im, fma(-x.im, y.re, a.im))};
}

// ------------------------------------ unpivoted panel, one row per lane
// The unpivoted panel of at most 64 columns with its rows in registers:
// lane i of every wavefront holds row i, wavefront w the columns
// [w C, w C + C) of it (C = 64 / W: W = 1 for complex64, 2 for complex128,
// whose 64 columns would not fit one lane's registers).  A step reads the
// pivot row's values from lane j through the scalar unit (v_readlane), so
// only the multipliers of the step's column cross wavefronts (through LDS,
// one barrier per column; none for W = 1).  Against the block-in-LDS
// kernels above, a 64-column panel of the 12 428-unknown preconditioner
// takes a fraction of the time; both stay on the latency-bound chain of the
// factorisation, whose late panels no longer hide behind the update GEMM.

// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>), each
// call inlined with its index a compile-time constant (register arrays stay
// in registers: no loop the unroller might keep rolled).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// The b x b diagonal block (b <= 64; rows and columns past b act as the
// identity) factored in place, A = L U, by one workgroup of 64 W threads.
// sym: the block of a complex-symmetric matrix whose updates reached only its
// lower triangle; its upper one is read mirrored.  The body is instantiated
// per wavefront (WV), so that every column index is a compile-time constant.
template <typename R, int W, int WV>
__device__ __forceinline__ void diag_rows_body(cplx<R>* A, int64_t lda, int k0, int b, int32_t* ipiv, int32_t* info,
                                               int sym, cplx<R> (*lx)[64]) {
  constexpr int B = 64, C = B / W, C0 = WV * C;
  const int i = threadIdx.x & 63;
  constexpr cplx<R> zero = {(R)0, (R)0}, one = {(R)1, (R)0};
  cplx<R> a[C];
  static_for<C>([&](auto cc) {
    const int c = C0 + cc;
    cplx<R> v = {(R)(i == c), (R)0};
    if (i < b && c < b) {
      const bool up = sym && c > i;
      v = A[(int64_t)(k0 + (up ? c : i)) * lda + k0 + (up ? i : c)];
    }
    a[cc] = v;
  });
  if (WV == 0 && i < b) ipiv[k0 + i] = k0 + i;
  static_for<B>([&](auto jc) {
    constexpr int j = jc;
    cplx<R> l = zero;
    if constexpr (j / C == WV) {
      constexpr int jj = j % C;
      const cplx<R> p = lane_value(a[jj], j);  // u_jj, final since step j - 1
      if (i == 0 && j < b && p.re == (R)0 && p.im == (R)0 && info[0] == 0) info[0] = k0 + j + 1;
      if (i > j) {
        l = a[jj] * cdiv(one, nonzero_pivot(p));
        a[jj] = l;
      } else if (i == j) {
        a[jj] = nonzero_pivot(p);
      }
      if constexpr (W > 1) lx[j & 1][i] = l;
    }
    if constexpr (W > 1) {
      __syncthreads();
      if constexpr (j / C != WV) l = lx[j & 1][i];
    }
    if (i > j) {
      static_for<C>([&](auto cc) {
        if constexpr (C0 + cc > j) a[cc] = fms(a[cc], l, lane_value(a[cc], j));
      });
    }
    __builtin_amdgcn_sched_barrier(0);
  });
  static_for<C>([&](auto cc) {
    const int c = C0 + cc;
    if (i < b && c < b) A[(int64_t)(k0 + i) * lda + k0 + c] = a[cc];
  });
}

template <typename R, int W>
__global__ __launch_bounds__(64 * W) void diag_rows_kernel(cplx<R>* A, int64_t lda, int k0, int b, int32_t* ipiv,
                                                         int32_t* info, int sym) {
  // on the factorisation's critical path, beside the look-ahead's update
  // GEMM on the same CUs: its instructions go first
  __builtin_amdgcn_s_setprio(3);
  __shared__ cplx<R> lx[2][64];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  static_for<W>([&](auto wv) {
    if (w == wv) diag_rows_body<R, W, wv>(A, lda, k0, b, ipiv, info, sym, lx);
  });
}

// L21 = A21 U11^-1 for the rows [k0 + b, n) below a factored diagonal block
// (b <= 64): 64 rows per workgroup of 64 W threads, row i in lane i as in
// diag_rows_kernel; U11 and the reciprocals of its diagonal are staged in LDS
// and read at uniform addresses (broadcast).  Column c is finished in step c,
// x_c = a_c / u_cc, and the columns after it updated, a_k -= x_c u_ck, in
// groups of 16 whose reads of U stay next to their use.
template <typename R, int W, int WV>
__device__ __forceinline__ void below_rows_body(cplx<R>* A, int64_t lda, int n, int k0, int b, const cplx<R>* U,
                                                const cplx<R>* dinv, cplx<R> (*lx)[64]) {
  constexpr int B = 64, C = B / W, C0 = WV * C, G = 16;
  const int i = threadIdx.x & 63;
  constexpr cplx<R> zero = {(R)0, (R)0};
  const int r = k0 + b + blockIdx.x * B + i;
  const bool row = r < n;
  cplx<R> a[C];
  static_for<C>([&](auto cc) {
    const int c = C0 + cc;
    cplx<R> v = zero;
    if (row && c < b) v = A[(int64_t)r * lda + k0 + c];
    a[cc] = v;
  });
  static_for<B>([&](auto cs) {
    constexpr int c = cs;
    cplx<R> x = zero;
    if constexpr (c / C == WV) {
      constexpr int cl = c % C;
      x = a[cl] * dinv[c];
      a[cl] = x;
      if constexpr (W > 1) lx[c & 1][i] = x;
    }
    if constexpr (W > 1) {
      __syncthreads();
      if constexpr (c / C != WV) x = lx[c & 1][i];
    }
    static_for<(C + G - 1) / G>([&](auto g) {
      // an offset the compiler cannot see to be zero, so that the group's
      // reads of U stay behind it (hoisted, as invariant loads, they would
      // all be live at once)
      int z;
      asm volatile("s_mov_b32 %0, 0" : "=s"(z));
      static_for<G>([&](auto e) {
        constexpr int cc = g * G + e;
        if constexpr (cc < C && C0 + cc > c) a[cc] = fms(a[cc], x, U[c * B + C0 + cc + z]);
      });
      __builtin_amdgcn_sched_barrier(0);
    });
  });
  static_for<C>([&](auto cc) {
    const int c = C0 + cc;
    if (row && c < b) A[(int64_t)r * lda + k0 + c] = a[cc];
  });
}

template <typename R, int W>
__global__ __launch_bounds__(64 * W) void below_rows_kernel(cplx<R>* A, int64_t lda, int n, int k0, int b) {
  constexpr int B = 64;
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* U = reinterpret_cast<cplx<R>*>(smem_raw);  // [B][B]
  cplx<R>* dinv = U + B * B;                           // [B]
  __shared__ cplx<R> lx[2][B];
  __builtin_amdgcn_s_setprio(3);  // on the critical path, as diag_rows_kernel
  const int tid = threadIdx.x;
  for (int e = tid; e < B * B; e += 64 * W) {
    const int r = e / B, c = e % B;
    U[e] = (r < b && c < b) ? A[(int64_t)(k0 + r) * lda + k0 + c] : cplx<R>{(R)(r == c), (R)0};
  }
  __syncthreads();
  for (int c = tid; c < B; c += 64 * W) dinv[c] = cdiv(cplx<R>{(R)1, (R)0}, nonzero_pivot(U[c * B + c]));
  __syncthreads();
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  static_for<W>([&](auto wv) {
    if (w == wv) below_rows_body<R, W, wv>(A, lda, n, k0, b, U, dinv, lx);
  });
}

// Set-up of a solve in one launch: the two block tickets cleared (words 0
// and 2), X = P B (rows gathered by perm, n x nrhs, rounded to R) and Y (the
// lower solve's output) filled with the sentinel.  The failure word (word 4)
// is left alone: it is sticky, set by any solve whose wait ran out and
// cleared only by the caller, so that a check after a run of solves (a GMRES
// solve calls this hundreds of times) still sees a failure in any of them.
template <typename R, typename TB>
__global__ void getrs_prep_kernel(int n, int nrhs, const int32_t* perm, const cplx<TB>* B, cplx<R>* X, cplx<R>* Y,
                                  unsigned* flags) {
  gmres_step_priority();
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q == 0 || q == 2) flags[q] = 0u;
  if (q >= n * nrhs) return;
  const int i = q / nrhs, c = q % nrhs;
  const cplx<TB> b = B[(int64_t)perm[i] * nrhs + c];
  X[q] = {(R)b.re, (R)b.im};
  Y[q] = {Sentinel<R>::value(), Sentinel<R>::value()};
}

// Solves with the right-hand sides B (n x NR) into Xout (n x NR, filled with
// the sentinel); *ticket hands out the blocks, *error is set to 1 when a
// bounded wait ran out.  With INV the diagonal blocks are applied as the
// precomputed inverses Dinv (64 x 64 row-major per block, trtri_blocks_kernel)
// by a product spread over the workgroup instead of the substitution.  When
// B_reset is B, each block sets the rows of B it has read to the sentinel
// (B then serves as the next solve's output; no other block reads them).
// With out64 the solution is also written there in complex128.  Blocks of
// BS rows (64, or 128 with inverted blocks) and BS / 16 column segments of
// 16 per row: BS^2 / 16 threads.  With Mchain (64-row inverted blocks: the
// products M_k = Dinv_k A_k,prev of each block's inverse and its coupling
// to the block solved just before it, chain_matrix_kernel) a block forms
// y_k = Dinv_k (b_k - sum of all earlier couplings but that one) while it
// waits for the previous block, and then only x_k = y_k - M_k x_prev: one
// product on the chain's critical path instead of two.
template <typename R, bool LOWER, int NR, bool INV, int BS = kTrsvBS>
__global__ __launch_bounds__(BS * BS / 16) void trsv_kernel(int n, const cplx<R>* __restrict__ A, int64_t lda,
                                                            const cplx<R>* B, cplx<R>* Xout,
                                                            const cplx<R>* __restrict__ Dinv,
                                                            const cplx<R>* __restrict__ Mchain, cplx<R>* B_reset,
                                                            z64* __restrict__ out64, unsigned* ticket,
                                                            unsigned* error, unsigned spin_limit) {
  gmres_step_priority();
  static_assert(INV || BS == kTrsvBS, "the substitution works on 64-row blocks");
  constexpr int Q = BS / 16, NT = BS * Q;
  constexpr int TRI = INV ? 1 : (LOWER ? BS * (BS - 1) / 2 : BS * (BS + 1) / 2);
  __shared__ cplx<R> Dk[TRI];
  __shared__ cplx<R> xs[BS][NR];
  __shared__ cplx<R> part[Q][BS][NR];
  __shared__ cplx<R> ys[INV ? BS : 1][NR];
  __shared__ int s_blk;
  __shared__ int s_abort;
  const int tid = threadIdx.x;
  const int r = tid % BS, q = tid / BS;  // row of the block, column segment
  const int nblk = (n + BS - 1) / BS;
  if (tid == 0) {
    const unsigned t = atomicAdd(ticket, 1u);
    // a ticket past the grid means that two solves share this workspace at
    // once: report it instead of indexing outside the factors
    if (t >= (unsigned)nblk) {
      atomicOr(error, 1u);
      s_blk = -1;
    } else {
      s_blk = LOWER ? (int)t : nblk - 1 - (int)t;
    }
    s_abort = 0;
  }
  __syncthreads();
  if (s_blk < 0) return;
  const int blk = s_blk;
  const int k0 = blk * BS, b = min(BS, n - k0);
  const int row = k0 + r;
  // this thread's 16 entries of the inverse's row r: loaded up front in the
  // 64-row form; in the 128-row form (128 VGPRs per thread at 1024 threads)
  // only after the dependency loop, so that they are not live beside the
  // prefetched tile of A
  constexpr bool EARLY_DV = BS == kTrsvBS;
  cplx<R> dv[INV ? 16 : 1];
  auto load_dv = [&]() {
    const cplx<R>* drow = Dinv + (int64_t)blk * BS * BS + r * BS + 16 * q;
#pragma unroll
    for (int e = 0; e < 16; ++e) dv[e] = drow[e];
  };
  auto partsum = [&](int i, int c) {
    cplx<R> t = part[0][i][c];
#pragma unroll
    for (int k = 1; k < Q; ++k) t += part[k][i][c];
    return t;
  };
  if (INV && EARLY_DV) load_dv();
  for (int e = tid; !INV && e < BS * BS; e += NT) {
    const int i = e / BS, k = e % BS;
    if (LOWER ? i > k : i <= k) {
      cplx<R> v = (i < b && k < b) ? A[(int64_t)(k0 + i) * lda + k0 + k] : cplx<R>{(R)(i == k), 0};
      if (!LOWER && i == k) v = cdiv(cplx<R>{1, 0}, v);
      Dk[packed_index<LOWER>(i, k)] = v;
    }
  }
  cplx<R> acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = {0, 0};
  const int need = LOWER ? blk : nblk - 1 - blk;  // blocks before this one in solve order
  // This thread's 16 entries of A[row, block d] for dependency d, loaded
  // before the workgroup waits for block d (the wait covers their latency).
  cplx<R> la[16];
  auto fetch = [&](int d) {
    const int j0 = (LOWER ? d : nblk - 1 - d) * BS + 16 * q;
    const cplx<R>* arow = A + (int64_t)row * lda + j0;  // one base, immediate offsets
    if (r < b && j0 + 16 <= n) {
#pragma unroll
      for (int e = 0; e < 16; ++e) la[e] = arow[e];
    } else {
#pragma unroll
      for (int e = 0; e < 16; ++e) la[e] = (r < b && j0 + e < n) ? arow[e] : cplx<R>{0, 0};
    }
  };
  // x of dependency d into xs; false when a bounded wait ran out (the block
  // before this one never arrived: the kernel gives up)
  auto wait_dep = [&](int d) {
    const int j0 = (LOWER ? d : nblk - 1 - d) * BS;
    for (int t = tid; t < BS * NR; t += NT) {
      const int i = t / NR, c = t % NR;
      cplx<R> v = {0, 0};
      if (j0 + i < n) {
        unsigned spins = 0;
        for (;;) {
          v = load_coherent_c(Xout + (int64_t)(j0 + i) * NR + c);
          if (!is_sentinel(v)) break;
          poll_pause();
          if (++spins > spin_limit) {
            s_abort = 1;
            break;
          }
        }
      }
      xs[i][c] = v;
    }
    __syncthreads();
    if (s_abort) {
      if (tid == 0) atomicOr(error, 1u);
      return false;
    }
    return true;
  };
  // the chain form leaves the last dependency (the previous block) to M_k
  const bool chain = INV && BS == kTrsvBS && Mchain != nullptr && need > 0;
  for (int d = 0; d < (chain ? need - 1 : need); ++d) {
    fetch(d);
    if (!wait_dep(d)) return;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] += la[e] * xs[16 * q + e][c];
      // in the 128-row form keep the LDS reads to 4 columns in flight
      // (hoisting all 16 exceeds 128 VGPRs at 1024 threads and spills)
      if (!EARLY_DV && (e & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
  }
  if (INV && !EARLY_DV) load_dv();
#pragma unroll
  for (int c = 0; c < NR; ++c) part[q][r][c] = acc[c];
  __syncthreads();
  if (INV) {
    // x_k = Dinv_k (b_k - sum): v into LDS, then one product per (row, quarter)
    for (int t = tid; t < BS * NR; t += NT) {
      const int i = t / NR, c = t % NR;
      xs[i][c] = i < b ? B[(int64_t)(k0 + i) * NR + c] - partsum(i, c) : cplx<R>{0, 0};
    }
    __syncthreads();
    for (int t = tid; B_reset && t < BS * NR; t += NT)
      if (t / NR < b) B_reset[(int64_t)(k0 + t / NR) * NR + t % NR] = {Sentinel<R>::value(), Sentinel<R>::value()};
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = {0, 0};
#pragma unroll
    for (int e = 0; e < 16; ++e) {
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] += dv[e] * xs[16 * q + e][c];
      if (!EARLY_DV && (e & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) part[q][r][c] = acc[c];
    __syncthreads();
    if (chain) {
      // y = Dinv (b - earlier couplings) kept in LDS; then x = y - M x_prev
      for (int t = tid; t < BS * NR; t += NT) ys[t / NR][t % NR] = partsum(t / NR, t % NR);
      cplx<R> mv[16];
      const cplx<R>* mrow = Mchain + (int64_t)blk * BS * BS + r * BS + 16 * q;
#pragma unroll
      for (int e = 0; e < 16; ++e) mv[e] = mrow[e];
      if (!wait_dep(need - 1)) return;
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = {0, 0};
#pragma unroll
      for (int e = 0; e < 16; ++e) {
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[c] += mv[e] * xs[16 * q + e][c];
      }
#pragma unroll
      for (int c = 0; c < NR; ++c) part[q][r][c] = acc[c];
      __syncthreads();
    }
    for (int t = tid; t < BS * NR; t += NT) {
      const int i = t / NR, c = t % NR;
      if (i < b) {
        const cplx<R> x = chain ? ys[i][c] - partsum(i, c) : partsum(i, c);
        store_coherent_c(Xout + (int64_t)(k0 + i) * NR + c, x);
        if (out64) out64[(int64_t)(k0 + i) * NR + c] = {(double)x.re, (double)x.im};
      }
    }
    return;
  }
  if (tid >= 64) return;
  // wavefront 0: v = b - sum, then substitution with the diagonal block; x_k
  // is read from lane k by every lane (uniform k: scalar broadcast)
  const int i = tid;
  cplx<R> v[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c)
    v[c] = i < b ? B[(int64_t)(k0 + i) * NR + c] - partsum(i, c) : cplx<R>{0, 0};
  if (B_reset && i < b) {
#pragma unroll
    for (int c = 0; c < NR; ++c)
      B_reset[(int64_t)(k0 + i) * NR + c] = {Sentinel<R>::value(), Sentinel<R>::value()};
  }
  if (LOWER) {
#pragma unroll
    for (int k = 0; k < BS - 1; ++k) {
      const cplx<R> l = i > k ? Dk[packed_index<true>(i, k)] : cplx<R>{0, 0};
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const cplx<R> xk = {lane_value(v[c].re, k), lane_value(v[c].im, k)};
        v[c] -= l * xk;
      }
    }
  } else {
#pragma unroll
    for (int k = BS - 1; k >= 0; --k) {
      const cplx<R> rd = Dk[packed_index<false>(k, k)];
      const cplx<R> u = i < k ? Dk[packed_index<false>(i, k)] : cplx<R>{0, 0};
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const cplx<R> xk = cplx<R>{lane_value(v[c].re, k), lane_value(v[c].im, k)} * rd;
        v[c] = i == k ? xk : v[c] - u * xk;
      }
    }
  }
  if (i < b) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      store_coherent_c(Xout + (int64_t)(k0 + i) * NR + c, v[c]);
      if (out64) out64[(int64_t)(k0 + i) * NR + c] = {(double)v[c].re, (double)v[c].im};
    }
  }
}

// R = B - Z X (FP64); when rowsum is given it also receives
// d_r sum_k (|Re z_rk| + |Im z_rk|) d_k (d = 1 when dsc is null), the row
// sums of the diagonally scaled matrix (for its norm).  HBM-bound on the one
// pass over Z: each wavefront takes 4 rows, so one load of x_k serves 4 rows,
// and keeps 4 independent 16-byte streams of Z in flight (non-temporal: Z
// is read once, x stays in the caches).  SCALED: Rout = D Z D X instead
// (D = diag(dsc); B and rowsum unused), the GMRES product of the scaled system.
template <int NR, bool ROWSUM, bool SCALED = false>
__global__ __launch_bounds__(256) void residual_kernel(int n, const z64* __restrict__ Z, int64_t ldz,
                                                       const z64* __restrict__ X, const z64* __restrict__ B,
                                                       z64* __restrict__ Rout, double* __restrict__ rowsum,
                                                       const double* __restrict__ dsc) {
  constexpr int RW = 4;
  const int lane = threadIdx.x & 63;
  const int r0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * RW;
  if (r0 >= n) return;
  typedef double dv2 __attribute__((ext_vector_type(2)));
  const dv2* zr[RW];
#pragma unroll
  for (int q = 0; q < RW; ++q) zr[q] = reinterpret_cast<const dv2*>(Z + (int64_t)min(r0 + q, n - 1) * ldz);
  z64 s[RW][NR];
  double a[RW];
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    a[q] = 0.0;
#pragma unroll
    for (int c = 0; c < NR; ++c) s[q][c] = {0, 0};
  }
#pragma unroll 2
  for (int k = lane; k < n; k += 64) {
    z64 x[NR];
    const double dk = ((ROWSUM || SCALED) && dsc) ? dsc[k] : 1.0;
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      x[c] = X[(int64_t)k * NR + c];
      if (SCALED) x[c] = {x[c].re * dk, x[c].im * dk};
    }
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      const dv2 v = __builtin_nontemporal_load(zr[q] + k);
      const z64 zz = {v.x, v.y};
      if (ROWSUM) a[q] += (fabs(zz.re) + fabs(zz.im)) * dk;
#pragma unroll
      for (int c = 0; c < NR; ++c) s[q][c] += zz * x[c];
    }
  }
#pragma unroll
  for (int q = 0; q < RW; ++q) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      double re = s[q][c].re, im = s[q][c].im;
      for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_down(re, off, 64);
        im += __shfl_down(im, off, 64);
      }
      if (lane == 0 && r0 + q < n) {
        if (SCALED)
          Rout[(int64_t)(r0 + q) * NR + c] = {dsc[r0 + q] * re, dsc[r0 + q] * im};
        else
          Rout[(int64_t)(r0 + q) * NR + c] = B[(int64_t)(r0 + q) * NR + c] - z64{re, im};
      }
    }
    if (ROWSUM) {
      double t = a[q];
      for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
      if (lane == 0 && r0 + q < n) rowsum[r0 + q] = t * (dsc ? dsc[r0 + q] : 1.0);
    }
  }
}

// GMRES orthogonalisation helpers over a basis V (j1 x n x nrhs, complex128):
// h[i, c] = sum_r conj(V[i, r, c]) w[r, c]   (one workgroup per basis vector)
__global__ __launch_bounds__(256) void basis_dot_kernel(int n, int nrhs, const z64* __restrict__ V,
                                                        const z64* __restrict__ w, z64* __restrict__ h) {
  const int i = blockIdx.x, tid = threadIdx.x;
  z64 acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  const z64* Vi = V + (int64_t)i * n * nrhs;
  for (int r = tid; r < n; r += 256)
    for (int c = 0; c < nrhs; ++c) {
      const z64 v = Vi[(int64_t)r * nrhs + c], x = w[(int64_t)r * nrhs + c];
      acc[c] += z64{v.re * x.re + v.im * x.im, v.re * x.im - v.im * x.re};
    }
  __shared__ z64 part[4][4];
  const int lane = tid & 63, wave = tid >> 6;
  for (int c = 0; c < nrhs; ++c) {
    double re = acc[c].re, im = acc[c].im;
    for (int off = 32; off > 0; off >>= 1) {
      re += __shfl_down(re, off, 64);
      im += __shfl_down(im, off, 64);
    }
    if (lane == 0) part[wave][c] = {re, im};
  }
  __syncthreads();
  if (tid < nrhs) {
    z64 t = part[0][tid];
    t += part[1][tid];
    t += part[2][tid];
    t += part[3][tid];
    h[(int64_t)i * nrhs + tid] = t;
  }
}

// w[r, c] -= sum_{i < j1} V[i, r, c] h[i, c]
__global__ void basis_sub_kernel(int n, int nrhs, int j1, const z64* __restrict__ V, const z64* __restrict__ h,
                                 z64* __restrict__ w) {
  gmres_step_priority();
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * nrhs) return;
  const int c = q % nrhs;
  z64 s = {0, 0};
  for (int i = 0; i < j1; ++i) s += V[(int64_t)i * n * nrhs + q] * h[(int64_t)i * nrhs + c];
  w[q] -= s;
}

__global__ void cvt_z_to_c_kernel(int64_t n, const double* src, float* dst) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// X += D (D complex64), and column maxima of |D| and |X| (cabs1) for the stopping test.
__global__ void refine_update_kernel(int n, int nrhs, const c32* D, z64* X, unsigned long long* dmax,
                                     unsigned long long* xmax) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * nrhs) return;
  int c = q % nrhs;
  z64 d = {(double)D[q].re, (double)D[q].im};
  z64 x = X[q] + d;
  X[q] = x;
  double ad = fabs(d.re) + fabs(d.im), ax = fabs(x.re) + fabs(x.im);
  atomicMax(dmax + c, (unsigned long long)__double_as_longlong(ad));
  atomicMax(xmax + c, (unsigned long long)__double_as_longlong(ax));
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ----------------------------------------------------------------- drivers
static int g_ncu = 0;

// Optional event timing of selected launches (the benchmark's roofline
// report): while a kind is on, each of its launches is bracketed by a pair
// of HIP events on the launch's own stream, and its algorithmic work (flops
// or bytes) is kept beside them.  Launches come from several host threads
// (one per factorisation in flight, and the lanes' solves), so slots are
// taken under a lock.
enum TimingKind { kTimeGemm32 = 0, kTimeSymv = 1, kTimeChain32 = 2, kTimeChain64 = 3, kTimeGemm64 = 4, kTimeKinds = 5 };
struct KernelTiming {
  bool on = false;
  std::vector<hipEvent_t> ev;
  std::vector<double> work;
  size_t used = 0;
};
static KernelTiming g_kt[kTimeKinds];
static std::mutex g_kt_mu;

// slot of a timed launch of `kind` (its start event recorded), or -1
static int timing_begin(int kind, hipStream_t s) {
  KernelTiming& t = g_kt[kind];
  int slot;
  {
    std::lock_guard<std::mutex> lock(g_kt_mu);
    if (!t.on || t.used + 2 > t.ev.size()) return -1;
    slot = (int)t.used;
    t.used += 2;
    t.work.push_back(0.0);
  }
  (void)hipEventRecord(t.ev[slot], s);
  return slot;
}
static void timing_end(int kind, int slot, hipStream_t s, double work) {
  if (slot < 0) return;
  KernelTiming& t = g_kt[kind];
  (void)hipEventRecord(t.ev[slot + 1], s);
  std::lock_guard<std::mutex> lock(g_kt_mu);
  t.work[slot / 2] = work;
}

static int device_cus() {
  if (g_ncu == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&g_ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) g_ncu = 0;
  }
  return g_ncu;
}

// Panel publication buffers for up to 256 workgroups: a slot per (column,
// workgroup), a candidate row per (column, workgroup) and row c per column.
template <typename R>
static size_t getrf_work_bytes(int n, int jb) {
  size_t b = 0;
  b += sizeof(PivotSlot) * (size_t)jb * 256 + 256;
  b += sizeof(cplx<R>) * ((size_t)jb * 256 * jb + (size_t)jb * jb) + 256;
  return b + 1024;
}

// Side stream and events of the look-ahead, one set per caller stream
// (created on first use and kept until the process ends), so that
// factorisations enqueued on different streams never share a side stream.
struct Lookahead {
  hipStream_t side = nullptr;
  std::vector<hipEvent_t> ev;
};
static std::mutex g_la_mu;
static std::map<hipStream_t, Lookahead> g_la;

static int lookahead_for(hipStream_t s, size_t need, Lookahead** out) {
  std::lock_guard<std::mutex> lock(g_la_mu);
  Lookahead& la = g_la[s];
  if (!la.side) MOM_TRY(hipStreamCreateWithFlags(&la.side, hipStreamNonBlocking));
  while (la.ev.size() < need) {
    hipEvent_t e;
    MOM_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    la.ev.push_back(e);
  }
  *out = &la;
  return 0;
}

template <typename R>
static int launch_panel(cplx<R>* A, int64_t lda, int n, int k0, int b, int32_t* ipiv, int32_t* info,
                        PivotSlot* slots, cplx<R>* rowbuf, int gmax, hipStream_t s) {
  const int m = n - k0;
  const size_t lds_max = 152 * 1024;  // of the 160 KB per workgroup, beside the static rows
  if (int rc0 = reserve_lds(reinterpret_cast<const void*>(panel_kernel<R>), lds_max, "panel_kernel")) return rc0;
  // as few workgroups as the LDS allows (fewer slots to poll per column)
  int rp = std::max(1, (int)(lds_max / ((size_t)b * sizeof(cplx<R>))));
  int G = (m + rp - 1) / rp;
  if (G > gmax) return fail("mom_getrf: matrix too large for the LDS-resident panel (reduce the panel width)");
  G = std::max(G, std::min(gmax, (m + 255) / 256));  // at least one row per thread block of work
  rp = (m + G - 1) / G;
  G = (m + rp - 1) / rp;
  // publication buffers of this panel: [b][G] slots, [b][G][b] candidate
  // rows, then [b][b] diagonal rows, all set to the sentinel first
  const int64_t nslot = (int64_t)b * G, nrow = (int64_t)b * G * b + (int64_t)b * b;
  const unsigned rgrid = (unsigned)std::min<int64_t>(2048, (nrow + 255) / 256);
  hipLaunchKernelGGL(panel_reset_kernel<R>, dim3(rgrid), dim3(256), 0, s, slots, nslot, rowbuf, nrow);
  int rc = check_launch("panel_reset_kernel");
  if (rc) return rc;
  PanelArgs<R> pa{A, lda, n, k0, b, rp, g_spin_limit, ipiv, info, slots, rowbuf, rowbuf + (int64_t)b * G * b};
  // An ordinary launch, not a cooperative one: cooperative launches are
  // serialised across streams, which would stop the panels of concurrent
  // factorisations from overlapping.  Every workgroup of the panel is
  // resident together all the same: the grid is at most G <= max_groups <=
  // CUs / (factorisations in flight) workgroups; the other kernels that can
  // hold a CU meanwhile either run to completion on their own (updates,
  // swaps, fills) or wait only for workgroups that are already running (the
  // ticket-ordered triangular solves), so they release their CUs and the
  // dispatcher places the rest of the panel.
  hipLaunchKernelGGL(panel_kernel<R>, dim3(G), dim3(256), (size_t)rp * b * sizeof(cplx<R>), s, pa);
  return check_launch("panel_kernel");
}

// Block width of the trailing updates, per precision (0: the panel width).
// A block of nb columns is factored as panels of jb columns, each followed by
// the update of the block's remaining columns; the rest of the matrix then
// takes one update of depth nb.  Set by mom_set_lu_block (the autotuner).
static int g_lu_nb[2] = {0, 0};

// Row blocks of the complex64 solves with inverted diagonal blocks (64 or
// 128), set by mom_set_solve_rows.
static int g_solve_rows = 64;
static int g_chain = 1;  // chain form of the 64-row inverted-block solves (mom_set_solve_chain)

// U12 = L11^-1 A12 for the block row [k0, k0 + b) of columns [c0, c1), with
// b up to 256: 128-row triangular solves and the update between them.
template <typename R>
static int trsm_block(cplx<R>* A, int64_t lda, int k0, int b, int c0, int c1, hipStream_t s) {
  int rc;
  for (int q0 = k0; q0 < k0 + b; q0 += 128) {
    const int bq = std::min(128, k0 + b - q0);
    hipLaunchKernelGGL(trsm_lower_unit_kernel<R>, dim3((c1 - c0 + 15) / 16), dim3(256), 0, s, A, lda, q0, bq, c0,
                       c1);
    if ((rc = check_launch("trsm_lower_unit_kernel"))) return rc;
    const int q1 = q0 + bq;
    if (q1 < k0 + b &&
        (rc = launch_cgemm<R>(k0 + b - q1, c1 - c0, bq, A + (int64_t)q1 * lda + q0, lda, A + (int64_t)q0 * lda + c0,
                              lda, A + (int64_t)q1 * lda + c0, lda, s)))
      return rc;
  }
  return 0;
}

// The pivots [k0, k0 + b) applied to columns [c0, c1), in groups of <= 128
// (the row interchanges are sequential, so groups in order are the same).
template <typename R>
static int laswp_block(cplx<R>* A, int64_t lda, int n, int c0, int c1, int k0, int b, const int32_t* ipiv,
                       hipStream_t s) {
  if (c1 <= c0) return 0;
  for (int q0 = k0; q0 < k0 + b; q0 += 128) {
    hipLaunchKernelGGL(laswp_kernel<R>, dim3((c1 - c0 + 15) / 16), dim3(256), 0, s, A, lda, n, c0, c1, q0,
                       std::min(128, k0 + b - q0), ipiv);
    int rc = check_launch("laswp_kernel");
    if (rc) return rc;
  }
  return 0;
}

// Trailing update of rows [r0, n) x columns [c0, c1) by the factored block
// column [k0, k0 + b): U12 = L11^-1 A12, then A22 -= L21 U12 (depth b).
// sym (a complex-symmetric matrix without interchanges): U12 = D L21^T from
// the final L21 instead, and only the tiles on or below the diagonal updated
// (the upper triangle is never read: every later U12 comes from L21 too).
template <typename R>
static int update_block(cplx<R>* A, int64_t lda, int n, int k0, int b, int r0, int c0, int c1, hipStream_t s,
                        bool sym = false) {
  if (c1 <= c0) return 0;
  int rc;
  if (sym) {
    hipLaunchKernelGGL(u12_from_l21_kernel<R>, dim3((c1 - c0 + 31) / 32, (b + 31) / 32), dim3(256), 0, s, A, lda, k0,
                       b, c0, c1);
    if ((rc = check_launch("u12_from_l21_kernel"))) return rc;
  } else if ((rc = trsm_block<R>(A, lda, k0, b, c0, c1, s))) {
    return rc;
  }
  const int mm = n - r0, nn = c1 - c0;
  if (mm <= 0) return 0;
  const int tkind = sizeof(R) == sizeof(double) ? kTimeGemm64 : kTimeGemm32;
  const int tslot = timing_begin(tkind, s);
  if ((rc = launch_cgemm<R>(mm, nn, b, A + (int64_t)r0 * lda + k0, lda, A + (int64_t)k0 * lda + c0, lda,
                            A + (int64_t)r0 * lda + c0, lda, s, sym ? c0 - r0 : -1)))
    return rc;
  timing_end(tkind, tslot, s, g_last_gemm_flops);
  return 0;
}

// The unpivoted panel of columns [k0, k0 + b) (b <= kNoPivPanelMax): its
// diagonal block, then the rows below.
template <typename R>
static int launch_panel_nopiv(cplx<R>* A, int64_t lda, int n, int k0, int b, int32_t* ipiv, int32_t* info,
                              hipStream_t s, bool sym = false) {
  if (b > kNoPivPanelMax<R>()) return fail("mom_getrf_nopiv: panel wider than the diagonal-block kernel");
  if (b <= 64) {  // rows in registers
    // wavefronts sharing a row: its 64 columns (and, below, a row of U11 per
    // step) within one lane's registers
    constexpr int WD = sizeof(R) == sizeof(double) ? 2 : 1, WB = 2;
    hipLaunchKernelGGL((diag_rows_kernel<R, WD>), dim3(1), dim3(64 * WD), 0, s, A, lda, k0, b, ipiv, info, (int)sym);
    int rc = check_launch("diag_rows_kernel");
    if (rc) return rc;
    const int below = n - k0 - b;
    if (below <= 0) return 0;
    const size_t lds = sizeof(cplx<R>) * (64 * 64 + 64);
    if ((rc = reserve_lds((const void*)below_rows_kernel<R, WB>, lds, "below_rows_kernel"))) return rc;
    hipLaunchKernelGGL((below_rows_kernel<R, WB>), dim3((below + 63) / 64), dim3(64 * WB), lds, s, A, lda, n, k0, b);
    return check_launch("below_rows_kernel");
  }
  const size_t most = sizeof(cplx<R>) * (size_t)kNoPivPanelMax<R>() * kNoPivPanelMax<R>();
  if (int rc0 = reserve_lds((const void*)diag_lu_kernel<R>, most, "diag_lu_kernel")) return rc0;
  hipLaunchKernelGGL(diag_lu_kernel<R>, dim3(1), dim3(1024), sizeof(cplx<R>) * (size_t)b * b, s, A, lda, k0, b, ipiv,
                     info, (int)sym);
  int rc = check_launch("diag_lu_kernel");
  if (rc) return rc;
  const int below = n - k0 - b;
  if (below <= 0) return 0;
  hipLaunchKernelGGL(panel_below_kernel<R>, dim3((below + 15) / 16), dim3(256), 0, s, A, lda, n, k0, b);
  return check_launch("panel_below_kernel");
}

struct PanelWork {
  PivotSlot* slots;
  void* rowbuf;
  int gmax;
};

// Factor the block column [k0, k0 + nb) (rows k0..n) with panels of width jb:
// after each panel, its interchanges are applied to the block's other
// columns and the block's remaining columns are updated.
template <typename R>
static int factor_block(cplx<R>* A, int64_t lda, int n, int k0, int nb, int jb, int32_t* ipiv, int32_t* info,
                        const PanelWork& pw, bool pivot, hipStream_t s, bool sym = false) {
  int rc;
  const int kend = k0 + nb;
  for (int q0 = k0; q0 < kend; q0 += jb) {
    const int bq = std::min(jb, kend - q0), q1 = q0 + bq;
    if (pivot) {
      if ((rc = launch_panel<R>(A, lda, n, q0, bq, ipiv, info, pw.slots, (cplx<R>*)pw.rowbuf, pw.gmax, s)))
        return rc;
      if ((rc = laswp_block<R>(A, lda, n, k0, q0, q0, bq, ipiv, s))) return rc;
    } else if ((rc = launch_panel_nopiv<R>(A, lda, n, q0, bq, ipiv, info, s, sym))) {
      return rc;
    }
    if (q1 < kend) {
      if (pivot && (rc = laswp_block<R>(A, lda, n, q1, kend, q0, bq, ipiv, s))) return rc;
      if ((rc = update_block<R>(A, lda, n, q0, bq, q1, q1, kend, s, sym))) return rc;
    }
  }
  return 0;
}

// Right-looking blocked LU with one block of look-ahead: while the trailing
// update of block p runs on the caller's stream, block p+1 (whose columns
// were updated first) is factored on the side stream.
template <typename R>
static int getrf(int n, cplx<R>* A, int64_t lda, int32_t* ipiv, int32_t* info, void* work, size_t work_bytes, int jb,
                 int max_groups, hipStream_t s, bool pivot = true, bool sym = false) {
  if (sym && pivot) return fail("getrf: the symmetric form has no row interchanges");
  if (n <= 0) return 0;
  if (jb <= 0 || jb > 128 || (jb % 8) != 0) return fail("mom_getrf: panel width must be a multiple of 8 in [8, 128]");
  if (work_bytes < getrf_work_bytes<R>(n, jb)) return fail("mom_getrf: workspace too small");
  const int ncu = device_cus();
  if (ncu <= 0) return fail("mom_getrf: no device");
  int nb = g_lu_nb[sizeof(R) == sizeof(double) ? 0 : 1];
  nb = nb <= 0 ? jb : std::max(jb, nb - nb % jb);
  // unpivoted: the diagonal block of a panel lives in one workgroup's LDS
  // (64 columns in complex128); the update block keeps its width
  const int jp = pivot ? jb : std::min(jb, kNoPivPanelMax<R>());
  // slot arrays hold 256 workgroups
  const int gmax = std::min(std::min(ncu, 256), max_groups > 0 ? max_groups : 256);
  unsigned char* w = reinterpret_cast<unsigned char*>(work);
  auto take = [&](size_t bytes) {
    unsigned char* p = w;
    w += (bytes + 255) & ~size_t(255);
    return p;
  };
  const int nblock = (n + nb - 1) / nb;
  PanelWork pw;
  pw.slots = reinterpret_cast<PivotSlot*>(take(sizeof(PivotSlot) * (size_t)jb * 256));
  pw.rowbuf = take(sizeof(cplx<R>) * ((size_t)jb * 256 * jb + (size_t)jb * jb));
  pw.gmax = gmax;
  int rc;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(info), 2);
  if ((rc = check_launch("zero_u32_kernel"))) return rc;
  Lookahead* la = nullptr;
  if ((rc = lookahead_for(s, 2 * (size_t)nblock + 2, &la))) return rc;
  hipStream_t side = la->side;
  // block 0 on the caller's stream
  if ((rc = factor_block<R>(A, lda, n, 0, std::min(nb, n), jp, ipiv, info, pw, pivot, s, sym))) return rc;
  for (int p = 0; p < nblock; ++p) {
    const int k0 = p * nb, b = std::min(nb, n - k0);
    const int c0 = k0 + b;
    if (p > 0) MOM_TRY(hipStreamWaitEvent(s, la->ev[2 * p], 0));  // block p done on the side stream
    if (pivot && (rc = laswp_block<R>(A, lda, n, 0, k0, k0, b, ipiv, s))) return rc;
    if (c0 >= n) break;
    if (pivot && (rc = laswp_block<R>(A, lda, n, c0, n, k0, b, ipiv, s))) return rc;
    const int b1 = std::min(nb, n - c0);
    // next block's columns first, then hand the block to the side stream
    if ((rc = update_block<R>(A, lda, n, k0, b, c0, c0, c0 + b1, s, sym))) return rc;
    MOM_TRY(hipEventRecord(la->ev[2 * p + 1], s));
    MOM_TRY(hipStreamWaitEvent(side, la->ev[2 * p + 1], 0));
    if ((rc = factor_block<R>(A, lda, n, c0, b1, jp, ipiv, info, pw, pivot, side, sym))) return rc;
    MOM_TRY(hipEventRecord(la->ev[2 * (p + 1)], side));
    // the rest of the trailing matrix
    if ((rc = update_block<R>(A, lda, n, k0, b, c0, c0 + b1, n, s, sym))) return rc;
  }
  return 0;
}

// Inverses of the 64 x 64 diagonal blocks of the factors: blockIdx.y = 0 the
// unit lower L_kk, 1 the upper U_kk (a partial last block is padded with the
// identity).  Lane j solves T x = e_j by substitution in FP64 (T and the
// columns in LDS) and the result is rounded to R: Dinv[(y * nblk + k) * 4096
// + i * 64 + j] = (T_kk^-1)[i][j].
template <typename R>
__global__ __launch_bounds__(64) void trtri_blocks_kernel(int n, const cplx<R>* __restrict__ A, int64_t lda,
                                                          cplx<R>* Dinv) {
  constexpr int BS = kTrsvBS;
  extern __shared__ unsigned char smem_raw[];
  z64* T = reinterpret_cast<z64*>(smem_raw);  // [BS][BS]
  z64* Xc = T + BS * BS;                      // [BS][BS], column j in lane j
  const int blk = blockIdx.x, upper = blockIdx.y, j = threadIdx.x;
  const int nblk = gridDim.x;
  const int k0 = blk * BS, b = min(BS, n - k0);
  for (int e = j; e < BS * BS; e += BS) {
    const int i = e / BS, k = e % BS;
    const bool in = upper ? i <= k : i > k;
    z64 v = {(double)(i == k), 0.0};
    if (in && i < b && k < b) {
      const cplx<R> a = A[(int64_t)(k0 + i) * lda + k0 + k];
      v = {(double)a.re, (double)a.im};
    } else if (!in && i != k) {
      v = {0.0, 0.0};
    }
    T[e] = v;
  }
  __syncthreads();
  if (!upper) {
    for (int i = 0; i < BS; ++i) {
      z64 x = {(double)(i == j), 0.0};
      for (int k = j; k < i; ++k) x -= T[i * BS + k] * Xc[k * BS + j];
      Xc[i * BS + j] = i < j ? z64{0.0, 0.0} : x;
    }
  } else {
    for (int i = BS - 1; i >= 0; --i) {
      z64 x = {(double)(i == j), 0.0};
      for (int k = i + 1; k <= j; ++k) x -= T[i * BS + k] * Xc[k * BS + j];
      Xc[i * BS + j] = i > j ? z64{0.0, 0.0} : cdiv(x, T[i * BS + i]);
    }
  }
  cplx<R>* out = Dinv + ((int64_t)upper * nblk + blk) * BS * BS;
  for (int i = 0; i < BS; ++i) {
    const z64 x = Xc[i * BS + j];
    out[i * BS + j] = {(R)x.re, (R)x.im};
  }
}

// Inverse of a 128-row diagonal block of L or U from the inverses of its two
// 64-row blocks (trtri_blocks_kernel): [[A, 0], [C, D]]^-1 =
// [[A^-1, 0], [-D^-1 C A^-1, D^-1]] and [[A, B], [0, D]]^-1 =
// [[A^-1, -A^-1 B D^-1], [0, D^-1]].  blockIdx.x: the 128-block, blockIdx.y:
// 0 lower, 1 upper.  A second 64-block past n is the identity (as the padding
// of the 64-block inverses).  The two 64 x 64 products run through LDS.
constexpr int kTrsvBS2 = 128;

template <typename R>
__global__ __launch_bounds__(256) void compose_inverse_kernel(int n, const cplx<R>* __restrict__ A, int64_t lda,
                                                              const cplx<R>* __restrict__ D64, int nblk64,
                                                              cplx<R>* __restrict__ D128) {
  constexpr int H = kTrsvBS;  // 64
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* S1 = reinterpret_cast<cplx<R>*>(smem_raw);  // [H][H]
  cplx<R>* S2 = S1 + H * H;                             // [H][H]
  const int K = blockIdx.x, upper = blockIdx.y, tid = threadIdx.x;
  const int nblk128 = gridDim.x;
  const int a = 2 * K, b = 2 * K + 1;
  const bool has_b = b < nblk64;
  const cplx<R>* Ia = D64 + ((int64_t)upper * nblk64 + a) * H * H;
  const cplx<R>* Ib = has_b ? D64 + ((int64_t)upper * nblk64 + b) * H * H : nullptr;
  auto inv_b = [&](int i, int k) { return has_b ? Ib[i * H + k] : cplx<R>{(R)(i == k), 0}; };
  // off-diagonal block: lower C = L[rows b, cols a]; upper B = U[rows a, cols b]
  const int r0 = (upper ? a : b) * H, c0 = (upper ? b : a) * H;
  for (int e = tid; e < H * H; e += 256) {
    const int i = e / H, k = e % H;
    S1[e] = (r0 + i < n && c0 + k < n) ? A[(int64_t)(r0 + i) * lda + c0 + k] : cplx<R>{0, 0};
    S2[e] = upper ? inv_b(i, k) : Ia[e];
  }
  __syncthreads();
  // thread: row i, 16 columns j0 .. j0 + 15 of a 64 x 64 product
  const int i = tid / 4, j0 = (tid % 4) * 16;
  z64 acc[16];
  auto product = [&]() {  // acc = S1 (left) x S2 (right) in FP64
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = {0, 0};
    for (int k = 0; k < H; ++k) {
      const cplx<R> x = S1[i * H + k];
      const z64 xd = {(double)x.re, (double)x.im};
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const cplx<R> y = S2[k * H + j0 + e];
        acc[e] += xd * z64{(double)y.re, (double)y.im};
      }
    }
  };
  product();  // T = C A^-1 (lower) or B D^-1 (upper)
  __syncthreads();
#pragma unroll
  for (int e = 0; e < 16; ++e) S2[i * H + j0 + e] = {(R)acc[e].re, (R)acc[e].im};  // T into S2 (right operand)
  for (int e = tid; e < H * H; e += 256) S1[e] = upper ? Ia[e] : inv_b(e / H, e % H);  // left: A^-1 / D^-1
  __syncthreads();
  product();  // D^-1 T (lower) or A^-1 T (upper)
  cplx<R>* out = D128 + ((int64_t)upper * nblk128 + K) * kTrsvBS2 * kTrsvBS2;
  const int mr = upper ? 0 : H, mc = upper ? H : 0;  // quadrant of the product
#pragma unroll
  for (int e = 0; e < 16; ++e)
    out[(int64_t)(mr + i) * kTrsvBS2 + mc + j0 + e] = {(R)(-acc[e].re), (R)(-acc[e].im)};
  for (int e = tid; e < H * H; e += 256) {
    const int ii = e / H, k = e % H;
    out[(int64_t)ii * kTrsvBS2 + k] = Ia[e];                                  // top-left: A^-1
    out[(int64_t)(H + ii) * kTrsvBS2 + H + k] = inv_b(ii, k);                // bottom-right: D^-1
    out[(int64_t)(upper ? H + ii : ii) * kTrsvBS2 + (upper ? k : H + k)] = {0, 0};  // the zero quadrant
  }
}

// M_k = Dinv_k C_k for the chain form of the inverted-block solves, with C_k
// the coupling of block k to the block solved just before it: lower
// (blockIdx.y = 0) C_k = L[rows of block k, columns of block k - 1], upper
// (1) C_k = U[rows of block k, columns of block k + 1]; zero for the first
// block of each solve.  Formed in FP64 from the factors and inverses in
// precision R, rounded to R; rows and columns beyond n are zero.
template <typename R>
__global__ __launch_bounds__(256) void chain_matrix_kernel(int n, const cplx<R>* __restrict__ A, int64_t lda,
                                                           const cplx<R>* __restrict__ D, int nblk,
                                                           cplx<R>* __restrict__ M) {
  constexpr int H = kTrsvBS;
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* SD = reinterpret_cast<cplx<R>*>(smem_raw);  // Dinv_k [H][H]
  cplx<R>* SC = SD + H * H;                             // C_k [H][H]
  const int k = blockIdx.x, upper = blockIdx.y, tid = threadIdx.x;
  const int prev = upper ? k + 1 : k - 1;
  cplx<R>* out = M + ((int64_t)upper * nblk + k) * H * H;
  if (prev < 0 || prev >= nblk) {
    for (int e = tid; e < H * H; e += 256) out[e] = {0, 0};
    return;
  }
  const cplx<R>* Dk = D + ((int64_t)upper * nblk + k) * H * H;
  const int r0 = k * H, c0 = prev * H;
  for (int e = tid; e < H * H; e += 256) {
    const int i = e / H, j = e % H;
    SD[e] = Dk[e];
    SC[e] = (r0 + i < n && c0 + j < n) ? A[(int64_t)(r0 + i) * lda + c0 + j] : cplx<R>{0, 0};
  }
  __syncthreads();
  const int i = tid / 4, j0 = (tid % 4) * 16;  // row i, columns j0 .. j0 + 15
  z64 acc[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) acc[e] = {0, 0};
  for (int kk = 0; kk < H; ++kk) {
    const cplx<R> x = SD[i * H + kk];
    const z64 xd = {(double)x.re, (double)x.im};
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const cplx<R> y = SC[kk * H + j0 + e];
      acc[e] += xd * z64{(double)y.re, (double)y.im};
    }
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) out[i * H + j0 + e] = {(R)acc[e].re, (R)acc[e].im};
}

// Workspace of getrs: the block tickets of the two solves (words 0 and 2),
// the failure flag (word 4), then the lower solve's output and the
// precision-R copy of the solution (n x 4 complex each).
template <typename R>
static size_t getrs_work_bytes(int n) {
  return 256 + 2 * sizeof(cplx<R>) * kMaxRhs * (size_t)std::max(n, 1);
}

// getrs_inv adds the inverted 64-row diagonal blocks of L and U after that,
// and in complex64 the inverted 128-row blocks composed from them.
template <typename R>
static constexpr bool inv128() {
  return sizeof(R) == sizeof(float);
}
// After those, the chain matrices M_k of the lower and upper solves (64 x 64
// per block, chain_matrix_kernel).
template <typename R>
static size_t getrs_inv_chain_offset(int n) {
  const size_t nblk = (size_t)(std::max(n, 1) + kTrsvBS - 1) / kTrsvBS;
  const size_t nblk2 = (size_t)(std::max(n, 1) + kTrsvBS2 - 1) / kTrsvBS2;
  size_t b = getrs_work_bytes<R>(n) + sizeof(cplx<R>) * 2 * nblk * kTrsvBS * kTrsvBS;
  if (inv128<R>()) b += sizeof(cplx<R>) * 2 * nblk2 * kTrsvBS2 * kTrsvBS2;
  return b;
}
template <typename R>
static size_t getrs_inv_work_bytes(int n) {
  const size_t nblk = (size_t)(std::max(n, 1) + kTrsvBS - 1) / kTrsvBS;
  return getrs_inv_chain_offset<R>(n) + sizeof(cplx<R>) * 2 * nblk * kTrsvBS * kTrsvBS;
}
template <typename R>
static cplx<R>* chain_matrices(void* work, int n) {
  return reinterpret_cast<cplx<R>*>(reinterpret_cast<unsigned char*>(work) + getrs_inv_chain_offset<R>(n));
}

template <typename R>
static int trtri_blocks(int n, const cplx<R>* LU, int64_t lda, void* work, hipStream_t s) {
  if (n <= 0) return 0;
  const int nblk = (n + kTrsvBS - 1) / kTrsvBS;
  cplx<R>* D = reinterpret_cast<cplx<R>*>(reinterpret_cast<unsigned char*>(work) + getrs_work_bytes<R>(n));
  const size_t lds = 2 * sizeof(z64) * kTrsvBS * kTrsvBS;
  if (int rc0 = reserve_lds((const void*)trtri_blocks_kernel<R>, lds, "trtri_blocks_kernel")) return rc0;
  hipLaunchKernelGGL(trtri_blocks_kernel<R>, dim3(nblk, 2), dim3(kTrsvBS), lds, s, n, LU, lda, D);
  int rc = check_launch("trtri_blocks_kernel");
  if (rc) return rc;
  const size_t ldsm = 2 * sizeof(cplx<R>) * kTrsvBS * kTrsvBS;
  if ((rc = reserve_lds((const void*)chain_matrix_kernel<R>, ldsm, "chain_matrix_kernel"))) return rc;
  hipLaunchKernelGGL(chain_matrix_kernel<R>, dim3(nblk, 2), dim3(256), ldsm, s, n, LU, lda, D, nblk,
                     chain_matrices<R>(work, n));
  if ((rc = check_launch("chain_matrix_kernel")) || !inv128<R>()) return rc;
  const int nblk2 = (n + kTrsvBS2 - 1) / kTrsvBS2;
  cplx<R>* D2 = D + (int64_t)2 * nblk * kTrsvBS * kTrsvBS;
  const size_t lds2 = 2 * sizeof(cplx<R>) * kTrsvBS * kTrsvBS;
  if ((rc = reserve_lds((const void*)compose_inverse_kernel<R>, lds2, "compose_inverse_kernel"))) return rc;
  hipLaunchKernelGGL(compose_inverse_kernel<R>, dim3(nblk2, 2), dim3(256), lds2, s, n, LU, lda, D, nblk, D2);
  return check_launch("compose_inverse_kernel");
}

// X = (LU)^-1 P B.  B in precision TB; the solution goes to X (precision R)
// or, when X64 is given, to X64 in complex128 (X is then the workspace's
// copy).
// The lower solve P B (in X) -> Y, clearing X for the upper solve, then the
// upper solve Y -> X (and X64), with blocks of BSX rows.
template <typename R, bool INV, int BSX>
static int trsv_pair(int n, int nrhs, const cplx<R>* LU, int64_t lda, cplx<R>* X, cplx<R>* Y, const cplx<R>* DL,
                     const cplx<R>* DU, const cplx<R>* ML, const cplx<R>* MU, z64* X64, unsigned* flags,
                     unsigned* error, hipStream_t s) {
  const int nblkx = (n + BSX - 1) / BSX;
  const dim3 block(BSX * BSX / 16);
  int rc;
#define MOM_TRSV(NR)                                                                                                  \
  case NR:                                                                                                            \
    hipLaunchKernelGGL((trsv_kernel<R, true, NR, INV, BSX>), dim3(nblkx), block, 0, s, n, LU, lda, X, Y, DL, ML, X,  \
                       (z64*)nullptr, flags, error, g_spin_limit);                                                    \
    if ((rc = check_launch("trsv_kernel"))) return rc;                                                                \
    hipLaunchKernelGGL((trsv_kernel<R, false, NR, INV, BSX>), dim3(nblkx), block, 0, s, n, LU, lda, Y, X, DU, MU,    \
                       (cplx<R>*)nullptr, X64, flags + 2, error, g_spin_limit);                                       \
    break;
  switch (nrhs) {
    MOM_TRSV(1)
    MOM_TRSV(2)
    MOM_TRSV(3)
    MOM_TRSV(4)
  }
  if constexpr (BSX == kTrsvBS) {  // 5 to 8 columns on 64-row blocks only
    switch (nrhs) {
      MOM_TRSV(5)
      MOM_TRSV(6)
      MOM_TRSV(7)
      MOM_TRSV(8)
    }
  }
#undef MOM_TRSV
  return check_launch("trsv_kernel");
}

// X = (LU)^-1 P B.  B in precision TB; the solution goes to X (precision R)
// or, when X64 is given, to X64 in complex128 (X is then the workspace's
// copy).
// The failure word is word 4 of the workspace unless `error` points
// elsewhere.
template <typename R, bool INV, typename TB>
static int getrs(int n, int nrhs, const cplx<R>* LU, int64_t lda, const int32_t* perm, const cplx<TB>* B, cplx<R>* X,
                 z64* X64, unsigned* flags, hipStream_t s, unsigned* error = nullptr) {
  if (n <= 0 || nrhs <= 0) return 0;
  if (!error) error = flags + 4;
  if (nrhs > kMaxRhs) return fail("mom_getrs: at most 8 right-hand sides");
  const int nblk = (n + kTrsvBS - 1) / kTrsvBS;
  cplx<R>* Y = reinterpret_cast<cplx<R>*>(reinterpret_cast<unsigned char*>(flags) + 256);
  if (X64) X = Y + kMaxRhs * (int64_t)std::max(n, 1);
  const int64_t nx = (int64_t)n * nrhs;
  hipLaunchKernelGGL((getrs_prep_kernel<R, TB>), dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, s, n, nrhs, perm,
                     B, X, Y, flags);
  int rc = check_launch("getrs_prep_kernel");
  if (rc) return rc;
  const cplx<R>* DL = INV ? reinterpret_cast<const cplx<R>*>(reinterpret_cast<const unsigned char*>(flags) +
                                                             getrs_work_bytes<R>(n))
                          : nullptr;
  const cplx<R>* DU = INV ? DL + (int64_t)nblk * kTrsvBS * kTrsvBS : nullptr;
  // complex64 with inverted blocks and g_solve_rows = 128: the composed
  // 128-row inverses (half the chain's blocks)
  if constexpr (INV && inv128<R>()) {
    if (g_solve_rows == kTrsvBS2 && nrhs <= 4) {  // (more columns would spill at 1024 threads)
      const int nblk2 = (n + kTrsvBS2 - 1) / kTrsvBS2;
      const cplx<R>* D2L = DU + (int64_t)nblk * kTrsvBS * kTrsvBS;
      const cplx<R>* D2U = D2L + (int64_t)nblk2 * kTrsvBS2 * kTrsvBS2;
      return trsv_pair<R, INV, kTrsvBS2>(n, nrhs, LU, lda, X, Y, D2L, D2U, nullptr, nullptr, X64, flags, error, s);
    }
  }
  const cplx<R>* ML = INV && g_chain ? chain_matrices<R>(flags, n) : nullptr;
  const cplx<R>* MU = ML ? ML + (int64_t)nblk * kTrsvBS * kTrsvBS : nullptr;
  return trsv_pair<R, INV, kTrsvBS>(n, nrhs, LU, lda, X, Y, DL, DU, ML, MU, X64, flags, error, s);
}

// ------------------------------------------------- super-block triangular solves
// The preconditioner solve of the mixed path as a chain of super-blocks of
// sb = 512 rows instead of 64, so that the chain has 8x fewer hops.  After a
// factorisation (once), with D_k the inverse of the k-th diagonal super-block
// of L (of U) and the couplings scaled by it,
//     T_kj = -D_k L_kj (j < k),   T_kj = -D_k U_kj (j > k),
// the two solves read
//     x_k = D_k b_k + sum_{j before k} T_kj x_j
// where "before" is the solve order (increasing k for L, decreasing for U).
// Each row of x_k is an independent dot product: a workgroup owns a few rows
// of one super-block, forms D_k b_k and the terms of the super-blocks solved
// earlier as they appear (streaming its rows of T from HBM), and only the
// last term, that of the super-block solved just before, sits on the chain's
// critical path (its tile of T is loaded before the wait).  The x values are
// their own ready flags (sentinel, as in the 64-row chain), and workgroups
// take their super-block by ticket, so that they only wait for workgroups
// already running.
//
// Workspace (mom_sbsolve_work_bytes): 256 status bytes (tickets in words 0
// and 2, the sticky failure word 4), Y (the lower solve's output), X2 (the
// upper solve's), X1 (P b), each n x 4 complex in precision R; the 64-row
// inverses of L's and U's diagonal blocks (scratch of the set-up); D for L
// and for U (nsb x sb x sb); T (n x ldt, ldt = nsb sb).
constexpr int kSbMax = 512;

struct SbLayout {
  size_t y, x2, x1, d64, dl, du, t, total;
  int nsb, ldt, nblk64;
};

template <typename R>
static SbLayout sb_layout(int n, int sb) {
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  SbLayout L;
  n = std::max(n, 1);
  L.nsb = (n + sb - 1) / sb;
  L.ldt = L.nsb * sb;
  L.nblk64 = (n + kTrsvBS - 1) / kTrsvBS;
  const size_t vec = al(sizeof(cplx<R>) * 4 * (size_t)n);
  size_t off = 256;
  L.y = off;
  off += vec;
  L.x2 = off;
  off += vec;
  L.x1 = off;
  off += vec;
  L.d64 = off;
  off += al(sizeof(cplx<R>) * 2 * (size_t)L.nblk64 * kTrsvBS * kTrsvBS);
  L.dl = off;
  off += al(sizeof(cplx<R>) * (size_t)L.nsb * sb * sb);
  L.du = off;
  off += al(sizeof(cplx<R>) * (size_t)L.nsb * sb * sb);
  L.t = off;
  off += al(sizeof(cplx<R>) * (size_t)n * L.ldt);
  L.total = off;
  return L;
}

// Inverse of the k-th diagonal super-block of L (blockIdx.z = 0, unit lower)
// or U (1), 64 columns per workgroup (blockIdx.y), by block substitution over
// its 64-row sub-blocks with the 64-row inverses D64 (trtri_blocks_kernel):
// lower X_i = D64_i (E_i - sum_{c <= m < i} L_im X_m) for i = c .. (the
// identity column block E), upper the mirror; products in FP64 through LDS,
// results rounded to R.  Sub-blocks past n are the identity (rows and
// columns of the factors there read as zero), so a partial last super-block
// gets a finite padding that the solve multiplies by zero.
template <typename R>
__global__ __launch_bounds__(256) void sb_inverse_kernel(int n, const cplx<R>* __restrict__ A, int64_t lda,
                                                         const cplx<R>* __restrict__ D64, int nblk64, int sb,
                                                         cplx<R>* DL, cplx<R>* DU) {
  constexpr int H = kTrsvBS;
  extern __shared__ unsigned char smem_raw[];
  z64* SX = reinterpret_cast<z64*>(smem_raw);  // [H][H] right operand (FP64)
  cplx<R>* SA = reinterpret_cast<cplx<R>*>(SX + H * H);  // [H][H] left operand
  const int k = blockIdx.x, c = blockIdx.y, upper = blockIdx.z, tid = threadIdx.x;
  const int nsub = sb / H, base = k * sb;
  cplx<R>* out = (upper ? DU : DL) + (int64_t)k * sb * sb;
  const int ir = tid / 4, j0 = (tid % 4) * 16;  // thread: row ir, columns j0 .. j0 + 15 of a 64 x 64 block
  auto at = [&](int i, int r, int col) { return out + (int64_t)(i * H + r) * sb + c * H + col; };
  for (int i = 0; i < nsub; ++i)  // the sub-blocks of this column block that are zero
    if (upper ? i > c : i < c)
      for (int e = tid; e < H * H; e += 256) *at(i, e / H, e % H) = {0, 0};
  const int step = upper ? -1 : 1;
  for (int i = c; upper ? i >= 0 : i < nsub; i += step) {
    z64 acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = {(double)(i == c && ir == j0 + e), 0.0};
    for (int mm = c; mm != i; mm += step) {
      __syncthreads();
      for (int e = tid; e < H * H; e += 256) {
        const int r = e / H, q = e % H;
        const int gr = base + i * H + r, gc = base + mm * H + q;
        SA[e] = (gr < n && gc < n) ? A[(int64_t)gr * lda + gc] : cplx<R>{0, 0};
        const cplx<R> x = load_coherent_c(at(mm, r, q));  // written by this workgroup (stores drained below)
        SX[e] = {(double)x.re, (double)x.im};
      }
      __syncthreads();
      for (int kk = 0; kk < H; ++kk) {
        const cplx<R> a = SA[ir * H + kk];
        const z64 ad = {(double)a.re, (double)a.im};
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[e] -= ad * SX[kk * H + j0 + e];
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) SX[ir * H + j0 + e] = acc[e];
    const int g64 = base / H + i;
    for (int e = tid; e < H * H; e += 256)
      SA[e] = g64 < nblk64 ? D64[((int64_t)upper * nblk64 + g64) * H * H + e] : cplx<R>{(R)(e / H == e % H), 0};
    __syncthreads();
    z64 x[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) x[e] = {0, 0};
    for (int kk = 0; kk < H; ++kk) {
      const cplx<R> a = SA[ir * H + kk];
      const z64 ad = {(double)a.re, (double)a.im};
#pragma unroll
      for (int e = 0; e < 16; ++e) x[e] += ad * SX[kk * H + j0 + e];
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) *at(i, ir, j0 + e) = {(R)x[e].re, (R)x[e].im};
    __builtin_amdgcn_s_waitcnt(0);  // the stores have reached L2 before another thread reads them
  }
}

// Set-up of a solve: tickets cleared (words 0 and 2; the failure word stays
// for the caller), X1 = P B rounded to R, Y and X2 (the two solves' outputs)
// filled with the sentinel.
template <typename R, typename TB>
__global__ void sb_prep_kernel(int n, int nrhs, const int32_t* perm, const cplx<TB>* B, cplx<R>* X1, cplx<R>* Y,
                               cplx<R>* X2, unsigned* flags) {
  gmres_step_priority();
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q == 0 || q == 2) flags[q] = 0u;
  if (q >= n * nrhs) return;
  const int i = q / nrhs, c = q % nrhs;
  const cplx<TB> b = B[(int64_t)perm[i] * nrhs + c];
  X1[q] = {(R)b.re, (R)b.im};
  const cplx<R> sen = {Sentinel<R>::value(), Sentinel<R>::value()};
  Y[q] = sen;
  X2[q] = sen;
}

// DPP move of a 32-bit value (lanes without a source, or outside row_mask, read 0).
template <int CTRL, int ROWS>
__device__ inline int dpp32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, true);
}
template <int CTRL, int ROWS>
__device__ inline float dpp_move(float v) {
  return __int_as_float(dpp32<CTRL, ROWS>(__float_as_int(v)));
}
template <int CTRL, int ROWS>
__device__ inline double dpp_move(double v) {
  return __hiloint2double(dpp32<CTRL, ROWS>(__double2hiint(v)), dpp32<CTRL, ROWS>(__double2loint(v)));
}
// Sum over the 64 lanes of a wavefront, returned in every lane: shifts by 1,
// 2, 4 and 8 within each row of 16 lanes (lane 15 of a row then holds the
// row's sum), row 0 into row 1 and row 2 into row 3 (row_bcast:15), rows 0-1
// into rows 2-3 (row_bcast:31), and lane 63 read out through the scalar unit.
template <typename T>
__device__ inline T wave_total(T v) {
  v += dpp_move<0x111, 0xf>(v);
  v += dpp_move<0x112, 0xf>(v);
  v += dpp_move<0x114, 0xf>(v);
  v += dpp_move<0x118, 0xf>(v);
  v += dpp_move<0x142, 0xa>(v);
  v += dpp_move<0x143, 0xc>(v);
  return lane_value(v, 63);
}

// One triangular solve (lower: UPPER false) over super-blocks of 512 rows
// with NR right-hand sides.  Workgroup: 4 wavefronts, RPW rows each (2 in
// complex64, 1 in complex128); lane l covers the 8 columns 128 q + 2 l + h
// (q < 4, h < 2) of a super-block, so that every load instruction of a
// wavefront reads 64 consecutive pieces of one row.  Per-lane partial sums
// are reduced across the wavefront once, after the last term.
// rows per wavefront of the super-block chain: one tile of T (RPW x 8
// complex per lane) in flight besides the next keeps it at two workgroups
// per CU without spills
template <typename R>
constexpr int kSbRowsPerWave() { return sizeof(R) == sizeof(float) ? 2 : 1; }

template <typename R, int NR, bool UPPER>
__global__ __launch_bounds__(256, 2) void sb_chain_kernel(int n, int nsb, const cplx<R>* __restrict__ T, int64_t ldt,
                                                       const cplx<R>* __restrict__ D, const cplx<R>* __restrict__ in,
                                                       cplx<R>* out, z64* __restrict__ out64, unsigned* ticket,
                                                       unsigned* error, unsigned spin_limit) {
  gmres_step_priority();
  constexpr int SB = kSbMax, CPL = 8;
  constexpr int RPW = kSbRowsPerWave<R>();
  constexpr int ROWS = 4 * RPW, G = SB / ROWS;
  // x of up to NB earlier super-blocks at once in LDS (32 KB): their values
  // are fetched with one round of device-coherent loads, so a workgroup that
  // is behind the chain catches up one memory round trip per NB super-blocks
  constexpr int XB = 32768 / (SB * NR * (int)sizeof(cplx<R>));
  constexpr int NB = XB >= 4 ? 4 : (XB >= 1 ? XB : 1);
  __shared__ cplx<R> xs[NB][SB][NR];
  __shared__ int s_ticket, s_abort;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) {
    const unsigned t = atomicAdd(ticket, 1u);
    // a ticket past the grid means that two solves share this workspace at
    // once: report it instead of indexing outside the factors
    if (t >= (unsigned)(nsb * G)) {
      atomicOr(error, 1u);
      s_ticket = -1;
    } else {
      s_ticket = (int)t;
    }
    s_abort = 0;
  }
  __syncthreads();
  if (s_ticket < 0) return;
  const int o = s_ticket / G, g = s_ticket % G;
  const int k = UPPER ? nsb - 1 - o : o;
  const int lrow0 = g * ROWS + wave * RPW;  // first row of this wavefront within the super-block
  const int row0 = k * SB + lrow0;
  auto colof = [&](int e) { return 128 * (e >> 1) + 2 * lane + (e & 1); };
  cplx<R> acc[RPW][NR];
#pragma unroll
  for (int r = 0; r < RPW; ++r)
#pragma unroll
    for (int cc = 0; cc < NR; ++cc) acc[r][cc] = {0, 0};
  // (columns past n read memory inside the allocation and are then zeroed;
  // T's tiles below are loaded under a row / column guard, measured faster
  // than clamped unconditional loads: 0.42 against 0.58 ms for the two
  // complex64 solves at n = 12 428)
  const cplx<R> zero = {0, 0};
  {  // D_k b_k (b: the input, complete before this launch; D is padded to whole super-blocks)
    cplx<R> bv[CPL][NR];
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      const int col = k * SB + colof(e);
#pragma unroll
      for (int cc = 0; cc < NR; ++cc) {
        const cplx<R> b = in[(int64_t)min(col, n - 1) * NR + cc];
        bv[e][cc] = col < n ? b : zero;
      }
    }
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const cplx<R>* drow = D + ((int64_t)k * SB + lrow0 + r) * SB;
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        const cplx<R> d = drow[colof(e)];
#pragma unroll
        for (int cc = 0; cc < NR; ++cc) acc[r][cc] += d * bv[e][cc];
      }
    }
  }
  const int ndep = UPPER ? nsb - 1 - k : k;
  auto dep = [&](int q) { return UPPER ? nsb - 1 - q : q; };  // super-blocks in solve order
  // this wavefront's rows of T_kj
  auto load_tile = [&](int j, cplx<R> (&tv)[RPW][CPL]) {
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const cplx<R>* trow = T + (int64_t)(row0 + r) * ldt + (int64_t)j * SB;
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        const int col = j * SB + colof(e);
        tv[r][e] = (row0 + r < n && col < n) ? trow[colof(e)] : cplx<R>{0, 0};
      }
    }
  };
  auto fma_tile = [&](const cplx<R> (&tv)[RPW][CPL], int b) {
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
      for (int e = 0; e < CPL; ++e)
#pragma unroll
        for (int cc = 0; cc < NR; ++cc) acc[r][cc] += tv[r][e] * xs[b][colof(e)][cc];
  };
  // x of super-blocks dep(q0) .. dep(q0 + nb - 1) into xs[0 .. nb): every
  // thread issues its (up to VPT) device-coherent loads together, then
  // reissues together, round after round, the ones that still read the
  // sentinel (bounded), so that a wait costs one memory round trip per round
  // rather than one per value; false when a wait ran out.  (Completion
  // counters per super-block, read before one pass over the values, were
  // slower: 0.64 against 0.42 ms for the two complex64 solves at n = 12 428.)
  constexpr int VPT = NB * SB * NR / 256;  // values per thread (<= 16: the batch is <= 32 KB)
  static_assert(VPT * 256 == NB * SB * NR && VPT <= 32, "batch layout");
  auto fetch_x = [&](int q0, int nb) {
    const int total = nb * SB * NR;
    auto addr = [&](int idx, int& b, int& c, int& cc) {
      b = idx / (SB * NR);
      const int rem = idx % (SB * NR);
      c = rem / NR;
      cc = rem % NR;
      return dep(q0 + b) * SB + c;  // column of x
    };
    uint32_t pending = 0;
    cplx<R> v[VPT];
#pragma unroll
    for (int t = 0; t < VPT; ++t) {
      const int idx = tid + 256 * t;
      int b, c, cc;
      const int col = idx < total ? addr(idx, b, c, cc) : n;
      v[t] = col < n ? load_coherent_c(out + (int64_t)col * NR + cc) : cplx<R>{0, 0};
    }
#pragma unroll
    for (int t = 0; t < VPT; ++t) {
      const int idx = tid + 256 * t;
      if (idx >= total) continue;
      int b, c, cc;
      addr(idx, b, c, cc);
      xs[b][c][cc] = v[t];
      if (is_sentinel(v[t])) pending |= 1u << t;
    }
    unsigned spins = 0;
    while (pending) {
      poll_pause();
      if (++spins > spin_limit) {
        s_abort = 1;
        break;
      }
#pragma unroll
      for (int t = 0; t < VPT; ++t) {
        int b, c, cc;
        const int col = addr(tid + 256 * t, b, c, cc);
        if ((pending >> t) & 1u) v[t] = load_coherent_c(out + (int64_t)col * NR + cc);
      }
#pragma unroll
      for (int t = 0; t < VPT; ++t) {
        if (!((pending >> t) & 1u) || is_sentinel(v[t])) continue;
        int b, c, cc;
        addr(tid + 256 * t, b, c, cc);
        xs[b][c][cc] = v[t];
        pending &= ~(1u << t);
      }
    }
    __syncthreads();
    return s_abort == 0;
  };
  // the super-blocks solved before the previous one, NB at a time
  cplx<R> tcur[RPW][CPL];
  for (int q0 = 0; q0 < ndep - 1; q0 += NB) {
    const int nb = min(NB, ndep - 1 - q0);
    load_tile(dep(q0), tcur);
    if (!fetch_x(q0, nb)) {
      if (tid == 0) atomicOr(error, 1u);
      return;
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (b >= nb) break;
      cplx<R> tnext[RPW][CPL];
      if (b + 1 < nb) load_tile(dep(q0 + b + 1), tnext);
      fma_tile(tcur, b);
      if (b + 1 < nb) {
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
          for (int e = 0; e < CPL; ++e) tcur[r][e] = tnext[r][e];
      }
    }
    __syncthreads();  // every wavefront has read xs before the next batch overwrites it
  }
  // the previous super-block: the chain's critical path (its tile loaded before the wait)
  if (ndep > 0) {
    load_tile(dep(ndep - 1), tcur);
    if (!fetch_x(ndep - 1, 1)) {
      if (tid == 0) atomicOr(error, 1u);
      return;
    }
    fma_tile(tcur, 0);
  }
#pragma unroll
  for (int r = 0; r < RPW; ++r)
#pragma unroll
    for (int cc = 0; cc < NR; ++cc) {
      acc[r][cc].re = wave_total(acc[r][cc].re);
      acc[r][cc].im = wave_total(acc[r][cc].im);
    }
#pragma unroll
  for (int r = 0; r < RPW; ++r)
#pragma unroll
    for (int cc = 0; cc < NR; ++cc)
      if (lane == r * NR + cc && row0 + r < n) {
        store_coherent_c(out + (int64_t)(row0 + r) * NR + cc, acc[r][cc]);
        if (out64) out64[(int64_t)(row0 + r) * NR + cc] = {(double)acc[r][cc].re, (double)acc[r][cc].im};
      }
}

template <typename R>
static int sbsolve_setup(int n, const cplx<R>* LU, int64_t lda, int sb, void* work, size_t work_bytes, hipStream_t s) {
  if (sb != kSbMax) return fail("mom_sbsolve_setup: super-block rows must be 512");
  if (n <= 0) return 0;
  const SbLayout L = sb_layout<R>(n, sb);
  if (work_bytes < L.total) return fail("mom_sbsolve_setup: workspace too small");
  unsigned char* w = reinterpret_cast<unsigned char*>(work);
  cplx<R>* D64 = reinterpret_cast<cplx<R>*>(w + L.d64);
  cplx<R>* DL = reinterpret_cast<cplx<R>*>(w + L.dl);
  cplx<R>* DU = reinterpret_cast<cplx<R>*>(w + L.du);
  cplx<R>* T = reinterpret_cast<cplx<R>*>(w + L.t);
  const size_t lds64 = 2 * sizeof(z64) * kTrsvBS * kTrsvBS;
  if (int rc0 = reserve_lds((const void*)trtri_blocks_kernel<R>, lds64, "trtri_blocks_kernel")) return rc0;
  if (int rc0 = reserve_lds((const void*)sb_inverse_kernel<R>, lds64, "sb_inverse_kernel")) return rc0;
  hipLaunchKernelGGL(trtri_blocks_kernel<R>, dim3(L.nblk64, 2), dim3(kTrsvBS), lds64, s, n, LU, lda, D64);
  int rc = check_launch("trtri_blocks_kernel");
  if (rc) return rc;
  const size_t ldsi = sizeof(z64) * kTrsvBS * kTrsvBS + sizeof(cplx<R>) * kTrsvBS * kTrsvBS;
  hipLaunchKernelGGL(sb_inverse_kernel<R>, dim3(L.nsb, sb / kTrsvBS, 2), dim3(256), ldsi, s, n, LU, lda,
                     (const cplx<R>*)D64, L.nblk64, sb, DL, DU);
  if ((rc = check_launch("sb_inverse_kernel"))) return rc;
  for (int k = 0; k < L.nsb; ++k) {
    const int r0 = k * sb, rows = std::min(sb, n - r0);
    if (k > 0 &&
        (rc = launch_cgemm<R, true>(rows, r0, rows, DL + (int64_t)k * sb * sb, sb, LU + (int64_t)r0 * lda, lda,
                                    T + (int64_t)r0 * L.ldt, L.ldt, s)))
      return rc;
    const int c0 = r0 + sb;
    if (c0 < n &&
        (rc = launch_cgemm<R, true>(rows, n - c0, rows, DU + (int64_t)k * sb * sb, sb, LU + (int64_t)r0 * lda + c0,
                                    lda, T + (int64_t)r0 * L.ldt + c0, L.ldt, s)))
      return rc;
  }
  return 0;
}

template <typename R, typename TB>
static int sbsolve(int n, int nrhs, int sb, const int32_t* perm, const cplx<TB>* B, z64* X64, void* work,
                   hipStream_t s) {
  if (sb != kSbMax) return fail("mom_sbsolve: super-block rows must be 512");
  if (n <= 0 || nrhs <= 0) return 0;
  if (nrhs > 4) return fail("mom_sbsolve: at most 4 right-hand sides");
  const SbLayout L = sb_layout<R>(n, sb);
  unsigned char* w = reinterpret_cast<unsigned char*>(work);
  unsigned* flags = reinterpret_cast<unsigned*>(w);
  cplx<R>* Y = reinterpret_cast<cplx<R>*>(w + L.y);
  cplx<R>* X2 = reinterpret_cast<cplx<R>*>(w + L.x2);
  cplx<R>* X1 = reinterpret_cast<cplx<R>*>(w + L.x1);
  const cplx<R>* DL = reinterpret_cast<const cplx<R>*>(w + L.dl);
  const cplx<R>* DU = reinterpret_cast<const cplx<R>*>(w + L.du);
  const cplx<R>* T = reinterpret_cast<const cplx<R>*>(w + L.t);
  const int64_t nx = (int64_t)n * nrhs;
  hipLaunchKernelGGL((sb_prep_kernel<R, TB>), dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, s, n, nrhs, perm, B,
                     X1, Y, X2, flags);
  int rc = check_launch("sb_prep_kernel");
  if (rc) return rc;
  constexpr int ROWS = 4 * kSbRowsPerWave<R>();
  const dim3 grid(L.nsb * (kSbMax / ROWS));
  // timed per chain: algorithmic bytes, one triangle of the factors and the columns in and out
  const int tkind = sizeof(R) == sizeof(double) ? kTimeChain64 : kTimeChain32;
  const double tbytes = sizeof(cplx<R>) * (0.5 * n * (n + 1.0) + 2.0 * n * nrhs);
  int ts;
#define MOM_SB(NR)                                                                                                    \
  case NR:                                                                                                            \
    ts = timing_begin(tkind, s);                                                                                      \
    hipLaunchKernelGGL((sb_chain_kernel<R, NR, false>), grid, dim3(256), 0, s, n, L.nsb, T, (int64_t)L.ldt, DL,      \
                       (const cplx<R>*)X1, Y, (z64*)nullptr, flags, flags + 4, g_spin_limit);                         \
    timing_end(tkind, ts, s, tbytes);                                                                                 \
    if ((rc = check_launch("sb_chain_kernel"))) return rc;                                                            \
    ts = timing_begin(tkind, s);                                                                                      \
    hipLaunchKernelGGL((sb_chain_kernel<R, NR, true>), grid, dim3(256), 0, s, n, L.nsb, T, (int64_t)L.ldt, DU,       \
                       (const cplx<R>*)Y, X2, X64, flags + 2, flags + 4, g_spin_limit);                               \
    timing_end(tkind, ts, s, tbytes);                                                                                 \
    break;
  switch (nrhs) {
    MOM_SB(1)
    MOM_SB(2)
    MOM_SB(3)
    MOM_SB(4)
  }
#undef MOM_SB
  return check_launch("sb_chain_kernel");
}

}  // namespace mom

using namespace mom;

extern "C" int64_t mom_getrf_work_bytes(int prec, int n, int jb) {
  return prec == 64 ? (int64_t)getrf_work_bytes<double>(n, jb) : (int64_t)getrf_work_bytes<float>(n, jb);
}

extern "C" int mom_getrf(int prec, int n, void* A, int64_t lda, int32_t* ipiv, int32_t* info, void* work,
                         int64_t work_bytes, int jb, int max_groups, void* stream) {
  if (!A || !ipiv || !info || !work) return fail("mom_getrf: null device pointer");
  if (lda < n) return fail("mom_getrf: lda < n");
  hipStream_t s = (hipStream_t)stream;
  if (prec == 64) return getrf<double>(n, (z64*)A, lda, ipiv, info, work, (size_t)work_bytes, jb, max_groups, s);
  if (prec == 32) return getrf<float>(n, (c32*)A, lda, ipiv, info, work, (size_t)work_bytes, jb, max_groups, s);
  return fail("mom_getrf: prec must be 64 (complex128) or 32 (complex64)");
}

extern "C" int mom_getrf_nopiv(int prec, int n, void* A, int64_t lda, int32_t* ipiv, int32_t* info, void* work,
                               int64_t work_bytes, int jb, void* stream) {
  if (!A || !ipiv || !info || !work) return fail("mom_getrf_nopiv: null device pointer");
  if (lda < n) return fail("mom_getrf_nopiv: lda < n");
  hipStream_t s = (hipStream_t)stream;
  if (prec == 64)
    return getrf<double>(n, (z64*)A, lda, ipiv, info, work, (size_t)work_bytes, jb, 0, s, false);
  if (prec == 32) return getrf<float>(n, (c32*)A, lda, ipiv, info, work, (size_t)work_bytes, jb, 0, s, false);
  return fail("mom_getrf_nopiv: prec must be 64 (complex128) or 32 (complex64)");
}

extern "C" int mom_sytrf_nopiv(int prec, int n, void* A, int64_t lda, int32_t* ipiv, int32_t* info, void* work,
                               int64_t work_bytes, int jb, void* stream) {
  if (!A || !ipiv || !info || !work) return fail("mom_sytrf_nopiv: null device pointer");
  if (lda < n) return fail("mom_sytrf_nopiv: lda < n");
  hipStream_t s = (hipStream_t)stream;
  if (prec == 64)
    return getrf<double>(n, (z64*)A, lda, ipiv, info, work, (size_t)work_bytes, jb, 0, s, false, true);
  if (prec == 32) return getrf<float>(n, (c32*)A, lda, ipiv, info, work, (size_t)work_bytes, jb, 0, s, false, true);
  return fail("mom_sytrf_nopiv: prec must be 64 (complex128) or 32 (complex64)");
}

extern "C" int mom_build_perm(int n, const int32_t* ipiv, int32_t* perm, void* stream) {
  if (n <= 0) return 0;
  size_t lds = n <= 16000 ? sizeof(int32_t) * (size_t)n : 0;
  hipLaunchKernelGGL(build_perm_kernel, dim3(1), dim3(256), lds, (hipStream_t)stream, n, ipiv, perm);
  return check_launch("build_perm_kernel");
}

// Workspace of mom_getrs: ticket and failure flag of each triangular solve
// and the lower solve's output.
extern "C" int64_t mom_getrs_work_bytes(int prec, int n) {
  return prec == 64 ? (int64_t)getrs_work_bytes<double>(n) : (int64_t)getrs_work_bytes<float>(n);
}

extern "C" int mom_getrs(int prec, int n, int nrhs, const void* LU, int64_t lda, const int32_t* perm, const void* B,
                         void* X, void* work, void* stream) {
  if (!LU || !perm || !B || !X || !work) return fail("mom_getrs: null device pointer");
  if (B == X) return fail("mom_getrs: B and X must not alias");
  hipStream_t s = (hipStream_t)stream;
  unsigned* flags = reinterpret_cast<unsigned*>(work);
  if (prec == 64)
    return getrs<double, false, double>(n, nrhs, (const z64*)LU, lda, perm, (const z64*)B, (z64*)X, nullptr, flags, s);
  if (prec == 32)
    return getrs<float, false, float>(n, nrhs, (const c32*)LU, lda, perm, (const c32*)B, (c32*)X, nullptr, flags, s);
  return fail("mom_getrs: prec must be 64 or 32");
}

// ---------------------------------------------------------------------------
// mom_getrs_noalloc: mom_getrs with its workspace (the two block tickets and
// the lower solve's output) in a file-scope device array instead of a
// caller buffer, for callers that cannot allocate.  Each device holds its
// own copy of `scratch` until the process ends; host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the
// symbol), but the address hipGetSymbolAddress returns for the current
// device, looked up once per device.  Every call on a device shares that
// device's array: calls that overlap on one device race.
constexpr int kNoAllocMaxN = 16384;
constexpr size_t kNoAllocBytes = 256 + 2 * sizeof(z64) * kMaxRhs * (size_t)kNoAllocMaxN;  // getrs_work_bytes<double>
static_assert(kNoAllocBytes >= 256 + 2 * sizeof(c32) * kMaxRhs * (size_t)kNoAllocMaxN, "complex64 fits too");
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static int scratch_for_current_device(void** out) {
  int dev = -1;
  MOM_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("mom_getrs_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    MOM_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail("mom_getrs_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// Workspace of mom_trtri_blocks / mom_getrs_inv: that of mom_getrs followed
// by the inverted diagonal blocks of L and U.
extern "C" int64_t mom_getrs_inv_work_bytes(int prec, int n) {
  return prec == 64 ? (int64_t)getrs_inv_work_bytes<double>(n) : (int64_t)getrs_inv_work_bytes<float>(n);
}

extern "C" int mom_getrs_noalloc(int prec, int n, int nrhs, const void* LU, int64_t lda, const int32_t* perm,
                                 const void* B, void* X, uint32_t* status, void* stream) {
  if (!LU || !perm || !B || !X || !status) return fail("mom_getrs_noalloc: null device pointer");
  if (B == X) return fail("mom_getrs_noalloc: B and X must not alias");
  if (prec != 64 && prec != 32) return fail("mom_getrs_noalloc: prec must be 64 or 32");
  if (n > kNoAllocMaxN)
    return fail("mom_getrs_noalloc: n = " + std::to_string(n) + " exceeds the static workspace's limit of " +
                std::to_string(kNoAllocMaxN) + " unknowns (use mom_getrs with a caller workspace)");
  if (nrhs > kMaxRhs) return fail("mom_getrs_noalloc: at most 8 right-hand sides");
  if (n <= 0 || nrhs <= 0) return 0;
  void* ws = nullptr;
  int rc = scratch_for_current_device(&ws);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  unsigned* flags = reinterpret_cast<unsigned*>(ws);  // tickets cleared by the solve's own set-up kernel
  unsigned* err = reinterpret_cast<unsigned*>(status);
  if (prec == 64)
    return getrs<double, false, double>(n, nrhs, (const z64*)LU, lda, perm, (const z64*)B, (z64*)X, nullptr, flags, s,
                                        err);
  return getrs<float, false, float>(n, nrhs, (const c32*)LU, lda, perm, (const c32*)B, (c32*)X, nullptr, flags, s,
                                    err);
}

extern "C" int mom_trtri_blocks(int prec, int n, const void* LU, int64_t lda, void* work, void* stream) {
  if (!LU || !work) return fail("mom_trtri_blocks: null device pointer");
  hipStream_t s = (hipStream_t)stream;
  if (prec == 64) return trtri_blocks<double>(n, (const z64*)LU, lda, work, s);
  if (prec == 32) return trtri_blocks<float>(n, (const c32*)LU, lda, work, s);
  return fail("mom_trtri_blocks: prec must be 64 or 32");
}

extern "C" int mom_getrs_inv(int prec, int n, int nrhs, const void* LU, int64_t lda, const int32_t* perm,
                             const void* B, void* X, void* work, void* stream) {
  if (!LU || !perm || !B || !X || !work) return fail("mom_getrs_inv: null device pointer");
  if (B == X) return fail("mom_getrs_inv: B and X must not alias");
  hipStream_t s = (hipStream_t)stream;
  unsigned* flags = reinterpret_cast<unsigned*>(work);
  if (prec == 64)
    return getrs<double, true, double>(n, nrhs, (const z64*)LU, lda, perm, (const z64*)B, (z64*)X, nullptr, flags, s);
  if (prec == 32)
    return getrs<float, true, float>(n, nrhs, (const c32*)LU, lda, perm, (const c32*)B, (c32*)X, nullptr, flags, s);
  return fail("mom_getrs_inv: prec must be 64 or 32");
}

extern "C" int mom_getrs_inv_z(int prec, int n, int nrhs, const void* LU, int64_t lda, const int32_t* perm,
                               const void* B, void* X, void* work, void* stream) {
  if (!LU || !perm || !B || !X || !work) return fail("mom_getrs_inv_z: null device pointer");
  if (B == X) return fail("mom_getrs_inv_z: B and X must not alias");
  hipStream_t s = (hipStream_t)stream;
  unsigned* flags = reinterpret_cast<unsigned*>(work);
  if (prec == 64)
    return getrs<double, true, double>(n, nrhs, (const z64*)LU, lda, perm, (const z64*)B, nullptr, (z64*)X, flags, s);
  if (prec == 32)
    return getrs<float, true, double>(n, nrhs, (const c32*)LU, lda, perm, (const z64*)B, nullptr, (z64*)X, flags, s);
  return fail("mom_getrs_inv_z: prec must be 64 or 32");
}

extern "C" int64_t mom_sbsolve_work_bytes(int prec, int n, int sb) {
  if (sb != kSbMax) return -1;
  return prec == 64 ? (int64_t)sb_layout<double>(n, sb).total : (int64_t)sb_layout<float>(n, sb).total;
}

extern "C" int mom_sbsolve_setup(int prec, int n, const void* LU, int64_t lda, int sb, void* work, int64_t work_bytes,
                                 void* stream) {
  if (!LU || !work) return fail("mom_sbsolve_setup: null device pointer");
  if (lda < n) return fail("mom_sbsolve_setup: lda < n");
  hipStream_t s = (hipStream_t)stream;
  if (prec == 64) return sbsolve_setup<double>(n, (const z64*)LU, lda, sb, work, (size_t)work_bytes, s);
  if (prec == 32) return sbsolve_setup<float>(n, (const c32*)LU, lda, sb, work, (size_t)work_bytes, s);
  return fail("mom_sbsolve_setup: prec must be 64 or 32");
}

extern "C" int mom_sbsolve_z(int prec, int n, int nrhs, int sb, const int32_t* perm, const void* B, void* X,
                             void* work, void* stream) {
  if (!perm || !B || !X || !work) return fail("mom_sbsolve_z: null device pointer");
  if (B == X) return fail("mom_sbsolve_z: B and X must not alias");
  hipStream_t s = (hipStream_t)stream;
  if (prec == 64) return sbsolve<double, double>(n, nrhs, sb, perm, (const z64*)B, (z64*)X, work, s);
  if (prec == 32) return sbsolve<float, double>(n, nrhs, sb, perm, (const z64*)B, (z64*)X, work, s);
  return fail("mom_sbsolve_z: prec must be 64 or 32");
}

extern "C" int mom_set_solve_rows(int rows) {
  if (rows != 64 && rows != 128) return fail("mom_set_solve_rows: rows must be 64 or 128");
  g_solve_rows = rows;
  return 0;
}

extern "C" int mom_get_solve_rows(void) { return g_solve_rows; }

extern "C" int mom_set_solve_chain(int on) {
  g_chain = on ? 1 : 0;
  return 0;
}

extern "C" int mom_get_solve_chain(void) { return g_chain; }

extern "C" int mom_set_spin_limit(unsigned polls) {
  if (polls == 0) return fail("mom_set_spin_limit: the limit must be positive");
  g_spin_limit = polls;
  return 0;
}

extern "C" int mom_set_lu_block(int prec, int nb) {
  if (prec != 64 && prec != 32) return fail("mom_set_lu_block: prec must be 64 or 32");
  if (nb < 0 || nb > 256 || nb % 8) return fail("mom_set_lu_block: block width must be 0 or a multiple of 8 up to 256");
  g_lu_nb[prec == 64 ? 0 : 1] = nb;
  return 0;
}

extern "C" int mom_get_lu_block(int prec) { return prec == 64 ? g_lu_nb[0] : g_lu_nb[1]; }

extern "C" int mom_set_gemm_tile(int prec, int bm, int bn) {
  if (prec != 64 && prec != 32) return fail("mom_set_gemm_tile: prec must be 64 or 32");
  if (!((bm == 64 || bm == 128) && (bn == 64 || bn == 128)))
    return fail("mom_set_gemm_tile: tile must be 64x64, 64x128, 128x64 or 128x128");
  g_tile[prec == 64 ? 0 : 1][0] = bm;
  g_tile[prec == 64 ? 0 : 1][1] = bn;
  return 0;
}

extern "C" int mom_get_gemm_tile(int prec, int* bm, int* bn) {
  if (prec != 64 && prec != 32) return fail("mom_get_gemm_tile: prec must be 64 or 32");
  *bm = g_tile[prec == 64 ? 0 : 1][0];
  *bn = g_tile[prec == 64 ? 0 : 1][1];
  return 0;
}

extern "C" int mom_cgemm_sub(int prec, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                             void* C, int64_t ldc, void* stream) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (prec == 64) return launch_cgemm<double>(m, n, k, (const z64*)A, lda, (const z64*)B, ldb, (z64*)C, ldc, s);
  if (prec == 32) return launch_cgemm<float>(m, n, k, (const c32*)A, lda, (const c32*)B, ldb, (c32*)C, ldc, s);
  return fail("mom_cgemm_sub: prec must be 64 or 32");
}

extern "C" int mom_residual(int n, int nrhs, const void* Z, int64_t ldz, const void* X, const void* B, void* R,
                            double* rowsum, const double* dsc, void* stream) {
  if (n <= 0) return 0;
  if (nrhs < 1 || nrhs > 4) return fail("mom_residual: 1 to 4 right-hand sides");
  if ((reinterpret_cast<uintptr_t>(Z) | ((uintptr_t)ldz * sizeof(z64))) % 16)
    return fail("mom_residual: Z rows must be 16-byte aligned");
  const dim3 grid((unsigned)((n + 15) / 16)), block(256);  // 4 rows per wavefront
  hipStream_t s = (hipStream_t)stream;
  const z64 *z = (const z64*)Z, *x = (const z64*)X, *b = (const z64*)B;
  z64* r = (z64*)R;
#define MOM_RES(NR)                                                                                     \
  case NR:                                                                                              \
    if (rowsum)                                                                                         \
      hipLaunchKernelGGL((residual_kernel<NR, true>), grid, block, 0, s, n, z, ldz, x, b, r, rowsum, dsc); \
    else                                                                                                \
      hipLaunchKernelGGL((residual_kernel<NR, false>), grid, block, 0, s, n, z, ldz, x, b, r, rowsum, dsc); \
    break;
  switch (nrhs) {
    MOM_RES(1)
    MOM_RES(2)
    MOM_RES(3)
    MOM_RES(4)
  }
#undef MOM_RES
  return check_launch("residual_kernel");
}

extern "C" int mom_scaled_matvec(int n, int nrhs, const void* Z, int64_t ldz, const double* dsc, const void* X,
                                 void* W, void* stream) {
  if (n <= 0) return 0;
  if (nrhs < 1 || nrhs > 4) return fail("mom_scaled_matvec: 1 to 4 right-hand sides");
  if (!Z || !dsc || !X || !W) return fail("mom_scaled_matvec: null device pointer");
  if ((reinterpret_cast<uintptr_t>(Z) | ((uintptr_t)ldz * sizeof(z64))) % 16)
    return fail("mom_scaled_matvec: Z rows must be 16-byte aligned");
  const dim3 grid((unsigned)((n + 15) / 16)), block(256);
  hipStream_t s = (hipStream_t)stream;
  const z64 *z = (const z64*)Z, *x = (const z64*)X;
  z64* w = (z64*)W;
#define MOM_SMV(NR)                                                                                      \
  case NR:                                                                                               \
    hipLaunchKernelGGL((residual_kernel<NR, false, true>), grid, block, 0, s, n, z, ldz, x, nullptr, w, nullptr, \
                       dsc);                                                                             \
    break;
  switch (nrhs) {
    MOM_SMV(1)
    MOM_SMV(2)
    MOM_SMV(3)
    MOM_SMV(4)
  }
#undef MOM_SMV
  return check_launch("residual_kernel");
}

// W = D Z D X for a symmetric Z (Z = Z^T, as the fill leaves it) from the
// 64 x 64 tiles on and above the diagonal only: half the HBM traffic of the
// row-wise product.  symv_strip_kernel: workgroup (c, I) streams the tiles
// Z_IJ, J = I + 8c .. I + 8c + 7, of block row I.  Each tile gives its row
// product Z_IJ (D X)_J, summed over the strip in registers, and, off the
// diagonal, its column product Z_IJ^T (D X)_I, which stands for the tile
// Z_JI below the diagonal.  Wavefront w takes the tile's columns 8w .. 8w + 7
// and then 32 + 8w .. 32 + 8w + 7 (two half-tiles) over all 64 rows: lane
// (a, b) = (lane / 8, lane % 8) holds the entries of rows a + 8u (u < 8) in
// one column, so that a column product reduces over the 8 lanes of equal b
// inside the wavefront (no LDS, no barrier per tile) and each load
// instruction reads 8 rows of 8 consecutive entries.  The half-tiles pass
// through a ring of registers with the next ones' loads in flight; two
// workgroups of 256 threads share a CU, one streaming while the other starts
// or ends its strip.  The partial products go to Q[K][S] (64 x NR each):
// the column product of tile (I, J) to Q[J][I] (S = I < J), the row product
// of strip c of block row I to Q[I][I + c], so that block K's partials are
// Q[K][0 .. K + nstrips(K)) -- summed in a fixed order by
// symv_reduce_kernel, which also applies D (deterministic: no atomics).
constexpr int kSymvT = 64;
constexpr int kSymvChunk = 8;         // tiles per workgroup along a block row
constexpr int kSymvLanesPerRow = 8;   // lanes that read one row's segment per load
// steps (column groups of 4 LPR) in the register ring of
// symv_strip_kernel<NR>: as many as fit in 256 VGPRs without spilling
template <int NR>
constexpr int kSymvBufs = NR == 1 ? 4 : 2;

template <int NR, int NBUF, int LPR>
__global__ __launch_bounds__(256) void symv_strip_kernel(int n, const z64* __restrict__ Z, int64_t ldz,
                                                         const z64* __restrict__ X, int ldx,
                                                         const double* __restrict__ dsc, z64* __restrict__ Q) {
  gmres_step_priority();
  constexpr int RPI = 64 / LPR;              // rows per load instruction
  constexpr int U = kSymvT / RPI;            // rows per lane
  constexpr int SPT = kSymvT / (4 * LPR);    // steps per tile
  constexpr int CH = kSymvChunk, T = SPT * CH;
  const int I = blockIdx.y, nb = gridDim.y;
  const int J0 = I + (int)blockIdx.x * CH;
  if (J0 >= nb) return;  // uniform per workgroup
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int a = lane / LPR, b = lane % LPR;
  __shared__ z64 ys[kSymvT * NR];      // (D X) of block row I, 0 beyond n
  __shared__ z64 part[4][kSymvT][NR];  // the wavefronts' row products
  // The loop body has no branches (a branch around a load makes the
  // compiler wait for every load in flight at the join): indices beyond n
  // are clamped to n - 1, and such an entry (finite) meets a zero of (D X)
  // in both products or lands in a row of W that is never written; tiles
  // past the end of the block row (J >= nb) are read at column n - 1 and
  // contribute zeros; a column product that is not to be kept goes to the
  // workspace's spare slot.
  typedef double dv2 __attribute__((ext_vector_type(2)));
  const z64* zrow[U];
#pragma unroll
  for (int u = 0; u < U; ++u) zrow[u] = Z + (int64_t)min(I * kSymvT + a + RPI * u, n - 1) * ldz;
  z64 zb[NBUF][U], xb[NBUF][NR];
  double db[NBUF];
  auto column = [&](int t) { return (J0 + t / SPT) * kSymvT + (t % SPT) * 4 * LPR + LPR * w + b; };
  auto load = [&](int t, z64 (&z)[U], z64 (&xj)[NR], double& dc) {
    const int c = column(t), cc = min(c, n - 1);
    dc = dsc[cc] * (c < n ? 1.0 : 0.0);
#pragma unroll
    for (int k = 0; k < NR; ++k) xj[k] = X[(int64_t)cc * ldx + k];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // plain loads: measured faster here than non-temporal ones
      const dv2 v = *reinterpret_cast<const dv2*>(zrow[u] + cc);
      z[u] = {v.x, v.y};
    }
  };
  // (D X)_I: one entry per thread (threads past 64 NR repeat the last one)
  const int ye = min(tid, kSymvT * NR - 1), yr = I * kSymvT + ye / NR, yrc = min(yr, n - 1);
  const z64 yx = X[(int64_t)yrc * ldx + ye % NR];
  const double yd = dsc[yrc] * (yr < n ? 1.0 : 0.0);
#pragma unroll
  for (int t = 0; t < NBUF - 1; ++t) load(t, zb[t], xb[t], db[t]);
  __builtin_amdgcn_sched_barrier(0);
  ys[ye] = yx * yd;
  __syncthreads();
  z64 rp[U][NR];
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int k = 0; k < NR; ++k) rp[u][k] = {0.0, 0.0};
  z64* const spare = Q + (int64_t)nb * nb * kSymvT * NR;
#pragma unroll
  for (int t = 0; t < T; ++t) {
    if (t + NBUF - 1 < T)
      load(t + NBUF - 1, zb[(t + NBUF - 1) % NBUF], xb[(t + NBUF - 1) % NBUF], db[(t + NBUF - 1) % NBUF]);
    // keeps the loads of later steps and the reads of ys below in this
    // iteration (hoisted, they would take the registers of the ring)
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    const z64(&z)[U] = zb[t % NBUF];
    z64 xj[NR], cp[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      xj[k] = xb[t % NBUF][k] * db[t % NBUF];
      cp[k] = {0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        rp[u][k] += z[u] * xj[k];
        cp[k] += z[u] * ys[(a + RPI * u) * NR + k];
        // the row sums are due here (left to float to the end of the strip,
        // their updates would keep every step's registers live)
        asm volatile("" : "+v"(rp[u][k].re), "+v"(rp[u][k].im));
      }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      double re = cp[k].re, im = cp[k].im;
#pragma unroll
      for (int off = LPR; off < 64; off <<= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
      }
      cp[k] = {re, im};
    }
    // every lane of the column holds the sum; all write it (same value).
    // The diagonal tile is whole: its row product covers it.
    const int J = J0 + t / SPT, cl = column(t) - J * kSymvT;
    const bool keep = J < nb && J != I;
    z64* q = (keep ? Q + ((int64_t)J * nb + I) * kSymvT * NR : spare) + cl * NR;
#pragma unroll
    for (int k = 0; k < NR; ++k) q[k] = cp[k];
  }
  // row products: the LPR lanes of equal a, then the 4 wavefronts in order
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      double re = rp[u][k].re, im = rp[u][k].im;
#pragma unroll
      for (int off = 1; off < LPR; off <<= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
      }
      if (b == 0) part[w][a + RPI * u][k] = {re, im};
    }
  __syncthreads();
  z64* q = Q + ((int64_t)I * nb + I + (int)blockIdx.x) * kSymvT * NR;  // slot S = I + c
  for (int e = tid; e < kSymvT * NR; e += 256) {
    const int r = e / NR, k = e % NR;
    z64 t = part[0][r][k];
#pragma unroll
    for (int v = 1; v < 4; ++v) t += part[v][r][k];
    q[e] = t;
  }
}

// block K of W: thread (p, e) sums the partials S = p, p + PS, ... < K +
// nchunks(K) of entry e (independent chains of loads per entry), then the
// PS sums in order.
template <int NR>
__global__ __launch_bounds__(1024) void symv_reduce_kernel(int n, int nb, int chunk, const z64* __restrict__ Q,
                                                           const double* __restrict__ dsc, z64* __restrict__ W,
                                                           int ldw) {
  gmres_step_priority();
  constexpr int E = kSymvT * NR, PS = 1024 / E;  // entries per block, partitions of S
  __shared__ z64 part[PS][E];
  const int K = blockIdx.x, e = threadIdx.x % E, p = threadIdx.x / E;
  const int ns = K + (nb - K + chunk - 1) / chunk;
  const z64* q = Q + (int64_t)K * nb * E + e;
  z64 t = {0.0, 0.0};
  for (int S = p; S < ns; S += PS) t += q[(int64_t)S * E];
  part[p][e] = t;
  __syncthreads();
  if (p != 0) return;
  z64 sum = part[0][e];
  for (int k = 1; k < PS; ++k) sum += part[k][e];
  const int row = K * kSymvT + e / NR;
  if (row < n) W[(int64_t)row * ldw + e % NR] = sum * dsc[row];
}

extern "C" int64_t mom_scaled_symv_work_bytes(int n, int nrhs) {
  // the partials of the passes (of at most 2 right-hand sides each) and a
  // spare slot for the column products that are not kept
  const int64_t nb = (n + kSymvT - 1) / kSymvT;
  return (nb * nb + 1) * kSymvT * (int64_t)std::max(1, std::min(nrhs, 2)) * (int64_t)sizeof(z64);
}

extern "C" int mom_scaled_symv(int n, int nrhs, const void* Z, int64_t ldz, const double* dsc, const void* X,
                               void* W, void* work, int64_t work_bytes, void* stream) {
  if (n <= 0) return 0;
  if (nrhs < 1 || nrhs > 4) return fail("mom_scaled_symv: 1 to 4 right-hand sides");
  if (!Z || !dsc || !X || !W || !work) return fail("mom_scaled_symv: null device pointer");
  if (ldz < n) return fail("mom_scaled_symv: ldz < n");
  if ((reinterpret_cast<uintptr_t>(Z) | ((uintptr_t)ldz * sizeof(z64))) % 16)
    return fail("mom_scaled_symv: Z rows must be 16-byte aligned");
  if (work_bytes < mom_scaled_symv_work_bytes(n, nrhs)) return fail("mom_scaled_symv: workspace too small");
  const int nb = (n + kSymvT - 1) / kSymvT;
  if (nb > 65535) return fail("mom_scaled_symv: n too large");
  hipStream_t s = (hipStream_t)stream;
  const z64 *z = (const z64*)Z, *x = (const z64*)X;
  z64 *w = (z64*)W, *q = (z64*)work;
  const int ch = kSymvChunk, nc = (nb + ch - 1) / ch;
  // right-hand sides in passes of at most 2 (the strip kernel's registers)
  for (int k0 = 0; k0 < nrhs; k0 += 2) {
    const int nr = std::min(2, nrhs - k0);
    // timed: the strip pass (the product's pass over Z); algorithmic bytes: the
    // symmetric half of Z and the columns in and out
    const double bytes = 8.0 * n * (n + 1.0) + 32.0 * n * nr;
    if (nr == 2) {
      const int ts = timing_begin(kTimeSymv, s);
      hipLaunchKernelGGL((symv_strip_kernel<2, kSymvBufs<2>, kSymvLanesPerRow>), dim3(nc, nb), dim3(256), 0, s, n, z, ldz, x + k0, nrhs,
                         dsc, q);
      timing_end(kTimeSymv, ts, s, bytes);
      if (int rc = check_launch("symv_strip_kernel")) return rc;
      hipLaunchKernelGGL((symv_reduce_kernel<2>), dim3(nb), dim3(1024), 0, s, n, nb, ch, (const z64*)q, dsc, w + k0,
                         nrhs);
    } else {
      const int ts = timing_begin(kTimeSymv, s);
      hipLaunchKernelGGL((symv_strip_kernel<1, kSymvBufs<1>, kSymvLanesPerRow>), dim3(nc, nb), dim3(256), 0, s, n, z, ldz, x + k0, nrhs,
                         dsc, q);
      timing_end(kTimeSymv, ts, s, bytes);
      if (int rc = check_launch("symv_strip_kernel")) return rc;
      hipLaunchKernelGGL((symv_reduce_kernel<1>), dim3(nb), dim3(1024), 0, s, n, nb, ch, (const z64*)q, dsc, w + k0,
                         nrhs);
    }
    if (int rc = check_launch("symv_reduce_kernel")) return rc;
  }
  return 0;
}

extern "C" int mom_cvt_z_to_c(int64_t n_reals, const double* src, float* dst, void* stream) {
  if (n_reals <= 0) return 0;
  hipLaunchKernelGGL(cvt_z_to_c_kernel, dim3((unsigned)((n_reals + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     n_reals, src, dst);
  return check_launch("cvt_z_to_c_kernel");
}

extern "C" int mom_refine_update(int n, int nrhs, const void* D, void* X, void* dmax, void* xmax, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(refine_update_kernel, dim3((n * nrhs + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, nrhs,
                     (const c32*)D, (z64*)X, (unsigned long long*)dmax, (unsigned long long*)xmax);
  return check_launch("refine_update_kernel");
}

// Timing of the LU's update GEMMs: begin(capacity) creates the events and
// starts recording; end() waits for the recorded launches and returns, in
// out[0..2], the number of launches, their total duration [ms] and their
// total algorithmic flops (8 m n k per complex update).
extern "C" int mom_timing_begin(int kind, int capacity) {
  if (kind < 0 || kind >= kTimeKinds) return fail("mom_timing_begin: kind must be 0..4");
  if (capacity < 0) return fail("mom_timing_begin: negative capacity");
  KernelTiming& t = g_kt[kind];
  std::lock_guard<std::mutex> lock(g_kt_mu);
  for (hipEvent_t e : t.ev) (void)hipEventDestroy(e);
  t.ev.assign(2 * (size_t)capacity, nullptr);
  for (auto& e : t.ev) MOM_TRY(hipEventCreate(&e));
  t.work.clear();
  t.used = 0;
  t.on = true;
  return 0;
}

extern "C" int mom_timing_end(int kind, double* out) {
  if (kind < 0 || kind >= kTimeKinds) return fail("mom_timing_end: kind must be 0..4");
  if (!out) return fail("mom_timing_end: null output");
  KernelTiming& t = g_kt[kind];
  size_t used;
  {
    std::lock_guard<std::mutex> lock(g_kt_mu);
    t.on = false;
    used = t.used;
  }
  double ms = 0.0, work = 0.0;
  for (size_t i = 0; i + 1 < used; i += 2) {
    MOM_TRY(hipEventSynchronize(t.ev[i + 1]));
    float e = 0.f;
    MOM_TRY(hipEventElapsedTime(&e, t.ev[i], t.ev[i + 1]));
    ms += e;
    work += t.work[i / 2];
  }
  out[0] = (double)(used / 2);
  out[1] = ms;
  out[2] = work;
  return 0;
}

extern "C" int mom_basis_dot(int n, int nrhs, int j1, const void* V, const void* w, void* h, void* stream) {
  if (j1 <= 0 || n <= 0) return 0;
  if (nrhs < 1 || nrhs > 4) return fail("mom_basis_dot: 1..4 columns");
  hipLaunchKernelGGL(basis_dot_kernel, dim3(j1), dim3(256), 0, (hipStream_t)stream, n, nrhs, (const z64*)V,
                     (const z64*)w, (z64*)h);
  return check_launch("basis_dot_kernel");
}

extern "C" int mom_basis_sub(int n, int nrhs, int j1, const void* V, const void* h, void* w, void* stream) {
  if (j1 <= 0 || n <= 0) return 0;
  if (nrhs < 1 || nrhs > 4) return fail("mom_basis_sub: 1..4 columns");
  hipLaunchKernelGGL(basis_sub_kernel, dim3((n * nrhs + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, nrhs, j1,
                     (const z64*)V, (const z64*)h, (z64*)w);
  return check_launch("basis_sub_kernel");
}

// ----------------------------------------------------------- GMRES step
// One Arnoldi step of the mixed path's right-preconditioned GMRES in a
// single call (the host then only reads the new Hessenberg column): W_j =
// M^-1 V_j, w = (D Z D) W_j into V_{j+1}, classical Gram-Schmidt twice
// against V_0 .. V_j with the coefficients summed into H[0 .. j], H[j+1] =
// |w| per column, V_{j+1} = w / |w|, and H[0 .. j+1] copied to the host
// buffer Hh asynchronously on the stream.
namespace {

constexpr int kOrthSplit = 8;  // row ranges of the step's dot products and norms

// part[s][i][c] = sum over the rows of range s of conj(V[i, r, c]) w[r, c]:
// workgroup (i, s), the ranges cutting n into kOrthSplit pieces.
__global__ __launch_bounds__(256) void dot_part_kernel(int n, int nrhs, int j1, const z64* __restrict__ V,
                                                       const z64* __restrict__ w, z64* __restrict__ part) {
  gmres_step_priority();
  const int i = blockIdx.x, sp = blockIdx.y, tid = threadIdx.x;
  const int chunk = (n + kOrthSplit - 1) / kOrthSplit, r0 = sp * chunk, r1 = min(n, r0 + chunk);
  z64 acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  const z64* Vi = V + (int64_t)i * n * nrhs;
  for (int r = r0 + tid; r < r1; r += 256)
    for (int c = 0; c < nrhs; ++c) {
      const z64 v = Vi[(int64_t)r * nrhs + c], x = w[(int64_t)r * nrhs + c];
      acc[c] += z64{v.re * x.re + v.im * x.im, v.re * x.im - v.im * x.re};
    }
  __shared__ z64 red[4][4];
  const int lane = tid & 63, wave = tid >> 6;
  for (int c = 0; c < nrhs; ++c) {
    double re = acc[c].re, im = acc[c].im;
    for (int off = 32; off > 0; off >>= 1) {
      re += __shfl_down(re, off, 64);
      im += __shfl_down(im, off, 64);
    }
    if (lane == 0) red[wave][c] = {re, im};
  }
  __syncthreads();
  if (tid < nrhs) {
    z64 t = red[0][tid];
    t += red[1][tid];
    t += red[2][tid];
    t += red[3][tid];
    part[((int64_t)sp * j1 + i) * nrhs + tid] = t;
  }
}

// h[i][c] = sum_s part[s][i][c] in order; with H also H[i][c] += h[i][c]
__global__ void dot_reduce_kernel(int cnt, const z64* __restrict__ part, z64* __restrict__ h, z64* __restrict__ H) {
  gmres_step_priority();
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= cnt) return;
  z64 t = part[q];
  for (int sp = 1; sp < kOrthSplit; ++sp) t += part[(int64_t)sp * cnt + q];
  h[q] = t;
  if (H) H[q] += t;
}

// npart[s][c] = sum over the rows of range s of |w[r, c]|^2: workgroup (c, s)
__global__ __launch_bounds__(256) void norm_part_kernel(int n, int nrhs, const z64* __restrict__ w,
                                                        double* __restrict__ npart) {
  gmres_step_priority();
  const int c = blockIdx.x, sp = blockIdx.y, tid = threadIdx.x;
  const int chunk = (n + kOrthSplit - 1) / kOrthSplit, r0 = sp * chunk, r1 = min(n, r0 + chunk);
  double t = 0.0;
  for (int r = r0 + tid; r < r1; r += 256) {
    const z64 v = w[(int64_t)r * nrhs + c];
    t += v.re * v.re + v.im * v.im;
  }
  for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
  __shared__ double red[4];
  if ((tid & 63) == 0) red[tid >> 6] = t;
  __syncthreads();
  if (tid == 0) npart[sp * nrhs + c] = red[0] + red[1] + red[2] + red[3];
}

// |w[:, c]| from the partials (summed in order by every thread), w[:, c] /= it,
// and hn[c] = (|w[:, c]|, 0) written once
__global__ void col_scale_kernel(int n, int nrhs, z64* __restrict__ w, const double* __restrict__ npart,
                                 z64* __restrict__ hn) {
  gmres_step_priority();
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * nrhs) return;
  const int c = q % nrhs;
  double t = npart[c];
  for (int sp = 1; sp < kOrthSplit; ++sp) t += npart[sp * nrhs + c];
  const double nrm = sqrt(t);
  if (q < nrhs) hn[q] = {nrm, 0.0};
  w[q] = w[q] * (1.0 / nrm);
}

}  // namespace

extern "C" int64_t mom_gmres_work_elems(int m, int nrhs) {
  return (int64_t)(kOrthSplit + 2) * (m + 1) * nrhs + (kOrthSplit * nrhs + 1) / 2 + 1;
}

// The step after its preconditioner solve W_j = M^-1 V_j: product, Gram-Schmidt
// twice, normalisation, Hessenberg column to the host.
static int gmres_step_tail(int n, int nrhs, int j, int m, const void* Z, int64_t ldz, const double* dsc,
                           void* symv_work, int64_t symv_bytes, void* V, void* W, void* H, void* Hh, hipStream_t s) {
  void* stream = (void*)s;
  const int64_t vec = (int64_t)n * nrhs;
  z64* Vj = (z64*)V + (int64_t)j * vec;
  z64* Vn = Vj + vec;  // w, then V_{j+1}
  z64* Wj = (z64*)W + (int64_t)j * vec;
  z64* Hc = (z64*)H;                            // (m + 1) x nrhs: this column
  z64* T = Hc + (int64_t)(m + 1) * nrhs;        // (m + 1) x nrhs: second Gram-Schmidt pass
  z64* part = T + (int64_t)(m + 1) * nrhs;      // kOrthSplit x (m + 1) x nrhs: partial dot products
  double* npart = reinterpret_cast<double*>(part + (int64_t)kOrthSplit * (m + 1) * nrhs);  // kOrthSplit x nrhs
  int rc;
  if (symv_work)
    rc = mom_scaled_symv(n, nrhs, Z, ldz, dsc, Wj, Vn, symv_work, symv_bytes, stream);
  else
    rc = mom_scaled_matvec(n, nrhs, Z, ldz, dsc, Wj, Vn, stream);
  if (rc) return rc;
  const int j1 = j + 1, cnt = j1 * nrhs;
  const z64* Vc = (const z64*)V;
  for (int pass = 0; pass < 2; ++pass) {  // classical Gram-Schmidt, twice
    z64* h = pass == 0 ? Hc : T;
    hipLaunchKernelGGL(dot_part_kernel, dim3(j1, kOrthSplit), dim3(256), 0, s, n, nrhs, j1, Vc, (const z64*)Vn, part);
    if ((rc = check_launch("dot_part_kernel"))) return rc;
    hipLaunchKernelGGL(dot_reduce_kernel, dim3((cnt + 255) / 256), dim3(256), 0, s, cnt, (const z64*)part, h,
                       pass == 0 ? (z64*)nullptr : Hc);
    if ((rc = check_launch("dot_reduce_kernel"))) return rc;
    if ((rc = mom_basis_sub(n, nrhs, j1, V, h, Vn, stream))) return rc;
  }
  hipLaunchKernelGGL(norm_part_kernel, dim3(nrhs, kOrthSplit), dim3(256), 0, s, n, nrhs, (const z64*)Vn, npart);
  if ((rc = check_launch("norm_part_kernel"))) return rc;
  hipLaunchKernelGGL(col_scale_kernel, dim3((unsigned)((vec + 255) / 256)), dim3(256), 0, s, n, nrhs, Vn,
                     (const double*)npart, Hc + (int64_t)j1 * nrhs);
  if ((rc = check_launch("col_scale_kernel"))) return rc;
  MOM_TRY(hipMemcpyAsync(Hh, Hc, sizeof(z64) * (size_t)(j + 2) * nrhs, hipMemcpyDeviceToHost, s));
  return 0;
}

extern "C" int mom_gmres_arnoldi(int prec, int n, int nrhs, int j, int m, const void* LU, int64_t lda,
                                 const int32_t* perm, void* solve_work, const void* Z, int64_t ldz, const double* dsc,
                                 void* symv_work, int64_t symv_bytes, void* V, void* W, void* H, void* Hh,
                                 void* stream) {
  if (n <= 0) return 0;
  if (nrhs < 1 || nrhs > 4) return fail("mom_gmres_arnoldi: 1 to 4 right-hand sides");
  if (j < 0 || j >= m) return fail("mom_gmres_arnoldi: step outside the restart length");
  if (!V || !W || !H || !Hh) return fail("mom_gmres_arnoldi: null pointer");
  const int64_t vec = (int64_t)n * nrhs;
  z64* Vj = (z64*)V + (int64_t)j * vec;
  z64* Wj = (z64*)W + (int64_t)j * vec;
  int rc;
  if ((rc = mom_getrs_inv_z(prec, n, nrhs, LU, lda, perm, Vj, Wj, solve_work, stream))) return rc;
  return gmres_step_tail(n, nrhs, j, m, Z, ldz, dsc, symv_work, symv_bytes, V, W, H, Hh, (hipStream_t)stream);
}

extern "C" int mom_gmres_arnoldi_sb(int prec, int n, int nrhs, int j, int m, int sb, const int32_t* perm,
                                    void* sb_work, const void* Z, int64_t ldz, const double* dsc, void* symv_work,
                                    int64_t symv_bytes, void* V, void* W, void* H, void* Hh, void* stream) {
  if (n <= 0) return 0;
  if (nrhs < 1 || nrhs > 4) return fail("mom_gmres_arnoldi_sb: 1 to 4 right-hand sides");
  if (j < 0 || j >= m) return fail("mom_gmres_arnoldi_sb: step outside the restart length");
  if (!V || !W || !H || !Hh || !perm || !sb_work) return fail("mom_gmres_arnoldi_sb: null pointer");
  const int64_t vec = (int64_t)n * nrhs;
  z64* Vj = (z64*)V + (int64_t)j * vec;
  z64* Wj = (z64*)W + (int64_t)j * vec;
  int rc;
  if ((rc = mom_sbsolve_z(prec, n, nrhs, sb, perm, Vj, Wj, sb_work, stream))) return rc;
  return gmres_step_tail(n, nrhs, j, m, Z, ldz, dsc, symv_work, symv_bytes, V, W, H, Hh, (hipStream_t)stream);
}
