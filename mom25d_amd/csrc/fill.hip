// FP64 MPIE impedance-matrix fill for RWG bases on triangles.
//
// Work is organised as in Engine/mom25d/_fast.py:82-158 (group_vals +
// scatter) and Engine/mom25d/solver.py:429-524 (assemble): test triangles
// come in Morton-ordered blocks of up to 32, and every source triangle has
// one of three tiers per (test block, source triangle):
//   near  7-point source rule on the smooth part, plus the closed-form
//         integrals of 1/R and (r'-r)/R (Wilton et al. 1984) at each test point
//   mid   3-point source rule on the full kernel
//   far   1-point (centroid) source rule on the full kernel
// The test side always uses the 7-point Dunavant rule.  For a triangle pair
// (p, s) the kernel forms the 3x3 half-basis block
//   vals[i][j] = jw mu0 cA_p,i cA_s,j IA[i][j] + cV_p,i cV_s,j IV / (jw eps0)
// with IA[i][j] = sum_k w_k (r_k - a_i).(K1A(r_k) - b_j K0A(r_k)),
//      IV       = sum_k w_k K0V(r_k),
// and adds it to Z[basis(p,i), basis(s,j)] (boundary edges carry no basis).
//
// GPU mapping: one workgroup (4 wavefronts) per work item = (test block,
// source layer, tier, chunk of <= 64 source triangles).  Lane l of every
// wavefront owns source triangle l of the chunk; wavefront w walks test
// triangles w, w+4, ... of the block, whose data is staged in LDS.  The
// smooth kernel is a cubic spline read from global memory (L1/L2 resident);
// its interval comes from a uniform bucket table whose entries hold the
// interval at the start of the bucket and the one node inside it, so the
// lookup is one load and one comparison (no data-dependent loop), giving
// exactly scipy's find_interval result (clamped end intervals).
// Contributions are summed into Z with hardware FP64 atomics, except for the
// entries that only one work item of a launch touches (basis functions whose
// triangles all sit in the item's test block and source chunk), which are
// added with plain loads and stores; the matrix is symmetrised by a later
// pass, as the reference does.

#include "fill.h"

#include <type_traits>

namespace mom {

constexpr int kStageMax = 64;        // spline intervals one workgroup can stage
constexpr int kStageStride = 18;     // doubles per staged interval (16 + pad: spreads LDS banks)
constexpr int kStageBuckets = 2048;  // lookup buckets one workgroup can stage (one byte each)

// Entry b of a bucket table covers the distances R with (int)(R * inv_dx)
// == b (the last one every larger R too): i0 is the spline interval at the
// start of that range and split the node inside it (+inf if none), the host
// having placed at most one node per range (width <= half the smallest
// spacing) and computed the ranges for this exact double arithmetic.
__device__ inline int spline_interval(const Bucket* __restrict__ bucket, int nbucket, double inv_dx, double R) {
  int b = (int)(R * inv_dx);
  b = b < 0 ? 0 : (b >= nbucket ? nbucket - 1 : b);
  const Bucket e = bucket[b];
  return e.i0 + (R >= e.split ? 1 : 0);
}

// S_A(R), S_V(R) of the tabulated smooth part (scipy PPoly evaluation).
__device__ inline void spline_eval(const PairTab& pt, const double* __restrict__ x, const double* __restrict__ c,
                                   const Bucket* __restrict__ bucket, double R, z64& sa, z64& sv) {
  const int i = spline_interval(bucket, pt.nbucket, pt.inv_dx, R);
  double d = R - x[i];
  const double* q = c + (int64_t)i * 16;  // [k][kernel][re,im], k = 0..3 (cubic .. constant)
  double z = d;
  double ar = q[12] + q[8] * z, ai = q[13] + q[9] * z;
  double vr = q[14] + q[10] * z, vi = q[15] + q[11] * z;
  z *= d;
  ar += q[4] * z; ai += q[5] * z; vr += q[6] * z; vi += q[7] * z;
  z *= d;
  ar += q[0] * z; ai += q[1] * z; vr += q[2] * z; vi += q[3] * z;
  sa = {ar, ai};
  sv = {vr, vi};
}

// 1 / sqrt(x) for x > 0: the hardware estimate refined by one Newton step
// (y (3 - x y^2) / 2, written as y + y e / 2 with e = 1 - x y^2 by FMA).
__device__ inline double rsqrt_refined(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// The same from the intervals [base, base + nst) staged in LDS: nodes
// sx[0 .. nst] (x[base ..]), coefficients sc (kStageStride doubles per
// interval) and, for the buckets [b0, b0 + nbk), the interval at each
// bucket's start relative to base (sb).  The interval of R is then
// sb[b] + (R >= sx[sb[b] + 1]), clamped to the last staged interval: the
// global table's i0 + (R >= split), since a bucket holds at most one node
// and its split is that node, or +inf when there is none or crossing it
// leaves the clamped interval unchanged.  Identical arithmetic, so identical
// values; R lies in the item's bounds, so every index is in the staged range.
__device__ inline void spline_eval_staged(const PairTab& pt, const double* sx, const double* sc,
                                          const unsigned char* sb, int b0, int nbk, int nst, double R, z64& sa,
                                          z64& sv) {
  int b = (int)(R * pt.inv_dx);
  b = (b < 0 ? 0 : (b >= pt.nbucket ? pt.nbucket - 1 : b)) - b0;
  b = b < 0 ? 0 : (b >= nbk ? nbk - 1 : b);
  int i = sb[b];
  i += R >= sx[i + 1] ? 1 : 0;
  i = i >= nst ? nst - 1 : i;
  double d = R - sx[i];
  const double* q = sc + i * kStageStride;
  double z = d;
  double ar = q[12] + q[8] * z, ai = q[13] + q[9] * z;
  double vr = q[14] + q[10] * z, vi = q[15] + q[11] * z;
  z *= d;
  ar += q[4] * z; ai += q[5] * z; vr += q[6] * z; vi += q[7] * z;
  z *= d;
  ar += q[0] * z; ai += q[1] * z; vr += q[2] * z; vi += q[3] * z;
  sa = {ar, ai};
  sv = {vr, vi};
}

// Closed-form I0 = int_T 1/R dS' and I1 = int_T (r'-r)/R dS' (in-plane part)
// for an observer at (rx, ry) a height d above the triangle's plane
// (Engine/mom25d/integrals.py:24-67).
__device__ inline void wilton(const double* Vs, double rx, double ry, double d, double& I0, double& I1x,
                              double& I1y) {
  const double d2 = d * d;
  I0 = 0.0; I1x = 0.0; I1y = 0.0;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int a = (e + 1) % 3, b = (e + 2) % 3;
    double pmx = Vs[2 * a], pmy = Vs[2 * a + 1];
    double ppx = Vs[2 * b], ppy = Vs[2 * b + 1];
    double Lx = ppx - pmx, Ly = ppy - pmy;
    double ln = sqrt(Lx * Lx + Ly * Ly);
    double lx = Lx / ln, ly = Ly / ln;
    double ux = ly, uy = -lx;
    double dmx = pmx - rx, dmy = pmy - ry;
    double dpx = ppx - rx, dpy = ppy - ry;
    double lm = dmx * lx + dmy * ly;
    double lp = dpx * lx + dpy * ly;
    double t = dmx * ux + dmy * uy;
    double Rm = sqrt(dmx * dmx + dmy * dmy + d2);
    double Rp = sqrt(dpx * dpx + dpy * dpy + d2);
    double R02 = t * t + d2;
    double logt = 0.0;
    if (!(sqrt(R02) <= 1e-12 * ln)) {
      double fp = lp >= 0.0 ? Rp + lp : R02 / (Rp - lp);
      double fm = lm >= 0.0 ? Rm + lm : R02 / (Rm - lm);
      logt = log(fp / fm);
    }
    I0 += t * logt;
    if (d != 0.0) I0 -= d * (atan2(t * lp, R02 + d * Rp) - atan2(t * lm, R02 + d * Rm));
    double g = 0.5 * (R02 * logt + lp * Rp - lm * Rm);
    I1x += ux * g;
    I1y += uy * g;
  }
}

// One workgroup per work item (test block, source chunk of one tier): lane l
// of every wavefront owns source triangle l, wavefront w walks the block's
// test triangles w, w + NW, ...  The 3 x 3 contributions of every pair are
// summed in LDS per (block row, item column) -- the item's distinct test and
// source basis functions -- and each entry goes to Z with one atomic add, so
// that Z sees a third of the scattered updates of a per-pair scatter.
// The LDS accumulator allows one workgroup per CU, so a workgroup holds as
// many wavefronts as the registers allow: 16 on the far tier (4 per SIMD,
// 128 registers), 12 on the mid tier, 8 on the near tier, whose closed
// forms need about 250 registers per lane.
template <int TIER>
struct FillShape {
  static constexpr int threads = TIER == 0 ? 512 : (TIER == 1 ? 768 : 1024);
};

template <int TIER>
__global__ __launch_bounds__(FillShape<TIER>::threads) void fill_kernel(FillArgs a) {
  constexpr int L = TIER == 0 ? 7 : (TIER == 1 ? 3 : 1);
  constexpr int kFillThreads = FillShape<TIER>::threads;
  constexpr int NW = kFillThreads / 64;
  extern __shared__ double acc[];  // (nr, nc) complex, interleaved
  __shared__ double s_qp[kBlock][14];
  __shared__ double s_qw[kBlock][7];
  __shared__ double s_V[kBlock][6];
  __shared__ double s_cA[kBlock][3];
  __shared__ double s_cV[kBlock][3];
  __shared__ int32_t s_rl[kBlock][3];
  __shared__ int32_t s_tri[kBlock];
  __shared__ int32_t s_rows[3 * kBlock];
  __shared__ int32_t s_cols[3 * 64];
  __shared__ double s_sx[kStageMax + 1];             // staged spline nodes
  __shared__ double s_sc[kStageMax * kStageStride];  // and coefficients
  __shared__ unsigned char s_sb[kStageBuckets];      // and buckets

  const int* item = a.items + 6 * (int64_t)blockIdx.x;
  const int blk = item[0], pp = item[1], soff = item[2], scnt = item[3], coff = item[4], nc = item[5];
  const int r0 = a.blk_row_off[blk], nr = a.blk_row_off[blk + 1] - r0;
  const int tid = threadIdx.x;
  const PairTab pt = a.pairs[pp];
  const double* __restrict__ tx = a.tab_x + pt.x_off;
  const double* __restrict__ tc = a.tab_c + pt.c_off;
  const Bucket* __restrict__ tbk = a.tab_b + pt.b_off;
  // the spline intervals and lookup buckets of this item's distances
  // (uniform per workgroup); base is one interval early, so that the
  // buckets' start intervals (at most one node before the item's smallest
  // distance) are staged too
  const double rlo = a.item_r[2 * (int64_t)blockIdx.x], rhi = a.item_r[2 * (int64_t)blockIdx.x + 1];
  const int base = max(spline_interval(tbk, pt.nbucket, pt.inv_dx, rlo) - 1, 0);
  const int nst = spline_interval(tbk, pt.nbucket, pt.inv_dx, rhi) - base + 1;
  const int b0 = min((int)(rlo * pt.inv_dx), pt.nbucket - 1);
  const int nbk = min((int)(rhi * pt.inv_dx), pt.nbucket - 1) - b0 + 1;
  const bool staged = nst <= min(a.stage_max, kStageMax) && nbk <= kStageBuckets;
  if (staged) {
    for (int e = tid; e < nst * 16; e += kFillThreads)
      s_sc[(e / 16) * kStageStride + e % 16] = tc[(int64_t)base * 16 + e];
    for (int e = tid; e <= nst; e += kFillThreads) s_sx[e] = tx[base + e];
    for (int e = tid; e < nbk; e += kFillThreads) s_sb[e] = (unsigned char)(tbk[b0 + e].i0 - base);
  }
  if (tid < kBlock) {
    int p = a.blk_tris[blk * kBlock + tid];
    s_tri[tid] = p;
    if (p >= 0) {
      for (int q = 0; q < 14; ++q) s_qp[tid][q] = a.qp7[14 * (int64_t)p + q];
      for (int q = 0; q < 7; ++q) s_qw[tid][q] = a.qw7[7 * (int64_t)p + q];
      for (int q = 0; q < 6; ++q) s_V[tid][q] = a.V[6 * (int64_t)p + q];
      for (int q = 0; q < 3; ++q) {
        s_cA[tid][q] = a.cA[3 * (int64_t)p + q];
        s_cV[tid][q] = a.cV[3 * (int64_t)p + q];
        s_rl[tid][q] = a.blk_rloc[(blk * kBlock + tid) * 3 + q];
      }
    }
  }
  // rows and columns with bit 31 set when another work item of this launch
  // may add to their entries too
  const bool owned = a.row_own != nullptr && a.col_own != nullptr;
  for (int e = tid; e < nr; e += kFillThreads)
    s_rows[e] = a.blk_rows[r0 + e] | (owned && a.row_own[r0 + e] ? 0 : INT32_MIN);
  for (int e = tid; e < nc; e += kFillThreads)
    s_cols[e] = a.item_cols[coff + e] | (owned && a.col_own[coff + e] ? 0 : INT32_MIN);
  for (int e = tid; e < 2 * nr * nc; e += kFillThreads) acc[e] = 0.0;
  __syncthreads();

  const z64 kA = {pt.cA_re, pt.cA_im}, kV = {pt.cV_re, pt.cV_im};
  const double dz = pt.dz;
  const z64 jwmu = {a.jwmu_re, a.jwmu_im}, ijwe = {a.ijwe_re, a.ijwe_im};

  const int lane = tid & 63, wave = tid >> 6;
  // the pair loop, instantiated for staged and global spline reads
  auto pairs = [&](auto staged_tag) {
    constexpr bool STAGED = decltype(staged_tag)::value;
    if (lane >= scnt) return;
    const int s = a.src[soff + lane];
    double Vs[6], px[L], py[L], w[L], cAs[3], cVs[3];
    int cl[3];
#pragma unroll
    for (int q = 0; q < 6; ++q) Vs[q] = a.V[6 * (int64_t)s + q];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      if (TIER == 0) {
        px[l] = a.qp7[14 * (int64_t)s + 2 * l];
        py[l] = a.qp7[14 * (int64_t)s + 2 * l + 1];
        w[l] = a.qw7[7 * (int64_t)s + l];
      } else if (TIER == 1) {
        px[l] = a.qp3[6 * (int64_t)s + 2 * l];
        py[l] = a.qp3[6 * (int64_t)s + 2 * l + 1];
        w[l] = a.qw3[3 * (int64_t)s + l];
      } else {
        px[l] = a.cen[2 * (int64_t)s];
        py[l] = a.cen[2 * (int64_t)s + 1];
        w[l] = a.area[s];
      }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      cAs[q] = a.cA[3 * (int64_t)s + q];
      cVs[q] = a.cV[3 * (int64_t)s + q];
      cl[q] = a.src_cloc[3 * (int64_t)(soff + lane) + q];
    }
    const bool pads = a.IVpad != nullptr;
    const double* near_st = TIER == 0 ? a.near_static + a.item_soff[blockIdx.x] + lane : nullptr;

    for (int pl = wave; pl < kBlock; pl += NW) {
      const int p = s_tri[pl];
      if (p < 0) break;
      // Positions relative to the test triangle's first vertex o, so that
      // the moments below stay of the size of the triangles and their
      // distance (no cancellation of absolute coordinates).  With a_ik =
      // w_k (r_k - v_i), IA[i][j] = sum_k a_ik . (K1_k - v_j K0_k) is formed
      // from six sums over the test points:
      //   A0 = sum w_k K0A_k, A = sum w_k r_k K0A_k, C = sum w_k K1_k,
      //   B = sum w_k r_k . K1_k
      // as M_i = A - v_i A0 (= sum_k a_ik K0A_k) and T_i = B - v_i . C
      // (= sum_k a_ik . K1_k), IA[i][j] = T_i - v_j . M_i.  On the far tier
      // (one source point p') K1 = p' K0A, so T_i = p' . M_i and
      // IA[i][j] = (p' - v_j) . M_i.
      const double ox = s_V[pl][0], oy = s_V[pl][1];
      z64 A0 = {0.0, 0.0}, Ax = A0, Ay = A0, Bs = A0, Cx = A0, Cy = A0, IV = A0;
      // fully unrolled, the far tier would need more registers than
      // occupancy allows; two test points at a time keep two spline
      // lookups in flight there
      constexpr int KU = TIER == 2 ? 2 : 1;
#pragma unroll KU
      for (int k = 0; k < 7; ++k) {
        const double rx = s_qp[pl][2 * k], ry = s_qp[pl][2 * k + 1];
        z64 k0a, k1x, k1y, k0v;
        if (TIER == 0) {
          const double* st = near_st + (int64_t)(3 * (7 * pl + k)) * scnt;
          const double S0 = st[0], S1x = st[scnt], S1y = st[2 * scnt];
          k0a = kA * S0; k1x = kA * S1x; k1y = kA * S1y; k0v = kV * S0;
        } else {
          k0a = {0.0, 0.0}; k1x = k0a; k1y = k0a; k0v = k0a;
        }
#pragma unroll
        for (int l = 0; l < L; ++l) {
          const double ddx = rx - px[l], ddy = ry - py[l];
          const double r2 = ddx * ddx + ddy * ddy;
          double R;
          z64 sa, sv;
          if (TIER != 0) {
            // 1 / R and R from one reciprocal square root refined by a
            // Newton step (within an ulp or two of the rounded sqrt and
            // divide; R > 0 off the near tier)
            const double y = rsqrt_refined(r2);
            R = r2 * y;
            const double inv = (dz == 0.0 ? y : rsqrt_refined(r2 + dz * dz)) * (1.0 / kFourPi);
            if (STAGED)
              spline_eval_staged(pt, s_sx, s_sc, s_sb, b0, nbk, nst, R, sa, sv);
            else
              spline_eval(pt, tx, tc, tbk, R, sa, sv);
            sa += kA * inv;
            sv += kV * inv;
          } else {
            R = sqrt(r2);
            if (STAGED)
              spline_eval_staged(pt, s_sx, s_sc, s_sb, b0, nbk, nst, R, sa, sv);
            else
              spline_eval(pt, tx, tc, tbk, R, sa, sv);
          }
          const z64 saw = sa * w[l];
          k0a += saw;
          if (TIER != 2) {
            k1x += saw * (px[l] - ox);
            k1y += saw * (py[l] - oy);
          }
          k0v += sv * w[l];
        }
        const double wk = s_qw[pl][k], lx = rx - ox, ly = ry - oy;
        IV += wk * k0v;
        const z64 a0 = wk * k0a;
        A0 += a0;
        Ax += lx * a0;
        Ay += ly * a0;
        if (TIER != 2) {
          const z64 cx = wk * k1x, cy = wk * k1y;
          Cx += cx;
          Cy += cy;
          Bs += lx * cx + ly * cy;
        }
      }
      if (pads) {
        const int pr = a.pad_pos[p];
        if (pr >= 0) {
          z64* dst = reinterpret_cast<z64*>(a.IVpad) + (int64_t)pr * a.nt + s;
          *dst = IV;
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int rl = s_rl[pl][i];
        if (rl < 0) continue;
        const double vx = s_V[pl][2 * i] - ox, vy = s_V[pl][2 * i + 1] - oy;
        const z64 Mx = Ax - vx * A0, My = Ay - vy * A0;
        const z64 Ti = TIER == 2 ? z64{0.0, 0.0} : Bs - (vx * Cx + vy * Cy);
        const double cAi = s_cA[pl][i], cVi = s_cV[pl][i];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          if (cl[j] < 0) continue;
          const double ux = Vs[2 * j] - ox, uy = Vs[2 * j + 1] - oy;
          const z64 ia = TIER == 2 ? (px[0] - Vs[2 * j]) * Mx + (py[0] - Vs[2 * j + 1]) * My
                                   : Ti - (ux * Mx + uy * My);
          const z64 v = (jwmu * (cAi * cAs[j])) * ia + (ijwe * (cVi * cVs[j])) * IV;
          double* dst = acc + 2 * (rl * nc + cl[j]);
          atomicAdd(dst, v.re);
          atomicAdd(dst + 1, v.im);
        }
      }
    }
  };
  if (staged)
    pairs(std::true_type{});
  else
    pairs(std::false_type{});
  __syncthreads();
  // FU entries per thread at a time: the Z values of the owned ones are all
  // loaded before the first is stored, so that a workgroup's flush waits one
  // memory round trip rather than one per entry
  constexpr int FU = 8;
  for (int e0 = tid; e0 < nr * nc; e0 += FU * kFillThreads) {
    double2 z[FU];
#pragma unroll
    for (int u = 0; u < FU; ++u) {
      const int e = e0 + u * kFillThreads;
      if (e >= nr * nc) break;
      const int32_t rr = s_rows[e / nc], cc = s_cols[e % nc];
      if ((rr | cc) >= 0)  // this work item is the launch's only writer of the entry
        z[u] = *reinterpret_cast<const double2*>(a.Z + 2 * ((int64_t)rr * a.ldz + cc));
    }
#pragma unroll
    for (int u = 0; u < FU; ++u) {
      const int e = e0 + u * kFillThreads;
      if (e >= nr * nc) break;
      const double re = acc[2 * e], im = acc[2 * e + 1];
      const int32_t rr = s_rows[e / nc], cc = s_cols[e % nc];
      double* dst = a.Z + 2 * ((int64_t)(rr & INT32_MAX) * a.ldz + (cc & INT32_MAX));
      if ((rr | cc) >= 0) {
        z[u].x += re;
        z[u].y += im;
        *reinterpret_cast<double2*>(dst) = z[u];
      } else if (re != 0.0 || im != 0.0) {
        unsafeAtomicAdd(dst, re);
        unsafeAtomicAdd(dst + 1, im);
      }
    }
  }
}

// The near tier's closed-form parts, once per mesh (Engine/mom25d/solver.py:
// 373-376 likewise computes them in its set-up): for every near work item,
// test triangle pl of its block, test point k and source triangle (lane l),
//   S0 = I0 / 4 pi,   S1 = ((r_k - o) I0 + I1) / 4 pi,
// with I0, I1 the integrals of 1/R and (r'-r)/R over the source triangle at
// the pair's level distance and o the test triangle's first vertex.  The
// same arithmetic the pair kernel would do in line, so Z is unchanged.
constexpr int kStaticThreads = 256;

__global__ __launch_bounds__(kStaticThreads) void near_static_kernel(FillArgs a, const double* __restrict__ pair_dz,
                                                                      double* __restrict__ out) {
  const int* item = a.items + 6 * (int64_t)blockIdx.x;
  const int blk = item[0], pp = item[1], soff = item[2], scnt = item[3];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane >= scnt) return;
  const int s = a.src[soff + lane];
  double Vs[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) Vs[q] = a.V[6 * (int64_t)s + q];
  const double dz = pair_dz[pp];
  double* dst = out + a.item_soff[blockIdx.x] + lane;
  for (int pl = wave; pl < kBlock; pl += kStaticThreads / 64) {
    const int p = a.blk_tris[blk * kBlock + pl];
    if (p < 0) break;
    const double ox = a.V[6 * (int64_t)p], oy = a.V[6 * (int64_t)p + 1];
    for (int k = 0; k < 7; ++k) {
      const double rx = a.qp7[14 * (int64_t)p + 2 * k], ry = a.qp7[14 * (int64_t)p + 2 * k + 1];
      double I0, I1x, I1y;
      wilton(Vs, rx, ry, dz, I0, I1x, I1y);
      double* st = dst + (int64_t)(3 * (7 * pl + k)) * scnt;
      st[0] = I0 / kFourPi;
      st[scnt] = ((rx - ox) * I0 + I1x) / kFourPi;
      st[2 * scnt] = ((ry - oy) * I0 + I1y) / kFourPi;
    }
  }
}

// Z[b_i, b_j] += zs * g[t][i][j]: conductor loss through the Gram matrix of
// the RWG functions on each triangle (Engine/mom25d/solver.py:385-391, 511-513).
__global__ void gram_kernel(int nt, const double* __restrict__ g, const int32_t* __restrict__ tb, double zs_re,
                            double zs_im, double* __restrict__ Z, int64_t ldz) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  for (int i = 0; i < 3; ++i) {
    int bi = tb[3 * t + i];
    if (bi < 0) continue;
    for (int j = 0; j < 3; ++j) {
      int bj = tb[3 * t + j];
      if (bj < 0) continue;
      double v = g[9 * t + 3 * i + j];
      double* dst = Z + 2 * ((int64_t)bi * ldz + bj);
      unsafeAtomicAdd(dst, zs_re * v);
      unsafeAtomicAdd(dst + 1, zs_im * v);
    }
  }
}

// Via rows / columns (Engine/mom25d/solver.py:527-566):
//   Zvb[k, n] = 1/(jw eps0) sum_pad D[pad, k] (cV+ IVpad[pad, T+_n] + cV- IVpad[pad, T-_n])
// written to Z[nb + k, n] and Z[n, nb + k].
__global__ void via_rows_kernel(int nb, int nv, int nt, const double* __restrict__ IVpad,
                                const int32_t* __restrict__ d_off, const int32_t* __restrict__ d_pad,
                                const double* __restrict__ d_val, const int32_t* __restrict__ tp,
                                const int32_t* __restrict__ vp, const int32_t* __restrict__ tm,
                                const int32_t* __restrict__ vm, const double* __restrict__ cV, double ijwe_re,
                                double ijwe_im, double* __restrict__ Z, int64_t ldz) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (n >= nb) return;
  const z64* IV = reinterpret_cast<const z64*>(IVpad);
  const int a = tp[n], b = tm[n];
  const double ca = cV[3 * a + vp[n]], cb = cV[3 * b + vm[n]];
  z64 acc = {0.0, 0.0};
  for (int q = d_off[k]; q < d_off[k + 1]; ++q) {
    const int64_t row = (int64_t)d_pad[q] * nt;
    z64 y = ca * IV[row + a] + cb * IV[row + b];
    acc += d_val[q] * y;
  }
  const z64 v = z64{ijwe_re, ijwe_im} * acc;
  z64* Zc = reinterpret_cast<z64*>(Z);
  Zc[(int64_t)(nb + k) * ldz + n] = v;
  Zc[(int64_t)n * ldz + nb + k] = v;
}

//   Zvv[k, m] = 1/(jw eps0) sum D[pad,k] IVpad[pad, tri(pad')] D[pad',m] + E[k, m]
__global__ void via_block_kernel(int nb, int nv, int nt, const double* __restrict__ IVpad,
                                 const int32_t* __restrict__ d_off, const int32_t* __restrict__ d_pad,
                                 const double* __restrict__ d_val, const int32_t* __restrict__ pad_tri,
                                 const double* __restrict__ E, double ijwe_re, double ijwe_im,
                                 double* __restrict__ Z, int64_t ldz) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (m >= nv) return;
  const z64* IV = reinterpret_cast<const z64*>(IVpad);
  z64 acc = {0.0, 0.0};
  for (int q = d_off[k]; q < d_off[k + 1]; ++q) {
    const int64_t row = (int64_t)d_pad[q] * nt;
    z64 inner = {0.0, 0.0};
    for (int r = d_off[m]; r < d_off[m + 1]; ++r) inner += d_val[r] * IV[row + pad_tri[d_pad[r]]];
    acc += d_val[q] * inner;
  }
  const z64 e = reinterpret_cast<const z64*>(E)[(int64_t)k * nv + m];
  reinterpret_cast<z64*>(Z)[(int64_t)(nb + k) * ldz + nb + m] = z64{ijwe_re, ijwe_im} * acc + e;
}

// Z <- 0.5 (Z + Z^T) in place (Engine/mom25d/solver.py:522-524); optionally
// also writes a complex64 copy Z32 (leading dimension ld32) of the result,
// scaled to d_i Z_ij d_j when a scaling vector d is given.
__global__ __launch_bounds__(256) void symmetrize_kernel(int n, double* __restrict__ Z, int64_t ldz,
                                                         float* __restrict__ Z32, int64_t ld32,
                                                         const double* __restrict__ dsc) {
  const int I = blockIdx.y, J = blockIdx.x;
  if (I > J) return;
  __shared__ z64 A[32][33];
  __shared__ z64 B[32][33];
  z64* Zc = reinterpret_cast<z64*>(Z);
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    int gi = I * 32 + r, gj = J * 32 + tx;
    if (gi < n && gj < n) A[r][tx] = Zc[(int64_t)gi * ldz + gj];
    gi = J * 32 + r; gj = I * 32 + tx;
    if (gi < n && gj < n) B[r][tx] = Zc[(int64_t)gi * ldz + gj];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    int gi = I * 32 + r, gj = J * 32 + tx;
    if (gi < n && gj < n) {
      z64 a = A[r][tx], b = B[tx][r];
      z64 v = {0.5 * (a.re + b.re), 0.5 * (a.im + b.im)};
      Zc[(int64_t)gi * ldz + gj] = v;
      if (Z32) {
        const double sc = dsc ? dsc[gi] * dsc[gj] : 1.0;
        Z32[2 * ((int64_t)gi * ld32 + gj)] = (float)(sc * v.re);
        Z32[2 * ((int64_t)gi * ld32 + gj) + 1] = (float)(sc * v.im);
      }
    }
    gi = J * 32 + r; gj = I * 32 + tx;
    if (I != J && gi < n && gj < n) {
      z64 a = B[r][tx], b = A[tx][r];
      z64 v = {0.5 * (a.re + b.re), 0.5 * (a.im + b.im)};
      Zc[(int64_t)gi * ldz + gj] = v;
      if (Z32) {
        const double sc = dsc ? dsc[gi] * dsc[gj] : 1.0;
        Z32[2 * ((int64_t)gi * ld32 + gj)] = (float)(sc * v.re);
        Z32[2 * ((int64_t)gi * ld32 + gj) + 1] = (float)(sc * v.im);
      }
    }
  }
}

// d[i] = 1 / sqrt(|Z[i, i]|) (1 where the diagonal vanishes): the symmetric
// diagonal scaling applied to the complex64 copy of Z.
__global__ void diag_scale_kernel(int n, const double* __restrict__ Z, int64_t ldz, double* __restrict__ d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double re = Z[2 * ((int64_t)i * ldz + i)], im = Z[2 * ((int64_t)i * ldz + i) + 1];
  const double a = sqrt(re * re + im * im);
  d[i] = a > 0.0 ? 1.0 / sqrt(a) : 1.0;
}

// Line voltage to the plane below at observation points on level `li`
// (Engine/mom25d/solver.py:1101-1125):
//   phi(pt) = sum_t [ sum_k S_W(|pt - r_k|) w_k + cW I0(t, pt) / (4 pi) ] q_t / eps0
// One workgroup per point; threads stride over the triangles.  The S_W
// spline has 8 doubles per interval (c0..c3, re / im) at c_off / 2.
constexpr int kMaxRhs = 4;

__global__ __launch_bounds__(256) void line_voltage_kernel(int npts, const double* __restrict__ pts, int li, int nl,
                                                           int nt, int nrhs, const int32_t* __restrict__ tri_layer,
                                                           const double* __restrict__ V,
                                                           const double* __restrict__ qp7,
                                                           const double* __restrict__ qw7,
                                                           const PairTab* __restrict__ pairs,
                                                           const double* __restrict__ tab_x,
                                                           const double* __restrict__ tab_cw,
                                                           const Bucket* __restrict__ tab_b,
                                                           const double* __restrict__ q, double inv_eps0,
                                                           double* __restrict__ out) {
  const int m = blockIdx.x, tid = threadIdx.x;
  const double rx = pts[2 * m], ry = pts[2 * m + 1];
  z64 acc[kMaxRhs];
#pragma unroll
  for (int c = 0; c < kMaxRhs; ++c) acc[c] = {0.0, 0.0};
  const z64* Q = reinterpret_cast<const z64*>(q);
  for (int t = tid; t < nt; t += 256) {
    const PairTab pt = pairs[li * nl + tri_layer[t]];
    if (pt.zero) continue;
    const double* x = tab_x + pt.x_off;
    const double* cw = tab_cw + pt.c_off / 2;
    const Bucket* bk = tab_b + pt.b_off;
    z64 k = {0.0, 0.0};
    for (int l = 0; l < 7; ++l) {
      const double dx = rx - qp7[14 * (int64_t)t + 2 * l], dy = ry - qp7[14 * (int64_t)t + 2 * l + 1];
      const double R = sqrt(dx * dx + dy * dy);
      const int i = spline_interval(bk, pt.nbucket, pt.inv_dx, R);
      const double d = R - x[i];
      const double* c = cw + (int64_t)i * 8;
      double z = d;
      double sr = c[6] + c[4] * z, si = c[7] + c[5] * z;
      z *= d;
      sr += c[2] * z; si += c[3] * z;
      z *= d;
      sr += c[0] * z; si += c[1] * z;
      const double w = qw7[7 * (int64_t)t + l];
      k += z64{sr * w, si * w};
    }
    if (pt.cW_re != 0.0 || pt.cW_im != 0.0) {
      double I0, I1x, I1y;
      wilton(V + 6 * (int64_t)t, rx, ry, fabs(pt.dz), I0, I1x, I1y);
      k += z64{pt.cW_re, pt.cW_im} * (I0 / kFourPi);
    }
    for (int c = 0; c < nrhs; ++c) acc[c] += k * Q[(int64_t)t * nrhs + c];
  }
  __shared__ z64 part[4][kMaxRhs];
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
    z64 s = part[0][tid];
    s += part[1][tid];
    s += part[2][tid];
    s += part[3][tid];
    reinterpret_cast<z64*>(out)[(int64_t)m * nrhs + tid] = s * inv_eps0;
  }
}

}  // namespace mom

using namespace mom;

template <int TIER>
static int launch_fill(int n_items, const FillArgs& a, hipStream_t s) {
  const size_t lds = (size_t)a.acc_bytes;
  if (int rc = reserve_lds(reinterpret_cast<const void*>(fill_kernel<TIER>), lds, "fill_kernel")) return rc;
  hipLaunchKernelGGL(fill_kernel<TIER>, dim3(n_items), dim3(FillShape<TIER>::threads), lds, s, a);
  return check_launch("fill_kernel");
}

extern "C" int mom_fill_tier(int tier, int n_items, const void* args, void* stream) {
  if (n_items <= 0) return n_items == 0 ? 0 : fail("mom_fill_tier: negative item count");
  if (!args) return fail("mom_fill_tier: null argument block");
  const FillArgs a = *reinterpret_cast<const FillArgs*>(args);
  if (!a.Z || !a.items || !a.src || !a.pairs || !a.blk_row_off || !a.blk_rows || !a.blk_rloc || !a.item_cols ||
      !a.src_cloc)
    return fail("mom_fill_tier: null device pointer");
  // LDS: the accumulator plus about 18 KB of block data and staged spline,
  // within the 160 KB of a CU
  if (a.acc_bytes < 0 || a.acc_bytes > 136 * 1024) return fail("mom_fill_tier: LDS accumulator out of range");
  if ((uintptr_t)a.Z % 16) return fail("mom_fill_tier: Z must be 16-byte aligned (complex128)");
  if (!a.item_r) return fail("mom_fill_tier: null item bounds");
  if (tier == 0 && (!a.near_static || !a.item_soff)) return fail("mom_fill_tier: near tier without its statics");
  hipStream_t s = (hipStream_t)stream;
  switch (tier) {
    case 0: return launch_fill<0>(n_items, a, s);
    case 1: return launch_fill<1>(n_items, a, s);
    case 2: return launch_fill<2>(n_items, a, s);
    default: return fail("mom_fill_tier: tier must be 0 (near), 1 (mid) or 2 (far)");
  }
}

extern "C" int mom_near_statics(int n_items, const void* args, const double* pair_dz, double* out, void* stream) {
  if (n_items <= 0) return n_items == 0 ? 0 : fail("mom_near_statics: negative item count");
  if (!args) return fail("mom_near_statics: null argument block");
  FillArgs a = *reinterpret_cast<const FillArgs*>(args);
  if (!a.items || !a.src || !a.V || !a.qp7 || !a.blk_tris || !a.item_soff || !pair_dz || !out)
    return fail("mom_near_statics: null device pointer");
  hipLaunchKernelGGL(near_static_kernel, dim3(n_items), dim3(kStaticThreads), 0, (hipStream_t)stream, a, pair_dz,
                     out);
  return check_launch("near_static_kernel");
}

extern "C" int mom_fill_args_size(void) { return (int)sizeof(FillArgs); }
extern "C" int mom_pair_tab_size(void) { return (int)sizeof(PairTab); }

extern "C" int mom_gram_add(int nt, const double* g, const int32_t* tb, double zs_re, double zs_im, double* Z,
                            int64_t ldz, void* stream) {
  if (nt <= 0) return 0;
  hipLaunchKernelGGL(gram_kernel, dim3((nt + 255) / 256), dim3(256), 0, (hipStream_t)stream, nt, g, tb, zs_re,
                     zs_im, Z, ldz);
  return check_launch("gram_kernel");
}

extern "C" int mom_via_blocks(int nb, int nv, int nt, const double* IVpad, const int32_t* d_off,
                              const int32_t* d_pad, const double* d_val, const int32_t* pad_tri, const int32_t* tp,
                              const int32_t* vp, const int32_t* tm, const int32_t* vm, const double* cV,
                              const double* E, double ijwe_re, double ijwe_im, double* Z, int64_t ldz,
                              void* stream) {
  if (nv <= 0) return 0;
  if (nv > 65535) return fail("mom_via_blocks: more than 65535 via segments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(via_rows_kernel, dim3((nb + 255) / 256, nv), dim3(256), 0, s, nb, nv, nt, IVpad, d_off,
                     d_pad, d_val, tp, vp, tm, vm, cV, ijwe_re, ijwe_im, Z, ldz);
  int rc = check_launch("via_rows_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(via_block_kernel, dim3((nv + 63) / 64, nv), dim3(64), 0, s, nb, nv, nt, IVpad, d_off, d_pad,
                     d_val, pad_tri, E, ijwe_re, ijwe_im, Z, ldz);
  return check_launch("via_block_kernel");
}

extern "C" int mom_symmetrize(int n, double* Z, int64_t ldz, float* Z32, int64_t ld32, const double* d,
                              void* stream) {
  if (n <= 0) return 0;
  int T = (n + 31) / 32;
  if (T > 65535) return fail("mom_symmetrize: matrix too large");
  hipLaunchKernelGGL(symmetrize_kernel, dim3(T, T), dim3(256), 0, (hipStream_t)stream, n, Z, ldz, Z32, ld32, d);
  return check_launch("symmetrize_kernel");
}

extern "C" int mom_diag_scale(int n, const double* Z, int64_t ldz, double* d, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(diag_scale_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, Z, ldz, d);
  return check_launch("diag_scale_kernel");
}

extern "C" int mom_line_voltage(int npts, const double* pts, int li, int nl, int nt, int nrhs,
                                const int32_t* tri_layer, const double* V, const double* qp7, const double* qw7,
                                const void* pairs, const double* tab_x, const double* tab_cw, const void* tab_b,
                                const double* q, double inv_eps0, double* out, void* stream) {
  if (npts <= 0) return 0;
  if (nrhs < 1 || nrhs > kMaxRhs) return fail("mom_line_voltage: 1..4 right-hand sides");
  hipLaunchKernelGGL(line_voltage_kernel, dim3(npts), dim3(256), 0, (hipStream_t)stream, npts, pts, li, nl, nt, nrhs,
                     tri_layer, V, qp7, qw7, (const PairTab*)pairs, tab_x, tab_cw, (const Bucket*)tab_b, q, inv_eps0,
                     out);
  return check_launch("line_voltage_kernel");
}
