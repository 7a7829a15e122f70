// Surrogate of the smooth layered-Green remainder: an MLP
//   6 -> 128 -> 128 -> 128 -> 4, SiLU after the three hidden layers,
// trained and evaluated on the GPU.
//
// Training step (mlp_kernel<true>): every workgroup takes one tile of 64
// samples (16 per wavefront), runs the forward pass, the squared-error
// gradient and the backward pass, and writes the tile's gradient as one
// partial gradient vector.
// mom_mlp_adam then sums the partials (fixed order: deterministic) and
// applies Adam to the FP32 master weights, refreshing the BF16 copies.
//
// Arithmetic: the two 128 x 128 layers run on the matrix cores
// (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation), the 6-wide
// input layer and the 4-wide output layer in FP32 on the vector ALUs.
// Activations are kept in LDS as BF16 pre-activations z (the MFMA operand
// silu(z) is formed from them when loaded); gradients flow back as BF16 dz
// written over z.  mom_mlp_forward_f32 evaluates the same network in FP32
// throughout (no BF16 rounding), for table generation.
//
// Parameter vector (FP32, row-major): W1[128][6] b1 W2[128][128] b2
// W3[128][128] b3 W4[4][128] b4 (34 436 values); BF16 copies of W2 and W3
// (row-major [out][in]).

#include "common.h"

namespace mom {
namespace mlp {

constexpr int DIN = 6, DH = 128, DOUT = 4;
constexpr int TS = 64;          // samples per tile, 16 per wavefront
constexpr int LDH = DH + 8;     // LDS row stride (elements) of BF16 rows
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + DH * DIN;
constexpr int OFF_W2 = OFF_B1 + DH;
constexpr int OFF_B2 = OFF_W2 + DH * DH;
constexpr int OFF_W3 = OFF_B2 + DH;
constexpr int OFF_B3 = OFF_W3 + DH * DH;
constexpr int OFF_W4 = OFF_B3 + DH;
constexpr int OFF_B4 = OFF_W4 + DOUT * DH;
constexpr int NPARAM = OFF_B4 + DOUT;   // 34436
constexpr int PART = NPARAM + 4;        // partial-gradient slot: parameters, loss, padding

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (inputs are finite).
__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct Smem {
  uint16_t W[2][DH][LDH];   // W2, W3 (BF16, [out][in])
  uint16_t Z[3][TS][LDH];   // pre-activations of the hidden layers, then their gradients
  float W1[DH][DIN];
  float b[3][DH];
  float W4[DOUT][DH];
  float b4[DOUT];
  float X[TS][DIN];
  float Y[TS][DOUT];        // targets, then d loss / d output
  float O[TS][DOUT];        // outputs
  float red[4];
};

struct Args {
  int n;                     // samples
  const float* X;            // [n][6]
  const float* Y;            // [n][4] targets (training)
  const float* wout;         // [4] loss weight of each output
  float grad_scale;          // d loss / d out = grad_scale * wout * (out - y)
  const float* P;            // FP32 parameters
  const uint16_t* Wbf;       // BF16 W2, W3
  float* out;                // [n][4] outputs (forward) or null
  float* part;               // [gridDim.x][PART] partial gradients (training)
};

// One 128 -> 128 layer for the 16 samples of this wavefront:
// Z[dst][s][:] = bf16(W silu(Z[src][s][:]) + b).
__device__ __forceinline__ void hidden_forward(Smem& sm, int src, int dst, int wsel, int wave, int lane) {
  f32x4 acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) {
    const uint16_t* zr = &sm.Z[src][16 * wave + r][16 * kk + 4 * q];
    bf16x4 a;
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = (short)f2bf(silu(bf2f(zr[i])));
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
      const bf16x4 bw = *reinterpret_cast<const bf16x4*>(&sm.W[wsel][16 * nt + r][16 * kk + 4 * q]);
      acc[nt] = mfma(a, bw, acc[nt]);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 8; ++nt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int s = 16 * wave + 4 * q + v, j = 16 * nt + r;
      sm.Z[dst][s][j] = f2bf(acc[nt][v] + sm.b[dst][j]);
    }
}

// Wavefront part of the backward pass through one 128 -> 128 layer, for the
// 16 samples of this wavefront: dz_in[s][j] = (sum_n dZ[s][n] W[n][j])
// silu'(z_in[s][j]), written over z_in (each element is read and written
// by the same lane).
__device__ __forceinline__ void hidden_backward_rows(Smem& sm, int dz, int zin, int wsel, int wave, int lane) {
  const int r = lane & 15, q = lane >> 4;
  f32x4 g[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) g[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(&sm.Z[dz][16 * wave + r][16 * kk + 4 * q]);
#pragma unroll
    for (int ct = 0; ct < 8; ++ct) {
      bf16x4 bw;
#pragma unroll
      for (int i = 0; i < 4; ++i) bw[i] = (short)sm.W[wsel][16 * kk + 4 * q + i][16 * ct + r];
      g[ct] = mfma(a, bw, g[ct]);
    }
  }
#pragma unroll
  for (int ct = 0; ct < 8; ++ct)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uint16_t& z = sm.Z[zin][16 * wave + 4 * q + v][16 * ct + r];
      z = f2bf(g[ct][v] * dsilu(bf2f(z)));
    }
}

// Weight gradient of one 128 x 128 layer over the tile:
// gw[n][j] += sum_s dZ[s][n] silu(z_in[s][j]); wavefront w owns rows
// n in [32 w, 32 w + 32) (2 x 8 MFMA tiles).
__device__ __forceinline__ void hidden_weight_grad(const Smem& sm, int dz, int zin, int wave, int lane,
                                                   f32x4 (&gw)[2][8]) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int ks = 0; ks < TS / 16; ++ks) {
    bf16x4 a[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) a[mt][i] = (short)sm.Z[dz][16 * ks + 4 * q + i][32 * wave + 16 * mt + r];
#pragma unroll
    for (int ct = 0; ct < 8; ++ct) {
      bf16x4 bh;
#pragma unroll
      for (int i = 0; i < 4; ++i) bh[i] = (short)f2bf(silu(bf2f(sm.Z[zin][16 * ks + 4 * q + i][16 * ct + r])));
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) gw[mt][ct] = mfma(a[mt], bh, gw[mt][ct]);
    }
  }
}

template <bool TRAIN>
__global__ __launch_bounds__(256) void mlp_kernel(Args a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem& sm = *reinterpret_cast<Smem*>(smem_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  // weights into LDS
  for (int e = tid; e < 2 * DH * DH; e += 256) {
    const int l = e / (DH * DH), rem = e % (DH * DH);
    sm.W[l][rem / DH][rem % DH] = a.Wbf[e];
  }
  for (int e = tid; e < DH * DIN; e += 256) sm.W1[e / DIN][e % DIN] = a.P[OFF_W1 + e];
  for (int e = tid; e < DH; e += 256) {
    sm.b[0][e] = a.P[OFF_B1 + e];
    sm.b[1][e] = a.P[OFF_B2 + e];
    sm.b[2][e] = a.P[OFF_B3 + e];
  }
  for (int e = tid; e < DOUT * DH; e += 256) sm.W4[e / DH][e % DH] = a.P[OFF_W4 + e];
  if (tid < DOUT) sm.b4[tid] = a.P[OFF_B4 + tid];

  const int ntile = (a.n + TS - 1) / TS;
  // Training: one tile per workgroup (gridDim.x == ntile), partial gradient
  // of the tile in slot blockIdx.x.  Forward: tiles strided over the grid.
  for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const int s0 = tile * TS;
    __syncthreads();  // weights loaded / previous tile finished with the LDS
    for (int e = tid; e < TS * DIN; e += 256) {
      const int s = e / DIN;
      sm.X[s][e % DIN] = s0 + s < a.n ? a.X[(int64_t)s0 * DIN + e] : 0.f;
    }
    if (TRAIN)
      for (int e = tid; e < TS * DOUT; e += 256) {
        const int s = e / DOUT;
        sm.Y[s][e % DOUT] = s0 + s < a.n ? a.Y[(int64_t)s0 * DOUT + e] : 0.f;
      }
    __syncthreads();
    // ---- forward (each wavefront its 16 samples)
    for (int e = lane; e < 16 * DH; e += 64) {
      const int s = 16 * wave + e / DH, j = e % DH;
      float z = sm.b[0][j];
#pragma unroll
      for (int k = 0; k < DIN; ++k) z += sm.X[s][k] * sm.W1[j][k];
      sm.Z[0][s][j] = f2bf(z);
    }
    __syncthreads();
    hidden_forward(sm, 0, 1, 0, wave, lane);
    __syncthreads();
    hidden_forward(sm, 1, 2, 1, wave, lane);
    __syncthreads();
    {
      const int s = 16 * wave + (lane >> 2), o = lane & 3;
      float y = sm.b4[o];
      for (int k = 0; k < DH; ++k) y += sm.W4[o][k] * silu(bf2f(sm.Z[2][s][k]));
      sm.O[s][o] = y;
      if (!TRAIN && s0 + s < a.n) a.out[(int64_t)(s0 + s) * DOUT + o] = y;
    }
    if (!TRAIN) continue;
    float* part = a.part + (int64_t)blockIdx.x * PART;
    __syncthreads();
    // ---- d loss / d out
    float loss;
    {
      const int s = tid >> 2, o = tid & 3;
      const float d = s0 + s < a.n ? sm.O[s][o] - sm.Y[s][o] : 0.f;
      loss = a.wout[o] * d * d;
      sm.Y[s][o] = a.grad_scale * a.wout[o] * d;
    }
    __syncthreads();
    // ---- output layer: weight gradient, then dz3 over z3
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int p = tid + 256 * u, o = p / DH, k = p % DH;
      float acc = 0.f;
      for (int s = 0; s < TS; ++s) acc += sm.Y[s][o] * silu(bf2f(sm.Z[2][s][k]));
      part[OFF_W4 + p] = acc;
    }
    if (tid < DOUT) {
      float acc = 0.f;
      for (int s = 0; s < TS; ++s) acc += sm.Y[s][tid];
      part[OFF_B4 + tid] = acc;
    }
    __syncthreads();
    for (int e = lane; e < 16 * DH; e += 64) {
      const int s = 16 * wave + e / DH, k = e % DH;
      float dh = 0.f;
#pragma unroll
      for (int o = 0; o < DOUT; ++o) dh += sm.Y[s][o] * sm.W4[o][k];
      sm.Z[2][s][k] = f2bf(dh * dsilu(bf2f(sm.Z[2][s][k])));
    }
    __syncthreads();
    // ---- hidden layers: W3 (z2 -> z3), then W2 (z1 -> z2)
#pragma unroll
    for (int l = 1; l >= 0; --l) {
      const int dz = l + 1, zin = l;
      const int boff = l == 1 ? OFF_B3 : OFF_B2, woff = l == 1 ? OFF_W3 : OFF_W2;
      {
        f32x4 gw[2][8];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int ct = 0; ct < 8; ++ct) gw[mt][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
        hidden_weight_grad(sm, dz, zin, wave, lane, gw);
        const int r = lane & 15, q = lane >> 4;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int ct = 0; ct < 8; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              part[woff + (32 * wave + 16 * mt + 4 * q + v) * DH + 16 * ct + r] = gw[mt][ct][v];
      }
      if (tid < DH) {
        float acc = 0.f;
        for (int s = 0; s < TS; ++s) acc += bf2f(sm.Z[dz][s][tid]);
        part[boff + tid] = acc;
      }
      __syncthreads();  // every wavefront done reading z_in of all samples
      hidden_backward_rows(sm, dz, zin, l, wave, lane);
      __syncthreads();
    }
    // ---- input layer
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int p = tid + 256 * u, j = p / DIN, k = p % DIN;
      float acc = 0.f;
      for (int s = 0; s < TS; ++s) acc += bf2f(sm.Z[0][s][j]) * sm.X[s][k];
      part[OFF_W1 + p] = acc;
    }
    if (tid < DH) {
      float acc = 0.f;
      for (int s = 0; s < TS; ++s) acc += bf2f(sm.Z[0][s][tid]);
      part[OFF_B1 + tid] = acc;
    }
    // loss of the tile
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    if (lane == 0) sm.red[wave] = loss;
    __syncthreads();
    if (tid == 0) {
      part[NPARAM] = sm.red[0] + sm.red[1] + sm.red[2] + sm.red[3];
      part[NPARAM + 1] = part[NPARAM + 2] = part[NPARAM + 3] = 0.f;
    }
  }
}

// g = sum of the partials (fixed order), then Adam on the FP32 parameters and
// the BF16 copies of W2 / W3.  step_lr = lr sqrt(1 - b2^t) / (1 - b1^t).
__global__ void adam_kernel(int nslots, const float* __restrict__ part, float* __restrict__ P,
                            float* __restrict__ m1, float* __restrict__ m2, uint16_t* __restrict__ Wbf,
                            float* __restrict__ grad, float* __restrict__ loss, float b1, float b2, float eps,
                            float step_lr) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > NPARAM) return;
  float g = 0.f;
  for (int s = 0; s < nslots; ++s) g += part[(int64_t)s * PART + p];
  if (p == NPARAM) {
    if (loss) *loss = g;
    return;
  }
  if (grad) grad[p] = g;
  if (step_lr == 0.f) return;
  const float a = b1 * m1[p] + (1.f - b1) * g;
  const float v = b2 * m2[p] + (1.f - b2) * g * g;
  m1[p] = a;
  m2[p] = v;
  const float w = P[p] - step_lr * a / (sqrtf(v) + eps);
  P[p] = w;
  if (p >= OFF_W2 && p < OFF_B2) Wbf[p - OFF_W2] = f2bf(w);
  if (p >= OFF_W3 && p < OFF_B3) Wbf[DH * DH + p - OFF_W3] = f2bf(w);
}

// BF16 copies of W2 and W3 from the FP32 parameters.
__global__ void pack_bf16_kernel(const float* __restrict__ P, uint16_t* __restrict__ Wbf) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * DH * DH) return;
  Wbf[e] = f2bf(P[(e < DH * DH ? OFF_W2 : OFF_W3 - DH * DH) + e]);
}

// FP32 forward pass, no BF16 rounding: 32 samples per workgroup.
constexpr int TSF = 32;
__global__ __launch_bounds__(256) void forward_f32_kernel(int n, const float* __restrict__ X,
                                                          const float* __restrict__ P, float* __restrict__ out) {
  __shared__ float H[2][TSF][DH + 1];
  __shared__ float Xs[TSF][DIN];
  const int tid = threadIdx.x, s0 = blockIdx.x * TSF;
  for (int e = tid; e < TSF * DIN; e += 256) Xs[e / DIN][e % DIN] = s0 + e / DIN < n ? X[(int64_t)s0 * DIN + e] : 0.f;
  __syncthreads();
  for (int e = tid; e < TSF * DH; e += 256) {
    const int s = e / DH, j = e % DH;
    float z = P[OFF_B1 + j];
    for (int k = 0; k < DIN; ++k) z += Xs[s][k] * P[OFF_W1 + j * DIN + k];
    H[0][s][j] = silu(z);
  }
  __syncthreads();
  for (int l = 0; l < 2; ++l) {
    const float* W = P + (l == 0 ? OFF_W2 : OFF_W3);
    const float* bb = P + (l == 0 ? OFF_B2 : OFF_B3);
    for (int e = tid; e < TSF * DH; e += 256) {
      const int s = e / DH, j = e % DH;
      float z = bb[j];
      for (int k = 0; k < DH; ++k) z += H[l][s][k] * W[j * DH + k];
      H[l ^ 1][s][j] = silu(z);
    }
    __syncthreads();
  }
  for (int e = tid; e < TSF * DOUT; e += 256) {
    const int s = e / DOUT, o = e % DOUT;
    if (s0 + s >= n) continue;
    float y = P[OFF_B4 + o];
    for (int k = 0; k < DH; ++k) y += H[0][s][k] * P[OFF_W4 + o * DH + k];
    out[(int64_t)(s0 + s) * DOUT + o] = y;
  }
}

static int launch(bool train, int grid, const Args& a, hipStream_t s) {
  const size_t lds = sizeof(Smem);
  if (int rc = mom::reserve_lds(train ? reinterpret_cast<const void*>(mlp_kernel<true>)
                                      : reinterpret_cast<const void*>(mlp_kernel<false>),
                                lds, "mlp_kernel"))
    return rc;
  if (train)
    hipLaunchKernelGGL(mlp_kernel<true>, dim3(grid), dim3(256), lds, s, a);
  else
    hipLaunchKernelGGL(mlp_kernel<false>, dim3(grid), dim3(256), lds, s, a);
  return check_launch("mlp_kernel");
}

}  // namespace mlp
}  // namespace mom

using namespace mom;

extern "C" int mom_mlp_param_count() { return mlp::NPARAM; }
extern "C" int mom_mlp_partial_stride() { return mlp::PART; }

extern "C" int mom_mlp_pack(const float* P, uint16_t* Wbf, void* stream) {
  if (!P || !Wbf) return fail("mom_mlp_pack: null device pointer");
  hipLaunchKernelGGL(mlp::pack_bf16_kernel, dim3((2 * mlp::DH * mlp::DH + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, P, Wbf);
  return check_launch("pack_bf16_kernel");
}

extern "C" int mom_mlp_forward(int precision, int n, const float* X, const float* P, const uint16_t* Wbf,
                               float* out, void* stream) {
  if (n <= 0) return 0;
  if (!X || !P || !out) return fail("mom_mlp_forward: null device pointer");
  hipStream_t s = (hipStream_t)stream;
  const int ntile = (n + mlp::TS - 1) / mlp::TS;
  if (precision == 32) {
    hipLaunchKernelGGL(mlp::forward_f32_kernel, dim3((n + mlp::TSF - 1) / mlp::TSF), dim3(256), 0, s, n, X, P, out);
    return check_launch("forward_f32_kernel");
  }
  if (precision != 16) return fail("mom_mlp_forward: precision must be 16 (BF16 matrix cores) or 32 (FP32)");
  if (!Wbf) return fail("mom_mlp_forward: null BF16 weights");
  mlp::Args a{n, X, nullptr, nullptr, 0.f, P, Wbf, out, nullptr};
  return mlp::launch(false, ntile < 1024 ? ntile : 1024, a, s);
}

// One partial-gradient slot per 64-sample tile: nslots = ceil(n / 64).
extern "C" int mom_mlp_train_step(int n, const float* X, const float* Y, const float* wout, float grad_scale,
                                  const float* P, const uint16_t* Wbf, float* part, int nslots, void* stream) {
  if (n <= 0) return fail("mom_mlp_train_step: empty batch");
  if (!X || !Y || !wout || !P || !Wbf || !part) return fail("mom_mlp_train_step: null device pointer");
  if (nslots != (n + mlp::TS - 1) / mlp::TS) return fail("mom_mlp_train_step: nslots must be ceil(n / 64)");
  mlp::Args a{n, X, Y, wout, grad_scale, P, Wbf, nullptr, part};
  return mlp::launch(true, nslots, a, (hipStream_t)stream);
}

extern "C" int mom_mlp_adam(int nslots, const float* part, float* P, float* m1, float* m2, uint16_t* Wbf,
                            float* grad, float* loss, float beta1, float beta2, float eps, float step_lr,
                            void* stream) {
  if (!part || !P || !Wbf || (step_lr != 0.f && (!m1 || !m2))) return fail("mom_mlp_adam: null device pointer");
  hipLaunchKernelGGL(mlp::adam_kernel, dim3((mlp::NPARAM + 1 + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     nslots, part, P, m1, m2, Wbf, grad, loss, beta1, beta2, eps, step_lr);
  return check_launch("adam_kernel");
}
