// ice3_mlp.hpp -- the per-point MLP of the rain_ice surrogate on CDNA4 matrix cores.
//
// Network: 18 features (padded to MLP_IN = 32) -> 4 hidden layers of 128 units with
// SiLU -> 8 outputs (padded to MLP_OUT = 32).  Weights are bf16 on the matrix cores
// with fp32 accumulation (v_mfma_f32_16x16x32_bf16); master weights, biases and the
// optimiser state are fp32.
//
// Layer l multiplies a (points x K_l) activation tile by W_l (K_l x N_l, row-major).
// Matrix-core operand layouts of v_mfma_f32_16x16x32_bf16 (wave64):
//   A (16 x 32): lane l holds A[l % 16][8*(l/16) + i], i = 0..7
//   B (32 x 16): lane l holds B[8*(l/16) + i][l % 16], i = 0..7
//   D (16 x 16): lane l holds D[4*(l/16) + i][l % 16], i = 0..3
// The forward pass reads W_l as pre-packed B fragments (64 lanes x 8 bf16 per 32x16
// block, so each lane loads 16 contiguous bytes); the backward pass reads the
// row-major copy.
#pragma once
#include <hip/hip_runtime.h>

namespace ice3 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MLP_NFEAT = 18;   // features used
constexpr int MLP_NTARG = 8;    // outputs used
constexpr int MLP_IN = 32;      // padded input width
constexpr int MLP_H = 128;      // hidden width
constexpr int MLP_OUT = 32;     // padded output width
constexpr int MLP_LAYERS = 5;
constexpr int MLP_NW = MLP_IN * MLP_H + 3 * MLP_H * MLP_H + MLP_H * MLP_OUT;  // 57344
constexpr int MLP_NB = 4 * MLP_H + MLP_OUT;                                     // 544
constexpr int MLP_NP = MLP_NW + MLP_NB;  // parameters: weights then biases

constexpr int MLP_TILE = 128;                          // points per workgroup tile
constexpr int MLP_WAVES = 4;                           // 256-thread workgroups
constexpr int MLP_MT = MLP_TILE / (16 * MLP_WAVES);    // 16-row m-tiles per wave (2)
constexpr int MLP_ROWS = MLP_TILE / MLP_WAVES;         // rows owned by one wave (32)
constexpr int MLP_HS = MLP_H + 8;                      // padded LDS row strides (bf16)
constexpr int MLP_XS = MLP_IN + 8;
constexpr int MLP_YS = MLP_OUT + 4;                    // fp32 output rows

__host__ __device__ constexpr int mlp_k(int l) { return l == 0 ? MLP_IN : MLP_H; }
__host__ __device__ constexpr int mlp_n(int l) { return l == MLP_LAYERS - 1 ? MLP_OUT : MLP_H; }
__host__ __device__ constexpr int mlp_woff(int l) {
  return l == 0 ? 0 : MLP_IN * MLP_H + (l - 1) * MLP_H * MLP_H;
}
__host__ __device__ constexpr int mlp_boff(int l) { return MLP_NW + l * MLP_H; }

// Offset of row-major element (k, n) of layer l in the packed-fragment copy.
__host__ __device__ inline int mlp_pack_index(int l, int k, int n) {
  const int N = mlp_n(l);
  const int kk = k / 32, kr = k % 32, nn = n / 16, lane = (kr / 8) * 16 + n % 16, i = kr % 8;
  return mlp_woff(l) + ((kk * (N / 16) + nn) * 64 + lane) * 8 + i;
}

__device__ __forceinline__ f32x4 mfma_bf16(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// SiLU and its derivative from the hardware exp2 and reciprocal (about 1 ulp in fp32;
// the activations are rounded to bf16 afterwards).
__device__ __forceinline__ float sigmoid_fast(float z) {
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.44269504f * z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid_fast(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid_fast(z);
  return s * (1.f + z * (1.f - s));
}

// acc[mt][nn] = In[rows of this wave] (MLP_ROWS x K) * W (K x N), In row-major with row
// stride S (bf16) in LDS, W as packed fragments (LDS).
template <int K, int N, int S>
__device__ __forceinline__ void mlp_layer(const __bf16 *sIn, int row0, const bf16x8 *sWpk,
                                          f32x4 (&acc)[MLP_MT][N / 16], int lane) {
#pragma unroll
  for (int mt = 0; mt < MLP_MT; ++mt)
#pragma unroll
    for (int nn = 0; nn < N / 16; ++nn) acc[mt][nn] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < K / 32; ++kk) {
    bf16x8 a[MLP_MT];
#pragma unroll
    for (int mt = 0; mt < MLP_MT; ++mt)
      a[mt] = *(const bf16x8 *)(sIn + (row0 + mt * 16 + lane % 16) * S + kk * 32 + 8 * (lane / 16));
#pragma unroll
    for (int nn = 0; nn < N / 16; ++nn) {
      const bf16x8 b = sWpk[(kk * (N / 16) + nn) * 64 + lane];
#pragma unroll
      for (int mt = 0; mt < MLP_MT; ++mt) acc[mt][nn] = mfma_bf16(a[mt], b, acc[mt][nn]);
    }
  }
}

// Copy n elements global -> LDS with the whole workgroup (16-byte granules).
__device__ __forceinline__ void block_copy16(void *dst, const void *src, int bytes) {
  const int4 *s = (const int4 *)src;
  int4 *d = (int4 *)dst;
  for (int q = threadIdx.x; q < bytes / 16; q += blockDim.x) d[q] = s[q];
}

// Launch the inference kernel k_mlp_infer (ice3_mlp.hip): y[npts][MLP_NTARG] from the bf16
// inputs x[npts][MLP_IN], the packed bf16 weights wpk and the fp32 parameters (biases), on
// nblk persistent workgroups of 256 threads.
hipError_t mlp_infer_launch(const void *wpk, const float *params, const void *x, float *y,
                            unsigned int npts, unsigned int nblk, hipStream_t s);

}  // namespace ice3
