// ice3_mlp.hip -- kernels of the rain_ice surrogate MLP (bf16 matrix cores, fp32
// accumulation): training, with its C ABI (declared in include/ice3.h), and inference.
//
//   k_mlp_forward   one 128-point tile per workgroup: 5 layers on v_mfma_f32_16x16x32_bf16
//                   with every weight resident in LDS; saves pre-activations (fp32) and
//                   activations (bf16) for the backward pass; MSE loss and its gradient.
//   k_mlp_backward  one tile per workgroup: dW = H^T dZ and dH = dZ W^T on the matrix
//                   cores, SiLU derivative, bias gradients; one fp32 partial gradient per
//                   workgroup (no atomics: the reduction is deterministic).
//   k_mlp_reduce    sum of the partial gradients.
//   k_mlp_adam      Adam on the fp32 master parameters; rewrites the bf16 row-major and
//                   packed-fragment copies the matrix cores read.
//   k_mlp_infer     inference over all points of the surrogate step: persistent grid,
//                   weights resident in LDS (launched by mlp_infer_launch).
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/ice3.h"
#include "ice3_mlp.hpp"

namespace ice3 {

struct MlpFwdArgs {
  int npts;
  const bf16x8 *wpk;      // packed bf16 weights
  const float *params;    // fp32 master parameters (biases at MLP_NW)
  const __bf16 *x;        // [npts][MLP_IN]
  const float *target;    // [npts][MLP_NTARG] or NULL
  float *y;               // [npts][MLP_NTARG] or NULL
  float *zsave;           // [4][npts][MLP_H] or NULL
  __bf16 *hsave;          // [4][npts][MLP_H] or NULL
  float *dy;              // [npts][MLP_OUT] or NULL
  float *loss_part;       // [gridDim.x] or NULL
  float inv_count;        // 1 / (npts * MLP_NTARG)
};

// bias + activation epilogue of a hidden layer: z -> (save z, h) -> h into sOut
template <int N>
__device__ __forceinline__ void hidden_epilogue(const MlpFwdArgs &a, int l, int p0, int row0,
                                                const f32x4 (&acc)[MLP_MT][N / 16],
                                                const float *sB, __bf16 *sOut, int lane) {
#pragma unroll
  for (int mt = 0; mt < MLP_MT; ++mt)
#pragma unroll
    for (int nn = 0; nn < N / 16; ++nn)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = row0 + mt * 16 + 4 * (lane / 16) + i, col = nn * 16 + lane % 16;
        const float z = acc[mt][nn][i] + sB[l * MLP_H + col];
        const __bf16 h = (__bf16)silu(z);
        sOut[row * MLP_HS + col] = h;
        const int p = p0 + row;
        if (a.zsave && p < a.npts) {
          const size_t o = ((size_t)l * a.npts + p) * MLP_H + col;
          a.zsave[o] = z;
          a.hsave[o] = h;
        }
      }
}

__global__ void __launch_bounds__(256) k_mlp_forward(MlpFwdArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  bf16x8 *sW = (bf16x8 *)smem;                                   // MLP_NW bf16
  float *sB = (float *)(smem + MLP_NW * 2);                      // MLP_NB fp32
  __bf16 *sH = (__bf16 *)(smem + MLP_NW * 2 + MLP_NB * 4);       // [MLP_TILE][MLP_HS]
  __bf16 *sX = sH + MLP_TILE * MLP_HS;                           // [MLP_TILE][MLP_XS]
  __shared__ float sLoss[MLP_WAVES];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64, row0 = w * MLP_ROWS;
  const int p0 = blockIdx.x * MLP_TILE;
  block_copy16(sW, a.wpk, MLP_NW * 2);
  for (int q = threadIdx.x; q < MLP_NB; q += blockDim.x) sB[q] = a.params[MLP_NW + q];
  // input tile: 4 x 16 bytes per row
  for (int q = threadIdx.x; q < MLP_TILE * 4; q += blockDim.x) {
    const int r = q / 4, c = q % 4, p = p0 + r;
    int4 v = {0, 0, 0, 0};
    if (p < a.npts) v = ((const int4 *)(a.x + (size_t)p * MLP_IN))[c];
    *(int4 *)(sX + r * MLP_XS + c * 8) = v;
  }
  __syncthreads();
  {
    f32x4 acc[MLP_MT][MLP_H / 16];
    mlp_layer<MLP_IN, MLP_H, MLP_XS>(sX, row0, sW + mlp_woff(0) / 8, acc, lane);
    hidden_epilogue<MLP_H>(a, 0, p0, row0, acc, sB, sH, lane);
  }
#pragma unroll 1
  for (int l = 1; l <= 3; ++l) {
    f32x4 acc[MLP_MT][MLP_H / 16];
    mlp_layer<MLP_H, MLP_H, MLP_HS>(sH, row0, sW + mlp_woff(l) / 8, acc, lane);
    hidden_epilogue<MLP_H>(a, l, p0, row0, acc, sB, sH, lane);
  }
  f32x4 acc[MLP_MT][MLP_OUT / 16];
  mlp_layer<MLP_H, MLP_OUT, MLP_HS>(sH, row0, sW + mlp_woff(4) / 8, acc, lane);
  float loss = 0.f;
#pragma unroll
  for (int mt = 0; mt < MLP_MT; ++mt)
#pragma unroll
    for (int nn = 0; nn < MLP_OUT / 16; ++nn)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = row0 + mt * 16 + 4 * (lane / 16) + i, col = nn * 16 + lane % 16;
        const int p = p0 + row;
        if (p >= a.npts) continue;
        const float y = acc[mt][nn][i] + sB[4 * MLP_H + col];
        float g = 0.f;
        if (col < MLP_NTARG) {
          if (a.y) a.y[(size_t)p * MLP_NTARG + col] = y;
          if (a.target) {
            const float d = y - a.target[(size_t)p * MLP_NTARG + col];
            loss += d * d;
            g = 2.f * d * a.inv_count;
          }
        }
        if (a.dy) a.dy[(size_t)p * MLP_OUT + col] = g;
      }
  if (a.loss_part) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) loss += __shfl_xor(loss, m, 64);
    if (lane == 0) sLoss[w] = loss;
    __syncthreads();
    if (threadIdx.x == 0) a.loss_part[blockIdx.x] = sLoss[0] + sLoss[1] + sLoss[2] + sLoss[3];
  }
}

constexpr size_t MLP_FWD_LDS = (size_t)MLP_NW * 2 + MLP_NB * 4 + MLP_TILE * MLP_HS * 2 +
                               MLP_TILE * MLP_XS * 2;

struct MlpBwdArgs {
  int npts;
  const __bf16 *wrm;      // row-major bf16 weights
  const __bf16 *x;        // [npts][MLP_IN]
  const float *zsave;     // [4][npts][MLP_H]
  const __bf16 *hsave;    // [4][npts][MLP_H]
  const float *dy;        // [npts][MLP_OUT]
  float *gpart;           // [gridDim.x][MLP_NP]
};

constexpr int BW_S = MLP_H + 8;          // row stride of the backward LDS tiles (bf16)
constexpr int BW_TILE = MLP_H * BW_S;    // elements of one 128 x BW_S tile

// One layer of the backward pass.  On entry sDZ ([point][BW_S], row-major) and sDZT
// ([unit][BW_S], transposed) hold dZ_L in bf16 for the tile; on exit (L > 0) they hold
// dZ_{L-1}.
template <int L>
__device__ __forceinline__ void bwd_layer(const MlpBwdArgs &a, int p0, __bf16 *sW, __bf16 *sDZ,
                                          __bf16 *sDZT, __bf16 *sHT, int lane, int w) {
  constexpr int K = mlp_k(L), N = mlp_n(L);
  float *gp = a.gpart + (size_t)blockIdx.x * MLP_NP;
  // layer input, transposed: sHT[k][p]
  const __bf16 *hin = L == 0 ? a.x : a.hsave + (size_t)(L - 1) * a.npts * MLP_H;
  for (int q = threadIdx.x; q < MLP_TILE * (K / 8); q += blockDim.x) {
    const int r = q / (K / 8), c = q % (K / 8), p = p0 + r;
    bf16x8 v;
    if (p < a.npts) v = *(const bf16x8 *)(hin + (size_t)p * K + c * 8);
    else v = bf16x8{};
#pragma unroll
    for (int i = 0; i < 8; ++i) sHT[(c * 8 + i) * BW_S + r] = v[i];
  }
  if (L > 0)  // W_L row-major: sW[k][n]
    for (int q = threadIdx.x; q < K * (N / 8); q += blockDim.x) {
      const int k = q / (N / 8), c = q % (N / 8);
      *(bf16x8 *)(sW + k * BW_S + c * 8) = *(const bf16x8 *)(a.wrm + mlp_woff(L) + k * N + c * 8);
    }
  __syncthreads();
  // dW_L = H^T dZ (K x N): 16x16 output tiles spread over the waves
  for (int t = w; t < (K / 16) * (N / 16); t += MLP_WAVES) {
    const int m0 = (t / (N / 16)) * 16, n0 = (t % (N / 16)) * 16;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < MLP_TILE / 32; ++kk) {
      const bf16x8 av = *(const bf16x8 *)(sHT + (m0 + lane % 16) * BW_S + kk * 32 + 8 * (lane / 16));
      const bf16x8 bv = *(const bf16x8 *)(sDZT + (n0 + lane % 16) * BW_S + kk * 32 + 8 * (lane / 16));
      acc = mfma_bf16(av, bv, acc);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
      gp[mlp_woff(L) + (m0 + 4 * (lane / 16) + i) * N + n0 + lane % 16] = acc[i];
  }
  // bias gradient: column sums of dZ
  for (int n = threadIdx.x; n < N; n += blockDim.x) {
    float s = 0.f;
    for (int p = 0; p < MLP_TILE; ++p) s += (float)sDZT[n * BW_S + p];
    gp[mlp_boff(L) + n] = s;
  }
  if (L == 0) return;
  // dH = dZ W^T (points x K), then dZ_{L-1} = dH * SiLU'(Z_{L-1})
  const int row0 = w * MLP_ROWS;
  f32x4 acc[MLP_MT][K / 16];
#pragma unroll
  for (int mt = 0; mt < MLP_MT; ++mt)
#pragma unroll
    for (int nn = 0; nn < K / 16; ++nn) acc[mt][nn] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < N / 32; ++kk) {
    bf16x8 av[MLP_MT];
#pragma unroll
    for (int mt = 0; mt < MLP_MT; ++mt)
      av[mt] = *(const bf16x8 *)(sDZ + (row0 + mt * 16 + lane % 16) * BW_S + kk * 32 + 8 * (lane / 16));
#pragma unroll
    for (int nn = 0; nn < K / 16; ++nn) {
      const bf16x8 bv = *(const bf16x8 *)(sW + (nn * 16 + lane % 16) * BW_S + kk * 32 + 8 * (lane / 16));
#pragma unroll
      for (int mt = 0; mt < MLP_MT; ++mt) acc[mt][nn] = mfma_bf16(av[mt], bv, acc[mt][nn]);
    }
  }
  const float *zin = a.zsave + (size_t)(L - 1) * a.npts * MLP_H;
#pragma unroll
  for (int mt = 0; mt < MLP_MT; ++mt)
#pragma unroll
    for (int nn = 0; nn < K / 16; ++nn)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = row0 + mt * 16 + 4 * (lane / 16) + i, col = nn * 16 + lane % 16;
        const int p = p0 + row;
        acc[mt][nn][i] = p < a.npts ? acc[mt][nn][i] * silu_grad(zin[(size_t)p * MLP_H + col]) : 0.f;
      }
  __syncthreads();  // every wave is done reading sDZ, sDZT, sHT and sW
#pragma unroll
  for (int mt = 0; mt < MLP_MT; ++mt)
#pragma unroll
    for (int nn = 0; nn < K / 16; ++nn)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = row0 + mt * 16 + 4 * (lane / 16) + i, col = nn * 16 + lane % 16;
        const __bf16 v = (__bf16)acc[mt][nn][i];
        sDZ[row * BW_S + col] = v;
        sDZT[col * BW_S + row] = v;
      }
  __syncthreads();
}

__global__ void __launch_bounds__(256) k_mlp_backward(MlpBwdArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __bf16 *sW = (__bf16 *)smem;
  __bf16 *sDZ = sW + BW_TILE;
  __bf16 *sDZT = sDZ + BW_TILE;
  __bf16 *sHT = sDZT + BW_TILE;
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int p0 = blockIdx.x * MLP_TILE;
  // dZ_4 = dY (the output layer is linear)
  for (int q = threadIdx.x; q < MLP_TILE * MLP_OUT; q += blockDim.x) {
    const int r = q / MLP_OUT, c = q % MLP_OUT, p = p0 + r;
    const __bf16 v = (__bf16)(p < a.npts ? a.dy[(size_t)p * MLP_OUT + c] : 0.f);
    sDZ[r * BW_S + c] = v;
    sDZT[c * BW_S + r] = v;
  }
  __syncthreads();
  bwd_layer<4>(a, p0, sW, sDZ, sDZT, sHT, lane, w);
  bwd_layer<3>(a, p0, sW, sDZ, sDZT, sHT, lane, w);
  bwd_layer<2>(a, p0, sW, sDZ, sDZT, sHT, lane, w);
  bwd_layer<1>(a, p0, sW, sDZ, sDZT, sHT, lane, w);
  bwd_layer<0>(a, p0, sW, sDZ, sDZT, sHT, lane, w);
}

constexpr size_t MLP_BWD_LDS = (size_t)4 * BW_TILE * 2;

__global__ void __launch_bounds__(256) k_mlp_reduce(const float *gpart, int nblk, float *grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= MLP_NP) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += gpart[(size_t)b * MLP_NP + e];
  grad[e] = s;
}

// Layer, row and column of weight element e (row-major per layer).
__device__ __forceinline__ void weight_coords(int e, int &l, int &k, int &n) {
  l = 0;
#pragma unroll
  for (int q = 1; q < MLP_LAYERS; ++q)
    if (e >= mlp_woff(q)) l = q;
  const int r = e - mlp_woff(l);
  k = r / mlp_n(l);
  n = r % mlp_n(l);
}

__global__ void __launch_bounds__(256) k_mlp_adam(int step, float lr, float b1, float b2, float eps,
                                                  float *params, float *m, float *v,
                                                  const float *grad, __bf16 *wrm, __bf16 *wpk) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= MLP_NP) return;
  float p = params[e];
  if (grad) {
    const float g = grad[e];
    const float mm = b1 * m[e] + (1.f - b1) * g;
    const float vv = b2 * v[e] + (1.f - b2) * g * g;
    m[e] = mm;
    v[e] = vv;
    const float mh = mm / (1.f - __powf(b1, (float)step));
    const float vh = vv / (1.f - __powf(b2, (float)step));
    p -= lr * mh / (sqrtf(vh) + eps);
    params[e] = p;
  }
  if (e < MLP_NW) {
    int l, k, n;
    weight_coords(e, l, k, n);
    const __bf16 b = (__bf16)p;
    wrm[e] = b;
    wpk[mlp_pack_index(l, k, n)] = b;
  }
}

// Inference over all points (the surrogate step's network stage).  Persistent grid, all
// weights resident in LDS,
// 128-point tiles (32 rows per wave); layer-0 operands come straight from HBM as matrix-
// core fragments and the next tile's are fetched while the current one is computed.
__global__ void __launch_bounds__(256) k_mlp_infer(const bf16x8 *__restrict__ wpk,
                                                   const float *__restrict__ params,
                                                   const __bf16 *__restrict__ x,
                                                   float *__restrict__ y, unsigned int npts) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  bf16x8 *sW = (bf16x8 *)smem;
  float *sB = (float *)(smem + MLP_NW * 2);
  __bf16 *sH = (__bf16 *)(smem + MLP_NW * 2 + MLP_NB * 4);
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64, row0 = w * MLP_ROWS;
  block_copy16(sW, wpk, MLP_NW * 2);
  for (int q = threadIdx.x; q < MLP_NB; q += blockDim.x) sB[q] = params[MLP_NW + q];
  __syncthreads();
  const unsigned int ntiles = (npts + MLP_TILE - 1) / MLP_TILE;
  // layer-0 A fragments of a tile: rows row0 + mt*16 + lane%16, inputs 8*(lane/16)..+7
  auto load_a = [&](unsigned int tile, bf16x8 (&a0)[MLP_MT]) {
#pragma unroll
    for (int mt = 0; mt < MLP_MT; ++mt) {
      const unsigned int p = tile * MLP_TILE + row0 + mt * 16 + lane % 16;
      if (p < npts) a0[mt] = *(const bf16x8 *)(x + (size_t)p * MLP_IN + 8 * (lane / 16));
      else a0[mt] = bf16x8{};
    }
  };
  bf16x8 a0[MLP_MT];
  unsigned int tile = blockIdx.x;
  if (tile < ntiles) load_a(tile, a0);
  for (; tile < ntiles; tile += gridDim.x) {
    f32x4 acc[MLP_MT][MLP_H / 16];
#pragma unroll
    for (int mt = 0; mt < MLP_MT; ++mt)
#pragma unroll
      for (int nn = 0; nn < MLP_H / 16; ++nn) acc[mt][nn] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nn = 0; nn < MLP_H / 16; ++nn) {
      const bf16x8 b = sW[mlp_woff(0) / 8 + nn * 64 + lane];
#pragma unroll
      for (int mt = 0; mt < MLP_MT; ++mt) acc[mt][nn] = mfma_bf16(a0[mt], b, acc[mt][nn]);
    }
    if (tile + gridDim.x < ntiles) load_a(tile + gridDim.x, a0);  // prefetch the next tile
#pragma unroll
    for (int mt = 0; mt < MLP_MT; ++mt)
#pragma unroll
      for (int nn = 0; nn < MLP_H / 16; ++nn)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = row0 + mt * 16 + 4 * (lane / 16) + i, col = nn * 16 + lane % 16;
          sH[row * MLP_HS + col] = (__bf16)silu(acc[mt][nn][i] + sB[col]);
        }
#pragma unroll 1
    for (int l = 1; l <= 3; ++l) {
      mlp_layer<MLP_H, MLP_H, MLP_HS>(sH, row0, sW + mlp_woff(l) / 8, acc, lane);
#pragma unroll
      for (int mt = 0; mt < MLP_MT; ++mt)
#pragma unroll
        for (int nn = 0; nn < MLP_H / 16; ++nn)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = row0 + mt * 16 + 4 * (lane / 16) + i, col = nn * 16 + lane % 16;
            sH[row * MLP_HS + col] = (__bf16)silu(acc[mt][nn][i] + sB[l * MLP_H + col]);
          }
    }
    f32x4 out[MLP_MT][MLP_OUT / 16];
    mlp_layer<MLP_H, MLP_OUT, MLP_HS>(sH, row0, sW + mlp_woff(4) / 8, out, lane);
#pragma unroll
    for (int mt = 0; mt < MLP_MT; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const unsigned int p = tile * MLP_TILE + row0 + mt * 16 + 4 * (lane / 16) + i;
        const int col = lane % 16;
        if (col < MLP_NTARG && p < npts) y[(size_t)p * MLP_NTARG + col] = out[mt][0][i] + sB[4 * MLP_H + col];
      }
  }
}

constexpr size_t INFER_LDS = (size_t)MLP_NW * 2 + MLP_NB * 4 + MLP_TILE * MLP_HS * 2;

hipError_t mlp_infer_launch(const void *wpk, const float *params, const void *x, float *y,
                            unsigned int npts, unsigned int nblk, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_infer, dim3(nblk), dim3(256), INFER_LDS, s, (const bf16x8 *)wpk,
                     params, (const __bf16 *)x, y, npts);
  return hipGetLastError();
}

}  // namespace ice3

using namespace ice3;

#define HIPCHK(x)                                 \
  do {                                            \
    if ((x) != hipSuccess) return ICE3_EDEVICE;   \
  } while (0)

extern "C" int ice3_mlp_forward(int npts, const void *wpk, const float *params, const void *x,
                                const float *target, float *y, float *zsave, void *hsave,
                                float *dy, float *loss_part, void *stream) {
  if (npts <= 0 || !wpk || !params || !x) return ICE3_EINVAL;
  if ((zsave == nullptr) != (hsave == nullptr) || (zsave && !dy) || (loss_part && !target))
    return ICE3_EINVAL;
  MlpFwdArgs a;
  a.npts = npts; a.wpk = (const bf16x8 *)wpk; a.params = params; a.x = (const __bf16 *)x;
  a.target = target; a.y = y; a.zsave = zsave; a.hsave = (__bf16 *)hsave; a.dy = dy;
  a.loss_part = loss_part; a.inv_count = 1.f / ((float)npts * MLP_NTARG);
  const int nblk = (npts + MLP_TILE - 1) / MLP_TILE;
  hipLaunchKernelGGL(k_mlp_forward, dim3(nblk), dim3(256), MLP_FWD_LDS, (hipStream_t)stream, a);
  HIPCHK(hipGetLastError());
  return ICE3_OK;
}

extern "C" int ice3_mlp_backward(int npts, const void *wrm, const void *x, const float *zsave,
                                 const void *hsave, const float *dy, float *gpart, float *grad,
                                 void *stream) {
  if (npts <= 0 || !wrm || !x || !zsave || !hsave || !dy || !gpart || !grad) return ICE3_EINVAL;
  MlpBwdArgs a;
  a.npts = npts; a.wrm = (const __bf16 *)wrm; a.x = (const __bf16 *)x; a.zsave = zsave;
  a.hsave = (const __bf16 *)hsave; a.dy = dy; a.gpart = gpart;
  const int nblk = (npts + MLP_TILE - 1) / MLP_TILE;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mlp_backward, dim3(nblk), dim3(256), MLP_BWD_LDS, s, a);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NP + 255) / 256), dim3(256), 0, s,
                     (const float *)gpart, nblk, grad);
  HIPCHK(hipGetLastError());
  return ICE3_OK;
}

extern "C" int ice3_mlp_adam(int step, float lr, float beta1, float beta2, float eps,
                             float *params, float *m, float *v, const float *grad, void *wrm,
                             void *wpk, void *stream) {
  if (!params || !wrm || !wpk || (grad && (!m || !v || step < 1))) return ICE3_EINVAL;
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     step, lr, beta1, beta2, eps, params, m, v, grad, (__bf16 *)wrm,
                     (__bf16 *)wpk);
  HIPCHK(hipGetLastError());
  return ICE3_OK;
}

extern "C" int ice3_mlp_sizes(int *n_params, int *n_weights, int *tile, int *in_width,
                              int *out_width) {
  if (n_params) *n_params = MLP_NP;
  if (n_weights) *n_weights = MLP_NW;
  if (tile) *tile = MLP_TILE;
  if (in_width) *in_width = MLP_IN;
  if (out_width) *out_width = MLP_OUT;
  return ICE3_OK;
}
