// ice3_kernels.hip -- HIP kernels for gfx950 and the C ABI of libice3_hip.so
// (declared in include/ice3.h).
//
// Layout: every field is float64 (nlev, ld), columns contiguous, so a 64-lane
// wavefront that owns 64 consecutive columns reads one level of a field as a
// single 512-byte coalesced transaction.
//
// Kernels:
//   k_ice_adjust   one thread per (level, column) point, pointwise ICE_ADJUST;
//   k_adjust_sedim one thread per column: ICE_ADJUST and the sedimentation in one
//                  top-down walk (the FP64 fused step);
//   k_sedim<R>     one thread per column, walking the levels top-down: the
//                  statistical sedimentation flux recurrence, whose fluxes
//                  never leave registers, and the surface precipitation;
//   k_micro<R>     persistent, one workgroup per compute unit, lookup tables in LDS:
//                  the ICE4 time-split microphysics, total tendencies and negativity
//                  correction per point, with the points still being time-stepped
//                  compacted into full wavefronts (ballot, LDS parking);
//   k_sg_features, k_mlp_infer, k_sg_decode
//                  the learned surrogate of k_micro: features and range check per point;
//                  the MLP on the bf16 matrix cores (k_mlp_infer, in ice3_mlp.hip);
//                  decode, water and sign checks per point.  Points it rejects are
//                  queued for k_micro_segments, the plain FP64 microphysics over the queue.
// rain_ice is k_sedim then k_micro on one stream; the fused time step is
// k_ice_adjust, k_sedim, k_micro.  R is the arithmetic of the process rates
// and lookups (double, or float for the mixed path).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/ice3.h"
#include "ice3_params.h"
#include "ice3_physics.hpp"
#include "ice3_stages.hpp"
#include "ice3_mlp.hpp"

namespace ice3 {

// ------------------------------------------------------------------ kernels

__global__ void __launch_bounds__(256) k_ice_adjust(AdjustArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  adjust_at(a, i, (size_t)blockIdx.y * a.ld + i);
}

// ICE_ADJUST with the OCND2 branch: one thread per point, blockIdx.y the level (the layer
// thickness reads the heights of the level above).
__global__ void __launch_bounds__(256) k_ice_adjust_ocnd2(AdjustArgs a, Ocnd2Args x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  adjust_at<true>(a, i, (size_t)blockIdx.y * a.ld + i, (int)blockIdx.y, &x);
}

// The FP64 fused step's ICE_ADJUST and sedimentation in one column walk: the adjustment's
// memory traffic overlaps the sedimentation's FP64 arithmetic.
__global__ void __launch_bounds__(64) k_adjust_sedim(AdjustArgs aa, SedimArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  adjust_sedim_column<double>(aa, a, i);
}

template <class R, bool ZZZ = false>
__global__ void __launch_bounds__(256) k_sedim(SedimArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  sedim_column<R, ZZZ>(a, i);
}

// ---- hail (KRR=7), FP64: ICE_ADJUST with the hail heat capacity (one thread per point),
// the sedimentation with hail (one thread per column), the two in one column walk for the
// fused step; the microphysics is k_micro<double, false, true> (k_micro_points<double, false,
// true> when the context sets another sub-step cap).
__global__ void __launch_bounds__(256) k_ice_adjust_hail(AdjustArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  adjust_at<false, true>(a, i, (size_t)blockIdx.y * a.ld + i);
}

__global__ void __launch_bounds__(256) k_sedim_hail(SedimArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  sedim_column<double, false, true>(a, i);
}

__global__ void __launch_bounds__(64) k_adjust_sedim_hail(AdjustArgs aa, SedimArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  adjust_sedim_column<double, true>(aa, a, i);
}

// Sum of v over the 64 lanes of the wavefront (every lane gets the total).
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// ---- k_count_nonfinite: input validation (ice3_count_nonfinite).  blockIdx.y is the
// field; each thread takes one column (grid-stride) and walks the field's rows, so that a
// wavefront's loads are 64 consecutive doubles.  A value is non-finite when its exponent
// bits are all ones (NaN or an infinity).  Only a wavefront that found some adds to the
// field's counter.
struct CountArgs {
  const double *f[ICE3_MAX_CHECK_FIELDS];
  int rows[ICE3_MAX_CHECK_FIELDS];
  int ncol, ld;
  unsigned long long *counts;
};

__global__ void __launch_bounds__(256) k_count_nonfinite(CountArgs a) {
  const int q = blockIdx.y;
  const double *__restrict__ f = a.f[q];
  const int rows = a.rows[q];
  unsigned long long n = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.ncol; i += gridDim.x * blockDim.x)
#pragma unroll 8
    for (int r = 0; r < rows; ++r) {
      const unsigned long long b = __double_as_longlong(f[(size_t)r * a.ld + i]);
      n += (b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull;
    }
  if (__ballot(n != 0) == 0ull) return;  // uniform: every lane of the wavefront sees it
  n = wave_sum(n);
  if ((threadIdx.x & 63) == 0) atomicAdd(&a.counts[q], n);
}

// ---- k_sedim_split: ICE4_SEDIMENTATION with CSEDIM='SPLI' (micro/mode_ice4_sedimentation.F90:
// 196-219, micro/mode_ice4_sedimentation_split.F90), FP64.
//
// Each species of a column is integrated in sub-steps of one common length, the largest that
// keeps every level's CFL number under XSPLIT_MAXCFL, until the time step is spent.  One
// wavefront holds one column in registers, level k in lane k % 64, slot k / 64 (L slots):
// a sub-step's fluxes are computed by all lanes at once, its length is a minimum over the
// wavefront, and the flux entering level k from above comes from the next lane.  A workgroup
// of SPLIT_COLS wavefronts takes as many adjacent columns: their inputs are read in rows of
// SPLIT_COLS doubles, a batch of up to SPLIT_BATCH fields at a time (every load of a batch
// issued before the first wait), and handed through LDS to the columns' wavefronts; the
// outputs go back the same way.  The two workgroups that take the halves of a 16-column row
// run on the same XCD, so that the row's cache lines are fetched once.  While a wavefront's
// state fits 128 registers (one slot, or two without hail) two workgroups share a compute
// unit, one's loads overlapping the other's sub-steps.
constexpr int SPLIT_COLS = 8, SPLIT_BATCH = 8;
template <int L, bool HAIL> constexpr int split_wpe() { return L == 1 || (L == 2 && !HAIL) ? 4 : 2; }
constexpr int SPLIT_MAX_SUBSTEPS = 1 << 16;  // per species and column; the reference has no
                                             // bound (its loop ends when the time is spent)

// Lane moves of an FP64 value inside the VALU, as two dword moves: a DPP move (lanes whose
// source lane is outside the wavefront keep old) and the gfx950 permlane swaps, which give
// a the lower and b the upper 16-lane rows of each 32-lane half (16) or the lower and upper
// halves (32) of the wavefront.  Every lane of the wavefront must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v, double old) {
  const int lo = __builtin_amdgcn_update_dpp((int)__double2loint(old), (int)__double2loint(v), CTRL,
                                             0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)__double2hiint(old), (int)__double2hiint(v), CTRL,
                                             0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
constexpr int DPP_QUAD_1032 = 0xB1, DPP_QUAD_2301 = 0x4E, DPP_ROW_MIRROR = 0x140,
              DPP_ROW_HALF_MIRROR = 0x141, DPP_WAVE_SHL1 = 0x130;
__device__ __forceinline__ void swap16_f64(double v, double &a, double &b) {
  const auto l = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
  const auto h = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
  a = __hiloint2double(h[0], l[0]);
  b = __hiloint2double(h[1], l[1]);
}
__device__ __forceinline__ void swap32_f64(double v, double &a, double &b) {
  const auto l = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
  const auto h = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
  a = __hiloint2double(h[0], l[0]);
  b = __hiloint2double(h[1], l[1]);
}

// Minimum over the wavefront, in every lane: within quads, half-rows and rows by DPP, then
// across rows and halves by the permlane swaps (min is exact, so the order is immaterial).
__device__ __forceinline__ double wave_min(double v) {
  v = dmin(v, dpp_f64<DPP_QUAD_1032>(v, v));
  v = dmin(v, dpp_f64<DPP_QUAD_2301>(v, v));
  v = dmin(v, dpp_f64<DPP_ROW_HALF_MIRROR>(v, v));
  v = dmin(v, dpp_f64<DPP_ROW_MIRROR>(v, v));
  double a, b;
  swap16_f64(v, a, b);
  v = dmin(a, b);
  swap32_f64(v, a, b);
  return dmin(a, b);
}

// Lane 0's value, in a scalar register.
__device__ __forceinline__ double lane0_f64(double v) {
  return __hiloint2double(__builtin_amdgcn_readlane((int)__double2hiint(v), 0),
                          __builtin_amdgcn_readlane((int)__double2loint(v), 0));
}

// Column group of block b: blocks b and b + 8 of a full group of 16 blocks, which the
// dispatcher sends to the same XCD, take the two halves of a 16-column row; the blocks of the
// last, partial group keep their order.
__device__ __forceinline__ int split_group(int b, int nblk) {
  if (b >= (nblk & ~15)) return b;
  return (b & ~15) + ((b & 7) << 1) + ((b >> 3) & 1);
}

// Stage NF fields (field f times scale[f]) of the workgroup's SPLIT_COLS columns from c0 on
// through LDS: v[f][q] of this lane receives level lane + 64 q of its wavefront's column (0
// outside the domain).  Every thread of the workgroup calls it.  With KEEP, fields 1 to 3 are
// also stored as read into a.keep[0..2] (when set), in the same rows.
template <int L, int NF, bool KEEP = false>
__device__ __forceinline__ void split_load(double (*tile)[L * 64][SPLIT_COLS + 1],
                                           const SplitArgs &a, const double *const (&f)[NF],
                                           const double (&scale)[NF], int c0, double (&v)[NF][L]) {
  // every load of the batch is issued unconditionally (outside the domain, element 0 is read
  // and dropped after the barrier), so that none waits for another
  size_t idx[L];
  bool in[L];
#pragma unroll
  for (int p = 0; p < L; ++p) {
    const int e = threadIdx.x + p * 64 * SPLIT_COLS, k = e / SPLIT_COLS, col = c0 + e - k * SPLIT_COLS;
    in[p] = k < a.nlev && col < a.ncol;
    idx[p] = in[p] ? (size_t)k * a.ld + col : 0;
  }
  double x[NF][L];
#pragma unroll
  for (int g = 0; g < NF; ++g)
#pragma unroll
    for (int p = 0; p < L; ++p) x[g][p] = f[g][idx[p]];
  __syncthreads();  // the previous batch has been read from the tile
#pragma unroll
  for (int g = 0; g < NF; ++g)
#pragma unroll
    for (int p = 0; p < L; ++p) {
      const int e = threadIdx.x + p * 64 * SPLIT_COLS, k = e / SPLIT_COLS;
      tile[g][k][e - k * SPLIT_COLS] = in[p] ? x[g][p] * scale[g] : 0.;
    }
  if constexpr (KEEP) {
    if (a.keep[0]) {  // uniform
#pragma unroll
      for (int g = 0; g < 3; ++g)
#pragma unroll
        for (int p = 0; p < L; ++p)
          if (in[p]) a.keep[g][idx[p]] = x[g + 1][p];
    }
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int g = 0; g < NF; ++g)
#pragma unroll
    for (int q = 0; q < L; ++q) v[g][q] = tile[g][lane + 64 * q][w];
}

// Mass flux ZWSED of species JRR at mixing ratio r above its threshold; rf is the density
// factor rho^(1-XCEXVT) (pristine ice) or rho^(XEXSED-XCEXVT) (the power-law species), fixed
// over the sub-steps (micro/mode_ice4_sedimentation_split.F90:574-578,630-633,747-749).
template <int JRR>
__device__ __forceinline__ double split_flux(double r, double rho, double rf) {
  if constexpr (JRR == 4) {
    return P.XFSEDI * r * rf *
           rpow<double>(dmax(0.05E6, -0.15319E6 - 0.021454E6 * rlog<double>(rho * r)), P.XEXCSEDI);
  } else {
    constexpr double fsed = JRR == 3 ? P.XFSEDR : (JRR == 5 ? P.XFSEDS : (JRR == 6 ? P.XFSEDG : P.XFSEDH));
    constexpr double exsed = JRR == 3 ? P.XEXSEDR : (JRR == 5 ? P.XEXSEDS : (JRR == 6 ? P.XEXSEDG : P.XEXSEDH));
    return fsed * rpow<double>(r, exsed) * rf;
  }
}

// INTERNAL_SEDIM_SPLI (micro/mode_ice4_sedimentation_split.F90:359-849) of species JRR for the
// wavefront's column: r (the state at t on entry, the working state after), s (the source,
// updated); returns the surface rate PINPRX.  Sets *capped when the sub-step cap is reached.
template <int L, int JRR>
__device__ __forceinline__ double split_species(const SplitArgs &a, int lane, const double (&rho)[L],
                                                const double (&dz)[L], const double (&oorhodz)[L],
                                                double (&r)[L], double (&s)[L], bool *capped) {
  constexpr double thr = JRR == 4 ? (P.XRTMIN[4] > 1.0E-7 ? P.XRTMIN[4] : 1.0E-7) : P.XRTMIN[JRR];
  const double dt = a.dt, zinv = 1. / dt;
  bool any = false;
#pragma unroll
  for (int q = 0; q < L; ++q) any = any || (lane + 64 * q < a.nlev && r[q] > thr);
  // nothing above the fall threshold: one sub-step of dt changes no source (:786-826)
  if (!__any(any)) return 0.;
  double prs[L], rf[L];
#pragma unroll
  for (int q = 0; q < L; ++q) {
    prs[q] = s[q] - r[q] * zinv;  // external tendency (:241-253)
    rf[q] = lane + 64 * q < a.nlev
                ? rpow<double>(rho[q], JRR == 4 ? 1. - P.XCEXVT
                                                : (JRR == 3 ? P.XEXSEDR : JRR == 5 ? P.XEXSEDS
                                                   : JRR == 6 ? P.XEXSEDG : P.XEXSEDH) - P.XCEXVT)
                : 0.;
  }
  double remaint = dt, inpr = 0.;
  for (int n = 0; remaint > 0.; ++n) {  // wave-uniform: remaint is the column's
    if (n == SPLIT_MAX_SUBSTEPS) {
      *capped = true;
      break;
    }
    double w[L], m = INFINITY;
#pragma unroll
    for (int q = 0; q < L; ++q) {
      const bool in = lane + 64 * q < a.nlev;
      w[q] = in && r[q] > thr ? split_flux<JRR>(r[q], rho[q], rf[q]) : 0.;
      if (in && r[q] > P.XRTMIN[JRR] && w[q] > 1.E-20)
        m = dmin(m, a.maxcfl * rho[q] * r[q] * dz[q] / w[q]);  // :795-808
    }
    const double step = dmin(remaint, wave_min(m));
    remaint = remaint - step;  // :814-815
    inpr = inpr + lane0_f64(w[0]) / P.XRHOLW * (step * zinv);
#pragma unroll
    for (int q = 0; q < L; ++q) {  // :820-824; level k + 1 is the next lane (or slot)
      const double next = q + 1 < L ? lane0_f64(w[q + 1 < L ? q + 1 : q]) : 0.;
      const double up = dpp_f64<DPP_WAVE_SHL1>(w[q], next);
      const double mrchange = step * oorhodz[q] * (up - w[q]);
      r[q] = r[q] + mrchange + prs[q] * step;
      s[q] = s[q] + mrchange * zinv;
    }
  }
  return inpr;
}

template <int L, bool HAIL, bool ZZZ>
__global__ void __launch_bounds__(64 * SPLIT_COLS)
    __attribute__((amdgpu_waves_per_eu(split_wpe<L, HAIL>()))) k_sedim_split(SplitArgs a) {
  __shared__ double tile[SPLIT_BATCH][L * 64][SPLIT_COLS + 1];
  constexpr int NV = HAIL ? 8 : 7;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c0 = split_group(blockIdx.x, gridDim.x) * SPLIT_COLS, i = c0 + w;
  const double ts = a.tscale;
  double lv[L], ls[L], r[NV - 3][L], rho[L], dz[L], src[NV][L];
  // batch 1: exn and the state at t of th .. rg; batch 2: hail's state at t, rho, dz and
  // sources; batch 3: the last sources.  RAIN_ICE's L/cph factors of the state at t
  // (rain_ice.F90:335-345) as soon as that state is in.
  double b1[8][L];
  {
    const double *const f1[8] = {a.exn, a.t[0], a.t[1], a.t[2], a.t[3], a.t[4], a.t[5], a.t[6]};
    const double s1[8] = {1., ts, ts, ts, ts, ts, ts, ts};
    split_load<L, 8, true>(tile, a, f1, s1, c0, b1);
  }
  auto factors = [&](const double (&rh_t)[L]) {
#pragma unroll
    for (int q = 0; q < L; ++q) {
      double tq[8];
#pragma unroll
      for (int v = 0; v < 7; ++v) tq[v] = b1[v + 1][q];
      tq[7] = rh_t[q];
      rain_ice_factors<HAIL>(tq, b1[0][q], ls[q], lv[q]);
#pragma unroll
      for (int v = 3; v < 7; ++v) r[v - 3][q] = b1[v + 1][q];
      if constexpr (HAIL) r[4][q] = rh_t[q];
    }
  };
  if constexpr (!HAIL) {
    const double zero[L] = {};
    factors(zero);
  }
  {
    constexpr int N2 = HAIL ? 8 : 7;
    double b2[N2][L];
    const double *const f2[8] = {a.rhodref, a.dzz, a.src[0], a.src[1], a.src[2], a.src[3],
                                 a.src[4], a.t[7]};
    const double s2[8] = {1., 1., 1., 1., 1., 1., 1., ts};
    split_load<L, N2>(tile, a, reinterpret_cast<const double *const (&)[N2]>(f2),
                      reinterpret_cast<const double (&)[N2]>(s2), c0, b2);
    if constexpr (ZZZ) {  // RAIN_ICE's model level height (rain_ice.F90:387-395), in its order
      if (lane == 0 && i < a.ncol) {
        double zzsum = 0.;
        for (int k = a.nlev - 1; k >= 0; --k) {
          zzsum = zzsum + tile[1][k][w];
          a.zzz[(size_t)k * a.ld + i] = zzsum - tile[1][k][w] * 0.5;
        }
      }
    }
    if constexpr (HAIL) factors(b2[7]);
#pragma unroll
    for (int q = 0; q < L; ++q) {
      rho[q] = b2[0][q];
      dz[q] = b2[1][q];
#pragma unroll
      for (int v = 0; v < 5; ++v) src[v][q] = b2[v + 2][q];
    }
  }
  {
    constexpr int N3 = HAIL ? 3 : 2;
    double b3[N3][L];
    const double *const f3[3] = {a.src[5], a.src[6], a.src[7]};
    const double s3[3] = {1., 1., 1.};
    split_load<L, N3>(tile, a, reinterpret_cast<const double *const (&)[N3]>(f3),
                      reinterpret_cast<const double (&)[N3]>(s3), c0, b3);
#pragma unroll
    for (int v = 0; v < N3; ++v)
#pragma unroll
      for (int q = 0; q < L; ++q) src[v + 5][q] = b3[v][q];
  }
  double oorhodz[L];
#pragma unroll
  for (int q = 0; q < L; ++q)
    oorhodz[q] = lane + 64 * q < a.nlev ? 1. / (rho[q] * dz[q]) : 0.;  // ZW (:263)
  if (i < a.ncol) {
    bool capped = false;
    const double pr = split_species<L, 3>(a, lane, rho, dz, oorhodz, r[0], src[IRR], &capped);
    const double pi = split_species<L, 4>(a, lane, rho, dz, oorhodz, r[1], src[IRI], &capped);
    const double ps = split_species<L, 5>(a, lane, rho, dz, oorhodz, r[2], src[IRS], &capped);
    const double pg = split_species<L, 6>(a, lane, rho, dz, oorhodz, r[3], src[IRG], &capped);
    double ph = 0.;
    if constexpr (HAIL) ph = split_species<L, 7>(a, lane, rho, dz, oorhodz, r[4], src[IRH], &capped);
    if (lane == 0) {
      a.inprc[i] = 0.;
      a.inprr[i] = pr;
      a.inpri[i] = pi;
      a.inprs[i] = ps + pi;  // mode_ice4_sedimentation.F90:212
      a.inprg[i] = pg;
      if constexpr (HAIL) a.inprh[i] = ph;
      if (capped && a.capped) atomicAdd(a.capped, 1ull);
    }
#pragma unroll
    for (int q = 0; q < L; ++q) {  // ICE4_CORRECT_NEGATIVITIES (:216-218), every point
      double rh = HAIL ? src[NV - 1][q] : 0.;
      correct_negativities<HAIL>(src[ITH][q], src[IRV][q], src[IRC][q], src[IRR][q], src[IRI][q],
                                 src[IRS][q], src[IRG][q], lv[q], ls[q], &rh);
      if constexpr (HAIL) src[NV - 1][q] = rh;
    }
  }
  // the sources back through LDS, in rows of SPLIT_COLS columns
  __syncthreads();  // every wavefront has read the last batch
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int q = 0; q < L; ++q) tile[v][lane + 64 * q][w] = src[v][q];
  __syncthreads();
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int p = 0; p < L; ++p) {
      const int e = threadIdx.x + p * 64 * SPLIT_COLS, k = e / SPLIT_COLS,
                c = e - k * SPLIT_COLS, col = c0 + c;
      if (k < a.nlev && col < a.ncol) a.dst[v][(size_t)k * a.ld + col] = tile[v][k][c];
    }
}

// ---- k_micro: the ICE4 microphysics, persistent, with the time splitting compacted
// across each wavefront.
//
// Every point makes one hard tendency call, and 42 % of them (on the 512x512x90 synthetic
// workload) then make soft calls, most of them one, a few up to five.  Run one point per
// lane to the end and a wavefront makes as many soft passes as its slowest lane needs,
// nearly all with few lanes active.  Here a wavefront takes 64-point chunks in turn; after
// the chunk's hard call, the lanes whose time step goes on are joined by points parked in
// LDS by earlier chunks, and a soft pass runs over a full wavefront (64 points) whenever
// that many are at hand, or over more than MicroPark::CAP points when the park would
// otherwise overflow; unfinished points are parked again (ballot + prefix count give
// each its slot).  The per-point arithmetic is that of micro_at, only the order in which
// points are advanced differs, so the results are those of one thread per point.
// Eight wavefronts per workgroup, two per SIMD: the register budget of 256 per wavefront
// costs a few spills (FP64), but the second wavefront hides the memory latency the first
// one waits on.  The workgroup's wavefronts share the lookup tables, staged once into LDS.
// Wavefronts per workgroup.  The mixed path (183 VGPRs at two wavefronts per SIMD) measured
// slower with three (168 VGPRs, spills, 49-slot parks: +22 %) or four (+120 %) per SIMD.
// The hail state needs about 306 registers: one wavefront per SIMD (at two, 55 registers
// spill and the kernel is slower than one thread per point), with 49-slot parks.
template <class R, bool HAIL = false> constexpr int micro_waves() { return HAIL ? 4 : 8; }
// Lookup tables staged in LDS: the ICE3 block; the hail kernel reads the wet-growth kernels
// of hail, which follow it, from device memory.
template <class R, bool HAIL = false> constexpr size_t micro_tab_bytes() {
  return ((size_t)ICE3_T_ICE3 * sizeof(R) + 15) / 16 * 16;
}
// Variables of the time splitting (KRR + 1) and rates kept between tendency calls.
template <bool HAIL> constexpr int micro_nv() { return HAIL ? 8 : 7; }
template <class R, bool HAIL> constexpr int micro_ns() {
  return (int)(sizeof(typename StepStateFor<R, HAIL>::type) / sizeof(R));
}
// Park slots per wavefront: what the LDS holds beside the tables, at most 63 (31 in FP64,
// 63 in the mixed path, 49 with hail).
template <class R, bool HAIL = false> constexpr int micro_park_cap() {
  const size_t slot = (micro_nv<HAIL>() + 2) * sizeof(double) + micro_ns<R, HAIL>() * sizeof(R) +
                      2 * sizeof(int);
  const size_t n = (160 * 1024 - micro_tab_bytes<R, HAIL>()) / micro_waves<R, HAIL>() / slot;
  return n < 63 ? (int)n : 63;
}

// Points parked by one wavefront: state between tendency calls, one row per field.
// Capacity: what fits the LDS beside the tables (fewer than 64).
template <class R, bool HAIL = false> struct MicroPark {
  static constexpr int NV = micro_nv<HAIL>(), NS = micro_ns<R, HAIL>();
  static constexpr int CAP = micro_park_cap<R, HAIL>();
  double d[NV + 2][CAP];   // vart[NV], cit, time
  R s[NS][CAP];            // StepState (StepStateHail)
  unsigned int q[CAP];     // point index k * ncol + i
  int calls[CAP];
};
static_assert(sizeof(StepState<double>) == 32 * sizeof(double), "StepState layout");
static_assert(sizeof(StepState<float>) == 32 * sizeof(float), "StepState layout");
static_assert(sizeof(StepStateHail<double>) == 43 * sizeof(double), "StepStateHail layout");

template <class R, bool HAIL = false> constexpr size_t micro_lds_bytes() {
  return micro_tab_bytes<R, HAIL>() + micro_waves<R, HAIL>() * sizeof(MicroPark<R, HAIL>);
}
static_assert(MicroPark<double>::CAP == 31 && MicroPark<float>::CAP == 63 &&
              MicroPark<double, true>::CAP == 49, "park capacity");
static_assert(micro_lds_bytes<double>() <= 160 * 1024, "k_micro<double> LDS");
static_assert(micro_lds_bytes<float>() <= 160 * 1024, "k_micro<float> LDS");
static_assert(micro_lds_bytes<double, true>() <= 160 * 1024, "k_micro<double, false, true> LDS");

template <class R, bool HAIL, class SS>
__device__ __forceinline__ void park_store(MicroPark<R, HAIL> &pk, int e, unsigned int q,
                                           const double *vart, double cit, double time,
                                           const SS &S, int calls) {
  constexpr int NV = MicroPark<R, HAIL>::NV, NS = MicroPark<R, HAIL>::NS;
#pragma unroll
  for (int f = 0; f < NV; ++f) pk.d[f][e] = vart[f];
  pk.d[NV][e] = cit;
  pk.d[NV + 1][e] = time;
  const R *sv = reinterpret_cast<const R *>(&S);
#pragma unroll
  for (int f = 0; f < NS; ++f) pk.s[f][e] = sv[f];
  pk.q[e] = q;
  pk.calls[e] = calls;
}

template <class R, bool HAIL, class SS>
__device__ __forceinline__ unsigned int park_load(const MicroPark<R, HAIL> &pk, int e, double *vart,
                                                  double &cit, double &time, SS &S, int &calls) {
  constexpr int NV = MicroPark<R, HAIL>::NV, NS = MicroPark<R, HAIL>::NS;
#pragma unroll
  for (int f = 0; f < NV; ++f) vart[f] = pk.d[f][e];
  cit = pk.d[NV][e];
  time = pk.d[NV + 1][e];
  R *sv = reinterpret_cast<R *>(&S);
#pragma unroll
  for (int f = 0; f < NS; ++f) sv[f] = pk.s[f][e];
  calls = pk.calls[e];
  return pk.q[e];
}

template <class R, bool OCND2 = false, bool HAIL = false>
__global__ void __launch_bounds__((64 * micro_waves<R, HAIL>())) k_micro(MicroArgs a) {
  extern __shared__ __align__(16) unsigned char micro_lds[];
  constexpr int NV = micro_nv<HAIL>();
  R *tab = reinterpret_cast<R *>(micro_lds);
  for (int e = threadIdx.x; e < ICE3_T_ICE3; e += blockDim.x) tab[e] = (R)a.tab[e];
  __syncthreads();
  // hail: the wet-growth kernels of hail from device memory
  const double *htab = HAIL ? a.tab + ICE3_T_ICE3 : nullptr;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  MicroPark<R, HAIL> &pk =
      reinterpret_cast<MicroPark<R, HAIL> *>(micro_lds + micro_tab_bytes<R, HAIL>())[wid];
  const unsigned long long below = (1ull << lane) - 1ull;  // lanes under this one
  const int cpl = (a.ncol + 63) >> 6;                       // chunks per level
  const int nchunk = cpl * a.nlev;
  const double dt = a.dt;
  const PointConst<R> pc_soft{};  // soft calls read no point constants
  int parked = 0;                 // wave-uniform
  bool active = false;            // this lane holds a point whose time step goes on
  unsigned int q = 0;             // that point: index k * ncol + i, offset j
  size_t j = 0;
  StepPointN<NV> p;
  double vart[NV], cit = 0., time = 0.;
  typename StepStateFor<R, HAIL>::type S;
  int calls = 0;
  unsigned long long ncalls = 0, ncap = 0, nmicro = 0;
  for (int c = blockIdx.x * micro_waves<R, HAIL>() + wid;; c += gridDim.x * micro_waves<R, HAIL>()) {
    const bool more = c < nchunk;  // wave-uniform
    if (more) {
      const int k = c / cpl, i = (c - k * cpl) * 64 + lane;
      if (i < a.ncol) {
        q = (unsigned int)k * (unsigned int)a.ncol + (unsigned int)i;
        j = (size_t)k * a.ld + i;
        double t[NV], s[NV];
        micro_load<NV>(a, j, t, s);
        cit = a.src_ci[j];
        double rvheni = 0., evap3d = 0.;
        int st = 0;
        calls = 0;
        if (micro_mask<OCND2, HAIL>(t, a, j)) {
          ++nmicro;
          step_point(dt, a.exn[j], a.rhodref[j], a.pabs[j], a.cldfr[j], a.sigs[j], s, t, p);
#pragma unroll
          for (int f = 0; f < NV; ++f) vart[f] = t[f];
          time = 0.;
          const PointConst<R> pc = point_const<R>(p.rhodref);
          MicroOcnd2 xo;
          if constexpr (OCND2) xo = ocnd2_point(a, j);
          st = step_call<R, false, OCND2, HAIL>(dt, p, pc, vart, cit, time, S, calls,
                                                DEFAULT_MAX_CALLS, tab, &xo, htab);
          if (st == 0) evap3d = (double)S.rrevav;
        } else {
          micro_nucleation<R, OCND2, HAIL>(a, j, t, cit, rvheni);
#pragma unroll
          for (int f = 0; f < NV; ++f) vart[f] = t[f];
        }
        if (st > 0) {
          active = true;
        } else {
          ncalls += calls;
          ncap += st < 0;
          micro_finish<HAIL>(a, j, vart, cit, evap3d, rvheni, st < 0 ? -calls : calls);
        }
      }
    }
    // soft passes: over a full wavefront while 64 points are at hand, over all of them
    // while they would not fit the park, and over all of them once the chunks are done
    for (;;) {
      const unsigned long long am = __ballot(active);
      const int nact = __popcll(am), need = 64 - nact;
      if (more ? parked < need && nact + parked <= MicroPark<R, HAIL>::CAP : nact + parked == 0) break;
      if (!active) {
        const int r = __popcll(~am & below);  // rank among the idle lanes
        if (r < parked) {
          q = park_load(pk, parked - 1 - r, vart, cit, time, S, calls);
          const unsigned int k = q / (unsigned int)a.ncol;
          j = (size_t)k * a.ld + (q - k * (unsigned int)a.ncol);
          double t[NV], s[NV];
          micro_load<NV>(a, j, t, s);
          step_point(dt, a.exn[j], a.rhodref[j], a.pabs[j], a.cldfr[j], a.sigs[j], s, t, p);
          active = true;
        }
      }
      parked -= need < parked ? need : parked;
      if (active) {
        const int st = step_call<R, true, OCND2, HAIL>(dt, p, pc_soft, vart, cit, time, S, calls,
                                                       DEFAULT_MAX_CALLS, tab, nullptr, htab);
        if (st <= 0) {
          ncalls += calls;
          ncap += st < 0;
          micro_finish<HAIL>(a, j, vart, cit, st == 0 ? (double)S.rrevav : 0., 0.,
                             st < 0 ? -calls : calls);
          active = false;
        }
      }
    }
    if (!more) break;
    // park the points whose time step goes on (parked + their number <= CAP here)
    const unsigned long long am = __ballot(active);
    if (active) {
      park_store(pk, parked + __popcll(am & below), q, vart, cit, time, S, calls);
      active = false;
    }
    parked += __popcll(am);
    // the stores above are read by other lanes of this wavefront in later passes
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  // every lane of the wavefront takes part in the sums (uniform branches)
  const unsigned long long cc = wave_sum(ncap);
  if (cc != 0 && lane == 0 && a.capped) atomicAdd(a.capped, cc);
  if (a.counters) {
    const unsigned long long n = wave_sum(ncalls), m = wave_sum(nmicro);
    if (lane == 0) {
      atomicAdd(&a.counters[0], n);
      atomicAdd(&a.counters[1], cc);
      atomicAdd(&a.counters[2], m);
    }
  }
}

// The microphysics one thread per point, with the context's sub-step cap a.max_calls read at
// run time (k_micro has the default cap compiled in: the register allocator fails on
// k_micro<double> with a run-time bound).  Launched when a context sets another cap; the
// per-point arithmetic is k_micro's, so the results are the same for the same cap.  With HAIL
// (KRR=7) it is the microphysics of the hail path, with the wet-growth tables of hail.
template <class R, bool OCND2 = false, bool HAIL = false>
__global__ void __launch_bounds__(256) k_micro_points(MicroArgs a) {
  const unsigned int q = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long ncalls = 0, ncap = 0, nmicro = 0;
  if (q < (unsigned int)a.nlev * (unsigned int)a.ncol) {
    const unsigned int k = q / (unsigned int)a.ncol;
    const int c = micro_at<R, OCND2, HAIL>(a, (size_t)k * a.ld + (q - k * (unsigned int)a.ncol), a.tab);
    ncalls = c < 0 ? -c : c;
    ncap = c < 0;
    nmicro = c != 0;
  }
  const unsigned long long cc = wave_sum(ncap);
  if (cc != 0 && (threadIdx.x & 63) == 0 && a.capped) atomicAdd(a.capped, cc);
  if (a.counters) {
    const unsigned long long n = wave_sum(ncalls), m = wave_sum(nmicro);
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&a.counters[0], n);
      atomicAdd(&a.counters[1], cc);
      atomicAdd(&a.counters[2], m);
    }
  }
}

// ICE_ADJUST one thread per point with the record of the branches each point took (the test
// hook of ice3_ice_adjust_branches).
template <bool HAIL>
__global__ void __launch_bounds__(256) k_ice_adjust_branches(AdjustArgs a, unsigned *branches) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  const size_t j = (size_t)blockIdx.y * a.ld + i;
  unsigned br = 0;
  adjust_at<false, HAIL>(a, i, j, 0, nullptr, &br);
  branches[j] = br;
}

// The microphysics one thread per point with the branch record of each point (the test
// hook of ice3_rain_ice_branches; the counters of ice3_stats are not kept).
template <bool HAIL>
__global__ void __launch_bounds__(256) k_micro_branches(MicroArgs a) {
  const unsigned int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (unsigned int)a.nlev * (unsigned int)a.ncol) return;
  const unsigned int k = q / (unsigned int)a.ncol;
  const size_t j = (size_t)k * a.ld + (q - k * (unsigned int)a.ncol);
  unsigned br = 0;
  micro_at<double, false, HAIL>(a, j, a.tab, &br);
  a.branches[j] = br;
}

struct SurrogateArgs {
  MicroArgs m;
  unsigned int npts;                 // nlev * ncol; point q = k * ncol + i
  const ice3_surrogate_norm *norm;
  float budget_tol, neg_tol;
  __bf16 *x;                         // [npts][MLP_IN] network inputs
  float *y;                          // [npts][MLP_NTARG] network outputs
  unsigned char *flag;               // [npts] SG_* reason, SG_OK = ask the network
  unsigned int *queue;               // [0] count, [1 + q] point offsets
  unsigned long long *counters;      // 5 entries, see include/ice3.h
};

// Reasons a point is not taken from the network.
enum { SG_OK = 0, SG_RANGE = 1, SG_NEG = 2, SG_BUDGET = 3, SG_NOMICRO = 4 };

__device__ __forceinline__ size_t point_offset(const MicroArgs &m, unsigned int q) {
  return (size_t)(q / m.ncol) * m.ld + q % m.ncol;
}

// Stage 1 of the surrogate: features, range check, network inputs (one thread per point).
__global__ void __launch_bounds__(256) k_sg_features(SurrogateArgs a) {
  const unsigned int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.npts) return;
  const ice3_surrogate_norm &nm = *a.norm;
  float f[MLP_NFEAT];
  int flag = SG_NOMICRO;
  if (surrogate_features(a.m, point_offset(a.m, q), f)) {
    flag = SG_OK;
#pragma unroll
    for (int c = 0; c < MLP_NFEAT; ++c)
      if (!(f[c] >= nm.feat_lo[c] && f[c] <= nm.feat_hi[c])) flag = SG_RANGE;
  }
  a.flag[q] = (unsigned char)flag;
  bf16x8 v[MLP_IN / 8];
#pragma unroll
  for (int c = 0; c < MLP_IN; ++c)
    v[c / 8][c % 8] = (__bf16)(flag == SG_OK && c < MLP_NFEAT ? (f[c] - nm.feat_mean[c]) * nm.feat_scale[c] : 0.f);
  bf16x8 *xr = (bf16x8 *)(a.x + (size_t)q * MLP_IN);
#pragma unroll
  for (int c = 0; c < MLP_IN / 8; ++c) xr[c] = v[c];
}

// Stage 2: the network over all points, k_mlp_infer (ice3_mlp.hip, launched by
// mlp_infer_launch).


// Stage 3: decode, physical checks, results or fallback queue.  Persistent grid: block b
// decodes the 256-point chunks b, b + gridDim.x, ... (spread over all levels, so every
// block meets the same mix of cheap and costly fallback points) and appends its rejected
// points to its own queue segment queue[1 + b*seg ...], counting them in qcount[b]; no
// global atomics but the 5 counters.
// Decode blocks: a fixed count (8 per CU on a 256-CU MI355X), so that the scratch size
// depends on the problem size only.  Block b's queue segment holds seg =
// ceil(chunks / SG_DEC_BLOCKS) * 256 entries; all segments fit in npts + 256 * (SG_DEC_BLOCKS + 1).
constexpr unsigned int SG_DEC_BLOCKS = 2048;

__device__ __forceinline__ int sg_decode_point(const SurrogateArgs &a, unsigned int q, size_t j) {
  const MicroArgs &m = a.m;
  const ice3_surrogate_norm &nm = *a.norm;
  int flag = a.flag[q];
  if (flag != SG_OK) return flag;
  double out[7], y[7], rsum = 0., rmag = 0.;
  const float *yq = a.y + (size_t)q * MLP_NTARG;
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    y[v] = (double)(nm.targ_unit[v] * sinhf(nm.targ_mean[v] + nm.targ_std[v] * yq[v]));
    out[v] = m.src[v][j] + y[v];
    if (v >= 1) { rsum += y[v]; rmag += fabs(y[v]); }
  }
  // water residual against the predicted changes, with a floor of budget_tol times the
  // smallest typical change (median |change| of the training targets)
  double umin = nm.targ_unit[IRV];
#pragma unroll
  for (int v = IRC; v <= IRG; ++v) umin = fmin(umin, (double)nm.targ_unit[v]);
  if (!(fabs(rsum) <= (double)a.budget_tol * (rmag + umin))) return SG_BUDGET;
  out[IRV] -= rsum;  // move the water residual into vapour
  // a source may go negative by neg_tol of (its value + its change + the typical change)
#pragma unroll
  for (int v = IRC; v <= IRG; ++v)
    if (!(out[v] >= -(double)a.neg_tol * (fabs(m.src[v][j]) + fabs(y[v]) + (double)nm.targ_unit[v])))
      return SG_NEG;
  double lvfact, lsfact;
  latent_factors(m, j, lvfact, lsfact);
  correct_negativities(out[ITH], out[IRV], out[IRC], out[IRR], out[IRI], out[IRS], out[IRG],
                       lvfact, lsfact);
  if (!(out[IRV] >= 0.)) return SG_NEG;
  const double cinew =
      fmax(0., m.src_ci[j] + (double)(nm.targ_unit[7] * sinhf(nm.targ_mean[7] + nm.targ_std[7] * yq[7])));
#pragma unroll
  for (int v = 0; v < 7; ++v) m.dst[v][j] = out[v];
  m.ci[j] = cinew;
  m.evap3d[j] = 0.;
  m.rainfr[j] = 1.;
  if (m.calls) m.calls[j] = 0;
  return SG_OK;
}

__global__ void __launch_bounds__(256) k_sg_decode(SurrogateArgs a, unsigned int seg,
                                                   unsigned int *qcount) {
  __shared__ unsigned int sWave[MLP_WAVES];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const unsigned int q0 = blockIdx.x * seg;   // this block's queue segment
  unsigned int nq = 0;                 // rejected points of this block so far (uniform)
  unsigned int cnt[5] = {0, 0, 0, 0, 0};
  for (unsigned int base = blockIdx.x * blockDim.x; base < a.npts; base += gridDim.x * blockDim.x) {
    const unsigned int q = base + threadIdx.x;
    int flag = -1;
    size_t j = 0;
    if (q < a.npts) {
      j = point_offset(a.m, q);
      flag = sg_decode_point(a, q, j);
      ++cnt[flag];
    }
    const bool rej = flag > SG_OK;
    const unsigned long long mask = __ballot(rej);
    if (lane == 0) sWave[w] = (unsigned int)__popcll(mask);
    __syncthreads();
    unsigned int off = nq;
    for (int v = 0; v < w; ++v) off += sWave[v];
    if (rej) a.queue[1 + q0 + off + __popcll(mask & ((1ull << lane) - 1ull))] = (unsigned int)j;
    for (int v = 0; v < MLP_WAVES; ++v) nq += sWave[v];
    __syncthreads();
  }
  if (threadIdx.x == 0) qcount[blockIdx.x] = nq;
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    unsigned long long v = cnt[c];
#pragma unroll
    for (int mm = 32; mm >= 1; mm >>= 1) v += __shfl_xor(v, mm, 64);
    if (lane == 0 && v) atomicAdd(&a.counters[c], v);
  }
}

// Plain microphysics over the surrogate's rejected points: block b runs the qcount[b]
// points that decode block b queued.
template <class R>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) k_micro_segments(MicroArgs a, const unsigned int *queue,
                                                       const unsigned int *qcount, unsigned int seg) {
  const unsigned int n = qcount[blockIdx.x];
  const unsigned int *list = queue + 1 + (size_t)blockIdx.x * seg;
  for (unsigned int e = threadIdx.x; e < n; e += blockDim.x)
    if (micro_at<R>(a, (size_t)list[e], a.tab) < 0 && a.capped) atomicAdd(a.capped, 1ull);
}

// Scratch layout of the surrogate step (bytes, 256-aligned blocks).
struct SgWork {
  size_t x, y, flag, queue, qcount, total;
};
static SgWork sg_work_layout(size_t npts) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  SgWork w;
  w.x = 0;
  w.y = w.x + al(npts * MLP_IN * 2);
  w.flag = w.y + al(npts * MLP_NTARG * 4);
  w.queue = w.flag + al(npts);
  w.qcount = w.queue + al((npts + 1 + 256 * ((size_t)SG_DEC_BLOCKS + 1)) * 4);
  w.total = w.qcount + al((size_t)SG_DEC_BLOCKS * 4);
  return w;
}

// Surrogate features of the points of MicroArgs (training data and tests).
__global__ void __launch_bounds__(256) k_features(MicroArgs m, unsigned int npts,
                                                  const ice3_surrogate_norm *norm, float *raw,
                                                  __bf16 *x, unsigned char *micro) {
  const unsigned int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  const size_t j = (size_t)(q / m.ncol) * m.ld + q % m.ncol;
  float f[MLP_NFEAT];
  const bool mi = surrogate_features(m, j, f);
  if (micro) micro[q] = mi ? 1 : 0;
  if (raw)
    for (int c = 0; c < MLP_NFEAT; ++c) raw[(size_t)q * MLP_NFEAT + c] = f[c];
  if (x)
    for (int c = 0; c < MLP_IN; ++c)
      x[(size_t)q * MLP_IN + c] =
          (__bf16)(c < MLP_NFEAT ? (f[c] - norm->feat_mean[c]) * norm->feat_scale[c] : 0.f);
}

}  // namespace ice3

// Scratch of the allocation-free surrogate step (ice3_adjust_rain_ice_surrogate_noalloc):
// the same layout as the caller's work area of ice3_adjust_rain_ice_surrogate.  One copy per
// GPU, held until the process ends.  Host code never uses the name: a kernel argument made
// from it would be a host address; ice3_ctx_create takes the device address with
// hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[ICE3_SURROGATE_SCRATCH_BYTES];

// ------------------------------------------------------------------ C ABI

struct ice3_ctx {
  int device;
  int ncu;            // compute units of the device: k_micro runs one workgroup on each
  ice3_params params;
  double *d_tab;
  double h_tab[ICE3_T_TOTAL];
  double *d_tiwmx;    // TIWMX tables of the OCND2 branch (ice3_tiwmx_tables order)
  unsigned long long *d_capped;  // points that hit the sub-step cap since the last status query
  void *d_scratch;    // device address of this device's copy of scratch
  int max_calls;      // sub-step cap per point
  int sedim;          // ICE3_SEDIM_STAT or ICE3_SEDIM_SPLI
  double maxcfl;      // XSPLIT_MAXCFL of the split sedimentation
  void *d_work;       // work area of the fused step with the split sedimentation
  size_t work_bytes;
  int timing;         // record ev[] around the stages of each call
  hipEvent_t ev[7];   // before adjust, after adjust, after sedimentation, then after the
                      // microphysics (plain step: events 3..6 together) or, for the surrogate
                      // step, after the features, the network, the decode and the fallback
};

using namespace ice3;

#define HIPCHK(x)                                 \
  do {                                            \
    if ((x) != hipSuccess) return ICE3_EDEVICE;   \
  } while (0)

extern "C" int ice3_ctx_create(ice3_ctx **out) {
  if (!out) return ICE3_EINVAL;
  ice3_ctx *c = (ice3_ctx *)calloc(1, sizeof(ice3_ctx));
  if (!c) return ICE3_EINVAL;
  ice3_build_params(&c->params, c->h_tab);
  // the kernels carry the constants as literals (P); they must be the host builder's
  if (memcmp(&c->params, &P, sizeof(ice3_params)) != 0) {
    free(c);
    return ICE3_ECONFIG;
  }
  constexpr size_t TIWMX_BYTES = sizeof(double) * ICE3_TW_NTAB * ICE3_TIWMX_N;
  double *tiwmx = (double *)malloc(TIWMX_BYTES);
  if (!tiwmx) {
    free(c);
    return ICE3_EDEVICE;
  }
  ice3_build_tiwmx(&c->params, tiwmx);
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) !=
          hipSuccess || c->ncu <= 0 ||
      hipMalloc((void **)&c->d_tab, sizeof(double) * ICE3_T_TOTAL) != hipSuccess ||
      hipMemcpy(c->d_tab, c->h_tab, sizeof(double) * ICE3_T_TOTAL, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMalloc((void **)&c->d_tiwmx, TIWMX_BYTES) != hipSuccess ||
      hipMemcpy(c->d_tiwmx, tiwmx, TIWMX_BYTES, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMalloc((void **)&c->d_capped, sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(c->d_capped, 0, sizeof(unsigned long long)) != hipSuccess ||
      hipGetSymbolAddress(&c->d_scratch, HIP_SYMBOL(scratch)) != hipSuccess ||
      c->d_scratch == nullptr) {
    if (c->d_tab) (void)hipFree(c->d_tab);
    if (c->d_tiwmx) (void)hipFree(c->d_tiwmx);
    if (c->d_capped) (void)hipFree(c->d_capped);
    free(tiwmx);
    free(c);
    return ICE3_EDEVICE;
  }
  free(tiwmx);
  c->max_calls = DEFAULT_MAX_CALLS;
  c->sedim = ICE3_SEDIM_STAT;
  c->maxcfl = 0.8;  // PARAM_ICE XSPLIT_MAXCFL default (micro/modd_param_icen.F90:347)
  *out = c;
  return ICE3_OK;
}

extern "C" void ice3_ctx_destroy(ice3_ctx *c) {
  if (!c) return;
  if (c->d_tab) (void)hipFree(c->d_tab);
  if (c->d_tiwmx) (void)hipFree(c->d_tiwmx);
  if (c->d_capped) (void)hipFree(c->d_capped);
  if (c->d_work) (void)hipFree(c->d_work);
  if (c->timing)
    for (int q = 0; q < 7; ++q) (void)hipEventDestroy(c->ev[q]);
  free(c);
}

extern "C" int ice3_stage_timing(ice3_ctx *c, int enable) {
  if (!c) return ICE3_EINVAL;
  if (enable && !c->timing) {
    for (int q = 0; q < 7; ++q) HIPCHK(hipEventCreate(&c->ev[q]));
    c->timing = 1;
  } else if (!enable && c->timing) {
    for (int q = 0; q < 7; ++q) (void)hipEventDestroy(c->ev[q]);
    c->timing = 0;
  }
  return ICE3_OK;
}

extern "C" int ice3_stage_times(ice3_ctx *c, float ms[6]) {
  if (!c || !ms || !c->timing) return ICE3_EINVAL;
  HIPCHK(hipEventSynchronize(c->ev[6]));
  for (int q = 0; q < 6; ++q) HIPCHK(hipEventElapsedTime(&ms[q], c->ev[q], c->ev[q + 1]));
  return ICE3_OK;
}

extern "C" int ice3_ctx_set_substep_cap(ice3_ctx *c, int max_calls) {
  if (!c || max_calls < 1) return ICE3_EINVAL;
  c->max_calls = max_calls;
  return ICE3_OK;
}

extern "C" int ice3_ctx_set_sedimentation(ice3_ctx *c, int scheme, double max_cfl) {
  if (!c || (scheme != ICE3_SEDIM_STAT && scheme != ICE3_SEDIM_SPLI) || !(max_cfl > 0.) ||
      !(max_cfl < INFINITY))
    return ICE3_EINVAL;
  c->sedim = scheme;
  c->maxcfl = max_cfl;
  return ICE3_OK;
}

extern "C" int ice3_ctx_status(ice3_ctx *c, void *stream, unsigned long long *capped) {
  if (!c) return ICE3_EINVAL;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev != c->device) HIPCHK(hipSetDevice(c->device));
  const hipStream_t s = (hipStream_t)stream;
  unsigned long long h = 0;
  int rc = ICE3_OK;
  if (hipMemcpyAsync(&h, c->d_capped, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess ||
      (h != 0 && hipMemsetAsync(c->d_capped, 0, sizeof h, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess)
    rc = ICE3_EDEVICE;
  if (dev != c->device) (void)hipSetDevice(dev);
  if (rc != ICE3_OK) return rc;
  if (capped) *capped = h;
  return h != 0 ? ICE3_ESUBSTEP : ICE3_OK;
}

// Record stage event q of a timed call (no-op when timing is off).
static int mark(ice3_ctx *c, int q, hipStream_t s) {
  if (c->timing) HIPCHK(hipEventRecord(c->ev[q], s));
  return ICE3_OK;
}
#define MARK(q) do { int rc_ = mark(ctx, (q), s); if (rc_) return rc_; } while (0)

extern "C" const char *ice3_strerror(int code) {
  switch (code) {
    case ICE3_OK: return "ok";
    case ICE3_EINVAL: return "invalid argument (size, stride, precision or NULL pointer)";
    case ICE3_EDEVICE: return "HIP runtime error";
    case ICE3_ESUBSTEP: return "a point reached the sub-step cap of the ICE4 time splitting";
    case ICE3_ECONFIG: return "constants compiled into the kernels differ from the host builder (rebuild)";
    case ICE3_ETOOLARGE:
      return "the problem needs more scratch than the library's static area "
             "(ICE3_SURROGATE_SCRATCH_BYTES); use ice3_adjust_rain_ice_surrogate with a work area";
    default: return "unknown error";
  }
}

static void put(std::string &s, const char *name, const double *v, int n) {
  char b[64];
  snprintf(b, sizeof b, "%s %d", name, n);
  s += b;
  for (int k = 0; k < n; ++k) {
    snprintf(b, sizeof b, " %.17e", v[k]);
    s += b;
  }
  s += "\n";
}

extern "C" long ice3_dump_constants(const ice3_ctx *ctx, char *buf, long buflen) {
  // a NULL context builds the constants on the host only (no device needed)
  static ice3_ctx host;
  const ice3_ctx *c = ctx;
  if (!c) {
    ice3_build_params(&host.params, host.h_tab);
    c = &host;
  }
  const ice3_params &P = c->params;
  std::string s;
#define D(x) put(s, #x, &P.x, 1)
  D(XTT); D(XLVTT); D(XLSTT); D(XLMTT); D(XCPD); D(XCPV); D(XCL); D(XCI); D(XRD); D(XRV);
  D(XEPSILO); D(XALPW); D(XBETAW); D(XGAMW); D(XALPI); D(XBETAI); D(XGAMI); D(XESTT); D(XP00);
  D(XRHOLW); D(XPI); D(XG); D(XMNH_TINY);
  put(s, "XRTMIN", P.XRTMIN + 1, 7);
  put(s, "XFRMIN", P.XFRMIN + 1, 40);
  D(XCEXVT); D(XLBR); D(XLBEXR); D(XLBS); D(XLBEXS); D(XLBG); D(XLBEXG); D(XLBI); D(XLBEXI);
  D(XLBDAS_MAX); D(XDI); D(XBS); D(XCXS); D(XCXG); D(XDG); D(XCONC_LAND);
  D(XCRIAUTC); D(XCRIAUTI); D(XACRIAUTI); D(XBCRIAUTI); D(XT0CRIAUTI); D(XNU10); D(XALPHA1);
  D(XBETA1); D(XNU20); D(XALPHA2); D(XBETA2); D(XMNU0); D(XALPHA3); D(XBETA3); D(XHON);
  D(XSCFAC); D(X0DEPI); D(X2DEPI); D(X0DEPS); D(X1DEPS); D(XEX0DEPS); D(XEX1DEPS); D(X0DEPG);
  D(X1DEPG); D(XEX0DEPG); D(XEX1DEPG); D(XTIMAUTI); D(XTEXAUTI); D(XCOLEXIS); D(XFIAGGS);
  D(XEXIAGGS); D(XTIMAUTC); D(XFCACCR); D(XEXCACCR); D(X0EVAR); D(X1EVAR); D(XEX0EVAR);
  D(XEX1EVAR); D(XEXCRIMSS); D(XCRIMSS); D(XEXCRIMSG); D(XCRIMSG); D(XSRIMCG); D(XEXSRIMCG);
  D(XSRIMCG2); D(XSRIMCG3); D(XEXSRIMCG2); D(XRIMINTP1); D(XRIMINTP2);
  D(XFRACCSS); D(XLBRACCS1); D(XLBRACCS2); D(XLBRACCS3); D(XFSACCRG); D(XLBSACCR1);
  D(XLBSACCR2); D(XLBSACCR3); D(XACCINTP1S); D(XACCINTP2S); D(XACCINTP1R); D(XACCINTP2R);
  D(XFSCVMG); D(XEXRCFRI); D(XRCFRI); D(XEXICFRR); D(XICFRR); D(XFCDRYG); D(XCOLIG);
  D(XCOLEXIG); D(XFIDRYG); D(XCOLSG); D(XCOLEXSG); D(XFSDRYG); D(XLBSDRYG1); D(XLBSDRYG2);
  D(XLBSDRYG3); D(XFRDRYG); D(XLBRDRYG1); D(XLBRDRYG2); D(XLBRDRYG3); D(XDRYINTP1R);
  D(XDRYINTP2R); D(XDRYINTP1S); D(XDRYINTP2S); D(XDRYINTP1G); D(XDRYINTP2G); D(XFSEDR);
  D(XEXSEDR); D(XFSEDS); D(XEXSEDS); D(XFSEDG); D(XEXSEDG); D(XFSEDI); D(XEXCSEDI);
  D(XLBH); D(XLBEXH); D(XCXH); D(XDH); D(XBG); D(XFSEDH); D(XEXSEDH); D(X0DEPH); D(X1DEPH);
  D(XEX0DEPH); D(XEX1DEPH); D(XCOLIH); D(XCOLEXIH); D(XCOLSH); D(XCOLEXSH); D(XCOLGH);
  D(XCOLEXGH); D(XFWETH); D(XFSWETH); D(XLBSWETH1); D(XLBSWETH2); D(XLBSWETH3); D(XFGWETH);
  D(XLBGWETH1); D(XLBGWETH2); D(XLBGWETH3); D(XFRWETH); D(XLBRWETH1); D(XLBRWETH2);
  D(XLBRWETH3); D(XWETINTP1S); D(XWETINTP2S); D(XWETINTP1G); D(XWETINTP2G); D(XWETINTP1R);
  D(XWETINTP2R); D(XWETINTP1H); D(XWETINTP2H);
#undef D
  put(s, "XGAMINC_RIM1", c->h_tab + ICE3_T_RIM1, 80);
  put(s, "XGAMINC_RIM2", c->h_tab + ICE3_T_RIM2, 80);
  put(s, "XGAMINC_RIM4", c->h_tab + ICE3_T_RIM4, 80);
  put(s, "XKER_RACCSS", c->h_tab + ICE3_T_RACCSS, 1600);
  put(s, "XKER_RACCS", c->h_tab + ICE3_T_RACCS, 1600);
  put(s, "XKER_SACCRG", c->h_tab + ICE3_T_SACCRG, 1600);
  put(s, "XKER_SDRYG", c->h_tab + ICE3_T_SDRYG, 3200);
  put(s, "XKER_RDRYG", c->h_tab + ICE3_T_RDRYG, 1600);
  put(s, "XKER_SWETH", c->h_tab + ICE3_T_SWETH, 3200);
  put(s, "XKER_GWETH", c->h_tab + ICE3_T_GWETH, 1600);
  put(s, "XKER_RWETH", c->h_tab + ICE3_T_RWETH, 1600);
  long n = (long)s.size();
  if (buf && buflen > 0) {
    long m = n < buflen - 1 ? n : buflen - 1;
    memcpy(buf, s.data(), (size_t)m);
    buf[m] = 0;
  }
  return n;
}

static bool bad_shape(int ncol, int nlev, int ld, double dt) {
  return ncol <= 0 || nlev <= 0 || ld < ncol || !(dt > 0.) ||
         (double)ld * (double)nlev > 9.0e15 || (double)ncol * (double)nlev > 4294967295.0;
}

// A domain with no column or no level: a valid call with nothing to do (its field pointers
// may be NULL, as an empty array's are).
static bool empty_domain(int ncol, int nlev, int ld, double dt) {
  return (ncol == 0 || nlev == 0) && ncol >= 0 && nlev >= 0 && ld >= ncol && dt > 0.;
}

static bool known_prec(int prec) { return prec == ICE3_PREC_FP64 || prec == ICE3_PREC_MIXED; }

#define NN(p) if (!(p)) return ICE3_EINVAL

extern "C" int ice3_count_nonfinite(int nfield, const double *const *fields, const int *rows,
                                    int ncol, int ld, unsigned long long *counts, void *stream) {
  if (nfield < 0 || nfield > ICE3_MAX_CHECK_FIELDS || ncol < 0 || ld < ncol) return ICE3_EINVAL;
  if (nfield == 0 || ncol == 0) return ICE3_OK;
  if (!fields || !rows || !counts) return ICE3_EINVAL;
  CountArgs a = CountArgs{};
  a.ncol = ncol; a.ld = ld; a.counts = counts;
  int maxrows = 0;
  for (int q = 0; q < nfield; ++q) {
    if (rows[q] < 0 || (rows[q] > 0 && !fields[q])) return ICE3_EINVAL;
    if ((double)rows[q] * (double)ld > 9.0e15) return ICE3_EINVAL;
    a.f[q] = fields[q];
    a.rows[q] = rows[q];
    maxrows = rows[q] > maxrows ? rows[q] : maxrows;
  }
  if (maxrows == 0) return ICE3_OK;
  const int blocks = (ncol + 255) / 256;
  dim3 grid(blocks < 1024 ? blocks : 1024, nfield);
  hipLaunchKernelGGL(k_count_nonfinite, grid, dim3(256), 0, (hipStream_t)stream, a);
  HIPCHK(hipGetLastError());
  return ICE3_OK;
}

extern "C" int ice3_ice_adjust(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld, double dt,
                               const ice3_ice_adjust_in *in, const ice3_ice_adjust_out *out,
                               void *stream) {
  if (!ctx || !in || !out || !known_prec(prec)) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt)) return ICE3_EINVAL;
  NN(in->pabs); NN(in->sigs); NN(in->th); NN(in->exn); NN(in->exn_ref); NN(in->rho_dry_ref);
  NN(in->rv); NN(in->rc); NN(in->ri); NN(in->rr); NN(in->rs); NN(in->rg); NN(in->cf_mf);
  NN(in->rc_mf); NN(in->ri_mf); NN(in->rvs); NN(in->rcs); NN(in->ris); NN(in->ths);
  NN(in->sigqsat);
  NN(out->t); NN(out->rv_out); NN(out->rc_out); NN(out->ri_out); NN(out->cldfr);
  NN(out->hlc_hrc); NN(out->hlc_hcf); NN(out->hli_hri); NN(out->hli_hcf); NN(out->cph);
  NN(out->lv); NN(out->ls); NN(out->rvs); NN(out->rcs); NN(out->ris); NN(out->ths);
  if (nlev > 65535) return ICE3_EINVAL;
  AdjustArgs a = AdjustArgs{};  // ris_copy stays NULL
  a.ncol = ncol; a.nlev = nlev; a.ld = ld; a.dt = dt; a.in = *in; a.out = *out;
  dim3 grid((ncol + 255) / 256, nlev);
  hipStream_t s = (hipStream_t)stream;
  MARK(0);
  hipLaunchKernelGGL(k_ice_adjust, grid, dim3(256), 0, s, a);
  HIPCHK(hipGetLastError());
  MARK(1); MARK(2); MARK(3); MARK(4); MARK(5); MARK(6);
  return ICE3_OK;
}

extern "C" int ice3_ice_adjust_ocnd2(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                     const ice3_ice_adjust_in *in, const ice3_ocnd2_in *in2,
                                     const ice3_ice_adjust_out *out, const ice3_ocnd2_out *out2,
                                     void *stream) {
  if (!ctx || !in || !in2 || !out || !out2) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt) || nlev < 2 || nlev > 65535) return ICE3_EINVAL;
  NN(in->pabs); NN(in->sigs); NN(in->th); NN(in->exn); NN(in->exn_ref); NN(in->rho_dry_ref);
  NN(in->rv); NN(in->rc); NN(in->ri); NN(in->rr); NN(in->rs); NN(in->rg); NN(in->cf_mf);
  NN(in->rc_mf); NN(in->ri_mf); NN(in->rvs); NN(in->rcs); NN(in->ris); NN(in->ths);
  NN(in->sigqsat); NN(in2->zz); NN(in2->ice_cld_wgt);
  NN(out->t); NN(out->rv_out); NN(out->rc_out); NN(out->ri_out); NN(out->cldfr);
  NN(out->hlc_hrc); NN(out->hlc_hcf); NN(out->hli_hri); NN(out->hli_hcf); NN(out->cph);
  NN(out->lv); NN(out->ls); NN(out->rvs); NN(out->rcs); NN(out->ris); NN(out->ths);
  NN(out2->icldfr); NN(out2->wcldfr); NN(out2->ssio); NN(out2->ssiu); NN(out2->ifr);
  AdjustArgs a = AdjustArgs{};
  a.ncol = ncol; a.nlev = nlev; a.ld = ld; a.dt = dt; a.in = *in; a.out = *out;
  Ocnd2Args x{*in2, *out2, ctx->d_tiwmx};
  dim3 grid((ncol + 255) / 256, nlev);
  hipStream_t s = (hipStream_t)stream;
  MARK(0);
  hipLaunchKernelGGL(k_ice_adjust_ocnd2, grid, dim3(256), 0, s, a, x);
  HIPCHK(hipGetLastError());
  MARK(1); MARK(2); MARK(3); MARK(4); MARK(5); MARK(6);
  return ICE3_OK;
}

// Launch k_micro: one workgroup of eight wavefronts per compute unit (its LDS holds the
// tables and the parked points, so one workgroup fits a CU).
static hipError_t launch_micro(const ice3_ctx *ctx, int prec, const MicroArgs &m, hipStream_t s) {
  if (m.max_calls != DEFAULT_MAX_CALLS) {  // a context-set cap: the point kernel reads it
    const unsigned int npts = (unsigned int)m.nlev * (unsigned int)m.ncol;
    const dim3 g((npts + 255) / 256);
    if (prec == ICE3_PREC_FP64) hipLaunchKernelGGL(k_micro_points<double>, g, dim3(256), 0, s, m);
    else hipLaunchKernelGGL(k_micro_points<float>, g, dim3(256), 0, s, m);
    return hipGetLastError();
  }
  if (prec == ICE3_PREC_FP64)
    hipLaunchKernelGGL(k_micro<double>, dim3(ctx->ncu), dim3(64 * micro_waves<double>()),
                       micro_lds_bytes<double>(), s, m);
  else
    hipLaunchKernelGGL(k_micro<float>, dim3(ctx->ncu), dim3(64 * micro_waves<float>()),
                       micro_lds_bytes<float>(), s, m);
  return hipGetLastError();
}

// The OCND2 microphysics (FP64): one workgroup per compute unit, or the point kernel when the
// context sets another sub-step cap.
static hipError_t launch_micro_ocnd2(const ice3_ctx *ctx, const MicroArgs &m, hipStream_t s) {
  if (m.max_calls != DEFAULT_MAX_CALLS) {
    const unsigned int npts = (unsigned int)m.nlev * (unsigned int)m.ncol;
    hipLaunchKernelGGL((k_micro_points<double, true>), dim3((npts + 255) / 256), dim3(256), 0, s, m);
  } else {
    hipLaunchKernelGGL((k_micro<double, true>), dim3(ctx->ncu), dim3(64 * micro_waves<double>()),
                       micro_lds_bytes<double>(), s, m);
  }
  return hipGetLastError();
}

// Levels a wavefront of k_sedim_split holds (4 slots per lane).
constexpr int SPLIT_MAX_LEVELS = 4 * 64;

// Whether the context's sedimentation scheme can run nlev levels.
static bool sedim_fits(const ice3_ctx *ctx, int nlev) {
  return ctx->sedim != ICE3_SEDIM_SPLI || nlev <= SPLIT_MAX_LEVELS;
}

template <int L, bool HAIL, bool ZZZ>
static void launch_split_k(const SplitArgs &a, hipStream_t s) {
  hipLaunchKernelGGL((k_sedim_split<L, HAIL, ZZZ>), dim3((a.ncol + SPLIT_COLS - 1) / SPLIT_COLS),
                     dim3(64 * SPLIT_COLS), 0, s, a);
}

template <bool HAIL, bool ZZZ>
static void launch_split_l(const SplitArgs &a, hipStream_t s) {
  switch ((a.nlev + 63) / 64) {
    case 1: launch_split_k<1, HAIL, ZZZ>(a, s); break;
    case 2: launch_split_k<2, HAIL, ZZZ>(a, s); break;
    case 3: launch_split_k<3, HAIL, ZZZ>(a, s); break;
    default: launch_split_k<4, HAIL, ZZZ>(a, s); break;
  }
}

// The split sedimentation (CSEDIM='SPLI') of the stages described by sa (sedimentation) and m
// (the state at t and the th, rv, rc sources); sa.zzz selects the OCND2 model level height,
// sa.src_rh hail.  nlev <= SPLIT_MAX_LEVELS (checked by the entry points).
static hipError_t launch_split(const ice3_ctx *ctx, const SedimArgs &sa, const MicroArgs &m,
                               hipStream_t s, double *const *keep = nullptr) {
  SplitArgs a = split_args(sa, m, ctx->maxcfl, ctx->d_capped);
  if (keep)
    for (int q = 0; q < 3; ++q) a.keep[q] = keep[q];
  if (sa.src_rh) launch_split_l<true, false>(a, s);
  else if (sa.zzz) launch_split_l<false, true>(a, s);
  else launch_split_l<false, false>(a, s);
  return hipGetLastError();
}

// The fused step with the split sedimentation: its negativity correction updates the th, rv
// and rc sources, whose values before it, times dt, are the microphysics' state at t.  The
// split kernel stores them as it reads them into the context's work area (3 fields, grown on
// demand; growing it waits for the device), and m then reads its state at t there.
static int launch_split_keep(ice3_ctx *ctx, int nlev, int ld, const SedimArgs &sa, MicroArgs &m,
                             hipStream_t s) {
  const size_t n = (size_t)nlev * ld;
  const size_t bytes = 3 * n * sizeof(double);
  if (bytes > ctx->work_bytes) {
    HIPCHK(hipDeviceSynchronize());
    if (ctx->d_work) HIPCHK(hipFree(ctx->d_work));
    ctx->d_work = nullptr;
    ctx->work_bytes = 0;
    HIPCHK(hipMalloc(&ctx->d_work, bytes));
    ctx->work_bytes = bytes;
  }
  double *w = (double *)ctx->d_work;
  double *const keep[3] = {w, w + n, w + 2 * n};
  HIPCHK(launch_split(ctx, sa, m, s, keep));
  for (int q = 0; q < 3; ++q) m.t[q] = keep[q];
  return ICE3_OK;
}

// Launch the sedimentation then the microphysics stage on one stream.
static int launch_rain_ice(ice3_ctx *ctx, int prec, const SedimArgs &sa, const MicroArgs &ma,
                           void *stream) {
  hipStream_t s = (hipStream_t)stream;
  const dim3 gs((sa.ncol + 255) / 256);
  if (ctx->sedim == ICE3_SEDIM_SPLI) HIPCHK(launch_split(ctx, sa, ma, s));
  else if (prec == ICE3_PREC_FP64) hipLaunchKernelGGL(k_sedim<double>, gs, dim3(256), 0, s, sa);
  else hipLaunchKernelGGL(k_sedim<float>, gs, dim3(256), 0, s, sa);
  HIPCHK(hipGetLastError());
  MARK(2);
  HIPCHK(launch_micro(ctx, prec, ma, s));
  MARK(3); MARK(4); MARK(5); MARK(6);
  return ICE3_OK;
}

extern "C" int ice3_rain_ice(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld, double dt,
                             const ice3_rain_ice_in *in, const ice3_rain_ice_inout *io,
                             const ice3_rain_ice_out *out, const ice3_stats *stats,
                             void *stream) {
  if (!ctx || !in || !io || !out || !known_prec(prec)) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt) || nlev > 65535) return ICE3_EINVAL;
  NN(in->exn); NN(in->dzz); NN(in->rhodref); NN(in->exn_ref); NN(in->pabs); NN(in->cldfr);
  NN(in->sigs); NN(in->tht); NN(in->rvt); NN(in->rct); NN(in->rrt); NN(in->rit); NN(in->rst);
  NN(in->rgt);
  NN(io->ths); NN(io->rvs); NN(io->rcs); NN(io->rrs); NN(io->ris); NN(io->rss); NN(io->rgs);
  NN(io->ci);
  NN(out->evap3d); NN(out->rainfr); NN(out->inprc); NN(out->inprr); NN(out->inpri);
  NN(out->inprs); NN(out->inprg);
  if (!sedim_fits(ctx, nlev)) return ICE3_EINVAL;
  SedimArgs sa;
  MicroArgs m;
  rain_ice_args(ncol, nlev, ld, dt, ctx->d_tab, in, io, out, stats, sa, m);
  m.max_calls = ctx->max_calls; m.capped = ctx->d_capped;
  hipStream_t s = (hipStream_t)stream;
  MARK(0); MARK(1);
  return launch_rain_ice(ctx, prec, sa, m, stream);
}

extern "C" int ice3_rain_ice_ocnd2(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                   const ice3_rain_ice_in *in, const ice3_rain_ocnd2_in *in2,
                                   const ice3_rain_ice_inout *io, const ice3_rain_ice_out *out,
                                   const ice3_stats *stats, void *stream) {
  if (!ctx || !in || !in2 || !io || !out) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt) || nlev > 65535) return ICE3_EINVAL;
  NN(in->exn); NN(in->dzz); NN(in->rhodref); NN(in->exn_ref); NN(in->pabs); NN(in->cldfr);
  NN(in->sigs); NN(in->tht); NN(in->rvt); NN(in->rct); NN(in->rrt); NN(in->rit); NN(in->rst);
  NN(in->rgt); NN(in2->icldfr); NN(in2->ssio); NN(in2->ssiu); NN(in2->ifr);
  NN(io->ths); NN(io->rvs); NN(io->rcs); NN(io->rrs); NN(io->ris); NN(io->rss); NN(io->rgs);
  NN(io->ci);
  NN(out->evap3d); NN(out->rainfr); NN(out->inprc); NN(out->inprr); NN(out->inpri);
  NN(out->inprs); NN(out->inprg);
  SedimArgs sa;
  MicroArgs m;
  rain_ice_args(ncol, nlev, ld, dt, ctx->d_tab, in, io, out, stats, sa, m);
  m.max_calls = ctx->max_calls; m.capped = ctx->d_capped;
  m.icldfr = in2->icldfr; m.ssio = in2->ssio; m.ssiu = in2->ssiu; m.ifr = in2->ifr;
  m.tiwmx = ctx->d_tiwmx;
  sa.zzz = out->rainfr;  // the model level height, read by the microphysics, then overwritten
  if (!sedim_fits(ctx, nlev)) return ICE3_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  MARK(0); MARK(1);
  if (ctx->sedim == ICE3_SEDIM_SPLI)
    HIPCHK(launch_split(ctx, sa, m, s));
  else
    hipLaunchKernelGGL((k_sedim<double, true>), dim3((ncol + 255) / 256), dim3(256), 0, s, sa);
  HIPCHK(hipGetLastError());
  MARK(2);
  HIPCHK(launch_micro_ocnd2(ctx, m, s));
  MARK(3); MARK(4); MARK(5); MARK(6);
  return ICE3_OK;
}

static bool bad_pipeline_ptrs(const ice3_pipeline_in *in, const ice3_pipeline_out *out) {
  const void *const pin[] = {in->pabs, in->sigs, in->th, in->exn, in->exn_ref, in->rhodref,
                             in->dzz, in->rv, in->rc, in->ri, in->rr, in->rs, in->rg, in->ci,
                             in->cf_mf, in->rc_mf, in->ri_mf, in->ths, in->rvs, in->rcs,
                             in->rrs, in->ris, in->rss, in->rgs, in->sigqsat};
  const void *const pout[] = {out->ths, out->rvs, out->rcs, out->rrs, out->ris, out->rss,
                              out->rgs, out->cldfr, out->hlc_hrc, out->hlc_hcf, out->hli_hri,
                              out->hli_hcf, out->ci, out->evap3d, out->rainfr, out->inprc,
                              out->inprr, out->inpri, out->inprs, out->inprg};
  for (const void *p : pin)
    if (!p) return true;
  for (const void *p : pout)
    if (!p) return true;
  return false;
}

extern "C" int ice3_adjust_rain_ice_stages(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld,
                                           double dt, const ice3_pipeline_in *in,
                                           const ice3_pipeline_out *out,
                                           const ice3_stats *stats, int stages, void *stream) {
  if (!ctx || !in || !out || !known_prec(prec) || stages <= 0 || stages > 7) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt) || nlev > 65535) return ICE3_EINVAL;
  if (bad_pipeline_ptrs(in, out)) return ICE3_EINVAL;
  if ((stages & ICE3_STAGE_SEDIMENTATION) && !sedim_fits(ctx, nlev)) return ICE3_EINVAL;
  // with SPLI the microphysics' state at t is kept by the sedimentation stage of its call
  if (ctx->sedim == ICE3_SEDIM_SPLI && (stages & ICE3_STAGE_MICROPHYSICS) &&
      !(stages & ICE3_STAGE_SEDIMENTATION))
    return ICE3_EINVAL;
  AdjustArgs aa;
  SedimArgs sa;
  MicroArgs m;
  pipeline_args(ncol, nlev, ld, dt, ctx->d_tab, in, out, stats, aa, sa, m);
  m.max_calls = ctx->max_calls; m.capped = ctx->d_capped;
  hipStream_t s = (hipStream_t)stream;
  const dim3 gs((ncol + 255) / 256), gp((ncol + 255) / 256, nlev);
  const bool split = ctx->sedim == ICE3_SEDIM_SPLI;
  MARK(0);
  if (prec == ICE3_PREC_FP64 && !split && (stages & ICE3_STAGE_ADJUST) &&
      (stages & ICE3_STAGE_SEDIMENTATION)) {
    // one column walk for both stages (in the mixed path the float sedimentation is cheap
    // and the two kernels are faster apart)
    hipLaunchKernelGGL(k_adjust_sedim, dim3((ncol + 63) / 64), dim3(64), 0, s, aa, sa);
    HIPCHK(hipGetLastError());
    stages &= ~ICE3_STAGE_SEDIMENTATION;
  } else if (stages & ICE3_STAGE_ADJUST) {
    hipLaunchKernelGGL(k_ice_adjust, gp, dim3(256), 0, s, aa);
    HIPCHK(hipGetLastError());
  }
  MARK(1);
  if (stages & ICE3_STAGE_SEDIMENTATION) {
    if (split && (stages & ICE3_STAGE_MICROPHYSICS)) {
      const int rc = launch_split_keep(ctx, nlev, ld, sa, m, s);
      if (rc != ICE3_OK) return rc;
    } else if (split) {
      HIPCHK(launch_split(ctx, sa, m, s));
    } else if (prec == ICE3_PREC_FP64) hipLaunchKernelGGL(k_sedim<double>, gs, dim3(256), 0, s, sa);
    else hipLaunchKernelGGL(k_sedim<float>, gs, dim3(256), 0, s, sa);
    HIPCHK(hipGetLastError());
  }
  MARK(2);
  if (stages & ICE3_STAGE_MICROPHYSICS) {
    HIPCHK(launch_micro(ctx, prec, m, s));
  }
  MARK(3); MARK(4); MARK(5); MARK(6);
  return ICE3_OK;
}

extern "C" int ice3_adjust_rain_ice(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld,
                                    double dt, const ice3_pipeline_in *in,
                                    const ice3_pipeline_out *out, const ice3_stats *stats,
                                    void *stream) {
  return ice3_adjust_rain_ice_stages(ctx, prec, ncol, nlev, ld, dt, in, out, stats,
                                     ICE3_STAGE_ADJUST | ICE3_STAGE_SEDIMENTATION |
                                         ICE3_STAGE_MICROPHYSICS,
                                     stream);
}

extern "C" int ice3_adjust_rain_ice_ocnd2(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                          const ice3_pipeline_in *in, const ice3_ocnd2_in *in2,
                                          const ice3_pipeline_out *out,
                                          const ice3_ocnd2_out *out2, const ice3_stats *stats,
                                          void *stream) {
  if (!ctx || !in || !in2 || !out || !out2) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt) || nlev < 2 || nlev > 65535) return ICE3_EINVAL;
  if (bad_pipeline_ptrs(in, out)) return ICE3_EINVAL;
  NN(in2->zz); NN(in2->ice_cld_wgt);
  NN(out2->icldfr); NN(out2->wcldfr); NN(out2->ssio); NN(out2->ssiu); NN(out2->ifr);
  AdjustArgs aa;
  SedimArgs sa;
  MicroArgs m;
  pipeline_args(ncol, nlev, ld, dt, ctx->d_tab, in, out, stats, aa, sa, m);
  m.max_calls = ctx->max_calls; m.capped = ctx->d_capped;
  m.icldfr = out2->icldfr; m.ssio = out2->ssio; m.ssiu = out2->ssiu; m.ifr = out2->ifr;
  m.tiwmx = ctx->d_tiwmx;
  sa.zzz = out->rainfr;  // the model level height, read by the microphysics, then overwritten
  if (!sedim_fits(ctx, nlev)) return ICE3_EINVAL;
  const Ocnd2Args x{*in2, *out2, ctx->d_tiwmx};
  hipStream_t s = (hipStream_t)stream;
  MARK(0);
  hipLaunchKernelGGL(k_ice_adjust_ocnd2, dim3((ncol + 255) / 256, nlev), dim3(256), 0, s, aa, x);
  HIPCHK(hipGetLastError());
  MARK(1);
  if (ctx->sedim == ICE3_SEDIM_SPLI) {
    const int rc = launch_split_keep(ctx, nlev, ld, sa, m, s);
    if (rc != ICE3_OK) return rc;
  } else {
    hipLaunchKernelGGL((k_sedim<double, true>), dim3((ncol + 255) / 256), dim3(256), 0, s, sa);
  }
  HIPCHK(hipGetLastError());
  MARK(2);
  HIPCHK(launch_micro_ocnd2(ctx, m, s));
  MARK(3); MARK(4); MARK(5); MARK(6);
  return ICE3_OK;
}

// ---- hail (KRR=7) entry points, FP64

extern "C" int ice3_ice_adjust_hail(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                    const ice3_ice_adjust_in *in, const double *rh,
                                    const ice3_ice_adjust_out *out, void *stream) {
  if (!ctx || !in || !out) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt) || nlev > 65535) return ICE3_EINVAL;
  NN(in->pabs); NN(in->sigs); NN(in->th); NN(in->exn); NN(in->exn_ref); NN(in->rho_dry_ref);
  NN(in->rv); NN(in->rc); NN(in->ri); NN(in->rr); NN(in->rs); NN(in->rg); NN(in->cf_mf);
  NN(in->rc_mf); NN(in->ri_mf); NN(in->rvs); NN(in->rcs); NN(in->ris); NN(in->ths);
  NN(in->sigqsat); NN(rh);
  NN(out->t); NN(out->rv_out); NN(out->rc_out); NN(out->ri_out); NN(out->cldfr);
  NN(out->hlc_hrc); NN(out->hlc_hcf); NN(out->hli_hri); NN(out->hli_hcf); NN(out->cph);
  NN(out->lv); NN(out->ls); NN(out->rvs); NN(out->rcs); NN(out->ris); NN(out->ths);
  AdjustArgs a = AdjustArgs{};
  a.ncol = ncol; a.nlev = nlev; a.ld = ld; a.dt = dt; a.in = *in; a.out = *out; a.rh = rh;
  hipStream_t s = (hipStream_t)stream;
  MARK(0);
  hipLaunchKernelGGL(k_ice_adjust_hail, dim3((ncol + 255) / 256, nlev), dim3(256), 0, s, a);
  HIPCHK(hipGetLastError());
  MARK(1); MARK(2); MARK(3); MARK(4); MARK(5); MARK(6);
  return ICE3_OK;
}

// The hail microphysics (FP64): one workgroup per compute unit, or the point kernel when the
// context sets another sub-step cap (it reads the cap at run time).
static hipError_t launch_micro_hail(const ice3_ctx *ctx, const MicroArgs &m, hipStream_t s) {
  if (m.max_calls != DEFAULT_MAX_CALLS) {
    const unsigned int npts = (unsigned int)m.nlev * (unsigned int)m.ncol;
    hipLaunchKernelGGL((k_micro_points<double, false, true>), dim3((npts + 255) / 256), dim3(256),
                       0, s, m);
  } else {
    hipLaunchKernelGGL((k_micro<double, false, true>), dim3(ctx->ncu),
                       dim3(64 * (micro_waves<double, true>())), (micro_lds_bytes<double, true>()), s, m);
  }
  return hipGetLastError();
}

extern "C" int ice3_ice_adjust_branches(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                        const ice3_ice_adjust_in *in, const double *rh,
                                        const ice3_ice_adjust_out *out, unsigned *branches,
                                        void *stream) {
  if (!ctx || !in || !out || !branches) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt) || nlev > 65535) return ICE3_EINVAL;
  NN(in->pabs); NN(in->sigs); NN(in->th); NN(in->exn); NN(in->exn_ref); NN(in->rho_dry_ref);
  NN(in->rv); NN(in->rc); NN(in->ri); NN(in->rr); NN(in->rs); NN(in->rg); NN(in->cf_mf);
  NN(in->rc_mf); NN(in->ri_mf); NN(in->rvs); NN(in->rcs); NN(in->ris); NN(in->ths);
  NN(in->sigqsat);
  NN(out->t); NN(out->rv_out); NN(out->rc_out); NN(out->ri_out); NN(out->cldfr);
  NN(out->hlc_hrc); NN(out->hlc_hcf); NN(out->hli_hri); NN(out->hli_hcf); NN(out->cph);
  NN(out->lv); NN(out->ls); NN(out->rvs); NN(out->rcs); NN(out->ris); NN(out->ths);
  AdjustArgs a = AdjustArgs{};
  a.ncol = ncol; a.nlev = nlev; a.ld = ld; a.dt = dt; a.in = *in; a.out = *out; a.rh = rh;
  const dim3 grid((ncol + 255) / 256, nlev);
  hipStream_t s = (hipStream_t)stream;
  if (rh)
    hipLaunchKernelGGL(k_ice_adjust_branches<true>, grid, dim3(256), 0, s, a, branches);
  else
    hipLaunchKernelGGL(k_ice_adjust_branches<false>, grid, dim3(256), 0, s, a, branches);
  HIPCHK(hipGetLastError());
  return ICE3_OK;
}

extern "C" int ice3_rain_ice_branches(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                      const ice3_rain_ice_in *in, const ice3_rain_ice_inout *io,
                                      const ice3_rain_ice_out *out, const ice3_hail *hail,
                                      unsigned *branches, void *stream) {
  if (!ctx || !in || !io || !out || !branches) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt) || nlev > 65535) return ICE3_EINVAL;
  NN(in->exn); NN(in->dzz); NN(in->rhodref); NN(in->exn_ref); NN(in->pabs); NN(in->cldfr);
  NN(in->sigs); NN(in->tht); NN(in->rvt); NN(in->rct); NN(in->rrt); NN(in->rit); NN(in->rst);
  NN(in->rgt);
  NN(io->ths); NN(io->rvs); NN(io->rcs); NN(io->rrs); NN(io->ris); NN(io->rss); NN(io->rgs);
  NN(io->ci);
  NN(out->evap3d); NN(out->rainfr); NN(out->inprc); NN(out->inprr); NN(out->inpri);
  NN(out->inprs); NN(out->inprg);
  if (hail) { NN(hail->rht); NN(hail->rhs); NN(hail->inprh); }
  if (!sedim_fits(ctx, nlev)) return ICE3_EINVAL;
  SedimArgs sa;
  MicroArgs m;
  rain_ice_args(ncol, nlev, ld, dt, ctx->d_tab, in, io, out, nullptr, sa, m);
  m.max_calls = ctx->max_calls; m.capped = ctx->d_capped;
  m.branches = branches;
  if (hail) {
    sa.src_rh = hail->rhs; sa.dst_rh = hail->rhs; sa.inprh = hail->inprh;
    m.t[IRH] = hail->rht; m.src[IRH] = hail->rhs; m.dst[IRH] = hail->rhs;
  }
  hipStream_t s = (hipStream_t)stream;
  if (ctx->sedim == ICE3_SEDIM_SPLI)
    HIPCHK(launch_split(ctx, sa, m, s));
  else if (hail)
    hipLaunchKernelGGL(k_sedim_hail, dim3((ncol + 255) / 256), dim3(256), 0, s, sa);
  else
    hipLaunchKernelGGL(k_sedim<double>, dim3((ncol + 255) / 256), dim3(256), 0, s, sa);
  HIPCHK(hipGetLastError());
  const unsigned int npts = (unsigned int)nlev * (unsigned int)ncol;
  if (hail)
    hipLaunchKernelGGL(k_micro_branches<true>, dim3((npts + 255) / 256), dim3(256), 0, s, m);
  else
    hipLaunchKernelGGL(k_micro_branches<false>, dim3((npts + 255) / 256), dim3(256), 0, s, m);
  HIPCHK(hipGetLastError());
  return ICE3_OK;
}

extern "C" int ice3_rain_ice_hail(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                  const ice3_rain_ice_in *in, const ice3_rain_ice_inout *io,
                                  const ice3_rain_ice_out *out, const ice3_hail *hail,
                                  const ice3_stats *stats, void *stream) {
  if (!ctx || !in || !io || !out || !hail) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt) || nlev > 65535) return ICE3_EINVAL;
  NN(in->exn); NN(in->dzz); NN(in->rhodref); NN(in->exn_ref); NN(in->pabs); NN(in->cldfr);
  NN(in->sigs); NN(in->tht); NN(in->rvt); NN(in->rct); NN(in->rrt); NN(in->rit); NN(in->rst);
  NN(in->rgt);
  NN(io->ths); NN(io->rvs); NN(io->rcs); NN(io->rrs); NN(io->ris); NN(io->rss); NN(io->rgs);
  NN(io->ci);
  NN(out->evap3d); NN(out->rainfr); NN(out->inprc); NN(out->inprr); NN(out->inpri);
  NN(out->inprs); NN(out->inprg);
  NN(hail->rht); NN(hail->rhs); NN(hail->inprh);
  SedimArgs sa;
  MicroArgs m;
  rain_ice_args(ncol, nlev, ld, dt, ctx->d_tab, in, io, out, stats, sa, m);
  m.max_calls = ctx->max_calls; m.capped = ctx->d_capped;
  sa.src_rh = hail->rhs; sa.dst_rh = hail->rhs; sa.inprh = hail->inprh;
  m.t[IRH] = hail->rht; m.src[IRH] = hail->rhs; m.dst[IRH] = hail->rhs;
  if (!sedim_fits(ctx, nlev)) return ICE3_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  MARK(0); MARK(1);
  if (ctx->sedim == ICE3_SEDIM_SPLI)
    HIPCHK(launch_split(ctx, sa, m, s));
  else
    hipLaunchKernelGGL(k_sedim_hail, dim3((ncol + 255) / 256), dim3(256), 0, s, sa);
  HIPCHK(hipGetLastError());
  MARK(2);
  HIPCHK(launch_micro_hail(ctx, m, s));
  MARK(3); MARK(4); MARK(5); MARK(6);
  return ICE3_OK;
}

extern "C" int ice3_adjust_rain_ice_hail(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                         const ice3_pipeline_in *in, const ice3_pipeline_out *out,
                                         const ice3_pipeline_hail *hail, const ice3_stats *stats,
                                         void *stream) {
  if (!ctx || !in || !out || !hail) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt) || nlev > 65535) return ICE3_EINVAL;
  if (bad_pipeline_ptrs(in, out)) return ICE3_EINVAL;
  NN(hail->rh); NN(hail->rhs); NN(hail->rhs_out); NN(hail->inprh);
  if ((const double *)hail->rhs_out == hail->rhs) return ICE3_EINVAL;
  AdjustArgs aa;
  SedimArgs sa;
  MicroArgs m;
  pipeline_args(ncol, nlev, ld, dt, ctx->d_tab, in, out, stats, aa, sa, m);
  m.max_calls = ctx->max_calls; m.capped = ctx->d_capped;
  aa.rh = hail->rh;
  sa.src_rh = hail->rhs; sa.dst_rh = hail->rhs_out; sa.inprh = hail->inprh;
  m.t[IRH] = hail->rhs; m.src[IRH] = hail->rhs_out; m.dst[IRH] = hail->rhs_out;
  if (!sedim_fits(ctx, nlev)) return ICE3_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  MARK(0);
  if (ctx->sedim == ICE3_SEDIM_SPLI) {
    hipLaunchKernelGGL(k_ice_adjust_hail, dim3((ncol + 255) / 256, nlev), dim3(256), 0, s, aa);
    HIPCHK(hipGetLastError());
    MARK(1);
    const int rc = launch_split_keep(ctx, nlev, ld, sa, m, s);
    if (rc != ICE3_OK) return rc;
  } else {
    hipLaunchKernelGGL(k_adjust_sedim_hail, dim3((ncol + 63) / 64), dim3(64), 0, s, aa, sa);
    HIPCHK(hipGetLastError());
    MARK(1);
  }
  MARK(2);
  HIPCHK(launch_micro_hail(ctx, m, s));
  MARK(3); MARK(4); MARK(5); MARK(6);
  return ICE3_OK;
}

static int n_cus(int device) {
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      n <= 0)
    n = 256;
  return n;
}

extern "C" size_t ice3_surrogate_work_bytes(int ncol, int nlev) {
  if (ncol <= 0 || nlev <= 0) return 0;
  return sg_work_layout((size_t)ncol * nlev).total;
}

// Argument checks of the surrogate step (both entry points); ICE3_OK with *empty set when
// there is nothing to do.
static int surrogate_check(const ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                           const ice3_pipeline_in *in, const ice3_pipeline_out *out,
                           const ice3_surrogate *sg, const unsigned long long *sg_counters,
                           bool *empty) {
  *empty = false;
  if (!ctx || !in || !out || !sg || !sg_counters) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) {  // the work area is then 0 bytes
    *empty = true;
    return ICE3_OK;
  }
  if (bad_shape(ncol, nlev, ld, dt)) return ICE3_EINVAL;
  if (nlev > 65535 || (double)nlev * ncol >= 4.0e9 || (double)nlev * ld >= 4.0e9)
    return ICE3_EINVAL;
  if (!sg->weights_packed || !sg->params || !sg->norm || !(sg->budget_tol >= 0.f) ||
      !(sg->neg_tol >= 0.f))
    return ICE3_EINVAL;
  if (ctx->sedim != ICE3_SEDIM_STAT) return ICE3_EINVAL;  // trained on the statistical scheme
  if (bad_pipeline_ptrs(in, out)) return ICE3_EINVAL;
  return ICE3_OK;
}

// The surrogate step on a checked problem, with the scratch at work (sg_work_layout).
static int surrogate_step(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                          const ice3_pipeline_in *in, const ice3_pipeline_out *out,
                          const ice3_surrogate *sg, void *work,
                          unsigned long long *sg_counters, const ice3_stats *stats,
                          void *stream) {
  AdjustArgs aa;
  SedimArgs sa;
  MicroArgs m;
  pipeline_args(ncol, nlev, ld, dt, ctx->d_tab, in, out, stats, aa, sa, m);
  m.max_calls = ctx->max_calls; m.capped = ctx->d_capped;
  hipStream_t s = (hipStream_t)stream;
  const unsigned int npts = (unsigned int)nlev * (unsigned int)ncol;
  const SgWork wl = sg_work_layout(npts);
  unsigned char *wb = (unsigned char *)work;
  SurrogateArgs g;
  g.m = m;
  g.npts = npts;
  g.norm = sg->norm;
  g.budget_tol = sg->budget_tol;
  g.neg_tol = sg->neg_tol;
  g.x = (__bf16 *)(wb + wl.x);
  g.y = (float *)(wb + wl.y);
  g.flag = wb + wl.flag;
  g.queue = (unsigned int *)(wb + wl.queue);
  g.counters = sg_counters;
  const int ncu = n_cus(ctx->device);
  const unsigned int ntiles = (npts + MLP_TILE - 1) / MLP_TILE;
  const unsigned int nblk = ntiles < (unsigned int)ncu ? ntiles : (unsigned int)ncu;
  MARK(0);
  hipLaunchKernelGGL(k_adjust_sedim, dim3((ncol + 63) / 64), dim3(64), 0, s, aa, sa);
  HIPCHK(hipGetLastError());
  MARK(1);
  MARK(2);
  hipLaunchKernelGGL(k_sg_features, dim3((npts + 255) / 256), dim3(256), 0, s, g);
  HIPCHK(hipGetLastError());
  MARK(3);
  HIPCHK(mlp_infer_launch(sg->weights_packed, sg->params, g.x, g.y, npts, nblk, s));
  MARK(4);
  // decode: a persistent grid of SG_DEC_BLOCKS blocks; block b decodes chunks
  // b, b + ndec, ... of 256 points and owns a queue segment of seg entries
  unsigned int *qcount = (unsigned int *)(wb + wl.qcount);
  const unsigned int nchunks = (npts + 255) / 256;
  const unsigned int ndec = nchunks < SG_DEC_BLOCKS ? nchunks : SG_DEC_BLOCKS;
  const unsigned int seg = ((nchunks + ndec - 1) / ndec) * 256;
  hipLaunchKernelGGL(k_sg_decode, dim3(ndec), dim3(256), 0, s, g, seg, qcount);
  HIPCHK(hipGetLastError());
  MARK(5);
  hipLaunchKernelGGL(k_micro_segments<double>, dim3(ndec), dim3(64), 0, s, m,
                     (const unsigned int *)g.queue, (const unsigned int *)qcount, seg);
  HIPCHK(hipGetLastError());
  MARK(6);
  return ICE3_OK;
}

extern "C" int ice3_adjust_rain_ice_surrogate(ice3_ctx *ctx, int ncol, int nlev, int ld,
                                              double dt, const ice3_pipeline_in *in,
                                              const ice3_pipeline_out *out,
                                              const ice3_surrogate *sg, void *work,
                                              unsigned long long *sg_counters,
                                              const ice3_stats *stats, void *stream) {
  bool empty;
  const int rc = surrogate_check(ctx, ncol, nlev, ld, dt, in, out, sg, sg_counters, &empty);
  if (rc != ICE3_OK || empty) return rc;
  if (!work || ((uintptr_t)work & 255) != 0) return ICE3_EINVAL;
  return surrogate_step(ctx, ncol, nlev, ld, dt, in, out, sg, work, sg_counters, stats, stream);
}

extern "C" int ice3_adjust_rain_ice_surrogate_noalloc(ice3_ctx *ctx, int ncol, int nlev, int ld,
                                                      double dt, const ice3_pipeline_in *in,
                                                      const ice3_pipeline_out *out,
                                                      const ice3_surrogate *sg,
                                                      unsigned long long *sg_counters,
                                                      const ice3_stats *stats, void *stream) {
  bool empty;
  const int rc = surrogate_check(ctx, ncol, nlev, ld, dt, in, out, sg, sg_counters, &empty);
  if (rc != ICE3_OK || empty) return rc;
  if (sg_work_layout((size_t)ncol * nlev).total > ICE3_SURROGATE_SCRATCH_BYTES) return ICE3_ETOOLARGE;
  return surrogate_step(ctx, ncol, nlev, ld, dt, in, out, sg, ctx->d_scratch, sg_counters, stats,
                        stream);
}

extern "C" int ice3_surrogate_features(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                       const ice3_pipeline_in *in, const ice3_pipeline_out *out,
                                       const ice3_surrogate_norm *norm, float *raw, void *x,
                                       unsigned char *micro, void *stream) {
  if (!ctx || !in || !out || (x && !norm)) return ICE3_EINVAL;
  if (empty_domain(ncol, nlev, ld, dt)) return ICE3_OK;
  if (bad_shape(ncol, nlev, ld, dt)) return ICE3_EINVAL;
  if ((double)nlev * ncol >= 4.0e9 || bad_pipeline_ptrs(in, out)) return ICE3_EINVAL;
  AdjustArgs aa;
  SedimArgs sa;
  MicroArgs m;
  pipeline_args(ncol, nlev, ld, dt, ctx->d_tab, in, out, nullptr, aa, sa, m);
  const unsigned int npts = (unsigned int)nlev * (unsigned int)ncol;
  hipLaunchKernelGGL(k_features, dim3((npts + 255) / 256), dim3(256), 0, (hipStream_t)stream, m,
                     npts, norm, raw, (__bf16 *)x, micro);
  HIPCHK(hipGetLastError());
  return ICE3_OK;
}
