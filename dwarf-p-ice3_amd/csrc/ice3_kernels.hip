// ice3_kernels.hip -- HIP kernels for gfx950 and the C ABI of libice3_hip.so
// (declared in include/ice3.h).
//
// Layout: every field is float64 (nlev, ld), columns contiguous, so a 64-lane
// wavefront that owns 64 consecutive columns reads one level of a field as a
// single 512-byte coalesced transaction.
//
// Kernels:
//   k_ice_adjust   one thread per (level, column) point, pointwise ICE_ADJUST;
//   k_sedim<R>     one thread per column, walking the levels top-down: the
//                  statistical sedimentation flux recurrence, whose fluxes
//                  never leave registers, and the surface precipitation;
//   k_micro<R>     one thread per (level, column) point: the ICE4 time-split
//                  microphysics, total tendencies and negativity correction.
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

namespace ice3 {

// ------------------------------------------------------------------ kernels

__global__ void __launch_bounds__(256) k_ice_adjust(AdjustArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  adjust_at(a, i, (size_t)blockIdx.y * a.ld + i);
}

template <class R>
__global__ void __launch_bounds__(256) k_sedim(SedimArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  sedim_column<R>(a, i);
}

// Sum of v over the 64 lanes of the wavefront (every lane gets the total).
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

template <class R>
__global__ void __launch_bounds__(256) k_micro(MicroArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  int calls = 0;
  if (i < a.ncol) calls = micro_at<R>(a, (size_t)k * a.ld + i);
  if (a.counters) {  // uniform branch: every lane of the wave takes part in the sums
    const int n = wave_sum(calls < 0 ? -calls : calls);
    const int c = wave_sum(calls < 0 ? 1 : 0);
    const int p = wave_sum(calls != 0 ? 1 : 0);
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&a.counters[0], (unsigned long long)n);
      atomicAdd(&a.counters[1], (unsigned long long)c);
      atomicAdd(&a.counters[2], (unsigned long long)p);
    }
  }
}

}  // namespace ice3

// ------------------------------------------------------------------ C ABI

struct ice3_ctx {
  int device;
  ice3_params params;
  double *d_tab;
  double h_tab[ICE3_T_TOTAL];
  int timing;         // record ev[] around the stages of each call
  hipEvent_t ev[4];   // before adjust, after adjust, after sedimentation, after microphysics
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
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_P), &c->params, sizeof(ice3_params)) != hipSuccess ||
      hipMalloc((void **)&c->d_tab, sizeof(double) * ICE3_T_TOTAL) != hipSuccess ||
      hipMemcpy(c->d_tab, c->h_tab, sizeof(double) * ICE3_T_TOTAL, hipMemcpyHostToDevice) !=
          hipSuccess) {
    if (c->d_tab) (void)hipFree(c->d_tab);
    free(c);
    return ICE3_EDEVICE;
  }
  *out = c;
  return ICE3_OK;
}

extern "C" void ice3_ctx_destroy(ice3_ctx *c) {
  if (!c) return;
  if (c->d_tab) (void)hipFree(c->d_tab);
  if (c->timing)
    for (int q = 0; q < 4; ++q) (void)hipEventDestroy(c->ev[q]);
  free(c);
}

extern "C" int ice3_stage_timing(ice3_ctx *c, int enable) {
  if (!c) return ICE3_EINVAL;
  if (enable && !c->timing) {
    for (int q = 0; q < 4; ++q) HIPCHK(hipEventCreate(&c->ev[q]));
    c->timing = 1;
  } else if (!enable && c->timing) {
    for (int q = 0; q < 4; ++q) (void)hipEventDestroy(c->ev[q]);
    c->timing = 0;
  }
  return ICE3_OK;
}

extern "C" int ice3_stage_times(ice3_ctx *c, float ms[3]) {
  if (!c || !ms || !c->timing) return ICE3_EINVAL;
  HIPCHK(hipEventSynchronize(c->ev[3]));
  for (int q = 0; q < 3; ++q) HIPCHK(hipEventElapsedTime(&ms[q], c->ev[q], c->ev[q + 1]));
  return ICE3_OK;
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
  put(s, "XRTMIN", P.XRTMIN + 1, 6);
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
#undef D
  put(s, "XGAMINC_RIM1", c->h_tab + ICE3_T_RIM1, 80);
  put(s, "XGAMINC_RIM2", c->h_tab + ICE3_T_RIM2, 80);
  put(s, "XGAMINC_RIM4", c->h_tab + ICE3_T_RIM4, 80);
  put(s, "XKER_RACCSS", c->h_tab + ICE3_T_RACCSS, 1600);
  put(s, "XKER_RACCS", c->h_tab + ICE3_T_RACCS, 1600);
  put(s, "XKER_SACCRG", c->h_tab + ICE3_T_SACCRG, 1600);
  put(s, "XKER_SDRYG", c->h_tab + ICE3_T_SDRYG, 3200);
  put(s, "XKER_RDRYG", c->h_tab + ICE3_T_RDRYG, 1600);
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
         (double)ld * (double)nlev > 9.0e15;
}

#define NN(p) if (!(p)) return ICE3_EINVAL

extern "C" int ice3_ice_adjust(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld, double dt,
                               const ice3_ice_adjust_in *in, const ice3_ice_adjust_out *out,
                               void *stream) {
  if (!ctx || !in || !out || bad_shape(ncol, nlev, ld, dt)) return ICE3_EINVAL;
  if (prec != ICE3_PREC_FP64 && prec != ICE3_PREC_MIXED) return ICE3_EINVAL;
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
  MARK(1); MARK(2); MARK(3);
  return ICE3_OK;
}

// Launch the sedimentation then the microphysics stage on one stream.
static int launch_rain_ice(ice3_ctx *ctx, int prec, const SedimArgs &sa, const MicroArgs &ma,
                           void *stream) {
  hipStream_t s = (hipStream_t)stream;
  const dim3 gs((sa.ncol + 255) / 256), gm((ma.ncol + 255) / 256, ma.nlev);
  if (prec == ICE3_PREC_FP64) hipLaunchKernelGGL(k_sedim<double>, gs, dim3(256), 0, s, sa);
  else hipLaunchKernelGGL(k_sedim<float>, gs, dim3(256), 0, s, sa);
  HIPCHK(hipGetLastError());
  MARK(2);
  if (prec == ICE3_PREC_FP64) hipLaunchKernelGGL(k_micro<double>, gm, dim3(256), 0, s, ma);
  else hipLaunchKernelGGL(k_micro<float>, gm, dim3(256), 0, s, ma);
  HIPCHK(hipGetLastError());
  MARK(3);
  return ICE3_OK;
}

extern "C" int ice3_rain_ice(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld, double dt,
                             const ice3_rain_ice_in *in, const ice3_rain_ice_inout *io,
                             const ice3_rain_ice_out *out, const ice3_stats *stats,
                             void *stream) {
  if (!ctx || !in || !io || !out || bad_shape(ncol, nlev, ld, dt)) return ICE3_EINVAL;
  if (prec != ICE3_PREC_FP64 && prec != ICE3_PREC_MIXED) return ICE3_EINVAL;
  if (nlev > 65535) return ICE3_EINVAL;
  NN(in->exn); NN(in->dzz); NN(in->rhodref); NN(in->exn_ref); NN(in->pabs); NN(in->cldfr);
  NN(in->sigs); NN(in->tht); NN(in->rvt); NN(in->rct); NN(in->rrt); NN(in->rit); NN(in->rst);
  NN(in->rgt);
  NN(io->ths); NN(io->rvs); NN(io->rcs); NN(io->rrs); NN(io->ris); NN(io->rss); NN(io->rgs);
  NN(io->ci);
  NN(out->evap3d); NN(out->rainfr); NN(out->inprc); NN(out->inprr); NN(out->inpri);
  NN(out->inprs); NN(out->inprg);
  SedimArgs sa;
  MicroArgs m;
  rain_ice_args(ncol, nlev, ld, dt, ctx->d_tab, in, io, out, stats, sa, m);
  hipStream_t s = (hipStream_t)stream;
  MARK(0); MARK(1);
  return launch_rain_ice(ctx, prec, sa, m, stream);
}

extern "C" int ice3_adjust_rain_ice(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld,
                                    double dt, const ice3_pipeline_in *in,
                                    const ice3_pipeline_out *out, const ice3_stats *stats,
                                    void *stream) {
  if (!ctx || !in || !out || bad_shape(ncol, nlev, ld, dt)) return ICE3_EINVAL;
  if (prec != ICE3_PREC_FP64 && prec != ICE3_PREC_MIXED) return ICE3_EINVAL;
  if (nlev > 65535) return ICE3_EINVAL;
  NN(in->pabs); NN(in->sigs); NN(in->th); NN(in->exn); NN(in->exn_ref); NN(in->rhodref);
  NN(in->dzz); NN(in->rv); NN(in->rc); NN(in->ri); NN(in->rr); NN(in->rs);
  NN(in->rg); NN(in->ci); NN(in->cf_mf); NN(in->rc_mf); NN(in->ri_mf); NN(in->ths);
  NN(in->rvs); NN(in->rcs); NN(in->rrs); NN(in->ris); NN(in->rss); NN(in->rgs); NN(in->sigqsat);
  NN(out->ths); NN(out->rvs); NN(out->rcs); NN(out->rrs); NN(out->ris); NN(out->rss);
  NN(out->rgs); NN(out->cldfr); NN(out->hlc_hrc); NN(out->hlc_hcf); NN(out->hli_hri);
  NN(out->hli_hcf); NN(out->ci); NN(out->evap3d); NN(out->rainfr); NN(out->inprc);
  NN(out->inprr); NN(out->inpri); NN(out->inprs); NN(out->inprg);
  AdjustArgs aa;
  SedimArgs sa;
  MicroArgs m;
  pipeline_args(ncol, nlev, ld, dt, ctx->d_tab, in, out, stats, aa, sa, m);
  hipStream_t s = (hipStream_t)stream;
  MARK(0);
  hipLaunchKernelGGL(k_ice_adjust, dim3((ncol + 255) / 256, nlev), dim3(256), 0, s, aa);
  HIPCHK(hipGetLastError());
  MARK(1);
  return launch_rain_ice(ctx, prec, sa, m, stream);
}
