// ice3_kernels.hip -- HIP kernels for gfx950 and the C ABI of libice3_hip.so
// (declared in include/ice3.h).
//
// Layout: every field is float64 (nlev, ld), columns contiguous, so a 64-lane
// wavefront that owns 64 consecutive columns reads one level of a field as a
// single 512-byte coalesced transaction.
//
// Kernels:
//   k_ice_adjust       one thread per (level, column) point, pointwise ICE_ADJUST;
//   k_column<R, ADJ>   one thread per column, walking the levels top-down in
//                      registers: [ICE_ADJUST ->] statistical sedimentation
//                      (a top-down flux recurrence whose fluxes never leave
//                      registers) -> ICE4 time-split microphysics -> total
//                      tendencies -> negativity correction.  Sedimentation at
//                      level k only needs the fluxes from the levels above, so
//                      the microphysics of level k can run as soon as its
//                      sedimentation is done and the whole rain_ice (or the
//                      fused ice_adjust + rain_ice) is one pass over HBM.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/ice3.h"
#include "ice3_params.h"
#include "ice3_physics.hpp"

namespace ice3 {

// ------------------------------------------------------------------ kernels

struct AdjustArgs {
  int ncol, nlev, ld;
  double dt;
  ice3_ice_adjust_in in;
  ice3_ice_adjust_out out;
};

__global__ void __launch_bounds__(256) k_ice_adjust(AdjustArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= a.ncol) return;
  const size_t j = (size_t)k * a.ld + i;
  const ice3_ice_adjust_in &in = a.in;
  double rvs = in.rvs[j], rcs = in.rcs[j], ris = in.ris[j], ths = in.ths[j];
  const double exn = in.exn[j];
  AdjustOut o;
  ice_adjust_point(a.dt, in.sigqsat[i], in.pabs[j], in.sigs[j], in.th[j], exn, in.exn_ref[j],
                   in.rho_dry_ref[j], in.rv[j], in.rc[j], in.ri[j], in.rr[j], in.rs[j], in.rg[j],
                   in.cf_mf[j], in.rc_mf[j], in.ri_mf[j], rvs, rcs, ris, ths, o);
  const ice3_ice_adjust_out &out = a.out;
  out.t[j] = o.t;
  out.rv_out[j] = o.rv;
  out.rc_out[j] = o.rc;
  out.ri_out[j] = o.ri;
  out.cldfr[j] = o.cldfr;
  out.hlc_hrc[j] = o.hlc_hrc;
  out.hlc_hcf[j] = o.hlc_hcf;
  out.hli_hri[j] = o.hli_hri;
  out.hli_hcf[j] = o.hli_hcf;
  out.cph[j] = o.cph;
  out.lv[j] = o.lv;
  out.ls[j] = o.ls;
  out.rvs[j] = rvs;
  out.rcs[j] = rcs;
  out.ris[j] = ris;
  out.ths[j] = ths;
  if (out.srcs) out.srcs[j] = o.srcs;
}

// Column-kernel argument block: the union of the rain_ice and pipeline views.
struct ColumnArgs {
  int ncol, nlev, ld;
  double dt;
  const double *tab;
  // inputs common to both entry points
  const double *exn, *dzz, *rhodref, *exn_ref, *pabs, *sigs;
  // rain_ice only: cloud state and t-state
  const double *cldfr_in, *tht, *rvt, *rct, *rrt, *rit, *rst, *rgt;
  // pipeline only: ice_adjust inputs
  const double *th, *rv, *rc, *ri, *rr, *rs, *rg, *cf_mf, *rc_mf, *ri_mf, *sigqsat;
  // sources: read from src_*, written to dst_* (may alias)
  const double *src_ths, *src_rvs, *src_rcs, *src_rrs, *src_ris, *src_rss, *src_rgs, *src_ci;
  double *ths, *rvs, *rcs, *rrs, *ris, *rss, *rgs, *ci;
  // outputs
  double *cldfr, *hlc_hrc, *hlc_hcf, *hli_hri, *hli_hcf;  // pipeline only
  double *evap3d, *rainfr, *inprc, *inprr, *inpri, *inprs, *inprg;
  int *calls;
  unsigned long long *counters;
};

// Microphysics of one point after sedimentation: RAIN_ICE sections 1, 4, 5, 8
// (micro/rain_ice.F90:224-273, 391-431, 529-543 and the reinstated 649-714).
template <class R>
__device__ inline int micro_point(double dt, double exn, double rhodref, double exnref,
                                  double pabs, double cldfr, double sigs, double tht, double rvt,
                                  double rct, double rrt, double rit, double rst, double rgt,
                                  double s[7], double &cit, double &evap3d,
                                  const double *__restrict__ tab) {
  const ice3_params &P = c_P;
  const double zinv = 1. / dt;
  const double zrice = rit + rst + rgt;
  const double zdevide = P.XCPD + P.XCPV * rvt + P.XCL * (rct + rrt) + P.XCI * zrice;
  const double zt = tht * exn;
  double lsfact = (P.XLSTT + (P.XCPV - P.XCI) * (zt - P.XTT)) / zdevide;
  double lvfact = (P.XLVTT + (P.XCPV - P.XCL) * (zt - P.XTT)) / zdevide;
  const bool llmicro = rct > P.XRTMIN[2] || rrt > P.XRTMIN[3] || rit > P.XRTMIN[4] ||
                       rst > P.XRTMIN[5] || rgt > P.XRTMIN[6];
  double wr[7] = {tht, rvt, rct, rrt, rit, rst, rgt};
  double rvheni = 0.;
  evap3d = 0.;
  int calls = 0;
  if (!llmicro) {
    cit = 0.;
    nucleation<R>(true, tht, pabs, rhodref, exn, lsfact / exn, zt, rvt, cit, rvheni);
  }
  rvheni = dmin(s[IRV], rvheni / dt);
  if (llmicro) {
    double extpk[7] = {s[0], s[1], s[2], s[3], s[4], s[5], s[6]};
    calls = stepping_point<R>(dt, exn, rhodref, pabs, cldfr, sigs, cit, wr, extpk, evap3d, tab);
  }
  // 8.1 total tendencies
  lsfact = lsfact / exnref;
  lvfact = lvfact / exnref;
  const double wrv = (wr[IRV] - rvt) * zinv, wrc = (wr[IRC] - rct) * zinv;
  const double wrr = (wr[IRR] - rrt) * zinv, wri = (wr[IRI] - rit) * zinv;
  const double wrs = (wr[IRS] - rst) * zinv, wrg = (wr[IRG] - rgt) * zinv;
  const double wth = (wrc + wrr) * lvfact + (wri + wrs + wrg + 0.) * lsfact;
  double ths = s[ITH] + wth + rvheni * lsfact;
  double rvs = s[IRV] + wrv - rvheni;
  double rcs = s[IRC] + wrc;
  double rrs = s[IRR] + wrr;
  double ris = s[IRI] + wri + rvheni;
  double rss = s[IRS] + wrs;
  double rgs = s[IRG] + wrg;
  // 8.2 ICE4_CORRECT_NEGATIVITIES (micro/mode_ice4_correct_negativities.F90:51-171)
  double zw;
  zw = rcs - dmax(rcs, 0.); rvs = rvs + zw; ths = ths - zw * lvfact; rcs = rcs - zw;
  zw = rrs - dmax(rrs, 0.); rvs = rvs + zw; ths = ths - zw * lvfact; rrs = rrs - zw;
  zw = ris - dmax(ris, 0.); rvs = rvs + zw; ths = ths - zw * lsfact; ris = ris - zw;
  zw = rss - dmax(rss, 0.); rvs = rvs + zw; ths = ths - zw * lsfact; rss = rss - zw;
  zw = rgs - dmax(rgs, 0.); rvs = rvs + zw; ths = ths - zw * lsfact; rgs = rgs - zw;
  zw = dmin(1., dmax(P.XRTMIN[1] - rvs, 0.) / dmax(rcs + ris, 1.E-20));
  ths = ths - zw * (rcs * lvfact + ris * lsfact);
  rvs = rvs + zw * (rcs + ris);
  rcs = (1. - zw) * rcs;
  ris = (1. - zw) * ris;
  zw = dmin(dmax(rrs, 0.), dmax(P.XRTMIN[1] - rvs, 0.));
  rvs = rvs + zw; rrs = rrs - zw; ths = ths - zw * lvfact;
  zw = dmin(dmax(rss, 0.), dmax(P.XRTMIN[1] - rvs, 0.));
  rvs = rvs + zw; rss = rss - zw; ths = ths - zw * lsfact;
  zw = dmin(dmax(rgs, 0.), dmax(P.XRTMIN[1] - rvs, 0.));
  rvs = rvs + zw; rgs = rgs - zw; ths = ths - zw * lsfact;
  s[ITH] = ths; s[IRV] = rvs; s[IRC] = rcs; s[IRR] = rrs; s[IRI] = ris; s[IRS] = rss; s[IRG] = rgs;
  return calls;
}

template <class R, bool ADJ>
__global__ void __launch_bounds__(64) k_column(ColumnArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  const double dt = a.dt;
  double flux[7] = {0., 0., 0., 0., 0., 0., 0.};
  unsigned long long tot_calls = 0, capped = 0, npts = 0;
  for (int k = a.nlev - 1; k >= 0; --k) {
    const size_t j = (size_t)k * a.ld + i;
    const double exn = a.exn[j], rho = a.rhodref[j], exnref = a.exn_ref[j], pabs = a.pabs[j];
    const double sigs = a.sigs[j];
    double s[7] = {a.src_ths[j], a.src_rvs[j], a.src_rcs[j], a.src_rrs[j],
                   a.src_ris[j], a.src_rss[j], a.src_rgs[j]};
    double cit = a.src_ci[j];
    double cldfr, tht, rvt, rct, rrt, rit, rst, rgt;
    if (ADJ) {
      AdjustOut o;
      ice_adjust_point(dt, a.sigqsat[i], pabs, sigs, a.th[j], exn, exnref, rho, a.rv[j], a.rc[j],
                       a.ri[j], a.rr[j], a.rs[j], a.rg[j], a.cf_mf[j], a.rc_mf[j], a.ri_mf[j],
                       s[IRV], s[IRC], s[IRI], s[ITH], o);
      cldfr = o.cldfr;
      a.cldfr[j] = o.cldfr;
      a.hlc_hrc[j] = o.hlc_hrc;
      a.hlc_hcf[j] = o.hlc_hcf;
      a.hli_hri[j] = o.hli_hri;
      a.hli_hcf[j] = o.hli_hcf;
      // t-state of rain_ice: the adjusted sources times dt
      tht = s[ITH] * dt; rvt = s[IRV] * dt; rct = s[IRC] * dt; rrt = s[IRR] * dt;
      rit = s[IRI] * dt; rst = s[IRS] * dt; rgt = s[IRG] * dt;
    } else {
      cldfr = a.cldfr_in[j];
      tht = a.tht[j]; rvt = a.rvt[j]; rct = a.rct[j]; rrt = a.rrt[j];
      rit = a.rit[j]; rst = a.rst[j]; rgt = a.rgt[j];
    }
    sedimentation_level<R>(dt, rho, a.dzz[j], s[IRR], s[IRI], s[IRS], s[IRG], flux);
    double evap3d;
    const int calls = micro_point<R>(dt, exn, rho, exnref, pabs, cldfr, sigs, tht, rvt, rct, rrt,
                                     rit, rst, rgt, s, cit, evap3d, a.tab);
    a.ths[j] = s[ITH]; a.rvs[j] = s[IRV]; a.rcs[j] = s[IRC]; a.rrs[j] = s[IRR];
    a.ris[j] = s[IRI]; a.rss[j] = s[IRS]; a.rgs[j] = s[IRG]; a.ci[j] = cit;
    a.evap3d[j] = evap3d;
    a.rainfr[j] = 1.;
    if (a.calls) a.calls[j] = calls;
    tot_calls += calls < 0 ? -calls : calls;
    capped += calls < 0;
    npts += calls != 0;
  }
  a.inprc[i] = 0.;
  a.inprr[i] = flux[IRR] / c_P.XRHOLW;
  a.inpri[i] = flux[IRI] / c_P.XRHOLW;
  a.inprs[i] = flux[IRS] / c_P.XRHOLW + flux[IRI] / c_P.XRHOLW;
  a.inprg[i] = flux[IRG] / c_P.XRHOLW;
  if (a.counters) {
    atomicAdd(&a.counters[0], tot_calls);
    atomicAdd(&a.counters[1], capped);
    atomicAdd(&a.counters[2], npts);
  }
}

}  // namespace ice3

// ------------------------------------------------------------------ C ABI

struct ice3_ctx {
  int device;
  ice3_params params;
  double *d_tab;
  double h_tab[ICE3_T_TOTAL];
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
  free(c);
}

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
  AdjustArgs a;
  a.ncol = ncol; a.nlev = nlev; a.ld = ld; a.dt = dt; a.in = *in; a.out = *out;
  dim3 grid((ncol + 255) / 256, nlev);
  hipLaunchKernelGGL(k_ice_adjust, grid, dim3(256), 0, (hipStream_t)stream, a);
  HIPCHK(hipGetLastError());
  return ICE3_OK;
}

static int launch_column(ice3_ctx *ctx, int prec, bool adj, const ColumnArgs &a, void *stream) {
  dim3 grid((a.ncol + 63) / 64);
  hipStream_t s = (hipStream_t)stream;
  if (prec == ICE3_PREC_FP64) {
    if (adj) hipLaunchKernelGGL((k_column<double, true>), grid, dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_column<double, false>), grid, dim3(64), 0, s, a);
  } else {
    if (adj) hipLaunchKernelGGL((k_column<float, true>), grid, dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_column<float, false>), grid, dim3(64), 0, s, a);
  }
  HIPCHK(hipGetLastError());
  return ICE3_OK;
}

extern "C" int ice3_rain_ice(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld, double dt,
                             const ice3_rain_ice_in *in, const ice3_rain_ice_inout *io,
                             const ice3_rain_ice_out *out, const ice3_stats *stats,
                             void *stream) {
  if (!ctx || !in || !io || !out || bad_shape(ncol, nlev, ld, dt)) return ICE3_EINVAL;
  if (prec != ICE3_PREC_FP64 && prec != ICE3_PREC_MIXED) return ICE3_EINVAL;
  NN(in->exn); NN(in->dzz); NN(in->rhodref); NN(in->exn_ref); NN(in->pabs); NN(in->cldfr);
  NN(in->sigs); NN(in->tht); NN(in->rvt); NN(in->rct); NN(in->rrt); NN(in->rit); NN(in->rst);
  NN(in->rgt);
  NN(io->ths); NN(io->rvs); NN(io->rcs); NN(io->rrs); NN(io->ris); NN(io->rss); NN(io->rgs);
  NN(io->ci);
  NN(out->evap3d); NN(out->rainfr); NN(out->inprc); NN(out->inprr); NN(out->inpri);
  NN(out->inprs); NN(out->inprg);
  ColumnArgs a;
  memset(&a, 0, sizeof a);
  a.ncol = ncol; a.nlev = nlev; a.ld = ld; a.dt = dt; a.tab = ctx->d_tab;
  a.exn = in->exn; a.dzz = in->dzz; a.rhodref = in->rhodref; a.exn_ref = in->exn_ref;
  a.pabs = in->pabs; a.sigs = in->sigs; a.cldfr_in = in->cldfr;
  a.tht = in->tht; a.rvt = in->rvt; a.rct = in->rct; a.rrt = in->rrt; a.rit = in->rit;
  a.rst = in->rst; a.rgt = in->rgt;
  a.src_ths = a.ths = io->ths; a.src_rvs = a.rvs = io->rvs; a.src_rcs = a.rcs = io->rcs;
  a.src_rrs = a.rrs = io->rrs; a.src_ris = a.ris = io->ris; a.src_rss = a.rss = io->rss;
  a.src_rgs = a.rgs = io->rgs; a.src_ci = a.ci = io->ci;
  a.evap3d = out->evap3d; a.rainfr = out->rainfr; a.inprc = out->inprc; a.inprr = out->inprr;
  a.inpri = out->inpri; a.inprs = out->inprs; a.inprg = out->inprg;
  if (stats) { a.calls = stats->calls; a.counters = stats->counters; }
  return launch_column(ctx, prec, false, a, stream);
}

extern "C" int ice3_adjust_rain_ice(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld,
                                    double dt, const ice3_pipeline_in *in,
                                    const ice3_pipeline_out *out, const ice3_stats *stats,
                                    void *stream) {
  if (!ctx || !in || !out || bad_shape(ncol, nlev, ld, dt)) return ICE3_EINVAL;
  if (prec != ICE3_PREC_FP64 && prec != ICE3_PREC_MIXED) return ICE3_EINVAL;
  NN(in->pabs); NN(in->sigs); NN(in->th); NN(in->exn); NN(in->exn_ref); NN(in->rhodref);
  NN(in->dzz); NN(in->rv); NN(in->rc); NN(in->ri); NN(in->rr); NN(in->rs);
  NN(in->rg); NN(in->ci); NN(in->cf_mf); NN(in->rc_mf); NN(in->ri_mf); NN(in->ths);
  NN(in->rvs); NN(in->rcs); NN(in->rrs); NN(in->ris); NN(in->rss); NN(in->rgs); NN(in->sigqsat);
  NN(out->ths); NN(out->rvs); NN(out->rcs); NN(out->rrs); NN(out->ris); NN(out->rss);
  NN(out->rgs); NN(out->cldfr); NN(out->hlc_hrc); NN(out->hlc_hcf); NN(out->hli_hri);
  NN(out->hli_hcf); NN(out->ci); NN(out->evap3d); NN(out->rainfr); NN(out->inprc);
  NN(out->inprr); NN(out->inpri); NN(out->inprs); NN(out->inprg);
  ColumnArgs a;
  memset(&a, 0, sizeof a);
  a.ncol = ncol; a.nlev = nlev; a.ld = ld; a.dt = dt; a.tab = ctx->d_tab;
  a.exn = in->exn; a.dzz = in->dzz; a.rhodref = in->rhodref; a.exn_ref = in->exn_ref;
  a.pabs = in->pabs; a.sigs = in->sigs;
  a.th = in->th; a.rv = in->rv; a.rc = in->rc; a.ri = in->ri; a.rr = in->rr; a.rs = in->rs;
  a.rg = in->rg; a.cf_mf = in->cf_mf; a.rc_mf = in->rc_mf; a.ri_mf = in->ri_mf;
  a.sigqsat = in->sigqsat;
  a.src_ths = in->ths; a.src_rvs = in->rvs; a.src_rcs = in->rcs; a.src_rrs = in->rrs;
  a.src_ris = in->ris; a.src_rss = in->rss; a.src_rgs = in->rgs; a.src_ci = in->ci;
  a.ths = out->ths; a.rvs = out->rvs; a.rcs = out->rcs; a.rrs = out->rrs; a.ris = out->ris;
  a.rss = out->rss; a.rgs = out->rgs; a.ci = out->ci;
  a.cldfr = out->cldfr; a.hlc_hrc = out->hlc_hrc; a.hlc_hcf = out->hlc_hcf;
  a.hli_hri = out->hli_hri; a.hli_hcf = out->hli_hcf;
  a.evap3d = out->evap3d; a.rainfr = out->rainfr; a.inprc = out->inprc; a.inprr = out->inprr;
  a.inpri = out->inpri; a.inprs = out->inprs; a.inprg = out->inprg;
  if (stats) { a.calls = stats->calls; a.counters = stats->counters; }
  return launch_column(ctx, prec, true, a, stream);
}
