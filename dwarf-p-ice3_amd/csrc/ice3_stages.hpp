// ice3_stages.hpp -- the stages of the ice_adjust / rain_ice path as per-thread
// device functions:
//   adjust_at     one point of ICE_ADJUST;
//   sedim_column  one column, levels walked top-down: the statistical
//                 sedimentation flux recurrence (fluxes stay in registers) and
//                 the surface precipitation;
//   adjust_sedim_column  ICE_ADJUST and the sedimentation of one column in one walk;
//   micro_at      one point: RAIN_ICE sections 1, 4, 5 and 8 (nucleation
//                 outside the microphysics mask, the ICE4 time splitting, the
//                 total tendencies and the negativity correction).
// The kernels in ice3_kernels.hip launch one thread per column / per point on
// these functions.  Everything here is force-inlined and every array index is a
// compile-time constant after unrolling, so the state lives in registers.
#pragma once
#include "../../include/ice3.h"
#include "ice3_physics.hpp"

namespace ice3 {

struct AdjustArgs {
  int ncol, nlev, ld;
  double dt;
  ice3_ice_adjust_in in;
  ice3_ice_adjust_out out;  // t, rv_out, rc_out, ri_out, cph, lv, ls, srcs may be NULL
  double *ris_copy;         // optional second copy of the adjusted ris
  const double *rh;         // hail mixing ratio at t (KRR=7 only)
};

// The OCND2 branch's extra inputs and outputs (include/ice3.h: ice3_ocnd2_in/_out).
struct Ocnd2Args {
  ice3_ocnd2_in in;
  ice3_ocnd2_out out;
  const double *tiwmx;  // ESTABW then ESTABI, ICE3_TIWMX_N entries each
};

// ICE_ADJUST at point j (level k) of column i (micro/ice_adjust.F90:257-472); returns the
// adjusted ris.  With OCND2 the branch's extra fields come from x: the layer thickness is
// ZDZ(k) = PZZ(kp) - PZZ(kp-1), kp = min(k+1, nlev-1) (condensation.F90:366,376-378).
// HAIL (KRR=7): a.rh enters the heat capacity.
template <bool OCND2 = false, bool HAIL = false>
__device__ __forceinline__ double adjust_at(const AdjustArgs &a, int i, size_t j, int k = 0,
                                            const Ocnd2Args *x = nullptr, unsigned *br = nullptr) {
  const ice3_ice_adjust_in &in = a.in;
  double rvs = in.rvs[j], rcs = in.rcs[j], ris = in.ris[j], ths = in.ths[j];
  AdjustOut o;
  Ocnd2Point xp;
  if constexpr (OCND2) {
    const int kp = k + 1 < a.nlev - 1 ? k + 1 : a.nlev - 1;
    const double *zz = x->in.zz;
    xp.tiwmx = x->tiwmx;
    xp.z = zz[j];
    xp.dz = zz[(size_t)kp * a.ld + i] - zz[(size_t)(kp - 1) * a.ld + i];
    xp.cld_wgt = x->in.ice_cld_wgt[i];
    xp.bottom = k == 0;
  }
  ice_adjust_point<OCND2, HAIL>(a.dt, in.sigqsat[i], in.pabs[j], in.sigs[j], in.th[j], in.exn[j],
                                in.exn_ref[j], in.rho_dry_ref[j], in.rv[j], in.rc[j], in.ri[j],
                                in.rr[j], in.rs[j], in.rg[j], in.cf_mf[j], in.rc_mf[j], in.ri_mf[j],
                                rvs, rcs, ris, ths, o, &xp, HAIL ? a.rh[j] : 0., br);
  if constexpr (OCND2) {
    x->out.icldfr[j] = xp.icldfr;
    x->out.wcldfr[j] = xp.wcldfr;
    x->out.ssio[j] = xp.ssio;
    x->out.ssiu[j] = xp.ssiu;
    x->out.ifr[j] = 10.;  // PIFR(:,:) = 10. (condensation.F90:260)
  }
  const ice3_ice_adjust_out &out = a.out;
  if (out.t) out.t[j] = o.t;
  if (out.rv_out) out.rv_out[j] = o.rv;
  if (out.rc_out) out.rc_out[j] = o.rc;
  if (out.ri_out) out.ri_out[j] = o.ri;
  out.cldfr[j] = o.cldfr;
  out.hlc_hrc[j] = o.hlc_hrc;
  out.hlc_hcf[j] = o.hlc_hcf;
  out.hli_hri[j] = o.hli_hri;
  out.hli_hcf[j] = o.hli_hcf;
  if (out.cph) out.cph[j] = o.cph;
  if (out.lv) out.lv[j] = o.lv;
  if (out.ls) out.ls[j] = o.ls;
  out.rvs[j] = rvs;
  out.rcs[j] = rcs;
  out.ris[j] = ris;
  out.ths[j] = ths;
  if (out.srcs) out.srcs[j] = o.srcs;
  if (a.ris_copy) a.ris_copy[j] = ris;
  return ris;
}

// Sedimentation stage arguments.  src_* are the sources entering RAIN_ICE,
// dst_* receive them after sedimentation (dst may alias src).
struct SedimArgs {
  int ncol, nlev, ld;
  double dt;
  const double *rhodref, *dzz;
  const double *src_rr, *src_ri, *src_rs, *src_rg;
  double *dst_rr, *dst_ri, *dst_rs, *dst_rg;
  double *inprc, *inprr, *inpri, *inprs, *inprg;
  double *zzz;  // OCND2 branch: receives the model level height ZZZZ (micro/rain_ice.F90:387-395)
  const double *src_rh;  // hail (KRR=7): source in, source out, surface rate
  double *dst_rh, *inprh;
};

// ICE4_SEDIMENTATION with CSEDIM='SPLI' (micro/mode_ice4_sedimentation.F90:196-219): the
// split sedimentation of rain, pristine ice, snow, graupel (and hail) on the state at t, then
// ICE4_CORRECT_NEGATIVITIES of every source.  FP64 whatever the precision of the call.
struct SplitArgs {
  int ncol, nlev, ld;
  double dt, tscale, maxcfl;
  const double *rhodref, *dzz, *exn;
  const double *t[8];    // state at t (times tscale): th, rv, rc, rr, ri, rs, rg, rh
  const double *src[8];  // sources read
  double *dst[8];        // sources written (may alias src)
  double *inprc, *inprr, *inpri, *inprs, *inprg, *inprh;
  double *zzz;           // OCND2 branch: receives the model level height ZZZZ
  unsigned long long *capped;  // columns whose species hit the sub-step cap
  double *keep[3];       // optional: receive t[0..2] as read (th, rv, rc at t, unscaled)
};

// ICE4_SEDIMENTATION_STAT (micro/mode_ice4_sedimentation_stat.F90:143-229) for column i,
// followed by the surface precipitation of RAIN_ICE (PINPRS += PINPRI, PINPRC = 0 with
// LSEDIC=F).  With ZZZ the walk also writes RAIN_ICE's model level height ZZZZ into a.zzz:
// PDZZ accumulated from the top level down, minus half the level's own (rain_ice.F90:387-395,
// JK = IKTE..IKTB with level 1 the lowest).  HAIL: hail sediments too (species 7,
// micro/mode_ice4_sedimentation_stat.F90:184-188,208-210,220-222).
template <class R, bool ZZZ = false, bool HAIL = false>
__device__ __forceinline__ void sedim_column(const SedimArgs &a, int i) {
  const double dt = a.dt;
  double fr = 0., fi = 0., fs = 0., fg = 0., fh = 0., zzsum = 0.;
  for (int k = a.nlev - 1; k >= 0; --k) {
    const size_t j = (size_t)k * a.ld + i;
    const double rho = a.rhodref[j], dz = a.dzz[j];
    if constexpr (ZZZ) {
      zzsum = zzsum + dz;
      a.zzz[j] = zzsum - dz * 0.5;
    }
    const R lnrho = rlog<R>((R)rho);
    double rrs = a.src_rr[j], ris = a.src_ri[j], rss = a.src_rs[j], rgs = a.src_rg[j];
    sediment_species<R>(IRR, dt, rho, lnrho, dz, rrs, fr);
    sediment_species<R>(IRI, dt, rho, lnrho, dz, ris, fi);
    sediment_species<R>(IRS, dt, rho, lnrho, dz, rss, fs);
    sediment_species<R>(IRG, dt, rho, lnrho, dz, rgs, fg);
    a.dst_rr[j] = rrs; a.dst_ri[j] = ris; a.dst_rs[j] = rss; a.dst_rg[j] = rgs;
    if constexpr (HAIL) {
      double rhs = a.src_rh[j];
      sediment_species<R>(IRH, dt, rho, lnrho, dz, rhs, fh);
      a.dst_rh[j] = rhs;
    }
  }
  a.inprc[i] = 0.;
  a.inprr[i] = fr / P.XRHOLW;
  a.inpri[i] = fi / P.XRHOLW;
  a.inprs[i] = fs / P.XRHOLW + fi / P.XRHOLW;
  a.inprg[i] = fg / P.XRHOLW;
  if constexpr (HAIL) a.inprh[i] = fh / P.XRHOLW;
}

// ICE_ADJUST and the sedimentation of column i in one top-down walk (the FP64 fused step):
// at each level the point is adjusted, then its precipitating species sediment, ri from the
// adjusted source held in a register.  The sedimentation of a level needs only that level
// and the flux from above, so the results are those of adjust_at over all points followed
// by sedim_column.  HAIL: ICE_ADJUST with the hail heat capacity, and hail sediments.
template <class R, bool HAIL = false>
__device__ __forceinline__ void adjust_sedim_column(const AdjustArgs &aa, const SedimArgs &a, int i) {
  const double dt = a.dt;
  double fr = 0., fi = 0., fs = 0., fg = 0., fh = 0.;
  for (int k = a.nlev - 1; k >= 0; --k) {
    const size_t j = (size_t)k * a.ld + i;
    double ris = adjust_at<false, HAIL>(aa, i, j);
    const double rho = a.rhodref[j], dz = a.dzz[j];
    const R lnrho = rlog<R>((R)rho);
    double rrs = a.src_rr[j], rss = a.src_rs[j], rgs = a.src_rg[j];
    sediment_species<R>(IRR, dt, rho, lnrho, dz, rrs, fr);
    sediment_species<R>(IRI, dt, rho, lnrho, dz, ris, fi);
    sediment_species<R>(IRS, dt, rho, lnrho, dz, rss, fs);
    sediment_species<R>(IRG, dt, rho, lnrho, dz, rgs, fg);
    a.dst_rr[j] = rrs; a.dst_ri[j] = ris; a.dst_rs[j] = rss; a.dst_rg[j] = rgs;
    if constexpr (HAIL) {
      double rhs = a.src_rh[j];
      sediment_species<R>(IRH, dt, rho, lnrho, dz, rhs, fh);
      a.dst_rh[j] = rhs;
    }
  }
  a.inprc[i] = 0.;
  a.inprr[i] = fr / P.XRHOLW;
  a.inpri[i] = fi / P.XRHOLW;
  a.inprs[i] = fs / P.XRHOLW + fi / P.XRHOLW;
  a.inprg[i] = fg / P.XRHOLW;
  if constexpr (HAIL) a.inprh[i] = fh / P.XRHOLW;
}

// Microphysics stage arguments.  The t-state of species v is t[v][j] * tscale
// (tscale = 1 when the caller hands the t-state over, dt when it is the adjusted
// sources); src[v] are the sources after sedimentation, dst[v] the results (may
// alias src).  Species order: th, rv, rc, rr, ri, rs, rg, and hail (entry 7, read by the
// KRR=7 kernels only).
struct MicroArgs {
  int ncol, nlev, ld;
  double dt, tscale;
  const double *tab;
  const double *exn, *rhodref, *exn_ref, *pabs, *cldfr, *sigs;
  const double *t[8];
  const double *src[8];
  const double *src_ci;
  double *dst[8];
  double *ci, *evap3d, *rainfr;
  int *calls;
  unsigned long long *counters;  // [0] tendency calls, [1] capped points, [2] micro points
  int max_calls;                 // sub-step cap per point
  unsigned long long *capped;    // context counter of capped points (always counted)
  // OCND2 branch only: ICE_ADJUST's OCND2 outputs, the TIWMX tables, and the model level
  // height ZZZZ, which the sedimentation's column walk writes into rainfr before the
  // microphysics reads it and writes the rain fraction over it
  const double *icldfr, *ssio, *ssiu, *ifr, *tiwmx;
  unsigned *branches;  // test hook only (ice3_rain_ice_branches): per point, ICE3_BR_* bits
};

// ICE4_CORRECT_NEGATIVITIES (micro/mode_ice4_correct_negativities.F90:51-171) on one
// point's sources: negative condensate goes back to vapour with its latent heat, then
// vapour below XRTMIN(1) is restored from cloud + ice, rain, snow and graupel; with HAIL,
// hail takes part in both steps after graupel (:115-120,165-170).
template <bool HAIL = false>
__device__ __forceinline__ void correct_negativities(double &ths, double &rvs, double &rcs,
                                                     double &rrs, double &ris, double &rss,
                                                     double &rgs, double lvfact, double lsfact,
                                                     double *rhs = nullptr) {
  double zw;
  zw = rcs - dmax(rcs, 0.); rvs = rvs + zw; ths = ths - zw * lvfact; rcs = rcs - zw;
  zw = rrs - dmax(rrs, 0.); rvs = rvs + zw; ths = ths - zw * lvfact; rrs = rrs - zw;
  zw = ris - dmax(ris, 0.); rvs = rvs + zw; ths = ths - zw * lsfact; ris = ris - zw;
  zw = rss - dmax(rss, 0.); rvs = rvs + zw; ths = ths - zw * lsfact; rss = rss - zw;
  zw = rgs - dmax(rgs, 0.); rvs = rvs + zw; ths = ths - zw * lsfact; rgs = rgs - zw;
  if constexpr (HAIL) {
    zw = *rhs - dmax(*rhs, 0.); rvs = rvs + zw; ths = ths - zw * lsfact; *rhs = *rhs - zw;
  }
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
  if constexpr (HAIL) {
    zw = dmin(dmax(*rhs, 0.), dmax(P.XRTMIN[1] - rvs, 0.));
    rvs = rvs + zw; *rhs = *rhs - zw; ths = ths - zw * lsfact;
  }
}

// t-state (times tscale) and sources of the first NV variables at point j (7, or 8 with hail).
template <int NV = 7>
__device__ __forceinline__ void micro_load(const MicroArgs &a, size_t j, double *t, double *s) {
  const double ts = a.tscale;
#pragma unroll
  for (int v = 0; v < NV; ++v) t[v] = a.t[v][j] * ts;
#pragma unroll
  for (int v = 0; v < NV; ++v) s[v] = a.src[v][j];
}

// RAIN_ICE's latent heat factors L / cph (micro/rain_ice.F90:334-347) from the t-state t;
// with HAIL the ice mixing ratio includes hail (:337-338).
template <bool HAIL = false>
__device__ __forceinline__ void rain_ice_factors(const double *t, double exn, double &lsfact,
                                                 double &lvfact) {
  double zrice = t[IRI] + t[IRS] + t[IRG];
  if constexpr (HAIL) zrice = zrice + t[IRH];
  const double zdevide = P.XCPD + P.XCPV * t[IRV] + P.XCL * (t[IRC] + t[IRR]) + P.XCI * zrice;
  const double zt = t[ITH] * exn;
  lsfact = (P.XLSTT + (P.XCPV - P.XCI) * (zt - P.XTT)) / zdevide;
  lvfact = (P.XLVTT + (P.XCPV - P.XCL) * (zt - P.XTT)) / zdevide;
}

// RAIN_ICE's microphysics mask (micro/rain_ice.F90:365-381): a hydrometeor above XRTMIN;
// with OCND2 (:348-364), ice supersaturation in the ice cloud above XFRMIN(12) or a
// hydrometeor above XFRMIN(13).  HAIL: hail above XRTMIN(7) too (:366-372).
template <bool OCND2 = false, bool HAIL = false>
__device__ __forceinline__ bool micro_mask(const double *t, const MicroArgs &a, size_t j) {
  if constexpr (OCND2)
    return a.ssio[j] > P.XFRMIN[12] || t[IRC] > P.XFRMIN[13] || t[IRR] > P.XFRMIN[13] ||
           t[IRI] > P.XFRMIN[13] || t[IRS] > P.XFRMIN[13] || t[IRG] > P.XFRMIN[13];
  if constexpr (HAIL)
    return t[IRC] > P.XRTMIN[2] || t[IRR] > P.XRTMIN[3] || t[IRI] > P.XRTMIN[4] ||
           t[IRS] > P.XRTMIN[5] || t[IRG] > P.XRTMIN[6] || t[IRH] > P.XRTMIN[7];
  return t[IRC] > P.XRTMIN[2] || t[IRR] > P.XRTMIN[3] || t[IRI] > P.XRTMIN[4] ||
         t[IRS] > P.XRTMIN[5] || t[IRG] > P.XRTMIN[6];
}

// The OCND2 inputs of point j (the model level height is staged in rainfr).
__device__ __forceinline__ MicroOcnd2 ocnd2_point(const MicroArgs &a, size_t j) {
  return MicroOcnd2{a.tiwmx, a.icldfr[j], a.ssio[j], a.ssiu[j], a.ifr[j], a.rainfr[j]};
}

// Nucleation outside the microphysics mask (micro/rain_ice.F90:505-540) at point j with
// t-state t: cit is reset and updated, rvheni receives the nucleated vapour.
template <class R, bool OCND2 = false, bool HAIL = false>
__device__ __forceinline__ void micro_nucleation(const MicroArgs &a, size_t j, const double *t,
                                                 double &cit, double &rvheni) {
  const double exn = a.exn[j];
  double lsfact, lvfact;
  rain_ice_factors<HAIL>(t, exn, lsfact, lvfact);
  cit = 0.;
  MicroOcnd2 x;
  if constexpr (OCND2) x = ocnd2_point(a, j);
  nucleation<R, OCND2>(true, t[ITH], a.pabs[j], a.rhodref[j], exn, lsfact / exn, t[ITH] * exn,
                       t[IRV], cit, rvheni, &x);
}

// RAIN_ICE section 8 at point j (micro/rain_ice.F90:612-652 and the reinstated 754-823):
// total tendencies from the end state wr of the time splitting, ICE4_CORRECT_NEGATIVITIES,
// and the stores.  The t-state and sources are read again from memory (they are not
// written before this point's stores).  rvheni: vapour nucleated outside the mask.  HAIL:
// wr has hail last, and the hail source takes its tendency and negativity correction.
template <bool HAIL = false>
__device__ __forceinline__ void micro_finish(const MicroArgs &a, size_t j, const double *wr,
                                             double cit, double evap3d, double rvheni, int calls) {
  constexpr int NV = HAIL ? 8 : 7;
  const double dt = a.dt, zinv = 1. / dt;
  double t[NV], s[NV];
  micro_load<NV>(a, j, t, s);
  double lsfact, lvfact;
  rain_ice_factors<HAIL>(t, a.exn[j], lsfact, lvfact);
  rvheni = dmin(s[IRV], rvheni / dt);
  // 8.1 total tendencies
  const double exnref = a.exn_ref[j];
  lsfact = lsfact / exnref;
  lvfact = lvfact / exnref;
  const double wrv = (wr[IRV] - t[IRV]) * zinv, wrc = (wr[IRC] - t[IRC]) * zinv;
  const double wrr = (wr[IRR] - t[IRR]) * zinv, wri = (wr[IRI] - t[IRI]) * zinv;
  const double wrs = (wr[IRS] - t[IRS]) * zinv, wrg = (wr[IRG] - t[IRG]) * zinv;
  const double wrh = HAIL ? (wr[NV - 1] - t[NV - 1]) * zinv : 0.;
  const double wth = (wrc + wrr) * lvfact + (wri + wrs + wrg + wrh) * lsfact;
  double ths = s[ITH] + wth + rvheni * lsfact;
  double rvs = s[IRV] + wrv - rvheni;
  double rcs = s[IRC] + wrc;
  double rrs = s[IRR] + wrr;
  double ris = s[IRI] + wri + rvheni;
  double rss = s[IRS] + wrs;
  double rgs = s[IRG] + wrg;
  double rhs = HAIL ? s[NV - 1] + wrh : 0.;
  // 8.2 ICE4_CORRECT_NEGATIVITIES
  correct_negativities<HAIL>(ths, rvs, rcs, rrs, ris, rss, rgs, lvfact, lsfact, &rhs);
  a.dst[ITH][j] = ths; a.dst[IRV][j] = rvs; a.dst[IRC][j] = rcs; a.dst[IRR][j] = rrs;
  a.dst[IRI][j] = ris; a.dst[IRS][j] = rss; a.dst[IRG][j] = rgs;
  if constexpr (HAIL) a.dst[NV - 1][j] = rhs;
  a.ci[j] = cit;
  a.evap3d[j] = evap3d;
  a.rainfr[j] = 1.;
  if (a.calls) a.calls[j] = calls;
}

// RAIN_ICE sections 1, 4, 5, 8 (micro/rain_ice.F90:334-384, 505-540, 612-652 and the
// reinstated 754-823) for point j, all in one thread.  Returns the number of tendency
// calls, negative if the sub-step cap was hit.  HAIL: KRR=7, hail the eighth variable.
template <class R, bool OCND2 = false, bool HAIL = false, class TB>
__device__ __forceinline__ int micro_at(const MicroArgs &a, size_t j, const TB *__restrict__ tab,
                                        unsigned *br = nullptr) {
  constexpr int NV = HAIL ? 8 : 7;
  double wr[NV], s[NV];
  micro_load<NV>(a, j, wr, s);
  double cit = a.src_ci[j], rvheni = 0., evap3d = 0.;
  int calls = 0;
  if (micro_mask<OCND2, HAIL>(wr, a, j)) {
    MicroOcnd2 x;
    if constexpr (OCND2) x = ocnd2_point(a, j);
    calls = stepping_point<R, OCND2, HAIL>(a.dt, a.exn[j], a.rhodref[j], a.pabs[j], a.cldfr[j],
                                           a.sigs[j], cit, wr, s, evap3d, a.max_calls, tab, &x,
                                           HAIL ? a.tab + ICE3_T_ICE3 : nullptr, br);
  } else {
    micro_nucleation<R, OCND2, HAIL>(a, j, wr, cit, rvheni);
  }
  micro_finish<HAIL>(a, j, wr, cit, evap3d, rvheni, calls);
  return calls;
}

// Input features of the surrogate at point j (the microphysics inputs of micro_at):
// T, ln p, rhodref, ln(r + 1e-9) of the six water species of the t-state, ln(1 + ci),
// sigma_s, exner, dt, cloud fraction, and asinh(1e6 * dt * e) of the external tendency
// e = source - t-state/dt of rr, ri, rs, rg (the sedimentation, which the time
// splitting integrates beside the processes).  Returns false where RAIN_ICE would not
// run the ICE4 microphysics (no hydrometeor above XRTMIN), as micro_at decides it.
__device__ __forceinline__ bool surrogate_features(const MicroArgs &a, size_t j, float *f) {
  const double ts = a.tscale, exn = a.exn[j];
  double r[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) r[v] = a.t[v][j] * ts;
  f[0] = (float)(r[ITH] * exn);
  f[1] = logf((float)a.pabs[j]);
  f[2] = (float)a.rhodref[j];
#pragma unroll
  for (int v = 1; v < 7; ++v) f[2 + v] = logf((float)fmax(r[v], 0.) + 1e-9f);
  f[9] = logf(1.f + (float)fmax(a.src_ci[j], 0.));
  f[10] = (float)a.sigs[j];
  f[11] = (float)exn;
  f[12] = (float)a.dt;
  f[13] = (float)a.cldfr[j];
#pragma unroll
  for (int v = IRR; v <= IRG; ++v)
    f[14 + v - IRR] = asinhf((float)(1e6 * (a.src[v][j] * a.dt - r[v])));
  return r[IRC] > P.XRTMIN[2] || r[IRR] > P.XRTMIN[3] || r[IRI] > P.XRTMIN[4] ||
         r[IRS] > P.XRTMIN[5] || r[IRG] > P.XRTMIN[6];
}

// Latent heat factors L/(cph * exnref) of the total tendencies at point j (t-state), as
// micro_at computes them (micro/rain_ice.F90:334-347 divided by PEXNREF).
__device__ __forceinline__ void latent_factors(const MicroArgs &a, size_t j, double &lvfact,
                                               double &lsfact) {
  double t[7], s[7];
  micro_load(a, j, t, s);
  rain_ice_factors(t, a.exn[j], lsfact, lvfact);
  lsfact = lsfact / a.exn_ref[j];
  lvfact = lvfact / a.exn_ref[j];
}

// ---------------------------------------------------------------- argument set-up
// Host-side wiring of the C-ABI structs into the stage arguments, shared by the
// kernel launchers and by CPU builds of the stages.

// rain_ice: sedimentation updates the precipitating sources in place, then the
// microphysics reads the handed-over t-state and updates all sources in place.
// The split sedimentation's arguments for the stages whose sedimentation is described by sa
// and whose microphysics (the state at t, th/rv/rc sources) by m.
inline SplitArgs split_args(const SedimArgs &sa, const MicroArgs &m, double maxcfl,
                            unsigned long long *capped) {
  SplitArgs a = SplitArgs{};
  a.ncol = sa.ncol; a.nlev = sa.nlev; a.ld = sa.ld; a.dt = sa.dt; a.tscale = m.tscale;
  a.maxcfl = maxcfl;
  a.rhodref = sa.rhodref; a.dzz = sa.dzz; a.exn = m.exn;
  for (int q = 0; q < 8; ++q) a.t[q] = m.t[q];
  for (int q = 0; q < 3; ++q) { a.src[q] = m.src[q]; a.dst[q] = m.dst[q]; }
  a.src[3] = sa.src_rr; a.src[4] = sa.src_ri; a.src[5] = sa.src_rs; a.src[6] = sa.src_rg;
  a.src[7] = sa.src_rh;
  a.dst[3] = sa.dst_rr; a.dst[4] = sa.dst_ri; a.dst[5] = sa.dst_rs; a.dst[6] = sa.dst_rg;
  a.dst[7] = sa.dst_rh;
  a.inprc = sa.inprc; a.inprr = sa.inprr; a.inpri = sa.inpri; a.inprs = sa.inprs;
  a.inprg = sa.inprg; a.inprh = sa.inprh;
  a.zzz = sa.zzz;
  a.capped = capped;
  return a;
}

inline void rain_ice_args(int ncol, int nlev, int ld, double dt, const double *tab,
                          const ice3_rain_ice_in *in, const ice3_rain_ice_inout *io,
                          const ice3_rain_ice_out *out, const ice3_stats *stats,
                          SedimArgs &sa, MicroArgs &m) {
  sa = SedimArgs{};
  sa.ncol = ncol; sa.nlev = nlev; sa.ld = ld; sa.dt = dt;
  sa.rhodref = in->rhodref; sa.dzz = in->dzz;
  sa.src_rr = sa.dst_rr = io->rrs; sa.src_ri = sa.dst_ri = io->ris;
  sa.src_rs = sa.dst_rs = io->rss; sa.src_rg = sa.dst_rg = io->rgs;
  sa.inprc = out->inprc; sa.inprr = out->inprr; sa.inpri = out->inpri; sa.inprs = out->inprs;
  sa.inprg = out->inprg;
  m = MicroArgs{};
  m.ncol = ncol; m.nlev = nlev; m.ld = ld; m.dt = dt; m.tscale = 1.; m.tab = tab;
  m.max_calls = DEFAULT_MAX_CALLS;
  m.exn = in->exn; m.rhodref = in->rhodref; m.exn_ref = in->exn_ref; m.pabs = in->pabs;
  m.cldfr = in->cldfr; m.sigs = in->sigs;
  const double *t[7] = {in->tht, in->rvt, in->rct, in->rrt, in->rit, in->rst, in->rgt};
  double *v[7] = {io->ths, io->rvs, io->rcs, io->rrs, io->ris, io->rss, io->rgs};
  for (int q = 0; q < 7; ++q) { m.t[q] = t[q]; m.src[q] = v[q]; m.dst[q] = v[q]; }
  m.src_ci = m.ci = io->ci;
  m.evap3d = out->evap3d; m.rainfr = out->rainfr;
  if (stats) { m.calls = stats->calls; m.counters = stats->counters; }
}

// The fused time step.
// 1. ICE_ADJUST writes the adjusted th, rv, rc, ri sources into the outputs and
//    stages a copy of the adjusted ris in out->evap3d (the microphysics reads it
//    there as the ri t-state, then overwrites it with the evaporation).
// 2. Sedimentation of rr, rs, rg (input sources) and ri (adjusted) into the outputs.
// 3. Microphysics with the t-state = adjusted sources x dt.
inline void pipeline_args(int ncol, int nlev, int ld, double dt, const double *tab,
                          const ice3_pipeline_in *in, const ice3_pipeline_out *out,
                          const ice3_stats *stats, AdjustArgs &aa, SedimArgs &sa, MicroArgs &m) {
  aa = AdjustArgs{};
  aa.ncol = ncol; aa.nlev = nlev; aa.ld = ld; aa.dt = dt;
  aa.in.pabs = in->pabs; aa.in.sigs = in->sigs; aa.in.th = in->th; aa.in.exn = in->exn;
  aa.in.exn_ref = in->exn_ref; aa.in.rho_dry_ref = in->rhodref; aa.in.rv = in->rv;
  aa.in.rc = in->rc; aa.in.ri = in->ri; aa.in.rr = in->rr; aa.in.rs = in->rs; aa.in.rg = in->rg;
  aa.in.cf_mf = in->cf_mf; aa.in.rc_mf = in->rc_mf; aa.in.ri_mf = in->ri_mf;
  aa.in.rvs = in->rvs; aa.in.rcs = in->rcs; aa.in.ris = in->ris; aa.in.ths = in->ths;
  aa.in.sigqsat = in->sigqsat;
  aa.out.cldfr = out->cldfr; aa.out.hlc_hrc = out->hlc_hrc; aa.out.hlc_hcf = out->hlc_hcf;
  aa.out.hli_hri = out->hli_hri; aa.out.hli_hcf = out->hli_hcf;
  aa.out.rvs = out->rvs; aa.out.rcs = out->rcs; aa.out.ris = out->ris; aa.out.ths = out->ths;
  aa.ris_copy = out->evap3d;
  sa = SedimArgs{};
  sa.ncol = ncol; sa.nlev = nlev; sa.ld = ld; sa.dt = dt;
  sa.rhodref = in->rhodref; sa.dzz = in->dzz;
  sa.src_rr = in->rrs; sa.src_ri = out->ris; sa.src_rs = in->rss; sa.src_rg = in->rgs;
  sa.dst_rr = out->rrs; sa.dst_ri = out->ris; sa.dst_rs = out->rss; sa.dst_rg = out->rgs;
  sa.inprc = out->inprc; sa.inprr = out->inprr; sa.inpri = out->inpri; sa.inprs = out->inprs;
  sa.inprg = out->inprg;
  m = MicroArgs{};
  m.ncol = ncol; m.nlev = nlev; m.ld = ld; m.dt = dt; m.tscale = dt; m.tab = tab;
  m.max_calls = DEFAULT_MAX_CALLS;
  m.exn = in->exn; m.rhodref = in->rhodref; m.exn_ref = in->exn_ref; m.pabs = in->pabs;
  m.cldfr = out->cldfr; m.sigs = in->sigs;
  const double *t[7] = {out->ths, out->rvs, out->rcs, in->rrs, out->evap3d, in->rss, in->rgs};
  double *v[7] = {out->ths, out->rvs, out->rcs, out->rrs, out->ris, out->rss, out->rgs};
  for (int q = 0; q < 7; ++q) { m.t[q] = t[q]; m.src[q] = v[q]; m.dst[q] = v[q]; }
  m.src_ci = in->ci; m.ci = out->ci;
  m.evap3d = out->evap3d; m.rainfr = out->rainfr;
  if (stats) { m.calls = stats->calls; m.counters = stats->counters; }
}

}  // namespace ice3
