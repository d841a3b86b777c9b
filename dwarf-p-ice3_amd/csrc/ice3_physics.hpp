// ice3_physics.hpp -- device-side ICE3 point and column physics for CDNA4.
//
// Each function cites the PHYEX CY50T1 routine it implements (paths relative to
// PHYEX-IAL_CY50T1/).  Branch set: AROME ICE3, KRR=6, sigma_s subgrid
// condensation (LSUBG_COND, OSIGMAS, CB02, lambda3='CB', ice fraction 'T'),
// TRIANGLE mass-flux pdf, statistical sedimentation without cloud droplets,
// rc autoconversion pdf 'PDF'+'SIGM', ri 'NONE', accretion/evaporation 'NONE',
// LFEEDBACKT, LEVLIMIT, LCRFLIMIT, LNULLWETG, LWETGPOST, snow riming 'M90',
// XMRSTEP = XTSTEP_TS = 0 (per-point independent time splitting).
//
// Precision: template parameter R is the arithmetic of the process rates and
// table lookups (double: plain FP64; float: the mixed path).  The prognostic
// state, the time-splitting update, the saturation adjustment and the
// sedimentation fluxes are always double.
#pragma once
#include <hip/hip_runtime.h>

#include "ice3_params.h"
#include "ice3_params_gen.inc"

namespace ice3 {

// The constants of the pinned configuration are the compile-time object P
// (ice3_params_gen.inc, generated from ice3_build_params): every P.X below is a
// literal in the generated code.

enum { ITH = 0, IRV = 1, IRC = 2, IRR = 3, IRI = 4, IRS = 5, IRG = 6, IRH = 7 };

// ---------------------------------------------------------------- FP64 exp and log
// The physics evaluates about ten FP64 logarithms and as many exponentials per point and
// sedimentation level, and more in a hard tendency call.  The library versions keep
// extended-precision intermediates (about 100 VALU instructions for a log, 40 for an
// exp); these two stay within about 1 ulp in about 40 and 25, and the parity bar of the
// physics (1e-11 relative) is five orders of magnitude wider.  Every step is a correctly
// rounded IEEE operation (frexp, ldexp, rint, fma, division), except that the kernels are
// built with -fapprox-func, under which the device divides by a reciprocal refined with
// Newton steps (within 1 ulp; in practice the IEEE quotient).

// log x: x = 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s) = 2s + s z p(z),
// s = (m - 1)/(m + 1), z = s^2 <= 0.0295; p is a degree-6 near-minimax fit of
// (log m - 2s)/(s z) on that interval (interpolation at Chebyshev nodes in 60-digit
// arithmetic, fit_exp_log.py; the atanh series 2/3 + 2/5 z + ... needs three more terms
// for the same accuracy).  log x = e ln2 + log m, ln2 split into a high part with 42
// significant bits (e * LN2_HI is exact) and the rest.
__host__ __device__ __forceinline__ double log_f64(double x) {
  constexpr double LN2_HI = 0x1.62e42fefa3800p-1, LN2_LO = 0x1.ef35793c76730p-45;
  constexpr double SQRT_HALF = 0x1.6a09e667f3bcdp-1;
  int e;
  double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  if (m < SQRT_HALF) {
    m = m + m;
    e = e - 1;
  }
  const double f = m - 1.0;  // exact
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 0x1.2b5900de53b32p-3;
  p = fma(p, z, 0x1.39fe51a7c18f9p-3);
  p = fma(p, z, 0x1.7462b51cb66b1p-3);
  p = fma(p, z, 0x1.c71c62e3f11e6p-3);
  p = fma(p, z, 0x1.2492492df281ap-2);
  p = fma(p, z, 0x1.99999999952d7p-2);
  p = fma(p, z, 0x1.5555555555558p-1);
  const double r = fma(s * z, p, s + s);  // log m
  const double de = (double)e;
  const double y = fma(de, LN2_HI, fma(de, LN2_LO, r));
  if (x > 0. && x < INFINITY) return y;
  return x == 0. ? -INFINITY : (x == INFINITY ? x : NAN);  // log 0, log inf, x < 0 or NaN
}

// exp x: x = n ln2 + r, n = rint(x / ln2), |r| <= 0.347; e^r = 1 + r (1 + r q(r)) with q a
// degree-9 polynomial, q(0) = 1/2 exactly, fitted near-minimax to (e^r - 1 - r)/r^2 on
// |r| <= 0.3466 (fit_exp_log.py; the Taylor series needs two more terms for the same
// accuracy); e^x = 2^n e^r.  x is first clamped to [-760, 720], where e^x is 0 or infinite
// at either end.
__host__ __device__ __forceinline__ double exp_f64(double x) {
  constexpr double LN2_HI = 0x1.62e42fefa3800p-1, LN2_LO = 0x1.ef35793c76730p-45;
  constexpr double LOG2E = 0x1.71547652b82fep+0;
  const double xc = fmin(fmax(x, -760.), 720.);
  const double n = rint(xc * LOG2E);
  double r = fma(-n, LN2_HI, xc);  // exact: n LN2_HI is exact and close to xc
  r = fma(-n, LN2_LO, r);
  double p = 0x1.af236857dc4b7p-26;
  p = fma(p, r, 0x1.28803f36ecdd5p-22);
  p = fma(p, r, 0x1.71de17e4e5d5dp-19);
  p = fma(p, r, 0x1.a019c9a6c2942p-16);
  p = fma(p, r, 0x1.a01a01a4b2f7fp-13);
  p = fma(p, r, 0x1.6c16c1738faf6p-10);
  p = fma(p, r, 0x1.1111111110e6ap-7);
  p = fma(p, r, 0x1.5555555554cb7p-5);
  p = fma(p, r, 0x1.5555555555555p-3);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  const double y = ldexp(p, (int)n);
  return x == x ? y : x;  // NaN stays NaN
}

template <class R> __device__ __forceinline__ R rexp(R x);
template <class R> __device__ __forceinline__ R rlog(R x);
template <class R> __device__ __forceinline__ R rpow(R x, R y);
template <> __device__ __forceinline__ double rexp<double>(double x) { return exp_f64(x); }
template <> __device__ __forceinline__ double rlog<double>(double x) { return log_f64(x); }
template <> __device__ __forceinline__ double rpow<double>(double x, double y) {
  return exp_f64(y * log_f64(x));  // x >= 0 at every call site
}
template <> __device__ __forceinline__ float rexp<float>(float x) { return expf(x); }
template <> __device__ __forceinline__ float rlog<float>(float x) { return logf(x); }
template <> __device__ __forceinline__ float rpow<float>(float x, float y) { return powf(x, y); }

// MIN / MAX as the hardware's v_min / v_max (IEEE minNum / maxNum), one instruction where
// a < b ? a : b takes a compare and two selects per double.  The two agree except when b is
// NaN (minNum returns a) or on the sign of a zero result; the operands on the path are
// finite.  (FP64 outputs of the fused step: bitwise unchanged; step 3 % faster.)
__device__ __forceinline__ double dmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float dmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ float dmax(float a, float b) { return __builtin_fmaxf(a, b); }
template <class T> __device__ __forceinline__ T sq(T a) { return a * a; }

// ---------------------------------------------------------------- ice_adjust

struct AdjustOut {
  double t, rv, rc, ri, cldfr, hlc_hrc, hlc_hcf, hli_hri, hli_hcf, cph, lv, ls, srcs;
};

// ZSRC_1D(-22:11), micro/condensation.F90:208-216; entry n is ZSRC_1D(n - 22).  A table in
// memory: a switch over the index compiles to a branch tree that a wavefront with
// scattered indices walks almost whole, twice per point.
__device__ constexpr double ZSRC_1D[34] = {
    0.,            0.,            2.0094444E-04, 0.316670E-03,  4.9965648E-04, 0.785956E-03,
    1.2341294E-03, 0.193327E-02,  3.0190963E-03, 0.470144E-02,  7.2950651E-03, 0.112759E-01,
    1.7350994E-02, 0.265640E-01,  4.0427860E-02, 0.610997E-01,  9.1578111E-02, 0.135888E+00,
    0.1991484,     0.230756E+00,  0.2850565,     0.375050E+00,  0.5000000,     0.691489E+00,
    0.8413813,     0.933222E+00,  0.9772662,     0.993797E+00,  0.9986521,     0.999768E+00,
    0.9999684,     0.999997E+00,  1.0000000,     1.000000};

// ---- OCND2 branch (PARAM_ICE LOCND2 = .TRUE.)
// TIWMX saturation tables (micro/modd_tiwmx.F90: 0.01 K steps from 132 K to 373.16 K):
// ESATW(TIWMX, T) reads the entry NINT(100 T) (micro/mode_tiwmx.F90:9-13).  tab holds ESTABW
// then ESTABI, built on the host as INI_TIWMX does (ice3_build_tiwmx).  Indices outside the
// table are clamped (the reference reads out of bounds there).
__host__ __device__ __forceinline__ int tiwmx_index(double t) {
  const double jk = round(100. * t);  // NINT: half away from zero
  return (int)(jk < ICE3_TIWMX_START ? ICE3_TIWMX_START : (jk > ICE3_TIWMX_STOP ? ICE3_TIWMX_STOP : jk)) -
         ICE3_TIWMX_START;
}

// Per-point inputs and outputs of the OCND2 branch: the TIWMX tables, level height z, layer
// thickness dz (ZDZ, condensation.F90:376-378), first level of the column, PICE_CLD_WGT;
// PICLDFR, PWCLDFR, PSSIO, PSSIU out (PIFR is the constant 10, condensation.F90:260).
struct Ocnd2Point {
  const double *tiwmx;
  double z, dz, cld_wgt;
  bool bottom;
  double icldfr, wcldfr, ssio, ssiu;
};

// ICECLOUD (micro/mode_icecloud.F90:52-205) for one point as CONDENSATION calls it
// (condensation.F90:379-382): PR = rv, PTSTEP = 1, PPBLH = -1, PWCLD = -1.  Returns the ice
// cloud fraction and the super-/sub-saturation with respect to ice; esw, esi are the
// tabulated saturation pressures at t.
__device__ __forceinline__ void icecloud(double p, double z, double dz, double t, double rv,
                                         double esw, double esi, double &sifrc, double &ssio,
                                         double &ssiu) {
  const double sigx = 3.E-4, sigy = 3.E-4, sigz = 1.E-2, xdist = 2500., ydist = 2500.;
  const double r = dmax(0., rv * 1.);
  double s = 0.;
  const double za = r * p / (P.XEPSILO + r);
  const double rhw = za / esw, rhi = za / esi, i2w = esw / esi;
  double su = dmin(i2w, rhi), so = su;
  if (t > 273.1 || r <= 0. || esi >= p * 0.5) {
    su = su - 1.;
    so = su;
  } else {
    const double rhin = dmax(0.05, dmin(1., rhw));
    const double drhdz = rhin * P.XG / (t * P.XRD) * (P.XEPSILO * P.XLVTT / (P.XCPD * t) - 1.);
    const double zz = dmin(1., dmax(0., z * 0.001));  // PPBLH < 0
    double rhdist = sqrt(xdist * sq(sigx) + ydist * sq(sigy) + (1. - zz) * sq(dz * drhdz) +
                         zz * dz * sq(sigz));
    if (zz > 0.1) rhdist = rhdist / (1. + rhdist);
    const double rhlim = dmax(0.5, dmin(0.99, 1. - 0.5 * rhdist));
    double rhdif = (1. - rhw) / (1.0 - rhlim);        // PWCLD < 0
    rhdif = 1. - sqrt(dmax(0., rhdif));
    const double wcld = dmin(1., dmax(rhdif, 0.0));
    s = wcld;
    const double rhlimice = 1. + i2w * (rhlim - 1.);
    if (rhlim <= 0.999) {
      rhdif = (rhi - rhlimice) * (1. / (1. - rhlimice));
      if (wcld == 0.) {
        s = dmin(1., 0.5 * dmax(0., rhdif));
      } else {
        s = dmin(1., (1. - 1. / i2w) * 0.5 / (1. - rhlim));
        s = dmin(1., wcld + s);
      }
    }
    if (s > 0.01) {
      su = s + rhlimice * (1. - s);
      so = (rhi - (1. - s) * su) / s;
    } else {
      s = 0.;
      su = dmax(0., s + rhlimice * (1. - s) + 2. * dmin(0., rhi - rhlimice));
    }
    so = dmin(i2w, so);
    su = dmax(0., su) - 1.;
    so = so - 1.;
  }
  sifrc = s;
  ssio = so;
  ssiu = su;
}

// ICE_ADJUST (micro/ice_adjust.F90:257-472) + ITERATION (:488-565) + CONDENSATION
// (micro/condensation.F90:270-650) for one point; rvs..ths updated in place.  OCND2 selects
// the branch of PARAM_ICE LOCND2 = .TRUE. (x then carries its inputs and outputs):
// ZPRIFACT = 0 (condensation.F90:265), ICECLOUD and the tabulated saturation pressures
// (:373-394), no ice fraction (:412), and the liquid condensate with its noise check and the
// separate ice cloud (:588-637).  HAIL (KRR=7): the hail mixing ratio rh enters the heat
// capacity (micro/ice_adjust.F90:503-507).
template <bool OCND2 = false, bool HAIL = false>
__device__ __forceinline__ void ice_adjust_point(double dt, double sigqsat, double pabs, double sigs,
                                        double th, double exn, double exnref, double rhodref,
                                        double rv, double rc, double ri, double rr, double rs,
                                        double rg, double cf_mf, double rc_mf, double ri_mf,
                                        double &rvs, double &rcs, double &ris, double &ths,
                                        AdjustOut &o, Ocnd2Point *x = nullptr, double rh = 0.,
                                        unsigned *br = nullptr) {
  // br (a test hook; compiled out when null): the ICE3_ABR_* record of the branches taken
  double zt = th * exn;
  const double lv = P.XLVTT + (P.XCPV - P.XCL) * (zt - P.XTT);
  const double ls = P.XLSTT + (P.XCPV - P.XCI) * (zt - P.XTT);
  const double cph = HAIL ? P.XCPD + P.XCPV * rv + P.XCL * (rc + rr) + P.XCI * (ri + rs + rg + rh)
                          : P.XCPD + P.XCPV * rv + P.XCL * (rc + rr) + P.XCI * (ri + rs + rg);
  const double prifact = OCND2 ? 0. : 1.0;
  const double rt = rv + rc + ri * prifact;
  double pv, piv, frac, esw_t = 0.;
  if constexpr (OCND2) {
    const int it = tiwmx_index(zt);
    esw_t = x->tiwmx[ICE3_TW_ESW * ICE3_TIWMX_N + it];
    const double esi_t = x->tiwmx[ICE3_TW_ESI * ICE3_TIWMX_N + it];
    icecloud(pabs, x->z, x->dz, zt, rv, esw_t, esi_t, x->icldfr, x->ssio, x->ssiu);
    pv = dmin(esw_t, .99 * pabs);
    piv = dmin(esi_t, .99 * pabs);
    frac = 0.;
  } else {
    const double logt = log_f64(zt);
    pv = dmin(exp_f64(P.XALPW - P.XBETAW / zt - P.XGAMW * logt), .99 * pabs);
    piv = dmin(exp_f64(P.XALPI - P.XBETAI / zt - P.XGAMI * logt), .99 * pabs);
    frac = dmax(0., dmin(1., ((P.XTMAXMIX - zt) / (P.XTMAXMIX - P.XTMINMIX))));
  }
  double qsl = P.XRD / P.XRV * pv / (pabs - pv);
  const double qsi = P.XRD / P.XRV * piv / (pabs - piv);
  qsl = (1. - frac) * qsl + frac * qsi;
  const double lvs = (1. - frac) * lv + frac * ls;
  const double ah = lvs * qsl / (P.XRV * sq(zt)) * (P.XRV * qsl / P.XRD + 1.);
  const double a = 1. / (1. + lvs / cph * ah);
  const double sbar = a * (rt - qsl + ah * lvs * (rc + ri * prifact) / cph);
  double sigma = sigqsat != 0. ? sqrt(sq(2 * sigs) + sq(sigqsat * qsl * a)) : 2 * sigs;
  sigma = dmax(1.E-10, sigma);
  const double q1 = sbar / sigma;
  double cond;
  if (q1 > 0. && q1 <= 2)
    cond = dmin(exp(-1.) + .66 * q1 + .086 * sq(q1), 2.);
  else if (q1 > 2.)
    cond = q1;
  else
    cond = exp_f64(1.2 * q1 - 1.);
  cond = cond * sigma;
  double cldfr = cond < 1.E-12 ? 0. : dmax(0., dmin(1., 0.5 + 0.36 * atan(1.55 * q1)));
  if (cldfr == 0.) cond = 0.;
  int inq1 = (int)floor(dmin(100., dmax(-100., 2 * q1)));
  inq1 = inq1 < -22 ? -22 : (inq1 > 10 ? 10 : inq1);
  const double inc = 2. * q1 - (double)inq1;
  double sigrc = dmin(1., (1. - inc) * ZSRC_1D[inq1 + 22] + inc * ZSRC_1D[inq1 + 23]);
  if (br)
    *br = (q1 > 0. && q1 <= 2 ? 1u : q1 > 2. ? 2u : 0u) << ICE3_ABR_Q1 |
          (cldfr == 0. ? 1u : 0u) << ICE3_ABR_NOCLOUD |
          (unsigned)(inq1 + 22) << ICE3_ABR_INQ1;
  double rc_out = (1. - frac) * cond;
  double ri_out = frac * cond;
  if constexpr (OCND2) {
    // ZRSW = -9999 (condensation.F90:263): PRV_IN < ZRSW never holds, ZRSDIF = 0
    double rcold = rc;
    if (fabs(rc_out - rc) > 1.0E-12 && esw_t < pabs * 0.5) {
      rcold = rc_out;
      rc_out = cond - 0.;
    }
    x->wcldfr = cldfr;
    const double dum1 = dmin(1.0, 20. * rc_out * sqrt(x->dz) / qsl);
    const double dum3 = dmax(0., x->icldfr - x->wcldfr);
    double dum4 = x->bottom ? ri : ri + rs * 0.5 + rg * 0.25;
    dum4 = dmax(0., dmin(1., x->cld_wgt * dum4 * sqrt(x->dz) / qsi));
    const double dum2 = (0.8 * cldfr + 0.2) * dmin(1., dum1 + dum4 * cldfr);
    cldfr = dmin(1., dum2 + (0.5 * dum3 + 0.5) * dum4);
    ri_out = ri;
    zt = zt + ((rc_out - rcold) * lv + (ri_out - ri) * ls) / cph;
  } else {
    zt = zt + ((rc_out - rc) * lv + (ri_out - ri) * ls) / cph;
  }
  double rv_out = rt - rc_out - ri_out * prifact;
  sigrc = sigrc * dmin(3., dmax(1., 1. - q1));
  // 5.0-5.1 sources
  double w1 = (rc_out - rc) / dt;
  double w2 = (ri_out - ri) / dt;
  if (br) *br |= (w1 < 0.0 ? 1u << ICE3_ABR_ZW1NEG : 0u) | (w2 < 0.0 ? 1u << ICE3_ABR_ZW2NEG : 0u);
  w1 = w1 < 0.0 ? dmax(w1, -rcs) : dmin(w1, rvs);
  rvs = rvs - w1;
  rcs = rcs + w1;
  ths = ths + w1 * lv / (cph * exnref);
  w2 = w2 < 0.0 ? dmax(w2, -ris) : dmin(w2, rvs);
  rvs = rvs - w2;
  ris = ris + w2;
  ths = ths + w2 * ls / (cph * exnref);
  // 5.2 subgrid: mass-flux contribution, TRIANGLE pdf
  w1 = rc_mf / dt;
  w2 = ri_mf / dt;
  if (w1 + w2 > rvs) {
    if (br) *br |= 1u << ICE3_ABR_MFLIMIT;
    w1 = w1 * rvs / (w1 + w2);
    w2 = rvs - w1;
  }
  cldfr = dmin(1., cldfr + cf_mf);
  rcs = rcs + w1;
  ris = ris + w2;
  rvs = rvs - (w1 + w2);
  ths = ths + (w1 * lv + w2 * ls) / cph / exnref;
  double hlc_hcf = 0., hlc_hrc = 0., hli_hcf = 0., hli_hri = 0.;
  {
    const double criaut = P.XCRIAUTC / rhodref, x = w1 * dt, c = criaut * cf_mf;
    double hcf, hr;
    if (br) *br |= (x > cf_mf * criaut ? 0u : 2. * w1 * dt <= cf_mf * criaut ? 1u : 2u) << ICE3_ABR_PDFC;
    if (x > cf_mf * criaut) {
      hcf = 1. - .5 * sq(criaut * cf_mf / dmax(1.E-20, x));
      hr = x - (c * c * c) / (3 * sq(dmax(1.E-20, x)));
    } else if (2. * w1 * dt <= cf_mf * criaut) {
      hcf = 0.;
      hr = 0.;
    } else {
      hcf = sq(2. * w1 * dt - criaut * cf_mf) / (2. * sq(dmax(1.E-20, x)));
      hr = (4. * (x * x * x) - 3. * w1 * dt * sq(c) + (c * c * c)) / (3 * sq(dmax(1.E-20, x)));
    }
    hlc_hcf = dmin(1., hlc_hcf + hcf * cf_mf);
    hlc_hrc = hlc_hrc + hr;
  }
  // the 10**(aT+b) threshold only matters with mass-flux ice or cloud fraction (without
  // either the ice pdf takes its no-cloud case, 2 w2 dt = 0 <= 0)
  if (br && !(w2 != 0. || cf_mf != 0.)) *br |= 1u << ICE3_ABR_PDFI;
  if (w2 != 0. || cf_mf != 0.) {
    const double criaut = dmin(P.XCRIAUTI, exp_f64(P.LN_10 * (P.XACRIAUTI * (zt - P.XTT) + P.XBCRIAUTI)));
    const double x = w2 * dt, c = criaut * cf_mf;
    double hcf, hr;
    if (br) *br |= (x > cf_mf * criaut ? 0u : 2. * w2 * dt <= cf_mf * criaut ? 1u : 2u) << ICE3_ABR_PDFI;
    if (x > cf_mf * criaut) {
      hcf = 1. - .5 * sq(criaut * cf_mf / x);
      hr = x - (c * c * c) / (3 * sq(x));
    } else if (2. * w2 * dt <= cf_mf * criaut) {
      hcf = 0.;
      hr = 0.;
    } else {
      hcf = sq(2. * w2 * dt - criaut * cf_mf) / (2. * sq(x));
      hr = (4. * (x * x * x) - 3. * w2 * dt * sq(c) + (c * c * c)) / (3 * sq(x));
    }
    hli_hcf = dmin(1., hli_hcf + hcf * cf_mf);
    hli_hri = hli_hri + hr;
  }
  {  // adjusted state with the mass-flux condensate (POUT_*, :449-466)
    double v1 = rc_mf, v2 = ri_mf;
    if (br && v1 + v2 > rv_out) *br |= 1u << ICE3_ABR_MFOUT;
    if (v1 + v2 > rv_out) {
      v1 = v1 * rv_out / (v1 + v2);
      v2 = rv_out - v1;
    }
    rc_out = rc_out + v1;
    ri_out = ri_out + v2;
    rv_out = rv_out - (v1 + v2);
    zt = zt + (v1 * lv + v2 * ls) / cph;
  }
  o.t = zt; o.rv = rv_out; o.rc = rc_out; o.ri = ri_out; o.cldfr = cldfr;
  o.hlc_hrc = hlc_hrc; o.hlc_hcf = hlc_hcf; o.hli_hri = hli_hri; o.hli_hcf = hli_hcf;
  o.cph = cph; o.lv = lv; o.ls = ls; o.srcs = sigrc;
}

// ---------------------------------------------------------------- rain_ice

// Per-point inputs of RAIN_ICE's OCND2 branch: the TIWMX tables, ICE_ADJUST's ice cloud
// fraction and super-/sub-saturation over ice (PICLDFR, PSSIO, PSSIU, PIFR) and RAIN_ICE's
// model level height ZZZZ (micro/rain_ice.F90:387-395).
struct MicroOcnd2 {
  const double *tiwmx;
  double icldfr, ssio, ssiu, ifr, zzz;
};

// Entry of a TIWMX table at temperature t (the reference's NINT(XNDEGR * T) lookup).
__device__ __forceinline__ double tiwmx_at(const double *tiwmx, int table, double t) {
  return tiwmx[table * ICE3_TIWMX_N + tiwmx_index(t)];
}

// ICE4_NUCLEATION (micro/ice4_nucleation.func.h:68-143), LFEEDBACKT.  With OCND2 (FP64 only)
// the nucleated concentration is the Meyers function of the tables, reduced between 0 and
// -20 C and with height, times the ice cloud fraction, and no vapour is taken (:127-139).
template <class R, bool OCND2 = false>
__device__ __forceinline__ void nucleation(bool ldcompute, double tht, double pabst, double rhodref,
                                  double exn, double lsfact, double t, double rvt, double &cit,
                                  double &rvheni_mr, const MicroOcnd2 *x = nullptr) {
  const bool gnegt = ldcompute && t < P.XTT && rvt > P.XRTMIN[1];
  rvheni_mr = 0.;
  if (!gnegt) return;
  if constexpr (OCND2) {
    const double zredin = tiwmx_at(x->tiwmx, ICE3_TW_REDIN, t);
    const double zam3 = tiwmx_at(x->tiwmx, ICE3_TW_AM3, dmax(P.XFRMIN[27], t));
    const double q = (20000. - dmin(20000., x->zzz)) / 20000.;
    double zzw = zredin * dmax(0.1, sq(q) * sq(q)) * zam3 * (0.0001 + 0.9999 * x->icldfr);
    zzw = dmin(zzw, 50.E3);
    cit = dmax(zzw + cit, cit);
    return;
  }
  const R rt = (R)t;
  const R lt = rlog<R>(rt);
  R usw = rexp<R>((R)P.XALPW - (R)P.XBETAW / rt - (R)P.XGAMW * lt);
  R zzw = rexp<R>((R)P.XALPI - (R)P.XBETAI / rt - (R)P.XGAMI * lt);
  const R pab = (R)pabst;
  zzw = dmin(pab / (R)2., zzw);
  R ssi = (R)rvt * (pab - zzw) / ((R)P.XEPSILO * zzw) - (R)1.0;
  usw = dmin(pab / (R)2., usw);
  usw = (usw / zzw) * ((pab - zzw) / (pab - usw)) - (R)1.0;
  ssi = dmin(ssi, usw);
  R nn = 0.;
  if (t < P.XTT - 5.0 && ssi > (R)0.0)
    nn = (R)P.XNU20 * rexp<R>((R)P.XALPHA2 * ssi - (R)P.XBETA2);
  else if (t <= P.XTT - 2.0 && t >= P.XTT - 5.0 && ssi > (R)0.0)
    nn = dmax((R)P.XNU20 * rexp<R>(-(R)P.XBETA2),
              (R)P.XNU10 * rexp<R>(-(R)P.XBETA1 * (rt - (R)P.XTT)) * rpow<R>(ssi / usw, (R)P.XALPHA1));
  double zz = (double)nn - cit;
  zz = dmin(zz, 50.E3);
  double rv_mr = dmax(zz, 0.0) * P.XMNU0 / rhodref;
  rv_mr = dmin(rvt, rv_mr);
  const double zw = dmin(rv_mr, dmax(0., (P.XTT / exn - tht) / lsfact)) / dmax(rv_mr, 1.E-20);
  rvheni_mr = rv_mr * zw;
  zz = zz * zw;
  cit = dmax(zz + cit, cit);
}

// Per-point state persisting across the tendency calls of one time step
// (ZBU_INST, ZRS_TEND, ZRG_TEND and ZSSI of micro/mode_ice4_stepping.F90:132-158):
// only the rates that soft calls reuse are kept.
template <class R> struct StepState {
  R rchoni, rvdeps, riaggs, riauts, rvdepg, rcautr, rcaccr, rrevav, rcberi;
  R ricfrrg, rrcfrig, ricfrr, rsmltg, rcmltsr, rgmltr;
  R rs_rcrims, rs_rcrimss, rs_rsrimcg, rs_rraccs, rs_rraccss, rs_rsaccrg, rs_freez1, rs_freez2;
  R rg_rcdryg, rg_ridryg, rg_riwetg, rg_rsdryg, rg_rswetg, rg_rrdryg, rg_freez1, rg_freez2;
  R ssi;
};

// With hail (KRR=7) also ZRH_TEND (micro/mode_ice4_stepping.F90:156) and the hail melting
// rate of ZBU_INST.
template <class R> struct StepStateHail : StepState<R> {
  R rh_rcweth, rh_rrweth, rh_ridryh, rh_riweth, rh_rsdryh, rh_rsweth, rh_rgdryh, rh_rgweth,
      rh_freez1, rh_freez2, rhmltr;
};
template <class R, bool HAIL> struct StepStateFor { using type = StepState<R>; };
template <class R> struct StepStateFor<R, true> { using type = StepStateHail<R>; };

// INTERP_MICRO_1D (micro/interp_micro.func.h:34-88) for one point
// (the argument is ln x: the callers keep the logarithm of each slope parameter)
template <class R, class TB>
__device__ __forceinline__ R interp1d(R lnx, int knum, R p1, R p2, const TB *__restrict__ lt) {
  R idx = dmax((R)1.00001, dmin((R)knum - (R)0.00001, p1 * lnx + p2));
  int i = (int)idx;
  R f = idx - (R)i;
  return (R)lt[i] * f - (R)lt[i - 1] * (f - (R)1.0);
}

// INTERP_MICRO_2D (micro/interp_micro.func.h:159-223) for one point: indices + weights
template <class R> struct Idx2 { int i1, i2; R f1, f2; };
template <class R>
__device__ __forceinline__ Idx2<R> interp2d_idx(R lnx1, R lnx2, int n1, int n2, R p11, R p12, R p21,
                                                R p22) {
  Idx2<R> r;
  R a = dmax((R)1.00001, dmin((R)n1 - (R)0.00001, p11 * lnx1 + p12));
  r.i1 = (int)a;
  r.f1 = a - (R)r.i1;
  R b = dmax((R)1.00001, dmin((R)n2 - (R)0.00001, p21 * lnx2 + p22));
  r.i2 = (int)b;
  r.f2 = b - (R)r.i2;
  return r;
}
// table lt is (n1, n2) column-major, addressed lt(i1, i2) (1-based)
template <class R, class TB>
__device__ __forceinline__ R interp2d(const Idx2<R> &x, int n1, const TB *__restrict__ lt) {
#define T(i, j) ((R)lt[((i)-1) + ((j)-1) * n1])
  return (T(x.i1 + 1, x.i2 + 1) * x.f2 - T(x.i1 + 1, x.i2) * (x.f2 - (R)1.0)) * x.f1 -
         (T(x.i1, x.i2 + 1) * x.f2 - T(x.i1, x.i2) * (x.f2 - (R)1.0)) * (x.f1 - (R)1.0);
#undef T
}
// transposed addressing lt(i2, i1) of an (n2, n1) table (:211-216)
template <class R, class TB>
__device__ __forceinline__ R interp2d_t(const Idx2<R> &x, int n2, const TB *__restrict__ lt) {
#define T(i, j) ((R)lt[((i)-1) + ((j)-1) * n2])
  return (T(x.i2 + 1, x.i1 + 1) * x.f1 - T(x.i2 + 1, x.i1) * (x.f1 - (R)1.0)) * x.f2 -
         (T(x.i2, x.i1 + 1) * x.f1 - T(x.i2, x.i1) * (x.f1 - (R)1.0)) * (x.f2 - (R)1.0);
#undef T
}

// Powers of the reference density, constant over the time splitting of one point.
// x^c is evaluated as exp(c*ln x) throughout the hard tendency call: one logarithm
// per base (rho, T, each slope parameter) instead of one per power.
template <class R> struct PointConst {
  R lnrho;       // ln rhodref
  R rhomcexvt;   // rho^-XCEXVT
  R rhom;        // rho^(-XCEXVT-1)
  R rho03;       // rho^0.3
};

template <class R>
__device__ __forceinline__ PointConst<R> point_const(double rhodref) {
  PointConst<R> c;
  c.lnrho = rlog<R>((R)rhodref);
  c.rhomcexvt = rexp<R>(-(R)P.XCEXVT * c.lnrho);
  c.rhom = rexp<R>((-(R)P.XCEXVT - (R)1.) * c.lnrho);
  c.rho03 = rexp<R>((R)0.3 * c.lnrho);
  return c;
}

// x^c for x > 0 from ln x
template <class R> __device__ __forceinline__ R powl_(R lnx, R c) { return rexp<R>(c * lnx); }

// ICE4_TENDENCIES (micro/mode_ice4_tendencies.F90:155-717) for one point that is
// being integrated (LDCOMPUTE true), with ICE4_COMPUTE_PDF, ICE4_SLOW, ICE4_WARM,
// ICE4_FAST_RS, ICE4_FAST_RG and ICE4_FAST_RI inlined in call order.
// Outputs: pa (rates) and pb (instantaneous changes), both double.
// Soft calls (reuse of the hard call's rates) skip everything that only feeds
// rate evaluations guarded by .NOT.LDSOFT, such as the slope parameters.
//
// OCND2 (PARAM_ICE LOCND2 = .TRUE., FP64 only; x holds the point's OCND2 inputs, read by
// hard calls): the nucleation of the tables; ZESI and ZAI from the TIWMX tables (:304-310);
// the collision factors ZCOLF, ZACRF (:321-345) in the riming and the rain accretion of cloud
// water; the deposition tuning XFRMIN(40)/XFRMIN(39) (micro/mode_ice4_slow.F90:92-100); no
// snow-graupel collection (micro/mode_ice4_fast_rg.F90:185-188,207-209); ICE4_FAST_RI_RS
// (micro/mode_ice4_fast_ri_rs.F90) in place of ICE4_FAST_RI; their PA terms (:625-691).
//
// HAIL (KRR=7, HCLOUD='ICE4'; S is then a StepStateHail): pvart, pa and pb have an eighth
// entry, hail; the hail slope parameter (:426-432), the conversion of wet-growing graupel
// into hail in ICE4_FAST_RG (micro/mode_ice4_fast_rg.F90:330-351), ICE4_FAST_RH
// (micro/mode_ice4_fast_rh.F90) and their PA / PB terms (:619-716); htab then holds the
// wet-growth kernels of hail (the block that follows the ICE3 tables, ICE3_T_ICE3 on).
template <class R, bool OCND2 = false, bool HAIL = false, class TB, class SS>
__device__ __forceinline__ void tendencies(bool soft, double exn, double rhodref, double lvfact,
                                  double lsfact, double pres, double cf, double sigma_rc,
                                  double &cit, double pt, const double *pvart,
                                  SS &S, const PointConst<R> &pc,
                                  const TB *__restrict__ tab, double *pa, double *pb,
                                  const MicroOcnd2 *x = nullptr,
                                  const double *__restrict__ htab = nullptr,
                                  unsigned *br = nullptr) {
  static_assert(!OCND2 || sizeof(R) == sizeof(double), "the OCND2 branch is FP64");
  static_assert(!(OCND2 && HAIL), "OCND2 with hail is not built");
  constexpr int NV = HAIL ? 8 : 7;
  double zt = pt;
  double zv[NV];
#pragma unroll
  for (int jv = 0; jv < NV; ++jv) { zv[jv] = pvart[jv]; pa[jv] = 0.; pb[jv] = 0.; }
  double rrhong_mr = 0.;  // RRHONG of the hard call (graupel produced at once)
  R ka = 0, dv = 0, ai = 0, cj = 0, lnt = 0, esi = 0, esw = 0;
  double esi_tab = 0., colf = 1., acrf = 1.;  // OCND2: tabulated ZESI, ZCOLF, ZACRF
  if (!soft) {
    double rvheni, rrhong, rimltc;
    nucleation<R, OCND2>(true, zv[ITH], pres, rhodref, exn, lsfact, zt, zv[IRV], cit, rvheni, x);
    zv[ITH] = zv[ITH] + rvheni * lsfact;
    zt = zv[ITH] * exn;
    zv[IRV] = zv[IRV] - rvheni;
    zv[IRI] = zv[IRI] + rvheni;
    // RRHONG (micro/mode_ice4_rrhong.F90:65-75)
    if (zt < P.XTT - 35.0 && zv[IRR] > P.XRTMIN[3])
      rrhong = dmin(zv[IRR], dmax(0., ((P.XTT - 35.) / exn - zv[ITH]) / (lsfact - lvfact)));
    else
      rrhong = 0.;
    if (br && rrhong > 0.) *br |= 1u << ICE3_BR_RRHONG;
    zv[ITH] = zv[ITH] + rrhong * (lsfact - lvfact);
    zt = zv[ITH] * exn;
    zv[IRR] = zv[IRR] - rrhong;
    zv[IRG] = zv[IRG] + rrhong;
    rrhong_mr = rrhong;
    // RIMLTC (micro/mode_ice4_rimltc.F90:64-74)
    if (zv[IRI] > 0. && zt > P.XTT)
      rimltc = dmin(zv[IRI], dmax(0., (zv[ITH] - P.XTT / exn) / (lsfact - lvfact)));
    else
      rimltc = 0.;
    if (br && rimltc > 0.) *br |= 1u << ICE3_BR_RIMLTC;
    zv[ITH] = zv[ITH] - rimltc * (lsfact - lvfact);
    zt = zv[ITH] * exn;
    zv[IRC] = zv[IRC] + rimltc;
    zv[IRI] = zv[IRI] - rimltc;
    pb[ITH] = pb[ITH] + rvheni * lsfact;
    pb[ITH] = pb[ITH] + rrhong * (lsfact - lvfact);
    pb[ITH] = pb[ITH] - rimltc * (lsfact - lvfact);
    pb[IRV] = pb[IRV] - rvheni;
    pb[IRC] = pb[IRC] + rimltc;
    pb[IRR] = pb[IRR] - rrhong;
    pb[IRI] = pb[IRI] + rvheni;
    pb[IRI] = pb[IRI] - rimltc;
    pb[IRG] = pb[IRG] + rrhong;
    // derived fields (:312-320)
    const R t = (R)zt;
    lnt = rlog<R>(t);
    esi = rexp<R>((R)P.XALPI - (R)P.XBETAI / t - (R)P.XGAMI * lnt);
    esw = rexp<R>((R)P.XALPW - (R)P.XBETAW / t - (R)P.XGAMW * lnt);
    if constexpr (OCND2) {
      esi_tab = tiwmx_at(x->tiwmx, ICE3_TW_ESI, zt);
      S.ssi = (R)(zv[IRV] * (pres - esi_tab) / (P.XEPSILO * esi_tab) - 1.0);
      ka = (R)2.38E-2 + (R)0.0071E-2 * (t - (R)P.XTT);
      dv = (R)0.211E-4 * powl_<R>(lnt - (R)P.LN_XTT, (R)1.94) * ((R)P.XP00 / (R)pres);
      ai = (R)(tiwmx_at(x->tiwmx, ICE3_TW_A2, zt) + tiwmx_at(x->tiwmx, ICE3_TW_BB3, zt) * pres);
    } else {
      const R zzw = esi;
      S.ssi = (R)zv[IRV] * ((R)pres - zzw) / ((R)P.XEPSILO * zzw) - (R)1.0;
      ka = (R)2.38E-2 + (R)0.0071E-2 * (t - (R)P.XTT);
      dv = (R)0.211E-4 * powl_<R>(lnt - (R)P.LN_XTT, (R)1.94) * ((R)P.XP00 / (R)pres);
      ai = sq((R)P.XLSTT + (R)(P.XCPV - P.XCI) * (t - (R)P.XTT)) / (ka * (R)P.XRV * sq(t)) +
           ((R)P.XRV * t) / (dv * zzw);
    }
    cj = (R)P.XSCFAC * pc.rho03 / sqrt((R)1.718E-5 + (R)0.0049E-5 * (t - (R)P.XTT));
    if constexpr (OCND2) {  // collision factors, PCONC = XCONC_LAND (:321-345)
      colf = 0.00001;
      acrf = 0.00001;
      if (zv[IRC] > 1.0E-10) {
        const double conc = P.XCONC_LAND * 0.000001;
        const double rcw = 0.1 * rpow<double>(0.75 * zv[IRC] * rhodref / (P.XPI * conc), 0.333);
        double vt = rcw < 0.0065 ? 1.19E6 * sq(rcw) : 8000. * rcw;
        vt = dmin(10., vt);
        const double st = dmax(0.01, 2. * (100. - vt) * vt / (P.XG * 10.));
        if (st > 0.1) colf = dmax(0.01, dmin(1., 0.939 * rpow<double>(st, 2.657)));
        if (zv[IRR] > 1.0E-10 && rcw > 1.0E-5) {
          const double r20 = exp_f64(rcw * 2000.);
          acrf = (r20 - 1.) / (r20 + 1.);
        }
      }
    }
  }
  const R t = (R)zt, rho = (R)rhodref;
  const R rv = (R)zv[IRV], rc = (R)zv[IRC], rr = (R)zv[IRR], ri = (R)zv[IRI], rs = (R)zv[IRS],
          rg = (R)zv[IRG];
  const bool cold = zt < P.XTT, warm = zt > P.XTT;

  // ICE4_COMPUTE_PDF (micro/mode_ice4_compute_pdf.F90:88-390): rc 'PDF'+'SIGM', ri 'NONE'
  R hlc_hcf, hlc_hrc, hli_hcf, hli_hri, criauti;
  {
    const double rcrautc = P.XCRIAUTC / rhodref, zrc = zv[IRC];
    if (zrc > rcrautc + sigma_rc) {
      hlc_hcf = 1.; hlc_hrc = (R)zrc;
    } else if (zrc > (rcrautc - sigma_rc) && zrc <= (rcrautc + sigma_rc)) {
      hlc_hcf = (R)((zrc + sigma_rc - rcrautc) / (2. * sigma_rc));
      hlc_hrc = (R)((zrc + sigma_rc - rcrautc) * (zrc + sigma_rc + rcrautc) / (4. * sigma_rc));
    } else {
      hlc_hcf = 0.; hlc_hrc = 0.;
    }
    criauti = dmin((R)P.XCRIAUTI, rexp<R>((R)P.LN_10 * ((R)P.XACRIAUTI * (t - (R)P.XTT) + (R)P.XBCRIAUTI)));
    if (ri > criauti) {
      hli_hcf = 1.; hli_hri = ri;
    } else {
      hli_hcf = 0.; hli_hri = 0.;
    }
  }

  // slope parameters (:389-433), hard calls only (soft calls reuse rates)
  // ln(lbda) is kept beside each lbda: every power and table index below uses it.
  R lbdar = 0, lbdas = 0, lbdag = 0, llr = 0, lls = 0, llg = 0;
  if (!soft) {
    if (rr > (R)0.) {
      llr = (R)P.LN_XLBR + (R)P.XLBEXR * rlog<R>(rho * dmax(rr, (R)P.XRTMIN[3]));
      lbdar = rexp<R>(llr);
    }
    if (rs > (R)0.) {
      lls = (R)P.LN_XLBS + (R)P.XLBEXS * rlog<R>(rho * dmax(rs, (R)P.XRTMIN[5]));
      if (lls < (R)P.LN_XLBDAS_MAX) {
        lbdas = rexp<R>(lls);
      } else {
        lbdas = (R)P.XLBDAS_MAX;
        lls = (R)P.LN_XLBDAS_MAX;
      }
    }
    if (rg > (R)0.) {
      llg = (R)P.LN_XLBG + (R)P.XLBEXG * rlog<R>(rho * dmax(rg, (R)P.XRTMIN[6]));
      lbdag = rexp<R>(llg);
    }
  }
  R lbdah = 0, llh = 0, vent_h = 0;
  if constexpr (HAIL) {
    if (!soft && zv[IRH] > 0.) {
      llh = (R)P.LN_XLBH + (R)P.XLBEXH * rlog<R>(rho * dmax((R)zv[IRH], (R)P.XRTMIN[7]));
      lbdah = rexp<R>(llh);
    }
  }
  const R rhomcexvt = soft ? (R)0 : pc.rhomcexvt;
  // powers of the snow and graupel slopes used by several processes of the hard call
  // (vapour deposition, freezing budget, melting; accretion and dry growth)
  R vent_s = 0, vent_g = 0, lbs_cxs_bs = 0, lbg_cxg = 0;
  if (!soft) {
    if (rs > (R)0.) {
      vent_s = (R)P.X0DEPS * powl_<R>(lls, (R)P.XEX0DEPS) + (R)P.X1DEPS * cj * powl_<R>(lls, (R)P.XEX1DEPS);
      lbs_cxs_bs = powl_<R>(lls, (R)(P.XCXS - P.XBS));
    }
    if (rg > (R)0.) {
      vent_g = (R)P.X0DEPG * powl_<R>(llg, (R)P.XEX0DEPG) + (R)P.X1DEPG * cj * powl_<R>(llg, (R)P.XEX1DEPG);
      lbg_cxg = powl_<R>(llg, (R)P.XCXG);
    }
    if constexpr (HAIL)
      if (zv[IRH] > 0.)
        vent_h = (R)P.X0DEPH * powl_<R>(llh, (R)P.XEX0DEPH) + (R)P.X1DEPH * cj * powl_<R>(llh, (R)P.XEX1DEPH);
  }
  const R ssi = S.ssi;
  const bool rv_ok = zv[IRV] > P.XRTMIN[1], rc_ok = zv[IRC] > P.XRTMIN[2],
             rr_ok = zv[IRR] > P.XRTMIN[3], ri_ok = zv[IRI] > P.XRTMIN[4],
             rs_ok = zv[IRS] > P.XRTMIN[5], rg_ok = zv[IRG] > P.XRTMIN[6];

  // ICE4_SLOW (micro/mode_ice4_slow.F90:106-217)
  if (zt < P.XTT - 35.0 && rc_ok) {
    if (!soft) S.rchoni = dmin((R)1000., (R)P.XHON * rho * rc * rexp<R>((R)P.XALPHA3 * (t - (R)P.XTT) - (R)P.XBETA3));
  } else S.rchoni = 0;
  if (rv_ok && rs_ok) {
    if (!soft) {
      S.rvdeps = (ssi / (rho * ai)) * vent_s;
      if constexpr (OCND2) S.rvdeps = S.rvdeps * (R)P.XFRMIN[40];
    }
  } else S.rvdeps = 0;
  if (ri_ok && rs_ok) {
    if (!soft) S.riaggs = (R)P.XFIAGGS * rexp<R>((R)P.XCOLEXIS * (t - (R)P.XTT)) * ri * powl_<R>(lls, (R)P.XEXIAGGS) * rhomcexvt;
  } else S.riaggs = 0;
  if (hli_hri > (R)P.XRTMIN[4]) {
    if (!soft) {
      S.riauts = (R)P.XTIMAUTI * rexp<R>((R)P.XTEXAUTI * (t - (R)P.XTT)) * dmax(hli_hri - criauti * hli_hcf, (R)0.);
    }
  } else S.riauts = 0;
  if (rv_ok && rg_ok) {
    if (!soft) {
      S.rvdepg = (ssi / (rho * ai)) * vent_g;
      if constexpr (OCND2) S.rvdepg = S.rvdepg * (R)P.XFRMIN[39];
    }
  } else S.rvdepg = 0;
  // graupel deposition partly to snow with OCND2 (micro/mode_ice4_tendencies.F90:445-457)
  double zdepg_s = 0., znodepg_s = 1.;
  if constexpr (OCND2) {
    if (P.XFRMIN[5] > 1.0E-12 && P.XFRMIN[6] > 0.01) {
      zdepg_s = dmax(0., dmin(1., (P.XFRMIN[5] - zv[IRG] / P.XFRMIN[5]))) *
                dmax(0., dmin(1., ((double)ssi / P.XFRMIN[6])));
      znodepg_s = 1. - zdepg_s;
    }
  }

  // ICE4_WARM (micro/mode_ice4_warm.F90:100-197)
  if (hlc_hrc > (R)P.XRTMIN[2] && hlc_hcf > (R)0.) {
    if (!soft) S.rcautr = (R)P.XTIMAUTC * dmax(hlc_hrc - hlc_hcf * (R)P.XCRIAUTC / rho, (R)0.0);
  } else S.rcautr = 0;
  if (rc_ok && rr_ok) {
    if (!soft) S.rcaccr = (R)P.XFCACCR * rc * (R)acrf * powl_<R>(llr, (R)P.XEXCACCR) * rhomcexvt;
  } else S.rcaccr = 0;
  if (rr_ok && !rc_ok) {
    if (!soft) {
      const R ev = esw;
      const R usw = (R)1. - rv * ((R)pres - ev) / ((R)P.XEPSILO * ev);
      const R x = sq((R)P.XLVTT + (R)(P.XCPV - P.XCL) * (t - (R)P.XTT)) / (ka * (R)P.XRV * sq(t)) + ((R)P.XRV * t) / (dv * ev);
      S.rrevav = (dmax((R)0., usw) / (rho * x)) * ((R)P.X0EVAR * powl_<R>(llr, (R)P.XEX0EVAR) + (R)P.X1EVAR * cj * powl_<R>(llr, (R)P.XEX1EVAR));
    }
  } else S.rrevav = 0;

  // ICE4_FAST_RS (micro/mode_ice4_fast_rs.F90:109-441)
  R freez;
  if (rs_ok) {
    if (!soft) {
      R f1 = rv * (R)pres / ((R)P.XEPSILO + rv);
      f1 = dmin(f1, esi);
      f1 = ka * ((R)P.XTT - t) + (dv * ((R)P.XLVTT + (R)(P.XCPV - P.XCL) * (t - (R)P.XTT)) * ((R)P.XESTT - f1) / ((R)P.XRV * t));
      f1 = f1 * vent_s /
           (rho * ((R)P.XLMTT - (R)P.XCL * ((R)P.XTT - t)));
      S.rs_freez1 = f1;
      S.rs_freez2 = (rho * ((R)P.XLMTT + (R)(P.XCI - P.XCL) * ((R)P.XTT - t))) / (rho * ((R)P.XLMTT - (R)P.XCL * ((R)P.XTT - t)));
    }
    freez = dmax((R)0., dmax((R)0., S.rs_freez1 + S.rs_freez2 * S.riaggs) - S.riaggs);
  } else {
    S.rs_freez1 = 0; S.rs_freez2 = 0; freez = 0;
  }
  const bool grim = rc_ok && rs_ok;
  if (!grim) { S.rs_rcrims = 0; S.rs_rcrimss = 0; S.rs_rsrimcg = 0; }
  if (!soft && grim) {
    const R zzw1 = interp1d<R>(lls, P.NGAMINC, (R)P.XRIMINTP1, (R)P.XRIMINTP2, tab + ICE3_T_RIM1);
    const R zzw2 = interp1d<R>(lls, P.NGAMINC, (R)P.XRIMINTP1, (R)P.XRIMINTP2, tab + ICE3_T_RIM2);
    const R zzw3 = interp1d<R>(lls, P.NGAMINC, (R)P.XRIMINTP1, (R)P.XRIMINTP2, tab + ICE3_T_RIM4);
    const R lexc = powl_<R>(lls, (R)P.XEXCRIMSS);
    S.rs_rcrimss = (R)P.XCRIMSS * zzw1 * rc * (R)colf * lexc * rhomcexvt;
    S.rs_rcrims = (R)P.XCRIMSG * rc * (R)colf * lexc * rhomcexvt;
    const R zzw = S.rs_rcrims - S.rs_rcrimss;
    const R x = (R)P.XSRIMCG * powl_<R>(lls, (R)P.XEXSRIMCG) * ((R)1.0 - zzw2);
    S.rs_rsrimcg = zzw * x / dmax((R)1.E-20, (R)P.XSRIMCG3 * (R)P.XSRIMCG2 * powl_<R>(lls, (R)P.XEXSRIMCG2) * ((R)1. - zzw3) - (R)P.XSRIMCG3 * x);
  }
  R rcrimss = 0, rcrimsg = 0, rsrimcg = 0;
  if (grim && cold) {
    rcrimss = dmin(freez, S.rs_rcrimss);
    freez = dmax((R)0., freez - rcrimss);
    const R zzw0d = dmin((R)1., freez / dmax((R)1.E-20, S.rs_rcrims - rcrimss));
    rcrimsg = zzw0d * dmax((R)0., S.rs_rcrims - rcrimss);
    freez = dmax((R)0., freez - rcrimsg);
    rsrimcg = zzw0d * S.rs_rsrimcg;
    rsrimcg = rsrimcg * dmax((R)0., -copysign((R)1., -rcrimsg));
    rcrimsg = dmax((R)0., rcrimsg);
  }
  const bool gacc = rr_ok && rs_ok;
  if (!gacc) { S.rs_rraccs = 0; S.rs_rraccss = 0; S.rs_rsaccrg = 0; }
  if (!soft) {
    S.rs_rraccs = 0; S.rs_rraccss = 0; S.rs_rsaccrg = 0;
    if (gacc) {
      const Idx2<R> ix = interp2d_idx<R>(lls, llr, P.NACCLBDAS, P.NACCLBDAR, (R)P.XACCINTP1S,
                                         (R)P.XACCINTP2S, (R)P.XACCINTP1R, (R)P.XACCINTP2R);
      const R zzw1 = interp2d<R>(ix, P.NACCLBDAS, tab + ICE3_T_RACCSS);
      const R zzw2 = interp2d<R>(ix, P.NACCLBDAS, tab + ICE3_T_RACCS);
      const R zzw3 = interp2d_t<R>(ix, P.NACCLBDAR, tab + ICE3_T_SACCRG);
      const R rhom = pc.rhom;
      const R lr2 = lbdar * lbdar;
      const R zzw = (R)P.XFRACCSS * (powl_<R>(lls, (R)P.XCXS)) * rhom *
                    ((R)P.XLBRACCS1 / (sq(lbdas)) + (R)P.XLBRACCS2 / (lbdas * lbdar) + (R)P.XLBRACCS3 / (lr2)) / (lr2 * lr2);
      S.rs_rraccss = zzw1 * zzw;
      if constexpr (OCND2) S.rs_rraccss = S.rs_rraccss * (R)P.XFRMIN[7];
      S.rs_rraccs = zzw2 * zzw;
      if (!OCND2 || zv[IRS] > P.XFRMIN[1])
        S.rs_rsaccrg = (R)P.XFSACCRG * zzw3 * lbs_cxs_bs * rhom *
                       ((R)P.XLBSACCR1 / (lr2) + (R)P.XLBSACCR2 / (lbdar * lbdas) + (R)P.XLBSACCR3 / (sq(lbdas))) / lbdar;
    }
  }
  R rraccss = 0, rraccsg = 0, rsaccrg = 0;
  if (gacc && cold) {
    rraccss = dmin(freez, S.rs_rraccss);
    freez = dmax((R)0., freez - rraccss);
    const R zzw = dmin((R)1., freez / dmax((R)1.E-20, S.rs_rraccs - rraccss));
    rraccsg = zzw * dmax((R)0., S.rs_rraccs - rraccss);
    freez = dmax((R)0., freez - rraccsg);
    rsaccrg = zzw * S.rs_rsaccrg;
    rsaccrg = rsaccrg * dmax((R)0., -copysign((R)1., -rraccsg));
    rraccsg = dmax((R)0., rraccsg);
  }
  if (rs_ok && warm) {
    if (!soft) {
      R m = rv * (R)pres / ((R)P.XEPSILO + rv);
      m = dmin(m, esw);
      m = ka * ((R)P.XTT - t) + (dv * ((R)P.XLVTT + (R)(P.XCPV - P.XCL) * (t - (R)P.XTT)) * ((R)P.XESTT - m) / ((R)P.XRV * t));
      S.rsmltg = (R)P.XFSCVMG * dmax((R)0., (-m * vent_s -
                                              (S.rs_rcrims + S.rs_rraccs) * (rho * (R)P.XCL * ((R)P.XTT - t))) / (rho * (R)P.XLMTT));
      S.rcmltsr = S.rs_rcrims;
    }
  } else {
    S.rsmltg = 0; S.rcmltsr = 0;
  }

  // ICE4_FAST_RG (micro/mode_ice4_fast_rg.F90:121-410)
  if (ri_ok && rr_ok) {
    if (!soft) {
      R icf = (R)P.XICFRR * ri * powl_<R>(llr, (R)P.XEXICFRR) * rhomcexvt;
      R rcf = (R)P.XRCFRI * (R)cit * powl_<R>(llr, (R)P.XEXRCFRI) * pc.rhom;
      const R zzw0d = dmax((R)0., dmin((R)1., (icf * (R)P.XCI + rcf * (R)P.XCL) * ((R)P.XTT - t) / dmax((R)1.E-20, (R)P.XLVTT * rcf)));
      S.rrcfrig = zzw0d * rcf;
      S.ricfrr = ((R)1. - zzw0d) * icf;
      S.ricfrrg = zzw0d * icf;
    }
  } else {
    S.ricfrrg = 0; S.rrcfrig = 0; S.ricfrr = 0;
  }
  const R lgx = (!soft && rg_ok) ? powl_<R>(llg, (R)(P.XCXG - P.XDG - 2.)) * rhomcexvt : (R)0;
  if (rg_ok && rc_ok) {
    if (!soft) S.rg_rcdryg = (R)P.XFCDRYG * rc * lgx;
  } else S.rg_rcdryg = 0;
  if (rg_ok && ri_ok) {
    if (!soft) {
      const R e = rexp<R>((R)P.XCOLEXIG * (t - (R)P.XTT));
      S.rg_ridryg = (R)P.XFIDRYG * e * ri * lgx;
      S.rg_riwetg = S.rg_ridryg / ((R)P.XCOLIG * e);
    }
  } else {
    S.rg_ridryg = 0; S.rg_riwetg = 0;
  }
  const R rhom1 = (!soft && rg_ok) ? pc.rhom : (R)0;
  const bool gdrys = rs_ok && rg_ok;
  if (!gdrys || OCND2) { S.rg_rsdryg = 0; S.rg_rswetg = 0; }
  if (!OCND2 && !soft && gdrys) {
    const Idx2<R> ix = interp2d_idx<R>(llg, lls, P.NDRYLBDAG, P.NDRYLBDAS, (R)P.XDRYINTP1G,
                                       (R)P.XDRYINTP2G, (R)P.XDRYINTP1S, (R)P.XDRYINTP2S);
    const R zzw = interp2d<R>(ix, P.NDRYLBDAG, tab + ICE3_T_SDRYG);
    S.rg_rswetg = (R)P.XFSDRYG * zzw / (R)P.XCOLSG * lbs_cxs_bs * lbg_cxg * rhom1 *
                  ((R)P.XLBSDRYG1 / (sq(lbdag)) + (R)P.XLBSDRYG2 / (lbdag * lbdas) + (R)P.XLBSDRYG3 / (sq(lbdas)));
    S.rg_rsdryg = S.rg_rswetg * (R)P.XCOLSG * rexp<R>((R)P.XCOLEXSG * (t - (R)P.XTT));
  }
  const bool gdryr = rr_ok && rg_ok;
  if (!gdryr) S.rg_rrdryg = 0;
  if (!soft && gdryr) {
    const Idx2<R> ix = interp2d_idx<R>(llg, llr, P.NDRYLBDAG, P.NDRYLBDAR, (R)P.XDRYINTP1G,
                                       (R)P.XDRYINTP2G, (R)P.XDRYINTP1R, (R)P.XDRYINTP2R);
    const R zzw = interp2d<R>(ix, P.NDRYLBDAG, tab + ICE3_T_RDRYG);
    const R lr2 = lbdar * lbdar;
    S.rg_rrdryg = (R)P.XFRDRYG * zzw * ((R)1.0 / (lr2 * lr2)) * lbg_cxg * rhom1 *
                  ((R)P.XLBRDRYG1 / (sq(lbdag)) + (R)P.XLBRDRYG2 / (lbdag * lbdar) + (R)P.XLBRDRYG3 / (lr2));
  }
  const R rdryg_init = S.rg_rcdryg + S.rg_ridryg + S.rg_rsdryg + S.rg_rrdryg;
  R rwetg_init = 0;
  bool ldwetg = false, lldryg = false;
  if (rg_ok) {
    if (!soft) {
      R f1 = rv * (R)pres / ((R)P.XEPSILO + rv);
      f1 = dmin(f1, esi);
      f1 = ka * ((R)P.XTT - t) + (dv * ((R)P.XLVTT + (R)(P.XCPV - P.XCL) * (t - (R)P.XTT)) * ((R)P.XESTT - f1) / ((R)P.XRV * t));
      f1 = f1 * vent_g /
           (rho * ((R)P.XLMTT - (R)P.XCL * ((R)P.XTT - t)));
      S.rg_freez1 = f1;
      S.rg_freez2 = (rho * ((R)P.XLMTT + (R)(P.XCI - P.XCL) * ((R)P.XTT - t))) / (rho * ((R)P.XLMTT - (R)P.XCL * ((R)P.XTT - t)));
    }
    rwetg_init = dmax(S.rg_riwetg + S.rg_rswetg, dmax((R)0., S.rg_freez1 + S.rg_freez2 * (S.rg_riwetg + S.rg_rswetg)));
    const R wet = dmax((R)0., rwetg_init - S.rg_riwetg - S.rg_rswetg);
    const R dry = dmax((R)0., rdryg_init - S.rg_ridryg - S.rg_rsdryg);
    ldwetg = wet <= dry && rdryg_init > (R)0.;
    lldryg = cold && rdryg_init > (R)1.E-20 && wet > dry;
  } else {
    S.rg_freez1 = 0; S.rg_freez2 = 0;
  }
  // part of the wet growth converted into hail (micro/mode_ice4_fast_rg.F90:330-351), from
  // the graupel the other processes produce (micro/mode_ice4_tendencies.F90:509-514)
  R rwetgh = 0, rwetgh_mr = 0;
  if constexpr (HAIL) {
    if (ldwetg) {
      const R rgsi = S.rvdepg + S.rsmltg + rraccsg + rsaccrg + rcrimsg + rsrimcg;
      const R rgsi_mr = (R)rrhong_mr;
      rwetgh = (dmax((R)0., rgsi + S.ricfrrg + S.rrcfrig) + rwetg_init) * rdryg_init /
               (rwetg_init + rdryg_init);
      rwetgh_mr = dmax((R)0., rgsi_mr) * rdryg_init / (rwetg_init + rdryg_init);
    }
  }
  R rrwetg = 0, rcwetg = 0, riwetg = 0, rswetg = 0, rcdryg = 0, rrdryg = 0, ridryg = 0, rsdryg = 0;
  if (ldwetg) {
    rrwetg = -(S.rg_riwetg + S.rg_rswetg + S.rg_rcdryg - rwetg_init);
    rcwetg = S.rg_rcdryg; riwetg = S.rg_riwetg; rswetg = S.rg_rswetg;
  }
  if (lldryg) {
    rcdryg = S.rg_rcdryg; rrdryg = S.rg_rrdryg; ridryg = S.rg_ridryg; rsdryg = S.rg_rsdryg;
  }
  if (rg_ok && warm) {
    if (!soft) {
      R m = rv * (R)pres / ((R)P.XEPSILO + rv);
      m = dmin(m, esw);
      m = ka * ((R)P.XTT - t) + dv * ((R)P.XLVTT + (R)(P.XCPV - P.XCL) * (t - (R)P.XTT)) * ((R)P.XESTT - m) / ((R)P.XRV * t);
      S.rgmltr = dmax((R)0., (-m * vent_g -
                              (S.rg_rcdryg + S.rg_rrdryg) * (rho * (R)P.XCL * ((R)P.XTT - t))) / (rho * (R)P.XLMTT));
    }
  } else S.rgmltr = 0;

  // ICE4_FAST_RH (micro/mode_ice4_fast_rh.F90:109-405), KRR=7: LNULLWETH=T, LWETHPOST=T,
  // LCONVHG=T, LEVLIMIT=T, LNEWCOEFF=F.  Collection of c, i, s, g, r by hail, the wet/dry
  // growth choice, hail to graupel in dry growth, melting.
  R rcweth = 0, riweth = 0, rsweth = 0, rgweth = 0, rrweth = 0;
  R rcdryh = 0, ridryh = 0, rsdryh = 0, rrdryh = 0, rgdryh = 0, rdryhg = 0;
  if constexpr (HAIL) {
    const bool rh_ok = zv[IRH] > P.XRTMIN[7];
    const R lhx = (!soft && rh_ok) ? powl_<R>(llh, (R)(P.XCXH - P.XDH - 2.0)) * rhomcexvt : (R)0;
    const R lbh_cxh = (!soft && rh_ok) ? powl_<R>(llh, (R)P.XCXH) : (R)0;
    if (rh_ok && rc_ok) {
      if (!soft) S.rh_rcweth = (R)P.XFWETH * rc * lhx;
    } else S.rh_rcweth = 0;
    if (rh_ok && ri_ok) {
      if (!soft) {
        S.rh_riweth = (R)P.XFWETH * ri * lhx;
        S.rh_ridryh = S.rh_riweth * ((R)P.XCOLIH * rexp<R>((R)P.XCOLEXIH * (t - (R)P.XTT)));
      }
    } else {
      S.rh_riweth = 0; S.rh_ridryh = 0;
    }
    if (!(rh_ok && rs_ok)) { S.rh_rsweth = 0; S.rh_rsdryh = 0; }
    if (!soft && rh_ok && rs_ok) {
      const Idx2<R> ix = interp2d_idx<R>(llh, lls, P.NWETLBDAH, P.NWETLBDAS, (R)P.XWETINTP1H,
                                         (R)P.XWETINTP2H, (R)P.XWETINTP1S, (R)P.XWETINTP2S);
      const R zzw = interp2d<R>(ix, P.NWETLBDAH, htab + (ICE3_T_SWETH - ICE3_T_ICE3));
      S.rh_rsweth = (R)P.XFSWETH * zzw * lbs_cxs_bs * lbh_cxh * pc.rhom *
                    ((R)P.XLBSWETH1 / (sq(lbdah)) + (R)P.XLBSWETH2 / (lbdah * lbdas) + (R)P.XLBSWETH3 / (sq(lbdas)));
      S.rh_rsdryh = S.rh_rsweth * ((R)P.XCOLSH * rexp<R>((R)P.XCOLEXSH * (t - (R)P.XTT)));
    }
    if (!(rh_ok && rg_ok)) { S.rh_rgweth = 0; S.rh_rgdryh = 0; }
    if (!soft && rh_ok && rg_ok) {
      const Idx2<R> ix = interp2d_idx<R>(llh, llg, P.NWETLBDAH, P.NWETLBDAG, (R)P.XWETINTP1H,
                                         (R)P.XWETINTP2H, (R)P.XWETINTP1G, (R)P.XWETINTP2G);
      const R zzw = interp2d<R>(ix, P.NWETLBDAH, htab + (ICE3_T_GWETH - ICE3_T_ICE3));
      S.rh_rgweth = (R)P.XFGWETH * zzw * powl_<R>(llg, (R)(P.XCXG - P.XBG)) * lbh_cxh * pc.rhom *
                    ((R)P.XLBGWETH1 / (sq(lbdah)) + (R)P.XLBGWETH2 / (lbdah * lbdag) + (R)P.XLBGWETH3 / (sq(lbdag)));
      S.rh_rgdryh = S.rh_rgweth;
      // wet graupel keeps its collection efficiency (:221-223)
      if (!ldwetg) S.rh_rgdryh = S.rh_rgdryh * ((R)P.XCOLGH * rexp<R>((R)P.XCOLEXGH * (t - (R)P.XTT)));
    }
    if (!(rh_ok && rr_ok)) S.rh_rrweth = 0;
    if (!soft && rh_ok && rr_ok) {
      const Idx2<R> ix = interp2d_idx<R>(llh, llr, P.NWETLBDAH, P.NWETLBDAR, (R)P.XWETINTP1H,
                                         (R)P.XWETINTP2H, (R)P.XWETINTP1R, (R)P.XWETINTP2R);
      const R zzw = interp2d<R>(ix, P.NWETLBDAH, htab + (ICE3_T_RWETH - ICE3_T_ICE3));
      const R lr2 = lbdar * lbdar;
      S.rh_rrweth = (R)P.XFRWETH * zzw * ((R)1.0 / (lr2 * lr2)) * lbh_cxh * pc.rhom *
                    ((R)P.XLBRWETH1 / (sq(lbdah)) + (R)P.XLBRWETH2 / (lbdah * lbdar) + (R)P.XLBRWETH3 / (lr2));
    }
    const R rdryh_init = S.rh_rcweth + S.rh_ridryh + S.rh_rsdryh + S.rh_rrweth + S.rh_rgdryh;
    R rweth_init = 0;
    bool llweth = false, lldryh = false;
    if (rh_ok) {
      if (!soft) {
        R f1 = rv * (R)pres / ((R)P.XEPSILO + rv);
        f1 = dmin(f1, esi);
        f1 = ka * ((R)P.XTT - t) + (dv * ((R)P.XLVTT + (R)(P.XCPV - P.XCL) * (t - (R)P.XTT)) * ((R)P.XESTT - f1) / ((R)P.XRV * t));
        f1 = f1 * vent_h / (rho * ((R)P.XLMTT - (R)P.XCL * ((R)P.XTT - t)));
        S.rh_freez1 = f1;
        S.rh_freez2 = (rho * ((R)P.XLMTT + (R)(P.XCI - P.XCL) * ((R)P.XTT - t))) / (rho * ((R)P.XLMTT - (R)P.XCL * ((R)P.XTT - t)));
      }
      const R isg = S.rh_riweth + S.rh_rsweth + S.rh_rgweth;
      rweth_init = dmax(isg, dmax((R)0., S.rh_freez1 + S.rh_freez2 * isg));
      llweth = dmax((R)0., rweth_init - S.rh_riweth - S.rh_rsweth - S.rh_rgweth) <=
                   dmax((R)0., rdryh_init - S.rh_ridryh - S.rh_rsdryh - S.rh_rgdryh) &&
               rdryh_init > (R)0.;
      lldryh = cold && rdryh_init > (R)1.E-20 &&
               dmax((R)0., rweth_init - S.rh_riweth - S.rh_rsweth) >
                   dmax((R)0., rdryh_init - S.rh_ridryh - S.rh_rsdryh);
    } else {
      S.rh_freez1 = 0; S.rh_freez2 = 0;
    }
    if (llweth) {
      rcweth = S.rh_rcweth; riweth = S.rh_riweth; rsweth = S.rh_rsweth; rgweth = S.rh_rgweth;
      rrweth = (rweth_init - S.rh_riweth - S.rh_rsweth - S.rh_rgweth - S.rh_rcweth);
    }
    if (lldryh) {
      rcdryh = S.rh_rcweth; ridryh = S.rh_ridryh; rsdryh = S.rh_rsdryh; rrdryh = S.rh_rrweth;
      rgdryh = S.rh_rgdryh;
      rdryhg = rdryh_init * rweth_init / (rdryh_init + rweth_init);
    }
    if (br)
      *br |= (llweth ? 1u << ICE3_BR_WETH : 0u) | (lldryh ? 1u << ICE3_BR_DRYH : 0u) |
             (rdryhg > (R)0 ? 1u << ICE3_BR_DRYHG : 0u);
    if (rh_ok && warm) {
      if (!soft) {
        R m = rv * (R)pres / ((R)P.XEPSILO + rv);
        m = dmin(m, esw);
        m = ka * ((R)P.XTT - t) + dv * ((R)P.XLVTT + (R)(P.XCPV - P.XCL) * (t - (R)P.XTT)) * ((R)P.XESTT - m) / ((R)P.XRV * t);
        S.rhmltr = dmax((R)0., (-m * vent_h - (S.rh_rcweth + S.rh_rrweth) * (rho * (R)P.XCL * ((R)P.XTT - t))) /
                                   (rho * (R)P.XLMTT));
      }
    } else S.rhmltr = 0;
  }

  // ICE4_FAST_RI (micro/mode_ice4_fast_ri.F90:70-81); not called with OCND2 (tendencies :568-578)
  if (OCND2) {
    S.rcberi = 0;
  } else if (ssi > (R)0. && rc_ok && ri_ok && cit > 1.E-20) {
    if (!soft) {
      R lli = (R)P.LN_XLBI + (R)P.XLBEXI * rlog<R>(rho * ri / (R)cit);
      R l;
      if (lli < (R)P.LN_1E8) {
        l = rexp<R>(lli);
      } else {
        l = (R)1.E8;
        lli = (R)P.LN_1E8;
      }
      S.rcberi = (ssi / (rho * ai)) * (R)cit * ((R)P.X0DEPI / l + (R)P.X2DEPI * cj * cj / powl_<R>(lli, (R)(P.XDI + 2.0)));
    }
  } else S.rcberi = 0;

  // ICE4_FAST_RI_RS (micro/mode_ice4_fast_ri_rs.F90:95-250) with OCND2, LMODICEDEP=F: large
  // ice crystals to snow (RILARS) and deposition on non-spherical crystals in and out of the
  // ice cloud (RVDEPI); both are 0 in soft calls (:137,:168).
  double rilars = 0., rvdepi = 0.;
  if constexpr (OCND2) {
    if (!soft) {
      const double ztimesc = P.XFRMIN[14], zdicrit = P.XFRMIN[15], zwcitred = 0.1;
      const double ri_d = zv[IRI], ssi_d = (double)ssi, ai_d = (double)ai;
      if (ri_d > 0.0 && ssi_d > 0.001) rilars = dmin(ri_d, ri_d / (87.5 * sq(zdicrit) * ai_d / ssi_d));
      if ((pres * 0.5 - esi_tab) > 0. && cit > 0.) {
        const double zw2d = 1. / (x->ifr * x->icldfr + 1. - x->icldfr);
        const double ztc = dmax(-18., dmin(-1., zt - P.XTT));
        const double zhu = dmin(0.15, dmax(0., ssi_d));
        const double zcrysha = 1.1 + 3. * zhu * (1. + sin(0.64 * ztc - 1.3));
        const double zqimax = 366.5 * (zdicrit * zdicrit * zdicrit) * cit * zwcitred / rhodref;
        double zci2s = 0.;
        if (ri_d > 1.0e-12 && ztimesc > 0.)
          zci2s = ri_d * (1. - dmin(1., 0.5 * zqimax / ri_d)) * (1. - x->icldfr) * zw2d / (ztimesc * 0.5);
        const double zzwc = zcrysha * 0.878 / ai_d * rpow<double>(cit / rhodref, 0.667) *
                            rpow<double>(dmax(P.XRTMIN[4], ri_d) * zw2d, 0.333);
        if (x->ssio > 0. && x->icldfr > 0.02)
          rvdepi = rvdepi + zzwc * rpow<double>(x->ifr, 0.333) * x->ssio * x->icldfr;
        if (x->ssiu < 0. && x->icldfr < 0.98) {
          rilars = rilars + zci2s;
          rvdepi = rvdepi + zzwc * rpow<double>(zwcitred, 0.667) * x->ssiu * (1. - x->icldfr);
        }
      }
    }
  }

  // aggregation into PA / PB (micro/mode_ice4_tendencies.F90:606-717)
  const double lsl = lsfact - lvfact;
  double a;
  if (br) {  // the process branches this call took (a test hook; compiled out with br null)
    unsigned b = 0;
    if (rcrimsg > (R)0) b |= 1u << ICE3_BR_RIMING;
    if (rraccsg > (R)0) b |= 1u << ICE3_BR_RACCSG;
    if (S.rsmltg > (R)0) b |= 1u << ICE3_BR_SMLTG;
    if (ldwetg) b |= 1u << ICE3_BR_WETG;
    if (lldryg) b |= 1u << ICE3_BR_DRYG;
    if (S.rgmltr > (R)0) b |= 1u << ICE3_BR_GMLTR;
    if constexpr (HAIL) {
      if (rwetgh > (R)0) b |= 1u << ICE3_BR_WETGH;
      if (S.rhmltr > (R)0) b |= 1u << ICE3_BR_HMLTR;
    }
    *br |= b;
  }
  a = 0.;
  a = a + (double)S.rvdepg * lsfact;
  a = a + (double)S.rchoni * lsl;
  a = a + (double)S.rvdeps * lsfact;
  a = a - (double)S.rrevav * lvfact;
  a = a + (double)rcrimss * lsl;
  a = a + (double)rcrimsg * lsl;
  a = a + (double)rraccss * lsl;
  a = a + (double)rraccsg * lsl;
  a = a + ((double)S.rrcfrig - (double)S.ricfrr) * lsl;
  a = a + ((double)rcwetg + (double)rrwetg) * lsl;
  a = a + ((double)rcdryg + (double)rrdryg) * lsl;
  a = a - (double)S.rgmltr * lsl;
  if constexpr (HAIL) {
    a = a + ((double)rrweth + (double)rcweth) * lsl;
    a = a + ((double)rcdryh + (double)rrdryh) * lsl;
    a = a - (double)S.rhmltr * lsl;
  }
  a = a + (double)S.rcberi * lsl;
  if constexpr (OCND2) a = a + rvdepi * lsfact;
  pa[ITH] = a;
  a = 0.;
  a = a - (double)S.rvdepg;
  a = a - (double)S.rvdeps;
  a = a + (double)S.rrevav;
  if constexpr (OCND2) a = a - rvdepi;
  pa[IRV] = a;
  a = 0.;
  a = a - (double)S.rchoni;
  a = a - (double)S.rcautr;
  a = a - (double)S.rcaccr;
  a = a - (double)rcrimss;
  a = a - (double)rcrimsg;
  a = a - (double)S.rcmltsr;
  a = a - (double)rcwetg;
  a = a - (double)rcdryg;
  if constexpr (HAIL) {
    a = a - (double)rcweth;
    a = a - (double)rcdryh;
  }
  a = a - (double)S.rcberi;
  pa[IRC] = a;
  a = 0.;
  a = a + (double)S.rcautr;
  a = a + (double)S.rcaccr;
  a = a - (double)S.rrevav;
  a = a - (double)rraccss;
  a = a - (double)rraccsg;
  a = a + (double)S.rcmltsr;
  a = a - (double)S.rrcfrig + (double)S.ricfrr;
  a = a - (double)rrwetg;
  a = a - (double)rrdryg;
  a = a + (double)S.rgmltr;
  if constexpr (HAIL) {
    a = a - (double)rrweth;
    a = a - (double)rrdryh;
    a = a + (double)S.rhmltr;
  }
  pa[IRR] = a;
  a = 0.;
  a = a + (double)S.rchoni;
  a = a - (double)S.riaggs;
  a = a - (double)S.riauts;
  a = a - (double)S.ricfrrg - (double)S.ricfrr;
  a = a - (double)riwetg;
  a = a - (double)ridryg;
  if constexpr (HAIL) {
    a = a - (double)riweth;
    a = a - (double)ridryh;
  }
  a = a + (double)S.rcberi;
  if constexpr (OCND2) {
    a = a + rvdepi;
    a = a - rilars;
  }
  pa[IRI] = a;
  a = 0.;
  a = a + (double)S.rvdeps;
  if constexpr (OCND2) a = a + (double)S.rvdepg * zdepg_s;
  a = a + (double)S.riaggs;
  a = a + (double)S.riauts;
  a = a + (double)rcrimss;
  a = a - (double)rsrimcg;
  a = a + (double)rraccss;
  a = a - (double)rsaccrg;
  a = a - (double)S.rsmltg;
  a = a - (double)rswetg;
  a = a - (double)rsdryg;
  if constexpr (HAIL) {
    a = a - (double)rsweth;
    a = a - (double)rsdryh;
  }
  if constexpr (OCND2) a = a + rilars;
  pa[IRS] = a;
  a = 0.;
  a = a + (double)S.rvdepg * znodepg_s;
  a = a + (double)rcrimsg + (double)rsrimcg;
  a = a + (double)rraccsg + (double)rsaccrg;
  a = a + (double)S.rsmltg;
  a = a + (double)S.ricfrrg + (double)S.rrcfrig;
  a = a + (double)rcwetg + (double)riwetg + (double)rswetg + (double)rrwetg;
  if constexpr (HAIL) {
    a = a - (double)rwetgh;
    pb[IRG] = pb[IRG] - (double)rwetgh_mr;
  }
  a = a + (double)rcdryg + (double)ridryg + (double)rsdryg + (double)rrdryg;
  a = a - (double)S.rgmltr;
  if constexpr (HAIL) {
    a = a - (double)rgweth;
    a = a - (double)rgdryh + (double)rdryhg;
  }
  pa[IRG] = a;
  if constexpr (HAIL) {
    a = 0.;
    a = a + (double)rwetgh;
    pb[IRH] = pb[IRH] + (double)rwetgh_mr;
    a = a + (double)rcweth + (double)riweth + (double)rsweth + (double)rgweth + (double)rrweth;
    a = a + (double)rcdryh + (double)ridryh + (double)rsdryh + (double)rrdryh + (double)rgdryh -
        (double)rdryhg;
    a = a - (double)S.rhmltr;
    pa[IRH] = a;
  }
}

// Default sub-step cap per point (the reference's loop has none); a context can set another.
constexpr int DEFAULT_MAX_CALLS = 4096;

// Fixed inputs of ICE4_STEPPING for one point; NV = KRR + 1 variables (7, or 8 with hail).
template <int NV> struct StepPointN {
  double exn, rhodref, pres, cf, sigma_rc;
  double extpk[NV];  // external tendencies: sources minus t-state / dt
};
using StepPoint = StepPointN<7>;

// Set-up of ICE4_STEPPING (micro/mode_ice4_stepping.F90:199-240) for one point: the
// external tendencies from the sources s and the t-state vart.
template <int NV>
__device__ __forceinline__ void step_point(double dt, double exn, double rhodref, double pres,
                                           double cf, double sigs, const double *s,
                                           const double *vart, StepPointN<NV> &p) {
  const double zinv = 1. / dt;
  p.exn = exn; p.rhodref = rhodref; p.pres = pres; p.cf = cf; p.sigma_rc = sigs * 2.;
#pragma unroll
  for (int jv = 0; jv < NV; ++jv) p.extpk[jv] = s[jv] - vart[jv] * zinv;
}

// One pass of the inner loop of ICE4_STEPPING (:256-516): an ICE4_TENDENCIES call (hard
// when SOFT is false, else reusing the rates kept in S), the time-step limitation and the
// state update.  The point's state between calls: vart (th, rv, rc, rr, ri, rs, rg), cit,
// time (part of the step integrated so far), S (rates of the hard call) and calls.
// Returns 1 while the point's time step goes on, 0 when it is complete (vart then holds
// the end state, :520-537), -1 when the point would make more than max_calls tendency
// calls: vart then holds the state after the last completed call with the external
// tendencies integrated so far taken out again (as the complete path does at :520-537 for
// the whole step), so that the total tendencies count them once, through the sources.
// pc is read by hard calls only.  With HAIL, vart, p and S carry hail (8 variables).
template <class R, bool SOFT, bool OCND2 = false, bool HAIL = false, class TB, class SS, int NV>
__device__ __forceinline__ int step_call(double dt, const StepPointN<NV> &p, const PointConst<R> &pc,
                                         double *vart, double &cit, double &time_,
                                         SS &S, int &calls, int max_calls,
                                         const TB *__restrict__ tab,
                                         const MicroOcnd2 *x = nullptr,
                                         const double *__restrict__ htab = nullptr,
                                         unsigned *br = nullptr) {
  static_assert(NV == (HAIL ? 8 : 7), "variables of the time splitting");
  if (++calls > max_calls) {
#pragma unroll
    for (int jv = 0; jv < NV; ++jv) vart[jv] = vart[jv] - p.extpk[jv] * time_;
    return -1;
  }
  double za[NV], zb[NV];
  const double exn = p.exn;
  double sum2 = vart[IRI] + vart[IRS] + vart[IRG];
  if constexpr (HAIL) sum2 = sum2 + vart[IRH];
  const double zdevide =
      (P.XCPD + P.XCPV * vart[IRV] + P.XCL * (vart[IRC] + vart[IRR]) + P.XCI * sum2) * exn;
  const double zzt = vart[ITH] * exn;
  const double lsfact = (P.XLSTT + (P.XCPV - P.XCI) * (zzt - P.XTT)) / zdevide;
  const double lvfact = (P.XLVTT + (P.XCPV - P.XCL) * (zzt - P.XTT)) / zdevide;
  tendencies<R, OCND2, HAIL>(SOFT, exn, p.rhodref, lvfact, lsfact, p.pres, p.cf, p.sigma_rc, cit,
                             zzt, vart, S, pc, tab, za, zb, x, htab, br);
#pragma unroll
  for (int jv = 0; jv < NV; ++jv) za[jv] = za[jv] + p.extpk[jv];
  const double time = time_;
  double maxtime = dt - time;
  const double zx = P.XTT / exn;
  if ((vart[ITH] - zx) * (vart[ITH] + zb[ITH] - zx) < 0.) maxtime = 0.;
  if (fabs(za[ITH]) > 1.E-20) {
    const double tth = (zx - zb[ITH] - vart[ITH]) / za[ITH];
    if (tth > 0.) maxtime = dmin(maxtime, tth);
  }
#pragma unroll
  for (int jv = 1; jv < NV; ++jv)
    if (za[jv] < -1.E-20 && vart[jv] > P.XRTMIN[jv]) {
      maxtime = dmin(maxtime, -(zb[jv] + vart[jv]) / za[jv]);
      maxtime = dmax(maxtime, P.XMNH_TINY);
    }
  const bool done = time + maxtime >= dt;
#pragma unroll
  for (int jv = 0; jv < NV; ++jv) vart[jv] = vart[jv] + za[jv] * maxtime + zb[jv];
  if (vart[IRI] <= 0.) cit = 0.;
  time_ = time + maxtime;
  if (!done) return 1;
#pragma unroll
  for (int jv = 0; jv < NV; ++jv) vart[jv] = vart[jv] - p.extpk[jv] * dt;
  return 0;
}

// ICE4_STEPPING (micro/mode_ice4_stepping.F90:199-537) for one microphysics point, all
// calls in a row: vart (t-state in, end state out), cit and rrevav are updated, s holds
// the sources.  Each point's time step is one outer iteration: a hard call, then soft
// calls until the step is complete.  Returns the number of tendency calls, or -calls if
// the cap was hit.  With HAIL, vart and s have 8 entries (hail last).
template <class R, bool OCND2 = false, bool HAIL = false, class TB>
__device__ __forceinline__ int stepping_point(double dt, double exn, double rhodref, double pres,
                                              double cf, double sigs, double &cit, double *vart,
                                              const double *s, double &rrevav, int max_calls,
                                              const TB *__restrict__ tab,
                                              const MicroOcnd2 *x = nullptr,
                                              const double *__restrict__ htab = nullptr,
                                              unsigned *br = nullptr) {
  StepPointN<HAIL ? 8 : 7> p;
  step_point(dt, exn, rhodref, pres, cf, sigs, s, vart, p);
  typename StepStateFor<R, HAIL>::type S;
  double time = 0.;
  int calls = 0;
  const PointConst<R> pc = point_const<R>(rhodref);
  int st = step_call<R, false, OCND2, HAIL>(dt, p, pc, vart, cit, time, S, calls, max_calls, tab, x,
                                            htab, br);
  while (st > 0)
    st = step_call<R, true, OCND2, HAIL>(dt, p, pc, vart, cit, time, S, calls, max_calls, tab, x,
                                         htab, br);
  if (st == 0) rrevav = (double)S.rrevav;
  return st < 0 ? -calls : calls;
}

// Fall speed of one species at one level, OTHER_SPECIES / PRISTINE_ICE
// (micro/mode_ice4_sedimentation_stat.F90:295-431); jrr = 3 (r), 4 (i), 5 (s), 6 (g),
// 7 (h, :184-188).  lnrho = ln(rhodref) of the level; the powers are evaluated as exp(c*ln x).
template <class R>
__device__ __forceinline__ R fall_speed(int jrr, double rx, R lnrho) {
  if (jrr == 4) {
    if (!(rx > dmax(P.XRTMIN[4], 1.0E-7))) return (R)0;
    const R lnrx = rlog<R>((R)rx);
    return (R)P.XFSEDI * rexp<R>(-(R)P.XCEXVT * lnrho) *
           rexp<R>((R)P.XEXCSEDI *
                   rlog<R>(dmax((R)0.05E6, (R)-0.15319E6 - (R)0.021454E6 * (lnrho + lnrx))));
  }
  if (!(rx > P.XRTMIN[jrr])) return (R)0;
  const double fsed = jrr == 3 ? P.XFSEDR : (jrr == 5 ? P.XFSEDS : (jrr == 6 ? P.XFSEDG : P.XFSEDH));
  const double exsed = jrr == 3 ? P.XEXSEDR : (jrr == 5 ? P.XEXSEDS : (jrr == 6 ? P.XEXSEDG : P.XEXSEDH));
  return (R)fsed * rexp<R>((R)(exsed - 1) * rlog<R>((R)rx) + (R)(exsed - P.XCEXVT - 1) * lnrho);
}

// One species at one level of ICE4_SEDIMENTATION_STAT (:143-229) for one column:
// flux holds the flux leaving the level above on entry and this level's on exit
// (double); rs is the species' source, updated by the flux divergence.
// jrr = 3 (r), 4 (i), 5 (s), 6 (g), 7 (h); the species are independent, so the
// reference's species-outer loop order and this level-outer order agree.
template <class R>
__device__ __forceinline__ void sediment_species(int jrr, double dt, double rho, R lnrho, double dz,
                                                 double &rs, double &flux) {
  const double zinv = 1. / dt;
  const double tsorhodz = dt / (rho * dz);
  const double rx = rs * dt;
  const double qp = flux * tsorhodz;
  const double thr = P.XRTMIN[jrr];
  double cur;
  if (rx > thr || qp > thr) {
    const double w1 = (double)fall_speed<R>(jrr, rx, lnrho);
    const double w2 = (double)fall_speed<R>(jrr, qp, lnrho);
    const double f1 = dmin(rho * dz * rx * zinv, w1 * rho * rx);
    cur = w2 != 0. ? f1 + dmax(0., 1. - dz / (dt * w2)) * flux : f1;
  } else {
    cur = dmin(rho * dz * rx * zinv, 0. * rho * rx);
  }
  rs = rs + tsorhodz * (flux - cur) * zinv;
  flux = cur;
}

}  // namespace ice3
