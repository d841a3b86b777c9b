// ice3_params.cpp -- host-side constant and table builder.
//
// Restates INI_CST (aux/mode_ini_cst.F90:77-193) and the parts of INI_RAIN_ICE
// (micro/mode_ini_rain_ice.F90) that the ice_adjust / rain_ice path reads, for
// the pinned AROME ICE3 configuration: HPROGRAM='AROME', HCLOUD='ICE3' (KRR=6) and
// HCLOUD='ICE4' (KRR=7: XRTMIN(7) and the hail constants),
// LRED=F, LSNOW_T=F (so LNEWCOEFF=F), CPRISTINE_ICE='PLAT', LCRIAUTI=F,
// XFRACM90=0.1, XFRMIN(16)=0.  The special functions are restated from the same
// published methods the reference uses: GAMMA (aux/gamma.F90, 6-term Lanczos of
// Numerical Recipes with the reference's truncated pi), GAMMA_INC
// (aux/gamma_inc.F90, Numerical Recipes gammp series / continued fraction,
// EPS=3e-7, ITMAX=100) and MOMG (micro/momg.F90).  The collection kernels are
// the reference's tabulated data (xker_tables.inc).
#include "ice3_params.h"

#include <cmath>
#include <cstring>

namespace {

#include "xker_tables.inc"

// GAMMA_X0D, aux/gamma.F90:12-111
double gamma_ref(double px) {
  static const double coef[6] = {76.18009172947146,     -86.50532032941677,
                                 24.01409824083091,     -1.231739572450155,
                                 0.1208650973866179E-2, -0.5395239384953E-5};
  const double zstp = 2.5066282746310005;
  const double zpi = 3.141592654;
  double zx = px < 0. ? 1. - px : px;
  double zy = zx;
  double ztmp = zx + 5.5;
  ztmp = (zx + 0.5) * std::log(ztmp) - ztmp;
  double zser = 1.000000000190015;
  for (int j = 0; j < 6; ++j) {
    zy = zy + 1.0;
    zser = zser + coef[j] / zy;
  }
  if (px < 0.) return zpi / std::sin(zpi * px) / std::exp(ztmp + std::log(zstp * zser / zx));
  return std::exp(ztmp + std::log(zstp * zser / zx));
}

// MOMG_X0D, micro/momg.F90:57-99
double momg(double alpha, double nu, double p) { return gamma_ref(nu + p / alpha) / gamma_ref(nu); }

// GAMMA_INC, aux/gamma_inc.F90:5-137 (regularised lower incomplete gamma P(a,x))
double gamma_inc(double pa, double px) {
  const int itmax = 100;
  const double zeps = 3.E-7, zfpmin = 1.E-30, tiny = 2.2250738585072014e-308;
  if (px < pa + 1.0) {
    double zap = pa, zsum = 1.0 / pa, zdel = zsum;
    for (int jn = 1; jn <= itmax; ++jn) {
      zap = zap + 1.0;
      zdel = zdel * px / zap;
      zsum = zsum + zdel;
      if (std::fabs(zdel) < std::fabs(zsum) * zeps) break;
    }
    return zsum * std::exp(-px + pa * std::log(px) - std::log(gamma_ref(pa)));
  }
  double zb = px + 1.0 - pa, zc = 1.0 / tiny, zd = 1.0 / zb, zh = zd;
  for (int jn = 1; jn <= itmax; ++jn) {
    double zan = -(double)jn * ((double)jn - pa);
    zb = zb + 2.0;
    zd = zan * zd + zb;
    if (std::fabs(zd) < tiny) zd = zfpmin;
    zc = zb + zan / zc;
    if (std::fabs(zc) < tiny) zc = zfpmin;
    zd = 1.0 / zd;
    double zdel = zd * zc;
    zh = zh * zdel;
    if (std::fabs(zdel - 1.0) < zeps) break;
  }
  return 1.0 - zh * std::exp(-px + pa * std::log(px) - std::log(gamma_ref(pa)));
}

// x**n for an integer exponent, by binary powering (the lowering of Fortran
// REAL**INTEGER used for the GAMINC abscissae, micro/mode_ini_rain_ice.F90:786)
double powi(double a, int b) {
  const bool recip = b < 0;
  double r = 1;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1 / r : r;
}

}  // namespace

void ice3_build_params(ice3_params *P, double *tables) {
  std::memset(P, 0, sizeof(*P));
  // ---- INI_CST (aux/mode_ini_cst.F90:77-157,179)
  const double XBOLTZ = 1.380658E-23, XAVOGADRO = 6.0221367E+23;
  const double XMD = 28.9644E-3, XMV = 18.0153E-3;
  P->XPI = 2. * std::asin(1.);
  P->XG = 9.80665;
  P->XP00 = 1.E5;
  P->XRD = XAVOGADRO * XBOLTZ / XMD;
  P->XRV = XAVOGADRO * XBOLTZ / XMV;
  P->XEPSILO = XMV / XMD;
  P->XCPD = 7. * P->XRD / 2.;
  P->XCPV = 4. * P->XRV;
  P->XRHOLW = 1000.;
  P->XCL = 4.218E+3;
  P->XCI = 2.106E+3;
  P->XTT = 273.16;
  P->XLVTT = 2.5008E+6;
  P->XLSTT = 2.8345E+6;
  P->XLMTT = P->XLSTT - P->XLVTT;
  P->XESTT = 611.14;
  P->XGAMW = (P->XCL - P->XCPV) / P->XRV;
  P->XBETAW = (P->XLVTT / P->XRV) + (P->XGAMW * P->XTT);
  P->XALPW = std::log(P->XESTT) + (P->XBETAW / P->XTT) + (P->XGAMW * std::log(P->XTT));
  P->XGAMI = (P->XCI - P->XCPV) / P->XRV;
  P->XBETAI = (P->XLSTT / P->XRV) + (P->XGAMI * P->XTT);
  P->XALPI = std::log(P->XESTT) + (P->XBETAI / P->XTT) + (P->XGAMI * std::log(P->XTT));
  P->XMNH_TINY = 1.0e-80;
  // NEB_t pinned values (micro/modd_nebn.F90:178-179 semantics)
  P->XTMINMIX = 253.16;
  P->XTMAXMIX = 273.16;

  // ---- INI_RAIN_ICE: descriptors (micro/mode_ini_rain_ice.F90:252-463)
  const double XPI = P->XPI;
  P->XRTMIN[1] = 1.0E-20; P->XRTMIN[2] = 1.0E-20; P->XRTMIN[3] = 1.0E-20;
  P->XRTMIN[4] = 1.0E-20; P->XRTMIN[5] = 1.0E-15; P->XRTMIN[6] = 1.0E-15;
  P->XRTMIN[7] = 1.0E-15;  // HCLOUD='ICE4' (:258)
  const double XAR = (XPI / 6.0) * P->XRHOLW, XBR = 3.0, XCR = 842., XDR = 0.8, XCCR = 8.E6;
  const double ZXR = -1., XF0R = 1.00, XF1R = 0.26, XC1R = 1. / 2.;
  const double XAI = 0.82, XBI = 2.5, XC_I = 800., XDI = 1.0, XC1I = 1. / XPI, XF0I = 1.00,
               XF2I = 0.14;
  const double XAS = 0.02, XBS = 1.9, XCS = 5.1, XDS = 0.27, XCCS = 5.0, XCXS = 1.0, XF0S = 0.86,
               XF1S = 0.28, XC1S = 1. / XPI;
  const double XAG = 19.6, XBG = 2.8, XCG = 124., XDG = 0.66, XCCG = 5.E5, XCXG = -0.5,
               XF0G = 0.86, XF1G = 0.28, XC1G = 1. / 2.;
  const double XAH = 470., XBH = 3.0, XCH = 207., XDH = 0.64, XCCH = 4.E4, XCXH = -1.0,
               XF0H = 0.86, XF1H = 0.28, XC1H = 1. / 2.;  // hail (:375-392)
  const double XALPHAR = 1.0, XNUR = 1.0, XALPHAI = 3.0, XNUI = 3.0, XALPHAS = 1.0, XNUS = 1.0,
               XALPHAG = 1.0, XNUG = 1.0, XALPHAH = 1.0, XNUH = 8.0;
  P->XDI = XDI; P->XBS = XBS; P->XCXS = XCXS; P->XCXG = XCXG; P->XDG = XDG;
  P->XLBEXR = 1.0 / (-1.0 - XBR);
  P->XLBR = std::pow(XAR * XCCR * momg(XALPHAR, XNUR, XBR), -P->XLBEXR);
  P->XLBEXI = 1.0 / (-XBI);
  P->XLBI = std::pow(XAI * momg(XALPHAI, XNUI, XBI), -P->XLBEXI);
  P->XLBEXS = 1.0 / (XCXS - XBS);
  P->XLBS = std::pow(XAS * XCCS * momg(XALPHAS, XNUS, XBS), -P->XLBEXS);
  P->XLBEXG = 1.0 / (XCXG - XBG);
  P->XLBG = std::pow(XAG * XCCG * momg(XALPHAG, XNUG, XBG), -P->XLBEXG);
  P->XLBEXH = 1.0 / (XCXH - XBH);
  P->XLBH = std::pow(XAH * XCCH * momg(XALPHAH, XNUH, XBH), -P->XLBEXH);
  P->XCXH = XCXH; P->XDH = XDH; P->XBG = XBG;
  const double ZCONC_MAX = 1.E6;
  P->XLBDAS_MAX = 100000.0;
  if (XCCS > 0. && XCXS > 0.) P->XLBDAS_MAX = std::pow(ZCONC_MAX / XCCS, 1. / XCXS);
  P->XCONC_LAND = 3E8;
  // PARAM_ICE XFRMIN_NAM of the pinned AROME configuration, copied by INI_RAIN_ICE
  // (micro/mode_ini_rain_ice.F90:203): used by the OCND2 branch of RAIN_ICE
  for (int j = 0; j <= 40; ++j) P->XFRMIN[j] = 0.;
  P->XFRMIN[7] = 1.; P->XFRMIN[8] = 1.; P->XFRMIN[9] = 1.; P->XFRMIN[10] = 10.; P->XFRMIN[11] = 1.;
  P->XFRMIN[13] = 1.0E-15; P->XFRMIN[14] = 120.; P->XFRMIN[15] = 1.0E-4; P->XFRMIN[21] = 1.;
  P->XFRMIN[22] = 1.; P->XFRMIN[23] = 0.5; P->XFRMIN[24] = 1.5; P->XFRMIN[25] = 30.;
  P->XFRMIN[30] = 1.; P->XFRMIN[39] = 0.25; P->XFRMIN[40] = 0.15;
  P->XCEXVT = 0.4;
  const double XCEXVT = P->XCEXVT;
  double ZE = (50. / 100.) * std::exp(P->XALPW - P->XBETAW / 293.15 - P->XGAMW * std::log(293.15));
  double ZRV = (P->XRD / P->XRV) * ZE / (101325. - ZE);
  const double ZRHO00 = 101325. * (1. + ZRV) / (P->XRD + ZRV * P->XRV) / 293.15;
  const double ZRHO00_SNOW = ZRHO00;
  // sedimentation (:505-549)
  P->XEXSEDR = (XBR + XDR + 1.0) / (XBR + 1.0);
  P->XFSEDR = XCR * XAR * XCCR * momg(XALPHAR, XNUR, XBR + XDR) *
              std::pow(XAR * XCCR * momg(XALPHAR, XNUR, XBR), -P->XEXSEDR) * std::pow(ZRHO00, XCEXVT);
  const double XEXRSEDI = 1.9324;
  P->XFSEDI = 3.89745E11 * momg(XALPHAI, XNUI, 3.285) *
              std::pow(momg(XALPHAI, XNUI, 1.7), -XEXRSEDI) * std::pow(ZRHO00, XCEXVT);
  P->XEXCSEDI = -0.9324 * 3.0;
  P->XEXSEDS = (XBS + XDS - XCXS) / (XBS - XCXS);
  P->XFSEDS = XCS * XAS * XCCS * momg(XALPHAS, XNUS, XBS + XDS) *
              std::pow(XAS * XCCS * momg(XALPHAS, XNUS, XBS), -P->XEXSEDS) *
              std::pow(ZRHO00_SNOW, XCEXVT);
  P->XEXSEDG = (XBG + XDG - XCXG) / (XBG - XCXG);
  P->XFSEDG = XCG * XAG * XCCG * momg(XALPHAG, XNUG, XBG + XDG) *
              std::pow(XAG * XCCG * momg(XALPHAG, XNUG, XBG), -P->XEXSEDG) * std::pow(ZRHO00, XCEXVT);
  P->XEXSEDH = (XBH + XDH - XCXH) / (XBH - XCXH);
  P->XFSEDH = XCH * XAH * XCCH * momg(XALPHAH, XNUH, XBH + XDH) *
              std::pow(XAH * XCCH * momg(XALPHAH, XNUH, XBH), -P->XEXSEDH) * std::pow(ZRHO00, XCEXVT);
  // nucleation (:555-591)
  P->XNU10 = 50.; P->XALPHA1 = 4.5; P->XBETA1 = 0.6;
  P->XNU20 = 1000.; P->XALPHA2 = 12.96; P->XBETA2 = 0.639;
  P->XMNU0 = 6.88E-13;
  P->XALPHA3 = -3.075; P->XBETA3 = 81.00356;
  P->XHON = (XPI / 6.) * ((2.0 * 3.0 * 4.0 * 5.0 * 6.0) / (2.0 * 3.0)) * std::pow(1.1E5, -3.0);
  // deposition (:609-634)
  P->XSCFAC = std::pow(0.63, 1. / 3.) * std::sqrt(std::pow(ZRHO00, XCEXVT));
  P->X0DEPI = (4.0 * XPI) * XC1I * XF0I * momg(XALPHAI, XNUI, 1.);
  P->X2DEPI = (4.0 * XPI) * XC1I * XF2I * XC_I * momg(XALPHAI, XNUI, XDI + 2.0);
  P->X0DEPS = (4.0 * XPI) * XCCS * XC1S * XF0S * momg(XALPHAS, XNUS, 1.);
  P->X1DEPS = (4.0 * XPI) * XCCS * XC1S * XF1S * std::sqrt(XCS) * momg(XALPHAS, XNUS, 0.5 * XDS + 1.5);
  P->XEX0DEPS = XCXS - 1.0;
  P->XEX1DEPS = XCXS - 0.5 * (XDS + 3.0);
  P->X0DEPG = (4.0 * XPI) * XCCG * XC1G * XF0G * momg(XALPHAG, XNUG, 1.);
  P->X1DEPG = (4.0 * XPI) * XCCG * XC1G * XF1G * std::sqrt(XCG) * momg(XALPHAG, XNUG, 0.5 * XDG + 1.5);
  P->XEX0DEPG = XCXG - 1.0;
  P->XEX1DEPG = XCXG - 0.5 * (XDG + 3.0);
  P->X0DEPH = (4.0 * XPI) * XCCH * XC1H * XF0H * momg(XALPHAH, XNUH, 1.);
  P->X1DEPH = (4.0 * XPI) * XCCH * XC1H * XF1H * std::sqrt(XCH) * momg(XALPHAH, XNUH, 0.5 * XDH + 1.5);
  P->XEX0DEPH = XCXH - 1.0;
  P->XEX1DEPH = XCXH - 0.5 * (XDH + 3.0);
  // aggregation / autoconversion (:648-717)
  P->XTIMAUTI = 1.E-3;
  P->XTEXAUTI = 0.015;
  P->XCRIAUTI = 0.2E-4;
  P->XACRIAUTI = 0.06;
  P->XBCRIAUTI = -3.5;
  P->XT0CRIAUTI = (std::log10(P->XCRIAUTI) - P->XBCRIAUTI) / 0.06;
  const double XCOLIS = 0.25;
  P->XCOLEXIS = 0.05;
  P->XFIAGGS = (XPI / 4.0) * XCOLIS * XCCS * XCS * std::pow(ZRHO00_SNOW, XCEXVT) *
               momg(XALPHAS, XNUS, XDS + 2.0);
  P->XEXIAGGS = XCXS - XDS - 2.0;
  P->XTIMAUTC = 1.E-3;
  P->XCRIAUTC = 0.5E-3;
  // accretion, evaporation (:736-760)
  P->XFCACCR = (XPI / 4.0) * XCCR * XCR * std::pow(ZRHO00, XCEXVT) * momg(XALPHAR, XNUR, XDR + 2.0);
  P->XEXCACCR = -XDR - 3.0;
  P->X0EVAR = (4.0 * XPI) * XCCR * XC1R * XF0R * momg(XALPHAR, XNUR, 1.);
  P->X1EVAR = (4.0 * XPI) * XCCR * XC1R * XF1R * std::sqrt(XCR) * momg(XALPHAR, XNUR, 0.5 * XDR + 1.5);
  P->XEX0EVAR = -2.0;
  P->XEX1EVAR = -1.0 - 0.5 * (XDR + 3.0);
  // riming (:762-795)
  const double XDCSLIM = 0.007, XCOLCS = 1.0;
  P->XEXCRIMSS = XCXS - XDS - 2.0;
  P->XCRIMSS = (XPI / 4.0) * XCOLCS * XCCS * XCS * std::pow(ZRHO00_SNOW, XCEXVT) *
               momg(XALPHAS, XNUS, XDS + 2.0);
  P->XEXCRIMSG = P->XEXCRIMSS;
  P->XCRIMSG = P->XCRIMSS;
  P->XSRIMCG = XCCS * XAS * momg(XALPHAS, XNUS, XBS);
  P->XEXSRIMCG = XCXS - XBS;
  P->XSRIMCG2 = XCCS * XAG * momg(XALPHAS, XNUS, XBG);
  P->XSRIMCG3 = 0.1;  // XFRACM90
  P->XEXSRIMCG2 = XCXS - XBG;
  P->NGAMINC = 80;
  const double BMIN = 1.0E-1, BMAX = 1.0E7;
  double zrate = std::exp(std::log(BMAX / BMIN) / (double)(P->NGAMINC - 1));
  for (int j = 1; j <= P->NGAMINC; ++j) {
    double zbound = BMIN * powi(zrate, j - 1);
    tables[ICE3_T_RIM1 + j - 1] = gamma_inc(XNUS + (2.0 + XDS) / XALPHAS, zbound);
    tables[ICE3_T_RIM2 + j - 1] = gamma_inc(XNUS + XBS / XALPHAS, zbound);
    tables[ICE3_T_RIM4 + j - 1] = gamma_inc(XNUS + XBG / XALPHAS, zbound);
  }
  P->XRIMINTP1 = XALPHAS / std::log(zrate);
  P->XRIMINTP2 = 1.0 + P->XRIMINTP1 * std::log(XDCSLIM / std::pow(BMIN, 1.0 / XALPHAS));
  // rain-snow accretion (:805-996)
  P->XFRACCSS = ((XPI * XPI) / 24.0) * XCCS * XCCR * P->XRHOLW * std::pow(ZRHO00_SNOW, XCEXVT);
  P->XLBRACCS1 = momg(XALPHAS, XNUS, 2.) * momg(XALPHAR, XNUR, 3.);
  P->XLBRACCS2 = 2. * momg(XALPHAS, XNUS, 1.) * momg(XALPHAR, XNUR, 4.);
  P->XLBRACCS3 = momg(XALPHAR, XNUR, 5.);
  P->XFSACCRG = (XPI / 4.0) * XAS * XCCS * XCCR * std::pow(ZRHO00_SNOW, XCEXVT);
  P->XLBSACCR1 = momg(XALPHAR, XNUR, 2.) * momg(XALPHAS, XNUS, XBS);
  P->XLBSACCR2 = 2. * momg(XALPHAR, XNUR, 1.) * momg(XALPHAS, XNUS, XBS + 1.);
  P->XLBSACCR3 = momg(XALPHAS, XNUS, XBS + 2.);
  P->NACCLBDAS = 40;
  zrate = std::log(5.0E5 / 5.0E1) / (double)(P->NACCLBDAS - 1);
  P->XACCINTP1S = 1.0 / zrate;
  P->XACCINTP2S = 1.0 - std::log(5.0E1) / zrate;
  P->NACCLBDAR = 40;
  zrate = std::log(1.0E7 / 1.0E3) / (double)(P->NACCLBDAR - 1);
  P->XACCINTP1R = 1.0 / zrate;
  P->XACCINTP2R = 1.0 - std::log(1.0E3) / zrate;
  P->XFSCVMG = 2.0;
  // rain contact freezing (:1000-1030)
  const double XCOLIR = 1.0;
  P->XEXRCFRI = -XDR - 5.0 + ZXR;
  P->XRCFRI = ((XPI * XPI) / 24.0) * XCCR * P->XRHOLW * XCOLIR * XCR * std::pow(ZRHO00, XCEXVT) *
              momg(XALPHAR, XNUR, XDR + 5.0);
  P->XEXICFRR = -XDR - 2.0 + ZXR;
  P->XICFRR = (XPI / 4.0) * XCOLIR * XCR * std::pow(ZRHO00, XCEXVT) * XCCR * momg(XALPHAR, XNUR, XDR + 2.0);
  // graupel dry growth (:1035-1160)
  P->XFCDRYG = (XPI / 4.0) * XCCG * XCG * std::pow(ZRHO00, XCEXVT) * momg(XALPHAG, XNUG, XDG + 2.0);
  P->XCOLIG = 0.01;
  P->XCOLEXIG = 0.1;
  P->XFIDRYG = (XPI / 4.0) * P->XCOLIG * XCCG * XCG * std::pow(ZRHO00, XCEXVT) *
               momg(XALPHAG, XNUG, XDG + 2.0);
  P->XCOLSG = 0.01;
  P->XCOLEXSG = 0.1;
  P->XFSDRYG = (XPI / 4.0) * P->XCOLSG * XCCG * XCCS * XAS * std::pow(ZRHO00_SNOW, XCEXVT);
  P->XLBSDRYG1 = momg(XALPHAG, XNUG, 2.) * momg(XALPHAS, XNUS, XBS);
  P->XLBSDRYG2 = 2. * momg(XALPHAG, XNUG, 1.) * momg(XALPHAS, XNUS, XBS + 1.);
  P->XLBSDRYG3 = momg(XALPHAS, XNUS, XBS + 2.);
  P->XFRDRYG = ((XPI * XPI) / 24.0) * XCCG * XCCR * P->XRHOLW * std::pow(ZRHO00, XCEXVT);
  P->XLBRDRYG1 = momg(XALPHAG, XNUG, 2.) * momg(XALPHAR, XNUR, 3.);
  P->XLBRDRYG2 = 2. * momg(XALPHAG, XNUG, 1.) * momg(XALPHAR, XNUR, 4.);
  P->XLBRDRYG3 = momg(XALPHAR, XNUR, 5.);
  P->NDRYLBDAR = 40;
  zrate = std::log(1.0E7 / 1.0E3) / (double)(P->NDRYLBDAR - 1);
  P->XDRYINTP1R = 1.0 / zrate;
  P->XDRYINTP2R = 1.0 - std::log(1.0E3) / zrate;
  P->NDRYLBDAS = 80;
  zrate = std::log(2.5E9 / 2.5E1) / (double)(P->NDRYLBDAS - 1);
  P->XDRYINTP1S = 1.0 / zrate;
  P->XDRYINTP2S = 1.0 - std::log(2.5E1) / zrate;
  P->NDRYLBDAG = 40;
  zrate = std::log(1.0E7 / 1.0E3) / (double)(P->NDRYLBDAG - 1);
  P->XDRYINTP1G = 1.0 / zrate;
  P->XDRYINTP2G = 1.0 - std::log(1.0E3) / zrate;
  // hail wet and dry growth (:1225-1296)
  P->XCOLIH = 0.01;
  P->XCOLEXIH = 0.1;
  P->XFWETH = (XPI / 4.0) * XCCH * XCH * std::pow(ZRHO00, XCEXVT) * momg(XALPHAH, XNUH, XDH + 2.0);
  P->XCOLSH = 0.01;
  P->XCOLEXSH = 0.1;
  P->XFSWETH = (XPI / 4.0) * XCCH * XCCS * XAS * std::pow(ZRHO00_SNOW, XCEXVT);
  P->XLBSWETH1 = momg(XALPHAH, XNUH, 2.) * momg(XALPHAS, XNUS, XBS);
  P->XLBSWETH2 = 2. * momg(XALPHAH, XNUH, 1.) * momg(XALPHAS, XNUS, XBS + 1.);
  P->XLBSWETH3 = momg(XALPHAS, XNUS, XBS + 2.);
  P->XCOLGH = 0.01;
  P->XCOLEXGH = 0.1;
  P->XFGWETH = (XPI / 4.0) * XCCH * XCCG * XAG * std::pow(ZRHO00, XCEXVT);
  P->XLBGWETH1 = momg(XALPHAH, XNUH, 2.) * momg(XALPHAG, XNUG, XBG);
  P->XLBGWETH2 = 2. * momg(XALPHAH, XNUH, 1.) * momg(XALPHAG, XNUG, XBG + 1.);
  P->XLBGWETH3 = momg(XALPHAG, XNUG, XBG + 2.);
  P->XFRWETH = (XPI / 4.0) * XCCH * XCCR * XAR * std::pow(ZRHO00, XCEXVT);
  P->XLBRWETH1 = momg(XALPHAH, XNUH, 2.) * momg(XALPHAR, XNUR, XBR);
  P->XLBRWETH2 = 2. * momg(XALPHAH, XNUH, 1.) * momg(XALPHAR, XNUR, XBR + 1.);
  P->XLBRWETH3 = momg(XALPHAR, XNUR, XBR + 2.);
  P->NWETLBDAS = 80;
  zrate = std::log(2.5E9 / 2.5E1) / (double)(P->NWETLBDAS - 1);
  P->XWETINTP1S = 1.0 / zrate;
  P->XWETINTP2S = 1.0 - std::log(2.5E1) / zrate;
  P->NWETLBDAG = 40;
  zrate = std::log(1.0E7 / 1.0E3) / (double)(P->NWETLBDAG - 1);
  P->XWETINTP1G = 1.0 / zrate;
  P->XWETINTP2G = 1.0 - std::log(1.0E3) / zrate;
  P->NWETLBDAR = 40;
  zrate = std::log(1.0E7 / 1.0E3) / (double)(P->NWETLBDAR - 1);
  P->XWETINTP1R = 1.0 / zrate;
  P->XWETINTP2R = 1.0 - std::log(1.0E3) / zrate;
  P->NWETLBDAH = 40;
  zrate = std::log(1.0E7 / 1.0E3) / (double)(P->NWETLBDAH - 1);
  P->XWETINTP1H = 1.0 / zrate;
  P->XWETINTP2H = 1.0 - std::log(1.0E3) / zrate;
  // tabulated collection kernels
  std::memcpy(tables + ICE3_T_RACCSS, XKER_RACCSS_DATA, sizeof(XKER_RACCSS_DATA));
  std::memcpy(tables + ICE3_T_RACCS, XKER_RACCS_DATA, sizeof(XKER_RACCS_DATA));
  std::memcpy(tables + ICE3_T_SACCRG, XKER_SACCRG_DATA, sizeof(XKER_SACCRG_DATA));
  std::memcpy(tables + ICE3_T_SDRYG, XKER_SDRYG_DATA, sizeof(XKER_SDRYG_DATA));
  std::memcpy(tables + ICE3_T_RDRYG, XKER_RDRYG_DATA, sizeof(XKER_RDRYG_DATA));
  std::memcpy(tables + ICE3_T_SWETH, XKER_SWETH_DATA, sizeof(XKER_SWETH_DATA));
  std::memcpy(tables + ICE3_T_GWETH, XKER_GWETH_DATA, sizeof(XKER_GWETH_DATA));
  std::memcpy(tables + ICE3_T_RWETH, XKER_RWETH_DATA, sizeof(XKER_RWETH_DATA));
  // logarithms used by the device code's power evaluations
  P->LN_XLBR = std::log(P->XLBR);
  P->LN_XLBS = std::log(P->XLBS);
  P->LN_XLBG = std::log(P->XLBG);
  P->LN_XLBI = std::log(P->XLBI);
  P->LN_XLBDAS_MAX = std::log(P->XLBDAS_MAX);
  P->LN_XTT = std::log(P->XTT);
  P->LN_1E8 = std::log(1.E8);
  P->LN_10 = std::log(10.);
  P->LN_XLBH = std::log(P->XLBH);
}

// INI_TIWMX (micro/mode_ini_tiwmx.F90:14-49) for the tables the OCND2 branch reads, with the
// functions of micro/mode_tiwmx_fun.F90:36-106 at T = jk / XNDEGR.  Above 273.15 K the ice
// table holds the liquid-water pressure and AM3 is 0.  ESATI (the function) is the pressure
// over ice below XTT and over water from XTT up; x**2 and x**3 are products, as the
// reference's integer powers.
void ice3_build_tiwmx(const ice3_params *P, double *tab) {
  double *esw = tab + ICE3_TW_ESW * ICE3_TIWMX_N, *esi = tab + ICE3_TW_ESI * ICE3_TIWMX_N;
  double *a2 = tab + ICE3_TW_A2 * ICE3_TIWMX_N, *bb3 = tab + ICE3_TW_BB3 * ICE3_TIWMX_N;
  double *am3 = tab + ICE3_TW_AM3 * ICE3_TIWMX_N, *redin = tab + ICE3_TW_REDIN * ICE3_TIWMX_N;
  for (int jk = ICE3_TIWMX_START; jk <= ICE3_TIWMX_STOP; ++jk) {
    const int e = jk - ICE3_TIWMX_START;
    const double t = (double)jk / 100.;
    const double w = std::exp(P->XALPW - P->XBETAW / t - P->XGAMW * std::log(t));
    const double i = std::exp(P->XALPI - P->XBETAI / t - P->XGAMI * std::log(t));
    const double sgn = t - P->XTT >= 0. ? 0.5 : -0.5;  // SIGN(0.5, TT - XTT)
    const double esati = (0.5 + sgn) * w - (sgn - 0.5) * i;
    esw[e] = w;
    esi[e] = t > 273.15 ? w : esati;
    const double l = P->XLSTT + (P->XCPV - P->XCI) * (t - P->XTT);
    a2[e] = l * l / (2.38E-2 + 0.0071E-2 * (t - P->XTT)) / (t * t * P->XRV);
    bb3[e] = P->XRV / (0.211E-4 * std::pow(t / P->XTT, 1.94) * P->XP00) * t / esati;
    am3[e] = t > 273.15 ? 0. : 1000. * std::exp(12.96 * (w / esati - 1.) - 0.639);
    const double zzt = std::fmax(0., std::fmin(1., (P->XTT - t) / 20.));
    redin[e] = 1. - (1. - zzt) / std::pow(zzt * zzt * zzt + (1. - zzt) * (1. - zzt) * (1. - zzt), .333);
  }
}
