// ice3_params.h -- constants and lookup tables of the AROME ICE3 configuration
// (KRR=6) and of its hail variant (KRR=7, HCLOUD='ICE4'), as INI_CST
// (aux/mode_ini_cst.F90:77-193) and INI_RAIN_ICE (micro/mode_ini_rain_ice.F90:10-1571)
// define them.  Plain data shared by the host builder (ice3_params.cpp) and the device
// kernels.
#pragma once

struct ice3_params {
  // CST_t
  double XTT, XLVTT, XLSTT, XLMTT, XCPD, XCPV, XCL, XCI, XRD, XRV, XEPSILO;
  double XALPW, XBETAW, XGAMW, XALPI, XBETAI, XGAMI, XESTT, XP00, XRHOLW, XPI, XG, XMNH_TINY;
  // NEB_t (pinned: 'T' ice fraction between these temperatures)
  double XTMINMIX, XTMAXMIX;
  // RAIN_ICE_DESCR_t
  double XRTMIN[8];  // 1-based, [1..7]; [7] (hail) is read by the KRR=7 kernels only
  double XCEXVT, XLBR, XLBEXR, XLBS, XLBEXS, XLBG, XLBEXG, XLBI, XLBEXI, XLBDAS_MAX, XDI, XBS,
      XCXS, XCXG, XDG, XCONC_LAND;
  // RAIN_ICE_PARAM_t
  double XCRIAUTC, XCRIAUTI, XACRIAUTI, XBCRIAUTI, XT0CRIAUTI, XNU10, XALPHA1, XBETA1, XNU20,
      XALPHA2, XBETA2, XMNU0, XALPHA3, XBETA3, XHON, XSCFAC, X0DEPI, X2DEPI, X0DEPS, X1DEPS,
      XEX0DEPS, XEX1DEPS, X0DEPG, X1DEPG, XEX0DEPG, XEX1DEPG, XTIMAUTI, XTEXAUTI, XCOLEXIS,
      XFIAGGS, XEXIAGGS, XTIMAUTC, XFCACCR, XEXCACCR, X0EVAR, X1EVAR, XEX0EVAR, XEX1EVAR,
      XEXCRIMSS, XCRIMSS, XEXCRIMSG, XCRIMSG, XSRIMCG, XEXSRIMCG, XSRIMCG2, XSRIMCG3, XEXSRIMCG2,
      XRIMINTP1, XRIMINTP2, XFRACCSS, XLBRACCS1, XLBRACCS2, XLBRACCS3, XFSACCRG, XLBSACCR1,
      XLBSACCR2, XLBSACCR3, XACCINTP1S, XACCINTP2S, XACCINTP1R, XACCINTP2R, XFSCVMG, XEXRCFRI,
      XRCFRI, XEXICFRR, XICFRR, XFCDRYG, XCOLIG, XCOLEXIG, XFIDRYG, XCOLSG, XCOLEXSG, XFSDRYG,
      XLBSDRYG1, XLBSDRYG2, XLBSDRYG3, XFRDRYG, XLBRDRYG1, XLBRDRYG2, XLBRDRYG3, XDRYINTP1R,
      XDRYINTP2R, XDRYINTP1S, XDRYINTP2S, XDRYINTP1G, XDRYINTP2G, XFSEDR, XEXSEDR, XFSEDS,
      XEXSEDS, XFSEDG, XEXSEDG, XFSEDI, XEXCSEDI;
  // hail, KRR=7 (RAIN_ICE_DESCR_t, RAIN_ICE_PARAM_t; micro/mode_ini_rain_ice.F90:375-392,
  // 436-437,462-463,547-549,633-636,1225-1296)
  double XLBH, XLBEXH, XCXH, XDH, XBG, XFSEDH, XEXSEDH, X0DEPH, X1DEPH, XEX0DEPH, XEX1DEPH,
      XCOLIH, XCOLEXIH, XCOLSH, XCOLEXSH, XCOLGH, XCOLEXGH, XFWETH, XFSWETH, XLBSWETH1,
      XLBSWETH2, XLBSWETH3, XFGWETH, XLBGWETH1, XLBGWETH2, XLBGWETH3, XFRWETH, XLBRWETH1,
      XLBRWETH2, XLBRWETH3, XWETINTP1S, XWETINTP2S, XWETINTP1G, XWETINTP2G, XWETINTP1R,
      XWETINTP2R, XWETINTP1H, XWETINTP2H;
  double XFRMIN[41];  // 1-based, [1..40]: the PARAM_ICE XFRMIN_NAM tuning values (OCND2 branch)
  int NGAMINC, NACCLBDAS, NACCLBDAR, NDRYLBDAR, NDRYLBDAS, NDRYLBDAG;
  int NWETLBDAS, NWETLBDAG, NWETLBDAR, NWETLBDAH;
  // natural logarithms of constants, for powers evaluated as exp(c*ln x) on the device
  double LN_XLBR, LN_XLBS, LN_XLBG, LN_XLBI, LN_XLBDAS_MAX, LN_XTT, LN_1E8, LN_10, LN_XLBH;
};

// Table block: one contiguous float64 array, offsets in doubles.
enum ice3_table_offsets {
  ICE3_T_RIM1 = 0,                      // GAMINC_RIM1(80)
  ICE3_T_RIM2 = 80,                     // GAMINC_RIM2(80)
  ICE3_T_RIM4 = 160,                    // GAMINC_RIM4(80)
  ICE3_T_RACCSS = 240,                  // (40,40) lbda_s x lbda_r
  ICE3_T_RACCS = ICE3_T_RACCSS + 1600,  // (40,40)
  ICE3_T_SACCRG = ICE3_T_RACCS + 1600,  // (40,40) lbda_r x lbda_s
  ICE3_T_SDRYG = ICE3_T_SACCRG + 1600,  // (40,80) lbda_g x lbda_s
  ICE3_T_RDRYG = ICE3_T_SDRYG + 3200,   // (40,40) lbda_g x lbda_r
  ICE3_T_SWETH = ICE3_T_RDRYG + 1600,   // (40,80) lbda_h x lbda_s, hail wet growth
  ICE3_T_GWETH = ICE3_T_SWETH + 3200,   // (40,40) lbda_h x lbda_g
  ICE3_T_RWETH = ICE3_T_GWETH + 1600,   // (40,40) lbda_h x lbda_r
  ICE3_T_TOTAL = ICE3_T_RWETH + 1600,   // 16240 doubles
  ICE3_T_ICE3 = ICE3_T_SWETH            // 9840: the part the KRR=6 kernels read (staged in LDS)
};

// Builds every constant and the table block (tables must hold ICE3_T_TOTAL doubles).
void ice3_build_params(ice3_params *p, double *tables);

// TIWMX tables of the OCND2 branch (micro/modd_tiwmx.F90: XNDEGR = 100, NSTART = 13200,
// NSTOP = 37316): entry jk - ICE3_TIWMX_START of each table holds its function at jk / 100 K,
// computed as INI_TIWMX does (micro/mode_ini_tiwmx.F90:14-49).  The tables follow one another
// in one block, ICE3_TIWMX_N doubles each, in the order of ice3_tiwmx_tables.
enum { ICE3_TIWMX_START = 13200, ICE3_TIWMX_STOP = 37316,
       ICE3_TIWMX_N = ICE3_TIWMX_STOP - ICE3_TIWMX_START + 1 };
enum ice3_tiwmx_tables {
  ICE3_TW_ESW = 0,  // ESTABW: saturation pressure over water
  ICE3_TW_ESI,      // ESTABI: over ice below 273.15 K, over water above
  ICE3_TW_A2,       // A2TAB: AA2, the heat-conduction term of the ice growth function
  ICE3_TW_BB3,      // BB3TAB: BB3, the vapour-diffusion term (times pressure)
  ICE3_TW_AM3,      // AM3TAB: Meyers ice-nucleus concentration, 0 above 273.15 K
  ICE3_TW_REDIN,    // REDINTAB: reduction of ice nuclei between 0 and -20 C
  ICE3_TW_NTAB
};
// tab must hold ICE3_TW_NTAB * ICE3_TIWMX_N doubles; p are the constants ice3_build_params made.
void ice3_build_tiwmx(const ice3_params *p, double *tab);
