/*
 * ice3_oracle.h -- C restatement of the PHYEX CY50T1 ice_adjust / rain_ice path.
 * Test infrastructure only (see ice3_oracle.c).  All arrays are float64,
 * shape (nlev, ncol) with the column index contiguous; per-column arrays have
 * length ncol.
 */
#ifndef ICE3_ORACLE_H
#define ICE3_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ice_adjust inputs (JAX IceAdjustJAX argument order minus sigqsat/timestep) */
enum {
  IA_PABS, IA_SIGS, IA_TH, IA_EXN, IA_EXNREF, IA_RHODREF, IA_RV, IA_RC, IA_RI, IA_RR, IA_RS,
  IA_RG, IA_CF_MF, IA_RC_MF, IA_RI_MF, IA_RVS, IA_RCS, IA_RIS, IA_THS, ORACLE_ADJ_NIN
};
/* ice_adjust outputs: the 16-tuple of IceAdjustJAX (t .. ths) + srcs + th_out */
enum {
  OA_T, OA_RV, OA_RC, OA_RI, OA_CLDFR, OA_HLC_HRC, OA_HLC_HCF, OA_HLI_HRI, OA_HLI_HCF, OA_CPH,
  OA_LV, OA_LS, OA_SRCS, OA_TH, ORACLE_ADJ_NOUT_POINT,
  OA_RVS = ORACLE_ADJ_NOUT_POINT, OA_RCS, OA_RIS, OA_THS, ORACLE_ADJ_NOUT
};

/* rain_ice inputs (per point) */
enum {
  RI_EXN, RI_DZZ, RI_RHODJ, RI_RHODREF, RI_EXNREF, RI_PABS, RI_CLDFR, RI_HLC_HRC, RI_HLC_HCF,
  RI_HLI_HRI, RI_HLI_HCF, RI_THT, RI_RVT, RI_RCT, RI_RRT, RI_RIT, RI_RST, RI_RGT, RI_SIGS,
  ORACLE_RAIN_NIN
};
/* rain_ice in/out per point: sources, ci; out per point: evap3d, rainfr */
enum {
  RO_THS, RO_RVS, RO_RCS, RO_RRS, RO_RIS, RO_RSS, RO_RGS, RO_CI, RO_EVAP3D, RO_RAINFR,
  ORACLE_RAIN_NIO
};
/* per-column outputs */
enum { RC_INPRC, RC_INPRR, RC_INPRI, RC_INPRS, RC_INPRG, ORACLE_RAIN_NCOL };

/* Reads the constants dump of the compiled reference (tests/golden/constants.txt). */
int oracle_load_constants(const char *path);

/* ICE_ADJUST; returns 0, or -1 when constants are not loaded. */
int oracle_ice_adjust(int ncol, int nlev, double dt, const double *sigqsat,
                      const double *const in[ORACLE_ADJ_NIN], double *const out[ORACLE_ADJ_NOUT]);

/* extra outputs of ICE_ADJUST with OCND2 */
enum { OX_ICLDFR, OX_WCLDFR, OX_SSIO, OX_SSIU, OX_IFR, ORACLE_OCND2_NX };

/* ICE_ADJUST with PARAMI%LOCND2 = .TRUE.; zz (nlev, ncol) level heights, ice_cld_wgt per
 * column.  Returns 0, -1 without constants, -2 when nlev < 2. */
int oracle_ice_adjust_ocnd2(int ncol, int nlev, double dt, const double *sigqsat,
                            const double *ice_cld_wgt, const double *zz,
                            const double *const in[ORACLE_ADJ_NIN],
                            double *const out[ORACLE_ADJ_NOUT], double *const xout[ORACLE_OCND2_NX]);

/* RAIN_ICE with sedimentation, total tendencies and negativity correction;
 * io[RO_THS..RO_CI] are read and updated in place, io[RO_EVAP3D/RAINFR] written.
 * iters (optional, ncol*nlev ints) receives each point's tendency-call count.
 * Returns 0, -1 without constants, -2 if a point exceeded the sub-step cap. */
int oracle_rain_ice(int ncol, int nlev, double dt, const double *const in[ORACLE_RAIN_NIN],
                    double *const io[ORACLE_RAIN_NIO], double *const col[ORACLE_RAIN_NCOL],
                    int *iters);

/* ICE_ADJUST with KRR = 7: the hail mixing ratio rh (nlev, ncol) enters the heat capacity. */
int oracle_ice_adjust_hail(int ncol, int nlev, double dt, const double *sigqsat,
                           const double *const in[ORACLE_ADJ_NIN], const double *rh,
                           double *const out[ORACLE_ADJ_NOUT]);

/* RAIN_ICE with KRR = 7 (hail): as oracle_rain_ice, plus the hail state rht, its source rhs
 * (updated in place) and the per-column surface hail rate inprh.  Returns as
 * oracle_rain_ice; -1 also when the constants hold no hail tables. */
int oracle_rain_ice_hail(int ncol, int nlev, double dt, const double *const in[ORACLE_RAIN_NIN],
                         const double *rht, double *rhs, double *inprh,
                         double *const io[ORACLE_RAIN_NIO], double *const col[ORACLE_RAIN_NCOL],
                         int *iters);

/* Branch record for coverage tests: while flags (ncol*nlev unsigned) is set, every rain_ice
 * call of the calling thread writes there, per point, the ORACLE_BR_* bits of the branches
 * its tendency calls took (ORed over the calls); NULL stops the recording.  The setting is
 * per thread (so is oracle_record_adjust_branches'). */
enum {
  ORACLE_BR_RRHONG, ORACLE_BR_RIMLTC, ORACLE_BR_RIMING, ORACLE_BR_RACCSG, ORACLE_BR_SMLTG,
  ORACLE_BR_WETG, ORACLE_BR_DRYG, ORACLE_BR_GMLTR, ORACLE_BR_WETGH, ORACLE_BR_WETH,
  ORACLE_BR_DRYH, ORACLE_BR_DRYHG, ORACLE_BR_HMLTR,
  ORACLE_BR_TIE,  /* a graupel or hail wet/dry choice decided by rounding residuals */
  ORACLE_BR_SHARE_TIE  /* a graupel share of riming or rain accretion that is a rounding residual */
};
void oracle_record_branches(unsigned *flags);

/* ICE_ADJUST's branch record: while flags is not NULL, oracle_ice_adjust and
 * oracle_ice_adjust_hail write per point (nlev, ncol) the mask of the branches it took: the Q1
 * interval of the CB02 condensate (2 bits at ORACLE_ABR_Q1: 0 for Q1 <= 0, 1 for 0 < Q1 <= 2,
 * 2 for Q1 > 2), no cloud (the condensate below 1e-12 or a zero cloud fraction), ZW1 < 0, ZW2 < 0, the mass-flux limit
 * on the sources (ice_adjust.F90:377-381) and on the adjusted state (:452-456), the case of
 * each TRIANGLE pdf (2 bits each, droplets :393-406, ice :420-433: 0 above the
 * autoconversion threshold, 1 at or below half of it, 2 between) and INQ1 + 22 (6 bits). */
enum {
  ORACLE_ABR_Q1 = 0, ORACLE_ABR_NOCLOUD = 2, ORACLE_ABR_ZW1NEG = 3, ORACLE_ABR_ZW2NEG = 4,
  ORACLE_ABR_MFLIMIT = 5, ORACLE_ABR_MFOUT = 6, ORACLE_ABR_PDFC = 7, ORACLE_ABR_PDFI = 9,
  ORACLE_ABR_INQ1 = 11
};
void oracle_record_adjust_branches(unsigned *flags);

/* RAIN_ICE with PARAMI%LOCND2 = .TRUE.: as oracle_rain_ice, plus ICE_ADJUST's OCND2 outputs
 * x4 = {PICLDFR, PSSIO, PSSIU, PIFR}, each (nlev, ncol).  Returns as oracle_rain_ice, or -3
 * when a work array cannot be allocated. */
enum { RX_ICLDFR, RX_SSIO, RX_SSIU, RX_IFR, ORACLE_RAIN_NOCND2 };
int oracle_rain_ice_ocnd2(int ncol, int nlev, double dt, const double *const in[ORACLE_RAIN_NIN],
                          const double *const x4[ORACLE_RAIN_NOCND2],
                          double *const io[ORACLE_RAIN_NIO], double *const col[ORACLE_RAIN_NCOL],
                          int *iters);

/* Sedimentation scheme of every rain_ice call that follows (PARAM_ICE CSEDIM and
 * XSPLIT_MAXCFL): ORACLE_SEDIM_STAT, the default, or ORACLE_SEDIM_SPLI with the CFL bound
 * maxcfl (> 0) of its sub-steps.  Returns 0, or -1 for an unknown scheme or maxcfl <= 0. */
enum { ORACLE_SEDIM_STAT, ORACLE_SEDIM_SPLI };
int oracle_set_sedimentation(int scheme, double maxcfl);

/* ICE4_SEDIMENTATION_SPLIT alone: rt[q] the mixing ratios at t and rs[q] the sources
 * (updated in place) of rain, pristine ice, snow, graupel and, with krr = 7, hail
 * (q = 0..4; rt[4], rs[4], inpr[4] unused with krr = 6), inpr[q] (ncol each) their surface
 * rates; nsub (optional, 5 x ncol ints) the number of sub-steps of each species and column.
 * Returns 0, -1 without constants, -2 for a bad krr or maxcfl, -3 when out of memory. */
int oracle_sedim_split(int ncol, int nlev, double dt, int krr, double maxcfl, const double *dzz,
                       const double *rhodref, const double *const rt[5], double *const rs[5],
                       double *const inpr[5], int *nsub);

/* ICE4_SEDIMENTATION with CSEDIM='SPLI' as RAIN_ICE calls it before the microphysics: the
 * split sedimentation on the state at t, the ice rate added to the snow rate, then
 * ICE4_CORRECT_NEGATIVITIES of every source.  t = {th, rv, rc, rr, ri, rs, rg, rh} at t,
 * src their sources (updated in place), exn the Exner function; the rh entries are read
 * with krr = 7 only.  inpr = {inprc, inprr, inpri, inprs, inprg, inprh} (ncol each; inprh
 * with krr = 7).  nsub as for oracle_sedim_split.  Returns as oracle_sedim_split. */
int oracle_sedimentation_spli(int ncol, int nlev, double dt, int krr, double maxcfl,
                              const double *dzz, const double *rhodref, const double *exn,
                              const double *const t[8], double *const src[8],
                              double *const inpr[6], int *nsub);

#ifdef __cplusplus
}
#endif
#endif
