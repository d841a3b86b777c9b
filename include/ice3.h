/*
 * ice3.h -- C ABI of libice3_hip.so, the MI355X implementation of the ICE3
 * microphysics column kernels ice_adjust and rain_ice (PHYEX CY50T1 numerics,
 * AROME ICE3 configuration, KRR=6).
 *
 * Plain C: pointers, sizes and return codes only.  Every field pointer is a
 * device pointer to float64 data laid out (nlev, ld): element (level k,
 * column i) is at p[k*ld + i], columns contiguous, level 0 the lowest model
 * level.  This is byte-identical to the reference's Fortran (NIJT, NKT) arrays
 * (PHYEX-IAL_CY50T1/bridge/phyex_bridge.F90:96-118, NKL=1 at :128).
 * Per-column arrays have ncol entries.  Calls are asynchronous on the given
 * HIP stream (NULL = default stream); buffers stay owned by the caller.  A call
 * on an empty domain (ncol or nlev 0, with valid context, precision and dt)
 * does nothing and returns ICE3_OK; its field pointers may be NULL.
 *
 * Entry points and the reference interfaces they replace:
 *   ice3_ice_adjust      c_ice_adjust   (bridge/phyex_bridge.F90:27-34)  and the JAX
 *                        IceAdjustJAX.__call__ (src/ice3/jax/ice_adjust.py:139-230)
 *   ice3_rain_ice        c_rain_ice     (bridge/phyex_bridge.F90:242-254) and
 *                        RainIceJAX.__call__ (src/ice3/jax/rain_ice.py:181-420)
 *   ice3_adjust_rain_ice the two calls in sequence, fused on the device
 *   ..._ocnd2, ..._hail  the same with PARAM_ICE LOCND2 = .TRUE., or with hail (KRR=7):
 *                        the Fortran ICE_ADJUST / RAIN_ICE with those options
 *   ice3_ctx_create      c_ini_rain_ice (bridge/phyex_bridge.F90:215-239): INI_CST +
 *                        INI_RAIN_ICE constants and lookup tables, uploaded once
 */
#ifndef ICE3_H
#define ICE3_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define ICE3_OK 0
#define ICE3_EINVAL (-1)     /* bad size, stride, precision or NULL pointer */
#define ICE3_EDEVICE (-2)    /* HIP runtime error (launch, allocation, copy) */
#define ICE3_ESUBSTEP (-3)   /* a point hit the sub-step cap (ice3_ctx_status, ice3_stats) */
#define ICE3_ECONFIG (-4)    /* kernels built with constants other than the host builder's */
#define ICE3_ETOOLARGE (-5)  /* problem larger than the static scratch of a _noalloc entry point */

/* arithmetic of the microphysics kernels */
#define ICE3_PREC_FP64 0     /* FP64 everywhere: parity with the FP64 reference */
#define ICE3_PREC_MIXED 1    /* FP32 process rates and lookups; FP64 saturation adjustment,
                                state update, sedimentation fluxes and budgets */

typedef struct ice3_ctx ice3_ctx;

/* INI_CST + INI_RAIN_ICE for the pinned configuration; uploads constants and the
 * collection-kernel / incomplete-gamma tables to the current HIP device. */
int ice3_ctx_create(ice3_ctx **ctx);
void ice3_ctx_destroy(ice3_ctx *ctx);
const char *ice3_strerror(int code);

/* ---- sub-step cap of the ICE4 time splitting (no counterpart in the reference: its loop
 * has no bound, micro/mode_ice4_stepping.F90:256-516) ----
 * A microphysics point stops after max_calls ICE4_TENDENCIES calls (default 4096): its
 * outputs are then the processes integrated up to the time reached, with the external
 * tendencies counted once (through the sources).  Every such point is counted in a
 * context-owned device counter, whatever the stats argument of the call.
 * ice3_ctx_set_substep_cap sets max_calls (>= 1) for the calls made after it.  The
 * microphysics kernels that compact the time splitting across a wavefront have the default
 * compiled in; under another cap the microphysics runs one thread per point, with the same
 * results, more slowly.
 * ice3_ctx_status waits for the work queued on stream, returns in *capped (may be NULL)
 * the number of capped points of all calls on ctx since the previous query, resets the
 * counter, and returns ICE3_ESUBSTEP if that number is not 0, else ICE3_OK.  Calls on
 * other streams of the same context that overlap a query may be counted in it or in the
 * next one. */
int ice3_ctx_set_substep_cap(ice3_ctx *ctx, int max_calls);
int ice3_ctx_status(ice3_ctx *ctx, void *stream, unsigned long long *capped);

/* ---- sedimentation scheme: PARAM_ICE CSEDIM and XSPLIT_MAXCFL (micro/modd_param_icen.F90:83,
 * 347; the dispatch of micro/mode_ice4_sedimentation.F90:167-223) ----
 * ICE3_SEDIM_STAT (the default, the AROME configuration): ICE4_SEDIMENTATION_STAT.
 * ICE3_SEDIM_SPLI: ICE4_SEDIMENTATION_SPLIT (micro/mode_ice4_sedimentation_split.F90) on the
 * state at t, in sub-steps whose CFL number stays under max_cfl (> 0; the reference's
 * default is 0.8), followed by ICE4_CORRECT_NEGATIVITIES of every source.  The reference's
 * Python package has no split scheme, and its Fortran file for it is outside its build.
 * Applies to the rain_ice calls made after it on ctx (ice3_rain_ice*, ice3_adjust_rain_ice*,
 * ICE3_STAGE_SEDIMENTATION).  With SPLI: nlev <= 256 (else ICE3_EINVAL); the sedimentation
 * is FP64 in both precisions; the surrogate entry points return ICE3_EINVAL (the network was
 * trained with the statistical scheme); a species of a column that would take more than 65536
 * sub-steps stops there and the column is counted by ice3_ctx_status.  The fused entry points
 * keep the microphysics' state at t of th, rv and rc (the adjusted sources, which the
 * sedimentation's negativity correction updates) in a work area of the context, 24 bytes per
 * point of nlev x ld, allocated on first use and grown when a larger call needs it (growing
 * waits for the device); fused calls with SPLI on one context must not overlap, and
 * ice3_adjust_rain_ice_stages returns ICE3_EINVAL for a MICROPHYSICS stage without the
 * SEDIMENTATION stage in the same call. */
#define ICE3_SEDIM_STAT 0
#define ICE3_SEDIM_SPLI 1
int ice3_ctx_set_sedimentation(ice3_ctx *ctx, int scheme, double max_cfl);

/* Host copy of the derived constants, written as text lines "NAME n v1 .. vn"
 * (the format of the reference constants dump) into buf; returns the number
 * of bytes needed (excluding the terminating NUL).  ctx may be NULL: the
 * constants are then built on the host without touching a device. */
long ice3_dump_constants(const ice3_ctx *ctx, char *buf, long buflen);

/* ---- input validation (no counterpart in the reference: its Fortran and JAX paths take
 * NaN and infinite inputs without a check; a NaN stays in its own column here, but a
 * MIN/MAX or a threshold can turn it into a finite, wrong value) ----
 * Counts the non-finite values (NaN, +-Inf) of nfield float64 device arrays: field q
 * has rows[q] rows (nlev for a per-point field, 1 for a per-column one, 0 to skip it)
 * of ncol values at row stride ld, and counts[q] (device) is increased by the number
 * found (the caller zeroes counts).  fields and rows are host arrays.  One HBM pass over
 * the fields; the Python drop-ins call it with check_inputs=True. */
#define ICE3_MAX_CHECK_FIELDS 32
int ice3_count_nonfinite(int nfield, const double *const *fields, const int *rows, int ncol,
                         int ld, unsigned long long *counts, void *stream);

/* ---- ice_adjust: ICE_ADJUST (micro/ice_adjust.F90:7) with CONDENSATION ---- */
typedef struct {
  const double *pabs, *sigs, *th, *exn, *exn_ref, *rho_dry_ref;
  const double *rv, *rc, *ri, *rr, *rs, *rg;
  const double *cf_mf, *rc_mf, *ri_mf;
  const double *rvs, *rcs, *ris, *ths;  /* sources at entry */
  const double *sigqsat;                /* per column */
} ice3_ice_adjust_in;

typedef struct {   /* the IceAdjustJAX 16-tuple, same order */
  double *t, *rv_out, *rc_out, *ri_out, *cldfr;
  double *hlc_hrc, *hlc_hcf, *hli_hri, *hli_hcf;
  double *cph, *lv, *ls;
  double *rvs, *rcs, *ris, *ths;        /* updated sources; may alias the inputs */
  double *srcs;                         /* optional (NULL): s'rc'/2sigma_s^2 * lambda3 */
} ice3_ice_adjust_out;

int ice3_ice_adjust(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld, double dt,
                    const ice3_ice_adjust_in *in, const ice3_ice_adjust_out *out,
                    void *stream);

/* ---- ice_adjust with OCND2: ICE_ADJUST called with PARAM_ICE LOCND2 = .TRUE.
 * (micro/ice_adjust.F90:533-542 passing PARAMI%LOCND2 to CONDENSATION): liquid-only
 * statistical condensation, its noise check, and the separate ice cloud of ICECLOUD
 * (micro/mode_icecloud.F90) over the TIWMX saturation tables (micro/mode_tiwmx.F90); the
 * OCND2 branch of micro/condensation.F90:263-265,373-394,582-637.  The reference reaches it
 * through the Fortran ICE_ADJUST only (its JAX component pins OCND2 = False,
 * src/ice3/jax/ice_adjust.py:116).  FP64; nlev >= 2 (the layer thickness needs two levels). */
typedef struct {
  const double *zz;           /* PZZ: level heights (m), (nlev, ld), increasing upward */
  const double *ice_cld_wgt;  /* PICE_CLD_WGT, per column */
} ice3_ocnd2_in;

typedef struct {   /* PICLDFR, PWCLDFR, PSSIO, PSSIU, PIFR */
  double *icldfr, *wcldfr, *ssio, *ssiu, *ifr;
} ice3_ocnd2_out;

int ice3_ice_adjust_ocnd2(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                          const ice3_ice_adjust_in *in, const ice3_ocnd2_in *in2,
                          const ice3_ice_adjust_out *out, const ice3_ocnd2_out *out2,
                          void *stream);

/* ---- rain_ice: RAIN_ICE (micro/rain_ice.F90:7) with statistical sedimentation,
 *      total tendencies and negativity correction ---- */
typedef struct {
  const double *exn, *dzz, *rhodj, *rhodref, *exn_ref, *pabs;
  const double *cldfr, *hlc_hrc, *hlc_hcf, *hli_hri, *hli_hcf, *sigs;
  const double *tht, *rvt, *rct, *rrt, *rit, *rst, *rgt;   /* state at t */
} ice3_rain_ice_in;

typedef struct {   /* read and updated in place */
  double *ths, *rvs, *rcs, *rrs, *ris, *rss, *rgs, *ci;
} ice3_rain_ice_inout;

typedef struct {
  double *evap3d, *rainfr;                           /* per point */
  double *inprc, *inprr, *inpri, *inprs, *inprg;     /* per column, surface precipitation */
} ice3_rain_ice_out;

typedef struct {   /* optional diagnostics, device pointers, each may be NULL */
  int *calls;               /* per point: number of ICE4_TENDENCIES calls */
  unsigned long long *counters; /* [0] total tendency calls, [1] points that hit the cap,
                                   [2] microphysics points; accumulated (caller zeroes) */
} ice3_stats;

int ice3_rain_ice(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld, double dt,
                  const ice3_rain_ice_in *in, const ice3_rain_ice_inout *io,
                  const ice3_rain_ice_out *out, const ice3_stats *stats, void *stream);

/* ---- rain_ice with OCND2: RAIN_ICE called with PARAM_ICE LOCND2 = .TRUE. (FP64).  The
 * microphysics mask on PSSIO and XFRMIN(12:13) (micro/rain_ice.F90:348-364), nucleation from
 * the tabulated Meyers and reduction functions (micro/ice4_nucleation.func.h:127-139), and in
 * ICE4_TENDENCIES the tabulated ZESI and ZAI, the OCND2 collision factors, ICE4_FAST_RI_RS in
 * place of ICE4_FAST_RI and the OCND2 terms (micro/mode_ice4_tendencies.F90:304-346,
 * 445-457,568-599,625-691).  The reference reaches it through the Fortran RAIN_ICE only.
 * Its extra inputs are ICE_ADJUST's OCND2 outputs (ice3_ice_adjust_ocnd2). */
typedef struct {   /* PICLDFR, PSSIO, PSSIU, PIFR, (nlev, ld) */
  const double *icldfr, *ssio, *ssiu, *ifr;
} ice3_rain_ocnd2_in;

int ice3_rain_ice_ocnd2(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                        const ice3_rain_ice_in *in, const ice3_rain_ocnd2_in *in2,
                        const ice3_rain_ice_inout *io, const ice3_rain_ice_out *out,
                        const ice3_stats *stats, void *stream);

/* ---- hail: the configuration with HCLOUD='ICE4' (KRR=7), FP64.  The reference reaches it
 * through the Fortran ICE_ADJUST and RAIN_ICE with KRR=7 and their optional hail arguments
 * (micro/ice_adjust.F90:18,193; micro/rain_ice.F90:18,247-249): the hail mixing ratio in the
 * heat capacities and the microphysics mask (micro/ice_adjust.F90:503-507,
 * micro/rain_ice.F90:337-372), hail sedimentation (micro/mode_ice4_sedimentation_stat.F90:
 * 184-188,208-222), ICE4_FAST_RH with the SWETH/GWETH/RWETH wet-growth kernels
 * (micro/mode_ice4_fast_rh.F90, micro/mode_read_xker_{sweth,gweth,rweth}.F90), the conversion
 * of wet-growing graupel into hail (micro/mode_ice4_fast_rg.F90:330-351), the hail terms of
 * the tendencies (micro/mode_ice4_tendencies.F90:619-716), of the total tendencies
 * (micro/rain_ice.F90:771-791) and of the negativity correction
 * (micro/mode_ice4_correct_negativities.F90:115-120,165-170).  The bridge's c_ice_adjust and
 * c_rain_ice (bridge/phyex_bridge.F90:27-34,242-254) pass no hail array. */
typedef struct {
  const double *rht;    /* hail mixing ratio at t, (nlev, ld) */
  double *rhs;          /* hail source, read and updated in place */
  double *inprh;        /* per column: surface hail precipitation rate (m/s) */
} ice3_hail;

/* ICE_ADJUST with KRR=7: rh (hail mixing ratio, (nlev, ld)) enters the heat capacity. */
int ice3_ice_adjust_hail(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                         const ice3_ice_adjust_in *in, const double *rh,
                         const ice3_ice_adjust_out *out, void *stream);

/* RAIN_ICE with KRR=7: as ice3_rain_ice, plus the hail arrays. */
int ice3_rain_ice_hail(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                       const ice3_rain_ice_in *in, const ice3_rain_ice_inout *io,
                       const ice3_rain_ice_out *out, const ice3_hail *hail,
                       const ice3_stats *stats, void *stream);

/* ---- branch record of the microphysics (no counterpart in the reference: a test hook for
 *      the per-point branch accounting of the parity tests) ----
 * As ice3_rain_ice in FP64 (ice3_rain_ice_hail when hail is not NULL), with the microphysics
 * run one thread per point; branches ((nlev, ld) unsigned) receives for each point the
 * ICE3_BR_* bits of the process branches its tendency calls took, ORed over the calls (0
 * outside the microphysics mask): RRHONG and RIMLTC (hard calls); snow riming into graupel,
 * rain accreted by snow into graupel, snow melting into graupel (ICE4_FAST_RS); graupel wet
 * growth, dry growth and melting (ICE4_FAST_RG); with hail, wet-growing graupel converted into
 * hail, hail wet and dry growth, hail converted into graupel and hail melting (ICE4_FAST_RH).
 * A bit is set when the process's rate is positive, or for the wet/dry choices when the
 * choice is made. */
enum {
  ICE3_BR_RRHONG, ICE3_BR_RIMLTC, ICE3_BR_RIMING, ICE3_BR_RACCSG, ICE3_BR_SMLTG, ICE3_BR_WETG,
  ICE3_BR_DRYG, ICE3_BR_GMLTR, ICE3_BR_WETGH, ICE3_BR_WETH, ICE3_BR_DRYH, ICE3_BR_DRYHG,
  ICE3_BR_HMLTR
};
int ice3_rain_ice_branches(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                           const ice3_rain_ice_in *in, const ice3_rain_ice_inout *io,
                           const ice3_rain_ice_out *out, const ice3_hail *hail,
                           unsigned *branches, void *stream);

/* The same for ICE_ADJUST (FP64; with the hail heat capacity when rh is not NULL, KRR=6 when
 * it is): per point the Q1 interval of the CB02 condensate (2 bits at ICE3_ABR_Q1: 0 for
 * Q1 <= 0, 1 for 0 < Q1 <= 2, 2 for Q1 > 2), no cloud (the condensate below 1e-12 or a zero
 * cloud fraction before the mass-flux part), ZW1 < 0 and ZW2 < 0, the mass-flux limits on the
 * sources and on the adjusted state, the case of each TRIANGLE pdf (droplets at
 * ICE3_ABR_PDFC, ice at ICE3_ABR_PDFI, 2 bits each: 0 above the autoconversion threshold, 1 at
 * or below half of it, 2 between) and INQ1 + 22 (6 bits at ICE3_ABR_INQ1).  It also writes
 * out as ice3_ice_adjust (ice3_ice_adjust_hail) does, one thread per point. */
enum {
  ICE3_ABR_Q1 = 0, ICE3_ABR_NOCLOUD = 2, ICE3_ABR_ZW1NEG = 3, ICE3_ABR_ZW2NEG = 4,
  ICE3_ABR_MFLIMIT = 5, ICE3_ABR_MFOUT = 6, ICE3_ABR_PDFC = 7, ICE3_ABR_PDFI = 9,
  ICE3_ABR_INQ1 = 11
};
int ice3_ice_adjust_branches(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                             const ice3_ice_adjust_in *in, const double *rh,
                             const ice3_ice_adjust_out *out, unsigned *branches, void *stream);

/* ---- fused ice_adjust + rain_ice (the "adjust then microphysics" time step) ----
 * ice_adjust runs on the state at t; its updated sources, times dt, are the
 * t-state of rain_ice (the synthetic and AROME-like inputs carry the state at t
 * in the sources, rxs = rx/dt); sedimentation and microphysics follow. */
typedef struct {
  const double *pabs, *sigs, *th, *exn, *exn_ref, *rhodref, *rhodj, *dzz;
  const double *rv, *rc, *ri, *rr, *rs, *rg, *ci;
  const double *cf_mf, *rc_mf, *ri_mf;
  const double *ths, *rvs, *rcs, *rrs, *ris, *rss, *rgs;
  const double *sigqsat;                 /* per column */
} ice3_pipeline_in;

typedef struct {
  double *ths, *rvs, *rcs, *rrs, *ris, *rss, *rgs;   /* final sources */
  double *cldfr, *hlc_hrc, *hlc_hcf, *hli_hri, *hli_hcf, *ci, *evap3d, *rainfr;
  double *inprc, *inprr, *inpri, *inprs, *inprg;     /* per column */
} ice3_pipeline_out;

int ice3_adjust_rain_ice(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld, double dt,
                         const ice3_pipeline_in *in, const ice3_pipeline_out *out,
                         const ice3_stats *stats, void *stream);

/* The fused step with hail (KRR=7): rh is the hail mixing ratio at t (ICE_ADJUST's heat
 * capacity), rhs the hail source at entry (times dt, the hail t-state of rain_ice, as for the
 * other species), rhs_out the final hail source (must not alias rhs) and inprh the surface hail
 * rate.  FP64. */
typedef struct {
  const double *rh, *rhs;
  double *rhs_out;
  double *inprh;        /* per column */
} ice3_pipeline_hail;

int ice3_adjust_rain_ice_hail(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                              const ice3_pipeline_in *in, const ice3_pipeline_out *out,
                              const ice3_pipeline_hail *hail, const ice3_stats *stats,
                              void *stream);

/* The fused step with OCND2 in both halves: ice3_ice_adjust_ocnd2's ICE_ADJUST (in2: level
 * heights and PICE_CLD_WGT; out2: its ice cloud outputs), then the sedimentation and
 * ice3_rain_ice_ocnd2's microphysics on the adjusted state, with out2 as its OCND2 inputs.
 * FP64; nlev >= 2. */
int ice3_adjust_rain_ice_ocnd2(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                               const ice3_pipeline_in *in, const ice3_ocnd2_in *in2,
                               const ice3_pipeline_out *out, const ice3_ocnd2_out *out2,
                               const ice3_stats *stats, void *stream);

/* Stages of the fused step, for ice3_adjust_rain_ice_stages: run any subset, in this order,
 * on the same in/out buffers (the outputs of one call are the state the next call reads). */
#define ICE3_STAGE_ADJUST 1          /* ICE_ADJUST into out (adjusted sources, cloud state) */
#define ICE3_STAGE_SEDIMENTATION 2   /* ICE4_SEDIMENTATION_STAT into out (+ surface precip) */
#define ICE3_STAGE_MICROPHYSICS 4    /* ICE4 time splitting, total tendencies, negativities */
int ice3_adjust_rain_ice_stages(ice3_ctx *ctx, int prec, int ncol, int nlev, int ld, double dt,
                                const ice3_pipeline_in *in, const ice3_pipeline_out *out,
                                const ice3_stats *stats, int stages, void *stream);

/* ---- learned surrogate of the ICE4 microphysics (no counterpart in the reference) ----
 * A per-point MLP (18 features -> 4 x 128 SiLU -> 8 outputs, bf16 weights on the matrix
 * cores, fp32 accumulation) replaces the microphysics stage of the fused step.  Its
 * features are T, ln p, rhodref, ln(r + 1e-9) of rv, rc, rr, ri, rs, rg (t-state),
 * ln(1 + ci), sigma_s, exner, dt, the cloud fraction and asinh(1e6 dt e) of the
 * sedimentation tendency e of rr, ri, rs, rg; its outputs are the changes the
 * microphysics makes to the sources of th, rv, rc, rr, ri, rs, rg and to ci, each as
 * unit * sinh(mean + std * y).  A point goes back to the plain FP64 microphysics when a
 * feature lies outside the training range, when the water residual sum(d rv..d rg)
 * exceeds budget_tol * (sum|d r| + u_min) (u_min: the smallest targ_unit of rv..rg),
 * when a resulting source of rc..rg is negative by more than
 * neg_tol * (|source| + |change| + targ_unit), or when it has no hydrometeor (the plain
 * kernel's nucleation-only branch).  Accepted points have their water residual moved into rv and
 * then pass ICE4_CORRECT_NEGATIVITIES, as the plain kernel's results do, so total water
 * is conserved and every source is non-negative. */
typedef struct {
  float feat_mean[18], feat_scale[18];   /* x = (f - mean) * scale */
  float feat_lo[18], feat_hi[18];        /* training range of f */
  float targ_unit[8], targ_mean[8], targ_std[8];
} ice3_surrogate_norm;

typedef struct {
  const void *weights_packed;            /* bf16 packed fragments (ice3_mlp_adam output) */
  const float *params;                   /* fp32 parameters (biases) */
  const ice3_surrogate_norm *norm;       /* device pointer */
  float budget_tol;
  float neg_tol;
} ice3_surrogate;

/* Fused step with the surrogate microphysics.  work: device scratch of
 * ice3_surrogate_work_bytes(ncol, nlev) bytes, 256-byte aligned (network inputs and
 * outputs, reasons, fallback queue).  sg_counters (device, 5 entries, accumulated):
 * [0] points taken from the network, [1] out of range, [2] negative source,
 * [3] water residual, [4] no hydrometeor.  stats as for ice3_adjust_rain_ice (fallback
 * points only; network points report 0 calls). */
size_t ice3_surrogate_work_bytes(int ncol, int nlev);
int ice3_adjust_rain_ice_surrogate(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                   const ice3_pipeline_in *in, const ice3_pipeline_out *out,
                                   const ice3_surrogate *sg, void *work,
                                   unsigned long long *sg_counters, const ice3_stats *stats,
                                   void *stream);

/* The same step with no work argument and no allocation (no counterpart in the reference:
 * for callers that cannot allocate device memory, such as code under HIP graph capture or
 * serving loops that forbid it).  Its scratch is a static device array of
 * ICE3_SURROGATE_SCRATCH_BYTES per GPU (64 MiB), in place from the first ice3_ctx_create
 * on that GPU until the process ends; a problem whose ice3_surrogate_work_bytes(ncol, nlev)
 * exceeds it (about 650,000 points: 7,000 columns of 90 levels) returns ICE3_ETOOLARGE.
 * Every call on one GPU shares that array: calls that overlap on one GPU (other streams,
 * other contexts, other threads) race and must be ordered by the caller.  Nothing in the
 * scratch needs to start at zero, and no call leaves state there that a later call reads. */
#define ICE3_SURROGATE_SCRATCH_BYTES ((size_t)64 << 20)
int ice3_adjust_rain_ice_surrogate_noalloc(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                                           const ice3_pipeline_in *in,
                                           const ice3_pipeline_out *out, const ice3_surrogate *sg,
                                           unsigned long long *sg_counters,
                                           const ice3_stats *stats, void *stream);

/* Surrogate features of the state between the sedimentation and microphysics stages
 * (in/out as after ice3_adjust_rain_ice_stages(ADJUST | SEDIMENTATION)).  raw: [npts][18]
 * fp32 (may be NULL); x: [npts][32] bf16 normalised with norm (both may be NULL);
 * micro: [npts] 1 where the point runs the microphysics (may be NULL).  npts = nlev*ncol,
 * point q = k*ncol + i. */
int ice3_surrogate_features(ice3_ctx *ctx, int ncol, int nlev, int ld, double dt,
                            const ice3_pipeline_in *in, const ice3_pipeline_out *out,
                            const ice3_surrogate_norm *norm, float *raw, void *x,
                            unsigned char *micro, void *stream);

/* Training kernels of the MLP (device pointers, fp32 parameters in ice3_mlp_sizes order:
 * row-major weights of the 5 layers, then the biases).  x: [npts][32] bf16 features;
 * target: [npts][8] fp32 normalised outputs.  ice3_mlp_forward: y [npts][8] (may be
 * NULL); for training also zsave [4][npts][128] fp32, hsave [4][npts][128] bf16,
 * dy [npts][32] fp32 and loss_part [ceil(npts/128)] (sum of squared errors per tile).
 * ice3_mlp_backward: gpart [ceil(npts/128)][n_params] scratch, grad [n_params].
 * ice3_mlp_adam: one Adam step (grad NULL: only rewrite the bf16 copies wrm, wpk). */
int ice3_mlp_sizes(int *n_params, int *n_weights, int *tile, int *in_width, int *out_width);
int ice3_mlp_forward(int npts, const void *wpk, const float *params, const void *x,
                     const float *target, float *y, float *zsave, void *hsave, float *dy,
                     float *loss_part, void *stream);
int ice3_mlp_backward(int npts, const void *wrm, const void *x, const float *zsave,
                      const void *hsave, const float *dy, float *gpart, float *grad,
                      void *stream);
int ice3_mlp_adam(int step, float lr, float beta1, float beta2, float eps, float *params,
                  float *m, float *v, const float *grad, void *wrm, void *wpk, void *stream);

/* ---- per-stage timing (no counterpart in the reference: a measurement hook) ----
 * ice3_stage_timing(ctx, 1) makes the three entry points above record HIP events
 * around each of their kernels on the call's stream; ice3_stage_times waits for
 * the last such call and returns its stage durations in milliseconds:
 * ms[0] ICE_ADJUST, ms[1] sedimentation, ms[2] microphysics (in the FP64 fused step and
 * the surrogate step the first two stages run as one kernel: ms[0] covers both and ms[1]
 * is about 0); for the surrogate step
 * ms[2] is the feature kernel and ms[3] the network, ms[4] the decode and ms[5] the
 * fallback microphysics (0 for a stage the call did not run).  Off by default; a
 * context's events serve one call at a time. */
int ice3_stage_timing(ice3_ctx *ctx, int enable);
int ice3_stage_times(ice3_ctx *ctx, float ms[6]);

#ifdef __cplusplus
}
#endif
#endif /* ICE3_H */
