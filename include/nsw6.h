/*
 * nsw6.h -- C ABI of the MI355X-native NICAM NSW6 single-moment microphysics step.
 *
 * This header is the drop-in boundary for NICAM's `mp_driver` call chain
 * (mp_driver -> MP_NSW6 rates -> PRECIP_transport/vadv1d -> SATURATION_adjustment
 * -> THRMDYN).  Every array is FP64, in NICAM's (ij,k,l) Fortran layout, i.e. a
 * row-major C array [l][k][ij] with ij fastest.  Tracer arrays are [l][nq][k][ij]
 * (NICAM's *_Lswp ordering), per-column arrays [l][ij], two-row precip [l][2][ij]
 * and one-row diagnostics (rctop, rwtop, tctop) [l][1][ij].
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   nsw6_init            <- MP_DRIVER::mp_init("NSW6") + GRD_Setup + CVW_CPW_Setup +
 *                           SATURATION_Setup (src/mod_mp_driver.cpp:11-24,
 *                           src/mod_debug.cpp:76-187, src/mod_satadjust.cpp:24-47)
 *   nsw6_params_default  <- MP_NSW6::mp_nsw6_init (src/mod_mp_nsw6.cpp:276-434) and the
 *                           effective-radius constants of mp_nsw6_new::init
 *                           (src/mod_mp_nsw6.cpp:506-517); struct GlobalParams
 *                           (include/mod_mp_nsw6.h:16-123)
 *   nsw6_mp_driver       <- MP_DRIVER::mp_driver (include/mod_mp_driver.h:71-124,
 *                           src/mod_mp_driver.cpp:335-613), same argument order
 *   nsw6_mp_driver_fields<- the same call with the 49 arrays gathered in a struct
 *   nsw6_mp_driver_adj   <- no counterpart: the reverse mode (vector-Jacobian product)
 *                           of one FP64 mp_driver step
 *   nsw6_finalize        <- the end of the reference's process (views released)
 *
 * All device pointers handed to nsw6_mp_driver* must be HIP device memory on the
 * device the context was created on.  Calls are stream-ordered and asynchronous;
 * errors detected on the device (Newton non-convergence, NaN temperature) are
 * reported by nsw6_last_status() after the stream has been synchronised.
 */
#ifndef NSW6_H
#define NSW6_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSW6_KALL  96   /* ADM_kall (include/problemsize.h:64)                  */
#define NSW6_KMIN  1    /* ADM_kmin (include/problemsize.h:66), 0-based          */
#define NSW6_KMAX  94   /* ADM_kmax (include/problemsize.h:67), 0-based          */
#define NSW6_NQMAX 6    /* QV,QC,QR,QI,QS,QG (include/problemsize.h:226-231)    */

enum nsw6_status {
    NSW6_OK             = 0,
    NSW6_NOT_CONVERGED  = 1,  /* satadjust Newton hit itelim=100 (src/mod_satadjust.cpp:883-891) */
    NSW6_NAN_DETECTED   = 2,  /* tem*0 != 0 guard fired (src/mod_satadjust.cpp:880)             */
    NSW6_BAD_ARG        = 3,
    NSW6_HIP_ERROR      = 4,
    NSW6_NO_DEVICE      = 5,
    NSW6_UNSUPPORTED    = 6,  /* the reverse mode was asked for a scheme it does not  */
                              /* cover (nsw6_mp_driver_adj)                           */
    NSW6_IO_ERROR       = 7   /* a snapshot file is missing, unreadable or of the     */
                              /* wrong size                                           */
};

enum nsw6_precision {
    NSW6_FP64  = 0,   /* every operation in FP64 (parity form)                            */
    NSW6_MIXED = 1    /* FP32 process rates / fall speeds / slope+gamma terms, FP64 energy, */
                      /* Newton inversion and column flux / mass sums                      */
};

/* Scheme coefficients: GlobalParams after mp_nsw6_init, plus the two literals of the
 * rate code that are exposed as tunable coefficients (kk_aut_rate, icenuc_coef) and
 * the effective-radius constants of mp_nsw6_new::init. */
typedef struct nsw6_params {
    double N0r, N0s, N0g, rho_s, rho_g, C_d, Cr, Cs;
    double Ar, As, Ag, Br, Bs, Bg, Cg, Dr, Ds, Dg;
    double GAM, GAM_2, GAM_3;
    double GAM_1br, GAM_2br, GAM_3br, GAM_3dr, GAM_6dr, GAM_1brdr, GAM_5dr_h;
    double GAM_1bs, GAM_2bs, GAM_3bs, GAM_3ds, GAM_1bsds, GAM_5ds_h;
    double GAM_1bg, GAM_3dg, GAM_1bgdg, GAM_5dg_h;
    double sw_kk2000, sw_roh2014, ln10;
    double coef_a[10], coef_b[10];
    double Eiw, Erw, Esw, Egw, Eri, Esi, Egi, Esr, Egr, Egs;
    double gamma_sacr, gamma_gacs, mi;
    double Nc_def;
    double beta_saut, gamma_saut, beta_gaut, gamma_gaut, qicrt_saut, qscrt_gaut;
    double f1r, f2r, f1s, f2s, f1g, f2g;
    double A_frz, B_frz;
    double mi40, mi50, vti50, Ri50;
    double sw_expice, Nc_ihtr, Di_max, Di_a;
    double sw_constVti, CONST_Vti;
    double kk_aut_rate;   /* 1350.0  in Praut_kk  (src/mod_mp_nsw6.cpp:1154) */
    double icenuc_coef;   /* 4.92E-11 in Qi0      (src/mod_mp_nsw6.cpp:1083) */
    double dgamma_a, GAM_dgam, GAM_dgam23, coef_dgam, coef_xf;  /* src/mod_mp_nsw6.cpp:511-516 */
    /* Cloud-top diagnostics rctop, tctop, rwtop.  0 (default): the reference as
     * written (src/mod_mp_nsw6.cpp:1575-1582): every level writes sw*rceff +
     * (1-sw)*UNDEF with sw = [qc+qr > q_min]*zerosw, and the serial order leaves level
     * kmax.  1: the cloud top its comments describe: sw = [qc+qr > q_min]*(1-zerosw)
     * ("if qc+qr > 0.01[g/kg], sw=1"), and the highest level with sw = 1 gives
     * rctop = rceff and tctop = T there (UNDEF when there is none); rwtop as
     * TagUpdateCLD (src/mod_mp_nsw6.cpp:717-720).  Deterministic in both modes. */
    double cloudtop_intended;
    /* OPT_INDIR (src/mod_mp_nsw6.cpp:131,475-484): 1 = aerosol indirect effect, the
     * droplet number is Nc = max(unccn * 1e-6, Nc_def) per point (TagAerosolT,
     * :676-679) instead of Nc_def; the reference default is 0. */
    double opt_indir;
    /* PRCIP_TRN_ECORRECT (src/problemsize.cpp:39): 0 = "KIN2EIN" (the reference
     * default: the kinetic energy lost by the sedimentation flux goes to internal
     * energy), 1 = "KIN2KIN" (src/mod_precip_transport.cpp:1012-1068: the velocities
     * are rescaled to keep the transported kinetic energy; internal energy untouched). */
    double kin2kin;
    /* precip_transport_type (src/mod_mp_nsw6.cpp:127, 593-600): 0 = "3WATER" (the
     * reference default: QR, QS, QG sediment), 1 = "4WATER" (QI sediments too, with
     * Vt(QI) = Vti of src/mod_mp_nsw6.cpp:1073-1075; sw_constVti = 1 with CONST_Vti > 0
     * forces the constant fall speed -CONST_Vti, as mp_nsw6_init sets it, :399-407). */
    double four_water;
    /* dtemp_criteria of the saturation adjustment's Newton (src/mod_satadjust.cpp:788):
     * it stops when |dtemp| < this [K]; the reference fixes 10^(-(RP_PREC+1)/2) = 1e-8
     * (RP_PREC = 15, integer division).  A caller may tighten it; 0 makes every Newton
     * point run to itelim = 100 and report NSW6_NOT_CONVERGED. */
    double satadjust_dtemp_criteria;
} nsw6_params;

/* The 8 tunable NSW6 coefficients, in this order, for gradients and perturbations:
 * 0 kk_aut_rate (KK2000 autoconversion rate), 1 qicrt_saut (autoconversion threshold),
 * 2 Esi (snow-ice collection efficiency), 3 Esr (snow-rain collection efficiency),
 * 4 Cr, 5 Cs, 6 Cg (rain/snow/graupel fall-speed prefactors),
 * 7 icenuc_coef (ice nucleation coefficient). */
#define NSW6_NCOEF 8

/* Vertical grid (GRD_* of include/problemsize.h:143-163, filled by GRD_Setup). */
typedef struct nsw6_grid {
    double gz[NSW6_KALL], gzh[NSW6_KALL];
    double dgz[NSW6_KALL], dgzh[NSW6_KALL], rdgz[NSW6_KALL], rdgzh[NSW6_KALL];
    double afact[NSW6_KALL], bfact[NSW6_KALL], cfact[NSW6_KALL], dfact[NSW6_KALL];
} nsw6_grid;

/* The 49 arrays of mp_driver, in reference order.  RW = read and written in place. */
typedef struct nsw6_fields {
    double *rhog, *rhogvx, *rhogvy, *rhogvz, *rhogw, *rhoge;   /* RW [l][k][ij]        */
    double *rhogq;                                              /* RW [l][nq][k][ij]    */
    double *vx, *vy, *vz, *w, *unccn;                           /* R  [l][k][ij]        */
    double *rho, *tem, *pre;                                    /* RW [l][k][ij]        */
    double *q;                                                  /* RW [l][nq][k][ij]    */
    double *qd;                                                 /* W  [l][k][ij]        */
    double *precip, *ISO1_precip, *ISO2_precip;                 /* W  [l][2][ij]        */
    double *precip_rhoe, *precip_lh_heat, *precip_rhophi, *precip_rhokin; /* W [l][ij] */
    double *re_liquid, *re_solid, *re_cld;                      /* W  [l][k][ij]        */
    double *rctop, *rwtop, *tctop;                              /* W  [l][1][ij]        */
    double *frhoge_af, *frhogqv_af, *frhoge_rad, *rhogqke;      /* unused [l][k][ij]    */
    double *gsgam2, *gsgam2h, *gam2, *gam2h;                    /* R  [l][k][ij]        */
    double *ix, *iy, *iz, *jx, *jy, *jz;                        /* R  [l][ij]           */
    double *z, *zh;                                             /* R  [l][k][ij]        */
    double *GDCLW, *GDCFRC, *GPREC, *CBMFX;                     /* W (CBMFX unused)     */
} nsw6_fields;

/* Device-side error report of the last completed step (all regions). */
typedef struct nsw6_report {
    int32_t status;          /* enum nsw6_status                                   */
    int32_t n_not_converged; /* Newton points that hit itelim                      */
    int32_t n_nan;           /* Newton points that produced a NaN temperature      */
    int32_t first_l, first_k, first_ij;  /* first failing point (-1 if none)        */
    int32_t n_general_regions; /* regions that took the exact general sedimentation */
                               /* path (departure beyond one cell / upward / exit)  */
} nsw6_report;

typedef struct nsw6_ctx nsw6_ctx;

/* Fill *p with the mp_nsw6_init defaults (Roh & Satoh 2014 + KK2000, 3WATER,
 * EIN_TYPE=SIMPLE, Nc_def = Nc_ocn = 50). */
void nsw6_params_default(nsw6_params *p);

/* GRD_Setup (src/mod_debug.cpp:92-187): derive the metric factors from gz and gzh. */
void nsw6_grid_setup(const double *gz, const double *gzh, nsw6_grid *g);

/* Create a context on the current HIP device.  `grid` holds gz/gzh (the rest is
 * derived), `params` may be NULL for the defaults.  `max_columns` = the largest
 * lall*ijdim the context will be called with; workspace is allocated once here. */
int nsw6_init(const nsw6_grid *grid, const nsw6_params *params, int64_t max_columns,
              nsw6_ctx **out);

/* Replace the scheme coefficients of an existing context (e.g. perturbed tunables). */
int nsw6_set_params(nsw6_ctx *ctx, const nsw6_params *params);

/* Choose the arithmetic of subsequent steps (enum nsw6_precision). */
int nsw6_set_precision(nsw6_ctx *ctx, int precision);

/* MP_DIV_NUM (include/problemsize.h:175, a compile-time 1 in the reference): each
 * later mp_driver call runs `mp_div_num` NSW6 sub-steps of dt_mp = dt * (1/mp_div_num)
 * and returns precip, precip_rhoe/lh_heat/rhophi/rhokin, GPREC, GDCLW, GDCFRC and
 * the ISO precipitations as the sub-step sums times 1/mp_div_num
 * (src/mod_mp_driver.cpp:420-612); the other diagnostics are the last sub-step's.
 * 1 <= mp_div_num <= NSW6_MAX_DIV_NUM.  The first call with mp_div_num > 1 allocates
 * the sums ((96 + 6) doubles per column).  nsw6_mp_driver_adj then differentiates
 * the whole sub-cycled call (see there). */
#define NSW6_MAX_DIV_NUM 1000
int nsw6_set_div_num(nsw6_ctx *ctx, int mp_div_num);

/* One mp_driver step over `lall` regions of `ijdim` columns each.  `fields` holds
 * device pointers.  `stream` is a hipStream_t (NULL = the default stream). */
int nsw6_mp_driver_fields(nsw6_ctx *ctx, int l_region, int lall, int ijdim,
                          const nsw6_fields *fields, double dt, double ct, void *stream);

/* Same call with the reference's flat argument list (include/mod_mp_driver.h:71-124),
 * preceded by the context and the batch shape and followed by the stream. */
int nsw6_mp_driver(nsw6_ctx *ctx, int lall, int ijdim, int l_region,
    double *rhog, double *rhogvx, double *rhogvy, double *rhogvz, double *rhogw,
    double *rhoge, double *rhogq, double *vx, double *vy, double *vz, double *w,
    double *unccn, double *rho, double *tem, double *pre, double *q, double *qd,
    double *precip, double *ISO1_precip, double *ISO2_precip, double *precip_rhoe,
    double *precip_lh_heat, double *precip_rhophi, double *precip_rhokin,
    double *re_liquid, double *re_solid, double *re_cld, double *rctop, double *rwtop,
    double *tctop, double *frhoge_af, double *frhogqv_af, double *frhoge_rad,
    double *rhogqke, double *gsgam2, double *gsgam2h, double *gam2, double *gam2h,
    double *ix, double *iy, double *iz, double *jx, double *jy, double *jz,
    double *z, double *zh, double dt, double ct, double *GDCLW, double *GDCFRC,
    double *GPREC, double *CBMFX, void *stream);

/* No counterpart in the reference: the same step as nsw6_mp_driver_fields for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or other
 * allocating routine: the step's workspace (post-rate state, fall speeds, wall fluxes,
 * partial extremes, kcell bounds, region flags, the device report and the MP_DIV_NUM
 * sums) lives in one device array of the library per GPU, sized for
 * NSW6_NOALLOC_MAX_COLUMNS columns in at most NSW6_NOALLOC_MAX_REGIONS regions (about
 * 330 MB), which each GPU the library runs on holds from nsw6_init until the process
 * ends.  Larger calls return NSW6_BAD_ARG.  `ctx` (any context of the device, from
 * nsw6_init) supplies the coefficients, grid, precision and MP_DIV_NUM; kernel timing
 * is not recorded.  Every call on a GPU shares that GPU's array: calls that overlap on
 * one GPU (other streams, other contexts, other threads) race and must be ordered by
 * the caller.  nsw6_last_status and nsw6_kcell_bounds read the report of the last
 * call of `ctx`, which a later call on the same GPU overwrites. */
#define NSW6_NOALLOC_MAX_COLUMNS 16900
#define NSW6_NOALLOC_MAX_REGIONS 64
int nsw6_mp_driver_noalloc(nsw6_ctx *ctx, int l_region, int lall, int ijdim,
                           const nsw6_fields *fields, double dt, double ct, void *stream);

/* Reverse mode of one FP64 step (Roh & Satoh snow moments, or the exponential ones with
 * sw_roh2014 = 0; KK2000, or Berry's autoconversion with sw_kk2000 = 0; explicit ice
 * nucleation, sw_expice = 1, or without it, sw_expice = 0: the Bergeron growth and the
 * mixed-phase saturation adjustment; KIN2EIN or KIN2KIN; 3WATER, or 4WATER with QI
 * sedimenting at Vti (or CONST_Vti under sw_constVti); each with or without OPT_INDIR,
 * which then reads `state->unccn`): given the step's INPUT state `state`
 * (not modified; the diagnostics a step would write are kept in context scratch)
 * and cotangents of the step's outputs, write the cotangents of its inputs,
 *     grad = (d outputs / d inputs)^T cot,   grad_coef = (d outputs / d coef)^T cot.
 * Output cotangents `cot` (NULL members = 0): rhog, rhogvx, rhogvy, rhogvz, rhogw,
 *   rhoge, rhogq, rho, tem, pre, q, qd (after the step), precip, precip_rhoe,
 *   precip_lh_heat, precip_rhophi, precip_rhokin, GPREC; plus `cot_sflux`
 *   [l][NT][ij] (NULL = 0) for the surface mass fluxes sflux_t = -f_t(kmin)/dt of the
 *   NT sedimenting tracers: QR, QS, QG (3WATER) or QR, QI, QS, QG (4WATER).  The other
 *   diagnostics carry no cotangent.
 * Input cotangents `grad` (NULL members are not written): rhog, rhogvx, rhogvy,
 *   rhogvz, rhogw, rhoge (nonzero on the two halo levels only: the interior internal
 *   energy is recomputed from tem), rhogq, vx, vy, vz, tem, q (QV only; other
 *   species get 0), unccn (under OPT_INDIR through the point's Nc: Nc^-1.79 of KK2000, Nc and
 *   Dc(Nc) of Berry; 0 otherwise).  `grad_coef` (device, NSW6_NCOEF doubles, may be NULL): the
 *   cotangents of the tunable coefficients in NSW6_NCOEF order.
 * Discrete choices (switches, limiters, kcell bounds, departure cells) are taken at
 * the input state, and the saturation adjustment is differentiated at its converged
 * solution (implicit-function theorem).  Regions whose departures leave the next cell
 * (the general wall fluxes) are differentiated like the others; nsw6_last_status()
 * counts them in n_general_regions.
 * Any other scheme (a switch other than 0 or 1) returns NSW6_UNSUPPORTED at once
 * (nothing is launched).
 * With mp_div_num = n > 1 (nsw6_set_div_num) the call is differentiated as the
 * forward call runs it: n sub-steps of dt/n, precip, precip_rhoe/lh_heat/rhophi/
 * rhokin, GPREC and the surface mass fluxes averaged over them (their cotangents reach
 * every sub-step times 1/n), the state after the last sub-step.  The sub-step inputs
 * are rebuilt by n-1 forward sub-steps into context scratch of 21*96 doubles per
 * column for each of them, plus 2*21*96 + 7*96 + 10 doubles per column, a second step
 * workspace of about 24*96 doubles per column, and the region flags (grown on demand).
 * The last sub-step's forward part goes to the second workspace, so that sub-step
 * n-2 reuses the forward part its rebuild left in the first.  A region that takes the
 * general path in any sub-step is counted as above.  Timing ids 6/7 then cover the
 * forward rebuild and the reversed sub-steps. */
int nsw6_mp_driver_adj(nsw6_ctx *ctx, int l_region, int lall, int ijdim,
                       const nsw6_fields *state, const nsw6_fields *cot,
                       const double *cot_sflux, const nsw6_fields *grad,
                       double *grad_coef, double dt, double ct, void *stream);

/* Synchronise the context's last stream and return the device report of the last
 * step (or of the forward part recomputed by the last nsw6_mp_driver_adj).  With
 * sub-steps the Newton counts cover all of them; n_general_regions counts the last. */
int nsw6_last_status(nsw6_ctx *ctx, nsw6_report *report);

/* kcell_min / kcell_max of vadv1d_prep (src/mod_vadv1d.cpp:317-352), i.e. the
 * per-region departure-cell bounds that the last nsw6_mp_driver* step (its last
 * sub-step) or the forward part of the last nsw6_mp_driver_adj computed on the
 * device, for regions 0..lall-1: out[l][NT][2 (full level, half level)][NSW6_KALL]
 * [2 (kcell_min, kcell_max)], NT = 3 ("3WATER": QR, QS, QG) or 4 ("4WATER": QR, QI,
 * QS, QG), 0-based levels.  The reference keeps them in host Views of
 * precip_transport_new; here they are a read-out for checks (synchronises the last
 * stream).  n_out = elements available at `out`. */
int nsw6_kcell_bounds(nsw6_ctx *ctx, int lall, int32_t *out, int64_t n_out);

/* Average duration (ms) of the named kernel over the launches recorded since the
 * last reset, measured with HIP events on the stream the kernel ran on.
 * which: 0 = whole step, 1 = process rates, 2 = fall-speed extremes + kcell bounds,
 * 3 = sedimentation fluxes of the single-cell regions, 4 = flux divergence + KIN2EIN +
 * saturation adjustment tail, 8 = the wall fluxes of the general regions (an empty
 * launch when there are none); for nsw6_mp_driver_adj: 5 = whole call, 6 = its forward
 * part, 7 = the reverse passes.
 * Recording is off until nsw6_timing_enable(ctx, 1). */
int nsw6_timing_enable(nsw6_ctx *ctx, int on);
int nsw6_timing_read(nsw6_ctx *ctx, int which, double *avg_ms, int64_t *count);
int nsw6_timing_reset(nsw6_ctx *ctx);

void nsw6_finalize(nsw6_ctx *ctx);

/* NICAM snapshot files of the reference's stand-alone run, on HOST arrays in the
 * nsw6_fields layout.  `dir` is a run directory holding data/<name>.dat (the inputs
 * src/main.cpp:128-174 reads), data/vgrid/GRD_gz.dat and GRD_gzh.dat
 * (src/mod_debug.cpp:189-195) and ref_verify/calculated_<name>_DP.dat
 * (src/main.cpp:326-337).  Raw native doubles, row-major in the reader's shapes
 * (src/data_io.cpp:207-402): [l][96][ij] fields, [l][6][96][ij] rhogq_Lswp/q_Lswp,
 * [2][l][1][ij] precip_mp/precip1_mp/precip2_mp (precip, ISO1_precip, ISO2_precip),
 * [l][1][ij] rhoein/lh/rhophi/rhokin_precip_mp, [l][ij] ix..jz; qke is rhogqke.
 * Every file must hold exactly its shape's bytes.  NULL members of `fields` (and
 * NULL gz/gzh) are skipped.  Returns NSW6_OK, NSW6_BAD_ARG or NSW6_IO_ERROR. */
int nsw6_snapshot_read(const char *dir, int lall, int ijdim, const nsw6_fields *fields,
                       double *gz, double *gzh);
/* Write the same files (creating data/, data/vgrid/ under an existing `dir`). */
int nsw6_snapshot_write(const char *dir, int lall, int ijdim, const nsw6_fields *fields,
                        const double *gz, const double *gzh);
/* The verification set: rhog, rhogvx, rhogvy, rhogvz, rhogw, rhoge and rhogq of
 * QV..QG (calculated_QV1_DP.dat .. calculated_QG6_DP.dat), each [l][96][ij].
 * `check` receives NSW6_VERIFY_FIELDS such arrays back to back. */
#define NSW6_VERIFY_FIELDS 12
int nsw6_verify_read(const char *dir, int lall, int ijdim, double *check);
/* Write the verification set of a state (creating ref_verify/ under `dir`). */
int nsw6_verify_write(const char *dir, int lall, int ijdim, const nsw6_fields *fields);

const char *nsw6_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif /* NSW6_H */
