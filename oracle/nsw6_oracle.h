/*
 * nsw6_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the reference NSW6 microphysics step, used as the
 * correctness oracle by tests/, by __graft_entry__.smoke() and as the CPU baseline
 * of bench.py.  The product library never links, loads or calls anything here.
 *
 * Every function follows the reference source phase by phase (file:line cited at
 * each definition in nsw6_oracle.c).  Arrays use the public layout of nsw6.h.
 */
#ifndef NSW6_ORACLE_H
#define NSW6_ORACLE_H

#include "../include/nsw6.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Outputs of one TagBigLoop point (src/mod_mp_nsw6.cpp:760-1594). */
typedef struct ora_bigloop_out {
    double drhogq[6];     /* drhogqv..drhogqg, tracer order QV,QC,QR,QI,QS,QG */
    double Vt[6];         /* Vt(QR), Vt(QI), Vt(QS), Vt(QG); QV, QC left 0     */
    double rceff, rceff_cld;
    double rctop, tctop, rctop_cld, tctop_cld;
    double cloud_sw;      /* [qc+qr > q_min]*(1-zerosw): the cloud switch of the
                             comment at src/mod_mp_nsw6.cpp:1577 (cloudtop_intended) */
    double wk[49];        /* final process-rate work array                    */
    double q_t[6];        /* final tendencies qv_t..qg_t                       */
} ora_bigloop_out;

double ora_gammafunc(double x);
void   ora_params_init(nsw6_params *p);
void   ora_grid_setup(const double *gz, const double *gzh, nsw6_grid *g);
double ora_psat_liq(double tem);
double ora_psat_ice(double tem);
void   ora_bergeron_point(double tem, double *a1, double *a2, double *ma2);
void   ora_bigloop_point(const nsw6_params *P, double dt, double dens, double temp,
                         const double q6[6], double qsatl, double qsati, double pre,
                         double rhog, double Nc, double a1, double a2, double ma2,
                         ora_bigloop_out *o);

/* vadv1d on one region, arrays [k][ij] (src/mod_vadv1d.cpp:231-450). */
void ora_vadv1d_prep(int ijdim, int mkmin, int mkmax, const double *dz, const double *zh,
                     const double *wp, double *zdis, int *kcell, int *kcell_max,
                     int *kcell_min, double dt);
void ora_vadv1d_getflux(int ijdim, int mkmin, int mkmax, const double *dz,
                        const double *rhof, const double *zdis0, const int *kcell,
                        const int *kcell_max, const int *kcell_min, double *frhof);

/* satadjust_liq Newton on a single point: returns iterations used (100 = not
 * converged), updates *tem, *qv, *qc (src/mod_satadjust.cpp:828-894). */
int ora_satadjust_liq_point(double rho, double Emoist, double qsum, double qd_loc,
                            const double q6[6], double *tem, double *qv, double *qc);
/* One point of satadjust_all (mixed-phase Newton, src/mod_satadjust.cpp:620-769);
 * returns 0 / iterations / -1 NaN / -2 itelim. */
int ora_satadjust_all_point(double rho, double Emoist, double qsum, double qd_loc, const double q6[6],
                            double UL, double LL, double *tem, double *qv, double *qc, double *qi);

/* mp_driver over lall regions of host arrays (src/mod_mp_driver.cpp:335-613).
 * `ice_adjust` < 0 keeps the reference choice (satadjust_all exactly when
 * sw_expice != 1, i.e. without explicit ice generation); 0/1 force
 * satadjust_liq/satadjust_all.  Returns NSW6_OK or a failure status
 * and fills *rep (may be NULL). */
int ora_mp_driver(const nsw6_params *P, const nsw6_grid *G, int lall, int ijdim,
                  const nsw6_fields *f, double dt, double ct, int ice_adjust,
                  nsw6_report *rep);

/* ora_mp_driver with the reference's MP_DIV_NUM sub-cycling (include/problemsize.h:175,
 * src/mod_mp_driver.cpp:420-612): div_num >= 1 steps of dt / div_num; the surface
 * outputs, GPREC, GDCLW and GDCFRC are the sub-step sums times 1 / div_num. */
int ora_mp_driver_div(const nsw6_params *P, const nsw6_grid *G, int lall, int ijdim,
                      const nsw6_fields *f, double dt, double ct, int ice_adjust, int div_num,
                      nsw6_report *rep);

/* Test hook: while `buf` is not NULL, every ora_mp_driver* call records the kcell
 * bounds of its vadv1d_prep calls (src/mod_vadv1d.cpp:317-352) as
 * buf[l][NT][2 (full level, half level)][NSW6_KALL][2 (kcell_min, kcell_max)], NT = 3
 * ("3WATER") or 4 ("4WATER"); with sub-steps the last one's are kept. */
void ora_record_kcell_bounds(int *buf);

/* Test hook: while `buf` is not NULL, every ora_mp_driver* call records for each
 * point of the saturation adjustment's Newton (src/mod_satadjust.cpp:828-894,
 * 700-769) buf[l][k][ij] = the stopping margin min(| |dtemp_n| - c | / c,
 * | |dtemp_n-1| - c | / c) over its last two steps, c = dtemp_criteria = 1e-8 K;
 * 1e300 where no iteration ran or it did not converge.  A small margin marks a point
 * where a last-bit change of the inputs can change the iteration count (and so
 * q(QV) = qsat of the previous iterate by dqsat/dT * |dtemp|).  Interior levels
 * only; the caller initialises the halo levels. */
void ora_record_newton_margin(double *buf);

/* Single phases of the step on one region ([k][ij] arrays, tracers [6][k][ij]), for
 * the pins of the tests: THRMDYN_qd/cv/tempre (src/mod_thrmdyn.cpp:98-180);
 * negative_filter (src/mod_mp_nsw6.cpp:1750-1832, in place); cnvvar_rhogkin_in
 * (src/mod_debug.cpp:269-330) with precip_transport_new's C2W/W2C factors;
 * precip_transport_new (src/mod_precip_transport.cpp:573-1087) on region 0 of `f` in
 * place, Vterm [6][k][ij]; SATURATION_adjustment (src/mod_satadjust.cpp:274-362) in
 * place, with the scheme's range. */
void ora_thrmdyn(int ijdim, const double *q, const double *ein, const double *rho, double *qd, double *cv,
                 double *tem, double *pre);
void ora_negative_filter(int ijdim, double *rhog, double *rhoge, double *rhogq, double *rho, const double *tem,
                         double *pre, double *q, const double *gsgam2);
void ora_rhogkin(const nsw6_grid *G, int ijdim, const double *rhog, const double *rhogvx, const double *rhogvy,
                 const double *rhogvz, const double *rhogw, const double *gsgam2, const double *gsgam2h,
                 double *rhogkin, double *rhogkin_h, double *rhogkin_v);
void ora_precip_transport(const nsw6_params *P, const nsw6_grid *G, int ijdim, const nsw6_fields *f,
                          const double *Vterm, double dt);
void ora_saturation_adjustment(const nsw6_params *P, int ijdim, double *rhog, double *rhoge, double *rhogq,
                               double *tem, double *q, double *qd, double *gsgam2);

/* Seconds of the last ora_mp_driver call (wall clock, like the reference's
 * "NSW6 Elapsed Time", src/mod_mp_driver.cpp:500-552). */
double ora_last_elapsed(void);
int    ora_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
