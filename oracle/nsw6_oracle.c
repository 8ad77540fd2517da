/*
 * nsw6_oracle.c -- TEST INFRASTRUCTURE ONLY (the correctness oracle and the CPU
 * baseline).  Nothing in the product library links or calls this file.
 *
 * A plain-C, phase-by-phase restatement of the reference NICAM NSW6 microphysics
 * step as found in the Kokkos source (paths relative to the reference repository).
 * Every reference pass over a full (k,ij) array is kept as a pass here, with the
 * same loop bounds (interior k in [1,94] vs all k), the same operation order and
 * every port quirk of the source (see the comments marked QUIRK).  Work arrays are
 * zero-initialised like freshly constructed Kokkos Views.  OpenMP parallelises
 * each pass over its (k,ij) range; reductions that the reference does serially on
 * the host (vadv1d_prep) are done serially here too.
 *
 * Build: gcc -O2 -fno-fast-math -ffp-contract=off -fopenmp (oracle/Makefile).
 *
 * Every function is pinned by checks independent of this text (tests/pin_checks.py,
 * tests/test_oracle_*.py): closed forms, quadratures of the physical integrals,
 * budgets, the reference's tables, vgrid files and parameters; each pin catches a
 * mistake planted under ORA_PLANT.  Parity unpinned: the last bits of exp, log and pow
 * (the reference's Kokkos math is absent; the parity bar absorbs them) and the reading
 * of the racy cloud-top diagnostics in the cloudtop_intended mode.
 */
#define _POSIX_C_SOURCE 199309L   /* clock_gettime under -std=c11 */
#include "nsw6_oracle.h"

#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants: include/problemsize.h:74-137, src/problemsize.cpp ---------- */
#define KALL   96
#define KMIN   1
#define KMAX   94
#define NQMAX  6
#define I_QV 0
#define I_QC 1
#define I_QR 2
#define I_QI 3
#define I_QS 4
#define I_QG 5
#define NQW_STR 1
#define NQW_END 5

/* ORA_PLANT = n (n >= 1) builds the oracle with mistake n planted, for
 * tests/test_oracle_plants.py, which checks that the pin of each function catches its
 * mistake.  The oracle proper is built with ORA_PLANT = 0. */
#ifndef ORA_PLANT
#define ORA_PLANT 0
#endif

static const double CONST_PI    = 3.141592653589793;
static const double CONST_EPS   = 2.220446E-16;
static const double CONST_UNDEF = -9.9999E+30;
static const double CONST_GRAV  = 9.79764;
static const double CONST_Rdry  = 287.04;
static const double CONST_CPdry = 1004.64;
static const double CONST_CVdry = 717.56;
static const double CONST_Rvap  = 461.46;
static const double CONST_CL    = 4218.0;
static const double CONST_PSAT0 = 610.7;
static const double CONST_LHV0  = 2501000.0;
static const double CONST_LHS0  = 2834000.0;
static const double CONST_LHF0  = 333000.0;
static const double CONST_Pstd  = 101325.0;
static const double CONST_TEM00 = 273.15;
static const double CONST_DWATR = 1000.0;
#define LHV CONST_LHV0
#define LHF CONST_LHF0
#define RP_PREC 15

/* EIN_TYPE = "SIMPLE" (src/problemsize.cpp:30): CVW(:) = CVdry, CPW(0) = CPdry,
 * CPW(1:5) = CVdry (src/mod_debug.cpp:76-90); CPovR = CVovR = 0
 * (src/mod_satadjust.cpp:37-43); LovR = L/Rvap (:45-46). */
static const double CVW[NQMAX] = {717.56, 717.56, 717.56, 717.56, 717.56, 717.56};
static const double CPovR_liq = 0.0, CPovR_ice = 0.0, CVovR_liq = 0.0, CVovR_ice = 0.0;
#define LovR_liq (CONST_LHV0 / CONST_Rvap)
#define LovR_ice (CONST_LHS0 / CONST_Rvap)

/* mod_mp_nsw6.cpp private constants (:8, :117-125) */
static const double dens00 = 1.28;
static const double Nc_ocn = 50.0;
static const double Ka0 = 2.428E-2, dKa_dT = 7.47E-5;
static const double Kd0 = 2.222E-5, dKd_dT = 1.37E-7;
static const double nu0 = 1.718E-5, dnu_dT = 5.28E-8;
/* satadjust private constants (src/mod_satadjust.cpp:6-7) */
static const double TEM_MIN = 10.0;

static double g_elapsed = 0.0;
/* Test hook (ora_record_kcell_bounds): when set, precip_transport_new records the
 * kcell_min/kcell_max of each of its vadv1d_prep calls into
 * g_bounds_rec[l][NT][2 (full, half level)][KALL][2 (min, max)], NT = the number of
 * sedimenting tracers, for region g_bounds_l. */
static int *g_bounds_rec = NULL;
static int g_bounds_l = 0;
/* Test hook (ora_record_newton_margin): when set, saturation_adjustment records per
 * point g_newton_rec[l][k][ij] the relative distance of the Newton's stopping test
 * from its criterion over the last two iterations (see ora_record_newton_margin). */
static double *g_newton_rec = NULL;
/* the Newton's dtemp_criteria of the current step (nsw6_params.satadjust_dtemp_criteria;
 * the reference's 10^(-(RP_PREC+1)/2) = 1e-8, src/mod_satadjust.cpp:788) */
static double g_dtemp_criteria = 1.0e-8;

/* Stopping margin of a Newton that ran `n` iterations, the last two steps being
 * d_prev (only when n >= 2) and d_last: min over those of | |d| - crit | / crit. */
static double newton_margin(int n, int converged, double d_prev, double d_last, double crit)
{
    if (n <= 0 || !converged) return 1e300;
    double m = fabs(fabs(d_last) - crit) / crit;
    if (n >= 2) {
        double mp = fabs(fabs(d_prev) - crit) / crit;
        if (mp < m) m = mp;
    }
    return m;
}

#define IDX(k, ij)      ((size_t)(k) * (size_t)ijdim + (size_t)(ij))
#define QIDX(nq, k, ij) (((size_t)(nq) * KALL + (size_t)(k)) * (size_t)ijdim + (size_t)(ij))

static inline double dmax(double a, double b) { return (a < b) ? b : a; }
static inline double dmin(double a, double b) { return (b < a) ? b : a; }

static double *zalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }

/* ---- MISC_gammafunc: src/mod_debug.cpp:332-361 ------------------------------ */
double ora_gammafunc(double xx)
{
    static const double coef[6] = {+76.18009172947146, -86.50532032941677,
                                   +24.01409824083091, -1.231739572450155,
                                   +0.1208650973866179E-2, -0.5395239384953E-5};
    double x = xx, y = x;
    double tmp = x + 5.5;
    tmp = tmp - (x + 0.5) * log(tmp);
    double ser = 1.000000000190015;
    for (int j = 0; j < 6; j++) {
        y += 1;
        ser += coef[j] / y;
    }
    return exp(-tmp + log(2.5066282746310005 * ser / x));
}

/* ---- GlobalParams defaults + mp_nsw6_init: include/mod_mp_nsw6.h:16-123,
 *      src/mod_mp_nsw6.cpp:276-434; effective radius constants :506-517 ---------- */
void ora_params_init(nsw6_params *p)
{
    memset(p, 0, sizeof(*p));
    p->N0r = 8.E+6; p->N0s = 3.E+6; p->N0g = 4.E+6;
    p->rho_s = 100.0; p->rho_g = 400.0; p->C_d = 0.6; p->Cr = 130.0; p->Cs = 4.84;
    p->sw_kk2000 = 0.0; p->sw_roh2014 = 0.0;
    {
        const double a[10] = {5.065339, -0.062659, -3.032362, 0.029469, -0.000285,
                              0.31255, 0.000204, 0.003199, 0.0, -0.015952};
        const double b[10] = {0.476221, -0.015896, 0.165977, 0.007468, -0.000141,
                              0.060366, 0.000079, 0.000594, 0.0, -0.003577};
        memcpy(p->coef_a, a, sizeof a);
        memcpy(p->coef_b, b, sizeof b);
    }
    p->Eiw = 1.0; p->Erw = 1.0; p->Esw = 1.0; p->Egw = 1.0; p->Eri = 1.0;
    p->Esi = 1.0; p->Egi = 0.1; p->Esr = 1.0; p->Egr = 1.0; p->Egs = 1.0;
    p->gamma_sacr = 25.E-3; p->gamma_gacs = 90.E-3; p->mi = 4.19E-13;
    p->beta_saut = 1.E-3; p->gamma_saut = 25.E-3; p->beta_gaut = 1.E-3;
    p->gamma_gaut = 90.E-3; p->qicrt_saut = 0.0; p->qscrt_gaut = 6.E-4;
    p->f1r = 0.78; p->f2r = 0.27; p->f1s = 0.65; p->f2s = 0.39; p->f1g = 0.78; p->f2g = 0.27;
    p->A_frz = 0.66; p->B_frz = 100.0;
    p->mi40 = 2.46E-10; p->mi50 = 4.80E-10; p->vti50 = 1.0; p->Ri50 = 5.E-5;
    p->sw_expice = 0.0; p->Nc_ihtr = 300.0; p->Di_max = 500.E-6; p->Di_a = 11.9;
    p->sw_constVti = 0.0;

    /* mp_nsw6_init */
    p->Nc_def = Nc_ocn;
    p->CONST_Vti = CONST_UNDEF;
    p->gamma_sacr = 6.0E-2;
    p->gamma_saut = 6.0E-2;
    p->beta_saut = 5.0E-3;           /* PSAUT_BETA0 */
    p->Ar = CONST_PI * CONST_DWATR / 6.0;
    p->As = CONST_PI * p->rho_s / 6.0;
    p->Ag = CONST_PI * p->rho_g / 6.0;
    p->Br = 3.0; p->Bs = 3.0; p->Bg = 3.0;
    p->Cg = sqrt((4.0 * p->rho_g * CONST_GRAV) / (3.0 * dens00 * p->C_d));
    p->Dr = 0.50; p->Ds = 0.25; p->Dg = 0.50;
    /* Roh_flag = true */
    p->sw_roh2014 = 1.0;
    p->N0g = 4.0E+8; p->As = 0.069; p->Bs = 2.0;
    p->Esi = 0.25; p->Egi = 0.0; p->Egs = 0.0;
    p->sw_expice = 1.0;              /* OPT_EXPLICIT_ICEGEN */
    p->sw_kk2000 = 1.0;              /* qr_aut_acc_type = KK2000 */
    p->beta_gaut = 0.0;              /* INC_PGAUT = false */
    p->sw_constVti = 0.0;            /* CONST_Vti == UNDEF */
    p->GAM = 1.0; p->GAM_2 = 1.0; p->GAM_3 = 2.0;
    p->GAM_1br = ora_gammafunc(1.0 + p->Br);
    p->GAM_2br = ora_gammafunc(2.0 + p->Br);
    p->GAM_3br = ora_gammafunc(3.0 + p->Br);
    p->GAM_3dr = ora_gammafunc(3.0 + p->Dr);
    p->GAM_6dr = ora_gammafunc(6.0 + p->Dr);
    p->GAM_1brdr = ora_gammafunc(1.0 + p->Br + p->Dr);
    p->GAM_5dr_h = ora_gammafunc(0.5 * (5.0 + p->Dr));
    p->GAM_1bs = ora_gammafunc(1.0 + p->Bs);
    p->GAM_2bs = ora_gammafunc(2.0 + p->Bs);
    p->GAM_3bs = ora_gammafunc(3.0 + p->Bs);
    p->GAM_3ds = ora_gammafunc(3.0 + p->Ds);
    p->GAM_1bsds = ora_gammafunc(1.0 + p->Bs + p->Ds);
    p->GAM_5ds_h = ora_gammafunc(0.5 * (5.0 + p->Ds));
    p->GAM_1bg = ora_gammafunc(1.0 + p->Bg);
    p->GAM_3dg = ora_gammafunc(3.0 + p->Dg);
    p->GAM_1bgdg = ora_gammafunc(1.0 + p->Bg + p->Dg);
    p->GAM_5dg_h = ora_gammafunc(0.5 * (5.0 + p->Dg));
    p->ln10 = log(10.0);
    /* literals of the rate code exposed as tunables */
    p->kk_aut_rate = 1350.0;
    p->icenuc_coef = 4.92E-11;
    p->satadjust_dtemp_criteria = pow(10.0, (double)(-(RP_PREC + 1) / 2));   /* = 1e-8 */
    /* mp_nsw6_new::init effective-radius work (src/mod_mp_nsw6.cpp:511-516) */
    {
        double Dc = 0.146 - 5.964E-2 * log(p->Nc_def / 2000.0);
        p->dgamma_a = 0.1444 / (Dc * Dc);
        p->GAM_dgam = ora_gammafunc(p->dgamma_a);
        p->GAM_dgam23 = ora_gammafunc(p->dgamma_a + 2.0 / 3.0);
        p->coef_dgam = p->GAM_dgam23 / p->GAM_dgam * pow(p->dgamma_a, (-2.0 / 3.0));
        p->coef_xf = 3.0 / 4.0 / CONST_PI / CONST_DWATR;
    }
}

/* ---- GRD_Setup: src/mod_debug.cpp:92-187 ------------------------------------ */
void ora_grid_setup(const double *gz, const double *gzh, nsw6_grid *g)
{
    memset(g, 0, sizeof(*g));
    for (int k = 0; k < KALL; k++) { g->gz[k] = gz[k]; g->gzh[k] = gzh[k]; }
    for (int k = KMIN - 1; k <= KMAX; k++) g->dgz[k] = g->gzh[k + 1] - g->gzh[k];
    g->dgz[KMAX + 1] = g->dgz[KMAX];
    for (int k = KMIN; k <= KMAX + 1; k++) g->dgzh[k] = g->gz[k] - g->gz[k - 1];
    g->dgzh[KMIN - 1] = g->dgzh[KMIN];
    for (int k = 0; k < KALL; k++) {
        g->rdgz[k] = 1.0 / g->dgz[k];
        g->rdgzh[k] = 1.0 / g->dgzh[k];
    }
    for (int k = KMIN; k <= KMAX + 1; k++)
        g->afact[k] = (g->gzh[k] - g->gz[k - 1]) / (g->gz[k] - g->gz[k - 1]);
    g->afact[KMIN - 1] = 1.0;
    for (int k = 0; k < KALL; k++) g->bfact[k] = 1.0 - g->afact[k];
    for (int k = KMIN; k <= KMAX; k++)
        g->cfact[k] = (g->gz[k] - g->gzh[k]) / (g->gzh[k + 1] - g->gzh[k]);
    g->cfact[KMIN - 1] = 1.0;
    g->cfact[KMAX + 1] = 0.0;
    for (int k = 0; k < KALL; k++) g->dfact[k] = 1.0 - g->cfact[k];
}

/* ---- SATURATION_psat_liq / _ice: src/mod_satadjust.cpp:80-106, 133-160 ------- */
double ora_psat_liq(double tem)
{
    const double RTEM00 = 1.0 / CONST_TEM00;
    double rtem = 1.0 / (dmax(tem, TEM_MIN));
    return CONST_PSAT0 * pow((tem * RTEM00), CPovR_liq) * exp(LovR_liq * (RTEM00 - rtem));
}
double ora_psat_ice(double tem)
{
    const double RTEM00 = 1.0 / CONST_TEM00;
    double rtem = 1.0 / (dmax(tem, TEM_MIN));
    return CONST_PSAT0 * pow((tem * RTEM00), CPovR_ice) * exp(LovR_ice * (RTEM00 - rtem));
}

/* ---- THRMDYN: src/mod_thrmdyn.cpp:98-180 (all k) ----------------------------- */
static void thrmdyn_qd(int ijdim, const double *q, double *qd)
{
#pragma omp parallel for collapse(2)
    for (int k = 0; k < KALL; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            double v = 1.0;
            for (int nq = (ORA_PLANT == 4 ? 0 : NQW_STR); nq <= NQW_END; nq++) v = v - q[QIDX(nq, k, ij)];
            qd[IDX(k, ij)] = v;
        }
}
static void thrmdyn_cv(int ijdim, const double *qd, const double *q, double *cv)
{
#pragma omp parallel for collapse(2)
    for (int k = 0; k < KALL; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            double v = qd[IDX(k, ij)] * CONST_CVdry;
            for (int nq = NQW_STR; nq <= NQW_END; nq++) v = v + q[QIDX(nq, k, ij)] * CVW[nq];
            cv[IDX(k, ij)] = v;
        }
}
static void thrmdyn_tempre(int ijdim, const double *ein, const double *rho, const double *q,
                           double *tem, double *pre)
{
#pragma omp parallel for collapse(2)
    for (int k = 0; k < KALL; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            double cv = 0.0, qd = 1.0;
            for (int nq = NQW_STR; nq <= NQW_END; nq++) {
                cv = cv + q[QIDX(nq, k, ij)] * CVW[nq];
                qd = qd - q[QIDX(nq, k, ij)];
            }
            cv = cv + qd * CONST_CVdry;
            tem[IDX(k, ij)] = ein[IDX(k, ij)] / cv;
#if ORA_PLANT == 5
            pre[IDX(k, ij)] = rho[IDX(k, ij)] * tem[IDX(k, ij)] *
                              (qd * CONST_Rvap + q[QIDX(I_QV, k, ij)] * CONST_Rdry);
#else
            pre[IDX(k, ij)] = rho[IDX(k, ij)] * tem[IDX(k, ij)] *
                              (qd * CONST_Rdry + q[QIDX(I_QV, k, ij)] * CONST_Rvap);
#endif
        }
}

/* ---- negative_filter: src/mod_mp_nsw6.cpp:1750-1832 --------------------------- */
static void negative_filter(int ijdim, double *rhog, double *rhoge, double *rhogq, double *rho,
                            const double *tem, double *pre, double *q, const double *gsgam2)
{
    double *qd = zalloc((size_t)KALL * ijdim);
    double *cva = zalloc((size_t)KALL * ijdim);
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            double diffq = 0.0;
            const int nq0 = (ORA_PLANT == 2) ? NQW_STR : NQW_STR + 1;
            for (int nq = nq0; nq <= NQW_END; nq++) {
                diffq += rhogq[QIDX(nq, k, ij)];
                rhogq[QIDX(nq, k, ij)] = dmax(rhogq[QIDX(nq, k, ij)], 0.0);
            }
            for (int nq = nq0; nq <= NQW_END; nq++) diffq -= rhogq[QIDX(nq, k, ij)];
#if ORA_PLANT != 1
            rhogq[QIDX(I_QV, k, ij)] += diffq;
#endif
        }
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            double diffq = rhogq[QIDX(I_QV, k, ij)];
            rhogq[QIDX(I_QV, k, ij)] = dmax(rhogq[QIDX(I_QV, k, ij)], 0.0);
            diffq -= rhogq[QIDX(I_QV, k, ij)];
            rhog[IDX(k, ij)] = rhog[IDX(k, ij)] * (1.0 - (ORA_PLANT == 3 ? -diffq : diffq));
            rho[IDX(k, ij)] = rhog[IDX(k, ij)] / gsgam2[IDX(k, ij)];
        }
    for (int nq = NQW_STR; nq <= NQW_END; nq++) {
#pragma omp parallel for collapse(2)
        for (int k = KMIN; k <= KMAX; k++)
            for (int ij = 0; ij < ijdim; ij++)
                q[QIDX(nq, k, ij)] = rhogq[QIDX(nq, k, ij)] / rhog[IDX(k, ij)];
    }
    thrmdyn_qd(ijdim, q, qd);
    thrmdyn_cv(ijdim, qd, q, cva);
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            rhoge[IDX(k, ij)] = tem[IDX(k, ij)] * rhog[IDX(k, ij)] * cva[IDX(k, ij)];
            pre[IDX(k, ij)] = rho[IDX(k, ij)] *
                              (qd[IDX(k, ij)] * CONST_Rdry + q[QIDX(I_QV, k, ij)] * CONST_Rvap) *
                              tem[IDX(k, ij)];
        }
    free(qd);
    free(cva);
}

/* ---- Bergeron_param: src/mod_mp_nsw6.cpp:1834-1890 ---------------------------- */
static const double a1_tab[32] = {
    0.0001E-7, 0.7939E-7, 0.7841E-6, 0.3369E-5, 0.4336E-5, 0.5285E-5, 0.3728E-5, 0.1852E-5,
    0.2991E-6, 0.4248E-6, 0.7434E-6, 0.1812E-5, 0.4394E-5, 0.9145E-5, 0.1725E-4, 0.3348E-4,
    0.1725E-4, 0.9175E-5, 0.4412E-5, 0.2252E-5, 0.9115E-6, 0.4876E-6, 0.3473E-6, 0.4758E-6,
    0.6306E-6, 0.8573E-6, 0.7868E-6, 0.7192E-6, 0.6513E-6, 0.5956E-6, 0.5333E-6, 0.4834E-6};
static const double a2_tab[32] = {
    0.0100, 0.4006, 0.4831, 0.5320, 0.5307, 0.5319, 0.5249, 0.4888, 0.3849, 0.4047, 0.4318,
    0.4771, 0.5183, 0.5463, 0.5651, 0.5813, 0.5655, 0.5478, 0.5203, 0.4906, 0.4447, 0.4126,
    0.3960, 0.4149, 0.4320, 0.4506, 0.4483, 0.4460, 0.4433, 0.4413, 0.4382, 0.4361};

void ora_bergeron_point(double tem, double *a1, double *a2, double *ma2)
{
    double temc = dmin(dmax(tem - CONST_TEM00, -30.99), 0.0);
    /* int(-temc) of a NaN is undefined (the reference indexes the tables with it); the
     * index is taken from a NaN-free copy, the NaN still reaches a1, a2 through fact */
    int itemc = (int)(-((temc == temc) ? temc : 0.0));
    double fact = -(temc + (double)(itemc - 1));   /* QUIRK: off-by-one kept (A-5) */
    double v1 = (1.0 - fact) * a1_tab[itemc] + (fact)*a1_tab[itemc + 1];
    double v2 = (1.0 - fact) * a2_tab[itemc] + (fact)*a2_tab[itemc + 1];
    double m2 = 1.0 - v2;
    *a1 = v1 * pow(1.0E-3, m2);
    *a2 = v2;
    *ma2 = m2;
}

/* ---- TagBigLoop: src/mod_mp_nsw6.cpp:760-1594 (one point) --------------------- */
enum {
    I_dqv_dt, I_dqc_dt, I_dqr_dt, I_dqi_dt, I_dqs_dt, I_dqg_dt, I_delta1, I_delta2, I_spsati,
    I_iceflg, I_RLMDr, I_RLMDs, I_RLMDg, I_Piacr, I_Psacr, I_Praci, I_Pigen, I_Pidep, I_Psdep,
    I_Pgdep, I_Praut, I_Pracw, I_Pihom, I_Pihtr, I_Psacw, I_Psfw, I_Pgacw, I_Prevp, I_Piacr_s,
    I_Psacr_s, I_Piacr_g, I_Psacr_g, I_Pgacr, I_Pgfrz, I_Pisub, I_Pimlt, I_Psaut, I_Praci_s,
    I_Psaci, I_Psfi, I_Praci_g, I_Pgaci, I_Pssub, I_Psmlt, I_Pgaut, I_Pracs, I_Pgacs, I_Pgsub,
    I_Pgmlt
};

void ora_bigloop_point(const nsw6_params *P, double dt, double dens, double temp,
                       const double q6[6], double qsatl, double qsati, double pre, double rhog,
                       double Nc, double a1, double a2, double ma2, ora_bigloop_out *o)
{
    const double PI = CONST_PI, EPS = CONST_EPS, UNDEF = CONST_UNDEF;
    const double Rvap = CONST_Rvap, LHV0 = CONST_LHV0, LHS0 = CONST_LHS0, LHF0 = CONST_LHF0;
    const double PRE00 = CONST_Pstd, TEM00 = CONST_TEM00, CL = CONST_CL, rho_w = CONST_DWATR;
    const double r2_min = 1.0E-10, q_min = 1.0E-5;
    double wk[49];
    double qv, qc, qr, qi, qs, qg, qv_t, qc_t, qr_t, qi_t, qs_t, qg_t;
    double Sliq, Sice, Rdens, rho_fact, temc;
    double RLMDr, RLMDr_2, RLMDr_3, RLMDs, RLMDs_2, RLMDs_3, RLMDg, RLMDg_2, RLMDg_3;
    double RLMDr_1br, RLMDr_2br, RLMDr_3br, RLMDs_1bs, RLMDs_2bs, RLMDs_3bs;
    double RLMDr_dr, RLMDr_3dr, RLMDr_5dr, RLMDs_ds, RLMDs_3ds, RLMDs_5ds;
    double RLMDg_dg, RLMDg_3dg, RLMDg_5dg, RLMDr_7, RLMDr_6dr;
    double tems, Xs2, MOMs_0, MOMs_1, MOMs_2, MOMs_0bs, MOMs_1bs, MOMs_2bs;
    double MOMs_2ds, MOMs_5ds_h, RMOMs_Vt, coef_at[4], coef_bt[4], loga_, b_, nm;
    double Vti, Vtr, Vts, Vtg, Esi_mod, Egs_mod, rhoqc, Pracw_orig, Pracw_kk;
    double Praut_berry, Praut_kk, Dc, betai, betas, Ka, Kd, Nu, Glv, Giv, Gil;
    double ventr, vents, ventg, net, fac, fac_sw, zerosw, tmp;
    double sw_bergeron, dt1, Ni50, sw, rhoqi, XNi, XMi, Di, Ni0, Qi0;
    double xf_qc, rf_qc, r2_qc, r2_qr, r3_qc, r3_qr;

    const double Ar = P->Ar, As = P->As, Ag = P->Ag, Cr = P->Cr, Cs = P->Cs, Cg = P->Cg;
    const double N0r = P->N0r, N0s = P->N0s, N0g = P->N0g;
    const double GAM = P->GAM, GAM_2 = P->GAM_2, GAM_3 = P->GAM_3;
    const double GAM_1br = P->GAM_1br, GAM_2br = P->GAM_2br, GAM_3br = P->GAM_3br;
    const double GAM_3dr = P->GAM_3dr, GAM_6dr = P->GAM_6dr, GAM_1brdr = P->GAM_1brdr;
    const double GAM_5dr_h = P->GAM_5dr_h;
    const double GAM_1bs = P->GAM_1bs, GAM_2bs = P->GAM_2bs, GAM_3bs = P->GAM_3bs;
    const double GAM_3ds = P->GAM_3ds, GAM_1bsds = P->GAM_1bsds, GAM_5ds_h = P->GAM_5ds_h;
    const double GAM_1bg = P->GAM_1bg, GAM_3dg = P->GAM_3dg, GAM_1bgdg = P->GAM_1bgdg;
    const double GAM_5dg_h = P->GAM_5dg_h;
    const double Eiw = P->Eiw, Erw = P->Erw, Esw = P->Esw, Egw = P->Egw, Eri = P->Eri;
    const double Esi = P->Esi, Egi = P->Egi, Esr = P->Esr, Egr = P->Egr, Egs = P->Egs;
    const double gamma_sacr = P->gamma_sacr, gamma_gacs = P->gamma_gacs, mi = P->mi;
    const double beta_saut = P->beta_saut, gamma_saut = P->gamma_saut;
    const double beta_gaut = P->beta_gaut, gamma_gaut = P->gamma_gaut;
    const double qicrt_saut = P->qicrt_saut, qscrt_gaut = P->qscrt_gaut;
    const double f1r = P->f1r, f2r = P->f2r, f1s = P->f1s, f2s = P->f2s, f1g = P->f1g, f2g = P->f2g;
    const double A_frz = P->A_frz, B_frz = P->B_frz;
    const double mi40 = P->mi40, mi50 = P->mi50, vti50 = P->vti50, Ri50 = P->Ri50;
    const double sw_expice = P->sw_expice, Nc_ihtr = P->Nc_ihtr, Di_max = P->Di_max, Di_a = P->Di_a;
    const double sw_kk2000 = P->sw_kk2000, sw_constVti = P->sw_constVti, CONST_Vti = P->CONST_Vti;
    const double sw_roh2014 = P->sw_roh2014, ln10 = P->ln10;

    qv = dmax(q6[I_QV], 0.0);
    qc = dmax(q6[I_QC], 0.0);
    qr = dmax(q6[I_QR], 0.0);
    qi = dmax(q6[I_QI], 0.0);
    qs = dmax(q6[I_QS], 0.0);
    qg = dmax(q6[I_QG], 0.0);

    Sliq = qv / (dmax(qsatl, EPS));
    Sice = qv / (dmax(qsati, EPS));
    Rdens = 1.0 / dens;
    rho_fact = sqrt(dens00 * Rdens);
    temc = temp - TEM00;

    wk[I_delta1] = (0.5 + copysign(0.5, qr - (ORA_PLANT == 36 ? 1.0E-5 : 1.0E-4)));
    wk[I_delta2] = (0.5 + copysign(0.5, 1.0E-4 - qr)) * (0.5 + copysign(0.5, 1.0E-4 - qs));
    wk[I_spsati] = 0.5 + copysign(0.5, Sice - 1.0);
    wk[I_iceflg] = 0.5 - copysign(0.5, temc);
    wk[I_dqv_dt] = qv / dt;
    wk[I_dqc_dt] = qc / dt;
    wk[I_dqr_dt] = qr / dt;
    wk[I_dqi_dt] = qi / dt;
    wk[I_dqs_dt] = qs / dt;
    wk[I_dqg_dt] = qg / dt;

    sw_bergeron = (0.5 + copysign(0.5, temc + 30.0)) * (0.5 + copysign(0.5, 0.0 - temc)) *
                  (1.0 - sw_expice);

    zerosw = 0.5 - copysign(0.5, qr - 1.0E-12);
    RLMDr = sqrt(sqrt(dens * qr / (Ar * N0r * GAM_1br) + zerosw)) * (1.0 - zerosw);
    RLMDr_dr = sqrt(RLMDr);
    RLMDr_2 = pow(RLMDr, 2);
    RLMDr_3 = pow(RLMDr, 3);
    RLMDr_7 = pow(RLMDr, 7);
    RLMDr_1br = pow(RLMDr, 4);
    RLMDr_2br = pow(RLMDr, 5);
    RLMDr_3br = pow(RLMDr, 6);
    RLMDr_3dr = pow(RLMDr, 3) * (ORA_PLANT == 14 ? 1.0 : RLMDr_dr);
    RLMDr_5dr = pow(RLMDr, 5) * RLMDr_dr;
    RLMDr_6dr = pow(RLMDr, 6) * RLMDr_dr;

    zerosw = 0.5 - copysign(0.5, qs - 1.0E-12);
    RLMDs = sqrt(sqrt(dens * qs / (As * N0s * GAM_1bs) + zerosw)) * (1.0 - zerosw);
    RLMDs_ds = sqrt(sqrt(RLMDs));
    RLMDs_2 = pow(RLMDs, 2);
    RLMDs_3 = pow(RLMDs, 3);
    RLMDs_1bs = pow(RLMDs, 4);
    RLMDs_2bs = pow(RLMDs, 5);
    RLMDs_3bs = pow(RLMDs, 6);
    RLMDs_3ds = pow(RLMDs, 3) * RLMDs_ds;
    RLMDs_5ds = pow(RLMDs, 5) * RLMDs_ds;

    MOMs_0 = N0s * GAM * RLMDs;
    MOMs_1 = N0s * GAM_2 * RLMDs_2;
    MOMs_2 = N0s * GAM_3 * RLMDs_3;
    MOMs_0bs = N0s * GAM_1bs * RLMDs_1bs;
    MOMs_1bs = N0s * GAM_2bs * RLMDs_2bs;
    MOMs_2bs = N0s * GAM_3bs * RLMDs_3bs;
    MOMs_2ds = N0s * GAM_3ds * RLMDs_3ds;
    MOMs_5ds_h = N0s * GAM_5ds_h * sqrt(RLMDs_5ds);
    RMOMs_Vt = GAM_1bsds / GAM_1bs * RLMDs_ds;

    Xs2 = dens * qs / As;
    zerosw = 0.5 - copysign(0.5, Xs2 - 1.0E-12);
    tems = dmin(-0.1, temc);
    coef_at[0] = P->coef_a[0] + tems * (P->coef_a[1] + tems * (P->coef_a[4] + tems * P->coef_a[8]));
    coef_at[1] = P->coef_a[2] + tems * (P->coef_a[3] + tems * P->coef_a[6]);
    coef_at[2] = P->coef_a[5] + tems * P->coef_a[7];
    coef_at[3] = P->coef_a[9];
    coef_bt[0] = P->coef_b[0] + tems * (P->coef_b[1] + tems * (P->coef_b[4] + tems * P->coef_b[8]));
    coef_bt[1] = P->coef_b[2] + tems * (P->coef_b[3] + tems * P->coef_b[6]);
    coef_bt[2] = P->coef_b[5] + tems * P->coef_b[7];
    coef_bt[3] = P->coef_b[9];

#define ROH_MOM(NM)                                                                            \
    nm = (NM);                                                                                 \
    loga_ = coef_at[0] + nm * (coef_at[1] + nm * (coef_at[2] + nm * coef_at[3]));              \
    b_ = coef_bt[0] + nm * (coef_bt[1] + nm * (coef_bt[2] + nm * coef_bt[3]));

    loga_ = coef_at[0];
    b_ = coef_bt[0];
    MOMs_0 = sw_roh2014 * exp(ln10 * loga_) * exp(log(Xs2 + zerosw) * b_) * (1.0 - zerosw) +
             (1.0 - sw_roh2014) * MOMs_0;
    ROH_MOM(1.0)
    MOMs_1 = sw_roh2014 * exp(ln10 * loga_) * exp(log(Xs2 + zerosw) * b_) * (1.0 - zerosw) +
             (1.0 - sw_roh2014) * MOMs_1;
    MOMs_2 = sw_roh2014 * Xs2 + (1.0 - sw_roh2014) * MOMs_2;
    ROH_MOM(2.0)
    MOMs_0bs = sw_roh2014 * exp(ln10 * loga_) * exp(log(Xs2 + zerosw) * b_) * (1.0 - zerosw) +
               (1.0 - sw_roh2014) * MOMs_0bs;
    ROH_MOM(3.0)
    MOMs_1bs = sw_roh2014 * exp(ln10 * loga_) * exp(log(Xs2 + zerosw) * b_) * (1.0 - zerosw) +
               (1.0 - sw_roh2014) * MOMs_1bs;
    ROH_MOM(4.0)
    MOMs_2bs = sw_roh2014 * exp(ln10 * loga_) * exp(log(Xs2 + zerosw) * b_) * (1.0 - zerosw) +
               (1.0 - sw_roh2014) * MOMs_2bs;
    ROH_MOM(ORA_PLANT == 15 ? 2.0 : 2.25)
    MOMs_2ds = sw_roh2014 * exp(ln10 * loga_) * exp(log(Xs2 + zerosw) * b_) * (1.0 - zerosw) +
               (1.0 - sw_roh2014) * MOMs_2ds;
    ROH_MOM(1.625)
    MOMs_5ds_h = sw_roh2014 * exp(ln10 * loga_) * exp(log(Xs2 + zerosw) * b_) * (1.0 - zerosw) +
                 (1.0 - sw_roh2014) * MOMs_5ds_h;
    ROH_MOM(2.25)
    RMOMs_Vt = sw_roh2014 * exp(ln10 * loga_) * exp(log(Xs2 + zerosw) * b_) * (1.0 - zerosw) /
                   (MOMs_0bs + zerosw) +
               (1.0 - sw_roh2014) * RMOMs_Vt;
#undef ROH_MOM

    zerosw = 0.5 - copysign(0.5, qg - 1.0E-12);
    RLMDg = sqrt(sqrt(dens * qg / (Ag * (ORA_PLANT == 37 ? N0r : N0g) * GAM_1bg) + zerosw)) * (1.0 - zerosw);
    RLMDg_dg = sqrt(RLMDg);
    RLMDg_2 = pow(RLMDg, 2);
    RLMDg_3 = pow(RLMDg, 3);
    RLMDg_3dg = pow(RLMDg, 3) * RLMDg_dg;
    RLMDg_5dg = pow(RLMDg, 5) * RLMDg_dg;

    wk[I_RLMDr] = RLMDr;
    wk[I_RLMDs] = RLMDs;
    wk[I_RLMDg] = RLMDg;

    zerosw = 0.5 - copysign(0.5, qi - 1.0E-8);
    Vti = (1.0 - sw_constVti) * (-3.29) * exp(log(dens * qi + zerosw) * 0.16) * (1.0 - zerosw) +
          (sw_constVti) * (-CONST_Vti);
    Vtr = -Cr * rho_fact * GAM_1brdr / GAM_1br * RLMDr_dr;
    Vts = -Cs * rho_fact * RMOMs_Vt;
    Vtg = -Cg * rho_fact * GAM_1bgdg / GAM_1bg * (ORA_PLANT == 41 ? RLMDg : RLMDg_dg);

    Ni0 = dmax(exp(-0.1 * temc), 1.0) * 1000.0;
    Qi0 = P->icenuc_coef * exp(log(Ni0) * (ORA_PLANT == 30 ? 1.3 : 1.33)) * Rdens;
    wk[I_Pigen] = dmax(dmin(Qi0 - qi, qv - qsati), 0.0) / dt;

    Esi_mod = dmin(Esi, Esi * exp(gamma_sacr * temc));
    Egs_mod = dmin(Egs, Egs * exp(gamma_gacs * temc));

    Pracw_orig = qc * 0.25 * PI * Erw * N0r * Cr * (ORA_PLANT == 38 ? GAM_3 : GAM_3dr) * RLMDr_3dr * rho_fact;
    zerosw = 0.5 - copysign(0.5, qc * qr - 1.0E-12);
    Pracw_kk = 67.0 * exp(log(qc * qr + zerosw) * (ORA_PLANT == 40 ? 1.5 : 1.15)) * (1.0 - zerosw);
    wk[I_Pracw] = (1.0 - sw_kk2000) * Pracw_orig + (sw_kk2000)*Pracw_kk;

    wk[I_Psacw] = qc * 0.25 * PI * Esw * Cs * MOMs_2ds * rho_fact;
    wk[I_Pgacw] = qc * 0.25 * PI * Egw * N0g * Cg * (ORA_PLANT == 20 ? GAM_3 : GAM_3dg) * RLMDg_3dg * rho_fact;
    wk[I_Praci] = qi * 0.25 * PI * Eri * N0r * Cr * GAM_3dr * RLMDr_3dr * rho_fact;
    wk[I_Psaci] = qi * 0.25 * PI * (ORA_PLANT == 22 ? Esi : Esi_mod) * Cs * MOMs_2ds * rho_fact;
    wk[I_Pgaci] = qi * 0.25 * PI * (ORA_PLANT == 21 ? Egw : Egi) * N0g * Cg * GAM_3dg * RLMDg_3dg * rho_fact;
    wk[I_Piacr] = qi * Ar / mi * 0.25 * PI * Eri * N0r * Cr * (ORA_PLANT == 23 ? GAM_3dr : GAM_6dr) * RLMDr_6dr * rho_fact;
    wk[I_Psacr] = Ar * 0.25 * PI * Rdens * Esr * N0r * fabs(Vtr - Vts) *
                  (GAM_1br * RLMDr_1br * MOMs_2 + (ORA_PLANT == 24 ? 1.0 : 2.0) * GAM_2br * RLMDr_2br * MOMs_1 +
                   GAM_3br * RLMDr_3br * MOMs_0);
    wk[I_Pgacr] = Ar * 0.25 * PI * Rdens * Egr * N0g * N0r * fabs(Vtg - Vtr) *
                  (GAM_1br * RLMDr_1br * GAM_3 * RLMDg_3 + (ORA_PLANT == 16 ? 1.0 : 2.0) * GAM_2br * RLMDr_2br * GAM_2 * RLMDg_2 +
                   GAM_3br * RLMDr_3br * GAM * RLMDg);
    wk[I_Pracs] = As * 0.25 * PI * Rdens * Esr * N0r * fabs(Vtr - Vts) *
                  (MOMs_0bs * GAM_3 * RLMDr_3 + 2.0 * MOMs_1bs * GAM_2 * RLMDr_2 +
                   (ORA_PLANT == 25 ? MOMs_1bs : MOMs_2bs) * GAM * RLMDr);
    wk[I_Pgacs] = As * 0.25 * PI * Rdens * Egs_mod * N0g * fabs(Vtg - (ORA_PLANT == 26 ? Vtr : Vts)) *
                  (MOMs_0bs * GAM_3 * RLMDg_3 + 2.0 * MOMs_1bs * GAM_2 * RLMDg_2 + MOMs_2bs * GAM * RLMDg);

    rhoqc = dens * qc * 1000.0;
    Dc = 0.146 - 5.964E-2 * log(Nc / 2000.0);
    Praut_berry = Rdens * 1.67E-5 * rhoqc * rhoqc / (5.0 + 3.66E-2 * Nc / (Dc * rhoqc + EPS));
    zerosw = 0.5 - copysign(0.5, qc - 1.0E-12);
    Praut_kk = P->kk_aut_rate * exp(log(qc + zerosw) * 2.47) * (1.0 - zerosw) * exp(log(Nc) * (-1.79));
    Praut_kk = P->kk_aut_rate * pow(qc, 2.47) * pow(Nc, -1.79);   /* QUIRK A-9: pow form wins */
    wk[I_Praut] = (1.0 - sw_kk2000) * Praut_berry + sw_kk2000 * Praut_kk;

    betai = dmin(beta_saut, beta_saut * exp(gamma_saut * temc));
    wk[I_Psaut] = dmax(betai * (qi - qicrt_saut), 0.0);
    betas = dmin(beta_gaut, beta_gaut * exp(gamma_gaut * temc));
    wk[I_Pgaut] = dmax(betas * (qs - (ORA_PLANT == 34 ? qicrt_saut : qscrt_gaut)), 0.0);

    Ka = (Ka0 + dKa_dT * temc);
    Kd = (Kd0 + dKd_dT * temc) * PRE00 / pre;
    Nu = (nu0 + dnu_dT * temc) * Rdens;
    Glv = 1.0 / (LHV0 / (Ka * temp) * (LHV0 / (Rvap * temp) - 1.0) + 1.0 / (Kd * dens * qsatl));
    Giv = 1.0 / ((ORA_PLANT == 27 ? LHV0 : LHS0) / (Ka * temp) * (LHS0 / (Rvap * temp) - 1.0) + 1.0 / (Kd * dens * qsati));
    Gil = 1.0 / ((ORA_PLANT == 18 ? LHV0 : LHF0) / (Ka * temc));   /* IEEE: temc == 0 -> 1/inf = 0 */

    ventr = f1r * GAM_2 * (ORA_PLANT == 17 ? RLMDr_3 : RLMDr_2) + f2r * sqrt(Cr * rho_fact / Nu * RLMDr_5dr) * GAM_5dr_h;
    wk[I_Prevp] = 2.0 * PI * Rdens * N0r * (1.0 - dmin(Sliq, 1.0)) * Glv * ventr;

    rhoqi = dmax(dens * qi, EPS);
    XNi = dmin(dmax(5.38E+7 * exp(log(rhoqi) * (ORA_PLANT == 29 ? 0.7 : 0.75)), 1.0E+3), 1.0E+6);
    XMi = rhoqi / XNi;
    Di = dmin(Di_a * sqrt(XMi), Di_max);
    tmp = 4.0 * Di * XNi * Rdens * (Sice - 1.0) * Giv;
    wk[I_Pidep] = (wk[I_spsati]) * (tmp);
    wk[I_Pisub] = (1.0 - wk[I_spsati]) * (-tmp);

    sw = (0.5 - copysign(0.5, temc + (ORA_PLANT == 32 ? 38.0 : 40.0)));
    wk[I_Pihom] = sw * qc / dt;
    sw = (0.5 + copysign(0.5, temc + 40.0)) * (0.5 - copysign(0.5, temc));
    wk[I_Pihtr] = sw * (dens / rho_w * (ORA_PLANT == 31 ? qc : qc * qc) / (Nc_ihtr * 1.0E+6)) * B_frz * (exp(-A_frz * temc) - 1.0);
    sw = (0.5 + copysign(0.5, temc));
    wk[I_Pimlt] = sw * (ORA_PLANT == 39 ? qs : qi) / dt;

    vents = f1s * MOMs_1 + f2s * sqrt(Cs * rho_fact / Nu) * MOMs_5ds_h;
    tmp = 2.0 * PI * Rdens * (Sice - 1.0) * Giv * vents;
    wk[I_Psdep] = (wk[I_spsati]) * (tmp);
    wk[I_Pssub] = (1.0 - wk[I_spsati]) * (-tmp);

    wk[I_Psmlt] = 2.0 * PI * Rdens * Gil * vents + CL * temc / LHF0 * (wk[I_Psacw] + wk[I_Psacr]);
    wk[I_Psmlt] = dmax(wk[I_Psmlt], 0.0);

    ventg = f1g * GAM_2 * RLMDg_2 + f2g * sqrt(Cg * rho_fact / Nu * (ORA_PLANT == 28 ? RLMDg_3dg : RLMDg_5dg)) * GAM_5dg_h;
    tmp = 2.0 * PI * Rdens * N0g * (Sice - 1.0) * Giv * ventg;
    wk[I_Pgdep] = (wk[I_spsati]) * (tmp);
    wk[I_Pgsub] = (1.0 - wk[I_spsati]) * (-tmp);

    wk[I_Pgmlt] = 2.0 * PI * Rdens * N0g * Gil * ventg + CL * temc / LHF0 * ((ORA_PLANT == 33 ? 0.0 : wk[I_Pgacw]) + wk[I_Pgacr]);
    wk[I_Pgmlt] = dmax(wk[I_Pgmlt], 0.0);

    wk[I_Pgfrz] = 2.0 * PI * Rdens * N0r * 60.0 * B_frz * Ar * (exp(-A_frz * temc) - 1.0) * (ORA_PLANT == 19 ? pow(RLMDr, 6) : RLMDr_7);

    dt1 = (exp(log(mi50) * ma2) - exp(log(mi40) * ma2)) / (a1 * ma2);
    Ni50 = qi * dt / (mi50 * dt1);
    wk[I_Psfw] = Ni50 * (a1 * pow(mi50, a2) + PI * Eiw * dens * qc * Ri50 * (ORA_PLANT == 35 ? 1.0 : Ri50) * vti50);
    wk[I_Psfi] = qi / dt1;

    /* limiter */
    wk[I_Pigen] = dmin(wk[I_Pigen], wk[I_dqv_dt]) * (wk[I_iceflg]) * sw_expice;
    wk[I_Pidep] = dmin(wk[I_Pidep], wk[I_dqv_dt]) * (wk[I_iceflg]) * sw_expice;
    wk[I_Psdep] = dmin(wk[I_Psdep], wk[I_dqv_dt]) * (wk[I_iceflg]);
    wk[I_Pgdep] = dmin(wk[I_Pgdep], wk[I_dqv_dt]) * (wk[I_iceflg]);

    wk[I_Pracw] = wk[I_Pracw] + wk[I_Psacw] * (1.0 - wk[I_iceflg]) + wk[I_Pgacw] * (1.0 - wk[I_iceflg]);

    wk[I_Praut] = dmin(wk[I_Praut], wk[I_dqc_dt]);
    wk[I_Pracw] = dmin(wk[I_Pracw], wk[I_dqc_dt]);
    wk[I_Pihom] = dmin(wk[I_Pihom], wk[I_dqc_dt]) * (wk[I_iceflg]) * sw_expice;
    wk[I_Pihtr] = dmin(wk[I_Pihtr], wk[I_dqc_dt]) * (wk[I_iceflg]) * sw_expice;
    wk[I_Psacw] = dmin(wk[I_Psacw], wk[I_dqc_dt]) * (wk[I_iceflg]);
    wk[I_Psfw] = dmin(wk[I_Psfw], wk[I_dqc_dt]) * (wk[I_iceflg]) * sw_bergeron;
    wk[I_Pgacw] = dmin(wk[I_Pgacw], wk[I_dqc_dt]) * (wk[I_iceflg]);

    wk[I_Prevp] = dmin(wk[I_Prevp], wk[I_dqr_dt]);
    wk[I_Piacr] = dmin(wk[I_Piacr], wk[I_dqr_dt]) * (wk[I_iceflg]);
    wk[I_Psacr] = dmin(wk[I_Psacr], wk[I_dqr_dt]) * (wk[I_iceflg]);
    wk[I_Pgacr] = dmin(wk[I_Pgacr], wk[I_dqr_dt]) * (wk[I_iceflg]);
    wk[I_Pgfrz] = dmin(wk[I_Pgfrz], wk[I_dqr_dt]) * (wk[I_iceflg]);

    wk[I_Pisub] = dmin(wk[I_Pisub], wk[I_dqi_dt]) * (wk[I_iceflg]) * sw_expice;
    wk[I_Pimlt] = dmin(wk[I_Pimlt], wk[I_dqi_dt]) * (1.0 - wk[I_iceflg]) * sw_expice;
    wk[I_Psaut] = dmin(wk[I_Psaut], wk[I_dqi_dt]) * (wk[I_iceflg]);
    wk[I_Praci] = dmin(wk[I_Praci], wk[I_dqi_dt]) * (wk[I_iceflg]);
    wk[I_Psaci] = dmin(wk[I_Psaci], wk[I_dqi_dt]) * (wk[I_iceflg]);
    wk[I_Psfi] = dmin(wk[I_Psfi], wk[I_dqi_dt]) * (wk[I_iceflg]) * sw_bergeron;
    wk[I_Pgaci] = dmin(wk[I_Pgaci], wk[I_dqi_dt]) * (wk[I_iceflg]);

    wk[I_Pssub] = dmin(wk[I_Pssub], wk[I_dqs_dt]) * (wk[I_iceflg]);
    wk[I_Psmlt] = dmin(wk[I_Psmlt], wk[I_dqs_dt]) * (1.0 - wk[I_iceflg]);
    wk[I_Pgaut] = dmin(wk[I_Pgaut], wk[I_dqs_dt]) * (wk[I_iceflg]);
    wk[I_Pracs] = dmin(wk[I_Pracs], wk[I_dqs_dt]) * (wk[I_iceflg]);
    wk[I_Pgacs] = dmin(wk[I_Pgacs], wk[I_dqs_dt]);

    wk[I_Pgsub] = dmin(wk[I_Pgsub], wk[I_dqg_dt]) * (wk[I_iceflg]);
    wk[I_Pgmlt] = dmin(wk[I_Pgmlt], wk[I_dqg_dt]) * (1.0 - wk[I_iceflg]);

    wk[I_Piacr_s] = (1.0 - wk[I_delta1]) * wk[I_Piacr];
    wk[I_Piacr_g] = (wk[I_delta1]) * wk[I_Piacr];
    wk[I_Praci_s] = (1.0 - wk[I_delta1]) * wk[I_Praci];
    wk[I_Praci_g] = (wk[I_delta1]) * wk[I_Praci];
    wk[I_Psacr_s] = (wk[I_delta2]) * wk[I_Psacr];
    wk[I_Psacr_g] = (1.0 - wk[I_delta2]) * wk[I_Psacr];
    wk[I_Pracs] = (1.0 - wk[I_delta2]) * wk[I_Pracs];

    /* [QC] */
    net = +wk[I_Pimlt] - wk[I_Praut] - wk[I_Pracw] - wk[I_Pihom] - wk[I_Pihtr] - wk[I_Psacw] -
          wk[I_Psfw] - wk[I_Pgacw];
    fac_sw = 0.5 + copysign(0.5, net + EPS);
    fac = fac_sw + (1.0 - fac_sw) * dmin(-wk[I_dqc_dt] / (net - fac_sw), 1.0);
    wk[I_Pimlt] *= fac; wk[I_Praut] *= fac; wk[I_Pracw] *= fac; wk[I_Pihom] *= fac;
    wk[I_Pihtr] *= fac; wk[I_Psacw] *= fac; wk[I_Psfw] *= fac; wk[I_Pgacw] *= fac;

    /* [QI] */
    net = +wk[I_Pigen] + wk[I_Pidep] + wk[I_Pihom] + wk[I_Pihtr] - wk[I_Pisub] - wk[I_Pimlt] -
          wk[I_Psaut] - wk[I_Praci_s] - wk[I_Psaci] - wk[I_Psfi] - wk[I_Praci_g] - wk[I_Pgaci];
    fac_sw = 0.5 + copysign(0.5, net + EPS);
    fac = fac_sw + (1.0 - fac_sw) * dmin(-wk[I_dqi_dt] / (net - fac_sw), 1.0);
    wk[I_Pigen] *= fac; wk[I_Pidep] *= fac; wk[I_Pihom] *= fac; wk[I_Pihtr] *= fac;
    wk[I_Pisub] *= fac; wk[I_Pimlt] *= fac; wk[I_Psaut] *= fac; wk[I_Praci_s] *= fac;
    wk[I_Psaci] *= fac; wk[I_Psfi] *= fac; wk[I_Praci_g] *= fac; wk[I_Pgaci] *= fac;

    /* [QR] */
    net = +wk[I_Praut] + wk[I_Pracw] + wk[I_Psmlt] + wk[I_Pgmlt] - wk[I_Prevp] - wk[I_Piacr_s] -
          wk[I_Psacr_s] - wk[I_Piacr_g] - wk[I_Psacr_g] - wk[I_Pgacr] - wk[I_Pgfrz];
    fac_sw = 0.5 + copysign(0.5, net + EPS);
    fac = fac_sw + (1.0 - fac_sw) * dmin(-wk[I_dqr_dt] / (net - fac_sw), 1.0);
    wk[I_Praut] *= fac; wk[I_Pracw] *= fac; wk[I_Psmlt] *= fac; wk[I_Pgmlt] *= fac;
    wk[I_Prevp] *= fac; wk[I_Piacr_s] *= fac; wk[I_Psacr_s] *= fac; wk[I_Piacr_g] *= fac;
    wk[I_Psacr_g] *= fac; wk[I_Pgacr] *= fac; wk[I_Pgfrz] *= fac;

    /* [QV] */
    net = +wk[I_Prevp] + wk[I_Pisub] + wk[I_Pssub] + wk[I_Pgsub] - wk[I_Pigen] - wk[I_Pidep] -
          wk[I_Psdep] - wk[I_Pgdep];
    fac_sw = 0.5 + copysign(0.5, net + EPS);
    fac = fac_sw + (1.0 - fac_sw) * dmin(-wk[I_dqv_dt] / (net - fac_sw), 1.0);
    wk[I_Prevp] *= fac; wk[I_Pisub] *= fac; wk[I_Pssub] *= fac; wk[I_Pgsub] *= fac;
    wk[I_Pigen] *= fac; wk[I_Pidep] *= fac; wk[I_Psdep] *= fac; wk[I_Pgdep] *= fac;

    /* [QS] */
    net = +wk[I_Psdep] + wk[I_Psacw] + wk[I_Psfw] + wk[I_Piacr_s] + wk[I_Psacr_s] + wk[I_Psaut] +
          wk[I_Praci_s] + wk[I_Psaci] + wk[I_Psfi] - wk[I_Pssub] - wk[I_Psmlt] - wk[I_Pgaut] -
          wk[I_Pracs] - wk[I_Pgacs];
    fac_sw = 0.5 + copysign(0.5, net + EPS);
    fac = fac_sw + (1.0 - fac_sw) * dmin(-wk[I_dqs_dt] / (net - fac_sw), 1.0);
    wk[I_Psdep] *= fac; wk[I_Psacw] *= fac; wk[I_Psfw] *= fac; wk[I_Piacr_s] *= fac;
    wk[I_Psacr_s] *= fac; wk[I_Psaut] *= fac; wk[I_Praci_s] *= fac; wk[I_Psaci] *= fac;
    wk[I_Psfi] *= fac; wk[I_Pssub] *= fac; wk[I_Psmlt] *= fac; wk[I_Pgaut] *= fac;
    wk[I_Pracs] *= fac; wk[I_Pgacs] *= fac;

    /* [QG] */
    net = +wk[I_Pgdep] + wk[I_Pgacw] + wk[I_Piacr_g] + wk[I_Psacr_g] + wk[I_Pgacr] + wk[I_Pgfrz] +
          wk[I_Praci_g] + wk[I_Pgaci] + wk[I_Pgaut] + wk[I_Pracs] + wk[I_Pgacs] - wk[I_Pgsub] -
          wk[I_Pgmlt];
    fac_sw = 0.5 + copysign(0.5, net + EPS);
    fac = fac_sw + (1.0 - fac_sw) * dmin(-wk[I_dqg_dt] / (net - fac_sw), 1.0);
    wk[I_Pgdep] *= fac; wk[I_Pgacw] *= fac; wk[I_Piacr_g] *= fac; wk[I_Psacr_g] *= fac;
    wk[I_Pgacr] *= fac; wk[I_Pgfrz] *= fac; wk[I_Praci_g] *= fac; wk[I_Pgaci] *= fac;
    wk[I_Pgaut] *= fac; wk[I_Pracs] *= fac; wk[I_Pgacs] *= fac; wk[I_Pgsub] *= fac;
    wk[I_Pgmlt] *= fac;

    qc_t = +wk[I_Pimlt] - wk[I_Praut] - wk[I_Pracw] - wk[I_Pihom] - wk[I_Pihtr] - wk[I_Psacw] -
           wk[I_Psfw] - wk[I_Pgacw];
    qr_t = +wk[I_Praut] + wk[I_Pracw] + wk[I_Psmlt] + wk[I_Pgmlt] - wk[I_Prevp] - wk[I_Piacr_s] -
           wk[I_Psacr_s] - wk[I_Piacr_g] - wk[I_Psacr_g] - wk[I_Pgacr] - wk[I_Pgfrz];
    qi_t = +wk[I_Pigen] + wk[I_Pidep] + wk[I_Pihom] + wk[I_Pihtr] - wk[I_Pisub] - wk[I_Pimlt] -
           wk[I_Psaut] - wk[I_Praci_s] - wk[I_Psaci] - wk[I_Psfi] - wk[I_Praci_g] - wk[I_Pgaci];
    qs_t = +wk[I_Psdep] + wk[I_Psacw] + wk[I_Psfw] + wk[I_Piacr_s] + wk[I_Psacr_s] + wk[I_Psaut] +
           wk[I_Praci_s] + wk[I_Psaci] + wk[I_Psfi] - wk[I_Pssub] - wk[I_Psmlt] - wk[I_Pgaut] -
           wk[I_Pracs] - wk[I_Pgacs];
    qg_t = +wk[I_Pgdep] + wk[I_Pgacw] + wk[I_Piacr_g] + wk[I_Psacr_g] + wk[I_Pgacr] + wk[I_Pgfrz] +
           wk[I_Praci_g] + wk[I_Pgaci] + wk[I_Pgaut] + wk[I_Pracs] + wk[I_Pgacs] - wk[I_Pgsub] -
           wk[I_Pgmlt];

    qc_t = dmax(qc_t, -wk[I_dqc_dt]);
    qr_t = dmax(qr_t, -wk[I_dqr_dt]);
    qi_t = dmax(qi_t, -wk[I_dqi_dt]);
    qs_t = dmax(qs_t, -wk[I_dqs_dt]);
    qg_t = dmax(qg_t, -wk[I_dqg_dt]);
    qv_t = -(qc_t + qr_t + qi_t + qs_t + qg_t);

    o->drhogq[I_QV] = rhog * qv_t * dt;
    o->drhogq[I_QC] = rhog * qc_t * dt;
    o->drhogq[I_QR] = rhog * qr_t * dt;
    o->drhogq[I_QI] = rhog * qi_t * dt;
    o->drhogq[I_QS] = rhog * qs_t * dt;
    o->drhogq[I_QG] = rhog * qg_t * dt;
    o->q_t[I_QV] = qv_t; o->q_t[I_QC] = qc_t; o->q_t[I_QR] = qr_t;
    o->q_t[I_QI] = qi_t; o->q_t[I_QS] = qs_t; o->q_t[I_QG] = qg_t;

    o->Vt[I_QV] = 0.0; o->Vt[I_QC] = 0.0;
    o->Vt[I_QR] = Vtr; o->Vt[I_QI] = Vti; o->Vt[I_QS] = Vts; o->Vt[I_QG] = Vtg;

    xf_qc = rhoqc / Nc * 1.0E-9;
    rf_qc = pow((P->coef_xf * xf_qc + EPS), 0.33333333);
    r2_qc = P->coef_dgam * (rf_qc * rf_qc) * (Nc * 1.0E+6);
    r2_qr = 0.25 * N0r * GAM_3 * RLMDr_3;
    r3_qc = P->coef_xf * dens * qc;
    r3_qr = P->coef_xf * dens * qr;
    zerosw = 0.5 - copysign(0.5, (r2_qc + r2_qr) - r2_min);
    o->rceff = (r3_qc + r3_qr) / (r2_qc + r2_qr + zerosw) * (1.0 - zerosw);
    sw = (0.5 + copysign(0.5, (qc + qr) - q_min)) * zerosw;
    o->rctop = (sw)*o->rceff + (1.0 - sw) * UNDEF;
    o->tctop = (sw)*temp + (1.0 - sw) * UNDEF;
    o->cloud_sw = (0.5 + copysign(0.5, (qc + qr) - q_min)) * (1.0 - zerosw);
    zerosw = 0.5 - copysign(0.5, r2_qc - r2_min);
    o->rceff_cld = r3_qc / (r2_qc + zerosw) * (1.0 - zerosw);
    sw = (0.5 + copysign(0.5, qc - q_min)) * zerosw;
    o->rctop_cld = (sw)*o->rceff_cld + (1.0 - sw) * UNDEF;
    o->tctop_cld = (sw)*temp + (1.0 - sw) * UNDEF;
    memcpy(o->wk, wk, sizeof wk);
}

/* ---- vadv1d_prep: src/mod_vadv1d.cpp:231-391 (host code in the reference) ----- */
void ora_vadv1d_prep(int ijdim, int mkmin, int mkmax, const double *dz, const double *zh,
                     const double *wp, double *zdis, int *kcell, int *kcell_max, int *kcell_min,
                     double dt)
{
    double *wh = zalloc((size_t)KALL * ijdim);
    double dt2 = pow(dt, 2);
    double dt3 = pow(dt, 3);
    for (int k = mkmin + 1; k <= mkmax; k++)
        for (int ij = 0; ij < ijdim; ij++)
            wh[IDX(k, ij)] = 0.5 * (wp[IDX(k - 1, ij)] + wp[IDX(k, ij)]);
    for (int ij = 0; ij < ijdim; ij++) {
        wh[IDX(mkmin, ij)] = wp[IDX(mkmin, ij)];
        wh[IDX(mkmin - 1, ij)] = wp[IDX(mkmin - 1, ij)];
        wh[IDX(mkmax + 1, ij)] = wp[IDX(mkmax, ij)];
    }
    for (int k = mkmin + 1; k <= mkmax; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            double whk = wh[IDX(k, ij)], whp = wh[IDX(k + 1, ij)], whm = wh[IDX(k - 1, ij)];
            zdis[IDX(k, ij)] =
                dt * whk - dt2 * whk * (whp - whm) / (dz[k - 1] + dz[k]) / 2.0 +
                dt3 * whk *
                    (pow((whp - whm) / (dz[k - 1] + dz[k]), 2) +
                     whk * (((whp - whk) / dz[k] - (whk - whm) / dz[k - 1]) / (dz[k - 1] + dz[k]) * 2.0)) /
                    6.0;
        }
    for (int ij = 0; ij < ijdim; ij++) {
        zdis[IDX(mkmin - 1, ij)] = 0.0;
        zdis[IDX(mkmin, ij)] = dt * wh[IDX(mkmin, ij)] -
                               dt2 * wh[IDX(mkmin, ij)] * (wh[IDX(mkmin + 1, ij)] - wh[IDX(mkmin, ij)]) /
                                   dz[mkmin] / 2.0;
        zdis[IDX(mkmax + 1, ij)] = dt * wh[IDX(mkmax + 1, ij)] -
                                   dt2 * wh[IDX(mkmax + 1, ij)] *
                                       (wh[IDX(mkmax + 1, ij)] - wh[IDX(mkmax, ij)]) / dz[mkmax] / 2.0;
    }
    for (int k = 0; k < KALL; k++) {
        for (int ij = 0; ij < ijdim; ij++) kcell[IDX(k, ij)] = k;
        kcell_min[k] = k;
        kcell_max[k] = k;
    }
    for (int k = mkmin; k <= mkmax; k++) {
        double zzmax = DBL_MIN;     /* QUIRK A-4: numeric_limits<double>::min() */
        double zzmin = DBL_MIN;
        for (int ij = 0; ij < ijdim; ij++) {
            double val = zdis[IDX(k, ij)];
            zzmax = val > zzmax ? val : zzmax;
            zzmin = val < zzmin ? val : zzmin;
        }
        if (zzmax > 0.0) {
            for (int k2 = k; k2 >= mkmin; k2--) {
                if ((zh[k2] <= zh[k] - zzmax) && (zh[k2 + 1] > zh[k] - zzmax)) {
                    kcell_min[k] = k2;
                    break;
                }
            }
        }
        if (zzmin < 0.0) {
            for (int k2 = k; k2 <= mkmax; k2++) {
                if ((zh[k2] <= zh[k] - zzmin) && (zh[k2 + 1] > zh[k] - zzmin)) {
                    kcell_max[k] = k2;
                    break;
                }
            }
        }
    }
    for (int k = mkmin; k <= mkmax; k++) {
        for (int ij = 0; ij < ijdim; ij++) {
            if (kcell_min[k] == k && kcell_max[k] == k) {
                kcell[IDX(k, ij)] = k;
            } else {
                int kc = 0;
                for (int k2 = kcell_min[k]; k2 <= kcell_max[k]; k2++) {
                    double x = zh[k] - zdis[IDX(k, ij)];
                    int tmp = (int)((double)k2 * copysign(1.0, x - zh[k2]) * copysign(1.0, zh[k2 + 1] - x));
                    kc = kc > tmp ? kc : tmp;
                }
                kcell[IDX(k, ij)] = kc;
            }
        }
    }
    for (int k = 0; k < KALL; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            if (kcell[IDX(k, ij)] == 0) kcell[IDX(k, ij)] = mkmin;
            if (kcell[IDX(k, ij)] == mkmax + 1) kcell[IDX(k, ij)] = mkmax;
        }
    free(wh);
}

/* ---- vadv1d_getflux_new: src/mod_vadv1d.cpp:393-450 --------------------------- */
void ora_vadv1d_getflux(int ijdim, int mkmin, int mkmax, const double *dz, const double *rhof,
                        const double *zdis0, const int *kcell, const int *kcell_max,
                        const int *kcell_min, double *frhof)
{
    for (size_t i = 0; i < (size_t)KALL * ijdim; i++) frhof[i] = 0.0;
#pragma omp parallel for collapse(2)
    for (int k = mkmin; k <= mkmax; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            double zdis = 0.0;
            double acc = frhof[IDX(k, ij)];
            int kc = kcell[IDX(k, ij)];
            if (kcell_min[k] == k && kcell_max[k] == k) {
                zdis = zdis0[IDX(k, ij)];
            } else {
                for (int k2 = kcell_min[k]; k2 <= kcell_max[k]; k2++) {
                    double f1 = (double)(k2 - (kc + 1));
                    double f2 = (double)((k - 1) - k2);
                    double f3 = (double)(k2 - k);
                    double f4 = (double)((kc - 1) - k2);
                    double fact = dz[k2] * 0.25 *
                                  ((copysign(1.0, f1) + 1.0) * (copysign(1.0, f2) + 1.0) -
                                   (copysign(1.0, f3) + 1.0) * (copysign(1.0, f4) + 1.0));
                    acc = acc + rhof[IDX(k2, ij)] * fact;
                    zdis = zdis0[IDX(k, ij)] - fact;   /* QUIRK A-2: assignment */
                }
            }
            acc = acc + rhof[IDX(kc, ij)] * zdis;
            frhof[IDX(k, ij)] = acc;
        }
#pragma omp parallel for
    for (size_t i = 0; i < (size_t)KALL * ijdim; i++)
        frhof[i] = frhof[i] * (0.5 + copysign(0.5, fabs(frhof[i]) - CONST_EPS));
}

/* ---- cnvvar_rhogkin_in: src/mod_debug.cpp:269-330 ------------------------------ */
static void cnvvar_rhogkin_in(int ijdim, const double *rhog, const double *rhogvx,
                              const double *rhogvy, const double *rhogvz, const double *rhogw,
                              const double *C2Wfact, const double *W2Cfact, double *rhogkin,
                              double *rhogkin_h, double *rhogkin_v)
{
    const size_t K2 = (size_t)KALL * ijdim;
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int g = 0; g < ijdim; g++)
            rhogkin_h[IDX(k, g)] = 0.5 *
                                   (rhogvx[IDX(k, g)] * rhogvx[IDX(k, g)] + rhogvy[IDX(k, g)] * rhogvy[IDX(k, g)] +
                                    rhogvz[IDX(k, g)] * rhogvz[IDX(k, g)]) /
                                   rhog[IDX(k, g)];
    for (int g = 0; g < ijdim; g++) {
        rhogkin_h[IDX(KMIN - 1, g)] = 0.0;
        rhogkin_h[IDX(KMAX + 1, g)] = 0.0;
    }
#pragma omp parallel for collapse(2)
    for (int k = KMIN + 1; k <= KMAX; k++)
        for (int g = 0; g < ijdim; g++)
#if ORA_PLANT == 6
            rhogkin_v[IDX(k, g)] = 0.5 * (rhogw[IDX(k, g)] * rhogw[IDX(k, g)]) /
                                   (C2Wfact[K2 + IDX(k, g)] * rhog[IDX(k, g)] +
                                    C2Wfact[IDX(k, g)] * rhog[IDX(k - 1, g)]);
#else
            rhogkin_v[IDX(k, g)] = 0.5 * (rhogw[IDX(k, g)] * rhogw[IDX(k, g)]) /
                                   (C2Wfact[IDX(k, g)] * rhog[IDX(k, g)] +
                                    C2Wfact[K2 + IDX(k, g)] * rhog[IDX(k - 1, g)]);
#endif
    for (int g = 0; g < ijdim; g++) {
        rhogkin_v[IDX(KMIN - 1, g)] = 0.0;
        rhogkin_v[IDX(KMIN, g)] = 0.0;
        rhogkin_v[IDX(KMAX + 1, g)] = 0.0;
    }
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int g = 0; g < ijdim; g++)
#if ORA_PLANT == 7
            rhogkin[IDX(k, g)] = rhogkin_h[IDX(k, g)] +
                                 (W2Cfact[K2 + IDX(k, g)] * rhogkin_v[IDX(k + 1, g)] +
                                  W2Cfact[IDX(k, g)] * rhogkin_v[IDX(k, g)]);
#else
            rhogkin[IDX(k, g)] = rhogkin_h[IDX(k, g)] +
                                 (W2Cfact[IDX(k, g)] * rhogkin_v[IDX(k + 1, g)] +
                                  W2Cfact[K2 + IDX(k, g)] * rhogkin_v[IDX(k, g)]);
#endif
    for (int g = 0; g < ijdim; g++) {
        rhogkin[IDX(KMIN - 1, g)] = 0.0;
        rhogkin[IDX(KMAX + 1, g)] = 0.0;
    }
}

/* ---- precip_transport_new: src/mod_precip_transport.cpp:573-1087 -------------- */
typedef struct region_ptrs {
    double *rhog, *rhogvx, *rhogvy, *rhogvz, *rhogw, *rhoge, *rhogq, *rho, *tem, *pre;
    double *vx, *vy, *vz, *w, *unccn, *q, *qd, *z, *precip, *precip_rhoe, *precip_lh_heat;
    double *precip_rhophi, *precip_rhokin, *gprec, *gsgam2, *gsgam2h, *gam2, *gam2h;
    double *ix, *iy, *iz, *jx, *jy, *jz;
} region_ptrs;

/* C2Wfact [2][k][ij] and W2Cfact [2][k][ij] of precip_transport_new
 * (src/mod_precip_transport.cpp:700-738): cell-to-wall and wall-to-cell weights. */
static void c2w_w2c_fact(const nsw6_grid *G, int ijdim, const double *gsgam2, const double *gsgam2h,
                         double *C2Wfact, double *W2Cfact)
{
    const size_t K2 = (size_t)KALL * ijdim;
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX + 1; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            C2Wfact[IDX(k, ij)] = G->afact[k] / gsgam2[IDX(k, ij)] * gsgam2h[IDX(k, ij)];
            C2Wfact[K2 + IDX(k, ij)] = G->bfact[k] / gsgam2[IDX(k - 1, ij)] * gsgam2h[IDX(k, ij)];
        }
    for (int ij = 0; ij < ijdim; ij++) {
        C2Wfact[IDX(KMIN - 1, ij)] = 0.0;
        C2Wfact[K2 + IDX(KMIN - 1, ij)] = 0.0;
    }
#pragma omp parallel for collapse(2)
    for (int k = KMIN - 1; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            W2Cfact[IDX(k, ij)] = G->cfact[k] * gsgam2[IDX(k, ij)] / gsgam2h[IDX(k + 1, ij)];
            W2Cfact[K2 + IDX(k, ij)] = G->dfact[k] * gsgam2[IDX(k, ij)] / gsgam2h[IDX(k + 1, ij)];
        }
    for (int ij = 0; ij < ijdim; ij++) {
        W2Cfact[IDX(KMAX + 1, ij)] = 0.0;
        W2Cfact[K2 + IDX(KMAX + 1, ij)] = 0.0;
    }
}

static void record_bounds(int t, int nt, int type, const int *kmin, const int *kmax)
{
    if (!g_bounds_rec) return;
    int *b = g_bounds_rec + (((size_t)g_bounds_l * nt + t) * 2 + type) * KALL * 2;
    for (int k = 0; k < KALL; k++) {
        b[k * 2 + 0] = kmin[k];
        b[k * 2 + 1] = kmax[k];
    }
}

void ora_record_kcell_bounds(int *buf) { g_bounds_rec = buf; }
void ora_record_newton_margin(double *buf) { g_newton_rec = buf; }

static void precip_transport_new(const nsw6_grid *G, int ijdim, region_ptrs *R, const double *Vterm,
                                 const int precipitating_flag[NQMAX], const double *rgs,
                                 const double *rgsh, double dt, int kin2kin)
{
    const size_t K2 = (size_t)KALL * ijdim;
    double *rhogkin = zalloc(K2), *rhogkin_h = zalloc(K2), *rhogkin_v = zalloc(K2);
    double *rhoq = zalloc(K2), *rhoeq = zalloc(K2), *rhophiq = zalloc(K2), *rhokin_h = zalloc(K2);
    double *rhokin_v = zalloc(K2), *rhouq = zalloc(K2), *rhovq = zalloc(K2), *rhowq = zalloc(K2);
    double *fprec_q = zalloc(K2), *fprec_rhoe = zalloc(K2), *fprec_rhophi = zalloc(K2);
    double *fprec_rhokin_h = zalloc(K2), *fprec_rhokin_v = zalloc(K2), *fprec_rhou = zalloc(K2);
    double *fprec_rhov = zalloc(K2), *fprec_rhow = zalloc(K2);
    double *drhoq = zalloc((size_t)NQMAX * K2), *drhoe = zalloc(K2), *drhophi = zalloc(K2);
    double *drhokin_h = zalloc(K2), *drhokin_v = zalloc(K2), *drhogu = zalloc(K2);
    double *drhogv = zalloc(K2), *drhogw = zalloc(K2);
    double *zdis0 = zalloc(K2), *zdis0h = zalloc(K2);
    int *kcell = (int *)calloc(K2, sizeof(int)), *kcellh = (int *)calloc(K2, sizeof(int));
    int kcell_max[KALL] = {0}, kcell_min[KALL] = {0}, kcellh_max[KALL] = {0}, kcellh_min[KALL] = {0};
    double *Vtermh = zalloc(K2), *qh = zalloc(K2), *ein = zalloc(K2), *tmp = zalloc(K2);
    double *tmp_h = zalloc(K2), *tmp_v = zalloc(K2), *tmp2 = zalloc(K2);
    double *C2Wfact = zalloc(2 * K2), *W2Cfact = zalloc(2 * K2);
    double GRD_gz_shift[KALL] = {0};

    double *rhog = R->rhog, *rhoge = R->rhoge, *rhogq = R->rhogq, *q = R->q, *tem = R->tem;

    for (int k = KMIN; k <= KMAX + 1; k++) GRD_gz_shift[k] = G->gz[k - 1];

    c2w_w2c_fact(G, ijdim, R->gsgam2, R->gsgam2h, C2Wfact, W2Cfact);

    cnvvar_rhogkin_in(ijdim, rhog, R->rhogvx, R->rhogvy, R->rhogvz, R->rhogw, C2Wfact, W2Cfact,
                      rhogkin, rhogkin_h, rhogkin_v);

    for (int ij = 0; ij < ijdim; ij++) {
        R->precip[0 * ijdim + ij] = 0.0;
        R->precip[0 * ijdim + ij] = 0.0;       /* QUIRK A-6: precip(1) not zeroed here */
        R->precip_rhoe[ij] = 0.0;
        R->precip_lh_heat[ij] = 0.0;
        R->precip_rhophi[ij] = 0.0;
        R->precip_rhokin[ij] = 0.0;
    }

    int n_sed = 0;
    for (int nq = 0; nq < NQMAX; nq++) n_sed += precipitating_flag[nq] != 0;
    int nt_rec = -1;   /* sedimenting-tracer index of nq */
    for (int nq = 0; nq < NQMAX; nq++) {
        if (!precipitating_flag[nq]) continue;
        nt_rec++;
        const double *Vt_nq = Vterm + (size_t)nq * K2;
        ora_vadv1d_prep(ijdim, KMIN, KMAX, G->dgz, G->gzh, Vt_nq, zdis0, kcell, kcell_max,
                        kcell_min, dt);
        record_bounds(nt_rec, n_sed, 0, kcell_min, kcell_max);
#pragma omp parallel for collapse(2)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++) rhoq[IDX(k, ij)] = rhogq[QIDX(nq, k, ij)] * rgs[IDX(k, ij)];
        ora_vadv1d_getflux(ijdim, KMIN, KMAX, G->dgz, rhoq, zdis0, kcell, kcell_max, kcell_min, fprec_q);
#pragma omp parallel for collapse(2)
        for (int k = KMIN; k <= KMAX; k++)
            for (int ij = 0; ij < ijdim; ij++)
                drhoq[(size_t)nq * K2 + IDX(k, ij)] =
                    -(fprec_q[IDX(k + 1, ij)] - fprec_q[IDX(k, ij)]) * (ORA_PLANT == 11 ? G->rdgzh[k] : G->rdgz[k]);

        if (nq >= NQW_STR && nq <= NQW_END) {
#pragma omp parallel for collapse(2)
            for (int k = 0; k < KALL; k++)
                for (int ij = 0; ij < ijdim; ij++) {
                    size_t p = IDX(k, ij);
                    rhoeq[p] = rhogq[QIDX(nq, k, ij)] * rgs[p] * CVW[nq] * tem[p];
                    rhophiq[p] = rhogq[QIDX(nq, k, ij)] * rgs[p] * CONST_GRAV * R->z[p];
                    rhokin_h[p] = q[QIDX(nq, k, ij)] * rgs[p] * rhogkin_h[p];
                    rhouq[p] = rhogq[QIDX(nq, k, ij)] * rgs[p] *
                               (R->vx[p] * R->ix[ij] + R->vy[p] * R->iy[ij] + R->vz[p] * R->iz[ij]);
                    rhovq[p] = rhogq[QIDX(nq, k, ij)] * rgs[p] *
                               (R->vx[p] * R->jx[ij] + R->vy[p] * R->jy[ij] + R->vz[p] * R->jz[ij]);
                }
            ora_vadv1d_getflux(ijdim, KMIN, KMAX, G->dgz, rhoeq, zdis0, kcell, kcell_max, kcell_min, fprec_rhoe);
            ora_vadv1d_getflux(ijdim, KMIN, KMAX, G->dgz, rhophiq, zdis0, kcell, kcell_max, kcell_min, fprec_rhophi);
            ora_vadv1d_getflux(ijdim, KMIN, KMAX, G->dgz, rhokin_h, zdis0, kcell, kcell_max, kcell_min, fprec_rhokin_h);
            ora_vadv1d_getflux(ijdim, KMIN, KMAX, G->dgz, rhouq, zdis0, kcell, kcell_max, kcell_min, fprec_rhou);
            ora_vadv1d_getflux(ijdim, KMIN, KMAX, G->dgz, rhovq, zdis0, kcell, kcell_max, kcell_min, fprec_rhov);
#pragma omp parallel for collapse(2)
            for (int k = KMIN; k <= KMAX; k++)
                for (int ij = 0; ij < ijdim; ij++) {
                    size_t p = IDX(k, ij), pp = IDX(k + 1, ij);
                    drhoe[p] = drhoe[p] - (fprec_rhoe[pp] - fprec_rhoe[p]) * G->rdgz[k];
                    drhophi[p] = drhophi[p] - (fprec_rhophi[pp] - fprec_rhophi[p]) * G->rdgz[k];
                    drhokin_h[p] = drhokin_h[p] - (fprec_rhokin_h[pp] - fprec_rhokin_h[p]) * G->rdgz[k];
                    drhogu[p] = drhogu[p] - (fprec_rhou[pp] - fprec_rhou[p]) * G->rdgz[k];
                    drhogv[p] = drhogv[p] - (fprec_rhov[pp] - fprec_rhov[p]) * G->rdgz[k];
                }

            /* half level */
            for (int k = KMIN + 1; k <= KMAX - 1; k++)
                for (int ij = 0; ij < ijdim; ij++)
                    Vtermh[IDX(k, ij)] = 0.5 * (Vt_nq[IDX(k, ij)] + Vt_nq[IDX(k - 1, ij)]);
            for (int ij = 0; ij < ijdim; ij++) {
                Vtermh[IDX(KMIN - 1, ij)] = 0.0;
                Vtermh[IDX(KMIN, ij)] = 0.0;
                Vtermh[IDX(KMAX, ij)] = 0.0;
                Vtermh[IDX(KMAX + 1, ij)] = 0.0;
            }
            ora_vadv1d_prep(ijdim, KMIN + 1, KMAX, G->dgzh, GRD_gz_shift, Vtermh, zdis0h, kcellh,
                            kcellh_max, kcellh_min, dt);
            record_bounds(nt_rec, n_sed, 1, kcellh_min, kcellh_max);
            for (int k = KMIN + 1; k <= KMAX; k++)
                for (int ij = 0; ij < ijdim; ij++)
                    qh[IDX(k, ij)] = 0.5 * (q[QIDX(nq, k, ij)] + q[QIDX(nq, k - 1, ij)]) * rgsh[IDX(k, ij)];
            for (int ij = 0; ij < ijdim; ij++) {
                qh[IDX(KMIN - 1, ij)] = 0.0;
                qh[IDX(KMIN, ij)] = 0.0;
                qh[IDX(KMAX + 1, ij)] = 0.0;
            }
#pragma omp parallel for collapse(2)
            for (int k = KMIN + 1; k <= KMAX; k++)
                for (int ij = 0; ij < ijdim; ij++) {
                    rhokin_v[IDX(k, ij)] = qh[IDX(k, ij)] * rhogkin_v[IDX(k, ij)];
                    rhowq[IDX(k, ij)] = qh[IDX(k, ij)] * R->rhogw[IDX(k, ij)];
                }
            /* QUIRK A-3: the full-level kcell_max bounds the half-level flux */
            ora_vadv1d_getflux(ijdim, KMIN + 1, KMAX, G->dgzh, rhokin_v, zdis0h, kcellh, kcell_max,
                               kcellh_min, fprec_rhokin_v);
            ora_vadv1d_getflux(ijdim, KMIN + 1, KMAX, G->dgzh, rhowq, zdis0h, kcellh, kcell_max,
                               kcellh_min, fprec_rhow);
#pragma omp parallel for collapse(2)
            for (int k = KMIN + 1; k <= KMAX; k++)
                for (int ij = 0; ij < ijdim; ij++) {
                    size_t p = IDX(k, ij), pm = IDX(k - 1, ij);
                    drhokin_v[p] = drhokin_v[p] - (fprec_rhokin_v[p] - fprec_rhokin_v[pm]) * G->rdgzh[k];
                    drhogw[p] = drhogw[p] - (fprec_rhow[p] - fprec_rhow[pm]) * G->rdgzh[k];
                }

            if (nq == I_QC) {
                for (int ij = 0; ij < ijdim; ij++)
                    R->precip[0 * ijdim + ij] = R->precip[0 * ijdim + ij] - fprec_q[IDX(KMIN, ij)] / dt;
            } else if (nq == I_QR) {
                for (int ij = 0; ij < ijdim; ij++)
                    R->precip[0 * ijdim + ij] = R->precip[0 * ijdim + ij] - fprec_q[IDX(KMIN, ij)] / dt;
                for (int k = KMIN; k <= KMAX; k++)
                    for (int ij = 0; ij < ijdim; ij++) R->gprec[IDX(k, ij)] = -fprec_q[IDX(k, ij)] / dt;
                for (int ij = 0; ij < ijdim; ij++) {
                    R->gprec[IDX(KMIN - 1, ij)] = 0.0;
                    R->gprec[IDX(KMAX + 1, ij)] = 0.0;
                }
            } else if (nq == I_QI || nq == I_QS || nq == I_QG) {
                for (int ij = 0; ij < ijdim; ij++) {
                    R->precip[1 * ijdim + ij] = R->precip[1 * ijdim + ij] - fprec_q[IDX(KMIN, ij)] / dt;
                    R->precip_lh_heat[ij] = R->precip_lh_heat[ij] + fprec_q[IDX(KMIN, ij)] * LHF / dt;
                }
            }
            for (int ij = 0; ij < ijdim; ij++) {
                R->precip_rhoe[ij] = R->precip_rhoe[ij] - fprec_rhoe[IDX(KMIN, ij)] / dt;
                R->precip_rhophi[ij] = R->precip_rhophi[ij] - fprec_rhophi[IDX(KMIN, ij)] / dt;
                R->precip_rhokin[ij] = R->precip_rhokin[ij] - fprec_rhokin_h[IDX(KMIN, ij)] / dt -
                                       fprec_rhokin_v[IDX(KMIN, ij)] / dt;
            }
        }
    }

#pragma omp parallel for collapse(2)
    for (int k = 0; k < KALL; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            size_t p = IDX(k, ij);
            rhoge[p] = rhoge[p] + drhoe[p] + drhophi[p];
            rhogkin_h[p] = rhogkin_h[p] + drhokin_h[p];
            rhogkin_v[p] = rhogkin_v[p] + drhokin_v[p];
#if ORA_PLANT == 10
            R->rhogvx[p] = R->rhogvx[p] + drhogu[p] * R->jx[ij] + drhogv[p] * R->ix[ij];
#else
            R->rhogvx[p] = R->rhogvx[p] + drhogu[p] * R->ix[ij] + drhogv[p] * R->jx[ij];
#endif
            R->rhogvy[p] = R->rhogvy[p] + drhogu[p] * R->iy[ij] + drhogv[p] * R->jy[ij];
            R->rhogvz[p] = R->rhogvz[p] + drhogu[p] * R->iz[ij] + drhogv[p] * R->jz[ij];
            R->rhogw[p] = R->rhogw[p] + drhogw[p];
        }
    for (int nq = 0; nq < NQMAX; nq++) {
        if (nq >= NQW_STR && nq <= NQW_END) {
#pragma omp parallel for collapse(2)
            for (int k = 0; k < KALL; k++)
                for (int ij = 0; ij < ijdim; ij++) {
                    size_t p = IDX(k, ij);
                    double d = drhoq[(size_t)nq * K2 + p];
                    rhogq[QIDX(nq, k, ij)] = rhogq[QIDX(nq, k, ij)] + d;
                    rhog[p] = rhog[p] + d;
#if ORA_PLANT != 8
                    rhoge[p] = rhoge[p] - d * CONST_GRAV * R->z[p];
#endif
                }
        } else {
#pragma omp parallel for collapse(2)
            for (int k = 0; k < KALL; k++)
                for (int ij = 0; ij < ijdim; ij++)
                    rhogq[QIDX(nq, k, ij)] = rhogq[QIDX(nq, k, ij)] + drhoq[(size_t)nq * K2 + IDX(k, ij)];
        }
    }

    if (!kin2kin) {
    /* PRCIP_TRN_ECORRECT = "KIN2EIN" (src/problemsize.cpp:39) */
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++)
            tmp2[IDX(k, ij)] = rhogkin_h[IDX(k, ij)] + (W2Cfact[IDX(k, ij)] * rhogkin_v[IDX(k + 1, ij)] +
                                                        W2Cfact[K2 + IDX(k, ij)] * rhogkin_v[IDX(k, ij)]);
    for (int ij = 0; ij < ijdim; ij++) {
        tmp2[IDX(KMIN - 1, ij)] = 0.0;
        tmp2[IDX(KMAX + 1, ij)] = 0.0;
    }
    cnvvar_rhogkin_in(ijdim, rhog, R->rhogvx, R->rhogvy, R->rhogvz, R->rhogw, C2Wfact, W2Cfact, tmp,
                      tmp_h, tmp_v);
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++)
            rhoge[IDX(k, ij)] = rhoge[IDX(k, ij)] + (ORA_PLANT == 9 ? tmp[IDX(k, ij)] - tmp2[IDX(k, ij)]
                                                                    : tmp2[IDX(k, ij)] - tmp[IDX(k, ij)]);
    } else {
    /* "KIN2KIN" (src/mod_precip_transport.cpp:1012-1068): velocities rescaled to the
     * transported kinetic energy, horizontal on cells [kmin,kmax], vertical on walls
     * [kmin,kmax+1] (the wall density from the densities after sedimentation) */
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            const size_t p = IDX(k, ij);
            const double kin_h0 = rhogkin_h[p] / rhog[p];
            double vx_t = R->rhogvx[p] / rhog[p];
            double vy_t = R->rhogvy[p] / rhog[p];
            double vz_t = R->rhogvz[p] / rhog[p];
            const double kin_h = 0.5 * (vx_t * vx_t + vy_t * vy_t + vz_t * vz_t);
            if (kin_h > 1.0E-20) {
                vx_t = vx_t * sqrt(fabs(kin_h0 / kin_h));
                vy_t = vy_t * sqrt(fabs(kin_h0 / kin_h));
                vz_t = vz_t * sqrt(fabs(kin_h0 / kin_h));
            }
            R->rhogvx[p] = vx_t * rhog[p];
            R->rhogvy[p] = vy_t * rhog[p];
            R->rhogvz[p] = vz_t * rhog[p];
        }
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX + 1; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            const size_t p = IDX(k, ij);
            const double rhogh = C2Wfact[p] * rhog[p] + C2Wfact[K2 + p] * rhog[IDX(k - 1, ij)];
            const double kin_v0 = rhogkin_v[p] / rhogh;
            double w_t = R->rhogw[p] / rhogh;
            const double kin_v = 0.5 * pow(w_t, 2);
            if (kin_v > 1.0E-20) w_t = w_t * sqrt(fabs(kin_v0 / kin_v));
            R->rhogw[p] = w_t * rhogh;
        }
    }

#pragma omp parallel for collapse(2)
    for (int nq = 0; nq < NQMAX; nq++)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++)
                q[QIDX(nq, k, ij)] = rhogq[QIDX(nq, k, ij)] / rhog[IDX(k, ij)];
#pragma omp parallel for collapse(2)
    for (int k = 0; k < KALL; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            R->rho[IDX(k, ij)] = rhog[IDX(k, ij)] / R->gsgam2[IDX(k, ij)];
            ein[IDX(k, ij)] = rhoge[IDX(k, ij)] / rhog[IDX(k, ij)];
        }
    thrmdyn_qd(ijdim, q, R->qd);
    thrmdyn_tempre(ijdim, ein, R->rho, q, tem, R->pre);

    free(rhogkin); free(rhogkin_h); free(rhogkin_v); free(rhoq); free(rhoeq); free(rhophiq);
    free(rhokin_h); free(rhokin_v); free(rhouq); free(rhovq); free(rhowq); free(fprec_q);
    free(fprec_rhoe); free(fprec_rhophi); free(fprec_rhokin_h); free(fprec_rhokin_v);
    free(fprec_rhou); free(fprec_rhov); free(fprec_rhow); free(drhoq); free(drhoe); free(drhophi);
    free(drhokin_h); free(drhokin_v); free(drhogu); free(drhogv); free(drhogw); free(zdis0);
    free(zdis0h); free(kcell); free(kcellh); free(Vtermh); free(qh); free(ein); free(tmp);
    free(tmp_h); free(tmp_v); free(tmp2); free(C2Wfact); free(W2Cfact);
}

/* ---- satadjust_liq Newton (one point): src/mod_satadjust.cpp:828-894 -----------
 * Returns 0 when the point is not supersaturated (no iteration), n > 0 when the
 * Newton converged after n iterations, -1 when the NaN guard fired and -2 when
 * itelim = 100 iterations did not converge (the reference aborts in both cases). */
static int satadjust_liq_core(double rho, double Emoist, double qsum, double qd_loc, const double q6[6],
                              double *tem, double *qv, double *qc, double *margin);

int ora_satadjust_liq_point(double rho, double Emoist, double qsum, double qd_loc,
                            const double q6[6], double *tem, double *qv, double *qc)
{
    return satadjust_liq_core(rho, Emoist, qsum, qd_loc, q6, tem, qv, qc, NULL);
}

static int satadjust_liq_core(double rho, double Emoist, double qsum, double qd_loc, const double q6[6],
                              double *tem, double *qv, double *qc, double *margin)
{
    const double RTEM00 = 1.0 / CONST_TEM00;
    const double dtemp_criteria = g_dtemp_criteria;   /* 10^(-(RP_PREC+1)/2) = 1e-8 by default */
    const int itelim = 100;
    double qq[6];
    for (int i = 0; i < 6; i++) qq[i] = q6[i];
    double t = *tem;
    double rtem = 1.0 / (dmax(t, TEM_MIN));
    double psat = CONST_PSAT0 * (pow(t * RTEM00, CPovR_liq)) * exp(LovR_liq * (RTEM00 - rtem));
    double qsat = psat / (rho * CONST_Rvap * t);
    int code = 0;
    double d_prev = 0.0, d_last = 0.0;
    if (margin) *margin = 1e300;
    if (qsum - qsat > CONST_EPS) {
        int converged = 0, nan_break = 0, ite;
        for (ite = 0; ite < itelim; ite++) {
            rtem = 1.0 / (dmax(t, TEM_MIN));
            psat = CONST_PSAT0 * (pow(t * RTEM00, CPovR_liq)) * exp(LovR_liq * (RTEM00 - rtem));
            qsat = psat / (rho * CONST_Rvap * t);
            qq[I_QV] = qsat;
            qq[I_QC] = qsum - qsat;
            double CVtot = qd_loc * CONST_CVdry;
            for (int nq = NQW_STR; nq <= NQW_END; nq++) CVtot = CVtot + qq[nq] * CVW[nq];
            double Emoist_new = t * CVtot + qq[I_QV] * LHV;
            double dqsat_dT = (LovR_liq / (pow(t, 2)) + CVovR_liq / t) * qsat;
            double dCVtot_dT = dqsat_dT * (CVW[I_QV] - CVW[I_QC]);
            double dEmoist_dT = t * dCVtot_dT + (ORA_PLANT == 42 ? 0.0 : CVtot) + dqsat_dT * LHV;
            double dtemp = (Emoist_new - Emoist) / dEmoist_dT;
            t = t - dtemp;
            d_prev = d_last;
            d_last = dtemp;
            if (fabs(dtemp) < dtemp_criteria) {
                converged = 1;
                break;
            }
            if (t * 0.0 != 0.0) { nan_break = 1; break; }
        }
        code = converged ? ite + 1 : (nan_break ? -1 : -2);
        if (margin) *margin = newton_margin(ite + 1, converged, d_prev, d_last, dtemp_criteria);
    }
    *tem = t;
    *qv = qq[I_QV];
    *qc = qq[I_QC];
    return code;
}

/* One point of satadjust_all (src/mod_satadjust.cpp:620-769): the mixed-phase Newton
 * with the liquid fraction alpha = clamp((T - LL) / (UL - LL), 0, 1).  Returns 0 (not
 * supersaturated), the iteration count, -1 (NaN) or -2 (itelim); writes the final
 * tem and the last iterate's q(QV), q(QC), q(QI). */
static int satadjust_all_core(double rho, double Emoist, double qsum, double qd_loc, const double q6[6],
                              double UL, double LL, double *tem, double *qv, double *qc, double *qi,
                              double *margin);

int ora_satadjust_all_point(double rho, double Emoist, double qsum, double qd_loc, const double q6[6],
                            double UL, double LL, double *tem, double *qv, double *qc, double *qi)
{
    return satadjust_all_core(rho, Emoist, qsum, qd_loc, q6, UL, LL, tem, qv, qc, qi, NULL);
}

static int satadjust_all_core(double rho, double Emoist, double qsum, double qd_loc, const double q6[6],
                              double UL, double LL, double *tem, double *qv, double *qc, double *qi,
                              double *margin)
{
    const double RTEM00 = 1.0 / CONST_TEM00;
    const double dtemp_criteria = g_dtemp_criteria;
    double qq[6];
    for (int i = 0; i < 6; i++) qq[i] = q6[i];
    double t = *tem;
    double alpha = (t - LL) / (UL - LL);
    alpha = dmin(dmax(alpha, 0.0), 1.0);
    double rtem = 1.0 / (dmax(t, TEM_MIN));
    double psatl = CONST_PSAT0 * pow((t * RTEM00), CPovR_liq) * exp(LovR_liq * (RTEM00 - rtem));
    double psati = CONST_PSAT0 * pow((t * RTEM00), CPovR_ice) * exp(LovR_ice * (RTEM00 - rtem));
    double psat = psatl * alpha + psati * (1.0 - alpha);
    double qsat = psat / (rho * CONST_Rvap * t);
    int code = 0;
    double d_prev = 0.0, d_last = 0.0;
    if (margin) *margin = 1e300;
    if (qsum - qsat > CONST_EPS) {
        int converged = 0, nan_break = 0, ite;
        for (ite = 0; ite < 100; ite++) {
            alpha = (t - LL) / (UL - LL);
            alpha = dmin(dmax(alpha, 0.0), 1.0);
            rtem = 1.0 / (dmax(t, TEM_MIN));
            psatl = CONST_PSAT0 * pow((t * RTEM00), CPovR_liq) * exp(LovR_liq * (RTEM00 - rtem));
            psati = CONST_PSAT0 * pow((t * RTEM00), CPovR_ice) * exp(LovR_ice * (RTEM00 - rtem));
            psat = psatl * alpha + psati * (1.0 - alpha);
            double qsatl = psatl / (rho * CONST_Rvap * t);
            double qsati = psati / (rho * CONST_Rvap * t);
            qsat = psat / (rho * CONST_Rvap * t);
            qq[I_QV] = qsat;
            qq[I_QC] = (qsum - qsat) * alpha;
            qq[I_QI] = (qsum - qsat) * (1.0 - alpha);
            double CVt = qd_loc * CONST_CVdry;
            for (int nq = NQW_STR; nq <= NQW_END; nq++) CVt = CVt + qq[nq] * CVW[nq];
            double Emoist_new = t * CVt + qq[I_QV] * LHV - qq[I_QI] * LHF;
            double lim1 = 0.5 + copysign(0.5, UL - t);
            double lim2 = 0.5 + copysign(0.5, t - LL);
            double dalpha_dT = lim1 * lim2 / (UL - LL);
            double dqsatl_dT = (LovR_liq / (pow(t, 2)) + CVovR_liq / t) * qsatl;
            double dqsati_dT = (LovR_ice / (pow(t, 2)) + CVovR_ice / t) * qsati;
            double dqsat_dT = qsatl * dalpha_dT + dqsatl_dT * alpha - qsati * dalpha_dT +
                              dqsati_dT * (1.0 - alpha);
            double dqc_dT = (qsum - qsat) * dalpha_dT - dqsat_dT * alpha;
            double dqi_dT = -(qsum - qsat) * dalpha_dT - dqsat_dT * (1.0 - alpha);
            double dCVtot_dT = dqsat_dT * CVW[I_QV] + dqc_dT * CVW[I_QC] + dqi_dT * CVW[I_QI];
            double dEmoist_dT = t * dCVtot_dT + CVt + dqsat_dT * LHV - (ORA_PLANT == 43 ? 0.0 : dqi_dT * LHF);
            double dtemp = (Emoist_new - Emoist) / dEmoist_dT;
            t = t - dtemp;
            d_prev = d_last;
            d_last = dtemp;
            if (fabs(dtemp) < dtemp_criteria) { converged = 1; break; }
            if (t * 0.0 != 0.0) { nan_break = 1; break; }
        }
        code = converged ? ite + 1 : (nan_break ? -1 : -2);
        if (margin) *margin = newton_margin(ite + 1, converged, d_prev, d_last, dtemp_criteria);
    }
    *tem = t;
    *qv = qq[I_QV];
    *qc = qq[I_QC];
    *qi = qq[I_QI];
    return code;
}

/* ---- SATURATION_adjustment: src/mod_satadjust.cpp:274-362, satadjust_liq :771-896,
 *      satadjust_all :620-769 ---------------------------------------------------- */
static double g_sat_ulimit = 273.15, g_sat_llimit = 233.15;  /* SATURATION_Setrange state */

static void saturation_adjustment(int ijdim, double *rhog, double *rhoge, double *rhogq, double *tem,
                                  double *q, double *qd, double *gsgam2, int ice_adjust,
                                  nsw6_report *rep, int l)
{
    const size_t K2 = (size_t)KALL * ijdim;
    double *ein_moist = zalloc(K2), *qsum = zalloc(K2), *CVtot = zalloc(K2), *rho = zalloc(K2);
    double *qd_loc = zalloc(K2);
    if (ice_adjust) {
#pragma omp parallel for collapse(2)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++) {
                size_t p = IDX(k, ij);
                ein_moist[p] = rhoge[p] / rhog[p] + q[QIDX(I_QV, k, ij)] * LHV - q[QIDX(I_QI, k, ij)] * LHF;
                qsum[p] = q[QIDX(I_QV, k, ij)] + q[QIDX(I_QC, k, ij)] + q[QIDX(I_QI, k, ij)];
                q[QIDX(I_QV, k, ij)] = qsum[p];
                q[QIDX(I_QC, k, ij)] = 0.0;
                q[QIDX(I_QI, k, ij)] = 0.0;
            }
    } else {
#pragma omp parallel for collapse(2)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++) {
                size_t p = IDX(k, ij);
                ein_moist[p] = rhoge[p] / rhog[p] + (ORA_PLANT == 12 ? 0.0 : q[QIDX(I_QV, k, ij)] * LHV);
                qsum[p] = q[QIDX(I_QV, k, ij)] + q[QIDX(I_QC, k, ij)];
                if (ORA_PLANT == 13 && (k < KMIN || k > KMAX)) continue;
                q[QIDX(I_QV, k, ij)] = qsum[p];
                q[QIDX(I_QC, k, ij)] = 0.0;
            }
    }
    thrmdyn_cv(ijdim, qd, q, CVtot);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < KALL; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            size_t p = IDX(k, ij);
            rho[p] = rhog[p] / gsgam2[p];
            tem[p] = (ein_moist[p] - q[QIDX(I_QV, k, ij)] * LHV) / CVtot[p];
        }

    thrmdyn_qd(ijdim, q, qd_loc);
    int n_nc = 0, n_nan = 0, fk = -1, fij = -1;
    if (!ice_adjust) {
#pragma omp parallel for collapse(2) reduction(+ : n_nc, n_nan)
        for (int k = KMIN; k <= KMAX; k++)
            for (int ij = 0; ij < ijdim; ij++) {
                size_t p = IDX(k, ij);
                double q6[6], t = tem[p], qv, qc;
                for (int nq = 0; nq < 6; nq++) q6[nq] = q[QIDX(nq, k, ij)];
                qv = q6[I_QV];
                qc = q6[I_QC];
                double mg;
                int ite = satadjust_liq_core(rho[p], ein_moist[p], qsum[p], qd_loc[p], q6, &t, &qv, &qc, &mg);
                if (g_newton_rec) g_newton_rec[(size_t)l * K2 + p] = mg;
                tem[p] = t;
                q[QIDX(I_QV, k, ij)] = qv;
                q[QIDX(I_QC, k, ij)] = qc;
                if (ite < 0) {
                    if (ite == -1) n_nan++; else n_nc++;
#pragma omp critical
                    {
                        if (fk < 0 || k < fk || (k == fk && ij < fij)) { fk = k; fij = ij; }
                    }
                }
            }
    } else {
        const double UL = g_sat_ulimit, LL = g_sat_llimit;
#pragma omp parallel for collapse(2) reduction(+ : n_nc, n_nan)
        for (int k = KMIN; k <= KMAX; k++)
            for (int ij = 0; ij < ijdim; ij++) {
                size_t p = IDX(k, ij);
                double q6[6], t = tem[p], qv, qc, qi;
                for (int nq = 0; nq < 6; nq++) q6[nq] = q[QIDX(nq, k, ij)];
                qv = q6[I_QV];
                qc = q6[I_QC];
                qi = q6[I_QI];
                double mg;
                int ite = satadjust_all_core(rho[p], ein_moist[p], qsum[p], qd_loc[p], q6, UL, LL, &t, &qv,
                                             &qc, &qi, &mg);
                if (g_newton_rec) g_newton_rec[(size_t)l * K2 + p] = mg;
                tem[p] = t;
                q[QIDX(I_QV, k, ij)] = qv;
                q[QIDX(I_QC, k, ij)] = qc;
                q[QIDX(I_QI, k, ij)] = qi;
                if (ite < 0) {
                    if (ite == -1) n_nan++; else n_nc++;
#pragma omp critical
                    {
                        if (fk < 0 || k < fk || (k == fk && ij < fij)) { fk = k; fij = ij; }
                    }
                }
            }
    }
    if (rep && (n_nc || n_nan)) {
        rep->n_not_converged += n_nc;
        rep->n_nan += n_nan;
        if (rep->first_l < 0) { rep->first_l = l; rep->first_k = fk; rep->first_ij = fij; }
    }

#pragma omp parallel for collapse(2)
    for (int k = 0; k < KALL; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            rhogq[QIDX(I_QV, k, ij)] = rhog[IDX(k, ij)] * q[QIDX(I_QV, k, ij)];
            rhogq[QIDX(I_QC, k, ij)] = rhog[IDX(k, ij)] * q[QIDX(I_QC, k, ij)];
        }
    if (ice_adjust) {
#pragma omp parallel for collapse(2)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++)
                rhogq[QIDX(I_QI, k, ij)] = rhog[IDX(k, ij)] * q[QIDX(I_QI, k, ij)];
    }
    thrmdyn_cv(ijdim, qd, q, CVtot);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < KALL; k++)
        for (int ij = 0; ij < ijdim; ij++)
            rhoge[IDX(k, ij)] = rhog[IDX(k, ij)] * tem[IDX(k, ij)] * CVtot[IDX(k, ij)];
    free(ein_moist); free(qsum); free(CVtot); free(rho); free(qd_loc);
    (void)K2;
}

/* ---- mp_nsw6_new::init + compute: src/mod_mp_nsw6.cpp:439-674 ------------------ */
static void mp_nsw6(const nsw6_params *P, const nsw6_grid *G, int ijdim, region_ptrs *R,
                    double *re_liquid, double *re_cld, double *rctop, double *rwtop, double *tctop,
                    double dt, int ice_adjust_override, nsw6_report *rep, int l)
{
    const size_t K2 = (size_t)KALL * ijdim;
    double *drhogq = zalloc((size_t)NQMAX * K2);
    double *psatl = zalloc(K2), *qsatl = zalloc(K2), *psati = zalloc(K2), *qsati = zalloc(K2);
    double *Nc = zalloc(K2), *a1 = zalloc(K2), *a2 = zalloc(K2), *ma2 = zalloc(K2);
    double *Vt = zalloc((size_t)NQMAX * K2), *cva = zalloc(K2), *rgs = zalloc(K2), *rgsh = zalloc(K2);
    double *rctop_cld = zalloc((size_t)ijdim), *tctop_cld = zalloc((size_t)ijdim);
    const int intended = P->cloudtop_intended != 0.0;
    double *csw = intended ? zalloc(K2) : NULL;

    /* init() */
    negative_filter(ijdim, R->rhog, R->rhoge, R->rhogq, R->rho, R->tem, R->pre, R->q, R->gsgam2);
    if (P->opt_indir != 0.0) {
        /* TagAerosolT (src/mod_mp_nsw6.cpp:676-679): Kokkos::max(UNCCN*1e-6, Nc_def) */
#pragma omp parallel for
        for (size_t i = 0; i < K2; i++) {
            const double a = R->unccn[i] * 1.0E-6;
            Nc[i] = (a < P->Nc_def) ? P->Nc_def : a;
        }
    } else {
#pragma omp parallel for
        for (size_t i = 0; i < K2; i++) Nc[i] = P->Nc_def;       /* TagAerosolF */
    }
#pragma omp parallel for
    for (size_t i = 0; i < K2; i++) {
        psatl[i] = ora_psat_liq(R->tem[i]);
        psati[i] = ora_psat_ice(R->tem[i]);
    }
#pragma omp parallel for
    for (size_t i = 0; i < K2; i++) {                            /* TagSATWater */
        qsatl[i] = psatl[i] / (R->rho[i] * CONST_Rvap * R->tem[i]);
        qsati[i] = psati[i] / (R->rho[i] * CONST_Rvap * R->tem[i]);
    }
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++)
            ora_bergeron_point(R->tem[IDX(k, ij)], &a1[IDX(k, ij)], &a2[IDX(k, ij)], &ma2[IDX(k, ij)]);

    /* compute(): TagBigLoop over k in [kmin,kmax] */
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            size_t p = IDX(k, ij);
            double q6[6];
            ora_bigloop_out o;
            for (int nq = 0; nq < 6; nq++) q6[nq] = R->q[QIDX(nq, k, ij)];
            ora_bigloop_point(P, dt, R->rho[p], R->tem[p], q6, qsatl[p], qsati[p], R->pre[p],
                              R->rhog[p], Nc[p], a1[p], a2[p], ma2[p], &o);
            for (int nq = 0; nq < 6; nq++) drhogq[(size_t)nq * K2 + p] = o.drhogq[nq];
            Vt[(size_t)I_QR * K2 + p] = o.Vt[I_QR];
            Vt[(size_t)I_QI * K2 + p] = o.Vt[I_QI];
            Vt[(size_t)I_QS * K2 + p] = o.Vt[I_QS];
            Vt[(size_t)I_QG * K2 + p] = o.Vt[I_QG];
            re_liquid[p] = o.rceff;
            re_cld[p] = o.rceff_cld;
            if (intended) csw[p] = o.cloud_sw;
            /* QUIRK A-10: every k writes the cloud-top row; serial semantics (k = kmax wins) */
            else if (k == KMAX) {
                rctop[ij] = o.rctop;
                tctop[ij] = o.tctop;
                rctop_cld[ij] = o.rctop_cld;
                tctop_cld[ij] = o.tctop_cld;
            }
        }
    /* TagUpdateRHOGQ, TagUpdateRHOGE */
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            size_t p = IDX(k, ij);
            for (int nq = 0; nq < 6; nq++) R->rhogq[QIDX(nq, k, ij)] += drhogq[(size_t)nq * K2 + p];
            R->rhoge[p] = R->rhoge[p] - LHV * drhogq[(size_t)I_QV * K2 + p] +
                          LHF * drhogq[(size_t)I_QI * K2 + p] + LHF * drhogq[(size_t)I_QS * K2 + p] +
                          LHF * drhogq[(size_t)I_QG * K2 + p];
        }
    if (intended) {
        /* cloud top as a serial update over the levels, bottom to top:
         * x = sw*value + (1-sw)*x from x = UNDEF, so the highest cloudy level wins */
        for (int ij = 0; ij < ijdim; ij++) {
            double rct = CONST_UNDEF, tct = CONST_UNDEF;
            for (int k = KMIN; k <= KMAX; k++) {
                const double s = csw[IDX(k, ij)];
                rct = (s)*re_liquid[IDX(k, ij)] + (1.0 - s) * rct;
                tct = (s)*R->tem[IDX(k, ij)] + (1.0 - s) * tct;
            }
            rctop[ij] = rct;
            tctop[ij] = tct;
        }
        free(csw);
    }
    /* TagUpdateCLD */
    for (int ij = 0; ij < ijdim; ij++) {
        re_liquid[IDX(KMIN - 1, ij)] = 0.0;
        re_cld[IDX(KMIN - 1, ij)] = 0.0;
        re_liquid[IDX(KMAX + 1, ij)] = 0.0;
        re_cld[IDX(KMAX + 1, ij)] = 0.0;
        double sw = (0.5 + copysign(0.5, tctop[ij] - CONST_TEM00));
        rwtop[ij] = (sw)*rctop[ij] + (1.0 - sw) * CONST_UNDEF;
    }
    /* TagPrecip */
    for (int nq = 0; nq < NQMAX; nq++)
        for (int ij = 0; ij < ijdim; ij++) {
            Vt[(size_t)nq * K2 + IDX(KMIN - 1, ij)] = 0.0;
            Vt[(size_t)nq * K2 + IDX(KMAX + 1, ij)] = 0.0;
        }
    /* TagUpdateMassConc */
#pragma omp parallel for collapse(2)
    for (int nq = NQW_STR; nq <= NQW_END; nq++)
        for (int k = KMIN; k <= KMAX; k++)
            for (int ij = 0; ij < ijdim; ij++)
                R->q[QIDX(nq, k, ij)] = R->rhogq[QIDX(nq, k, ij)] / R->rhog[IDX(k, ij)];
    thrmdyn_qd(ijdim, R->q, R->qd);
    thrmdyn_cv(ijdim, R->qd, R->q, cva);
    /* TagUpdateTEM */
#pragma omp parallel for collapse(2)
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            size_t p = IDX(k, ij);
            R->tem[p] = R->rhoge[p] / (R->rhog[p] * cva[p]);
            rgs[p] = R->gam2[p] / R->gsgam2[p];
            rgsh[p] = R->gam2h[p] / R->gsgam2h[p];
        }
    /* preciptation_flag (src/mod_mp_nsw6.cpp:593-600): QR, QS, QG; QI too under "4WATER" */
    int flags[NQMAX] = {0, 0, 1, P->four_water != 0.0, 1, 1};
    precip_transport_new(G, ijdim, R, Vt, flags, rgs, rgsh, dt, P->kin2kin != 0.0);

    /* src/mod_mp_nsw6.cpp:647-660: OPT_EXPLICIT_ICEGEN (sw_expice = 1) takes
     * satadjust_liq with SATURATION_Setrange(100, 90); otherwise satadjust_all with
     * SATURATION_Setrange(273.16, 233.16).  ice_adjust_override 0 / 1 forces a path. */
    int ice_adjust = (P->sw_expice != 1.0);
    if (ice_adjust_override == 0) ice_adjust = 0;
    if (ice_adjust_override == 1) ice_adjust = 1;
    g_sat_ulimit = ice_adjust ? 273.16 : 100.0;
    g_sat_llimit = ice_adjust ? 233.16 : 90.0;
    g_dtemp_criteria = P->satadjust_dtemp_criteria;
    saturation_adjustment(ijdim, R->rhog, R->rhoge, R->rhogq, R->tem, R->q, R->qd, R->gsgam2,
                          ice_adjust, rep, l);
    negative_filter(ijdim, R->rhog, R->rhoge, R->rhogq, R->rho, R->tem, R->pre, R->q, R->gsgam2);

    free(drhogq); free(psatl); free(qsatl); free(psati); free(qsati); free(Nc); free(a1); free(a2);
    free(ma2); free(Vt); free(cva); free(rgs); free(rgsh); free(rctop_cld); free(tctop_cld);
}

/* Region l of the arrays of nsw6_fields. */
static region_ptrs region_of(const nsw6_fields *f, int l, int ijdim)
{
    const size_t K2 = (size_t)KALL * ijdim;
    region_ptrs R;
    size_t o2 = (size_t)l * K2, o3 = (size_t)l * NQMAX * K2, o1 = (size_t)l * ijdim;
    R.rhog = f->rhog + o2; R.rhogvx = f->rhogvx + o2; R.rhogvy = f->rhogvy + o2;
    R.rhogvz = f->rhogvz + o2; R.rhogw = f->rhogw + o2; R.rhoge = f->rhoge + o2;
    R.rhogq = f->rhogq + o3; R.rho = f->rho + o2; R.tem = f->tem + o2; R.pre = f->pre + o2;
    R.vx = f->vx + o2; R.vy = f->vy + o2; R.vz = f->vz + o2; R.w = f->w ? f->w + o2 : NULL; R.q = f->q + o3;
    R.unccn = f->unccn ? f->unccn + o2 : NULL;
    R.qd = f->qd + o2; R.z = f->z + o2; R.precip = f->precip + (size_t)l * 2 * ijdim;
    R.precip_rhoe = f->precip_rhoe + o1; R.precip_lh_heat = f->precip_lh_heat + o1;
    R.precip_rhophi = f->precip_rhophi + o1; R.precip_rhokin = f->precip_rhokin + o1;
    R.gprec = f->GPREC + o2; R.gsgam2 = f->gsgam2 + o2; R.gsgam2h = f->gsgam2h + o2;
    R.gam2 = f->gam2 + o2; R.gam2h = f->gam2h + o2;
    R.ix = f->ix + o1; R.iy = f->iy + o1; R.iz = f->iz + o1;
    R.jx = f->jx + o1; R.jy = f->jy + o1; R.jz = f->jz + o1;
    return R;
}

/* ---- mp_driver: src/mod_mp_driver.cpp:335-613 (MP_DIV_NUM = 1) ---------------- */
int ora_mp_driver(const nsw6_params *P, const nsw6_grid *G, int lall, int ijdim,
                  const nsw6_fields *f, double dt, double ct, int ice_adjust, nsw6_report *rep)
{
    return ora_mp_driver_div(P, G, lall, ijdim, f, dt, ct, ice_adjust, 1, rep);
}

/* The MP_DIV_NUM loop of src/mod_mp_driver.cpp:420-612: div_num NSW6 steps of
 * dt / div_num, each after zeroing the step outputs; the outputs are summed over the
 * sub-steps and scaled by fraction_mp = 1 / div_num at the end. */
int ora_mp_driver_div(const nsw6_params *P, const nsw6_grid *G, int lall, int ijdim,
                      const nsw6_fields *f, double dt, double ct, int ice_adjust, int div_num,
                      nsw6_report *rep)
{
    (void)ct;
    struct timespec t0, t1;
    nsw6_report local;
    memset(&local, 0, sizeof local);
    local.first_l = local.first_k = local.first_ij = -1;
    if (!rep) rep = &local;
    memset(rep, 0, sizeof *rep);
    rep->first_l = rep->first_k = rep->first_ij = -1;
    if (lall <= 0 || ijdim <= 0 || div_num < 1 || (P->opt_indir != 0.0 && !f->unccn)) {
        rep->status = NSW6_BAD_ARG;
        return NSW6_BAD_ARG;
    }

    const double fraction_mp = 1.0 / (double)div_num;
    const double dt_mp = dt * fraction_mp;
    const size_t K2 = (size_t)KALL * ijdim;
    double elapsed = 0.0;

    for (int l = 0; l < lall; l++) {
        region_ptrs R = region_of(f, l, ijdim);
        size_t o2 = (size_t)l * K2, o1 = (size_t)l * ijdim;

        double *precip_sum = zalloc(2 * (size_t)ijdim);
        double *prh_sum = zalloc(ijdim), *plh_sum = zalloc(ijdim), *pphi_sum = zalloc(ijdim),
               *pkin_sum = zalloc(ijdim);
        double *gdclw_sum = zalloc(K2), *gdcfrc_sum = zalloc(K2), *gprec_sum = zalloc(K2);
        double *iso1 = f->ISO1_precip + (size_t)l * 2 * ijdim, *iso2 = f->ISO2_precip + (size_t)l * 2 * ijdim;

        for (int m = 0; m < div_num; m++) {
            for (size_t i = 0; i < 2 * (size_t)ijdim; i++) { R.precip[i] = 0.0; iso1[i] = 0.0; iso2[i] = 0.0; }
            for (size_t i = 0; i < K2; i++) {
                f->GDCLW[o2 + i] = 0.0;
                f->GDCFRC[o2 + i] = 0.0;
                f->GPREC[o2 + i] = 0.0;
            }
            for (int ij = 0; ij < ijdim; ij++) {
                R.precip_rhoe[ij] = 0.0; R.precip_lh_heat[ij] = 0.0;
                R.precip_rhophi[ij] = 0.0; R.precip_rhokin[ij] = 0.0;
            }
            clock_gettime(CLOCK_MONOTONIC, &t0);
            g_bounds_l = l;
            mp_nsw6(P, G, ijdim, &R, f->re_liquid + o2, f->re_cld + o2, f->rctop + o1, f->rwtop + o1,
                    f->tctop + o1, dt_mp, ice_adjust, rep, l);
            clock_gettime(CLOCK_MONOTONIC, &t1);
            elapsed += (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
            for (size_t i = 0; i < K2; i++) f->re_solid[o2 + i] = 20.0E-6;

            for (size_t i = 0; i < 2 * (size_t)ijdim; i++) precip_sum[i] += R.precip[i];
            for (int ij = 0; ij < ijdim; ij++) {
                prh_sum[ij] += R.precip_rhoe[ij]; plh_sum[ij] += R.precip_lh_heat[ij];
                pphi_sum[ij] += R.precip_rhophi[ij]; pkin_sum[ij] += R.precip_rhokin[ij];
            }
            for (size_t i = 0; i < K2; i++) {
                gdclw_sum[i] += f->GDCLW[o2 + i];
                gdcfrc_sum[i] += f->GDCFRC[o2 + i];
                gprec_sum[i] += f->GPREC[o2 + i];
            }
        }
        for (size_t i = 0; i < 2 * (size_t)ijdim; i++) {
            R.precip[i] = precip_sum[i] * fraction_mp;
            iso1[i] = 0.0 * fraction_mp;
            iso2[i] = 0.0 * fraction_mp;
        }
        for (int ij = 0; ij < ijdim; ij++) {
            R.precip_rhoe[ij] = prh_sum[ij] * fraction_mp;
            R.precip_lh_heat[ij] = plh_sum[ij] * fraction_mp;
            R.precip_rhophi[ij] = pphi_sum[ij] * fraction_mp;
            R.precip_rhokin[ij] = pkin_sum[ij] * fraction_mp;
        }
        for (size_t i = 0; i < K2; i++) {
            f->GDCLW[o2 + i] = gdclw_sum[i] * fraction_mp;
            f->GDCFRC[o2 + i] = gdcfrc_sum[i] * fraction_mp;
            f->GPREC[o2 + i] = gprec_sum[i] * fraction_mp;
        }
        free(precip_sum); free(prh_sum); free(plh_sum); free(pphi_sum); free(pkin_sum);
        free(gdclw_sum); free(gdcfrc_sum); free(gprec_sum);
    }
    g_elapsed = elapsed;
    if (rep->n_nan) rep->status = NSW6_NAN_DETECTED;
    else if (rep->n_not_converged) rep->status = NSW6_NOT_CONVERGED;
    else rep->status = NSW6_OK;
    return rep->status;
}

/* ---- single phases of the step, for the pins of tests/ (one region, [k][ij]) ------- */
void ora_thrmdyn(int ijdim, const double *q, const double *ein, const double *rho, double *qd, double *cv,
                 double *tem, double *pre)
{
    thrmdyn_qd(ijdim, q, qd);
    thrmdyn_cv(ijdim, qd, q, cv);
    thrmdyn_tempre(ijdim, ein, rho, q, tem, pre);
}

void ora_negative_filter(int ijdim, double *rhog, double *rhoge, double *rhogq, double *rho, const double *tem,
                         double *pre, double *q, const double *gsgam2)
{
    negative_filter(ijdim, rhog, rhoge, rhogq, rho, tem, pre, q, gsgam2);
}

void ora_rhogkin(const nsw6_grid *G, int ijdim, const double *rhog, const double *rhogvx, const double *rhogvy,
                 const double *rhogvz, const double *rhogw, const double *gsgam2, const double *gsgam2h,
                 double *rhogkin, double *rhogkin_h, double *rhogkin_v)
{
    const size_t K2 = (size_t)KALL * ijdim;
    double *C2W = zalloc(2 * K2), *W2C = zalloc(2 * K2);
    c2w_w2c_fact(G, ijdim, gsgam2, gsgam2h, C2W, W2C);
    cnvvar_rhogkin_in(ijdim, rhog, rhogvx, rhogvy, rhogvz, rhogw, C2W, W2C, rhogkin, rhogkin_h, rhogkin_v);
    free(C2W);
    free(W2C);
}

/* precip_transport_new on region 0 of `f` (in place) with the fall speeds Vterm
 * [6][k][ij], after mp_driver's zeroing of the surface outputs and GPREC, with rgs and
 * rgsh as TagUpdateTEM forms them (interior levels). */
void ora_precip_transport(const nsw6_params *P, const nsw6_grid *G, int ijdim, const nsw6_fields *f,
                          const double *Vterm, double dt)
{
    const size_t K2 = (size_t)KALL * ijdim;
    region_ptrs R = region_of(f, 0, ijdim);
    for (size_t i = 0; i < 2 * (size_t)ijdim; i++) R.precip[i] = 0.0;
    for (int ij = 0; ij < ijdim; ij++) {
        R.precip_rhoe[ij] = 0.0; R.precip_lh_heat[ij] = 0.0;
        R.precip_rhophi[ij] = 0.0; R.precip_rhokin[ij] = 0.0;
    }
    for (size_t i = 0; i < K2; i++) R.gprec[i] = 0.0;
    double *rgs = zalloc(K2), *rgsh = zalloc(K2);
    for (int k = KMIN; k <= KMAX; k++)
        for (int ij = 0; ij < ijdim; ij++) {
            rgs[IDX(k, ij)] = R.gam2[IDX(k, ij)] / R.gsgam2[IDX(k, ij)];
            rgsh[IDX(k, ij)] = R.gam2h[IDX(k, ij)] / R.gsgam2h[IDX(k, ij)];
        }
    int flags[NQMAX] = {0, 0, 1, P->four_water != 0.0, 1, 1};
    precip_transport_new(G, ijdim, &R, Vterm, flags, rgs, rgsh, dt, P->kin2kin != 0.0);
    free(rgs);
    free(rgsh);
}

/* SATURATION_adjustment on one region (in place) with the range mp_nsw6 sets for the
 * scheme (src/mod_mp_nsw6.cpp:647-660). */
void ora_saturation_adjustment(const nsw6_params *P, int ijdim, double *rhog, double *rhoge, double *rhogq,
                               double *tem, double *q, double *qd, double *gsgam2)
{
    const int ice_adjust = (P->sw_expice != 1.0);
    g_sat_ulimit = ice_adjust ? 273.16 : 100.0;
    g_sat_llimit = ice_adjust ? 233.16 : 90.0;
    g_dtemp_criteria = P->satadjust_dtemp_criteria;
    saturation_adjustment(ijdim, rhog, rhoge, rhogq, tem, q, qd, gsgam2, ice_adjust, NULL, 0);
}

double ora_last_elapsed(void) { return g_elapsed; }

int ora_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
