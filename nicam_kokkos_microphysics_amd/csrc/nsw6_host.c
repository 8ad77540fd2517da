/*
 * nsw6_host.c -- host-side set-up of the NSW6 step: scheme coefficients and the
 * vertical-grid metrics.  Runs once per context; nothing here is on the hot path.
 */
#include <math.h>
#include <string.h>

#include "../../include/nsw6.h"

/* Lanczos Gamma function, MISC_gammafunc (src/mod_debug.cpp:332-361). */
static double lanczos_gamma(double xx)
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

/* GlobalParams defaults (include/mod_mp_nsw6.h:16-123) updated the way
 * mp_nsw6_init does it (src/mod_mp_nsw6.cpp:276-434): Roh & Satoh (2014) snow,
 * explicit ice generation, KK2000 autoconversion, no Pgaut, Nc = Nc_ocn.  The
 * effective-radius constants follow mp_nsw6_new::init (src/mod_mp_nsw6.cpp:506-517). */
void nsw6_params_default(nsw6_params *p)
{
    static const double coef_a[10] = {5.065339, -0.062659, -3.032362, 0.029469, -0.000285,
                                      0.31255,  0.000204,  0.003199,  0.0,      -0.015952};
    static const double coef_b[10] = {0.476221, -0.015896, 0.165977, 0.007468, -0.000141,
                                      0.060366, 0.000079,  0.000594, 0.0,      -0.003577};
    const double PI = 3.141592653589793, GRAV = 9.79764, DWATR = 1000.0, dens00 = 1.28;

    memset(p, 0, sizeof(*p));
    p->N0r = 8.E+6;
    p->N0s = 3.E+6;
    p->rho_s = 100.0;
    p->rho_g = 400.0;
    p->C_d = 0.6;
    p->Cr = 130.0;
    p->Cs = 4.84;
    memcpy(p->coef_a, coef_a, sizeof coef_a);
    memcpy(p->coef_b, coef_b, sizeof coef_b);
    p->Eiw = 1.0; p->Erw = 1.0; p->Esw = 1.0; p->Egw = 1.0; p->Eri = 1.0;
    p->Esr = 1.0; p->Egr = 1.0;
    p->gamma_gacs = 90.E-3;
    p->mi = 4.19E-13;
    p->gamma_gaut = 90.E-3;
    p->qicrt_saut = 0.0;
    p->qscrt_gaut = 6.E-4;
    p->f1r = 0.78; p->f2r = 0.27; p->f1s = 0.65; p->f2s = 0.39; p->f1g = 0.78; p->f2g = 0.27;
    p->A_frz = 0.66;
    p->B_frz = 100.0;
    p->mi40 = 2.46E-10; p->mi50 = 4.80E-10; p->vti50 = 1.0; p->Ri50 = 5.E-5;
    p->Nc_ihtr = 300.0; p->Di_max = 500.E-6; p->Di_a = 11.9;

    /* mp_nsw6_init */
    p->Nc_def = 50.0;                 /* Nc_ocn */
    p->CONST_Vti = -9.9999E+30;       /* UNDEF -> sw_constVti = 0 */
    p->sw_constVti = 0.0;
    p->gamma_sacr = 6.0E-2;
    p->gamma_saut = 6.0E-2;
    p->beta_saut = 5.0E-3;
    p->beta_gaut = 0.0;               /* INC_PGAUT = false */
    p->Ar = PI * DWATR / 6.0;
    p->Ag = PI * p->rho_g / 6.0;
    p->Br = 3.0; p->Bs = 3.0; p->Bg = 3.0;
    p->Cg = sqrt((4.0 * p->rho_g * GRAV) / (3.0 * dens00 * p->C_d));
    p->Dr = 0.50; p->Ds = 0.25; p->Dg = 0.50;
    p->sw_roh2014 = 1.0;              /* Roh & Satoh (2014) overrides */
    p->N0g = 4.0E+8;
    p->As = 0.069;
    p->Bs = 2.0;
    p->Esi = 0.25;
    p->Egi = 0.0;
    p->Egs = 0.0;
    p->sw_expice = 1.0;               /* OPT_EXPLICIT_ICEGEN */
    p->sw_kk2000 = 1.0;               /* KK2000 */

    p->GAM = 1.0; p->GAM_2 = 1.0; p->GAM_3 = 2.0;
    p->GAM_1br = lanczos_gamma(1.0 + p->Br);
    p->GAM_2br = lanczos_gamma(2.0 + p->Br);
    p->GAM_3br = lanczos_gamma(3.0 + p->Br);
    p->GAM_3dr = lanczos_gamma(3.0 + p->Dr);
    p->GAM_6dr = lanczos_gamma(6.0 + p->Dr);
    p->GAM_1brdr = lanczos_gamma(1.0 + p->Br + p->Dr);
    p->GAM_5dr_h = lanczos_gamma(0.5 * (5.0 + p->Dr));
    p->GAM_1bs = lanczos_gamma(1.0 + p->Bs);
    p->GAM_2bs = lanczos_gamma(2.0 + p->Bs);
    p->GAM_3bs = lanczos_gamma(3.0 + p->Bs);
    p->GAM_3ds = lanczos_gamma(3.0 + p->Ds);
    p->GAM_1bsds = lanczos_gamma(1.0 + p->Bs + p->Ds);
    p->GAM_5ds_h = lanczos_gamma(0.5 * (5.0 + p->Ds));
    p->GAM_1bg = lanczos_gamma(1.0 + p->Bg);
    p->GAM_3dg = lanczos_gamma(3.0 + p->Dg);
    p->GAM_1bgdg = lanczos_gamma(1.0 + p->Bg + p->Dg);
    p->GAM_5dg_h = lanczos_gamma(0.5 * (5.0 + p->Dg));
    p->ln10 = log(10.0);

    p->kk_aut_rate = 1350.0;
    p->icenuc_coef = 4.92E-11;
    /* dtemp_criteria = 10^(-(RP_PREC+1)/2), integer division (src/mod_satadjust.cpp:788) */
    p->satadjust_dtemp_criteria = pow(10.0, (double)(-(15 + 1) / 2));

    {
        double Dc = 0.146 - 5.964E-2 * log(p->Nc_def / 2000.0);
        p->dgamma_a = 0.1444 / (Dc * Dc);
        p->GAM_dgam = lanczos_gamma(p->dgamma_a);
        p->GAM_dgam23 = lanczos_gamma(p->dgamma_a + 2.0 / 3.0);
        p->coef_dgam = p->GAM_dgam23 / p->GAM_dgam * pow(p->dgamma_a, (-2.0 / 3.0));
        p->coef_xf = 3.0 / 4.0 / PI / DWATR;
    }
}

/* GRD_Setup (src/mod_debug.cpp:92-187). */
void nsw6_grid_setup(const double *gz, const double *gzh, nsw6_grid *g)
{
    const int kmin = NSW6_KMIN, kmax = NSW6_KMAX, kall = NSW6_KALL;
    memset(g, 0, sizeof(*g));
    for (int k = 0; k < kall; k++) {
        g->gz[k] = gz[k];
        g->gzh[k] = gzh[k];
    }
    for (int k = kmin - 1; k <= kmax; k++) g->dgz[k] = g->gzh[k + 1] - g->gzh[k];
    g->dgz[kmax + 1] = g->dgz[kmax];
    for (int k = kmin; k <= kmax + 1; k++) g->dgzh[k] = g->gz[k] - g->gz[k - 1];
    g->dgzh[kmin - 1] = g->dgzh[kmin];
    for (int k = 0; k < kall; k++) {
        g->rdgz[k] = 1.0 / g->dgz[k];
        g->rdgzh[k] = 1.0 / g->dgzh[k];
    }
    for (int k = kmin; k <= kmax + 1; k++)
        g->afact[k] = (g->gzh[k] - g->gz[k - 1]) / (g->gz[k] - g->gz[k - 1]);
    g->afact[kmin - 1] = 1.0;
    for (int k = 0; k < kall; k++) g->bfact[k] = 1.0 - g->afact[k];
    for (int k = kmin; k <= kmax; k++)
        g->cfact[k] = (g->gz[k] - g->gzh[k]) / (g->gzh[k + 1] - g->gzh[k]);
    g->cfact[kmin - 1] = 1.0;
    g->cfact[kmax + 1] = 0.0;
    for (int k = 0; k < kall; k++) g->dfact[k] = 1.0 - g->cfact[k];
}

const char *nsw6_status_string(int status)
{
    switch (status) {
    case NSW6_OK: return "ok";
    case NSW6_NOT_CONVERGED: return "saturation adjustment did not converge in 100 iterations";
    case NSW6_NAN_DETECTED: return "saturation adjustment produced a NaN temperature";
    case NSW6_BAD_ARG: return "bad argument";
    case NSW6_HIP_ERROR: return "HIP runtime error";
    case NSW6_NO_DEVICE: return "no HIP device";
    case NSW6_UNSUPPORTED: return "not supported by the reverse mode (non-default scheme, sub-cycling, or a region needing the general sedimentation)";
    case NSW6_IO_ERROR: return "snapshot file missing, unreadable or of the wrong size";
    default: return "unknown status";
    }
}
