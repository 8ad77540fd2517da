// nsw6_setup.h -- host-side preparation shared by the C ABI and the host gradient
// check: scheme coefficients as the kernels see them and the device vertical grid.
#pragma once

#include <math.h>
#include <string.h>

#include "../../include/nsw6.h"
#include "nsw6_internal.h"

namespace nsw6 {

// nsw6_params -> DevParams, plus the constants the reference evaluates per point
// (src/mod_mp_nsw6.cpp:1083,1154,1245; src/mod_satadjust.cpp:24-47,788).
inline void to_dev_params(const nsw6_params &p, DevParams &d)
{
    memset(&d, 0, sizeof d);
    d.N0r = p.N0r; d.N0s = p.N0s; d.N0g = p.N0g;
    d.Ar = p.Ar; d.As = p.As; d.Ag = p.Ag; d.Cr = p.Cr; d.Cs = p.Cs; d.Cg = p.Cg;
    d.GAM = p.GAM; d.GAM_2 = p.GAM_2; d.GAM_3 = p.GAM_3;
    d.GAM_1br = p.GAM_1br; d.GAM_2br = p.GAM_2br; d.GAM_3br = p.GAM_3br; d.GAM_3dr = p.GAM_3dr;
    d.GAM_6dr = p.GAM_6dr; d.GAM_1brdr = p.GAM_1brdr; d.GAM_5dr_h = p.GAM_5dr_h;
    d.GAM_1bs = p.GAM_1bs; d.GAM_2bs = p.GAM_2bs; d.GAM_3bs = p.GAM_3bs; d.GAM_3ds = p.GAM_3ds;
    d.GAM_1bsds = p.GAM_1bsds; d.GAM_5ds_h = p.GAM_5ds_h;
    d.GAM_1bg = p.GAM_1bg; d.GAM_3dg = p.GAM_3dg; d.GAM_1bgdg = p.GAM_1bgdg; d.GAM_5dg_h = p.GAM_5dg_h;
    d.sw_kk2000 = p.sw_kk2000; d.sw_roh2014 = p.sw_roh2014; d.ln10 = p.ln10;
    for (int i = 0; i < 10; i++) { d.coef_a[i] = p.coef_a[i]; d.coef_b[i] = p.coef_b[i]; }
    d.Eiw = p.Eiw; d.Erw = p.Erw; d.Esw = p.Esw; d.Egw = p.Egw; d.Eri = p.Eri;
    d.Esi = p.Esi; d.Egi = p.Egi; d.Esr = p.Esr; d.Egr = p.Egr; d.Egs = p.Egs;
    d.gamma_sacr = p.gamma_sacr; d.gamma_gacs = p.gamma_gacs; d.mi = p.mi;
    d.Nc_def = p.Nc_def; d.beta_saut = p.beta_saut; d.gamma_saut = p.gamma_saut;
    d.beta_gaut = p.beta_gaut; d.gamma_gaut = p.gamma_gaut; d.qicrt_saut = p.qicrt_saut;
    d.qscrt_gaut = p.qscrt_gaut;
    d.f1r = p.f1r; d.f2r = p.f2r; d.f1s = p.f1s; d.f2s = p.f2s; d.f1g = p.f1g; d.f2g = p.f2g;
    d.A_frz = p.A_frz; d.B_frz = p.B_frz;
    d.mi40 = p.mi40; d.mi50 = p.mi50; d.vti50 = p.vti50; d.Ri50 = p.Ri50;
    d.sw_expice = p.sw_expice; d.Nc_ihtr = p.Nc_ihtr; d.Di_max = p.Di_max; d.Di_a = p.Di_a;
    d.sw_constVti = p.sw_constVti; d.CONST_Vti = p.CONST_Vti;
    d.kk_aut_rate = p.kk_aut_rate; d.icenuc_coef = p.icenuc_coef;
    d.coef_dgam = p.coef_dgam; d.coef_xf = p.coef_xf;
    // constants the reference evaluates per point, computed once here
    d.ln_mi40 = log(p.mi40);
    d.ln_mi50 = log(p.mi50);
    d.Nc_pow_m179 = pow(p.Nc_def, -1.79);
    d.Dc = 0.146 - 5.964E-2 * log(p.Nc_def / 2000.0);
    d.LovR_liq = C_LHV0 / C_Rvap;
    d.LovR_ice = C_LHS0 / C_Rvap;
    d.RTEM00 = 1.0 / C_TEM00;
    d.dtemp_criteria = p.satadjust_dtemp_criteria;   // the reference: 10^(-(15+1)/2) = 1e-8
    d.rK_r = 1.0 / (p.Ar * p.N0r * p.GAM_1br);
    d.rK_s = 1.0 / (p.As * p.N0s * p.GAM_1bs);
    d.rK_g = 1.0 / (p.Ag * p.N0g * p.GAM_1bg);
    d.rAs = 1.0 / p.As;
    d.rmi = 1.0 / p.mi;
    d.rNc_ihtr = 1.0 / (p.Nc_ihtr * 1.0E+6);
    d.Vtr_fac = p.GAM_1brdr / p.GAM_1br;
    d.Vtg_fac = p.GAM_1bgdg / p.GAM_1bg;
    d.ice_adjust = (p.sw_expice != 1.0) ? 1.0 : 0.0;
    d.sat_ulimit = d.ice_adjust != 0.0 ? 273.16 : 100.0;
    d.sat_llimit = d.ice_adjust != 0.0 ? 233.16 : 90.0;
    d.cloudtop_intended = (p.cloudtop_intended != 0.0) ? 1.0 : 0.0;
    d.opt_indir = (p.opt_indir != 0.0) ? 1.0 : 0.0;
    d.kin2kin = (p.kin2kin != 0.0) ? 1.0 : 0.0;
    d.four_water = (p.four_water != 0.0) ? 1.0 : 0.0;
}

// GRD_Setup (src/mod_debug.cpp:92-187) plus GRD_gz_shift of precip_transport_new
// (src/mod_precip_transport.cpp:705-708) and the reciprocal sums used by zdis.
inline void make_grid_dev(const double *gz, const double *gzh, GridDev &gh)
{
    nsw6_grid g;
    nsw6_grid_setup(gz, gzh, &g);
    for (int k = 0; k < KALL; k++) {
        gh.gz[k] = g.gz[k]; gh.gzh[k] = g.gzh[k];
        gh.dgz[k] = g.dgz[k]; gh.dgzh[k] = g.dgzh[k];
        gh.rdgz[k] = g.rdgz[k]; gh.rdgzh[k] = g.rdgzh[k];
        gh.afact[k] = g.afact[k]; gh.bfact[k] = g.bfact[k];
        gh.cfact[k] = g.cfact[k]; gh.dfact[k] = g.dfact[k];
        gh.gz_shift[k] = 0.0;
    }
    for (int k = KMIN; k <= KMAX + 1; k++) gh.gz_shift[k] = g.gz[k - 1];
    for (int k = 0; k < KALL; k++) {
        gh.rsum_gz[k] = (k >= 1) ? 1.0 / (g.dgz[k - 1] + g.dgz[k]) : 0.0;
        gh.rsum_gzh[k] = (k >= 1) ? 1.0 / (g.dgzh[k - 1] + g.dgzh[k]) : 0.0;
    }
}

}  // namespace nsw6
