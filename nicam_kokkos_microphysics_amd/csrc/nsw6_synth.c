/*
 * nsw6_synth.c -- seeded synthetic NICAM column states for the NSW6 step.
 *
 * The reference drives mp_driver from NICAM snapshot files (src/main.cpp:128-174),
 * which are not available; this generator produces self-consistent states of the
 * same shape and layout instead.  Every random number is a pure function of
 * (seed, stream, l, ij, k) (SplitMix64 finaliser as a counter-based hash), so any
 * subset of columns is bit-identical to the same columns of a full-size run.
 *
 * Profiles (per column):
 *   grid     : power-law walls dz_k = 80 + 920 (k/93)^p, k = 0..93, p solved so the
 *              94 layers span 0-40 km exactly; gzh[1] = 0, halo wall gzh[0] = -dz_0,
 *              gz = cell midpoints, gz[95] one top-cell height above gz[94] -- the
 *              reference vgrid convention (run/data/vgrid/GRD_gz*.dat).
 *   T        : 300 K + column offset U[-2,2] K, -6.5 K/km up to 12 km, isothermal
 *              above (the tropopause keeps the stated lapse rate and height).
 *   p        : hydrostatic from 1000 hPa through that T(z) with dry R.
 *   qv       : RH * qsat(T) with qsat from psat_liq and the reference's equation of
 *              state; RH ~ U[rh_lo, rh_hi] per point (supersaturated columns use
 *              U[1.0, 1.2]).
 *   qc / qi  : lognormal (median 1e-4, sigma 1, cap 1e-3) where RH > 0.95, warm /
 *              cold side of 0 C.
 *   qr       : lognormal (median 5e-5, cap 5e-4) below the 0 C level;
 *   qs, qg   : the same above the 0 C level and below the tropopause.
 *   towers   : deep convective columns, RH ~ U[1.0,1.1] and qc or qi lognormal
 *              (median 1e-3, cap 3e-3) over 2-12 km.
 *   rain shafts: columns of heavy precipitation: qr lognormal (median 2e-3, sigma 0.5,
 *              cap 6e-3) below the 0 C level, qg (median 1e-3, cap 3e-3) above it up
 *              to the tropopause.  Rain of 2-6 g/kg falls about 6-7 m/s, i.e. through
 *              more than one 72-80 m bottom layer in dt = 15 s.
 *   winds    : u, v ~ U[-20,20] m/s rotated to Cartesian with random (lon, lat);
 *              w ~ U[-2,2] m/s at walls 2..94, 0 at walls 0, 1, 95.
 *   metrics  : gsgam2, gsgam2h, gam2, gam2h = 1 + U[-1e-3, 1e-3].
 *   consistency: rho from pre = rho T (qd Rdry + qv Rvap); rhog = rho gsgam2;
 *              rhogq = rhog q; rhoge = rhog cv T (THRMDYN_cv order); rhogv = rhog v;
 *              rhogw = (C2W0 rhog_k + C2W1 rhog_{k-1}) w.  Halo levels 0 and 95 copy
 *              levels 1 and 94.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/nsw6.h"

#define KALL 96

typedef struct nsw6_synth_config {
    uint64_t seed;
    double rh_lo, rh_hi;      /* default 0.5, 1.05                                */
    double frac_supersat;     /* fraction of columns with RH ~ U[1.0,1.2]          */
    double frac_towers;       /* fraction of deep convective columns               */
    double tsfc;              /* surface temperature (300 K)                       */
    double frac_rainshafts;   /* fraction of heavy-precipitation columns           */
} nsw6_synth_config;

static const double Rdry = 287.04, Rvap = 461.46, CVdry = 717.56, GRAV = 9.79764;
static const double PSAT0 = 610.7, TEM00 = 273.15, LHV0 = 2501000.0;

static inline uint64_t mix64(uint64_t x)
{
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27; x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}
/* uniform in [0,1) from (seed, stream, l, ij, k) */
static inline double urand(uint64_t seed, uint32_t stream, uint32_t l, uint32_t ij, uint32_t k)
{
    uint64_t h = mix64(seed + 0x9e3779b97f4a7c15ULL * (uint64_t)(stream + 1));
    h = mix64(h ^ ((uint64_t)l << 40) ^ ((uint64_t)ij << 8) ^ (uint64_t)k);
    h = mix64(h + 0x632be59bd9b4e019ULL);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
static inline double lognormal(double u1, double u2, double median, double sigma, double cap)
{
    double a = u1 < 1e-300 ? 1e-300 : u1;
    double n = sqrt(-2.0 * log(a)) * cos(6.283185307179586 * u2);
    double v = median * exp(sigma * n);
    return v < cap ? v : cap;
}

enum { S_TCOL = 1, S_RH, S_QC1, S_QC2, S_QR1, S_QR2, S_QS1, S_QS2, S_QG1, S_QG2, S_U, S_V, S_W,
       S_LON, S_LAT, S_G1, S_G2, S_G3, S_G4, S_SUPER, S_TOWER, S_SHAFT, S_QH1, S_QH2 };

void nsw6_synth_default_config(nsw6_synth_config *c)
{
    c->seed = 20240101ULL;
    c->rh_lo = 0.5;
    c->rh_hi = 1.05;
    c->frac_supersat = 0.0;
    c->frac_towers = 0.0;
    c->tsfc = 300.0;
    c->frac_rainshafts = 0.0;
}

/* Power-law grid: 94 layers over 0..40 km with dz from 80 m to 1 km. */
void nsw6_synth_grid(double *gz, double *gzh)
{
    double lo = 0.5, hi = 4.0;
    for (int it = 0; it < 200; it++) {
        double p = 0.5 * (lo + hi), s = 0.0;
        for (int k = 0; k < 94; k++) s += 80.0 + 920.0 * pow(k / 93.0, p);
        if (s > 40000.0) lo = p; else hi = p;
    }
    double p = 0.5 * (lo + hi);
    double dz[KALL];
    for (int k = 1; k <= 94; k++) dz[k] = 80.0 + 920.0 * pow((k - 1) / 93.0, p);
    dz[0] = dz[1];
    dz[95] = dz[94];
    gzh[1] = 0.0;
    gzh[0] = -dz[0];
    for (int k = 1; k <= 94; k++) gzh[k + 1] = gzh[k] + dz[k];
    for (int k = 0; k <= 95; k++) {
        double top = (k < 95) ? gzh[k + 1] : gzh[95] + dz[95];
        gz[k] = 0.5 * (gzh[k] + top);
    }
}

static double psat_liq(double t)
{
    const double RT0 = 1.0 / TEM00, LovR = LHV0 / Rvap;
    double rt = 1.0 / (t > 10.0 ? t : 10.0);
    return PSAT0 * pow(t * RT0, 0.0) * exp(LovR * (RT0 - rt));
}

/* Fill host arrays (layout of nsw6.h) for lall regions x ijdim columns.  Column
 * identity for the RNG is (l + l_offset, ij + ij_offset).  Only the fields that
 * mp_driver reads are set to meaningful values; outputs are zeroed. */
int nsw6_synth_fill(const nsw6_synth_config *cfg, const double *gz, const double *gzh,
                    int lall, int ijdim, int l_offset, int ij_offset, const nsw6_fields *f)
{
    if (!cfg || !gz || !gzh || !f || lall <= 0 || ijdim <= 0) return NSW6_BAD_ARG;
    const size_t K2 = (size_t)KALL * ijdim;
    nsw6_grid g;
    nsw6_grid_setup(gz, gzh, &g);

#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l < lall; l++)
        for (int ij = 0; ij < ijdim; ij++) {
            const uint32_t L = (uint32_t)(l + l_offset), IJ = (uint32_t)(ij + ij_offset);
            const uint64_t sd = cfg->seed;
            const size_t o2 = (size_t)l * K2, o3 = (size_t)l * 6 * K2, o1 = (size_t)l * ijdim;
            double T[KALL], P[KALL], Q[6][KALL], RH[KALL], RHO[KALL], U[KALL], V[KALL], W[KALL];
            double G1[KALL], G2[KALL], G3[KALL], G4[KALL];
            int supersat = urand(sd, S_SUPER, L, IJ, 0) < cfg->frac_supersat;
            int tower = urand(sd, S_TOWER, L, IJ, 0) < cfg->frac_towers;
            int shaft = urand(sd, S_SHAFT, L, IJ, 0) < cfg->frac_rainshafts;
            double tsfc = cfg->tsfc + 4.0 * (urand(sd, S_TCOL, L, IJ, 0) - 0.5);
            const double gamma = 0.0065, ztp = 12000.0;
            const double ttp = tsfc - gamma * ztp;
            const double ptp = 100000.0 * pow(ttp / tsfc, GRAV / (Rdry * gamma));
            for (int k = 1; k <= 94; k++) {
                double z = gz[k];
                if (z <= ztp) {
                    T[k] = tsfc - gamma * z;
                    P[k] = 100000.0 * pow(T[k] / tsfc, GRAV / (Rdry * gamma));
                } else {
                    T[k] = ttp;
                    P[k] = ptp * exp(-GRAV * (z - ztp) / (Rdry * ttp));
                }
                double rh;
                if (supersat) rh = 1.0 + 0.2 * urand(sd, S_RH, L, IJ, k);
                else if (tower && z >= 2000.0 && z <= ztp) rh = 1.0 + 0.1 * urand(sd, S_RH, L, IJ, k);
                else rh = cfg->rh_lo + (cfg->rh_hi - cfg->rh_lo) * urand(sd, S_RH, L, IJ, k);
                RH[k] = rh;
                for (int n = 0; n < 6; n++) Q[n][k] = 0.0;
                int warm = T[k] > TEM00;
                double u1 = urand(sd, S_QC1, L, IJ, k), u2 = urand(sd, S_QC2, L, IJ, k);
                if (tower && z >= 2000.0 && z <= ztp) {
                    double v = lognormal(u1, u2, 1.0e-3, 1.0, 3.0e-3);
                    if (warm) Q[1][k] = v; else Q[3][k] = v;
                } else if (rh > 0.95) {
                    double v = lognormal(u1, u2, 1.0e-4, 1.0, 1.0e-3);
                    if (warm) Q[1][k] = v; else Q[3][k] = v;
                }
                if (warm) {
                    Q[2][k] = lognormal(urand(sd, S_QR1, L, IJ, k), urand(sd, S_QR2, L, IJ, k), 5.0e-5, 1.0, 5.0e-4);
                } else if (z <= ztp) {
                    Q[4][k] = lognormal(urand(sd, S_QS1, L, IJ, k), urand(sd, S_QS2, L, IJ, k), 5.0e-5, 1.0, 5.0e-4);
                    Q[5][k] = lognormal(urand(sd, S_QG1, L, IJ, k), urand(sd, S_QG2, L, IJ, k), 5.0e-5, 1.0, 5.0e-4);
                }
                if (shaft) {
                    const double h1 = urand(sd, S_QH1, L, IJ, k), h2 = urand(sd, S_QH2, L, IJ, k);
                    if (warm) Q[2][k] = lognormal(h1, h2, 2.0e-3, 0.5, 6.0e-3);
                    else if (z <= ztp) Q[5][k] = lognormal(h1, h2, 1.0e-3, 0.5, 3.0e-3);
                }
                /* rho from pre = rho T (qd Rdry + qv Rvap) with qv = RH psat / (rho Rv T) */
                double qd = 1.0;
                for (int n = 1; n <= 5; n++) qd -= Q[n][k];
                double ps = psat_liq(T[k]);
                RHO[k] = (P[k] - rh * ps) / (T[k] * qd * Rdry);
                Q[0][k] = rh * ps / (RHO[k] * Rvap * T[k]);
                U[k] = 40.0 * (urand(sd, S_U, L, IJ, k) - 0.5);
                V[k] = 40.0 * (urand(sd, S_V, L, IJ, k) - 0.5);
                G1[k] = 1.0 + 2.0e-3 * (urand(sd, S_G1, L, IJ, k) - 0.5);
                G2[k] = 1.0 + 2.0e-3 * (urand(sd, S_G2, L, IJ, k) - 0.5);
                G3[k] = 1.0 + 2.0e-3 * (urand(sd, S_G3, L, IJ, k) - 0.5);
                G4[k] = 1.0 + 2.0e-3 * (urand(sd, S_G4, L, IJ, k) - 0.5);
            }
            /* halo levels copy their neighbours */
            T[0] = T[1]; T[95] = T[94]; P[0] = P[1]; P[95] = P[94];
            RH[0] = RH[1]; RH[95] = RH[94]; RHO[0] = RHO[1]; RHO[95] = RHO[94];
            U[0] = U[1]; U[95] = U[94]; V[0] = V[1]; V[95] = V[94];
            G1[0] = G1[1]; G1[95] = G1[94]; G2[0] = G2[1]; G2[95] = G2[94];
            G3[0] = G3[1]; G3[95] = G3[94]; G4[0] = G4[1]; G4[95] = G4[94];
            for (int n = 0; n < 6; n++) { Q[n][0] = Q[n][1]; Q[n][95] = Q[n][94]; }
            for (int k = 0; k < KALL; k++)
                W[k] = (k >= 2 && k <= 94) ? 4.0 * (urand(sd, S_W, L, IJ, k) - 0.5) : 0.0;

            double lon = 6.283185307179586 * urand(sd, S_LON, L, IJ, 0);
            double lat = 3.141592653589793 * (urand(sd, S_LAT, L, IJ, 0) - 0.5);
            double ix = -sin(lon), iy = cos(lon), iz = 0.0;
            double jx = -sin(lat) * cos(lon), jy = -sin(lat) * sin(lon), jz = cos(lat);
            f->ix[o1 + ij] = ix; f->iy[o1 + ij] = iy; f->iz[o1 + ij] = iz;
            f->jx[o1 + ij] = jx; f->jy[o1 + ij] = jy; f->jz[o1 + ij] = jz;

            double rhog_col[KALL];
            for (int k = 0; k < KALL; k++) {
                size_t p = o2 + (size_t)k * ijdim + ij;
                double rho = RHO[k];
                double rhog = rho * G1[k];
                rhog_col[k] = rhog;
                double qd = 1.0;
                for (int n = 1; n <= 5; n++) qd = qd - Q[n][k];
                double cv = qd * CVdry;
                for (int n = 1; n <= 5; n++) cv = cv + Q[n][k] * CVdry;
                double vx = U[k] * ix + V[k] * jx, vy = U[k] * iy + V[k] * jy, vz = U[k] * iz + V[k] * jz;
                f->rho[p] = rho;
                f->tem[p] = T[k];
                f->pre[p] = rho * T[k] * (qd * Rdry + Q[0][k] * Rvap);
                f->rhog[p] = rhog;
                f->rhoge[p] = rhog * cv * T[k];
                f->vx[p] = vx; f->vy[p] = vy; f->vz[p] = vz;
                f->rhogvx[p] = rhog * vx; f->rhogvy[p] = rhog * vy; f->rhogvz[p] = rhog * vz;
                f->w[p] = W[k];
                f->gsgam2[p] = G1[k]; f->gsgam2h[p] = G2[k]; f->gam2[p] = G3[k]; f->gam2h[p] = G4[k];
                f->z[p] = gz[k];
                f->zh[p] = gzh[k];
                f->qd[p] = qd;
                for (int n = 0; n < 6; n++) {
                    size_t pq = o3 + ((size_t)n * KALL + k) * ijdim + ij;
                    f->q[pq] = Q[n][k];
                    f->rhogq[pq] = rhog * Q[n][k];
                }
                if (f->unccn) f->unccn[p] = 0.0;
                if (f->frhoge_af) f->frhoge_af[p] = 0.0;
                if (f->frhogqv_af) f->frhogqv_af[p] = 0.0;
                if (f->frhoge_rad) f->frhoge_rad[p] = 0.0;
                if (f->rhogqke) f->rhogqke[p] = 0.0;
                if (f->CBMFX) f->CBMFX[p] = 0.0;
                f->re_liquid[p] = 0.0; f->re_solid[p] = 0.0; f->re_cld[p] = 0.0;
                f->GDCLW[p] = 0.0; f->GDCFRC[p] = 0.0; f->GPREC[p] = 0.0;
            }
            for (int k = 0; k < KALL; k++) {
                size_t p = o2 + (size_t)k * ijdim + ij;
                double rw = 0.0;
                if (k >= 1) {
                    double c0 = g.afact[k] / G1[k] * G2[k];
                    double c1 = g.bfact[k] / G1[k - 1] * G2[k];
                    rw = (c0 * rhog_col[k] + c1 * rhog_col[k - 1]) * W[k];
                }
                f->rhogw[p] = rw;
            }
            for (int r = 0; r < 2; r++) {
                f->precip[(size_t)l * 2 * ijdim + (size_t)r * ijdim + ij] = 0.0;
                f->ISO1_precip[(size_t)l * 2 * ijdim + (size_t)r * ijdim + ij] = 0.0;
                f->ISO2_precip[(size_t)l * 2 * ijdim + (size_t)r * ijdim + ij] = 0.0;
            }
            f->precip_rhoe[o1 + ij] = 0.0; f->precip_lh_heat[o1 + ij] = 0.0;
            f->precip_rhophi[o1 + ij] = 0.0; f->precip_rhokin[o1 + ij] = 0.0;
            f->rctop[o1 + ij] = 0.0; f->rwtop[o1 + ij] = 0.0; f->tctop[o1 + ij] = 0.0;
        }
    return NSW6_OK;
}
