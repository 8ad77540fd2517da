/*
 * nsw6_snapshot.c -- the NICAM snapshot files of the reference's stand-alone kernel
 * run (src/main.cpp:128-174 reads them, :326-337 reads the verification set), on
 * host arrays in the layout of nsw6_fields.
 *
 * A run directory holds
 *   data/<name>.dat               inputs of mp_driver,
 *   data/vgrid/GRD_gz.dat, GRD_gzh.dat   the vertical grid (96 doubles each,
 *                                 src/mod_debug.cpp:189-195),
 *   ref_verify/calculated_<name>_DP.dat  the state after the run.
 * Every file is a raw array of native (little-endian) doubles without a header,
 * row-major in the shape the reference's reader assumes (src/data_io.cpp:207-402):
 *   3-D fields   [l][96][ij]          (read_data_3d, BUF_3D)
 *   tracers      [l][6][96][ij]       (read_data_4d, BUF_4D: rhogq_Lswp, q_Lswp)
 *   precip*_mp   [2][l][1][ij]        (read_data_4d, BUF_4D2: View4D(2, lall, 1, ij),
 *                                      the precipitation row is the slowest axis)
 *   per column   [l][1][ij]           (read_data_3d, BUF_3D2)
 *   unit vectors [l][ij]              (read_data_2d, BUF_2D)
 * The reference stops on a file it cannot open or that is too short; here every file
 * must hold exactly the bytes of its shape (NSW6_IO_ERROR otherwise), so a wrong
 * lall or ijdim is reported instead of read.
 */
#define _POSIX_C_SOURCE 200809L   /* mkdir */
#include <errno.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "../../include/nsw6.h"

enum shape { S3D, S4D, SPRECIP, SCOL, S2D };

typedef struct {
    const char *file;   /* data/<file>.dat */
    size_t offset;      /* byte offset of the member in nsw6_fields */
    enum shape shape;
} entry;

#define F(member) offsetof(nsw6_fields, member)
/* The read order of src/main.cpp:128-174 (rhogq_Lswp/q_Lswp are the tracer arrays,
 * precip1/precip2 the ISO precipitations, qke the rhogqke argument). */
static const entry INPUTS[] = {
    {"rhog", F(rhog), S3D},           {"rhogvx", F(rhogvx), S3D},
    {"rhogvy", F(rhogvy), S3D},       {"rhogvz", F(rhogvz), S3D},
    {"rhogw", F(rhogw), S3D},         {"rhoge", F(rhoge), S3D},
    {"vx", F(vx), S3D},               {"vy", F(vy), S3D},
    {"vz", F(vz), S3D},               {"w", F(w), S3D},
    {"unccn", F(unccn), S3D},         {"rho", F(rho), S3D},
    {"pre", F(pre), S3D},             {"tem", F(tem), S3D},
    {"rhogq_Lswp", F(rhogq), S4D},    {"q_Lswp", F(q), S4D},
    {"precip_mp", F(precip), SPRECIP}, {"precip1_mp", F(ISO1_precip), SPRECIP},
    {"precip2_mp", F(ISO2_precip), SPRECIP},
    {"rhoein_precip_mp", F(precip_rhoe), SCOL}, {"lh_precip_mp", F(precip_lh_heat), SCOL},
    {"rhophi_precip_mp", F(precip_rhophi), SCOL}, {"rhokin_precip_mp", F(precip_rhokin), SCOL},
    {"frhoge_af", F(frhoge_af), S3D}, {"frhogqv_af", F(frhogqv_af), S3D},
    {"frhoge_rad", F(frhoge_rad), S3D}, {"qke", F(rhogqke), S3D},
    {"gsgam2", F(gsgam2), S3D},       {"gsgam2h", F(gsgam2h), S3D},
    {"gam2", F(gam2), S3D},           {"gam2h", F(gam2h), S3D},
    {"ix", F(ix), S2D}, {"iy", F(iy), S2D}, {"iz", F(iz), S2D},
    {"jx", F(jx), S2D}, {"jy", F(jy), S2D}, {"jz", F(jz), S2D},
    {"z", F(z), S3D},   {"zh", F(zh), S3D},
    {"GPREC", F(GPREC), S3D}, {"CBMFX", F(CBMFX), S3D},
};
#undef F
#define N_INPUTS (sizeof INPUTS / sizeof INPUTS[0])

/* ref_verify names of src/main.cpp:326-337: six 3-D fields, then rhogq per tracer */
static const char *VERIFY_NAMES[NSW6_VERIFY_FIELDS] = {
    "rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "QV1", "QC2", "QR3", "QI4", "QS5", "QG6"};

static size_t shape_count(enum shape s, size_t lall, size_t ijdim)
{
    switch (s) {
    case S3D: return lall * NSW6_KALL * ijdim;
    case S4D: return lall * NSW6_NQMAX * NSW6_KALL * ijdim;
    case SPRECIP: return 2 * lall * ijdim;
    case SCOL: return lall * ijdim;
    default: return lall * ijdim;
    }
}

static int path_of(char *buf, size_t n, const char *dir, const char *sub, const char *name, const char *suffix)
{
    const int w = snprintf(buf, n, "%s/%s/%s%s", dir, sub, name, suffix);
    return (w > 0 && (size_t)w < n) ? NSW6_OK : NSW6_BAD_ARG;
}

/* dir/sub (one level below an existing dir), created when missing */
static int make_dir(const char *dir, const char *sub)
{
    char path[4096];
    const int w = snprintf(path, sizeof path, "%s/%s", dir, sub);
    if (w <= 0 || (size_t)w >= sizeof path) return NSW6_BAD_ARG;
    if (mkdir(path, 0755) != 0 && errno != EEXIST) return NSW6_IO_ERROR;
    return NSW6_OK;
}

static int read_exact(const char *path, double *dst, size_t count)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return NSW6_IO_ERROR;
    int st = NSW6_OK;
    if (fseek(fp, 0, SEEK_END) != 0 || ftell(fp) != (long)(count * sizeof(double)) ||
        fseek(fp, 0, SEEK_SET) != 0 || fread(dst, sizeof(double), count, fp) != count)
        st = NSW6_IO_ERROR;
    fclose(fp);
    return st;
}

static int write_exact(const char *path, const double *src, size_t count)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return NSW6_IO_ERROR;
    int st = (fwrite(src, sizeof(double), count, fp) == count) ? NSW6_OK : NSW6_IO_ERROR;
    if (fclose(fp) != 0) st = NSW6_IO_ERROR;
    return st;
}

static double *member(const nsw6_fields *f, size_t offset)
{
    double *p;
    memcpy(&p, (const char *)f + offset, sizeof p);
    return p;
}

/* precip files are [2][l][ij]; nsw6_fields holds [l][2][ij] */
static void precip_to_fields(const double *file, double *dst, size_t lall, size_t ijdim)
{
    for (size_t r = 0; r < 2; r++)
        for (size_t l = 0; l < lall; l++)
            memcpy(dst + (l * 2 + r) * ijdim, file + (r * lall + l) * ijdim, ijdim * sizeof(double));
}
static void precip_to_file(const double *src, double *file, size_t lall, size_t ijdim)
{
    for (size_t r = 0; r < 2; r++)
        for (size_t l = 0; l < lall; l++)
            memcpy(file + (r * lall + l) * ijdim, src + (l * 2 + r) * ijdim, ijdim * sizeof(double));
}

static int vgrid_io(const char *dir, double *gz, double *gzh, const double *wgz, const double *wgzh)
{
    char path[4096];
    const char *names[2] = {"GRD_gz", "GRD_gzh"};
    for (int i = 0; i < 2; i++) {
        int st = path_of(path, sizeof path, dir, "data/vgrid", names[i], ".dat");
        if (st != NSW6_OK) return st;
        if (wgz) st = write_exact(path, i == 0 ? wgz : wgzh, NSW6_KALL);
        else st = read_exact(path, i == 0 ? gz : gzh, NSW6_KALL);
        if (st != NSW6_OK) return st;
    }
    return NSW6_OK;
}

int nsw6_snapshot_read(const char *dir, int lall, int ijdim, const nsw6_fields *f, double *gz, double *gzh)
{
    if (!dir || !f || lall <= 0 || ijdim <= 0) return NSW6_BAD_ARG;
    char path[4096];
    double *tmp = NULL;
    for (size_t i = 0; i < N_INPUTS; i++) {
        double *dst = member(f, INPUTS[i].offset);
        if (!dst) continue;   /* the caller does not want this array */
        const size_t n = shape_count(INPUTS[i].shape, (size_t)lall, (size_t)ijdim);
        int st = path_of(path, sizeof path, dir, "data", INPUTS[i].file, ".dat");
        if (st != NSW6_OK) { free(tmp); return st; }
        if (INPUTS[i].shape == SPRECIP) {
            if (!tmp && !(tmp = (double *)malloc(n * sizeof(double)))) return NSW6_IO_ERROR;
            st = read_exact(path, tmp, n);
            if (st == NSW6_OK) precip_to_fields(tmp, dst, (size_t)lall, (size_t)ijdim);
        } else {
            st = read_exact(path, dst, n);
        }
        if (st != NSW6_OK) { free(tmp); return st; }
    }
    free(tmp);
    if (gz && gzh) return vgrid_io(dir, gz, gzh, NULL, NULL);
    return NSW6_OK;
}

int nsw6_snapshot_write(const char *dir, int lall, int ijdim, const nsw6_fields *f, const double *gz,
                        const double *gzh)
{
    if (!dir || !f || lall <= 0 || ijdim <= 0) return NSW6_BAD_ARG;
    int st0 = make_dir(dir, "data");
    if (st0 == NSW6_OK && gz && gzh) st0 = make_dir(dir, "data/vgrid");
    if (st0 != NSW6_OK) return st0;
    char path[4096];
    double *tmp = NULL;
    for (size_t i = 0; i < N_INPUTS; i++) {
        const double *src = member(f, INPUTS[i].offset);
        if (!src) continue;
        const size_t n = shape_count(INPUTS[i].shape, (size_t)lall, (size_t)ijdim);
        int st = path_of(path, sizeof path, dir, "data", INPUTS[i].file, ".dat");
        if (st != NSW6_OK) { free(tmp); return st; }
        if (INPUTS[i].shape == SPRECIP) {
            if (!tmp && !(tmp = (double *)malloc(n * sizeof(double)))) return NSW6_IO_ERROR;
            precip_to_file(src, tmp, (size_t)lall, (size_t)ijdim);
            st = write_exact(path, tmp, n);
        } else {
            st = write_exact(path, src, n);
        }
        if (st != NSW6_OK) { free(tmp); return st; }
    }
    free(tmp);
    if (gz && gzh) return vgrid_io(dir, NULL, NULL, gz, gzh);
    return NSW6_OK;
}

int nsw6_verify_read(const char *dir, int lall, int ijdim, double *check)
{
    if (!dir || !check || lall <= 0 || ijdim <= 0) return NSW6_BAD_ARG;
    const size_t n = (size_t)lall * NSW6_KALL * (size_t)ijdim;
    char path[4096], name[64];
    for (int i = 0; i < NSW6_VERIFY_FIELDS; i++) {
        snprintf(name, sizeof name, "calculated_%s", VERIFY_NAMES[i]);
        int st = path_of(path, sizeof path, dir, "ref_verify", name, "_DP.dat");
        if (st == NSW6_OK) st = read_exact(path, check + (size_t)i * n, n);
        if (st != NSW6_OK) return st;
    }
    return NSW6_OK;
}

int nsw6_verify_write(const char *dir, int lall, int ijdim, const nsw6_fields *f)
{
    if (!dir || !f || !f->rhog || !f->rhogvx || !f->rhogvy || !f->rhogvz || !f->rhogw || !f->rhoge ||
        !f->rhogq || lall <= 0 || ijdim <= 0)
        return NSW6_BAD_ARG;
    const int st0 = make_dir(dir, "ref_verify");
    if (st0 != NSW6_OK) return st0;
    const size_t K2 = (size_t)NSW6_KALL * (size_t)ijdim, n = (size_t)lall * K2;
    const double *fields6[6] = {f->rhog, f->rhogvx, f->rhogvy, f->rhogvz, f->rhogw, f->rhoge};
    char path[4096], name[64];
    double *tmp = NULL;
    for (int i = 0; i < NSW6_VERIFY_FIELDS; i++) {
        snprintf(name, sizeof name, "calculated_%s", VERIFY_NAMES[i]);
        int st = path_of(path, sizeof path, dir, "ref_verify", name, "_DP.dat");
        if (st != NSW6_OK) { free(tmp); return st; }
        if (i < 6) {
            st = write_exact(path, fields6[i], n);
        } else {
            /* tracer i-6 of rhogq [l][6][K][ij] gathered to [l][K][ij] */
            if (!tmp && !(tmp = (double *)malloc(n * sizeof(double)))) return NSW6_IO_ERROR;
            for (size_t l = 0; l < (size_t)lall; l++)
                memcpy(tmp + l * K2, f->rhogq + (l * NSW6_NQMAX + (size_t)(i - 6)) * K2, K2 * sizeof(double));
            st = write_exact(path, tmp, n);
        }
        if (st != NSW6_OK) { free(tmp); return st; }
    }
    free(tmp);
    return NSW6_OK;
}
