/* snapshot.c -- Gadget-2 snapshot/IC files, as the reference reads and writes them
 * (src/snapshot.c:5-22 header, 64-120 read_GadgetHeader, 211-293
 * read_Particle_Gadget2, 397-503 write_Particle_Gadget2).
 *
 * File: Fortran-style records (int length, data, int length): the 256-byte header,
 * then positions (float32 x 3 per particle, all particle types in order), then
 * velocities (float32 x 3, stored as v / a^1.5 with a = 1/(1+z): the reader
 * multiplies by a^1.5).  IDs, when present, follow and are not read (the reference
 * skips them too).  Host code, no GPU.                                             */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "photons_fmm.h"

/* the 256-byte header, field for field (src/snapshot.c:5-22) */
typedef struct {
    int32_t npart[6];
    double mass[6];
    double time;
    double redshift;
    int32_t flag_sfr;
    int32_t flag_feedback;
    int32_t npartTotal[6];
    int32_t flag_cooling;
    int32_t num_files;
    double BoxSize;
    double Omega0;
    double OmegaLambda;
    double HubbleParam;
    char fill[256 - 6 * 4 - 6 * 8 - 2 * 8 - 2 * 4 - 6 * 4 - 2 * 4 - 4 * 8];
} gadget_header;

_Static_assert(sizeof(gadget_header) == 256, "the Gadget-2 header is 256 bytes");

static _Thread_local char g_err[256];

const char *photons_snapshot_last_error(void) { return g_err; }

static int fail(const char *m, const char *path)
{
    snprintf(g_err, sizeof g_err, "%s: %s", m, path);
    return -1;
}

static int read_header(FILE *f, gadget_header *h)
{
    int32_t a, b;
    if (fread(&a, 4, 1, f) != 1 || fread(h, sizeof *h, 1, f) != 1 || fread(&b, 4, 1, f) != 1) return -1;
    return 0;
}

static void to_info(const gadget_header *h, photons_snapshot_info *info)
{
    memset(info, 0, sizeof *info);
    info->n = 0;
    for (int m = 0; m < 6; m++) {
        info->npart[m] = h->npart[m];
        info->npart_total[m] = h->npartTotal[m];
        info->mass[m] = h->mass[m];
        info->n += h->npart[m];
    }
    info->time = h->time;
    info->redshift = h->redshift;
    info->boxsize = h->BoxSize;
    info->omega_m = h->Omega0;
    info->omega_l = h->OmegaLambda;
    info->hubble = h->HubbleParam;
    info->num_files = h->num_files;
}

int photons_snapshot_header(const char *path, photons_snapshot_info *info)
{
    FILE *f = fopen(path, "rb");
    if (!f) return fail("cannot open", path);
    gadget_header h;
    int rc = read_header(f, &h);
    fclose(f);
    if (rc) return fail("short header", path);
    to_info(&h, info);
    return 0;
}

/* particles [n_start, n_start + n_count) of the file's particle order, positions and
 * velocities as doubles [n_count][3] (either may be NULL) */
int photons_snapshot_read(const char *path, int64_t n_start, int64_t n_count, double *pos, double *vel)
{
    FILE *f = fopen(path, "rb");
    if (!f) return fail("cannot open", path);
    gadget_header h;
    if (read_header(f, &h)) {
        fclose(f);
        return fail("short header", path);
    }
    int64_t n = 0;
    for (int m = 0; m < 6; m++) n += h.npart[m];
    if (n_start < 0 || n_count < 0 || n_start + n_count > n) {
        fclose(f);
        return fail("particle range outside the file", path);
    }
    const double gdt2unit = pow(1.0 / (1.0 + h.redshift), 1.5);
    float d[3];
    int32_t mark;
    for (int block = 0; block < 2; block++) {
        double *out = block == 0 ? pos : vel;
        if (fread(&mark, 4, 1, f) != 1) goto short_file;
        for (int64_t c = 0; c < n; c++) {
            if (fread(d, sizeof(float), 3, f) != 3) goto short_file;
            if (out && c >= n_start && c < n_start + n_count) {
                const int64_t p = c - n_start;
                for (int k = 0; k < 3; k++) out[3 * p + k] = block == 0 ? (double)d[k] : (double)d[k] * gdt2unit;
            }
        }
        if (fread(&mark, 4, 1, f) != 1) goto short_file;
    }
    fclose(f);
    return 0;
short_file:
    fclose(f);
    return fail("short particle block", path);
}

/* one file of n type-1 particles; info supplies the cosmology, box, redshift and
 * masses (npart/npart_total are set from n and npart_total_all) */
int photons_snapshot_write(const char *path, const photons_snapshot_info *info, int64_t n, int64_t npart_total_all,
                           const double *pos, const double *vel)
{
    if (n < 0 || n > INT32_MAX || npart_total_all > INT32_MAX) return fail("particle count out of range", path);
    gadget_header h;
    memset(&h, 0, sizeof h);
    h.num_files = 1;
    h.BoxSize = info->boxsize;
    h.Omega0 = info->omega_m;
    h.OmegaLambda = info->omega_l;
    h.HubbleParam = info->hubble;
    h.npart[1] = (int32_t)n;
    h.npartTotal[1] = (int32_t)npart_total_all;
    for (int m = 0; m < 6; m++) h.mass[m] = info->mass[m];
    h.redshift = info->redshift;
    h.time = 1.0 / (info->redshift + 1.0);
    FILE *f = fopen(path, "wb");
    if (!f) return fail("cannot create", path);
    /* record markers hold the record lengths (the reference leaves them unset) */
    int32_t mark = (int32_t)sizeof h;
    int ok = fwrite(&mark, 4, 1, f) == 1 && fwrite(&h, sizeof h, 1, f) == 1 && fwrite(&mark, 4, 1, f) == 1;
    const double gdt2unit = pow(1.0 / (1.0 + h.redshift), 1.5);
    mark = (int32_t)(12 * n);
    for (int block = 0; block < 2 && ok; block++) {
        const double *src = block == 0 ? pos : vel;
        ok = fwrite(&mark, 4, 1, f) == 1;
        for (int64_t c = 0; c < n && ok; c++) {
            float d[3];
            for (int k = 0; k < 3; k++)
                d[k] = block == 0 ? (float)src[3 * c + k] : (float)((double)(float)src[3 * c + k] / gdt2unit);
            ok = fwrite(d, sizeof(float), 3, f) == 3;
        }
        ok = ok && fwrite(&mark, 4, 1, f) == 1;
    }
    if (fclose(f) != 0) ok = 0;
    return ok ? 0 : fail("write failed", path);
}
