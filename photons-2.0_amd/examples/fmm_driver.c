/* fmm_driver.c -- the reference driver's force call, linked against
 * libphotons_fmm.so in place of src/fmm.c, src/operator.c and src/remotes.c.
 *
 * The calls and globals are those of src/photoNs.c:97-116 and 222-252: the caller
 * owns part[], zeroes part[].acc, sets the parameters and calls fmm_construct,
 * fmm_prepare, fmm_task and fmm_ext; part[].acc then holds the short-range force
 * (part[] permuted into tree order, so each body carries its input index in vel[0]).
 *
 *   fmm_driver IN OUT [fp32]
 *   IN   int32 n, maxleaf, 0, 0; double boxsize, mass, soft, rs, rc, theta;
 *        n x 3 double positions
 *   OUT  n x 3 double accelerations in input order, then the int64 counters
 *        idxP2P, idxM2L, walk_m2l_count, p2p_count_remote, leaves, nodes, then the
 *        double timers dtime_prep, dtime_task, dtime_ext, dtime_fmm, dtime_p2p,
 *        dtime_m2l [s] and the library's phase times of the call [ms]
 *        (PHOTONS_NPHASE doubles).  The timers are printed as src/photoNs.c:122-127
 *        prints them.                                                            */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "photons_fmm.h"

/* the reference's globals (inc/photoNs.h, inc/typesdef.h), defined by the library */
extern photons_body *part;
extern int NPART, MAXLEAF, first_leaf, last_leaf, first_node, last_node;
extern double MASSPART, SoftenScale, splitRadius, cutoffRadius, open_angle, BOXSIZE;
extern unsigned long idxP2P, idxM2L, walk_m2l_count, p2p_count_remote;
extern double dtime_prep, dtime_task, dtime_ext, dtime_fmm, dtime_p2p, dtime_m2l;

/* the drop-in layer exits on a fatal error, as the reference does; report its message */
static void report(void)
{
    const char *m = photons_gpu_last_error();
    if (m[0]) fprintf(stderr, "fmm_driver: the force path stopped: %s\n", m);
}

static void die(const char *m)
{
    fprintf(stderr, "fmm_driver: %s\n", m);
    exit(1);
}

int main(int argc, char **argv)
{
    if (argc < 3) die("usage: fmm_driver IN OUT [fp32]");
    atexit(report);
    FILE *f = fopen(argv[1], "rb");
    if (!f) die("cannot open the input");
    int32_t hdr[4];
    double prm[6];
    if (fread(hdr, sizeof hdr, 1, f) != 1 || fread(prm, sizeof prm, 1, f) != 1) die("short input");
    const int n = hdr[0];
    if (n <= 0) die("no particles");
    double *pos = malloc(sizeof(double) * 3 * (size_t)n);
    if (fread(pos, sizeof(double) * 3, (size_t)n, f) != (size_t)n) die("short input");
    fclose(f);

    part = calloc((size_t)n, sizeof(photons_body));
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) part[i].pos[k] = pos[3 * i + k];
        part[i].vel[0] = i; /* tag: fmm_prepare permutes part[] */
    }
    NPART = n;
    MAXLEAF = hdr[1];
    BOXSIZE = prm[0];
    MASSPART = prm[1];
    SoftenScale = prm[2];
    splitRadius = prm[3];
    cutoffRadius = prm[4];
    open_angle = prm[5];
    if (argc > 3) photons_compat_set_precision(PHOTONS_FP32_MIXED);

    /* src/photoNs.c:222-252 */
    for (int i = 0; i < NPART; i++) part[i].acc[0] = part[i].acc[1] = part[i].acc[2] = 0.0;
    fmm_construct();
    fmm_prepare();
    fmm_task();
    fmm_ext();

    double *acc = malloc(sizeof(double) * 3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        const int id = (int)part[i].vel[0];
        for (int k = 0; k < 3; k++) acc[3 * id + k] = part[i].acc[k];
    }
    int64_t cnt[6] = {(int64_t)idxP2P, (int64_t)idxM2L, (int64_t)walk_m2l_count, (int64_t)p2p_count_remote,
                      (int64_t)(last_leaf - first_leaf), (int64_t)(last_node - first_node + 1)};
    /* src/photoNs.c:122-127 */
    printf(" dtime_m2l = %lf, dtime_p2p = %lf [sec]\n", dtime_m2l, dtime_p2p);
    printf(" dtime_fmm/task = %lf, task = %lf [sec]\n", dtime_fmm, dtime_task);
    printf(" dtime_prep = %lf, ext = %lf [sec]\n", dtime_prep, dtime_ext);
    double tm[6] = {dtime_prep, dtime_task, dtime_ext, dtime_fmm, dtime_p2p, dtime_m2l};
    double ph[PHOTONS_NPHASE];
    if (photons_get_phase_ms(photons_compat_context(), ph) != 0) die(photons_last_error());
    fmm_deconstruct();
    f = fopen(argv[2], "wb");
    if (!f) die("cannot open the output");
    fwrite(acc, sizeof(double) * 3, (size_t)n, f);
    fwrite(cnt, sizeof cnt, 1, f);
    fwrite(tm, sizeof tm, 1, f);
    fwrite(ph, sizeof ph, 1, f);
    fclose(f);
    free(acc);
    free(pos);
    free(part);
    return 0;
}
