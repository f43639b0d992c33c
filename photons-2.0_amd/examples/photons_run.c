/* photons_run.c -- the reference program's run on one GPU: a Gadget-2 initial
 * condition in, Nstep TreePM steps in ln a up to a_f, a Gadget-2 snapshot out
 * (src/photoNs.c:35-313 driver, :331-355 main; src/initial.c:60-345 parameters).
 *
 *   photons_run PARAMFILE [--nstep N] [--af A] [--fp32] [--order P] [--maxleaf L] [--no-pm]
 *
 * PARAMFILE has the reference's "Name value" lines (# starts a comment).  Read here:
 * InputPath, OutputPath, OutputName, OPENANGLE, NumMeshSide, GravConstant,
 * SOFTENING, SPLITSCALE, MaxPackage (the leaf size; 32 when absent, --maxleaf
 * overrides it), SnapFormat (2, Gadget-2, only) and SnapNumber (1 only); the other
 * names are accepted and ignored.  As the reference: theta 0.3 unless
 * OPENANGLE is given, rs = 1.25 L/NumMeshSide, soft = 0.03 L/N^(1/3), rc = 4.5 rs,
 * the cosmology, box, mass and start redshift from the IC header, 512 steps to
 * a = 1, output OutputPath/OutputName_3.0.
 *
 * Particles stay in HBM for the whole run; each step is photons_step_device (the
 * reference's loop body with the PM force beside the short-range one).  Host C;
 * device memory through the HIP runtime.                                        */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "photons_fmm.h"

static void die(const char *what, const char *detail)
{
    fprintf(stderr, "photons_run: %s%s%s\n", what, detail ? ": " : "", detail ? detail : "");
    exit(1);
}

static void hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess) die(what, hipGetErrorString(e));
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

typedef struct {
    char input[768], out_path[768], out_name[768];
    double theta, grav, soft, split;
    int nside, format, nfiles, maxleaf;
} run_params;

/* src/initial.c:60-175: one "name value" pair per line */
static void read_params(const char *path, run_params *p)
{
    FILE *f = fopen(path, "r");
    if (!f) die("cannot open the parameter file", path);
    char line[1024], name[256], value[768];
    while (fgets(line, sizeof line, f)) {
        char *hash = strchr(line, '#');
        if (hash) *hash = 0;
        if (sscanf(line, "%255s %767s", name, value) != 2) continue;
        if (!strcmp(name, "InputPath")) snprintf(p->input, sizeof p->input, "%s", value);
        else if (!strcmp(name, "OutputPath")) snprintf(p->out_path, sizeof p->out_path, "%s", value);
        else if (!strcmp(name, "OutputName")) snprintf(p->out_name, sizeof p->out_name, "%s", value);
        else if (!strcmp(name, "OPENANGLE")) p->theta = atof(value);
        else if (!strcmp(name, "NumMeshSide")) p->nside = atoi(value);
        else if (!strcmp(name, "GravConstant")) p->grav = atof(value);
        else if (!strcmp(name, "SOFTENING")) p->soft = atof(value);
        else if (!strcmp(name, "SPLITSCALE")) p->split = atof(value);
        else if (!strcmp(name, "SnapFormat")) p->format = atoi(value);
        else if (!strcmp(name, "SnapNumber")) p->nfiles = atoi(value);
        else if (!strcmp(name, "MaxPackage")) p->maxleaf = atoi(value);
    }
    fclose(f);
}

int main(int argc, char **argv)
{
    if (argc < 2) die("usage: photons_run PARAMFILE [--nstep N] [--af A] [--fp32] [--order P] [--maxleaf L] [--no-pm]", NULL);
    run_params rp = {.theta = 0.3, .grav = 43007.105732, .soft = -1.0, .split = -1.0, .nside = 0, .format = 2,
                     .nfiles = 1, .maxleaf = 32};
    snprintf(rp.out_path, sizeof rp.out_path, ".");
    snprintf(rp.out_name, sizeof rp.out_name, "snapshot");
    read_params(argv[1], &rp);
    int nstep = 512, order = 3, maxleaf = rp.maxleaf, with_pm = 1, precision = PHOTONS_FP64;
    double af = 1.0;
    for (int i = 2; i < argc; i++) {
        if (!strcmp(argv[i], "--nstep") && i + 1 < argc) nstep = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--af") && i + 1 < argc) af = atof(argv[++i]);
        else if (!strcmp(argv[i], "--order") && i + 1 < argc) order = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--maxleaf") && i + 1 < argc) maxleaf = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--fp32")) precision = PHOTONS_FP32_MIXED;
        else if (!strcmp(argv[i], "--no-pm")) with_pm = 0;
        else die("unknown option", argv[i]);
    }
    if (rp.format != 2) die("only Gadget-2 initial conditions (SnapFormat 2) are read", NULL);
    if (rp.nfiles != 1) die("only single-file initial conditions (SnapNumber 1) are read", NULL);
    if (rp.nside <= 0) die("NumMeshSide must be set", NULL);
    if (nstep < 1) die("--nstep must be positive", NULL);

    photons_snapshot_info info;
    if (photons_snapshot_header(rp.input, &info)) die("IC header", photons_snapshot_last_error());
    const int64_t n = info.n;
    const double L = info.boxsize;
    const double ai = 1.0 / (1.0 + info.redshift);
    if (n <= 0 || !(L > 0.0)) die("the IC holds no particles or no box", rp.input);
    if (!(af > ai)) die("a_f must be after the IC's expansion factor", NULL);

    /* src/initial.c:314-345 */
    photons_params prm;
    memset(&prm, 0, sizeof prm);
    prm.boxsize = L;
    prm.mass = info.mass[1];
    prm.rs = rp.split > 0.0 ? rp.split : 1.25 * L / (double)rp.nside;
    prm.soft = rp.soft >= 0.0 ? rp.soft : 0.03 * L / pow((double)n, 0.3333333);
    prm.rc = 4.5 * prm.rs;
    prm.theta = rp.theta;
    prm.maxleaf = maxleaf;
    prm.order = order;
    prm.periodic = 1;
    prm.precision = precision;
    printf(" photons_run: N = %lld, L = %g, z_init = %g, a %g -> %g in %d steps\n", (long long)n, L,
           info.redshift, ai, af, nstep);
    printf("   rs = %g, rc = %g, soft = %g, theta = %g, p = %d, leaf %d, %s, PM %s (NSIDE %d)\n", prm.rs, prm.rc,
           prm.soft, prm.theta, order, maxleaf, precision == PHOTONS_FP64 ? "FP64" : "FP32-mixed",
           with_pm ? "on" : "off", rp.nside);

    const size_t bytes = sizeof(double) * 3 * (size_t)n;
    double *pos = malloc(bytes), *vel = malloc(bytes);
    if (!pos || !vel) die("out of host memory", NULL);
    if (photons_snapshot_read(rp.input, 0, n, pos, vel)) die("IC particles", photons_snapshot_last_error());

    photons_ctx *ctx = NULL;
    photons_pm *pm = NULL;
    if (photons_create(&ctx, &prm, 0, NULL)) die("photons_create", photons_last_error());
    if (with_pm && photons_pm_create(&pm, rp.nside, L, prm.rs, prm.mass, 0, NULL))
        die("photons_pm_create", photons_pm_last_error());
    double *d_pos, *d_vel, *d_acc, *d_apm = NULL;
    hip_ok(hipMalloc((void **)&d_pos, bytes), "hipMalloc");
    hip_ok(hipMalloc((void **)&d_vel, bytes), "hipMalloc");
    hip_ok(hipMalloc((void **)&d_acc, bytes), "hipMalloc");
    if (pm) hip_ok(hipMalloc((void **)&d_apm, bytes), "hipMalloc");
    hip_ok(hipMemcpy(d_pos, pos, bytes, hipMemcpyHostToDevice), "upload");
    hip_ok(hipMemcpy(d_vel, vel, bytes, hipMemcpyHostToDevice), "upload");

    /* the forces of the initial positions (src/photoNs.c:97-116) */
    const double t0 = now_s();
    if (pm && photons_pm_force_device(pm, n, d_pos, d_apm)) die("PM", photons_pm_last_error());
    if (photons_force_device(ctx, n, d_pos, d_acc)) die("force", photons_last_error());
    if (pm && photons_pm_sync(pm)) die("PM", photons_pm_last_error());
    if (photons_sync(ctx)) die("force", photons_last_error());

    /* src/photoNs.c:136-254 */
    const double dloga = (log(af) - log(ai)) / nstep;
    for (int loop = 0; loop < nstep; loop++) {
        const double ts = now_s();
        const double loga_i = loop * dloga + log(ai), loga_f = (loop + 1) * dloga + log(ai);
        const double dk = photons_kick_factor(loga_i, loga_f, info.omega_m, info.omega_l);
        const double dd = photons_drift_factor(loga_i, loga_f, info.omega_m, info.omega_l);
        const double dkh = 0.5 * dk * rp.grav;
        if (photons_step_device(ctx, pm, n, d_pos, d_vel, d_acc, d_apm, dkh, dd))
            die("step", photons_last_error());
        if (photons_sync(ctx)) die("step", photons_last_error());
        printf(" LOOP %5d  a = (%lf to %lf)  %.3f s\n", loop + 1, exp(loga_i), exp(loga_f), now_s() - ts);
        fflush(stdout);
    }
    const double total = now_s() - t0;

    hip_ok(hipMemcpy(pos, d_pos, bytes, hipMemcpyDeviceToHost), "download");
    hip_ok(hipMemcpy(vel, d_vel, bytes, hipMemcpyDeviceToHost), "download");
    /* src/photoNs.c:316-322, src/snapshot.c:538-549: OutputPath/OutputName_3.0 at z = 1/a_f - 1 */
    char out[1600];
    snprintf(out, sizeof out, "%s/%s_%d.%d", rp.out_path, rp.out_name, 3, 0);
    photons_snapshot_info oi = info;
    oi.redshift = 1.0 / af - 1.0;
    if (photons_snapshot_write(out, &oi, n, n, pos, vel)) die("snapshot", photons_snapshot_last_error());
    printf(" wrote %s; %d steps in %.2f s (%.3g particles/s per step)\n", out, nstep, total,
           (double)n * nstep / total);

    hipFree(d_pos);
    hipFree(d_vel);
    hipFree(d_acc);
    if (d_apm) hipFree(d_apm);
    photons_pm_destroy(pm);
    photons_destroy(ctx);
    free(pos);
    free(vel);
    return 0;
}
