/*
 * compat.c -- the reference's force-call entry points and globals, on the GPU.
 *
 * A driver written against the reference headers (inc/photoNs.h, inc/fmm.h,
 * inc/operator.h, inc/kernels.h) calls fmm_construct / fmm_prepare / fmm_task /
 * fmm_ext / fmm_deconstruct on the global particle array `part` (src/photoNs.c:
 * 97-116, 231-252).  Here those calls drive one GPU context of the native API:
 *
 *   fmm_prepare  part[].pos -> GPU; tree, P2M, M2M; part[] is permuted into tree
 *                order as build_localtree does (src/fmm.c:93); host mirrors of the
 *                reference's leaf[] and btree[] arrays are filled (p = 3 layout)
 *   fmm_task     walk, M2L, L2L, P2P and L2P on the GPU; result kept aside
 *   fmm_ext      part[].acc += result (the reference adds the image part and L2P
 *                here, src/fmm.c:987-1075; the sum over both calls is the same)
 *
 * The single operators (p2m, m2m, m2l, l2l, l2p), the pair kernels of kernels.h
 * and the walk entry points run the GPU kernels on one item or one walk.  Errors
 * print a message and exit, like the reference (src/fmm.c:411-414); the message
 * of the native layer is available from photons_last_error().
 */
#include <math.h>
#include <stdarg.h>
#include <time.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "photons_fmm.h"

#define NM3 20 /* NMULTI for p = 3 (inc/typesdef.h:13) */

/* Node and Pack of inc/typesdef.h:25-45 (p = 3) */
typedef struct {
    int updated;
    int npart;
    int son[2];
    double split;
    double width[3];
    double center[3];
    double M[NM3];
    double L[NM3];
} ref_node;

typedef struct {
    int npart;
    int ipart;
    double width[3];
    double center[3];
    double M[NM3];
    double L[NM3];
} ref_pack;

/* ---- globals of inc/photoNs.h / inc/typesdef.h read or written by the force call */
photons_body *part;
ref_node *btree;
ref_pack *leaf;
int NPART, MAXLEAF, NLEAF, NNODE;
int first_node, last_node, first_leaf, last_leaf;
double MASSPART, SoftenScale, splitRadius, cutoffRadius, open_angle, BOXSIZE;
unsigned long idxP2P, idxM2L, walk_m2l_count, p2p_count_remote;
/* phase timers of inc/photoNs.h:95-111, in seconds, written as the reference writes them:
 *   dtime_prep   wall time of fmm_prepare (src/fmm.c:725-764)
 *   dtime_fmm    fmm_prepare's time, plus L2L in fmm_ext (src/fmm.c:766-767, 1061)
 *   dtime_task   wall time of fmm_task (src/fmm.c:911, 981)
 *   dtime_ext    wall time of fmm_ext (src/fmm.c:995, 1073)
 *   dtime_p2p    P2P (+ L2P) kernel time, reset by build_localtree (src/fmm.c:182, 871)
 *   dtime_m2l    M2L kernel time, accumulated over calls (src/fmm.c:906)
 *   dtime_fmm_remote  the image walks' own time: they run fused with the home walk
 *                here, so it stays 0 (src/fmm.c:1071)
 *   dtime_pm     wall time of partmesh (src/partmesh.c:808, 1576-1577)
 * The kernel times are GPU times of the context's stream (HIP events). */
double dtime_prep, dtime_task, dtime_ext, dtime_fmm, dtime_p2p, dtime_m2l, dtime_fmm_remote, dtime_pm;

int photons_prepare_host(photons_ctx *ctx, int64_t n, const double *pos);
int photons_evaluate_host(photons_ctx *ctx, double *acc_tree_order);
int photons_acceptance(const double *wi, const double *wj, const double *dist, double theta, double rc);
int photons_walk_eval(photons_ctx *x, int start_target, int start_source, int shift_lo, int shift_hi, int do_p2p,
                      int do_m2l, double *acc_tree, int64_t *n_p2p, int64_t *n_m2l);
int photons_p2p_pairs(photons_ctx *x, int count, const int32_t *tgt, const int32_t *src, const int8_t *shift,
                      int precision, double *acc_tree);
int photons_local_expansions(photons_ctx *x, double *L, int write);
void *photons_scratch(photons_ctx *x, int slot, size_t bytes);
int photons_copy(photons_ctx *x, void *dst, const void *src, size_t bytes, int to_device);

static photons_ctx *g_ctx;
static int g_precision = PHOTONS_FP64;
static double g_ph_ms[PHOTONS_NPHASE]; /* phase times of the last fmm_task */
static double *g_acc;       /* result of fmm_task, tree order */
static ref_pack *g_leaf_mem; /* mirrors, unshifted base pointers */
static ref_node *g_node_mem;
static int g_nleaf, g_nnode;
static int g_shift_id;

/* the reference's failure behaviour (src/fmm.c:411-414, src/utility.c:78-81): a message,
 * then exit; the message stays readable through photons_gpu_last_error() for exit
 * handlers */
static char g_fatal[1024];

const char *photons_gpu_last_error(void) { return g_fatal; }

static void fatal(const char *fmt, ...) __attribute__((noreturn, format(printf, 1, 2)));
static void fatal(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_fatal, sizeof g_fatal, fmt, ap);
    va_end(ap);
    fprintf(stderr, "photons: %s\n", g_fatal);
    fflush(stderr);
    exit(1);
}

static void die(const char *where) { fatal("%s: %s", where, photons_last_error()); }

static void check(int rc, const char *where)
{
    if (rc != 0) die(where);
}

/* the reference's dtime() (src/utility.c): wall clock in seconds */
static double wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int photons_compat_set_precision(int precision)
{
    if (precision != PHOTONS_FP64 && precision != PHOTONS_FP32_MIXED) return -1;
    g_precision = precision;
    return 0;
}

photons_ctx *photons_compat_context(void) { return g_ctx; }

static photons_params current_params(void)
{
    photons_params p;
    p.boxsize = BOXSIZE;
    p.mass = MASSPART;
    p.soft = SoftenScale;
    p.rs = splitRadius;
    p.rc = cutoffRadius;
    p.theta = open_angle;
    p.maxleaf = MAXLEAF;
    p.order = 3;
    p.periodic = 1;
    p.precision = g_precision;
    return p;
}

static photons_ctx *ctx(void)
{
    photons_params p = current_params();
    if (!g_ctx) {
        check(photons_create(&g_ctx, &p, 0, NULL), "photons_create");
        check(photons_set_timing(g_ctx, 1), "photons_set_timing"); /* feeds the dtime_* timers */
    } else {
        check(photons_set_params(g_ctx, &p), "photons_set_params");
    }
    return g_ctx;
}

/* reference index space (src/fmm.c:200-212): leaves [first_leaf, last_leaf),
 * nodes [first_node, last_node]; cells: leaves 0..nleaf-1, nodes after them */
static int to_cell(int idx)
{
    if (idx >= first_leaf && idx < last_leaf) return idx - first_leaf;
    if (idx >= first_node && idx <= last_node) return g_nleaf + (idx - first_node);
    fatal("index %d is neither a leaf nor a node", idx);
}

static int to_ref(int cell) { return cell < g_nleaf ? first_leaf + cell : first_node + (cell - g_nleaf); }

static void free_mirrors(void)
{
    free(g_leaf_mem);
    free(g_node_mem);
    g_leaf_mem = NULL;
    g_node_mem = NULL;
    leaf = NULL;
    btree = NULL;
}

static void pull_local_expansions(void)
{
    if (g_precision != PHOTONS_FP64 || !g_leaf_mem) return;
    int ncell = g_nleaf + g_nnode;
    double *L = (double *)malloc(sizeof(double) * NM3 * (size_t)ncell);
    check(photons_local_expansions(g_ctx, L, 0), "local expansions");
    for (int k = 0; k < ncell; k++) {
        double *dst = k < g_nleaf ? g_leaf_mem[k].L : g_node_mem[k - g_nleaf].L;
        memcpy(dst, L + (size_t)NM3 * k, sizeof(double) * NM3);
    }
    free(L);
}

void fmm_construct(void)
{
    /* one rank: the domain is the whole periodic box [0, BOXSIZE)^3 (src/fmm.c:329-360) */
    if (!(BOXSIZE > 0.0)) {
        fatal("fmm_construct: BOXSIZE must be positive");
    }
}

void build_localtree(void)
{
    photons_ctx *x = ctx();
    dtime_p2p = 0.0; /* src/fmm.c:182 */
    free_mirrors();
    double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)(NPART > 0 ? NPART : 1));
    for (int i = 0; i < NPART; i++) memcpy(pos + 3 * (size_t)i, part[i].pos, 3 * sizeof(double));
    check(photons_prepare_host(x, NPART, pos), "fmm_prepare");
    free(pos);
    photons_stats st;
    check(photons_get_stats(x, &st), "stats");
    g_nleaf = st.nleaf;
    g_nnode = st.nnode;
    /* permute part[] into tree order, as bksort_inplace leaves it */
    int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)(NPART > 0 ? NPART : 1));
    check(photons_tree_perm(x, perm), "tree order");
    photons_body *tmp = (photons_body *)malloc(sizeof(photons_body) * (size_t)(NPART > 0 ? NPART : 1));
    for (int i = 0; i < NPART; i++) tmp[i] = part[perm[i]];
    memcpy(part, tmp, sizeof(photons_body) * (size_t)NPART);
    free(tmp);
    free(perm);
    /* index spaces of src/fmm.c:203-212, widened when the tree is larger */
    int cap = (int)(2.0 * (double)NPART / (double)MAXLEAF);
    if (cap > NPART) cap = NPART + 1;
    NLEAF = cap > g_nleaf ? cap : g_nleaf;
    NNODE = cap > g_nnode ? cap : g_nnode;
    first_leaf = last_leaf = NPART;
    first_node = NPART + NLEAF;
    last_leaf = first_leaf + g_nleaf;
    last_node = first_node + g_nnode - 1;
    /* host mirrors of leaf[] and btree[] */
    int ncell = g_nleaf + g_nnode;
    if (ncell == 0) return;
    int32_t *npart = (int32_t *)malloc(sizeof(int32_t) * ncell);
    int32_t *ipart = (int32_t *)malloc(sizeof(int32_t) * ncell);
    int32_t *son = (int32_t *)malloc(sizeof(int32_t) * 2 * ncell);
    double *center = (double *)malloc(sizeof(double) * 3 * ncell);
    double *width = (double *)malloc(sizeof(double) * 3 * ncell);
    double *M = (double *)malloc(sizeof(double) * NM3 * (size_t)ncell);
    check(photons_tree_cells(x, npart, ipart, son, center, width), "tree cells");
    check(photons_expansions(x, M, NULL), "multipoles");
    g_leaf_mem = (ref_pack *)calloc((size_t)NLEAF, sizeof(ref_pack));
    g_node_mem = (ref_node *)calloc((size_t)NNODE, sizeof(ref_node));
    for (int k = 0; k < ncell; k++) {
        if (k < g_nleaf) {
            ref_pack *p = &g_leaf_mem[k];
            p->npart = npart[k];
            p->ipart = ipart[k];
            memcpy(p->width, width + 3 * k, sizeof p->width);
            memcpy(p->center, center + 3 * k, sizeof p->center);
            memcpy(p->M, M + (size_t)NM3 * k, sizeof p->M);
        } else {
            ref_node *p = &g_node_mem[k - g_nleaf];
            p->npart = npart[k];
            p->son[0] = to_ref(son[2 * k]);
            p->son[1] = to_ref(son[2 * k + 1]);
            memcpy(p->width, width + 3 * k, sizeof p->width);
            memcpy(p->center, center + 3 * k, sizeof p->center);
            memcpy(p->M, M + (size_t)NM3 * k, sizeof p->M);
        }
    }
    /* split plane: lower face of the right child along the node's axis, the axis
     * cycling x, y, z with depth from the root (src/fmm.c:99) */
    int *depth = (int *)calloc((size_t)g_nnode, sizeof(int));
    for (int j = 0; j < g_nnode; j++) { /* pre-order: parents come first */
        int ax = depth[j] % 3;
        int c1 = son[2 * (g_nleaf + j) + 1];
        g_node_mem[j].split = center[3 * c1 + ax] - 0.5 * width[3 * c1 + ax];
        for (int k = 0; k < 2; k++) {
            int c = son[2 * (g_nleaf + j) + k];
            if (c >= g_nleaf) depth[c - g_nleaf] = depth[j] + 1;
        }
    }
    free(depth);
    leaf = g_leaf_mem - first_leaf;
    btree = g_node_mem - first_node;
    free(npart); free(ipart); free(son); free(center); free(width); free(M);
}

void fmm_prepare(void)
{
    const double t0 = wall();
    build_localtree();
    idxP2P = idxM2L = 0;
    dtime_prep = wall() - t0;
    dtime_fmm = dtime_prep;
}

void fmm_task(void)
{
    const double t0 = wall();
    photons_ctx *x = ctx();
    free(g_acc);
    g_acc = (double *)calloc(3 * (size_t)(NPART > 0 ? NPART : 1), sizeof(double));
    check(photons_evaluate_host(x, g_acc), "fmm_task");
    check(photons_get_phase_ms(x, g_ph_ms), "fmm_task");
    dtime_p2p += 1e-3 * g_ph_ms[PHOTONS_PH_P2P];
    dtime_m2l += 1e-3 * g_ph_ms[PHOTONS_PH_M2L];
    photons_stats st;
    check(photons_get_stats(x, &st), "stats");
    idxP2P = (unsigned long)st.p2p_local_pairs;
    idxM2L = (unsigned long)st.m2l_local_pairs;
    walk_m2l_count += (unsigned long)st.m2l_pairs;
    p2p_count_remote += (unsigned long)st.p2p_image_interactions;
    pull_local_expansions();
    dtime_task = wall() - t0;
}

void fmm_ext(void)
{
    const double t0 = wall();
    if (!g_acc) return;
    for (int i = 0; i < NPART; i++)
        for (int d = 0; d < 3; d++) part[i].acc[d] += g_acc[3 * (size_t)i + d];
    free(g_acc);
    g_acc = NULL;
    dtime_fmm += 1e-3 * g_ph_ms[PHOTONS_PH_L2L]; /* walk_l2l; L2P runs fused with P2P */
    dtime_fmm_remote = 0.0;
    dtime_ext = wall() - t0;
}

void fmm_deconstruct(void) { free_mirrors(); }

/* ---- single operators (p = 3), src/operator.c */

void p2m(int iPart, int npart, double center[3], double M[])
{
    photons_ctx *x = ctx();
    double *dpos = (double *)photons_scratch(x, 0, sizeof(double) * 3 * (size_t)(npart > 0 ? npart : 1));
    int32_t hv[2] = {0, npart};
    int32_t *dint = (int32_t *)photons_scratch(x, 1, sizeof(hv));
    double *dc = (double *)photons_scratch(x, 2, 3 * sizeof(double));
    double *dM = (double *)photons_scratch(x, 3, NM3 * sizeof(double));
    double *hp = (double *)malloc(sizeof(double) * 3 * (size_t)(npart > 0 ? npart : 1));
    for (int i = 0; i < npart; i++) memcpy(hp + 3 * i, part[iPart + i].pos, 3 * sizeof(double));
    check(photons_copy(x, dpos, hp, sizeof(double) * 3 * (size_t)npart, 1), "p2m");
    check(photons_copy(x, dint, hv, sizeof hv, 1), "p2m");
    check(photons_copy(x, dc, center, 3 * sizeof(double), 1), "p2m");
    check(photons_p2m_batch(x, 3, 1, dpos, dint, dint + 1, dc, MASSPART, dM), "p2m");
    check(photons_copy(x, M, dM, NM3 * sizeof(double), 0), "p2m");
    free(hp);
}

static void shift_op(int kind, double dx, double dy, double dz, const double *X, double *Y)
{
    photons_ctx *x = ctx();
    double d[3] = {dx, dy, dz};
    double *dd = (double *)photons_scratch(x, 0, sizeof d);
    double *dX = (double *)photons_scratch(x, 1, NM3 * sizeof(double));
    double *dY = (double *)photons_scratch(x, 2, NM3 * sizeof(double));
    check(photons_copy(x, dd, d, sizeof d, 1), "shift");
    check(photons_copy(x, dX, X, NM3 * sizeof(double), 1), "shift");
    check(photons_copy(x, dY, Y, NM3 * sizeof(double), 1), "shift");
    if (kind == 0) check(photons_m2m_batch(x, 3, 1, dd, dX, dY), "m2m");
    else if (kind == 1) check(photons_l2l_batch(x, 3, 1, dd, dX, dY), "l2l");
    else check(photons_m2l_batch(x, 3, 1, splitRadius, dd, dX, dY), "m2l");
    check(photons_copy(x, Y, dY, NM3 * sizeof(double), 0), "shift");
}

void m2m(double dx, double dy, double dz, double M[], double tM[]) { shift_op(0, dx, dy, dz, M, tM); }
void l2l(double dx, double dy, double dz, double L[], double toL[]) { shift_op(1, dx, dy, dz, L, toL); }

void m2l(double dx, double dy, double dz, double M[], double toL[])
{
    shift_op(2, dx, dy, dz, M, toL);
    walk_m2l_count++;
}

void l2p(int ileaf)
{
    photons_ctx *x = ctx();
    ref_pack *lf = &leaf[ileaf];
    int np = lf->npart;
    if (np <= 0) return;
    double *hp = (double *)malloc(sizeof(double) * 3 * (size_t)np);
    double *ha = (double *)calloc(3 * (size_t)np, sizeof(double));
    int32_t *own = (int32_t *)calloc((size_t)np, sizeof(int32_t));
    for (int i = 0; i < np; i++) memcpy(hp + 3 * i, part[lf->ipart + i].pos, 3 * sizeof(double));
    double *dp = (double *)photons_scratch(x, 0, sizeof(double) * 3 * (size_t)np);
    double *da = (double *)photons_scratch(x, 1, sizeof(double) * 3 * (size_t)np);
    int32_t *dow = (int32_t *)photons_scratch(x, 2, sizeof(int32_t) * (size_t)np);
    double *dc = (double *)photons_scratch(x, 3, 3 * sizeof(double));
    double *dL = (double *)photons_scratch(x, 4, NM3 * sizeof(double));
    check(photons_copy(x, dp, hp, sizeof(double) * 3 * (size_t)np, 1), "l2p");
    check(photons_copy(x, da, ha, sizeof(double) * 3 * (size_t)np, 1), "l2p");
    check(photons_copy(x, dow, own, sizeof(int32_t) * (size_t)np, 1), "l2p");
    check(photons_copy(x, dc, lf->center, 3 * sizeof(double), 1), "l2p");
    check(photons_copy(x, dL, lf->L, NM3 * sizeof(double), 1), "l2p");
    check(photons_l2p_batch(x, 3, np, dp, dow, dc, dL, da), "l2p");
    check(photons_copy(x, ha, da, sizeof(double) * 3 * (size_t)np, 0), "l2p");
    for (int i = 0; i < np; i++)
        for (int d = 0; d < 3; d++) part[lf->ipart + i].acc[d] += ha[3 * i + d];
    free(hp); free(ha); free(own);
}

/* walk_m2m / walk_l2l (src/operator.c:165-194, 498-528) on the host mirrors:
 * one GPU shift per child, in the reference's order */
static double *cell_M(int ref) { return ref < first_node ? leaf[ref].M : btree[ref].M; }
static double *cell_L(int ref) { return ref < first_node ? leaf[ref].L : btree[ref].L; }
static double *cell_c(int ref) { return ref < first_node ? leaf[ref].center : btree[ref].center; }

void walk_m2m(int iNode)
{
    for (int n = 0; n < 2; n++) {
        int idx = btree[iNode].son[n];
        if (idx >= first_node) walk_m2m(idx);
        double *c = cell_c(idx);
        m2m(btree[iNode].center[0] - c[0], btree[iNode].center[1] - c[1], btree[iNode].center[2] - c[2],
            cell_M(idx), btree[iNode].M);
    }
}

void walk_l2l(int inode)
{
    if (inode < first_node) return;
    for (int n = 0; n < 2; n++) {
        int idx = btree[inode].son[n];
        if (idx < first_leaf) return;
        double *c = cell_c(idx);
        l2l(c[0] - btree[inode].center[0], c[1] - btree[inode].center[1], c[2] - btree[inode].center[2],
            btree[inode].L, cell_L(idx));
        if (idx >= first_node) walk_l2l(idx);
    }
}

int acceptance(double wi[3], double wj[3], double dist[3])
{
    return photons_acceptance(wi, wj, dist, open_angle, cutoffRadius);
}

/* ---- pair kernels of inc/kernels.h */

static void pair_p2p(int inode, int jnode, int precision)
{
    photons_ctx *x = ctx();
    int32_t t = inode - first_leaf, s = jnode - first_leaf;
    int8_t sh = 0;
    if (t < 0 || t >= g_nleaf || s < 0 || s >= g_nleaf) {
        fatal("p2p_kernel needs two leaves");
    }
    double *acc = (double *)calloc(3 * (size_t)NPART, sizeof(double));
    check(photons_p2p_pairs(x, 1, &t, &s, &sh, precision, acc), "p2p_kernel");
    int ip = leaf[inode].ipart;
    for (int i = ip; i < ip + leaf[inode].npart; i++)
        for (int d = 0; d < 3; d++) part[i].acc[d] += acc[3 * (size_t)i + d];
    free(acc);
}

void p2p_kernel(int inode, int jnode) { pair_p2p(inode, jnode, PHOTONS_FP64); }
void p2p_kernel_inleaf(int inode, int jnode) { pair_p2p(inode, jnode, PHOTONS_FP64); }
void p2p_kernel_rinvsqrt(int inode, int jnode) { pair_p2p(inode, jnode, PHOTONS_FP32_MIXED); }

/* ---- walks */

static void walk_eval(int im, int jm, int shift_lo, int shift_hi, int do_p2p, int do_m2l)
{
    photons_ctx *x = ctx();
    if (im == -1 || jm == -1) return;
    int a = to_cell(im), b = to_cell(jm);
    double *acc = NULL;
    if (do_p2p) acc = (double *)calloc(3 * (size_t)NPART, sizeof(double));
    if (do_m2l && g_precision == PHOTONS_FP64 && g_leaf_mem) {
        /* the device expansions start from the mirror's current L */
        int ncell = g_nleaf + g_nnode;
        double *L = (double *)malloc(sizeof(double) * NM3 * (size_t)ncell);
        for (int k = 0; k < ncell; k++)
            memcpy(L + (size_t)NM3 * k, k < g_nleaf ? g_leaf_mem[k].L : g_node_mem[k - g_nleaf].L,
                   sizeof(double) * NM3);
        check(photons_local_expansions(x, L, 1), "local expansions");
        free(L);
    }
    int64_t np = 0, nm = 0;
    check(photons_walk_eval(x, a, b, shift_lo, shift_hi, do_p2p, do_m2l, acc, &np, &nm), "walk");
    if (do_p2p) {
        for (int i = 0; i < NPART; i++)
            for (int d = 0; d < 3; d++) part[i].acc[d] += acc[3 * (size_t)i + d];
        free(acc);
        if (shift_lo == 0) idxP2P += (unsigned long)np;
    }
    if (do_m2l) {
        if (shift_lo == 0) idxM2L += (unsigned long)nm;
        walk_m2l_count += (unsigned long)nm;
        pull_local_expansions();
    }
}

void walk_task_p2p(int im, int jm) { walk_eval(im, jm, 0, 1, 1, 0); }
void walk_task_m2l(int im, int jm) { walk_eval(im, jm, 0, 1, 0, 1); }

static void need_root_source(int jm)
{
    if (jm != 0) {
        fatal("image walks start at the image tree root (jm = 0)");
    }
}

void walk_task_p2p_ext(int im, int jm)
{
    need_root_source(jm);
    walk_eval(im, first_node, g_shift_id, g_shift_id + 1, 1, 0);
}

void walk_task_m2l_ext(int im, int jm)
{
    need_root_source(jm);
    walk_eval(im, first_node, g_shift_id, g_shift_id + 1, 0, 1);
}

void fmm_remote(int idx, double shift[3])
{
    /* one rank: idx 0 is the home rank; shift is one of the 26 box periods */
    if (idx != 0) {
        fatal("fmm_remote: one rank, idx must be 0");
    }
    int m[3];
    for (int d = 0; d < 3; d++) m[d] = (int)lround(shift[d] / BOXSIZE);
    int k = (m[0] + 1) * 9 + (m[1] + 1) * 3 + (m[2] + 1);
    if (k == 13 || k < 0 || k > 26) return; /* no shift: nothing to add */
    g_shift_id = k < 13 ? k + 1 : k;
    /* fmm_remote_task: the image P2P walk, then the image M2L walk (src/remotes.c:631-682) */
    walk_eval(first_node, first_node, g_shift_id, g_shift_id + 1, 1, 1);
}

/* ---- PM: src/partmesh.c:800-1550 (partmesh), :11-16 (pm_thread) on one rank ------ */
int NSIDE; /* inc/photoNs.h:58 */

static photons_pm *g_pm;
static int g_pm_nside;
static double g_pm_box, g_pm_rs, g_pm_mass;
static double *g_pm_io;
static int64_t g_pm_cap;

/* part[].acc_pm = long-range force of part[] (overwritten, no G), on the GPU */
void partmesh(void)
{
    const double t0 = wall();
    if (!g_pm || g_pm_nside != NSIDE || g_pm_box != BOXSIZE || g_pm_rs != splitRadius || g_pm_mass != MASSPART) {
        if (g_pm) photons_pm_destroy(g_pm);
        g_pm = NULL;
        if (photons_pm_create(&g_pm, NSIDE, BOXSIZE, splitRadius, MASSPART, 0, NULL) != 0) {
            fatal("partmesh: %s", photons_pm_last_error());
        }
        g_pm_nside = NSIDE;
        g_pm_box = BOXSIZE;
        g_pm_rs = splitRadius;
        g_pm_mass = MASSPART;
    }
    if (NPART <= 0) return;
    if (NPART > g_pm_cap) {
        free(g_pm_io);
        g_pm_io = (double *)malloc(sizeof(double) * 6 * (size_t)NPART);
        if (!g_pm_io) {
            fatal("partmesh: out of host memory");
        }
        g_pm_cap = NPART;
    }
    double *pos = g_pm_io, *acc = g_pm_io + 3 * (size_t)NPART;
    for (int i = 0; i < NPART; i++)
        for (int k = 0; k < 3; k++) pos[3 * i + k] = part[i].pos[k];
    if (photons_pm_force_host(g_pm, NPART, pos, acc) != 0) {
        fatal("partmesh: %s", photons_pm_last_error());
    }
    for (int i = 0; i < NPART; i++)
        for (int k = 0; k < 3; k++) part[i].acc_pm[k] = acc[3 * i + k];
    dtime_pm = wall() - t0;
}

void partmesh_thread(void) { partmesh(); }

void *pm_thread(void *arg)
{
    (void)arg;
    partmesh();
    return NULL;
}
