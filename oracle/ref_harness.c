/*
 * oracle/ref_harness.c -- driver that runs the reference FMM subset on one rank.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from the
 * reference sources where they lie (src/fmm.c, src/operator.c, src/remotes.c,
 * src/toptree.c, src/utility.c) and linked against the MPICH installed in the
 * image.  It sets the globals the force call reads, calls the force path in the
 * driver's order (src/photoNs.c:97-116 without PM), and writes raw outputs.
 *
 *   refharness force <in> <out>   one short-range force evaluation
 *   refharness tree  <in> <out>   fmm_prepare only: the multipoles of every node and leaf
 *   refharness ops   <in> <out>   single operator calls (p2m, m2m, m2l, l2l, l2p, acceptance)
 *
 * force input : int32 n, maxleaf, pad, pad; double boxsize, mass, soft, rs, rc, theta; double pos[n][3]
 * force output: double acc[n][3] (input order); int64 idxP2P, idxM2L, walk_m2l_count,
 *               p2p_count_remote, nleaf, nnode
 * tree input  : as force; output: int64 nleaf, nnode; double M[nnode][NMULTI] of btree[]
 *               from first_node on, then M[nleaf][NMULTI] of leaf[] from first_leaf on
 * ops input   : int32 count, then records (see run_ops); ops output: raw doubles / int32.
 */
#include "photoNs.h"
#include "fmm.h"
#include "operator.h"

#include <stdint.h>
#include <string.h>

static void die(const char *m)
{
    fprintf(stderr, "refharness: %s\n", m);
    MPI_Abort(MPI_COMM_WORLD, 2);
}

static void rd(void *p, size_t sz, size_t n, FILE *f)
{
    if (fread(p, sz, n, f) != n) die("short read");
}

static void setup_one_rank(void)
{
    MPI_Comm_rank(MPI_COMM_WORLD, &PROC_RANK);
    MPI_Comm_size(MPI_COMM_WORLD, &PROC_SIZE);
    if (PROC_SIZE != 1) die("run on one rank");
    /* setup_domain_index() for one rank (src/initial.c:199-221) */
    mostleft = 0;
    this_domain = 0;
    first_domain = 0;
    last_domain = 0;
    /* MPI datatypes used by the image exchange (src/initial.c:236-240) */
    MPI_Type_contiguous(sizeof(RemoteNode), MPI_CHAR, &strReNode);
    MPI_Type_commit(&strReNode);
    MPI_Type_contiguous(sizeof(RemoteBody), MPI_CHAR, &strReBody);
    MPI_Type_commit(&strReBody);
    domtree = (DomainNode *)calloc(1, sizeof(DomainNode));
}

static int run_force(FILE *in, FILE *out, int tree_only)
{
    int32_t hdr[4];
    double prm[6];
    rd(hdr, sizeof(int32_t), 4, in);
    rd(prm, sizeof(double), 6, in);
    int n = hdr[0];
    NPART = n;
    NPART_TOTAL = n;
    MAXLEAF = hdr[1];
    BOXSIZE = prm[0];
    MASSPART = prm[1];
    SoftenScale = prm[2];
    splitRadius = prm[3];
    cutoffRadius = prm[4];
    open_angle = prm[5];

    part = (Body *)calloc((size_t)n, sizeof(Body));
    for (int i = 0; i < n; i++) {
        rd(part[i].pos, sizeof(double), 3, in);
        part[i].vel[0] = (double)i; /* carried through the in-place permutation */
    }
    idxP2P = idxM2L = 0;
    walk_m2l_count = 0;
    p2p_count_remote = 0;

    fmm_construct();
    fmm_prepare();
    if (tree_only) {
        /* build_localtree, p2m of every leaf and walk_m2m(first_node) have run */
        int64_t sz[2] = {(int64_t)(last_leaf - first_leaf), (int64_t)(last_node - first_node + 1)};
        fwrite(sz, sizeof(int64_t), 2, out);
        for (int k = first_node; k <= last_node; k++) fwrite(btree[k].M, sizeof(double), NMULTI, out);
        for (int k = first_leaf; k < last_leaf; k++) fwrite(leaf[k].M, sizeof(double), NMULTI, out);
        fmm_deconstruct();
        free(part);
        return 0;
    }
    fmm_task();
    int64_t cnt[6];
    cnt[0] = (int64_t)idxP2P;
    cnt[1] = (int64_t)idxM2L;
    fmm_ext();
    cnt[2] = (int64_t)walk_m2l_count;
    cnt[3] = (int64_t)p2p_count_remote;
    cnt[4] = (int64_t)(last_leaf - first_leaf);
    cnt[5] = (int64_t)(last_node - first_node + 1);

    double *acc = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        int id = (int)part[i].vel[0];
        memcpy(acc + 3 * (size_t)id, part[i].acc, 3 * sizeof(double));
    }
    fwrite(acc, sizeof(double), 3 * (size_t)n, out);
    fwrite(cnt, sizeof(int64_t), 6, out);
    fmm_deconstruct();
    free(acc);
    free(part);
    return 0;
}

/* ops records: int32 kind, then
 *   kind 0 p2m : int32 np; double mass, center[3], pos[np][3]      -> double M[NMULTI]
 *   kind 1 m2m : double d[3], M[NMULTI]                             -> double tM[NMULTI]
 *   kind 2 m2l : double rs, d[3], M[NMULTI]                         -> double L[NMULTI]
 *   kind 3 l2l : double d[3], L[NMULTI]                             -> double tL[NMULTI]
 *   kind 4 l2p : int32 np; double center[3], L[NMULTI], pos[np][3]  -> double acc[np][3]
 *   kind 5 acc : double theta, rc, wi[3], wj[3], dist[3]            -> int32 flag          */
static int run_ops(FILE *in, FILE *out)
{
    int32_t count;
    rd(&count, sizeof(int32_t), 1, in);
    for (int k = 0; k < count; k++) {
        int32_t kind;
        rd(&kind, sizeof(int32_t), 1, in);
        double M[NMULTI], L[NMULTI], d[3];
        if (kind == 0 || kind == 4) {
            int32_t np;
            rd(&np, sizeof(int32_t), 1, in);
            double mass = 0.0, c[3];
            if (kind == 0) rd(&mass, sizeof(double), 1, in);
            rd(c, sizeof(double), 3, in);
            if (kind == 4) rd(L, sizeof(double), NMULTI, in);
            part = (Body *)calloc((size_t)(np > 0 ? np : 1), sizeof(Body));
            for (int i = 0; i < np; i++) rd(part[i].pos, sizeof(double), 3, in);
            if (kind == 0) {
                MASSPART = mass;
                p2m(0, np, c, M);
                fwrite(M, sizeof(double), NMULTI, out);
            } else {
                Pack *lf = (Pack *)calloc(1, sizeof(Pack));
                lf->npart = np;
                lf->ipart = 0;
                memcpy(lf->center, c, sizeof c);
                memcpy(lf->L, L, sizeof L);
                leaf = lf;
                for (int i = 0; i < np; i++) part[i].acc[0] = part[i].acc[1] = part[i].acc[2] = 0.0;
                l2p(0);
                for (int i = 0; i < np; i++) fwrite(part[i].acc, sizeof(double), 3, out);
                free(lf);
                leaf = NULL;
            }
            free(part);
            part = NULL;
        } else if (kind == 1 || kind == 3) {
            rd(d, sizeof(double), 3, in);
            rd(M, sizeof(double), NMULTI, in);
            double t[NMULTI];
            memset(t, 0, sizeof t);
            if (kind == 1) m2m(d[0], d[1], d[2], M, t);
            else l2l(d[0], d[1], d[2], M, t);
            fwrite(t, sizeof(double), NMULTI, out);
        } else if (kind == 2) {
            double rs;
            rd(&rs, sizeof(double), 1, in);
            rd(d, sizeof(double), 3, in);
            rd(M, sizeof(double), NMULTI, in);
            splitRadius = rs;
            memset(L, 0, sizeof L);
            m2l(d[0], d[1], d[2], M, L);
            fwrite(L, sizeof(double), NMULTI, out);
        } else if (kind == 5) {
            double v[11];
            rd(v, sizeof(double), 11, in);
            open_angle = v[0];
            cutoffRadius = v[1];
            int32_t f = acceptance(v + 2, v + 5, v + 8);
            fwrite(&f, sizeof(int32_t), 1, out);
        } else {
            die("bad op kind");
        }
    }
    return 0;
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    if (argc != 4) die("usage: refharness force|tree|ops <in> <out>");
    setup_one_rank();
    FILE *in = fopen(argv[2], "rb");
    FILE *out = fopen(argv[3], "wb");
    if (!in || !out) die("cannot open files");
    int rc;
    if (strcmp(argv[1], "force") == 0) rc = run_force(in, out, 0);
    else if (strcmp(argv[1], "tree") == 0) rc = run_force(in, out, 1);
    else rc = run_ops(in, out);
    fclose(in);
    fclose(out);
    MPI_Finalize();
    return rc;
}
