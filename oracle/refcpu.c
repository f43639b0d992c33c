/*
 * oracle/refcpu.c -- FP64 CPU restatement of photoNs-2.0's short-range force path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the MI355X library, never part of it.
 * It follows, function by function (citations are to the reference tree):
 *   kd tree build ............ src/fmm.c:30-177 (bksort_inplace, build_kdtree, center_kdtree)
 *   acceptance (MAC) ......... src/fmm.c:267-326
 *   local dual-tree walk ..... src/fmm.c:406-538 (P2P), 569-712 (M2L)
 *   P2P pair kernel .......... src/fmm.c:823-867, src/remotes.c:26-56
 *   image LET prune .......... src/remotes.c:60-169 (prepare_sendtree2)
 *   image dual-tree walk ..... src/remotes.c:213-372 (P2P), 405-552 (M2L)
 *   operators ................ src/operator.c:13-528 (p2m, m2m, walk_m2m, l2p, m2l, l2l, walk_l2l)
 *   force call order ......... src/fmm.c:719-770, 909-984, 987-1075
 * The multipole operators are written for a general order p with multi-index
 * loops; for p = 3 they reduce to the reference's hand-expanded formulas.
 */
#include "refcpu.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXP 6
#define ORC_MAXT 84 /* nterms(6) */
#define ORC_NONE (-2147483647 - 1)

/* ------------------------------------------------------------------------ */
/* multi-index tables: terms ordered by total degree, then a descending, then
 * b descending (inc/operator.h:24-164 for degrees 0..4).                    */
static int g_tab_ready = 0;
static int mia[ORC_MAXT], mib[ORC_MAXT], mic[ORC_MAXT];
static double g_fact[2 * ORC_MAXP + 2];

static int mi_index(int a, int b, int c)
{
    int n = a + b + c, m = b + c;
    return n * (n + 1) * (n + 2) / 6 + m * (m + 1) / 2 + c;
}

static void tab_init(void)
{
    if (g_tab_ready) return;
    int k = 0;
    for (int n = 0; n <= ORC_MAXP; n++)
        for (int a = n; a >= 0; a--)
            for (int b = n - a; b >= 0; b--) {
                mia[k] = a; mib[k] = b; mic[k] = n - a - b;
                k++;
            }
    g_fact[0] = 1.0;
    for (int i = 1; i < 2 * ORC_MAXP + 2; i++) g_fact[i] = g_fact[i - 1] * i;
    g_tab_ready = 1;
}

int orc_nterms(int order) { return (order + 1) * (order + 2) * (order + 3) / 6; }

static double afact(int k) { return g_fact[mia[k]] * g_fact[mib[k]] * g_fact[mic[k]]; }

/* ------------------------------------------------------------------------ */
/* operators, src/operator.c                                                 */

void orc_p2m(int order, int n, const double *pos, double mass, const double *center, double *M)
{
    tab_init();
    int nt = orc_nterms(order);
    for (int k = 0; k < nt; k++) M[k] = 0.0; /* p2m overwrites, operator.c:18-19 */
    double px[ORC_MAXP + 1], py[ORC_MAXP + 1], pz[ORC_MAXP + 1];
    for (int i = 0; i < n; i++) {
        /* M_a = sum m (-dx)^a / a!, dx = x - center (operator.c:24-64) */
        double dx = -(pos[3 * i + 0] - center[0]);
        double dy = -(pos[3 * i + 1] - center[1]);
        double dz = -(pos[3 * i + 2] - center[2]);
        px[0] = py[0] = pz[0] = 1.0;
        for (int e = 1; e <= order; e++) {
            px[e] = px[e - 1] * dx; py[e] = py[e - 1] * dy; pz[e] = pz[e - 1] * dz;
        }
        for (int k = 0; k < nt; k++) M[k] += mass * px[mia[k]] * py[mib[k]] * pz[mic[k]] / afact(k);
    }
}

void orc_m2m(int order, const double *d, const double *M, double *tM)
{
    /* M'_a = sum_{b<=a} M_b d^(a-b)/(a-b)!, d = parent - child (operator.c:96-160) */
    tab_init();
    int nt = orc_nterms(order);
    double out[ORC_MAXT];
    for (int k = 0; k < nt; k++) {
        double s = 0.0;
        for (int j = 0; j < nt; j++) {
            int ga = mia[k] - mia[j], gb = mib[k] - mib[j], gc = mic[k] - mic[j];
            if (ga < 0 || gb < 0 || gc < 0) continue;
            s += M[j] * pow(d[0], ga) * pow(d[1], gb) * pow(d[2], gc) /
                 (g_fact[ga] * g_fact[gb] * g_fact[gc]);
        }
        out[k] = s;
    }
    for (int k = 0; k < nt; k++) tM[k] += out[k];
}

void orc_radial(double r, double rs, int nmax, double *fac)
{
    /* B_0 = erfc(r/2rs)/r,
     * B_n = [(2n-1) B_{n-1} + (2a^2)^n/(a sqrt(pi)) exp(-a^2 r^2)] / r^2,  a = 1/(2 rs),
     * fac[n] = (-1)^n B_n = (r^-1 d/dr)^n B_0.  For n <= 4 these equal the
     * closed forms of operator.c:301-305.                                   */
    double a = 0.5 / rs;
    double u = a * r;
    double e = exp(-u * u) / sqrt(M_PI);
    double ir2 = 1.0 / (r * r);
    double b = erfc(u) / r;
    double t = e / a;
    fac[0] = b;
    for (int n = 1; n <= nmax; n++) {
        t *= 2.0 * a * a;
        b = ((2 * n - 1) * b + t) * ir2;
        fac[n] = (n & 1) ? -b : b;
    }
}

static double hcoef(int a, int k)
{
    /* coefficient of the k-th contraction in d^a of a radial function:
     * a! / (2^k k! (a-2k)!) */
    return g_fact[a] / (ldexp(1.0, k) * g_fact[k] * g_fact[a - 2 * k]);
}

void orc_dtensor(int order, double rs, const double *R, double *D)
{
    tab_init();
    int nt = orc_nterms(order);
    double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    double f[ORC_MAXP + 1];
    orc_radial(r, rs, order, f);
    for (int k = 0; k < nt; k++) {
        int a = mia[k], b = mib[k], c = mic[k], n = a + b + c;
        double s = 0.0;
        for (int kx = 0; 2 * kx <= a; kx++)
            for (int ky = 0; 2 * ky <= b; ky++)
                for (int kz = 0; 2 * kz <= c; kz++) {
                    int kk = kx + ky + kz;
                    s += f[n - kk] * hcoef(a, kx) * hcoef(b, ky) * hcoef(c, kz) *
                         pow(R[0], a - 2 * kx) * pow(R[1], b - 2 * ky) * pow(R[2], c - 2 * kz);
                }
        D[k] = s;
    }
}

void orc_m2l(int order, double rs, const double *d, const double *M, double *L)
{
    /* L_a += sum_{|a|+|b|<=p} M_b D_{a+b}(d), d = c_target - c_source
     * (operator.c:255-392) */
    tab_init();
    int nt = orc_nterms(order);
    double D[ORC_MAXT];
    orc_dtensor(order, rs, d, D);
    for (int k = 0; k < nt; k++) {
        int nk = mia[k] + mib[k] + mic[k];
        double s = 0.0;
        for (int j = 0; j < nt; j++) {
            int nj = mia[j] + mib[j] + mic[j];
            if (nk + nj > order) continue;
            s += M[j] * D[mi_index(mia[k] + mia[j], mib[k] + mib[j], mic[k] + mic[j])];
        }
        L[k] += s;
    }
}

void orc_l2l(int order, const double *d, const double *L, double *tL)
{
    /* L'_a += sum_g L_{a+g} d^g/g!, d = child - parent (operator.c:395-494) */
    tab_init();
    int nt = orc_nterms(order);
    for (int k = 0; k < nt; k++) {
        int nk = mia[k] + mib[k] + mic[k];
        double s = 0.0;
        for (int g = 0; g < nt; g++) {
            int ng = mia[g] + mib[g] + mic[g];
            if (nk + ng > order) continue;
            s += L[mi_index(mia[k] + mia[g], mib[k] + mib[g], mic[k] + mic[g])] *
                 pow(d[0], mia[g]) * pow(d[1], mib[g]) * pow(d[2], mic[g]) / afact(g);
        }
        tL[k] += s;
    }
}

void orc_l2p(int order, const double *L, const double *center, int n, const double *pos, double *acc)
{
    /* acc_i += sum_{|a|<=p-1} L_{a+e_i} dx^a/a!, dx = x - center (operator.c:197-251) */
    tab_init();
    int nlow = orc_nterms(order - 1);
    for (int i = 0; i < n; i++) {
        double d[3] = {pos[3 * i] - center[0], pos[3 * i + 1] - center[1], pos[3 * i + 2] - center[2]};
        double f[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < nlow; k++) {
            double w = pow(d[0], mia[k]) * pow(d[1], mib[k]) * pow(d[2], mic[k]) / afact(k);
            f[0] += L[mi_index(mia[k] + 1, mib[k], mic[k])] * w;
            f[1] += L[mi_index(mia[k], mib[k] + 1, mic[k])] * w;
            f[2] += L[mi_index(mia[k], mib[k], mic[k] + 1)] * w;
        }
        acc[3 * i + 0] += f[0];
        acc[3 * i + 1] += f[1];
        acc[3 * i + 2] += f[2];
    }
}

/* ------------------------------------------------------------------------ */
/* pair kernel, src/fmm.c:834-852 (local) and src/remotes.c:30-50 (image)   */

static inline void pair_add(double dx, double dy, double dz, double mass, double soft, double rs,
                            double *ax, double *ay, double *az)
{
    double dr = sqrt(dx * dx + dy * dy + dz * dz);
    double k;
    if (dr < soft)
        k = mass / (soft * soft * soft);
    else
        k = mass / (dr * dr * dr);
    double u = 0.5 * dr / rs;
    k *= (erfc(u) + (2.0 / sqrt(M_PI)) * u * exp(-u * u));
    *ax += dx * k;
    *ay += dy * k;
    *az += dz * k;
}

void orc_p2p_pair(const double *xt, const double *xs, double mass, double soft, double rs, double *acc)
{
    pair_add(xs[0] - xt[0], xs[1] - xt[1], xs[2] - xt[2], mass, soft, rs, &acc[0], &acc[1], &acc[2]);
}

/* ------------------------------------------------------------------------ */
/* acceptance, src/fmm.c:267-326 (LONGSHORT build)                           */

int orc_acceptance(const double *wi, const double *wj, const double *dist, double theta, double rc)
{
    double w[3], gap[3];
    for (int i = 0; i < 3; i++) w[i] = (wi[i] + wj[i]) * 0.5;
    double dd2 = dist[0] * dist[0] + dist[1] * dist[1] + dist[2] * dist[2];
    for (int i = 0; i < 3; i++) {
        double g = dist[i] < 0.0 ? -dist[i] : dist[i];
        g -= w[i];
        gap[i] = g <= 0.0 ? 0.0 : g;
    }
    if (gap[0] + gap[1] + gap[2] < 0.0001) return 0; /* touching boxes always open */
    double dm2 = gap[0] * gap[0] + gap[1] * gap[1] + gap[2] * gap[2];
    double c2 = rc * rc;
    if (dm2 >= c2) return -1;
    if (dd2 > c2) return 0;
    double wmax = w[0];
    if (w[1] > wmax) wmax = w[1];
    if (w[2] > wmax) wmax = w[2];
    wmax *= 2;
    return (wmax * wmax < theta * theta * dd2) ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* tree                                                                      */

typedef struct {
    double x[3];
    int32_t id;
} opart;

typedef struct {
    int npart, ipart;
    double center[3], width[3];
} oleaf;

typedef struct {
    int npart;
    int son[2]; /* >=0 node id, <0 leaf -(id+1), ORC_NONE none */
    double split;
    double lo[3], hi[3];
    double center[3], width[3];
} onode;

struct orc_tree {
    orc_params prm;
    int n;
    opart *p;
    oleaf *leaf;
    int nleaf, capleaf;
    onode *node;
    int nnode, capnode;
    int maxdepth;
};

static int isleaf(int c) { return c < 0 && c != ORC_NONE; }
static int leafid(int c) { return -(c + 1); }

/* bksort_inplace (src/fmm.c:30-78): mean along axis D of the segment, then the
 * reference's in-place two-pointer partition; positions <= mean end on the
 * left.  The same scan order is kept so that the permutation, and with it the
 * summation order of every later mean, is the reference's.                  */
static void kd_partition(opart *p, int len, int D, int *nl, int *nr, double *split)
{
    *nl = *nr = 0;
    if (len < 2) {
        *nr = len;
        return;
    }
    if (len == 2) {
        *nl = *nr = 1;
        *split = 0.5 * (p[0].x[D] + p[1].x[D]);
        if (p[0].x[D] > p[1].x[D]) {
            opart t = p[0]; p[0] = p[1]; p[1] = t;
        }
        return;
    }
    double mean = 0.0;
    for (int i = 0; i < len; i++) mean += p[i].x[D];
    mean /= (double)len;
    int j = len - 1;
    for (int i = 0; i < j; i++) {
        if (p[i].x[D] > mean) {
            opart t = p[i];
            while (p[j].x[D] > mean && j > i) j--;
            p[i] = p[j];
            p[j] = t;
        }
    }
    *nl = j;
    *nr = len - j;
    *split = mean;
}

static int new_leaf(orc_tree *t)
{
    if (t->nleaf == t->capleaf) {
        t->capleaf = t->capleaf ? 2 * t->capleaf : 1024;
        t->leaf = (oleaf *)realloc(t->leaf, sizeof(oleaf) * t->capleaf);
    }
    memset(&t->leaf[t->nleaf], 0, sizeof(oleaf));
    return t->nleaf++;
}

static int new_node(orc_tree *t)
{
    if (t->nnode == t->capnode) {
        t->capnode = t->capnode ? 2 * t->capnode : 1024;
        t->node = (onode *)realloc(t->node, sizeof(onode) * t->capnode);
    }
    memset(&t->node[t->nnode], 0, sizeof(onode));
    t->node[t->nnode].son[0] = t->node[t->nnode].son[1] = ORC_NONE;
    return t->nnode++;
}

/* build_kdtree (src/fmm.c:81-120): nodes numbered in pre-order, leaves in the
 * order they are met, a child with npart <= maxleaf becomes a leaf.         */
static void kd_build(orc_tree *t, int D, int ip, int len, int inode, int depth)
{
    if (len == 0) return;
    if (depth > t->maxdepth) t->maxdepth = depth;
    t->node[inode].npart = len;
    int cnt[2];
    double split = 0.0;
    kd_partition(t->p + ip, len, D, &cnt[0], &cnt[1], &split);
    t->node[inode].split = split;
    for (int s = 0; s < 2; s++) {
        if (cnt[s] <= t->prm.maxleaf) {
            int l = new_leaf(t);
            t->leaf[l].npart = cnt[s];
            t->leaf[l].ipart = ip;
            t->node[inode].son[s] = -(l + 1);
            if (depth + 1 > t->maxdepth) t->maxdepth = depth + 1;
        } else {
            int c = new_node(t);
            t->node[inode].son[s] = c;
            kd_build(t, (D + 1) % 3, ip, cnt[s], c, depth + 1);
        }
        ip += cnt[s];
    }
}

/* center_kdtree (src/fmm.c:123-177): geometric boxes cut by the split planes */
static void kd_boxes(orc_tree *t, int D, int inode, const double *lo, const double *hi)
{
    onode *nd = &t->node[inode];
    for (int i = 0; i < 3; i++) {
        nd->lo[i] = lo[i];
        nd->hi[i] = hi[i];
        nd->width[i] = hi[i] - lo[i];
        nd->center[i] = 0.5 * (hi[i] + lo[i]);
    }
    for (int s = 0; s < 2; s++) {
        double clo[3], chi[3];
        memcpy(clo, lo, sizeof clo);
        memcpy(chi, hi, sizeof chi);
        if (s == 0) chi[D] = t->node[inode].split;
        else clo[D] = t->node[inode].split;
        int c = t->node[inode].son[s];
        if (c == ORC_NONE) continue;
        if (isleaf(c)) {
            oleaf *lf = &t->leaf[leafid(c)];
            for (int i = 0; i < 3; i++) {
                lf->width[i] = chi[i] - clo[i];
                lf->center[i] = 0.5 * (chi[i] + clo[i]);
            }
        } else {
            kd_boxes(t, (D + 1) % 3, c, clo, chi);
        }
    }
}

orc_tree *orc_build(const orc_params *prm, int n, const double *pos)
{
    tab_init();
    if (n < 0 || prm->maxleaf < 2) return NULL;
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    t->prm = *prm;
    t->n = n;
    t->p = (opart *)malloc(sizeof(opart) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        t->p[i].x[0] = pos[3 * i];
        t->p[i].x[1] = pos[3 * i + 1];
        t->p[i].x[2] = pos[3 * i + 2];
        t->p[i].id = i;
    }
    int root = new_node(t);
    kd_build(t, 0, 0, n, root, 0); /* axis 0 at the domain root on one rank (toptree.c:161-162) */
    double lo[3] = {0.0, 0.0, 0.0};
    double hi[3] = {prm->boxsize, prm->boxsize, prm->boxsize};
    kd_boxes(t, 0, root, lo, hi);
    return t;
}

void orc_free(orc_tree *t)
{
    if (!t) return;
    free(t->p);
    free(t->leaf);
    free(t->node);
    free(t);
}

void orc_tree_sizes(const orc_tree *t, int32_t *nleaf, int32_t *nnode, int32_t *maxdepth)
{
    *nleaf = t->nleaf;
    *nnode = t->nnode;
    *maxdepth = t->maxdepth;
}

void orc_tree_export(const orc_tree *t, int32_t *perm, int32_t *leaf_ipart, int32_t *leaf_npart,
                     double *leaf_center, double *leaf_width, int32_t *node_son, int32_t *node_npart,
                     double *node_split, double *node_center, double *node_width)
{
    for (int i = 0; i < t->n; i++)
        if (perm) perm[i] = t->p[i].id;
    for (int l = 0; l < t->nleaf; l++) {
        if (leaf_ipart) leaf_ipart[l] = t->leaf[l].ipart;
        if (leaf_npart) leaf_npart[l] = t->leaf[l].npart;
        for (int i = 0; i < 3; i++) {
            if (leaf_center) leaf_center[3 * l + i] = t->leaf[l].center[i];
            if (leaf_width) leaf_width[3 * l + i] = t->leaf[l].width[i];
        }
    }
    for (int k = 0; k < t->nnode; k++) {
        if (node_son) {
            node_son[2 * k] = t->node[k].son[0];
            node_son[2 * k + 1] = t->node[k].son[1];
        }
        if (node_npart) node_npart[k] = t->node[k].npart;
        if (node_split) node_split[k] = t->node[k].split;
        for (int i = 0; i < 3; i++) {
            if (node_center) node_center[3 * k + i] = t->node[k].center[i];
            if (node_width) node_width[3 * k + i] = t->node[k].width[i];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* walks                                                                     */

typedef struct {
    int32_t *t, *s;
    int8_t *sh;
    int64_t n, cap;
} olist;

static void list_push(olist *l, int32_t t, int32_t s, int8_t sh)
{
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 4096;
        l->t = (int32_t *)realloc(l->t, sizeof(int32_t) * l->cap);
        l->s = (int32_t *)realloc(l->s, sizeof(int32_t) * l->cap);
        l->sh = (int8_t *)realloc(l->sh, sizeof(int8_t) * l->cap);
    }
    l->t[l->n] = t;
    l->s[l->n] = s;
    l->sh[l->n] = sh;
    l->n++;
}

static void list_free(olist *l)
{
    free(l->t);
    free(l->s);
    free(l->sh);
    memset(l, 0, sizeof *l);
}

typedef struct {
    const orc_tree *t;
    olist p2p, m2l;
    int shift_id;
    double shift[3];
    double top_c[3], top_w[3]; /* domain box (toptree[this_domain]) */
    const uint8_t *tflag;      /* NULL, or by cell id: walk only targets flagged here */
} owalk;

static int cell_id(const orc_tree *t, int c) { return isleaf(c) ? leafid(c) : t->nleaf + c; }
static const double *cw(const orc_tree *t, int c) { return isleaf(c) ? t->leaf[leafid(c)].width : t->node[c].width; }
static const double *cc(const orc_tree *t, int c) { return isleaf(c) ? t->leaf[leafid(c)].center : t->node[c].center; }
static double wsum(const double *w) { return w[0] + w[1] + w[2]; }

/* walk_task_p2p + walk_task_m2l (src/fmm.c:406-538, 569-712).  Both walks make
 * the same decisions, so one recursion emits both lists in the order each of
 * the reference walks emits them.                                          */
static void walk_local(owalk *w, int im, int jm)
{
    const orc_tree *t = w->t;
    if (im == ORC_NONE || jm == ORC_NONE) return;
    if (w->tflag && !w->tflag[cell_id(t, im)]) return; /* no selected leaf below the target */
    if (im == jm) {
        if (isleaf(im)) {
            list_push(&w->p2p, leafid(im), leafid(jm), 0);
        } else {
            int a0 = t->node[im].son[0], a1 = t->node[im].son[1];
            walk_local(w, a0, a0);
            walk_local(w, a0, a1);
            walk_local(w, a1, a0);
            walk_local(w, a1, a1);
        }
        return;
    }
    if (isleaf(im) && isleaf(jm)) { /* leaf-leaf pairs are always P2P */
        list_push(&w->p2p, leafid(im), leafid(jm), 0);
        return;
    }
    double dist[3];
    const double *ci = cc(t, im), *cj = cc(t, jm);
    for (int i = 0; i < 3; i++) dist[i] = ci[i] - cj[i];
    int flag = orc_acceptance(cw(t, im), cw(t, jm), dist, t->prm.theta, t->prm.rc);
    if (flag == 1) {
        list_push(&w->m2l, cell_id(t, im), cell_id(t, jm), 0);
        return;
    }
    if (flag == -1) return;
    int open_target;
    if (isleaf(im)) open_target = 0;
    else if (isleaf(jm)) open_target = 1;
    else open_target = wsum(t->node[im].width) > wsum(t->node[jm].width);
    if (open_target) {
        walk_local(w, t->node[im].son[0], jm);
        walk_local(w, t->node[im].son[1], jm);
    } else {
        walk_local(w, im, t->node[jm].son[0]);
        walk_local(w, im, t->node[jm].son[1]);
    }
}

/* prepare_sendtree2 (src/remotes.c:97-158): a node of the shifted tree keeps
 * no children when its box is rc or more away from the domain box, or when it
 * is small enough that 0.95*theta*distance exceeds its largest width.       */
static int let_pruned(const owalk *w, int jm)
{
    const orc_tree *t = w->t;
    const onode *nd = &t->node[jm];
    double d[3];
    for (int i = 0; i < 3; i++) {
        double x = w->top_c[i] - nd->center[i] - w->shift[i];
        if (x < 0.0) x = -x;
        d[i] = x - (w->top_w[i] + nd->width[i]) * 0.5;
    }
    double dr = 0.0;
    for (int i = 0; i < 3; i++)
        if (d[i] > 0.0) dr += d[i] * d[i];
    dr = sqrt(dr);
    double wmax = nd->width[0];
    if (wmax < nd->width[1]) wmax = nd->width[1];
    if (wmax < nd->width[2]) wmax = nd->width[2];
    if (dr >= t->prm.rc) return 1;
    return wmax < 0.95 * t->prm.theta * dr;
}

/* walk_task_p2p_ext + walk_task_m2l_ext (src/remotes.c:213-372, 405-552):
 * target = home tree, source = tree shifted by one box period.            */
static void walk_ext(owalk *w, int im, int jm)
{
    const orc_tree *t = w->t;
    int8_t sh = (int8_t)w->shift_id;
    if (w->tflag && !w->tflag[cell_id(t, im)]) return; /* no selected leaf below the target */
    if (isleaf(im) && isleaf(jm)) {
        list_push(&w->p2p, leafid(im), leafid(jm), sh);
        return;
    }
    double dist[3];
    const double *ci = cc(t, im), *cj = cc(t, jm);
    for (int i = 0; i < 3; i++) dist[i] = ci[i] - (cj[i] + w->shift[i]);
    int flag = orc_acceptance(cw(t, im), cw(t, jm), dist, t->prm.theta, t->prm.rc);
    if (flag == -1) return;
    int pruned = isleaf(jm) ? 0 : let_pruned(w, jm);
    if (isleaf(im)) { /* source is a node here */
        if (flag == 1 || pruned) {
            list_push(&w->m2l, cell_id(t, im), cell_id(t, jm), sh);
        } else {
            walk_ext(w, im, t->node[jm].son[0]);
            walk_ext(w, im, t->node[jm].son[1]);
        }
        return;
    }
    if (flag == 1) {
        list_push(&w->m2l, cell_id(t, im), cell_id(t, jm), sh);
        return;
    }
    int open_target;
    if (isleaf(jm)) open_target = 1;
    else open_target = wsum(t->node[im].width) > wsum(t->node[jm].width) || pruned;
    if (open_target) {
        walk_ext(w, t->node[im].son[0], jm);
        walk_ext(w, t->node[im].son[1], jm);
    } else {
        walk_ext(w, im, t->node[jm].son[0]);
        walk_ext(w, im, t->node[jm].son[1]);
    }
}

static void shift_vec(int sid, double L, double *s)
{
    /* ids 1..26 in loop order mi, mj, mk over {-1,0,1}, skipping (0,0,0) */
    int k = 0;
    for (int mi = -1; mi <= 1; mi++)
        for (int mj = -1; mj <= 1; mj++)
            for (int mk = -1; mk <= 1; mk++) {
                if (mi == 0 && mj == 0 && mk == 0) continue;
                k++;
                if (k == sid) {
                    s[0] = (double)mi * L;
                    s[1] = (double)mj * L;
                    s[2] = (double)mk * L;
                    return;
                }
            }
    s[0] = s[1] = s[2] = 0.0;
}

static void run_walks(const orc_tree *t, owalk *w, int sid, const uint8_t *tflag)
{
    memset(w, 0, sizeof *w);
    w->tflag = tflag;
    w->t = t;
    w->shift_id = sid;
    shift_vec(sid, t->prm.boxsize, w->shift);
    for (int i = 0; i < 3; i++) {
        w->top_c[i] = 0.5 * (t->prm.boxsize + 0.0);
        w->top_w[i] = t->prm.boxsize - 0.0;
    }
    if (t->n == 0) return;
    if (sid == 0) walk_local(w, 0, 0);
    else walk_ext(w, 0, 0);
}

static int64_t collect(const orc_tree *t, int which, int32_t *tg, int32_t *sr, int8_t *sh, int64_t cap)
{
    int64_t total = 0;
    int nshift = t->prm.periodic ? 27 : 1;
    for (int sid = 0; sid < nshift; sid++) {
        owalk w;
        run_walks(t, &w, sid, NULL);
        olist *l = which ? &w.m2l : &w.p2p;
        for (int64_t k = 0; k < l->n; k++) {
            if (tg && total < cap) {
                tg[total] = l->t[k];
                sr[total] = l->s[k];
                sh[total] = l->sh[k];
            }
            total++;
        }
        list_free(&w.p2p);
        list_free(&w.m2l);
    }
    return total;
}

/* Size of the shifted tree prepare_sendtree2 would send for image sid
 * (src/remotes.c:60-169): every leaf reached copies its bodies; a pruned node
 * keeps no children.                                                        */
static void let_count(const owalk *w, int c, int64_t *nn, int64_t *nb)
{
    const orc_tree *t = w->t;
    (*nn)++;
    if (isleaf(c)) {
        *nb += t->leaf[leafid(c)].npart;
        return;
    }
    if (let_pruned(w, c)) return;
    for (int s = 0; s < 2; s++)
        if (t->node[c].son[s] != ORC_NONE) let_count(w, t->node[c].son[s], nn, nb);
}

void orc_let_size(const orc_tree *t, int sid, int64_t *nnodes, int64_t *nbodies)
{
    owalk w;
    memset(&w, 0, sizeof w);
    w.t = t;
    w.shift_id = sid;
    shift_vec(sid, t->prm.boxsize, w.shift);
    for (int i = 0; i < 3; i++) {
        w.top_c[i] = 0.5 * (t->prm.boxsize + 0.0);
        w.top_w[i] = t->prm.boxsize - 0.0;
    }
    *nnodes = *nbodies = 0;
    if (t->n > 0) let_count(&w, 0, nnodes, nbodies);
}

int64_t orc_p2p_list(const orc_tree *t, int32_t *tgt, int32_t *src, int8_t *shift, int64_t cap)
{
    return collect(t, 0, tgt, src, shift, cap);
}

int64_t orc_m2l_list(const orc_tree *t, int32_t *tgt, int32_t *src, int8_t *shift, int64_t cap)
{
    return collect(t, 1, tgt, src, shift, cap);
}

/* ------------------------------------------------------------------------ */
/* the force call                                                            */

static void walk_m2m(const orc_tree *t, int order, int inode, double *Mleaf, double *Mnode)
{
    /* walk_m2m (src/operator.c:165-194): post-order, children in son order */
    int nt = orc_nterms(order);
    const onode *nd = &t->node[inode];
    for (int s = 0; s < 2; s++) {
        int c = nd->son[s];
        if (c == ORC_NONE) continue;
        double d[3];
        const double *Mc;
        if (isleaf(c)) {
            const oleaf *lf = &t->leaf[leafid(c)];
            for (int i = 0; i < 3; i++) d[i] = nd->center[i] - lf->center[i];
            Mc = Mleaf + (size_t)nt * leafid(c);
        } else {
            walk_m2m(t, order, c, Mleaf, Mnode);
            for (int i = 0; i < 3; i++) d[i] = nd->center[i] - t->node[c].center[i];
            Mc = Mnode + (size_t)nt * c;
        }
        orc_m2m(order, d, Mc, Mnode + (size_t)nt * inode);
    }
}

static void walk_l2l(const orc_tree *t, int order, int inode, double *Lleaf, double *Lnode)
{
    /* walk_l2l (src/operator.c:498-528): pre-order */
    int nt = orc_nterms(order);
    const onode *nd = &t->node[inode];
    for (int s = 0; s < 2; s++) {
        int c = nd->son[s];
        if (c == ORC_NONE) return;
        double d[3];
        if (isleaf(c)) {
            const oleaf *lf = &t->leaf[leafid(c)];
            for (int i = 0; i < 3; i++) d[i] = lf->center[i] - nd->center[i];
            orc_l2l(order, d, Lnode + (size_t)nt * inode, Lleaf + (size_t)nt * leafid(c));
        } else {
            for (int i = 0; i < 3; i++) d[i] = t->node[c].center[i] - nd->center[i];
            orc_l2l(order, d, Lnode + (size_t)nt * inode, Lnode + (size_t)nt * c);
            walk_l2l(t, order, c, Lleaf, Lnode);
        }
    }
}

static double *cell_vec(const orc_tree *t, int nt, double *leafv, double *nodev, int cid)
{
    return cid < t->nleaf ? leafv + (size_t)nt * cid : nodev + (size_t)nt * (cid - t->nleaf);
}

static const double *cell_center(const orc_tree *t, int cid)
{
    return cid < t->nleaf ? t->leaf[cid].center : t->node[cid - t->nleaf].center;
}

int orc_force(const orc_params *prm, int n, const double *pos, double *acc, orc_stats *st)
{
    tab_init();
    if (st) memset(st, 0, sizeof *st);
    if (n < 0 || prm->order < 1 || prm->order > ORC_MAXP || prm->maxleaf < 2) return -1;
    if (n == 0) return 0;
    if (prm->periodic && n <= prm->maxleaf) return -2; /* the image tree root would read as a leaf */
    orc_tree *t = orc_build(prm, n, pos);
    int order = prm->order, nt = orc_nterms(order);
    double *Mleaf = (double *)calloc((size_t)nt * t->nleaf, sizeof(double));
    double *Lleaf = (double *)calloc((size_t)nt * t->nleaf, sizeof(double));
    double *Mnode = (double *)calloc((size_t)nt * t->nnode, sizeof(double));
    double *Lnode = (double *)calloc((size_t)nt * t->nnode, sizeof(double));
    double *a = (double *)calloc((size_t)3 * n, sizeof(double)); /* tree order */
    double *xyz = (double *)malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) xyz[3 * i + k] = t->p[i].x[k];

    /* fmm_prepare: P2M per leaf, M2M up the tree (src/fmm.c:741-744) */
    for (int l = 0; l < t->nleaf; l++)
        orc_p2m(order, t->leaf[l].npart, xyz + 3 * t->leaf[l].ipart, prm->mass, t->leaf[l].center,
                Mleaf + (size_t)nt * l);
    walk_m2m(t, order, 0, Mleaf, Mnode);

    int nshift = prm->periodic ? 27 : 1;
    for (int sid = 0; sid < nshift; sid++) {
        owalk w;
        run_walks(t, &w, sid, NULL);
        /* P2P tasks */
        for (int64_t k = 0; k < w.p2p.n; k++) {
            const oleaf *lt = &t->leaf[w.p2p.t[k]];
            const oleaf *ls = &t->leaf[w.p2p.s[k]];
            for (int q = 0; q < lt->npart; q++) {
                int ip = lt->ipart + q;
                const double *xt = xyz + 3 * ip;
                if (sid == 0) {
                    /* task_compute_p2p: per-task sum, then added (src/fmm.c:823-867) */
                    double sx = 0.0, sy = 0.0, sz = 0.0;
                    for (int jp = ls->ipart; jp < ls->ipart + ls->npart; jp++) {
                        if (jp == ip) continue;
                        const double *xs = xyz + 3 * jp;
                        pair_add(xs[0] - xt[0], xs[1] - xt[1], xs[2] - xt[2], prm->mass, prm->soft, prm->rs,
                                 &sx, &sy, &sz);
                        if (st) st->p2p_inter++;
                    }
                    a[3 * ip] += sx;
                    a[3 * ip + 1] += sy;
                    a[3 * ip + 2] += sz;
                } else {
                    /* p2p_kernel_ex: image body = pos + shift, accumulated directly */
                    for (int jp = ls->ipart; jp < ls->ipart + ls->npart; jp++) {
                        const double *xs = xyz + 3 * jp;
                        double bx = xs[0] + w.shift[0], by = xs[1] + w.shift[1], bz = xs[2] + w.shift[2];
                        pair_add(bx - xt[0], by - xt[1], bz - xt[2], prm->mass, prm->soft, prm->rs, &a[3 * ip],
                                 &a[3 * ip + 1], &a[3 * ip + 2]);
                        if (st) {
                            st->p2p_inter++;
                            st->p2p_ext_inter++;
                        }
                    }
                }
            }
        }
        /* M2L tasks */
        for (int64_t k = 0; k < w.m2l.n; k++) {
            int ct = w.m2l.t[k], cs = w.m2l.s[k];
            const double *c1 = cell_center(t, ct), *c2 = cell_center(t, cs);
            double d[3];
            for (int i = 0; i < 3; i++) d[i] = c1[i] - (c2[i] + w.shift[i]);
            orc_m2l(order, prm->rs, d, cell_vec(t, nt, Mleaf, Mnode, cs), cell_vec(t, nt, Lleaf, Lnode, ct));
        }
        if (st) {
            if (sid == 0) {
                st->p2p_local_pairs += w.p2p.n;
                st->m2l_local_pairs += w.m2l.n;
            } else {
                st->p2p_ext_pairs += w.p2p.n;
                st->m2l_ext_pairs += w.m2l.n;
            }
            st->m2l_calls += w.m2l.n;
        }
        list_free(&w.p2p);
        list_free(&w.m2l);
    }
    /* L2L down the tree, then L2P per leaf (src/fmm.c:1054-1057) */
    walk_l2l(t, order, 0, Lleaf, Lnode);
    for (int l = 0; l < t->nleaf; l++)
        orc_l2p(order, Lleaf + (size_t)nt * l, t->leaf[l].center, t->leaf[l].npart, xyz + 3 * t->leaf[l].ipart,
                a + 3 * t->leaf[l].ipart);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) acc[3 * t->p[i].id + k] = a[3 * i + k];
    if (st) {
        st->nleaf = t->nleaf;
        st->nnode = t->nnode;
        st->maxdepth = t->maxdepth;
    }
    free(Mleaf); free(Lleaf); free(Mnode); free(Lnode); free(a); free(xyz);
    orc_free(t);
    return 0;
}

/* flag[cell] = 1 for the selected leaves and every node above one of them */
static int flag_up(const orc_tree *t, int c, uint8_t *flag)
{
    if (c == ORC_NONE) return 0;
    if (isleaf(c)) return flag[leafid(c)];
    int f = flag_up(t, t->node[c].son[0], flag) | flag_up(t, t->node[c].son[1], flag);
    flag[t->nleaf + c] = (uint8_t)f;
    return f;
}

/* orc_force restricted to selected target leaves: the whole tree and every multipole,
 * the walks of orc_force entered only for targets with a selected leaf below them (the
 * same decisions, so the selected leaves get the same lists), M2L into those targets,
 * L2L down, then P2P and L2P of the selected leaves alone.  acc_sel: the selected
 * leaves' particles in tree order, leaf after leaf; id_sel: their caller indices;
 * p2p_len / m2l_len: per selected leaf, the entries of its P2P list (local and image)
 * and of the M2L lists of the leaf and its ancestors.  Returns the particle count, or
 * < 0 on bad arguments. */
int64_t orc_force_leaves(const orc_params *prm, int n, const double *pos, int nsel, const int32_t *sel,
                         double *acc_sel, int32_t *id_sel, int64_t *p2p_len, int64_t *m2l_len)
{
    tab_init();
    if (n < 1 || prm->order < 1 || prm->order > ORC_MAXP || prm->maxleaf < 2) return -1;
    if (prm->periodic && n <= prm->maxleaf) return -2;
    orc_tree *t = orc_build(prm, n, pos);
    for (int k = 0; k < nsel; k++)
        if (sel[k] < 0 || sel[k] >= t->nleaf) {
            orc_free(t);
            return -3;
        }
    const int ncell = t->nleaf + t->nnode;
    int order = prm->order, nt = orc_nterms(order);
    uint8_t *flag = (uint8_t *)calloc((size_t)ncell, 1);
    for (int k = 0; k < nsel; k++) flag[sel[k]] = 1;
    flag_up(t, 0, flag);
    double *Mleaf = (double *)calloc((size_t)nt * t->nleaf, sizeof(double));
    double *Lleaf = (double *)calloc((size_t)nt * t->nleaf, sizeof(double));
    double *Mnode = (double *)calloc((size_t)nt * t->nnode, sizeof(double));
    double *Lnode = (double *)calloc((size_t)nt * t->nnode, sizeof(double));
    double *a = (double *)calloc((size_t)3 * n, sizeof(double)); /* tree order */
    double *xyz = (double *)malloc(sizeof(double) * 3 * n);
    int64_t *plen = (int64_t *)calloc((size_t)t->nleaf, sizeof(int64_t));
    int64_t *mlen = (int64_t *)calloc((size_t)ncell, sizeof(int64_t));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) xyz[3 * i + k] = t->p[i].x[k];
    for (int l = 0; l < t->nleaf; l++)
        orc_p2m(order, t->leaf[l].npart, xyz + 3 * t->leaf[l].ipart, prm->mass, t->leaf[l].center,
                Mleaf + (size_t)nt * l);
    walk_m2m(t, order, 0, Mleaf, Mnode);
    int nshift = prm->periodic ? 27 : 1;
    for (int sid = 0; sid < nshift; sid++) {
        owalk w;
        run_walks(t, &w, sid, flag);
        for (int64_t k = 0; k < w.p2p.n; k++) {
            const oleaf *lt = &t->leaf[w.p2p.t[k]];
            const oleaf *ls = &t->leaf[w.p2p.s[k]];
            plen[w.p2p.t[k]]++;
            for (int q = 0; q < lt->npart; q++) {
                int ip = lt->ipart + q;
                const double *xt = xyz + 3 * ip;
                if (sid == 0) {
                    double sx = 0.0, sy = 0.0, sz = 0.0;
                    for (int jp = ls->ipart; jp < ls->ipart + ls->npart; jp++) {
                        if (jp == ip) continue;
                        const double *xs = xyz + 3 * jp;
                        pair_add(xs[0] - xt[0], xs[1] - xt[1], xs[2] - xt[2], prm->mass, prm->soft, prm->rs,
                                 &sx, &sy, &sz);
                    }
                    a[3 * ip] += sx;
                    a[3 * ip + 1] += sy;
                    a[3 * ip + 2] += sz;
                } else {
                    for (int jp = ls->ipart; jp < ls->ipart + ls->npart; jp++) {
                        const double *xs = xyz + 3 * jp;
                        double bx = xs[0] + w.shift[0], by = xs[1] + w.shift[1], bz = xs[2] + w.shift[2];
                        pair_add(bx - xt[0], by - xt[1], bz - xt[2], prm->mass, prm->soft, prm->rs, &a[3 * ip],
                                 &a[3 * ip + 1], &a[3 * ip + 2]);
                    }
                }
            }
        }
        for (int64_t k = 0; k < w.m2l.n; k++) {
            int ct = w.m2l.t[k], cs = w.m2l.s[k];
            const double *c1 = cell_center(t, ct), *c2 = cell_center(t, cs);
            double d[3];
            for (int i = 0; i < 3; i++) d[i] = c1[i] - (c2[i] + w.shift[i]);
            orc_m2l(order, prm->rs, d, cell_vec(t, nt, Mleaf, Mnode, cs), cell_vec(t, nt, Lleaf, Lnode, ct));
            mlen[ct]++;
        }
        list_free(&w.p2p);
        list_free(&w.m2l);
    }
    walk_l2l(t, order, 0, Lleaf, Lnode);
    /* the M2L entries above each leaf: its own and those of its ancestors */
    int64_t *above = (int64_t *)calloc((size_t)ncell, sizeof(int64_t));
    for (int v = 0; v < t->nnode; v++) { /* pre-order: a parent before its children */
        int64_t mine = above[t->nleaf + v] + mlen[t->nleaf + v];
        for (int s2 = 0; s2 < 2; s2++) {
            int c = t->node[v].son[s2];
            if (c == ORC_NONE) continue;
            above[cell_id(t, c)] = mine;
        }
    }
    int64_t m = 0;
    for (int k = 0; k < nsel; k++) {
        const oleaf *lf = &t->leaf[sel[k]];
        orc_l2p(order, Lleaf + (size_t)nt * sel[k], lf->center, lf->npart, xyz + 3 * lf->ipart, a + 3 * lf->ipart);
        for (int q = 0; q < lf->npart; q++) {
            int ip = lf->ipart + q;
            for (int d = 0; d < 3; d++) acc_sel[3 * m + d] = a[3 * ip + d];
            id_sel[m] = t->p[ip].id;
            m++;
        }
        if (p2p_len) p2p_len[k] = plen[sel[k]];
        if (m2l_len) m2l_len[k] = above[sel[k]] + mlen[sel[k]];
    }
    free(above); free(plen); free(mlen); free(flag);
    free(Mleaf); free(Lleaf); free(Mnode); free(Lnode); free(a); free(xyz);
    orc_free(t);
    return m;
}

/* ------------------------------------------------------------------------ */
/* direct sum with a periodic cell list                                      */

typedef struct {
    const orc_params *prm;
    const double *pos;
    const int32_t *targets;
    double *acc;
    int nt, t0, t1;
    double rmax;
    int ncell;
    const int32_t *cstart, *cidx;
    int n;
} dsum_job;

static int cell_of(double x, double L, int nc)
{
    int c = (int)floor(x / L * nc);
    c %= nc;
    if (c < 0) c += nc;
    return c;
}

static void *dsum_worker(void *arg)
{
    dsum_job *j = (dsum_job *)arg;
    const orc_params *p = j->prm;
    double L = p->boxsize, r2max = j->rmax * j->rmax;
    int nc = j->ncell;
    for (int k = j->t0; k < j->t1; k++) {
        int i = j->targets[k];
        const double *xt = j->pos + 3 * i;
        int ci[3];
        for (int d = 0; d < 3; d++) ci[d] = cell_of(xt[d], L, nc);
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int ox = -1; ox <= 1; ox++)
            for (int oy = -1; oy <= 1; oy++)
                for (int oz = -1; oz <= 1; oz++) {
                    int cx = (ci[0] + ox + nc) % nc, cy = (ci[1] + oy + nc) % nc, cz = (ci[2] + oz + nc) % nc;
                    int c = (cx * nc + cy) * nc + cz;
                    for (int q = j->cstart[c]; q < j->cstart[c + 1]; q++) {
                        int s = j->cidx[q];
                        if (s == i) continue;
                        double dx = j->pos[3 * s] - xt[0], dy = j->pos[3 * s + 1] - xt[1], dz = j->pos[3 * s + 2] - xt[2];
                        dx -= L * nearbyint(dx / L);
                        dy -= L * nearbyint(dy / L);
                        dz -= L * nearbyint(dz / L);
                        if (dx * dx + dy * dy + dz * dz >= r2max) continue;
                        pair_add(dx, dy, dz, p->mass, p->soft, p->rs, &ax, &ay, &az);
                    }
                }
        j->acc[3 * k] = ax;
        j->acc[3 * k + 1] = ay;
        j->acc[3 * k + 2] = az;
    }
    return NULL;
}

int orc_direct(const orc_params *prm, int n, const double *pos, int nt, const int32_t *targets, double rmax,
               double *acc, int nthreads)
{
    double L = prm->boxsize;
    if (!(rmax > 0.0) || 3.0 * rmax > L) return -1;
    int nc = (int)floor(L / rmax);
    if (nc < 3) return -1;
    if (nc > 256) nc = 256;
    int ncell = nc * nc * nc;
    int32_t *cnt = (int32_t *)calloc((size_t)ncell + 1, sizeof(int32_t));
    int32_t *cidx = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    int32_t *cof = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        int c = (cell_of(pos[3 * i], L, nc) * nc + cell_of(pos[3 * i + 1], L, nc)) * nc + cell_of(pos[3 * i + 2], L, nc);
        cof[i] = c;
        cnt[c + 1]++;
    }
    for (int c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * ncell);
    memcpy(fill, cnt, sizeof(int32_t) * ncell);
    for (int i = 0; i < n; i++) cidx[fill[cof[i]]++] = i;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    dsum_job jobs[256];
    for (int k = 0; k < nthreads; k++) {
        jobs[k].prm = prm;
        jobs[k].pos = pos;
        jobs[k].targets = targets;
        jobs[k].acc = acc;
        jobs[k].nt = nt;
        jobs[k].t0 = (int)((long long)nt * k / nthreads);
        jobs[k].t1 = (int)((long long)nt * (k + 1) / nthreads);
        jobs[k].rmax = rmax;
        jobs[k].ncell = nc;
        jobs[k].cstart = cnt;
        jobs[k].cidx = cidx;
        jobs[k].n = n;
        pthread_create(&th[k], NULL, dsum_worker, &jobs[k]);
    }
    for (int k = 0; k < nthreads; k++) pthread_join(th[k], NULL);
    free(cnt); free(cidx); free(cof); free(fill);
    return 0;
}

/* ---------------------------------------------------------------------------- time integration */

/* 128-block composite Simpson in ln a of a^-k / H(a), H(a) = 0.1 sqrt(Om a^-3 + OL),
 * k = 1 for the kick and 2 for the drift (src/initial.c:650-683: weights 1, 4, 2,
 * ..., 4, 1 over 3, the integrand at ln a_i + n dloga, the end points from exp). */
static double simpson_loga(double loga_i, double loga_f, double om, double ox, int k)
{
    const int nblock = 128;
    const double dloga = (loga_f - loga_i) / nblock;
    double z = 1.0 / exp(loga_i);
    double h = 0.1 * sqrt(om * z * z * z + ox);
    double zk = k == 1 ? z : z * z;
    double sum = dloga * zk / h;
    for (int j = 1; j < nblock; j++) {
        z = 1.0 / exp(loga_i + dloga * j);
        h = 0.1 * sqrt(om * z * z * z + ox);
        zk = k == 1 ? z : z * z;
        sum += 2.0 * (1 + j % 2) * dloga * zk / h;
    }
    z = 1.0 / exp(loga_f);
    h = 0.1 * sqrt(om * z * z * z + ox);
    zk = k == 1 ? z : z * z;
    sum += dloga * zk / h;
    return sum / 3.0;
}

double orc_kick_loga(double loga_i, double loga_f, double omega_m, double omega_x)
{
    return simpson_loga(loga_i, loga_f, omega_m, omega_x, 1);
}

double orc_drift_loga(double loga_i, double loga_f, double omega_m, double omega_x)
{
    return simpson_loga(loga_i, loga_f, omega_m, omega_x, 2);
}

/* src/photoNs.c:208-268 without the PM terms: half kick with the previous force,
 * drift, wrap (the reference's while loops), new force, half kick */
int orc_step(const orc_params *prm, int n, double *pos, double *vel, double *acc, double dkh, double dd)
{
    const double L = prm->boxsize;
    for (int i = 0; i < 3 * n; i++) vel[i] += acc[i] * dkh;
    for (int i = 0; i < 3 * n; i++) pos[i] += vel[i] * dd;
    for (int i = 0; i < 3 * n; i++) {
        while (pos[i] < 0.0) pos[i] += L;
        while (pos[i] >= L) pos[i] -= L;
    }
    int rc = orc_force(prm, n, pos, acc, NULL);
    if (rc) return rc;
    for (int i = 0; i < 3 * n; i++) vel[i] += acc[i] * dkh;
    return 0;
}
