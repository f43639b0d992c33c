/*
 * oracle/refcpu.h -- CPU restatement of photoNs-2.0's short-range FMM force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the MI355X path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Everything is FP64 and single threaded, except orc_direct() (pthreads).
 * Index conventions shared with the GPU library's exports:
 *   - leaves are numbered 0..nleaf-1 in creation (depth-first) order, which is
 *     also increasing particle order;
 *   - nodes are numbered 0..nnode-1 in pre-order, root = 0;
 *   - a "cell" id in interaction lists is the leaf id for a leaf and
 *     nleaf + node id for a node;
 *   - shift id 0 is the home box, ids 1..26 are the periodic images in the
 *     order of the reference loop (mi, mj, mk) in {-1,0,1}^3 minus (0,0,0),
 *     mi slowest (src/fmm.c:1029-1044).
 */
#ifndef PHOTONS_ORACLE_REFCPU_H
#define PHOTONS_ORACLE_REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Force parameters.  Layout is mirrored by ctypes in the tests. */
typedef struct {
    double boxsize;   /* BOXSIZE, periodic box is [0, boxsize)^3          */
    double mass;      /* MASSPART                                         */
    double soft;      /* SoftenScale                                      */
    double rs;        /* splitRadius                                      */
    double rc;        /* cutoffRadius                                     */
    double theta;     /* open_angle                                       */
    int32_t maxleaf;  /* MAXLEAF                                          */
    int32_t order;    /* Cartesian expansion order p (reference: 3)       */
    int32_t periodic; /* 1: add the 26 image walks (PERIODIC_CONDITION)   */
    int32_t pad;
} orc_params;

typedef struct {
    int64_t p2p_local_pairs;   /* idxP2P: local (target leaf, source leaf) tasks   */
    int64_t m2l_local_pairs;   /* idxM2L: local M2L tasks                          */
    int64_t p2p_ext_pairs;     /* image (target leaf, source leaf) pairs           */
    int64_t m2l_ext_pairs;     /* image M2L pairs                                  */
    int64_t m2l_calls;         /* walk_m2l_count: all m2l() calls                  */
    int64_t p2p_ext_inter;     /* p2p_count_remote: image particle interactions    */
    int64_t p2p_inter;         /* all particle-particle interactions (local+image) */
    int32_t nleaf, nnode, maxdepth, pad;
} orc_stats;

typedef struct orc_tree orc_tree;

/* Multipole bookkeeping: number of Cartesian terms up to order p. */
int orc_nterms(int order);

/* ---- tree (src/fmm.c:30-264) ------------------------------------------- */
orc_tree *orc_build(const orc_params *prm, int n, const double *pos);
void orc_free(orc_tree *t);
void orc_tree_sizes(const orc_tree *t, int32_t *nleaf, int32_t *nnode, int32_t *maxdepth);
/* perm[i] = original index of the i-th particle in tree order.
 * node_son: 2 per node, >=0 node id, <0 leaf id -(son+1).                 */
void orc_tree_export(const orc_tree *t, int32_t *perm,
                     int32_t *leaf_ipart, int32_t *leaf_npart,
                     double *leaf_center, double *leaf_width,
                     int32_t *node_son, int32_t *node_npart, double *node_split,
                     double *node_center, double *node_width);

/* ---- interaction lists (src/fmm.c:406-712, src/remotes.c:60-169,213-552) */
/* Returns the number of entries; arrays may be NULL to only count.
 * Lists are in the reference's emission order: local walk first, then the
 * 26 image walks in loop order.                                           */
int64_t orc_p2p_list(const orc_tree *t, int32_t *tgt, int32_t *src, int8_t *shift, int64_t cap);
int64_t orc_m2l_list(const orc_tree *t, int32_t *tgt, int32_t *src, int8_t *shift, int64_t cap);

/* nodes and bodies in the pruned image tree of shift sid (src/remotes.c:60-169) */
void orc_let_size(const orc_tree *t, int sid, int64_t *nnodes, int64_t *nbodies);

/* ---- the whole short-range force (fmm_prepare + fmm_task + fmm_ext) ---- */
/* acc is written (not accumulated) in the caller's particle order.        */
int orc_force(const orc_params *prm, int n, const double *pos, double *acc, orc_stats *st);

/* The same force for the particles of selected target leaves only (leaf ids of the
 * tree; it may hold a leaf more than once only if the caller wants it twice): the
 * whole tree and multipoles, the walks entered only for targets above a selected
 * leaf, P2P and L2P of the selected leaves.  acc_sel/id_sel: their particles leaf
 * after leaf in tree order (acceleration, caller index); p2p_len[k]: entries of leaf
 * sel[k]'s P2P list; m2l_len[k]: M2L entries of the leaf and its ancestors.
 * Returns the number of particles written (the caller sizes acc_sel for the sum of
 * the selected leaves' counts), < 0 on bad arguments.                      */
int64_t orc_force_leaves(const orc_params *prm, int n, const double *pos, int nsel, const int32_t *sel,
                         double *acc_sel, int32_t *id_sel, int64_t *p2p_len, int64_t *m2l_len);

/* ---- single operators (src/operator.c) --------------------------------- */
void orc_p2m(int order, int n, const double *pos, double mass, const double *center, double *M);
void orc_m2m(int order, const double *d, const double *M, double *tM);
void orc_m2l(int order, double rs, const double *d, const double *M, double *L);
void orc_l2l(int order, const double *d, const double *L, double *tL);
void orc_l2p(int order, const double *L, const double *center, int n, const double *pos, double *acc);
/* radial factors fac[n] = (r^-1 d/dr)^n erfc(r/2rs)/r, n = 0..nmax        */
void orc_radial(double r, double rs, int nmax, double *fac);
/* derivative tensor D_a = d^a erfc(|R|/2rs)/|R| for |a| <= order          */
void orc_dtensor(int order, double rs, const double *R, double *D);
/* acceptance() (src/fmm.c:267-326) */
int orc_acceptance(const double *wi, const double *wj, const double *dist, double theta, double rc);
/* one softened erfc-split pair force, src/fmm.c:834-852                  */
void orc_p2p_pair(const double *xt, const double *xs, double mass, double soft, double rs, double *acc);

/* ---- FP64 direct-sum short-range oracle --------------------------------- */
/* Sum over every source within rmax of each target under the minimum image
 * convention (boxsize > 0) using a cell list; rmax >= boxsize/2 is refused.
 * acc[3*k] for target k = targets[k].  nthreads pthreads.                 */
int orc_direct(const orc_params *prm, int n, const double *pos, int nt, const int32_t *targets,
               double rmax, double *acc, int nthreads);

/* ---- time integration (the reference's step, src/photoNs.c:136-268) ----- */
/* Kick and drift factors over [ln a_i, ln a_f] by 128-block Simpson in ln a,
 * H(a) = 0.1 sqrt(Om a^-3 + OL) (src/initial.c:650-683):
 *   kick  = int dln a / (a H),   drift = int dln a / (a^2 H).             */
double orc_kick_loga(double loga_i, double loga_f, double omega_m, double omega_x);
double orc_drift_loga(double loga_i, double loga_f, double omega_m, double omega_x);
/* One step of n particles of one periodic box, in the reference's order:
 * v += acc*dkh; x += v*dd; wrap into [0, L); acc = short-range force(x);
 * v += acc*dkh (dkh = kick/2 * G; the PM force is not part of this path).
 * acc holds the previous force on entry and the new one on return.        */
int orc_step(const orc_params *prm, int n, double *pos, double *vel, double *acc, double dkh, double dd);

#ifdef __cplusplus
}
#endif
#endif
