/*
 * photons_fmm.h -- C ABI of the MI355X short-range FMM force library
 * (libphotons_fmm.so).  Plain pointers and sizes only.
 *
 * Two layers:
 *  1. the native handle API (photons_*): a context owns device-resident particles,
 *     tree, expansions and interaction lists; one call runs the whole short-range
 *     force on the GPU;
 *  2. the drop-in layer: the reference's own entry points and globals
 *     (inc/fmm.h, inc/operator.h, inc/kernels.h, inc/photoNs.h), implemented on
 *     top of layer 1, so the reference driver's force call links unchanged.
 *
 * All functions of layer 1 return 0 on success and a negative code on failure;
 * photons_last_error() describes the last failure of the calling thread.
 * Nothing here falls back to the CPU: without a usable GPU every call fails.
 */
#ifndef PHOTONS_FMM_H
#define PHOTONS_FMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================= */
/* 1. native handle API                                                     */
/* ======================================================================= */

enum { PHOTONS_FP64 = 0, PHOTONS_FP32_MIXED = 1 };

/* Force parameters; the names are the reference globals they stand for
 * (inc/photoNs.h:20-45, 125-126; src/initial.c:316-345).                   */
typedef struct {
    double boxsize;   /* BOXSIZE: periodic box [0, boxsize)^3                 */
    double mass;      /* MASSPART                                             */
    double soft;      /* SoftenScale                                          */
    double rs;        /* splitRadius                                          */
    double rc;        /* cutoffRadius                                         */
    double theta;     /* open_angle                                           */
    int32_t maxleaf;  /* MAXLEAF, MaxPackage of the parameter file (2 .. 256)   */
    int32_t order;    /* expansion order p: 3 (reference), 4 or 5             */
    int32_t periodic; /* 1: the 26 image walks of fmm_ext (PERIODIC_CONDITION)*/
    int32_t precision;/* PHOTONS_FP64 or PHOTONS_FP32_MIXED                   */
} photons_params;

typedef struct {
    int64_t n;                 /* particles                                         */
    int32_t nleaf, nnode;      /* cells                                             */
    int32_t depth;             /* levels below the root                             */
    int32_t pad;
    int64_t p2p_pairs;         /* (target leaf, source leaf, shift) entries         */
    int64_t p2p_local_pairs;   /* idxP2P: those with shift 0                        */
    int64_t m2l_pairs;         /* (target cell, source cell, shift) entries         */
    int64_t m2l_local_pairs;   /* idxM2L                                            */
    int64_t p2p_interactions;  /* particle pairs i != j evaluated                   */
    int64_t p2p_image_interactions; /* p2p_count_remote                             */
} photons_stats;

/* Phases of one force evaluation, milliseconds of GPU time of the last call
 * (filled when timing is enabled with photons_set_timing).                */
enum {
    PHOTONS_PH_UPLOAD = 0, PHOTONS_PH_TREE, PHOTONS_PH_UPWARD, PHOTONS_PH_WALK,
    PHOTONS_PH_M2L, PHOTONS_PH_L2L, PHOTONS_PH_P2P, PHOTONS_PH_DOWNLOAD, PHOTONS_NPHASE
};

typedef struct photons_ctx photons_ctx;

const char *photons_last_error(void);
int photons_version(void);

/* stream: a hipStream_t to run on, or NULL for a private stream. */
int photons_create(photons_ctx **ctx, const photons_params *prm, int device, void *stream);
int photons_destroy(photons_ctx *ctx);
int photons_set_params(photons_ctx *ctx, const photons_params *prm);
int photons_get_params(const photons_ctx *ctx, photons_params *prm);
int photons_set_timing(photons_ctx *ctx, int enabled);
/* Plan reuse (no counterpart in the reference; on by default): the first evaluation of a
 * box set reads counts back to size its buffers and plans the next ones, which then run
 * without host round trips and wait once at the end for the overflow check (particles
 * that outgrew the plan are evaluated again the first way).  Same result bit for bit;
 * off, every evaluation sizes its buffers as it goes.                            */
int photons_set_plan_reuse(photons_ctx *ctx, int enabled);
int photons_get_phase_ms(const photons_ctx *ctx, double *ms /* [PHOTONS_NPHASE] */);
int photons_get_stats(const photons_ctx *ctx, photons_stats *st);

/* One short-range force evaluation (tree build through L2P) of n particles.
 * pos and acc are [n][3] doubles in the caller's order; acc is overwritten with
 * the short-range acceleration without G (the reference's part[].acc, which the
 * caller zeroes before the force call: src/photoNs.c:222-226).
 *   _host:   host arrays (copies in and out over PCIe),
 *   _device: device arrays on the context's device, work queued on its stream;
 *            returns after the work is queued, photons_sync() waits.        */
int photons_force_host(photons_ctx *ctx, int64_t n, const double *pos, double *acc);
int photons_force_device(photons_ctx *ctx, int64_t n, const double *d_pos, double *d_acc);
int photons_sync(photons_ctx *ctx);

/* Many independent periodic boxes in one evaluation (no counterpart in the
 * reference, whose force call covers one domain: src/fmm.c:909-1075 per rank).
 * Box b holds counts[b] >= 1 particles (more than maxleaf when periodic) in
 * [0, lengths[b])^3; the boxes' particles are concatenated in pos/acc, box 0
 * first.  All other parameters (mass, soft, rs, rc, theta, order, leaf size,
 * precision) are the context's and shared by the boxes.  Each box's result is
 * that of the box alone: trees, lists and shifts never cross boxes.  One tree
 * build, walk and P2P/M2L launch covers the whole batch.                      */
int photons_force_boxes_host(photons_ctx *ctx, int nbox, const int64_t *counts, const double *lengths,
                             const double *pos, double *acc);
int photons_force_boxes_device(photons_ctx *ctx, int nbox, const int64_t *counts, const double *lengths,
                               const double *d_pos, double *d_acc);
/* photons_force_device without any allocation, for callers that cannot allocate
 * (tight serving loops, code that must not call hipMalloc).  No counterpart in the
 * reference.  One box of n <= 32768 particles (else an error): every buffer of the
 * evaluation comes from a static 96 MiB device array that the library holds on each
 * GPU it runs on until the process ends; the context's own buffers and its last
 * ordinary evaluation are left as they were.  Returns when d_acc is written.  All
 * calls on one GPU share the array: calls that overlap in time on one GPU (from other
 * threads) race.                                                              */
int photons_force_device_static(photons_ctx *ctx, int64_t n, const double *d_pos, double *d_acc);

/* Tuning (no counterpart in the reference, which takes MAXLEAF from its parameter
 * file, src/initial.c:172-173, untuned).  The FP32-mixed P2P kernel stages 256 or
 * 512 source particles per batch in LDS (default 256).  photons_autotune_boxes times leaf
 * sizes 16/32/64 x staging 256/512 (staging only for FP32-mixed) on the
 * caller's batch, keeps the fastest in the context and leaves its result in
 * d_acc; ms[6] (or NULL) gets the mean ms per evaluation of each candidate,
 * leaf-major, -1 where a candidate was skipped (boxes of <= leaf particles). */
int photons_set_p2p_stage(photons_ctx *ctx, int stage);
int photons_get_p2p_stage(const photons_ctx *ctx);
int photons_autotune_boxes(photons_ctx *ctx, int nbox, const int64_t *counts, const double *lengths,
                           const double *d_pos, double *d_acc, int reps, int *best_leaf, int *best_stage,
                           double *ms);

/* A planned evaluation of a fixed set of boxes, captured in a hipGraph.  No
 * counterpart in the reference, whose task pipeline never waits on a count
 * (src/fmm.c:381-404, 909-984): here, too, one evaluation runs from the upload to
 * P2P + L2P without a host round trip, an allocation or a memset, so that it can be
 * captured and replayed.  photons_graph_create measures one ordinary evaluation of
 * the boxes at d_pos, allocates every buffer `margin` (e.g. 0.25) above what it
 * measured, runs the planned evaluation once and, with capture != 0, captures it into
 * a hipGraph on the context's stream; d_acc then holds the force of d_pos.
 * photons_graph_launch queues one replay (one graph launch; without capture the same
 * launches one by one): it reads d_pos and writes d_acc (caller order), whose contents
 * may change between launches (a drift), not their addresses or the box set.
 * photons_graph_check waits and reports the overflow bits of the last launch: 0, or
 * the particles have moved so far that the tree or the lists outgrew the plan, the
 * result is incomplete and the graph must be created again; ms gets its GPU time when
 * timing is on (else -1).  While a planned evaluation is held, the context refuses
 * other evaluations; photons_graph_destroy releases it.                          */
int photons_graph_create(photons_ctx *ctx, int nbox, const int64_t *counts, const double *lengths,
                         const double *d_pos, double *d_acc, double margin, int capture);
int photons_graph_launch(photons_ctx *ctx);
int photons_graph_check(photons_ctx *ctx, int *overflow, double *ms);
int photons_graph_destroy(photons_ctx *ctx);

/* The same force call in two steps, as the reference splits it:
 *   photons_prepare_host   upload, tree, P2M and M2M   (fmm_prepare)
 *   photons_evaluate_host  walk, M2L, L2L, P2P and L2P (fmm_task + fmm_ext);
 *                          acc is written in tree order (photons_tree_perm). */
int photons_prepare_host(photons_ctx *ctx, int64_t n, const double *pos);
int photons_evaluate_host(photons_ctx *ctx, double *acc_tree_order);

/* Inspection of the last force call (tree order; see photons_tree_perm).
 * Leaves are numbered in depth-first order, nodes in pre-order; a cell id is
 * the leaf id for a leaf and nleaf + node id for a node; list entries are
 * (target cell, source cell, shift id), shift id 0 = home box and 1..26 the
 * images in the order of src/fmm.c:1029-1044.                               */
int photons_tree_perm(const photons_ctx *ctx, int32_t *perm /* [n]: caller index of slot i */);
int photons_tree_cells(const photons_ctx *ctx, int32_t *npart, int32_t *ipart, int32_t *son /* [ncell][2] */,
                       double *center /* [ncell][3] */, double *width /* [ncell][3] */);
int photons_p2p_list(const photons_ctx *ctx, int64_t cap, int32_t *tgt, int32_t *src, int8_t *shift, int64_t *count);
int photons_m2l_list(const photons_ctx *ctx, int64_t cap, int32_t *tgt, int32_t *src, int8_t *shift, int64_t *count);
/* entries per target: len[t] for every target leaf (which = 0, P2P) or cell
 * (which = 1, M2L) of the last force call                                    */
int photons_list_lengths(const photons_ctx *ctx, int which, int64_t *len);
/* expansions of every cell, [ncell][nterms(p)] doubles (FP32-mixed values are
 * returned converted to kpc/h units)                                         */
int photons_expansions(const photons_ctx *ctx, double *M, double *L);

/* Batched operators on device arrays (order p, FP64), one item per thread.
 * Layouts: positions [n][3]; expansions [count][nterms(p)]; offsets [count][3].
 *   p2m:  M[k] = P2M of particles first[k] .. first[k]+npart[k]-1 about center[k]
 *   m2m:  tM[k] += shift(M[k]) by d[k] = parent - child
 *   m2l:  L[k] += M2L(M[k]) at R[k] = c_target - c_source
 *   l2l:  tL[k] += shift(L[k]) by d[k] = child - parent
 *   l2p:  acc[i] += grad of L[owner[i]] at pos[i] - center[owner[i]]        */
int photons_p2m_batch(photons_ctx *ctx, int order, int count, const double *d_pos, const int32_t *d_first,
                      const int32_t *d_npart, const double *d_center, double mass, double *d_M);
int photons_m2m_batch(photons_ctx *ctx, int order, int count, const double *d_d, const double *d_M, double *d_tM);
int photons_m2l_batch(photons_ctx *ctx, int order, int count, double rs, const double *d_R, const double *d_M,
                      double *d_L);
int photons_l2l_batch(photons_ctx *ctx, int order, int count, const double *d_d, const double *d_L, double *d_tL);
int photons_l2p_batch(photons_ctx *ctx, int order, int n, const double *d_pos, const int32_t *d_owner,
                      const double *d_center, const double *d_L, double *d_acc);

/* ---- serving: resident boxes and continuous batching of step requests ----
 * No counterpart in the reference (its driver steps one domain in a loop,
 * src/photoNs.c:136-313).  A step request advances one box by the reference's
 * loop body (src/photoNs.c:136-268), in FP64 on the GPU:
 *   v += acc_pm*dkh + acc*dkh; x += v*dd; wrap into [0, L);
 *   acc = force(x); acc_pm = pm(x); v += acc*dkh + acc_pm*dkh
 * with dkh = G*kick/2 and dd = drift (photons_kick_factor/_drift_factor over
 * the step's ln a interval, src/initial.c:650-683); the acc_pm terms only for a
 * box added with a mesh (photons_server_add_box_pm: nside^3 cells, the PM of
 * photons_pm_* with the server's rs and mass), the others take the short-range
 * force alone.  Boxes stay in HBM between steps and are admitted while
 * pos/vel/acc (72 B per particle, 96 B with acc_pm) fit the HBM budget minus
 * what the lanes reserve for evaluation buffers (about 1 KB per particle of
 * max_batch_particles, per lane; a lane's PM meshes come on top).  Each of nlanes lanes (own
 * context and stream) takes, whenever it is free, every queued request whose
 * box is idle (arrival order, one request per box at a time, up to
 * max_batch_particles) and evaluates them as one batch of boxes; the lanes'
 * work overlaps on the device.  Thread safe; requests of one box run in order. */
typedef struct photons_server photons_server;
typedef struct {
    int64_t requests_done, particles_stepped, batches, queued;
    double mean_boxes_per_batch, span_s, resident_bytes, reserved_bytes;
    int32_t nboxes, max_lanes_busy;  /* lanes that evaluated batches at the same time, at most */
} photons_server_stats;

const char *photons_server_last_error(void);
int photons_server_create(photons_server **srv, const photons_params *prm, int device, int nlanes,
                          double hbm_budget_bytes, int64_t max_batch_particles);
int photons_server_destroy(photons_server *srv);
/* pos, vel [n][3] in [0, boxsize); acc the current force or NULL (then the box's
 * first request computes it, as the reference does before its first step)  */
int photons_server_add_box(photons_server *srv, int64_t n, double boxsize, const double *pos, const double *vel,
                           const double *acc, int *box_id);
/* the same with a PM mesh of nside^3 cells (even, in [8, 1624]); acc and acc_pm
 * are both given or both NULL (then the first request computes both)          */
int photons_server_add_box_pm(photons_server *srv, int64_t n, double boxsize, int nside, const double *pos,
                              const double *vel, const double *acc, const double *acc_pm, int *box_id);
int photons_server_remove_box(photons_server *srv, int box_id);
int photons_server_submit(photons_server *srv, int box_id, double dkh, double dd, int64_t *request_id);
/* waits for a request and reports it (latency: submit to completion, host clock);
 * each request is waited on once, the report releases it                     */
int photons_server_wait(photons_server *srv, int64_t request_id, double *latency_ms);
int photons_server_drain(photons_server *srv);
/* paused != 0: the lanes finish their batches but take no new one until resumed
 * (requests submitted meanwhile are batched together on resume); drain waits
 * forever on a paused server with queued requests                             */
int photons_server_pause(photons_server *srv, int paused);
int photons_server_read_box(photons_server *srv, int box_id, double *pos, double *vel, double *acc);
int photons_server_read_pm(photons_server *srv, int box_id, double *acc_pm);  /* a box with a mesh */
int photons_server_get_stats(photons_server *srv, photons_server_stats *st);
double photons_kick_factor(double loga_i, double loga_f, double omega_m, double omega_x);
double photons_drift_factor(double loga_i, double loga_f, double omega_m, double omega_x);

/* ---- PM long-range force (one rank's whole periodic box, FP64) -------------
 * The reference's partmesh() and its convolution on one rank (src/partmesh.c:
 * 800-1550, src/conv.f90:171-229): CIC deposit with mesh points at cell centres,
 * FFT, Green's function pref exp(-k^2 (2 pi rs/L)^2) sinc^-4 / k^2 with
 * pref = L^2/(pi N^3), gf(0) = pref and pi = 3.1415926 as the reference
 * truncates it, inverse FFT, 4th-order central-difference gradient, CIC
 * interpolation.  acc_pm (no G) in the caller's order.  No counterpart of the
 * reference's multi-rank slab exchange: the whole mesh lives on the GPU.      */
typedef struct photons_pm photons_pm;
const char *photons_pm_last_error(void);
/* nside even, in [8, 1624]: the mesh indices are 32-bit (N^3 < 2^32)          */
int photons_pm_create(photons_pm **pm, int nside, double boxsize, double rs, double mass, int device, void *stream);
int photons_pm_destroy(photons_pm *pm);
/* the box length of the next evaluations, the mesh shape unchanged (one context for
 * boxes of one mesh size; no counterpart in the reference, one box per run)    */
int photons_pm_set_boxsize(photons_pm *pm, double boxsize);
int photons_pm_force_device(photons_pm *pm, int64_t n, const double *d_pos, double *d_acc);
int photons_pm_force_host(photons_pm *pm, int64_t n, const double *pos, double *acc);
int photons_pm_sync(photons_pm *pm);

/* ---- one TreePM step on device-resident particles --------------------------
 * Replaces one pass of the reference's time loop (src/photoNs.c:136-254) with
 * the particles kept in HBM: v += acc_pm dkh; v += acc dkh; x += v dd; wrap into
 * [0, boxsize); acc = short-range force (fmm_construct .. fmm_ext, the tree
 * rebuilt as the reference does every step); acc_pm = PM force (partmesh, run on
 * the PM handle's stream alongside); v += acc dkh; v += acc_pm dkh.
 * dkh = G kick/2 and dd = drift over the step's ln a interval
 * (photons_kick_factor / photons_drift_factor).  pm NULL: no long-range part
 * (d_acc_pm is then unused).  The PM handle must be for the context's boxsize.
 * Before the first step acc (and acc_pm) hold the forces of the initial
 * positions, as the reference computes them before its loop
 * (src/photoNs.c:97-116).  Returns with the last kick queued on the context's
 * stream (photons_sync waits).  All arrays live on the context's device.  A
 * failed step leaves the particles part-way through it (the first kicks and the
 * drift may have been applied).                                              */
int photons_step_device(photons_ctx *ctx, photons_pm *pm, int64_t n, double *d_pos, double *d_vel, double *d_acc,
                        double *d_acc_pm, double dkh, double dd);

/* ---- Gadget-2 snapshot/IC files (host only) -------------------------------
 * The reference's format and conventions (src/snapshot.c:5-22, 64-120, 211-293,
 * 397-503): 256-byte header, float32 positions, float32 velocities stored as
 * v / a^1.5 (a = 1/(1+z)); the reader returns v.  The writer writes one file of
 * type-1 particles with the record lengths set (the reference leaves them unset;
 * neither reader looks at them).                                              */
typedef struct {
    int64_t n;              /* particles in this file (sum of npart) */
    int32_t npart[6], npart_total[6], num_files, pad;
    double mass[6], time, redshift, boxsize, omega_m, omega_l, hubble;
} photons_snapshot_info;

const char *photons_snapshot_last_error(void);
int photons_snapshot_header(const char *path, photons_snapshot_info *info);
int photons_snapshot_read(const char *path, int64_t n_start, int64_t n_count, double *pos, double *vel);
int photons_snapshot_write(const char *path, const photons_snapshot_info *info, int64_t n, int64_t npart_total_all,
                           const double *pos, const double *vel);

/* ======================================================================= */
/* 2. drop-in layer: the reference's entry points                            */
/* ======================================================================= */
/* Globals read and written by the force call have the reference's names and
 * types (inc/photoNs.h, inc/typesdef.h); a caller built against the reference
 * headers shares them.  Only p = 3 (NMULTI 20) layouts are exported here.
 * The layer keeps one GPU context (device 0, FP64) behind the globals;
 * photons_compat_set_precision() switches it to FP32-mixed.               */

typedef struct {
    double pos[3];
    double acc[3];
    double vel[3];
    double acc_pm[3];
} photons_body; /* Body, inc/typesdef.h:48-57 */

int photons_compat_set_precision(int precision);
photons_ctx *photons_compat_context(void);
/* The drop-in entry points keep the reference's void signatures and its failure
 * behaviour (a message, then exit: src/fmm.c:411-414, src/utility.c:78-81); the
 * message of the failure stays readable here, e.g. from an atexit handler.
 * Empty while nothing has failed.                                              */
const char *photons_gpu_last_error(void);

/* phase timers of inc/photoNs.h:95-111 [s], written by the calls below as the
 * reference writes them: dtime_prep (fmm_prepare), dtime_task (fmm_task),
 * dtime_ext (fmm_ext), dtime_fmm (fmm_prepare + L2L), dtime_p2p (P2P + L2P kernel
 * time; reset by fmm_prepare, src/fmm.c:182), dtime_m2l (M2L kernel time, summed
 * over calls, src/fmm.c:906), dtime_fmm_remote (0: the image walks are fused with
 * the home walk), dtime_pm (partmesh).  Kernel times are GPU times (HIP events). */
extern double dtime_prep, dtime_task, dtime_ext, dtime_fmm, dtime_p2p, dtime_m2l, dtime_fmm_remote, dtime_pm;

/* the force call, inc/fmm.h:11-22 and src/fmm.c:180-264, 329-360, 719-1087:
 *   fmm_construct   domain box [0, BOXSIZE)^3 (one rank)
 *   fmm_prepare     upload part[].pos, build the tree, P2M and M2M on the GPU;
 *                   part[] is permuted into tree order like the reference
 *   fmm_task        dual-tree walk, P2P and M2L on the GPU (home box)
 *   fmm_ext         the 26 periodic images, L2L and L2P; part[].acc += result
 *   fmm_deconstruct release the host mirrors (device buffers stay for reuse) */
void fmm_construct(void);
void fmm_prepare(void);
void fmm_task(void);
void fmm_ext(void);
void fmm_deconstruct(void);
void build_localtree(void);

/* operators, inc/operator.h:6-19 and inc/photoNs.h:325-330 (p = 3) */
void p2m(int iPart, int npart, double center[3], double M[]);
void m2m(double dx, double dy, double dz, double M[], double tM[]);
void m2l(double dx, double dy, double dz, double M[], double toL[]);
void l2l(double dx, double dy, double dz, double L[], double toL[]);
void l2p(int ileaf);
void walk_m2m(int iNode);
void walk_l2l(int inode);
int acceptance(double wi[3], double wj[3], double dist[3]);

/* pair kernels, inc/kernels.h:4-8 (declared, never defined by the reference):
 * part[i].acc += P2P from the particles of leaf jnode onto those of leaf inode
 *   p2p_kernel           FP64
 *   p2p_kernel_inleaf    FP64, skipping i == j (inode == jnode)
 *   p2p_kernel_rinvsqrt  FP32-mixed (rsqrt pair kernel)                       */
void p2p_kernel(int inode, int jnode);
void p2p_kernel_inleaf(int inode, int jnode);
void p2p_kernel_rinvsqrt(int inode, int jnode);

/* dual-tree walks, src/fmm.c:406,569 and src/remotes.c:213,405: from the pair
 * (im, jm) (reference index spaces), evaluate the P2P (resp. M2L) work the
 * reference walk would emit: P2P into part[].acc, M2L into the cell L.  The
 * _ext forms use the image set by the last fmm_remote() call.              */
void walk_task_p2p(int im, int jm);
void walk_task_m2l(int im, int jm);
void walk_task_p2p_ext(int im, int jm);
void walk_task_m2l_ext(int im, int jm);
/* image exchange, src/remotes.c:684-751 (one rank: idx = 0 only) */
void fmm_remote(int idx, double shift[3]);

/* PM, src/partmesh.c:11-16, 800-1550 (inc/partmesh.h:16-22), on the GPU through
 * photons_pm_*: part[].acc_pm = long-range force of part[] (no G) on an
 * NSIDE^3 mesh with splitRadius, BOXSIZE and MASSPART.  pm_thread may run
 * beside fmm_task, as the reference's PMTHREAD build does.  A program that
 * still links the reference's partmesh.c keeps its own definitions.         */
void partmesh(void);
void partmesh_thread(void);
void *pm_thread(void *arg);

#ifdef __cplusplus
}
#endif
#endif /* PHOTONS_FMM_H */
