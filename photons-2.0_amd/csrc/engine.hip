// engine.hip -- context lifetime, force pipeline and the native C API.
//
// One force evaluation (photons_force_device), all on the context's stream:
//   upload    AoS positions -> SoA tree buffers
//   tree      kd build (tree.hip)
//   upward    P2M + M2M (kernels.hip)
//   walk      per-target dual-tree walk -> P2P and M2L lists (walk.hip)
//   m2l       M2L over the list, one thread per target cell
//   l2l       L2L level by level
//   p2p       P2P + L2P, one wavefront per target leaf, result in caller order
// The tree build and the walk read a few counts back to size the next step.
#include <algorithm>
#include <cstddef>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "engine.h"
#include "fmm_math.h"
#include "kick.h"

namespace photons {
void p2p_l2p_launch(Ctx &c, bool with_l2p, bool accumulate, bool caller_order, double *out);
void run_walk_from(Ctx &c, int shift_lo, int shift_hi, int start_target, int start_source);
void op_p2m(hipStream_t st, int order, int count, const double *pos, const int32_t *first, const int32_t *npart,
            const double *center, double mass, double *M);
void op_shift(hipStream_t st, int order, int kind, int count, const double *d, const double *X, double *Y);
void op_m2l(hipStream_t st, int order, int count, double rs, const double *R, const double *M, double *L);
void op_l2p(hipStream_t st, int order, int n, const double *pos, const int32_t *owner, const double *center,
            const double *L, double *acc);

// ---- the static scratch of photons_force_device_static: a file-scope device array of
// kScratchBytes on every GPU the library runs on, held until the process ends.  Its
// address on a device comes from hipGetSymbolAddress (per device, at context creation);
// the array's name is not a device address in host code.
constexpr size_t kScratchBytes = (size_t)96 << 20;
constexpr int64_t kStaticMaxParticles = 32768;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static char *g_scratch_ptr[kMaxDevices];

static char *scratch_on(int device)
{
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device id out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        PH_CHECK(hipGetDevice(&cur));
        PH_CHECK(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        PH_CHECK(hipSetDevice(cur));
        PH_CHECK(e);
        g_scratch_ptr[device] = static_cast<char *>(p);
    }
    return g_scratch_ptr[device];
}

// bump allocation over the scratch during one static evaluation (this thread only)
struct Arena {
    char *base;
    size_t cap, used;
};
static thread_local Arena *g_arena = nullptr;

bool arena_active() { return g_arena != nullptr; }

void *dev_alloc(size_t bytes)
{
    if (g_arena) {
        const size_t off = (g_arena->used + 255) & ~(size_t)255;
        if (off + bytes > g_arena->cap)
            throw std::runtime_error("the evaluation needs more than the " + std::to_string(kScratchBytes >> 20) +
                                     " MiB static scratch");
        g_arena->used = off + bytes;
        return g_arena->base + off;
    }
    void *p = nullptr;
    PH_CHECK(hipMalloc(&p, bytes));
    return p;
}

void dev_free(void *p)
{
    if (p && !g_arena) PH_CHECK(hipFree(p));
}

void *grow(DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 1;
    if (bytes <= b.cap) return b.p;
    dev_free(b.p);
    b.p = nullptr;
    b.cap = 0;
    const size_t want = arena_active() ? bytes : bytes + bytes / 8;
    b.p = dev_alloc(want);
    b.cap = want;
    return b.p;
}

void need(const DevBuf &b, size_t bytes, const char *what)
{
    if (!b.p || b.cap < bytes) throw std::runtime_error(std::string(what) + " is not available for this tree");
}

void *ctx_buf(Ctx &c, DevBuf &b, size_t bytes, const char *what)
{
    if (c.plan.active && !c.plan.alloc) {
        need(b, bytes == 0 ? 1 : bytes, what);
        return b.p;
    }
    return grow(b, bytes);
}

// zeros in place of hipMemsetAsync (a captured evaluation records no memset)
__global__ void k_zero_words(uint32_t *__restrict__ p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// the buffers of one force evaluation (what photons_force_device_static takes from the
// static scratch instead)
static std::vector<DevBuf *> eval_bufs(Ctx &c)
{
    return {&c.px, &c.py, &c.pz, &c.pidx, &c.qx, &c.qy, &c.qz, &c.qidx, &c.ox, &c.oy, &c.oz, &c.geom, &c.info,
            &c.level_nodes, &c.M, &c.L, &c.p2p_off, &c.p2p_src, &c.m2l_off, &c.m2l_src, &c.cnt_tmp, &c.scan_tmp,
            &c.t_start, &c.t_count, &c.t_split, &c.t_lo, &c.t_hi, &c.t_son, &c.t_depth, &c.t_parent, &c.l_start,
            &c.l_count, &c.l_lo, &c.l_hi, &c.l_parent, &c.seg_chunk, &c.chunk_part, &c.chunk_rcnt, &c.seg_nleft,
            &c.seg_flag, &c.seg_nn, &c.seg_non, &c.t_size, &c.t_lsize, &c.t_dfs, &c.t_lfirst, &c.l_final,
            &c.walk_rec, &c.walk_trec, &c.dl_a, &c.dl_b, &c.pair_cnt, &c.big_pairs, &c.walk_tlist, &c.walk_tcount, &c.walk_qoff, &c.walk_ccount, &c.walk_cidx, &c.m2l_item_off, &c.m2l_item_tgt, &c.m2l_part, &c.d_box_off, &c.d_box_len,
            &c.d_box_root, &c.cell_box, &c.p2p_order, &c.p2p_order_tmp, &c.p2p_part};
}

static void release(DevBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

void set_boxes(Ctx &c, int nbox, const int64_t *counts, const double *lengths)
{
    c.nbox = nbox;
    c.box_off.assign(nbox + 1, 0);
    c.box_len.assign(lengths, lengths + nbox);
    for (int b = 0; b < nbox; b++) c.box_off[b + 1] = c.box_off[b] + counts[b];
    void *o = grow(c.d_box_off, sizeof(int64_t) * (size_t)(nbox + 1));
    void *l = grow(c.d_box_len, sizeof(double) * (size_t)nbox);
    PH_CHECK(hipMemcpyAsync(o, c.box_off.data(), sizeof(int64_t) * (nbox + 1), hipMemcpyHostToDevice, c.stream));
    PH_CHECK(hipMemcpyAsync(l, c.box_len.data(), sizeof(double) * nbox, hipMemcpyHostToDevice, c.stream));
}

void phase_mark(Ctx &c)
{
    if (!c.timing.enabled || c.timing.nev >= 16) return;
    PH_CHECK(hipEventRecord(c.timing.ev[c.timing.nev++], c.stream));
}

}  // namespace photons

using namespace photons;

struct photons_ctx {
    Ctx c;
    DevBuf io_pos, io_acc, op_a, op_b, op_c, op_d, op_e;
};

static thread_local std::string g_last_error;

static int fail(const char *where, const std::exception &e)
{
    g_last_error = std::string(where) + ": " + e.what();
    return -1;
}

static int fail_msg(const std::string &m)
{
    g_last_error = m;
    return -2;
}

static std::string check_params(const photons_params *p)
{
    if (!p) return "null parameters";
    if (!(p->boxsize > 0.0)) return "boxsize must be positive";
    if (!(p->rs > 0.0) || !(p->rc > 0.0)) return "rs and rc must be positive";
    if (!(p->soft >= 0.0)) return "soft must be non-negative";
    if (!(p->theta > 0.0)) return "theta must be positive";
    // a source leaf must fit one staging batch of the P2P kernels (256 particles); target
    // leaves of more than 64 run in chunks of 64
    if (p->maxleaf < 2 || p->maxleaf > 256) return "maxleaf must be in [2, 256]";
    if (p->order < 3 || p->order > 5) return "order must be 3, 4 or 5";
    if (p->precision != PHOTONS_FP64 && p->precision != PHOTONS_FP32_MIXED) return "unknown precision";
    return "";
}

extern "C" {

const char *photons_last_error(void) { return g_last_error.c_str(); }
int photons_version(void) { return 100; }

int photons_create(photons_ctx **out, const photons_params *prm, int device, void *stream)
{
    try {
        std::string m = check_params(prm);
        if (!m.empty()) return fail_msg(m);
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
            return fail_msg("no usable GPU for this context");
        PH_CHECK(hipSetDevice(device));
        photons_ctx *x = new photons_ctx();
        x->c.prm = *prm;
        x->c.device = device;
        try {
            if (stream) {
                x->c.stream = reinterpret_cast<hipStream_t>(stream);
            } else {
                PH_CHECK(hipStreamCreateWithFlags(&x->c.stream, hipStreamNonBlocking));
                x->c.own_stream = true;
            }
            PH_CHECK(hipHostMalloc(&x->c.pinned.p, 4096, hipHostMallocDefault));
            x->c.pinned.cap = 4096;
            grow(x->c.err, sizeof(int));
            grow(x->c.sizes, sizeof(DevSizes));
            scratch_on(device);  // the static scratch's address on this device, once
            for (int i = 0; i < 16; i++) PH_CHECK(hipEventCreate(&x->c.timing.ev[i]));
            PH_CHECK(hipEventCreateWithFlags(&x->c.counts_ready, hipEventDisableTiming));
        } catch (...) {
            photons_destroy(x);  // releases whatever was created; null members are skipped
            throw;
        }
        *out = x;
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_create", e);
    }
}

int photons_destroy(photons_ctx *x)
{
    if (!x) return 0;
    try {
        Ctx &c = x->c;
        PH_CHECK(hipSetDevice(c.device));
        PH_CHECK(hipStreamSynchronize(c.stream));
        if (c.graph_exec) (void)hipGraphExecDestroy(c.graph_exec);
        if (c.graph) (void)hipGraphDestroy(c.graph);
        for (DevBuf *b : eval_bufs(c)) release(*b);
        DevBuf *bufs[] = {&c.sizes, &c.err, &c.stats, &x->io_pos, &x->io_acc, &x->op_a, &x->op_b, &x->op_c,
                          &x->op_d, &x->op_e};
        for (DevBuf *b : bufs) release(*b);
        if (c.pinned.p) (void)hipHostFree(c.pinned.p);
        for (int i = 0; i < 16; i++) (void)hipEventDestroy(c.timing.ev[i]);
        if (c.counts_ready) (void)hipEventDestroy(c.counts_ready);
        if (c.own_stream) (void)hipStreamDestroy(c.stream);
        delete x;
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_destroy", e);
    }
}

int photons_set_params(photons_ctx *x, const photons_params *prm)
{
    std::string m = check_params(prm);
    if (!m.empty()) return fail_msg(m);
    x->c.prm = *prm;
    return 0;
}

int photons_get_params(const photons_ctx *x, photons_params *prm)
{
    *prm = x->c.prm;
    return 0;
}

int photons_set_timing(photons_ctx *x, int enabled)
{
    x->c.timing.enabled = enabled != 0;
    return 0;
}

int photons_set_plan_reuse(photons_ctx *x, int enabled)
{
    x->c.reuse.on = enabled != 0;
    if (!x->c.reuse.on) x->c.reuse.have = false;
    return 0;
}

int photons_get_phase_ms(const photons_ctx *x, double *ms)
{
    for (int i = 0; i < PHOTONS_NPHASE; i++) ms[i] = x->c.phase_ms[i];
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------- pipeline

namespace photons {

// phases up to and including the multipoles (fmm_prepare)
static void prepare(Ctx &c, const double *d_pos)
{
    c.timing.nev = 0;
    phase_mark(c);
    upload_positions(c, d_pos);
    phase_mark(c);
    c.lists_valid = false;
    build_tree(c);
    if (c.prm.precision == PHOTONS_FP32_MIXED) compute_offsets_f32(c);
    phase_mark(c);
    upward_pass(c);
    // local expansions start at zero for this tree (the drop-in walks accumulate into them)
    const size_t lwords = (c.prm.precision == PHOTONS_FP32_MIXED ? 1 : 2) * (size_t)ncell_bound(c) *
                          nterms_of(c.prm.order);
    uint32_t *L = static_cast<uint32_t *>(ctx_buf(c, c.L, 4 * lwords, "local expansions"));
    k_zero_words<<<(unsigned)((lwords + 255) / 256), 256, 0, c.stream>>>(L, lwords);
    phase_mark(c);
}

// walk, M2L, L2L, P2P + L2P (fmm_task + fmm_ext)
static void evaluate(Ctx &c, double *d_out, bool caller_order)
{
    run_walk(c, 0, c.prm.periodic ? 27 : 1);
    phase_mark(c);
    m2l_pass(c, false);
    phase_mark(c);
    downward_pass(c);
    phase_mark(c);
    p2p_l2p_launch(c, true, false, caller_order, d_out);
    phase_mark(c);
}

// Phase times from the marks, interval i giving phase first + i (marks: start, after
// upload, after tree, after upward, after walk, after m2l, after l2l, after p2p).  The
// two-step form (prepare, then evaluate) times its second call from a mark of its own
// (first = PHOTONS_PH_WALK), keeping the first call's phases.
static void collect_timing(Ctx &c, int first = PHOTONS_PH_UPLOAD)
{
    for (int i = first; i < PHOTONS_NPHASE; i++) c.phase_ms[i] = 0.0;
    if (!c.timing.enabled || c.timing.nev == 0) return;
    PH_CHECK(hipEventSynchronize(c.timing.ev[c.timing.nev - 1]));
    const int ph[] = {PHOTONS_PH_UPLOAD, PHOTONS_PH_TREE, PHOTONS_PH_UPWARD, PHOTONS_PH_WALK,
                      PHOTONS_PH_M2L, PHOTONS_PH_L2L, PHOTONS_PH_P2P};
    for (int i = 0; i + 1 < c.timing.nev && first + i <= PHOTONS_PH_P2P; i++) {
        float ms = 0.f;
        PH_CHECK(hipEventElapsedTime(&ms, c.timing.ev[i], c.timing.ev[i + 1]));
        c.phase_ms[ph[first + i]] = ms;
    }
}

static std::string check_n(const Ctx &c, int64_t n)
{
    if (n < 0) return "negative particle count";
    if (n > (int64_t)0x7fffffff / 2) return "too many particles for one context";
    if (n > 0 && c.prm.periodic && n <= c.prm.maxleaf)
        return "periodic force needs more than maxleaf particles (the image tree root would read as a leaf)";
    return "";
}

static std::string check_boxes(const Ctx &c, int nbox, const int64_t *counts, const double *lengths)
{
    if (nbox < 1 || !counts || !lengths) return "need at least one box";
    int64_t total = 0;
    for (int b = 0; b < nbox; b++) {
        if (counts[b] < 1) return "box " + std::to_string(b) + " holds no particles";
        if (!(lengths[b] > 0.0)) return "box " + std::to_string(b) + " has a non-positive length";
        std::string m = check_n(c, counts[b]);
        if (!m.empty()) return "box " + std::to_string(b) + ": " + m;
        total += counts[b];
    }
    return check_n(c, total);
}

// one box of the context's own length
static void single_box(Ctx &c, int64_t n) { set_boxes(c, 1, &n, &c.prm.boxsize); }

// A context that holds a planned evaluation keeps its buffers and boxes for it: other
// evaluations would move them under the captured graph.
static const char *kGraphHeld = "the context holds a planned evaluation (photons_graph_destroy releases it)";

static void plan_from_measure(Ctx &c, double margin);
static void planned_evaluation(Ctx &c, const double *d_pos, double *d_acc);
static int read_sizes(Ctx &c);

// capacities of a kept plan above the measured counts (particles move between steps)
constexpr double kReuseMargin = 0.125;

// One force evaluation of the context's boxes (set_boxes done).  The first one for a box
// set is ordinary (it reads counts back to size its buffers) and plans the next ones;
// those run planned, without a host round trip, and wait once at the end for the
// overflow bits: particles that outgrew the plan are evaluated again the ordinary way
// (and planned anew).  Both forms run the same kernels on the same counts, so they give
// the same result bit for bit.
static void force_eval(Ctx &c, const double *d_pos, double *d_acc)
{
    Reuse &ru = c.reuse;
    if (ru.on && ru.have && ru.box_off == c.box_off && ru.box_len == c.box_len) {
        c.plan = ru.plan;
        c.plan.alloc = !ru.allocated;
        try {
            planned_evaluation(c, d_pos, d_acc);
        } catch (...) {
            c.plan = Plan();
            ru.have = false;
            throw;
        }
        ru.allocated = true;
        c.plan = Plan();
        if (read_sizes(c) == 0) return;
        ru.have = false;  // outgrown: measured again below
    }
    prepare(c, d_pos);
    evaluate(c, d_acc, true);
    if (ru.on) {
        plan_from_measure(c, kReuseMargin);
        ru.plan = c.plan;
        c.plan = Plan();
        ru.have = true;
        ru.allocated = false;
        ru.box_off = c.box_off;
        ru.box_len = c.box_len;
    }
}

}  // namespace photons

extern "C" {

int photons_force_device(photons_ctx *x, int64_t n, const double *d_pos, double *d_acc)
{
    try {
        Ctx &c = x->c;
        if (c.graph_ready) return fail_msg(kGraphHeld);
        std::string m = check_n(c, n);
        if (!m.empty()) return fail_msg(m);
        PH_CHECK(hipSetDevice(c.device));
        c.n = n;
        if (n == 0) {
            c.nleaf = c.nnode = c.ncell = 0;
            return 0;
        }
        single_box(c, n);
        force_eval(c, d_pos, d_acc);
        collect_timing(c);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_force_device", e);
    }
}

int photons_force_boxes_device(photons_ctx *x, int nbox, const int64_t *counts, const double *lengths,
                               const double *d_pos, double *d_acc)
{
    try {
        Ctx &c = x->c;
        if (c.graph_ready) return fail_msg(kGraphHeld);
        std::string m = check_boxes(c, nbox, counts, lengths);
        if (!m.empty()) return fail_msg(m);
        PH_CHECK(hipSetDevice(c.device));
        set_boxes(c, nbox, counts, lengths);
        c.n = c.box_off[nbox];
        force_eval(c, d_pos, d_acc);
        collect_timing(c);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_force_boxes_device", e);
    }
}
// The force of one box of n <= kStaticMaxParticles particles without allocating: every
// buffer of the evaluation comes from the static scratch, the context's own buffers stand
// aside meanwhile (and keep their contents).  Synchronous; two calls that overlap on one
// GPU (other threads) share the scratch and race.
int photons_force_device_static(photons_ctx *x, int64_t n, const double *d_pos, double *d_acc)
{
    try {
        Ctx &c = x->c;
        if (c.graph_ready) return fail_msg(kGraphHeld);
        std::string m = check_n(c, n);
        if (!m.empty()) return fail_msg(m);
        if (n > kStaticMaxParticles)
            return fail_msg("photons_force_device_static: at most " + std::to_string(kStaticMaxParticles) +
                            " particles (the " + std::to_string(kScratchBytes >> 20) + " MiB static scratch)");
        if (n == 0) return 0;  // nothing to compute, nothing written
        PH_CHECK(hipSetDevice(c.device));
        Arena arena{scratch_on(c.device), kScratchBytes, 0};
        // the context as it was (buffers, counts, lists): restored afterwards, so its last
        // ordinary evaluation stays inspectable
        const Ctx kept = c;
        c.n = n;
        for (DevBuf *b : eval_bufs(c)) *b = DevBuf();
        c.tcap = c.lcap = 0;
        g_arena = &arena;
        try {
            single_box(c, n);
            prepare(c, d_pos);
            evaluate(c, d_acc, true);
            PH_CHECK(hipStreamSynchronize(c.stream));
            collect_timing(c);
        } catch (...) {
            g_arena = nullptr;
            (void)hipStreamSynchronize(c.stream);
            c = kept;
            throw;
        }
        g_arena = nullptr;
        double ph[16];
        std::memcpy(ph, c.phase_ms, sizeof ph);
        c = kept;
        std::memcpy(c.phase_ms, ph, sizeof ph);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_force_device_static", e);
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------- planned evaluation, hipGraph

namespace photons {

// capacities `margin` above what the ordinary evaluation measured (plus a fixed slack)
static void plan_from_measure(Ctx &c, double margin)
{
    const Measure &m = c.meas;
    auto up = [&](int64_t v, int64_t slack) { return (int64_t)std::ceil((double)v * (1.0 + margin)) + slack; };
    Plan p;
    p.levels = std::min(m.levels + 2, kMaxDepth);  // moving particles may deepen the tree
    p.depths = p.levels + 1;                      // the deepest targets are leaves
    p.scap = (int)up(m.scap, 64);
    p.tcap = (int)up(m.ntmp, 1024);
    p.lcap = (int)up(m.nleaf, 1024);
    p.dl_cap = up(m.dl, 1024);
    // per depth; depths past the measured ones (a deeper tree) as the last measured one
    p.rec_depth.assign(p.depths, 1024);
    for (int d = 0; d < p.depths && !m.dl_depth.empty(); d++)
        p.rec_depth[d] = std::min(p.dl_cap, up(m.dl_depth[std::min<size_t>(d, m.dl_depth.size() - 1)], 1024));
    p.p2p_cap = up(m.p2p, 4096);
    p.m2l_cap = up(m.m2l, 4096);
    p.item_cap = (int)up(m.items, 1024);
    if ((int64_t)p.tcap + p.lcap >= (1 << 27)) throw std::runtime_error("planned cells exceed 2^27");
    c.plan = p;
}

// the planned evaluation: every launch, no host round trip, no allocation, no memset
static void planned_evaluation(Ctx &c, const double *d_pos, double *d_acc)
{
    c.plan.active = true;
    try {
        prepare(c, d_pos);
        evaluate(c, d_acc, true);
    } catch (...) {
        c.plan.active = false;
        throw;
    }
    c.plan.active = false;
}

// the sizes the device wrote: the host's copies of the counts (stats, list exports)
static int read_sizes(Ctx &c)
{
    DevSizes h;
    PH_CHECK(hipMemcpyAsync(&h, c.sizes.p, offsetof(DevSizes, lvl), hipMemcpyDeviceToHost, c.stream));
    PH_CHECK(hipStreamSynchronize(c.stream));
    if (!h.overflow) {
        c.nleaf = h.nleaf;
        c.nnode = h.nnode;
        c.ncell = h.ncell;
        c.nlevel = c.maxdepth = h.nlevel;
        c.n_p2p = h.n_p2p;
        c.n_m2l = h.n_m2l;
        c.lists_valid = true;
    } else {
        c.lists_valid = false;
    }
    return h.overflow;
}

static void release_graph(Ctx &c)
{
    if (c.graph_exec) (void)hipGraphExecDestroy(c.graph_exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c.graph_exec = nullptr;
    c.graph = nullptr;
    c.graph_ready = c.graph_captured = false;
    c.plan = Plan();
}

}  // namespace photons

extern "C" {

// Plans the evaluation of a fixed set of boxes: one ordinary evaluation of d_pos
// measures the tree and the lists, a first planned evaluation allocates the capacities,
// and (capture != 0) a third is captured into a hipGraph on the context's stream.
// d_acc holds the force of d_pos afterwards.
int photons_graph_create(photons_ctx *x, int nbox, const int64_t *counts, const double *lengths,
                         const double *d_pos, double *d_acc, double margin, int capture)
{
    try {
        Ctx &c = x->c;
        if (!(margin >= 0.0) || margin > 16.0) return fail_msg("margin must be in [0, 16]");
        std::string m = check_boxes(c, nbox, counts, lengths);
        if (!m.empty()) return fail_msg(m);
        PH_CHECK(hipSetDevice(c.device));
        PH_CHECK(hipStreamSynchronize(c.stream));
        release_graph(c);
        set_boxes(c, nbox, counts, lengths);
        c.n = c.box_off[nbox];
        // 1. measure
        prepare(c, d_pos);
        evaluate(c, d_acc, true);
        PH_CHECK(hipStreamSynchronize(c.stream));
        // 2. plan and allocate
        plan_from_measure(c, margin);
        c.plan.alloc = true;
        try {
            planned_evaluation(c, d_pos, d_acc);
        } catch (...) {
            c.plan.alloc = false;
            throw;
        }
        c.plan.alloc = false;
        if (const int ovf = read_sizes(c)) {
            release_graph(c);
            return fail_msg("planned evaluation overflowed its own measurement (bits " + std::to_string(ovf) + ")");
        }
        // 3. capture
        if (capture) {
            const bool timing = c.timing.enabled;
            c.timing.enabled = false;  // no event records inside the graph
            PH_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            try {
                planned_evaluation(c, d_pos, d_acc);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(c.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                c.timing.enabled = timing;
                release_graph(c);
                throw;
            }
            c.timing.enabled = timing;
            PH_CHECK(hipStreamEndCapture(c.stream, &c.graph));
            PH_CHECK(hipGraphInstantiate(&c.graph_exec, c.graph, nullptr, nullptr, 0));
            c.graph_captured = true;
        }
        c.graph_pos = d_pos;
        c.graph_acc = d_acc;
        c.graph_ready = true;
        return 0;
    } catch (const std::exception &e) {
        release_graph(x->c);
        return fail("photons_graph_create", e);
    }
}

// Queues one evaluation of the planned boxes (the captured graph, or the same launches
// one by one when nothing was captured) on the context's stream; returns at once.
int photons_graph_launch(photons_ctx *x)
{
    try {
        Ctx &c = x->c;
        if (!c.graph_ready) return fail_msg("no planned evaluation (photons_graph_create)");
        PH_CHECK(hipSetDevice(c.device));
        if (c.timing.enabled) PH_CHECK(hipEventRecord(c.timing.ev[0], c.stream));
        if (c.graph_captured) {
            PH_CHECK(hipGraphLaunch(c.graph_exec, c.stream));
        } else {
            const bool timing = c.timing.enabled;
            c.timing.enabled = false;
            try {
                planned_evaluation(c, c.graph_pos, c.graph_acc);
            } catch (...) {
                c.timing.enabled = timing;
                throw;
            }
            c.timing.enabled = timing;
        }
        if (c.timing.enabled) PH_CHECK(hipEventRecord(c.timing.ev[1], c.stream));
        c.timing.nev = c.timing.enabled ? 2 : 0;
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_graph_launch", e);
    }
}

// Waits for the launched evaluations; *overflow = the kOvf bits of the last one (0: its
// result is complete; else the boxes outgrew the plan and the graph must be created
// again); *ms = its GPU time when timing is on (else -1).
int photons_graph_check(photons_ctx *x, int *overflow, double *ms)
{
    try {
        Ctx &c = x->c;
        if (!c.graph_ready) return fail_msg("no planned evaluation (photons_graph_create)");
        PH_CHECK(hipSetDevice(c.device));
        const int ovf = read_sizes(c);
        if (overflow) *overflow = ovf;
        if (ms) {
            *ms = -1.0;
            if (c.timing.enabled && c.timing.nev == 2) {
                float t = 0.f;
                PH_CHECK(hipEventElapsedTime(&t, c.timing.ev[0], c.timing.ev[1]));
                *ms = t;
            }
        }
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_graph_check", e);
    }
}

int photons_graph_destroy(photons_ctx *x)
{
    try {
        Ctx &c = x->c;
        PH_CHECK(hipSetDevice(c.device));
        PH_CHECK(hipStreamSynchronize(c.stream));
        release_graph(c);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_graph_destroy", e);
    }
}

}  // extern "C"

extern "C" {


int photons_set_p2p_stage(photons_ctx *x, int stage)
{
    if (stage != 256 && stage != 512) return fail_msg("the P2P stage must hold 256 or 512 particles");
    x->c.p2p_stage = stage;
    return 0;
}

int photons_get_p2p_stage(const photons_ctx *x) { return x->c.p2p_stage; }

// Leaf size (16, 32, 64) and, for FP32-mixed, the P2P staging size (256, 512) timed
// on the caller's own batch: each candidate runs reps + 1 evaluations (the first
// untimed), wall clock from queueing to the end of the last one; the fastest is
// kept in the context and d_acc holds its result.
int photons_autotune_boxes(photons_ctx *x, int nbox, const int64_t *counts, const double *lengths,
                           const double *d_pos, double *d_acc, int reps, int *best_leaf, int *best_stage,
                           double *ms /* [6] or NULL: leaf-major, stage 256 then 512 */)
{
    try {
        Ctx &c = x->c;
        if (c.graph_ready) return fail_msg(kGraphHeld);
        if (reps < 1) return fail_msg("reps must be positive");
        const photons_params keep = c.prm;
        const int keep_stage = c.p2p_stage;
        const int leaves[3] = {16, 32, 64}, stages[2] = {256, 512};
        const bool f32 = c.prm.precision == PHOTONS_FP32_MIXED;
        double best = 1e300;
        int bl = keep.maxleaf, bs = c.p2p_stage;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 2; j++) {
                if (ms) ms[2 * i + j] = -1.0;
                if (!f32 && j == 0) continue;  // the staging size is an FP32 choice
                c.prm.maxleaf = leaves[i];
                c.p2p_stage = stages[j];
                if (!check_boxes(c, nbox, counts, lengths).empty()) continue;  // boxes too small for this leaf
                int rc = photons_force_boxes_device(x, nbox, counts, lengths, d_pos, d_acc);
                if (rc == 0) rc = photons_sync(x);
                if (rc != 0) {
                    c.prm = keep;
                    c.p2p_stage = keep_stage;
                    return rc;
                }
                const auto t0 = std::chrono::steady_clock::now();
                for (int r = 0; r < reps && rc == 0; r++) rc = photons_force_boxes_device(x, nbox, counts, lengths, d_pos, d_acc);
                if (rc == 0) rc = photons_sync(x);
                if (rc != 0) {
                    c.prm = keep;
                    c.p2p_stage = keep_stage;
                    return rc;
                }
                const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
                if (ms) ms[2 * i + j] = t;
                if (t < best) {
                    best = t;
                    bl = leaves[i];
                    bs = stages[j];
                }
            }
        c.prm = keep;
        c.prm.maxleaf = bl;
        c.p2p_stage = bs;
        int rc = photons_force_boxes_device(x, nbox, counts, lengths, d_pos, d_acc);
        if (rc == 0) rc = photons_sync(x);
        if (best_leaf) *best_leaf = bl;
        if (best_stage) *best_stage = bs;
        return rc;
    } catch (const std::exception &e) {
        return fail("photons_autotune_boxes", e);
    }
}

int photons_force_boxes_host(photons_ctx *x, int nbox, const int64_t *counts, const double *lengths,
                             const double *pos, double *acc)
{
    try {
        Ctx &c = x->c;
        if (c.graph_ready) return fail_msg(kGraphHeld);
        std::string m = check_boxes(c, nbox, counts, lengths);
        if (!m.empty()) return fail_msg(m);
        PH_CHECK(hipSetDevice(c.device));
        set_boxes(c, nbox, counts, lengths);
        const int64_t n = c.n = c.box_off[nbox];
        double *dp = reinterpret_cast<double *>(grow(x->io_pos, sizeof(double) * 3 * (size_t)n));
        double *da = reinterpret_cast<double *>(grow(x->io_acc, sizeof(double) * 3 * (size_t)n));
        PH_CHECK(hipMemcpyAsync(dp, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, c.stream));
        force_eval(c, dp, da);
        PH_CHECK(hipMemcpyAsync(acc, da, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, c.stream));
        PH_CHECK(hipStreamSynchronize(c.stream));
        collect_timing(c);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_force_boxes_host", e);
    }
}

int photons_force_host(photons_ctx *x, int64_t n, const double *pos, double *acc)
{
    try {
        Ctx &c = x->c;
        if (c.graph_ready) return fail_msg(kGraphHeld);
        std::string m = check_n(c, n);
        if (!m.empty()) return fail_msg(m);
        PH_CHECK(hipSetDevice(c.device));
        c.n = n;
        if (n == 0) return 0;
        single_box(c, n);
        double *dp = reinterpret_cast<double *>(grow(x->io_pos, sizeof(double) * 3 * (size_t)n));
        double *da = reinterpret_cast<double *>(grow(x->io_acc, sizeof(double) * 3 * (size_t)n));
        PH_CHECK(hipMemcpyAsync(dp, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, c.stream));
        force_eval(c, dp, da);
        PH_CHECK(hipMemcpyAsync(acc, da, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, c.stream));
        PH_CHECK(hipStreamSynchronize(c.stream));
        collect_timing(c);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_force_host", e);
    }
}

// two-phase form used by the drop-in layer: prepare (tree + multipoles), then
// evaluate into a host array in tree order
int photons_prepare_host(photons_ctx *x, int64_t n, const double *pos)
{
    try {
        Ctx &c = x->c;
        if (c.graph_ready) return fail_msg(kGraphHeld);
        std::string m = check_n(c, n);
        if (!m.empty()) return fail_msg(m);
        PH_CHECK(hipSetDevice(c.device));
        c.n = n;
        if (n == 0) return 0;
        single_box(c, n);
        double *dp = reinterpret_cast<double *>(grow(x->io_pos, sizeof(double) * 3 * (size_t)n));
        PH_CHECK(hipMemcpyAsync(dp, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, c.stream));
        prepare(c, dp);
        PH_CHECK(hipStreamSynchronize(c.stream));
        collect_timing(c);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_prepare_host", e);
    }
}

int photons_evaluate_host(photons_ctx *x, double *acc_tree_order)
{
    try {
        Ctx &c = x->c;
        if (c.graph_ready) return fail_msg(kGraphHeld);
        PH_CHECK(hipSetDevice(c.device));
        if (c.n == 0) return 0;
        double *da = reinterpret_cast<double *>(grow(x->io_acc, sizeof(double) * 3 * (size_t)c.n));
        c.timing.nev = 0;  // this call's phases are timed from here
        phase_mark(c);
        evaluate(c, da, false);
        PH_CHECK(hipMemcpyAsync(acc_tree_order, da, sizeof(double) * 3 * (size_t)c.n, hipMemcpyDeviceToHost,
                                c.stream));
        PH_CHECK(hipStreamSynchronize(c.stream));
        collect_timing(c, PHOTONS_PH_WALK);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_evaluate_host", e);
    }
}

// One step of the reference's time loop (src/photoNs.c:136-254) on device-resident
// particles: half kicks with the previous forces (acc_pm, then acc), drift, wrap, a new
// tree and short-range force (the reference rebuilds every step), the PM force on the
// PM handle's stream at the same time, half kicks with the new forces (acc, then acc_pm).
int photons_step_device(photons_ctx *x, photons_pm *pm, int64_t n, double *d_pos, double *d_vel, double *d_acc,
                        double *d_acc_pm, double dkh, double dd)
{
    try {
        if (!x) return fail_msg("photons_step_device: null context");
        Ctx &c = x->c;
        std::string m = check_n(c, n);
        if (!m.empty()) return fail_msg(m);
        if (n == 0) return 0;
        if (!d_pos || !d_vel || !d_acc) return fail_msg("photons_step_device: pos, vel and acc are required");
        if (pm && !d_acc_pm) return fail_msg("photons_step_device: a PM handle needs the acc_pm array");
        PH_CHECK(hipSetDevice(c.device));
        const int64_t n3 = 3 * n;
        const double *apm = pm ? d_acc_pm : nullptr;
        k_kick_drift<<<kick_blocks(n3), 256, 0, c.stream>>>(n3, d_pos, d_vel, d_acc, apm, dkh, dd, c.prm.boxsize, 1,
                                                           nullptr);
        PH_CHECK(hipGetLastError());
        if (pm) {
            // the PM reads the new positions on its own stream
            PH_CHECK(hipStreamSynchronize(c.stream));
            if (photons_pm_force_device(pm, n, d_pos, d_acc_pm) != 0)
                return fail_msg(std::string("photons_step_device: PM: ") + photons_pm_last_error());
        }
        const int rc = photons_force_device(x, n, d_pos, d_acc);
        if (pm && photons_pm_sync(pm) != 0 && rc == 0)
            return fail_msg(std::string("photons_step_device: PM: ") + photons_pm_last_error());
        if (rc != 0) return rc;
        k_kick<<<kick_blocks(n3), 256, 0, c.stream>>>(n3, d_vel, d_acc, nullptr, apm, dkh, 1);
        PH_CHECK(hipGetLastError());
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_step_device", e);
    }
}

int photons_sync(photons_ctx *x)
{
    try {
        PH_CHECK(hipStreamSynchronize(x->c.stream));
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_sync", e);
    }
}

// ---------------------------------------------------------------------------- inspection

__global__ void k_list_stats(int nleaf, const CellInfo *info, const int64_t *off, const int32_t *src,
                             unsigned long long *out)
{
    int T = blockIdx.x * blockDim.x + threadIdx.x;
    if (T >= nleaf) return;
    long long nT = info[T].npart;
    unsigned long long local = 0, inter = 0, image = 0;
    for (int64_t e = span_beg(off, T); e < span_end(off, T); e++) {
        int s = src[e], S = entry_cell(s), sid = entry_shift(s);
        long long nS = info[S].npart;
        long long k = nT * nS;
        if (sid == 0) {
            local++;
            if (S == T) k -= nT;
        } else {
            image += k;
        }
        inter += k;
    }
    atomicAdd(&out[0], local);
    atomicAdd(&out[1], inter);
    atomicAdd(&out[2], image);
}

__global__ void k_m2l_local(int ncell, const int64_t *off, const int32_t *src, unsigned long long *out)
{
    int X = blockIdx.x * blockDim.x + threadIdx.x;
    if (X >= ncell) return;
    unsigned long long local = 0;
    for (int64_t e = span_beg(off, X); e < span_end(off, X); e++) local += entry_shift(src[e]) == 0;
    atomicAdd(&out[3], local);
}

int photons_get_stats(const photons_ctx *xc, photons_stats *st)
{
    try {
        photons_ctx *x = const_cast<photons_ctx *>(xc);
        Ctx &c = x->c;
        std::memset(st, 0, sizeof *st);
        st->n = c.n;
        st->nleaf = c.nleaf;
        st->nnode = c.nnode;
        st->depth = c.maxdepth;
        st->p2p_pairs = c.n_p2p;
        st->m2l_pairs = c.n_m2l;
        if (c.n == 0 || c.ncell == 0 || !c.lists_valid) return 0;
        need(c.p2p_off, sizeof(int64_t) * 2 * (size_t)c.nleaf, "P2P list");
        need(c.p2p_src, sizeof(int32_t) * (size_t)c.n_p2p, "P2P list");
        need(c.m2l_off, sizeof(int64_t) * 2 * (size_t)c.ncell, "M2L list");
        need(c.m2l_src, sizeof(int32_t) * (size_t)c.n_m2l, "M2L list");
        unsigned long long *d = reinterpret_cast<unsigned long long *>(grow(c.stats, 4 * sizeof(unsigned long long)));
        PH_CHECK(hipMemsetAsync(d, 0, 4 * sizeof(unsigned long long), c.stream));
        k_list_stats<<<(c.nleaf + 127) / 128, 128, 0, c.stream>>>(c.nleaf, dptr<CellInfo>(c.info),
                                                                  dptr<int64_t>(c.p2p_off), dptr<int32_t>(c.p2p_src), d);
        k_m2l_local<<<(c.ncell + 127) / 128, 128, 0, c.stream>>>(c.ncell, dptr<int64_t>(c.m2l_off),
                                                                 dptr<int32_t>(c.m2l_src), d);
        unsigned long long h[4];
        PH_CHECK(hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, c.stream));
        PH_CHECK(hipStreamSynchronize(c.stream));
        st->p2p_local_pairs = (int64_t)h[0];
        st->p2p_interactions = (int64_t)h[1];
        st->p2p_image_interactions = (int64_t)h[2];
        st->m2l_local_pairs = (int64_t)h[3];
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_get_stats", e);
    }
}

int photons_tree_perm(const photons_ctx *xc, int32_t *perm)
{
    try {
        photons_ctx *x = const_cast<photons_ctx *>(xc);
        Ctx &c = x->c;
        if (c.n == 0) return 0;
        PH_CHECK(hipMemcpyAsync(perm, c.pidx.p, sizeof(int32_t) * c.n, hipMemcpyDeviceToHost, c.stream));
        PH_CHECK(hipStreamSynchronize(c.stream));
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_tree_perm", e);
    }
}

int photons_tree_cells(const photons_ctx *xc, int32_t *npart, int32_t *ipart, int32_t *son, double *center,
                       double *width)
{
    try {
        photons_ctx *x = const_cast<photons_ctx *>(xc);
        Ctx &c = x->c;
        if (c.ncell == 0) return 0;
        std::vector<CellInfo> I(c.ncell);
        std::vector<CellGeom> G(c.ncell);
        PH_CHECK(hipMemcpyAsync(I.data(), c.info.p, sizeof(CellInfo) * c.ncell, hipMemcpyDeviceToHost, c.stream));
        PH_CHECK(hipMemcpyAsync(G.data(), c.geom.p, sizeof(CellGeom) * c.ncell, hipMemcpyDeviceToHost, c.stream));
        PH_CHECK(hipStreamSynchronize(c.stream));
        for (int k = 0; k < c.ncell; k++) {
            if (npart) npart[k] = I[k].npart;
            if (ipart) ipart[k] = I[k].ipart;
            if (son) {
                son[2 * k] = I[k].son0;
                son[2 * k + 1] = I[k].son1;
            }
            for (int d = 0; d < 3; d++) {
                if (center) center[3 * k + d] = G[k].c[d];
                if (width) width[3 * k + d] = G[k].w[d];
            }
        }
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_tree_cells", e);
    }
}

// Lists are exported target by target, each target's entries as the list holds them:
// in the reference's order, (shift id, depth-first rank of the source cell), the
// order its depth-first walks emit them in (the tests check it entry by entry).
static int export_list(Ctx &c, int ntarget, DevBuf &offb, DevBuf &srcb, int64_t total, int64_t cap, int32_t *tgt,
                       int32_t *src, int8_t *shift, int64_t *count)
{
    if (!c.lists_valid) total = 0;
    *count = total;
    if (!tgt || total == 0) return 0;
    need(offb, sizeof(int64_t) * 2 * (size_t)ntarget, "list spans");
    need(srcb, sizeof(int32_t) * (size_t)total, "list entries");
    std::vector<int64_t> off(2 * (size_t)ntarget);
    std::vector<int32_t> s(total);
    PH_CHECK(hipMemcpyAsync(off.data(), offb.p, sizeof(int64_t) * 2 * ntarget, hipMemcpyDeviceToHost, c.stream));
    PH_CHECK(hipMemcpyAsync(s.data(), srcb.p, sizeof(int32_t) * total, hipMemcpyDeviceToHost, c.stream));
    PH_CHECK(hipStreamSynchronize(c.stream));
    int64_t k = 0;
    for (int t = 0; t < ntarget && k < cap; t++)
        for (int64_t e = off[2 * (size_t)t]; e < off[2 * (size_t)t + 1] && k < cap; e++, k++) {
            tgt[k] = t;
            src[k] = entry_cell(s[e]);
            shift[k] = (int8_t)entry_shift(s[e]);
        }
    return 0;
}

int photons_p2p_list(const photons_ctx *xc, int64_t cap, int32_t *tgt, int32_t *src, int8_t *shift, int64_t *count)
{
    try {
        photons_ctx *x = const_cast<photons_ctx *>(xc);
        return export_list(x->c, x->c.nleaf, x->c.p2p_off, x->c.p2p_src, x->c.n_p2p, cap, tgt, src, shift, count);
    } catch (const std::exception &e) {
        return fail("photons_p2p_list", e);
    }
}

int photons_m2l_list(const photons_ctx *xc, int64_t cap, int32_t *tgt, int32_t *src, int8_t *shift, int64_t *count)
{
    try {
        photons_ctx *x = const_cast<photons_ctx *>(xc);
        return export_list(x->c, x->c.ncell, x->c.m2l_off, x->c.m2l_src, x->c.n_m2l, cap, tgt, src, shift, count);
    } catch (const std::exception &e) {
        return fail("photons_m2l_list", e);
    }
}

int photons_list_lengths(const photons_ctx *xc, int which, int64_t *len)
{
    try {
        photons_ctx *x = const_cast<photons_ctx *>(xc);
        Ctx &c = x->c;
        if (which != 0 && which != 1) return fail_msg("which must be 0 (P2P) or 1 (M2L)");
        const int nt = which == 0 ? c.nleaf : c.ncell;
        if (!c.lists_valid) {
            std::fill(len, len + nt, (int64_t)0);
            return 0;
        }
        DevBuf &offb = which == 0 ? c.p2p_off : c.m2l_off;
        need(offb, sizeof(int64_t) * 2 * (size_t)nt, "list spans");
        std::vector<int64_t> off(2 * (size_t)nt);
        PH_CHECK(hipMemcpyAsync(off.data(), offb.p, sizeof(int64_t) * 2 * nt, hipMemcpyDeviceToHost, c.stream));
        PH_CHECK(hipStreamSynchronize(c.stream));
        for (int t = 0; t < nt; t++) len[t] = off[2 * (size_t)t + 1] - off[2 * (size_t)t];
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_list_lengths", e);
    }
}

int photons_expansions(const photons_ctx *xc, double *M, double *L)
{
    try {
        photons_ctx *x = const_cast<photons_ctx *>(xc);
        Ctx &c = x->c;
        const int NT = nterms_of(c.prm.order);
        size_t cnt = (size_t)c.ncell * NT;
        if (cnt == 0) return 0;
        const bool f32 = c.prm.precision == PHOTONS_FP32_MIXED;
        need(c.M, sizeof(double) * cnt, "multipoles");
        need(c.L, (f32 ? sizeof(float) : sizeof(double)) * cnt, "local expansions");
        if (M) PH_CHECK(hipMemcpyAsync(M, c.M.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, c.stream));
        std::vector<float> lf;
        if (L) {
            if (f32) {
                lf.resize(cnt);
                PH_CHECK(hipMemcpyAsync(lf.data(), c.L.p, sizeof(float) * cnt, hipMemcpyDeviceToHost, c.stream));
            } else {
                PH_CHECK(hipMemcpyAsync(L, c.L.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, c.stream));
            }
        }
        PH_CHECK(hipStreamSynchronize(c.stream));
        if (L && f32) {
            // L_a = L'_a / rs^(|a|+1)
            std::vector<double> sc(NT);
            int k = 0;
            for (int n = 0; n <= c.prm.order; n++)
                for (int a = n; a >= 0; a--)
                    for (int b = n - a; b >= 0; b--) sc[k++] = 1.0 / std::pow(c.prm.rs, n + 1);
            for (size_t i = 0; i < cnt; i++) L[i] = lf[i] * sc[i % NT];
        }
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_expansions", e);
    }
}

// ---------------------------------------------------------------------------- batched operators

int photons_p2m_batch(photons_ctx *x, int order, int count, const double *d_pos, const int32_t *d_first,
                      const int32_t *d_npart, const double *d_center, double mass, double *d_M)
{
    try {
        if (count <= 0) return 0;
        op_p2m(x->c.stream, order, count, d_pos, d_first, d_npart, d_center, mass, d_M);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_p2m_batch", e);
    }
}

int photons_m2m_batch(photons_ctx *x, int order, int count, const double *d_d, const double *d_M, double *d_tM)
{
    try {
        if (count <= 0) return 0;
        op_shift(x->c.stream, order, 0, count, d_d, d_M, d_tM);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_m2m_batch", e);
    }
}

int photons_m2l_batch(photons_ctx *x, int order, int count, double rs, const double *d_R, const double *d_M,
                      double *d_L)
{
    try {
        if (count <= 0) return 0;
        op_m2l(x->c.stream, order, count, rs, d_R, d_M, d_L);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_m2l_batch", e);
    }
}

int photons_l2l_batch(photons_ctx *x, int order, int count, const double *d_d, const double *d_L, double *d_tL)
{
    try {
        if (count <= 0) return 0;
        op_shift(x->c.stream, order, 1, count, d_d, d_L, d_tL);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_l2l_batch", e);
    }
}

int photons_l2p_batch(photons_ctx *x, int order, int n, const double *d_pos, const int32_t *d_owner,
                      const double *d_center, const double *d_L, double *d_acc)
{
    try {
        if (n <= 0) return 0;
        op_l2p(x->c.stream, order, n, d_pos, d_owner, d_center, d_L, d_acc);
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_l2p_batch", e);
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------- drop-in support
// Entry points used by the reference-named functions (compat.c).  Cell ids are
// those of photons_tree_cells; accumulation targets are host arrays in tree order.

namespace photons {
__global__ void k_zero_f64(double *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
}  // namespace photons

extern "C" {

int photons_acceptance(const double *wi, const double *wj, const double *dist, double theta, double rc)
{
    return mac(wi, wj, dist[0], dist[1], dist[2], theta * theta, rc * rc);
}

// Walk from the pair (start_target, start_source) over shifts [shift_lo, shift_hi) and
// evaluate what it emits: P2P added into acc_tree (host, [n][3], tree order) and M2L
// added into the context's L.  Returns the numbers of P2P and M2L entries.
int photons_walk_eval(photons_ctx *x, int start_target, int start_source, int shift_lo, int shift_hi, int do_p2p,
                      int do_m2l, double *acc_tree, int64_t *n_p2p, int64_t *n_m2l)
{
    try {
        Ctx &c = x->c;
        if (c.graph_ready) return fail_msg(kGraphHeld);
        PH_CHECK(hipSetDevice(c.device));
        if (c.ncell == 0) return 0;
        if (start_target < 0 || start_target >= c.ncell || start_source < 0 || start_source >= c.ncell)
            return fail_msg("walk start pair outside the tree");
        run_walk_from(c, shift_lo, shift_hi, start_target, start_source);
        if (n_p2p) *n_p2p = c.n_p2p;
        if (n_m2l) *n_m2l = c.n_m2l;
        if (do_m2l) {
            if (!c.L.p) {
                const int NT = nterms_of(c.prm.order);
                size_t bytes = (c.prm.precision == PHOTONS_FP32_MIXED ? 4 : 8) * (size_t)c.ncell * NT;
                grow(c.L, bytes);
                PH_CHECK(hipMemsetAsync(c.L.p, 0, bytes, c.stream));
            }
            m2l_pass(c, true);
        }
        if (do_p2p && acc_tree) {
            double *da = reinterpret_cast<double *>(grow(x->io_acc, sizeof(double) * 3 * (size_t)c.n));
            PH_CHECK(hipMemcpyAsync(da, acc_tree, sizeof(double) * 3 * c.n, hipMemcpyHostToDevice, c.stream));
            if (c.prm.precision == PHOTONS_FP32_MIXED && !c.ox.p) compute_offsets_f32(c);
            p2p_l2p_launch(c, false, true, false, da);
            PH_CHECK(hipMemcpyAsync(acc_tree, da, sizeof(double) * 3 * c.n, hipMemcpyDeviceToHost, c.stream));
        }
        PH_CHECK(hipStreamSynchronize(c.stream));
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_walk_eval", e);
    }
}

// P2P of one (target leaf, source leaf, shift) pair list, added into acc_tree
// (host, tree order); precision selects the FP64 or the FP32-mixed pair kernel.
int photons_p2p_pairs(photons_ctx *x, int count, const int32_t *tgt, const int32_t *src, const int8_t *shift,
                      int precision, double *acc_tree)
{
    try {
        Ctx &c = x->c;
        if (c.graph_ready) return fail_msg(kGraphHeld);
        PH_CHECK(hipSetDevice(c.device));
        if (count <= 0 || c.nleaf == 0) return 0;
        // CSR over all leaves from the (unsorted) pairs, keeping their order per target
        std::vector<int64_t> off(c.nleaf + 1, 0);
        for (int k = 0; k < count; k++) {
            if (tgt[k] < 0 || tgt[k] >= c.nleaf || src[k] < 0 || src[k] >= c.nleaf)
                return fail_msg("p2p pair outside the leaves");
            off[tgt[k] + 1]++;
        }
        for (int t = 0; t < c.nleaf; t++) off[t + 1] += off[t];
        std::vector<int32_t> ent(count);
        std::vector<int64_t> fill(off.begin(), off.end() - 1);
        for (int k = 0; k < count; k++) ent[fill[tgt[k]]++] = pack_entry(src[k], shift[k]);
        std::vector<int64_t> span(2 * (size_t)c.nleaf);
        for (int t = 0; t < c.nleaf; t++) {
            span[2 * (size_t)t] = off[t];
            span[2 * (size_t)t + 1] = off[t + 1];
        }
        int64_t *doff = reinterpret_cast<int64_t *>(grow(c.p2p_off, sizeof(int64_t) * 2 * (size_t)c.nleaf));
        int32_t *dsrc = reinterpret_cast<int32_t *>(grow(c.p2p_src, sizeof(int32_t) * count));
        PH_CHECK(hipMemcpyAsync(doff, span.data(), sizeof(int64_t) * 2 * c.nleaf, hipMemcpyHostToDevice, c.stream));
        PH_CHECK(hipMemcpyAsync(dsrc, ent.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice, c.stream));
        c.n_p2p = count;
        c.lists_valid = false;  // the M2L list no longer matches this P2P list
        double *da = reinterpret_cast<double *>(grow(x->io_acc, sizeof(double) * 3 * (size_t)c.n));
        PH_CHECK(hipMemcpyAsync(da, acc_tree, sizeof(double) * 3 * c.n, hipMemcpyHostToDevice, c.stream));
        int saved = c.prm.precision;
        c.prm.precision = precision;
        if (precision == PHOTONS_FP32_MIXED) compute_offsets_f32(c);
        try {
            p2p_l2p_launch(c, false, true, false, da);
        } catch (...) {
            c.prm.precision = saved;
            throw;
        }
        c.prm.precision = saved;
        PH_CHECK(hipMemcpyAsync(acc_tree, da, sizeof(double) * 3 * c.n, hipMemcpyDeviceToHost, c.stream));
        PH_CHECK(hipStreamSynchronize(c.stream));
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_p2p_pairs", e);
    }
}

// the context's local expansions, read or overwritten (double, [ncell][nterms])
int photons_local_expansions(photons_ctx *x, double *L, int write)
{
    try {
        Ctx &c = x->c;
        const int NT = nterms_of(c.prm.order);
        size_t cnt = (size_t)c.ncell * NT;
        if (cnt == 0) return 0;
        if (c.prm.precision == PHOTONS_FP32_MIXED) return fail_msg("local expansions are float in this context");
        need(c.L, sizeof(double) * cnt, "local expansions");
        if (write) {
            grow(c.L, sizeof(double) * cnt);
            PH_CHECK(hipMemcpyAsync(c.L.p, L, sizeof(double) * cnt, hipMemcpyHostToDevice, c.stream));
        } else {
            PH_CHECK(hipMemcpyAsync(L, c.L.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, c.stream));
        }
        PH_CHECK(hipStreamSynchronize(c.stream));
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_local_expansions", e);
    }
}

// device scratch of the context for the single-operator entry points
void *photons_scratch(photons_ctx *x, int slot, size_t bytes)
{
    try {
        DevBuf *b[] = {&x->op_a, &x->op_b, &x->op_c, &x->op_d, &x->op_e};
        if (slot < 0 || slot > 4) return nullptr;
        return grow(*b[slot], bytes);
    } catch (const std::exception &e) {
        fail("photons_scratch", e);
        return nullptr;
    }
}

int photons_copy(photons_ctx *x, void *dst, const void *src, size_t bytes, int to_device)
{
    try {
        PH_CHECK(hipMemcpyAsync(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost,
                                x->c.stream));
        PH_CHECK(hipStreamSynchronize(x->c.stream));
        return 0;
    } catch (const std::exception &e) {
        return fail("photons_copy", e);
    }
}

}  // extern "C"
