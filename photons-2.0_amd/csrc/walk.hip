// walk.hip -- device dual-tree walk emitting flat interaction lists.
//
// The reference walks (target, source) cell pairs depth-first from (root, root)
// and emits P2P and M2L tasks (src/fmm.c:406-712 for the home box,
// src/remotes.c:213-552 against each of the 26 shifted, pruned copies of the tree
// built by prepare_sendtree2, src/remotes.c:60-169).  Which pairs it emits depends
// only on the rule applied to each pair, so any traversal that applies the same
// rules emits the same set.
//
// MI355X formulation: a breadth-first walk over the whole pair space.  Each level
// is two launches with one thread per pair (count, then write at scanned offsets):
// the thread applies the rule and appends its children (1, 2 or 4 pairs) or its
// emitted task.  No stack, no atomics, no per-thread divergence beyond one rule.
// The emitted tasks (target, packed source entry) are then grouped by target with a
// stable radix sort on the target id alone, so every list is deterministic.  Within
// one target the entries keep the walk's emission order; the list exports restore
// the reference's depth-first order, which is (shift, source rank in depth-first
// order): the reference finishes a source subtree before its sibling's starts.
//
// Image walks: a shifted node keeps no children when its box is rc or more from
// the domain box or when 0.95*theta*distance exceeds its largest width
// (src/remotes.c:145-158); the rule is evaluated on the fly, no copy is built.
#include <hipcub/hipcub.hpp>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <stdexcept>

#include "engine.h"
#include "fmm_math.h"

namespace photons {

// pair: A in bits 0-26, B in bits 27-53, shift id in bits 54-58
__device__ __forceinline__ uint64_t pair_of(uint32_t A, uint32_t B, uint32_t sid)
{
    return (uint64_t)A | ((uint64_t)B << 27) | ((uint64_t)sid << 54);
}

// task: target cell in bits 32-63, packed source entry (cell, shift id) in bits 0-31
__device__ __forceinline__ uint64_t task_of(uint32_t target, int cell, int sid)
{
    return ((uint64_t)target << 32) | (uint64_t)(uint32_t)pack_entry(cell, sid);
}

struct BfsArgs {
    const CellGeom *geom;
    const CellInfo *info;
    int nleaf;
    double theta, theta2, rc, rc2;
    BoxView bx;  // box lengths: shifts and the domain box [0, L]^3 of the image prune
    const uint64_t *in;
    int64_t n_in;
    uint64_t *out;
    uint32_t *p2p_t, *m2l_t;  // task targets
    int32_t *p2p_s, *m2l_s;   // task source entries
    int64_t p2p_base, m2l_base;  // tasks emitted by earlier levels
    uint8_t *act;                // [n_in] the rule's decision per pair, from the count pass
};

__host__ __device__ inline void shift_of(int sid, double L, double &sx, double &sy, double &sz)
{
    // ids 1..26: (mi, mj, mk) over {-1,0,1}^3 without (0,0,0), mi slowest
    if (sid <= 0) {
        sx = sy = sz = 0.0;
        return;
    }
    int k = sid <= 13 ? sid - 1 : sid;  // skip the centre (index 13 of the 27)
    sx = (double)(k / 9 - 1) * L;
    sy = (double)((k / 3) % 3 - 1) * L;
    sz = (double)(k % 3 - 1) * L;
}

__device__ __forceinline__ bool let_pruned(const CellGeom &root, const CellGeom &g, double sx, double sy, double sz,
                                           double rc, double theta)
{
    double s[3] = {sx, sy, sz};
    double dr = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double d = root.c[i] - g.c[i] - s[i];
        if (d < 0.0) d = -d;
        d -= (root.w[i] + g.w[i]) * 0.5;
        if (d > 0.0) dr += d * d;
    }
    dr = sqrt(dr);
    double wmax = g.w[0];
    if (wmax < g.w[1]) wmax = g.w[1];
    if (wmax < g.w[2]) wmax = g.w[2];
    if (dr >= rc) return true;
    return wmax < 0.95 * theta * dr;
}

// Outputs of the rule for one pair, in four fixed slots (static indices keep them in
// registers): slot c holds a frontier pair when bit c of `kid` is set, a P2P task
// (target, entry) when bit c of `p2p` is set; `m2l` holds an M2L task when has_m2l.
struct RuleOut {
    uint64_t slot[4];
    unsigned kid, p2p;
    uint64_t m2l;
    int has_m2l;
};

// A child pair of two leaves is always a P2P task (the rule below), so it is emitted
// as one straight away instead of passing through the next level's frontier.
__device__ __forceinline__ void put_child(const BfsArgs &a, RuleOut &o, int c, int A, int B, int sid)
{
    if (A < a.nleaf && B < a.nleaf) {
        o.slot[c] = task_of(A, B, sid);
        o.p2p |= 1u << c;
    } else {
        o.slot[c] = pair_of(A, B, sid);
        o.kid |= 1u << c;
    }
}

// The rule applied to one pair (src/fmm.c:406-712, src/remotes.c:213-552), in two parts:
// bfs_decide reads the geometry and returns what the pair does; bfs_emit turns that
// into its children (up to 4, leaf-leaf children as P2P tasks), its P2P task or its M2L
// task, reading only child ids.  The count pass stores the decision (one byte per pair),
// so the write pass does not load the geometry or evaluate the acceptance again.
enum : uint8_t {
    kActNone = 0,  // pruned
    kActSelfLeaf,  // a leaf with itself: P2P
    kActSelfNode,  // a node with itself: its 4 child pairs
    kActLeafLeaf,  // two leaves: P2P, whatever the acceptance says
    kActM2L,       // accepted (or a target leaf against a pruned image node): M2L
    kActOpenA,     // open the target
    kActOpenB      // open the source
};

__device__ __forceinline__ int bfs_decide(const BfsArgs &a, uint64_t pr)
{
    const int A = (int)(pr & 0x7ffffff), B = (int)((pr >> 27) & 0x7ffffff), sid = (int)(pr >> 54);
    const int nleaf = a.nleaf;
    const bool aleaf = A < nleaf, bleaf = B < nleaf;
    if (sid == 0 && A == B) return aleaf ? kActSelfLeaf : kActSelfNode;
    if (aleaf && bleaf) return kActLeafLeaf;  // leaf-leaf pairs are always P2P
    double sx = 0.0, sy = 0.0, sz = 0.0, L = 0.0;
    if (sid != 0) {
        L = a.bx.len_of(A);
        shift_of(sid, L, sx, sy, sz);
    }
    const CellGeom gA = a.geom[A], gB = a.geom[B];
    const double d0 = gA.c[0] - (gB.c[0] + sx);
    const double d1 = gA.c[1] - (gB.c[1] + sy);
    const double d2 = gA.c[2] - (gB.c[2] + sz);
    const int flag = mac(gA.w, gB.w, d0, d1, d2, a.theta2, a.rc2);
    if (flag < 0) return kActNone;
    bool pruned = false;
    if (sid != 0 && !bleaf) {
        const CellGeom box = {{0.5 * L, 0.5 * L, 0.5 * L}, {L, L, L}};
        pruned = let_pruned(box, gB, sx, sy, sz, a.rc, a.theta);
    }
    // accepted; in the image walk a target leaf against a pruned node takes M2L whatever
    // the acceptance says (src/remotes.c:442)
    if (flag == 1 || (aleaf && pruned)) return kActM2L;
    if (aleaf) return kActOpenB;
    if (bleaf) return kActOpenA;
    const bool bigger = (gA.w[0] + gA.w[1] + gA.w[2]) > (gB.w[0] + gB.w[1] + gB.w[2]);
    return (bigger || pruned) ? kActOpenA : kActOpenB;  // ties open the source
}

__device__ __forceinline__ void bfs_emit(const BfsArgs &a, uint64_t pr, int act, RuleOut &o)
{
    o.kid = o.p2p = 0;
    o.has_m2l = 0;
    const int A = (int)(pr & 0x7ffffff), B = (int)((pr >> 27) & 0x7ffffff), sid = (int)(pr >> 54);
    if (act == kActSelfLeaf || act == kActLeafLeaf) {
        o.slot[0] = task_of(A, B, sid);
        o.p2p = 1;
    } else if (act == kActSelfNode) {
        const CellInfo IA = a.info[A];
        put_child(a, o, 0, IA.son0, IA.son0, 0);
        put_child(a, o, 1, IA.son0, IA.son1, 0);
        put_child(a, o, 2, IA.son1, IA.son0, 0);
        put_child(a, o, 3, IA.son1, IA.son1, 0);
    } else if (act == kActM2L) {
        o.m2l = task_of(A, B, sid);
        o.has_m2l = 1;
    } else if (act == kActOpenA) {
        const CellInfo IA = a.info[A];
        put_child(a, o, 0, IA.son0, B, sid);
        put_child(a, o, 1, IA.son1, B, sid);
    } else if (act == kActOpenB) {
        const CellInfo IB = a.info[B];
        put_child(a, o, 0, A, IB.son0, sid);
        put_child(a, o, 1, A, IB.son1, sid);
    }
}

// Each level runs in two passes without atomics: k_bfs_count sums the outputs of
// every tile of kBfsTile pairs, one device scan turns those sums into tile offsets (and
// the exact output sizes, which the host reads), and k_bfs_write re-applies the rule and
// writes each output at its scanned position.
// Within a tile the outputs keep the input order (item k * kBfsThreads + thread), so
// neighbouring threads store to neighbouring addresses.
constexpr int kBfsThreads = 256;
constexpr int kBfsPerThread = 4;
constexpr int kBfsTile = kBfsThreads * kBfsPerThread;

// output counts of one tile (child pairs, P2P tasks, M2L tasks), scanned together
struct Cnt3 {
    int64_t k, p, m;
};
struct AddCnt3 {
    __host__ __device__ Cnt3 operator()(const Cnt3 &x, const Cnt3 &y) const { return {x.k + y.k, x.p + y.p, x.m + y.m}; }
};

__global__ void __launch_bounds__(kBfsThreads) k_bfs_count(BfsArgs a, Cnt3 *__restrict__ tile_cnt, int ntile)
{
    typedef hipcub::BlockReduce<int, kBfsThreads> Reduce;
    __shared__ typename Reduce::TempStorage tmp;
    int ck = 0, cp = 0, cm = 0;
    // all items' pairs and decisions first, the byte stores last: a byte store may alias
    // anything, and between the loads it would keep them from overlapping
    uint64_t pr[kBfsPerThread];
    int act[kBfsPerThread];
#pragma unroll
    for (int k = 0; k < kBfsPerThread; k++) {
        const int64_t i = (int64_t)blockIdx.x * kBfsTile + k * kBfsThreads + threadIdx.x;
        pr[k] = i < a.n_in ? a.in[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kBfsPerThread; k++) {
        const int64_t i = (int64_t)blockIdx.x * kBfsTile + k * kBfsThreads + threadIdx.x;
        act[k] = i < a.n_in ? bfs_decide(a, pr[k]) : kActNone;
    }
#pragma unroll
    for (int k = 0; k < kBfsPerThread; k++) {
        const int64_t i = (int64_t)blockIdx.x * kBfsTile + k * kBfsThreads + threadIdx.x;
        if (i < a.n_in) {
            RuleOut o;
            bfs_emit(a, pr[k], act[k], o);
            ck += __popc(o.kid);
            cp += __popc(o.p2p);
            cm += o.has_m2l;
        }
    }
#pragma unroll
    for (int k = 0; k < kBfsPerThread; k++) {
        const int64_t i = (int64_t)blockIdx.x * kBfsTile + k * kBfsThreads + threadIdx.x;
        if (i < a.n_in) a.act[i] = (uint8_t)act[k];
    }
    int tk = Reduce(tmp).Sum(ck);
    __syncthreads();
    int tp = Reduce(tmp).Sum(cp);
    __syncthreads();
    int tm = Reduce(tmp).Sum(cm);
    if (threadIdx.x == 0) {
        tile_cnt[blockIdx.x] = Cnt3{tk, tp, tm};
        if (blockIdx.x == 0) tile_cnt[ntile] = Cnt3{0, 0, 0};  // the scan's last entry is the total
    }
}

// The counts of a thread's kBfsPerThread items sit in 16-bit fields of one 64-bit
// word (a field's tile sum is at most 4 * kBfsThreads), so one block scan gives the
// offset of every item within the tile's output.
__device__ __forceinline__ int field(uint64_t v, int k) { return (int)((v >> (16 * k)) & 0xffff); }

__global__ void __launch_bounds__(kBfsThreads) k_bfs_write(BfsArgs a, const Cnt3 *__restrict__ tile_off)
{
    static_assert(kBfsPerThread <= 4 && 4 * kBfsThreads < 65536, "16-bit count fields");
    typedef hipcub::BlockScan<uint64_t, kBfsThreads> Scan;
    __shared__ typename Scan::TempStorage tmp;
    RuleOut o[kBfsPerThread];
    uint64_t ck = 0, cp = 0, cm = 0;
#pragma unroll
    for (int k = 0; k < kBfsPerThread; k++) {
        const int64_t i = (int64_t)blockIdx.x * kBfsTile + k * kBfsThreads + threadIdx.x;
        o[k].kid = o[k].p2p = 0;
        o[k].has_m2l = 0;
        if (i < a.n_in) bfs_emit(a, a.in[i], a.act[i], o[k]);
        ck |= (uint64_t)__popc(o[k].kid) << (16 * k);
        cp |= (uint64_t)__popc(o[k].p2p) << (16 * k);
        cm |= (uint64_t)o[k].has_m2l << (16 * k);
    }
    uint64_t ok, op, om, tk, tp, tm;
    Scan(tmp).ExclusiveSum(ck, ok, tk);
    __syncthreads();
    Scan(tmp).ExclusiveSum(cp, op, tp);
    __syncthreads();
    Scan(tmp).ExclusiveSum(cm, om, tm);
    const Cnt3 to = tile_off[blockIdx.x];
    const int64_t bk = to.k;
    const int64_t bp = a.p2p_base + to.p;
    const int64_t bm = a.m2l_base + to.m;
    int sk = 0, sp = 0, sm = 0;  // outputs of the tile's earlier items
#pragma unroll
    for (int k = 0; k < kBfsPerThread; k++) {
        const int64_t qk = bk + sk + field(ok, k), qp = bp + sp + field(op, k);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const unsigned below = (1u << c) - 1u;
            if ((o[k].kid >> c) & 1u) a.out[qk + __popc(o[k].kid & below)] = o[k].slot[c];
            if ((o[k].p2p >> c) & 1u) {
                const int64_t q = qp + __popc(o[k].p2p & below);
                a.p2p_t[q] = (uint32_t)(o[k].slot[c] >> 32);
                a.p2p_s[q] = (int32_t)(uint32_t)o[k].slot[c];
            }
        }
        if (o[k].has_m2l) {
            const int64_t q = bm + sm + field(om, k);
            a.m2l_t[q] = (uint32_t)(o[k].m2l >> 32);
            a.m2l_s[q] = (int32_t)(uint32_t)o[k].m2l;
        }
        sk += field(tk, k);
        sp += field(tp, k);
        sm += field(tm, k);
    }
}

// first frontier: (root, root, shift) of every box for the shifts [lo, hi), or one
// given (A, B) pair per shift when box_root is null
__global__ void k_init_frontier(uint64_t *f, const int *__restrict__ box_root, int nbox, int A, int B, int shift_lo,
                                int shift_hi)
{
    const int nshift = shift_hi - shift_lo;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (box_root ? nbox : 1) * nshift) return;
    const int s = shift_lo + i % nshift;
    if (box_root) {
        const int r = box_root[i / nshift];
        f[i] = pair_of(r, r, s);
    } else {
        f[i] = pair_of(A, B, s);
    }
}

// CSR offsets of the target-sorted tasks: off[t] = first task with target >= t
__global__ void k_csr_offsets(const uint32_t *__restrict__ tgt, int64_t n, int ntarget, int64_t *__restrict__ off)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntarget) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (tgt[mid] < (uint32_t)t) lo = mid + 1;
        else hi = mid;
    }
    off[t] = lo;
}

// ---------------------------------------------------------------------------- host side

static int bits_for(int64_t v)
{
    int b = 1;
    while (b < 40 && ((int64_t)1 << b) <= v) b++;
    return b;
}

// grow to `need` bytes keeping the first `keep` bytes
static void *grow_keep(DevBuf &b, size_t need, size_t keep, hipStream_t st)
{
    if (need <= b.cap) return b.p;
    size_t want = need + need / 2;
    void *nw = nullptr;
    PH_CHECK(hipMalloc(&nw, want));
    if (b.p && keep) PH_CHECK(hipMemcpyAsync(nw, b.p, keep, hipMemcpyDeviceToDevice, st));
    PH_CHECK(hipStreamSynchronize(st));
    if (b.p) PH_CHECK(hipFree(b.p));
    b.p = nw;
    b.cap = want;
    return nw;
}

// Stable LSD radix sort of (target, entry) tasks by target (20-21 bits at 256^3
// particles).  7-bit digits: three passes, measured 1.4 ms faster per evaluation than
// 8-bit digits (also three passes) at 256^3; 10-bit digits (two passes) match 8, and
// 11-bit ones are slower because their onesweep blocks hold 139 KB of LDS.  The item
// count is a size_t, so lists beyond 2^31 entries sort too.
constexpr unsigned kTaskSortBits = 7;
using TaskSortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 8>, rocprim::kernel_config<1024, 8>,
                                        kTaskSortBits, rocprim::block_radix_rank_algorithm::match>>;

// the spare buffers take the alternate halves and may be swapped with the task buffers
static void sort_tasks(Ctx &c, DevBuf &tgt, DevBuf &src, DevBuf &spare_t, DevBuf &spare_s, int64_t n, int end_bit)
{
    if (n <= 1) return;
    uint32_t *t0 = reinterpret_cast<uint32_t *>(tgt.p);
    int32_t *s0 = reinterpret_cast<int32_t *>(src.p);
    uint32_t *t1 = reinterpret_cast<uint32_t *>(grow(spare_t, sizeof(uint32_t) * n));
    int32_t *s1 = reinterpret_cast<int32_t *>(grow(spare_s, sizeof(int32_t) * n));
    rocprim::double_buffer<uint32_t> kb(t0, t1);
    rocprim::double_buffer<int32_t> vb(s0, s1);
    size_t bytes = 0;
    PH_CHECK(rocprim::radix_sort_pairs<TaskSortConfig>(nullptr, bytes, kb, vb, (size_t)n, 0, end_bit, c.stream));
    void *tmp = grow(c.scan_tmp, bytes);
    PH_CHECK(rocprim::radix_sort_pairs<TaskSortConfig>(tmp, bytes, kb, vb, (size_t)n, 0, end_bit, c.stream));
    if (kb.current() != t0) std::swap(tgt, spare_t);
    if (vb.current() != s0) std::swap(src, spare_s);
}

void run_walk_from(Ctx &c, int shift_lo, int shift_hi, int start_target, int start_source)
{
    hipStream_t st = c.stream;
    const int nleaf = c.nleaf, ncell = c.ncell;
    if (ncell >= (1 << 27)) throw std::runtime_error("more than 2^27 cells");

    BfsArgs a{};
    a.geom = dptr<CellGeom>(c.geom);
    a.info = dptr<CellInfo>(c.info);
    a.nleaf = nleaf;
    a.theta = c.prm.theta;
    a.theta2 = c.prm.theta * c.prm.theta;
    a.rc = c.prm.rc;
    a.rc2 = c.prm.rc * c.prm.rc;
    a.bx = box_view(c);
    need(c.cell_box, sizeof(int) * (size_t)ncell, "cell boxes");

    // first frontier: the roots of all boxes, or the given pair (start_target >= 0)
    const bool forest = start_target < 0;
    const int nshift = shift_hi - shift_lo;
    const int64_t n0 = (int64_t)(forest ? c.nbox : 1) * nshift;
    if (forest) need(c.d_box_root, sizeof(int) * (size_t)c.nbox, "box roots");
    grow_keep(c.bfs_a, sizeof(uint64_t) * std::max<int64_t>(n0, 64), 0, st);
    int64_t *pin = reinterpret_cast<int64_t *>(c.pinned.p) + 16;
    k_init_frontier<<<(int)((n0 + 255) / 256), 256, 0, st>>>(reinterpret_cast<uint64_t *>(c.bfs_a.p),
                                                             forest ? dptr<int>(c.d_box_root) : nullptr, c.nbox,
                                                             start_target, start_source, shift_lo, shift_hi);
    int64_t n_in = n0, n_p2p = 0, n_m2l = 0;
    int levels = 0;
    while (n_in > 0) {
        const int ntile = (int)((n_in + kBfsTile - 1) / kBfsTile);
        Cnt3 *tile_cnt = reinterpret_cast<Cnt3 *>(grow(c.cnt_tmp, sizeof(Cnt3) * 2 * (size_t)(ntile + 1)));
        Cnt3 *tile_off = tile_cnt + (ntile + 1);
        a.in = reinterpret_cast<uint64_t *>(c.bfs_a.p);
        a.n_in = n_in;
        a.act = static_cast<uint8_t *>(grow(c.bfs_act, (size_t)n_in));
        // one count pass, one scan of the three counts, one copy of the three totals
        k_bfs_count<<<ntile, kBfsThreads, 0, st>>>(a, tile_cnt, ntile);
        {
            size_t bytes = 0;
            PH_CHECK(hipcub::DeviceScan::ExclusiveScan(nullptr, bytes, tile_cnt, tile_off, AddCnt3(), Cnt3{0, 0, 0},
                                                       ntile + 1, st));
            void *tmp = grow(c.scan_tmp, bytes);
            PH_CHECK(hipcub::DeviceScan::ExclusiveScan(tmp, bytes, tile_cnt, tile_off, AddCnt3(), Cnt3{0, 0, 0},
                                                       ntile + 1, st));
        }
        PH_CHECK(hipMemcpyAsync(pin, tile_off + ntile, sizeof(Cnt3), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipEventRecord(c.counts_ready, st));
        a.p2p_base = n_p2p;
        a.m2l_base = n_m2l;
        // A pair has at most 4 children, 4 P2P tasks and 1 M2L task.  When the buffers
        // already hold that bound (repeated evaluations keep their capacity) the write is
        // queued at once and the host reads the counts while it runs; otherwise the
        // buffers grow to the exact counts first.
        const bool fits = c.bfs_b.cap >= sizeof(uint64_t) * 4 * (size_t)n_in &&
                          c.p2p_keys.cap >= 4 * (size_t)(n_p2p + 4 * n_in) &&
                          c.p2p_src.cap >= 4 * (size_t)(n_p2p + 4 * n_in) &&
                          c.m2l_keys.cap >= 4 * (size_t)(n_m2l + n_in) && c.m2l_src.cap >= 4 * (size_t)(n_m2l + n_in);
        if (fits) {
            a.out = reinterpret_cast<uint64_t *>(c.bfs_b.p);
            a.p2p_t = static_cast<uint32_t *>(c.p2p_keys.p);
            a.p2p_s = static_cast<int32_t *>(c.p2p_src.p);
            a.m2l_t = static_cast<uint32_t *>(c.m2l_keys.p);
            a.m2l_s = static_cast<int32_t *>(c.m2l_src.p);
            k_bfs_write<<<ntile, kBfsThreads, 0, st>>>(a, tile_off);
        }
        PH_CHECK(hipEventSynchronize(c.counts_ready));
        const int64_t n_out = pin[0], add_p2p = pin[1], add_m2l = pin[2];
        if (!fits) {
            grow_keep(c.bfs_b, sizeof(uint64_t) * std::max<int64_t>(n_out, 1), 0, st);
            const size_t np = std::max<int64_t>(n_p2p + add_p2p, 1), nm = std::max<int64_t>(n_m2l + add_m2l, 1);
            a.p2p_t = static_cast<uint32_t *>(grow_keep(c.p2p_keys, 4 * np, 4 * n_p2p, st));
            a.p2p_s = static_cast<int32_t *>(grow_keep(c.p2p_src, 4 * np, 4 * n_p2p, st));
            a.m2l_t = static_cast<uint32_t *>(grow_keep(c.m2l_keys, 4 * nm, 4 * n_m2l, st));
            a.m2l_s = static_cast<int32_t *>(grow_keep(c.m2l_src, 4 * nm, 4 * n_m2l, st));
            a.out = reinterpret_cast<uint64_t *>(c.bfs_b.p);
            k_bfs_write<<<ntile, kBfsThreads, 0, st>>>(a, tile_off);
        }
        n_in = n_out;
        n_p2p += add_p2p;
        n_m2l += add_m2l;
        std::swap(c.bfs_a, c.bfs_b);
        if (++levels > 4 * kMaxDepth + 8) throw std::runtime_error("dual-tree walk does not terminate");
    }
    c.n_p2p = n_p2p;
    c.n_m2l = n_m2l;
    c.lists_valid = true;

    // group by target, then CSR offsets
    sort_tasks(c, c.p2p_keys, c.p2p_src, c.bfs_a, c.bfs_b, n_p2p, bits_for(nleaf));
    sort_tasks(c, c.m2l_keys, c.m2l_src, c.bfs_a, c.bfs_b, n_m2l, bits_for(ncell));
    int64_t *p2p_off = reinterpret_cast<int64_t *>(grow(c.p2p_off, sizeof(int64_t) * (size_t)(nleaf + 1)));
    int64_t *m2l_off = reinterpret_cast<int64_t *>(grow(c.m2l_off, sizeof(int64_t) * (size_t)(ncell + 1)));
    grow(c.p2p_src, sizeof(int32_t));
    grow(c.m2l_src, sizeof(int32_t));
    k_csr_offsets<<<(nleaf + 1 + 255) / 256, 256, 0, st>>>(reinterpret_cast<uint32_t *>(c.p2p_keys.p), n_p2p, nleaf,
                                                           p2p_off);
    k_csr_offsets<<<(ncell + 1 + 255) / 256, 256, 0, st>>>(reinterpret_cast<uint32_t *>(c.m2l_keys.p), n_m2l, ncell,
                                                           m2l_off);
    PH_CHECK(hipGetLastError());
}

void run_walk(Ctx &c, int shift_lo, int shift_hi) { run_walk_from(c, shift_lo, shift_hi, -1, -1); }

}  // namespace photons
