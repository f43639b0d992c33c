// walk.hip -- device dual-tree walk emitting interaction lists grouped by target.
//
// The reference walks (target, source) cell pairs depth-first from (root, root)
// and emits P2P and M2L tasks (src/fmm.c:406-712 for the home box,
// src/remotes.c:213-552 against each of the 26 shifted, pruned copies of the tree
// built by prepare_sendtree2, src/remotes.c:60-169).  Which pairs it emits depends
// only on the rule applied to each pair, so any traversal that applies the same
// rules emits the same set.
//
// MI355X formulation: one step per depth of the target tree.  A pair (A, B) whose
// rule opens the target A passes the source B on to both children of A, and a node
// paired with itself passes its two children; so each node keeps one list of sources
// for its children (its down list), and the pairs whose target is a cell X at depth d
// are X against the down list of X's parent.  At depth d every such pair (one thread
// each) walks the source subtree of its entry in depth-first order without a stack:
// at source rank r the rule either descends (opens the source: r + 1) or settles the
// whole subtree (P2P, M2L, pass down or nothing: r + subtree size).  Two passes per
// depth (count, then write at scanned offsets) and no atomics.  The pairs of one target
// are contiguous and its sources come in (shift, depth-first rank) order, so every
// list is written grouped by target in the reference's order: no sort.
//
// Image walks: a shifted node keeps no children when its box is rc or more from
// the domain box or when 0.95*theta*distance exceeds its largest width
// (src/remotes.c:145-158); the rule is evaluated on the fly, no copy is built.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>

#include "engine.h"
#include "fmm_math.h"

namespace photons {

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

// what the rule does with one pair
enum : int {
    kActNone = 0,  // pruned
    kActP2P,       // a leaf with itself, or two leaves: P2P, whatever the acceptance says
    kActSelfNode,  // a node with itself: its children pass down
    kActM2L,       // accepted (or a target leaf against a pruned image node): M2L
    kActOpenA,     // open the target: the source passes down
    kActOpenB      // open the source
};

// by depth-first rank: what a traversal reads of a source cell, in one 64-byte record
struct alignas(16) WalkRec {
    CellGeom g;
    int32_t cell, dsize;  // cell id; subtree size (1: a leaf)
    int32_t kid[2];       // a node's children, first and second in depth-first order
};
static_assert(sizeof(WalkRec) == 64, "one 64-byte record per rank");

// a target of the current depth and what the rule needs of it, written by k_walk_prep
struct Target {
    CellGeom g;
    double L;              // its box length (shifts, the image prune's domain box)
    int64_t pbeg;          // its pending entries: pend[pbeg ..], one per pair
    int32_t X, rank;       // cell id, depth-first rank
    int32_t son0r, son1r;  // ranks of the children (nodes)
    int32_t leaf, pad;
};

// a traversal left to a wavefront: pair q of target j goes on at rank r (< rend, the end
// of its source subtree) after np, nm, nd outputs
struct Cont {
    int32_t q, j, r, rend;
    uint32_t np, nm, nd, pad;
};

struct WalkArgs {
    const CellInfo *info;
    const WalkRec *rec;       // by depth-first rank
    DevSizes *sz;             // cell counts, the targets of each depth, list bases, overflow
    double theta, theta2, rc, rc2;
    BoxView bx;
    // this depth: its targets are walk_tgt[tgt_off[d] .. tgt_off[d+1])
    const int *walk_tgt;
    int d;
    int64_t tgt_cap;          // launch bound of the targets
    Target *tgt;              // [tgt_cap]: the depth's targets (k_walk_prep)
    const int64_t *pair_off;  // [ntgt + 1]: first pair of each target (scan of the pending counts)
    const int32_t *tile_tgt;  // per tile of kWalkThreads pairs: the target of its first pair
    int64_t npair_cap;        // pairs the per-pair arrays and the launch hold
    const int32_t *pend;      // pending entries: the previous depth's down lists (or the start list)
    uint32_t *pcnt;           // [3 npair]: outputs of each pair (P2P, M2L, down), count pass -> write pass
    Cont *big;                // [npair]: the traversals continued by a wavefront each
    // outputs of the write pass
    int32_t *p2p_out, *m2l_out, *dl_out;  // list bases: sz->p2p_base, sz->m2l_base
    int64_t *p2p_span, *m2l_span, *dl_span;  // per leaf / cell / node (index cell - nleaf)
};

// Entries of the pending and down lists name the source by its depth-first rank (the
// traversal starts there without a cell lookup); the P2P and M2L lists name cells.
__device__ __forceinline__ int entry_rank(int32_t e) { return entry_cell(e); }

__device__ __forceinline__ int decide(const WalkArgs &a, const Target &t, const WalkRec &R, int r, int sid,
                                      double sx, double sy, double sz)
{
    const bool bleaf = R.dsize == 1;
    if (sid == 0 && r == t.rank) return t.leaf ? kActP2P : kActSelfNode;
    if (t.leaf && bleaf) return kActP2P;  // leaf-leaf pairs are always P2P
    const double d0 = t.g.c[0] - (R.g.c[0] + sx);
    const double d1 = t.g.c[1] - (R.g.c[1] + sy);
    const double d2 = t.g.c[2] - (R.g.c[2] + sz);
    const int flag = mac(t.g.w, R.g.w, d0, d1, d2, a.theta2, a.rc2);
    if (flag < 0) return kActNone;
    bool pruned = false;
    if (sid != 0 && !bleaf) {
        const double L = t.L;
        const CellGeom box = {{0.5 * L, 0.5 * L, 0.5 * L}, {L, L, L}};
        pruned = let_pruned(box, R.g, sx, sy, sz, a.rc, a.theta);
    }
    // accepted; in the image walk a target leaf against a pruned node takes M2L whatever
    // the acceptance says (src/remotes.c:442)
    if (flag == 1 || (t.leaf && pruned)) return kActM2L;
    if (t.leaf) return kActOpenB;
    if (bleaf) return kActOpenA;
    const bool bigger = (t.g.w[0] + t.g.w[1] + t.g.w[2]) > (R.g.w[0] + R.g.w[1] + R.g.w[2]);
    return (bigger || pruned) ? kActOpenA : kActOpenB;  // ties open the source
}

constexpr int kWalkThreads = 256;  // pairs per tile

// A thread applies the rule to at most kThreadVisits ranks of its pair; a traversal that
// goes on (a target leaf beside a dense cluster opens one source node after another: one
// pair in a few hundred, but a wavefront waits for its slowest lane) is continued from the
// rank it reached by a whole wavefront, 64 ranks at a time (wave_traverse).
constexpr int kThreadVisits = 16;

// The pairs of target t with the source subtree at rank r (record R, already loaded), in
// depth-first order, for at most kThreadVisits ranks; returns the rank it stopped at (>=
// the subtree's end when it finished).  WRITE = false counts the outputs; WRITE = true
// stores them from the given positions, never more than the count pass found (lp, lm,
// ld), so a store stays inside the ranges the scan gave this pair.
template <bool WRITE>
__device__ __forceinline__ int traverse(const WalkArgs &a, const Target &t, int r, int sid, WalkRec R, int &rend,
                                        uint32_t &np, uint32_t &nm, uint32_t &nd, int64_t qp, int64_t qm, int64_t qd,
                                        uint32_t lp = 0, uint32_t lm = 0, uint32_t ld = 0)
{
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (sid != 0) shift_of(sid, t.L, sx, sy, sz);
    rend = r + R.dsize;
    for (int visits = 1;; visits++) {
        const int act = decide(a, t, R, r, sid, sx, sy, sz);
        if (act == kActP2P) {
            if (WRITE && np < lp) a.p2p_out[qp + np] = pack_entry(R.cell, sid);
            np++;
        } else if (act == kActM2L) {
            if (WRITE && nm < lm) a.m2l_out[qm + nm] = pack_entry(R.cell, sid);
            nm++;
        } else if (act == kActOpenA) {
            if (WRITE && nd < ld) a.dl_out[qd + nd] = pack_entry(r, sid);
            nd++;
        } else if (act == kActSelfNode) {
            if (WRITE && nd + 1 < ld) {
                a.dl_out[qd + nd] = pack_entry(t.son0r, 0);
                a.dl_out[qd + nd + 1] = pack_entry(t.son1r, 0);
            }
            nd += 2;
        }
        if (act == kActOpenB && t.leaf) {
            // a target leaf against a leaf takes P2P whatever the geometry: the opened
            // node's leaf children are emitted here, in rank order, without a visit
            const int nleaf = a.sz->nleaf;
            r += 1;
            if (R.kid[0] < nleaf) {
                if (WRITE && np < lp) a.p2p_out[qp + np] = pack_entry(R.kid[0], sid);
                np++;
                r += 1;
                if (R.kid[1] < nleaf) {
                    if (WRITE && np < lp) a.p2p_out[qp + np] = pack_entry(R.kid[1], sid);
                    np++;
                    r += 1;
                }
            }
        } else {
            r += act == kActOpenB ? 1 : R.dsize;
        }
        if (r >= rend || visits == kThreadVisits) return r;
        R = a.rec[r];
    }
}

// The rest of a traversal, ranks [r, rend), by a wavefront: lane l takes rank carry + l
// and applies the rule; a rank is visited unless an earlier rank of the chunk settled a
// subtree containing it (an exclusive prefix max of the ranks each lane accounts for: a
// settled rank its whole subtree, an opened one itself; a rank inside a settled subtree
// accounts for no more than that subtree, and every subtree settled before r ends by r);
// the visited ranks' outputs are emitted in rank order, as the single-thread walk emits
// them.  A pair's own target is only ever met at its first rank (self pair), which the
// thread has taken, so no rank here passes two children down.  np, nm, nd: wave-uniform
// counts; WRITE stores from qp, qm, qd, never past lp, lm, ld.
template <bool WRITE>
__device__ __forceinline__ void wave_traverse(const WalkArgs &a, const Target &t, int r, int rend, int sid, int lane,
                                              uint32_t &np, uint32_t &nm, uint32_t &nd, int64_t qp = 0,
                                              int64_t qm = 0, int64_t qd = 0, uint32_t lp = 0, uint32_t lm = 0,
                                              uint32_t ld = 0)
{
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (sid != 0) shift_of(sid, t.L, sx, sy, sz);
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int carry = r;
    while (carry < rend) {
        const int s = carry + lane;
        int act = kActNone, cov = s + 1, v = 0;
        if (s < rend) {
            const WalkRec R = a.rec[s];
            v = R.cell;
            act = decide(a, t, R, s, sid, sx, sy, sz);
            cov = act == kActOpenB ? s + 1 : s + R.dsize;
        }
        int incl = cov;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(incl, d, 64);
            if (lane >= d) incl = max(incl, o);
        }
        int excl = __shfl_up(incl, 1, 64);
        excl = lane == 0 ? carry : max(excl, carry);
        const bool visited = s < rend && s >= excl;
        const bool isp = visited && act == kActP2P, ism = visited && act == kActM2L;
        const bool isd = visited && act == kActOpenA;
        const uint64_t bp = __ballot(isp), bm = __ballot(ism), bd = __ballot(isd);
        if (WRITE) {
            const uint32_t ip = np + __popcll(bp & below), im = nm + __popcll(bm & below);
            const uint32_t id = nd + __popcll(bd & below);
            if (isp && ip < lp) a.p2p_out[qp + ip] = pack_entry(v, sid);
            if (ism && im < lm) a.m2l_out[qm + im] = pack_entry(v, sid);
            if (isd && id < ld) a.dl_out[qd + id] = pack_entry(s, sid);
        }
        np += __popcll(bp);
        nm += __popcll(bm);
        nd += __popcll(bd);
        carry = max(carry + 64, __shfl(incl, 63, 64));
    }
}

// the target of pair q: the last j in [lo, hi) with pair_off[j] <= q
__device__ __forceinline__ int target_of(const int64_t *__restrict__ pair_off, int lo, int hi, int64_t q)
{
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pair_off[mid] <= q) lo = mid;
        else hi = mid;
    }
    return lo;
}

// the depth's targets, at most the launch bound (a depth with more flags the overflow in
// k_walk_prep, and the pair arrays hold only the bound's targets)
__device__ __forceinline__ int ntgt_all(const WalkArgs &a) { return (int)(a.sz->tgt_off[a.d + 1] - a.sz->tgt_off[a.d]); }
__device__ __forceinline__ int ntgt_of(const WalkArgs &a) { return (int)min((int64_t)ntgt_all(a), a.tgt_cap); }
__device__ __forceinline__ const int *tgt_of(const WalkArgs &a) { return a.walk_tgt + a.sz->tgt_off[a.d]; }
__device__ __forceinline__ int64_t npair_of(const WalkArgs &a, int ntgt) { return min(a.pair_off[ntgt], a.npair_cap); }

// The first pair offsets of the targets from the tile's first target on, in LDS: a tile
// of pairs rarely spans more than kWalkThreads targets.  Called by every thread of the
// block (it synchronises); `live` says whether the tile holds pairs (block-uniform).
__device__ __forceinline__ int tile_window(const WalkArgs &a, int ntgt, bool live, int64_t *s_off)
{
    int j0 = 0;
    if (live) {
        j0 = a.tile_tgt[blockIdx.x];
        for (int i = threadIdx.x; i <= kWalkThreads; i += kWalkThreads) s_off[i] = a.pair_off[min(j0 + i, ntgt)];
    }
    __syncthreads();
    return j0;
}

// the target of pair q of this tile: a search of the LDS window, else of the rest
__device__ __forceinline__ int tile_target(const WalkArgs &a, const int64_t *s_off, int j0, int ntgt, int64_t q)
{
    if (q < s_off[kWalkThreads]) {
        int lo = 0, hi = kWalkThreads;  // s_off[lo] <= q < s_off[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_off[mid] <= q) lo = mid;
            else hi = mid;
        }
        return j0 + lo;
    }
    // pair_off[j0 + kWalkThreads] <= q < pair_off[ntgt]
    return target_of(a.pair_off, j0 + kWalkThreads, ntgt, q);
}

// per tile, the target of its first pair: each target marks the tile starts in its range
__global__ void k_tile_targets(WalkArgs a, int32_t *__restrict__ tile_tgt)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int ntgt = ntgt_of(a);
    if (j >= ntgt) return;
    const int64_t npair = npair_of(a, ntgt);
    const int64_t lo = a.pair_off[j], hi = min(a.pair_off[j + 1], npair);
    for (int64_t k = (lo + kWalkThreads - 1) / kWalkThreads; k * kWalkThreads < hi; k++) tile_tgt[k] = j;
}

struct Cnt3 {
    int64_t p, m, d;
};
struct AddCnt3 {
    __host__ __device__ Cnt3 operator()(const Cnt3 &x, const Cnt3 &y) const { return {x.p + y.p, x.m + y.m, x.d + y.d}; }
};

__global__ void __launch_bounds__(kWalkThreads) k_walk_count(WalkArgs a, Cnt3 *__restrict__ tile_cnt, int ntile)
{
    typedef hipcub::BlockReduce<int64_t, kWalkThreads> Reduce;
    __shared__ typename Reduce::TempStorage tmp;
    __shared__ int64_t s_off[kWalkThreads + 1];
    const int ntgt = ntgt_of(a);
    const int64_t npair_all = a.pair_off[ntgt];
    const int64_t npair = min(npair_all, a.npair_cap);
    if (blockIdx.x == 0 && threadIdx.x == 0 && npair_all > a.npair_cap) a.sz->overflow |= kOvfPairs;
    const int64_t q0 = (int64_t)blockIdx.x * kWalkThreads, q = q0 + threadIdx.x;
    const int j0 = tile_window(a, ntgt, q0 < npair, s_off);
    uint32_t np = 0, nm = 0, nd = 0;
    bool more = false;
    Cont cont;
    if (q < npair) {
        const int j = tile_target(a, s_off, j0, ntgt, q);
        const Target t = a.tgt[j];
        const int32_t entry = a.pend[t.pbeg + (q - a.pair_off[j])];
        const int r = entry_rank(entry);
        int rend;
        const int stop = traverse<false>(a, t, r, entry_shift(entry), a.rec[r], rend, np, nm, nd, 0, 0, 0);
        more = stop < rend;
        if (more) cont = Cont{(int32_t)q, j, stop, rend, np, nm, nd, 0};
        a.pcnt[3 * q] = np;
        a.pcnt[3 * q + 1] = nm;
        a.pcnt[3 * q + 2] = nd;
    }
    {
        // the continued traversals, counted by k_walk_deep_count (which adds them to this
        // tile's sums): one append per wavefront
        const uint64_t bal = __ballot(more);
        if (bal) {
            const int lane = threadIdx.x % 64;
            unsigned long long base = 0;
            if (lane == __ffsll((unsigned long long)bal) - 1)
                base = atomicAdd(reinterpret_cast<unsigned long long *>(&a.sz->n_big), (unsigned long long)__popcll(bal));
            base = __shfl(base, __ffsll((unsigned long long)bal) - 1, 64);
            if (more) a.big[base + __popcll(bal & ((1ull << lane) - 1ull))] = cont;
        }
    }
    const int64_t tp = Reduce(tmp).Sum((int64_t)np);
    __syncthreads();
    const int64_t tm = Reduce(tmp).Sum((int64_t)nm);
    __syncthreads();
    const int64_t td = Reduce(tmp).Sum((int64_t)nd);
    if (threadIdx.x == 0) {
        tile_cnt[blockIdx.x] = Cnt3{tp, tm, td};
        if (blockIdx.x == 0) tile_cnt[ntile] = Cnt3{0, 0, 0};  // the scan's last entry is the total
    }
}

__global__ void __launch_bounds__(kWalkThreads) k_walk_write(WalkArgs a, const Cnt3 *__restrict__ tile_off)
{
    typedef hipcub::BlockScan<int64_t, kWalkThreads> Scan;
    __shared__ typename Scan::TempStorage tmp;
    __shared__ int64_t s_off[kWalkThreads + 1];
    if (a.sz->overflow) return;  // the same for every thread: no block scan is left half-entered
    const int ntgt = ntgt_of(a);
    const int64_t npair = npair_of(a, ntgt);
    const int64_t q0 = (int64_t)blockIdx.x * kWalkThreads, q = q0 + threadIdx.x;
    const bool live = q < npair;
    int64_t np = 0, nm = 0, nd = 0;
    if (live) {
        np = a.pcnt[3 * q];
        nm = a.pcnt[3 * q + 1];
        nd = a.pcnt[3 * q + 2];
    }
    int64_t op, om, od;
    Scan(tmp).ExclusiveSum(np, op);
    __syncthreads();
    Scan(tmp).ExclusiveSum(nm, om);
    __syncthreads();
    Scan(tmp).ExclusiveSum(nd, od);
    const int j0 = tile_window(a, ntgt, q0 < npair, s_off);
    if (!live) return;
    const Cnt3 to = tile_off[blockIdx.x];
    const int64_t qp = a.sz->p2p_base + to.p + op, qm = a.sz->m2l_base + to.m + om, qd = to.d + od;
    const int j = tile_target(a, s_off, j0, ntgt, q);
    const int64_t pj = a.pair_off[j];
    const Target t = a.tgt[j];
    if (np + nm + nd > 0) {
        // the same ranks as the count pass; a continued traversal's rest is k_walk_deep_write's
        const int32_t entry = a.pend[t.pbeg + (q - pj)];
        const int r = entry_rank(entry);
        int rend;
        uint32_t cp = 0, cm = 0, cd = 0;
        traverse<true>(a, t, r, entry_shift(entry), a.rec[r], rend, cp, cm, cd, qp, qm, qd, (uint32_t)np,
                       (uint32_t)nm, (uint32_t)nd);
    }
    // the target's first pair opens its spans, its last pair closes them
    const bool first = q == pj, last = q + 1 == a.pair_off[j + 1];
    const int nleaf = a.sz->nleaf;
    if (first) {
        if (t.leaf) a.p2p_span[2 * (size_t)t.X] = qp;
        else a.dl_span[2 * (size_t)(t.X - nleaf)] = qd;
        a.m2l_span[2 * (size_t)t.X] = qm;
    }
    if (last) {
        if (t.leaf) a.p2p_span[2 * (size_t)t.X + 1] = qp + np;
        else a.dl_span[2 * (size_t)(t.X - nleaf) + 1] = qd + nd;
        a.m2l_span[2 * (size_t)t.X + 1] = qm + nm;
    }
}

// The continued traversals, one wavefront each (a fixed grid walking the list of them):
// their counts into pcnt and, by atomic adds of integers (order-free), into their tiles'
// sums.
constexpr int kDeepWaves = 8192;

__global__ void __launch_bounds__(256) k_walk_deep_count(WalkArgs a, Cnt3 *__restrict__ tile_cnt)
{
    const int lane = threadIdx.x % 64;
    const int64_t nbig = a.sz->n_big;
    for (int64_t k = (int64_t)blockIdx.x * 4 + threadIdx.x / 64; k < nbig; k += (int64_t)gridDim.x * 4) {
        const Cont ct = a.big[k];
        const int64_t q = ct.q;
        const Target t = a.tgt[ct.j];
        const int32_t entry = a.pend[t.pbeg + (q - a.pair_off[ct.j])];
        uint32_t np = 0, nm = 0, nd = 0;
        wave_traverse<false>(a, t, ct.r, ct.rend, entry_shift(entry), lane, np, nm, nd);
        if (lane == 0) {
            a.pcnt[3 * q] = ct.np + np;
            a.pcnt[3 * q + 1] = ct.nm + nm;
            a.pcnt[3 * q + 2] = ct.nd + nd;
            unsigned long long *tc = reinterpret_cast<unsigned long long *>(tile_cnt + q / kWalkThreads);
            atomicAdd(tc, (unsigned long long)np);
            atomicAdd(tc + 1, (unsigned long long)nm);
            atomicAdd(tc + 2, (unsigned long long)nd);
        }
    }
}

// The continued traversals' outputs: the pair's offsets (the tile's offset plus the
// outputs of the tile's earlier pairs, summed by the wavefront) past what its thread
// wrote.
__global__ void __launch_bounds__(256) k_walk_deep_write(WalkArgs a, const Cnt3 *__restrict__ tile_off)
{
    if (a.sz->overflow) return;
    const int lane = threadIdx.x % 64;
    const int64_t nbig = a.sz->n_big;
    for (int64_t k = (int64_t)blockIdx.x * 4 + threadIdx.x / 64; k < nbig; k += (int64_t)gridDim.x * 4) {
        const Cont ct = a.big[k];
        const int64_t q = ct.q;
        const int64_t t0 = q / kWalkThreads * kWalkThreads;
        int64_t sp = 0, sm = 0, sd = 0;
        for (int64_t i = t0 + lane; i < q; i += 64) {
            sp += a.pcnt[3 * i];
            sm += a.pcnt[3 * i + 1];
            sd += a.pcnt[3 * i + 2];
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            sp += __shfl_xor(sp, o, 64);
            sm += __shfl_xor(sm, o, 64);
            sd += __shfl_xor(sd, o, 64);
        }
        const Cnt3 to = tile_off[q / kWalkThreads];
        const int64_t qp = a.sz->p2p_base + to.p + sp + ct.np, qm = a.sz->m2l_base + to.m + sm + ct.nm;
        const int64_t qd = to.d + sd + ct.nd;
        const Target t = a.tgt[ct.j];
        const int32_t entry = a.pend[t.pbeg + (q - a.pair_off[ct.j])];
        uint32_t np = 0, nm = 0, nd = 0;
        wave_traverse<true>(a, t, ct.r, ct.rend, entry_shift(entry), lane, np, nm, nd, qp, qm, qd,
                            a.pcnt[3 * q] - ct.np, a.pcnt[3 * q + 1] - ct.nm, a.pcnt[3 * q + 2] - ct.nd);
    }
}

// Pending entries of each target of a depth: its parent's down list, or at the start
// depth the start list (start_target < 0: target j is the root of box j and takes
// entries j * nstart .. ; else start_target alone takes all nstart entries); and the
// target's record.  Targets without pending entries get empty spans here (the write
// pass sets the others).
__global__ void k_walk_prep(WalkArgs a, int start, int start_target, int nstart, int64_t *__restrict__ cnt)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int ntgt = ntgt_of(a);
    if (j == 0) {
        a.sz->n_big = 0;  // the count pass lists this depth's continued traversals
        if (ntgt_all(a) > a.tgt_cap) a.sz->overflow |= kOvfTargets;
    }
    if (j > a.tgt_cap) return;
    if (j >= ntgt || a.sz->overflow) {
        cnt[j] = 0;  // zeros up to the bound: the scan's entry ntgt is the total
        return;
    }
    const int X = tgt_of(a)[j];
    const int nleaf = a.sz->nleaf;
    const CellInfo I = a.info[X];
    int64_t b = 0, e = 0;
    if (start) {
        if (start_target < 0) {
            b = (int64_t)j * nstart;
            e = b + nstart;
        } else if (X == start_target) {
            e = nstart;
        }
    } else {
        b = a.dl_span[2 * (size_t)(I.parent - nleaf)];
        e = a.dl_span[2 * (size_t)(I.parent - nleaf) + 1];
    }
    Target t;
    t.g = a.rec[I.dfs].g;
    t.L = a.bx.len_of(X);
    t.pbeg = b;
    t.X = X;
    t.rank = I.dfs;
    t.leaf = X < nleaf;
    t.son0r = t.leaf ? -1 : a.info[I.son0].dfs;
    t.son1r = t.leaf ? -1 : a.info[I.son1].dfs;
    t.pad = 0;
    a.tgt[j] = t;
    cnt[j] = e - b;
    if (X < nleaf) {
        a.p2p_span[2 * (size_t)X] = a.p2p_span[2 * (size_t)X + 1] = 0;
    } else {
        a.dl_span[2 * (size_t)(X - nleaf)] = a.dl_span[2 * (size_t)(X - nleaf) + 1] = 0;
    }
    a.m2l_span[2 * (size_t)X] = a.m2l_span[2 * (size_t)X + 1] = 0;
}

// The depth's totals (P2P, M2L and down entries: the tile scan's last entry) against the
// capacities: the list bases move past the previous depth's entries, and what would not
// fit flags the overflow, after which nothing more is written.
__global__ void k_walk_step(DevSizes *sz, const Cnt3 *__restrict__ tot, int64_t p2p_cap, int64_t m2l_cap,
                            int64_t dl_cap, int64_t pair_cap)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    sz->p2p_base = sz->n_p2p;
    sz->m2l_base = sz->n_m2l;
    const Cnt3 t = *tot;
    int ovf = 0;
    if (sz->n_p2p + t.p > p2p_cap) ovf |= kOvfP2P;
    if (sz->n_m2l + t.m > m2l_cap) ovf |= kOvfM2L;
    if (t.d > dl_cap) ovf |= kOvfDown;
    if (2 * t.d > pair_cap) ovf |= kOvfPairs;
    if (ovf || sz->overflow) {
        sz->overflow |= ovf;
        sz->dl_total = 0;
        return;
    }
    sz->n_p2p += t.p;
    sz->n_m2l += t.m;
    sz->dl_total = t.d;
}

// walk start: empty lists, bases at zero
__global__ void k_walk_begin(DevSizes *sz)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    sz->n_p2p = sz->n_m2l = sz->p2p_base = sz->m2l_base = 0;
    sz->dl_total = 0;
}

// after the last depth of a planned walk: down entries left over mean deeper targets
// that were never visited
__global__ void k_walk_end(DevSizes *sz)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (sz->dl_total > 0) sz->overflow |= kOvfDepths;
}

// cells by depth-first rank: geometry, cell id and subtree size, read by the traversals
__global__ void k_walk_records(const DevSizes *__restrict__ sz, const CellInfo *__restrict__ info,
                               const CellGeom *__restrict__ geom, WalkRec *__restrict__ rec)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (sz->overflow || c >= sz->ncell) return;
    const CellInfo I = info[c];
    WalkRec R;
    R.g = geom[c];
    R.cell = c;
    R.dsize = I.dsize;
    R.kid[0] = R.kid[1] = -1;
    if (I.son0 >= 0) {
        const bool first0 = info[I.son0].dfs == I.dfs + 1;
        R.kid[0] = first0 ? I.son0 : I.son1;
        R.kid[1] = first0 ? I.son1 : I.son0;
    }
    rec[I.dfs] = R;
}

// every span empty (cells a walk from a given pair never reaches keep no entries); n is
// the launch bound of the cells
__global__ void k_clear_spans(int n, int64_t *__restrict__ span)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * (int64_t)n) span[i] = 0;
}

// start list: (root of box b, shift) for every box and shift, or (source, shift)
__global__ void k_start_list(int32_t *f, const CellInfo *__restrict__ info, const int *__restrict__ box_root, int nbox,
                             int B, int shift_lo, int shift_hi)
{
    const int nshift = shift_hi - shift_lo;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (box_root ? nbox : 1) * nshift) return;
    const int s = shift_lo + i % nshift;
    f[i] = pack_entry(info[box_root ? box_root[i / nshift] : B].dfs, s);  // pending entries name ranks
}

// ---------------------------------------------------------------------------- host side

// grow to `need` bytes keeping the first `keep` bytes
static void *grow_keep(DevBuf &b, size_t need, size_t keep, hipStream_t st)
{
    if (need <= b.cap) return b.p;
    const size_t want = arena_active() ? need : need + need / 2;
    void *nw = dev_alloc(want);
    if (b.p && keep) PH_CHECK(hipMemcpyAsync(nw, b.p, keep, hipMemcpyDeviceToDevice, st));
    PH_CHECK(hipStreamSynchronize(st));
    dev_free(b.p);
    b.p = nw;
    b.cap = want;
    return nw;
}

// The walk from the pair (start_target, start_source) over the shifts [shift_lo, shift_hi),
// or from every box's (root, root) when start_target < 0.  Per depth: pending counts and
// their scan, the count pass, the scan of its tile sums (the host reads the depth's three
// totals to size the outputs), the write pass.
void run_walk_from(Ctx &c, int shift_lo, int shift_hi, int start_target, int start_source)
{
    hipStream_t st = c.stream;
    const bool planned = c.plan.active;
    const int nlb = nleaf_bound(c), ncb = ncell_bound(c), nnb = nnode_bound(c);
    if (ncb >= (1 << 27)) throw std::runtime_error("more than 2^27 cells");
    need(c.cell_box, sizeof(int) * (size_t)ncb, "cell boxes");
    need(c.walk_tgt, sizeof(int) * (size_t)ncb, "walk targets");
    const bool forest = start_target < 0;
    if (planned && !forest) throw std::runtime_error("a planned walk starts from the roots");
    // depths of the walk: the tree's (ordinary) or the plan's
    const int ndepth = planned ? c.plan.depths : (int)c.tgt_off.size() - 1;

    DevSizes *sz = dptr<DevSizes>(c.sizes);
    WalkArgs a{};
    a.info = dptr<CellInfo>(c.info);
    a.sz = sz;
    a.theta = c.prm.theta;
    a.theta2 = c.prm.theta * c.prm.theta;
    a.rc = c.prm.rc;
    a.rc2 = c.prm.rc * c.prm.rc;
    a.bx = box_view(c);
    a.walk_tgt = dptr<int>(c.walk_tgt);
    WalkRec *rec = static_cast<WalkRec *>(ctx_buf(c, c.walk_rec, sizeof(WalkRec) * (size_t)ncb, "walk records"));
    k_walk_records<<<(ncb + 255) / 256, 256, 0, st>>>(sz, dptr<CellInfo>(c.info), dptr<CellGeom>(c.geom), rec);
    a.rec = rec;
    a.p2p_span = static_cast<int64_t *>(ctx_buf(c, c.p2p_off, sizeof(int64_t) * 2 * (size_t)std::max(1, nlb), "P2P list"));
    a.m2l_span = static_cast<int64_t *>(ctx_buf(c, c.m2l_off, sizeof(int64_t) * 2 * (size_t)ncb, "M2L list"));
    a.dl_span = static_cast<int64_t *>(ctx_buf(c, c.dl_span, sizeof(int64_t) * 2 * (size_t)std::max(1, nnb), "down lists"));
    k_clear_spans<<<(2 * nlb + 255) / 256 + 1, 256, 0, st>>>(nlb, a.p2p_span);
    k_clear_spans<<<(2 * ncb + 255) / 256 + 1, 256, 0, st>>>(ncb, a.m2l_span);
    k_walk_begin<<<1, 1, 0, st>>>(sz);

    // start: every box's root against itself, or the given pair; shifts [lo, hi)
    const int nshift = shift_hi - shift_lo;
    const int64_t n0 = (int64_t)(forest ? c.nbox : 1) * nshift;
    if (forest) need(c.d_box_root, sizeof(int) * (size_t)c.nbox, "box roots");
    int32_t *slist = static_cast<int32_t *>(ctx_buf(c, c.start_list, sizeof(int32_t) * (size_t)std::max<int64_t>(n0, 1),
                                                    "walk start"));
    k_start_list<<<(int)((n0 + 255) / 256), 256, 0, st>>>(slist, dptr<CellInfo>(c.info),
                                                          forest ? dptr<int>(c.d_box_root) : nullptr, c.nbox,
                                                          start_source, shift_lo, shift_hi);
    int d0 = 0;
    if (!forest) {
        CellInfo I;
        PH_CHECK(hipMemcpyAsync(&I, dptr<CellInfo>(c.info) + start_target, sizeof I, hipMemcpyDeviceToHost, st));
        PH_CHECK(hipStreamSynchronize(st));
        d0 = I.depth;
    }

    int64_t *pin = reinterpret_cast<int64_t *>(c.pinned.p) + 16;
    int64_t npair = n0, n_p2p = 0, n_m2l = 0;  // host copies (ordinary evaluation)
    int64_t max_pairs = n0, max_dl = 0;
    int depths = 0;
    const int32_t *pend = slist;
    DevBuf *dl_out = &c.dl_b, *dl_other = &c.dl_a;
    for (int d = d0; d < ndepth && (planned || npair > 0); d++) {
        // launch bounds of this depth: targets and pairs
        const int64_t tcap = planned ? c.plan.tgt_cap : c.tgt_off[d + 1] - c.tgt_off[d];
        const int64_t pcap = planned ? c.plan.pair_cap : npair;
        a.d = d;
        a.tgt_cap = tcap;
        a.pend = pend;
        int64_t *poff = static_cast<int64_t *>(ctx_buf(c, c.pair_off, sizeof(int64_t) * (size_t)(tcap + 1), "walk"));
        if (pcap >= ((int64_t)1 << 31)) throw std::runtime_error("more than 2^31 walk pairs at one depth");
        int64_t *cnt = static_cast<int64_t *>(ctx_buf(c, c.pend_span, sizeof(int64_t) * (size_t)(tcap + 1), "walk"));
        a.pair_off = poff;
        a.tgt = static_cast<Target *>(ctx_buf(c, c.walk_trec, sizeof(Target) * (size_t)std::max<int64_t>(tcap, 1), "walk"));
        k_walk_prep<<<(int)((tcap + 1 + 255) / 256), 256, 0, st>>>(a, d == d0, forest ? -1 : start_target, nshift, cnt);
        {
            size_t bytes = 0;
            PH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt, poff, (int)(tcap + 1), st));
            void *tmp = scan_storage(c, bytes);
            PH_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnt, poff, (int)(tcap + 1), st));
        }
        const int ntile = (int)((pcap + kWalkThreads - 1) / kWalkThreads);  // >= 1: pcap >= 1
        a.pcnt = static_cast<uint32_t *>(ctx_buf(c, c.pair_cnt, sizeof(uint32_t) * 3 * (size_t)std::max<int64_t>(pcap, 1),
                                                 "walk pairs"));
        a.npair_cap = pcap;
        a.big = static_cast<Cont *>(ctx_buf(c, c.big_pairs, sizeof(Cont) * (size_t)std::max<int64_t>(pcap, 1),
                                            "walk pairs"));
        int32_t *tile_tgt = static_cast<int32_t *>(ctx_buf(c, c.walk_tiles, sizeof(int32_t) * (size_t)ntile, "walk tiles"));
        a.tile_tgt = tile_tgt;
        k_tile_targets<<<(int)((std::max<int64_t>(tcap, 1) + 255) / 256), 256, 0, st>>>(a, tile_tgt);
        Cnt3 *tile_cnt = static_cast<Cnt3 *>(ctx_buf(c, c.cnt_tmp, sizeof(Cnt3) * 2 * (size_t)(ntile + 1), "walk tiles"));
        Cnt3 *tile_off = tile_cnt + (ntile + 1);
        k_walk_count<<<ntile, kWalkThreads, 0, st>>>(a, tile_cnt, ntile);
        const int deep_blocks = (int)std::max<int64_t>(1, std::min<int64_t>(kDeepWaves / 4, (pcap + 3) / 4));
        k_walk_deep_count<<<deep_blocks, 256, 0, st>>>(a, tile_cnt);
        {
            size_t bytes = 0;
            PH_CHECK(hipcub::DeviceScan::ExclusiveScan(nullptr, bytes, tile_cnt, tile_off, AddCnt3(), Cnt3{0, 0, 0},
                                                       ntile + 1, st));
            void *tmp = scan_storage(c, bytes);
            PH_CHECK(hipcub::DeviceScan::ExclusiveScan(tmp, bytes, tile_cnt, tile_off, AddCnt3(), Cnt3{0, 0, 0},
                                                       ntile + 1, st));
        }
        int64_t p2p_cap, m2l_cap, dl_cap, next_pairs;
        if (planned) {
            p2p_cap = c.plan.p2p_cap;
            m2l_cap = c.plan.m2l_cap;
            dl_cap = c.plan.dl_cap;
            next_pairs = c.plan.pair_cap;
        } else {
            // the depth's totals size its outputs
            PH_CHECK(hipMemcpyAsync(pin, tile_off + ntile, sizeof(Cnt3), hipMemcpyDeviceToHost, st));
            PH_CHECK(hipEventRecord(c.counts_ready, st));
            PH_CHECK(hipEventSynchronize(c.counts_ready));
            const int64_t add_p2p = pin[0], add_m2l = pin[1], add_dl = pin[2];
            grow_keep(c.p2p_src, 4 * (size_t)std::max<int64_t>(n_p2p + add_p2p, 1), 4 * (size_t)n_p2p, st);
            grow_keep(c.m2l_src, 4 * (size_t)std::max<int64_t>(n_m2l + add_m2l, 1), 4 * (size_t)n_m2l, st);
            // the pending list (the other ping-pong buffer) is read by this write: only the
            // output buffer may move
            grow_keep(*dl_out, 4 * (size_t)std::max<int64_t>(add_dl, 1), 0, st);
            n_p2p += add_p2p;
            n_m2l += add_m2l;
            npair = 2 * add_dl;  // the children of this depth's nodes read their parents' lists
            max_pairs = std::max(max_pairs, npair);
            max_dl = std::max(max_dl, add_dl);
            depths = d + 1;
            p2p_cap = c.p2p_src.cap / 4;
            m2l_cap = c.m2l_src.cap / 4;
            dl_cap = dl_out->cap / 4;
            next_pairs = npair;
        }
        // at the plan's capacities (allocated by the first planned evaluation, checked by the
        // others); an ordinary evaluation grew them above
        ctx_buf(c, c.p2p_src, sizeof(int32_t) * (size_t)std::max<int64_t>(p2p_cap, 1), "P2P list");
        ctx_buf(c, c.m2l_src, sizeof(int32_t) * (size_t)std::max<int64_t>(m2l_cap, 1), "M2L list");
        ctx_buf(c, *dl_out, sizeof(int32_t) * (size_t)std::max<int64_t>(dl_cap, 1), "down lists");
        a.p2p_out = dptr<int32_t>(c.p2p_src);
        a.m2l_out = dptr<int32_t>(c.m2l_src);
        a.dl_out = dptr<int32_t>(*dl_out);
        k_walk_step<<<1, 1, 0, st>>>(sz, tile_off + ntile, p2p_cap, m2l_cap, dl_cap, next_pairs);
        k_walk_write<<<ntile, kWalkThreads, 0, st>>>(a, tile_off);
        k_walk_deep_write<<<deep_blocks, 256, 0, st>>>(a, tile_off);
        pend = a.dl_out;
        std::swap(dl_out, dl_other);
    }
    if (planned) {
        k_walk_end<<<1, 1, 0, st>>>(sz);
    } else {
        c.n_p2p = n_p2p;
        c.n_m2l = n_m2l;
        if (forest) {
            Measure &m = c.meas;
            m.depths = depths;
            m.pairs = max_pairs;
            m.dl = max_dl;
            m.p2p = n_p2p;
            m.m2l = n_m2l;
        }
    }
    c.lists_valid = true;
    PH_CHECK(hipGetLastError());
}

void run_walk(Ctx &c, int shift_lo, int shift_hi) { run_walk_from(c, shift_lo, shift_hi, -1, -1); }

}  // namespace photons
