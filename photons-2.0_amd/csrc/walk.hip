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
// paired with itself passes its two children; so each node writes one down record per
// source for its children, and the pairs of a depth are the previous depth's down
// records, each against both children it names (pair g the first, pair n + g the
// second: one target's pairs are contiguous and in its parent's order).  At each depth
// every pair (one thread) walks the source subtree of its entry in depth-first order
// without a stack: at source rank r the rule either descends (opens the source: r + 1)
// or settles the whole subtree (P2P, M2L, pass down or nothing: r + subtree size); a
// traversal longer than a few ranks is continued by a group of lanes.  Two passes per
// depth (count, then write at scanned offsets).  Every list is written grouped by target
// in the reference's order (shift, depth-first rank): no sort.

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

// by cell: what the rule needs of a target, in one 80-byte record
struct alignas(16) TgtRec {
    CellGeom g;
    double L;         // its box length (shifts, the image prune's domain box)
    int32_t rank;     // depth-first rank
    int32_t son[2];   // children (nodes)
    int32_t sonr[2];  // their ranks
    int32_t leaf;
};
static_assert(sizeof(TgtRec) == 80, "one 80-byte record per cell");

// A down record: a source passed on to the two children t0, t1 of a node (at the start
// depth t0 is the target itself and t1 = -1).  The entry holds the source's shift and its
// depth-first rank, where a traversal of its subtree starts; a leaf source, which needs
// no traversal, is named by its cell instead and flagged in t0.
struct DownRec {
    int32_t entry, t0, t1;
};
constexpr int32_t kSrcLeaf = 1 << 30;
__device__ __forceinline__ bool src_leaf(const DownRec &D) { return (D.t0 & kSrcLeaf) != 0; }
__device__ __forceinline__ int owner_of(const DownRec &D) { return D.t0 & kCellMask; }  // names the parent
__device__ __forceinline__ int target_of(const DownRec &D, int c) { return c ? D.t1 : (D.t0 & kCellMask); }

// A traversal continued by a group of lanes: pair q goes on at rank r (< rend, the end of
// its source subtree) after the outputs its thread found (`part`, packed); the count
// passes add up the pair's outputs (tp, tm, td), the write pass leaves where the rest
// goes (qp, qm, qd).
struct alignas(16) Cont {
    int32_t q, r, rend, pad;
    uint64_t part;
    uint32_t tp, tm, td, pad2;
    int64_t qp, qm, qd;
};
static_assert(sizeof(Cont) == 64, "one 64-byte record per continued traversal");

// a listed pair (a node source): the pair, its target and its source entry
struct TravRec {
    int32_t q, T, entry;
};

struct WalkArgs {
    const WalkRec *rec;       // by depth-first rank
    const TgtRec *trec;       // by cell
    DevSizes *sz;             // cell counts, records of the depth, list bases, overflow
    double theta, theta2, rc, rc2;
    const DownRec *din;       // this depth's records (sz->nrec of them, at most rec_cap)
    int64_t rec_cap;          // launch bound of the records; the pairs are twice as many
    uint64_t *pcnt;           // per pair: its outputs (packed) or its continuation
    TravRec *tlist;           // per tile, its pairs with a node source (a traversal): tlist[tile * 256 + i]
    int32_t *tcount;          // per tile, how many
    uint32_t *qoff;           // [3 per pair]: a listed pair's output offsets from its tile's base
    int trav_waves;           // wavefronts of the traversal passes (a multiple of kTileWaves)
    int trav_split;           // slices per super-chunk handed to the traversal wavefronts
    int visits;               // steps of a traversal on its lane before it is continued
    Cont *cont;               // continued traversals: kContPerWave slots per traversal wavefront
    int32_t *ccount;          // per traversal wavefront, its slots in use
    const int32_t *coff;      // their exclusive scan
    int32_t *cidx;            // the slots in use, in order
    // outputs of the write pass
    int32_t *p2p_out, *m2l_out;  // list bases: sz->p2p_base, sz->m2l_base
    DownRec *dl_out;             // the next depth's records
    int64_t *p2p_span, *m2l_span;  // per leaf / cell
};

// Entries of the walk's own records name the source by its depth-first rank (the
// traversal starts there without a cell lookup); the P2P and M2L lists name cells.
__device__ __forceinline__ int entry_rank(int32_t e) { return entry_cell(e); }

__device__ __forceinline__ int decide(const WalkArgs &a, const TgtRec &t, const WalkRec &R, int r, int sid,
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

constexpr int kTile = 256;       // pairs per tile, one wavefront each
constexpr int kTileWaves = 4;    // tiles (wavefronts) per block
constexpr int kWalkThreads = 64 * kTileWaves;

// A lane applies the rule to at most kThreadVisits ranks of a pair; a traversal that goes
// on (a target leaf beside a dense cluster opens one source node after another) is
// continued from the rank it reached by a group of kGroupLanes lanes, that many ranks at
// a time.  A whole wavefront per group loads 64 consecutive 64-byte records at once, where
// single lanes each chase their own subtree: at 256^3 groups of 8, 16, 32 and 64 lanes
// walked in 54, 48, 44 and 42 ms, and handing on after 32 steps rather than 64 (or 16)
// gained another half millisecond.
constexpr int kThreadVisits = 32;
// at a depth of few records (a small box, the first depths of any) the traversal passes
// run few wavefronts whose lanes each walk their records one dependent load at a time:
// there a traversal is handed on after kThreadVisitsFew steps (a 16^3 box's replay
// 1.45 -> 1.31 ms, 32^3 2.52 -> 2.13 ms; 8 steps: 1.35, 2.17 ms)
constexpr int kThreadVisitsFew = 4;
constexpr int kGroupLanes = 64;

// A pair's outputs as one 64-bit word: bit 0 clear, then down (19 bits), M2L (20), P2P
// (24); bit 0 set: the pair was continued, bits 1.. hold its continuation's slot.  A
// count past its field (a single pair's source subtree with 2^24 leaves) flags the
// evaluation (kOvfPairs) instead of wrapping.
__device__ __forceinline__ uint64_t pack_cnt(const WalkArgs &a, uint32_t p, uint32_t m, uint32_t d)
{
    if (p >= (1u << 24) || m >= (1u << 20) || d >= (1u << 19)) atomicOr(&a.sz->overflow, (int)kOvfPairs);
    return ((uint64_t)(p & 0xffffffu) << 40) | ((uint64_t)(m & 0xfffffu) << 20) | ((uint64_t)(d & 0x7ffffu) << 1);
}
__device__ __forceinline__ void unpack_cnt(uint64_t w, uint32_t &p, uint32_t &m, uint32_t &d)
{
    p = (uint32_t)(w >> 40);
    m = (uint32_t)(w >> 20) & 0xfffffu;
    d = (uint32_t)(w >> 1) & 0x7ffffu;
}

// the down record of source `src` (its cell if a leaf, else its rank) for t's children
__device__ __forceinline__ void write_down(const WalkArgs &a, int64_t at, int src, bool leaf, int sid, const TgtRec &t)
{
    a.dl_out[at] = DownRec{pack_entry(src, sid), t.son[0] | (leaf ? kSrcLeaf : 0), t.son[1]};
}

// One step of a pair's depth-first traversal of its source subtree: the rule at rank r
// (record R) of target t; returns the next rank (>= the subtree's end when done).
// WRITE = false counts the outputs; WRITE = true stores them from the given positions,
// never more than the count pass found (lp, lm, ld), so a store stays inside the ranges
// the scan gave this pair.
template <bool WRITE>
__device__ __forceinline__ int visit(const WalkArgs &a, const TgtRec &t, const WalkRec &R, int r, int sid, double sx,
                                     double sy, double sz, uint32_t &np, uint32_t &nm, uint32_t &nd, int64_t qp,
                                     int64_t qm, int64_t qd, uint32_t lp, uint32_t lm, uint32_t ld)
{
    const int act = decide(a, t, R, r, sid, sx, sy, sz);
    const bool isp = act == kActP2P, ism = act == kActM2L, isd = act == kActOpenA, iss = act == kActSelfNode;
    if (WRITE) {
        if (isp && np < lp) a.p2p_out[qp + np] = pack_entry(R.cell, sid);
        if (ism && nm < lm) a.m2l_out[qm + nm] = pack_entry(R.cell, sid);
        if (isd && nd < ld) {
            const bool lf = R.dsize == 1;
            write_down(a, qd + nd, lf ? R.cell : r, lf, sid, t);
        }
        if (iss && nd + 1 < ld) {
            const int nleaf = a.sz->nleaf;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const bool lf = t.son[k] < nleaf;
                write_down(a, qd + nd + k, lf ? t.son[k] : t.sonr[k], lf, 0, t);
            }
        }
    }
    // the counts by arithmetic: increments under a branch each were merged by the
    // compiler into one increment through a selected address, which put the three
    // counters in scratch memory
    np += isp ? 1u : 0u;
    nm += ism ? 1u : 0u;
    nd += isd ? 1u : (iss ? 2u : 0u);
    if (act != kActOpenB) return r + R.dsize;
    if (!t.leaf) return r + 1;
    // a target leaf against a leaf takes P2P whatever the geometry: the opened node's leaf
    // children are emitted here, in rank order, without a step of their own
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
    return r;
}

// The rest of a traversal, ranks [r, rend), by a group of G lanes (gl: the lane in the
// group): lane gl takes rank carry + gl and applies the rule; a rank is visited unless an
// earlier rank of the chunk settled a subtree containing it (an exclusive prefix max of
// the ranks each lane accounts for: a settled rank its whole subtree, an opened one
// itself; a rank inside a settled subtree accounts for no more than that subtree, and
// every subtree settled before r ends by r); the visited ranks' outputs are emitted in
// rank order, as the single-thread walk emits them.  A pair's own target is only ever met
// at its first rank (self pair), which the thread has taken, so no rank here passes two
// children down.  np, nm, nd: group-uniform counts; WRITE stores from qp, qm, qd, never
// past lp, lm, ld.  Every lane of the group runs this together.
template <bool WRITE, int G>
__device__ __forceinline__ void group_traverse(const WalkArgs &a, const TgtRec &t, int r, int rend, int sid, int gl,
                                               uint32_t &np, uint32_t &nm, uint32_t &nd, int64_t qp = 0,
                                               int64_t qm = 0, int64_t qd = 0, uint32_t lp = 0, uint32_t lm = 0,
                                               uint32_t ld = 0)
{
    static_assert(G <= 64 && (G & (G - 1)) == 0, "a power-of-two group of at most 64 lanes");
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (sid != 0) shift_of(sid, t.L, sx, sy, sz);
    const int gshift = (int)(threadIdx.x % 64) / G * G;  // the group's first lane in the wavefront
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const uint64_t below = (1ull << gl) - 1ull;
    int carry = r;
    while (carry < rend) {
        const int s = carry + gl;
        int act = kActNone, cov = s + 1, v = 0;
        bool lf = false;
        if (s < rend) {
            const WalkRec R = a.rec[s];
            v = R.cell;
            lf = R.dsize == 1;
            act = decide(a, t, R, s, sid, sx, sy, sz);
            cov = act == kActOpenB ? s + 1 : s + R.dsize;
        }
        int incl = cov;
#pragma unroll
        for (int d = 1; d < G; d <<= 1) {
            const int o = __shfl_up(incl, d, G);
            if (gl >= d) incl = max(incl, o);
        }
        int excl = __shfl_up(incl, 1, G);
        excl = gl == 0 ? carry : max(excl, carry);
        const bool visited = s < rend && s >= excl;
        const bool isp = visited && act == kActP2P, ism = visited && act == kActM2L;
        const bool isd = visited && act == kActOpenA;
        const uint64_t bp = (__ballot(isp) >> gshift) & gmask;
        const uint64_t bm = (__ballot(ism) >> gshift) & gmask;
        const uint64_t bd = (__ballot(isd) >> gshift) & gmask;
        if (WRITE) {
            const uint32_t ip = np + __popcll(bp & below), im = nm + __popcll(bm & below);
            const uint32_t id = nd + __popcll(bd & below);
            if (isp && ip < lp) a.p2p_out[qp + ip] = pack_entry(v, sid);
            if (ism && im < lm) a.m2l_out[qm + im] = pack_entry(v, sid);
            if (isd && id < ld) write_down(a, qd + id, lf ? v : s, lf, sid, t);
        }
        np += __popcll(bp);
        nm += __popcll(bm);
        nd += __popcll(bd);
        carry = max(carry + G, __shfl(incl, G - 1, G));
    }
}

// The pairs of a depth: record g yields pair g (target t0) and pair npad + g (target t1;
// npad: the records rounded up to whole tiles), so the pairs of one target are contiguous
// and in the order of its parent's down list.  Output positions follow this pair order
// (tile sums scanned in tile order); which wavefront runs which tile is free, and is
// chosen for the caches: workgroups go to the 8 XCDs in turn (block b to XCD b % 8), each
// with its own L2, so XCD x takes one contiguous run of tiles, and within it the tiles of
// both children of the same records one after the other (their sources are the same).
__device__ __forceinline__ int64_t nrec_of(const WalkArgs &a) { return min(a.sz->nrec, a.rec_cap); }
constexpr int kXcds = 8;

struct TileMap {
    int64_t n, npad, H;  // records, records rounded up to tiles, tiles per child stream
    int64_t tile;        // this wavefront's tile (pair order), or -1: no tile
    int64_t g0;          // the first record of the tile
    int c;               // the child stream
    int npair;           // pairs of the tile with a record
};

__device__ __forceinline__ TileMap tile_map(const WalkArgs &a)
{
    TileMap m;
    m.n = nrec_of(a);
    m.H = (m.n + kTile - 1) / kTile;
    m.npad = m.H * kTile;
    const int64_t per = (2 * m.H + kXcds * kTileWaves - 1) / (kXcds * kTileWaves);  // blocks per XCD
    const int x = blockIdx.x % kXcds;
    const int64_t i = blockIdx.x / kXcds;
    const int64_t p = ((int64_t)x * per + i) * kTileWaves + threadIdx.x / 64;
    m.tile = -1;
    m.g0 = 0;
    m.c = 0;
    m.npair = 0;
    if (i < per && p < 2 * m.H) {
        m.c = (int)(p & 1);
        m.g0 = (p >> 1) * kTile;
        m.tile = m.c * m.H + (p >> 1);
        m.npair = (int)min<int64_t>(kTile, m.n - m.g0);
    }
    return m;
}

// blocks for ntile tiles of pairs: a whole number per XCD
static inline int walk_blocks(int64_t ntile)
{
    const int64_t per = (ntile + kXcds * kTileWaves - 1) / (kXcds * kTileWaves);
    return (int)(per * kXcds);
}

struct Cnt3 {
    int64_t p, m, d;
};
struct AddCnt3 {
    __host__ __device__ Cnt3 operator()(const Cnt3 &x, const Cnt3 &y) const { return {x.p + y.p, x.m + y.m, x.d + y.d}; }
};

__device__ __forceinline__ uint64_t wave_sum(uint64_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// A lane's current pair: target record, source traversal state and outputs so far.
struct LanePair {
    int64_t q;           // the pair, -1: none
    TgtRec t;
    WalkRec R;           // the record at rank r
    int r, rend, sid, visits;
    double sx, sy, sz;   // the source's shift
    uint32_t np, nm, nd;
};

// start pair q on this lane (target T >= 0, a node source: entry names its rank)
__device__ __forceinline__ void lane_start(const WalkArgs &a, LanePair &L, int64_t q, int T, int32_t entry)
{
    L.q = q;
    L.t = a.trec[T];
    L.r = entry_rank(entry);
    L.sid = entry_shift(entry);
    L.R = a.rec[L.r];
    L.rend = L.r + L.R.dsize;
    L.sx = L.sy = L.sz = 0.0;
    if (L.sid != 0) shift_of(L.sid, L.t.L, L.sx, L.sy, L.sz);
    L.np = L.nm = L.nd = 0;
    L.visits = 0;
}

// Two pairs whose source is a leaf (use[i]; target T[i] >= 0), one rule each, no
// traversal: a target leaf takes P2P (leaf-leaf pairs always do); a node target an M2L
// entry, a down record or nothing, by the rule on the two cells.  The loads of both pairs
// go out before either is decided.  Counts only: the write pass stores each pair's one
// output from its count.
__device__ __forceinline__ void leaf_source_pairs(const WalkArgs &a, int nleaf, const bool (&use)[2],
                                                  const int (&T)[2], const DownRec (&D)[2], uint32_t (&np)[2],
                                                  uint32_t (&nm)[2], uint32_t (&nd)[2])
{
    TgtRec t[2];
    CellGeom g[2];
    bool node[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        node[i] = use[i] && T[i] >= nleaf;
        t[i] = a.trec[node[i] ? T[i] : 0];
        g[i] = a.trec[node[i] ? entry_cell(D[i].entry) : 0].g;
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
        int act = kActNone;
        if (node[i]) {
            const int sid = entry_shift(D[i].entry);
            WalkRec R;
            R.g = g[i];
            R.cell = entry_cell(D[i].entry);
            R.dsize = 1;
            R.kid[0] = R.kid[1] = -1;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            if (sid != 0) shift_of(sid, t[i].L, sx, sy, sz);
            act = decide(a, t[i], R, -1, sid, sx, sy, sz);  // rank -1: not the target itself
        }
        // (selects: stores under branches were merged into one store through a selected
        // address, which put the counters in scratch memory)
        np[i] = use[i] && !node[i] ? 1u : 0u;
        nm[i] = act == kActM2L ? 1u : 0u;
        nd[i] = act == kActOpenA ? 1u : 0u;
    }
}

// the lanes of one wavefront see each other's LDS stores from here on
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the target and the source entry of pair q (npad: records rounded up to tiles)
__device__ __forceinline__ void pair_of(const WalkArgs &a, int64_t npad, int64_t q, int &T, int32_t &entry)
{
    const DownRec D = a.din[q < npad ? q : q - npad];
    T = target_of(D, q < npad ? 0 : 1);
    entry = D.entry;
}
__device__ __forceinline__ int64_t npad_of(const WalkArgs &a)
{
    return (nrec_of(a) + kTile - 1) / kTile * kTile;
}

// the traversal passes' grid: persistent wavefronts, each with its own continuation slots
constexpr int kTravWavesMax = 16384;
constexpr int kTravWavesFill = 2048;  // 8 wavefronts per CU
constexpr int kTravSplitMax = 16;     // slices of at least 64 of a super-chunk's 1024 positions
constexpr int kContPerWave = 256;

// a pair's outputs: its packed counts, or its continuation's totals
__device__ __forceinline__ void pair_counts(const WalkArgs &a, uint64_t w, uint32_t &p, uint32_t &m, uint32_t &d)
{
    if (w & 1u) {
        const Cont &ct = a.cont[w >> 1];  // a slot in use
        p = ct.tp;
        m = ct.tm;
        d = ct.td;
    } else {
        unpack_cnt(w, p, m, d);
    }
}

// Count pass, pairs with a leaf source (most of them): one wavefront per tile of 256
// pairs, two pairs per lane at a time, all lanes in step; their packed counts.  The pairs
// with a node source (a traversal) are listed per tile instead, for k_walk_count_trav.
__global__ void __launch_bounds__(kWalkThreads) k_walk_count(WalkArgs a)
{
    const int lane = threadIdx.x % 64;
    const uint64_t below = (1ull << lane) - 1ull;
    const TileMap tm = tile_map(a);
    if (tm.tile < 0) return;  // the same for the whole wavefront
    const int64_t q0 = tm.tile * kTile;
    const int nleaf = a.sz->nleaf;
    bool trav[kTile / 64];
    uint64_t bt[kTile / 64];
    int ntrav = 0;
#pragma unroll
    for (int k = 0; k < kTile / 64; k += 2) {
        int j[2], T[2];
        DownRec D[2];
        bool use[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            j[i] = lane + 64 * (k + i);
            D[i] = j[i] < tm.npair ? a.din[tm.g0 + j[i]] : DownRec{0, 0, -1};
            T[i] = j[i] < tm.npair ? target_of(D[i], tm.c) : -1;
            use[i] = T[i] >= 0 && src_leaf(D[i]);
            trav[k + i] = T[i] >= 0 && !src_leaf(D[i]);
            bt[k + i] = __ballot(trav[k + i]);
            ntrav += __popcll(bt[k + i]);
        }
        uint32_t np[2], nm[2], nd[2];
        leaf_source_pairs(a, nleaf, use, T, D, np, nm, nd);
#pragma unroll
        for (int i = 0; i < 2; i++)
            if (j[i] < tm.npair && !trav[k + i]) a.pcnt[q0 + j[i]] = pack_cnt(a, np[i], nm[i], nd[i]);
    }
    int at = 0;
#pragma unroll
    for (int k = 0; k < kTile / 64; k++) {
        if (trav[k]) {
            const DownRec D = a.din[tm.g0 + lane + 64 * k];
            a.tlist[q0 + at + __popcll(bt[k] & below)] = TravRec{(int32_t)(q0 + lane + 64 * k), target_of(D, tm.c), D.entry};
        }
        at += __popcll(bt[k]);
    }
    if (lane == 0) a.tcount[tm.tile] = ntrav;
}

// The listed pairs of a depth, by super-chunks of kSuper tiles: two consecutive tiles of
// the first child stream and the same two of the second.  Both streams hold the same
// records, so a record with a node source is listed in both (the two children of its
// target against the same source subtree), at the same place: the chunk hands them out
// interleaved, so that the two traversals of one subtree start side by side in a
// wavefront and their record loads coincide where the two children decide alike.  A
// wavefront walks its super-chunks' lists as one sequence.
constexpr int kSuper = 4;

struct SuperChunk {
    int64_t k0;           // first tile of the first stream (the second's: H + k0)
    int64_t H;            // tiles per stream
    int n0[2], m[2];      // per tile pair: listed in the first stream, min of both streams
    int pre[3];           // listed pairs before each tile pair of the chunk
};

__device__ __forceinline__ SuperChunk super_chunk(const WalkArgs &a, int64_t s, int64_t ntile_dev)
{
    SuperChunk sc;
    sc.H = ntile_dev / 2;
    sc.k0 = 2 * s;
    const int lane = threadIdx.x % 64;
    // lanes 0, 1: the first stream's tiles k0, k0 + 1; lanes 2, 3: the second stream's
    int cnt = 0;
    if (lane < kSuper && sc.k0 + (lane & 1) < sc.H) cnt = a.tcount[(lane >> 1) * sc.H + sc.k0 + (lane & 1)];
    sc.pre[0] = 0;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        // the same in every lane: kept in scalar registers
        const int c0 = __builtin_amdgcn_readfirstlane(__shfl(cnt, j, 64));
        const int c1 = __builtin_amdgcn_readfirstlane(__shfl(cnt, 2 + j, 64));
        sc.n0[j] = c0;
        sc.m[j] = min(c0, c1);
        sc.pre[j + 1] = sc.pre[j] + c0 + c1;
    }
    return sc;
}

// the listed pair at position idx (< sc.pre[2]) of a super-chunk: within a tile pair,
// first and second stream alternate while both have pairs, then the longer one's rest
__device__ __forceinline__ TravRec super_pair(const WalkArgs &a, const SuperChunk &sc, int idx)
{
    // (selects, not indexing: an indexed array would live in scratch memory)
    const int j = idx >= sc.pre[1] ? 1 : 0;
    const int r = idx - (j ? sc.pre[1] : sc.pre[0]), m = j ? sc.m[1] : sc.m[0];
    int st, i;
    if (r < 2 * m) {
        st = r & 1;
        i = r >> 1;
    } else {
        st = (j ? sc.n0[1] : sc.n0[0]) > m ? 0 : 1;
        i = m + (r - 2 * m);
    }
    return a.tlist[(st * sc.H + sc.k0 + j) * kTile + i];
}

// A wavefront's stream of listed pairs: its units u = gw, gw + trav_waves, ...  in turn,
// without waiting at the end of one for the lanes still busy with its pairs.  A unit is
// one of trav_split equal slices of a super-chunk's positions (more than one at a depth
// with few records, so that its traversals spread over more wavefronts).
struct TravStream {
    int64_t u, nunit, ntile_dev;
    SuperChunk sc;
    int cn, next;  // end of the current unit's positions, next position handed out
    bool more;     // a current unit
};

__device__ __forceinline__ void stream_load(const WalkArgs &a, TravStream &ts)
{
    ts.more = ts.u < ts.nunit;
    ts.cn = ts.next = 0;
    if (ts.more) {
        const int split = a.trav_split;
        const int part = (int)(ts.u % split);
        ts.sc = super_chunk(a, ts.u / split, ts.ntile_dev);
        const int n = ts.sc.pre[2];
        ts.next = n * part / split;
        ts.cn = n * (part + 1) / split;
    }
}

__device__ __forceinline__ void stream_open(const WalkArgs &a, TravStream &ts, int64_t gw)
{
    ts.ntile_dev = 2 * (npad_of(a) / kTile);
    ts.nunit = (ts.ntile_dev / 2 + 1) / 2 * a.trav_split;  // super-chunks: pairs of tiles of the first stream
    ts.u = gw;
    stream_load(a, ts);
}

// the next unit while the current one is handed out (the same in every lane)
__device__ __forceinline__ void stream_advance(const WalkArgs &a, TravStream &ts)
{
    while (ts.more && ts.next >= ts.cn) {
        ts.u += a.trav_waves;
        stream_load(a, ts);
    }
}

// Count pass, the listed pairs (a node source): super-chunks per wavefront; lanes take the
// pairs in turn and a lane that finishes one takes the next, so a long traversal holds up
// one lane only; each step of the loop is one rule (one record load) per busy lane.  A
// traversal that reaches a.visits steps is handed to k_walk_deep_count through one
// of the wavefront's continuation slots; when they are used up, its lane goes on alone.
__global__ void __launch_bounds__(kWalkThreads) __attribute__((amdgpu_waves_per_eu(6))) k_walk_count_trav(WalkArgs a)
{
    const int lane = threadIdx.x % 64;
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t gw = (int64_t)blockIdx.x * kTileWaves + __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    Cont *slots = a.cont + gw * kContPerWave;
    int ncont = 0;  // slots in use (the same in every lane)
    TravStream ts;
    stream_open(a, ts, gw);
    LanePair L;
    L.q = -1;
    for (;;) {
        stream_advance(a, ts);
        if (!ts.more && !__ballot(L.q >= 0)) break;
        const bool want = L.q < 0;
        const uint64_t bw = __ballot(want);
        if (want) {
            const int idx = ts.next + __popcll(bw & below);
            if (idx < ts.cn) {
                const TravRec tr = super_pair(a, ts.sc, idx);
                lane_start(a, L, tr.q, tr.T, tr.entry);
            }
        }
        ts.next = min(ts.next + __popcll(bw), ts.cn);
        bool fin = false, cap = false;
        if (L.q >= 0) {
            L.r = visit<false>(a, L.t, L.R, L.r, L.sid, L.sx, L.sy, L.sz, L.np, L.nm, L.nd, 0, 0, 0, 0, 0, 0);
            L.visits++;
            if (L.r >= L.rend) fin = true;
            else if (L.visits == a.visits) cap = true;
        }
        const uint64_t bh = __ballot(cap);
        const bool room = ncont + __popcll(bh) <= kContPerWave;  // the same in every lane
        const bool hit = cap && room;
        if (hit) {
            const int k = ncont + __popcll(bh & below);
            Cont ct{};
            ct.q = (int32_t)L.q;
            ct.r = L.r;
            ct.rend = L.rend;
            ct.part = pack_cnt(a, L.np, L.nm, L.nd);
            slots[k] = ct;
            a.pcnt[L.q] = ((uint64_t)(gw * kContPerWave + k) << 1) | 1u;
        }
        if (room) ncont += __popcll(bh);
        if (fin) a.pcnt[L.q] = pack_cnt(a, L.np, L.nm, L.nd);
        if (fin || hit) L.q = -1;
        else if (L.q >= 0) L.R = a.rec[L.r];
    }
    if (lane == 0) a.ccount[gw] = ncont;
}

// the continuation slots in use, listed in order (coff: the scan of ccount)
__global__ void k_walk_cont_index(WalkArgs a)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)a.trav_waves * kContPerWave) return;
    const int64_t g = k / kContPerWave, i = k % kContPerWave;
    if (i < a.ccount[g]) a.cidx[a.coff[g] + i] = (int32_t)k;
}

__device__ __forceinline__ int64_t cont_total(const WalkArgs &a)
{
    return (int64_t)a.coff[a.trav_waves - 1] + a.ccount[a.trav_waves - 1];
}

// the continued traversals, kGroupLanes lanes each, over a fixed grid: their totals
constexpr int kDeepBlocks = 2048;

__global__ void __launch_bounds__(256) k_walk_deep_count(WalkArgs a)
{
    constexpr int G = kGroupLanes;
    const int gl = threadIdx.x % G;
    const int64_t npad = npad_of(a), ntot = cont_total(a);
    const int64_t ngroup = (int64_t)gridDim.x * (blockDim.x / G);
    for (int64_t m = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; m < ntot; m += ngroup) {
        const int64_t k = a.cidx[m];
        const Cont ct = a.cont[k];
        int T;
        int32_t entry;
        pair_of(a, npad, ct.q, T, entry);
        const TgtRec t = a.trec[T];
        uint32_t np = 0, nm = 0, nd = 0;
        group_traverse<false, G>(a, t, ct.r, ct.rend, entry_shift(entry), gl, np, nm, nd);
        if (gl == 0) {
            uint32_t p0, m0, d0;
            unpack_cnt(ct.part, p0, m0, d0);
            a.cont[k].tp = p0 + np;
            a.cont[k].tm = m0 + nm;
            a.cont[k].td = d0 + nd;
        }
    }
}

// per tile, the sums of its pairs' outputs (scanned next); sums past the depth's tiles
// are zero (the scan runs to the launch bound ntile, whose last entry is the total)
__global__ void __launch_bounds__(kWalkThreads) k_walk_tile_sums(WalkArgs a, Cnt3 *__restrict__ tile_cnt, int ntile)
{
    const int lane = threadIdx.x % 64;
    const TileMap tm = tile_map(a);
    for (int64_t t = 2 * tm.H + (int64_t)blockIdx.x * kWalkThreads + threadIdx.x; t <= ntile;
         t += (int64_t)gridDim.x * kWalkThreads)
        tile_cnt[t] = Cnt3{0, 0, 0};
    if (tm.tile < 0) return;
    const int64_t q0 = tm.tile * kTile;
    uint64_t sp = 0, sm = 0, sd = 0;
#pragma unroll
    for (int k = 0; k < kTile / 64; k++) {
        const int j = lane + 64 * k;
        if (j < tm.npair) {
            uint32_t p, m, d;
            pair_counts(a, a.pcnt[q0 + j], p, m, d);
            sp += p;
            sm += m;
            sd += d;
        }
    }
    sp = wave_sum(sp);
    sm = wave_sum(sm);
    sd = wave_sum(sd);
    if (lane == 0) tile_cnt[tm.tile] = Cnt3{(int64_t)sp, (int64_t)sm, (int64_t)sd};
}

// Write pass, per tile: the pairs' outputs scanned by the wavefront into offsets from the
// tile's base (LDS, one more entry: the counts are the differences); the first and last
// pair of each target set its spans; the pairs with a leaf source are stored here, the
// listed ones keep their offsets (qoff) for k_walk_write_trav.
__global__ void __launch_bounds__(kWalkThreads) k_walk_write(WalkArgs a, const Cnt3 *__restrict__ tile_off)
{
    __shared__ int32_t s_own[kTileWaves][kTile];
    __shared__ uint32_t s_off[kTileWaves][3][kTile + 1];
    if (a.sz->overflow) return;  // the same for every thread
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const TileMap tm = tile_map(a);
    if (tm.tile < 0) return;  // the same for the whole wavefront; only wavefront barriers follow
    const int64_t q0 = tm.tile * kTile;
    for (int i = lane; i < tm.npair; i += 64) s_own[w][i] = owner_of(a.din[tm.g0 + i]);
    // each lane scans 4 consecutive pairs; the wavefront scans the lanes' sums
    {
        uint32_t c[3][4];
        uint32_t tot[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int j = 4 * lane + k;
            uint32_t p = 0, m = 0, d = 0;
            if (j < tm.npair) pair_counts(a, a.pcnt[q0 + j], p, m, d);
            c[0][k] = p;
            c[1][k] = m;
            c[2][k] = d;
            tot[0] += p;
            tot[1] += m;
            tot[2] += d;
        }
#pragma unroll
        for (int x = 0; x < 3; x++) {
            uint32_t incl = tot[x];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            uint32_t run = incl - tot[x];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                s_off[w][x][4 * lane + k] = run;
                run += c[x][k];
            }
            if (lane == 63) s_off[w][x][kTile] = run;
        }
    }
    int32_t own_prev = -1, own_next = -1;  // the parents named by the records either side of the tile
    if (lane == 0 && tm.g0 > 0) own_prev = owner_of(a.din[tm.g0 - 1]);
    if (lane == 0 && tm.g0 + kTile < tm.n) own_next = owner_of(a.din[tm.g0 + kTile]);
    own_prev = __shfl(own_prev, 0, 64);
    own_next = __shfl(own_next, 0, 64);
    wave_sync();
    const Cnt3 to = tile_off[tm.tile];
    const int64_t bp = a.sz->p2p_base + to.p, bm = a.sz->m2l_base + to.m, bd = to.d;
    const int nleaf = a.sz->nleaf;
#pragma unroll
    for (int k = 0; k < kTile / 64; k += 2) {
        int j[2], T[2];
        DownRec D[2];
        bool use[2];
        int64_t qp[2], qm[2], qd[2];
        uint32_t lp[2], lm[2], ld[2], o[2][3];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            j[i] = lane + 64 * (k + i);
            const bool valid = j[i] < tm.npair;
            D[i] = valid ? a.din[tm.g0 + j[i]] : DownRec{0, 0, -1};
            T[i] = valid ? target_of(D[i], tm.c) : -1;
            const int jj = valid ? j[i] : 0;
#pragma unroll
            for (int x = 0; x < 3; x++) o[i][x] = s_off[w][x][jj];
            lp[i] = s_off[w][0][jj + 1] - o[i][0];
            lm[i] = s_off[w][1][jj + 1] - o[i][1];
            ld[i] = s_off[w][2][jj + 1] - o[i][2];
            qp[i] = bp + o[i][0];
            qm[i] = bm + o[i][1];
            qd[i] = bd + o[i][2];
            use[i] = T[i] >= 0 && src_leaf(D[i]);
            if (T[i] >= 0 && !src_leaf(D[i]) && lp[i] + lm[i] + ld[i] > 0) {
                uint32_t *qo = a.qoff + 3 * (q0 + j[i]);
                qo[0] = o[i][0];
                qo[1] = o[i][1];
                qo[2] = o[i][2];
            }
            if (T[i] >= 0) {
                // spans: the records of one parent are contiguous, and t0 names it
                const int32_t own = owner_of(D[i]);
                const int32_t prev = j[i] > 0 ? s_own[w][j[i] - 1] : own_prev;
                const int32_t nxt = j[i] + 1 < tm.npair ? s_own[w][j[i] + 1] : own_next;
                if (prev != own) {
                    if (T[i] < nleaf) a.p2p_span[2 * (size_t)T[i]] = qp[i];
                    a.m2l_span[2 * (size_t)T[i]] = qm[i];
                }
                if (nxt != own) {
                    if (T[i] < nleaf) a.p2p_span[2 * (size_t)T[i] + 1] = qp[i] + lp[i];
                    a.m2l_span[2 * (size_t)T[i] + 1] = qm[i] + lm[i];
                }
            }
        }
        // a pair with a leaf source has one output at most, and its count (from the count
        // pass's rule) says which: stored from the record alone, without the rule again
#pragma unroll
        for (int i = 0; i < 2; i++) {
            if (!use[i]) continue;
            const int cell = entry_cell(D[i].entry), sid = entry_shift(D[i].entry);
            if (lp[i]) {
                a.p2p_out[qp[i]] = pack_entry(cell, sid);
            } else if (lm[i]) {
                a.m2l_out[qm[i]] = pack_entry(cell, sid);
            } else if (ld[i]) {
                const TgtRec &t = a.trec[T[i]];
                a.dl_out[qd[i]] = DownRec{pack_entry(cell, sid), t.son[0] | kSrcLeaf, t.son[1]};
            }
        }
    }
}

// Write pass, the listed pairs, as in k_walk_count_trav, storing at the offsets the tile
// pass left; a continued traversal (flagged in its pair's word) stops where the count
// pass handed it on and leaves where its rest goes.
__global__ void __launch_bounds__(kWalkThreads) __attribute__((amdgpu_waves_per_eu(5))) k_walk_write_trav(WalkArgs a, const Cnt3 *__restrict__ tile_off)
{
    if (a.sz->overflow) return;
    const int lane = threadIdx.x % 64;
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t gw = (int64_t)blockIdx.x * kTileWaves + __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int64_t p2p_base = a.sz->p2p_base, m2l_base = a.sz->m2l_base;
    TravStream ts;
    stream_open(a, ts, gw);
    LanePair L;
    L.q = -1;
    int64_t qp = 0, qm = 0, qd = 0;
    uint32_t lp = 0, lm = 0, ld = 0;
    uint64_t wd = 0;
    for (;;) {
        stream_advance(a, ts);
        if (!ts.more && !__ballot(L.q >= 0)) break;
        const bool want = L.q < 0;
        const uint64_t bw = __ballot(want);
        if (want) {
            const int idx = ts.next + __popcll(bw & below);
            if (idx < ts.cn) {
                const TravRec tr = super_pair(a, ts.sc, idx);
                wd = a.pcnt[tr.q];
                pair_counts(a, wd, lp, lm, ld);
                if (lp + lm + ld > 0) {
                    const Cnt3 to = tile_off[tr.q / kTile];
                    const uint32_t *qo = a.qoff + 3 * (int64_t)tr.q;
                    qp = p2p_base + to.p + qo[0];
                    qm = m2l_base + to.m + qo[1];
                    qd = to.d + qo[2];
                    lane_start(a, L, tr.q, tr.T, tr.entry);
                }
            }
        }
        ts.next = min(ts.next + __popcll(bw), ts.cn);
        if (L.q >= 0) {
            L.r = visit<true>(a, L.t, L.R, L.r, L.sid, L.sx, L.sy, L.sz, L.np, L.nm, L.nd, qp, qm, qd, lp, lm, ld);
            L.visits++;
            if (L.r >= L.rend) {
                L.q = -1;
            } else if (L.visits == a.visits && (wd & 1u)) {
                // the rest is k_walk_deep_write's, from here (flagged by the count pass at
                // the same step)
                Cont &ct = a.cont[wd >> 1];
                ct.qp = qp + L.np;
                ct.qm = qm + L.nm;
                ct.qd = qd + L.nd;
                L.q = -1;
            } else {
                L.R = a.rec[L.r];
            }
        }
    }
}

// the continued traversals' outputs, where the write pass left them
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_walk_deep_write(WalkArgs a)
{
    constexpr int G = kGroupLanes;
    if (a.sz->overflow) return;
    const int gl = threadIdx.x % G;
    const int64_t npad = npad_of(a), ntot = cont_total(a);
    const int64_t ngroup = (int64_t)gridDim.x * (blockDim.x / G);
    for (int64_t m = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; m < ntot; m += ngroup) {
        const int64_t k = a.cidx[m];
        const Cont ct = a.cont[k];
        int T;
        int32_t entry;
        pair_of(a, npad, ct.q, T, entry);
        const TgtRec t = a.trec[T];
        uint32_t p0, m0, d0;
        unpack_cnt(ct.part, p0, m0, d0);
        uint32_t np = 0, nm = 0, nd = 0;
        group_traverse<true, G>(a, t, ct.r, ct.rend, entry_shift(entry), gl, np, nm, nd, ct.qp, ct.qm, ct.qd,
                                ct.tp - p0, ct.tm - m0, ct.td - d0);
    }
}

// The depth's totals (P2P, M2L and down records: the tile scan's last entry) against the
// capacities: the list bases move past the previous depth's entries, and what would not
// fit flags the overflow, after which nothing more is written.
__global__ void k_walk_step(DevSizes *sz, const Cnt3 *__restrict__ tot, int64_t p2p_cap, int64_t m2l_cap,
                            int64_t dl_cap)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    sz->p2p_base = sz->n_p2p;
    sz->m2l_base = sz->n_m2l;
    const Cnt3 t = *tot;
    int ovf = 0;
    if (sz->n_p2p + t.p > p2p_cap) ovf |= kOvfP2P;
    if (sz->n_m2l + t.m > m2l_cap) ovf |= kOvfM2L;
    if (t.d > dl_cap) ovf |= kOvfDown;
    if (ovf || sz->overflow) {
        sz->overflow |= ovf;
        sz->dl_total = 0;
        return;
    }
    sz->n_p2p += t.p;
    sz->n_m2l += t.m;
    sz->dl_total = t.d;
}

// between depths: the down records written become the next depth's records
__global__ void k_walk_next(DevSizes *sz)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    sz->nrec = sz->dl_total;
}

// walk start: empty lists, bases at zero, n0 start records -- none when the tree has
// overflowed (a planned evaluation whose tree outgrew the plan wrote no cells and no walk
// records, and its buffers may be fresh: nothing of them is read)
__global__ void k_walk_begin(DevSizes *sz, int64_t n0)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    sz->n_p2p = sz->n_m2l = sz->p2p_base = sz->m2l_base = 0;
    sz->dl_total = 0;
    sz->nrec = sz->overflow ? 0 : n0;
}

// after the last depth of a planned walk: records left over mean deeper targets that
// were never visited
__global__ void k_walk_end(DevSizes *sz)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (sz->dl_total > 0) sz->overflow |= kOvfDepths;
}

// by depth-first rank, the cells as sources (geometry, id, subtree size, children in rank
// order); by cell, the cells as targets
__global__ void k_walk_records(const DevSizes *__restrict__ sz, const CellInfo *__restrict__ info,
                               const CellGeom *__restrict__ geom, BoxView bx, WalkRec *__restrict__ rec,
                               TgtRec *__restrict__ trec)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (sz->overflow || c >= sz->ncell) return;
    const CellInfo I = info[c];
    const CellGeom g = geom[c];
    WalkRec R;
    R.g = g;
    R.cell = c;
    R.dsize = I.dsize;
    R.kid[0] = R.kid[1] = -1;
    TgtRec t;
    t.g = g;
    t.L = bx.len_of(c);
    t.rank = I.dfs;
    t.son[0] = t.son[1] = t.sonr[0] = t.sonr[1] = -1;
    t.leaf = c < sz->nleaf;
    if (I.son0 >= 0) {
        const int r0 = info[I.son0].dfs, r1 = info[I.son1].dfs;
        const bool first0 = r0 == I.dfs + 1;
        R.kid[0] = first0 ? I.son0 : I.son1;
        R.kid[1] = first0 ? I.son1 : I.son0;
        t.son[0] = I.son0;
        t.son[1] = I.son1;
        t.sonr[0] = r0;
        t.sonr[1] = r1;
    }
    rec[I.dfs] = R;
    trec[c] = t;
}

// every span empty (cells a walk from a given pair never reaches keep no entries); n is
// the launch bound of the cells
__global__ void k_clear_spans(int n, int64_t *__restrict__ span)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * (int64_t)n) span[i] = 0;
}

// the start records: every box's root against itself (box_root given) or the given pair
// (target A, source B), one record per shift in [shift_lo, shift_hi)
__global__ void k_start_list(const DevSizes *__restrict__ sz, DownRec *f, const CellInfo *__restrict__ info,
                             const int *__restrict__ box_root, int nbox, int A, int B, int shift_lo, int shift_hi)
{
    const int nshift = shift_hi - shift_lo;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // after a tree overflow the roots and cells were not written: the walk starts empty
    if (sz->overflow || i >= (box_root ? nbox : 1) * nshift) return;
    const int s = shift_lo + i % nshift;
    const int tgt = box_root ? box_root[i / nshift] : A;
    const int src = box_root ? tgt : B;
    const bool lf = info[src].son0 < 0;
    f[i] = DownRec{pack_entry(lf ? src : info[src].dfs, s), tgt | (lf ? kSrcLeaf : 0), -1};
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
// or from every box's (root, root) when start_target < 0.  Per depth: the count pass and
// its continuations, the scan of the tile sums (an ordinary walk reads the depth's three
// totals to size the outputs), the write pass and its continuations.
void run_walk_from(Ctx &c, int shift_lo, int shift_hi, int start_target, int start_source)
{
    hipStream_t st = c.stream;
    const bool planned = c.plan.active;
    const int nlb = nleaf_bound(c), ncb = ncell_bound(c);
    if (ncb >= (1 << 27)) throw std::runtime_error("more than 2^27 cells");
    need(c.cell_box, sizeof(int) * (size_t)ncb, "cell boxes");
    const bool forest = start_target < 0;
    if (planned && !forest) throw std::runtime_error("a planned walk starts from the roots");
    // depths of the walk: the plan's, or (ordinary) until no records are left, at most
    // one more than the tree's levels
    const int ndepth = planned ? c.plan.depths : kMaxDepth + 1;

    DevSizes *sz = dptr<DevSizes>(c.sizes);
    WalkArgs a{};
    a.sz = sz;
    a.theta = c.prm.theta;
    a.theta2 = c.prm.theta * c.prm.theta;
    a.rc = c.prm.rc;
    a.rc2 = c.prm.rc * c.prm.rc;
    WalkRec *rec = static_cast<WalkRec *>(ctx_buf(c, c.walk_rec, sizeof(WalkRec) * (size_t)ncb, "walk records"));
    TgtRec *trec = static_cast<TgtRec *>(ctx_buf(c, c.walk_trec, sizeof(TgtRec) * (size_t)ncb, "walk records"));
    k_walk_records<<<(ncb + 255) / 256, 256, 0, st>>>(sz, dptr<CellInfo>(c.info), dptr<CellGeom>(c.geom), box_view(c),
                                                       rec, trec);
    a.rec = rec;
    a.trec = trec;
    a.p2p_span = static_cast<int64_t *>(ctx_buf(c, c.p2p_off, sizeof(int64_t) * 2 * (size_t)std::max(1, nlb), "P2P list"));
    a.m2l_span = static_cast<int64_t *>(ctx_buf(c, c.m2l_off, sizeof(int64_t) * 2 * (size_t)ncb, "M2L list"));
    k_clear_spans<<<(2 * nlb + 255) / 256 + 1, 256, 0, st>>>(nlb, a.p2p_span);
    k_clear_spans<<<(2 * ncb + 255) / 256 + 1, 256, 0, st>>>(ncb, a.m2l_span);

    // start: every box's root against itself, or the given pair; shifts [lo, hi)
    const int nshift = shift_hi - shift_lo;
    const int64_t n0 = (int64_t)(forest ? c.nbox : 1) * nshift;
    if (forest) need(c.d_box_root, sizeof(int) * (size_t)c.nbox, "box roots");
    if (planned && n0 > c.plan.dl_cap) throw std::runtime_error("the plan holds fewer walk records than the start");
    DevBuf *din = &c.dl_a, *dout = &c.dl_b;
    DownRec *slist = static_cast<DownRec *>(ctx_buf(c, *din, sizeof(DownRec) * (size_t)(planned ? c.plan.dl_cap : n0),
                                                    "walk records"));
    k_start_list<<<(int)((n0 + 255) / 256), 256, 0, st>>>(sz, slist, dptr<CellInfo>(c.info),
                                                          forest ? dptr<int>(c.d_box_root) : nullptr, c.nbox,
                                                          start_target, start_source, shift_lo, shift_hi);
    k_walk_begin<<<1, 1, 0, st>>>(sz, n0);

    int64_t *pin = reinterpret_cast<int64_t *>(c.pinned.p) + 16;
    int64_t nrec = n0, n_p2p = 0, n_m2l = 0;  // host copies (ordinary evaluation)
    int64_t max_rec = n0;
    std::vector<int64_t> rec_depth;
    int depths = 0;
    for (int d = 0; d < ndepth && (planned || nrec > 0); d++) {
        // launch bound of this depth's records: the plan's (per depth), or the count
        const int64_t rcap = planned ? c.plan.rec_depth[d] : nrec;
        if (!planned) rec_depth.push_back(nrec);
        if (2 * (rcap + kTile) >= ((int64_t)1 << 31))
            throw std::runtime_error("more than 2^31 walk pairs at one depth");
        a.din = static_cast<const DownRec *>(din->p);
        a.rec_cap = rcap;
        const int ntile = (int)(2 * ((rcap + kTile - 1) / kTile));  // >= 2: rcap >= 1
        a.pcnt = static_cast<uint64_t *>(ctx_buf(c, c.pair_cnt, sizeof(uint64_t) * kTile * (size_t)ntile, "walk pairs"));
        // the traversal lists, per tile; the continuation slots, a fixed number per
        // traversal wavefront
        a.tlist = static_cast<TravRec *>(ctx_buf(c, c.walk_tlist, sizeof(TravRec) * kTile * (size_t)ntile, "walk pairs"));
        a.tcount = static_cast<int32_t *>(ctx_buf(c, c.walk_tcount, sizeof(int32_t) * (size_t)ntile, "walk pairs"));
        a.qoff = static_cast<uint32_t *>(ctx_buf(c, c.walk_qoff, 3 * sizeof(uint32_t) * kTile * (size_t)ntile, "walk pairs"));
        // the traversal passes: about one unit per wavefront, at most kTravWavesMax; at a
        // depth of fewer than kTravWavesFill super-chunks, each is cut into slices (up to
        // kTravSplitMax) so that its traversals spread over more wavefronts
        const int64_t nsuper_cap = (ntile + kSuper - 1) / kSuper;
        const int split = (int)std::min<int64_t>(kTravSplitMax, std::max<int64_t>(1, kTravWavesFill / nsuper_cap));
        const int64_t nunit_cap = nsuper_cap * split;
        const int tw = (int)std::min<int64_t>(kTravWavesMax, (nunit_cap + kTileWaves - 1) / kTileWaves * kTileWaves);
        a.trav_split = split;
        a.visits = split > 1 ? kThreadVisitsFew : kThreadVisits;
        const int nslot = tw * kContPerWave;
        a.trav_waves = tw;
        a.cont = static_cast<Cont *>(ctx_buf(c, c.big_pairs, sizeof(Cont) * (size_t)nslot, "walk continuations"));
        int32_t *ccount = static_cast<int32_t *>(ctx_buf(c, c.walk_ccount, sizeof(int32_t) * 2 * (size_t)tw,
                                                         "walk continuations"));
        a.ccount = ccount;
        a.coff = ccount + tw;
        a.cidx = static_cast<int32_t *>(ctx_buf(c, c.walk_cidx, sizeof(int32_t) * (size_t)nslot, "walk continuations"));
        Cnt3 *tile_cnt = static_cast<Cnt3 *>(ctx_buf(c, c.cnt_tmp, sizeof(Cnt3) * 2 * (size_t)(ntile + 1), "walk tiles"));
        Cnt3 *tile_off = tile_cnt + (ntile + 1);
        const int nblk = walk_blocks(ntile);
        // the leaf-source pairs and the lists; the traversals and their continuations; the
        // tile sums, scanned
        k_walk_count<<<nblk, kWalkThreads, 0, st>>>(a);
        k_walk_count_trav<<<tw / kTileWaves, kWalkThreads, 0, st>>>(a);
        {
            size_t bytes = 0;
            PH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, ccount, ccount + tw, tw, st));
            void *tmp = scan_storage(c, bytes);
            PH_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, ccount, ccount + tw, tw, st));
        }
        k_walk_cont_index<<<(nslot + 255) / 256, 256, 0, st>>>(a);
        k_walk_deep_count<<<kDeepBlocks, 256, 0, st>>>(a);
        k_walk_tile_sums<<<nblk, kWalkThreads, 0, st>>>(a, tile_cnt, ntile);
        {
            size_t bytes = 0;
            PH_CHECK(hipcub::DeviceScan::ExclusiveScan(nullptr, bytes, tile_cnt, tile_off, AddCnt3(), Cnt3{0, 0, 0},
                                                       ntile + 1, st));
            void *tmp = scan_storage(c, bytes);
            PH_CHECK(hipcub::DeviceScan::ExclusiveScan(tmp, bytes, tile_cnt, tile_off, AddCnt3(), Cnt3{0, 0, 0},
                                                       ntile + 1, st));
        }
        int64_t p2p_cap, m2l_cap, dl_cap;
        if (planned) {
            p2p_cap = c.plan.p2p_cap;
            m2l_cap = c.plan.m2l_cap;
            // the records written here are the next depth's, within its launch bound
            dl_cap = d + 1 < ndepth ? std::min(c.plan.dl_cap, c.plan.rec_depth[d + 1]) : c.plan.dl_cap;
        } else {
            // the depth's totals size its outputs
            PH_CHECK(hipMemcpyAsync(pin, tile_off + ntile, sizeof(Cnt3), hipMemcpyDeviceToHost, st));
            PH_CHECK(hipMemcpyAsync(pin + 3, &sz->overflow, sizeof(int), hipMemcpyDeviceToHost, st));
            PH_CHECK(hipEventRecord(c.counts_ready, st));
            PH_CHECK(hipEventSynchronize(c.counts_ready));
            if (const int ovf = *reinterpret_cast<const int *>(pin + 3))
                throw std::runtime_error("walk: a pair's outputs exceed their packed counts (overflow bits " +
                                         std::to_string(ovf) + ")");
            const int64_t add_p2p = pin[0], add_m2l = pin[1], add_dl = pin[2];
            grow_keep(c.p2p_src, 4 * (size_t)std::max<int64_t>(n_p2p + add_p2p, 1), 4 * (size_t)n_p2p, st);
            grow_keep(c.m2l_src, 4 * (size_t)std::max<int64_t>(n_m2l + add_m2l, 1), 4 * (size_t)n_m2l, st);
            // this depth's records (the other ping-pong buffer) are read by the write: only
            // the output buffer may move
            grow_keep(*dout, sizeof(DownRec) * (size_t)std::max<int64_t>(add_dl, 1), 0, st);
            n_p2p += add_p2p;
            n_m2l += add_m2l;
            nrec = add_dl;
            max_rec = std::max(max_rec, nrec);
            depths = d + 1;
            p2p_cap = c.p2p_src.cap / 4;
            m2l_cap = c.m2l_src.cap / 4;
            dl_cap = dout->cap / sizeof(DownRec);
        }
        // at the plan's capacities (allocated by the first planned evaluation, checked by the
        // others); an ordinary evaluation grew them above
        ctx_buf(c, c.p2p_src, sizeof(int32_t) * (size_t)std::max<int64_t>(p2p_cap, 1), "P2P list");
        ctx_buf(c, c.m2l_src, sizeof(int32_t) * (size_t)std::max<int64_t>(m2l_cap, 1), "M2L list");
        ctx_buf(c, *dout, sizeof(DownRec) * (size_t)std::max<int64_t>(dl_cap, 1), "walk records");
        a.p2p_out = dptr<int32_t>(c.p2p_src);
        a.m2l_out = dptr<int32_t>(c.m2l_src);
        a.dl_out = dptr<DownRec>(*dout);
        k_walk_step<<<1, 1, 0, st>>>(sz, tile_off + ntile, p2p_cap, m2l_cap, dl_cap);
        k_walk_write<<<nblk, kWalkThreads, 0, st>>>(a, tile_off);
        k_walk_write_trav<<<tw / kTileWaves, kWalkThreads, 0, st>>>(a, tile_off);
        k_walk_deep_write<<<kDeepBlocks, 256, 0, st>>>(a);
        k_walk_next<<<1, 1, 0, st>>>(sz);
        std::swap(din, dout);
    }
    if (planned) {
        k_walk_end<<<1, 1, 0, st>>>(sz);
    } else {
        c.n_p2p = n_p2p;
        c.n_m2l = n_m2l;
        if (forest) {
            Measure &m = c.meas;
            m.depths = depths;
            m.dl = max_rec;
            m.dl_depth = rec_depth;
            m.p2p = n_p2p;
            m.m2l = n_m2l;
        }
    }
    c.lists_valid = true;
    PH_CHECK(hipGetLastError());
}

void run_walk(Ctx &c, int shift_lo, int shift_hi) { run_walk_from(c, shift_lo, shift_hi, -1, -1); }

}  // namespace photons
