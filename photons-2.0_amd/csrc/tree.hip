// tree.hip -- kd tree of the short-range force on the GPU.
//
// Same tree as build_kdtree/bksort_inplace/center_kdtree (src/fmm.c:30-177): a node
// is split at the mean of its particles' coordinate along the axis of its depth
// (x at the root, cycling), particles <= mean go left; a child with at most maxleaf
// particles is a leaf (possibly empty), boxes are cut at the split planes from
// the domain box [0, L)^3.  When no particle lies above the mean the last one goes
// right, and when none lies at or below it (equal coordinates whose mean rounds low)
// the left child is empty, as the reference's two-pointer partition leaves them.
// Only when the same particles have gone right whole on all three axes in a row (they
// coincide, and every mean rounds low), where the reference would split forever, the
// first particle goes left instead.
//
// MI355X formulation: one level of the tree per step over all nodes of that
// level.  Each node's particle range is cut into chunks of kChunk elements, one
// wavefront per chunk: a deterministic per-chunk sum (fixed lane order and
// butterfly), per-node mean from the chunk partials in chunk order, a stable
// partition whose per-chunk offsets come from one exclusive scan.  After the last
// level, subtree sizes give the depth-first numbering of the reference (leaves
// in creation order, nodes in pre-order).  The mean's rounding differs from the
// reference's sequential sum by a few ulps; the partition (which particles go
// left) is the same except for particles within those ulps of the plane.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>

#include "engine.h"

namespace photons {

constexpr int kChunk = 2048;
constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ int find_seg(const int *__restrict__ seg_chunk, int S, int c)
{
    int lo = 0, hi = S;  // seg_chunk[lo] <= c < seg_chunk[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (seg_chunk[mid] <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// t_depth holds a node's depth in its low 16 bits and, above them, how many splits in a
// row have passed its particles to the right child whole (the empty-left case)
__host__ __device__ __forceinline__ int depth_of(int d) { return d & 0xffff; }
__host__ __device__ __forceinline__ int streak_of(int d) { return d >> 16; }
constexpr int kStuckStreak = 3;  // all three axes: the split would repeat forever

// A level's nodes (segments) are the temporary nodes sz->lvl[d] .. sz->lvl[d+1] - 1, read
// on the device.  Scap bounds the segments of any level and C the chunks (Scap +
// ceil(n / kChunk)); the launches are sized by them, the kernels take the exact counts
// (S and seg_chunk[S]) from the device.
struct LevelArgs {
    const double *coord;  // the split axis of this level
    const int *t_start, *t_count, *t_depth;
    const DevSizes *sz;
    int d;                 // the level
    int Scap, C;
    const int *seg_chunk;  // [Scap+1], exclusive scan of the chunks per segment
    int axis;              // 0, 1, 2: which of x, y, z coord is
    __device__ __forceinline__ int seg0() const { return sz->lvl[d]; }
    __device__ __forceinline__ int nseg() const { return sz->lvl[d + 1] - sz->lvl[d]; }
};

__device__ __forceinline__ double axis_value(int axis, double x, double y, double z)
{
    return axis == 0 ? x : (axis == 1 ? y : z);
}

__global__ void __launch_bounds__(256) k_chunk_sum(LevelArgs a, double *__restrict__ partial)
{
    int c = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    int lane = threadIdx.x % kWave;
    const int S = a.nseg(), seg0 = a.seg0();
    if (c >= a.seg_chunk[S]) return;
    int s = find_seg(a.seg_chunk, S, c);
    int st = a.t_start[seg0 + s], ct = a.t_count[seg0 + s];
    int b = st + (c - a.seg_chunk[s]) * kChunk;
    int e = min(b + kChunk, st + ct);
    double v = 0.0;
    // four loads in flight per step, added in the same order as one at a time
    int i = b + lane;
    for (; i + 3 * kWave < e; i += 4 * kWave) {
        const double x0 = a.coord[i], x1 = a.coord[i + kWave], x2 = a.coord[i + 2 * kWave],
                     x3 = a.coord[i + 3 * kWave];
        v += x0;
        v += x1;
        v += x2;
        v += x3;
    }
    for (; i < e; i += kWave) v += a.coord[i];
    v = wave_sum(v);
    if (lane == 0) partial[c] = v;
}

// chunks per segment, zeros up to the bound (the exclusive scan then ends with the total)
__global__ void k_nchunks(LevelArgs a, int *__restrict__ out)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > a.Scap) return;
    out[s] = (s >= a.nseg()) ? 0 : (a.t_count[a.seg0() + s] + kChunk - 1) / kChunk;
}

// One wavefront per segment: lane l adds the chunk partials l, l + 64, ... in order and a
// fixed butterfly combines the lanes, so the mean is deterministic, a one-chunk segment
// gets its partial exactly, and the root's thousands of chunks are not one thread's
// chain of dependent loads.
__global__ void __launch_bounds__(256) k_mean(LevelArgs a, const double *__restrict__ partial,
                                              double *__restrict__ t_split)
{
    const int s = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int seg0 = a.seg0();
    if (s >= a.nseg()) return;
    double v = 0.0;
    for (int c = a.seg_chunk[s] + lane; c < a.seg_chunk[s + 1]; c += kWave) v += partial[c];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    if (lane == 0) t_split[seg0 + s] = v / (double)a.t_count[seg0 + s];
}

__global__ void __launch_bounds__(256) k_chunk_count(LevelArgs a, const double *__restrict__ t_split,
                                                     int *__restrict__ rcnt)
{
    int c = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    int lane = threadIdx.x % kWave;
    const int S = a.nseg(), seg0 = a.seg0();
    if (c > a.C) return;
    if (c >= a.seg_chunk[S]) {  // zeros up to the bound: the exclusive scan ends with the total
        if (lane == 0) rcnt[c] = 0;
        return;
    }
    int s = find_seg(a.seg_chunk, S, c);
    int st = a.t_start[seg0 + s], ct = a.t_count[seg0 + s];
    double split = t_split[seg0 + s];
    int b = st + (c - a.seg_chunk[s]) * kChunk;
    int e = min(b + kChunk, st + ct);
    int n = 0;
    int i = b + lane;
    for (; i + 3 * kWave < e; i += 4 * kWave)
        n += (a.coord[i] > split) + (a.coord[i + kWave] > split) + (a.coord[i + 2 * kWave] > split) +
             (a.coord[i + 3 * kWave] > split);
    for (; i < e; i += kWave) n += a.coord[i] > split;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) n += __shfl_xor(n, o, kWave);
    if (lane == 0) rcnt[c] = n;
}

// per segment: particles left of the split, the degenerate-partition flag, and the kinds
// of its two children (new nodes in the high 32 bits, new leaves in the low 32 bits, packed
// so that one scan gives both offsets; the leaf count stays below 2^32, so no carry crosses)
__global__ void k_seg_counts(LevelArgs a, const int *__restrict__ rpre, int maxleaf, int *__restrict__ nleft,
                             int *__restrict__ flag, uint64_t *__restrict__ kinds)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int S = a.nseg(), seg0 = a.seg0();
    if (s > a.Scap) return;
    if (s >= S) {  // zeros up to the bound: the scan's entry S is the total
        kinds[s] = 0;
        return;
    }
    int ct = a.t_count[seg0 + s];
    int nr = rpre[a.seg_chunk[s + 1]] - rpre[a.seg_chunk[s]];
    // nothing above the mean: the last particle goes right (f = 1: the order is kept);
    // nothing at or below it (nr = ct): the left child is empty, everything goes right,
    // unless that has happened on all three axes before (f = 2: the first goes left)
    int nl = ct - nr, f = 0;
    if (nr == 0) {
        nl = ct - 1;
        f = 1;
    } else if (nl == 0 && streak_of(a.t_depth[seg0 + s]) >= kStuckStreak) {
        nl = 1;
        f = 2;
    }
    nleft[s] = nl;
    flag[s] = f;
    const int k = (nl > maxleaf) + (ct - nl > maxleaf);
    kinds[s] = ((uint64_t)k << 32) | (uint64_t)(2 - k);
}

struct Particles {
    double *x, *y, *z;
    int *idx;
};

__global__ void __launch_bounds__(256) k_scatter(LevelArgs a, const double *__restrict__ t_split,
                                                 const int *__restrict__ rpre, const int *__restrict__ nleft,
                                                 const int *__restrict__ flag, Particles src, Particles dst)
{
    int c = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    int lane = threadIdx.x % kWave;
    const int S = a.nseg(), seg0 = a.seg0();
    if (c >= a.seg_chunk[S]) return;
    int s = find_seg(a.seg_chunk, S, c);
    int st = a.t_start[seg0 + s], ct = a.t_count[seg0 + s];
    int ci = c - a.seg_chunk[s];
    int b = st + ci * kChunk;
    int e = min(b + kChunk, st + ct);
    double split = t_split[seg0 + s];
    int f = flag[s];
    int nl = nleft[s];
    int rbase = rpre[c] - rpre[a.seg_chunk[s]];
    int lbase = ci * kChunk - rbase;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // four groups of 64 particles per pass: their loads are all in flight before the
    // first ballot, so a wave waits on memory once per 256 particles instead of per 64
    constexpr int G = 4;
    for (int i0 = b; i0 < e; i0 += G * kWave) {
        double px[G] = {}, py[G] = {}, pz[G] = {};
        int pi[G] = {};
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int i = i0 + g * kWave + lane;
            if (i < e) {
                px[g] = src.x[i];
                py[g] = src.y[i];
                pz[g] = src.z[i];
                pi[g] = src.idx[i];
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int i = i0 + g * kWave + lane;
            const bool active = i < e;
            const double v = axis_value(a.axis, px[g], py[g], pz[g]);
            const bool right = active && f == 0 && v > split;
            const uint64_t bal = __ballot(right);
            const int rr = __popcll(bal & lt);
            const int rl = lane - rr;
            if (active) {
                const int d = (f != 0) ? i : (right ? st + nl + rbase + rr : st + lbase + rl);
                dst.x[d] = px[g];
                dst.y[d] = py[g];
                dst.z[d] = pz[g];
                dst.idx[d] = pi[g];
            }
            const int nact = max(0, min(kWave, e - (i0 + g * kWave)));
            const int nr = __popcll(bal);
            rbase += nr;
            lbase += nact - nr;
        }
    }
}

struct TreeTmp {
    int *t_start, *t_count, *t_son, *t_depth, *t_parent;
    double *t_split, *t_lo, *t_hi;  // lo/hi: [3][cap] SoA
    int *l_start, *l_count, *l_parent;
    double *l_lo, *l_hi;
    int tcap, lcap;
};

// children of the level's nodes: the new nodes (depth d+1) from temporary id lvl[d+1] on,
// the new leaves from ldo[d+1] on; ids past the capacities are not written (the level
// step below flags the overflow and stops the tree there)
__global__ void k_make_children(LevelArgs a, TreeTmp T, int axis, int maxleaf, const int *__restrict__ nleft,
                                const uint64_t *__restrict__ kind_off)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nseg()) return;
    int v = a.seg0() + s;
    const int node_base = a.sz->lvl[a.d + 1], leaf_base = a.sz->ldo[a.d + 1];
    int st = T.t_start[v], ct = T.t_count[v];
    int cnt[2] = {nleft[s], ct - nleft[s]};
    int beg[2] = {st, st + nleft[s]};
    double split = T.t_split[v];
    const int depth = depth_of(T.t_depth[v]);
    // the right child inherits the streak of whole passes, plus this one if nothing went left
    const int streak = nleft[s] == 0 ? streak_of(T.t_depth[v]) + 1 : 0;
    const uint64_t ko = kind_off[s];
    int nextn = node_base + (int)(ko >> 32), nextl = leaf_base + (int)(uint32_t)ko;
    for (int k = 0; k < 2; k++) {
        double lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = T.t_lo[d * T.tcap + v];
            hi[d] = T.t_hi[d * T.tcap + v];
        }
        if (k == 0) hi[axis] = split;
        else lo[axis] = split;
        if (cnt[k] > maxleaf) {
            int id = nextn++;
            if (id >= T.tcap) continue;
            T.t_start[id] = beg[k];
            T.t_count[id] = cnt[k];
            T.t_depth[id] = (depth + 1) | ((k == 1 ? streak : 0) << 16);
            T.t_parent[id] = v;
            for (int d = 0; d < 3; d++) {
                T.t_lo[d * T.tcap + id] = lo[d];
                T.t_hi[d * T.tcap + id] = hi[d];
            }
            T.t_son[2 * v + k] = id;
        } else {
            int id = nextl++;
            if (id >= T.lcap) continue;
            T.l_start[id] = beg[k];
            T.l_count[id] = cnt[k];
            T.l_parent[id] = v;
            for (int d = 0; d < 3; d++) {
                T.l_lo[d * T.lcap + id] = lo[d];
                T.l_hi[d * T.lcap + id] = hi[d];
            }
            T.t_son[2 * v + k] = -(id + 1);
        }
    }
}

// After a level: the next level's node range and the leaves of depth d+2 from the scanned
// child kinds (kind_off[S] is the total: new nodes high, new leaves low).  A level that
// would pass the node capacities (temporary nodes, nodes per level) or the leaf capacity
// flags the overflow and ends the tree.
__global__ void k_level_step(DevSizes *sz, int d, const uint64_t *__restrict__ kind_off, int tcap, int lcap,
                             int scap)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int S = sz->lvl[d + 1] - sz->lvl[d];
    const uint64_t tot = kind_off[S];
    const int nn = (int)(tot >> 32), nl = (int)(uint32_t)tot;
    int ovf = 0;
    if ((int64_t)sz->lvl[d + 1] + nn > tcap) ovf |= kOvfNodes;
    if (nn > scap) ovf |= kOvfNodes;
    if ((int64_t)sz->ldo[d + 1] + nl > lcap) ovf |= kOvfLeaves;
    if (ovf) {
        sz->overflow |= ovf;
        sz->lvl[d + 2] = sz->lvl[d + 1];
        sz->ldo[d + 2] = sz->ldo[d + 1];
        return;
    }
    sz->lvl[d + 2] = sz->lvl[d + 1] + nn;
    sz->ldo[d + 2] = sz->ldo[d + 1] + nl;
}

// the roots: one per box, temporary ids 0 .. nbox-1, boxes [0, L_b]^3; the sizes start over
__global__ void k_tree_roots(int nbox, const int64_t *__restrict__ box_off, const double *__restrict__ box_len,
                             TreeTmp T, DevSizes *sz)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) {
        sz->overflow = 0;
        sz->lvl[0] = 0;
        sz->lvl[1] = nbox;
        sz->ldo[0] = sz->ldo[1] = 0;
    }
    if (b >= nbox) return;
    T.t_start[b] = (int)box_off[b];
    T.t_count[b] = (int)(box_off[b + 1] - box_off[b]);
    T.t_depth[b] = 0;
    T.t_parent[b] = -1;
    for (int d = 0; d < 3; d++) {
        T.t_lo[d * T.tcap + b] = 0.0;
        T.t_hi[d * T.tcap + b] = box_len[b];
    }
}

// after `levels` levels: the cell counts, the levels that hold nodes, and the walk targets
// of each depth (nodes, then leaves); nodes left at depth `levels` were never split
// (a tree deeper than planned): flagged
__global__ void k_tree_sizes(DevSizes *sz, int levels)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (sz->lvl[levels + 1] > sz->lvl[levels]) sz->overflow |= kOvfLevels;
    const int ntmp = sz->lvl[levels], nleaf = sz->ldo[levels + 1];
    sz->nnode = ntmp;
    sz->nleaf = nleaf;
    sz->ncell = ntmp + nleaf;
    int nlevel = 0;
    while (nlevel < levels && sz->lvl[nlevel + 1] > sz->lvl[nlevel]) nlevel++;
    sz->nlevel = nlevel;
}

// The levels scatter alternately into the two particle buffers and never copy back:
// the particles of a leaf at depth D are left in buffer D % 2.  This pass moves the
// leaves held by the second buffer into the first, one wavefront per leaf.
__global__ void __launch_bounds__(256) k_gather_leaves(TreeTmp T, const DevSizes *__restrict__ sz, Particles src,
                                                       Particles dst)
{
    int l = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    int lane = threadIdx.x % kWave;
    // leaf depth = parent depth + 1; an even leaf depth is already in the first buffer
    if (sz->overflow || l >= sz->nleaf || (depth_of(T.t_depth[T.l_parent[l]]) & 1) == 1) return;
    int b = T.l_start[l], e = b + T.l_count[l];
    for (int i = b + lane; i < e; i += kWave) {
        dst.x[i] = src.x[i];
        dst.y[i] = src.y[i];
        dst.z[i] = src.z[i];
        dst.idx[i] = src.idx[i];
    }
}

// ---- final numbering

__global__ void k_sizes(const int *__restrict__ t_son, const DevSizes *__restrict__ sz, int d, int *__restrict__ size,
                        int *__restrict__ lsize)
{
    int v = sz->lvl[d] + blockIdx.x * blockDim.x + threadIdx.x;
    if (sz->overflow || v >= sz->lvl[d + 1]) return;
    int cells = 1, ls = 0;
    for (int k = 0; k < 2; k++) {
        int c = t_son[2 * v + k];
        if (c < 0) {
            cells += 1;
            ls += 1;
        } else {
            cells += size[c];
            ls += lsize[c];
        }
    }
    size[v] = cells;
    lsize[v] = ls;
}

__global__ void k_number(const int *__restrict__ t_son, const DevSizes *__restrict__ sz, int d,
                         const int *__restrict__ size, const int *__restrict__ lsize, int *__restrict__ dfs,
                         int *__restrict__ lfirst, int *__restrict__ l_final, int *__restrict__ l_dfs)
{
    int v = sz->lvl[d] + blockIdx.x * blockDim.x + threadIdx.x;
    if (sz->overflow || v >= sz->lvl[d + 1]) return;
    int r = dfs[v], lf = lfirst[v];
    int c0 = t_son[2 * v], c1 = t_son[2 * v + 1];
    int s0 = c0 < 0 ? 1 : size[c0];
    int ls0 = c0 < 0 ? 1 : lsize[c0];
    int dd[2] = {r + 1, r + 1 + s0};
    int ll[2] = {lf, lf + ls0};
    int cs[2] = {c0, c1};
    for (int k = 0; k < 2; k++) {
        int c = cs[k];
        if (c < 0) {
            l_final[-c - 1] = ll[k];
            l_dfs[-c - 1] = dd[k];
        } else {
            dfs[c] = dd[k];
            lfirst[c] = ll[k];
        }
    }
}

// the roots (temporary ids 0 .. nbox-1, one per box, in box order) take consecutive
// depth-first ranges: dfs and lfirst of root b = cells and leaves of the earlier roots
__global__ void k_root_starts(int nbox, const DevSizes *__restrict__ sz, const int *__restrict__ size,
                              const int *__restrict__ lsize, int *__restrict__ dfs, int *__restrict__ lfirst,
                              int *__restrict__ box_root)
{
    if (blockIdx.x != 0 || threadIdx.x != 0 || sz->overflow) return;
    const int nleaf = sz->nleaf;
    int d = 0, lf = 0;
    for (int b = 0; b < nbox; b++) {
        dfs[b] = d;
        lfirst[b] = lf;
        box_root[b] = nleaf + d - lf;
        d += size[b];
        lf += lsize[b];
    }
}


__global__ void k_emit_nodes(const DevSizes *__restrict__ sz, const int *__restrict__ dfs,
                             const int *__restrict__ lfirst, const int *__restrict__ size, TreeTmp T,
                             const int *__restrict__ l_final, CellGeom *__restrict__ geom, CellInfo *__restrict__ info,
                             int *__restrict__ level_nodes, const int64_t *__restrict__ box_off, int nbox,
                             int *__restrict__ cell_box)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int nleaf = sz->nleaf;
    if (sz->overflow || v >= sz->nnode) return;
    int f = nleaf + dfs[v] - lfirst[v];
    CellGeom g;
    for (int d = 0; d < 3; d++) {
        double lo = T.t_lo[d * T.tcap + v], hi = T.t_hi[d * T.tcap + v];
        g.c[d] = 0.5 * (hi + lo);
        g.w[d] = hi - lo;
    }
    geom[f] = g;
    CellInfo I;
    I.npart = T.t_count[v];
    I.ipart = T.t_start[v];
    int cs[2];
    for (int k = 0; k < 2; k++) {
        int c = T.t_son[2 * v + k];
        cs[k] = c < 0 ? l_final[-c - 1] : nleaf + dfs[c] - lfirst[c];
    }
    I.son0 = cs[0];
    I.son1 = cs[1];
    I.dfs = dfs[v];
    I.dsize = size[v];
    int p = T.t_parent[v];
    I.parent = p < 0 ? -1 : nleaf + dfs[p] - lfirst[p];
    I.depth = depth_of(T.t_depth[v]);
    info[f] = I;
    level_nodes[v] = f;
    // nodes hold more than maxleaf >= 2 particles, so their first slot names the box
    cell_box[f] = box_of_slot(box_off, nbox, I.ipart);
}

__global__ void k_emit_leaves(const DevSizes *__restrict__ sz, const int *__restrict__ l_final,
                              const int *__restrict__ l_dfs, const int *__restrict__ dfs,
                              const int *__restrict__ lfirst, TreeTmp T, CellGeom *__restrict__ geom,
                              CellInfo *__restrict__ info, const int64_t *__restrict__ box_off, int nbox,
                              int *__restrict__ cell_box)
{
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int nleaf = sz->nleaf;
    if (sz->overflow || l >= nleaf) return;
    int f = l_final[l];
    CellGeom g;
    for (int d = 0; d < 3; d++) {
        double lo = T.l_lo[d * T.lcap + l], hi = T.l_hi[d * T.lcap + l];
        g.c[d] = 0.5 * (hi + lo);
        g.w[d] = hi - lo;
    }
    geom[f] = g;
    CellInfo I;
    I.npart = T.l_count[l];
    I.ipart = T.l_start[l];
    I.son0 = I.son1 = -1;
    I.dfs = l_dfs[l];
    I.dsize = 1;
    int p = T.l_parent[l];
    I.parent = nleaf + dfs[p] - lfirst[p];
    I.depth = depth_of(T.t_depth[p]) + 1;
    info[f] = I;
    // a leaf may be empty: its parent's first slot names the box (every box holds a particle)
    cell_box[f] = box_of_slot(box_off, nbox, T.t_start[p]);
}

__global__ void k_init_particles(int64_t n, const double *__restrict__ pos, Particles p)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p.x[i] = pos[3 * i];
    p.y[i] = pos[3 * i + 1];
    p.z[i] = pos[3 * i + 2];
    p.idx[i] = (int)i;
}

// ---------------------------------------------------------------------------- host side

static inline int blocks(int64_t n, int b) { return (int)((n + b - 1) / b); }

template <typename T>
static T *ensure(DevBuf &b, size_t count)
{
    return reinterpret_cast<T *>(grow(b, count * sizeof(T)));
}

// grow a level-order array keeping its first `keep` elements (SoA width `width`)
template <typename T>
static T *regrow(DevBuf &b, size_t oldcap, size_t newcap, int width, hipStream_t s)
{
    T *old = reinterpret_cast<T *>(b.p);
    T *nw = static_cast<T *>(dev_alloc(newcap * width * sizeof(T)));
    if (old) {
        for (int w = 0; w < width; w++)
            PH_CHECK(hipMemcpyAsync(nw + w * newcap, old + w * oldcap, oldcap * sizeof(T), hipMemcpyDeviceToDevice, s));
        PH_CHECK(hipStreamSynchronize(s));
        dev_free(old);
    }
    b.p = nw;
    b.cap = newcap * width * sizeof(T);
    return nw;
}

// exclusive scans whose temporary storage is the context's scan buffer (grown by an
// ordinary evaluation, sized by the plan for a planned one)
template <typename T>
static void exclusive_scan_t(Ctx &c, const T *in, T *out, int n)
{
    size_t bytes = 0;
    PH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, c.stream));
    void *tmp = scan_storage(c, bytes);
    PH_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, c.stream));
}
static void exclusive_scan(Ctx &c, const int *in, int *out, int n) { exclusive_scan_t<int>(c, in, out, n); }
static void exclusive_scan_u64(Ctx &c, const uint64_t *in, uint64_t *out, int n)
{
    exclusive_scan_t<uint64_t>(c, in, out, n);
}

void upload_positions(Ctx &c, const double *d_pos_aos)
{
    int64_t n = c.n;
    Particles p{static_cast<double *>(ctx_buf(c, c.px, sizeof(double) * n, "particles")),
                static_cast<double *>(ctx_buf(c, c.py, sizeof(double) * n, "particles")),
                static_cast<double *>(ctx_buf(c, c.pz, sizeof(double) * n, "particles")),
                static_cast<int *>(ctx_buf(c, c.pidx, sizeof(int) * n, "particles"))};
    if (n > 0) k_init_particles<<<blocks(n, 256), 256, 0, c.stream>>>(n, d_pos_aos, p);
}

// Tree buffers: grown to the build's needs (an ordinary evaluation), or present at the
// plan's capacities (a planned evaluation allocates nothing: `need` only checks).
template <typename T>
static T *tree_buf(Ctx &c, DevBuf &b, size_t count, const char *what)
{
    return static_cast<T *>(ctx_buf(c, b, count * sizeof(T), what));
}

// The kd tree, level by level, all counts on the device (DevSizes).  An ordinary
// evaluation reads each level's child counts back (one copy per level, overlapped with the
// partition) to grow the temporaries and to stop after the last level; a planned one runs
// plan.levels levels at the plan's capacities without reading anything.
void build_tree(Ctx &c)
{
    const bool planned = c.plan.active;
    const int64_t n = c.n;
    const int maxleaf = c.prm.maxleaf;
    hipStream_t st = c.stream;
    DevSizes *sz = dptr<DevSizes>(c.sizes);
    need(c.sizes, sizeof(DevSizes), "sizes");
    Particles A{dptr<double>(c.px), dptr<double>(c.py), dptr<double>(c.pz), dptr<int>(c.pidx)};
    Particles B{tree_buf<double>(c, c.qx, n, "particles"), tree_buf<double>(c, c.qy, n, "particles"),
                tree_buf<double>(c, c.qz, n, "particles"), tree_buf<int>(c, c.qidx, n, "particles")};

    TreeTmp T;
    auto alloc_nodes = [&](int newcap, int oldcap) {
        T.t_start = regrow<int>(c.t_start, oldcap, newcap, 1, st);
        T.t_count = regrow<int>(c.t_count, oldcap, newcap, 1, st);
        T.t_son = regrow<int>(c.t_son, 2 * (size_t)oldcap, 2 * (size_t)newcap, 1, st);
        T.t_depth = regrow<int>(c.t_depth, oldcap, newcap, 1, st);
        T.t_parent = regrow<int>(c.t_parent, oldcap, newcap, 1, st);
        T.t_split = regrow<double>(c.t_split, oldcap, newcap, 1, st);
        T.t_lo = regrow<double>(c.t_lo, oldcap, newcap, 3, st);
        T.t_hi = regrow<double>(c.t_hi, oldcap, newcap, 3, st);
        T.tcap = c.tcap = newcap;
    };
    auto alloc_leaves = [&](int newcap, int oldcap) {
        T.l_start = regrow<int>(c.l_start, oldcap, newcap, 1, st);
        T.l_count = regrow<int>(c.l_count, oldcap, newcap, 1, st);
        T.l_parent = regrow<int>(c.l_parent, oldcap, newcap, 1, st);
        T.l_lo = regrow<double>(c.l_lo, oldcap, newcap, 3, st);
        T.l_hi = regrow<double>(c.l_hi, oldcap, newcap, 3, st);
        T.lcap = c.lcap = newcap;
    };
    // level-order temporaries: keep the capacity of earlier builds, start at ~4N/maxleaf (an
    // ordinary evaluation) or at the plan's capacities (allocated by the first planned one)
    const int want = (int)std::max<int64_t>(std::max(1024, 2 * c.nbox + 16), 4 * n / std::max(1, maxleaf) + 16);
    const int want_nodes = planned ? c.plan.tcap : want, want_leaves = planned ? c.plan.lcap : want + 1;
    const bool may_grow = !planned || c.plan.alloc;
    if (!may_grow && (c.tcap < want_nodes || c.lcap < want_leaves))
        throw std::runtime_error("tree temporaries are smaller than the plan");
    if (may_grow && c.tcap < want_nodes) {
        alloc_nodes(want_nodes, 0);
    } else {
        T.t_start = dptr<int>(c.t_start); T.t_count = dptr<int>(c.t_count); T.t_son = dptr<int>(c.t_son);
        T.t_depth = dptr<int>(c.t_depth); T.t_parent = dptr<int>(c.t_parent); T.t_split = dptr<double>(c.t_split);
        T.t_lo = dptr<double>(c.t_lo); T.t_hi = dptr<double>(c.t_hi); T.tcap = c.tcap;
    }
    if (may_grow && c.lcap < want_leaves) {
        alloc_leaves(want_leaves, 0);
    } else {
        T.l_start = dptr<int>(c.l_start); T.l_count = dptr<int>(c.l_count); T.l_parent = dptr<int>(c.l_parent);
        T.l_lo = dptr<double>(c.l_lo); T.l_hi = dptr<double>(c.l_hi); T.lcap = c.lcap;
    }

    const int nbox = c.nbox;
    need(c.d_box_off, sizeof(int64_t) * (size_t)(nbox + 1), "boxes");
    need(c.d_box_len, sizeof(double) * (size_t)nbox, "boxes");
    k_tree_roots<<<blocks(nbox, 256), 256, 0, st>>>(nbox, dptr<int64_t>(c.d_box_off), dptr<double>(c.d_box_len), T,
                                                     sz);

    int *pin = reinterpret_cast<int *>(c.pinned.p);
    std::vector<int> lvl = {0, nbox};  // host copies of the level ranges (ordinary evaluation)
    std::vector<int> ldo = {0, 0};
    const int cextra = (int)((n + kChunk - 1) / kChunk);
    int depth = 0;
    for (;; depth++) {
        int Scap;
        if (planned) {
            if (depth >= c.plan.levels) break;
            Scap = c.plan.scap;
        } else {
            if (lvl[depth + 1] == lvl[depth]) break;
            if (depth >= kMaxDepth) throw std::runtime_error("tree deeper than the supported number of levels");
            Scap = lvl[depth + 1] - lvl[depth];
        }
        const int axis = depth % 3;
        // sum over segments of ceil(count / kChunk) <= S + ceil(n / kChunk)
        const int C = Scap + cextra;
        int *seg_chunk = tree_buf<int>(c, c.seg_chunk, Scap + 1, "tree level");
        int *nch = tree_buf<int>(c, c.seg_flag, 2 * (size_t)(Scap + 1), "tree level");
        int *flag = nch + (Scap + 1);
        double *partial = tree_buf<double>(c, c.chunk_part, C + 1, "tree level");
        int *rcnt = tree_buf<int>(c, c.chunk_rcnt, 2 * (size_t)(C + 1), "tree level");
        int *rpre = rcnt + (C + 1);
        int *nleft = tree_buf<int>(c, c.seg_nleft, Scap + 1, "tree level");
        uint64_t *kinds = tree_buf<uint64_t>(c, c.seg_nn, Scap + 1, "tree level");
        uint64_t *kind_off = tree_buf<uint64_t>(c, c.seg_non, Scap + 1, "tree level");
        const double *coord = axis == 0 ? A.x : (axis == 1 ? A.y : A.z);
        LevelArgs a{coord, T.t_start, T.t_count, T.t_depth, sz, depth, Scap, C, seg_chunk, axis};
        k_nchunks<<<blocks(Scap + 1, 256), 256, 0, st>>>(a, nch);
        exclusive_scan(c, nch, seg_chunk, Scap + 1);
        const int wb = blocks(C, kWavesPerBlock);
        k_chunk_sum<<<wb, 256, 0, st>>>(a, partial);
        k_mean<<<blocks(Scap, kWavesPerBlock), 256, 0, st>>>(a, partial, T.t_split);
        k_chunk_count<<<blocks(C + 1, kWavesPerBlock), 256, 0, st>>>(a, T.t_split, rcnt);
        exclusive_scan(c, rcnt, rpre, C + 1);
        k_seg_counts<<<blocks(Scap + 1, 256), 256, 0, st>>>(a, rpre, maxleaf, nleft, flag, kinds);
        exclusive_scan_u64(c, kinds, kind_off, Scap + 1);
        if (!planned) {
            PH_CHECK(hipMemcpyAsync(pin, kind_off + Scap, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
            PH_CHECK(hipEventRecord(c.counts_ready, st));
        }
        // the partition runs while the host reads the counts
        k_scatter<<<wb, 256, 0, st>>>(a, T.t_split, rpre, nleft, flag, A, B);
        std::swap(A, B);  // the children's particles are in B now
        if (!planned) {
            PH_CHECK(hipEventSynchronize(c.counts_ready));
            const uint64_t tot = *reinterpret_cast<const uint64_t *>(pin);
            const int new_nodes = (int)(tot >> 32), new_leaves = (int)(uint32_t)tot;
            const int ntmp = lvl[depth + 1], nleaf = ldo[depth + 1];
            if (ntmp + new_nodes > T.tcap) {
                alloc_nodes(std::max(2 * T.tcap, ntmp + new_nodes + 1024), T.tcap);
                a.t_start = T.t_start;
                a.t_count = T.t_count;
                a.t_depth = T.t_depth;
            }
            if (nleaf + new_leaves > T.lcap) alloc_leaves(std::max(2 * T.lcap, nleaf + new_leaves + 1024), T.lcap);
            lvl.push_back(ntmp + new_nodes);
            ldo.push_back(nleaf + new_leaves);
        }
        k_make_children<<<blocks(Scap, 256), 256, 0, st>>>(a, T, axis, maxleaf, nleft, kind_off);
        // the planned capacities (the arrays may be larger, from earlier evaluations)
        k_level_step<<<1, 1, 0, st>>>(sz, depth, kind_off, planned ? c.plan.tcap : T.tcap,
                                      planned ? c.plan.lcap : T.lcap, planned ? c.plan.scap : INT32_MAX);
    }
    const int levels = depth;  // levels split: nodes of depths 0 .. levels-1
    k_tree_sizes<<<1, 1, 0, st>>>(sz, levels);
    // launch bounds: the counts themselves (ordinary) or the plan's capacities
    const int ntmp_b = planned ? c.plan.tcap : lvl[levels];
    const int nleaf_b = planned ? c.plan.lcap : ldo[levels + 1];
    auto level_bound = [&](int d) { return planned ? c.plan.scap : lvl[d + 1] - lvl[d]; };

    // leaves at odd depth sit in the second buffer
    {
        Particles P{dptr<double>(c.px), dptr<double>(c.py), dptr<double>(c.pz), dptr<int>(c.pidx)};
        Particles Q{dptr<double>(c.qx), dptr<double>(c.qy), dptr<double>(c.qz), dptr<int>(c.qidx)};
        if (nleaf_b > 0) k_gather_leaves<<<blocks(nleaf_b, kWavesPerBlock), 256, 0, st>>>(T, sz, Q, P);
    }

    // subtree sizes, bottom-up
    int *size = tree_buf<int>(c, c.t_size, T.tcap, "tree numbering");
    int *lsize = tree_buf<int>(c, c.t_lsize, T.tcap, "tree numbering");
    int *dfs = tree_buf<int>(c, c.t_dfs, T.tcap, "tree numbering");
    int *lfirst = tree_buf<int>(c, c.t_lfirst, T.tcap, "tree numbering");
    int *l_final = tree_buf<int>(c, c.l_final, 2 * (size_t)T.lcap, "tree numbering");
    int *l_dfs = l_final + T.lcap;
    for (int d = levels - 1; d >= 0; d--)
        k_sizes<<<blocks(level_bound(d), 256), 256, 0, st>>>(T.t_son, sz, d, size, lsize);
    int *box_root = tree_buf<int>(c, c.d_box_root, nbox, "box roots");
    k_root_starts<<<1, 1, 0, st>>>(nbox, sz, size, lsize, dfs, lfirst, box_root);
    for (int d = 0; d < levels; d++)
        k_number<<<blocks(level_bound(d), 256), 256, 0, st>>>(T.t_son, sz, d, size, lsize, dfs, lfirst, l_final,
                                                              l_dfs);
    const int ncell_b = ntmp_b + nleaf_b;
    CellGeom *geom = tree_buf<CellGeom>(c, c.geom, ncell_b, "cells");
    CellInfo *info = tree_buf<CellInfo>(c, c.info, ncell_b, "cells");
    int *level_nodes = tree_buf<int>(c, c.level_nodes, std::max(1, ntmp_b), "cells");
    int *cell_box = tree_buf<int>(c, c.cell_box, ncell_b, "cells");
    const int64_t *box_off = dptr<int64_t>(c.d_box_off);
    k_emit_nodes<<<blocks(ntmp_b, 256), 256, 0, st>>>(sz, dfs, lfirst, size, T, l_final, geom, info, level_nodes,
                                                       box_off, nbox, cell_box);
    k_emit_leaves<<<blocks(nleaf_b, 256), 256, 0, st>>>(sz, l_final, l_dfs, dfs, lfirst, T, geom, info, box_off, nbox,
                                                        cell_box);
    PH_CHECK(hipGetLastError());

    if (!planned) {
        // host copies: cell counts, level ranges
        const int nleaf = ldo[levels + 1], ntmp = lvl[levels];
        c.nleaf = nleaf;
        c.nnode = ntmp;
        c.ncell = nleaf + ntmp;
        int nlevel = 0;
        while (nlevel < levels && lvl[nlevel + 1] > lvl[nlevel]) nlevel++;
        c.nlevel = nlevel;
        c.maxdepth = nlevel;  // leaves sit at most at depth nlevel
        c.level_off.assign(lvl.begin(), lvl.end());
        Measure &m = c.meas;
        m.levels = levels;
        m.ntmp = ntmp;
        m.nleaf = nleaf;
        m.scap = 0;
        for (int d = 0; d < levels; d++) m.scap = std::max(m.scap, lvl[d + 1] - lvl[d]);
    }
}

}  // namespace photons
