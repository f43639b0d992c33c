// tree.hip -- kd tree of the short-range force on the GPU.
//
// Same tree as build_kdtree/bksort_inplace/center_kdtree (src/fmm.c:30-177): a node
// is split at the mean of its particles' coordinate along the axis of its depth
// (x at the root, cycling), particles <= mean go left; a child with at most maxleaf
// particles is a leaf (possibly empty), boxes are cut at the split planes from
// the domain box [0, L)^3.  When no particle lies above the mean the last one goes
// right, as the reference partition does.
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

// C is an upper bound of the level's chunk count (S + ceil(n / kChunk)), so the host need
// not read the count back; the kernels take the exact count, seg_chunk[S], from the device
struct LevelArgs {
    const double *coord;  // the split axis of this level
    const int *t_start, *t_count;
    int seg0, S, C;
    const int *seg_chunk;  // [S+1], exclusive scan of the chunks per segment
    int axis;              // 0, 1, 2: which of x, y, z coord is
};

__device__ __forceinline__ double axis_value(int axis, double x, double y, double z)
{
    return axis == 0 ? x : (axis == 1 ? y : z);
}

__global__ void __launch_bounds__(256) k_chunk_sum(LevelArgs a, double *__restrict__ partial)
{
    int c = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    int lane = threadIdx.x % kWave;
    if (c >= a.seg_chunk[a.S]) return;
    int s = find_seg(a.seg_chunk, a.S, c);
    int st = a.t_start[a.seg0 + s], ct = a.t_count[a.seg0 + s];
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

__global__ void k_nchunks(const int *__restrict__ cnt, int seg0, int S, int *__restrict__ out)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > S) return;
    out[s] = (s == S) ? 0 : (cnt[seg0 + s] + kChunk - 1) / kChunk;
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
    if (s >= a.S) return;
    double v = 0.0;
    for (int c = a.seg_chunk[s] + lane; c < a.seg_chunk[s + 1]; c += kWave) v += partial[c];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    if (lane == 0) t_split[a.seg0 + s] = v / (double)a.t_count[a.seg0 + s];
}

__global__ void __launch_bounds__(256) k_chunk_count(LevelArgs a, const double *__restrict__ t_split,
                                                     int *__restrict__ rcnt)
{
    int c = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    int lane = threadIdx.x % kWave;
    if (c > a.C) return;
    if (c >= a.seg_chunk[a.S]) {  // zeros up to the bound: the exclusive scan ends with the total
        if (lane == 0) rcnt[c] = 0;
        return;
    }
    int s = find_seg(a.seg_chunk, a.S, c);
    int st = a.t_start[a.seg0 + s], ct = a.t_count[a.seg0 + s];
    double split = t_split[a.seg0 + s];
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
    if (s > a.S) return;
    if (s == a.S) {  // the scan's last entry is the total
        kinds[s] = 0;
        return;
    }
    int ct = a.t_count[a.seg0 + s];
    int nr = rpre[a.seg_chunk[s + 1]] - rpre[a.seg_chunk[s]];
    int nl = ct - nr, f = 0;
    if (nr == 0) {  // nothing above the mean: the last particle goes right
        nl = ct - 1;
        f = 1;
    } else if (nl == 0) {  // nothing at or below it (rounding): the first goes left
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
    if (c >= a.seg_chunk[a.S]) return;
    int s = find_seg(a.seg_chunk, a.S, c);
    int st = a.t_start[a.seg0 + s], ct = a.t_count[a.seg0 + s];
    int ci = c - a.seg_chunk[s];
    int b = st + ci * kChunk;
    int e = min(b + kChunk, st + ct);
    double split = t_split[a.seg0 + s];
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

__global__ void k_make_children(LevelArgs a, TreeTmp T, int axis, int maxleaf, const int *__restrict__ nleft,
                                const uint64_t *__restrict__ kind_off, int node_base, int leaf_base)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.S) return;
    int v = a.seg0 + s;
    int st = T.t_start[v], ct = T.t_count[v];
    int cnt[2] = {nleft[s], ct - nleft[s]};
    int beg[2] = {st, st + nleft[s]};
    double split = T.t_split[v];
    int depth = T.t_depth[v];
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
            T.t_start[id] = beg[k];
            T.t_count[id] = cnt[k];
            T.t_depth[id] = depth + 1;
            T.t_parent[id] = v;
            for (int d = 0; d < 3; d++) {
                T.t_lo[d * T.tcap + id] = lo[d];
                T.t_hi[d * T.tcap + id] = hi[d];
            }
            T.t_son[2 * v + k] = id;
        } else {
            int id = nextl++;
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

// The levels scatter alternately into the two particle buffers and never copy back:
// the particles of a leaf at depth D are left in buffer D % 2.  This pass moves the
// leaves held by the second buffer into the first, one wavefront per leaf.
__global__ void __launch_bounds__(256) k_gather_leaves(TreeTmp T, int nleaf, Particles src, Particles dst)
{
    int l = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    int lane = threadIdx.x % kWave;
    // leaf depth = parent depth + 1; an even leaf depth is already in the first buffer
    if (l >= nleaf || (T.t_depth[T.l_parent[l]] & 1) == 1) return;
    int b = T.l_start[l], e = b + T.l_count[l];
    for (int i = b + lane; i < e; i += kWave) {
        dst.x[i] = src.x[i];
        dst.y[i] = src.y[i];
        dst.z[i] = src.z[i];
        dst.idx[i] = src.idx[i];
    }
}

// ---- final numbering

__global__ void k_sizes(const int *__restrict__ t_son, int b, int e, int *__restrict__ size, int *__restrict__ lsize)
{
    int v = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= e) return;
    int sz = 1, ls = 0;
    for (int k = 0; k < 2; k++) {
        int c = t_son[2 * v + k];
        if (c < 0) {
            sz += 1;
            ls += 1;
        } else {
            sz += size[c];
            ls += lsize[c];
        }
    }
    size[v] = sz;
    lsize[v] = ls;
}

__global__ void k_number(const int *__restrict__ t_son, int b, int e, const int *__restrict__ size,
                         const int *__restrict__ lsize, int *__restrict__ dfs, int *__restrict__ lfirst,
                         int *__restrict__ l_final, int *__restrict__ l_dfs)
{
    int v = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= e) return;
    int d = dfs[v], lf = lfirst[v];
    int c0 = t_son[2 * v], c1 = t_son[2 * v + 1];
    int s0 = c0 < 0 ? 1 : size[c0];
    int ls0 = c0 < 0 ? 1 : lsize[c0];
    int dd[2] = {d + 1, d + 1 + s0};
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
__global__ void k_root_starts(int nbox, int nleaf, const int *__restrict__ size, const int *__restrict__ lsize,
                              int *__restrict__ dfs, int *__restrict__ lfirst, int *__restrict__ box_root)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int d = 0, lf = 0;
    for (int b = 0; b < nbox; b++) {
        dfs[b] = d;
        lfirst[b] = lf;
        box_root[b] = nleaf + d - lf;
        d += size[b];
        lf += lsize[b];
    }
}

// box owning particle slot i: the last b with box_off[b] <= i
__device__ __forceinline__ int box_of_slot(const int64_t *__restrict__ box_off, int nbox, int64_t i)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (box_off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void k_emit_nodes(int ntmp, int nleaf, const int *__restrict__ dfs, const int *__restrict__ lfirst,
                             const int *__restrict__ size, TreeTmp T, const int *__restrict__ l_final,
                             CellGeom *__restrict__ geom, CellInfo *__restrict__ info, int *__restrict__ level_nodes,
                             const int64_t *__restrict__ box_off, int nbox, int *__restrict__ cell_box)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= ntmp) return;
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
    I.depth = T.t_depth[v];
    info[f] = I;
    level_nodes[v] = f;
    // nodes hold more than maxleaf >= 2 particles, so their first slot names the box
    cell_box[f] = box_of_slot(box_off, nbox, I.ipart);
}

__global__ void k_emit_leaves(int nl, int nleaf, const int *__restrict__ l_final, const int *__restrict__ l_dfs,
                              const int *__restrict__ dfs, const int *__restrict__ lfirst, TreeTmp T,
                              CellGeom *__restrict__ geom, CellInfo *__restrict__ info,
                              const int64_t *__restrict__ box_off, int nbox, int *__restrict__ cell_box)
{
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nl) return;
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
    I.depth = T.t_depth[p] + 1;
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
    T *nw = nullptr;
    PH_CHECK(hipMalloc(&nw, newcap * width * sizeof(T)));
    if (old) {
        for (int w = 0; w < width; w++)
            PH_CHECK(hipMemcpyAsync(nw + w * newcap, old + w * oldcap, oldcap * sizeof(T), hipMemcpyDeviceToDevice, s));
        PH_CHECK(hipStreamSynchronize(s));
        PH_CHECK(hipFree(old));
    }
    b.p = nw;
    b.cap = newcap * width * sizeof(T);
    return nw;
}

static void exclusive_scan(Ctx &c, const int *in, int *out, int n)
{
    size_t bytes = 0;
    PH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, c.stream));
    void *tmp = grow(c.scan_tmp, bytes);
    PH_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, c.stream));
}

void upload_positions(Ctx &c, const double *d_pos_aos)
{
    int64_t n = c.n;
    Particles p{ensure<double>(c.px, n), ensure<double>(c.py, n), ensure<double>(c.pz, n), ensure<int>(c.pidx, n)};
    if (n > 0) k_init_particles<<<blocks(n, 256), 256, 0, c.stream>>>(n, d_pos_aos, p);
}

void build_tree(Ctx &c)
{
    const int64_t n = c.n;
    const int maxleaf = c.prm.maxleaf;
    hipStream_t st = c.stream;
    Particles A{dptr<double>(c.px), dptr<double>(c.py), dptr<double>(c.pz), dptr<int>(c.pidx)};
    Particles B{ensure<double>(c.qx, n), ensure<double>(c.qy, n), ensure<double>(c.qz, n), ensure<int>(c.qidx, n)};

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
    // level-order temporaries: keep the capacity of earlier builds, start at ~4N/maxleaf
    const int want = (int)std::max<int64_t>(std::max(1024, 2 * c.nbox + 16), 4 * n / std::max(1, maxleaf) + 16);
    if (c.tcap < want) {
        alloc_nodes(want, 0);
    } else {
        T.t_start = dptr<int>(c.t_start); T.t_count = dptr<int>(c.t_count); T.t_son = dptr<int>(c.t_son);
        T.t_depth = dptr<int>(c.t_depth); T.t_parent = dptr<int>(c.t_parent); T.t_split = dptr<double>(c.t_split);
        T.t_lo = dptr<double>(c.t_lo); T.t_hi = dptr<double>(c.t_hi); T.tcap = c.tcap;
    }
    if (c.lcap < want + 1) {
        alloc_leaves(want + 1, 0);
    } else {
        T.l_start = dptr<int>(c.l_start); T.l_count = dptr<int>(c.l_count); T.l_parent = dptr<int>(c.l_parent);
        T.l_lo = dptr<double>(c.l_lo); T.l_hi = dptr<double>(c.l_hi); T.lcap = c.lcap;
    }

    // roots: one per box, temporary ids 0 .. nbox-1, boxes [0, L_b]^3
    const int nbox = c.nbox;
    {
        std::vector<int> start(nbox), cnt(nbox), zero(nbox, 0), parent(nbox, -1);
        std::vector<double> lo(3 * (size_t)nbox, 0.0), hi(3 * (size_t)nbox);
        for (int b = 0; b < nbox; b++) {
            start[b] = (int)c.box_off[b];
            cnt[b] = (int)(c.box_off[b + 1] - c.box_off[b]);
            for (int d = 0; d < 3; d++) hi[d * (size_t)nbox + b] = c.box_len[b];
        }
        const size_t ib = sizeof(int) * nbox, db = sizeof(double) * nbox;
        PH_CHECK(hipMemcpyAsync(T.t_start, start.data(), ib, hipMemcpyHostToDevice, st));
        PH_CHECK(hipMemcpyAsync(T.t_count, cnt.data(), ib, hipMemcpyHostToDevice, st));
        PH_CHECK(hipMemcpyAsync(T.t_depth, zero.data(), ib, hipMemcpyHostToDevice, st));
        PH_CHECK(hipMemcpyAsync(T.t_parent, parent.data(), ib, hipMemcpyHostToDevice, st));
        for (int d = 0; d < 3; d++) {
            PH_CHECK(hipMemcpyAsync(T.t_lo + d * T.tcap, lo.data() + d * (size_t)nbox, db, hipMemcpyHostToDevice, st));
            PH_CHECK(hipMemcpyAsync(T.t_hi + d * T.tcap, hi.data() + d * (size_t)nbox, db, hipMemcpyHostToDevice, st));
        }
        PH_CHECK(hipStreamSynchronize(st));
    }

    int *pin = reinterpret_cast<int *>(c.pinned.p);
    std::vector<int> lvl = {0, nbox};  // temp node ids of each level
    int ntmp = nbox, nleaf = 0;
    int depth = 0;
    std::vector<int> seg_chunk_h;
    while (lvl[depth + 1] > lvl[depth]) {
        int seg0 = lvl[depth], S = lvl[depth + 1] - lvl[depth];
        int axis = depth % 3;
        int *seg_chunk = ensure<int>(c.seg_chunk, S + 1);
        {
            int *nch = ensure<int>(c.seg_nn, S + 1);
            k_nchunks<<<blocks(S + 1, 256), 256, 0, st>>>(T.t_count, seg0, S, nch);
            exclusive_scan(c, nch, seg_chunk, S + 1);
        }
        // sum over segments of ceil(count / kChunk) <= S + ceil(n / kChunk)
        const int C = S + (int)((c.n + kChunk - 1) / kChunk);
        const double *coord = axis == 0 ? A.x : (axis == 1 ? A.y : A.z);
        LevelArgs a{coord, T.t_start, T.t_count, seg0, S, C, seg_chunk, axis};
        double *partial = ensure<double>(c.chunk_part, C + 1);
        int *rcnt = ensure<int>(c.chunk_rcnt, 2 * (size_t)(C + 1));
        int *rpre = rcnt + (C + 1);
        int *nleft = ensure<int>(c.seg_nleft, S + 1);
        int *flag = ensure<int>(c.seg_flag, S + 1);
        int wb = blocks(C, kWavesPerBlock);
        k_chunk_sum<<<wb, 256, 0, st>>>(a, partial);
        k_mean<<<blocks(S, kWavesPerBlock), 256, 0, st>>>(a, partial, T.t_split);
        k_chunk_count<<<blocks(C + 1, kWavesPerBlock), 256, 0, st>>>(a, T.t_split, rcnt);
        exclusive_scan(c, rcnt, rpre, C + 1);
        uint64_t *kinds = ensure<uint64_t>(c.seg_nn, S + 1);
        uint64_t *kind_off = ensure<uint64_t>(c.seg_non, S + 1);
        k_seg_counts<<<blocks(S + 1, 256), 256, 0, st>>>(a, rpre, maxleaf, nleft, flag, kinds);
        {
            size_t bytes = 0;
            PH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, kinds, kind_off, S + 1, st));
            void *tmp = grow(c.scan_tmp, bytes);
            PH_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, kinds, kind_off, S + 1, st));
        }
        PH_CHECK(hipMemcpyAsync(pin, kind_off + S, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipEventRecord(c.counts_ready, st));
        // the partition runs while the host reads the counts
        k_scatter<<<wb, 256, 0, st>>>(a, T.t_split, rpre, nleft, flag, A, B);
        std::swap(A, B);  // the children's particles are in B now
        PH_CHECK(hipEventSynchronize(c.counts_ready));
        const uint64_t tot = *reinterpret_cast<const uint64_t *>(pin);
        int new_nodes = (int)(tot >> 32), new_leaves = (int)(uint32_t)tot;
        if (ntmp + new_nodes > T.tcap) {
            int ncap = std::max(2 * T.tcap, ntmp + new_nodes + 1024);
            alloc_nodes(ncap, T.tcap);
            a.t_start = T.t_start;
            a.t_count = T.t_count;
        }
        if (nleaf + new_leaves > T.lcap) alloc_leaves(std::max(2 * T.lcap, nleaf + new_leaves + 1024), T.lcap);
        k_make_children<<<blocks(S, 256), 256, 0, st>>>(a, T, axis, maxleaf, nleft, kind_off, ntmp, nleaf);
        ntmp += new_nodes;
        nleaf += new_leaves;
        lvl.push_back(ntmp);
        depth++;
        if (depth > kMaxDepth) throw std::runtime_error("tree deeper than the supported number of levels");
    }
    const int nlevels = depth;  // levels holding nodes: 0 .. depth-1
    // leaves at odd depth sit in the second buffer
    {
        Particles P{dptr<double>(c.px), dptr<double>(c.py), dptr<double>(c.pz), dptr<int>(c.pidx)};
        Particles Q{dptr<double>(c.qx), dptr<double>(c.qy), dptr<double>(c.qz), dptr<int>(c.qidx)};
        if (nleaf > 0) k_gather_leaves<<<blocks(nleaf, kWavesPerBlock), 256, 0, st>>>(T, nleaf, Q, P);
    }

    // subtree sizes, bottom-up
    int *size = ensure<int>(c.t_size, ntmp);
    int *lsize = ensure<int>(c.t_lsize, ntmp);
    int *dfs = ensure<int>(c.t_dfs, ntmp);
    int *lfirst = ensure<int>(c.t_lfirst, ntmp);
    int *l_final = ensure<int>(c.l_final, 2 * (size_t)nleaf);
    int *l_dfs = l_final + nleaf;
    for (int d = nlevels - 1; d >= 0; d--) {
        int b = lvl[d], e = lvl[d + 1];
        k_sizes<<<blocks(e - b, 256), 256, 0, st>>>(T.t_son, b, e, size, lsize);
    }
    int *box_root = ensure<int>(c.d_box_root, nbox);
    k_root_starts<<<1, 1, 0, st>>>(nbox, nleaf, size, lsize, dfs, lfirst, box_root);
    for (int d = 0; d < nlevels; d++) {
        int b = lvl[d], e = lvl[d + 1];
        k_number<<<blocks(e - b, 256), 256, 0, st>>>(T.t_son, b, e, size, lsize, dfs, lfirst, l_final, l_dfs);
    }
    c.nleaf = nleaf;
    c.nnode = ntmp;
    c.ncell = nleaf + ntmp;
    c.nlevel = nlevels;
    c.maxdepth = nlevels;  // leaves sit at most at depth nlevels
    CellGeom *geom = ensure<CellGeom>(c.geom, c.ncell);
    CellInfo *info = ensure<CellInfo>(c.info, c.ncell);
    int *level_nodes = ensure<int>(c.level_nodes, ntmp);
    int *cell_box = ensure<int>(c.cell_box, c.ncell);
    const int64_t *box_off = dptr<int64_t>(c.d_box_off);
    k_emit_nodes<<<blocks(ntmp, 256), 256, 0, st>>>(ntmp, nleaf, dfs, lfirst, size, T, l_final, geom, info,
                                                     level_nodes, box_off, nbox, cell_box);
    k_emit_leaves<<<blocks(nleaf, 256), 256, 0, st>>>(nleaf, nleaf, l_final, l_dfs, dfs, lfirst, T, geom, info,
                                                      box_off, nbox, cell_box);
    c.level_off.assign(lvl.begin(), lvl.end());
    PH_CHECK(hipGetLastError());
}

}  // namespace photons
