// kernels.hip -- P2M, M2M, M2L, L2L and the fused P2P + L2P kernel.
//
// FP64 path: every operator in double on global coordinates (the reference's
// arithmetic, src/operator.c, src/fmm.c:796-872, src/remotes.c:14-57).
// FP32-mixed path: P2M and M2M in double; M2L, L2L, L2P and P2P in float on
// relative coordinates formed in double, lengths in units of rs so that r^-(2p+1)
// stays inside the float range:
//   M'_a = M_a / rs^|a|,  L_a = L'_a / rs^(|a|+1),  acc = rs^-2 * grad L'(d/rs).
//
// P2P: one wavefront per target leaf (per chunk of 64 of its particles for leaves of
// more than 64).  Its source list is read in batches whose particles are staged in LDS
// once for the whole wavefront; every (target, source) pair of a batch goes to one lane
// by a two-pass layout that keeps all 64 lanes busy for any leaf size (LaneMap below);
// the per-lane partial sums are combined in a fixed order (deterministic results).  The
// split factor of the pair force comes from a table in LDS (split_table.h).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>

#include "engine.h"
#include "fmm_math.h"
#include "split_table.h"

namespace photons {

// a source leaf must fit one staging batch (the smallest stage holds 256 particles)
constexpr int kP2PMaxLeaf = 256;

// ---------------------------------------------------------------------------- upward

template <int P>
__global__ void __launch_bounds__(128) k_p2m(const DevSizes *__restrict__ sz, const CellGeom *__restrict__ geom,
                                             const CellInfo *__restrict__ info, const double *__restrict__ px,
                                             const double *__restrict__ py, const double *__restrict__ pz, double mass,
                                             double *__restrict__ M)
{
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (sz->overflow || l >= sz->nleaf) return;
    constexpr int NT = nterms(P);
    double m[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) m[k] = 0.0;
    const CellInfo I = info[l];
    const CellGeom g = geom[l];
    for (int i = I.ipart; i < I.ipart + I.npart; i++)
        p2m_add<P, double>(mass, px[i] - g.c[0], py[i] - g.c[1], pz[i] - g.c[2], m);
#pragma unroll
    for (int k = 0; k < NT; k++) M[(size_t)l * NT + k] = m[k];
}

// nodes of level d: level_nodes[lvl[d] .. lvl[d+1])
template <int P>
__global__ void __launch_bounds__(128) k_m2m_level(const int *__restrict__ nodes, const DevSizes *__restrict__ sz, int d,
                                                   const CellGeom *__restrict__ geom, const CellInfo *__restrict__ info,
                                                   double *__restrict__ M)
{
    int t = sz->lvl[d] + blockIdx.x * blockDim.x + threadIdx.x;
    if (sz->overflow || t >= sz->lvl[d + 1]) return;
    constexpr int NT = nterms(P);
    const int v = nodes[t];
    const CellInfo I = info[v];
    const CellGeom gv = geom[v];
    double m[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) m[k] = 0.0;
    const int sons[2] = {I.son0, I.son1};
    for (int s = 0; s < 2; s++) {
        const int c = sons[s];
        const CellGeom gc = geom[c];
        double mc[NT];
#pragma unroll
        for (int k = 0; k < NT; k++) mc[k] = M[(size_t)c * NT + k];
        m2m_add<P, double>(gv.c[0] - gc.c[0], gv.c[1] - gc.c[1], gv.c[2] - gc.c[2], mc, m);
    }
#pragma unroll
    for (int k = 0; k < NT; k++) M[(size_t)v * NT + k] = m[k];
}

template <int P>
__global__ void __launch_bounds__(128) k_scale_down(const DevSizes *__restrict__ sz, const double *__restrict__ M,
                                                    double inv_rs, float *__restrict__ Ms)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (sz->overflow || c >= sz->ncell) return;
    constexpr int NT = nterms(P);
    double s[P + 1];
    s[0] = 1.0;
#pragma unroll
    for (int e = 1; e <= P; e++) s[e] = s[e - 1] * inv_rs;
#pragma unroll
    for (int k = 0; k < NT; k++) Ms[(size_t)c * NT + k] = (float)(M[(size_t)c * NT + k] * s[mdeg(k)]);
}

// ---------------------------------------------------------------------------- M2L

// M2L work is cut into items of at most kM2LChunk list entries so that targets with
// thousands of sources (cells inside dense clusters) do not serialise one thread;
// each item writes a partial expansion and a second pass adds the partials of a
// target in item order (deterministic).
constexpr int kM2LChunk = 8;

__device__ __forceinline__ void shift_xyz(int sid, double L, double &sx, double &sy, double &sz)
{
    sx = sy = sz = 0.0;
    if (sid) {
        int k = sid <= 13 ? sid - 1 : sid;
        sx = (double)(k / 9 - 1) * L;
        sy = (double)((k / 3) % 3 - 1) * L;
        sz = (double)(k % 3 - 1) * L;
    }
}

// items per target cell, zeros up to the bound (the exclusive scan ends with the total)
__global__ void k_m2l_nitems(const int64_t *__restrict__ off, const DevSizes *__restrict__ sz, int bound,
                             int *__restrict__ nitems)
{
    int X = blockIdx.x * blockDim.x + threadIdx.x;
    if (X > bound) return;
    nitems[X] = (sz->overflow || X >= sz->ncell)
                    ? 0
                    : (int)((span_end(off, X) - span_beg(off, X) + kM2LChunk - 1) / kM2LChunk);
}

// the item count (item_off[ncell]) against the capacity of the item arrays
__global__ void k_m2l_check(const int *__restrict__ item_off, DevSizes *sz, int item_cap)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int n = sz->overflow ? 0 : item_off[sz->ncell];
    if (n > item_cap) sz->overflow |= kOvfItems;
    sz->n_items = sz->overflow ? 0 : n;
}

__global__ void k_m2l_item_targets(const int *__restrict__ item_off, const DevSizes *__restrict__ sz,
                                   int *__restrict__ item_tgt)
{
    int X = blockIdx.x * blockDim.x + threadIdx.x;
    if (sz->overflow || X >= sz->ncell) return;
    for (int j = item_off[X]; j < item_off[X + 1]; j++) item_tgt[j] = X;
}

template <int P, typename T>
__global__ void __launch_bounds__(128) k_m2l_items(const DevSizes *__restrict__ sz, const int *__restrict__ item_tgt,
                                                   const int *__restrict__ item_off, const CellGeom *__restrict__ geom,
                                                   const int64_t *__restrict__ off, const int32_t *__restrict__ src,
                                                   const T *__restrict__ M, double scale, double inv2rs,
                                                   BoxView bx, T *__restrict__ part)
{
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= sz->n_items) return;  // 0 after an overflow
    constexpr int NT = nterms(P);
    const int X = item_tgt[it];
    const int64_t e0 = span_beg(off, X) + (int64_t)(it - item_off[X]) * kM2LChunk;
    const int64_t e1 = min(span_end(off, X), e0 + kM2LChunk);
    T L[NT];
    static_for<0, NT>([&](auto k) { L[k.value] = T(0); });
    const CellGeom gX = geom[X];
    const double boxlen = bx.len_of(X);
    for (int64_t e = e0; e < e1; e++) {
        const int32_t s = src[e];
        const int B = entry_cell(s), sid = entry_shift(s);
        double sx, sy, sz;
        shift_xyz(sid, boxlen, sx, sy, sz);
        const CellGeom gB = geom[B];
        // relative coordinates formed in double (then scaled to units of rs for float)
        const T rx = (T)((gX.c[0] - (gB.c[0] + sx)) * scale);
        const T ry = (T)((gX.c[1] - (gB.c[1] + sy)) * scale);
        const T rz = (T)((gX.c[2] - (gB.c[2] + sz)) * scale);
        T f[P + 1], D[NT], m[NT];
        if constexpr (sizeof(T) == 8) {
            radial_f64<P>(sqrt(rx * rx + ry * ry + rz * rz), inv2rs, f);
        } else {
            radial_f32_unit<P>(__builtin_sqrtf(rx * rx + ry * ry + rz * rz), f);
        }
        dtensor<P, T>(rx, ry, rz, f, D);
        static_for<0, NT>([&](auto k) { m[k.value] = M[(size_t)B * NT + k.value]; });
        m2l_contract<P, T>(D, m, L);
    }
    static_for<0, NT>([&](auto k) { part[(size_t)it * NT + k.value] = L[k.value]; });
}

template <int P, typename T>
__global__ void __launch_bounds__(128) k_m2l_reduce(const DevSizes *__restrict__ sz, const int *__restrict__ item_off,
                                                    const T *__restrict__ part, T *__restrict__ Lout, int accumulate)
{
    int X = blockIdx.x * blockDim.x + threadIdx.x;
    if (sz->overflow || X >= sz->ncell) return;
    constexpr int NT = nterms(P);
    T L[NT];
    static_for<0, NT>([&](auto k) { L[k.value] = accumulate ? Lout[(size_t)X * NT + k.value] : T(0); });
    for (int it = item_off[X]; it < item_off[X + 1]; it++)
        static_for<0, NT>([&](auto k) { L[k.value] += part[(size_t)it * NT + k.value]; });
    static_for<0, NT>([&](auto k) { Lout[(size_t)X * NT + k.value] = L[k.value]; });
}

// ---------------------------------------------------------------------------- L2L

template <int P, typename T>
__global__ void __launch_bounds__(128) k_l2l_level(const int *__restrict__ nodes, const DevSizes *__restrict__ sz, int d,
                                                   const CellGeom *__restrict__ geom, const CellInfo *__restrict__ info,
                                                   double scale, T *__restrict__ L)
{
    int t = sz->lvl[d] + blockIdx.x * blockDim.x + threadIdx.x;
    if (sz->overflow || t >= sz->lvl[d + 1]) return;
    constexpr int NT = nterms(P);
    const int v = nodes[t];
    const CellInfo I = info[v];
    const CellGeom gv = geom[v];
    T lv[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) lv[k] = L[(size_t)v * NT + k];
    const int sons[2] = {I.son0, I.son1};
    for (int s = 0; s < 2; s++) {
        const int c = sons[s];
        const CellGeom gc = geom[c];
        T lc[NT];
#pragma unroll
        for (int k = 0; k < NT; k++) lc[k] = L[(size_t)c * NT + k];
        l2l_add<P, T>((T)((gc.c[0] - gv.c[0]) * scale), (T)((gc.c[1] - gv.c[1]) * scale),
                      (T)((gc.c[2] - gv.c[2]) * scale), lv, lc);
#pragma unroll
        for (int k = 0; k < NT; k++) L[(size_t)c * NT + k] = lc[k];
    }
}

// ---------------------------------------------------------------------------- offsets

__global__ void k_offsets_f32(const DevSizes *__restrict__ sz, const CellGeom *__restrict__ geom,
                              const CellInfo *__restrict__ info, const double *__restrict__ px,
                              const double *__restrict__ py, const double *__restrict__ pz, float *__restrict__ ox,
                              float *__restrict__ oy, float *__restrict__ oz)
{
    // one wavefront per leaf
    int l = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    int lane = threadIdx.x % 64;
    if (sz->overflow || l >= sz->nleaf) return;
    const CellInfo I = info[l];
    const CellGeom g = geom[l];
    for (int q = lane; q < I.npart; q += 64) {
        int i = I.ipart + q;
        ox[i] = (float)(px[i] - g.c[0]);
        oy[i] = (float)(py[i] - g.c[1]);
        oz[i] = (float)(pz[i] - g.c[2]);
    }
}

// ---------------------------------------------------------------------------- P2P + L2P

struct P2PArgs {
    const CellGeom *geom;
    const CellInfo *info;
    const int64_t *off;  // spans of the P2P list per target leaf
    const int32_t *src;
    const double *px, *py, *pz;
    const float *ox, *oy, *oz;
    const int32_t *pidx;
    const void *L;  // leaf expansions (double or float)
    double mass, soft, inv2rs, inv_rs;
    BoxView bx;
    int with_l2p;
    int accumulate;       // out += instead of out =
    int out_caller_order; // scatter through pidx into out[3*pidx]; else tree order
    double *out;          // [n][3]
    const int32_t *order; // target leaf of each launch slot (heaviest first; past nleaf: none)
    const DevSizes *sz;
    int nchunk;           // workgroups per target leaf: chunks of 64 targets (maxleaf / 64 up)
    int nslice;           // launch bound of the slices of a target leaf's list (p2p_slices of its box)
    int nbox;
    const int64_t *box_off;  // [nbox + 1]: particle slots of each box
    int64_t n;            // particles (the slices' forces: part[slice][n][3])
    double *part;
};

__device__ __forceinline__ void shift_vec(int sid, double L, double &sx, double &sy, double &sz)
{
    sx = sy = sz = 0.0;
    if (sid) {
        int k = sid <= 13 ? sid - 1 : sid;
        sx = (double)(k / 9 - 1) * L;
        sy = (double)((k / 3) % 3 - 1) * L;
        sz = (double)(k % 3 - 1) * L;
    }
}

__device__ __forceinline__ void store_acc(const P2PArgs &a, int i, double ax, double ay, double az)
{
    size_t o = 3 * (size_t)(a.out_caller_order ? a.pidx[i] : i);
    if (a.accumulate) {
        ax += a.out[o];
        ay += a.out[o + 1];
        az += a.out[o + 2];
    }
    a.out[o] = ax;
    a.out[o + 1] = ay;
    a.out[o + 2] = az;
}

// Target leaf of a P2P workgroup.  Workgroups go round-robin to the 8 XCDs, each with its
// own L2; neighbouring leaves (depth-first order) list mostly the same sources, so tiles of
// 8 consecutive leaves stay on one XCD (their source particles are fetched into that L2
// once), and the tiles rotate over the XCDs, which keeps dense clusters spread over all
// of them.  A bijection of [0, nleaf); the last partial round keeps the identity.
__device__ __forceinline__ int p2p_target(int b, int nb)
{
    constexpr int kXcd = 8, kTile = 8;
    const int full = nb / (kXcd * kTile) * (kXcd * kTile);
    if (b >= full) return b;
    const int x = b % kXcd, k = b / kXcd;
    return ((k / kTile) * kXcd + x) * kTile + k % kTile;
}

// Slices of the P2P lists of the target leaves of a box of nb particles: up to
// kP2PSliceMax, while slices x particles stay within kP2PSliceParticles (a 16^3 box: 8,
// 32^3: 4, 64^3 and more: 1)
constexpr int kP2PSliceMax = 8;
constexpr int64_t kP2PSliceParticles = (int64_t)1 << 17;

__host__ __device__ inline int p2p_slices(int64_t nb)
{
    int ns = 1;
    while (nb > 0 && ns < kP2PSliceMax && 2 * ns * nb <= kP2PSliceParticles) ns *= 2;
    return ns;  // an empty box has no leaves to slice
}

// A P2P workgroup's target leaf, its chunk of 64 of the leaf's particles and its slice of
// the leaf's list: a.nslice slices (the slowest index) of a.nchunk consecutive workgroups
// per launch slot
__device__ __forceinline__ int p2p_work(const P2PArgs &a, int &chunk, int &slice)
{
    const int nws = gridDim.x / a.nslice;  // workgroups per slice
    slice = blockIdx.x / nws;
    const int w = blockIdx.x - slice * nws;
    chunk = w % a.nchunk;
    return a.order[p2p_target(w / a.nchunk, nws / a.nchunk)];
}

// the slices of the lists of a box's target leaves
__device__ __forceinline__ int p2p_box_slices(const P2PArgs &a, int b)
{
    return p2p_slices(a.box_off[b + 1] - a.box_off[b]);
}

// the entries [b, e) of target leaf T's P2P list that slice `slice` of ns takes
__device__ __forceinline__ void p2p_slice(const P2PArgs &a, int T, int slice, int ns, int64_t &b, int64_t &e)
{
    const int64_t b0 = span_beg(a.off, T), len = span_end(a.off, T) - b0;
    b = b0 + len * slice / ns;
    e = b0 + len * (slice + 1) / ns;
}

// a target's force: stored, or its slice's part (of ns), summed by k_p2p_merge
__device__ __forceinline__ void p2p_out(const P2PArgs &a, int slice, int ns, int it, double ax, double ay,
                                        double az)
{
    if (ns == 1) {
        store_acc(a, it, ax, ay, az);
        return;
    }
    double *q = a.part + 3 * ((size_t)slice * a.n + it);
    q[0] = ax;
    q[1] = ay;
    q[2] = az;
}

// Launch order of the P2P workgroups: target leaves by decreasing work (list length times
// particle count) in bins of a factor 16, depth-first within a bin (a stable sort).
// The longest leaves of dense clusters take ~10-20 ms alone; started last they would
// leave the GPU part-idle at the end of the launch.  Keeping the depth-first order within
// a bin keeps neighbouring leaves, which list mostly the same sources, together in the
// launch (and on one XCD): a full sort by work scattered them and fetched ten times the
// bytes from HBM, bins of a factor 2 still more than twice.
// work bin of leaf slot T (launch slots past the leaves take the last bin)
__device__ __forceinline__ int work_bin(int T, const DevSizes *__restrict__ sz, const int64_t *__restrict__ off,
                                        const CellInfo *__restrict__ info)
{
    if (sz->overflow || T >= sz->nleaf) return 31;
    const uint64_t w = (uint64_t)(span_end(off, T) - span_beg(off, T)) * (uint64_t)info[T].npart;
    return (int)((uint32_t)__clzll(w | 1) >> 2);  // factor-16 bins of work, ascending = decreasing
}

// A stable counting sort of the leaves by work bin (32 bins), tiles of kOrderTile slots:
// per tile and bin the count (bin-major, so one exclusive scan gives every tile's first
// slot of each bin), then a stable scatter.  Within a wave, a slot's rank among equal
// bins comes from one ballot per bin; the waves of a tile add up in order.
constexpr int kOrderTile = 256;

__device__ __forceinline__ void tile_bins(int key, int wave, int lane, int (*wcnt)[32], int &rank)
{
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int b = 0; b < 32; b++) {
        const uint64_t m = __ballot(key == b);
        if (key == b) rank = __popcll(m & lt);
        if (lane == 0) wcnt[wave][b] = __popcll(m);
    }
}

__global__ void __launch_bounds__(kOrderTile) k_order_hist(int bound, const DevSizes *__restrict__ sz,
                                                           const int64_t *__restrict__ off,
                                                           const CellInfo *__restrict__ info, int ntile,
                                                           int *__restrict__ hist)
{
    __shared__ int wcnt[kOrderTile / 64][32];
    const int T = blockIdx.x * kOrderTile + threadIdx.x;
    const int key = T < bound ? work_bin(T, sz, off, info) : -1;
    int rank = 0;
    tile_bins(key, threadIdx.x / 64, threadIdx.x % 64, wcnt, rank);
    __syncthreads();
    if (threadIdx.x < 32) {
        int t = 0;
        for (int w = 0; w < kOrderTile / 64; w++) t += wcnt[w][threadIdx.x];
        hist[threadIdx.x * ntile + blockIdx.x] = t;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) hist[32 * ntile] = 0;  // the scan's last entry is the total
}

__global__ void __launch_bounds__(kOrderTile) k_order_scatter(int bound, const DevSizes *__restrict__ sz,
                                                              const int64_t *__restrict__ off,
                                                              const CellInfo *__restrict__ info, int ntile,
                                                              const int *__restrict__ base, int32_t *__restrict__ order)
{
    __shared__ int wcnt[kOrderTile / 64][32];
    const int T = blockIdx.x * kOrderTile + threadIdx.x;
    const int key = T < bound ? work_bin(T, sz, off, info) : -1;
    const int wave = threadIdx.x / 64;
    int rank = 0;
    tile_bins(key, wave, threadIdx.x % 64, wcnt, rank);
    __syncthreads();
    if (key < 0) return;
    int before = 0;
    for (int w = 0; w < wave; w++) before += wcnt[w][key];
    order[base[key * ntile + blockIdx.x] + before + rank] = T;
}

static void p2p_order(Ctx &c, int n)
{
    hipStream_t st = c.stream;
    const int ntile = (n + kOrderTile - 1) / kOrderTile;
    const size_t nh = 32 * (size_t)ntile + 1;
    int *hist = static_cast<int *>(ctx_buf(c, c.p2p_order_tmp, sizeof(int) * 2 * nh, "P2P order"));
    int *base = hist + nh;
    int32_t *order = static_cast<int32_t *>(ctx_buf(c, c.p2p_order, sizeof(int32_t) * (size_t)n, "P2P order"));
    const DevSizes *sz = dptr<DevSizes>(c.sizes);
    const int64_t *off = dptr<int64_t>(c.p2p_off);
    const CellInfo *info = dptr<CellInfo>(c.info);
    k_order_hist<<<ntile, kOrderTile, 0, st>>>(n, sz, off, info, ntile, hist);
    size_t bytes = 0;
    PH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, hist, base, (int)nh, st));
    void *tmp = scan_storage(c, bytes);
    PH_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, hist, base, (int)nh, st));
    k_order_scatter<<<ntile, kOrderTile, 0, st>>>(n, sz, off, info, ntile, base, order);
}

// FP32 pair force per interaction, 20 VALU instructions and one LDS read:
//   r2 = |d|^2 + 1e-30 (keeps rsqrt finite for coincident particles, whose d = 0)
//   1/max(r, eps)^3 = min(rsqrt(r2), 1/eps)^3   (the reference's softening branch)
//   G(u) = erfc(u) + 2u/sqrt(pi) exp(-u^2) from the cubic table in LDS (split_table.h)
// Only the rsqrt goes to the transcendental unit (the erfc/exp form took three: rsqrt,
// rcp and exp2, and that unit paced the loop).  Coordinates are in table units
// (u N / U per unit of r, scaled once when staged), so that r itself indexes the table;
// the particle mass and the factor the scaling leaves on the sum (the square of the
// scale) multiply it once at the end.
struct PairConsts {
    float ieps;  // 1/eps in table units, capped
    float vmax;  // the table's last row: G = 0 from there on
};

using SplitRow32 = float4;
constexpr int kSplitN32 = kSplitTab32N;

__device__ __forceinline__ void pair_f32(const PairConsts &k, const SplitRow32 *tab, float dx, float dy, float dz,
                                         float &ax, float &ay, float &az)
{
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, 1e-30f)));
    const float ri = __builtin_amdgcn_rsqf(r2);
    // min(ri, 1/eps) and min(r, vmax) on the bit patterns: all are non-negative floats,
    // whose order is that of their integer images (one v_min_i32, no NaN canonicalisation)
    const float rc = __int_as_float(min(__float_as_int(ri), __float_as_int(k.ieps)));
    const float v = __int_as_float(min(__float_as_int(r2 * ri), __float_as_int(k.vmax)));
    const SplitRow32 c = tab[(int)v];
    const float f = __builtin_amdgcn_fractf(v);
    const float g = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c.w, f, c.z), f, c.y), f, c.x);
    const float fr = rc * rc * (rc * g);
    ax = __builtin_fmaf(dx, fr, ax);
    ay = __builtin_fmaf(dy, fr, ay);
    az = __builtin_fmaf(dz, fr, az);
}

// Two interactions of one target, their steps interleaved so that each one's LDS
// table read is in flight while the other computes.
__device__ __forceinline__ void pair2_f32(const PairConsts &k, const SplitRow32 *tab, float dx, float dy, float dz,
                                          float ex, float ey, float ez, float &ax, float &ay, float &az)
{
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, 1e-30f)));
    const float s2 = __builtin_fmaf(ex, ex, __builtin_fmaf(ey, ey, __builtin_fmaf(ez, ez, 1e-30f)));
    const float ri = __builtin_amdgcn_rsqf(r2);
    const float si = __builtin_amdgcn_rsqf(s2);
    const float vr = __int_as_float(min(__float_as_int(r2 * ri), __float_as_int(k.vmax)));
    const float vs = __int_as_float(min(__float_as_int(s2 * si), __float_as_int(k.vmax)));
    const SplitRow32 cr = tab[(int)vr];
    const SplitRow32 cs = tab[(int)vs];
    const float rc = __int_as_float(min(__float_as_int(ri), __float_as_int(k.ieps)));
    const float sc = __int_as_float(min(__float_as_int(si), __float_as_int(k.ieps)));
    const float fr = __builtin_amdgcn_fractf(vr);
    const float fs = __builtin_amdgcn_fractf(vs);
    const float rc2 = rc * rc, sc2 = sc * sc;
    const float gr = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(cr.w, fr, cr.z), fr, cr.y), fr, cr.x);
    const float gs = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(cs.w, fs, cs.z), fs, cs.y), fs, cs.x);
    const float qr = rc2 * (rc * gr);
    const float qs = sc2 * (sc * gs);
    ax = __builtin_fmaf(ex, qs, __builtin_fmaf(dx, qr, ax));
    ay = __builtin_fmaf(ey, qs, __builtin_fmaf(dy, qr, ay));
    az = __builtin_fmaf(ez, qs, __builtin_fmaf(dz, qr, az));
}

// Source particles staged in LDS per batch, as float offsets from the target leaf's
// centre (3 x 4 bytes each): STAGE = 256 or 512 slots (a tuning choice; 512 is the
// default), at least 128 so the final 6 x 64 partial sums fit the same array.

// Lane layout for a target leaf of n particles (1 <= n <= 64); a = 64 / n, e = 64 - a n.
//   pass A: lane l serves target l % n and the sources q = l / n + i sA of the staged
//           batch, sA = a when e = 0, else a + 1 (the extra class a then exists only
//           for targets 0 .. e-1);
//   pass B: (e > 0 only) the m = n - e targets e .. n-1 still miss class a: each gets
//           b = 64 / m lanes, lane l < m b serving target e + l % m and the sources
//           q = a + (a + 1)(l / m + i b).
// Every (target, source) pair of the batch is computed once, and all 64 lanes run the
// same number of iterations to within one (about n * sources / 64 in total).
struct LaneMap {
    int a, e, m, b;
    int sA, kA, tA;      // pass A: stride, first source, target
    int sB, kB, tB;      // pass B: stride, first source (or a huge value: idle), target
};

__device__ __forceinline__ LaneMap lane_map(int n, int lane)
{
    LaneMap L;
    L.a = 64 / n;
    L.e = 64 - L.a * n;
    L.sA = L.e ? L.a + 1 : L.a;
    L.kA = lane / n;
    L.tA = lane % n;
    L.m = n - L.e;
    L.b = L.e ? 64 / L.m : 0;
    L.sB = (L.a + 1) * L.b;
    L.tB = L.e ? L.e + lane % L.m : 0;
    L.kB = (L.e && lane < L.m * L.b) ? L.a + (L.a + 1) * (lane / L.m) : (1 << 30);
    return L;
}

// One wavefront per target leaf.  The source list is consumed in batches: 64 list
// entries are read at once, one per lane (entry -> cell -> centre offset, formed in
// double), a wave prefix sum of their particle counts picks the leading entries that
// fit STAGE particles, and their particles are staged in LDS as float offsets
// from the target centre, all loads of a batch issued together.  The batch's entries
// stay in registers (the staging lanes read them by shuffle), so LDS holds only the
// staged particles and the split-factor table.  The pairs of the batch are then spread
// over the lanes by the two-pass layout above.  Partial sums are combined per target in a
// fixed lane order, so results are deterministic.
// Occupancy: amdgpu_waves_per_eu(6) caps the kernel at 80 VGPRs; with STAGE 256 a
// workgroup holds 7.2 KB of LDS (22 per 160 KB CU, 5.5 waves per SIMD), with STAGE 512
// 10.3 KB (15, 3.75 waves).  Measured at 256^3: 160 ms at STAGE 256 against 172 ms at
// 512 (and 165 ms at 192, 167 ms at 128: more batches per target), so 256 is the default.
template <int P, int STAGE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) k_p2p_f32(P2PArgs a)
{
    static_assert(STAGE >= kP2PMaxLeaf && STAGE % 64 == 0, "stage size: a source leaf fits one batch");
    __shared__ float stage[3 * STAGE];
    float *const sx = stage, *const sy = stage + STAGE, *const sz = stage + 2 * STAGE;
    __shared__ SplitRow32 tab[kSplitN32 + 1];
    // a workgroup per chunk of 64 of the target leaf's particles (one chunk for leaves of
    // up to 64) and slice of its list
    int chunk, slice;
    const int T = p2p_work(a, chunk, slice);
    const int lane = threadIdx.x;
    if (a.sz->overflow || T >= a.sz->nleaf) return;
    const CellInfo IT = a.info[T];
    const int c0 = 64 * chunk;
    if (c0 >= IT.npart) return;
    const int ns = p2p_box_slices(a, a.bx.cell_box[T]);
    if (slice >= ns) return;
    const int n = min(64, IT.npart - c0), i0 = IT.ipart + c0;  // this chunk's targets
    // the split-factor table into LDS (read after the first barrier below)
    for (int j = lane; j <= kSplitN32; j += 64) tab[j] = reinterpret_cast<const SplitRow32 *>(kSplitTab32)[j];
    const LaneMap L = lane_map(n, lane);
    const double lam_d = a.inv2rs * (kSplitTab32N / kSplitTab32U);  // length -> table units
    const float lam = (float)lam_d;
    const float txA = a.ox[i0 + L.tA] * lam, tyA = a.oy[i0 + L.tA] * lam,
                tzA = a.oz[i0 + L.tA] * lam;
    const float txB = a.ox[i0 + L.tB] * lam, tyB = a.oy[i0 + L.tB] * lam,
                tzB = a.oz[i0 + L.tB] * lam;
    const CellGeom gT = a.geom[T];
    const double boxlen = a.bx.len_of(T);
    PairConsts pc;
    // 1/eps in these units, capped so that its cube stays finite: with eps = 0 a particle
    // with itself (r2 = 1e-30, d = 0) then adds 0 * finite = 0, as the reference's skip
    // of that pair does; the cap only touches pairs closer than 1e-12 units
    pc.ieps = (float)fmin(1.0 / (a.soft * lam_d), 1e12);
    pc.vmax = (float)kSplitN32;
    float axA = 0.f, ayA = 0.f, azA = 0.f, axB = 0.f, ayB = 0.f, azB = 0.f;
    int64_t b1, e1;
    p2p_slice(a, T, slice, ns, b1, e1);
    for (int64_t base = b1; base < e1;) {
        const int nb = (int)min((int64_t)64, e1 - base);
        // entry `lane` of the batch: its centre offset, its particle count and the slot
        // offset of its first particle (in registers; other lanes read them by shuffle)
        int cnt = 0, first = 0;
        float ecx = 0.f, ecy = 0.f, ecz = 0.f;
        if (lane < nb) {
            const int32_t s = a.src[base + lane];
            const int S = entry_cell(s), sid = entry_shift(s);
            const CellInfo IS = a.info[S];
            const CellGeom gS = a.geom[S];
            double shx, shy, shz;
            shift_vec(sid, boxlen, shx, shy, shz);
            ecx = (float)(((gS.c[0] + shx) - gT.c[0]) * lam_d);
            ecy = (float)(((gS.c[1] + shy) - gT.c[1]) * lam_d);
            ecz = (float)(((gS.c[2] + shz) - gT.c[2]) * lam_d);
            first = IS.ipart;
            cnt = IS.npart;
        }
        int incl = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        // the counts are at most maxleaf <= STAGE, so at least one entry fits (the break
        // only keeps the loop finite if that were ever broken)
        const int ne = __popcll(__ballot(lane < nb && incl <= STAGE));
        if (ne == 0) break;
        const int tot = __shfl(incl, ne - 1, 64);
        const int eincl = lane < ne ? incl : 0x7fffffff;
        const int ebase = first - (incl - cnt);  // slot p of this entry is particle ebase + p
        // stage: slot p belongs to the first entry whose inclusive count exceeds p (the
        // shuffles run in every lane; only the loads depend on p < tot)
        constexpr int R = STAGE / 64;
        float vx[R], vy[R], vz[R];
        int js[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int p = lane + 64 * r;
            int j = 0;
#pragma unroll
            for (int s = 32; s >= 1; s >>= 1)
                if (__shfl(eincl, j + s - 1, 64) <= p) j += s;
            js[r] = j;
            const int gi = __shfl(ebase, j, 64) + p;
            if (p < tot) {
                vx[r] = a.ox[gi];
                vy[r] = a.oy[gi];
                vz[r] = a.oz[gi];
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int p = lane + 64 * r;
            const int j = js[r];
            const float cx = __shfl(ecx, j, 64), cy = __shfl(ecy, j, 64), cz = __shfl(ecz, j, 64);
            if (p < tot) {
                sx[p] = __builtin_fmaf(vx[r], lam, cx);
                sy[p] = __builtin_fmaf(vy[r], lam, cy);
                sz[p] = __builtin_fmaf(vz[r], lam, cz);
            }
        }
        __syncthreads();
        {
            // sources (q, q + s) of this lane's class while q + s < tot, then at most one
            // more; the LDS address is the only induction variable
            const int st = L.sA;
            const float *p = sx + min(L.kA, tot);  // idle lanes: start at the end
            const float *const pend = sx + max(tot - st, 0);
            for (; p < pend; p += 2 * st)
                pair2_f32(pc, tab, p[0] - txA, p[STAGE] - tyA, p[2 * STAGE] - tzA, p[st] - txA,
                          p[st + STAGE] - tyA, p[st + 2 * STAGE] - tzA, axA, ayA, azA);
            if (p < sx + tot) pair_f32(pc, tab, p[0] - txA, p[STAGE] - tyA, p[2 * STAGE] - tzA, axA, ayA, azA);
        }
        if (L.e) {
            const int st = L.sB;
            const float *p = sx + min(L.kB, tot);  // idle lanes: start at the end
            const float *const pend = sx + max(tot - st, 0);
            for (; p < pend; p += 2 * st)
                pair2_f32(pc, tab, p[0] - txB, p[STAGE] - tyB, p[2 * STAGE] - tzB, p[st] - txB,
                          p[st + STAGE] - tyB, p[st + 2 * STAGE] - tzB, axB, ayB, azB);
            if (p < sx + tot) pair_f32(pc, tab, p[0] - txB, p[STAGE] - tyB, p[2 * STAGE] - tzB, axB, ayB, azB);
        }
        base += ne;
        __syncthreads();  // the staging arrays are rewritten by the next batch
    }
    // per-target sums in a fixed order: pass A lanes t + n k, then pass B lanes
    float *red = stage;  // 6 x 64 floats, the staging array is free now
    red[lane] = axA;
    red[64 + lane] = ayA;
    red[128 + lane] = azA;
    red[192 + lane] = axB;
    red[256 + lane] = ayB;
    red[320 + lane] = azB;
    __syncthreads();
    if (lane < n) {
        const int t = lane;
        float bx = 0.f, by = 0.f, bz = 0.f;
        const int kmax = t < L.e ? L.a + 1 : L.a;
        for (int k = 0; k < kmax; k++) {
            bx += red[t + n * k];
            by += red[64 + t + n * k];
            bz += red[128 + t + n * k];
        }
        if (t >= L.e) {
            for (int j = 0; j < L.b; j++) {
                const int l = (t - L.e) + L.m * j;
                bx += red[192 + l];
                by += red[256 + l];
                bz += red[320 + l];
            }
        }
        const float m = (float)(a.mass * lam_d * lam_d);
        bx *= m;
        by *= m;
        bz *= m;
        const int it = i0 + t;
        if (a.with_l2p && slice == 0) {
            constexpr int NT = nterms(P);
            const float *Lp = reinterpret_cast<const float *>(a.L) + (size_t)T * NT;
            float Lr[NT];
#pragma unroll
            for (int q = 0; q < NT; q++) Lr[q] = Lp[q];
            const float is = (float)a.inv_rs;
            float lx, ly, lz;
            l2p_eval<P, float>(Lr, a.ox[it] * is, a.oy[it] * is, a.oz[it] * is, lx, ly, lz);
            const float is2 = is * is;
            bx += lx * is2;
            by += ly * is2;
            bz += lz * is2;
        }
        p2p_out(a, slice, ns, it, (double)bx, (double)by, (double)bz);
    }
}

// FP64 pair force as the reference forms it (src/fmm.c:834-852): d = x_s - x_t,
// r = |d|, m/eps^3 when r < eps else m/r^3, times G(r/2rs).  1/r comes from
// v_rsq_f64 and two Newton steps (r = r^2 / r, m/r^3 = m (1/r)^3): the same values
// to an ulp, without the IEEE square root and divide sequences.  G comes from the
// degree-9 table in LDS (split_table.h, 2e-15 from erfc/exp), read at
// v = min(r * vscale, N) with vscale = N / (2 rs U); pairs beyond the table (u >= 6,
// G < 1.6e-15: rare, in sparse regions) take the erfc/exp form, so that G stays
// relatively exact where the reference's is.
using SplitRow64 = double2;  // a row of the table: five of these
constexpr int kSplitN64 = kSplitTab64N;
constexpr int kSplitW64 = (kSplitTab64Deg + 1) / 2;
static_assert(kSplitTab64Deg % 2 == 1, "rows of an even number of coefficients");

__device__ __forceinline__ double split_table_f64(const SplitRow64 *tab, double v)
{
    const SplitRow64 *row = tab + kSplitW64 * (int)v;
    SplitRow64 c[kSplitW64];
#pragma unroll
    for (int k = 0; k < kSplitW64; k++) c[k] = row[k];
    const double f = __builtin_amdgcn_fract(v);
    double g = fma(c[kSplitW64 - 1].y, f, c[kSplitW64 - 1].x);
#pragma unroll
    for (int k = kSplitW64 - 2; k >= 0; k--) g = fma(fma(g, f, c[k].y), f, c[k].x);
    return g;
}

__device__ __forceinline__ void pair_f64(const SplitRow64 *tab, double dx, double dy, double dz, double m, double eps,
                                         double meps3, double vscale, double inv2rs, double &ax, double &ay,
                                         double &az)
{
    const double r2 = dx * dx + dy * dy + dz * dz;
    // floored so that coincident particles (d = 0) keep a finite 1/r and take the
    // eps branch with r = 0
    const double r2s = fmax(r2, 1e-280);
    double ri = __builtin_amdgcn_rsq(r2s);
    const double h = 0.5 * r2s;
    ri = ri * fma(-h * ri, ri, 1.5);
    ri = ri * fma(-h * ri, ri, 1.5);
    const double dr = r2 * ri;
    // a particle with itself (d = 0) adds exactly nothing, also with eps = 0, where the
    // reference skips the pair (src/fmm.c:829) and 1/r^3 would overflow
    double f = r2 == 0.0 ? 0.0 : (dr < eps ? meps3 : m * (ri * ri * ri));
    const double v = dr * vscale;
    double g = split_table_f64(tab, fmin(v, (double)kSplitN64));
    if (v >= (double)kSplitN64) g = split_factor_f64(dr * inv2rs);
    f *= g;
    ax += dx * f;
    ay += dy * f;
    az += dz * f;
}

// FP64 P2P + L2P: the batching and lane layout of k_p2p_f32, with source positions
// staged in double as the reference forms them (image body = pos + shift,
// src/remotes.c:84-86) and the reference's pair arithmetic.
template <int P>
__global__ void __launch_bounds__(64) k_p2p_f64(P2PArgs a)
{
    constexpr int STAGE = 256;
    static_assert(STAGE >= kP2PMaxLeaf, "a source leaf fits one batch");
    __shared__ double stage[3 * STAGE];
    double *const sx = stage, *const sy = stage + STAGE, *const sz = stage + 2 * STAGE;
    __shared__ double ecx[64], ecy[64], ecz[64];
    __shared__ int eincl[64], ecnt[64], eip[64];
    __shared__ SplitRow64 tab[kSplitW64 * (kSplitN64 + 1)];
    // a workgroup per chunk of 64 of the target leaf's particles and slice of its list
    int chunk, slice;
    const int T = p2p_work(a, chunk, slice);
    const int lane = threadIdx.x;
    if (a.sz->overflow || T >= a.sz->nleaf) return;
    const CellInfo IT = a.info[T];
    const int c0 = 64 * chunk;
    if (c0 >= IT.npart) return;
    const int ns = p2p_box_slices(a, a.bx.cell_box[T]);
    if (slice >= ns) return;
    const int n = min(64, IT.npart - c0), i0 = IT.ipart + c0;  // this chunk's targets
    // the split-factor table into LDS (read after the first barrier below)
    for (int j = lane; j < kSplitW64 * (kSplitN64 + 1); j += 64)
        tab[j] = reinterpret_cast<const SplitRow64 *>(kSplitTab64)[j];
    const LaneMap L = lane_map(n, lane);
    const double xA = a.px[i0 + L.tA], yA = a.py[i0 + L.tA], zA = a.pz[i0 + L.tA];
    const double xB = a.px[i0 + L.tB], yB = a.py[i0 + L.tB], zB = a.pz[i0 + L.tB];
    const double m = a.mass, eps = a.soft, meps3 = eps > 0.0 ? m / (eps * eps * eps) : 0.0;
    const double inv2rs = a.inv2rs, vscale = inv2rs * (kSplitTab64N / kSplitTab64U);
    const double boxlen = a.bx.len_of(T);
    double axA = 0.0, ayA = 0.0, azA = 0.0, axB = 0.0, ayB = 0.0, azB = 0.0;
    int64_t b1, e1;
    p2p_slice(a, T, slice, ns, b1, e1);
    for (int64_t base = b1; base < e1;) {
        const int nb = (int)min((int64_t)64, e1 - base);
        int cnt = 0;
        if (lane < nb) {
            const int32_t s = a.src[base + lane];
            const int S = entry_cell(s), sid = entry_shift(s);
            const CellInfo IS = a.info[S];
            double shx, shy, shz;
            shift_vec(sid, boxlen, shx, shy, shz);
            ecx[lane] = shx;
            ecy[lane] = shy;
            ecz[lane] = shz;
            eip[lane] = IS.ipart;
            cnt = IS.npart;
        }
        int incl = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        const int ne = __popcll(__ballot(lane < nb && incl <= STAGE));
        if (ne == 0) break;  // as in k_p2p_f32
        const int tot = __shfl(incl, ne - 1, 64);
        ecnt[lane] = cnt;
        eincl[lane] = lane < ne ? incl : 0x7fffffff;
        __syncthreads();
        // one staging round at a time (FP64 pairs cost enough that the load latency
        // is hidden by the other wavefronts; fewer live registers)
        for (int p = lane; p < tot; p += 64) {
            int j = 0;
#pragma unroll
            for (int s = 32; s >= 1; s >>= 1)
                if (eincl[j + s - 1] <= p) j += s;
            const int gi = eip[j] + p - (eincl[j] - ecnt[j]);
            sx[p] = a.px[gi] + ecx[j];
            sy[p] = a.py[gi] + ecy[j];
            sz[p] = a.pz[gi] + ecz[j];
        }
        __syncthreads();
        for (int q = L.kA; q < tot; q += L.sA)
            pair_f64(tab, sx[q] - xA, sy[q] - yA, sz[q] - zA, m, eps, meps3, vscale, inv2rs, axA, ayA, azA);
        if (L.e)
            for (int q = L.kB; q < tot; q += L.sB)
                pair_f64(tab, sx[q] - xB, sy[q] - yB, sz[q] - zB, m, eps, meps3, vscale, inv2rs, axB, ayB, azB);
        base += ne;
        __syncthreads();  // the staging and entry arrays are rewritten by the next batch
    }
    double *red = stage;  // 6 x 64 doubles, the staging array is free now
    red[lane] = axA;
    red[64 + lane] = ayA;
    red[128 + lane] = azA;
    red[192 + lane] = axB;
    red[256 + lane] = ayB;
    red[320 + lane] = azB;
    __syncthreads();
    if (lane < n) {
        const int t = lane;
        double bx = 0.0, by = 0.0, bz = 0.0;
        const int kmax = t < L.e ? L.a + 1 : L.a;
        for (int k = 0; k < kmax; k++) {
            bx += red[t + n * k];
            by += red[64 + t + n * k];
            bz += red[128 + t + n * k];
        }
        if (t >= L.e) {
            for (int j = 0; j < L.b; j++) {
                const int l = (t - L.e) + L.m * j;
                bx += red[192 + l];
                by += red[256 + l];
                bz += red[320 + l];
            }
        }
        const int it = i0 + t;
        if (a.with_l2p && slice == 0) {
            constexpr int NT = nterms(P);
            const double *Lp = reinterpret_cast<const double *>(a.L) + (size_t)T * NT;
            double Lr[NT];
#pragma unroll
            for (int q = 0; q < NT; q++) Lr[q] = Lp[q];
            const CellGeom g = a.geom[T];
            double lx, ly, lz;
            l2p_eval<P, double>(Lr, a.px[it] - g.c[0], a.py[it] - g.c[1], a.pz[it] - g.c[2], lx, ly, lz);
            bx += lx;
            by += ly;
            bz += lz;
        }
        p2p_out(a, slice, ns, it, bx, by, bz);
    }
}

// the slices' parts of the forces of the particles of sliced boxes, summed in slice order
// and stored
__global__ void k_p2p_merge(P2PArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a.sz->overflow || i >= a.n) return;
    const int ns = p2p_box_slices(a, box_of_slot(a.box_off, a.nbox, i));
    if (ns == 1) return;  // stored by the P2P kernel
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int k = 0; k < ns; k++) {
        const double *q = a.part + 3 * ((size_t)k * a.n + i);
        ax += q[0];
        ay += q[1];
        az += q[2];
    }
    store_acc(a, (int)i, ax, ay, az);
}

// ---------------------------------------------------------------------------- host side

static inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

int nterms_of(int order) { return nterms(order); }

#define PH_DISPATCH(order, F)                                      \
    switch (order) {                                               \
    case 3: F(3); break;                                           \
    case 4: F(4); break;                                           \
    case 5: F(5); break;                                           \
    default: throw std::runtime_error("expansion order must be 3, 4 or 5"); \
    }

void compute_offsets_f32(Ctx &c)
{
    float *ox = static_cast<float *>(ctx_buf(c, c.ox, sizeof(float) * c.n, "float offsets"));
    float *oy = static_cast<float *>(ctx_buf(c, c.oy, sizeof(float) * c.n, "float offsets"));
    float *oz = static_cast<float *>(ctx_buf(c, c.oz, sizeof(float) * c.n, "float offsets"));
    k_offsets_f32<<<nblk(nleaf_bound(c), 4), 256, 0, c.stream>>>(dptr<DevSizes>(c.sizes), dptr<CellGeom>(c.geom),
                                                                  dptr<CellInfo>(c.info), dptr<double>(c.px),
                                                                  dptr<double>(c.py), dptr<double>(c.pz), ox, oy, oz);
}

// nodes of level d (launch bound): the count (ordinary evaluation) or the plan's capacity
static int level_bound(const Ctx &c, int d) { return c.plan.active ? c.plan.scap : c.level_off[d + 1] - c.level_off[d]; }
static int levels_run(const Ctx &c) { return c.plan.active ? c.plan.levels : c.nlevel; }

void upward_pass(Ctx &c)
{
    const int NT = nterms(c.prm.order);
    const bool f32 = c.prm.precision == PHOTONS_FP32_MIXED;
    const int ncb = ncell_bound(c);
    // FP64 multipoles always; the FP32-mixed path stores a scaled float copy after them
    size_t mbytes = sizeof(double) * (size_t)ncb * NT;
    double *M = static_cast<double *>(
        ctx_buf(c, c.M, mbytes + (f32 ? sizeof(float) * (size_t)ncb * NT : 0), "multipoles"));
    hipStream_t st = c.stream;
    const DevSizes *sz = dptr<DevSizes>(c.sizes);
    const CellGeom *geom = dptr<CellGeom>(c.geom);
    const CellInfo *info = dptr<CellInfo>(c.info);
    const int *lv = dptr<int>(c.level_nodes);
    const int nlev = levels_run(c);
#define P2M_L(PP)                                                                                                 \
    k_p2m<PP><<<nblk(nleaf_bound(c), 128), 128, 0, st>>>(sz, geom, info, dptr<double>(c.px), dptr<double>(c.py), \
                                                         dptr<double>(c.pz), c.prm.mass, M);                      \
    for (int d = nlev - 1; d >= 0; d--)                                                                           \
        k_m2m_level<PP><<<nblk(level_bound(c, d), 128), 128, 0, st>>>(lv, sz, d, geom, info, M);                  \
    if (f32)                                                                                                      \
        k_scale_down<PP><<<nblk(ncb, 128), 128, 0, st>>>(sz, M, 1.0 / c.prm.rs,                                   \
                                                         reinterpret_cast<float *>(M + (size_t)ncb * NT));
    PH_DISPATCH(c.prm.order, P2M_L)
#undef P2M_L
    PH_CHECK(hipGetLastError());
}

// the FP32 copy of the multipoles follows the FP64 ones, whose rows the buffer holds for
// the launch bound of the cells
static size_t m32_offset(const Ctx &c) { return (size_t)ncell_bound(c) * nterms(c.prm.order); }

static void exclusive_scan_int(Ctx &c, const int *in, int *out, int n)
{
    size_t bytes = 0;
    PH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, c.stream));
    void *tmp = scan_storage(c, bytes);
    PH_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, c.stream));
}

void m2l_pass(Ctx &c, bool accumulate)
{
    const int NT = nterms(c.prm.order);
    const bool f32 = c.prm.precision == PHOTONS_FP32_MIXED;
    const size_t tsize = f32 ? sizeof(float) : sizeof(double);
    const int ncb = ncell_bound(c);
    need(c.L, tsize * (size_t)ncb * NT, "local expansions");
    need(c.M, (sizeof(double) + (f32 ? sizeof(float) : 0)) * (size_t)ncb * NT, "multipoles");
    need(c.m2l_off, sizeof(int64_t) * 2 * (size_t)ncb, "M2L list");
    need(c.m2l_src, sizeof(int32_t), "M2L list");
    void *L = c.L.p;
    hipStream_t st = c.stream;
    need(c.cell_box, sizeof(int) * (size_t)ncb, "cell boxes");
    DevSizes *sz = dptr<DevSizes>(c.sizes);
    const BoxView bx = box_view(c);
    const CellGeom *geom = dptr<CellGeom>(c.geom);
    const int64_t *off = dptr<int64_t>(c.m2l_off);
    const int32_t *src = dptr<int32_t>(c.m2l_src);
    double *M = dptr<double>(c.M);
    // work items: counted per target, scanned, their number checked on the device; an
    // ordinary evaluation reads it back to size the item arrays
    int *item_off = static_cast<int *>(ctx_buf(c, c.m2l_item_off, sizeof(int) * 2 * (size_t)(ncb + 1), "M2L items"));
    int *nitems = item_off + (ncb + 1);
    k_m2l_nitems<<<nblk(ncb + 1, 256), 256, 0, st>>>(off, sz, ncb, nitems);
    exclusive_scan_int(c, nitems, item_off, ncb + 1);
    int item_bound;
    if (c.plan.active) {
        item_bound = c.plan.item_cap;
    } else {
        int *pin = reinterpret_cast<int *>(c.pinned.p) + 64;
        PH_CHECK(hipMemcpyAsync(pin, item_off + ncb, sizeof(int), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipStreamSynchronize(st));
        item_bound = *pin;
        c.meas.items = item_bound;
    }
    int *item_tgt = static_cast<int *>(ctx_buf(c, c.m2l_item_tgt, sizeof(int) * (size_t)std::max(1, item_bound),
                                               "M2L items"));
    void *part = ctx_buf(c, c.m2l_part, tsize * (size_t)std::max(1, item_bound) * NT, "M2L items");
    k_m2l_check<<<1, 1, 0, st>>>(item_off, sz, item_bound);
    k_m2l_item_targets<<<nblk(ncb, 256), 256, 0, st>>>(item_off, sz, item_tgt);
#define M2L_L(PP)                                                                                                  \
    if (f32) {                                                                                                     \
        if (item_bound)                                                                                            \
            k_m2l_items<PP, float><<<nblk(item_bound, 128), 128, 0, st>>>(                                         \
                sz, item_tgt, item_off, geom, off, src, reinterpret_cast<const float *>(M + m32_offset(c)),        \
                1.0 / c.prm.rs, 0.5, bx, reinterpret_cast<float *>(part));                                        \
        k_m2l_reduce<PP, float><<<nblk(ncb, 128), 128, 0, st>>>(sz, item_off, reinterpret_cast<const float *>(part), \
                                                                reinterpret_cast<float *>(L), accumulate);         \
    } else {                                                                                                       \
        if (item_bound)                                                                                            \
            k_m2l_items<PP, double><<<nblk(item_bound, 128), 128, 0, st>>>(sz, item_tgt, item_off, geom, off, src,  \
                                                                           M, 1.0, 0.5 / c.prm.rs, bx,             \
                                                                           reinterpret_cast<double *>(part));      \
        k_m2l_reduce<PP, double><<<nblk(ncb, 128), 128, 0, st>>>(sz, item_off,                                     \
                                                                 reinterpret_cast<const double *>(part),           \
                                                                 reinterpret_cast<double *>(L), accumulate);       \
    }
    PH_DISPATCH(c.prm.order, M2L_L)
#undef M2L_L
    PH_CHECK(hipGetLastError());
}

void downward_pass(Ctx &c)
{
    const bool f32 = c.prm.precision == PHOTONS_FP32_MIXED;
    need(c.L, (f32 ? sizeof(float) : sizeof(double)) * (size_t)ncell_bound(c) * nterms(c.prm.order),
         "local expansions");
    hipStream_t st = c.stream;
    const DevSizes *sz = dptr<DevSizes>(c.sizes);
    const CellGeom *geom = dptr<CellGeom>(c.geom);
    const CellInfo *info = dptr<CellInfo>(c.info);
    const int *lv = dptr<int>(c.level_nodes);
    const int nlev = levels_run(c);
#define L2L_L(PP)                                                                                                  \
    for (int d = 0; d < nlev; d++) {                                                                               \
        const int nb = nblk(level_bound(c, d), 128);                                                              \
        if (f32)                                                                                                   \
            k_l2l_level<PP, float><<<nb, 128, 0, st>>>(lv, sz, d, geom, info, 1.0 / c.prm.rs, dptr<float>(c.L));   \
        else                                                                                                       \
            k_l2l_level<PP, double><<<nb, 128, 0, st>>>(lv, sz, d, geom, info, 1.0, dptr<double>(c.L));            \
    }
    PH_DISPATCH(c.prm.order, L2L_L)
#undef L2L_L
    PH_CHECK(hipGetLastError());
}

void p2p_l2p_launch(Ctx &c, bool with_l2p, bool accumulate, bool caller_order, double *out)
{
    if (c.prm.maxleaf > kP2PMaxLeaf) throw std::runtime_error("maxleaf above 256 is not supported");
    const bool f32p = c.prm.precision == PHOTONS_FP32_MIXED;
    const int nlb = nleaf_bound(c), ncb = ncell_bound(c);
    need(c.p2p_off, sizeof(int64_t) * 2 * (size_t)nlb, "P2P list");
    need(c.p2p_src, sizeof(int32_t), "P2P list");
    need(c.geom, sizeof(CellGeom) * (size_t)ncb, "cells");
    need(c.cell_box, sizeof(int) * (size_t)ncb, "cell boxes");
    need(c.px, sizeof(double) * (size_t)c.n, "positions");
    if (f32p) need(c.ox, sizeof(float) * (size_t)c.n, "float offsets");
    if (with_l2p) need(c.L, (f32p ? sizeof(float) : sizeof(double)) * (size_t)ncb * nterms(c.prm.order),
                       "local expansions");
    P2PArgs a;
    a.geom = dptr<CellGeom>(c.geom);
    a.info = dptr<CellInfo>(c.info);
    a.off = dptr<int64_t>(c.p2p_off);
    a.src = dptr<int32_t>(c.p2p_src);
    a.px = dptr<double>(c.px);
    a.py = dptr<double>(c.py);
    a.pz = dptr<double>(c.pz);
    a.ox = dptr<float>(c.ox);
    a.oy = dptr<float>(c.oy);
    a.oz = dptr<float>(c.oz);
    a.pidx = dptr<int32_t>(c.pidx);
    a.L = c.L.p;
    a.mass = c.prm.mass;
    a.soft = c.prm.soft;
    a.inv2rs = 0.5 / c.prm.rs;
    a.inv_rs = 1.0 / c.prm.rs;
    a.bx = box_view(c);
    a.with_l2p = with_l2p;
    a.nchunk = (c.prm.maxleaf + 63) / 64;
    a.accumulate = accumulate;
    a.out_caller_order = caller_order;
    a.out = out;
    a.sz = dptr<DevSizes>(c.sizes);
    if (nlb > 0) p2p_order(c, nlb);
    a.order = dptr<int32_t>(c.p2p_order);
    const bool f32 = c.prm.precision == PHOTONS_FP32_MIXED;
    hipStream_t st = c.stream;
    // Few particles (few target leaves, a long list each) leave most CUs idle: the list of
    // each target leaf of a small box is cut into slices, one workgroup each, whose parts
    // k_p2p_merge sums.  The slices depend on the box's particle count alone, so that a box
    // gives the same forces bit for bit alone or with other boxes, planned or not.
    a.nbox = c.nbox;
    a.box_off = dptr<int64_t>(c.d_box_off);
    a.n = c.n;
    a.nslice = 1;
    for (int b = 0; b < c.nbox; b++) a.nslice = std::max(a.nslice, p2p_slices(c.box_off[b + 1] - c.box_off[b]));
    a.part = a.nslice > 1 ? static_cast<double *>(ctx_buf(c, c.p2p_part, sizeof(double) * 3 * (size_t)a.nslice * c.n,
                                                           "P2P slices"))
                          : nullptr;
    if ((int64_t)nlb * a.nchunk * a.nslice >= ((int64_t)1 << 31)) throw std::runtime_error("too many P2P workgroups");
    const int nwg = nlb * a.nchunk * a.nslice;
#define P2P_L(PP)                                              \
    if (f32 && c.p2p_stage == 256) k_p2p_f32<PP, 256><<<nwg, 64, 0, st>>>(a); \
    else if (f32) k_p2p_f32<PP, 512><<<nwg, 64, 0, st>>>(a);          \
    else k_p2p_f64<PP><<<nwg, 64, 0, st>>>(a);
    if (nlb > 0) PH_DISPATCH(c.prm.order, P2P_L)
#undef P2P_L
    if (nlb > 0 && a.nslice > 1) k_p2p_merge<<<nblk(c.n, 256), 256, 0, st>>>(a);
    PH_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------- batched operators

template <int P>
__global__ void __launch_bounds__(128) k_op_p2m(int count, const double *pos, const int32_t *first, const int32_t *npart,
                         const double *center, double mass, double *M)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    constexpr int NT = nterms(P);
    double m[NT];
#pragma unroll
    for (int q = 0; q < NT; q++) m[q] = 0.0;
    for (int i = first[k]; i < first[k] + npart[k]; i++)
        p2m_add<P, double>(mass, pos[3 * i] - center[3 * k], pos[3 * i + 1] - center[3 * k + 1],
                           pos[3 * i + 2] - center[3 * k + 2], m);
#pragma unroll
    for (int q = 0; q < NT; q++) M[(size_t)k * NT + q] = m[q];
}

template <int P, int KIND>  // 0 m2m, 1 l2l
__global__ void __launch_bounds__(128) k_op_shift(int count, const double *d, const double *X, double *Y)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    constexpr int NT = nterms(P);
    double x[NT], y[NT];
#pragma unroll
    for (int q = 0; q < NT; q++) {
        x[q] = X[(size_t)k * NT + q];
        y[q] = Y[(size_t)k * NT + q];
    }
    if (KIND == 0) m2m_add<P, double>(d[3 * k], d[3 * k + 1], d[3 * k + 2], x, y);
    else l2l_add<P, double>(d[3 * k], d[3 * k + 1], d[3 * k + 2], x, y);
#pragma unroll
    for (int q = 0; q < NT; q++) Y[(size_t)k * NT + q] = y[q];
}

template <int P>
__global__ void __launch_bounds__(128) k_op_m2l(int count, double inv2rs, const double *R, const double *M, double *L)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    constexpr int NT = nterms(P);
    double m[NT], l[NT], f[P + 1], D[NT];
#pragma unroll
    for (int q = 0; q < NT; q++) {
        m[q] = M[(size_t)k * NT + q];
        l[q] = L[(size_t)k * NT + q];
    }
    const double rx = R[3 * k], ry = R[3 * k + 1], rz = R[3 * k + 2];
    radial_f64<P>(sqrt(rx * rx + ry * ry + rz * rz), inv2rs, f);
    dtensor<P, double>(rx, ry, rz, f, D);
    m2l_contract<P, double>(D, m, l);
#pragma unroll
    for (int q = 0; q < NT; q++) L[(size_t)k * NT + q] = l[q];
}

template <int P>
__global__ void __launch_bounds__(128) k_op_l2p(int n, const double *pos, const int32_t *owner, const double *center, const double *L,
                         double *acc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int NT = nterms(P);
    const int o = owner[i];
    double l[NT];
#pragma unroll
    for (int q = 0; q < NT; q++) l[q] = L[(size_t)o * NT + q];
    double ax, ay, az;
    l2p_eval<P, double>(l, pos[3 * i] - center[3 * o], pos[3 * i + 1] - center[3 * o + 1],
                        pos[3 * i + 2] - center[3 * o + 2], ax, ay, az);
    acc[3 * i] += ax;
    acc[3 * i + 1] += ay;
    acc[3 * i + 2] += az;
}

void op_p2m(hipStream_t st, int order, int count, const double *pos, const int32_t *first, const int32_t *npart,
            const double *center, double mass, double *M)
{
#define L_(PP) k_op_p2m<PP><<<nblk(count, 64), 64, 0, st>>>(count, pos, first, npart, center, mass, M);
    PH_DISPATCH(order, L_)
#undef L_
    PH_CHECK(hipGetLastError());
}

void op_shift(hipStream_t st, int order, int kind, int count, const double *d, const double *X, double *Y)
{
#define L_(PP)                                                                              \
    if (kind == 0) k_op_shift<PP, 0><<<nblk(count, 64), 64, 0, st>>>(count, d, X, Y);       \
    else k_op_shift<PP, 1><<<nblk(count, 64), 64, 0, st>>>(count, d, X, Y);
    PH_DISPATCH(order, L_)
#undef L_
    PH_CHECK(hipGetLastError());
}

void op_m2l(hipStream_t st, int order, int count, double rs, const double *R, const double *M, double *L)
{
#define L_(PP) k_op_m2l<PP><<<nblk(count, 64), 64, 0, st>>>(count, 0.5 / rs, R, M, L);
    PH_DISPATCH(order, L_)
#undef L_
    PH_CHECK(hipGetLastError());
}

void op_l2p(hipStream_t st, int order, int n, const double *pos, const int32_t *owner, const double *center,
            const double *L, double *acc)
{
#define L_(PP) k_op_l2p<PP><<<nblk(n, 64), 64, 0, st>>>(n, pos, owner, center, L, acc);
    PH_DISPATCH(order, L_)
#undef L_
    PH_CHECK(hipGetLastError());
}

}  // namespace photons
