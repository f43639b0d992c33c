// walk.hip -- device dual-tree walk emitting flat interaction lists.
//
// The reference walks (target, source) cell pairs depth-first from (root, root)
// and emits P2P and M2L tasks (src/fmm.c:406-712 for the home box,
// src/remotes.c:213-552 against each of the 26 shifted, pruned copies of the tree
// built by prepare_sendtree2, src/remotes.c:60-169).  Which pairs it emits depends
// only on the rules applied to each pair, so the same lists come out of any
// traversal that applies them.
//
// Here one thread owns one target cell X and replays the walk restricted to the
// pairs whose target is an ancestor-or-self of X: opening a target keeps only the
// child on the way to X, and a pair is emitted only when its target is X.  Each
// thread therefore writes its own list, in the reference's emission order for
// that target, with no atomics and no sort.  A count pass sizes the lists, an
// exclusive scan places them, a fill pass writes them.
//
// Image walks: a shifted node keeps no children when its box is rc or more from
// the domain box or when 0.95*theta*distance exceeds its largest width; that rule
// is evaluated on the fly.  An image whose shifted domain box is rc or more from
// the box of X's parent cannot reach X and is skipped (every pair reaching X
// passes through a pair (parent, source) that opened the target, which needs a
// gap below rc).
#include <hipcub/hipcub.hpp>

#include <stdexcept>

#include "engine.h"
#include "fmm_math.h"

namespace photons {

struct WalkArgs {
    const CellGeom *geom;
    const CellInfo *info;
    int nleaf, ncell;
    double theta, theta2, rc, rc2, boxsize;
    int shift_lo, shift_hi;
    int start_target, start_source;
    int64_t *p2p_cnt, *m2l_cnt;          // pass 0
    const int64_t *p2p_off, *m2l_off;    // pass 1
    int32_t *p2p_src, *m2l_src;          // pass 1
    int *err;
};

__host__ __device__ inline void shift_of(int sid, double L, double &sx, double &sy, double &sz)
{
    // ids 1..26: (mi, mj, mk) over {-1,0,1}^3 without (0,0,0), mi slowest
    if (sid <= 0) {
        sx = sy = sz = 0.0;
        return;
    }
    int k = sid <= 13 ? sid - 1 : sid;  // skip the centre (index 13 of the 27)
    int mi = k / 9 - 1, mj = (k / 3) % 3 - 1, mk = k % 3 - 1;
    sx = (double)mi * L;
    sy = (double)mj * L;
    sz = (double)mk * L;
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

__device__ __forceinline__ double box_gap(const CellGeom &a, const CellGeom &b, double sx, double sy, double sz)
{
    double s[3] = {sx, sy, sz}, g2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double d = a.c[i] - (b.c[i] + s[i]);
        if (d < 0.0) d = -d;
        d -= (a.w[i] + b.w[i]) * 0.5;
        if (d > 0.0) g2 += d * d;
    }
    return sqrt(g2);
}

template <int PASS>
__global__ void __launch_bounds__(128) k_walk(WalkArgs a)
{
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    if (X >= a.ncell) return;
    const int nleaf = a.nleaf;
    const CellInfo IX = a.info[X];
    const CellInfo IT = a.info[a.start_target];
    const bool inside = IX.dfs >= IT.dfs && IX.dfs < IT.dfs + IT.dsize;
    int64_t np = 0, nm = 0;
    int64_t po = 0, mo = 0;
    if (PASS == 1) {
        if (X < nleaf) po = a.p2p_off[X];
        mo = a.m2l_off[X];
    }
    const int dX = IX.dfs;
    int2 stack[kWalkStack];
    const int root = a.start_target;
    const CellGeom groot = a.geom[nleaf];  // domain box (tree root)

    for (int sid = inside ? a.shift_lo : a.shift_hi; sid < a.shift_hi; sid++) {
        double sx, sy, sz;
        shift_of(sid, a.boxsize, sx, sy, sz);
        if (sid > 0 && X != root && IX.parent >= 0) {
            const CellGeom gp = a.geom[IX.parent];
            if (box_gap(gp, groot, sx, sy, sz) >= a.rc * (1.0 + 1e-9) + 1e-9 * a.boxsize) continue;
        }
        int sp = 0;
        stack[sp++] = make_int2(root, a.start_source);
        while (sp > 0) {
            const int2 e = stack[--sp];
            const int A = e.x, B = e.y;
            const bool aleaf = A < nleaf, bleaf = B < nleaf;
            if (sp + 2 > kWalkStack) {
                atomicOr(a.err, 1);
                break;
            }
            if (sid == 0 && A == B) {  // self pair
                if (aleaf) {
                    if (PASS == 1) a.p2p_src[po + np] = B;
                    np++;
                } else if (A != X) {
                    const CellInfo IA = a.info[A];
                    const CellInfo I0 = a.info[IA.son0];
                    int toward = (dX >= I0.dfs && dX < I0.dfs + I0.dsize) ? IA.son0 : IA.son1;
                    stack[sp++] = make_int2(toward, IA.son1);
                    stack[sp++] = make_int2(toward, IA.son0);
                }
                continue;
            }
            if (aleaf && bleaf) {  // leaf-leaf pairs are always P2P
                if (PASS == 1) a.p2p_src[po + np] = pack_entry(B, sid);
                np++;
                continue;
            }
            const CellGeom gA = a.geom[A], gB = a.geom[B];
            const double d0 = gA.c[0] - (gB.c[0] + sx);
            const double d1 = gA.c[1] - (gB.c[1] + sy);
            const double d2 = gA.c[2] - (gB.c[2] + sz);
            const int flag = mac(gA.w, gB.w, d0, d1, d2, a.theta2, a.rc2);
            if (flag < 0) continue;
            bool open_target;
            if (sid == 0) {
                if (flag == 1) {
                    if (A == X) {
                        if (PASS == 1) a.m2l_src[mo + nm] = B;
                        nm++;
                    }
                    continue;
                }
                open_target = aleaf ? false : (bleaf ? true : (gA.w[0] + gA.w[1] + gA.w[2]) > (gB.w[0] + gB.w[1] + gB.w[2]));
            } else {
                const bool pruned = !bleaf && let_pruned(groot, gB, sx, sy, sz, a.rc, a.theta);
                if (aleaf) {  // target leaf against a shifted node
                    if (flag == 1 || pruned) {
                        if (PASS == 1) a.m2l_src[mo + nm] = pack_entry(B, sid);
                        nm++;
                    } else {
                        const CellInfo IB = a.info[B];
                        stack[sp++] = make_int2(A, IB.son1);
                        stack[sp++] = make_int2(A, IB.son0);
                    }
                    continue;
                }
                if (flag == 1) {
                    if (A == X) {
                        if (PASS == 1) a.m2l_src[mo + nm] = pack_entry(B, sid);
                        nm++;
                    }
                    continue;
                }
                open_target = bleaf ? true
                                    : ((gA.w[0] + gA.w[1] + gA.w[2]) > (gB.w[0] + gB.w[1] + gB.w[2]) || pruned);
            }
            if (open_target) {
                if (A != X) {
                    const CellInfo IA = a.info[A];
                    const CellInfo I0 = a.info[IA.son0];
                    int toward = (dX >= I0.dfs && dX < I0.dfs + I0.dsize) ? IA.son0 : IA.son1;
                    stack[sp++] = make_int2(toward, B);
                }
            } else {
                const CellInfo IB = a.info[B];
                stack[sp++] = make_int2(A, IB.son1);
                stack[sp++] = make_int2(A, IB.son0);
            }
        }
    }
    if (PASS == 0) {
        if (X < nleaf) a.p2p_cnt[X] = np;
        a.m2l_cnt[X] = nm;
    }
}

__global__ void k_zero_tail(int64_t *p, int idx) { p[idx] = 0; }

static void scan64(Ctx &c, const int64_t *in, int64_t *out, int n)
{
    size_t bytes = 0;
    PH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, c.stream));
    void *tmp = grow(c.scan_tmp, bytes);
    PH_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, c.stream));
}

void run_walk_from(Ctx &c, int shift_lo, int shift_hi, int start_target, int start_source)
{
    hipStream_t st = c.stream;
    const int nleaf = c.nleaf, ncell = c.ncell;
    if (2 * c.maxdepth + 2 > kWalkStack) throw std::runtime_error("tree too deep for the walk stack");
    int64_t *cnt = reinterpret_cast<int64_t *>(grow(c.cnt_tmp, sizeof(int64_t) * (size_t)(nleaf + 1 + ncell + 1)));
    int64_t *p2p_cnt = cnt, *m2l_cnt = cnt + nleaf + 1;
    int64_t *p2p_off = reinterpret_cast<int64_t *>(grow(c.p2p_off, sizeof(int64_t) * (size_t)(nleaf + 1)));
    int64_t *m2l_off = reinterpret_cast<int64_t *>(grow(c.m2l_off, sizeof(int64_t) * (size_t)(ncell + 1)));
    int *err = reinterpret_cast<int *>(c.err.p);
    PH_CHECK(hipMemsetAsync(err, 0, sizeof(int), st));

    WalkArgs a{};
    a.geom = dptr<CellGeom>(c.geom);
    a.info = dptr<CellInfo>(c.info);
    a.nleaf = nleaf;
    a.ncell = ncell;
    a.theta = c.prm.theta;
    a.theta2 = c.prm.theta * c.prm.theta;
    a.rc = c.prm.rc;
    a.rc2 = c.prm.rc * c.prm.rc;
    a.boxsize = c.prm.boxsize;
    a.shift_lo = shift_lo;
    a.shift_hi = shift_hi;
    a.start_target = start_target;
    a.start_source = start_source;
    a.p2p_cnt = p2p_cnt;
    a.m2l_cnt = m2l_cnt;
    a.err = err;
    const int bs = 128, nb = (ncell + bs - 1) / bs;
    k_walk<0><<<nb, bs, 0, st>>>(a);
    k_zero_tail<<<1, 1, 0, st>>>(p2p_cnt, nleaf);
    k_zero_tail<<<1, 1, 0, st>>>(m2l_cnt, ncell);
    scan64(c, p2p_cnt, p2p_off, nleaf + 1);
    scan64(c, m2l_cnt, m2l_off, ncell + 1);
    int64_t *pin = reinterpret_cast<int64_t *>(c.pinned.p) + 8;
    PH_CHECK(hipMemcpyAsync(pin, p2p_off + nleaf, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipMemcpyAsync(pin + 1, m2l_off + ncell, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipMemcpyAsync(pin + 2, err, sizeof(int), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
    if (*reinterpret_cast<int *>(pin + 2)) throw std::runtime_error("walk stack overflow");
    c.n_p2p = pin[0];
    c.n_m2l = pin[1];
    a.p2p_off = p2p_off;
    a.m2l_off = m2l_off;
    a.p2p_src = reinterpret_cast<int32_t *>(grow(c.p2p_src, sizeof(int32_t) * (size_t)std::max<int64_t>(1, c.n_p2p)));
    a.m2l_src = reinterpret_cast<int32_t *>(grow(c.m2l_src, sizeof(int32_t) * (size_t)std::max<int64_t>(1, c.n_m2l)));
    k_walk<1><<<nb, bs, 0, st>>>(a);
    PH_CHECK(hipGetLastError());
}

void run_walk(Ctx &c, int shift_lo, int shift_hi) { run_walk_from(c, shift_lo, shift_hi, c.nleaf, c.nleaf); }

}  // namespace photons
