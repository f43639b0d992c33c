// engine.h -- device-resident state of one short-range force context (internal).
//
// HBM layout (all arrays indexed in tree order unless noted):
//   particles  px, py, pz (double SoA), pidx (int32: caller index of each slot),
//              ox, oy, oz (float SoA: offset from the leaf centre, FP32-mixed path)
//   cells      leaves first (depth-first order, = increasing particle order), then
//              nodes in pre-order: CellGeom (centre, width; 48 B) and CellInfo (32 B)
//   expansions M, L: ncell x nterms(p) doubles (FP64 path) or floats in units of
//              rs (FP32-mixed path), one row per cell
//   lists      P2P: CSR over target leaves; M2L: CSR over target cells; an entry is
//              source cell | shift id << 27 (shift id 0 = home box, 1..26 images)
//   boxes      a context holds a forest: nbox independent periodic boxes, each with
//              its own length, whose particles are contiguous (box b owns caller
//              slots box_off[b] .. box_off[b+1]); every cell belongs to one box
//              (cell_box) and pairs never cross boxes
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/photons_fmm.h"

namespace photons {

constexpr int kShiftBits = 27;
constexpr int kCellMask = (1 << kShiftBits) - 1;
// list entries: source cell in the low 27 bits, shift id (0..26) in the top 5 bits
__host__ __device__ inline int32_t pack_entry(int cell, int sid)
{
    return (int32_t)((uint32_t)cell | ((uint32_t)sid << kShiftBits));
}
__host__ __device__ inline int entry_cell(int32_t e) { return (int)((uint32_t)e & (uint32_t)kCellMask); }
__host__ __device__ inline int entry_shift(int32_t e) { return (int)((uint32_t)e >> kShiftBits); }
// the entries of target t in a list: src[off[2 t]] .. src[off[2 t + 1] - 1]
__host__ __device__ inline int64_t span_beg(const int64_t *off, int t) { return off[2 * (size_t)t]; }
__host__ __device__ inline int64_t span_end(const int64_t *off, int t) { return off[2 * (size_t)t + 1]; }
constexpr int kMaxDepth = 240;  // levels of the kd tree

struct CellGeom {
    double c[3];
    double w[3];
};

struct CellInfo {
    int32_t npart, ipart;  // particle range (tree order)
    int32_t son0, son1;    // child cell ids, -1 for leaves
    int32_t dfs, dsize;    // pre-order rank over all cells and subtree cell count
    int32_t parent;        // parent cell id, -1 at the root
    int32_t depth;
};

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// Sizes of the current tree and lists, written on the device by the kernels that
// produce them and read there by every later kernel, so that an evaluation needs no
// host round trip (launch grids are sized by host-side bounds; a kernel's work past
// the device count is empty).
constexpr int kMaxLevels = kMaxDepth + 2;
enum : int {  // overflow bits: the evaluation exceeded the capacities it was planned for
    kOvfNodes = 1, kOvfLeaves = 2, kOvfLevels = 4, kOvfDepths = 8, kOvfPairs = 16, kOvfP2P = 32, kOvfM2L = 64,
    kOvfDown = 128, kOvfItems = 256, kOvfTargets = 512
};
struct DevSizes {
    int nleaf, nnode, ncell, nlevel;  // cells; nlevel = depths holding nodes (0 .. nlevel-1)
    int overflow;                     // kOvf* bits
    int n_items;                      // M2L work items
    int pad[2];
    int64_t n_p2p, n_m2l;             // list lengths
    int64_t nrec;                     // walk records of the current depth (its pairs: twice as many)
    int64_t p2p_base, m2l_base;       // list entries written by the walk's earlier depths
    int64_t dl_total;                 // down records written by the walk's current depth
    int lvl[kMaxLevels + 1];          // first temporary node of depth d
    int ldo[kMaxLevels + 1];          // first temporary leaf of depth d
};

// per-box data the kernels read: box length by cell
struct BoxView {
    const double *len;    // [nbox]
    const int *cell_box;  // [ncell]
    __device__ __forceinline__ double len_of(int cell) const { return len[cell_box[cell]]; }
};

// Capacities of a planned evaluation (photons_plan_*): every buffer is allocated for them
// before the evaluation is queued or captured, the launches are sized by them, and the
// loops over tree levels and walk depths run a fixed number of steps.  What does not fit
// sets an overflow bit in DevSizes instead of being written.
struct Plan {
    bool active = false;
    bool alloc = false;    // the first planned evaluation allocates the capacities it uses
    int levels = 0;        // tree levels split (depths 0 .. levels-1)
    int depths = 0;        // walk depths (0 .. depths-1)
    int scap = 0;          // nodes of one level
    int tcap = 0, lcap = 0;  // temporary nodes, leaves
    int64_t dl_cap = 0;    // walk records (down records) of one depth, the largest
    std::vector<int64_t> rec_depth;  // walk records of each depth (launch bounds)
    int64_t p2p_cap = 0, m2l_cap = 0;  // list entries
    int item_cap = 0;      // M2L work items
};

// A plan kept for the next evaluations of the same boxes: an ordinary evaluation measures
// and plans; the next ones run planned (no host round trip inside) and check the overflow
// bits once at the end, measuring again if the particles outgrew the plan.
struct Reuse {
    bool on = true;
    bool have = false;       // a plan for the box set below
    bool allocated = false;  // its buffers are allocated
    std::vector<int64_t> box_off;
    std::vector<double> box_len;
    Plan plan;
};

// What an ordinary evaluation measured, for planning: the largest count of each kind
struct Measure {
    int levels = 0, scap = 0, ntmp = 0, nleaf = 0, depths = 0, items = 0;
    int64_t dl = 0, p2p = 0, m2l = 0;
    std::vector<int64_t> dl_depth;  // walk records of each depth
};

struct Timing {
    bool enabled = false;
    hipEvent_t ev[16];
    int nev = 0;
};

struct Ctx {
    photons_params prm;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;

    int64_t n = 0;
    int nleaf = 0, nnode = 0, ncell = 0, nlevel = 0, maxdepth = 0;
    int nbox = 1;
    int p2p_stage = 256;           // FP32 P2P staging slots (256 or 512; tuned)
    std::vector<int64_t> box_off;  // particle offsets of the boxes (nbox + 1)
    std::vector<double> box_len;   // box lengths
    DevBuf d_box_off, d_box_len;   // device copies
    DevBuf d_box_root;             // root cell id of each box (int)
    DevBuf cell_box;               // box of each cell (int)
    std::vector<int> level_off;  // nodes of depth d: level_nodes[level_off[d] .. level_off[d+1])

    // persistent device buffers, grown on demand
    DevBuf px, py, pz, pidx;       // tree-ordered particles
    DevBuf qx, qy, qz, qidx;       // build double buffer
    DevBuf ox, oy, oz;             // FP32 offsets
    DevBuf geom, info;             // CellGeom / CellInfo [ncell]
    DevBuf level_nodes;            // node cell ids by depth
    DevBuf M, L;                   // expansions
    // lists: entries grouped by target; p2p_off / m2l_off hold one span (begin, end) of
    // entries per target leaf / cell: off[2 t] .. off[2 t + 1]
    DevBuf p2p_off, p2p_src, m2l_off, m2l_src, cnt_tmp;
    // walk: the cells by depth-first rank as sources and by cell as targets, the down
    // records of a depth (ping-pong), each pair's outputs, the continued traversals
    DevBuf walk_rec, walk_trec, dl_a, dl_b, pair_cnt, big_pairs, walk_tlist, walk_tcount, walk_qoff, walk_ccount,
        walk_cidx;
    DevBuf p2p_order, p2p_order_tmp;          // P2P launch order (heaviest target leaves first)
    DevBuf p2p_part;                          // P2P forces per slice of the lists (few particles)
    DevBuf m2l_item_off, m2l_item_tgt, m2l_part;  // balanced M2L work items
    DevBuf scan_tmp;
    // build temporaries (level order)
    int tcap = 0, lcap = 0;        // capacities of the level-order temporaries
    DevBuf t_start, t_count, t_split, t_lo, t_hi, t_son, t_depth, t_parent;
    DevBuf l_start, l_count, l_lo, l_hi, l_parent;
    DevBuf seg_chunk, chunk_part, chunk_rcnt, seg_nleft, seg_flag, seg_nn, seg_non;
    DevBuf t_size, t_lsize, t_dfs, t_lfirst, l_final;
    DevBuf err;                    // device error flag (int)
    DevBuf sizes;                  // DevSizes of the current evaluation
    DevBuf stats;                  // device counters (int64)
    DevBuf pinned;                 // small host-pinned staging
    hipEvent_t counts_ready = nullptr;  // a count copied to `pinned` while later work runs

    Plan plan;                     // active: the evaluation runs without host round trips
    Reuse reuse;                   // the plan kept for the next evaluations of these boxes
    Measure meas;                  // filled by ordinary evaluations
    hipGraph_t graph = nullptr;    // a captured planned evaluation (photons_graph_*)
    hipGraphExec_t graph_exec = nullptr;
    bool graph_ready = false;      // planned buffers allocated (captured or not)
    bool graph_captured = false;
    const double *graph_pos = nullptr;
    double *graph_acc = nullptr;

    int64_t n_p2p = 0, n_m2l = 0;  // list lengths
    bool lists_valid = false;      // CSR lists built by a walk on the current tree
    int64_t n_p2p_local = 0, n_m2l_local = 0;

    Timing timing;
    double phase_ms[16] = {0};
    int phase_count = 0;

    std::string err_msg;
};

// ---- helpers shared by the .hip files
#define PH_CHECK(x)                                                                 \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
        }                                                                           \
    } while (0)

void *grow(DevBuf &b, size_t bytes);
// Device memory of the evaluations: hipMalloc, or, inside photons_force_device_static, a
// piece of the library's static scratch array (dev_free of such a piece does nothing).
void *dev_alloc(size_t bytes);
void dev_free(void *p);
bool arena_active();  // the static scratch serves dev_alloc on this thread
// a buffer of the current evaluation: grown (ordinary) or checked against the plan
// (a planned evaluation allocates nothing)
void *ctx_buf(Ctx &c, DevBuf &b, size_t bytes, const char *what);
inline void *scan_storage(Ctx &c, size_t bytes) { return ctx_buf(c, c.scan_tmp, bytes, "scan storage"); }
// launch bounds of the cells: the counts of the tree (ordinary) or the plan's capacities
inline int nleaf_bound(const Ctx &c) { return c.plan.active ? c.plan.lcap : c.nleaf; }
inline int nnode_bound(const Ctx &c) { return c.plan.active ? c.plan.tcap : c.nnode; }
inline int ncell_bound(const Ctx &c) { return nleaf_bound(c) + nnode_bound(c); }
// throws unless b holds at least `bytes` (guards every launch that reads b)
void need(const DevBuf &b, size_t bytes, const char *what);
void upload_positions(Ctx &c, const double *d_pos_aos);  // AoS [n][3] -> px, py, pz, pidx
template <typename T>
T *dptr(DevBuf &b)
{
    return reinterpret_cast<T *>(b.p);
}
// box owning particle slot i (tree order keeps each box's particles in its input range):
// the last b with box_off[b] <= i
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

inline BoxView box_view(Ctx &c) { return BoxView{dptr<double>(c.d_box_len), dptr<int>(c.cell_box)}; }
// sets the boxes of the next evaluation (counts and lengths; uploads them)
void set_boxes(Ctx &c, int nbox, const int64_t *counts, const double *lengths);

// phases (tree.hip, walk.hip, kernels.hip)
void build_tree(Ctx &c);                 // px.. sorted, cells and levels built
void compute_offsets_f32(Ctx &c);        // ox, oy, oz
void upward_pass(Ctx &c);                // P2M + M2M (+ FP32 conversion)
void run_walk(Ctx &c, int shift_lo, int shift_hi);  // lists for shifts [lo, hi)
void m2l_pass(Ctx &c, bool accumulate);  // L (+)= M2L over the M2L list
void downward_pass(Ctx &c);              // L2L
void p2p_l2p_pass(Ctx &c, bool with_l2p, bool accumulate);  // acc (+)= P2P (+ L2P)

int nterms_of(int order);
void phase_mark(Ctx &c);                 // records a timing event (when enabled)

}  // namespace photons
