// serve.hip -- a request loop over resident boxes: continuous batching of
// single-step (force + kick/drift) requests on one GPU.
//
// Each box keeps pos, vel and the last force (acc) in HBM, caller order, FP64.  A
// step is the reference's loop body (src/photoNs.c:136-268): half kicks with the
// previous forces, drift, wrap into [0, L), new forces, half kicks; a box added with
// a mesh (photons_server_add_box_pm) also keeps its PM force (acc_pm) and takes the
// long-range half kicks too, its PM force computed on the evaluating lane's stream
// (one PM context per lane and mesh shape, made when first needed); a box without
// one is stepped with the short-range force alone.  Requests queue in arrival order.  Each of nlanes worker lanes (its own
// force context and HIP stream) takes, whenever it is free, every queued request
// whose box is idle, up to its particle capacity, and evaluates them as one forest
// (photons_force_boxes_device), so requests that arrive while others run are taken
// as soon as a lane frees up, and the lanes' work overlaps on the device.  Boxes
// are admitted while the resident bytes stay within the HBM budget minus what the
// lanes reserve for their evaluation buffers.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <algorithm>
#include <utility>

#include "../../include/photons_fmm.h"
#include "kick.h"

namespace {

#define SV_CHECK(x)                                                                   \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// evaluation buffers per particle of a batch: tree, expansions, lists (~37 entries
// per particle at the 256^3 density, 8 bytes each plus sort buffers) and the
// batch's own pos/acc copies, rounded up
constexpr double kTransientBytesPerParticle = 1024.0;
constexpr double kResidentBytesPerParticle = 72.0;  // pos, vel, acc (3 doubles each)
constexpr int kLaneMeshes = 8;                      // PM contexts kept per lane

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Box {
    int64_t n = 0;
    double L = 0.0;
    double *pos = nullptr, *vel = nullptr, *acc = nullptr;
    int nside = 0;              // PM mesh cells per side, 0: no long-range part
    double *acc_pm = nullptr;   // the PM force (nside > 0)
    bool busy = false;      // a request of this box is being evaluated
};

enum Kind { kStep = 0, kForce = 1 };

struct Request {
    int box = -1;
    int kind = kStep;
    double dkh = 0.0, dd = 0.0;
    double t_submit = 0.0, t_done = 0.0;
    bool done = false;
    std::string error;
};

struct Lane {
    photons_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    double *in = nullptr, *out = nullptr;  // batch pos / acc, [cap][3]
    int64_t cap = 0;
    // PM contexts by mesh side, on this lane's stream (the least recently used goes
    // beyond kLaneMeshes); a context takes each box's length in turn
    std::map<int, std::pair<photons_pm *, int64_t>> pm;
    int64_t pm_clock = 0;
    std::thread th;
};

}  // namespace

struct photons_server {
    photons_params prm;
    int device = 0;
    double budget = 0.0, reserve = 0.0, resident = 0.0, peak = 0.0;
    int64_t lane_cap = 0;  // particles per batch
    int max_boxes_per_batch = 256;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;  // request ids, arrival order
    std::map<int64_t, Request> reqs;
    std::map<int, Box> boxes;
    int next_box = 0;
    int64_t next_req = 0;
    bool stop = false;
    bool paused = false;  // lanes take no new batch (a burst of requests queued at once)
    std::vector<Lane> lanes;
    // statistics
    int64_t n_done = 0, particles_stepped = 0, batches = 0, batch_boxes = 0;
    int lanes_busy = 0, max_lanes_busy = 0;  // lanes evaluating a batch, now and at most at once
    double t_first = -1.0, t_last = 0.0;
    std::string fatal;
};

namespace {

thread_local std::string g_err;

int fail(const std::string &m)
{
    g_err = m;
    return -1;
}

// the lane's PM context for meshes of nside^3 cells, made when first needed; the least
// recently used one is released past kLaneMeshes (its stream, the lane's, holds no
// work of it: the lane's batches run one at a time and end with a wait)
photons_pm *lane_mesh(photons_server *s, Lane &ln, int nside)
{
    auto it = ln.pm.find(nside);
    if (it == ln.pm.end()) {
        if ((int)ln.pm.size() >= kLaneMeshes) {
            auto lru = ln.pm.begin();
            for (auto jt = ln.pm.begin(); jt != ln.pm.end(); ++jt)
                if (jt->second.second < lru->second.second) lru = jt;
            photons_pm_destroy(lru->second.first);
            ln.pm.erase(lru);
        }
        photons_pm *pm = nullptr;
        if (photons_pm_create(&pm, nside, 1.0, s->prm.rs, s->prm.mass, s->device, ln.stream) != 0)
            throw std::runtime_error(std::string("PM mesh: ") + photons_pm_last_error());
        it = ln.pm.emplace(nside, std::make_pair(pm, (int64_t)0)).first;
    }
    it->second.second = ++ln.pm_clock;
    return it->second.first;
}

// one batch on one lane: requests `ids` (their boxes marked busy by the caller)
void run_batch(photons_server *s, Lane &ln, const std::vector<int64_t> &ids)
{
    std::vector<int64_t> counts;
    std::vector<double> lengths;
    std::vector<Request> rq;
    std::vector<Box> bx;
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (int64_t id : ids) {
            rq.push_back(s->reqs[id]);
            bx.push_back(s->boxes[rq.back().box]);
        }
    }
    int64_t off = 0;
    for (size_t k = 0; k < ids.size(); k++) {
        const int64_t n3 = 3 * bx[k].n;
        const int step = rq[k].kind == kStep;
        k_kick_drift<<<kick_blocks(n3), 256, 0, ln.stream>>>(n3, bx[k].pos, bx[k].vel, bx[k].acc, bx[k].acc_pm,
                                                            rq[k].dkh, rq[k].dd, bx[k].L, step, ln.in + 3 * off);
        counts.push_back(bx[k].n);
        lengths.push_back(bx[k].L);
        off += bx[k].n;
    }
    SV_CHECK(hipGetLastError());
    if (photons_force_boxes_device(ln.ctx, (int)counts.size(), counts.data(), lengths.data(), ln.in, ln.out) != 0)
        throw std::runtime_error(std::string("force: ") + photons_last_error());
    // the PM force of the boxes with a mesh, at the drifted positions, on this lane's stream
    off = 0;
    for (size_t k = 0; k < ids.size(); k++) {
        if (bx[k].nside > 0) {
            photons_pm *pm = lane_mesh(s, ln, bx[k].nside);
            if (photons_pm_set_boxsize(pm, bx[k].L) != 0 ||
                photons_pm_force_device(pm, bx[k].n, ln.in + 3 * off, bx[k].acc_pm) != 0)
                throw std::runtime_error(std::string("PM force: ") + photons_pm_last_error());
        }
        off += bx[k].n;
    }
    off = 0;
    for (size_t k = 0; k < ids.size(); k++) {
        const int64_t n3 = 3 * bx[k].n;
        k_kick<<<kick_blocks(n3), 256, 0, ln.stream>>>(n3, bx[k].vel, bx[k].acc, ln.out + 3 * off, bx[k].acc_pm,
                                                      rq[k].dkh, rq[k].kind == kStep);
        off += bx[k].n;
    }
    SV_CHECK(hipGetLastError());
    SV_CHECK(hipStreamSynchronize(ln.stream));
}

void lane_loop(photons_server *s, int lane_id)
{
    Lane &ln = s->lanes[lane_id];
    (void)hipSetDevice(s->device);
    for (;;) {
        std::vector<int64_t> ids;
        {
            std::unique_lock<std::mutex> g(s->mu);
            for (;;) {
                if (s->stop) return;
                if (s->paused) {
                    s->cv_work.wait(g);
                    continue;
                }
                // every queued request whose box is idle, in arrival order, within capacity
                int64_t total = 0;
                std::vector<int> taken;
                for (auto it = s->queue.begin(); it != s->queue.end();) {
                    Request &r = s->reqs[*it];
                    Box &b = s->boxes[r.box];
                    bool box_taken = false;
                    for (int t : taken) box_taken |= (t == r.box);
                    if (b.busy || box_taken || (int)ids.size() >= s->max_boxes_per_batch ||
                        (total > 0 && total + b.n > ln.cap)) {
                        ++it;
                        continue;
                    }
                    ids.push_back(*it);
                    taken.push_back(r.box);
                    total += b.n;
                    it = s->queue.erase(it);
                }
                if (!ids.empty()) {
                    for (int t : taken) s->boxes[t].busy = true;
                    s->lanes_busy++;
                    s->max_lanes_busy = std::max(s->max_lanes_busy, s->lanes_busy);
                    break;
                }
                s->cv_work.wait(g);
            }
        }
        std::string err;
        try {
            run_batch(s, ln, ids);
        } catch (const std::exception &e) {
            err = e.what();
        }
        const double t = now_s();
        {
            std::lock_guard<std::mutex> g(s->mu);
            for (int64_t id : ids) {
                Request &r = s->reqs[id];
                Box &b = s->boxes[r.box];
                b.busy = false;
                r.done = true;
                r.t_done = t;
                r.error = err;
                if (err.empty()) {
                    s->n_done++;
                    if (r.kind == kStep) s->particles_stepped += b.n;
                }
                // a box's initial force request has no id a caller could wait on
                if (r.kind == kForce) s->reqs.erase(id);
            }
            s->lanes_busy--;
            s->batches++;
            s->batch_boxes += (int64_t)ids.size();
            s->t_last = t;
            if (!err.empty() && s->fatal.empty()) s->fatal = err;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();  // boxes freed: other lanes may take their next requests
    }
}

int64_t enqueue(photons_server *s, int box, int kind, double dkh, double dd)
{
    Request r;
    r.box = box;
    r.kind = kind;
    r.dkh = dkh;
    r.dd = dd;
    r.t_submit = now_s();
    const int64_t id = s->next_req++;
    s->reqs[id] = r;
    s->queue.push_back(id);
    if (s->t_first < 0.0) s->t_first = r.t_submit;
    return id;
}

}  // namespace

extern "C" {

const char *photons_server_last_error(void) { return g_err.c_str(); }

int photons_server_create(photons_server **out, const photons_params *prm, int device, int nlanes,
                          double hbm_budget_bytes, int64_t max_batch_particles)
{
    try {
        if (!out || !prm) return fail("null argument");
        if (nlanes < 1 || nlanes > 8) return fail("nlanes must be in [1, 8]");
        if (!(hbm_budget_bytes > 0.0)) return fail("the HBM budget must be positive");
        if (max_batch_particles < 1) return fail("max_batch_particles must be positive");
        SV_CHECK(hipSetDevice(device));
        size_t free_b = 0, total_b = 0;
        SV_CHECK(hipMemGetInfo(&free_b, &total_b));
        if (hbm_budget_bytes > (double)free_b) return fail("the HBM budget exceeds the free device memory");
        photons_server *s = new photons_server();
        s->prm = *prm;
        s->device = device;
        s->budget = hbm_budget_bytes;
        s->lane_cap = max_batch_particles;
        s->reserve = nlanes * (double)max_batch_particles * kTransientBytesPerParticle;
        if (s->reserve >= s->budget) {
            delete s;
            return fail("the lanes' evaluation buffers alone exceed the HBM budget");
        }
        s->lanes.resize(nlanes);
        try {
            for (int l = 0; l < nlanes; l++) {
                Lane &ln = s->lanes[l];
                SV_CHECK(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking));
                if (photons_create(&ln.ctx, prm, device, ln.stream) != 0)
                    throw std::runtime_error(std::string("lane context: ") + photons_last_error());
                ln.cap = max_batch_particles;
                SV_CHECK(hipMalloc(&ln.in, sizeof(double) * 3 * (size_t)ln.cap));
                SV_CHECK(hipMalloc(&ln.out, sizeof(double) * 3 * (size_t)ln.cap));
            }
        } catch (...) {
            // no lane thread runs yet: release the lanes built so far
            for (Lane &ln : s->lanes) {
                if (ln.ctx) photons_destroy(ln.ctx);
                if (ln.in) (void)hipFree(ln.in);
                if (ln.out) (void)hipFree(ln.out);
                if (ln.stream) (void)hipStreamDestroy(ln.stream);
            }
            delete s;
            throw;
        }
        for (int l = 0; l < nlanes; l++) s->lanes[l].th = std::thread(lane_loop, s, l);
        *out = s;
        return 0;
    } catch (const std::exception &e) {
        return fail(std::string("photons_server_create: ") + e.what());
    }
}

// Waits for the queued and running requests, then stops the lanes.  No other call
// on the server may run concurrently with or after this one.
int photons_server_destroy(photons_server *s)
{
    if (!s) return 0;
    {
        std::unique_lock<std::mutex> g(s->mu);
        s->paused = false;  // the queued requests run first
        s->cv_work.notify_all();
        s->cv_done.wait(g, [&] {
            if (!s->fatal.empty()) return true;  // the lanes take no more work
            if (!s->queue.empty()) return false;
            for (auto &kv : s->boxes)
                if (kv.second.busy) return false;
            return true;
        });
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (Lane &ln : s->lanes)
        if (ln.th.joinable()) ln.th.join();
    (void)hipSetDevice(s->device);
    for (Lane &ln : s->lanes) {
        if (ln.stream) (void)hipStreamSynchronize(ln.stream);
        for (auto &kv : ln.pm) photons_pm_destroy(kv.second.first);
        if (ln.ctx) photons_destroy(ln.ctx);
        if (ln.in) (void)hipFree(ln.in);
        if (ln.out) (void)hipFree(ln.out);
        if (ln.stream) (void)hipStreamDestroy(ln.stream);
    }
    for (auto &kv : s->boxes) {
        (void)hipFree(kv.second.pos);
        (void)hipFree(kv.second.vel);
        (void)hipFree(kv.second.acc);
        if (kv.second.acc_pm) (void)hipFree(kv.second.acc_pm);
    }
    delete s;
    return 0;
}

}  // extern "C"

namespace {

int add_box(photons_server *s, int64_t n, double boxsize, int nside, const double *pos, const double *vel,
            const double *acc, const double *acc_pm, int *box_id)
{
    if (!s || !pos || !vel || !box_id) return fail("null argument");
    if (!(boxsize > 0.0)) return fail("boxsize must be positive");
    if (n <= (int64_t)s->prm.maxleaf) return fail("a periodic box needs more than maxleaf particles");
    if (n > s->lane_cap) return fail("the box is larger than a lane's batch capacity");
    if (nside > 0 && (acc != nullptr) != (acc_pm != nullptr))
        return fail("a box with a mesh takes both forces or neither");
    const double bytes = (kResidentBytesPerParticle + (nside > 0 ? 24.0 : 0.0)) * (double)n;
    std::unique_lock<std::mutex> g(s->mu);
    if (s->resident + bytes > s->budget - s->reserve) return fail("HBM budget exhausted: box not admitted");
    s->resident += bytes;
    s->peak = std::max(s->peak, s->resident + s->reserve);
    g.unlock();
    Box b;
    b.n = n;
    b.L = boxsize;
    b.nside = nside;
    try {
        SV_CHECK(hipSetDevice(s->device));
        const size_t nb = sizeof(double) * 3 * (size_t)n;
        SV_CHECK(hipMalloc(&b.pos, nb));
        SV_CHECK(hipMalloc(&b.vel, nb));
        SV_CHECK(hipMalloc(&b.acc, nb));
        if (nside > 0) SV_CHECK(hipMalloc(&b.acc_pm, nb));
        SV_CHECK(hipMemcpy(b.pos, pos, nb, hipMemcpyHostToDevice));
        SV_CHECK(hipMemcpy(b.vel, vel, nb, hipMemcpyHostToDevice));
        if (acc) SV_CHECK(hipMemcpy(b.acc, acc, nb, hipMemcpyHostToDevice));
        if (acc_pm && nside > 0) SV_CHECK(hipMemcpy(b.acc_pm, acc_pm, nb, hipMemcpyHostToDevice));
    } catch (...) {
        // give back what this box took: its allocations and its share of the budget
        for (double *p : {b.pos, b.vel, b.acc, b.acc_pm})
            if (p) (void)hipFree(p);
        std::lock_guard<std::mutex> r(s->mu);
        s->resident -= bytes;
        throw;
    }
    g.lock();
    const int id = s->next_box++;
    s->boxes[id] = b;
    // without a given force the box's first request computes it (the forces before the
    // reference's first step, src/photoNs.c:97-127)
    if (!acc) enqueue(s, id, kForce, 0.0, 0.0);
    *box_id = id;
    g.unlock();
    s->cv_work.notify_all();
    return 0;
}

double resident_bytes(const Box &b) { return (kResidentBytesPerParticle + (b.nside > 0 ? 24.0 : 0.0)) * (double)b.n; }

}  // namespace

extern "C" {

int photons_server_add_box(photons_server *s, int64_t n, double boxsize, const double *pos, const double *vel,
                           const double *acc, int *box_id)
{
    try {
        return add_box(s, n, boxsize, 0, pos, vel, acc, nullptr, box_id);
    } catch (const std::exception &e) {
        return fail(std::string("photons_server_add_box: ") + e.what());
    }
}

int photons_server_add_box_pm(photons_server *s, int64_t n, double boxsize, int nside, const double *pos,
                              const double *vel, const double *acc, const double *acc_pm, int *box_id)
{
    try {
        // the mesh shape photons_pm_create accepts (checked here, before the box is admitted)
        if (nside < 8 || nside > 1624 || nside % 2) return fail("nside must be even, in [8, 1624]");
        return add_box(s, n, boxsize, nside, pos, vel, acc, acc_pm, box_id);
    } catch (const std::exception &e) {
        return fail(std::string("photons_server_add_box_pm: ") + e.what());
    }
}

int photons_server_remove_box(photons_server *s, int box_id)
{
    std::unique_lock<std::mutex> g(s->mu);
    if (s->boxes.find(box_id) == s->boxes.end()) return fail("no such box");
    // wait until no request of the box is queued or running (looked up again after
    // every wake-up: another thread may remove the box meanwhile)
    s->cv_done.wait(g, [&] {
        auto it = s->boxes.find(box_id);
        if (it == s->boxes.end()) return true;
        if (it->second.busy) return false;
        for (int64_t q : s->queue)
            if (s->reqs[q].box == box_id) return false;
        return true;
    });
    auto it = s->boxes.find(box_id);
    if (it == s->boxes.end()) return fail("no such box");
    Box b = it->second;
    s->boxes.erase(it);
    s->resident -= resident_bytes(b);
    g.unlock();
    (void)hipSetDevice(s->device);
    (void)hipFree(b.pos);
    (void)hipFree(b.vel);
    (void)hipFree(b.acc);
    if (b.acc_pm) (void)hipFree(b.acc_pm);
    return 0;
}

int photons_server_submit(photons_server *s, int box_id, double dkh, double dd, int64_t *request_id)
{
    std::unique_lock<std::mutex> g(s->mu);
    if (!s->fatal.empty()) return fail("server stopped after an error: " + s->fatal);
    if (s->boxes.find(box_id) == s->boxes.end()) return fail("no such box");
    const int64_t id = enqueue(s, box_id, kStep, dkh, dd);
    if (request_id) *request_id = id;
    g.unlock();
    s->cv_work.notify_all();
    return 0;
}

// Each request is waited on once: the wait that reports it releases it, so a server's
// bookkeeping does not grow with the requests it has served.
int photons_server_wait(photons_server *s, int64_t request_id, double *latency_ms)
{
    std::unique_lock<std::mutex> g(s->mu);
    if (s->reqs.find(request_id) == s->reqs.end()) return fail("no such request (or already waited on)");
    bool gone = false;
    s->cv_done.wait(g, [&] {
        auto it = s->reqs.find(request_id);
        gone = it == s->reqs.end();  // another wait reported it meanwhile
        return gone || it->second.done;
    });
    if (gone) return fail("request already waited on");
    auto it = s->reqs.find(request_id);
    const Request r = it->second;
    s->reqs.erase(it);
    if (latency_ms) *latency_ms = 1e3 * (r.t_done - r.t_submit);
    if (!r.error.empty()) return fail(r.error);
    return 0;
}

int photons_server_pause(photons_server *s, int paused)
{
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->paused = paused != 0;
    }
    s->cv_work.notify_all();
    return 0;
}

int photons_server_drain(photons_server *s)
{
    std::unique_lock<std::mutex> g(s->mu);
    s->cv_done.wait(g, [&] {
        if (!s->queue.empty()) return false;
        for (auto &kv : s->boxes)
            if (kv.second.busy) return false;
        return true;
    });
    if (!s->fatal.empty()) return fail(s->fatal);
    return 0;
}

int photons_server_read_box(photons_server *s, int box_id, double *pos, double *vel, double *acc)
{
    try {
        std::unique_lock<std::mutex> g(s->mu);
        auto it = s->boxes.find(box_id);
        if (it == s->boxes.end()) return fail("no such box");
        s->cv_done.wait(g, [&] { return !it->second.busy; });
        const Box b = it->second;
        it->second.busy = true;  // no lane takes the box while it is read
        g.unlock();
        SV_CHECK(hipSetDevice(s->device));
        const size_t nb = sizeof(double) * 3 * (size_t)b.n;
        std::string err;
        try {
            if (pos) SV_CHECK(hipMemcpy(pos, b.pos, nb, hipMemcpyDeviceToHost));
            if (vel) SV_CHECK(hipMemcpy(vel, b.vel, nb, hipMemcpyDeviceToHost));
            if (acc) SV_CHECK(hipMemcpy(acc, b.acc, nb, hipMemcpyDeviceToHost));
        } catch (const std::exception &e) {
            err = e.what();
        }
        g.lock();
        s->boxes[box_id].busy = false;
        g.unlock();
        s->cv_work.notify_all();
        s->cv_done.notify_all();
        if (!err.empty()) return fail(err);
        return 0;
    } catch (const std::exception &e) {
        return fail(std::string("photons_server_read_box: ") + e.what());
    }
}

int photons_server_read_pm(photons_server *s, int box_id, double *acc_pm)
{
    try {
        std::unique_lock<std::mutex> g(s->mu);
        auto it = s->boxes.find(box_id);
        if (it == s->boxes.end()) return fail("no such box");
        if (!it->second.acc_pm) return fail("the box has no mesh");
        s->cv_done.wait(g, [&] { return !it->second.busy; });
        const Box b = it->second;
        it->second.busy = true;  // no lane takes the box while it is read
        g.unlock();
        std::string err;
        try {
            SV_CHECK(hipSetDevice(s->device));
            if (acc_pm) SV_CHECK(hipMemcpy(acc_pm, b.acc_pm, sizeof(double) * 3 * (size_t)b.n, hipMemcpyDeviceToHost));
        } catch (const std::exception &e) {
            err = e.what();
        }
        g.lock();
        s->boxes[box_id].busy = false;
        g.unlock();
        s->cv_work.notify_all();
        s->cv_done.notify_all();
        if (!err.empty()) return fail(err);
        return 0;
    } catch (const std::exception &e) {
        return fail(std::string("photons_server_read_pm: ") + e.what());
    }
}

int photons_server_get_stats(photons_server *s, photons_server_stats *st)
{
    std::lock_guard<std::mutex> g(s->mu);
    st->requests_done = s->n_done;
    st->particles_stepped = s->particles_stepped;
    st->batches = s->batches;
    st->mean_boxes_per_batch = s->batches ? (double)s->batch_boxes / s->batches : 0.0;
    st->span_s = s->t_first >= 0.0 ? s->t_last - s->t_first : 0.0;
    st->nboxes = (int)s->boxes.size();
    st->max_lanes_busy = s->max_lanes_busy;
    st->resident_bytes = s->resident;
    st->reserved_bytes = s->reserve;
    st->queued = (int64_t)s->queue.size();
    return 0;
}

// kick and drift factors of the reference (src/initial.c:650-683): 128-block
// composite Simpson in ln a of a^-1/H and a^-2/H, H(a) = 0.1 sqrt(Om a^-3 + OL)
static double simpson_loga(double li, double lf, double om, double ox, int k)
{
    const int nblock = 128;
    const double h = (lf - li) / nblock;
    auto f = [&](double z) {
        const double H = 0.1 * std::sqrt(om * z * z * z + ox);
        return (k == 1 ? z : z * z) / H;
    };
    double s = h * f(1.0 / std::exp(li));
    for (int j = 1; j < nblock; j++) s += 2.0 * (1 + j % 2) * h * f(1.0 / std::exp(li + h * j));
    s += h * f(1.0 / std::exp(lf));
    return s / 3.0;
}

double photons_kick_factor(double loga_i, double loga_f, double omega_m, double omega_x)
{
    return simpson_loga(loga_i, loga_f, omega_m, omega_x, 1);
}

double photons_drift_factor(double loga_i, double loga_f, double omega_m, double omega_x)
{
    return simpson_loga(loga_i, loga_f, omega_m, omega_x, 2);
}

}  // extern "C"
