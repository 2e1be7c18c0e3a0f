// serve.cpp -- continuous-batching request server over the labeling kernels.
//
// One scheduler thread per server.  Each loop iteration
//   1. retires groups whose completion event has fired (results are already in
//      the callers' host buffers; their arena space and stream are released),
//   2. admits requests from the FIFO while a stream is free and the arena has
//      room: a run of consecutive 2D bleq requests becomes one batched launch
//      (the Many-source kernels of bleq2d.hip), any other request runs alone;
//      each group is copied in, labelled and copied out on its own stream, so
//      groups overlap on the device.
// Admission is by device footprint: input (16-byte pitched rows), labels and
// union-find workspace of every request come from one arena allocated at
// creation (first-fit with coalescing), so admitting never calls hipMalloc.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <new>
#include <thread>

#include "bleq2d.hpp"

using namespace lbl;

namespace {

class Arena {
public:
    void init(uint64_t cap)
    {
        free_.clear();
        if (cap) free_[0] = cap;
    }
    uint64_t largest() const
    {
        uint64_t m = 0;
        for (const auto &kv : free_) m = std::max(m, kv.second);
        return m;
    }
    bool alloc(uint64_t bytes, uint64_t &off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const uint64_t rem = it->second - bytes, noff = it->first + bytes;
            free_.erase(it);
            if (rem) free_[noff] = rem;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<uint64_t, uint64_t> free_;
};

struct Req {
    uint64_t ticket;
    lbl_request r;
    lbl_label_mode mode = LBL_LABELS_ROOT;   // the context's label mode at submission
    double t_submit = 0, t_start = 0;
    bool is3d() const { return r.d != 0; }
    bool canonical() const { return mode == LBL_LABELS_CANONICAL; }
    // fused 2D bleq launches; canonical numbering runs per request
    bool batchable() const { return !is3d() && r.alg == LBL_BLEQ && !canonical(); }
    uint64_t elems() const { return (uint64_t)(r.d ? r.d : 1) * r.h * r.w; }
    uint64_t in_bytes() const { return is3d() ? align256(elems()) : align256((uint64_t)r.h * ((r.w + 15u) & ~15u)); }
    uint64_t lab_bytes() const { return align256(elems() * 4); }
    uint64_t ws_bytes() const
    {
        const bool c = canonical();
        if (is3d())
            return align256(r.alg == LBL_BLEQ ? bleq3d_workspace(r.d, r.h, r.w, c) : lbeq3d_workspace(r.d, r.h, r.w, c));
        return align256(r.alg == LBL_BLEQ ? bleq2d_workspace(r.h, r.w, c) : lbeq2d_workspace(r.h, r.w, c));
    }
    uint64_t bytes() const { return in_bytes() + lab_bytes() + ws_bytes(); }
};

constexpr uint64_t kGroupExtra(uint32_t n) { return align256((uint64_t)n * sizeof(Geom2D)) + align256(4ull * n + 4); }

struct Group {
    std::vector<Req *> reqs;
    int slot = -1;
    uint64_t off = 0, bytes = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Group *group = nullptr;
    Geom2D *h_geoms = nullptr;   // pinned
    uint32_t *h_counts = nullptr;  // pinned: n counts + status word
};

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct lbl_server {
    lbl_ctx *ctx = nullptr;
    lbl_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    uint8_t *arena_base = nullptr;
    Arena arena;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> pending;
    std::vector<lbl_completion> finished;
    uint64_t next_ticket = 1, submitted = 0, completed = 0;
    bool stop = false;
    std::thread th;
    std::string fatal;

    void loop();
    bool admit_one(std::unique_lock<std::mutex> &lk);
    void launch(Group *g);
    void retire(Slot &s);
};

void lbl_server::launch(Group *g)
{
    Slot &sl = slots[g->slot];
    hipStream_t s = sl.stream;
    uint8_t *base = arena_base + g->off;
    const uint32_t n = (uint32_t)g->reqs.size();
    uint32_t *d_counts = reinterpret_cast<uint32_t *>(base + align256((uint64_t)n * sizeof(Geom2D)));
    uint32_t *d_status = d_counts + n;
    uint64_t cur = kGroupExtra(n);
    Options opt = ctx->opt;
    opt.status = d_status;
    (void)hipMemsetAsync(d_counts, 0, 4ull * n + 4, s);
    const double t = seconds_since(t0);
    for (Req *r : g->reqs) r->t_start = t;

    if (n > 1 || g->reqs[0]->batchable()) {
        // batched 2D bleq: per-image geometry table, then the four batched kernels
        uint32_t tb = 0, fb = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const lbl_request &q = g->reqs[i]->r;
            const uint32_t pitch = (q.w + 15u) & ~15u;
            uint8_t *d_in = base + cur;
            uint32_t *d_lab = reinterpret_cast<uint32_t *>(d_in + g->reqs[i]->in_bytes());
            uint8_t *ws = reinterpret_cast<uint8_t *>(d_lab) + g->reqs[i]->lab_bytes();
            uint32_t *d_P = reinterpret_cast<uint32_t *>(ws + align256(kCountBytes));
            cur += g->reqs[i]->bytes();
            (void)upload2d(d_in, pitch, q.px, q.w, q.w, q.h, reinterpret_cast<uint8_t *>(d_lab), g->reqs[i]->lab_bytes(), s);
            Geom2D &gm = sl.h_geoms[i];
            gm = Geom2D{};
            gm.px = d_in;
            gm.labels = d_lab;
            gm.count = d_counts + i;
            gm.P = d_P;
            gm.pitch = pitch;
            gm.h = q.h;
            gm.w = q.w;
            gm.flags = g->reqs[i]->mode == LBL_LABELS_REFERENCE ? kRefLabels : 0u;
            bleq2d_geometry(gm, 0);
            gm.slots = reinterpret_cast<uint32_t *>(ws);   // count slots first, zeroed below
            (void)hipMemsetAsync(ws, 0, kCountBytes, s);
            gm.tile_base = tb;
            gm.final_base = fb;
            tb += gm.ntiles;
            fb += gm.nfinal;
        }
        Geom2D *d_geoms = reinterpret_cast<Geom2D *>(base);
        (void)hipMemcpyAsync(d_geoms, sl.h_geoms, sizeof(Geom2D) * n, hipMemcpyHostToDevice, s);
        (void)bleq2d_run_batch(d_geoms, n, tb, fb, 0, 0, true, s, nullptr);
        for (uint32_t i = 0; i < n; ++i) {
            const lbl_request &q = g->reqs[i]->r;
            (void)hipMemcpyAsync(q.labels, sl.h_geoms[i].labels, (size_t)q.h * q.w * 4, hipMemcpyDeviceToHost, s);
        }
    } else {
        Req *rq = g->reqs[0];
        const lbl_request &q = rq->r;
        opt.mode = rq->mode;
        uint8_t *d_in = base + cur;
        uint32_t *d_lab = reinterpret_cast<uint32_t *>(d_in + rq->in_bytes());
        void *ws = reinterpret_cast<uint8_t *>(d_lab) + rq->lab_bytes();
        (void)hipMemsetAsync(ws, 0, kCountBytes, s);   // count slots (workspace offset 0) start at zero
        if (rq->is3d()) {
            (void)hipMemcpyAsync(d_in, q.px, rq->elems(), hipMemcpyHostToDevice, s);
            Image3D im{d_in, q.d, q.h, q.w, d_lab, d_counts};
            if (q.alg == LBL_BLEQ) (void)bleq3d_run(im, opt, ws, s, nullptr);
            else (void)lbeq3d_run(im, opt, ws, s, nullptr);
        } else {
            const uint32_t pitch = (q.w + 15u) & ~15u;
            (void)upload2d(d_in, pitch, q.px, q.w, q.w, q.h, reinterpret_cast<uint8_t *>(d_lab), rq->lab_bytes(), s);
            Image2D im{d_in, pitch, q.h, q.w, d_lab, d_counts};
            if (q.alg == LBL_BLEQ) (void)bleq2d_run(im, opt, ws, s, nullptr);   // canonical numbering
            else (void)lbeq2d_run(im, q.conn == 8 ? 8 : 4, opt, ws, s, nullptr);
        }
        (void)hipMemcpyAsync(q.labels, d_lab, rq->elems() * 4, hipMemcpyDeviceToHost, s);
    }
    (void)hipMemcpyAsync(sl.h_counts, d_counts, 4ull * n + 4, hipMemcpyDeviceToHost, s);
    (void)hipEventRecord(sl.done, s);
    sl.group = g;
}

// Admit the next group from the FIFO if a stream is free and it fits.
bool lbl_server::admit_one(std::unique_lock<std::mutex> &)
{
    if (pending.empty()) return false;
    int slot = -1;
    for (size_t i = 0; i < slots.size(); ++i)
        if (!slots[i].group) {
            slot = (int)i;
            break;
        }
    if (slot < 0) return false;
    const uint64_t room = arena.largest();
    Group *g = new Group;
    uint64_t bytes = 0;
    const bool batch = pending.front()->batchable();
    for (size_t k = 0; k < pending.size(); ++k) {
        Req *r = pending[k];
        if (k > 0 && (!batch || !r->batchable() || g->reqs.size() >= cfg.max_batch)) break;
        const uint64_t nb = bytes + r->bytes();
        if (nb + kGroupExtra((uint32_t)g->reqs.size() + 1) > room) break;
        bytes = nb;
        g->reqs.push_back(r);
    }
    if (g->reqs.empty()) {
        delete g;
        return false;   // head of line waits for memory
    }
    g->bytes = bytes + kGroupExtra((uint32_t)g->reqs.size());
    if (!arena.alloc(g->bytes, g->off)) {
        delete g;
        return false;
    }
    pending.erase(pending.begin(), pending.begin() + (long)g->reqs.size());
    g->slot = slot;
    launch(g);
    return true;
}

void lbl_server::retire(Slot &s)
{
    Group *g = s.group;
    const double t = seconds_since(t0);
    const uint32_t n = (uint32_t)g->reqs.size();
    const bool iter_cap = (s.h_counts[n] & 1u) != 0;
    for (uint32_t i = 0; i < n; ++i) {
        Req *r = g->reqs[i];
        finished.push_back(lbl_completion{r->ticket, iter_cap ? LBL_EITER : LBL_OK, s.h_counts[i], r->t_submit,
                                          r->t_start, t});
        delete r;
    }
    completed += n;
    arena.release(g->off, g->bytes);
    s.group = nullptr;
    delete g;
}

void lbl_server::loop()
{
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool progressed = false;
        for (Slot &s : slots) {
            if (!s.group) continue;
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess && fatal.empty()) fatal = hipGetErrorString(q);
            retire(s);
            progressed = true;
        }
        if (progressed) cv_done.notify_all();
        while (admit_one(lk)) progressed = true;
        bool busy = false;
        for (Slot &s : slots) busy |= s.group != nullptr;
        if (stop && pending.empty() && !busy) break;
        if (!progressed) {
            if (!busy && pending.empty()) {
                cv_work.wait(lk);
            } else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            }
        }
    }
}

extern "C" lbl_status lbl_server_create(lbl_ctx *ctx, const lbl_server_config *cfg, lbl_server **out)
{
    if (!ctx || !cfg || !out) return LBL_EINVAL;
    *out = nullptr;
    if (cfg->streams < 1 || cfg->streams > 16 || cfg->max_batch < 1 || cfg->hbm_budget_bytes < (1u << 20)) {
        ctx->last_error = "lbl_server_create: need 1..16 streams, max_batch >= 1 and a budget of at least 1 MiB";
        return LBL_EINVAL;
    }
    lbl_server *srv = new (std::nothrow) lbl_server;
    if (!srv) return LBL_ENOMEM;
    srv->ctx = ctx;
    srv->cfg = *cfg;
    srv->t0 = std::chrono::steady_clock::now();
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&srv->arena_base), cfg->hbm_budget_bytes);
    srv->slots.resize(cfg->streams);
    for (Slot &s : srv->slots) {
        if (e != hipSuccess) break;
        if ((e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking)) != hipSuccess) break;
        if ((e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming)) != hipSuccess) break;
        if ((e = hipHostMalloc(reinterpret_cast<void **>(&s.h_geoms), sizeof(Geom2D) * cfg->max_batch, 0)) != hipSuccess) break;
        if ((e = hipHostMalloc(reinterpret_cast<void **>(&s.h_counts), 4ull * cfg->max_batch + 4, 0)) != hipSuccess) break;
    }
    if (prev >= 0) (void)hipSetDevice(prev);
    if (e != hipSuccess) {
        ctx->last_error = std::string("lbl_server_create: ") + hipGetErrorString(e);
        (void)hipGetLastError();
        lbl_server_destroy(srv);
        return e == hipErrorOutOfMemory ? LBL_ENOMEM : LBL_EDEVICE;
    }
    srv->arena.init(cfg->hbm_budget_bytes);
    srv->th = std::thread([srv] { srv->loop(); });
    *out = srv;
    return LBL_OK;
}

extern "C" lbl_status lbl_server_submit(lbl_server *srv, const lbl_request *req, uint64_t *ticket)
{
    if (!srv || !req || !ticket) return LBL_EINVAL;
    const lbl_request &q = *req;
    if (!q.px || !q.labels || !q.h || !q.w) return LBL_EINVAL;
    if (q.alg != LBL_BLEQ && q.alg != LBL_LBEQ) return LBL_EINVAL;
    if (q.d == 0) {
        if (q.alg == LBL_BLEQ && q.conn != 0 && q.conn != 8) return q.conn == 4 ? LBL_EUNSUPPORTED : LBL_EINVAL;
        if (q.alg == LBL_LBEQ && q.conn != 0 && q.conn != 4 && q.conn != 8) return LBL_EINVAL;
        if (q.alg == LBL_BLEQ && ((q.h + 1ull) / 2) * ((q.w + 1ull) / 2) > kMaxBlocks2D) return LBL_EUNSUPPORTED;
    } else if (q.conn != 0 && q.conn != 26) {
        return LBL_EUNSUPPORTED;
    }
    const lbl_label_mode mode = srv->ctx->opt.mode;
    if (mode == LBL_LABELS_REFERENCE && (q.d != 0 || q.alg != LBL_BLEQ)) return LBL_EUNSUPPORTED;
    Req *r = new (std::nothrow) Req;
    if (!r) return LBL_ENOMEM;
    r->r = q;
    r->mode = mode;
    if ((uint64_t)(q.d ? q.d : 1) * q.h * q.w > kMaxElems) {
        delete r;
        return LBL_EUNSUPPORTED;
    }
    if (r->bytes() + kGroupExtra(1) > srv->cfg.hbm_budget_bytes) {
        delete r;
        return LBL_ENOMEM;
    }
    std::lock_guard<std::mutex> lk(srv->mu);
    if (!srv->fatal.empty()) {
        delete r;
        return LBL_EDEVICE;
    }
    r->ticket = srv->next_ticket++;
    r->t_submit = seconds_since(srv->t0);
    *ticket = r->ticket;
    srv->pending.push_back(r);
    ++srv->submitted;
    srv->cv_work.notify_one();
    return LBL_OK;
}

extern "C" lbl_status lbl_server_poll(lbl_server *srv, lbl_completion *out, uint32_t cap, uint32_t *n)
{
    if (!srv || !n || (cap && !out)) return LBL_EINVAL;
    std::lock_guard<std::mutex> lk(srv->mu);
    const uint32_t k = (uint32_t)std::min<size_t>(cap, srv->finished.size());
    std::copy(srv->finished.begin(), srv->finished.begin() + k, out);
    srv->finished.erase(srv->finished.begin(), srv->finished.begin() + k);
    *n = k;
    return srv->fatal.empty() ? LBL_OK : LBL_EDEVICE;
}

extern "C" lbl_status lbl_server_drain(lbl_server *srv)
{
    if (!srv) return LBL_EINVAL;
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [srv] { return srv->completed == srv->submitted; });
    return srv->fatal.empty() ? LBL_OK : LBL_EDEVICE;
}

extern "C" double lbl_server_now(const lbl_server *srv)
{
    return srv ? seconds_since(srv->t0) : 0.0;
}

extern "C" void lbl_server_destroy(lbl_server *srv)
{
    if (!srv) return;
    if (srv->th.joinable()) {
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->stop = true;
        }
        srv->cv_work.notify_all();
        srv->th.join();
    }
    for (Req *r : srv->pending) delete r;
    (void)hipSetDevice(srv->ctx->device);
    for (Slot &s : srv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.group) {
            for (Req *r : s.group->reqs) delete r;
            delete s.group;
        }
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
        if (s.h_geoms) (void)hipHostFree(s.h_geoms);
        if (s.h_counts) (void)hipHostFree(s.h_counts);
    }
    if (srv->arena_base) (void)hipFree(srv->arena_base);
    delete srv;
}

extern "C" void *lbl_host_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, 0) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

extern "C" void lbl_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}
