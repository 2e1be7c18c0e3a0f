// serve.cpp -- continuous-batching request server: the HIP backend of the
// scheduler in serve_core.hpp, and the lbl_server_* entry points.
//
// A group (one request, or a run of fusable 2D bleq requests) runs on a slot:
// one HIP stream with a pinned result area and a completion event.  The
// group's input (16-byte pitched rows), labels and union-find workspace come
// from one device arena allocated at creation, so admission never calls
// hipMalloc.  Fused requests are labelled by the batched kernels (the
// Many-source kernels of bleq2d.hip) in one launch sequence.  Every enqueue's
// status is checked: the first failure stops the group's enqueueing and is
// reported as LBL_EDEVICE in each of its requests' completions.
#include <new>

#include "bleq2d.hpp"
#include "serve_core.hpp"

using namespace lbl;

namespace {

// A request with the context settings captured when it was submitted.
struct HipReq : serve::Req {
    Options opt;            // label mode, iteration cap, single-image tile shape and flatten
    uint32_t batch_tile = 0;   // launch configuration when fused (batch tuning of its size class)
    bool batch_flatten = false;
    bool is3d() const { return r.d != 0; }
    bool canonical() const { return mode == LBL_LABELS_CANONICAL; }
    uint64_t elems() const { return (uint64_t)(r.d ? r.d : 1) * r.h * r.w; }
    uint64_t in_bytes() const { return is3d() ? align256(elems()) : align256((uint64_t)r.h * ((r.w + 15u) & ~15u)); }
    uint64_t lab_bytes() const { return align256(elems() * 4); }
    uint64_t table_bytes = 0;   // fusable: its share of the batch id tables, for any tile shape
    uint64_t ws_bytes() const
    {
        const bool c = canonical();
        if (is3d())
            return align256(r.alg == LBL_BLEQ ? bleq3d_workspace(r.d, r.h, r.w, c) : lbeq3d_workspace(r.d, r.h, r.w, c));
        return align256(r.alg == LBL_BLEQ ? bleq2d_workspace(r.h, r.w, c) : lbeq2d_workspace(r.h, r.w, c));
    }
    uint64_t data_bytes() const { return in_bytes() + lab_bytes() + ws_bytes(); }
};

// Words of the batch id tables one 2D image adds (its tiles and final
// workgroups), for the tile shape with the most tiles.
uint64_t table_words_of(uint32_t h, uint32_t w)
{
    uint64_t m = 0;
    for (uint32_t t = 0; t < plan::kTileCount; ++t) {
        const plan::Geometry g = plan::geometry2d(h, w, t, plan::kBatchFinalChunks);
        m = std::max<uint64_t>(m, (uint64_t)g.ntiles + g.nfinal);
    }
    return m;
}

// Arena bytes of a group beyond its requests: the batch table, then n counts
// and the status word.
constexpr uint64_t group_extra_bytes(uint32_t n)
{
    return align256((uint64_t)n * sizeof(Geom2D)) + align256(4ull * n + 4);
}

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Geom2D *h_geoms = nullptr;     // pinned
    uint32_t *h_counts = nullptr;  // pinned: n counts + status word
    uint32_t *h_tables = nullptr;  // pinned: batch id tables (grown on demand)
    uint64_t tables_cap = 0;       // words
};

class HipBackend final : public serve::ServeBackend {
public:
    int device = 0;
    uint32_t max_batch = 1;
    uint8_t *arena_base = nullptr;
    std::vector<Slot> slots;

    void bind_thread() override { (void)hipSetDevice(device); }
    uint64_t group_extra(uint32_t n) const override { return group_extra_bytes(n); }

    int launch(serve::Group &g, bool fail_kernels) override
    {
        Slot &sl = slots[g.slot];
        hipStream_t s = sl.stream;
        uint8_t *base = arena_base + g.off;
        const uint32_t n = (uint32_t)g.reqs.size();
        uint32_t *d_counts = reinterpret_cast<uint32_t *>(base + align256((uint64_t)n * sizeof(Geom2D)));
        uint32_t *d_status = d_counts + n;
        hipError_t err = hipSuccess;
        // first failure wins; the sticky launch error is cleared so that later
        // groups are not blamed for it
        const auto ok = [&err](hipError_t e) {
            if (e != hipSuccess && err == hipSuccess) {
                err = e;
                (void)hipGetLastError();
            }
            return err == hipSuccess;
        };
        const auto kernels = [&](hipError_t e) { return ok(fail_kernels ? hipErrorLaunchFailure : e); };
        uint64_t cur = group_extra_bytes(n);
        if (!ok(hipMemsetAsync(d_counts, 0, 4ull * n + 4, s))) return err;

        if (n > 1 || g.reqs[0]->fusable) {
            // fused 2D bleq: per-image geometry table and id tables (batch_plan.hpp),
            // then the batched kernels, in the launch configuration tuned for the
            // group's largest request.  Arena: [Geom2D table][counts][id tables][requests]
            const HipReq *lead = static_cast<const HipReq *>(g.reqs[0]);
            for (serve::Req *r : g.reqs)
                if (r->r.h * (uint64_t)r->r.w > lead->r.h * (uint64_t)lead->r.w) lead = static_cast<const HipReq *>(r);
            const uint32_t tile = lead->batch_tile;
            std::vector<uint32_t> hs(n), ws(n);
            uint64_t table_room = 0;
            for (uint32_t i = 0; i < n; ++i) {
                hs[i] = g.reqs[i]->r.h;
                ws[i] = g.reqs[i]->r.w;
                table_room += static_cast<const HipReq *>(g.reqs[i])->table_bytes;
            }
            plan::Batch pl;
            if (!plan::plan_batch(hs.data(), ws.data(), n, tile, false, pl) || pl.table_words() * 4 > table_room) {
                ok(hipErrorInvalidValue);
                return err;
            }
            if (pl.table_words() > sl.tables_cap) {   // grow the slot's pinned staging
                if (sl.h_tables) (void)hipHostFree(sl.h_tables);
                sl.h_tables = nullptr;
                sl.tables_cap = 0;
                if (!ok(hipHostMalloc(reinterpret_cast<void **>(&sl.h_tables), pl.table_words() * 4, 0))) return err;
                sl.tables_cap = pl.table_words();
            }
            uint32_t *d_tab = reinterpret_cast<uint32_t *>(base + cur);
            cur += table_room;
            std::copy(pl.tile_img.begin(), pl.tile_img.end(), sl.h_tables + pl.tile_tab());
            std::copy(pl.final_img.begin(), pl.final_img.end(), sl.h_tables + pl.final_tab());
            for (uint32_t i = 0; i < n; ++i) {
                const HipReq *rq = static_cast<const HipReq *>(g.reqs[i]);
                const lbl_request &q = rq->r;
                const uint32_t pitch = (q.w + 15u) & ~15u;
                uint8_t *d_in = base + cur;
                uint32_t *d_lab = reinterpret_cast<uint32_t *>(d_in + rq->in_bytes());
                uint8_t *ws_i = reinterpret_cast<uint8_t *>(d_lab) + rq->lab_bytes();
                cur += rq->data_bytes();
                if (!ok(upload2d(d_in, pitch, q.px, q.w, q.w, q.h, reinterpret_cast<uint8_t *>(d_lab), rq->lab_bytes(), s)))
                    return err;
                if (!ok(hipMemsetAsync(ws_i, 0, kCountBytes, s))) return err;   // count slots start at zero
                Geom2D &gm = sl.h_geoms[i];
                gm = Geom2D{};
                gm.px = d_in;
                gm.labels = d_lab;
                gm.count = d_counts + i;
                gm.slots = reinterpret_cast<uint32_t *>(ws_i);
                gm.P = reinterpret_cast<uint32_t *>(ws_i + align256(kCountBytes));
                gm.pitch = pitch;
                gm.h = q.h;
                gm.w = q.w;
                gm.flags = rq->mode == LBL_LABELS_REFERENCE ? kRefLabels : 0u;
                bleq2d_geometry(gm, tile, true);
                gm.tile_base = pl.images[i].tile_base;
                gm.final_base = pl.images[i].final_base;
            }
            Geom2D *d_geoms = reinterpret_cast<Geom2D *>(base);
            if (!ok(hipMemcpyAsync(d_geoms, sl.h_geoms, sizeof(Geom2D) * n, hipMemcpyHostToDevice, s))) return err;
            if (!ok(hipMemcpyAsync(d_tab, sl.h_tables, pl.table_words() * 4, hipMemcpyHostToDevice, s))) return err;
            const BatchTables tabs{d_tab + pl.tile_tab(), d_tab + pl.final_tab(), d_tab + pl.chunk_tab()};
            if (!kernels(fail_kernels ? hipSuccess
                                      : bleq2d_run_batch(d_geoms, n, tabs, pl.total_tiles, pl.total_final, 0, tile,
                                                         lead->batch_flatten, s, nullptr)))
                return err;
            for (uint32_t i = 0; i < n; ++i) {
                const lbl_request &q = g.reqs[i]->r;
                if (!ok(hipMemcpyAsync(q.labels, sl.h_geoms[i].labels, (size_t)q.h * q.w * 4, hipMemcpyDeviceToHost, s)))
                    return err;
            }
        } else {
            const HipReq *rq = static_cast<const HipReq *>(g.reqs[0]);
            const lbl_request &q = rq->r;
            Options opt = rq->opt;
            opt.status = d_status;
            uint8_t *d_in = base + cur;
            uint32_t *d_lab = reinterpret_cast<uint32_t *>(d_in + rq->in_bytes());
            void *ws = reinterpret_cast<uint8_t *>(d_lab) + rq->lab_bytes();
            if (!ok(hipMemsetAsync(ws, 0, kCountBytes, s))) return err;   // count slots start at zero
            if (rq->is3d()) {
                if (!ok(hipMemcpyAsync(d_in, q.px, rq->elems(), hipMemcpyHostToDevice, s))) return err;
                Image3D im{d_in, q.d, q.h, q.w, d_lab, d_counts};
                if (!kernels(fail_kernels ? hipSuccess
                                          : (q.alg == LBL_BLEQ ? bleq3d_run(im, opt, ws, s, nullptr)
                                                               : lbeq3d_run(im, opt, ws, s, nullptr))))
                    return err;
            } else {
                const uint32_t pitch = (q.w + 15u) & ~15u;
                if (!ok(upload2d(d_in, pitch, q.px, q.w, q.w, q.h, reinterpret_cast<uint8_t *>(d_lab), rq->lab_bytes(), s)))
                    return err;
                Image2D im{d_in, pitch, q.h, q.w, d_lab, d_counts};
                if (!kernels(fail_kernels ? hipSuccess
                                          : (q.alg == LBL_BLEQ ? bleq2d_run(im, opt, ws, s, nullptr)
                                                               : lbeq2d_run(im, q.conn == 8 ? 8 : 4, opt, ws, s, nullptr))))
                    return err;
            }
            if (!ok(hipMemcpyAsync(q.labels, d_lab, rq->elems() * 4, hipMemcpyDeviceToHost, s))) return err;
        }
        if (!ok(hipMemcpyAsync(sl.h_counts, d_counts, 4ull * n + 4, hipMemcpyDeviceToHost, s))) return err;
        ok(hipEventRecord(sl.done, s));
        return err;
    }

    int query(int slot) override
    {
        const hipError_t q = hipEventQuery(slots[slot].done);
        return q == hipErrorNotReady ? kNotReady : (int)q;
    }
    void wait(int slot) override { (void)hipStreamSynchronize(slots[slot].stream); }
    uint32_t count(int slot, uint32_t i) const override { return slots[slot].h_counts[i]; }
    bool iter_cap(int slot, uint32_t n) const override { return (slots[slot].h_counts[n] & 1u) != 0; }
    std::string describe(int err) const override
    {
        return std::string("HIP error ") + std::to_string(err) + ": " + hipGetErrorString((hipError_t)err);
    }
};

}  // namespace

struct lbl_server {
    lbl_ctx *ctx = nullptr;
    lbl_server_config cfg{};
    HipBackend backend;
    serve::Scheduler *sched = nullptr;
};

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
    HipBackend &be = srv->backend;
    be.device = ctx->device;
    be.max_batch = cfg->max_batch;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&be.arena_base), cfg->hbm_budget_bytes);
    be.slots.resize(cfg->streams);
    for (Slot &s : be.slots) {
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
    srv->sched = new (std::nothrow) serve::Scheduler(be, cfg->streams, cfg->hbm_budget_bytes, cfg->max_batch);
    if (!srv->sched) {
        lbl_server_destroy(srv);
        return LBL_ENOMEM;
    }
    srv->sched->start();
    *out = srv;
    return LBL_OK;
}

extern "C" lbl_status lbl_server_submit(lbl_server *srv, const lbl_request *req, uint64_t *ticket)
{
    if (!srv || !req || !ticket) return LBL_EINVAL;
    const lbl_request &q = *req;
    if (!q.px || !q.labels || !q.h || !q.w) return LBL_EINVAL;
    if (q.alg != LBL_BLEQ && q.alg != LBL_LBEQ)   // the server runs the two union-find methods
        return (int)q.alg > (int)LBL_LBEQ && (int)q.alg <= (int)LBL_OCV ? LBL_EUNSUPPORTED : LBL_EINVAL;
    if (q.d == 0) {
        if (q.alg == LBL_BLEQ && q.conn != 0 && q.conn != 8) return q.conn == 4 ? LBL_EUNSUPPORTED : LBL_EINVAL;
        if (q.alg == LBL_LBEQ && q.conn != 0 && q.conn != 4 && q.conn != 8) return LBL_EINVAL;
        if (q.alg == LBL_BLEQ && ((q.h + 1ull) / 2) * ((q.w + 1ull) / 2) > kMaxBlocks2D) return LBL_EUNSUPPORTED;
    } else if (q.conn != 0 && q.conn != 26) {
        return LBL_EUNSUPPORTED;
    }
    if ((uint64_t)(q.d ? q.d : 1) * q.h * q.w > kMaxElems) return LBL_EUNSUPPORTED;
    HipReq *r = new (std::nothrow) HipReq;
    if (!r) return LBL_ENOMEM;
    r->r = q;
    {
        // the context's settings as of this call: label mode, caps, tuning
        lbl_ctx *ctx = srv->ctx;
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        r->opt = ctx->opt;
        r->mode = ctx->opt.mode;
        const int cls = size_class_of((uint64_t)q.h * q.w);
        if (q.d == 0) {
            const int talg = (q.alg == LBL_LBEQ && q.conn == 8) ? (int)LBL_BLEQ : (int)q.alg;
            auto it = ctx->tuning.find(TuneKey{talg, 0, cls});
            if (it != ctx->tuning.end()) {
                r->opt.tile2d = it->second.tile;
                r->opt.flatten = it->second.flatten;
            } else {
                r->opt.tile2d = plan::default_tile2d((uint64_t)q.h * q.w);
            }
            auto bt = ctx->tuning.find(TuneKey{(int)LBL_BLEQ, 1, cls});
            if (bt != ctx->tuning.end()) {
                r->batch_tile = bt->second.tile;
                r->batch_flatten = bt->second.flatten;
            }
        }
    }
    if (r->mode == LBL_LABELS_REFERENCE_UNCROPPED ||   // server results are dense, one label per input element
        (r->mode == LBL_LABELS_REFERENCE && q.d == 0 && q.alg == LBL_LBEQ && r->opt.iter_cap)) {
        delete r;
        return LBL_EUNSUPPORTED;
    }
    // fused launches: 2D bleq in root or reference values (canonical numbering runs per request)
    r->fusable = !r->is3d() && q.alg == LBL_BLEQ && !r->canonical();
    r->table_bytes = r->fusable ? align256(table_words_of(q.h, q.w) * 4) : 0;
    r->bytes = r->data_bytes() + r->table_bytes;
    return srv->sched->submit(r, ticket);
}

extern "C" lbl_status lbl_server_poll(lbl_server *srv, lbl_completion *out, uint32_t cap, uint32_t *n)
{
    if (!srv || !n || (cap && !out)) return LBL_EINVAL;
    return srv->sched->poll(out, cap, n);
}

extern "C" lbl_status lbl_server_drain(lbl_server *srv)
{
    if (!srv) return LBL_EINVAL;
    return srv->sched->drain();
}

extern "C" lbl_status lbl_server_inject_faults(lbl_server *srv, uint32_t n)
{
    if (!srv) return LBL_EINVAL;
    srv->sched->inject_faults(n);
    return LBL_OK;
}

extern "C" double lbl_server_now(const lbl_server *srv)
{
    return srv && srv->sched ? srv->sched->now() : 0.0;
}

extern "C" void lbl_server_destroy(lbl_server *srv)
{
    if (!srv) return;
    if (srv->sched) {
        srv->sched->stop();
        srv->sched->release_all();
    }
    HipBackend &be = srv->backend;
    (void)hipSetDevice(srv->ctx->device);
    for (Slot &s : be.slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
        if (s.h_geoms) (void)hipHostFree(s.h_geoms);
        if (s.h_counts) (void)hipHostFree(s.h_counts);
        if (s.h_tables) (void)hipHostFree(s.h_tables);
    }
    if (be.arena_base) (void)hipFree(be.arena_base);
    delete srv->sched;
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
