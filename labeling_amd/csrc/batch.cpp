// batch.cpp -- batches of 2D images as one hipGraph launch, the autotuner and
// its tuning table.
//
// A batch is planned once (batch_plan.hpp, pure host code tested on the CPU):
// every image gets a Geom2D (its tile grid, block-word slice of one shared
// workspace, first tile / final-workgroup id in the batch) in a device table,
// and tables give the image of every tile, final workgroup and scan chunk id.
// The batched kernels (tile, merge, flatten, final) are captured into a
// hipGraph; lbl_batch_run replays it: thousands of images cost four kernel
// launches instead of four per image.  The workspace and the tables are sized
// for the largest plan of every tile shape the batch will launch (all seven
// when autotuning), and every upload checks its plan against those sizes.
//
// The autotuner times every tile shape (kBleq2dTiles) with and without the
// flatten pass on the caller's data and keeps the fastest, per workload class
// (algorithm, single image or batch, floor(log2(pixels per image))).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <new>
#include <sstream>
#include <type_traits>

#include "bleq2d.hpp"
#include "uf_device.hpp"

using namespace lbl;

struct lbl_batch {
    lbl_ctx *ctx = nullptr;
    lbl_alg alg = LBL_BLEQ;
    uint32_t n = 0;
    std::vector<lbl_image2d> images;
    uint32_t *d_ncomp = nullptr;
    std::vector<uint32_t> hs, ws_;              // image heights and widths
    plan::Batch plan;                          // the plan of the current tile shape
    Geom2D *d_geoms = nullptr;
    void *ws = nullptr;
    size_t ws_cap = 0;                         // bytes
    uint32_t *d_tables = nullptr;
    uint64_t tables_cap = 0;                   // words
    BatchTables tables{};
    uint32_t total_tiles = 0, total_final = 0, total_chunks = 0;   // chunks: canonical scan (0: off)
    bool canonical = false;
    uint32_t tile = 0;
    bool flatten = false;
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace lbl {

int size_class_of(uint64_t pixels)
{
    int c = 0;
    while (pixels > 1) {
        pixels >>= 1;
        ++c;
    }
    return c;
}

}  // namespace lbl

namespace {

lbl_status bfail(lbl_ctx *ctx, lbl_status st, const std::string &msg)
{
    if (ctx) ctx->last_error = msg;
    return st;
}

lbl_status bhip(lbl_ctx *ctx, const char *what, hipError_t e)
{
    (void)hipGetLastError();
    return bfail(ctx, e == hipErrorOutOfMemory ? LBL_ENOMEM : LBL_EDEVICE,
                 std::string(what) + ": " + hipGetErrorString(e));
}

#define B_HIP(ctx, expr)                                  \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return bhip(ctx, #expr, e_); \
    } while (0)

struct DevGuard {
    int prev = -1;
    explicit DevGuard(int d)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != d) (void)hipSetDevice(d);
    }
    ~DevGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// Plan the batch for one tile shape (host only).
lbl_status make_plan(lbl_batch *b, uint32_t tile, plan::Batch &out)
{
    if (!plan::plan_batch(b->hs.data(), b->ws_.data(), b->n, tile, b->canonical, out))
        return bfail(b->ctx, LBL_EUNSUPPORTED, "batch too large: tile or workgroup ids exceed 32 bits");
    return LBL_OK;
}

// Make `pl` the current plan: its Geom2D table and id tables go to the device.
// The workspace and table buffers must already hold the plan.
lbl_status upload(lbl_batch *b, plan::Batch &pl)
{
    lbl_ctx *ctx = b->ctx;
    if (pl.ws_bytes > b->ws_cap || pl.table_words() > b->tables_cap)
        return bfail(ctx, LBL_EINVAL, "batch plan exceeds its buffers (" + std::to_string(pl.ws_bytes) + " > " +
                                          std::to_string(b->ws_cap) + " bytes or " + std::to_string(pl.table_words()) +
                                          " > " + std::to_string(b->tables_cap) + " table words)");
    uint8_t *base = static_cast<uint8_t *>(b->ws);
    std::vector<Geom2D> dev(b->n);
    for (uint32_t i = 0; i < b->n; ++i) {
        const lbl_image2d &im = b->images[i];
        const plan::Image &pi = pl.images[i];
        Geom2D g{};
        g.px = im.px;
        g.labels = im.labels;
        g.count = b->d_ncomp ? b->d_ncomp + i : nullptr;
        g.pitch = im.pitch_bytes;
        g.h = im.h;
        g.w = im.w;
        g.flags = b->ctx->opt.mode == LBL_LABELS_REFERENCE ? kRefLabels : 0u;
        if (b->canonical) g.flags |= kKeys | kCanonLabels;   // bit-position keys
        bleq2d_geometry(g, pl.tile_id, true);
        if (g.ntiles != pi.g.ntiles || g.nfinal != pi.g.nfinal || g.bw != pi.g.bw || g.bh != pi.g.bh)
            return bfail(ctx, LBL_EINVAL, "batch plan and kernel geometry disagree");
        g.tile_base = pi.tile_base;
        g.final_base = pi.final_base;
        g.chunk_base = pi.chunk_base;
        g.slots = reinterpret_cast<uint32_t *>(base + pi.ws_off);
        g.P = reinterpret_cast<uint32_t *>(base + pi.ws_off + align256(kCountBytes));
        if (b->canonical) g.canon = canon_layout(base + pi.canon_off, pi.nkeys, pi.nnodes);
        dev[i] = g;
    }
    std::vector<uint32_t> tab;
    tab.reserve(pl.table_words());
    tab.insert(tab.end(), pl.tile_img.begin(), pl.tile_img.end());
    tab.insert(tab.end(), pl.final_img.begin(), pl.final_img.end());
    tab.insert(tab.end(), pl.chunk_img.begin(), pl.chunk_img.end());
    B_HIP(ctx, hipMemcpyAsync(b->d_geoms, dev.data(), dev.size() * sizeof(Geom2D), hipMemcpyHostToDevice, ctx->stream));
    if (!tab.empty())
        B_HIP(ctx, hipMemcpyAsync(b->d_tables, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                  ctx->stream));
    B_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->tables = BatchTables{b->d_tables + pl.tile_tab(), b->d_tables + pl.final_tab(), b->d_tables + pl.chunk_tab()};
    b->total_tiles = pl.total_tiles;
    b->total_final = pl.total_final;
    b->total_chunks = pl.total_chunks;
    b->tile = pl.tile_id;
    b->plan = std::move(pl);
    return LBL_OK;
}

lbl_status launch(lbl_batch *b, hipStream_t s)
{
    hipError_t e = bleq2d_run_batch(b->d_geoms, b->n, b->tables, b->total_tiles, b->total_final, b->total_chunks,
                                    b->tile, b->flatten, s, nullptr);
    if (e != hipSuccess) return bhip(b->ctx, "bleq2d batch launch", e);
    return LBL_OK;
}

lbl_status build_graph(lbl_batch *b)
{
    lbl_ctx *ctx = b->ctx;
    hipStream_t cs;
    B_HIP(ctx, hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = bleq2d_run_batch(b->d_geoms, b->n, b->tables, b->total_tiles, b->total_final,
                                         b->total_chunks, b->tile, b->flatten, cs, nullptr);
        e = hipStreamEndCapture(cs, &b->graph);
        if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) return bhip(ctx, "batch graph capture", e);
    return LBL_OK;
}

float time_launch(lbl_ctx *ctx, const std::function<lbl_status()> &run, int reps)
{
    hipEvent_t a, c;
    if (hipEventCreate(&a) != hipSuccess) return -1.f;
    if (hipEventCreate(&c) != hipSuccess) { (void)hipEventDestroy(a); return -1.f; }
    float best = -1.f;
    if (run() == LBL_OK && hipStreamSynchronize(ctx->stream) == hipSuccess) {   // warm-up
        for (int r = 0; r < reps; ++r) {
            (void)hipEventRecord(a, ctx->stream);
            if (run() != LBL_OK) { best = -1.f; break; }
            (void)hipEventRecord(c, ctx->stream);
            float ms = 0.f;
            if (hipEventSynchronize(c) != hipSuccess || hipEventElapsedTime(&ms, a, c) != hipSuccess) {
                best = -1.f;
                break;
            }
            if (best < 0.f || ms < best) best = ms;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(c);
    return best < 0.f ? -1.f : best * 1000.f;
}

uint64_t mean_pixels(const lbl_batch *b)
{
    uint64_t px = 0;
    for (const auto &im : b->images) px += (uint64_t)im.h * im.w;
    return b->n ? px / b->n : 0;
}

// Buffers for the largest plan of the given tile shapes: the workspace (zeroed:
// the count slots start at zero) and the id tables.
lbl_status alloc_for(lbl_batch *b, const std::vector<uint32_t> &tiles)
{
    size_t ws = 256;
    uint64_t words = 1;
    for (uint32_t t : tiles) {
        plan::Batch pl;
        lbl_status st = make_plan(b, t, pl);
        if (st != LBL_OK) return st;
        ws = std::max<size_t>(ws, pl.ws_bytes);
        words = std::max<uint64_t>(words, pl.table_words());
    }
    B_HIP(b->ctx, hipMalloc(&b->ws, ws));
    b->ws_cap = ws;
    B_HIP(b->ctx, hipMemset(b->ws, 0, ws));
    B_HIP(b->ctx, hipMalloc(reinterpret_cast<void **>(&b->d_tables), words * sizeof(uint32_t)));
    b->tables_cap = words;
    return LBL_OK;
}

}  // namespace

extern "C" lbl_status lbl_batch_create(lbl_ctx *ctx, lbl_alg alg, const lbl_image2d *images, uint32_t n, int conn,
                                       uint32_t *d_ncomp, unsigned flags, lbl_batch **out)
{
    if (!ctx || !out) return LBL_EINVAL;
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!images || n == 0) return bfail(ctx, LBL_EINVAL, "lbl_batch_create: empty batch");
    if ((int)alg > (int)LBL_BLEQ && (int)alg <= (int)LBL_OCV)
        return bfail(ctx, LBL_EUNSUPPORTED, "lbl_batch_create: batches run bleq only");
    if (alg != LBL_BLEQ) return bfail(ctx, LBL_EINVAL, "unknown algorithm id");
    if (conn == 4) return bfail(ctx, LBL_EUNSUPPORTED, "bleq: method does not support 4-connectivity");
    if (conn != 0 && conn != 8) return bfail(ctx, LBL_EINVAL, "connectivity must be 0, 4 or 8 for 2D");
    for (uint32_t i = 0; i < n; ++i) {
        const lbl_image2d &im = images[i];
        if (!im.px || !im.labels || im.h == 0 || im.w == 0 || im.pitch_bytes < im.w)
            return bfail(ctx, LBL_EINVAL, "lbl_batch_create: image " + std::to_string(i) + " is invalid");
        if (((im.h + 1ull) / 2) * ((im.w + 1ull) / 2) > kMaxBlocks2D)
            return bfail(ctx, LBL_EUNSUPPORTED, "bleq 2D supports at most 2^28 2x2 blocks per image");
    }
    DevGuard dg(ctx->device);
    lbl_batch *b = new (std::nothrow) lbl_batch;
    if (!b) return bfail(ctx, LBL_ENOMEM, "host allocation failed");
    b->ctx = ctx;
    b->alg = alg;
    b->n = n;
    b->canonical = ctx->opt.mode == LBL_LABELS_CANONICAL;
    b->images.assign(images, images + n);
    for (uint32_t i = 0; i < n; ++i) {
        b->hs.push_back(images[i].h);
        b->ws_.push_back(images[i].w);
    }
    b->d_ncomp = d_ncomp;
    b->use_graph = !(flags & LBL_BATCH_NO_GRAPH);
    auto cleanup = [&](lbl_status st) {
        lbl_batch_destroy(b);
        return st;
    };
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&b->d_geoms), sizeof(Geom2D) * n);
    if (e != hipSuccess) return cleanup(bhip(ctx, "hipMalloc(batch table)", e));

    const TuneKey key{(int)alg, 1, size_class_of(mean_pixels(b))};
    uint32_t tile = 0;
    bool flatten = false;
    auto it = ctx->tuning.find(key);
    if (it != ctx->tuning.end()) {
        tile = it->second.tile;
        flatten = it->second.flatten;
    }
    lbl_status st;
    std::vector<uint32_t> shapes;
    if (flags & LBL_BATCH_AUTOTUNE)
        for (uint32_t t = 0; t < kBleq2dTileCount; ++t) shapes.push_back(t);
    else
        shapes.push_back(tile);
    if ((st = alloc_for(b, shapes)) != LBL_OK) return cleanup(st);
    if (flags & LBL_BATCH_AUTOTUNE) {   // time every configuration on this batch
        float best = -1.f;
        for (uint32_t t : shapes) {
            plan::Batch pl;
            if ((st = make_plan(b, t, pl)) != LBL_OK || (st = upload(b, pl)) != LBL_OK) return cleanup(st);
            for (int fl = 1; fl >= 0; --fl) {
                b->flatten = fl != 0;
                const float us = time_launch(ctx, [&] { return launch(b, ctx->stream); }, 3);
                if (us > 0.f && (best < 0.f || us < best)) {
                    best = us;
                    tile = t;
                    flatten = b->flatten;
                }
            }
        }
        if (best < 0.f) return cleanup(bfail(ctx, LBL_EDEVICE, "autotune: no configuration ran"));
        ctx->tuning[key] = TuneVal{tile, flatten, best};
    }
    b->flatten = flatten;
    {
        plan::Batch pl;
        if ((st = make_plan(b, tile, pl)) != LBL_OK || (st = upload(b, pl)) != LBL_OK) return cleanup(st);
    }
    if (b->use_graph && (st = build_graph(b)) != LBL_OK) return cleanup(st);
    *out = b;
    return LBL_OK;
}

extern "C" lbl_status lbl_batch_run(lbl_batch *b, void *stream)
{
    if (!b) return LBL_EINVAL;
    lbl_ctx *ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DevGuard dg(ctx->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (b->use_graph) {
        B_HIP(ctx, hipGraphLaunch(b->exec, s));
        return LBL_OK;
    }
    return launch(b, s);
}

extern "C" lbl_status lbl_batch_config(const lbl_batch *b, uint32_t *tile_id, uint32_t *tile_w, uint32_t *tile_h,
                                       int *flatten)
{
    if (!b) return LBL_EINVAL;
    if (tile_id) *tile_id = b->tile;
    if (tile_w) *tile_w = kBleq2dTiles[b->tile].bw;
    if (tile_h) *tile_h = kBleq2dTiles[b->tile].bh;
    if (flatten) *flatten = b->flatten ? 1 : 0;
    return LBL_OK;
}

extern "C" void lbl_batch_destroy(lbl_batch *b)
{
    if (!b) return;
    {
        DevGuard dg(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        (void)hipDeviceSynchronize();
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        if (b->d_geoms) (void)hipFree(b->d_geoms);
        if (b->d_tables) (void)hipFree(b->d_tables);
        if (b->ws) (void)hipFree(b->ws);
    }
    delete b;
}

// ---- single-image autotuning ------------------------------------------------------

extern "C" lbl_status lbl_autotune2d(lbl_ctx *ctx, lbl_alg alg, const uint8_t *d_px, uint32_t h, uint32_t w,
                                     size_t pitch_bytes, uint32_t *d_labels, void *stream)
{
    if (!ctx) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (alg != LBL_BLEQ) return bfail(ctx, LBL_EUNSUPPORTED, "autotuning covers bleq 2D");
    if (!d_px || !d_labels || !h || !w || pitch_bytes < w) return bfail(ctx, LBL_EINVAL, "lbl_autotune2d: bad image");
    if (((h + 1ull) / 2) * ((w + 1ull) / 2) > kMaxBlocks2D)
        return bfail(ctx, LBL_EUNSUPPORTED, "bleq 2D supports at most 2^28 2x2 blocks");
    DevGuard dg(ctx->device);
    // The timed launches run on the context's stream (its events time them),
    // ordered after the caller's stream, on which the image is ready; the
    // caller's stream then waits for them (they write d_labels).
    struct EventGuard {
        hipEvent_t e = nullptr;
        ~EventGuard()
        {
            if (e) (void)hipEventDestroy(e);
        }
    } order;
    hipStream_t caller = static_cast<hipStream_t>(stream);
    if (caller == ctx->stream) caller = nullptr;
    if (caller) {
        B_HIP(ctx, hipEventCreateWithFlags(&order.e, hipEventDisableTiming));
        B_HIP(ctx, hipEventRecord(order.e, caller));
        B_HIP(ctx, hipStreamWaitEvent(ctx->stream, order.e, 0));
    }
    const size_t need = bleq2d_workspace(h, w);
    if (ctx->ws_bytes < need) {
        if (ctx->ws) {
            B_HIP(ctx, hipDeviceSynchronize());
            B_HIP(ctx, hipFree(ctx->ws));
            ctx->ws = nullptr;
            ctx->ws_bytes = 0;
        }
        B_HIP(ctx, hipMalloc(&ctx->ws, need));
        ctx->ws_bytes = need;
        B_HIP(ctx, hipMemsetAsync(ctx->ws, 0, kCountBytes, ctx->stream));   // count slots start at zero
    }
    Image2D im{d_px, pitch_bytes, h, w, d_labels, ctx->d_misc};
    Options opt = ctx->opt;
    opt.mode = LBL_LABELS_ROOT;   // the label mode adds passes outside the tuned launches
    float best = -1.f;
    TuneVal bestv{0, true, 0.f};
    for (uint32_t t = 0; t < kBleq2dTileCount; ++t)
        for (int fl = 1; fl >= 0; --fl) {
            opt.tile2d = t;
            opt.flatten = fl != 0;
            const float us = time_launch(ctx, [&] {
                return bleq2d_run(im, opt, ctx->ws, ctx->stream, nullptr) == hipSuccess ? LBL_OK : LBL_EDEVICE;
            }, 5);
            if (us > 0.f && (best < 0.f || us < best)) {
                best = us;
                bestv = TuneVal{t, fl != 0, us};
            }
        }
    if (best < 0.f) return bfail(ctx, LBL_EDEVICE, "autotune: no configuration ran");
    ctx->tuning[TuneKey{(int)alg, 0, size_class_of((uint64_t)h * w)}] = bestv;
    if (caller) {
        B_HIP(ctx, hipEventRecord(order.e, ctx->stream));
        B_HIP(ctx, hipStreamWaitEvent(caller, order.e, 0));
    }
    return LBL_OK;
}

extern "C" lbl_status lbl_tuning_save(lbl_ctx *ctx, const char *path)
{
    if (!ctx || !path) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    std::ofstream f(path);
    if (!f) return bfail(ctx, LBL_EINVAL, std::string("cannot write ") + path);
    f << "{";
    bool first = true;
    for (const auto &kv : ctx->tuning) {
        f << (first ? "\n" : ",\n") << "  \"" << (kv.first.alg == LBL_BLEQ ? "bleq" : "lbeq") << "/"
          << (kv.first.kind == 0 ? "single2d" : "batch2d") << "/" << kv.first.size_class << "\": [" << kv.second.tile
          << ", " << (kv.second.flatten ? 1 : 0) << ", " << kv.second.us << "]";
        first = false;
    }
    f << "\n}\n";
    return f ? LBL_OK : bfail(ctx, LBL_EINVAL, std::string("write failed: ") + path);
}

extern "C" lbl_status lbl_tuning_load(lbl_ctx *ctx, const char *path)
{
    if (!ctx || !path) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    std::ifstream f(path);
    if (!f) return bfail(ctx, LBL_EINVAL, std::string("cannot read ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    size_t pos = 0;
    int loaded = 0;
    while ((pos = text.find('"', pos)) != std::string::npos) {
        const size_t end = text.find('"', pos + 1);
        if (end == std::string::npos) break;
        const std::string key = text.substr(pos + 1, end - pos - 1);
        char alg[16] = {0}, kind[16] = {0};
        int cls = 0;
        unsigned tile = 0;
        int fl = 1;
        float us = 0.f;
        const size_t lb = text.find('[', end);
        if (lb == std::string::npos) break;
        if (std::sscanf(key.c_str(), "%15[^/]/%15[^/]/%d", alg, kind, &cls) == 3 &&
            std::sscanf(text.c_str() + lb, "[%u, %d, %f]", &tile, &fl, &us) >= 2 && tile < kBleq2dTileCount) {
            const int a = !std::strcmp(alg, "bleq") ? LBL_BLEQ : LBL_LBEQ;
            const int k = !std::strcmp(kind, "single2d") ? 0 : 1;
            ctx->tuning[TuneKey{a, k, cls}] = TuneVal{tile, fl != 0, us};
            ++loaded;
        }
        pos = text.find(']', lb);
        if (pos == std::string::npos) break;
    }
    return loaded ? LBL_OK : bfail(ctx, LBL_EINVAL, std::string("no tuning entries in ") + path);
}
