// batch.cpp -- batches of 2D images as one hipGraph launch, the autotuner and
// its tuning table.
//
// A batch is planned once: every image gets a Geom2D (its tile grid, block-word
// slice of one shared workspace, first tile / final-workgroup id in the batch)
// in a device table, and the four batched kernels (tile, merge, flatten, final)
// are captured into a hipGraph.  lbl_batch_run replays the graph: thousands of
// images cost four kernel launches instead of four per image.
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
    std::vector<Geom2D> geoms;
    Geom2D *d_geoms = nullptr;
    void *ws = nullptr;
    uint32_t total_tiles = 0, total_final = 0, total_chunks = 0;   // chunks: canonical scan (0: off)
    bool canonical = false;
    uint32_t tile = 0;
    bool flatten = true;
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

// Geometry of every image for one tile shape; returns the workspace bytes.
size_t plan(lbl_batch *b, uint32_t tile)
{
    b->geoms.resize(b->n);
    uint64_t off = 0;
    uint32_t tb = 0, fb = 0, cb = 0;
    for (uint32_t i = 0; i < b->n; ++i) {
        const lbl_image2d &im = b->images[i];
        Geom2D g{};
        g.px = im.px;
        g.labels = im.labels;
        g.count = b->d_ncomp ? b->d_ncomp + i : nullptr;
        g.pitch = im.pitch_bytes;
        g.h = im.h;
        g.w = im.w;
        g.flags = b->ctx->opt.mode == LBL_LABELS_REFERENCE ? kRefLabels : 0u;
        if (b->canonical) g.flags |= kKeys | kCanonLabels;   // bit-position keys
        bleq2d_geometry(g, tile);
        g.tile_base = tb;
        g.final_base = fb;
        g.chunk_base = cb;
        g.slots = reinterpret_cast<uint32_t *>(off);   // offsets, rebased once ws exists
        g.P = reinterpret_cast<uint32_t *>(off + align256(kCountBytes));
        tb += g.ntiles;
        fb += g.nfinal;
        off += align256(kCountBytes) + align256((uint64_t)g.bw * g.bh * sizeof(uint32_t));
        if (b->canonical) {   // key array, bits, prefix and partial sums of this image
            const uint64_t nkeys = (uint64_t)g.h * g.bw, nodes = (uint64_t)g.bw * g.bh;
            g.canon = canon_layout(reinterpret_cast<void *>(off), nkeys, nodes);
            cb += canon_chunks(g.canon.nwords);
            off += canon_workspace(nkeys, nodes);
        }
        b->geoms[i] = g;
    }
    b->total_tiles = tb;
    b->total_final = fb;
    b->total_chunks = cb;
    b->tile = tile;
    return (size_t)off;
}

lbl_status upload(lbl_batch *b)
{
    lbl_ctx *ctx = b->ctx;
    std::vector<Geom2D> dev = b->geoms;
    uint8_t *base = static_cast<uint8_t *>(b->ws);
    // every planned pointer is an offset into the workspace (the first image's
    // count slots are at offset 0, so a null pointer here is offset 0)
    const auto rebase = [base](auto *&p) {
        p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(base + reinterpret_cast<uintptr_t>(p));
    };
    for (auto &g : dev) {
        rebase(g.P);
        rebase(g.slots);
        if (b->canonical) {
            rebase(g.canon.key);
            rebase(g.canon.bits);
            rebase(g.canon.prefix);
            rebase(g.canon.partial);
        }
    }
    B_HIP(ctx, hipMemcpyAsync(b->d_geoms, dev.data(), dev.size() * sizeof(Geom2D), hipMemcpyHostToDevice, ctx->stream));
    B_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LBL_OK;
}

lbl_status launch(lbl_batch *b, hipStream_t s)
{
    hipError_t e = bleq2d_run_batch(b->d_geoms, b->n, b->total_tiles, b->total_final, b->total_chunks, b->tile,
                                    b->flatten, s, nullptr);
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
        hipError_t le = bleq2d_run_batch(b->d_geoms, b->n, b->total_tiles, b->total_final, b->total_chunks, b->tile,
                                         b->flatten, cs, nullptr);
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

lbl_status alloc_ws(lbl_batch *b, size_t bytes)
{
    if (b->ws) {
        B_HIP(b->ctx, hipFree(b->ws));
        b->ws = nullptr;
    }
    B_HIP(b->ctx, hipMalloc(&b->ws, std::max<size_t>(bytes, 256)));
    B_HIP(b->ctx, hipMemset(b->ws, 0, std::max<size_t>(bytes, 256)));   // count slots start at zero
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
    if (alg == LBL_LBEQ) return bfail(ctx, LBL_EUNSUPPORTED, "lbl_batch_create: batches run bleq only");
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
    bool flatten = true;
    auto it = ctx->tuning.find(key);
    if (it != ctx->tuning.end()) {
        tile = it->second.tile;
        flatten = it->second.flatten;
    }
    lbl_status st;
    if (flags & LBL_BATCH_AUTOTUNE) {
        // workspace for the largest layout, then time every configuration
        size_t need = 0;
        for (uint32_t t = 0; t < kBleq2dTileCount; ++t) need = std::max(need, plan(b, t));
        if ((st = alloc_ws(b, need)) != LBL_OK) return cleanup(st);
        float best = -1.f;
        for (uint32_t t = 0; t < kBleq2dTileCount; ++t) {
            plan(b, t);
            if ((st = upload(b)) != LBL_OK) return cleanup(st);
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
    const size_t need = plan(b, tile);
    b->flatten = flatten;
    if (!b->ws || !(flags & LBL_BATCH_AUTOTUNE))
        if ((st = alloc_ws(b, need)) != LBL_OK) return cleanup(st);
    if ((st = upload(b)) != LBL_OK) return cleanup(st);
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
    (void)stream;
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
