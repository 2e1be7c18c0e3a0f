// capi.cpp -- extern "C" entry points declared in include/labeling_amd.h.
//
// Mirrors the reference's label-call contract (IOCLLabeling::Label,
// LabelingTools.cpp:98-132; IOCLLabeling3D::Label, :180-221) without its
// padding, per-call allocation and blocking per-iteration transfers: device
// buffers are cached on the context, the labeling kernels are enqueued
// back-to-back on one stream, and errors are returned as status codes.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "batch_plan.hpp"
#include "lbl_internal.hpp"

using namespace lbl;

namespace {

thread_local std::string g_create_error;

lbl_status fail(lbl_ctx *ctx, lbl_status st, const std::string &msg)
{
    if (ctx) ctx->last_error = msg;
    else g_create_error = msg;
    return st;
}

lbl_status hip_fail(lbl_ctx *ctx, const char *what, hipError_t e)
{
    return fail(ctx, e == hipErrorOutOfMemory ? LBL_ENOMEM : LBL_EDEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define LBL_HIP(ctx, expr)                                  \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(ctx, #expr, e_); \
    } while (0)

// Restores the caller's current device on scope exit.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// Grow a cached device buffer; waits for the device before freeing the old one.
lbl_status ensure(lbl_ctx *ctx, void **buf, size_t *cap, size_t need, bool *fresh)
{
    if (fresh) *fresh = false;
    if (*cap >= need && *buf) return LBL_OK;
    if (*buf) {
        LBL_HIP(ctx, hipDeviceSynchronize());
        LBL_HIP(ctx, hipFree(*buf));
        *buf = nullptr;
        *cap = 0;
    }
    need = align256(need < 256 ? 256 : need);
    hipError_t e = hipMalloc(buf, need);
    if (e != hipSuccess) {
        *buf = nullptr;
        (void)hipGetLastError();
        return fail(ctx, LBL_ENOMEM, "device allocation of " + std::to_string(need) + " bytes failed: " +
                                         hipGetErrorString(e));
    }
    *cap = need;
    if (fresh) *fresh = true;
    return LBL_OK;
}

// Connectivity defaults and support, as in the reference:
// bleq 2D: 8 (COH_4 rejected, LabelingAlgs.cpp:338,1070); lbeq 2D: default 4
// (LabelingAlgs.cpp:213), 8 on request; 3D: default only = 26 (LabelingTools.cpp:185).
// gr-block and runeq reject 4 (LabelingAlgs.cpp:51,537,1242; runeq then joins
// runs by shared columns, its own rule); he-run and ocv use 8 only when asked
// (LabelingAlgs.cpp:42,140); bin ignores it.
lbl_status normalize_conn2d(lbl_ctx *ctx, lbl_alg alg, int conn, int *out)
{
    if (conn != 0 && conn != 4 && conn != 8)
        return fail(ctx, LBL_EINVAL, "connectivity must be 0, 4 or 8 for 2D, got " + std::to_string(conn));
    switch (alg) {
    case LBL_BLEQ:
    case LBL_GRBLOCK:
    case LBL_RUNEQ:
        if (conn == 4)
            return fail(ctx, LBL_EUNSUPPORTED, std::string(alg == LBL_BLEQ ? "bleq" : alg == LBL_GRBLOCK ? "gr-block" : "runeq") +
                                                   ": method does not support 4-connectivity");
        *out = 8;
        return LBL_OK;
    case LBL_LBEQ:
    case LBL_HERUN:
    case LBL_OCV: *out = conn == 8 ? 8 : 4; return LBL_OK;
    case LBL_BIN: *out = conn == 0 ? 8 : conn; return LBL_OK;
    }
    return fail(ctx, LBL_EINVAL, "unknown algorithm id " + std::to_string((int)alg));
}

lbl_status normalize_conn3d(lbl_ctx *ctx, lbl_alg alg, int conn)
{
    if ((int)alg < (int)LBL_BLEQ || (int)alg > (int)LBL_OCV)
        return fail(ctx, LBL_EINVAL, "unknown algorithm id " + std::to_string((int)alg));
    if (alg != LBL_BLEQ && alg != LBL_LBEQ && alg != LBL_BIN)   // the registry's 3D entries (main.cpp:103-130)
        return fail(ctx, LBL_EUNSUPPORTED, "this method has no 3D variant");
    if (conn == 0 || conn == 26) return LBL_OK;
    return fail(ctx, LBL_EUNSUPPORTED, "3D labeling supports only the default 26-connectivity");
}

// The forest a 2D method runs on: 2x2 blocks (8-connectivity) or 1x2 blocks (4).
bool on_2x2_blocks(lbl_alg alg, int conn)
{
    return alg == LBL_BLEQ || alg == LBL_GRBLOCK || ((alg == LBL_HERUN || alg == LBL_OCV || alg == LBL_LBEQ) && conn == 8);
}

lbl_status check_mode(lbl_ctx *ctx, lbl_alg alg, bool is3d)
{
    const lbl_label_mode m = ctx->opt.mode;
    if (m == LBL_LABELS_REFERENCE_UNCROPPED && !is3d)
        return fail(ctx, LBL_EUNSUPPORTED, "LBL_LABELS_REFERENCE_UNCROPPED is defined for 3D only (2D output is cropped)");
    if (m == LBL_LABELS_REFERENCE && !is3d && alg == LBL_LBEQ && ctx->opt.iter_cap)
        return fail(ctx, LBL_EUNSUPPORTED, "LBL_LABELS_REFERENCE is not available on the lbeq sweep path (iteration cap set)");
    return LBL_OK;
}

// Elements of a 3D label buffer: the input's, or the reference's padded volume.
uint64_t labels3d_elems(lbl_label_mode mode, uint32_t d, uint32_t h, uint32_t w)
{
    if (mode != LBL_LABELS_REFERENCE_UNCROPPED) return (uint64_t)d * h * w;
    return (uint64_t)ref_pad3(d) * ref_pad3(h) * ref_pad3(w);
}

lbl_status check_size2d(lbl_ctx *ctx, lbl_alg alg, int conn, uint32_t h, uint32_t w, size_t pitch)
{
    if (h == 0 || w == 0) return fail(ctx, LBL_EINVAL, "input image is empty");
    if (pitch < w) return fail(ctx, LBL_EINVAL, "pitch is smaller than the width");
    if ((uint64_t)h * w > kMaxElems) return fail(ctx, LBL_EUNSUPPORTED, "image has more than 2^32-2 pixels");
    // lbeq falls back to its sweeps beyond the block forests' limit; the others run on the forests only
    if (alg != LBL_LBEQ && alg != LBL_BIN) {
        const uint64_t blocks = on_2x2_blocks(alg, conn) ? ((h + 1ull) / 2) * ((w + 1ull) / 2) : h * ((w + 1ull) / 2);
        if (blocks > kMaxBlocks2D)
            return fail(ctx, LBL_EUNSUPPORTED, "this method supports at most 2^28 blocks (2x2, or 1x2 with 4-connectivity)");
    }
    if ((alg == LBL_RUNEQ || alg == LBL_HERUN) && (uint64_t)h * ((w + 63ull) / 64 * 64) > kMaxElems)
        return fail(ctx, LBL_EUNSUPPORTED, "run numbering needs h * (w rounded up to 64) below 2^32");
    // reference values index the reference's padded image (or its block grid)
    if (ctx->opt.mode == LBL_LABELS_REFERENCE && ref_pad2(h) * ref_pad2(w) > kMaxElems)
        return fail(ctx, LBL_EUNSUPPORTED, "reference label values need a padded image of at most 2^32-2 pixels");
    return LBL_OK;
}

lbl_status check_size3d(lbl_ctx *ctx, uint32_t d, uint32_t h, uint32_t w)
{
    if (d == 0 || h == 0 || w == 0) return fail(ctx, LBL_EINVAL, "input volume is empty");
    if ((uint64_t)d * h * w > kMaxElems) return fail(ctx, LBL_EUNSUPPORTED, "volume has more than 2^32-2 voxels");
    const lbl_label_mode m = ctx->opt.mode;
    if ((m == LBL_LABELS_REFERENCE || m == LBL_LABELS_REFERENCE_UNCROPPED) &&
        labels3d_elems(LBL_LABELS_REFERENCE_UNCROPPED, d, h, w) > kMaxElems)
        return fail(ctx, LBL_EUNSUPPORTED, "reference label values need a padded volume of at most 2^32-2 voxels");
    return LBL_OK;
}

size_t workspace2d(lbl_alg alg, int conn, uint32_t h, uint32_t w, lbl_label_mode mode)
{
    const bool canon = mode == LBL_LABELS_CANONICAL;
    switch (alg) {
    case LBL_BLEQ: return bleq2d_workspace(h, w, canon);
    case LBL_GRBLOCK: return grblock2d_workspace(h, w);
    case LBL_RUNEQ: return runs2d_workspace(h, w, false, canon);
    case LBL_HERUN: return runs2d_workspace(h, w, conn == 8, canon);
    case LBL_OCV: return conn == 8 ? bleq2d_workspace(h, w, true) : bleq2d_conn4_workspace(h, w, true);
    case LBL_BIN: return align256(kCountBytes);
    default: return lbeq2d_workspace(h, w, canon);
    }
}

size_t workspace3d(lbl_alg alg, uint32_t d, uint32_t h, uint32_t w, lbl_label_mode mode)
{
    const bool canon = mode == LBL_LABELS_CANONICAL;
    if (alg == LBL_BIN) return align256(kCountBytes);
    return alg == LBL_BLEQ ? bleq3d_workspace(d, h, w, canon) : lbeq3d_workspace(d, h, w, canon);
}

hipStream_t pick_stream(lbl_ctx *ctx, void *stream)
{
    return stream ? static_cast<hipStream_t>(stream) : ctx->stream;
}

}  // namespace

// ---- KernelTimer ---------------------------------------------------------------

void lbl::KernelTimer::begin(hipStream_t s)
{
    used = 0;
    names.clear();
    mark(s, nullptr);
}

void lbl::KernelTimer::mark(hipStream_t s, const char *name)
{
    if (used == events.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        events.push_back(e);
    }
    if (hipEventRecord(events[used], s) != hipSuccess) return;
    ++used;
    if (name) names.push_back(name);
}

// ---- context -----------------------------------------------------------------------

extern "C" lbl_status lbl_create(lbl_ctx **out, int device_ordinal)
{
    if (!out) return fail(nullptr, LBL_EINVAL, "lbl_create: out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(nullptr, LBL_EDEVICE, std::string("no HIP device available: ") + hipGetErrorString(e));
    if (device_ordinal < 0 || device_ordinal >= n)
        return fail(nullptr, LBL_EINVAL, "device ordinal out of range");
    lbl_ctx *ctx = new (std::nothrow) lbl_ctx;
    if (!ctx) return fail(nullptr, LBL_ENOMEM, "host allocation failed");
    ctx->device = device_ordinal;
    DeviceGuard g(device_ordinal);
    lbl_status st = LBL_OK;
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&ctx->ev0)) != hipSuccess || (e = hipEventCreate(&ctx->ev1)) != hipSuccess ||
        (e = hipMalloc(reinterpret_cast<void **>(&ctx->d_misc), 256)) != hipSuccess ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&ctx->h_misc), 256, hipHostMallocDefault)) != hipSuccess ||
        (e = hipMemsetAsync(ctx->d_misc, 0, 256, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) {
        st = fail(nullptr, LBL_EDEVICE, std::string("context setup failed: ") + hipGetErrorString(e));
        lbl_destroy(ctx);
        return st;
    }
    // the device address of this device's fixed scratch (lbl_label2d_dev_noalloc)
    if ((e = scratch_setup(device_ordinal, &ctx->scratch)) != hipSuccess) {
        ctx->scratch = nullptr;
        ctx->scratch_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        (void)hipGetLastError();
    }
    *out = ctx;
    return LBL_OK;
}

extern "C" void lbl_destroy(lbl_ctx *ctx)
{
    if (!ctx) return;
    {
        DeviceGuard g(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        if (ctx->ws) (void)hipFree(ctx->ws);
        if (ctx->d_in) (void)hipFree(ctx->d_in);
        if (ctx->d_lab) (void)hipFree(ctx->d_lab);
        if (ctx->d_misc) (void)hipFree(ctx->d_misc);
        if (ctx->h_misc) (void)hipHostFree(ctx->h_misc);
        if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
        if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
        for (hipEvent_t e : ctx->timer.events) (void)hipEventDestroy(e);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

extern "C" const char *lbl_last_error(const lbl_ctx *ctx)
{
    return ctx ? ctx->last_error.c_str() : g_create_error.c_str();
}

extern "C" const char *lbl_status_string(lbl_status s)
{
    switch (s) {
    case LBL_OK: return "LBL_OK";
    case LBL_EINVAL: return "LBL_EINVAL";
    case LBL_EUNSUPPORTED: return "LBL_EUNSUPPORTED";
    case LBL_ENOMEM: return "LBL_ENOMEM";
    case LBL_EDEVICE: return "LBL_EDEVICE";
    case LBL_EITER: return "LBL_EITER";
    case LBL_EBUSY: return "LBL_EBUSY";
    }
    return "unknown status";
}

extern "C" lbl_status lbl_set_label_mode(lbl_ctx *ctx, lbl_label_mode mode)
{
    if (!ctx) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (mode != LBL_LABELS_ROOT && mode != LBL_LABELS_REFERENCE && mode != LBL_LABELS_CANONICAL &&
        mode != LBL_LABELS_REFERENCE_UNCROPPED)
        return fail(ctx, LBL_EINVAL, "unknown label mode");
    ctx->opt.mode = mode;
    return LBL_OK;
}

extern "C" lbl_status lbl_set_iteration_cap(lbl_ctx *ctx, uint32_t cap)
{
    if (!ctx) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->opt.iter_cap = cap;
    return LBL_OK;
}

extern "C" lbl_status lbl_alg_from_name(const char *name, lbl_alg *out)
{
    if (!name || !out) return LBL_EINVAL;
    static const struct {
        const char *name;
        lbl_alg alg;
    } list[] = {{"bleq", LBL_BLEQ}, {"lbeq", LBL_LBEQ}, {"bin", LBL_BIN}, {"runeq", LBL_RUNEQ},
                {"he-run", LBL_HERUN}, {"gr-block", LBL_GRBLOCK}, {"ocv", LBL_OCV}};   // main.cpp:103-130
    for (const auto &e : list)
        if (!std::strcmp(name, e.name)) {
            *out = e.alg;
            return LBL_OK;
        }
    return LBL_EINVAL;
}

extern "C" lbl_status lbl_set_profiling(lbl_ctx *ctx, int enable)
{
    if (!ctx) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->profiling = enable != 0;
    return LBL_OK;
}

extern "C" int lbl_last_kernel_times(lbl_ctx *ctx, float *us, const char **names, int cap)
{
    if (!ctx) return -1;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    KernelTimer &kt = ctx->timer;
    if (kt.used < 2) return 0;
    if (hipEventSynchronize(kt.events[kt.used - 1]) != hipSuccess) return -1;
    int n = (int)kt.used - 1;
    for (int i = 0; i < n && i < cap; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, kt.events[i], kt.events[i + 1]) != hipSuccess) ms = -1.f;
        if (us) us[i] = ms * 1000.f;
        if (names) names[i] = i < (int)kt.names.size() ? kt.names[i] : "";
    }
    return n;
}

// ---- device-buffer entry points ----------------------------------------------

// status: device word for the kernels' error bits (host-buffer calls, which
// clear it before and read it after their own launch), or null (device-buffer
// calls: not reported).
// Enqueue the 2D pipeline with workspace ws (at least workspace2d bytes, count
// slots at zero).
static lbl_status launch2d(lbl_ctx *ctx, lbl_alg alg, int conn, const Image2D &im, hipStream_t s, uint32_t *status,
                           void *ws, KernelTimer *kt)
{
    Options opt = ctx->opt;
    opt.status = status;
    // the methods that run the bleq 2x2-block pipeline (lbeq, he-run and ocv
    // with 8-connectivity, gr-block) take bleq's tuning
    const bool blocks2x2 = alg == LBL_GRBLOCK || ((alg == LBL_LBEQ || alg == LBL_HERUN || alg == LBL_OCV) && conn == 8);
    const int talg = blocks2x2 ? (int)LBL_BLEQ : (int)alg;
    auto it = ctx->tuning.find(TuneKey{talg, 0, size_class_of((uint64_t)im.h * im.w)});
    if (it != ctx->tuning.end()) {   // autotuned launch configuration for this size class
        opt.tile2d = it->second.tile;
        opt.flatten = it->second.flatten;
    } else {
        opt.tile2d = plan::default_tile2d((uint64_t)im.h * im.w);
    }
    hipError_t e = hipSuccess;
    switch (alg) {
    case LBL_BLEQ: e = bleq2d_run(im, opt, ws, s, kt); break;
    case LBL_LBEQ: e = lbeq2d_run(im, conn, opt, ws, s, kt); break;
    case LBL_GRBLOCK: e = grblock2d_run(im, opt, ws, s, kt); break;
    case LBL_RUNEQ: e = runs2d_run(im, opt, ws, s, kt, false, conn); break;
    case LBL_HERUN: e = runs2d_run(im, opt, ws, s, kt, true, conn); break;
    case LBL_OCV: {   // 1..N in raster order of first pixels, whatever the mode
        opt.mode = LBL_LABELS_CANONICAL;
        e = conn == 8 ? bleq2d_run(im, opt, ws, s, kt) : bleq2d_run_conn4(im, opt, ws, s, kt);
        break;
    }
    case LBL_BIN: e = bin2d_run(im, ws, s, kt); break;
    }
    if (e != hipSuccess) return hip_fail(ctx, "2D labeling launch", e);
    return LBL_OK;
}

static lbl_status run2d(lbl_ctx *ctx, lbl_alg alg, int conn, const Image2D &im, hipStream_t s, uint32_t *status)
{
    bool fresh = false;
    lbl_status st = ensure(ctx, &ctx->ws, &ctx->ws_bytes, workspace2d(alg, conn, im.h, im.w, ctx->opt.mode), &fresh);
    if (st != LBL_OK) return st;
    // every workspace starts with the count slots, which runs leave at zero
    if (fresh) LBL_HIP(ctx, hipMemsetAsync(ctx->ws, 0, kCountBytes, s));
    return launch2d(ctx, alg, conn, im, s, status, ctx->ws, ctx->profiling ? &ctx->timer : nullptr);
}

// lbl_label2d_dev's checks, then the pipeline in a caller-provided workspace:
// nothing is allocated.  `prologue` (enqueued first, on the same stream)
// prepares the workspace.
lbl_status lbl::label2d_in_workspace(lbl_ctx *ctx, lbl_alg alg, const uint8_t *d_px, uint32_t h, uint32_t w,
                                     size_t pitch_bytes, int conn, uint32_t *d_labels, uint32_t *d_ncomp, void *stream,
                                     void *ws, size_t ws_bytes, hipError_t (*prologue)(void *, hipStream_t))
{
    if (!ctx) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!d_px || !d_labels) return fail(ctx, LBL_EINVAL, "null buffer");
    if (!ws) return fail(ctx, LBL_EDEVICE, "no workspace: " + ctx->scratch_error);
    if (alg != LBL_BLEQ && alg != LBL_LBEQ)
        return fail(ctx, LBL_EUNSUPPORTED, "the allocation-free entry point runs bleq and lbeq only");
    int c = 0;
    lbl_status st;
    if ((st = normalize_conn2d(ctx, alg, conn, &c)) != LBL_OK) return st;
    if ((st = check_size2d(ctx, alg, c, h, w, pitch_bytes)) != LBL_OK) return st;
    if ((st = check_mode(ctx, alg, false)) != LBL_OK) return st;
    const size_t need = workspace2d(alg, c, h, w, ctx->opt.mode);
    if (need > ws_bytes)
        return fail(ctx, LBL_ENOMEM, "image needs " + std::to_string(need) + " bytes of workspace, more than the " +
                                         std::to_string(ws_bytes) + " available");
    DeviceGuard g(ctx->device);
    hipStream_t s = pick_stream(ctx, stream);
    if (prologue) LBL_HIP(ctx, prologue(ws, s));
    Image2D im{d_px, pitch_bytes, h, w, d_labels, d_ncomp};
    return launch2d(ctx, alg, c, im, s, nullptr, ws, nullptr);
}

static lbl_status run3d(lbl_ctx *ctx, lbl_alg alg, const Image3D &im, hipStream_t s, uint32_t *status)
{
    bool fresh = false;
    lbl_status st = ensure(ctx, &ctx->ws, &ctx->ws_bytes, workspace3d(alg, im.d, im.h, im.w, ctx->opt.mode), &fresh);
    if (st != LBL_OK) return st;
    // every workspace starts with the count slots, which runs leave at zero
    if (fresh) LBL_HIP(ctx, hipMemsetAsync(ctx->ws, 0, kCountBytes, s));
    KernelTimer *kt = ctx->profiling ? &ctx->timer : nullptr;
    Options opt = ctx->opt;
    opt.status = status;
    hipError_t e = alg == LBL_BLEQ  ? bleq3d_run(im, opt, ctx->ws, s, kt)
                   : alg == LBL_BIN ? bin3d_run(im, opt, ctx->ws, s, kt)
                                    : lbeq3d_run(im, opt, ctx->ws, s, kt);
    if (e != hipSuccess) return hip_fail(ctx, "3D labeling launch", e);
    return LBL_OK;
}

extern "C" lbl_status lbl_label2d_dev(lbl_ctx *ctx, lbl_alg alg, const uint8_t *d_px, uint32_t h, uint32_t w,
                                      size_t pitch_bytes, int conn, uint32_t *d_labels, uint32_t *d_ncomp,
                                      void *stream)
{
    if (!ctx) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!d_px || !d_labels) return fail(ctx, LBL_EINVAL, "lbl_label2d_dev: null buffer");
    int c = 0;
    lbl_status st;
    if ((st = normalize_conn2d(ctx, alg, conn, &c)) != LBL_OK) return st;
    if ((st = check_size2d(ctx, alg, c, h, w, pitch_bytes)) != LBL_OK) return st;
    if ((st = check_mode(ctx, alg, false)) != LBL_OK) return st;
    DeviceGuard g(ctx->device);
    Image2D im{d_px, pitch_bytes, h, w, d_labels, d_ncomp};
    return run2d(ctx, alg, c, im, pick_stream(ctx, stream), nullptr);
}

extern "C" lbl_status lbl_label3d_dev(lbl_ctx *ctx, lbl_alg alg, const uint8_t *d_vox, uint32_t d, uint32_t h,
                                      uint32_t w, int conn, uint32_t *d_labels, uint32_t *d_ncomp, void *stream)
{
    if (!ctx) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!d_vox || !d_labels) return fail(ctx, LBL_EINVAL, "lbl_label3d_dev: null buffer");
    lbl_status st;
    if ((st = normalize_conn3d(ctx, alg, conn)) != LBL_OK) return st;
    if ((st = check_size3d(ctx, d, h, w)) != LBL_OK) return st;
    if ((st = check_mode(ctx, alg, true)) != LBL_OK) return st;
    DeviceGuard g(ctx->device);
    Image3D im{d_vox, d, h, w, d_labels, d_ncomp};
    return run3d(ctx, alg, im, pick_stream(ctx, stream), nullptr);
}

// ---- host-buffer entry points --------------------------------------------------

static lbl_status finish_host(lbl_ctx *ctx, uint32_t *labels, size_t n, uint32_t *n_components, float *elapsed_us)
{
    hipStream_t s = ctx->stream;
    LBL_HIP(ctx, hipEventRecord(ctx->ev1, s));
    LBL_HIP(ctx, hipMemcpyAsync(labels, ctx->d_lab, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    LBL_HIP(ctx, hipMemcpyAsync(ctx->h_misc, ctx->d_misc, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    LBL_HIP(ctx, hipStreamSynchronize(s));
    if (elapsed_us) {
        float ms = 0.f;
        LBL_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *elapsed_us = ms * 1000.f;
    }
    if (n_components) *n_components = ctx->h_misc[0];
    if (ctx->h_misc[1] & 1u) return fail(ctx, LBL_EITER, "iteration cap reached");
    return LBL_OK;
}

extern "C" lbl_status lbl_label2d(lbl_ctx *ctx, lbl_alg alg, const uint8_t *px, uint32_t h, uint32_t w,
                                  size_t pitch_bytes, int conn, uint32_t *labels, uint32_t *n_components,
                                  float *elapsed_us)
{
    if (!ctx) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!px || !labels) return fail(ctx, LBL_EINVAL, "lbl_label2d: null buffer");
    int c = 0;
    lbl_status st;
    if ((st = normalize_conn2d(ctx, alg, conn, &c)) != LBL_OK) return st;
    if ((st = check_size2d(ctx, alg, c, h, w, pitch_bytes)) != LBL_OK) return st;
    if ((st = check_mode(ctx, alg, false)) != LBL_OK) return st;
    DeviceGuard g(ctx->device);
    const size_t n = (size_t)h * w;
    const size_t dpitch = (w + 15u) & ~size_t(15);   // 16-byte rows for vector loads
    if ((st = ensure(ctx, reinterpret_cast<void **>(&ctx->d_in), &ctx->in_bytes, dpitch * h, nullptr)) != LBL_OK) return st;
    if ((st = ensure(ctx, reinterpret_cast<void **>(&ctx->d_lab), &ctx->lab_bytes, n * sizeof(uint32_t), nullptr)) != LBL_OK)
        return st;
    hipStream_t s = ctx->stream;
    // the label buffer (4 B/px) stages the rows until the kernels write it
    LBL_HIP(ctx, upload2d(ctx->d_in, dpitch, px, pitch_bytes, w, h, reinterpret_cast<uint8_t *>(ctx->d_lab),
                          n * sizeof(uint32_t), s));
    LBL_HIP(ctx, hipMemsetAsync(ctx->d_misc + 1, 0, sizeof(uint32_t), s));   // this call's status bits
    LBL_HIP(ctx, hipEventRecord(ctx->ev0, s));
    Image2D im{ctx->d_in, dpitch, h, w, ctx->d_lab, ctx->d_misc};
    if ((st = run2d(ctx, alg, c, im, s, ctx->d_misc + 1)) != LBL_OK) return st;
    return finish_host(ctx, labels, n, n_components, elapsed_us);
}

extern "C" lbl_status lbl_label3d(lbl_ctx *ctx, lbl_alg alg, const uint8_t *vox, uint32_t d, uint32_t h,
                                  uint32_t w, int conn, uint32_t *labels, uint32_t *n_components, float *elapsed_us)
{
    if (!ctx) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!vox || !labels) return fail(ctx, LBL_EINVAL, "lbl_label3d: null buffer");
    lbl_status st;
    if ((st = normalize_conn3d(ctx, alg, conn)) != LBL_OK) return st;
    if ((st = check_size3d(ctx, d, h, w)) != LBL_OK) return st;
    if ((st = check_mode(ctx, alg, true)) != LBL_OK) return st;
    DeviceGuard g(ctx->device);
    const size_t n = (size_t)d * h * w, nl = labels3d_elems(ctx->opt.mode, d, h, w);
    if ((st = ensure(ctx, reinterpret_cast<void **>(&ctx->d_in), &ctx->in_bytes, n, nullptr)) != LBL_OK) return st;
    if ((st = ensure(ctx, reinterpret_cast<void **>(&ctx->d_lab), &ctx->lab_bytes, nl * sizeof(uint32_t), nullptr)) != LBL_OK)
        return st;
    hipStream_t s = ctx->stream;
    LBL_HIP(ctx, hipMemcpyAsync(ctx->d_in, vox, n, hipMemcpyHostToDevice, s));
    LBL_HIP(ctx, hipMemsetAsync(ctx->d_misc + 1, 0, sizeof(uint32_t), s));   // this call's status bits
    LBL_HIP(ctx, hipEventRecord(ctx->ev0, s));
    Image3D im{ctx->d_in, d, h, w, ctx->d_lab, ctx->d_misc};
    if ((st = run3d(ctx, alg, im, s, ctx->d_misc + 1)) != LBL_OK) return st;
    return finish_host(ctx, labels, nl, n_components, elapsed_us);
}

extern "C" lbl_status lbl_label2d_each(lbl_ctx *ctx, lbl_alg alg, const lbl_image2d *images, uint32_t n, int conn,
                                       uint32_t *d_ncomp, void *stream)
{
    if (!ctx) return LBL_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!images || n == 0) return fail(ctx, LBL_EINVAL, "lbl_label2d_each: empty image list");
    int c = 0;
    lbl_status st;
    if ((st = normalize_conn2d(ctx, alg, conn, &c)) != LBL_OK) return st;
    if ((st = check_mode(ctx, alg, false)) != LBL_OK) return st;
    size_t need = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const lbl_image2d &im = images[i];
        if (!im.px || !im.labels) return fail(ctx, LBL_EINVAL, "lbl_label2d_each: null buffer");
        if ((st = check_size2d(ctx, alg, c, im.h, im.w, im.pitch_bytes)) != LBL_OK) return st;
        need = std::max(need, workspace2d(alg, c, im.h, im.w, ctx->opt.mode));
    }
    DeviceGuard g(ctx->device);
    hipStream_t s = pick_stream(ctx, stream);
    bool fresh = false;
    if ((st = ensure(ctx, &ctx->ws, &ctx->ws_bytes, need, &fresh)) != LBL_OK) return st;
    if (fresh) LBL_HIP(ctx, hipMemsetAsync(ctx->ws, 0, kCountBytes, s));
    for (uint32_t i = 0; i < n; ++i) {
        const lbl_image2d &im = images[i];
        Image2D one{im.px, im.pitch_bytes, im.h, im.w, im.labels, d_ncomp ? d_ncomp + i : nullptr};
        if ((st = run2d(ctx, alg, c, one, s, nullptr)) != LBL_OK) return st;   // the workspace is reused in stream order
    }
    return LBL_OK;
}
