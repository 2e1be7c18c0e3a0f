// noalloc.hip -- lbl_label2d_dev_noalloc: 2D labeling of device buffers that
// allocates nothing, for callers that cannot call hipMalloc (code under HIP
// graph capture, tight serving loops, harnesses that forbid allocation).
//
// The union-find workspace (the count slots, then the block words) lives in
// `scratch`, a file-scope device array of kScratchBytes, one copy per device,
// held until the process ends.  Host code never passes the array's name to a
// kernel or a copy: the device address of the current device's copy comes
// from hipGetSymbolAddress, resolved once per device when a context is
// created (scratch_setup), and only that address is used.
//
// Every call on a device shares its scratch, and the scratch keeps the last
// call's contents: each call first clears the count slots with a small kernel
// on its own stream (no hipMemsetAsync, which must not be recorded into a
// captured graph).  Calls that overlap on one device race on the scratch:
// callers serialise them (one stream, or events).
#include <mutex>

#include "lbl_internal.hpp"
#include "uf_device.hpp"

namespace lbl {

namespace {

// 72 MiB: room for the count slots and the block words of a 2D image of up
// to 18,873,344 2x2 blocks (8688 x 8688 pixels) with root or reference values.
constexpr size_t kScratchSize = size_t(72) << 20;

static __device__ __align__(256) uint8_t scratch[kScratchSize];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

__global__ __launch_bounds__(kCountSlots) void scratch_clear_counts(uint32_t *slots)
{
    count_clear(slots);
}

hipError_t clear_counts(void *ws, hipStream_t s)
{
    hipLaunchKernelGGL(scratch_clear_counts, dim3(1), dim3(kCountSlots), 0, s, static_cast<uint32_t *>(ws));
    return hipGetLastError();
}

}  // namespace

const size_t kScratchBytes = kScratchSize;

// Called with `device` current.
hipError_t scratch_setup(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        void *p = nullptr;
        g_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[device] = g_err[device] == hipSuccess ? p : nullptr;
        g_done[device] = true;
    }
    *ptr = g_ptr[device];
    return g_err[device];
}

}  // namespace lbl

extern "C" lbl_status lbl_label2d_dev_noalloc(lbl_ctx *ctx, lbl_alg alg, const uint8_t *d_px, uint32_t h, uint32_t w,
                                              size_t pitch_bytes, int conn, uint32_t *d_labels, uint32_t *d_ncomp,
                                              void *stream)
{
    if (!ctx) return LBL_EINVAL;
    return lbl::label2d_in_workspace(ctx, alg, d_px, h, w, pitch_bytes, conn, d_labels, d_ncomp, stream, ctx->scratch,
                                     lbl::kScratchBytes, lbl::clear_counts);
}
