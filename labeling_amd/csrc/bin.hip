// bin.hip -- the reference's "bin" method (TBinLabeling / TOCLBinLabeling /
// TOCLBinLabeling3D, LabelingAlgs.cpp:23-36,815-862; BinLabelingKernel and
// Bin3D_LabelingKernel, LabelingAlgs.cl:18-25,622-628): every foreground pixel
// (voxel) gets label 1, background 0, whatever the label mode.  Component
// count: 1 when there is a foreground pixel, else 0.
#include "lbl_internal.hpp"
#include "uf_device.hpp"

namespace lbl {
namespace {

// Foreground flag of the whole launch: a workgroup that holds a foreground
// pixel stores 1 into its count slot (a plain store of the same value by every
// such workgroup: no atomics, which on one address serialize across the whole
// grid).  count_any_kernel turns the slots into the count and clears them.
__device__ __forceinline__ void flag_any(uint32_t *slots, bool any)
{
    if (__syncthreads_or(any) && threadIdx.x == 0) slots[(blockIdx.x % kCountSlots) * kCountStride] = 1u;
}

__global__ __launch_bounds__(kCountSlots) void count_any_kernel(uint32_t *slots, uint32_t *count)
{
    const uint32_t v = slots[threadIdx.x * kCountStride];
    slots[threadIdx.x * kCountStride] = 0;
    const bool any = __ballot(v != 0) != 0;
    if (threadIdx.x == 0) *count = any ? 1u : 0u;
}

// 4 labels (0/1) from the 4 bytes at p .. p + 3 of which the first n < 4 or
// all exist; one 32-bit load when whole and 4-byte aligned.
__device__ __forceinline__ void bin4(const uint8_t *p, uint32_t n, uint32_t (&v)[4])
{
    if (n >= 4 && (reinterpret_cast<uintptr_t>(p) & 3u) == 0) {
        const uint32_t nz = nz_bytes(*reinterpret_cast<const uint32_t *>(p));
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) v[k] = (nz >> (8 * k)) & 1u;
        return;
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) v[k] = k < n && p[k] != 0 ? 1u : 0u;
}

// 4 labels at out .. out + 3 of which the first n < 4 or all exist.
__device__ __forceinline__ void store4(uint32_t *out, uint32_t n, const uint32_t (&v)[4])
{
    if (n >= 4 && (reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
        store_nt4(out, v[0], v[1], v[2], v[3]);
        return;
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
        if (k < n) out[k] = v[k];
}

constexpr uint32_t kBinThreads = 256, kBinSpan = 4 * kBinThreads;   // pixels of a row per workgroup

// Workgroup g: row g / chunks, pixels (g % chunks) * 1024 + 4 * thread .. + 3.
__global__ __launch_bounds__(kBinThreads) void bin2d_kernel(const uint8_t *__restrict__ px, size_t pitch, uint32_t w,
                                                           uint32_t chunks, uint32_t *__restrict__ labels,
                                                           uint32_t *slots)
{
    const uint32_t y = blockIdx.x / chunks, x0 = (blockIdx.x % chunks) * kBinSpan + 4 * threadIdx.x;
    bool any = false;
    if (x0 < w) {
        uint32_t v[4];
        bin4(px + (size_t)y * pitch + x0, w - x0, v);
        store4(labels + (size_t)y * w + x0, w - x0, v);
        any = (v[0] | v[1] | v[2] | v[3]) != 0;
    }
    if (slots) flag_any(slots, any);
}

// The same over output rows of a volume; the uncropped output is the
// reference's padded volume pd x ph x pw with the input at offset +2 in every
// dimension (off = 2), rows r = z * ph + y.
__global__ __launch_bounds__(kBinThreads) void bin3d_kernel(const uint8_t *__restrict__ vox, uint32_t d, uint32_t h,
                                                           uint32_t w, uint32_t ph, uint32_t pw, uint32_t off,
                                                           uint32_t chunks, uint32_t *__restrict__ labels,
                                                           uint32_t *slots)
{
    const uint32_t r = blockIdx.x / chunks, x0 = (blockIdx.x % chunks) * kBinSpan + 4 * threadIdx.x;
    const uint32_t sy = r % ph - off, sz = r / ph - off;   // unsigned wrap -> outside
    bool any = false;
    if (x0 < pw) {
        uint32_t v[4] = {0, 0, 0, 0};
        if (sy < h && sz < d) {
            const uint8_t *row = vox + ((size_t)sz * h + sy) * w;
            if (off == 0) {
                bin4(row + x0, w - x0, v);
            } else {
#pragma unroll
                for (uint32_t k = 0; k < 4; ++k) {
                    const uint32_t sx = x0 + k - off;
                    v[k] = sx < w && row[sx] != 0 ? 1u : 0u;
                }
            }
        }
        store4(labels + (size_t)r * pw + x0, pw - x0, v);
        any = (v[0] | v[1] | v[2] | v[3]) != 0;
    }
    if (slots) flag_any(slots, any);
}

}  // namespace

hipError_t bin2d_run(const Image2D &im, void *ws, hipStream_t s, KernelTimer *kt)
{
    uint32_t *slots = im.count ? static_cast<uint32_t *>(ws) : nullptr;   // count slots start the workspace
    const uint32_t chunks = (im.w + kBinSpan - 1) / kBinSpan;
    const uint64_t groups = (uint64_t)im.h * chunks;
    if (groups > 0x7FFFFFFFu) return hipErrorInvalidValue;
    if (kt) kt->begin(s);
    hipLaunchKernelGGL(bin2d_kernel, dim3((uint32_t)groups), dim3(kBinThreads), 0, s, im.px, im.pitch, im.w, chunks,
                       im.labels, slots);
    if (kt) kt->mark(s, "bin2d");
    if (slots) hipLaunchKernelGGL(count_any_kernel, dim3(1), dim3(kCountSlots), 0, s, slots, im.count);
    return hipGetLastError();
}

hipError_t bin3d_run(const Image3D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt)
{
    uint32_t *slots = im.count ? static_cast<uint32_t *>(ws) : nullptr;
    const bool padded = opt.mode == LBL_LABELS_REFERENCE_UNCROPPED;
    const uint32_t pd = padded ? (uint32_t)ref_pad3(im.d) : im.d, ph = padded ? (uint32_t)ref_pad3(im.h) : im.h;
    const uint32_t pw = padded ? (uint32_t)ref_pad3(im.w) : im.w, off = padded ? 2u : 0u;
    const uint32_t chunks = (pw + kBinSpan - 1) / kBinSpan;
    const uint64_t groups = (uint64_t)pd * ph * chunks;
    if (groups > 0x7FFFFFFFu) return hipErrorInvalidValue;
    if (kt) kt->begin(s);
    hipLaunchKernelGGL(bin3d_kernel, dim3((uint32_t)groups), dim3(kBinThreads), 0, s, im.vox, im.d, im.h, im.w, ph, pw,
                       off, chunks, im.labels, slots);
    if (kt) kt->mark(s, "bin3d");
    if (slots) hipLaunchKernelGGL(count_any_kernel, dim3(1), dim3(kCountSlots), 0, s, slots, im.count);
    return hipGetLastError();
}

}  // namespace lbl
