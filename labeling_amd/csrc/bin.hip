// bin.hip -- the reference's "bin" method (TBinLabeling / TOCLBinLabeling /
// TOCLBinLabeling3D, LabelingAlgs.cpp:23-36,815-862; BinLabelingKernel and
// Bin3D_LabelingKernel, LabelingAlgs.cl:18-25,622-628): every foreground pixel
// (voxel) gets label 1, background 0, whatever the label mode.  Component
// count: 1 when there is a foreground pixel, else 0 (one atomicOr per wave
// that holds one into a count slot, summed and cleared by count_reduce).
#include "lbl_internal.hpp"
#include "uf_device.hpp"

namespace lbl {
namespace {

// One thread per 4 pixels of a row: label 1 where the byte is non-zero.
__global__ __launch_bounds__(256) void bin2d_kernel(const uint8_t *__restrict__ px, size_t pitch, uint32_t h,
                                                    uint32_t w, uint32_t *__restrict__ labels, uint32_t *slots)
{
    const uint32_t quads = (w + 3) / 4;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = (uint32_t)(t / quads), x0 = (uint32_t)(t % quads) * 4;
    bool any = false;
    if (y < h) {
        const uint8_t *row = px + (size_t)y * pitch;
        uint32_t *out = labels + (size_t)y * w;
        uint32_t v[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            v[k] = x0 + k < w && row[x0 + k] != 0 ? 1u : 0u;
            any = any || v[k];
        }
        if (x0 + 4 <= w && (reinterpret_cast<uintptr_t>(out + x0) & 15u) == 0) {
            store_nt4(out + x0, v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k)
                if (x0 + k < w) out[x0 + k] = v[k];
        }
    }
    if (slots && __ballot(any) && (threadIdx.x & 63u) == 0) atomicOr(slots, 1u);
}

// One thread per output voxel; the uncropped output is the reference's padded
// volume pd x ph x pw with the input at offset +2 in every dimension.
__global__ __launch_bounds__(256) void bin3d_kernel(const uint8_t *__restrict__ vox, uint32_t d, uint32_t h,
                                                    uint32_t w, uint32_t pd, uint32_t ph, uint32_t pw, uint32_t off,
                                                    uint32_t *__restrict__ labels, uint32_t *slots)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool any = false;
    if (t < (uint64_t)pd * ph * pw) {
        const uint32_t x = (uint32_t)(t % pw), y = (uint32_t)((t / pw) % ph), z = (uint32_t)(t / ((uint64_t)pw * ph));
        const uint32_t sx = x - off, sy = y - off, sz = z - off;   // unsigned wrap -> outside
        uint32_t v = 0;
        if (sx < w && sy < h && sz < d) v = vox[((size_t)sz * h + sy) * w + sx] != 0 ? 1u : 0u;
        labels[t] = v;
        any = v != 0;
    }
    if (slots && __ballot(any) && (threadIdx.x & 63u) == 0) atomicOr(slots, 1u);
}

}  // namespace

hipError_t bin2d_run(const Image2D &im, void *ws, hipStream_t s, KernelTimer *kt)
{
    uint32_t *slots = im.count ? static_cast<uint32_t *>(ws) : nullptr;   // count slots start the workspace
    const uint64_t threads = (uint64_t)im.h * ((im.w + 3) / 4);
    if (kt) kt->begin(s);
    hipLaunchKernelGGL(bin2d_kernel, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s, im.px, im.pitch, im.h,
                       im.w, im.labels, slots);
    if (kt) kt->mark(s, "bin2d");
    if (slots) hipLaunchKernelGGL(count_reduce_kernel<0>, dim3(1), dim3(kCountSlots), 0, s, slots, im.count);
    return hipGetLastError();
}

hipError_t bin3d_run(const Image3D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt)
{
    uint32_t *slots = im.count ? static_cast<uint32_t *>(ws) : nullptr;
    const bool padded = opt.mode == LBL_LABELS_REFERENCE_UNCROPPED;
    const uint32_t pd = padded ? (uint32_t)ref_pad3(im.d) : im.d, ph = padded ? (uint32_t)ref_pad3(im.h) : im.h;
    const uint32_t pw = padded ? (uint32_t)ref_pad3(im.w) : im.w, off = padded ? 2u : 0u;
    const uint64_t n = (uint64_t)pd * ph * pw;
    if (kt) kt->begin(s);
    hipLaunchKernelGGL(bin3d_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, im.vox, im.d, im.h, im.w, pd,
                       ph, pw, off, im.labels, slots);
    if (kt) kt->mark(s, "bin3d");
    if (slots) hipLaunchKernelGGL(count_reduce_kernel<0>, dim3(1), dim3(kCountSlots), 0, s, slots, im.count);
    return hipGetLastError();
}

}  // namespace lbl
