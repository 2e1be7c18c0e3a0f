// upload.hip -- host image rows to pitched device rows.
//
// hipMemcpy2DAsync from host memory with a source pitch unlike the destination
// pitch runs at well under 1 GB/s on this platform (row by row), while a
// contiguous copy runs at PCIe speed.  Images whose row pitch needs changing
// are therefore copied contiguously into a staging buffer on the device and
// re-pitched there by a small kernel.
#include "lbl_internal.hpp"

namespace lbl {
namespace {

constexpr uint32_t kRepitchThreads = 256;
constexpr uint32_t kChunk = 16;   // destination bytes per thread

__global__ __launch_bounds__(kRepitchThreads) void repitch(uint8_t *__restrict__ dst, uint64_t dpitch,
                                                          const uint8_t *__restrict__ src, uint64_t spitch,
                                                          uint32_t w, uint32_t h, uint32_t chunks_per_row)
{
    const uint64_t i = (uint64_t)blockIdx.x * kRepitchThreads + threadIdx.x;
    const uint32_t y = (uint32_t)(i / chunks_per_row), x0 = (uint32_t)(i % chunks_per_row) * kChunk;
    if (y >= h) return;
    const uint8_t *s = src + y * spitch + x0;
    uint8_t *d = dst + y * dpitch + x0;
    if (x0 + kChunk <= w) {
        uint32_t q[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k)
            q[k] = (uint32_t)s[4 * k] | ((uint32_t)s[4 * k + 1] << 8) | ((uint32_t)s[4 * k + 2] << 16) |
                   ((uint32_t)s[4 * k + 3] << 24);
        if (((reinterpret_cast<uintptr_t>(d)) & 15u) == 0) {
            *reinterpret_cast<uint4 *>(d) = make_uint4(q[0], q[1], q[2], q[3]);
            return;
        }
        for (uint32_t k = 0; k < kChunk; ++k) d[k] = (uint8_t)(q[k >> 2] >> (8 * (k & 3)));
        return;
    }
    for (uint32_t k = 0; x0 + k < w; ++k) d[k] = s[k];
}

}  // namespace

hipError_t upload2d(uint8_t *dst, size_t dpitch, const uint8_t *src, size_t spitch, uint32_t w, uint32_t h,
                    uint8_t *staging, size_t staging_bytes, hipStream_t s)
{
    if (h == 0 || w == 0) return hipSuccess;
    if (spitch == dpitch) return hipMemcpyAsync(dst, src, dpitch * (h - 1) + w, hipMemcpyHostToDevice, s);
    const size_t span = spitch * (h - 1) + w;
    if (!staging || span > staging_bytes)
        return hipMemcpy2DAsync(dst, dpitch, src, spitch, w, h, hipMemcpyHostToDevice, s);
    hipError_t e = hipMemcpyAsync(staging, src, span, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    const uint32_t chunks = (w + kChunk - 1) / kChunk;
    const uint64_t threads = (uint64_t)chunks * h;
    hipLaunchKernelGGL(repitch, dim3((uint32_t)((threads + kRepitchThreads - 1) / kRepitchThreads)),
                       dim3(kRepitchThreads), 0, s, dst, (uint64_t)dpitch, staging, (uint64_t)spitch, w, h, chunks);
    return hipGetLastError();
}

}  // namespace lbl
