// canon.hip -- the scan behind canonical label numbering: set bits per chunk of
// the key bit array, then per word the set bits before it.
#include "batch_plan.hpp"
#include "canon.hpp"

namespace lbl {
namespace {

__global__ __launch_bounds__(kCanonThreads) void canon_partial(Canon c) { canon_partial_chunk(c, blockIdx.x); }

__global__ __launch_bounds__(kCanonThreads) void canon_prefix(Canon c) { canon_prefix_chunk(c, blockIdx.x); }

__global__ __launch_bounds__(kCanonThreads) void canon_clear(Canon c)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kCanonThreads + threadIdx.x; i < c.nwords; i += (uint64_t)gridDim.x * kCanonThreads)
        c.bits[i] = 0ull;
}

uint64_t words_of(uint64_t nkeys) { return (nkeys + 63) / 64; }
uint64_t chunks_of(uint64_t nwords) { return (nwords + kCanonChunkWords - 1) / kCanonChunkWords; }

}  // namespace

size_t canon_workspace(uint64_t nkeys, uint64_t nnodes)
{
    return plan::canon_bytes(nkeys, nnodes);   // the layout below, as the batch planner computes it
}

Canon canon_layout(void *base, uint64_t nkeys, uint64_t nnodes)
{
    const uint64_t nw = words_of(nkeys);
    uint8_t *p = static_cast<uint8_t *>(base);
    Canon c{};
    c.key = nnodes ? reinterpret_cast<uint32_t *>(p) : nullptr;
    p += align256(nnodes * sizeof(uint32_t));
    c.bits = reinterpret_cast<unsigned long long *>(p);
    p += align256(nw * sizeof(unsigned long long));
    c.prefix = reinterpret_cast<uint32_t *>(p);
    p += align256(nw * sizeof(uint32_t));
    c.partial = reinterpret_cast<uint32_t *>(p);
    c.nwords = (uint32_t)nw;
    return c;
}

hipError_t canon_clear_launch(const Canon &c, hipStream_t s, KernelTimer *kt)
{
    const uint64_t blocks = (c.nwords + kCanonThreads - 1) / kCanonThreads;
    hipLaunchKernelGGL(canon_clear, dim3((uint32_t)(blocks < 4096 ? blocks : 4096)), dim3(kCanonThreads), 0, s, c);
    if (kt) kt->mark(s, "canon_clear");
    return hipGetLastError();
}

hipError_t canon_scan_launch(const Canon &c, hipStream_t s, KernelTimer *kt)
{
    const dim3 grid((uint32_t)chunks_of(c.nwords));
    hipLaunchKernelGGL(canon_partial, grid, dim3(kCanonThreads), 0, s, c);
    if (kt) kt->mark(s, "canon_partial");
    hipLaunchKernelGGL(canon_prefix, grid, dim3(kCanonThreads), 0, s, c);
    if (kt) kt->mark(s, "canon_prefix");
    return hipGetLastError();
}

}  // namespace lbl
