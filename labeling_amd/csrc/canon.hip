// canon.hip -- the scan behind canonical label numbering: set bits per chunk of
// the key bit array, then per word the set bits before it.
#include "canon.hpp"

namespace lbl {
namespace {

// Inclusive sum over the 256 threads of a workgroup; `total` gets the sum of all.
__device__ __forceinline__ uint32_t block_inclusive_sum(uint32_t v, uint32_t &total)
{
    __shared__ uint32_t wave_sums[kCanonThreads / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    if (lane == 63) wave_sums[wave] = v;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (uint32_t i = 0; i < kCanonThreads / 64; ++i) {
        const uint32_t s = wave_sums[i];
        before += i < wave ? s : 0u;
        total += s;
    }
    __syncthreads();   // wave_sums reusable
    return v + before;
}

// Set bits of this thread's kCanonWordsPerThread words of chunk blockIdx.x.
__device__ __forceinline__ uint32_t thread_bits(const Canon &c, uint32_t counts[kCanonWordsPerThread])
{
    const uint64_t w0 = (uint64_t)blockIdx.x * kCanonChunkWords + (uint64_t)threadIdx.x * kCanonWordsPerThread;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < kCanonWordsPerThread; ++i) {
        counts[i] = w0 + i < c.nwords ? (uint32_t)__popcll(c.bits[w0 + i]) : 0u;
        sum += counts[i];
    }
    return sum;
}

__global__ __launch_bounds__(kCanonThreads) void canon_partial(Canon c)
{
    uint32_t counts[kCanonWordsPerThread], total;
    block_inclusive_sum(thread_bits(c, counts), total);
    if (threadIdx.x == 0) c.partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(kCanonThreads) void canon_prefix(Canon c)
{
    // set bits in the chunks before this one
    uint32_t base = 0, total;
    for (uint32_t i = threadIdx.x; i < blockIdx.x; i += kCanonThreads) base += c.partial[i];
    block_inclusive_sum(base, total);
    base = total;
    uint32_t counts[kCanonWordsPerThread];
    const uint32_t mine = thread_bits(c, counts);
    uint32_t run = base + block_inclusive_sum(mine, total) - mine;
    const uint64_t w0 = (uint64_t)blockIdx.x * kCanonChunkWords + (uint64_t)threadIdx.x * kCanonWordsPerThread;
#pragma unroll
    for (uint32_t i = 0; i < kCanonWordsPerThread; ++i) {
        if (w0 + i < c.nwords) c.prefix[w0 + i] = run;
        run += counts[i];
    }
}

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
    const uint64_t nw = words_of(nkeys);
    return align256(nnodes * sizeof(uint32_t)) + align256(nw * sizeof(unsigned long long)) +
           align256(nw * sizeof(uint32_t)) + align256(chunks_of(nw) * sizeof(uint32_t));
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
