// canon.hpp -- canonical label numbering shared by every path: components are
// numbered 1, 2, 3, ... in the raster order of their first pixel (voxel), x
// fastest; background stays 0.
//
// Each component gets a key, an integer that orders components as their first
// pixels do: lbeq uses the root itself (the minimum pixel index), bleq the
// first pixel's row (2D; z * h + y in 3D) times the block-grid width plus its
// block column, unique per component because the pixels of one block always
// belong to one component.  One bit per possible key is set at the keys of the
// components; a scan gives, per 64-bit word, the set bits before it, and a
// component's label is 1 + the number of set bits below its key.
#pragma once

#include "lbl_internal.hpp"

namespace lbl {

struct Canon {
    uint32_t *key;              // per union-find node (bleq blocks): key of the component rooted there
    unsigned long long *bits;   // one bit per possible key
    uint32_t *prefix;           // per word of bits: set bits in the words before it
    uint32_t *partial;          // per scan chunk: set bits in the chunk
    uint32_t nwords;            // words of bits
};

constexpr uint32_t kCanonThreads = 256;
constexpr uint32_t kCanonWordsPerThread = 8;
constexpr uint32_t kCanonChunkWords = kCanonThreads * kCanonWordsPerThread;   // words per scan workgroup

// Workspace bytes for nkeys possible keys and nnodes key slots (0 = no key
// array), and the layout of a Canon in such a workspace.
size_t canon_workspace(uint64_t nkeys, uint64_t nnodes);
Canon canon_layout(void *base, uint64_t nkeys, uint64_t nnodes);

// Enqueue a fill of c.bits with zeros.
hipError_t canon_clear_launch(const Canon &c, hipStream_t s, KernelTimer *kt);

// Enqueue the scan over c.bits (after every bit is set): partial counts per
// chunk, then prefix per word.
hipError_t canon_scan_launch(const Canon &c, hipStream_t s, KernelTimer *kt);

// Canonical label of the component with key k (its bit is set; a key outside
// the bit array, which no component has, gives 0).
__device__ __forceinline__ uint32_t canon_label(const Canon &c, uint32_t k)
{
    const uint32_t w = k >> 6;
    if (w >= c.nwords) return 0u;
    const unsigned long long below = c.bits[w] & ((1ull << (k & 63u)) - 1ull);
    return 1u + c.prefix[w] + (uint32_t)__popcll(below);
}

// ---- the scan, per chunk of kCanonChunkWords words (one 256-thread workgroup) ----

// Inclusive sum over the 256 threads of a workgroup; `total` gets the sum of all.
__device__ __forceinline__ uint32_t canon_block_sum(uint32_t v, uint32_t &total)
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

// Set bits of this thread's kCanonWordsPerThread words of a chunk.
__device__ __forceinline__ uint32_t canon_thread_bits(const Canon &c, uint32_t chunk,
                                                      uint32_t counts[kCanonWordsPerThread])
{
    const uint64_t w0 = (uint64_t)chunk * kCanonChunkWords + (uint64_t)threadIdx.x * kCanonWordsPerThread;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < kCanonWordsPerThread; ++i) {
        counts[i] = w0 + i < c.nwords ? (uint32_t)__popcll(c.bits[w0 + i]) : 0u;
        sum += counts[i];
    }
    return sum;
}

// partial[chunk] = set bits in the chunk (all 256 threads call it).
__device__ __forceinline__ void canon_partial_chunk(const Canon &c, uint32_t chunk)
{
    uint32_t counts[kCanonWordsPerThread], total;
    canon_block_sum(canon_thread_bits(c, chunk, counts), total);
    if (threadIdx.x == 0) c.partial[chunk] = total;
}

// prefix[w] for the chunk's words, after every partial is written.
__device__ __forceinline__ void canon_prefix_chunk(const Canon &c, uint32_t chunk)
{
    uint32_t base = 0, total;
    for (uint32_t i = threadIdx.x; i < chunk; i += kCanonThreads) base += c.partial[i];   // chunks before
    canon_block_sum(base, total);
    base = total;
    uint32_t counts[kCanonWordsPerThread];
    const uint32_t mine = canon_thread_bits(c, chunk, counts);
    uint32_t run = base + canon_block_sum(mine, total) - mine;
    const uint64_t w0 = (uint64_t)chunk * kCanonChunkWords + (uint64_t)threadIdx.x * kCanonWordsPerThread;
#pragma unroll
    for (uint32_t i = 0; i < kCanonWordsPerThread; ++i) {
        if (w0 + i < c.nwords) c.prefix[w0 + i] = run;
        run += counts[i];
    }
}

// Zero the chunk's words of the bit array.
__device__ __forceinline__ void canon_clear_chunk(const Canon &c, uint32_t chunk)
{
    const uint64_t w0 = (uint64_t)chunk * kCanonChunkWords;
    for (uint32_t i = threadIdx.x; i < kCanonChunkWords; i += kCanonThreads)
        if (w0 + i < c.nwords) c.bits[w0 + i] = 0ull;
}

inline uint32_t canon_chunks(uint32_t nwords) { return (nwords + kCanonChunkWords - 1) / kCanonChunkWords; }

}  // namespace lbl
