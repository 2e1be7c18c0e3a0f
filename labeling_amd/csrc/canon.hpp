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

}  // namespace lbl
