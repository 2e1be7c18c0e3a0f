// uf_device.hpp -- lock-free union-find primitives used by every labeling kernel.
//
// Invariant everywhere: parent(x) <= x, so the root of a set is its minimum
// element and links only ever decrease.  Linking uses atomicMin on the larger
// root: if the target stopped being a root concurrently, atomicMin keeps the
// smaller of the two parents and the loop continues with the displaced parent,
// so no equivalence is lost.  max(a, b) strictly decreases on every retry, so
// the loops terminate.  Stale reads are harmless: values only decrease and any
// value ever stored is an element of the same set.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbl {

// ---- loads that observe concurrent atomics --------------------------------

__device__ __forceinline__ uint32_t lds_load(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_store(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint32_t dev_load(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void dev_store(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t dev_min(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t lds_min(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---- LDS forest: parent of element x stored at par[Map(x)] ------------------

// Identity placement.
struct LinearMap {
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const { return x; }
};

// Element x = SEG * t + k (block k of thread t) stored at k * NT + t: when a
// wave touches block k of each of its threads the 64 words are consecutive,
// i.e. in distinct LDS banks.  With x in place, the 64 lanes would hit
// 32 / SEG banks only.
template <uint32_t SEG, uint32_t NT>
struct SwizzleMap {
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const { return (x % SEG) * NT + x / SEG; }
};

// Root of x with path halving: x's entry is lowered to its grandparent (by
// atomicMin, so an entry never grows) before stepping there.  Runs of a
// dense tile link into long chains without it.
template <class Map = LinearMap>
__device__ __forceinline__ uint32_t lds_find(uint32_t *par, uint32_t x, Map at = Map())
{
    uint32_t p = lds_load(&par[at(x)]);
    while (p != x) {
        const uint32_t g = lds_load(&par[at(p)]);
        if (g == p) return p;
        lds_min(&par[at(x)], g);
        x = g;
        p = lds_load(&par[at(x)]);
    }
    return x;
}

template <class Map = LinearMap>
__device__ __forceinline__ void lds_union(uint32_t *par, uint32_t a, uint32_t b, Map at = Map())
{
    for (;;) {
        a = lds_find(par, a, at);
        b = lds_find(par, b, at);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }
        uint32_t old = lds_min(&par[at(a)], b);
        if (old == a) return;
        a = old;
    }
}

// ---- global forest with a packed low field ---------------------------------
// word = parent << SHIFT | low, where `low` (the block's foreground mask) never
// changes.  SHIFT = 0 with an all-zero low field is the plain parent array.
//
// Callers start find/unite at tile-local roots (one hop from any element, as
// the tile kernels leave it), so path halving only ever rewrites root entries:
// an ordinary element keeps pointing at its tile-local root until compress.

template <int SHIFT>
struct PackedForest {
    static constexpr uint32_t kLowMask = (1u << SHIFT) - 1u;

    // Root of x by a read-only chase; the starting element is then pointed at
    // the root, so the next find from it (typically another border union of
    // the same tile-local root) takes one hop.  (Halving every visited entry
    // costs a device-scope atomic per hop, and the entries near a large
    // component's root are visited by thousands of finds at once.)
    __device__ static uint32_t find(uint32_t *w, uint32_t x)
    {
        const uint32_t x0 = x;
        const uint32_t v0 = dev_load(&w[x]);
        uint32_t v = v0, root;
        for (;;) {
            const uint32_t p = v >> SHIFT;
            if (p == x) { root = x; break; }
            const uint32_t vp = dev_load(&w[p]);
            const uint32_t gp = vp >> SHIFT;
            if (gp == p) { root = p; break; }
            x = gp;
            v = dev_load(&w[x]);
        }
        if (root != x0 && root != (v0 >> SHIFT)) dev_min(&w[x0], (root << SHIFT) | (v0 & kLowMask));
        return root;
    }

    // Root of x when nothing writes the forest any more (plain cached loads).
    __device__ static uint32_t find_final(const uint32_t *__restrict__ w, uint32_t x)
    {
        uint32_t v = w[x];
        while ((v >> SHIFT) != x) {
            x = v >> SHIFT;
            v = w[x];
        }
        return x;
    }

    __device__ static void unite(uint32_t *w, uint32_t a, uint32_t b)
    {
        for (;;) {
            a = find(w, a);
            b = find(w, b);
            if (a == b) return;
            if (a < b) { uint32_t t = a; a = b; b = t; }
            uint32_t va = dev_load(&w[a]);
            if ((va >> SHIFT) != a) continue;   // a was linked meanwhile: re-find
            uint32_t old = dev_min(&w[a], (b << SHIFT) | (va & kLowMask));
            if ((old >> SHIFT) == a) return;
            a = old >> SHIFT;
        }
    }
};

// ---- label forest: label = index + 1, 0 = background -----------------------
// Used by label equivalence, where the output label image itself is the
// equivalence array (as in the reference, LabelingAlgs.cl:105-106,275-276).

struct LabelForest {
    __device__ static uint32_t find(uint32_t *L, uint32_t x)
    {
        uint32_t v = dev_load(&L[x]);
        for (;;) {
            uint32_t p = v - 1u;
            if (p == x) return x;
            uint32_t vp = dev_load(&L[p]);
            uint32_t gp = vp - 1u;
            if (gp == p) return p;
            dev_min(&L[x], vp);
            x = gp;
            v = dev_load(&L[x]);
        }
    }

    __device__ static void unite(uint32_t *L, uint32_t a, uint32_t b)
    {
        for (;;) {
            a = find(L, a);
            b = find(L, b);
            if (a == b) return;
            if (a < b) { uint32_t t = a; a = b; b = t; }
            uint32_t old = dev_min(&L[a], b + 1u);
            if (old == a + 1u) return;
            a = old - 1u;
        }
    }

    // As PackedForest::compress: x and its tile-local root -> final root.
    __device__ static void compress(uint32_t *L, uint32_t x)
    {
        const uint32_t r = dev_load(&L[x]) - 1u;
        const uint32_t vr = dev_load(&L[r]);
        if (vr - 1u == r) return;
        const uint32_t root = find(L, r);
        dev_min(&L[r], root + 1u);
        if (r != x) dev_min(&L[x], root + 1u);
    }
};

// ---- vector memory helpers --------------------------------------------------

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16-byte streaming store (bypasses cache retention): for label output that is
// written once and not read back by the kernels.
__device__ __forceinline__ void store_nt4(uint32_t *p, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    u32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(p));
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void store_nt2(uint32_t *p, uint32_t a, uint32_t b)
{
    u32x2 v = {a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x2 *>(p));
}

// ---- wave helpers ----------------------------------------------------------

// Component counting.  One counter word hit by every wave serialises tens of
// thousands of atomics on one L2 line, so counts go to kCountSlots words 64 B
// apart (a workgroup adds its total to slot blockIdx % kCountSlots) and
// count_reduce sums the slots afterwards.
constexpr uint32_t kCountSlots = 64;
constexpr uint32_t kCountStride = 16;   // words between slots
constexpr size_t kCountBytes = kCountSlots * kCountStride * sizeof(uint32_t);

// Workgroup-wide sum of v, added to slots[blockIdx % kCountSlots] by thread 0.
// Every thread of the workgroup must call it.
__device__ __forceinline__ void count_add(uint32_t *slots, uint32_t v)
{
    __shared__ uint32_t part[16];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const uint32_t wave = threadIdx.x >> 6, nwaves = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63u) == 0) part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t i = 0; i < nwaves; ++i) s += part[i];
        if (s) atomicAdd(&slots[(blockIdx.x % kCountSlots) * kCountStride], s);
    }
}

// Sum the slots into *count (called by one wave of kCountSlots threads).
__device__ __forceinline__ void count_reduce(const uint32_t *slots, uint32_t *count)
{
    uint32_t v = threadIdx.x < kCountSlots ? slots[threadIdx.x * kCountStride] : 0u;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (threadIdx.x == 0) *count = v;
}

template <int Dummy = 0>
__global__ __launch_bounds__(kCountSlots) void count_reduce_kernel(const uint32_t *slots, uint32_t *count)
{
    count_reduce(slots, count);
}

// First lane of a run of consecutive lanes holding the same (domain, key), and
// that lane's index (lanes whose `active` is false never lead and never match;
// domain separates images of a batch, whose indices overlap).  All lanes of
// the wave must call it.
__device__ __forceinline__ bool run_leader(uint64_t key, uint32_t domain, bool active, uint32_t &leader_lane)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t prev = __shfl_up(key, 1, 64);
    const uint32_t prev_dom = __shfl_up(domain, 1, 64);
    const bool prev_active = __shfl_up((int)active, 1, 64) != 0;
    const bool lead = active && (lane == 0 || !prev_active || prev != key || prev_dom != domain);
    const uint64_t leaders = __ballot(lead);
    const uint64_t below = leaders & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
    leader_lane = below ? 63u - (uint32_t)__clzll(below) : lane;
    return lead;
}

// Zero the slots (called by one workgroup before any count_add).
__device__ __forceinline__ void count_clear(uint32_t *slots)
{
    if (threadIdx.x < kCountSlots) slots[threadIdx.x * kCountStride] = 0;
}

// Nonzero-byte test: bit 8k of the result is set iff byte k of v is nonzero.
__device__ __forceinline__ uint32_t nz_bytes(uint32_t v)
{
    uint32_t t = ((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v;
    return (t >> 7) & 0x01010101u;
}

// Two 4-pixel words of consecutive rows -> the 2x2 masks of the two blocks
// they cover, as two nibbles (bit0 = top-left, bit1 = top-right,
// bit2 = bottom-left, bit3 = bottom-right).
__device__ __forceinline__ uint32_t quad_masks(uint32_t top, uint32_t bot)
{
    uint32_t t = nz_bytes(top), b = nz_bytes(bot);
    return (t & 1u) | ((t >> 7) & 2u) | ((b & 1u) << 2) | ((b >> 5) & 8u) |
           ((t >> 12) & 0x10u) | ((t >> 19) & 0x20u) | ((b >> 10) & 0x40u) | ((b >> 17) & 0x80u);
}

}  // namespace lbl
