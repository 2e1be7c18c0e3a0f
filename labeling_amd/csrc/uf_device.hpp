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

#include "lbl_internal.hpp"

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

// A load that may be served by the caches of this CU / XCD and so may return
// an earlier value of the location than other CUs have written (a plain
// global_load, no device-coherence bits); atomic only so that concurrent
// writes are not a data race.
__device__ __forceinline__ uint32_t cached_load(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

__device__ __forceinline__ uint32_t dev_min(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t lds_min(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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

    // Root of x by a chase with cached loads, which may return earlier parents:
    // every value a parent entry ever held is an element of the node's final
    // set and below it, so the chase ends at an element of x's set (possibly a
    // root that has been linked since), and pointing x at it keeps parent <=
    // self and the partition.  Faster than device-coherent loads, which miss
    // the XCD's L2.
    __device__ static uint32_t find_cached(uint32_t *w, uint32_t x)
    {
        const uint32_t x0 = x;
        const uint32_t v0 = cached_load(&w[x]);
        uint32_t v = v0, root;
        for (;;) {
            const uint32_t p = v >> SHIFT;
            if (p == x) { root = x; break; }
            const uint32_t vp = cached_load(&w[p]);
            const uint32_t gp = vp >> SHIFT;
            if (gp == p) { root = p; break; }
            x = gp;
            v = cached_load(&w[x]);
        }
        if (root != x0 && root != (v0 >> SHIFT)) dev_min(&w[x0], (root << SHIFT) | (v0 & kLowMask));
        return root;
    }

    // Unite the sets of a and b; true iff this call made the link that joined
    // two different sets (exactly one call per merge of two sets returns
    // true).  Finds by find_cached; whether a candidate root still is one is
    // checked with a device-coherent load; if it is not, the union continues
    // from its fresh parent (finding it again could return the same stale
    // root forever).  Each pass either ends or strictly lowers max(a, b).
    __device__ static bool unite_cached(uint32_t *w, uint32_t a, uint32_t b)
    {
        for (;;) {
            a = find_cached(w, a);
            b = find_cached(w, b);
            if (a == b) return false;
            if (a < b) { uint32_t t = a; a = b; b = t; }
            const uint32_t va = dev_load(&w[a]);
            if ((va >> SHIFT) != a) {
                a = va >> SHIFT;
                continue;
            }
            const uint32_t old = dev_min(&w[a], (b << SHIFT) | (va & kLowMask));
            if ((old >> SHIFT) == a) return true;
            a = old >> SHIFT;
        }
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

// count_add for per-lane values below 32 (a tile lane's roots: at most one per
// block of its 16): the wave's sum by five ballots and popcounts on the scalar
// unit, then one atomic per wave; no LDS, no barrier, and the waves of a
// workgroup spread over the slots.  All lanes of the wave must call it.
__device__ __forceinline__ void count_add_wave(uint32_t *slots, uint32_t v)
{
    uint32_t s = 0;
#pragma unroll
    for (uint32_t bit = 0; bit < 5; ++bit) s += (uint32_t)__popcll(__ballot((v >> bit) & 1u)) << bit;
    const uint32_t slot = (blockIdx.x * ((blockDim.x + 63) >> 6) + (threadIdx.x >> 6)) % kCountSlots;
    if ((threadIdx.x & 63u) == 0 && s) atomicAdd(&slots[slot * kCountStride], s);
}

// The previous lane's v (lane 0: its own): a DPP wave shift right by one lane
// (wave_shr:1, a GFX9-family DPP control), a VALU move with no LDS round trip.
// Every lane of the wave must execute it.
__device__ __forceinline__ uint32_t lane_prev(uint32_t v)
{
    constexpr int kWaveShr1 = 0x138;
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, kWaveShr1, 0xF, 0xF, false);
}

// Bit k of the result: bit 2k or bit 2k + 1 of s (a 2-column block holds a
// set bit of a 32-column row mask), k < 16.
__device__ __forceinline__ uint32_t pair_any16(uint32_t s)
{
    uint32_t x = (s | (s >> 1)) & 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    return (x | (x >> 8)) & 0xFFFFu;
}

// Sum the slots into *count and zero them again (called by one full wave,
// threads 0..63, after every kernel that adds to them has finished).  Slots
// are zero between pipeline runs: workspaces zero them when allocated and
// every run that counts ends here.
__device__ __forceinline__ void count_reduce(uint32_t *slots, uint32_t *count)
{
    uint32_t v = slots[threadIdx.x * kCountStride];
    slots[threadIdx.x * kCountStride] = 0;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (threadIdx.x == 0) *count = v;
}

// n links each joined two components: the count of the image whose slots
// these are drops by n (unsigned slot arithmetic wraps; only the sum over the
// slots is meaningful).  The lanes of a wave add up first when they all count
// for one image (one atomic per wave); slots may be null where n is 0.  All
// lanes of the wave must call it.
__device__ __forceinline__ void count_links(uint32_t *slots, uint32_t n)
{
    const uint64_t has = __ballot(n != 0);
    if (!has) return;
    const uint32_t lane = threadIdx.x & 63u, first = (uint32_t)__builtin_ctzll(has);
    const uint64_t key = reinterpret_cast<uintptr_t>(slots);
    const uint64_t key0 = __shfl(key, (int)first, 64);
    uint32_t *slot = &slots[(blockIdx.x % kCountSlots) * kCountStride];
    if (__ballot(n != 0 && key != key0) == 0) {
        uint32_t sum = n;
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
        if (lane == first) atomicSub(slot, sum);
    } else if (n) {
        atomicSub(slot, n);
    }
}

template <int Dummy = 0>
__global__ __launch_bounds__(kCountSlots) void count_reduce_kernel(uint32_t *slots, uint32_t *count)
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

// ---- distinct pairs, collected in LDS by a workgroup ------------------------------
//
// An open-addressing hash set of (a, b) pairs (a, b < 2^32 - 1) plus the list of
// its filled slots in insertion order, so that a workgroup can unite each
// distinct pair once, spread evenly over its threads:
//   init() by every thread, __syncthreads(), insert() by any thread,
//   __syncthreads(), then pairs 0..size()-1.
// insert() returns false when its probe sequence finds no room, or the list of
// filled slots is full (LIST = CAP - 2 entries, so that the whole set takes
// exactly 10 * CAP bytes of LDS): the caller handles that pair itself.
template <uint32_t CAP, uint32_t PROBES = 16>
struct LdsPairSet {
    static_assert((CAP & (CAP - 1)) == 0 && CAP >= 64 && CAP <= 65536, "CAP: power of two, 64..65536");
    static constexpr unsigned long long kEmpty = ~0ull;
    static constexpr uint32_t LIST = CAP - 2;
    unsigned long long key[CAP];
    uint16_t filled[LIST];
    uint32_t n;

    __device__ __forceinline__ void init()
    {
        for (uint32_t i = threadIdx.x; i < CAP; i += blockDim.x) key[i] = kEmpty;
        if (threadIdx.x == 0) n = 0;
    }
    __device__ __forceinline__ bool insert(uint32_t a, uint32_t b)
    {
        const unsigned long long k = (unsigned long long)a << 32 | b;
        uint32_t h = (a * 0x9E3779B1u ^ b * 0x85EBCA6Bu) >> (32 - __builtin_ctz(CAP));
        bool placed = false;
        for (uint32_t t = 0; t < PROBES; ++t, h = (h + 1) & (CAP - 1)) {
            const unsigned long long old = atomicCAS(&key[h], kEmpty, k);
            if (old == k) return true;   // already there
            if (old == kEmpty) {
                placed = true;
                break;
            }
        }
        if (!placed) return false;
        // first insertion of k: list it, or hand it back when the list is full
        const uint32_t i = atomicAdd(&n, 1u);
        if (i >= LIST) return false;
        filled[i] = (uint16_t)h;
        return true;
    }
    __device__ __forceinline__ uint32_t size() const { return n < LIST ? n : LIST; }
    __device__ __forceinline__ uint2 pair(uint32_t i) const
    {
        const unsigned long long k = key[filled[i]];
        return make_uint2((uint32_t)(k >> 32), (uint32_t)k);
    }
};

// ---- bit masks of pixel rows ------------------------------------------------

// bit i set <=> byte i of w is non-zero (i = 0..3)
__device__ __forceinline__ uint32_t nz_bits4(uint32_t w)
{
    const uint32_t nz = (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
    return (nz * 0x00204081u) >> 28;
}

__device__ __forceinline__ uint32_t nz_bits16(uint4 q)
{
    return nz_bits4(q.x) | (nz_bits4(q.y) << 4) | (nz_bits4(q.z) << 8) | (nz_bits4(q.w) << 12);
}

__device__ __forceinline__ uint32_t hibit(uint32_t v) { return 31u - __builtin_clz(v); }   // v != 0
__device__ __forceinline__ uint32_t upto(uint32_t x) { return (2u << x) - 1u; }           // bits 0..x

// 32 pixels of one row as a bit mask (bit k = column x0 + k), zero past the
// width; x0 < w.  vec: row + x0 is 16-byte aligned.  Then the segment is read
// as its (at most two) 16-byte chunks holding pixels of the row, and the bits
// of columns >= w are masked off: an aligned 16-byte chunk never crosses a
// page, so the bytes it holds after the row's last pixel can be read without
// a fault (they belong to the padding, the next row or the end of the
// buffer's last page).  The second chunk, when wholly past the row end, is
// not read.
__device__ __forceinline__ uint32_t tail_mask(uint32_t x0, uint32_t w)
{
    return w - x0 >= 32 ? ~0u : (1u << (w - x0)) - 1u;
}

// FULL: every segment is whole (the image width is a multiple of 32), so no
// row end is handled.
template <bool FULL = false>
__device__ __forceinline__ uint32_t row_bits32(const uint8_t *row, uint32_t x0, uint32_t w, bool vec)
{
    if (vec) {
        const uint4 *q = reinterpret_cast<const uint4 *>(row + x0);
        if (FULL) return nz_bits16(q[0]) | (nz_bits16(q[1]) << 16);
        const uint32_t lo = nz_bits16(q[0]);
        const uint32_t hi = x0 + 16 < w ? nz_bits16(q[1]) : 0u;   // a branch, not an address select
        return (lo | (hi << 16)) & tail_mask(x0, w);
    }
    uint32_t m = 0;
    for (uint32_t k = 0; k < 32 && x0 + k < w; ++k) m |= (uint32_t)(row[x0 + k] != 0) << k;
    return m;
}

// The same for N rows at byte offsets off[i] from row0 (mask 0 where
// !use[i]).  On the vector path every row's loads are issued before the first
// is converted, so a lane waits for one memory round trip instead of N in a
// row; an unused row re-reads row0 (cached) instead of branching around its
// load.
template <uint32_t N, bool FULL = false>
__device__ __forceinline__ void rows_bits32(const uint8_t *row0, const size_t (&off)[N], const bool (&use)[N], uint32_t x0,
                                            uint32_t w, bool vec, uint32_t (&m)[N])
{
    if (vec) {
        // the second chunk is read unless wholly past the row end (a branch
        // on the lane's own condition, uniform in waves that hold no row end:
        // a select of its address would split its load into dword loads)
        // each row's two chunks are requested back to back
        const bool second = FULL || x0 + 16 < w;
        uint4 lo[N], hi[N];
#pragma unroll
        for (uint32_t i = 0; i < N; ++i) {
            const uint4 *q = reinterpret_cast<const uint4 *>(row0 + (use[i] ? off[i] : 0) + x0);
            lo[i] = q[0];
            hi[i] = make_uint4(0u, 0u, 0u, 0u);
            if (second) hi[i] = q[1];
        }
        // Needed here, unconditionally: keeps the compiler from sinking a
        // load into the branch of its `use` test, behind the other rows' waits.
#pragma unroll
        for (uint32_t i = 0; i < N; ++i) {
            asm volatile("" : "+v"(lo[i].x), "+v"(lo[i].y), "+v"(lo[i].z), "+v"(lo[i].w));
            asm volatile("" : "+v"(hi[i].x), "+v"(hi[i].y), "+v"(hi[i].z), "+v"(hi[i].w));
        }
#pragma unroll
        for (uint32_t i = 0; i < N; ++i) m[i] = use[i] ? nz_bits16(lo[i]) | (nz_bits16(hi[i]) << 16) : 0u;
        if (!FULL && x0 + 32 > w) {   // the row's last segment (a branch: uniform in waves that hold no row end)
            const uint32_t keep = tail_mask(x0, w);
#pragma unroll
            for (uint32_t i = 0; i < N; ++i) m[i] &= keep;
        }
        return;
    }
#pragma unroll
    for (uint32_t i = 0; i < N; ++i) m[i] = use[i] ? row_bits32<false>(row0 + off[i], x0, w, false) : 0u;
}

// ---- run forests of bit-mask segments ------------------------------------------
// The band / bar tile kernels give each lane a 32-pixel segment (16 blocks) of
// one row of blocks; all lanes of a row are consecutive lanes of one wave.  A
// segment-local run of set bits is a node named by the block of its first
// column.  Node names are byte offsets into the tile's LDS forest (4 x the
// block's local index), so that an entry's address is its name: a parent read
// is used as the next address with no shift.  The entry of node n is
// node_entry(par, n), holding its parent's name, parent <= n.

constexpr uint32_t kLaneBlocks = 16;          // blocks per lane (32 pixel columns)
constexpr uint32_t kResolved = 0x80000000u;   // forest tag: entry holds the global root
constexpr uint32_t kNodeBytes = 4;            // node name = kNodeBytes * local block index

__device__ __forceinline__ uint32_t *node_entry(uint32_t *par, uint32_t n)
{
    return reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(par) + n);
}

// Name of the node of block s >> 1 of a segment whose first block is named base
// (s: a column of the segment's 32).
__device__ __forceinline__ uint32_t node_at(uint32_t base, uint32_t s)
{
    return base + ((s >> 1) * kNodeBytes);
}

// Node standing for the first run of segment j of a row whose first block is
// row0 (o: the segment's run mask): the segment's own first block, or, when
// the run continues from the segment to the left, the first block of the
// whole run -- found with a ballot over the row's lanes (the nearest segment
// to the left that is not all set).  All lanes of the wave must call it.
// (row0: the name of the row's first block; returns a name.)
__device__ __forceinline__ uint32_t run_head(uint32_t o, uint32_t j, uint32_t row0)
{
    constexpr uint32_t SEG = kLaneBlocks * kNodeBytes;     // name distance of consecutive segments
    const uint32_t lw = threadIdx.x & 63u, base = lw - j;   // lane of segment 0 of this row
    const uint64_t nonfull = __ballot(o != 0xFFFFFFFFu);
    const uint64_t below = nonfull & ((1ull << lw) - 1ull) & ~((1ull << base) - 1ull);
    const uint32_t k = below ? 63u - (uint32_t)__builtin_clzll(below) : lw;
    const uint32_t o_left = __shfl(o, (int)((lw + 63u) & 63u), 64);
    const uint32_t o_k = __shfl(o, (int)k, 64);
    if (!(j > 0 && (o & 1u) && (o_left >> 31))) return row0 + j * SEG;
    const uint32_t sk = row0 + (k - base) * SEG;   // segment k's first block
    return !below ? row0 : ((o_k >> 31) ? node_at(sk, hibit(o_k & ~(o_k << 1))) : sk + SEG);
}

// One step of uniting the sets of x and y by Rem's algorithm with splicing:
// of the two nodes, the one with the larger parent takes the other's parent
// (atomicMin keeps parent <= self; it moves the node, with its subtree, into
// the other set) and continues from its old parent, until both have the same
// parent or a root is linked.  It stops short of the roots when the sets are
// already one, which most contacts of a dense image are.  Returns true when
// done, with `top` the node where the sets met: a member of the joined set,
// from which a lane's next union on the same node may start (no node moves).
//
// A splice cuts the node's old set in two, and the steps that follow join the
// halves again: every cut has a pending join.  No node may be moved without
// one.  Pointing the two nodes a union started from at the node where their
// sets met, once it is done, breaks this: the moved node can be where another
// lane's union still stands, the only pending join of the set the node leaves;
// that union then finds its two nodes together and stops, and the rest of
// that set stays apart.
__device__ __forceinline__ bool lds_rem_step(uint32_t *par, uint32_t &x, uint32_t &y, uint32_t &top)
{
    const uint32_t px = lds_load(node_entry(par, x)), py = lds_load(node_entry(par, y));
    top = px < py ? px : py;
    if (px == py) return true;
    if (px < py) {   // the node with the larger parent moves
        const uint32_t t = x;
        x = y;
        y = t;
    }
    const uint32_t old = lds_min(node_entry(par, x), top);
    if (old == x) return true;   // x was a root: now under the other parent
    x = old;   // x's old parent (or whatever lowered it meanwhile) still has to join
    return false;
}

// Contacts between a segment's facing row A and the facing row B of a
// neighbouring row of segments, same columns (Bl / Br: that row's segments to
// the left / right), where facing pixels touch when their columns differ by at
// most one:
//   e0      straight contacts: the first column of each run of A & B, and of
//           those only the first after a run start of either row (st, stn;
//           carry trick: adding the boundary bits to the complement runs each
//           carry up to the next contact of its segment);
//   el, er  diagonal contacts (B at x-1 / x+1), only where neither the
//           straight contact nor the neighbouring A pixel exists (otherwise
//           they repeat a straight contact's pair of runs).
// Popped in column order (neighbour column x-1 before x before x+1), the pairs
// of runs they join are monotone in both rows, so equal pairs are adjacent.
struct SegContacts {
    uint32_t e0, el, er;
    __device__ __forceinline__ SegContacts(uint32_t A, uint32_t B, uint32_t Bl, uint32_t Br, uint32_t st,
                                           uint32_t stn)
    {
        const uint32_t v0 = A & B, f0 = v0 & ~(v0 << 1), bnd = st | stn | 1u;
        const uint32_t y = ~f0 & ~(bnd >> 1);
        e0 = ((y + bnd) & ~y) & f0;
        el = A & ((B << 1) | (Bl >> 31)) & ~B & ~(A << 1);
        er = A & ((B >> 1) | (Br << 31)) & ~B & ~(A >> 1);
    }
    __device__ __forceinline__ uint32_t any() const { return e0 | el | er; }
    // pop any contact: the straight ones, then the left, then the right
    // diagonals (fewer instructions than pop; the tile kernels' unions do not
    // depend on the order); returns the neighbour column
    __device__ __forceinline__ int pop_any(uint32_t &xo)
    {
        const bool c0 = e0 != 0, cl = el != 0;
        const uint32_t m = c0 ? e0 : (cl ? el : er);
        const uint32_t x = __builtin_ctz(m), rest = m & (m - 1u);
        e0 = c0 ? rest : e0;
        el = !c0 && cl ? rest : el;
        er = !c0 && !cl ? rest : er;
        xo = x;
        return (int)x + (c0 ? 0 : (cl ? -1 : 1));
    }
    // pop the next contact: its own column x, returns the neighbour column (-1..32)
    __device__ __forceinline__ int pop(uint32_t &xo)
    {
        const uint32_t x = __builtin_ctz(any()), bit = 1u << x;
        const bool tl = (el & bit) != 0, t0 = (e0 & bit) != 0;
        el ^= tl ? bit : 0u;
        e0 ^= (!tl && t0) ? bit : 0u;
        er ^= (!tl && !t0) ? bit : 0u;
        xo = x;
        return (int)x + (tl ? -1 : (t0 ? 0 : 1));
    }
};

// Nodes of a row of segments: the run containing column p of segment j (p in
// -1..32 reaches into the segments to the left / right).
struct SegNodes {
    uint32_t st;      // run starts of segment j
    uint32_t base;    // name of segment j's first block
    uint32_t first;   // node standing for segment j's first run
    uint32_t left;    // node of the left segment's last run
    uint32_t right;   // node standing for the right segment's first run
    __device__ __forceinline__ uint32_t node(int p) const
    {
        if (p < 0) return left;
        if (p > 31) return right;
        const uint32_t s = hibit(st & upto((uint32_t)p));
        return s ? node_at(base, s) : first;
    }
};

// Node of the last run of a segment with run mask o (o != 0) whose first block
// is named base, `first` standing for its first run.
__device__ __forceinline__ uint32_t last_run_node(uint32_t o, uint32_t base, uint32_t first)
{
    const uint32_t s = o ? hibit(o & ~(o << 1)) : 0u;
    return s ? node_at(base, s) : first;
}

// Chase the nodes of blocks k (the bits of pend) of the segment whose first
// block is named base to their roots and store kResolved | glob(root) in their
// entries (glob: local block index -> global block index).  A chase stops at
// an entry another lane already resolved.  No unions may run concurrently.
template <class Glob>
__device__ __forceinline__ void resolve_nodes(uint32_t *par, uint32_t base, uint32_t pend, const Glob &glob)
{
    uint32_t cur = 0, orig = 0;
    bool busy = false;
    for (;;) {
        if (!busy) {
            if (!pend) break;
            orig = cur = base + __builtin_ctz(pend) * kNodeBytes;
            pend &= pend - 1;
            busy = true;
        }
        const uint32_t p = lds_load(node_entry(par, cur));
        uint32_t g;
        if (p & kResolved) {
            g = p;
        } else if (p == cur) {
            g = kResolved | glob(cur / kNodeBytes);
        } else {
            cur = p;
            continue;
        }
        lds_store(node_entry(par, orig), g);
        busy = false;
    }
}

// Nonzero-byte test: bit 8k of the result is set iff byte k of v is nonzero.
__device__ __forceinline__ uint32_t nz_bytes(uint32_t v)
{
    uint32_t t = ((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v;
    return (t >> 7) & 0x01010101u;
}

}  // namespace lbl
