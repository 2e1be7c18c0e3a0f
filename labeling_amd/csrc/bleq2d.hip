// bleq2d.hip -- 2D block-equivalence labeling (2x2 blocks, 8-connectivity) for
// gfx950: single images and batches of images of uneven sizes.
//
// Same problem as the reference's TOCLLabelEquivalenceX2 (LabelingAlgs.cpp:948-1084,
// kernels LBEQ2_* LabelingAlgs.cl:133-319): pixels are grouped into 2x2 blocks,
// blocks with a foreground pixel are the nodes, two blocks are equivalent when a
// foreground pixel of one is 8-adjacent to a foreground pixel of the other, and
// each component is labelled by its minimum block.  The reference iterates
// Scan/Analyze passes over the whole block grid with a blocking host round trip
// per pass; here the equivalences are resolved by union-find in three or four
// launches and no host synchronisation:
//
//   1. tile    one workgroup per TBW x TBH-block tile: each lane turns two
//              32-pixel row segments into bit masks; union-find over the
//              runs of each 2-row band in LDS; writes one packed word per
//              block, parent << 4 | mask, whose parent is the tile-local root
//              (global block index).
//   2. merge   unions across tile borders: each tile's top band against the
//              band above (bit-mask contacts per 16-block segment) and its
//              first pixel column against the column to its left; pairs of
//              tile-local roots collected in LDS, united by all threads.
//   3. flatten (optional, chosen by the autotuner) one thread per tile-border
//              block: the block and its tile-local root are pointed straight
//              at the global root.
//   4. final   one thread per 2 blocks: root of the block, writes the labels
//              of the 2x2 pixels of each block, 0 for background.
//
// Components are counted at the tile-local roots by (1), each link made in (2)
// subtracts one, and the image's first workgroup of (4) sums the 64 spread
// count slots (zeroing them for the next run).
//
// Every kernel reads its image geometry from a "source": One (a single image
// passed by value) or Many (a device table of images, and the image of every
// tile / final workgroup / scan chunk id in tables planned on the host,
// batch_plan.hpp), so a batch of thousands of images runs as the same four
// launches.
//
// HBM traffic per pixel: 1 B image in + 1 B block words out in (1), 1 B block
// words in + 4 B labels out in (4), plus the border blocks in (2)-(3).
#include <algorithm>
#include <type_traits>

#include "bleq2d.hpp"
#include "uf_device.hpp"

namespace lbl {
namespace {

constexpr uint32_t FINAL_BLOCKS = plan::kFinalBlocksPerThread;   // blocks per thread in the final kernel: one
                                                                // 16-byte store per pixel row, so each wave
                                                                // store covers 1 KiB contiguously
constexpr uint32_t NT_FINAL = plan::kFinalThreads;

using Forest = PackedForest<4>;

// Block mask bits: 1 = (0,0), 2 = (1,0), 4 = (0,1), 8 = (1,1).

// ---- geometry sources ----------------------------------------------------------

struct One {
    Geom2D g;
    static constexpr bool kSingle = true;
    __device__ __forceinline__ uint32_t image_of_tile(uint32_t) const { return 0; }
    __device__ __forceinline__ const Geom2D &image(uint32_t) const { return g; }
    __device__ __forceinline__ const Geom2D &by_tile(uint32_t gt, uint32_t &lt) const
    {
        lt = gt;
        return g;
    }
    __device__ __forceinline__ const Geom2D &by_final(uint32_t gw, uint32_t &lw) const
    {
        lw = gw;
        return g;
    }
    __device__ __forceinline__ const Geom2D &by_chunk(uint32_t gc, uint32_t &lc) const
    {
        lc = gc;
        return g;
    }
};

struct Many {
    const Geom2D *g;
    uint32_t n;
    BatchTables t;
    static constexpr bool kSingle = false;
    __device__ __forceinline__ uint32_t image_of_tile(uint32_t gt) const { return t.tile_img[gt]; }
    __device__ __forceinline__ const Geom2D &image(uint32_t i) const { return g[i]; }
    __device__ __forceinline__ const Geom2D &by_tile(uint32_t gt, uint32_t &lt) const
    {
        const Geom2D &r = g[t.tile_img[gt]];
        lt = gt - r.tile_base;
        return r;
    }
    __device__ __forceinline__ const Geom2D &by_final(uint32_t gw, uint32_t &lw) const
    {
        const Geom2D &r = g[t.final_img[gw]];
        lw = gw - r.final_base;
        return r;
    }
    __device__ __forceinline__ const Geom2D &by_chunk(uint32_t gc, uint32_t &lc) const
    {
        const Geom2D &r = g[t.chunk_img[gc]];
        lc = gc - r.chunk_base;
        return r;
    }
};

// Final-kernel chunks per workgroup (plan::geometry2d).
template <class Src>
constexpr uint32_t final_chunks()
{
    return Src::kSingle ? 1u : plan::kBatchFinalChunks;
}

// The FINAL_BLOCKS consecutive blocks, in raster order, of this thread in chunk
// `chunk` of its image (chunk c takes blocks c * NT_FINAL * FINAL_BLOCKS on,
// whatever the rows): the first block k0 lies at (by, bx); n of the blocks lie
// in the image; they all lie in row by unless bx + FINAL_BLOCKS > bw.
struct Span {
    uint32_t k0, by, bx, n;
};

// n / d for the divisor of g (plan::div_magic; the same arithmetic as plan::div_by)
__device__ __forceinline__ uint32_t div_by(uint32_t n, const plan::DivMagic &g)
{
    const uint32_t t = __umulhi(g.m, n);
    return (t + ((n - t) >> g.s1)) >> g.s2;
}

// The chunk's first block row is found once per workgroup (uniform: scalar
// unit); a lane's offset from there wraps into the next row at most once when
// rows hold a workgroup's blocks, else by another division.
__device__ __forceinline__ Span final_span(const Geom2D &G, uint32_t chunk)
{
    Span sp;
    const uint32_t bw = G.bw, k_wg = chunk * (NT_FINAL * FINAL_BLOCKS);
    const uint32_t by_wg = div_by(k_wg, G.bdiv), off = threadIdx.x * FINAL_BLOCKS;
    sp.k0 = k_wg + off;
    const uint32_t nb = bw * G.bh;
    sp.n = sp.k0 >= nb ? 0u : min(FINAL_BLOCKS, nb - sp.k0);
    uint32_t by = by_wg, bx = k_wg - by_wg * bw + off;
    if (bw >= NT_FINAL * FINAL_BLOCKS) {
        if (bx >= bw) {
            bx -= bw;
            ++by;
        }
    } else {
        const uint32_t q = div_by(bx, G.bdiv);
        by += q;
        bx -= q * bw;
    }
    sp.by = by;
    sp.bx = bx;
    return sp;
}

// Row and column of block k of a span.
__device__ __forceinline__ void span_block(const Span &sp, uint32_t k, uint32_t bw, uint32_t &by, uint32_t &bx)
{
    const bool wrap = sp.bx + k >= bw;
    by = wrap ? sp.by + 1 : sp.by;
    bx = wrap ? sp.bx + k - bw : sp.bx + k;
}

// First-pixel key of block (bx, by) with mask m != 0 (bit 1 << (dx + 2 dy)):
// its first pixel's index (kPixelKeys), or first-pixel row * bw + bx, the bit
// position of canonical numbering.  Both order blocks' first pixels as the
// raster does, and no two blocks share a key.
__device__ __forceinline__ uint32_t block_key(const Geom2D &G, uint32_t by, uint32_t bx, uint32_t m)
{
    const uint32_t bit = __builtin_ctz(m), y = 2 * by + (bit >> 1);
    return (G.flags & kPixelKeys) ? y * G.w + 2 * bx + (bit & 1u) : y * G.bw + bx;
}

// ---- 1. tile kernel -------------------------------------------------------------
//
// A tile is TBW x TBH blocks; one thread ("lane") per 32-pixel-wide segment of
// one band (a block row = two pixel rows), TBW / 16 lanes per band, all lanes
// of a band in one wave.  A lane holds its segment as two 32-bit masks, `a`
// (top row) and `b` (bottom row).  Inside a band, 8-connectivity is exactly
// "same run of set bits of o = a | b" (two pixels in adjacent columns of a
// 2-row band always touch), so the union-find nodes are the segment-local runs,
// each named by the block of its first column:
//   * a run continuing the left segment's run is linked straight to the head
//     of the whole band run, found with a ballot over the band's lanes (the
//     nearest segment to the left that is not all foreground), and stands for
//     that head in every union;
//   * vertical contacts (a top-row pixel of this band touching a bottom-row
//     pixel of the band above, straight or diagonally) are found with bit
//     operations; of the straight ones only the first after each lower or
//     upper run boundary is kept (the others repeat its pair of runs).
// Each lane feeds its contacts through one union loop (a single loop body for
// all of a lane's contacts keeps divergent lanes in step; a finished union
// points both of its nodes at the root, which keeps later finds short), then
// resolves its nodes' roots and writes the 16 block words of its segment.

constexpr uint32_t LANE_BLOCKS = kLaneBlocks;

// dst[k] = v[k] for lo <= k < hi: compile-time indices only (a runtime index
// into a register array would place it in scratch memory or LDS).
template <uint32_t N>
__device__ __forceinline__ void store_words(uint32_t *dst, const uint32_t (&v)[N], uint32_t lo, uint32_t hi)
{
#pragma unroll
    for (uint32_t k = 0; k < N; ++k)
        if (k >= lo && k < hi) dst[k] = v[k];
}

// The same for four words held in scalars.
__device__ __forceinline__ void store_words4(uint32_t *dst, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3,
                                             uint32_t lo, uint32_t hi)
{
    if (lo <= 0 && 0 < hi) dst[0] = v0;
    if (lo <= 1 && 1 < hi) dst[1] = v1;
    if (lo <= 2 && 2 < hi) dst[2] = v2;
    if (lo <= 3 && 3 < hi) dst[3] = v3;
}

// x_o (o < 4) by two selects on the bits of o: a choice between register
// values that the compiler cannot fold into a runtime index.
__device__ __forceinline__ uint32_t by_offset(uint32_t o, uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3)
{
    const uint32_t even = (o & 2u) ? x2 : x0, odd = (o & 2u) ? x3 : x1;
    return (o & 1u) ? odd : even;
}

// Word t of the 16-byte-aligned store windows of a lane whose words start o
// words past an aligned address: its own out[t - o], or for t < o the previous
// lane's last words (p13, p14, p15 = its out[13..15]).
__device__ __forceinline__ uint32_t window_word(const uint32_t (&out)[kLaneBlocks], uint32_t p13, uint32_t p14,
                                                uint32_t p15, uint32_t t, uint32_t o)
{
    const auto at = [&](int i) { return i >= 0 ? out[i] : i == -1 ? p15 : i == -2 ? p14 : p13; };
    const int ti = (int)t;
    return by_offset(o, at(ti), at(ti - 1), at(ti - 2), at(ti - 3));
}

// Store a tile lane's LANE_BLOCKS block words at word index `at` of P, the
// lane's first block being column bxs of its row (inside: the lane holds
// blocks of the image; min(LANE_BLOCKS, bw - bxs) of its words are), lane j of
// its band.  ALIGNED: bw % 16 == 0, so every lane inside the image is whole.
// When the
// rows are not 16-byte aligned (bw % 4 != 0), each full lane stores the
// aligned 4-word windows that end inside its own words, the first one
// completed with the previous lane's last words (by shuffle); a band's first
// lane starts at its own first word, and a lane whose next lane is not full
// stores its own last words itself.  Every lane of the wave must call it.
template <uint32_t LPB, bool ALIGNED>
__device__ __forceinline__ void store_block_words(uint32_t *P, size_t at, uint32_t bxs, bool inside, uint32_t j,
                                                  uint32_t bw, const uint32_t (&out)[kLaneBlocks])
{
    uint32_t *dst = P + at;
    const bool full = inside && bxs + kLaneBlocks <= bw;
    const auto store_whole = [&]() {
        uint4 *d4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
        for (uint32_t q = 0; q < kLaneBlocks / 4; ++q) d4[q] = make_uint4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
    };
    if (ALIGNED) {   // an exact image: every lane inside it is whole
        if (inside) store_whole();
        return;
    }
    if ((bw % 4u) == 0) {   // aligned rows (uniform per image)
        if (full) {
            store_whole();
            // keeps the compiler from merging the last words' store with the
            // partial lanes' (one more store instruction per whole lane)
            asm volatile("" ::: "memory");
        } else if (inside) {
            store_words(dst, out, 0, bw - bxs);
        }
        return;
    }
    const uint32_t p13 = lane_prev(out[13]), p14 = lane_prev(out[14]), p15 = lane_prev(out[15]);
    if (full) {
        const uint32_t o = (uint32_t)(at & 3u);
        uint4 *base = reinterpret_cast<uint4 *>(dst - o);
        const bool has_prev = j > 0;   // a full or partial lane's previous lane in the band is full
#pragma unroll
        for (uint32_t q = 0; q < kLaneBlocks / 4; ++q) {
            const uint32_t w0 = window_word(out, p13, p14, p15, 4 * q, o);
            const uint32_t w1 = window_word(out, p13, p14, p15, 4 * q + 1, o);
            const uint32_t w2 = window_word(out, p13, p14, p15, 4 * q + 2, o);
            const uint32_t w3 = window_word(out, p13, p14, p15, 4 * q + 3, o);
            if (q == 0 && !has_prev && o != 0) {   // a band's first lane: its own words only
                store_words(dst, out, 0, 4 - o);
            } else {
                base[q] = make_uint4(w0, w1, w2, w3);
            }
        }
        const bool next_full = j + 1 < LPB && bxs + 2 * kLaneBlocks <= bw;
        if (o && !next_full) store_words(dst, out, kLaneBlocks - o, kLaneBlocks);
    } else if (inside) {
        store_words(dst, out, 0, bw - bxs);
    }
}

// CONN4: 4-connectivity (the label equivalence path): blocks are 1x2 pixels and
// a band is one pixel row; a row run is 4-connected, and only straight
// contacts join rows.
// Tiles of 128 blocks and wider: at most 64 VGPRs (8 waves per SIMD); single
// images need 65 otherwise, which leaves three 512-lane workgroups per CU where
// LDS holds four (4% faster); batches fit as well without spilling (4096-image
// batch 10.93 -> 10.47 ms).
// EXACT: a single image flagged kExact (no row ends, aligned rows): the
// kernel holds no code for them.
template <uint32_t TBW, uint32_t TBH, class Src, bool CONN4, bool EXACT>
__global__ __launch_bounds__((TBW / LANE_BLOCKS) * TBH, TBW >= 128 ? 8 : 1) void bleq2d_tile(Src src)
{
    constexpr uint32_t LPB = TBW / LANE_BLOCKS;   // lanes per band
    static_assert((TBW & (TBW - 1)) == 0 && LPB >= 1 && LPB <= 64, "TBW: power of two, 16..1024");
    constexpr uint32_t LOG_TBW = __builtin_ctz(TBW);
    __shared__ uint32_t par[TBH * TBW];            // node (run start block) -> parent, local indices
    __shared__ uint32_t so[TBH + 1][LPB + 2];      // o = a | b per band/segment; guard row 0, guard columns
    __shared__ uint32_t sb[TBH + 1][LPB + 2];      // b (bottom row)
    __shared__ uint32_t sh[TBH + 1][LPB + 2];      // node standing for the segment's first run

    uint32_t lt;
    const Geom2D &G = src.by_tile(blockIdx.x, lt);
    const uint32_t h = G.h, w = G.w, bw = G.bw, bh = G.bh;
    const uint32_t t = threadIdx.x, r = t / LPB, j = t % LPB;
    const uint32_t bx0 = (lt % G.tiles_x) * TBW, by0 = (lt / G.tiles_x) * TBH;
    const uint32_t by = by0 + r, bxs = bx0 + j * LANE_BLOCKS;

    uint32_t a = 0, b = 0;
    if (by < bh && bxs < bw) {
        const uint32_t y0 = CONN4 ? by : 2 * by, x0 = 2 * bxs;
        const uint8_t *row0 = G.px + (size_t)y0 * G.pitch;
        const bool vec = EXACT || (G.flags & kVecIn) != 0;
        if (CONN4) {
            a = row_bits32<EXACT>(row0, x0, w, vec);
        } else {
            const size_t off[2] = {0, G.pitch};
            const bool use[2] = {true, y0 + 1 < h};
            uint32_t m[2];
            rows_bits32<2, EXACT>(row0, off, use, x0, w, vec, m);
            a = m[0]; b = m[1];
        }
        // the reference's label equivalence never labels pixel (0,0) and reads
        // no neighbour through it (DistrInitKernel, GetLabel: LabelingAlgs.cl:37,44-52)
        if ((G.flags & kDropOrigin) && (by | bxs) == 0) a &= ~1u;
    }
    const uint32_t o = a | b;
    const uint32_t st = o & ~(o << 1);               // first column of each segment-local run
    const uint32_t nb = r * TBW + j * LANE_BLOCKS;   // local index of the segment's first block
    const uint32_t nbn = nb * kNodeBytes;            // its node name

    // node standing for the segment's first run (the band run's head if it continues from the left)
    const uint32_t first = run_head(o, j, r * TBW * kNodeBytes);

    so[r + 1][j + 1] = o;
    sb[r + 1][j + 1] = CONN4 ? a : b;   // the row the band below touches
    sh[r + 1][j + 1] = first;
    if (j == 0) so[r + 1][0] = sb[r + 1][0] = 0;
    if (j == LPB - 1) so[r + 1][LPB + 1] = sb[r + 1][LPB + 1] = 0;
    if (r == 0) {
        so[0][j + 1] = sb[0][j + 1] = 0;
        if (j == 0) so[0][0] = sb[0][0] = 0;
        if (j == LPB - 1) so[0][LPB + 1] = sb[0][LPB + 1] = 0;
    }
    {
        constexpr uint32_t K = kNodeBytes;
        uint4 *p4 = reinterpret_cast<uint4 *>(&par[nb]);
        p4[0] = make_uint4(first, nbn + K, nbn + 2 * K, nbn + 3 * K);
#pragma unroll
        for (uint32_t q = 1; q < LANE_BLOCKS / 4; ++q)
            p4[q] = make_uint4(nbn + 4 * q * K, nbn + (4 * q + 1) * K, nbn + (4 * q + 2) * K, nbn + (4 * q + 3) * K);
    }
    __syncthreads();

    // unions along the vertical contacts with the band above
    {
        const uint32_t ou = so[r][j + 1];
        const uint32_t upn = nbn - TBW * kNodeBytes;   // name of the upper segment's first block
        const SegNodes up{ou & ~(ou << 1), upn, sh[r][j + 1],
                          last_run_node(so[r][j], upn - LANE_BLOCKS * kNodeBytes, sh[r][j]),
                          sh[r][j + 2]};   // only used when a contact exists (r > 0)
        SegContacts cs(a, sb[r][j + 1], sb[r][j], sb[r][j + 2], st, up.st);
        if (CONN4) cs.el = cs.er = 0;   // straight contacts only
        // a contact sharing a node with the lane's previous union starts that
        // side from the node where the previous union met
        uint32_t ua = 0, ub = 0, la = ~0u, lb = ~0u, ltop = 0;
        bool busy = false;
        for (;;) {
            if (!busy) {
                if (!cs.any()) break;
                uint32_t x;
                const int p = cs.pop_any(x);
                const uint32_t s = hibit(st & upto(x));
                const uint32_t na = s ? node_at(nbn, s) : first, nu = up.node(p);
                ua = na == la ? ltop : na;
                ub = nu == lb ? ltop : nu;
                la = na;
                lb = nu;
                busy = true;
            }
            if (lds_rem_step(par, ua, ub, ltop)) busy = false;
        }
    }
    __syncthreads();

    // Roots of this lane's nodes.  A node still pointing at itself is a root
    // (its global index is g0 + block); the others are chased and their entry
    // replaced by the tagged global root.
    const uint32_t g0 = by * bw + bxs;   // global index of local block nb
    uint32_t v[LANE_BLOCKS];
    const auto load_nodes = [&]() {
        const uint4 *p4 = reinterpret_cast<const uint4 *>(&par[nb]);
#pragma unroll
        for (uint32_t q = 0; q < LANE_BLOCKS / 4; ++q) {
            const uint4 x = p4[q];
            v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
        }
    };
    load_nodes();
    uint32_t pend = 0;
#pragma unroll
    for (uint32_t kk = 0; kk < LANE_BLOCKS; ++kk)
        if ((st & (3u << (2 * kk))) && v[kk] != nbn + kk * kNodeBytes) pend |= 1u << kk;
    if (pend) {
        resolve_nodes(par, nbn, pend, [&](uint32_t l) { return (by0 + (l >> LOG_TBW)) * bw + bx0 + (l & (TBW - 1)); });
        load_nodes();
    }

    // block words: root << 4 | mask (background: own index << 4); the tile's
    // components are counted at their roots (the merge subtracts its links)
    // every lane computes its words (selects only: no per-block branch; a
    // lane outside the image has no pixels, so no nodes and no roots, and
    // stores nothing).  A background block (mask 0) carries the parent of the
    // run before it (or 0), <= its own index: no reader uses the parent of a
    // block whose mask is 0.
    const bool inside = by < bh && bxs < bw;
    uint32_t out[LANE_BLOCKS], cur = 0;
#pragma unroll
    for (uint32_t kk = 0; kk < LANE_BLOCKS; ++kk) {
        const uint32_t m = ((a >> (2 * kk)) & 3u) | (((b >> (2 * kk)) & 3u) << 2);
        const bool node = (st & (3u << (2 * kk))) != 0;
        const uint32_t nv = (v[kk] & kResolved) ? v[kk] : g0 + kk;
        cur = node ? nv : cur;
        out[kk] = (cur << 4) | m;   // << 4 drops the tag bit
    }
    // the tile-local roots: the nodes no chase started from (their entries
    // still hold their own index; resolution writes only the chased nodes')
    static_assert(LANE_BLOCKS < 32, "count_add_wave takes per-lane counts below 32");
    const uint32_t rootmask = pair_any16(st) & ~pend;
    const uint32_t roots = __builtin_popcount(rootmask);
    // first-pixel keys start at the root block's own first pixel (global
    // roots are tile-local roots)
    if (G.flags & kKeys)
        for (uint32_t rm = rootmask; rm; rm &= rm - 1) {
            const uint32_t kk = __builtin_ctz(rm);
            const uint32_t m = ((a >> (2 * kk)) & 3u) | (((b >> (2 * kk)) & 3u) << 2);
            G.canon.key[g0 + kk] = block_key(G, by, bxs + kk, m);
        }
    store_block_words<LPB, EXACT>(G.P, (size_t)by * bw + bxs, bxs, inside, j, bw, out);
    if (G.count) count_add_wave(G.slots, roots);   // one image per workgroup: uniform branch
}

// ---- tile-border block jobs (the flatten kernel) ------------------------------------

template <uint32_t TBW, uint32_t TBH>
struct Border {
    static constexpr uint32_t JOBS = 2 * TBW + 2 * (TBH - 2);

    // j-th border block of a tile: top row, bottom row, then the left and right
    // columns without their corner blocks.
    __device__ static void block(uint32_t j, uint32_t bx0, uint32_t by0, uint32_t &bx, uint32_t &by)
    {
        if (j < TBW) {
            bx = bx0 + j;
            by = by0;
        } else if (j < 2 * TBW) {
            bx = bx0 + (j - TBW);
            by = by0 + TBH - 1;
        } else if (j < 2 * TBW + (TBH - 2)) {
            bx = bx0;
            by = by0 + 1 + (j - 2 * TBW);
        } else {
            bx = bx0 + TBW - 1;
            by = by0 + 1 + (j - 2 * TBW - (TBH - 2));
        }
    }
};

// Border job of thread gid: its image, block and packed word (valid = the job
// exists and the block is foreground).
template <uint32_t TBW, uint32_t TBH, class Src>
__device__ __forceinline__ bool border_job(const Src &src, uint32_t gid, uint32_t total_tiles, const Geom2D *&G,
                                           uint32_t &bx, uint32_t &by, uint32_t &bx0, uint32_t &by0, uint32_t &vb)
{
    using B = Border<TBW, TBH>;
    const uint32_t gt = gid / B::JOBS, j = gid % B::JOBS;
    if (gt >= total_tiles) return false;
    uint32_t lt;
    G = &src.by_tile(gt, lt);
    bx0 = (lt % G->tiles_x) * TBW;
    by0 = (lt / G->tiles_x) * TBH;
    B::block(j, bx0, by0, bx, by);
    if (bx >= G->bw || by >= G->bh) return false;
    vb = G->P[by * G->bw + bx];
    return (vb & 15u) != 0;
}

// ---- 2. merge kernel ------------------------------------------------------------
//
// Unions across tile borders, between tile-local roots read from the block
// words (whose mask nibbles give back the pixel rows as bit masks):
//   * top border: one lane per 16-block segment of a tile's top band, against
//     the band above (the last band of the tile above), with the tile
//     kernel's contact logic;
//   * left border: one lane per band of a tile, its first pixel column against
//     the last column of the tile to the left, in the same band and
//     diagonally in the bands above and below.
// Jobs: every tile's top-border lanes, then every tile's left-border lanes.
// Lanes collect pairs of roots in LDS; then every thread of the workgroup
// unites pairs from there (global unions are chains of dependent loads: many
// short chains in flight beat a few long ones).  A single image collects them
// in a hash set, so each distinct pair is united once per workgroup (in dense
// images every border segment of a tile yields the same pair); a batch, whose
// pairs also carry an image, in a queue.

// A single image's workgroup takes MERGE_JOBS border lanes but unites their
// pairs with all MERGE_THREADS threads: the top border of a tile in a dense
// image yields many distinct pairs, and each union is a chain of dependent
// loads, so spreading them over 8 threads per border lane shortens the unions
// done in a row by one thread (8192^2, 256x32-block tiles: merge 28.8 -> 23.3
// us; the 3D merge, whose workgroups dedupe pairs across more bars, does not
// gain from this).  Batches keep one thread per lane: their many small tiles
// give short queues, and 8x the threads cost more than they unite (4096
// images: 310 -> 268 Gpixel/s).
constexpr uint32_t MERGE_THREADS = 256;
constexpr uint32_t MERGE_JOBS = 32;
constexpr uint32_t QUEUE_CAP = 4 * MERGE_THREADS;   // pairs per workgroup (batches; 12 KB of LDS, more
                                                    // workgroups per CU than 2048 pairs: 17.38 -> 17.27 ms)
constexpr uint32_t HASH_CAP = 16 * MERGE_JOBS;      // pair set slots (single image)

template <class Src>
constexpr uint32_t merge_jobs()
{
    return Src::kSingle ? MERGE_JOBS : MERGE_THREADS;
}

// Block words of a 16-block segment starting at block bxs of a row (zero past bw).
__device__ __forceinline__ void load_segment(const uint32_t *row, uint32_t bxs, uint32_t bw, uint32_t v[LANE_BLOCKS])
{
    if (bxs + LANE_BLOCKS <= bw && (reinterpret_cast<uintptr_t>(row + bxs) & 15u) == 0) {
        const uint4 *q = reinterpret_cast<const uint4 *>(row + bxs);
#pragma unroll
        for (uint32_t i = 0; i < LANE_BLOCKS / 4; ++i) {
            const uint4 x = q[i];
            v[4 * i] = x.x; v[4 * i + 1] = x.y; v[4 * i + 2] = x.z; v[4 * i + 3] = x.w;
        }
        return;
    }
#pragma unroll
    for (uint32_t k = 0; k < LANE_BLOCKS; ++k) v[k] = bxs + k < bw ? row[bxs + k] : 0u;
}

// Pixel-row masks of a segment's band from its block words: top row = mask
// bits 0-1 of each block, bottom row = bits 2-3.
__device__ __forceinline__ void band_rows(const uint32_t v[LANE_BLOCKS], uint32_t &top, uint32_t &bot)
{
    top = bot = 0;
#pragma unroll
    for (uint32_t k = 0; k < LANE_BLOCKS; ++k) {
        top |= (v[k] & 3u) << (2 * k);
        bot |= ((v[k] >> 2) & 3u) << (2 * k);
    }
}

template <uint32_t TBW, uint32_t TBH, class Src, bool CONN4>
__global__ __launch_bounds__(MERGE_THREADS) void bleq2d_merge(Src src, uint32_t total_tiles)
{
    constexpr uint32_t LPB = TBW / LANE_BLOCKS;
    constexpr bool kSet = Src::kSingle;
    struct Queue {                 // batches: pairs of roots and their image
        uint2 pair[QUEUE_CAP];
        uint32_t img[QUEUE_CAP];
        uint32_t n;
    };
    // one or the other: the single-image set takes 40 KB, four workgroups per CU
    __shared__ std::conditional_t<kSet, LdsPairSet<HASH_CAP>, Queue> store;
    auto &set = reinterpret_cast<LdsPairSet<HASH_CAP> &>(store);
    auto &queue = reinterpret_cast<Queue &>(store);
    if constexpr (kSet) set.init();
    else if (threadIdx.x == 0) queue.n = 0;
    __syncthreads();

    constexpr uint32_t JOBS = merge_jobs<Src>();
    const uint32_t gid = blockIdx.x * JOBS + threadIdx.x, n_top = total_tiles * LPB;
    const bool job = threadIdx.x < JOBS;
    const Geom2D *G = &src.image(0);   // a valid image for every lane (no null pointer phi)
    uint32_t img = 0, links = 0, la = ~0u, lb = ~0u;
    const auto push = [&](uint32_t ra, uint32_t rb) {
        if (ra == rb || (ra == la && rb == lb)) return;
        la = ra;
        lb = rb;
        if constexpr (kSet) {
            if (set.insert(ra, rb)) return;
        } else {
            const uint32_t i = atomicAdd(&queue.n, 1u);
            if (i < QUEUE_CAP) {
                queue.pair[i] = make_uint2(ra, rb);
                queue.img[i] = img;
                return;
            }
        }
        links += Forest::unite_cached(G->P, ra, rb);   // no room: unite here
    };

    if (job && gid < n_top) {
        const uint32_t gt = gid / LPB, j = gid % LPB;
        uint32_t lt;
        img = src.image_of_tile(gt);
        G = &src.by_tile(gt, lt);
        const uint32_t bw = G->bw, tx = lt % G->tiles_x, ty = lt / G->tiles_x;
        const uint32_t by0 = ty * TBH, bxs = tx * TBW + j * LANE_BLOCKS;
        if (ty > 0 && bxs < bw) {
            const uint32_t *own = G->P + (size_t)by0 * bw + bxs, *up = own - bw;
            // the blocks left and right of the upper segment, loaded unconditionally
            // (the segment's first block where there is none) and ahead of the
            // segments, so that all the loads are in flight at once
            const bool hl = !CONN4 && bxs > 0, hr = !CONN4 && bxs + LANE_BLOCKS < bw;
            uint32_t wl = *(hl ? up - 1 : up), wr = *(hr ? up + LANE_BLOCKS : up);
            uint32_t vo[LANE_BLOCKS], vu[LANE_BLOCKS];
            load_segment(G->P + (size_t)by0 * bw, bxs, bw, vo);
            load_segment(G->P + (size_t)(by0 - 1) * bw, bxs, bw, vu);
            asm volatile("" : "+v"(wl), "+v"(wr));   // no sinking of those loads into their uses
            uint32_t a, b, au, bu;
            band_rows(vo, a, b);
            band_rows(vu, au, bu);
            const uint32_t o = a | b, ou = au | bu;
            const uint32_t bul = hl ? (wl & 8u) << 28 : 0u;   // bottom-right pixel -> bit 31
            const uint32_t bur = hr ? (wr >> 2) & 1u : 0u;    // bottom-left pixel -> bit 0
            // the band above touches with its bottom row (CONN4: its only row)
            SegContacts cs(a, CONN4 ? au : bu, bul, bur, o & ~(o << 1), ou & ~(ou << 1));
            if (CONN4) cs.el = cs.er = 0;
            while (cs.any()) {
                uint32_t x;
                const int p = cs.pop(x);
                push(own[x >> 1] >> 4, up[((p + 2) >> 1) - 1] >> 4);   // block of upper column p (-1 / 32: neighbours)
            }
        }
    } else if (job) {
        const uint32_t g2 = gid - n_top, gt = g2 / TBH, r = g2 % TBH;
        if (gt < total_tiles) {
            uint32_t lt;
            img = src.image_of_tile(gt);
            G = &src.by_tile(gt, lt);
            const uint32_t bw = G->bw, bh = G->bh, tx = lt % G->tiles_x, ty = lt / G->tiles_x;
            const uint32_t bx0 = tx * TBW, by = ty * TBH + r;
            if (tx > 0 && by < bh) {
                // block bits: 1 top-left, 2 top-right, 4 bottom-left, 8 bottom-right
                const uint32_t *w = G->P + (size_t)by * bw + bx0;
                const uint32_t vo = w[0], m = vo & 15u;
                if (CONN4) {
                    if ((m & 1u) && (w[-1] & 2u)) push(vo >> 4, w[-1] >> 4);   // left neighbour pixel
                } else if ((m & 5u) && (w[-1] & 10u)) {
                    push(vo >> 4, w[-1] >> 4);                              // same band
                }
                if (!CONN4 && r > 0 && (m & 1u)) {                           // band above
                    const uint32_t vu = w[-(int64_t)bw - 1];
                    if (vu & 8u) push(vo >> 4, vu >> 4);
                }
                if (!CONN4 && r + 1 < TBH && by + 1 < bh && (m & 4u)) {     // band below
                    const uint32_t vd = w[bw - 1];
                    if (vd & 2u) push(vo >> 4, vd >> 4);
                }
            }
        }
    }
    __syncthreads();
    if constexpr (kSet) {
        const Geom2D &G0 = src.image(0);
        const uint32_t n = set.size();
        for (uint32_t i = threadIdx.x; i < n; i += MERGE_THREADS) {
            const uint2 pr = set.pair(i);
            links += Forest::unite_cached(G0.P, pr.x, pr.y);
        }
        if (G0.count) count_links(G0.slots, links);   // kernel-uniform branch
    } else {
        const uint32_t n = queue.n < QUEUE_CAP ? queue.n : QUEUE_CAP;
        if (links && G->count) atomicSub(&G->slots[(blockIdx.x % kCountSlots) * kCountStride], links);
        for (uint32_t i = threadIdx.x; i < n; i += MERGE_THREADS) {
            const uint2 e = queue.pair[i];
            const Geom2D &Gi = src.image(queue.img[i]);
            if (Forest::unite_cached(Gi.P, e.x, e.y) && Gi.count) atomicSub(&Gi.slots[(blockIdx.x % kCountSlots) * kCountStride], 1u);
        }
    }
}

// Point each distinct tile-local root seen on the tile border (and the border
// block itself) at its final root.
template <uint32_t TBW, uint32_t TBH, class Src>
__global__ __launch_bounds__(256) void bleq2d_flatten(Src src, uint32_t total_tiles)
{
    const Geom2D *G = &src.image(0);   // a valid image for every lane (no null pointer phi)
    uint32_t bx = 0, by = 0, bx0 = 0, by0 = 0, vb = 0;
    const bool valid = border_job<TBW, TBH>(src, blockIdx.x * blockDim.x + threadIdx.x, total_tiles, G, bx, by,
                                            bx0, by0, vb);
    const uint32_t r = vb >> 4;
    uint32_t leader;
    const bool lead = run_leader(r, valid ? G->tile_base : 0u, valid, leader);
    uint32_t root = 0;
    if (lead) {
        uint32_t *P = G->P;
        const uint32_t vr = dev_load(&P[r]);
        root = vr >> 4;
        if (root != r) {
            root = Forest::find_cached(P, r);
            dev_min(&P[r], (root << 4) | (vr & 15u));
        }
    }
    root = __shfl(root, leader, 64);
    if (valid && root != r) {
        const uint32_t b = by * G->bw + bx;
        if (b != r) dev_min(&G->P[b], (root << 4) | (vb & 15u));   // border block one hop from its root
    }
}

// ---- 4. final kernel -------------------------------------------------------------------

// Block words of one lane's FINAL_BLOCKS blocks (zero outside the image).
__device__ __forceinline__ void load_span(const Geom2D &G, const Span &sp, uint32_t v[FINAL_BLOCKS])
{
    const uint32_t *__restrict__ P = G.P;
#pragma unroll
    for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) v[k] = 0;
    if (sp.n == FINAL_BLOCKS) {   // k0 even: 8-byte aligned
        const uint2 q = *reinterpret_cast<const uint2 *>(P + sp.k0);
        v[0] = q.x; v[1] = q.y;
    } else if (sp.n) {
        for (uint32_t k = 0; k < sp.n; ++k) v[k] = P[sp.k0 + k];
    }
}

// Roots of a lane's blocks (where the mask is not empty).  Neighbouring lanes
// mostly share their blocks' parent (a tile-local root): the first lane of each
// run of lanes with the same first parent chases it, the others take its root.
__device__ __forceinline__ void lane_roots(const Geom2D &G, const uint32_t v[FINAL_BLOCKS], uint32_t root[FINAL_BLOCKS])
{
    const uint32_t *__restrict__ P = G.P;
#pragma unroll
    for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) root[k] = 0;
    uint32_t key = 0;
    bool has = false;
#pragma unroll
    for (uint32_t k = FINAL_BLOCKS; k-- > 0;)
        if (v[k] & 15u) {
            key = v[k] >> 4;
            has = true;
        }
    uint32_t leader;
    const bool lead = run_leader(key, 0u, has, leader);
    uint32_t key_root = lead ? Forest::find_final(P, key) : 0u;
    key_root = __shfl(key_root, (int)leader, 64);
    uint32_t prev_parent = key, prev_root = key_root;
#pragma unroll
    for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
        if (!(v[k] & 15u)) continue;
        const uint32_t p = v[k] >> 4;
        root[k] = p == prev_parent ? prev_root : Forest::find_final(P, p);
        prev_parent = p;
        prev_root = root[k];
    }
}

// Label values the final kernel writes: fixed per launch (MODE), or per image
// from its flags (kModeAny, batches).
// kModeFirst: 1 + the component's first pixel index (the label equivalence values).
// kModeBlockRank: 1..N in the order of the components' root blocks (Grana's
//   block-based labeling numbers its components so, cvlabeling_imagelab.cpp:1523-1531).
// kModeRunRank: the number of runs up to and including the component's first
//   run that have no pixel of the row above in reach (the provisional label He
//   et al.'s run labeling keeps for the component, LabelingAlgs.cpp:80-130);
//   the run bits hold those runs' first pixels.
// kModeRunIndex: 1 + first-run row * run_w2 + the index of the first run in its
//   row (the run equivalence values, LabelingAlgs.cpp:573-588,611-691); the run
//   bits hold every run's first pixel.
enum : uint32_t {
    kModeRoot = 0, kModeRef = 1, kModeCanon = 2, kModeFirst = 3, kModeAny = 4,
    kModeBlockRank = 5, kModeRunRank = 6, kModeRunIndex = 7
};

// The label of a component of 2x2 blocks with root block r, for the label
// values MODE (kModeAny: from the image's flags): the reference's raw values
// are 1 + the root's block index in its padded grid (LabelingAlgs.cl:202,317);
// the label equivalence values 1 + the first pixel's index, or with kRefLabels
// the reference's 0-based index in its padded image (DistrInitKernel,
// LabelingAlgs.cl:37).
template <uint32_t MODE>
__device__ __forceinline__ uint32_t root_value(const Geom2D &G, uint32_t r)
{
    const bool ref = MODE == kModeRef || (MODE == kModeAny && (G.flags & kRefLabels));
    const bool canon = MODE == kModeCanon || (MODE == kModeAny && (G.flags & kCanonLabels));
    if (canon) return canon_label(G.canon, G.canon.key[r]);
    if (MODE == kModeBlockRank) return canon_label(G.canon, r);
    if (MODE == kModeRunRank) {   // the first pixel's key, in the run bits' row-padded layout
        const uint32_t k = G.canon.key[r], y = k / G.w;
        return canon_label(G.canon, y * G.w64 + (k - y * G.w));
    }
    if (MODE == kModeFirst) {
        const uint32_t k = G.canon.key[r];
        return (G.flags & kRefLabels) ? (k / G.w) * (2 * G.ref_bw) + k % G.w : k + 1u;
    }
    if (ref) {
        const uint32_t ry = div_by(r, G.bdiv);
        return 1u + (r - ry * G.bw) + ry * G.ref_bw;
    }
    return r + 1u;
}

// Store a lane's four labels v0..v3 of one pixel row, columns x0 .. x0 + 3
// (those < w), row starting at `row`.  A row whose lanes' labels are not
// 16-byte aligned (word offset o = address / 4 mod 4 != 0) is written in
// aligned 4-word windows: each lane stores the window that ends in its own
// labels, its first o words being the previous lane's last labels (shuffled);
// a lane without that previous lane (has_prev: the previous lane of the wave
// holds this row's 4 labels just left of x0) stores its own first 4 - o
// labels one by one, and a lane whose next lane will not store such a window
// stores its own last o labels.  write: the lane has labels in this row.
// Every lane of the wave must call it.
__device__ __forceinline__ void store_row_aligned(uint32_t *row, uint32_t x0, uint32_t w, uint32_t v0, uint32_t v1,
                                                  uint32_t v2, uint32_t v3, bool write, bool has_prev, bool next_in_row)
{
    const uint32_t p1 = lane_prev(v1), p2 = lane_prev(v2), p3 = lane_prev(v3);
    if (!write) return;
    uint32_t *dst = row + x0;   // x0 < w
    const uint32_t o = (uint32_t)((reinterpret_cast<uintptr_t>(dst) >> 2) & 3u), lim = min(4u, w - x0);
    if (o == 0) {
        if (lim == 4)
            store_nt4(dst, v0, v1, v2, v3);
        else
            store_words4(dst, v0, v1, v2, v3, 0, lim);
        return;
    }
    // the window: the previous lane's last o labels, then this lane's first 4 - o
    const uint32_t a0 = by_offset(o, 0, p3, p2, p1);
    const uint32_t a1 = by_offset(o, 0, v0, p3, p2);
    const uint32_t a2 = by_offset(o, 0, v1, v0, p3);
    const uint32_t a3 = by_offset(o, 0, v2, v1, v0);
    if (has_prev && x0 + 4 - o <= w)
        store_nt4(dst - o, a0, a1, a2, a3);
    else
        store_words4(dst, v0, v1, v2, v3, 0, min(4u - o, lim));
    if (!(next_in_row && x0 + 8 - o <= w)) store_words4(dst, v0, v1, v2, v3, 4 - o, lim);
}

// Labels of one lane's blocks of a span, written out.
template <uint32_t MODE, bool CONN4, bool EXACT>
__device__ __forceinline__ void final_span_labels(const Geom2D &G, const Span &sp, const uint32_t v[FINAL_BLOCKS])
{
    const uint32_t h = G.h, w = G.w, bw = G.bw;
    uint32_t root[FINAL_BLOCKS];
    lane_roots(G, v, root);
    uint32_t lab[FINAL_BLOCKS];
#pragma unroll
    for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) lab[k] = 0;
    if (sp.n) {
        const bool canon = MODE == kModeCanon || (MODE == kModeAny && (G.flags & kCanonLabels));
        uint32_t prev_root = ~0u, prev_lab = 0;
#pragma unroll
        for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
            lab[k] = 0;
            if (!(v[k] & 15u)) continue;
            const uint32_t r = root[k];
            if (r != prev_root) {
                prev_root = r;
                // label equivalence values (CONN4): 1 + the first pixel's index, or
                // with kRefLabels the reference's 0-based index in its padded image
                // (width 2 ref_bw; DistrInitKernel, LabelingAlgs.cl:37)
                if (CONN4) {   // the root block (1x2, the component's first) holds its first pixel
                    const uint32_t y = div_by(r, G.bdiv), x = 2 * (r - y * bw) + ((G.P[r] & 1u) ? 0u : 1u);
                    if (MODE == kModeRunRank)
                        prev_lab = canon_label(G.canon, y * G.w64 + x);
                    else if (MODE == kModeRunIndex)   // runs of row y before x: the bits between the row's start and x
                        prev_lab = 1u + y * G.run_w2 + canon_label(G.canon, y * G.w64 + x) -
                                   canon_label(G.canon, y * G.w64);
                    else
                        prev_lab = canon ? canon_label(G.canon, r)
                                         : ((G.flags & kRefLabels) ? y * (2 * G.ref_bw) + x : 1u + y * w + x);
                } else {
                    prev_lab = root_value<MODE>(G, r);
                }
            }
            lab[k] = prev_lab;
        }
    }
    // the lane's blocks in one block row: one 16-byte store per pixel row
    const bool row_path = sp.n && sp.bx + FINAL_BLOCKS <= bw;
    const uint32_t y0 = CONN4 ? sp.by : 2 * sp.by, x0 = 2 * sp.bx;
    uint32_t *row0 = G.labels + (size_t)y0 * w;
    const bool two_rows = !CONN4 && y0 + 1 < h;
    uint32_t top[2 * FINAL_BLOCKS], bot[2 * FINAL_BLOCKS];
#pragma unroll
    for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
        const uint32_t m = v[k] & 15u;
        top[2 * k] = (m & 1u) ? lab[k] : 0u;
        top[2 * k + 1] = (m & 2u) ? lab[k] : 0u;
        bot[2 * k] = (m & 4u) ? lab[k] : 0u;
        bot[2 * k + 1] = (m & 8u) ? lab[k] : 0u;
    }
    if (EXACT || (G.flags & kVecOut)) {   // aligned label rows (uniform)
        if (row_path) {
            const uint32_t lim = min(4u, w - x0);
            if (lim == 4) {
                store_nt4(row0 + x0, top[0], top[1], top[2], top[3]);
                if (two_rows) store_nt4(row0 + w + x0, bot[0], bot[1], bot[2], bot[3]);
            } else {
                store_words4(row0 + x0, top[0], top[1], top[2], top[3], 0, lim);
                if (two_rows) store_words4(row0 + w + x0, bot[0], bot[1], bot[2], bot[3], 0, lim);
            }
        }
    } else {   // aligned windows across lanes (shuffles by every lane of the wave)
        const uint32_t lane = threadIdx.x & 63u;
        const bool has_prev = lane != 0 && sp.bx >= FINAL_BLOCKS;                // the previous lane: this row, just left
        const bool next_in_row = lane != 63 && sp.bx + 2 * FINAL_BLOCKS <= bw;   // the next lane: this row, just right
        store_row_aligned(row0, x0, w, top[0], top[1], top[2], top[3], row_path, has_prev, next_in_row);
        if (!CONN4)
            store_row_aligned(row0 + w, x0, w, bot[0], bot[1], bot[2], bot[3], row_path && two_rows, has_prev, next_in_row);
    }
    if (sp.n && !row_path) {
        {   // the blocks run into the next block row (odd bw): pixel by pixel
            for (uint32_t k = 0; k < sp.n; ++k) {
                uint32_t by, bx;
                span_block(sp, k, bw, by, bx);
                const uint32_t m = v[k] & 15u;
                for (uint32_t dy = 0; dy < (CONN4 ? 1u : 2u); ++dy) {
                    const uint32_t y = CONN4 ? by : 2 * by + dy;
                    if (y >= h) continue;
                    for (uint32_t dx = 0; dx < 2; ++dx) {
                        const uint32_t x = 2 * bx + dx;
                        if (x < w) G.labels[(size_t)y * w + x] = ((m >> (2 * dy + dx)) & 1u) ? lab[k] : 0u;
                    }
                }
            }
        }
    }
}

template <class Src, uint32_t MODE, bool CONN4, bool EXACT>
__global__ __launch_bounds__(NT_FINAL) void bleq2d_final(Src src)
{
    constexpr uint32_t CH = final_chunks<Src>();
    uint32_t lw;
    const Geom2D &G = src.by_final(blockIdx.x, lw);
    // every chunk's block words are loaded before the first root chase
    Span sp[CH];
    uint32_t v[CH][FINAL_BLOCKS];
#pragma unroll
    for (uint32_t c = 0; c < CH; ++c) {
        sp[c] = final_span(G, lw * CH + c);
        load_span(G, sp[c], v[c]);
    }
#pragma unroll
    for (uint32_t c = 0; c < CH; ++c) final_span_labels<MODE, CONN4, EXACT>(G, sp[c], v[c]);
    // the tile and merge kernels have finished: the image's first workgroup sums its count
    if (lw == 0 && G.count && threadIdx.x < kCountSlots) count_reduce(G.slots, G.count);
}

// ---- first-pixel keys; canonical bits -------------------------------------------------
//
// A component's first pixel lies in its topmost block row, which is the block
// row of its root (the minimum block index); each block of that row in the
// component offers its own first pixel (its lowest mask bit), and the smallest
// offer is the component's key: the pixel index (PIXEL_KEY; 1 + that is the
// label equivalence value), or first-pixel row * bw + block column, the bit
// position of canonical numbering.  The keys kernel also clears the key bits
// (canon.nwords words); the mark kernel then sets the bit of every root's key.

// One lane's offers for a span of blocks (see bleq2d_keys).
__device__ __forceinline__ void keys_span(const Geom2D &G, const Span &sp)
{
    const Canon &c = G.canon;
    const uint32_t bw = G.bw;
    uint32_t v[FINAL_BLOCKS] = {};
    for (uint32_t k = 0; k < sp.n; ++k) v[k] = G.P[sp.k0 + k];
    // this lane's offers: (root, key) of its first candidate block, and of a
    // second one with another root.  Roots are never above parents (parent <=
    // self), so only a block whose parent lies in this block row can have its
    // root here: the others need no chase.  A root's key starts at its own
    // block's key, so only offers below the current key are made.
    uint32_t r0 = ~0u, k0 = ~0u, r1 = ~0u, k1 = ~0u, prev_p = ~0u, prev_r = 0, prev_k = 0;
#pragma unroll
    for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
        uint32_t by, bx;
        span_block(sp, k, bw, by, bx);
        const uint32_t row0 = by * bw;
        const uint32_t m = v[k] & 15u, p = v[k] >> 4;
        if (!m || p < row0) continue;
        if (p != prev_p) {
            prev_p = p;
            prev_r = Forest::find_final(G.P, p);
            prev_k = prev_r >= row0 ? c.key[prev_r] : 0u;
        }
        const uint32_t r = prev_r;
        if (r < row0) continue;   // not in its component's top block row
        const uint32_t key = block_key(G, by, bx, m);
        if (key >= prev_k) continue;
        if (r0 == ~0u || r0 == r) {
            r0 = r;
            k0 = min(k0, key);
        } else {
            r1 = r;
            k1 = key;
        }
    }
    // one atomic per wave when every offer of the wave is for one root
    const bool has = r0 != ~0u;
    const uint64_t any = __ballot(has);
    if (!any) return;
    const uint32_t R = __shfl(r0, (int)__builtin_ctzll(any), 64);
    if (__ballot(has && (r0 != R || r1 != ~0u)) == 0) {
        uint32_t kmin = k0;
        for (int off = 32; off > 0; off >>= 1) kmin = min(kmin, (uint32_t)__shfl_xor(kmin, off, 64));
        if ((threadIdx.x & 63u) == (uint32_t)__builtin_ctzll(any)) atomicMin(&c.key[R], kmin);
    } else {
        if (has) atomicMin(&c.key[r0], k0);
        if (r1 != ~0u) atomicMin(&c.key[r1], k1);
    }
}


template <class Src>
__global__ __launch_bounds__(NT_FINAL) void bleq2d_keys(Src src)
{
    constexpr uint32_t CH = final_chunks<Src>();
    uint32_t lw;
    const Geom2D &G = src.by_final(blockIdx.x, lw);
#pragma unroll
    for (uint32_t ch = 0; ch < CH; ++ch) {
        const uint64_t t = (uint64_t)(lw * CH + ch) * NT_FINAL + threadIdx.x;   // covers every word: bw * h / 64 < bw * bh / 2
        if (t < G.canon.nwords) G.canon.bits[t] = 0ull;
        keys_span(G, final_span(G, lw * CH + ch));
    }
}

// CONN4: the key is the root block index itself (a 1x2 root block holds the
// component's first pixel, so block order is first-pixel order).
// Bit k of the result: bit k of x at 2k (even positions), of y at 2k + 1.
__device__ __forceinline__ unsigned long long interleave32(uint32_t x, uint32_t y)
{
    const auto spread = [](unsigned long long v) {
        v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
        v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
        v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
        v = (v | (v << 2)) & 0x3333333333333333ull;
        return (v | (v << 1)) & 0x5555555555555555ull;
    };
    return spread(x) | (spread(y) << 1);
}

// Set the bit of every component's key.  KEY_IS_BLOCK: the key is the root
// block's own index (4-connectivity, block-order ranks), so the 512 blocks of
// a workgroup own bit words 8 c .. 8 c + 7 outright: each wave builds its two
// words from ballots of its lanes' root flags and stores them whole (every
// word is written, so the bits need no clearing first).  Otherwise the key is
// the first-pixel key, which lies in pixel row 2 by or 2 by + 1 of the root's
// block row by, at or right of the root's column: roots of the workgroup's
// first block row OR their bits into two 16-word LDS windows over those rows
// (from the workgroup's first column on), flushed with one device atomicOr per
// non-zero word; other keys go to memory directly.  One device atomic per
// root took 52 us at 8192^2.
template <class Src, bool KEY_IS_BLOCK>
__global__ __launch_bounds__(NT_FINAL) void bleq2d_mark(Src src)
{
    static_assert(final_chunks<Src>() == 1 && FINAL_BLOCKS == 2 && NT_FINAL * FINAL_BLOCKS == 512,
                  "one 512-block span per workgroup, two blocks per thread");
    constexpr uint32_t WIN = 16;   // words per LDS window
    __shared__ unsigned long long win[2 * WIN];
    uint32_t lw;
    const Geom2D &G = src.by_final(blockIdx.x, lw);
    const Span sp = final_span(G, lw);
    bool root[FINAL_BLOCKS];
#pragma unroll
    for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
        const uint32_t b = sp.k0 + k;
        const uint32_t vb = k < sp.n ? G.P[b] : 0u;
        root[k] = (vb & 15u) && (vb >> 4) == b;
    }
    if constexpr (KEY_IS_BLOCK) {
        const unsigned long long r0 = __ballot(root[0]), r1 = __ballot(root[1]);
        const uint32_t lane = threadIdx.x & 63u;
        if ((lane & 31u) == 0) {
            const uint32_t word = (sp.k0 >> 6);   // lane 0: blocks 128 v .. + 63, lane 32: the next 64
            const uint32_t sh = lane;             // this half's 32 lanes in the ballots
            if (word < G.canon.nwords)
                G.canon.bits[word] = interleave32((uint32_t)(r0 >> sh), (uint32_t)(r1 >> sh));
        }
    } else {
        const uint32_t bw = G.bw, k_wg = lw * (NT_FINAL * FINAL_BLOCKS);
        const uint32_t by0 = div_by(k_wg, G.bdiv), bx0 = k_wg - by0 * bw;
        const uint32_t w0 = (uint32_t)(((uint64_t)(2 * by0) * bw + bx0) >> 6);
        const uint32_t w1 = (uint32_t)(((uint64_t)(2 * by0 + 1) * bw + bx0) >> 6);
        if (threadIdx.x < 2 * WIN) win[threadIdx.x] = 0ull;
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
            if (!root[k]) continue;
            const uint32_t key = G.canon.key[sp.k0 + k], word = key >> 6;
            const unsigned long long bit = 1ull << (key & 63u);
            if (word - w0 < WIN) atomicOr(&win[word - w0], bit);
            else if (word - w1 < WIN) atomicOr(&win[WIN + word - w1], bit);
            else if (word < G.canon.nwords) atomicOr(&G.canon.bits[word], bit);
        }
        __syncthreads();
        if (threadIdx.x < 2 * WIN) {
            const unsigned long long v = win[threadIdx.x];
            const uint32_t word = threadIdx.x < WIN ? w0 + threadIdx.x : w1 + threadIdx.x - WIN;
            if (v && word < G.canon.nwords) atomicOr(&G.canon.bits[word], v);
        }
    }
}

// The canonical scan (canon.hpp) over every image's chunks: workgroup = chunk.
template <class Src>
__global__ __launch_bounds__(kCanonThreads) void bleq2d_canon_clear(Src src)
{
    uint32_t lc;
    const Geom2D &G = src.by_chunk(blockIdx.x, lc);
    canon_clear_chunk(G.canon, lc);
}

template <class Src>
__global__ __launch_bounds__(kCanonThreads) void bleq2d_canon_partial(Src src)
{
    uint32_t lc;
    const Geom2D &G = src.by_chunk(blockIdx.x, lc);
    canon_partial_chunk(G.canon, lc);
}

template <class Src>
__global__ __launch_bounds__(kCanonThreads) void bleq2d_canon_prefix(Src src)
{
    uint32_t lc;
    const Geom2D &G = src.by_chunk(blockIdx.x, lc);
    canon_prefix_chunk(G.canon, lc);
}

// ---- run bits: the He et al. and run equivalence values ----------------------------------
//
// One wave per pixel row y sets bit x of 64-bit word y * w64 / 64 + x / 64 of
// the canonical bit array (rows padded to whole words, so every word belongs
// to one lane and no atomics are needed) where a run of row y starts at column
// x.  LC: only the runs with no pixel of row y - 1 in columns start - conpix ..
// end + conpix, the runs for which He et al.'s run labeling opens a new
// provisional label (TRunLabeling::SetRunLabel, LabelingAlgs.cpp:80-130).  The
// canonical scan over these bits then ranks a component's first run among them.
//
// "Touched" (a pixel of the run has a pixel of the row above in reach) is found
// bit-parallel: contact bits are filled towards bit 0 through each run's set
// bits by log-step shifts, and a run that continues into the next words takes
// their contacts by ballots over the wave (lanes = consecutive words); the
// row's words are walked in chunks of 64 from the right, the touched state of
// a run entering a chunk from its right carried over.

__device__ __forceinline__ unsigned long long row_bits64(const uint8_t *row, uint32_t x0, uint32_t w, bool vec)
{
    const unsigned long long lo = row_bits32(row, x0, w, vec);
    const unsigned long long hi = x0 + 32 < w ? row_bits32(row, x0 + 32, w, vec) : 0ull;
    return lo | (hi << 32);
}

template <bool LC>
__global__ __launch_bounds__(64) void runs2d_bits(Geom2D G, uint32_t conpix)
{
    const uint32_t y = blockIdx.x, lane = threadIdx.x, w = G.w, nw = G.w64 / 64;
    const bool vec = (G.flags & kVecIn) != 0;
    const uint8_t *row = G.px + (size_t)y * G.pitch;
    const uint8_t *up = row - G.pitch;   // read only when y > 0
    unsigned long long *out = G.canon.bits + (size_t)y * nw;
    bool carry = false, carry_b0 = false;   // the run entering the current chunk from its right: touched / exists
    for (uint32_t c = (nw + 63) / 64; c-- > 0;) {
        const uint32_t j = c * 64 + lane, x0 = 64 * j;
        unsigned long long r = 0, u = 0;
        uint32_t prev = 0, ul = 0, ur = 0;
        if (j < nw) {
            r = row_bits64(row, x0, w, vec);
            prev = x0 > 0 && row[x0 - 1] != 0;
            if (LC && y > 0) {
                u = row_bits64(up, x0, w, vec);
                ul = x0 > 0 && up[x0 - 1] != 0;
                ur = x0 + 64 < w && up[x0 + 64] != 0;
            }
        }
        const unsigned long long starts = r & ~((r << 1) | prev);
        unsigned long long res = starts;
        if (LC) {
            unsigned long long reach = u;
            if (conpix) reach |= (u << 1) | ul | (u >> 1) | ((unsigned long long)ur << 63);
            // m: bits of r with a contact at or above them in their run (within the word)
            unsigned long long m = r & reach, p = r;
#pragma unroll
            for (uint32_t sh = 1; sh < 64; sh <<= 1) {
                m |= (m >> sh) & p;
                p &= p >> sh;
            }
            const bool full = r == ~0ull, b0 = (r & 1ull) != 0, m0 = (m & 1ull) != 0;
            const unsigned long long F = __ballot(full), M = __ballot(m0), B0 = __ballot(b0);
            const unsigned long long above = lane < 63 ? ~0ull << (lane + 1) : 0ull;
            const bool next_b0 = lane < 63 ? ((B0 >> (lane + 1)) & 1ull) != 0 : carry_b0;
            // the run holding bit 63 continues into the next word: its part there,
            // up to the first word that is not all foreground, decides
            bool t_in = false;
            if ((r >> 63) && next_b0) {
                const unsigned long long nf = ~F & above;
                const uint32_t k = nf ? (uint32_t)__builtin_ctzll(nf) : 64u;
                const unsigned long long span = (k < 64 ? ((2ull << k) - 1ull) : ~0ull) & above;
                t_in = (M & span) != 0 || (k == 64 && carry);
            }
            // the run holding bit 63 starts at the highest start bit, if it starts here
            const unsigned long long top = (r >> 63) && starts ? 1ull << (63 - __builtin_clzll(starts)) : 0ull;
            res = starts & ~m & ~(t_in ? top : 0ull);
            // for the chunk to the left: lane 0's run at bit 0, touched (with its continuation) / exists
            const bool bottom = m0 || (full && t_in);
            carry = __shfl((int)bottom, 0, 64) != 0;
            carry_b0 = __shfl((int)b0, 0, 64) != 0;
        }
        if (j < nw) out[j] = res;
    }
}

// ---- host side --------------------------------------------------------------------

template <uint32_t TBW, uint32_t TBH>
struct Variant {
    static constexpr uint32_t THREADS = (TBW / LANE_BLOCKS) * TBH;

    template <class Src, bool CONN4, bool EXACT>
    static void launch_final(const Src &src, uint32_t total_final, hipStream_t s, uint32_t mode)
    {
        const dim3 grid(total_final), block(NT_FINAL);
        switch (mode) {
        case kModeRoot: hipLaunchKernelGGL((bleq2d_final<Src, kModeRoot, CONN4, EXACT>), grid, block, 0, s, src); break;
        case kModeRef: hipLaunchKernelGGL((bleq2d_final<Src, kModeRef, CONN4, EXACT>), grid, block, 0, s, src); break;
        case kModeCanon: hipLaunchKernelGGL((bleq2d_final<Src, kModeCanon, CONN4, EXACT>), grid, block, 0, s, src); break;
        case kModeFirst: hipLaunchKernelGGL((bleq2d_final<Src, kModeFirst, CONN4, EXACT>), grid, block, 0, s, src); break;
        // the values of the other registry methods: single images only
        case kModeBlockRank:
            if constexpr (Src::kSingle && !CONN4)
                hipLaunchKernelGGL((bleq2d_final<Src, kModeBlockRank, CONN4, EXACT>), grid, block, 0, s, src);
            break;
        case kModeRunRank:
            if constexpr (Src::kSingle)
                hipLaunchKernelGGL((bleq2d_final<Src, kModeRunRank, CONN4, EXACT>), grid, block, 0, s, src);
            break;
        case kModeRunIndex:
            if constexpr (Src::kSingle && CONN4)
                hipLaunchKernelGGL((bleq2d_final<Src, kModeRunIndex, CONN4, EXACT>), grid, block, 0, s, src);
            break;
        default: hipLaunchKernelGGL((bleq2d_final<Src, kModeAny, CONN4, EXACT>), grid, block, 0, s, src); break;
        }
    }

    template <class Src, bool CONN4 = false>
    static hipError_t launch(const Src &src, uint32_t total_tiles, uint32_t total_final, hipStream_t s,
                             KernelTimer *kt, bool flatten, uint32_t canon_chunks, uint32_t mode)
    {
        // run bits (he-run / runeq values) and block-order ranks (gr-block) use
        // the canonical scan over their own bits; canon: first-pixel canonical
        // numbering (some image asks for canonical labels)
        const bool run_bits = mode == kModeRunRank || mode == kModeRunIndex;
        const bool block_rank = mode == kModeBlockRank;
        const bool canon = canon_chunks != 0 && !run_bits && !block_rank;
        const uint32_t merge_grid =
            (uint32_t)(((uint64_t)total_tiles * (TBW / LANE_BLOCKS + TBH) + merge_jobs<Src>() - 1) / merge_jobs<Src>());
        const uint32_t border_grid = (uint32_t)(((uint64_t)total_tiles * Border<TBW, TBH>::JOBS + 255) / 256);
        bool exact = false;
        if constexpr (Src::kSingle) exact = (src.g.flags & kExact) != 0;
        if constexpr (Src::kSingle) {
            if (exact)
                hipLaunchKernelGGL((bleq2d_tile<TBW, TBH, Src, CONN4, true>), dim3(total_tiles), dim3(THREADS), 0, s, src);
            else
                hipLaunchKernelGGL((bleq2d_tile<TBW, TBH, Src, CONN4, false>), dim3(total_tiles), dim3(THREADS), 0, s, src);
        } else {
            hipLaunchKernelGGL((bleq2d_tile<TBW, TBH, Src, CONN4, false>), dim3(total_tiles), dim3(THREADS), 0, s, src);
        }
        if (kt) kt->mark(s, "bleq2d_tile");
        hipLaunchKernelGGL((bleq2d_merge<TBW, TBH, Src, CONN4>), dim3(merge_grid), dim3(MERGE_THREADS), 0, s, src,
                           total_tiles);
        if (kt) kt->mark(s, "bleq2d_merge");
        if (flatten) {
            hipLaunchKernelGGL((bleq2d_flatten<TBW, TBH, Src>), dim3(border_grid), dim3(256), 0, s, src, total_tiles);
            if (kt) kt->mark(s, "bleq2d_flatten");
        }
        if (!CONN4 && (mode == kModeFirst || mode == kModeRunRank || canon)) {
            hipLaunchKernelGGL((bleq2d_keys<Src>), dim3(total_final), dim3(NT_FINAL), 0, s, src);   // clears the bits
            if (kt) kt->mark(s, "bleq2d_keys");
        }
        if (canon) {
            // 4-connectivity keys are root blocks: the mark kernel writes every bit word
            hipLaunchKernelGGL((bleq2d_mark<Src, CONN4>), dim3(total_final), dim3(NT_FINAL), 0, s, src);
            if (kt) kt->mark(s, "bleq2d_mark");
            hipLaunchKernelGGL((bleq2d_canon_partial<Src>), dim3(canon_chunks), dim3(kCanonThreads), 0, s, src);
            if (kt) kt->mark(s, "canon_partial");
            hipLaunchKernelGGL((bleq2d_canon_prefix<Src>), dim3(canon_chunks), dim3(kCanonThreads), 0, s, src);
            if (kt) kt->mark(s, "canon_prefix");
        }
        if (block_rank || run_bits) {
            if (block_rank) {   // one bit per root, at the root block's own index (every word written)
                hipLaunchKernelGGL((bleq2d_mark<Src, true>), dim3(total_final), dim3(NT_FINAL), 0, s, src);
                if (kt) kt->mark(s, "bleq2d_mark");
            } else if constexpr (Src::kSingle) {   // every bit word is written
                const uint32_t conpix = CONN4 ? 0u : 1u;   // he-run 8-connectivity runs on 2x2 blocks
                if (mode == kModeRunRank)
                    hipLaunchKernelGGL(runs2d_bits<true>, dim3(src.g.h), dim3(64), 0, s, src.g, conpix);
                else
                    hipLaunchKernelGGL(runs2d_bits<false>, dim3(src.g.h), dim3(64), 0, s, src.g, conpix);
                if (kt) kt->mark(s, "runs2d_bits");
            }
            hipLaunchKernelGGL((bleq2d_canon_partial<Src>), dim3(canon_chunks), dim3(kCanonThreads), 0, s, src);
            if (kt) kt->mark(s, "canon_partial");
            hipLaunchKernelGGL((bleq2d_canon_prefix<Src>), dim3(canon_chunks), dim3(kCanonThreads), 0, s, src);
            if (kt) kt->mark(s, "canon_prefix");
        }
        if constexpr (Src::kSingle) {
            if (exact)
                launch_final<Src, CONN4, true>(src, total_final, s, mode);
            else
                launch_final<Src, CONN4, false>(src, total_final, s, mode);
        } else {
            launch_final<Src, CONN4, false>(src, total_final, s, mode);
        }
        if (kt) kt->mark(s, "bleq2d_final");
        return hipGetLastError();
    }
};

// canon_chunks: canonical scan chunks of the images that ask for canonical
// labels (0: none); mode: the label values (kMode*)
template <class Src>
hipError_t dispatch(uint32_t tile_id, const Src &src, uint32_t total_tiles, uint32_t total_final, hipStream_t s,
                    KernelTimer *kt, bool flatten, uint32_t canon_chunks, uint32_t mode)
{
    switch (tile_id) {
    case 1: return Variant<64, 64>::launch(src, total_tiles, total_final, s, kt, flatten, canon_chunks, mode);
    case 2: return Variant<128, 64>::launch(src, total_tiles, total_final, s, kt, flatten, canon_chunks, mode);
    case 3: return Variant<256, 32>::launch(src, total_tiles, total_final, s, kt, flatten, canon_chunks, mode);
    case 4: return Variant<256, 64>::launch(src, total_tiles, total_final, s, kt, flatten, canon_chunks, mode);
    case 5: return Variant<512, 32>::launch(src, total_tiles, total_final, s, kt, flatten, canon_chunks, mode);
    case 6: return Variant<128, 128>::launch(src, total_tiles, total_final, s, kt, flatten, canon_chunks, mode);
    default: return Variant<128, 32>::launch(src, total_tiles, total_final, s, kt, flatten, canon_chunks, mode);
    }
}

// the instantiations above follow the planner's tile table
constexpr bool tile_is(uint32_t id, uint32_t bw, uint32_t bh) { return plan::kTiles[id].bw == bw && plan::kTiles[id].bh == bh; }
static_assert(tile_is(0, 128, 32) && tile_is(1, 64, 64) && tile_is(2, 128, 64) && tile_is(3, 256, 32) &&
                  tile_is(4, 256, 64) && tile_is(5, 512, 32) && tile_is(6, 128, 128),
              "dispatch() and plan::kTiles disagree");
static_assert(NT_FINAL * FINAL_BLOCKS == plan::kFinalWgBlocks && kCountBytes == plan::kCountBytes &&
                  kCanonChunkWords == plan::kCanonChunkWords,
              "kernel constants and the planner disagree");

}  // namespace

void bleq2d_geometry(Geom2D &g, uint32_t tile_id, bool batch)
{
    const plan::Geometry pg = plan::geometry2d(g.h, g.w, tile_id, batch ? plan::kBatchFinalChunks : 1u);
    g.bw = pg.bw;
    g.bh = pg.bh;
    g.bdiv = plan::div_magic(g.bw);
    g.tiles_x = pg.tiles_x;
    g.ntiles = pg.ntiles;
    g.nfinal = pg.nfinal;
    g.ref_bw = (uint32_t)((((uint64_t)g.w >> 5) << 5) + 32) / 2;   // LabelingTools.cpp:103
    g.flags &= kRefLabels | kKeys | kPixelKeys | kCanonLabels | kDropOrigin;
    if (g.pitch % 16 == 0 && reinterpret_cast<uintptr_t>(g.px) % 16 == 0) g.flags |= kVecIn;
    if (g.w % 4 == 0 && reinterpret_cast<uintptr_t>(g.labels) % 16 == 0) g.flags |= kVecOut;
    if ((g.flags & kVecIn) && (g.flags & kVecOut) && g.w % 32 == 0) g.flags |= kExact;
}

size_t bleq2d_workspace(uint32_t h, uint32_t w, bool canonical)
{
    const uint64_t bw = (w + 1ull) / 2, bh = (h + 1ull) / 2;
    return align256(kCountBytes) + align256(bw * bh * sizeof(uint32_t)) +
           (canonical ? canon_workspace((uint64_t)h * bw, bw * bh) : 0);
}

hipError_t bleq2d_run(const Image2D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt, bool first)
{
    One one;
    Geom2D &g = one.g;
    g = Geom2D{};
    g.px = im.px;
    g.pitch = im.pitch;
    g.h = im.h;
    g.w = im.w;
    g.labels = im.labels;
    g.count = im.count;
    g.slots = static_cast<uint32_t *>(ws);   // count slots first: zero between runs
    g.P = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(ws) + align256(kCountBytes));
    g.flags = opt.mode == LBL_LABELS_REFERENCE ? (first ? kRefLabels | kDropOrigin : kRefLabels) : 0u;
    bleq2d_geometry(g, opt.tile2d);
    const bool canon = opt.mode == LBL_LABELS_CANONICAL;
    if (canon || first) {
        g.canon = canon_layout(reinterpret_cast<uint8_t *>(g.P) + align256((size_t)g.bw * g.bh * sizeof(uint32_t)),
                               (uint64_t)g.h * g.bw, (uint64_t)g.bw * g.bh);
        if (!canon) g.canon.nwords = 0;   // keys only
        // the tile kernel initialises the keys of roots; canonical numbering uses
        // bit-position keys for both label sets (the same components)
        g.flags |= canon ? kKeys : (kKeys | kPixelKeys);
    }
    if (kt) kt->begin(s);
    const uint32_t mode = canon ? kModeCanon : (first ? kModeFirst : ((g.flags & kRefLabels) ? kModeRef : kModeRoot));
    return dispatch(opt.tile2d, one, g.ntiles, g.nfinal, s, kt, opt.flatten, canon ? canon_chunks(g.canon.nwords) : 0u,
                    mode);
}

size_t bleq2d_conn4_workspace(uint32_t h, uint32_t w, bool canonical)
{
    const uint64_t bw = (w + 1ull) / 2;
    return align256(kCountBytes) + align256(bw * h * sizeof(uint32_t)) + (canonical ? canon_workspace(bw * h, 0) : 0);
}

// Label equivalence with 4-connectivity on 1x2-pixel blocks: labels 1 + the
// component's first pixel index (or canonical numbering).  The minimum block
// of a component is its first block in raster order, on the component's first
// row, so it holds the first pixel.
hipError_t bleq2d_run_conn4(const Image2D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt)
{
    constexpr uint32_t TBW = 128, TBH = 64;
    One one;
    Geom2D &g = one.g;
    g = Geom2D{};
    g.px = im.px;
    g.pitch = im.pitch;
    g.h = im.h;
    g.w = im.w;
    g.labels = im.labels;
    g.count = im.count;
    g.slots = static_cast<uint32_t *>(ws);
    g.P = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(ws) + align256(kCountBytes));
    g.flags = opt.mode == LBL_LABELS_REFERENCE ? kRefLabels | kDropOrigin : 0u;
    bleq2d_geometry(g, 2);   // 128 x 64 blocks, then one pixel row per block row:
    g.bh = g.h;
    g.tiles_x = (g.bw + TBW - 1) / TBW;
    g.ntiles = g.tiles_x * ((g.bh + TBH - 1) / TBH);
    g.nfinal = (uint32_t)(((uint64_t)g.bw * g.bh + NT_FINAL * FINAL_BLOCKS - 1) / (NT_FINAL * FINAL_BLOCKS));
    const bool canon = opt.mode == LBL_LABELS_CANONICAL;
    if (canon) g.canon = canon_layout(reinterpret_cast<uint8_t *>(g.P) + align256((size_t)g.bw * g.bh * sizeof(uint32_t)),
                                      (uint64_t)g.bw * g.bh, 0);
    if (kt) kt->begin(s);
    return Variant<TBW, TBH>::launch<One, true>(one, g.ntiles, g.nfinal, s, kt, false,
                                                canon ? canon_chunks(g.canon.nwords) : 0u, canon ? kModeCanon : kModeFirst);
}

// ---- the other registry methods' label values, on the same block forests ------------

size_t grblock2d_workspace(uint32_t h, uint32_t w) { return bleq2d_workspace(h, w, true); }

// Grana's block-based labeling ("gr-block", TBlockGranaLabeling, LabelingAlgs.cpp:47-56):
// 8-connectivity, components numbered 1..N in the order of their first 2x2
// block (the root block): a bit per root at its block index, ranked by the
// canonical scan.  LBL_LABELS_CANONICAL: first-pixel order instead.
hipError_t grblock2d_run(const Image2D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt)
{
    if (opt.mode == LBL_LABELS_CANONICAL) return bleq2d_run(im, opt, ws, s, kt);
    One one;
    Geom2D &g = one.g;
    g = Geom2D{};
    g.px = im.px;
    g.pitch = im.pitch;
    g.h = im.h;
    g.w = im.w;
    g.labels = im.labels;
    g.count = im.count;
    g.slots = static_cast<uint32_t *>(ws);
    g.P = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(ws) + align256(kCountBytes));
    bleq2d_geometry(g, opt.tile2d);
    g.canon = canon_layout(reinterpret_cast<uint8_t *>(g.P) + align256((size_t)g.bw * g.bh * sizeof(uint32_t)),
                           (uint64_t)g.bw * g.bh, 0);
    if (kt) kt->begin(s);
    return dispatch(opt.tile2d, one, g.ntiles, g.nfinal, s, kt, opt.flatten, canon_chunks(g.canon.nwords),
                    kModeBlockRank);
}

size_t runs2d_workspace(uint32_t h, uint32_t w, bool conn8, bool canonical)
{
    const uint64_t bw = (w + 1ull) / 2, w64 = (w + 63ull) / 64 * 64;
    const uint64_t nodes = conn8 ? bw * ((h + 1ull) / 2) : bw * h;
    const size_t own = align256(kCountBytes) + align256(nodes * sizeof(uint32_t)) +
                       canon_workspace((uint64_t)h * w64, conn8 ? nodes : 0);
    if (!canonical) return own;
    return std::max(own, conn8 ? bleq2d_workspace(h, w, true) : bleq2d_conn4_workspace(h, w, true));
}

// He et al.'s run labeling ("he-run", TRunLabeling, LabelingAlgs.cpp:58-205;
// he = true) and the run equivalence labeling ("runeq", TRunEqivLabeling /
// TOCLRunEquivLabeling, LabelingAlgs.cpp:535-810,1090-1268; he = false).
// Components: he-run 4-connectivity on 1x2 blocks, 8 on 2x2 blocks with
// first-pixel keys; runeq joins runs of adjacent rows that share a column
// (IsNeib, LabelingAlgs.cpp:605-613), which is 4-connectivity.  Values from the
// run bits (runs2d_bits): he-run the rank of the component's first run among
// the runs that open a provisional label; runeq 1 + row * run_w2 + the index
// of the first run in its row, run_w2 = ceil(w / 2) (the reference CPU's
// w >> 1 run slots per row, for even w) or, with LBL_LABELS_REFERENCE, half the
// reference's padded width (the GPU run slots, LabelingAlgs.cpp:1253).
// LBL_LABELS_CANONICAL: first-pixel order on the same forests.
hipError_t runs2d_run(const Image2D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt, bool he,
                      int conn)
{
    const bool conn8 = he && conn == 8;
    if (opt.mode == LBL_LABELS_CANONICAL)
        return conn8 ? bleq2d_run(im, opt, ws, s, kt) : bleq2d_run_conn4(im, opt, ws, s, kt);
    One one;
    Geom2D &g = one.g;
    g = Geom2D{};
    g.px = im.px;
    g.pitch = im.pitch;
    g.h = im.h;
    g.w = im.w;
    g.labels = im.labels;
    g.count = im.count;
    g.slots = static_cast<uint32_t *>(ws);
    g.P = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(ws) + align256(kCountBytes));
    g.flags = conn8 ? kKeys | kPixelKeys : 0u;
    const uint32_t tile = conn8 ? opt.tile2d : 2u;   // 1x2 blocks: the label equivalence tile (below)
    bleq2d_geometry(g, tile);
    if (!conn8) {   // 1x2 blocks, one pixel row per block row (as bleq2d_run_conn4)
        g.bh = g.h;
        g.tiles_x = (g.bw + 128 - 1) / 128;
        g.ntiles = g.tiles_x * ((g.bh + 64 - 1) / 64);
        g.nfinal = (uint32_t)(((uint64_t)g.bw * g.bh + NT_FINAL * FINAL_BLOCKS - 1) / (NT_FINAL * FINAL_BLOCKS));
    }
    g.w64 = (g.w + 63u) / 64 * 64;
    g.run_w2 = opt.mode == LBL_LABELS_REFERENCE ? (uint32_t)(ref_pad2(g.w) / 2) : (g.w + 1u) / 2;
    const uint64_t nodes = (uint64_t)g.bw * g.bh;
    g.canon = canon_layout(reinterpret_cast<uint8_t *>(g.P) + align256(nodes * sizeof(uint32_t)),
                           (uint64_t)g.h * g.w64, conn8 ? nodes : 0);
    const uint32_t mode = he ? kModeRunRank : kModeRunIndex, chunks = canon_chunks(g.canon.nwords);
    if (kt) kt->begin(s);
    if (conn8) return dispatch(opt.tile2d, one, g.ntiles, g.nfinal, s, kt, opt.flatten, chunks, mode);
    return Variant<128, 64>::launch<One, true>(one, g.ntiles, g.nfinal, s, kt, false, chunks, mode);
}

hipError_t bleq2d_run_batch(const Geom2D *d_geoms, uint32_t n, const BatchTables &tables, uint32_t total_tiles,
                            uint32_t total_final, uint32_t total_chunks, uint32_t tile_id, bool flatten, hipStream_t s,
                            KernelTimer *kt)
{
    Many many{d_geoms, n, tables};
    if (kt) kt->begin(s);
    return dispatch(tile_id, many, total_tiles, total_final, s, kt, flatten, total_chunks, kModeAny);
}

}  // namespace lbl
