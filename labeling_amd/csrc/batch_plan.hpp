// batch_plan.hpp -- the host plan of a batch of 2D images: per image its block
// grid, tile grid, final-kernel workgroups and canonical scan chunks, its slice
// of the shared workspace, and the device tables giving the image of every
// tile, final workgroup and scan chunk id the batched kernels launch.
//
// Pure host code with no HIP dependency, so that the CPU tests can check the
// plan on its own (tests/cpp/batch_plan_check.cpp): every id a kernel can
// compute maps to an image inside its table, each image's ids are exactly
// its own, and the workspace slices fit.  The single-image path uses the same
// geometry (bleq2d_geometry calls geometry2d).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace lbl {
namespace plan {

struct TileShape {
    uint32_t bw, bh;   // tile size in 2x2 blocks
};

// The 2D tile shapes, by tile id (bleq2d.hip instantiates its kernels for each).
constexpr uint32_t kTileCount = 7;
constexpr TileShape kTiles[kTileCount] = {{128, 32}, {64, 64}, {128, 64}, {256, 32}, {256, 64}, {512, 32}, {128, 128}};

// Tile shape of a single image with no tuning-table entry for its size class,
// from a sweep of every shape on MI355X (square images of 2^18 .. 2^27 pixels,
// Bernoulli p = 0.1 .. 0.9 and blobs): 256x32 blocks from 2^23 pixels up (the
// best there, or within 4%), 128x32 below (within 0.5% of the best).  The
// flatten pass is off by default: it cost 3 to 40 us at every size.
constexpr uint32_t default_tile2d(uint64_t pixels) { return pixels >= (1ull << 23) ? 3u : 0u; }

// Final-kernel (and key / mark kernel) workgroups: 256 threads x 2 blocks each
// per chunk, over the image's blocks in raster order; a single image's
// workgroups take one chunk each, a batch's kBatchFinalChunks consecutive
// chunks of their image (so that the image lookup at the start of a workgroup
// is paid once per 2048 blocks).
constexpr uint32_t kFinalThreads = 256;
constexpr uint32_t kFinalBlocksPerThread = 2;
constexpr uint32_t kFinalWgBlocks = kFinalThreads * kFinalBlocksPerThread;
constexpr uint32_t kBatchFinalChunks = 1;

// Count slots at the start of every workspace (64 counters, a 64-byte line each).
constexpr uint64_t kCountBytes = 64 * 16 * 4;

// Canonical numbering: words of 64 key bits per scan workgroup.
constexpr uint32_t kCanonChunkWords = 256 * 8;

inline uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

// Exact unsigned 32-bit division by a divisor d >= 1 fixed at launch, by a
// multiply-high and shifts (Granlund and Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994, figure 4.1): with l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, t = mulhi(m, n),
// n / d = (t + ((n - t) >> s1)) >> s2, s1 = min(l, 1), s2 = max(l - 1, 0).
// The kernels evaluate it on the scalar unit when n is uniform.
struct DivMagic {
    uint32_t m, s1, s2;
};

inline DivMagic div_magic(uint32_t d)
{
    uint32_t l = 0;
    while (l < 32 && (uint64_t(1) << l) < d) ++l;
    DivMagic r;
    r.m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
    r.s1 = l < 1 ? l : 1;
    r.s2 = l > 0 ? l - 1 : 0;
    return r;
}

inline uint32_t div_by(uint32_t n, const DivMagic &g)
{
    const uint32_t t = (uint32_t)(((uint64_t)g.m * n) >> 32);
    return (t + ((n - t) >> g.s1)) >> g.s2;
}

struct Geometry {
    uint32_t bw, bh;        // block grid
    uint32_t tiles_x, ntiles;
    uint32_t nfinal;        // final workgroups
};

// chunks: final-kernel chunks per workgroup (1 for single images,
// kBatchFinalChunks in batches)
inline Geometry geometry2d(uint32_t h, uint32_t w, uint32_t tile_id, uint32_t chunks = 1)
{
    const TileShape ts = kTiles[tile_id < kTileCount ? tile_id : 0];
    Geometry g{};
    g.bw = (w + 1u) / 2;
    g.bh = (h + 1u) / 2;
    g.tiles_x = (g.bw + ts.bw - 1) / ts.bw;
    g.ntiles = g.tiles_x * ((g.bh + ts.bh - 1) / ts.bh);
    const uint64_t per_wg = (uint64_t)kFinalWgBlocks * chunks;
    g.nfinal = (uint32_t)(((uint64_t)g.bw * g.bh + per_wg - 1) / per_wg);
    return g;
}

// Canonical numbering state for nkeys possible keys and nnodes key slots.
inline uint64_t canon_words(uint64_t nkeys) { return (nkeys + 63) / 64; }
inline uint64_t canon_chunk_count(uint64_t nwords) { return (nwords + kCanonChunkWords - 1) / kCanonChunkWords; }
inline uint64_t canon_bytes(uint64_t nkeys, uint64_t nnodes)
{
    const uint64_t nw = canon_words(nkeys);
    return align256(nnodes * 4) + align256(nw * 8) + align256(nw * 4) + align256(canon_chunk_count(nw) * 4);
}

struct Image {
    uint32_t h, w;
    Geometry g;
    uint32_t tile_base, final_base, chunk_base, nchunks;
    uint64_t ws_off;      // count slots; block words at ws_off + align256(kCountBytes)
    uint64_t canon_off;   // canonical state (canonical batches)
    uint64_t nkeys, nnodes;
};

struct Batch {
    uint32_t tile_id = 0;
    bool canonical = false;
    std::vector<Image> images;
    uint32_t total_tiles = 0, total_final = 0, total_chunks = 0;
    uint64_t ws_bytes = 0;
    // image index of every tile id, final workgroup id and scan chunk id
    std::vector<uint32_t> tile_img, final_img, chunk_img;
    // element offsets of the three tables in one device buffer of table_words() words
    uint64_t tile_tab() const { return 0; }
    uint64_t final_tab() const { return tile_img.size(); }
    uint64_t chunk_tab() const { return tile_img.size() + final_img.size(); }
    uint64_t table_words() const { return tile_img.size() + final_img.size() + chunk_img.size(); }
};

// Plan n images (heights h[i], widths w[i]) for one tile shape.  Returns false
// when the ids would not fit 32 bits.
inline bool plan_batch(const uint32_t *h, const uint32_t *w, uint32_t n, uint32_t tile_id, bool canonical, Batch &b)
{
    b = Batch{};
    b.tile_id = tile_id < kTileCount ? tile_id : 0;
    b.canonical = canonical;
    b.images.resize(n);
    uint64_t off = 0, tb = 0, fb = 0, cb = 0;
    for (uint32_t i = 0; i < n; ++i) {
        Image &im = b.images[i];
        im.h = h[i];
        im.w = w[i];
        im.g = geometry2d(h[i], w[i], b.tile_id, kBatchFinalChunks);
        im.tile_base = (uint32_t)tb;
        im.final_base = (uint32_t)fb;
        im.chunk_base = (uint32_t)cb;
        im.ws_off = off;
        off += align256(kCountBytes) + align256((uint64_t)im.g.bw * im.g.bh * 4);
        im.nchunks = 0;
        if (canonical) {   // key array, bits, prefix and partial sums of this image
            im.nkeys = (uint64_t)im.h * im.g.bw;
            im.nnodes = (uint64_t)im.g.bw * im.g.bh;
            im.canon_off = off;
            off += canon_bytes(im.nkeys, im.nnodes);
            im.nchunks = (uint32_t)canon_chunk_count(canon_words(im.nkeys));
        }
        tb += im.g.ntiles;
        fb += im.g.nfinal;
        cb += im.nchunks;
        if (tb > 0xFFFFFFFFull || fb > 0xFFFFFFFFull || cb > 0xFFFFFFFFull) return false;
    }
    b.total_tiles = (uint32_t)tb;
    b.total_final = (uint32_t)fb;
    b.total_chunks = (uint32_t)cb;
    b.ws_bytes = off;
    b.tile_img.resize(tb);
    b.final_img.resize(fb);
    b.chunk_img.resize(cb);
    for (uint32_t i = 0; i < n; ++i) {
        const Image &im = b.images[i];
        std::fill(b.tile_img.begin() + im.tile_base, b.tile_img.begin() + im.tile_base + im.g.ntiles, i);
        std::fill(b.final_img.begin() + im.final_base, b.final_img.begin() + im.final_base + im.g.nfinal, i);
        std::fill(b.chunk_img.begin() + im.chunk_base, b.chunk_img.begin() + im.chunk_base + im.nchunks, i);
    }
    return true;
}

}  // namespace plan
}  // namespace lbl
