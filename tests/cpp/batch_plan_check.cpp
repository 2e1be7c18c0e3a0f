// batch_plan_check.cpp -- CPU check of the batch planner (batch_plan.hpp) that
// the batched kernels index by: for the given image shapes, every tile shape
// and both canonical settings, in the order the autotuner re-plans them
// (tile shapes 0..6 into buffers sized once for the largest plan):
//   * each image's tile / final-workgroup / scan-chunk ids are one contiguous
//     range, the ranges tile [0, total) in image order, and the id tables map
//     every id a kernel grid launches (id < total) to its own image;
//   * the local id (id - the image's base) stays inside the image's grid;
//   * final workgroups cover the image's blocks exactly (and clear every word
//     of the canonical bit array in the keys kernel);
//   * workspace slices are ordered, 256-byte aligned, disjoint and inside
//     ws_bytes; every plan fits the buffers sized for the largest one;
//   * the server's per-request table share bounds the request's table words
//     for every tile shape;
//   * the division by the block-grid width that the final kernels use
//     (div_magic / div_by) is exact: every divisor up to 5000 and random ones
//     up to 2^32 - 1, against multiples of the divisor, their neighbours, the
//     ends of the 32-bit range and random numerators.
// Input: a text file of "h w" lines.  Exit code 0 = all checks pass.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "labeling_amd/csrc/batch_plan.hpp"

using namespace lbl::plan;

static long failures = 0;
#define CHECK(cond, ...)                                                \
    do {                                                                \
        if (!(cond)) {                                                  \
            if (++failures <= 20) {                                     \
                std::printf("FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
                std::printf(__VA_ARGS__);                               \
                std::printf("\n");                                      \
            }                                                           \
        }                                                               \
    } while (0)

static void check_plan(const Batch &b, const std::vector<uint32_t> &h, const std::vector<uint32_t> &w)
{
    const uint32_t n = (uint32_t)h.size();
    CHECK(b.images.size() == n, "%zu images", b.images.size());
    CHECK(b.tile_img.size() == b.total_tiles && b.final_img.size() == b.total_final &&
              b.chunk_img.size() == b.total_chunks,
          "table sizes");
    uint64_t tb = 0, fb = 0, cb = 0, ws_end = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const Image &im = b.images[i];
        const TileShape ts = kTiles[b.tile_id];
        CHECK(im.g.bw == (w[i] + 1) / 2 && im.g.bh == (h[i] + 1) / 2, "image %u block grid", i);
        CHECK((uint64_t)im.g.tiles_x * ts.bw >= im.g.bw && (uint64_t)(im.g.tiles_x - 1) * ts.bw < im.g.bw,
              "image %u tiles_x", i);
        const uint64_t ty = im.g.ntiles / im.g.tiles_x;
        CHECK(ty * im.g.tiles_x == im.g.ntiles && ty * ts.bh >= im.g.bh && (ty - 1) * ts.bh < im.g.bh,
              "image %u tile rows", i);
        const uint64_t nb = (uint64_t)im.g.bw * im.g.bh;
        const uint64_t per_wg = (uint64_t)kFinalWgBlocks * kBatchFinalChunks;
        CHECK((uint64_t)im.g.nfinal * per_wg >= nb && (uint64_t)(im.g.nfinal - 1) * per_wg < nb,
              "image %u final workgroups cover its blocks", i);
        CHECK(im.tile_base == tb && im.final_base == fb && im.chunk_base == cb, "image %u bases", i);
        for (uint32_t t = 0; t < im.g.ntiles; ++t)
            CHECK(b.tile_img[im.tile_base + t] == i, "tile %u of image %u", t, i);
        for (uint32_t f = 0; f < im.g.nfinal; ++f)
            CHECK(b.final_img[im.final_base + f] == i, "final %u of image %u", f, i);
        for (uint32_t c = 0; c < im.nchunks; ++c)
            CHECK(b.chunk_img[im.chunk_base + c] == i, "chunk %u of image %u", c, i);
        tb += im.g.ntiles;
        fb += im.g.nfinal;
        cb += im.nchunks;
        // workspace: count slots + block words (+ canonical state), in order
        CHECK(im.ws_off % 256 == 0 && im.ws_off >= ws_end, "image %u ws_off", i);
        uint64_t end = im.ws_off + align256(kCountBytes) + align256(nb * 4);
        if (b.canonical) {
            CHECK(im.canon_off == end && im.canon_off % 256 == 0, "image %u canon_off", i);
            CHECK(im.nkeys == (uint64_t)h[i] * im.g.bw && im.nnodes == nb, "image %u canonical sizes", i);
            const uint64_t words = canon_words(im.nkeys);
            CHECK(im.nchunks == canon_chunk_count(words) && (uint64_t)im.nchunks * kCanonChunkWords >= words,
                  "image %u chunks", i);
            CHECK((uint64_t)im.g.nfinal * kBatchFinalChunks * kFinalThreads >= words,
                  "image %u keys kernel clears every bit word", i);
            end = im.canon_off + canon_bytes(im.nkeys, im.nnodes);
        } else {
            CHECK(im.nchunks == 0, "image %u chunks without canonical", i);
        }
        ws_end = end;
    }
    CHECK(tb == b.total_tiles && fb == b.total_final && cb == b.total_chunks, "totals");
    CHECK(ws_end <= b.ws_bytes, "workspace end %llu > %llu", (unsigned long long)ws_end,
          (unsigned long long)b.ws_bytes);
    // every id a grid launches maps inside the image it names
    for (uint64_t g = 0; g < b.total_tiles; ++g) {
        const uint32_t i = b.tile_img[g];
        CHECK(i < n && g - b.images[i].tile_base < b.images[i].g.ntiles, "tile id %llu", (unsigned long long)g);
    }
    for (uint64_t g = 0; g < b.total_final; ++g) {
        const uint32_t i = b.final_img[g];
        CHECK(i < n && g - b.images[i].final_base < b.images[i].g.nfinal, "final id %llu", (unsigned long long)g);
    }
    for (uint64_t g = 0; g < b.total_chunks; ++g) {
        const uint32_t i = b.chunk_img[g];
        CHECK(i < n && g - b.images[i].chunk_base < b.images[i].nchunks, "chunk id %llu", (unsigned long long)g);
    }
    CHECK(b.final_tab() == b.tile_img.size() && b.chunk_tab() == b.tile_img.size() + b.final_img.size(),
          "table offsets");
}

static void check_division()
{
    std::mt19937_64 rng(7);
    const auto check_d = [&](uint32_t d) {
        const DivMagic g = div_magic(d);
        const auto one = [&](uint32_t n) { CHECK(div_by(n, g) == n / d, "%u / %u: %u", n, d, div_by(n, g)); };
        one(0);
        one(0xFFFFFFFFu);
        one(0xFFFFFFFEu);
        for (uint64_t k = 1; k < 40; ++k) {
            const uint64_t m = k * d;
            if (m - 1 <= 0xFFFFFFFFu) one((uint32_t)(m - 1));
            if (m <= 0xFFFFFFFFu) one((uint32_t)m);
            if (m + 1 <= 0xFFFFFFFFu) one((uint32_t)(m + 1));
        }
        const uint64_t top = 0xFFFFFFFFull / d * d;   // the largest multiple and its neighbours
        one((uint32_t)top);
        if (top) one((uint32_t)(top - 1));
        for (int i = 0; i < 64; ++i) one((uint32_t)rng());
    };
    for (uint32_t d = 1; d <= 5000; ++d) check_d(d);
    for (int i = 0; i < 20000; ++i) check_d(1u + (uint32_t)(rng() % 0xFFFFFFFFull));
    check_d(0xFFFFFFFFu);
    check_d(0x80000000u);
    check_d(0x80000001u);
}

int main(int argc, char **argv)
{
    check_division();
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s shapes.txt\n", argv[0]);
        return 2;
    }
    std::vector<uint32_t> h, w;
    FILE *f = std::fopen(argv[1], "r");
    if (!f) return 2;
    unsigned a, c;
    while (std::fscanf(f, "%u %u", &a, &c) == 2) {
        h.push_back(a);
        w.push_back(c);
    }
    std::fclose(f);
    const uint32_t n = (uint32_t)h.size();
    std::printf("%u images\n", n);
    for (int canonical = 0; canonical < 2; ++canonical) {
        // the autotuner sizes the buffers once for the largest plan of all shapes ...
        uint64_t ws_cap = 0, tab_cap = 0;
        for (uint32_t t = 0; t < kTileCount; ++t) {
            Batch b;
            CHECK(plan_batch(h.data(), w.data(), n, t, canonical != 0, b), "plan %u", t);
            ws_cap = std::max<uint64_t>(ws_cap, b.ws_bytes);
            tab_cap = std::max<uint64_t>(tab_cap, b.table_words());
        }
        // ... then re-plans and uploads every shape in turn
        for (uint32_t t = 0; t < kTileCount; ++t) {
            Batch b;
            CHECK(plan_batch(h.data(), w.data(), n, t, canonical != 0, b), "plan %u", t);
            check_plan(b, h, w);
            CHECK(b.ws_bytes <= ws_cap && b.table_words() <= tab_cap, "shape %u fits the buffers", t);
            std::printf("canonical=%d tile %u (%ux%u): %u tiles, %u final, %u chunks, ws %.1f MB, tables %.1f MB\n",
                        canonical, t, kTiles[t].bw, kTiles[t].bh, b.total_tiles, b.total_final, b.total_chunks,
                        b.ws_bytes / 1e6, b.table_words() * 4 / 1e6);
        }
    }
    // the server reserves, per fused request, the table words of its worst tile shape
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t share = 0;
        for (uint32_t t = 0; t < kTileCount; ++t) {
            const Geometry g = geometry2d(h[i], w[i], t, kBatchFinalChunks);
            share = std::max<uint64_t>(share, (uint64_t)g.ntiles + g.nfinal);
        }
        for (uint32_t t = 0; t < kTileCount; ++t) {
            Batch b;
            CHECK(plan_batch(&h[i], &w[i], 1, t, false, b), "single plan");
            CHECK(b.table_words() <= share, "server share of image %u, shape %u", i, t);
        }
    }
    std::printf("%s (%ld failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
