/* oracle_sanitize_check.c -- the oracle's C restatements (oracle/oracle.c,
 * oracle/oracle_registry.c) built with AddressSanitizer and UBSan and run over
 * random images and volumes of every small shape (1-pixel rows and columns,
 * odd sizes, row pitches wider than the row), every entry point and mode:
 * no out-of-bounds access or undefined behaviour may occur, and the exact
 * labelers must agree with the breadth-first labeler after canonical
 * relabelling.  Exit code 0 and "ok (0 failures)" on success. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle/oracle.h"

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (uint32_t)(rng_state >> 11);
}

static int failures = 0;
#define CHECK(c, ...)                                   \
    do {                                                \
        if (!(c)) {                                     \
            ++failures;                                 \
            if (failures < 20) {                        \
                printf("FAIL %s:%d: ", __FILE__, __LINE__); \
                printf(__VA_ARGS__);                    \
                printf("\n");                           \
            }                                           \
        }                                               \
    } while (0)

/* canonical copy of labels (n elements) equals the reference canonical labels */
static int same_partition(const uint32_t *lab, const uint32_t *ref, int64_t n, int64_t nref)
{
    uint32_t *c = malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    memcpy(c, lab, (size_t)n * sizeof(uint32_t));
    const int64_t k = orc_canonicalize(c, n);
    const int ok = k == nref && memcmp(c, ref, (size_t)n * sizeof(uint32_t)) == 0;
    free(c);
    return ok;
}

static void check2d(int64_t h, int64_t w, int64_t pitch, uint32_t pct)
{
    const int64_t n = h * w;
    uint8_t *px = malloc((size_t)(h * pitch));
    uint8_t *dense = malloc((size_t)n);
    for (int64_t i = 0; i < h * pitch; ++i) px[i] = (rnd() % 100 < pct) ? (uint8_t)(1 + rnd() % 255) : 0;
    for (int64_t y = 0; y < h; ++y) memcpy(dense + y * w, px + y * pitch, (size_t)w);
    uint32_t *b4 = malloc((size_t)n * 4), *b8 = malloc((size_t)n * 4), *lab = malloc((size_t)n * 4);
    const int64_t n4 = orc_bfs2d(px, h, w, pitch, 4, b4), n8 = orc_bfs2d(px, h, w, pitch, 8, b8);
    CHECK(n4 >= 0 && n8 >= 0, "bfs2d %lldx%lld", (long long)h, (long long)w);

    uint32_t it = 0;
    int oob = 0;
    CHECK(orc_bleq2d(px, h, w, pitch, lab, 1, 2, &it, &oob, NULL) == 0, "bleq2d padded");
    CHECK(same_partition(lab, b8, n, n8), "bleq2d padded %lldx%lld", (long long)h, (long long)w);
    CHECK(orc_bleq2d(px, h, w, pitch, lab, 0, 2, &it, &oob, NULL) == 0, "bleq2d as written");
    for (int conn = 4; conn <= 8; conn += 4) {
        CHECK(orc_lbeq2d(px, h, w, pitch, conn, lab, 0, 2, &it, NULL) == 0, "lbeq2d");
        CHECK(same_partition(lab, conn == 4 ? b4 : b8, n, conn == 4 ? n4 : n8), "lbeq2d conn %d %lldx%lld", conn,
              (long long)h, (long long)w);
        CHECK(orc_lbeq2d(px, h, w, pitch, conn, lab, 1, 2, &it, NULL) == 0, "lbeq2d quirks 1");
        CHECK(orc_lbeq2d(px, h, w, pitch, conn, lab, 2, 2, &it, NULL) == 0, "lbeq2d quirks 2");
    }
    /* the registry's methods take dense rows */
    const int rr = orc_runeq2d(dense, h, w, (w + 1) / 2, lab, &it);
    CHECK(rr == 0, "runeq2d %d", rr);
    if (rr == 0) CHECK(same_partition(lab, b4, n, n4), "runeq2d %lldx%lld", (long long)h, (long long)w);
    for (int conpix = 0; conpix <= 1; ++conpix) {
        CHECK(orc_herun2d(dense, h, w, conpix, 0, lab) == 0, "herun2d");
        CHECK(same_partition(lab, conpix ? b8 : b4, n, conpix ? n8 : n4), "herun2d conpix %d %lldx%lld", conpix,
              (long long)h, (long long)w);
        CHECK(orc_herun2d(dense, h, w, conpix, 1, lab) == 0, "herun2d quirks");
    }
    const int64_t ng = orc_grblock2d(dense, h, w, lab);
    CHECK(ng == n8 && same_partition(lab, b8, n, n8), "grblock2d %lldx%lld", (long long)h, (long long)w);
    free(px);
    free(dense);
    free(b4);
    free(b8);
    free(lab);
}

static void check3d(int64_t d, int64_t h, int64_t w, uint32_t pct)
{
    const int64_t n = d * h * w;
    uint8_t *v = malloc((size_t)n);
    for (int64_t i = 0; i < n; ++i) v[i] = (rnd() % 100 < pct) ? 255 : 0;
    uint32_t *b = malloc((size_t)n * 4), *lab = malloc((size_t)n * 4);
    const int64_t nb = orc_bfs3d(v, d, h, w, 26, b);
    CHECK(nb >= 0, "bfs3d");
    uint32_t it = 0;
    CHECK(orc_bleq3d(v, d, h, w, lab, 2, &it, NULL) == 0, "bleq3d");
    CHECK(same_partition(lab, b, n, nb), "bleq3d %lldx%lldx%lld", (long long)d, (long long)h, (long long)w);
    CHECK(orc_lbeq3d(v, d, h, w, lab, 2, &it, NULL) == 0, "lbeq3d");
    CHECK(same_partition(lab, b, n, nb), "lbeq3d %lldx%lldx%lld", (long long)d, (long long)h, (long long)w);
    for (int conn = 6; conn <= 18; conn += 12) CHECK(orc_bfs3d(v, d, h, w, conn, lab) >= 0, "bfs3d conn %d", conn);
    free(v);
    free(b);
    free(lab);
}

int main(int argc, char **argv)
{
    const int trials = argc > 1 ? atoi(argv[1]) : 400;
    for (int t = 0; t < trials; ++t) {
        const int64_t h = 1 + rnd() % 40, w = 1 + rnd() % 40, pitch = w + rnd() % 8;
        check2d(h, w, pitch, 5 + rnd() % 91);
    }
    for (int t = 0; t < trials / 4; ++t) check3d(1 + rnd() % 10, 1 + rnd() % 12, 1 + rnd() % 14, 5 + rnd() % 91);
    uint64_t first = 0;
    CHECK(orc_exhaustive2d(3, 4, 0, 8, &first) == 0, "exhaustive bleq2d 3x4");
    CHECK(orc_exhaustive2d(4, 3, 1, 4, &first) == 0, "exhaustive lbeq2d 4x3");
    CHECK(orc_exhaustive3d(2, 2, 3, 0, &first) == 0, "exhaustive bleq3d 2x2x3");
    CHECK(orc_pad2(0) >= 0 && orc_pad3(1) > 1, "pads");
    CHECK(orc_bfs2d(NULL, 1, 1, 1, 8, NULL) < 0, "null arguments rejected");
    printf("ok (%d failures)\n", failures);
    return failures ? 1 : 0;
}
