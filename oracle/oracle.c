/*
 * oracle.c -- CPU restatement of szavalishin/Labeling's block-equivalence
 * (bleq) and label-equivalence (lbeq) labelers.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker / timed CPU baseline, never by the product library.
 *
 * Each routine keeps the reference's loop structure and OpenMP placement so that
 * it can be timed as the CPU baseline.  The reference's benign data races (a
 * non-atomic min store, a shared "no changes" flag) are expressed with relaxed
 * atomics, which compile to the same plain loads/stores on x86 but are defined
 * behaviour in C.
 *
 * Pinning (see tests/test_oracle.py): every labeler here is checked against the
 * independent BFS labeler on all 4x4, 4x5 and 5x4 images (2D) and all 2x2x2 and
 * 2x3x3 volumes (3D), against scipy.ndimage.label golden fixtures committed
 * under tests/golden/, and against the known-answer images derived from the
 * reference source (right-edge wrap of the unpadded CPU bleq, pixel (0,0) of the
 * CPU lbeq, corner adjacency, degenerate shapes).
 */
#include "oracle.h"

#include <limits.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define LD(p) __atomic_load_n((p), __ATOMIC_RELAXED)
#define ST(p, v) __atomic_store_n((p), (v), __ATOMIC_RELAXED)

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ILabeling::SetupThreads, LabelingTools.cpp:53-65 (0 = MAX_THREADS). */
static void setup_threads(int threads)
{
    if (threads > 0) {
        omp_set_dynamic(0);
        omp_set_num_threads(threads);
    } else {
        omp_set_dynamic(0);
        omp_set_num_threads(omp_get_num_procs());
    }
}

int64_t orc_pad2(int64_t n) { return ((n >> 5) << 5) + 32; }           /* LabelingTools.cpp:103 */
int64_t orc_pad3(int64_t n) { return (((n + 4) >> 5) << 5) + 32; }     /* LabelingTools.cpp:165 (padding 2, align 5) */

/* Binary copy of a pitched image into a zeroed wp x hp buffer at (ox, oy):
 * the padding step of IOCLLabeling::Label (LabelingTools.cpp:103-104). */
static uint8_t *pad_image2(const uint8_t *px, int64_t h, int64_t w, int64_t pitch, int64_t hp, int64_t wp)
{
    uint8_t *out = (uint8_t *)calloc((size_t)(hp * wp), 1);
    if (!out) return NULL;
    for (int64_t y = 0; y < h; ++y)
        for (int64_t x = 0; x < w; ++x)
            out[y * wp + x] = px[y * pitch + x] ? 255 : 0;
    return out;
}

/* CopyAlignImg, LabelingTools.cpp:160-176: zero volume of orc_pad3 size per dim,
 * data at offset +2 in every dimension. */
static uint8_t *pad_volume3(const uint8_t *vox, int64_t d, int64_t h, int64_t w, int64_t s0, int64_t s1, int64_t s2)
{
    uint8_t *out = (uint8_t *)calloc((size_t)(s0 * s1 * s2), 1);
    if (!out) return NULL;
    for (int64_t z = 0; z < d; ++z)
        for (int64_t y = 0; y < h; ++y)
            for (int64_t x = 0; x < w; ++x)
                out[((z + 2) * s1 + (y + 2)) * s2 + (x + 2)] = vox[(z * h + y) * w + x] ? 255 : 0;
    return out;
}

/* ------------------------------------------------------------------------ */
/* 2D block equivalence: TLabelEquivalenceX2, LabelingAlgs.cpp:336-529       */
/* ------------------------------------------------------------------------ */

typedef struct {      /* TSPixel, LabelingAlgs.hpp:119-126 */
    uint32_t lb;
    uint8_t conn;     /* neighbour order: 0 1 2 / 3 x 4 / 5 6 7 */
} spix_t;

typedef struct {      /* the pixel buffer exactly as the reference indexes it */
    const uint8_t *pix;
    int64_t w, h, n;
    int oob;
} img2_t;

/* TestBit / CHECK_PIXEL read pix[px + dx + (py + dy) * w] (LabelingAlgs.cpp:353-364):
 * a 1D offset, so x past an edge wraps into the neighbouring row.  Reads outside
 * the buffer are undefined in the reference; here they read background and are
 * reported. */
static inline int pix2(img2_t *im, int64_t idx)
{
    if (idx < 0 || idx >= im->n) {
        ST(&im->oob, 1);
        return 0;
    }
    return im->pix[idx] != 0;
}

/* SPT / BPT, LabelingAlgs.cpp:360-361: 4x4 window, bit = wx + 4*wy, the block's
 * pixel A at window (1,1).  SPT is A's 3x3 neighbourhood. */
#define SPT2 0x777u
static inline uint32_t bpt2(int x, int y) { return (SPT2 << x) << (4 * y); }

/* RemoveBorderBlocks, LabelingAlgs.cpp:368-376 (called with pixel coordinates). */
static inline uint32_t remove_border2(uint32_t p, int64_t x, int64_t y, int64_t w, int64_t h)
{
    if (x == 0) p &= 0xEEEEu;
    if (y == 0) p &= 0xFFF0u;
    if (x == w - 1) p &= 0x7777u;
    if (y == h - 1) p &= 0x0FFFu;
    return p;
}

/* Neighbour block k (0..7) in the order 0 1 2 / 3 x 4 / 5 6 7. */
static const int NB2_DX[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
static const int NB2_DY[8] = {-1, -1, -1, 0, 0, 1, 1, 1};

/* Connectivity byte of one block (the `if (testPattern)` body of InitSPixels,
 * LabelingAlgs.cpp:403-426).  For neighbour direction (nx, ny) the pixels of that
 * neighbour block inside the 4x4 window are, per axis, offset -1 for n=-1, {0,1}
 * for n=0 and 2 for n=+1; the bit is set iff one of them is a window bit of
 * testPattern and is foreground. */
static uint8_t block_conn2(img2_t *im, uint32_t tp, int64_t px, int64_t py)
{
    uint8_t conn = 0;
    for (int k = 0; k < 8; ++k) {
        int nx = NB2_DX[k], ny = NB2_DY[k];
        int x0 = nx < 0 ? -1 : (nx == 0 ? 0 : 2), x1 = nx == 0 ? 1 : x0;
        int y0 = ny < 0 ? -1 : (ny == 0 ? 0 : 2), y1 = ny == 0 ? 1 : y0;
        int hit = 0;
        for (int oy = y0; oy <= y1 && !hit; ++oy)
            for (int ox = x0; ox <= x1 && !hit; ++ox) {
                int wbit = (ox + 1) + 4 * (oy + 1);
                if ((tp >> wbit) & 1u)
                    hit = pix2(im, px + ox + (py + oy) * im->w);
            }
        if (hit) conn |= (uint8_t)(1u << k);
    }
    return conn;
}

/* InitSPixels, LabelingAlgs.cpp:380-433 (0-based lb = spos). */
static void bleq2_init(img2_t *im, spix_t *sp, int64_t sw, int64_t sh)
{
    const int64_t w = im->w, h = im->h;
    #pragma omp parallel for schedule(static)
    for (int64_t spy = 0; spy < sh; ++spy) {
        for (int64_t spx = 0; spx < sw; ++spx) {
            int64_t spos = spx + spy * sw;
            int64_t px = spx * 2, py = spy * 2;
            int64_t ppos = px + py * w;
            spix_t s = {0, 0};
            uint32_t tp = 0;
            if (pix2(im, ppos)) tp |= bpt2(0, 0);
            if (pix2(im, ppos + w)) tp |= bpt2(0, 1);
            if (pix2(im, ppos + 1)) tp |= bpt2(1, 0);
            if (pix2(im, ppos + 1 + w)) tp |= bpt2(1, 1);
            tp = remove_border2(tp, px, py, w, h) & 0xFFFFu;   /* ushort testPattern */
            if (tp) {
                s.lb = (uint32_t)spos;
                s.conn = block_conn2(im, tp, px, py);
            }
            sp[spos] = s;
        }
    }
}

/* Scan + MinSPixLabel + GetBlockLabel, LabelingAlgs.cpp:437-486.  Returns 1 when
 * nothing changed.  Block indices are 1D in the block grid like the reference;
 * an index outside the grid (only reachable through the unpadded wrap) counts
 * as an out-of-bounds read and contributes nothing. */
static int bleq2_scan(spix_t *sp, int64_t sw, int64_t sh, int *oob)
{
    int no_changes = 1;
    const int64_t nb = sw * sh;
    #pragma omp parallel for schedule(static)
    for (int64_t y = 0; y < sh; ++y) {
        for (int64_t x = 0; x < sw; ++x) {
            uint32_t label = LD(&sp[x + y * sw].lb);
            if (!label) continue;
            uint8_t conn = sp[x + y * sw].conn;
            uint32_t min_label = UINT32_MAX;
            for (int k = 0; k < 8; ++k) {
                if (!(conn & (1u << k))) continue;
                int64_t q = x + NB2_DX[k] + (y + NB2_DY[k]) * sw;
                if (q < 0 || q >= nb) { ST(oob, 1); continue; }
                uint32_t l = LD(&sp[q].lb);
                if (l < min_label) min_label = l;
            }
            if (min_label < label) {
                uint32_t tmp = LD(&sp[label].lb);
                ST(&sp[label].lb, tmp < min_label ? tmp : min_label);
                ST(&no_changes, 0);
            }
        }
    }
    return no_changes;
}

/* Analyze, LabelingAlgs.cpp:490-506. */
static void bleq2_analyze(spix_t *sp, int64_t nb)
{
    #pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < nb; ++s) {
        uint32_t label = LD(&sp[s].lb);
        if (!label) continue;
        uint32_t cur = LD(&sp[label].lb);
        while (cur != label) {
            label = LD(&sp[cur].lb);
            cur = LD(&sp[label].lb);
        }
        ST(&sp[s].lb, label);
    }
}

int orc_bleq2d(const uint8_t *px, int64_t h, int64_t w, int64_t pitch,
               uint32_t *labels, int padded, int threads,
               uint32_t *iterations, int *oob_reads, double *seconds)
{
    if (!px || !labels || h <= 0 || w <= 0 || pitch < w) return -1;
    setup_threads(threads);
    int64_t wp = padded ? orc_pad2(w) : w;
    int64_t hp = padded ? orc_pad2(h) : h;
    uint8_t *img = pad_image2(px, h, w, pitch, hp, wp);
    uint32_t *lab = (uint32_t *)calloc((size_t)(wp * hp), sizeof(uint32_t));   /* labels = zeros, LabelingTools.cpp:25 */
    if (!img || !lab) { free(img); free(lab); return -2; }
    img2_t im = {img, wp, hp, wp * hp, 0};

    double t0 = now_s();
    /* DoLabel, LabelingAlgs.cpp:336-349 */
    int64_t sw = (wp + 1) / 2, sh = (hp + 1) / 2;     /* ceil(cols / 2), ceil(rows / 2) */
    spix_t *sp = (spix_t *)calloc((size_t)(sw * sh), sizeof(spix_t));
    if (!sp) { free(img); free(lab); return -2; }
    bleq2_init(&im, sp, sw, sh);
    uint32_t iters = 0;
    for (;;) {
        ++iters;
        if (bleq2_scan(sp, sw, sh, &im.oob)) break;
        bleq2_analyze(sp, sw * sh);
    }
    /* SetFinalLabels, LabelingAlgs.cpp:510-529 */
    #pragma omp parallel for schedule(static)
    for (int64_t y = 0; y < hp; ++y)
        for (int64_t x = 0; x < wp; ++x) {
            int64_t spos = x / 2 + (y / 2) * sw, pos = x + y * wp;
            if (img[pos]) lab[pos] = sp[spos].lb + 1;
        }
    double t1 = now_s();

    for (int64_t y = 0; y < h; ++y)                   /* crop, LabelingTools.cpp:129 */
        memcpy(labels + y * w, lab + y * wp, (size_t)w * sizeof(uint32_t));
    if (iterations) *iterations = iters;
    if (oob_reads) *oob_reads = im.oob;
    if (seconds) *seconds = t1 - t0;
    free(sp); free(img); free(lab);
    return 0;
}

int64_t orc_bleq2d_conn(const uint8_t *px, int64_t h, int64_t w, int64_t pitch, uint8_t *conn)
{
    if (!px || !conn || h <= 0 || w <= 0 || pitch < w) return -1;
    int64_t wp = orc_pad2(w), hp = orc_pad2(h);
    uint8_t *img = pad_image2(px, h, w, pitch, hp, wp);
    if (!img) return -2;
    img2_t im = {img, wp, hp, wp * hp, 0};
    int64_t sw = wp / 2, sh = hp / 2;
    spix_t *sp = (spix_t *)calloc((size_t)(sw * sh), sizeof(spix_t));
    if (!sp) { free(img); return -2; }
    bleq2_init(&im, sp, sw, sh);
    for (int64_t i = 0; i < sw * sh; ++i) conn[i] = sp[i].conn;
    free(sp); free(img);
    return sw * sh;
}

/* ------------------------------------------------------------------------ */
/* 3D block equivalence: BLEQ3D_* kernels, LabelingAlgs.cl:737-979           */
/* ------------------------------------------------------------------------ */

/* SPT3D / BPT3D, LabelingAlgs.cl:750-751: 4x4x4 window, bit = w0 + 4*w1 + 16*w2
 * where dim 0 is the slowest memory axis and dim 2 the fastest. */
#define SPT3 0x77707770777ull
static inline uint64_t bpt3(int x, int y, int z) { return ((SPT3 << x) << (4 * y)) << (16 * z); }

/* RemoveBorderBlocks3D, LabelingAlgs.cl:791-801 (block coordinates). */
static inline uint64_t remove_border3(uint64_t p, const int64_t b[3], const int64_t s[3])
{
    if (b[0] == 0) p &= 0xEEEEEEEEEEEEEEEEull;
    if (b[1] == 0) p &= 0xFFF0FFF0FFF0FFF0ull;
    if (b[2] == 0) p &= 0xFFFFFFFFFFFF0000ull;
    if (b[0] == s[0] - 1) p &= 0x7777777777777777ull;
    if (b[1] == s[1] - 1) p &= 0x0FFF0FFF0FFF0FFFull;
    if (b[2] == s[2] - 1) p &= 0x0000FFFFFFFFFFFFull;
    return p;
}

typedef struct {
    const uint8_t *vox;
    int64_t s0, s1, s2;   /* padded voxel extents */
    int64_t b0, b1, b2;   /* block extents = s / 2 */
} vol3_t;

/* BLEQ3D_Init, LabelingAlgs.cl:805-860, for one block.  Connectivity bit for
 * neighbour (n0, n1, n2) is (n0+1) + 3(n1+1) + 9(n2+1) (bit 13 = self unused);
 * the tested voxels are those of the neighbour block that lie in the window
 * (per axis offset -1, {0,1} or 2), each gated by its window bit. */
static void bleq3_init_block(const vol3_t *v, int64_t b0, int64_t b1, int64_t b2,
                             uint32_t *lab, uint32_t *conn_out)
{
    const int64_t bb[3] = {b0, b1, b2}, ss[3] = {v->b0, v->b1, v->b2};
    const int64_t st0 = v->s1 * v->s2, st1 = v->s2;
    int64_t spos = (b0 * v->b1 + b1) * v->b2 + b2;
    int64_t ppos = (2 * b0) * st0 + (2 * b1) * st1 + 2 * b2;
    uint64_t tp = 0;
    for (int z = 0; z < 2; ++z)          /* CHECK_SLICE_3D(0), CHECK_SLICE_3D(1) */
        for (int y = 0; y < 2; ++y)
            for (int x = 0; x < 2; ++x)
                if (v->vox[ppos + x * st0 + y * st1 + z]) tp |= bpt3(x, y, z);
    tp = remove_border3(tp, bb, ss);
    uint32_t conn = 0;
    if (tp) {
        lab[spos] = (uint32_t)(spos + 1);
        for (int n2 = -1; n2 <= 1; ++n2)
            for (int n1 = -1; n1 <= 1; ++n1)
                for (int n0 = -1; n0 <= 1; ++n0) {
                    if (!n0 && !n1 && !n2) continue;
                    int lo[3], hi[3], nn[3] = {n0, n1, n2};
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = nn[a] < 0 ? -1 : (nn[a] == 0 ? 0 : 2);
                        hi[a] = nn[a] == 0 ? 1 : lo[a];
                    }
                    int hit = 0;
                    for (int o2 = lo[2]; o2 <= hi[2] && !hit; ++o2)
                        for (int o1 = lo[1]; o1 <= hi[1] && !hit; ++o1)
                            for (int o0 = lo[0]; o0 <= hi[0] && !hit; ++o0) {
                                int wbit = (o0 + 1) + 4 * (o1 + 1) + 16 * (o2 + 1);
                                if ((tp >> wbit) & 1ull)
                                    hit = v->vox[ppos + o0 * st0 + o1 * st1 + o2] != 0;
                            }
                    if (hit) conn |= 1u << ((n0 + 1) + 3 * (n1 + 1) + 9 * (n2 + 1));
                }
    } else {
        lab[spos] = 0;   /* the reference leaves it uninitialised (clCreateBuffer, LabelingAlgs.cpp:1451) */
    }
    conn_out[spos] = conn;
}

/* BLEQ3D_Scan + MinSPixLabel3D, LabelingAlgs.cl:864-932. */
static int bleq3_scan(const vol3_t *v, uint32_t *lab, const uint32_t *conn)
{
    int no_changes = 1;
    #pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < v->b0; ++x)
        for (int64_t y = 0; y < v->b1; ++y)
            for (int64_t z = 0; z < v->b2; ++z) {
                int64_t spos = (x * v->b1 + y) * v->b2 + z;
                uint32_t label = LD(&lab[spos]);
                if (!label) continue;
                uint32_t c = conn[spos], min_label = UINT32_MAX;
                for (int bit = 0; bit < 27; ++bit) {
                    if (!(c & (1u << bit))) continue;
                    int n0 = bit % 3 - 1, n1 = (bit / 3) % 3 - 1, n2 = bit / 9 - 1;
                    uint32_t l = LD(&lab[((x + n0) * v->b1 + (y + n1)) * v->b2 + (z + n2)]);
                    if (l < min_label) min_label = l;
                }
                if (min_label < label) {
                    uint32_t tmp = LD(&lab[label - 1]);
                    ST(&lab[label - 1], tmp < min_label ? tmp : min_label);
                    ST(&no_changes, 0);
                }
            }
    return no_changes;
}

/* BLEQ3D_Analyze / LBEQ3D_AnalyzeKernel, LabelingAlgs.cl:718-734,936-952 (1-based). */
static void analyze_1based(uint32_t *lab, int64_t n)
{
    #pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < n; ++s) {
        uint32_t label = LD(&lab[s]);
        if (!label) continue;
        uint32_t cur = LD(&lab[label - 1]);
        while (cur != label) {
            label = LD(&lab[cur - 1]);
            cur = LD(&lab[label - 1]);
        }
        ST(&lab[s], label);
    }
}

int orc_bleq3d(const uint8_t *vox, int64_t d, int64_t h, int64_t w,
               uint32_t *labels, int threads, uint32_t *iterations, double *seconds)
{
    if (!vox || !labels || d <= 0 || h <= 0 || w <= 0) return -1;
    setup_threads(threads);
    int64_t s0 = orc_pad3(d), s1 = orc_pad3(h), s2 = orc_pad3(w);
    uint8_t *pv = pad_volume3(vox, d, h, w, s0, s1, s2);
    vol3_t v = {pv, s0, s1, s2, s0 / 2, s1 / 2, s2 / 2};
    int64_t nb = v.b0 * v.b1 * v.b2;
    uint32_t *lab = (uint32_t *)malloc((size_t)nb * sizeof(uint32_t));
    uint32_t *conn = (uint32_t *)malloc((size_t)nb * sizeof(uint32_t));
    if (!pv || !lab || !conn) { free(pv); free(lab); free(conn); return -2; }

    double t0 = now_s();
    /* InitSPixels: BLEQ3D_Init over {b0, b1, b2}, LabelingAlgs.cpp:1443-1463 */
    #pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < v.b0; ++x)
        for (int64_t y = 0; y < v.b1; ++y)
            for (int64_t z = 0; z < v.b2; ++z)
                bleq3_init_block(&v, x, y, z, lab, conn);
    /* LabelSPixels, LabelingAlgs.cpp:1467-1492 */
    uint32_t iters = 0;
    for (;;) {
        ++iters;
        if (bleq3_scan(&v, lab, conn)) break;
        analyze_1based(lab, nb);
    }
    /* BLEQ3D_SetFinalLabels, LabelingAlgs.cl:956-979, restricted to the
     * cropped region (the reference returns the padded volume, F6). */
    #pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < d; ++z)
        for (int64_t y = 0; y < h; ++y)
            for (int64_t x = 0; x < w; ++x) {
                int64_t q0 = z + 2, q1 = y + 2, q2 = x + 2;
                int64_t pos = (q0 * s1 + q1) * s2 + q2;
                int64_t spos = ((q0 >> 1) * v.b1 + (q1 >> 1)) * v.b2 + (q2 >> 1);
                labels[(z * h + y) * w + x] = pv[pos] ? lab[spos] : 0;
            }
    double t1 = now_s();
    if (iterations) *iterations = iters;
    if (seconds) *seconds = t1 - t0;
    free(pv); free(lab); free(conn);
    return 0;
}

int64_t orc_bleq3d_conn(const uint8_t *vox, int64_t d, int64_t h, int64_t w, uint32_t *conn)
{
    if (!vox || !conn || d <= 0 || h <= 0 || w <= 0) return -1;
    int64_t s0 = orc_pad3(d), s1 = orc_pad3(h), s2 = orc_pad3(w);
    uint8_t *pv = pad_volume3(vox, d, h, w, s0, s1, s2);
    vol3_t v = {pv, s0, s1, s2, s0 / 2, s1 / 2, s2 / 2};
    int64_t nb = v.b0 * v.b1 * v.b2;
    uint32_t *lab = (uint32_t *)malloc((size_t)nb * sizeof(uint32_t));
    if (!pv || !lab) { free(pv); free(lab); return -2; }
    for (int64_t x = 0; x < v.b0; ++x)
        for (int64_t y = 0; y < v.b1; ++y)
            for (int64_t z = 0; z < v.b2; ++z)
                bleq3_init_block(&v, x, y, z, lab, conn);
    free(pv); free(lab);
    return nb;
}

/* ------------------------------------------------------------------------ */
/* 2D label equivalence: TLabelDistribution, LabelingAlgs.cpp:211-330        */
/* ------------------------------------------------------------------------ */

/* The MinLabel chain (MinLabel, LabelingAlgs.cpp:242-250, folded over the
 * neighbours by MinNWSELabel :254-270) is the minimum over the nonzero
 * neighbour labels, UINT_MAX when there is none. */

int orc_lbeq2d(const uint8_t *px, int64_t h, int64_t w, int64_t pitch,
               int conn, uint32_t *labels, int quirks, int threads,
               uint32_t *iterations, double *seconds)
{
    if (!px || !labels || h <= 0 || w <= 0 || pitch < w) return -1;
    if (conn != 0 && conn != 4 && conn != 8) return -1;
    int c8 = (conn == 8);                                     /* COH_DEFAULT -> COH_4, LabelingAlgs.cpp:213 */
    setup_threads(threads);
    if (quirks < 0 || quirks > 2) return -1;
    /* quirks 1: unpadded (rows wrap); 0 and 2: the GPU padding contract */
    int64_t wp = quirks == 1 ? w : orc_pad2(w), hp = quirks == 1 ? h : orc_pad2(h);
    uint8_t *img = pad_image2(px, h, w, pitch, hp, wp);
    uint32_t *lab = (uint32_t *)malloc((size_t)(wp * hp) * sizeof(uint32_t));
    if (!img || !lab) { free(img); free(lab); return -2; }
    const int64_t n = wp * hp;
    /* quirks: GetLabel takes a 32-bit unsigned position (LabelingAlgs.cpp:274-277) */
    if (quirks && n > (int64_t)UINT32_MAX) { free(img); free(lab); return -1; }
    /* quirks 1, 2: 0-based labels, so position 0 is never labelled nor read
     * (GetLabel: pos && pos < maxSize, LabelingAlgs.cpp:274-277, LabelingAlgs.cl:44-52);
     * quirks 0: 1-based as the reference's 3D variant (LabelingAlgs.cl:641) */
    const uint32_t base = quirks ? 0u : 1u;

    double t0 = now_s();
    /* InitMap, LabelingAlgs.cpp:226-238 / DistrInitKernel, LabelingAlgs.cl:31-38 */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) lab[i] = img[i] ? (uint32_t)i + base : 0u;

    static const int DX8[8] = {-1, 0, 0, -1, 1, -1, 1, 1};
    static const int DY8[8] = {0, -1, 1, -1, 1, 1, -1, 0};
    const int nnb = c8 ? 8 : 4;
    static const int DX4[4] = {-1, 0, 0, 1};
    static const int DY4[4] = {0, -1, 1, 0};
    uint32_t iters = 0;
    for (;;) {
        /* Scan, LabelingAlgs.cpp:281-305, MinNWSELabel :254-270, GetLabel :274-277 */
        int no_changes = 1;
        ++iters;
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint32_t label = LD(&lab[i]);
            if (!label) continue;
            uint32_t m = UINT32_MAX;
            for (int k = 0; k < nnb; ++k) {
                int64_t dx = c8 ? DX8[k] : DX4[k], dy = c8 ? DY8[k] : DY4[k];
                uint32_t l = 0;
                if (quirks) {
                    uint32_t pos = (uint32_t)i + (uint32_t)dx + (uint32_t)dy * (uint32_t)wp;   /* uint arithmetic wraps */
                    if (pos && pos < (uint32_t)n) l = LD(&lab[pos]);
                } else {
                    uint64_t pos = (uint64_t)i + (uint64_t)dx + (uint64_t)dy * (uint64_t)wp;
                    if (pos < (uint64_t)n) l = LD(&lab[pos]);
                }
                if (l && l < m) m = l;   /* MinLabel chain: min over nonzero labels */
            }
            if (m < label) {
                uint32_t *slot = &lab[label - base];
                uint32_t tmp = LD(slot);
                ST(slot, tmp < m ? tmp : m);
                ST(&no_changes, 0);
            }
        }
        if (no_changes) break;
        /* Analyze, LabelingAlgs.cpp:309-330 */
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint32_t label = LD(&lab[i]);
            if (!label) continue;
            uint32_t cur = LD(&lab[label - base]);
            while (cur != label) {
                label = LD(&lab[cur - base]);
                cur = LD(&lab[label - base]);
            }
            ST(&lab[i], label);
        }
    }
    double t1 = now_s();
    for (int64_t y = 0; y < h; ++y)
        memcpy(labels + y * w, lab + y * wp, (size_t)w * sizeof(uint32_t));
    if (iterations) *iterations = iters;
    if (seconds) *seconds = t1 - t0;
    free(img); free(lab);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 3D label equivalence: LBEQ3D_* kernels, LabelingAlgs.cl:635-734           */
/* ------------------------------------------------------------------------ */

int orc_lbeq3d(const uint8_t *vox, int64_t d, int64_t h, int64_t w,
               uint32_t *labels, int threads, uint32_t *iterations, double *seconds)
{
    if (!vox || !labels || d <= 0 || h <= 0 || w <= 0) return -1;
    setup_threads(threads);
    int64_t s0 = orc_pad3(d), s1 = orc_pad3(h), s2 = orc_pad3(w);
    uint8_t *pv = pad_volume3(vox, d, h, w, s0, s1, s2);
    const int64_t n = s0 * s1 * s2;
    uint32_t *lab = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    if (!pv || !lab) { free(pv); free(lab); return -2; }

    double t0 = now_s();
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) lab[i] = pv[i] ? (uint32_t)(i + 1) : 0u;   /* LBEQ3D_InitKernel */
    uint32_t iters = 0;
    for (;;) {
        int no_changes = 1;
        ++iters;
        /* LBEQ3D_ScanKernel + Min3DLabel + GetLabel3D (bounds-checked), :646-714 */
        #pragma omp parallel for schedule(static)
        for (int64_t x = 0; x < s0; ++x)
            for (int64_t y = 0; y < s1; ++y)
                for (int64_t z = 0; z < s2; ++z) {
                    int64_t pos = (x * s1 + y) * s2 + z;
                    uint32_t label = LD(&lab[pos]);
                    if (!label) continue;
                    uint32_t m = UINT32_MAX;
                    for (int dx = -1; dx <= 1; ++dx)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dz = -1; dz <= 1; ++dz) {
                                if (!dx && !dy && !dz) continue;
                                int64_t X = x + dx, Y = y + dy, Z = z + dz;
                                if (X < 0 || Y < 0 || Z < 0 || X >= s0 || Y >= s1 || Z >= s2) continue;
                                uint32_t l = LD(&lab[(X * s1 + Y) * s2 + Z]);
                                if (l && l < m) m = l;
                            }
                    if (m < label) {
                        uint32_t tmp = LD(&lab[label - 1]);
                        ST(&lab[label - 1], tmp < m ? tmp : m);
                        ST(&no_changes, 0);
                    }
                }
        if (no_changes) break;
        analyze_1based(lab, n);   /* LBEQ3D_AnalyzeKernel */
    }
    #pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < d; ++z)
        for (int64_t y = 0; y < h; ++y)
            for (int64_t x = 0; x < w; ++x)
                labels[(z * h + y) * w + x] = lab[((z + 2) * s1 + (y + 2)) * s2 + (x + 2)];
    double t1 = now_s();
    if (iterations) *iterations = iters;
    if (seconds) *seconds = t1 - t0;
    free(pv); free(lab);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Independent checkers                                                      */
/* ------------------------------------------------------------------------ */

int64_t orc_bfs2d(const uint8_t *px, int64_t h, int64_t w, int64_t pitch, int conn, uint32_t *labels)
{
    if (!px || !labels || h <= 0 || w <= 0 || pitch < w || (conn != 4 && conn != 8)) return -1;
    int64_t n = h * w;
    int64_t *queue = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!queue) return -2;
    memset(labels, 0, (size_t)n * sizeof(uint32_t));
    uint32_t next = 0;
    for (int64_t s = 0; s < n; ++s) {
        if (labels[s] || !px[(s / w) * pitch + s % w]) continue;
        labels[s] = ++next;
        int64_t head = 0, tail = 0;
        queue[tail++] = s;
        while (head < tail) {
            int64_t q = queue[head++], y = q / w, x = q % w;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dx && !dy) continue;
                    if (conn == 4 && dx && dy) continue;
                    int64_t yy = y + dy, xx = x + dx;
                    if (yy < 0 || xx < 0 || yy >= h || xx >= w) continue;
                    int64_t r = yy * w + xx;
                    if (labels[r] || !px[yy * pitch + xx]) continue;
                    labels[r] = next;
                    queue[tail++] = r;
                }
        }
    }
    free(queue);
    return next;
}

int64_t orc_bfs3d(const uint8_t *vox, int64_t d, int64_t h, int64_t w, int conn, uint32_t *labels)
{
    if (!vox || !labels || d <= 0 || h <= 0 || w <= 0 || (conn != 6 && conn != 18 && conn != 26)) return -1;
    int64_t n = d * h * w;
    int64_t *queue = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!queue) return -2;
    memset(labels, 0, (size_t)n * sizeof(uint32_t));
    uint32_t next = 0;
    for (int64_t s = 0; s < n; ++s) {
        if (labels[s] || !vox[s]) continue;
        labels[s] = ++next;
        int64_t head = 0, tail = 0;
        queue[tail++] = s;
        while (head < tail) {
            int64_t q = queue[head++];
            int64_t z = q / (h * w), y = (q / w) % h, x = q % w;
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int nz = (dx != 0) + (dy != 0) + (dz != 0);
                        if (!nz) continue;
                        if (conn == 6 && nz > 1) continue;
                        if (conn == 18 && nz > 2) continue;
                        int64_t zz = z + dz, yy = y + dy, xx = x + dx;
                        if (zz < 0 || yy < 0 || xx < 0 || zz >= d || yy >= h || xx >= w) continue;
                        int64_t r = (zz * h + yy) * w + xx;
                        if (labels[r] || !vox[r]) continue;
                        labels[r] = next;
                        queue[tail++] = r;
                    }
        }
    }
    free(queue);
    return next;
}

int64_t orc_canonicalize(uint32_t *labels, int64_t n)
{
    if (!labels || n < 0) return -1;
    /* open-addressing map label -> canonical id, capacity a power of two >= 2n */
    int64_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    uint32_t *keys = (uint32_t *)calloc((size_t)cap, sizeof(uint32_t));
    uint32_t *vals = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
    if (!keys || !vals) { free(keys); free(vals); return -2; }
    uint32_t next = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t l = labels[i];
        if (!l) continue;
        uint64_t hsh = (uint64_t)l * 0x9E3779B97F4A7C15ull;
        int64_t slot = (int64_t)(hsh >> 20) & (cap - 1);
        while (keys[slot] && keys[slot] != l) slot = (slot + 1) & (cap - 1);
        if (!keys[slot]) { keys[slot] = l; vals[slot] = ++next; }
        labels[i] = vals[slot];
    }
    free(keys); free(vals);
    return next;
}

/* ------------------------------------------------------------------------ */
/* Exhaustive self-checks                                                     */
/* ------------------------------------------------------------------------ */

static void note_fail(uint64_t *first, uint64_t code)
{
    #pragma omp critical(orc_first_fail)
    {
        if (code < *first) *first = code;
    }
}

int64_t orc_exhaustive2d(int64_t h, int64_t w, int alg, int conn, uint64_t *first_fail)
{
    const int64_t n = h * w;
    if (h <= 0 || w <= 0 || n > 24 || alg < 0 || alg > 2) return -1;
    const int bconn = alg == 1 ? (conn == 8 ? 8 : 4) : 8;
    const uint64_t total = 1ull << n;
    uint64_t first = UINT64_MAX;
    int64_t bad = 0;
    #pragma omp parallel reduction(+ : bad)
    {
        uint8_t img[24];
        uint32_t a[24], b[24];
        #pragma omp for schedule(dynamic, 256)
        for (uint64_t code = 0; code < total; ++code) {
            for (int64_t i = 0; i < n; ++i) img[i] = ((code >> i) & 1u) ? 255 : 0;
            orc_bfs2d(img, h, w, w, bconn, b);
            int rc = alg == 1 ? orc_lbeq2d(img, h, w, w, bconn, a, 0, 1, NULL, NULL)
                              : orc_bleq2d(img, h, w, w, a, alg == 0, 1, NULL, NULL, NULL);
            if (rc == 0) orc_canonicalize(a, n);
            if (rc != 0 || memcmp(a, b, (size_t)n * sizeof(uint32_t)) != 0) {
                ++bad;
                note_fail(&first, code);
            }
        }
    }
    if (first_fail) *first_fail = first;
    return bad;
}

int64_t orc_exhaustive3d(int64_t d, int64_t h, int64_t w, int alg, uint64_t *first_fail)
{
    const int64_t n = d * h * w;
    if (d <= 0 || h <= 0 || w <= 0 || n > 24 || alg < 0 || alg > 1) return -1;
    const uint64_t total = 1ull << n;
    uint64_t first = UINT64_MAX;
    int64_t bad = 0;
    #pragma omp parallel reduction(+ : bad)
    {
        uint8_t vol[24];
        uint32_t a[24], b[24];
        #pragma omp for schedule(dynamic, 64)
        for (uint64_t code = 0; code < total; ++code) {
            for (int64_t i = 0; i < n; ++i) vol[i] = ((code >> i) & 1u) ? 255 : 0;
            orc_bfs3d(vol, d, h, w, 26, b);
            int rc = alg == 0 ? orc_bleq3d(vol, d, h, w, a, 1, NULL, NULL) : orc_lbeq3d(vol, d, h, w, a, 1, NULL, NULL);
            if (rc == 0) orc_canonicalize(a, n);
            if (rc != 0 || memcmp(a, b, (size_t)n * sizeof(uint32_t)) != 0) {
                ++bad;
                note_fail(&first, code);
            }
        }
    }
    if (first_fail) *first_fail = first;
    return bad;
}
