/*
 * oracle_registry.c -- CPU restatements of the reference registry's other
 * methods (ALG_LIST, main.cpp:103-130): run equivalence, He et al.'s run
 * labeling and Grana et al.'s block labeling.  TEST INFRASTRUCTURE ONLY (see
 * oracle.h): only tests/ load it, as the checker.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---- run equivalence (TRunEqivLabeling, LabelingAlgs.cpp:535-810) ------------ */

typedef struct {
    uint32_t l, r;   /* TRunSize: first and last run-array index of the neighbours */
} RunSize;

typedef struct {
    uint32_t lb, l, r;   /* label (1 + run-array index of the representative), first / last column */
    RunSize top, bot;
} Run;

/* IsNeib (LabelingAlgs.cpp:605-613), including its repeated last clause. */
static int is_neib(const Run *r1, const Run *r2)
{
    return (r1->l >= r2->l && r1->l <= r2->r) || (r1->r >= r2->l && r1->r <= r2->r) ||
           (r2->r >= r1->l && r2->r <= r1->r) || (r2->r >= r1->l && r2->r <= r1->r);
}

/* FindNeibRuns (LabelingAlgs.cpp:617-651): walks the neighbour row from *pos. */
static void find_neib(const Run *cur, RunSize *size, const Run *row, uint32_t *pos, uint32_t n)
{
    int no_neib = 0;
    size->l = 1;
    size->r = 0;
    while (*pos < n && !no_neib) {
        if (is_neib(cur, row + *pos)) {
            if (size->l > size->r) size->l = row[*pos].lb - 1;
            size->r = row[*pos].lb - 1;
            if (*pos + 1 < n && row[*pos + 1].lb && row[*pos + 1].l <= cur->r) ++*pos;
            else no_neib = 1;
        } else if (row[*pos].r < cur->l) {
            ++*pos;
        } else {
            no_neib = 1;
        }
    }
}

int orc_runeq2d(const uint8_t *px, int64_t h, int64_t w, int64_t slots, uint32_t *labels, uint32_t *iterations)
{
    if (!px || !labels || h <= 0 || w <= 0 || slots <= 0) return -1;
    Run *runs = calloc((size_t)(h * slots), sizeof(Run));
    uint32_t *num = calloc((size_t)h, sizeof(uint32_t));
    if (!runs || !num) {
        free(runs);
        free(num);
        return -1;
    }
    int rc = 0;
    /* FindRuns (LabelingAlgs.cpp:569-604); the sentinel lb = 0 after a row's
     * last run is kept inside the row's slots */
    for (int64_t row = 0; row < h && rc == 0; ++row) {
        const uint32_t row_pos = (uint32_t)(row * slots);
        Run *cur = runs + row_pos;
        uint32_t run_pos = 0;
        cur->lb = 0;
        for (int64_t pos = 0; pos < w; ++pos) {
            if (px[row * w + pos]) {
                if (!cur->lb) {
                    if (run_pos >= (uint32_t)slots) { rc = -2; break; }
                    cur->lb = row_pos + ++run_pos;
                    cur->l = (uint32_t)pos;
                }
                if (pos == w - 1) {
                    cur->r = (uint32_t)pos;
                    if (run_pos < (uint32_t)slots) (++cur)->lb = 0;
                }
            } else if (cur->lb) {
                cur->r = (uint32_t)pos - 1;
                if (run_pos < (uint32_t)slots) (++cur)->lb = 0;
            }
        }
        num[row] = run_pos;
    }
    /* FindNeibRuns (LabelingAlgs.cpp:655-691) */
    for (int64_t row = 0; row < h && rc == 0; ++row) {
        Run *cur = runs + row * slots;
        const Run *top = cur - slots, *bot = cur + slots;   /* the neighbour rows' first slots */
        uint32_t tp = 0, bp = 0;
        for (uint32_t pos = 0; pos < num[row]; ++pos, ++cur) {
            if (row > 0) find_neib(cur, &cur->top, top, &tp, num[row - 1]);
            else { cur->top.l = 1; cur->top.r = 0; }
            if (row < h - 1) find_neib(cur, &cur->bot, bot, &bp, num[row + 1]);
            else { cur->bot.l = 1; cur->bot.r = 0; }
        }
    }
    /* Scan / Analyze until no change (LabelingAlgs.cpp:719-790), sequentially */
    uint32_t it = 0;
    for (;;) {
        if (rc) break;
        int no_changes = 1;
        ++it;
        for (int64_t row = 0; row < h; ++row)
            for (uint32_t pos = 0; pos < num[row]; ++pos) {
                const Run *cr = &runs[row * slots + pos];
                const uint32_t label = cr->lb;
                if (!label) continue;
                uint32_t min_label = UINT32_MAX;   /* MinRunLabel, :697-715 */
                if (cr->top.l <= cr->top.r)
                    for (uint32_t i = cr->top.l; i <= cr->top.r; ++i) min_label = runs[i].lb < min_label ? runs[i].lb : min_label;
                if (cr->bot.l <= cr->bot.r)
                    for (uint32_t i = cr->bot.l; i <= cr->bot.r; ++i) min_label = runs[i].lb < min_label ? runs[i].lb : min_label;
                if (min_label < label) {
                    const uint32_t tmp = runs[label - 1].lb;
                    runs[label - 1].lb = tmp < min_label ? tmp : min_label;
                    no_changes = 0;
                }
            }
        if (no_changes) break;
        for (int64_t row = 0; row < h; ++row)
            for (uint32_t pos = 0; pos < num[row]; ++pos) {
                Run *cr = &runs[row * slots + pos];
                uint32_t label = cr->lb;
                if (!label) continue;
                uint32_t cur_label = runs[label - 1].lb;
                while (cur_label != label) {
                    label = runs[cur_label - 1].lb;
                    cur_label = runs[label - 1].lb;
                }
                cr->lb = label;
            }
    }
    /* SetFinalLabels (LabelingAlgs.cpp:794-810) */
    if (rc == 0) {
        memset(labels, 0, (size_t)(h * w) * sizeof(uint32_t));
        for (int64_t row = 0; row < h; ++row)
            for (uint32_t k = 0; k < num[row]; ++k) {
                const Run *cr = &runs[row * slots + k];
                for (uint32_t i = cr->l; i <= cr->r; ++i) labels[row * w + i] = cr->lb;
            }
    }
    if (iterations) *iterations = it;
    free(runs);
    free(num);
    return rc;
}

/* ---- He et al. run labeling (TRunLabeling, LabelingAlgs.cpp:58-205) -------------- */

typedef struct {
    uint32_t l, r, row, label;
} HeRun;

int orc_herun2d(const uint8_t *px, int64_t h, int64_t w, int conpix, int quirks, uint32_t *labels)
{
    if (!px || !labels || h <= 0 || w <= 0) return -1;
    const size_t cap = (size_t)(h * ((w + 1) / 2)) + 1;
    HeRun *runs = malloc(cap * sizeof(HeRun));
    uint32_t *objects = malloc((cap + 1) * sizeof(uint32_t));
    uint32_t *stack = malloc(((size_t)w + 2) * sizeof(uint32_t));
    uint32_t *last = malloc(((size_t)w + 2) * sizeof(uint32_t)), *cur_row = malloc(((size_t)w + 2) * sizeof(uint32_t));
    if (!runs || !objects || !stack || !last || !cur_row) {
        free(runs); free(objects); free(stack); free(last); free(cur_row);
        return -1;
    }
    size_t nruns = 0, nobj = 1, nlast = 0, ncur = 0;
    objects[0] = 0;
    /* SetRunLabel (LabelingAlgs.cpp:80-130) for run index c */
#define SET_RUN_LABEL(c)                                                                          \
    do {                                                                                          \
        HeRun *cr = &runs[c];                                                                     \
        cr->label = UINT32_MAX;                                                                   \
        size_t ns = 0;                                                                            \
        for (size_t i = 0; i < nlast; ++i) {                                                      \
            const HeRun *lr = &runs[last[i]];                                                     \
            if ((int64_t)lr->r >= (int64_t)cr->l - conpix && (int64_t)lr->l <= (int64_t)cr->r + conpix) { \
                if (objects[lr->label] < cr->label) {                                             \
                    cr->label = objects[lr->label];                                               \
                    stack[ns++] = cr->label;                                                      \
                } else {                                                                          \
                    stack[ns++] = objects[lr->label];                                             \
                }                                                                                 \
            }                                                                                     \
        }                                                                                         \
        if (cr->label == UINT32_MAX) {                                                            \
            objects[nobj] = (uint32_t)nobj;                                                       \
            cr->label = (uint32_t)nobj++;                                                         \
        } else {                                                                                  \
            for (size_t i = 1; i < nobj; ++i)                                                     \
                for (size_t j = 0; j < ns; ++j)                                                   \
                    if (objects[i] == stack[j]) objects[i] = cr->label;                           \
        }                                                                                         \
    } while (0)

    const int64_t rows = quirks ? h - 1 : h;   /* Row < Bottom = rows - 1 (:150-153) */
    for (int64_t row = 0; row < rows; ++row) {
        int64_t open = -1;   /* CurRun */
        for (int64_t i = 0; i < w; ++i) {
            const int fg = px[row * w + i] != 0;
            if (!fg || (quirks && i == w - 1)) {
                if (i > 0 && open >= 0) {   /* (:160-171) */
                    runs[open].r = (uint32_t)i - 1;
                    SET_RUN_LABEL(open);
                    open = -1;
                }
            } else if (open < 0) {
                open = (int64_t)nruns++;
                runs[open].l = (uint32_t)i;
                runs[open].row = (uint32_t)row;
                cur_row[ncur++] = (uint32_t)open;
            }
        }
        if (!quirks && open >= 0) {   /* a run reaching the last column */
            runs[open].r = (uint32_t)w - 1;
            SET_RUN_LABEL(open);
        }
        memcpy(last, cur_row, ncur * sizeof(uint32_t));
        nlast = ncur;
        ncur = 0;
    }
#undef SET_RUN_LABEL
    /* setting up labels (:193-200); with quirks a run opened and never closed
     * (it started in the last column) has no label and is not written */
    memset(labels, 0, (size_t)(h * w) * sizeof(uint32_t));
    for (size_t k = 0; k < nruns; ++k) {
        const HeRun *rr = &runs[k];
        if (quirks && rr->l == (uint32_t)w - 1) continue;
        for (uint32_t j = rr->l; j <= rr->r; ++j) labels[(int64_t)rr->row * w + j] = objects[rr->label];
    }
    free(runs); free(objects); free(stack); free(last); free(cur_row);
    return 0;
}

/* ---- block-based two-pass labeling (Grana et al., cvlabeling_imagelab.cpp) ------ */

static uint32_t uf_root(uint32_t *p, uint32_t x)
{
    while (p[x] != x) x = p[x] = p[p[x]];
    return x;
}

int64_t orc_grblock2d(const uint8_t *px, int64_t h, int64_t w, uint32_t *labels)
{
    if (!px || !labels || h <= 0 || w <= 0) return -1;
    const int64_t bw = (w + 1) / 2, bh = (h + 1) / 2;
    uint32_t *lab = calloc((size_t)(bw * bh), sizeof(uint32_t));
    uint32_t *par = malloc(((size_t)(bw * bh) + 1) * sizeof(uint32_t));
    if (!lab || !par) { free(lab); free(par); return -1; }
#define FG(x, y) ((x) >= 0 && (x) < w && (y) >= 0 && (y) < h && px[(y) * w + (x)] != 0)
    uint32_t nl = 0;
    for (int64_t by = 0; by < bh; ++by)
        for (int64_t bx = 0; bx < bw; ++bx) {
            const int64_t x0 = 2 * bx, y0 = 2 * by;
            if (!(FG(x0, y0) || FG(x0 + 1, y0) || FG(x0, y0 + 1) || FG(x0 + 1, y0 + 1))) continue;
            /* earlier neighbour blocks: up-left, up, up-right, left */
            const int64_t nb[4][2] = {{bx - 1, by - 1}, {bx, by - 1}, {bx + 1, by - 1}, {bx - 1, by}};
            uint32_t got = 0;
            for (int k = 0; k < 4; ++k) {
                const int64_t nx = nb[k][0], ny = nb[k][1];
                if (nx < 0 || ny < 0 || nx >= bw) continue;
                const uint32_t ln = lab[ny * bw + nx];
                if (!ln) continue;
                int touch = 0;   /* a foreground pixel of each block, 8-adjacent */
                for (int a = 0; a < 4 && !touch; ++a) {
                    const int64_t ax = x0 + (a & 1), ay = y0 + (a >> 1);
                    if (!FG(ax, ay)) continue;
                    for (int b = 0; b < 4 && !touch; ++b) {
                        const int64_t qx = 2 * nx + (b & 1), qy = 2 * ny + (b >> 1);
                        if (FG(qx, qy) && llabs(ax - qx) <= 1 && llabs(ay - qy) <= 1) touch = 1;
                    }
                }
                if (!touch) continue;
                const uint32_t rn = uf_root(par, ln);
                if (!got) got = rn;
                else {   /* merge: the smaller provisional label stays the representative */
                    const uint32_t rg = uf_root(par, got);
                    if (rg < rn) par[rn] = rg;
                    else par[rg] = rn;
                    got = rg < rn ? rg : rn;
                }
            }
            if (!got) {
                got = ++nl;
                par[got] = got;
            }
            lab[by * bw + bx] = got;
        }
    /* renumber 1..N by representative order (:1523-1531), then write the pixels */
    uint32_t *map = calloc((size_t)nl + 1, sizeof(uint32_t));
    if (!map) { free(lab); free(par); return -1; }
    uint32_t n = 0;
    for (uint32_t k = 1; k <= nl; ++k)
        if (uf_root(par, k) == k) map[k] = ++n;
    for (int64_t y = 0; y < h; ++y)
        for (int64_t x = 0; x < w; ++x) {
            const uint32_t l = lab[(y / 2) * bw + x / 2];
            labels[y * w + x] = (px[y * w + x] && l) ? map[uf_root(par, l)] : 0u;
        }
#undef FG
    free(map);
    free(lab);
    free(par);
    return n;
}
