/*
 * synth.c -- deterministic synthetic binary images and volumes for the
 * labeling workloads (host code, OpenMP).
 *
 * Every value is a pure function of (seed, stream, index): a SplitMix64 state
 * seed*GAMMA + stream, advanced by GAMMA per element and finalised by the
 * SplitMix64 mixer.  No global state, so any thread count and any machine give
 * identical bytes.  Foreground bytes are 255, background 0.
 *
 *  - Bernoulli: foreground iff (h >> 40) < round(p * 2^24).
 *  - Blobs: N(0,1) white noise by Box-Muller (doubles), separable Gaussian of
 *    standard deviation sigma and radius ceil(4 sigma) with mirror-reflect
 *    borders (d c b a | a b c d | d c b a), fixed summation order, then
 *    foreground iff value > the exact q-quantile (nth element of rank
 *    floor(q * n)).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define GAMMA 0x9E3779B97F4A7C15ull

uint64_t lbl_synth_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t stream_key(uint64_t seed, uint64_t stream)
{
    return lbl_synth_mix64(seed * GAMMA + stream);
}

/* i-th 64-bit value of stream (seed, stream). */
uint64_t lbl_synth_u64(uint64_t seed, uint64_t stream, uint64_t i)
{
    return lbl_synth_mix64(stream_key(seed, stream) + GAMMA * (i + 1));
}

void lbl_synth_bernoulli(uint8_t *out, int64_t n, double p, uint64_t seed, uint64_t stream)
{
    if (!out || n <= 0) return;
    const uint64_t key = stream_key(seed, stream);
    double t = p * 16777216.0;
    uint64_t thr = t <= 0.0 ? 0 : (t >= 16777216.0 ? 16777216ull : (uint64_t)llround(t));
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        uint64_t hsh = lbl_synth_mix64(key + GAMMA * (uint64_t)(i + 1));
        out[i] = (hsh >> 40) < thr ? 255 : 0;
    }
}

static void white_noise(double *v, int64_t n, uint64_t seed)
{
    const uint64_t key = stream_key(seed, 0x6E6F697365ull);   /* "noise" */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        uint64_t a = lbl_synth_mix64(key + GAMMA * (uint64_t)(2 * i + 1));
        uint64_t b = lbl_synth_mix64(key + GAMMA * (uint64_t)(2 * i + 2));
        double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);   /* (0, 1] */
        double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);           /* [0, 1) */
        v[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

static inline int64_t reflect(int64_t i, int64_t n)
{
    /* period 2n mirror: -1 -> 0, n -> n-1 */
    int64_t m = 2 * n;
    i %= m;
    if (i < 0) i += m;
    return i < n ? i : m - 1 - i;
}

/* Convolve along one axis of a row-major array viewed as [outer][len][inner]. */
static void convolve_axis(const double *src, double *dst, int64_t outer, int64_t len, int64_t inner,
                          const double *k, int r)
{
    #pragma omp parallel for schedule(static)
    for (int64_t o = 0; o < outer; ++o)
        for (int64_t i = 0; i < len; ++i)
            for (int64_t j = 0; j < inner; ++j) {
                double s = 0.0;
                for (int t = -r; t <= r; ++t)
                    s += k[t + r] * src[(o * len + reflect(i + t, len)) * inner + j];
                dst[(o * len + i) * inner + j] = s;
            }
}

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* k-th smallest (0-based) by quickselect; deterministic pivots. */
static double select_kth(double *a, int64_t n, int64_t k)
{
    int64_t lo = 0, hi = n - 1;
    while (hi - lo > 16) {
        int64_t mid = lo + (hi - lo) / 2;
        double x = a[lo], y = a[mid], z = a[hi], piv;
        if (x < y) piv = (y < z) ? y : (x < z ? z : x);
        else piv = (x < z) ? x : (y < z ? z : y);
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (a[i] < piv) ++i;
            while (a[j] > piv) --j;
            if (i <= j) { double t = a[i]; a[i] = a[j]; a[j] = t; ++i; --j; }
        }
        if (k <= j) hi = j;
        else if (k >= i) lo = i;
        else return a[k];
    }
    qsort(a + lo, (size_t)(hi - lo + 1), sizeof(double), cmp_double);
    return a[k];
}

static int blobs_nd(uint8_t *out, const int64_t *dims, int nd, double sigma, double q, uint64_t seed)
{
    if (!out || sigma <= 0.0 || q < 0.0 || q >= 1.0) return -1;
    int64_t n = 1;
    for (int a = 0; a < nd; ++a) { if (dims[a] <= 0) return -1; n *= dims[a]; }
    int r = (int)ceil(4.0 * sigma);
    double *k = (double *)malloc(sizeof(double) * (size_t)(2 * r + 1));
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    double *t = (double *)malloc(sizeof(double) * (size_t)n);
    if (!k || !v || !t) { free(k); free(v); free(t); return -2; }
    double ks = 0.0;
    for (int i = -r; i <= r; ++i) { k[i + r] = exp(-0.5 * (double)(i * i) / (sigma * sigma)); ks += k[i + r]; }
    for (int i = 0; i <= 2 * r; ++i) k[i] /= ks;
    white_noise(v, n, seed);
    /* innermost (fastest) axis first */
    for (int a = nd - 1; a >= 0; --a) {
        int64_t outer = 1, inner = 1;
        for (int b = 0; b < a; ++b) outer *= dims[b];
        for (int b = a + 1; b < nd; ++b) inner *= dims[b];
        convolve_axis(v, t, outer, dims[a], inner, k, r);
        double *sw = v; v = t; t = sw;
    }
    memcpy(t, v, sizeof(double) * (size_t)n);
    int64_t rank = (int64_t)floor(q * (double)n);
    if (rank >= n) rank = n - 1;
    double thr = select_kth(t, n, rank);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = v[i] > thr ? 255 : 0;
    free(k); free(v); free(t);
    return 0;
}

int lbl_synth_blobs2d(uint8_t *out, int64_t h, int64_t w, double sigma, double quantile, uint64_t seed)
{
    int64_t dims[2] = {h, w};
    return blobs_nd(out, dims, 2, sigma, quantile, seed);
}

int lbl_synth_blobs3d(uint8_t *out, int64_t d, int64_t h, int64_t w, double sigma, double quantile, uint64_t seed)
{
    int64_t dims[3] = {d, h, w};
    return blobs_nd(out, dims, 3, sigma, quantile, seed);
}
