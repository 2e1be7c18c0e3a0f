// bleq2d.hip -- 2D block-equivalence labeling (2x2 blocks, 8-connectivity) for
// gfx950: single images and batches of images of uneven sizes.
//
// Same problem as the reference's TOCLLabelEquivalenceX2 (LabelingAlgs.cpp:948-1084,
// kernels LBEQ2_* LabelingAlgs.cl:133-319): pixels are grouped into 2x2 blocks,
// blocks with a foreground pixel are the nodes, two blocks are equivalent when a
// foreground pixel of one is 8-adjacent to a foreground pixel of the other, and
// each component is labelled by its minimum block.  The reference iterates
// Scan/Analyze passes over the whole block grid with a blocking host round trip
// per pass; here the equivalences are resolved by union-find in four launches
// and no host synchronisation:
//
//   1. tile    one workgroup per TBW x TBH-block tile: 16-byte row loads ->
//              4-bit block masks; union-find of the tile's blocks in LDS;
//              writes one packed word per block, parent << 4 | mask, whose
//              parent is the tile-local root (global block index).
//   2. merge   one thread per block on a tile border: unions with each
//              8-neighbour block in another tile that it touches (atomicMin
//              links between tile-local roots).
//   3. flatten one thread per tile-border block: the block and its tile-local
//              root are pointed straight at the global root.
//   4. final   one thread per 4 blocks: root of the block (<= 2 hops), writes
//              the 8 labels of the 2x2 pixels of each block, 0 for background,
//              and counts roots.
//
// Every kernel reads its image geometry from a "source": One (a single image
// passed by value) or Many (a device table of images; a workgroup finds its
// image by binary search over the images' first tile / workgroup ids), so a
// batch of thousands of images runs as the same four launches.
//
// HBM traffic per pixel: 1 B image + 1 B block words in (1), 1 B block words +
// 4 B labels in (4), plus the border blocks in (2)-(3).
#include "bleq2d.hpp"
#include "uf_device.hpp"

namespace lbl {
namespace {

constexpr uint32_t SEG = 8;              // blocks per thread in the tile kernel
constexpr uint32_t FINAL_BLOCKS = 4;     // blocks per thread in the final kernel
constexpr uint32_t NT_FINAL = 256;

using Forest = PackedForest<4>;

// Block mask bits: 1 = (0,0), 2 = (1,0), 4 = (0,1), 8 = (1,1).
// Pair tests for a block `c` and an earlier neighbour `n` in direction
// left, up-left, up, up-right: a foreground pixel of c is 8-adjacent to a
// foreground pixel of n.
__device__ __forceinline__ bool adj_left(uint32_t c, uint32_t n) { return (c & 0x5u) && (n & 0xAu); }
__device__ __forceinline__ bool adj_upleft(uint32_t c, uint32_t n) { return (c & 0x1u) && (n & 0x8u); }
__device__ __forceinline__ bool adj_up(uint32_t c, uint32_t n) { return (c & 0x3u) && (n & 0xCu); }
__device__ __forceinline__ bool adj_upright(uint32_t c, uint32_t n) { return (c & 0x2u) && (n & 0x4u); }

// ---- geometry sources ----------------------------------------------------------

struct One {
    Geom2D g;
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
    __device__ __forceinline__ const Geom2D &by_image(uint32_t i, uint32_t &li) const
    {
        li = i;
        return g;
    }
    uint32_t images() const { return g.count ? 1u : 0u; }
};

struct Many {
    const Geom2D *g;
    uint32_t n;
    // last image whose first id is <= id
    template <int FIELD>
    __device__ __forceinline__ uint32_t find(uint32_t id) const
    {
        uint32_t lo = 0, hi = n;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            const uint32_t base = FIELD == 0 ? g[mid].tile_base : g[mid].final_base;
            if (base <= id) lo = mid;
            else hi = mid;
        }
        return lo;
    }
    __device__ __forceinline__ const Geom2D &by_tile(uint32_t gt, uint32_t &lt) const
    {
        const Geom2D &r = g[find<0>(gt)];
        lt = gt - r.tile_base;
        return r;
    }
    __device__ __forceinline__ const Geom2D &by_final(uint32_t gw, uint32_t &lw) const
    {
        const Geom2D &r = g[find<1>(gw)];
        lw = gw - r.final_base;
        return r;
    }
    __device__ __forceinline__ const Geom2D &by_image(uint32_t i, uint32_t &li) const
    {
        li = i;
        return g[i];
    }
    uint32_t images() const { return n; }
};

// ---- 1. tile kernel -------------------------------------------------------------

template <uint32_t TBW, uint32_t TBH, class Src>
__global__ __launch_bounds__((TBW / SEG) * TBH) void bleq2d_tile(Src src)
{
    constexpr uint32_t SEGS = TBW / SEG;
    __shared__ uint32_t par[TBH * TBW];
    const SwizzleMap<SEG, (TBW / SEG) * TBH> at;   // conflict-free placement of the forest
    __shared__ uint32_t mw[TBH][SEGS + 2];   // 8 nibbles per thread, guard column on each side

    uint32_t lt;
    const Geom2D &G = src.by_tile(blockIdx.x, lt);
    const uint32_t h = G.h, w = G.w, bw = G.bw, bh = G.bh;
    const uint32_t t = threadIdx.x, r = t / SEGS, s = t % SEGS;
    const uint32_t tx = lt % G.tiles_x, ty = lt / G.tiles_x;
    const uint32_t bx0 = tx * TBW, by0 = ty * TBH;
    const uint32_t by = by0 + r, bxs = bx0 + s * SEG;
    if (lt == 0 && G.count) count_clear(G.slots);

    // masks of this thread's 8 blocks
    uint32_t m8 = 0;
    if (by < bh && bxs < bw) {
        const uint32_t y0 = 2 * by, x0 = 2 * bxs;
        const uint8_t *row0 = G.px + (size_t)y0 * G.pitch;
        const uint8_t *row1 = row0 + G.pitch;
        const bool two_rows = y0 + 1 < h;
        if ((G.flags & kVecIn) && two_rows && x0 + 16 <= w) {
            const uint4 a = *reinterpret_cast<const uint4 *>(row0 + x0);
            const uint4 b = *reinterpret_cast<const uint4 *>(row1 + x0);
            m8 = quad_masks(a.x, b.x) | (quad_masks(a.y, b.y) << 8) | (quad_masks(a.z, b.z) << 16) |
                 (quad_masks(a.w, b.w) << 24);
        } else {
            for (uint32_t q = 0; q < 4; ++q) {
                uint32_t tw = 0, bwd = 0;
                for (uint32_t j = 0; j < 4; ++j) {
                    const uint32_t x = x0 + 4 * q + j;
                    if (x < w) {
                        tw |= (uint32_t)(row0[x] != 0) << (8 * j);
                        if (two_rows) bwd |= (uint32_t)(row1[x] != 0) << (8 * j);
                    }
                }
                m8 |= quad_masks(tw, bwd) << (8 * q);
            }
        }
    }
    mw[r][s + 1] = m8;
    if (s == 0) {
        mw[r][0] = 0;
        mw[r][SEGS + 1] = 0;
    }

    // horizontal runs inside the thread's segment
    const uint32_t li0 = r * TBW + s * SEG;
    {
        uint32_t head = li0, prev = 0;
#pragma unroll
        for (uint32_t k = 0; k < SEG; ++k) {
            const uint32_t m = (m8 >> (4 * k)) & 15u, li = li0 + k;
            if (m && adj_left(m, prev)) {
                par[at(li)] = head;
            } else {
                par[at(li)] = li;
                head = li;
            }
            prev = m;
        }
    }
    __syncthreads();

    // unions with blocks of other threads inside the tile
    if (s > 0) {
        const uint32_t m0 = m8 & 15u, left = mw[r][s] >> 28;
        if (adj_left(m0, left)) lds_union(par, li0, li0 - 1, at);
    }
    if (r > 0) {
        const uint32_t up = mw[r - 1][s + 1];
        const uint32_t up_l = mw[r - 1][s] >> 28;         // upper column -1
        const uint32_t up_r = mw[r - 1][s + 2] & 15u;     // upper column 8
        const uint32_t ubase = (r - 1) * TBW + s * SEG;   // local index of upper column 0
        int last = -3;                                    // last upper column joined by the current run
        uint32_t prev = 0;
        for (int k = 0; k < (int)SEG; ++k) {
            const uint32_t m = (m8 >> (4 * k)) & 15u;
            if (!m) {
                prev = 0;
                last = -3;
                continue;
            }
            if (!adj_left(m, prev)) last = -3;            // a new run starts
            prev = m;
            for (int c = k - 1; c <= k + 1; ++c) {
                const uint32_t um = c < 0 ? up_l : (c >= (int)SEG ? up_r : ((up >> (4 * c)) & 15u));
                const bool adj = c < k ? adj_upleft(m, um) : (c == k ? adj_up(m, um) : adj_upright(m, um));
                if (!adj || c == last) continue;
                if (c == last + 1) {
                    // upper blocks last and c joined horizontally: same set already
                    const uint32_t lm = last < 0 ? up_l : ((up >> (4 * last)) & 15u);
                    if (adj_left(um, lm)) {
                        last = c;
                        continue;
                    }
                }
                lds_union(par, li0 + k, (uint32_t)((int)ubase + c), at);
                last = c;
            }
        }
    }
    __syncthreads();

    // resolve tile-local roots and write the block words
    if (by < bh && bxs < bw) {
        uint32_t out[SEG];
#pragma unroll
        for (uint32_t k = 0; k < SEG; ++k) {
            const uint32_t m = (m8 >> (4 * k)) & 15u, li = li0 + k;
            const uint32_t root = m ? lds_find(par, li, at) : li;
            const uint32_t g = (by0 + root / TBW) * bw + bx0 + root % TBW;
            out[k] = (g << 4) | m;
        }
        uint32_t *dst = G.P + (size_t)by * bw + bxs;
        if ((bw % 4u) == 0 && bxs + SEG <= bw) {
            reinterpret_cast<uint4 *>(dst)[0] = make_uint4(out[0], out[1], out[2], out[3]);
            reinterpret_cast<uint4 *>(dst)[1] = make_uint4(out[4], out[5], out[6], out[7]);
        } else {
            for (uint32_t k = 0; k < SEG; ++k)
                if (bxs + k < bw) dst[k] = out[k];
        }
    }
}

// ---- 2./3. tile-border kernels -----------------------------------------------------

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

// Consecutive border blocks of a tile usually repeat the same union (same
// tile-local roots on both sides); only the first lane of such a run unites.
template <uint32_t TBW, uint32_t TBH, class Src>
__global__ __launch_bounds__(256) void bleq2d_merge(Src src, uint32_t total_tiles)
{
    const Geom2D *G = nullptr;
    uint32_t bx = 0, by = 0, bx0 = 0, by0 = 0, vb = 0;
    const bool valid = border_job<TBW, TBH>(src, blockIdx.x * blockDim.x + threadIdx.x, total_tiles, G, bx, by,
                                            bx0, by0, vb);
    // earlier neighbours: left, up-left, up, up-right; distinct (own root, neighbour root) pairs
    uint32_t pr[4], np = 0;
    if (valid) {
        const uint32_t bw = G->bw, bh = G->bh, m = vb & 15u;
        const int ndx[4] = {-1, -1, 0, 1}, ndy[4] = {0, -1, -1, -1};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t nx = bx + ndx[k], ny = by + ndy[k];   // unsigned wrap -> out of range
            if (nx >= bw || ny >= bh) continue;
            if (nx - bx0 < TBW && ny - by0 < TBH) continue;    // same tile: done in the tile kernel
            const uint32_t vn = G->P[ny * bw + nx];
            const uint32_t nm = vn & 15u;
            const bool adj = k == 0 ? adj_left(m, nm)
                                    : (k == 1 ? adj_upleft(m, nm) : (k == 2 ? adj_up(m, nm) : adj_upright(m, nm)));
            if (adj && (np == 0 || pr[np - 1] != (vn >> 4))) pr[np++] = vn >> 4;
        }
    }
    const uint64_t key = np ? ((uint64_t)(vb >> 4) << 32) | pr[0] : 0;
    uint32_t leader;
    const bool lead = run_leader(key, valid ? G->tile_base : 0u, np != 0, leader);
    for (uint32_t i = lead ? 0u : 1u; i < np; ++i) Forest::unite(G->P, vb >> 4, pr[i]);   // tile-local roots
}

// Point each distinct tile-local root seen on the tile border (and the border
// block itself) at its final root.
template <uint32_t TBW, uint32_t TBH, class Src>
__global__ __launch_bounds__(256) void bleq2d_flatten(Src src, uint32_t total_tiles)
{
    const Geom2D *G = nullptr;
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
            root = Forest::find(P, r);
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

// Eight labels of one pixel row starting at column x0 (x0 % 8 == 0): the widest
// stores the row's actual address alignment allows (16, 8 or 4 bytes), so
// images of any width keep vector stores on most rows.
__device__ __forceinline__ void store_row8(uint32_t *dst, const uint32_t v[8], uint32_t x0, uint32_t w)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(dst);
    if (x0 + 8 <= w) {
        if ((a & 15u) == 0) {
            store_nt4(dst, v[0], v[1], v[2], v[3]);
            store_nt4(dst + 4, v[4], v[5], v[6], v[7]);
            return;
        }
        if ((a & 7u) == 0) {
#pragma unroll
            for (int i = 0; i < 8; i += 2) store_nt2(dst + i, v[i], v[i + 1]);
            return;
        }
    }
    for (uint32_t i = 0; i < 8 && x0 + i < w; ++i) dst[i] = v[i];
}

template <class Src>
__global__ __launch_bounds__(NT_FINAL) void bleq2d_final(Src src)
{
    uint32_t lw;
    const Geom2D &G = src.by_final(blockIdx.x, lw);
    const uint32_t h = G.h, w = G.w, bw = G.bw, bh = G.bh;
    const uint32_t by = lw / G.xgroups;
    const uint32_t bxs = ((lw % G.xgroups) * NT_FINAL + threadIdx.x) * FINAL_BLOCKS;
    const uint32_t *__restrict__ P = G.P;
    uint32_t roots = 0;
    if (by < bh && bxs < bw) {
        const uint32_t *src_w = P + (size_t)by * bw + bxs;
        uint32_t v[FINAL_BLOCKS];
        if ((reinterpret_cast<uintptr_t>(src_w) & 15u) == 0 && bxs + FINAL_BLOCKS <= bw) {
            const uint4 q = *reinterpret_cast<const uint4 *>(src_w);
            v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
            for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) v[k] = bxs + k < bw ? src_w[k] : 0u;
        }
        uint32_t lab[FINAL_BLOCKS];
        uint32_t prev_parent = 0xFFFFFFFFu, prev_root = 0;
        const uint32_t b0 = by * bw + bxs;
        const bool ref = (G.flags & kRefLabels) != 0;
#pragma unroll
        for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
            lab[k] = 0;
            if (!(v[k] & 15u)) continue;
            const uint32_t p = v[k] >> 4;
            const uint32_t root = p == prev_parent ? prev_root : Forest::find_final(P, p);
            prev_parent = p;
            prev_root = root;
            roots += root == b0 + k;
            // reference raw values: 1 + block index in the padded grid (LabelingAlgs.cl:202,317)
            lab[k] = ref ? 1u + (root % bw) + (root / bw) * G.ref_bw : root + 1u;
        }
        const uint32_t y0 = 2 * by, x0 = 2 * bxs;
        uint32_t *row0 = G.labels + (size_t)y0 * w + x0;
        uint32_t *row1 = row0 + w;
        const bool two_rows = y0 + 1 < h;
        uint32_t top[2 * FINAL_BLOCKS], bot[2 * FINAL_BLOCKS];
#pragma unroll
        for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
            const uint32_t m = v[k] & 15u;
            top[2 * k] = (m & 1u) ? lab[k] : 0u;
            top[2 * k + 1] = (m & 2u) ? lab[k] : 0u;
            bot[2 * k] = (m & 4u) ? lab[k] : 0u;
            bot[2 * k + 1] = (m & 8u) ? lab[k] : 0u;
        }
        store_row8(row0, top, x0, w);
        if (two_rows) store_row8(row1, bot, x0, w);
    }
    if (G.count) count_add(G.slots, roots);   // one image per workgroup: uniform branch
}

template <class Src>
__global__ __launch_bounds__(kCountSlots) void bleq2d_count(Src src)
{
    uint32_t li;
    const Geom2D &G = src.by_image(blockIdx.x, li);
    if (G.count) count_reduce(G.slots, G.count);
}

// ---- host side --------------------------------------------------------------------

template <uint32_t TBW, uint32_t TBH>
struct Variant {
    static constexpr uint32_t THREADS = (TBW / SEG) * TBH;

    template <class Src>
    static hipError_t launch(const Src &src, uint32_t total_tiles, uint32_t total_final, hipStream_t s,
                             KernelTimer *kt, bool flatten)
    {
        const uint32_t border_grid = (uint32_t)(((uint64_t)total_tiles * Border<TBW, TBH>::JOBS + 255) / 256);
        hipLaunchKernelGGL((bleq2d_tile<TBW, TBH, Src>), dim3(total_tiles), dim3(THREADS), 0, s, src);
        if (kt) kt->mark(s, "bleq2d_tile");
        hipLaunchKernelGGL((bleq2d_merge<TBW, TBH, Src>), dim3(border_grid), dim3(256), 0, s, src, total_tiles);
        if (kt) kt->mark(s, "bleq2d_merge");
        if (flatten) {
            hipLaunchKernelGGL((bleq2d_flatten<TBW, TBH, Src>), dim3(border_grid), dim3(256), 0, s, src, total_tiles);
            if (kt) kt->mark(s, "bleq2d_flatten");
        }
        hipLaunchKernelGGL((bleq2d_final<Src>), dim3(total_final), dim3(NT_FINAL), 0, s, src);
        if (kt) kt->mark(s, "bleq2d_final");
        if (src.images()) {
            hipLaunchKernelGGL((bleq2d_count<Src>), dim3(src.images()), dim3(kCountSlots), 0, s, src);
            if (kt) kt->mark(s, "bleq2d_count");
        }
        return hipGetLastError();
    }
};

template <class Src>
hipError_t dispatch(uint32_t tile_id, const Src &src, uint32_t total_tiles, uint32_t total_final, hipStream_t s,
                    KernelTimer *kt, bool flatten)
{
    switch (tile_id) {
    case 1: return Variant<64, 16>::launch(src, total_tiles, total_final, s, kt, flatten);
    case 2: return Variant<128, 32>::launch(src, total_tiles, total_final, s, kt, flatten);
    case 3: return Variant<256, 8>::launch(src, total_tiles, total_final, s, kt, flatten);
    case 4: return Variant<64, 32>::launch(src, total_tiles, total_final, s, kt, flatten);
    case 5: return Variant<128, 64>::launch(src, total_tiles, total_final, s, kt, flatten);
    case 6: return Variant<256, 32>::launch(src, total_tiles, total_final, s, kt, flatten);
    default: return Variant<128, 16>::launch(src, total_tiles, total_final, s, kt, flatten);
    }
}

}  // namespace

const TileShape2D kBleq2dTiles[kBleq2dTileCount] = {{128, 16}, {64, 16}, {128, 32}, {256, 8},
                                                    {64, 32},  {128, 64}, {256, 32}};

void bleq2d_geometry(Geom2D &g, uint32_t tile_id)
{
    const TileShape2D ts = kBleq2dTiles[tile_id < kBleq2dTileCount ? tile_id : 0];
    g.bw = (g.w + 1u) / 2;
    g.bh = (g.h + 1u) / 2;
    g.tiles_x = (g.bw + ts.bw - 1) / ts.bw;
    g.ntiles = g.tiles_x * ((g.bh + ts.bh - 1) / ts.bh);
    g.xgroups = (g.bw + NT_FINAL * FINAL_BLOCKS - 1) / (NT_FINAL * FINAL_BLOCKS);
    g.nfinal = g.xgroups * g.bh;
    g.ref_bw = (uint32_t)((((uint64_t)g.w >> 5) << 5) + 32) / 2;   // LabelingTools.cpp:103
    g.flags &= kRefLabels;
    if (g.pitch % 16 == 0 && reinterpret_cast<uintptr_t>(g.px) % 16 == 0) g.flags |= kVecIn;
    if (g.w % 8 == 0 && reinterpret_cast<uintptr_t>(g.labels) % 16 == 0) g.flags |= kVecOut;
}

size_t bleq2d_workspace(uint32_t h, uint32_t w)
{
    const uint64_t bw = (w + 1ull) / 2, bh = (h + 1ull) / 2;
    return align256(bw * bh * sizeof(uint32_t)) + kCountBytes;
}

hipError_t bleq2d_run(const Image2D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt)
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
    g.P = static_cast<uint32_t *>(ws);
    g.flags = opt.mode == LBL_LABELS_REFERENCE ? kRefLabels : 0u;
    bleq2d_geometry(g, opt.tile2d);
    g.slots = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(ws) + align256((size_t)g.bw * g.bh * 4));
    if (kt) kt->begin(s);
    return dispatch(opt.tile2d, one, g.ntiles, g.nfinal, s, kt, opt.flatten);
}

hipError_t bleq2d_run_batch(const Geom2D *d_geoms, uint32_t n, uint32_t total_tiles, uint32_t total_final,
                            uint32_t tile_id, bool flatten, hipStream_t s, KernelTimer *kt)
{
    Many many{d_geoms, n};
    if (kt) kt->begin(s);
    return dispatch(tile_id, many, total_tiles, total_final, s, kt, flatten);
}

}  // namespace lbl
