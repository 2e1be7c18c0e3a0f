// bleq2d.hip -- 2D block-equivalence labeling (2x2 blocks, 8-connectivity) for
// gfx950.
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
//   1. tile    one workgroup per 128x16-block tile (256x32 px): 16-byte row
//              loads -> 4-bit block masks; union-find of the tile's blocks in
//              LDS; writes one packed word per block, parent << 4 | mask, whose
//              parent is the tile-local root (global block index).
//   2. merge   one thread per block on a tile border: unions with each
//              8-neighbour block in another tile that it touches (atomicMin
//              links in the global forest).
//   3. flatten one thread per tile-border block: points the block's tile-local
//              root straight at its global root.
//   4. final   one thread per 4 blocks: root of the block (<= 2 hops), writes
//              the 8 labels of the 2x2 pixels of each block, 0 for background,
//              and counts roots.
//
// HBM traffic per pixel: 1 B image + 1 B block words in (1), 1 B block words +
// 4 B labels in (4), plus the border blocks in (2)-(3).
#include "lbl_internal.hpp"
#include "uf_device.hpp"

namespace lbl {
namespace {

constexpr uint32_t TBW = 128;            // tile width in blocks
constexpr uint32_t TBH = 16;             // tile height in blocks
constexpr uint32_t SEG = 8;              // blocks per thread in the tile kernel
constexpr uint32_t SEGS = TBW / SEG;     // threads per block row
constexpr uint32_t NT_TILE = SEGS * TBH; // 256
constexpr uint32_t JOBS = 2 * TBW + 2 * (TBH - 2);  // border blocks per tile
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

// ---- 1. tile kernel ---------------------------------------------------------

template <bool VEC>
__global__ __launch_bounds__(NT_TILE) void bleq2d_tile(const uint8_t *__restrict__ px, size_t pitch,
                                                        uint32_t h, uint32_t w, uint32_t bw, uint32_t bh,
                                                        uint32_t tiles_x, uint32_t *__restrict__ P,
                                                        uint32_t *count)
{
    __shared__ uint32_t par[TBH * TBW];
    __shared__ uint32_t mw[TBH][SEGS + 2];   // 8 nibbles per thread, guard column on each side

    const uint32_t t = threadIdx.x, r = t / SEGS, s = t % SEGS;
    const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const uint32_t bx0 = tx * TBW, by0 = ty * TBH;
    const uint32_t by = by0 + r, bxs = bx0 + s * SEG;
    if (blockIdx.x == 0 && t == 0 && count) *count = 0;

    // masks of this thread's 8 blocks
    uint32_t m8 = 0;
    if (by < bh && bxs < bw) {
        const uint32_t y0 = 2 * by, x0 = 2 * bxs;
        const uint8_t *row0 = px + (size_t)y0 * pitch;
        const uint8_t *row1 = row0 + pitch;
        const bool two_rows = y0 + 1 < h;
        if (VEC && two_rows && x0 + 16 <= w) {
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
                par[li] = head;
            } else {
                par[li] = li;
                head = li;
            }
            prev = m;
        }
    }
    __syncthreads();

    // unions with blocks of other threads inside the tile
    if (s > 0) {
        const uint32_t m0 = m8 & 15u, left = mw[r][s] >> 28;
        if (adj_left(m0, left)) lds_union(par, li0, li0 - 1);
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
                lds_union(par, li0 + k, (uint32_t)((int)ubase + c));
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
            const uint32_t root = m ? lds_find(par, li) : li;
            const uint32_t g = (by0 + root / TBW) * bw + bx0 + root % TBW;
            out[k] = (g << 4) | m;
        }
        uint32_t *dst = P + (size_t)by * bw + bxs;
        if ((bw % 4u) == 0 && bxs + SEG <= bw) {
            reinterpret_cast<uint4 *>(dst)[0] = make_uint4(out[0], out[1], out[2], out[3]);
            reinterpret_cast<uint4 *>(dst)[1] = make_uint4(out[4], out[5], out[6], out[7]);
        } else {
            for (uint32_t k = 0; k < SEG; ++k)
                if (bxs + k < bw) dst[k] = out[k];
        }
    }
}

// ---- 2./3. tile-border kernels -----------------------------------------------

// j-th border block of a tile: top row, bottom row, then the left and right
// columns without their corner blocks.
__device__ __forceinline__ void border_block(uint32_t j, uint32_t bx0, uint32_t by0, uint32_t &bx, uint32_t &by)
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

__global__ __launch_bounds__(256) void bleq2d_merge(uint32_t *P, uint32_t bw, uint32_t bh, uint32_t tiles_x,
                                                     uint32_t ntiles)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t tile = g / JOBS, j = g % JOBS;
    if (tile >= ntiles) return;
    const uint32_t bx0 = (tile % tiles_x) * TBW, by0 = (tile / tiles_x) * TBH;
    uint32_t bx, by;
    border_block(j, bx0, by0, bx, by);
    if (bx >= bw || by >= bh) return;
    const uint32_t b = by * bw + bx;
    const uint32_t vb = P[b];
    const uint32_t m = vb & 15u;
    if (!m) return;
    // earlier neighbours: left, up-left, up, up-right
    const int ndx[4] = {-1, -1, 0, 1}, ndy[4] = {0, -1, -1, -1};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t nx = bx + ndx[k], ny = by + ndy[k];   // unsigned wrap -> out of range
        if (nx >= bw || ny >= bh) continue;
        if (nx - bx0 < TBW && ny - by0 < TBH) continue;    // same tile: done in the tile kernel
        const uint32_t n = ny * bw + nx;
        const uint32_t vn = P[n];
        const uint32_t nm = vn & 15u;
        const bool adj = k == 0 ? adj_left(m, nm) : (k == 1 ? adj_upleft(m, nm) : (k == 2 ? adj_up(m, nm) : adj_upright(m, nm)));
        if (adj) Forest::unite(P, vb >> 4, vn >> 4);   // tile-local roots
    }
}

__global__ __launch_bounds__(256) void bleq2d_flatten(uint32_t *P, uint32_t bw, uint32_t bh, uint32_t tiles_x,
                                                       uint32_t ntiles)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t tile = g / JOBS, j = g % JOBS;
    if (tile >= ntiles) return;
    const uint32_t bx0 = (tile % tiles_x) * TBW, by0 = (tile / tiles_x) * TBH;
    uint32_t bx, by;
    border_block(j, bx0, by0, bx, by);
    if (bx >= bw || by >= bh) return;
    const uint32_t b = by * bw + bx;
    if (!(P[b] & 15u)) return;
    Forest::compress(P, b);
}

// ---- 4. final kernel -----------------------------------------------------------

__device__ __forceinline__ uint32_t label_of(uint32_t root, uint32_t mode, uint32_t bw, uint32_t ref_bw)
{
    if (mode == LBL_LABELS_REFERENCE) return 1u + (root % bw) + (root / bw) * ref_bw;
    return root + 1u;
}

template <bool VEC>
__global__ __launch_bounds__(NT_FINAL) void bleq2d_final(const uint32_t *__restrict__ P, uint32_t *__restrict__ labels,
                                                          uint32_t h, uint32_t w, uint32_t bw, uint32_t bh,
                                                          uint32_t xgroups, uint32_t *count, uint32_t mode,
                                                          uint32_t ref_bw)
{
    const uint32_t by = blockIdx.x / xgroups;
    const uint32_t bxs = ((blockIdx.x % xgroups) * NT_FINAL + threadIdx.x) * FINAL_BLOCKS;
    uint32_t roots = 0;
    if (by < bh && bxs < bw) {
        const uint32_t *src = P + (size_t)by * bw + bxs;
        uint32_t v[FINAL_BLOCKS];
        if (VEC && bxs + FINAL_BLOCKS <= bw) {
            const uint4 q = *reinterpret_cast<const uint4 *>(src);
            v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
            for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) v[k] = bxs + k < bw ? src[k] : 0u;
        }
        uint32_t lab[FINAL_BLOCKS];
        uint32_t prev_parent = 0xFFFFFFFFu, prev_root = 0;
        const uint32_t b0 = by * bw + bxs;
#pragma unroll
        for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
            lab[k] = 0;
            if (!(v[k] & 15u)) continue;
            const uint32_t p = v[k] >> 4;
            const uint32_t root = p == prev_parent ? prev_root : Forest::find_final(P, p);
            prev_parent = p;
            prev_root = root;
            roots += root == b0 + k;
            lab[k] = label_of(root, mode, bw, ref_bw);
        }
        const uint32_t y0 = 2 * by, x0 = 2 * bxs;
        uint32_t *row0 = labels + (size_t)y0 * w + x0;
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
        if (VEC && two_rows && x0 + 2 * FINAL_BLOCKS <= w) {
            store_nt4(row0, top[0], top[1], top[2], top[3]);
            store_nt4(row0 + 4, top[4], top[5], top[6], top[7]);
            store_nt4(row1, bot[0], bot[1], bot[2], bot[3]);
            store_nt4(row1 + 4, bot[4], bot[5], bot[6], bot[7]);
        } else {
            for (uint32_t i = 0; i < 2 * FINAL_BLOCKS; ++i) {
                if (x0 + i >= w) break;
                row0[i] = top[i];
                if (two_rows) row1[i] = bot[i];
            }
        }
    }
    if (count) wave_add_atomic(count, roots);
}

}  // namespace

size_t bleq2d_workspace(uint32_t h, uint32_t w)
{
    const uint64_t bw = (w + 1ull) / 2, bh = (h + 1ull) / 2;
    return align256(bw * bh * sizeof(uint32_t));
}

hipError_t bleq2d_run(const Image2D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt)
{
    const uint32_t bw = (im.w + 1u) / 2, bh = (im.h + 1u) / 2;
    const uint32_t tiles_x = (bw + TBW - 1) / TBW, tiles_y = (bh + TBH - 1) / TBH;
    const uint32_t ntiles = tiles_x * tiles_y;
    uint32_t *P = static_cast<uint32_t *>(ws);
    const bool vec_in = (im.pitch % 16 == 0) && (reinterpret_cast<uintptr_t>(im.px) % 16 == 0);
    const bool vec_out = (im.w % 8 == 0) && (reinterpret_cast<uintptr_t>(im.labels) % 16 == 0);
    const uint32_t ref_bw = (uint32_t)((((uint64_t)im.w >> 5) << 5) + 32) / 2;   // LabelingTools.cpp:103
    const uint32_t border_grid = (uint32_t)(((uint64_t)ntiles * JOBS + 255) / 256);
    const uint32_t xgroups = (bw + NT_FINAL * FINAL_BLOCKS - 1) / (NT_FINAL * FINAL_BLOCKS);

    if (kt) kt->begin(s);
    if (vec_in)
        hipLaunchKernelGGL(bleq2d_tile<true>, dim3(ntiles), dim3(NT_TILE), 0, s, im.px, im.pitch, im.h, im.w, bw,
                           bh, tiles_x, P, im.count);
    else
        hipLaunchKernelGGL(bleq2d_tile<false>, dim3(ntiles), dim3(NT_TILE), 0, s, im.px, im.pitch, im.h, im.w, bw,
                           bh, tiles_x, P, im.count);
    if (kt) kt->mark(s, "bleq2d_tile");
    hipLaunchKernelGGL(bleq2d_merge, dim3(border_grid), dim3(256), 0, s, P, bw, bh, tiles_x, ntiles);
    if (kt) kt->mark(s, "bleq2d_merge");
    hipLaunchKernelGGL(bleq2d_flatten, dim3(border_grid), dim3(256), 0, s, P, bw, bh, tiles_x, ntiles);
    if (kt) kt->mark(s, "bleq2d_flatten");
    const dim3 fgrid((uint32_t)((uint64_t)xgroups * bh));
    if (vec_out)
        hipLaunchKernelGGL(bleq2d_final<true>, fgrid, dim3(NT_FINAL), 0, s, P, im.labels, im.h, im.w, bw, bh,
                           xgroups, im.count, (uint32_t)opt.mode, ref_bw);
    else
        hipLaunchKernelGGL(bleq2d_final<false>, fgrid, dim3(NT_FINAL), 0, s, P, im.labels, im.h, im.w, bw, bh,
                           xgroups, im.count, (uint32_t)opt.mode, ref_bw);
    if (kt) kt->mark(s, "bleq2d_final");
    return hipGetLastError();
}

}  // namespace lbl
