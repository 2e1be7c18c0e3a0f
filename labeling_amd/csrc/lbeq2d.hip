// lbeq2d.hip -- 2D label equivalence (Kalentev et al.) on pixels, 4- or
// 8-connectivity, for gfx950.
//
// Same problem as the reference's TOCLLabelDistribution (LabelingAlgs.cpp:867-942,
// kernels DistrInitKernel / DistrScanKernel / DistrAnalizeKernel,
// LabelingAlgs.cl:31-130): every foreground pixel starts with its own label, a
// scan pass lowers the label of a pixel's root to the smallest neighbouring
// label, an analyze pass points every pixel at its root, and the two repeat
// until a scan changes nothing.  The label image itself is the equivalence
// array.  Here the fixed point is reached per tile in LDS, and only the
// equivalences that cross tile borders go through global memory:
//
//   1. tile    one workgroup per 128x32-pixel tile: 16-byte row loads; Kalentev
//              scan/analyze sweeps in LDS until no change (workgroup vote);
//              writes label = 1 + global index of the tile-local root.
//   2. merge   one thread per tile-border pixel: atomicMin union with the
//              earlier neighbours that lie in other tiles.
//   3. flatten one thread per tile-border pixel: the pixel's tile-local root is
//              pointed straight at its global root.
//   4. final   the analyze pass over the whole image: label = label of the root
//              (at most one extra read), rewritten only where it changed;
//              counts roots.
//
// Labels: 1 + raster index of the component's minimum pixel (1-based like the
// reference's 3D variant, LabelingAlgs.cl:641; the 2D reference is 0-based and
// loses pixel (0,0), LabelingAlgs.cl:37,48-50).
//
// These kernels run 4-connectivity (and 8-connectivity beyond the 2^28-block
// limit of the block forest); 8-connectivity otherwise runs on the bleq block
// pipeline with a first-pixel key pass (bleq2d.hip), which gives the same
// labels.
#include <algorithm>

#include "lbl_internal.hpp"
#include "canon.hpp"
#include "uf_device.hpp"

namespace lbl {
namespace {

constexpr uint32_t TW = 128;   // tile width, pixels
constexpr uint32_t TH = 32;    // tile height, pixels
constexpr uint32_t NT = 256;
constexpr uint32_t JOBS = 2 * TW + 2 * (TH - 2);

// ---- 1. tile kernel -----------------------------------------------------------

template <bool VEC, int CONN>
__global__ __launch_bounds__(NT) void lbeq2d_tile(const uint8_t *__restrict__ px, size_t pitch, uint32_t h,
                                                  uint32_t w, uint32_t tiles_x, uint32_t *__restrict__ L,
                                                  uint32_t *slots, uint32_t cap, uint32_t *status)
{
    __shared__ uint32_t S[TH * TW];
    const uint32_t t = threadIdx.x;
    const uint32_t x0 = (blockIdx.x % tiles_x) * TW, y0 = (blockIdx.x / tiles_x) * TH;
    if (blockIdx.x == 0 && slots) count_clear(slots);

    // load: thread -> 16 consecutive pixels of one row; init label = local index + 1
    {
        const uint32_t r = t / (TW / 16), xs = (t % (TW / 16)) * 16;
        const uint32_t y = y0 + r, gx = x0 + xs;
        uint32_t fg = 0;   // bit k = pixel xs + k is foreground
        if (y < h) {
            const uint8_t *row = px + (size_t)y * pitch;
            if (VEC && gx + 16 <= w) {
                const uint4 a = *reinterpret_cast<const uint4 *>(row + gx);
                const uint32_t q[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t nz = nz_bytes(q[i]);
                    fg |= ((nz & 1u) | ((nz >> 7) & 2u) | ((nz >> 14) & 4u) | ((nz >> 21) & 8u)) << (4 * i);
                }
            } else {
                for (uint32_t k = 0; k < 16; ++k)
                    if (gx + k < w && row[gx + k]) fg |= 1u << k;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) {
            const uint32_t li = r * TW + xs + k;
            S[li] = ((fg >> k) & 1u) ? li + 1 : 0u;
        }
    }
    __syncthreads();

    // Kalentev sweeps; thread -> column cx of rows ry, ry + 2, ... (conflict-free LDS rows)
    const uint32_t cx = t % TW, ry = t / TW;
    uint32_t iters = 0;
    for (;;) {
        int changed = 0;
#pragma unroll 4
        for (uint32_t j = 0; j < TH / 2; ++j) {
            const uint32_t y = ry + 2 * j, li = y * TW + cx;
            const uint32_t lab = lds_load(&S[li]);
            if (!lab) continue;
            uint32_t m = lab, v;
            if (cx > 0 && (v = lds_load(&S[li - 1])) && v < m) m = v;
            if (cx + 1 < TW && (v = lds_load(&S[li + 1])) && v < m) m = v;
            if (y > 0 && (v = lds_load(&S[li - TW])) && v < m) m = v;
            if (y + 1 < TH && (v = lds_load(&S[li + TW])) && v < m) m = v;
            if (CONN == 8) {
                if (y > 0 && cx > 0 && (v = lds_load(&S[li - TW - 1])) && v < m) m = v;
                if (y > 0 && cx + 1 < TW && (v = lds_load(&S[li - TW + 1])) && v < m) m = v;
                if (y + 1 < TH && cx > 0 && (v = lds_load(&S[li + TW - 1])) && v < m) m = v;
                if (y + 1 < TH && cx + 1 < TW && (v = lds_load(&S[li + TW + 1])) && v < m) m = v;
            }
            if (m < lab) {
                lds_min(&S[lab - 1], m);
                changed = 1;
            }
        }
        if (!__syncthreads_or(changed)) break;
        if (cap && ++iters >= cap) {
            if (t == 0 && status) atomicOr(status, 1u);
            break;
        }
#pragma unroll 4
        for (uint32_t j = 0; j < TH / 2; ++j) {
            const uint32_t li = (ry + 2 * j) * TW + cx;
            uint32_t lab = lds_load(&S[li]);
            if (!lab) continue;
            uint32_t cur = lds_load(&S[lab - 1]);
            while (cur != lab) {
                lab = lds_load(&S[cur - 1]);
                cur = lds_load(&S[lab - 1]);
            }
            lds_store(&S[li], lab);
        }
        __syncthreads();
    }

    // write 1 + global index of the tile-local root
    for (uint32_t j = 0; j < TH / 2; ++j) {
        const uint32_t y = ry + 2 * j, gy = y0 + y, gx = x0 + cx;
        if (gy >= h || gx >= w) continue;
        const uint32_t lab = S[y * TW + cx];
        uint32_t out = 0;
        if (lab) {
            const uint32_t lr = lab - 1;
            out = (y0 + lr / TW) * w + x0 + lr % TW + 1u;
        }
        L[(size_t)gy * w + gx] = out;
    }
}

// ---- 2./3. tile-border kernels -------------------------------------------------

__device__ __forceinline__ void border_pixel(uint32_t j, uint32_t x0, uint32_t y0, uint32_t &x, uint32_t &y)
{
    if (j < TW) {
        x = x0 + j;
        y = y0;
    } else if (j < 2 * TW) {
        x = x0 + (j - TW);
        y = y0 + TH - 1;
    } else if (j < 2 * TW + (TH - 2)) {
        x = x0;
        y = y0 + 1 + (j - 2 * TW);
    } else {
        x = x0 + TW - 1;
        y = y0 + 1 + (j - 2 * TW - (TH - 2));
    }
}

template <int CONN>
__global__ __launch_bounds__(256) void lbeq2d_merge(uint32_t *L, uint32_t h, uint32_t w, uint32_t tiles_x,
                                                     uint32_t ntiles)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t tile = g / JOBS, j = g % JOBS;
    if (tile >= ntiles) return;
    const uint32_t x0 = (tile % tiles_x) * TW, y0 = (tile / tiles_x) * TH;
    uint32_t x, y;
    border_pixel(j, x0, y0, x, y);
    if (x >= w || y >= h) return;
    const uint32_t p = y * w + x;
    const uint32_t vp = L[p];
    if (!vp) return;
    // earlier neighbours: left, up (4-conn); up-left, up-right (8-conn)
    const int ndx[4] = {-1, 0, -1, 1}, ndy[4] = {0, -1, -1, -1};
#pragma unroll
    for (int k = 0; k < (CONN == 8 ? 4 : 2); ++k) {
        const uint32_t nx = x + ndx[k], ny = y + ndy[k];
        if (nx >= w || ny >= h) continue;
        if (nx - x0 < TW && ny - y0 < TH) continue;
        const uint32_t vn = L[ny * w + nx];
        if (vn) LabelForest::unite(L, vp - 1u, vn - 1u);   // tile-local roots
    }
}

__global__ __launch_bounds__(256) void lbeq2d_flatten(uint32_t *L, uint32_t h, uint32_t w, uint32_t tiles_x,
                                                       uint32_t ntiles)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t tile = g / JOBS, j = g % JOBS;
    if (tile >= ntiles) return;
    const uint32_t x0 = (tile % tiles_x) * TW, y0 = (tile / tiles_x) * TH;
    uint32_t x, y;
    border_pixel(j, x0, y0, x, y);
    if (x >= w || y >= h) return;
    const uint32_t p = y * w + x;
    if (L[p]) LabelForest::compress(L, p);
}

// ---- 4. final analyze pass -------------------------------------------------------

// MARK: also write the canonical key bits, one bit per pixel set at roots (a
// root is its component's minimum pixel index, i.e. its first pixel); the 16
// lanes holding the 64 labels of one bits word assemble it.
template <bool VEC, bool MARK>
__global__ __launch_bounds__(256) void lbeq_final(uint32_t *__restrict__ L, uint64_t n, uint32_t *slots,
                                                  unsigned long long *bits)
{
    const uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    uint32_t roots = 0, rmask = 0;
    if (i0 < n) {
        uint32_t v[4];
        const bool full = VEC && i0 + 4 <= n;
        if (full) {
            const uint4 q = *reinterpret_cast<const uint4 *>(L + i0);
            v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
            for (uint32_t k = 0; k < 4; ++k) v[k] = i0 + k < n ? L[i0 + k] : 0u;
        }
        bool dirty = false;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            if (!v[k]) continue;
            if ((uint64_t)v[k] == i0 + k + 1) {
                ++roots;
                rmask |= 1u << k;
                continue;
            }
            uint32_t r = v[k], vr = L[r - 1u];
            while (vr != r) {   // one hop after the flatten pass
                r = vr;
                vr = L[r - 1u];
            }
            if (r != v[k]) {
                v[k] = r;
                dirty = true;
            }
        }
        if (dirty) {
            if (full) *reinterpret_cast<uint4 *>(L + i0) = make_uint4(v[0], v[1], v[2], v[3]);
            else
                for (uint32_t k = 0; k < 4; ++k)
                    if (i0 + k < n) L[i0 + k] = v[k];
        }
    }
    if (MARK) {
        const uint32_t sh = (uint32_t)(i0 & 63u);   // lanes 16k .. 16k+15 share a word
        uint32_t lo = sh < 32 ? rmask << sh : 0u, hi = sh >= 32 ? rmask << (sh - 32) : 0u;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            lo |= __shfl_xor(lo, off, 64);
            hi |= __shfl_xor(hi, off, 64);
        }
        if (sh == 0 && i0 < n) bits[i0 >> 6] = ((unsigned long long)hi << 32) | lo;
    }
    if (slots) count_add(slots, roots);
}

// Canonical values in place of root labels (every label is its root's).
template <bool VEC>
__global__ __launch_bounds__(256) void lbeq_relabel(uint32_t *__restrict__ L, uint64_t n, Canon c)
{
    const uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i0 >= n) return;
    uint32_t v[4];
    const bool full = VEC && i0 + 4 <= n;
    if (full) {
        const uint4 q = *reinterpret_cast<const uint4 *>(L + i0);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
        for (uint32_t k = 0; k < 4; ++k) v[k] = i0 + k < n ? L[i0 + k] : 0u;
    }
    uint32_t prev = 0, prev_lab = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (!v[k]) continue;
        if (v[k] != prev) {
            prev = v[k];
            prev_lab = canon_label(c, v[k] - 1u);
        }
        v[k] = prev_lab;
    }
    if (full) *reinterpret_cast<uint4 *>(L + i0) = make_uint4(v[0], v[1], v[2], v[3]);
    else
        for (uint32_t k = 0; k < 4; ++k)
            if (i0 + k < n) L[i0 + k] = v[k];
}

}  // namespace

size_t lbeq2d_workspace(uint32_t h, uint32_t w, bool canonical)
{
    const size_t own = align256(kCountBytes) + (canonical ? canon_workspace((uint64_t)h * w, 0) : 0);
    // both connectivities run on block forests (below the block limit)
    return std::max({own, bleq2d_workspace(h, w, true), bleq2d_conn4_workspace(h, w, canonical)});
}

// Shared with the 3D label-equivalence path: the analyze pass over n labels,
// then the component count (slots -> *count) when count is not null, then,
// with `canon_ws` (canonical numbering state after the count slots), the
// canonical values.
hipError_t lbeq_final_launch(uint32_t *labels, uint64_t n, uint32_t *count, uint32_t *slots, void *canon_ws,
                             hipStream_t s, KernelTimer *kt)
{
    const uint64_t threads = (n + 3) / 4;
    const dim3 grid((uint32_t)((threads + 255) / 256));
    uint32_t *sl = count ? slots : nullptr;
    const bool vec = reinterpret_cast<uintptr_t>(labels) % 16 == 0;
    const Canon c = canon_ws ? canon_layout(canon_ws, n, 0) : Canon{};
    if (canon_ws) {
        if (vec) hipLaunchKernelGGL((lbeq_final<true, true>), grid, dim3(256), 0, s, labels, n, sl, c.bits);
        else hipLaunchKernelGGL((lbeq_final<false, true>), grid, dim3(256), 0, s, labels, n, sl, c.bits);
    } else {
        if (vec) hipLaunchKernelGGL((lbeq_final<true, false>), grid, dim3(256), 0, s, labels, n, sl, nullptr);
        else hipLaunchKernelGGL((lbeq_final<false, false>), grid, dim3(256), 0, s, labels, n, sl, nullptr);
    }
    if (kt) kt->mark(s, "lbeq_final");
    if (count) {
        hipLaunchKernelGGL(count_reduce_kernel<0>, dim3(1), dim3(kCountSlots), 0, s, slots, count);
        if (kt) kt->mark(s, "count_reduce");
    }
    if (canon_ws) {
        const hipError_t e = canon_scan_launch(c, s, kt);
        if (e != hipSuccess) return e;
        if (vec) hipLaunchKernelGGL(lbeq_relabel<true>, grid, dim3(256), 0, s, labels, n, c);
        else hipLaunchKernelGGL(lbeq_relabel<false>, grid, dim3(256), 0, s, labels, n, c);
        if (kt) kt->mark(s, "lbeq_relabel");
    }
    return hipGetLastError();
}

hipError_t lbeq2d_run(const Image2D &im, int conn, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt)
{
    // 8-connectivity: the pixels of a 2x2 block are always connected, so the
    // bleq block forest has the same components; its first-pixel key pass gives
    // the label equivalence values (1 + minimum pixel index).  The sweeps below
    // run for the iteration-cap debug knob, and beyond the block forests'
    // 2^28-block limit.
    if (conn == 8 && opt.iter_cap == 0 && ((im.h + 1ull) / 2) * ((im.w + 1ull) / 2) <= kMaxBlocks2D)
        return bleq2d_run(im, opt, ws, s, kt, true);
    // 4-connectivity: 1x2 blocks, whose two pixels are 4-adjacent
    if (conn == 4 && opt.iter_cap == 0 && (uint64_t)im.h * ((im.w + 1ull) / 2) <= kMaxBlocks2D)
        return bleq2d_run_conn4(im, opt, ws, s, kt);
    const uint32_t tiles_x = (im.w + TW - 1) / TW, tiles_y = (im.h + TH - 1) / TH;
    const uint32_t ntiles = tiles_x * tiles_y;
    const bool vec = (im.pitch % 16 == 0) && (reinterpret_cast<uintptr_t>(im.px) % 16 == 0);
    const uint32_t border_grid = (uint32_t)(((uint64_t)ntiles * JOBS + 255) / 256);
    uint32_t *slots = static_cast<uint32_t *>(ws);

    if (kt) kt->begin(s);
    auto tile = conn == 8 ? (vec ? lbeq2d_tile<true, 8> : lbeq2d_tile<false, 8>)
                          : (vec ? lbeq2d_tile<true, 4> : lbeq2d_tile<false, 4>);
    hipLaunchKernelGGL(tile, dim3(ntiles), dim3(NT), 0, s, im.px, im.pitch, im.h, im.w, tiles_x, im.labels,
                       im.count ? slots : nullptr, opt.iter_cap, opt.status);
    if (kt) kt->mark(s, "lbeq2d_tile");
    if (conn == 8)
        hipLaunchKernelGGL(lbeq2d_merge<8>, dim3(border_grid), dim3(256), 0, s, im.labels, im.h, im.w, tiles_x, ntiles);
    else
        hipLaunchKernelGGL(lbeq2d_merge<4>, dim3(border_grid), dim3(256), 0, s, im.labels, im.h, im.w, tiles_x, ntiles);
    if (kt) kt->mark(s, "lbeq2d_merge");
    hipLaunchKernelGGL(lbeq2d_flatten, dim3(border_grid), dim3(256), 0, s, im.labels, im.h, im.w, tiles_x, ntiles);
    if (kt) kt->mark(s, "lbeq2d_flatten");
    void *canon_ws = opt.mode == LBL_LABELS_CANONICAL ? static_cast<uint8_t *>(ws) + align256(kCountBytes) : nullptr;
    hipError_t e = lbeq_final_launch(im.labels, (uint64_t)im.h * im.w, im.count, slots, canon_ws, s, kt);
    return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace lbl
