// lbeq3d.hip -- 3D label equivalence (Kalentev) on voxels, 26-connectivity,
// for gfx950.
//
// Same problem as the reference's TOCLLabelEquivalence3D (LabelingAlgs.cpp:1324-1396,
// kernels LBEQ3D_InitKernel / LBEQ3D_ScanKernel / LBEQ3D_AnalyzeKernel,
// LabelingAlgs.cl:635-734): 1-based voxel labels, scan = minimum over the 26
// neighbours stored into the label's root, analyze = pointer chase, repeated to
// a fixed point.  As in 2D the fixed point is reached per tile in LDS and only
// the equivalences crossing tile faces use global atomics:
//
//   1. tile    one workgroup per 64x8x8-voxel tile: Kalentev sweeps in LDS,
//              label = 1 + global index of the tile-local root.
//   2. merge   one thread per tile-face voxel: unions with the earlier
//              26-neighbours in other tiles.
//   3. flatten one thread per tile-face voxel: tile-local root -> global root.
//   4. final   the analyze pass over the volume (shared with 2D).
#include "lbl_internal.hpp"
#include "canon.hpp"
#include "uf_device.hpp"

namespace lbl {

hipError_t lbeq_final_launch(uint32_t *labels, uint64_t n, uint32_t *count, uint32_t *slots, void *canon_ws,
                             hipStream_t s, KernelTimer *kt);

namespace {

constexpr uint32_t TX = 64, TY = 8, TZ = 8;   // tile, voxels
constexpr uint32_t NT = 256;
constexpr uint32_t TILE = TX * TY * TZ;       // 4096
constexpr uint32_t ROWS = TY * TZ;            // 64 rows of TX voxels
constexpr uint32_t FACE_Z = TX * TY, FACE_Y = TX * (TZ - 2), FACE_X = (TY - 2) * (TZ - 2);
constexpr uint32_t JOBS = 2 * FACE_Z + 2 * FACE_Y + 2 * FACE_X;

template <bool VEC>
__global__ __launch_bounds__(NT) void lbeq3d_tile(const uint8_t *__restrict__ vox, uint32_t d, uint32_t h, uint32_t w,
                                                  uint32_t tiles_x, uint32_t tiles_y, uint32_t *__restrict__ L,
                                                  uint32_t *slots, uint32_t cap, uint32_t *status)
{
    __shared__ uint32_t S[TILE];
    const uint32_t t = threadIdx.x;
    const uint32_t x0 = (blockIdx.x % tiles_x) * TX, y0 = ((blockIdx.x / tiles_x) % tiles_y) * TY;
    const uint32_t z0 = (blockIdx.x / (tiles_x * tiles_y)) * TZ;
    if (blockIdx.x == 0 && slots) count_clear(slots);

    {   // load: thread -> 16 voxels of one row
        const uint32_t row = t / (TX / 16), xs = (t % (TX / 16)) * 16;
        const uint32_t ly = row % TY, lz = row / TY;
        const uint32_t y = y0 + ly, z = z0 + lz, gx = x0 + xs;
        uint32_t fg = 0;
        if (y < h && z < d) {
            const uint8_t *src = vox + ((size_t)z * h + y) * w;
            if (VEC && gx + 16 <= w) {
                const uint4 a = *reinterpret_cast<const uint4 *>(src + gx);
                const uint32_t q[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t nz = nz_bytes(q[i]);
                    fg |= ((nz & 1u) | ((nz >> 7) & 2u) | ((nz >> 14) & 4u) | ((nz >> 21) & 8u)) << (4 * i);
                }
            } else {
                for (uint32_t k = 0; k < 16; ++k)
                    if (gx + k < w && src[gx + k]) fg |= 1u << k;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) {
            const uint32_t li = row * TX + xs + k;
            S[li] = ((fg >> k) & 1u) ? li + 1 : 0u;
        }
    }
    __syncthreads();

    const uint32_t cx = t % TX, q0 = t / TX;   // column; rows q0, q0 + 4, ...
    uint32_t iters = 0;
    for (;;) {
        int changed = 0;
        for (uint32_t j = 0; j < ROWS / (NT / TX); ++j) {
            const uint32_t q = q0 + (NT / TX) * j, ly = q % TY, lz = q / TY, li = q * TX + cx;
            const uint32_t lab = lds_load(&S[li]);
            if (!lab) continue;
            uint32_t m = lab;
            for (int dz = -1; dz <= 1; ++dz) {
                if ((int)lz + dz < 0 || (int)lz + dz >= (int)TZ) continue;
                for (int dy = -1; dy <= 1; ++dy) {
                    if ((int)ly + dy < 0 || (int)ly + dy >= (int)TY) continue;
                    const uint32_t rb = (uint32_t)((int)li + (dz * (int)TY + dy) * (int)TX);
#pragma unroll
                    for (int dx = -1; dx <= 1; ++dx) {
                        if ((int)cx + dx < 0 || (int)cx + dx >= (int)TX) continue;
                        const uint32_t v = lds_load(&S[rb + dx]);
                        if (v && v < m) m = v;
                    }
                }
            }
            if (m < lab) {
                lds_min(&S[lab - 1], m);
                changed = 1;
            }
        }
        if (!__syncthreads_or(changed)) break;
        if (cap && ++iters >= cap) {
            if (t == 0) atomicOr(status, 1u);
            break;
        }
        for (uint32_t j = 0; j < ROWS / (NT / TX); ++j) {
            const uint32_t li = (q0 + (NT / TX) * j) * TX + cx;
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

    for (uint32_t j = 0; j < ROWS / (NT / TX); ++j) {
        const uint32_t q = q0 + (NT / TX) * j, ly = q % TY, lz = q / TY;
        const uint32_t gx = x0 + cx, gy = y0 + ly, gz = z0 + lz;
        if (gx >= w || gy >= h || gz >= d) continue;
        const uint32_t lab = S[q * TX + cx];
        uint32_t out = 0;
        if (lab) {
            const uint32_t lr = lab - 1;
            out = ((z0 + lr / (TX * TY)) * h + y0 + (lr / TX) % TY) * w + x0 + lr % TX + 1u;
        }
        L[((size_t)gz * h + gy) * w + gx] = out;
    }
}

struct Grid3 {
    uint32_t d, h, w, tiles_x, tiles_y, ntiles;
};

__device__ __forceinline__ bool face_voxel(const Grid3 &g, uint32_t gid, uint32_t &x, uint32_t &y, uint32_t &z,
                                           uint32_t &x0, uint32_t &y0, uint32_t &z0)
{
    const uint32_t tile = gid / JOBS;
    uint32_t j = gid % JOBS;
    if (tile >= g.ntiles) return false;
    x0 = (tile % g.tiles_x) * TX;
    y0 = ((tile / g.tiles_x) % g.tiles_y) * TY;
    z0 = (tile / (g.tiles_x * g.tiles_y)) * TZ;
    uint32_t lx, ly, lz;
    if (j < 2 * FACE_Z) {
        lz = j < FACE_Z ? 0 : TZ - 1;
        ly = (j % FACE_Z) / TX;
        lx = (j % FACE_Z) % TX;
    } else if ((j -= 2 * FACE_Z) < 2 * FACE_Y) {
        ly = j < FACE_Y ? 0 : TY - 1;
        lz = 1 + (j % FACE_Y) / TX;
        lx = (j % FACE_Y) % TX;
    } else {
        j -= 2 * FACE_Y;
        lx = j < FACE_X ? 0 : TX - 1;
        lz = 1 + (j % FACE_X) / (TY - 2);
        ly = 1 + (j % FACE_X) % (TY - 2);
    }
    x = x0 + lx;
    y = y0 + ly;
    z = z0 + lz;
    return x < g.w && y < g.h && z < g.d;
}

__global__ __launch_bounds__(256) void lbeq3d_merge(uint32_t *L, Grid3 g)
{
    uint32_t x, y, z, x0, y0, z0;
    if (!face_voxel(g, blockIdx.x * blockDim.x + threadIdx.x, x, y, z, x0, y0, z0)) return;
    const uint32_t p = (z * g.h + y) * g.w + x;
    const uint32_t vp = L[p];
    if (!vp) return;
    for (int e = 0; e < 13; ++e) {   // earlier neighbours in raster order
        const int dz = e < 9 ? -1 : 0;
        const int dy = e < 9 ? e / 3 - 1 : (e < 12 ? -1 : 0);
        const int dx = e < 9 ? e % 3 - 1 : (e < 12 ? e - 10 : -1);
        const uint32_t nx = x + dx, ny = y + dy, nz = z + dz;
        if (nx >= g.w || ny >= g.h || nz >= g.d) continue;
        if (nx - x0 < TX && ny - y0 < TY && nz - z0 < TZ) continue;
        const uint32_t vn = L[(nz * g.h + ny) * g.w + nx];
        if (vn) LabelForest::unite(L, vp - 1u, vn - 1u);
    }
}

__global__ __launch_bounds__(256) void lbeq3d_flatten(uint32_t *L, Grid3 g)
{
    uint32_t x, y, z, x0, y0, z0;
    if (!face_voxel(g, blockIdx.x * blockDim.x + threadIdx.x, x, y, z, x0, y0, z0)) return;
    const uint32_t p = (z * g.h + y) * g.w + x;
    if (L[p]) LabelForest::compress(L, p);
}

}  // namespace

size_t lbeq3d_workspace(uint32_t d, uint32_t h, uint32_t w, bool canonical)
{
    return align256(kCountBytes) + (canonical ? canon_workspace((uint64_t)d * h * w, 0) : 0);
}

hipError_t lbeq3d_run(const Image3D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt)
{
    Grid3 g{im.d, im.h, im.w, (im.w + TX - 1) / TX, (im.h + TY - 1) / TY, 0};
    g.ntiles = g.tiles_x * g.tiles_y * ((im.d + TZ - 1) / TZ);
    const bool vec = (im.w % 16 == 0) && (reinterpret_cast<uintptr_t>(im.vox) % 16 == 0);
    const uint32_t face_grid = (uint32_t)(((uint64_t)g.ntiles * JOBS + 255) / 256);
    uint32_t *slots = static_cast<uint32_t *>(ws);
    if (kt) kt->begin(s);
    if (vec)
        hipLaunchKernelGGL(lbeq3d_tile<true>, dim3(g.ntiles), dim3(NT), 0, s, im.vox, im.d, im.h, im.w, g.tiles_x,
                           g.tiles_y, im.labels, im.count ? slots : nullptr, opt.iter_cap, opt.status);
    else
        hipLaunchKernelGGL(lbeq3d_tile<false>, dim3(g.ntiles), dim3(NT), 0, s, im.vox, im.d, im.h, im.w, g.tiles_x,
                           g.tiles_y, im.labels, im.count ? slots : nullptr, opt.iter_cap, opt.status);
    if (kt) kt->mark(s, "lbeq3d_tile");
    hipLaunchKernelGGL(lbeq3d_merge, dim3(face_grid), dim3(256), 0, s, im.labels, g);
    if (kt) kt->mark(s, "lbeq3d_merge");
    hipLaunchKernelGGL(lbeq3d_flatten, dim3(face_grid), dim3(256), 0, s, im.labels, g);
    if (kt) kt->mark(s, "lbeq3d_flatten");
    void *canon_ws = opt.mode == LBL_LABELS_CANONICAL ? static_cast<uint8_t *>(ws) + align256(kCountBytes) : nullptr;
    hipError_t e = lbeq_final_launch(im.labels, (uint64_t)im.d * im.h * im.w, im.count, slots, canon_ws, s, kt);
    return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace lbl
