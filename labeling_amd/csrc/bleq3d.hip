// bleq3d.hip -- 3D block-equivalence labeling (2x2x2 blocks, 26-connectivity)
// for gfx950.
//
// Same problem as the reference's TOCLBlockEquivalence3D (LabelingAlgs.cpp:1402-1536,
// kernels BLEQ3D_*, LabelingAlgs.cl:737-979): voxels are grouped into 2x2x2
// blocks, blocks holding a foreground voxel are the nodes, two blocks are
// equivalent when a foreground voxel of one is 26-adjacent to a foreground voxel
// of the other, and each component is labelled by its minimum block.  The
// reference iterates whole-grid Scan/Analyze passes with a host round trip per
// pass; here three launches (and a one-wave count) resolve everything with
// union-find:
//
//   1. tile    one workgroup (512 threads) per 64x8x8-block tile
//              (128x16x16 voxels): four 16-byte row loads per thread -> 8-bit
//              block masks; union-find of the tile's blocks in LDS; writes the
//              tile-local root (global block index) and the mask of every block.
//   2. merge   one thread per block on a tile face: unions with each earlier
//              26-neighbour block in another tile that it touches.
//   3. final   one thread per 2 blocks: root of each block (neighbouring lanes
//              share one chase per tile-local root), writes the 16 voxel
//              labels of the 2 blocks, counts roots.
//
// Block mask bit of voxel (dx, dy, dz) in {0,1}^3 is dx + 2 dy + 4 dz.
#include "lbl_internal.hpp"
#include "uf_device.hpp"

namespace lbl {
namespace {

constexpr uint32_t TX = 64, TY = 8, TZ = 8;       // tile, blocks
constexpr uint32_t SEG = 8;                      // blocks per thread along x
constexpr uint32_t SEGS = TX / SEG;              // 8 threads per block row
constexpr uint32_t NT_TILE = SEGS * TY * TZ;     // 512
constexpr uint32_t TILE_BLOCKS = TX * TY * TZ;   // 4096
constexpr uint32_t FACE_Z = TX * TY, FACE_Y = TX * (TZ - 2), FACE_X = (TY - 2) * (TZ - 2);
constexpr uint32_t JOBS = 2 * FACE_Z + 2 * FACE_Y + 2 * FACE_X;
constexpr uint32_t FINAL_BLOCKS = 2;                // one 16-byte store per voxel row per thread
constexpr uint32_t NT_FINAL = 256;

using Forest = PackedForest<0>;

// Voxels of a block on the side facing neighbour direction d (per axis:
// d = -1 -> coordinate 0, d = +1 -> coordinate 1, d = 0 -> both), and of the
// neighbour on the side facing back.  Two blocks in direction d are 26-adjacent
// iff (cur & side_cur) && (nb & side_nb): any such pair of voxels is within one
// step on every axis.
__host__ __device__ constexpr uint32_t side_mask(int dx, int dy, int dz, bool neighbour)
{
    uint32_t m = 0;
    for (int v = 0; v < 8; ++v) {
        const int c[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
        const int d[3] = {dx, dy, dz};
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            if (d[a] == 0) continue;
            const int want = (d[a] < 0) != neighbour ? 0 : 1;
            ok = ok && c[a] == want;
        }
        if (ok) m |= 1u << v;
    }
    return m;
}

__device__ __forceinline__ bool adjacent(uint32_t cur, uint32_t nb, int dx, int dy, int dz)
{
    return (cur & side_mask(dx, dy, dz, false)) && (nb & side_mask(dx, dy, dz, true));
}

static_assert(side_mask(-1, 0, 0, false) == 0x55 && side_mask(-1, 0, 0, true) == 0xAA, "x faces");
static_assert(side_mask(0, -1, 0, false) == 0x33 && side_mask(0, -1, 0, true) == 0xCC, "y faces");
static_assert(side_mask(0, 0, -1, false) == 0x0F && side_mask(0, 0, -1, true) == 0xF0, "z faces");
static_assert(side_mask(-1, -1, -1, false) == 0x01 && side_mask(-1, -1, -1, true) == 0x80, "corner");

// Two 16-voxel rows of one z-plane pair -> masks of 8 blocks (8 bits each),
// split into the low (blocks 0-3) and high (blocks 4-7) 32-bit halves.
__device__ __forceinline__ void block_masks(const uint4 &r00, const uint4 &r01, const uint4 &r10, const uint4 &r11,
                                            uint32_t &lo, uint32_t &hi)
{
    const uint32_t a[4] = {r00.x, r00.y, r00.z, r00.w}, b[4] = {r01.x, r01.y, r01.z, r01.w};
    const uint32_t c[4] = {r10.x, r10.y, r10.z, r10.w}, d[4] = {r11.x, r11.y, r11.z, r11.w};
    uint32_t out[2] = {0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t z0 = quad_masks(a[q], b[q]);   // nibble per block, plane z0
        const uint32_t z1 = quad_masks(c[q], d[q]);   // plane z1
        const uint32_t m0 = (z0 & 15u) | ((z1 & 15u) << 4);
        const uint32_t m1 = (z0 >> 4) | (z1 & 0xF0u);
        out[q >> 1] |= (m0 | (m1 << 8)) << (16 * (q & 1));
    }
    lo = out[0];
    hi = out[1];
}

__device__ __forceinline__ uint4 load_row16(const uint8_t *row, uint32_t x0, uint32_t w, bool vec)
{
    if (vec && x0 + 16 <= w) return *reinterpret_cast<const uint4 *>(row + x0);
    uint32_t q[4] = {0, 0, 0, 0};
    for (uint32_t k = 0; k < 16; ++k)
        if (x0 + k < w && row[x0 + k]) q[k >> 2] |= 0xFFu << (8 * (k & 3));
    return make_uint4(q[0], q[1], q[2], q[3]);
}

__device__ __forceinline__ uint32_t mask_at(uint32_t lo, uint32_t hi, int k)
{
    return ((k < 4 ? lo : hi) >> (8 * (k & 3))) & 0xFFu;
}

// ---- 1. tile kernel -------------------------------------------------------------

template <bool VEC>
__global__ __launch_bounds__(NT_TILE) void bleq3d_tile(const uint8_t *__restrict__ vox, uint32_t d, uint32_t h,
                                                        uint32_t w, uint32_t bd, uint32_t bh, uint32_t bw,
                                                        uint32_t tiles_x, uint32_t tiles_y, uint32_t *__restrict__ P,
                                                        uint8_t *__restrict__ M, uint32_t *count, uint32_t *slots)
{
    __shared__ uint32_t par[TILE_BLOCKS];
    const SwizzleMap<SEG, NT_TILE> at;   // conflict-free placement of the forest
    __shared__ uint2 mw[TZ][TY][SEGS + 2];   // guard column on each side

    const uint32_t t = threadIdx.x, s = t % SEGS, ly = (t / SEGS) % TY, lz = t / (SEGS * TY);
    const uint32_t tx = blockIdx.x % tiles_x, ty = (blockIdx.x / tiles_x) % tiles_y, tz = blockIdx.x / (tiles_x * tiles_y);
    const uint32_t bx0 = tx * TX, by0 = ty * TY, bz0 = tz * TZ;
    const uint32_t bxs = bx0 + s * SEG, by = by0 + ly, bz = bz0 + lz;
    if (blockIdx.x == 0 && count) count_clear(slots);

    uint32_t lo = 0, hi = 0;
    if (bz < bd && by < bh && bxs < bw) {
        const uint32_t x0 = 2 * bxs, y0 = 2 * by, z0 = 2 * bz;
        const bool y1 = y0 + 1 < h, z1 = z0 + 1 < d;
        const uint8_t *p00 = vox + ((size_t)z0 * h + y0) * w;
        const uint4 zero = make_uint4(0, 0, 0, 0);
        const uint4 r00 = load_row16(p00, x0, w, VEC);
        const uint4 r01 = y1 ? load_row16(p00 + w, x0, w, VEC) : zero;
        const uint4 r10 = z1 ? load_row16(p00 + (size_t)h * w, x0, w, VEC) : zero;
        const uint4 r11 = (z1 && y1) ? load_row16(p00 + (size_t)h * w + w, x0, w, VEC) : zero;
        block_masks(r00, r01, r10, r11, lo, hi);
    }
    mw[lz][ly][s + 1] = make_uint2(lo, hi);
    if (s == 0) {
        mw[lz][ly][0] = make_uint2(0, 0);
        mw[lz][ly][SEGS + 1] = make_uint2(0, 0);
    }

    const uint32_t li0 = (lz * TY + ly) * TX + s * SEG;
    {
        uint32_t head = li0, prev = 0;
#pragma unroll
        for (int k = 0; k < (int)SEG; ++k) {
            const uint32_t m = mask_at(lo, hi, k), li = li0 + k;
            if (m && adjacent(m, prev, -1, 0, 0)) {
                par[at(li)] = head;
            } else {
                par[at(li)] = li;
                head = li;
            }
            prev = m;
        }
    }
    __syncthreads();

    if (s > 0) {
        const uint32_t m0 = mask_at(lo, hi, 0);
        const uint32_t left = mw[lz][ly][s].y >> 24;
        if (adjacent(m0, left, -1, 0, 0)) lds_union(par, li0, li0 - 1, at);
    }
    // earlier neighbour rows inside the tile: (dy, dz) = (-1,0), (-1,-1), (0,-1), (+1,-1)
    const int rdy[4] = {-1, -1, 0, 1}, rdz[4] = {0, -1, -1, -1};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int ny = (int)ly + rdy[q], nz = (int)lz + rdz[q];
        if (ny < 0 || ny >= (int)TY || nz < 0) continue;
        const uint2 up = mw[nz][ny][s + 1];
        const uint32_t up_l = mw[nz][ny][s].y >> 24;
        const uint32_t up_r = mw[nz][ny][s + 2].x & 0xFFu;
        const int ubase = (int)(((uint32_t)nz * TY + (uint32_t)ny) * TX + s * SEG);
        int last = -3;
        uint32_t prev = 0;
        for (int k = 0; k < (int)SEG; ++k) {
            const uint32_t m = mask_at(lo, hi, k);
            if (!m) {
                prev = 0;
                last = -3;
                continue;
            }
            if (!adjacent(m, prev, -1, 0, 0)) last = -3;
            prev = m;
            for (int c = k - 1; c <= k + 1; ++c) {
                const uint32_t um = c < 0 ? up_l : (c >= (int)SEG ? up_r : mask_at(up.x, up.y, c));
                if (!adjacent(m, um, c - k, rdy[q], rdz[q]) || c == last) continue;
                if (c == last + 1) {
                    const uint32_t lm = last < 0 ? up_l : mask_at(up.x, up.y, last);
                    if (adjacent(um, lm, -1, 0, 0)) {
                        last = c;
                        continue;
                    }
                }
                lds_union(par, li0 + k, (uint32_t)(ubase + c), at);
                last = c;
            }
        }
    }
    __syncthreads();

    if (bz < bd && by < bh && bxs < bw) {
        uint32_t out[SEG];
        uint32_t mk[2] = {lo, hi};
#pragma unroll
        for (int k = 0; k < (int)SEG; ++k) {
            const uint32_t m = mask_at(lo, hi, k), li = li0 + k;
            const uint32_t root = m ? lds_find(par, li, at) : li;
            const uint32_t rx = root % TX, ry = (root / TX) % TY, rz = root / (TX * TY);
            out[k] = ((bz0 + rz) * bh + by0 + ry) * bw + bx0 + rx;
        }
        const size_t row = ((size_t)bz * bh + by) * bw + bxs;
        uint32_t *dst = P + row;
        uint8_t *mdst = M + row;
        if ((bw % 4u) == 0 && bxs + SEG <= bw) {
            reinterpret_cast<uint4 *>(dst)[0] = make_uint4(out[0], out[1], out[2], out[3]);
            reinterpret_cast<uint4 *>(dst)[1] = make_uint4(out[4], out[5], out[6], out[7]);
            *reinterpret_cast<uint32_t *>(mdst) = mk[0];
            *reinterpret_cast<uint32_t *>(mdst + 4) = mk[1];
        } else {
            for (int k = 0; k < (int)SEG; ++k)
                if (bxs + k < bw) {
                    dst[k] = out[k];
                    mdst[k] = (uint8_t)mask_at(lo, hi, k);
                }
        }
    }
}

// ---- 2. tile-face merge -----------------------------------------------------

__device__ __forceinline__ void face_block(uint32_t j, uint32_t &lx, uint32_t &ly, uint32_t &lz)
{
    if (j < 2 * FACE_Z) {
        lz = j < FACE_Z ? 0 : TZ - 1;
        const uint32_t r = j % FACE_Z;
        ly = r / TX;
        lx = r % TX;
        return;
    }
    j -= 2 * FACE_Z;
    if (j < 2 * FACE_Y) {
        ly = j < FACE_Y ? 0 : TY - 1;
        const uint32_t r = j % FACE_Y;
        lz = 1 + r / TX;
        lx = r % TX;
        return;
    }
    j -= 2 * FACE_Y;
    lx = j < FACE_X ? 0 : TX - 1;
    const uint32_t r = j % FACE_X;
    lz = 1 + r / (TY - 2);
    ly = 1 + r % (TY - 2);
}

struct Grid3 {
    uint32_t bd, bh, bw, tiles_x, tiles_y, ntiles;
};

__device__ __forceinline__ bool job_block(const Grid3 &g, uint32_t gid, uint32_t &bx, uint32_t &by, uint32_t &bz,
                                          uint32_t &bx0, uint32_t &by0, uint32_t &bz0)
{
    const uint32_t tile = gid / JOBS, j = gid % JOBS;
    if (tile >= g.ntiles) return false;
    bx0 = (tile % g.tiles_x) * TX;
    by0 = ((tile / g.tiles_x) % g.tiles_y) * TY;
    bz0 = (tile / (g.tiles_x * g.tiles_y)) * TZ;
    uint32_t lx, ly, lz;
    face_block(j, lx, ly, lz);
    bx = bx0 + lx;
    by = by0 + ly;
    bz = bz0 + lz;
    return bx < g.bw && by < g.bh && bz < g.bd;
}

__global__ __launch_bounds__(256) void bleq3d_merge(uint32_t *P, const uint8_t *__restrict__ M, Grid3 g)
{
    uint32_t bx, by, bz, bx0, by0, bz0;
    if (!job_block(g, blockIdx.x * blockDim.x + threadIdx.x, bx, by, bz, bx0, by0, bz0)) return;
    const uint32_t b = (bz * g.bh + by) * g.bw + bx;
    const uint32_t m = M[b];
    if (!m) return;
    // the 13 earlier neighbours: dz = -1 (9), dz = 0 & dy = -1 (3), dz = dy = 0 & dx = -1
    for (int e = 0; e < 13; ++e) {
        const int dz = e < 9 ? -1 : 0;
        const int dy = e < 9 ? e / 3 - 1 : (e < 12 ? -1 : 0);
        const int dx = e < 9 ? e % 3 - 1 : (e < 12 ? e - 10 : -1);
        const uint32_t nx = bx + dx, ny = by + dy, nz = bz + dz;
        if (nx >= g.bw || ny >= g.bh || nz >= g.bd) continue;
        if (nx - bx0 < TX && ny - by0 < TY && nz - bz0 < TZ) continue;
        const uint32_t n = (nz * g.bh + ny) * g.bw + nx;
        if (adjacent(m, M[n], dx, dy, dz)) Forest::unite(P, P[b], P[n]);   // tile-local roots
    }
}

// ---- 3. final kernel --------------------------------------------------------------

template <bool VEC>
__global__ __launch_bounds__(NT_FINAL) void bleq3d_final(const uint32_t *__restrict__ P, const uint8_t *__restrict__ M,
                                                          uint32_t *__restrict__ labels, uint32_t d, uint32_t h,
                                                          uint32_t w, uint32_t bd, uint32_t bh, uint32_t bw,
                                                          uint32_t xgroups, uint32_t *count, uint32_t *slots)
{
    const uint32_t row = blockIdx.x / xgroups;           // block row index bz * bh + by
    const uint32_t by = row % bh, bz = row / bh;
    const uint32_t bxs = ((blockIdx.x % xgroups) * NT_FINAL + threadIdx.x) * FINAL_BLOCKS;
    const bool inside = bz < bd && bxs < bw;
    const size_t b0 = ((size_t)bz * bh + by) * bw + bxs;
    uint32_t pv[FINAL_BLOCKS] = {}, mv[FINAL_BLOCKS] = {};
    if (inside) {
        if (VEC && bxs + FINAL_BLOCKS <= bw) {
            const uint2 q = *reinterpret_cast<const uint2 *>(P + b0);
            const uint32_t mm = *reinterpret_cast<const uint16_t *>(M + b0);
            pv[0] = q.x; pv[1] = q.y;
            mv[0] = mm & 0xFFu; mv[1] = mm >> 8;
        } else {
            for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
                pv[k] = bxs + k < bw ? P[b0 + k] : 0u;
                mv[k] = bxs + k < bw ? M[b0 + k] : 0u;
            }
        }
    }
    // Neighbouring lanes mostly share their blocks' parent (a tile-local root):
    // the first lane of each run of lanes with the same first parent chases
    // it, the others take its root.
    uint32_t key = 0;
    bool has = false;
#pragma unroll
    for (uint32_t k = FINAL_BLOCKS; k-- > 0;)
        if (mv[k]) {
            key = pv[k];
            has = true;
        }
    uint32_t leader;
    const bool lead = run_leader(key, 0u, has, leader);
    uint32_t key_root = lead ? Forest::find_final(P, key) : 0u;
    key_root = __shfl(key_root, (int)leader, 64);

    uint32_t roots = 0;
    if (inside) {
        uint32_t lab[FINAL_BLOCKS];
        uint32_t prev_parent = key, prev_root = key_root;
#pragma unroll
        for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
            lab[k] = 0;
            if (!mv[k]) continue;
            const uint32_t root = pv[k] == prev_parent ? prev_root : Forest::find_final(P, pv[k]);
            prev_parent = pv[k];
            prev_root = root;
            roots += root == (uint32_t)(b0 + k);
            lab[k] = root + 1u;
        }
        const uint32_t x0 = 2 * bxs;
#pragma unroll
        for (uint32_t v = 0; v < 4; ++v) {           // (dy, dz) rows of the blocks
            const uint32_t dy = v & 1u, dz = v >> 1;
            const uint32_t y = 2 * by + dy, z = 2 * bz + dz;
            if (y >= h || z >= d) continue;
            uint32_t r[2 * FINAL_BLOCKS];
#pragma unroll
            for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
                const uint32_t bit = 2 * dy + 4 * dz;
                r[2 * k] = ((mv[k] >> bit) & 1u) ? lab[k] : 0u;
                r[2 * k + 1] = ((mv[k] >> (bit + 1)) & 1u) ? lab[k] : 0u;
            }
            uint32_t *dst = labels + ((size_t)z * h + y) * w + x0;
            if (VEC && x0 + 2 * FINAL_BLOCKS <= w) {
                store_nt4(dst, r[0], r[1], r[2], r[3]);
            } else {
                for (uint32_t i = 0; i < 2 * FINAL_BLOCKS && x0 + i < w; ++i) dst[i] = r[i];
            }
        }
    }
    if (count) count_add(slots, roots);
}

}  // namespace

size_t bleq3d_workspace(uint32_t d, uint32_t h, uint32_t w)
{
    const uint64_t nb = ((d + 1ull) / 2) * ((h + 1ull) / 2) * ((w + 1ull) / 2);
    return align256(nb * sizeof(uint32_t)) + align256(nb) + kCountBytes;
}

hipError_t bleq3d_run(const Image3D &im, const Options &, void *ws, hipStream_t s, KernelTimer *kt)
{
    const uint32_t bd = (im.d + 1u) / 2, bh = (im.h + 1u) / 2, bw = (im.w + 1u) / 2;
    const uint64_t nb = (uint64_t)bd * bh * bw;
    Grid3 g{bd, bh, bw, (bw + TX - 1) / TX, (bh + TY - 1) / TY, 0};
    const uint32_t tiles_z = (bd + TZ - 1) / TZ;
    g.ntiles = g.tiles_x * g.tiles_y * tiles_z;
    uint32_t *P = static_cast<uint32_t *>(ws);
    uint8_t *M = static_cast<uint8_t *>(ws) + align256(nb * sizeof(uint32_t));
    uint32_t *slots = reinterpret_cast<uint32_t *>(M + align256(nb));
    const bool vec_in = (im.w % 16 == 0) && (reinterpret_cast<uintptr_t>(im.vox) % 16 == 0);
    const bool vec_out = (im.w % 8 == 0) && (reinterpret_cast<uintptr_t>(im.labels) % 16 == 0);
    const uint32_t face_grid = (uint32_t)(((uint64_t)g.ntiles * JOBS + 255) / 256);
    const uint32_t xgroups = (bw + NT_FINAL * FINAL_BLOCKS - 1) / (NT_FINAL * FINAL_BLOCKS);

    if (kt) kt->begin(s);
    if (vec_in)
        hipLaunchKernelGGL(bleq3d_tile<true>, dim3(g.ntiles), dim3(NT_TILE), 0, s, im.vox, im.d, im.h, im.w, bd, bh,
                           bw, g.tiles_x, g.tiles_y, P, M, im.count, slots);
    else
        hipLaunchKernelGGL(bleq3d_tile<false>, dim3(g.ntiles), dim3(NT_TILE), 0, s, im.vox, im.d, im.h, im.w, bd, bh,
                           bw, g.tiles_x, g.tiles_y, P, M, im.count, slots);
    if (kt) kt->mark(s, "bleq3d_tile");
    hipLaunchKernelGGL(bleq3d_merge, dim3(face_grid), dim3(256), 0, s, P, M, g);
    if (kt) kt->mark(s, "bleq3d_merge");
    const dim3 fgrid((uint32_t)((uint64_t)xgroups * bh * bd));
    if (vec_out)
        hipLaunchKernelGGL(bleq3d_final<true>, fgrid, dim3(NT_FINAL), 0, s, P, M, im.labels, im.d, im.h, im.w, bd, bh,
                           bw, xgroups, im.count, slots);
    else
        hipLaunchKernelGGL(bleq3d_final<false>, fgrid, dim3(NT_FINAL), 0, s, P, M, im.labels, im.d, im.h, im.w, bd,
                           bh, bw, xgroups, im.count, slots);
    if (kt) kt->mark(s, "bleq3d_final");
    if (im.count) {
        hipLaunchKernelGGL(count_reduce_kernel<0>, dim3(1), dim3(kCountSlots), 0, s, slots, im.count);
        if (kt) kt->mark(s, "count_reduce");
    }
    return hipGetLastError();
}

}  // namespace lbl
