// bleq3d.hip -- 3D block-equivalence labeling (2x2x2 blocks, 26-connectivity)
// for gfx950.
//
// Same problem as the reference's TOCLBlockEquivalence3D (LabelingAlgs.cpp:1402-1536,
// kernels BLEQ3D_*, LabelingAlgs.cl:737-979): voxels are grouped into 2x2x2
// blocks, blocks holding a foreground voxel are the nodes, two blocks are
// equivalent when a foreground voxel of one is 26-adjacent to a foreground voxel
// of the other, and each component is labelled by its minimum block.  The
// reference iterates whole-grid Scan/Analyze passes with a host round trip per
// pass; here three launches resolve everything with union-find (components
// counted at tile-local roots, minus the merge's links, summed by the final
// kernel's first workgroup):
//
//   1. tile    one workgroup (512 threads) per 64x16x8-block tile
//              (128x32x16 voxels): each lane turns four 32-voxel rows into bit
//              masks; union-find over the runs of each 2x2 bar of voxel rows
//              in LDS; writes the tile-local root (global block index) and the
//              mask of every block.
//   2. merge   unions across tile faces: bar segments against their earlier
//              neighbour bars in other tiles (bit-mask contacts), and the
//              tile's first block column against the column to its left;
//              distinct pairs of tile-local roots collected in LDS, united
//              by all threads.
//   3. final   one thread per 2 blocks: root of each block (neighbouring lanes
//              share one chase per tile-local root), writes the 16 voxel
//              labels of the 2 blocks, counts roots.
//
// Block mask bit of voxel (dx, dy, dz) in {0,1}^3 is dx + 2 dy + 4 dz.
#include <type_traits>

#include "lbl_internal.hpp"
#include "canon.hpp"
#include "uf_device.hpp"

namespace lbl {
namespace {

constexpr uint32_t TX = 64, TY = 16, TZ = 8;      // tile, blocks (128 x 32 x 16 voxels)
// merge lanes per tile: segments of the bars whose earlier neighbour bars lie
// in another tile (z = 0 plane, and the y = 0 / y = TY - 1 rows above it), and
// one lane per bar for the tile's first block column
constexpr uint32_t MERGE_BARS = TY + 2 * (TZ - 1);
constexpr uint32_t BAR_LANES = MERGE_BARS * (TX / kLaneBlocks);
constexpr uint32_t X_LANES = TY * TZ;
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

// ---- 1. tile kernel -------------------------------------------------------------
//
// A tile is TBW x TBY x TBZ blocks.  A "bar" is one row of blocks along x; one
// lane holds a 32-voxel segment of a bar (16 blocks) as the masks of its four
// voxel rows: a = (y0, z0), b = (y1, z0), c = (y0, z1), e = (y1, z1).  All
// lanes of a bar are consecutive lanes of one wave.  Within a bar,
// 26-connectivity is exactly "same run of set bits of o = a | b | c | e" (two
// voxels in adjacent columns of a 2 x 2 bar always touch), so, as in 2D, the
// union-find nodes are the segment-local runs, a run continuing the left
// segment stands for the head of the whole bar run, and contacts with the four
// earlier neighbour bars are found with bit operations on facing rows:
//   (y-1):      a | c  against the neighbour's b | e
//   (z-1):      a | b  against the neighbour's c | e
//   (y-1, z-1): a      against the neighbour's e
//   (y+1, z-1): b      against the neighbour's c
// (the later neighbours meet this bar as one of their earlier ones).  Each lane
// feeds its contacts through one union loop, then resolves its nodes' roots and
// writes its 16 block words and masks.

template <uint32_t TBW, uint32_t TBY, uint32_t TBZ>
struct Tile3 {
    static constexpr uint32_t LPB = TBW / kLaneBlocks;          // lanes per bar
    static constexpr uint32_t THREADS = LPB * TBY * TBZ;
    static constexpr uint32_t BLOCKS = TBW * TBY * TBZ;
    static_assert((TBW & (TBW - 1)) == 0 && (TBY & (TBY - 1)) == 0 && LPB >= 1 && LPB <= 64, "tile shape");
    static constexpr uint32_t LOG_TBW = __builtin_ctz(TBW), LOG_TBY = __builtin_ctz(TBY);
    // per-bar LDS tables: [z + 1][y + 1][segment + 1], zero guard entries around
    static constexpr uint32_t GZ = TBZ + 1, GY = TBY + 2, GX = LPB + 2, ENTRIES = GZ * GY * GX;
    static __device__ __forceinline__ uint32_t at(uint32_t gz, uint32_t gy, uint32_t gx) { return (gz * GY + gy) * GX + gx; }
};

// First-voxel key of block (bx, by, bz) with mask m != 0 (bit 1 << (dx + 2 dy
// + 4 dz), so the lowest bit is the first voxel): its voxel index (voxel_key),
// or (z * h + y) * bw + bx, the bit position of canonical numbering.  Both
// order blocks' first voxels as the raster does, and no two blocks share a key.
__device__ __forceinline__ uint32_t block_key3(uint32_t h, uint32_t w, uint32_t bw, uint32_t bz, uint32_t by,
                                               uint32_t bx, uint32_t m, bool voxel_key)
{
    const uint32_t bit = __builtin_ctz(m), dx = bit & 1u, dy = (bit >> 1) & 1u, dz = bit >> 2;
    const uint32_t zy = (2 * bz + dz) * h + 2 * by + dy;
    return voxel_key ? zy * w + 2 * bx + dx : zy * bw + bx;
}

// x-boundary planes: for each boundary k = 1 .. tiles_x - 1 between tile
// columns k - 1 and k (block column x = k * TX), the (tile-local root, mask) of
// every block of the column just right of it (west plane: tile column k's
// first) and just left of it (east plane: tile column k - 1's last), entry
// z * bh + y.  The tile kernel writes them with its block words; the merge's
// x-face lanes read them instead of those columns of P and M, whose blocks
// each lie in a cache line of their own.
__device__ __forceinline__ uint2 *xplane_west(uint2 *xf, uint32_t bd, uint32_t bh, uint32_t k)
{
    return xf + (size_t)(k - 1) * bd * bh;
}

__device__ __forceinline__ uint2 *xplane_east(uint2 *xf, uint32_t tiles_x, uint32_t bd, uint32_t bh, uint32_t k)
{
    return xf + (size_t)(tiles_x - 1 + k - 1) * bd * bh;
}

// VEC: 16-byte vector loads of the rows; FULL: w % 32 == 0, so every lane's
// segments are whole and the kernel holds no code for row ends.
template <uint32_t TBW, uint32_t TBY, uint32_t TBZ, bool VEC, bool FULL>
__global__ __launch_bounds__((TBW / kLaneBlocks) * TBY * TBZ) void bleq3d_tile(
    const uint8_t *__restrict__ vox, uint32_t d, uint32_t h, uint32_t w, uint32_t bd, uint32_t bh, uint32_t bw,
    uint32_t tiles_x, uint32_t tiles_y, uint32_t *__restrict__ P, uint8_t *__restrict__ M, uint2 *__restrict__ xf,
    uint32_t *count, uint32_t *slots, uint32_t *ckey, bool voxel_key)
{
    using T = Tile3<TBW, TBY, TBZ>;
    constexpr uint32_t LPB = T::LPB;
    __shared__ uint32_t par[T::BLOCKS];   // node (run start block) -> parent, local indices
    __shared__ uint32_t so[T::ENTRIES];   // o = a | b | c | e
    __shared__ uint32_t sbe[T::ENTRIES];  // b | e (the y1 rows)
    __shared__ uint32_t sc[T::ENTRIES];   // c
    __shared__ uint32_t se[T::ENTRIES];   // e
    __shared__ uint32_t sh[T::ENTRIES];   // node standing for the segment's first run

    const uint32_t t = threadIdx.x, j = t % LPB, bar = t / LPB, ly = bar % TBY, lz = bar / TBY;
    const uint32_t tx = blockIdx.x % tiles_x, ty = (blockIdx.x / tiles_x) % tiles_y, tz = blockIdx.x / (tiles_x * tiles_y);
    const uint32_t bx0 = tx * TBW, by0 = ty * TBY, bz0 = tz * TBZ;
    const uint32_t bxs = bx0 + j * kLaneBlocks, by = by0 + ly, bz = bz0 + lz;

    uint32_t a = 0, b = 0, c = 0, e = 0;
    if (bz < bd && by < bh && bxs < bw) {
        const uint32_t x0 = 2 * bxs, y0 = 2 * by, z0 = 2 * bz;
        const bool y1 = y0 + 1 < h, z1 = z0 + 1 < d;
        const uint8_t *r00 = vox + ((size_t)z0 * h + y0) * w;
        const size_t plane = (size_t)h * w, off[4] = {0, w, plane, plane + w};
        const bool use[4] = {true, y1, z1, z1 && y1};
        uint32_t m[4];
        rows_bits32<4, FULL>(r00, off, use, x0, w, VEC, m);
        a = m[0]; b = m[1]; c = m[2]; e = m[3];
    }
    const uint32_t o = a | b | c | e;
    const uint32_t st = o & ~(o << 1);                 // first column of each segment-local run
    const uint32_t bar0 = bar * TBW, nb = bar0 + j * kLaneBlocks, nbn = nb * kNodeBytes;   // nbn: node name
    const uint32_t first = run_head(o, j, bar0 * kNodeBytes);   // node standing for the segment's first run

    // guard entries (outside the tile) read as empty segments
    for (uint32_t i = t; i < T::ENTRIES; i += T::THREADS) {
        const uint32_t gx = i % T::GX, gy = (i / T::GX) % T::GY, gz = i / (T::GX * T::GY);
        if (gz == 0 || gy == 0 || gy == TBY + 1 || gx == 0 || gx == LPB + 1) so[i] = sbe[i] = sc[i] = se[i] = 0;
    }
    const uint32_t me = T::at(lz + 1, ly + 1, j + 1);
    so[me] = o;
    sbe[me] = b | e;
    sc[me] = c;
    se[me] = e;
    sh[me] = first;
    {
        constexpr uint32_t K = kNodeBytes;
        uint4 *p4 = reinterpret_cast<uint4 *>(&par[nb]);
        p4[0] = make_uint4(first, nbn + K, nbn + 2 * K, nbn + 3 * K);
#pragma unroll
        for (uint32_t q = 1; q < kLaneBlocks / 4; ++q)
            p4[q] = make_uint4(nbn + 4 * q * K, nbn + (4 * q + 1) * K, nbn + (4 * q + 2) * K, nbn + (4 * q + 3) * K);
    }
    __syncthreads();

    // unions along the contacts with the four earlier neighbour bars, one direction after the other
    {
        // direction q: neighbour bar offset (dy, dz) and the facing rows
        const auto direction = [&](uint32_t q, SegContacts &cs, SegNodes &nn) {
            const int dy = q == 1 ? 0 : (q == 3 ? 1 : -1), dz = q == 0 ? 0 : -1;
            const uint32_t gz = lz + 1 + dz, gy = ly + 1 + dy;
            const uint32_t i0 = T::at(gz, gy, j), i1 = i0 + 1, i2 = i0 + 2;   // left, same, right segment
            uint32_t A, B, Bl, Br;
            if (q == 0) {
                A = a | c; B = sbe[i1]; Bl = sbe[i0]; Br = sbe[i2];
            } else if (q == 1) {
                A = a | b; B = sc[i1] | se[i1]; Bl = sc[i0] | se[i0]; Br = sc[i2] | se[i2];
            } else if (q == 2) {
                A = a; B = se[i1]; Bl = se[i0]; Br = se[i2];
            } else {
                A = b; B = sc[i1]; Bl = sc[i0]; Br = sc[i2];
            }
            const uint32_t on = so[i1];
            const uint32_t nbar0 = (uint32_t)((int)bar + dy + dz * (int)TBY) * TBW;   // only used on a contact
            nn.st = on & ~(on << 1);
            nn.base = (nbar0 + j * kLaneBlocks) * kNodeBytes;
            nn.first = sh[i1];
            nn.left = last_run_node(so[i0], nn.base - kLaneBlocks * kNodeBytes, sh[i0]);
            nn.right = sh[i2];
            cs = SegContacts(A, B, Bl, Br, st, nn.st);
        };
        SegContacts cs(0, 0, 0, 0, 0, 0);
        SegNodes nn{};
        uint32_t q = 0;
        direction(0, cs, nn);
        // (starting a contact that shares a node with the lane's previous union
        // from where that union met, as the 2D tile kernel does, made this
        // kernel slower: 134 -> 141 us at 512^3)
        uint32_t ua = 0, ub = 0, top = 0;
        bool busy = false;
        for (;;) {
            if (!busy) {
                while (!cs.any() && q < 3) direction(++q, cs, nn);
                if (!cs.any()) break;
                uint32_t x;
                const int p = cs.pop_any(x);
                const uint32_t s = hibit(st & upto(x));
                ua = s ? node_at(nbn, s) : first;
                ub = nn.node(p);
                busy = true;
            }
            if (lds_rem_step(par, ua, ub, top)) busy = false;
        }
    }
    __syncthreads();

    // Roots of this lane's nodes (self-roots: global index g0 + block; the
    // others chased and replaced by the tagged global root).
    const uint32_t g0 = (bz * bh + by) * bw + bxs;   // global index of local block nb
    uint32_t v[kLaneBlocks];
    const auto load_nodes = [&]() {
        const uint4 *p4 = reinterpret_cast<const uint4 *>(&par[nb]);
#pragma unroll
        for (uint32_t q = 0; q < kLaneBlocks / 4; ++q) {
            const uint4 x = p4[q];
            v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
        }
    };
    load_nodes();
    uint32_t pend = 0;
#pragma unroll
    for (uint32_t k = 0; k < kLaneBlocks; ++k)
        if ((st & (3u << (2 * k))) && v[k] != nbn + k * kNodeBytes) pend |= 1u << k;
    if (pend) {
        resolve_nodes(par, nbn, pend, [&](uint32_t l) {
            const uint32_t q = l >> T::LOG_TBW;
            return ((bz0 + (q >> T::LOG_TBY)) * bh + by0 + (q & (TBY - 1))) * bw + bx0 + (l & (TBW - 1));
        });
        load_nodes();
    }

    // block words (root; a background block, mask 0, carries the parent of the
    // run before it, or 0: no reader uses it) and masks (bit dx + 2 dy + 4 dz);
    // the tile's components are counted at their roots (the merge subtracts
    // its links): the nodes no chase started from
    static_assert(kLaneBlocks < 32, "count_add_wave takes per-lane counts below 32");
    const uint32_t rootmask = pair_any16(st) & ~pend;
    const uint32_t roots = __builtin_popcount(rootmask);
    if (bz < bd && by < bh && bxs < bw) {
        uint32_t out[kLaneBlocks], mk[kLaneBlocks / 4] = {0, 0, 0, 0}, cur = 0;
        const auto mask_of = [&](uint32_t k) {
            return ((a >> (2 * k)) & 3u) | (((b >> (2 * k)) & 3u) << 2) | (((c >> (2 * k)) & 3u) << 4) |
                   (((e >> (2 * k)) & 3u) << 6);
        };
#pragma unroll
        for (uint32_t k = 0; k < kLaneBlocks; ++k) {   // selects only: no per-block branch
            const uint32_t m = mask_of(k);
            const bool node = (st & (3u << (2 * k))) != 0;
            const uint32_t nv = (v[k] & kResolved) ? (v[k] & ~kResolved) : g0 + k;
            cur = node ? nv : cur;
            out[k] = cur;
            mk[k >> 2] |= m << (8 * (k & 3));
        }
        // first-voxel keys start at the root block's own first voxel (global
        // roots are tile-local roots)
        if (ckey)
            for (uint32_t rm = rootmask; rm; rm &= rm - 1) {
                const uint32_t k = __builtin_ctz(rm);
                ckey[g0 + k] = block_key3(h, w, bw, bz, by, bxs + k, mask_of(k), voxel_key);
            }
        const size_t row = ((size_t)bz * bh + by) * bw + bxs;
        uint32_t *dst = P + row;
        uint8_t *mdst = M + row;
        if ((bw % kLaneBlocks) == 0 && bxs + kLaneBlocks <= bw) {
            uint4 *d4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
            for (uint32_t q = 0; q < kLaneBlocks / 4; ++q) d4[q] = make_uint4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
            *reinterpret_cast<uint4 *>(mdst) = make_uint4(mk[0], mk[1], mk[2], mk[3]);
        } else {
            for (uint32_t k = 0; k < kLaneBlocks && bxs + k < bw; ++k) {
                dst[k] = out[k];
                mdst[k] = (uint8_t)(mk[k >> 2] >> (8 * (k & 3)));
            }
        }
        // the tile's first and last block columns, also into the x-boundary
        // planes for the merge (xplane_west / xplane_east)
        if (j == 0 && tx > 0) xplane_west(xf, bd, bh, tx)[(size_t)bz * bh + by] = make_uint2(out[0], mask_of(0));
        if (j == LPB - 1 && tx + 1 < tiles_x)
            xplane_east(xf, tiles_x, bd, bh, tx + 1)[(size_t)bz * bh + by] =
                make_uint2(out[kLaneBlocks - 1], mask_of(kLaneBlocks - 1));
    }
    if (count) count_add_wave(slots, roots);
}

// ---- 2. tile-face merge -----------------------------------------------------------
//
// Unions across tile faces, between tile-local roots (P holds each block's
// tile-local root, M its mask).  Two kinds of lanes:
//   * bar lanes: a 16-block segment of a bar whose earlier neighbour bars lie
//     in another tile (y-1, z-1, (y-1, z-1), (y+1, z-1)); the bars' voxel rows
//     come back from M as bit masks and the tile kernel's contact logic finds
//     the contacts (x-boundary contacts excepted: the next kind has them);
//   * x lanes: one per bar at the tile's first block column, against the nine
//     blocks at the column to the left.
// Lanes put pairs of roots into a workgroup hash set in LDS, then every thread
// of the workgroup unites the distinct pairs (a lane with many contacts does
// not serialize its unions, and the many contacts of two large components
// across a face become one union).

struct Grid3 {
    uint32_t bd, bh, bw, tiles_x, tiles_y, ntiles;
};

// A workgroup takes MERGE_JOBS face lanes and unites their distinct pairs with
// all MERGE_THREADS threads (unions are chains of dependent loads).
constexpr uint32_t MERGE_THREADS = 256;
constexpr uint32_t MERGE_JOBS = 256;
// Pair set slots per workgroup: 2048 (20 KB of LDS) lets eight workgroups share
// a CU, so the whole merge grid is resident at once; pairs that find the set
// full are united by their lane (512^3: merge 71.8 -> 64.6 us against 4096
// slots; 1024 slots: 68.0).
constexpr uint32_t PAIR_CAP = 8 * MERGE_JOBS;

// The four voxel rows of 4 blocks (mask bit dx + 2 dy + 4 dz) as 8 bits each:
// bit 2k + dx of row (dy, dz) for block k.
__device__ __forceinline__ uint32_t row_bits8(uint32_t four_masks, uint32_t row_shift)
{
    const uint32_t x = (four_masks >> row_shift) & 0x03030303u;
    return (x | (x >> 6) | (x >> 12) | (x >> 18)) & 0xFFu;
}

// Row masks a (y0, z0), b (y1, z0), c (y0, z1), e (y1, z1) of the 16-block
// segment at m[0..15] (zero past bw).
__device__ __forceinline__ void bar_rows(const uint8_t *m, uint32_t bxs, uint32_t bw, uint32_t &a, uint32_t &b,
                                         uint32_t &c, uint32_t &e)
{
    uint32_t w[4];
    if (bxs + kLaneBlocks <= bw && (reinterpret_cast<uintptr_t>(m) & 15u) == 0) {
        const uint4 q = *reinterpret_cast<const uint4 *>(m);
        w[0] = q.x; w[1] = q.y; w[2] = q.z; w[3] = q.w;
    } else {
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            w[i] = 0;
            for (uint32_t k = 0; k < 4; ++k)
                if (bxs + 4 * i + k < bw) w[i] |= (uint32_t)m[4 * i + k] << (8 * k);
        }
    }
    a = b = c = e = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        a |= row_bits8(w[i], 0) << (8 * i);
        b |= row_bits8(w[i], 2) << (8 * i);
        c |= row_bits8(w[i], 4) << (8 * i);
        e |= row_bits8(w[i], 6) << (8 * i);
    }
}

__global__ __launch_bounds__(MERGE_THREADS) void bleq3d_merge(uint32_t *P, const uint8_t *__restrict__ M, uint2 *xf,
                                                              Grid3 g, uint32_t *count, uint32_t *slots)
{
    constexpr uint32_t LPB = TX / kLaneBlocks;
    __shared__ LdsPairSet<PAIR_CAP> set;   // distinct pairs of roots
    set.init();
    __syncthreads();

    const uint64_t gid = (uint64_t)blockIdx.x * MERGE_JOBS + threadIdx.x;
    const bool job = threadIdx.x < MERGE_JOBS;
    const uint64_t bar_jobs = (uint64_t)g.ntiles * BAR_LANES;
    uint32_t links = 0;
    uint32_t la = ~0u, lb = ~0u;   // the lane's previous pair
    const auto push = [&](uint32_t ra, uint32_t rb) {
        if (ra == rb || (ra == la && rb == lb)) return;
        la = ra;
        lb = rb;
        if (!set.insert(ra, rb)) links += Forest::unite_cached(P, ra, rb);   // no room: unite here
    };

    if (job && gid < bar_jobs) {
        const uint32_t tile = (uint32_t)(gid / BAR_LANES), k = (uint32_t)(gid % BAR_LANES);
        const uint32_t bar = k / LPB, j = k % LPB;
        const uint32_t ly = bar < TY ? bar : ((bar - TY) & 1u) * (TY - 1), lz = bar < TY ? 0u : 1u + (bar - TY) / 2;
        const uint32_t bx0 = (tile % g.tiles_x) * TX, by = ((tile / g.tiles_x) % g.tiles_y) * TY + ly;
        const uint32_t bz = (tile / (g.tiles_x * g.tiles_y)) * TZ + lz, bxs = bx0 + j * kLaneBlocks;
        if (by < g.bh && bz < g.bd && bxs < g.bw) {
            const size_t row = ((size_t)bz * g.bh + by) * g.bw;
            uint32_t a, b, c, e;
            bar_rows(M + row + bxs, bxs, g.bw, a, b, c, e);
            const uint32_t o = a | b | c | e, st = o & ~(o << 1);
            for (uint32_t q = 0; q < 4 && o; ++q) {
                // direction q: (y-1), (z-1), (y-1, z-1), (y+1, z-1); only those leaving the tile
                const int dy = q == 1 ? 0 : (q == 3 ? 1 : -1), dz = q == 0 ? 0 : -1;
                const bool out = (dy < 0 && ly == 0) || (dy > 0 && ly == TY - 1) || (dz < 0 && lz == 0);
                const uint32_t ny = by + dy, nz = bz + dz;   // unsigned wrap -> outside
                if (!out || ny >= g.bh || nz >= g.bd) continue;
                const size_t nrow = ((size_t)nz * g.bh + ny) * g.bw;
                uint32_t na, nbm, nc, ne;
                bar_rows(M + nrow + bxs, bxs, g.bw, na, nbm, nc, ne);
                // facing rows: ours A, theirs B (row shifts of B in a block mask)
                const uint32_t A = q == 0 ? (a | c) : (q == 1 ? (a | b) : (q == 2 ? a : b));
                const uint32_t B = q == 0 ? (nbm | ne) : (q == 1 ? (nc | ne) : (q == 2 ? ne : nc));
                const uint32_t bmask = q == 0 ? 0xCCu : (q == 1 ? 0xF0u : (q == 2 ? 0xC0u : 0x30u));
                // neighbour segments' edge columns; none across the tile's x boundary
                const uint32_t ml = (j > 0) ? M[nrow + bxs - 1] : 0u;
                const uint32_t mr = (j + 1 < LPB && bxs + kLaneBlocks < g.bw) ? M[nrow + bxs + kLaneBlocks] : 0u;
                const uint32_t Bl = (ml & bmask & 0xAAu) ? 0x80000000u : 0u;   // dx = 1 voxels of the left block
                const uint32_t Br = (mr & bmask & 0x55u) ? 1u : 0u;             // dx = 0 voxels of the right block
                const uint32_t on = na | nbm | nc | ne;
                SegContacts cs(A, B, Bl, Br, st, on & ~(on << 1));
                while (cs.any()) {
                    uint32_t x;
                    const int p = cs.pop(x);
                    push(P[row + bxs + (x >> 1)], P[nrow + bxs + ((p + 2) >> 1) - 1]);
                }
            }
        }
    } else if (job) {
        const uint64_t g2 = gid - bar_jobs;
        const uint32_t tile = (uint32_t)(g2 / X_LANES), k = (uint32_t)(g2 % X_LANES);
        if (tile < g.ntiles) {
            const uint32_t ly = k % TY, lz = k / TY;
            const uint32_t bx0 = (tile % g.tiles_x) * TX, by = ((tile / g.tiles_x) % g.tiles_y) * TY + ly;
            const uint32_t bz = (tile / (g.tiles_x * g.tiles_y)) * TZ + lz;
            if (bx0 > 0 && bx0 < g.bw && by < g.bh && bz < g.bd) {
                const uint32_t kx = tile % g.tiles_x;   // the boundary on this tile's left
                const uint2 *W = xplane_west(xf, g.bd, g.bh, kx);
                const uint2 *E = xplane_east(xf, g.tiles_x, g.bd, g.bh, kx);
                const uint2 w0 = W[(size_t)bz * g.bh + by];
                const uint32_t m = w0.y;
                if (m & 0x55u) {
                    for (int e9 = 0; e9 < 9; ++e9) {
                        const int dy = e9 % 3 - 1, dz = e9 / 3 - 1;
                        const uint32_t ny = by + dy, nz = bz + dz;
                        if (ny >= g.bh || nz >= g.bd) continue;
                        const uint2 en = E[(size_t)nz * g.bh + ny];
                        if (adjacent(m, en.y, -1, dy, dz)) push(w0.x, en.x);
                    }
                }
            }
        }
    }
    __syncthreads();
    const uint32_t n = set.size();
    for (uint32_t i = threadIdx.x; i < n; i += MERGE_THREADS) {
        const uint2 pr = set.pair(i);
        links += Forest::unite_cached(P, pr.x, pr.y);
    }
    if (count) count_links(slots, links);   // kernel-uniform branch
}

// ---- 3. final kernel --------------------------------------------------------------

// Parents (P) and masks (M) of one lane's FINAL_BLOCKS blocks (zero outside
// the volume) and the blocks' roots (where the mask is not empty).
// Neighbouring lanes mostly share their blocks' parent (a tile-local root):
// the first lane of each run of lanes with the same first parent chases it,
// the others take its root.
template <bool VEC>
__device__ __forceinline__ void lane_blocks3(const uint32_t *__restrict__ P, const uint8_t *__restrict__ M,
                                             size_t b0, bool inside, uint32_t bxs, uint32_t bw,
                                             uint32_t pv[FINAL_BLOCKS], uint32_t mv[FINAL_BLOCKS],
                                             uint32_t root[FINAL_BLOCKS])
{
#pragma unroll
    for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) pv[k] = mv[k] = root[k] = 0;
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
    if (inside) {
        uint32_t prev_parent = key, prev_root = key_root;
#pragma unroll
        for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
            if (!mv[k]) continue;
            root[k] = pv[k] == prev_parent ? prev_root : Forest::find_final(P, pv[k]);
            prev_parent = pv[k];
            prev_root = root[k];
        }
    }
}

// Label values the final kernel writes.
enum : uint32_t {
    kModeRoot = 0,    // 1 + root block index
    kModeCanon = 1,   // canonical numbering (keys: bit positions)
    kModeFirst = 2,   // 1 + the component's first voxel index (keys: voxel indices): the lbeq values
    kModeRef = 3,     // the reference's bleq values: 1 + padded block index of the root
    kModeRefFirst = 4,   // the reference's lbeq values: 1 + padded voxel index of the first voxel
};

// The reference's values and output layout (LBL_LABELS_REFERENCE[_UNCROPPED]):
// its volume is padded to ref_pad3 per dimension with the data at +2, i.e. +1
// block, and the values are 1 + the padded block index (bleq, BLEQ3D_Init,
// LabelingAlgs.cl:828) or 1 + the padded voxel index (lbeq, LBEQ3D_InitKernel,
// LabelingAlgs.cl:641) of the component's minimum, the same block or voxel as
// ours since padding keeps the raster order.
struct Out3 {
    uint32_t off;        // 2: the reference's uncropped volume (data at +2), 0: dense output
    uint32_t oh, ow;     // output rows per plane and row length (h, w when dense)
    uint32_t rbh, rbw;   // padded block grid: ref_pad3(h) / 2, ref_pad3(w) / 2
    uint32_t rvh, rvw;   // padded volume: ref_pad3(h), ref_pad3(w)
};

template <bool VEC, uint32_t MODE>
__global__ __launch_bounds__(NT_FINAL) void bleq3d_final(const uint32_t *__restrict__ P, const uint8_t *__restrict__ M,
                                                          uint32_t *__restrict__ labels, uint32_t d, uint32_t h,
                                                          uint32_t w, uint32_t bd, uint32_t bh, uint32_t bw,
                                                          uint32_t xgroups, uint32_t *count, uint32_t *slots,
                                                          Canon canon, Out3 o)
{
    const uint32_t row = blockIdx.x / xgroups;           // block row index bz * bh + by
    const uint32_t by = row % bh, bz = row / bh;
    const uint32_t bxs = ((blockIdx.x % xgroups) * NT_FINAL + threadIdx.x) * FINAL_BLOCKS;
    const bool inside = bz < bd && bxs < bw;
    const size_t b0 = ((size_t)bz * bh + by) * bw + bxs;
    uint32_t pv[FINAL_BLOCKS], mv[FINAL_BLOCKS], root[FINAL_BLOCKS];
    lane_blocks3<VEC>(P, M, b0, inside, bxs, bw, pv, mv, root);

    if (inside) {
        uint32_t lab[FINAL_BLOCKS];
        uint32_t prev_root = ~0u, prev_lab = 0;
#pragma unroll
        for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
            lab[k] = 0;
            if (!mv[k]) continue;
            if (root[k] != prev_root) {
                const uint32_t r = prev_root = root[k];
                if (MODE == kModeCanon) {
                    prev_lab = canon_label(canon, canon.key[r]);
                } else if (MODE == kModeFirst) {   // key: first voxel index
                    prev_lab = canon.key[r] + 1u;
                } else if (MODE == kModeRefFirst) {
                    const uint32_t key = canon.key[r], plane = h * w;
                    const uint32_t z = key / plane, y = (key % plane) / w, x = key % w;
                    prev_lab = 1u + ((z + 2) * o.rvh + y + 2) * o.rvw + x + 2;
                } else if (MODE == kModeRef) {
                    const uint32_t plane = bh * bw, z = r / plane, y = (r % plane) / bw, x = r % bw;
                    prev_lab = 1u + ((z + 1) * o.rbh + y + 1) * o.rbw + x + 1;
                } else {
                    prev_lab = r + 1u;
                }
            }
            lab[k] = prev_lab;
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
            uint32_t *dst = labels + ((size_t)(z + o.off) * o.oh + y + o.off) * o.ow + x0 + o.off;
            if (VEC && x0 + 2 * FINAL_BLOCKS <= w) {
                if (o.off) {   // 8-byte aligned in the padded volume
                    store_nt2(dst, r[0], r[1]);
                    store_nt2(dst + 2, r[2], r[3]);
                } else {
                    store_nt4(dst, r[0], r[1], r[2], r[3]);
                }
            } else {
                for (uint32_t i = 0; i < 2 * FINAL_BLOCKS && x0 + i < w; ++i) dst[i] = r[i];
            }
        }
    }
    // the tile and merge kernels have finished: the first workgroup sums the count
    if (blockIdx.x == 0 && count && threadIdx.x < kCountSlots) count_reduce(slots, count);
}

// Zeros of the reference's uncropped output outside the input's voxels (the
// final kernel writes those): one workgroup per output row.
__global__ __launch_bounds__(256) void bleq3d_pad_zero(uint32_t *__restrict__ out, uint32_t d, uint32_t h, uint32_t w,
                                                       uint32_t ph, uint32_t pw)
{
    const uint32_t z = blockIdx.x / ph, y = blockIdx.x % ph;
    const bool inner = z >= 2 && z < d + 2 && y >= 2 && y < h + 2;
    uint32_t *row = out + (size_t)blockIdx.x * pw;
    for (uint32_t x = threadIdx.x; x < pw; x += 256)
        if (!inner || x < 2 || x >= w + 2) row[x] = 0;
}

// ---- first-voxel keys; canonical bits -----------------------------------------------
//
// A component's first voxel ([z][y][x] raster order) lies in its front block
// plane, the plane of its root (the minimum block index); each block of that
// plane in the component offers its own first voxel (the lowest mask bit, as
// bit dx + 2 dy + 4 dz orders the block's voxels), and the smallest offer is
// the component's key: its voxel index (VOXEL_KEY; 1 + that is the label
// equivalence value), or (z * h + y) * bw + block column, the bit position of
// canonical numbering.  The keys kernel also clears the key bits (c.nwords
// words); the mark kernel then sets the bit of every root's key.

template <bool VOXEL_KEY>
__global__ __launch_bounds__(NT_FINAL) void bleq3d_keys(const uint32_t *__restrict__ P, const uint8_t *__restrict__ M,
                                                         uint32_t h, uint32_t w, uint32_t bd, uint32_t bh, uint32_t bw,
                                                         uint32_t xgroups, Canon c)
{
    const uint64_t t = (uint64_t)blockIdx.x * NT_FINAL + threadIdx.x;   // covers every word
    if (t < c.nwords) c.bits[t] = 0ull;
    const uint32_t row = blockIdx.x / xgroups;
    const uint32_t by = row % bh, bz = row / bh;
    const uint32_t bxs = ((blockIdx.x % xgroups) * NT_FINAL + threadIdx.x) * FINAL_BLOCKS;
    const bool inside = bz < bd && bxs < bw;
    const size_t b0 = ((size_t)bz * bh + by) * bw + bxs;
    uint32_t pv[FINAL_BLOCKS] = {}, mv[FINAL_BLOCKS] = {};
    if (inside)
        for (uint32_t k = 0; k < FINAL_BLOCKS; ++k)
            if (bxs + k < bw) {
                pv[k] = P[b0 + k];
                mv[k] = M[b0 + k];
            }

    // Roots are never above parents (parent <= self), so only a block whose
    // parent lies in this plane can have its root here: the others need no
    // chase.  A root's key starts at its own block's key, so only offers below
    // the current key are made.
    const uint32_t plane0 = bz * (bh * bw);
    uint32_t r0 = ~0u, k0 = ~0u, r1 = ~0u, k1 = ~0u, prev_p = ~0u, prev_r = 0, prev_k = 0;
#pragma unroll
    for (uint32_t k = 0; k < FINAL_BLOCKS; ++k) {
        const uint32_t m = mv[k];
        if (!m || pv[k] < plane0) continue;
        if (pv[k] != prev_p) {
            prev_p = pv[k];
            prev_r = Forest::find_final(P, pv[k]);
            prev_k = prev_r >= plane0 ? c.key[prev_r] : 0u;
        }
        const uint32_t r = prev_r;
        if (r < plane0) continue;   // not in its component's front plane
        const uint32_t key = block_key3(h, w, bw, bz, by, bxs + k, m, VOXEL_KEY);
        if (key >= prev_k) continue;
        if (r0 == ~0u || r0 == r) {
            r0 = r;
            k0 = min(k0, key);
        } else {
            r1 = r;
            k1 = key;
        }
    }
    const bool has = r0 != ~0u;
    const uint64_t any = __ballot(has);
    if (!any) return;
    const uint32_t R = __shfl(r0, (int)__builtin_ctzll(any), 64);
    if (__ballot(has && (r0 != R || r1 != ~0u)) == 0) {   // one root for the whole wave: one atomic
        uint32_t kmin = k0;
        for (int off = 32; off > 0; off >>= 1) kmin = min(kmin, (uint32_t)__shfl_xor(kmin, off, 64));
        if ((threadIdx.x & 63u) == (uint32_t)__builtin_ctzll(any)) atomicMin(&c.key[R], kmin);
    } else {
        if (has) atomicMin(&c.key[r0], k0);
        if (r1 != ~0u) atomicMin(&c.key[r1], k1);
    }
}

// Set the bit of every component's key.  One thread per 16 consecutive blocks
// (four 16-byte loads of P, one of M; P and M start 256-byte aligned), one
// device atomicOr per root.
__global__ __launch_bounds__(256) void bleq3d_mark(const uint32_t *__restrict__ P, const uint8_t *__restrict__ M,
                                                   uint64_t nb, Canon c)
{
    const uint64_t b0 = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * 16;
    if (b0 >= nb) return;
    uint32_t p[16];
    uint8_t m[16];
    if (b0 + 16 <= nb) {
        const uint4 *p4 = reinterpret_cast<const uint4 *>(P + b0);
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint4 x = p4[q];
            p[4 * q] = x.x; p[4 * q + 1] = x.y; p[4 * q + 2] = x.z; p[4 * q + 3] = x.w;
        }
        const uint4 mm = *reinterpret_cast<const uint4 *>(M + b0);
        const uint32_t mw[4] = {mm.x, mm.y, mm.z, mm.w};
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) m[k] = (uint8_t)(mw[k >> 2] >> (8 * (k & 3)));
    } else {
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) {
            p[k] = b0 + k < nb ? P[b0 + k] : 0u;
            m[k] = b0 + k < nb ? M[b0 + k] : (uint8_t)0;
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
        if (m[k] && p[k] == b0 + k) {   // a root
            const uint32_t key = c.key[b0 + k];
            if ((key >> 6) < c.nwords) atomicOr(&c.bits[key >> 6], 1ull << (key & 63u));
        }
    }
}

}  // namespace

// Bytes of the x-boundary planes (west and east for each of tiles_x - 1 boundaries).
static uint64_t xplane_bytes(uint64_t bd, uint64_t bh, uint64_t bw)
{
    const uint64_t tiles_x = (bw + TX - 1) / TX;
    return 2 * (tiles_x - 1) * bd * bh * sizeof(uint2);
}

size_t bleq3d_workspace(uint32_t d, uint32_t h, uint32_t w, bool canonical)
{
    const uint64_t bd = (d + 1ull) / 2, bh = (h + 1ull) / 2, bw = (w + 1ull) / 2, nb = bd * bh * bw;
    return align256(kCountBytes) + align256(nb * sizeof(uint32_t)) + align256(nb) + align256(xplane_bytes(bd, bh, bw)) +
           (canonical ? canon_workspace((uint64_t)d * h * ((w + 1ull) / 2), nb) : 0);
}

// The block pipeline for both 3D paths: bleq values (1 + root block), or with
// `first` the label-equivalence values (1 + first voxel index: for
// 26-connectivity the voxels of a 2x2x2 block are always connected, so the
// block forest has the voxel forest's components); canonical numbering for
// either in LBL_LABELS_CANONICAL mode.
static hipError_t run_blocks3d(const Image3D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt,
                               bool first)
{
    const uint32_t bd = (im.d + 1u) / 2, bh = (im.h + 1u) / 2, bw = (im.w + 1u) / 2;
    const uint64_t nb = (uint64_t)bd * bh * bw;
    Grid3 g{bd, bh, bw, (bw + TX - 1) / TX, (bh + TY - 1) / TY, 0};
    const uint32_t tiles_z = (bd + TZ - 1) / TZ;
    g.ntiles = g.tiles_x * g.tiles_y * tiles_z;
    uint32_t *slots = static_cast<uint32_t *>(ws);   // count slots first: zero between runs
    uint32_t *P = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(ws) + align256(kCountBytes));
    uint8_t *M = reinterpret_cast<uint8_t *>(P) + align256(nb * sizeof(uint32_t));
    uint2 *xf = reinterpret_cast<uint2 *>(M + align256(nb));
    uint8_t *after_planes = reinterpret_cast<uint8_t *>(xf) + align256(xplane_bytes(bd, bh, bw));
    const bool canonical = opt.mode == LBL_LABELS_CANONICAL, keys = canonical || first;
    Canon canon = keys ? canon_layout(after_planes, (uint64_t)im.d * im.h * bw, nb) : Canon{};
    if (!canonical) canon.nwords = 0;   // keys only
    const bool ref = opt.mode == LBL_LABELS_REFERENCE || opt.mode == LBL_LABELS_REFERENCE_UNCROPPED;
    const bool uncropped = opt.mode == LBL_LABELS_REFERENCE_UNCROPPED;
    const uint32_t pd = (uint32_t)ref_pad3(im.d), ph = (uint32_t)ref_pad3(im.h), pw = (uint32_t)ref_pad3(im.w);
    const Out3 out{uncropped ? 2u : 0u, uncropped ? ph : im.h, uncropped ? pw : im.w, ph / 2, pw / 2, ph, pw};
    const bool vec_in = (im.w % 16 == 0) && (reinterpret_cast<uintptr_t>(im.vox) % 16 == 0);
    const bool vec_out = (uncropped || im.w % 8 == 0) && (reinterpret_cast<uintptr_t>(im.labels) % 16 == 0);
    const uint32_t merge_grid = (uint32_t)(((uint64_t)g.ntiles * (BAR_LANES + X_LANES) + MERGE_JOBS - 1) / MERGE_JOBS);
    const uint32_t xgroups = (bw + NT_FINAL * FINAL_BLOCKS - 1) / (NT_FINAL * FINAL_BLOCKS);

    if (kt) kt->begin(s);
    constexpr uint32_t NT_TILE = Tile3<TX, TY, TZ>::THREADS;
    if (vec_in && im.w % 32 == 0)
        hipLaunchKernelGGL((bleq3d_tile<TX, TY, TZ, true, true>), dim3(g.ntiles), dim3(NT_TILE), 0, s, im.vox, im.d,
                           im.h, im.w, bd, bh, bw, g.tiles_x, g.tiles_y, P, M, xf, im.count, slots, canon.key, !canonical);
    else if (vec_in)
        hipLaunchKernelGGL((bleq3d_tile<TX, TY, TZ, true, false>), dim3(g.ntiles), dim3(NT_TILE), 0, s, im.vox, im.d,
                           im.h, im.w, bd, bh, bw, g.tiles_x, g.tiles_y, P, M, xf, im.count, slots, canon.key, !canonical);
    else
        hipLaunchKernelGGL((bleq3d_tile<TX, TY, TZ, false, false>), dim3(g.ntiles), dim3(NT_TILE), 0, s, im.vox, im.d,
                           im.h, im.w, bd, bh, bw, g.tiles_x, g.tiles_y, P, M, xf, im.count, slots, canon.key, !canonical);
    if (kt) kt->mark(s, "bleq3d_tile");
    hipLaunchKernelGGL(bleq3d_merge, dim3(merge_grid), dim3(MERGE_THREADS), 0, s, P, M, xf, g, im.count, slots);
    if (kt) kt->mark(s, "bleq3d_merge");
    const dim3 fgrid((uint32_t)((uint64_t)xgroups * bh * bd));
    if (keys) {
        // canonical numbering uses bit-position keys (the same for both paths); otherwise voxel indices
        auto kern = canonical ? bleq3d_keys<false> : bleq3d_keys<true>;
        hipLaunchKernelGGL(kern, fgrid, dim3(NT_FINAL), 0, s, P, M, im.h, im.w, bd, bh, bw, xgroups, canon);
        if (kt) kt->mark(s, "bleq3d_keys");
    }
    if (canonical) {
        hipLaunchKernelGGL(bleq3d_mark, dim3((uint32_t)((nb + 16 * 256 - 1) / (16 * 256))), dim3(256), 0, s, P, M, nb,
                           canon);
        if (kt) kt->mark(s, "bleq3d_mark");
        const hipError_t e = canon_scan_launch(canon, s, kt);
        if (e != hipSuccess) return e;
    }
    const uint32_t mode = canonical ? kModeCanon : (first ? (ref ? kModeRefFirst : kModeFirst) : (ref ? kModeRef : kModeRoot));
    auto pick = [mode](auto vec) {
        constexpr bool V = decltype(vec)::value;
        switch (mode) {
        case kModeCanon: return bleq3d_final<V, kModeCanon>;
        case kModeFirst: return bleq3d_final<V, kModeFirst>;
        case kModeRef: return bleq3d_final<V, kModeRef>;
        case kModeRefFirst: return bleq3d_final<V, kModeRefFirst>;
        default: return bleq3d_final<V, kModeRoot>;
        }
    };
    auto fin = vec_out ? pick(std::true_type{}) : pick(std::false_type{});
    if (uncropped) {
        hipLaunchKernelGGL(bleq3d_pad_zero, dim3(pd * ph), dim3(256), 0, s, im.labels, im.d, im.h, im.w, ph, pw);
        if (kt) kt->mark(s, "bleq3d_pad_zero");
    }
    hipLaunchKernelGGL(fin, fgrid, dim3(NT_FINAL), 0, s, P, M, im.labels, im.d, im.h, im.w, bd, bh, bw, xgroups,
                       im.count, slots, canon, out);
    if (kt) kt->mark(s, "bleq3d_final");
    return hipGetLastError();
}

hipError_t bleq3d_run(const Image3D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt)
{
    return run_blocks3d(im, opt, ws, s, kt, false);
}

// Label equivalence in 3D (TOCLLabelEquivalence3D, LabelingAlgs.cpp:1362-1396;
// LBEQ3D_*, LabelingAlgs.cl:635-734) has the same components and labels them
// 1 + minimum voxel index, which is the first voxel in raster order.
hipError_t lbeq3d_run(const Image3D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt)
{
    return run_blocks3d(im, opt, ws, s, kt, true);
}

size_t lbeq3d_workspace(uint32_t d, uint32_t h, uint32_t w, bool)
{
    return bleq3d_workspace(d, h, w, true);
}

}  // namespace lbl
