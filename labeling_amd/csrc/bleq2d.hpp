// bleq2d.hpp -- launch geometry of the 2D block-equivalence kernels, shared by
// the single-image entry points and the batch planner.
#pragma once

#include "batch_plan.hpp"
#include "canon.hpp"

namespace lbl {

enum : uint32_t {
    kVecIn = 1u,       // image rows start 16-byte aligned (pitch and base)
    kRefLabels = 4u,   // write the reference's raw label values
    kKeys = 8u,        // keep first-pixel keys of roots (canonical or label equivalence values)
    kPixelKeys = 16u,  // keys are pixel indices (else canonical bit positions)
    kCanonLabels = 32u,  // write canonical labels (batches: per image)
    kDropOrigin = 64u,   // pixel (0,0) reads as background (the reference's 2D label-equivalence values)
    kVecOut = 128u,      // label rows start 16-byte aligned (w % 4 == 0 and base)
    kExact = 256u,       // kVecIn, kVecOut and w % 32 == 0: every lane's segments and stores are whole
};

// One image as the kernels see it (also the element of a batch's device table).
struct Geom2D {
    const uint8_t *px;
    uint32_t *labels;
    uint32_t *count;       // component count (may be null)
    uint32_t *P;           // block words: bw * bh uint32
    uint32_t *slots;       // kCountBytes of partial counts (workspace, before P; zero between runs)
    uint64_t pitch;
    uint32_t h, w, bw, bh;
    plan::DivMagic bdiv;        // division by bw (the final kernels' row of a raster block index)
    uint32_t tiles_x, ntiles;
    uint32_t nfinal;            // final-kernel workgroups: chunks of 512 blocks in raster order (plan::geometry2d)
    uint32_t tile_base;         // first global tile id of this image (batch)
    uint32_t final_base;        // first global final workgroup id (batch)
    uint32_t chunk_base;        // first global canonical scan chunk id (batch)
    uint32_t flags;
    uint32_t ref_bw;            // block-grid width of the reference's padded image
    uint32_t w64;               // run-bit key row width: w rounded up to 64 (he-run / runeq values)
    uint32_t run_w2;            // run numbers per row of the runeq values
    Canon canon;                // canonical numbering state (kCanon)
};

using TileShape2D = plan::TileShape;
constexpr uint32_t kBleq2dTileCount = plan::kTileCount;
constexpr const TileShape2D *kBleq2dTiles = plan::kTiles;

// Device tables of a batch: the image of every tile id, final workgroup id and
// canonical scan chunk id (plan::Batch).
struct BatchTables {
    const uint32_t *tile_img, *final_img, *chunk_img;
};

// Fill bw..nfinal and the alignment flags for a tile shape (px, labels, pitch,
// h, w and the kRefLabels / kCanon flags must be set); batch: an image of a
// batch (the batched kernels' final workgroups take kBatchFinalChunks chunks).
void bleq2d_geometry(Geom2D &g, uint32_t tile_id, bool batch = false);

// total_chunks: the images' canonical scan chunks (0: no image asks for canonical labels)
hipError_t bleq2d_run_batch(const Geom2D *d_geoms, uint32_t n, const BatchTables &tables, uint32_t total_tiles,
                            uint32_t total_final, uint32_t total_chunks, uint32_t tile_id, bool flatten, hipStream_t s,
                            KernelTimer *kt);

}  // namespace lbl
