/*
 * oracle.h -- CPU restatement of the reference labeler (szavalishin/Labeling),
 * used ONLY as test infrastructure.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker (or as the timed CPU baseline).  The
 * product path (labeling_amd) never links, loads or calls it.
 *
 * Every function names the reference file:line it restates.  Labels are uint32,
 * background 0.  Images are row-major [y][x] (2D) or [d][h][w] (3D, w fastest,
 * which is the reference's cv::Mat {rows, cols, slices} memory order with the
 * slice index fastest, main.cpp:191,203).  "Nonzero byte = foreground".
 */
#ifndef LBL_ORACLE_H_
#define LBL_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference padding rules.
 * 2D GPU contract: ((n >> 5) << 5) + 32          (LabelingTools.cpp:103)
 * 3D GPU contract: (((n + 4) >> 5) << 5) + 32,
 *                  data placed at offset +2      (LabelingTools.cpp:160-176,187-195) */
int64_t orc_pad2(int64_t n);
int64_t orc_pad3(int64_t n);

/* Block equivalence 2D (TLabelEquivalenceX2, LabelingAlgs.cpp:336-529).
 * padded=1: run on the image zero-padded by orc_pad2 (the GPU contract), which
 *           makes the algorithm an exact 8-connected labeling; output cropped.
 *           Label = 1 + min padded-block raster index (LabelingAlgs.cpp:404,525),
 *           identical to the GPU raw value (LabelingAlgs.cl:202,317).
 * padded=0: run the reference CPU code on the raw image exactly as it indexes it
 *           (1D indexing with row wrap, LabelingAlgs.cpp:353-376).  Reads that
 *           would fall outside the pixel buffer (undefined in the reference) read
 *           as background and set *oob_reads.
 * threads: OpenMP threads, 0 = all (SetupThreads, LabelingTools.cpp:53-65).
 * iterations (nullable): number of Scan passes run.
 * seconds (nullable): wall time of the DoLabel-equivalent region
 *           (LabelingTools.cpp:27-34): block init, Scan/Analyze loop, write-back.
 * Returns 0 on success, negative on bad arguments. */
int orc_bleq2d(const uint8_t *px, int64_t h, int64_t w, int64_t pitch,
               uint32_t *labels, int padded, int threads,
               uint32_t *iterations, int *oob_reads, double *seconds);

/* Block equivalence 3D (BLEQ3D_* kernels, LabelingAlgs.cl:737-979, host loop
 * LabelingAlgs.cpp:1443-1536) on the volume padded by orc_pad3 with data at +2,
 * output cropped to d*h*w.  Label = 1 + min block index in the padded block
 * grid (LabelingAlgs.cl:828).  No reference CPU 3D labeler exists
 * (main.cpp:103-130), so this restates the OpenCL kernels' arithmetic. */
int orc_bleq3d(const uint8_t *vox, int64_t d, int64_t h, int64_t w,
               uint32_t *labels, int threads, uint32_t *iterations,
               double *seconds);

/* Label equivalence 2D (TLabelDistribution, LabelingAlgs.cpp:211-330;
 * DistrScanKernel etc., LabelingAlgs.cl:31-130).  conn: 0 (default = 4),
 * 4 or 8 (LabelingAlgs.cpp:213, LabelingAlgs.cl:68).
 * quirks=0: the GPU contract (orc_pad2 padding, so 1D offsets never wrap) with
 *           1-based labels as in the reference's 3D lbeq (LabelingAlgs.cl:641):
 *           label = 1 + min padded pixel index.  Exact 4/8-connected labeling.
 * quirks=1: the reference CPU code as written: no padding (rows wrap), 0-based
 *           labels, and index 0 excluded from neighbour reads
 *           (LabelingAlgs.cpp:236,274-277).  Used to pin the documented
 *           deviations (pixel (0,0) gets label 0).
 * quirks=2: the reference GPU kernels' raw values (DistrInitKernel, GetLabel,
 *           DistrScanKernel, LabelingAlgs.cl:31-110): orc_pad2 padding, 0-based
 *           labels = min padded pixel index, position 0 never labelled nor read
 *           as a neighbour; output cropped as IOCLLabeling::Label does
 *           (LabelingTools.cpp:129). */
int orc_lbeq2d(const uint8_t *px, int64_t h, int64_t w, int64_t pitch,
               int conn, uint32_t *labels, int quirks, int threads,
               uint32_t *iterations, double *seconds);

/* Label equivalence 3D (LBEQ3D_* kernels, LabelingAlgs.cl:635-734, host loop
 * LabelingAlgs.cpp:1362-1396) on the orc_pad3 volume, 26-connectivity with
 * bounds-checked neighbours, label = 1 + min padded voxel index, cropped. */
int orc_lbeq3d(const uint8_t *vox, int64_t d, int64_t h, int64_t w,
               uint32_t *labels, int threads, uint32_t *iterations,
               double *seconds);

/* Independent breadth-first labeler (not from the reference): canonical labels
 * 1..N in raster order of each component's first pixel.  conn 2D: 4 or 8;
 * 3D: 6, 18 or 26.  Returns N, or a negative value on bad arguments. */
int64_t orc_bfs2d(const uint8_t *px, int64_t h, int64_t w, int64_t pitch,
                  int conn, uint32_t *labels);
int64_t orc_bfs3d(const uint8_t *vox, int64_t d, int64_t h, int64_t w,
                  int conn, uint32_t *labels);

/* Canonical relabelling in place: scan in memory order, map each unseen
 * nonzero label to 1, 2, 3, ...; background stays 0.  Returns the number of
 * components. */
int64_t orc_canonicalize(uint32_t *labels, int64_t n);

/* Block connectivity bits the reference computes in LBEQ2_Init / InitSPixels
 * for block (bx, by) of the orc_pad2-padded image (LabelingAlgs.cpp:380-433):
 * writes conn[by*bw+bx] (bw = orc_pad2(w)/2, bh = orc_pad2(h)/2) and returns
 * the number of blocks.  Exposed so tests can check the symmetry invariant. */
int64_t orc_bleq2d_conn(const uint8_t *px, int64_t h, int64_t w, int64_t pitch,
                        uint8_t *conn);
/* Same for 3D (BLEQ3D_Init, LabelingAlgs.cl:805-860): 27-bit masks per block of
 * the orc_pad3 volume, block grid (orc_pad3(d)/2, orc_pad3(h)/2, orc_pad3(w)/2). */
int64_t orc_bleq3d_conn(const uint8_t *vox, int64_t d, int64_t h, int64_t w,
                        uint32_t *conn);

/* Exhaustive self-checks: every h x w image (h*w <= 24) through a restated
 * labeler, canonicalised, compared with orc_bfs2d.  alg 0 = orc_bleq2d padded
 * (8-conn), 1 = orc_lbeq2d (conn 4 or 8), 2 = orc_bleq2d unpadded (the
 * reference CPU as written; mismatches expected, F1).  Returns the number of
 * mismatching images and the smallest failing image code (bit i = pixel i). */
int64_t orc_exhaustive2d(int64_t h, int64_t w, int alg, int conn, uint64_t *first_fail);
/* Every d x h x w volume (<= 24 voxels) through orc_bleq3d (alg 0) or
 * orc_lbeq3d (alg 1) versus orc_bfs3d with 26-connectivity. */
int64_t orc_exhaustive3d(int64_t d, int64_t h, int64_t w, int alg, uint64_t *first_fail);

/* ---- the registry's other methods (oracle_registry.c; main.cpp:103-130) ---- */

/* Run equivalence (TRunEqivLabeling, LabelingAlgs.cpp:535-810; its OpenCL
 * kernels REFindRunsKernel .. RELabelKernel, LabelingAlgs.cl:348-616, compute
 * the same on the padded image).  Runs of each row are numbered in slots of
 * `slots` per row (the reference: w >> 1; a row with more runs returns -2
 * instead of spilling into the next row's slots); runs of adjacent rows join
 * where IsNeib (:605-613) finds a shared column; the Scan / Analyze loop runs
 * sequentially to its fixed point.  labels: 1 + row * slots + the index of the
 * component's first run in its row. */
int orc_runeq2d(const uint8_t *px, int64_t h, int64_t w, int64_t slots, uint32_t *labels, uint32_t *iterations);

/* He et al.'s run labeling (TRunLabeling::DoLabel / SetRunLabel,
 * LabelingAlgs.cpp:58-205): conpix 1 for 8-connectivity (coh == COH_8), 0 for
 * 4.  quirks = 1: as written: the last row is not scanned (Row < Bottom =
 * rows - 1, :150-153), a run reaching the last column loses its last pixel and
 * a run starting there is dropped (:158-171).  quirks = 0: every pixel. */
int orc_herun2d(const uint8_t *px, int64_t h, int64_t w, int conpix, int quirks, uint32_t *labels);

/* Block-based two-pass 8-connected labeling after Grana et al. (the decisions
 * of cvLabelingImageLab, cvlabeling_imagelab.cpp:57-1567, without its decision
 * tree): 2x2 blocks in raster order take the provisional label of the earlier
 * neighbour blocks they touch and merge the others (the smallest label stays
 * the representative), else a new label; labels are renumbered 1..N in order
 * of their representative (:1523-1531).  Returns N. */
int64_t orc_grblock2d(const uint8_t *px, int64_t h, int64_t w, uint32_t *labels);

#ifdef __cplusplus
}
#endif

#endif /* LBL_ORACLE_H_ */
