/*
 * labeling_amd.h -- C ABI of the MI355X-native connected-component labeler.
 *
 * Drop-in boundary for szavalishin/Labeling's algorithm-selection and
 * label-call interface, restricted to the two GPU methods of that project:
 *   "bleq"  block equivalence (2D on 2x2 blocks, 8-connectivity;
 *           3D on 2x2x2 blocks, 26-connectivity)
 *   "lbeq"  label equivalence after Kalentev (2D 4/8-connectivity, 3D 26)
 * A binary image (or a slice stack) goes in, a uint32 label image comes out.
 *
 * Conventions
 *  - Input: uint8, any nonzero byte is foreground.  2D images are row-major
 *    [y][x] with a byte pitch; 3D volumes are dense [d][h][w] with w fastest.
 *    The reference's 3D cv::Mat {rows, cols, slices} (slice fastest,
 *    main.cpp:191,203) is passed as d = rows, h = cols, w = slices.
 *  - Output: dense uint32, same shape as the input (3D is cropped, unlike the
 *    reference, LabelingTools.cpp:217-220), background 0, allocated by the
 *    caller.  Label values follow lbl_label_mode.
 *  - Errors: every entry point returns lbl_status and never throws; the
 *    message of the last failure on a context is lbl_last_error().  The
 *    reference throws std::exception instead (LabelingTools.hpp:20-27).
 *  - Threading: a context serialises its calls (the reference's labeler
 *    objects are not re-entrant, LabelingAlgs.hpp:247-252).  Use one context
 *    per host thread for concurrency.
 *  - Streams: `stream` arguments are hipStream_t values passed as void*;
 *    NULL selects the context's own stream.
 */
#ifndef LABELING_AMD_H_
#define LABELING_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbl_ctx lbl_ctx;

/* Registry names of ALG_LIST (main.cpp:103-130; lbl_alg_from_name).  Every
 * method runs on the GPU; the reference's CPU-only methods (he-run, gr-block,
 * ocv) are computed from the same block forests with their label values.
 *   LBL_BLEQ    "bleq"     block equivalence; 2D 8-connectivity (4 rejected), 3D 26.
 *   LBL_LBEQ    "lbeq"     label equivalence; 2D 4 (default) or 8, 3D 26.
 *   LBL_BIN     "bin"      binarization: label 1 on every foreground pixel, every
 *                          mode (TBinLabeling, LabelingAlgs.cpp:23-36;
 *                          BinLabelingKernel, LabelingAlgs.cl:18-25); 2D and 3D;
 *                          count 1 if there is foreground, else 0.
 *   LBL_RUNEQ   "runeq"    run equivalence (TRunEqivLabeling / TOCLRunEquivLabeling,
 *                          LabelingAlgs.cpp:535-810,1090-1268): connectivity 4
 *                          rejected like the reference, yet runs of adjacent
 *                          rows join only where they share a column (IsNeib,
 *                          LabelingAlgs.cpp:605-613), i.e. 4-connected components.
 *                          Root mode: 1 + y * ceil(w/2) + k for the component's
 *                          first run, the k-th (0-based) run of row y (the CPU
 *                          values, w >> 1 run slots per row for even w);
 *                          reference mode: 1 + y * (P2(w)/2) + k (the GPU values
 *                          on the padded image).  2D only.
 *   LBL_HERUN   "he-run"   He et al. run labeling (TRunLabeling, LabelingAlgs.cpp:58-205):
 *                          4-connectivity by default, 8 on request; root and
 *                          reference mode: the provisional label the reference
 *                          keeps for a component, i.e. the number of runs up to
 *                          and including its first run that have no pixel of
 *                          the row above in reach.  Labels every pixel (the
 *                          reference skips the last row and the last column's
 *                          runs, LabelingAlgs.cpp:153,162-171).  2D only.
 *   LBL_GRBLOCK "gr-block" Grana et al. block labeling (TBlockGranaLabeling,
 *                          LabelingAlgs.cpp:47-56): 8-connectivity (4 rejected);
 *                          root and reference mode: 1..N in the order of each
 *                          component's first 2x2 block in block raster order
 *                          (cvlabeling_imagelab.cpp:1523-1531).  2D only.
 *   LBL_OCV     "ocv"      cv::connectedComponents (TOpenCVLabeling, LabelingAlgs.cpp:39-45):
 *                          8-connectivity when 8 is asked for, else 4; every mode:
 *                          1..N in raster order of first pixels.  2D only.
 * LBL_LABELS_CANONICAL gives 1..N in raster order of first pixels for every
 * method except bin. */
typedef enum {
    LBL_BLEQ = 0,
    LBL_LBEQ = 1,
    LBL_BIN = 2,
    LBL_RUNEQ = 3,
    LBL_HERUN = 4,
    LBL_GRBLOCK = 5,
    LBL_OCV = 6
} lbl_alg;

typedef enum {
    LBL_OK = 0,
    LBL_EINVAL = 1,          /* bad argument (null pointer, zero size, bad pitch) */
    LBL_EUNSUPPORTED = 2,    /* e.g. bleq with 4-connectivity (LabelingAlgs.cpp:338,1070),
                                3D with a connectivity other than 26 (LabelingTools.cpp:185),
                                images beyond the index limits */
    LBL_ENOMEM = 3,          /* device allocation failed, or a request exceeds a budget */
    LBL_EDEVICE = 4,         /* a HIP runtime call failed */
    LBL_EITER = 5,           /* an iteration cap was reached (debug knob) */
    LBL_EBUSY = 6            /* resource temporarily unavailable (serving admission) */
} lbl_status;

/* Values written to foreground pixels.  All modes are deterministic. */
typedef enum {
    /* 1 + raster index (in the unpadded input) of the component's root:
     * bleq: root = min 2x2 (2x2x2) block index, block grid ceil(h/2) x ceil(w/2);
     * lbeq: root = min pixel (voxel) index.  Fastest mode; default. */
    LBL_LABELS_ROOT = 0,
    /* The reference OpenCL output values, computed from the roots above, with
     * the reference's padding ((n>>5)<<5)+32 in 2D, LabelingTools.cpp:103, and
     * (((n+4)>>5)<<5)+32 with the data at offset +2 in 3D, :160-176,187-195):
     *  2D bleq: 1 + min block index in the padded block grid
     *           (LabelingAlgs.cl:202,317);
     *  2D lbeq: the min pixel index in the padded image, 0-based, and pixel
     *           (0,0) is never labelled: it keeps 0 (background) and neighbours
     *           do not connect through it (DistrInitKernel / GetLabel,
     *           LabelingAlgs.cl:31-52);
     *  3D bleq: 1 + min block index in the padded [d][h][w] block grid
     *           (BLEQ3D_Init, LabelingAlgs.cl:828);
     *  3D lbeq: 1 + min voxel index in the padded volume (LabelingAlgs.cl:641).
     * The output stays dense and cropped to the input's shape; the component
     * count is the number of distinct nonzero labels written. */
    LBL_LABELS_REFERENCE = 1,
    /* 1..N in raster order of each component's first pixel (voxel), x
     * fastest: the canonical form used for parity, computed on the device.
     * Every entry point: single images and volumes of both algorithms,
     * lbl_label2d_each, batches (per image), and the server (which then runs
     * each request on its own).  Costs a
     * key pass over the block words plus a bit-array scan (bleq), or a relabel
     * pass over the labels (lbeq), and a larger workspace. */
    LBL_LABELS_CANONICAL = 2,
    /* 3D only (lbl_label3d, lbl_label3d_dev): the LBL_LABELS_REFERENCE values
     * written into the reference's uncropped output volume
     * (IOCLLabeling3D::Label returns the padded volume, LabelingTools.cpp:196-220):
     * P3(d) x P3(h) x P3(w) labels, P3(n) = (((n+4)>>5)<<5)+32, the input's
     * labels at offset +2 in every dimension and 0 elsewhere.  The label buffer
     * must hold P3(d)*P3(h)*P3(w) elements.  2D calls return LBL_EUNSUPPORTED
     * (the reference crops 2D output, LabelingTools.cpp:129). */
    LBL_LABELS_REFERENCE_UNCROPPED = 3
} lbl_label_mode;

/* Connectivity argument: 0 = default (bleq 2D: 8, lbeq 2D: 4, 3D: 26),
 * otherwise 4, 8 (2D) or 26 (3D).  Mirrors TCoherence (LabelingTools.hpp:45-50). */

/* ---- context ------------------------------------------------------------ */

/* Replaces the reference's labeler construction + IOCLLabeling::Init
 * (LabelingTools.cpp:81-94): binds a device, creates a stream and events.
 * Kernels are compiled offline for gfx950 (no run-time program build). */
lbl_status lbl_create(lbl_ctx **out, int device_ordinal);

/* Replaces ~IOCLLabeling / TerminateOCL (LabelingTools.cpp:136-154). */
void lbl_destroy(lbl_ctx *ctx);

/* Message of the last failed call on ctx (replaces the exception text of
 * THROW_IF / THROW_IF_OCL, LabelingTools.hpp:20-27).  ctx may be NULL (then the
 * message of the last failed lbl_create on this thread). */
const char *lbl_last_error(const lbl_ctx *ctx);

const char *lbl_status_string(lbl_status s);

lbl_status lbl_set_label_mode(lbl_ctx *ctx, lbl_label_mode mode);

/* Debug knob for 2D lbeq.  cap > 0 selects the Kalentev sweep kernels (per-tile
 * scan/analyze passes to a fixed point in LDS, then border unions; the
 * reference's DistrScanKernel / DistrAnalizeKernel, LabelingAlgs.cl:86-130)
 * instead of the run-based block forests, and caps the sweeps per tile;
 * exceeding the cap returns LBL_EITER from host-buffer calls.  0 = no cap and
 * the faster block-forest path (default).  LBL_LABELS_REFERENCE is not
 * available on the sweep path (LBL_EUNSUPPORTED). */
lbl_status lbl_set_iteration_cap(lbl_ctx *ctx, uint32_t cap);

/* Registry lookup by the reference's algorithm names, mirroring
 * SetLabelingAlg (main.cpp:373-395): "bleq", "lbeq", "bin", "runeq", "he-run",
 * "gr-block" or "ocv" (every ALG_LIST entry, main.cpp:103-130).  Returns
 * LBL_EINVAL for unknown names. */
lbl_status lbl_alg_from_name(const char *name, lbl_alg *out);

/* ---- single image, host buffers ------------------------------------------ */

/* 2D labeling of a host image.  Replaces IOCLLabeling::Label
 * (LabelingTools.hpp:110, LabelingTools.cpp:98-132) running
 * TOCLLabelEquivalenceX2::DoOCLLabel (LabelingAlgs.cpp:1067-1084) for bleq or
 * TOCLLabelDistribution::DoOCLLabel (LabelingAlgs.cpp:905-942) for lbeq.
 * labels: h*w dense, caller-allocated.  n_components (nullable): number of
 * components.  elapsed_us (nullable): device time of the labeling kernels
 * only, excluding transfers and allocation -- the reference's timed region
 * (LabelingTools.cpp:118-125). */
lbl_status lbl_label2d(lbl_ctx *ctx, lbl_alg alg, const uint8_t *px, uint32_t h, uint32_t w,
                       size_t pitch_bytes, int conn, uint32_t *labels, uint32_t *n_components,
                       float *elapsed_us);

/* 3D labeling of a host volume.  Replaces IOCLLabeling3D::Label
 * (LabelingTools.hpp:154, LabelingTools.cpp:180-221) running
 * TOCLBlockEquivalence3D::DoOCLLabel3D (LabelingAlgs.cpp:1520-1536) for bleq or
 * TOCLLabelEquivalence3D::DoOCLLabel3D (LabelingAlgs.cpp:1362-1396) for lbeq,
 * TOCLBinLabeling3D (LabelingAlgs.cpp:1270-1320) for bin; the registry has no
 * 3D variant of the other methods (LBL_EUNSUPPORTED).
 * labels: d*h*w dense (cropped). */
lbl_status lbl_label3d(lbl_ctx *ctx, lbl_alg alg, const uint8_t *vox, uint32_t d, uint32_t h,
                       uint32_t w, int conn, uint32_t *labels, uint32_t *n_components,
                       float *elapsed_us);

/* ---- single image, device buffers ----------------------------------------- */

/* Same as lbl_label2d on device-resident buffers, enqueued on `stream` without
 * any host synchronisation (the reference's DoOCLLabel body with its blocking
 * per-iteration flag transfers removed, LabelingAlgs.cpp:1028-1038).
 * d_ncomp (nullable, device): receives the component count. */
lbl_status lbl_label2d_dev(lbl_ctx *ctx, lbl_alg alg, const uint8_t *d_px, uint32_t h, uint32_t w,
                           size_t pitch_bytes, int conn, uint32_t *d_labels, uint32_t *d_ncomp,
                           void *stream);

lbl_status lbl_label3d_dev(lbl_ctx *ctx, lbl_alg alg, const uint8_t *d_vox, uint32_t d, uint32_t h,
                           uint32_t w, int conn, uint32_t *d_labels, uint32_t *d_ncomp, void *stream);

/* lbl_label2d_dev that allocates nothing.  No counterpart in the reference
 * (whose labelers create their device buffers per call,
 * LabelingAlgs.cpp:996-997); it is for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The union-find workspace is a fixed 72 MiB device array
 * per device, held until the process ends: images whose workspace exceeds it
 * return LBL_ENOMEM (2D bleq with root or reference values: at most
 * 18,873,344 2x2 blocks, e.g. 8688 x 8688 pixels; lbeq and canonical values
 * need more per pixel).  Every call on a device shares that array
 * and clears what must start at zero on its own stream, so calls that overlap
 * on one device race: serialise them (one stream, or events).  Labels, count
 * and errors are those of lbl_label2d_dev.  bleq and lbeq only (other methods:
 * LBL_EUNSUPPORTED). */
lbl_status lbl_label2d_dev_noalloc(lbl_ctx *ctx, lbl_alg alg, const uint8_t *d_px, uint32_t h, uint32_t w,
                                   size_t pitch_bytes, int conn, uint32_t *d_labels, uint32_t *d_ncomp,
                                   void *stream);

/* Device time in microseconds of each kernel of the most recent *_dev call on
 * ctx when profiling is enabled (lbl_set_profiling).  Writes up to `cap`
 * entries to us[] and the kernel names to names[] (static strings), returns
 * the number of kernels.  Blocks until that call's work is done. */
lbl_status lbl_set_profiling(lbl_ctx *ctx, int enable);
int lbl_last_kernel_times(lbl_ctx *ctx, float *us, const char **names, int cap);

/* ---- batches of 2D images ------------------------------------------------------ */

/* One device-resident 2D image of a batch. */
typedef struct {
    const uint8_t *px;      /* device, h rows of pitch_bytes */
    uint32_t *labels;       /* device, h*w dense */
    uint32_t h, w;
    size_t pitch_bytes;
} lbl_image2d;

typedef struct lbl_batch lbl_batch;

enum {
    LBL_BATCH_AUTOTUNE = 1,   /* time the tile shapes and merge schedules on this batch, keep the fastest */
    LBL_BATCH_NO_GRAPH = 2    /* lbl_batch_run launches the kernels directly instead of replaying a hipGraph */
};

/* Prepares n images of uneven sizes for labeling as one unit: a device table of
 * per-image geometry, one workspace, the launch configuration (autotuned per
 * size class, or from the context's tuning table), and a hipGraph holding the
 * four batched kernels.  No reference counterpart: the reference labels a
 * directory with a sequential loop of Label() calls (ProcessImages,
 * main.cpp:309-340), which is what one lbl_label2d_dev call per image mirrors.
 * Each image is labelled exactly as lbl_label2d_dev would label it alone.
 * d_ncomp (nullable, device): n per-image component counts.  The device
 * buffers must stay valid while the batch exists.  bleq only (LBL_EUNSUPPORTED
 * for lbeq). */
lbl_status lbl_batch_create(lbl_ctx *ctx, lbl_alg alg, const lbl_image2d *images, uint32_t n, int conn,
                            uint32_t *d_ncomp, unsigned flags, lbl_batch **out);

/* Labels every image of the batch: one hipGraph launch on `stream`. */
lbl_status lbl_batch_run(lbl_batch *b, void *stream);

/* Launch configuration chosen for the batch: tile shape id and size in 2x2
 * blocks, whether the flatten pass runs. */
lbl_status lbl_batch_config(const lbl_batch *b, uint32_t *tile_id, uint32_t *tile_w, uint32_t *tile_h,
                            int *flatten);

void lbl_batch_destroy(lbl_batch *b);

/* Labels n device images one after another with their own per-image launch
 * sequences (what n lbl_label2d_dev calls do, without per-call overhead in the
 * caller): the device-side form of the reference's sequential directory loop
 * (ProcessImages, main.cpp:309-340), and the baseline a batch is measured
 * against.  d_ncomp (nullable, device): n counts. */
lbl_status lbl_label2d_each(lbl_ctx *ctx, lbl_alg alg, const lbl_image2d *images, uint32_t n, int conn,
                            uint32_t *d_ncomp, void *stream);

/* ---- autotuning ----------------------------------------------------------------- */

/* Times every tile shape and merge schedule of the 2D path on one device image
 * and records the fastest for images of its size class (floor(log2(h*w))) in
 * the context's tuning table, which later 2D calls of that class use.  The
 * timed launches run on the context's stream, after the work enqueued on
 * `stream` (where d_px is ready) so far; on return `stream` waits for them
 * (they write d_labels).  stream NULL: the context's stream. */
lbl_status lbl_autotune2d(lbl_ctx *ctx, lbl_alg alg, const uint8_t *d_px, uint32_t h, uint32_t w,
                          size_t pitch_bytes, uint32_t *d_labels, void *stream);

/* Tuning table as JSON ({"<alg>/<kind>/<size class>": [tile, flatten, us], ...}). */
lbl_status lbl_tuning_save(lbl_ctx *ctx, const char *path);
lbl_status lbl_tuning_load(lbl_ctx *ctx, const char *path);

/* ---- serving ------------------------------------------------------------------------
 * No reference counterpart (the reference labels one image per call).  A server
 * owns a device arena of `hbm_budget_bytes` and `streams` device streams, and a
 * scheduler thread that does continuous batching: requests that arrive while
 * others are on the GPU are admitted as soon as a stream is free and their
 * input, labels and union-find workspace fit in the arena; consecutive 2D bleq
 * requests are fused into one batched launch.  Each result is identical to the
 * single-request entry points' result on the same input. */

typedef struct lbl_server lbl_server;

typedef struct {
    uint32_t streams;            /* concurrent device streams (1..16) */
    uint64_t hbm_budget_bytes;   /* device arena for in-flight requests */
    uint32_t max_batch;          /* most 2D requests fused into one launch (>= 1) */
} lbl_server_config;

typedef struct {
    lbl_alg alg;
    int conn;                    /* as for lbl_label2d / lbl_label3d */
    uint32_t d, h, w;            /* d = 0 for a 2D image */
    const uint8_t *px;           /* host input, dense; pinned memory (lbl_host_alloc) lets copies overlap */
    uint32_t *labels;            /* host output, dense */
} lbl_request;

typedef struct {
    uint64_t ticket;
    lbl_status status;
    uint32_t n_components;
    double t_submit, t_start, t_done;   /* seconds on the server's steady clock */
} lbl_completion;

lbl_status lbl_server_create(lbl_ctx *ctx, const lbl_server_config *cfg, lbl_server **out);
/* Queues a request (never blocks on the device).  LBL_ENOMEM when the request
 * alone exceeds the arena; the buffers must stay valid until it completes. */
lbl_status lbl_server_submit(lbl_server *srv, const lbl_request *req, uint64_t *ticket);
/* Moves up to `cap` finished requests into out[]; *n receives the count. */
lbl_status lbl_server_poll(lbl_server *srv, lbl_completion *out, uint32_t cap, uint32_t *n);
/* Blocks until every submitted request has finished. */
lbl_status lbl_server_drain(lbl_server *srv);
/* Testing knob: the next n groups (a fused launch or a single request) the
 * server launches fail at their kernel launch, before anything is enqueued
 * there, as a failing hipLaunchKernel would.  Their requests complete with
 * status LBL_EDEVICE and their label buffers are not written; later groups run
 * normally. */
lbl_status lbl_server_inject_faults(lbl_server *srv, uint32_t n);
/* Seconds on the server's clock (the clock of lbl_completion). */
double lbl_server_now(const lbl_server *srv);
void lbl_server_destroy(lbl_server *srv);

/* Page-locked host buffers for requests and results. */
void *lbl_host_alloc(size_t bytes);
void lbl_host_free(void *p);

#ifdef __cplusplus
}
#endif

#endif /* LABELING_AMD_H_ */
