// lbl_internal.hpp -- context, error plumbing and launch descriptors shared by
// the C ABI and the HIP kernels.  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/labeling_amd.h"

namespace lbl {

// Launch description of one 2D image, in the units the kernels use.
struct Image2D {
    const uint8_t *px;     // device
    size_t pitch;          // bytes
    uint32_t h, w;
    uint32_t *labels;      // device, dense h*w
    uint32_t *count;       // device, component count (written)
};

struct Image3D {
    const uint8_t *vox;    // device, dense d*h*w
    uint32_t d, h, w;
    uint32_t *labels;      // device, dense
    uint32_t *count;       // device
};

// Per-call settings shared by every algorithm.
struct Options {
    lbl_label_mode mode = LBL_LABELS_ROOT;
    uint32_t iter_cap = 0;       // lbeq tile sweeps cap, 0 = none
    uint32_t *status = nullptr;  // device word; kernels OR in error bits (bit0 = iteration cap hit)
    uint32_t tile2d = 0;         // bleq 2D tile shape id (kBleq2dTiles; per call: tuning table, else plan::default_tile2d)
    bool flatten = false;        // run the flatten pass (else the final pass chases full paths)
};

// Autotuned launch configuration for a workload class.
struct TuneKey {
    int alg, kind;        // kind: 0 = single 2D image, 1 = 2D batch
    int size_class;       // floor(log2(pixels per image))
    bool operator<(const TuneKey &o) const
    {
        return alg != o.alg ? alg < o.alg : (kind != o.kind ? kind < o.kind : size_class < o.size_class);
    }
};
struct TuneVal {
    uint32_t tile;
    bool flatten;
    float us;
};
int size_class_of(uint64_t pixels);

// A named kernel launch recorded for optional per-kernel timing.
struct KernelTimer {
    bool enabled = false;
    std::vector<hipEvent_t> events;      // events[i] .. events[i+1] brackets kernel i
    std::vector<const char *> names;
    size_t used = 0;
    void begin(hipStream_t s);
    void mark(hipStream_t s, const char *name);
};

// Workspace sizes (bytes) needed by each path; 256-byte aligned sub-buffers.
// `canonical`: room for canonical numbering (LBL_LABELS_CANONICAL) as well.
size_t bleq2d_workspace(uint32_t h, uint32_t w, bool canonical = false);
size_t bleq3d_workspace(uint32_t d, uint32_t h, uint32_t w, bool canonical = false);
size_t lbeq2d_workspace(uint32_t h, uint32_t w, bool canonical = false);
size_t lbeq3d_workspace(uint32_t d, uint32_t h, uint32_t w, bool canonical = false);

// Enqueue the full labeling pipeline of one image on `s`.  `ws` must hold the
// workspace size above.  Return hipSuccess or the first launch error.
// first: label equivalence values (1 + first pixel index) from the bleq block
// forest, for 8-connectivity (the pixels of a 2x2 block are 8-connected).
hipError_t bleq2d_run(const Image2D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt,
                      bool first = false);
// Label equivalence values with 4-connectivity from the same kernels on 1x2
// blocks (one pixel row per band), and its workspace.
hipError_t bleq2d_run_conn4(const Image2D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt);
size_t bleq2d_conn4_workspace(uint32_t h, uint32_t w, bool canonical);
hipError_t bleq3d_run(const Image3D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt);

// The reference registry's other methods (ALG_LIST, main.cpp:103-130), on the
// same block forests: gr-block (block-order numbering), he-run / runeq (run
// numbering; he: He et al., else run equivalence), and bin (foreground -> 1,
// 2D and 3D; ws holds the count slots).
size_t grblock2d_workspace(uint32_t h, uint32_t w);
hipError_t grblock2d_run(const Image2D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt);
size_t runs2d_workspace(uint32_t h, uint32_t w, bool conn8, bool canonical);
hipError_t runs2d_run(const Image2D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt, bool he,
                      int conn);
hipError_t bin2d_run(const Image2D &im, void *ws, hipStream_t s, KernelTimer *kt);
hipError_t bin3d_run(const Image3D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt);
hipError_t lbeq2d_run(const Image2D &im, int conn, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt);
hipError_t lbeq3d_run(const Image3D &im, const Options &opt, void *ws, hipStream_t s, KernelTimer *kt);

// Copy h rows of w bytes from host memory (row pitch spitch) to device rows of
// pitch dpitch on stream s.  Unequal pitches go through `staging` (device
// memory, at least spitch * (h - 1) + w bytes; otherwise a 2D copy is used)
// and a re-pitch kernel.
hipError_t upload2d(uint8_t *dst, size_t dpitch, const uint8_t *src, size_t spitch, uint32_t w, uint32_t h,
                    uint8_t *staging, size_t staging_bytes, hipStream_t s);

// Component counts: kCountSlots partial counters kCountStride words apart (a
// 64-byte line each) at the start of every workspace; zero between runs.
constexpr uint32_t kCountSlots = 64;
constexpr uint32_t kCountStride = 16;   // words between slots
constexpr size_t kCountBytes = kCountSlots * kCountStride * sizeof(uint32_t);

// Limits checked on the host before any launch.
constexpr uint64_t kMaxBlocks2D = (1ull << 28);   // 2D block word packs parent << 4 | mask
constexpr uint64_t kMaxElems = (1ull << 32) - 2;  // labels are uint32 with 0 = background

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Fixed device scratch of lbl_label2d_dev_noalloc (noalloc.hip): its size, and
// the device address of the calling device's copy (resolved once per device).
extern const size_t kScratchBytes;
hipError_t scratch_setup(int device, void **ptr);

// The reference's padded sizes, which fix its raw label values: 2D
// ((n >> 5) << 5) + 32 with the image at the origin (LabelingTools.cpp:103);
// 3D (((n + 4) >> 5) << 5) + 32 with the data at offset +2 (LabelingTools.cpp:160-176,187-195).
__host__ __device__ inline uint64_t ref_pad2(uint64_t n) { return ((n >> 5) << 5) + 32; }
__host__ __device__ inline uint64_t ref_pad3(uint64_t n) { return (((n + 4) >> 5) << 5) + 32; }

}  // namespace lbl

// The context object behind the opaque lbl_ctx handle.
struct lbl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    std::string last_error;
    lbl::Options opt;
    bool profiling = false;
    lbl::KernelTimer timer;

    // cached device buffers
    void *ws = nullptr;
    size_t ws_bytes = 0;
    uint8_t *d_in = nullptr;
    size_t in_bytes = 0;
    uint32_t *d_lab = nullptr;
    size_t lab_bytes = 0;
    uint32_t *d_misc = nullptr;   // [0] = component count, [1] = status bits
    uint32_t *h_misc = nullptr;   // pinned host mirror of d_misc
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::map<lbl::TuneKey, lbl::TuneVal> tuning;
    void *scratch = nullptr;        // this device's fixed scratch (device address), or null
    std::string scratch_error;      // why it is null
};

namespace lbl {
lbl_status label2d_in_workspace(lbl_ctx *ctx, lbl_alg alg, const uint8_t *d_px, uint32_t h, uint32_t w,
                                size_t pitch_bytes, int conn, uint32_t *d_labels, uint32_t *d_ncomp, void *stream,
                                void *ws, size_t ws_bytes, hipError_t (*prologue)(void *, hipStream_t));
}  // namespace lbl
