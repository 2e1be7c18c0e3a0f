// pending.cpp -- paths not built yet; removed as each lands.
#include "lbl_internal.hpp"
namespace lbl {
size_t bleq3d_workspace(uint32_t, uint32_t, uint32_t) { return 256; }
size_t lbeq2d_workspace(uint32_t, uint32_t) { return 256; }
size_t lbeq3d_workspace(uint32_t, uint32_t, uint32_t) { return 256; }
hipError_t bleq3d_run(const Image3D &, const Options &, void *, hipStream_t, KernelTimer *) { return hipErrorNotSupported; }
hipError_t lbeq2d_run(const Image2D &, int, const Options &, void *, hipStream_t, KernelTimer *) { return hipErrorNotSupported; }
hipError_t lbeq3d_run(const Image3D &, const Options &, void *, hipStream_t, KernelTimer *) { return hipErrorNotSupported; }
}
