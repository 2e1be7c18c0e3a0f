// LabelingTools.hpp -- C++ host interface mirroring the reference's labeling
// framework (LabelingTools.hpp:35-177) and algorithm registry (main.cpp:88-131,
// 373-395) with every method of that registry, implemented over the C ABI in
// include/labeling_amd.h.
//
// Same names, argument meaning and error behaviour as the reference:
//   ILabeling::Label(pixels, labels, threads, coh) -> elapsed microseconds of the
//   labeling kernels (LabelingTools.hpp:64, LabelingTools.cpp:98-132,180-221);
//   errors are thrown (LabelingTools.hpp:20-27), as std::runtime_error because
//   std::exception(const char*) is not standard C++.
// TImage replaces cv::Mat: a dense 2D (rows x cols) or 3D (size[0] x size[1] x
// size[2], last index fastest) array of uint8 pixels or uint32 labels.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/labeling_amd.h"

namespace LabelingTools {

typedef uint8_t TPixel;     // LabelingTools.hpp:38
typedef uint32_t TLabel;    // LabelingTools.hpp:39
typedef uint32_t TTime;     // microseconds, LabelingTools.hpp:41

enum TCoherence { COH_4, COH_8, COH_DEFAULT };   // LabelingTools.hpp:45-50
const char MAX_THREADS = 0;                       // LabelingTools.hpp:52

class TImage {
public:
    enum Type { U8, U32 };

    TImage() = default;
    TImage(int rows, int cols, Type t) { create({rows, cols}, t); }
    TImage(int d0, int d1, int d2, Type t) { create({d0, d1, d2}, t); }

    void create(const std::vector<int> &dims, Type t);
    bool empty() const { return total() == 0; }
    int dims() const { return (int)size_.size(); }
    int size(int i) const { return size_.at(i); }
    int rows() const { return size_.empty() ? 0 : size_[0]; }
    int cols() const { return size_.size() < 2 ? 0 : size_[1]; }
    size_t total() const;
    Type type() const { return type_; }
    uint8_t *data() { return bytes_.data(); }
    const uint8_t *data() const { return bytes_.data(); }
    template <typename T> T *ptr() { return reinterpret_cast<T *>(bytes_.data()); }
    template <typename T> const T *ptr() const { return reinterpret_cast<const T *>(bytes_.data()); }

private:
    std::vector<int> size_;
    Type type_ = U8;
    std::vector<uint8_t> bytes_;
};

// Basic labeling interface (LabelingTools.hpp:58-73).
class ILabeling {
public:
    virtual ~ILabeling() = default;
    virtual TTime Label(const TImage &pixels, TImage &labels, char threads = MAX_THREADS,
                        TCoherence coh = COH_DEFAULT) = 0;
    // v > 8 -> 255, else 0: what the reference's cv::threshold call computes
    // (LabelingTools.cpp:39-49, arguments THRESH_OTSU = 8 as threshold, type 0).
    static TImage RGB2Gray(const TImage &img);
};

// GPU labeler bound to one device context (IOCLLabeling, LabelingTools.hpp:100-137).
class IOCLLabeling : public ILabeling {
public:
    ~IOCLLabeling() override;
    TTime Label(const TImage &pixels, TImage &labels, char threads = MAX_THREADS,
                TCoherence coh = COH_DEFAULT) override;
    bool Initialized() const { return ctx_ != nullptr; }
    lbl_ctx *Context() const { return ctx_; }

protected:
    IOCLLabeling(lbl_alg alg, bool runOnGPU);
    lbl_alg alg_;
    lbl_ctx *ctx_ = nullptr;

private:
    IOCLLabeling(const IOCLLabeling &) = delete;
    IOCLLabeling &operator=(const IOCLLabeling &) = delete;
};

// 3D variant (IOCLLabeling3D, LabelingTools.hpp:143-177): only the default
// coherence (26-connectivity); labels are returned cropped to the input size.
class IOCLLabeling3D : public IOCLLabeling {
public:
    TTime Label(const TImage &pixels, TImage &labels, char threads = MAX_THREADS,
                TCoherence coh = COH_DEFAULT) override;

protected:
    IOCLLabeling3D(lbl_alg alg, bool runOnGPU) : IOCLLabeling(alg, runOnGPU) {}
};

// The GPU classes of the reference's registry (LabelingAlgs.hpp:196-390).
class TOCLBinLabeling final : public IOCLLabeling {
public:
    explicit TOCLBinLabeling(bool runOnGPU = true) : IOCLLabeling(LBL_BIN, runOnGPU) {}
};
class TOCLLabelDistribution final : public IOCLLabeling {
public:
    explicit TOCLLabelDistribution(bool runOnGPU = true) : IOCLLabeling(LBL_LBEQ, runOnGPU) {}
};
class TOCLLabelEquivalenceX2 final : public IOCLLabeling {
public:
    explicit TOCLLabelEquivalenceX2(bool runOnGPU = true) : IOCLLabeling(LBL_BLEQ, runOnGPU) {}
};
class TOCLRunEquivLabeling final : public IOCLLabeling {
public:
    explicit TOCLRunEquivLabeling(bool runOnGPU = true) : IOCLLabeling(LBL_RUNEQ, runOnGPU) {}
};
class TOCLBinLabeling3D final : public IOCLLabeling3D {
public:
    explicit TOCLBinLabeling3D(bool runOnGPU = true) : IOCLLabeling3D(LBL_BIN, runOnGPU) {}
};
class TOCLLabelEquivalence3D final : public IOCLLabeling3D {
public:
    explicit TOCLLabelEquivalence3D(bool runOnGPU = true) : IOCLLabeling3D(LBL_LBEQ, runOnGPU) {}
};
class TOCLBlockEquivalence3D final : public IOCLLabeling3D {
public:
    explicit TOCLBlockEquivalence3D(bool runOnGPU = true) : IOCLLabeling3D(LBL_BLEQ, runOnGPU) {}
};

// The reference's CPU classes (LabelingAlgs.hpp:29-190), computed on the GPU
// with the label values of the CPU methods (include/labeling_amd.h describes
// each method's values); `threads` is ignored (it sets OpenMP threads in the
// reference).  They are what SetLabelingAlg returns without -g / -u.
class TBinLabeling final : public IOCLLabeling {
public:
    TBinLabeling() : IOCLLabeling(LBL_BIN, true) {}
};
class TOpenCVLabeling final : public IOCLLabeling {
public:
    TOpenCVLabeling() : IOCLLabeling(LBL_OCV, true) {}
};
class TBlockGranaLabeling final : public IOCLLabeling {
public:
    TBlockGranaLabeling() : IOCLLabeling(LBL_GRBLOCK, true) {}
};
class TRunLabeling final : public IOCLLabeling {
public:
    TRunLabeling() : IOCLLabeling(LBL_HERUN, true) {}
};
class TLabelDistribution final : public IOCLLabeling {
public:
    TLabelDistribution() : IOCLLabeling(LBL_LBEQ, true) {}
};
class TLabelEquivalenceX2 final : public IOCLLabeling {
public:
    TLabelEquivalenceX2() : IOCLLabeling(LBL_BLEQ, true) {}
};
class TRunEqivLabeling final : public IOCLLabeling {
public:
    TRunEqivLabeling() : IOCLLabeling(LBL_RUNEQ, true) {}
};

// Registry (main.cpp:88-131): every entry of ALG_LIST with the reference's
// factories; a factory the reference leaves empty is empty here too.
struct Algs {
    std::string descr;
    std::function<std::shared_ptr<ILabeling>()> cpu;
    std::function<std::shared_ptr<IOCLLabeling>(bool)> ocl;
    std::function<std::shared_ptr<ILabeling>()> cpu3d;
    std::function<std::shared_ptr<IOCLLabeling3D>(bool)> ocl3d;
};

const std::map<std::string, Algs> &AlgList();   // ALG_LIST after InitAlgs

enum class Device { CPU, OCL_CPU, OCL_GPU };    // Options::useOCL, main.cpp:38

// SetLabelingAlg (main.cpp:373-395): nullptr when the algorithm does not
// support the requested capabilities; throws for an unknown name.
std::shared_ptr<ILabeling> SetLabelingAlg(const std::string &algName, Device device, bool label3D);

}  // namespace LabelingTools
