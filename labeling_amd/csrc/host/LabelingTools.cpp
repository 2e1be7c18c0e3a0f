// LabelingTools.cpp -- the C++ mirror of the reference's labeling classes over
// the C ABI (see LabelingTools.hpp for the reference correspondence).
#include "LabelingTools.hpp"

namespace LabelingTools {

namespace {

void throw_status(lbl_ctx *ctx, lbl_status st, const char *where)
{
    if (st == LBL_OK) return;
    throw std::runtime_error(std::string(where) + " : " + lbl_status_string(st) + " (" + lbl_last_error(ctx) + ")");
}

lbl_status to_conn(TCoherence coh, int *conn)
{
    *conn = coh == COH_4 ? 4 : (coh == COH_8 ? 8 : 0);
    return LBL_OK;
}

}  // namespace

void TImage::create(const std::vector<int> &dims, Type t)
{
    for (int d : dims)
        if (d < 0) throw std::invalid_argument("TImage: negative size");
    size_ = dims;
    type_ = t;
    bytes_.assign(total() * (t == U8 ? 1 : 4), 0);
}

size_t TImage::total() const
{
    if (size_.empty()) return 0;
    size_t n = 1;
    for (int d : size_) n *= (size_t)d;
    return n;
}

TImage ILabeling::RGB2Gray(const TImage &img)
{
    if (img.type() != TImage::U8) throw std::invalid_argument("ILabeling::RGB2Gray : input must be uint8");
    TImage out;
    std::vector<int> dims;
    for (int i = 0; i < img.dims(); ++i) dims.push_back(img.size(i));
    out.create(dims, TImage::U8);
    const uint8_t *s = img.data();
    uint8_t *d = out.data();
    for (size_t i = 0; i < img.total(); ++i) d[i] = s[i] > 8 ? 255 : 0;
    return out;
}

IOCLLabeling::IOCLLabeling(lbl_alg alg, bool runOnGPU) : alg_(alg)
{
    if (!runOnGPU) throw std::runtime_error("IOCLLabeling::Init : only GPU devices are supported");
    throw_status(nullptr, lbl_create(&ctx_, 0), "IOCLLabeling::Init");
}

IOCLLabeling::~IOCLLabeling()
{
    lbl_destroy(ctx_);
}

TTime IOCLLabeling::Label(const TImage &pixels, TImage &labels, char, TCoherence coh)
{
    if (!ctx_) throw std::runtime_error("IOCLLabeling::Label : OpenCL device is not initialized");
    if (pixels.empty()) throw std::runtime_error("IOCLLabeling::Label : Input image is empty");
    if (pixels.dims() != 2 || pixels.type() != TImage::U8)
        throw std::runtime_error("IOCLLabeling::Label : Input image must be a 2D uint8 image");
    int conn = 0;
    to_conn(coh, &conn);
    labels = TImage(pixels.rows(), pixels.cols(), TImage::U32);
    float us = 0.f;
    throw_status(ctx_, lbl_label2d(ctx_, alg_, pixels.data(), (uint32_t)pixels.rows(), (uint32_t)pixels.cols(),
                                   (size_t)pixels.cols(), conn, labels.ptr<uint32_t>(), nullptr, &us),
                 "IOCLLabeling::Label");
    return (TTime)us;
}

TTime IOCLLabeling3D::Label(const TImage &pixels, TImage &labels, char, TCoherence coh)
{
    if (!ctx_) throw std::runtime_error("IOCLLabeling3D::Label : OpenCL device is not initialized");
    if (pixels.empty()) throw std::runtime_error("IOCLLabeling3D::Label : Input image is empty");
    if (pixels.dims() != 3) throw std::runtime_error("IOCLLabeling3D::Label : Input image is not a 3D image");
    if (coh != COH_DEFAULT)
        throw std::runtime_error("IOCLLabeling3D::Label : Only default coherence is supported for 3D labeling");
    labels = TImage(pixels.size(0), pixels.size(1), pixels.size(2), TImage::U32);
    float us = 0.f;
    throw_status(ctx_, lbl_label3d(ctx_, alg_, pixels.data(), (uint32_t)pixels.size(0), (uint32_t)pixels.size(1),
                                   (uint32_t)pixels.size(2), 0, labels.ptr<uint32_t>(), nullptr, &us),
                 "IOCLLabeling3D::Label");
    return (TTime)us;
}

const std::map<std::string, Algs> &AlgList()
{
    // InitAlgs (main.cpp:99-131): names, descriptions and which factories exist
    static const std::map<std::string, Algs> list = [] {
        std::map<std::string, Algs> m;
        m.emplace("bin", Algs{"Simple binarization", [] { return std::make_shared<TBinLabeling>(); },
                              [](bool gpu) { return std::make_shared<TOCLBinLabeling>(gpu); }, nullptr,
                              [](bool gpu) { return std::make_shared<TOCLBinLabeling3D>(gpu); }});
        m.emplace("he-run", Algs{"Run labeling by He et.al. 2008", [] { return std::make_shared<TRunLabeling>(); },
                                 nullptr, nullptr, nullptr});
        m.emplace("gr-block", Algs{"Block labeling by Grana et.al. 2010",
                                   [] { return std::make_shared<TBlockGranaLabeling>(); }, nullptr, nullptr, nullptr});
        m.emplace("ocv", Algs{"OpenCV labeling", [] { return std::make_shared<TOpenCVLabeling>(); }, nullptr, nullptr,
                              nullptr});
        m.emplace("lbeq", Algs{"Label equivalence by Kalentev et.al. 2012",
                               [] { return std::make_shared<TLabelDistribution>(); },
                               [](bool gpu) { return std::make_shared<TOCLLabelDistribution>(gpu); }, nullptr,
                               [](bool gpu) { return std::make_shared<TOCLLabelEquivalence3D>(gpu); }});
        m.emplace("bleq", Algs{"Block equivalence by Zavalishin et.al. 2016",
                               [] { return std::make_shared<TLabelEquivalenceX2>(); },
                               [](bool gpu) { return std::make_shared<TOCLLabelEquivalenceX2>(gpu); }, nullptr,
                               [](bool gpu) { return std::make_shared<TOCLBlockEquivalence3D>(gpu); }});
        m.emplace("runeq", Algs{"Run equivalence by Bekhtin et.al. 2015",
                                [] { return std::make_shared<TRunEqivLabeling>(); },
                                [](bool gpu) { return std::make_shared<TOCLRunEquivLabeling>(gpu); }, nullptr, nullptr});
        return m;
    }();
    return list;
}

std::shared_ptr<ILabeling> SetLabelingAlg(const std::string &algName, Device device, bool label3D)
{
    const auto &list = AlgList();
    auto it = list.find(algName);
    if (it == list.end()) throw std::runtime_error("No labeling algorithm specified");
    const bool ocl = device != Device::CPU, gpu = device == Device::OCL_GPU;
    const Algs &a = it->second;
    if (!label3D) {
        if (!ocl) return a.cpu ? a.cpu() : nullptr;
        return a.ocl && gpu ? a.ocl(gpu) : nullptr;
    }
    if (!ocl) return a.cpu ? a.cpu() : nullptr;
    return a.ocl3d && gpu ? a.ocl3d(gpu) : nullptr;
}

}  // namespace LabelingTools
