// Exercises the C++ mirror of the reference's registry and labeler classes
// (labeling_amd/csrc/host/LabelingTools.hpp) the way main.cpp uses them.
// Usage: registry_check <in.raw> <h> <w> <out.raw> [3d-depth [alg [cpu]]]
// Writes the labels of the input (uint8 raw) as raw little-endian uint32 and
// prints "ok <time_us>"; checks the reference's error behaviour first.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>

#include "LabelingTools.hpp"

using namespace LabelingTools;

static int expect_throw(const char *what, void (*fn)())
{
    try {
        fn();
    } catch (const std::exception &e) {
        std::cout << "threw (" << what << "): " << e.what() << "\n";
        return 0;
    }
    std::cout << "did not throw: " << what << "\n";
    return 1;
}

int main(int argc, char **argv)
{
    if (argc < 5) return 2;
    const int h = std::atoi(argv[2]), w = std::atoi(argv[3]);
    const int d = argc > 5 ? std::atoi(argv[5]) : 0;
    int bad = 0;
    // SetLabelingAlg semantics (main.cpp:373-395)
    // the factories the reference's InitAlgs leaves empty (main.cpp:103-130) are empty here
    if (SetLabelingAlg("he-run", Device::OCL_GPU, false) != nullptr) { std::cout << "ocl he-run should be absent\n"; bad++; }
    if (SetLabelingAlg("runeq", Device::OCL_GPU, true) != nullptr) { std::cout << "ocl3d runeq should be absent\n"; bad++; }
    if (SetLabelingAlg("ocv", Device::OCL_GPU, false) != nullptr) { std::cout << "ocl ocv should be absent\n"; bad++; }
    if (SetLabelingAlg("lbeq", Device::OCL_CPU, true) != nullptr) { std::cout << "ocl-cpu should be absent\n"; bad++; }
    if (AlgList().size() != 7) { std::cout << "ALG_LIST has 7 entries\n"; bad++; }
    bad += expect_throw("gr-block with 4-connectivity", [] {
        auto a = SetLabelingAlg("gr-block", Device::CPU, false);
        TImage img(4, 4, TImage::U8), lab;
        a->Label(img, lab, MAX_THREADS, COH_4);
    });
    bad += expect_throw("unknown algorithm", [] { SetLabelingAlg("nope", Device::OCL_GPU, false); });
    bad += expect_throw("bleq with 4-connectivity", [] {
        auto a = SetLabelingAlg("bleq", Device::OCL_GPU, false);
        TImage img(4, 4, TImage::U8), lab;
        a->Label(img, lab, MAX_THREADS, COH_4);
    });
    bad += expect_throw("3D with COH_8", [] {
        auto a = SetLabelingAlg("bleq", Device::OCL_GPU, true);
        TImage img(4, 4, 4, TImage::U8), lab;
        a->Label(img, lab, MAX_THREADS, COH_8);
    });
    bad += expect_throw("empty image", [] {
        auto a = SetLabelingAlg("lbeq", Device::OCL_GPU, false);
        TImage img, lab;
        a->Label(img, lab);
    });

    TImage img = d ? TImage(d, h, w, TImage::U8) : TImage(h, w, TImage::U8), labels;
    std::ifstream f(argv[1], std::ios::binary);
    f.read(reinterpret_cast<char *>(img.data()), (std::streamsize)img.total());
    if (!f) { std::cout << "short input\n"; return 2; }
    const bool cpu = argc > 7 && !std::strcmp(argv[7], "cpu");
    auto alg = SetLabelingAlg(argc > 6 ? argv[6] : "bleq", cpu ? Device::CPU : Device::OCL_GPU, d != 0);
    if (!alg) { std::cout << "no such algorithm for this device\n"; return 1; }
    TTime us = alg->Label(ILabeling::RGB2Gray(img), labels);
    std::ofstream o(argv[4], std::ios::binary);
    o.write(reinterpret_cast<const char *>(labels.data()), (std::streamsize)(labels.total() * 4));
    std::cout << (bad ? "FAILED " : "ok ") << us << "\n";
    return bad ? 1 : 0;
}
