// imageio_fuzz_check.cpp -- the image file readers (labeling_amd/csrc/host/
// imageio.cpp) built with AddressSanitizer and UBSan and fed every file named
// on the command line: valid files of each format and damaged copies of them
// (truncated, bytes overwritten).  A reader may reject a file (it throws
// std::runtime_error); it must never read or write out of bounds, and a file
// it accepts must give an image whose buffer matches its size.
#include <cstdio>
#include <exception>
#include <string>

#include "labeling_amd/csrc/host/imageio.hpp"

namespace io = LabelingTools::io;

int main(int argc, char **argv)
{
    int accepted = 0, rejected = 0, bad = 0;
    for (int i = 1; i < argc; ++i) {
        for (int gray = 0; gray <= 1; ++gray) {
            try {
                const io::Image8 im = gray ? io::imread_gray(argv[i]) : io::imread_color(argv[i]);
                if (im.w <= 0 || im.h <= 0 || (im.c != 1 && im.c != 3) ||
                    im.px.size() != (size_t)im.w * im.h * im.c) {
                    std::printf("BAD SHAPE %s: %d x %d x %d, %zu bytes\n", argv[i], im.h, im.w, im.c, im.px.size());
                    ++bad;
                }
                ++accepted;
            } catch (const std::exception &) {
                ++rejected;
            }
        }
    }
    std::printf("accepted %d rejected %d bad %d\n", accepted, rejected, bad);
    std::printf("%s\n", bad ? "FAILED" : "ok");
    return bad ? 1 : 0;
}
