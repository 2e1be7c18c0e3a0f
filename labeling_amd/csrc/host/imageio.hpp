// imageio.hpp -- image files for the `labeling` command line, in the formats
// the reference's driver reads and writes through OpenCV (main.cpp:134,184,322,
// 477: cv::imread / cv::imwrite on .jpg .jpeg .bmp .png .tif .tiff), plus the
// PNM formats.  Host code only; no OpenCV.
//
// Decoding follows cv::imread's conversions:
//   * color reads (cv::imread(path), the 2D driver, main.cpp:322,477) return
//     8-bit BGR: gray replicated, palettes expanded, alpha dropped, 16-bit
//     samples reduced to their high byte, 1/2/4-bit gray scaled to 0..255;
//   * gray reads (cv::imread(path, IMREAD_GRAYSCALE), the 3D slice reader,
//     main.cpp:184) return 8-bit gray: color converted with the BT.601 weights
//     in 14-bit fixed point (JPEG: its luma channel as decoded).
// Writing picks the format from the file extension; label images are 8-bit BGR.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace LabelingTools {
namespace io {

struct Image8 {
    int w = 0, h = 0, c = 0;      // c: 1 (gray) or 3 (BGR)
    std::vector<uint8_t> px;      // h rows of w * c bytes
};

// Throw std::runtime_error on unreadable or unsupported files.
Image8 imread_color(const std::string &path);
Image8 imread_gray(const std::string &path);
void imwrite(const std::string &path, const Image8 &img);

// The extensions the reference's directory scan accepts (main.cpp:134), and the PNM ones.
bool is_image_name(const std::string &path);

}  // namespace io
}  // namespace LabelingTools

// C entry points of the same functions (for tests through ctypes).
// lt_image_read: gray = 0 for a color read (BGR), 1 for a gray read; fills
// *h, *w, *c and, when out is not NULL and cap is large enough, the pixels.
// Returns 0 on success, -1 on failure (message: lt_image_error()).
extern "C" int lt_image_read(const char *path, int gray, int *h, int *w, int *c, uint8_t *out, size_t cap);
extern "C" int lt_image_write(const char *path, const uint8_t *px, int h, int w, int c);
extern "C" const char *lt_image_error(void);
