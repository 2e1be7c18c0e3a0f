// cli.cpp -- `labeling` command line, mirroring the reference's main.cpp flags
// (-i -o -a -3 -g -u -j -l -c -h; main.cpp:344-369,399-447) for every method
// of its registry.
//
// Image files are read and written as the reference's driver does through
// OpenCV (imageio.hpp): a 2D input is read in color (cv::imread,
// main.cpp:322,477) and turned gray by ILabeling::RGB2Gray's cv::cvtColor
// (COLOR_RGB2GRAY applied to OpenCV's BGR order, LabelingTools.cpp:43-44);
// a 3D input is a directory of equally sized slices read as gray
// (IMREAD_GRAYSCALE, main.cpp:172-212) in sorted file-name order, stored as
// {rows, cols, slices} with the slice index fastest like the reference.
// Binarisation is the reference's v > 8 (LabelingTools.cpp:46).  Output (-o, an
// existing directory): colourised labels under the input's file name, in its
// format (main.cpp:329,486), and for 3D one PNG per slice, 1.png, 2.png, ...
// (Write3DLabels, main.cpp:290-305).  Timing lines follow
// main.cpp:334-339,451-462 (per-image ms, min/avg/max over -l cycles).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <dirent.h>
#include <fstream>
#include <iostream>
#include <sstream>
#include <sys/stat.h>

#include "LabelingTools.hpp"
#include "imageio.hpp"

using namespace LabelingTools;

namespace {

struct Options {   // main.cpp:28-42
    std::string inPath, outPath, algName;
    int numThreads = MAX_THREADS;
    int cycles = 1;
    Device device = Device::CPU;
    TCoherence coh = COH_DEFAULT;
    bool label3D = false;
    bool quickExit = false;
};

bool is_dir(const std::string &p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool is_file(const std::string &p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string base_name(const std::string &p)
{
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}

// A 2D image as the reference's driver hands it to Label(): cv::imread's BGR
// pixels through cvtColor(COLOR_RGB2GRAY), whose weights 0.299 / 0.587 / 0.114
// (14-bit fixed point) then fall on B, G, R (LabelingTools.cpp:43-44).
TImage read_2d(const std::string &path)
{
    const io::Image8 c = io::imread_color(path);
    TImage img(c.h, c.w, TImage::U8);
    uint8_t *d = img.data();
    for (size_t i = 0; i < (size_t)c.w * c.h; ++i) {
        const uint8_t *s = &c.px[i * 3];
        d[i] = (uint8_t)((s[0] * 4899 + s[1] * 9617 + s[2] * 1868 + 8192) >> 14);
    }
    return img;
}

std::vector<std::string> find_files(const std::string &dir)
{
    std::vector<std::string> out;
    if (DIR *d = opendir(dir.c_str())) {
        while (dirent *e = readdir(d)) {
            std::string n = e->d_name;
            if (io::is_image_name(n) && is_file(dir + "/" + n)) out.push_back(dir + "/" + n);   // main.cpp:134-152
        }
        closedir(d);
    }
    std::sort(out.begin(), out.end());
    return out;
}

TImage read_3d(const std::string &dir)
{
    const auto files = find_files(dir);
    if (files.empty()) throw std::runtime_error("Wrong input path (assumed directory with 3D image slices)");
    TImage vol;
    for (size_t k = 0; k < files.size(); ++k) {
        const io::Image8 s = io::imread_gray(files[k]);
        if (vol.empty()) vol = TImage(s.h, s.w, (int)files.size(), TImage::U8);
        if (s.h != vol.size(0) || s.w != vol.size(1))
            throw std::runtime_error("Cannot read 3D image: slice sizes do not match");
        for (int i = 0; i < s.h; ++i)
            for (int j = 0; j < s.w; ++j)
                vol.data()[((size_t)i * vol.size(1) + j) * vol.size(2) + k] = s.px[(size_t)i * s.w + j];
    }
    return vol;
}

void color_of(uint32_t l, uint8_t rgb[3])
{
    uint32_t h = l * 2654435761u;
    rgb[0] = l ? (uint8_t)(h >> 24) : 0;
    rgb[1] = l ? (uint8_t)(h >> 16) : 0;
    rgb[2] = l ? (uint8_t)(h >> 8) : 0;
    if (l && !(rgb[0] | rgb[1] | rgb[2])) rgb[0] = 1;   // black is the background's
}

// LabelsToRGB (main.cpp:250-286): a colour per label, black background; the
// reference draws colours with std::rand, here a hash of the label gives them.
io::Image8 labels_to_bgr(const uint32_t *lab, int rows, int cols, size_t stride, size_t step)
{
    io::Image8 o;
    o.w = cols;
    o.h = rows;
    o.c = 3;
    o.px.resize((size_t)rows * cols * 3);
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) color_of(lab[(size_t)i * stride + (size_t)j * step], &o.px[((size_t)i * cols + j) * 3]);
    return o;
}

struct ImgTime {   // main.cpp:46-84
    TTime mn = UINT32_MAX, mx = 0;
    double sum = 0;
    size_t c = 0;
    void Add(TTime t) { mn = std::min(mn, t); mx = std::max(mx, t); sum += t; ++c; }
    double Avg() const { return c ? sum / c : 0; }
    TTime Min() const { return c ? mn : 0; }
};

void PrintHelp()
{
    std::cout << "Usage: labeling [options]\n\nOptions:\n"
                 "  -i <input_path> : Input file or path (.jpg .jpeg .bmp .png .tif .tiff, PNM)\n"
                 "  -o <out_path>   : Output path (existing directory)\n"
                 "  -a <algorithm>  : Labeling algorithm:\n";
    for (const auto &a : AlgList())
    {
        std::string caps;
        if (a.second.cpu) caps += " CPU";
        if (a.second.ocl) caps += " GPU";
        if (a.second.ocl3d) caps += " GPU-3D";
        std::cout << "            " << a.first << "     : " << a.second.descr << " [" << caps.substr(1) << "]\n";
    }
    std::cout << "  -3           : Treat input sequence as a single 3D image\n"
                 "  -g           : Run algorithm on GPU\n"
                 "  -u           : Run algorithm in OpenCL mode on CPU (not available)\n"
                 "  -j <threads> : Set number of parallel threads for OpenMP (default 0)\n"
                 "  -l <cycles>  : Set number of cycles for each image (default 1)\n"
                 "  -c <connect> : Set connectivity (4 or 8 [default])\n"
                 "  -h           : Print this help\n\n";
}

Options ParseInput(int argc, char **argv)   // main.cpp:399-447
{
    Options o;
    auto data = [&](int &i) -> std::string {
        if (i + 1 < argc) return argv[++i];
        throw std::runtime_error(std::string("Wrong input parameters (no data following the key ") + argv[i] + ")");
    };
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "-i")) o.inPath = data(i);
        else if (!strcmp(argv[i], "-o")) o.outPath = data(i);
        else if (!strcmp(argv[i], "-a")) o.algName = data(i);
        else if (!strcmp(argv[i], "-3")) o.label3D = true;
        else if (!strcmp(argv[i], "-g")) o.device = Device::OCL_GPU;
        else if (!strcmp(argv[i], "-u")) o.device = Device::OCL_CPU;
        else if (!strcmp(argv[i], "-j")) o.numThreads = std::stoi(data(i));
        else if (!strcmp(argv[i], "-l")) o.cycles = std::stoi(data(i));
        else if (!strcmp(argv[i], "-c")) o.coh = data(i) == "4" ? COH_4 : COH_8;
        else if (!strcmp(argv[i], "-h")) { PrintHelp(); o.quickExit = true; return o; }
        else throw std::runtime_error(std::string("Wrong input parameters (unknown key ") + argv[i] + ")");
    }
    return o;
}

void PrintTime(const std::string &name, const ImgTime &t, const Options &o)   // main.cpp:451-462
{
    std::cout << "Image: " << name << "\n";
    if (o.cycles == 1) std::cout << "Processing time: " << t.Avg() / 1000 << " ms\n";
    else
        std::cout << "Processing time:\n Min = " << t.Min() / 1000.0 << "ms\n Avg = " << t.Avg() / 1000
                  << "ms\n Max = " << t.mx / 1000.0 << "ms\n";
}

int Run(const Options &o)
{
    if (o.quickExit) return 0;
    auto alg = SetLabelingAlg(o.algName, o.device, o.label3D);
    if (!alg) throw std::runtime_error("Chosen algorithm doesn't support specified capabilities");
    const bool want_out = !o.outPath.empty() && is_dir(o.outPath);
    if (o.label3D) {
        if (!is_dir(o.inPath)) { PrintHelp(); throw std::runtime_error("Wrong input path (assumed directory with 3D image slices)"); }
        TImage vol = ILabeling::RGB2Gray(read_3d(o.inPath)), labels;
        ImgTime t;
        for (int c = 0; c < o.cycles; ++c) t.Add(alg->Label(vol, labels, (char)o.numThreads, o.coh));
        PrintTime(o.inPath, t, o);
        if (want_out)
            for (int k = 0; k < labels.size(2); ++k)
                io::imwrite(o.outPath + "/" + std::to_string(k + 1) + ".png",
                            labels_to_bgr(labels.ptr<uint32_t>() + k, labels.size(0), labels.size(1),
                                          (size_t)labels.size(1) * labels.size(2), (size_t)labels.size(2)));
        return 0;
    }
    std::vector<std::string> files;
    if (is_dir(o.inPath)) files = find_files(o.inPath);
    else if (is_file(o.inPath)) files.push_back(o.inPath);
    else { PrintHelp(); throw std::runtime_error("Wrong input path"); }
    ImgTime all;
    size_t count = 0;
    for (const auto &f : files) {
        TImage img = ILabeling::RGB2Gray(read_2d(f)), labels;
        ImgTime t;
        for (int c = 0; c < o.cycles; ++c) {
            TTime us = alg->Label(img, labels, (char)o.numThreads, o.coh);
            t.Add(us);
            all.Add(us);
        }
        if (files.size() == 1) PrintTime(base_name(f), t, o);
        else std::cout << "Processing image " << ++count << "/" << files.size() << " (" << base_name(f) << ") "
                       << t.Avg() / 1000 << " ms\n";
        if (want_out)
            io::imwrite(o.outPath + "/" + base_name(f),
                        labels_to_bgr(labels.ptr<uint32_t>(), labels.rows(), labels.cols(), (size_t)labels.cols(), 1));
    }
    if (files.size() > 1)
        std::cout << "\nMin processing time: " << all.Min() / 1000.0 << " ms\nAvg processing time: "
                  << all.Avg() / 1000 << " ms\nMax processing time: " << all.mx / 1000.0 << " ms\n";
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        return Run(ParseInput(argc, argv));
    } catch (const std::exception &e) {
        // the reference prints the message and still exits 0 (main.cpp:530-547);
        // a nonzero status is returned here so scripts can see the failure
        std::cout << "Error: " << e.what() << "\n\n";
        return 1;
    }
}
