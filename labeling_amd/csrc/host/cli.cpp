// cli.cpp -- `labeling` command line, mirroring the reference's main.cpp flags
// (-i -o -a -3 -g -u -j -l -c -h; main.cpp:344-369,399-447) for the GPU
// methods "bleq" and "lbeq".
//
// Image files are binary PGM (P5, 8-bit) instead of OpenCV-decoded formats; a
// 3D image is a directory of equally sized PGM slices in sorted file-name order
// (Read3DImage, main.cpp:172-212), stored as {rows, cols, slices} with the slice
// index fastest like the reference.  Binarisation is the reference's v > 8
// (LabelingTools.cpp:46).  Output (-o, an existing directory): colourised PPM
// images (LabelsToRGB / Write3DLabels, main.cpp:250-305).  Timing lines follow
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

// Binary PGM (P5, maxval <= 255).
TImage read_pgm(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string magic;
    f >> magic;
    auto next_int = [&]() {
        int v;
        while (f >> std::ws && f.peek() == '#') f.ignore(1 << 20, '\n');
        f >> v;
        return v;
    };
    if (magic != "P5") throw std::runtime_error(path + ": not a binary PGM (P5) file");
    int w = next_int(), h = next_int(), maxv = next_int();
    if (w <= 0 || h <= 0 || maxv <= 0 || maxv > 255) throw std::runtime_error(path + ": unsupported PGM header");
    f.get();
    TImage img(h, w, TImage::U8);
    f.read(reinterpret_cast<char *>(img.data()), (std::streamsize)img.total());
    if (!f) throw std::runtime_error(path + ": truncated PGM data");
    return img;
}

std::vector<std::string> find_files(const std::string &dir)
{
    std::vector<std::string> out;
    if (DIR *d = opendir(dir.c_str())) {
        while (dirent *e = readdir(d)) {
            std::string n = e->d_name;
            if (n.size() > 4 && n.substr(n.size() - 4) == ".pgm" && is_file(dir + "/" + n)) out.push_back(dir + "/" + n);
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
        TImage s = read_pgm(files[k]);
        if (vol.empty()) vol = TImage(s.rows(), s.cols(), (int)files.size(), TImage::U8);
        if (s.rows() != vol.size(0) || s.cols() != vol.size(1))
            throw std::runtime_error("Cannot read 3D image: slice sizes do not match");
        for (int i = 0; i < s.rows(); ++i)
            for (int j = 0; j < s.cols(); ++j)
                vol.data()[((size_t)i * vol.size(1) + j) * vol.size(2) + k] = s.data()[(size_t)i * s.cols() + j];
    }
    return vol;
}

void color_of(uint32_t l, uint8_t rgb[3])
{
    uint32_t h = l * 2654435761u;
    rgb[0] = l ? (uint8_t)(h >> 24) : 0;
    rgb[1] = l ? (uint8_t)(h >> 16) : 0;
    rgb[2] = l ? (uint8_t)(h >> 8) : 0;
}

void write_ppm(const std::string &path, const uint32_t *lab, int rows, int cols, size_t stride, size_t step)
{
    std::ofstream f(path, std::ios::binary);
    f << "P6\n" << cols << " " << rows << "\n255\n";
    std::vector<uint8_t> row((size_t)cols * 3);
    for (int i = 0; i < rows; ++i) {
        for (int j = 0; j < cols; ++j) color_of(lab[(size_t)i * stride + (size_t)j * step], &row[(size_t)j * 3]);
        f.write(reinterpret_cast<const char *>(row.data()), (std::streamsize)row.size());
    }
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
                 "  -i <input_path> : Input file (binary PGM) or path\n"
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
                write_ppm(o.outPath + "/" + std::to_string(k + 1) + ".ppm", labels.ptr<uint32_t>() + k, labels.size(0),
                          labels.size(1), (size_t)labels.size(1) * labels.size(2), (size_t)labels.size(2));
        return 0;
    }
    std::vector<std::string> files;
    if (is_dir(o.inPath)) files = find_files(o.inPath);
    else if (is_file(o.inPath)) files.push_back(o.inPath);
    else { PrintHelp(); throw std::runtime_error("Wrong input path"); }
    ImgTime all;
    size_t count = 0;
    for (const auto &f : files) {
        TImage img = ILabeling::RGB2Gray(read_pgm(f)), labels;
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
            write_ppm(o.outPath + "/" + base_name(f) + ".ppm", labels.ptr<uint32_t>(), labels.rows(), labels.cols(),
                      (size_t)labels.cols(), 1);
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
