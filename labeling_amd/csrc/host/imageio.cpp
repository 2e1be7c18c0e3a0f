// imageio.cpp -- PNG, BMP, TIFF, JPEG and PNM files for the `labeling` command
// line (see imageio.hpp for the conversions, which follow cv::imread /
// cv::imwrite as the reference's driver calls them).  Written from the file
// format specifications: PNG (ISO/IEC 15948, zlib for its deflate streams),
// BMP (the Windows DIB headers), TIFF 6.0 baseline with LZW / PackBits /
// Deflate strips, JPEG baseline and extended sequential Huffman coding (ITU-T
// T.81, with the JFIF YCbCr conversion and the triangle-filter chroma
// upsampling of the IJG decoder), and Netpbm.
#include "imageio.hpp"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace LabelingTools {
namespace io {
namespace {

[[noreturn]] void bad(const std::string &what) { throw std::runtime_error(what); }

std::vector<uint8_t> read_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) bad("cannot open " + path);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

std::string lower_ext(const std::string &path)
{
    const size_t dot = path.find_last_of('.'), slash = path.find_last_of('/');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    std::string e = path.substr(dot);
    for (char &ch : e) ch = (char)std::tolower((unsigned char)ch);
    return e;
}

// Decoded samples before imread's conversions: c = 1 gray, 2 gray + alpha,
// 3 RGB, 4 RGBA (8 bits each); luma: a JPEG's decoded luma plane (gray reads
// take it instead of converting RGB).
struct Raw {
    int w = 0, h = 0, c = 0;
    std::vector<uint8_t> v;
    std::vector<uint8_t> luma;
    bool png = false;   // gray reads of a color PNG use libpng's 15-bit weights
};

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
uint16_t be16(const uint8_t *p) { return (uint16_t)(p[0] << 8 | p[1]); }
uint32_t le32(const uint8_t *p) { return (uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0]; }
uint16_t le16(const uint8_t *p) { return (uint16_t)(p[1] << 8 | p[0]); }

void check_size(int64_t w, int64_t h)
{
    if (w <= 0 || h <= 0 || w > (1 << 20) || h > (1 << 20) || w * h > (int64_t(1) << 31)) bad("unsupported image size");
}

// zlib inflate of a complete stream into exactly `out_size` bytes (or fewer: error)
std::vector<uint8_t> inflate_all(const uint8_t *p, size_t n, size_t out_size)
{
    std::vector<uint8_t> out(out_size);
    z_stream z{};
    if (inflateInit(&z) != Z_OK) bad("zlib init failed");
    z.next_in = const_cast<Bytef *>(p);
    z.avail_in = (uInt)n;
    z.next_out = out.data();
    z.avail_out = (uInt)out.size();
    const int rc = inflate(&z, Z_FINISH);
    const size_t got = out.size() - z.avail_out;
    inflateEnd(&z);
    if ((rc != Z_STREAM_END && rc != Z_OK && rc != Z_BUF_ERROR) || got < out_size) bad("corrupt deflate stream");
    return out;
}

// ---- PNG -------------------------------------------------------------------------

uint8_t paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (uint8_t)(pa <= pb && pa <= pc ? a : (pb <= pc ? b : c));
}

void unfilter(uint8_t *rows, int nrows, size_t rowbytes, int bpp)
{
    std::vector<uint8_t> zero(rowbytes, 0);
    const uint8_t *prev = zero.data();
    for (int y = 0; y < nrows; ++y) {
        uint8_t *r = rows + y * (rowbytes + 1);
        const int f = r[0];
        uint8_t *x = r + 1;
        for (size_t i = 0; i < rowbytes; ++i) {
            const int a = i >= (size_t)bpp ? x[i - bpp] : 0, b = prev[i], c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            switch (f) {
            case 0: break;
            case 1: x[i] = (uint8_t)(x[i] + a); break;
            case 2: x[i] = (uint8_t)(x[i] + b); break;
            case 3: x[i] = (uint8_t)(x[i] + ((a + b) >> 1)); break;
            case 4: x[i] = (uint8_t)(x[i] + paeth(a, b, c)); break;
            default: bad("PNG: bad filter type");
            }
        }
        prev = x;
    }
}

Raw decode_png(const std::vector<uint8_t> &f)
{
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8)) bad("not a PNG file");
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> plte, idat;
    while (pos + 12 <= f.size()) {
        const uint32_t len = be32(&f[pos]);
        if (pos + 12 + (size_t)len > f.size()) bad("PNG: truncated chunk");
        const char *type = reinterpret_cast<const char *>(&f[pos + 4]);
        const uint8_t *d = &f[pos + 8];
        if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(d, d + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + len;
    }
    check_size(w, h);
    int ch;
    switch (ctype) {
    case 0: ch = 1; break;
    case 2: ch = 3; break;
    case 3: ch = 1; break;
    case 4: ch = 2; break;
    case 6: ch = 4; break;
    default: bad("PNG: unsupported color type");
    }
    if (!(depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16) || (ctype != 0 && ctype != 3 && depth < 8) ||
        (ctype == 3 && depth == 16))
        bad("PNG: unsupported bit depth");
    const int bits = ch * depth, bpp = std::max(1, bits / 8);
    // Adam7 passes: x0, y0, dx, dy
    static const int pass[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const int npass = interlace ? 7 : 1;
    size_t total = 0;
    for (int p = 0; p < npass; ++p) {
        const int x0 = interlace ? pass[p][0] : 0, y0 = interlace ? pass[p][1] : 0;
        const int dx = interlace ? pass[p][2] : 1, dy = interlace ? pass[p][3] : 1;
        const int64_t pw = ((int64_t)w - x0 + dx - 1) / dx, ph = ((int64_t)h - y0 + dy - 1) / dy;
        if (pw > 0 && ph > 0) total += (size_t)ph * (1 + ((size_t)pw * bits + 7) / 8);
    }
    std::vector<uint8_t> raw = inflate_all(idat.data(), idat.size(), total);
    Raw r;
    r.png = true;
    r.w = (int)w;
    r.h = (int)h;
    r.c = ctype == 3 ? 3 : ch;
    r.v.assign((size_t)w * h * r.c, 0);
    size_t off = 0;
    for (int p = 0; p < npass; ++p) {
        const int x0 = interlace ? pass[p][0] : 0, y0 = interlace ? pass[p][1] : 0;
        const int dx = interlace ? pass[p][2] : 1, dy = interlace ? pass[p][3] : 1;
        const int64_t pw = ((int64_t)w - x0 + dx - 1) / dx, ph = ((int64_t)h - y0 + dy - 1) / dy;
        if (pw <= 0 || ph <= 0) continue;
        const size_t rowbytes = ((size_t)pw * bits + 7) / 8;
        unfilter(&raw[off], (int)ph, rowbytes, bpp);
        for (int64_t j = 0; j < ph; ++j) {
            const uint8_t *row = &raw[off + j * (rowbytes + 1) + 1];
            for (int64_t i = 0; i < pw; ++i) {
                uint8_t *o = &r.v[((size_t)(y0 + j * dy) * w + (size_t)(x0 + i * dx)) * r.c];
                if (depth < 8) {   // gray or palette index, MSB first
                    const size_t bit = (size_t)i * depth;
                    const int v = (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1 << depth) - 1);
                    if (ctype == 3) {
                        if ((size_t)v * 3 + 2 >= plte.size()) bad("PNG: palette index out of range");
                        o[0] = plte[v * 3];
                        o[1] = plte[v * 3 + 1];
                        o[2] = plte[v * 3 + 2];
                    } else {
                        o[0] = (uint8_t)(v * 255 / ((1 << depth) - 1));
                    }
                } else if (ctype == 3) {
                    const int v = row[i];
                    if ((size_t)v * 3 + 2 >= plte.size()) bad("PNG: palette index out of range");
                    o[0] = plte[v * 3];
                    o[1] = plte[v * 3 + 1];
                    o[2] = plte[v * 3 + 2];
                } else {
                    for (int k = 0; k < ch; ++k)   // 16-bit samples: their high byte
                        o[k] = depth == 16 ? row[(i * ch + k) * 2] : row[i * ch + k];
                }
            }
        }
        off += (size_t)ph * (rowbytes + 1);
    }
    return r;
}

void put_chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data)
{
    const uint32_t n = (uint32_t)data.size();
    const uint8_t len[4] = {(uint8_t)(n >> 24), (uint8_t)(n >> 16), (uint8_t)(n >> 8), (uint8_t)n};
    out.insert(out.end(), len, len + 4);
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    const uint32_t crc = (uint32_t)crc32(0, &out[start], (uInt)(out.size() - start));
    const uint8_t c[4] = {(uint8_t)(crc >> 24), (uint8_t)(crc >> 16), (uint8_t)(crc >> 8), (uint8_t)crc};
    out.insert(out.end(), c, c + 4);
}

std::vector<uint8_t> encode_png(const Image8 &img)
{
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    std::vector<uint8_t> ihdr(13, 0);
    for (int k = 0; k < 4; ++k) {
        ihdr[k] = (uint8_t)(img.w >> (24 - 8 * k));
        ihdr[4 + k] = (uint8_t)(img.h >> (24 - 8 * k));
    }
    ihdr[8] = 8;
    ihdr[9] = img.c == 3 ? 2 : 0;
    put_chunk(out, "IHDR", ihdr);
    const size_t rowbytes = (size_t)img.w * img.c;
    std::vector<uint8_t> rows((rowbytes + 1) * img.h);
    for (int y = 0; y < img.h; ++y) {
        uint8_t *r = &rows[y * (rowbytes + 1)];
        r[0] = 0;
        const uint8_t *s = &img.px[y * rowbytes];
        for (int x = 0; x < img.w; ++x)
            for (int k = 0; k < img.c; ++k) r[1 + x * img.c + k] = s[x * img.c + (img.c == 3 ? 2 - k : k)];   // BGR -> RGB
    }
    uLongf zn = compressBound((uLong)rows.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, rows.data(), (uLong)rows.size(), 6) != Z_OK) bad("PNG: deflate failed");
    z.resize(zn);
    put_chunk(out, "IDAT", z);
    put_chunk(out, "IEND", {});
    return out;
}

// ---- BMP -------------------------------------------------------------------------

int mask_shift(uint32_t m)
{
    int s = 0;
    while (m && !(m & 1)) { m >>= 1; ++s; }
    return s;
}

int mask_bits(uint32_t m)
{
    m >>= mask_shift(m);
    int b = 0;
    while (m & 1) { m >>= 1; ++b; }
    return b;
}

uint8_t from_mask(uint32_t v, uint32_t m)
{
    if (!m) return 0;
    const int s = mask_shift(m), b = mask_bits(m);
    const uint32_t x = (v & m) >> s;
    return (uint8_t)(b >= 8 ? x >> (b - 8) : (x * 255 + ((1u << b) - 1) / 2) / ((1u << b) - 1));
}

Raw decode_bmp(const std::vector<uint8_t> &f)
{
    if (f.size() < 26 || f[0] != 'B' || f[1] != 'M') bad("not a BMP file");
    const uint32_t data_off = le32(&f[10]), hsize = le32(&f[14]);
    if (14 + (size_t)hsize > f.size()) bad("BMP: truncated header");
    int64_t w, h;
    int bpp;
    uint32_t comp = 0, nclr = 0, rm = 0, gm = 0, bm = 0;
    size_t pal_off = 14 + hsize;
    int pal_entry = 4;
    if (hsize == 12) {   // BITMAPCOREHEADER
        w = le16(&f[18]);
        h = (int16_t)le16(&f[20]);
        bpp = le16(&f[24]);
        pal_entry = 3;
    } else {
        if (hsize < 40) bad("BMP: unsupported header");
        w = (int32_t)le32(&f[18]);
        h = (int32_t)le32(&f[22]);
        bpp = le16(&f[28]);
        comp = le32(&f[30]);
        nclr = le32(&f[46]);
        if (comp == 3 || comp == 6) {   // BI_BITFIELDS: masks in the header (v4/v5) or after it
            const size_t mo = hsize >= 52 ? 54 : 14 + hsize;
            if (mo + 12 > f.size()) bad("BMP: truncated masks");
            rm = le32(&f[mo]);
            gm = le32(&f[mo + 4]);
            bm = le32(&f[mo + 8]);
            if (hsize < 52) pal_off += comp == 6 ? 16 : 12;
        }
    }
    const bool top_down = h < 0;
    h = top_down ? -h : h;
    check_size(w, h);
    std::vector<uint8_t> pal;
    if (bpp <= 8) {
        const uint32_t n = nclr ? nclr : (1u << bpp);
        if (pal_off + (size_t)n * pal_entry > f.size()) bad("BMP: truncated palette");
        for (uint32_t i = 0; i < n; ++i) {   // stored BGR(0)
            pal.push_back(f[pal_off + i * pal_entry + 2]);
            pal.push_back(f[pal_off + i * pal_entry + 1]);
            pal.push_back(f[pal_off + i * pal_entry]);
        }
    }
    Raw r;
    r.w = (int)w;
    r.h = (int)h;
    r.c = 3;
    r.v.assign((size_t)w * h * 3, 0);
    auto put_index = [&](int64_t x, int64_t y, int idx) {
        if (x < 0 || x >= w || y < 0 || y >= h) return;
        if ((size_t)idx * 3 + 2 >= pal.size()) idx = 0;
        uint8_t *o = &r.v[((size_t)(top_down ? y : h - 1 - y) * w + x) * 3];
        o[0] = pal[idx * 3];
        o[1] = pal[idx * 3 + 1];
        o[2] = pal[idx * 3 + 2];
    };
    if (comp == 1 || comp == 2) {   // RLE8 / RLE4, bottom-up
        size_t p = data_off;
        int64_t x = 0, y = 0;
        while (p + 1 < f.size() && y < h) {
            const int n = f[p], v = f[p + 1];
            p += 2;
            if (n) {
                for (int i = 0; i < n; ++i)
                    put_index(x++, y, comp == 1 ? v : ((i & 1) ? v & 15 : v >> 4));
            } else if (v == 0) {
                x = 0;
                ++y;
            } else if (v == 1) {
                break;
            } else if (v == 2) {
                if (p + 1 >= f.size()) break;
                x += f[p];
                y += f[p + 1];
                p += 2;
            } else {   // absolute run of v pixels, padded to 16 bits
                const size_t bytes = comp == 1 ? (size_t)v : ((size_t)v + 1) / 2;
                if (p + bytes > f.size()) break;
                for (int i = 0; i < v; ++i)
                    put_index(x++, y, comp == 1 ? f[p + i] : ((i & 1) ? f[p + i / 2] & 15 : f[p + i / 2] >> 4));
                p += (bytes + 1) & ~size_t(1);
            }
        }
        return r;
    }
    if (comp != 0 && comp != 3 && comp != 6) bad("BMP: unsupported compression");
    const size_t stride = (((size_t)w * bpp + 31) / 32) * 4;
    if (data_off + stride * h > f.size()) bad("BMP: truncated pixel data");
    if (bpp == 16 && !rm) { rm = 0x7C00; gm = 0x03E0; bm = 0x001F; }
    for (int64_t y = 0; y < h; ++y) {
        const uint8_t *s = &f[data_off + stride * y];
        const int64_t oy = top_down ? y : h - 1 - y;
        for (int64_t x = 0; x < w; ++x) {
            uint8_t *o = &r.v[((size_t)oy * w + x) * 3];
            if (bpp <= 8) {
                const size_t bit = (size_t)x * bpp;
                int idx = (s[bit >> 3] >> (8 - bpp - (bit & 7))) & ((1 << bpp) - 1);
                if ((size_t)idx * 3 + 2 >= pal.size()) idx = 0;
                o[0] = pal[idx * 3];
                o[1] = pal[idx * 3 + 1];
                o[2] = pal[idx * 3 + 2];
            } else if (bpp == 24) {
                o[0] = s[x * 3 + 2];
                o[1] = s[x * 3 + 1];
                o[2] = s[x * 3];
            } else if (bpp == 16 || bpp == 32) {
                const uint32_t v = bpp == 16 ? le16(&s[x * 2]) : le32(&s[x * 4]);
                if (bpp == 32 && !rm) {
                    o[0] = s[x * 4 + 2];
                    o[1] = s[x * 4 + 1];
                    o[2] = s[x * 4];
                } else {
                    o[0] = from_mask(v, rm);
                    o[1] = from_mask(v, gm);
                    o[2] = from_mask(v, bm);
                }
            } else {
                bad("BMP: unsupported bit depth");
            }
        }
    }
    return r;
}

std::vector<uint8_t> encode_bmp(const Image8 &img)
{
    const size_t stride = ((size_t)img.w * 3 + 3) & ~size_t(3), size = 54 + stride * img.h;
    std::vector<uint8_t> out(size, 0);
    auto put32 = [&](size_t at, uint32_t v) { for (int k = 0; k < 4; ++k) out[at + k] = (uint8_t)(v >> (8 * k)); };
    out[0] = 'B';
    out[1] = 'M';
    put32(2, (uint32_t)size);
    put32(10, 54);
    put32(14, 40);
    put32(18, (uint32_t)img.w);
    put32(22, (uint32_t)img.h);
    out[26] = 1;
    out[28] = 24;
    put32(34, (uint32_t)(stride * img.h));
    for (int y = 0; y < img.h; ++y) {
        uint8_t *d = &out[54 + stride * (img.h - 1 - y)];
        for (int x = 0; x < img.w; ++x)
            for (int k = 0; k < 3; ++k) d[x * 3 + k] = img.px[((size_t)y * img.w + x) * img.c + (img.c == 3 ? k : 0)];
    }
    return out;
}

// ---- TIFF ------------------------------------------------------------------------

struct Tiff {
    const std::vector<uint8_t> &f;
    bool le;
    uint16_t u16(size_t p) const
    {
        if (p + 2 > f.size()) bad("TIFF: truncated");
        return le ? le16(&f[p]) : be16(&f[p]);
    }
    uint32_t u32(size_t p) const
    {
        if (p + 4 > f.size()) bad("TIFF: truncated");
        return le ? le32(&f[p]) : be32(&f[p]);
    }
    // values of an IFD entry as integers (BYTE, SHORT or LONG)
    std::vector<uint32_t> values(size_t entry) const
    {
        const uint16_t type = u16(entry + 2);
        const uint32_t count = u32(entry + 4);
        const int size = type == 3 ? 2 : (type == 4 ? 4 : 1);
        if (count > (1u << 26)) bad("TIFF: bad entry count");
        const size_t at = (size_t)count * size <= 4 ? entry + 8 : u32(entry + 8);
        std::vector<uint32_t> v(count);
        for (uint32_t i = 0; i < count; ++i) {
            const size_t p = at + (size_t)i * size;
            v[i] = size == 2 ? u16(p) : (size == 4 ? u32(p) : (p < f.size() ? f[p] : 0));
        }
        return v;
    }
};

std::vector<uint8_t> lzw_tiff(const uint8_t *p, size_t n, size_t want)
{
    std::vector<uint8_t> out;
    out.reserve(want);
    std::vector<std::vector<uint8_t>> dict;
    auto reset = [&] {
        dict.assign(258, {});
        for (int i = 0; i < 256; ++i) dict[i] = {(uint8_t)i};
    };
    reset();
    size_t bitpos = 0;
    int width = 9;
    int prev = -1;
    while (out.size() < want) {
        if ((bitpos + width + 7) / 8 > n) break;
        uint32_t code = 0;   // MSB first
        for (int k = 0; k < width; ++k, ++bitpos) code = (code << 1) | ((p[bitpos >> 3] >> (7 - (bitpos & 7))) & 1);
        if (code == 256) {
            reset();
            width = 9;
            prev = -1;
            continue;
        }
        if (code == 257) break;
        std::vector<uint8_t> entry;
        if (code < dict.size()) {
            entry = dict[code];
            if (prev >= 0) {
                std::vector<uint8_t> e = dict[prev];
                e.push_back(entry[0]);
                dict.push_back(e);
            }
        } else if (prev >= 0 && code == dict.size()) {
            entry = dict[prev];
            entry.push_back(dict[prev][0]);
            dict.push_back(entry);
        } else {
            bad("TIFF: corrupt LZW data");
        }
        out.insert(out.end(), entry.begin(), entry.end());
        prev = (int)code;
        // the code width grows one code early (TIFF 6.0, section 13)
        if (dict.size() + 1 >= (size_t)1 << width && width < 12) ++width;
    }
    out.resize(want, 0);
    return out;
}

std::vector<uint8_t> packbits(const uint8_t *p, size_t n, size_t want)
{
    std::vector<uint8_t> out;
    size_t i = 0;
    while (i < n && out.size() < want) {
        const int8_t c = (int8_t)p[i++];
        if (c >= 0) {
            for (int k = 0; k <= c && i < n; ++k) out.push_back(p[i++]);
        } else if (c != -128 && i < n) {
            out.insert(out.end(), (size_t)(1 - c), p[i++]);
        }
    }
    out.resize(want, 0);
    return out;
}

Raw decode_tiff(const std::vector<uint8_t> &f)
{
    if (f.size() < 8) bad("not a TIFF file");
    const bool le = f[0] == 'I' && f[1] == 'I';
    if (!le && !(f[0] == 'M' && f[1] == 'M')) bad("not a TIFF file");
    Tiff t{f, le};
    if (t.u16(2) != 42) bad("not a TIFF file");
    const size_t ifd = t.u32(4);
    const uint16_t n = t.u16(ifd);
    uint32_t w = 0, h = 0, comp = 1, photo = 1, spp = 1, rps = 0xFFFFFFFFu, planar = 1, pred = 1, fill = 1;
    std::vector<uint32_t> bps{1}, offs, counts, cmap;
    for (uint16_t i = 0; i < n; ++i) {
        const size_t e = ifd + 2 + (size_t)i * 12;
        const uint16_t tag = t.u16(e);
        switch (tag) {
        case 256: w = t.values(e)[0]; break;
        case 257: h = t.values(e)[0]; break;
        case 258: bps = t.values(e); break;
        case 259: comp = t.values(e)[0]; break;
        case 262: photo = t.values(e)[0]; break;
        case 266: fill = t.values(e)[0]; break;
        case 273: offs = t.values(e); break;
        case 277: spp = t.values(e)[0]; break;
        case 278: rps = t.values(e)[0]; break;
        case 279: counts = t.values(e); break;
        case 284: planar = t.values(e)[0]; break;
        case 317: pred = t.values(e)[0]; break;
        case 320: cmap = t.values(e); break;
        case 322: bad("TIFF: tiled images are not supported");
        default: break;
        }
    }
    check_size(w, h);
    const uint32_t depth = bps.empty() ? 1 : bps[0];
    if (planar != 1 && spp > 1) bad("TIFF: planar configuration 2 is not supported");
    if (!(depth == 1 || depth == 8 || depth == 16) || spp < 1 || spp > 4 || fill != 1) bad("TIFF: unsupported samples");
    if (offs.empty() || offs.size() != counts.size()) bad("TIFF: missing strips");
    rps = std::min<uint32_t>(rps, h);
    const size_t rowbytes = ((size_t)w * spp * depth + 7) / 8;
    std::vector<uint8_t> data;
    data.reserve(rowbytes * h);
    for (size_t s = 0; s < offs.size() && data.size() < rowbytes * h; ++s) {
        const size_t rows = std::min<size_t>(rps, h - data.size() / rowbytes), want = rows * rowbytes;
        if ((size_t)offs[s] + counts[s] > f.size()) bad("TIFF: truncated strip");
        const uint8_t *p = &f[offs[s]];
        std::vector<uint8_t> strip;
        if (comp == 1) strip.assign(p, p + std::min<size_t>(counts[s], want));
        else if (comp == 5) strip = lzw_tiff(p, counts[s], want);
        else if (comp == 32773) strip = packbits(p, counts[s], want);
        else if (comp == 8 || comp == 32946) {
            std::vector<uint8_t> z(want);
            uLongf zn = (uLongf)want;
            const int rc = uncompress(z.data(), &zn, p, counts[s]);
            if (rc != Z_OK && rc != Z_BUF_ERROR) bad("TIFF: corrupt deflate strip");
            strip = z;
        } else bad("TIFF: unsupported compression");
        strip.resize(want, 0);
        if (pred == 2) {   // horizontal differencing, per sample
            for (size_t y = 0; y < rows; ++y) {
                uint8_t *r = &strip[y * rowbytes];
                if (depth == 8)
                    for (size_t i = spp; i < (size_t)w * spp; ++i) r[i] = (uint8_t)(r[i] + r[i - spp]);
                else if (depth == 16)
                    for (size_t i = spp; i < (size_t)w * spp; ++i) {
                        const uint16_t a = le ? le16(&r[2 * i]) : be16(&r[2 * i]);
                        const uint16_t b = le ? le16(&r[2 * (i - spp)]) : be16(&r[2 * (i - spp)]);
                        const uint16_t v = (uint16_t)(a + b);
                        r[2 * i] = (uint8_t)(le ? v : v >> 8);
                        r[2 * i + 1] = (uint8_t)(le ? v >> 8 : v);
                    }
            }
        }
        data.insert(data.end(), strip.begin(), strip.end());
    }
    data.resize(rowbytes * h, 0);
    Raw r;
    r.w = (int)w;
    r.h = (int)h;
    const bool palette = photo == 3;
    r.c = palette ? 3 : (photo == 2 ? (spp >= 4 ? 4 : 3) : (spp >= 2 ? 2 : 1));
    if (photo == 2 && spp < 3) bad("TIFF: RGB with fewer than 3 samples");
    r.v.assign((size_t)w * h * r.c, 0);
    for (size_t y = 0; y < h; ++y) {
        const uint8_t *row = &data[y * rowbytes];
        for (size_t x = 0; x < w; ++x) {
            uint8_t *o = &r.v[(y * w + x) * r.c];
            auto sample = [&](uint32_t k) -> uint32_t {   // sample k of pixel x, at full precision
                const size_t i = x * spp + k;
                if (depth == 1) return (row[i >> 3] >> (7 - (i & 7))) & 1;
                if (depth == 8) return row[i];
                return le ? le16(&row[2 * i]) : be16(&row[2 * i]);
            };
            auto to8 = [&](uint32_t v) -> uint8_t { return depth == 1 ? (v ? 255 : 0) : (depth == 16 ? v >> 8 : v); };
            if (palette) {
                const uint32_t idx = sample(0), nc = (uint32_t)cmap.size() / 3;
                if (idx >= nc) bad("TIFF: palette index out of range");
                o[0] = (uint8_t)(cmap[idx] >> 8);
                o[1] = (uint8_t)(cmap[nc + idx] >> 8);
                o[2] = (uint8_t)(cmap[2 * nc + idx] >> 8);
            } else {
                for (int k = 0; k < r.c; ++k) {
                    uint8_t v = to8(sample(k));
                    if (photo == 0 && k == 0) v = (uint8_t)(255 - v);   // WhiteIsZero
                    o[k] = v;
                }
            }
        }
    }
    return r;
}

std::vector<uint8_t> encode_tiff(const Image8 &img)
{
    const uint32_t c = img.c == 3 ? 3 : 1, n = 10;
    const size_t pix = (size_t)img.w * img.h * c, ifd = 8 + pix + (pix & 1);
    std::vector<uint8_t> out(ifd + 2 + n * 12 + 4 + 8, 0);
    auto p16 = [&](size_t at, uint32_t v) { out[at] = (uint8_t)v; out[at + 1] = (uint8_t)(v >> 8); };
    auto p32 = [&](size_t at, uint32_t v) { for (int k = 0; k < 4; ++k) out[at + k] = (uint8_t)(v >> (8 * k)); };
    out[0] = out[1] = 'I';
    p16(2, 42);
    p32(4, (uint32_t)ifd);
    for (int y = 0; y < img.h; ++y)
        for (int x = 0; x < img.w; ++x)
            for (uint32_t k = 0; k < c; ++k)   // BGR -> RGB
                out[8 + ((size_t)y * img.w + x) * c + k] = img.px[((size_t)y * img.w + x) * img.c + (c == 3 ? 2 - k : 0)];
    const size_t bps_at = ifd + 2 + n * 12 + 4;   // three SHORTs of 8 (RGB)
    p16(bps_at, 8);
    p16(bps_at + 2, 8);
    p16(bps_at + 4, 8);
    p16(ifd, n);
    const uint32_t entries[n][4] = {{256, 4, 1, (uint32_t)img.w},
                                    {257, 4, 1, (uint32_t)img.h},
                                    {258, 3, c, c == 3 ? (uint32_t)bps_at : 8u},
                                    {259, 3, 1, 1},
                                    {262, 3, 1, c == 3 ? 2u : 1u},
                                    {273, 4, 1, 8},
                                    {277, 3, 1, c},
                                    {278, 4, 1, (uint32_t)img.h},
                                    {279, 4, 1, (uint32_t)pix},
                                    {284, 3, 1, 1}};
    for (uint32_t i = 0; i < n; ++i) {
        const size_t e = ifd + 2 + i * 12;
        p16(e, entries[i][0]);
        p16(e + 2, entries[i][1]);
        p32(e + 4, entries[i][2]);
        if (entries[i][1] == 3 && entries[i][2] == 1) p16(e + 8, entries[i][3]);
        else p32(e + 8, entries[i][3]);
    }
    return out;
}

// ---- PNM ---------------------------------------------------------------------------

Raw decode_pnm(const std::vector<uint8_t> &f)
{
    if (f.size() < 3 || f[0] != 'P' || f[1] < '1' || f[1] > '6') bad("not a PNM file");
    const int kind = f[1] - '0';
    size_t p = 2;
    auto next_int = [&]() -> long {
        for (;;) {
            while (p < f.size() && std::isspace(f[p])) ++p;
            if (p < f.size() && f[p] == '#') {
                while (p < f.size() && f[p] != '\n') ++p;
                continue;
            }
            break;
        }
        if (p >= f.size() || !std::isdigit(f[p])) bad("PNM: bad header");
        long v = 0;
        while (p < f.size() && std::isdigit(f[p])) v = v * 10 + (f[p++] - '0');
        return v;
    };
    const long w = next_int(), h = next_int();
    const long maxv = (kind == 1 || kind == 4) ? 1 : next_int();
    check_size(w, h);
    if (maxv < 1 || maxv > 65535) bad("PNM: bad maxval");
    const int c = (kind == 3 || kind == 6) ? 3 : 1;
    Raw r;
    r.w = (int)w;
    r.h = (int)h;
    r.c = c;
    r.v.assign((size_t)w * h * c, 0);
    auto scale = [&](long v) -> uint8_t {
        if (kind == 1 || kind == 4) return v ? 0 : 255;   // 1 = black
        if (maxv == 255) return (uint8_t)v;
        if (maxv > 255) return (uint8_t)(std::min(v, maxv) * 255 / maxv);
        return (uint8_t)((std::min(v, maxv) * 255 + maxv / 2) / maxv);
    };
    const size_t n = (size_t)w * h * c;
    if (kind <= 3) {
        for (size_t i = 0; i < n; ++i) {
            if (kind == 1) {   // plain bits may be unseparated
                while (p < f.size() && (std::isspace(f[p]) || f[p] == '#')) {
                    if (f[p] == '#') while (p < f.size() && f[p] != '\n') ++p;
                    else ++p;
                }
                if (p >= f.size()) bad("PNM: truncated data");
                r.v[i] = scale(f[p++] - '0');
            } else {
                r.v[i] = scale(next_int());
            }
        }
        return r;
    }
    ++p;   // the single whitespace after the header
    if (kind == 4) {
        const size_t rowbytes = ((size_t)w + 7) / 8;
        if (p + rowbytes * h > f.size()) bad("PNM: truncated data");
        for (long y = 0; y < h; ++y)
            for (long x = 0; x < w; ++x) r.v[(size_t)y * w + x] = scale((f[p + y * rowbytes + x / 8] >> (7 - x % 8)) & 1);
        return r;
    }
    const int bytes = maxv > 255 ? 2 : 1;
    if (p + n * bytes > f.size()) bad("PNM: truncated data");
    for (size_t i = 0; i < n; ++i) r.v[i] = scale(bytes == 2 ? be16(&f[p + 2 * i]) : f[p + i]);
    return r;
}

std::vector<uint8_t> encode_pnm(const Image8 &img, bool gray)
{
    const std::string head = std::string(gray ? "P5\n" : "P6\n") + std::to_string(img.w) + " " + std::to_string(img.h) + "\n255\n";
    std::vector<uint8_t> out(head.begin(), head.end());
    for (size_t i = 0; i < (size_t)img.w * img.h; ++i) {
        const uint8_t *s = &img.px[i * img.c];
        if (gray) out.push_back(img.c == 1 ? s[0] : (uint8_t)((s[0] * 1868 + s[1] * 9617 + s[2] * 4899 + 8192) >> 14));
        else for (int k = 0; k < 3; ++k) out.push_back(img.c == 3 ? s[2 - k] : s[0]);
    }
    return out;
}

// ---- JPEG ------------------------------------------------------------------------

const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                             41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                             30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huff {
    bool set = false;
    uint8_t vals[256] = {};
    int maxcode[18] = {}, valptr[17] = {}, mincode[17] = {};
    void build(const uint8_t *counts, const uint8_t *v, int nv)
    {
        std::memcpy(vals, v, (size_t)std::min(nv, 256));
        int code = 0, k = 0;
        for (int len = 1; len <= 16; ++len) {
            valptr[len] = k;
            mincode[len] = code;
            code += counts[len - 1];
            k += counts[len - 1];
            maxcode[len] = counts[len - 1] ? code - 1 : -1;
            code <<= 1;
        }
        maxcode[17] = 0x7FFFFFFF;
        set = true;
    }
};

// Separable 8x8 inverse DCT in double precision (T.81 A.3.3), rounded.
struct Idct {
    double c[8][8];
    Idct()
    {
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u) c[x][u] = (u ? 1.0 : std::sqrt(0.5)) * std::cos((2 * x + 1) * u * M_PI / 16) / 2;
    }
    void run(const int *in, uint8_t *out, int stride) const
    {
        double t[64];
        for (int y = 0; y < 8; ++y)
            for (int u = 0; u < 8; ++u) {
                double s = 0;
                for (int v = 0; v < 8; ++v) s += c[y][v] * in[v * 8 + u];
                t[y * 8 + u] = s;
            }
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                double s = 0;
                for (int u = 0; u < 8; ++u) s += c[x][u] * t[y * 8 + u];
                const int v = (int)std::floor(s + 128.5);
                out[y * stride + x] = (uint8_t)std::min(255, std::max(0, v));
            }
    }
};

struct JpegDecoder {
    const std::vector<uint8_t> &f;
    size_t pos = 0;
    uint16_t qt[4][64] = {};
    Huff dc[4], ac[4];
    struct Comp {
        int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0, pred = 0;
        int bw = 0, bh = 0;         // blocks across / down (padded to whole MCUs)
        std::vector<uint8_t> plane; // bw*8 x bh*8 samples
    };
    std::vector<Comp> comps;
    int W = 0, H = 0, hmax = 1, vmax = 1, restart = 0;
    bool adobe_rgb = false, have_frame = false;
    uint32_t bitbuf = 0;
    int bitcnt = 0;
    Idct idct;

    explicit JpegDecoder(const std::vector<uint8_t> &file) : f(file) {}

    uint8_t byte() { return pos < f.size() ? f[pos++] : 0; }
    int bit()
    {
        if (bitcnt == 0) {
            uint8_t b = byte();
            if (b == 0xFF) {
                const uint8_t nx = pos < f.size() ? f[pos] : 0;
                if (nx == 0) ++pos;                      // stuffed zero
                else { --pos; b = 0; }                   // a marker: feed zeros
            }
            bitbuf = b;
            bitcnt = 8;
        }
        --bitcnt;
        return (bitbuf >> bitcnt) & 1;
    }
    int bits(int n)
    {
        int v = 0;
        for (int i = 0; i < n; ++i) v = (v << 1) | bit();
        return v;
    }
    int decode(const Huff &t)
    {
        if (!t.set) bad("JPEG: missing Huffman table");
        int code = 0;
        for (int len = 1; len <= 16; ++len) {
            code = (code << 1) | bit();
            if (t.maxcode[len] >= 0 && code <= t.maxcode[len]) return t.vals[t.valptr[len] + code - t.mincode[len]];
        }
        bad("JPEG: bad Huffman code");
    }
    static int extend(int v, int s) { return s && v < (1 << (s - 1)) ? v - (1 << s) + 1 : v; }

    void block(Comp &c, int bx, int by)
    {
        int coef[64] = {};
        const uint16_t *q = qt[c.tq];
        // baseline 8-bit data: DC differences of at most 11 bits, AC values of
        // at most 10, a quantised DC within +-2047 (the level-shifted DC term
        // lies in [-1024, 1016]); anything larger is a damaged file
        const int s = decode(dc[c.td]);
        if (s > 11) bad("JPEG: bad DC difference size");
        c.pred += extend(bits(s), s);
        if (c.pred < -2048 || c.pred > 2047) bad("JPEG: DC value out of range");
        coef[0] = c.pred * q[0];
        for (int k = 1; k < 64;) {
            const int rs = decode(ac[c.ta]), r = rs >> 4, sz = rs & 15;
            if (!sz) {
                if (r != 15) break;
                k += 16;
                continue;
            }
            if (sz > 10) bad("JPEG: bad AC value size");
            k += r;
            if (k > 63) break;
            coef[kZigzag[k]] = extend(bits(sz), sz) * q[k];
            ++k;
        }
        if (bx < c.bw && by < c.bh) idct.run(coef, &c.plane[(size_t)by * 8 * c.bw * 8 + bx * 8], c.bw * 8);
    }

    void restart_marker()
    {
        bitcnt = 0;
        while (pos + 1 < f.size() && !(f[pos] == 0xFF && f[pos + 1] >= 0xD0 && f[pos + 1] <= 0xD7)) ++pos;
        if (pos + 1 < f.size()) pos += 2;
        for (Comp &c : comps) c.pred = 0;
    }

    void scan(const std::vector<int> &ids)
    {
        std::vector<Comp *> sc;
        for (int id : ids)
            for (Comp &c : comps)
                if (c.id == id) sc.push_back(&c);
        for (Comp &c : comps) c.pred = 0;
        bitcnt = 0;
        int mcus = 0;
        if (sc.size() == 1) {   // non-interleaved: the component's own blocks
            Comp &c = *sc[0];
            const int cw = (int)(((int64_t)W * c.h + hmax - 1) / hmax), chh = (int)(((int64_t)H * c.v + vmax - 1) / vmax);
            const int nbx = (cw + 7) / 8, nby = (chh + 7) / 8;
            for (int by = 0; by < nby; ++by)
                for (int bx = 0; bx < nbx; ++bx) {
                    if (restart && mcus && mcus % restart == 0) restart_marker();
                    block(c, bx, by);
                    ++mcus;
                }
        } else {
            const int mx = (W + 8 * hmax - 1) / (8 * hmax), my = (H + 8 * vmax - 1) / (8 * vmax);
            for (int y = 0; y < my; ++y)
                for (int x = 0; x < mx; ++x) {
                    if (restart && mcus && mcus % restart == 0) restart_marker();
                    for (Comp *c : sc)
                        for (int v = 0; v < c->v; ++v)
                            for (int h = 0; h < c->h; ++h) block(*c, x * c->h + h, y * c->v + v);
                    ++mcus;
                }
        }
        bitcnt = 0;
    }

    void run()
    {
        if (f.size() < 4 || f[0] != 0xFF || f[1] != 0xD8) bad("not a JPEG file");
        pos = 2;
        while (pos < f.size()) {
            if (f[pos] != 0xFF) { ++pos; continue; }
            const uint8_t m = f[pos + 1 < f.size() ? pos + 1 : pos];
            pos += 2;
            if (m == 0xD9) break;
            if (m == 0xFF || (m >= 0xD0 && m <= 0xD7) || m == 0x01) { if (m == 0xFF) --pos; continue; }
            if (pos + 2 > f.size()) bad("JPEG: truncated segment");
            const size_t len = be16(&f[pos]), end = pos + len;
            if (end > f.size() || len < 2) bad("JPEG: truncated segment");
            const uint8_t *d = &f[pos + 2];
            const size_t n = len - 2;   // bytes of the segment after its length
            if (m == 0xDB) {   // DQT
                size_t i = 0;
                while (i < n) {
                    const int pq = d[i] >> 4, tq = d[i] & 3;
                    if (i + 1 + (pq ? 128u : 64u) > n) bad("JPEG: bad DQT");
                    ++i;
                    for (int k = 0; k < 64; ++k) qt[tq][k] = pq ? be16(&d[i + 2 * k]) : d[i + k];
                    i += pq ? 128 : 64;
                }
            } else if (m == 0xC4) {   // DHT
                size_t i = 0;
                while (i + 17 <= n) {
                    const int tc = d[i] >> 4, th = d[i] & 3;
                    int nv = 0;
                    for (int k = 0; k < 16; ++k) nv += d[i + 1 + k];
                    if (i + 17 + nv > n) bad("JPEG: bad DHT");
                    (tc ? ac : dc)[th].build(&d[i + 1], &d[i + 17], nv);
                    i += 17 + nv;
                }
            } else if (m == 0xC0 || m == 0xC1) {   // baseline / extended sequential Huffman
                if (n < 6 || n < 6 + 3 * (size_t)d[5]) bad("JPEG: bad frame header");
                if (d[0] != 8) bad("JPEG: only 8-bit samples are supported");
                H = be16(&d[1]);
                W = be16(&d[3]);
                check_size(W, H);
                const int nc = d[5];
                if (nc != 1 && nc != 3) bad("JPEG: only 1 or 3 components are supported");
                comps.clear();
                for (int k = 0; k < nc; ++k) {
                    Comp c;
                    c.id = d[6 + 3 * k];
                    c.h = d[7 + 3 * k] >> 4;
                    c.v = d[7 + 3 * k] & 15;
                    c.tq = d[8 + 3 * k] & 3;
                    if (c.h < 1 || c.v < 1 || c.h > 4 || c.v > 4) bad("JPEG: bad sampling factors");
                    comps.push_back(c);
                }
                hmax = vmax = 1;
                for (const Comp &c : comps) { hmax = std::max(hmax, c.h); vmax = std::max(vmax, c.v); }
                const int mx = (W + 8 * hmax - 1) / (8 * hmax), my = (H + 8 * vmax - 1) / (8 * vmax);
                for (Comp &c : comps) {
                    c.bw = mx * c.h;
                    c.bh = my * c.v;
                    c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
                }
                have_frame = true;
            } else if (m >= 0xC2 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) {
                bad("JPEG: progressive, lossless and arithmetic-coded files are not supported");
            } else if (m == 0xDD) {
                if (n < 2) bad("JPEG: bad DRI");
                restart = be16(d);
            } else if (m == 0xEE && len >= 14 && !std::memcmp(d, "Adobe", 5)) {
                adobe_rgb = d[11] == 0;   // transform 0: RGB (or CMYK), no YCbCr
            } else if (m == 0xDA) {   // SOS
                if (!have_frame) bad("JPEG: scan before frame");
                if (n < 1 || n < 1 + 2 * (size_t)d[0]) bad("JPEG: bad scan header");
                const int ns = d[0];
                std::vector<int> ids;
                for (int k = 0; k < ns; ++k) {
                    const int id = d[1 + 2 * k], td = d[2 + 2 * k] >> 4, ta = d[2 + 2 * k] & 15;
                    if (td > 3 || ta > 3) bad("JPEG: bad Huffman table selector");
                    ids.push_back(id);
                    for (Comp &c : comps)
                        if (c.id == id) { c.td = td; c.ta = ta; }
                }
                pos = end;
                scan(ids);
                continue;
            }
            pos = end;
        }
        if (!have_frame) bad("JPEG: no frame");
    }

    // The component's samples at full resolution (row stride W): replicated,
    // or for 2:1 chroma the IJG triangle filter ("fancy upsampling").
    std::vector<uint8_t> full(const Comp &c) const
    {
        const int sw = c.bw * 8, sh = c.bh * 8;
        std::vector<uint8_t> out((size_t)W * H);
        const uint8_t *in = c.plane.data();
        const int fx = hmax / c.h, fy = vmax / c.v;
        const int cw = (int)(((int64_t)W * c.h + hmax - 1) / hmax), ch = (int)(((int64_t)H * c.v + vmax - 1) / vmax);
        if (fx == 1 && fy == 1) {
            for (int y = 0; y < H; ++y) std::memcpy(&out[(size_t)y * W], &in[(size_t)y * sw], (size_t)W);
            return out;
        }
        if (fx == 2 && fy == 1 && hmax % c.h == 0) {   // h2v1: out = (3 x near + far) / 4
            for (int y = 0; y < H; ++y) {
                const uint8_t *r = &in[(size_t)y * sw];
                uint8_t *o = &out[(size_t)y * W];
                for (int x = 0; x < cw; ++x) {
                    const int a = x > 0 ? (3 * r[x] + r[x - 1] + 1) >> 2 : r[x];
                    const int b = x + 1 < cw ? (3 * r[x] + r[x + 1] + 2) >> 2 : r[x];
                    if (2 * x < W) o[2 * x] = (uint8_t)a;
                    if (2 * x + 1 < W) o[2 * x + 1] = (uint8_t)b;
                }
            }
            return out;
        }
        if (fx == 2 && fy == 2 && hmax % c.h == 0 && vmax % c.v == 0) {   // h2v2: the same in both directions
            std::vector<int> colsum(cw);
            for (int oy = 0; oy < H; ++oy) {
                const int iy = oy / 2;   // the nearer input row weighs 3, the other 1 (edge rows repeat)
                const int ny = (oy & 1) ? std::min(iy + 1, ch - 1) : std::max(iy - 1, 0);
                for (int x = 0; x < cw; ++x) colsum[x] = 3 * in[(size_t)iy * sw + x] + in[(size_t)ny * sw + x];
                uint8_t *o = &out[(size_t)oy * W];
                for (int x = 0; x < cw; ++x) {
                    const int cur = colsum[x];
                    const int a = x > 0 ? (3 * cur + colsum[x - 1] + 8) >> 4 : (4 * cur + 8) >> 4;
                    const int b = x + 1 < cw ? (3 * cur + colsum[x + 1] + 7) >> 4 : (4 * cur + 7) >> 4;
                    if (2 * x < W) o[2 * x] = (uint8_t)a;
                    if (2 * x + 1 < W) o[2 * x + 1] = (uint8_t)b;
                }
            }
            return out;
        }
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) out[(size_t)y * W + x] = in[(size_t)std::min(y / fy, sh - 1) * sw + std::min(x / fx, sw - 1)];
        return out;
    }
};

Raw decode_jpeg(const std::vector<uint8_t> &f)
{
    JpegDecoder d(f);
    d.run();
    Raw r;
    r.w = d.W;
    r.h = d.H;
    if (d.comps.size() == 1) {
        r.c = 1;
        r.v = d.full(d.comps[0]);
        return r;
    }
    const std::vector<uint8_t> y = d.full(d.comps[0]), cb = d.full(d.comps[1]), cr = d.full(d.comps[2]);
    r.c = 3;
    r.v.resize((size_t)d.W * d.H * 3);
    if (d.adobe_rgb) {
        for (size_t i = 0; i < (size_t)d.W * d.H; ++i) {
            r.v[3 * i] = y[i];
            r.v[3 * i + 1] = cb[i];
            r.v[3 * i + 2] = cr[i];
        }
        return r;
    }
    r.luma = y;
    // JFIF YCbCr -> RGB in 16-bit fixed point, rounded (the IJG tables)
    auto fix = [](double v) { return (int)(v * 65536.0 + 0.5); };
    const int kr = fix(1.40200), kb = fix(1.77200), kgb = fix(0.34414), kgr = fix(0.71414);
    for (size_t i = 0; i < (size_t)d.W * d.H; ++i) {
        const int Y = y[i], Cb = cb[i] - 128, Cr = cr[i] - 128;
        const int R = Y + ((kr * Cr + 32768) >> 16);
        const int G = Y + ((-kgb * Cb - kgr * Cr + 32768) >> 16);
        const int B = Y + ((kb * Cb + 32768) >> 16);
        r.v[3 * i] = (uint8_t)std::min(255, std::max(0, R));
        r.v[3 * i + 1] = (uint8_t)std::min(255, std::max(0, G));
        r.v[3 * i + 2] = (uint8_t)std::min(255, std::max(0, B));
    }
    return r;
}

// Baseline JPEG writer: 4:2:0 YCbCr (gray: one component), quality 95 (the
// cv::imwrite default), the T.81 Annex K example tables scaled as the IJG
// quality setting does, floating-point forward DCT.
struct BitWriter {
    std::vector<uint8_t> &out;
    uint32_t buf = 0;
    int n = 0;
    void put(uint32_t code, int len)
    {
        for (int i = len - 1; i >= 0; --i) {
            buf = (buf << 1) | ((code >> i) & 1);
            if (++n == 8) {
                out.push_back((uint8_t)buf);
                if (buf == 0xFF) out.push_back(0);
                buf = 0;
                n = 0;
            }
        }
    }
    void flush()
    {
        while (n) put(1, 1);
    }
};

const uint8_t kStdLumQ[64] = {16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57,
                              69, 56, 14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64,
                              81, 104, 113, 92, 49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const uint8_t kStdChrQ[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99,
                              99, 99, 47, 66, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                              99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
// Annex K.3 tables: code-length counts and values
const uint8_t kDcLumBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcChrBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcLumBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcLumVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71, 0x14,
    0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72, 0x82, 0x09,
    0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a,
    0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64, 0x65,
    0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88,
    0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9,
    0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca,
    0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea,
    0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
const uint8_t kAcChrBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const uint8_t kAcChrVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22, 0x32,
    0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1, 0x0a, 0x16,
    0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36, 0x37, 0x38, 0x39,
    0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64,
    0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x82, 0x83, 0x84, 0x85, 0x86,
    0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7,
    0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8,
    0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9,
    0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

struct EncTable {
    uint16_t code[256] = {};
    uint8_t len[256] = {};
    EncTable(const uint8_t *bits, const uint8_t *vals)
    {
        int code_v = 0, k = 0;
        for (int l = 1; l <= 16; ++l) {
            for (int i = 0; i < bits[l - 1]; ++i, ++k) {
                code[vals[k]] = (uint16_t)code_v++;
                len[vals[k]] = (uint8_t)l;
            }
            code_v <<= 1;
        }
    }
};

std::vector<uint8_t> encode_jpeg(const Image8 &img, int quality = 95)
{
    const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
    uint8_t q[2][64];   // zigzag order
    for (int k = 0; k < 64; ++k) {
        q[0][k] = (uint8_t)std::min(255, std::max(1, (kStdLumQ[kZigzag[k]] * scale + 50) / 100));
        q[1][k] = (uint8_t)std::min(255, std::max(1, (kStdChrQ[kZigzag[k]] * scale + 50) / 100));
    }
    const bool color = img.c == 3;
    const int W = img.w, H = img.h, nc = color ? 3 : 1, f = color ? 2 : 1;
    std::vector<uint8_t> out = {0xFF, 0xD8, 0xFF, 0xE0, 0, 16, 'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0};
    for (int t = 0; t < (color ? 2 : 1); ++t) {
        const uint8_t head[5] = {0xFF, 0xDB, 0, 67, (uint8_t)t};
        out.insert(out.end(), head, head + 5);
        out.insert(out.end(), q[t], q[t] + 64);
    }
    const uint8_t sof[] = {0xFF, 0xC0, 0, (uint8_t)(8 + 3 * nc), 8, (uint8_t)(H >> 8), (uint8_t)H, (uint8_t)(W >> 8),
                           (uint8_t)W, (uint8_t)nc};
    out.insert(out.end(), sof, sof + 10);
    for (int k = 0; k < nc; ++k) {
        const uint8_t c[3] = {(uint8_t)(k + 1), (uint8_t)(k == 0 ? (f << 4 | f) : 0x11), (uint8_t)(k ? 1 : 0)};
        out.insert(out.end(), c, c + 3);
    }
    auto dht = [&](int cls, int id, const uint8_t *bits, const uint8_t *vals, int nv) {
        const uint8_t h[5] = {0xFF, 0xC4, (uint8_t)((19 + nv) >> 8), (uint8_t)(19 + nv), (uint8_t)(cls << 4 | id)};
        out.insert(out.end(), h, h + 5);
        out.insert(out.end(), bits, bits + 16);
        out.insert(out.end(), vals, vals + nv);
    };
    dht(0, 0, kDcLumBits, kDcVals, 12);
    dht(1, 0, kAcLumBits, kAcLumVals, 162);
    if (color) {
        dht(0, 1, kDcChrBits, kDcVals, 12);
        dht(1, 1, kAcChrBits, kAcChrVals, 162);
    }
    const uint8_t sos_head[] = {0xFF, 0xDA, 0, (uint8_t)(6 + 2 * nc), (uint8_t)nc};
    out.insert(out.end(), sos_head, sos_head + 5);
    for (int k = 0; k < nc; ++k) {
        const uint8_t c[2] = {(uint8_t)(k + 1), (uint8_t)(k ? 0x11 : 0x00)};
        out.insert(out.end(), c, c + 2);
    }
    const uint8_t sos_tail[3] = {0, 63, 0};
    out.insert(out.end(), sos_tail, sos_tail + 3);
    // planes (YCbCr of BGR input, the JFIF weights), padded by edge replication
    const int mw = 8 * f, mx = (W + mw - 1) / mw, my = (H + mw - 1) / mw, PW = mx * mw, PH = my * mw;
    std::vector<float> P[3];
    for (int k = 0; k < nc; ++k) P[k].assign((size_t)PW * PH, 0.f);
    for (int y = 0; y < PH; ++y)
        for (int x = 0; x < PW; ++x) {
            const uint8_t *s = &img.px[((size_t)std::min(y, H - 1) * W + std::min(x, W - 1)) * img.c];
            if (!color) {
                P[0][(size_t)y * PW + x] = s[0];
                continue;
            }
            const float B = s[0], G = s[1], R = s[2];
            P[0][(size_t)y * PW + x] = 0.299f * R + 0.587f * G + 0.114f * B;
            P[1][(size_t)y * PW + x] = -0.168736f * R - 0.331264f * G + 0.5f * B + 128.f;
            P[2][(size_t)y * PW + x] = 0.5f * R - 0.418688f * G - 0.081312f * B + 128.f;
        }
    const EncTable tdc[2] = {EncTable(kDcLumBits, kDcVals), EncTable(kDcChrBits, kDcVals)};
    const EncTable tac[2] = {EncTable(kAcLumBits, kAcLumVals), EncTable(kAcChrBits, kAcChrVals)};
    double cosv[8][8];
    for (int x = 0; x < 8; ++x)
        for (int u = 0; u < 8; ++u) cosv[x][u] = (u ? 1.0 : std::sqrt(0.5)) * std::cos((2 * x + 1) * u * M_PI / 16) / 2;
    BitWriter bw{out};
    int pred[3] = {0, 0, 0};
    auto encode_block = [&](int comp, const float *blk) {
        double t[64], F[64];
        for (int y = 0; y < 8; ++y)
            for (int u = 0; u < 8; ++u) {
                double s = 0;
                for (int x = 0; x < 8; ++x) s += cosv[x][u] * (blk[y * 8 + x] - 128.0);
                t[y * 8 + u] = s;
            }
        for (int v = 0; v < 8; ++v)
            for (int u = 0; u < 8; ++u) {
                double s = 0;
                for (int y = 0; y < 8; ++y) s += cosv[y][v] * t[y * 8 + u];
                F[v * 8 + u] = s;
            }
        const int tq = comp ? 1 : 0;
        int zz[64];
        for (int k = 0; k < 64; ++k) zz[k] = (int)std::lround(F[kZigzag[k]] / q[tq][k]);
        auto magnitude = [](int v, int &bits) {
            int a = v < 0 ? -v : v, n = 0;
            while (a) { ++n; a >>= 1; }
            bits = v < 0 ? v + (1 << n) - 1 : v;
            return n;
        };
        int bits, n = magnitude(zz[0] - pred[comp], bits);
        pred[comp] = zz[0];
        bw.put(tdc[tq].code[n], tdc[tq].len[n]);
        bw.put((uint32_t)bits, n);
        int run = 0;
        for (int k = 1; k < 64; ++k) {
            if (!zz[k]) { ++run; continue; }
            while (run > 15) { bw.put(tac[tq].code[0xF0], tac[tq].len[0xF0]); run -= 16; }
            n = magnitude(zz[k], bits);
            const int rs = run << 4 | n;
            bw.put(tac[tq].code[rs], tac[tq].len[rs]);
            bw.put((uint32_t)bits, n);
            run = 0;
        }
        if (run) bw.put(tac[tq].code[0], tac[tq].len[0]);
    };
    float blk[64];
    for (int my0 = 0; my0 < my; ++my0)
        for (int mx0 = 0; mx0 < mx; ++mx0) {
            for (int by = 0; by < f; ++by)
                for (int bx = 0; bx < f; ++bx) {
                    for (int y = 0; y < 8; ++y)
                        for (int x = 0; x < 8; ++x)
                            blk[y * 8 + x] = P[0][(size_t)(my0 * mw + by * 8 + y) * PW + mx0 * mw + bx * 8 + x];
                    encode_block(0, blk);
                }
            for (int k = 1; k < nc; ++k) {   // 2x2 averages
                for (int y = 0; y < 8; ++y)
                    for (int x = 0; x < 8; ++x) {
                        const size_t b = (size_t)(my0 * mw + 2 * y) * PW + mx0 * mw + 2 * x;
                        blk[y * 8 + x] = (P[k][b] + P[k][b + 1] + P[k][b + PW] + P[k][b + PW + 1]) * 0.25f;
                    }
                encode_block(k, blk);
            }
        }
    bw.flush();
    out.push_back(0xFF);
    out.push_back(0xD9);
    return out;
}

Raw decode_any(const std::string &path)
{
    const std::vector<uint8_t> f = read_file(path);
    if (f.size() >= 8 && f[0] == 0x89 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G') return decode_png(f);
    if (f.size() >= 2 && f[0] == 0xFF && f[1] == 0xD8) return decode_jpeg(f);
    if (f.size() >= 2 && f[0] == 'B' && f[1] == 'M') return decode_bmp(f);
    if (f.size() >= 4 && ((f[0] == 'I' && f[1] == 'I') || (f[0] == 'M' && f[1] == 'M'))) return decode_tiff(f);
    if (f.size() >= 2 && f[0] == 'P' && f[1] >= '1' && f[1] <= '6') return decode_pnm(f);
    bad(path + ": unrecognised image format");
}

}  // namespace

bool is_image_name(const std::string &path)
{
    static const char *ext[] = {".jpg", ".bmp", ".jpeg", ".png", ".tif", ".tiff", ".pgm", ".ppm", ".pbm", ".pnm"};
    const std::string e = lower_ext(path);
    for (const char *x : ext)
        if (e == x) return true;
    return false;
}

Image8 imread_color(const std::string &path)
{
    const Raw r = decode_any(path);
    Image8 o;
    o.w = r.w;
    o.h = r.h;
    o.c = 3;
    o.px.resize((size_t)r.w * r.h * 3);
    for (size_t i = 0; i < (size_t)r.w * r.h; ++i) {
        const uint8_t *s = &r.v[i * r.c];
        uint8_t *d = &o.px[i * 3];
        if (r.c <= 2) d[0] = d[1] = d[2] = s[0];
        else { d[0] = s[2]; d[1] = s[1]; d[2] = s[0]; }   // RGB -> BGR
    }
    return o;
}

Image8 imread_gray(const std::string &path)
{
    const Raw r = decode_any(path);
    Image8 o;
    o.w = r.w;
    o.h = r.h;
    o.c = 1;
    o.px.resize((size_t)r.w * r.h);
    for (size_t i = 0; i < (size_t)r.w * r.h; ++i) {
        const uint8_t *s = &r.v[i * r.c];
        if (r.c <= 2) o.px[i] = s[0];
        else if (!r.luma.empty()) o.px[i] = r.luma[i];
        else if (r.png) o.px[i] = (uint8_t)((s[0] * 9798 + s[1] * 19235 + s[2] * 3735 + 16384) >> 15);
        else o.px[i] = (uint8_t)((s[0] * 4899 + s[1] * 9617 + s[2] * 1868 + 8192) >> 14);
    }
    return o;
}

void imwrite(const std::string &path, const Image8 &img)
{
    if (img.w <= 0 || img.h <= 0 || (img.c != 1 && img.c != 3) || img.px.size() < (size_t)img.w * img.h * img.c)
        bad("imwrite: bad image");
    const std::string e = lower_ext(path);
    std::vector<uint8_t> data;
    if (e == ".png") data = encode_png(img);
    else if (e == ".bmp") data = encode_bmp(img);
    else if (e == ".tif" || e == ".tiff") data = encode_tiff(img);
    else if (e == ".jpg" || e == ".jpeg") data = encode_jpeg(img);
    else if (e == ".pgm" || e == ".pbm") data = encode_pnm(img, true);
    else if (e == ".ppm" || e == ".pnm") data = encode_pnm(img, false);
    else bad(path + ": no image writer for this extension");
    std::ofstream f(path, std::ios::binary);
    if (!f) bad("cannot write " + path);
    f.write(reinterpret_cast<const char *>(data.data()), (std::streamsize)data.size());
    if (!f) bad("write failed: " + path);
}

}  // namespace io
}  // namespace LabelingTools

namespace {
thread_local std::string g_image_error;
}

extern "C" int lt_image_read(const char *path, int gray, int *h, int *w, int *c, uint8_t *out, size_t cap)
{
    try {
        const auto img = gray ? LabelingTools::io::imread_gray(path) : LabelingTools::io::imread_color(path);
        *h = img.h;
        *w = img.w;
        *c = img.c;
        if (out) {
            if (cap < img.px.size()) throw std::runtime_error("output buffer too small");
            std::memcpy(out, img.px.data(), img.px.size());
        }
        return 0;
    } catch (const std::exception &e) {
        g_image_error = e.what();
        return -1;
    }
}

extern "C" int lt_image_write(const char *path, const uint8_t *px, int h, int w, int c)
{
    try {
        LabelingTools::io::Image8 img;
        img.h = h;
        img.w = w;
        img.c = c;
        img.px.assign(px, px + (size_t)h * w * c);
        LabelingTools::io::imwrite(path, img);
        return 0;
    } catch (const std::exception &e) {
        g_image_error = e.what();
        return -1;
    }
}

extern "C" const char *lt_image_error(void) { return g_image_error.c_str(); }
