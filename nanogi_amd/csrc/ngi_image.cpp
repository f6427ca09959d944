// Film output: replaces SaveImage (include/nanogi/basic.hpp:506-672).
//
// The film is row-major with row 0 at the bottom (raster y grows upwards,
// rt.hpp:1348-1378).  Every writer stores the image upright, as the
// reference's FreeImage (bottom-up scanlines) and tinyexr (explicit y flip,
// basic.hpp:590-597) paths do:
//   .hdr  Radiance RGBE, "-Y H +X W", run-length encoded scanlines
//   .exr  OpenEXR scanline image, FLOAT channels B, G, R, ZIP compression
//         (blocks of 16 scanlines, as tinyexr writes them)
//   .png  8-bit RGB, v -> clamp(int(v^(1/2.2) * 255), 0, 255) (basic.hpp:627-641),
//         zlib-deflated scanlines
//   .pfm  portable float map (bottom-up rows natively)
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ngi_internal.h"

namespace ngi {
namespace {

bool make_parent_dirs(const std::string& path) {
    size_t pos = 0;
    while ((pos = path.find('/', pos + 1)) != std::string::npos) {
        const std::string dir = path.substr(0, pos);
        if (dir.empty()) continue;
        struct stat st;
        if (stat(dir.c_str(), &st) != 0 && mkdir(dir.c_str(), 0755) != 0) return false;
    }
    return true;
}

void put_u32be(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        init = true;
    }
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c;
}

// zlib stream (RFC 1950) of `in`.
bool deflate_zlib(const std::vector<uint8_t>& in, std::vector<uint8_t>& out) {
    uLongf n = compressBound((uLong)in.size());
    out.resize(n);
    if (compress2(out.data(), &n, in.data(), (uLong)in.size(), 6) != Z_OK) return false;
    out.resize(n);
    return true;
}

int write_file(const std::string& path, const std::vector<uint8_t>& bytes) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return fail(NGI_ERR_IO, "Failed to save image : %s", path.c_str());
    const size_t w = fwrite(bytes.data(), 1, bytes.size(), f);
    const bool ok = fclose(f) == 0 && w == bytes.size();
    return ok ? NGI_OK : fail(NGI_ERR_IO, "Failed to save image : %s", path.c_str());
}

void rgbe(double r, double g, double b, uint8_t out[4]) {
    const double m = std::max(r, std::max(g, b));
    if (!(m > 1e-38)) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    int e;
    const double f = std::frexp(m, &e) * 256.0 / m;
    out[0] = (uint8_t)std::max(0.0, r * f);
    out[1] = (uint8_t)std::max(0.0, g * f);
    out[2] = (uint8_t)std::max(0.0, b * f);
    out[3] = (uint8_t)(e + 128);
}

int save_hdr(const std::string& path, const double* film, int w, int h) {
    std::vector<uint8_t> o;
    char head[128];
    const int n = snprintf(head, sizeof(head), "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w);
    o.insert(o.end(), head, head + n);
    std::vector<uint8_t> line(4 * (size_t)w);
    for (int y = h - 1; y >= 0; --y) {
        for (int x = 0; x < w; ++x) {
            const double* p = film + 3 * ((size_t)y * w + x);
            // Film values are rounded to float first, as FIT_RGBF stores them.
            rgbe((float)p[0], (float)p[1], (float)p[2], &line[4 * (size_t)x]);
        }
        if (w < 8 || w > 0x7FFF) {
            o.insert(o.end(), line.begin(), line.end());
            continue;
        }
        // New-style RLE: per-component runs.
        o.push_back(2), o.push_back(2), o.push_back((uint8_t)(w >> 8)), o.push_back((uint8_t)(w & 0xFF));
        for (int c = 0; c < 4; ++c) {
            int x = 0;
            while (x < w) {
                int run = 1;
                while (x + run < w && run < 127 && line[4 * (size_t)(x + run) + c] == line[4 * (size_t)x + c]) ++run;
                if (run >= 4) {
                    o.push_back((uint8_t)(128 + run));
                    o.push_back(line[4 * (size_t)x + c]);
                    x += run;
                    continue;
                }
                // Literal segment up to the next run of >= 4 (max 128).
                int len = 0;
                while (x + len < w && len < 128) {
                    int r = 1;
                    while (x + len + r < w && r < 4 && line[4 * (size_t)(x + len + r) + c] == line[4 * (size_t)(x + len) + c]) ++r;
                    if (r >= 4) break;
                    ++len;
                }
                o.push_back((uint8_t)len);
                for (int k = 0; k < len; ++k) o.push_back(line[4 * (size_t)(x + k) + c]);
                x += len;
            }
        }
    }
    return write_file(path, o);
}

void put_u32le(std::vector<uint8_t>& o, uint32_t v) {
    for (int k = 0; k < 4; ++k) o.push_back((uint8_t)(v >> (8 * k)));
}

void put_str(std::vector<uint8_t>& o, const char* s) { o.insert(o.end(), s, s + strlen(s) + 1); }

void put_attr(std::vector<uint8_t>& o, const char* name, const char* type, const std::vector<uint8_t>& val) {
    put_str(o, name);
    put_str(o, type);
    put_u32le(o, (uint32_t)val.size());
    o.insert(o.end(), val.begin(), val.end());
}

int save_exr(const std::string& path, const double* film, int w, int h) {
    std::vector<uint8_t> o;
    put_u32le(o, 20000630u);   // magic
    put_u32le(o, 2u);          // version 2, single-part scanline
    std::vector<uint8_t> ch;
    for (const char* n : {"B", "G", "R"}) {
        put_str(ch, n);
        put_u32le(ch, 2);   // FLOAT
        ch.push_back(0), ch.push_back(0), ch.push_back(0), ch.push_back(0);   // pLinear + reserved
        put_u32le(ch, 1), put_u32le(ch, 1);                                   // sampling
    }
    ch.push_back(0);
    put_attr(o, "channels", "chlist", ch);
    put_attr(o, "compression", "compression", {3});   // ZIP: 16 scanlines per block
    std::vector<uint8_t> box;
    put_u32le(box, 0), put_u32le(box, 0), put_u32le(box, (uint32_t)(w - 1)), put_u32le(box, (uint32_t)(h - 1));
    put_attr(o, "dataWindow", "box2i", box);
    put_attr(o, "displayWindow", "box2i", box);
    put_attr(o, "lineOrder", "lineOrder", {0});
    std::vector<uint8_t> par;
    const float one = 1.0f;
    uint32_t bits;
    memcpy(&bits, &one, 4);
    put_u32le(par, bits);
    put_attr(o, "pixelAspectRatio", "float", par);
    std::vector<uint8_t> ctr(8, 0);
    put_attr(o, "screenWindowCenter", "v2f", ctr);
    put_attr(o, "screenWindowWidth", "float", par);
    o.push_back(0);   // end of header
    // Offset table, filled in as the blocks are written.
    const int blocks = (h + 15) / 16;
    const size_t table = o.size();
    o.resize(table + 8 * (size_t)blocks, 0);
    std::vector<uint8_t> raw, tmp, z;
    for (int b = 0; b < blocks; ++b) {
        const int y0 = 16 * b, y1 = std::min(h, y0 + 16);
        raw.clear();
        for (int y = y0; y < y1; ++y) {
            const int fy = h - 1 - y;   // top row first
            for (int c : {2, 1, 0}) {   // B, G, R
                for (int x = 0; x < w; ++x) {
                    const float v = (float)film[3 * ((size_t)fy * w + x) + c];
                    memcpy(&bits, &v, 4);
                    put_u32le(raw, bits);
                }
            }
        }
        // OpenEXR's ZIP filter: even-indexed bytes then odd-indexed bytes,
        // then each byte replaced by its difference to the previous one + 128.
        const size_t n = raw.size(), half = (n + 1) / 2;
        tmp.resize(n);
        for (size_t i = 0; i < n; ++i) tmp[(i & 1) ? half + i / 2 : i / 2] = raw[i];
        for (size_t i = n; i-- > 1;) tmp[i] = (uint8_t)(tmp[i] - tmp[i - 1] + 128);
        if (!deflate_zlib(tmp, z)) return fail(NGI_ERR_IO, "Failed to save image : %s (zlib)", path.c_str());
        const std::vector<uint8_t>& data = z.size() < n ? z : raw;   // readers take raw data when it is not smaller
        const uint64_t off = o.size();
        for (int k = 0; k < 8; ++k) o[table + 8 * (size_t)b + k] = (uint8_t)(off >> (8 * k));
        put_u32le(o, (uint32_t)y0);
        put_u32le(o, (uint32_t)data.size());
        o.insert(o.end(), data.begin(), data.end());
    }
    return write_file(path, o);
}

int save_png(const std::string& path, const double* film, int w, int h) {
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * (3 * (size_t)w + 1));
    const double ex = 1.0 / 2.2;
    for (int y = h - 1; y >= 0; --y) {
        raw.push_back(0);   // filter: none
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) {
                const double v = (double)(float)film[3 * ((size_t)y * w + x) + c];
                int q = (int)(std::pow(v, ex) * 255.0);
                q = q < 0 ? 0 : (q > 255 ? 255 : q);
                raw.push_back((uint8_t)q);
            }
    }
    std::vector<uint8_t> z;
    if (!deflate_zlib(raw, z)) return fail(NGI_ERR_IO, "Failed to save image : %s (zlib)", path.c_str());
    std::vector<uint8_t> o = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    auto chunk = [&](const char* type, const std::vector<uint8_t>& data) {
        put_u32be(o, (uint32_t)data.size());
        const size_t start = o.size();
        o.insert(o.end(), type, type + 4);
        o.insert(o.end(), data.begin(), data.end());
        put_u32be(o, crc32(&o[start], o.size() - start) ^ 0xFFFFFFFFu);
    };
    std::vector<uint8_t> ihdr;
    put_u32be(ihdr, (uint32_t)w);
    put_u32be(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
    return write_file(path, o);
}

int save_pfm(const std::string& path, const double* film, int w, int h) {
    std::vector<uint8_t> o;
    char head[64];
    const int n = snprintf(head, sizeof(head), "PF\n%d %d\n-1.0\n", w, h);
    o.insert(o.end(), head, head + n);
    for (size_t i = 0; i < 3 * (size_t)w * h; ++i) {
        const float v = (float)film[i];
        uint32_t bits;
        memcpy(&bits, &v, 4);
        put_u32le(o, bits);
    }
    return write_file(path, o);
}

}  // namespace
}  // namespace ngi

using namespace ngi;

extern "C" int ngi_image_save(const char* path, const double* film, int width, int height) {
    clear_error();
    if (!path || !film || width <= 0 || height <= 0) return fail(NGI_ERR_INVALID_ARG, "bad image arguments");
    const std::string p(path);
    if (!make_parent_dirs(p)) return fail(NGI_ERR_IO, "Failed to create output directory for %s", path);
    const size_t dot = p.find_last_of('.');
    const std::string ext = dot == std::string::npos ? std::string() : p.substr(dot);
    if (ext == ".hdr") return save_hdr(p, film, width, height);
    if (ext == ".exr") return save_exr(p, film, width, height);
    if (ext == ".png") return save_png(p, film, width, height);
    if (ext == ".pfm") return save_pfm(p, film, width, height);
    return fail(NGI_ERR_INVALID_ARG, "Invalid extension: %s", ext.c_str());
}
