// film_io.cpp — image writers for resolved films: 8-bit sRGB PNG and float OpenEXR.
//
// The reference writes its film through OpenImageIO (Film::writePNG, src/film/Film.cpp:233-276;
// Film::writeEXR, Film.cpp:158-228), which is not available here.  These writers take the
// resolved buffers ahip_resolve produces and write the same images:
//   - PNG: 8-bit RGB, written to "<path>.writing.png" and renamed into place so a viewer never
//     reads a partial file (Film.cpp:261-275); zlib deflate, no row filters;
//   - EXR: one scanline image with float channels named "<prefix>R/G/B" per layer (the beauty
//     with the empty prefix, AOV layers as "albedo.", "normals." ...), channels in the sorted
//     order the format requires.  Stored without compression (the reference asks OIIO for ZIP; the
//     pixel values are the same).
#include "../../include/anacapa_hip.h"

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

void be32(std::vector<uint8_t>& b, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
}
template <typename T>
void le(std::vector<uint8_t>& b, T v) {
    uint8_t raw[sizeof(T)];
    std::memcpy(raw, &v, sizeof(T));  // the host is little-endian (x86-64)
    b.insert(b.end(), raw, raw + sizeof(T));
}
void cstr(std::vector<uint8_t>& b, const std::string& s) {
    b.insert(b.end(), s.begin(), s.end());
    b.push_back(0);
}

void pngChunk(std::vector<uint8_t>& out, const char type[4], const std::vector<uint8_t>& data) {
    be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    uLong crc = crc32(0L, Z_NULL, 0);
    crc = crc32(crc, out.data() + start, (uInt)(out.size() - start));
    be32(out, (uint32_t)crc);
}

bool writeAll(const std::string& path, const std::vector<uint8_t>& bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    return (std::fclose(f) == 0) && ok;
}

void attr(std::vector<uint8_t>& h, const char* name, const char* type, const std::vector<uint8_t>& value) {
    cstr(h, name);
    cstr(h, type);
    le<int32_t>(h, (int32_t)value.size());
    h.insert(h.end(), value.begin(), value.end());
}

}  // namespace

extern "C" {

int ahip_write_png(const char* path, uint32_t width, uint32_t height, const uint8_t* rgb) {
    if (!path || !rgb || width == 0 || height == 0) return AHIP_E_INVALID_ARG;
    std::vector<uint8_t> raw;
    raw.reserve((size_t)height * (1 + 3 * (size_t)width));
    for (uint32_t y = 0; y < height; ++y) {
        raw.push_back(0);  // filter type: none
        const uint8_t* row = rgb + (size_t)y * width * 3;
        raw.insert(raw.end(), row, row + 3 * (size_t)width);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return AHIP_E_OUT_OF_MEMORY;
    z.resize(zlen);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    be32(ihdr, width);
    be32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8 bits, truecolour, deflate, no filter, no interlace
    pngChunk(out, "IHDR", ihdr);
    pngChunk(out, "IDAT", z);
    pngChunk(out, "IEND", {});
    const std::string tmp = std::string(path) + ".writing.png";
    if (!writeAll(tmp, out)) return AHIP_E_INVALID_ARG;
    return std::rename(tmp.c_str(), path) == 0 ? AHIP_OK : AHIP_E_INVALID_ARG;
}

int ahip_write_exr(const char* path, uint32_t width, uint32_t height, uint32_t num_layers,
                   const char* const* layer_prefixes, const float* const* layers) {
    if (!path || !layer_prefixes || !layers || width == 0 || height == 0 || num_layers == 0) return AHIP_E_INVALID_ARG;
    // channels: (name, layer, component), sorted by name as the file format requires
    struct Ch {
        std::string name;
        uint32_t layer, comp;
    };
    std::vector<Ch> ch;
    for (uint32_t l = 0; l < num_layers; ++l) {
        if (!layer_prefixes[l] || !layers[l]) return AHIP_E_INVALID_ARG;
        for (uint32_t c = 0; c < 3; ++c) ch.push_back({std::string(layer_prefixes[l]) + "RGB"[c], l, c});
    }
    std::sort(ch.begin(), ch.end(), [](const Ch& a, const Ch& b) { return a.name < b.name; });
    for (size_t i = 1; i < ch.size(); ++i)
        if (ch[i].name == ch[i - 1].name) return AHIP_E_INVALID_ARG;

    std::vector<uint8_t> f;
    le<int32_t>(f, 20000630);  // magic
    le<int32_t>(f, 2);         // version 2, single-part scanline file, short names
    std::vector<uint8_t> v;
    for (const Ch& c : ch) {
        cstr(v, c.name);
        le<int32_t>(v, 2);  // pixel type FLOAT
        v.insert(v.end(), {0, 0, 0, 0});  // pLinear, reserved
        le<int32_t>(v, 1);  // x sampling
        le<int32_t>(v, 1);  // y sampling
    }
    v.push_back(0);
    attr(f, "channels", "chlist", v);
    attr(f, "compression", "compression", {0});  // NO_COMPRESSION
    v.clear();
    for (int32_t x : {0, 0, (int32_t)width - 1, (int32_t)height - 1}) le<int32_t>(v, x);
    attr(f, "dataWindow", "box2i", v);
    attr(f, "displayWindow", "box2i", v);
    attr(f, "lineOrder", "lineOrder", {0});  // INCREASING_Y
    v.clear();
    le<float>(v, 1.f);
    attr(f, "pixelAspectRatio", "float", v);
    v.clear();
    le<float>(v, 0.f);
    le<float>(v, 0.f);
    attr(f, "screenWindowCenter", "v2f", v);
    v.clear();
    le<float>(v, 1.f);
    attr(f, "screenWindowWidth", "float", v);
    f.push_back(0);  // end of header

    // offset table: one scanline per block without compression
    const size_t rowBytes = ch.size() * (size_t)width * 4;
    const size_t tableAt = f.size();
    f.resize(f.size() + 8 * (size_t)height);
    for (uint32_t y = 0; y < height; ++y) {
        const uint64_t at = f.size();
        std::memcpy(&f[tableAt + 8 * (size_t)y], &at, 8);
        le<int32_t>(f, (int32_t)y);
        le<int32_t>(f, (int32_t)rowBytes);
        for (const Ch& c : ch) {
            const float* src = layers[c.layer] + (size_t)y * width * 3 + c.comp;
            for (uint32_t x = 0; x < width; ++x) le<float>(f, src[3 * (size_t)x]);
        }
    }
    return writeAll(path, f) ? AHIP_OK : AHIP_E_INVALID_ARG;
}

}  // extern "C"
