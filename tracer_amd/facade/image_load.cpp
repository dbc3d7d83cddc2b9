// Image textures behind create_texture2d_from_file and create_texture_from_file
// (src/utility/image_file.cpp:100-180): the Radiance RGBE (.hdr) and PNG
// readers of this build and the texture class that materials and
// create_ibl_light use.
//
// The reference decodes images with stb_image (dep/stb_image.h); this file
// reads the Radiance picture format from its published description:
//   * a "#?RADIANCE" or "#?RGBE" first line, header lines up to an empty line
//     (a FORMAT line must name 32-bit_rle_rgbe), then the resolution line
//     "-Y <height> +X <width>" (rows top to bottom, columns left to right);
//   * per scanline either the adaptive run-length form (bytes 2, 2, then the
//     width as a 15-bit big-endian number, then the R, G, B and E planes, each
//     a sequence of runs: a count byte above 128 repeats the next byte
//     count - 128 times, else that many literal bytes follow) or flat
//     4-byte RGBE pixels;
//   * a pixel (r, g, b, e) is (r, g, b) * 2^(e - 136), and 0 when e = 0;
// and PNG from ISO/IEC 15948 (read_png below).  JPEG and BMP are not read.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "facade_impl.h"
#include "tracer/scenes.h"

namespace tracer {
namespace detail {

namespace {
[[noreturn]] void bad(const std::string& path, const std::string& what) {
    throw std::runtime_error("image " + path + ": " + what);
}
}  // namespace

void read_hdr(const std::string& path, int& width, int& height, std::vector<float>& rgb) {
    std::ifstream in(path, std::ios::binary);
    if (!in) bad(path, "cannot open");
    const std::vector<unsigned char> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    size_t pos = 0;
    auto line = [&]() {
        std::string s;
        while (pos < buf.size() && buf[pos] != '\n') s.push_back(char(buf[pos++]));
        if (pos >= buf.size()) bad(path, "truncated header");
        ++pos;
        return s;
    };
    const std::string magic = line();
    if (magic != "#?RADIANCE" && magic != "#?RGBE") bad(path, "not a Radiance picture");
    for (;;) {
        const std::string s = line();
        if (s.empty()) break;
        if (s.rfind("FORMAT=", 0) == 0 && s != "FORMAT=32-bit_rle_rgbe") bad(path, "unsupported " + s);
    }
    const std::string res = line();
    int h = 0, w = 0;
    char ys[3] = {0}, xs[3] = {0};
    if (std::sscanf(res.c_str(), "%2s %d %2s %d", ys, &h, xs, &w) != 4 || std::strcmp(ys, "-Y") != 0 ||
        std::strcmp(xs, "+X") != 0)
        bad(path, "only the -Y H +X W orientation is supported");
    if (w <= 0 || h <= 0 || int64_t(w) * h > (int64_t(1) << 28)) bad(path, "bad resolution");
    rgb.assign(size_t(w) * size_t(h) * 3, 0.0f);
    std::vector<unsigned char> scan(size_t(w) * 4);
    auto byte = [&]() -> unsigned char {
        if (pos >= buf.size()) bad(path, "truncated pixel data");
        return buf[pos++];
    };
    for (int y = 0; y < h; ++y) {
        const bool rle = w >= 8 && w < 32768 && pos + 4 <= buf.size() && buf[pos] == 2 && buf[pos + 1] == 2 &&
                         (buf[pos + 2] & 0x80) == 0;
        if (rle) {
            pos += 2;
            const int len = (int(buf[pos]) << 8) | int(buf[pos + 1]);
            pos += 2;
            if (len != w) bad(path, "scanline length differs from the width");
            for (int c = 0; c < 4; ++c) {
                int x = 0;
                while (x < w) {
                    int count = byte();
                    if (count > 128) {
                        count -= 128;
                        if (x + count > w) bad(path, "run past the end of a scanline");
                        const unsigned char v = byte();
                        for (int k = 0; k < count; ++k) scan[size_t(x++) * 4 + size_t(c)] = v;
                    } else {
                        if (count == 0 || x + count > w) bad(path, "bad literal count");
                        for (int k = 0; k < count; ++k) scan[size_t(x++) * 4 + size_t(c)] = byte();
                    }
                }
            }
        } else {
            for (size_t i = 0; i < scan.size(); ++i) scan[i] = byte();
        }
        for (int x = 0; x < w; ++x) {
            const unsigned char* p = &scan[size_t(x) * 4];
            float* o = &rgb[(size_t(y) * size_t(w) + size_t(x)) * 3];
            if (p[3] == 0) { o[0] = o[1] = o[2] = 0.0f; continue; }
            const float f = std::ldexp(1.0f, int(p[3]) - 136);
            o[0] = float(p[0]) * f;
            o[1] = float(p[1]) * f;
            o[2] = float(p[2]) * f;
        }
    }
    width = w;
    height = h;
}

// PNG (ISO/IEC 15948): signature, IHDR (8-bit, no interlace), PLTE, the
// concatenated IDAT zlib stream, then per row a filter byte (0 none, 1 sub,
// 2 up, 3 average, 4 Paeth) applied to the bytes of the previous pixel (a),
// the row above (b) and the pixel above-left (c).
void read_png(const std::string& path, int& width, int& height, std::vector<uint8_t>& rgb) {
    std::ifstream in(path, std::ios::binary);
    if (!in) bad(path, "cannot open");
    const std::vector<unsigned char> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (buf.size() < 8 || std::memcmp(buf.data(), sig, 8) != 0) bad(path, "not a PNG file");
    auto be32 = [&](size_t p) {
        return (uint32_t(buf[p]) << 24) | (uint32_t(buf[p + 1]) << 16) | (uint32_t(buf[p + 2]) << 8) | uint32_t(buf[p + 3]);
    };
    size_t pos = 8;
    int w = 0, h = 0, ctype = -1;
    std::vector<unsigned char> idat, plte;
    bool end = false;
    while (!end) {
        if (pos + 12 > buf.size()) bad(path, "truncated chunk");
        const uint32_t n = be32(pos);
        const std::string type(reinterpret_cast<const char*>(&buf[pos + 4]), 4);
        if (pos + 12 + size_t(n) > buf.size()) bad(path, "truncated chunk");
        const unsigned char* d = &buf[pos + 8];
        if (type == "IHDR") {
            if (n != 13) bad(path, "bad IHDR");
            w = int(be32(pos + 8));
            h = int(be32(pos + 12));
            if (d[8] != 8) bad(path, "only 8-bit channels are supported");
            ctype = d[9];
            if (d[12] != 0) bad(path, "interlaced images are not supported");
        } else if (type == "PLTE") {
            plte.assign(d, d + n);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + n);
        } else if (type == "IEND") {
            end = true;
        }
        pos += 12 + size_t(n);
    }
    if (w <= 0 || h <= 0 || int64_t(w) * h > (int64_t(1) << 28)) bad(path, "bad size");
    int ch;
    switch (ctype) {
        case 0: ch = 1; break;
        case 2: ch = 3; break;
        case 3: ch = 1; if (plte.empty()) bad(path, "palette image without PLTE"); break;
        case 6: ch = 4; break;
        case 4: bad(path, "invalid image component (grey + alpha)");
        default: bad(path, "unsupported colour type");
    }
    const size_t stride = size_t(w) * size_t(ch);
    std::vector<unsigned char> raw((stride + 1) * size_t(h));
    uLongf rl = uLongf(raw.size());
    if (uncompress(raw.data(), &rl, idat.data(), uLong(idat.size())) != Z_OK || rl != raw.size())
        bad(path, "corrupt image data");
    std::vector<unsigned char> prev(stride, 0), cur(stride);
    rgb.assign(size_t(w) * size_t(h) * 3, 0);
    for (int y = 0; y < h; ++y) {
        const unsigned char* r = &raw[size_t(y) * (stride + 1)];
        const int f = r[0];
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= size_t(ch) ? cur[i - size_t(ch)] : 0, b = prev[i], c = i >= size_t(ch) ? prev[i - size_t(ch)] : 0;
            int pred = 0;
            switch (f) {
                case 0: pred = 0; break;
                case 1: pred = a; break;
                case 2: pred = b; break;
                case 3: pred = (a + b) / 2; break;
                case 4: {
                    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
                    pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: bad(path, "bad filter type");
            }
            cur[i] = (unsigned char)((r[1 + i] + pred) & 255);
        }
        // flipped vertically: file row y becomes texture row h - 1 - y
        uint8_t* o = &rgb[size_t(h - 1 - y) * size_t(w) * 3];
        for (int x = 0; x < w; ++x) {
            const unsigned char* px = &cur[size_t(x) * size_t(ch)];
            if (ctype == 3) {
                if (size_t(px[0]) * 3 + 2 >= plte.size()) bad(path, "palette index out of range");
                o[3 * x] = plte[px[0] * 3]; o[3 * x + 1] = plte[px[0] * 3 + 1]; o[3 * x + 2] = plte[px[0] * 3 + 2];
            } else if (ch == 1) {
                o[3 * x] = o[3 * x + 1] = o[3 * x + 2] = px[0];
            } else {
                o[3 * x] = px[0]; o[3 * x + 1] = px[1]; o[3 * x + 2] = px[2];
            }
        }
        std::swap(prev, cur);
    }
    width = w;
    height = h;
}

// Texture2D::evaluate with wrap_repeat (core/texture.hpp:57-83) over
// LinearSampler::Sample2D (utility/sampler.hpp:12-25), in floating point.
Spectrum ImageTexture2D::evaluate(const Point2f& uv) const {
    double u = std::fmin(std::fmax(uv.x - std::floor(uv.x), 0.0), 1.0) * (w - 1);
    double v = std::fmin(std::fmax(uv.y - std::floor(uv.y), 0.0), 1.0) * (h - 1);
    const int u0 = std::min(int(u), w - 1), u1 = std::min(u0 + 1, w - 1);
    const int v0 = std::min(int(v), h - 1), v1 = std::min(v0 + 1, h - 1);
    const double du = u - u0, dv = v - v0;
    float out[3];
    for (int k = 0; k < 3; ++k) {
        auto at = [&](int x, int y) {
            const size_t i = (size_t(y) * size_t(w) + size_t(x)) * 3 + size_t(k);
            return is8() ? double(rgb8[i]) / 255.0 : double(rgb[i]);
        };
        out[k] = float((at(u0, v0) * (1.0 - du) + at(u1, v0) * du) * (1.0 - dv) +
                       (at(u0, v1) * (1.0 - du) + at(u1, v1) * du) * dv);
    }
    return Spectrum(out[0], out[1], out[2]);
}

}  // namespace detail

namespace detail {
namespace {
std::string extension(const std::string& name) {
    const size_t dot = name.find_last_of('.');
    return dot == std::string::npos ? std::string() : name.substr(dot);
}
RC<ImageTexture2D> load_image(const std::string& name) {
    auto t = newRC<ImageTexture2D>();
    const std::string ext = extension(name);
    if (ext == ".hdr") read_hdr(name, t->w, t->h, t->rgb);
    else if (ext == ".png") read_png(name, t->w, t->h, t->rgb8);
    else throw std::runtime_error("image " + name + ": only .hdr and .png images are read by this build");
    return t;
}
}  // namespace

// _create_texture_from_file (image_file.cpp:100-127) for a file name: an all-zero
// scale counts as 1; an 8-bit image is decoded to linear by
// uint8(255 (c / 255)^2.2) and then scaled as uint8(c * scale) (both
// truncating, as Point3<uint8_t> stores a float); a float image is scaled.
RC<Texture2D> texture_from_file_scaled(const std::string& name, const real* scale) {
    real s[3] = {scale[0], scale[1], scale[2]};
    if (!s[0] && !s[1] && !s[2]) s[0] = s[1] = s[2] = 1;
    auto t = load_image(name);
    if (t->is8()) {
        for (size_t i = 0; i < t->rgb8.size(); ++i) {
            const float lin = std::pow(float(t->rgb8[i]) / 255.0f, 2.2f);
            const uint8_t q = uint8_t(lin * 255.0f);
            t->rgb8[i] = uint8_t(float(q) * s[i % 3]);
        }
    } else {
        for (size_t i = 0; i < t->rgb.size(); ++i) t->rgb[i] *= s[i % 3];
    }
    return t;
}
}  // namespace detail

// image_file.cpp:168-180: .hdr as float texels, .png as 8-bit texels
// (value / 255, no decoding to linear).
RC<const Texture2D> create_texture2d_from_file(const std::string& filename) {
    return detail::load_image(filename);
}

}  // namespace tracer

extern "C" int32_t trs_read_png(const char* path, int32_t* width, int32_t* height, uint8_t* rgb, int64_t cap,
                                const char** err) {
    static thread_local std::string msg;
    try {
        if (!path || !width || !height) throw std::runtime_error("trs_read_png: NULL argument");
        int w = 0, h = 0;
        std::vector<uint8_t> data;
        tracer::detail::read_png(path, w, h, data);
        *width = w;
        *height = h;
        if (rgb) {
            if (cap < int64_t(data.size())) throw std::runtime_error("trs_read_png: output too small");
            std::memcpy(rgb, data.data(), data.size());
        }
        return 0;
    } catch (const std::exception& e) {
        msg = e.what();
        if (err) *err = msg.c_str();
        return 1;
    }
}

extern "C" int32_t trs_read_hdr(const char* path, int32_t* width, int32_t* height, float* rgb, int64_t cap,
                                const char** err) {
    static thread_local std::string msg;
    try {
        if (!path || !width || !height) throw std::runtime_error("trs_read_hdr: NULL argument");
        int w = 0, h = 0;
        std::vector<float> data;
        tracer::detail::read_hdr(path, w, h, data);
        *width = w;
        *height = h;
        if (rgb) {
            if (cap < int64_t(data.size())) throw std::runtime_error("trs_read_hdr: output too small");
            std::memcpy(rgb, data.data(), data.size() * sizeof(float));
        }
        return 0;
    } catch (const std::exception& e) {
        msg = e.what();
        if (err) *err = msg.c_str();
        return 1;
    }
}
