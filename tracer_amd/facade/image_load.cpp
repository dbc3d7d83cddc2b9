// Image textures behind create_texture2d_from_file (src/utility/image_file.cpp:168-180):
// the Radiance RGBE (.hdr) reader of this build, and the texture class that
// create_ibl_light turns into an image environment.
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
//   * a pixel (r, g, b, e) is (r, g, b) * 2^(e - 136), and 0 when e = 0.
// Only .hdr files are read: the GPU path has no 8-bit image textures.
#include <cmath>
#include <cstdio>
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
    throw std::runtime_error("read_hdr: " + path + ": " + what);
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

// Texture2D::evaluate with wrap_repeat (core/texture.hpp:57-83) over
// LinearSampler::Sample2D (utility/sampler.hpp:12-25).
Spectrum ImageTexture2D::evaluate(const Point2f& uv) const {
    double u = std::fmin(std::fmax(uv.x - std::floor(uv.x), 0.0), 1.0) * (w - 1);
    double v = std::fmin(std::fmax(uv.y - std::floor(uv.y), 0.0), 1.0) * (h - 1);
    const int u0 = std::min(int(u), w - 1), u1 = std::min(u0 + 1, w - 1);
    const int v0 = std::min(int(v), h - 1), v1 = std::min(v0 + 1, h - 1);
    const double du = u - u0, dv = v - v0;
    float out[3];
    for (int k = 0; k < 3; ++k) {
        auto at = [&](int x, int y) { return double(rgb[(size_t(y) * size_t(w) + size_t(x)) * 3 + size_t(k)]); };
        out[k] = float((at(u0, v0) * (1.0 - du) + at(u1, v0) * du) * (1.0 - dv) +
                       (at(u0, v1) * (1.0 - du) + at(u1, v1) * du) * dv);
    }
    return Spectrum(out[0], out[1], out[2]);
}

}  // namespace detail

RC<const Texture2D> create_texture2d_from_file(const std::string& filename) {
    const size_t dot = filename.find_last_of('.');
    const std::string ext = dot == std::string::npos ? std::string() : filename.substr(dot);
    if (ext != ".hdr")
        throw std::runtime_error("create_texture2d_from_file: only Radiance .hdr images are supported by the GPU path (" +
                                 filename + ")");
    auto t = newRC<detail::ImageTexture2D>();
    detail::read_hdr(filename, t->w, t->h, t->rgb);
    return t;
}

}  // namespace tracer

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
