// Image outputs and post processing behind the reference's
// write_image_to_hdr / write_image_to_png (src/utility/image_file.cpp:14-24)
// and create_gamma_corrector / create_aces_tone_mapper
// (src/post_processor/gamma.cpp, src/post_processor/aces.cpp).
//
// The reference writes with stb_image_write (dep/stb_image_write.h).  Here:
//   * .hdr: Radiance RGBE, header "#?RADIANCE", "FORMAT=32-bit_rle_rgbe", the
//     resolution line "-Y H +X W", then one adaptive run-length scanline per
//     row (runs of 3 to 127 equal bytes, literals of up to 128 bytes), each
//     pixel (m r, m g, m b, e + 128) with v = max(r, g, b) = f 2^e,
//     f in [0.5, 1), m = 256 f / v (zero below 1e-32);
//   * .png: 8-bit RGB, one IDAT chunk with every row under filter 0,
//     compressed by zlib, chunk CRCs by zlib's crc32.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "facade_impl.h"

namespace tracer {

namespace {
[[noreturn]] void fail(const std::string& what) { throw std::runtime_error(what); }

// One RGBE pixel.  Channels are first clamped into what the format holds:
// negative and NaN to 0, anything above 255/256 * 2^127 (infinity included)
// to that largest value, whose exponent byte is 255.
void put_rgbe(float r, float g, float b, unsigned char* o) {
    const float kMax = std::ldexp(255.0f / 256.0f, 127);
    auto clamp = [&](float c) { return c > 0.0f ? std::min(c, kMax) : 0.0f; };   // NaN > 0 is false
    r = clamp(r);
    g = clamp(g);
    b = clamp(b);
    const float v = std::max(r, std::max(g, b));
    if (!(v >= 1e-32f)) {
        o[0] = o[1] = o[2] = o[3] = 0;
        return;
    }
    int e;
    const float f = std::frexp(v, &e);   // v = f 2^e, f in [0.5, 1)
    const float m = f * 256.0f / v;
    o[0] = (unsigned char)(std::max(0.0f, r) * m);
    o[1] = (unsigned char)(std::max(0.0f, g) * m);
    o[2] = (unsigned char)(std::max(0.0f, b) * m);
    o[3] = (unsigned char)(e + 128);
}

// One plane of a scanline in the adaptive run-length form.
void put_plane(const unsigned char* p, int n, std::vector<unsigned char>& out) {
    int i = 0;
    while (i < n) {
        int run = 1;
        while (i + run < n && run < 127 && p[i + run] == p[i]) ++run;
        if (run >= 3) {
            out.push_back((unsigned char)(128 + run));
            out.push_back(p[i]);
            i += run;
            continue;
        }
        // literals up to the next run of at least 3
        int j = i;
        while (j < n && j - i < 128) {
            if (j + 2 < n && p[j] == p[j + 1] && p[j] == p[j + 2]) break;
            ++j;
        }
        out.push_back((unsigned char)(j - i));
        out.insert(out.end(), p + i, p + j);
        i = j;
    }
}

void put_be32(std::vector<unsigned char>& v, uint32_t x) {
    v.push_back((unsigned char)(x >> 24)); v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8)); v.push_back((unsigned char)x);
}

void png_chunk(std::ofstream& out, const char type[4], const std::vector<unsigned char>& data) {
    std::vector<unsigned char> buf;
    put_be32(buf, uint32_t(data.size()));
    buf.insert(buf.end(), type, type + 4);
    buf.insert(buf.end(), data.begin(), data.end());
    const uLong crc = crc32(crc32(0L, Z_NULL, 0), buf.data() + 4, uInt(buf.size() - 4));
    put_be32(buf, uint32_t(crc));
    out.write(reinterpret_cast<const char*>(buf.data()), std::streamsize(buf.size()));
}
}  // namespace

void write_image_to_hdr(const Image2D<Spectrum>& image, const std::string& filename) {
    const int w = image.width(), h = image.height();
    if (w <= 0 || h <= 0) fail("write_image_to_hdr: empty image");
    std::ofstream out(filename, std::ios::binary);
    if (!out) fail("write_image_to_hdr: cannot open " + filename);
    const std::string head = "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(h) + " +X " + std::to_string(w) + "\n";
    out.write(head.data(), std::streamsize(head.size()));
    std::vector<unsigned char> px(static_cast<size_t>(w) * 4), planes(static_cast<size_t>(w)), line;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const Spectrum& s = image.at(x, y);
            put_rgbe(s.r, s.g, s.b, &px[size_t(x) * 4]);
        }
        line.clear();
        if (w >= 8 && w < 32768) {
            line.push_back(2); line.push_back(2);
            line.push_back((unsigned char)(w >> 8)); line.push_back((unsigned char)(w & 255));
            for (int c = 0; c < 4; ++c) {
                for (int x = 0; x < w; ++x) planes[size_t(x)] = px[size_t(x) * 4 + size_t(c)];
                put_plane(planes.data(), w, line);
            }
        } else {
            line = px;
        }
        out.write(reinterpret_cast<const char*>(line.data()), std::streamsize(line.size()));
    }
    if (!out) fail("write_image_to_hdr: write failed for " + filename);
}

void write_image_to_png(const Image2D<Color3b>& image, const std::string& filename) {
    const int w = image.width(), h = image.height();
    if (w <= 0 || h <= 0) fail("write_image_to_png: empty image");
    std::vector<unsigned char> raw(size_t(h) * (size_t(w) * 3 + 1));
    for (int y = 0; y < h; ++y) {
        unsigned char* row = &raw[size_t(y) * (size_t(w) * 3 + 1)];
        row[0] = 0;   // filter type 0
        for (int x = 0; x < w; ++x) {
            const Color3b& c = image.at(x, y);
            row[1 + 3 * x] = c.x; row[2 + 3 * x] = c.y; row[3 + 3 * x] = c.z;
        }
    }
    uLongf zlen = compressBound(uLong(raw.size()));
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK) fail("write_image_to_png: zlib failed");
    z.resize(zlen);
    std::ofstream out(filename, std::ios::binary);
    if (!out) fail("write_image_to_png: cannot open " + filename);
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    out.write(reinterpret_cast<const char*>(sig), 8);
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, uint32_t(w));
    put_be32(ihdr, uint32_t(h));
    const unsigned char rest[5] = {8, 2, 0, 0, 0};   // 8-bit, RGB, deflate, filter method 0, no interlace
    ihdr.insert(ihdr.end(), rest, rest + 5);
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
    if (!out) fail("write_image_to_png: write failed for " + filename);
}

namespace {
// gamma.cpp: every channel to the power gamma (gamma <= 0 resets to 1)
class GammaCorrector : public PostProcessor {
public:
    explicit GammaCorrector(real g) : gamma(g > 0 ? g : 1) {}
    void process(RenderTarget& rt) override {
        for (int y = 0; y < rt.color.height(); ++y)
            for (int x = 0; x < rt.color.width(); ++x) {
                Spectrum& p = rt.color.at(x, y);
                p = Spectrum(std::pow(p.r, gamma), std::pow(p.g, gamma), std::pow(p.b, gamma));
            }
    }
    real gamma;
};
// aces.cpp: the fitted ACES curve x (2.51 x + 0.03) / (x (2.43 x + 0.59) + 0.14),
// clamped to [0, 1], of exposure * channel (exposure < 0 resets to 1)
class ACESToneMapper : public PostProcessor {
public:
    explicit ACESToneMapper(real e) : exposure(e < 0 ? 1 : e) {}
    static real curve(real x) {
        return std::clamp((x * (real(2.51) * x + real(0.03))) / (x * (real(2.43) * x + real(0.59)) + real(0.14)), real(0),
                          real(1));
    }
    void process(RenderTarget& rt) override {
        for (int y = 0; y < rt.color.height(); ++y)
            for (int x = 0; x < rt.color.width(); ++x) {
                Spectrum& p = rt.color.at(x, y);
                p = Spectrum(curve(p.r * exposure), curve(p.g * exposure), curve(p.b * exposure));
            }
    }
    real exposure;
};
}  // namespace

RC<PostProcessor> create_gamma_corrector(real gamma) { return newRC<GammaCorrector>(gamma); }
RC<PostProcessor> create_aces_tone_mapper(real exposure) { return newRC<ACESToneMapper>(exposure); }

}  // namespace tracer

extern "C" int32_t trs_write_hdr(const char* path, int32_t width, int32_t height, const float* rgb, const char** err) {
    static thread_local std::string msg;
    try {
        if (!path || !rgb || width <= 0 || height <= 0) throw std::runtime_error("trs_write_hdr: bad arguments");
        tracer::Image2D<tracer::Spectrum> img(width, height);
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const float* p = rgb + (size_t(y) * size_t(width) + size_t(x)) * 3;
                img.at(x, y) = tracer::Spectrum(p[0], p[1], p[2]);
            }
        tracer::write_image_to_hdr(img, path);
        return 0;
    } catch (const std::exception& e) {
        msg = e.what();
        if (err) *err = msg.c_str();
        return 1;
    }
}

extern "C" int32_t trs_write_png(const char* path, int32_t width, int32_t height, const uint8_t* rgb, const char** err) {
    static thread_local std::string msg;
    try {
        if (!path || !rgb || width <= 0 || height <= 0) throw std::runtime_error("trs_write_png: bad arguments");
        tracer::Image2D<tracer::Color3b> img(width, height);
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const uint8_t* p = rgb + (size_t(y) * size_t(width) + size_t(x)) * 3;
                img.at(x, y) = tracer::Color3b{p[0], p[1], p[2]};
            }
        tracer::write_image_to_png(img, path);
        return 0;
    } catch (const std::exception& e) {
        msg = e.what();
        if (err) *err = msg.c_str();
        return 1;
    }
}

extern "C" int32_t trs_postprocess(float* rgb, int32_t width, int32_t height, float aces_exposure, float gamma) {
    if (!rgb || width <= 0 || height <= 0) return 1;
    tracer::RenderTarget rt{tracer::Image2D<tracer::Spectrum>(width, height)};
    for (int i = 0; i < width * height; ++i) rt.color.get_raw_data()[i] = tracer::Spectrum(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    if (aces_exposure >= 0) tracer::create_aces_tone_mapper(aces_exposure)->process(rt);
    if (gamma > 0) tracer::create_gamma_corrector(gamma)->process(rt);
    for (int i = 0; i < width * height; ++i) {
        const tracer::Spectrum& s = rt.color.get_raw_data()[i];
        rgb[3 * i] = s.r; rgb[3 * i + 1] = s.g; rgb[3 * i + 2] = s.b;
    }
    return 0;
}
