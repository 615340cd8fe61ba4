// Target images from files and PNG output of renders (host code).
//
//   ds_image_load       ImageLoader::load    (src/io/image_io.cpp:23-78)
//   ds_image_save_png   ImageSaver::saveAsPNG (src/io/image_io.cpp:82-93)
//
// Decoding: PNG (bit depth 8 or 16, grey / RGB / palette / grey+alpha / RGBA,
// not interlaced; alpha is dropped) through zlib, JPEG (baseline and
// progressive Huffman, jpeg.cpp) and binary PPM (P6, maxval 255). The
// reference decodes with stb_image (src/io/image_io.cpp:26). Placement follows the reference exactly: scale =
// min((W - 2 pad) / w, (H - 2 pad) / h), fit size = (int)(w * scale) x
// (int)(h * scale), centred at ((W - fit_w) / 2, (H - fit_h) / 2), black
// elsewhere, value / 255. Resampling differs: the reference uses
// stb_image_resize2's default filter; here a shrinking axis averages the
// source pixels each output pixel covers (exact area weights) and a growing
// axis interpolates linearly between pixel centres, both on 8-bit values,
// rounded to nearest.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "../../../include/ds_api.h"
#include "../ds_internal.hpp"
#include "jpeg.hpp"

namespace {

struct Rgb8 {
    int w = 0, h = 0;
    std::vector<uint8_t> px;  // w * h * 3
};

uint32_t be32(const uint8_t *p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

ds_status decode_png(const std::vector<uint8_t> &f, Rgb8 &out) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    DS_REQUIRE(f.size() >= 8 && std::memcmp(f.data(), sig, 8) == 0, "not a PNG file");
    size_t o = 8;
    int w = 0, h = 0, depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte;
    bool seen_end = false;
    while (o + 12 <= f.size()) {
        const uint32_t len = be32(&f[o]);
        DS_REQUIRE(o + 12 + (size_t)len <= f.size(), "truncated PNG chunk");
        const uint8_t *type = &f[o + 4];
        const uint8_t *data = &f[o + 8];
        const uint32_t crc = be32(&f[o + 8 + len]);
        DS_REQUIRE(crc == (uint32_t)crc32(0L, type, len + 4), "PNG chunk CRC mismatch");
        if (!std::memcmp(type, "IHDR", 4)) {
            DS_REQUIRE(len == 13, "bad IHDR");
            w = (int)be32(data);
            h = (int)be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            seen_end = true;
            break;
        }
        o += 12 + (size_t)len;
    }
    DS_REQUIRE(seen_end, "PNG without IEND");
    DS_REQUIRE(w > 0 && h > 0 && w <= 65536 && h <= 65536, "bad PNG size");
    DS_REQUIRE(interlace == 0, "interlaced PNG is not supported");
    int chans = 0;
    switch (ctype) {
        case 0: chans = 1; break;
        case 2: chans = 3; break;
        case 3: chans = 1; break;
        case 4: chans = 2; break;
        case 6: chans = 4; break;
        default: return ds::set_error(DS_ERR_INVALID_ARG, "bad PNG colour type");
    }
    DS_REQUIRE(depth == 8 || (depth == 16 && ctype != 3), "PNG bit depth must be 8 (or 16)");
    DS_REQUIRE(ctype != 3 || plte.size() >= 3, "palette PNG without PLTE");
    const int bpp = chans * depth / 8;  // bytes per pixel
    const size_t stride = (size_t)w * bpp;
    std::vector<uint8_t> raw((stride + 1) * (size_t)h);
    uLongf rawlen = (uLongf)raw.size();
    DS_REQUIRE(uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) == Z_OK &&
                   rawlen == raw.size(),
               "PNG image data does not inflate to its size");
    // undo the per-row filters in place
    std::vector<uint8_t> prev(stride, 0), cur(stride);
    out.w = w;
    out.h = h;
    out.px.assign((size_t)w * h * 3, 0);
    for (int y = 0; y < h; ++y) {
        const uint8_t ft = raw[(stride + 1) * (size_t)y];
        const uint8_t *row = &raw[(stride + 1) * (size_t)y + 1];
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)bpp ? cur[i - bpp] : 0;
            const int b = prev[i];
            const int c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            int v = row[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) / 2; break;
                case 4: v += paeth(a, b, c); break;
                default: return ds::set_error(DS_ERR_INVALID_ARG, "bad PNG filter type");
            }
            cur[i] = (uint8_t)v;
        }
        for (int x = 0; x < w; ++x) {
            // most significant byte of each sample
            auto sample = [&](int ch) { return cur[(size_t)x * bpp + (size_t)ch * (depth / 8)]; };
            uint8_t *d = &out.px[((size_t)y * w + x) * 3];
            if (ctype == 0 || ctype == 4) {
                d[0] = d[1] = d[2] = sample(0);
            } else if (ctype == 3) {
                const size_t k = (size_t)sample(0) * 3;
                DS_REQUIRE(k + 2 < plte.size(), "PNG palette index out of range");
                d[0] = plte[k];
                d[1] = plte[k + 1];
                d[2] = plte[k + 2];
            } else {
                d[0] = sample(0);
                d[1] = sample(1);
                d[2] = sample(2);
            }
        }
        std::swap(prev, cur);
    }
    return DS_OK;
}

ds_status decode_ppm(const std::vector<uint8_t> &f, Rgb8 &out) {
    size_t o = 2;
    int vals[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        while (o < f.size()) {  // whitespace and comments
            if (f[o] == '#') {
                while (o < f.size() && f[o] != '\n') ++o;
            } else if (std::isspace(f[o])) {
                ++o;
            } else {
                break;
            }
        }
        DS_REQUIRE(o < f.size() && std::isdigit(f[o]), "bad PPM header");
        long v = 0;
        while (o < f.size() && std::isdigit(f[o])) v = v * 10 + (f[o++] - '0');
        vals[k] = (int)v;
    }
    ++o;  // one whitespace byte before the data
    DS_REQUIRE(vals[2] == 255, "PPM maxval must be 255");
    DS_REQUIRE(vals[0] > 0 && vals[1] > 0, "bad PPM size");
    const size_t n = (size_t)vals[0] * vals[1] * 3;
    DS_REQUIRE(o + n <= f.size(), "PPM data ends early");
    out.w = vals[0];
    out.h = vals[1];
    out.px.assign(f.begin() + (long)o, f.begin() + (long)(o + n));
    return DS_OK;
}

ds_status decode_file(const char *path, Rgb8 &img) {
    DS_REQUIRE(path, "NULL path");
    std::ifstream s(path, std::ios::binary);
    DS_REQUIRE(s.is_open(), std::string("cannot open ") + path);
    std::vector<uint8_t> f((std::istreambuf_iterator<char>(s)), std::istreambuf_iterator<char>());
    if (f.size() >= 2 && f[0] == 'P' && f[1] == '6') return decode_ppm(f, img);
    if (f.size() >= 8 && f[0] == 137 && f[1] == 'P') return decode_png(f, img);
    if (f.size() >= 2 && f[0] == 0xFF && f[1] == 0xD8) {
        dsh::JpegImage j;
        std::string err;
        if (!dsh::decode_jpeg(f.data(), f.size(), j, err))
            return ds::set_error(DS_ERR_INVALID_ARG, std::string(path) + ": " + err);
        img.w = j.w;
        img.h = j.h;
        img.px = std::move(j.rgb);
        return DS_OK;
    }
    return ds::set_error(DS_ERR_INVALID_ARG,
                         std::string("unsupported image format (PNG, JPEG or binary PPM): ") +
                             path);
}

// Weights of the source samples for each output sample along one axis: the
// cubic filters of the reference's resize (stbir_resize_uint8_linear,
// src/io/image_io.cpp:44, whose defaults are Catmull-Rom when an axis grows
// and Mitchell-Netravali B = C = 1/3 when it shrinks, with clamped edges),
// evaluated between pixel centres; a shrinking axis stretches the filter by
// the reduction factor. Weights are normalised to sum to 1.
struct Tap {
    int i;
    float w;
};
double cubic(double x, double B, double C) {
    x = std::fabs(x);
    if (x < 1.0)
        return ((12 - 9 * B - 6 * C) * x * x * x + (-18 + 12 * B + 6 * C) * x * x + (6 - 2 * B)) /
               6.0;
    if (x < 2.0)
        return ((-B - 6 * C) * x * x * x + (6 * B + 30 * C) * x * x + (-12 * B - 48 * C) * x +
                (8 * B + 24 * C)) /
               6.0;
    return 0.0;
}
std::vector<std::vector<Tap>> axis_weights(int src, int dst) {
    std::vector<std::vector<Tap>> out((size_t)dst);
    const double s = (double)dst / src;
    const bool up = s >= 1.0;
    const double B = up ? 0.0 : 1.0 / 3.0, Cc = up ? 0.5 : 1.0 / 3.0;
    const double support = up ? 2.0 : 2.0 / s;
    for (int o = 0; o < dst; ++o) {
        const double centre = (o + 0.5) / s;  // in source pixel units
        std::vector<double> w((size_t)src, 0.0);
        const int i0 = (int)std::floor(centre - support - 0.5), i1 = (int)std::ceil(centre + support);
        double total = 0.0;
        for (int i = i0; i <= i1; ++i) {
            const double x = (i + 0.5) - centre;
            const double k = up ? cubic(x, B, Cc) : cubic(x * s, B, Cc) * s;
            if (k == 0.0) continue;
            w[(size_t)std::min(std::max(i, 0), src - 1)] += k;
            total += k;
        }
        auto &t = out[(size_t)o];
        for (int i = 0; i < src; ++i)
            if (w[(size_t)i] != 0.0) t.push_back({i, (float)(w[(size_t)i] / total)});
    }
    return out;
}

Rgb8 resize(const Rgb8 &in, int w, int h) {
    if (w == in.w && h == in.h) return in;
    const auto wx = axis_weights(in.w, w), wy = axis_weights(in.h, h);
    std::vector<float> tmp((size_t)in.h * w * 3, 0.f);  // horizontal pass, values / 255
    for (int y = 0; y < in.h; ++y)
        for (int x = 0; x < w; ++x)
            for (const Tap &t : wx[(size_t)x])
                for (int c = 0; c < 3; ++c)
                    tmp[((size_t)y * w + x) * 3 + c] +=
                        t.w * (in.px[((size_t)y * in.w + t.i) * 3 + c] * (1.f / 255.f));
    Rgb8 out;
    out.w = w;
    out.h = h;
    out.px.assign((size_t)w * h * 3, 0);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) {
                float v = 0.f;
                for (const Tap &t : wy[(size_t)y]) v += t.w * tmp[((size_t)t.i * w + x) * 3 + c];
                out.px[((size_t)y * w + x) * 3 + c] =
                    (uint8_t)std::min(255.f, std::max(0.f, std::floor(v * 255.f + 0.5f)));
            }
    return out;
}

void put32(std::vector<uint8_t> &v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}

void chunk(std::vector<uint8_t> &png, const char *type, const std::vector<uint8_t> &data) {
    put32(png, (uint32_t)data.size());
    const size_t start = png.size();
    png.insert(png.end(), type, type + 4);
    png.insert(png.end(), data.begin(), data.end());
    put32(png, (uint32_t)crc32(0L, &png[start], (uInt)(data.size() + 4)));
}

}  // namespace

extern "C" {

ds_status ds_image_info(const char *path, int32_t *width, int32_t *height) {
    ds::clear_error();
    Rgb8 img;
    DS_TRY(decode_file(path, img));
    if (width) *width = img.w;
    if (height) *height = img.h;
    return DS_OK;
}

ds_status ds_image_load(const char *path, int target_w, int target_h, int padding,
                        float *host_rgb, int32_t *src_w, int32_t *src_h) {
    ds::clear_error();
    DS_REQUIRE(host_rgb, "NULL output");
    DS_REQUIRE(target_w > 0 && target_h > 0 && padding >= 0 && 2 * padding < target_w &&
                   2 * padding < target_h,
               "bad canvas size or padding");
    Rgb8 img;
    DS_TRY(decode_file(path, img));
    const float scale = std::min((float)(target_w - 2 * padding) / img.w,
                                 (float)(target_h - 2 * padding) / img.h);
    const int fit_w = (int)(img.w * scale), fit_h = (int)(img.h * scale);
    std::fill(host_rgb, host_rgb + (size_t)target_w * target_h * 3, 0.f);
    if (fit_w > 0 && fit_h > 0) {
        const Rgb8 r = resize(img, fit_w, fit_h);
        const int ox = (target_w - fit_w) / 2, oy = (target_h - fit_h) / 2;
        for (int y = 0; y < fit_h; ++y)
            for (int x = 0; x < fit_w; ++x)
                for (int c = 0; c < 3; ++c)
                    host_rgb[((size_t)(y + oy) * target_w + (x + ox)) * 3 + c] =
                        r.px[((size_t)y * fit_w + x) * 3 + c] / 255.f;
    }
    if (src_w) *src_w = img.w;
    if (src_h) *src_h = img.h;
    return DS_OK;
}

ds_status ds_image_save_png(const char *path, const float *host_rgb, int width, int height) {
    ds::clear_error();
    DS_REQUIRE(path && host_rgb, "NULL argument");
    DS_REQUIRE(width > 0 && height > 0, "bad image size");
    const size_t stride = (size_t)width * 3;
    std::vector<uint8_t> raw((stride + 1) * (size_t)height);
    for (int y = 0; y < height; ++y) {
        raw[(stride + 1) * (size_t)y] = 0;  // filter: none
        for (size_t i = 0; i < stride; ++i) {
            const float v = host_rgb[(size_t)y * stride + i] * 255.f;
            raw[(stride + 1) * (size_t)y + 1 + i] = (uint8_t)std::fmin(std::fmax(v, 0.f), 255.f);
        }
    }
    uLongf clen = compressBound((uLong)raw.size());
    std::vector<uint8_t> comp(clen);
    DS_REQUIRE(compress2(comp.data(), &clen, raw.data(), (uLong)raw.size(), 6) == Z_OK,
               "zlib compression failed");
    comp.resize(clen);
    std::vector<uint8_t> png = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)width);
    put32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, no filter/interlace
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", comp);
    chunk(png, "IEND", {});
    std::ofstream s(path, std::ios::binary);
    DS_REQUIRE(s.is_open(), std::string("cannot open for writing: ") + path);
    s.write(reinterpret_cast<const char *>(png.data()), (std::streamsize)png.size());
    DS_REQUIRE(s.good(), std::string("write failed: ") + path);
    return DS_OK;
}

ds_status ds_image_save_png_device(const char *path, const float *dev_rgb, int width,
                                   int height) {
    ds::clear_error();
    DS_REQUIRE(path && dev_rgb, "NULL argument");
    DS_REQUIRE(width > 0 && height > 0, "bad image size");
    std::vector<float> host((size_t)width * height * 3);
    DS_TRY_HIP(hipDeviceSynchronize());
    DS_TRY_HIP(hipMemcpy(host.data(), dev_rgb, sizeof(float) * host.size(),
                         hipMemcpyDeviceToHost));
    return ds_image_save_png(path, host.data(), width, height);
}

}  // extern "C"
