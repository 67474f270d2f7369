// rc_gpu.hpp -- C++17 host API over librcgpu that mirrors rust-compress's Rust API for the JPEG target-quality
// path, so a caller of the reference finds the same names, argument meanings and error behaviour:
//
//   rc_core::image::Image, ColorType, Image::flatten        core/src/image.rs:6-11,104-190
//   rc_core::error::Error (Invalid / Unsupported) + Result   core/src/error.rs:4-22
//   rc_jpeg::encode::{Subsampling, EncodeOptions, CoefComponent, CoefImage, build_coefs}
//                                                            jpeg/src/encode.rs:12-93,426
//   rc_jpeg::decode::reconstruct                             jpeg/src/decode.rs:334
//   rc_core::ssimulacra2::{ssimulacra2, MIN_SIDE}            core/src/ssimulacra2.rs:25,471
//   rc_jpeg::target::{Found, find_quality_with}              jpeg/src/target.rs:11-56
//
// Header-only; link with librcgpu.so. Every device call goes through the C ABI of include/rcgpu.h. Trellis
// quantisation (EncodeOptions::trellis, on by default as in the reference) runs on the host between device passes.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <functional>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rcgpu.h"

namespace rc_gpu {

// core/src/error.rs:4-22, plus the two device-side failures of the C ABI.
struct Error {
    enum Kind { Invalid, Unsupported, Device, Oom } kind;
    std::string message;
    static Error from_status(int status) {
        const char* m = rcg_last_error();
        Kind k = status == RCG_INVALID ? Invalid
                 : status == RCG_UNSUPPORTED ? Unsupported
                 : status == RCG_OOM ? Oom
                                     : Device;
        return Error{k, m ? m : ""};
    }
};

// Result<T>: either a value or an Error, like std::result::Result.
template <class T>
class Result {
   public:
    Result(T v) : v_(std::move(v)) {}
    Result(Error e) : e_(std::move(e)) {}
    bool is_ok() const { return v_.has_value(); }
    bool is_err() const { return !v_.has_value(); }
    const T& value() const {
        if (!v_) throw std::runtime_error("called value() on an Err: " + e_.message);
        return *v_;
    }
    T unwrap_or(T fallback) const { return v_ ? *v_ : fallback; }
    const Error& error() const { return e_; }

   private:
    std::optional<T> v_;
    Error e_{Error::Invalid, ""};
};

// core/src/image.rs:6-11
enum class ColorType { Gray, GrayAlpha, Rgb, Rgba };
inline size_t channels(ColorType c) {
    return c == ColorType::Gray ? 1 : c == ColorType::GrayAlpha ? 2 : c == ColorType::Rgb ? 3 : 4;
}

// core/src/image.rs:104-113 (8-bit only on this path; 16-bit images are brought to 8 bits with to_depth8 first).
struct Image {
    uint32_t width = 0, height = 0;
    ColorType color = ColorType::Rgb;
    uint8_t depth = 8;
    std::vector<uint8_t> data;

    static Result<Image> create(uint32_t w, uint32_t h, ColorType c, uint8_t depth, std::vector<uint8_t> data) {
        if (depth != 8 && depth != 16) return Error{Error::Unsupported, "sample depth " + std::to_string(depth)};
        Image img{w, h, c, depth, std::move(data)};
        if (img.data.size() != size_t(w) * h * channels(c) * (depth / 8))
            return Error{Error::Invalid, "pixel buffer size does not match dimensions"};
        return img;
    }

    // core/src/image.rs:144-157: drop the low byte of 16-bit samples, rounding to nearest.
    Image to_depth8() const {
        if (depth == 8) return *this;
        Image o{width, height, color, 8, {}};
        o.data.resize(data.size() / 2);
        for (size_t i = 0; i < o.data.size(); i++) {
            const uint32_t v = (uint32_t(data[2 * i]) << 8) | data[2 * i + 1];
            o.data[i] = uint8_t((v * 255 + 32767) / 65535);
        }
        return o;
    }

    // core/src/image.rs:160-190: composite onto an opaque background, 8-bit output.
    Image flatten(std::array<uint8_t, 3> bg) const {
        Image src = to_depth8();
        auto blend = [](uint32_t c, uint32_t a, uint32_t b) { return uint8_t((c * a + b * (255 - a) + 127) / 255); };
        if (src.color == ColorType::Gray || src.color == ColorType::Rgb) return src;
        Image o{src.width, src.height, src.color == ColorType::GrayAlpha ? ColorType::Gray : ColorType::Rgb, 8, {}};
        const size_t n = size_t(src.width) * src.height;
        if (src.color == ColorType::GrayAlpha) {
            o.data.resize(n);
            for (size_t i = 0; i < n; i++) o.data[i] = blend(src.data[2 * i], src.data[2 * i + 1], bg[0]);
        } else {
            o.data.resize(n * 3);
            for (size_t i = 0; i < n; i++)
                for (int c = 0; c < 3; c++) o.data[3 * i + c] = blend(src.data[4 * i + c], src.data[4 * i + 3], bg[c]);
        }
        return o;
    }
};

// jpeg/src/encode.rs:12-37
enum class Subsampling { Auto, S444, S422, S420 };

struct EncodeOptions {
    uint8_t quality = 80;
    Subsampling subsampling = Subsampling::Auto;
    std::array<uint8_t, 3> background{255, 255, 255};
    bool trellis = true;
};

// jpeg/src/encode.rs:56-83
struct CoefComponent {
    uint8_t id = 0;
    size_t h = 1, v = 1, tq = 0;
    size_t blocks_w = 0;
    std::vector<int16_t> coefs;
};

struct CoefImage {
    size_t width = 0, height = 0;
    std::vector<CoefComponent> comps;
    std::array<std::optional<std::array<uint8_t, 64>>, 4> qtables;
    std::optional<uint8_t> adobe_transform;
};

// jpeg/src/target.rs:11-16
struct Found {
    uint8_t quality = 0;
    double score = 0.0;
    CoefImage coefs;
};

constexpr uint32_t MIN_SIDE = 8;  // core/src/ssimulacra2.rs:25

// A device context (one HIP stream and workspace on one GPU).
class Device {
   public:
    explicit Device(int device = 0) {
        const int rc = rcg_create(device, &ctx_);
        if (rc != RCG_OK) throw std::runtime_error(std::string("rcg_create: ") + rcg_last_error());
    }
    ~Device() { rcg_destroy(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    rcg_ctx* raw() const { return ctx_; }

   private:
    rcg_ctx* ctx_ = nullptr;
};

inline int subsampling_code(Subsampling s) {
    return s == Subsampling::Auto ? RCG_AUTO : s == Subsampling::S444 ? RCG_S444 : s == Subsampling::S422 ? RCG_S422
                                                                                                       : RCG_S420;
}

// The C ABI's view of a CoefImage (borrows its buffers).
inline Result<rcg_coef_image> to_c(const CoefImage& ci);

// jpeg/src/encode.rs:426. `src` must be Gray or Rgb at 8 bits (callers flatten first, as encode.rs:416-421 does).
// With opts.trellis (the reference's default) and quality < 100, the device computes the plain coefficients and the
// hand-off values and the host runs the reference's trellis pass over them (encode.rs:527-537).
inline Result<CoefImage> build_coefs(Device& dev, const Image& src, const EncodeOptions& opts) {
    const bool trellis = opts.trellis && opts.quality < 100;
    if (src.depth != 8 || (src.color != ColorType::Gray && src.color != ColorType::Rgb))
        return Error{Error::Unsupported, "build_coefs takes flattened 8-bit Gray or Rgb pixels"};
    const uint32_t ch = uint32_t(channels(src.color));
    rcg_layout L;
    int rc = rcg_coef_layout(src.width, src.height, ch, opts.quality, subsampling_code(opts.subsampling), &L);
    if (rc != RCG_OK) return Error::from_status(rc);
    CoefImage ci;
    ci.width = src.width;
    ci.height = src.height;
    int16_t* ptrs[3] = {nullptr, nullptr, nullptr};
    for (uint32_t c = 0; c < L.ncomp; c++) {
        CoefComponent k;
        k.id = uint8_t(c + 1);
        k.h = c == 0 ? L.hmax : 1;
        k.v = c == 0 ? L.vmax : 1;
        k.tq = c == 0 ? 0 : 1;
        k.blocks_w = L.blocks_w[c];
        k.coefs.resize(size_t(L.blocks_w[c]) * L.blocks_h[c] * 64);
        ci.comps.push_back(std::move(k));
    }
    for (uint32_t c = 0; c < L.ncomp; c++) ptrs[c] = ci.comps[c].coefs.data();
    uint8_t qt[128];
    std::vector<std::vector<float>> scaled(trellis ? L.ncomp : 0);
    if (trellis) {
        float* sp[3] = {nullptr, nullptr, nullptr};
        for (uint32_t c = 0; c < L.ncomp; c++) {
            scaled[c].resize(ci.comps[c].coefs.size());
            sp[c] = scaled[c].data();
        }
        rc = rcg_build_scaled(dev.raw(), src.data.data(), src.width, src.height, ch, opts.quality,
                              subsampling_code(opts.subsampling), ptrs, sp, qt);
    } else {
        rc = rcg_build_coefs(dev.raw(), src.data.data(), src.width, src.height, ch, opts.quality,
                             subsampling_code(opts.subsampling), ptrs, qt);
    }
    if (rc != RCG_OK) return Error::from_status(rc);
    std::array<uint8_t, 64> a, b;
    std::copy(qt, qt + 64, a.begin());
    std::copy(qt + 64, qt + 128, b.begin());
    ci.qtables[0] = a;
    if (L.ncomp == 3) ci.qtables[1] = b;
    if (trellis) {
        // The trellis pass reads the plain coefficients (its rate model) and the hand-off values.
        CoefImage out = ci;
        Result<rcg_coef_image> plain = to_c(ci);
        if (plain.is_err()) return plain.error();
        const float* sp[4] = {nullptr, nullptr, nullptr, nullptr};
        int16_t* op[4] = {nullptr, nullptr, nullptr, nullptr};
        for (uint32_t c = 0; c < L.ncomp; c++) {
            sp[c] = scaled[c].data();
            op[c] = out.comps[c].coefs.data();
        }
        rc = rcg_trellis(&plain.value(), sp, op);
        if (rc != RCG_OK) return Error::from_status(rc);
        return out;
    }
    return ci;
}

inline Result<rcg_coef_image> to_c(const CoefImage& ci) {
    rcg_coef_image c{};
    c.width = uint32_t(ci.width);
    c.height = uint32_t(ci.height);
    c.ncomp = uint32_t(ci.comps.size());
    if (c.ncomp < 1 || c.ncomp > 4) return Error{Error::Invalid, "1 to 4 components"};
    for (size_t i = 0; i < ci.comps.size(); i++) {
        const CoefComponent& k = ci.comps[i];
        if (k.blocks_w == 0) return Error{Error::Invalid, "empty component"};
        c.comps[i].id = k.id;
        c.comps[i].h = uint8_t(k.h);
        c.comps[i].v = uint8_t(k.v);
        c.comps[i].tq = uint8_t(k.tq);
        c.comps[i].blocks_w = uint32_t(k.blocks_w);
        c.comps[i].blocks_h = uint32_t(k.coefs.size() / 64 / k.blocks_w);
        c.comps[i].coefs = k.coefs.data();
    }
    for (int s = 0; s < 4; s++) c.qtables[s] = ci.qtables[s] ? ci.qtables[s]->data() : nullptr;
    c.adobe_transform = ci.adobe_transform ? int(*ci.adobe_transform) : -1;
    return c;
}

// Entropy coding on the host (no device): jpeg/src/encode.rs:595 write_jpeg, jpeg/src/progressive.rs:386
// write_progressive, jpeg/src/encode.rs:543 write_best. The reference's functions cannot fail; these return an
// error only for a coefficient image the reference could not have built (a missing table, a short block grid).
inline Result<std::vector<uint8_t>> write_with(const CoefImage& ci, int mode) {
    Result<rcg_coef_image> c = to_c(ci);
    if (c.is_err()) return c.error();
    uint8_t* out = nullptr;
    size_t len = 0;
    const int rc = rcg_write_jpeg(&c.value(), mode, &out, &len);
    if (rc != RCG_OK) return Error::from_status(rc);
    std::vector<uint8_t> bytes(out, out + len);
    rcg_free(out);
    return bytes;
}
inline Result<std::vector<uint8_t>> write_jpeg(const CoefImage& ci) { return write_with(ci, RCG_WRITE_BASELINE); }
inline Result<std::vector<uint8_t>> write_progressive(const CoefImage& ci) {
    return write_with(ci, RCG_WRITE_PROGRESSIVE);
}
inline Result<std::vector<uint8_t>> write_best(const CoefImage& ci) { return write_with(ci, RCG_WRITE_BEST); }

// jpeg/src/encode.rs:111: CoefImage::requantize(quality, trellis), in place (host).
inline Result<bool> requantize(CoefImage& ci, uint8_t quality, bool trellis) {
    Result<rcg_coef_image> c = to_c(ci);
    if (c.is_err()) return c.error();
    std::vector<std::vector<int16_t>> outs(ci.comps.size());
    rcg_requantized r{};
    for (size_t i = 0; i < ci.comps.size(); i++) {
        outs[i].resize(ci.comps[i].coefs.size());
        r.coefs[i] = outs[i].data();
    }
    const int rc = rcg_requantize(&c.value(), quality, trellis ? 1 : 0, &r);
    if (rc != RCG_OK) return Error::from_status(rc);
    for (size_t i = 0; i < ci.comps.size(); i++) {
        ci.comps[i].coefs = std::move(outs[i]);
        ci.comps[i].tq = r.tq[i];
    }
    for (int s = 0; s < 4; s++) {
        if (r.present[s]) {
            std::array<uint8_t, 64> t;
            std::copy(r.qtables[s], r.qtables[s] + 64, t.begin());
            ci.qtables[s] = t;
        } else {
            ci.qtables[s].reset();
        }
    }
    return true;
}

// The JPEG -> JPEG target search of cli/src/pipeline.rs:369-386 on the device (rcg_requant_find_quality).
inline Result<Found> find_jpeg_quality_requant(Device& dev, const CoefImage& ci, double target, bool trellis) {
    Result<rcg_coef_image> c = to_c(ci);
    if (c.is_err()) return c.error();
    std::vector<std::vector<int16_t>> outs(ci.comps.size());
    rcg_requantized r{};
    for (size_t i = 0; i < ci.comps.size(); i++) {
        outs[i].resize(ci.comps[i].coefs.size());
        r.coefs[i] = outs[i].data();
    }
    rcg_found f{};
    const int rc = rcg_requant_find_quality(dev.raw(), &c.value(), target, 100, trellis ? 1 : 0, &f, &r);
    if (rc != RCG_OK) return Error::from_status(rc);
    Found out;
    out.quality = uint8_t(f.quality);
    out.score = f.score;
    out.coefs = ci;
    for (size_t i = 0; i < ci.comps.size(); i++) {
        out.coefs.comps[i].coefs = std::move(outs[i]);
        out.coefs.comps[i].tq = r.tq[i];
    }
    for (int s = 0; s < 4; s++) {
        if (r.present[s]) {
            std::array<uint8_t, 64> t;
            std::copy(r.qtables[s], r.qtables[s] + 64, t.begin());
            out.coefs.qtables[s] = t;
        } else {
            out.coefs.qtables[s].reset();
        }
    }
    return out;
}

// jpeg/src/decode.rs:334
inline Result<Image> reconstruct(Device& dev, const CoefImage& ci) {
    Result<rcg_coef_image> cc = to_c(ci);
    if (cc.is_err()) return cc.error();
    const rcg_coef_image& c = cc.value();
    Image img;
    img.width = c.width;
    img.height = c.height;
    img.color = c.ncomp == 1 ? ColorType::Gray : ColorType::Rgb;
    img.data.resize(size_t(c.width) * c.height * (c.ncomp == 1 ? 1 : 3));
    const int rc = rcg_reconstruct(dev.raw(), &c, img.data.data());
    if (rc != RCG_OK) return Error::from_status(rc);
    return img;
}

// core/src/ssimulacra2.rs:471. Alpha is composited onto mid-grey first and 16-bit images are reduced to 8 bits
// (this path scores 8-bit pixels).
inline Result<double> ssimulacra2(Device& dev, const Image& reference, const Image& distorted,
                                  int precision = RCG_PREC_FP32) {
    if (reference.width != distorted.width || reference.height != distorted.height)
        return Error{Error::Invalid, "images differ in size"};
    if (reference.width < MIN_SIDE || reference.height < MIN_SIDE)
        return Error{Error::Invalid, "SSIMULACRA 2 needs at least 8x8 pixels"};
    Image a = reference.flatten({128, 128, 128}), b = distorted.flatten({128, 128, 128});
    // Gray against Rgb: score both as Rgb (the reference replicates grey into three channels).
    auto to_rgb = [](const Image& g) {
        Image o{g.width, g.height, ColorType::Rgb, 8, {}};
        o.data.resize(g.data.size() * 3);
        for (size_t i = 0; i < g.data.size(); i++) o.data[3 * i] = o.data[3 * i + 1] = o.data[3 * i + 2] = g.data[i];
        return o;
    };
    if (a.color != b.color) {
        if (a.color == ColorType::Gray) a = to_rgb(a);
        if (b.color == ColorType::Gray) b = to_rgb(b);
    }
    double score = 0.0;
    const int rc = rcg_ssimulacra2(dev.raw(), a.data.data(), b.data.data(), a.width, a.height,
                                   uint32_t(channels(a.color)), precision, &score, nullptr, nullptr);
    if (rc != RCG_OK) return Error::from_status(rc);
    return score;
}

// jpeg/src/target.rs:29-56, generic over `make` and `metric` exactly like the reference: serial bisection over
// 1..=max_quality, max_quality first; a candidate that cannot be reconstructed scores f64::MIN.
inline Found find_quality_with(Device& dev, const Image& reference, double target, uint8_t max_quality,
                               const std::function<CoefImage(uint8_t)>& make,
                               const std::function<double(const Image&, const Image&)>& metric) {
    auto score = [&](const CoefImage& ci) -> double {
        Result<Image> img = reconstruct(dev, ci);
        return img.is_ok() ? metric(reference, img.value()) : -1.7976931348623157e308;
    };
    uint8_t lo = 1, hi = std::max<uint8_t>(max_quality, 1);
    CoefImage top = make(hi);
    const double top_score = score(top);
    Found best{hi, top_score, std::move(top)};
    if (top_score < target) return best;
    while (lo < hi) {
        const uint8_t mid = uint8_t(lo + (hi - lo) / 2);
        CoefImage ci = make(mid);
        const double s = score(ci);
        if (s >= target) {
            hi = mid;
            best = Found{mid, s, std::move(ci)};
        } else {
            lo = uint8_t(mid + 1);
        }
    }
    return best;
}

// The CLI's --target-ssimulacra2 search (cli/src/pipeline.rs:207-247) with make = build_coefs(opts) and
// metric = ssimulacra2, run device-resident: rcg_find_quality with trellis off (one upload, one score read back per
// step), rcg_find_quality_trellis with the reference's default trellis (the host trellis between the device encode
// and the device scoring of every step).
inline Result<Found> find_jpeg_quality(Device& dev, const Image& reference, double target, const EncodeOptions& opts) {
    if (reference.width < MIN_SIDE || reference.height < MIN_SIDE)
        return Error{Error::Unsupported, "--target-ssimulacra2 needs an image of at least 8x8 pixels"};
    const Image src = reference.flatten(opts.background);
    const uint32_t ch = uint32_t(channels(src.color));
    // The chosen quality's layout is only known afterwards: size for every layout the search can meet (under Auto
    // both 4:4:4 and 4:2:0, whose 16-pixel MCUs can give luma an extra block row or column).
    rcg_layout L, L2;
    int rc = rcg_coef_layout(src.width, src.height, ch, 100, subsampling_code(opts.subsampling), &L);
    if (rc == RCG_OK) rc = rcg_coef_layout(src.width, src.height, ch, 1, subsampling_code(opts.subsampling), &L2);
    if (rc != RCG_OK) return Error::from_status(rc);
    std::vector<std::vector<int16_t>> bufs(L.ncomp);
    int16_t* ptrs[3] = {nullptr, nullptr, nullptr};
    for (uint32_t c = 0; c < L.ncomp; c++) {
        bufs[c].resize(std::max(size_t(L.blocks_w[c]) * L.blocks_h[c], size_t(L2.blocks_w[c]) * L2.blocks_h[c]) * 64);
        ptrs[c] = bufs[c].data();
    }
    rcg_found f{};
    uint8_t qt[128];
    rc = (opts.trellis ? rcg_find_quality_trellis : rcg_find_quality)(dev.raw(), src.data.data(), src.width,
                                                                       src.height, ch, target, 100,
                                                                       subsampling_code(opts.subsampling), &f, ptrs,
                                                                       qt);
    if (rc != RCG_OK) return Error::from_status(rc);
    rcg_layout Lq;
    rcg_coef_layout(src.width, src.height, ch, f.quality, subsampling_code(opts.subsampling), &Lq);
    Found out;
    out.quality = uint8_t(f.quality);
    out.score = f.score;
    out.coefs.width = src.width;
    out.coefs.height = src.height;
    for (uint32_t c = 0; c < Lq.ncomp; c++) {
        CoefComponent k;
        k.id = uint8_t(c + 1);
        k.h = c == 0 ? Lq.hmax : 1;
        k.v = c == 0 ? Lq.vmax : 1;
        k.tq = c == 0 ? 0 : 1;
        k.blocks_w = Lq.blocks_w[c];
        k.coefs.assign(bufs[c].begin(), bufs[c].begin() + size_t(Lq.blocks_w[c]) * Lq.blocks_h[c] * 64);
        out.coefs.comps.push_back(std::move(k));
    }
    std::array<uint8_t, 64> a, b;
    std::copy(qt, qt + 64, a.begin());
    std::copy(qt + 64, qt + 128, b.begin());
    out.coefs.qtables[0] = a;
    if (Lq.ncomp == 3) out.coefs.qtables[1] = b;
    return out;
}

}  // namespace rc_gpu
