// C++ tests of the host API (rust-compress_amd/host/rc_gpu.hpp) on the GPU, written after the reference's own tests
// of this path so they read alike:
//   core/src/ssimulacra2.rs:541-633 (identity, size checks), cli/tests/ssimulacra2.rs:40-123 (target search),
//   jpeg/tests/jpeg.rs:65-77 (odd and tiny sizes), :97-119 (search meets the target, q-1 misses), :132-152 (4:2:2).
// Prints one line per test and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../rust-compress_amd/host/rc_gpu.hpp"

using namespace rc_gpu;

#define CHECK(cond, ...)                                            \
    do {                                                            \
        if (!(cond)) {                                              \
            std::printf("FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
            std::printf(__VA_ARGS__);                               \
            std::printf("\n");                                      \
            std::exit(1);                                           \
        }                                                           \
    } while (0)

static Image smooth_image(uint32_t w, uint32_t h, ColorType color) {
    const size_t ch = channels(color);
    std::vector<uint8_t> data;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++)
            for (size_t c = 0; c < ch; c++) data.push_back(uint8_t((x * 3 + y * 2 + uint32_t(c) * 40) % 256));
    return Image::create(w, h, color, 8, data).value();
}

static Image textured(uint32_t w, uint32_t h) {
    std::vector<uint8_t> data(size_t(w) * h * 3);
    if (rcg_synth_rgb8(5, w, h, data.data()) != RCG_OK) std::exit(2);
    return Image::create(w, h, ColorType::Rgb, 8, data).value();
}

static double psnr(const std::vector<uint8_t>& a, const std::vector<uint8_t>& b) {
    double mse = 0;
    for (size_t i = 0; i < a.size(); i++) mse += (double(a[i]) - b[i]) * (double(a[i]) - b[i]);
    mse /= double(a.size());
    return mse == 0 ? 99.0 : 10.0 * std::log10(255.0 * 255.0 / mse);
}

static EncodeOptions opts(uint8_t q, Subsampling s) {
    EncodeOptions o;
    o.quality = q;
    o.subsampling = s;
    o.trellis = false;
    return o;
}

int main() {
    Device dev(0);

    {  // identical images score exactly 100
        Image img = textured(96, 72);
        Result<double> s = ssimulacra2(dev, img, img, RCG_PREC_FP64);
        CHECK(s.is_ok() && s.value() == 100.0, "score %f", s.unwrap_or(-1));
        std::printf("ok identical_images_score_exactly_100\n");
    }
    {  // size checks
        Image a = textured(8, 7), b = textured(8, 8);
        CHECK(ssimulacra2(dev, a, a).is_err() && ssimulacra2(dev, a, a).error().kind == Error::Invalid, "small");
        CHECK(ssimulacra2(dev, b, a).is_err(), "size mismatch");
        std::printf("ok rejects_small_and_mismatched\n");
    }
    {  // odd and tiny sizes round-trip with the right dimensions
        const uint32_t sizes[][2] = {{1, 1}, {7, 9}, {8, 8}, {9, 9}, {17, 1}, {1, 17}, {33, 31}, {250, 3}};
        for (auto& wh : sizes)
            for (ColorType c : {ColorType::Gray, ColorType::Rgb})
                for (Subsampling s : {Subsampling::S444, Subsampling::S422, Subsampling::S420}) {
                    Image img = smooth_image(wh[0], wh[1], c);
                    Result<CoefImage> ci = build_coefs(dev, img, opts(90, s));
                    CHECK(ci.is_ok(), "build_coefs %ux%u", wh[0], wh[1]);
                    Result<Image> back = reconstruct(dev, ci.value());
                    CHECK(back.is_ok() && back.value().width == wh[0] && back.value().height == wh[1], "dims");
                    CHECK(back.value().color == c, "colour type");
                }
        std::printf("ok odd_and_tiny_sizes\n");
    }
    {  // 4:2:2 layout and fidelity
        const uint32_t w = 203, h = 77;
        std::vector<uint8_t> data;
        for (uint32_t y = 0; y < h; y++)
            for (uint32_t x = 0; x < w; x++) {
                data.push_back(uint8_t(x * 5 / 4));
                data.push_back(uint8_t(y * 3));
                data.push_back(uint8_t(128 + x / 4 + y));
            }
        Image img = Image::create(w, h, ColorType::Rgb, 8, data).value();
        CoefImage ci = build_coefs(dev, img, opts(92, Subsampling::S422)).value();
        CHECK(ci.comps[0].h == 2 && ci.comps[0].v == 1 && ci.comps[1].h == 1, "sampling factors");
        const double p = psnr(reconstruct(dev, ci).value().data, img.data);
        CHECK(p > 30.0, "PSNR %.2f", p);
        std::printf("ok s422_layout_and_fidelity\n");
    }
    {  // the search meets its target, is monotone in it, and one step lower misses
        Image ref = textured(192, 160);
        auto make = [&](uint8_t q) { return build_coefs(dev, ref, opts(q, Subsampling::S420)).value(); };
        auto metric = [&](const Image& a, const Image& b) { return ssimulacra2(dev, a, b).unwrap_or(-1.7976931348623157e308); };
        uint8_t last = 0;
        for (double target : {50.0, 65.0, 80.0, 90.0}) {
            Found f = find_quality_with(dev, ref, target, 100, make, metric);
            CHECK(f.score >= target || f.quality == 100, "target %.1f: %f", target, f.score);
            CHECK(f.quality >= last, "quality dropped");
            last = f.quality;
            if (f.quality > 1) {
                Image lower = reconstruct(dev, make(uint8_t(f.quality - 1))).value();
                CHECK(metric(ref, lower) < target + 0.5, "q-1 would also pass");
            }
            // The device-resident search picks the same quality.
            Result<Found> g = find_jpeg_quality(dev, ref, target, opts(100, Subsampling::S420));
            CHECK(g.is_ok() && g.value().quality == f.quality, "device search %d vs %d", g.value().quality, f.quality);
            CHECK(g.value().coefs.comps[0].coefs == make(f.quality).comps[0].coefs, "winner coefficients");
        }
        std::printf("ok quality_search_meets_the_target_and_is_monotonic_in_it\n");
    }
    {  // flag-style errors of the search (cli/tests/ssimulacra2.rs:111-122)
        Image tiny = textured(7, 7);
        Result<Found> r = find_jpeg_quality(dev, tiny, 70.0, opts(100, Subsampling::S420));
        CHECK(r.is_err() && r.error().kind == Error::Unsupported, "tiny image");
        std::printf("ok search_rejects_images_under_8x8\n");
    }
    {  // trellis on (the reference's default): the device hands the unrounded values to the host trellis. DC stays the
       // plain DC, every AC is zero or one of its two candidates, and the file shrinks (jpeg/src/rdo.rs:1-7)
        Image img = textured(200, 136);
        EncodeOptions t = opts(80, Subsampling::S420);
        t.trellis = true;
        CoefImage tr = build_coefs(dev, img, t).value();
        CoefImage pl = build_coefs(dev, img, opts(80, Subsampling::S420)).value();
        size_t changed = 0;
        for (size_t c = 0; c < tr.comps.size(); c++)
            for (size_t i = 0; i < tr.comps[c].coefs.size(); i++) {
                const int a = tr.comps[c].coefs[i], p = pl.comps[c].coefs[i];
                if (i % 64 == 0) CHECK(a == p, "DC changed");
                // plain rounds half to even, the trellis's candidates come from round-half-up: allow one step
                const bool ok = a == 0 || (p == 0 ? std::abs(a) == 1 : ((a > 0) == (p > 0) && std::abs(std::abs(a) - std::abs(p)) <= 1));
                CHECK(ok, "candidate set: %d vs %d", a, p);
                changed += a != p;
            }
        CHECK(changed > 0, "trellis changed nothing");
        CHECK(write_jpeg(tr).value().size() < write_jpeg(pl).value().size(), "trellis file not smaller");
        t.quality = 100;  // the reference skips the trellis at q100 (encode.rs:527)
        CHECK(build_coefs(dev, img, t).value().comps[0].coefs == build_coefs(dev, img, opts(100, Subsampling::S420)).value().comps[0].coefs,
              "q100 is plain");
        std::printf("ok trellis_hand_off\n");
    }
    {  // the device search with the trellis (the CLI default) equals the generic bisection over make = build_coefs
       // with trellis, and returns the winner's trellis coefficients
        Image ref = textured(144, 104);
        EncodeOptions t = opts(100, Subsampling::Auto);
        t.trellis = true;
        auto make = [&](uint8_t q) {
            EncodeOptions o = t;
            o.quality = q;
            return build_coefs(dev, ref, o).value();
        };
        auto metric = [&](const Image& a, const Image& b) { return ssimulacra2(dev, a, b).unwrap_or(-1.7976931348623157e308); };
        for (double target : {60.0, 85.0}) {
            Found f = find_quality_with(dev, ref, target, 100, make, metric);
            Result<Found> g = find_jpeg_quality(dev, ref, target, t);
            CHECK(g.is_ok() && g.value().quality == f.quality, "trellis search %d vs %d", g.value().quality, f.quality);
            for (size_t c = 0; c < f.coefs.comps.size(); c++)
                CHECK(g.value().coefs.comps[c].coefs == f.coefs.comps[c].coefs, "winner coefficients, component %zu", c);
        }
        std::printf("ok trellis_search_equals_the_generic_bisection\n");
    }
    {  // JPEG -> JPEG (cli/src/pipeline.rs:369-386): requantising a q92 file towards a target; q100 is lossless
       // (cli/tests/cmyk.rs:38-54) and the device search's winner equals requantize at the chosen quality
        Image img = textured(176, 120);
        CoefImage src = build_coefs(dev, img, opts(92, Subsampling::S420)).value();
        CoefImage same = src;
        CHECK(requantize(same, 100, true).is_ok(), "requantize q100");
        CHECK(same.comps[0].coefs == src.comps[0].coefs && same.qtables[0] == src.qtables[0], "q100 lossless");
        for (bool trellis : {false, true}) {
            Found f = find_jpeg_quality_requant(dev, src, 70.0, trellis).value();
            CHECK(f.score >= 70.0 || f.quality == 100, "requant target: %f", f.score);
            CoefImage w = src;
            requantize(w, f.quality, trellis).value();
            CHECK(w.comps[0].coefs == f.coefs.comps[0].coefs && w.comps[1].coefs == f.coefs.comps[1].coefs,
                  "winner");
            const Image decoded = reconstruct(dev, src).value();
            const double s = ssimulacra2(dev, decoded, reconstruct(dev, f.coefs).value()).value();
            CHECK(std::fabs(s - f.score) < 0.05, "score %f vs %f", s, f.score);
        }
        std::printf("ok requantisation_search\n");
    }
    {  // the search's winner, written: the same file as writing build_coefs at the chosen quality (the CLI's
       // encode_jpeg_pixels, cli/src/pipeline.rs:225-247), and write_best picks one of the two encodings
        Image ref = textured(160, 96);
        Found f = find_jpeg_quality(dev, ref, 75.0, opts(100, Subsampling::S420)).value();
        const std::vector<uint8_t> a = write_jpeg(f.coefs).value();
        const std::vector<uint8_t> b = write_jpeg(build_coefs(dev, ref, opts(f.quality, Subsampling::S420)).value()).value();
        CHECK(a == b, "winner file differs");
        CHECK(a.size() > 4 && a[0] == 0xFF && a[1] == 0xD8 && a[a.size() - 2] == 0xFF && a[a.size() - 1] == 0xD9,
              "SOI/EOI");
        const std::vector<uint8_t> p = write_progressive(f.coefs).value();
        const std::vector<uint8_t> best = write_best(f.coefs).value();
        CHECK(best == a || best == p, "write_best is one of the two");
        CoefImage broken = f.coefs;
        broken.qtables[0].reset();
        CHECK(write_jpeg(broken).is_err(), "missing table");
        std::printf("ok search_winner_writes_the_same_file\n");
    }
    std::printf("all reference-API tests passed\n");
    return 0;
}
