// Seeded synthetic RGB8 images for the benchmark and parity workloads.
//
// For each channel: 128 + the sum of 4 2-D sinusoids A*sin(2*pi*(x*cos(t) + y*sin(t))/p + phi) with p in
// U[16,256] px, A in U[10,60], t in U[0,pi), phi in U[0,2*pi); then 12 axis-aligned rectangles of uniform random
// colour painted in order with hard edges (top-left corner uniform in the image, sides U[W/32, W/4] x U[H/32, H/4],
// clipped); then i.i.d. Gaussian noise of sigma 4 per sample (Box-Muller); finally floor(v + 0.5) clamped to
// [0, 255]. The parameters come from one SplitMix64 stream seeded with `seed`; the noise of sample n is a hash of
// (seed, n), so the image does not depend on how rows are split across threads.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "rcg_host.h"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct SplitMix64 {
    uint64_t s;
    uint64_t next() { return mix64(s += kGolden); }
    double uniform() { return double(next() >> 11) * 0x1.0p-53; }
};

struct Wave {
    double fx, fy, amp, phase;
};
struct Rect {
    uint32_t x0, y0, x1, y1;
    uint8_t rgb[3];
};

}  // namespace

extern "C" int rcg_synth_rgb8(uint64_t seed, uint32_t width, uint32_t height, uint8_t* out) {
    return rcg::guarded([&]() -> int {
        if (!out || width == 0 || height == 0) return RCG_INVALID;
        const double kPi = 3.14159265358979323846;
        SplitMix64 rng{seed};
        Wave waves[3][4];
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 4; k++) {
                const double p = 16.0 + 240.0 * rng.uniform();
                const double a = 10.0 + 50.0 * rng.uniform();
                const double t = kPi * rng.uniform();
                const double phi = 2.0 * kPi * rng.uniform();
                waves[c][k] = Wave{2.0 * kPi * std::cos(t) / p, 2.0 * kPi * std::sin(t) / p, a, phi};
            }
        Rect rects[12];
        for (int i = 0; i < 12; i++) {
            const double W = width, H = height;
            const uint32_t x0 = std::min(width - 1, uint32_t(std::floor(rng.uniform() * W)));
            const uint32_t y0 = std::min(height - 1, uint32_t(std::floor(rng.uniform() * H)));
            const double rw = W / 32.0 + rng.uniform() * (W / 4.0 - W / 32.0);
            const double rh = H / 32.0 + rng.uniform() * (H / 4.0 - H / 32.0);
            const uint32_t w = std::max(1u, uint32_t(std::floor(rw))), h = std::max(1u, uint32_t(std::floor(rh)));
            Rect r;
            r.x0 = x0;
            r.y0 = y0;
            r.x1 = uint32_t(std::min<uint64_t>(width, uint64_t(x0) + w));
            r.y1 = uint32_t(std::min<uint64_t>(height, uint64_t(y0) + h));
            for (int c = 0; c < 3; c++) r.rgb[c] = uint8_t(std::min(255.0, std::floor(rng.uniform() * 256.0)));
            rects[i] = r;
        }
        const uint64_t noise_key = mix64(seed ^ 0xD1B54A32D192ED03ull);

        auto rows = [&](uint32_t ya, uint32_t yb) {
            for (uint32_t y = ya; y < yb; y++)
                for (uint32_t x = 0; x < width; x++) {
                    double v[3];
                    int top = -1;
                    for (int i = 0; i < 12; i++)
                        if (x >= rects[i].x0 && x < rects[i].x1 && y >= rects[i].y0 && y < rects[i].y1) top = i;
                    for (int c = 0; c < 3; c++) {
                        if (top >= 0) {
                            v[c] = rects[top].rgb[c];
                        } else {
                            double s = 128.0;
                            for (int k = 0; k < 4; k++) {
                                const Wave& wv = waves[c][k];
                                s += wv.amp * std::sin(wv.fx * double(x) + wv.fy * double(y) + wv.phase);
                            }
                            v[c] = s;
                        }
                    }
                    const uint64_t n = (uint64_t(y) * width + x) * 3;
                    for (int c = 0; c < 3; c++) {
                        const uint64_t h1 = mix64(noise_key + (2 * (n + c) + 1) * kGolden);
                        const uint64_t h2 = mix64(noise_key + (2 * (n + c) + 2) * kGolden);
                        const double u1 = double((h1 >> 11) + 1) * 0x1.0p-53;
                        const double u2 = double(h2 >> 11) * 0x1.0p-53;
                        const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
                        const double val = std::floor(v[c] + 4.0 * g + 0.5);
                        out[n + c] = uint8_t(std::min(255.0, std::max(0.0, val)));
                    }
                }
        };
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const uint32_t nthreads = std::min<uint32_t>(hw, std::max<uint32_t>(1, height / 16));
        if (nthreads <= 1) {
            rows(0, height);
            return RCG_OK;
        }
        std::vector<std::thread> pool;
        for (uint32_t t = 0; t < nthreads; t++) {
            const uint32_t a = uint32_t(uint64_t(height) * t / nthreads), b = uint32_t(uint64_t(height) * (t + 1) / nthreads);
            pool.emplace_back(rows, a, b);
        }
        for (auto& th : pool) th.join();
        return RCG_OK;
    });
}
