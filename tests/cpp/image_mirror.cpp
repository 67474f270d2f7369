// Host-only driver of the C++ mirror's Image (rust-compress_amd/host/rc_gpu.hpp; core/src/image.rs:104-190): reads
// one image from stdin, flattens it, and writes the result to stdout, so tests/test_image_mirror.py can check it
// against the command-line driver's own flattening. Nothing here touches a device.
//
// stdin:  u32 width, u32 height, u32 color (0 Gray, 1 GrayAlpha, 2 Rgb, 3 Rgba), u32 depth, u8 bg[3], then the
//         samples (16-bit ones big-endian, as the reference stores them).
// stdout: u32 status (0 = ok, else 1 + Error::Kind), then on success u32 width, u32 height, u32 color, u32 depth and
//         the flattened samples.
#include <cstdio>
#include <vector>

#include "../../rust-compress_amd/host/rc_gpu.hpp"

static bool read_all(void* p, size_t n) { return std::fread(p, 1, n, stdin) == n; }
static void write_u32(uint32_t v) { std::fwrite(&v, 4, 1, stdout); }

int main() {
    uint32_t hdr[4];
    std::array<uint8_t, 3> bg;
    if (!read_all(hdr, sizeof(hdr)) || !read_all(bg.data(), 3)) return 2;
    std::vector<uint8_t> data;
    uint8_t buf[4096];
    for (size_t n; (n = std::fread(buf, 1, sizeof(buf), stdin)) > 0;) data.insert(data.end(), buf, buf + n);
    if (hdr[2] > 3) return 2;
    auto img = rc_gpu::Image::create(hdr[0], hdr[1], rc_gpu::ColorType(hdr[2]), uint8_t(hdr[3]), std::move(data));
    if (img.is_err()) {
        write_u32(1 + uint32_t(img.error().kind));
        return 0;
    }
    const rc_gpu::Image out = img.value().flatten(bg);
    write_u32(0);
    write_u32(out.width);
    write_u32(out.height);
    write_u32(uint32_t(out.color));
    write_u32(out.depth);
    std::fwrite(out.data.data(), 1, out.data.size(), stdout);
    return 0;
}
