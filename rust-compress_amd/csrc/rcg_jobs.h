// Work descriptors shared by the host launchers and the kernels.
//
// A launch processes a list of jobs (one image, or one image at one scale), so one launch can cover many images of
// uneven size (the batched search) or a single one (the plain drop-in), with the same kernel. The encode kernel
// numbers the tiles of its jobs consecutively (job j owns workgroups [tile_begin, tile_begin + tiles), found by a
// binary search). The SSIMULACRA2 kernels read their (job, tile) from a per-launch unit table instead, which the
// host orders for the L2 caches: see Unit.
#pragma once
#include <stdint.h>

namespace rcg {

constexpr int kMaxCand = 16;       // candidate qualities evaluated together for one image
constexpr int kEncTileW = 512;     // encode tile width in pixels (64 luma blocks)
constexpr int kEncThreads = 192;   // threads per encode workgroup: one per 8x8 block of a 4:2:0/4:4:4 tile

// One candidate quality of one image: its quantisation data and where its outputs go.
struct EncCand {
    int16_t* coef[3];     // zigzag coefficients per component, or null when not wanted
    uint8_t* rec[3];      // reconstructed (IDCT) sample planes per component, or null
    float* scaled[3];     // trellis hand-off: the f32 values a[k] = dct[f]*div[f] before rounding, zigzag order, or null
    float div[2][64];     // reciprocal divisors, transposed [u*8+v] layout (jpeg/src/encode.rs:175-183)
    float mult[2][64];    // dequantisation multipliers, natural [v*8+u] layout (jpeg/src/decode.rs:1317-1322)
};

struct EncJob {
    const uint8_t* src;   // interleaved Gray8 or RGB8, row stride w*channels
    uint64_t src_bytes;   // w*h*channels
    uint32_t w, h, channels;
    uint32_t hmax, vmax;
    uint32_t pw, ph;      // padded to whole MCUs
    uint32_t tiles_x, tiles_y;
    uint32_t tile_begin;
    uint32_t ncand;
    const EncCand* cand; // ncand entries
};

// A plane of IDCT output samples (stride = blocks_w*8) and how it maps onto the full-size image.
struct PlaneView {
    const uint8_t* p;
    uint32_t stride;
    uint32_t cw, ch;      // real (unpadded) component size: ceil(w*h_c/hmax) x ceil(h*v_c/vmax)
    uint32_t fx, fy;      // upsampling factors hmax/h_c, vmax/v_c
};

// Source of linear-light RGB for one image at one SSIMULACRA2 scale.
// kSrcRefCache: the reference's per-scale cache (9 floats per pixel: the shifted XYB, its blur and sigma_aa + C2, in
// the plane layout of Loader<kSrcRefCache> in ssim2.hip), read by candidate kernels in place of the reference's pixels.
enum SrcKind : uint32_t { kSrcBytes = 0, kSrcRecon = 1, kSrcPlanes = 2, kSrcRefCache = 3 };

struct SsimSrc {
    uint32_t kind;
    uint32_t channels;      // kSrcBytes: 1 or 3
    uint32_t ncomp;         // kSrcRecon: 1 or 3
    uint32_t rgb_direct;    // kSrcRecon: components are R,G,B (no colour conversion)
    const uint8_t* bytes;   // kSrcBytes
    PlaneView comp[3];      // kSrcRecon
    const float* planes;    // kSrcPlanes: [3][h][w]
};

struct SsimJob {
    uint32_t w, h;          // size at this scale
    uint32_t th;            // output rows per tile (even)
    uint32_t tiles_x, tiles_y;
    uint32_t x_begin;       // first output column of the job's tiles (a tail job covers the image's last columns)
    SsimSrc a, b;           // a = reference, b = distorted
    float* a_next;          // next scale's planes [3][ceil(h/2)][ceil(w/2)] or null
    float* b_next;
    double* partials;       // [tiles_x*tiles_y][18] per-tile sums, written by this launch
};

// One image at one scale for the reference-cache kernel: its pixels in, its cache (and next pyramid level) out.
struct RefJob {
    uint32_t w, h, th;
    uint32_t tiles_x, tiles_y;
    SsimSrc a;
    float* a_next;          // next scale's planes [3][ceil(h/2)][ceil(w/2)] or null
    float* cache;           // 9 * w * h floats (layout of Loader<kSrcRefCache>)
};

// Workgroup b of an SSIMULACRA2 launch processes tile units[b].tile of job units[b].job (job == kIdleJob: none); a
// launch of one job may pass no table, and then workgroup b processes tile b.
// Workgroups are dispatched round-robin over the GPU's 8 XCDs (b mod 8), each with its own L2. The host places the
// candidates of one image tile in consecutive slots of one XCD, so the reference cache rows they all read come from
// HBM once and from that XCD's L2 for the others.
struct Unit {
    uint32_t job, tile;
};
constexpr uint32_t kIdleJob = 0xFFFFFFFFu;
constexpr int kXcds = 8;

}  // namespace rcg
