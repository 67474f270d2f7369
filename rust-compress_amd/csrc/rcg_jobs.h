// Work descriptors shared by the host launchers and the kernels.
//
// A launch processes a list of jobs (one image, or one image at one scale) whose tiles are numbered
// consecutively: job j owns workgroups [tile_begin, tile_begin + tiles). A workgroup finds its job by a binary
// search over tile_begin, so one launch can cover many images of uneven size (the batched search) or a single one
// (the plain drop-in), with the same kernel.
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
enum SrcKind : uint32_t { kSrcBytes = 0, kSrcRecon = 1, kSrcPlanes = 2 };

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
    uint32_t tile_begin;
    SsimSrc a, b;           // a = reference, b = distorted
    float* a_next;          // next scale's planes [3][ceil(h/2)][ceil(w/2)] or null
    float* b_next;
    double* partials;       // [tiles_x*tiles_y][18] per-tile sums, written by this launch
};

}  // namespace rcg
