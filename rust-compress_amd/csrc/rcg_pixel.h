// Pixel-domain helpers of the JPEG reconstruction, shared by the reconstruct kernels and the SSIMULACRA2 loader
// (which upsamples and colour-converts reconstructed planes on the fly).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rcg_jobs.h"

namespace rcg {

// One full-resolution sample of a component (decode.rs:1769-1829).
__device__ __forceinline__ int32_t upsample_at(const PlaneView& v, uint32_t x, uint32_t y) {
    const uint8_t* p = v.p;
    const uint32_t s = v.stride;
    if (v.fx == 1 && v.fy == 1) return p[uint64_t(y) * s + x];
    if (v.fx == 2 && v.fy == 1) {
        const uint32_t i = x >> 1;
        const int32_t nearv = p[uint64_t(y) * s + i];
        const uint32_t fi = (x & 1) == 0 ? (i > 0 ? i - 1 : 0) : min(i + 1, v.cw - 1);
        const int32_t far = p[uint64_t(y) * s + fi];
        return (3 * nearv + far + ((x & 1) ? 2 : 1)) >> 2;
    }
    if (v.fx == 1 && v.fy == 2) {
        const uint32_t jj = y >> 1;
        const uint32_t fy = (y & 1) == 0 ? (jj > 0 ? jj - 1 : 0) : min(jj + 1, v.ch - 1);
        return (3 * int32_t(p[uint64_t(jj) * s + x]) + int32_t(p[uint64_t(fy) * s + x]) + ((y & 1) ? 2 : 1)) >> 2;
    }
    if (v.fx == 2 && v.fy == 2) {
        const uint32_t jj = y >> 1, i = x >> 1;
        const uint32_t fy = (y & 1) == 0 ? (jj > 0 ? jj - 1 : 0) : min(jj + 1, v.ch - 1);
        const uint32_t fi = (x & 1) == 0 ? (i > 0 ? i - 1 : 0) : min(i + 1, v.cw - 1);
        const uint8_t* r0 = p + uint64_t(jj) * s;
        const uint8_t* r1 = p + uint64_t(fy) * s;
        const int32_t nearv = 3 * int32_t(r0[i]) + int32_t(r1[i]);
        const int32_t far = 3 * int32_t(r0[fi]) + int32_t(r1[fi]);
        return (3 * nearv + far + ((x & 1) ? 7 : 8)) >> 4;
    }
    return p[uint64_t(y / v.fy) * s + x / v.fx];
}

// k * x + c for a constant k and x that fit in 24 signed bits, as v_mad_i32_i24 (full rate). A 32-bit multiply is
// quarter rate, and the compiler emits one wherever it cannot see that x is small (x computed in another basic
// block: the SSIMULACRA2 loader's chroma comes out of a branch per sampling layout). k is read from an SGPR and c
// from a VGPR (one scalar operand per instruction on this GPU).
__device__ __forceinline__ int32_t mad_i24(int32_t k, int32_t x, int32_t c) {
    int32_t r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "s"(k), "v"(x), "v"(c));
    return r;
}

// Fixed-point YCbCr -> RGB with clamping (decode.rs:1434-1441). Every factor and offset sample fits in 24 signed bits
// and every product in 32, so 24-bit multiply-adds give the reference's sums exactly.
__device__ __forceinline__ void ycc_to_rgb(int32_t a, int32_t b, int32_t c, int32_t& r, int32_t& g, int32_t& bl) {
    const int32_t cb = b - 128, cr = c - 128;
    r = a + (mad_i24(91881, cr, 32768) >> 16);
    g = a - (mad_i24(46802, cr, mad_i24(22554, cb, 32768)) >> 16);
    bl = a + (mad_i24(116130, cb, 32768) >> 16);
    r = r < 0 ? 0 : (r > 255 ? 255 : r);
    g = g < 0 ? 0 : (g > 255 ? 255 : g);
    bl = bl < 0 ? 0 : (bl > 255 ? 255 : bl);
}

// Four components to RGB (cmyk_to_rgb, decode.rs:1454-1478): with an Adobe marker the values are stored inverted
// (255 = no ink) and transform 2 means the first three are YCbCr-coded; without a marker they are ink amounts. The
// light an ink lets through is multiplied by the black's, (v*k + 127) / 255.
__device__ __forceinline__ void cmyk_to_rgb(int32_t c, int32_t m, int32_t y, int32_t k, int adobe_transform,
                                            int32_t& r, int32_t& g, int32_t& b) {
    if (adobe_transform == 2) {
        int32_t rr, gg, bb;
        ycc_to_rgb(c, m, y, rr, gg, bb);
        c = 255 - rr;
        m = 255 - gg;
        y = 255 - bb;
    }
    const bool inverted = adobe_transform >= 0;
    const int32_t lk = inverted ? k : 255 - k;
    r = ((inverted ? c : 255 - c) * lk + 127) / 255;
    g = ((inverted ? m : 255 - m) * lk + 127) / 255;
    b = ((inverted ? y : 255 - y) * lk + 127) / 255;
}

}  // namespace rcg
