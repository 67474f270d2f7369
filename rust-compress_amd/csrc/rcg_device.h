// Device-side constants and helpers shared by the JPEG and SSIMULACRA2 kernels (gfx950).
//
// Everything on the bit-exact JPEG path is f32 in the reference's operation order. The library is compiled with
// -ffp-contract=off, so no multiply-add is fused unless a kernel asks for it with an explicit fmaf() or a
// "#pragma clang fp contract(fast)" scope (the SSIMULACRA2 kernels do; the DCT kernels never do).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rcg {

// Natural index of the k-th zigzag coefficient (jpeg/src/lib.rs:15-19).
__host__ __device__ constexpr int zigzag(int k) {
    constexpr int z[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                           12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                           35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                           58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
    return z[k];
}

// AAN per-frequency scale factors (jpeg/src/dct.rs:8), each the f32 nearest its decimal literal.
__host__ __device__ constexpr float aan(int i) {
    constexpr float a[8] = {1.0f, 1.3870398f, 1.306563f, 1.1758756f, 1.0f, 0.78569495f, 0.5411961f, 0.27589938f};
    return a[i];
}

// Round half to even through the 1.5*2^23 magic constant (jpeg/src/dct.rs:127-129); valid for |x| < 2^22.
__device__ __forceinline__ int32_t round_f32(float x) {
    return __float_as_int(__fadd_rn(x, 12582912.0f)) - 0x4B400000;
}

// One AAN forward pass over 8 values (jpeg/src/dct.rs:20-57), in place: d[0..7] along the pass direction.
// Explicit __fadd_rn/__fmul_rn pin every operation to a separately rounded IEEE op.
__device__ __forceinline__ void fdct8(float* d) {
    float tmp0 = __fadd_rn(d[0], d[7]), tmp7 = __fsub_rn(d[0], d[7]);
    float tmp1 = __fadd_rn(d[1], d[6]), tmp6 = __fsub_rn(d[1], d[6]);
    float tmp2 = __fadd_rn(d[2], d[5]), tmp5 = __fsub_rn(d[2], d[5]);
    float tmp3 = __fadd_rn(d[3], d[4]), tmp4 = __fsub_rn(d[3], d[4]);
    float tmp10 = __fadd_rn(tmp0, tmp3), tmp13 = __fsub_rn(tmp0, tmp3);
    float tmp11 = __fadd_rn(tmp1, tmp2), tmp12 = __fsub_rn(tmp1, tmp2);
    d[0] = __fadd_rn(tmp10, tmp11);
    d[4] = __fsub_rn(tmp10, tmp11);
    float z1 = __fmul_rn(__fadd_rn(tmp12, tmp13), 0.70710677f);
    d[2] = __fadd_rn(tmp13, z1);
    d[6] = __fsub_rn(tmp13, z1);
    float o10 = __fadd_rn(tmp4, tmp5), o11 = __fadd_rn(tmp5, tmp6), o12 = __fadd_rn(tmp6, tmp7);
    float z5 = __fmul_rn(__fsub_rn(o10, o12), 0.38268343f);
    float z2 = __fadd_rn(__fmul_rn(0.5411961f, o10), z5);
    float z4 = __fadd_rn(__fmul_rn(1.306563f, o12), z5);
    float z3 = __fmul_rn(o11, 0.70710677f);
    float z11 = __fadd_rn(tmp7, z3), z13 = __fsub_rn(tmp7, z3);
    d[5] = __fadd_rn(z13, z2);
    d[3] = __fsub_rn(z13, z2);
    d[1] = __fadd_rn(z11, z4);
    d[7] = __fsub_rn(z11, z4);
}

// One AAN inverse pass (jpeg/src/dct.rs:67-109).
__device__ __forceinline__ void idct8(float* d) {
    const float SQRT2 = 1.41421356237309504880f;
    float t0 = d[0], t1 = d[2], t2 = d[4], t3 = d[6];
    float tmp10 = __fadd_rn(t0, t2), tmp11 = __fsub_rn(t0, t2);
    float tmp13 = __fadd_rn(t1, t3);
    float tmp12 = __fsub_rn(__fmul_rn(__fsub_rn(t1, t3), SQRT2), tmp13);
    float e0 = __fadd_rn(tmp10, tmp13), e3 = __fsub_rn(tmp10, tmp13);
    float e1 = __fadd_rn(tmp11, tmp12), e2 = __fsub_rn(tmp11, tmp12);
    float t4 = d[1], t5 = d[3], t6 = d[5], t7 = d[7];
    float z13 = __fadd_rn(t6, t5), z10 = __fsub_rn(t6, t5);
    float z11 = __fadd_rn(t4, t7), z12 = __fsub_rn(t4, t7);
    float o7 = __fadd_rn(z11, z13);
    float o11 = __fmul_rn(__fsub_rn(z11, z13), SQRT2);
    float z5 = __fmul_rn(__fadd_rn(z10, z12), 1.8477591f);
    float o10 = __fsub_rn(__fmul_rn(1.0823922f, z12), z5);
    float o12 = __fadd_rn(__fmul_rn(-2.613126f, z10), z5);
    float o6 = __fsub_rn(o12, o7);
    float o5 = __fsub_rn(o11, o6);
    float o4 = __fadd_rn(o10, o5);
    d[0] = __fadd_rn(e0, o7);
    d[7] = __fsub_rn(e0, o7);
    d[1] = __fadd_rn(e1, o6);
    d[6] = __fsub_rn(e1, o6);
    d[2] = __fadd_rn(e2, o5);
    d[5] = __fsub_rn(e2, o5);
    d[4] = __fadd_rn(e3, o4);
    d[3] = __fsub_rn(e3, o4);
}

// Forward DCT of a block b[y*8+x] held in registers (jpeg/src/dct.rs:61-65): a pass down every column, a
// transpose, a pass down every column again. The result is b[u*8+v] (u horizontal, v vertical frequency).
// Expressed without the transposes: pass 1 transforms columns in place, pass 2 transforms rows and stores the
// result transposed, which is the same set of operations on the same operands.
__device__ __forceinline__ void fdct_block(float* b) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float col[8];
#pragma unroll
        for (int i = 0; i < 8; i++) col[i] = b[i * 8 + j];
        fdct8(col);
#pragma unroll
        for (int i = 0; i < 8; i++) b[i * 8 + j] = col[i];
    }
    // After the transpose, column j of the working block is row j of the current one.
    float t[64];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float row[8];
#pragma unroll
        for (int i = 0; i < 8; i++) row[i] = b[j * 8 + i];
        fdct8(row);
#pragma unroll
        for (int i = 0; i < 8; i++) t[i * 8 + j] = row[i];
    }
#pragma unroll
    for (int i = 0; i < 64; i++) b[i] = t[i];
}

// Inverse DCT (jpeg/src/dct.rs:113-123): pass, transpose, pass, transpose, times 0.125. Input b[v*8+u]
// dequantised and AAN-scaled; output b[y*8+x] level-shifted samples, unrounded. kScale = false leaves out the final
// exact multiplication by 0.125 (the caller folds it into its rounding).
template <bool kScale = true>
__device__ __forceinline__ void idct_block(float* b) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float col[8];
#pragma unroll
        for (int i = 0; i < 8; i++) col[i] = b[i * 8 + j];
        idct8(col);
#pragma unroll
        for (int i = 0; i < 8; i++) b[i * 8 + j] = col[i];
    }
    // transpose + pass down columns + transpose == pass along rows in place.
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float row[8];
#pragma unroll
        for (int i = 0; i < 8; i++) row[i] = b[j * 8 + i];
        idct8(row);
#pragma unroll
        for (int i = 0; i < 8; i++) b[j * 8 + i] = kScale ? __fmul_rn(row[i], 0.125f) : row[i];
    }
}

}  // namespace rcg
