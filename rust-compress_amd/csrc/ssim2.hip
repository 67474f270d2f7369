// SSIMULACRA2 kernels for gfx950 (core/src/ssimulacra2.rs:471-512).
//
// k_ssim_scale<T, NT> computes one scale of one or many image pairs. A workgroup of NT threads owns a tile of
// OW = NT-12 output columns and TH output rows. Thread t owns column x0-6+t and walks down rows y0-5 .. y0+TH+4:
//   1. loads the linear-light RGB of the distorted image at its column (from RGB8 through the sRGB table, from the
//      reconstructed JPEG sample planes with fused upsampling and colour conversion, or from the previous scale's
//      f32 planes), rows above/below the image mirrored (ssimulacra2.rs:385-397);
//   2. converts to XYB (ssimulacra2.rs:251-258) and keeps the last 11 rows in a register ring, beside the ring of
//      the reference image's XYB;
//   3. blurs the per-channel quantities down its column from the rings, then exchanges one row of them through LDS
//      and blurs across the row; columns outside the image hold zeros, which is the reference's horizontal border
//      rule (ssimulacra2.rs:34-40, 278-290);
//   4. evaluates the SSIM, artifact and detail maps and accumulates their sums and 4th-power sums;
//   5. writes the 2x2 linear-light mean for the next scale (ssimulacra2.rs:225-247, summed in f64 in the
//      reference's order, so the pyramid is bit-identical to the reference's f32 planes).
// Each tile writes its 18 sums as f64; k_ssim_finalize reduces them in a fixed order and pools the scales.
//
// T = float is the precision form: f32 XYB shifted by per-channel constants, blurs and maps, f32 sums within a tile
// column and f64 above that. Its reference side (XYB, blur and variance of the reference image) comes from a
// per-image cache that k_ssim_ref writes once per scale, so a candidate costs one colour transform and three blurs
// per channel instead of two and five.
// T = double follows the reference's f64 arithmetic on both images and is used to settle candidates whose f32 score
// lies within the guard band of a search target.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "rcg_host.h"
#include "rcg_jobs.h"
#include "rcg_pixel.h"

namespace rcg {

template <typename T>
struct Taps {
    T k[6];  // k[0] centre, k[t] at distance t (normalised 11-tap Gaussian, sigma 1.5; ssimulacra2.rs:261-271)
};

__constant__ double c_weights[108];

template <typename T>
struct Xyb;

// Multiply-add with one rounding. The blurs and the colour transform use it explicitly (the file is compiled without
// FP contraction), so the reference-cache kernel and the candidate kernel round every shared quantity alike.
__device__ __forceinline__ float mad(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double mad(double a, double b, double c) { return fma(a, b, c); }

// Quantities blurred per channel and the SSIM map computed from them (core/src/ssimulacra2.rs:364-382, 424-433).
//
// f64 follows the reference: mu_a, mu_b, blur(a^2), blur(b^2), blur(ab) and
//   d = 1 - (1 - (mu_a - mu_b)^2) * (2*sigma_ab + C2) / (sigma_aa + sigma_bb + C2).
// f32 takes the reference image's side from its per-scale cache (a, mu_a and sigma_aa + C2, with sigma_aa =
// blur(a^2) - mu_a^2, built once per image by k_ssim_ref) and blurs three quantities of the candidate: mu_b, blur(b^2)
// and R = blur((a-b)^2).
// With dmu = mu_a - mu_b, V = R - dmu^2 (= sigma_aa + sigma_bb - 2*sigma_ab, the local variance of a - b) and
// D = sigma_aa + (blur(b^2) - mu_b^2) + C2 (= sigma_aa + sigma_bb + C2), the same map is
//   d = (V + dmu^2 * (D - V)) / D,
// since 2*sigma_ab + C2 = D - V. V is formed from differences of the two images, not by cancelling large blurred
// products, so the f32 form keeps its precision where the images are close, and V = 0 exactly for identical ones.
template <typename T>
struct Form;
template <>
struct Form<double> {
    static constexpr int NQ = 5;
    static constexpr int MU_B = 1;  // index of mu_b among the blurred quantities
    static __device__ __forceinline__ void centre(double k0, double a0, double b0, double q[5]) {
        q[0] = k0 * a0;
        q[1] = k0 * b0;
        q[2] = k0 * (a0 * a0);
        q[3] = k0 * (b0 * b0);
        q[4] = k0 * (a0 * b0);
    }
    static __device__ __forceinline__ void pair(double kt, double au, double ad, double bu, double bd, double q[5]) {
        q[0] = mad(kt, au + ad, q[0]);
        q[1] = mad(kt, bu + bd, q[1]);
        q[2] = mad(kt, mad(au, au, ad * ad), q[2]);
        q[3] = mad(kt, mad(bu, bu, bd * bd), q[3]);
        q[4] = mad(kt, mad(au, bu, ad * bd), q[4]);
    }
    static __device__ __forceinline__ double ssim(const double v[5]) {
        const double m1 = v[0], m2 = v[1], s11 = v[2], s22 = v[3], s12 = v[4];
        const double C2 = 0.0009;
        const double num_m = 1.0 - (m1 - m2) * (m1 - m2);
        const double num_s = 2.0 * (s12 - m1 * m2) + C2;
        const double denom_s = (s11 - m1 * m1) + (s22 - m2 * m2) + C2;
        return 1.0 - num_m * num_s / denom_s;
    }
};
template <>
struct Form<float> {
    static constexpr int NQ = 3;
    static constexpr int MU_B = 0;
    // The first ring holds e = (a - b)^2 (the reference side comes from the cache), the second b.
    static __device__ __forceinline__ void centre(float k0, float e0, float b0, float q[3]) {
        q[0] = k0 * b0;
        q[1] = k0 * (b0 * b0);
        q[2] = k0 * e0;
    }
    static __device__ __forceinline__ void pair(float kt, float eu, float ed, float bu, float bd, float q[3]) {
        q[0] = mad(kt, bu + bd, q[0]);
        q[1] = mad(kt, mad(bu, bu, bd * bd), q[1]);
        q[2] = mad(kt, eu + ed, q[2]);
    }
    // v: the candidate's blurred quantities; ma, sa: mu_a and sigma_aa + C2 (C2 = 0.0009) from the reference cache.
    static __device__ __forceinline__ float ssim(const float v[3], float ma, float sa) {
        const float m2 = v[0], B2 = v[1], R = v[2];
        const float dmu = ma - m2, dmu2 = dmu * dmu;
        const float V = R - dmu2;
        const float D = sa + mad(-m2, m2, B2);  // sa holds sigma_aa + C2
        return mad(dmu2, D - V, V) * __builtin_amdgcn_rcpf(D);
    }
};

// The edge quotient of the maps (core/src/ssimulacra2.rs:433):
//   d1 = (1 + |b - mu2|) / (1 + |a - mu1|) - 1.
// f64 evaluates it as the reference writes it. f32 uses the equal form (|b - mu2| - |a - mu1|) / (1 + |a - mu1|)
// with a hardware reciprocal: exactly 0 for identical images, and no 1 - ratio cancellation.
template <typename T>
struct Maps;
template <>
struct Maps<float> {
    static __device__ __forceinline__ float edge(float da, float db) {
        return (db - da) * __builtin_amdgcn_rcpf(1.0f + da);
    }
};
template <>
struct Maps<double> {
    static __device__ __forceinline__ double edge(double da, double db) { return (1.0 + db) / (1.0 + da) - 1.0; }
};

// The f32 form works on XYB shifted by per-channel constants (X - 0.42, Y - 0.51, B - 0.55). Every map depends on
// the values only through differences and variances (mu1 - mu2, a - mu1, blur(a^2) - mu1^2, blur(ab) - mu1*mu2),
// which a constant shift leaves unchanged provided the samples outside the image hold -offset instead of 0; the
// shift keeps blur(a^2) and mu1^2 near 0.01 instead of 0.3, so their f32 difference keeps ~5 more bits.
template <>
struct Xyb<float> {
    static __device__ __forceinline__ float off(int c) { return c == 0 ? 0.42f : (c == 1 ? 0.51f : 0.55f); }
    static __device__ __forceinline__ float cbrt_pos(float x) {
        // x >= 0; exp2(log2(x)/3) on the transcendental unit. log2(0) = -inf gives 0.
        return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 3.0f));
    }
    static __device__ __forceinline__ void eval(float r, float g, float b, float& X, float& Y, float& B) {
        const float M00 = 0.30f, M02 = 0.078f, M01 = float(1.0 - 0.078 - 0.30);
        const float M10 = 0.23f, M12 = 0.078f, M11 = float(1.0 - 0.078 - 0.23);
        const float M20 = float(0.2434226892454782), M21 = float(0.2047674442449682);
        const float M22 = float(1.0 - 0.2434226892454782 - 0.2047674442449682);
        const float BIAS = float(0.0037930732552754493);
        const float CB = float(0.15595420054924863);  // cbrt(BIAS), rounded to f32
        // With c_i the cube roots and m_i = c_i - CB (the reference's opsin values): x = (m0 - m1)/2, y = (m0 + m1)/2,
        // X = 14x, Y = y + 0.01, B = (m2 - y) + 0.55. CB cancels from X and B, and the shifts fold into constants.
        const float c0 = cbrt_pos(fmaxf(mad(M00, r, mad(M01, g, mad(M02, b, BIAS))), 0.0f));
        const float c1 = cbrt_pos(fmaxf(mad(M10, r, mad(M11, g, mad(M12, b, BIAS))), 0.0f));
        const float c2 = cbrt_pos(fmaxf(mad(M20, r, mad(M21, g, mad(M22, b, BIAS))), 0.0f));
        const float s = c0 + c1;
        X = 7.0f * (c0 - c1);          // 14x, shifted: (14x + 0.42) - 0.42
        Y = mad(0.5f, s, -(CB + 0.5f));  // (y + 0.01) - 0.51
        B = mad(-0.5f, s, c2);           // ((m2 - y) + 0.55) - 0.55
    }
};

template <>
struct Xyb<double> {
    static __device__ __forceinline__ double off(int) { return 0.0; }
    static __device__ __forceinline__ void eval(double r, double g, double b, double& X, double& Y, double& B) {
        const double M00 = 0.30, M02 = 0.078, M01 = 1.0 - M02 - M00;
        const double M10 = 0.23, M12 = 0.078, M11 = 1.0 - M12 - M10;
        const double M20 = 0.2434226892454782, M21 = 0.2047674442449682, M22 = 1.0 - M20 - M21;
        const double BIAS = 0.0037930732552754493;
        const double CB = cbrt(BIAS);
        const double m0 = cbrt(fmax(M00 * r + M01 * g + M02 * b + BIAS, 0.0)) - CB;
        const double m1 = cbrt(fmax(M10 * r + M11 * g + M12 * b + BIAS, 0.0)) - CB;
        const double m2 = cbrt(fmax(M20 * r + M21 * g + M22 * b + BIAS, 0.0)) - CB;
        const double x = 0.5 * (m0 - m1), y = 0.5 * (m0 + m1);
        X = x * 14.0 + 0.42;
        Y = y + 0.01;
        B = (m2 - y) + 0.55;
    }
};

#define RCG_GLOBAL __attribute__((address_space(1)))

// Loads and stores through the global address space (global_load/global_store rather than flat accesses: the
// pointers live in job tables, so the compiler cannot infer their address space by itself).
template <class V>
__device__ __forceinline__ V gload(const V* p) {
    return *(const RCG_GLOBAL V*)p;
}
template <class V>
__device__ __forceinline__ void gstore(V* p, V v) {
    *(RCG_GLOBAL V*)p = v;
}
// A pointer that is uniform over the wave (a row pointer built from the workgroup's row and job fields), passed
// through readfirstlane so that it is an opaque scalar value: the compiler then keeps each access as that base plus
// a 32-bit offset (the scalar-base form of the global instructions) instead of reassociating base + plane stride +
// offset into one 64-bit vector add per access. Only for uniform pointers: a per-lane pointer would take lane 0's
// value.
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return reinterpret_cast<T*>((uint64_t(hi) << 32) | lo);
}
// Element at byte offset `byte_off` of a uniform row pointer.
template <class V>
__device__ __forceinline__ V gload_b(const V* row, uint32_t byte_off) {
    return *(const RCG_GLOBAL V*)(reinterpret_cast<const char*>(uniform_ptr(row)) + byte_off);
}
template <class V>
__device__ __forceinline__ void gstore_b(V* row, uint32_t byte_off, V v) {
    *(RCG_GLOBAL V*)(reinterpret_cast<char*>(uniform_ptr(row)) + byte_off) = v;
}
// gstore_b with the non-temporal hint (the nt bit of the store).
template <class V>
__device__ __forceinline__ void gstore_nt_b(V* row, uint32_t byte_off, V v) {
    __builtin_nontemporal_store(v, (RCG_GLOBAL V*)(reinterpret_cast<char*>(uniform_ptr(row)) + byte_off));
}

// Row loaders of linear-light RGB. begin_row() takes the row the whole workgroup is on (uniform), so row pointers
// are scalar. prep(x) computes the thread's byte offsets within a row once, for its column x; fresh() re-issues them
// at the top of every row (an empty asm, no instruction), so the compiler forms each access as a uniform row
// pointer plus a 32-bit offset (scalar-base addressing) rather than hoisting 64-bit offsets out of the walk and
// adding them to the row pointer with a 64-bit vector add per access. load() reads the thread's column of the
// current row. The source kind is a template parameter: each kernel instance carries only the loader it uses.
template <uint32_t K>
struct Loader;

// Gray8/RGB8 bytes through the sRGB table (core/src/ssimulacra2.rs:186-207).
template <>
struct Loader<kSrcBytes> {
    const uint8_t* base;
    const uint8_t* row;
    uint32_t w, ch;
    uint32_t ox;  // byte offset of the thread's column
    __device__ __forceinline__ void init(const SsimSrc& S, uint32_t w_, uint32_t) {
        base = S.bytes;
        ch = S.channels;
        w = w_;
    }
    __device__ __forceinline__ void prep(uint32_t x) { ox = ch == 1 ? x : 3 * x; }
    __device__ __forceinline__ void fresh() { asm volatile("" : "+v"(ox)); }
    __device__ __forceinline__ void begin_row(uint32_t r) { row = base + uint64_t(r) * w * ch; }
    __device__ __forceinline__ void load(const float* lut, float out[3]) const {
        if (ch == 1) {
            const float l = lut[gload_b(row, ox)];
            out[0] = out[1] = out[2] = l;
        } else {
            out[0] = lut[gload_b(row, ox)];
            out[1] = lut[gload_b(row + 1, ox)];
            out[2] = lut[gload_b(row + 2, ox)];
        }
    }
};

// The previous scale's planar f32 pyramid level.
template <>
struct Loader<kSrcPlanes> {
    const float* base;
    const float* row;
    uint64_t n;
    uint32_t w;
    uint32_t ox;
    __device__ __forceinline__ void init(const SsimSrc& S, uint32_t w_, uint32_t h) {
        base = S.planes;
        w = w_;
        n = uint64_t(w_) * h;
    }
    __device__ __forceinline__ void prep(uint32_t x) { ox = 4 * x; }
    __device__ __forceinline__ void fresh() { asm volatile("" : "+v"(ox)); }
    __device__ __forceinline__ void begin_row(uint32_t r) { row = base + uint64_t(r) * w; }
    __device__ __forceinline__ void load(const float*, float out[3]) const {
        out[0] = gload_b(row, ox);
        out[1] = gload_b(row + n, ox);
        out[2] = gload_b(row + 2 * n, ox);
    }
};

// Reconstructed JPEG sample planes of the search path with reconstruct's upsampling (jpeg/src/decode.rs:1769-1829)
// and fixed-point colour conversion (decode.rs:1434-1441) applied on the fly, then the sRGB table. These planes come
// from build_coefs layouts: luma is full resolution and both chroma planes share one geometry (1x1, 2x1 or 2x2).
template <>
struct Loader<kSrcRecon> {
    const uint8_t* py;
    const uint8_t* pc[2];
    uint32_t stride_y, stride_c, cw, ch, fx, fy, ncomp;
    const uint8_t* ry;
    const uint8_t* ra[2];   // chroma rows of the current row: nearer
    const uint8_t* rb[2];   // and farther (equal to ra unless vertically subsampled)
    // The thread's luma column, and its chroma column and the farther chroma column packed as 16-bit halves: the
    // candidate kernel holds its registers at the limit of four waves, and unpacking (and the rounding term of the
    // column's parity, from ox) costs a few instructions per row.
    uint32_t ox, ocp;
    __device__ __forceinline__ void init(const SsimSrc& S, uint32_t, uint32_t) {
        ncomp = S.ncomp;
        py = S.comp[0].p;
        stride_y = S.comp[0].stride;
        pc[0] = S.comp[1].p;
        pc[1] = S.comp[2].p;
        stride_c = S.comp[1].stride;
        cw = S.comp[1].cw;
        ch = S.comp[1].ch;
        fx = S.comp[1].fx;
        fy = S.comp[1].fy;
    }
    __device__ __forceinline__ void prep(uint32_t x) {
        ox = x;
        uint32_t oi = x, ofi = x;
        if (fx != 1) {
            oi = x >> 1;
            ofi = (x & 1) == 0 ? (oi > 0 ? oi - 1 : 0) : min(oi + 1, cw - 1);
        }
        ocp = oi | (ofi << 16);  // chroma planes are at most 32768 samples wide
    }
    __device__ __forceinline__ void fresh() { asm volatile("" : "+v"(ox), "+v"(ocp)); }
    __device__ __forceinline__ void begin_row(uint32_t r) {
        ry = py + uint64_t(r) * stride_y;
        uint64_t c0, c1;
        if (fy == 2) {
            const uint32_t jj = r >> 1;
            const uint32_t far = (r & 1) == 0 ? (jj > 0 ? jj - 1 : 0) : min(jj + 1, ch - 1);
            c0 = uint64_t(jj) * stride_c;
            c1 = uint64_t(far) * stride_c;
        } else {
            c0 = c1 = uint64_t(r) * stride_c;
        }
        for (int k = 0; k < 2; k++) {
            ra[k] = pc[k] + c0;
            rb[k] = pc[k] + c1;
        }
    }
    __device__ __forceinline__ int32_t chroma(int k, uint32_t oi, uint32_t ofi) const {
        const uint8_t* a = ra[k];
        const uint8_t* b = rb[k];
        const int32_t odd = int32_t(ox & 1);
        if (fx == 1) return gload_b(a, oi);
        if (fy == 2) {
            const int32_t nearv = 3 * int32_t(gload_b(a, oi)) + int32_t(gload_b(b, oi));
            const int32_t far = 3 * int32_t(gload_b(a, ofi)) + int32_t(gload_b(b, ofi));
            return (3 * nearv + far + 8 - odd) >> 4;
        }
        return (3 * int32_t(gload_b(a, oi)) + int32_t(gload_b(a, ofi)) + 1 + odd) >> 2;
    }
    __device__ __forceinline__ void load(const float* lut, float out[3]) const {
        const int32_t y = gload_b(ry, ox);
        if (ncomp == 1) {
            out[0] = out[1] = out[2] = lut[y];
            return;
        }
        int32_t r, g, b;
        const uint32_t oi = ocp & 0xFFFFu, ofi = ocp >> 16;
        ycc_to_rgb(y, chroma(0, oi, ofi), chroma(1, oi, ofi), r, g, b);
        out[0] = lut[r];
        out[1] = lut[g];
        out[2] = lut[b];
    }
};

// The reference image's per-scale cache written by k_ssim_ref, 9 floats per pixel in planes of w x h elements: M =
// float4 (mu_X, mu_Y, mu_B, s_X), S = float2 (s_Y, s_B), then A_X, A_Y, A_B = float, at kCacheS, kCacheA floats per
// pixel from the start; a is the shifted XYB, mu = blur(a) and s = sigma_aa + C2 = blur(a^2) - mu^2 + 0.0009. Every
// store and read of a lane covers 16, 8 or 4 consecutive bytes, so each access of a wave spans whole lines. The input
// rows read a alone (three planes, as the ring needs), the output rows all nine values in five loads.
constexpr uint32_t kCacheS = 4, kCacheA = 6;  // plane offsets in floats per pixel of the image
typedef float f32x4 __attribute__((ext_vector_type(4)));  // native vectors (HIP's float4 and float2 are structs)
typedef float f32x2 __attribute__((ext_vector_type(2)));
template <>
struct Loader<kSrcRefCache> {
    const float* base;
    const float* row;
    uint64_t n;
    uint32_t w;
    uint32_t ox;    // byte offset of the lane's pixel in the float planes (twice and four times it in S and M)
    uint64_t orow;  // element index of the current output row's first pixel
    __device__ __forceinline__ void init(const SsimSrc& S, uint32_t w_, uint32_t h) {
        base = S.planes;
        w = w_;
        n = uint64_t(w_) * h;
    }
    __device__ __forceinline__ void prep(uint32_t x) { ox = 4 * x; }
    __device__ __forceinline__ void fresh() { asm volatile("" : "+v"(ox)); }
    __device__ __forceinline__ void begin_row(uint32_t r) { row = base + kCacheA * n + uint64_t(r) * w; }
    __device__ __forceinline__ void load(const float*, float out[3]) const {
        out[0] = gload_b(row, ox);
        out[1] = gload_b(row + n, ox);
        out[2] = gload_b(row + 2 * n, ox);
    }
    // a, mu_a and sigma_aa + C2 at an output row.
    __device__ __forceinline__ void begin_out_row(uint32_t r) { orow = uint64_t(r) * w; }
    __device__ __forceinline__ void load_stats(float a[3], float mu[3], float sig[3]) const {
        const f32x4 m = gload_b(reinterpret_cast<const f32x4*>(base) + orow, ox << 2);
        const f32x2 sv = gload_b(reinterpret_cast<const f32x2*>(base + kCacheS * n) + orow, ox << 1);
        const float* ar = base + kCacheA * n + orow;
#pragma unroll
        for (int c = 0; c < 3; c++) a[c] = gload_b(ar + c * n, ox);
        mu[0] = m.x;
        mu[1] = m.y;
        mu[2] = m.z;
        sig[0] = m.w;
        sig[1] = sv.x;
        sig[2] = sv.y;
    }
};

// Exchange with the neighbouring lane of the pair (x ^ 1) through DPP (quad_perm [1,0,3,2]).
__device__ __forceinline__ float swap_pair(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}

// Where a thread keeps the first row of its current 2x2 pyramid blocks (k = c: left column of channel c, k = 3 + c:
// the right one). The candidate kernels keep it in a per-thread LDS slot, registers being their scarcer resource;
// the reference kernel, short of LDS occupancy rather than registers, keeps it in registers.
template <int NT>
struct PyrLds {
    float* p;  // the thread's slot: element k at p[k * NT]
    __device__ __forceinline__ void set(int k, float v) { p[k * NT] = v; }
    __device__ __forceinline__ float get(int k) const { return p[k * NT]; }
};
struct PyrRegs {
    float v[6];
    __device__ __forceinline__ void set(int k, float x) { v[k] = x; }
    __device__ __forceinline__ float get(int k) const { return v[k]; }
};

// LDS layout of the vertically blurred quantities exchanged for the horizontal blur (two row buffers). The f32
// kernels interleave them per column (quantity q of column x at x * QS + q, QS odd: the lanes of a wave hit distinct
// banks), so all taps of all quantities lie within the immediate-offset reach of one base address; the f64 kernel
// keeps one plane per quantity.
template <int NQS, int NT, bool kInterleaved>
struct QBuf {
    static constexpr int QS = kInterleaved ? (NQS | 1) : NQS;     // column stride (interleaved) or plane count
    static constexpr int kBuf = QS * NT;                         // elements per row buffer
    static constexpr int kCol = kInterleaved ? QS : 1;           // distance between neighbouring columns
    static __device__ __forceinline__ int at(int buf, int q, int x) {
        return buf * kBuf + (kInterleaved ? x * QS + q : q * NT + x);
    }
    static constexpr int qoff(int q) { return kInterleaved ? q : q * NT; }  // at(buf, q, x) - at(buf, 0, x)
    // Index of the lowest element a horizontal 11-tap window of column x reads (valid for x >= 5), made opaque to the
    // compiler: taps addressed from it have non-negative constant offsets, which the LDS instructions take as
    // immediates, where offsets below a centre column would each need a base register of their own.
    static __device__ __forceinline__ int window(int buf, int x) {
        int lo = at(buf, 0, x) - 5 * kCol;
        asm volatile("" : "+v"(lo));
        return lo;
    }
};

// Per-workgroup constants and per-thread state of the scale kernel's row walk. The f32 kernel reads the reference
// side from its cache (KA = kSrcRefCache); the f64 kernel computes both sides from pixels or pyramid planes.
template <typename T, int NT, uint32_t KA, uint32_t KB>
struct Walk {
    static constexpr bool kCached = KA == kSrcRefCache;
    static_assert(kCached == (sizeof(T) == 4), "the f32 kernel reads the reference cache; the f64 kernel does not");
    Loader<KA> la;
    Loader<KB> lb;
    float* a_next;
    float* b_next;
    const float* lut;
    const T* taps;
    using Q = QBuf<3 * Form<T>::NQ, NT, sizeof(T) == 4>;
    T* s_q;
    int tid, w, h, x0, y0, gx, y_end, ylast, nw, nh;
    uint32_t po;  // byte offset of the lane's column pair in the next pyramid level
    bool in_img, is_out, want_next, ds_writer;
    int yy;
    T ra[11][3], rb[11][3];  // XYB rings of the reference (cached: (a - b)^2 instead) and of the distorted image
    T acc[18];
    PyrLds<NT> pv;  // first row of the current 2x2 pyramid blocks
};

// Next-scale 2x2 mean of one image row (core/src/ssimulacra2.rs:225-247). x0v/x1v hold the pair's left and right
// columns; the first row of a block is kept in pv and the block is written on its second row (or on the last row of
// an odd-height image), summed in f64 row by row, left to right, as the reference does.
template <bool kUniformDst, class P>
__device__ __forceinline__ void pyramid_row(const float x0v[3], const float x1v[3], P& pv, bool first_row,
                                            bool emit, int gx, int yy, int w, int nw, int nh, uint32_t po,
                                            float* dst) {
    if (first_row) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            pv.set(c, x0v[c]);
            pv.set(3 + c, x1v[c]);
        }
    }
    if (!emit || !dst) return;
    const int gx0 = gx & ~1;
    const bool two_cols = gx0 + 1 < w;
    const int oy = yy >> 1;  // uniform: yy is the workgroup's row; po = 4 * (gx0 / 2) is the lane's byte offset
    const uint64_t n2 = uint64_t(nw) * nh;
    float* orow = dst + uint64_t(oy) * nw;
    // Dividing by the pixel count (1, 2 or 4), as the reference does, is the exact multiplication by its inverse.
    const double inv = first_row ? (two_cols ? 0.5 : 1.0) : (two_cols ? 0.25 : 0.5);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double sum = double(pv.get(c));
        if (two_cols) sum += double(pv.get(3 + c));
        if (!first_row) {
            sum += double(x0v[c]);
            if (two_cols) sum += double(x1v[c]);
        }
        if constexpr (kUniformDst)
            gstore_b(orow + c * n2, po, float(sum * inv));
        else
            gstore(reinterpret_cast<float*>(reinterpret_cast<char*>(orow + c * n2) + po), float(sum * inv));
    }
}

// One input row of the walk, in ring slot S (compile-time, so the ring stays in registers). Returns false once the
// walk is past its last row.
template <int S, typename T, int NT, uint32_t KA, uint32_t KB>
__device__ __forceinline__ bool walk_row(Walk<T, NT, KA, KB>& st) {
#pragma clang fp contract(fast)
    constexpr bool kCached = Walk<T, NT, KA, KB>::kCached;
    const int yy = st.yy;
    if (yy > st.ylast) return false;
    const int h = st.h;
    const int r = yy < 0 ? -yy - 1 : (yy >= h ? 2 * h - 1 - yy : yy);
    float la[3] = {0, 0, 0}, lb[3] = {0, 0, 0};
    st.la.begin_row(uint32_t(r));
    st.lb.begin_row(uint32_t(r));
    st.la.fresh();
    st.lb.fresh();
    asm volatile("" : "+v"(st.po));
    // Beyond the left/right edges the reference's samples are 0, i.e. -offset in the shifted domain, for both images
    // (so their difference is 0).
    T xa[3], xb[3];
#pragma unroll
    for (int c = 0; c < 3; c++) xa[c] = xb[c] = -Xyb<T>::off(c);
    if (st.in_img) {
        st.la.load(st.lut, la);
        st.lb.load(st.lut, lb);
        Xyb<T>::eval(T(lb[0]), T(lb[1]), T(lb[2]), xb[0], xb[1], xb[2]);
        if constexpr (kCached) {
#pragma unroll
            for (int c = 0; c < 3; c++) xa[c] = la[c];
        } else {
            Xyb<T>::eval(T(la[0]), T(la[1]), T(la[2]), xa[0], xa[1], xa[2]);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        st.rb[S][c] = xb[c];
        if constexpr (kCached) {
            const T e = xa[c] - xb[c];
            st.ra[S][c] = e * e;
        } else {
            st.ra[S][c] = xa[c];
        }
    }
    // Next-scale 2x2 mean (uniform condition: the row is one of this tile's own rows). Uncached, the two lanes of a
    // column pair split the work: the even lane sums the reference's block, the odd lane the distorted one's, each
    // after one DPP exchange per channel of the value its partner needs. Cached, only the distorted image's level is
    // written, by the even lane.
    if (st.want_next && yy >= st.y0 && yy < st.y_end) {
        const bool odd = (st.gx & 1) != 0;
        float x0v[3], x1v[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float recv = swap_pair(kCached ? lb[c] : (odd ? la[c] : lb[c]));
            const float own = (kCached || odd) ? lb[c] : la[c];
            x0v[c] = odd ? recv : own;
            x1v[c] = odd ? own : recv;
        }
        const bool first_row = ((yy - st.y0) & 1) == 0;
        const bool emit = (!first_row || yy == h - 1) && st.ds_writer;
        // Cached, every lane writes the distorted image's level (a uniform pointer); uncached, each lane of a pair
        // writes its own image's level.
        float* dst = odd ? st.b_next : (kCached ? st.b_next : st.a_next);
        pyramid_row<kCached>(x0v, x1v, st.pv, first_row, emit, st.gx, yy, st.w, st.nw, st.nh, st.po, dst);
    }
    if (yy >= st.y0 + 5) {
        // Vertical blur for output row yy-5: centre in slot (S+6)%11, distance t in (S+6-t)%11 and (S+6+t)%11.
        constexpr int SC = (S + 6) % 11;
        constexpr int NQ = Form<T>::NQ;
        const int buf = (yy - 5) & 1;
        const T* k = st.taps;
        const int tid = st.tid;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            T q[NQ];
            Form<T>::centre(k[0], st.ra[SC][c], st.rb[SC][c], q);
#pragma unroll
            for (int t = 1; t <= 5; t++) {
                const int su = (S + 6 - t + 11) % 11, sd = (S + 6 + t) % 11;
                Form<T>::pair(k[t], st.ra[su][c], st.ra[sd][c], st.rb[su][c], st.rb[sd][c], q);
            }
#pragma unroll
            for (int i = 0; i < NQ; i++) st.s_q[Walk<T, NT, KA, KB>::Q::at(buf, c * NQ + i, tid)] = q[i];
        }
        __syncthreads();
        if (st.is_out) {
            float ac[3] = {0, 0, 0}, ma[3] = {0, 0, 0}, sa[3] = {0, 0, 0};
            if constexpr (kCached) {
                // The reference's a, mu_a and sigma_aa at this output pixel, loaded ahead of the horizontal blur.
                st.la.begin_out_row(uint32_t(yy - 5));
                st.la.load_stats(ac, ma, sa);
            }
            using Q = typename Walk<T, NT, KA, KB>::Q;
            const T* win = &st.s_q[Q::window(buf, tid)];  // output lanes only: tid >= 6
#pragma unroll
            for (int c = 0; c < 3; c++) {
                // One channel at a time: the scheduler may not pull the next channel's LDS reads ahead of this
                // channel's arithmetic, which would keep all 99 taps live at once.
                __builtin_amdgcn_sched_barrier(0);
                T v[NQ];
#pragma unroll
                for (int i = 0; i < NQ; i++) {
                    const T* row = win + Q::qoff(c * NQ + i) + 5 * Q::kCol;
                    T sum = k[0] * row[0];
#pragma unroll
                    for (int t = 1; t <= 5; t++) sum = mad(k[t], row[-t * Q::kCol] + row[t * Q::kCol], sum);
                    v[i] = sum;
                }
                T a0, m1, d;
                if constexpr (kCached) {
                    a0 = ac[c];
                    m1 = ma[c];
                    d = Form<T>::ssim(v, ma[c], sa[c]);
                } else {
                    a0 = st.ra[SC][c];
                    m1 = v[0];
                    d = Form<T>::ssim(v);
                }
                const T m2 = v[Form<T>::MU_B];
                d = d > T(0) ? d : T(0);
                const T d1 = Maps<T>::edge(fabs(a0 - m1), fabs(st.rb[SC][c] - m2));
                const T art = d1 > T(0) ? d1 : T(0);
                const T det = -d1 > T(0) ? -d1 : T(0);
                const T d2 = d * d, a2 = art * art, e2 = det * det;
                st.acc[c * 6 + 0] += d;
                st.acc[c * 6 + 1] += d2 * d2;
                st.acc[c * 6 + 2] += art;
                st.acc[c * 6 + 3] += a2 * a2;
                st.acc[c * 6 + 4] += det;
                st.acc[c * 6 + 5] += e2 * e2;
            }
        }
    }
    st.yy = yy + 1;
    return true;
}

template <typename T, int NT, uint32_t KA, uint32_t KB>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1))) void k_ssim_scale(
    const SsimJob* __restrict__ jobs, const Unit* __restrict__ units, Taps<T> taps, const float* __restrict__ lut_g) {
    constexpr int OW = NT - 12;
    constexpr int NWAVES = NT / 64;
    __shared__ T s_q[2 * Walk<T, NT, KA, KB>::Q::kBuf];
    __shared__ float s_lut[256];
    __shared__ double s_red[NWAVES][18];
    __shared__ float s_pyr[6 * NT];

    const Unit u = units ? units[blockIdx.x] : Unit{0u, blockIdx.x};  // no table: one job, tiles in order
    if (u.job == kIdleJob) return;  // uniform over the workgroup, before any barrier
    const int tid = threadIdx.x;
    for (int i = tid; i < 256; i += NT) s_lut[i] = lut_g[i];

    const SsimJob& J = jobs[u.job];
    const uint32_t tile = u.tile;
    const uint32_t tx = tile % J.tiles_x, ty = tile / J.tiles_x;
    Walk<T, NT, KA, KB> st;
    st.la.init(J.a, J.w, J.h);
    st.lb.init(J.b, J.w, J.h);
    st.a_next = J.a_next;
    st.b_next = J.b_next;
    st.lut = s_lut;
    st.taps = taps.k;
    st.s_q = s_q;
    st.tid = tid;
    st.w = int(J.w);
    st.h = int(J.h);
    st.x0 = int(J.x_begin) + int(tx) * OW;
    const int TH = int(J.th);
    st.y0 = int(ty) * TH;
    st.gx = st.x0 - 6 + tid;
    st.in_img = st.gx >= 0 && st.gx < st.w;
    st.is_out = tid >= 6 && tid < 6 + OW && st.gx < st.w;
    // Offsets of the lane's column (used only where the lane is inside the image).
    st.la.prep(uint32_t(st.gx));
    st.lb.prep(uint32_t(st.gx));
    st.po = 4u * (uint32_t(st.gx & ~1) >> 1);
    st.y_end = min(st.y0 + TH, st.h);
    st.ylast = st.y_end + 4;
    st.want_next = st.a_next != nullptr || st.b_next != nullptr;
    st.nw = (st.w + 1) >> 1;
    st.nh = (st.h + 1) >> 1;
    // A column pair's lanes write when the pair's left column is one of this tile's output columns inside the image:
    // uncached both (the even one the reference's plane, the odd one the distorted one's), cached the even one.
    st.ds_writer = st.want_next && tid >= 6 && tid < 6 + OW && (st.gx & ~1) < st.w &&
                   (!Walk<T, NT, KA, KB>::kCached || (st.gx & 1) == 0);
    st.yy = st.y0 - 5;
#pragma unroll
    for (int i = 0; i < 18; i++) st.acc[i] = T(0);
    st.pv.p = s_pyr + tid;
#pragma unroll
    for (int k = 0; k < 6; k++) st.pv.set(k, 0.0f);
    __syncthreads();

    while (walk_row<0>(st) && walk_row<1>(st) && walk_row<2>(st) && walk_row<3>(st) && walk_row<4>(st) &&
           walk_row<5>(st) && walk_row<6>(st) && walk_row<7>(st) && walk_row<8>(st) && walk_row<9>(st) &&
           walk_row<10>(st)) {
    }

    // Workgroup reduction of the 18 sums in f64.
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int i = 0; i < 18; i++) {
        double v = double(st.acc[i]);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) s_red[wave][i] = v;
    }
    __syncthreads();
    if (tid < 18) {
        double v = 0.0;
#pragma unroll
        for (int wv = 0; wv < NWAVES; wv++) v += s_red[wv][tid];
        gstore(J.partials + uint64_t(tile) * 18 + tid, v);
    }
}

// ----------------------------------------------------------------------------------------------------------------
// Reference cache. k_ssim_ref<NT, KA> walks one image at one scale with the same tiles and border rules as the scale
// kernel and writes, for every pixel, the quantities of the reference side that do not depend on a candidate: the
// shifted XYB a (planes 0-2), mu_a = blur(a) (3-5) and sigma_aa + C2 = blur(a^2) - mu_a^2 + 0.0009 (6-8); it also
// writes the image's next pyramid level. Every f32 candidate evaluation of the image then reads these 9 floats per
// pixel instead of recomputing the reference's colour transform and two of its five blurs. The values at a pixel do
// not depend on the tiling, so the cache is the same whichever tile policy built it.

template <int NT, uint32_t KA>
struct RefWalk {
    Loader<KA> la;
    float* a_next;
    float* cache;
    const float* lut;
    const float* taps;
    using Q = QBuf<6, NT, false>;  // plane per quantity: the interleaved form needs more registers than it saves here
    float* s_q;
    uint64_t n;
    int tid, w, h, x0, y0, gx, y_end, ylast, nw, nh;
    uint32_t oc, po;  // byte offsets of the lane's column in the cache's float planes and in the next pyramid level
    bool in_img, is_out, ds_writer;
    int yy;
    float ra[11][3];
    PyrRegs pv;  // first row of the current 2x2 pyramid blocks
};

template <int S, int NT, uint32_t KA>
__device__ __forceinline__ bool ref_row(RefWalk<NT, KA>& st) {
    const int yy = st.yy;
    if (yy > st.ylast) return false;
    const int h = st.h;
    const int r = yy < 0 ? -yy - 1 : (yy >= h ? 2 * h - 1 - yy : yy);
    float la[3] = {0, 0, 0};
    st.la.begin_row(uint32_t(r));
    st.la.fresh();
    asm volatile("" : "+v"(st.oc), "+v"(st.po));
    if (st.in_img) {
        st.la.load(st.lut, la);
        Xyb<float>::eval(la[0], la[1], la[2], st.ra[S][0], st.ra[S][1], st.ra[S][2]);
    } else {
#pragma unroll
        for (int c = 0; c < 3; c++) st.ra[S][c] = -Xyb<float>::off(c);
    }
    if (st.a_next && yy >= st.y0 && yy < st.y_end) {
        // The even lane of a column pair sums the 2x2 block after one DPP exchange per channel.
        float x0v[3], x1v[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            x0v[c] = la[c];
            x1v[c] = swap_pair(la[c]);
        }
        const bool first_row = ((yy - st.y0) & 1) == 0;
        const bool emit = (!first_row || yy == h - 1) && st.ds_writer;
        pyramid_row<true>(x0v, x1v, st.pv, first_row, emit, st.gx, yy, st.w, st.nw, st.nh, st.po, st.a_next);
    }
    if (yy >= st.y0 + 5) {
        constexpr int SC = (S + 6) % 11;
        const int buf = (yy - 5) & 1;
        const float* k = st.taps;
        const int tid = st.tid;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float a0 = st.ra[SC][c];
            float m = k[0] * a0, q = k[0] * (a0 * a0);
#pragma unroll
            for (int t = 1; t <= 5; t++) {
                const float au = st.ra[(S + 6 - t + 11) % 11][c], ad = st.ra[(S + 6 + t) % 11][c];
                m = mad(k[t], au + ad, m);
                q = mad(k[t], mad(au, au, ad * ad), q);
            }
            st.s_q[RefWalk<NT, KA>::Q::at(buf, 2 * c, tid)] = m;
            st.s_q[RefWalk<NT, KA>::Q::at(buf, 2 * c + 1, tid)] = q;
        }
        __syncthreads();
        if (st.is_out) {
            const uint64_t orow = uint64_t(yy - 5) * st.w;  // uniform row; the lane's column is the offset
            float mu[3], sg[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float v[2];
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    // Planes of NT floats: a plane's offset does not fit the paired reads' immediates, so each
                    // quantity keeps its own base (window() does not help here).
                    using Q = typename RefWalk<NT, KA>::Q;
                    const float* row = &st.s_q[Q::at(buf, 2 * c + i, tid)];
                    float sum = k[0] * row[0];
#pragma unroll
                    for (int t = 1; t <= 5; t++) sum = mad(k[t], row[-t * Q::kCol] + row[t * Q::kCol], sum);
                    v[i] = sum;
                }
                mu[c] = v[0];
                sg[c] = mad(-v[0], v[0], v[1]) + 0.0009f;  // sigma_aa + C2
            }
            // Five stores in the layout of Loader<kSrcRefCache>, each covering whole lines across the wave (nine 4-byte
            // planes took the launch 8.5 ms on the 256-image batch, these five 7.0 ms). Non-temporal: the cache is
            // written once and read back only by later launches (on one 4096x4096 image with six candidates, where this
            // launch runs beside the encode launch, the whole step was about 2% shorter than with ordinary stores).
            f32x4 m;
            m.x = mu[0];
            m.y = mu[1];
            m.z = mu[2];
            m.w = sg[0];
            f32x2 sv;
            sv.x = sg[1];
            sv.y = sg[2];
            gstore_nt_b(reinterpret_cast<f32x4*>(st.cache) + orow, st.oc << 2, m);
            gstore_nt_b(reinterpret_cast<f32x2*>(st.cache + kCacheS * st.n) + orow, st.oc << 1, sv);
#pragma unroll
            for (int c = 0; c < 3; c++) gstore_nt_b(st.cache + (kCacheA + c) * st.n + orow, st.oc, st.ra[SC][c]);
        }
    }
    st.yy = yy + 1;
    return true;
}

template <int NT, uint32_t KA>
__global__ __launch_bounds__(NT) void k_ssim_ref(
    const RefJob* __restrict__ jobs, const Unit* __restrict__ units, Taps<float> taps, const float* __restrict__ lut_g) {
    constexpr int OW = NT - 12;
    __shared__ float s_q[2 * RefWalk<NT, KA>::Q::kBuf];
    __shared__ float s_lut[256];
    const Unit u = units ? units[blockIdx.x] : Unit{0u, blockIdx.x};  // no table: one job, tiles in order
    if (u.job == kIdleJob) return;
    const int tid = threadIdx.x;
    for (int i = tid; i < 256; i += NT) s_lut[i] = lut_g[i];

    const RefJob& J = jobs[u.job];
    const uint32_t tile = u.tile;
    const uint32_t tx = tile % J.tiles_x, ty = tile / J.tiles_x;
    RefWalk<NT, KA> st;
    st.la.init(J.a, J.w, J.h);
    st.a_next = J.a_next;
    st.cache = J.cache;
    st.lut = s_lut;
    st.taps = taps.k;
    st.s_q = s_q;
    st.n = uint64_t(J.w) * J.h;
    st.tid = tid;
    st.w = int(J.w);
    st.h = int(J.h);
    st.x0 = int(tx) * OW;
    st.y0 = int(ty) * int(J.th);
    st.gx = st.x0 - 6 + tid;
    st.in_img = st.gx >= 0 && st.gx < st.w;
    st.is_out = tid >= 6 && tid < 6 + OW && st.gx < st.w;
    st.la.prep(uint32_t(st.gx));
    st.oc = 4u * uint32_t(st.gx);
    st.po = 4u * (uint32_t(st.gx & ~1) >> 1);
    st.y_end = min(st.y0 + int(J.th), st.h);
    st.ylast = st.y_end + 4;
    st.nw = (st.w + 1) >> 1;
    st.nh = (st.h + 1) >> 1;
    st.ds_writer = tid >= 6 && tid < 6 + OW && st.gx < st.w && (st.gx & 1) == 0;
    st.yy = st.y0 - 5;
#pragma unroll
    for (int k = 0; k < 6; k++) st.pv.set(k, 0.0f);
    __syncthreads();
    while (ref_row<0>(st) && ref_row<1>(st) && ref_row<2>(st) && ref_row<3>(st) && ref_row<4>(st) &&
           ref_row<5>(st) && ref_row<6>(st) && ref_row<7>(st) && ref_row<8>(st) && ref_row<9>(st) &&
           ref_row<10>(st)) {
    }
}

// One workgroup per image pair, one wave per scale: fixed-order f64 reduction of each scale's tile sums, then the
// mean and 4-norm of each map (ssimulacra2.rs:457-465), the weighted pooling and the final mapping (:491-511).
constexpr int kFinThreads = 384;

__global__ __launch_bounds__(kFinThreads) void k_ssim_finalize(const FinalDesc* __restrict__ fj) {
    const FinalDesc& F = fj[blockIdx.x];
    __shared__ double s_part[6][18][65];  // per scale, statistic and lane (odd stride: no bank conflicts)
    __shared__ double s_sum[6][18];
    __shared__ double s_stat[6][18];
    const int tid = threadIdx.x, lane = tid & 63, sc = tid >> 6;
    const uint32_t ns = F.nscales;
    // Lane l of wave s sums tiles l, l+64, ... of scale s in that order (each tile's 18 sums are one contiguous
    // 144-byte record; four tiles are loaded ahead of their additions). The lanes' sums are then added in lane order.
    if (uint32_t(sc) < ns) {
        const double* p = F.partials[sc];
        const uint32_t nt = F.ntiles[sc];
        double acc[18];
#pragma unroll
        for (int i = 0; i < 18; i++) acc[i] = 0.0;
        uint32_t t = lane;
        for (; t + 3 * 64 < nt; t += 4 * 64) {
            double2 v[4][9];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const double2* rec = reinterpret_cast<const double2*>(p + uint64_t(t + u * 64) * 18);
#pragma unroll
                for (int i = 0; i < 9; i++) v[u][i] = rec[i];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int i = 0; i < 9; i++) {
                    acc[2 * i] += v[u][i].x;
                    acc[2 * i + 1] += v[u][i].y;
                }
        }
        for (; t < nt; t += 64) {
            const double2* rec = reinterpret_cast<const double2*>(p + uint64_t(t) * 18);
#pragma unroll
            for (int i = 0; i < 9; i++) {
                const double2 v = rec[i];
                acc[2 * i] += v.x;
                acc[2 * i + 1] += v.y;
            }
        }
#pragma unroll
        for (int i = 0; i < 18; i++) s_part[sc][i][lane] = acc[i];
    }
    __syncthreads();
    if (tid < int(ns) * 18) {
        const int s2 = tid / 18, i = tid % 18;
        const double* row = s_part[s2][i];
        double v = 0.0;
        for (int l = 0; l < 64; l++) v += row[l];
        s_sum[s2][i] = v;
    }
    __syncthreads();
    // One thread per (scale, statistic): mean for even entries, 4-norm for odd ones.
    for (int k = tid; k < int(ns) * 18; k += kFinThreads) {
        const int sc = k / 18, i = k % 18;
        const double v = s_sum[sc][i] / F.count[sc];
        s_stat[sc][i] = (i & 1) ? pow(v, 0.25) : v;
    }
    __syncthreads();
    if (tid != 0) return;
    double ssim = 0.0;
    int w = 0;
    for (int c = 0; c < 3; c++)
        for (uint32_t sc = 0; sc < ns; sc++)
            for (int n = 0; n < 2; n++) {
                ssim += c_weights[w++] * fabs(s_stat[sc][c * 6 + n]);
                ssim += c_weights[w++] * fabs(s_stat[sc][c * 6 + 2 + n]);
                ssim += c_weights[w++] * fabs(s_stat[sc][c * 6 + 4 + n]);
            }
    ssim *= 0.9562382616834844;
    ssim = 2.326765642916932 * ssim - 0.020884521182843837 * ssim * ssim + 6.248496625763138e-5 * ssim * ssim * ssim;
    const double score = ssim > 0.0 ? 100.0 - 10.0 * pow(ssim, 0.6276336467831387) : 100.0;
    *F.score_out = score;
    if (F.stats_out)
        for (uint32_t sc = 0; sc < ns; sc++)
            for (int i = 0; i < 18; i++) F.stats_out[sc * 18 + i] = s_stat[sc][i];
}

// ----------------------------------------------------------------------------------------------------------------
// Launchers.

int upload_weights() {
    static const double w[108] = {
#include "ssim2_weights.inc"
    };
    RCG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_weights), w, sizeof(w)));
    return RCG_OK;
}

template <typename T, int NT, uint32_t KA, uint32_t KB>
static void launch_one(rcg_ctx* ctx, hipStream_t stream, const SsimJob* d_jobs, const Unit* d_units, uint32_t nunits,
                       const Taps<T>& t) {
    hipLaunchKernelGGL((k_ssim_scale<T, NT, KA, KB>), dim3(nunits), dim3(NT), 0, stream, d_jobs, d_units, t,
                       ctx->d_lut);
}

template <typename T, int NT>
static int launch_kinds(rcg_ctx* ctx, hipStream_t st, const SsimJob* d_jobs, const Unit* d_units, uint32_t nunits,
                        const Taps<T>& t, uint32_t ka, uint32_t kb) {
    if constexpr (sizeof(T) == 4) {
        if (ka == kSrcRefCache && kb == kSrcRecon)
            launch_one<T, NT, kSrcRefCache, kSrcRecon>(ctx, st, d_jobs, d_units, nunits, t);
        else if (ka == kSrcRefCache && kb == kSrcBytes)
            launch_one<T, NT, kSrcRefCache, kSrcBytes>(ctx, st, d_jobs, d_units, nunits, t);
        else if (ka == kSrcRefCache && kb == kSrcPlanes)
            launch_one<T, NT, kSrcRefCache, kSrcPlanes>(ctx, st, d_jobs, d_units, nunits, t);
        else {
            set_error("internal: unsupported SSIMULACRA2 source combination");
            return RCG_DEVICE;
        }
    } else {
        if (ka == kSrcBytes && kb == kSrcRecon)
            launch_one<T, NT, kSrcBytes, kSrcRecon>(ctx, st, d_jobs, d_units, nunits, t);
        else if (ka == kSrcBytes && kb == kSrcBytes)
            launch_one<T, NT, kSrcBytes, kSrcBytes>(ctx, st, d_jobs, d_units, nunits, t);
        else if (ka == kSrcPlanes && kb == kSrcPlanes)
            launch_one<T, NT, kSrcPlanes, kSrcPlanes>(ctx, st, d_jobs, d_units, nunits, t);
        else {
            set_error("internal: unsupported SSIMULACRA2 source combination");
            return RCG_DEVICE;
        }
    }
    return RCG_OK;
}

int launch_ssim_scale(rcg_ctx* ctx, int precision, const SsimJob* d_jobs, const Unit* d_units, uint32_t nunits, int nt,
                      uint32_t kind_a, uint32_t kind_b, hipStream_t stream) {
    if (nunits == 0) return RCG_OK;
    const hipStream_t st = stream ? stream : ctx->stream;
    ProfScope ps(ctx, precision == RCG_PREC_FP64 ? "k_ssim_scale_f64" : "k_ssim_scale", st);
    int rc;
    if (precision == RCG_PREC_FP64) {
        Taps<double> t;
        for (int i = 0; i < 6; i++) t.k[i] = ctx->taps64[i];
        rc = launch_kinds<double, 128>(ctx, st, d_jobs, d_units, nunits, t, kind_a, kind_b);
    } else {
        Taps<float> t;
        for (int i = 0; i < 6; i++) t.k[i] = ctx->taps32[i];
        rc = nt == 64    ? launch_kinds<float, 64>(ctx, st, d_jobs, d_units, nunits, t, kind_a, kind_b)
             : nt == 128 ? launch_kinds<float, 128>(ctx, st, d_jobs, d_units, nunits, t, kind_a, kind_b)
                         : launch_kinds<float, 256>(ctx, st, d_jobs, d_units, nunits, t, kind_a, kind_b);
    }
    if (rc != RCG_OK) return rc;
    RCG_HIP(hipGetLastError());
    return RCG_OK;
}

int ref_kernel_slots(int device, uint32_t* wide, uint32_t* narrow) {
    hipDeviceProp_t prop;
    RCG_HIP(hipGetDeviceProperties(&prop, device));
    int b256 = 0, b128 = 0;
    RCG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b256, k_ssim_ref<256, kSrcBytes>, 256, 0));
    RCG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b128, k_ssim_ref<128, kSrcBytes>, 128, 0));
    *wide = uint32_t(std::max(1, b256 * prop.multiProcessorCount));
    *narrow = uint32_t(std::max(1, b128 * prop.multiProcessorCount));
    return RCG_OK;
}

int launch_ssim_ref(rcg_ctx* ctx, const RefJob* d_jobs, const Unit* d_units, uint32_t nunits, int nt,
                    uint32_t kind_a, hipStream_t stream) {
    if (nunits == 0) return RCG_OK;
    if (kind_a != kSrcBytes && kind_a != kSrcPlanes) {
        set_error("internal: unsupported reference-cache source");
        return RCG_DEVICE;
    }
    ProfScope ps(ctx, "k_ssim_ref", stream);
    Taps<float> t;
    for (int i = 0; i < 6; i++) t.k[i] = ctx->taps32[i];
    const dim3 grid(nunits);
    if (nt == 128) {
        if (kind_a == kSrcBytes)
            hipLaunchKernelGGL((k_ssim_ref<128, kSrcBytes>), grid, dim3(128), 0, stream, d_jobs, d_units, t,
                               ctx->d_lut);
        else
            hipLaunchKernelGGL((k_ssim_ref<128, kSrcPlanes>), grid, dim3(128), 0, stream, d_jobs, d_units,
                               t, ctx->d_lut);
    } else {
        if (kind_a == kSrcBytes)
            hipLaunchKernelGGL((k_ssim_ref<256, kSrcBytes>), grid, dim3(256), 0, stream, d_jobs, d_units, t,
                               ctx->d_lut);
        else
            hipLaunchKernelGGL((k_ssim_ref<256, kSrcPlanes>), grid, dim3(256), 0, stream, d_jobs, d_units,
                               t, ctx->d_lut);
    }
    RCG_HIP(hipGetLastError());
    return RCG_OK;
}

int launch_finalize(rcg_ctx* ctx, const FinalDesc* d_jobs, int n) {
    if (n == 0) return RCG_OK;
    ProfScope ps(ctx, "k_ssim_finalize");
    hipLaunchKernelGGL(k_ssim_finalize, dim3(n), dim3(kFinThreads), 0, ctx->stream, d_jobs);
    RCG_HIP(hipGetLastError());
    return RCG_OK;
}

}  // namespace rcg
