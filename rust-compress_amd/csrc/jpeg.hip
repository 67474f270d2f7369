// JPEG pixel-domain kernels for gfx950: the encode stage of build_coefs and the pixel stage of reconstruct.
//
//   k_encode         RGB8/Gray8 -> (edge-replicated padding) -> YCbCr -> chroma box downsample -> f32 AAN FDCT ->
//                    reciprocal quantisation -> zigzag i16 coefficients; optionally, for each of up to kMaxCand
//                    qualities, dequantisation -> f32 AAN IDCT -> u8 sample planes (the search's reconstruction,
//                    without a coefficient round trip through HBM). jpeg/src/encode.rs:426-539, decode.rs:1312-1343.
//   k_idct           zigzag i16 coefficients of any CoefImage component -> u8 sample plane (decode.rs:1312-1343).
//   k_upsample_color sample planes -> full-size Gray8 or RGB8 ("fancy" triangle upsampling, fixed-point
//                    YCbCr->RGB; decode.rs:1404-1445, 1769-1829).
//
// Layout: one encode workgroup owns a 512-pixel-wide, one-MCU-tall strip. Its RGB rows are read with aligned
// dword loads into LDS, converted there, and each of the 192 threads then transforms one 8x8 block held entirely
// in registers (the AAN butterflies need no cross-lane traffic). Coefficients leave as 16-byte stores, 128 bytes
// per block, in the reference's raster block order so the host wraps them without a permutation.
#include "rcg_device.h"
#include "rcg_host.h"
#include "rcg_jobs.h"
#include "rcg_pixel.h"

namespace rcg {

__device__ __forceinline__ int find_job_by_tile(const uint32_t* begins, int njobs, uint32_t bid) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (begins[mid] <= bid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Byte-exact load of the dword at absolute address a (4-aligned), reading only bytes inside [lo, hi).
__device__ __forceinline__ uint32_t load_dword_guarded(uintptr_t a, uintptr_t lo, uintptr_t hi) {
    if (a >= lo && a + 4 <= hi) return *reinterpret_cast<const uint32_t*>(a);
    uint32_t v = 0;
    for (int k = 0; k < 4; k++) {
        uintptr_t b = a + k;
        if (b >= lo && b < hi) v |= uint32_t(*reinterpret_cast<const uint8_t*>(b)) << (8 * k);
    }
    return v;
}

constexpr int kRawDw = (kEncTileW * 3) / 4 + 2;  // dwords per staged source row (1536 bytes + alignment slack)

__global__ __launch_bounds__(kEncThreads) void k_encode(const EncJob* __restrict__ jobs,
                                                         const uint32_t* __restrict__ tile_begins, int njobs) {
    __shared__ __align__(16) uint32_t s_raw[16][kRawDw];
    __shared__ __align__(16) uint8_t s_y[16][kEncTileW];
    __shared__ __align__(16) uint8_t s_c[2][8][kEncTileW];
    __shared__ uint32_t s_off[16];

    const int tid = threadIdx.x;
    const int j = find_job_by_tile(tile_begins, njobs, blockIdx.x);
    const EncJob& J = jobs[j];
    const uint32_t tile = blockIdx.x - J.tile_begin;
    const uint32_t tx = tile % J.tiles_x, ty = tile / J.tiles_x;
    const int C = J.channels;
    const int hmax = J.hmax, vmax = J.vmax;
    const int rows = 8 * vmax;
    const uint32_t x0 = tx * kEncTileW, y0 = ty * rows;
    const uint32_t w = J.w, h = J.h;
    const uint32_t n = min(uint32_t(kEncTileW), w - x0);  // real source pixels in this strip (x0 < w always)

    // Phase A: stage the strip's source rows (bottom rows replicated) in LDS with aligned dword loads.
    const uintptr_t base = reinterpret_cast<uintptr_t>(J.src), end = base + J.src_bytes;
    const uint32_t ndw = (n * C + 3 + 3) / 4;
    if (tid < rows) {
        const uint32_t sy = min(y0 + tid, h - 1);
        const uintptr_t g0 = base + (uint64_t(sy) * w + x0) * C;
        s_off[tid] = uint32_t(g0 & 3);
    }
    for (uint32_t idx = tid; idx < uint32_t(rows) * ndw; idx += kEncThreads) {
        const uint32_t r = idx / ndw, i = idx - r * ndw;
        const uint32_t sy = min(y0 + r, h - 1);
        const uintptr_t g0 = base + (uint64_t(sy) * w + x0) * C;
        s_raw[r][i] = load_dword_guarded((g0 & ~uintptr_t(3)) + 4 * uintptr_t(i), base, end);
    }
    __syncthreads();

    // Phase B: colour conversion (16-bit fixed point, encode.rs:466-477) and box downsampling (encode.rs:478-499)
    // of one hmax x vmax pixel group per iteration; right-edge replication by clamping the column.
    const uint8_t* raw = reinterpret_cast<const uint8_t*>(&s_raw[0][0]);
    const int gw = kEncTileW / hmax;
    for (int idx = tid; idx < gw * 8; idx += kEncThreads) {
        const int gx = idx % gw, gy = idx / gw;
        int sb = 0, sr = 0;
        for (int dy = 0; dy < vmax; dy++) {
            const int py = gy * vmax + dy;
            const uint8_t* row = raw + py * (kRawDw * 4) + s_off[py];
            for (int dx = 0; dx < hmax; dx++) {
                const int px = gx * hmax + dx;
                const int lx = min(uint32_t(px), n - 1);
                if (C == 1) {
                    s_y[py][px] = row[lx];
                } else {
                    const int32_t r = row[lx * 3], g = row[lx * 3 + 1], b = row[lx * 3 + 2];
                    s_y[py][px] = uint8_t((19595 * r + 38470 * g + 7471 * b + 32768) >> 16);
                    const int32_t cb = (-11059 * r - 21709 * g + 32768 * b + (128 << 16) + 32767) >> 16;
                    const int32_t cr = (32768 * r - 27439 * g - 5329 * b + (128 << 16) + 32767) >> 16;
                    sb += cb & 0xFF;
                    sr += cr & 0xFF;
                }
            }
        }
        if (C == 3) {
            const int nn = hmax * vmax;
            s_c[0][gy][gx] = uint8_t((sb + nn / 2) / nn);
            s_c[1][gy][gx] = uint8_t((sr + nn / 2) / nn);
        }
    }
    __syncthreads();

    // Phase C: one 8x8 block per thread.
    const int nY = 64 * vmax;
    const int cpr = 64 / hmax;  // chroma blocks per component in this strip
    const int nblocks = (C == 1) ? 64 : nY + 2 * cpr;
    if (tid >= nblocks) return;
    const int cls = __builtin_amdgcn_readfirstlane(tid >= nY ? 1 : 0);  // waves never mix luma and chroma
    const uint8_t* srcrow;
    int comp;
    uint32_t bx, by, bw_comp;
    if (tid < nY) {
        const int br = tid / 64, bc = tid % 64;
        comp = 0;
        bx = x0 / 8 + bc;
        by = y0 / 8 + br;
        bw_comp = J.pw / 8;
        if (bx >= bw_comp) return;
        srcrow = &s_y[br * 8][bc * 8];
    } else {
        const int u = tid - nY;
        comp = 1 + u / cpr;
        const int bc = u % cpr;
        bx = x0 / (8 * hmax) + bc;
        by = ty;
        bw_comp = J.pw / hmax / 8;
        if (bx >= bw_comp) return;
        srcrow = &s_c[comp - 1][0][bc * 8];
    }

    float blk[64];
#pragma unroll
    for (int y = 0; y < 8; y++) {
        const uint2 v = *reinterpret_cast<const uint2*>(srcrow + y * kEncTileW);
#pragma unroll
        for (int x = 0; x < 8; x++) {
            const uint32_t word = x < 4 ? v.x : v.y;
            blk[y * 8 + x] = __fsub_rn(float((word >> (8 * (x & 3))) & 0xFF), 128.0f);
        }
    }
    fdct_block(blk);  // blk[u*8+v]

    const uint64_t blk_index = uint64_t(by) * bw_comp + bx;
    for (uint32_t ci = 0; ci < J.ncand; ci++) {
        const EncCand& K = J.cand[ci];
        const float* div = K.div[cls];
        // Quantise: natural position n = v*8+u reads the transposed DCT output at u*8+v (encode.rs:45-54,203-205).
        int16_t* co = K.coef[comp];
        if (co) {
            // Coefficients in zigzag order, packed two per dword as they are computed (no 64-entry array beside blk).
            uint4* dst = reinterpret_cast<uint4*>(co + blk_index * 64);
#pragma unroll
            for (int k8 = 0; k8 < 8; k8++) {
                uint32_t pk[4];
#pragma unroll
                for (int h = 0; h < 4; h++) {
                    uint32_t pair = 0;
#pragma unroll
                    for (int e = 0; e < 2; e++) {
                        const int nidx = zigzag(8 * k8 + 2 * h + e), f = (nidx % 8) * 8 + nidx / 8;
                        int32_t r = round_f32(__fmul_rn(blk[f], div[f]));
                        r = r < -2047 ? -2047 : (r > 2047 ? 2047 : r);
                        pair |= (uint32_t(r) & 0xFFFFu) << (16 * e);
                    }
                    pk[h] = pair;
                }
                dst[k8] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
            }
        }
        float* sp = K.scaled[comp];
        if (sp) {
            // The trellis branch of quantize_plane (encode.rs:206-211): the same f32 products, unrounded.
            float4* dst = reinterpret_cast<float4*>(sp + blk_index * 64);
#pragma unroll
            for (int k = 0; k < 16; k++) {
                float a[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int n = zigzag(4 * k + j), f = (n % 8) * 8 + n / 8;
                    a[j] = __fmul_rn(blk[f], div[f]);
                }
                dst[k] = make_float4(a[0], a[1], a[2], a[3]);
            }
        }
        uint8_t* rp = K.rec[comp];
        if (rp) {
            // Dequantise straight from the rounded values (decode.rs:1317-1322), then the IDCT.
            const float* mult = K.mult[cls];
            float b2[64];
#pragma unroll
            for (int nidx = 0; nidx < 64; nidx++) {
                const int f = (nidx % 8) * 8 + nidx / 8;
                int32_t r = round_f32(__fmul_rn(blk[f], div[f]));
                r = r < -2047 ? -2047 : (r > 2047 ? 2047 : r);
                b2[nidx] = __fmul_rn(float(r), mult[nidx]);
            }
            idct_block(b2);
            const uint32_t stride = bw_comp * 8;
            uint8_t* out = rp + uint64_t(by) * 8 * stride + bx * 8;
#pragma unroll
            for (int y = 0; y < 8; y++) {
                uint32_t lo = 0, hi = 0;
#pragma unroll
                for (int x = 0; x < 8; x++) {
                    int32_t v = round_f32(b2[y * 8 + x]) + 128;
                    v = v < 0 ? 0 : (v > 255 ? 255 : v);
                    if (x < 4) lo |= uint32_t(v) << (8 * x);
                    else hi |= uint32_t(v) << (8 * (x - 4));
                }
                *reinterpret_cast<uint2*>(out + uint64_t(y) * stride) = make_uint2(lo, hi);
            }
        }
    }
}

// ----------------------------------------------------------------------------------------------------------------
// General reconstruct: IDCT of one component plane per grid row y (decode.rs:1312-1343).

struct IdctComp {
    const int16_t* coefs;
    uint8_t* plane;          // stride blocks_w*8
    uint32_t blocks_w, blocks_h;
    float mult[64];          // natural order
    int requant;             // requantise first (jpeg/src/encode.rs:130-136)
    float ratio[64];         // old/new step per zigzag position
    int16_t* requant_out;    // the requantised coefficients, or null
};
struct IdctArgs {
    IdctComp c[4];
};

__global__ __launch_bounds__(256) void k_idct(IdctArgs args) {
    const IdctComp& C = args.c[blockIdx.y];
    const uint32_t nb = C.blocks_w * C.blocks_h;
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nb) return;
    const uint32_t by = b / C.blocks_w, bx = b - by * C.blocks_w;
    const uint4* src = reinterpret_cast<const uint4*>(C.coefs + uint64_t(b) * 64);
    uint32_t pk[32];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        uint4 v = src[k];
        pk[4 * k] = v.x;
        pk[4 * k + 1] = v.y;
        pk[4 * k + 2] = v.z;
        pk[4 * k + 3] = v.w;
    }
    if (C.requant) {
        // DCT-domain requantisation: round_f32(c * old/new), clamped (encode.rs:130-136).
#pragma unroll
        for (int k = 0; k < 64; k++) {
            const int16_t q = int16_t((pk[k / 2] >> (16 * (k & 1))) & 0xFFFF);
            int32_t r = round_f32(__fmul_rn(float(q), C.ratio[k]));
            r = r < -2047 ? -2047 : (r > 2047 ? 2047 : r);
            const uint32_t sh = 16 * (k & 1);
            pk[k / 2] = (pk[k / 2] & ~(0xFFFFu << sh)) | ((uint32_t(r) & 0xFFFFu) << sh);
        }
        if (C.requant_out) {
            uint4* dst = reinterpret_cast<uint4*>(C.requant_out + uint64_t(b) * 64);
#pragma unroll
            for (int k = 0; k < 8; k++) dst[k] = make_uint4(pk[4 * k], pk[4 * k + 1], pk[4 * k + 2], pk[4 * k + 3]);
        }
    }
    float blk[64];
#pragma unroll
    for (int k = 0; k < 64; k++) {
        const int16_t q = int16_t((pk[k / 2] >> (16 * (k & 1))) & 0xFFFF);
        const int nn = zigzag(k);
        blk[nn] = __fmul_rn(float(q), C.mult[nn]);
    }
    idct_block(blk);
    const uint32_t stride = C.blocks_w * 8;
    uint8_t* out = C.plane + uint64_t(by) * 8 * stride + bx * 8;
#pragma unroll
    for (int y = 0; y < 8; y++) {
        uint32_t lo = 0, hi = 0;
#pragma unroll
        for (int x = 0; x < 8; x++) {
            int32_t v = round_f32(blk[y * 8 + x]) + 128;
            v = v < 0 ? 0 : (v > 255 ? 255 : v);
            if (x < 4) lo |= uint32_t(v) << (8 * x);
            else hi |= uint32_t(v) << (8 * (x - 4));
        }
        *reinterpret_cast<uint2*>(out + uint64_t(y) * stride) = make_uint2(lo, hi);
    }
}

struct UpArgs {
    PlaneView comp[3];
    uint32_t ncomp, rgb_direct;
    uint32_t w, h;
    uint8_t* out;
};

// Four horizontally adjacent output pixels per thread.
__global__ __launch_bounds__(256) void k_upsample_color(UpArgs a) {
    const uint32_t y = blockIdx.y;
    const uint32_t xs = (blockIdx.x * 256 + threadIdx.x) * 4;
    if (xs >= a.w) return;
    const uint32_t xe = min(xs + 4, a.w);
    if (a.ncomp == 1) {
        uint8_t* o = a.out + uint64_t(y) * a.w;
        for (uint32_t x = xs; x < xe; x++) o[x] = uint8_t(upsample_at(a.comp[0], x, y));
        return;
    }
    uint8_t* o = a.out + uint64_t(y) * a.w * 3;
    for (uint32_t x = xs; x < xe; x++) {
        const int32_t p0 = upsample_at(a.comp[0], x, y);
        const int32_t p1 = upsample_at(a.comp[1], x, y);
        const int32_t p2 = upsample_at(a.comp[2], x, y);
        int32_t r, g, b;
        if (a.rgb_direct) {
            r = p0; g = p1; b = p2;
        } else {
            ycc_to_rgb(p0, p1, p2, r, g, b);
        }
        o[x * 3] = uint8_t(r);
        o[x * 3 + 1] = uint8_t(g);
        o[x * 3 + 2] = uint8_t(b);
    }
}

// ----------------------------------------------------------------------------------------------------------------
// Launchers.

int launch_encode(rcg_ctx* ctx, const EncJob* d_jobs, const uint32_t* d_begins, int njobs, uint32_t total_tiles) {
    if (total_tiles == 0) return RCG_OK;
    ProfScope ps(ctx, "k_encode");
    hipLaunchKernelGGL(k_encode, dim3(total_tiles), dim3(kEncThreads), 0, ctx->stream, d_jobs, d_begins, njobs);
    RCG_HIP(hipGetLastError());
    return RCG_OK;
}

int launch_idct(rcg_ctx* ctx, const IdctPlane* planes, int n) {
    IdctArgs a;
    std::memset(&a, 0, sizeof(a));
    uint32_t maxb = 0;
    for (int i = 0; i < n; i++) {
        a.c[i].coefs = planes[i].coefs;
        a.c[i].plane = planes[i].plane;
        a.c[i].blocks_w = planes[i].blocks_w;
        a.c[i].blocks_h = planes[i].blocks_h;
        std::memcpy(a.c[i].mult, planes[i].mult, sizeof(a.c[i].mult));
        a.c[i].requant = planes[i].requant;
        std::memcpy(a.c[i].ratio, planes[i].ratio, sizeof(a.c[i].ratio));
        a.c[i].requant_out = planes[i].requant_out;
        maxb = std::max(maxb, planes[i].blocks_w * planes[i].blocks_h);
    }
    if (maxb == 0) return RCG_OK;
    ProfScope ps(ctx, "k_idct");
    hipLaunchKernelGGL(k_idct, dim3((maxb + 255) / 256, n), dim3(256), 0, ctx->stream, a);
    RCG_HIP(hipGetLastError());
    return RCG_OK;
}

int launch_upsample_color(rcg_ctx* ctx, const PlaneView* comps, int ncomp, int rgb_direct, uint32_t w, uint32_t h,
                          uint8_t* out) {
    UpArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int i = 0; i < ncomp; i++) a.comp[i] = comps[i];
    a.ncomp = ncomp;
    a.rgb_direct = rgb_direct;
    a.w = w;
    a.h = h;
    a.out = out;
    ProfScope ps(ctx, "k_upsample_color");
    hipLaunchKernelGGL(k_upsample_color, dim3((w + 1023) / 1024, h), dim3(256), 0, ctx->stream, a);
    RCG_HIP(hipGetLastError());
    return RCG_OK;
}

}  // namespace rcg
