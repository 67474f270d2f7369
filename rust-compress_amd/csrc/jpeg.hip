// JPEG pixel-domain kernels for gfx950: the encode stage of build_coefs and the pixel stage of reconstruct.
//
//   k_encode         RGB8/Gray8 -> (edge-replicated padding) -> YCbCr -> chroma box downsample -> f32 AAN FDCT ->
//                    reciprocal quantisation -> zigzag i16 coefficients; optionally, for each of up to kMaxCand
//                    qualities, dequantisation -> f32 AAN IDCT -> u8 sample planes (the search's reconstruction,
//                    without a coefficient round trip through HBM). jpeg/src/encode.rs:426-539, decode.rs:1312-1343.
//   k_idct           zigzag i16 coefficients of any CoefImage component -> u8 sample plane (decode.rs:1312-1343).
//   k_upsample_color sample planes -> full-size Gray8 or RGB8 ("fancy" triangle upsampling, fixed-point
//                    YCbCr->RGB, or cmyk_to_rgb for four components; decode.rs:1404-1478, 1769-1829).
//
// Layout: one encode workgroup owns a 512-pixel-wide, one-MCU-tall strip. Its RGB rows are read as aligned 16-byte
// chunks into LDS (all of a thread's loads in flight at once), converted there by a per-layout routine, and each of
// the 192 threads then transforms one 8x8 block held entirely in registers (the AAN butterflies need no cross-lane
// traffic). Coefficients leave as 16-byte stores, 128 bytes per block, in the reference's raster block order so the
// host wraps them without a permutation.
#include "rcg_device.h"
#include "rcg_host.h"
#include "rcg_jobs.h"
#include "rcg_pixel.h"

namespace rcg {

// A descriptor pointer that is uniform over the wave, seen through the constant address space: loads through it are
// invariant for the kernel, so the compiler issues them as scalar loads into SGPRs (the scalar cache) instead of
// per-lane vector loads that each wait for a round trip to L2. Only for read-only descriptors and uniform addresses.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* const_desc(const T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return reinterpret_cast<const __attribute__((address_space(4))) T*>((uint64_t(hi) << 32) | lo);
}

__device__ __forceinline__ int find_job_by_tile(const uint32_t* begins, int njobs, uint32_t bid) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (begins[mid] <= bid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Byte-exact load of the 16 bytes at absolute address a (16-aligned), reading only bytes inside [lo, hi); the
// bytes outside read as zero.
__device__ __forceinline__ uint4 load_chunk_guarded(uintptr_t a, uintptr_t lo, uintptr_t hi) {
    uint32_t v[4] = {0, 0, 0, 0};
    for (int k = 0; k < 16; k++) {
        const uintptr_t b = a + k;
        if (b >= lo && b < hi) v[k >> 2] |= uint32_t(*reinterpret_cast<const uint8_t*>(b)) << (8 * (k & 3));
    }
    return make_uint4(v[0], v[1], v[2], v[3]);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // a native 16-byte vector: kept in registers when
                                                             // held in a constant-indexed array

constexpr int kRawChunks = (kEncTileW * 3 + 15 + 15) / 16;  // 16-byte chunks per staged source row (97)
constexpr int kRawRow = kRawChunks * 16;                     // LDS bytes per staged source row
constexpr int kRawLoads = (16 * kRawChunks + kEncThreads - 1) / kEncThreads;  // chunk loads per thread, 16 rows
// LDS of one workgroup, carved per subsampling mode: the Y rows (rows x 512), the two chroma planes (2 x 8 x 512/hmax),
// then the staged source rows (rows x kRawRow). 4:2:0 is the largest: 16 Y rows, 256-wide chroma, 16 source rows.
// Its last 16 bytes lie beyond every carve-up: the discard slot of the source staging.
constexpr int kEncLds = 16 * kEncTileW + 2 * 8 * (kEncTileW / 2) + 16 * kRawRow + 16;
static_assert(kEncLds - 16 >= 8 * kEncTileW + 2 * 8 * kEncTileW + 8 * kRawRow, "4:4:4 carve-up must fit");
static_assert(kEncLds + 64 <= 40960, "four encode workgroups per CU");

// The 4 bytes at LDS byte offset o (any alignment), from the two dwords that hold them.
__device__ __forceinline__ uint32_t lds_bytes4(const uint8_t* base, uint32_t o) {
    const uint32_t* d = reinterpret_cast<const uint32_t*>(base + (o & ~3u));
    return __builtin_amdgcn_alignbyte(d[1], d[0], o & 3u);
}

// Gray8: the staged rows are the Y rows; 4 samples per item.
__device__ __forceinline__ void convert_gray(const uint8_t* s_raw, const uint32_t* s_off, uint8_t* s_y, int tid) {
    for (int it = tid; it < 8 * (kEncTileW / 4); it += kEncThreads) {
        const int py = it / (kEncTileW / 4), q = it % (kEncTileW / 4);
        *reinterpret_cast<uint32_t*>(s_y + py * kEncTileW + 4 * q) =
            lds_bytes4(s_raw, uint32_t(py) * kRawRow + s_off[py] + 4 * uint32_t(q));
    }
}

// RGB8 -> YCbCr in 16-bit fixed point (encode.rs:466-477) and the HM x VM box mean of Cb and Cr (encode.rs:478-499).
// One item is 4 chroma samples: 4*HM pixels of each of VM rows, read as dwords and realigned with v_alignbyte; its Y
// bytes and chroma samples leave as dword stores.
template <int HM, int VM>
__device__ __forceinline__ void convert_rgb(const uint8_t* s_raw, const uint32_t* s_off, uint8_t* s_y, uint8_t* s_c,
                                            int tid) {
    constexpr int CW = kEncTileW / HM;  // chroma samples per row
    constexpr int QPR = CW / 4;         // items per chroma row
    constexpr int DW = 3 * HM;          // dwords of RGB per row of one item
    constexpr int NSH = (HM - 1) + (VM - 1);
    constexpr int HALF = (1 << NSH) >> 1;
    for (int it = tid; it < 8 * QPR; it += kEncThreads) {
        const int gy = it / QPR, q = it % QPR;
        int sb[4] = {0, 0, 0, 0}, sr[4] = {0, 0, 0, 0};
#pragma unroll
        for (int dy = 0; dy < VM; dy++) {
            const int py = gy * VM + dy;
            const uint32_t o = uint32_t(py) * kRawRow + s_off[py] + uint32_t(q) * (12 * HM);
            const uint32_t* d = reinterpret_cast<const uint32_t*>(s_raw + (o & ~3u));
            uint32_t wd[DW + 1], a[DW];
#pragma unroll
            for (int k = 0; k <= DW; k++) wd[k] = d[k];
#pragma unroll
            for (int k = 0; k < DW; k++) a[k] = __builtin_amdgcn_alignbyte(wd[k + 1], wd[k], o & 3u);
            uint32_t ypk[HM];
#pragma unroll
            for (int p = 0; p < 4 * HM; p++) {
                const int32_t r = (a[(3 * p) / 4] >> (8 * ((3 * p) % 4))) & 0xFF;
                const int32_t g = (a[(3 * p + 1) / 4] >> (8 * ((3 * p + 1) % 4))) & 0xFF;
                const int32_t b = (a[(3 * p + 2) / 4] >> (8 * ((3 * p + 2) % 4))) & 0xFF;
                const uint32_t y = uint32_t((19595 * r + 38470 * g + 7471 * b + 32768) >> 16) & 0xFF;
                const int32_t cb = (-11059 * r - 21709 * g + 32768 * b + (128 << 16) + 32767) >> 16;
                const int32_t cr = (32768 * r - 27439 * g - 5329 * b + (128 << 16) + 32767) >> 16;
                ypk[p / 4] = (p % 4 == 0) ? y : (ypk[p / 4] | (y << (8 * (p % 4))));
                sb[p / HM] += cb & 0xFF;
                sr[p / HM] += cr & 0xFF;
            }
#pragma unroll
            for (int k = 0; k < HM; k++)
                *reinterpret_cast<uint32_t*>(s_y + py * kEncTileW + q * (4 * HM) + 4 * k) = ypk[k];
        }
        uint32_t cbp = 0, crp = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            // The rounded mean of HM*VM samples (a power of two; the sums are non-negative).
            cbp |= (uint32_t((sb[j] + HALF) >> NSH) & 0xFF) << (8 * j);
            crp |= (uint32_t((sr[j] + HALF) >> NSH) & 0xFF) << (8 * j);
        }
        *reinterpret_cast<uint32_t*>(s_c + gy * CW + 4 * q) = cbp;
        *reinterpret_cast<uint32_t*>(s_c + 8 * CW + gy * CW + 4 * q) = crp;
    }
}

__global__ __launch_bounds__(kEncThreads) void k_encode(const EncJob* __restrict__ jobs,
                                                         const uint32_t* __restrict__ tile_begins, int njobs) {
    __shared__ __align__(16) uint8_t s_mem[kEncLds];
    __shared__ uint32_t s_off[16];

    const int tid = threadIdx.x;
    const int j = __builtin_amdgcn_readfirstlane(find_job_by_tile(tile_begins, njobs, blockIdx.x));
    const auto& J = const_desc(jobs)[j];
    const uint32_t tile = blockIdx.x - J.tile_begin;
    const uint32_t tx = tile % J.tiles_x, ty = tile / J.tiles_x;
    const int C = J.channels;
    const int hmax = J.hmax, vmax = J.vmax;  // each 1 or 2
    const int rows = 8 * vmax;
    const uint32_t x0 = tx * kEncTileW, y0 = ty * rows;
    const uint32_t w = J.w, h = J.h;
    const uint32_t n = min(uint32_t(kEncTileW), w - x0);  // real source pixels in this strip (x0 < w always)
    const int cw = kEncTileW / hmax;                       // chroma samples per row of this strip
    uint8_t* const s_y = s_mem;
    uint8_t* const s_c = s_mem + rows * kEncTileW;
    uint8_t* const s_raw = s_c + 2 * 8 * cw;

    // Phase A: stage the strip's source rows (bottom rows replicated) in LDS. Row r starts at byte s_off[r] of its
    // LDS row and is read as the aligned 16-byte chunks that cover it.
    const uintptr_t base = reinterpret_cast<uintptr_t>(J.src), end = base + J.src_bytes;
    const uint32_t nbytes = n * C;
    const uint32_t nchunks = uint32_t(rows) * kRawChunks;
    auto row_start = [&](uint32_t r) {
        return base + (uint64_t(min(y0 + r, h - 1)) * w + x0) * C;
    };
    const uintptr_t first = row_start(0) & ~uintptr_t(15);
    const uintptr_t last = (row_start(rows - 1) + nbytes + 15) & ~uintptr_t(15);
    if (first >= base && last <= end) {
        // Every chunk of the strip lies inside the source buffer: unconditional 16-byte loads, all in flight before
        // the first is used (a lane without a chunk of its own reloads the strip's first and discards it).
        const uint32_t raw_off = uint32_t(s_raw - s_mem);
        u32x4 v[kRawLoads];
        uint32_t lds[kRawLoads];
#pragma unroll
        for (int k = 0; k < kRawLoads; k++) {
            const uint32_t c = tid + k * kEncThreads;
            const uint32_t r = c / kRawChunks, i = c - r * kRawChunks;
            const uintptr_t g0 = row_start(r);
            const uintptr_t a = (g0 & ~uintptr_t(15)) + 16 * uintptr_t(i);
            const bool ok = c < nchunks && a < g0 + nbytes;
            lds[k] = ok ? raw_off + r * kRawRow + 16 * i : kEncLds - 16;  // else the discard slot
            v[k] = *reinterpret_cast<const u32x4*>(ok ? a : first);
        }
#pragma unroll
        for (int k = 0; k < kRawLoads; k++) *reinterpret_cast<u32x4*>(s_mem + lds[k]) = v[k];
    } else {
        // A strip at either end of the source buffer: a chunk that reaches outside it goes byte by byte.
        for (uint32_t c = tid; c < nchunks; c += kEncThreads) {
            const uint32_t r = c / kRawChunks, i = c - r * kRawChunks;
            const uintptr_t g0 = row_start(r);
            const uintptr_t a = (g0 & ~uintptr_t(15)) + 16 * uintptr_t(i);
            if (a < g0 + nbytes)
                *reinterpret_cast<uint4*>(s_raw + r * kRawRow + 16 * i) =
                    (a >= base && a + 16 <= end) ? *reinterpret_cast<const uint4*>(a) : load_chunk_guarded(a, base, end);
        }
    }
    if (tid < rows) s_off[tid] = uint32_t(row_start(tid) & 15);
    __syncthreads();
    if (n < uint32_t(kEncTileW)) {
        // The strip ends inside the tile: replicate each row's last pixel to the tile's width (the right-edge padding,
        // x.min(w - 1), of encode.rs:453-458 and 466-470), so that the conversion reads whole groups of pixels.
        const uint32_t pad = (kEncTileW - n) * C;
        for (uint32_t idx = tid; idx < uint32_t(rows) * pad; idx += kEncThreads) {
            const uint32_t r = idx / pad, k = idx - r * pad;
            uint8_t* row = s_raw + r * kRawRow + s_off[r];
            row[nbytes + k] = row[nbytes - C + k % C];
        }
        __syncthreads();
    }

    // Phase B: colour conversion and chroma box downsampling into the Y rows and chroma planes.
    if (C == 1)
        convert_gray(s_raw, s_off, s_y, tid);
    else if (hmax == 1)
        convert_rgb<1, 1>(s_raw, s_off, s_y, s_c, tid);
    else if (vmax == 1)
        convert_rgb<2, 1>(s_raw, s_off, s_y, s_c, tid);
    else
        convert_rgb<2, 2>(s_raw, s_off, s_y, s_c, tid);
    __syncthreads();

    // Phase C: one 8x8 block per thread.
    const int nY = 64 * vmax;
    const int cpr = 64 / hmax;  // chroma blocks per component in this strip
    const int nblocks = (C == 1) ? 64 : nY + 2 * cpr;
    if (tid >= nblocks) return;
    const int cls = __builtin_amdgcn_readfirstlane(tid >= nY ? 1 : 0);  // waves never mix luma and chroma
    const uint8_t* srcrow;
    int comp, lds_stride;
    uint32_t bx, by, bw_comp;
    if (tid < nY) {
        const int br = tid / 64, bc = tid % 64;
        comp = 0;
        bx = x0 / 8 + bc;
        by = y0 / 8 + br;
        bw_comp = J.pw / 8;
        if (bx >= bw_comp) return;
        srcrow = s_y + br * 8 * kEncTileW + bc * 8;
        lds_stride = kEncTileW;
    } else {
        const int u = tid - nY;
        comp = 1 + u / cpr;
        const int bc = u % cpr;
        bx = x0 / (8 * hmax) + bc;
        by = ty;
        bw_comp = J.pw / hmax / 8;
        if (bx >= bw_comp) return;
        srcrow = s_c + (comp - 1) * 8 * cw + bc * 8;
        lds_stride = cw;
    }

    float blk[64];
#pragma unroll
    for (int y = 0; y < 8; y++) {
        const uint2 px8 = *reinterpret_cast<const uint2*>(srcrow + y * lds_stride);
#pragma unroll
        for (int x = 0; x < 8; x++) {
            const uint32_t word = x < 4 ? px8.x : px8.y;
            blk[y * 8 + x] = __fsub_rn(float((word >> (8 * (x & 3))) & 0xFF), 128.0f);
        }
    }
    fdct_block(blk);  // blk[u*8+v]

    const uint64_t blk_index = uint64_t(by) * bw_comp + bx;
    const auto* cands = const_desc(J.cand);
    for (uint32_t ci = 0; ci < J.ncand; ci++) {
        const auto& K = cands[ci];
        const auto* div = K.div[cls];
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
            const auto* mult = K.mult[cls];
            float b2[64];
#pragma unroll
            for (int nidx = 0; nidx < 64; nidx++) {
                // float(clamp(round_f32(x), -2047, 2047)) computed in f32: (x + 1.5*2^23) - 1.5*2^23 is x rounded half
                // to even, exactly, for |x| < 2^22 (the quantised DCT values here stay below 2^15).
                const int f = (nidx % 8) * 8 + nidx / 8;
                const float r = __fsub_rn(__fadd_rn(__fmul_rn(blk[f], div[f]), 12582912.0f), 12582912.0f);
                b2[nidx] = __fmul_rn(__builtin_amdgcn_fmed3f(r, -2047.0f, 2047.0f), mult[nidx]);
            }
            idct_block<false>(b2);  // 8x the samples: the 1/8 is folded into the rounding below
            const uint32_t stride = bw_comp * 8;
            uint8_t* out = rp + uint64_t(by) * 8 * stride + bx * 8;
#pragma unroll
            for (int y = 0; y < 8; y++) {
                uint32_t lo = 0, hi = 0;
#pragma unroll
                for (int x = 0; x < 8; x++) {
                    // round_f32(x * 0.125) + 128: x + 1.5*2^26 rounds x to a multiple of 8, ties to the even multiple
                    // of 16, which is round-half-even of x/8; its bits less those of 1.5*2^26 count eighths. Exact for
                    // |x| < 2^25; here |x| <= 57172 (any block of 8-bit samples, any table of 8-bit steps).
                    int32_t v = __float_as_int(__fadd_rn(b2[y * 8 + x], 100663296.0f)) - (0x4CC00000 - 128);
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
    PlaneView comp[4];
    uint32_t ncomp, rgb_direct;
    int adobe_transform;     // four components: -1 without an Adobe marker
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
    if (a.ncomp == 4) {
        for (uint32_t x = xs; x < xe; x++) {
            int32_t r, g, b;
            cmyk_to_rgb(upsample_at(a.comp[0], x, y), upsample_at(a.comp[1], x, y), upsample_at(a.comp[2], x, y),
                        upsample_at(a.comp[3], x, y), a.adobe_transform, r, g, b);
            o[x * 3] = uint8_t(r);
            o[x * 3 + 1] = uint8_t(g);
            o[x * 3 + 2] = uint8_t(b);
        }
        return;
    }
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
    if (n < 1 || n > 4) {
        set_error("internal: the IDCT launch takes 1 to 4 components");
        return RCG_DEVICE;
    }
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
                          uint8_t* out, int adobe_transform) {
    if (ncomp != 1 && ncomp != 3 && ncomp != 4) {
        set_error("internal: pixels are made from 1, 3 or 4 components");
        return RCG_DEVICE;
    }
    UpArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int i = 0; i < ncomp; i++) a.comp[i] = comps[i];
    a.ncomp = ncomp;
    a.rgb_direct = rgb_direct;
    a.adobe_transform = adobe_transform;
    a.w = w;
    a.h = h;
    a.out = out;
    ProfScope ps(ctx, "k_upsample_color");
    hipLaunchKernelGGL(k_upsample_color, dim3((w + 1023) / 1024, h), dim3(256), 0, ctx->stream, a);
    RCG_HIP(hipGetLastError());
    return RCG_OK;
}

}  // namespace rcg
