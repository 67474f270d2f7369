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
// LDS of one workgroup: the staged source rows (rows x kRawRow) and, over the same bytes once the conversion has read
// them into registers, the Y rows (rows x 512) and the two chroma planes (2 x 8 x 512/hmax). 16 staged rows (4:2:0)
// are the largest part; the last 16 bytes lie beyond it: the discard slot of the source staging.
constexpr int kEncLds = 16 * kRawRow + 16;
static_assert(16 * kRawRow >= 16 * kEncTileW + 2 * 8 * (kEncTileW / 2), "4:2:0 Y rows and chroma over the staging");
static_assert(16 * kRawRow >= 8 * kEncTileW + 2 * 8 * kEncTileW, "4:4:4 Y rows and chroma over the staging");
static_assert(kEncLds + 64 <= 160 * 1024 / 6, "six encode workgroups per CU");

// The 4 bytes at LDS byte offset o (any alignment), from the two dwords that hold them.
__device__ __forceinline__ uint32_t lds_bytes4(const uint8_t* base, uint32_t o) {
    const uint32_t* d = reinterpret_cast<const uint32_t*>(base + (o & ~3u));
    return __builtin_amdgcn_alignbyte(d[1], d[0], o & 3u);
}

// Gray8: the staged rows are the Y rows; 4 samples per item. Every item is read into registers before the barrier
// after which the Y rows overwrite the staged rows (the same LDS bytes).
__device__ __forceinline__ void convert_gray(const uint8_t* s_raw, const uint32_t* s_off, uint8_t* s_y, int tid) {
    constexpr int NI = 8 * (kEncTileW / 4), NIT = (NI + kEncThreads - 1) / kEncThreads;
    uint32_t v[NIT];
#pragma unroll
    for (int k = 0; k < NIT; k++) {
        const int it = tid + k * kEncThreads;
        if (it < NI) {
            const int py = it / (kEncTileW / 4), q = it % (kEncTileW / 4);
            v[k] = lds_bytes4(s_raw, uint32_t(py) * kRawRow + s_off[py] + 4 * uint32_t(q));
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NIT; k++) {
        const int it = tid + k * kEncThreads;
        if (it < NI) {
            const int py = it / (kEncTileW / 4), q = it % (kEncTileW / 4);
            *reinterpret_cast<uint32_t*>(s_y + py * kEncTileW + 4 * q) = v[k];
        }
    }
}

// RGB8 -> YCbCr in 16-bit fixed point (encode.rs:466-477) and the HM x VM box mean of Cb and Cr (encode.rs:478-499).
// One item is 4 chroma samples: 4*HM pixels of each of VM rows, read as dwords and realigned with v_alignbyte. A
// thread converts all its items into registers first; after a barrier its Y bytes and chroma samples leave as dword
// stores over the staged rows (the same LDS bytes).
template <int HM, int VM>
__device__ __forceinline__ void convert_rgb(const uint8_t* s_raw, const uint32_t* s_off, uint8_t* s_y, uint8_t* s_c,
                                            int tid) {
    constexpr int CW = kEncTileW / HM;  // chroma samples per row
    constexpr int QPR = CW / 4;         // items per chroma row
    constexpr int NI = 8 * QPR;         // items per strip
    constexpr int NIT = (NI + kEncThreads - 1) / kEncThreads;
    constexpr int DW = 3 * HM;          // dwords of RGB per row of one item
    constexpr int NSH = (HM - 1) + (VM - 1);
    constexpr int HALF = (1 << NSH) >> 1;
    uint32_t ypk[NIT][VM][HM], cbp[NIT], crp[NIT];
#pragma unroll
    for (int n = 0; n < NIT; n++) {
        const int it = tid + n * kEncThreads;
        if (it >= NI) break;
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
#pragma unroll
            for (int p = 0; p < 4 * HM; p++) {
                const int32_t r = (a[(3 * p) / 4] >> (8 * ((3 * p) % 4))) & 0xFF;
                const int32_t g = (a[(3 * p + 1) / 4] >> (8 * ((3 * p + 1) % 4))) & 0xFF;
                const int32_t b = (a[(3 * p + 2) / 4] >> (8 * ((3 * p + 2) % 4))) & 0xFF;
                const uint32_t y = uint32_t((19595 * r + 38470 * g + 7471 * b + 32768) >> 16) & 0xFF;
                const int32_t cb = (-11059 * r - 21709 * g + 32768 * b + (128 << 16) + 32767) >> 16;
                const int32_t cr = (32768 * r - 27439 * g - 5329 * b + (128 << 16) + 32767) >> 16;
                ypk[n][dy][p / 4] = (p % 4 == 0) ? y : (ypk[n][dy][p / 4] | (y << (8 * (p % 4))));
                sb[p / HM] += cb & 0xFF;
                sr[p / HM] += cr & 0xFF;
            }
        }
        cbp[n] = crp[n] = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            // The rounded mean of HM*VM samples (a power of two; the sums are non-negative).
            cbp[n] |= (uint32_t((sb[j] + HALF) >> NSH) & 0xFF) << (8 * j);
            crp[n] |= (uint32_t((sr[j] + HALF) >> NSH) & 0xFF) << (8 * j);
        }
    }
    __syncthreads();
#pragma unroll
    for (int n = 0; n < NIT; n++) {
        const int it = tid + n * kEncThreads;
        if (it >= NI) break;
        const int gy = it / QPR, q = it % QPR;
#pragma unroll
        for (int dy = 0; dy < VM; dy++)
#pragma unroll
            for (int k = 0; k < HM; k++)
                *reinterpret_cast<uint32_t*>(s_y + (gy * VM + dy) * kEncTileW + q * (4 * HM) + 4 * k) = ypk[n][dy][k];
        *reinterpret_cast<uint32_t*>(s_c + gy * CW + 4 * q) = cbp[n];
        *reinterpret_cast<uint32_t*>(s_c + 8 * CW + gy * CW + 4 * q) = crp[n];
    }
}

// kFull: candidates may ask for the trellis hand-off values and the reconstruction planes besides the coefficients;
// without it the launch writes coefficients only, and the kernel holds far fewer registers.
template <bool kFull>
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
    uint8_t* const s_raw = s_mem;  // staged rows; later the Y rows and chroma planes over them
    uint8_t* const s_y = s_mem;
    uint8_t* const s_c = s_mem + rows * kEncTileW;

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
        float* sp = kFull ? K.scaled[comp] : nullptr;
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
        uint8_t* rp = kFull ? K.rec[comp] : nullptr;
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
// Fused reconstruct of the usual layouts: decode.rs:1312-1343 (IDCT), 1769-1829 (upsampling) and 1427-1445 (colour)
// in one pass, with no sample planes in HBM. Component 0 is at full resolution; components 1 and 2 share the
// upsampling factors (FX, FY), each 1 or 2. A 256-thread workgroup owns a tile of 32 x 8 luma blocks (256 x 64
// pixels):
//   phase 1 (subsampled chroma only): the IDCT of the tile's chroma blocks into two LDS planes, with a ring of one
//           block in each direction whose factor is 2 (the triangle filter reads one chroma sample beyond the tile);
//           the sample just outside the component's real extent (column cw, row ch, column/row -1 at the image's
//           edges) is then overwritten with the edge sample, which is what the reference's clamped index reads, so
//           phase 2 reads its window without any clamping;
//   phase 2: one luma block per thread: its IDCT in registers, then for each of its rows the chroma of its 8 pixels
//           from the LDS window, the fixed-point colour conversion, and 24 (or 8) bytes stored straight into the
//           output image. Without subsampling a thread runs the IDCT of all three components of its block instead.
constexpr int kRcBX = 32, kRcBY = 8;  // luma blocks per tile
constexpr int kRcThreads = kRcBX * kRcBY;

// IDCT of one block of zigzag coefficients (dequantisation with mult, natural order) into its 64 level-shifted,
// clamped samples, packed four per dword: row y in out[2y] (x 0-3) and out[2y+1] (x 4-7). The final 1/8 is folded
// into the rounding as in k_encode (round-half-even of x/8, exact for |x| < 2^25).
__device__ __forceinline__ void load_block(const int16_t* co, uint32_t pk[32]) {
    const uint4* src = reinterpret_cast<const uint4*>(co);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint4 v = src[k];
        pk[4 * k] = v.x;
        pk[4 * k + 1] = v.y;
        pk[4 * k + 2] = v.z;
        pk[4 * k + 3] = v.w;
    }
}
__device__ __forceinline__ void idct_packed_u8(const uint32_t pk[32], const float* mult, uint32_t out[16]) {
    float blk[64];
#pragma unroll
    for (int k = 0; k < 64; k++) {
        const int16_t q = int16_t((pk[k / 2] >> (16 * (k & 1))) & 0xFFFF);
        const int nn = zigzag(k);
        blk[nn] = __fmul_rn(float(q), mult[nn]);
    }
    idct_block<false>(blk);
#pragma unroll
    for (int y = 0; y < 8; y++) {
        uint32_t lo = 0, hi = 0;
#pragma unroll
        for (int x = 0; x < 8; x++) {
            int32_t v = __float_as_int(__fadd_rn(blk[y * 8 + x], 100663296.0f)) - (0x4CC00000 - 128);
            v = v < 0 ? 0 : (v > 255 ? 255 : v);
            if (x < 4) lo |= uint32_t(v) << (8 * x);
            else hi |= uint32_t(v) << (8 * (x - 4));
        }
        out[2 * y] = lo;
        out[2 * y + 1] = hi;
    }
    // Make the packed samples opaque so the compiler forms all of them here, rather than sinking parts of the IDCT
    // into the per-row store branches of the caller, where its 64 values would stay live across them.
#pragma unroll
    for (int k = 0; k < 16; k++) asm volatile("" : "+v"(out[k]));
}
__device__ __forceinline__ void idct_block_u8(const int16_t* co, const float* mult, uint32_t out[16]) {
    uint32_t pk[32];
    load_block(co, pk);
    idct_packed_u8(pk, mult, out);
}

__device__ __forceinline__ int32_t byte_of(const uint32_t* d, int i) { return int32_t((d[i >> 2] >> (8 * (i & 3))) & 0xFF); }

// Store the first n bytes of d (little-endian, n <= 4 * ND) at p, which may have any alignment: aligned dword stores
// for whole rows that start on a dword, else the dwords realigned with v_alignbyte and byte stores at either end.
template <int ND>
__device__ __forceinline__ void store_bytes(uint8_t* p, const uint32_t (&d)[ND], uint32_t n) {
    const uint32_t s = uint32_t(reinterpret_cast<uintptr_t>(p) & 3);
    if (n == 4 * ND) {
        if (s == 0) {
            uint32_t* q = reinterpret_cast<uint32_t*>(p);
#pragma unroll
            for (int k = 0; k < ND; k++) q[k] = d[k];
            return;
        }
        // Bytes 4 - s .. 4 * ND - 1 - s fill the ND - 1 aligned dwords from p + 4 - s; the 4 - s bytes before them
        // come from d[0], the s after them from d[ND - 1].
        uint32_t* q = reinterpret_cast<uint32_t*>(p + 4 - s);
#pragma unroll
        for (int k = 0; k + 1 < ND; k++) q[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], 4 - s);
        for (uint32_t i = 0; i < 4 - s; i++) p[i] = uint8_t(d[0] >> (8 * i));
        for (uint32_t i = 4 * ND - s; i < 4 * ND; i++) p[i] = uint8_t(d[ND - 1] >> (8 * (i & 3)));
        return;
    }
#pragma unroll
    for (int i = 0; i < 4 * ND; i++)
        if (uint32_t(i) < n) p[i] = uint8_t(byte_of(d, i));
}

struct FusedRecon {
    const int16_t* coefs[3];
    uint32_t bw[3];        // block grid widths
    float mult[3][64];     // dequantisation per component, natural order
    uint32_t w, h;         // image size
    uint32_t cw, ch;       // real size of components 1 and 2
    uint32_t nbx_c, nby_c; // their blocks that hold real samples: ceil(cw/8), ceil(ch/8)
    uint8_t* out;          // Gray8 (one component) or RGB8, w x h, rows of w*channels bytes
};

// One output row of 8 pixels (24 bytes) from the luma row and the 8 chroma samples of each plane.
template <bool kRgbDirect>
__device__ __forceinline__ void color_row(const int32_t* yv, const int32_t* cb, const int32_t* cr, uint32_t (&d)[6]) {
#pragma unroll
    for (int k = 0; k < 6; k++) d[k] = 0;
#pragma unroll
    for (int x = 0; x < 8; x++) {
        int32_t r, g, b;
        if constexpr (kRgbDirect) {
            r = yv[x];
            g = cb[x];
            b = cr[x];
        } else {
            ycc_to_rgb(yv[x], cb[x], cr[x], r, g, b);
        }
        const int i = 3 * x;
        d[i >> 2] |= uint32_t(r) << (8 * (i & 3));
        d[(i + 1) >> 2] |= uint32_t(g) << (8 * ((i + 1) & 3));
        d[(i + 2) >> 2] |= uint32_t(b) << (8 * ((i + 2) & 3));
    }
}

// The 8 chroma samples of output row yy of a block from the thread's window of one plane (rows and columns from one
// below the block's first chroma sample where the factor is 2). win[r] holds window row r as bytes: 6 bytes from byte
// 3 of three dwords (FX = 2, columns i0-1 .. i0+4) or 8 bytes of two dwords (FX = 1).
template <int FX, int FY, int WD>
__device__ __forceinline__ void chroma_row(const uint32_t (*win)[WD], int yy, int32_t out[8]) {
    constexpr int c0 = FX == 2 ? 3 : 0;  // byte of the window's first column
    if constexpr (FY == 2) {
        const int jr = (yy >> 1) + 1, fr = (yy & 1) ? jr + 1 : jr - 1;
        if constexpr (FX == 2) {
            int32_t vert[6];
#pragma unroll
            for (int k = 0; k < 6; k++) vert[k] = 3 * byte_of(win[jr], c0 + k) + byte_of(win[fr], c0 + k);
#pragma unroll
            for (int x = 0; x < 8; x++) {
                const int n = (x >> 1) + 1, f = (x & 1) ? n + 1 : n - 1;
                out[x] = (3 * vert[n] + vert[f] + ((x & 1) ? 7 : 8)) >> 4;
            }
        } else {
#pragma unroll
            for (int x = 0; x < 8; x++)
                out[x] = (3 * byte_of(win[jr], x) + byte_of(win[fr], x) + ((yy & 1) ? 2 : 1)) >> 2;
        }
    } else {
        if constexpr (FX == 2) {
#pragma unroll
            for (int x = 0; x < 8; x++) {
                const int n = (x >> 1) + 1, f = (x & 1) ? n + 1 : n - 1;
                out[x] = (3 * byte_of(win[yy], c0 + n) + byte_of(win[yy], c0 + f) + ((x & 1) ? 2 : 1)) >> 2;
            }
        } else {
#pragma unroll
            for (int x = 0; x < 8; x++) out[x] = byte_of(win[yy], x);
        }
    }
}

// NC = 1: Gray8 from one component. NC = 3, FX = FY = 1: each thread reconstructs all three components of its block.
// NC = 3 with a factor of 2: the chroma goes through LDS (phases above). RD: RGB-coded components (no colour
// conversion).
template <int NC, int FX, int FY, bool RD>
__global__ __launch_bounds__(kRcThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_reconstruct(FusedRecon a) {
    constexpr bool kLds = NC == 3 && (FX == 2 || FY == 2);
    constexpr int HX = FX - 1, HY = FY - 1;                     // ring blocks per side
    constexpr int NBX = kRcBX / FX + 2 * HX, NBY = kRcBY / FY + 2 * HY;  // chroma blocks of the LDS window
    constexpr int PW = 8 * NBX, PH = 8 * NBY;                   // LDS plane width and height in samples
    constexpr int NB = NBX * NBY;                               // chroma blocks per plane
    constexpr int NBP = (NB + 63) & ~63;                        // per plane, padded to whole waves
    __shared__ __align__(16) uint8_t s_c[kLds ? 2 * PW * PH : 4];

    const int tid = threadIdx.x;
    const uint32_t cbx0 = blockIdx.x * (kRcBX / FX), cby0 = blockIdx.y * (kRcBY / FY);  // first chroma block
    const int lbx = tid % kRcBX, lby = tid / kRcBX;
    const uint32_t x0 = (blockIdx.x * kRcBX + lbx) * 8, y0 = (blockIdx.y * kRcBY + lby) * 8;
    const bool has_px = x0 < a.w && y0 < a.h;
    const uint32_t gbx = x0 / 8, gby = y0 / 8;
    // The thread's luma coefficients are requested first, so that they arrive while the chroma phase runs.
    uint32_t ypk[32];
    if (kLds && has_px) load_block(a.coefs[0] + (uint64_t(gby) * a.bw[0] + gbx) * 64, ypk);
    if constexpr (kLds) {
        // Phase 1: one chroma block per thread and pass; a wave never mixes the two planes.
        for (int k = tid; k < 2 * NBP; k += kRcThreads) {
            const int comp = __builtin_amdgcn_readfirstlane(k / NBP);
            const int lb = k - comp * NBP;
            if (lb >= NB) continue;
            const int lx = lb % NBX, ly = lb / NBX;
            const int gbx = int(cbx0) - HX + lx, gby = int(cby0) - HY + ly;
            if (gbx < 0 || gby < 0 || gbx >= int(a.nbx_c) || gby >= int(a.nby_c)) continue;
            uint32_t px[16];
            idct_block_u8(a.coefs[1 + comp] + (uint64_t(gby) * a.bw[1 + comp] + uint32_t(gbx)) * 64, a.mult[1 + comp],
                          px);
            uint8_t* dst = s_c + comp * PW * PH + (8 * ly) * PW + 8 * lx;
#pragma unroll
            for (int r = 0; r < 8; r++) *reinterpret_cast<uint2*>(dst + r * PW) = make_uint2(px[2 * r], px[2 * r + 1]);
        }
        __syncthreads();
        // Clamped columns, then clamped rows (so the corners take both).
        if constexpr (HX) {
            const int ox = (int(cbx0) - HX) * 8;  // global chroma column of LDS column 0
            const bool left = cbx0 == 0;
            const int rc = int(a.cw) - ox;        // LDS column of global column cw
            const bool right = rc < PW;
            if (left || right) {
                for (int k = tid; k < 2 * PH; k += kRcThreads) {
                    uint8_t* row = s_c + (k / PH) * PW * PH + (k % PH) * PW;
                    if (left) row[8 * HX - 1] = row[8 * HX];
                    if (right) row[rc] = row[rc - 1];
                }
            }
            __syncthreads();
        }
        if constexpr (HY) {
            const int oy = (int(cby0) - HY) * 8;
            const bool top = cby0 == 0;
            const int rr = int(a.ch) - oy;
            const bool bottom = rr < PH;
            if (top || bottom) {
                for (int k = tid; k < 2 * PW; k += kRcThreads) {
                    uint8_t* col = s_c + (k / PW) * PW * PH + (k % PW);
                    if (top) col[(8 * HY - 1) * PW] = col[(8 * HY) * PW];
                    if (bottom) col[rr * PW] = col[(rr - 1) * PW];
                }
            }
            __syncthreads();
        }
    }

    // Phase 2: one luma block per thread.
    if (!has_px) return;
    uint32_t ypx[16];
    if constexpr (!kLds) load_block(a.coefs[0] + (uint64_t(gby) * a.bw[0] + gbx) * 64, ypk);
    idct_packed_u8(ypk, a.mult[0], ypx);
    // One block's IDCT at a time: its 64 values and its coefficients are most of the thread's registers.
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t ncols = min(8u, a.w - x0), nrows = min(8u, a.h - y0);
    if constexpr (NC == 1) {
#pragma unroll
        for (int y = 0; y < 8; y++) {
            if (uint32_t(y) >= nrows) break;
            const uint32_t d[2] = {ypx[2 * y], ypx[2 * y + 1]};
            store_bytes(a.out + uint64_t(y0 + y) * a.w + x0, d, ncols);
        }
    } else {
        uint32_t cpx[2][16];  // FX = FY = 1: the block's Cb and Cr samples
        constexpr int WR = FY == 2 ? 6 : 8, WD = FX == 2 ? 3 : 2;
        uint32_t win[2][WR][WD];
        if constexpr (!kLds) {
#pragma unroll
            for (int c = 0; c < 2; c++) {
                idct_block_u8(a.coefs[1 + c] + (uint64_t(gby) * a.bw[1 + c] + gbx) * 64, a.mult[1 + c], cpx[c]);
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
            // Window origin in the LDS planes: one row/column above/left of the block's first chroma sample where the
            // factor is 2 (the ring block puts that at 8 - 1), the dword holding it for FX = 2.
            const int wx = FX == 2 ? 4 * lbx + 4 : 8 * lbx;
            const int wy = FY == 2 ? 4 * lby + 7 : 8 * lby;
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int r = 0; r < WR; r++) {
                    const uint32_t* src = reinterpret_cast<const uint32_t*>(s_c + c * PW * PH + (wy + r) * PW + wx);
#pragma unroll
                    for (int k = 0; k < WD; k++) win[c][r][k] = src[k];
                }
        }
        uint8_t* orow = a.out + (uint64_t(y0) * a.w + x0) * 3;
#pragma unroll
        for (int y = 0; y < 8; y++) {
            if (uint32_t(y) >= nrows) break;
            int32_t yv[8], cb[8], cr[8];
#pragma unroll
            for (int x = 0; x < 8; x++) yv[x] = byte_of(&ypx[2 * y], x);
            if constexpr (!kLds) {
#pragma unroll
                for (int x = 0; x < 8; x++) {
                    cb[x] = byte_of(&cpx[0][2 * y], x);
                    cr[x] = byte_of(&cpx[1][2 * y], x);
                }
            } else {
                chroma_row<FX, FY, WD>(win[0], y, cb);
                chroma_row<FX, FY, WD>(win[1], y, cr);
            }
            uint32_t d[6];
            color_row<RD>(yv, cb, cr, d);
            store_bytes(orow + uint64_t(y) * a.w * 3, d, 3 * ncols);
        }
    }
}

// ----------------------------------------------------------------------------------------------------------------
// Launchers.

int launch_encode(rcg_ctx* ctx, const EncJob* d_jobs, const uint32_t* d_begins, int njobs, uint32_t total_tiles,
                  bool coef_only) {
    if (total_tiles == 0) return RCG_OK;
    ProfScope ps(ctx, "k_encode");
    if (coef_only)
        hipLaunchKernelGGL(k_encode<false>, dim3(total_tiles), dim3(kEncThreads), 0, ctx->stream, d_jobs, d_begins,
                           njobs);
    else
        hipLaunchKernelGGL(k_encode<true>, dim3(total_tiles), dim3(kEncThreads), 0, ctx->stream, d_jobs, d_begins,
                           njobs);
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

template <int FX, int FY>
static void launch_rc(bool rgb_direct, dim3 grid, hipStream_t stream, const FusedRecon& a) {
    if (rgb_direct)
        hipLaunchKernelGGL((k_reconstruct<3, FX, FY, true>), grid, dim3(kRcThreads), 0, stream, a);
    else
        hipLaunchKernelGGL((k_reconstruct<3, FX, FY, false>), grid, dim3(kRcThreads), 0, stream, a);
}

bool fused_reconstruct_layout(const rcg_coef_image* ci, uint32_t hmax, uint32_t vmax, int* fx, int* fy) {
    if (ci->ncomp != 1 && ci->ncomp != 3) return false;
    if (ci->comps[0].h != hmax || ci->comps[0].v != vmax) return false;
    *fx = *fy = 1;
    if (ci->ncomp == 1) return true;
    const rcg_coef_component &c1 = ci->comps[1], &c2 = ci->comps[2];
    if (c1.h != c2.h || c1.v != c2.v) return false;
    *fx = int(hmax / c1.h);
    *fy = int(vmax / c1.v);
    return (*fx == 1 || *fx == 2) && (*fy == 1 || *fy == 2);
}

int launch_reconstruct_fused(rcg_ctx* ctx, const rcg_coef_image* ci, const int16_t* const d_coefs[4], uint8_t* d_out,
                             uint32_t hmax, uint32_t vmax, int rgb_direct) {
    int fx, fy;
    if (!fused_reconstruct_layout(ci, hmax, vmax, &fx, &fy)) {
        set_error("internal: layout without a fused reconstruct");
        return RCG_DEVICE;
    }
    FusedRecon a;
    std::memset(&a, 0, sizeof(a));
    for (uint32_t c = 0; c < ci->ncomp; c++) {
        a.coefs[c] = d_coefs[c];
        a.bw[c] = ci->comps[c].blocks_w;
        mult_table(ci->qtables[ci->comps[c].tq], a.mult[c]);
    }
    a.w = ci->width;
    a.h = ci->height;
    if (ci->ncomp == 3) {
        a.cw = (ci->width * ci->comps[1].h + hmax - 1) / hmax;
        a.ch = (ci->height * ci->comps[1].v + vmax - 1) / vmax;
        a.nbx_c = (a.cw + 7) / 8;
        a.nby_c = (a.ch + 7) / 8;
    }
    a.out = d_out;
    const dim3 grid((ci->width + 8 * kRcBX - 1) / (8 * kRcBX), (ci->height + 8 * kRcBY - 1) / (8 * kRcBY));
    ProfScope ps(ctx, "k_reconstruct");
    const bool rd = rgb_direct != 0;
    if (ci->ncomp == 1)
        hipLaunchKernelGGL((k_reconstruct<1, 1, 1, false>), grid, dim3(kRcThreads), 0, ctx->stream, a);
    else if (fx == 1 && fy == 1)
        launch_rc<1, 1>(rd, grid, ctx->stream, a);
    else if (fx == 2 && fy == 1)
        launch_rc<2, 1>(rd, grid, ctx->stream, a);
    else if (fx == 1 && fy == 2)
        launch_rc<1, 2>(rd, grid, ctx->stream, a);
    else
        launch_rc<2, 2>(rd, grid, ctx->stream, a);
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
