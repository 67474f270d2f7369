/*
 * oracle.c -- CPU restatement of rust-compress's JPEG target-quality inner loop.
 *
 * TEST INFRASTRUCTURE. This file is the checker for the HIP product path: only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py load it (through oracle/oracle.py). The product library
 * never links, loads or calls it.
 *
 * Each function restates one reference function and cites it (paths relative to the rust-compress tree). The
 * arithmetic follows the reference operation by operation:
 *   - the DCT and quantisation are f32, in the reference's evaluation order, with no FMA contraction
 *     (build with -ffp-contract=off; x86-64 SSE gives IEEE single precision with no excess precision);
 *   - colour conversion, downsampling and upsampling are 32-bit integer arithmetic;
 *   - SSIMULACRA2 keeps linear planes in f32 between scales and does everything else in f64, with the same
 *     96-row stripes and the same summation order, and calls glibc's cbrt/pow/exp as Rust's std does on Linux.
 * Threads (OpenMP) only split work the way rayon does; every reduction is done serially in the reference's order,
 * so results do not depend on the thread count.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   orc_build_coefs   - hierarchical fixture frame-0 entropy data (colour 4:2:0, gray, right/bottom padding),
 *                       DQT bytes of the q85 table, DCT-vs-naive and flat-block closed forms.
 *   orc_reconstruct   - the reference decoder's recorded decodes of two hierarchical fixtures, sample for sample
 *                       (f32 IDCT outputs over three frames each, the (2,2) upsampling and YCbCr -> RGB through
 *                       orc_idct_plane / orc_upsample / orc_ycc_to_rgb, tests/test_hier_golden.py, with planted
 *                       mistakes that must break the match); DC-only closed form, naive f64 IDCT agreement,
 *                       round-trip PSNR floors. Changes of about one ulp in the IDCT (a re-association) are too
 *                       small for those files: the exact f32 operation order is pinned by restatement.
 *   orc_cmyk_to_rgb   - Pillow's conversions of the reference's CMYK/YCCK fixtures within the reference's own
 *                       tolerance (cli/tests/cmyk.rs:14-80); the formula itself is pinned by restatement.
 *   orc_ssimulacra2   - the 54 py-ssimulacra2 golden pairs (fixtures/ssim2/expected.csv) at 1e-4, identity = 100.
 *   orc_find_quality  - bisection semantics against brute-force search over monotone score tables.
 *   orc_build_lengths - Kraft completeness, the length limit, single-symbol and textbook-Huffman optimality.
 *   orc_rdo_quantize_ac / orc_trellis - the two unit tests of jpeg/src/rdo.rs and brute-force optimality of the
 *                       trellis choice over every candidate assignment of small blocks.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#include "orc_weights.h"

/* ------------------------------------------------------------------------------------------------------------ */
/* jpeg/src/lib.rs:15-19 -- zigzag scan order (natural index of the k-th coefficient).                           */
static const int ZIGZAG[64] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
    30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

/* jpeg/src/quality.rs:26-35 -- the base table used for both luma and chroma. */
static const uint8_t PERCEPTUAL_Q[64] = {
    16, 16, 16,  18,  25,  37,  56,  85,  16, 17, 20,  27,  34,  40,  53,  75,
    16, 20, 24,  31,  43,  62,  91,  135, 18, 27, 31,  40,  53,  74,  106, 156,
    25, 34, 43,  53,  69,  94,  131, 189, 37, 40, 62,  74,  94,  124, 169, 238,
    56, 53, 91,  106, 131, 169, 226, 255, 85, 75, 135, 156, 189, 238, 255, 255};

/* jpeg/src/dct.rs:8 -- AAN scale factors, each the f32 nearest its literal. */
static const float AAN[8] = {1.0f, 1.3870398f, 1.306563f, 1.1758756f, 1.0f, 0.78569495f, 0.5411961f, 0.27589938f};

/* jpeg/src/quality.rs:6-14 -- IJG scaling of a base table. */
void orc_scale_base(const uint8_t base[64], int quality, uint8_t out[64]) {
    unsigned q = (unsigned)(quality < 1 ? 1 : (quality > 100 ? 100 : quality));
    unsigned scale = q < 50 ? 5000 / q : 200 - 2 * q;
    for (int i = 0; i < 64; i++) {
        unsigned t = ((unsigned)base[i] * scale + 50) / 100;
        out[i] = (uint8_t)(t < 1 ? 1 : (t > 255 ? 255 : t));
    }
}

/* jpeg/src/quality.rs:37-43 -- luma and chroma tables both scale PERCEPTUAL_Q. */
void orc_scale_table(int quality, uint8_t out[64]) { orc_scale_base(PERCEPTUAL_Q, quality, out); }

/* jpeg/src/dct.rs:127-129 -- round half to even through the 1.5*2^23 magic number. */
static inline int32_t round_f32(float x) {
    float y = x + 12582912.0f;
    int32_t bits;
    memcpy(&bits, &y, 4);
    return bits - 0x4B400000;
}

typedef float Block[8][8];

/* jpeg/src/dct.rs:10-18 */
static void transpose(Block b) {
    for (int i = 0; i < 8; i++)
        for (int j = i + 1; j < 8; j++) {
            float t = b[i][j];
            b[i][j] = b[j][i];
            b[j][i] = t;
        }
}

/* jpeg/src/dct.rs:20-57 -- one AAN pass down every column j. */
static void fdct_pass(Block d) {
    for (int j = 0; j < 8; j++) {
        float d0 = d[0][j], d1 = d[1][j], d2 = d[2][j], d3 = d[3][j];
        float d4 = d[4][j], d5 = d[5][j], d6 = d[6][j], d7 = d[7][j];
        float tmp0 = d0 + d7, tmp7 = d0 - d7;
        float tmp1 = d1 + d6, tmp6 = d1 - d6;
        float tmp2 = d2 + d5, tmp5 = d2 - d5;
        float tmp3 = d3 + d4, tmp4 = d3 - d4;
        float tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3;
        float tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
        d[0][j] = tmp10 + tmp11;
        d[4][j] = tmp10 - tmp11;
        float z1 = (tmp12 + tmp13) * 0.70710677f;
        d[2][j] = tmp13 + z1;
        d[6][j] = tmp13 - z1;
        float o10 = tmp4 + tmp5, o11 = tmp5 + tmp6, o12 = tmp6 + tmp7;
        float z5 = (o10 - o12) * 0.38268343f;
        float z2 = 0.5411961f * o10 + z5;
        float z4 = 1.306563f * o12 + z5;
        float z3 = o11 * 0.70710677f;
        float z11 = tmp7 + z3, z13 = tmp7 - z3;
        d[5][j] = z13 + z2;
        d[3][j] = z13 - z2;
        d[1][j] = z11 + z4;
        d[7][j] = z11 - z4;
    }
}

/* jpeg/src/dct.rs:61-65 */
void orc_fdct(Block b) {
    fdct_pass(b);
    transpose(b);
    fdct_pass(b);
}

/* jpeg/src/dct.rs:67-109 */
static void idct_pass(Block d) {
    const float SQRT2 = 1.41421356237309504880f;
    for (int j = 0; j < 8; j++) {
        float tmp0 = d[0][j], tmp1 = d[2][j], tmp2 = d[4][j], tmp3 = d[6][j];
        float tmp10 = tmp0 + tmp2, tmp11 = tmp0 - tmp2;
        float tmp13 = tmp1 + tmp3;
        float tmp12 = (tmp1 - tmp3) * SQRT2 - tmp13;
        float e0 = tmp10 + tmp13, e3 = tmp10 - tmp13;
        float e1 = tmp11 + tmp12, e2 = tmp11 - tmp12;
        float tmp4 = d[1][j], tmp5 = d[3][j], tmp6 = d[5][j], tmp7 = d[7][j];
        float z13 = tmp6 + tmp5, z10 = tmp6 - tmp5;
        float z11 = tmp4 + tmp7, z12 = tmp4 - tmp7;
        float o7 = z11 + z13;
        float o11 = (z11 - z13) * SQRT2;
        float z5 = (z10 + z12) * 1.8477591f;
        float o10 = 1.0823922f * z12 - z5;
        float o12 = -2.613126f * z10 + z5;
        float o6 = o12 - o7;
        float o5 = o11 - o6;
        float o4 = o10 + o5;
        d[0][j] = e0 + o7;
        d[7][j] = e0 - o7;
        d[1][j] = e1 + o6;
        d[6][j] = e1 - o6;
        d[2][j] = e2 + o5;
        d[5][j] = e2 - o5;
        d[4][j] = e3 + o4;
        d[3][j] = e3 - o4;
    }
}

/* jpeg/src/dct.rs:113-123 */
void orc_idct(Block b) {
    idct_pass(b);
    transpose(b);
    idct_pass(b);
    transpose(b);
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) b[y][x] *= 0.125f;
}

/* Exposed for the KAT tests: forward / inverse transform of one block given row-major. */
void orc_fdct_block(float* b) { orc_fdct((float(*)[8])b); }
void orc_idct_block(float* b) { orc_idct((float(*)[8])b); }
int32_t orc_round_f32(float x) { return round_f32(x); }

/* jpeg/src/encode.rs:175-183 -- reciprocal divisors in the transposed ([u][v]) layout. */
static void fdct_divisors(const uint8_t q[64], float d[64]) {
    for (int v = 0; v < 8; v++)
        for (int u = 0; u < 8; u++) d[u * 8 + v] = 1.0f / ((float)q[v * 8 + u] * AAN[u] * AAN[v] * 8.0f);
}

/* jpeg/src/encode.rs:45-54 */
static int ZZ_T(int k) {
    int n = ZIGZAG[k];
    return (n % 8) * 8 + n / 8;
}

/* jpeg/src/encode.rs:185-219. out receives the plain coefficients; scaled (may be NULL) the f32 values
 * a[k] = flat[f] * div[f] that the trellis branch (encode.rs:206-211) hands to Rdo::quantize_ac. */
static void quantize_plane(const uint8_t* plane, size_t stride, int bw, int bh, const float div[64], int16_t* out,
                           float* scaled) {
    int zzt[64];
    for (int k = 0; k < 64; k++) zzt[k] = ZZ_T(k);
#pragma omp parallel for schedule(dynamic, 4)
    for (int by = 0; by < bh; by++) {
        for (int bx = 0; bx < bw; bx++) {
            Block blk;
            for (int y = 0; y < 8; y++) {
                const uint8_t* src = plane + ((size_t)by * 8 + y) * stride + (size_t)bx * 8;
                for (int x = 0; x < 8; x++) blk[y][x] = (float)src[x] - 128.0f;
            }
            orc_fdct(blk);
            const float* flat = &blk[0][0];
            int16_t* dst = out + ((size_t)by * bw + bx) * 64;
            for (int k = 0; k < 64; k++) {
                int f = zzt[k];
                int32_t r = round_f32(flat[f] * div[f]);
                dst[k] = (int16_t)(r < -2047 ? -2047 : (r > 2047 ? 2047 : r));
            }
            if (scaled) {
                float* sd = scaled + ((size_t)by * bw + bx) * 64;
                for (int k = 0; k < 64; k++) sd[k] = flat[zzt[k]] * div[zzt[k]];
            }
        }
    }
}

/* Subsampling codes, in the order of jpeg/src/encode.rs:12-20: Auto, S444, S422, S420. */
enum { ORC_AUTO = 0, ORC_S444 = 1, ORC_S422 = 2, ORC_S420 = 3 };

/* Sampling factors chosen by build_coefs (jpeg/src/encode.rs:428-439). */
void orc_sampling(int channels, int quality, int subsampling, int* hmax, int* vmax) {
    if (channels == 1) {
        *hmax = 1;
        *vmax = 1;
        return;
    }
    switch (subsampling) {
        case ORC_S420: *hmax = 2; *vmax = 2; break;
        case ORC_S422: *hmax = 2; *vmax = 1; break;
        case ORC_S444: *hmax = 1; *vmax = 1; break;
        default:
            if (quality < 90) { *hmax = 2; *vmax = 2; }
            else { *hmax = 1; *vmax = 1; }
    }
}

/*
 * jpeg/src/encode.rs:426-539 with trellis off: pad with edge replication, RGB->YCbCr (16-bit fixed point),
 * box-downsample chroma, FDCT and quantise every block.
 * pixels: interleaved Gray8 (channels=1) or RGB8 (channels=3), row stride w*channels.
 * out_y/out_cb/out_cr: 64 zigzag coefficients per block, raster block order over the padded grid;
 *   Y has (pw/8)*(ph/8) blocks, chroma (pw/hmax/8)*(ph/vmax/8).
 * qt_out (128 bytes, may be NULL): luma then chroma table in natural order.
 * Returns 0, or 1 for invalid arguments.
 */
int orc_build_coefs_tables(const uint8_t* pixels, int w, int h, int channels, int quality, int subsampling,
                           const uint8_t* qy, const uint8_t* qc, int16_t* out_y, int16_t* out_cb, int16_t* out_cr);

int orc_build_coefs(const uint8_t* pixels, int w, int h, int channels, int quality, int subsampling, int16_t* out_y,
                    int16_t* out_cb, int16_t* out_cr, uint8_t* qt_out) {
    uint8_t qy[64], qc[64];
    orc_scale_table(quality, qy);
    orc_scale_table(quality, qc);
    if (qt_out) {
        memcpy(qt_out, qy, 64);
        memcpy(qt_out + 64, qc, 64);
    }
    return orc_build_coefs_tables(pixels, w, h, channels, quality, subsampling, qy, qc, out_y, out_cb, out_cr);
}

/* The same with explicit natural-order tables (quality only picks Auto subsampling); lets the tests feed the
 * tables that a fixture file carries. */
static int build_impl(const uint8_t* pixels, int w, int h, int channels, int quality, int subsampling,
                      const uint8_t* qy, const uint8_t* qc, int16_t* out_y, int16_t* out_cb, int16_t* out_cr,
                      float* sc_y, float* sc_cb, float* sc_cr);

int orc_build_coefs_tables(const uint8_t* pixels, int w, int h, int channels, int quality, int subsampling,
                           const uint8_t* qy, const uint8_t* qc, int16_t* out_y, int16_t* out_cb, int16_t* out_cr) {
    return build_impl(pixels, w, h, channels, quality, subsampling, qy, qc, out_y, out_cb, out_cr, NULL, NULL, NULL);
}

/* As orc_build_coefs, also writing the trellis hand-off values (64 f32 per block, same layout). */
int orc_build_scaled(const uint8_t* pixels, int w, int h, int channels, int quality, int subsampling, int16_t* out_y,
                     int16_t* out_cb, int16_t* out_cr, float* sc_y, float* sc_cb, float* sc_cr) {
    uint8_t qy[64], qc[64];
    orc_scale_table(quality, qy);
    orc_scale_table(quality, qc);
    return build_impl(pixels, w, h, channels, quality, subsampling, qy, qc, out_y, out_cb, out_cr, sc_y, sc_cb, sc_cr);
}

static int build_impl(const uint8_t* pixels, int w, int h, int channels, int quality, int subsampling,
                      const uint8_t* qy, const uint8_t* qc, int16_t* out_y, int16_t* out_cb, int16_t* out_cr,
                      float* sc_y, float* sc_cb, float* sc_cr) {
    if (w < 1 || h < 1 || (channels != 1 && channels != 3)) return 1;
    int hmax, vmax;
    orc_sampling(channels, quality, subsampling, &hmax, &vmax);
    int mcu_w = 8 * hmax, mcu_h = 8 * vmax;
    int pw = (w + mcu_w - 1) / mcu_w * mcu_w, ph = (h + mcu_h - 1) / mcu_h * mcu_h;
    float dy[64], dc[64];
    fdct_divisors(qy, dy);
    fdct_divisors(qc, dc);
    if (channels == 1) {
        uint8_t* yp = (uint8_t*)malloc((size_t)pw * ph);
        if (!yp) return 4;
        for (int r = 0; r < ph; r++) {
            int sr = r < h - 1 ? r : h - 1;
            for (int x = 0; x < pw; x++) yp[(size_t)r * pw + x] = pixels[(size_t)sr * w + (x < w - 1 ? x : w - 1)];
        }
        quantize_plane(yp, pw, pw / 8, ph / 8, dy, out_y, sc_y);
        free(yp);
        return 0;
    }
    size_t n = (size_t)pw * ph;
    uint8_t* yp = (uint8_t*)malloc(n);
    uint8_t* cbp = (uint8_t*)malloc(n);
    uint8_t* crp = (uint8_t*)malloc(n);
    if (!yp || !cbp || !crp) { free(yp); free(cbp); free(crp); return 4; }
#pragma omp parallel for schedule(static)
    for (int r = 0; r < ph; r++) {
        int sr = r < h - 1 ? r : h - 1;
        for (int x = 0; x < pw; x++) {
            size_t i = ((size_t)sr * w + (x < w - 1 ? x : w - 1)) * 3;
            int32_t rr = pixels[i], gg = pixels[i + 1], bb = pixels[i + 2];
            yp[(size_t)r * pw + x] = (uint8_t)((19595 * rr + 38470 * gg + 7471 * bb + 32768) >> 16);
            cbp[(size_t)r * pw + x] = (uint8_t)((-11059 * rr - 21709 * gg + 32768 * bb + (128 << 16) + 32767) >> 16);
            crp[(size_t)r * pw + x] = (uint8_t)((32768 * rr - 27439 * gg - 5329 * bb + (128 << 16) + 32767) >> 16);
        }
    }
    int cw = pw, ch = ph;
    uint8_t *cb = cbp, *cr = crp;
    if (hmax * vmax > 1) {
        cw = pw / hmax;
        ch = ph / vmax;
        cb = (uint8_t*)malloc((size_t)cw * ch);
        cr = (uint8_t*)malloc((size_t)cw * ch);
        unsigned nn = (unsigned)(hmax * vmax);
        for (int pl = 0; pl < 2; pl++) {
            const uint8_t* p = pl ? crp : cbp;
            uint8_t* o = pl ? cr : cb;
#pragma omp parallel for schedule(static)
            for (int r = 0; r < ch; r++)
                for (int x = 0; x < cw; x++) {
                    unsigned s = 0;
                    for (int ddy = 0; ddy < vmax; ddy++)
                        for (int ddx = 0; ddx < hmax; ddx++) s += p[((size_t)vmax * r + ddy) * pw + hmax * x + ddx];
                    o[(size_t)r * cw + x] = (uint8_t)((s + nn / 2) / nn);
                }
        }
    }
    quantize_plane(yp, pw, pw / 8, ph / 8, dy, out_y, sc_y);
    quantize_plane(cb, cw, cw / 8, ch / 8, dc, out_cb, sc_cb);
    quantize_plane(cr, cw, cw / 8, ch / 8, dc, out_cr, sc_cr);
    if (cb != cbp) { free(cb); free(cr); }
    free(yp);
    free(cbp);
    free(crp);
    return 0;
}

/* jpeg/src/decode.rs:1317-1322 -- dequantisation multipliers q[v*8+u]*AAN[u]*AAN[v] in f32, in that product order. */
static void dequant_mult(const uint8_t qt[64], float mult[8][8]) {
    for (int v = 0; v < 8; v++)
        for (int u = 0; u < 8; u++) mult[v][u] = (float)qt[v * 8 + u] * AAN[u] * AAN[v];
}

/* jpeg/src/decode.rs:1327-1337 -- one block: zigzag -> natural dequantisation, the f32 AAN IDCT and round_f32, before
 * the level shift and the clamp. */
static void idct_block(const int16_t* src, const float mult[8][8], int32_t out[64]) {
    Block blk;
    for (int k = 0; k < 64; k++) {
        int nn = ZIGZAG[k];
        blk[nn / 8][nn % 8] = (float)src[k] * mult[nn / 8][nn % 8];
    }
    orc_idct(blk);
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) out[y * 8 + x] = round_f32(blk[y][x]);
}

/* jpeg/src/decode.rs:1312-1343 -- idct_planes::<T>(center, lo, max) for one component: sample = round_f32(x) + center
 * clamped to [lo, hi], in a plane of (blocks_w*8) x (blocks_h*8) int32 samples. An 8-bit picture uses (128, 0, 255)
 * (decode.rs:1406); the frames of a hierarchical image are not clamped, (128 or 0, i32::MIN/4, i32::MAX/4)
 * (decode.rs:836-844). */
void orc_idct_plane(const int16_t* coefs, int blocks_w, int blocks_h, const uint8_t qt[64], int32_t center,
                    int32_t lo, int32_t hi, int32_t* plane) {
    float mult[8][8];
    dequant_mult(qt, mult);
    size_t stride = (size_t)blocks_w * 8;
#pragma omp parallel for schedule(dynamic, 4)
    for (int by = 0; by < blocks_h; by++)
        for (int bx = 0; bx < blocks_w; bx++) {
            int32_t r[64];
            idct_block(coefs + ((size_t)by * blocks_w + bx) * 64, mult, r);
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++) {
                    int32_t v = r[y * 8 + x] + center;
                    plane[((size_t)by * 8 + y) * stride + (size_t)bx * 8 + x] = v < lo ? lo : (v > hi ? hi : v);
                }
        }
}

/* The same with the 8-bit picture's arguments (128, 0, 255), stored as u8 (T = u8, decode.rs:1406). */
static void idct_plane(const int16_t* coefs, int blocks_w, int blocks_h, const uint8_t qt[64], uint8_t* plane) {
    float mult[8][8];
    dequant_mult(qt, mult);
    size_t stride = (size_t)blocks_w * 8;
#pragma omp parallel for schedule(dynamic, 4)
    for (int by = 0; by < blocks_h; by++)
        for (int bx = 0; bx < blocks_w; bx++) {
            int32_t r[64];
            idct_block(coefs + ((size_t)by * blocks_w + bx) * 64, mult, r);
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++) {
                    int32_t v = r[y * 8 + x] + 128;
                    plane[((size_t)by * 8 + y) * stride + (size_t)bx * 8 + x] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
                }
        }
}

/* jpeg/src/decode.rs:1769-1829 -- bring one plane to w x h ("fancy" triangle filter for 2x factors). */
void orc_upsample(const uint8_t* p, size_t stride, int cw, int ch, int fx, int fy, int w, int h, uint8_t* out) {
#define PX(xx, yy) ((int32_t)p[(size_t)(yy) * stride + (xx)])
    if (fx == 1 && fy == 1) {
        for (int y = 0; y < h; y++) memcpy(out + (size_t)y * w, p + (size_t)y * stride, (size_t)w);
    } else if (fx == 2 && fy == 1) {
#pragma omp parallel for schedule(static)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int i = x / 2;
                int32_t nearv = PX(i, y);
                int32_t far = (x % 2 == 0) ? PX(i > 0 ? i - 1 : 0, y) : PX(i + 1 < cw - 1 ? i + 1 : cw - 1, y);
                out[(size_t)y * w + x] = (uint8_t)((3 * nearv + far + ((x % 2 == 0) ? 1 : 2)) >> 2);
            }
    } else if (fx == 1 && fy == 2) {
#pragma omp parallel for schedule(static)
        for (int y = 0; y < h; y++) {
            int j = y / 2;
            int far_y = (y % 2 == 0) ? (j > 0 ? j - 1 : 0) : (j + 1 < ch - 1 ? j + 1 : ch - 1);
            for (int x = 0; x < w; x++)
                out[(size_t)y * w + x] = (uint8_t)((3 * PX(x, j) + PX(x, far_y) + ((y % 2 == 0) ? 1 : 2)) >> 2);
        }
    } else if (fx == 2 && fy == 2) {
#pragma omp parallel for schedule(static)
        for (int y = 0; y < h; y++) {
            int j = y / 2;
            int far_y = (y % 2 == 0) ? (j > 0 ? j - 1 : 0) : (j + 1 < ch - 1 ? j + 1 : ch - 1);
            for (int x = 0; x < w; x++) {
                int i = x / 2;
                int32_t nearv = 3 * PX(i, j) + PX(i, far_y);
                int fi = (x % 2 == 0) ? (i > 0 ? i - 1 : 0) : (i + 1 < cw - 1 ? i + 1 : cw - 1);
                int32_t far = 3 * PX(fi, j) + PX(fi, far_y);
                out[(size_t)y * w + x] = (uint8_t)((3 * nearv + far + ((x % 2 == 0) ? 8 : 7)) >> 4);
            }
        }
    } else {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) out[(size_t)y * w + x] = p[(size_t)(y / fy) * stride + x / fx];
    }
#undef PX
}

/* jpeg/src/decode.rs:1427-1445 -- three full-resolution planes to interleaved RGB8: copied when the components hold
 * R, G, B directly (Adobe transform 0, or no Adobe marker and ids "RGB"), else fixed-point YCbCr -> RGB. */
void orc_ycc_to_rgb(const uint8_t* p0, const uint8_t* p1, const uint8_t* p2, size_t n, int rgb_direct, uint8_t* out) {
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)n; ii++) {
        size_t i = (size_t)ii;
        int32_t a = p0[i], b = p1[i], c = p2[i];
        uint8_t* px = out + i * 3;
        if (rgb_direct) {
            px[0] = (uint8_t)a; px[1] = (uint8_t)b; px[2] = (uint8_t)c;
        } else {
            int32_t cb = b - 128, cr = c - 128;
            int32_t r = a + ((91881 * cr + 32768) >> 16);
            int32_t g = a - ((22554 * cb + 46802 * cr + 32768) >> 16);
            int32_t bl = a + ((116130 * cb + 32768) >> 16);
            px[0] = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
            px[1] = (uint8_t)(g < 0 ? 0 : (g > 255 ? 255 : g));
            px[2] = (uint8_t)(bl < 0 ? 0 : (bl > 255 ? 255 : bl));
        }
    }
}

/* jpeg/src/decode.rs:1454-1478 (cmyk_to_rgb) -- four full-resolution planes to RGB8. adobe_transform < 0: no Adobe
 * marker, the values are ink amounts; with a marker they are stored inverted (255 = no ink), and transform 2 means
 * the first three planes are YCbCr-coded (converted, clamped and inverted to CMY first). */
void orc_cmyk_to_rgb(const uint8_t* p0, const uint8_t* p1, const uint8_t* p2, const uint8_t* p3, size_t n,
                     int adobe_transform, uint8_t* out) {
    int inverted = adobe_transform >= 0, ycck = adobe_transform == 2;
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)n; ii++) {
        size_t i = (size_t)ii;
        int32_t c = p0[i], m = p1[i], y = p2[i], k = p3[i];
        if (ycck) {
            int32_t yy = c, cb = m - 128, cr = y - 128;
            int32_t r = yy + ((91881 * cr + 32768) >> 16);
            int32_t g = yy - ((22554 * cb + 46802 * cr + 32768) >> 16);
            int32_t b = yy + ((116130 * cb + 32768) >> 16);
            r = r < 0 ? 0 : (r > 255 ? 255 : r);
            g = g < 0 ? 0 : (g > 255 ? 255 : g);
            b = b < 0 ? 0 : (b > 255 ? 255 : b);
            c = 255 - r; m = 255 - g; y = 255 - b;
        }
        int32_t v3[3] = {c, m, y};
        for (int j = 0; j < 3; j++) {
            int32_t lv = inverted ? v3[j] : 255 - v3[j];
            int32_t lk = inverted ? k : 255 - k;
            out[i * 3 + j] = (uint8_t)((lv * lk + 127) / 255);
        }
    }
}

/*
 * jpeg/src/decode.rs:334-388 + 1390-1451 -- pixels a viewer would show for a coefficient image.
 * ncomp: 1, 3 or 4. comp_h/comp_v: sampling factors; blocks_w/blocks_h: block grid of each component;
 * coefs[c]: zigzag coefficients; qt[c]: 64-entry natural-order table of component c.
 * rgb_direct: the components hold R,G,B directly (ids "RGB" or Adobe transform 0).
 * adobe_transform: the Adobe APP14 transform, -1 without the marker (four components: cmyk_to_rgb's rule).
 * out: Gray8 (ncomp=1) or RGB8 (ncomp=3 or 4), w*h*(ncomp == 1 ? 1 : 3) bytes.
 * Returns 0, 1 (invalid) or 2 (unsupported: non-integral sampling ratio or another component count).
 */
int orc_reconstruct(int w, int h, int ncomp, const int* comp_h, const int* comp_v, const int* blocks_w,
                    const int* blocks_h, const int16_t* const* coefs, const uint8_t* const* qt, int rgb_direct,
                    int adobe_transform, uint8_t* out) {
    if (ncomp != 1 && ncomp != 3 && ncomp != 4) return 2;
    int hmax = 1, vmax = 1;
    for (int c = 0; c < ncomp; c++) {
        if (comp_h[c] > hmax) hmax = comp_h[c];
        if (comp_v[c] > vmax) vmax = comp_v[c];
    }
    uint8_t* full[4] = {0, 0, 0, 0};
    int status = 0;
    for (int c = 0; c < ncomp; c++) {
        size_t stride = (size_t)blocks_w[c] * 8;
        uint8_t* plane = (uint8_t*)malloc(stride * blocks_h[c] * 8);
        idct_plane(coefs[c], blocks_w[c], blocks_h[c], qt[c], plane);
        int cw = (w * comp_h[c] + hmax - 1) / hmax, ch = (h * comp_v[c] + vmax - 1) / vmax;
        if (hmax % comp_h[c] != 0 || vmax % comp_v[c] != 0) {
            free(plane);
            status = 2;
            break;
        }
        full[c] = (uint8_t*)malloc((size_t)w * h);
        orc_upsample(plane, stride, cw, ch, hmax / comp_h[c], vmax / comp_v[c], w, h, full[c]);
        free(plane);
    }
    if (status == 0) {
        if (ncomp == 1)
            memcpy(out, full[0], (size_t)w * h);
        else if (ncomp == 3)
            orc_ycc_to_rgb(full[0], full[1], full[2], (size_t)w * h, rgb_direct, out);
        else
            orc_cmyk_to_rgb(full[0], full[1], full[2], full[3], (size_t)w * h, adobe_transform, out);
    }
    for (int c = 0; c < 4; c++) free(full[c]);
    return status;
}

/* ============================================================================================================ */
/* SSIMULACRA2, core/src/ssimulacra2.rs                                                                          */

#define NUM_SCALES 6
#define MIN_SIDE 8
#define RADIUS 5
#define STRIPE 96
/* ------------------------------------------------------------------------------------------------------------ */
/* Trellis quantisation (jpeg/src/rdo.rs, jpeg/src/encode.rs:222-265,527-537) and the Huffman code lengths it takes
 * its rate model from (core/src/util/huffman.rs:10-88). */

/* core/src/util/huffman.rs:10-88: length-limited Huffman code lengths. */
typedef struct { uint32_t f; int i; } FreqSym;
static int cmp_freqsym(const void* pa, const void* pb) {
    const FreqSym* a = (const FreqSym*)pa;
    const FreqSym* b = (const FreqSym*)pb;
    if (a->f != b->f) return a->f < b->f ? -1 : 1;
    return a->i > b->i ? -1 : (a->i < b->i ? 1 : 0); /* higher symbol first among equal frequencies */
}

void orc_build_lengths(const uint32_t* freqs, int nsym, int max_len, uint8_t* lengths) {
    memset(lengths, 0, (size_t)nsym);
    FreqSym* syms = (FreqSym*)malloc(sizeof(FreqSym) * (size_t)(nsym > 0 ? nsym : 1));
    int n = 0;
    for (int i = 0; i < nsym; i++)
        if (freqs[i] > 0) { syms[n].f = freqs[i]; syms[n].i = i; n++; }
    if (n == 0) { free(syms); return; }
    if (n == 1) { lengths[syms[0].i] = 1; free(syms); return; }
    qsort(syms, (size_t)n, sizeof(FreqSym), cmp_freqsym);
    uint64_t* weight = (uint64_t*)calloc((size_t)(2 * n - 1), sizeof(uint64_t));
    int* parent = (int*)calloc((size_t)(2 * n - 1), sizeof(int));
    uint32_t* depth = (uint32_t*)calloc((size_t)(2 * n - 1), sizeof(uint32_t));
    for (int i = 0; i < n; i++) weight[i] = syms[i].f;
    int leaf = 0, node = n, next = n;
    while (next < 2 * n - 1) {
        int pair[2];
        for (int t = 0; t < 2; t++) {
            if (leaf < n && (node >= next || weight[leaf] <= weight[node])) pair[t] = leaf++;
            else pair[t] = node++;
        }
        weight[next] = weight[pair[0]] + weight[pair[1]];
        parent[pair[0]] = next;
        parent[pair[1]] = next;
        next++;
    }
    for (int i = 2 * n - 3; i >= 0; i--) depth[i] = depth[parent[i]] + 1;
    uint32_t count[65] = {0};
    for (int i = 0; i < n; i++) count[depth[i] < (uint32_t)max_len ? depth[i] : (uint32_t)max_len]++;
    uint64_t total = 0;
    for (int l = 1; l <= max_len; l++) total += (uint64_t)count[l] << (max_len - l);
    while (total > ((uint64_t)1 << max_len)) {
        count[max_len]--;
        for (int l = max_len - 1; l >= 1; l--)
            if (count[l] > 0) { count[l]--; count[l + 1] += 2; break; }
        total--;
    }
    int idx = 0;
    for (int l = max_len; l >= 1; l--)
        for (uint32_t k = 0; k < count[l]; k++) lengths[syms[idx++].i] = (uint8_t)l;
    free(weight); free(parent); free(depth); free(syms);
}

/* jpeg/src/encode.rs:240-265: code sizes of an optimal table (a dummy symbol reserves the all-ones code). */
static void optimal_sizes(const uint32_t freq[256], uint8_t size[256]) {
    uint32_t f[257];
    uint8_t lens[257];
    memcpy(f, freq, sizeof(uint32_t) * 256);
    f[256] = 1;
    orc_build_lengths(f, 257, 16, lens);
    int any = 0;
    for (int s = 0; s < 256; s++) { size[s] = lens[s]; any |= lens[s] > 0; }
    if (!any) size[0] = 1;
}

/* jpeg/src/encode.rs:337-370,376-396: AC symbol counts of the plain coefficients, per table class. */
static void ac_counts(int ncomp, const int16_t* const* coefs, const int* comp_h, const int* comp_v,
                      const int* blocks_w, int mcus_x, int mcus_y, uint32_t ac[2][256]) {
    memset(ac, 0, sizeof(uint32_t) * 512);
    for (int my = 0; my < mcus_y; my++)
        for (int mx = 0; mx < mcus_x; mx++)
            for (int c = 0; c < ncomp; c++)
                for (int by = 0; by < comp_v[c]; by++)
                    for (int bx = 0; bx < comp_h[c]; bx++) {
                        const size_t b = ((size_t)my * comp_v[c] + by) * blocks_w[c] + (size_t)mx * comp_h[c] + bx;
                        const int16_t* blk = coefs[c] + b * 64;
                        const int tbl = c < 1 ? c : 1;
                        uint32_t run = 0;
                        for (int k = 1; k < 64; k++) {
                            if (blk[k] == 0) { run++; continue; }
                            while (run >= 16) { ac[tbl][0xF0]++; run -= 16; }
                            int32_t v = blk[k];
                            uint32_t a = (uint32_t)(v < 0 ? -v : v), cat = 0;
                            while (a) { cat++; a >>= 1; }
                            ac[tbl][(run << 4) | cat]++;
                            run = 0;
                        }
                        if (run > 0) ac[tbl][0]++;
                    }
}

typedef struct { float q[64], lambda, w[64], len[256]; } Rdo;

static int cmp_float(const void* a, const void* b) {
    float x = *(const float*)a, y = *(const float*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* jpeg/src/rdo.rs:30-48 (FREQ_WEIGHT_GAMMA = 1). */
static void rdo_new(Rdo* r, const uint8_t table[64], float kappa, const uint8_t sizes[256]) {
    for (int k = 0; k < 64; k++) r->q[k] = (float)table[ZIGZAG[k]];
    float sorted[63];
    memcpy(sorted, r->q + 1, sizeof(sorted));
    qsort(sorted, 63, sizeof(float), cmp_float);
    const float median = sorted[31];
    for (int s = 0; s < 256; s++) r->len[s] = sizes[s] == 0 ? 18.0f : (float)sizes[s];
    r->w[0] = 1.0f;
    for (int k = 1; k < 64; k++) r->w[k] = powf(median / r->q[k], 1.0f);
    r->lambda = kappa * median * median;
}

/* jpeg/src/rdo.rs:50-55 */
static float rdo_rate(const Rdo* r, int run, uint32_t v) {
    uint32_t size = 0, t = v;
    while (t) { size++; t >>= 1; }
    int zrl = run / 16;
    return (float)zrl * r->len[0xF0] + r->len[((run % 16) << 4) | (int)size] + (float)size;
}

/* jpeg/src/rdo.rs:57-113 */
void orc_rdo_quantize_ac(const uint8_t table[64], float kappa, const uint8_t sizes[256], const float a[64],
                         int16_t out[64]);
static void rdo_quantize_ac(const Rdo* r, const float a[64], int16_t out[64]) {
    typedef struct { int pos; uint32_t v; float cost; int prev; } Node;
    Node nodes[126];
    int nn = 0;
    float zd[64];
    zd[0] = 0.0f;
    for (int k = 1; k < 64; k++) {
        float e = a[k] * r->q[k];
        zd[k] = zd[k - 1] + e * e * r->w[k];
    }
    for (int i = 1; i < 64; i++) {
        float m = fabsf(a[i]);
        uint32_t v0 = (uint32_t)(m + 0.5f);
        if (v0 == 0) continue;
        uint32_t lo = v0 > 1 ? v0 - 1 : v0;
        for (uint32_t v = lo; v <= v0; v++) {
            float e = (m - (float)v) * r->q[i];
            float dnz = e * e * r->w[i];
            float best = zd[i - 1] + r->lambda * rdo_rate(r, i - 1, v) + dnz;
            int prev = -1;
            for (int ni = 0; ni < nn; ni++) {
                if (nodes[ni].pos >= i) continue;
                float c = nodes[ni].cost + (zd[i - 1] - zd[nodes[ni].pos]) +
                          r->lambda * rdo_rate(r, i - nodes[ni].pos - 1, v) + dnz;
                if (c < best) { best = c; prev = ni; }
            }
            nodes[nn].pos = i; nodes[nn].v = v; nodes[nn].cost = best; nodes[nn].prev = prev;
            nn++;
        }
    }
    float eob = r->lambda * r->len[0];
    float best_cost = zd[63] + eob;
    int best_end = -1;
    for (int ni = 0; ni < nn; ni++) {
        float c = nodes[ni].cost + (zd[63] - zd[nodes[ni].pos]) + (nodes[ni].pos < 63 ? eob : 0.0f);
        if (c < best_cost) { best_cost = c; best_end = ni; }
    }
    for (int k = 1; k < 64; k++) out[k] = 0;
    for (int cur = best_end; cur != -1; cur = nodes[cur].prev) {
        int16_t v = (int16_t)nodes[cur].v;
        out[nodes[cur].pos] = a[nodes[cur].pos] < 0.0f ? (int16_t)-v : v;
    }
}

void orc_rdo_quantize_ac(const uint8_t table[64], float kappa, const uint8_t sizes[256], const float a[64],
                         int16_t out[64]) {
    Rdo r;
    rdo_new(&r, table, kappa, sizes);
    rdo_quantize_ac(&r, a, out);
}

/* jpeg/src/encode.rs:527-537 with TRELLIS_KAPPA (encode.rs:42): code lengths of the plain pass, then every block's
 * DC rounded and clamped and its AC re-quantised. Components: 1 (gray) or 3 (Y, Cb, Cr); class = min(c, 1).
 * width/height give the MCU grid the code-length statistics walk. */
int orc_trellis(int width, int height, int ncomp, const int* comp_h, const int* comp_v, const int* blocks_w,
                const int* blocks_h, const int16_t* const* plain, const float* const* scaled, const uint8_t* qy,
                const uint8_t* qc, int16_t* const* out) {
    int hmax = 1, vmax = 1;
    for (int c = 0; c < ncomp; c++) {
        if (comp_h[c] > hmax) hmax = comp_h[c];
        if (comp_v[c] > vmax) vmax = comp_v[c];
    }
    const int mcus_x = (width + 8 * hmax - 1) / (8 * hmax), mcus_y = (height + 8 * vmax - 1) / (8 * vmax);
    uint32_t ac[2][256];
    ac_counts(ncomp, plain, comp_h, comp_v, blocks_w, mcus_x, mcus_y, ac);
    uint8_t sizes[2][256];
    optimal_sizes(ac[0], sizes[0]);
    optimal_sizes(ac[1], sizes[1]);
    Rdo r[2];
    rdo_new(&r[0], qy, 0.02f, sizes[0]);
    rdo_new(&r[1], qc, 0.02f, sizes[1]);
    for (int c = 0; c < ncomp; c++) {
        const Rdo* rr = &r[c < 1 ? c : 1];
        const long nb = (long)blocks_w[c] * blocks_h[c];
#pragma omp parallel for schedule(dynamic, 64)
        for (long b = 0; b < nb; b++) {
            const float* a = scaled[c] + b * 64;
            int16_t* o = out[c] + b * 64;
            int32_t dc = round_f32(a[0]);
            o[0] = (int16_t)(dc < -2047 ? -2047 : (dc > 2047 ? 2047 : dc));
            rdo_quantize_ac(rr, a, o);
        }
    }
    return 0;
}

/* jpeg/src/encode.rs:107-171 -- CoefImage::requantize(quality, trellis). Components c = 0..ncomp-1 with table slot
 * tq[c] (qtabs[slot] natural order or NULL = all ones), sampling comp_h/comp_v, grid blocks_w x blocks_h; adobe is
 * the Adobe transform (-1 = none). out[c] receives the coefficients, new_tabs[4][64] the tables by slot,
 * present[4] which slots hold one, new_tq[c] each component's slot. */
int orc_requantize(int width, int height, int ncomp, const int* tq, const int* comp_h, const int* comp_v,
                   const int* blocks_w, const int* blocks_h, const int16_t* const* coefs, const uint8_t* const* qtabs,
                   int adobe, int quality, int trellis, int16_t* const* out, uint8_t* new_tabs, int* present,
                   int* new_tq) {
    uint8_t std_tab[64];
    orc_scale_table(quality, std_tab); /* luma_table and chroma_table share the base (quality.rs:37-43) */
    float ratio[4][64];
    uint8_t tabs[4][64];
    int changed = 0;
    memset(present, 0, sizeof(int) * 4);
    memset(new_tabs, 0, 4 * 64);
    for (int c = 0; c < ncomp; c++) {
        uint8_t old[64];
        if (qtabs[tq[c]]) memcpy(old, qtabs[tq[c]], 64);
        else memset(old, 1, 64);
        for (int n = 0; n < 64; n++) {
            tabs[c][n] = std_tab[n] > old[n] ? std_tab[n] : old[n];
            if (tabs[c][n] != old[n]) changed = 1;
        }
        for (int k = 0; k < 64; k++) ratio[c][k] = (float)old[ZIGZAG[k]] / (float)tabs[c][ZIGZAG[k]];
        const long n = (long)blocks_w[c] * blocks_h[c] * 64;
        for (long i = 0; i < n; i++) {
            int32_t r = round_f32((float)coefs[c][i] * ratio[c][i % 64]);
            out[c][i] = (int16_t)(r < -2047 ? -2047 : (r > 2047 ? 2047 : r));
        }
        int slot = -1;
        for (int s2 = 0; s2 < 4 && slot < 0; s2++)
            if (present[s2] && memcmp(new_tabs + s2 * 64, tabs[c], 64) == 0) slot = s2;
        if (slot < 0) {
            for (int s2 = 0; s2 < 4 && slot < 0; s2++)
                if (!present[s2]) slot = s2;
            present[slot] = 1;
            memcpy(new_tabs + slot * 64, tabs[c], 64);
        }
        new_tq[c] = slot;
    }
    (void)adobe;
    if (!trellis || !changed) return 0;
    int hmax = 1, vmax = 1;
    for (int c = 0; c < ncomp; c++) {
        if (comp_h[c] > hmax) hmax = comp_h[c];
        if (comp_v[c] > vmax) vmax = comp_v[c];
    }
    uint32_t ac[2][256];
    ac_counts(ncomp, (const int16_t* const*)out, comp_h, comp_v, blocks_w, (width + 8 * hmax - 1) / (8 * hmax),
              (height + 8 * vmax - 1) / (8 * vmax), ac);
    uint8_t sizes[2][256];
    optimal_sizes(ac[0], sizes[0]);
    optimal_sizes(ac[1], sizes[1]);
    for (int c = 0; c < ncomp; c++) {
        Rdo r;
        rdo_new(&r, tabs[c], 0.02f, sizes[c < 1 ? c : 1]);
        const long nb = (long)blocks_w[c] * blocks_h[c];
#pragma omp parallel for schedule(dynamic, 64)
        for (long b = 0; b < nb; b++) {
            float a[64];
            for (int k = 0; k < 64; k++) a[k] = (float)coefs[c][b * 64 + k] * ratio[c][k];
            rdo_quantize_ac(&r, a, out[c] + b * 64);
        }
    }
    return 0;
}

static const double C2 = 0.0009;
static const double SIGMA = 1.5;
static const double M02 = 0.078, M00 = 0.30, M12 = 0.078, M10 = 0.23;
static const double M20 = 0.2434226892454782, M21 = 0.2047674442449682;
static const double BIAS = 0.0037930732552754493;

typedef struct {
    int w, h;
    float* ch[3];
} Planes;

/* core/src/ssimulacra2.rs:173-176 */
static double srgb_to_linear(double v) { return v <= 0.04045 ? v / 12.92 : pow((v + 0.055) / 1.055, 2.4); }

/* core/src/ssimulacra2.rs:179-221 (8-bit Gray and RGB; Gray is replicated to three channels). */
static int linear_planes(const uint8_t* px, int w, int h, int channels, Planes* p) {
    float lut[256];
    for (int v = 0; v < 256; v++) lut[v] = (float)srgb_to_linear((double)v / 255.0);
    size_t n = (size_t)w * h;
    p->w = w;
    p->h = h;
    for (int c = 0; c < 3; c++) p->ch[c] = (float*)malloc(n * sizeof(float));
    for (size_t i = 0; i < n; i++) {
        if (channels == 1) {
            float l = lut[px[i]];
            p->ch[0][i] = l; p->ch[1][i] = l; p->ch[2][i] = l;
        } else {
            p->ch[0][i] = lut[px[i * 3]];
            p->ch[1][i] = lut[px[i * 3 + 1]];
            p->ch[2][i] = lut[px[i * 3 + 2]];
        }
    }
    return 0;
}

static void free_planes(Planes* p) {
    for (int c = 0; c < 3; c++) { free(p->ch[c]); p->ch[c] = 0; }
}

/* core/src/ssimulacra2.rs:225-247 -- 2x2 mean in linear light, partial blocks at odd edges. */
static Planes downsample(const Planes* p) {
    Planes o;
    int w = p->w, h = p->h;
    o.w = (w + 1) / 2;
    o.h = (h + 1) / 2;
    for (int c = 0; c < 3; c++) {
        const float* src = p->ch[c];
        float* out = (float*)malloc((size_t)o.w * o.h * sizeof(float));
        for (int oy = 0; oy < o.h; oy++) {
            int y0 = oy * 2, y1 = oy * 2 + 2 < h ? oy * 2 + 2 : h;
            for (int ox = 0; ox < o.w; ox++) {
                int x0 = ox * 2, x1 = ox * 2 + 2 < w ? ox * 2 + 2 : w;
                double sum = 0.0;
                for (int y = y0; y < y1; y++)
                    for (int x = x0; x < x1; x++) sum += (double)src[(size_t)y * w + x];
                out[(size_t)oy * o.w + ox] = (float)(sum / (double)((y1 - y0) * (x1 - x0)));
            }
        }
        o.ch[c] = out;
    }
    return o;
}

/* core/src/ssimulacra2.rs:251-258 */
static void xyb(double r, double g, double b, double out[3]) {
    const double M01 = 1.0 - M02 - M00, M11 = 1.0 - M12 - M10, M22 = 1.0 - M20 - M21;
    double cb = cbrt(BIAS);
    double t0 = M00 * r + M01 * g + M02 * b + BIAS;
    double t1 = M10 * r + M11 * g + M12 * b + BIAS;
    double t2 = M20 * r + M21 * g + M22 * b + BIAS;
    double m0 = cbrt(t0 > 0.0 ? t0 : 0.0) - cb;
    double m1 = cbrt(t1 > 0.0 ? t1 : 0.0) - cb;
    double m2 = cbrt(t2 > 0.0 ? t2 : 0.0) - cb;
    double x = 0.5 * (m0 - m1);
    double y = 0.5 * (m0 + m1);
    out[0] = x * 14.0 + 0.42;
    out[1] = y + 0.01;
    out[2] = (m2 - y) + 0.55;
}

/* core/src/ssimulacra2.rs:261-271 */
static void kernel(double k[2 * RADIUS + 1]) {
    double sum = 0.0;
    for (int i = 0; i < 2 * RADIUS + 1; i++) {
        double x = (double)i - (double)RADIUS;
        k[i] = exp(-0.5 / (SIGMA * SIGMA) * x * x);
        sum += k[i];
    }
    for (int i = 0; i < 2 * RADIUS + 1; i++) k[i] /= sum;
}

/* core/src/ssimulacra2.rs:275-319 -- horizontal blur, zero beyond the left/right edges. */
static void blur_row(const double* src, double* dst, int w, const double k[11]) {
    if (w > 2 * RADIUS) {
        int lo = RADIUS, hi = w - RADIUS;
        double centre = k[RADIUS];
        for (int x = lo; x < hi; x++) dst[x] = centre * src[x];
        for (int t = 1; t <= RADIUS; t++) {
            double kt = k[RADIUS + t];
            for (int x = lo; x < hi; x++) dst[x] += kt * (src[x - t] + src[x + t]);
        }
        for (int x = 0; x < w; x++) {
            if (x >= lo && x < hi) continue;
            double s = 0.0;
            for (int t = 0; t < 11; t++) {
                int i = x + t - RADIUS;
                s += k[t] * ((i < 0 || i > w - 1) ? 0.0 : src[i]);
            }
            dst[x] = s;
        }
    } else {
        for (int x = 0; x < w; x++) {
            double s = 0.0;
            for (int t = 0; t < 11; t++) {
                int i = x + t - RADIUS;
                s += k[t] * ((i < 0 || i > w - 1) ? 0.0 : src[i]);
            }
            dst[x] = s;
        }
    }
}

/* core/src/ssimulacra2.rs:338-444 -- sums of the SSIM, artifact and detail maps (and their 4th powers) over
 * output rows y0..y1, per channel. */
static void stripe_sums(const Planes* a, const Planes* b, int y0, int y1, const double k[11], double sums[3][6]) {
    int w = a->w, h = a->h;
    int lo = y0 > RADIUS ? y0 - RADIUS : 0;
    int hi = y1 + RADIUS < h ? y1 + RADIUS : h;
    int rows = hi - lo;
    size_t rw = (size_t)rows * w;
    double* xa[3];
    double* xb[3];
    for (int c = 0; c < 3; c++) {
        xa[c] = (double*)malloc(rw * sizeof(double));
        xb[c] = (double*)malloc(rw * sizeof(double));
    }
    for (int r = 0; r < rows; r++) {
        size_t base = (size_t)(lo + r) * w;
        for (int x = 0; x < w; x++) {
            size_t i = base + x;
            double pa[3], pb[3];
            xyb((double)a->ch[0][i], (double)a->ch[1][i], (double)a->ch[2][i], pa);
            xyb((double)b->ch[0][i], (double)b->ch[1][i], (double)b->ch[2][i], pb);
            for (int c = 0; c < 3; c++) {
                xa[c][(size_t)r * w + x] = pa[c];
                xb[c][(size_t)r * w + x] = pb[c];
            }
        }
    }
    memset(sums, 0, sizeof(double) * 18);
    double* h5[5];
    double* acc[5];
    for (int q = 0; q < 5; q++) {
        h5[q] = (double*)malloc(rw * sizeof(double));
        acc[q] = (double*)malloc((size_t)w * sizeof(double));
    }
    double* tmp = (double*)malloc((size_t)w * sizeof(double));
    for (int c = 0; c < 3; c++) {
        for (int r = 0; r < rows; r++) {
            const double* ra = xa[c] + (size_t)r * w;
            const double* rb = xb[c] + (size_t)r * w;
            blur_row(ra, h5[0] + (size_t)r * w, w, k);
            blur_row(rb, h5[1] + (size_t)r * w, w, k);
            for (int x = 0; x < w; x++) tmp[x] = ra[x] * ra[x];
            blur_row(tmp, h5[2] + (size_t)r * w, w, k);
            for (int x = 0; x < w; x++) tmp[x] = rb[x] * rb[x];
            blur_row(tmp, h5[3] + (size_t)r * w, w, k);
            for (int x = 0; x < w; x++) tmp[x] = ra[x] * rb[x];
            blur_row(tmp, h5[4] + (size_t)r * w, w, k);
        }
        double* s = sums[c];
        for (int y = y0; y < y1; y++) {
            for (int q = 0; q < 5; q++) {
                /* Rows beyond the top/bottom edges are mirrored (core/src/ssimulacra2.rs:385-397). */
#define ROW_OF(yy_)                                                                          \
    (h5[q] + (size_t)(((yy_) < 0 ? -(yy_)-1 : ((yy_) >= h ? 2 * h - 1 - (yy_) : (yy_))) - lo) * w)
                const double* centre = ROW_OF(y);
                for (int x = 0; x < w; x++) acc[q][x] = k[RADIUS] * centre[x];
                for (int t = 1; t <= RADIUS; t++) {
                    double kt = k[RADIUS + t];
                    const double* up = ROW_OF(y - t);
                    const double* down = ROW_OF(y + t);
                    for (int x = 0; x < w; x++) acc[q][x] += kt * (up[x] + down[x]);
                }
#undef ROW_OF
            }
            const double* ra = xa[c] + (size_t)(y - lo) * w;
            const double* rb = xb[c] + (size_t)(y - lo) * w;
            for (int x = 0; x < w; x++) {
                double m1 = acc[0][x], m2 = acc[1][x];
                double s11 = acc[2][x], s22 = acc[3][x], s12 = acc[4][x];
                double num_m = 1.0 - (m1 - m2) * (m1 - m2);
                double num_s = 2.0 * (s12 - m1 * m2) + C2;
                double denom_s = (s11 - m1 * m1) + (s22 - m2 * m2) + C2;
                double d = 1.0 - num_m * num_s / denom_s;
                d = d > 0.0 ? d : 0.0;
                double d1 = (1.0 + fabs(rb[x] - m2)) / (1.0 + fabs(ra[x] - m1)) - 1.0;
                double art = d1 > 0.0 ? d1 : 0.0;
                double det = -d1 > 0.0 ? -d1 : 0.0;
                s[0] += d;
                s[1] += (d * d) * (d * d);
                s[2] += art;
                s[3] += (art * art) * (art * art);
                s[4] += det;
                s[5] += (det * det) * (det * det);
            }
        }
    }
    free(tmp);
    for (int q = 0; q < 5; q++) { free(h5[q]); free(acc[q]); }
    for (int c = 0; c < 3; c++) { free(xa[c]); free(xb[c]); }
}

/* core/src/ssimulacra2.rs:448-466 */
static void scale_stats(const Planes* a, const Planes* b, double out[3][6]) {
    double k[11];
    kernel(k);
    int n = (a->h + STRIPE - 1) / STRIPE;
    double(*parts)[3][6] = (double(*)[3][6])malloc(sizeof(double) * 18 * (size_t)n);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; i++) {
        int y1 = (i + 1) * STRIPE < a->h ? (i + 1) * STRIPE : a->h;
        stripe_sums(a, b, i * STRIPE, y1, k, parts[i]);
    }
    double total[3][6];
    memset(total, 0, sizeof(total));
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++)
            for (int m = 0; m < 6; m++) total[c][m] += parts[i][c][m];
    free(parts);
    double count = (double)((size_t)a->w * a->h);
    for (int c = 0; c < 3; c++)
        for (int m = 0; m < 3; m++) {
            out[c][2 * m] = total[c][2 * m] / count;
            out[c][2 * m + 1] = pow(total[c][2 * m + 1] / count, 0.25);
        }
}

/* core/src/ssimulacra2.rs:491-511 -- weighted pooling of the per-scale statistics and the final mapping. */
double orc_ssim2_pool(const double* stats, int nscales) {
    double ssim = 0.0;
    int i = 0;
    for (int c = 0; c < 3; c++)
        for (int sc = 0; sc < nscales; sc++)
            for (int nrm = 0; nrm < 2; nrm++) {
                const double* st = stats + ((size_t)sc * 3 + c) * 6;
                ssim += ORC_WEIGHTS[i] * fabs(st[nrm]);
                i++;
                ssim += ORC_WEIGHTS[i] * fabs(st[2 + nrm]);
                i++;
                ssim += ORC_WEIGHTS[i] * fabs(st[4 + nrm]);
                i++;
            }
    ssim *= 0.9562382616834844;
    ssim = 2.326765642916932 * ssim - 0.020884521182843837 * ssim * ssim + 6.248496625763138e-5 * ssim * ssim * ssim;
    return ssim > 0.0 ? 100.0 - 10.0 * pow(ssim, 0.6276336467831387) : 100.0;
}

/*
 * core/src/ssimulacra2.rs:471-512. a/b: Gray8 or RGB8 images of the same w x h.
 * stats_out (may be NULL): [scale][channel][6] for the scales computed, nscales_out their count.
 * Returns 0 and the score in *score, or 1 (invalid: size below 8x8 or a non-finite score).
 */
int orc_ssimulacra2(const uint8_t* a, const uint8_t* b, int w, int h, int channels, double* score, double* stats_out,
                    int* nscales_out) {
    if (w < MIN_SIDE || h < MIN_SIDE || (channels != 1 && channels != 3)) return 1;
    Planes pa, pb;
    linear_planes(a, w, h, channels, &pa);
    linear_planes(b, w, h, channels, &pb);
    double scales[NUM_SCALES][3][6];
    int ns = 0;
    for (int s = 0; s < NUM_SCALES; s++) {
        if (pa.w < MIN_SIDE || pa.h < MIN_SIDE) break;
        scale_stats(&pa, &pb, scales[ns]);
        ns++;
        if (s + 1 < NUM_SCALES) {
            Planes na = downsample(&pa), nb = downsample(&pb);
            free_planes(&pa);
            free_planes(&pb);
            pa = na;
            pb = nb;
        }
    }
    free_planes(&pa);
    free_planes(&pb);
    if (stats_out) memcpy(stats_out, scales, sizeof(double) * 18 * (size_t)ns);
    if (nscales_out) *nscales_out = ns;
    double sc = orc_ssim2_pool(&scales[0][0][0], ns);
    if (!isfinite(sc)) return 1;
    *score = sc;
    return 0;
}

/* ============================================================================================================ */
/* jpeg/src/target.rs:29-56 with make = build_coefs (trellis off) and metric = ssimulacra2, as
 * cli/src/pipeline.rs:207-247 wires them. Returns 0; *quality and *score describe the chosen candidate. */
static double score_candidate(const uint8_t* px, int w, int h, int channels, int q, int subsampling) {
    int hmax, vmax;
    orc_sampling(channels, q, subsampling, &hmax, &vmax);
    int pw = (w + 8 * hmax - 1) / (8 * hmax) * 8 * hmax, ph = (h + 8 * vmax - 1) / (8 * vmax) * 8 * vmax;
    int bw[3] = {pw / 8, pw / hmax / 8, pw / hmax / 8}, bh[3] = {ph / 8, ph / vmax / 8, ph / vmax / 8};
    int16_t* co[3];
    for (int c = 0; c < 3; c++) co[c] = (int16_t*)malloc((size_t)bw[c] * bh[c] * 64 * sizeof(int16_t));
    uint8_t qt[128];
    orc_build_coefs(px, w, h, channels, q, subsampling, co[0], co[1], co[2], qt);
    int ch[3] = {hmax, 1, 1}, cv[3] = {vmax, 1, 1};
    const int16_t* cc[3] = {co[0], co[1], co[2]};
    const uint8_t* qq[3] = {qt, qt + 64, qt + 64};
    uint8_t* rec = (uint8_t*)malloc((size_t)w * h * channels);
    double s = -DBL_MAX;
    if (orc_reconstruct(w, h, channels, ch, cv, bw, bh, cc, qq, 0, -1, rec) == 0) {
        double v;
        if (orc_ssimulacra2(px, rec, w, h, channels, &v, NULL, NULL) == 0) s = v;
    }
    free(rec);
    for (int c = 0; c < 3; c++) free(co[c]);
    return s;
}

/* Bisection over a caller-given score table, for testing the search semantics in isolation:
 * scores[q-1] is the score of quality q. */
int orc_bisect_table(const double* scores, double target, int max_quality, int* quality, double* score) {
    int lo = 1, hi = max_quality > 1 ? max_quality : 1;
    double top = scores[hi - 1];
    *quality = hi;
    *score = top;
    if (top < target) return 0;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        double s = scores[mid - 1];
        if (s >= target) {
            hi = mid;
            *quality = mid;
            *score = s;
        } else {
            lo = mid + 1;
        }
    }
    return 0;
}

int orc_find_quality(const uint8_t* px, int w, int h, int channels, double target, int max_quality, int subsampling,
                     int* quality, double* score, int* steps) {
    if (w < MIN_SIDE || h < MIN_SIDE) return 2;
    int lo = 1, hi = max_quality > 1 ? max_quality : 1;
    int n = 1;
    double top = score_candidate(px, w, h, channels, hi, subsampling);
    *quality = hi;
    *score = top;
    if (top >= target) {
        while (lo < hi) {
            int mid = lo + (hi - lo) / 2;
            double s = score_candidate(px, w, h, channels, mid, subsampling);
            n++;
            if (s >= target) {
                hi = mid;
                *quality = mid;
                *score = s;
            } else {
                lo = mid + 1;
            }
        }
    }
    if (steps) *steps = n;
    return 0;
}
