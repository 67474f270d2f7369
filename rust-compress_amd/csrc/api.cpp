// C ABI of librcgpu (include/rcgpu.h): context management, layouts and tables, and the drop-in entry points for
// build_coefs, reconstruct, ssimulacra2 and the target-quality search, in host-buffer and device-resident forms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rcg_host.h"

namespace rcg {

static thread_local std::string g_last_error;
static thread_local const char* g_last_fixed = nullptr;  // set when the message itself could not be stored
void set_error(const std::string& msg) noexcept {
    try {
        g_last_error = msg;
        g_last_fixed = nullptr;
    } catch (...) {
        g_last_fixed = "out of host memory (while recording an error)";
    }
}

// Fault injection for tests: with RCG_TEST_ARENA_LIMIT=<bytes> in the environment, a call whose working set needs
// more than that many bytes of workspace fails with RCG_OOM, as if the device could not hold it.
static size_t test_arena_limit() {
    const char* v = std::getenv("RCG_TEST_ARENA_LIMIT");
    return v ? size_t(std::strtoull(v, nullptr, 10)) : 0;
}

int Arena::ensure(size_t bytes) {
    const size_t limit = test_arena_limit();
    if (limit && bytes > limit) {
        set_error("device workspace of " + std::to_string(bytes) + " bytes: above RCG_TEST_ARENA_LIMIT");
        return RCG_OOM;
    }
    if (bytes <= cap) return RCG_OK;
    if (base) (void)hipFree(base);
    base = nullptr;
    cap = 0;
    size_t want = std::max(bytes, size_t(64) << 20);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&base), want);
    if (e != hipSuccess) {
        base = nullptr;
        set_error(std::string("device workspace of ") + std::to_string(want) + " bytes: " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? RCG_OOM : RCG_DEVICE;
    }
    cap = want;
    return RCG_OK;
}

int Pinned::ensure(size_t bytes) {
    if (bytes <= cap) return RCG_OK;
    if (base) (void)hipHostFree(base);
    base = nullptr;
    cap = 0;
    size_t want = std::max(bytes, size_t(1) << 20);
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&base), want, hipHostMallocDefault);
    if (e != hipSuccess) {
        base = nullptr;
        set_error(std::string("pinned staging: ") + hipGetErrorString(e));
        return RCG_OOM;
    }
    cap = want;
    return RCG_OK;
}

// jpeg/src/quality.rs:26-35
static const uint8_t kPerceptualQ[64] = {
    16, 16, 16, 18,  25,  37,  56,  85,  16, 17, 20,  27,  34,  40,  53,  75,  16, 20, 24,  31,  43,  62,
    91, 135, 18, 27, 31,  40,  53,  74,  106, 156, 25, 34, 43,  53,  69,  94,  131, 189, 37, 40, 62,  74,
    94, 124, 169, 238, 56, 53, 91,  106, 131, 169, 226, 255, 85, 75, 135, 156, 189, 238, 255, 255};

// jpeg/src/quality.rs:6-14 (IJG scaling); luma and chroma scale the same base (quality.rs:37-43).
void quant_tables(int quality, uint8_t luma[64], uint8_t chroma[64]) {
    const unsigned q = unsigned(std::min(100, std::max(1, quality)));
    const unsigned scale = q < 50 ? 5000 / q : 200 - 2 * q;
    for (int i = 0; i < 64; i++) {
        unsigned t = (unsigned(kPerceptualQ[i]) * scale + 50) / 100;
        t = std::min(255u, std::max(1u, t));
        luma[i] = chroma[i] = uint8_t(t);
    }
}

static const float kAan[8] = {1.0f, 1.3870398f, 1.306563f, 1.1758756f, 1.0f, 0.78569495f, 0.5411961f, 0.27589938f};

// Divisors (jpeg/src/encode.rs:175-183) and dequantisation multipliers (jpeg/src/decode.rs:1317-1322), f32 in the
// reference's order; this file is compiled with -ffp-contract=off.
static void tables_for(const uint8_t q[64], float div[64], float mult[64]) {
    for (int v = 0; v < 8; v++)
        for (int u = 0; u < 8; u++) {
            div[u * 8 + v] = 1.0f / (float(q[v * 8 + u]) * kAan[u] * kAan[v] * 8.0f);
            mult[v * 8 + u] = float(q[v * 8 + u]) * kAan[u] * kAan[v];
        }
}

void fill_cand_tables(const uint8_t qy[64], const uint8_t qc[64], EncCand& c) {
    tables_for(qy, c.div[0], c.mult[0]);
    tables_for(qc, c.div[1], c.mult[1]);
}

// jpeg/src/encode.rs:428-443
Layout make_layout(uint32_t w, uint32_t h, uint32_t channels, int quality, int subsampling) {
    Layout L{};
    L.w = w;
    L.h = h;
    L.channels = channels;
    if (channels == 1) {
        L.hmax = L.vmax = 1;
    } else {
        switch (subsampling) {
            case RCG_S420: L.hmax = 2; L.vmax = 2; break;
            case RCG_S422: L.hmax = 2; L.vmax = 1; break;
            case RCG_S444: L.hmax = 1; L.vmax = 1; break;
            default:
                L.hmax = L.vmax = quality < 90 ? 2 : 1;
        }
    }
    const uint32_t mw = 8 * L.hmax, mh = 8 * L.vmax;
    L.pw = (w + mw - 1) / mw * mw;
    L.ph = (h + mh - 1) / mh * mh;
    L.ncomp = channels == 1 ? 1 : 3;
    L.bw[0] = L.pw / 8;
    L.bh[0] = L.ph / 8;
    for (int c = 1; c < 3; c++) {
        L.bw[c] = channels == 1 ? 0 : L.pw / L.hmax / 8;
        L.bh[c] = channels == 1 ? 0 : L.ph / L.vmax / 8;
    }
    return L;
}

// The largest block grids a search over every quality can produce. Under Auto a search meets both 4:4:4 (q >= 90)
// and 4:2:0 (q < 90), and 4:2:0 pads luma to 16-pixel MCUs, so its luma grid can exceed the 4:4:4 one by a block
// row or column (9x8 pixels: 2x2 blocks against 2x1). Buffers of a search are sized by this layout.
Layout max_layout(uint32_t w, uint32_t h, uint32_t channels, int subsampling) {
    Layout L = make_layout(w, h, channels, 100, subsampling);
    if (subsampling == RCG_AUTO && channels != 1) {
        const Layout S = make_layout(w, h, channels, 1, subsampling);
        L.pw = std::max(L.pw, S.pw);
        L.ph = std::max(L.ph, S.ph);
        for (int c = 0; c < 3; c++) {
            L.bw[c] = std::max(L.bw[c], S.bw[c]);
            L.bh[c] = std::max(L.bh[c], S.bh[c]);
        }
    }
    return L;
}

constexpr size_t kTailMinPixels = size_t(1) << 19;

PairPlan plan_pair(uint32_t w, uint32_t h, const SsimTune& T) {
    PairPlan P{};
    P.w = w;
    P.h = h;
    uint32_t sw = w, sh = h;
    for (int s = 0; s < 6; s++) {
        if (sw < 8 || sh < 8) break;
        P.sw[s] = sw;
        P.sh[s] = sh;
        P.nt[s] = T.fixed_nt ? uint32_t(T.fixed_nt) : (sw <= T.narrow_w ? 128u : uint32_t(T.wide_nt));
        P.th[s] = uint32_t(T.th[T.nclass - 1]);
        for (int i = 0; i < T.nclass; i++)
            if (sh >= T.th_min_h[i]) {
                P.th[s] = uint32_t(T.th[i]);
                break;
            }
        P.tiles_y[s] = ssim_tiles_y(sh, int(P.th[s]));
        // FP32: the column tiling with the fewest lanes (a lane per column plus 12 halo lanes per tile; every lane
        // walks the same rows, so lanes measure the work) among whole tiles of one thread count (the scale's class:
        // 256 or 128 threads, or narrower) and, on scales of at least kTailMinPixels, whole tiles followed by one
        // column of narrower tail tiles for the columns left over (at most 52 or 116). A tail is a second launch on
        // a second stream, joined before the next scale: on a smaller scale, whose tiles leave most of the GPU idle
        // when one image is searched alone, that costs more than its lanes save (measured: the per-image search over
        // 256 images of 64..3000 px was 4% slower with tails on every scale, the batched pass 0.25% faster). On a
        // tie the single thread count wins, then the wider tiles. The FP64 kernel has one thread count for every tile.
        const uint32_t nt_class = P.nt[s];
        P.tnt[s] = 0;
        P.tiles_x[s] = ssim_tiles_x(sw, int(nt_class));
        if (!T.fixed_nt) {
            uint32_t best = P.tiles_x[s] * nt_class, best_launches = 1;
            auto consider = [&](uint32_t lanes, uint32_t launches, uint32_t nt, uint32_t tiles_x, uint32_t tnt) {
                if (lanes > best || (lanes == best && launches >= best_launches)) return;
                best = lanes;
                best_launches = launches;
                P.nt[s] = nt;
                P.tiles_x[s] = tiles_x;
                P.tnt[s] = tnt;
            };
            for (uint32_t nt : {256u, 128u, 64u}) {
                if (nt > nt_class) continue;
                const uint32_t ow = nt - 12, full = sw / ow, rest = sw - full * ow;
                const uint32_t whole = ssim_tiles_x(sw, int(nt));
                consider(whole * nt, 1, nt, whole, 0);
                if (size_t(sw) * sh < kTailMinPixels || full == 0 || rest == 0) continue;
                for (uint32_t tnt : {128u, 64u})
                    if (tnt < nt && rest <= tnt - 12) consider(full * nt + tnt, 2, nt, full, tnt);
            }
        }
        P.tiles[s] = P.tiles_x[s] * P.tiles_y[s] + (P.tnt[s] ? P.tiles_y[s] : 0);
        // The reference-cache kernel's output does not depend on its tiling: whole tiles of the scale's class thread
        // count, as tall as one round of the GPU's workgroup slots allows for this scale (at least 4 rows), so one
        // image keeps the GPU full without a partly idle last round, and small scales walk few rows.
        P.nt_ref[s] = nt_class;
        P.tiles_x_ref[s] = ssim_tiles_x(sw, int(P.nt_ref[s]));
        const uint32_t slots = P.nt_ref[s] == 128 ? T.ref_slots_narrow : T.ref_slots_wide;
        const uint32_t rows_of_tiles = std::max(1u, slots / P.tiles_x_ref[s]);
        uint32_t tr = (sh + rows_of_tiles - 1) / rows_of_tiles;
        tr = std::max(4u, (tr + 1) & ~1u);
        P.th_ref[s] = tr;
        P.tiles_ref[s] = P.tiles_x_ref[s] * ssim_tiles_y(sh, int(tr));
        P.nscales++;
        sw = (sw + 1) / 2;
        sh = (sh + 1) / 2;
    }
    return P;
}

}  // namespace rcg

using namespace rcg;

namespace {

bool valid_channels(uint32_t c) { return c == 1 || c == 3; }

// sRGB -> linear table of core/src/ssimulacra2.rs:173-176,191: f64 pow, stored as f32.
void make_lut(float lut[256]) {
    for (int v = 0; v < 256; v++) {
        const double x = double(v) / 255.0;
        const double l = x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4);
        lut[v] = float(l);
    }
}

// core/src/ssimulacra2.rs:261-271
void make_taps(double out[6]) {
    double k[11], sum = 0.0;
    for (int i = 0; i < 11; i++) {
        const double x = double(i) - 5.0;
        k[i] = std::exp(-0.5 / (1.5 * 1.5) * x * x);
        sum += k[i];
    }
    for (int i = 0; i < 11; i++) k[i] /= sum;
    for (int t = 0; t <= 5; t++) out[t] = k[5 + t];
}

// Bytes one SSIMULACRA2 evaluation of a w x h pair needs in the arena: the distorted image's pyramid, the reference
// side (FP32: its pyramid and cache; FP64: its pyramid), the tile partials and the results.
size_t ssim_bytes(const PairPlan& P, int precision) {
    size_t b = 0;
    for (uint32_t s = 1; s < P.nscales; s++) b += Arena::align(size_t(P.sw[s]) * P.sh[s] * 3 * sizeof(float));
    if (precision == RCG_PREC_FP64)
        b *= 2;
    else
        b += ref_state_bytes(P);
    for (uint32_t s = 0; s < P.nscales; s++) b += Arena::align(size_t(P.tiles[s]) * 18 * sizeof(double));
    b += 4096;
    return b;
}

struct SsimBuffers {
    RefState ref{};  // the reference side (FP64 uses ref.planes only)
    float* b_planes[6] = {};
    double* partials[6] = {};
    double* d_score = nullptr;
    double* d_stats = nullptr;
};

SsimBuffers take_ssim(Arena& A, const PairPlan& P, int precision) {
    SsimBuffers B;
    if (precision == RCG_PREC_FP64) {
        for (uint32_t s = 1; s < P.nscales; s++)
            B.ref.planes[s] = static_cast<float*>(A.take(size_t(P.sw[s]) * P.sh[s] * 3 * sizeof(float)));
    } else {
        take_ref_state(A, P, B.ref);
    }
    for (uint32_t s = 1; s < P.nscales; s++)
        B.b_planes[s] = static_cast<float*>(A.take(size_t(P.sw[s]) * P.sh[s] * 3 * sizeof(float)));
    for (uint32_t s = 0; s < P.nscales; s++)
        B.partials[s] = static_cast<double*>(A.take(size_t(P.tiles[s]) * 18 * sizeof(double)));
    B.d_score = static_cast<double*>(A.take(sizeof(double)));
    B.d_stats = static_cast<double*>(A.take(6 * 18 * sizeof(double)));
    return B;
}

// Stage the descriptors for every scale of one pair and its final pooling; launch_staged_ssim() enqueues them after
// Desc::upload(). a0/b0 are the scale-0 sources (a0 as pixels or pyramid, never the cache). FP32 builds the reference
// state (pyramid and cache) unless ref_ready and reads the reference side from the cache; FP64 computes both sides
// and writes the reference pyramid unless ref_ready.
struct StagedSsim {
    uint32_t kind_a0, kind_b0;
    const RefJob* ref_jobs[6];
    const SsimJob* jobs[6];
    const SsimJob* tail_jobs[6];  // the scale's tail tiles (PairPlan::tnt), or null
    const FinalDesc* fin;
    PairPlan P;
    int precision;
    bool build_ref;
};

StagedSsim stage_ssim(Desc& D, const PairPlan& P, const SsimSrc& a0, const SsimSrc& b0, const SsimBuffers& B,
                      bool ref_ready, int precision) {
    const bool f64 = precision == RCG_PREC_FP64;
    StagedSsim st;
    st.P = P;
    st.precision = precision;
    st.kind_a0 = a0.kind;
    st.kind_b0 = b0.kind;
    st.build_ref = !f64 && !ref_ready;
    for (uint32_t s = 0; s < P.nscales; s++) {
        st.ref_jobs[s] = nullptr;
        if (st.build_ref) {
            const RefJob r = make_ref_job(P, s, a0, B.ref);
            st.ref_jobs[s] = D.put(&r, 1);
        }
        SsimJob j;
        std::memset(&j, 0, sizeof(j));
        j.w = P.sw[s];
        j.h = P.sh[s];
        j.th = P.th[s];
        j.tiles_x = P.tiles_x[s];
        j.tiles_y = P.tiles_y[s];
        if (!f64)
            j.a = cache_src(B.ref.cache[s]);
        else
            j.a = s == 0 ? a0 : planes_src(B.ref.planes[s]);
        j.b = s == 0 ? b0 : planes_src(B.b_planes[s]);
        const bool next = s + 1 < P.nscales;
        j.a_next = (f64 && next && !ref_ready) ? B.ref.planes[s + 1] : nullptr;
        j.b_next = next ? B.b_planes[s + 1] : nullptr;
        j.partials = B.partials[s];
        st.jobs[s] = D.put(&j, 1);
        st.tail_jobs[s] = nullptr;
        if (P.tnt[s]) {
            j.x_begin = P.x_tail(int(s));
            j.tiles_x = 1;
            j.partials = B.partials[s] + size_t(P.tiles_main(int(s))) * 18;
            st.tail_jobs[s] = D.put(&j, 1);
        }
    }
    FinalDesc f;
    std::memset(&f, 0, sizeof(f));
    f.nscales = P.nscales;
    for (uint32_t s = 0; s < P.nscales; s++) {
        f.ntiles[s] = P.tiles[s];
        f.partials[s] = B.partials[s];
        f.count[s] = double(size_t(P.sw[s]) * P.sh[s]);
    }
    f.stats_out = B.d_stats;
    f.score_out = B.d_score;
    st.fin = D.put(&f, 1);
    return st;
}

// The reference-cache launches (when the pair builds the reference state) go to the side stream, which the caller
// has forked from the context's stream (fork_side()) when st.build_ref; work enqueued on the context's stream after
// the fork runs beside them. launch_staged_ssim() starts scale 0 once the scale-0 cache is built (ev_mid) and joins
// the side stream before scale 1.
int launch_staged_refs(rcg_ctx* ctx, const StagedSsim& st) {
    if (!st.build_ref) return RCG_OK;
    for (uint32_t s = 0; s < st.P.nscales; s++) {
        RCG_TRY(launch_ssim_ref(ctx, st.ref_jobs[s], nullptr, st.P.tiles_ref[s], int(st.P.nt_ref[s]),
                                s == 0 ? st.kind_a0 : uint32_t(kSrcPlanes), ctx->side));
        if (s == 0) RCG_TRY(mark_side(ctx));
    }
    if (st.P.nscales == 0) RCG_TRY(mark_side(ctx));  // launch_staged_ssim() waits for the mark in any case
    return RCG_OK;
}

// A scale's tail tiles run on the side stream beside its main tiles; the two streams join before the next scale,
// whose tiles read the pyramid level both wrote.
int launch_staged_ssim(rcg_ctx* ctx, const StagedSsim& st) {
    const bool f64 = st.precision == RCG_PREC_FP64;
    if (st.build_ref) RCG_TRY(wait_side_mark(ctx));
    bool side_open = st.build_ref;  // the side stream holds work the context's stream has not joined
    for (uint32_t s = 0; s < st.P.nscales; s++) {
        if (s > 0 && side_open) {
            RCG_TRY(join_side(ctx));
            side_open = false;
        }
        const uint32_t ka = !f64 ? uint32_t(kSrcRefCache) : (s == 0 ? st.kind_a0 : uint32_t(kSrcPlanes));
        const uint32_t kb = s == 0 ? st.kind_b0 : uint32_t(kSrcPlanes);
        if (st.tail_jobs[s]) {
            RCG_TRY(fork_side(ctx));
            RCG_TRY(launch_ssim_scale(ctx, st.precision, st.tail_jobs[s], nullptr, st.P.tiles_y[s],
                                      int(st.P.tnt[s]), ka, kb, ctx->side));
            side_open = true;
        }
        RCG_TRY(launch_ssim_scale(ctx, st.precision, st.jobs[s], nullptr, st.P.tiles_main(int(s)), int(st.P.nt[s]),
                                  ka, kb));
    }
    if (side_open) RCG_TRY(join_side(ctx));
    return launch_finalize(ctx, st.fin, 1);
}

constexpr size_t kDescBytes = 64 << 10;

// Begin a call: lock the context, size the arena and the descriptor staging.
struct Call {
    rcg_ctx* ctx;
    std::lock_guard<std::mutex> lock;
    Desc D;
    explicit Call(rcg_ctx* c) : ctx(c), lock(c->mu) {}
    int begin(size_t arena_bytes) {
        RCG_HIP(hipSetDevice(ctx->device));
        ctx->arena.reset();
        RCG_TRY(ctx->arena.ensure(arena_bytes + kDescBytes + 4096));
        RCG_TRY(ctx->pinned.ensure(kDescBytes));
        D.host = ctx->pinned.base;
        D.dev = static_cast<char*>(ctx->arena.take(kDescBytes));
        D.cap = kDescBytes;
        D.off = 0;
        return RCG_OK;
    }
};

// Encode (and optionally reconstruct) one image at one quality on the device.
EncJob make_enc_job(const Layout& L, const uint8_t* d_src) {
    EncJob j;
    std::memset(&j, 0, sizeof(j));
    j.src = d_src;
    j.src_bytes = uint64_t(L.w) * L.h * L.channels;
    j.w = L.w;
    j.h = L.h;
    j.channels = L.channels;
    j.hmax = L.hmax;
    j.vmax = L.vmax;
    j.pw = L.pw;
    j.ph = L.ph;
    j.tiles_x = (L.pw + kEncTileW - 1) / kEncTileW;
    j.tiles_y = L.ph / (8 * L.vmax);
    j.tile_begin = 0;
    return j;
}

// One encode job staged in D (not yet uploaded).
struct StagedEncode {
    const EncJob* job;
    const uint32_t* begins;
    uint32_t tiles;
};

StagedEncode stage_encode(Desc& D, const Layout& L, const uint8_t* d_src, int quality, int16_t* const d_coefs[3],
                          uint8_t* const d_rec[3], float* const d_scaled[3] = nullptr) {
    EncCand c;
    std::memset(&c, 0, sizeof(c));
    uint8_t qy[64], qc[64];
    quant_tables(quality, qy, qc);
    fill_cand_tables(qy, qc, c);
    for (int k = 0; k < 3; k++) {
        c.coef[k] = d_coefs ? d_coefs[k] : nullptr;
        c.rec[k] = d_rec ? d_rec[k] : nullptr;
        c.scaled[k] = d_scaled ? d_scaled[k] : nullptr;
    }
    EncJob j = make_enc_job(L, d_src);
    j.ncand = 1;
    j.cand = D.put(&c, 1);
    const uint32_t zero = 0;
    StagedEncode se;
    se.begins = D.put(&zero, 1);
    se.job = D.put(&j, 1);
    se.tiles = j.tiles_x * j.tiles_y;
    return se;
}

int stage_and_launch_encode(rcg_ctx* ctx, Desc& D, const Layout& L, const uint8_t* d_src, int quality,
                            int16_t* const d_coefs[3], uint8_t* const d_rec[3], float* const d_scaled[3] = nullptr) {
    const StagedEncode se = stage_encode(D, L, d_src, quality, d_coefs, d_rec, d_scaled);
    RCG_TRY(D.upload(ctx->stream));
    return launch_encode(ctx, se.job, se.begins, 1, se.tiles, !d_rec && !d_scaled);
}

size_t coef_bytes(const Layout& L) {
    size_t b = 0;
    for (uint32_t c = 0; c < L.ncomp; c++) b += Arena::align(L.coef_count(c) * sizeof(int16_t));
    return b;
}
size_t plane_bytes(const Layout& L) {
    size_t b = 0;
    for (uint32_t c = 0; c < L.ncomp; c++) b += Arena::align(L.plane_bytes(c));
    return b;
}

}  // namespace

int rcg::check_image(uint32_t w, uint32_t h, uint32_t channels) {
    if (!valid_channels(channels)) {
        set_error("pixels must be Gray8 (1 channel) or RGB8 (3 channels)");
        return RCG_UNSUPPORTED;
    }
    if (w == 0 || h == 0) {
        set_error("empty image");
        return RCG_INVALID;
    }
    if (w > 65535 || h > 65535) {
        set_error("JPEG dimensions are limited to 65535 pixels");
        return RCG_UNSUPPORTED;
    }
    return RCG_OK;
}

// Recon planes of the encode stage seen as reconstruct's sample planes (jpeg/src/decode.rs:1404-1415). Luma is
// full resolution and both chroma planes share one geometry, which the SSIMULACRA2 loader relies on.
SsimSrc rcg::recon_src(const Layout& L, uint8_t* const rec[3]) {
    SsimSrc s;
    std::memset(&s, 0, sizeof(s));
    s.kind = kSrcRecon;
    s.ncomp = L.ncomp;
    s.rgb_direct = 0;
    for (uint32_t c = 0; c < L.ncomp; c++) {
        const uint32_t hc = c == 0 ? L.hmax : 1, vc = c == 0 ? L.vmax : 1;
        s.comp[c].p = rec[c];
        s.comp[c].stride = L.bw[c] * 8;
        s.comp[c].cw = (L.w * hc + L.hmax - 1) / L.hmax;
        s.comp[c].ch = (L.h * vc + L.vmax - 1) / L.vmax;
        s.comp[c].fx = L.hmax / hc;
        s.comp[c].fy = L.vmax / vc;
    }
    return s;
}

SsimSrc rcg::bytes_src(const uint8_t* d, uint32_t channels) {
    SsimSrc s;
    std::memset(&s, 0, sizeof(s));
    s.kind = kSrcBytes;
    s.channels = channels;
    s.bytes = d;
    return s;
}

SsimSrc rcg::planes_src(const float* planes) {
    SsimSrc s;
    std::memset(&s, 0, sizeof(s));
    s.kind = kSrcPlanes;
    s.planes = planes;
    return s;
}

SsimSrc rcg::cache_src(const float* cache) {
    SsimSrc s;
    std::memset(&s, 0, sizeof(s));
    s.kind = kSrcRefCache;
    s.planes = cache;
    return s;
}

namespace {

int ssim_pair_dev(rcg_ctx* ctx, Call& call, const uint8_t* d_a, const uint8_t* d_b, uint32_t w, uint32_t h,
                  uint32_t channels, int precision, double* score, double* per_scale, int* nscales) {
    const SsimTune& T = precision == RCG_PREC_FP64 ? ctx->tune64 : ctx->tune32;
    const PairPlan P = plan_pair(w, h, T);
    SsimBuffers B = take_ssim(ctx->arena, P, precision);
    StagedSsim st = stage_ssim(call.D, P, bytes_src(d_a, channels), bytes_src(d_b, channels), B, false, precision);
    RCG_TRY(call.D.upload(ctx->stream));
    if (st.build_ref) RCG_TRY(fork_side(ctx));
    RCG_TRY(launch_staged_refs(ctx, st));
    RCG_TRY(launch_staged_ssim(ctx, st));
    double stats[6 * 18];
    double sc = 0.0;
    RCG_HIP(hipMemcpyAsync(&sc, B.d_score, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RCG_HIP(hipMemcpyAsync(stats, B.d_stats, sizeof(stats), hipMemcpyDeviceToHost, ctx->stream));
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    if (!std::isfinite(sc)) {
        set_error("SSIMULACRA 2 produced a non-finite score");
        return RCG_INVALID;
    }
    *score = sc;
    if (per_scale) {
        std::memset(per_scale, 0, sizeof(double) * 6 * 18);
        std::memcpy(per_scale, stats, sizeof(double) * 18 * P.nscales);
    }
    if (nscales) *nscales = int(P.nscales);
    return RCG_OK;
}

}  // namespace

// ================================================================================================================
// extern "C"

extern "C" {

const char* rcg_last_error(void) { return g_last_fixed ? g_last_fixed : g_last_error.c_str(); }

const char* rcg_version(void) { return "rcgpu 0.1 gfx950"; }

int rcg_create(int device, rcg_ctx** out) {
    return rcg::guarded([&]() -> int {
        if (!out) return RCG_INVALID;
        *out = nullptr;
        int n = 0;
        RCG_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) {
            set_error("no such HIP device");
            return RCG_INVALID;
        }
        RCG_HIP(hipSetDevice(device));
        rcg_ctx* c = new (std::nothrow) rcg_ctx();
        if (!c) return RCG_OOM;
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_mid, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
        if (e != hipSuccess) {
            set_error(std::string("stream/event creation: ") + hipGetErrorString(e));
            rcg_destroy(c);
            return RCG_DEVICE;
        }
        float lut[256];
        make_lut(lut);
        e = hipMalloc(reinterpret_cast<void**>(&c->d_lut), sizeof(lut));
        if (e == hipSuccess) e = hipMemcpy(c->d_lut, lut, sizeof(lut), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_error(std::string("lookup table: ") + hipGetErrorString(e));
            rcg_destroy(c);
            return RCG_DEVICE;
        }
        make_taps(c->taps64);
        for (int i = 0; i < 6; i++) c->taps32[i] = float(c->taps64[i]);
        c->tune64.fixed_nt = 128;
        int rc = ref_kernel_slots(device, &c->tune32.ref_slots_wide, &c->tune32.ref_slots_narrow);
        if (rc != RCG_OK) {
            rcg_destroy(c);
            return rc;
        }
        c->tune64.ref_slots_wide = c->tune32.ref_slots_wide;
        c->tune64.ref_slots_narrow = c->tune32.ref_slots_narrow;
        rc = upload_weights();
        if (rc == RCG_OK) rc = noalloc_setup(device);
        if (rc != RCG_OK) {
            rcg_destroy(c);
            return rc;
        }
        *out = c;
        return RCG_OK;
    });
}

void rcg_destroy(rcg_ctx* ctx) {
    rcg::guarded_void([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        if (ctx->side) (void)hipStreamSynchronize(ctx->side);
        if (ctx->arena.base) (void)hipFree(ctx->arena.base);
        if (ctx->pinned.base) (void)hipHostFree(ctx->pinned.base);
        if (ctx->d_lut) (void)hipFree(ctx->d_lut);
        if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
        if (ctx->ev_mid) (void)hipEventDestroy(ctx->ev_mid);
        if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
        if (ctx->side) (void)hipStreamDestroy(ctx->side);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int rcg_coef_layout(uint32_t width, uint32_t height, uint32_t channels, int quality, int subsampling,
                    rcg_layout* out) {
    return rcg::guarded([&]() -> int {
        if (!out) return RCG_INVALID;
        RCG_TRY(check_image(width, height, channels));
        const Layout L = make_layout(width, height, channels, quality, subsampling);
        out->width = width;
        out->height = height;
        out->ncomp = L.ncomp;
        out->hmax = L.hmax;
        out->vmax = L.vmax;
        for (int c = 0; c < 3; c++) {
            out->blocks_w[c] = L.bw[c];
            out->blocks_h[c] = L.bh[c];
        }
        return RCG_OK;
    });
}

int rcg_quant_tables(int quality, uint8_t out[128]) {
    return rcg::guarded([&]() -> int {
        if (!out) return RCG_INVALID;
        quant_tables(quality, out, out + 64);
        return RCG_OK;
    });
}

int rcg_build_coefs_dev(rcg_ctx* ctx, const uint8_t* d_pixels, uint32_t width, uint32_t height, uint32_t channels,
                        int quality, int subsampling, int16_t* d_coefs[3]) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !d_pixels || !d_coefs) return RCG_INVALID;
        RCG_TRY(check_image(width, height, channels));
        const Layout L = make_layout(width, height, channels, quality, subsampling);
        for (uint32_t c = 0; c < L.ncomp; c++)
            if (!d_coefs[c] || (reinterpret_cast<uintptr_t>(d_coefs[c]) & 15)) {
                set_error("coefficient buffers must be 16-byte aligned device pointers");
                return RCG_INVALID;
            }
        Call call(ctx);
        RCG_TRY(call.begin(0));
        int16_t* co[3] = {d_coefs[0], L.ncomp == 3 ? d_coefs[1] : nullptr, L.ncomp == 3 ? d_coefs[2] : nullptr};
        RCG_TRY(stage_and_launch_encode(ctx, call.D, L, d_pixels, quality, co, nullptr));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        return RCG_OK;
    });
}

int rcg_build_coefs(rcg_ctx* ctx, const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                    int quality, int subsampling, int16_t* coefs[3], uint8_t qtables[128]) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !pixels || !coefs) return RCG_INVALID;
        RCG_TRY(check_image(width, height, channels));
        const Layout L = make_layout(width, height, channels, quality, subsampling);
        for (uint32_t c = 0; c < L.ncomp; c++)
            if (!coefs[c]) return RCG_INVALID;
        const size_t src_bytes = size_t(width) * height * channels;
        Call call(ctx);
        RCG_TRY(call.begin(Arena::align(src_bytes) + coef_bytes(L)));
        uint8_t* d_src = static_cast<uint8_t*>(ctx->arena.take(src_bytes));
        int16_t* d_co[3] = {nullptr, nullptr, nullptr};
        for (uint32_t c = 0; c < L.ncomp; c++)
            d_co[c] = static_cast<int16_t*>(ctx->arena.take(L.coef_count(c) * sizeof(int16_t)));
        RCG_HIP(hipMemcpyAsync(d_src, pixels, src_bytes, hipMemcpyHostToDevice, ctx->stream));
        RCG_TRY(stage_and_launch_encode(ctx, call.D, L, d_src, quality, d_co, nullptr));
        for (uint32_t c = 0; c < L.ncomp; c++)
            RCG_HIP(hipMemcpyAsync(coefs[c], d_co[c], L.coef_count(c) * sizeof(int16_t), hipMemcpyDeviceToHost,
                                   ctx->stream));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        if (qtables) quant_tables(quality, qtables, qtables + 64);
        return RCG_OK;
    });
}

int rcg_build_scaled(rcg_ctx* ctx, const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                     int quality, int subsampling, int16_t* coefs[3], float* scaled[3], uint8_t qtables[128]) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !pixels || !coefs || !scaled) return RCG_INVALID;
        RCG_TRY(check_image(width, height, channels));
        const Layout L = make_layout(width, height, channels, quality, subsampling);
        for (uint32_t c = 0; c < L.ncomp; c++)
            if (!coefs[c] || !scaled[c]) return RCG_INVALID;
        const size_t src_bytes = size_t(width) * height * channels;
        size_t sc_bytes = 0;
        for (uint32_t c = 0; c < L.ncomp; c++) sc_bytes += Arena::align(L.coef_count(c) * sizeof(float));
        Call call(ctx);
        RCG_TRY(call.begin(Arena::align(src_bytes) + coef_bytes(L) + sc_bytes));
        uint8_t* d_src = static_cast<uint8_t*>(ctx->arena.take(src_bytes));
        int16_t* d_co[3] = {nullptr, nullptr, nullptr};
        float* d_sc[3] = {nullptr, nullptr, nullptr};
        for (uint32_t c = 0; c < L.ncomp; c++) {
            d_co[c] = static_cast<int16_t*>(ctx->arena.take(L.coef_count(c) * sizeof(int16_t)));
            d_sc[c] = static_cast<float*>(ctx->arena.take(L.coef_count(c) * sizeof(float)));
        }
        RCG_HIP(hipMemcpyAsync(d_src, pixels, src_bytes, hipMemcpyHostToDevice, ctx->stream));
        RCG_TRY(stage_and_launch_encode(ctx, call.D, L, d_src, quality, d_co, nullptr, d_sc));
        for (uint32_t c = 0; c < L.ncomp; c++) {
            RCG_HIP(hipMemcpyAsync(coefs[c], d_co[c], L.coef_count(c) * sizeof(int16_t), hipMemcpyDeviceToHost,
                                   ctx->stream));
            RCG_HIP(hipMemcpyAsync(scaled[c], d_sc[c], L.coef_count(c) * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx->stream));
        }
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        if (qtables) quant_tables(quality, qtables, qtables + 64);
        return RCG_OK;
    });
}

}  // extern "C"

// Validate a coefficient image and derive its sample planes (jpeg/src/decode.rs:1314,1404-1415). rgb_direct: three
// components stored as R, G, B (Adobe transform 0, or no Adobe marker and ids "RGB"; decode.rs:1423-1426).
int rcg::check_coef_image(const rcg_coef_image* ci, uint32_t& hmax, uint32_t& vmax, int& rgb_direct) {
    if (!ci || ci->width == 0 || ci->height == 0) return RCG_INVALID;
    if (ci->ncomp != 1 && ci->ncomp != 3 && ci->ncomp != 4) {
        set_error("only 1-, 3- and 4-component images are reconstructed");
        return RCG_UNSUPPORTED;
    }
    hmax = 1;
    vmax = 1;
    for (uint32_t c = 0; c < ci->ncomp; c++) {
        const rcg_coef_component& k = ci->comps[c];
        if (k.h == 0 || k.v == 0 || k.tq > 3 || !k.coefs || k.blocks_w == 0 || k.blocks_h == 0) return RCG_INVALID;
        hmax = std::max<uint32_t>(hmax, k.h);
        vmax = std::max<uint32_t>(vmax, k.v);
    }
    for (uint32_t c = 0; c < ci->ncomp; c++) {
        const rcg_coef_component& k = ci->comps[c];
        if (!ci->qtables[k.tq]) {
            set_error("missing quantization table");
            return RCG_INVALID;
        }
        if (hmax % k.h || vmax % k.v) {
            set_error("non-integral chroma subsampling ratio");
            return RCG_UNSUPPORTED;
        }
        const uint32_t cw = (ci->width * k.h + hmax - 1) / hmax, ch = (ci->height * k.v + vmax - 1) / vmax;
        if (uint64_t(k.blocks_w) * 8 < cw || uint64_t(k.blocks_h) * 8 < ch) {
            set_error("coefficient grid smaller than the image");
            return RCG_INVALID;
        }
    }
    rgb_direct = ci->ncomp == 3 &&
                 (ci->adobe_transform == 0 || (ci->adobe_transform < 0 && ci->comps[0].id == 'R' &&
                                               ci->comps[1].id == 'G' && ci->comps[2].id == 'B'));
    return RCG_OK;
}

// Bytes per pixel of reconstruct's output: Gray8 for one component, RGB8 otherwise (decode.rs:1416-1451).
uint32_t rcg::out_channels(const rcg_coef_image* ci) { return ci->ncomp == 1 ? 1 : 3; }

void rcg::mult_table(const uint8_t q[64], float mult[64]) {
    for (int v = 0; v < 8; v++)
        for (int u = 0; u < 8; u++) mult[v * 8 + u] = float(q[v * 8 + u]) * kAan[u] * kAan[v];
}

// IDCT + upsample + colour of a coefficient image whose coefficient arrays are already on the device: the fused
// kernel for the usual layouts (when fused_ok), else the IDCT into `planes_out` (one per component, blocks_w*8 x
// blocks_h*8 bytes each) and the colour stage.
int rcg::reconstruct_on_device(rcg_ctx* ctx, const rcg_coef_image* ci, const int16_t* const d_coefs[4],
                          uint8_t* const planes_out[4], uint8_t* d_out, uint32_t hmax, uint32_t vmax, int rgb_direct,
                          bool fused_ok) {
    int fx, fy;
    if (fused_ok && fused_reconstruct_layout(ci, hmax, vmax, &fx, &fy))
        return launch_reconstruct_fused(ctx, ci, d_coefs, d_out, hmax, vmax, rgb_direct);
    IdctPlane planes[4] = {};
    PlaneView views[4];
    for (uint32_t c = 0; c < ci->ncomp; c++) {
        const rcg_coef_component& k = ci->comps[c];
        planes[c].coefs = d_coefs[c];
        planes[c].blocks_w = k.blocks_w;
        planes[c].blocks_h = k.blocks_h;
        planes[c].plane = planes_out[c];
        mult_table(ci->qtables[k.tq], planes[c].mult);
        views[c].p = planes[c].plane;
        views[c].stride = k.blocks_w * 8;
        views[c].cw = (ci->width * k.h + hmax - 1) / hmax;
        views[c].ch = (ci->height * k.v + vmax - 1) / vmax;
        views[c].fx = hmax / k.h;
        views[c].fy = vmax / k.v;
    }
    RCG_TRY(launch_idct(ctx, planes, int(ci->ncomp)));
    return launch_upsample_color(ctx, views, int(ci->ncomp), rgb_direct, ci->width, ci->height, d_out,
                                 ci->adobe_transform < 0 ? -1 : ci->adobe_transform);
}

// Sample-plane bytes of a coefficient image (what reconstruct_on_device writes before the colour stage).
size_t rcg::sample_plane_bytes(const rcg_coef_image* ci) {
    size_t b = 0;
    for (uint32_t c = 0; c < ci->ncomp; c++)
        b += Arena::align(size_t(ci->comps[c].blocks_w) * ci->comps[c].blocks_h * 64);
    return b;
}

extern "C" {

int rcg_reconstruct(rcg_ctx* ctx, const rcg_coef_image* ci, uint8_t* pixels) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !pixels) return RCG_INVALID;
        uint32_t hmax, vmax;
        int rgb_direct;
        RCG_TRY(check_coef_image(ci, hmax, vmax, rgb_direct));
        const size_t out_bytes = size_t(ci->width) * ci->height * out_channels(ci);
        int fx, fy;
        const bool fused = fused_reconstruct_layout(ci, hmax, vmax, &fx, &fy);  // no sample planes then
        size_t need = Arena::align(out_bytes) + (fused ? 0 : sample_plane_bytes(ci));
        for (uint32_t c = 0; c < ci->ncomp; c++)
            need += Arena::align(size_t(ci->comps[c].blocks_w) * ci->comps[c].blocks_h * 64 * sizeof(int16_t));
        Call call(ctx);
        RCG_TRY(call.begin(need));
        const int16_t* d_co[4] = {nullptr, nullptr, nullptr, nullptr};
        uint8_t* planes[4] = {nullptr, nullptr, nullptr, nullptr};
        for (uint32_t c = 0; c < ci->ncomp; c++) {
            const size_t nb = size_t(ci->comps[c].blocks_w) * ci->comps[c].blocks_h;
            int16_t* d = static_cast<int16_t*>(ctx->arena.take(nb * 64 * sizeof(int16_t)));
            RCG_HIP(hipMemcpyAsync(d, ci->comps[c].coefs, nb * 64 * sizeof(int16_t), hipMemcpyHostToDevice, ctx->stream));
            d_co[c] = d;
            if (!fused) planes[c] = static_cast<uint8_t*>(ctx->arena.take(nb * 64));
        }
        uint8_t* d_out = static_cast<uint8_t*>(ctx->arena.take(out_bytes));
        RCG_TRY(reconstruct_on_device(ctx, ci, d_co, planes, d_out, hmax, vmax, rgb_direct));
        RCG_HIP(hipMemcpyAsync(pixels, d_out, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        return RCG_OK;
    });
}

int rcg_reconstruct_dev(rcg_ctx* ctx, const rcg_coef_image* d_ci, uint8_t* d_pixels) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !d_pixels) return RCG_INVALID;
        uint32_t hmax, vmax;
        int rgb_direct;
        RCG_TRY(check_coef_image(d_ci, hmax, vmax, rgb_direct));
        const int16_t* d_co[4] = {nullptr, nullptr, nullptr, nullptr};
        for (uint32_t c = 0; c < d_ci->ncomp; c++) {
            if (reinterpret_cast<uintptr_t>(d_ci->comps[c].coefs) & 15) {
                set_error("coefficient buffers must be 16-byte aligned device pointers");
                return RCG_INVALID;
            }
            d_co[c] = d_ci->comps[c].coefs;
        }
        int fx, fy;
        const bool fused = fused_reconstruct_layout(d_ci, hmax, vmax, &fx, &fy);  // no sample planes then
        Call call(ctx);
        RCG_TRY(call.begin(fused ? 0 : sample_plane_bytes(d_ci)));
        uint8_t* planes[4] = {nullptr, nullptr, nullptr, nullptr};
        for (uint32_t c = 0; !fused && c < d_ci->ncomp; c++)
            planes[c] = static_cast<uint8_t*>(ctx->arena.take(size_t(d_ci->comps[c].blocks_w) * d_ci->comps[c].blocks_h * 64));
        RCG_TRY(reconstruct_on_device(ctx, d_ci, d_co, planes, d_pixels, hmax, vmax, rgb_direct));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        return RCG_OK;
    });
}

int rcg_ssimulacra2_dev(rcg_ctx* ctx, const uint8_t* d_reference, const uint8_t* d_distorted, uint32_t width,
                        uint32_t height, uint32_t channels, int precision, double* score, double* per_scale,
                        int* nscales) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !d_reference || !d_distorted || !score) return RCG_INVALID;
        if (!valid_channels(channels)) return RCG_UNSUPPORTED;
        if (width < 8 || height < 8) {
            set_error("SSIMULACRA 2 needs at least 8x8 pixels");
            return RCG_INVALID;
        }
        const SsimTune& T = precision == RCG_PREC_FP64 ? ctx->tune64 : ctx->tune32;
        const PairPlan P = plan_pair(width, height, T);
        Call call(ctx);
        RCG_TRY(call.begin(ssim_bytes(P, precision)));
        return ssim_pair_dev(ctx, call, d_reference, d_distorted, width, height, channels, precision, score, per_scale,
                             nscales);
    });
}

int rcg_ssimulacra2(rcg_ctx* ctx, const uint8_t* reference, const uint8_t* distorted, uint32_t width,
                    uint32_t height, uint32_t channels, int precision, double* score, double* per_scale,
                    int* nscales) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !reference || !distorted || !score) return RCG_INVALID;
        if (!valid_channels(channels)) return RCG_UNSUPPORTED;
        if (width < 8 || height < 8) {
            set_error("SSIMULACRA 2 needs at least 8x8 pixels");
            return RCG_INVALID;
        }
        const SsimTune& T = precision == RCG_PREC_FP64 ? ctx->tune64 : ctx->tune32;
        const PairPlan P = plan_pair(width, height, T);
        const size_t n = size_t(width) * height * channels;
        Call call(ctx);
        RCG_TRY(call.begin(ssim_bytes(P, precision) + 2 * Arena::align(n)));
        uint8_t* d_a = static_cast<uint8_t*>(ctx->arena.take(n));
        uint8_t* d_b = static_cast<uint8_t*>(ctx->arena.take(n));
        RCG_HIP(hipMemcpyAsync(d_a, reference, n, hipMemcpyHostToDevice, ctx->stream));
        RCG_HIP(hipMemcpyAsync(d_b, distorted, n, hipMemcpyHostToDevice, ctx->stream));
        return ssim_pair_dev(ctx, call, d_a, d_b, width, height, channels, precision, score, per_scale, nscales);
    });
}

}  // extern "C"

namespace {

constexpr double kGuard = 0.05;  // |f32 score - target| below this is settled with the f64 scorer

// Device bytes of one target search over a width x height image (sources, candidate planes, coefficients, hand-off
// values with trellis, and the FP32 and FP64 SSIMULACRA2 state), not counting the descriptor staging.
size_t search_need(rcg_ctx* ctx, uint32_t width, uint32_t height, uint32_t channels, int subsampling, bool trellis) {
    const Layout Lmax = max_layout(width, height, channels, subsampling);
    const PairPlan P32 = plan_pair(width, height, ctx->tune32);
    const PairPlan P64 = plan_pair(width, height, ctx->tune64);
    size_t need = Arena::align(size_t(width) * height * channels) + plane_bytes(Lmax) + coef_bytes(Lmax) +
                  ssim_bytes(P32, RCG_PREC_FP32) + ssim_bytes(P64, RCG_PREC_FP64);
    if (trellis)
        for (uint32_t c = 0; c < Lmax.ncomp; c++) need += Arena::align(Lmax.coef_count(c) * sizeof(float));
    return need;
}

// Score one candidate quality on the device: encode + reconstruct planes, SSIMULACRA2 against the source.
// The reference state of each precision is built by its first evaluation and reused by the later ones.
int score_candidate(rcg_ctx* ctx, Call& call, const Layout& L, const uint8_t* d_src, int quality,
                    uint8_t* const rec[3], const SsimBuffers& B32, const SsimBuffers& B64, bool& a32_ready,
                    bool& a64_ready, double target, double& score, bool& settled) {
    call.D.off = 0;
    const StagedEncode se = stage_encode(call.D, L, d_src, quality, nullptr, rec);
    const PairPlan P = plan_pair(L.w, L.h, ctx->tune32);
    StagedSsim st = stage_ssim(call.D, P, bytes_src(d_src, L.channels), recon_src(L, rec), B32, a32_ready,
                               RCG_PREC_FP32);
    RCG_TRY(call.D.upload(ctx->stream));
    // The first evaluation builds the reference state on the side stream, beside the encode.
    if (st.build_ref) RCG_TRY(fork_side(ctx));
    RCG_TRY(launch_encode(ctx, se.job, se.begins, 1, se.tiles));
    RCG_TRY(launch_staged_refs(ctx, st));
    RCG_TRY(launch_staged_ssim(ctx, st));
    double s = 0.0;
    RCG_HIP(hipMemcpyAsync(&s, B32.d_score, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    a32_ready = true;
    settled = false;
    if (std::isfinite(s) && std::fabs(s - target) < kGuard) {
        call.D.off = 0;
        const PairPlan P64 = plan_pair(L.w, L.h, ctx->tune64);
        StagedSsim s64 = stage_ssim(call.D, P64, bytes_src(d_src, L.channels), recon_src(L, rec), B64, a64_ready,
                                    RCG_PREC_FP64);
        RCG_TRY(call.D.upload(ctx->stream));
        RCG_TRY(launch_staged_ssim(ctx, s64));
        RCG_HIP(hipMemcpyAsync(&s, B64.d_score, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        a64_ready = true;
        settled = true;
    }
    // target.rs:36 / pipeline.rs:218: a candidate that cannot be scored counts as f64::MIN.
    score = std::isfinite(s) ? s : -DBL_MAX;
    return RCG_OK;
}

int find_quality_impl(rcg_ctx* ctx, const uint8_t* pixels, bool host_src, uint32_t width, uint32_t height,
                      uint32_t channels, double target, int max_quality, int subsampling, rcg_found* out,
                      int16_t* coefs[3], uint8_t qtables[128]) {
    if (!ctx || !pixels || !out) return RCG_INVALID;
    RCG_TRY(check_image(width, height, channels));
    if (width < 8 || height < 8) {
        set_error("--target-ssimulacra2 needs an image of at least 8x8 pixels");
        return RCG_UNSUPPORTED;
    }
    // Auto subsampling depends on the quality: size for every layout the search can meet.
    const Layout Lmax = max_layout(width, height, channels, subsampling);
    const size_t src_bytes = size_t(width) * height * channels;
    const PairPlan P32 = plan_pair(width, height, ctx->tune32);
    const PairPlan P64 = plan_pair(width, height, ctx->tune64);
    const size_t need = search_need(ctx, width, height, channels, subsampling, false);
    Call call(ctx);
    RCG_TRY(call.begin(need));
    const uint8_t* d_src = pixels;
    if (host_src) {
        uint8_t* d = static_cast<uint8_t*>(ctx->arena.take(src_bytes));
        RCG_HIP(hipMemcpyAsync(d, pixels, src_bytes, hipMemcpyHostToDevice, ctx->stream));
        d_src = d;
    }
    uint8_t* rec[3] = {nullptr, nullptr, nullptr};
    for (uint32_t c = 0; c < Lmax.ncomp; c++) rec[c] = static_cast<uint8_t*>(ctx->arena.take(Lmax.plane_bytes(c)));
    int16_t* d_co[3] = {nullptr, nullptr, nullptr};
    for (uint32_t c = 0; c < Lmax.ncomp; c++)
        d_co[c] = static_cast<int16_t*>(ctx->arena.take(Lmax.coef_count(c) * sizeof(int16_t)));
    const SsimBuffers B32 = take_ssim(ctx->arena, P32, RCG_PREC_FP32);
    const SsimBuffers B64 = take_ssim(ctx->arena, P64, RCG_PREC_FP64);
    bool a32 = false, a64 = false;

    // jpeg/src/target.rs:37-55
    int lo = 1, hi = std::max(1, max_quality);
    int steps = 0;
    double s = 0.0;
    bool settled = false;
    Layout L = make_layout(width, height, channels, hi, subsampling);
    RCG_TRY(score_candidate(ctx, call, L, d_src, hi, rec, B32, B64, a32, a64, target, s, settled));
    steps++;
    int best_q = hi;
    double best_s = s;
    if (s >= target) {
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            L = make_layout(width, height, channels, mid, subsampling);
            RCG_TRY(score_candidate(ctx, call, L, d_src, mid, rec, B32, B64, a32, a64, target, s, settled));
            steps++;
            if (s >= target) {
                hi = mid;
                best_q = mid;
                best_s = s;
            } else {
                lo = mid + 1;
            }
        }
    }
    out->quality = best_q;
    out->score = best_s;
    out->steps = steps;
    if (coefs) {
        L = make_layout(width, height, channels, best_q, subsampling);
        call.D.off = 0;
        int16_t* co[3] = {d_co[0], L.ncomp == 3 ? d_co[1] : nullptr, L.ncomp == 3 ? d_co[2] : nullptr};
        RCG_TRY(stage_and_launch_encode(ctx, call.D, L, d_src, best_q, co, nullptr));
        for (uint32_t c = 0; c < L.ncomp; c++)
            if (coefs[c])
                RCG_HIP(hipMemcpyAsync(coefs[c], d_co[c], L.coef_count(c) * sizeof(int16_t), hipMemcpyDeviceToHost,
                                       ctx->stream));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (qtables) quant_tables(best_q, qtables, qtables + 64);
    return RCG_OK;
}

}  // namespace

extern "C" {

int rcg_find_quality(rcg_ctx* ctx, const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                     double target, int max_quality, int subsampling, rcg_found* out, int16_t* coefs[3],
                     uint8_t qtables[128]) {
    return rcg::guarded([&]() -> int {
        return find_quality_impl(ctx, pixels, true, width, height, channels, target, max_quality, subsampling, out,
                                 coefs, qtables);
    });
}

int rcg_find_quality_dev(rcg_ctx* ctx, const uint8_t* d_pixels, uint32_t width, uint32_t height,
                         uint32_t channels, double target, int max_quality, int subsampling, rcg_found* out) {
    return rcg::guarded([&]() -> int {
        return find_quality_impl(ctx, d_pixels, false, width, height, channels, target, max_quality, subsampling, out,
                                 nullptr, nullptr);
    });
}

}  // extern "C"

// ================================================================================================================
// The search with trellis quantisation (the CLI default, encode.rs:527-537): make = build_coefs(trellis on)

namespace {

// Host side of one trellis candidate: the plain coefficients and hand-off values read back, the trellis output.
struct TrellisHost {
    std::vector<int16_t> plain[3], tr[3], best[3];
    std::vector<float> scaled[3];
};

// FP64 re-score of a candidate whose FP32 score fell inside the guard band around the target.
int rescore_fp64(rcg_ctx* ctx, Call& call, const Layout& L, const uint8_t* d_ref, uint8_t* const rec[3],
                 const SsimBuffers& B64, bool& a64_ready, double& s) {
    call.D.off = 0;
    const PairPlan P64 = plan_pair(L.w, L.h, ctx->tune64);
    StagedSsim s64 = stage_ssim(call.D, P64, bytes_src(d_ref, L.channels), recon_src(L, rec), B64, a64_ready,
                                RCG_PREC_FP64);
    RCG_TRY(call.D.upload(ctx->stream));
    RCG_TRY(launch_staged_ssim(ctx, s64));
    RCG_HIP(hipMemcpyAsync(&s, B64.d_score, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    a64_ready = true;
    return RCG_OK;
}

// One candidate: encode on the device with the hand-off values, read back; the host trellis; upload; IDCT and
// SSIMULACRA2 on the device. On the first candidate the reference state is built on the side stream while the
// encode, the read-back and the host trellis run. At quality 100 there is no trellis pass (encode.rs:527): the
// encode reconstructs its planes itself and only the coefficients are read back.
int score_trellis(rcg_ctx* ctx, Call& call, const Layout& L, const uint8_t* d_src, int quality,
                  int16_t* const d_co[3], float* const d_sc[3], uint8_t* const rec[3], TrellisHost& H,
                  const SsimBuffers& B32, const SsimBuffers& B64, bool& a32_ready, bool& a64_ready, double target,
                  double& score) {
    const bool use_trellis = quality < 100;
    call.D.off = 0;
    int16_t* co[3] = {d_co[0], L.ncomp == 3 ? d_co[1] : nullptr, L.ncomp == 3 ? d_co[2] : nullptr};
    float* sc[3] = {d_sc[0], L.ncomp == 3 ? d_sc[1] : nullptr, L.ncomp == 3 ? d_sc[2] : nullptr};
    uint8_t* rc3[3] = {rec[0], L.ncomp == 3 ? rec[1] : nullptr, L.ncomp == 3 ? rec[2] : nullptr};
    const StagedEncode se = stage_encode(call.D, L, d_src, quality, co, use_trellis ? nullptr : rc3,
                                         use_trellis ? sc : nullptr);
    const PairPlan P32 = plan_pair(L.w, L.h, ctx->tune32);
    StagedSsim st = stage_ssim(call.D, P32, bytes_src(d_src, L.channels), recon_src(L, rec), B32, a32_ready,
                               RCG_PREC_FP32);
    RCG_TRY(call.D.upload(ctx->stream));
    if (st.build_ref) RCG_TRY(fork_side(ctx));
    RCG_TRY(launch_encode(ctx, se.job, se.begins, 1, se.tiles));
    RCG_TRY(launch_staged_refs(ctx, st));
    for (uint32_t c = 0; c < L.ncomp; c++) H.tr[c].resize(L.coef_count(c));
    if (use_trellis) {
        for (uint32_t c = 0; c < L.ncomp; c++) {
            H.plain[c].resize(L.coef_count(c));
            H.scaled[c].resize(L.coef_count(c));
            RCG_HIP(hipMemcpyAsync(H.plain[c].data(), d_co[c], L.coef_count(c) * sizeof(int16_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
            RCG_HIP(hipMemcpyAsync(H.scaled[c].data(), d_sc[c], L.coef_count(c) * sizeof(float),
                                   hipMemcpyDeviceToHost, ctx->stream));
        }
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        uint8_t qt[2][64];
        quant_tables(quality, qt[0], qt[1]);
        rcg_coef_image ci;
        std::memset(&ci, 0, sizeof(ci));
        ci.width = L.w;
        ci.height = L.h;
        ci.ncomp = L.ncomp;
        ci.adobe_transform = -1;
        ci.qtables[0] = qt[0];
        ci.qtables[1] = L.ncomp == 3 ? qt[1] : nullptr;
        const float* sp[4] = {nullptr, nullptr, nullptr, nullptr};
        int16_t* op[4] = {nullptr, nullptr, nullptr, nullptr};
        for (uint32_t c = 0; c < L.ncomp; c++) {
            rcg_coef_component& k = ci.comps[c];
            k.id = uint8_t(c + 1);
            k.h = uint8_t(c == 0 ? L.hmax : 1);
            k.v = uint8_t(c == 0 ? L.vmax : 1);
            k.tq = uint8_t(c == 0 ? 0 : 1);
            k.blocks_w = L.bw[c];
            k.blocks_h = L.bh[c];
            k.coefs = H.plain[c].data();
            sp[c] = H.scaled[c].data();
            op[c] = H.tr[c].data();
        }
        RCG_TRY(rcg_trellis(&ci, sp, op));
        IdctPlane ip[3] = {};
        for (uint32_t c = 0; c < L.ncomp; c++) {
            // The plain coefficients on the device are no longer needed: the trellis output takes their place.
            RCG_HIP(hipMemcpyAsync(d_co[c], H.tr[c].data(), L.coef_count(c) * sizeof(int16_t), hipMemcpyHostToDevice,
                                   ctx->stream));
            ip[c].coefs = d_co[c];
            ip[c].blocks_w = L.bw[c];
            ip[c].blocks_h = L.bh[c];
            ip[c].plane = rec[c];
            mult_table(qt[c == 0 ? 0 : 1], ip[c].mult);
        }
        RCG_TRY(launch_idct(ctx, ip, int(L.ncomp)));
    } else {
        for (uint32_t c = 0; c < L.ncomp; c++)
            RCG_HIP(hipMemcpyAsync(H.tr[c].data(), d_co[c], L.coef_count(c) * sizeof(int16_t), hipMemcpyDeviceToHost,
                                   ctx->stream));
    }
    RCG_TRY(launch_staged_ssim(ctx, st));
    double s = 0.0;
    RCG_HIP(hipMemcpyAsync(&s, B32.d_score, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    a32_ready = true;
    if (std::isfinite(s) && std::fabs(s - target) < kGuard)
        RCG_TRY(rescore_fp64(ctx, call, L, d_src, rec, B64, a64_ready, s));
    score = std::isfinite(s) ? s : -DBL_MAX;
    return RCG_OK;
}

}  // namespace

extern "C" int rcg_find_quality_trellis(rcg_ctx* ctx, const uint8_t* pixels, uint32_t width, uint32_t height,
                                        uint32_t channels, double target, int max_quality, int subsampling,
                                        rcg_found* out, int16_t* coefs[3], uint8_t qtables[128]) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !pixels || !out) return RCG_INVALID;
        RCG_TRY(check_image(width, height, channels));
        if (width < 8 || height < 8) {
            set_error("--target-ssimulacra2 needs an image of at least 8x8 pixels");
            return RCG_UNSUPPORTED;
        }
        const Layout Lmax = max_layout(width, height, channels, subsampling);
        const PairPlan P32 = plan_pair(width, height, ctx->tune32);
        const PairPlan P64 = plan_pair(width, height, ctx->tune64);
        const size_t src_bytes = size_t(width) * height * channels;
        const size_t need = search_need(ctx, width, height, channels, subsampling, true);
        try {
            Call call(ctx);
            RCG_TRY(call.begin(need));
            uint8_t* d_src = static_cast<uint8_t*>(ctx->arena.take(src_bytes));
            RCG_HIP(hipMemcpyAsync(d_src, pixels, src_bytes, hipMemcpyHostToDevice, ctx->stream));
            uint8_t* rec[3] = {nullptr, nullptr, nullptr};
            int16_t* d_co[3] = {nullptr, nullptr, nullptr};
            float* d_sc[3] = {nullptr, nullptr, nullptr};
            for (uint32_t c = 0; c < Lmax.ncomp; c++) {
                rec[c] = static_cast<uint8_t*>(ctx->arena.take(Lmax.plane_bytes(c)));
                d_co[c] = static_cast<int16_t*>(ctx->arena.take(Lmax.coef_count(c) * sizeof(int16_t)));
                d_sc[c] = static_cast<float*>(ctx->arena.take(Lmax.coef_count(c) * sizeof(float)));
            }
            const SsimBuffers B32 = take_ssim(ctx->arena, P32, RCG_PREC_FP32);
            const SsimBuffers B64 = take_ssim(ctx->arena, P64, RCG_PREC_FP64);
            bool a32 = false, a64 = false;
            TrellisHost H;
            auto keep = [&](const Layout& Lq) {
                for (uint32_t c = 0; c < Lq.ncomp; c++) H.best[c] = H.tr[c];
            };
            // jpeg/src/target.rs:37-55
            int lo = 1, hi = std::max(1, std::min(100, max_quality));
            int steps = 0;
            double s = 0.0;
            Layout L = make_layout(width, height, channels, hi, subsampling);
            RCG_TRY(score_trellis(ctx, call, L, d_src, hi, d_co, d_sc, rec, H, B32, B64, a32, a64, target, s));
            steps++;
            keep(L);
            int best_q = hi;
            double best_s = s;
            if (s >= target) {
                while (lo < hi) {
                    const int mid = lo + (hi - lo) / 2;
                    L = make_layout(width, height, channels, mid, subsampling);
                    RCG_TRY(score_trellis(ctx, call, L, d_src, mid, d_co, d_sc, rec, H, B32, B64, a32, a64, target, s));
                    steps++;
                    if (s >= target) {
                        hi = mid;
                        best_q = mid;
                        best_s = s;
                        keep(L);
                    } else {
                        lo = mid + 1;
                    }
                }
            }
            out->quality = best_q;
            out->score = best_s;
            out->steps = steps;
            if (coefs) {
                const Layout Lb = make_layout(width, height, channels, best_q, subsampling);
                for (uint32_t c = 0; c < Lb.ncomp; c++)
                    if (coefs[c]) std::memcpy(coefs[c], H.best[c].data(), Lb.coef_count(c) * sizeof(int16_t));
            }
            if (qtables) quant_tables(best_q, qtables, qtables + 64);
        } catch (const std::bad_alloc&) {
            set_error("out of host memory");
            return RCG_OOM;
        }
        return RCG_OK;
    });
}

// ================================================================================================================
// JPEG -> JPEG: the DCT-domain requantisation search (cli/src/pipeline.rs:369-386)

namespace {

// Device bytes of the requantisation search over the block grid L: per component the source and candidate
// coefficients and sample planes, the decoded reference image, and the FP32 and FP64 SSIMULACRA2 state.
// Device bytes of one requantisation search: the decoded source, the FP32 and FP64 scoring state, per component the
// source and trellis coefficients and two sets of sample planes, and the RGB8 candidate of the pixel route (counted
// always: an upper bound).
size_t requant_need(rcg_ctx* ctx, const Layout& L) {
    const PairPlan P32 = plan_pair(L.w, L.h, ctx->tune32);
    const PairPlan P64 = plan_pair(L.w, L.h, ctx->tune64);
    size_t need = 2 * Arena::align(size_t(L.w) * L.h * L.channels) + ssim_bytes(P32, RCG_PREC_FP32) +
                  ssim_bytes(P64, RCG_PREC_FP64);
    for (uint32_t c = 0; c < L.ncomp; c++) {
        const size_t nb = size_t(L.bw[c]) * L.bh[c];
        need += 2 * Arena::align(nb * 64 * sizeof(int16_t)) + 2 * Arena::align(nb * 64);
    }
    return need;
}

// Where a requantisation search reads its candidates from.
//   fused: the SSIMULACRA2 loader upsamples and colour-converts the IDCT's sample planes as it reads them (luma at
//          full resolution, chroma planes of one geometry, ratios 1 or 2, YCbCr or gray);
//   pixel: any other layout reconstruct accepts (four components, RGB-coded components, other sampling ratios):
//          the IDCT's planes go through the reconstruct colour stage to an RGB8 candidate, which is scored as pixels.
struct RequantRoute {
    bool fused;
    uint8_t* d_cand;  // pixel route: w*h*3 bytes
    PlaneView views[4];
    int rgb_direct;
    int adobe_transform;
};

// One requantised candidate scored against the decoded source: its sample planes (requantised on the device inside
// the IDCT, or uploaded after the host trellis), then SSIMULACRA2 with the FP64 re-score inside the guard band.
int score_requant(rcg_ctx* ctx, Call& call, const rcg_coef_image* ci, const Layout& L, const int16_t* const d_src[4],
                  const uint8_t* d_ref, uint32_t ref_ch, uint8_t* const planes[4], int16_t* const d_tr[4],
                  const std::vector<std::vector<int16_t>>& host_src, int quality, bool trellis,
                  const RequantRoute& route, const SsimBuffers& B32, const SsimBuffers& B64, bool& a32_ready,
                  bool& a64_ready, double target, double& score) {
    RequantPlan P;
    RCG_TRY(requant_plan(ci, quality, P));
    IdctPlane ip[4] = {};
    for (uint32_t c = 0; c < ci->ncomp; c++) {
        const rcg_coef_component& k = ci->comps[c];
        ip[c].blocks_w = k.blocks_w;
        ip[c].blocks_h = k.blocks_h;
        ip[c].plane = planes[c];
        mult_table(P.table[c], ip[c].mult);
    }
    const bool host_trellis = trellis && P.changed;
    if (host_trellis) {
        // The trellis runs on the host from the source coefficients (encode.rs:146-169); its output is uploaded.
        std::vector<std::vector<int16_t>> rq(ci->ncomp);
        rcg_requantized out;
        std::memset(&out, 0, sizeof(out));
        for (uint32_t c = 0; c < ci->ncomp; c++) {
            rq[c].resize(host_src[c].size());
            out.coefs[c] = rq[c].data();
        }
        RCG_TRY(rcg_requantize(ci, quality, 1, &out));
        for (uint32_t c = 0; c < ci->ncomp; c++) {
            RCG_HIP(hipMemcpyAsync(d_tr[c], rq[c].data(), rq[c].size() * sizeof(int16_t), hipMemcpyHostToDevice,
                                   ctx->stream));
            ip[c].coefs = d_tr[c];
        }
        RCG_HIP(hipStreamSynchronize(ctx->stream));  // rq leaves scope
    } else {
        for (uint32_t c = 0; c < ci->ncomp; c++) {
            ip[c].coefs = d_src[c];
            ip[c].requant = 1;
            std::memcpy(ip[c].ratio, P.ratio[c], sizeof(ip[c].ratio));
        }
    }
    uint8_t* rec[3] = {planes[0], planes[1], planes[2]};
    const SsimSrc cand = route.fused ? recon_src(L, rec) : bytes_src(route.d_cand, 3);
    const PairPlan P32 = plan_pair(L.w, L.h, ctx->tune32);
    call.D.off = 0;
    StagedSsim st = stage_ssim(call.D, P32, bytes_src(d_ref, ref_ch), cand, B32, a32_ready, RCG_PREC_FP32);
    RCG_TRY(call.D.upload(ctx->stream));
    if (st.build_ref) RCG_TRY(fork_side(ctx));  // the reference state beside the IDCT, on the first step
    RCG_TRY(launch_idct(ctx, ip, int(ci->ncomp)));
    if (!route.fused)
        RCG_TRY(launch_upsample_color(ctx, route.views, int(ci->ncomp), route.rgb_direct, L.w, L.h, route.d_cand,
                                      route.adobe_transform));
    RCG_TRY(launch_staged_refs(ctx, st));
    RCG_TRY(launch_staged_ssim(ctx, st));
    double s = 0.0;
    RCG_HIP(hipMemcpyAsync(&s, B32.d_score, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    a32_ready = true;
    if (std::isfinite(s) && std::fabs(s - target) < kGuard) {
        call.D.off = 0;
        const PairPlan P64 = plan_pair(L.w, L.h, ctx->tune64);
        StagedSsim s64 = stage_ssim(call.D, P64, bytes_src(d_ref, ref_ch), cand, B64, a64_ready, RCG_PREC_FP64);
        RCG_TRY(call.D.upload(ctx->stream));
        RCG_TRY(launch_staged_ssim(ctx, s64));
        RCG_HIP(hipMemcpyAsync(&s, B64.d_score, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        a64_ready = true;
    }
    score = std::isfinite(s) ? s : -DBL_MAX;
    return RCG_OK;
}

}  // namespace

extern "C" int rcg_requant_find_quality(rcg_ctx* ctx, const rcg_coef_image* ci, double target, int max_quality,
                                        int trellis, rcg_found* out, rcg_requantized* winner) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !ci || !out) return RCG_INVALID;
        uint32_t hmax, vmax;
        int rgb_direct;
        RCG_TRY(check_coef_image(ci, hmax, vmax, rgb_direct));
        if (ci->width < 8 || ci->height < 8) {
            set_error("--target-ssimulacra2 needs an image of at least 8x8 pixels");
            return RCG_UNSUPPORTED;
        }
        const rcg_coef_component* k = ci->comps;
        RequantRoute route;
        std::memset(&route, 0, sizeof(route));
        route.fused = (ci->ncomp == 1 || ci->ncomp == 3) && !rgb_direct && k[0].h == hmax && k[0].v == vmax &&
                      hmax <= 2 && vmax <= 2;
        if (ci->ncomp == 3)
            route.fused = route.fused && k[1].h == 1 && k[1].v == 1 && k[2].h == 1 && k[2].v == 1 &&
                          k[1].blocks_w == k[2].blocks_w && k[1].blocks_h == k[2].blocks_h;
        route.rgb_direct = rgb_direct;
        route.adobe_transform = ci->adobe_transform < 0 ? -1 : ci->adobe_transform;
        Layout L;
        std::memset(&L, 0, sizeof(L));
        L.w = ci->width;
        L.h = ci->height;
        L.channels = out_channels(ci);
        L.hmax = hmax;
        L.vmax = vmax;
        L.ncomp = ci->ncomp;
        for (uint32_t c = 0; c < ci->ncomp; c++) {
            L.bw[c] = ci->comps[c].blocks_w;
            L.bh[c] = ci->comps[c].blocks_h;
        }
        const PairPlan P32 = plan_pair(L.w, L.h, ctx->tune32);
        const PairPlan P64 = plan_pair(L.w, L.h, ctx->tune64);
        const size_t ref_bytes = size_t(L.w) * L.h * L.channels;
        const size_t need = requant_need(ctx, L);
        std::vector<std::vector<int16_t>> host_src(ci->ncomp);
        for (uint32_t c = 0; c < ci->ncomp; c++) {
            const size_t nb = size_t(ci->comps[c].blocks_w) * ci->comps[c].blocks_h;
            if (trellis) host_src[c].assign(ci->comps[c].coefs, ci->comps[c].coefs + nb * 64);
        }
        Call call(ctx);
        RCG_TRY(call.begin(need));
        int16_t* d_src[4] = {nullptr, nullptr, nullptr, nullptr};
        int16_t* d_tr[4] = {nullptr, nullptr, nullptr, nullptr};
        uint8_t* planes[4] = {nullptr, nullptr, nullptr, nullptr};
        uint8_t* src_planes[4] = {nullptr, nullptr, nullptr, nullptr};
        for (uint32_t c = 0; c < ci->ncomp; c++) {
            const size_t nb = size_t(ci->comps[c].blocks_w) * ci->comps[c].blocks_h;
            d_src[c] = static_cast<int16_t*>(ctx->arena.take(nb * 64 * sizeof(int16_t)));
            d_tr[c] = static_cast<int16_t*>(ctx->arena.take(nb * 64 * sizeof(int16_t)));
            planes[c] = static_cast<uint8_t*>(ctx->arena.take(nb * 64));
            src_planes[c] = static_cast<uint8_t*>(ctx->arena.take(nb * 64));
            RCG_HIP(hipMemcpyAsync(d_src[c], ci->comps[c].coefs, nb * 64 * sizeof(int16_t), hipMemcpyHostToDevice,
                                   ctx->stream));
            const rcg_coef_component& kc = ci->comps[c];
            route.views[c].p = planes[c];
            route.views[c].stride = kc.blocks_w * 8;
            route.views[c].cw = (ci->width * kc.h + hmax - 1) / hmax;
            route.views[c].ch = (ci->height * kc.v + vmax - 1) / vmax;
            route.views[c].fx = hmax / kc.h;
            route.views[c].fy = vmax / kc.v;
        }
        uint8_t* d_ref = static_cast<uint8_t*>(ctx->arena.take(ref_bytes));
        if (!route.fused) route.d_cand = static_cast<uint8_t*>(ctx->arena.take(size_t(L.w) * L.h * 3));
        const SsimBuffers B32 = take_ssim(ctx->arena, P32, RCG_PREC_FP32);
        const SsimBuffers B64 = take_ssim(ctx->arena, P64, RCG_PREC_FP64);
        // reference = reconstruct(&ci) (pipeline.rs:373), decoded once on the device.
        RCG_TRY(reconstruct_on_device(ctx, ci, d_src, src_planes, d_ref, hmax, vmax, rgb_direct));
        bool a32 = false, a64 = false;
        // jpeg/src/target.rs:37-55
        int lo = 1, hi = std::max(1, std::min(100, max_quality));
        int steps = 0;
        double s = 0.0;
        RCG_TRY(score_requant(ctx, call, ci, L, d_src, d_ref, L.channels, planes, d_tr, host_src, hi, trellis != 0,
                              route, B32, B64, a32, a64, target, s));
        steps++;
        int best_q = hi;
        double best_s = s;
        if (s >= target) {
            while (lo < hi) {
                const int mid = lo + (hi - lo) / 2;
                RCG_TRY(score_requant(ctx, call, ci, L, d_src, d_ref, L.channels, planes, d_tr, host_src, mid,
                                      trellis != 0, route, B32, B64, a32, a64, target, s));
                steps++;
                if (s >= target) {
                    hi = mid;
                    best_q = mid;
                    best_s = s;
                } else {
                    lo = mid + 1;
                }
            }
        }
        out->quality = best_q;
        out->score = best_s;
        out->steps = steps;
        if (winner) RCG_TRY(rcg_requantize(ci, best_q, trellis, winner));
        return RCG_OK;
    });
}

extern "C" int rcg_search_bytes(rcg_ctx* ctx, uint32_t width, uint32_t height, uint32_t channels, int subsampling,
                                int mode, size_t* out) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !out) return RCG_INVALID;
        if (mode == RCG_SEARCH_REQUANT && channels == 4) {
            // A four-component source: every component's grid is at most the 4:2:0 luma grid of its size.
            RCG_TRY(check_image(width, height, 3));
            Layout L = max_layout(width, height, 3, RCG_AUTO);
            L.ncomp = 4;
            for (int c = 1; c < 4; c++) {
                L.bw[c] = L.bw[0];
                L.bh[c] = L.bh[0];
            }
            *out = requant_need(ctx, L) + kDescBytes + 4096;
            return RCG_OK;
        }
        RCG_TRY(check_image(width, height, channels));
        size_t need = 0;
        switch (mode) {
            case RCG_SEARCH_PLAIN: need = search_need(ctx, width, height, channels, subsampling, false); break;
            case RCG_SEARCH_TRELLIS: need = search_need(ctx, width, height, channels, subsampling, true); break;
            case RCG_SEARCH_REQUANT: {
                // The search takes sources with sampling ratios 1 or 2: their grids are at most the larger of the
                // 4:4:4 and 4:2:0 ones of that size, per component.
                const Layout L = max_layout(width, height, channels, RCG_AUTO);
                need = requant_need(ctx, L);
                break;
            }
            default:
                set_error("unknown search mode");
                return RCG_INVALID;
        }
        *out = need + kDescBytes + 4096;
        return RCG_OK;
    });
}

// ================================================================================================================
// Tile policy and autotuning

namespace {

void tune_to_c(const SsimTune& T, rcg_tune* out) {
    out->wide_nt = T.wide_nt;
    out->narrow_w = T.narrow_w;
    out->nclass = T.nclass;
    for (int i = 0; i < 8; i++) {
        out->min_h[i] = T.th_min_h[i];
        out->th[i] = T.th[i];
    }
}

int tune_from_c(const rcg_tune* in, SsimTune& T) {
    if (!in || in->nclass < 1 || in->nclass > 8 || (in->wide_nt != 128 && in->wide_nt != 256)) return RCG_INVALID;
    for (int i = 0; i < in->nclass; i++) {
        if (in->th[i] < 2 || in->th[i] > 512 || (in->th[i] & 1)) return RCG_INVALID;
        if (i > 0 && in->min_h[i] > in->min_h[i - 1]) return RCG_INVALID;
    }
    if (in->min_h[in->nclass - 1] != 0) return RCG_INVALID;
    T.wide_nt = in->wide_nt;
    T.narrow_w = in->narrow_w;
    T.nclass = in->nclass;
    for (int i = 0; i < 8; i++) {
        T.th[i] = i < in->nclass ? in->th[i] : in->th[in->nclass - 1];
        T.th_min_h[i] = i < in->nclass ? in->min_h[i] : 0;
    }
    return RCG_OK;
}

}  // namespace

extern "C" int rcg_get_tune(rcg_ctx* ctx, rcg_tune* out) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !out) return RCG_INVALID;
        std::lock_guard<std::mutex> lock(ctx->mu);
        tune_to_c(ctx->tune32, out);
        return RCG_OK;
    });
}

extern "C" int rcg_set_tune(rcg_ctx* ctx, const rcg_tune* tune) {
    return rcg::guarded([&]() -> int {
        if (!ctx) return RCG_INVALID;
        std::lock_guard<std::mutex> lock(ctx->mu);
        SsimTune T = ctx->tune32;
        const int rc = tune_from_c(tune, T);
        if (rc != RCG_OK) {
            set_error("invalid tile policy");
            return rc;
        }
        ctx->tune32 = T;
        return RCG_OK;
    });
}

// Times k_ssim_scale (f32, the reference cache and f32 planes in, no pyramid out) on `jobs` pairs of S x S synthetic
// planes in one launch for every tile height, for S = 4096 ... 64, and installs the fastest tile height of each size
// class.
extern "C" int rcg_autotune(rcg_ctx* ctx, int jobs, rcg_tune* out) {
    return rcg::guarded([&]() -> int {
        // jobs = 0: throughput mode, for launches of very many pairs (a batch of many images). Each timing launch then
        // holds groups of 8 pairs, as many as make about 8 rounds of the GPU's workgroup slots, and the time per pair is
        // compared: a partly idle last round, which a launch of a few pairs pays and a large batch does not, no longer
        // favours short tiles.
        const bool throughput = jobs == 0;
        const int candidates = jobs;
        if (!ctx || candidates < 0 || candidates > 64) return RCG_INVALID;
        std::lock_guard<std::mutex> lock(ctx->mu);
        RCG_HIP(hipSetDevice(ctx->device));
        const uint32_t sizes[] = {4096, 2048, 1024, 512, 256, 128, 64};
        // Tile heights: a fine grid (a launch's last round of workgroups is the less idle the closer its tile count comes
        // to a multiple of the GPU's workgroup slots, and one step of the grid moves the count by a few percent). Small scales are latency-bound
        // (one workgroup walks TH + 10 rows one after another), so very short tiles are tried too; a launch of many
        // rounds (throughput) is best served by tall ones, whose 10-row halo costs least.
        const int ths[] = {2,  4,  6,  8,   12,  16,  24,  32,  40,  48,  56,  64,  72,  80,  88,  96,
                           104, 112, 128, 144, 160, 176, 208, 256, 320, 384, 512};
        const int nsz = 7;
        const size_t plane_bytes = size_t(4096) * 4096 * 3 * sizeof(float);
        ctx->arena.reset();
        RCG_TRY(ctx->arena.ensure(5 * Arena::align(plane_bytes) + (size_t(32) << 20) + (size_t(4) << 20)));
        RCG_TRY(ctx->pinned.ensure(size_t(4) << 20));
        float* a = static_cast<float*>(ctx->arena.take(plane_bytes));
        float* b = static_cast<float*>(ctx->arena.take(plane_bytes));
        float* cache = static_cast<float*>(ctx->arena.take(3 * plane_bytes));
        double* partials = static_cast<double*>(ctx->arena.take(size_t(32) << 20));
        char* ddesc = static_cast<char*>(ctx->arena.take(size_t(4) << 20));
        // Synthetic linear-light planes: a smooth pattern and a perturbed copy (the timing does not depend on values).
        {
            std::vector<float> h(size_t(4096) * 4096 * 3);
            uint64_t z = 0x9E3779B97F4A7C15ull;
            for (size_t i = 0; i < h.size(); i++) {
                z = z * 6364136223846793005ull + 1442695040888963407ull;
                h[i] = float((z >> 40) & 0xFFFF) / 65535.0f;
            }
            RCG_HIP(hipMemcpy(a, h.data(), plane_bytes, hipMemcpyHostToDevice));
            for (size_t i = 0; i < h.size(); i++) h[i] = std::min(1.0f, h[i] * 0.97f + 0.01f);
            RCG_HIP(hipMemcpy(b, h.data(), plane_bytes, hipMemcpyHostToDevice));
        }
        hipEvent_t e0, e1;
        RCG_HIP(hipEventCreate(&e0));
        RCG_HIP(hipEventCreate(&e1));
        SsimTune T = ctx->tune32;
        int best_th[nsz];
        const bool prof = ctx->prof.on;
        ctx->prof.on = false;
        for (int si = 0; si < nsz; si++) {
            const uint32_t S = sizes[si];
            double best = 1e30;
            best_th[si] = 8;
            for (int th : ths) {
                if (uint32_t(th) > S) continue;
                const int nt = S <= T.narrow_w ? 128 : T.wide_nt;
                const uint32_t tx = ssim_tiles_x(S, nt), ty = ssim_tiles_y(S, th);
                const uint32_t tiles = tx * ty;
                const int group = throughput ? 8 : candidates;
                const int pairs = throughput ? int(std::min<uint32_t>(4096u, ((8192u + tiles - 1) / tiles + 7) / 8 * 8))
                                             : candidates;
                if (size_t(tiles) * 18 * 8 * pairs > (size_t(32) << 20)) continue;
                if ((sizeof(SsimJob) + 64) * size_t(pairs) + sizeof(Unit) * (size_t(tiles) * pairs + 64 * kXcds) + 4096 >
                    (size_t(4) << 20))
                    continue;
                // The reference cache of the S x S plane `a` (the same for every tile height; rebuilt per size).
                RefJob rj;
                std::memset(&rj, 0, sizeof(rj));
                rj.w = S;
                rj.h = S;
                rj.th = uint32_t(th);
                rj.tiles_x = tx;
                rj.tiles_y = ty;
                rj.a = planes_src(a);
                rj.cache = cache;
                std::vector<SsimJob> jobs(pairs);
                for (int c = 0; c < pairs; c++) {
                    SsimJob& j = jobs[c];
                    std::memset(&j, 0, sizeof(j));
                    j.w = S;
                    j.h = S;
                    j.th = uint32_t(th);
                    j.tiles_x = tx;
                    j.tiles_y = ty;
                    j.a = cache_src(cache);
                    j.b = planes_src(b);
                    j.partials = partials + size_t(c) * tx * ty * 18;
                }
                // The candidates of a tile side by side on one XCD, as the scoring passes launch them.
                std::vector<UnitGroup> groups;
                for (int g = 0; g < pairs; g += group) groups.push_back(UnitGroup{uint32_t(g), uint32_t(group), tiles});
                const std::vector<Unit> units = xcd_units(groups);
                Desc D;
                D.host = ctx->pinned.base;
                D.dev = ddesc;
                D.cap = size_t(4) << 20;
                const SsimJob* dj = D.put(jobs.data(), jobs.size());
                const Unit* du = D.put(units.data(), units.size());
                const RefJob* drj = D.put(&rj, 1);
                RCG_TRY(D.upload(ctx->stream));
                RCG_TRY(launch_ssim_ref(ctx, drj, nullptr, tx * ty, nt, kSrcPlanes, ctx->stream));
                const uint32_t nu = uint32_t(units.size());
                RCG_TRY(launch_ssim_scale(ctx, RCG_PREC_FP32, dj, du, nu, nt, kSrcRefCache, kSrcPlanes));
                const size_t work = size_t(S) * S * size_t(pairs);
                const int reps = work >= (size_t(64) << 20) ? 2 : (work >= (size_t(16) << 20) ? 3 : 10);
                RCG_HIP(hipEventRecord(e0, ctx->stream));
                for (int r = 0; r < reps; r++)
                    RCG_TRY(launch_ssim_scale(ctx, RCG_PREC_FP32, dj, du, nu, nt, kSrcRefCache, kSrcPlanes));
                RCG_HIP(hipEventRecord(e1, ctx->stream));
                RCG_HIP(hipEventSynchronize(e1));
                float ms = 0;
                RCG_HIP(hipEventElapsedTime(&ms, e0, e1));
                const double per = double(ms) / reps / pairs;  // per pair; the pair count is fixed unless throughput
                if (per < best) {
                    best = per;
                    best_th[si] = th;
                }
            }
        }
        ctx->prof.on = prof;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        // Size classes: class i covers heights in [sizes[i], sizes[i-1]); adjacent classes with the same choice merge.
        SsimTune N = T;
        N.nclass = 0;
        for (int si = 0; si < nsz; si++) {
            const uint32_t lo = si == nsz - 1 ? 0 : sizes[si];
            if (N.nclass > 0 && N.th[N.nclass - 1] == best_th[si]) {
                N.th_min_h[N.nclass - 1] = lo;
                continue;
            }
            N.th[N.nclass] = best_th[si];
            N.th_min_h[N.nclass] = lo;
            N.nclass++;
        }
        for (int i = N.nclass; i < 8; i++) {
            N.th[i] = N.th[N.nclass - 1];
            N.th_min_h[i] = 0;
        }
        ctx->tune32 = N;
        if (out) tune_to_c(N, out);
        return RCG_OK;
    });
}

// ================================================================================================================
// Device buffers

extern "C" int rcg_device_alloc(rcg_ctx* ctx, size_t bytes, void** d_ptr) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !d_ptr) return RCG_INVALID;
        *d_ptr = nullptr;
        RCG_HIP(hipSetDevice(ctx->device));
        RCG_HIP(hipMalloc(d_ptr, bytes ? bytes : 1));
        return RCG_OK;
    });
}

extern "C" int rcg_device_free(rcg_ctx* ctx, void* d_ptr) {
    return rcg::guarded([&]() -> int {
        if (!ctx) return RCG_INVALID;
        if (!d_ptr) return RCG_OK;
        std::lock_guard<std::mutex> lock(ctx->mu);
        RCG_HIP(hipSetDevice(ctx->device));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        RCG_HIP(hipFree(d_ptr));
        return RCG_OK;
    });
}

extern "C" int rcg_memcpy_h2d(rcg_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
    return rcg::guarded([&]() -> int {
        if (!ctx || (!d_dst && bytes) || (!src && bytes)) return RCG_INVALID;
        std::lock_guard<std::mutex> lock(ctx->mu);
        RCG_HIP(hipSetDevice(ctx->device));
        RCG_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        return RCG_OK;
    });
}

extern "C" int rcg_memcpy_d2h(rcg_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
    return rcg::guarded([&]() -> int {
        if (!ctx || (!dst && bytes) || (!d_src && bytes)) return RCG_INVALID;
        std::lock_guard<std::mutex> lock(ctx->mu);
        RCG_HIP(hipSetDevice(ctx->device));
        RCG_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        return RCG_OK;
    });
}

extern "C" int rcg_synchronize(rcg_ctx* ctx) {
    return rcg::guarded([&]() -> int {
        if (!ctx) return RCG_INVALID;
        std::lock_guard<std::mutex> lock(ctx->mu);
        RCG_HIP(hipSetDevice(ctx->device));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        return RCG_OK;
    });
}

extern "C" int rcg_device_copy(rcg_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    return rcg::guarded([&]() -> int {
        if (!ctx || (bytes && (!d_dst || !d_src))) return RCG_INVALID;
        std::lock_guard<std::mutex> lock(ctx->mu);
        RCG_HIP(hipSetDevice(ctx->device));
        RCG_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        return RCG_OK;
    });
}
