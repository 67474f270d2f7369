// Host-side internals of librcgpu: the context (device, stream, workspace arena, descriptor staging) and the
// kernel launchers. Not part of the public C ABI (include/rcgpu.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <exception>
#include <map>
#include <new>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rcgpu.h"
#include "rcg_jobs.h"

namespace rcg {

// Message of the last failure on the calling thread (rcg_last_error). Never throws: without memory for the message
// it keeps a fixed one.
void set_error(const std::string& msg) noexcept;

// The exception barrier of every exported entry point: the ABI never unwinds. Host allocation failure becomes
// RCG_OOM, any other exception RCG_DEVICE, each with a message.
template <class F>
int guarded(F&& body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        set_error("out of host memory");
        return RCG_OOM;
    } catch (const std::exception& e) {
        set_error(std::string("internal error: ") + e.what());
        return RCG_DEVICE;
    } catch (...) {
        set_error("internal error");
        return RCG_DEVICE;
    }
}
template <class F>
void guarded_void(F&& body) noexcept {
    (void)guarded([&]() -> int {
        body();
        return RCG_OK;
    });
}

#define RCG_HIP(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            ::rcg::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                       \
            return e_ == hipErrorOutOfMemory ? RCG_OOM : RCG_DEVICE;                                   \
        }                                                                                              \
    } while (0)

// Device workspace: one growing allocation carved by a bump pointer. A call first adds up what it needs
// (reserve), then takes pointers (take); nothing is freed until the context is destroyed or grows.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    static size_t align(size_t n) { return (n + 255) & ~size_t(255); }
    int ensure(size_t bytes);
    void reset() { used = 0; }
    void* take(size_t bytes) {
        void* p = base + used;
        used += align(bytes);
        return p;
    }
};

// Pinned host staging for job descriptors and small results.
struct Pinned {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    int ensure(size_t bytes);
    void reset() { used = 0; }
    void* take(size_t bytes) {
        void* p = base + used;
        used += (bytes + 255) & ~size_t(255);
        return p;
    }
};

// Per-kernel HIP-event timing on the context's stream.
struct Prof {
    bool on = false;
    struct Pending {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, std::pair<double, int>> totals;
    hipEvent_t get();
    void collect();  // stream must be synchronised
};

// Tile policy of the SSIMULACRA2 scale kernel. A scale of w x h pixels runs with NT threads per workgroup
// (NT - 12 output columns; 128 when w <= narrow_w, else wide_nt) and TH output rows per tile, the largest of
// th[0..3] whose height threshold h >= th_min_h[i] holds. Small scales get short tiles: their cost is the latency of
// one workgroup's row walk, not throughput. The policy depends only on the image's own size, so an image's tiling
// (and its score, to the last bit) does not depend on what else is in a batch.
struct SsimTune {
    int wide_nt = 256;
    uint32_t narrow_w = 160;
    int nclass = 4;
    int th[8] = {64, 32, 16, 8, 8, 8, 8, 8};
    uint32_t th_min_h[8] = {2048, 512, 128, 0, 0, 0, 0, 0};  // descending; the last class must have 0
    int fixed_nt = 0;  // FP64 kernel: one thread count for every scale
    // Workgroups of the reference-cache kernel the GPU holds at once (256 and 128 threads); its tiles are sized so
    // that one image fills them in one round.
    uint32_t ref_slots_wide = 1792, ref_slots_narrow = 3584;
};

}  // namespace rcg

struct rcg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Second stream (default priority) for the reference-cache launches, which run beside the encode launch: forked
    // from `stream` (ev_fork), marked after the scale-0 reference launches (ev_mid: the candidates' scale-0 launch
    // waits only for those) and joined back into `stream` (ev_join).
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_mid = nullptr, ev_join = nullptr;
    std::mutex mu;
    rcg::Arena arena;
    rcg::Pinned pinned;
    float* d_lut = nullptr;  // sRGB -> linear table, 256 f32
    rcg::SsimTune tune32, tune64;
    rcg::Prof prof;
    bool prof_suspended = false;  // set during a call that must not create profiling events
    double taps64[6];
    float taps32[6];
};

namespace rcg {

// Sampling layout of build_coefs (jpeg/src/encode.rs:428-443).
struct Layout {
    uint32_t w, h, channels;
    uint32_t hmax, vmax, pw, ph;
    uint32_t ncomp;
    uint32_t bw[4], bh[4];  // per component (a fourth only for CMYK/YCCK coefficient images)
    size_t coef_count(int c) const { return size_t(bw[c]) * bh[c] * 64; }
    size_t plane_bytes(int c) const { return size_t(bw[c]) * 8 * bh[c] * 8; }
};
Layout make_layout(uint32_t w, uint32_t h, uint32_t channels, int quality, int subsampling);
// Per-component maxima of the block grids a search over every quality meets (Auto: 4:4:4 and 4:2:0).
Layout max_layout(uint32_t w, uint32_t h, uint32_t channels, int subsampling);
void quant_tables(int quality, uint8_t luma[64], uint8_t chroma[64]);
void fill_cand_tables(const uint8_t qy[64], const uint8_t qc[64], EncCand& c);

// Brackets one kernel launch with profiling events when profiling is on (on `stream`, by default the context's).
struct ProfScope {
    rcg_ctx* ctx;
    const char* name;
    hipStream_t stream;
    hipEvent_t b = nullptr;
    ProfScope(rcg_ctx* c, const char* n, hipStream_t s = nullptr);
    ~ProfScope();
};

// Launchers (defined next to their kernels).
// coef_only: no candidate of the launch asks for reconstruction planes or trellis hand-off values.
int launch_encode(rcg_ctx* ctx, const EncJob* d_jobs, const uint32_t* d_begins, int njobs, uint32_t total_tiles,
                  bool coef_only = false);
struct IdctPlane {
    const int16_t* coefs;
    uint8_t* plane;
    uint32_t blocks_w, blocks_h;
    float mult[64];
    // Requantisation fused ahead of the dequantisation (when requant): each coefficient becomes
    // clamp(round_f32(c * ratio[k]), +-2047) (k: zigzag position), is written to requant_out if not null, and is
    // dequantised with mult (the new table's).
    int requant;
    float ratio[64];
    int16_t* requant_out;
};
int launch_idct(rcg_ctx* ctx, const IdctPlane* planes, int n);
// Full-size pixels from sample planes: Gray8 for one component, RGB8 for three (YCbCr, or copied when rgb_direct)
// and for four (cmyk_to_rgb with the Adobe transform, -1 without the marker).
int launch_upsample_color(rcg_ctx* ctx, const PlaneView* comps, int ncomp, int rgb_direct, uint32_t w, uint32_t h,
                          uint8_t* out, int adobe_transform = -1);
int launch_ssim_scale(rcg_ctx* ctx, int precision, const SsimJob* d_jobs, const Unit* d_units, uint32_t nunits, int nt,
                      uint32_t kind_a, uint32_t kind_b, hipStream_t stream = nullptr);  // null: the context's stream
int launch_ssim_ref(rcg_ctx* ctx, const RefJob* d_jobs, const Unit* d_units, uint32_t nunits, int nt,
                    uint32_t kind_a, hipStream_t stream);
// Fork the side stream from the context's stream (everything enqueued there so far is visible to it), and join it
// back (the context's stream waits for everything enqueued on the side stream so far).
int fork_side(rcg_ctx* ctx);
int join_side(rcg_ctx* ctx);
int mark_side(rcg_ctx* ctx);       // record ev_mid on the side stream
int wait_side_mark(rcg_ctx* ctx);  // the context's stream waits for ev_mid

// Jobs job0 .. job0+njobs-1 of a launch that share one tiling of `tiles` tiles (the candidates of one image at one
// scale, or a single job).
struct UnitGroup {
    uint32_t job0, njobs, tiles;
    uint32_t chunk = 1;  // consecutive tiles kept together on one XCD
};
// The unit table of a launch (see Unit): each `chunk` consecutive tiles of a group become a chunk of chunk*njobs
// consecutive units on one XCD, and chunks go to the XCD with the fewest units so far (round-robin when the groups
// are alike). XCDs left shorter than the longest are padded with idle units. A launch of single-job, single-tile-chunk
// groups gets units in plain tile order.
std::vector<Unit> xcd_units(const std::vector<UnitGroup>& groups);
struct FinalDesc {
    uint32_t nscales;
    uint32_t ntiles[6];
    const double* partials[6];
    double count[6];
    double* stats_out;
    double* score_out;
};
int launch_finalize(rcg_ctx* ctx, const FinalDesc* d_jobs, int n);
int upload_weights();
// Resident workgroups of the reference-cache kernel on `device` (256- and 128-thread instances).
int ref_kernel_slots(int device, uint32_t* wide, uint32_t* narrow);

// DCT-domain requantisation of a coefficient image towards the tables of `quality` (jpeg/src/encode.rs:107-145):
// per component the new table (the quality's, never finer than the old one), the old/new ratio per zigzag position
// and the new table slot (components with equal tables share one).
struct RequantPlan {
    uint32_t ncomp;
    uint8_t table[4][64];   // natural order, per component
    float ratio[4][64];     // zigzag order, per component
    uint8_t tq[4];          // new slot per component
    uint8_t slots[4][64];   // new tables by slot
    int present[4];
    bool changed;           // some component's table changed
};
int requant_plan(const rcg_coef_image* ci, int quality, RequantPlan& P);

// reconstruct (jpeg/src/decode.rs:334, 1390-1478) of a coefficient image whose coefficients are on the device.
// check_coef_image validates it (RCG_INVALID / RCG_UNSUPPORTED as the reference's errors) and derives the sampling
// maxima and the RGB-coded flag; out_channels is 1 for gray, else 3; sample_plane_bytes is what the IDCT's planes
// take (256-byte aligned per component); reconstruct_on_device enqueues the IDCT into planes_out and the colour stage
// into d_out on the context's stream, or, when fused_ok and the layout allows it, the fused kernel alone (planes_out
// unused).
int check_coef_image(const rcg_coef_image* ci, uint32_t& hmax, uint32_t& vmax, int& rgb_direct);
uint32_t out_channels(const rcg_coef_image* ci);
size_t sample_plane_bytes(const rcg_coef_image* ci);
void mult_table(const uint8_t q[64], float mult[64]);
int reconstruct_on_device(rcg_ctx* ctx, const rcg_coef_image* ci, const int16_t* const d_coefs[4],
                          uint8_t* const planes_out[4], uint8_t* d_out, uint32_t hmax, uint32_t vmax, int rgb_direct,
                          bool fused_ok = true);
// The fused reconstruct (one kernel, no sample planes) takes component 0 at full resolution and, for colour,
// components 1 and 2 sharing upsampling factors of 1 or 2 each: Gray, 4:4:4, 4:2:2, 4:4:0 and 4:2:0.
bool fused_reconstruct_layout(const rcg_coef_image* ci, uint32_t hmax, uint32_t vmax, int* fx, int* fy);
int launch_reconstruct_fused(rcg_ctx* ctx, const rcg_coef_image* ci, const int16_t* const d_coefs[4], uint8_t* d_out,
                             uint32_t hmax, uint32_t vmax, int rgb_direct);
// Set-up of the allocation-free entry points on `device` (current): finds the device address of the static scratch
// array once per device. Called by rcg_create.
int noalloc_setup(int device);

// Tiles of the SSIMULACRA2 scale kernel for a w x h scale.
inline uint32_t ssim_tiles_x(uint32_t w, int nt) { return (w + (nt - 12) - 1) / (nt - 12); }
inline uint32_t ssim_tiles_y(uint32_t h, int th) { return (h + th - 1) / th; }

// One SSIMULACRA2 evaluation of an image pair through all of its scales, described for the batched launcher.
// The FP32 candidate kernel covers a scale with tiles_x tiles of nt threads (nt - 12 output columns each; nt is 256,
// 128 or 64) from column 0, then, when tnt is not 0, one column of narrower tail tiles of tnt threads for the columns
// left over, so that a width that is not a multiple of nt - 12 leaves few lanes idle. Partial sums are stored main
// tiles first, then tail tiles: tiles[s] counts both. The reference-cache kernel has a tiling of its own (nt_ref
// threads, the scale's class: 128 up to the policy's narrow width, else 256; its output does not depend on it).
struct PairPlan {
    uint32_t w, h, nscales;
    uint32_t sw[6], sh[6];
    uint32_t nt[6], th[6];
    uint32_t tiles[6], tiles_x[6], tiles_y[6];
    uint32_t tnt[6];                   // threads of the tail tiles, 0 for none
    uint32_t nt_ref[6], tiles_x_ref[6];
    uint32_t th_ref[6], tiles_ref[6];  // tiling of the reference-cache kernel (its own columns and rows)
    uint32_t tiles_main(int s) const { return tiles_x[s] * tiles_y[s]; }
    uint32_t x_tail(int s) const { return tiles_x[s] * (nt[s] - 12); }
};
PairPlan plan_pair(uint32_t w, uint32_t h, const SsimTune& tune);

// Descriptor staging: descriptors are written to pinned memory at the offsets they will have in a device region
// and uploaded with one copy before the launches that read them.
struct Desc {
    char* host = nullptr;
    char* dev = nullptr;
    size_t off = 0, cap = 0;
    bool overflow = false;
    template <class T>
    T* put(const T* items, size_t n) {
        off = (off + 63) & ~size_t(63);
        if (off + sizeof(T) * n > cap) {
            overflow = true;
            return reinterpret_cast<T*>(dev);
        }
        T* d = reinterpret_cast<T*>(dev + off);
        std::memcpy(host + off, items, sizeof(T) * n);
        off += sizeof(T) * n;
        return d;
    }
    int upload(hipStream_t s) { return upload_range(0, off, s); }
    // Bytes [begin, end) of what was staged.
    int upload_range(size_t begin, size_t end, hipStream_t s) {
        if (overflow) {
            set_error("internal: descriptor staging overflow");
            return RCG_DEVICE;
        }
        if (end > begin) RCG_HIP(hipMemcpyAsync(dev + begin, host + begin, end - begin, hipMemcpyHostToDevice, s));
        return RCG_OK;
    }
};

// Sources of the SSIMULACRA2 loader.
SsimSrc recon_src(const Layout& L, uint8_t* const rec[3]);
SsimSrc bytes_src(const uint8_t* d, uint32_t channels);
SsimSrc planes_src(const float* planes);
SsimSrc cache_src(const float* cache);
int check_image(uint32_t w, uint32_t h, uint32_t channels);

// The reference side of an image's f32 scoring, built once per image and read by every candidate: its linear-light
// pyramid (planes[s], [3][h_s][w_s] for scales 1..; the f64 kernel reads it too) and its per-scale cache (cache[s],
// 9 * w_s * h_s floats, written by k_ssim_ref).
struct RefState {
    float* planes[6];
    float* cache[6];
};
size_t ref_state_bytes(const PairPlan& P);
void take_ref_state(Arena& A, const PairPlan& P, RefState& R);
// The reference-cache job of scale s; a0 is the scale-0 source of the reference image.
RefJob make_ref_job(const PairPlan& P, uint32_t s, const SsimSrc& a0, const RefState& R);

// One image of a scoring pass: its resident source, its reference state (built by the pass when ref_ready is false)
// and the candidates to score. Scores come back in scores[]; settled[c] marks FP64 ones.
struct Item {
    const uint8_t* d_src;
    uint32_t w, h, ch;
    uint32_t ncand;
    int quality[RCG_MAX_CAND];
    int subsampling[RCG_MAX_CAND];
    double target;
    RefState ref;
    bool ref_ready;
    double scores[RCG_MAX_CAND];
    int settled[RCG_MAX_CAND];
};
size_t score_items_bytes(rcg_ctx* ctx, const Item* items, int n);
int score_items(rcg_ctx* ctx, Item* items, int n);

// The chosen candidate's coefficients of some items (Found.coefs, jpeg/src/target.rs:11-16): build_coefs at that
// candidate's quality and subsampling, encoded once more in one launch over the items. dest[k] receives item
// which[k]'s candidate chosen[k]: device pointers written by the kernel (16-byte aligned) when host_dest is false,
// else host pointers filled by copies from arena buffers. winner_bytes() is the arena space encode_winners() takes
// and winner_pinned() the pinned staging it needs. Synchronous.
struct WinnerDest {
    int16_t* coefs[3];
};
size_t winner_bytes(const Item* items, const int* which, int n, bool host_dest);
size_t winner_pinned(int n);
int encode_winners(rcg_ctx* ctx, Arena& A, Pinned& pin, const Item* items, const int* which, const int* chosen,
                   const WinnerDest* dest, int n, bool host_dest);
int select_candidate(const int* quality, const double* scores, int n, double target, int* sweep);

#define RCG_TRY(expr)                    \
    do {                                 \
        int rc_ = (expr);                \
        if (rc_ != RCG_OK) return rc_;   \
    } while (0)

}  // namespace rcg
