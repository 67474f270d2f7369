/*
 * rcgpu.h -- C ABI of librcgpu, the MI355X (gfx950) implementation of rust-compress's JPEG target-quality inner
 * loop: RGB->YCbCr with chroma downsampling, 8x8 f32 AAN FDCT and quantisation, dequantisation + IDCT back to
 * 8-bit pixels, and SSIMULACRA2 scoring, driven by the quality search.
 *
 * Every entry point takes plain pointers and sizes, returns an rcg_status, never aborts and never unwinds. The
 * functions without a _dev suffix take HOST buffers (the drop-in for the Rust API, which owns its Vec<u8>/Vec<i16>);
 * the _dev variants take DEVICE pointers that stay resident in HBM and enqueue on the context's stream.
 *
 * Layouts are the reference's own:
 *   pixels        interleaved Gray8 (channels = 1) or RGB8 (channels = 3), row stride width*channels
 *                 (rc_core::image::Image::data, core/src/image.rs:104-113);
 *   coefficients  per component, 64 int16 per 8x8 block in zigzag order, blocks in raster order over the padded
 *                 block grid (rc_jpeg::encode::CoefComponent::coefs, jpeg/src/encode.rs:56-69), so the Rust side
 *                 wraps the buffers with no permutation;
 *   qtables       64 uint8 in natural (row-major) order (CoefImage::qtables, jpeg/src/encode.rs:77-78).
 *
 * Thread safety: calls on one context are serialised by the context; use one context per thread for overlap.
 */
#ifndef RCGPU_H
#define RCGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. They map rc_core::error::Error (core/src/error.rs:4-22): Invalid -> RCG_INVALID,
 * Unsupported -> RCG_UNSUPPORTED; device failures and allocation failures get their own codes. */
typedef enum {
    RCG_OK = 0,
    RCG_INVALID = 1,
    RCG_UNSUPPORTED = 2,
    RCG_DEVICE = 3,
    RCG_OOM = 4
} rcg_status;

/* rc_jpeg::encode::Subsampling (jpeg/src/encode.rs:12-20), same order. */
typedef enum { RCG_AUTO = 0, RCG_S444 = 1, RCG_S422 = 2, RCG_S420 = 3 } rcg_subsampling;

/* Arithmetic of the SSIMULACRA2 scorer: FP32 maps with FP64 pooling, or the reference's FP64 arithmetic. */
typedef enum { RCG_PREC_FP32 = 0, RCG_PREC_FP64 = 1 } rcg_precision;

typedef struct rcg_ctx rcg_ctx;

/* Create a context on HIP device `device` (its own stream and device workspace). */
int rcg_create(int device, rcg_ctx** out);
void rcg_destroy(rcg_ctx* ctx);
/* Message of the last failure on the calling thread ("" if none). */
const char* rcg_last_error(void);
/* Library version string. */
const char* rcg_version(void);

/* Device buffers for the _dev entry points (resident sources, outputs): plain hipMalloc'd memory on the
 * context's device, copies ordered on the context's stream and complete on return. */
int rcg_device_alloc(rcg_ctx* ctx, size_t bytes, void** d_ptr);
int rcg_device_free(rcg_ctx* ctx, void* d_ptr);
int rcg_memcpy_h2d(rcg_ctx* ctx, void* d_dst, const void* src, size_t bytes);
int rcg_memcpy_d2h(rcg_ctx* ctx, void* dst, const void* d_src, size_t bytes);
int rcg_device_copy(rcg_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
/* Wait for everything enqueued on the context's stream. */
int rcg_synchronize(rcg_ctx* ctx);

/* Coefficient layout build_coefs produces (jpeg/src/encode.rs:428-443): sampling factors and block grids. */
typedef struct {
    uint32_t width, height;
    uint32_t ncomp;                 /* 1 (Gray) or 3 (YCbCr) */
    uint32_t hmax, vmax;            /* luma sampling factors; chroma are 1x1 */
    uint32_t blocks_w[3], blocks_h[3];
} rcg_layout;
int rcg_coef_layout(uint32_t width, uint32_t height, uint32_t channels, int quality, int subsampling,
                    rcg_layout* out);

/* Quantisation tables for `quality` (jpeg/src/quality.rs:6-43): luma then chroma, natural order. */
int rcg_quant_tables(int quality, uint8_t out[128]);

/*
 * Replaces rc_jpeg::encode::build_coefs(&Image, &EncodeOptions) -> CoefImage (jpeg/src/encode.rs:426) with
 * trellis off: padding by edge replication, fixed-point RGB->YCbCr, box downsampling, f32 AAN FDCT and reciprocal
 * quantisation. coefs[c] receives layout.blocks_w[c]*blocks_h[c]*64 int16 (coefs[1], coefs[2] unused for Gray).
 * qtables (may be NULL) receives the two tables written to DQT. Bit-exact with the reference.
 */
int rcg_build_coefs(rcg_ctx* ctx, const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                    int quality, int subsampling, int16_t* coefs[3], uint8_t qtables[128]);
int rcg_build_coefs_dev(rcg_ctx* ctx, const uint8_t* d_pixels, uint32_t width, uint32_t height, uint32_t channels,
                        int quality, int subsampling, int16_t* d_coefs[3]);

/* A coefficient image as rc_jpeg::encode::CoefImage holds it (jpeg/src/encode.rs:56-83). */
typedef struct {
    uint8_t id;
    uint8_t h, v;                   /* sampling factors */
    uint8_t tq;                     /* quantisation table slot 0..3 */
    uint32_t blocks_w, blocks_h;
    const int16_t* coefs;           /* blocks_w*blocks_h*64 zigzag int16 */
} rcg_coef_component;
typedef struct {
    uint32_t width, height;
    uint32_t ncomp;                 /* 1, 3 or 4 (CMYK/YCCK) */
    rcg_coef_component comps[4];
    const uint8_t* qtables[4];      /* natural-order tables by slot, NULL if absent */
    int adobe_transform;            /* -1 when the file has no Adobe marker */
} rcg_coef_image;

/*
 * Replaces rc_jpeg::decode::reconstruct(&CoefImage) -> Result<Image> (jpeg/src/decode.rs:334): dequantise, f32
 * AAN IDCT, round and clamp, "fancy" upsampling of subsampled components, fixed-point YCbCr->RGB (or a copy for
 * RGB-coded components), and for four components cmyk_to_rgb with the Adobe transform (decode.rs:1446-1478). pixels
 * receives width*height bytes for one component (Gray8) and width*height*3 (RGB8) for three or four. RCG_INVALID for a
 * missing table; RCG_UNSUPPORTED for a non-integral sampling ratio (decode.rs:1314,1410-1412). Bit-exact with the
 * reference.
 */
int rcg_reconstruct(rcg_ctx* ctx, const rcg_coef_image* ci, uint8_t* pixels);
int rcg_reconstruct_dev(rcg_ctx* ctx, const rcg_coef_image* d_ci, uint8_t* d_pixels);

/*
 * Replaces rc_core::ssimulacra2::ssimulacra2(&Image, &Image) -> Result<f64> (core/src/ssimulacra2.rs:471).
 * RCG_INVALID when a side is below 8 pixels or the score is not finite (ssimulacra2.rs:472-477,511).
 * per_scale (may be NULL) receives [6][3][6] doubles: per scale, per channel X/Y/B, the SSIM mean and 4-norm,
 * artifact mean and 4-norm, detail mean and 4-norm; nscales (may be NULL) the number of scales used.
 */
int rcg_ssimulacra2(rcg_ctx* ctx, const uint8_t* reference, const uint8_t* distorted, uint32_t width,
                    uint32_t height, uint32_t channels, int precision, double* score, double* per_scale,
                    int* nscales);
int rcg_ssimulacra2_dev(rcg_ctx* ctx, const uint8_t* d_reference, const uint8_t* d_distorted, uint32_t width,
                        uint32_t height, uint32_t channels, int precision, double* score, double* per_scale,
                        int* nscales);

/* Result of a quality search: rc_jpeg::target::Found (jpeg/src/target.rs:11-16) without the coefficients, which
 * the caller gets through the coefs argument. */
typedef struct {
    int quality;
    double score;                   /* score of the chosen quality (FP64 when it was within 0.05 of the target) */
    int steps;                      /* candidates evaluated */
} rcg_found;

/*
 * Replaces rc_jpeg::target::find_quality_with(reference, target, max_quality, make, metric) as the CLI wires it
 * for --target-ssimulacra2 (cli/src/pipeline.rs:207-247): make = build_coefs(trellis off), metric = ssimulacra2,
 * serial bisection over 1..=max_quality starting at max_quality (target.rs:37-55). Scores use FP32 maps; a score
 * within 0.05 of the target is re-scored in FP64 before the comparison, so the choice equals the reference's.
 * coefs (may be NULL) receive the chosen quality's coefficients, qtables (may be NULL) its tables. Each coefs[c]
 * must hold the largest grid the search can meet: rcg_coef_layout at quality 100, and under RCG_AUTO also at
 * quality 1 (4:2:0 pads luma to 16-pixel MCUs, so its luma grid can exceed the 4:4:4 one).
 * RCG_UNSUPPORTED below 8x8 pixels (pipeline.rs:215-217).
 */
int rcg_find_quality(rcg_ctx* ctx, const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                     double target, int max_quality, int subsampling, rcg_found* out, int16_t* coefs[3],
                     uint8_t qtables[128]);
int rcg_find_quality_dev(rcg_ctx* ctx, const uint8_t* d_pixels, uint32_t width, uint32_t height,
                         uint32_t channels, double target, int max_quality, int subsampling, rcg_found* out);

/*
 * Batched candidate scoring: many images of uneven size, each with up to 16 candidate (quality, subsampling)
 * pairs, scored in one sequence of launches (one encode launch for every image and candidate, one launch per
 * SSIMULACRA2 scale for every pair, one pooling launch). Candidates of one image that share a subsampling share
 * its colour conversion and FDCT; the reference image's pyramid is built once per image. The selection follows
 * find_quality_with (jpeg/src/target.rs:37-55) over the candidate list sorted by quality: the highest candidate
 * first, then bisection over list indices; chosen_* report that choice and sweep_* the exhaustive one (the lowest
 * passing candidate). Candidates whose FP32 score lies within 0.05 of the target are re-scored in FP64 first.
 * Pixels are DEVICE pointers for rcg_score_batch_dev and HOST pointers for rcg_score_batch.
 * winner (optional, winner[0] NULL when not wanted) receives the chosen candidate's coefficients, Found.coefs of
 * jpeg/src/target.rs:11-16 (build_coefs at the chosen quality and subsampling, coefficients as rcg_build_coefs):
 * DEVICE buffers, 16-byte aligned, for rcg_score_batch_dev and plans, HOST buffers for rcg_score_batch. winner[c]
 * must hold component c's grid for every candidate of the job (rcg_coef_layout); winner[1..2] are unused for Gray.
 * The chosen candidate is encoded once more after the selection, in one launch for every job that asked.
 * This entry point has no single counterpart in the reference: it replaces the serial find_quality_with loop of
 * cli/src/pipeline.rs:207-247 run once per file by cli/src/main.rs:57-75.
 */
#define RCG_MAX_CAND 16
typedef struct {
    const uint8_t* pixels;
    uint32_t width, height, channels;
    uint32_t ncand;
    int quality[RCG_MAX_CAND];
    int subsampling[RCG_MAX_CAND];
    double target;
    int16_t* winner[3];
} rcg_batch_job;
typedef struct {
    double scores[RCG_MAX_CAND];    /* per candidate, FP64-settled near the target */
    int settled[RCG_MAX_CAND];      /* 1 if that score is the FP64 one */
    int chosen_index, chosen_quality;
    int sweep_index, sweep_quality;
    int status;                     /* rcg_status of this job */
    int winner_written;             /* 1 when the job's winner buffers hold the chosen candidate's coefficients */
} rcg_batch_result;
int rcg_score_batch_dev(rcg_ctx* ctx, const rcg_batch_job* jobs, int njobs, rcg_batch_result* results);
int rcg_score_batch(rcg_ctx* ctx, const rcg_batch_job* jobs, int njobs, rcg_batch_result* results);

/*
 * The batched scoring of rcg_score_batch_dev captured once as a hipGraph: rcg_batch_plan_create allocates every
 * buffer, uploads every descriptor and records the launch sequence (encode, six scale launches, pooling, one
 * copy of the scores to pinned host memory); rcg_batch_plan_run replays it with one graph launch, then settles
 * scores near a target in FP64 and applies the selection. The jobs' device pixels must stay valid and unchanged in
 * place between runs (new pixel values at the same addresses are fine). No counterpart in the reference.
 */
typedef struct rcg_plan rcg_plan;
int rcg_batch_plan_create(rcg_ctx* ctx, const rcg_batch_job* jobs, int njobs, rcg_plan** out);
int rcg_batch_plan_run(rcg_plan* plan, rcg_batch_result* results);
void rcg_batch_plan_destroy(rcg_plan* plan);

/*
 * Tile policy of the SSIMULACRA2 scale kernel: a scale of height h uses th[i] output rows per tile for the first
 * class i with h >= min_h[i] (classes in descending min_h, the last with min_h = 0), and 128-thread workgroups when
 * its width is at most narrow_w (else wide_nt = 256). The policy depends only on an image's own size, so a score
 * never depends on what else shares a launch; different policies change scores only in the last bits.
 * rcg_autotune times the scale kernel for th in {2, 4, 6, 8, 12, 16, 24, 32, 40, ..., 112 (steps of 8), 128, 144,
 * 160, 176, 208, 256, 320, 384, 512} at each size class
 * (4096 down to 64 rows) and installs the fastest per class; it returns the installed policy. `jobs` (1..64) is how
 * many same-size pairs share a launch in the workload the policy is for: the candidates of one image for a single
 * search or sweep, more for a batch of many images (their launches hold many pairs of every size, so short tiles,
 * which only shorten the latency of a lone small image, cost halo work instead). `jobs` = 0 tunes for throughput:
 * each timing launch holds enough pairs for about eight rounds of the GPU's workgroup slots and the time per pair
 * decides, which suits a large batch, whose launches run many rounds.
 */
typedef struct {
    int wide_nt;
    uint32_t narrow_w;
    int nclass;
    uint32_t min_h[8];
    int th[8];
} rcg_tune;
int rcg_get_tune(rcg_ctx* ctx, rcg_tune* out);
int rcg_set_tune(rcg_ctx* ctx, const rcg_tune* tune);
int rcg_autotune(rcg_ctx* ctx, int jobs, rcg_tune* out);

/*
 * Serving loop with continuous batching. Two lanes, each a worker thread with a context of its own, serve the
 * requests below 2 MP and those of 2 MP or more, so that a small request's ticks do not wait for a large image's;
 * the lanes' ticks run side by side on the GPU. Requests (an image and a --target-ssimulacra2 goal) are admitted in
 * arrival order within their lane while the device footprint of both lanes fits the HBM budget, the
 * GPU analogue of cli/src/budget.rs:11-50: a request that alone exceeds the budget runs, but by itself. Every tick
 * scores the next candidate(s) of every admitted request's bisection (jpeg/src/target.rs:37-55) in one batched
 * launch sequence; requests that arrive meanwhile join at the next tick. Sources and reference pyramids stay
 * resident in HBM for the life of a request; a finished request's device blocks are kept for reuse by later ones
 * (admitted footprints and kept blocks together stay within hbm_budget), so a steady stream of requests allocates
 * nothing. Each result equals the one the request gets alone. A tick that fails (a working set that cannot be
 * allocated, say) is retried request by request, so only the requests that fail alone fail; an exception never leaves
 * the worker (host allocation failure is RCG_OOM for the requests of that tick).
 * speculate = 1 scores one bisection midpoint per tick; 3 also scores both possible next midpoints, so a request
 * takes about half the ticks for 1.5x the work.
 * This entry point has no single counterpart in the reference (it replaces the per-file rayon tasks of
 * cli/src/main.rs:57-75 gated by Budget::acquire, cli/src/pipeline.rs:881).
 */
typedef struct rcg_server rcg_server;
typedef struct {
    size_t hbm_budget;              /* bytes for admitted requests; 0 = 16 GiB */
    int max_active;                 /* requests in flight at once; 0 = no limit */
    int speculate;                  /* 1 or 3 */
    const rcg_tune* tune;           /* tile policy of the workers' contexts, or NULL for the default one */
} rcg_server_opts;
typedef struct {
    uint64_t id;                    /* caller's tag, returned in the completion */
    const uint8_t* pixels;          /* host Gray8/RGB8, copied by rcg_server_submit */
    uint32_t width, height, channels;
    double target;
    int max_quality;                /* 1..100 (the CLI uses 100) */
    int subsampling;
    /* Optional HOST buffers (coefs[0] NULL: not wanted) for the chosen quality's coefficients, Found.coefs of
     * jpeg/src/target.rs:11-16, sized as rcg_find_quality's; they must stay valid until the completion is polled. */
    int16_t* coefs[3];
} rcg_request;
typedef struct {
    uint64_t id;
    int status;                     /* rcg_status of this request */
    int quality;
    double score;
    int steps;                      /* candidates scored */
    int ticks;                      /* batched launches it took part in */
    double submit_s, admit_s, done_s;   /* steady-clock seconds */
    int coefs_written;              /* 1 when the request's coefs buffers hold the chosen quality's coefficients */
} rcg_completion;
typedef struct {
    uint64_t ticks, requests_done;
    double mean_batch;              /* mean requests per tick */
    size_t peak_footprint;          /* largest admitted footprint, bytes */
    double busy_s;                  /* time spent inside ticks, summed over the two lanes */
} rcg_server_stats;
int rcg_server_create(int device, const rcg_server_opts* opts, rcg_server** out);
/* RCG_INVALID / RCG_UNSUPPORTED for a request that can never run (bad size, under 8x8); nothing is queued then. */
int rcg_server_submit(rcg_server* server, const rcg_request* request);
/* Up to max completions, waiting at most timeout_ms for the first one; *n receives the count. */
int rcg_server_poll(rcg_server* server, rcg_completion* out, int max, int timeout_ms, int* n);
int rcg_server_get_stats(rcg_server* server, rcg_server_stats* out);
void rcg_server_destroy(rcg_server* server);

/* Per-kernel timing with HIP events recorded around every launch, on the stream the launch runs on. enable = 1
 * starts recording (and clears the totals), 0 stops. rcg_profile_get returns the summed milliseconds and the launch
 * count of one kernel ("k_encode", "k_reconstruct", "k_idct", "k_upsample_color", "k_ssim_ref", "k_ssim_scale" (FP32),
 * "k_ssim_scale_f64", "k_ssim_finalize"). */
int rcg_profile(rcg_ctx* ctx, int enable);
int rcg_profile_get(rcg_ctx* ctx, const char* kernel, double* total_ms, int* launches);

/* Fill `out` (width*height*3 bytes, host) with the seeded synthetic RGB8 test image: per channel 4 random 2-D
 * sinusoids (periods 16-256 px, amplitudes 10-60) around 128, 12 hard-edged rectangles, Gaussian noise sigma 4,
 * rounded and clamped. Deterministic for (seed, width, height) on any thread count. */
int rcg_synth_rgb8(uint64_t seed, uint32_t width, uint32_t height, uint8_t* out);

/* ------------------------------------------------------------------------------------------------------------
 * Trellis hand-off (jpeg/src/encode.rs:527-537, jpeg/src/rdo.rs). The reference's default encode re-quantises
 * every block with a rate-distortion trellis that runs on the CPU. The device supplies what that pass needs, so the
 * host trellis skips its own FDCT:
 *   rcg_build_scaled  the plain coefficients (as rcg_build_coefs) and, per block, the 64 f32 values
 *                     a[k] = dct[f]*div[f] of quantize_plane's trellis branch (encode.rs:206-211), zigzag order,
 *                     same block layout as the coefficients. Bit-exact with the reference.
 *   rcg_trellis       the trellis pass on the host: code lengths measured on the plain coefficients
 *                     (ac_code_lengths, encode.rs:222-227), then per block the DC rounded and clamped and
 *                     Rdo::quantize_ac (rdo.rs:52-113) with TRELLIS_KAPPA. plain: host coefficient image with the
 *                     tables in slots 0 (luma) and 1 (chroma); scaled[c]: that component's values; out[c] receives
 *                     the trellis coefficients (may alias nothing in plain). Bit-exact with the reference.
 *   rcg_rdo_quantize_ac  rc_jpeg::rdo::Rdo::new(table, kappa, sizes).quantize_ac(a, out) for one block (positions
 *                     1..63 of out are written).
 */
int rcg_build_scaled(rcg_ctx* ctx, const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                     int quality, int subsampling, int16_t* coefs[3], float* scaled[3], uint8_t qtables[128]);
int rcg_trellis(const rcg_coef_image* plain, const float* const scaled[4], int16_t* out[4]);
int rcg_rdo_quantize_ac(const uint8_t table[64], float kappa, const uint8_t sizes[256], const float a[64],
                        int16_t out[64]);

/* The target search of rcg_find_quality with the reference's default make = build_coefs(trellis on)
 * (cli/src/pipeline.rs:238 with EncodeOptions::trellis, encode.rs:527-537): every bisection step encodes on the
 * device with the hand-off values, runs rcg_trellis on the host, uploads its coefficients, and reconstructs and
 * scores them on the device against a reference state built once per call (on a side stream, while the first
 * step's encode and trellis run). Same arguments, guard band and results as rcg_find_quality; coefs receive the
 * chosen quality's trellis coefficients. */
int rcg_find_quality_trellis(rcg_ctx* ctx, const uint8_t* pixels, uint32_t width, uint32_t height,
                             uint32_t channels, double target, int max_quality, int subsampling, rcg_found* out,
                             int16_t* coefs[3], uint8_t qtables[128]);

/* ------------------------------------------------------------------------------------------------------------
 * Allocation-free reconstruct. No counterpart in the reference: this is for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops and harnesses that forbid hipMalloc). Same arguments, results
 * and errors as rcg_reconstruct_dev, and the same kernels, but the IDCT's sample planes live in a static device array
 * of RCG_NOALLOC_SCRATCH_BYTES per GPU (every GPU the library runs on holds its copy until the process ends) instead of
 * the context's growing workspace, and the call allocates nothing: no hipMalloc/hipFree, no profiling events. The
 * context must have been created (rcg_create, which allocates, and looks up the array's device address). An image
 * whose sample planes (per component blocks_w*blocks_h*64 bytes, rounded up to 256) exceed the array is refused with
 * RCG_UNSUPPORTED before anything is enqueued: about 44 MP at 4:2:0, 22 MP at 4:4:4. Calls on one GPU share its
 * array, so overlapping calls would race; the library serialises them per GPU. The array keeps the last call's
 * planes. Synchronous, as rcg_reconstruct_dev. */
#define RCG_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int rcg_reconstruct_dev_noalloc(rcg_ctx* ctx, const rcg_coef_image* d_ci, uint8_t* d_pixels);

/* ------------------------------------------------------------------------------------------------------------
 * DCT-domain requantisation (JPEG -> JPEG). */
typedef struct {
    int16_t* coefs[4];          /* caller buffers, one per component, as large as the source component's */
    uint8_t qtables[4][64];     /* the new tables by slot, natural order */
    int present[4];             /* slot holds a table */
    uint8_t tq[4];              /* each component's new slot */
} rcg_requantized;

/* Replaces rc_jpeg::encode::CoefImage::requantize(quality, trellis) (jpeg/src/encode.rs:111): every component moves
 * to the quality's table, never finer than its old one (so q100 changes nothing), by round_f32(c * old/new);
 * with trellis and a changed table the blocks are re-quantised by the trellis from c * old/new. Host; 1-4
 * components. Bit-exact with the reference. */
int rcg_requantize(const rcg_coef_image* ci, int quality, int trellis, rcg_requantized* out);

/* The JPEG -> JPEG target search of cli/src/pipeline.rs:369-386: reference = reconstruct(ci), make(q) =
 * requantize(ci, q, trellis), metric = SSIMULACRA2, bisection as rcg_find_quality. The source coefficients are
 * uploaded once; the reference image and its SSIMULACRA2 state are built once; each step requantises and
 * reconstructs on the device (the requantisation fused into the IDCT) or, with trellis, requantises on the host.
 * winner (may be NULL) receives requantize(ci, chosen quality, trellis). Any coefficient image reconstruct takes
 * (1, 3 or 4 components, integral sampling ratios): YCbCr and gray images with luma at full resolution and ratios 1 or
 * 2 are scored straight from the IDCT's planes; the others (CMYK/YCCK, RGB-coded, other ratios) through reconstruct's
 * colour stage. */
int rcg_requant_find_quality(rcg_ctx* ctx, const rcg_coef_image* ci, double target, int max_quality, int trellis,
                             rcg_found* out, rcg_requantized* winner);

/* Device bytes that one search call takes from its context (the arena it sizes, descriptors included), for
 * admission against an HBM budget. mode: RCG_SEARCH_PLAIN (rcg_find_quality), RCG_SEARCH_TRELLIS
 * (rcg_find_quality_trellis), RCG_SEARCH_REQUANT (rcg_requant_find_quality over a source of that size with
 * `channels` components, 1, 3 or 4; an upper bound over the sampling layouts). No counterpart in the reference: it feeds the CLI's replacement of
 * cli/src/budget.rs:117-119 (estimate_bytes) with device footprints. */
typedef enum { RCG_SEARCH_PLAIN = 0, RCG_SEARCH_TRELLIS = 1, RCG_SEARCH_REQUANT = 2 } rcg_search_mode;
int rcg_search_bytes(rcg_ctx* ctx, uint32_t width, uint32_t height, uint32_t channels, int subsampling, int mode,
                     size_t* out);

/* ------------------------------------------------------------------------------------------------------------
 * Reading a JPEG source on the host (no device): what the CLI needs to route a JPEG -> JPEG job and to feed the
 * requantisation search. */
typedef struct {
    uint32_t width, height, ncomp;  /* ncomp 1, 3 or 4 */
    uint8_t id[4], h[4], v[4], tq[4];
    uint32_t blocks_w[4], blocks_h[4];  /* MCU-padded block grid of each component */
    uint8_t qtables[4][64];         /* natural order: each slot's table as its components use it (clamped to 255) */
    int present[4];
    int adobe_transform;            /* -1: no Adobe APP14 marker */
    int source_quality;             /* IJG quality estimate of the luma table (jpeg/src/quality.rs:17-22) */
    int coded_rgb;                  /* three components stored as RGB (decode.rs:481-482) */
    int progressive, arithmetic, lossless, hierarchical, precision;
    int orientation;                /* EXIF orientation 1..8, 1 when absent (decode.rs:572-594) */
    int table16;                    /* a used table has an entry above 255 or equal to 0 */
} rcg_jpeg_info;

/* Replaces rc_jpeg::decode::parse(data)?.info() (jpeg/src/decode.rs:397-489) plus exif_orientation (:572): the
 * frame facts, walking the markers without decoding the scans. For a hierarchical file only the DHP facts are
 * filled. RCG_INVALID for a malformed file, RCG_UNSUPPORTED for a precision or component count the reference
 * rejects. */
int rcg_jpeg_parse(const uint8_t* data, size_t len, rcg_jpeg_info* info);
/* Replaces Parsed::into_coefs (decode.rs:529-569) for sequential and progressive Huffman-coded 8-bit DCT frames:
 * coefs[c] (blocks_w*blocks_h*64 int16, zigzag order per block) receive component c's quantised coefficients,
 * decoded as the reference's decoder does (T.81 Annex F/G). RCG_UNSUPPORTED for arithmetic-coded, lossless,
 * hierarchical and 12-bit frames; RCG_INVALID for corrupt entropy-coded data. */
int rcg_jpeg_read_coefs(const uint8_t* data, size_t len, int16_t* coefs[4]);

/* ------------------------------------------------------------------------------------------------------------
 * Entropy coding on the host (no device, no context): the stage after the search, which serialises the chosen
 * coefficients. Byte-identical to the reference's writers. ci points to HOST coefficient buffers; 1-4 components.
 * Descriptive metadata (EXIF, XMP, ICC) is not carried by rcg_coef_image and is not written.
 */
typedef enum {
    RCG_WRITE_BASELINE = 0,     /* rc_jpeg::encode::write_jpeg (jpeg/src/encode.rs:595): baseline, optimised tables */
    RCG_WRITE_PROGRESSIVE = 1,  /* rc_jpeg::progressive::write_progressive (jpeg/src/progressive.rs:386) */
    RCG_WRITE_BEST = 2          /* rc_jpeg::encode::write_best (jpeg/src/encode.rs:543): the smaller estimate */
} rcg_write_mode;

/* *out receives a buffer of *len bytes allocated by the library; release it with rcg_free. */
int rcg_write_jpeg(const rcg_coef_image* ci, int mode, uint8_t** out, size_t* len);
void rcg_free(void* p);

/* Replaces rc_core::util::huffman::build_lengths (core/src/util/huffman.rs:10): length-limited Huffman code
 * lengths of n symbol frequencies (0 for unused symbols; a single used symbol gets length 1). */
int rcg_huffman_lengths(const uint32_t* freqs, size_t n, int max_len, uint8_t* lengths);

#ifdef __cplusplus
}
#endif
#endif /* RCGPU_H */
