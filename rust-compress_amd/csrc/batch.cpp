// Batched candidate scoring (rcg_score_batch[_dev]) and the per-kernel event profiler.
//
// A batch is many images of uneven size, each with a list of candidate (quality, subsampling) pairs. The whole
// batch is scored by a fixed sequence of launches whose workgroups are spread over every image at once:
//   k_encode        one launch; per image one job per distinct sampling layout, whose candidates share the
//                   colour conversion and FDCT and differ only in quantisation + IDCT;
//   k_ssim_ref      per scale, one launch over the images whose reference state is not built yet: the reference
//                   image's cache and next pyramid level; these run on a second stream, beside k_encode;
//   k_ssim_scale    per scale, one launch covering every (image, candidate) pair that has that scale, reading the
//                   reference side from the cache;
//   k_ssim_finalize one launch pooling every pair.
// Pairs whose FP32 score falls within the guard band of the image's target are re-scored with the FP64 kernel
// (reading the already-built pyramids), then the selection of jpeg/src/target.rs:37-55 is replayed per image.
#include <hip/hip_runtime.h>

#include <memory>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "rcg_host.h"

using namespace rcg;

// ---------------------------------------------------------------------------------------------------------------
// Profiler

hipEvent_t Prof::get() {
    if (!free_events.empty()) {
        hipEvent_t e = free_events.back();
        free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void Prof::collect() {
    for (const Pending& p : pending) {
        float ms = 0.0f;
        if (p.a && p.b && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto& t = totals[p.name];
            t.first += ms;
            t.second += 1;
        }
        if (p.a) free_events.push_back(p.a);
        if (p.b) free_events.push_back(p.b);
    }
    pending.clear();
}

ProfScope::ProfScope(rcg_ctx* c, const char* n, hipStream_t s) : ctx(c), name(n), stream(s ? s : c->stream) {
    if (!ctx->prof.on || ctx->prof_suspended) return;
    hipEvent_t a = ctx->prof.get();
    b = ctx->prof.get();
    if (a) (void)hipEventRecord(a, stream);
    ctx->prof.pending.push_back(Prof::Pending{name, a, b});
}

ProfScope::~ProfScope() {
    if (ctx->prof.on && b) (void)hipEventRecord(b, stream);
}

namespace rcg {

int fork_side(rcg_ctx* ctx) {
    RCG_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
    RCG_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
    return RCG_OK;
}

int join_side(rcg_ctx* ctx) {
    RCG_HIP(hipEventRecord(ctx->ev_join, ctx->side));
    RCG_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    return RCG_OK;
}

int mark_side(rcg_ctx* ctx) {
    RCG_HIP(hipEventRecord(ctx->ev_mid, ctx->side));
    return RCG_OK;
}

int wait_side_mark(rcg_ctx* ctx) {
    RCG_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_mid, 0));
    return RCG_OK;
}

}  // namespace rcg

extern "C" int rcg_profile(rcg_ctx* ctx, int enable) {
    return rcg::guarded([&]() -> int {
        if (!ctx) return RCG_INVALID;
        std::lock_guard<std::mutex> lock(ctx->mu);
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        ctx->prof.collect();
        if (enable) ctx->prof.totals.clear();
        ctx->prof.on = enable != 0;
        return RCG_OK;
    });
}

extern "C" int rcg_profile_get(rcg_ctx* ctx, const char* kernel, double* total_ms, int* launches) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !kernel) return RCG_INVALID;
        std::lock_guard<std::mutex> lock(ctx->mu);
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        ctx->prof.collect();
        auto it = ctx->prof.totals.find(kernel);
        if (total_ms) *total_ms = it == ctx->prof.totals.end() ? 0.0 : it->second.first;
        if (launches) *launches = it == ctx->prof.totals.end() ? 0 : it->second.second;
        return RCG_OK;
    });
}

// ---------------------------------------------------------------------------------------------------------------
// Batch core

namespace rcg {

// Linear-light planes of scales 1.. of one image.
static size_t pyramid_bytes(const PairPlan& P) {
    size_t b = 0;
    for (uint32_t s = 1; s < P.nscales; s++) b += Arena::align(size_t(P.sw[s]) * P.sh[s] * 3 * sizeof(float));
    return b;
}

size_t ref_state_bytes(const PairPlan& P) {
    size_t b = pyramid_bytes(P);
    for (uint32_t s = 0; s < P.nscales; s++) b += Arena::align(size_t(P.sw[s]) * P.sh[s] * 9 * sizeof(float));
    return b;
}

static size_t partial_bytes(const PairPlan& P) {
    size_t b = 0;
    for (uint32_t s = 0; s < P.nscales; s++) b += Arena::align(size_t(P.tiles[s]) * 18 * sizeof(double));
    return b;
}

void take_ref_state(Arena& A, const PairPlan& P, RefState& R) {
    for (int s = 0; s < 6; s++) R.planes[s] = R.cache[s] = nullptr;
    for (uint32_t s = 1; s < P.nscales; s++)
        R.planes[s] = static_cast<float*>(A.take(size_t(P.sw[s]) * P.sh[s] * 3 * sizeof(float)));
    for (uint32_t s = 0; s < P.nscales; s++)
        R.cache[s] = static_cast<float*>(A.take(size_t(P.sw[s]) * P.sh[s] * 9 * sizeof(float)));
}

RefJob make_ref_job(const PairPlan& P, uint32_t s, const SsimSrc& a0, const RefState& R) {
    RefJob j;
    std::memset(&j, 0, sizeof(j));
    j.w = P.sw[s];
    j.h = P.sh[s];
    j.th = P.th_ref[s];
    j.tiles_x = P.tiles_x_ref[s];
    j.tiles_y = ssim_tiles_y(P.sh[s], int(P.th_ref[s]));
    j.a = s == 0 ? a0 : planes_src(R.planes[s]);
    j.a_next = s + 1 < P.nscales ? R.planes[s + 1] : nullptr;
    j.cache = R.cache[s];
    return j;
}

std::vector<Unit> xcd_units(const std::vector<UnitGroup>& groups) {
    // Two passes, so that each unit is written once, straight into its slot: the first assigns the chunks to XCDs
    // from the running unit counts, the second places the units. (Building per-XCD lists and interleaving them took
    // ~100 us of host time per call for the ~15,000 units of one 4096x4096 image with six candidates.)
    uint32_t count[kXcds] = {};
    std::vector<uint8_t> xcd_of;
    size_t nchunks = 0;
    for (const UnitGroup& g : groups) {
        const uint32_t chunk = std::max(1u, g.chunk);
        nchunks += (g.tiles + chunk - 1) / chunk;
    }
    xcd_of.reserve(nchunks);
    for (const UnitGroup& g : groups) {
        const uint32_t chunk = std::max(1u, g.chunk);
        for (uint32_t t0 = 0; t0 < g.tiles; t0 += chunk) {
            // The lowest count, the lowest index among equals, as selects rather than branches.
            uint32_t best = count[0];
            int x = 0;
            for (int i = 1; i < kXcds; i++) {
                const bool lt = count[i] < best;
                best = lt ? count[i] : best;
                x = lt ? i : x;
            }
            xcd_of.push_back(uint8_t(x));
            count[x] += (std::min(g.tiles, t0 + chunk) - t0) * g.njobs;
        }
    }
    size_t len = 0;
    for (uint32_t c : count) len = std::max(len, size_t(c));
    std::vector<Unit> out(len * kXcds, Unit{kIdleJob, 0});
    size_t pos[kXcds] = {};
    size_t k = 0;
    for (const UnitGroup& g : groups) {
        const uint32_t chunk = std::max(1u, g.chunk);
        for (uint32_t t0 = 0; t0 < g.tiles; t0 += chunk) {
            const int x = xcd_of[k++];
            const uint32_t t1 = std::min(g.tiles, t0 + chunk);
            Unit* o = out.data() + pos[x] * kXcds + x;
            for (uint32_t t = t0; t < t1; t++)
                for (uint32_t c = 0; c < g.njobs; c++, o += kXcds) *o = Unit{g.job0 + c, t};
            pos[x] += size_t(t1 - t0) * g.njobs;
        }
    }
    return out;
}

namespace {

constexpr double kGuard = 0.05;

// xcd_units through a per-thread memo of recent tables, keyed by the groups. Callers repeat launch geometries: the
// same image and candidate count call after call, every step of a bisection over one image, every tick of a served
// request. A repeated launch then costs a lookup instead of a rebuild (about 30-100 us of host time per call for the
// ~15,000 units of one 4096x4096 image with six candidates). Tables larger than kMemoUnits are built every time.
// The returned reference is valid until the thread's next call.
const std::vector<Unit>& xcd_units_memo(const std::vector<UnitGroup>& groups) {
    struct Entry {
        std::vector<UnitGroup> groups;
        std::vector<Unit> units;
    };
    constexpr size_t kMemoEntries = 32, kMemoUnits = size_t(1) << 16;  // at most 16 MiB of tables per thread
    static_assert(sizeof(UnitGroup) == 4 * sizeof(uint32_t), "groups are compared bytewise");
    thread_local std::vector<Entry> memo;
    thread_local size_t next = 0;
    thread_local std::vector<Unit> scratch;
    for (const Entry& e : memo)
        if (e.groups.size() == groups.size() &&
            std::memcmp(e.groups.data(), groups.data(), groups.size() * sizeof(UnitGroup)) == 0)
            return e.units;
    std::vector<Unit> units = xcd_units(groups);
    if (units.size() > kMemoUnits) {
        scratch = std::move(units);
        return scratch;
    }
    if (memo.size() < kMemoEntries) {
        memo.push_back(Entry{groups, std::move(units)});
        return memo.back().units;
    }
    Entry& e = memo[next];
    next = (next + 1) % kMemoEntries;
    e.groups = groups;
    e.units = std::move(units);
    return e.units;
}

// Transient per-item device buffers of one scoring pass.
struct ItemBufs {
    PairPlan P32, P64;
    Layout L[RCG_MAX_CAND];
    uint8_t* rec[RCG_MAX_CAND][3] = {};
    float* dist_planes[RCG_MAX_CAND][6] = {};
    double* part32[RCG_MAX_CAND][6] = {};
    double* part64[RCG_MAX_CAND][6] = {};
};

struct PairRef {
    int item, cand;
};

size_t desc_bytes(rcg_ctx* ctx, const Item* items, int n) {
    // Fixed part: alignment of every put and the idle units that pad each launch's unit table (at most
    // kXcds * RCG_MAX_CAND per launch, 30 launches per pass: per scale three candidate and two reference launches).
    size_t d = (1 << 16) + 2 * 32 * kXcds * RCG_MAX_CAND * sizeof(Unit);
    for (int i = 0; i < n; i++) {
        const Item& it = items[i];
        d += sizeof(EncJob) * RCG_MAX_CAND + sizeof(EncCand) * it.ncand + 64 * (RCG_MAX_CAND + 2);
        // FP32 pass plus, after it, an FP64 settling pass of up to every pair.
        d += 2 * ((sizeof(SsimJob) + 64) * 12 * it.ncand + sizeof(FinalDesc) * it.ncand + 256);  // main + tail jobs
        d += (sizeof(RefJob) + 64) * 6;
        d += 4 * 6 * 128;
        const PairPlan P32 = plan_pair(it.w, it.h, ctx->tune32);
        const PairPlan P64 = plan_pair(it.w, it.h, ctx->tune64);
        size_t units = 0;
        // A reference launch keeps tile rows together on one XCD: its idle padding is at most kXcds - 1 tile rows.
        for (uint32_t s = 0; s < P32.nscales; s++)
            units += size_t(P32.tiles[s]) * it.ncand + P32.tiles_ref[s] + size_t(kXcds) * P32.tiles_x_ref[s];
        for (uint32_t s = 0; s < P64.nscales; s++) units += size_t(P64.tiles[s]) * it.ncand;
        d += units * sizeof(Unit);
    }
    return d;
}

// Staged SSIMULACRA2 launches of a set of pairs: per (scale, thread count) one reference-cache launch (FP32, for the
// items whose reference state is not built) and one scale launch; then one pooling launch.
struct StagedLaunch {
    bool ref;
    int scale;
    const void* jobs;
    const Unit* units;
    uint32_t nunits;
    int nt;
    uint32_t ka, kb;
};
struct StagedPairs {
    std::vector<StagedLaunch> refs;    // reference-cache launches, scale by scale
    std::vector<StagedLaunch> scales;  // candidate launches, scale by scale
    const FinalDesc* fin = nullptr;
    int npairs = 0;
    int precision = RCG_PREC_FP32;
};

// Stages the descriptors to score every pair in `pairs` at `precision` into D. FP32 first builds the reference state
// of every item that lacks it, then writes each candidate's pyramid; FP64 reads both pyramids from the FP32 pass.
// The pairs of one item must be consecutive in `pairs` (they form one unit group per scale). `parts` selects what
// is staged: kStageRefs the reference launches, kStageScales the candidate and pooling launches.
constexpr int kStageRefs = 1, kStageScales = 2, kStageAll = 3;
StagedPairs stage_pairs(Desc& D, const Item* items, int nitems, std::vector<ItemBufs>& bufs,
                        const std::vector<PairRef>& pairs, int precision, double* d_scores, int parts = kStageAll) {
    const bool f64 = precision == RCG_PREC_FP64;
    StagedPairs out;
    out.precision = precision;
    out.npairs = int(pairs.size());
    std::vector<FinalDesc> fins(pairs.size());
    for (size_t i = 0; i < pairs.size(); i++) std::memset(&fins[i], 0, sizeof(FinalDesc));
    // Items whose reference state this pass builds: those with a pair here and no state yet.
    std::vector<char> build(nitems, 0);
    if (!f64)
        for (const PairRef& pr : pairs) build[pr.item] = !items[pr.item].ref_ready;
    for (int s = 0; s < 6; s++) {
        for (int nt : {256, 128, 64}) {
            if (!f64 && (parts & kStageRefs)) {
                std::vector<RefJob> rj;
                std::vector<UnitGroup> groups;
                for (int i = 0; i < nitems; i++) {
                    if (!build[i]) continue;
                    const Item& it = items[i];
                    const PairPlan& P = bufs[i].P32;
                    if (uint32_t(s) >= P.nscales || int(P.nt_ref[s]) != nt) continue;
                    // A tile row on one XCD, so that neighbouring tiles write the two halves of the cache lines at
                    // their seams (a tile row segment is 976 bytes) through one L2. Measured on one 4096x4096 image
                    // with six candidates: the whole step 2.3% faster.
                    groups.push_back(UnitGroup{uint32_t(rj.size()), 1, P.tiles_ref[s], P.tiles_x_ref[s]});
                    rj.push_back(make_ref_job(P, uint32_t(s), bytes_src(it.d_src, it.ch), it.ref));
                }
                if (!rj.empty()) {
                    const std::vector<Unit>& units = xcd_units_memo(groups);
                    StagedLaunch g;
                    g.ref = true;
                    g.scale = s;
                    g.jobs = D.put(rj.data(), rj.size());
                    g.units = D.put(units.data(), units.size());
                    g.nunits = uint32_t(units.size());
                    g.nt = nt;
                    g.ka = s == 0 ? uint32_t(kSrcBytes) : uint32_t(kSrcPlanes);
                    g.kb = 0;
                    out.refs.push_back(g);
                }
            }
            if (!(parts & kStageScales)) continue;
            std::vector<SsimJob> sj;
            std::vector<UnitGroup> groups;
            for (size_t i = 0; i < pairs.size(); i++) {
                const Item& it = items[pairs[i].item];
                ItemBufs& b = bufs[pairs[i].item];
                const int c = pairs[i].cand;
                const PairPlan& P = f64 ? b.P64 : b.P32;
                if (uint32_t(s) >= P.nscales) continue;
                // This launch holds the pair's main tiles (nt threads) or its tail tiles (tnt threads), if any.
                const bool main = int(P.nt[s]) == nt && P.tiles_x[s] > 0;
                const bool tail = int(P.tnt[s]) == nt;
                if (!main && !tail) continue;
                SsimJob j;
                std::memset(&j, 0, sizeof(j));
                j.w = P.sw[s];
                j.h = P.sh[s];
                j.th = P.th[s];
                j.tiles_x = main ? P.tiles_x[s] : 1;
                j.x_begin = main ? 0 : P.x_tail(s);
                j.tiles_y = P.tiles_y[s];
                if (!f64)
                    j.a = cache_src(it.ref.cache[s]);
                else
                    j.a = s == 0 ? bytes_src(it.d_src, it.ch) : planes_src(it.ref.planes[s]);
                j.b = s == 0 ? recon_src(b.L[c], b.rec[c]) : planes_src(b.dist_planes[c][s]);
                if (!f64 && uint32_t(s + 1) < P.nscales) j.b_next = b.dist_planes[c][s + 1];
                double* const part = f64 ? b.part64[c][s] : b.part32[c][s];
                j.partials = main ? part : part + size_t(P.tiles_main(s)) * 18;  // tail tiles after the main ones
                FinalDesc& f = fins[i];
                f.nscales = P.nscales;
                f.ntiles[s] = P.tiles[s];
                f.partials[s] = part;
                f.count[s] = double(size_t(P.sw[s]) * P.sh[s]);
                // The FP32 candidates of one item share the item's cache: one group, so they run side by side on one
                // XCD. FP64 jobs read no shared data and stay single.
                const bool same_item = !f64 && !groups.empty() && i > 0 && pairs[i - 1].item == pairs[i].item &&
                                       groups.back().job0 + groups.back().njobs == sj.size();
                if (same_item)
                    groups.back().njobs++;
                else
                    groups.push_back(UnitGroup{uint32_t(sj.size()), 1, main ? P.tiles_main(s) : P.tiles_y[s]});
                sj.push_back(j);
            }
            if (!sj.empty()) {
                const std::vector<Unit>& units = xcd_units_memo(groups);
                StagedLaunch g;
                g.ref = false;
                g.scale = s;
                g.jobs = D.put(sj.data(), sj.size());
                g.units = D.put(units.data(), units.size());
                g.nunits = uint32_t(units.size());
                g.nt = nt;
                g.ka = !f64 ? uint32_t(kSrcRefCache) : (s == 0 ? uint32_t(kSrcBytes) : uint32_t(kSrcPlanes));
                g.kb = s == 0 ? uint32_t(kSrcRecon) : uint32_t(kSrcPlanes);
                out.scales.push_back(g);
            }
        }
    }
    if (!(parts & kStageScales)) return out;
    for (size_t i = 0; i < pairs.size(); i++) {
        fins[i].score_out = d_scores + i;
        fins[i].stats_out = nullptr;
    }
    out.fin = D.put(fins.data(), fins.size());
    return out;
}

// The reference-cache launches go to the side stream, which the caller has forked from the context's stream: the
// scale-0 ones, then ev_mid (mark_side), then the later scales. The caller enqueues work that does not need them (the
// encode) on the context's stream, then launch_pairs() with `refs` set.
int launch_refs_side(rcg_ctx* ctx, const StagedPairs& sp) {
    size_t k = 0;
    for (; k < sp.refs.size() && sp.refs[k].scale == 0; k++) {
        const StagedLaunch& g = sp.refs[k];
        RCG_TRY(launch_ssim_ref(ctx, static_cast<const RefJob*>(g.jobs), g.units, g.nunits, g.nt, g.ka, ctx->side));
    }
    RCG_TRY(mark_side(ctx));
    for (; k < sp.refs.size(); k++) {
        const StagedLaunch& g = sp.refs[k];
        RCG_TRY(launch_ssim_ref(ctx, static_cast<const RefJob*>(g.jobs), g.units, g.nunits, g.nt, g.ka, ctx->side));
    }
    return RCG_OK;
}

// The candidate launches, scale by scale, then the pooling launch. With `refs` (launch_refs_side() enqueued this
// pass's reference state), the scale-0 launches wait for the scale-0 reference caches only and the later ones for
// the whole side stream, so the later reference scales are built while scale 0 is scored.
//
// A scale with launches of several thread counts runs the widest on the context's stream and the narrower ones on
// the side stream, beside it: they hold the tail tiles of wide images and the images narrow at that scale, a small
// share of the work whose tiles walk many rows one after another, so alone they would leave the GPU mostly idle for
// their duration. The two streams join before the next scale (an image's pyramid level is written by its main and
// tail tiles, and an image can move between launches from one scale to the next).
int launch_pairs(rcg_ctx* ctx, const StagedPairs& sp, bool refs) {
    auto launch = [&](const StagedLaunch& g, hipStream_t st) {
        return launch_ssim_scale(ctx, sp.precision, static_cast<const SsimJob*>(g.jobs), g.units, g.nunits, g.nt,
                                 g.ka, g.kb, st);
    };
    bool side_open = refs;  // the side stream holds work the context's stream has not joined yet
    size_t k = 0;
    while (k < sp.scales.size()) {
        const int s = sp.scales[k].scale;
        size_t e = k;
        while (e < sp.scales.size() && sp.scales[e].scale == s) e++;
        // [k, e): this scale's launches, widest first. The widest runs on the context's stream, the narrower ones
        // (the images tiled narrower at this scale, and the tail tiles of the others) on the side stream.
        if (s > 0 && side_open) {
            RCG_TRY(join_side(ctx));  // the previous scale's narrow launches (at scale 1 also the reference chain)
            side_open = false;
        }
        if (e - k > 1) {
            // The side stream runs them after everything enqueued here so far (the encode, earlier scales) and, at
            // scale 0, after the reference chain it already holds.
            RCG_TRY(fork_side(ctx));
            for (size_t i = k + 1; i < e; i++) RCG_TRY(launch(sp.scales[i], ctx->side));
            side_open = true;
        }
        if (s == 0 && refs) RCG_TRY(wait_side_mark(ctx));
        RCG_TRY(launch(sp.scales[k], nullptr));
        k = e;
    }
    if (side_open) RCG_TRY(join_side(ctx));
    return launch_finalize(ctx, sp.fin, sp.npairs);
}

}  // namespace

size_t score_items_bytes(rcg_ctx* ctx, const Item* items, int n) {
    size_t need = 0, npairs = 0;
    for (int i = 0; i < n; i++) {
        const Item& it = items[i];
        const PairPlan P32 = plan_pair(it.w, it.h, ctx->tune32);
        const PairPlan P64 = plan_pair(it.w, it.h, ctx->tune64);
        for (uint32_t c = 0; c < it.ncand; c++) {
            const Layout L = make_layout(it.w, it.h, it.ch, it.quality[c], it.subsampling[c]);
            for (uint32_t k = 0; k < L.ncomp; k++) need += Arena::align(L.plane_bytes(k));
            need += pyramid_bytes(P32) + partial_bytes(P32) + partial_bytes(P64);
        }
        npairs += it.ncand;
    }
    return need + Arena::align(desc_bytes(ctx, items, n)) + 2 * Arena::align(npairs * sizeof(double) + 8) + (1 << 12);
}

// The device state of one scoring pass: per-item buffers, staged descriptors (already uploaded) and the FP32
// launch sequence. prepare_items() does all allocation and uploading; enqueue_items() only enqueues kernels and one
// device-to-host copy, so a stream capture of it replays as a graph; finish_items() reads the scores back and
// settles the ones near a target with the FP64 kernel.
struct ScoreState {
    std::vector<ItemBufs> bufs;
    std::vector<PairRef> pairs;
    const EncJob* d_ej = nullptr;
    const uint32_t* d_eb = nullptr;
    int nej = 0;
    uint32_t etiles = 0;
    StagedPairs fp32;
    double* d_s32 = nullptr;
    double* d_s64 = nullptr;
    double* h_scores = nullptr;
    Desc D;
    size_t fp32_desc_end = 0;
};

// The reference-cache kernel's tiling, per launch. Its output does not depend on the tiling, and each image's own plan
// fills one round of the GPU's workgroup slots with that image's tiles, so a launch over many images would walk tiles
// of a few rows, mostly halo (10 rows per tile). Over the images a launch builds together, tiles are as tall as one
// round of slots for all their tile columns allows, at most kRefMaxRows, and never shorter than the image's own plan
// (a single image keeps it). Only lowers tile counts, so the descriptor sizing from the per-image plans still holds.
constexpr uint32_t kRefMaxRows = 256;

void retile_refs(const SsimTune& T, const Item* items, int n, std::vector<ItemBufs>& bufs) {
    for (uint32_t s = 0; s < 6; s++) {
        for (uint32_t nt : {256u, 128u}) {
            auto builds = [&](int i) {
                const PairPlan& P = bufs[i].P32;
                return !items[i].ref_ready && items[i].ncand > 0 && s < P.nscales && P.nt_ref[s] == nt;
            };
            uint32_t cols = 0;
            for (int i = 0; i < n; i++)
                if (builds(i)) cols += bufs[i].P32.tiles_x_ref[s];
            if (cols == 0) continue;
            const uint32_t slots = nt == 128 ? T.ref_slots_narrow : T.ref_slots_wide;
            const uint32_t rows_of_tiles = std::max(1u, slots / cols);
            for (int i = 0; i < n; i++) {
                if (!builds(i)) continue;
                PairPlan& P = bufs[i].P32;
                uint32_t tr = (P.sh[s] + rows_of_tiles - 1) / rows_of_tiles;
                tr = std::min(kRefMaxRows, std::max(4u, (tr + 1) & ~1u));
                if (tr > P.th_ref[s]) {
                    P.th_ref[s] = tr;
                    P.tiles_ref[s] = P.tiles_x_ref[s] * ssim_tiles_y(P.sh[s], int(tr));
                }
            }
        }
    }
}

// The encode descriptors: arena buffers of every pair, then each item's encode jobs, staged at the start of S.D.
int prepare_encode(rcg_ctx* ctx, Arena& A, Pinned& pin, Item* items, int n, ScoreState& S) {
    size_t npairs = 0;
    for (int i = 0; i < n; i++) npairs += items[i].ncand;
    const size_t dbytes = desc_bytes(ctx, items, n);
    RCG_TRY(pin.ensure(Arena::align(dbytes) + Arena::align(npairs * sizeof(double) + 8)));
    S.D.host = pin.base;
    S.D.dev = static_cast<char*>(A.take(dbytes));
    S.D.cap = dbytes;
    S.D.off = 0;
    S.d_s32 = static_cast<double*>(A.take(npairs * sizeof(double) + 8));
    S.d_s64 = static_cast<double*>(A.take(npairs * sizeof(double) + 8));
    S.h_scores = reinterpret_cast<double*>(pin.base + Arena::align(dbytes));
    S.bufs.assign(n, ItemBufs());
    S.pairs.clear();
    std::vector<EncJob> ej;
    std::vector<std::vector<EncCand>> ecands;
    for (int i = 0; i < n; i++) {
        Item& it = items[i];
        ItemBufs& b = S.bufs[i];
        b.P32 = plan_pair(it.w, it.h, ctx->tune32);
        b.P64 = plan_pair(it.w, it.h, ctx->tune64);
        const int first_group = int(ej.size());
        for (uint32_t c = 0; c < it.ncand; c++) {
            it.scores[c] = -DBL_MAX;
            it.settled[c] = 0;
            b.L[c] = make_layout(it.w, it.h, it.ch, it.quality[c], it.subsampling[c]);
            for (uint32_t k = 0; k < b.L[c].ncomp; k++)
                b.rec[c][k] = static_cast<uint8_t*>(A.take(b.L[c].plane_bytes(k)));
            for (uint32_t s = 1; s < b.P32.nscales; s++)
                b.dist_planes[c][s] =
                    static_cast<float*>(A.take(size_t(b.P32.sw[s]) * b.P32.sh[s] * 3 * sizeof(float)));
            for (uint32_t s = 0; s < b.P32.nscales; s++)
                b.part32[c][s] = static_cast<double*>(A.take(size_t(b.P32.tiles[s]) * 18 * 8));
            for (uint32_t s = 0; s < b.P64.nscales; s++)
                b.part64[c][s] = static_cast<double*>(A.take(size_t(b.P64.tiles[s]) * 18 * 8));
            S.pairs.push_back(PairRef{i, int(c)});
            // Candidates with the same sampling layout share one encode job (colour + FDCT once).
            int g = -1;
            for (int e = first_group; e < int(ej.size()); e++)
                if (ej[e].hmax == b.L[c].hmax && ej[e].vmax == b.L[c].vmax && ej[e].ncand < uint32_t(kMaxCand)) g = e;
            if (g < 0) {
                EncJob e;
                std::memset(&e, 0, sizeof(e));
                e.src = it.d_src;
                e.src_bytes = uint64_t(it.w) * it.h * it.ch;
                e.w = it.w;
                e.h = it.h;
                e.channels = it.ch;
                e.hmax = b.L[c].hmax;
                e.vmax = b.L[c].vmax;
                e.pw = b.L[c].pw;
                e.ph = b.L[c].ph;
                e.tiles_x = (e.pw + kEncTileW - 1) / kEncTileW;
                e.tiles_y = e.ph / (8 * e.vmax);
                ej.push_back(e);
                ecands.emplace_back();
                g = int(ej.size()) - 1;
            }
            EncCand k;
            std::memset(&k, 0, sizeof(k));
            uint8_t qy[64], qc[64];
            quant_tables(it.quality[c], qy, qc);
            fill_cand_tables(qy, qc, k);
            for (uint32_t q = 0; q < b.L[c].ncomp; q++) k.rec[q] = b.rec[c][q];
            ecands[g].push_back(k);
            ej[g].ncand++;
        }
    }
    std::vector<uint32_t> ebegins(ej.size());
    S.etiles = 0;
    for (size_t e = 0; e < ej.size(); e++) {
        ej[e].cand = S.D.put(ecands[e].data(), ecands[e].size());
        ej[e].tile_begin = S.etiles;
        ebegins[e] = S.etiles;
        S.etiles += ej[e].tiles_x * ej[e].tiles_y;
    }
    S.nej = int(ej.size());
    if (!ej.empty()) {
        S.d_ej = S.D.put(ej.data(), ej.size());
        S.d_eb = S.D.put(ebegins.data(), ebegins.size());
    }
    return RCG_OK;
}

// The FP32 SSIMULACRA2 descriptors, staged after the encode ones.
void prepare_pairs(rcg_ctx* ctx, Item* items, int n, ScoreState& S) {
    retile_refs(ctx->tune32, items, n, S.bufs);
    S.fp32 = stage_pairs(S.D, items, n, S.bufs, S.pairs, RCG_PREC_FP32, S.d_s32);
    S.fp32_desc_end = S.D.off;
}

int prepare_items(rcg_ctx* ctx, Arena& A, Pinned& pin, Item* items, int n, ScoreState& S) {
    RCG_TRY(prepare_encode(ctx, A, pin, items, n, S));
    prepare_pairs(ctx, items, n, S);
    return S.D.upload(ctx->stream);
}

// Everything after the encode launch: the reference launches (on the side stream, forked already when `refs`; unless
// the caller launched them already), the candidate launches, the pooling launch and the copy of the FP32 scores.
int enqueue_after_encode(rcg_ctx* ctx, const ScoreState& S, bool refs, bool launch_refs = true) {
    if (refs && launch_refs) RCG_TRY(launch_refs_side(ctx, S.fp32));
    if (S.pairs.empty()) {
        if (refs) RCG_TRY(join_side(ctx));
    } else {
        RCG_TRY(launch_pairs(ctx, S.fp32, refs));
        RCG_HIP(hipMemcpyAsync(S.h_scores, S.d_s32, S.pairs.size() * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
    }
    return RCG_OK;
}

int enqueue_items(rcg_ctx* ctx, const ScoreState& S) {
    // The reference caches are built on the side stream beside the encode; the encode is enqueued first so that it
    // starts while the host is still enqueueing the reference launches.
    const bool refs = !S.fp32.refs.empty();
    if (refs) RCG_TRY(fork_side(ctx));
    if (S.nej) RCG_TRY(launch_encode(ctx, S.d_ej, S.d_eb, S.nej, S.etiles));
    return enqueue_after_encode(ctx, S, refs);
}

// The stream must be synchronised (the FP32 scores are in S.h_scores).
int finish_items(rcg_ctx* ctx, ScoreState& S, Item* items, int n) {
    std::vector<PairRef> settle;
    for (size_t i = 0; i < S.pairs.size(); i++) {
        const double s = S.h_scores[i];
        Item& it = items[S.pairs[i].item];
        it.scores[S.pairs[i].cand] = std::isfinite(s) ? s : -DBL_MAX;
        it.settled[S.pairs[i].cand] = 0;
        if (std::isfinite(s) && std::fabs(s - it.target) < kGuard) settle.push_back(S.pairs[i]);
    }
    for (int i = 0; i < n; i++) items[i].ref_ready = true;
    if (settle.empty()) return RCG_OK;
    // FP64 descriptors go after the FP32 ones, which stay valid for later replays.
    S.D.off = S.fp32_desc_end;
    const StagedPairs f64 = stage_pairs(S.D, items, n, S.bufs, settle, RCG_PREC_FP64, S.d_s64);
    RCG_TRY(S.D.upload(ctx->stream));
    RCG_TRY(launch_pairs(ctx, f64, false));
    RCG_HIP(hipMemcpyAsync(S.h_scores, S.d_s64, settle.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < settle.size(); i++) {
        const double s = S.h_scores[i];
        Item& it = items[settle[i].item];
        it.scores[settle[i].cand] = std::isfinite(s) ? s : -DBL_MAX;
        it.settled[settle[i].cand] = 1;
    }
    return RCG_OK;
}

// Score every candidate of every item: one encode launch, one launch per scale, one pooling launch; then FP64
// settling of the scores inside the guard band of each item's target. The arena must already hold
// score_items_bytes() beyond what the caller took.
//
// Calls of 2^25 candidate-pixels or more upload their descriptors in two parts. The encode and reference-launch
// descriptors go first, and the encode is launched (and the reference launches enqueued on the side stream, forked
// before the encode so that they do not wait for it) as soon as they are up. The host then stages the candidate
// launches' descriptors (tens of microseconds for one large image) while those run, and uploads them on the
// context's stream behind the encode, which the candidate launches follow.
static int score_items_impl(rcg_ctx* ctx, Item* items, int n) {
    ScoreState S;
    // Below ~2^25 candidate-pixels the staging is short and the second upload costs more than it hides (a plain
    // search over images up to 3000x3000, one candidate per step, measured 3% slower with it): one upload.
    size_t work = 0;
    for (int i = 0; i < n; i++) work += size_t(items[i].w) * items[i].h * items[i].ncand;
    if (work < (size_t(1) << 25)) {
        RCG_TRY(prepare_items(ctx, ctx->arena, ctx->pinned, items, n, S));
        RCG_TRY(enqueue_items(ctx, S));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        return finish_items(ctx, S, items, n);
    }
    RCG_TRY(prepare_encode(ctx, ctx->arena, ctx->pinned, items, n, S));
    retile_refs(ctx->tune32, items, n, S.bufs);
    const StagedPairs refs_sp = stage_pairs(S.D, items, n, S.bufs, S.pairs, RCG_PREC_FP32, S.d_s32, kStageRefs);
    const size_t first_end = S.D.off;
    RCG_TRY(S.D.upload_range(0, first_end, ctx->stream));
    const bool refs = !refs_sp.refs.empty();
    if (refs) RCG_TRY(fork_side(ctx));
    if (S.nej) RCG_TRY(launch_encode(ctx, S.d_ej, S.d_eb, S.nej, S.etiles));
    if (refs) RCG_TRY(launch_refs_side(ctx, refs_sp));
    S.fp32 = stage_pairs(S.D, items, n, S.bufs, S.pairs, RCG_PREC_FP32, S.d_s32, kStageScales);
    S.fp32.refs = refs_sp.refs;
    S.fp32_desc_end = S.D.off;
    RCG_TRY(S.D.upload_range(first_end, S.D.off, ctx->stream));
    RCG_TRY(enqueue_after_encode(ctx, S, refs, false));
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    return finish_items(ctx, S, items, n);
}

// A failure part way leaves work in flight on either stream: it is waited for before the error returns, so that no
// launch of this call still reads or writes the caller's buffers (a server reuses a failed request's blocks).
int score_items(rcg_ctx* ctx, Item* items, int n) {
    int rc;
    try {
        rc = score_items_impl(ctx, items, n);
    } catch (...) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamSynchronize(ctx->stream);
        throw;
    }
    if (rc != RCG_OK) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamSynchronize(ctx->stream);
    }
    return rc;
}

// Coefficient bytes of each component of the largest grid among an item's candidates.
static void winner_grid_bytes(const Item& it, size_t out[3]) {
    out[0] = out[1] = out[2] = 0;
    for (uint32_t c = 0; c < it.ncand; c++) {
        const Layout L = make_layout(it.w, it.h, it.ch, it.quality[c], it.subsampling[c]);
        for (uint32_t k = 0; k < L.ncomp; k++) out[k] = std::max(out[k], L.coef_count(int(k)) * sizeof(int16_t));
    }
}

size_t winner_pinned(int n) {
    return size_t(n) * (sizeof(EncCand) + sizeof(EncJob) + sizeof(uint32_t) + 3 * 64) + 4096;
}

size_t winner_bytes(const Item* items, const int* which, int n, bool host_dest) {
    size_t b = Arena::align(winner_pinned(n));
    if (host_dest)
        for (int k = 0; k < n; k++) {
            size_t g[3];
            winner_grid_bytes(items[which[k]], g);
            for (int c = 0; c < 3; c++) b += Arena::align(g[c]);
        }
    return b;
}

int encode_winners(rcg_ctx* ctx, Arena& A, Pinned& pin, const Item* items, const int* which, const int* chosen,
                   const WinnerDest* dest, int n, bool host_dest) {
    if (n <= 0) return RCG_OK;
    const size_t dbytes = winner_pinned(n);
    RCG_TRY(pin.ensure(dbytes));
    Desc D;
    D.host = pin.base;
    D.dev = static_cast<char*>(A.take(dbytes));
    D.cap = dbytes;
    D.off = 0;
    std::vector<EncJob> ej(static_cast<size_t>(n));
    std::vector<uint32_t> begins(static_cast<size_t>(n));
    std::vector<Layout> lay(static_cast<size_t>(n));
    std::vector<WinnerDest> dev(static_cast<size_t>(n));
    uint32_t tiles = 0;
    for (int k = 0; k < n; k++) {
        const Item& it = items[which[k]];
        const int q = it.quality[chosen[k]];
        lay[k] = make_layout(it.w, it.h, it.ch, q, it.subsampling[chosen[k]]);
        const Layout& L = lay[k];
        for (int c = 0; c < 3; c++) {
            dev[k].coefs[c] = nullptr;
            if (uint32_t(c) >= L.ncomp) continue;
            dev[k].coefs[c] = host_dest ? static_cast<int16_t*>(A.take(L.coef_count(c) * sizeof(int16_t)))
                                        : dest[k].coefs[c];
        }
        EncCand cand;
        std::memset(&cand, 0, sizeof(cand));
        uint8_t qy[64], qc[64];
        quant_tables(q, qy, qc);
        fill_cand_tables(qy, qc, cand);
        for (int c = 0; c < 3; c++) cand.coef[c] = dev[k].coefs[c];
        EncJob& e = ej[k];
        std::memset(&e, 0, sizeof(e));
        e.src = it.d_src;
        e.src_bytes = uint64_t(it.w) * it.h * it.ch;
        e.w = it.w;
        e.h = it.h;
        e.channels = it.ch;
        e.hmax = L.hmax;
        e.vmax = L.vmax;
        e.pw = L.pw;
        e.ph = L.ph;
        e.tiles_x = (e.pw + kEncTileW - 1) / kEncTileW;
        e.tiles_y = e.ph / (8 * e.vmax);
        e.ncand = 1;
        e.cand = D.put(&cand, 1);
        e.tile_begin = tiles;
        begins[k] = tiles;
        tiles += e.tiles_x * e.tiles_y;
    }
    const EncJob* d_ej = D.put(ej.data(), ej.size());
    const uint32_t* d_eb = D.put(begins.data(), begins.size());
    RCG_TRY(D.upload(ctx->stream));
    RCG_TRY(launch_encode(ctx, d_ej, d_eb, n, tiles, true));  // the winners' coefficients only
    if (host_dest)
        for (int k = 0; k < n; k++)
            for (uint32_t c = 0; c < lay[k].ncomp; c++)
                RCG_HIP(hipMemcpyAsync(dest[k].coefs[c], dev[k].coefs[c], lay[k].coef_count(int(c)) * sizeof(int16_t),
                                       hipMemcpyDeviceToHost, ctx->stream));
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    return RCG_OK;
}

// Selection of jpeg/src/target.rs:37-55 over a candidate list sorted by quality: the highest first, then bisection
// over list indices. Returns the chosen index; *sweep receives the exhaustive choice (lowest passing candidate).
int select_candidate(const int* quality, const double* scores, int n, double target, int* sweep) {
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return quality[a] < quality[b]; });
    int lo = 0, hi = n - 1, best = hi;
    if (scores[order[hi]] >= target) {
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (scores[order[mid]] >= target) {
                hi = mid;
                best = mid;
            } else {
                lo = mid + 1;
            }
        }
    }
    int sw = n - 1;
    for (int k = 0; k < n; k++)
        if (scores[order[k]] >= target) {
            sw = k;
            break;
        }
    if (sweep) *sweep = order[sw];
    return order[best];
}

}  // namespace rcg

// ---------------------------------------------------------------------------------------------------------------
// rcg_score_batch

namespace {

// Validate one batch job into an Item; returns its status.
int job_item(const rcg_batch_job& j, bool device_winner, Item& it) {
    int st = check_image(j.width, j.height, j.channels);
    if (st == RCG_OK && (j.width < 8 || j.height < 8)) st = RCG_UNSUPPORTED;
    if (st == RCG_OK && (j.ncand == 0 || j.ncand > RCG_MAX_CAND || !j.pixels)) st = RCG_INVALID;
    if (st == RCG_OK && j.winner[0]) {
        const uint32_t ncomp = j.channels == 1 ? 1 : 3;
        for (uint32_t c = 0; c < ncomp; c++)
            if (!j.winner[c] || (device_winner && (reinterpret_cast<uintptr_t>(j.winner[c]) & 15))) {
                set_error("winner buffers: one per component, 16-byte aligned on the device");
                st = RCG_INVALID;
            }
    }
    if (st != RCG_OK) return st;
    std::memset(&it, 0, sizeof(it));
    it.d_src = j.pixels;
    it.w = j.width;
    it.h = j.height;
    it.ch = j.channels;
    it.ncand = j.ncand;
    for (uint32_t c = 0; c < j.ncand; c++) {
        it.quality[c] = j.quality[c];
        it.subsampling[c] = j.subsampling[c];
    }
    it.target = j.target;
    return RCG_OK;
}

// The jobs (by item index) that asked for their winner's coefficients.
void winner_items(const rcg_batch_job* jobs, const std::vector<int>& job_of, std::vector<int>& which,
                  std::vector<WinnerDest>& dest) {
    which.clear();
    dest.clear();
    for (size_t k = 0; k < job_of.size(); k++) {
        const rcg_batch_job& j = jobs[job_of[k]];
        if (!j.winner[0]) continue;
        which.push_back(int(k));
        WinnerDest d;
        for (int c = 0; c < 3; c++) d.coefs[c] = j.winner[c];
        dest.push_back(d);
    }
}

// Selection per item into the results, then the winners' coefficients for the jobs that asked.
int report_items(rcg_ctx* ctx, Arena& A, Pinned& pin, const Item* items, const std::vector<int>& job_of,
                 const std::vector<int>& which, const std::vector<WinnerDest>& dest, bool host_dest,
                 rcg_batch_result* results) {
    for (size_t k = 0; k < job_of.size(); k++) {
        const Item& it = items[k];
        rcg_batch_result& r = results[job_of[k]];
        for (uint32_t c = 0; c < it.ncand; c++) {
            r.scores[c] = it.scores[c];
            r.settled[c] = it.settled[c];
        }
        r.chosen_index = select_candidate(it.quality, it.scores, int(it.ncand), it.target, &r.sweep_index);
        r.chosen_quality = it.quality[r.chosen_index];
        r.sweep_quality = it.quality[r.sweep_index];
    }
    if (which.empty()) return RCG_OK;
    std::vector<int> chosen(which.size());
    for (size_t k = 0; k < which.size(); k++) chosen[k] = results[job_of[which[k]]].chosen_index;
    RCG_TRY(encode_winners(ctx, A, pin, items, which.data(), chosen.data(), dest.data(), int(which.size()),
                           host_dest));
    for (int k : which) results[job_of[k]].winner_written = 1;
    return RCG_OK;
}

int score_batch_impl(rcg_ctx* ctx, const rcg_batch_job* jobs, int njobs, rcg_batch_result* results, bool host_src) {
    if (!ctx || (!jobs && njobs) || (!results && njobs) || njobs < 0) return RCG_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    RCG_HIP(hipSetDevice(ctx->device));
    std::vector<Item> items;
    std::vector<int> job_of;
    size_t own = 0;
    for (int i = 0; i < njobs; i++) {
        const rcg_batch_job& j = jobs[i];
        rcg_batch_result& r = results[i];
        std::memset(&r, 0, sizeof(r));
        r.chosen_index = r.sweep_index = -1;
        Item it;
        r.status = job_item(j, !host_src, it);
        if (r.status != RCG_OK) continue;
        items.push_back(it);
        job_of.push_back(i);
        if (host_src) own += Arena::align(size_t(j.width) * j.height * j.channels);
        own += ref_state_bytes(plan_pair(j.width, j.height, ctx->tune32));
    }
    std::vector<int> which;
    std::vector<WinnerDest> dest;
    winner_items(jobs, job_of, which, dest);
    const size_t need = own + score_items_bytes(ctx, items.data(), int(items.size())) +
                        winner_bytes(items.data(), which.data(), int(which.size()), host_src);
    ctx->arena.reset();
    RCG_TRY(ctx->arena.ensure(need));
    for (size_t k = 0; k < items.size(); k++) {
        Item& it = items[k];
        if (host_src) {
            const size_t n = size_t(it.w) * it.h * it.ch;
            uint8_t* d = static_cast<uint8_t*>(ctx->arena.take(n));
            RCG_HIP(hipMemcpyAsync(d, jobs[job_of[k]].pixels, n, hipMemcpyHostToDevice, ctx->stream));
            it.d_src = d;
        }
        take_ref_state(ctx->arena, plan_pair(it.w, it.h, ctx->tune32), it.ref);
        it.ref_ready = false;
    }
    RCG_TRY(score_items(ctx, items.data(), int(items.size())));
    return report_items(ctx, ctx->arena, ctx->pinned, items.data(), job_of, which, dest, host_src, results);
}

}  // namespace

extern "C" int rcg_score_batch_dev(rcg_ctx* ctx, const rcg_batch_job* jobs, int njobs, rcg_batch_result* results) {
    return rcg::guarded([&]() -> int {
        return score_batch_impl(ctx, jobs, njobs, results, false);
    });
}

extern "C" int rcg_score_batch(rcg_ctx* ctx, const rcg_batch_job* jobs, int njobs, rcg_batch_result* results) {
    return rcg::guarded([&]() -> int {
        return score_batch_impl(ctx, jobs, njobs, results, true);
    });
}

// ---------------------------------------------------------------------------------------------------------------
// rcg_batch_plan: the batched launch sequence captured once as a hipGraph and replayed per run.

struct rcg_plan {
    rcg_ctx* ctx = nullptr;
    Arena arena;
    Pinned pinned;
    Pinned winner_pinned;       // staging of the winners' encode descriptors (the pass's own stays as uploaded)
    size_t winner_mark = 0;     // arena offset where each run's winner encode takes its descriptors
    std::vector<rcg_batch_job> jobs;
    std::vector<Item> items;
    std::vector<int> job_of;
    std::vector<int> status;
    std::vector<int> which;
    std::vector<WinnerDest> dest;
    ScoreState S;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // The caller has waited for every use of the plan's memory on the device.
    ~rcg_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (arena.base) (void)hipFree(arena.base);
        if (pinned.base) (void)hipHostFree(pinned.base);
        if (winner_pinned.base) (void)hipHostFree(winner_pinned.base);
    }
};

extern "C" void rcg_batch_plan_destroy(rcg_plan* p) {
    rcg::guarded_void([&] {
        if (!p) return;
        {
            std::lock_guard<std::mutex> lock(p->ctx->mu);
            (void)hipSetDevice(p->ctx->device);
            (void)hipStreamSynchronize(p->ctx->stream);
            (void)hipStreamSynchronize(p->ctx->side);
        }
        delete p;
    });
}

extern "C" int rcg_batch_plan_create(rcg_ctx* ctx, const rcg_batch_job* jobs, int njobs, rcg_plan** out) {
    return rcg::guarded([&]() -> int {
        if (!ctx || !out || njobs < 0 || (!jobs && njobs)) return RCG_INVALID;
        *out = nullptr;
        std::unique_ptr<rcg_plan> p(new rcg_plan());
        p->ctx = ctx;
        p->jobs.assign(jobs, jobs + njobs);
        p->status.assign(njobs, RCG_OK);
        size_t own = 0;
        for (int i = 0; i < njobs; i++) {
            Item it;
            p->status[i] = job_item(jobs[i], true, it);
            if (p->status[i] != RCG_OK) continue;
            p->items.push_back(it);
            p->job_of.push_back(i);
            own += ref_state_bytes(plan_pair(jobs[i].width, jobs[i].height, ctx->tune32));
        }
        winner_items(jobs, p->job_of, p->which, p->dest);
        std::lock_guard<std::mutex> lock(ctx->mu);
        RCG_HIP(hipSetDevice(ctx->device));
        // Whatever fails below, nothing of this plan is left running on the device when it is freed.
        struct Drain {
            rcg_ctx* c;
            bool armed = true;
            ~Drain() {
                if (armed) {
                    (void)hipStreamSynchronize(c->stream);
                    (void)hipStreamSynchronize(c->side);
                }
            }
        } drain{ctx};
        RCG_TRY(p->arena.ensure(own + score_items_bytes(ctx, p->items.data(), int(p->items.size())) +
                                winner_bytes(p->items.data(), p->which.data(), int(p->which.size()), false)));
        for (Item& it : p->items) {
            take_ref_state(p->arena, plan_pair(it.w, it.h, ctx->tune32), it.ref);
            it.ref_ready = false;
        }
        RCG_TRY(prepare_items(ctx, p->arena, p->pinned, p->items.data(), int(p->items.size()), p->S));
        p->winner_mark = p->arena.used;
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        // Capture the FP32 launch sequence (kernels and one device-to-host copy; no memsets, no allocation).
        const bool prof = ctx->prof.on;
        ctx->prof.on = false;
        int rc = RCG_OK;
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            rc = enqueue_items(ctx, p->S);
            hipError_t e2 = hipStreamEndCapture(ctx->stream, &p->graph);
            if (rc == RCG_OK && e2 != hipSuccess) {
                set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
                rc = RCG_DEVICE;
            }
        } else {
            set_error(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
            rc = RCG_DEVICE;
        }
        ctx->prof.on = prof;
        RCG_TRY(rc);
        RCG_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
        drain.armed = false;
        *out = p.release();
        return RCG_OK;
    });
}

extern "C" int rcg_batch_plan_run(rcg_plan* p, rcg_batch_result* results) {
    return rcg::guarded([&]() -> int {
        if (!p || (!results && !p->jobs.empty())) return RCG_INVALID;
        rcg_ctx* ctx = p->ctx;
        std::lock_guard<std::mutex> lock(ctx->mu);
        RCG_HIP(hipSetDevice(ctx->device));
        if (p->exec) {
            ProfScope ps(ctx, "batch_graph");
            RCG_HIP(hipGraphLaunch(p->exec, ctx->stream));
        }
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        RCG_TRY(finish_items(ctx, p->S, p->items.data(), int(p->items.size())));
        for (size_t i = 0; i < p->jobs.size(); i++) {
            std::memset(&results[i], 0, sizeof(rcg_batch_result));
            results[i].status = p->status[i];
            results[i].chosen_index = results[i].sweep_index = -1;
        }
        p->arena.used = p->winner_mark;
        return report_items(ctx, p->arena, p->winner_pinned, p->items.data(), p->job_of, p->which, p->dest, false,
                            results);
    });
}
