// Batched candidate scoring (rcg_score_batch[_dev]) and the per-kernel event profiler.
//
// A batch is many images of uneven size, each with a list of candidate (quality, subsampling) pairs. The whole
// batch is scored by a fixed sequence of launches whose workgroups are spread over every image at once:
//   k_encode        one launch; per image one job per distinct sampling layout, whose candidates share the
//                   colour conversion and FDCT and differ only in quantisation + IDCT;
//   k_ssim_scale    one launch per scale (6), covering every (image, candidate) pair that has that scale; the
//                   first candidate of an image also writes the reference pyramid, which every candidate of the
//                   image then reads;
//   k_ssim_finalize one launch pooling every pair.
// Pairs whose FP32 score falls within the guard band of the image's target are re-scored with the FP64 kernel
// (reading the already-built pyramids), then the selection of jpeg/src/target.rs:37-55 is replayed per image.
#include <hip/hip_runtime.h>

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

ProfScope::ProfScope(rcg_ctx* c, const char* n) : ctx(c), name(n) {
    if (!ctx->prof.on) return;
    hipEvent_t a = ctx->prof.get();
    b = ctx->prof.get();
    if (a) (void)hipEventRecord(a, ctx->stream);
    ctx->prof.pending.push_back(Prof::Pending{name, a, b});
}

ProfScope::~ProfScope() {
    if (ctx->prof.on && b) (void)hipEventRecord(b, ctx->stream);
}

extern "C" int rcg_profile(rcg_ctx* ctx, int enable) {
    if (!ctx) return RCG_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->prof.collect();
    if (enable) ctx->prof.totals.clear();
    ctx->prof.on = enable != 0;
    return RCG_OK;
}

extern "C" int rcg_profile_get(rcg_ctx* ctx, const char* kernel, double* total_ms, int* launches) {
    if (!ctx || !kernel) return RCG_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->prof.collect();
    auto it = ctx->prof.totals.find(kernel);
    if (total_ms) *total_ms = it == ctx->prof.totals.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == ctx->prof.totals.end() ? 0 : it->second.second;
    return RCG_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Batch

namespace {

constexpr double kGuard = 0.05;

struct PairRef {
    int job, cand;
};

struct JobPlan {
    bool ok = false;
    const uint8_t* d_src = nullptr;
    PairPlan P32, P64;
    float* ref_planes[6] = {};
    Layout L[RCG_MAX_CAND];
    int group[RCG_MAX_CAND];  // encode job index of each candidate
    uint8_t* rec[RCG_MAX_CAND][3] = {};
    float* dist_planes[RCG_MAX_CAND][6] = {};
    double* part32[RCG_MAX_CAND][6] = {};
    double* part64[RCG_MAX_CAND][6] = {};
};

size_t pyramid_bytes(const PairPlan& P) {
    size_t b = 0;
    for (uint32_t s = 1; s < P.nscales; s++) b += Arena::align(size_t(P.sw[s]) * P.sh[s] * 3 * sizeof(float));
    return b;
}
size_t partial_bytes(const PairPlan& P) {
    size_t b = 0;
    for (uint32_t s = 0; s < P.nscales; s++) b += Arena::align(size_t(P.tiles[s]) * 18 * sizeof(double));
    return b;
}

// Scores every pair of `pairs` at `precision`. For FP32 the pyramids are written (reference by the image's first
// candidate); for FP64 both pyramids are read from the FP32 pass.
int run_pairs(rcg_ctx* ctx, Desc& D, std::vector<JobPlan>& plans, const rcg_batch_job* jobs,
              const std::vector<PairRef>& pairs, int precision, double* d_scores) {
    const bool f64 = precision == RCG_PREC_FP64;
    const SsimTune& T = f64 ? ctx->tune64 : ctx->tune32;
    std::vector<FinalDesc> fins(pairs.size());
    for (size_t i = 0; i < pairs.size(); i++) std::memset(&fins[i], 0, sizeof(FinalDesc));
    struct Staged {
        const SsimJob* jobs;
        const uint32_t* begins;
        int n;
        uint32_t tiles;
    };
    Staged st[6] = {};
    for (int s = 0; s < 6; s++) {
        std::vector<SsimJob> sj;
        std::vector<uint32_t> begins;
        uint32_t total = 0;
        for (size_t i = 0; i < pairs.size(); i++) {
            JobPlan& jp = plans[pairs[i].job];
            const int c = pairs[i].cand;
            const PairPlan& P = f64 ? jp.P64 : jp.P32;
            if (uint32_t(s) >= P.nscales) continue;
            SsimJob j;
            std::memset(&j, 0, sizeof(j));
            j.w = P.sw[s];
            j.h = P.sh[s];
            j.tiles_x = P.tiles_x[s];
            j.tiles_y = P.tiles_y[s];
            j.tile_begin = total;
            const rcg_batch_job& bj = jobs[pairs[i].job];
            if (s == 0) {
                j.a = bytes_src(jp.d_src, bj.channels);
                j.b = recon_src(jp.L[c], jp.rec[c]);
            } else {
                j.a = planes_src(jp.ref_planes[s]);
                j.b = planes_src(jp.dist_planes[c][s]);
            }
            const bool next = uint32_t(s + 1) < P.nscales;
            if (!f64 && next) {
                j.b_next = jp.dist_planes[c][s + 1];
                if (c == 0) j.a_next = jp.ref_planes[s + 1];
            }
            j.partials = f64 ? jp.part64[c][s] : jp.part32[c][s];
            FinalDesc& f = fins[i];
            f.nscales = P.nscales;
            f.ntiles[s] = P.tiles[s];
            f.partials[s] = j.partials;
            f.count[s] = double(size_t(P.sw[s]) * P.sh[s]);
            sj.push_back(j);
            begins.push_back(total);
            total += P.tiles[s];
        }
        if (!sj.empty()) {
            st[s].jobs = D.put(sj.data(), sj.size());
            st[s].begins = D.put(begins.data(), begins.size());
            st[s].n = int(sj.size());
            st[s].tiles = total;
        }
    }
    for (size_t i = 0; i < pairs.size(); i++) {
        fins[i].score_out = d_scores + i;
        fins[i].stats_out = nullptr;
    }
    const FinalDesc* dfin = D.put(fins.data(), fins.size());
    RCG_TRY(D.upload(ctx->stream));
    for (int s = 0; s < 6; s++)
        if (st[s].n) RCG_TRY(launch_ssim_scale(ctx, precision, st[s].jobs, st[s].begins, st[s].n, st[s].tiles, T.nt, T.th));
    return launch_finalize(ctx, dfin, int(pairs.size()));
}

// Selection of jpeg/src/target.rs:37-55 over the candidate list sorted by quality.
void select(const rcg_batch_job& j, rcg_batch_result& r) {
    std::vector<int> order(j.ncand);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return j.quality[a] < j.quality[b]; });
    const int n = int(j.ncand);
    int lo = 0, hi = n - 1;
    int best = hi;
    if (r.scores[order[hi]] >= j.target) {
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (r.scores[order[mid]] >= j.target) {
                hi = mid;
                best = mid;
            } else {
                lo = mid + 1;
            }
        }
    }
    r.chosen_index = order[best];
    r.chosen_quality = j.quality[order[best]];
    int sweep = n - 1;
    for (int k = 0; k < n; k++)
        if (r.scores[order[k]] >= j.target) {
            sweep = k;
            break;
        }
    r.sweep_index = order[sweep];
    r.sweep_quality = j.quality[order[sweep]];
}

int score_batch_impl(rcg_ctx* ctx, const rcg_batch_job* jobs, int njobs, rcg_batch_result* results, bool host_src) {
    if (!ctx || (!jobs && njobs) || (!results && njobs) || njobs < 0) return RCG_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    RCG_HIP(hipSetDevice(ctx->device));
    std::vector<JobPlan> plans(njobs);
    size_t need = 0, desc = 1 << 16;
    size_t npairs = 0;
    for (int i = 0; i < njobs; i++) {
        const rcg_batch_job& j = jobs[i];
        rcg_batch_result& r = results[i];
        std::memset(&r, 0, sizeof(r));
        r.chosen_index = r.sweep_index = -1;
        int st = check_image(j.width, j.height, j.channels);
        if (st == RCG_OK && (j.width < 8 || j.height < 8)) st = RCG_UNSUPPORTED;
        if (st == RCG_OK && (j.ncand == 0 || j.ncand > RCG_MAX_CAND || !j.pixels)) st = RCG_INVALID;
        r.status = st;
        if (st != RCG_OK) continue;
        JobPlan& p = plans[i];
        p.ok = true;
        p.P32 = plan_pair(j.width, j.height, ctx->tune32.nt, ctx->tune32.th);
        p.P64 = plan_pair(j.width, j.height, ctx->tune64.nt, ctx->tune64.th);
        if (host_src) need += Arena::align(size_t(j.width) * j.height * j.channels);
        need += pyramid_bytes(p.P32);
        for (uint32_t c = 0; c < j.ncand; c++) {
            p.L[c] = make_layout(j.width, j.height, j.channels, j.quality[c], j.subsampling[c]);
            for (uint32_t k = 0; k < p.L[c].ncomp; k++) need += Arena::align(p.L[c].plane_bytes(k));
            need += pyramid_bytes(p.P32) + partial_bytes(p.P32) + partial_bytes(p.P64);
        }
        npairs += j.ncand;
        desc += sizeof(EncJob) * RCG_MAX_CAND + sizeof(EncCand) * j.ncand + 64 * (RCG_MAX_CAND + 2);
        desc += (sizeof(SsimJob) + 64) * 6 * j.ncand + sizeof(FinalDesc) * 2 * j.ncand + 256;
    }
    desc += 16 * 64 + npairs * 64;
    need += Arena::align(npairs * sizeof(double)) * 2 + Arena::align(desc) + (1 << 16);
    ctx->arena.reset();
    RCG_TRY(ctx->arena.ensure(need));
    RCG_TRY(ctx->pinned.ensure(desc + Arena::align(npairs * sizeof(double)) * 2));
    Desc D;
    D.host = ctx->pinned.base;
    D.dev = static_cast<char*>(ctx->arena.take(desc));
    D.cap = desc;
    double* d_s32 = static_cast<double*>(ctx->arena.take(npairs * sizeof(double) + 8));
    double* d_s64 = static_cast<double*>(ctx->arena.take(npairs * sizeof(double) + 8));
    double* h_scores = reinterpret_cast<double*>(ctx->pinned.base + ((desc + 255) & ~size_t(255)));

    // Buffers and the encode jobs.
    std::vector<EncJob> ej;
    std::vector<std::vector<EncCand>> ecands;
    std::vector<PairRef> pairs;
    for (int i = 0; i < njobs; i++) {
        JobPlan& p = plans[i];
        if (!p.ok) continue;
        const rcg_batch_job& j = jobs[i];
        if (host_src) {
            const size_t n = size_t(j.width) * j.height * j.channels;
            uint8_t* d = static_cast<uint8_t*>(ctx->arena.take(n));
            RCG_HIP(hipMemcpyAsync(d, j.pixels, n, hipMemcpyHostToDevice, ctx->stream));
            p.d_src = d;
        } else {
            p.d_src = j.pixels;
        }
        for (uint32_t s = 1; s < p.P32.nscales; s++)
            p.ref_planes[s] = static_cast<float*>(ctx->arena.take(size_t(p.P32.sw[s]) * p.P32.sh[s] * 3 * 4));
        for (uint32_t c = 0; c < j.ncand; c++) {
            for (uint32_t k = 0; k < p.L[c].ncomp; k++)
                p.rec[c][k] = static_cast<uint8_t*>(ctx->arena.take(p.L[c].plane_bytes(k)));
            for (uint32_t s = 1; s < p.P32.nscales; s++)
                p.dist_planes[c][s] =
                    static_cast<float*>(ctx->arena.take(size_t(p.P32.sw[s]) * p.P32.sh[s] * 3 * 4));
            for (uint32_t s = 0; s < p.P32.nscales; s++)
                p.part32[c][s] = static_cast<double*>(ctx->arena.take(size_t(p.P32.tiles[s]) * 18 * 8));
            for (uint32_t s = 0; s < p.P64.nscales; s++)
                p.part64[c][s] = static_cast<double*>(ctx->arena.take(size_t(p.P64.tiles[s]) * 18 * 8));
            pairs.push_back(PairRef{i, int(c)});
        }
        // One encode job per distinct sampling layout of this image.
        int first_group = int(ej.size());
        for (uint32_t c = 0; c < j.ncand; c++) {
            int g = -1;
            for (int e = first_group; e < int(ej.size()); e++)
                if (ej[e].hmax == p.L[c].hmax && ej[e].vmax == p.L[c].vmax && ej[e].ncand < uint32_t(kMaxCand)) g = e;
            if (g < 0) {
                EncJob e;
                std::memset(&e, 0, sizeof(e));
                e.src = p.d_src;
                e.src_bytes = uint64_t(j.width) * j.height * j.channels;
                e.w = j.width;
                e.h = j.height;
                e.channels = j.channels;
                e.hmax = p.L[c].hmax;
                e.vmax = p.L[c].vmax;
                e.pw = p.L[c].pw;
                e.ph = p.L[c].ph;
                e.tiles_x = (e.pw + kEncTileW - 1) / kEncTileW;
                e.tiles_y = e.ph / (8 * e.vmax);
                ej.push_back(e);
                ecands.emplace_back();
                g = int(ej.size()) - 1;
            }
            EncCand k;
            std::memset(&k, 0, sizeof(k));
            uint8_t qy[64], qc[64];
            quant_tables(j.quality[c], qy, qc);
            fill_cand_tables(qy, qc, k);
            for (uint32_t q = 0; q < p.L[c].ncomp; q++) k.rec[q] = p.rec[c][q];
            ecands[g].push_back(k);
            ej[g].ncand++;
            p.group[c] = g;
        }
    }
    std::vector<uint32_t> ebegins(ej.size());
    uint32_t etiles = 0;
    for (size_t e = 0; e < ej.size(); e++) {
        ej[e].cand = D.put(ecands[e].data(), ecands[e].size());
        ej[e].tile_begin = etiles;
        ebegins[e] = etiles;
        etiles += ej[e].tiles_x * ej[e].tiles_y;
    }
    const EncJob* d_ej = ej.empty() ? nullptr : D.put(ej.data(), ej.size());
    const uint32_t* d_eb = ej.empty() ? nullptr : D.put(ebegins.data(), ebegins.size());
    if (!ej.empty()) {
        RCG_TRY(D.upload(ctx->stream));
        RCG_TRY(launch_encode(ctx, d_ej, d_eb, int(ej.size()), etiles));
    }
    if (!pairs.empty()) {
        RCG_TRY(run_pairs(ctx, D, plans, jobs, pairs, RCG_PREC_FP32, d_s32));
        RCG_HIP(hipMemcpyAsync(h_scores, d_s32, pairs.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<PairRef> settle;
    std::vector<size_t> settle_idx;
    for (size_t i = 0; i < pairs.size(); i++) {
        const double s = h_scores[i];
        rcg_batch_result& r = results[pairs[i].job];
        r.scores[pairs[i].cand] = std::isfinite(s) ? s : -DBL_MAX;
        if (std::isfinite(s) && std::fabs(s - jobs[pairs[i].job].target) < kGuard) {
            settle.push_back(pairs[i]);
            settle_idx.push_back(i);
        }
    }
    if (!settle.empty()) {
        D.off = 0;
        RCG_TRY(run_pairs(ctx, D, plans, jobs, settle, RCG_PREC_FP64, d_s64));
        RCG_HIP(hipMemcpyAsync(h_scores, d_s64, settle.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < settle.size(); i++) {
            const double s = h_scores[i];
            rcg_batch_result& r = results[settle[i].job];
            r.scores[settle[i].cand] = std::isfinite(s) ? s : -DBL_MAX;
            r.settled[settle[i].cand] = 1;
        }
    }
    for (int i = 0; i < njobs; i++)
        if (plans[i].ok) select(jobs[i], results[i]);
    return RCG_OK;
}

}  // namespace

extern "C" int rcg_score_batch_dev(rcg_ctx* ctx, const rcg_batch_job* jobs, int njobs, rcg_batch_result* results) {
    return score_batch_impl(ctx, jobs, njobs, results, false);
}

extern "C" int rcg_score_batch(rcg_ctx* ctx, const rcg_batch_job* jobs, int njobs, rcg_batch_result* results) {
    return score_batch_impl(ctx, jobs, njobs, results, true);
}
