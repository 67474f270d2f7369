// Serving loop (rcg_server_*): continuous batching of target-quality searches on one GPU.
//
// The worker thread repeats ticks. Before a tick it admits queued requests in arrival order while the admitted
// footprint (resident source + reference pyramid + the largest per-tick working set of its candidates) fits the
// HBM budget; like cli/src/budget.rs:34-43, a request larger than the whole budget is admitted only when nothing
// else is running. A tick gathers the next candidate(s) of every admitted request's bisection
// (jpeg/src/target.rs:37-55) and scores them all with one score_items pass (one encode launch, one launch per
// SSIMULACRA2 scale, one pooling launch, FP64 settling near each target). Finished requests leave the batch and
// free their memory; requests that arrived meanwhile join at the next tick. A pair's score depends only on its own
// image and candidate (fixed tiling, fixed-order reductions), so each result equals the one it gets alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "rcg_host.h"

using namespace rcg;

#define RCG_HIP_VOID(expr) (void)(expr)

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Bisection of jpeg/src/target.rs:37-55 as an explicit state machine.
struct Search {
    int max_q = 100;
    double target = 0;
    bool top_done = false;
    int lo = 1, hi = 100;
    int best_q = 0;
    double best_s = 0;
    bool done = false;

    // The candidate whose score decides the next step.
    int next() const { return !top_done ? max_q : lo + (hi - lo) / 2; }
    // Candidates worth scoring now: the next one and, speculatively, the ones that follow either outcome.
    void wanted(int depth, std::vector<int>& out) const {
        out.clear();
        const int q = next();
        out.push_back(q);
        if (depth < 3) return;
        if (!top_done) {
            if (max_q > 1) out.push_back(1 + (max_q - 1) / 2);
            return;
        }
        const int mid = q;
        // pass: hi = mid -> lo + (mid - lo)/2 ; fail: lo = mid + 1 -> (mid+1) + (hi-mid-1)/2
        if (lo < mid) out.push_back(lo + (mid - lo) / 2);
        if (mid + 1 < hi) out.push_back((mid + 1) + (hi - mid - 1) / 2);
    }
    // Apply the score of the current candidate; returns false if `q` is not the current one.
    bool apply(int q, double s) {
        if (done || q != next()) return false;
        if (!top_done) {
            top_done = true;
            best_q = max_q;
            best_s = s;
            if (s < target || max_q <= 1) {
                done = true;
                return true;
            }
            lo = 1;
            hi = max_q;
            if (lo >= hi) done = true;
            return true;
        }
        const int mid = q;
        if (s >= target) {
            hi = mid;
            best_q = mid;
            best_s = s;
        } else {
            lo = mid + 1;
        }
        if (lo >= hi) done = true;
        return true;
    }
};

struct Request {
    rcg_request req;
    std::vector<uint8_t> pixels;
    double submit_s = 0, admit_s = 0;
    size_t footprint = 0;
    uint8_t* d_src = nullptr;
    size_t src_bytes = 0, ref_bytes = 0;
    RefState ref{};
    void* d_ref = nullptr;
    bool ref_ready = false;
    Search search;
    int steps = 0, ticks = 0;
    int status = RCG_OK;
};

}  // namespace

struct rcg_server {
    rcg_ctx* ctx = nullptr;
    rcg_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::vector<Request*> active;
    std::deque<rcg_completion> done;
    size_t used = 0;
    // Device blocks of finished requests, kept for reuse by size class: hipMalloc and hipFree of blocks of many
    // megabytes take milliseconds each, and on the worker thread they stalled the ticks (a sustained 200 requests/s
    // spent two thirds of its time there). Touched by the worker thread only; at most hbm_budget bytes are kept.
    std::multimap<size_t, void*> spare;
    size_t spare_bytes = 0;
    bool stop = false;
    std::thread worker;
    rcg_server_stats stats{};
    uint64_t batch_sum = 0;

    size_t footprint_of(const rcg_request& r) const {
        const PairPlan P = plan_pair(r.width, r.height, ctx->tune32);
        size_t b = Arena::align(size_t(r.width) * r.height * r.channels) + ref_state_bytes(P);
        // Under Auto a tick's candidates are 4:4:4 (q >= 90) or 4:2:0 (q < 90): charge the larger working set.
        size_t work = 0;
        for (int q : {100, 1}) {
            Item it;
            std::memset(&it, 0, sizeof(it));
            it.w = r.width;
            it.h = r.height;
            it.ch = r.channels;
            it.ncand = uint32_t(std::max(1, opts.speculate));
            for (uint32_t c = 0; c < it.ncand; c++) {
                it.quality[c] = q;
                it.subsampling[c] = r.subsampling;
            }
            work = std::max(work, score_items_bytes(ctx, &it, 1));
        }
        return b + work;
    }

    // Eighth-of-an-octave size classes (at most 12.5% slack), 1 MiB at least.
    static size_t size_class(size_t n) {
        if (n <= (size_t(1) << 20)) return size_t(1) << 20;
        const int b = 63 - __builtin_clzll(uint64_t(n));
        const size_t step = size_t(1) << (b - 3);
        return (n + step - 1) / step * step;
    }
    int device_alloc(size_t n, void** p) {
        const size_t c = size_class(n);
        auto it = spare.find(c);
        if (it != spare.end()) {
            *p = it->second;
            spare.erase(it);
            spare_bytes -= c;
            return RCG_OK;
        }
        RCG_HIP(hipMalloc(p, c));
        return RCG_OK;
    }
    void device_release(void* p, size_t n) {
        if (!p) return;
        const size_t c = size_class(n);
        if (spare_bytes + c > opts.hbm_budget) {
            (void)hipFree(p);
            return;
        }
        spare.emplace(c, p);
        spare_bytes += c;
    }

    void finish(Request* r, double t) {
        rcg_completion c{};
        c.id = r->req.id;
        c.status = r->status;
        c.quality = r->search.best_q;
        c.score = r->search.best_s;
        c.steps = r->steps;
        c.ticks = r->ticks;
        c.submit_s = r->submit_s;
        c.admit_s = r->admit_s;
        c.done_s = t;
        device_release(r->d_src, r->src_bytes);
        device_release(r->d_ref, r->ref_bytes);
        used -= std::min(used, r->footprint);
        done.push_back(c);
        stats.requests_done++;
        delete r;
    }

    // Admit queued requests in order while they fit (mu held).
    void admit() {
        while (!pending.empty()) {
            Request* r = pending.front();
            const size_t budget = opts.hbm_budget;
            const size_t need = std::min(r->footprint, budget);
            if (opts.max_active > 0 && int(active.size()) >= opts.max_active) break;
            if (!active.empty() && used + need > budget) break;
            pending.pop_front();
            r->admit_s = now_s();
            used += need;
            r->footprint = need;
            stats.peak_footprint = std::max(stats.peak_footprint, used);
            active.push_back(r);
        }
    }

    // Take resident buffers and enqueue the source upload of a newly admitted request (worker thread, mu not held).
    // The upload is ordered before the request's first tick on the context's stream; the host pixels are released
    // once that tick has synchronised the stream.
    int materialise(Request* r) {
        r->src_bytes = size_t(r->req.width) * r->req.height * r->req.channels;
        RCG_TRY(device_alloc(r->src_bytes, reinterpret_cast<void**>(&r->d_src)));
        RCG_HIP(hipMemcpyAsync(r->d_src, r->pixels.data(), r->src_bytes, hipMemcpyHostToDevice, ctx->stream));
        const PairPlan P = plan_pair(r->req.width, r->req.height, ctx->tune32);
        r->ref_bytes = ref_state_bytes(P);
        RCG_TRY(device_alloc(r->ref_bytes, &r->d_ref));
        Arena tmp;
        tmp.base = static_cast<char*>(r->d_ref);
        tmp.cap = r->ref_bytes;
        take_ref_state(tmp, P, r->ref);
        return RCG_OK;
    }

    void run() {
        RCG_HIP_VOID(hipSetDevice(ctx->device));
        std::vector<Item> items;
        std::vector<std::vector<int>> want;
        for (;;) {
            std::vector<Request*> fresh;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
                if (stop) return;
                const size_t before = active.size();
                admit();
                for (size_t i = before; i < active.size(); i++) fresh.push_back(active[i]);
            }
            for (Request* r : fresh)
                if (r->d_src == nullptr && materialise(r) != RCG_OK) r->status = RCG_DEVICE;
            // One tick over every active request (the worker is the only writer of `active`'s entries).
            std::vector<Request*> batch;
            {
                std::lock_guard<std::mutex> lk(mu);
                batch = active;
            }
            items.clear();
            want.assign(batch.size(), {});
            std::vector<int> item_of(batch.size(), -1);
            for (size_t i = 0; i < batch.size(); i++) {
                Request* r = batch[i];
                if (r->status != RCG_OK) continue;
                r->search.wanted(std::max(1, opts.speculate), want[i]);
                Item it;
                std::memset(&it, 0, sizeof(it));
                it.d_src = r->d_src;
                it.w = r->req.width;
                it.h = r->req.height;
                it.ch = r->req.channels;
                it.ncand = uint32_t(want[i].size());
                for (uint32_t c = 0; c < it.ncand; c++) {
                    it.quality[c] = want[i][c];
                    it.subsampling[c] = r->req.subsampling;
                }
                it.target = r->req.target;
                it.ref = r->ref;
                it.ref_ready = r->ref_ready;
                item_of[i] = int(items.size());
                items.push_back(it);
            }
            const double t0 = now_s();
            int rc = RCG_OK;
            if (!items.empty()) {
                std::lock_guard<std::mutex> clk(ctx->mu);
                ctx->arena.reset();
                rc = ctx->arena.ensure(score_items_bytes(ctx, items.data(), int(items.size())));
                if (rc == RCG_OK) rc = score_items(ctx, items.data(), int(items.size()));
            }
            const double t1 = now_s();
            if (!fresh.empty()) {
                // A tick that succeeded read its results back, so the uploads ahead of it have completed. A tick that
                // scored nothing (every fresh request failed to materialise) or failed part way may not have waited.
                if (items.empty() || rc != RCG_OK) (void)hipStreamSynchronize(ctx->stream);
                for (Request* r : fresh) std::vector<uint8_t>().swap(r->pixels);
            }
            std::lock_guard<std::mutex> lk(mu);
            stats.ticks++;
            stats.busy_s += t1 - t0;
            batch_sum += items.size();
            for (size_t i = 0; i < batch.size(); i++) {
                Request* r = batch[i];
                if (r->status == RCG_OK && rc != RCG_OK) r->status = rc;
                if (r->status == RCG_OK) {
                    const Item& it = items[item_of[i]];
                    r->ref_ready = true;
                    r->ticks++;
                    r->steps += int(it.ncand);
                    // Apply the scores in decision order: the current candidate, then whichever speculative one
                    // became current.
                    bool progressed = true;
                    while (progressed && !r->search.done) {
                        progressed = false;
                        for (uint32_t c = 0; c < it.ncand; c++)
                            if (r->search.apply(it.quality[c], it.scores[c])) {
                                progressed = true;
                                break;
                            }
                    }
                }
                if (r->status != RCG_OK || r->search.done) {
                    active.erase(std::find(active.begin(), active.end(), r));
                    finish(r, t1);
                }
            }
            if (stats.ticks) stats.mean_batch = double(batch_sum) / double(stats.ticks);
            cv_done.notify_all();
        }
    }
};

extern "C" {

int rcg_server_create(int device, const rcg_server_opts* opts, rcg_server** out) {
    if (!out) return RCG_INVALID;
    *out = nullptr;
    rcg_server* s = new (std::nothrow) rcg_server();
    if (!s) return RCG_OOM;
    if (opts) s->opts = *opts;
    if (s->opts.hbm_budget == 0) s->opts.hbm_budget = size_t(16) << 30;
    if (s->opts.speculate != 3) s->opts.speculate = 1;
    int rc = rcg_create(device, &s->ctx);
    if (rc != RCG_OK) {
        delete s;
        return rc;
    }
    if (s->opts.tune) {
        rc = rcg_set_tune(s->ctx, s->opts.tune);
        if (rc != RCG_OK) {
            rcg_destroy(s->ctx);
            delete s;
            return rc;
        }
        s->opts.tune = nullptr;
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return RCG_OK;
}

int rcg_server_submit(rcg_server* s, const rcg_request* req) {
    if (!s || !req || !req->pixels) return RCG_INVALID;
    RCG_TRY(check_image(req->width, req->height, req->channels));
    if (req->width < 8 || req->height < 8) {
        set_error("--target-ssimulacra2 needs an image of at least 8x8 pixels");
        return RCG_UNSUPPORTED;
    }
    Request* r = new (std::nothrow) Request();
    if (!r) return RCG_OOM;
    r->req = *req;
    r->pixels.assign(req->pixels, req->pixels + size_t(req->width) * req->height * req->channels);
    r->req.pixels = nullptr;
    r->submit_s = now_s();
    r->search.max_q = std::max(1, std::min(100, req->max_quality));
    r->search.target = req->target;
    r->search.hi = r->search.max_q;
    r->footprint = s->footprint_of(*req);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->pending.push_back(r);
    }
    s->cv_work.notify_one();
    return RCG_OK;
}

int rcg_server_poll(rcg_server* s, rcg_completion* out, int max, int timeout_ms, int* n) {
    if (!s || (!out && max > 0) || !n) return RCG_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && timeout_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->done.empty(); });
    int k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *n = k;
    return RCG_OK;
}

int rcg_server_get_stats(rcg_server* s, rcg_server_stats* out) {
    if (!s || !out) return RCG_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return RCG_OK;
}

void rcg_server_destroy(rcg_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (Request* r : s->active) {
        if (r->d_src) (void)hipFree(r->d_src);
        if (r->d_ref) (void)hipFree(r->d_ref);
        delete r;
    }
    for (auto& b : s->spare) (void)hipFree(b.second);
    for (Request* r : s->pending) delete r;
    rcg_destroy(s->ctx);
    delete s;
}

}  // extern "C"
