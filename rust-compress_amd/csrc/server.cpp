// Serving loop (rcg_server_*): continuous batching of target-quality searches on one GPU.
//
// Two lanes, each a context (stream pair) and a worker thread, serve requests below and above kLargePixels; their
// ticks run side by side on the GPU. A lane's worker repeats ticks. Before a tick it admits its class's queued
// requests in arrival order while the admitted footprint of both lanes (resident source + reference pyramid + the
// largest per-tick working set of its candidates) fits the HBM budget; like cli/src/budget.rs:34-43, a request larger
// than the whole budget is admitted only when nothing else is running. A tick gathers the next candidate(s) of every admitted request's bisection
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
#include <memory>
#include <new>
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
    int coefs_written = 0;
};

// Fault injection for tests: with RCG_TEST_THROW_TICK=<n> in the environment, the worker's n-th tick (counting from
// 1) throws std::bad_alloc once, as a host allocation inside the tick would.
int test_throw_tick() {
    const char* v = std::getenv("RCG_TEST_THROW_TICK");
    return v ? std::atoi(v) : 0;
}

}  // namespace

// Requests of at least kLargePixels run on a lane of their own: a second context, stream pair and worker thread. A
// tick lasts as long as its largest image takes, so small requests scored in the same ticks as a 24 MP image waited
// for it at every step of their bisection; on their own lane their ticks run beside the large ones on the GPU.
constexpr size_t kLargePixels = size_t(2) << 20;
constexpr int kLanes = 2;

struct rcg_server {
    struct Lane {
        rcg_ctx* ctx = nullptr;
        std::vector<Request*> active;  // admitted requests of this lane (mu held)
        std::thread worker;
    };
    Lane lanes[kLanes];
    rcg_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::deque<rcg_completion> done;
    size_t used = 0;      // admitted footprints (mu held)
    size_t nactive = 0;   // admitted requests over both lanes (mu held)
    uint64_t epoch = 0;   // bumped when a request is submitted or finishes (mu held): a lane that could admit nothing
                          // waits for it to change
    // Device blocks of finished requests, kept for reuse by size class: hipMalloc and hipFree of blocks of many
    // megabytes take milliseconds each, and on a worker thread they stalled the ticks (a sustained 200 requests/s
    // spent two thirds of its time there). Admitted footprints plus kept blocks stay within hbm_budget (mu held;
    // hipMalloc and hipFree are called outside the lock).
    std::multimap<size_t, void*> spare;
    size_t spare_bytes = 0;
    bool stop = false;
    rcg_server_stats stats{};
    uint64_t batch_sum = 0;
    int throw_tick = 0;

    static int lane_of(const rcg_request& r) { return size_t(r.width) * r.height >= kLargePixels ? 1 : 0; }

    // Eighth-of-an-octave size classes (at most 12.5% slack), 1 MiB at least.
    static size_t size_class(size_t n) {
        if (n <= (size_t(1) << 20)) return size_t(1) << 20;
        const int b = 63 - __builtin_clzll(uint64_t(n));
        const size_t step = size_t(1) << (b - 3);
        return (n + step - 1) / step * step;
    }

    // Device bytes a request holds while admitted: its resident blocks as allocated (size classes) and the largest
    // per-tick working set of its candidates. Both lanes' contexts carry the same tile policy.
    size_t footprint_of(const rcg_request& r) const {
        rcg_ctx* ctx = lanes[0].ctx;
        const PairPlan P = plan_pair(r.width, r.height, ctx->tune32);
        size_t b = size_class(size_t(r.width) * r.height * r.channels) + size_class(ref_state_bytes(P));
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

    // Give kept blocks back, the largest first, until at most `keep` bytes of them remain.
    void free_spares(size_t keep = 0) {
        std::vector<void*> out;
        {
            std::lock_guard<std::mutex> lk(mu);
            while (spare_bytes > keep && !spare.empty()) {
                auto it = std::prev(spare.end());
                spare_bytes -= it->first;
                out.push_back(it->second);
                spare.erase(it);
            }
        }
        for (void* p : out) (void)hipFree(p);
    }
    int device_alloc(size_t n, void** p) {
        const size_t c = size_class(n);
        {
            std::lock_guard<std::mutex> lk(mu);
            auto it = spare.find(c);
            if (it != spare.end()) {
                *p = it->second;
                spare.erase(it);
                spare_bytes -= c;
                return RCG_OK;
            }
        }
        if (hipMalloc(p, c) == hipSuccess) return RCG_OK;
        (void)hipGetLastError();
        // The kept blocks of other sizes may be what stands in the way: give them back and try once more.
        free_spares();
        RCG_HIP(hipMalloc(p, c));
        return RCG_OK;
    }
    // A finished request's block: kept for reuse while admitted footprints and kept blocks fit the budget (mu held).
    void device_release(void* p, size_t n) {
        if (!p) return;
        const size_t c = size_class(n);
        if (used + spare_bytes + c > opts.hbm_budget) {
            (void)hipFree(p);
            return;
        }
        spare.emplace(c, p);
        spare_bytes += c;
    }

    // mu held.
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
        c.coefs_written = r->coefs_written;
        used -= std::min(used, r->footprint);
        nactive--;
        epoch++;
        device_release(r->d_src, r->src_bytes);
        device_release(r->d_ref, r->ref_bytes);
        done.push_back(c);
        stats.requests_done++;
        delete r;
    }

    // mu held.
    bool has_pending(int lane) const {
        for (const Request* r : pending)
            if (lane_of(r->req) == lane) return true;
        return false;
    }

    // Admit queued requests of the lane's class in arrival order while they fit (mu held); the footprints of both
    // lanes share the budget, and a request larger than the whole budget is admitted only when no request is admitted
    // on either lane. Admission stays first come, first served across the lanes: an older request of the other lane
    // keeps the budget it needs (it is admitted by its own lane's worker), and one that does not fit stops the scan.
    // Kept blocks give way to admitted requests: the return value is how many bytes of kept blocks may remain
    // (free_spares() before the next tick).
    size_t admit(int lane) {
        Lane& L = lanes[lane];
        const size_t budget = opts.hbm_budget;
        size_t ahead = 0;  // budget the older requests of the other lane will take
        for (auto it = pending.begin(); it != pending.end();) {
            Request* r = *it;
            const size_t need = std::min(r->footprint, budget);
            const bool fits = (nactive == 0 && ahead == 0) || used + ahead + need <= budget;
            if (!fits) break;
            if (lane_of(r->req) != lane) {
                ahead += need;
                ++it;
                continue;
            }
            if (opts.max_active > 0 && int(nactive) >= opts.max_active) break;
            it = pending.erase(it);
            r->admit_s = now_s();
            used += need;
            r->footprint = need;
            nactive++;
            stats.peak_footprint = std::max(stats.peak_footprint, used);
            L.active.push_back(r);
        }
        return used >= opts.hbm_budget ? 0 : opts.hbm_budget - used;
    }

    // Take resident buffers and enqueue the source upload of a newly admitted request (worker thread, mu not held).
    // The upload is ordered before the request's first tick on the lane context's stream; the host pixels are
    // released once that tick has synchronised the stream.
    int materialise(rcg_ctx* ctx, Request* r) {
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

    // One scoring pass over `items` on a lane's context.
    static int score(rcg_ctx* ctx, std::vector<Item>& items) {
        std::lock_guard<std::mutex> clk(ctx->mu);
        ctx->arena.reset();
        RCG_TRY(ctx->arena.ensure(score_items_bytes(ctx, items.data(), int(items.size()))));
        return score_items(ctx, items.data(), int(items.size()));
    }

    // The chosen quality's coefficients of the requests in `fin` that asked for them, in one encode launch.
    static int write_winners(rcg_ctx* ctx, const std::vector<Request*>& fin) {
        std::vector<Item> items;
        std::vector<int> which, chosen;
        std::vector<WinnerDest> dest;
        for (Request* r : fin) {
            if (!r->req.coefs[0]) continue;
            Item it;
            std::memset(&it, 0, sizeof(it));
            it.d_src = r->d_src;
            it.w = r->req.width;
            it.h = r->req.height;
            it.ch = r->req.channels;
            it.ncand = 1;
            it.quality[0] = r->search.best_q;
            it.subsampling[0] = r->req.subsampling;
            which.push_back(int(items.size()));
            chosen.push_back(0);
            WinnerDest d;
            for (int c = 0; c < 3; c++) d.coefs[c] = r->req.coefs[c];
            dest.push_back(d);
            items.push_back(it);
        }
        if (items.empty()) return RCG_OK;
        std::lock_guard<std::mutex> clk(ctx->mu);
        ctx->arena.reset();
        RCG_TRY(ctx->arena.ensure(winner_bytes(items.data(), which.data(), int(which.size()), true)));
        return encode_winners(ctx, ctx->arena, ctx->pinned, items.data(), which.data(), chosen.data(), dest.data(),
                              int(which.size()), true);
    }

    // A tick: score the wanted candidates of every request in `batch` (the failing ones are marked), apply the
    // scores, and finish the requests that are done. Returns the number of pairs scored.
    size_t tick(Lane& L, const std::vector<Request*>& batch, const std::vector<Request*>& fresh, double& t0,
                double& t1) {
        rcg_ctx* const ctx = L.ctx;
        std::vector<Item> items;
        std::vector<std::vector<int>> want(batch.size());
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
        t0 = now_s();
        std::vector<int> status(items.size(), RCG_OK);
        if (!items.empty()) {
            bool throw_now = false;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (throw_tick > 0 && stats.ticks + 1 == uint64_t(throw_tick)) {
                    throw_tick = 0;
                    throw_now = true;
                }
            }
            if (throw_now) throw std::bad_alloc();
            int rc = score(ctx, items);
            if (rc != RCG_OK && items.size() > 1) {
                // Isolate the failure: each request alone, so that only those that fail alone fail.
                for (size_t k = 0; k < items.size(); k++) {
                    std::vector<Item> one(1, items[k]);
                    status[k] = score(ctx, one);
                    items[k] = one[0];
                }
            } else if (rc != RCG_OK) {
                status[0] = rc;
            }
        }
        t1 = now_s();
        if (!fresh.empty()) {
            // A tick that succeeded read its results back, so the uploads ahead of it have completed; a failed one
            // waited for its streams before returning. A tick that scored nothing may not have waited.
            (void)hipStreamSynchronize(ctx->stream);
            for (Request* r : fresh) std::vector<uint8_t>().swap(r->pixels);
        }
        std::vector<Request*> fin;
        for (size_t i = 0; i < batch.size(); i++) {
            Request* r = batch[i];
            if (r->status == RCG_OK && status[item_of[i]] != RCG_OK) r->status = status[item_of[i]];
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
            if (r->status != RCG_OK || r->search.done) fin.push_back(r);
        }
        std::vector<Request*> ok;
        for (Request* r : fin)
            if (r->status == RCG_OK && r->req.coefs[0]) ok.push_back(r);
        const int wrc = write_winners(ctx, ok);
        for (Request* r : ok) {
            if (wrc == RCG_OK)
                r->coefs_written = 1;
            else
                r->status = wrc;
        }
        std::lock_guard<std::mutex> lk(mu);
        stats.ticks++;
        stats.busy_s += t1 - t0;
        batch_sum += items.size();
        for (Request* r : fin) {
            L.active.erase(std::find(L.active.begin(), L.active.end(), r));
            finish(r, t1);
        }
        if (stats.ticks) stats.mean_batch = double(batch_sum) / double(stats.ticks);
        return items.size();
    }

    void run(int lane) {
        Lane& L = lanes[lane];
        rcg_ctx* const ctx = L.ctx;
        RCG_HIP_VOID(hipSetDevice(ctx->device));
        uint64_t seen = ~uint64_t(0);  // the epoch at which queued requests of this lane last could not be admitted
        for (;;) {
            std::vector<Request*> fresh, batch;
            size_t room = 0;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !L.active.empty() || (epoch != seen && has_pending(lane)); });
                if (stop) return;
                const size_t before = L.active.size();
                room = admit(lane);
                for (size_t i = before; i < L.active.size(); i++) fresh.push_back(L.active[i]);
                batch = L.active;
                if (batch.empty()) {
                    // Its queued requests wait for budget that the other lane holds: until a request finishes or
                    // arrives.
                    seen = epoch;
                    continue;
                }
            }
            free_spares(room);
            double t0 = now_s(), t1 = t0;
            try {
                for (Request* r : fresh)
                    if (r->d_src == nullptr) {
                        const int rc = materialise(ctx, r);
                        if (rc != RCG_OK) r->status = rc;
                    }
                tick(L, batch, fresh, t0, t1);
            } catch (...) {
                // Nothing may leave the worker: the requests of this tick fail, and the loop goes on.
                const int st = [] {
                    try {
                        throw;
                    } catch (const std::bad_alloc&) {
                        return int(RCG_OOM);
                    } catch (...) {
                        return int(RCG_DEVICE);
                    }
                }();
                set_error(st == RCG_OOM ? "out of host memory in a serving tick" : "internal error in a serving tick");
                (void)hipStreamSynchronize(ctx->side);
                (void)hipStreamSynchronize(ctx->stream);
                for (Request* r : fresh) std::vector<uint8_t>().swap(r->pixels);
                std::lock_guard<std::mutex> lk(mu);
                for (Request* r : batch) {
                    if (std::find(L.active.begin(), L.active.end(), r) == L.active.end()) continue;
                    if (r->status == RCG_OK) r->status = st;
                    L.active.erase(std::find(L.active.begin(), L.active.end(), r));
                    finish(r, now_s());
                }
            }
            // Finished requests free budget the other lane may be waiting for.
            cv_work.notify_all();
            cv_done.notify_all();
        }
    }
};

extern "C" {

int rcg_server_create(int device, const rcg_server_opts* opts, rcg_server** out) {
    return rcg::guarded([&]() -> int {
        if (!out) return RCG_INVALID;
        *out = nullptr;
        std::unique_ptr<rcg_server> s(new rcg_server());
        if (opts) s->opts = *opts;
        if (s->opts.hbm_budget == 0) s->opts.hbm_budget = size_t(16) << 30;
        if (s->opts.speculate != 3) s->opts.speculate = 1;
        s->throw_tick = test_throw_tick();
        struct Owner {  // destroys the contexts unless the workers took them over
            rcg_server* s;
            bool armed = true;
            ~Owner() {
                if (!armed) return;
                for (auto& L : s->lanes)
                    if (L.ctx) rcg_destroy(L.ctx);
            }
        } owner{s.get()};
        for (auto& L : s->lanes) {
            RCG_TRY(rcg_create(device, &L.ctx));
            if (s->opts.tune) RCG_TRY(rcg_set_tune(L.ctx, s->opts.tune));
        }
        s->opts.tune = nullptr;
        rcg_server* raw = s.get();
        int started = 0;
        try {
            for (; started < kLanes; started++) {
                const int lane = started;
                s->lanes[lane].worker = std::thread([raw, lane] { raw->run(lane); });
            }
        } catch (...) {
            // A worker that did start is stopped before the contexts go.
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->stop = true;
            }
            s->cv_work.notify_all();
            for (int i = 0; i < started; i++) s->lanes[i].worker.join();
            throw;
        }
        owner.armed = false;
        *out = s.release();
        return RCG_OK;
    });
}

int rcg_server_submit(rcg_server* s, const rcg_request* req) {
    return rcg::guarded([&]() -> int {
        if (!s || !req || !req->pixels) return RCG_INVALID;
        RCG_TRY(check_image(req->width, req->height, req->channels));
        if (req->width < 8 || req->height < 8) {
            set_error("--target-ssimulacra2 needs an image of at least 8x8 pixels");
            return RCG_UNSUPPORTED;
        }
        std::unique_ptr<Request> r(new Request());
        r->req = *req;
        r->pixels.assign(req->pixels, req->pixels + size_t(req->width) * req->height * req->channels);
        r->req.pixels = nullptr;
        r->submit_s = now_s();
        r->search.max_q = std::max(1, std::min(100, req->max_quality));
        r->search.target = req->target;
        r->search.hi = r->search.max_q;
        const uint32_t ncomp = req->channels == 1 ? 1 : 3;
        for (uint32_t c = 0; req->coefs[0] && c < ncomp; c++)
            if (!req->coefs[c]) return RCG_INVALID;
        r->footprint = s->footprint_of(*req);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->pending.push_back(r.get());
            s->epoch++;
        }
        r.release();
        s->cv_work.notify_all();
        return RCG_OK;
    });
}

int rcg_server_poll(rcg_server* s, rcg_completion* out, int max, int timeout_ms, int* n) {
    return rcg::guarded([&]() -> int {
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
    });
}

int rcg_server_get_stats(rcg_server* s, rcg_server_stats* out) {
    return rcg::guarded([&]() -> int {
        if (!s || !out) return RCG_INVALID;
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->stats;
        return RCG_OK;
    });
}

void rcg_server_destroy(rcg_server* s) {
    rcg::guarded_void([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        for (auto& L : s->lanes)
            if (L.worker.joinable()) L.worker.join();
        for (auto& L : s->lanes) {
            (void)hipSetDevice(L.ctx->device);
            (void)hipStreamSynchronize(L.ctx->side);
            (void)hipStreamSynchronize(L.ctx->stream);
            for (Request* r : L.active) {
                if (r->d_src) (void)hipFree(r->d_src);
                if (r->d_ref) (void)hipFree(r->d_ref);
                delete r;
            }
        }
        for (auto& b : s->spare) (void)hipFree(b.second);
        for (Request* r : s->pending) delete r;
        for (auto& L : s->lanes) rcg_destroy(L.ctx);
        delete s;
    });
}

}  // extern "C"
