// The serving queue under concurrent callers, built with the host side of the library instrumented by
// AddressSanitizer and UndefinedBehaviorSanitizer (rust-compress_amd/build.py). Four threads submit requests while the
// main thread polls completions and another thread reads the statistics, against one server on device 0: submission,
// admission, the worker's ticks, completion and shutdown all run concurrently. Every request must complete exactly
// once with status 0, and requests of the same image and target must get the same quality and score; a sanitizer
// report ends the process. Prints "OK" and the counts, exit 0.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/rcgpu.h"

namespace {

struct Img {
    uint32_t w, h;
    std::vector<uint8_t> px;
};

}  // namespace

// Ends the process once the server is destroyed and the verdict printed, without running the HIP runtime's own
// teardown: under AddressSanitizer that teardown can free runtime memory after the sanitizer's device allocator has
// marked the runtime unloaded, and the sanitizer then aborts on its own internal check after main has returned.
static int finish(int rc) {
    std::fflush(stdout);
    std::_Exit(rc);
}

int main() {
    constexpr int kThreads = 4, kPerThread = 12, kTotal = kThreads * kPerThread;
    // The last size is above the server's large-request threshold (2 MiB pixels): both lanes serve requests.
    const uint32_t sizes[][2] = {{96, 80}, {128, 128}, {640, 480}, {64, 64}, {333, 250}, {1024, 768}, {1600, 1400}};
    constexpr int kImages = 7;
    const double targets[] = {60.0, 70.0, 80.0};
    std::vector<Img> imgs;
    for (int i = 0; i < kImages; i++) {
        Img m{sizes[i][0], sizes[i][1], {}};
        m.px.resize(size_t(m.w) * m.h * 3);
        if (rcg_synth_rgb8(100 + i, m.w, m.h, m.px.data()) != RCG_OK) return finish(2);
        imgs.push_back(std::move(m));
    }
    rcg_server_opts opts{};
    opts.speculate = 3;
    rcg_server* srv = nullptr;
    if (rcg_server_create(0, &opts, &srv) != RCG_OK) {
        std::printf("rcg_server_create failed: %s\n", rcg_last_error());
        return finish(2);
    }
    // What each request is: (image, target index).
    std::vector<std::pair<int, int>> what(kTotal);
    for (int t = 0; t < kThreads; t++)
        for (int k = 0; k < kPerThread; k++) what[t * kPerThread + k] = {(t + k) % kImages, k % 3};

    std::atomic<int> submit_failures{0};
    std::atomic<bool> stop{false};
    std::vector<std::thread> submitters;
    for (int t = 0; t < kThreads; t++)
        submitters.emplace_back([&, t] {
            for (int k = 0; k < kPerThread; k++) {
                const int id = t * kPerThread + k;
                const Img& m = imgs[what[id].first];
                rcg_request r{};
                r.id = uint64_t(id);
                r.pixels = m.px.data();
                r.width = m.w;
                r.height = m.h;
                r.channels = 3;
                r.target = targets[what[id].second];
                r.max_quality = 100;
                r.subsampling = RCG_S420;
                if (rcg_server_submit(srv, &r) != RCG_OK) submit_failures++;
                // Arrivals spread over about 150 ms, so that most arrive while ticks run.
                std::this_thread::sleep_for(std::chrono::milliseconds((id * 7919) % 25));
            }
        });
    std::thread reader([&] {
        while (!stop.load()) {
            rcg_server_stats s;
            rcg_server_get_stats(srv, &s);
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
    });

    std::vector<int> seen(kTotal, 0);
    std::vector<rcg_completion> got(kTotal);
    int received = 0, bad = 0;
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::seconds(150);
    while (received < kTotal && std::chrono::steady_clock::now() < deadline) {
        rcg_completion buf[16];
        int n = 0;
        if (rcg_server_poll(srv, buf, 16, 200, &n) != RCG_OK) break;
        for (int i = 0; i < n; i++) {
            const uint64_t id = buf[i].id;
            if (id >= uint64_t(kTotal)) {
                bad++;
                continue;
            }
            seen[id]++;
            got[id] = buf[i];
            received++;
        }
    }
    for (auto& s : submitters) s.join();
    stop = true;
    reader.join();
    rcg_server_stats st{};
    rcg_server_get_stats(srv, &st);
    rcg_server_destroy(srv);

    if (submit_failures.load() || bad || received != kTotal) {
        std::printf("FAIL: %d submissions refused, %d unknown ids, %d of %d completions\n", submit_failures.load(), bad,
                    received, kTotal);
        return finish(1);
    }
    std::map<std::pair<int, int>, std::pair<int, double>> result;
    for (int id = 0; id < kTotal; id++) {
        if (seen[id] != 1 || got[id].status != RCG_OK) {
            std::printf("FAIL: request %d completed %d times, status %d\n", id, seen[id], got[id].status);
            return finish(1);
        }
        const auto r = std::make_pair(got[id].quality, got[id].score);
        auto it = result.find(what[id]);
        if (it == result.end())
            result[what[id]] = r;
        else if (it->second != r) {
            std::printf("FAIL: request %d (image %d, target %.0f) got q%d %.9f, another got q%d %.9f\n", id,
                        what[id].first, targets[what[id].second], r.first, r.second, it->second.first,
                        it->second.second);
            return finish(1);
        }
    }
    std::printf("OK %d requests, %zu distinct searches, %llu ticks, mean batch %.2f\n", kTotal, result.size(),
                (unsigned long long)st.ticks, st.mean_batch);
    return finish(0);
}
