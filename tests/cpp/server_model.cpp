// A host model of everything the serving loop (rust-compress_amd/csrc/server.cpp) asks of the device, so that the
// real server code -- submission, admission against the HBM budget, the two lanes' workers and their ticks, the
// reuse of finished requests' blocks, completion and shutdown -- runs on a machine without a GPU, under
// ThreadSanitizer (rust-compress_amd/build.py links it as tests/cpp/server_queue_tsan with tests/cpp/server_race.cpp).
//
// It replaces, for this one test binary, the HIP runtime calls the server makes (device memory is host memory,
// copies are memcpy, streams are already idle) and the library's scoring pass: score_items gives each candidate a
// score that depends only on the uploaded source bytes and the quality, and grows with the quality, so the server's
// bisection (jpeg/src/target.rs:37-55) runs its real course, and a pass takes a few hundred microseconds of wall time
// so that the lanes' ticks overlap the submitting threads. The scores are not SSIMULACRA2: what the test checks is the
// queue (every request completes once, the same request always gets the same result, no data race), not the metric.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>

#include "../../rust-compress_amd/csrc/rcg_host.h"

// ---------------------------------------------------------------------------------------------- the HIP calls
extern "C" {
hipError_t hipMalloc(void** p, size_t n) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t) {
    std::memcpy(dst, src, n);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "model device error"; }
}

namespace {
thread_local std::string g_error;

// A stand-in for a pass's device time: proportional to the candidate-pixels scored, at most a few milliseconds.
void device_time(double cand_pixels) {
    const double us = std::min(3000.0, 50.0 + cand_pixels / 2.0e4);
    std::this_thread::sleep_for(std::chrono::microseconds(int64_t(us)));
}
}  // namespace

// ------------------------------------------------------------------------------------- the library's host side
namespace rcg {

void set_error(const std::string& msg) noexcept {
    try {
        g_error = msg;
    } catch (...) {
    }
}

int check_image(uint32_t w, uint32_t h, uint32_t channels) {
    if (w == 0 || h == 0 || w > 65535 || h > 65535) return RCG_INVALID;
    return channels == 1 || channels == 3 ? RCG_OK : RCG_UNSUPPORTED;
}

int Arena::ensure(size_t bytes) {
    if (bytes <= cap) return RCG_OK;
    std::free(base);
    base = static_cast<char*>(std::malloc(bytes));
    cap = base ? bytes : 0;
    return base ? RCG_OK : RCG_OOM;
}

int Pinned::ensure(size_t bytes) {
    if (bytes <= cap) return RCG_OK;
    std::free(base);
    base = static_cast<char*>(std::malloc(bytes));
    cap = base ? bytes : 0;
    return base ? RCG_OK : RCG_OOM;
}

PairPlan plan_pair(uint32_t w, uint32_t h, const SsimTune&) {
    PairPlan P;
    std::memset(&P, 0, sizeof(P));
    P.w = w;
    P.h = h;
    return P;
}

// The reference state's footprint and carve-up as the library sizes them: about 48 bytes per source pixel.
size_t ref_state_bytes(const PairPlan& P) { return Arena::align(size_t(P.w) * P.h * 48); }
void take_ref_state(Arena& A, const PairPlan& P, RefState& R) {
    std::memset(&R, 0, sizeof(R));
    R.cache[0] = static_cast<float*>(A.take(size_t(P.w) * P.h * 48));
}

size_t score_items_bytes(rcg_ctx*, const Item* items, int n) {
    size_t b = 0;
    for (int i = 0; i < n; i++) b += Arena::align(size_t(items[i].w) * items[i].h * (2 + items[i].ncand));
    return b;
}

// Scores that grow with the quality: 100 - (100 - q) * k, where k in [0.5, 1.5) follows from the source bytes the
// server uploaded (so a missing or torn upload changes them). Writes the item's reference state on its first pass,
// as the real pass does.
int score_items(rcg_ctx* ctx, Item* items, int n) {
    double work = 0;
    for (int i = 0; i < n; i++) {
        Item& it = items[i];
        const size_t bytes = size_t(it.w) * it.h * it.ch;
        uint64_t sum = 1469598103934665603ull;
        for (size_t k = 0; k < bytes; k += 97) sum = (sum ^ it.d_src[k]) * 1099511628211ull;
        if (!it.ref_ready && it.ref.cache[0]) std::memset(it.ref.cache[0], 0, 64);
        const double kq = 0.5 + double(sum % 1000) / 1000.0;
        for (uint32_t c = 0; c < it.ncand; c++) {
            it.scores[c] = 100.0 - (100.0 - it.quality[c]) * kq;
            it.settled[c] = 0;
        }
        work += double(it.w) * it.h * it.ncand;
        (void)ctx->arena.take(size_t(it.w) * it.h);
    }
    device_time(work);
    return RCG_OK;
}

size_t winner_bytes(const Item*, const int*, int n, bool) { return size_t(n) * 256 + 256; }

// The chosen candidates' "coefficients": each destination plane filled with the chosen quality, so a test can tell
// which candidate was written.
int encode_winners(rcg_ctx*, Arena&, Pinned&, const Item* items, const int* which, const int* chosen,
                   const WinnerDest* dest, int n, bool) {
    for (int k = 0; k < n; k++) {
        const Item& it = items[which[k]];
        const int16_t q = int16_t(it.quality[chosen[k]]);
        for (int c = 0; c < (it.ch == 1 ? 1 : 3); c++)
            if (dest[k].coefs[c]) dest[k].coefs[c][0] = q;
    }
    return RCG_OK;
}

}  // namespace rcg

// ------------------------------------------------------------------------------------------ the C entry points
extern "C" {
const char* rcg_last_error(void) { return g_error.c_str(); }

int rcg_create(int device, rcg_ctx** out) {
    if (!out) return RCG_INVALID;
    rcg_ctx* c = new rcg_ctx();
    c->device = device;
    *out = c;
    return RCG_OK;
}

void rcg_destroy(rcg_ctx* c) {
    if (!c) return;
    std::free(c->arena.base);
    std::free(c->pinned.base);
    delete c;
}

int rcg_set_tune(rcg_ctx* c, const rcg_tune*) { return c ? RCG_OK : RCG_INVALID; }
}
