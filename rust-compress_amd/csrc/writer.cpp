// The host side of the JPEG path, after and around the device search:
//   - entropy coding of a coefficient image: baseline with optimised Huffman tables (jpeg/src/encode.rs:240-413,
//     568-673), progressive with a planned scan script (jpeg/src/progressive.rs), and the smaller of the two
//     (write_best, encode.rs:543-565); canonical length-limited Huffman code lengths (core/src/util/huffman.rs:10-88);
//   - the trellis quantiser (jpeg/src/rdo.rs) and the trellis pass of build_coefs (encode.rs:527-537) over the
//     device's hand-off values;
//   - DCT-domain requantisation (CoefImage::requantize, encode.rs:107-171).
//
// The output is byte-identical to the reference's: the same symbol statistics, the same code-length construction and
// tie-breaking, the same segment order and the same bit packing. This is the stage after the device search: it
// serialises the coefficients the search chose (rcg_find_quality's coefs, or rcg_build_coefs') into a file.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <system_error>
#include <thread>
#include <tuple>
#include <vector>

#include "rcg_host.h"

namespace {

using Bytes = std::vector<uint8_t>;

// Natural-order index of zigzag position k (jpeg/src/lib.rs:15-19 holds the same table).
const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                             41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                             30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

inline uint32_t bit_length(uint32_t v) { return v == 0 ? 0u : 32u - uint32_t(__builtin_clz(v)); }

// ------------------------------------------------------------------------------------------------------ Huffman

// Length-limited Huffman code lengths (core/src/util/huffman.rs:10-88): a two-queue construction over the symbols
// sorted rarest first (ties: the higher symbol first, so it gets the longer code), depths clamped to max_len, the
// Kraft sum restored by moving codes down a level, and the lengths handed out longest-first in that order.
std::vector<uint8_t> build_lengths(const uint32_t* freqs, size_t nsym, int max_len) {
    std::vector<uint8_t> lengths(nsym, 0);
    std::vector<std::pair<uint32_t, size_t>> syms;
    for (size_t i = 0; i < nsym; i++)
        if (freqs[i] > 0) syms.emplace_back(freqs[i], i);
    if (syms.empty()) return lengths;
    if (syms.size() == 1) {
        lengths[syms[0].second] = 1;
        return lengths;
    }
    std::sort(syms.begin(), syms.end(), [](const std::pair<uint32_t, size_t>& a, const std::pair<uint32_t, size_t>& b) {
        return a.first != b.first ? a.first < b.first : a.second > b.second;
    });
    const size_t n = syms.size();
    std::vector<uint64_t> weight(2 * n - 1, 0);
    for (size_t i = 0; i < n; i++) weight[i] = syms[i].first;
    std::vector<size_t> parent(2 * n - 1, 0);
    size_t leaf = 0, node = n, next = n;
    auto pick = [&]() -> size_t {
        if (leaf < n && (node >= next || weight[leaf] <= weight[node])) return leaf++;
        return node++;
    };
    while (next < 2 * n - 1) {
        const size_t a = pick();
        const size_t b = pick();
        weight[next] = weight[a] + weight[b];
        parent[a] = next;
        parent[b] = next;
        next++;
    }
    std::vector<uint32_t> depth(2 * n - 1, 0);
    for (size_t i = 2 * n - 2; i-- > 0;) depth[i] = depth[parent[i]] + 1;
    const size_t maxl = size_t(max_len);
    std::vector<uint32_t> count(maxl + 1, 0);
    for (size_t i = 0; i < n; i++) count[std::min<size_t>(depth[i], maxl)]++;
    uint64_t total = 0;
    for (size_t l = 1; l <= maxl; l++) total += uint64_t(count[l]) << (maxl - l);
    while (total > (uint64_t(1) << maxl)) {
        count[maxl]--;
        for (size_t l = maxl - 1; l >= 1; l--) {
            if (count[l] > 0) {
                count[l]--;
                count[l + 1] += 2;
                break;
            }
        }
        total--;
    }
    size_t idx = 0;
    for (size_t l = maxl; l >= 1; l--)
        for (uint32_t k = 0; k < count[l]; k++) lengths[syms[idx++].second] = uint8_t(l);
    return lengths;
}

struct HuffTable {
    uint8_t bits[17];  // number of codes of each length 1..16 (index 0 unused)
    std::vector<uint8_t> vals;
    uint16_t code[256];
    uint8_t size[256];
};

// An optimal JPEG table for 256 symbol frequencies (encode.rs:240-265): a dummy symbol of frequency 1 reserves the
// all-ones code; codes are assigned canonically in (length, symbol) order.
HuffTable optimal_table(const uint32_t freq[256]) {
    uint32_t f[257];
    std::memcpy(f, freq, 256 * sizeof(uint32_t));
    f[256] = 1;
    const std::vector<uint8_t> lens = build_lengths(f, 257, 16);
    std::vector<std::pair<uint8_t, uint8_t>> syms;
    for (int s = 0; s < 256; s++)
        if (lens[s] > 0) syms.emplace_back(lens[s], uint8_t(s));
    if (syms.empty()) syms.emplace_back(1, 0);
    std::sort(syms.begin(), syms.end());
    HuffTable t;
    std::memset(t.bits, 0, sizeof(t.bits));
    std::memset(t.code, 0, sizeof(t.code));
    std::memset(t.size, 0, sizeof(t.size));
    uint32_t code = 0;
    uint8_t prev = syms[0].first;
    for (const auto& ls : syms) {
        code <<= (ls.first - prev);
        prev = ls.first;
        t.bits[ls.first]++;
        t.vals.push_back(ls.second);
        t.code[ls.second] = uint16_t(code);
        t.size[ls.second] = ls.first;
        code++;
    }
    return t;
}

// MSB-first bit packing with 0xFF byte stuffing; the last byte is padded with ones (encode.rs:289-318).
struct MsbWriter {
    Bytes out;
    uint64_t acc = 0;
    uint32_t n = 0;
    void put(uint32_t v, uint32_t len) {
        if (len == 0) return;
        acc = (acc << len) | (uint64_t(v) & ((uint64_t(1) << len) - 1));
        n += len;
        while (n >= 8) {
            const uint8_t byte = uint8_t(acc >> (n - 8));
            out.push_back(byte);
            if (byte == 0xFF) out.push_back(0);
            n -= 8;
        }
    }
    void finish() {
        if (n > 0) {
            const uint32_t pad = 8 - n;
            put((1u << pad) - 1, pad);
        }
    }
};

// ------------------------------------------------------------------------------------------------------ Image view

struct Comp {
    uint8_t id, h, v, tq;
    size_t blocks_w, blocks_h;
    const int16_t* coefs;
};

struct Coefs {
    size_t width, height;
    std::vector<Comp> comps;
    const uint8_t* qtables[4];
    int adobe_transform;
    size_t hmax() const {
        size_t m = 1;
        for (const Comp& c : comps) m = std::max<size_t>(m, c.h);
        return m;
    }
    size_t vmax() const {
        size_t m = 1;
        for (const Comp& c : comps) m = std::max<size_t>(m, c.v);
        return m;
    }
    size_t mcus_x() const { return (width + 8 * hmax() - 1) / (8 * hmax()); }
    size_t mcus_y() const { return (height + 8 * vmax() - 1) / (8 * vmax()); }
};

// ------------------------------------------------------------------------------------------------------ baseline

struct Stats {
    uint32_t dc[2][256];
    uint32_t ac[2][256];
    uint64_t raw_bits;
    Stats() { std::memset(this, 0, sizeof(*this)); }
    void on_dc(int tbl, uint32_t cat, uint32_t) {
        dc[tbl][cat]++;
        raw_bits += cat;
    }
    void on_ac(int tbl, uint8_t sym, uint32_t, uint32_t nbits) {
        ac[tbl][sym]++;
        raw_bits += nbits;
    }
};

struct Emitter {
    MsbWriter w;
    const HuffTable* dc;
    const HuffTable* ac;
    void on_dc(int tbl, uint32_t cat, uint32_t extra) {
        w.put(dc[tbl].code[cat], dc[tbl].size[cat]);
        w.put(extra, cat);
    }
    void on_ac(int tbl, uint8_t sym, uint32_t extra, uint32_t nbits) {
        w.put(ac[tbl].code[sym], ac[tbl].size[sym]);
        w.put(extra, nbits);
    }
};

// Magnitude category and the extra bits of a coefficient (encode.rs:337-343).
inline void category(int32_t v, uint32_t& cat, uint32_t& extra) {
    const uint32_t a = uint32_t(v < 0 ? -int64_t(v) : int64_t(v));
    cat = bit_length(a);
    const uint32_t mask = cat >= 32 ? 0xFFFFFFFFu : ((1u << cat) - 1);
    extra = (v < 0 ? uint32_t(v - 1) : uint32_t(v)) & mask;
}

// One block: the DC difference, then AC run/size symbols with ZRL and EOB (encode.rs:345-370).
template <class S>
void code_block(S& s, const int16_t* blk, int16_t& pred, int tbl) {
    const int32_t diff = int32_t(blk[0]) - int32_t(pred);
    pred = blk[0];
    uint32_t cat, extra;
    category(diff, cat, extra);
    s.on_dc(tbl, cat, extra);
    uint32_t run = 0;
    for (int k = 1; k < 64; k++) {
        const int16_t c = blk[k];
        if (c == 0) {
            run++;
            continue;
        }
        while (run >= 16) {
            s.on_ac(tbl, 0xF0, 0, 0);
            run -= 16;
        }
        category(c, cat, extra);
        s.on_ac(tbl, uint8_t((run << 4) | cat), extra, cat);
        run = 0;
    }
    if (run > 0) s.on_ac(tbl, 0x00, 0, 0);
}

// Interleaved MCU order over every component (encode.rs:376-396); `differential` codes DC without prediction.
template <class S>
void scan_all(const Coefs& ci, S& sink, bool differential) {
    int16_t pred[4] = {0, 0, 0, 0};
    const size_t mx_n = ci.mcus_x(), my_n = ci.mcus_y();
    for (size_t my = 0; my < my_n; my++)
        for (size_t mx = 0; mx < mx_n; mx++)
            for (size_t k = 0; k < ci.comps.size(); k++) {
                const Comp& c = ci.comps[k];
                for (size_t by = 0; by < c.v; by++)
                    for (size_t bx = 0; bx < c.h; bx++) {
                        const size_t b = (my * c.v + by) * c.blocks_w + mx * c.h + bx;
                        if (differential) pred[k] = 0;
                        code_block(sink, c.coefs + b * 64, pred[k], int(std::min<size_t>(k, 1)));
                    }
            }
}

void segment(Bytes& out, uint8_t marker, const Bytes& body) {
    out.push_back(0xFF);
    out.push_back(marker);
    const size_t n = body.size() + 2;
    out.push_back(uint8_t(n >> 8));
    out.push_back(uint8_t(n & 0xFF));
    out.insert(out.end(), body.begin(), body.end());
}

void write_dht(Bytes& out, uint8_t class_id, const HuffTable& t) {
    Bytes body;
    body.push_back(class_id);
    body.insert(body.end(), t.bits + 1, t.bits + 17);
    body.insert(body.end(), t.vals.begin(), t.vals.end());
    segment(out, 0xC4, body);
}

// SOI, JFIF, Adobe (4 components with a transform), DQT and SOF (encode.rs:647-673). The C ABI's coefficient image
// carries no descriptive metadata, so no EXIF/XMP/ICC segments are written.
Bytes write_headers(const Coefs& ci, uint8_t sof_marker) {
    Bytes out = {0xFF, 0xD8};
    segment(out, 0xE0, Bytes{'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0});
    if (ci.adobe_transform >= 0 && ci.comps.size() == 4)
        segment(out, 0xEE, Bytes{'A', 'd', 'o', 'b', 'e', 0, 100, 0, 0, 0, 0, uint8_t(ci.adobe_transform)});
    Bytes dqt;
    for (int slot = 0; slot < 4; slot++)
        if (ci.qtables[slot]) {
            dqt.push_back(uint8_t(slot));
            for (int k = 0; k < 64; k++) dqt.push_back(ci.qtables[slot][kZigzag[k]]);
        }
    segment(out, 0xDB, dqt);
    Bytes sof = {8, uint8_t(ci.height >> 8), uint8_t(ci.height & 0xFF), uint8_t(ci.width >> 8),
                 uint8_t(ci.width & 0xFF), uint8_t(ci.comps.size())};
    for (const Comp& c : ci.comps) {
        sof.push_back(c.id);
        sof.push_back(uint8_t((c.h << 4) | c.v));
        sof.push_back(c.tq);
    }
    segment(out, sof_marker, sof);
    return out;
}

// Baseline sequential JPEG with optimised tables (encode.rs:595-628).
Bytes write_baseline(const Coefs& ci, bool differential) {
    const bool gray = ci.comps.size() == 1;
    const int ntables = gray ? 1 : 2;
    Stats st;
    scan_all(ci, st, differential);
    const HuffTable dc[2] = {optimal_table(st.dc[0]), optimal_table(st.dc[1])};
    const HuffTable ac[2] = {optimal_table(st.ac[0]), optimal_table(st.ac[1])};
    Emitter em;
    em.w.out.reserve(ci.width * ci.height / 4 + 1024);
    em.dc = dc;
    em.ac = ac;
    scan_all(ci, em, differential);
    em.w.finish();
    Bytes out = write_headers(ci, differential ? 0xC5 : 0xC0);
    for (int t = 0; t < ntables; t++) {
        write_dht(out, uint8_t(t), dc[t]);
        write_dht(out, uint8_t(0x10 | t), ac[t]);
    }
    Bytes sos = {uint8_t(ci.comps.size())};
    for (size_t i = 0; i < ci.comps.size(); i++) {
        const uint8_t s = uint8_t(std::min<size_t>(i, 1));
        sos.push_back(ci.comps[i].id);
        sos.push_back(uint8_t((s << 4) | s));
    }
    sos.push_back(0);
    sos.push_back(63);
    sos.push_back(0);
    segment(out, 0xDA, sos);
    out.insert(out.end(), em.w.out.begin(), em.w.out.end());
    out.push_back(0xFF);
    out.push_back(0xD9);
    return out;
}

// Estimated size of the baseline file from its symbol statistics (encode.rs:549-566).
uint64_t baseline_estimate(const Coefs& ci) {
    Stats st;
    scan_all(ci, st, false);
    uint64_t bits = st.raw_bits, total = 0;
    const int ntab = ci.comps.size() == 1 ? 1 : 2;
    for (int cls = 0; cls < 2; cls++)
        for (int t = 0; t < ntab; t++) {
            const uint32_t* freq = cls == 0 ? st.dc[t] : st.ac[t];
            const HuffTable tab = optimal_table(freq);
            for (int s = 0; s < 256; s++) bits += uint64_t(freq[s]) * tab.size[s];
            total += 5 + 16 + tab.vals.size();
        }
    return total + (bits + 7) / 8 + 20;
}

// ------------------------------------------------------------------------------------------------------ progressive

struct Scan {
    std::vector<size_t> comps;
    size_t ss, se;
    uint32_t ah, al;
    bool operator<(const Scan& o) const {
        return std::tie(comps, ss, se, ah, al) < std::tie(o.comps, o.ss, o.se, o.ah, o.al);
    }
};

struct PStats {
    uint32_t freq[2][2][256];
    uint64_t raw_bits;
    PStats() { std::memset(this, 0, sizeof(*this)); }
    void sym(int cls, int slot, uint8_t s) { freq[cls][slot][s]++; }
    void bits(uint32_t, uint32_t n) { raw_bits += n; }
};

struct PTables {
    bool have[2][2] = {{false, false}, {false, false}};
    HuffTable t[2][2];
};

struct PEmit {
    MsbWriter w;
    const PTables* tables;
    void sym(int cls, int slot, uint8_t s) {
        const HuffTable& t = tables->t[cls][slot];
        w.put(t.code[s], t.size[s]);
    }
    void bits(uint32_t v, uint32_t n) { w.put(v, n); }
};

constexpr size_t kMaxCorrBits = 1000;

// Entropy-coder state of one scan (progressive.rs:53-63).
struct PState {
    int32_t last_dc[4] = {0, 0, 0, 0};
    bool differential = false;
    uint32_t eobrun = 0;
    std::vector<uint8_t> pending;  // correction bits owed to the pending end-of-band run
    std::vector<uint8_t> br;       // correction bits of the current block
};

template <class S>
void emit_eobrun(PState& st, S& sink) {
    if (st.eobrun == 0) return;
    const uint32_t nb = bit_length(st.eobrun) - 1;
    sink.sym(1, 0, uint8_t(nb << 4));
    if (nb > 0) sink.bits(st.eobrun & ((1u << nb) - 1), nb);
    st.eobrun = 0;
    for (uint8_t b : st.pending) sink.bits(b, 1);
    st.pending.clear();
}

template <class S>
void dc_first(PState& st, S& sink, const int16_t* blk, size_t comp, int slot, uint32_t al) {
    const int32_t v = int32_t(int16_t(blk[0] >> al));
    if (st.differential) st.last_dc[comp] = 0;
    const int32_t diff = v - st.last_dc[comp];
    st.last_dc[comp] = v;
    const uint32_t nb = bit_length(uint32_t(diff < 0 ? -diff : diff));
    sink.sym(0, slot, uint8_t(nb));
    if (nb > 0) sink.bits((diff < 0 ? uint32_t(diff - 1) : uint32_t(diff)) & ((1u << nb) - 1), nb);
}

template <class S>
void dc_refine(S& sink, const int16_t* blk, uint32_t al) {
    sink.bits(uint32_t((blk[0] >> al) & 1), 1);
}

template <class S>
void ac_first(PState& st, S& sink, const int16_t* blk, const Scan& sc) {
    uint32_t r = 0;
    for (size_t k = sc.ss; k <= sc.se; k++) {
        const int32_t v = blk[k];
        if (v == 0) {
            r++;
            continue;
        }
        int32_t mag, ext;
        if (v < 0) {
            mag = (-v) >> sc.al;
            ext = ~mag;
        } else {
            mag = v >> sc.al;
            ext = mag;
        }
        if (mag == 0) {
            r++;
            continue;
        }
        emit_eobrun(st, sink);
        while (r > 15) {
            sink.sym(1, 0, 0xF0);
            r -= 16;
        }
        const uint32_t nb = bit_length(uint32_t(mag));
        sink.sym(1, 0, uint8_t((r << 4) + nb));
        sink.bits(uint32_t(ext) & ((1u << nb) - 1), nb);
        r = 0;
    }
    if (r > 0) {
        st.eobrun++;
        if (st.eobrun == 0x7FFF) emit_eobrun(st, sink);
    }
}

template <class S>
void ac_refine(PState& st, S& sink, const int16_t* blk, const Scan& sc) {
    int32_t absv[64] = {0};
    size_t eob = 0;
    for (size_t k = sc.ss; k <= sc.se; k++) {
        const int32_t v = blk[k];
        const int32_t t = (v < 0 ? -v : v) >> sc.al;
        absv[k] = t;
        if (t == 1) eob = k;
    }
    uint32_t r = 0;
    st.br.clear();
    for (size_t k = sc.ss; k <= sc.se; k++) {
        const int32_t t = absv[k];
        if (t == 0) {
            r++;
            continue;
        }
        while (r > 15 && k <= eob) {
            emit_eobrun(st, sink);
            sink.sym(1, 0, 0xF0);
            r -= 16;
            for (uint8_t b : st.br) sink.bits(b, 1);
            st.br.clear();
        }
        if (t > 1) {
            st.br.push_back(uint8_t(t & 1));
            continue;
        }
        emit_eobrun(st, sink);
        sink.sym(1, 0, uint8_t((r << 4) + 1));
        sink.bits(blk[k] < 0 ? 0u : 1u, 1);
        for (uint8_t b : st.br) sink.bits(b, 1);
        st.br.clear();
        r = 0;
    }
    if (r > 0 || !st.br.empty()) {
        st.eobrun++;
        st.pending.insert(st.pending.end(), st.br.begin(), st.br.end());
        st.br.clear();
        if (st.eobrun == 0x7FFF || st.pending.size() > kMaxCorrBits - 63) emit_eobrun(st, sink);
    }
}

// One scan over the image (progressive.rs:191-244): interleaved MCU order for several components, otherwise only
// the blocks that overlap the image.
template <class S>
void run_scan(const Coefs& ci, const Scan& sc, S& sink, bool differential) {
    PState st;
    st.differential = differential;
    const size_t hmax = ci.hmax(), vmax = ci.vmax();
    const bool dc_scan = sc.ss == 0;
    auto visit = [&](size_t k, const int16_t* blk) {
        const int slot = int(std::min<size_t>(k, 1));
        if (dc_scan) {
            if (sc.ah == 0)
                dc_first(st, sink, blk, k, slot, sc.al);
            else
                dc_refine(sink, blk, sc.al);
        } else if (sc.ah == 0) {
            ac_first(st, sink, blk, sc);
        } else {
            ac_refine(st, sink, blk, sc);
        }
    };
    if (sc.comps.size() > 1) {
        const size_t mx_n = ci.mcus_x(), my_n = ci.mcus_y();
        for (size_t my = 0; my < my_n; my++)
            for (size_t mx = 0; mx < mx_n; mx++)
                for (size_t k : sc.comps) {
                    const Comp& c = ci.comps[k];
                    for (size_t by = 0; by < c.v; by++)
                        for (size_t bx = 0; bx < c.h; bx++) {
                            const size_t b = (my * c.v + by) * c.blocks_w + mx * c.h + bx;
                            visit(k, c.coefs + b * 64);
                        }
                }
    } else {
        const size_t k = sc.comps[0];
        const Comp& c = ci.comps[k];
        const size_t real_w = ((ci.width * c.h + hmax - 1) / hmax + 7) / 8;
        const size_t real_h = ((ci.height * c.v + vmax - 1) / vmax + 7) / 8;
        for (size_t by = 0; by < real_h; by++)
            for (size_t bx = 0; bx < real_w; bx++) visit(k, c.coefs + (by * c.blocks_w + bx) * 64);
    }
    if (!dc_scan) emit_eobrun(st, sink);
}

// The family of scan scripts searched by the plan (progressive.rs:246-295): DC precision, luma band split,
// luma/chroma successive-approximation depth, chroma band split.
std::vector<std::vector<Scan>> scripts(size_t ncomp) {
    auto s = [](std::vector<size_t> comps, size_t ss, size_t se, uint32_t ah, uint32_t al) {
        return Scan{std::move(comps), ss, se, ah, al};
    };
    std::vector<size_t> all;
    for (size_t i = 0; i < ncomp; i++) all.push_back(i);
    const std::vector<uint32_t> chroma_als = ncomp >= 3 ? std::vector<uint32_t>{0, 1} : std::vector<uint32_t>{0};
    const std::vector<int> chroma_splits = ncomp >= 3 ? std::vector<int>{-1, 2, 5} : std::vector<int>{-1};
    std::vector<std::vector<Scan>> out;
    for (uint32_t dc_al : {0u, 1u})
        for (int split : {-1, 2, 5, 8})
            for (uint32_t luma_al : {0u, 1u, 2u})
                for (uint32_t chroma_al : chroma_als)
                    for (int chroma_split : chroma_splits) {
                        std::vector<Scan> v = {s(all, 0, 0, 0, dc_al)};
                        if (split < 0) {
                            v.push_back(s({0}, 1, 63, 0, luma_al));
                        } else {
                            v.push_back(s({0}, 1, size_t(split), 0, luma_al));
                            v.push_back(s({0}, size_t(split) + 1, 63, 0, luma_al));
                        }
                        if (ncomp >= 3)
                            for (size_t c = 1; c < ncomp; c++) {
                                if (chroma_split < 0) {
                                    v.push_back(s({c}, 1, 63, 0, chroma_al));
                                } else {
                                    v.push_back(s({c}, 1, size_t(chroma_split), 0, chroma_al));
                                    v.push_back(s({c}, size_t(chroma_split) + 1, 63, 0, chroma_al));
                                }
                            }
                        if (dc_al > 0) v.push_back(s(all, 0, 0, 1, 0));
                        for (uint32_t ah = luma_al; ah >= 1; ah--) v.push_back(s({0}, 1, 63, ah, ah - 1));
                        for (uint32_t ah = chroma_al; ah >= 1; ah--)
                            for (size_t c = 1; c < ncomp; c++) v.push_back(s({c}, 1, 63, ah, ah - 1));
                        out.push_back(std::move(v));
                    }
    return out;
}

PTables build_tables(const PStats& st) {
    PTables t;
    for (int cls = 0; cls < 2; cls++)
        for (int slot = 0; slot < 2; slot++) {
            bool any = false;
            for (int s = 0; s < 256 && !any; s++) any = st.freq[cls][slot][s] > 0;
            if (any) {
                t.have[cls][slot] = true;
                t.t[cls][slot] = optimal_table(st.freq[cls][slot]);
            }
        }
    return t;
}

// Estimated size of one scan: DHT segments, SOS and entropy data (progressive.rs:310-329).
uint64_t scan_cost(const Coefs& ci, const Scan& sc) {
    PStats st;
    run_scan(ci, sc, st, false);
    const PTables t = build_tables(st);
    uint64_t bits = st.raw_bits, total = 0;
    for (int cls = 0; cls < 2; cls++)
        for (int slot = 0; slot < 2; slot++)
            if (t.have[cls][slot]) {
                for (int s = 0; s < 256; s++) bits += uint64_t(st.freq[cls][slot][s]) * t.t[cls][slot].size[s];
                total += 5 + 16 + t.t[cls][slot].vals.size();
            }
    return total + (bits + 7) / 8 + 14;
}

// The cheapest script and its estimated size (progressive.rs:337-352). Every distinct scan is costed once, on host
// threads; scripts are compared by their summed costs and the first cheapest wins.
std::pair<uint64_t, std::vector<Scan>> plan(const Coefs& ci) {
    const std::vector<std::vector<Scan>> cands = scripts(ci.comps.size());
    std::vector<Scan> distinct;
    std::map<Scan, size_t> index;
    for (const auto& script : cands)
        for (const Scan& sc : script)
            if (index.emplace(sc, distinct.size()).second) distinct.push_back(sc);
    std::vector<uint64_t> cost(distinct.size(), 0);
    const size_t nthreads = std::max<size_t>(1, std::min<size_t>(distinct.size(), std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nthreads; t++)
        pool.emplace_back([&, t] {
            for (size_t i = t; i < distinct.size(); i += nthreads) cost[i] = scan_cost(ci, distinct[i]);
        });
    for (std::thread& th : pool) th.join();
    size_t best = 0;
    uint64_t best_total = UINT64_MAX;
    for (size_t i = 0; i < cands.size(); i++) {
        uint64_t total = 0;
        for (const Scan& sc : cands[i]) total += cost[index[sc]];
        if (total < best_total) {
            best_total = total;
            best = i;
        }
    }
    return {best_total, cands[best]};
}

// Every scan of a script with its own optimised tables (progressive.rs:354-384), framed as SOF2 (or SOF6).
Bytes emit_progressive(const Coefs& ci, const std::vector<Scan>& script, bool differential) {
    Bytes body;
    for (const Scan& sc : script) {
        PStats st;
        run_scan(ci, sc, st, differential);
        const PTables tables = build_tables(st);
        for (int cls = 0; cls < 2; cls++)
            for (int slot = 0; slot < 2; slot++)
                if (tables.have[cls][slot]) write_dht(body, uint8_t((cls << 4) | slot), tables.t[cls][slot]);
        Bytes sos = {uint8_t(sc.comps.size())};
        for (size_t k : sc.comps) {
            const uint8_t slot = uint8_t(std::min<size_t>(k, 1));
            const uint8_t td = sc.ss == 0 ? slot : 0;
            sos.push_back(ci.comps[k].id);
            sos.push_back(uint8_t(td << 4));
        }
        sos.push_back(uint8_t(sc.ss));
        sos.push_back(uint8_t(sc.se));
        sos.push_back(uint8_t((sc.ah << 4) | sc.al));
        segment(body, 0xDA, sos);
        PEmit em;
        em.tables = &tables;
        run_scan(ci, sc, em, differential);
        em.w.finish();
        body.insert(body.end(), em.w.out.begin(), em.w.out.end());
    }
    Bytes out = write_headers(ci, differential ? 0xC6 : 0xC2);
    out.insert(out.end(), body.begin(), body.end());
    out.push_back(0xFF);
    out.push_back(0xD9);
    return out;
}

int view(const rcg_coef_image* ci, Coefs& out) {
    if (!ci || ci->ncomp < 1 || ci->ncomp > 4 || ci->width == 0 || ci->height == 0 || ci->width > 65535 ||
        ci->height > 65535) {
        rcg::set_error("coefficient image: 1-4 components and sides of 1..65535 pixels");
        return RCG_INVALID;
    }
    out.width = ci->width;
    out.height = ci->height;
    out.adobe_transform = ci->adobe_transform;
    for (int s = 0; s < 4; s++) out.qtables[s] = ci->qtables[s];
    size_t hmax = 1, vmax = 1;
    for (uint32_t c = 0; c < ci->ncomp; c++) {
        const rcg_coef_component& k = ci->comps[c];
        if (!k.coefs || k.h < 1 || k.h > 4 || k.v < 1 || k.v > 4 || k.tq > 3) {
            rcg::set_error("coefficient image: a component needs coefficients, sampling 1..4 and a table slot 0..3");
            return RCG_INVALID;
        }
        hmax = std::max<size_t>(hmax, k.h);
        vmax = std::max<size_t>(vmax, k.v);
        out.comps.push_back(Comp{k.id, k.h, k.v, k.tq, k.blocks_w, k.blocks_h, k.coefs});
    }
    // Every MCU the scans visit must lie inside each component's block grid.
    const size_t mx = (out.width + 8 * hmax - 1) / (8 * hmax), my = (out.height + 8 * vmax - 1) / (8 * vmax);
    for (const Comp& c : out.comps)
        if (c.blocks_w < mx * c.h || c.blocks_h < my * c.v) {
            rcg::set_error("coefficient image: a component's block grid is smaller than its MCUs");
            return RCG_INVALID;
        }
    for (const Comp& c : out.comps)
        if (!out.qtables[c.tq]) {
            rcg::set_error("coefficient image: a component refers to a missing quantization table");
            return RCG_INVALID;
        }
    return RCG_OK;
}

// ------------------------------------------------------------------------------------------------------ trellis

// Round to nearest, ties to even, through the magic-constant add of jpeg/src/dct.rs:126-129 (|x| < 2^22).
inline int32_t round_f32(float x) {
    const float y = x + 12582912.0f;
    int32_t bits;
    std::memcpy(&bits, &y, 4);
    return bits - 0x4B400000;
}

// Rate-distortion quantiser of one component class (jpeg/src/rdo.rs:15-114). Every f32 expression keeps the
// reference's operation order (this file is compiled without FP contraction), so its choices are the reference's.
struct Rdo {
    float q[64];       // quantisation step per zigzag position
    float lambda;
    float w[64];       // distortion weight per zigzag position
    float len[256];    // code length of each AC symbol; 18 for a symbol the first pass never used

    Rdo(const uint8_t table[64], float kappa, const uint8_t sizes[256]) {
        for (int k = 0; k < 64; k++) q[k] = float(table[kZigzag[k]]);
        float sorted[63];
        std::memcpy(sorted, q + 1, sizeof(sorted));
        std::sort(sorted, sorted + 63);
        const float median = sorted[63 / 2];
        for (int s = 0; s < 256; s++) len[s] = sizes[s] == 0 ? 18.0f : float(sizes[s]);
        w[0] = 1.0f;
        // (median / q)^gamma with gamma = 1 (rdo.rs:12): the quotient itself.
        for (int k = 1; k < 64; k++) w[k] = median / q[k];
        lambda = kappa * median * median;
    }

    float rate(size_t run, uint32_t v) const {
        const uint32_t size = bit_length(v);
        const size_t zrl = run / 16;
        return float(zrl) * len[0xF0] + len[((run % 16) << 4) | size] + float(size);
    }

    // Positions 1..63 of a (values in quantisation steps); DC is the caller's (rdo.rs:52-113).
    void quantize_ac(const float a[64], int16_t out[64]) const {
        struct Node {
            size_t pos;
            uint32_t v;
            float cost;
            size_t prev;
        };
        const size_t START = SIZE_MAX;
        float zd[64];
        zd[0] = 0.0f;
        for (int k = 1; k < 64; k++) {
            const float e = a[k] * q[k];
            zd[k] = zd[k - 1] + e * e * w[k];
        }
        Node nodes[126];
        size_t nn = 0;
        for (size_t i = 1; i < 64; i++) {
            const float m = std::fabs(a[i]);
            const uint32_t v0 = uint32_t(m + 0.5f);
            if (v0 == 0) continue;
            const uint32_t lo = v0 > 1 ? v0 - 1 : v0;
            for (uint32_t v = lo; v <= v0; v++) {
                const float e = (m - float(v)) * q[i];
                const float dnz = e * e * w[i];
                float best = zd[i - 1] + lambda * rate(i - 1, v) + dnz;
                size_t prev = START;
                for (size_t ni = 0; ni < nn; ni++) {
                    const Node& n = nodes[ni];
                    if (n.pos >= i) continue;
                    const float c = n.cost + (zd[i - 1] - zd[n.pos]) + lambda * rate(i - n.pos - 1, v) + dnz;
                    if (c < best) {
                        best = c;
                        prev = ni;
                    }
                }
                nodes[nn++] = Node{i, v, best, prev};
            }
        }
        const float eob = lambda * len[0];
        float best_cost = zd[63] + eob;
        size_t best_end = START;
        for (size_t ni = 0; ni < nn; ni++) {
            const Node& n = nodes[ni];
            const float c = n.cost + (zd[63] - zd[n.pos]) + (n.pos < 63 ? eob : 0.0f);
            if (c < best_cost) {
                best_cost = c;
                best_end = ni;
            }
        }
        for (int k = 1; k < 64; k++) out[k] = 0;
        for (size_t cur = best_end; cur != START; cur = nodes[cur].prev) {
            const Node& n = nodes[cur];
            const int16_t v = int16_t(n.v);
            out[n.pos] = a[n.pos] < 0.0f ? int16_t(-v) : v;
        }
    }
};

constexpr float kTrellisKappa = 0.02f;  // TRELLIS_KAPPA, jpeg/src/encode.rs:42

// AC code lengths of the plainly quantised coefficients, luma and chroma (encode.rs:222-227): the rate model.
void ac_code_lengths(const Coefs& ci, uint8_t sizes[2][256]) {
    Stats st;
    scan_all(ci, st, false);
    for (int t = 0; t < 2; t++) {
        const HuffTable tab = optimal_table(st.ac[t]);
        std::memcpy(sizes[t], tab.size, 256);
    }
}

// The trellis pass of build_coefs (encode.rs:527-537) over the scaled values of the plain pass.
void trellis(const Coefs& plain, const float* const scaled[4], int16_t* const out[4]) {
    uint8_t sizes[2][256];
    ac_code_lengths(plain, sizes);
    const uint8_t* tabs[2] = {plain.qtables[0], plain.qtables[1] ? plain.qtables[1] : plain.qtables[0]};
    const Rdo rdo[2] = {Rdo(tabs[0], kTrellisKappa, sizes[0]), Rdo(tabs[1], kTrellisKappa, sizes[1])};
    for (size_t c = 0; c < plain.comps.size(); c++) {
        const Comp& k = plain.comps[c];
        const Rdo& r = rdo[std::min<int>(k.tq, 1)];
        const size_t nblocks = k.blocks_w * k.blocks_h;
        const size_t nth = std::max<size_t>(1, std::min<size_t>(std::thread::hardware_concurrency(), nblocks / 256 + 1));
        std::vector<std::thread> pool;
        for (size_t t = 0; t < nth; t++)
            pool.emplace_back([&, t] {
                for (size_t b = t; b < nblocks; b += nth) {
                    const float* a = scaled[c] + b * 64;
                    int16_t* o = out[c] + b * 64;
                    const int32_t dc = round_f32(a[0]);
                    o[0] = int16_t(dc < -2047 ? -2047 : (dc > 2047 ? 2047 : dc));
                    r.quantize_ac(a, o);
                }
            });
        for (std::thread& th : pool) th.join();
    }
}

// CoefImage::requantize (encode.rs:111-171) on the host: the plain requantisation of every block and, with trellis
// and a changed table, the trellis re-quantisation of every block from its exact source values a = c * ratio.
void requantize_host(const Coefs& src, const rcg::RequantPlan& P, bool use_trellis, int16_t* const out[4]) {
    for (size_t c = 0; c < src.comps.size(); c++) {
        const Comp& k = src.comps[c];
        const size_t n = k.blocks_w * k.blocks_h * 64;
        for (size_t i = 0; i < n; i++) {
            const int32_t r = round_f32(float(k.coefs[i]) * P.ratio[c][i % 64]);
            out[c][i] = int16_t(r < -2047 ? -2047 : (r > 2047 ? 2047 : r));
        }
    }
    if (!use_trellis || !P.changed) return;
    // The rate model: AC code lengths of the plainly requantised image, per table class min(c, 1).
    Coefs plain = src;
    for (size_t c = 0; c < plain.comps.size(); c++) {
        plain.comps[c].coefs = out[c];
        plain.comps[c].tq = P.tq[c];
    }
    uint8_t sizes[2][256];
    ac_code_lengths(plain, sizes);
    for (size_t c = 0; c < src.comps.size(); c++) {
        const Rdo rdo(P.table[c], kTrellisKappa, sizes[std::min<size_t>(c, 1)]);
        const Comp& k = src.comps[c];
        const size_t nblocks = k.blocks_w * k.blocks_h;
        const size_t nth = std::max<size_t>(1, std::min<size_t>(std::thread::hardware_concurrency(), nblocks / 256 + 1));
        std::vector<std::thread> pool;
        for (size_t t = 0; t < nth; t++)
            pool.emplace_back([&, t, c] {
                for (size_t b = t; b < nblocks; b += nth) {
                    float a[64];
                    for (int i = 0; i < 64; i++) a[i] = float(k.coefs[b * 64 + i]) * P.ratio[c][i];
                    rdo.quantize_ac(a, out[c] + b * 64);
                }
            });
        for (std::thread& th : pool) th.join();
    }
}

int hand_over(const Bytes& b, uint8_t** out, size_t* len) {
    uint8_t* p = static_cast<uint8_t*>(std::malloc(b.size()));
    if (!p) return RCG_OOM;
    std::memcpy(p, b.data(), b.size());
    *out = p;
    *len = b.size();
    return RCG_OK;
}

}  // namespace

extern "C" int rcg_write_jpeg(const rcg_coef_image* ci, int mode, uint8_t** out, size_t* len) {
    return rcg::guarded([&]() -> int {
        if (!out || !len) return RCG_INVALID;
        *out = nullptr;
        *len = 0;
        Coefs c;
        RCG_TRY(view(ci, c));
        try {
            switch (mode) {
                case RCG_WRITE_BASELINE:
                    return hand_over(write_baseline(c, false), out, len);
                case RCG_WRITE_PROGRESSIVE:
                    return hand_over(emit_progressive(c, plan(c).second, false), out, len);
                case RCG_WRITE_BEST: {
                    // encode.rs:543-547: both sizes are estimated first; only the winner is written.
                    const auto p = plan(c);
                    if (baseline_estimate(c) <= p.first) return hand_over(write_baseline(c, false), out, len);
                    return hand_over(emit_progressive(c, p.second, false), out, len);
                }
                default:
                    rcg::set_error("unknown write mode");
                    return RCG_INVALID;
            }
        } catch (const std::bad_alloc&) {
            rcg::set_error("out of host memory");
            return RCG_OOM;
        }
    });
}

extern "C" void rcg_free(void* p) {
    rcg::guarded_void([&] {
        std::free(p);
    });
}

extern "C" int rcg_huffman_lengths(const uint32_t* freqs, size_t n, int max_len, uint8_t* lengths) {
    return rcg::guarded([&]() -> int {
        if ((!freqs && n) || (!lengths && n) || max_len < 1 || max_len > 32) return RCG_INVALID;
        try {
            const std::vector<uint8_t> l = build_lengths(freqs, n, max_len);
            if (n) std::memcpy(lengths, l.data(), n);
        } catch (const std::bad_alloc&) {
            return RCG_OOM;
        }
        return RCG_OK;
    });
}

extern "C" int rcg_trellis(const rcg_coef_image* plain, const float* const scaled[4], int16_t* out[4]) {
    return rcg::guarded([&]() -> int {
        if (!plain || !scaled || !out) return RCG_INVALID;
        Coefs c;
        RCG_TRY(view(plain, c));
        if (c.comps.size() > 3 || !c.qtables[0]) {
            rcg::set_error("trellis: 1 or 3 components with a luma table in slot 0");
            return RCG_INVALID;
        }
        for (size_t k = 0; k < c.comps.size(); k++)
            if (!scaled[k] || !out[k]) return RCG_INVALID;
        try {
            trellis(c, scaled, out);
        } catch (const std::bad_alloc&) {
            return RCG_OOM;
        } catch (const std::system_error&) {
            rcg::set_error("trellis: could not start host threads");
            return RCG_DEVICE;
        }
        return RCG_OK;
    });
}

extern "C" int rcg_rdo_quantize_ac(const uint8_t table[64], float kappa, const uint8_t sizes[256], const float a[64],
                                   int16_t out[64]) {
    return rcg::guarded([&]() -> int {
        if (!table || !sizes || !a || !out) return RCG_INVALID;
        const Rdo r(table, kappa, sizes);
        r.quantize_ac(a, out);
        return RCG_OK;
    });
}

namespace rcg {

int requant_plan(const rcg_coef_image* ci, int quality, RequantPlan& P) {
    std::memset(&P, 0, sizeof(P));
    if (!ci || ci->ncomp < 1 || ci->ncomp > 4 || quality < 1 || quality > 100) {
        set_error("requantize: 1-4 components and a quality of 1..100");
        return RCG_INVALID;
    }
    P.ncomp = ci->ncomp;
    uint8_t std_tab[2][64];
    quant_tables(quality, std_tab[0], std_tab[1]);
    for (uint32_t c = 0; c < ci->ncomp; c++) {
        const rcg_coef_component& k = ci->comps[c];
        if (k.tq > 3) {
            set_error("requantize: table slot out of range");
            return RCG_INVALID;
        }
        // Table class (encode.rs:96-104): in a 4-component file only YCCK's colour-difference channels are chroma.
        const int cls = ci->ncomp == 4 ? int(ci->adobe_transform == 2 && (c == 1 || c == 2)) : int(std::min(c, 1u));
        uint8_t old[64];
        if (ci->qtables[k.tq])
            std::memcpy(old, ci->qtables[k.tq], 64);
        else
            std::memset(old, 1, 64);
        bool differs = false;
        for (int n = 0; n < 64; n++) {
            P.table[c][n] = std::max(std_tab[cls][n], old[n]);
            differs |= P.table[c][n] != old[n];
        }
        P.changed |= differs;
        for (int z = 0; z < 64; z++) P.ratio[c][z] = float(old[kZigzag[z]]) / float(P.table[c][kZigzag[z]]);
        // Components with the same table share a slot; a different table takes the next free one.
        int slot = -1;
        for (int s = 0; s < 4 && slot < 0; s++)
            if (P.present[s] && std::memcmp(P.slots[s], P.table[c], 64) == 0) slot = s;
        if (slot < 0) {
            for (int s = 0; s < 4 && slot < 0; s++)
                if (!P.present[s]) slot = s;
            P.present[slot] = 1;
            std::memcpy(P.slots[slot], P.table[c], 64);
        }
        P.tq[c] = uint8_t(slot);
    }
    return RCG_OK;
}

}  // namespace rcg

extern "C" int rcg_requantize(const rcg_coef_image* ci, int quality, int trellis_on, rcg_requantized* out) {
    return rcg::guarded([&]() -> int {
        if (!ci || !out) return RCG_INVALID;
        rcg::RequantPlan P;
        RCG_TRY(rcg::requant_plan(ci, quality, P));
        // The coefficient grid and sampling checks of the writer's view, without requiring the tables it will replace.
        rcg_coef_image probe = *ci;
        uint8_t ones[64];
        std::memset(ones, 1, 64);
        for (int s = 0; s < 4; s++)
            if (!probe.qtables[s]) probe.qtables[s] = ones;
        Coefs c;
        RCG_TRY(view(&probe, c));
        int16_t* outs[4] = {nullptr, nullptr, nullptr, nullptr};
        for (uint32_t k = 0; k < ci->ncomp; k++) {
            if (!out->coefs[k]) return RCG_INVALID;
            outs[k] = out->coefs[k];
        }
        try {
            requantize_host(c, P, trellis_on != 0, outs);
        } catch (const std::bad_alloc&) {
            return RCG_OOM;
        } catch (const std::system_error&) {
            rcg::set_error("requantize: could not start host threads");
            return RCG_DEVICE;
        }
        for (int s = 0; s < 4; s++) {
            out->present[s] = P.present[s];
            std::memcpy(out->qtables[s], P.slots[s], 64);
        }
        for (uint32_t k = 0; k < 4; k++) out->tq[k] = k < ci->ncomp ? P.tq[k] : 0;
        return RCG_OK;
    });
}
