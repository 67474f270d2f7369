// The JPEG reader in front of the DCT-domain path: the frame facts that choose between requantisation and a pixel
// round trip (rc_jpeg::decode::Parsed::info, jpeg/src/decode.rs:311-331,447-489), the EXIF orientation
// (decode.rs:572-620), and the quantised coefficients of a sequential or progressive Huffman-coded 8-bit DCT frame
// (Parsed::into_coefs, decode.rs:529-569). Arithmetic-coded, lossless and hierarchical frames are described but not
// decoded here (RCG_UNSUPPORTED); 12-bit frames likewise.
//
// Entropy decoding follows ITU-T T.81 Annex F (sequential) and Annex G (progressive spectral selection and
// successive approximation), with the reference's conventions wherever the standard leaves room: the bit reader
// feeds zero bits at a marker or at the end of the data; a restart discards buffered bits and consumes the next RSTn;
// non-interleaved scans cover each component's own block grid (ceil of its sample size over 8), interleaved ones the
// MCU grid; coefficients are stored in zigzag order in the MCU-padded block grid; the table a component uses is the
// one in effect at the first scan that contains it.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "rcg_host.h"

namespace {

using rcg::set_error;

const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                             41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                             30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// ITU-T T.81 Annex K table K.1, natural order: the base of the IJG quality estimate (jpeg/src/quality.rs:17-22).
const uint8_t kAnnexKLuma[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                                 14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                                 18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                                 49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};

constexpr uint64_t kMaxPixels = 150000000ull;  // rc_core::image::MAX_PIXELS (core/src/image.rs:33)
constexpr int kMaxScans = 1000;                 // decode.rs:283

struct Fail {
    int status;
    std::string msg;
};
[[noreturn]] void invalid(const char* m) { throw Fail{RCG_INVALID, m}; }
[[noreturn]] void unsupported(const std::string& m) { throw Fail{RCG_UNSUPPORTED, m}; }

// The mean ratio of the table to Annex K, mapped back through the IJG scaling (quality.rs:17-22).
int estimate_quality(const uint8_t t[64]) {
    double s = 0.0;
    for (int k = 0; k < 64; k++) s += double(t[k]) * 100.0 / double(kAnnexKLuma[k]);
    s /= 64.0;
    const double q = s <= 100.0 ? (200.0 - s) / 2.0 : 5000.0 / s;
    const double r = std::round(q);
    return int(r < 1.0 ? 1.0 : (r > 100.0 ? 100.0 : r));
}

// MSB-first bit reader over entropy-coded data with byte unstuffing.
struct Bits {
    const uint8_t* d;
    size_t n, pos;
    uint64_t acc = 0;
    uint32_t have = 0;
    bool marker = false;
    Bits(const uint8_t* data, size_t len, size_t at) : d(data), n(len), pos(at) {}

    void fill() {
        while (have <= 56) {
            uint32_t byte = 0;
            if (!marker) {
                if (pos >= n) {
                    marker = true;
                } else if (d[pos] == 0xFF) {
                    if (pos + 1 < n && d[pos + 1] == 0) {
                        byte = 0xFF;
                        pos += 2;
                    } else {
                        marker = true;
                    }
                } else {
                    byte = d[pos++];
                }
            }
            acc = (acc << 8) | byte;
            have += 8;
        }
    }
    uint32_t peek(uint32_t k) {
        if (have < k) fill();
        return uint32_t((acc >> (have - k)) & ((1ull << k) - 1));
    }
    void skip(uint32_t k) { have -= k; }
    uint32_t get(uint32_t k) {
        if (k == 0) return 0;
        const uint32_t v = peek(k);
        skip(k);
        return v;
    }
    // Drop buffered bits, pass over anything up to the next marker and consume it if it is RSTn.
    void restart() {
        acc = 0;
        have = 0;
        marker = false;
        while (pos + 1 < n) {
            if (d[pos] != 0xFF) {
                pos++;
                continue;
            }
            const uint8_t m = d[pos + 1];
            if (m == 0) {
                pos += 2;
            } else if (m >= 0xD0 && m <= 0xD7) {
                pos += 2;
                return;
            } else {
                return;
            }
        }
    }
};

// Canonical Huffman decoder (T.81 F.2.2.3) with a 9-bit first-level lookup.
struct Huff {
    static constexpr uint32_t kFast = 9;
    uint16_t fast[1 << kFast];  // (length << 8) | symbol, 0 when the code is longer
    int32_t mincode[17], maxcode[17], valptr[17];
    uint8_t vals[256];
    bool defined = false;

    void build(const uint8_t bits[16], const uint8_t* v, size_t total) {
        std::memset(fast, 0, sizeof(fast));
        std::memcpy(vals, v, total);
        int32_t code = 0;
        size_t k = 0;
        for (int len = 1; len <= 16; len++) {
            const int32_t cnt = bits[len - 1];
            valptr[len] = int32_t(k);
            mincode[len] = code;
            maxcode[len] = -1;
            if (cnt > 0) {
                if (code + cnt > (1 << len)) invalid("oversubscribed JPEG huffman table");
                if (uint32_t(len) <= kFast)
                    for (int32_t i = 0; i < cnt; i++) {
                        const uint32_t shift = kFast - uint32_t(len);
                        const uint32_t c = uint32_t(code + i) << shift;
                        for (uint32_t f = 0; f < (1u << shift); f++)
                            fast[c | f] = uint16_t((len << 8) | vals[k + size_t(i)]);
                    }
                code += cnt;
                k += size_t(cnt);
                maxcode[len] = code - 1;
            }
            code <<= 1;
        }
        defined = true;
    }
    uint8_t decode(Bits& b) const {
        const uint16_t e = fast[b.peek(kFast)];
        if (e) {
            b.skip(e >> 8);
            return uint8_t(e);
        }
        for (int len = 1; len <= 16; len++) {
            const int32_t c = int32_t(b.peek(uint32_t(len)));
            if (maxcode[len] >= 0 && c <= maxcode[len] && c >= mincode[len]) {
                b.skip(uint32_t(len));
                return vals[valptr[len] + c - mincode[len]];
            }
        }
        invalid("bad JPEG huffman code");
    }
};

inline int32_t extend(uint32_t v, uint32_t t) {
    if (t == 0) return 0;
    return v < (1u << (t - 1)) ? int32_t(v) - (1 << t) + 1 : int32_t(v);
}

struct Comp {
    uint8_t id, h, v, tq;
    uint32_t blocks_w, blocks_h, real_bw, real_bh;
    bool table_fixed = false;
    uint16_t table[64];  // natural order, captured at the first scan holding the component
    int16_t* coefs = nullptr;
};

struct ScanComp {
    size_t idx;
    int dc, ac;
};

struct Reader {
    const uint8_t* d;
    size_t n, pos = 2;
    bool decode;                 // entropy-decode scans (false: skip them to the next marker)
    int16_t* const* out = nullptr;
    uint16_t qt[4][64];
    bool qt_set[4] = {false, false, false, false};
    Huff dc[4], ac[4];
    std::vector<Comp> comps;
    uint32_t width = 0, height = 0, hmax = 1, vmax = 1, mcus_x = 0, mcus_y = 0;
    int precision = 8, adobe = -1, scans = 0, restart_interval = 0;
    bool progressive = false, arith = false, lossless = false, hier = false, sof_seen = false;
    // DHP (hierarchical) frame description.
    uint32_t dhp_w = 0, dhp_h = 0;
    int dhp_precision = 8;
    std::vector<std::pair<uint8_t, uint8_t>> dhp_sampling;

    Reader(const uint8_t* data, size_t len, bool dec) : d(data), n(len), decode(dec) {}

    uint32_t be16(size_t at) const {
        if (at + 2 > n) invalid("truncated JPEG");
        return uint32_t(d[at]) << 8 | d[at + 1];
    }

    uint8_t next_marker() {
        for (;;) {
            while (pos < n && d[pos] != 0xFF) pos++;
            while (pos < n && d[pos] == 0xFF) pos++;
            if (pos >= n) invalid("unexpected end of JPEG");
            const uint8_t m = d[pos++];
            if (m != 0 && !(m >= 0xD0 && m <= 0xD7)) return m;
        }
    }

    void run() {
        for (;;) {
            const uint8_t m = next_marker();
            if (m == 0xD9) return;
            if (m == 0x01) continue;
            const size_t len = be16(pos);
            if (len < 2 || pos + len > n) {
                // A truncated segment: keep what was decoded if there was a scan (decode.rs:651-654).
                if (scans > 0) return;
                invalid("truncated JPEG segment");
            }
            const uint8_t* body = d + pos + 2;
            const size_t blen = len - 2;
            pos += len;
            switch (m) {
                case 0xDB: dqt(body, blen); break;
                case 0xC4: dht(body, blen); break;
                case 0xDD:
                    if (blen < 2) invalid("bad DRI");
                    restart_interval = int(uint32_t(body[0]) << 8 | body[1]);
                    break;
                case 0xDE: dhp(body, blen); break;
                case 0xEE:
                    if (blen >= 12 && std::memcmp(body, "Adobe", 5) == 0) adobe = body[11];
                    break;
                case 0xDA:
                    if (hier) return;  // frames of a hierarchical image are not decoded here
                    if (scans >= kMaxScans) unsupported("too many scans in a JPEG frame");
                    sos(body, blen);
                    scans++;
                    break;
                default:
                    if ((m >= 0xC0 && m <= 0xC3) || (m >= 0xC5 && m <= 0xC7) || (m >= 0xC9 && m <= 0xCB) ||
                        (m >= 0xCD && m <= 0xCF)) {
                        sof(m, body, blen);
                        if (hier) return;
                    }
                    break;
            }
        }
    }

    void dqt(const uint8_t* b, size_t len) {
        while (len > 0) {
            const uint32_t pq = b[0] >> 4, tq = b[0] & 15;
            if (tq > 3 || pq > 1) invalid("bad DQT");
            const size_t need = 1 + 64 * (pq + 1);
            if (len < need) invalid("truncated DQT");
            for (int k = 0; k < 64; k++)
                qt[tq][kZigzag[k]] = pq == 0 ? b[1 + k] : uint16_t(uint32_t(b[1 + 2 * k]) << 8 | b[2 + 2 * k]);
            qt_set[tq] = true;
            b += need;
            len -= need;
        }
    }

    void dht(const uint8_t* b, size_t len) {
        while (len > 0) {
            if (len < 17) invalid("truncated DHT");
            const uint32_t tc = b[0] >> 4, th = b[0] & 15;
            if (tc > 1 || th > 3) invalid("bad DHT");
            size_t total = 0;
            for (int i = 0; i < 16; i++) total += b[1 + i];
            if (total > 256) invalid("bad DHT symbol count");
            if (len < 17 + total) invalid("truncated DHT");
            (tc == 0 ? dc : ac)[th].build(b + 1, b + 17, total);
            b += 17 + total;
            len -= 17 + total;
        }
    }

    void dhp(const uint8_t* b, size_t len) {
        if (hier || sof_seen) invalid("DHP must be the first frame-level marker, once");
        if (len < 6) invalid("bad DHP");
        dhp_precision = b[0];
        dhp_h = uint32_t(b[1]) << 8 | b[2];
        dhp_w = uint32_t(b[3]) << 8 | b[4];
        const size_t nc = b[5];
        if (dhp_w == 0 || dhp_h == 0) invalid("zero hierarchical JPEG dimension");
        if (uint64_t(dhp_w) * dhp_h > kMaxPixels) invalid("unreasonable image dimensions");
        if (dhp_precision < 2 || dhp_precision > 16) unsupported(std::to_string(dhp_precision) + "-bit JPEG");
        if (nc != 1 && nc != 3 && nc != 4) unsupported(std::to_string(nc) + "-component JPEG");
        if (len < 6 + 3 * nc) invalid("bad DHP");
        for (size_t i = 0; i < nc; i++) dhp_sampling.emplace_back(b[7 + 3 * i] >> 4, b[7 + 3 * i] & 15);
        hier = true;
    }

    void sof(uint8_t m, const uint8_t* b, size_t len) {
        if (hier) return;
        if (sof_seen) invalid("multiple SOF markers");
        if ((m >= 0xC5 && m <= 0xC7) || (m >= 0xCD && m <= 0xCF))
            invalid("differential frame outside a hierarchical image");
        if (len < 6) invalid("bad SOF");
        arith = m >= 0xC9;
        progressive = m == 0xC2 || m == 0xCA;
        lossless = m == 0xC3 || m == 0xCB;
        const int p = b[0];
        if (lossless ? (p < 2 || p > 16) : (p != 8 && p != 12)) unsupported(std::to_string(p) + "-bit JPEG");
        precision = p;
        height = uint32_t(b[1]) << 8 | b[2];
        width = uint32_t(b[3]) << 8 | b[4];
        const size_t nc = b[5];
        if (width == 0 || height == 0) invalid("zero JPEG dimension (DNL not supported)");
        if (uint64_t(width) * height > kMaxPixels) invalid("unreasonable image dimensions");
        if (nc != 1 && nc != 3 && nc != 4) unsupported(std::to_string(nc) + "-component JPEG");
        if (len < 6 + 3 * nc) invalid("bad SOF");
        comps.clear();
        for (size_t i = 0; i < nc; i++) {
            const uint8_t* c = b + 6 + 3 * i;
            Comp k;
            k.id = c[0];
            k.h = c[1] >> 4;
            k.v = c[1] & 15;
            k.tq = c[2];
            if (k.h < 1 || k.h > 4 || k.v < 1 || k.v > 4 || k.tq > 3) invalid("bad component description");
            comps.push_back(k);
        }
        if (nc == 1) comps[0].h = comps[0].v = 1;  // one component is always 1x1 sampled
        hmax = vmax = 1;
        for (const Comp& k : comps) {
            hmax = std::max<uint32_t>(hmax, k.h);
            vmax = std::max<uint32_t>(vmax, k.v);
        }
        const uint32_t unit = lossless ? 1 : 8;
        mcus_x = (width + unit * hmax - 1) / (unit * hmax);
        mcus_y = (height + unit * vmax - 1) / (unit * vmax);
        for (Comp& k : comps) {
            k.blocks_w = mcus_x * k.h;
            k.blocks_h = mcus_y * k.v;
            const uint32_t cw = (width * k.h + hmax - 1) / hmax, ch = (height * k.v + vmax - 1) / vmax;
            k.real_bw = (cw + unit - 1) / unit;
            k.real_bh = (ch + unit - 1) / unit;
        }
        sof_seen = true;
    }

    void sos(const uint8_t* b, size_t len) {
        if (comps.empty()) invalid("SOS before SOF");
        if (len < 1) invalid("bad SOS");
        const size_t ns = b[0];
        if (ns == 0 || ns > comps.size() || len < 1 + 2 * ns + 3) invalid("bad SOS");
        std::vector<ScanComp> sc;
        for (size_t i = 0; i < ns; i++) {
            const uint8_t cid = b[1 + 2 * i], t = b[2 + 2 * i];
            size_t idx = comps.size();
            for (size_t j = 0; j < comps.size(); j++)
                if (comps[j].id == cid) {
                    idx = j;
                    break;
                }
            if (idx == comps.size()) invalid("SOS references unknown component");
            for (const ScanComp& s : sc)
                if (s.idx == idx) invalid("duplicate component in scan");
            sc.push_back(ScanComp{idx, (t >> 4) & 3, t & 3});
        }
        for (const ScanComp& s : sc) {
            Comp& c = comps[s.idx];
            if (!c.table_fixed && qt_set[c.tq]) {
                std::memcpy(c.table, qt[c.tq], sizeof(c.table));
                c.table_fixed = true;
            }
        }
        const uint8_t* t = b + 1 + 2 * ns;
        uint32_t ss = 0, se = 63, ah = 0, al = 0;
        if (progressive && !lossless) {
            ss = t[0];
            se = t[1];
            ah = t[2] >> 4;
            al = t[2] & 15;
            if (ss > se || se > 63 || al > 13 || (ss == 0 && se != 0) || (ss > 0 && ns != 1))
                invalid("bad progressive scan parameters");
        }
        if (!decode || arith || lossless || precision != 8) {
            // Not decoded: the scan's data is passed over by the marker search.
            return;
        }
        pos = scan(sc, ss, se, ah, al);
    }

    size_t scan(const std::vector<ScanComp>& sc, uint32_t ss, uint32_t se, uint32_t ah, uint32_t al) {
        const bool need_dc = !progressive || (ss == 0 && ah == 0);
        const bool need_ac = !progressive || ss > 0;
        for (const ScanComp& s : sc)
            if ((need_dc && !dc[s.dc].defined) || (need_ac && !ac[s.ac].defined))
                invalid("scan uses undefined huffman table");
        Bits br(d, n, pos);
        int32_t pred[4] = {0, 0, 0, 0};
        uint32_t eobrun = 0;
        const bool inter = sc.size() > 1;
        const uint32_t mx_n = inter ? mcus_x : comps[sc[0].idx].real_bw;
        const uint32_t my_n = inter ? mcus_y : comps[sc[0].idx].real_bh;
        const uint64_t total = uint64_t(mx_n) * my_n;
        const uint64_t ri = uint64_t(restart_interval);
        for (uint64_t mcu = 0; mcu < total; mcu++) {
            if (ri > 0 && mcu > 0 && mcu % ri == 0) {
                br.restart();
                pred[0] = pred[1] = pred[2] = pred[3] = 0;
                eobrun = 0;
            }
            const uint32_t mx = uint32_t(mcu % mx_n), my = uint32_t(mcu / mx_n);
            for (size_t si = 0; si < sc.size(); si++) {
                const Comp& c = comps[sc[si].idx];
                const uint32_t bh = inter ? c.h : 1, bv = inter ? c.v : 1;
                for (uint32_t by = 0; by < bv; by++)
                    for (uint32_t bx = 0; bx < bh; bx++) {
                        const size_t col = size_t(mx) * bh + bx, row = size_t(my) * bv + by;
                        int16_t* blk = out[sc[si].idx] + (row * c.blocks_w + col) * 64;
                        const Huff& hd = dc[sc[si].dc];
                        const Huff& ha = ac[sc[si].ac];
                        if (!progressive)
                            sequential(br, hd, ha, blk, pred[si]);
                        else if (ss == 0) {
                            if (ah == 0)
                                dc_first(br, hd, blk, pred[si], al);
                            else
                                blk[0] = int16_t(blk[0] | int16_t(br.get(1) << al));
                        } else if (ah == 0) {
                            ac_first(br, ha, blk, eobrun, ss, se, al);
                        } else {
                            ac_refine(br, ha, blk, eobrun, ss, se, al);
                        }
                    }
            }
        }
        return br.pos;
    }

    static void sequential(Bits& br, const Huff& hd, const Huff& ha, int16_t* blk, int32_t& pred) {
        const uint32_t t = hd.decode(br);
        if (t > 16) invalid("bad DC category");
        pred = int32_t(uint32_t(pred) + uint32_t(extend(br.get(t), t)));
        blk[0] = int16_t(pred);
        for (uint32_t k = 1; k < 64;) {
            const uint8_t rs = ha.decode(br);
            const uint32_t r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r == 15) {
                    k += 16;
                    continue;
                }
                break;
            }
            k += r;
            if (k > 63) invalid("AC coefficient index out of range");
            blk[k] = int16_t(extend(br.get(s), s));
            k++;
        }
    }

    static void dc_first(Bits& br, const Huff& hd, int16_t* blk, int32_t& pred, uint32_t al) {
        const uint32_t t = hd.decode(br);
        if (t > 16) invalid("bad DC category");
        pred = int32_t(uint32_t(pred) + uint32_t(extend(br.get(t), t)));
        blk[0] = int16_t(uint32_t(pred) << al);
    }

    static void ac_first(Bits& br, const Huff& ha, int16_t* blk, uint32_t& eobrun, uint32_t ss, uint32_t se,
                         uint32_t al) {
        if (eobrun > 0) {
            eobrun--;
            return;
        }
        for (uint32_t k = ss; k <= se;) {
            const uint8_t rs = ha.decode(br);
            const uint32_t r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r < 15) {
                    eobrun = (1u << r) - 1;
                    if (r > 0) eobrun += br.get(r);
                    break;
                }
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) invalid("AC coefficient index out of range");
            blk[k] = int16_t(uint32_t(extend(br.get(s), s)) << al);
            k++;
        }
    }

    static void ac_refine(Bits& br, const Huff& ha, int16_t* blk, uint32_t& eobrun, uint32_t ss, uint32_t se,
                          uint32_t al) {
        const int16_t p1 = int16_t(1 << al), m1 = int16_t(-(1 << al));
        uint32_t k = ss;
        auto refine = [&](int16_t& c) {
            if (br.get(1) && (c & p1) == 0) c = int16_t(c + (c >= 0 ? p1 : m1));
        };
        if (eobrun == 0) {
            while (k <= se) {
                const uint8_t rs = ha.decode(br);
                int32_t r = rs >> 4;
                const uint32_t s = rs & 15;
                int16_t value = 0;
                if (s != 0) {
                    value = br.get(1) ? p1 : m1;
                } else if (r != 15) {
                    eobrun = 1u << r;
                    if (r > 0) eobrun += br.get(uint32_t(r));
                    break;
                }
                // Pass over r zero coefficients, refining the nonzero ones met on the way.
                while (k <= se) {
                    int16_t& c = blk[k];
                    if (c != 0) {
                        refine(c);
                    } else {
                        r--;
                        if (r < 0) break;
                    }
                    k++;
                }
                if (value != 0) {
                    if (k > 63) invalid("AC refine index out of range");
                    blk[k] = value;
                }
                k++;
            }
        }
        if (eobrun > 0) {
            for (; k <= se; k++)
                if (blk[k] != 0) refine(blk[k]);
            eobrun--;
        }
    }
};

// EXIF orientation from the first APP1 "Exif" segment before the first scan (decode.rs:572-620); 1 when absent.
int tiff_orientation(const uint8_t* t, size_t n) {
    if (n < 2) return 0;
    bool le;
    if (t[0] == 'I' && t[1] == 'I')
        le = true;
    else if (t[0] == 'M' && t[1] == 'M')
        le = false;
    else
        return 0;
    auto u16 = [&](size_t o, uint32_t& v) {
        if (o + 2 > n) return false;
        v = le ? (uint32_t(t[o]) | uint32_t(t[o + 1]) << 8) : (uint32_t(t[o]) << 8 | t[o + 1]);
        return true;
    };
    auto u32 = [&](size_t o, uint32_t& v) {
        if (o + 4 > n) return false;
        v = le ? (uint32_t(t[o]) | uint32_t(t[o + 1]) << 8 | uint32_t(t[o + 2]) << 16 | uint32_t(t[o + 3]) << 24)
               : (uint32_t(t[o]) << 24 | uint32_t(t[o + 1]) << 16 | uint32_t(t[o + 2]) << 8 | t[o + 3]);
        return true;
    };
    uint32_t ifd, cnt;
    if (!u32(4, ifd) || !u16(ifd, cnt)) return 0;
    for (uint32_t i = 0; i < cnt; i++) {
        const size_t e = size_t(ifd) + 2 + size_t(i) * 12;
        uint32_t tag, v;
        if (!u16(e, tag)) return 0;
        if (tag == 0x0112) {
            if (!u16(e + 8, v)) return 0;
            return v >= 1 && v <= 8 ? int(v) : 0;
        }
    }
    return 0;
}

int exif_orientation(const uint8_t* d, size_t n) {
    size_t pos = 2;
    while (pos + 4 <= n && d[pos] == 0xFF) {
        const uint8_t m = d[pos + 1];
        if (m == 0xD8 || m == 0x01 || (m >= 0xD0 && m <= 0xD7)) {
            pos += 2;
            continue;
        }
        if (m == 0xDA || m == 0xD9) break;
        const size_t len = size_t(d[pos + 2]) << 8 | d[pos + 3];
        if (len < 2 || pos + 2 + len > n) break;
        const uint8_t* seg = d + pos + 4;
        const size_t slen = len - 2;
        if (m == 0xE1 && slen > 14 && std::memcmp(seg, "Exif\0\0", 6) == 0) {
            const int o = tiff_orientation(seg + 6, slen - 6);
            return o ? o : 1;
        }
        pos += 2 + len;
    }
    return 1;
}

void describe(const Reader& r, rcg_jpeg_info* info) {
    std::memset(info, 0, sizeof(*info));
    info->adobe_transform = r.adobe;
    info->arithmetic = r.arith ? 1 : 0;
    if (r.hier) {
        info->width = r.dhp_w;
        info->height = r.dhp_h;
        info->ncomp = uint32_t(r.dhp_sampling.size());
        for (size_t c = 0; c < r.dhp_sampling.size(); c++) {
            info->h[c] = r.dhp_sampling[c].first;
            info->v[c] = r.dhp_sampling[c].second;
        }
        info->source_quality = 100;
        info->precision = r.dhp_precision;
        info->hierarchical = 1;
        return;
    }
    info->width = r.width;
    info->height = r.height;
    info->ncomp = uint32_t(r.comps.size());
    info->precision = r.precision;
    info->progressive = r.progressive ? 1 : 0;
    info->lossless = r.lossless ? 1 : 0;
    bool is_rgb_ids = r.comps.size() == 3 && r.comps[0].id == 'R' && r.comps[1].id == 'G' && r.comps[2].id == 'B';
    info->coded_rgb = r.comps.size() == 3 && (r.adobe == 0 || (r.adobe < 0 && is_rgb_ids)) ? 1 : 0;
    for (size_t c = 0; c < r.comps.size(); c++) {
        const Comp& k = r.comps[c];
        info->id[c] = k.id;
        info->h[c] = k.h;
        info->v[c] = k.v;
        info->tq[c] = k.tq;
        info->blocks_w[c] = k.blocks_w;
        info->blocks_h[c] = k.blocks_h;
    }
    // Each slot's table as its components use it (into_coefs, decode.rs:544-554); a later component wins a slot.
    for (const Comp& k : r.comps) {
        const uint16_t* t = k.table_fixed ? k.table : (r.qt_set[k.tq] ? r.qt[k.tq] : nullptr);
        if (!t) continue;
        bool wide = false;
        for (int i = 0; i < 64; i++) {
            wide = wide || t[i] > 255 || t[i] == 0;
            info->qtables[k.tq][i] = uint8_t(t[i] > 255 ? 255 : t[i]);
        }
        info->present[k.tq] = 1;
        info->table16 = info->table16 || wide;
    }
    // The quality estimate reads the luma slot as the frame ends (decode.rs:461-478).
    if (r.precision == 8 && !r.lossless && !r.comps.empty() && r.qt_set[r.comps[0].tq]) {
        uint8_t t8[64];
        for (int i = 0; i < 64; i++) t8[i] = uint8_t(r.qt[r.comps[0].tq][i] > 255 ? 255 : r.qt[r.comps[0].tq][i]);
        info->source_quality = estimate_quality(t8);
    } else {
        info->source_quality = 100;
    }
}

int parse(const uint8_t* data, size_t len, bool decode, int16_t* const* out, rcg_jpeg_info* info) {
    if (!data || len < 4 || data[0] != 0xFF || data[1] != 0xD8) {
        set_error("not a JPEG file");
        return RCG_INVALID;
    }
    try {
        Reader r(data, len, false);
        r.run();
        if (!r.hier && (r.scans == 0 || r.comps.empty())) invalid("JPEG contains no image data");
        if (info) {
            describe(r, info);
            info->orientation = exif_orientation(data, len);
        }
        if (!decode) return RCG_OK;
        if (r.hier) unsupported("hierarchical JPEG in the DCT domain");
        if (r.lossless) unsupported("lossless JPEG in the DCT domain");
        if (r.precision != 8) unsupported("12-bit JPEG in the DCT domain");
        if (r.arith) unsupported("arithmetic-coded JPEG input");
        for (size_t c = 0; c < r.comps.size(); c++) {
            if (!out[c]) invalid("missing coefficient buffer");
            std::memset(out[c], 0, size_t(r.comps[c].blocks_w) * r.comps[c].blocks_h * 64 * sizeof(int16_t));
        }
        Reader dec(data, len, true);
        dec.out = out;
        dec.run();
    } catch (const Fail& f) {
        set_error(f.msg);
        return f.status;
    } catch (const std::bad_alloc&) {
        set_error("out of host memory");
        return RCG_OOM;
    }
    return RCG_OK;
}

}  // namespace

extern "C" int rcg_jpeg_parse(const uint8_t* data, size_t len, rcg_jpeg_info* info) {
    return rcg::guarded([&]() -> int {
        if (!info) return RCG_INVALID;
        return parse(data, len, false, nullptr, info);
    });
}

extern "C" int rcg_jpeg_read_coefs(const uint8_t* data, size_t len, int16_t* coefs[4]) {
    return rcg::guarded([&]() -> int {
        if (!coefs) return RCG_INVALID;
        return parse(data, len, true, coefs, nullptr);
    });
}
