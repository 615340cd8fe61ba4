// JPEG decoding for target images (host code). The reference reads JPEG
// through stb_image (src/io/image_io.cpp:26); this is a decoder written from
// ITU-T T.81 (ISO/IEC 10918-1):
//   - baseline and extended sequential Huffman (SOF0, SOF1) and progressive
//     Huffman (SOF2) frames, 8-bit samples, 1 or 3 components, sampling
//     factors 1 or 2, restart intervals;
//   - progressive scans per Annex G: DC first / refinement, AC first (with
//     end-of-band runs) / refinement (successive approximation);
//   - dequantisation and an 8x8 inverse DCT in double precision (the
//     definition in A.3.3), level shift, rounding and clamping;
//   - chroma upsampling of 2x-subsampled axes by the triangle filter about
//     the JFIF sample centres (weights 3/4 and 1/4 per axis, edges
//     replicated), and the JFIF YCbCr -> RGB transform (an Adobe APP14
//     marker with transform 0 marks RGB components).
// Arithmetic coding, 12-bit samples and lossless or hierarchical frames are
// rejected with an error.
#include "jpeg.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

namespace dsh {
namespace {

// Natural (row-major) index of the k-th coefficient in zig-zag order (T.81
// figure A.6).
constexpr uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
    // canonical code tables (T.81 C.2, F.2.2.3): per length the largest code
    // (-1: none), and the index of its first value
    int32_t maxcode[18];
    int32_t valptr[17];
    int32_t mincode[17];
    uint8_t vals[256];
    bool present = false;
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int bw = 0, bh = 0;          // blocks covering the component's own samples
    int bw_pad = 0, bh_pad = 0;  // blocks of the MCU grid
    std::vector<int16_t> coef;   // bw_pad * bh_pad blocks, natural order
    int dc_pred = 0;
    int td = 0, ta = 0;          // table selectors of the current scan
};

class Decoder {
public:
    Decoder(const uint8_t *d, size_t n) : data(d), size(n) {}
    bool run(JpegImage &out, std::string &err);

private:
    const uint8_t *data;
    size_t size;
    size_t pos = 0;
    std::string error;
    uint16_t qt[4][64];  // natural order
    bool qt_present[4] = {false, false, false, false};
    Huffman dc_tab[4], ac_tab[4];
    std::vector<Component> comp;
    int W = 0, H = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool progressive = false, frame_seen = false, adobe_rgb = false;
    int restart_interval = 0;
    // scan state
    uint32_t bitbuf = 0;
    int bitcnt = 0;
    bool marker_hit = false;
    int eobrun = 0;

    bool fail(const std::string &m) {
        if (error.empty()) error = m;
        return false;
    }
    bool need(size_t n) { return pos + n <= size ? true : fail("JPEG data ends early"); }
    int u16() {
        const int v = (data[pos] << 8) | data[pos + 1];
        pos += 2;
        return v;
    }
    bool read_dqt(int len);
    bool read_dht(int len);
    bool read_sof(int len, int marker);
    bool read_sos(int len);
    bool decode_scan(const std::vector<int> &sc, int ss, int se, int ah, int al);
    bool decode_block(Component &c, int16_t *blk, int ss, int se, int ah, int al);
    bool restart();
    // entropy-coded segment
    void fill();
    int bit();
    int bits(int n);
    int decode(const Huffman &t);
    static int extend(int v, int t) { return v < (1 << (t - 1)) ? v - (1 << t) + 1 : v; }
    void output(JpegImage &out);
};

bool Decoder::read_dqt(int len) {
    const size_t end = pos + len;
    while (pos < end) {
        if (!need(1)) return false;
        const int pq = data[pos] >> 4, tq = data[pos] & 15;
        ++pos;
        if (tq > 3 || pq > 1) return fail("bad quantisation table");
        if (!need(pq ? 128 : 64)) return false;
        for (int k = 0; k < 64; ++k) {
            const int q = pq ? u16() : data[pos++];
            qt[tq][kZigzag[k]] = (uint16_t)q;
        }
        qt_present[tq] = true;
    }
    return pos == end ? true : fail("bad DQT length");
}

bool Decoder::read_dht(int len) {
    const size_t end = pos + len;
    while (pos < end) {
        if (!need(17)) return false;
        const int tc = data[pos] >> 4, th = data[pos] & 15;
        ++pos;
        if (tc > 1 || th > 3) return fail("bad Huffman table");
        uint8_t counts[17];
        int total = 0;
        for (int l = 1; l <= 16; ++l) {
            counts[l] = data[pos++];
            total += counts[l];
        }
        if (total > 256 || !need((size_t)total)) return fail("bad Huffman table");
        Huffman &t = tc ? ac_tab[th] : dc_tab[th];
        std::memcpy(t.vals, data + pos, (size_t)total);
        pos += (size_t)total;
        int code = 0, k = 0;
        for (int l = 1; l <= 16; ++l) {
            t.valptr[l] = k;
            t.mincode[l] = code;
            code += counts[l];
            k += counts[l];
            t.maxcode[l] = counts[l] ? code - 1 : -1;
            if (code > (1 << l)) return fail("bad Huffman code lengths");
            code <<= 1;
        }
        t.maxcode[17] = 0x7fffffff;
        t.present = true;
    }
    return pos == end ? true : fail("bad DHT length");
}

bool Decoder::read_sof(int len, int marker) {
    if (frame_seen) return fail("more than one frame");
    if (!need((size_t)len) || len < 6) return fail("bad SOF");
    frame_seen = true;
    progressive = marker == 0xC2;
    const int prec = data[pos++];
    H = u16();
    W = u16();
    const int nc = data[pos++];
    if (prec != 8) return fail("only 8-bit JPEG samples are supported");
    if (W <= 0 || H <= 0 || W > 65535 || H > 65535) return fail("bad JPEG size");
    if (nc != 1 && nc != 3) return fail("JPEG with " + std::to_string(nc) + " components");
    if (len != 6 + 3 * nc) return fail("bad SOF length");
    comp.resize((size_t)nc);
    for (Component &c : comp) {
        c.id = data[pos];
        c.h = data[pos + 1] >> 4;
        c.v = data[pos + 1] & 15;
        c.tq = data[pos + 2];
        pos += 3;
        if (c.h < 1 || c.h > 2 || c.v < 1 || c.v > 2 || c.tq > 3)
            return fail("unsupported sampling factors or table");
        hmax = std::max(hmax, c.h);
        vmax = std::max(vmax, c.v);
    }
    mcux = (W + 8 * hmax - 1) / (8 * hmax);
    mcuy = (H + 8 * vmax - 1) / (8 * vmax);
    for (Component &c : comp) {
        const int cw = (W * c.h + hmax - 1) / hmax, ch = (H * c.v + vmax - 1) / vmax;
        c.bw = (cw + 7) / 8;
        c.bh = (ch + 7) / 8;
        c.bw_pad = mcux * c.h;
        c.bh_pad = mcuy * c.v;
        c.coef.assign((size_t)c.bw_pad * c.bh_pad * 64, 0);
    }
    return true;
}

void Decoder::fill() {
    while (bitcnt <= 24) {
        int b = 0;
        if (!marker_hit && pos < size) {
            b = data[pos];
            if (b == 0xFF) {
                const int nx = pos + 1 < size ? data[pos + 1] : 0;
                if (nx == 0x00) {
                    pos += 2;
                } else {
                    marker_hit = true;  // leave the marker for the caller
                    b = 0;
                }
            } else {
                ++pos;
            }
        }
        bitbuf |= (uint32_t)b << (24 - bitcnt);
        bitcnt += 8;
    }
}

int Decoder::bit() {
    if (bitcnt < 1) fill();
    const int b = (int)(bitbuf >> 31);
    bitbuf <<= 1;
    --bitcnt;
    return b;
}

int Decoder::bits(int n) {
    if (n == 0) return 0;
    if (bitcnt < n) fill();
    const int v = (int)(bitbuf >> (32 - n));
    bitbuf <<= n;
    bitcnt -= n;
    return v;
}

int Decoder::decode(const Huffman &t) {
    int code = bit();
    int l = 1;
    while (code > t.maxcode[l]) {
        code = (code << 1) | bit();
        if (++l > 16) {
            fail("bad Huffman code");
            return 0;
        }
    }
    return t.vals[t.valptr[l] + code - t.mincode[l]];
}

// At a restart marker: drop the bit buffer, expect RSTn, reset predictions.
bool Decoder::restart() {
    bitbuf = 0;
    bitcnt = 0;
    marker_hit = false;
    if (pos + 1 < size && data[pos] == 0xFF && data[pos + 1] >= 0xD0 && data[pos + 1] <= 0xD7)
        pos += 2;
    else
        return fail("missing restart marker");
    for (Component &c : comp) c.dc_pred = 0;
    eobrun = 0;
    return true;
}

bool Decoder::decode_block(Component &c, int16_t *blk, int ss, int se, int ah, int al) {
    if (!progressive) {  // sequential: DC then AC 1..63
        const int t = decode(dc_tab[c.td]);
        if (t > 11) return fail("bad DC magnitude");
        const int diff = t ? extend(bits(t), t) : 0;
        c.dc_pred += diff;
        blk[0] = (int16_t)c.dc_pred;
        for (int k = 1; k < 64;) {
            const int rs = decode(ac_tab[c.ta]);
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r != 15) break;  // end of block
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) return fail("AC run past the block");
            blk[kZigzag[k]] = (int16_t)extend(bits(s), s);
            ++k;
        }
        return true;
    }
    if (ss == 0) {  // DC scans
        if (ah == 0) {
            const int t = decode(dc_tab[c.td]);
            if (t > 11) return fail("bad DC magnitude");
            const int diff = t ? extend(bits(t), t) : 0;
            c.dc_pred += diff;
            blk[0] = (int16_t)(c.dc_pred * (1 << al));
        } else if (bit()) {
            blk[0] = (int16_t)(blk[0] | (1 << al));
        }
        return true;
    }
    if (ah == 0) {  // AC first scan
        if (eobrun > 0) {
            --eobrun;
            return true;
        }
        for (int k = ss; k <= se;) {
            const int rs = decode(ac_tab[c.ta]);
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r < 15) {
                    eobrun = (1 << r) - 1;
                    if (r) eobrun += bits(r);
                    break;
                }
                k += 16;
                continue;
            }
            k += r;
            if (k > se) return fail("AC run past the band");
            blk[kZigzag[k]] = (int16_t)(extend(bits(s), s) * (1 << al));
            ++k;
        }
        return true;
    }
    // AC refinement (G.1.2.3): each nonzero coefficient of the band gets a
    // correction bit; new coefficients are +-1 << al after runs of zero ones
    const int p1 = 1 << al, m1 = -(1 << al);
    auto refine = [&](int16_t &x) {
        if (bit() && (x & p1) == 0) x = (int16_t)(x + (x >= 0 ? p1 : m1));
    };
    int k = ss;
    if (eobrun == 0) {
        for (; k <= se; ++k) {
            const int rs = decode(ac_tab[c.ta]);
            int r = rs >> 4;
            const int s = rs & 15;
            int val = 0;
            if (s) {
                if (s != 1) return fail("bad refinement magnitude");
                val = bit() ? p1 : m1;
            } else if (r != 15) {
                eobrun = 1 << r;
                if (r) eobrun += bits(r);
                break;
            }
            // skip r zero coefficients (refining the nonzero ones passed)
            for (; k <= se; ++k) {
                int16_t &x = blk[kZigzag[k]];
                if (x != 0) {
                    refine(x);
                } else {
                    if (r == 0) break;
                    --r;
                }
            }
            if (val && k <= se) blk[kZigzag[k]] = (int16_t)val;
        }
    }
    if (eobrun > 0) {
        for (; k <= se; ++k) {
            int16_t &x = blk[kZigzag[k]];
            if (x != 0) refine(x);
        }
        --eobrun;
    }
    return true;
}

bool Decoder::decode_scan(const std::vector<int> &sc, int ss, int se, int ah, int al) {
    bitbuf = 0;
    bitcnt = 0;
    marker_hit = false;
    eobrun = 0;
    for (int i : sc) comp[(size_t)i].dc_pred = 0;
    int since_restart = 0;
    auto next_unit = [&]() {
        if (restart_interval && since_restart == restart_interval) {
            if (!restart()) return false;
            since_restart = 0;
        }
        ++since_restart;
        return true;
    };
    if (sc.size() == 1) {  // non-interleaved: the component's own blocks
        Component &c = comp[(size_t)sc[0]];
        for (int by = 0; by < c.bh; ++by)
            for (int bx = 0; bx < c.bw; ++bx) {
                if (!next_unit()) return false;
                int16_t *blk = &c.coef[((size_t)by * c.bw_pad + bx) * 64];
                if (!decode_block(c, blk, ss, se, ah, al)) return false;
            }
    } else {
        for (int my = 0; my < mcuy; ++my)
            for (int mx = 0; mx < mcux; ++mx) {
                if (!next_unit()) return false;
                for (int i : sc) {
                    Component &c = comp[(size_t)i];
                    for (int y = 0; y < c.v; ++y)
                        for (int x = 0; x < c.h; ++x) {
                            const int by = my * c.v + y, bx = mx * c.h + x;
                            int16_t *blk = &c.coef[((size_t)by * c.bw_pad + bx) * 64];
                            if (!decode_block(c, blk, ss, se, ah, al)) return false;
                        }
                }
            }
    }
    if (!error.empty()) return false;
    // skip to the next marker
    bitbuf = 0;
    bitcnt = 0;
    marker_hit = false;
    while (pos + 1 < size && !(data[pos] == 0xFF && data[pos + 1] != 0x00 &&
                               !(data[pos + 1] >= 0xD0 && data[pos + 1] <= 0xD7)))
        ++pos;
    return true;
}

bool Decoder::read_sos(int len) {
    if (!frame_seen) return fail("scan before frame");
    if (!need((size_t)len) || len < 6) return fail("bad SOS");
    const int ns = data[pos++];
    if (ns < 1 || ns > 4 || len != 4 + 2 * ns) return fail("bad SOS");
    std::vector<int> sc;
    for (int k = 0; k < ns; ++k) {
        const int id = data[pos], t = data[pos + 1];
        pos += 2;
        int idx = -1;
        for (size_t i = 0; i < comp.size(); ++i)
            if (comp[i].id == id) idx = (int)i;
        if (idx < 0) return fail("scan names an unknown component");
        comp[(size_t)idx].td = t >> 4;
        comp[(size_t)idx].ta = t & 15;
        if (comp[(size_t)idx].td > 3 || comp[(size_t)idx].ta > 3) return fail("bad table");
        sc.push_back(idx);
    }
    const int ss = data[pos], se = data[pos + 1], ah = data[pos + 2] >> 4, al = data[pos + 2] & 15;
    pos += 3;
    if (progressive) {
        if (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && ns != 1) || al > 13)
            return fail("bad progressive scan");
    } else if (ss != 0 || se != 63) {
        return fail("bad sequential scan");
    }
    for (int i : sc) {
        const Component &c = comp[(size_t)i];
        if ((ss == 0 && ah == 0 && !dc_tab[c.td].present) || (se > 0 && !ac_tab[c.ta].present))
            return fail("scan uses an undefined Huffman table");
    }
    return decode_scan(sc, ss, se, ah, al);
}

// 8x8 inverse DCT (T.81 A.3.3) in double, + 128, rounded and clamped.
void idct8x8(const int16_t *in, const uint16_t *q, uint8_t *out, int stride) {
    struct Basis {  // c[x][u] = C(u)/2 cos((2x+1) u pi / 16)
        double c[8][8];
        Basis() {
            for (int x = 0; x < 8; ++x)
                for (int u = 0; u < 8; ++u)
                    c[x][u] = (u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 *
                              std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0);
        }
    };
    static const Basis basis;
    const auto &cs = basis.c;
    double F[64], tmp[64];
    for (int k = 0; k < 64; ++k) F[k] = (double)in[k] * q[k];
    for (int y = 0; y < 8; ++y)  // rows of frequency v -> spatial x
        for (int x = 0; x < 8; ++x) {
            double s = 0;
            for (int u = 0; u < 8; ++u) s += cs[x][u] * F[y * 8 + u];
            tmp[y * 8 + x] = s;
        }
    for (int x = 0; x < 8; ++x)
        for (int y = 0; y < 8; ++y) {
            double s = 0;
            for (int v = 0; v < 8; ++v) s += cs[y][v] * tmp[v * 8 + x];
            const long r = std::lround(s + 128.0);
            out[y * stride + x] = (uint8_t)std::min(255L, std::max(0L, r));
        }
}

// Upsamples a plane by 2 along x and/or y with the triangle filter about the
// sample centres (3/4 of the nearer, 1/4 of the farther sample; replicated at
// the edges), rounded to nearest; integer arithmetic in units of 1/16.
std::vector<uint8_t> upsample(const std::vector<uint8_t> &p, int pw, int ph, int fx, int fy,
                              int ow, int oh) {
    std::vector<uint8_t> out((size_t)ow * oh);
    auto at = [&](int x, int y) {
        x = std::min(std::max(x, 0), pw - 1);
        y = std::min(std::max(y, 0), ph - 1);
        return (int)p[(size_t)y * pw + x];
    };
    for (int y = 0; y < oh; ++y) {
        const int sy = fy == 2 ? y / 2 : y, ny = fy == 2 ? (y % 2 ? sy + 1 : sy - 1) : sy;
        for (int x = 0; x < ow; ++x) {
            const int sx = fx == 2 ? x / 2 : x, nx = fx == 2 ? (x % 2 ? sx + 1 : sx - 1) : sx;
            const int wy0 = fy == 2 ? 3 : 4, wy1 = fy == 2 ? 1 : 0;
            const int wx0 = fx == 2 ? 3 : 4, wx1 = fx == 2 ? 1 : 0;
            const int v = wy0 * (wx0 * at(sx, sy) + wx1 * at(nx, sy)) +
                          wy1 * (wx0 * at(sx, ny) + wx1 * at(nx, ny));
            out[(size_t)y * ow + x] = (uint8_t)((v + 8) >> 4);
        }
    }
    return out;
}

void Decoder::output(JpegImage &out) {
    const int nc = (int)comp.size();
    std::vector<std::vector<uint8_t>> planes((size_t)nc);
    for (int i = 0; i < nc; ++i) {
        const Component &c = comp[(size_t)i];
        const int pw = c.bw_pad * 8, ph = c.bh_pad * 8;
        std::vector<uint8_t> plane((size_t)pw * ph);
        for (int by = 0; by < c.bh_pad; ++by)
            for (int bx = 0; bx < c.bw_pad; ++bx)
                idct8x8(&c.coef[((size_t)by * c.bw_pad + bx) * 64], qt[c.tq],
                        &plane[(size_t)by * 8 * pw + bx * 8], pw);
        // the component's own samples, then up to the frame's sampling
        const int cw = (W * c.h + hmax - 1) / hmax, ch = (H * c.v + vmax - 1) / vmax;
        std::vector<uint8_t> own((size_t)cw * ch);
        for (int y = 0; y < ch; ++y)
            std::memcpy(&own[(size_t)y * cw], &plane[(size_t)y * pw], (size_t)cw);
        const int fx = hmax / c.h, fy = vmax / c.v;
        if (fx == 1 && fy == 1)
            planes[(size_t)i] = std::move(own);
        else
            planes[(size_t)i] = upsample(own, cw, ch, fx, fy, W, H);
    }
    out.w = W;
    out.h = H;
    out.rgb.assign((size_t)W * H * 3, 0);
    for (size_t p = 0; p < (size_t)W * H; ++p) {
        uint8_t *o = &out.rgb[3 * p];
        if (nc == 1) {
            o[0] = o[1] = o[2] = planes[0][p];
        } else if (adobe_rgb) {
            o[0] = planes[0][p];
            o[1] = planes[1][p];
            o[2] = planes[2][p];
        } else {  // JFIF YCbCr -> RGB
            const double y = planes[0][p], cb = planes[1][p] - 128.0, cr = planes[2][p] - 128.0;
            const double rgb[3] = {y + 1.402 * cr, y - 0.344136286 * cb - 0.714136286 * cr,
                                   y + 1.772 * cb};
            for (int k = 0; k < 3; ++k)
                o[k] = (uint8_t)std::min(255L, std::max(0L, std::lround(rgb[k])));
        }
    }
}

bool Decoder::run(JpegImage &out, std::string &err) {
    if (size < 4 || data[0] != 0xFF || data[1] != 0xD8) {
        err = "not a JPEG file";
        return false;
    }
    pos = 2;
    bool done = false;
    while (!done) {
        while (pos < size && data[pos] != 0xFF) ++pos;  // tolerate fill
        while (pos < size && data[pos] == 0xFF) ++pos;
        if (pos >= size) {
            fail("JPEG ends without EOI");
            break;
        }
        const int m = data[pos++];
        if (m == 0xD9) {
            done = true;
            break;
        }
        if (m >= 0xD0 && m <= 0xD7) continue;
        if (!need(2)) break;
        const int len = u16() - 2;
        if (len < 0 || !need((size_t)len)) {
            fail("bad JPEG segment length");
            break;
        }
        bool ok = true;
        const size_t seg = pos;
        switch (m) {
            case 0xDB: ok = read_dqt(len); break;
            case 0xC4: ok = read_dht(len); break;
            case 0xC0:
            case 0xC1:
            case 0xC2: ok = read_sof(len, m); break;
            case 0xDD:
                restart_interval = len >= 2 ? (data[pos] << 8) | data[pos + 1] : 0;
                pos += (size_t)len;
                break;
            case 0xDA: ok = read_sos(len); break;
            case 0xEE:  // Adobe APP14: transform flag
                if (len >= 12 && std::memcmp(data + pos, "Adobe", 5) == 0)
                    adobe_rgb = data[pos + 11] == 0;
                pos += (size_t)len;
                break;
            default:
                if ((m >= 0xC3 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
                    fail("unsupported JPEG process (arithmetic, lossless or hierarchical)");
                    ok = false;
                } else {
                    pos = seg + (size_t)len;  // APPn, COM, ...
                }
        }
        if (!ok) break;
    }
    if (error.empty() && !frame_seen) fail("JPEG without a frame");
    if (error.empty())
        for (const Component &c : comp)
            if (!qt_present[c.tq]) fail("component uses an undefined quantisation table");
    if (!error.empty()) {
        err = error;
        return false;
    }
    output(out);
    return true;
}

}  // namespace

bool decode_jpeg(const uint8_t *data, size_t size, JpegImage &out, std::string &err) {
    Decoder d(data, size);
    return d.run(out, err);
}

}  // namespace dsh
