// dpe_jpeg.cpp -- baseline JPEG decoding for the driver (the reference reads its
// images with cv::imread: grayscale for PatchMatch, DPE.cpp:713, and colour for
// the point colours, DPE.cpp:1422).  OpenCV decodes through libjpeg(-turbo),
// whose grayscale output of a YCbCr file is its Y component, reconstructed with
// the "islow" integer IDCT (Loeffler-Ligtenberg-Moschytz, 13-bit constants,
// 2 extra bits in the first pass) and the range-limit table of IJG libjpeg 6b;
// colour adds its triangle-filter ("fancy") chroma upsampling and the 16-bit
// fixed-point YCbCr -> RGB tables.  Those published algorithms are restated
// here; sequential Huffman (baseline / extended, 8-bit) only: progressive and
// arithmetic-coded files are rejected.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dpe_host.h"

namespace {

const int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                         41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                         30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huff {
    bool present = false;
    int maxcode[18], valptr[17], mincode[17];
    uint8_t vals[256];
};

struct Comp {
    int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
    int bw = 0, bh = 0;  // blocks per line / column (padded to whole MCUs)
    std::vector<uint8_t> plane;
    int pred = 0;
};

struct Decoder {
    const uint8_t *p = nullptr, *end = nullptr;
    uint16_t q[4][64] = {};
    Huff dc[4], ac[4];
    std::vector<Comp> comps;
    int W = 0, H = 0, hmax = 1, vmax = 1, restart = 0;
    bool frame = false;
    // bit reader
    uint32_t bits = 0;
    int nbits = 0;
    bool hit_marker = false;

    int byte() { return p < end ? *p++ : -1; }
    int u16() {
        const int a = byte(), b = byte();
        return (a < 0 || b < 0) ? -1 : (a << 8) | b;
    }
    void fill() {
        while (nbits <= 24) {
            int c = 0;
            if (!hit_marker && p < end) {
                c = *p;
                if (c == 0xFF) {
                    const int n = p + 1 < end ? p[1] : 0;
                    if (n == 0x00) {
                        p += 2;
                    } else {
                        hit_marker = true;  // a marker: feed zeros from here
                        c = 0;
                    }
                } else {
                    ++p;
                }
            }
            bits |= (uint32_t)c << (24 - nbits);
            nbits += 8;
        }
    }
    int getbits(int n) {
        if (n == 0) return 0;
        fill();
        const int v = (int)(bits >> (32 - n));
        bits <<= n;
        nbits -= n;
        return v;
    }
    int decode(const Huff &t) {
        fill();
        int code = 0;
        for (int l = 1; l <= 16; ++l) {
            code = (code << 1) | (int)(bits >> 31);
            bits <<= 1;
            nbits--;
            if (code <= t.maxcode[l]) return t.vals[t.valptr[l] + code - t.mincode[l]];
        }
        return -1;
    }
    static int extend(int v, int n) { return v < (1 << (n - 1)) ? v - (1 << n) + 1 : v; }

    bool dht(int len) {
        const uint8_t *e = p + len;
        while (p < e) {
            const int tc = byte();
            if (tc < 0) return false;
            Huff &t = ((tc >> 4) ? ac : dc)[tc & 3];
            int counts[17] = {0}, total = 0;
            for (int i = 1; i <= 16; ++i) total += (counts[i] = byte());
            if (total > 256 || p + total > end) return false;
            for (int i = 0; i < total; ++i) t.vals[i] = (uint8_t)byte();
            int code = 0, k = 0;
            for (int l = 1; l <= 16; ++l) {
                t.valptr[l] = k;
                t.mincode[l] = code;
                code += counts[l];
                k += counts[l];
                t.maxcode[l] = counts[l] ? code - 1 : -1;
                code <<= 1;
            }
            t.maxcode[17] = 0x7FFFFFFF;
            t.present = true;
        }
        return p == e;
    }
    bool dqt(int len) {
        const uint8_t *e = p + len;
        while (p < e) {
            const int pq = byte();
            if (pq < 0) return false;
            for (int i = 0; i < 64; ++i) q[pq & 3][kZigzag[i]] = (uint16_t)((pq >> 4) ? u16() : byte());
        }
        return p == e;
    }
    bool sof(int len) {
        const uint8_t *e = p + len;
        if (byte() != 8) return false;  // 8-bit samples only
        H = u16();
        W = u16();
        const int n = byte();
        if (H <= 0 || W <= 0 || (n != 1 && n != 3)) return false;
        comps.assign(n, Comp());
        for (Comp &c : comps) {
            c.id = byte();
            const int hv = byte();
            c.h = hv >> 4;
            c.v = hv & 15;
            c.tq = byte() & 3;
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4) return false;
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        const int mx = (W + 8 * hmax - 1) / (8 * hmax), my = (H + 8 * vmax - 1) / (8 * vmax);
        for (Comp &c : comps) {
            c.bw = mx * c.h;
            c.bh = my * c.v;
            c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
        }
        frame = true;
        return p == e;
    }

    // jpeg_idct_islow of IJG libjpeg 6b (as libjpeg-turbo keeps it)
    static void idct(const int16_t *coef, const uint16_t *qt, uint8_t *out, int stride) {
        enum { CB = 13, P1 = 2 };
        const long F0_298 = 2446, F0_390 = 3196, F0_541 = 4433, F0_765 = 6270, F0_899 = 7373, F1_175 = 9633,
                   F1_501 = 12299, F1_847 = 15137, F1_961 = 16069, F2_053 = 16819, F2_562 = 20995, F3_072 = 25172;
        auto descale = [](long x, int n) { return (x + (1L << (n - 1))) >> n; };
        int ws[64];
        for (int c = 0; c < 8; ++c) {
            const int16_t *in = coef + c;
            const uint16_t *qq = qt + c;
            int *w = ws + c;
            if (!in[8] && !in[16] && !in[24] && !in[32] && !in[40] && !in[48] && !in[56]) {
                const int dc = (int)(in[0] * qq[0]) << P1;
                for (int r = 0; r < 8; ++r) w[8 * r] = dc;
                continue;
            }
            long z2 = in[16] * qq[16], z3 = in[48] * qq[48];
            long z1 = (z2 + z3) * F0_541;
            long tmp2 = z1 + z3 * -F1_847, tmp3 = z1 + z2 * F0_765;
            z2 = in[0] * qq[0];
            z3 = in[32] * qq[32];
            long tmp0 = (z2 + z3) << CB, tmp1 = (z2 - z3) << CB;
            const long t10 = tmp0 + tmp3, t13 = tmp0 - tmp3, t11 = tmp1 + tmp2, t12 = tmp1 - tmp2;
            tmp0 = in[56] * qq[56];
            tmp1 = in[40] * qq[40];
            tmp2 = in[24] * qq[24];
            tmp3 = in[8] * qq[8];
            z1 = tmp0 + tmp3;
            z2 = tmp1 + tmp2;
            z3 = tmp0 + tmp2;
            long z4 = tmp1 + tmp3;
            const long z5 = (z3 + z4) * F1_175;
            tmp0 *= F0_298;
            tmp1 *= F2_053;
            tmp2 *= F3_072;
            tmp3 *= F1_501;
            z1 *= -F0_899;
            z2 *= -F2_562;
            z3 *= -F1_961;
            z4 *= -F0_390;
            z3 += z5;
            z4 += z5;
            tmp0 += z1 + z3;
            tmp1 += z2 + z4;
            tmp2 += z2 + z3;
            tmp3 += z1 + z4;
            w[0] = (int)descale(t10 + tmp3, CB - P1);
            w[56] = (int)descale(t10 - tmp3, CB - P1);
            w[8] = (int)descale(t11 + tmp2, CB - P1);
            w[48] = (int)descale(t11 - tmp2, CB - P1);
            w[16] = (int)descale(t12 + tmp1, CB - P1);
            w[40] = (int)descale(t12 - tmp1, CB - P1);
            w[24] = (int)descale(t13 + tmp0, CB - P1);
            w[32] = (int)descale(t13 - tmp0, CB - P1);
        }
        // the post-IDCT range limit: wrap into [-512, 511], add 128, clamp
        auto lim = [](long x) {
            const int v = (int)(((x + 512) & 1023) - 512) + 128;
            return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
        };
        for (int r = 0; r < 8; ++r) {
            const int *w = ws + 8 * r;
            uint8_t *o = out + (size_t)r * stride;
            if (!w[1] && !w[2] && !w[3] && !w[4] && !w[5] && !w[6] && !w[7]) {
                const uint8_t v = lim(descale(w[0], P1 + 3));
                for (int k = 0; k < 8; ++k) o[k] = v;
                continue;
            }
            long z2 = w[2], z3 = w[6];
            long z1 = (z2 + z3) * F0_541;
            long tmp2 = z1 + z3 * -F1_847, tmp3 = z1 + z2 * F0_765;
            long tmp0 = ((long)w[0] + w[4]) << CB, tmp1 = ((long)w[0] - w[4]) << CB;
            const long t10 = tmp0 + tmp3, t13 = tmp0 - tmp3, t11 = tmp1 + tmp2, t12 = tmp1 - tmp2;
            tmp0 = w[7];
            tmp1 = w[5];
            tmp2 = w[3];
            tmp3 = w[1];
            z1 = tmp0 + tmp3;
            z2 = tmp1 + tmp2;
            z3 = tmp0 + tmp2;
            long z4 = tmp1 + tmp3;
            const long z5 = (z3 + z4) * F1_175;
            tmp0 *= F0_298;
            tmp1 *= F2_053;
            tmp2 *= F3_072;
            tmp3 *= F1_501;
            z1 *= -F0_899;
            z2 *= -F2_562;
            z3 *= -F1_961;
            z4 *= -F0_390;
            z3 += z5;
            z4 += z5;
            tmp0 += z1 + z3;
            tmp1 += z2 + z4;
            tmp2 += z2 + z3;
            tmp3 += z1 + z4;
            const int sh = CB + P1 + 3;
            o[0] = lim(descale(t10 + tmp3, sh));
            o[7] = lim(descale(t10 - tmp3, sh));
            o[1] = lim(descale(t11 + tmp2, sh));
            o[6] = lim(descale(t11 - tmp2, sh));
            o[2] = lim(descale(t12 + tmp1, sh));
            o[5] = lim(descale(t12 - tmp1, sh));
            o[3] = lim(descale(t13 + tmp0, sh));
            o[4] = lim(descale(t13 - tmp0, sh));
        }
    }

    bool block(Comp &c, int bx, int by) {
        int16_t coef[64] = {0};
        const Huff &hd = dc[c.td], &ha = ac[c.ta];
        if (!hd.present || !ha.present) return false;
        const int s = decode(hd);
        if (s < 0 || s > 11) return false;
        c.pred += s ? extend(getbits(s), s) : 0;
        coef[0] = (int16_t)c.pred;
        for (int k = 1; k < 64;) {
            const int rs = decode(ha);
            if (rs < 0) return false;
            const int r = rs >> 4, sz = rs & 15;
            if (sz == 0) {
                if (r != 15) break;  // end of block
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) return false;
            coef[kZigzag[k++]] = (int16_t)extend(getbits(sz), sz);
        }
        const int stride = c.bw * 8;
        idct(coef, q[c.tq], c.plane.data() + (size_t)by * 8 * stride + bx * 8, stride);
        return true;
    }

    bool scan(int len) {
        const int n = byte();
        if (n < 1 || n > (int)comps.size()) return false;
        std::vector<Comp *> sc;
        for (int i = 0; i < n; ++i) {
            const int id = byte(), t = byte();
            Comp *c = nullptr;
            for (Comp &k : comps)
                if (k.id == id) c = &k;
            if (!c) return false;
            c->td = t >> 4 & 3;
            c->ta = t & 3;
            sc.push_back(c);
        }
        const int ss = byte(), se = byte(), a = byte();
        if (ss != 0 || se != 63 || a != 0 || len != 4 + 2 * n) return false;  // sequential only
        bits = 0;
        nbits = 0;
        hit_marker = false;
        for (Comp &c : comps) c.pred = 0;
        const int mx = (W + 8 * hmax - 1) / (8 * hmax), my = (H + 8 * vmax - 1) / (8 * vmax);
        int todo = restart;
        if (n == 1) {
            // non-interleaved: the component's own block grid (its real size)
            Comp &c = *sc[0];
            const int cw = (W * c.h + 8 * hmax - 1) / (8 * hmax), ch = (H * c.v + 8 * vmax - 1) / (8 * vmax);
            for (int by = 0; by < ch; ++by)
                for (int bx = 0; bx < cw; ++bx) {
                    if (restart && todo == 0 && !next_restart()) return false;
                    if (!block(c, bx, by)) return false;
                    if (restart) todo = (todo == 0 ? restart : todo) - 1;
                }
        } else {
            for (int my_ = 0; my_ < my; ++my_)
                for (int mx_ = 0; mx_ < mx; ++mx_) {
                    if (restart && todo == 0 && !next_restart()) return false;
                    for (Comp *c : sc)
                        for (int v = 0; v < c->v; ++v)
                            for (int h = 0; h < c->h; ++h)
                                if (!block(*c, mx_ * c->h + h, my_ * c->v + v)) return false;
                    if (restart) todo = (todo == 0 ? restart : todo) - 1;
                }
        }
        // skip to the next marker
        while (p + 1 < end && !(p[0] == 0xFF && p[1] != 0x00 && !(p[1] >= 0xD0 && p[1] <= 0xD7))) ++p;
        return true;
    }
    // a restart marker: realign, reset the predictors
    bool next_restart() {
        while (p + 1 < end && !(p[0] == 0xFF && p[1] >= 0xD0 && p[1] <= 0xD7)) ++p;
        if (p + 1 >= end) return false;
        p += 2;
        bits = 0;
        nbits = 0;
        hit_marker = false;
        for (Comp &c : comps) c.pred = 0;
        return true;
    }

    bool run() {
        if (byte() != 0xFF || byte() != 0xD8) return false;
        bool scanned = false;
        while (p < end) {
            int m = byte();
            if (m != 0xFF) continue;
            while (m == 0xFF) m = byte();
            if (m == 0xD9 || m < 0) break;
            const int len = u16() - 2;
            if (len < 0 || p + len > end) return false;
            if (m == 0xC4) {
                if (!dht(len)) return false;
            } else if (m == 0xDB) {
                if (!dqt(len)) return false;
            } else if (m == 0xC0 || m == 0xC1) {
                if (!sof(len)) return false;
            } else if ((m >= 0xC2 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
                return false;  // progressive, lossless or arithmetic coding
            } else if (m == 0xDD) {
                restart = u16();
            } else if (m == 0xDA) {
                if (!frame || !scan(len)) return false;
                scanned = true;
            } else {
                p += len;
            }
        }
        return frame && scanned;
    }
};

inline uint8_t clamp8(int v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

// libjpeg's triangle-filter ("fancy") upsampling of one chroma plane to the
// luma grid for h2v1 and h2v2 when the plane is more than 2 samples wide (as
// libjpeg-turbo decides), plain replication otherwise
std::vector<uint8_t> upsample(const Comp &c, int hmax, int vmax, int W, int H) {
    const int cw = (W * c.h + hmax - 1) / hmax, chh = (H * c.v + vmax - 1) / vmax;
    const int stride = c.bw * 8;
    const int fx = hmax / c.h, fy = vmax / c.v;
    std::vector<uint8_t> out((size_t)W * H);
    auto in = [&](int x, int y) { return (int)c.plane[(size_t)std::min(std::max(y, 0), chh - 1) * stride + x]; };
    if (fx == 2 && (fy == 1 || fy == 2) && cw > 2) {
        std::vector<int> row((size_t)cw);
        std::vector<uint8_t> line((size_t)2 * cw);
        for (int oy = 0; oy < H; ++oy) {
            const int iy = fy == 2 ? oy / 2 : oy;
            if (fy == 2) {  // column sums: 3 * nearer row + farther row
                const int other = (oy & 1) ? iy + 1 : iy - 1;
                for (int x = 0; x < cw; ++x) row[x] = in(x, iy) * 3 + in(x, other);
                int last = row[0], cur = row[0], next = row[1];
                line[0] = (uint8_t)((cur * 4 + 8) >> 4);
                line[1] = (uint8_t)((cur * 3 + next + 7) >> 4);
                for (int x = 1; x < cw - 1; ++x) {
                    last = cur;
                    cur = row[x];
                    next = row[x + 1];
                    line[2 * x] = (uint8_t)((cur * 3 + last + 8) >> 4);
                    line[2 * x + 1] = (uint8_t)((cur * 3 + next + 7) >> 4);
                }
                last = row[cw - 2];
                cur = row[cw - 1];
                line[2 * cw - 2] = (uint8_t)((cur * 3 + last + 8) >> 4);
                line[2 * cw - 1] = (uint8_t)((cur * 4 + 7) >> 4);
            } else {
                line[0] = (uint8_t)in(0, iy);
                line[1] = (uint8_t)((in(0, iy) * 3 + in(1, iy) + 2) >> 2);
                for (int x = 1; x < cw - 1; ++x) {
                    const int v = in(x, iy) * 3;
                    line[2 * x] = (uint8_t)((v + in(x - 1, iy) + 1) >> 2);
                    line[2 * x + 1] = (uint8_t)((v + in(x + 1, iy) + 2) >> 2);
                }
                line[2 * cw - 2] = (uint8_t)((in(cw - 1, iy) * 3 + in(cw - 2, iy) + 1) >> 2);
                line[2 * cw - 1] = (uint8_t)in(cw - 1, iy);
            }
            for (int x = 0; x < W; ++x) out[(size_t)oy * W + x] = line[std::min(x, 2 * cw - 1)];
        }
        return out;
    }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) out[(size_t)y * W + x] = c.plane[(size_t)(y / fy) * stride + x / fx];
    return out;
}

}  // namespace

extern "C" {

int dpe_host_read_jpeg(const char *path, int color, uint8_t *out, size_t capacity, int *width, int *height) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return -1;
    std::vector<uint8_t> buf;
    uint8_t tmp[65536];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    Decoder d;
    d.p = buf.data();
    d.end = buf.data() + buf.size();
    if (!d.run()) return -2;
    if (width) *width = d.W;
    if (height) *height = d.H;
    const size_t need = (size_t)d.W * d.H * (color ? 3 : 1);
    if (!out) return 0;
    if (capacity < need) return -3;
    const Comp &y = d.comps[0];
    const int ys = y.bw * 8;
    const int fyx = d.hmax / y.h, fyy = d.vmax / y.v;
    if (!color || d.comps.size() == 1) {
        for (int r = 0; r < d.H; ++r)
            for (int c = 0; c < d.W; ++c) {
                const uint8_t v = y.plane[(size_t)(r / fyy) * ys + c / fyx];
                if (color) {
                    uint8_t *o = out + ((size_t)r * d.W + c) * 3;
                    o[0] = o[1] = o[2] = v;
                } else {
                    out[(size_t)r * d.W + c] = v;
                }
            }
        return 0;
    }
    // jdcolor.c's YCbCr -> RGB tables (16-bit fixed point), written as BGR
    const std::vector<uint8_t> cb = upsample(d.comps[1], d.hmax, d.vmax, d.W, d.H);
    const std::vector<uint8_t> cr = upsample(d.comps[2], d.hmax, d.vmax, d.W, d.H);
    auto fix = [](double x) { return (long)(x * 65536.0 + 0.5); };
    int cr_r[256], cb_b[256];
    long cr_g[256], cb_g[256];
    for (int i = 0; i < 256; ++i) {
        const long x = i - 128;
        cr_r[i] = (int)((fix(1.40200) * x + 32768) >> 16);
        cb_b[i] = (int)((fix(1.77200) * x + 32768) >> 16);
        cr_g[i] = -fix(0.71414) * x;
        cb_g[i] = -fix(0.34414) * x + 32768;
    }
    for (int r = 0; r < d.H; ++r)
        for (int c = 0; c < d.W; ++c) {
            const size_t i = (size_t)r * d.W + c;
            const int yy = y.plane[(size_t)(r / fyy) * ys + c / fyx];
            uint8_t *o = out + i * 3;
            o[2] = clamp8(yy + cr_r[cr[i]]);
            o[1] = clamp8(yy + (int)((cb_g[cb[i]] + cr_g[cr[i]]) >> 16));
            o[0] = clamp8(yy + cb_b[cb[i]]);
        }
    return 0;
}

}  // extern "C"
