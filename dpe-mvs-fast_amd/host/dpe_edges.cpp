// dpe_edges.cpp -- the reference's support-map segmentation (EdgeSegment,
// Roberts, Connect; DPE.cpp:9-291) without OpenCV: the OpenCV routines it
// calls (Canny, HoughLinesP, threshold, resize, line) are restated from their
// published algorithms (OpenCV >= 3.3, the version the reference recommends;
// which release it was built with is not pinned).  Exact agreement with an
// OpenCV build cannot be checked here: Roberts, threshold and Connect are exact
// restatements; Canny and HoughLinesP follow OpenCV's algorithms including
// their fixed-point tests and random point order; the 8-bit INTER_LINEAR
// resize is computed in float and rounded (OpenCV uses 11-bit fixed point).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "dpe_host.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// cv::threshold(THRESH_BINARY) on 8-bit data: src > floor(thr) ? 255 : 0
void threshold_u8(std::vector<uint8_t> &img, double thr) {
    const int t = (int)std::floor(thr);
    for (uint8_t &v : img) v = (int)v > t ? 255 : 0;
}

// cv::resize(INTER_LINEAR) of 8-bit data: half-pixel centres, border taps
// clamped, rounded to nearest
std::vector<uint8_t> resize_u8(const std::vector<uint8_t> &src, int w, int h, int dw, int dh) {
    std::vector<uint8_t> dst((size_t)dw * dh);
    if (dw == w && dh == h) {
        dst = src;
        return dst;
    }
    const double sx = (double)w / dw, sy = (double)h / dh;
    for (int y = 0; y < dh; ++y) {
        double fy = (y + 0.5) * sy - 0.5;
        int y0 = (int)std::floor(fy);
        double b = fy - y0;
        if (y0 < 0) y0 = 0, b = 0;
        if (y0 >= h - 1) y0 = h - 1, b = 0;
        const int y1 = std::min(y0 + 1, h - 1);
        for (int x = 0; x < dw; ++x) {
            double fx = (x + 0.5) * sx - 0.5;
            int x0 = (int)std::floor(fx);
            double a = fx - x0;
            if (x0 < 0) x0 = 0, a = 0;
            if (x0 >= w - 1) x0 = w - 1, a = 0;
            const int x1 = std::min(x0 + 1, w - 1);
            const double top = src[(size_t)y0 * w + x0] * (1 - a) + src[(size_t)y0 * w + x1] * a;
            const double bot = src[(size_t)y1 * w + x0] * (1 - a) + src[(size_t)y1 * w + x1] * a;
            dst[(size_t)y * dw + x] = (uint8_t)clampi((int)std::lrint(top * (1 - b) + bot * b), 0, 255);
        }
    }
    return dst;
}

// Roberts (DPE.cpp:9-25); the reference's (uchar) of sqrt keeps the low 8 bits
// of the truncated value (gradients above 255 wrap), as x86 code does
void roberts(const uint8_t *src, int w, int h, uint8_t *dst) {
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            int t1 = 50, t2 = 50;
            if (i > 0 && i < h - 1 && j > 0 && j < w - 1) {
                t1 = src[(size_t)i * w + j] - src[(size_t)(i + 1) * w + j + 1];
                t2 = src[(size_t)(i + 1) * w + j] - src[(size_t)i * w + j + 1];
            }
            dst[(size_t)i * w + j] = (uint8_t)(unsigned)(int)std::sqrt((double)(t1 * t1 + t2 * t2));
        }
}

// Connect (DPE.cpp:27-139): 4-connected regions of 0-pixels with the
// reference's union-find and renumbering; 255-pixels get label 0.  Returns
// the per-label counts (index = label).
std::vector<int> connect(const uint8_t *img, int w, int h, int32_t *label) {
    std::vector<int> conn{0};
    int cnt = 1;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            if (img[i] == 255) {
                label[i] = 0;
                continue;
            }
            const bool left_n = x > 0 && img[i] == 0 && img[i - 1] == 0;
            const bool up_n = y > 0 && img[i] == 0 && img[i - w] == 0;
            if (left_n) label[i] = label[i - 1];
            if (up_n) label[i] = label[i - w];
            if (!left_n && !up_n) {
                label[i] = cnt;
                conn.push_back(cnt);
                cnt++;
            } else if (left_n && up_n) {
                const int l = label[i - 1], u = label[i - w];
                if (l > u) {
                    conn[l] = u;
                    label[i] = u;
                } else if (l < u) {
                    conn[u] = l;
                    label[i] = l;
                }
            }
        }
    for (size_t i = 1; i < conn.size(); ++i) {
        int cur = conn[i], pre = conn[cur];
        while (pre != cur) {
            cur = pre;
            pre = conn[pre];
        }
        conn[i] = cur;
    }
    int num = 1;
    std::vector<int> map(conn.size(), 0);
    for (size_t i = 1; i < conn.size(); ++i)
        if (conn[i] == (int)i) map[i] = num++;
    for (size_t i = 1; i < conn.size(); ++i) conn[i] = map[conn[i]];
    std::vector<int> counts(num, 0);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        label[i] = conn[label[i]];
        counts[label[i]]++;
    }
    return counts;
}

// cv::Canny(aperture 3, L2 gradient): 3x3 Sobel with replicated borders,
// squared magnitudes against squared thresholds, non-maximum suppression with
// OpenCV's fixed-point sector tests (tan 22.5 deg in 1.15 fixed point), then
// hysteresis over the 8-neighbourhood from pixels above the high threshold.
void canny(const uint8_t *src, int w, int h, double lo_t, double hi_t, uint8_t *out) {
    if (lo_t > hi_t) std::swap(lo_t, hi_t);
    lo_t = std::min(32767.0, lo_t);
    hi_t = std::min(32767.0, hi_t);
    if (lo_t > 0) lo_t *= lo_t;
    if (hi_t > 0) hi_t *= hi_t;
    const int low = (int)std::floor(lo_t), high = (int)std::floor(hi_t);
    auto px = [&](int x, int y) { return (int)src[(size_t)clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)]; };
    const size_t n = (size_t)w * h;
    std::vector<int> dx(n), dy(n), mag((size_t)(w + 2) * (h + 2), 0);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int gx = (px(x + 1, y - 1) - px(x - 1, y - 1)) + 2 * (px(x + 1, y) - px(x - 1, y)) +
                           (px(x + 1, y + 1) - px(x - 1, y + 1));
            const int gy = (px(x - 1, y + 1) - px(x - 1, y - 1)) + 2 * (px(x, y + 1) - px(x, y - 1)) +
                           (px(x + 1, y + 1) - px(x + 1, y - 1));
            dx[(size_t)y * w + x] = gx;
            dy[(size_t)y * w + x] = gy;
            mag[(size_t)(y + 1) * (w + 2) + x + 1] = gx * gx + gy * gy;
        }
    const int TG22 = (int)(0.4142135623730950488016887242097 * (1 << 15) + 0.5);
    auto M = [&](int x, int y) { return mag[(size_t)(y + 1) * (w + 2) + x + 1]; };
    std::vector<uint8_t> state(n, 1);  // 0 candidate, 1 not an edge, 2 edge
    std::vector<int> stack;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int m = M(x, y);
            if (m <= low) continue;
            const int xs = dx[(size_t)y * w + x], ys = dy[(size_t)y * w + x];
            const long ax = std::abs(xs), ay = (long)std::abs(ys) << 15;
            const long tg22x = ax * TG22;
            bool keep;
            if (ay < tg22x) {
                keep = m > M(x - 1, y) && m >= M(x + 1, y);
            } else {
                const long tg67x = tg22x + (ax << 16);
                if (ay > tg67x) {
                    keep = m > M(x, y - 1) && m >= M(x, y + 1);
                } else {
                    const int s = (xs ^ ys) < 0 ? -1 : 1;
                    keep = m > M(x - s, y - 1) && m > M(x + s, y + 1);
                }
            }
            if (!keep) continue;
            const size_t i = (size_t)y * w + x;
            if (m > high) {
                state[i] = 2;
                stack.push_back((int)i);
            } else {
                state[i] = 0;
            }
        }
    while (!stack.empty()) {
        const int i = stack.back();
        stack.pop_back();
        const int x = i % w, y = i / w;
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int xx = x + ox, yy = y + oy;
                if ((ox == 0 && oy == 0) || xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
                const size_t j = (size_t)yy * w + xx;
                if (state[j] == 0) {
                    state[j] = 2;
                    stack.push_back((int)j);
                }
            }
    }
    for (size_t i = 0; i < n; ++i) out[i] = state[i] == 2 ? 255 : 0;
}

// cv::RNG (multiply-with-carry, state (uint64)-1 as HoughLinesP seeds it)
struct CvRng {
    uint64_t state = ~(uint64_t)0;
    unsigned next() {
        state = (uint64_t)(unsigned)state * 4164903690u + (unsigned)(state >> 32);
        return (unsigned)state;
    }
    int uniform(int a, int b) { return a == b ? a : (int)(next() % (unsigned)(b - a)) + a; }
};

// cv::HoughLinesP (progressive probabilistic Hough transform): points taken in
// the RNG's random order vote; a bin reaching the threshold is followed along
// its line in both directions (16-bit fixed point, gaps up to max_gap); a
// segment of length >= min_len is output and its points withdraw their votes.
std::vector<int> hough_lines_p(const uint8_t *img, int w, int h, double rho, double theta, int threshold,
                               int min_len, int max_gap) {
    const float irho = (float)(1.0 / rho);
    const int numangle = (int)std::lrint(kPi / theta);
    const int numrho = (int)std::lrint(((w + h) * 2 + 1) / rho);
    std::vector<int> accum((size_t)numangle * numrho, 0);
    std::vector<uint8_t> mask((size_t)w * h, 0);
    std::vector<float> tab(2 * (size_t)numangle);
    float ang = 0.0f;
    for (int k = 0; k < numangle; ++k, ang += (float)theta) {
        tab[2 * k] = (float)(std::cos((double)ang) * irho);
        tab[2 * k + 1] = (float)(std::sin((double)ang) * irho);
    }
    std::vector<std::pair<int, int>> nz;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            if (img[(size_t)y * w + x]) {
                mask[(size_t)y * w + x] = 1;
                nz.push_back({x, y});
            }
    std::vector<int> lines;
    CvRng rng;
    const int shift = 16;
    for (int count = (int)nz.size(); count > 0; --count) {
        const int idx = rng.uniform(0, count);
        const int j = nz[idx].first, i = nz[idx].second;
        nz[idx] = nz[count - 1];
        if (!mask[(size_t)i * w + j]) continue;
        int max_val = threshold - 1, max_n = 0;
        for (int k = 0; k < numangle; ++k) {
            int r = (int)std::lrint(j * tab[2 * k] + i * tab[2 * k + 1]);
            r += (numrho - 1) / 2;
            const int val = ++accum[(size_t)k * numrho + r];
            if (max_val < val) {
                max_val = val;
                max_n = k;
            }
        }
        if (max_val < threshold) continue;
        const float a = -tab[2 * max_n + 1], b = tab[2 * max_n];
        int x0 = j, y0 = i, dx0, dy0;
        const bool xflag = std::fabs(a) > std::fabs(b);
        if (xflag) {
            dx0 = a > 0 ? 1 : -1;
            dy0 = (int)std::lrint(b * (1 << shift) / std::fabs(a));
            y0 = (y0 << shift) + (1 << (shift - 1));
        } else {
            dy0 = b > 0 ? 1 : -1;
            dx0 = (int)std::lrint(a * (1 << shift) / std::fabs(b));
            x0 = (x0 << shift) + (1 << (shift - 1));
        }
        int end_x[2] = {j, j}, end_y[2] = {i, i};
        for (int k = 0; k < 2; ++k) {
            int gap = 0, x = x0, y = y0, dx = k ? -dx0 : dx0, dy = k ? -dy0 : dy0;
            for (;; x += dx, y += dy) {
                const int j1 = xflag ? x : x >> shift, i1 = xflag ? y >> shift : y;
                if (j1 < 0 || j1 >= w || i1 < 0 || i1 >= h) break;
                if (mask[(size_t)i1 * w + j1]) {
                    gap = 0;
                    end_x[k] = j1;
                    end_y[k] = i1;
                } else if (++gap > max_gap) {
                    break;
                }
            }
        }
        const bool good = std::abs(end_x[1] - end_x[0]) >= min_len || std::abs(end_y[1] - end_y[0]) >= min_len;
        for (int k = 0; k < 2; ++k) {
            int x = x0, y = y0, dx = k ? -dx0 : dx0, dy = k ? -dy0 : dy0;
            for (;; x += dx, y += dy) {
                const int j1 = xflag ? x : x >> shift, i1 = xflag ? y >> shift : y;
                if (j1 < 0 || j1 >= w || i1 < 0 || i1 >= h) break;
                uint8_t &md = mask[(size_t)i1 * w + j1];
                if (md) {
                    if (good)
                        for (int q = 0; q < numangle; ++q) {
                            int r = (int)std::lrint(j1 * tab[2 * q] + i1 * tab[2 * q + 1]);
                            r += (numrho - 1) / 2;
                            accum[(size_t)q * numrho + r]--;
                        }
                    md = 0;
                }
                if (i1 == end_y[k] && j1 == end_x[k]) break;
            }
        }
        if (good) lines.insert(lines.end(), {end_x[0], end_y[0], end_x[1], end_y[1]});
    }
    return lines;
}

// cv::line(thickness 1, 8-connected) of value 255: Bresenham, both ends drawn
void draw_line(std::vector<uint8_t> &img, int w, int h, int x0, int y0, int x1, int y1) {
    const int dx = std::abs(x1 - x0), dy = -std::abs(y1 - y0);
    const int sx = x0 < x1 ? 1 : -1, sy = y0 < y1 ? 1 : -1;
    int err = dx + dy;
    for (;;) {
        if (x0 >= 0 && y0 >= 0 && x0 < w && y0 < h) img[(size_t)y0 * w + x0] = 255;
        if (x0 == x1 && y0 == y1) break;
        const int e2 = 2 * err;
        if (e2 >= dy) {
            err += dy;
            x0 += sx;
        }
        if (e2 <= dx) {
            err += dx;
            y0 += sy;
        }
    }
}

// DPE.cpp:262-273: the first / last column and row take the "not edge" state
// of their inner neighbour
void border_fix(std::vector<uint8_t> &e, int w, int h) {
    if (w < 2 || h < 2) return;
    for (int y = 0; y < h; ++y) {
        if (e[(size_t)y * w + 1] == 0) e[(size_t)y * w] = 0;
        if (e[(size_t)y * w + w - 2] == 0) e[(size_t)y * w + w - 1] = 0;
    }
    for (int x = 0; x < w; ++x) {
        if (e[(size_t)w + x] == 0) e[x] = 0;
        if (e[(size_t)(h - 2) * w + x] == 0) e[(size_t)(h - 1) * w + x] = 0;
    }
}

}  // namespace

extern "C" {

void dpe_host_roberts(const uint8_t *src, int w, int h, uint8_t *dst) { roberts(src, w, h, dst); }

int dpe_host_connect(const uint8_t *bin, int w, int h, int32_t *labels, int32_t *counts, int capacity) {
    const std::vector<int> c = connect(bin, w, h, labels);
    if (counts)
        for (int i = 0; i < std::min((int)c.size(), capacity); ++i) counts[i] = c[i];
    return (int)c.size();
}

void dpe_host_canny(const uint8_t *src, int w, int h, double low, double high, uint8_t *edges) {
    canny(src, w, h, low, high, edges);
}

int dpe_host_hough_lines_p(const uint8_t *img, int w, int h, double rho, double theta, int threshold, int min_len,
                           int max_gap, int32_t *lines, int max_lines) {
    const std::vector<int> l = hough_lines_p(img, w, h, rho, theta, threshold, min_len, max_gap);
    const int n = (int)l.size() / 4;
    if (lines)
        for (int i = 0; i < std::min(n, max_lines) * 4; ++i) lines[i] = l[i];
    return n;
}

int dpe_host_edge_segment(int scale, const uint8_t *src, int w, int h, int mode, int use_canny, int high_res,
                          uint8_t *edges, int32_t *labels, int *out_w, int *out_h) {
    if (!src || w < 4 || h < 4 || (mode != 0 && mode != 1)) return -1;
    const int robthr = high_res ? 4 : 6;
    const int weak_tex_num = (int)(1.0 * h * w / (1024 << scale << scale));
    std::vector<uint8_t> down(src, src + (size_t)w * h);
    int dw = w, dh = h;
    if (high_res) {
        down = resize_u8(down, dw, dh, dw / 2, dh / 2);
        dw /= 2;
        dh /= 2;
    }
    std::vector<uint8_t> dst;
    int cw = 0, ch = 0;
    if (!use_canny) {
        down = resize_u8(down, dw, dh, dw / 2, dh / 2);
        dw /= 2;
        dh /= 2;
        const int houthr = (int)(std::min(dw, dh) / 30.0);
        dst.resize((size_t)dw * dh);
        roberts(down.data(), dw, dh, dst.data());
        threshold_u8(dst, robthr);
        std::vector<int32_t> lab0((size_t)dw * dh);
        const std::vector<int> cnt0 = connect(dst.data(), dw, dh, lab0.data());
        for (size_t k = 1; k < cnt0.size(); ++k) {
            if (cnt0[k] < weak_tex_num) continue;
            std::vector<uint8_t> border((size_t)dw * dh, 0);
            for (int y = 0; y < dh; ++y)
                for (int x = 0; x < dw; ++x) {
                    const size_t i = (size_t)y * dw + x;
                    if (lab0[i] == (int)k) continue;
                    const bool b = (x > 0 && lab0[i - 1] == (int)k) || (x < dw - 1 && lab0[i + 1] == (int)k) ||
                                   (y > 0 && lab0[i - dw] == (int)k) || (y < dh - 1 && lab0[i + dw] == (int)k);
                    if (b) border[i] = 255;
                }
            const std::vector<int> l = hough_lines_p(border.data(), dw, dh, 1.0, kPi / 180, houthr, houthr, houthr);
            for (size_t q = 0; q + 3 < l.size(); q += 4) draw_line(dst, dw, dh, l[q], l[q + 1], l[q + 2], l[q + 3]);
        }
        cw = dw;
        ch = dh;
    } else {
        // median of the image's histogram (DPE.cpp:198-214)
        std::vector<int> hist(256, 0);
        for (size_t i = 0; i < (size_t)w * h; ++i) hist[src[i]]++;
        const int half = w * h / 2;
        int median = -1, sum = 0;
        for (int i = 0; i < 255; ++i) {
            sum += hist[i];
            if (sum > half) {
                median = i;
                break;
            }
        }
        const float sigma = 0.67f;
        const int t1 = (int)((1 - sigma) * median), t2 = median;
        dst.resize((size_t)w * h);
        canny(src, w, h, t1, t2, dst.data());
        cw = w;
        ch = h;
    }
    int ow = w, oh = h;
    if (mode == 1) {
        const float factor = 1.0f / (float)(1 << scale);
        ow = (int)std::round(w * factor);
        oh = (int)std::round(h * factor);
    }
    dst = resize_u8(dst, cw, ch, ow, oh);
    threshold_u8(dst, robthr);
    border_fix(dst, ow, oh);
    if (out_w) *out_w = ow;
    if (out_h) *out_h = oh;
    if (mode == 0) {
        if (edges) std::memcpy(edges, dst.data(), dst.size());
        return 0;
    }
    std::vector<int32_t> lab((size_t)ow * oh);
    const std::vector<int> cnt = connect(dst.data(), ow, oh, lab.data());
    for (size_t i = 0; i < lab.size(); ++i)
        if (lab[i] != 0 && cnt[lab[i]] <= weak_tex_num) lab[i] = -1;
    if (labels) std::memcpy(labels, lab.data(), lab.size() * sizeof(int32_t));
    return 0;
}

}  // extern "C"
