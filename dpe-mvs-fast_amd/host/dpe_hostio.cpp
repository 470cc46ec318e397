// dpe_hostio.cpp -- libdpe_host.so: the reference's file formats, image pyramid,
// state rescaling and support maps for the C++ driver, without OpenCV or Boost
// (declarations and citations in include/dpe_host.h).
#include "dpe_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

extern "C" {

int dpe_host_type_size(int type) {
    switch (type) {
        case DPE_HOST_CV_8UC1: return 1;
        case DPE_HOST_CV_32SC1: return 4;
        case DPE_HOST_CV_32FC1: return 4;
        case DPE_HOST_CV_32FC3: return 12;
        default: return 0;
    }
}

int dpe_host_write_bin_mat(const char *path, int rows, int cols, int type, const void *data) {
    const int es = dpe_host_type_size(type);
    if (!path || rows < 0 || cols < 0 || es == 0 || (!data && rows * cols > 0)) return -1;
    FILE *f = std::fopen(path, "wb");
    if (!f) return -1;
    const int32_t hdr[4] = {1, rows, cols, type};
    bool ok = std::fwrite(hdr, sizeof(hdr), 1, f) == 1;
    const size_t bytes = (size_t)rows * cols * es;
    if (ok && bytes) ok = std::fwrite(data, 1, bytes, f) == bytes;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : -1;
}

int dpe_host_read_bin_mat(const char *path, int *rows, int *cols, int *type, void *data, size_t capacity) {
    if (!path) return -1;
    FILE *f = std::fopen(path, "rb");
    if (!f) return -1;
    int32_t hdr[4];
    if (std::fread(hdr, sizeof(hdr), 1, f) != 1) {
        std::fclose(f);
        return -1;
    }
    if (hdr[0] != 1) {
        std::fclose(f);
        return -2;
    }
    if (rows) *rows = hdr[1];
    if (cols) *cols = hdr[2];
    if (type) *type = hdr[3];
    int st = 0;
    if (data) {
        const int es = dpe_host_type_size(hdr[3]);
        const size_t bytes = (size_t)hdr[1] * hdr[2] * es;
        if (es == 0 || hdr[1] < 0 || hdr[2] < 0) st = -1;
        else if (bytes > capacity) st = -3;
        else if (bytes && std::fread(data, 1, bytes, f) != bytes) st = -1;
    }
    std::fclose(f);
    return st;
}

int dpe_host_read_camera(const char *path, dpe_camera *cam) {
    if (!path || !cam) return -1;
    std::ifstream in(path);
    if (!in) return -1;
    dpe_camera c{};
    std::string word;
    in >> word;  // "extrinsic"
    for (int i = 0; i < 3; ++i) in >> c.R[3 * i + 0] >> c.R[3 * i + 1] >> c.R[3 * i + 2] >> c.t[i];
    float last_row[4];
    in >> last_row[0] >> last_row[1] >> last_row[2] >> last_row[3];
    in >> word;  // "intrinsic"
    for (int i = 0; i < 9; ++i) in >> c.K[i];
    // camera centre in world coordinates, in double (DPE.cpp:364-367)
    for (int j = 0; j < 3; ++j)
        c.c[j] = -(float)((double)c.R[0 + j] * (double)c.t[0] + (double)c.R[3 + j] * (double)c.t[1] +
                          (double)c.R[6 + j] * (double)c.t[2]);
    float interval = 0, depth_num = 0;
    in >> c.depth_min >> interval >> depth_num >> c.depth_max;
    if (in.fail()) return -1;
    *cam = c;
    return 0;
}

int dpe_host_read_pairs(const char *path, int max_problems, int32_t *ref_ids, int32_t *src_counts, int32_t *srcs) {
    std::ifstream file(path ? path : "");
    if (!file) return -1;
    std::string line;
    if (!std::getline(file, line)) return -1;
    int num = 0;
    {
        std::istringstream iss(line);
        iss >> num;
    }
    int n = 0;
    for (int i = 0; i < num; ++i) {
        int ref = 0, count = 0;
        if (!std::getline(file, line)) break;
        {
            std::istringstream iss(line);
            iss >> ref;
        }
        if (!std::getline(file, line)) break;
        std::istringstream iss(line);
        iss >> count;
        int kept = 0;
        std::vector<int32_t> ids;
        for (int j = 0; j < count; ++j) {
            int id = 0;
            float score = 0.0f;
            iss >> id >> score;
            if (score <= 0.0f) continue;  // GenerateSampleList drops non-positive scores
            ids.push_back(id);
        }
        if (n < max_problems) {
            if (ref_ids) ref_ids[n] = ref;
            for (int32_t id : ids) {
                if (kept >= DPE_MAX_IMAGES) break;
                if (srcs) srcs[(size_t)n * DPE_MAX_IMAGES + kept] = id;
                kept++;
            }
            if (src_counts) src_counts[n] = (int32_t)ids.size();
        }
        n++;
    }
    return n;
}

void dpe_host_resize_linear(const float *src, int w, int h, float *dst, int dw, int dh) {
    if (dw == w && dh == h) {
        std::memcpy(dst, src, sizeof(float) * (size_t)w * h);
        return;
    }
    if (dw * 2 == w && dh * 2 == h) {
        // exact factor 2: the area (2x2 mean) path OpenCV takes for this case
        for (int y = 0; y < dh; ++y)
            for (int x = 0; x < dw; ++x) {
                const float *s0 = src + (size_t)(2 * y) * w + 2 * x;
                const float *s1 = s0 + w;
                dst[(size_t)y * dw + x] = (s0[0] + s0[1] + s1[0] + s1[1]) * 0.25f;
            }
        return;
    }
    const double sx = (double)w / dw, sy = (double)h / dh;
    std::vector<int> xo(dw);
    std::vector<float> xa(dw);
    for (int x = 0; x < dw; ++x) {
        float fx = (float)((x + 0.5) * sx - 0.5);
        int ix = (int)std::floor(fx);
        fx -= ix;
        if (ix < 0) ix = 0, fx = 0.0f;
        if (ix >= w - 1) ix = w - 1, fx = 0.0f;
        xo[x] = ix;
        xa[x] = fx;
    }
    std::vector<float> r0(dw), r1(dw);
    for (int y = 0; y < dh; ++y) {
        float fy = (float)((y + 0.5) * sy - 0.5);
        int iy = (int)std::floor(fy);
        fy -= iy;
        if (iy < 0) iy = 0, fy = 0.0f;
        if (iy >= h - 1) iy = h - 1, fy = 0.0f;
        const int iy1 = std::min(iy + 1, h - 1);
        for (int x = 0; x < dw; ++x) {
            const int ix1 = std::min(xo[x] + 1, w - 1);
            const float *a = src + (size_t)iy * w, *b = src + (size_t)iy1 * w;
            r0[x] = a[xo[x]] * (1.0f - xa[x]) + a[ix1] * xa[x];
            r1[x] = b[xo[x]] * (1.0f - xa[x]) + b[ix1] * xa[x];
        }
        for (int x = 0; x < dw; ++x) dst[(size_t)y * dw + x] = r0[x] * (1.0f - fy) + r1[x] * fy;
    }
}

void dpe_host_rescale_nearest(const void *src, int w, int h, void *dst, int dw, int dh, int elem_size) {
    const char *s = (const char *)src;
    char *d = (char *)dst;
    if (dw == w && dh == h) {
        std::memcpy(d, s, (size_t)w * h * elem_size);
        return;
    }
    const float scale_x = dw / (float)w, scale_y = dh / (float)h;
    for (int r = 0; r < dh; ++r)
        for (int c = 0; c < dw; ++c) {
            const int o_r = (int)(r / scale_y), o_c = (int)(c / scale_x);
            char *out = d + ((size_t)r * dw + c) * elem_size;
            if (o_r < 0 || o_c < 0 || o_r >= h || o_c >= w) {
                std::memset(out, 0, elem_size);  // the reference leaves these uninitialised
                continue;
            }
            std::memcpy(out, s + ((size_t)o_r * w + o_c) * elem_size, elem_size);
        }
}

}  // extern "C"

namespace {

// DPE.cpp:262-273: the first / last column and row copy the "not edge" state inwards
void border_rule(std::vector<uint8_t> &e, int w, int h) {
    for (int y = 0; y < h; ++y)
        if (e[(size_t)y * w + 1 % w] == 0) e[(size_t)y * w] = 0;
    for (int y = 0; y < h; ++y)
        if (e[(size_t)y * w + std::max(w - 2, 0)] == 0) e[(size_t)y * w + w - 1] = 0;
    for (int x = 0; x < w; ++x)
        if (e[(size_t)std::min(1, h - 1) * w + x] == 0) e[x] = 0;
    for (int x = 0; x < w; ++x)
        if (e[(size_t)std::max(h - 2, 0) * w + x] == 0) e[(size_t)(h - 1) * w + x] = 0;
}

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Sobel magnitude (derivative [-1, 0, 1] along one axis, smoothing [1, 2, 1]
// along the other, borders by nearest), threshold at an order statistic.
void edge_map(const uint8_t *img, int w, int h, uint8_t *out) {
    const size_t n = (size_t)w * h;
    std::vector<float> f(n), t(n), gx(n), gy(n);
    for (size_t i = 0; i < n; ++i) f[i] = (float)img[i];
    auto at = [&](const std::vector<float> &a, int x, int y) { return a[(size_t)clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)]; };
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) t[(size_t)y * w + x] = at(f, x + 1, y) - at(f, x - 1, y);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) gx[(size_t)y * w + x] = at(t, x, y - 1) + 2.0f * at(t, x, y) + at(t, x, y + 1);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) t[(size_t)y * w + x] = at(f, x, y + 1) - at(f, x, y - 1);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) gy[(size_t)y * w + x] = at(t, x - 1, y) + 2.0f * at(t, x, y) + at(t, x + 1, y);
    std::vector<float> mag(n);
    for (size_t i = 0; i < n; ++i) mag[i] = std::hypot(gx[i], gy[i]);
    std::vector<float> sorted(mag);
    const size_t k = (size_t)(0.85 * (double)(n - 1));
    std::nth_element(sorted.begin(), sorted.begin() + k, sorted.end());
    const float thr = std::max(sorted[k], 16.0f);
    std::vector<uint8_t> e(n);
    for (size_t i = 0; i < n; ++i) e[i] = mag[i] > thr ? 255 : 0;
    border_rule(e, w, h);
    std::memcpy(out, e.data(), n);
}

// Roberts edges (threshold 4) on a 4x box-downsampled image, nearest upsampling,
// border rule, 4-connected regions of non-edge pixels numbered in raster order
// of their first pixel; regions of at most rows*cols / (1024 << scale << scale)
// pixels become -1, edge pixels 0.
void label_map(const uint8_t *img, int w, int h, int scale, int32_t *out) {
    const int h4 = std::max(1, h / 4), w4 = std::max(1, w / 4);
    std::vector<int32_t> s((size_t)h4 * w4);
    for (int y = 0; y < h4; ++y)
        for (int x = 0; x < w4; ++x) {
            float sum = 0.0f;
            for (int dy = 0; dy < 4; ++dy)
                for (int dx = 0; dx < 4; ++dx) {
                    const int yy = std::min(4 * y + dy, h - 1), xx = std::min(4 * x + dx, w - 1);
                    sum += (float)img[(size_t)yy * w + xx];
                }
            s[(size_t)y * w4 + x] = (int32_t)std::nearbyint(sum / 16.0f);
        }
    std::vector<int32_t> rob((size_t)h4 * w4, 0);
    for (int y = 0; y < h4; ++y)
        for (int x = 0; x < w4; ++x) {
            int32_t v = 0;
            if (y < h4 - 1 && x < w4 - 1) {
                const int32_t t1 = s[(size_t)y * w4 + x] - s[(size_t)(y + 1) * w4 + x + 1];
                const int32_t t2 = s[(size_t)(y + 1) * w4 + x] - s[(size_t)y * w4 + x + 1];
                v = (int32_t)std::sqrt((float)(t1 * t1 + t2 * t2));
            }
            if (y == 0 || y == h4 - 1 || x == 0 || x == w4 - 1) v = 50;
            rob[(size_t)y * w4 + x] = v;
        }
    const size_t n = (size_t)w * h;
    std::vector<uint8_t> e(n);
    for (int y = 0; y < h; ++y) {
        const int yi = std::min((int)((long)y * h4 / h), h4 - 1);
        for (int x = 0; x < w; ++x) {
            const int xi = std::min((int)((long)x * w4 / w), w4 - 1);
            e[(size_t)y * w + x] = rob[(size_t)yi * w4 + xi] > 4 ? 255 : 0;
        }
    }
    border_rule(e, w, h);
    // union-find over 4-connected non-edge pixels
    std::vector<int32_t> parent(n);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int32_t a) {
        while (parent[a] != a) {
            parent[a] = parent[parent[a]];
            a = parent[a];
        }
        return a;
    };
    auto unite = [&](int32_t a, int32_t b) {
        a = find(a);
        b = find(b);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    };
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int32_t i = y * w + x;
            if (e[i]) continue;
            if (x > 0 && !e[i - 1]) unite(i, i - 1);
            if (y > 0 && !e[i - w]) unite(i, i - w);
        }
    std::vector<int32_t> label(n, 0), root_label(n, 0);
    int32_t next = 0;
    for (size_t i = 0; i < n; ++i) {
        if (e[i]) continue;
        const int32_t r = find((int32_t)i);
        if (root_label[r] == 0) root_label[r] = ++next;
        label[i] = root_label[r];
    }
    std::vector<int64_t> counts((size_t)next + 1, 0);
    for (size_t i = 0; i < n; ++i) counts[label[i]]++;
    const int64_t weak_tex_num = (int64_t)(1.0 * h * w / (double)(1024 << scale << scale));
    for (size_t i = 0; i < n; ++i)
        out[i] = (label[i] != 0 && counts[label[i]] <= weak_tex_num) ? -1 : label[i];
}

}  // namespace

extern "C" {

void dpe_host_support_maps(const uint8_t *image, int w, int h, int scale, uint8_t *edges, int32_t *labels) {
    if (!image || w <= 0 || h <= 0) return;
    if (edges) edge_map(image, w, h, edges);
    if (labels) label_map(image, w, h, scale, labels);
}

int dpe_host_write_ply(const char *path, const dpe_point *points, int64_t count) {
    FILE *f = std::fopen(path ? path : "", "wb");
    if (!f) return -1;
    std::fprintf(f,
                 "ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
                 "property float z\nproperty uchar diffuse_blue\nproperty uchar diffuse_green\n"
                 "property uchar diffuse_red\nend_header\n",
                 (int)count);
    bool ok = true;
    std::vector<char> buf;
    buf.reserve(15 * 4096);
    for (int64_t i = 0; i < count && ok; ++i) {
        const dpe_point &p = points[i];
        char rec[15];
        std::memcpy(rec, &p.x, 4);
        std::memcpy(rec + 4, &p.y, 4);
        std::memcpy(rec + 8, &p.z, 4);
        rec[12] = (char)p.b;
        rec[13] = (char)p.g;
        rec[14] = (char)p.r;
        buf.insert(buf.end(), rec, rec + 15);
        if (buf.size() >= 15 * 4096) {
            ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
            buf.clear();
        }
    }
    if (ok && !buf.empty()) ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : -1;
}

}  // extern "C"
