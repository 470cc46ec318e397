// dpe_oracle.cpp -- CPU restatement of DPE-MVS-fast's per-reference-image
// PatchMatch (DPE.cu:9-3262) and depth-map fusion (DPE.cpp:1290-1548).
//
// TEST INFRASTRUCTURE ONLY.  This is the correctness oracle: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// as the checker or the reported CPU baseline.  The product never links it.
//
// Every function follows the reference function named in its comment, statement
// by statement, with these documented resolutions of reference behaviour that
// is undefined, racy or NVIDIA-specific (shared with the HIP path):
//   * numerics contract: oracle_numerics.h;
//   * textures (DPE.cpp:1083-1094): clamp-to-edge, exact fp32 bilinear weights,
//     sample position (x, y) = tex2D coordinate minus 0.5;
//   * same-colour reads in the edge-adaptive strong branch (DPE.cu:1279-1282):
//     every read of another pixel's cost, plane or selected views during a
//     strong half-iteration sees the value from before that half-iteration;
//   * reads the reference makes outside its buffers return 0;
//   * the radius map is initialised to strong_radius at the start of a call
//     (the reference reads it uninitialised for weak pixels with < 3 neighbours);
//   * GenNeighbours' two unsequenced curand() calls (DPE.cu:2186-2187) draw the
//     sign first, then the magnitude.
//
// Every function is pinned by an independent check (tests/test_oracle*.py):
// closed forms, float64 restatements of definitions, brute force, library
// routines, invariants, known-answer vectors.  Unpinnable here: the reference's
// exact outputs (clock64-seeded curand), OpenCV's preprocessing, curand's
// XORWOW streams and NVIDIA's texture rounding, none of which this file uses.
#include "dpe_oracle.h"
#include "oracle_numerics.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace oracle {
namespace {

struct F4 { float x, y, z, w; };
struct F3 { float x, y, z; };
struct F2 { float x, y; };
struct S2 { int16_t x, y; };
struct I2 { int x, y; };

constexpr int MAXIMG = DPE_MAX_IMAGES;
constexpr int NNUM = DPE_NEIGHBOUR_NUM;

// NCC families (dpe_stats.family_* of the GPU library) plus the evaluations
// whose results the reference never reads; tl_family names the family of the
// work a thread is doing.
constexpr int kFamFull = DPE_FAMILY_FULL, kFamRefine = DPE_FAMILY_REFINE, kFamPost = DPE_FAMILY_POST,
              kFamWeak = DPE_FAMILY_WEAK, kFamDead = DPE_NCC_FAMILIES, kFamilies = DPE_NCC_FAMILIES + 1;
thread_local int tl_family = kFamFull;
// Index of the parallel_for worker running on this thread: each worker counts
// into its own cache line of Ctx::tally (no shared atomics in the hot loop).
thread_local int tl_worker = 0;
struct alignas(64) Tally {
    int64_t hyps[kFamilies] = {}, taps[kFamilies] = {};
};

inline F4 mk4(float x, float y, float z, float w) { return F4{x, y, z, w}; }

// ---------------------------------------------------------------- call context
struct Ctx {
    int W = 0, H = 0, N = 0, P = 0;
    dpe_params prm;
    dpe_camera cams[MAXIMG];
    const float *img[MAXIMG];
    const float *gdepth[MAXIMG];
    const uint8_t *edge = nullptr, *edge_low = nullptr;
    int low_w = 0, low_h = 0;
    const int32_t *label = nullptr;
    int nthreads = 1;
    // state (DataPassHelper, DPE.h:104-138)
    std::vector<F4> planes, fit_planes;
    std::vector<float> costs, complex_map;
    std::vector<uint32_t> selected;
    std::vector<uint8_t> view_weight, weak, weak_reliable;
    std::vector<S2> nearest, neighbours, edge_neigh, label_bound;
    std::vector<int> nmap, radius;
    int weak_count = 0;
    // pre-half-iteration copies (snapshot semantics)
    std::vector<F4> snap_planes;
    std::vector<float> snap_costs;
    std::vector<uint32_t> snap_selected;
    // views whose NCC cost for the pixel's current plane an earlier evaluation
    // already gave (the strong propagation's re-evaluation of the current plane
    // counts as dead work for them; values are not reused here)
    std::vector<uint32_t> known;
    // Per candidate slot (1..16: GPU slot 1 + d of the 8 scanned candidates, 9 + d
    // of the 8 fixed-step ones) and pixel, the plane of the slot's last evaluated
    // NCC row in this call (all-ones bits: none).  A row is a pure function of
    // (pixel, plane, view), so a candidate repeating that plane gives the same
    // costs; the GPU keeps such rows from the pixel's previous half-iteration
    // and does not evaluate them, so they count as dead work here.
    mutable std::vector<F4> last_cand;
    // oracle_process_fp64: the multi-view cost is evaluated in double
    bool cost64 = false;
    // GenNeighbours constants computed as DPE.cu:2151-2157 does (host double math)
    float gn_cos_angle = 0, gn_sin_angle = 0, gn_threshold = 0;
    int gn_shift_range = 1;
    // counters, in total and per NCC family (see tl_family)
    std::vector<Tally> tally;  // one per worker
};

template <class F>
void parallel_for(int n, int nthreads, F &&fn) {
    if (nthreads <= 1 || n < 256) {
        tl_worker = 0;
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<int> next{0};
    const int chunk = 64;
    for (int t = 0; t < nthreads; ++t) {
        pool.emplace_back([&, t]() {
            tl_worker = t;
            for (;;) {
                const int s = next.fetch_add(chunk);
                if (s >= n) break;
                const int e = std::min(n, s + chunk);
                for (int i = s; i < e; ++i) fn(i);
            }
        });
    }
    for (auto &th : pool) th.join();
}

// ------------------------------------------------------------ small helpers
// sort_small, DPE.cu:9-18
void sort_small(float *d, int n) {
    int j;
    for (int i = 1; i < n; i++) {
        float tmp = d[i];
        for (j = i; j >= 1 && tmp < d[j - 1]; j--) d[j] = d[j - 1];
        d[j] = tmp;
    }
}

// sort_small_weighted(short2), DPE.cu:20-33
void sort_small_weighted(S2 *points, float *w, int n) {
    int j;
    for (int i = 1; i < n; i++) {
        S2 tmp = points[i];
        float tmp_w = w[i];
        for (j = i; j >= 1 && tmp_w < w[j - 1]; j--) {
            points[j] = points[j - 1];
            w[j] = w[j - 1];
        }
        points[j] = tmp;
        w[j] = tmp_w;
    }
}

// FindMinCostIndex, DPE.cu:50-61 (ties -> last)
int FindMinCostIndex(const float *costs, int n) {
    float min_cost = costs[0];
    int idx_min = 0;
    for (int idx = 1; idx < n; ++idx) {
        if (costs[idx] <= min_cost) {
            min_cost = costs[idx];
            idx_min = idx;
        }
    }
    return idx_min;
}

// setBit / unSetBit / isSet, DPE.cu:76-89 (unSetBit clears bits 0..n)
inline void setBit(uint32_t *input, uint32_t n) { *input |= (uint32_t)(1u << n); }
inline void unSetBit(uint32_t *input, uint32_t n) { *input &= (uint32_t)(0xFFFFFFFEu << n); }
inline int isSet(uint32_t input, uint32_t n) { return (input >> n) & 1; }

inline float Vec3Dot(const F4 &a, const F4 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float Vec3Dot(const F3 &a, const F3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// NormalizeVec3/2, DPE.cu:272-295 (rsqrtf restated as 1/sqrtf)
inline void NormalizeVec3(F4 *v) {
    const float n2 = v->x * v->x + v->y * v->y + v->z * v->z;
    const float inv = 1.0f / std::sqrt(n2);
    v->x *= inv; v->y *= inv; v->z *= inv;
}
inline void NormalizeVec2(F2 *v) {
    const float n2 = v->x * v->x + v->y * v->y;
    const float inv = 1.0f / std::sqrt(n2);
    v->x *= inv; v->y *= inv;
}

// Float to int conversion with defined NaN / overflow behaviour (value clamped
// to [lo, hi] first, NaN -> lo).
inline int f2i_clamp(float f, int lo, int hi) {
    if (!(f >= (float)lo)) return lo;
    if (f > (float)hi) return hi;
    return (int)f;
}

// TransformPDFToCDF, DPE.cu:297-311
void TransformPDFToCDF(float *probs, int num) {
    float sum = 0.0f;
    for (int i = 0; i < num; ++i) sum += probs[i];
    const float inv = 1.0f / sum;
    float cum = 0.0f;
    for (int i = 0; i < num; ++i) {
        const float prob = probs[i] * inv;
        cum += prob;
        probs[i] = cum;
    }
}

// Get3DPoint, DPE.cu:313-325
inline void Get3DPoint(const dpe_camera &cam, int px, int py, float depth, float *X) {
    X[0] = depth * ((float)px - cam.K[2]) / cam.K[0];
    X[1] = depth * ((float)py - cam.K[5]) / cam.K[4];
    X[2] = depth;
}

// GetViewDirection, DPE.cu:327-339
F4 GetViewDirection(const dpe_camera &cam, int px, int py, float depth) {
    float X[3];
    Get3DPoint(cam, px, py, depth, X);
    const float norm = std::sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    return mk4(X[0] / norm, X[1] / norm, X[2] / norm, 0.0f);
}

// GetDistance2Origin, DPE.cu:341-346
float GetDistance2Origin(const dpe_camera &cam, int px, int py, float depth, const F4 &n) {
    float X[3];
    Get3DPoint(cam, px, py, depth, X);
    return -(n.x * X[0] + n.y * X[1] + n.z * X[2]);
}

// ComputeDepthfromPlaneHypothesis, DPE.cu:360-363
float ComputeDepthfromPlaneHypothesis(const dpe_camera &cam, const F4 &plane, int px, int py) {
    return -plane.w * cam.K[0] /
           (((float)px - cam.K[2]) * plane.x + (cam.K[0] / cam.K[4]) * ((float)py - cam.K[5]) * plane.y +
            cam.K[0] * plane.z);
}

// GenerateRandomNormal, DPE.cu:365-391
F4 GenerateRandomNormal(const dpe_camera &cam, int px, int py, Rng &rng, float depth) {
    F4 normal;
    float q1 = 1.0f, q2 = 1.0f, s = 2.0f;
    while (s >= 1.0f) {
        q1 = 2.0f * rng.uniform() - 1.0f;
        q2 = 2.0f * rng.uniform() - 1.0f;
        s = q1 * q1 + q2 * q2;
    }
    const float sq = std::sqrt(1.0f - s);
    normal.x = 2.0f * q1 * sq;
    normal.y = 2.0f * q2 * sq;
    normal.z = 1.0f - 2.0f * s;
    normal.w = 0;
    const F4 vd = GetViewDirection(cam, px, py, depth);
    const float dot = normal.x * vd.x + normal.y * vd.y + normal.z * vd.z;
    if (dot > 0.0f) {
        normal.x = -normal.x;
        normal.y = -normal.y;
        normal.z = -normal.z;
    }
    NormalizeVec3(&normal);
    return normal;
}

// GeneratePerturbedNormal, DPE.cu:393-428
F4 GeneratePerturbedNormal(const dpe_camera &cam, int px, int py, const F4 &normal, Rng &rng,
                           float perturbation) {
    const F4 vd = GetViewDirection(cam, px, py, 1.0f);
    const float a1 = (rng.uniform() - 0.5f) * perturbation;
    const float a2 = (rng.uniform() - 0.5f) * perturbation;
    const float a3 = (rng.uniform() - 0.5f) * perturbation;
    float s1, c1, s2, c2, s3, c3;
    o_sincosf(a1, &s1, &c1);
    o_sincosf(a2, &s2, &c2);
    o_sincosf(a3, &s3, &c3);
    float R[9];
    R[0] = c2 * c3;
    R[1] = c3 * s1 * s2 - c1 * s3;
    R[2] = s1 * s3 + c1 * c3 * s2;
    R[3] = c2 * s3;
    R[4] = c1 * c3 + s1 * s2 * s3;
    R[5] = c1 * s2 * s3 - c3 * s1;
    R[6] = -s2;
    R[7] = c2 * s1;
    R[8] = c1 * c2;
    F4 np;
    np.x = R[0] * normal.x + R[1] * normal.y + R[2] * normal.z;
    np.y = R[3] * normal.x + R[4] * normal.y + R[5] * normal.z;
    np.z = R[6] * normal.x + R[7] * normal.y + R[8] * normal.z;
    np.w = normal.w;  // Mat33DotVec3 leaves w untouched (DPE.cu:91-96)
    if (Vec3Dot(np, vd) >= 0.0f) np = normal;
    NormalizeVec3(&np);
    return np;
}

// GenerateRandomPlaneHypothesis, DPE.cu:430-436
F4 GenerateRandomPlaneHypothesis(const dpe_camera &cam, int px, int py, Rng &rng, float dmin, float dmax) {
    const float depth = rng.uniform() * (dmax - dmin) + dmin;
    F4 plane = GenerateRandomNormal(cam, px, py, rng, depth);
    plane.w = GetDistance2Origin(cam, px, py, depth, plane);
    return plane;
}

// ComputeHomography, DPE.cu:457-517
void ComputeHomography(const dpe_camera &ref, const dpe_camera &src, const F4 &plane, float *H) {
    float ref_C[3], src_C[3];
    ref_C[0] = -(ref.R[0] * ref.t[0] + ref.R[3] * ref.t[1] + ref.R[6] * ref.t[2]);
    ref_C[1] = -(ref.R[1] * ref.t[0] + ref.R[4] * ref.t[1] + ref.R[7] * ref.t[2]);
    ref_C[2] = -(ref.R[2] * ref.t[0] + ref.R[5] * ref.t[1] + ref.R[8] * ref.t[2]);
    src_C[0] = -(src.R[0] * src.t[0] + src.R[3] * src.t[1] + src.R[6] * src.t[2]);
    src_C[1] = -(src.R[1] * src.t[0] + src.R[4] * src.t[1] + src.R[7] * src.t[2]);
    src_C[2] = -(src.R[2] * src.t[0] + src.R[5] * src.t[1] + src.R[8] * src.t[2]);
    float Rr[9], Cr[3], tr[3];
    Rr[0] = src.R[0] * ref.R[0] + src.R[1] * ref.R[1] + src.R[2] * ref.R[2];
    Rr[1] = src.R[0] * ref.R[3] + src.R[1] * ref.R[4] + src.R[2] * ref.R[5];
    Rr[2] = src.R[0] * ref.R[6] + src.R[1] * ref.R[7] + src.R[2] * ref.R[8];
    Rr[3] = src.R[3] * ref.R[0] + src.R[4] * ref.R[1] + src.R[5] * ref.R[2];
    Rr[4] = src.R[3] * ref.R[3] + src.R[4] * ref.R[4] + src.R[5] * ref.R[5];
    Rr[5] = src.R[3] * ref.R[6] + src.R[4] * ref.R[7] + src.R[5] * ref.R[8];
    Rr[6] = src.R[6] * ref.R[0] + src.R[7] * ref.R[1] + src.R[8] * ref.R[2];
    Rr[7] = src.R[6] * ref.R[3] + src.R[7] * ref.R[4] + src.R[8] * ref.R[5];
    Rr[8] = src.R[6] * ref.R[6] + src.R[7] * ref.R[7] + src.R[8] * ref.R[8];
    Cr[0] = ref_C[0] - src_C[0];
    Cr[1] = ref_C[1] - src_C[1];
    Cr[2] = ref_C[2] - src_C[2];
    tr[0] = src.R[0] * Cr[0] + src.R[1] * Cr[1] + src.R[2] * Cr[2];
    tr[1] = src.R[3] * Cr[0] + src.R[4] * Cr[1] + src.R[5] * Cr[2];
    tr[2] = src.R[6] * Cr[0] + src.R[7] * Cr[1] + src.R[8] * Cr[2];
    // divisions as x * RN(1/y): the reference's --use_fast_math build
    // (CMakeLists.txt:26) divides as x * rcp(y)
    const float iw = 1.0f / plane.w, ik0 = 1.0f / ref.K[0], ik4 = 1.0f / ref.K[4];
    H[0] = Rr[0] - tr[0] * plane.x * iw;
    H[1] = Rr[1] - tr[0] * plane.y * iw;
    H[2] = Rr[2] - tr[0] * plane.z * iw;
    H[3] = Rr[3] - tr[1] * plane.x * iw;
    H[4] = Rr[4] - tr[1] * plane.y * iw;
    H[5] = Rr[5] - tr[1] * plane.z * iw;
    H[6] = Rr[6] - tr[2] * plane.x * iw;
    H[7] = Rr[7] - tr[2] * plane.y * iw;
    H[8] = Rr[8] - tr[2] * plane.z * iw;
    float tmp[9];
    tmp[0] = H[0] * ik0;
    tmp[1] = H[1] * ik4;
    tmp[2] = -H[0] * ref.K[2] * ik0 - H[1] * ref.K[5] * ik4 + H[2];
    tmp[3] = H[3] * ik0;
    tmp[4] = H[4] * ik4;
    tmp[5] = -H[3] * ref.K[2] * ik0 - H[4] * ref.K[5] * ik4 + H[5];
    tmp[6] = H[6] * ik0;
    tmp[7] = H[7] * ik4;
    tmp[8] = -H[6] * ref.K[2] * ik0 - H[7] * ref.K[5] * ik4 + H[8];
    H[0] = src.K[0] * tmp[0] + src.K[2] * tmp[6];
    H[1] = src.K[0] * tmp[1] + src.K[2] * tmp[7];
    H[2] = src.K[0] * tmp[2] + src.K[2] * tmp[8];
    H[3] = src.K[4] * tmp[3] + src.K[5] * tmp[6];
    H[4] = src.K[4] * tmp[4] + src.K[5] * tmp[7];
    H[5] = src.K[4] * tmp[5] + src.K[5] * tmp[8];
    H[6] = src.K[8] * tmp[6];
    H[7] = src.K[8] * tmp[7];
    H[8] = src.K[8] * tmp[8];
}

// ComputeCorrespondingPoint, DPE.cu:519-526 (projection written with fmaf).
// Numerics contract: the divisions by w are multiplications by the correctly
// rounded reciprocal, (u * r, v * r) with r = 1.0f / w (the reference's
// --use_fast_math build, CMakeLists.txt:26, does not divide correctly rounded
// either).
inline F2 ComputeCorrespondingPoint(const float *H, int px, int py) {
    const float x = (float)px, y = (float)py;
    const float u = std::fma(H[0], x, std::fma(H[1], y, H[2]));
    const float v = std::fma(H[3], x, std::fma(H[4], y, H[5]));
    const float w = std::fma(H[6], x, std::fma(H[7], y, H[8]));
    const float r = 1.0f / w;
    return F2{u * r, v * r};
}

// TransformNormal, DPE.cu:528-536 (camera -> world)
F4 TransformNormal(const dpe_camera &cam, const F4 &p) {
    F4 r;
    r.x = cam.R[0] * p.x + cam.R[3] * p.y + cam.R[6] * p.z;
    r.y = cam.R[1] * p.x + cam.R[4] * p.y + cam.R[7] * p.z;
    r.z = cam.R[2] * p.x + cam.R[5] * p.y + cam.R[8] * p.z;
    r.w = p.w;
    return r;
}

// TransformNormal2RefCam, DPE.cu:538-546 (world -> camera)
F4 TransformNormal2RefCam(const dpe_camera &cam, const F4 &p) {
    F4 r;
    r.x = cam.R[0] * p.x + cam.R[1] * p.y + cam.R[2] * p.z;
    r.y = cam.R[3] * p.x + cam.R[4] * p.y + cam.R[5] * p.z;
    r.z = cam.R[6] * p.x + cam.R[7] * p.y + cam.R[8] * p.z;
    r.w = p.w;
    return r;
}

// --------------------------------------------------------------- texturing
// Texel at integer coordinates with clamp-to-edge (tex2D at (x+0.5, y+0.5)).
inline float texel(const float *im, int W, int H, int x, int y) {
    x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
    y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
    return im[(size_t)y * W + x];
}

// Bilinear sample at texel-space position (u, v) (tex2D at (u+0.5, v+0.5),
// linear filter, clamp-to-edge), exact fp32 weights.
inline float bilinear(const float *im, int W, int H, float u, float v) {
    const float fx = std::floor(u), fy = std::floor(v);
    const float a = u - fx, b = v - fy;
    const int ix = f2i_clamp(fx, -1, W - 1);
    const int iy = f2i_clamp(fy, -1, H - 1);
    const float t00 = texel(im, W, H, ix, iy);
    const float t10 = texel(im, W, H, ix + 1, iy);
    const float t01 = texel(im, W, H, ix, iy + 1);
    const float t11 = texel(im, W, H, ix + 1, iy + 1);
    const float top = std::fma(a, t10 - t00, t00);
    const float bot = std::fma(a, t11 - t01, t01);
    return std::fma(b, bot - top, top);
}

// ComputeBilateralWeight, DPE.cu:554-559
inline float ComputeBilateralWeight(float x_dist, float y_dist, float pix, float center_pix, float ss,
                                    float sc) {
    const float spatial = std::sqrt(x_dist * x_dist + y_dist * y_dist);
    const float color = std::fabs(pix - center_pix);
    return o_expf(-spatial / (2.0f * ss * ss) - color / (2.0f * sc * sc));
}

// One bilateral NCC window (the shared body of DPE.cu:613-672 and 719-779).
// Centre (cx, cy) of the window, ref_center_pix the colour weight reference.
float BilateralNCCWindow(const Ctx &C, int src_idx, const float *H, int cx, int cy, float ref_center_pix,
                         int radius, int increment) {
    const dpe_camera &src = C.cams[src_idx];
    const float *refim = C.img[0];
    const float *srcim = C.img[src_idx];
    const float ss = C.prm.sigma_spatial, sc = C.prm.sigma_color;
    float sum_ref = 0.0f, sum_ref_ref = 0.0f, sum_src = 0.0f, sum_src_src = 0.0f, sum_ref_src = 0.0f;
    float wsum = 0.0f;
    int taps = 0;
    for (int i = -radius; i <= radius; i += increment) {
        float r_ref = 0.0f, r_src = 0.0f, r_rr = 0.0f, r_ss = 0.0f, r_rs = 0.0f, r_w = 0.0f;
        for (int j = -radius; j <= radius; j += increment) {
            const int rx = cx + i, ry = cy + j;
            const float ref_pix = texel(refim, C.W, C.H, rx, ry);
            const F2 sp = ComputeCorrespondingPoint(H, rx, ry);
            const float src_pix = bilinear(srcim, src.width, src.height, sp.x, sp.y);
            const float w = ComputeBilateralWeight((float)i, (float)j, ref_pix, ref_center_pix, ss, sc);
            const float wr = w * ref_pix;
            r_ref = std::fma(w, ref_pix, r_ref);
            r_rr = std::fma(wr, ref_pix, r_rr);
            r_src = std::fma(w, src_pix, r_src);
            r_ss = std::fma(w * src_pix, src_pix, r_ss);
            r_rs = std::fma(wr, src_pix, r_rs);
            r_w = r_w + w;
            ++taps;
        }
        sum_ref += r_ref;
        sum_ref_ref += r_rr;
        sum_src += r_src;
        sum_src_src += r_ss;
        sum_ref_src += r_rs;
        wsum += r_w;
    }
    const_cast<Ctx &>(C).tally[tl_worker].taps[tl_family] += taps;
    const float inv = 1.0f / wsum;
    sum_ref *= inv;
    sum_ref_ref *= inv;
    sum_src *= inv;
    sum_src_src *= inv;
    sum_ref_src *= inv;
    const float var_ref = sum_ref_ref - sum_ref * sum_ref;
    const float var_src = sum_src_src - sum_src * sum_src;
    const float kMinVar = 1e-5f;
    if (var_ref < kMinVar || var_src < kMinVar) return 2.0f;
    const float covar = sum_ref_src - sum_ref * sum_src;
    const float var_rs = std::sqrt(var_ref * var_src);
    return std::fmax(0.0f, std::fmin(2.0f, 1.0f - covar / var_rs));
}

float ComputeBilateralNCCOld64(const Ctx &C, int px, int py, int src_idx, const F4 &plane);
float ComputeBilateralNCCNew64(const Ctx &C, int px, int py, int src_idx, const F4 &plane,
                               const std::vector<uint32_t> &selected);
float ComputeGeomConsistencyCost64(const Ctx &C, int px, int py, int src_idx, const F4 &plane);

// ComputeBilateralNCCOld, DPE.cu:696-782
float ComputeBilateralNCCOld(const Ctx &C, int px, int py, int src_idx, const F4 &plane) {
    if (C.cost64) return ComputeBilateralNCCOld64(C, px, py, src_idx, plane);
    const dpe_camera &src = C.cams[src_idx];
    float H[9];
    ComputeHomography(C.cams[0], src, plane, H);
    const_cast<Ctx &>(C).tally[tl_worker].hyps[tl_family]++;
    const F2 pt = ComputeCorrespondingPoint(H, px, py);
    if (pt.x >= (float)src.width || pt.x < 0.0f || pt.y >= (float)src.height || pt.y < 0.0f) return 2.0f;
    const float rc = texel(C.img[0], C.W, C.H, px, py);
    return BilateralNCCWindow(C, src_idx, H, px, py, rc, C.prm.strong_radius, C.prm.strong_increment);
}

inline S2 GetNeighbourPoint(const Ctx &C, int center, int index) {
    return C.neighbours[(size_t)C.nmap[center] * NNUM + index];
}

// ComputeBilateralNCCNew, DPE.cu:561-694 (WEAK centre pixels only)
float ComputeBilateralNCCNew(const Ctx &C, int px, int py, int src_idx, const F4 &plane,
                             const std::vector<uint32_t> &selected) {
    if (C.cost64) return ComputeBilateralNCCNew64(C, px, py, src_idx, plane, selected);
    const dpe_camera &src = C.cams[src_idx];
    const int center = px + py * C.W;
    const float cost_max = 2.0f;
    float H[9];
    ComputeHomography(C.cams[0], src, plane, H);
    const_cast<Ctx &>(C).tally[tl_worker].hyps[tl_family]++;
    const F2 pt = ComputeCorrespondingPoint(H, px, py);
    if (pt.x >= (float)src.width || pt.x < 0.0f || pt.y >= (float)src.height || pt.y < 0.0f) return cost_max;
    float cost = 0.0f;
    if (C.weak[center] == DPE_WEAK) {
        const float rc = texel(C.img[0], C.W, C.H, px, py);
        float center_cost = 0.0f, strong_cost = 0.0f;
        int strong_count = 0;
        for (int k = 0; k < NNUM; ++k) {
            const S2 np = GetNeighbourPoint(C, center, k);
            if (np.x == -1 || np.y == -1) continue;
            const F2 nsp = ComputeCorrespondingPoint(H, np.x, np.y);
            if (nsp.x < 0 || nsp.y < 0 || nsp.x >= (float)C.W || nsp.y >= (float)C.H) {
                if (k != 0) {
                    const uint32_t view_info = selected[np.x + np.y * C.W];
                    if (isSet(view_info, src_idx - 1)) {
                        strong_cost += cost_max;
                        strong_count++;
                    }
                    continue;
                } else {
                    return cost_max;
                }
            }
            int radius = (k == 0 ? C.prm.strong_radius : C.prm.weak_radius);
            int increment = (k == 0 ? C.prm.strong_increment : C.prm.weak_increment);
            if (C.prm.use_radius && k == 0) {
                radius = C.radius[center];
                increment = std::max(2, (int)(2.0 * radius / 5.0));
            }
            const float tc = BilateralNCCWindow(C, src_idx, H, np.x, np.y, rc, radius, increment);
            if (k == 0) {
                center_cost = tc;
            } else {
                strong_cost += tc;
                strong_count++;
            }
        }
        if (strong_count == 0) {
            cost = center_cost;
        } else {
            strong_cost /= strong_count;
            strong_cost = std::min(strong_cost, cost_max);
            cost = (float)(0.25 * center_cost + 0.75 * strong_cost);
        }
    }
    return cost;
}

// ComputeMultiViewCostVectorOld / New, DPE.cu:863-883
void CostVectorOld(const Ctx &C, int px, int py, const F4 &plane, float *cv) {
    for (int i = 1; i < C.N; ++i) cv[i - 1] = ComputeBilateralNCCOld(C, px, py, i, plane);
}
void CostVectorNew(const Ctx &C, int px, int py, const F4 &plane, float *cv, const std::vector<uint32_t> &sel) {
    for (int i = 1; i < C.N; ++i) cv[i - 1] = ComputeBilateralNCCNew(C, px, py, i, plane, sel);
}

// Get3DPointonWorld_cu, DPE.cu:885-905
F3 Get3DPointonWorld_cu(float x, float y, float depth, const dpe_camera &cam) {
    F3 X, T;
    X.x = depth * (x - cam.K[2]) / cam.K[0];
    X.y = depth * (y - cam.K[5]) / cam.K[4];
    X.z = depth;
    T.x = cam.R[0] * X.x + cam.R[3] * X.y + cam.R[6] * X.z;
    T.y = cam.R[1] * X.x + cam.R[4] * X.y + cam.R[7] * X.z;
    T.z = cam.R[2] * X.x + cam.R[5] * X.y + cam.R[8] * X.z;
    X.x = T.x + cam.c[0];
    X.y = T.y + cam.c[1];
    X.z = T.z + cam.c[2];
    return X;
}

// ProjectonCamera_cu, DPE.cu:907-917
void ProjectonCamera_cu(const F3 &X, const dpe_camera &cam, F2 &pt, float &depth) {
    F3 t;
    t.x = cam.R[0] * X.x + cam.R[1] * X.y + cam.R[2] * X.z + cam.t[0];
    t.y = cam.R[3] * X.x + cam.R[4] * X.y + cam.R[5] * X.z + cam.t[1];
    t.z = cam.R[6] * X.x + cam.R[7] * X.y + cam.R[8] * X.z + cam.t[2];
    depth = cam.K[6] * t.x + cam.K[7] * t.y + cam.K[8] * t.z;
    pt.x = (cam.K[0] * t.x + cam.K[1] * t.y + cam.K[2] * t.z) / depth;
    pt.y = (cam.K[3] * t.x + cam.K[4] * t.y + cam.K[5] * t.z) / depth;
}

// ComputeGeomConsistencyCost, DPE.cu:919-957 (depth lookup at truncated coords)
float ComputeGeomConsistencyCost(const Ctx &C, int px, int py, int src_idx, const F4 &plane) {
    if (C.cost64) return ComputeGeomConsistencyCost64(C, px, py, src_idx, plane);
    const dpe_camera &ref = C.cams[0];
    const dpe_camera &src = C.cams[src_idx];
    const float max_cost = 3.0f;
    const float depth = ComputeDepthfromPlaneHypothesis(ref, plane, px, py);
    const F3 fwd = Get3DPointonWorld_cu((float)px, (float)py, depth, ref);
    F2 sp;
    float sd;
    ProjectonCamera_cu(fwd, src, sp, sd);
    const int ix = f2i_clamp(sp.x, -1, C.W), iy = f2i_clamp(sp.y, -1, C.H);
    const float src_depth = texel(C.gdepth[src_idx], C.W, C.H, ix, iy);
    if (src_depth == 0.0f) return max_cost;
    const F3 s3 = Get3DPointonWorld_cu(sp.x, sp.y, src_depth, src);
    F2 bp;
    float rd;
    ProjectonCamera_cu(s3, ref, bp, rd);
    const float dc = (float)px - bp.x;
    const float dr = (float)py - bp.y;
    const float cc = std::sqrt(dc * dc + dr * dr);
    return std::fmin(max_cost, cc);
}

// ------------------------------------------------------------ FP64 cost path
// oracle_process_fp64 runs the same PatchMatch with the multi-view cost --
// homography, projection, bilinear sample, bilateral weights, window sums, NCC
// and the geometric-consistency cost -- evaluated in double from the fp32
// inputs (images, cameras and the fp32 plane hypotheses), with exact divisions
// and square roots; a cost is rounded to fp32 where the algorithm stores it.
// States, random streams and decision rules are the fp32 oracle's.  It is the
// reference the precision form's depth error is measured against.

// ComputeHomography (DPE.cu:457-517) in double.
void ComputeHomography64(const dpe_camera &ref, const dpe_camera &src, const F4 &plane, double *H) {
    double rR[9], sR[9], rt[3], st[3], rK[9], sK[9];
    for (int i = 0; i < 9; ++i) {
        rR[i] = ref.R[i];
        sR[i] = src.R[i];
        rK[i] = ref.K[i];
        sK[i] = src.K[i];
    }
    for (int i = 0; i < 3; ++i) {
        rt[i] = ref.t[i];
        st[i] = src.t[i];
    }
    double ref_C[3], src_C[3];
    for (int i = 0; i < 3; ++i) {
        ref_C[i] = -(rR[i] * rt[0] + rR[3 + i] * rt[1] + rR[6 + i] * rt[2]);
        src_C[i] = -(sR[i] * st[0] + sR[3 + i] * st[1] + sR[6 + i] * st[2]);
    }
    double Rr[9], tr[3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            Rr[3 * r + c] = sR[3 * r] * rR[3 * c] + sR[3 * r + 1] * rR[3 * c + 1] + sR[3 * r + 2] * rR[3 * c + 2];
    for (int r = 0; r < 3; ++r)
        tr[r] = sR[3 * r] * (ref_C[0] - src_C[0]) + sR[3 * r + 1] * (ref_C[1] - src_C[1]) +
                sR[3 * r + 2] * (ref_C[2] - src_C[2]);
    const double n[3] = {plane.x, plane.y, plane.z}, dist = plane.w;
    double M[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[3 * r + c] = Rr[3 * r + c] - tr[r] * n[c] / dist;
    // M K_ref^-1 (K_ref = [[fx 0 cx] [0 fy cy] [0 0 1]])
    double T[9];
    for (int r = 0; r < 3; ++r) {
        T[3 * r] = M[3 * r] / rK[0];
        T[3 * r + 1] = M[3 * r + 1] / rK[4];
        T[3 * r + 2] = -M[3 * r] * rK[2] / rK[0] - M[3 * r + 1] * rK[5] / rK[4] + M[3 * r + 2];
    }
    for (int c = 0; c < 3; ++c) {
        H[c] = sK[0] * T[c] + sK[2] * T[6 + c];
        H[3 + c] = sK[4] * T[3 + c] + sK[5] * T[6 + c];
        H[6 + c] = sK[8] * T[6 + c];
    }
}

inline void Project64(const double *H, double x, double y, double &u, double &v) {
    const double w = H[6] * x + H[7] * y + H[8];
    u = (H[0] * x + H[1] * y + H[2]) / w;
    v = (H[3] * x + H[4] * y + H[5]) / w;
}

inline int d2i_clamp(double f, int lo, int hi) {
    if (!(f >= (double)lo)) return lo;
    if (f > (double)hi) return hi;
    return (int)f;
}

// bilinear() in double: same texel addressing (clamp to edge), exact weights.
inline double bilinear64(const float *im, int W, int H, double u, double v) {
    const double fx = std::floor(u), fy = std::floor(v);
    const double a = u - fx, b = v - fy;
    const int ix = d2i_clamp(fx, -1, W - 1), iy = d2i_clamp(fy, -1, H - 1);
    const double t00 = texel(im, W, H, ix, iy), t10 = texel(im, W, H, ix + 1, iy);
    const double t01 = texel(im, W, H, ix, iy + 1), t11 = texel(im, W, H, ix + 1, iy + 1);
    const double top = t00 + a * (t10 - t00), bot = t01 + a * (t11 - t01);
    return top + b * (bot - top);
}

// BilateralNCCWindow in double.
double BilateralNCCWindow64(const Ctx &C, int src_idx, const double *H, int cx, int cy, double ref_center_pix,
                            int radius, int increment) {
    const dpe_camera &src = C.cams[src_idx];
    const float *refim = C.img[0];
    const float *srcim = C.img[src_idx];
    const double ss = C.prm.sigma_spatial, sc = C.prm.sigma_color;
    double s_r = 0, s_rr = 0, s_s = 0, s_ss = 0, s_rs = 0, s_w = 0;
    int taps = 0;
    for (int i = -radius; i <= radius; i += increment) {
        for (int j = -radius; j <= radius; j += increment) {
            const int rx = cx + i, ry = cy + j;
            const double r = texel(refim, C.W, C.H, rx, ry);
            double u, v;
            Project64(H, rx, ry, u, v);
            const double s = bilinear64(srcim, src.width, src.height, u, v);
            const double w = o_exp(-std::sqrt((double)(i * i + j * j)) / (2.0 * ss * ss) -
                                   std::fabs(r - ref_center_pix) / (2.0 * sc * sc));
            s_r += w * r;
            s_rr += w * r * r;
            s_s += w * s;
            s_ss += w * s * s;
            s_rs += w * r * s;
            s_w += w;
            ++taps;
        }
    }
    const_cast<Ctx &>(C).tally[tl_worker].taps[tl_family] += taps;
    s_r /= s_w;
    s_rr /= s_w;
    s_s /= s_w;
    s_ss /= s_w;
    s_rs /= s_w;
    const double var_ref = s_rr - s_r * s_r, var_src = s_ss - s_s * s_s;
    if (var_ref < 1e-5 || var_src < 1e-5) return 2.0;
    const double covar = s_rs - s_r * s_s;
    return std::fmax(0.0, std::fmin(2.0, 1.0 - covar / std::sqrt(var_ref * var_src)));
}

float ComputeBilateralNCCOld64(const Ctx &C, int px, int py, int src_idx, const F4 &plane) {
    const dpe_camera &src = C.cams[src_idx];
    double H[9];
    ComputeHomography64(C.cams[0], src, plane, H);
    const_cast<Ctx &>(C).tally[tl_worker].hyps[tl_family]++;
    double u, v;
    Project64(H, px, py, u, v);
    if (u >= src.width || u < 0.0 || v >= src.height || v < 0.0) return 2.0f;
    const double rc = texel(C.img[0], C.W, C.H, px, py);
    return (float)BilateralNCCWindow64(C, src_idx, H, px, py, rc, C.prm.strong_radius, C.prm.strong_increment);
}

float ComputeBilateralNCCNew64(const Ctx &C, int px, int py, int src_idx, const F4 &plane,
                               const std::vector<uint32_t> &selected) {
    const dpe_camera &src = C.cams[src_idx];
    const int center = px + py * C.W;
    const double cost_max = 2.0;
    double H[9];
    ComputeHomography64(C.cams[0], src, plane, H);
    const_cast<Ctx &>(C).tally[tl_worker].hyps[tl_family]++;
    double u, v;
    Project64(H, px, py, u, v);
    if (u >= src.width || u < 0.0 || v >= src.height || v < 0.0) return 2.0f;
    if (C.weak[center] != DPE_WEAK) return 0.0f;
    const double rc = texel(C.img[0], C.W, C.H, px, py);
    double center_cost = 0.0, strong_cost = 0.0;
    int strong_count = 0;
    for (int k = 0; k < NNUM; ++k) {
        const S2 np = GetNeighbourPoint(C, center, k);
        if (np.x == -1 || np.y == -1) continue;
        double nu, nv;
        Project64(H, np.x, np.y, nu, nv);
        if (nu < 0 || nv < 0 || nu >= C.W || nv >= C.H) {
            if (k == 0) return 2.0f;
            if (isSet(selected[np.x + np.y * C.W], src_idx - 1)) {
                strong_cost += cost_max;
                strong_count++;
            }
            continue;
        }
        int radius = (k == 0 ? C.prm.strong_radius : C.prm.weak_radius);
        int increment = (k == 0 ? C.prm.strong_increment : C.prm.weak_increment);
        if (C.prm.use_radius && k == 0) {
            radius = C.radius[center];
            increment = std::max(2, (int)(2.0 * radius / 5.0));
        }
        const double tc = BilateralNCCWindow64(C, src_idx, H, np.x, np.y, rc, radius, increment);
        if (k == 0) {
            center_cost = tc;
        } else {
            strong_cost += tc;
            strong_count++;
        }
    }
    if (strong_count == 0) return (float)center_cost;
    strong_cost = std::min(strong_cost / strong_count, cost_max);
    return (float)(0.25 * center_cost + 0.75 * strong_cost);
}

// ComputeGeomConsistencyCost (DPE.cu:919-957) in double; the source depth is
// looked up at the truncated projection as in the fp32 path.
float ComputeGeomConsistencyCost64(const Ctx &C, int px, int py, int src_idx, const F4 &plane) {
    const dpe_camera &ref = C.cams[0];
    const dpe_camera &src = C.cams[src_idx];
    auto world = [](double x, double y, double depth, const dpe_camera &cam, double *X) {
        const double cx = depth * (x - cam.K[2]) / cam.K[0], cy = depth * (y - cam.K[5]) / cam.K[4];
        for (int i = 0; i < 3; ++i)
            X[i] = (double)cam.R[i] * cx + (double)cam.R[3 + i] * cy + (double)cam.R[6 + i] * depth + cam.c[i];
    };
    auto project = [](const double *X, const dpe_camera &cam, double &u, double &v) {
        double t[3];
        for (int i = 0; i < 3; ++i)
            t[i] = (double)cam.R[3 * i] * X[0] + (double)cam.R[3 * i + 1] * X[1] + (double)cam.R[3 * i + 2] * X[2] +
                   cam.t[i];
        const double z = (double)cam.K[6] * t[0] + (double)cam.K[7] * t[1] + (double)cam.K[8] * t[2];
        u = ((double)cam.K[0] * t[0] + (double)cam.K[1] * t[1] + (double)cam.K[2] * t[2]) / z;
        v = ((double)cam.K[3] * t[0] + (double)cam.K[4] * t[1] + (double)cam.K[5] * t[2]) / z;
    };
    const double depth = -(double)plane.w * ref.K[0] /
                         ((px - (double)ref.K[2]) * plane.x + ((double)ref.K[0] / ref.K[4]) * (py - (double)ref.K[5]) *
                                                                  plane.y +
                          (double)ref.K[0] * plane.z);
    double X[3], u, v;
    world(px, py, depth, ref, X);
    project(X, src, u, v);
    const int ix = d2i_clamp(u, -1, C.W), iy = d2i_clamp(v, -1, C.H);
    const double src_depth = texel(C.gdepth[src_idx], C.W, C.H, ix, iy);
    if (src_depth == 0.0) return 3.0f;
    double S[3], bu, bv;
    world(u, v, src_depth, src, S);
    project(S, ref, bu, bv);
    const double dc = px - bu, dr = py - bv;
    return (float)std::fmin(3.0, std::sqrt(dc * dc + dr * dr));
}

// ComputeMultiViewInitialCostandSelectedViews, DPE.cu:784-830
float InitialCostAndSelectedViews(Ctx &C, int px, int py) {
    const int center = px + py * C.W;
    const F4 plane = C.planes[center];
    const float cost_max = 2.0f;
    float cv[32] = {2.0f};
    float cvc[32] = {2.0f};
    int cost_count = 0, num_valid = 0;
    for (int i = 1; i < C.N; ++i) {
        const float c = ComputeBilateralNCCOld(C, px, py, i, plane);
        cv[i - 1] = c;
        cvc[i - 1] = c;
        cost_count++;
        if (c < cost_max) num_valid++;
    }
    sort_small(cv, cost_count);
    C.selected[center] = 0;
    const int top_k = std::min(num_valid, C.prm.top_k);
    if (top_k > 0) {
        float cost = 0.0f;
        for (int i = 0; i < top_k; ++i) cost += cv[i];
        const float thr = cv[top_k - 1];
        for (int i = 0; i < C.N - 1; ++i)
            if (cvc[i] <= thr) setBit(&C.selected[center], i);
        return cost / top_k;
    }
    return cost_max;
}

// The view loop of ComputeMultiViewInitialCost, DPE.cu:832-861: `view_cost(i)`
// is the NCC cost of source view i (evaluated only for prior-selected views);
// an invalid view clears bits 0..i-1 of the mask (unSetBit, Q1).
template <class CostFn>
float InitialCostViews(int N, uint32_t *selected, CostFn view_cost) {
    const float cost_max = 2.0f;
    int cost_count = 0;
    float cost = 0.0f;
    for (int i = 1; i < N; ++i) {
        if (isSet(*selected, i - 1)) {
            const float c = view_cost(i);
            if (c < cost_max) {
                cost_count++;
                cost += c;
            } else {
                unSetBit(selected, i - 1);
            }
        }
    }
    return cost_count == 0 ? cost_max : cost / cost_count;
}

// ComputeMultiViewInitialCost, DPE.cu:832-861
float InitialCost(Ctx &C, int px, int py) {
    const int center = px + py * C.W;
    const F4 plane = C.planes[center];
    return InitialCostViews(C.N, &C.selected[center],
                            [&](int i) { return ComputeBilateralNCCOld(C, px, py, i, plane); });
}

uint32_t all_views(const Ctx &C) { return C.N - 1 >= 32 ? 0xFFFFFFFFu : (1u << (C.N - 1)) - 1u; }
bool same_plane(const F4 &a, const F4 &b) { return std::memcmp(&a, &b, sizeof(F4)) == 0; }

constexpr int kCandSlots = 17;
// NCC family of candidate slot `slot`'s row at `center` with plane pl (see
// Ctx::last_cand); records pl as the slot's last evaluated plane.
// A candidate whose plane equals, bit for bit, the pixel's current plane or the
// plane of a candidate taken before it in this half-iteration (one that did not
// itself repeat an earlier plane) gives the same costs as that row; the GPU
// reads that row instead of evaluating (and clears the slot's last-plane record,
// its own row then being stale), so it counts as dead work here too.
// tl_own holds this pixel's candidate planes that repeat no earlier one
// (StrongCandidates resets it).
thread_local int tl_own_n = 0;
thread_local F4 tl_own[16];
int candidate_family(const Ctx &C, int slot, int center, const F4 &pl) {
    F4 &l = C.last_cand[(size_t)slot * C.P + center];
    F4 none;
    std::memset(&none, 0xFF, sizeof(F4));
    const bool repeat = same_plane(l, pl) && !same_plane(l, none);
    bool alias = same_plane(pl, C.planes[center]);
    for (int k = 0; k < tl_own_n; ++k) alias = alias || same_plane(tl_own[k], pl);
    if (!alias && tl_own_n < 16) tl_own[tl_own_n++] = pl;
    if (repeat) return kFamDead;
    l = alias ? none : pl;
    return alias ? kFamDead : kFamFull;
}

// RandomInitialization, DPE.cu:1039-1067
void RandomInitialization(Ctx &C, int center) {
    tl_family = kFamFull;
    const int px = center % C.W, py = center / C.W;
    const dpe_camera &cam = C.cams[0];
    if (C.prm.state == DPE_FIRST_INIT) {
        Rng rng(C.prm.seed, center, PH_RANDOM_INIT, C.prm.call_index);
        C.planes[center] = GenerateRandomPlaneHypothesis(cam, px, py, rng, C.prm.depth_min, C.prm.depth_max);
        C.costs[center] = InitialCostAndSelectedViews(C, px, py);
        C.known[center] = all_views(C);
    } else {
        F4 plane = TransformNormal2RefCam(cam, C.planes[center]);
        const float depth = plane.w;
        plane.w = GetDistance2Origin(cam, px, py, depth, plane);
        C.planes[center] = plane;
        C.known[center] = C.selected[center] & all_views(C);  // the views InitialCost evaluates
        C.costs[center] = InitialCost(C, px, py);
    }
}

// View selection shared by both propagations (DPE.cu:1551-1631 / 1714-1783).
// Returns weight_norm; writes view weights and the selected bitmask.
float JointViewSelection(Ctx &C, int iter, const float cost_array[8][32], const float *priors, Rng &rng,
                         uint8_t *view_weights, uint32_t *temp_selected) {
    const int NV = C.N - 1;
    float sampling_probs[32] = {0.0f};
    const float cost_threshold = (float)(0.8 * (double)o_expf((float)(iter * iter) / (-90.0f)));
    for (int i = 0; i < NV; i++) {
        float count = 0;
        int count_false = 0;
        float tmpw = 0;
        for (int j = 0; j < 8; j++) {
            if (cost_array[j][i] < cost_threshold) {
                tmpw += o_expf(cost_array[j][i] * cost_array[j][i] / (-0.18f));
                count++;
            }
            if (cost_array[j][i] > 1.2f) count_false++;
        }
        if (count > 2 && count_false < 3) {
            sampling_probs[i] = tmpw / count;
        } else if (count_false < 3) {
            sampling_probs[i] = o_expf(cost_threshold * cost_threshold / (-0.32f));
        }
        sampling_probs[i] = sampling_probs[i] * priors[i];
    }
    TransformPDFToCDF(sampling_probs, NV);
    for (int sample = 0; sample < 15; ++sample) {
        const float rand_prob = rng.uniform() - FLT_EPSILON;
        for (int id = 0; id < NV; ++id) {
            if (sampling_probs[id] > rand_prob) {
                view_weights[id] += 1;
                break;
            }
        }
    }
    uint32_t tsel = 0;
    float weight_norm = 0;
    for (int i = 0; i < NV; ++i) {
        if (view_weights[i] > 0) {
            setBit(&tsel, i);
            weight_norm += view_weights[i];
        }
    }
    *temp_selected = tsel;
    return weight_norm;
}

// PlaneHypothesisRefinementStrong, DPE.cu:1069-1122
void RefinementStrong(Ctx &C, F4 *plane, float *depth, float *cost, Rng &rng, const uint8_t *vw, float wnorm,
                      int px, int py) {
    const float depth_perturbation = 0.02f;
    const float normal_perturbation = 0.02f;
    const dpe_camera &cam = C.cams[0];
    const float dmin = C.prm.depth_min, dmax = C.prm.depth_max;
    const float depth_rand = rng.uniform() * (dmax - dmin) + dmin;
    const F4 plane_rand = GenerateRandomNormal(cam, px, py, rng, *depth);
    float depth_perturbed = *depth;
    const float dminp = (1 - depth_perturbation) * depth_perturbed;
    const float dmaxp = (1 + depth_perturbation) * depth_perturbed;
    depth_perturbed = rng.uniform() * (dmaxp - dminp) + dminp;  // the do/while runs once (DPE.cu:1092-1094)
    const F4 plane_pert = GeneratePerturbedNormal(cam, px, py, *plane, rng, (float)(normal_perturbation * M_PI));
    const float depths[5] = {depth_rand, *depth, depth_rand, *depth, depth_perturbed};
    const F4 normals[5] = {*plane, plane_rand, plane_rand, plane_pert, *plane};
    for (int i = 0; i < 5; ++i) {
        float cv[32] = {2.0f};
        F4 tp = normals[i];
        tp.w = GetDistance2Origin(cam, px, py, depths[i], tp);
        for (int j = 0; j < C.N - 1; ++j) {  // CostVectorOld; views without weight are never read
            tl_family = vw[j] > 0 ? kFamRefine : kFamDead;
            cv[j] = ComputeBilateralNCCOld(C, px, py, j + 1, tp);
        }
        float tc = 0.0f;
        for (int j = 0; j < C.N - 1; ++j)
            if (vw[j] > 0) tc += vw[j] * cv[j];
        tc /= wnorm;
        const float db = ComputeDepthfromPlaneHypothesis(cam, tp, px, py);
        if (db >= dmin && db <= dmax && tc < *cost) {
            *depth = db;
            *plane = tp;
            *cost = tc;
        }
    }
}

// PlaneHypothesisRefinementWeak, DPE.cu:1124-1216
void RefinementWeak(Ctx &C, F4 *plane, float *depth, float *cost, Rng &rng, const uint8_t *vw, float wnorm,
                    int px, int py) {
    const float depth_perturbation = 0.02f;
    const float normal_perturbation = 0.02f;
    const dpe_camera &cam = C.cams[0];
    const float dmin = C.prm.depth_min, dmax = C.prm.depth_max;
    const int center = px + py * C.W;
    const int NV = C.N - 1;
    auto weighted = [&](const F4 &tp) {
        float cv[32] = {2.0f};
        for (int j = 0; j < NV; ++j) {  // CostVectorNew; views without weight are never read
            tl_family = vw[j] > 0 ? kFamWeak : kFamDead;
            cv[j] = ComputeBilateralNCCNew(C, px, py, j + 1, tp, C.selected);
        }
        tl_family = kFamWeak;
        float tc = 0.0f;
        for (int j = 0; j < NV; ++j) {
            if (vw[j] > 0) {
                if (C.prm.geom_consistency)
                    tc += vw[j] * (cv[j] + C.prm.geom_factor * ComputeGeomConsistencyCost(C, px, py, j + 1, tp));
                else
                    tc += vw[j] * cv[j];
            }
        }
        return tc / wnorm;
    };
    if (C.weak[center] == DPE_WEAK) {
        const F4 fit = C.fit_planes[center];
        if (fit.x == 0 && fit.y == 0 && fit.z == 0) return;
        const float tc = weighted(fit);
        const float db = ComputeDepthfromPlaneHypothesis(cam, fit, px, py);
        if (db >= dmin && db <= dmax && tc < *cost) {
            *depth = db;
            *plane = fit;
            *cost = tc;
        }
    }
    const float depth_rand = rng.uniform() * (dmax - dmin) + dmin;
    const F4 plane_rand = GenerateRandomNormal(cam, px, py, rng, *depth);
    float depth_perturbed = *depth;
    const float dminp = (1 - depth_perturbation) * depth_perturbed;
    const float dmaxp = (1 + depth_perturbation) * depth_perturbed;
    depth_perturbed = rng.uniform() * (dmaxp - dminp) + dminp;
    const F4 plane_pert = GeneratePerturbedNormal(cam, px, py, *plane, rng, (float)(normal_perturbation * M_PI));
    const float depths[5] = {depth_rand, *depth, depth_rand, *depth, depth_perturbed};
    const F4 normals[5] = {*plane, plane_rand, plane_rand, plane_pert, *plane};
    for (int i = 0; i < 5; ++i) {
        F4 tp = normals[i];
        tp.w = GetDistance2Origin(cam, px, py, depths[i], tp);
        const float tc = weighted(tp);
        const float db = ComputeDepthfromPlaneHypothesis(cam, tp, px, py);
        if (db >= dmin && db <= dmax && tc < *cost) {
            *depth = db;
            *plane = tp;
            *cost = tc;
        }
    }
}

// Candidate search of CheckerboardPropagationStrong, DPE.cu:1239-1549.  Reads
// the snapshot buffers.  flag/positions as in the reference.
void StrongCandidates(const Ctx &C, int px, int py, int iter, bool flag[8], int positions[8],
                      float cost_array[8][32]) {
    tl_own_n = 0;
    const int W = C.W, H = C.H;
    const int center = py * W + px;
    const std::vector<float> &costs = C.snap_costs;
    const std::vector<F4> &planes = C.snap_planes;
    if (C.prm.use_edge) {
        static const int dir[8][2] = {{0, -1}, {0, 1}, {-1, 0}, {1, 0}, {-1, -1}, {1, 1}, {-1, 1}, {1, -1}};
        const S2 *edge_neigh = &C.edge_neigh[(size_t)center * 8];
        const uint8_t *edge = C.edge;
        const float max_edge_dist = std::max(H, W) / 30.0f;
        const int min_step_len = 2;
        for (int d = 0; d < 8; ++d) {
            const int dx = dir[d][0], dy = dir[d][1];
            const int sx = std::max(1, 5 - 2 * iter) * dx;
            const int sy = std::max(1, 5 - 2 * iter) * dy;
            const S2 ep = edge_neigh[d];
            float dist = (float)std::sqrt((double)(ep.x - px) * (double)(ep.x - px) +
                                          (double)(ep.y - py) * (double)(ep.y - py));
            if (d >= 4) dist = (float)(dist / std::sqrt(2.0));
            if (edge[center]) {
                dist = 11 * min_step_len;
            } else if (ep.x == -1 || ep.y == -1 || dist > max_edge_dist) {
                dist = max_edge_dist;
                if (d >= 4) dist = (float)(dist / std::sqrt(2.0));
            }
            const int step_num = std::min(std::max(11, (int)(1.0f * dist / min_step_len)), 22);
            int step_len = std::max((int)(1.0f * dist / step_num), min_step_len);
            if (d < 4 && step_len % 2 == 1) step_len -= 1;
            int mcx = 0, mcy = 0;
            float min_cost = FLT_MAX;
            for (int step = 0; step < step_num; ++step) {
                int fx = 0, fy = 0;
                if (d > 4) {
                    if (d % 2) fx = dx; else fy = dy;
                }
                const int tx = px + sx + step * step_len * dx + fx;
                const int ty = py + sy + step * step_len * dy + fy;
                if (!(tx >= 0 && ty >= 0 && tx < W && ty < H)) continue;
                const int tcen = tx + ty * W;
                if (min_cost > costs[tcen]) {
                    mcx = tx; mcy = ty;
                    min_cost = costs[tcen];
                }
            }
            if (min_cost < FLT_MAX) {
                flag[d] = true;
                positions[d] = mcx + mcy * W;
                tl_family = candidate_family(C, 1 + d, center, planes[positions[d]]);
                CostVectorOld(C, px, py, planes[positions[d]], cost_array[d]);
                tl_family = kFamFull;
            }
        }
        if (!edge[center]) {
            const float good_threshold = 0.8f * o_expf((float)(iter * iter) / (-90.0f));
            const float bad_threshold = 1.2f;
            for (int d = 0; d < 8; ++d) {
                const int dx = dir[d][0], dy = dir[d][1];
                const int sx = std::max(1, 5 - 2 * iter) * dx;
                const int sy = std::max(1, 5 - 2 * iter) * dy;
                const bool had = flag[d];
                float tca[32] = {2.0f};
                int mcx = 0, mcy = 0;
                float min_cost = FLT_MAX;
                for (int step = 0; step < 11; ++step) {
                    int fx = 0, fy = 0;
                    if (d > 4) {
                        if (d % 2) fx = dx; else fy = dy;
                    }
                    const int tx = px + sx + step * min_step_len * dx + fx;
                    const int ty = py + sy + step * min_step_len * dy + fy;
                    if (!(tx >= 0 && ty >= 0 && tx < W && ty < H)) continue;
                    const int tcen = tx + ty * W;
                    if (min_cost > costs[tcen]) {
                        mcx = tx; mcy = ty;
                        min_cost = costs[tcen];
                    }
                }
                if (min_cost < FLT_MAX) {
                    flag[d] = true;
                    const int tpos = mcx + mcy * W;
                    // the pixel the adaptive scan already chose: equal costs, which
                    // never replace it below (a result the reference never uses)
                    tl_family = (had && tpos == positions[d]) ? kFamDead : candidate_family(C, 9 + d, center, planes[tpos]);
                    CostVectorOld(C, px, py, planes[tpos], tca);
                    tl_family = kFamFull;
                    int good[2] = {0, 0}, bad[2] = {0, 0};
                    for (int i = 0; i < 2; i++) {
                        for (int j = 0; j < C.N - 1; j++) {
                            const float val = (i == 0 ? cost_array[d][j] : tca[j]);
                            if (val < good_threshold) good[i]++;
                            if (val > bad_threshold) bad[i]++;
                        }
                    }
                    if (!had || good[1] > good[0] || (good[1] == good[0] && bad[1] < bad[0])) {
                        positions[d] = tpos;
                        for (int j = 0; j < C.N - 1; j++) cost_array[d][j] = tca[j];
                    }
                }
            }
        }
    } else {
        float costMin;
        int costMinPoint;
        int left_near = center - 1, left_far = center - 3, right_near = center + 1, right_far = center + 3;
        int up_near = center - W, up_far = center - 3 * W, down_near = center + W, down_far = center + 3 * W;
        if (py > 2) {
            flag[1] = true;
            costMin = costs[up_far];
            costMinPoint = up_far;
            for (int i = 1; i < 11; ++i) {
                if (py > 2 + 2 * i) {
                    const int pt = up_far - 2 * i * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            up_far = costMinPoint;
            tl_family = candidate_family(C, 2, center, planes[up_far]);
            CostVectorOld(C, px, py, planes[up_far], cost_array[1]);
            tl_family = kFamFull;
        }
        if (py < H - 3) {
            flag[3] = true;
            costMin = costs[down_far];
            costMinPoint = down_far;
            for (int i = 1; i < 11; ++i) {
                if (py < H - 3 - 2 * i) {
                    const int pt = down_far + 2 * i * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            down_far = costMinPoint;
            tl_family = candidate_family(C, 4, center, planes[down_far]);
            CostVectorOld(C, px, py, planes[down_far], cost_array[3]);
            tl_family = kFamFull;
        }
        if (px > 2) {
            flag[5] = true;
            costMin = costs[left_far];
            costMinPoint = left_far;
            for (int i = 1; i < 11; ++i) {
                if (px > 2 + 2 * i) {
                    const int pt = left_far - 2 * i;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            left_far = costMinPoint;
            tl_family = candidate_family(C, 6, center, planes[left_far]);
            CostVectorOld(C, px, py, planes[left_far], cost_array[5]);
            tl_family = kFamFull;
        }
        if (px < W - 3) {
            flag[7] = true;
            costMin = costs[right_far];
            costMinPoint = right_far;
            for (int i = 1; i < 11; ++i) {
                if (px < W - 3 - 2 * i) {
                    const int pt = right_far + 2 * i;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            right_far = costMinPoint;
            tl_family = candidate_family(C, 8, center, planes[right_far]);
            CostVectorOld(C, px, py, planes[right_far], cost_array[7]);
            tl_family = kFamFull;
        }
        if (py > 0) {
            flag[0] = true;
            costMin = costs[up_near];
            costMinPoint = up_near;
            for (int i = 0; i < 3; ++i) {
                if (py > 1 + i && px > i) {
                    const int pt = up_near - (1 + i) * W - (1 + i);
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
                if (py > 1 + i && px < W - 1 - i) {
                    const int pt = up_near - (1 + i) * W + (1 + i);
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            up_near = costMinPoint;
            tl_family = candidate_family(C, 1, center, planes[up_near]);
            CostVectorOld(C, px, py, planes[up_near], cost_array[0]);
            tl_family = kFamFull;
        }
        if (py < H - 1) {
            flag[2] = true;
            costMin = costs[down_near];
            costMinPoint = down_near;
            for (int i = 0; i < 3; ++i) {
                if (py < H - 2 - i && px > i) {
                    const int pt = down_near + (1 + i) * W - (1 + i);
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
                if (py < H - 2 - i && px < W - 1 - i) {
                    const int pt = down_near + (1 + i) * W + (1 + i);
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            down_near = costMinPoint;
            tl_family = candidate_family(C, 3, center, planes[down_near]);
            CostVectorOld(C, px, py, planes[down_near], cost_array[2]);
            tl_family = kFamFull;
        }
        if (px > 0) {
            flag[4] = true;
            costMin = costs[left_near];
            costMinPoint = left_near;
            for (int i = 0; i < 3; ++i) {
                if (px > 1 + i && py > i) {
                    const int pt = left_near - (1 + i) - (1 + i) * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
                if (px > 1 + i && py < H - 1 - i) {
                    const int pt = left_near - (1 + i) + (1 + i) * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            left_near = costMinPoint;
            tl_family = candidate_family(C, 5, center, planes[left_near]);
            CostVectorOld(C, px, py, planes[left_near], cost_array[4]);
            tl_family = kFamFull;
        }
        if (px < W - 1) {
            flag[6] = true;
            costMin = costs[right_near];
            costMinPoint = right_near;
            for (int i = 0; i < 3; ++i) {
                if (px < W - 2 - i && py > i) {
                    const int pt = right_near + (1 + i) - (1 + i) * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
                if (px < W - 2 - i && py < H - 1 - i) {
                    const int pt = right_near + (1 + i) + (1 + i) * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            right_near = costMinPoint;
            tl_family = candidate_family(C, 7, center, planes[right_near]);
            CostVectorOld(C, px, py, planes[right_near], cost_array[6]);
            tl_family = kFamFull;
        }
        positions[0] = up_near;
        positions[1] = up_far;
        positions[2] = down_near;
        positions[3] = down_far;
        positions[4] = left_near;
        positions[5] = left_far;
        positions[6] = right_near;
        positions[7] = right_far;
    }
}

inline uint32_t snap_selected_at(const Ctx &C, long idx) {
    if (idx < 0 || idx >= (long)C.P) return 0u;
    return C.snap_selected[idx];
}

// CheckerboardPropagationStrong, DPE.cu:1218-1670
void CheckerboardPropagationStrong(Ctx &C, int px, int py, int iter) {
    tl_family = kFamFull;
    const int W = C.W;
    const int center = py * W + px;
    const int NV = C.N - 1;
    float cost_array[8][32] = {{2.0f}};
    bool flag[8] = {false};
    int positions[8] = {0};
    StrongCandidates(C, px, py, iter, flag, positions, cost_array);

    uint8_t *vw = &C.view_weight[(size_t)center * MAXIMG];
    for (int i = 0; i < MAXIMG; ++i) vw[i] = 0;
    float priors[32] = {0.0f};
    const long nbr[4] = {(long)center - W, (long)center + W, (long)center - 1, (long)center + 1};
    for (int i = 0; i < 4; ++i) {
        if (flag[2 * i]) {
            const uint32_t sv = snap_selected_at(C, nbr[i]);
            for (int j = 0; j < NV; ++j) priors[j] += isSet(sv, j) == 1 ? 0.9f : 0.1f;
        }
    }
    Rng rng(C.prm.seed, center, ph_strong(iter), C.prm.call_index);
    uint32_t tsel = 0;
    const float wnorm = JointViewSelection(C, iter, cost_array, priors, rng, vw, &tsel);

    float final_costs[8] = {0.0f};
    for (int i = 0; i < 8; ++i) {
        for (int j = 0; j < NV; ++j)
            if (vw[j] > 0) final_costs[i] += vw[j] * cost_array[i][j];
        final_costs[i] /= wnorm;
    }
    const int mi = FindMinCostIndex(final_costs, 8);
    float cvn[32] = {2.0f};
    const F4 cur = C.planes[center];
    for (int i = 1; i < C.N; ++i) {  // CostVectorOld
        tl_family = isSet(C.known[center], i - 1) ? kFamDead : kFamFull;
        cvn[i - 1] = ComputeBilateralNCCOld(C, px, py, i, cur);
    }
    tl_family = kFamFull;
    float cost_now = 0.0f;
    for (int i = 0; i < NV; ++i) cost_now += vw[i] * cvn[i];
    cost_now /= wnorm;
    C.costs[center] = cost_now;
    float depth_now = ComputeDepthfromPlaneHypothesis(C.cams[0], cur, px, py);
    F4 plane_now = cur;
    if (flag[mi]) {
        const F4 cand = C.snap_planes[positions[mi]];
        const float db = ComputeDepthfromPlaneHypothesis(C.cams[0], cand, px, py);
        if (db >= C.prm.depth_min && db <= C.prm.depth_max && final_costs[mi] < cost_now) {
            depth_now = db;
            plane_now = cand;
            cost_now = final_costs[mi];
            C.selected[center] = tsel;
        }
    }
    const F4 adopted = plane_now;  // every view of it was evaluated (cvn or cost_array[mi])
    RefinementStrong(C, &plane_now, &depth_now, &cost_now, rng, vw, wnorm, px, py);
    if (C.prm.state == DPE_REFINE_INIT) {
        if ((double)cost_now < (double)C.costs[center] - 0.1) {
            C.costs[center] = cost_now;
            C.planes[center] = plane_now;
        }
    } else {
        C.costs[center] = cost_now;
        C.planes[center] = plane_now;
    }
    const F4 fin = C.planes[center];
    C.known[center] = (same_plane(fin, adopted) || same_plane(fin, cur)) ? all_views(C) : 0u;
}

// CheckerboardPropagationWeak, DPE.cu:1672-1866
void CheckerboardPropagationWeak(Ctx &C, int px, int py, int iter) {
    tl_family = kFamWeak;
    const int W = C.W;
    const int center = py * W + px;
    const int NV = C.N - 1;
    float cost_array[8][32] = {{2.0f}};
    bool flag[8] = {false};
    int positions[8] = {0};
    F4 new_plane[8];
    for (int i = 0; i < 8; ++i) {
        const S2 np = GetNeighbourPoint(C, center, i + 1);
        if (np.x == -1 || np.y == -1 || C.weak[np.x + np.y * W] != DPE_STRONG) {
            flag[i] = false;
            continue;
        }
        positions[i] = np.x + np.y * W;
        flag[i] = true;
        // a candidate at the same pixel as an earlier one has the same plane and
        // costs: computed as the reference does, counted apart as dead
        bool repeated = false;
        for (int k = 0; k < i; ++k) repeated = repeated || (flag[k] && positions[k] == positions[i]);
        tl_family = repeated ? kFamDead : kFamWeak;
        CostVectorNew(C, px, py, C.planes[positions[i]], cost_array[i], C.selected);
        tl_family = kFamWeak;
        new_plane[i] = C.planes[positions[i]];
    }
    uint8_t *vw = &C.view_weight[(size_t)center * MAXIMG];
    for (int i = 0; i < MAXIMG; ++i) vw[i] = 0;
    float priors[32] = {0.0f};
    for (int i = 0; i < 8; ++i) {
        const S2 np = GetNeighbourPoint(C, center, i + 1);
        if (np.x == -1 || np.y == -1) continue;
        const uint32_t sv = C.selected[np.x + np.y * W];
        for (int j = 0; j < NV; ++j) priors[j] += isSet(sv, j) == 1 ? 0.9f : 0.1f;
    }
    Rng rng(C.prm.seed, center, ph_weak(iter), C.prm.call_index);
    uint32_t tsel = 0;
    const float wnorm = JointViewSelection(C, iter, cost_array, priors, rng, vw, &tsel);
    const float gf = C.prm.geom_factor;
    float final_costs[8] = {0.0f};
    for (int i = 0; i < 8; ++i) {
        for (int j = 0; j < NV; ++j) {
            if (vw[j] > 0) {
                if (C.prm.geom_consistency) {
                    if (flag[i])
                        final_costs[i] += vw[j] * (cost_array[i][j] +
                                                   gf * ComputeGeomConsistencyCost(C, px, py, j + 1,
                                                                                   C.planes[positions[i]]));
                    else
                        final_costs[i] += vw[j] * (cost_array[i][j] + gf * 3.0f);
                } else {
                    final_costs[i] += vw[j] * cost_array[i][j];
                }
            }
        }
        final_costs[i] /= wnorm;
    }
    const int mi = FindMinCostIndex(final_costs, 8);
    float cvn[32] = {2.0f};
    const F4 cur = C.planes[center];
    // views of weight 0 add exactly 0 to the two weighted sums below (finite
    // costs): computed as the reference does, counted apart as dead
    for (int i = 1; i < C.N; ++i) {
        tl_family = vw[i - 1] > 0 ? kFamWeak : kFamDead;
        cvn[i - 1] = ComputeBilateralNCCNew(C, px, py, i, cur, C.selected);
    }
    tl_family = kFamWeak;
    float cost_now = 0.0f;
    for (int i = 0; i < NV; ++i) {
        if (C.prm.geom_consistency)
            cost_now += vw[i] * (cvn[i] + gf * ComputeGeomConsistencyCost(C, px, py, i + 1, cur));
        else
            cost_now += vw[i] * cvn[i];
    }
    cost_now /= wnorm;
    C.costs[center] = cost_now;
    float depth_now = ComputeDepthfromPlaneHypothesis(C.cams[0], cur, px, py);
    F4 plane_now = cur;
    if (flag[mi]) {
        const float db = ComputeDepthfromPlaneHypothesis(C.cams[0], new_plane[mi], px, py);
        if (db >= C.prm.depth_min && db <= C.prm.depth_max && final_costs[mi] < cost_now) {
            depth_now = db;
            plane_now = new_plane[mi];
            cost_now = final_costs[mi];
            C.selected[center] = tsel;
        }
    }
    RefinementWeak(C, &plane_now, &depth_now, &cost_now, rng, vw, wnorm, px, py);
    if (C.prm.state == DPE_REFINE_INIT) {
        if ((double)cost_now < (double)C.costs[center] - 0.1) {
            C.costs[center] = cost_now;
            C.planes[center] = plane_now;
        }
    } else {
        C.costs[center] = cost_now;
        C.planes[center] = plane_now;
    }
    // final cost with the Old NCC (DPE.cu:1849-1865); the radius swap there has no effect
    float cvo[32] = {2.0f};
    for (int i = 1; i < C.N; ++i) {
        tl_family = vw[i - 1] > 0 ? kFamWeak : kFamDead;
        cvo[i - 1] = ComputeBilateralNCCOld(C, px, py, i, C.planes[center]);
    }
    tl_family = kFamWeak;
    float c2 = 0.0f;
    for (int i = 0; i < NV; ++i) c2 += vw[i] * cvo[i];
    c2 /= wnorm;
    C.costs[center] = c2;
}

// GetDepthandNormal, DPE.cu:1944-1959
void GetDepthandNormal(Ctx &C, int center) {
    const int px = center % C.W, py = center / C.W;
    F4 p = C.planes[center];
    p.w = ComputeDepthfromPlaneHypothesis(C.cams[0], p, px, py);
    C.planes[center] = TransformNormal(C.cams[0], p);
}

// CheckerboardFilterStrong, DPE.cu:1961-2071
void CheckerboardFilterStrong(Ctx &C, int px, int py) {
    const int W = C.W, H = C.H;
    const int center = py * W + px;
    std::vector<F4> &pl = C.planes;
    const std::vector<uint8_t> &wk = C.weak;
    float filter[21];
    int index = 0;
    filter[index++] = pl[center].w;
    const int left = center - 1, leftleft = center - 3;
    const int up = center - W, upup = center - 3 * W;
    const int down = center + W, downdown = center + 3 * W;
    const int right = center + 1, rightright = center + 3;
    if (C.costs[center] < 0.001f) return;
    if (py > 0 && wk[up] == DPE_STRONG) filter[index++] = pl[up].w;
    if (py > 2 && wk[upup] == DPE_STRONG) filter[index++] = pl[upup].w;
    if (py > 4 && wk[upup - W * 2] == DPE_STRONG) filter[index++] = pl[upup - W * 2].w;
    if (py < H - 1 && wk[down] == DPE_STRONG) filter[index++] = pl[down].w;
    if (py < H - 3 && wk[downdown] == DPE_STRONG) filter[index++] = pl[downdown].w;
    if (py < H - 5 && wk[downdown + W * 2] == DPE_STRONG) filter[index++] = pl[downdown + W * 2].w;
    if (px > 0 && wk[left] == DPE_STRONG) filter[index++] = pl[left].w;
    if (px > 2 && wk[leftleft] == DPE_STRONG) filter[index++] = pl[leftleft].w;
    if (px > 4 && wk[leftleft - 2] == DPE_STRONG) filter[index++] = pl[leftleft - 2].w;
    if (px < W - 1 && wk[right] == DPE_STRONG) filter[index++] = pl[right].w;
    if (px < W - 3 && wk[rightright] == DPE_STRONG) filter[index++] = pl[rightright].w;
    if (px < W - 5 && wk[rightright + 2] == DPE_STRONG) filter[index++] = pl[rightright + 2].w;
    if (py > 0 && px < W - 2 && wk[up + 2] == DPE_STRONG) filter[index++] = pl[up + 2].w;
    if (py < H - 1 && px < W - 2 && wk[down + 2] == DPE_STRONG) filter[index++] = pl[down + 2].w;
    if (py > 0 && px > 1 && wk[up - 2] == DPE_STRONG) filter[index++] = pl[up - 2].w;
    if (py < H - 1 && px > 1 && wk[down - 2] == DPE_STRONG) filter[index++] = pl[down - 2].w;
    if (px > 0 && py > 2 && wk[left - W * 2] == DPE_STRONG) filter[index++] = pl[left - W * 2].w;
    if (px < W - 1 && py > 2 && wk[right - W * 2] == DPE_STRONG) filter[index++] = pl[right - W * 2].w;
    if (px > 0 && py < H - 2 && wk[left + W * 2] == DPE_STRONG) filter[index++] = pl[left + W * 2].w;
    if (px < W - 1 && py < H - 2 && wk[right + W * 2] == DPE_STRONG) filter[index++] = pl[right + W * 2].w;
    sort_small(filter, index);
    const int m = index / 2;
    if (index % 2 == 0)
        pl[center].w = (filter[m - 1] + filter[m]) / 2;
    else
        pl[center].w = filter[m];
}

// PointinTriangle, DPE.cu:139-160
bool PointinTriangle(S2 A, S2 B, S2 Cc, int px, int py) {
    const F2 AB{(float)(B.x - A.x), (float)(B.y - A.y)};
    const F2 BC{(float)(Cc.x - B.x), (float)(Cc.y - B.y)};
    const F2 CA{(float)(A.x - Cc.x), (float)(A.y - Cc.y)};
    const float ab = std::sqrt(AB.x * AB.x + AB.y * AB.y);
    const float bc = std::sqrt(BC.x * BC.x + BC.y * BC.y);
    const float ca = std::sqrt(CA.x * CA.x + CA.y * CA.y);
    if (ab <= 2 || bc <= 2 || ca <= 2) return false;
    if (!(ab + bc > ca && bc + ca > ab && ab + ca > bc)) return false;
    const F2 PA{(float)(A.x - px), (float)(A.y - py)};
    const F2 PB{(float)(B.x - px), (float)(B.y - py)};
    const F2 PC{(float)(Cc.x - px), (float)(Cc.y - py)};
    const float t1 = PA.x * PB.y - PA.y * PB.x;
    const float t2 = PB.x * PC.y - PB.y * PC.x;
    const float t3 = PC.x * PA.y - PC.y * PA.x;
    return t1 * t2 >= 0 && t1 * t3 >= 0;
}

// test access only: when set, BresenhamLine appends the low-resolution (x, y) of
// every pixel its walks test, and the step count of each walk
thread_local std::vector<long> *tl_bres_trace = nullptr;
thread_local int tl_bres_steps[2];

inline uint8_t edge_low_at(const Ctx &C, long idx) {
    if (idx < 0 || idx >= (long)C.low_w * C.low_h) return 0;
    return C.edge_low[idx];
}

// BresenhamLine, DPE.cu:162-248
bool BresenhamLine(const Ctx &C, int Ax, int Ay, int Bx, int By) {
    if (C.edge[Ax + Ay * C.W] || C.edge[Bx + By * C.W]) return false;
    const float scale_x = 1.0f * C.low_w / (float)C.W;
    const float scale_y = 1.0f * C.low_h / (float)C.H;
    const int height = C.low_h, width = C.low_w;
    const int max_step = C.prm.high_res_img ? (int)std::round(std::max(height, width) / 60.0) : std::max(height, width);
    for (int pass = 0; pass < 2; ++pass) {
        if (tl_bres_trace) tl_bres_steps[pass] = 0;
        const int fromx = pass == 0 ? Bx : Ax, fromy = pass == 0 ? By : Ay;
        const int tox = pass == 0 ? Ax : Bx, toy = pass == 0 ? Ay : By;
        int x0 = std::min((int)std::round(fromx * scale_x), width - 1);
        int y0 = std::min((int)std::round(fromy * scale_y), height - 1);
        const int x1 = std::min((int)std::round(tox * scale_x), width - 1);
        const int y1 = std::min((int)std::round(toy * scale_y), height - 1);
        const int dx = std::abs(x1 - x0), sx = x0 < x1 ? 1 : -1;
        const int dy = std::abs(y1 - y0), sy = y0 < y1 ? 1 : -1;
        int erro = (dx > dy ? dx : dy) / 2;
        int step = 0;
        bool tagx = true, tagy = true;
        while (tagx || tagy) {
            if (x0 == x1) tagx = false;
            if (y0 == y1) tagy = false;
            const int e2 = erro;
            if (e2 > -dx) {
                erro -= dy;
                x0 += sx;
            }
            if (e2 < dy) {
                erro += dx;
                y0 += sy;
            }
            if (tl_bres_trace) {
                tl_bres_trace->push_back(x0);
                tl_bres_trace->push_back(y0);
                tl_bres_steps[pass]++;
            }
            if (edge_low_at(C, (long)x0 + (long)y0 * width)) return true;
            step += 1;
            if (step >= max_step) break;
        }
    }
    return false;
}

// GenEdgeInform, DPE.cu:2487-2595
void GenEdgeInform(Ctx &C, int center) {
    const int W = C.W, H = C.H;
    const int px = center % W, py = center / W;
    static const int dir[8][2] = {{0, -1}, {0, 1}, {-1, 0}, {1, 0}, {-1, -1}, {1, 1}, {-1, 1}, {1, -1}};
    if (C.prm.use_edge) {
        S2 *en = &C.edge_neigh[(size_t)center * 8];
        for (int i = 0; i < 8; i++) {
            en[i].x = -1;
            en[i].y = -1;
            const int dx = dir[i][0], dy = dir[i][1];
            int nx = px + dx, ny = py + dy;
            while (true) {
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) break;
                if (C.edge[nx + ny * W]) {
                    en[i].x = (int16_t)nx;
                    en[i].y = (int16_t)ny;
                    break;
                }
                nx += dx;
                ny += dy;
            }
        }
        const int radius = C.prm.strong_radius;
        int edge_pix = 0, tot_pix = 0;
        for (int i = -radius; i <= radius; i++) {
            for (int j = -radius; j <= radius; j++) {
                const int nx = px + i, ny = py + j;
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                if (C.edge[ny * W + nx]) edge_pix++;
                tot_pix++;
            }
        }
        float density = 1.0f * edge_pix / tot_pix;
        if (C.prm.use_label) {
            int bound_pix = 0;
            for (int i = -radius; i <= radius; i++) {
                for (int j = -radius; j <= radius; j++) {
                    const int nx = px + i, ny = py + j;
                    if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                    if (C.label[ny * W + nx] == 0) bound_pix++;
                }
            }
            density = std::max(density, (float)(bound_pix / tot_pix));
        }
        C.complex_map[center] = (float)(1.0f / (1.0f + o_exp(-25.0 * ((double)density - 0.35))));
    }
    if (C.prm.use_label && C.weak[center] == DPE_WEAK) {
        S2 *lb = &C.label_bound[(size_t)C.nmap[center] * 8];
        const int cl = C.label[center];
        if (cl > 0) {
            for (int i = 0; i < 8; i++) {
                const int dx = dir[i][0], dy = dir[i][1];
                int nx = px + dx, ny = py + dy;
                int lx = -1, ly = -1;
                while (true) {
                    if (nx < 0 || nx >= W || ny < 0 || ny >= H) break;
                    const int nl = C.label[nx + ny * W];
                    if (nl == cl) {
                        lx = nx;
                        ly = ny;
                    } else if (nl == -1) {
                        break;
                    }
                    nx += dx;
                    ny += dy;
                }
                lb[i].x = (int16_t)lx;
                lb[i].y = (int16_t)ly;
            }
        }
    }
}

// FindNearestStrongPoint, DPE.cu:2859-2893
void FindNearestStrongPoint(Ctx &C, int center) {
    const int W = C.W, H = C.H;
    const int px = center % W, py = center / W;
    C.nearest[center] = S2{-1, -1};
    if (C.weak[center] != DPE_WEAK) return;
    for (int r = 0; r <= 100; ++r) {
        for (int x = -r; x <= r; ++x) {
            for (int y = -r; y <= r; ++y) {
                if (std::abs(x) != r && std::abs(y) != r) continue;
                const int nx = px + x, ny = py + y;
                if (nx < 0 || ny < 0 || nx >= W || ny >= H) continue;
                if (C.weak[nx + ny * W] == DPE_STRONG) {
                    C.nearest[center] = S2{(int16_t)nx, (int16_t)ny};
                    return;
                }
            }
        }
    }
}

// GenNeighbours, DPE.cu:2107-2467
// The ray search of GenNeighbours, DPE.cu:2170-2222: 8 base directions x
// rotate_time rotations; per direction, radii 2, 4, 8, ... (min(2r, r + 25)) up
// to MAX_SEARCH_RADIUS while the ray point is inside the image, 4 jittered
// probes per radius; a probe's pixel, or its nearest strong pixel when it is
// not STRONG, is taken if it lies within half the rotation angle of the
// direction and (edge_limit) no edge separates it from the centre.  Fills
// strong_points / dir_valid at index 4 * base direction + rotation; returns how
// many directions found a point.
int DirectionSearch(const Ctx &C, int px, int py, Rng &rng, bool edge_limit, S2 *strong_points, bool *dir_valid) {
    const int W = C.W, H = C.H;
    const int min_margin = 6;
    int odi = -1;
    int strong_point_size = 0;
    const int rotate_time = C.prm.rotate_time;
    const float cos_angle = C.gn_cos_angle, sin_angle = C.gn_sin_angle;
    const float threshhold = C.gn_threshold;
    const int shift_range = C.gn_shift_range;
    for (int odx = -1; odx <= 1; ++odx) {
        for (int ody = -1; ody <= 1; ++ody) {
            if (odx == 0 && ody == 0) continue;
            F2 od{(float)odx, (float)ody};
            NormalizeVec2(&od);
            odi++;
            for (int rot = 0; rot < rotate_time; ++rot) {
                const int di = odi * 4 + rot;
                for (int radius = 2; radius <= DPE_MAX_SEARCH_RADIUS; radius = std::min(radius * 2, radius + 25)) {
                    const float tx = (float)px + od.x * (float)radius;
                    const float ty = (float)py + od.y * (float)radius;
                    if (tx < 0 || ty < 0 || tx >= (float)W || ty >= (float)H) break;
                    for (int ri = 0; ri < 4; ++ri) {
                        const uint32_t s1 = rng.next();
                        const uint32_t m1 = rng.next();
                        const int rxs = (int)(((s1 % 2 == 0) ? 1u : 0xFFFFFFFFu) * m1 % (uint32_t)shift_range);
                        const uint32_t s2 = rng.next();
                        const uint32_t m2 = rng.next();
                        const int rys = (int)(((s2 % 2 == 0) ? 1u : 0xFFFFFFFFu) * m2 % (uint32_t)shift_range);
                        F2 dirv{od.x * 20 + (float)rxs, od.y * 20 + (float)rys};
                        NormalizeVec2(&dirv);
                        S2 np{(int16_t)(int)((float)px + dirv.x * (float)radius),
                              (int16_t)(int)((float)py + dirv.y * (float)radius)};
                        if (np.x < min_margin || np.y < min_margin || np.x >= W - min_margin || np.y >= H - min_margin)
                            continue;
                        int nc = np.x + np.y * W;
                        if (C.weak[nc] != DPE_STRONG) {
                            np = C.nearest[nc];
                            if (np.x == -1 || np.y == -1) continue;
                            nc = np.x + np.y * W;
                        }
                        F2 td{(float)(np.x - px), (float)(np.y - py)};
                        NormalizeVec2(&td);
                        const float ca = td.x * od.x + td.y * od.y;
                        if (ca > threshhold && (!edge_limit || !BresenhamLine(C, px, py, np.x, np.y))) {
                            strong_points[di] = np;
                            dir_valid[di] = true;
                            strong_point_size++;
                            break;
                        }
                    }
                    if (dir_valid[di]) break;
                }
                F2 rd;
                rd.x = od.x * cos_angle - od.y * sin_angle;
                rd.y = od.x * sin_angle + od.y * cos_angle;
                NormalizeVec2(&rd);
                od = rd;
            }
        }
    }
    return strong_point_size;
}

void GenNeighbours(Ctx &C, int center) {
    const int W = C.W, H = C.H;
    const int px = center % W, py = center / W;
    if (C.weak[center] != DPE_WEAK) return;
    const int max_pt_num = 64;
    const int min_margin = 6;
    const float depth_diff = C.prm.depth_max - C.prm.depth_min;
    const dpe_camera &cam = C.cams[0];
    S2 *neighbours = &C.neighbours[(size_t)C.nmap[center] * NNUM];
    Rng rng(C.prm.seed, center, PH_GEN_NEIGHBOURS, C.prm.call_index);
    for (int i = 0; i < NNUM; ++i) neighbours[i] = S2{-1, -1};
    neighbours[0] = S2{(int16_t)px, (int16_t)py};
    S2 strong_points[max_pt_num];
    bool dir_valid[max_pt_num];
    for (int i = 0; i < max_pt_num; ++i) {
        strong_points[i] = S2{-1, -1};
        dir_valid[i] = false;
    }
    const int rotate_time = C.prm.rotate_time;
    const float ransac_threshold = C.prm.ransac_threshold * depth_diff;
    bool edge_limit = false;
    if (C.prm.use_limit) {
        edge_limit = true;
        if (C.prm.use_edge) {
            const float cv = C.complex_map[center];
            const float rp = rng.uniform() - FLT_EPSILON;
            if (rp < cv) edge_limit = false;
            else C.complex_map[center] = std::max(0.99f, cv);
        }
    }
    int strong_point_size = DirectionSearch(C, px, py, rng, edge_limit, strong_points, dir_valid);
    int extend_index = 31;
    if (C.prm.use_label && C.label[center] > 0) {
        static const int dir[8][2] = {{0, -1}, {0, 1}, {-1, 0}, {1, 0}, {-1, -1}, {1, 1}, {-1, 1}, {1, -1}};
        const S2 *lb = &C.label_bound[(size_t)C.nmap[center] * 8];
        float bound_dist[8] = {0};
        int dir_step[8] = {0};
        for (int i = 0; i < 8; ++i) {
            const S2 bp = lb[i];
            float dist = 0.0f;
            if (bp.x != -1 && bp.y != -1) {
                dist = (float)std::sqrt((double)(px - bp.x) * (double)(px - bp.x) +
                                        (double)(py - bp.y) * (double)(py - bp.y));
                if (i >= 4) dist = (float)(dist / std::sqrt(2.0));
            }
            bound_dist[i] = dist;
            if (i % 2 == 1) {
                const float opposite = bound_dist[i - 1];
                const float ratio = 2 * rotate_time * dist / (dist + opposite);
                const int ri = (ratio != ratio) ? 0 : f2i_clamp(ratio, -1000000, 1000000);
                const int step = std::min(1, std::max(2 * rotate_time - 1, ri));
                dir_step[i - 1] = 2 * rotate_time - step;
                dir_step[i] = step;
            }
        }
        for (int i = 0; i < 8; ++i) {
            const float dist = bound_dist[i];
            const int gap = dir_step[i] + 1;
            const int step_len = std::max(1, (int)std::floor(1.0 * dist / gap));
            for (int step = 1; step <= dir_step[i]; ++step) {
                S2 np{(int16_t)(px + step * step_len * dir[i][0]), (int16_t)(py + step * step_len * dir[i][1])};
                if (np.x < min_margin || np.y < min_margin || np.x >= W - min_margin || np.y >= H - min_margin)
                    continue;
                int nc = np.x + np.y * W;
                if (C.weak[nc] != DPE_STRONG) {
                    np = C.nearest[nc];
                    if (np.x == -1 || np.y == -1) continue;
                    nc = np.x + np.y * W;
                }
                if (C.label[nc] != 0 && C.label[nc] != C.label[center]) continue;
                if (extend_index >= max_pt_num - 1) continue;
                extend_index++;
                strong_points[extend_index] = np;
                dir_valid[extend_index] = true;
                strong_point_size++;
            }
        }
    }
    uint8_t *rel = &C.weak_reliable[center];
    if (strong_point_size <= 3) {
        *rel = 0;
        return;
    }
    F4 best_plane{0, 0, 0, 0};
    bool has_valid_plane = false;
    S2 sv[max_pt_num];
    F3 s3[max_pt_num];
    F3 sn[max_pt_num];
    int vc = 0;
    float X[3];
    Get3DPoint(cam, px, py, C.planes[center].w, X);
    const F3 cpw{X[0], X[1], X[2]};
    for (int i = 0; i < max_pt_num; ++i) {
        sv[i] = S2{-1, -1};
        if (dir_valid[i]) {
            const S2 sp = strong_points[i];
            const int sc = sp.x + sp.y * W;
            sv[vc] = sp;
            Get3DPoint(cam, sp.x, sp.y, C.planes[sc].w, X);
            s3[vc] = F3{X[0], X[1], X[2]};
            const F4 n4 = TransformNormal2RefCam(cam, C.planes[sc]);
            sn[vc] = F3{n4.x, n4.y, n4.z};
            vc++;
        }
    }
    {
        int iteration = 50, max_iter = C.prm.high_res_img ? 200 : 125, max_count = 3;
        float min_cost = FLT_MAX;
        float residuals[max_pt_num] = {0.0f};
        float temp_thr = ransac_threshold;
        bool has_consist = false;
        bool must_in_triangle = (C.prm.use_label && C.label[center] > 0 && edge_limit) ? false : true;
        while (iteration > 0 && max_iter > 0) {
            max_iter--;
            const int a = (int)(rng.next() % (uint32_t)vc);
            const int b = (int)(rng.next() % (uint32_t)vc);
            const int c = (int)(rng.next() % (uint32_t)vc);
            if (a == b || b == c || a == c) continue;
            if (must_in_triangle && !PointinTriangle(sv[a], sv[b], sv[c], px, py)) continue;
            if (edge_limit) {
                if (BresenhamLine(C, sv[a].x, sv[a].y, sv[b].x, sv[b].y) ||
                    BresenhamLine(C, sv[b].x, sv[b].y, sv[c].x, sv[c].y) ||
                    BresenhamLine(C, sv[c].x, sv[c].y, sv[a].x, sv[a].y))
                    continue;
            }
            bool normal_consistency = false;
            if (C.prm.geom_consistency && edge_limit) {
                normal_consistency = true;
                if ((double)Vec3Dot(sn[a], sn[b]) < 0.8660254 || (double)Vec3Dot(sn[a], sn[c]) < 0.8660254 ||
                    (double)Vec3Dot(sn[b], sn[c]) < 0.8660254)
                    normal_consistency = false;
                if (has_consist && !normal_consistency) continue;
            }
            iteration--;
            const F3 &A = s3[a], &B = s3[b], &Cc = s3[c];
            const F3 AC{A.x - Cc.x, A.y - Cc.y, A.z - Cc.z};
            const F3 BC{B.x - Cc.x, B.y - Cc.y, B.z - Cc.z};
            F4 cr;
            cr.x = AC.y * BC.z - BC.y * AC.z;
            cr.y = -(AC.x * BC.z - BC.x * AC.z);
            cr.z = AC.x * BC.y - BC.x * AC.y;
            if ((cr.x == 0 && cr.y == 0 && cr.z == 0) || std::isnan(cr.x) || std::isnan(cr.y) || std::isnan(cr.z))
                continue;
            NormalizeVec3(&cr);
            cr.w = -(cr.x * A.x + cr.y * A.y + cr.z * A.z);
            int temp_count = 0;
            float strong_dist = 0.0f;
            for (int si = 0; si < vc; ++si) {
                const F3 &tp = s3[si];
                const S2 &pos = sv[si];
                const float fx = ((float)pos.x - cam.K[2]) / cam.K[0];
                const float fy = ((float)pos.y - cam.K[5]) / cam.K[4];
                const float fit_depth = -cr.w / (cr.x * fx + cr.y * fy + cr.z);
                const float distance = std::fabs(fit_depth - tp.z);
                residuals[si] = distance;
                if (distance < temp_thr) {
                    temp_count++;
                    strong_dist += distance;
                }
            }
            if (temp_count < 6) continue;
            if (temp_count > max_count) {
                if (!must_in_triangle && PointinTriangle(sv[a], sv[b], sv[c], px, py)) must_in_triangle = true;
                if (!has_consist && normal_consistency) has_consist = true;
                const float fx = ((float)px - cam.K[2]) / cam.K[0];
                const float fy = ((float)py - cam.K[5]) / cam.K[4];
                const float fit_depth = -cr.w / (cr.x * fx + cr.y * fy + cr.z);
                const float cd = std::fabs(fit_depth - cpw.z);
                best_plane = cr;
                max_count = temp_count;
                strong_dist /= temp_count;
                min_cost = cd;
                has_valid_plane = true;
                if ((double)temp_thr > (C.prm.high_res_img ? 0.05 : 0.005)) {
                    sort_small(residuals, vc);
                    if (temp_thr < residuals[NNUM]) continue;
                    temp_thr = (float)((double)residuals[NNUM] - 1e-6);
                    temp_count = 0;
                    for (int i = 0; i < vc; ++i) {
                        if (residuals[i] < temp_thr) temp_count++;
                        else break;
                    }
                    max_count = temp_count;
                }
            } else if (temp_count == max_count) {
                if (!must_in_triangle && PointinTriangle(sv[a], sv[b], sv[c], px, py)) must_in_triangle = true;
                const float fx = ((float)px - cam.K[2]) / cam.K[0];
                const float fy = ((float)py - cam.K[5]) / cam.K[4];
                const float fit_depth = -cr.w / (cr.x * fx + cr.y * fy + cr.z);
                const float cd = std::fabs(fit_depth - cpw.z);
                if (cd < min_cost) {
                    best_plane = cr;
                    max_count = temp_count;
                    strong_dist /= temp_count;
                    min_cost = cd;
                }
            }
        }
    }
    if (!has_valid_plane) {
        *rel = 0;
        return;
    }
    float weight[max_pt_num];
    for (int i = 0; i < vc; ++i) {
        const F3 &tp = s3[i];
        const S2 &pos = sv[i];
        const float fx = ((float)pos.x - cam.K[2]) / cam.K[0];
        const float fy = ((float)pos.y - cam.K[5]) / cam.K[4];
        const float fit_depth = -best_plane.w / (best_plane.x * fx + best_plane.y * fy + best_plane.z);
        const float distance = std::fabs(fit_depth - tp.z);
        if (distance >= ransac_threshold) {
            sv[i] = S2{-1, -1};
            weight[i] = FLT_MAX;
            continue;
        }
        weight[i] = distance;
    }
    sort_small_weighted(sv, weight, vc);
    for (int i = 1; i < NNUM; ++i) neighbours[i] = sv[i - 1];
    *rel = 1;
}

// RANSACToGetFitPlane, DPE.cu:2895-3128
void RANSACToGetFitPlane(Ctx &C, int center, int iter) {
    const int W = C.W;
    const int px = center % W, py = center / W;
    if (C.weak[center] != DPE_WEAK) {
        C.fit_planes[center] = C.planes[center];
        return;
    }
    Rng rng(C.prm.seed, center, ph_ransac(iter), C.prm.call_index);
    const dpe_camera &cam = C.cams[0];
    bool edge_limit = false;
    if (C.prm.use_limit) {
        edge_limit = true;
        if (C.prm.use_edge) {
            const float cv = C.complex_map[center];
            const float rp = rng.uniform() - FLT_EPSILON;
            if (rp < cv) edge_limit = false;
        }
    }
    S2 sp[NNUM - 1];
    F3 s3[NNUM - 1];
    F3 sn[NNUM - 1];
    int count = 0;
    float X[3];
    for (int i = 1; i < NNUM; ++i) {
        const S2 tp = GetNeighbourPoint(C, center, i);
        if (tp.x == -1 || tp.y == -1) continue;
        sp[count] = tp;
        const int tc = tp.x + tp.y * W;
        const float depth = ComputeDepthfromPlaneHypothesis(cam, C.planes[tc], tp.x, tp.y);
        Get3DPoint(cam, tp.x, tp.y, depth, X);
        s3[count] = F3{X[0], X[1], X[2]};
        const F4 n4 = C.planes[tc];
        sn[count] = F3{n4.x, n4.y, n4.z};
        count++;
    }
    if (count < 3) {
        C.fit_planes[center] = C.planes[center];
        return;
    }
    int iteration = 50;
    int ua = -1, ub = -1, uc = -1;
    float min_cost = FLT_MAX;
    F4 best{0, 0, 0, 0};
    bool has_best = false, has_strong = false;
    bool must_in_triangle = (C.prm.use_label && C.label[center] > 0 && edge_limit) ? false : true;
    while (iteration--) {
        const int a = (int)(rng.next() % (uint32_t)count);
        const int b = (int)(rng.next() % (uint32_t)count);
        const int c = (int)(rng.next() % (uint32_t)count);
        if (a == b || b == c || a == c) continue;
        bool is_strong = false;
        if (C.prm.geom_consistency && edge_limit) {
            is_strong = true;
            if ((double)Vec3Dot(sn[a], sn[b]) < 0.8660254 || (double)Vec3Dot(sn[a], sn[c]) < 0.8660254 ||
                (double)Vec3Dot(sn[b], sn[c]) < 0.8660254)
                is_strong = false;
            if (has_strong && !is_strong) continue;
        }
        if (must_in_triangle && !PointinTriangle(sp[a], sp[b], sp[c], px, py)) continue;
        if (edge_limit) {
            if (BresenhamLine(C, sp[a].x, sp[a].y, sp[b].x, sp[b].y) ||
                BresenhamLine(C, sp[b].x, sp[b].y, sp[c].x, sp[c].y) ||
                BresenhamLine(C, sp[c].x, sp[c].y, sp[a].x, sp[a].y))
                continue;
        }
        const F3 &A = s3[a], &B = s3[b], &Cc = s3[c];
        const F3 AC{A.x - Cc.x, A.y - Cc.y, A.z - Cc.z};
        const F3 BC{B.x - Cc.x, B.y - Cc.y, B.z - Cc.z};
        F4 cr;
        cr.x = AC.y * BC.z - BC.y * AC.z;
        cr.y = -(AC.x * BC.z - BC.x * AC.z);
        cr.z = AC.x * BC.y - BC.x * AC.y;
        if ((cr.x == 0 && cr.y == 0 && cr.z == 0) || std::isnan(cr.x) || std::isnan(cr.y) || std::isnan(cr.z))
            continue;
        NormalizeVec3(&cr);
        cr.w = -(cr.x * A.x + cr.y * A.y + cr.z * A.z);
        if (!has_strong && is_strong) has_strong = true;
        float tcost = 0.0f;
        for (int si = 0; si < count; ++si) {
            if (si == a || si == b || si == c) continue;
            const F3 &tp = s3[si];
            const S2 &pix = sp[si];
            const float fx = ((float)pix.x - cam.K[2]) / cam.K[0];
            const float fy = ((float)pix.y - cam.K[5]) / cam.K[4];
            const float fit_depth = -cr.w / (cr.x * fx + cr.y * fy + cr.z);
            tcost += std::fabs(fit_depth - tp.z);
        }
        if (tcost < min_cost) {
            if (!must_in_triangle && PointinTriangle(sp[a], sp[b], sp[c], px, py)) must_in_triangle = true;
            min_cost = tcost;
            best = cr;
            has_best = true;
            ua = a; ub = b; uc = c;
        }
    }
    if (has_best) {
        const float depth = ComputeDepthfromPlaneHypothesis(cam, C.planes[center], px, py);
        const F4 vd = GetViewDirection(cam, px, py, depth);
        const float dot = best.x * vd.x + best.y * vd.y + best.z * vd.z;
        if (dot > 0) {
            best.x = -best.x; best.y = -best.y; best.z = -best.z; best.w = -best.w;
        }
        C.fit_planes[center] = best;
        if (C.prm.use_radius) {
            if (must_in_triangle) {
                const S2 &A = sp[ua], &B = sp[ub], &Cc = sp[uc];
                const float a = std::sqrt((float)((A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y)));
                const float b = std::sqrt((float)((B.x - Cc.x) * (B.x - Cc.x) + (B.y - Cc.y) * (B.y - Cc.y)));
                const float c = std::sqrt((float)((Cc.x - A.x) * (Cc.x - A.x) + (Cc.y - A.y) * (Cc.y - A.y)));
                const float p = (float)((a + b + c) / 2.0);
                const float S = std::sqrt(p * (p - a) * (p - b) * (p - c));
                // (int)floor of NaN (degenerate triangle) is defined as 0 on both paths
                const double rr = std::floor(std::sqrt(S) / 2.0);
                int radius = (rr != rr) ? 0 : (int)std::fmax(-1e9, std::fmin(1e9, rr));
                const float A_dis = std::sqrt((float)((A.x - px) * (A.x - px) + (A.y - py) * (A.y - py)));
                const float B_dis = std::sqrt((float)((B.x - px) * (B.x - px) + (B.y - py) * (B.y - py)));
                const float C_dis = std::sqrt((float)((Cc.x - px) * (Cc.x - px) + (Cc.y - py) * (Cc.y - py)));
                const float min_dis = std::min(std::min(A_dis, B_dis), C_dis);
                if (2.5 * min_dis < radius) radius = (int)min_dis;
                if (edge_limit) {
                    if (C.prm.use_edge) {
                        float med = FLT_MAX;
                        const S2 *en = &C.edge_neigh[(size_t)center * 8];
                        for (int d = 0; d < 8; ++d) {
                            const S2 ep = en[d];
                            if (ep.x == -1 || ep.y == -1) continue;
                            const float dist =
                                std::sqrt((float)((ep.x - px) * (ep.x - px) + (ep.y - py) * (ep.y - py)));
                            med = std::min(med, dist);
                        }
                        if (med < radius) radius = (int)med;
                    }
                    if (C.prm.use_label && C.label[center] > 0) {
                        float mbd = FLT_MAX;
                        const S2 *lb = &C.label_bound[(size_t)C.nmap[center] * 8];
                        for (int d = 0; d < 8; ++d) {
                            const S2 bp = lb[d];
                            if (bp.x == -1 || bp.y == -1) continue;
                            const float dist = (float)std::sqrt((double)(px - bp.x) * (double)(px - bp.x) +
                                                                (double)(py - bp.y) * (double)(py - bp.y));
                            mbd = std::min(mbd, dist);
                        }
                        if (mbd < radius) radius = (int)mbd;
                    }
                }
                while ((radius << 1) % 5 != 0) radius--;
                if (!edge_limit)
                    C.radius[center] = radius > C.prm.strong_radius ? 0 : C.prm.strong_radius;
                else
                    C.radius[center] = radius > C.prm.strong_radius ? radius : C.prm.strong_radius;
            } else {
                C.radius[center] = C.prm.strong_radius;
            }
        }
    } else {
        C.fit_planes[center] = F4{0, 0, 0, 0};
        if (C.prm.use_radius) C.radius[center] = C.prm.strong_radius;
    }
}

// Baseline of DepthToWeak / LocalRefine, DPE.cu:2644-2650 (float sqrt of a float sum)
inline float CameraDistance(const Ctx &C, int src_index) {
    float d[3];
    d[0] = C.cams[0].c[0] - C.cams[src_index].c[0];
    d[1] = C.cams[0].c[1] - C.cams[src_index].c[1];
    d[2] = C.cams[0].c[2] - C.cams[src_index].c[2];
    return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

// The peak analysis of DepthToWeak (DPE.cu:2699-2750) over the 61 sampled
// costs (disparity offsets -30..30): WEAK unless the lowest strict local
// minimum sits within weak_peak_radius of the centre with cost <= 0.5, and then
// STRONG if it is the only one at <= 0.15, or if the other minima spread enough.
uint8_t DepthToWeakState(const float *pc, int weak_peak_radius) {
    const int radius = 30;
    const int n = 2 * radius + 1;
    bool is_peak[n];
    for (int i = 0; i < n; ++i) is_peak[i] = false;
    int peak_count = 0, min_peak = 0;
    float min_cost = 2.0f;
    for (int i = 2; i < n - 2; ++i) {
        if (pc[i - 1] > pc[i] && pc[i + 1] > pc[i]) {
            is_peak[i] = true;
            peak_count++;
            if (pc[i] < min_cost) {
                min_peak = i;
                min_cost = pc[i];
            }
        }
    }
    if (std::abs(min_peak - radius) > weak_peak_radius || pc[min_peak] > 0.5f) {
        return DPE_WEAK;
    }
    if (peak_count == 1) {
        return pc[min_peak] <= 0.15f ? DPE_STRONG : DPE_WEAK;
    }
    float var = 0.0f;
    for (int i = 2; i < n - 2; ++i) {
        if (is_peak[i] && i != min_peak) {
            const float dist = pc[i] - min_cost;
            var += dist * dist;
        }
    }
    var = std::sqrt(var);
    var /= (peak_count - 1);
    return var > 0.2f ? DPE_STRONG : DPE_WEAK;
}

// DepthToWeak, DPE.cu:2597-2751
void DepthToWeak(Ctx &C, int center) {
    // The cost of the current plane below is never read afterwards in the
    // reference (DPE.cu:2629-2658): counted apart as dead evaluations.
    tl_family = kFamDead;
    const int W = C.W, H = C.H;
    const int px = center % W, py = center / W;
    const int min_margin = 6;
    if (px < min_margin || py < min_margin || px >= W - min_margin || py >= H - min_margin) {
        C.weak[center] = DPE_UNKNOWN;
        return;
    }
    const dpe_camera &cam = C.cams[0];
    const uint32_t sel = C.selected[center];
    const uint8_t *vw = &C.view_weight[(size_t)MAXIMG * center];
    F4 op = TransformNormal2RefCam(cam, C.planes[center]);
    const float od = op.w;
    if (od == 0) {
        C.weak[center] = DPE_UNKNOWN;
        return;
    }
    const bool geom = C.prm.geom_consistency != 0;
    float cost_now = 0.0f, base_line = 0;
    int valid = 0;
    float wn = 0.0f;
    for (int s = 1; s < C.N; ++s) {
        const int v = s - 1;
        if (isSet(sel, v)) {
            F4 tp = op;
            tp.w = GetDistance2Origin(cam, px, py, od, tp);
            float tc = ComputeBilateralNCCOld(C, px, py, s, tp);
            if (geom) tc += C.prm.geom_factor * ComputeGeomConsistencyCost(C, px, py, s, tp);
            cost_now += (tc * vw[v]);
            wn += vw[v];
            base_line += CameraDistance(C, s);
            valid++;
        }
    }
    if (valid == 0) {
        C.weak[center] = DPE_UNKNOWN;
        return;
    }
    cost_now /= wn;
    base_line /= valid;
    tl_family = kFamPost;
    const float disp = cam.K[0] * base_line / od;
    const int radius = 30;
    const int n = 2 * radius + 1;
    float pc[n];
    for (int d = -radius; d <= radius; d += 1) {
        // samples the peak analysis never reads (d = +30 always, d = -30 unless
        // weak_peak_radius >= 30) are computed as the reference does but counted
        // apart as dead evaluations
        tl_family = (d == radius || (d == -radius && C.prm.weak_peak_radius < 30)) ? kFamDead : kFamPost;
        const float pd = cam.K[0] * base_line / (disp + d);
        if (pd < C.prm.depth_min || pd > C.prm.depth_max) {
            pc[d + radius] = 2.0f;
            continue;
        }
        F4 tp = op;
        tp.w = GetDistance2Origin(cam, px, py, pd, tp);
        float p_cost = 0.0f;
        for (int s = 1; s < C.N; ++s) {
            const int v = s - 1;
            float tc = 0.0f;
            if (isSet(sel, v)) {
                tc += ComputeBilateralNCCOld(C, px, py, s, tp);
                if (geom) tc += C.prm.geom_factor * ComputeGeomConsistencyCost(C, px, py, s, tp);
                p_cost += (tc * vw[v]);
            }
        }
        p_cost /= wn;
        pc[d + radius] = std::min(2.0f, p_cost);
    }
    C.weak[center] = DepthToWeakState(pc, C.prm.weak_peak_radius);
}

// LocalRefine, DPE.cu:2753-2839
void LocalRefine(Ctx &C, int center) {
    const int W = C.W;
    const int px = center % W, py = center / W;
    const dpe_camera &cam = C.cams[0];
    const uint32_t sel = C.selected[center];
    const uint8_t *vw = &C.view_weight[(size_t)MAXIMG * center];
    {
        // with zero total weight the cost of the current plane is never read (early return below)
        float wsum = 0.0f;
        for (int s = 1; s < C.N; ++s)
            if (isSet(sel, s - 1)) wsum += vw[s - 1];
        tl_family = wsum == 0 ? kFamDead : kFamPost;
    }
    F4 op = TransformNormal2RefCam(cam, C.planes[center]);
    const float od = op.w;
    if (od == 0) return;
    const bool geom = C.prm.geom_consistency != 0;
    float cost_now = 0.0f, base_line = 0;
    int valid = 0;
    float wn = 0.0f;
    for (int s = 1; s < C.N; ++s) {
        const int v = s - 1;
        if (isSet(sel, v)) {
            F4 tp = op;
            tp.w = GetDistance2Origin(cam, px, py, od, tp);
            float tc = ComputeBilateralNCCOld(C, px, py, s, tp);
            if (geom) tc += C.prm.geom_factor * ComputeGeomConsistencyCost(C, px, py, s, tp);
            cost_now += (tc * vw[v]);
            wn += vw[v];
            base_line += CameraDistance(C, s);
            valid++;
        }
    }
    if (wn == 0 || valid == 0) return;
    tl_family = kFamPost;
    // where DepthToWeak ran its samples, its d = -5..5 planes are these: the
    // per-view costs are the same values, so they count once (as DepthToWeak's)
    const int min_margin = 6;
    const bool dtw_ran = !(px < min_margin || py < min_margin || px >= W - min_margin || py >= C.H - min_margin);
    cost_now /= wn;
    base_line /= valid;
    const float disp = cam.K[0] * base_line / od;
    const int radius = 5;
    float min_cost = 2.0f;
    float best_depth = od;
    tl_family = dtw_ran ? kFamDead : kFamPost;
    for (int d = -radius; d <= radius; ++d) {
        const float pd = cam.K[0] * base_line / (disp + d);
        if (pd < C.prm.depth_min || pd > C.prm.depth_max) continue;
        F4 tp = op;
        tp.w = GetDistance2Origin(cam, px, py, pd, tp);
        float tc = 0.0f;
        for (int s = 1; s < C.N; ++s) {
            const int v = s - 1;
            if (isSet(sel, v)) {
                tc += (ComputeBilateralNCCOld(C, px, py, s, tp) * vw[v]);
                if (geom) tc += (C.prm.geom_factor * ComputeGeomConsistencyCost(C, px, py, s, tp) * vw[v]);
            }
        }
        tc /= wn;
        if (tc < min_cost) {
            min_cost = tc;
            best_depth = pd;
        }
    }
    if ((double)(cost_now - min_cost) > 0.1) C.planes[center].w = best_depth;
}

// Pixels of one checkerboard colour as the HALF launch covers them
// (DPE.cu:3148-3155, 1868-1942): colour 0 = Black = (x+y) even.
void colour_pixels(const Ctx &C, int colour, std::vector<int> &out) {
    out.clear();
    const int rows = std::min(C.H, 2 * (((C.H / 2) + 15) / 16) * 16);
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < C.W; ++x)
            if (((x + y) & 1) == colour) out.push_back(x + y * C.W);
}

// DPE.cu:2151-2157, evaluated on the host in double as the device code does
void set_gen_neighbours_constants(Ctx &C) {
    const float angle = 45.0f / C.prm.rotate_time;
    C.gn_cos_angle = (float)std::cos(angle * M_PI / 180.f);
    C.gn_sin_angle = (float)std::sin(angle * M_PI / 180.f);
    C.gn_threshold = (float)std::cos((angle / 2.0f) * M_PI / 180.0f);
    C.gn_shift_range = std::max((int)(std::tan((angle / 2.0f) * M_PI / 180.0f) * 20), 1);
}

int setup(Ctx &C, const dpe_params *params, const dpe_inputs *in, int nthreads) {
    if (!params || !in) return DPE_ERR_INVALID_ARGUMENT;
    if (in->num_images > MAXIMG) return DPE_ERR_TOO_MANY_IMAGES;
    if (in->num_images < 2 || in->width <= 0 || in->height <= 0) return DPE_ERR_INVALID_ARGUMENT;
    if (in->memory != DPE_MEM_HOST) return DPE_ERR_UNSUPPORTED;
    C.W = in->width;
    C.H = in->height;
    C.N = in->num_images;
    C.P = C.W * C.H;
    C.prm = *params;
    C.prm.num_images = C.N;
    C.nthreads = nthreads;
    C.tally.assign(nthreads, Tally{});
    for (int i = 0; i < C.N; ++i) {
        C.cams[i] = in->cameras[i];
        C.img[i] = in->images[i];
        C.gdepth[i] = in->geom_depths[i];
        if (!C.img[i]) return DPE_ERR_MISSING_INPUT;
        if (C.prm.geom_consistency && !C.gdepth[i]) return DPE_ERR_MISSING_INPUT;
    }
    const size_t P = (size_t)C.P;
    C.planes.assign(P, F4{0, 0, 0, 0});
    C.costs.assign(P, 0.0f);
    C.selected.assign(P, 0u);
    C.known.assign(P, 0u);
    {
        F4 none;
        std::memset(&none, 0xFF, sizeof(F4));
        C.last_cand.assign((size_t)kCandSlots * P, none);
    }
    C.view_weight.assign(P * MAXIMG, 0);
    C.fit_planes.assign(P, F4{0, 0, 0, 0});
    C.weak.assign(P, DPE_STRONG);
    C.weak_reliable.assign(P, 0);
    C.nearest.assign(P, S2{-1, -1});
    C.nmap.assign(P, 0);
    C.radius.assign(P, C.prm.strong_radius);
    C.weak_count = 0;
    if (C.prm.use_APD) {
        if (!in->weak_info) return DPE_ERR_MISSING_INPUT;
        for (size_t i = 0; i < P; ++i) {
            C.weak[i] = in->weak_info[i];
            if (C.weak[i] == DPE_WEAK) C.nmap[i] = C.weak_count++;
        }
    }
    C.neighbours.assign((size_t)C.weak_count * NNUM + 1, S2{-1, -1});
    if (C.prm.state != DPE_FIRST_INIT) {
        if (!in->prior_planes || !in->prior_selected) return DPE_ERR_MISSING_INPUT;
        for (size_t i = 0; i < P; ++i) {
            C.planes[i] = F4{in->prior_planes[4 * i], in->prior_planes[4 * i + 1], in->prior_planes[4 * i + 2],
                             in->prior_planes[4 * i + 3]};
            C.selected[i] = in->prior_selected[i];
        }
    }
    if (C.prm.use_edge || C.prm.use_limit) {
        if (!in->edges || !in->edges_low) return DPE_ERR_MISSING_INPUT;
        C.edge = in->edges;
        C.edge_low = in->edges_low;
        C.low_w = in->low_width;
        C.low_h = in->low_height;
    }
    if (C.prm.use_edge) {
        C.edge_neigh.assign(P * 8, S2{-1, -1});
        C.complex_map.assign(P, 0.0f);
    }
    if (C.prm.use_label) {
        if (!in->labels) return DPE_ERR_MISSING_INPUT;
        C.label = in->labels;
        C.label_bound.assign((size_t)C.weak_count * 8 + 1, S2{-1, -1});
    }
    set_gen_neighbours_constants(C);
    return DPE_OK;
}

void take_snapshot(Ctx &C) {
    C.snap_planes = C.planes;
    C.snap_costs = C.costs;
    C.snap_selected = C.selected;
}

// DPE::RunPatchMatch schedule, DPE.cu:3130-3262
void run_patchmatch(Ctx &C) {
    const int P = C.P, nt = C.nthreads;
    parallel_for(P, nt, [&](int i) { GenEdgeInform(C, i); });
    parallel_for(P, nt, [&](int i) { FindNearestStrongPoint(C, i); });
    parallel_for(P, nt, [&](int i) { GenNeighbours(C, i); });
    for (int i = 0; i < P; ++i)  // NeigbourUpdate, DPE.cu:2469-2485
        if (C.weak[i] == DPE_WEAK && C.weak_reliable[i] != 1) C.weak[i] = DPE_UNKNOWN;
    parallel_for(P, nt, [&](int i) { RandomInitialization(C, i); });
    std::vector<int> colour[2];
    colour_pixels(C, 0, colour[0]);
    colour_pixels(C, 1, colour[1]);
    for (int it = 0; it < C.prm.max_iterations; ++it) {
        for (int c = 0; c < 2; ++c) {
            take_snapshot(C);
            const std::vector<int> &px = colour[c];
            parallel_for((int)px.size(), nt, [&](int k) {
                const int i = px[k];
                if (C.weak[i] == DPE_WEAK) return;
                CheckerboardPropagationStrong(C, i % C.W, i / C.W, it);
            });
        }
        parallel_for(P, nt, [&](int i) { RANSACToGetFitPlane(C, i, it); });
        for (int c = 0; c < 2; ++c) {
            const std::vector<int> &px = colour[c];
            parallel_for((int)px.size(), nt, [&](int k) {
                const int i = px[k];
                if (C.weak[i] != DPE_WEAK) return;
                CheckerboardPropagationWeak(C, i % C.W, i / C.W, it);
            });
        }
    }
    parallel_for(P, nt, [&](int i) { GetDepthandNormal(C, i); });
    for (int c = 0; c < 2; ++c) {
        const std::vector<int> &px = colour[c];
        parallel_for((int)px.size(), nt, [&](int k) {
            const int i = px[k];
            if (C.weak[i] != DPE_WEAK) CheckerboardFilterStrong(C, i % C.W, i / C.W);
        });
    }
    parallel_for(P, nt, [&](int i) { DepthToWeak(C, i); });
    parallel_for(P, nt, [&](int i) { LocalRefine(C, i); });
}

// ------------------------------------------------------------------ fusion
// Get3DPointonWorld (host), DPE.cpp:1337-1361
F3 FuseGet3DPointonWorld(int x, int y, float depth, const dpe_camera &cam) {
    F3 X, T, Cc;
    X.x = depth * ((float)x - cam.K[2]) / cam.K[0];
    X.y = depth * ((float)y - cam.K[5]) / cam.K[4];
    X.z = depth;
    T.x = cam.R[0] * X.x + cam.R[3] * X.y + cam.R[6] * X.z;
    T.y = cam.R[1] * X.x + cam.R[4] * X.y + cam.R[7] * X.z;
    T.z = cam.R[2] * X.x + cam.R[5] * X.y + cam.R[8] * X.z;
    Cc.x = -(cam.R[0] * cam.t[0] + cam.R[3] * cam.t[1] + cam.R[6] * cam.t[2]);
    Cc.y = -(cam.R[1] * cam.t[0] + cam.R[4] * cam.t[1] + cam.R[7] * cam.t[2]);
    Cc.z = -(cam.R[2] * cam.t[0] + cam.R[5] * cam.t[1] + cam.R[8] * cam.t[2]);
    X.x = T.x + Cc.x;
    X.y = T.y + Cc.y;
    X.z = T.z + Cc.z;
    return X;
}

// int(v + 0.5f) of RunFusion (DPE.cpp:1491-1492); NaN and huge values map to -1
// (outside every image).
template <typename T>
inline int fuse_round(T v) {
    const T r = v + (T)0.5;
    if (!(r > (T)-1.0e9 && r < (T)1.0e9)) return -1;
    return (int)r;
}

// The fusion geometry in scalar type T (float: the reference; double: the
// precision form's FP64 reprojection check, camera entries widened, the same
// operations in the same order).
template <typename T> struct FV3 { T x, y, z; };
template <typename T> struct FV2 { T x, y; };
template <typename T>
FV3<T> fuse_world(int x, int y, T depth, const dpe_camera &cam) {
    const float *R = cam.R, *K = cam.K, *t = cam.t;
    FV3<T> X, Tr, Cc;
    X.x = depth * ((T)x - (T)K[2]) / (T)K[0];
    X.y = depth * ((T)y - (T)K[5]) / (T)K[4];
    X.z = depth;
    Tr.x = (T)R[0] * X.x + (T)R[3] * X.y + (T)R[6] * X.z;
    Tr.y = (T)R[1] * X.x + (T)R[4] * X.y + (T)R[7] * X.z;
    Tr.z = (T)R[2] * X.x + (T)R[5] * X.y + (T)R[8] * X.z;
    Cc.x = -((T)R[0] * (T)t[0] + (T)R[3] * (T)t[1] + (T)R[6] * (T)t[2]);
    Cc.y = -((T)R[1] * (T)t[0] + (T)R[4] * (T)t[1] + (T)R[7] * (T)t[2]);
    Cc.z = -((T)R[2] * (T)t[0] + (T)R[5] * (T)t[1] + (T)R[8] * (T)t[2]);
    X.x = Tr.x + Cc.x;
    X.y = Tr.y + Cc.y;
    X.z = Tr.z + Cc.z;
    return X;
}
template <typename T>
void fuse_proj(const FV3<T> &X, const dpe_camera &cam, FV2<T> &pt, T &depth) {
    const float *R = cam.R, *K = cam.K, *t0 = cam.t;
    FV3<T> t;
    t.x = (T)R[0] * X.x + (T)R[1] * X.y + (T)R[2] * X.z + (T)t0[0];
    t.y = (T)R[3] * X.x + (T)R[4] * X.y + (T)R[5] * X.z + (T)t0[1];
    t.z = (T)R[6] * X.x + (T)R[7] * X.y + (T)R[8] * X.z + (T)t0[2];
    depth = (T)K[6] * t.x + (T)K[7] * t.y + (T)K[8] * t.z;
    pt.x = ((T)K[0] * t.x + (T)K[1] * t.y + (T)K[2] * t.z) / depth;
    pt.y = ((T)K[3] * t.x + (T)K[4] * t.y + (T)K[5] * t.z) / depth;
}

int process_impl(const dpe_params *params, const dpe_inputs *in, dpe_outputs *out, int nthreads, int64_t *counters,
                 bool cost64) {
    Ctx C;
    C.cost64 = cost64;
    const int st = setup(C, params, in, nthreads < 1 ? 1 : nthreads);
    if (st != DPE_OK) return st;
    if (!out || !out->planes || !out->costs || !out->pixel_states || !out->selected_views)
        return DPE_ERR_INVALID_ARGUMENT;
    run_patchmatch(C);
    for (int i = 0; i < C.P; ++i) {
        out->planes[4 * i + 0] = C.planes[i].x;
        out->planes[4 * i + 1] = C.planes[i].y;
        out->planes[4 * i + 2] = C.planes[i].z;
        out->planes[4 * i + 3] = C.planes[i].w;
        out->costs[i] = C.costs[i];
        out->pixel_states[i] = C.weak[i];
        out->selected_views[i] = C.selected[i];
    }
    if (counters) {
        counters[0] = counters[1] = 0;
        for (int f = 0; f < kFamilies; ++f) {
            int64_t h = 0, t = 0;
            for (const Tally &w : C.tally) {
                h += w.hyps[f];
                t += w.taps[f];
            }
            counters[2 + 2 * f] = h;
            counters[3 + 2 * f] = t;
            counters[0] += h;
            counters[1] += t;
        }
    }
    return DPE_OK;
}
}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

// One process call (DPE.cu:3130-3262) on the fp32 oracle.
int oracle_process(const dpe_params *params, const dpe_inputs *in, dpe_outputs *out, int nthreads,
                   int64_t *counters) {
    return process_impl(params, in, out, nthreads, counters, false);
}

// The same call with the multi-view cost evaluated in double (FP64 cost path).
int oracle_process_fp64(const dpe_params *params, const dpe_inputs *in, dpe_outputs *out, int nthreads,
                        int64_t *counters) {
    return process_impl(params, in, out, nthreads, counters, true);
}

// ComputeMultiViewCostVectorOld over given camera-frame planes (test access).
static int ncc_cost_impl(const dpe_params *params, const dpe_inputs *in, int num_hyps, const float *planes,
                         float *costs, int nthreads, bool cost64) {
    dpe_params p = *params;
    p.use_APD = 0; p.use_edge = 0; p.use_limit = 0; p.use_label = 0; p.geom_consistency = 0;
    p.state = DPE_FIRST_INIT;
    Ctx C;
    C.cost64 = cost64;
    const int st = setup(C, &p, in, nthreads < 1 ? 1 : nthreads);
    if (st != DPE_OK) return st;
    const int NV = C.N - 1;
    const size_t P = (size_t)C.P;
    parallel_for(C.P * num_hyps, C.nthreads, [&](int k) {
        tl_family = kFamFull;
        const int h = k / C.P, i = k % C.P;
        const float *pl = planes + ((size_t)h * P + i) * 4;
        const F4 plane{pl[0], pl[1], pl[2], pl[3]};
        float cv[32];
        CostVectorOld(C, i % C.W, i / C.W, plane, cv);
        for (int v = 0; v < NV; ++v) costs[((size_t)h * NV + v) * P + i] = cv[v];
    });
    return DPE_OK;
}

int oracle_ncc_cost(const dpe_params *params, const dpe_inputs *in, int num_hyps, const float *planes, float *costs,
                    int nthreads) {
    return ncc_cost_impl(params, in, num_hyps, planes, costs, nthreads, false);
}

// The same cost vectors through the FP64 cost path.
int oracle_ncc_cost_fp64(const dpe_params *params, const dpe_inputs *in, int num_hyps, const float *planes,
                         float *costs, int nthreads) {
    return ncc_cost_impl(params, in, num_hyps, planes, costs, nthreads, true);
}

}  // extern "C"

namespace oracle {
namespace {
// Fusion, RunFusion DPE.cpp:1387-1548 (sequential, pair order, raster order);
// T = double: the precision form's FP64 reprojection geometry (the errors are
// rounded to fp32 for the threshold tests and the weighting).
template <typename T>
int fuse_impl(const dpe_fuse_view *views, int num_views, const dpe_fuse_params *fp, dpe_point *points,
              int64_t capacity, int64_t *num_points) {
    std::vector<std::vector<uint8_t>> masks(num_views);
    for (int i = 0; i < num_views; ++i) masks[i].assign((size_t)views[i].width * views[i].height, 0);
    int64_t count = 0;
    for (int i = 0; i < num_views; ++i) {
        const dpe_fuse_view &rv = views[i];
        const int cols = rv.width, rows = rv.height;
        const int num_ngb = rv.num_src;
        for (int r = 0; r < rows; ++r) {
            for (int c = 0; c < cols; ++c) {
                const size_t rp = (size_t)r * cols + c;
                if (masks[i][rp] == 1) continue;
                const float ref_depth = rv.depth[rp];
                if (ref_depth <= 0.0) continue;
                const float *rn = rv.normal + 3 * rp;
                const FV3<T> PointX = fuse_world<T>(c, r, (T)ref_depth, rv.camera);
                int num_consistent = 0;
                float dyn = 0.0f;
                int used[DPE_MAX_IMAGES][2];
                for (int j = 0; j < num_ngb; ++j) used[j][0] = used[j][1] = -1;
                for (int j = 0; j < num_ngb; ++j) {
                    const int si = rv.src_index[j];
                    const dpe_fuse_view &sv = views[si];
                    FV2<T> pt;
                    T proj_depth;
                    fuse_proj<T>(PointX, sv.camera, pt, proj_depth);
                    const int src_r = fuse_round<T>(pt.y);
                    const int src_c = fuse_round<T>(pt.x);
                    if (src_c >= 0 && src_c < sv.width && src_r >= 0 && src_r < sv.height) {
                        const size_t sp = (size_t)src_r * sv.width + src_c;
                        if (masks[si][sp] == 1) continue;
                        const float src_depth = sv.depth[sp];
                        if (src_depth <= 0.0) continue;
                        const float *sn = sv.normal + 3 * sp;
                        const FV3<T> tX = fuse_world<T>(src_c, src_r, (T)src_depth, sv.camera);
                        FV2<T> tp;
                        fuse_proj<T>(tX, rv.camera, tp, proj_depth);
                        // sqrt(pow(c - x, 2) + pow(r - y, 2)) evaluates in double (DPE.cpp:1506)
                        const T dx = (T)c - tp.x, dy = (T)r - tp.y;
                        const float reproj = (float)std::sqrt((double)dx * (double)dx + (double)dy * (double)dy);
                        const float rel = (float)(std::fabs(proj_depth - (T)ref_depth) / (T)ref_depth);
                        const float dot = rn[0] * sn[0] + rn[1] * sn[1] + rn[2] * sn[2];
                        const float angle = o_acosf(dot);
                        const bool angle_ok = (angle != angle) ? true : (angle < fp->max_angle_rad);
                        if (reproj < fp->max_reproj_px && rel < fp->max_rel_depth && angle_ok) {
                            used[j][0] = src_c;
                            used[j][1] = src_r;
                            const float ang = (angle != angle) ? 0.0f : angle;
                            const float idx = reproj + 200 * rel + ang * 10;
                            dyn += o_expf(-idx);
                            num_consistent++;
                        }
                    }
                }
                const float factor = rv.pixel_states[rp] == DPE_WEAK ? fp->weak_factor : fp->strong_factor;
                if (num_consistent >= fp->min_consistent && num_consistent >= 1 && (dyn > factor * num_consistent)) {
                    float col[3] = {(float)rv.color[3 * rp], (float)rv.color[3 * rp + 1], (float)rv.color[3 * rp + 2]};
                    for (int j = 0; j < num_ngb; ++j) {
                        if (used[j][0] == -1) continue;
                        const int si = rv.src_index[j];
                        const size_t sp = (size_t)used[j][1] * views[si].width + used[j][0];
                        masks[si][sp] = 1;
                        col[0] += views[si].color[3 * sp];
                        col[1] += views[si].color[3 * sp + 1];
                        col[2] += views[si].color[3 * sp + 2];
                    }
                    col[0] /= (num_consistent + 1);
                    col[1] /= (num_consistent + 1);
                    col[2] /= (num_consistent + 1);
                    if (points && count < capacity) {
                        dpe_point &o = points[count];
                        o.x = (float)PointX.x;
                        o.y = (float)PointX.y;
                        o.z = (float)PointX.z;
                        o.b = (uint8_t)col[0];
                        o.g = (uint8_t)col[1];
                        o.r = (uint8_t)col[2];
                        o.pad = 0;
                    }
                    count++;
                }
            }
        }
    }
    *num_points = count;
    return DPE_OK;
}

}  // namespace
}  // namespace oracle

extern "C" {

int oracle_fuse(const dpe_fuse_view *views, int num_views, const dpe_fuse_params *fp, dpe_point *points,
                int64_t capacity, int64_t *num_points) {
    if (!views || num_views <= 0 || !fp || !num_points) return DPE_ERR_INVALID_ARGUMENT;
    if (fp->fp64_reprojection == 1) return fuse_impl<double>(views, num_views, fp, points, capacity, num_points);
    if (fp->fp64_reprojection != 0) return DPE_ERR_INVALID_ARGUMENT;
    return fuse_impl<float>(views, num_views, fp, points, capacity, num_points);
}

// RunFusion_TAT_Intermediate (DPE.cpp:1550-1718) and RunFusion_TAT_advanced
// (DPE.cpp:1720-1867), sequential: views in pair order, pixels in raster order,
// one CostData record per source kept across the pixels of a view (DPE.cpp:1637,
// 1788) and overwritten only where the source pixel passes its checks.
int oracle_fuse_tat(const dpe_fuse_view *views, int num_views, const uint8_t *const *blocks,
                    const dpe_fuse_tat_params *tp, dpe_point *points, int64_t capacity, int64_t *num_points) {
    if (!views || num_views <= 0 || !tp || !num_points) return DPE_ERR_INVALID_ARGUMENT;
    if (tp->variant != DPE_TAT_INTERMEDIATE && tp->variant != DPE_TAT_ADVANCED) return DPE_ERR_INVALID_ARGUMENT;
    const bool inter = tp->variant == DPE_TAT_INTERMEDIATE;
    struct CostData {
        float dist = FLT_MAX, depth = FLT_MAX, angle = FLT_MAX;
        int src_r = 0, src_c = 0;
        bool use = false;
    };
    std::vector<std::vector<uint8_t>> masks(num_views);
    for (int i = 0; i < num_views; ++i) masks[i].assign((size_t)views[i].width * views[i].height, 0);
    int64_t count = 0;
    for (int i = 0; i < num_views; ++i) {
        const dpe_fuse_view &rv = views[i];
        const int cols = rv.width, rows = rv.height;
        const int num_ngb = rv.num_src;
        const uint8_t *block = blocks ? blocks[i] : nullptr;
        std::vector<CostData> diff(num_ngb);
        for (int r = 0; r < rows; ++r) {
            for (int c = 0; c < cols; ++c) {
                const size_t rp = (size_t)r * cols + c;
                if (block && block[rp] < 128) continue;
                const float ref_depth = rv.depth[rp];
                if (ref_depth <= 0.0) continue;
                const float *rn = rv.normal + 3 * rp;
                const F3 PointX = FuseGet3DPointonWorld(c, r, ref_depth, rv.camera);
                for (int j = 0; j < num_ngb; ++j) {
                    const int si = rv.src_index[j];
                    const dpe_fuse_view &sv = views[si];
                    F2 pt;
                    float proj_depth;
                    ProjectonCamera_cu(PointX, sv.camera, pt, proj_depth);
                    const int src_r = fuse_round(pt.y);
                    const int src_c = fuse_round(pt.x);
                    if (src_c >= 0 && src_c < sv.width && src_r >= 0 && src_r < sv.height) {
                        const size_t sp = (size_t)src_r * sv.width + src_c;
                        if (masks[si][sp] == 1) continue;
                        const float src_depth = sv.depth[sp];
                        if (src_depth <= 0.0) continue;
                        const float *sn = sv.normal + 3 * sp;
                        const F3 tX = FuseGet3DPointonWorld(src_c, src_r, src_depth, sv.camera);
                        F2 tp2;
                        ProjectonCamera_cu(tX, rv.camera, tp2, proj_depth);
                        const float dx = (float)c - tp2.x, dy = (float)r - tp2.y;
                        diff[j].dist = (float)std::sqrt((double)dx * (double)dx + (double)dy * (double)dy);
                        diff[j].depth = std::fabs(proj_depth - ref_depth) / ref_depth;
                        const float angle = o_acosf(rn[0] * sn[0] + rn[1] * sn[1] + rn[2] * sn[2]);
                        diff[j].angle = angle != angle ? 0.0f : angle;  // GetAngle, DPE.cpp:1375-1385
                        diff[j].src_r = src_r;
                        diff[j].src_c = src_c;
                    }
                }
                for (int k = 2; k <= num_ngb; ++k) {
                    int n = 0;
                    for (int j = 0; j < num_ngb; ++j) {
                        diff[j].use = diff[j].dist < k * tp->dist_base && diff[j].depth < k * tp->depth_base &&
                                      (!inter || diff[j].angle < (k * tp->angle_grad + tp->angle_base));
                        n += diff[j].use;
                    }
                    if (n < k) continue;
                    float col[3] = {(float)rv.color[3 * rp], (float)rv.color[3 * rp + 1], (float)rv.color[3 * rp + 2]};
                    if (inter) {
                        for (int j = 0; j < num_ngb; ++j) {
                            if (!diff[j].use) continue;
                            const dpe_fuse_view &sv = views[rv.src_index[j]];
                            const size_t sp = (size_t)diff[j].src_r * sv.width + diff[j].src_c;
                            col[0] += (float)sv.color[3 * sp];
                            col[1] += (float)sv.color[3 * sp + 1];
                            col[2] += (float)sv.color[3 * sp + 2];
                        }
                        col[0] /= (n + 1.0f);
                        col[1] /= (n + 1.0f);
                        col[2] /= (n + 1.0f);
                    }
                    if (points && count < capacity) {
                        dpe_point &o = points[count];
                        o.x = PointX.x;
                        o.y = PointX.y;
                        o.z = PointX.z;
                        o.b = (uint8_t)col[0];
                        o.g = (uint8_t)col[1];
                        o.r = (uint8_t)col[2];
                        o.pad = 0;
                    }
                    count++;
                    masks[i][rp] = 1;
                    break;
                }
            }
        }
    }
    *num_points = count;
    return DPE_OK;
}

// RANSACToGetFitPlane (DPE.cu:2895-3128) of one WEAK pixel (px, py) on a given
// state (test access): camera-frame planes of every pixel (float4 n, d), the
// pixel's 9 neighbour slots (x, y; slot 0 the pixel itself, (-1, -1) none), all
// other pixels STRONG, use_limit / use_label / geom_consistency / use_radius
// off.  Writes the fitted plane (0 0 0 0: none) to out.
void oracle_fit_plane(const dpe_camera *cam, int W, int H, const float *planes, const int16_t *neighbours, int px,
                      int py, uint64_t seed, float *out) {
    Ctx C;
    C.W = W;
    C.H = H;
    C.P = W * H;
    C.N = 2;
    std::memset(&C.prm, 0, sizeof(C.prm));
    C.prm.seed = seed;
    C.cams[0] = *cam;
    C.planes.resize((size_t)C.P);
    for (int i = 0; i < C.P; ++i)
        C.planes[i] = F4{planes[4 * i], planes[4 * i + 1], planes[4 * i + 2], planes[4 * i + 3]};
    C.fit_planes.assign((size_t)C.P, F4{0, 0, 0, 0});
    const int center = px + py * W;
    C.weak.assign((size_t)C.P, DPE_STRONG);
    C.weak[center] = DPE_WEAK;
    C.nmap.assign((size_t)C.P, 0);
    C.neighbours.resize(NNUM);
    for (int k = 0; k < NNUM; ++k) C.neighbours[k] = S2{neighbours[2 * k], neighbours[2 * k + 1]};
    C.complex_map.assign((size_t)C.P, 0.0f);
    C.radius.assign((size_t)C.P, 5);
    RANSACToGetFitPlane(C, center, 0);
    const F4 f = C.fit_planes[center];
    out[0] = f.x;
    out[1] = f.y;
    out[2] = f.z;
    out[3] = f.w;
}

// GenEdgeInform (DPE.cu:2487-2595) over a whole image (test access): edge map
// (u8), label map (i32, or NULL = use_label off), pixel states; edge_neigh: 8
// (x, y) pairs per pixel, complex_map: per pixel, label_bound: 8 (x, y) pairs
// per pixel (written for WEAK pixels with a label > 0, else (-1, -1)).
void oracle_gen_edge_inform(int W, int H, const uint8_t *edge, const int32_t *label, const uint8_t *weak,
                            int strong_radius, int16_t *edge_neigh, float *complex_map, int16_t *label_bound) {
    Ctx C;
    C.W = W;
    C.H = H;
    C.P = W * H;
    std::memset(&C.prm, 0, sizeof(C.prm));
    C.prm.use_edge = 1;
    C.prm.use_label = label != nullptr;
    C.prm.strong_radius = strong_radius;
    C.edge = edge;
    C.label = label;
    C.weak.assign(weak, weak + C.P);
    C.edge_neigh.assign((size_t)C.P * 8, S2{-1, -1});
    C.complex_map.assign((size_t)C.P, 0.0f);
    C.nmap.assign((size_t)C.P, 0);
    for (int i = 0; i < C.P; ++i) C.nmap[i] = i;  // one label_bound row per pixel
    C.label_bound.assign((size_t)C.P * 8, S2{-1, -1});
    for (int i = 0; i < C.P; ++i) GenEdgeInform(C, i);
    for (size_t k = 0; k < (size_t)C.P * 8; ++k) {
        edge_neigh[2 * k] = C.edge_neigh[k].x;
        edge_neigh[2 * k + 1] = C.edge_neigh[k].y;
        label_bound[2 * k] = C.label_bound[k].x;
        label_bound[2 * k + 1] = C.label_bound[k].y;
    }
    for (int i = 0; i < C.P; ++i) complex_map[i] = C.complex_map[i];
}

// ---- test access to single functions of the restatement (each checked by
// tests/test_oracle_pins.py against closed forms, brute force or invariants)

static float geom_cost_impl(const dpe_camera *ref, const dpe_camera *src, int W, int H, const float *src_depth,
                            int px, int py, const float *plane, bool cost64) {
    Ctx C;
    C.cost64 = cost64;
    C.W = W;
    C.H = H;
    C.P = W * H;
    C.N = 2;
    C.cams[0] = *ref;
    C.cams[1] = *src;
    C.gdepth[0] = nullptr;
    C.gdepth[1] = src_depth;
    return ComputeGeomConsistencyCost(C, px, py, 1, F4{plane[0], plane[1], plane[2], plane[3]});
}
float oracle_geom_cost(const dpe_camera *ref, const dpe_camera *src, int W, int H, const float *src_depth, int px,
                       int py, const float *plane) {
    return geom_cost_impl(ref, src, W, H, src_depth, px, py, plane, false);
}
float oracle_geom_cost_fp64(const dpe_camera *ref, const dpe_camera *src, int W, int H, const float *src_depth, int px,
                            int py, const float *plane) {
    return geom_cost_impl(ref, src, W, H, src_depth, px, py, plane, true);
}

float oracle_initial_cost(const float *view_costs, int num_images, uint32_t *selected) {
    return InitialCostViews(num_images, selected, [&](int i) { return view_costs[i - 1]; });
}

void oracle_random_normal(const dpe_camera *cam, int px, int py, float depth, uint64_t seed, uint32_t pixel,
                          uint32_t phase, uint32_t call, float *out) {
    Rng rng(seed, pixel, phase, call);
    const F4 n = GenerateRandomNormal(*cam, px, py, rng, depth);
    out[0] = n.x;
    out[1] = n.y;
    out[2] = n.z;
    out[3] = n.w;
}

void oracle_perturbed_normal(const dpe_camera *cam, int px, int py, const float *normal, float perturbation,
                             uint64_t seed, uint32_t pixel, uint32_t phase, uint32_t call, float *out) {
    Rng rng(seed, pixel, phase, call);
    const F4 n = GeneratePerturbedNormal(*cam, px, py, F4{normal[0], normal[1], normal[2], normal[3]}, rng,
                                         perturbation);
    out[0] = n.x;
    out[1] = n.y;
    out[2] = n.z;
    out[3] = n.w;
}

void oracle_transform_normal(const dpe_camera *cam, const float *n, int to_world, float *out) {
    const F4 p{n[0], n[1], n[2], n[3]};
    const F4 r = to_world ? TransformNormal(*cam, p) : TransformNormal2RefCam(*cam, p);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

void oracle_depth_and_normal(const dpe_camera *cam, int px, int py, const float *plane, float *out) {
    Ctx C;
    C.W = px + 1;
    C.H = py + 1;
    C.P = C.W * C.H;
    C.cams[0] = *cam;
    C.planes.assign((size_t)C.P, F4{0, 0, 0, 0});
    const int center = px + py * C.W;
    C.planes[center] = F4{plane[0], plane[1], plane[2], plane[3]};
    GetDepthandNormal(C, center);
    const F4 r = C.planes[center];
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

int oracle_bresenham(int W, int H, const uint8_t *edge, int low_w, int low_h, const uint8_t *edge_low, int high_res,
                     int ax, int ay, int bx, int by, int64_t *trace, int cap, int *steps) {
    Ctx C;
    C.W = W;
    C.H = H;
    C.P = W * H;
    std::memset(&C.prm, 0, sizeof(C.prm));
    C.prm.high_res_img = high_res;
    C.edge = edge;
    C.edge_low = edge_low;
    C.low_w = low_w;
    C.low_h = low_h;
    std::vector<long> tr;
    tl_bres_trace = &tr;
    tl_bres_steps[0] = tl_bres_steps[1] = 0;
    const bool r = BresenhamLine(C, ax, ay, bx, by);
    tl_bres_trace = nullptr;
    for (size_t i = 0; i < tr.size() && (int)i < cap; ++i) trace[i] = tr[i];
    steps[0] = tl_bres_steps[0];
    steps[1] = tl_bres_steps[1];
    return r ? 1 : 0;
}

int oracle_direction_search(int W, int H, const uint8_t *weak, const int16_t *nearest, const uint8_t *edge,
                            const uint8_t *edge_low, int low_w, int low_h, int high_res, int rotate_time,
                            int edge_limit, int px, int py, const uint32_t *draws, int n_draws, int16_t *points,
                            uint8_t *valid) {
    Ctx C;
    C.W = W;
    C.H = H;
    C.P = W * H;
    std::memset(&C.prm, 0, sizeof(C.prm));
    C.prm.high_res_img = high_res;
    C.prm.rotate_time = rotate_time;
    set_gen_neighbours_constants(C);
    C.weak.assign(weak, weak + C.P);
    C.nearest.resize((size_t)C.P);
    for (int i = 0; i < C.P; ++i) C.nearest[i] = S2{nearest[2 * i], nearest[2 * i + 1]};
    C.edge = edge;
    C.edge_low = edge_low;
    C.low_w = low_w;
    C.low_h = low_h;
    Rng rng(0, 0, 0, 0);
    rng.replay = draws;
    rng.replay_n = (uint32_t)n_draws;
    S2 pts[64];
    bool ok[64];
    for (int i = 0; i < 64; ++i) {
        pts[i] = S2{-1, -1};
        ok[i] = false;
    }
    const int n = DirectionSearch(C, px, py, rng, edge_limit != 0, pts, ok);
    for (int i = 0; i < 64; ++i) {
        points[2 * i] = pts[i].x;
        points[2 * i + 1] = pts[i].y;
        valid[i] = ok[i] ? 1 : 0;
    }
    return n;
}

uint8_t oracle_depth_to_weak_state(const float *costs61, int weak_peak_radius) {
    return DepthToWeakState(costs61, weak_peak_radius);
}

// Numerics test access.
float oracle_expf(float x) { return o_expf(x); }
float oracle_acosf(float x) { return o_acosf(x); }
double oracle_exp(double x) { return o_exp(x); }
void oracle_sincosf(float x, float *s, float *c) { o_sincosf(x, s, c); }
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox4x32_10(ctr, key, out); }
float oracle_bilinear(const float *im, int W, int H, float u, float v) { return bilinear(im, W, H, u, v); }
void oracle_homography(const dpe_camera *ref, const dpe_camera *src, const float *plane, float *H) {
    ComputeHomography(*ref, *src, F4{plane[0], plane[1], plane[2], plane[3]}, H);
}
int oracle_point_in_triangle(const int16_t *A, const int16_t *B, const int16_t *Cc, int px, int py) {
    return PointinTriangle(S2{A[0], A[1]}, S2{B[0], B[1]}, S2{Cc[0], Cc[1]}, px, py) ? 1 : 0;
}
void oracle_sort_small(float *d, int n) { sort_small(d, n); }
void oracle_sort_small_weighted(int16_t *xy, float *w, int n) {
    std::vector<S2> pts(n);
    for (int i = 0; i < n; ++i) pts[i] = S2{xy[2 * i], xy[2 * i + 1]};
    sort_small_weighted(pts.data(), w, n);
    for (int i = 0; i < n; ++i) {
        xy[2 * i] = pts[i].x;
        xy[2 * i + 1] = pts[i].y;
    }
}
int oracle_find_min_cost_index(const float *costs, int n) { return FindMinCostIndex(costs, n); }
void oracle_pdf_to_cdf(float *probs, int n) { TransformPDFToCDF(probs, n); }
void oracle_nearest_strong(const uint8_t *weak, int W, int H, int16_t *out) {
    Ctx C;
    C.W = W;
    C.H = H;
    C.P = W * H;
    C.weak.assign(weak, weak + (size_t)W * H);
    C.nearest.assign((size_t)W * H, S2{-1, -1});
    for (int i = 0; i < W * H; ++i) FindNearestStrongPoint(C, i);
    for (int i = 0; i < W * H; ++i) {
        out[2 * i] = C.nearest[i].x;
        out[2 * i + 1] = C.nearest[i].y;
    }
}
float oracle_filter_strong(const float *depth, const float *costs, const uint8_t *weak, int W, int H, int px,
                           int py) {
    Ctx C;
    C.W = W;
    C.H = H;
    C.P = W * H;
    C.planes.assign((size_t)W * H, F4{0, 0, 0, 0});
    for (int i = 0; i < W * H; ++i) C.planes[i].w = depth[i];
    C.costs.assign(costs, costs + (size_t)W * H);
    C.weak.assign(weak, weak + (size_t)W * H);
    CheckerboardFilterStrong(C, px, py);
    return C.planes[(size_t)py * W + px].w;
}

}  // extern "C"
