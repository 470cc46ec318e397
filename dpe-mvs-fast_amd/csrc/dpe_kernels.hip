// dpe_kernels.hip -- gfx950 kernels of DPE-MVS's per-reference-image PatchMatch.
//
// Work decomposition (MI355X-first, not the reference's one-thread-per-pixel
// monolith):
//   * The bilateral NCC window (DPE.cu:696-782), the dominant cost, runs in two
//     kernels: ncc_full writes the per-view cost matrix of hypothesis slots,
//     ncc_weighted folds the views with the pixel's view weights.  A workgroup
//     is 64 consecutive pixels x 4 waves; the reference-only half of each window
//     (bilateral weights, weighted texels, their sums) is computed once per
//     pixel into LDS (dpe_ncc.h) and shared by the waves, which loop over slots.
//     Each tap is then one 16-byte quad load + projection.
//   * Control kernels (candidate search, view selection, acceptance) are
//     per-pixel and touch O(slots) memory; they never evaluate a window.
//   * The weak-texture path (GenNeighbours, RANSAC, weak propagation) keeps the
//     reference's per-pixel structure, on compacted lists of weak pixels.
// Every kernel cites the reference function it restates.
#include "dpe_ncc.h"

#include <hipcub/hipcub.hpp>

namespace dpe {

enum { MODE_REFINE = 0, MODE_POST = 1 };

// ------------------------------------------------------------------ images
__global__ void k_make_quads(const float *__restrict__ img, float4 *__restrict__ quads, int W, int H) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = (W + 1) * (H + 1);
    if (idx >= n) return;
    const int iy = idx / (W + 1) - 1;
    const int ix = idx % (W + 1) - 1;
    quads[idx] = make_float4(texel(img, W, H, ix, iy), texel(img, W, H, ix + 1, iy), texel(img, W, H, ix, iy + 1),
                             texel(img, W, H, ix + 1, iy + 1));
}

// fp32 -> fp16 (round to nearest even, as the conversion instruction does) or
// bf16 (round to nearest even on the upper 16 bits; NaN stays NaN)
__device__ __forceinline__ uint32_t to16(float v, int fmt) {
    if (fmt == QUAD_BF16) {
        const uint32_t u = __builtin_bit_cast(uint32_t, v);
        if (v != v) return (u >> 16) | 0x40u;
        return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
    }
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v);
}
__global__ void k_make_quads_h(const float *__restrict__ img, uint2 *__restrict__ quads, int W, int H, int fmt) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = (W + 1) * (H + 1);
    if (idx >= n) return;
    const int iy = idx / (W + 1) - 1;
    const int ix = idx % (W + 1) - 1;
    // column pairs: word 0 = (t00, t01), word 1 = (t10, t11), so one packed
    // subtraction gives both rows' x differences (sample_fast)
    quads[idx] = make_uint2(to16(texel(img, W, H, ix, iy), fmt) | (to16(texel(img, W, H, ix, iy + 1), fmt) << 16),
                            to16(texel(img, W, H, ix + 1, iy), fmt) | (to16(texel(img, W, H, ix + 1, iy + 1), fmt) << 16));
}
// Bit 0 of *flag: some texel is not fp16-exact.  Bit 1: some texel is off the
// quarter grid {k/4 : |k/4| < 256} (8-bit images and their factor-2 pyramid
// levels are on it), on which texel differences are exact in fp16 as well.
__global__ void k_check_fp16_exact(const float *__restrict__ img, size_t n, int *__restrict__ flag) {
    bool bad = false, off_grid = false;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = img[i];
        bad = bad || !((float)(_Float16)v == v);
        off_grid = off_grid || !(fabsf(v) < 256.0f && floorf(v * 4.0f) == v * 4.0f);
    }
    const int bits = (bad ? 1 : 0) | (off_grid ? 2 : 0);
    if (bits) atomicOr(flag, bits);
}

// ---- weak-pixel index map and lists (the input side of DPE.cpp:989-1030 and
// of the weak launches, DPE.cu:1868-1902), by stable selection on the device
struct WeakSelect {  // a WEAK pixel, of colour `colour` in rows < half_rows (colour -1: any)
    const uint8_t *weak;
    int W, half_rows, colour;
    __device__ __forceinline__ bool operator()(int i) const {
        if (i < 0 || weak[i] != DPE_WEAK) return false;
        if (colour < 0) return true;
        const int x = i % W, y = i / W;
        return y < half_rows && ((x + y) & 1) == colour;
    }
};
struct WeakFlag {
    const uint8_t *weak;
    __device__ __forceinline__ int operator()(int i) const { return weak[i] == DPE_WEAK ? 1 : 0; }
};
using PixelIter = hipcub::CountingInputIterator<int>;
using WeakFlagIter = hipcub::TransformInputIterator<int, WeakFlag, PixelIter>;
// The weak lists enumerate pixels tile by tile (kWeakTileW x kWeakTileH tiles
// in raster order, raster order inside a tile), so the 64 lanes of a wave hold
// a compact 2-D patch of weak pixels, whose GenNeighbours searches, points and
// windows overlap, rather than a run of one row (at 1600x1200: weak update
// -7..-9 % for 8x8 to 32x32 tiles, 16x16 kept; GenNeighbours and RANSAC
// together -80 ms per REFINE call over 20 views).
// Item i of the padded tile grid maps to its pixel index, -1 = padding.
#ifndef DPE_WEAK_TILE_W
#define DPE_WEAK_TILE_W 16
#endif
#ifndef DPE_WEAK_TILE_H
#define DPE_WEAK_TILE_H 16
#endif
constexpr int kWeakTileW = DPE_WEAK_TILE_W, kWeakTileH = DPE_WEAK_TILE_H;
struct TiledPixel {
    int W, H, tiles_x;
    __host__ __device__ __forceinline__ int operator()(int i) const {
        const int t = i / (kWeakTileW * kWeakTileH), r = i - t * (kWeakTileW * kWeakTileH);
        const int ty = t / tiles_x, tx = t - ty * tiles_x;
        const int x = tx * kWeakTileW + r % kWeakTileW, y = ty * kWeakTileH + r / kWeakTileW;
        return (x < W && y < H) ? x + y * W : -1;
    }
};
using TiledIter = hipcub::TransformInputIterator<int, TiledPixel, PixelIter>;
static inline int tiled_items(int W, int H) {
    return ((W + kWeakTileW - 1) / kWeakTileW) * ((H + kWeakTileH - 1) / kWeakTileH) * kWeakTileW * kWeakTileH;
}
__global__ void k_nmap_clear_others(const uint8_t *__restrict__ weak, int *__restrict__ nmap, int P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P && weak[i] != DPE_WEAK) nmap[i] = 0;
}

// Sums of one bilateral window computed directly (any radius/increment).
template <bool REFINED>
__device__ __forceinline__ float ncc_window_sums(const Dev &d, int v, const float H[9], int cx, int cy, float rc,
                                                 int radius, int inc, unsigned long long &taps) {
    const float ss = d.prm.sigma_spatial, sc = d.prm.sigma_color;
    float sum_ref = 0.0f, sum_ref_ref = 0.0f, sum_src = 0.0f, sum_src_src = 0.0f, sum_ref_src = 0.0f, wsum = 0.0f;
    for (int i = -radius; i <= radius; i += inc) {
        float r_ref = 0.0f, r_src = 0.0f, r_rr = 0.0f, r_ss = 0.0f, r_rs = 0.0f, r_w = 0.0f;
        for (int j = -radius; j <= radius; j += inc) {
            const int rx = cx + i, ry = cy + j;
            const float ref = texel(d.ref, d.W, d.H, rx, ry);
            const float2 sp = REFINED ? project_refined(H, rx, ry) : project(H, rx, ry);
            const float src = sample_src_any(d, v, sp.x, sp.y);
            const float fi = (float)i, fj = (float)j;
            const float spatial = sqrtf(fi * fi + fj * fj);
            const float color = fabsf(ref - rc);
            const float w = d_expf(-spatial / (2.0f * ss * ss) - color / (2.0f * sc * sc));
            const float wr = w * ref;
            r_ref = fmaf(w, ref, r_ref);
            r_rr = fmaf(wr, ref, r_rr);
            r_src = fmaf(w, src, r_src);
            r_ss = fmaf(w * src, src, r_ss);
            r_rs = fmaf(wr, src, r_rs);
            r_w = r_w + w;
            taps++;
        }
        sum_ref += r_ref;
        sum_ref_ref += r_rr;
        sum_src += r_src;
        sum_src_src += r_ss;
        sum_ref_src += r_rs;
        wsum += r_w;
    }
    return ncc_from_sums(sum_ref, sum_ref_ref, sum_src, sum_src_src, sum_ref_src, wsum);
}

// One bilateral window computed directly (any radius/increment): the shared
// body of ComputeBilateralNCCNew (DPE.cu:613-672) and ComputeBilateralNCCOld.
__device__ float ncc_window_direct(const Dev &d, int v, const float H[9], int cx, int cy, float rc, int radius,
                                   int inc, unsigned long long &taps) {
    const int last = -radius + ((2 * radius) / inc) * inc;  // offset of the last tap
    if (refined_division_ok(H, cx - radius, cx + last, cy - radius, cy + last))
        return ncc_window_sums<true>(d, v, H, cx, cy, rc, radius, inc, taps);
    return ncc_window_sums<false>(d, v, H, cx, cy, rc, radius, inc, taps);
}

// ---- weak-pixel windows with the reference half precomputed
// The windows of ComputeBilateralNCCNew (DPE.cu:613-672) at a weak pixel -- the
// centre window (adaptive radius) and up to 8 neighbour windows -- depend on
// neither the hypothesis nor the view.  k_weak_patch stores, per weak pixel of
// a colour list (entry i of n), the (w, w*ref) of every tap at
// wpre[(t0_k + t) * n + i] and the reference sums (sum_ref, sum_ref_ref, wsum)
// of window k at wsum[(3k + j) * n + i], computed with the same operations in
// the same order as ncc_window_sums; wsum[2n + i] < 0 marks a pixel whose
// centre window does not fit kWeakCentreTaps (it takes the direct path).
constexpr int kWeakCentreTaps = 36;
constexpr int kWeakNeighbourTaps = 9;
static_assert(kWeakCentreTaps + 8 * kWeakNeighbourTaps == kWeakPatchTaps, "weak patch layout");
__device__ __forceinline__ int weak_window_first_tap(int k) { return k == 0 ? 0 : kWeakCentreTaps + (k - 1) * kWeakNeighbourTaps; }

__device__ __forceinline__ void weak_window_shape(const Dev &d, int center, int k, int &radius, int &inc) {
    radius = (k == 0 ? d.prm.strong_radius : d.prm.weak_radius);
    inc = (k == 0 ? d.prm.strong_increment : d.prm.weak_increment);
    if (d.prm.use_radius && k == 0) {
        radius = d.radius[center];
        inc = max(2, (int)(2.0 * radius / 5.0));
    }
}

// Measured and not kept here: the clamp-free gather (18.0 -> 24.4 ms per weak
// launch) and unrolled 6x6 centre windows (3 waves per SIMD at 158 VGPRs, or 2
// with the 3x3 ones).  This kernel waits on its scattered neighbour-window
// gathers (about 6 TB/s counted fabric traffic, 4x its algorithmic bytes).
// NTW > 0: the window has NTW taps per side (2 radius / inc + 1 == NTW) and is
// unrolled, so that its reads are in flight together (the neighbour windows,
// 3x3: k_weak_update 19.4 -> 18.5 ms per launch at 1600x1200 at 3 waves per
// SIMD, where unrolling only the rows at 4 waves, or capping the unrolled
// kernel at 4 waves, measured slower).
template <bool REFINED, int FMT, int NTW = 0>
__device__ __forceinline__ float ncc_window_pre(const Dev &d, int v, const float H[9], int cx, int cy, int radius,
                                                int inc, const float *__restrict__ tw, int n,
                                                const float *__restrict__ ws, unsigned long long &taps) {
    float sum_src = 0.0f, sum_src_src = 0.0f, sum_ref_src = 0.0f;
    int t = 0;
    if constexpr (NTW > 0) {
#pragma unroll
        for (int ii = 0; ii < NTW; ++ii) {
            float r_src = 0.0f, r_ss = 0.0f, r_rs = 0.0f;
#pragma unroll
            for (int jj = 0; jj < NTW; ++jj) {
                const int rx = cx - radius + ii * inc, ry = cy - radius + jj * inc;
                const float2 sp = REFINED ? project_refined(H, rx, ry) : project(H, rx, ry);
                const float src = sample_src<FMT>(d, v, sp.x, sp.y);
                const float wt = tw[(size_t)(ii * NTW + jj) * n];
                const float2 e = make_float2(wt, wt * texel(d.ref, d.W, d.H, rx, ry));
                r_src = fmaf(e.x, src, r_src);
                r_ss = fmaf(e.x * src, src, r_ss);
                r_rs = fmaf(e.y, src, r_rs);
            }
            sum_src += r_src;
            sum_src_src += r_ss;
            sum_ref_src += r_rs;
        }
        taps += NTW * NTW;
        return ncc_from_sums(ws[0], ws[n], sum_src, sum_src_src, sum_ref_src, ws[2 * (size_t)n]);
    }
    for (int i = -radius; i <= radius; i += inc) {
        float r_src = 0.0f, r_ss = 0.0f, r_rs = 0.0f;
        for (int j = -radius; j <= radius; j += inc) {
            const int rx = cx + i, ry = cy + j;
            const float2 sp = REFINED ? project_refined(H, rx, ry) : project(H, rx, ry);
            const float src = sample_src<FMT>(d, v, sp.x, sp.y);
            const float wt = tw[(size_t)t * n];
            const float2 e = make_float2(wt, wt * texel(d.ref, d.W, d.H, rx, ry));
            r_src = fmaf(e.x, src, r_src);
            r_ss = fmaf(e.x * src, src, r_ss);
            r_rs = fmaf(e.y, src, r_rs);
            t++;
        }
        sum_src += r_src;
        sum_src_src += r_ss;
        sum_ref_src += r_rs;
    }
    taps += t;
    return ncc_from_sums(ws[0], ws[n], sum_src, sum_src_src, sum_ref_src, ws[2 * (size_t)n]);
}

// ncc_window_pre unrolled for 3x3 windows (the neighbour windows, radius 5 step 5)
template <bool REFINED, int FMT>
__device__ __forceinline__ float ncc_window_pre_unrolled(const Dev &d, int v, const float H[9], int cx, int cy, int radius,
                                                     int inc, const float *__restrict__ tw, int n,
                                                     const float *__restrict__ ws, unsigned long long &taps) {
    const int nt = 2 * radius / inc + 1;
    if (nt == 3) return ncc_window_pre<REFINED, FMT, 3>(d, v, H, cx, cy, radius, inc, tw, n, ws, taps);
    return ncc_window_pre<REFINED, FMT>(d, v, H, cx, cy, radius, inc, tw, n, ws, taps);
}

__device__ float ncc_window_pre_any(const Dev &d, int v, const float H[9], int cx, int cy, int radius, int inc,
                                    const float *__restrict__ tw, int n, const float *__restrict__ ws,
                                    unsigned long long &taps) {
    const int last = -radius + ((2 * radius) / inc) * inc;
    const bool refined = refined_division_ok(H, cx - radius, cx + last, cy - radius, cy + last);
    if (d.quad_fmt == QUAD_FP16) {
        if (refined) return ncc_window_pre_unrolled<true, QUAD_FP16>(d, v, H, cx, cy, radius, inc, tw, n, ws, taps);
        return ncc_window_pre<false, QUAD_FP16>(d, v, H, cx, cy, radius, inc, tw, n, ws, taps);
    }
    if (d.quad_fmt == QUAD_BF16) {
        if (refined) return ncc_window_pre<true, QUAD_BF16>(d, v, H, cx, cy, radius, inc, tw, n, ws, taps);
        return ncc_window_pre<false, QUAD_BF16>(d, v, H, cx, cy, radius, inc, tw, n, ws, taps);
    }
    if (refined) return ncc_window_pre<true, QUAD_FP32>(d, v, H, cx, cy, radius, inc, tw, n, ws, taps);
    return ncc_window_pre<false, QUAD_FP32>(d, v, H, cx, cy, radius, inc, tw, n, ws, taps);
}

__device__ float ncc_old_direct(const Dev &d, int x, int y, int v, const float4 &pl, unsigned long long &hyps,
                                unsigned long long &taps) {
    const ViewConst &vc = d.vc[v];
    float H[9];
    homography(vc, d.cams[0], pl, H);
    hyps++;
    const float2 pt = project(H, x, y);
    if (pt.x >= (float)vc.width || pt.x < 0.0f || pt.y >= (float)vc.height || pt.y < 0.0f) return 2.0f;
    const float rc = texel(d.ref, d.W, d.H, x, y);
    return ncc_window_direct(d, v, H, x, y, rc, d.prm.strong_radius, d.prm.strong_increment, taps);
}

__device__ __forceinline__ short2 neighbour_point(const Dev &d, int center, int k) {
    return d.neighbours[(size_t)d.nmap[center] * kNeighbourNum + k];
}

// ComputeBilateralNCCNew, DPE.cu:561-694 (weak centre pixels)
__device__ float ncc_new_direct(const Dev &d, int x, int y, int v, const float4 &pl, unsigned long long &hyps,
                                unsigned long long &taps) {
    const ViewConst &vc = d.vc[v];
    const int center = x + y * d.W;
    const float cost_max = 2.0f;
    float H[9];
    homography(vc, d.cams[0], pl, H);
    hyps++;
    const float2 pt = project(H, x, y);
    if (pt.x >= (float)vc.width || pt.x < 0.0f || pt.y >= (float)vc.height || pt.y < 0.0f) return cost_max;
    float cost = 0.0f;
    if (d.weak[center] == DPE_WEAK) {
        const float rc = texel(d.ref, d.W, d.H, x, y);
        float center_cost = 0.0f, strong_cost = 0.0f;
        int strong_count = 0;
        for (int k = 0; k < kNeighbourNum; ++k) {
            const short2 np = neighbour_point(d, center, k);
            if (np.x == -1 || np.y == -1) continue;
            const float2 nsp = project(H, np.x, np.y);
            if (nsp.x < 0 || nsp.y < 0 || nsp.x >= (float)d.W || nsp.y >= (float)d.H) {
                if (k != 0) {
                    const uint32_t vi = d.selected[np.x + np.y * d.W];
                    if (isSet(vi, v - 1)) {
                        strong_cost += cost_max;
                        strong_count++;
                    }
                    continue;
                } else {
                    return cost_max;
                }
            }
            int radius = (k == 0 ? d.prm.strong_radius : d.prm.weak_radius);
            int inc = (k == 0 ? d.prm.strong_increment : d.prm.weak_increment);
            if (d.prm.use_radius && k == 0) {
                radius = d.radius[center];
                inc = max(2, (int)(2.0 * radius / 5.0));
            }
            const float tc = ncc_window_direct(d, v, H, np.x, np.y, rc, radius, inc, taps);
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
            strong_cost = fminf(strong_cost, cost_max);
            cost = (float)(0.25 * center_cost + 0.75 * strong_cost);
        }
    }
    return cost;
}

// ComputeBilateralNCCNew with the precomputed windows of colour-list entry i
// (same result as ncc_new_direct).
__device__ float ncc_new_pre(const Dev &d, int x, int y, int v, const float4 &pl, int colour, int i,
                             unsigned long long &hyps, unsigned long long &taps) {
    const int n = d.weak_colour_count[colour];
    const float *ws = d.weak_wsum[colour] + i;
    if (ws[2 * (size_t)n] < 0.0f) return ncc_new_direct(d, x, y, v, pl, hyps, taps);
    const ViewConst &vc = d.vc[v];
    const int center = x + y * d.W;
    const float cost_max = 2.0f;
    float H[9];
    homography(vc, d.cams[0], pl, H);
    hyps++;
    const float2 pt = project(H, x, y);
    if (pt.x >= (float)vc.width || pt.x < 0.0f || pt.y >= (float)vc.height || pt.y < 0.0f) return cost_max;
    float center_cost = 0.0f, strong_cost = 0.0f;
    int strong_count = 0;
    const float *tw = d.weak_wpre[colour] + i;
    for (int k = 0; k < kNeighbourNum; ++k) {
        const short2 np = neighbour_point(d, center, k);
        if (np.x == -1 || np.y == -1) continue;
        const float2 nsp = project(H, np.x, np.y);
        if (nsp.x < 0 || nsp.y < 0 || nsp.x >= (float)d.W || nsp.y >= (float)d.H) {
            if (k != 0) {
                const uint32_t vi = d.selected[np.x + np.y * d.W];
                if (isSet(vi, v - 1)) {
                    strong_cost += cost_max;
                    strong_count++;
                }
                continue;
            } else {
                return cost_max;
            }
        }
        int radius, inc;
        weak_window_shape(d, center, k, radius, inc);
        const float tc = ncc_window_pre_any(d, v, H, np.x, np.y, radius, inc, tw + (size_t)weak_window_first_tap(k) * n,
                                            n, ws + (size_t)3 * k * n, taps);
        if (k == 0) {
            center_cost = tc;
        } else {
            strong_cost += tc;
            strong_count++;
        }
    }
    if (strong_count == 0) return center_cost;
    strong_cost /= strong_count;
    strong_cost = fminf(strong_cost, cost_max);
    return (float)(0.25 * center_cost + 0.75 * strong_cost);
}

// Reference halves of the weak windows of every entry of a colour list
// (the layout described at ncc_window_pre).
__global__ __launch_bounds__(256) void k_weak_patch(Dev d, int colour) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = d.weak_colour_count[colour];
    if (i >= n) return;
    const int p = d.weak_colour_list[colour][i];
    const int x = p % d.W, y = p / d.W;
    float *tw = d.weak_wpre[colour] + i;
    float *ws = d.weak_wsum[colour] + i;
    if (d.weak[p] != DPE_WEAK) return;
    const float ss = d.prm.sigma_spatial, sc = d.prm.sigma_color;
    const float rc = texel(d.ref, d.W, d.H, x, y);
    for (int k = 0; k < kNeighbourNum; ++k) {
        const short2 np = neighbour_point(d, p, k);
        if (np.x == -1 || np.y == -1) continue;
        int radius, inc;
        weak_window_shape(d, p, k, radius, inc);
        const int side = (2 * radius) / inc + 1;
        if (side * side > (k == 0 ? kWeakCentreTaps : kWeakNeighbourTaps)) {
            ws[2 * (size_t)n] = -1.0f;  // centre window or neighbour windows too large: direct path
            return;
        }
        int t = weak_window_first_tap(k);
        float sum_ref = 0.0f, sum_ref_ref = 0.0f, wsum = 0.0f;
        for (int a = -radius; a <= radius; a += inc) {
            float r_ref = 0.0f, r_rr = 0.0f, r_w = 0.0f;
            for (int b = -radius; b <= radius; b += inc) {
                const float ref = texel(d.ref, d.W, d.H, np.x + a, np.y + b);
                const float fi = (float)a, fj = (float)b;
                const float spatial = sqrtf(fi * fi + fj * fj);
                const float color = fabsf(ref - rc);
                const float w = d_expf(-spatial / (2.0f * ss * ss) - color / (2.0f * sc * sc));
                const float wr = w * ref;
                r_ref = fmaf(w, ref, r_ref);
                r_rr = fmaf(wr, ref, r_rr);
                r_w = r_w + w;
                tw[(size_t)t * n] = w;
                t++;
            }
            sum_ref += r_ref;
            sum_ref_ref += r_rr;
            wsum += r_w;
        }
        ws[(size_t)(3 * k) * n] = sum_ref;
        ws[(size_t)(3 * k + 1) * n] = sum_ref_ref;
        if (k != 0) ws[(size_t)(3 * k + 2) * n] = wsum;
        else ws[2 * (size_t)n] = wsum;
    }
}

// Get3DPointonWorld_cu / ProjectonCamera_cu, DPE.cu:885-917
__device__ __forceinline__ float3 point_world(float x, float y, float depth, const dpe_camera &cam) {
    float3 X, T;
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
__device__ __forceinline__ void project_camera(const float3 &X, const dpe_camera &cam, float2 &pt, float &depth) {
    float3 t;
    t.x = cam.R[0] * X.x + cam.R[1] * X.y + cam.R[2] * X.z + cam.t[0];
    t.y = cam.R[3] * X.x + cam.R[4] * X.y + cam.R[5] * X.z + cam.t[1];
    t.z = cam.R[6] * X.x + cam.R[7] * X.y + cam.R[8] * X.z + cam.t[2];
    depth = cam.K[6] * t.x + cam.K[7] * t.y + cam.K[8] * t.z;
    pt.x = (cam.K[0] * t.x + cam.K[1] * t.y + cam.K[2] * t.z) / depth;
    pt.y = (cam.K[3] * t.x + cam.K[4] * t.y + cam.K[5] * t.z) / depth;
}

// ComputeGeomConsistencyCost, DPE.cu:919-957
__device__ float geom_cost(const Dev &d, int x, int y, int v, const float4 &pl) {
    const dpe_camera &ref = d.cams[0];
    const dpe_camera &src = d.cams[v];
    const float depth = depth_from_plane(ref, pl, x, y);
    const float3 fwd = point_world((float)x, (float)y, depth, ref);
    float2 sp;
    float sd;
    project_camera(fwd, src, sp, sd);
    const int ix = f2i_clamp(sp.x, -1, d.W), iy = f2i_clamp(sp.y, -1, d.H);
    const float src_depth = texel(d.gdepth + (size_t)v * d.P, d.W, d.H, ix, iy);
    if (src_depth == 0.0f) return 3.0f;
    const float3 s3 = point_world(sp.x, sp.y, src_depth, src);
    float2 bp;
    float rd;
    project_camera(s3, ref, bp, rd);
    const float dc = (float)x - bp.x, dr = (float)y - bp.y;
    return fminf(3.0f, sqrtf(dc * dc + dr * dr));
}

// ---------------------------------------------------- NCC cost kernels (hot)
// The wave's lanes see each other's LDS writes made before the call.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Set bits of the wave mask b in the lanes below this one.
__device__ __forceinline__ int lanes_below(unsigned long long b) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
}
// Bitwise equality of two planes.
__device__ __forceinline__ bool same_plane(const float4 &a, const float4 &b) {
    return __float_as_uint(a.x) == __float_as_uint(b.x) && __float_as_uint(a.y) == __float_as_uint(b.y) &&
           __float_as_uint(a.z) == __float_as_uint(b.z) && __float_as_uint(a.w) == __float_as_uint(b.w);
}
// A candidate row (the NCC costs of one plane at one pixel over the views) is a
// pure function of (pixel, plane, view).  last[slot * Qn + q] holds the plane of
// the row that slot `slot` of pixel q last wrote into its colour's cost matrix
// in this call (all-ones bits: none), so a candidate repeating it is not
// evaluated: its row is still there.  Returns true for such a repeat and
// records pl otherwise.
__device__ __forceinline__ bool repeated_row(float4 *__restrict__ last, size_t at, const float4 &pl) {
    const float4 l = last[at];
    const bool none = (__float_as_uint(l.x) & __float_as_uint(l.y) & __float_as_uint(l.z) & __float_as_uint(l.w)) ==
                      0xFFFFFFFFu;
    if (!none && same_plane(l, pl)) return true;
    last[at] = pl;
    return false;
}

// A strong candidate whose plane repeats another row of its pixel's
// half-iteration (cand_alias, k_strong_candidates) is read from that row and
// not evaluated; its own row is then stale, so its last-plane record is
// cleared (all-ones bits: none).  Returns true for such a slot.
__device__ __forceinline__ bool aliased_row(const Dev &d, float4 *__restrict__ last, int slot, int Qn, int q) {
    if (d.cand_alias[(size_t)slot * Qn + q] == slot) return false;
    const float none = __uint_as_float(0xFFFFFFFFu);
    last[(size_t)slot * Qn + q] = make_float4(none, none, none, none);
    return true;
}

// Per-view cost matrix out[(slot*NV + v-1)*Qn + q] of the plane of every
// (pixel, slot): plane_src[cand_pos[slot*Qn+q]] if cand_pos is given (a -1
// entry skips the slot), else plane_src[slot*Qn + q].  With view_mask (indexed
// by pixel, Qn == P) only the views whose bit is set are evaluated.  With
// `cached` (the strong launches; slot 0 = the pixel's current plane) the views
// of cur_valid take their cost from cur_cost instead of a window: the same
// plane was evaluated at this pixel when it was set, so the value is the same.
// With `last` (the strong launches), candidate slots 1.. whose plane repeats
// the slot's last evaluated one keep their row (repeated_row).
// One workgroup = 64 pixels x WAVES waves; wave w takes slots w, w+WAVES, ...
// The wave's (slot, pixel) items that still need windows are packed onto its
// lanes (pixels whose slot has no candidate, or repeats its row, leave no idle
// lane), 64 items per pass; a pass loops over the views, the same view in
// every lane, so the view's constants and quad base stay wave-uniform.
template <int NT, bool PREC, int WAVES>
__global__ __launch_bounds__(kLanes * WAVES) void k_ncc_full(Dev d, int colour, int Qn, int nslot,
                                                              const int *__restrict__ cand_pos,
                                                              const float4 *__restrict__ plane_src,
                                                              const uint32_t *__restrict__ view_mask,
                                                              float *__restrict__ out, bool cached,
                                                              float4 *__restrict__ last) {
    __shared__ SharedPatch<NT> S;
    extern __shared__ float2 full_lists[];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int q = blockIdx.x * kLanes + lane;
    int x = 0, y = 0;
    const bool valid = q < Qn && colour_pixel(d, colour, q, x, y);
    bool any_slot = false;
    if (valid && cand_pos) {
        for (int s = 0; s < nslot; ++s) any_slot = any_slot || cand_pos[(size_t)s * Qn + q] >= 0;
    } else {
        any_slot = valid;
    }
    fill_shared_patch<NT, WAVES>(d, any_slot, x, y, S);
    unsigned long long hyps = 0, taps = 0;
    const int NV = d.N - 1;
    const uint32_t all = NV >= 32 ? 0xFFFFFFFFu : (1u << NV) - 1u;
    const uint32_t mask = (valid && view_mask ? view_mask[q] : 0xFFFFFFFFu) & all;
    const int p = x + y * d.W;
    // the wave's item list: (slot index within the wave << 6) | pixel lane; without
    // cand_pos every valid pixel of every slot is an item and the list is implicit,
    // and with fewer slots than waves every wave takes every slot and the views
    // v with (v - 1) % WAVES == wave instead
    const bool dense = cand_pos == nullptr;
    const bool split = dense && nslot < WAVES;
    uint32_t wave_views = 0xFFFFFFFFu;
    if (split) {
        wave_views = 0u;
        for (int v = wave; v < 32; v += WAVES) wave_views |= 1u << v;
    }
    uint16_t *items = (uint16_t *)((char *)full_lists + (size_t)wave * full_list_bytes(nslot, WAVES));
    int total = dense ? (split ? nslot : (nslot - wave + WAVES - 1) / WAVES) * kLanes : 0;
    for (int k = 0, slot = wave; !dense && slot < nslot; ++k, slot += WAVES) {
        bool need = false;
        if (any_slot) {
            const int pos = cand_pos ? cand_pos[(size_t)slot * Qn + q] : 0;
            if (pos >= 0) {
                const float4 pl = cand_pos ? plane_src[pos] : plane_src[(size_t)slot * Qn + q];
                need = !(last && slot > 0 && repeated_row(last, (size_t)slot * Qn + q, pl));
                if (need && last && slot > 0) need = !aliased_row(d, last, slot, Qn, q);
                if (need && cached && slot == 0) {
                    const uint32_t known = mask & d.cur_valid[p];
                    for (uint32_t b = known; b; b &= b - 1) {
                        const int v = __builtin_ctz(b);
                        out[((size_t)slot * NV + v) * Qn + q] = d.cur_cost[(size_t)v * d.P + p];
                    }
                    need = (mask & ~known) != 0;
                }
            }
        }
        const unsigned long long b = __ballot(need);
        if (need) items[total + lanes_below(b)] = (uint16_t)((k << 6) | lane);
        total += __popcll(b);
    }
    // With item lists, the workgroup's lists are dealt to its waves 64 items at a
    // time (item i of the concatenation to wave (i / 64) % WAVES): a wave working
    // through its own list would end on a partial pass, and the waves' lists
    // differ in length (wave 0 takes one slot more when WAVES does not divide
    // the slot count).  All waves hold the same 64 pixels, so a wave reads an
    // item's pixel from its own lane `src` as before.
    __shared__ int list_first[WAVES + 1];
    int n_items = total;
    if (!dense) {
        if (lane == 0) list_first[wave + 1] = total;
        __syncthreads();
        if (threadIdx.x == 0) {
            list_first[0] = 0;
            for (int w = 1; w <= WAVES; ++w) list_first[w] += list_first[w - 1];
        }
        __syncthreads();
        n_items = list_first[WAVES];
    } else {
        wave_sync();
    }
    for (int base = dense ? 0 : wave * kLanes; base < n_items; base += dense ? kLanes : WAVES * kLanes) {
        const int i = base + lane;
        bool act;
        uint32_t e;
        int sw = wave;  // the wave whose list holds the item
        if (dense) {
            act = valid;
            e = (uint32_t)i;
        } else {
            act = i < n_items;
            sw = 0;
            for (int w = 1; w < WAVES; ++w)
                if (i >= list_first[w]) sw = w;
            const uint16_t *lst = (const uint16_t *)((char *)full_lists + (size_t)sw * full_list_bytes(nslot, WAVES));
            e = act ? lst[i - list_first[sw]] : 0u;
        }
        const int src = (int)(e & 63u), slot = split ? (int)(e >> 6) : sw + WAVES * (int)(e >> 6);
        const int ix = __shfl(x, src, kLanes), iy = __shfl(y, src, kLanes);
        const int iq = blockIdx.x * kLanes + src;
        uint32_t views = __shfl(mask, src, kLanes) & wave_views;
        float4 pl = make_float4(0, 0, 0, 0);
        if (act) {
            pl = cand_pos ? plane_src[cand_pos[(size_t)slot * Qn + iq]] : plane_src[(size_t)slot * Qn + iq];
            if (cached && slot == 0) views &= ~d.cur_valid[ix + iy * d.W];
        } else {
            views = 0;
        }
        for (int v = 1; v < d.N; ++v) {
            if (!isSet(views, v - 1)) continue;
            out[((size_t)slot * NV + (v - 1)) * Qn + iq] = ncc_shared<NT, PREC>(d, S, src, ix, iy, v, pl, taps);
            hyps++;
        }
    }
    count_add(d.counters, kFamFull, hyps, taps);
}

// Generic window (any radius/increment): one (pixel, slot) per thread, direct.
__global__ __launch_bounds__(256) void k_ncc_full_direct(Dev d, int colour, int Qn, const int *__restrict__ cand_pos,
                                                         const float4 *__restrict__ plane_src,
                                                         const uint32_t *__restrict__ view_mask,
                                                         float *__restrict__ out, float4 *__restrict__ last) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = blockIdx.y;
    int x = 0, y = 0;
    bool active = q < Qn && colour_pixel(d, colour, q, x, y);
    float4 pl = make_float4(0, 0, 0, 0);
    if (active) {
        if (cand_pos) {
            const int pos = cand_pos[(size_t)slot * Qn + q];
            active = pos >= 0;
            if (active) pl = plane_src[pos];
        } else {
            pl = plane_src[(size_t)slot * Qn + q];
        }
        // a candidate repeating its slot's last plane keeps its row, one repeating
        // another row of the pixel reads that one (k_ncc_full)
        if (active && last && slot > 0) active = !repeated_row(last, (size_t)slot * Qn + q, pl);
        if (active && last && slot > 0) active = !aliased_row(d, last, slot, Qn, q);
    }
    unsigned long long hyps = 0, taps = 0;
    if (active) {
        const uint32_t mask = view_mask ? view_mask[q] : 0xFFFFFFFFu;
        for (int v = 1; v < d.N; ++v) {
            if (!isSet(mask, v - 1)) continue;
            unsigned long long h = 0;
            out[((size_t)slot * d.NV + (v - 1)) * Qn + q] = ncc_old_direct(d, x, y, v, pl, h, taps);
            hyps += h;
        }
    }
    count_add(d.counters, kFamFull, hyps, taps);
}

// Plane and weight set-up of one weighted slot; returns false if the slot is
// not evaluated.  REFINE reads the refinement hypotheses; POST reads the sample
// planes k_post_planes materialised (w = NaN: sample outside the depth range,
// its cost already written).
template <int MODE>
__device__ __forceinline__ bool weighted_slot(const Dev &d, int Qn, int q, int p, int slot, float4 &pl, float &wn,
                                              uint32_t &sel) {
    if (MODE == MODE_REFINE) {
        if (d.cand_pos[q] < 0) return false;
        pl = d.rplane[(size_t)slot * Qn + q];
        wn = d.wnorm[q];
        return true;
    }
    const uint8_t fl = d.post_flags[p];
    const bool dtw = slot < 61;
    if (!((dtw && (fl & 1)) || (!dtw && (fl & 2)))) return false;
    pl = d.post_planes[(size_t)slot * d.P + p];
    if (pl.w != pl.w) return false;
    wn = d.post_wn[p];
    sel = d.selected[p];
    return true;
}

// Post-pass samples whose costs the peak analysis never reads: DepthToWeak's
// d = +30 (its scan stops at d = +29, DPE.cu:2700-2712) and d = -30 unless
// weak_peak_radius >= 30 (read only as the minimum when there is no peak,
// which the radius test already calls WEAK otherwise).
__device__ __forceinline__ bool post_sample_unread(const Dev &d, int slot) {
    return slot == 60 || (slot == 0 && d.prm.weak_peak_radius < 30);
}
// LocalRefine's samples d = -5..5 (slots 62..72) use the planes of DepthToWeak's
// d = -5..5 (slots 25..35): same normal, same depth expression.  Where a pixel
// runs both steps the DepthToWeak wave evaluates each view once and folds it
// in both steps' orders (kPostTwin = 37 slots apart).
constexpr int kPostTwin = 37;
__device__ __forceinline__ bool post_has_twin(int slot, uint8_t fl) { return slot >= 25 && slot <= 35 && (fl & 2); }
__device__ __forceinline__ bool post_is_twin(int slot, uint8_t fl) { return slot >= 62 && (fl & 1); }

// Weighted sums over views.
//   MODE_REFINE: slot s of the refinement hypotheses (PlaneHypothesisRefinementStrong,
//     DPE.cu:1101-1113): sum over views with weight > 0 of w*cost, / weight_norm.
//   MODE_POST: DepthToWeak samples (DPE.cu:2667-2690) and LocalRefine
//     (DPE.cu:2776-2835), all over the pixel's selected views.
//   With `twin` (MODE_POST, DepthToWeak slots 25..35) the LocalRefine sum of
//   the same plane goes to *twin as well.
template <int NT, int MODE, bool PREC>
__device__ __forceinline__ float weighted_sum(const Dev &d, const void *Sv, int lane, int x, int y,
                                              int slot, const uint8_t *vw, uint32_t sel, const float4 &pl, float wn,
                                              unsigned long long &hyps, unsigned long long &taps,
                                              float *twin = nullptr) {
    float acc = 0.0f, acc2 = 0.0f;
    const bool geom = d.prm.geom_consistency != 0;
    const float gf = d.prm.geom_factor;
    for (int v = 1; v < d.N; ++v) {
        const float wv = (float)vw[v - 1];
        const bool use = (MODE == MODE_REFINE) ? (vw[v - 1] > 0) : (isSet(sel, v - 1) != 0);
        if (!use) continue;
        float c;
        if constexpr (NT > 0) {
            c = ncc_shared<NT, PREC>(d, *(const SharedPatch<NT> *)Sv, lane, x, y, v, pl, taps);
            hyps++;
        } else {
            c = ncc_old_direct(d, x, y, v, pl, hyps, taps);
        }
        if (MODE == MODE_REFINE) {
            acc += wv * c;
        } else if (slot <= 61) {
            const float g = geom ? geom_cost(d, x, y, v, pl) : 0.0f;
            float tc = c;
            if (geom) tc += gf * g;
            acc += tc * wv;
            if (twin) {  // LocalRefine's order (DPE.cu:2822-2825)
                acc2 += c * wv;
                if (geom) acc2 += gf * g * wv;
            }
        } else {
            acc += c * wv;
            if (geom) acc += gf * geom_cost(d, x, y, v, pl) * wv;
        }
    }
    acc /= wn;
    if (twin) *twin = acc2 / wn;
    if (MODE == MODE_POST && slot < 61) acc = fminf(2.0f, acc);
    return acc;
}


// One workgroup = 64 pixels x WAVES waves; wave w evaluates slot WAVES*blockIdx.y + w.
// The windows of a wave are its (pixel, view) pairs, packed densely: a pixel
// sums over its own views only (the weighted ones, or the selected ones), and
// the 64 pixels of a wave select different subsets (at 1600x1200 a pixel uses
// ~6.6 of 10 views while the 64 use ~9.5 between them), so a loop over the
// views with idle lanes runs ~9.5 windows per lane, the packed list ~7.  The
// list is view-major (view 1's pixels in lane order, then view 2's, ...), so
// the lanes of a round gather from one or two source views at neighbouring
// pixels, as the per-view loop does (a pixel-major list spreads a round over
// all views and touches ~2x the cache lines).  Round r evaluates pairs
// 64r..64r+63, one per lane, the pixel's reference patch read from its LDS
// column; then each pixel's lane folds its pairs of the round in view order,
// the order of the sums above.
// Dynamic LDS per wave (wtd_list_bytes): the round's 64 (cost, geometric
// cost) results, per view the ballot of its pixels and its first list index,
// and the list, entry = (pixel lane << 5) | (view - 1).
template <int NT, int MODE, bool PREC, int WAVES>
__global__ __launch_bounds__(kLanes * WAVES) void k_ncc_weighted(Dev d, int colour, int Qn, int nslot,
                                                                  float *__restrict__ out) {
    __shared__ SharedPatch<NT> S;
    extern __shared__ float2 wtd_lists[];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int q = blockIdx.x * kLanes + lane;
    const int slot = blockIdx.y * WAVES + wave;
    int x = 0, y = 0;
    const bool valid = q < Qn && colour_pixel(d, colour, q, x, y);
    const int p = x + y * d.W;
    bool any = false;
    if (valid) any = MODE == MODE_REFINE ? d.cand_pos[q] >= 0 : d.post_flags[p] != 0;
    fill_shared_patch<NT, WAVES>(d, any, x, y, S);
    unsigned long long hyps = 0, taps = 0;
    float4 pl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float wn = 1.0f;
    uint32_t sel = 0;
    const uint8_t *vw = d.vw + (size_t)p * kMaxImages;
    // views: bit v - 1 = view v enters this pixel's sum
    uint32_t views = 0;
    bool act = false, twin = false;
    if (any && slot < nslot && weighted_slot<MODE>(d, Qn, q, p, slot, pl, wn, sel)) {
        const uint8_t fl = MODE == MODE_POST ? d.post_flags[p] : 0;
        if (!(MODE == MODE_POST && post_is_twin(slot, fl))) {
            act = true;
            twin = MODE == MODE_POST && post_has_twin(slot, fl);
            for (int v = 1; v < d.N; ++v) {
                const bool use = (MODE == MODE_REFINE) ? (vw[v - 1] > 0) : (isSet(sel, v - 1) != 0);
                if (use) views |= 1u << (v - 1);
            }
        }
    }
    const int NV = d.N - 1;
    char *lds = (char *)wtd_lists + (size_t)wave * wtd_wave_bytes(d.N);
    float2 *res = (float2 *)lds;
    unsigned long long *vball = (unsigned long long *)(lds + kLanes * sizeof(float2));
    uint16_t *vstart = (uint16_t *)(vball + NV);
    uint16_t *pairs = vstart + NV;
    int total = 0;
    for (int v = 0; v < NV; ++v) {
        const bool mine = (views >> v) & 1u;
        const unsigned long long b = __ballot(mine);
        if (mine) pairs[total + lanes_below(b)] = (uint16_t)((lane << 5) | v);
        if (lane == 0) {
            vball[v] = b;
            vstart[v] = (uint16_t)total;
        }
        total += __popcll(b);
    }
    wave_sync();
    const bool geom = d.prm.geom_consistency != 0;
    const float gf = d.prm.geom_factor;
    float acc = 0.0f, acc2 = 0.0f;
    uint32_t left = views;  // views not folded yet
    for (int base = 0; base < total; base += kLanes) {
        const int i = base + lane;
        const uint32_t e = i < total ? pairs[i] : 0u;
        const int src = (int)(e >> 5), v = (int)(e & 31u) + 1;
        const int ix = __shfl(x, src, kLanes), iy = __shfl(y, src, kLanes);
        const float4 ipl = make_float4(__shfl(pl.x, src, kLanes), __shfl(pl.y, src, kLanes),
                                       __shfl(pl.z, src, kLanes), __shfl(pl.w, src, kLanes));
        float c = 0.0f, g = 0.0f;
        if (i < total) {
            c = ncc_shared<NT, PREC>(d, S, src, ix, iy, v, ipl, taps);
            hyps++;
            if (MODE == MODE_POST && geom) g = geom_cost(d, ix, iy, v, ipl);
        }
        res[lane] = make_float2(c, g);
        wave_sync();
        // this lane's pairs of the round: its next views while their list index is below base + 64
        while (left) {
            const int vi = __builtin_ctz(left);
            const int at = (int)vstart[vi] + lanes_below(vball[vi]) - base;
            if (at >= kLanes) break;
            left &= left - 1;
            const float2 r = res[at];
            const float wv = (float)vw[vi];
            if (MODE == MODE_REFINE) {
                acc += wv * r.x;
            } else if (slot <= 61) {
                float tc = r.x;
                if (geom) tc += gf * r.y;
                acc += tc * wv;
                if (twin) {  // LocalRefine's order (DPE.cu:2822-2825)
                    acc2 += r.x * wv;
                    if (geom) acc2 += gf * r.y * wv;
                }
            } else {
                acc += r.x * wv;
                if (geom) acc += gf * r.y * wv;
            }
        }
        wave_sync();
    }
    if (act) {
        acc /= wn;
        if (MODE == MODE_POST && slot < 61) out[(size_t)slot * Qn + q] = fminf(2.0f, acc);
        else out[(size_t)slot * Qn + q] = acc;
        if (twin) out[(size_t)(slot + kPostTwin) * Qn + q] = acc2 / wn;
    }
    count_add(d.counters, (MODE == MODE_REFINE ? kFamRefine : kFamPost), hyps, taps);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_ncc_weighted_direct(Dev d, int colour, int Qn, float *__restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = blockIdx.y;
    int x = 0, y = 0;
    bool active = q < Qn && colour_pixel(d, colour, q, x, y);
    const int p = x + y * d.W;
    unsigned long long hyps = 0, taps = 0;
    float4 pl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float wn = 1.0f;
    uint32_t sel = 0;
    if (active) active = weighted_slot<MODE>(d, Qn, q, p, slot, pl, wn, sel);
    if (active) {
        const uint8_t *vw = d.vw + (size_t)p * kMaxImages;
        const uint8_t fl = MODE == MODE_POST ? d.post_flags[p] : 0;
        if (MODE == MODE_POST && post_has_twin(slot, fl)) {
            float tw;
            out[(size_t)slot * Qn + q] =
                weighted_sum<0, MODE, false>(d, nullptr, 0, x, y, slot, vw, sel, pl, wn, hyps, taps, &tw);
            out[(size_t)(slot + kPostTwin) * Qn + q] = tw;
        } else if (!(MODE == MODE_POST && post_is_twin(slot, fl))) {
            out[(size_t)slot * Qn + q] = weighted_sum<0, MODE, false>(d, nullptr, 0, x, y, slot, vw, sel, pl, wn, hyps, taps);
        }
    }
    count_add(d.counters, (MODE == MODE_REFINE ? kFamRefine : kFamPost), hyps, taps);
}

// ------------------------------------------------------ random initialisation
// RandomInitialization, DPE.cu:1039-1067 (plane part)
__device__ float4 random_normal(const dpe_camera &cam, int x, int y, Rng &rng, float depth) {
    float4 n;
    float q1 = 1.0f, q2 = 1.0f, s = 2.0f;
    while (s >= 1.0f) {
        q1 = 2.0f * rng.uniform() - 1.0f;
        q2 = 2.0f * rng.uniform() - 1.0f;
        s = q1 * q1 + q2 * q2;
    }
    const float sq = sqrtf(1.0f - s);
    n.x = 2.0f * q1 * sq;
    n.y = 2.0f * q2 * sq;
    n.z = 1.0f - 2.0f * s;
    n.w = 0;
    const float4 vd = view_direction(cam, x, y, depth);
    const float dot = n.x * vd.x + n.y * vd.y + n.z * vd.z;
    if (dot > 0.0f) {
        n.x = -n.x;
        n.y = -n.y;
        n.z = -n.z;
    }
    normalize3(&n);
    return n;
}

// GeneratePerturbedNormal, DPE.cu:393-428
__device__ float4 perturbed_normal(const dpe_camera &cam, int x, int y, const float4 &normal, Rng &rng,
                                   float perturbation) {
    const float4 vd = view_direction(cam, x, y, 1.0f);
    const float a1 = (rng.uniform() - 0.5f) * perturbation;
    const float a2 = (rng.uniform() - 0.5f) * perturbation;
    const float a3 = (rng.uniform() - 0.5f) * perturbation;
    float s1, c1, s2, c2, s3, c3;
    d_sincosf(a1, &s1, &c1);
    d_sincosf(a2, &s2, &c2);
    d_sincosf(a3, &s3, &c3);
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
    float4 np;
    np.x = R[0] * normal.x + R[1] * normal.y + R[2] * normal.z;
    np.y = R[3] * normal.x + R[4] * normal.y + R[5] * normal.z;
    np.z = R[6] * normal.x + R[7] * normal.y + R[8] * normal.z;
    np.w = normal.w;
    if (np.x * vd.x + np.y * vd.y + np.z * vd.z >= 0.0f) np = normal;
    normalize3(&np);
    return np;
}

__global__ void k_random_init_planes(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    const int x = p % d.W, y = p / d.W;
    const dpe_camera &cam = d.cams[0];
    if (d.prm.state == DPE_FIRST_INIT) {
        Rng rng(d.prm.seed, p, PH_RANDOM_INIT, d.prm.call_index);
        const float depth = rng.uniform() * (d.prm.depth_max - d.prm.depth_min) + d.prm.depth_min;
        const float4 n = random_normal(cam, x, y, rng, depth);
        d.planes[p] = make_float4(n.x, n.y, n.z, distance_to_origin(cam, x, y, depth, n));
    } else {
        const float4 n = to_camera(cam, d.planes[p]);
        d.planes[p] = make_float4(n.x, n.y, n.z, distance_to_origin(cam, x, y, n.w, n));
    }
}

// ComputeMultiViewInitialCostandSelectedViews (DPE.cu:784-830) /
// ComputeMultiViewInitialCost (DPE.cu:832-861) from the current planes' costs
// (cur_cost, every view in FIRST_INIT, the prior-selected ones otherwise).
__global__ void k_random_init_costs(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    const int NV = d.NV;
    const float cost_max = 2.0f;
    const float *cm = d.cur_cost;  // [NV][P]
    const uint32_t all = NV >= 32 ? 0xFFFFFFFFu : (1u << NV) - 1u;
    d.cur_valid[p] = d.prm.state == DPE_FIRST_INIT ? all : (d.selected[p] & all);
    if (d.prm.state == DPE_FIRST_INIT) {
        float cv[32];
        int num_valid = 0;
        for (int i = 0; i < NV; ++i) {
            const float c = cm[(size_t)i * d.P + p];
            cv[i] = c;
            if (c < cost_max) num_valid++;
        }
        int j;
        for (int i = 1; i < NV; i++) {  // sort_small, DPE.cu:9-18
            const float tmp = cv[i];
            for (j = i; j >= 1 && tmp < cv[j - 1]; j--) cv[j] = cv[j - 1];
            cv[j] = tmp;
        }
        uint32_t sel = 0;
        const int top_k = min(num_valid, d.prm.top_k);
        float cost = cost_max;
        if (top_k > 0) {
            float sum = 0.0f;
            for (int i = 0; i < top_k; ++i) sum += cv[i];
            const float thr = cv[top_k - 1];
            for (int i = 0; i < NV; ++i)
                if (cm[(size_t)i * d.P + p] <= thr) sel |= (1u << i);
            cost = sum / top_k;
        }
        d.selected[p] = sel;
        d.costs[p] = cost;
    } else {
        uint32_t sel = d.selected[p];
        int count = 0;
        float cost = 0.0f;
        for (int i = 1; i < d.N; ++i) {
            if (isSet(sel, i - 1)) {
                const float c = cm[(size_t)(i - 1) * d.P + p];
                if (c < cost_max) {
                    count++;
                    cost += c;
                } else {
                    sel &= (uint32_t)(0xFFFFFFFEu << (i - 1));  // unSetBit, DPE.cu:81-84
                }
            }
        }
        d.selected[p] = sel;
        d.costs[p] = count == 0 ? cost_max : cost / count;
    }
}

// ------------------------------------------------- strong half-iteration
// Candidate search of CheckerboardPropagationStrong (DPE.cu:1239-1549).
// Slots: 0 current plane, 1+d adaptive/near-far candidate d, 9+d fixed-step
// candidate d (edge branch).  Also snapshots the four neighbours' selected
// views for the view-selection priors (DPE.cu:1558-1570).
__global__ void k_strong_candidates(Dev d, int colour, int iter) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.Q) return;
    const int Q = d.Q;
    const int nslot = d.prm.use_edge ? kStrongSlots : 9;
    int *cp = d.cand_pos + q;
    int x, y;
    const bool valid = colour_pixel(d, colour, q, x, y);
    const int W = d.W, H = d.H;
    const int p = x + y * W;
    if (!valid || d.weak[p] == DPE_WEAK) {
        for (int s = 0; s < nslot; ++s) cp[(size_t)s * Q] = -1;
        return;
    }
    cp[0] = p;
    const float *costs = d.costs;
    if (d.prm.use_edge) {
        const int dirx[8] = {0, 0, -1, 1, -1, 1, -1, 1};
        const int diry[8] = {-1, 1, 0, 0, -1, 1, 1, -1};
        const short2 *en = d.edge_neigh + (size_t)p * 8;
        const bool on_edge = d.edge[p] != 0;
        const float max_edge_dist = max(H, W) / 30.0f;
        const int min_step_len = 2;
        for (int dd = 0; dd < 8; ++dd) {
            const int dx = dirx[dd], dy = diry[dd];
            const int sx = max(1, 5 - 2 * iter) * dx;
            const int sy = max(1, 5 - 2 * iter) * dy;
            const short2 ep = en[dd];
            float dist = (float)sqrt((double)(ep.x - x) * (double)(ep.x - x) + (double)(ep.y - y) * (double)(ep.y - y));
            if (dd >= 4) dist = (float)(dist / sqrt(2.0));
            if (on_edge) {
                dist = 11 * min_step_len;
            } else if (ep.x == -1 || ep.y == -1 || dist > max_edge_dist) {
                dist = max_edge_dist;
                if (dd >= 4) dist = (float)(dist / sqrt(2.0));
            }
            const int step_num = min(max(11, (int)(1.0f * dist / min_step_len)), 22);
            int step_len = max((int)(1.0f * dist / step_num), min_step_len);
            if (dd < 4 && step_len % 2 == 1) step_len -= 1;
            int fx = 0, fy = 0;
            if (dd > 4) {
                if (dd % 2) fx = dx; else fy = dy;
            }
            int best = -1;
            float min_cost = 3.402823466e+38f;
            for (int step = 0; step < step_num; ++step) {
                const int tx = x + sx + step * step_len * dx + fx;
                const int ty = y + sy + step * step_len * dy + fy;
                if (!(tx >= 0 && ty >= 0 && tx < W && ty < H)) continue;
                const int tc = tx + ty * W;
                if (min_cost > costs[tc]) {
                    best = tc;
                    min_cost = costs[tc];
                }
            }
            const int adaptive = min_cost < 3.402823466e+38f ? best : -1;
            cp[(size_t)(1 + dd) * Q] = adaptive;
            int bestf = -1;
            if (!on_edge) {
                float mc = 3.402823466e+38f;
                for (int step = 0; step < 11; ++step) {
                    const int tx = x + sx + step * min_step_len * dx + fx;
                    const int ty = y + sy + step * min_step_len * dy + fy;
                    if (!(tx >= 0 && ty >= 0 && tx < W && ty < H)) continue;
                    const int tc = tx + ty * W;
                    if (mc > costs[tc]) {
                        bestf = tc;
                        mc = costs[tc];
                    }
                }
                if (!(mc < 3.402823466e+38f)) bestf = -1;
            }
            // The fixed scan (11 steps of min_step_len) is a prefix of the adaptive
            // one whenever step_len == min_step_len (at 1600x1200 always off edges),
            // so both often pick the same pixel.  Its costs then equal the adaptive
            // candidate's and never replace it in k_strong_select (the test needs
            // more good or fewer bad views), so it is not evaluated: -1, which the
            // selection skips the same way.
            cp[(size_t)(9 + dd) * Q] = (bestf >= 0 && bestf == adaptive) ? -1 : bestf;
        }
    } else {
        int pos[8];
        for (int i = 0; i < 8; ++i) pos[i] = -1;
        float costMin;
        int costMinPoint;
        const int center = p;
        if (y > 2) {
            const int up_far = center - 3 * W;
            costMin = costs[up_far];
            costMinPoint = up_far;
            for (int i = 1; i < 11; ++i)
                if (y > 2 + 2 * i) {
                    const int pt = up_far - 2 * i * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            pos[1] = costMinPoint;
        }
        if (y < H - 3) {
            const int down_far = center + 3 * W;
            costMin = costs[down_far];
            costMinPoint = down_far;
            for (int i = 1; i < 11; ++i)
                if (y < H - 3 - 2 * i) {
                    const int pt = down_far + 2 * i * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            pos[3] = costMinPoint;
        }
        if (x > 2) {
            const int left_far = center - 3;
            costMin = costs[left_far];
            costMinPoint = left_far;
            for (int i = 1; i < 11; ++i)
                if (x > 2 + 2 * i) {
                    const int pt = left_far - 2 * i;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            pos[5] = costMinPoint;
        }
        if (x < W - 3) {
            const int right_far = center + 3;
            costMin = costs[right_far];
            costMinPoint = right_far;
            for (int i = 1; i < 11; ++i)
                if (x < W - 3 - 2 * i) {
                    const int pt = right_far + 2 * i;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            pos[7] = costMinPoint;
        }
        if (y > 0) {
            const int up_near = center - W;
            costMin = costs[up_near];
            costMinPoint = up_near;
            for (int i = 0; i < 3; ++i) {
                if (y > 1 + i && x > i) {
                    const int pt = up_near - (1 + i) * W - (1 + i);
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
                if (y > 1 + i && x < W - 1 - i) {
                    const int pt = up_near - (1 + i) * W + (1 + i);
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            pos[0] = costMinPoint;
        }
        if (y < H - 1) {
            const int down_near = center + W;
            costMin = costs[down_near];
            costMinPoint = down_near;
            for (int i = 0; i < 3; ++i) {
                if (y < H - 2 - i && x > i) {
                    const int pt = down_near + (1 + i) * W - (1 + i);
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
                if (y < H - 2 - i && x < W - 1 - i) {
                    const int pt = down_near + (1 + i) * W + (1 + i);
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            pos[2] = costMinPoint;
        }
        if (x > 0) {
            const int left_near = center - 1;
            costMin = costs[left_near];
            costMinPoint = left_near;
            for (int i = 0; i < 3; ++i) {
                if (x > 1 + i && y > i) {
                    const int pt = left_near - (1 + i) - (1 + i) * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
                if (x > 1 + i && y < H - 1 - i) {
                    const int pt = left_near - (1 + i) + (1 + i) * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            pos[4] = costMinPoint;
        }
        if (x < W - 1) {
            const int right_near = center + 1;
            costMin = costs[right_near];
            costMinPoint = right_near;
            for (int i = 0; i < 3; ++i) {
                if (x < W - 2 - i && y > i) {
                    const int pt = right_near + (1 + i) - (1 + i) * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
                if (x < W - 2 - i && y < H - 1 - i) {
                    const int pt = right_near + (1 + i) + (1 + i) * W;
                    if (costs[pt] < costMin) { costMin = costs[pt]; costMinPoint = pt; }
                }
            }
            pos[6] = costMinPoint;
        }
        for (int i = 0; i < 8; ++i) cp[(size_t)(1 + i) * Q] = pos[i];
    }
    // Candidate rows that repeat a plane of this pixel's half-iteration: a row
    // is a pure function of (pixel, plane, view), so a slot whose plane equals,
    // bit for bit, the current plane's or that of a slot before it (in the
    // order the candidates are taken: the scanned ones, then the fixed-step
    // ones; far before near without edges) reads that row; cand_alias holds
    // the row each slot's costs are read from (the slot itself otherwise).
    {
        const float4 cur = d.planes[p];
        uint8_t *al = d.cand_alias + q;
        al[0] = 0;
        constexpr int order_plain[8] = {2, 4, 6, 8, 1, 3, 5, 7};
        float4 own[kStrongSlots - 1];
        bool is_own[kStrongSlots - 1];
#pragma unroll
        for (int k = 0; k < kStrongSlots - 1; ++k) {
            is_own[k] = false;
            own[k] = make_float4(0, 0, 0, 0);
            if (k >= nslot - 1) continue;
            const int slot = d.prm.use_edge ? 1 + k : order_plain[k];
            const int pos = cp[(size_t)slot * Q];
            int a = slot;
            if (pos >= 0) {
                const float4 pl = d.planes[pos];
                if (same_plane(pl, cur)) {
                    a = 0;
                } else {
#pragma unroll
                    for (int j = 0; j < k; ++j)
                        if (a == slot && is_own[j] && same_plane(own[j], pl)) a = d.prm.use_edge ? 1 + j : order_plain[j];
                }
                own[k] = pl;
                is_own[k] = a == slot;
            }
            al[(size_t)slot * Q] = (uint8_t)a;
        }
    }
    const long nb[4] = {(long)p - W, (long)p + W, (long)p - 1, (long)p + 1};
    for (int i = 0; i < 4; ++i) d.nsel[(size_t)i * Q + q] = (nb[i] >= 0 && nb[i] < (long)d.P) ? d.selected[nb[i]] : 0u;
}

// Per-view sample counts of the joint view selection (each at most 15: 15
// samples), 4 bits per view in two 64-bit words.  Read with a shift, so a
// count indexed by a runtime view stays in registers (a byte array indexed so
// lives in private memory).
struct ViewCounts {
    uint64_t lo = 0, hi = 0;
    __device__ __forceinline__ int operator[](int v) const {
        return (int)((v < 16 ? lo >> (4 * v) : hi >> (4 * (v - 16))) & 15u);
    }
    __device__ __forceinline__ void add(int v) {
        if (v < 16) lo += 1ull << (4 * v);
        else hi += 1ull << (4 * (v - 16));
    }
};

// Multi-hypothesis joint view selection (DPE.cu:1551-1631), shared by the
// strong select kernel and the weak propagation.  `cost(i, v)` reads row i of
// the 8 x NV candidate cost matrix, `prior(v)` gives view v's prior.  The
// sampling probabilities live in a fixed-size register array: a runtime view
// writes its entry through a select over all entries, and the CDF and the
// sampling scans are unrolled over every view with a guard, so no array is
// indexed by a runtime value (the same operations in the same order).
template <class CostFn, class PriorFn>
__device__ float joint_view_selection(const Dev &d, int iter, CostFn cost, PriorFn prior, Rng &rng, ViewCounts &vc,
                                      uint32_t *tsel_out) {
    constexpr int kV = kMaxImages - 1;
    const int NV = d.NV;
    float probs[kV];
#pragma unroll
    for (int k = 0; k < kV; ++k) probs[k] = 0.0f;
    const float thr = (float)(0.8 * (double)d_expf((float)(iter * iter) / (-90.0f)));
    for (int i = 0; i < NV; i++) {
        float count = 0;
        int count_false = 0;
        float tmpw = 0;
        for (int j = 0; j < 8; j++) {
            const float c = cost(j, i);
            if (c < thr) {
                tmpw += d_expf(c * c / (-0.18f));
                count++;
            }
            if (c > 1.2f) count_false++;
        }
        float pr = 0.0f;
        if (count > 2 && count_false < 3) pr = tmpw / count;
        else if (count_false < 3) pr = d_expf(thr * thr / (-0.32f));
        const float val = pr * prior(i);
#pragma unroll
        for (int k = 0; k < kV; ++k) probs[k] = k == i ? val : probs[k];
    }
    float sum = 0.0f;  // TransformPDFToCDF, DPE.cu:297-311
#pragma unroll
    for (int k = 0; k < kV; ++k)
        if (k < NV) sum += probs[k];
    const float inv = 1.0f / sum;
    float cum = 0.0f;
#pragma unroll
    for (int k = 0; k < kV; ++k) {
        if (k < NV) {
            const float pr = probs[k] * inv;
            cum += pr;
            probs[k] = cum;
        }
    }
    vc = ViewCounts{};
    for (int s = 0; s < 15; ++s) {
        const float rp = rng.uniform() - 1.19209290e-07f;
        // the first view whose CDF value exceeds the draw (none: no count)
        int sel = -1;
#pragma unroll
        for (int k = kV - 1; k >= 0; --k)
            if (k < NV && probs[k] > rp) sel = k;
        if (sel >= 0) vc.add(sel);
    }
    uint32_t tsel = 0;
    float wn = 0;
    for (int i = 0; i < NV; ++i) {
        if (vc[i] > 0) {
            tsel |= (1u << i);
            wn += vc[i];
        }
    }
    *tsel_out = tsel;
    return wn;
}

// View selection, candidate adoption and refinement hypotheses of
// CheckerboardPropagationStrong (DPE.cu:1551-1658, 1069-1099).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) void k_strong_select(Dev d, int colour, int iter) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.Q) return;
    const int Q = d.Q;
    if (d.cand_pos[q] < 0) return;
    int x, y;
    colour_pixel(d, colour, q, x, y);
    const int p = x + y * d.W;
    const int NV = d.NV;
    const dpe_camera &cam = d.cams[0];
    const float *cm = d.cost_mat;
    // chosen slot of every candidate row (-1: no candidate, the row stays {2,0,..} / 0);
    // a slot's costs are read from the row cand_alias names (the same plane's)
    const uint8_t *alias = d.cand_alias + q;
    int row_slot[8];
    for (int i = 0; i < 8; ++i) row_slot[i] = d.cand_pos[(size_t)(1 + i) * Q + q] >= 0 ? 1 + i : -1;
    if (d.prm.use_edge && !d.edge[p]) {
        const float good_thr = 0.8f * d_expf((float)(iter * iter) / (-90.0f));
        const float bad_thr = 1.2f;
        for (int i = 0; i < 8; ++i) {
            if (d.cand_pos[(size_t)(9 + i) * Q + q] < 0) continue;
            const bool had = row_slot[i] >= 0;
            const int r0 = had ? alias[(size_t)row_slot[i] * Q] : 0, r1 = alias[(size_t)(9 + i) * Q];
            int good0 = 0, bad0 = 0, good1 = 0, bad1 = 0;
            for (int j = 0; j < NV; ++j) {
                const float v0 = had ? cm[((size_t)r0 * NV + j) * Q + q] : ((i == 0 && j == 0) ? 2.0f : 0.0f);
                const float v1 = cm[((size_t)r1 * NV + j) * Q + q];
                if (v0 < good_thr) good0++;
                if (v0 > bad_thr) bad0++;
                if (v1 < good_thr) good1++;
                if (v1 > bad_thr) bad1++;
            }
            if (!had || good1 > good0 || (good1 == good0 && bad1 < bad0)) row_slot[i] = 9 + i;
        }
    }
    int row_read[8];
    for (int i = 0; i < 8; ++i) row_read[i] = row_slot[i] >= 0 ? (int)alias[(size_t)row_slot[i] * Q] : -1;
    auto cost = [&](int i, int v) -> float {
        const int s = row_read[i];
        if (s < 0) return (i == 0 && v == 0) ? 2.0f : 0.0f;
        return cm[((size_t)s * NV + v) * Q + q];
    };
    // view priors from the neighbours' selected views (each view's sum in neighbour order)
    uint32_t nsv[4];
    bool nuse[4];
    for (int i = 0; i < 4; ++i) {
        nuse[i] = row_slot[2 * i] >= 0;
        nsv[i] = nuse[i] ? d.nsel[(size_t)i * Q + q] : 0u;
    }
    auto prior = [&](int j) -> float {
        float pj = 0.0f;
        for (int i = 0; i < 4; ++i)
            if (nuse[i]) pj += isSet(nsv[i], j) == 1 ? 0.9f : 0.1f;
        return pj;
    };
    Rng rng(d.prm.seed, p, ph_strong(iter), d.prm.call_index);
    ViewCounts vwl;
    uint32_t tsel = 0;
    const float wn = joint_view_selection(d, iter, cost, prior, rng, vwl, &tsel);
    float final_costs[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        float fc = 0.0f;
        for (int j = 0; j < NV; ++j)
            if (vwl[j] > 0) fc += vwl[j] * cost(i, j);
        final_costs[i] = fc / wn;
    }
    // FindMinCostIndex, DPE.cu:50-61; the minimum's slot and row are carried
    // along (unrolled), so no array is indexed by the runtime minimum
    float fc_mi = final_costs[0];
    int slot_mi = row_slot[0], read_mi = row_read[0];
#pragma unroll
    for (int i = 1; i < 8; ++i)
        if (final_costs[i] <= fc_mi) {
            fc_mi = final_costs[i];
            slot_mi = row_slot[i];
            read_mi = row_read[i];
        }
    float cost_now = 0.0f;
    for (int i = 0; i < NV; ++i) cost_now += vwl[i] * cm[((size_t)0 * NV + i) * Q + q];
    cost_now /= wn;
    d.costs[p] = cost_now;
    const float4 cur = d.planes[p];
    float depth_now = depth_from_plane(cam, cur, x, y);
    float4 plane_now = cur;
    int now_row = 0;
    if (slot_mi >= 0) {
        const float4 cand = d.planes[d.cand_pos[(size_t)slot_mi * Q + q]];
        const float db = depth_from_plane(cam, cand, x, y);
        if (db >= d.prm.depth_min && db <= d.prm.depth_max && fc_mi < cost_now) {
            depth_now = db;
            plane_now = cand;
            now_row = read_mi;
            cost_now = fc_mi;
            d.selected[p] = tsel;
        }
    }
    uint8_t *vwo = d.vw + (size_t)p * kMaxImages;
    for (int i = 0; i < kMaxImages; ++i) vwo[i] = (uint8_t)vwl[i];
    // hypotheses of PlaneHypothesisRefinementStrong (DPE.cu:1080-1104)
    const float dmin = d.prm.depth_min, dmax = d.prm.depth_max;
    const float depth_rand = rng.uniform() * (dmax - dmin) + dmin;
    const float4 plane_rand = random_normal(cam, x, y, rng, depth_now);
    const float dminp = (1 - 0.02f) * depth_now;
    const float dmaxp = (1 + 0.02f) * depth_now;
    const float depth_pert = rng.uniform() * (dmaxp - dminp) + dminp;
    const float4 plane_pert = perturbed_normal(cam, x, y, plane_now, rng, (float)(0.02f * 3.14159265358979323846));
    const float depths[5] = {depth_rand, depth_now, depth_rand, depth_now, depth_pert};
    const float4 normals[5] = {plane_now, plane_rand, plane_rand, plane_pert, plane_now};
    for (int i = 0; i < 5; ++i) {
        const float4 n = normals[i];
        d.rplane[(size_t)i * Q + q] = make_float4(n.x, n.y, n.z, distance_to_origin(cam, x, y, depths[i], n));
    }
    d.now_plane[q] = plane_now;
    d.now_depth[q] = depth_now;
    d.now_cost[q] = cost_now;
    d.now_row[q] = now_row;
    d.wnorm[q] = wn;
}

// Acceptance of the refinement hypotheses and the RunState write-back
// (DPE.cu:1115-1121, 1660-1669).
__global__ void k_strong_accept(Dev d, int colour) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.Q) return;
    if (d.cand_pos[q] < 0) return;
    const int Q = d.Q;
    int x, y;
    colour_pixel(d, colour, q, x, y);
    const int p = x + y * d.W;
    const dpe_camera &cam = d.cams[0];
    float4 plane = d.now_plane[q];
    float depth = d.now_depth[q];
    float cost = d.now_cost[q];
    for (int i = 0; i < 5; ++i) {
        const float tc = d.rcost[(size_t)i * Q + q];
        const float4 tp = d.rplane[(size_t)i * Q + q];
        const float db = depth_from_plane(cam, tp, x, y);
        if (db >= d.prm.depth_min && db <= d.prm.depth_max && tc < cost) {
            depth = db;
            plane = tp;
            cost = tc;
        }
    }
    const float4 cur = d.planes[p];
    float4 fin = cur;
    if (d.prm.state == DPE_REFINE_INIT) {
        if ((double)cost < (double)d.costs[p] - 0.1) {
            d.costs[p] = cost;
            d.planes[p] = plane;
            fin = plane;
        }
    } else {
        d.costs[p] = cost;
        d.planes[p] = plane;
        fin = plane;
    }
    // per-view costs of the plane the pixel now holds: the current plane's
    // (cost_mat row 0) or the adopted candidate's (row now_row), all views;
    // a refinement hypothesis was evaluated on the weighted views only
    const float4 np = d.now_plane[q];
    const int row = same_plane(fin, np) ? d.now_row[q] : same_plane(fin, cur) ? 0 : -1;
    const int NV = d.NV;
    if (row < 0) {
        d.cur_valid[p] = 0u;
        return;
    }
    for (int v = 0; v < NV; ++v) d.cur_cost[(size_t)v * d.P + p] = d.cost_mat[((size_t)row * NV + v) * Q + q];
    d.cur_valid[p] = NV >= 32 ? 0xFFFFFFFFu : (1u << NV) - 1u;
}

// ------------------------------------------------------ weak-texture path
__device__ __forceinline__ uint8_t edge_low_at(const Dev &d, long idx) {
    if (idx < 0 || idx >= (long)d.low_w * d.low_h) return 0;
    return d.edge_low[idx];
}

// BresenhamLine, DPE.cu:162-248 (reads outside the low-res map return 0).  The
// result is whether any pixel a walk tests is an edge, and which pixels a walk
// tests does not depend on the map, so each walk is taken in chunks of 8 steps:
// the 8 positions first, then their 8 reads together (the reference's read,
// test and branch per step would leave one read in flight at a time).
__device__ bool bresenham(const Dev &d, int Ax, int Ay, int Bx, int By) {
    if (d.edge[Ax + Ay * d.W] || d.edge[Bx + By * d.W]) return false;
    const float scale_x = 1.0f * d.low_w / (float)d.W;
    const float scale_y = 1.0f * d.low_h / (float)d.H;
    const int height = d.low_h, width = d.low_w;
    const int max_step = d.prm.high_res_img ? (int)round(max(height, width) / 60.0) : max(height, width);
    for (int pass = 0; pass < 2; ++pass) {
        const int fx = pass == 0 ? Bx : Ax, fy = pass == 0 ? By : Ay;
        const int tx = pass == 0 ? Ax : Bx, ty = pass == 0 ? Ay : By;
        int x0 = min((int)roundf(fx * scale_x), width - 1);
        int y0 = min((int)roundf(fy * scale_y), height - 1);
        const int x1 = min((int)roundf(tx * scale_x), width - 1);
        const int y1 = min((int)roundf(ty * scale_y), height - 1);
        const int dx = abs(x1 - x0), sx = x0 < x1 ? 1 : -1;
        const int dy = abs(y1 - y0), sy = y0 < y1 ? 1 : -1;
        int erro = (dx > dy ? dx : dy) / 2;
        int step = 0;
        bool tagx = true, tagy = true;
        bool go = true;  // the walk's loop runs another step
        while (go) {
            long idx[8];
            bool use[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                use[k] = go;
                idx[k] = 0;
                if (go) {
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
                    idx[k] = (long)x0 + (long)y0 * width;
                    step += 1;
                    go = step < max_step && (tagx || tagy);
                }
            }
            uint8_t hit = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (use[k]) hit |= edge_low_at(d, idx[k]);
            if (hit) return true;
        }
    }
    return false;
}

// Bit position of the unordered pair (i, j), i != j, of compacted RANSAC points
// in a 2-bit-per-pair table (2 (hi (hi - 1) / 2 + lo)): the reference keeps the
// BresenhamLine result of every pair it has walked (edge_test, DPE.cu:2311,
// 2962; 0 = not walked, 1 = an edge between, 2 = none) and walks each pair
// once.  BresenhamLine is symmetric in its endpoints (it tests both and walks
// both ways), so one entry serves both orders.
__device__ __forceinline__ int edge_pair_bit(int i, int j) {
    const int lo = min(i, j), hi = max(i, j);
    return hi * (hi - 1) + 2 * lo;
}

// PointinTriangle, DPE.cu:139-160
__device__ bool point_in_triangle(short2 A, short2 B, short2 C, int px, int py) {
    const float abx = (float)(B.x - A.x), aby = (float)(B.y - A.y);
    const float bcx = (float)(C.x - B.x), bcy = (float)(C.y - B.y);
    const float cax = (float)(A.x - C.x), cay = (float)(A.y - C.y);
    const float ab = sqrtf(abx * abx + aby * aby);
    const float bc = sqrtf(bcx * bcx + bcy * bcy);
    const float ca = sqrtf(cax * cax + cay * cay);
    if (ab <= 2 || bc <= 2 || ca <= 2) return false;
    if (!(ab + bc > ca && bc + ca > ab && ab + ca > bc)) return false;
    const float pax = (float)(A.x - px), pay = (float)(A.y - py);
    const float pbx = (float)(B.x - px), pby = (float)(B.y - py);
    const float pcx = (float)(C.x - px), pcy = (float)(C.y - py);
    const float t1 = pax * pby - pay * pbx;
    const float t2 = pbx * pcy - pby * pcx;
    const float t3 = pcx * pay - pcy * pax;
    return t1 * t2 >= 0 && t1 * t3 >= 0;
}

// GenEdgeInform's (2r + 1)^2 window counts (DPE.cu:2527-2557), taken in two
// passes: rc[p] = the edge pixels and rc[P + p] the label-0 pixels among
// (x - r .. x + r, y) inside the image; the window's counts are the sums of
// 2r + 1 such row counts (integers: the same counts as the cell-by-cell loop).
__global__ void k_edge_row_counts(Dev d, uint16_t *__restrict__ rc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    const int W = d.W, x = p % W, r = d.prm.strong_radius;
    const int x0 = max(x - r, 0), x1 = min(x + r, W - 1);
    int e = 0, b = 0;
    for (int nx = x0; nx <= x1; ++nx) {
        const int q = p - x + nx;
        e += d.edge[q] != 0;
        if (d.prm.use_label) b += d.label[q] == 0;
    }
    rc[p] = (uint16_t)e;
    rc[(size_t)d.P + p] = (uint16_t)b;
}

// GenEdgeInform, DPE.cu:2487-2595
__global__ void k_gen_edge_inform(Dev d, const uint16_t *__restrict__ rc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    const int W = d.W, H = d.H;
    const int x = p % W, y = p / W;
    const int dirx[8] = {0, 0, -1, 1, -1, 1, -1, 1};
    const int diry[8] = {-1, 1, 0, 0, -1, 1, 1, -1};
    if (d.prm.use_edge) {
        short2 *en = d.edge_neigh + (size_t)p * 8;
        for (int i = 0; i < 8; i++) {
            short2 r = make_short2(-1, -1);
            int nx = x + dirx[i], ny = y + diry[i];
            while (true) {
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) break;
                if (d.edge[nx + ny * W]) {
                    r = make_short2((short)nx, (short)ny);
                    break;
                }
                nx += dirx[i];
                ny += diry[i];
            }
            en[i] = r;
        }
        // the window's in-image pixels, edge pixels and label-0 pixels: sums of
        // the row counts over the window's rows
        const int radius = d.prm.strong_radius;
        const int y0 = max(y - radius, 0), y1 = min(y + radius, H - 1);
        const int tot_pix = max(0, min(x + radius, W - 1) - max(x - radius, 0) + 1) * max(0, y1 - y0 + 1);
        int edge_pix = 0, bound_pix = 0;
        for (int ny = y0; ny <= y1; ++ny) {
            edge_pix += rc[(size_t)ny * W + x];
            bound_pix += rc[(size_t)d.P + (size_t)ny * W + x];
        }
        float density = 1.0f * edge_pix / tot_pix;
        if (d.prm.use_label) density = fmaxf(density, (float)(bound_pix / tot_pix));
        d.complexity[p] = (float)(1.0f / (1.0f + d_exp(-25.0 * ((double)density - 0.35))));
    }
}

// GenEdgeInform's label bounds of the WEAK pixels (DPE.cu:2564-2594), one thread
// per entry of the WEAK list: their walks are long and one thread per pixel
// left the lanes of the other pixels idle.
__global__ void k_label_bounds(Dev d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.weak_count) return;
    const int p = d.weak_list[i];
    const int W = d.W, H = d.H;
    const int x = p % W, y = p / W;
    const int dirx[8] = {0, 0, -1, 1, -1, 1, -1, 1};
    const int diry[8] = {-1, 1, 0, 0, -1, 1, 1, -1};
    if (d.weak[p] == DPE_WEAK) {
        short2 *lb = d.label_bound + (size_t)d.nmap[p] * 8;
        const int cl = d.label[p];
        if (cl > 0) {
            for (int i = 0; i < 8; i++) {
                int nx = x + dirx[i], ny = y + diry[i];
                int lx = -1, ly = -1;
                while (true) {
                    if (nx < 0 || nx >= W || ny < 0 || ny >= H) break;
                    const int nl = d.label[nx + ny * W];
                    if (nl == cl) {
                        lx = nx;
                        ly = ny;
                    } else if (nl == -1) {
                        break;
                    }
                    nx += dirx[i];
                    ny += diry[i];
                }
                lb[i] = make_short2((short)lx, (short)ly);
            }
        }
    }
}

// FindNearestStrongPoint, DPE.cu:2859-2893: the first STRONG pixel on the
// Chebyshev rings r = 0..100 around a WEAK pixel, rings scanned x-major
// (x = -r..r, then y = -r..r, ring cells only).  Restated exactly in two steps:
// the ring radius r* = min over STRONG q of max(|dx|, |dy|) comes from a
// row-wise distance map g (k_row_strong_distance), then only ring r* is scanned
// in the reference's order.  Cost O(r*) per pixel instead of O(r*^2).
// g[y][x] = min(127, |x - x'|) over the STRONG pixels x' of row y (127 when
// none lies within 126 pixels): one thread per pixel, scanning outwards from x
// (neighbouring lanes read neighbouring bytes).
__global__ void k_row_strong_distance(Dev d, int16_t *__restrict__ g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    const int W = d.W;
    const int x = p % W;
    const uint8_t *wk = d.weak + (size_t)(p - x);
    int best = 127;
    for (int k = 0; k < 127; ++k) {
        if ((x - k >= 0 && wk[x - k] == DPE_STRONG) || (x + k < W && wk[x + k] == DPE_STRONG)) {
            best = k;
            break;
        }
    }
    g[p] = (int16_t)best;
}

// Pixels that are not WEAK have no nearest strong point: (-1, -1).
__global__ void k_nearest_clear(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < d.P && d.weak[p] != DPE_WEAK) d.nearest[p] = make_short2(-1, -1);
}

// One thread per entry of the WEAK list (the other pixels were cleared).
__global__ void k_find_nearest_strong(Dev d, const int16_t *__restrict__ g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.weak_count) return;
    const int p = d.weak_list[i];
    const int W = d.W, H = d.H;
    const int x = p % W, y = p / W;
    short2 result = make_short2(-1, -1);
    if (d.weak[p] == DPE_WEAK) {
        const int max_radius = 100;
        int best = max_radius + 1;
        for (int k = 0; k <= max_radius && k < best; ++k) {
            if (y - k >= 0) best = min(best, max(k, (int)g[(size_t)(y - k) * W + x]));
            if (k > 0 && y + k < H) best = min(best, max(k, (int)g[(size_t)(y + k) * W + x]));
        }
        if (best <= max_radius) {
            const int r = best;
            bool found = false;
            for (int i = -r; i <= r && !found; ++i) {
                const int nx = x + i;
                if (nx < 0 || nx >= W) continue;
                if (abs(i) == r) {
                    for (int j = -r; j <= r; ++j) {
                        const int ny = y + j;
                        if (ny < 0 || ny >= H) continue;
                        if (d.weak[nx + ny * W] == DPE_STRONG) {
                            result = make_short2((short)nx, (short)ny);
                            found = true;
                            break;
                        }
                    }
                } else {
                    if (y - r >= 0 && d.weak[nx + (y - r) * W] == DPE_STRONG) {
                        result = make_short2((short)nx, (short)(y - r));
                        found = true;
                    } else if (y + r < H && d.weak[nx + (y + r) * W] == DPE_STRONG) {
                        result = make_short2((short)nx, (short)(y + r));
                        found = true;
                    }
                }
            }
        }
    }
    d.nearest[p] = result;
}

// A pixel (x, y) packed into a float's bits (moved, never computed on).
__device__ __forceinline__ float pack_pixel(short2 s) {
    return __builtin_bit_cast(float, (uint32_t)(uint16_t)s.x | ((uint32_t)(uint16_t)s.y << 16));
}
__device__ __forceinline__ short2 unpack_pixel(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return make_short2((short)(u & 0xFFFFu), (short)(u >> 16));
}
__device__ __forceinline__ bool pixel_valid(float f) { return __builtin_bit_cast(uint32_t, f) != 0xFFFFFFFFu; }

// GenNeighbours, DPE.cu:2107-2467
// cache_words: 32-bit words per lane of the dynamic LDS pair table
// (edge_pair_bit; word w of lane l at [w * 64 + l]), 0 = no table.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_gen_neighbours(Dev d, int cache_words) {
    extern __shared__ uint32_t edge_cache[];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.weak_count) return;
    const int p = d.weak_list[i];
    if (d.weak[p] != DPE_WEAK) return;
    const int W = d.W, H = d.H;
    const int px = p % W, py = p / W;
    constexpr int max_pt_num = 64;
    const int min_margin = 6;
    const float depth_diff = d.prm.depth_max - d.prm.depth_min;
    const dpe_camera &cam = d.cams[0];
    short2 *neighbours = d.neighbours + (size_t)d.nmap[p] * kNeighbourNum;
    Rng rng(d.prm.seed, p, PH_GEN_NEIGHBOURS, d.prm.call_index);
    for (int i = 0; i < kNeighbourNum; ++i) neighbours[i] = make_short2(-1, -1);
    neighbours[0] = make_short2((short)px, (short)py);
    // One 16-B record per candidate strong point (private memory): first the
    // point only (.w = packed pixel, (-1, -1) = none), after compaction also the
    // iteration-invariant terms of the RANSAC residual: fx = (x - cx) / fx_K,
    // fy likewise, and the point's depth (the reference recomputes them per
    // iteration and keeps 3D points and normals in arrays; the values are the same).
    float4 pts[max_pt_num];
    for (int i = 0; i < max_pt_num; ++i) pts[i] = make_float4(0.0f, 0.0f, 0.0f, pack_pixel(make_short2(-1, -1)));
    int odi = -1;
    int strong_point_size = 0;
    const UMod shift_mod((uint32_t)d.gn_shift);
    const int rotate_time = d.prm.rotate_time;
    const float ransac_threshold = d.prm.ransac_threshold * depth_diff;
    bool edge_limit = false;
    if (d.prm.use_limit) {
        edge_limit = true;
        if (d.prm.use_edge) {
            const float cv = d.complexity[p];
            const float rp = rng.uniform() - 1.19209290e-07f;
            if (rp < cv) edge_limit = false;
            else d.complexity[p] = fmaxf(0.99f, cv);
        }
    }
    for (int odx = -1; odx <= 1; ++odx) {
        for (int ody = -1; ody <= 1; ++ody) {
            if (odx == 0 && ody == 0) continue;
            float2 od = make_float2((float)odx, (float)ody);
            normalize2(&od);
            odi++;
            for (int rot = 0; rot < rotate_time; ++rot) {
                const int di = odi * 4 + rot;
                for (int radius = 2; radius <= DPE_MAX_SEARCH_RADIUS; radius = min(radius * 2, radius + 25)) {
                    const float tx = (float)px + od.x * (float)radius;
                    const float ty = (float)py + od.y * (float)radius;
                    if (tx < 0 || ty < 0 || tx >= (float)W || ty >= (float)H) break;
                    // The 4 jittered probes of this radius: their pixels and the state
                    // and nearest-strong reads of all 4 first (the reads then overlap),
                    // then the tests in probe order.  Each probe takes 4 draws and the
                    // probes after the accepted one take none, so the stream position
                    // is set back to just after the accepted probe.
                    const uint32_t n0 = rng.n;
                    short2 probe[4];
                    bool inside[4];
#pragma unroll
                    for (int ri = 0; ri < 4; ++ri) {
                        const uint32_t s1 = rng.next();
                        const uint32_t m1 = rng.next();
                        const int rxs = (int)shift_mod(((s1 % 2 == 0) ? 1u : 0xFFFFFFFFu) * m1);
                        const uint32_t s2 = rng.next();
                        const uint32_t m2 = rng.next();
                        const int rys = (int)shift_mod(((s2 % 2 == 0) ? 1u : 0xFFFFFFFFu) * m2);
                        float2 dv = make_float2(od.x * 20 + (float)rxs, od.y * 20 + (float)rys);
                        normalize2(&dv);
                        probe[ri] = make_short2((short)(int)((float)px + dv.x * (float)radius),
                                                (short)(int)((float)py + dv.y * (float)radius));
                        inside[ri] = !(probe[ri].x < min_margin || probe[ri].y < min_margin ||
                                       probe[ri].x >= W - min_margin || probe[ri].y >= H - min_margin);
                    }
                    uint8_t state[4];
#pragma unroll
                    for (int ri = 0; ri < 4; ++ri)
                        state[ri] = inside[ri] ? d.weak[probe[ri].x + probe[ri].y * W] : (uint8_t)DPE_STRONG;
#pragma unroll
                    for (int ri = 0; ri < 4; ++ri)
                        if (inside[ri] && state[ri] != DPE_STRONG) probe[ri] = d.nearest[probe[ri].x + probe[ri].y * W];
                    int accepted = -1;
                    for (int ri = 0; ri < 4; ++ri) {
                        const short2 np = probe[ri];
                        if (!inside[ri] || np.x == -1 || np.y == -1) continue;
                        float2 td = make_float2((float)(np.x - px), (float)(np.y - py));
                        normalize2(&td);
                        const float ca = td.x * od.x + td.y * od.y;
                        if (ca > d.gn_thr && (!edge_limit || !bresenham(d, px, py, np.x, np.y))) {
                            pts[di].w = pack_pixel(np);
                            strong_point_size++;
                            accepted = ri;
                            break;
                        }
                    }
                    if (accepted >= 0) rng.n = n0 + 4u * (uint32_t)(accepted + 1);
                    if (pixel_valid(pts[di].w)) break;
                }
                float2 rd;
                rd.x = od.x * d.gn_cos - od.y * d.gn_sin;
                rd.y = od.x * d.gn_sin + od.y * d.gn_cos;
                normalize2(&rd);
                od = rd;
            }
        }
    }
    int extend_index = 31;
    if (d.prm.use_label && d.label[p] > 0) {
        const int dirx[8] = {0, 0, -1, 1, -1, 1, -1, 1};
        const int diry[8] = {-1, 1, 0, 0, -1, 1, 1, -1};
        const short2 *lb = d.label_bound + (size_t)d.nmap[p] * 8;
        float bound_dist[8];
        int dir_step[8];
        for (int i = 0; i < 8; ++i) {
            bound_dist[i] = 0;
            dir_step[i] = 0;
        }
        for (int i = 0; i < 8; ++i) {
            const short2 bp = lb[i];
            float dist = 0.0f;
            if (bp.x != -1 && bp.y != -1) {
                dist = (float)sqrt((double)(px - bp.x) * (double)(px - bp.x) + (double)(py - bp.y) * (double)(py - bp.y));
                if (i >= 4) dist = (float)(dist / sqrt(2.0));
            }
            bound_dist[i] = dist;
            if (i % 2 == 1) {
                const float opposite = bound_dist[i - 1];
                const float ratio = 2 * rotate_time * dist / (dist + opposite);
                const int rint_ = (ratio != ratio) ? 0 : f2i_clamp(ratio, -1000000, 1000000);
                const int step = min(1, max(2 * rotate_time - 1, rint_));
                dir_step[i - 1] = 2 * rotate_time - step;
                dir_step[i] = step;
            }
        }
        for (int i = 0; i < 8; ++i) {
            const float dist = bound_dist[i];
            const int gap = dir_step[i] + 1;
            const int step_len = max(1, (int)floor(1.0 * dist / gap));
            for (int step = 1; step <= dir_step[i]; ++step) {
                short2 np = make_short2((short)(px + step * step_len * dirx[i]), (short)(py + step * step_len * diry[i]));
                if (np.x < min_margin || np.y < min_margin || np.x >= W - min_margin || np.y >= H - min_margin) continue;
                int nc = np.x + np.y * W;
                if (d.weak[nc] != DPE_STRONG) {
                    np = d.nearest[nc];
                    if (np.x == -1 || np.y == -1) continue;
                    nc = np.x + np.y * W;
                }
                if (d.label[nc] != 0 && d.label[nc] != d.label[p]) continue;
                if (extend_index >= max_pt_num - 1) continue;
                extend_index++;
                pts[extend_index].w = pack_pixel(np);
                strong_point_size++;
            }
        }
    }
    if (strong_point_size <= 3) {
        d.weak_rel[p] = 0;
        return;
    }
    float4 best_plane = make_float4(0, 0, 0, 0);
    bool has_valid_plane = false;
    float residuals[max_pt_num];
    int vc = 0;
    float X[3];
    get3d(cam, px, py, d.planes[p].w, X);
    const float cpz = X[2];
    // compact the valid points to the front in index order (the rest: none)
    for (int i = 0; i < max_pt_num; ++i) {
        const float pw = pts[i].w;
        pts[i].w = pack_pixel(make_short2(-1, -1));
        if (pixel_valid(pw)) {
            const short2 sp = unpack_pixel(pw);
            const float fx = ((float)sp.x - cam.K[2]) / cam.K[0];
            const float fy = ((float)sp.y - cam.K[5]) / cam.K[4];
            pts[vc] = make_float4(fx, fy, d.planes[sp.x + sp.y * W].w, pw);
            vc++;
        }
    }
    // the 3D point (s3) and camera-frame normal (sn) of a compacted point
    auto point3 = [&](int k) {
        const short2 sp = unpack_pixel(pts[k].w);
        float Y[3];
        get3d(cam, sp.x, sp.y, pts[k].z, Y);
        return make_float3(Y[0], Y[1], Y[2]);
    };
    auto normal3 = [&](int k) {
        const short2 sp = unpack_pixel(pts[k].w);
        const float4 n4 = to_camera(cam, d.planes[sp.x + sp.y * W]);
        return make_float3(n4.x, n4.y, n4.z);
    };
    {
        int iteration = 50, max_iter = d.prm.high_res_img ? 200 : 125, max_count = 3;
        float min_cost = 3.402823466e+38f;
        // residuals of the current plane; only read where the threshold is
        // tightened, so they are written there (entries >= vc stay 0)
        for (int i = 0; i < max_pt_num; ++i) residuals[i] = 0.0f;
        float temp_thr = ransac_threshold;
        bool has_consist = false;
        bool must_in_triangle = (d.prm.use_label && d.label[p] > 0 && edge_limit) ? false : true;
        // the pair table of this lane, when its points fit
        const bool cached = cache_words > 0 && vc * (vc - 1) <= 32 * cache_words;
        uint32_t *table = edge_cache + (threadIdx.x & 63);
        if (cached)
            for (int w = 0; w < cache_words; ++w) table[w * 64] = 0u;
        auto edge_between = [&](int i, int j, short2 P, short2 Qp) -> bool {
            if (!cached) return bresenham(d, P.x, P.y, Qp.x, Qp.y);
            const int pb = edge_pair_bit(i, j);
            uint32_t &w = table[(pb >> 5) * 64];
            const uint32_t e = (w >> (pb & 31)) & 3u;
            if (e) return e == 1u;
            const bool r = bresenham(d, P.x, P.y, Qp.x, Qp.y);
            w |= (r ? 1u : 2u) << (pb & 31);
            return r;
        };
        const UMod pick((uint32_t)vc);
        while (iteration > 0 && max_iter > 0) {
            max_iter--;
            const int a = (int)pick(rng.next());
            const int b = (int)pick(rng.next());
            const int c = (int)pick(rng.next());
            if (a == b || b == c || a == c) continue;
            const short2 sa = unpack_pixel(pts[a].w), sb = unpack_pixel(pts[b].w), sc3 = unpack_pixel(pts[c].w);
            if (must_in_triangle && !point_in_triangle(sa, sb, sc3, px, py)) continue;
            if (edge_limit) {
                if (edge_between(a, b, sa, sb) || edge_between(b, c, sb, sc3) || edge_between(c, a, sc3, sa)) continue;
            }
            bool normal_consistency = false;
            if (d.prm.geom_consistency && edge_limit) {
                normal_consistency = true;
                const float3 na = normal3(a), nb = normal3(b), nc = normal3(c);
                const float ab = na.x * nb.x + na.y * nb.y + na.z * nb.z;
                const float ac = na.x * nc.x + na.y * nc.y + na.z * nc.z;
                const float bc = nb.x * nc.x + nb.y * nc.y + nb.z * nc.z;
                if ((double)ab < 0.8660254 || (double)ac < 0.8660254 || (double)bc < 0.8660254) normal_consistency = false;
                if (has_consist && !normal_consistency) continue;
            }
            iteration--;
            const float3 A = point3(a), B = point3(b), Cc = point3(c);
            const float3 AC = make_float3(A.x - Cc.x, A.y - Cc.y, A.z - Cc.z);
            const float3 BC = make_float3(B.x - Cc.x, B.y - Cc.y, B.z - Cc.z);
            float4 cr;
            cr.x = AC.y * BC.z - BC.y * AC.z;
            cr.y = -(AC.x * BC.z - BC.x * AC.z);
            cr.z = AC.x * BC.y - BC.x * AC.y;
            if ((cr.x == 0 && cr.y == 0 && cr.z == 0) || cr.x != cr.x || cr.y != cr.y || cr.z != cr.z) continue;
            normalize3(&cr);
            cr.w = -(cr.x * A.x + cr.y * A.y + cr.z * A.z);
            int temp_count = 0;
            float strong_dist = 0.0f;
            auto residual = [&](int si) {
                const float4 q = pts[si];
                const float fit_depth = -cr.w / (cr.x * q.x + cr.y * q.y + cr.z);
                return fabsf(fit_depth - q.z);
            };
            // A count below max(6, max_count) changes nothing below (no plane, no
            // flag, no threshold), so the scan stops once the remaining points
            // cannot lift the count to it: the same decisions, fewer residuals.
            const int needed = max(6, max_count);
            for (int si = 0; si < vc && temp_count + (vc - si) >= needed; ++si) {
                const float distance = residual(si);
                if (distance < temp_thr) {
                    temp_count++;
                    strong_dist += distance;
                }
            }
            if (temp_count < 6) continue;
            if (temp_count > max_count) {
                if (!must_in_triangle && point_in_triangle(sa, sb, sc3, px, py)) must_in_triangle = true;
                if (!has_consist && normal_consistency) has_consist = true;
                const float fx = ((float)px - cam.K[2]) / cam.K[0];
                const float fy = ((float)py - cam.K[5]) / cam.K[4];
                const float fit_depth = -cr.w / (cr.x * fx + cr.y * fy + cr.z);
                const float cd = fabsf(fit_depth - cpz);
                best_plane = cr;
                max_count = temp_count;
                strong_dist /= temp_count;
                min_cost = cd;
                has_valid_plane = true;
                if ((double)temp_thr > (d.prm.high_res_img ? 0.05 : 0.005)) {
                    for (int si = 0; si < vc; ++si) residuals[si] = residual(si);
                    int j;
                    for (int i = 1; i < vc; i++) {
                        const float tmp = residuals[i];
                        for (j = i; j >= 1 && tmp < residuals[j - 1]; j--) residuals[j] = residuals[j - 1];
                        residuals[j] = tmp;
                    }
                    if (temp_thr < residuals[kNeighbourNum]) continue;
                    temp_thr = (float)((double)residuals[kNeighbourNum] - 1e-6);
                    temp_count = 0;
                    for (int i = 0; i < vc; ++i) {
                        if (residuals[i] < temp_thr) temp_count++;
                        else break;
                    }
                    max_count = temp_count;
                }
            } else if (temp_count == max_count) {
                if (!must_in_triangle && point_in_triangle(sa, sb, sc3, px, py)) must_in_triangle = true;
                const float fx = ((float)px - cam.K[2]) / cam.K[0];
                const float fy = ((float)py - cam.K[5]) / cam.K[4];
                const float fit_depth = -cr.w / (cr.x * fx + cr.y * fy + cr.z);
                const float cd = fabsf(fit_depth - cpz);
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
        d.weak_rel[p] = 0;
        return;
    }
    float *weight = residuals;  // no longer needed as residuals
    for (int i = 0; i < vc; ++i) {
        const float4 q = pts[i];
        const float fit_depth = -best_plane.w / (best_plane.x * q.x + best_plane.y * q.y + best_plane.z);
        const float distance = fabsf(fit_depth - q.z);
        if (distance >= ransac_threshold) {
            pts[i].w = pack_pixel(make_short2(-1, -1));
            weight[i] = 3.402823466e+38f;
            continue;
        }
        weight[i] = distance;
    }
    {
        int j;  // sort_small_weighted, DPE.cu:20-33
        for (int i = 1; i < vc; i++) {
            const float tmp = pts[i].w;
            const float tw = weight[i];
            for (j = i; j >= 1 && tw < weight[j - 1]; j--) {
                pts[j].w = pts[j - 1].w;
                weight[j] = weight[j - 1];
            }
            pts[j].w = tmp;
            weight[j] = tw;
        }
    }
    for (int i = 1; i < kNeighbourNum; ++i) neighbours[i] = unpack_pixel(pts[i - 1].w);
    d.weak_rel[p] = 1;
}

// NeigbourUpdate, DPE.cu:2469-2485
__global__ void k_neighbour_update(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    if (d.weak[p] == DPE_WEAK && d.weak_rel[p] != 1) d.weak[p] = DPE_UNKNOWN;
}

// RANSACToGetFitPlane, DPE.cu:2895-3128
__global__ void k_fit_copy(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    if (d.weak[p] != DPE_WEAK) d.fit[p] = d.planes[p];
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) void k_ransac(Dev d, int iter) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.weak_count) return;
    const int p = d.weak_list[i];
    const int W = d.W;
    const int px = p % W, py = p / W;
    if (d.weak[p] != DPE_WEAK) return;
    Rng rng(d.prm.seed, p, ph_ransac(iter), d.prm.call_index);
    const dpe_camera &cam = d.cams[0];
    bool edge_limit = false;
    if (d.prm.use_limit) {
        edge_limit = true;
        if (d.prm.use_edge) {
            const float cv = d.complexity[p];
            const float rp = rng.uniform() - 1.19209290e-07f;
            if (rp < cv) edge_limit = false;
        }
    }
    short2 sp[kNeighbourNum - 1];
    float3 s3[kNeighbourNum - 1];
    float3 sn[kNeighbourNum - 1];
    int count = 0;
    float X[3];
    for (int i = 1; i < kNeighbourNum; ++i) {
        const short2 tp = neighbour_point(d, p, i);
        if (tp.x == -1 || tp.y == -1) continue;
        sp[count] = tp;
        const int tc = tp.x + tp.y * W;
        const float4 pl = d.planes[tc];
        const float depth = depth_from_plane(cam, pl, tp.x, tp.y);
        get3d(cam, tp.x, tp.y, depth, X);
        s3[count] = make_float3(X[0], X[1], X[2]);
        sn[count] = make_float3(pl.x, pl.y, pl.z);
        count++;
    }
    if (count < 3) {
        d.fit[p] = d.planes[p];
        return;
    }
    int iteration = 50;
    int ua = -1, ub = -1, uc = -1;
    float min_cost = 3.402823466e+38f;
    float4 best = make_float4(0, 0, 0, 0);
    bool has_best = false, has_strong = false;
    bool must_in_triangle = (d.prm.use_label && d.label[p] > 0 && edge_limit) ? false : true;
    uint64_t edge_table = 0;  // edge_pair_bit entries of this pixel's points
    const UMod pick((uint32_t)count);
    while (iteration--) {
        const int a = (int)pick(rng.next());
        const int b = (int)pick(rng.next());
        const int c = (int)pick(rng.next());
        if (a == b || b == c || a == c) continue;
        bool is_strong = false;
        if (d.prm.geom_consistency && edge_limit) {
            is_strong = true;
            const float ab = sn[a].x * sn[b].x + sn[a].y * sn[b].y + sn[a].z * sn[b].z;
            const float ac = sn[a].x * sn[c].x + sn[a].y * sn[c].y + sn[a].z * sn[c].z;
            const float bc = sn[b].x * sn[c].x + sn[b].y * sn[c].y + sn[b].z * sn[c].z;
            if ((double)ab < 0.8660254 || (double)ac < 0.8660254 || (double)bc < 0.8660254) is_strong = false;
            if (has_strong && !is_strong) continue;
        }
        if (must_in_triangle && !point_in_triangle(sp[a], sp[b], sp[c], px, py)) continue;
        if (edge_limit) {
            // the pair table (edge_pair_bit) of at most 8 points in one 64-bit word
            auto edge_between = [&](int i, int j) -> bool {
                const int pb = edge_pair_bit(i, j);
                const uint32_t e = (uint32_t)(edge_table >> pb) & 3u;
                if (e) return e == 1u;
                const bool r = bresenham(d, sp[i].x, sp[i].y, sp[j].x, sp[j].y);
                edge_table |= (uint64_t)(r ? 1u : 2u) << pb;
                return r;
            };
            if (edge_between(a, b) || edge_between(b, c) || edge_between(c, a)) continue;
        }
        const float3 A = s3[a], B = s3[b], Cc = s3[c];
        const float3 AC = make_float3(A.x - Cc.x, A.y - Cc.y, A.z - Cc.z);
        const float3 BC = make_float3(B.x - Cc.x, B.y - Cc.y, B.z - Cc.z);
        float4 cr;
        cr.x = AC.y * BC.z - BC.y * AC.z;
        cr.y = -(AC.x * BC.z - BC.x * AC.z);
        cr.z = AC.x * BC.y - BC.x * AC.y;
        if ((cr.x == 0 && cr.y == 0 && cr.z == 0) || cr.x != cr.x || cr.y != cr.y || cr.z != cr.z) continue;
        normalize3(&cr);
        cr.w = -(cr.x * A.x + cr.y * A.y + cr.z * A.z);
        if (!has_strong && is_strong) has_strong = true;
        // The partial sums of non-negative terms never decrease (and NaN stays),
        // so once tcost < min_cost fails it fails at the end: the scan stops there.
        float tcost = 0.0f;
        for (int si = 0; si < count && tcost < min_cost; ++si) {
            if (si == a || si == b || si == c) continue;
            const float fx = ((float)sp[si].x - cam.K[2]) / cam.K[0];
            const float fy = ((float)sp[si].y - cam.K[5]) / cam.K[4];
            const float fit_depth = -cr.w / (cr.x * fx + cr.y * fy + cr.z);
            tcost += fabsf(fit_depth - s3[si].z);
        }
        if (tcost < min_cost) {
            if (!must_in_triangle && point_in_triangle(sp[a], sp[b], sp[c], px, py)) must_in_triangle = true;
            min_cost = tcost;
            best = cr;
            has_best = true;
            ua = a;
            ub = b;
            uc = c;
        }
    }
    if (has_best) {
        const float depth = depth_from_plane(cam, d.planes[p], px, py);
        const float4 vd = view_direction(cam, px, py, depth);
        const float dot = best.x * vd.x + best.y * vd.y + best.z * vd.z;
        if (dot > 0) {
            best.x = -best.x;
            best.y = -best.y;
            best.z = -best.z;
            best.w = -best.w;
        }
        d.fit[p] = best;
        if (d.prm.use_radius) {
            if (must_in_triangle) {
                const short2 A = sp[ua], B = sp[ub], Cc = sp[uc];
                const float a = sqrtf((float)((A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y)));
                const float b = sqrtf((float)((B.x - Cc.x) * (B.x - Cc.x) + (B.y - Cc.y) * (B.y - Cc.y)));
                const float c = sqrtf((float)((Cc.x - A.x) * (Cc.x - A.x) + (Cc.y - A.y) * (Cc.y - A.y)));
                const float pp = (float)((a + b + c) / 2.0);
                const float S = sqrtf(pp * (pp - a) * (pp - b) * (pp - c));
                // (int)floor of NaN (degenerate triangle) is defined as 0 on both paths
                const double rr = floor(sqrtf(S) / 2.0);
                int radius = (rr != rr) ? 0 : (int)fmax(-1e9, fmin(1e9, rr));
                const float A_dis = sqrtf((float)((A.x - px) * (A.x - px) + (A.y - py) * (A.y - py)));
                const float B_dis = sqrtf((float)((B.x - px) * (B.x - px) + (B.y - py) * (B.y - py)));
                const float C_dis = sqrtf((float)((Cc.x - px) * (Cc.x - px) + (Cc.y - py) * (Cc.y - py)));
                const float min_dis = fminf(fminf(A_dis, B_dis), C_dis);
                if (2.5 * min_dis < radius) radius = (int)min_dis;
                if (edge_limit) {
                    if (d.prm.use_edge) {
                        float med = 3.402823466e+38f;
                        const short2 *en = d.edge_neigh + (size_t)p * 8;
                        for (int k = 0; k < 8; ++k) {
                            const short2 ep = en[k];
                            if (ep.x == -1 || ep.y == -1) continue;
                            med = fminf(med, sqrtf((float)((ep.x - px) * (ep.x - px) + (ep.y - py) * (ep.y - py))));
                        }
                        if (med < radius) radius = (int)med;
                    }
                    if (d.prm.use_label && d.label[p] > 0) {
                        float mbd = 3.402823466e+38f;
                        const short2 *lb = d.label_bound + (size_t)d.nmap[p] * 8;
                        for (int k = 0; k < 8; ++k) {
                            const short2 bp = lb[k];
                            if (bp.x == -1 || bp.y == -1) continue;
                            const float dist = (float)sqrt((double)(px - bp.x) * (double)(px - bp.x) +
                                                           (double)(py - bp.y) * (double)(py - bp.y));
                            mbd = fminf(mbd, dist);
                        }
                        if (mbd < radius) radius = (int)mbd;
                    }
                }
                while ((radius << 1) % 5 != 0) radius--;
                if (!edge_limit)
                    d.radius[p] = radius > d.prm.strong_radius ? 0 : d.prm.strong_radius;
                else
                    d.radius[p] = radius > d.prm.strong_radius ? radius : d.prm.strong_radius;
            } else {
                d.radius[p] = d.prm.strong_radius;
            }
        }
    } else {
        d.fit[p] = make_float4(0, 0, 0, 0);
        if (d.prm.use_radius) d.radius[p] = d.prm.strong_radius;
    }
}

// The candidate planes of CheckerboardPropagationWeak (DPE.cu:1690-1703): the
// STRONG ones among the 8 GenNeighbours points of weak pixel p.  src[i] = the
// first candidate at the same pixel as candidate i (i itself when distinct):
// different search directions often reach the same nearest strong pixel, and
// a candidate at the same pixel has the same plane and so the same costs.
// Returns the number of distinct candidates.
__device__ __forceinline__ int weak_candidates(const Dev &d, int p, int positions[8], bool flag[8], int src[8]) {
    int u = 0;
    for (int i = 0; i < 8; ++i) {
        flag[i] = false;
        positions[i] = 0;
        src[i] = i;
        const short2 np = neighbour_point(d, p, i + 1);
        if (np.x == -1 || np.y == -1 || d.weak[np.x + np.y * d.W] != DPE_STRONG) continue;
        positions[i] = np.x + np.y * d.W;
        flag[i] = true;
        for (int k = 0; k < i; ++k)
            if (flag[k] && positions[k] == positions[i]) {
                src[i] = k;
                break;
            }
        if (src[i] == i) u++;
    }
    return u;
}

// Candidate cost rows of the weak propagation (ComputeMultiViewCostVectorNew
// of the STRONG GenNeighbours points, DPE.cu:1700-1711), evaluated apart from
// the rest of the weak update so that every lane has a window to evaluate: a
// wave takes 64 consecutive entries of the colour's weak list, lists their
// distinct candidates as (entry, slot) items in LDS, entry by entry (the
// candidates of one pixel share its neighbour-window positions, so their
// gathers fall on nearby source texels), and evaluates 64 items at a time,
// each lane looping over the views.  Row (slot i, view v) of entry li goes to
// weak_cost[(i * NV + v - 1) * n + li]; k_weak_update reads it back.
constexpr int kWeakCandWaves = 4;
__global__ __launch_bounds__(kLanes * kWeakCandWaves) void k_weak_cand(Dev d, int colour) {
    __shared__ uint16_t items[kWeakCandWaves][kLanes * 8];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int n = d.weak_colour_count[colour];
    const int *list = d.weak_colour_list[colour];
    const int li0 = (xcd_block(blockIdx.x, gridDim.x) * kWeakCandWaves + wave) * kLanes;
    const int li = li0 + lane;
    uint32_t mask = 0;  // bit i: slot i holds a distinct STRONG candidate
    if (li < n && d.weak[list[li]] == DPE_WEAK) {
        bool flag[8];
        int positions[8], src[8];
        weak_candidates(d, list[li], positions, flag, src);
        for (int i = 0; i < 8; ++i)
            if (flag[i] && src[i] == i) mask |= 1u << i;
    }
    // exclusive prefix sum of the per-lane counts (0..8, four bits) over the wave
    const uint32_t cnt = (uint32_t)__popc(mask);
    int first = 0, total = 0;
    for (int b = 0; b < 4; ++b) {
        const unsigned long long bb = __ballot((cnt >> b) & 1u);
        first += lanes_below(bb) << b;
        total += __popcll(bb) << b;
    }
    uint16_t *it = items[wave];
    for (uint32_t m = mask; m; m &= m - 1) it[first++] = (uint16_t)((lane << 3) | __builtin_ctz(m));
    // the waves' lists, concatenated in entry order, are dealt to the waves 64
    // items at a time, so that no wave ends on a partial pass of its own
    __shared__ int list_first[kWeakCandWaves + 1];
    if (lane == 0) list_first[wave + 1] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        list_first[0] = 0;
        for (int w = 1; w <= kWeakCandWaves; ++w) list_first[w] += list_first[w - 1];
    }
    __syncthreads();
    const int n_items = list_first[kWeakCandWaves];
    const int block_li0 = xcd_block(blockIdx.x, gridDim.x) * kWeakCandWaves * kLanes;
    unsigned long long hyps = 0, taps = 0;
    const int NV = d.NV;
    for (int base = wave * kLanes; base < n_items; base += kWeakCandWaves * kLanes) {
        const int k = base + lane;
        if (k < n_items) {
            int sw = 0;
            for (int w = 1; w < kWeakCandWaves; ++w)
                if (k >= list_first[w]) sw = w;
            const uint32_t e = items[sw][k - list_first[sw]];
            const int lk = block_li0 + sw * kLanes + (int)(e >> 3), i = (int)(e & 7u);
            const int p = list[lk];
            const int x = p % d.W, y = p / d.W;
            const short2 np = neighbour_point(d, p, i + 1);
            const float4 pl = d.planes[np.x + np.y * d.W];
            float *row = d.weak_cost + (size_t)i * NV * n + lk;
            for (int v = 1; v < d.N; ++v) row[(size_t)(v - 1) * n] = ncc_new_pre(d, x, y, v, pl, colour, lk, hyps, taps);
        }
    }
    count_add(d.counters, kFamWeak, hyps, taps);
}

// CheckerboardPropagationWeak + PlaneHypothesisRefinementWeak (DPE.cu:1672-1866,
// 1124-1216), split where the windows are evaluated so that every window is a
// lane's own item (a weak pixel's window counts vary with its candidates,
// selected views and fitted plane, and one lane per pixel left half the lanes
// idle):
//   k_weak_cand    candidate cost rows (above)
//   k_weak_select  view selection over the candidates, their final costs, argmin
//   k_weak_terms   window terms of the current and the fitted plane
//   k_weak_mid     current cost, candidate and fitted-plane acceptance, the 5
//                  refinement hypotheses
//   k_weak_terms   window terms of the hypotheses
//   k_weak_final   hypothesis acceptance, write-back, final cost (NCC Old)
// Every value is computed by the same operations in the same order as in one
// pass over the pixel; the sums of the terms run over the views in order.

// One thread per entry of the colour's weak list (XCD-contiguous ranges of the
// tile-ordered list, so that nearby weak pixels share source cache lines);
// false for entries past the list or not WEAK.
__device__ __forceinline__ bool weak_entry(const Dev &d, int colour, int &li, int &p, int &x, int &y) {
    li = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (li >= d.weak_colour_count[colour]) return false;
    p = d.weak_colour_list[colour][li];
    x = p % d.W;
    y = p / d.W;
    return d.weak[p] == DPE_WEAK;
}

// Multi-hypothesis joint view selection and the candidates' final costs
// (DPE.cu:1713-1807).
__global__ __launch_bounds__(64) void k_weak_select(Dev d, int colour, int iter) {
    int li, p, x, y;
    const bool act = weak_entry(d, colour, li, p, x, y);
    if (li >= d.weak_colour_count[colour]) return;
    WeakState &st = d.weak_st[li];
    if (!act) {
        st.flags = 0;
        return;
    }
    const int W = d.W;
    const int NV = d.NV;
    const bool geom = d.prm.geom_consistency != 0;
    const float gf = d.prm.geom_factor;
    bool flag[8];
    int positions[8], src[8];
    weak_candidates(d, p, positions, flag, src);
    // the candidate cost matrix: k_weak_cand's rows, read into a local array in
    // one pass (independent loads; read where they are used, each would wait on
    // its own load); a repeated candidate takes the row of the first one at its
    // pixel; rows without a candidate are {2, 0, ...} and 0 as the reference
    // initialises them (DPE.cu:1694)
    float cost_array[8][32];
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 32; ++j) cost_array[i][j] = 0.0f;
    cost_array[0][0] = 2.0f;
    {
        const int n = d.weak_colour_count[colour];
        for (int i = 0; i < 8; ++i) {
            if (!flag[i]) continue;
            const float *row = d.weak_cost + (size_t)src[i] * NV * n + li;
            for (int v = 0; v < NV; ++v) cost_array[i][v] = row[(size_t)v * n];
        }
    }
    auto cost = [&](int i, int v) -> float { return cost_array[i][v]; };
    // view priors from the GenNeighbours points' selected views (each view's sum in point order)
    uint32_t nsv[8];
    bool nuse[8];
    for (int i = 0; i < 8; ++i) {
        const short2 np = neighbour_point(d, p, i + 1);
        nuse[i] = !(np.x == -1 || np.y == -1);
        nsv[i] = nuse[i] ? d.selected[np.x + np.y * W] : 0u;
    }
    auto prior = [&](int j) -> float {
        float pj = 0.0f;
        for (int i = 0; i < 8; ++i)
            if (nuse[i]) pj += isSet(nsv[i], j) == 1 ? 0.9f : 0.1f;
        return pj;
    };
    Rng rng(d.prm.seed, p, ph_weak(iter), d.prm.call_index);
    ViewCounts vwl;
    uint32_t tsel = 0;
    const float wn = joint_view_selection(d, iter, cost, prior, rng, vwl, &tsel);
    float final_costs[8];
    for (int i = 0; i < 8; ++i) {
        float fc = 0.0f;
        for (int j = 0; j < NV; ++j) {
            if (vwl[j] > 0) {
                if (geom) {
                    if (flag[i])
                        fc += vwl[j] * (cost(i, j) + gf * geom_cost(d, x, y, j + 1, d.planes[positions[i]]));
                    else
                        fc += vwl[j] * (cost(i, j) + gf * 3.0f);
                } else {
                    fc += vwl[j] * cost(i, j);
                }
            }
        }
        final_costs[i] = fc / wn;
    }
    int mi = 0;
    {
        float mc = final_costs[0];
        for (int i = 1; i < 8; ++i)
            if (final_costs[i] <= mc) { mc = final_costs[i]; mi = i; }
    }
    // the view weights (the reference's selection writes them in place,
    // DPE.cu:1761-1771)
    uint8_t *vwo = d.vw + (size_t)p * kMaxImages;
    for (int i = 0; i < kMaxImages; ++i) vwo[i] = (uint8_t)vwl[i];
    const float4 fit = d.fit[p];
    uint32_t fl = kWeakActive;
    if (flag[mi]) fl |= kWeakCand;
    if (!(fit.x == 0 && fit.y == 0 && fit.z == 0)) fl |= kWeakFit;
    st.cand = flag[mi] ? d.planes[positions[mi]] : make_float4(0, 0, 0, 0);
    st.fc_mi = final_costs[mi];
    st.wn = wn;
    st.tsel = tsel;
    st.rng_n = rng.n;
    st.flags = fl;
}

// Window terms vw_v * (cost_v + geom_factor * geometric cost_v) (or vw_v * cost_v)
// of the weak entries' planes over their selected views: slots 0 (the current
// plane) and 1 (the fitted plane, where there is one) in PHASE 0, the 5
// refinement hypotheses (where the entry draws them) in PHASE 1.  A wave takes
// 64 consecutive entries and deals their (view, slot) items, entry by entry, to
// its lanes 64 at a time; term (slot, v) of entry li goes to
// weak_terms[(slot * NV + v - 1) * n + li].
template <int PHASE>
__global__ __launch_bounds__(kLanes * kWeakCandWaves) __attribute__((amdgpu_waves_per_eu(4))) void k_weak_terms(Dev d, int colour) {
    __shared__ int pref[kWeakCandWaves][kLanes + 1];
    __shared__ uint32_t vmask[kWeakCandWaves][kLanes];
    __shared__ uint8_t nslot[kWeakCandWaves][kLanes];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int n = d.weak_colour_count[colour];
    const int *list = d.weak_colour_list[colour];
    const int li0 = (xcd_block(blockIdx.x, gridDim.x) * kWeakCandWaves + wave) * kLanes;
    const int li = li0 + lane;
    const int NV = d.NV;
    uint32_t views = 0;
    int ns = 0;
    if (li < n) {
        const uint32_t fl = d.weak_st[li].flags;
        if (fl & kWeakActive) {
            const uint8_t *vw = d.vw + (size_t)list[li] * kMaxImages;
            for (int v = 0; v < NV; ++v)
                if (vw[v] > 0) views |= 1u << v;
            ns = PHASE == 0 ? ((fl & kWeakFit) ? 2 : 1) : ((fl & kWeakRandom) ? 5 : 0);
        }
    }
    int incl = ns * __popc(views);
    for (int o = 1; o < kLanes; o <<= 1) {
        const int t = __shfl_up(incl, o, kLanes);
        if (lane >= o) incl += t;
    }
    pref[wave][lane + 1] = incl;
    if (lane == 0) pref[wave][0] = 0;
    vmask[wave][lane] = views;
    nslot[wave][lane] = (uint8_t)ns;
    const int total = __shfl(incl, kLanes - 1, kLanes);
    wave_sync();
    unsigned long long hyps = 0, taps = 0;
    const bool geom = d.prm.geom_consistency != 0;
    const float gf = d.prm.geom_factor;
    for (int base = 0; base < total; base += kLanes) {
        const int k = base + lane;
        if (k < total) {
            // the entry: the last lane whose items start at or before k
            int lo = 0, hi = kLanes - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pref[wave][mid] <= k) lo = mid;
                else hi = mid - 1;
            }
            const int j = k - pref[wave][lo];
            const int nsl = nslot[wave][lo];
            const int vi = j / nsl;
            const int slot = (PHASE == 0 ? 0 : 2) + (j - vi * nsl);
            uint32_t vm = vmask[wave][lo];
            for (int t = 0; t < vi; ++t) vm &= vm - 1;
            const int v = __builtin_ctz(vm) + 1;
            const int lk = li0 + lo;
            const int p = list[lk];
            const int x = p % d.W, y = p / d.W;
            const float4 pl = slot == 0 ? d.planes[p] : (slot == 1 ? d.fit[p] : d.weak_st[lk].hyp[slot - 2]);
            const float c = ncc_new_pre(d, x, y, v, pl, colour, lk, hyps, taps);
            const uint8_t w = d.vw[(size_t)p * kMaxImages + v - 1];
            const float term = geom ? w * (c + gf * geom_cost(d, x, y, v, pl)) : w * c;
            d.weak_terms[((size_t)slot * NV + v - 1) * n + lk] = term;
        }
    }
    count_add(d.counters, kFamWeak, hyps, taps);
}

// sum of an entry's terms of one slot over its selected views, in view order
__device__ __forceinline__ float weak_term_sum(const Dev &d, int n, int li, int slot, const uint8_t *vwl) {
    const float *t = d.weak_terms + (size_t)slot * d.NV * n + li;
    float s = 0.0f;
    for (int v = 0; v < d.NV; ++v)
        if (vwl[v] > 0) s += t[(size_t)v * n];
    return s;
}

// The current plane's cost, the candidate's acceptance (DPE.cu:1809-1834), the
// fitted plane's (PlaneHypothesisRefinementWeak, DPE.cu:1143-1172) and the
// refinement hypotheses (DPE.cu:1174-1192).
__global__ __launch_bounds__(64) void k_weak_mid(Dev d, int colour, int iter) {
    int li, p, x, y;
    if (!weak_entry(d, colour, li, p, x, y)) return;
    WeakState &st = d.weak_st[li];
    const uint32_t fl = st.flags;
    const int n = d.weak_colour_count[colour];
    const dpe_camera &cam = d.cams[0];
    const float wn = st.wn;
    uint8_t vwl[kMaxImages];
    const uint8_t *vw = d.vw + (size_t)p * kMaxImages;
    for (int i = 0; i < kMaxImages; ++i) vwl[i] = vw[i];
    const float4 cur = d.planes[p];
    float cost_now = weak_term_sum(d, n, li, 0, vwl);
    cost_now /= wn;
    d.costs[p] = cost_now;
    float depth_now = depth_from_plane(cam, cur, x, y);
    float4 plane_now = cur;
    if (fl & kWeakCand) {
        const float db = depth_from_plane(cam, st.cand, x, y);
        if (db >= d.prm.depth_min && db <= d.prm.depth_max && st.fc_mi < cost_now) {
            depth_now = db;
            plane_now = st.cand;
            cost_now = st.fc_mi;
            d.selected[p] = st.tsel;
        }
    }
    const float dmin = d.prm.depth_min, dmax = d.prm.depth_max;
    bool do_random = true;
    if (!(fl & kWeakFit)) {
        do_random = false;
    } else {
        const float4 fit = d.fit[p];
        const float tc = weak_term_sum(d, n, li, 1, vwl) / wn;
        const float db = depth_from_plane(cam, fit, x, y);
        if (db >= dmin && db <= dmax && tc < cost_now) {
            depth_now = db;
            plane_now = fit;
            cost_now = tc;
        }
    }
    if (do_random) {
        Rng rng(d.prm.seed, p, ph_weak(iter), d.prm.call_index);
        rng.n = st.rng_n;
        const float depth_rand = rng.uniform() * (dmax - dmin) + dmin;
        const float4 plane_rand = random_normal(cam, x, y, rng, depth_now);
        const float dminp = (1 - 0.02f) * depth_now;
        const float dmaxp = (1 + 0.02f) * depth_now;
        const float depth_pert = rng.uniform() * (dmaxp - dminp) + dminp;
        const float4 plane_pert = perturbed_normal(cam, x, y, plane_now, rng, (float)(0.02f * 3.14159265358979323846));
        const float depths[5] = {depth_rand, depth_now, depth_rand, depth_now, depth_pert};
        const float4 normals[5] = {plane_now, plane_rand, plane_rand, plane_pert, plane_now};
        for (int i = 0; i < 5; ++i) {
            const float4 nn = normals[i];
            st.hyp[i] = make_float4(nn.x, nn.y, nn.z, distance_to_origin(cam, x, y, depths[i], nn));
        }
        st.flags = fl | kWeakRandom;
    }
    st.plane_now = plane_now;
    st.cost_now = cost_now;
}

// The hypotheses' acceptance (DPE.cu:1193-1213), the write-back
// (DPE.cu:1838-1847) and the final cost with the Old NCC (DPE.cu:1849-1865).
__global__ __launch_bounds__(64) void k_weak_final(Dev d, int colour) {
    int li, p, x, y;
    const bool act = weak_entry(d, colour, li, p, x, y);
    unsigned long long hyps = 0, taps = 0;
    if (act) {
        const WeakState &st = d.weak_st[li];
        const int n = d.weak_colour_count[colour];
        const dpe_camera &cam = d.cams[0];
        const float wn = st.wn;
        uint8_t vwl[kMaxImages];
        const uint8_t *vw = d.vw + (size_t)p * kMaxImages;
        for (int i = 0; i < kMaxImages; ++i) vwl[i] = vw[i];
        const float dmin = d.prm.depth_min, dmax = d.prm.depth_max;
        float cost_now = st.cost_now;
        float4 plane_now = st.plane_now;
        if (st.flags & kWeakRandom) {
            for (int i = 0; i < 5; ++i) {
                const float4 tp = st.hyp[i];
                const float tc = weak_term_sum(d, n, li, 2 + i, vwl) / wn;
                const float db = depth_from_plane(cam, tp, x, y);
                if (db >= dmin && db <= dmax && tc < cost_now) {
                    plane_now = tp;
                    cost_now = tc;
                }
            }
        }
        float4 final_plane = d.planes[p];
        if (d.prm.state == DPE_REFINE_INIT) {
            if ((double)cost_now < (double)d.costs[p] - 0.1) {
                d.costs[p] = cost_now;
                d.planes[p] = plane_now;
                final_plane = plane_now;
            }
        } else {
            d.costs[p] = cost_now;
            d.planes[p] = plane_now;
            final_plane = plane_now;
        }
        float c2 = 0.0f;
        for (int v = 1; v < d.N; ++v)
            if (vwl[v - 1] > 0) c2 += vwl[v - 1] * ncc_old_direct(d, x, y, v, final_plane, hyps, taps);
        d.costs[p] = c2 / wn;
    }
    count_add(d.counters, kFamWeak, hyps, taps);
}

// ------------------------------------------------------------ post-pass
// GetDepthandNormal, DPE.cu:1944-1959
__global__ void k_depth_and_normal(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    const int x = p % d.W, y = p / d.W;
    const float4 pl = d.planes[p];
    d.planes[p] = to_world(d.cams[0], make_float4(pl.x, pl.y, pl.z, depth_from_plane(d.cams[0], pl, x, y)));
}

// CheckerboardFilterStrong, DPE.cu:1961-2105
__global__ void k_filter(Dev d, int colour) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    int x, y;
    if (!colour_pixel(d, colour, q, x, y)) return;
    const int W = d.W, H = d.H;
    const int center = x + y * W;
    const uint8_t *wk = d.weak;
    if (wk[center] == DPE_WEAK) return;
    float4 *pl = d.planes;
    if (d.costs[center] < 0.001f) return;
    float f[21];
    int n = 0;
    f[n++] = pl[center].w;
    const int left = center - 1, leftleft = center - 3;
    const int up = center - W, upup = center - 3 * W;
    const int down = center + W, downdown = center + 3 * W;
    const int right = center + 1, rightright = center + 3;
    if (y > 0 && wk[up] == DPE_STRONG) f[n++] = pl[up].w;
    if (y > 2 && wk[upup] == DPE_STRONG) f[n++] = pl[upup].w;
    if (y > 4 && wk[upup - W * 2] == DPE_STRONG) f[n++] = pl[upup - W * 2].w;
    if (y < H - 1 && wk[down] == DPE_STRONG) f[n++] = pl[down].w;
    if (y < H - 3 && wk[downdown] == DPE_STRONG) f[n++] = pl[downdown].w;
    if (y < H - 5 && wk[downdown + W * 2] == DPE_STRONG) f[n++] = pl[downdown + W * 2].w;
    if (x > 0 && wk[left] == DPE_STRONG) f[n++] = pl[left].w;
    if (x > 2 && wk[leftleft] == DPE_STRONG) f[n++] = pl[leftleft].w;
    if (x > 4 && wk[leftleft - 2] == DPE_STRONG) f[n++] = pl[leftleft - 2].w;
    if (x < W - 1 && wk[right] == DPE_STRONG) f[n++] = pl[right].w;
    if (x < W - 3 && wk[rightright] == DPE_STRONG) f[n++] = pl[rightright].w;
    if (x < W - 5 && wk[rightright + 2] == DPE_STRONG) f[n++] = pl[rightright + 2].w;
    if (y > 0 && x < W - 2 && wk[up + 2] == DPE_STRONG) f[n++] = pl[up + 2].w;
    if (y < H - 1 && x < W - 2 && wk[down + 2] == DPE_STRONG) f[n++] = pl[down + 2].w;
    if (y > 0 && x > 1 && wk[up - 2] == DPE_STRONG) f[n++] = pl[up - 2].w;
    if (y < H - 1 && x > 1 && wk[down - 2] == DPE_STRONG) f[n++] = pl[down - 2].w;
    if (x > 0 && y > 2 && wk[left - W * 2] == DPE_STRONG) f[n++] = pl[left - W * 2].w;
    if (x < W - 1 && y > 2 && wk[right - W * 2] == DPE_STRONG) f[n++] = pl[right - W * 2].w;
    if (x > 0 && y < H - 2 && wk[left + W * 2] == DPE_STRONG) f[n++] = pl[left + W * 2].w;
    if (x < W - 1 && y < H - 2 && wk[right + W * 2] == DPE_STRONG) f[n++] = pl[right + W * 2].w;
    int j;
    for (int i = 1; i < n; i++) {
        const float tmp = f[i];
        for (j = i; j >= 1 && tmp < f[j - 1]; j--) f[j] = f[j - 1];
        f[j] = tmp;
    }
    const int m = n / 2;
    pl[center].w = (n % 2 == 0) ? (f[m - 1] + f[m]) / 2 : f[m];
}

// Shared set-up of DepthToWeak (DPE.cu:2605-2661) and LocalRefine
// (DPE.cu:2761-2808): camera-frame plane, baseline, disparity, weight norm.
__global__ void k_post_prep(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    const int W = d.W, H = d.H;
    const int x = p % W, y = p / W;
    const dpe_camera &cam = d.cams[0];
    const uint32_t sel = d.selected[p];
    const uint8_t *vw = d.vw + (size_t)p * kMaxImages;
    const float4 op = to_camera(cam, d.planes[p]);
    const float od = op.w;
    float bl = 0.0f, wn = 0.0f;
    int valid = 0;
    for (int s = 1; s < d.N; ++s) {
        if (isSet(sel, s - 1)) {
            wn += vw[s - 1];
            bl += d.vc[s].base;
            valid++;
        }
    }
    uint8_t flags = 0;
    const bool margin = x < 6 || y < 6 || x >= W - 6 || y >= H - 6;
    if (margin || od == 0 || valid == 0) {
        d.weak[p] = DPE_UNKNOWN;  // DepthToWeak's early exits
    } else {
        flags |= 1;
    }
    if (!(od == 0) && !(wn == 0 || valid == 0)) flags |= 2;
    d.post_flags[p] = flags;
    if (valid > 0) bl /= valid;
    d.post_op[p] = op;
    d.post_bl[p] = bl;
    d.post_disp[p] = cam.K[0] * bl / od;
    d.post_wn[p] = wn;
}

// Sample planes of DepthToWeak (DPE.cu:2667-2675) and LocalRefine
// (DPE.cu:2813-2819): slot 0..60 disparity offsets -30..30, slot 61 the current
// depth, slots 62..72 offsets -5..5.  A sample outside the depth range gets its
// final cost here (DepthToWeak 2.0; LocalRefine +inf, i.e. skipped) and w = NaN.
__global__ void k_post_planes(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    const uint8_t fl = d.post_flags[p];
    if (!fl) return;
    const int x = p % d.W, y = p / d.W;
    const dpe_camera &cam = d.cams[0];
    const float4 op = d.post_op[p];
    const float bl = d.post_bl[p];
    const float disp = d.post_disp[p];
    const size_t P = d.P;
    for (int slot = 0; slot < kPostSlots; ++slot) {
        const bool dtw = slot < 61;
        if (!((dtw && (fl & 1)) || (!dtw && (fl & 2)))) continue;
        if (post_sample_unread(d, slot)) {
            d.pcost[(size_t)slot * P + p] = 2.0f;
            d.post_planes[(size_t)slot * P + p] = make_float4(0.0f, 0.0f, 0.0f, __builtin_nanf(""));
            continue;
        }
        float pd = op.w;
        if (slot != 61) {
            const int k = dtw ? slot - 30 : slot - 67;
            pd = cam.K[0] * bl / (disp + (float)k);
            if (pd < d.prm.depth_min || pd > d.prm.depth_max) {
                d.pcost[(size_t)slot * P + p] = dtw ? 2.0f : __builtin_inff();
                d.post_planes[(size_t)slot * P + p] = make_float4(0.0f, 0.0f, 0.0f, __builtin_nanf(""));
                continue;
            }
        }
        d.post_planes[(size_t)slot * P + p] = make_float4(op.x, op.y, op.z, distance_to_origin(cam, x, y, pd, op));
    }
}

// Peak analysis of DepthToWeak (DPE.cu:2699-2750) and LocalRefine's update
// (DPE.cu:2811-2838).
__global__ void k_post_final(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    const uint8_t flags = d.post_flags[p];
    const size_t P = d.P;
    if (flags & 1) {
        const int radius = 30, n = 61;
        int peak_count = 0, min_peak = 0;
        float min_cost = 2.0f;
        uint64_t peaks = 0;
        float prev = d.pcost[1 * P + p], cur = d.pcost[2 * P + p];
        for (int i = 2; i < n - 2; ++i) {
            const float nxt = d.pcost[(size_t)(i + 1) * P + p];
            if (prev > cur && nxt > cur) {
                peaks |= (1ull << i);
                peak_count++;
                if (cur < min_cost) {
                    min_peak = i;
                    min_cost = cur;
                }
            }
            prev = cur;
            cur = nxt;
        }
        const float pmin = d.pcost[(size_t)min_peak * P + p];
        uint8_t st;
        if (abs(min_peak - radius) > d.prm.weak_peak_radius || pmin > 0.5f) {
            st = DPE_WEAK;
        } else if (peak_count == 1) {
            st = pmin <= 0.15f ? DPE_STRONG : DPE_WEAK;
        } else {
            float var = 0.0f;
            for (int i = 2; i < n - 2; ++i) {
                if (((peaks >> i) & 1) && i != min_peak) {
                    const float dist = d.pcost[(size_t)i * P + p] - min_cost;
                    var += dist * dist;
                }
            }
            var = sqrtf(var);
            var /= (peak_count - 1);
            st = var > 0.2f ? DPE_STRONG : DPE_WEAK;
        }
        d.weak[p] = st;
    }
    if (flags & 2) {
        const float cost_now = d.pcost[61 * P + p];
        float min_cost = 2.0f;
        const dpe_camera &cam = d.cams[0];
        const float bl = d.post_bl[p], disp = d.post_disp[p];
        float best_depth = d.post_op[p].w;
        for (int k = -5; k <= 5; ++k) {
            const float tc = d.pcost[(size_t)(67 + k) * P + p];
            if (tc < min_cost) {
                min_cost = tc;
                best_depth = cam.K[0] * bl / (disp + (float)k);
            }
        }
        if ((double)(cost_now - min_cost) > 0.1) d.planes[p].w = best_depth;
    }
}

// ------------------------------------------------------------ launchers
static inline dim3 grid1(int n, int b = 256) { return dim3((n + b - 1) / b); }

void launch_make_quads(const Dev &d, const float *img, float4 *quads, hipStream_t s) {
    const int n = (d.W + 1) * (d.H + 1);
    hipLaunchKernelGGL(k_make_quads, grid1(n), dim3(256), 0, s, img, quads, d.W, d.H);
}
void launch_make_quads_h(const Dev &d, const float *img, uint2 *quads, int fmt, hipStream_t s) {
    const int n = (d.W + 1) * (d.H + 1);
    hipLaunchKernelGGL(k_make_quads_h, grid1(n), dim3(256), 0, s, img, quads, d.W, d.H, fmt);
}
void launch_check_fp16_exact(const float *img, size_t n, int *flag, hipStream_t s) {
    hipLaunchKernelGGL(k_check_fp16_exact, dim3(1024), dim3(256), 0, s, img, n, flag);
}
hipError_t weak_lists_temp_bytes(int W, int H, size_t *bytes) {
    const int P = W * H;
    size_t a = 0, b = 0, c = 0;
    hipError_t e = hipcub::DeviceSelect::If(nullptr, a, PixelIter(0), (int *)nullptr, (int *)nullptr, P,
                                            WeakSelect{nullptr, 1, 0, -1});
    if (e != hipSuccess) return e;
    e = hipcub::DeviceSelect::If(nullptr, c, TiledIter(PixelIter(0), TiledPixel{W, H, 1}), (int *)nullptr,
                                 (int *)nullptr, tiled_items(W, H), WeakSelect{nullptr, 1, 0, 0});
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, b, WeakFlagIter(PixelIter(0), WeakFlag{nullptr}), (int *)nullptr, P);
    *bytes = std::max(std::max(a, b), c);
    return e;
}
hipError_t launch_weak_lists(const uint8_t *weak, int W, int H, int half_rows, int *nmap, int *all, int *col0,
                             int *col1, int *counts, void *temp, size_t temp_bytes, hipStream_t s) {
    const int P = W * H;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, WeakFlagIter(PixelIter(0), WeakFlag{weak}),
                                                    nmap, P, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_nmap_clear_others, grid1(P), dim3(256), 0, s, weak, nmap, P);
    const TiledIter tiled(PixelIter(0), TiledPixel{W, H, (W + kWeakTileW - 1) / kWeakTileW});
    int *out[3] = {all, col0, col1};
    for (int k = 0; k < 3; ++k) {
        e = hipcub::DeviceSelect::If(temp, temp_bytes, tiled, out[k], counts + k, tiled_items(W, H),
                                     WeakSelect{weak, W, half_rows, k - 1}, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}
void launch_gen_edge_inform(const Dev &d, hipStream_t s) {
    if (d.prm.use_edge) {
        // the row counts (2 x P 16-bit words) in the post-pass scratch, free before the iterations
        uint16_t *rc = (uint16_t *)d.pcost;
        hipLaunchKernelGGL(k_edge_row_counts, grid1(d.P), dim3(256), 0, s, d, rc);
        hipLaunchKernelGGL(k_gen_edge_inform, grid1(d.P), dim3(256), 0, s, d, (const uint16_t *)rc);
    }
    if (d.prm.use_label && d.weak_count > 0)
        hipLaunchKernelGGL(k_label_bounds, grid1(d.weak_count), dim3(256), 0, s, d);
}
void launch_find_nearest_strong(const Dev &d, hipStream_t s) {
    int16_t *g = (int16_t *)d.pcost;  // the post-pass scratch is free before the iterations
    hipLaunchKernelGGL(k_row_strong_distance, grid1(d.P), dim3(256), 0, s, d, g);
    hipLaunchKernelGGL(k_nearest_clear, grid1(d.P), dim3(256), 0, s, d);
    if (d.weak_count > 0)
        hipLaunchKernelGGL(k_find_nearest_strong, grid1(d.weak_count), dim3(256), 0, s, d, (const int16_t *)g);
}
void launch_gen_neighbours(const Dev &d, hipStream_t s) {
    if (d.weak_count > 0) {
        // GenNeighbours collects at most 8 rotate_time points (16 rotate_time with
        // labels, at most 63); the pair table is kept in LDS when 32 points fit.
        const int max_points = min(63, 8 * d.prm.rotate_time * (d.prm.use_label ? 2 : 1));
        const int words = max_points <= 32 ? (max_points * (max_points - 1) + 31) / 32 : 0;
        hipLaunchKernelGGL(k_gen_neighbours, grid1(d.weak_count, 64), dim3(64), (size_t)words * 64 * 4, s, d,
                           words);
    }
}
void launch_neighbour_update(const Dev &d, hipStream_t s) {
    hipLaunchKernelGGL(k_neighbour_update, grid1(d.P), dim3(256), 0, s, d);
}
void launch_random_init_planes(const Dev &d, hipStream_t s) {
    hipLaunchKernelGGL(k_random_init_planes, grid1(d.P), dim3(256), 0, s, d);
}

static inline dim3 grid_direct(int n, int nslot) { return dim3((n + 255) / 256, nslot); }

// Shared-patch NCC launches with the call's tile (Dev::full_tile / wtd_tile:
// waves per workgroup and extra dynamic LDS that caps workgroups per CU).
// Every tile gives bit-identical results; only the schedule differs.
template <bool PREC, int WAVES>
static void full_t(const Dev &d, hipStream_t s, int colour, int Qn, int nslot, const int *cand, const float4 *src,
                   const uint32_t *mask, float *out, bool cached, float4 *last) {
    // the item lists (strong launches only) sit at the start of the dynamic LDS, inside the tile's pad
    const size_t lists = cand ? (size_t)WAVES * full_list_bytes(nslot, WAVES) : 0;
    const size_t lds = std::max(lists, (size_t)d.full_tile.lds_pad);
    hipLaunchKernelGGL((k_ncc_full<6, PREC, WAVES>), dim3((Qn + kLanes - 1) / kLanes), dim3(kLanes * WAVES), lds, s,
                       d, colour, Qn, nslot, cand, src, mask, out, cached, last);
}
static void launch_full(const Dev &d, hipStream_t s, int colour, int Qn, int nslot, const int *cand,
                        const float4 *src, const uint32_t *mask, float *out, bool cached = false,
                        float4 *last = nullptr) {
    const int w = d.full_tile.waves;
    if (d.prec) {
        if (w == 2) full_t<true, 2>(d, s, colour, Qn, nslot, cand, src, mask, out, cached, last);
        else if (w == 8) full_t<true, 8>(d, s, colour, Qn, nslot, cand, src, mask, out, cached, last);
        else full_t<true, 4>(d, s, colour, Qn, nslot, cand, src, mask, out, cached, last);
    } else {
        if (w == 2) full_t<false, 2>(d, s, colour, Qn, nslot, cand, src, mask, out, cached, last);
        else if (w == 8) full_t<false, 8>(d, s, colour, Qn, nslot, cand, src, mask, out, cached, last);
        else full_t<false, 4>(d, s, colour, Qn, nslot, cand, src, mask, out, cached, last);
    }
}
template <int MODE, bool PREC, int WAVES>
static void wtd_t(const Dev &d, hipStream_t s, int colour, int Qn, int nslot, float *out) {
    // the pair lists sit at the start of the dynamic LDS, inside the tile's pad
    const size_t lds = std::max(wtd_list_bytes(WAVES, d.N), (size_t)d.wtd_tile.lds_pad);
    hipLaunchKernelGGL((k_ncc_weighted<6, MODE, PREC, WAVES>),
                       dim3((Qn + kLanes - 1) / kLanes, (nslot + WAVES - 1) / WAVES), dim3(kLanes * WAVES), lds, s, d,
                       colour, Qn, nslot, out);
}
template <int MODE>
static void launch_wtd(const Dev &d, hipStream_t s, int colour, int Qn, int nslot, float *out) {
    int w = d.wtd_tile.waves;
    // The refinement's default tile is one wave per hypothesis (5): with 4 its
    // 5 slots take two workgroups per 64 pixels, the second with one busy wave
    // and a second patch fill (refinement launches -5 % at 1600x1200).  The
    // post-pass keeps 4 (5 waves measured 8-14 % slower there).
    static_assert(kRefineSlots == 5, "refinement tile");
    if (MODE == MODE_REFINE && w == 4 && d.wtd_tile.lds_pad == 0) w = kRefineSlots;
    if (d.prec) {
        if (w == 2) wtd_t<MODE, true, 2>(d, s, colour, Qn, nslot, out);
        else if (w == 8) wtd_t<MODE, true, 8>(d, s, colour, Qn, nslot, out);
        else if (w == 5) wtd_t<MODE, true, 5>(d, s, colour, Qn, nslot, out);
        else wtd_t<MODE, true, 4>(d, s, colour, Qn, nslot, out);
    } else {
        if (w == 2) wtd_t<MODE, false, 2>(d, s, colour, Qn, nslot, out);
        else if (w == 8) wtd_t<MODE, false, 8>(d, s, colour, Qn, nslot, out);
        else if (w == 5) wtd_t<MODE, false, 5>(d, s, colour, Qn, nslot, out);
        else wtd_t<MODE, false, 4>(d, s, colour, Qn, nslot, out);
    }
}

void launch_ncc_current(const Dev &d, hipStream_t s) {
    // FIRST_INIT needs every view (top-k selection); the REFINE states only the
    // prior-selected ones (ComputeMultiViewInitialCost, DPE.cu:832-861)
    const uint32_t *mask = d.prm.state == DPE_FIRST_INIT ? nullptr : d.selected;
    if (d.ntap == 6)
        launch_full(d, s, -1, d.P, 1, nullptr, d.planes, mask, d.cur_cost);
    else
        hipLaunchKernelGGL(k_ncc_full_direct, grid_direct(d.P, 1), dim3(256), 0, s, d, -1, d.P, nullptr, d.planes,
                           mask, d.cur_cost, (float4 *)nullptr);
}
void launch_random_init_costs(const Dev &d, hipStream_t s) {
    hipLaunchKernelGGL(k_random_init_costs, grid1(d.P), dim3(256), 0, s, d);
}
void launch_strong_candidates(const Dev &d, int colour, int iter, hipStream_t s) {
    hipLaunchKernelGGL(k_strong_candidates, grid1(d.Q), dim3(256), 0, s, d, colour, iter);
}
void launch_ncc_strong(const Dev &d, int colour, hipStream_t s) {
    const int nslot = d.prm.use_edge ? kStrongSlots : 9;
    float4 *last = d.last_cand + (size_t)colour * kStrongSlots * d.Q;
    if (d.ntap == 6)
        launch_full(d, s, colour, d.Q, nslot, d.cand_pos, d.planes, nullptr, d.cost_mat, true, last);
    else
        hipLaunchKernelGGL(k_ncc_full_direct, grid_direct(d.Q, nslot), dim3(256), 0, s, d, colour, d.Q, d.cand_pos,
                           d.planes, (const uint32_t *)nullptr, d.cost_mat, last);
}
void launch_strong_select(const Dev &d, int colour, int iter, hipStream_t s) {
    hipLaunchKernelGGL(k_strong_select, grid1(d.Q, 64), dim3(64), 0, s, d, colour, iter);
}
void launch_ncc_refine(const Dev &d, int colour, hipStream_t s) {
    if (d.ntap == 6)
        launch_wtd<MODE_REFINE>(d, s, colour, d.Q, kRefineSlots, d.rcost);
    else
        hipLaunchKernelGGL(k_ncc_weighted_direct<MODE_REFINE>, grid_direct(d.Q, kRefineSlots), dim3(256), 0, s, d,
                           colour, d.Q, d.rcost);
}
void launch_strong_accept(const Dev &d, int colour, hipStream_t s) {
    hipLaunchKernelGGL(k_strong_accept, grid1(d.Q), dim3(256), 0, s, d, colour);
}
void launch_ransac(const Dev &d, int iter, hipStream_t s) {
    hipLaunchKernelGGL(k_fit_copy, grid1(d.P), dim3(256), 0, s, d);
    if (d.weak_count > 0) hipLaunchKernelGGL(k_ransac, grid1(d.weak_count, 64), dim3(64), 0, s, d, iter);
}
void launch_weak_half(const Dev &d, int colour, int iter, hipStream_t s) {
    if (d.weak_colour_count[colour] > 0) {
        hipLaunchKernelGGL(k_weak_patch, grid1(d.weak_colour_count[colour], 256), dim3(256), 0, s, d, colour);
        hipLaunchKernelGGL(k_weak_cand, grid1(d.weak_colour_count[colour], kLanes * kWeakCandWaves),
                           dim3(kLanes * kWeakCandWaves), 0, s, d, colour);
        const int n = d.weak_colour_count[colour];
        hipLaunchKernelGGL(k_weak_select, grid1(n, 64), dim3(64), 0, s, d, colour, iter);
        hipLaunchKernelGGL(k_weak_terms<0>, grid1(n, kLanes * kWeakCandWaves), dim3(kLanes * kWeakCandWaves), 0, s,
                           d, colour);
        hipLaunchKernelGGL(k_weak_mid, grid1(n, 64), dim3(64), 0, s, d, colour, iter);
        hipLaunchKernelGGL(k_weak_terms<1>, grid1(n, kLanes * kWeakCandWaves), dim3(kLanes * kWeakCandWaves), 0, s,
                           d, colour);
        hipLaunchKernelGGL(k_weak_final, grid1(n, 64), dim3(64), 0, s, d, colour);
    }
}
void launch_depth_and_normal(const Dev &d, hipStream_t s) {
    hipLaunchKernelGGL(k_depth_and_normal, grid1(d.P), dim3(256), 0, s, d);
}
void launch_filter(const Dev &d, int colour, hipStream_t s) {
    hipLaunchKernelGGL(k_filter, grid1(d.Q), dim3(256), 0, s, d, colour);
}
void launch_post_prep(const Dev &d, hipStream_t s) {
    hipLaunchKernelGGL(k_post_prep, grid1(d.P), dim3(256), 0, s, d);
}
void launch_post_planes(const Dev &d, hipStream_t s) {
    hipLaunchKernelGGL(k_post_planes, grid1(d.P), dim3(256), 0, s, d);
}
void launch_ncc_post(const Dev &d, hipStream_t s) {
    if (d.ntap == 6)
        launch_wtd<MODE_POST>(d, s, -1, d.P, kPostSlots, d.pcost);
    else
        hipLaunchKernelGGL(k_ncc_weighted_direct<MODE_POST>, grid_direct(d.P, kPostSlots), dim3(256), 0, s, d, -1, d.P,
                           d.pcost);
}
void launch_post_final(const Dev &d, hipStream_t s) {
    hipLaunchKernelGGL(k_post_final, grid1(d.P), dim3(256), 0, s, d);
}
void launch_ncc_given(const Dev &d, int num_hyps, const float4 *planes, float *costs, hipStream_t s) {
    if (d.ntap == 6)
        launch_full(d, s, -1, d.P, num_hyps, nullptr, planes, nullptr, costs);
    else
        hipLaunchKernelGGL(k_ncc_full_direct, grid_direct(d.P, num_hyps), dim3(256), 0, s, d, -1, d.P, nullptr, planes,
                           (const uint32_t *)nullptr, costs, (float4 *)nullptr);
}

// An empty launch: HIP loads this file's code object at its first kernel launch
// (tens of ms), so a context does it once when it is created.
__global__ void k_preload_kernels() {}
void preload_kernels(hipStream_t s) { hipLaunchKernelGGL(k_preload_kernels, dim3(1), dim3(1), 0, s); }

}  // namespace dpe
