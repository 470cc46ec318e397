// dpe_ncc.h -- the shared pieces of the NCC kernels (dpe_kernels.hip,
// dpe_adjoint.hip): checkerboard pixel indexing, wave-reduced counters, the
// workgroup-shared reference patch and the source sums of one window.
#pragma once
#include "dpe_state.h"

namespace dpe {

__device__ __forceinline__ bool colour_pixel(const Dev &d, int colour, int q, int &x, int &y) {
    if (colour < 0) {
        if (q >= d.P) return false;
        y = q / d.W;
        x = q - y * d.W;
        return true;
    }
    if (q >= d.Q) return false;
    y = q / d.Wh;
    const int k = q - y * d.Wh;
    x = 2 * k + ((y + colour) & 1);
    return y < d.half_rows && x < d.W;
}

// Wave-reduced NCC counts of one kernel family into the wave's counter stripe;
// every lane of the wave must call.
__device__ __forceinline__ void count_add(unsigned long long *counters, int family, unsigned long long hyps,
                                          unsigned long long taps) {
    if (!counters) return;
    const uint32_t wave_id = (blockIdx.x + gridDim.x * blockIdx.y) * ((blockDim.x + 63) / 64) + threadIdx.x / 64;
    unsigned long long *c = counters + (size_t)(wave_id % kCounterStripes) * kCounterStride + 2 * family;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        hyps += __shfl_down(hyps, off, 64);
        taps += __shfl_down(taps, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (hyps) atomicAdd(c, hyps);
        if (taps) atomicAdd(c + 1, taps);
    }
}


// ------------------------------------------------------------ NCC windows
// ---- workgroup-shared reference patches (the hot-path layout)
// A workgroup covers 64 consecutive pixels (one per lane) and 4 waves.  The
// waves fill the 64 patches into LDS together (tap k of lane l at [k][l], so a
// wave reads 64 consecutive float2: no bank conflicts), then every wave loops
// over hypothesis slots of the same 64 pixels, reading the weights from LDS.
// Taps of one window row evaluated together (registers vs. loads in flight).
#ifndef DPE_TAP_UNROLL
#define DPE_TAP_UNROLL 6
#endif
// Bilinear taps read the quad layout of the source view: the 4 texels of one
// cell per load, 16 B in fp32 or 8 B in fp16.  The fp16 layout is used when
// every texel of the call converts to fp16 and back unchanged (8-bit images,
// and their factor-2 pyramid levels, do), so the samples are bit-identical.
// (Measured at 1600x1200 with 10 sources: two 8-B pairs from the plain fp32
// image, a quarter of the footprint, ran 4-15 % slower than the fp32 quads.)
// Quad formats (Dev::quad_fmt, QuadFormat): 16-bit formats hold 2 texels per 32-bit word.
template <int FMT>
__device__ __forceinline__ float texel16(uint32_t h) {
    if constexpr (FMT == QUAD_BF16) return __builtin_bit_cast(float, h << 16);
    else return (float)__builtin_bit_cast(_Float16, (uint16_t)h);
}
template <int FMT>
__device__ __forceinline__ float4 quad_at(const Dev &d, int v, int ix, int iy) {
    // 32-bit byte offsets: (W + 1) * (H + 1) * 16 < 2^32 is checked at set-up; both
    // factors are below 2^24 (W <= 65536), so the full-rate 24-bit multiply serves
    const uint32_t cell = __umul24((uint32_t)(iy + 1), (uint32_t)(d.W + 1)) + (uint32_t)(ix + 1);
    if constexpr (FMT != QUAD_FP32) {
        const uint2 h = *(const uint2 *)((const char *)(d.quads_h + (size_t)v * d.qs) + cell * 8u);
        // word 0 = (t00, t01), word 1 = (t10, t11)
        return make_float4(texel16<FMT>(h.x & 0xFFFFu), texel16<FMT>(h.y & 0xFFFFu), texel16<FMT>(h.x >> 16),
                           texel16<FMT>(h.y >> 16));
    } else {
        return *(const float4 *)((const char *)(d.quads + (size_t)v * d.qs) + cell * 16u);
    }
}
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
template <int FMT>
__device__ __forceinline__ float sample_src(const Dev &d, int v, float u, float w) {
    const float fx = floorf(u), fy = floorf(w);
    const float a = u - fx, b = w - fy;
    const float4 t = quad_at<FMT>(d, v, f2i_clamp(fx, -1, d.W - 1), f2i_clamp(fy, -1, d.H - 1));
    const float top = fmaf(a, t.y - t.x, t.x);
    const float bot = fmaf(a, t.w - t.z, t.z);
    return fmaf(b, bot - top, top);
}
__device__ __forceinline__ float sample_src_any(const Dev &d, int v, float u, float w) {
    if (d.quad_fmt == QUAD_FP16) return sample_src<QUAD_FP16>(d, v, u, w);
    if (d.quad_fmt == QUAD_BF16) return sample_src<QUAD_BF16>(d, v, u, w);
    return sample_src<QUAD_FP32>(d, v, u, w);
}

// Clamp-free bilinear gather (Dev::fast_gather, a window inside the source):
// the window's taps lie in [1/64, W - 1 - 1/64] x [1/64, H - 1 - 1/64], so
// floor is truncation, the cell index is a 24-bit multiply-add of the truncated
// coordinates, and u - floor(u) is exact, which v_fract_f32 returns (5
// instructions where floor, subtract and a float cell offset took 6); the
// texel differences t.y - t.x, t.w - t.z are taken in packed fp16, exact on the
// quarter grid, so the sample equals sample_src's bit for bit.
// base = cell (0, 0) of view 0 (wave-uniform), voff = the view's byte offset
// from it: all views' quads lie within 4 GiB of base (checked at set-up), so
// the gather takes a scalar base and one 32-bit offset per tap even where the
// view varies across the wave.
__device__ __forceinline__ float sample_fast(const char *__restrict__ base, uint32_t voff, uint32_t row_cells, float u,
                                             float w) {
    const float a = __builtin_amdgcn_fractf(u), b = __builtin_amdgcn_fractf(w);
    // one v_mad_u32_u24 for the cell index (the compiler otherwise folds voff
    // into the sum and spends a third instruction on the address)
    uint32_t cell;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(cell) : "v"((uint32_t)w), "s"(row_cells), "v"((uint32_t)u));
    const uint2 h = *(const uint2 *)(base + (voff + cell * 8u));
    const half2_t p = __builtin_bit_cast(half2_t, h.x);  // (t.x, t.z)
    const half2_t q = __builtin_bit_cast(half2_t, h.y);  // (t.y, t.w)
    const half2_t dd = q - p;
    const float top = fmaf(a, (float)dd.x, (float)p.x);
    const float bot = fmaf(a, (float)dd.y, (float)p.y);
    return fmaf(b, bot - top, top);
}
// Every tap of the window [x0, x1] x [y0, y1] projects to [1/64, W-1-1/64] x
// [1/64, H-1-1/64]: the projection is linear-fractional with w of one sign
// (refined_division_ok), so its extremes over the window are at the corner
// taps, and the margin covers rounding.  No tap then needs the clamp.  Called
// only where refined_division_ok holds for the window, so the corners are
// projected with the refined reciprocal (equal to project() there).
__device__ __forceinline__ bool window_inside(const Dev &d, const float H[9], int x0, int x1, int y0, int y1) {
    const float m = 1.0f / 64.0f;
    const float ux = (float)(d.W - 1) - m, uy = (float)(d.H - 1) - m;
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float2 p = project_refined(H, (c & 1) ? x1 : x0, (c & 2) ? y1 : y0);
        ok = ok && p.x >= m && p.x <= ux && p.y >= m && p.y <= uy;
    }
    return ok;
}

constexpr int kLanes = 64;
constexpr int kWaves = 4;

template <int NT>
struct SharedPatch {
    float2 wwr[NT * NT][kLanes];  // (w, w*ref) per tap
    float sum_ref[kLanes], sum_ref_ref[kLanes], wsum[kLanes];
    // precision form: sums of w (ref - rc) and w (ref - rc)^2, rc = the centre texel
    float sum_refc[kLanes], sum_refc2[kLanes];
    // the reference-only terms of ncc_from_sums: 1 / wsum, the weighted mean and
    // variance of the reference patch
    float inv_wsum[kLanes], mean_ref[kLanes], var_ref[kLanes];
};

// Dynamic LDS of a weighted-sum launch over N images, per wave: the round's
// 64 (cost, geometric cost) results, per view the 64-bit ballot of its pixels
// and its first list index (16 bits), and up to 64 (N - 1) 16-bit (pixel,
// view) list entries; 8-byte aligned.
__host__ __device__ constexpr size_t wtd_wave_bytes(int N) {
    return ((size_t)kLanes * sizeof(float2) + (size_t)(N - 1) * (8 + 2) + (size_t)kLanes * (N - 1) * 2 + 7) / 8 * 8;
}
__host__ __device__ constexpr size_t wtd_list_bytes(int waves, int N) { return (size_t)waves * wtd_wave_bytes(N); }
// Dynamic LDS of a k_ncc_full launch per wave: one 16-bit (slot, pixel lane)
// item per pixel of each of the wave's slots; 8-byte aligned.
__host__ __device__ constexpr size_t full_list_bytes(int nslot, int waves) {
    return ((size_t)(nslot + waves - 1) / waves * kLanes * 2 + 7) / 8 * 8;
}
static_assert(sizeof(SharedPatch<6>) + wtd_list_bytes(8, kMaxImages) <= 65536, "weighted tile exceeds 64 KiB of LDS");

// WAVES = waves of the workgroup (the tile's slot width); they fill the taps together.
template <int NT, int WAVES = kWaves>
__device__ __forceinline__ void fill_shared_patch(const Dev &d, bool valid, int x, int y, SharedPatch<NT> &S) {
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const float sc = d.prm.sigma_color;
    const float cden = 2.0f * sc * sc;
    const float rc = valid ? texel(d.ref, d.W, d.H, x, y) : 0.0f;
    for (int k = wave; k < NT * NT; k += WAVES) {
        const int ki = k / NT, kj = k - ki * NT;
        float2 e = make_float2(0.0f, 0.0f);
        if (valid) {
            const float ref = texel(d.ref, d.W, d.H, x + d.tap_off0 + ki * d.tap_inc, y + d.tap_off0 + kj * d.tap_inc);
            const float w = d_expf(d.spatial[k] - fabsf(ref - rc) / cden);
            // precision form: w (ref - rc), the reference centred at its centre texel
            e = make_float2(w, w * (d.prec ? ref - rc : ref));
        }
        S.wwr[k][lane] = e;
    }
    __syncthreads();
    if (wave == 0 && valid) {
        // hypothesis-independent sums in the reference's row-partial order
        float s_r = 0.0f, s_rr = 0.0f, s_w = 0.0f, s_c = 0.0f, s_c2 = 0.0f;
        for (int ki = 0; ki < NT; ++ki) {
            float r_r = 0.0f, r_rr = 0.0f, r_w = 0.0f;
#pragma unroll
            for (int kj = 0; kj < NT; ++kj) {
                const float2 e = S.wwr[ki * NT + kj][lane];
                const float ref = texel(d.ref, d.W, d.H, x + d.tap_off0 + ki * d.tap_inc, y + d.tap_off0 + kj * d.tap_inc);
                if (d.prec) {
                    s_c += e.y;
                    s_c2 = fmaf(e.y, ref - rc, s_c2);
                } else {
                    r_r = fmaf(e.x, ref, r_r);
                    r_rr = fmaf(e.y, ref, r_rr);
                }
                r_w = r_w + e.x;
            }
            s_r += r_r;
            s_rr += r_rr;
            s_w += r_w;
        }
        S.sum_ref[lane] = s_r;
        S.sum_ref_ref[lane] = s_rr;
        S.wsum[lane] = s_w;
        S.sum_refc[lane] = s_c;
        S.sum_refc2[lane] = s_c2;
        const float inv = 1.0f / s_w, mr = s_r * inv;
        S.inv_wsum[lane] = inv;
        S.mean_ref[lane] = mr;
        S.var_ref[lane] = s_rr * inv - mr * mr;
    }
    __syncthreads();
}

// Source sums of one window (DPE.cu:728-758) with the reference weights from
// LDS; REFINED: the projection uses the refined division (exact, see
// refined_division_ok).
template <int NT, bool REFINED, int FMT, bool FAST = false>
__device__ __forceinline__ void shared_window_sums(const Dev &d, const SharedPatch<NT> &S, int lane, int x, int y,
                                                   int v, const float H[9], float &s_s, float &s_ss, float &s_rs) {
    // FAST: cell (0, 0) of view 0's fp16 quads, view v's byte offset from it, the cells per row
    const char *base = (const char *)d.quads_h + (size_t)(d.W + 2) * 8;
    const uint32_t voff = (uint32_t)v * (uint32_t)d.qs * 8u;
    const uint32_t row_cells = (uint32_t)(d.W + 1);
#pragma unroll 1
    for (int ki = 0; ki < NT; ++ki) {
        float r_s = 0.0f, r_ss = 0.0f, r_rs = 0.0f;
        const int rx = x + d.tap_off0 + ki * d.tap_inc;
#pragma unroll DPE_TAP_UNROLL
        for (int kj = 0; kj < NT; ++kj) {
            const int ry = y + d.tap_off0 + kj * d.tap_inc;
            const float2 sp = REFINED ? project_refined(H, rx, ry) : project(H, rx, ry);
            const float src = FAST ? sample_fast(base, voff, row_cells, sp.x, sp.y) : sample_src<FMT>(d, v, sp.x, sp.y);
            const float2 e = S.wwr[ki * NT + kj][lane];
            r_s = fmaf(e.x, src, r_s);
            r_ss = fmaf(e.x * src, src, r_ss);
            r_rs = fmaf(e.y, src, r_rs);
        }
        s_s += r_s;
        s_ss += r_ss;
        s_rs += r_rs;
    }
}

// Precision form: the source sums shifted by the sample at the window centre's
// projection (s' = s - sc) and the reference by the centre texel (LDS holds
// w (r - rc)), so the one-pass variances E[x^2] - E[x]^2 act on centred values
// and do not cancel; 16-bit quads, fp32 accumulation.
template <int NT, bool REFINED, int FMT>
__device__ __forceinline__ void shifted_window_sums(const Dev &d, const SharedPatch<NT> &S, int lane, int x, int y,
                                                    int v, const float H[9], float sc, float &s_s, float &s_ss,
                                                    float &s_rs) {
#pragma unroll 1
    for (int ki = 0; ki < NT; ++ki) {
        float r_s = 0.0f, r_ss = 0.0f, r_rs = 0.0f;
        const int rx = x + d.tap_off0 + ki * d.tap_inc;
#pragma unroll DPE_TAP_UNROLL
        for (int kj = 0; kj < NT; ++kj) {
            const int ry = y + d.tap_off0 + kj * d.tap_inc;
            const float2 sp = REFINED ? project_refined(H, rx, ry) : project(H, rx, ry);
            const float src = sample_src<FMT>(d, v, sp.x, sp.y) - sc;
            const float2 e = S.wwr[ki * NT + kj][lane];
            r_s = fmaf(e.x, src, r_s);
            r_ss = fmaf(e.x * src, src, r_ss);
            r_rs = fmaf(e.y, src, r_rs);
        }
        s_s += r_s;
        s_ss += r_ss;
        s_rs += r_rs;
    }
}

// The window sums with the projection and quad layout this window and call
// allow; `refined` = refined_division_ok of the window.
template <int NT>
__device__ __forceinline__ void window_sums_any(const Dev &d, const SharedPatch<NT> &S, int lane, int x, int y, int v,
                                                const float H[9], bool refined, float &s_s, float &s_ss,
                                                float &s_rs) {
    const int x0 = x + d.tap_off0, y0 = y + d.tap_off0, span = (NT - 1) * d.tap_inc;
    if (d.fast_gather && refined && window_inside(d, H, x0, x0 + span, y0, y0 + span)) {
        shared_window_sums<NT, true, QUAD_FP16, true>(d, S, lane, x, y, v, H, s_s, s_ss, s_rs);
    } else if (d.quad_fmt == QUAD_FP16) {
        if (refined) shared_window_sums<NT, true, QUAD_FP16>(d, S, lane, x, y, v, H, s_s, s_ss, s_rs);
        else shared_window_sums<NT, false, QUAD_FP16>(d, S, lane, x, y, v, H, s_s, s_ss, s_rs);
    } else if (d.quad_fmt == QUAD_BF16) {
        if (refined) shared_window_sums<NT, true, QUAD_BF16>(d, S, lane, x, y, v, H, s_s, s_ss, s_rs);
        else shared_window_sums<NT, false, QUAD_BF16>(d, S, lane, x, y, v, H, s_s, s_ss, s_rs);
    } else {
        if (refined) shared_window_sums<NT, true, QUAD_FP32>(d, S, lane, x, y, v, H, s_s, s_ss, s_rs);
        else shared_window_sums<NT, false, QUAD_FP32>(d, S, lane, x, y, v, H, s_s, s_ss, s_rs);
    }
}

// The precision form's NCC of one window: FP64 homography, shifted sums.
template <int NT>
__device__ __forceinline__ float ncc_precise(const Dev &d, const SharedPatch<NT> &S, int lane, int x, int y, int v,
                                          const float4 &pl, unsigned long long &taps) {
    const ViewConst &vc = d.vc[v];
    float H[9];
    homography64(d.vc64[v], pl, H);
    const float2 pt = project(H, x, y);
    if (pt.x >= (float)vc.width || pt.x < 0.0f || pt.y >= (float)vc.height || pt.y < 0.0f) return 2.0f;
    const float sc = sample_src_any(d, v, pt.x, pt.y);
    const int x0 = x + d.tap_off0, y0 = y + d.tap_off0, span = (NT - 1) * d.tap_inc;
    const bool refined = refined_division_ok(H, x0, x0 + span, y0, y0 + span);
    float s_s = 0.0f, s_ss = 0.0f, s_rs = 0.0f;
    if (d.quad_fmt == QUAD_BF16) {
        if (refined) shifted_window_sums<NT, true, QUAD_BF16>(d, S, lane, x, y, v, H, sc, s_s, s_ss, s_rs);
        else shifted_window_sums<NT, false, QUAD_BF16>(d, S, lane, x, y, v, H, sc, s_s, s_ss, s_rs);
    } else {
        if (refined) shifted_window_sums<NT, true, QUAD_FP16>(d, S, lane, x, y, v, H, sc, s_s, s_ss, s_rs);
        else shifted_window_sums<NT, false, QUAD_FP16>(d, S, lane, x, y, v, H, sc, s_s, s_ss, s_rs);
    }
    taps += NT * NT;
    return ncc_from_sums(S.sum_refc[lane], S.sum_refc2[lane], s_s, s_ss, s_rs, S.wsum[lane]);
}

// ComputeBilateralNCCOld (DPE.cu:696-782) with the reference half from LDS;
// PREC: the precision form's window (the kernels are instantiated per form).
template <int NT, bool PREC>
__device__ __forceinline__ float ncc_shared(const Dev &d, const SharedPatch<NT> &S, int lane, int x, int y, int v,
                                            const float4 &pl, unsigned long long &taps) {
    if constexpr (PREC) return ncc_precise<NT>(d, S, lane, x, y, v, pl, taps);
    const ViewConst &vc = d.vc[v];
    float H[9];
    homography(vc, d.cams[0], pl, H);
    const int x0 = x + d.tap_off0, y0 = y + d.tap_off0, span = (NT - 1) * d.tap_inc;
    const bool refined = refined_division_ok(H, x0, x0 + span, y0, y0 + span);
    // the centre lies in the window's box when the last tap offset is >= 0, so
    // the refined reciprocal is exact there as well
    float2 pt;
    if (refined && span >= -d.tap_off0) pt = project_refined(H, x, y);
    else pt = project(H, x, y);
    if (pt.x >= (float)vc.width || pt.x < 0.0f || pt.y >= (float)vc.height || pt.y < 0.0f) return 2.0f;
    float s_s = 0.0f, s_ss = 0.0f, s_rs = 0.0f;
    window_sums_any<NT>(d, S, lane, x, y, v, H, refined, s_s, s_ss, s_rs);
    taps += NT * NT;
    return ncc_from_ref_terms(S.mean_ref[lane], S.var_ref[lane], S.inv_wsum[lane], s_s, s_ss, s_rs);
}


}  // namespace dpe
