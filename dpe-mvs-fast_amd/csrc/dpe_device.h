// dpe_device.h -- device-side numerics and per-pixel geometry of the gfx950 path.
//
// Numerics contract (the HIP path and the CPU oracle follow it independently):
//   * IEEE fp32, correctly rounded divide and sqrt, no contraction; the tap
//     projection multiplies by the correctly rounded 1/w (see project())
//     (-ffp-contract=off): fused multiply-adds only where written as fmaf
//     (NCC tap loop, homography projection, bilinear lerp, polynomials).
//   * exp: Cody-Waite reduction by ln2 (16-bit high part) and a degree-7 Taylor
//     polynomial in float, degree 13 in double; exact power-of-two scaling.
//   * sin/cos: three-part Cody-Waite reduction by pi/2 and Taylor polynomials.
//   * rsqrtf (DPE.cu:275) is 1.0f / sqrtf.
//   * random numbers: Philox4x32-10 (Salmon et al., SC'11), key = 64-bit seed,
//     counter = (draw block, pixel, phase, call); uniform = ((x>>8)+1) * 2^-24.
//   * textures: clamp-to-edge, exact fp32 bilinear weights, sample position =
//     tex2D coordinate - 0.5 (DPE.cpp:1083-1094 asks for wrap + linear filter on
//     unnormalised coordinates, where wrap does not apply).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/dpe_hip.h"

#define DPE_HD __host__ __device__ __forceinline__

namespace dpe {

constexpr int kMaxImages = DPE_MAX_IMAGES;
constexpr int kNeighbourNum = DPE_NEIGHBOUR_NUM;
constexpr int kMaxTaps = 11;  // taps per window dimension held in registers

// The dispatcher deals workgroups round-robin over the 8 XCDs (workgroup b runs
// on XCD b % 8), and each XCD has its own L2.  xcd_block(b, nb) renumbers the
// nb workgroups of a 1-D grid so that XCD x works through one contiguous range
// of logical workgroups (XCDs 0..nb%8-1 take one more); the mapping is a
// bijection on [0, nb).  Used where neighbouring workgroups read overlapping
// data that would otherwise be fetched into eight L2s.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// ----------------------------------------------------------------- numerics
__device__ __forceinline__ float d_expf(float x) {
    if (x != x) return x;
    if (!(x > -87.0f)) return 0.0f;
    if (x > 88.0f) return __builtin_inff();
    const float k = rintf(x * 1.44269502162933349609375f);
    float r = fmaf(-k, 0.693145751953125f, x);
    r = fmaf(-k, 1.428606820309417232e-06f, r);
    float p = 1.0f / 5040.0f;
    p = fmaf(p, r, 1.0f / 720.0f);
    p = fmaf(p, r, 1.0f / 120.0f);
    p = fmaf(p, r, 1.0f / 24.0f);
    p = fmaf(p, r, 1.0f / 6.0f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    const int ki = (int)k;
    return p * __uint_as_float((uint32_t)(ki + 127) << 23);
}

__device__ __forceinline__ double d_exp(double x) {
    if (x != x) return x;
    if (!(x > -708.0)) return 0.0;
    if (x > 709.0) return __builtin_inf();
    const double k = rint(x * 1.4426950408889634074);
    double r = fma(-k, 6.93147180369123816490e-01, x);
    r = fma(-k, 1.90821492927058770002e-10, r);
    double p = 1.0 / 6227020800.0;
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    const long long ki = (long long)k;
    return p * __longlong_as_double((long long)((unsigned long long)(ki + 1023) << 52));
}

__device__ __forceinline__ void d_sincosf(float x, float *s, float *c) {
    const float q = rintf(x * 0.636619746685028076171875f);
    float r = fmaf(-q, 1.5703125f, x);
    r = fmaf(-q, 4.83751296997070312500e-04f, r);
    r = fmaf(-q, 7.549790126404332e-08f, r);
    const float r2 = r * r;
    float ps = 1.0f / 362880.0f;
    ps = fmaf(ps, r2, -1.0f / 5040.0f);
    ps = fmaf(ps, r2, 1.0f / 120.0f);
    ps = fmaf(ps, r2, -1.0f / 6.0f);
    ps = fmaf(ps * r2, r, r);
    float pc = -1.0f / 3628800.0f;
    pc = fmaf(pc, r2, 1.0f / 40320.0f);
    pc = fmaf(pc, r2, -1.0f / 720.0f);
    pc = fmaf(pc, r2, 1.0f / 24.0f);
    pc = fmaf(pc, r2, -0.5f);
    pc = fmaf(pc, r2, 1.0f);
    const int qi = ((int)q) & 3;
    float sv, cv;
    if (qi == 0) { sv = ps; cv = pc; }
    else if (qi == 1) { sv = pc; cv = -ps; }
    else if (qi == 2) { sv = -ps; cv = -pc; }
    else { sv = -pc; cv = ps; }
    *s = sv;
    *c = cv;
}

__device__ __forceinline__ float d_asinf_small(float z) {
    const float z2 = z * z;
    float p = 13749310575.0f / 1880240947200.0f;
    p = fmaf(p, z2, 654729075.0f / 78033715200.0f);
    p = fmaf(p, z2, 34459425.0f / 3530096640.0f);
    p = fmaf(p, z2, 2027025.0f / 175472640.0f);
    p = fmaf(p, z2, 135135.0f / 9676800.0f);
    p = fmaf(p, z2, 10395.0f / 599040.0f);
    p = fmaf(p, z2, 945.0f / 42240.0f);
    p = fmaf(p, z2, 105.0f / 3456.0f);
    p = fmaf(p, z2, 15.0f / 336.0f);
    p = fmaf(p, z2, 3.0f / 40.0f);
    p = fmaf(p, z2, 1.0f / 6.0f);
    p = fmaf(p, z2, 1.0f);
    return p * z;
}

__device__ __forceinline__ float d_acosf(float x) {
    if (x != x || x > 1.0f || x < -1.0f) return __builtin_nanf("");
    if (x >= 0.5f) return 2.0f * d_asinf_small(sqrtf((1.0f - x) * 0.5f));
    if (x <= -0.5f) return 3.14159274101257324f - 2.0f * d_asinf_small(sqrtf((1.0f + x) * 0.5f));
    return 1.57079637050628662f - d_asinf_small(x);
}

// --------------------------------------------------------------- Philox4x32-10
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

enum : uint32_t { PH_GEN_NEIGHBOURS = 1, PH_RANDOM_INIT = 2 };
__device__ __forceinline__ uint32_t ph_strong(int it) { return 8u + 4u * (uint32_t)it; }
__device__ __forceinline__ uint32_t ph_ransac(int it) { return 9u + 4u * (uint32_t)it; }
__device__ __forceinline__ uint32_t ph_weak(int it) { return 10u + 4u * (uint32_t)it; }

struct Rng {
    uint32_t k0, k1, pixel, phase, call, n, cur;
    uint32_t blk[4];
    __device__ Rng(uint64_t seed, uint32_t pixel_, uint32_t phase_, uint32_t call_)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), pixel(pixel_), phase(phase_), call(call_), n(0),
          cur(0xFFFFFFFFu) {}
    __device__ uint32_t next() {
        const uint32_t b = n >> 2;
        if (b != cur) {
            philox4x32_10(b, pixel, phase, call, k0, k1, blk);
            cur = b;
        }
        const uint32_t w = n & 3;
        n++;
        return w == 0 ? blk[0] : (w == 1 ? blk[1] : (w == 2 ? blk[2] : blk[3]));
    }
    __device__ float uniform() { return (float)((next() >> 8) + 1u) * 5.9604644775390625e-08f; }
};

// x % d (d >= 1) with one multiply-high: for d >= 2, m = floor((2^32 - 1) / d) + 1
// exceeds 2^32 / d by less than 1, so mulhi(x, m) is floor(x / d) or one more
// (x (m - 2^32 / d) / 2^32 < 1) and a wrapped remainder is put back by adding d.
// Equal to x % d for every 32-bit x and d = 2..64 (checked exhaustively).
struct UMod {
    uint32_t d, m;
    __device__ explicit UMod(uint32_t d_) : d(d_), m(d_ > 1u ? 0xFFFFFFFFu / d_ + 1u : 0u) {}
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
        const uint32_t r = x - __umulhi(x, m) * d;
        return d == 1u ? 0u : (r >= d ? r + d : r);
    }
};

// ---------------------------------------------------------------- geometry
struct ViewConst {
    float Rr[9];  // R_relative = R_src R_ref^T   (DPE.cu:471-479)
    float tr[3];  // t_relative = R_src (C_ref - C_src) (DPE.cu:480-485)
    float Ks[9];
    float base;   // |c_ref - c_src| (DPE.cu:2644-2649)
    int width, height;
    int pad[2];
};

__device__ __forceinline__ float4 mk4(float x, float y, float z, float w) { return make_float4(x, y, z, w); }

// Get3DPoint, DPE.cu:313-325
__device__ __forceinline__ void get3d(const dpe_camera &cam, int px, int py, float depth, float X[3]) {
    X[0] = depth * ((float)px - cam.K[2]) / cam.K[0];
    X[1] = depth * ((float)py - cam.K[5]) / cam.K[4];
    X[2] = depth;
}

// GetViewDirection, DPE.cu:327-339
__device__ __forceinline__ float4 view_direction(const dpe_camera &cam, int px, int py, float depth) {
    float X[3];
    get3d(cam, px, py, depth, X);
    const float norm = sqrtf(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    return mk4(X[0] / norm, X[1] / norm, X[2] / norm, 0.0f);
}

// GetDistance2Origin, DPE.cu:341-346
__device__ __forceinline__ float distance_to_origin(const dpe_camera &cam, int px, int py, float depth,
                                                    const float4 &n) {
    float X[3];
    get3d(cam, px, py, depth, X);
    return -(n.x * X[0] + n.y * X[1] + n.z * X[2]);
}

// ComputeDepthfromPlaneHypothesis, DPE.cu:360-363
__device__ __forceinline__ float depth_from_plane(const dpe_camera &cam, const float4 &pl, int px, int py) {
    return -pl.w * cam.K[0] /
           (((float)px - cam.K[2]) * pl.x + (cam.K[0] / cam.K[4]) * ((float)py - cam.K[5]) * pl.y + cam.K[0] * pl.z);
}

__device__ __forceinline__ void normalize3(float4 *v) {
    const float n2 = v->x * v->x + v->y * v->y + v->z * v->z;
    const float inv = 1.0f / sqrtf(n2);
    v->x *= inv; v->y *= inv; v->z *= inv;
}
__device__ __forceinline__ void normalize2(float2 *v) {
    const float n2 = v->x * v->x + v->y * v->y;
    const float inv = 1.0f / sqrtf(n2);
    v->x *= inv; v->y *= inv;
}

// TransformNormal (camera -> world), DPE.cu:528-536
__device__ __forceinline__ float4 to_world(const dpe_camera &cam, const float4 &p) {
    return mk4(cam.R[0] * p.x + cam.R[3] * p.y + cam.R[6] * p.z, cam.R[1] * p.x + cam.R[4] * p.y + cam.R[7] * p.z,
               cam.R[2] * p.x + cam.R[5] * p.y + cam.R[8] * p.z, p.w);
}
// TransformNormal2RefCam (world -> camera), DPE.cu:538-546
__device__ __forceinline__ float4 to_camera(const dpe_camera &cam, const float4 &p) {
    return mk4(cam.R[0] * p.x + cam.R[1] * p.y + cam.R[2] * p.z, cam.R[3] * p.x + cam.R[4] * p.y + cam.R[5] * p.z,
               cam.R[6] * p.x + cam.R[7] * p.y + cam.R[8] * p.z, p.w);
}

// Plane-dependent half of ComputeHomography (DPE.cu:487-516); the
// plane-independent half (R_relative, t_relative) is precomputed per view.
// The reference's build (--use_fast_math, CMakeLists.txt:26) divides as
// x * rcp(y); here x * RN(1/y), with one correctly rounded reciprocal per plane
// (1/d) and per focal length, instead of 19 correctly rounded quotients
// (~290 -> ~90 instructions per window, the loop-invariant reciprocals hoisted).
__device__ __forceinline__ void homography(const ViewConst &v, const dpe_camera &ref, const float4 &pl, float H[9]) {
    const float iw = 1.0f / pl.w, ik0 = 1.0f / ref.K[0], ik4 = 1.0f / ref.K[4];
    float G[9];
    G[0] = v.Rr[0] - v.tr[0] * pl.x * iw;
    G[1] = v.Rr[1] - v.tr[0] * pl.y * iw;
    G[2] = v.Rr[2] - v.tr[0] * pl.z * iw;
    G[3] = v.Rr[3] - v.tr[1] * pl.x * iw;
    G[4] = v.Rr[4] - v.tr[1] * pl.y * iw;
    G[5] = v.Rr[5] - v.tr[1] * pl.z * iw;
    G[6] = v.Rr[6] - v.tr[2] * pl.x * iw;
    G[7] = v.Rr[7] - v.tr[2] * pl.y * iw;
    G[8] = v.Rr[8] - v.tr[2] * pl.z * iw;
    float t[9];
    t[0] = G[0] * ik0;
    t[1] = G[1] * ik4;
    t[2] = -G[0] * ref.K[2] * ik0 - G[1] * ref.K[5] * ik4 + G[2];
    t[3] = G[3] * ik0;
    t[4] = G[4] * ik4;
    t[5] = -G[3] * ref.K[2] * ik0 - G[4] * ref.K[5] * ik4 + G[5];
    t[6] = G[6] * ik0;
    t[7] = G[7] * ik4;
    t[8] = -G[6] * ref.K[2] * ik0 - G[7] * ref.K[5] * ik4 + G[8];
    H[0] = v.Ks[0] * t[0] + v.Ks[2] * t[6];
    H[1] = v.Ks[0] * t[1] + v.Ks[2] * t[7];
    H[2] = v.Ks[0] * t[2] + v.Ks[2] * t[8];
    H[3] = v.Ks[4] * t[3] + v.Ks[5] * t[6];
    H[4] = v.Ks[4] * t[4] + v.Ks[5] * t[7];
    H[5] = v.Ks[4] * t[5] + v.Ks[5] * t[8];
    H[6] = v.Ks[8] * t[6];
    H[7] = v.Ks[8] * t[7];
    H[8] = v.Ks[8] * t[8];
}

// Double-precision view constants and homography (precision form).  The
// plane-induced homography H = Ks (Rr - tr n^T / d) Kr^-1 is factored as
// A - b c^T with A = Ks Rr Kr^-1 and b = Ks tr per view (host, double) and
// c = Kr^-T n / d per plane: 9 FP64 FMAs per view, each entry rounded to fp32 once.
struct ViewConst64 {
    double A[9], b[3];
    double kinv[4];  // reference 1/fx, 1/fy, cx/fx, cy/fy
};
__device__ __forceinline__ void homography64(const ViewConst64 &v, const float4 &pl, float H[9]) {
    const double inv_d = 1.0 / (double)pl.w;
    const double nx = pl.x, ny = pl.y, nz = pl.z;
    const double c[3] = {nx * v.kinv[0] * inv_d, ny * v.kinv[1] * inv_d,
                         fma(-ny, v.kinv[3], fma(-nx, v.kinv[2], nz)) * inv_d};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) H[3 * r + k] = (float)fma(-v.b[r], c[k], v.A[3 * r + k]);
}

// ComputeCorrespondingPoint, DPE.cu:519-526, projection written with fmaf.  The
// divisions by w are multiplications by the correctly rounded 1/w, shared by
// both coordinates: (u * r, v * r), r = RN(1 / w).  The reference's build
// (--use_fast_math, CMakeLists.txt:26) does not divide correctly rounded
// either; this form costs one reciprocal per tap instead of two quotients.
__device__ __forceinline__ float2 project(const float H[9], int px, int py) {
    const float x = (float)px, y = (float)py;
    const float u = fmaf(H[0], x, fmaf(H[1], y, H[2]));
    const float v = fmaf(H[3], x, fmaf(H[4], y, H[5]));
    const float w = fmaf(H[6], x, fmaf(H[7], y, H[8]));
    const float r = 1.0f / w;
    return make_float2(u * r, v * r);
}

// ---- exact reciprocal in three instructions
// On gfx950 one Newton step from the hardware reciprocal estimate,
// y1 = fma(fma(-w, y, 1), y, y) with y = v_rcp_f32(w), equals the correctly
// rounded 1 / w for every float w with |w| in [2^-40, 2^40), both signs: checked
// exhaustively (all 1.34e9 such values) by dpe_rcp_check, which the GPU tests
// run.  IEEE 1.0f / w as the compiler emits it takes about 12 instructions
// (div_scale x2, rcp, 7 FMA/MUL, div_fmas, div_fixup).
__device__ __forceinline__ float recip_refined(float w) {
    const float y = __builtin_amdgcn_rcpf(w);
    return fmaf(fmaf(-w, y, 1.0f), y, y);
}

// project() with the three-instruction reciprocal; equal to project() wherever
// refined_division_ok() holds for a window containing (px, py).
__device__ __forceinline__ float2 project_refined(const float H[9], int px, int py) {
    const float x = (float)px, y = (float)py;
    const float u = fmaf(H[0], x, fmaf(H[1], y, H[2]));
    const float v = fmaf(H[3], x, fmaf(H[4], y, H[5]));
    const float w = fmaf(H[6], x, fmaf(H[7], y, H[8]));
    const float r = recip_refined(w);
    return make_float2(u * r, v * r);
}

// True when recip_refined gives RN(1 / w) at every tap (X, Y) of the window
// [x0, x1] x [y0, y1] (|X|, |Y| <= 2^17):
//  - each of H[0..5] is 0 or has magnitude in [2^-60, 2^20] (u and v are then
//    multiples of 2^-83 below 2^38; kept so the window test also bounds the
//    projected coordinates);
//  - w keeps one sign with |w| in [2^-40, 2^40]: w is affine in (X, Y), so its
//    extremes sit at the window corners, and computed values differ from exact
//    ones by less than 2^-22 M, M = |H6| max|X| + |H7| max|Y| + |H8|.
// Hence every computed w lies in the checked range [2^-40, 2^40).
__device__ __forceinline__ bool refined_division_ok(const float H[9], int x0, int x1, int y0, int y1) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float a = fabsf(H[i]);
        ok = ok && (a == 0.0f || (a >= 0x1p-60f && a <= 0x1p20f));
    }
    const float X0 = (float)x0, X1 = (float)x1, Y0 = (float)y0, Y1 = (float)y1;
    const float i0 = fmaf(H[7], Y0, H[8]), i1 = fmaf(H[7], Y1, H[8]);
    const float w00 = fmaf(H[6], X0, i0), w10 = fmaf(H[6], X1, i0);
    const float w01 = fmaf(H[6], X0, i1), w11 = fmaf(H[6], X1, i1);
    const float lo = fminf(fminf(w00, w10), fminf(w01, w11));
    const float hi = fmaxf(fmaxf(w00, w10), fmaxf(w01, w11));
    const float M = fabsf(H[6]) * fmaxf(fabsf(X0), fabsf(X1)) + fabsf(H[7]) * fmaxf(fabsf(Y0), fabsf(Y1)) + fabsf(H[8]);
    const float mn = lo > 0.0f ? lo : -hi;  // min |w| when the corners share a sign
    return ok && (lo > 0.0f || hi < 0.0f) && M <= 0x1p39f && mn >= 0x1p-19f * M + 0x1p-39f;
}

// (int) of a clamped float; NaN gives lo.  |lo|, |hi| <= 2^24 (exact floats).
__device__ __forceinline__ int f2i_clamp(float f, int lo, int hi) {
    return (int)fminf(fmaxf(f, (float)lo), (float)hi);
}

__device__ __forceinline__ float texel(const float *im, int W, int H, int x, int y) {
    x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
    y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
    return im[(size_t)y * W + x];
}

// Bilinear sample from the quad layout: quad (ix, iy) for ix in [-1, W-1],
// iy in [-1, H-1] holds the four clamped texels (x,y), (x+1,y), (x,y+1),
// (x+1,y+1), so one 16-byte load serves one tap.
__device__ __forceinline__ float bilinear_quad(const float4 *__restrict__ q, int W, int H, float u, float v) {
    const float fx = floorf(u), fy = floorf(v);
    const float a = u - fx, b = v - fy;
    const int ix = f2i_clamp(fx, -1, W - 1);
    const int iy = f2i_clamp(fy, -1, H - 1);
    // 32-bit byte offset: (W + 1) * (H + 1) * 16 < 2^32 is checked at set-up
    const uint32_t off = ((uint32_t)(iy + 1) * (uint32_t)(W + 1) + (uint32_t)(ix + 1)) * 16u;
    const float4 t = *(const float4 *)((const char *)q + off);
    const float top = fmaf(a, t.y - t.x, t.x);
    const float bot = fmaf(a, t.w - t.z, t.z);
    return fmaf(b, bot - top, top);
}

// Final NCC from the window sums (DPE.cu:760-778).
__device__ __forceinline__ float ncc_from_sums(float sum_ref, float sum_ref_ref, float sum_src, float sum_src_src,
                                               float sum_ref_src, float wsum) {
    const float inv = 1.0f / wsum;
    sum_ref *= inv;
    sum_ref_ref *= inv;
    sum_src *= inv;
    sum_src_src *= inv;
    sum_ref_src *= inv;
    const float var_ref = sum_ref_ref - sum_ref * sum_ref;
    const float var_src = sum_src_src - sum_src * sum_src;
    if (var_ref < 1e-5f || var_src < 1e-5f) return 2.0f;
    const float covar = sum_ref_src - sum_ref * sum_src;
    const float var_rs = sqrtf(var_ref * var_src);
    return fmaxf(0.0f, fminf(2.0f, 1.0f - covar / var_rs));
}
// ncc_from_sums with the reference-only terms computed once per pixel:
// inv = 1 / wsum, mean_ref = sum_ref * inv, var_ref = sum_ref_ref * inv - mean_ref^2
// (the same operations, so the same result).
__device__ __forceinline__ float ncc_from_ref_terms(float mean_ref, float var_ref, float inv, float sum_src,
                                                    float sum_src_src, float sum_ref_src) {
    sum_src *= inv;
    sum_src_src *= inv;
    sum_ref_src *= inv;
    const float var_src = sum_src_src - sum_src * sum_src;
    if (var_ref < 1e-5f || var_src < 1e-5f) return 2.0f;
    const float covar = sum_ref_src - mean_ref * sum_src;
    const float var_rs = sqrtf(var_ref * var_src);
    return fmaxf(0.0f, fminf(2.0f, 1.0f - covar / var_rs));
}

__device__ __forceinline__ int isSet(uint32_t input, uint32_t n) { return (input >> n) & 1; }

}  // namespace dpe
