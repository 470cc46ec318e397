// dpe_adjoint.hip -- the differentiation form: the aggregated multi-view NCC
// cost of each pixel's plane and its gradient with respect to the plane's
// parameters (theta, phi, depth).  The reference has no derivatives; the cost
// is its ComputeBilateralNCCOld (DPE.cu:696-782) folded over views:
//   C(p) = sum_v w_v c_v / sum_v w_v,
//   n = (sin t cos f, sin t sin f, -cos t)      (reference camera frame)
//   d = -n . X(depth), X(depth) = depth K^-1 [x, y, 1]   (GetDistance2Origin)
//   H = Ms (Rr - tr n^T / d) Mr                  (ComputeHomography: Ms = K_src, Mr = K_ref^-1)
// The forward pass is the plain path's window sums (same taps and bits as
// dpe_ncc_cost for the same plane).  The backward pass walks the taps again:
//   dc/ds_t = -(1/W) [ (w_t r_t - w_t mu_r) / sqrt(B S) - A w_t (s_t - mu_s) / (sqrt(B S) S) ]
// (A = covariance, B, S = variances of the normalised sums, W = sum of
// weights), then through the bilinear sample (ds/du, ds/dv), the projection
// q = (u / w, v / w) and H.  A view whose centre leaves the source image, whose
// variance is under the floor or whose cost is clamped to [0, 2] contributes
// its cost and no gradient.
#include "dpe_ncc.h"

namespace dpe {

// Cost of one (pixel, view) and its gradient with respect to nu = n / d.
template <int NT>
__device__ __forceinline__ float ncc_adjoint_view(const Dev &d, const SharedPatch<NT> &S, int lane, int x, int y,
                                                  int v, const float4 &pl, float gnu[3], unsigned long long &taps) {
    gnu[0] = gnu[1] = gnu[2] = 0.0f;
    const ViewConst &vc = d.vc[v];
    float H[9];
    homography(vc, d.cams[0], pl, H);
    const float2 pt = project(H, x, y);
    if (pt.x >= (float)vc.width || pt.x < 0.0f || pt.y >= (float)vc.height || pt.y < 0.0f) return 2.0f;
    float s_s = 0.0f, s_ss = 0.0f, s_rs = 0.0f;
    const int x0 = x + d.tap_off0, y0 = y + d.tap_off0, span = (NT - 1) * d.tap_inc;
    window_sums_any<NT>(d, S, lane, x, y, v, H, refined_division_ok(H, x0, x0 + span, y0, y0 + span), s_s, s_ss,
                        s_rs);
    taps += NT * NT;
    // ncc_from_sums with its intermediates
    const float inv = 1.0f / S.wsum[lane];
    const float mu_r = S.sum_ref[lane] * inv, rr = S.sum_ref_ref[lane] * inv;
    const float mu_s = s_s * inv, ss = s_ss * inv, rs = s_rs * inv;
    const float var_r = rr - mu_r * mu_r, var_s = ss - mu_s * mu_s;
    if (var_r < 1e-5f || var_s < 1e-5f) return 2.0f;
    const float cov = rs - mu_r * mu_s;
    const float vrs = sqrtf(var_r * var_s);
    const float raw = 1.0f - cov / vrs;
    const float c = fmaxf(0.0f, fminf(2.0f, raw));
    if (!(raw > 0.0f && raw < 2.0f)) return c;
    const float alpha = inv / vrs, beta = inv * cov / (vrs * var_s);
    taps += NT * NT;
    // dC/dH accumulated over the taps
    float G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int ki = 0; ki < NT; ++ki) {
        const int rx = x + d.tap_off0 + ki * d.tap_inc;
        const float X = (float)rx;
        for (int kj = 0; kj < NT; ++kj) {
            const int ry = y + d.tap_off0 + kj * d.tap_inc;
            const float Y = (float)ry;
            const float u = fmaf(H[0], X, fmaf(H[1], Y, H[2]));
            const float vv = fmaf(H[3], X, fmaf(H[4], Y, H[5]));
            const float w = fmaf(H[6], X, fmaf(H[7], Y, H[8]));
            const float iw = 1.0f / w;
            const float qx = u * iw, qy = vv * iw;
            const float fx = floorf(qx), fy = floorf(qy);
            const float a = qx - fx, b = qy - fy;
            const int ix = f2i_clamp(fx, -1, d.W - 1), iy = f2i_clamp(fy, -1, d.H - 1);
            const float4 t = d.quad_fmt == QUAD_FP16   ? quad_at<QUAD_FP16>(d, v, ix, iy)
                             : d.quad_fmt == QUAD_BF16 ? quad_at<QUAD_BF16>(d, v, ix, iy)
                                                       : quad_at<QUAD_FP32>(d, v, ix, iy);
            const float top = fmaf(a, t.y - t.x, t.x), bot = fmaf(a, t.w - t.z, t.z);
            const float s = fmaf(b, bot - top, top);
            const float s_u = (t.y - t.x) + b * ((t.w - t.z) - (t.y - t.x));
            const float s_v = bot - top;
            const float2 e = S.wwr[ki * NT + kj][lane];
            const float gt = -(alpha * (e.y - e.x * mu_r) - beta * e.x * (s - mu_s));
            const float cu = gt * s_u * iw, cv = gt * s_v * iw, cw = -(cu * qx + cv * qy);
            G[0] += cu * X; G[1] += cu * Y; G[2] += cu;
            G[3] += cv * X; G[4] += cv * Y; G[5] += cv;
            G[6] += cw * X; G[7] += cw * Y; G[8] += cw;
        }
    }
    // dC/dGm = Ms^T (dC/dH) Mr^T with Ms = K_src (zero skew) and Mr = K_ref^-1
    const float *Ks = vc.Ks;
    const dpe_camera &ref = d.cams[0];
    float T[9];
    for (int j = 0; j < 3; ++j) {
        T[j] = Ks[0] * G[j];
        T[3 + j] = Ks[4] * G[3 + j];
        T[6 + j] = Ks[2] * G[j] + Ks[5] * G[3 + j] + Ks[8] * G[6 + j];
    }
    float U[9];
    for (int r = 0; r < 3; ++r) {
        U[3 * r + 0] = (T[3 * r + 0] - T[3 * r + 2] * ref.K[2]) / ref.K[0];
        U[3 * r + 1] = (T[3 * r + 1] - T[3 * r + 2] * ref.K[5]) / ref.K[4];
        U[3 * r + 2] = T[3 * r + 2];
    }
    // Gm = Rr - tr nu^T
    for (int j = 0; j < 3; ++j) gnu[j] = -(U[j] * vc.tr[0] + U[3 + j] * vc.tr[1] + U[6 + j] * vc.tr[2]);
    return c;
}

// Chain rule from g_nu = dC/dnu (nu = n / d) to (dC/dtheta, dC/dphi, dC/ddepth)
// at the plane n(theta, phi), d = -depth (n . r), r = K^-1 [x, y, 1].
__device__ __forceinline__ float4 grad_theta_phi_depth(const dpe_camera &cam, int x, int y, float st, float ct,
                                                      float sp, float cp, float D, const float4 &n, float dd,
                                                      const float gnu[3]) {
    float4 grad = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (dd == 0.0f) return grad;
    const float id = 1.0f / dd;
    const float gd = -(gnu[0] * n.x + gnu[1] * n.y + gnu[2] * n.z) * id * id;
    const float r0 = ((float)x - cam.K[2]) / cam.K[0], r1 = ((float)y - cam.K[5]) / cam.K[4];
    const float gn0 = gnu[0] * id - gd * D * r0;
    const float gn1 = gnu[1] * id - gd * D * r1;
    const float gn2 = gnu[2] * id - gd * D;
    grad.z = -gd * (n.x * r0 + n.y * r1 + n.z);
    grad.x = gn0 * (ct * cp) + gn1 * (ct * sp) + gn2 * st;
    grad.y = gn0 * (-st * sp) + gn1 * (st * cp);
    return grad;
}

// One workgroup = 64 pixels x 4 waves; wave w takes views 1+w, 5+w, ...; the
// waves' partial sums meet in LDS.
template <int NT>
__global__ __launch_bounds__(kLanes * kWaves) void k_ncc_adjoint(Dev d, const float4 *__restrict__ tpd,
                                                                 const float *__restrict__ view_w,
                                                                 float *__restrict__ cost_out,
                                                                 float4 *__restrict__ grad_out) {
    __shared__ SharedPatch<NT> S;
    __shared__ float red[kWaves][5][kLanes];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int p = blockIdx.x * kLanes + lane;
    const bool valid = p < d.P;
    const int y = valid ? p / d.W : 0, x = valid ? p - y * d.W : 0;
    fill_shared_patch<NT>(d, valid, x, y, S);
    const float4 par = valid ? tpd[p] : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    const float st = sinf(par.x), ct = cosf(par.x), sp = sinf(par.y), cp = cosf(par.y);
    const float D = par.z;
    const dpe_camera &cam = d.cams[0];
    const float4 n = make_float4(st * cp, st * sp, -ct, 0.0f);
    const float dd = distance_to_origin(cam, x, y, D, n);
    const float4 pl = make_float4(n.x, n.y, n.z, dd);
    float acc[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};  // sum w c, sum w, sum w dc/dnu
    unsigned long long hyps = 0, taps = 0;
    if (valid) {
        for (int v = 1 + wave; v < d.N; v += kWaves) {
            const float wv = view_w ? view_w[(size_t)(v - 1) * d.P + p] : 1.0f;
            if (!(wv > 0.0f)) continue;
            float g[3];
            const float c = ncc_adjoint_view<NT>(d, S, lane, x, y, v, pl, g, taps);
            hyps++;
            acc[0] += wv * c;
            acc[1] += wv;
            acc[2] += wv * g[0];
            acc[3] += wv * g[1];
            acc[4] += wv * g[2];
        }
    }
    for (int k = 0; k < 5; ++k) red[wave][k][lane] = acc[k];
    __syncthreads();
    if (wave == 0 && valid) {
        float s[5];
        for (int k = 0; k < 5; ++k) s[k] = red[0][k][lane] + red[1][k][lane] + red[2][k][lane] + red[3][k][lane];
        float C = 2.0f;
        float4 grad = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (s[1] > 0.0f) {
            C = s[0] / s[1];
            const float gnu[3] = {s[2] / s[1], s[3] / s[1], s[4] / s[1]};
            grad = grad_theta_phi_depth(cam, x, y, st, ct, sp, cp, D, n, dd, gnu);
        }
        cost_out[p] = C;
        grad_out[p] = grad;
    }
    count_add(d.counters, kFamFull, hyps, taps);
}

// One Gauss-Newton step for the refinement of the strong half-iteration, from
// the camera-frame plane src[q] to dst[q]: the per-view costs c_v are the
// residuals, weighted by the pixel's view-selection weights w_v, and the step
// solves (J^T W J + lambda diag) delta = -J^T W c in the scaled parameters
// (theta, phi, depth / depth0), J the per-view cost gradients (the adjoint), so
// it vanishes where sum_v w_v c_v^2 is stationary.  Limited to 0.05 rad and
// 3 % of the depth; the refinement's acceptance test (the hypothesis must lower
// the aggregated cost) keeps only improvements.  Pixels without a usable step
// keep the source plane.
template <int NT>
__global__ __launch_bounds__(kLanes * kWaves) void k_gn_step(Dev d, int colour, const float4 *__restrict__ src,
                                                             float4 *__restrict__ dst) {
    constexpr int kAcc = 10;  // A = sum w g g^T (6), b = sum w c g (3), sum w
    __shared__ SharedPatch<NT> S;
    __shared__ float red[kWaves][kAcc][kLanes];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int q = blockIdx.x * kLanes + lane;
    int x = 0, y = 0;
    const bool valid = q < d.Q && colour_pixel(d, colour, q, x, y) && d.cand_pos[q] >= 0;
    fill_shared_patch<NT>(d, valid, x, y, S);
    const dpe_camera &cam = d.cams[0];
    const float4 pl = valid ? src[q] : make_float4(0.0f, 0.0f, -1.0f, 1.0f);
    const int p = x + y * d.W;
    float acc[kAcc];
    for (int k = 0; k < kAcc; ++k) acc[k] = 0.0f;
    unsigned long long hyps = 0, taps = 0;
    if (valid) {
        const uint8_t *vw = d.vw + (size_t)p * kMaxImages;
        for (int v = 1 + wave; v < d.N; v += kWaves) {
            const float wv = (float)vw[v - 1];
            if (!(wv > 0.0f)) continue;
            float g[3];
            const float c = ncc_adjoint_view<NT>(d, S, lane, x, y, v, pl, g, taps);
            hyps++;
            acc[0] += wv * g[0] * g[0];
            acc[1] += wv * g[0] * g[1];
            acc[2] += wv * g[0] * g[2];
            acc[3] += wv * g[1] * g[1];
            acc[4] += wv * g[1] * g[2];
            acc[5] += wv * g[2] * g[2];
            acc[6] += wv * c * g[0];
            acc[7] += wv * c * g[1];
            acc[8] += wv * c * g[2];
            acc[9] += wv;
        }
    }
    for (int k = 0; k < kAcc; ++k) red[wave][k][lane] = acc[k];
    __syncthreads();
    if (wave == 0 && valid) {
        float s[kAcc];
        for (int k = 0; k < kAcc; ++k) s[k] = red[0][k][lane] + red[1][k][lane] + red[2][k][lane] + red[3][k][lane];
        float4 out = pl;
        const float D = depth_from_plane(cam, pl, x, y);
        if (s[9] > 0.0f && D > 0.0f) {
            const float th = acosf(fminf(1.0f, fmaxf(-1.0f, -pl.z)));
            const float ph = atan2f(pl.y, pl.x);
            const float st = sinf(th), ct = cosf(th), sp = sinf(ph), cp = cosf(ph);
            const float4 n = make_float4(st * cp, st * sp, -ct, 0.0f);
            const float dd = distance_to_origin(cam, x, y, D, n);
            // J^T (nu -> scaled parameters) column by column: T(e_i) = J^T e_i
            float Jt[3][3];  // Jt[i] = J^T e_i (a 3-vector over (theta, phi, depth * D))
            for (int i = 0; i < 3; ++i) {
                const float e[3] = {i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f};
                const float4 g = grad_theta_phi_depth(cam, x, y, st, ct, sp, cp, D, n, dd, e);
                Jt[i][0] = g.x;
                Jt[i][1] = g.y;
                Jt[i][2] = g.z * D;
            }
            // N = J^T A J, r = J^T b with A, b in nu coordinates
            const float A[3][3] = {{s[0], s[1], s[2]}, {s[1], s[3], s[4]}, {s[2], s[4], s[5]}};
            const float bnu[3] = {s[6], s[7], s[8]};
            float N[3][3], r[3];
            for (int a = 0; a < 3; ++a) {
                r[a] = Jt[0][a] * bnu[0] + Jt[1][a] * bnu[1] + Jt[2][a] * bnu[2];
                for (int b = 0; b < 3; ++b) {
                    float acc2 = 0.0f;
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) acc2 += Jt[i][a] * A[i][j] * Jt[j][b];
                    N[a][b] = acc2;
                }
            }
            for (int a = 0; a < 3; ++a) N[a][a] *= 1.0f + 1e-3f;  // Levenberg damping
            // solve N delta = -r (Cramer's rule on the 3x3 system)
            const float det = N[0][0] * (N[1][1] * N[2][2] - N[1][2] * N[2][1]) -
                              N[0][1] * (N[1][0] * N[2][2] - N[1][2] * N[2][0]) +
                              N[0][2] * (N[1][0] * N[2][1] - N[1][1] * N[2][0]);
            if (det > 0.0f || det < 0.0f) {
                float delta[3];
                for (int k = 0; k < 3; ++k) {
                    float M[3][3];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) M[a][b] = b == k ? -r[a] : N[a][b];
                    delta[k] = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                                M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                                M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) / det;
                }
                const float lim = fmaxf(fmaxf(fabsf(delta[0]) / 0.05f, fabsf(delta[1]) / 0.05f), fabsf(delta[2]) / 0.03f);
                if (lim > 1.0f)
                    for (int k = 0; k < 3; ++k) delta[k] /= lim;
                if (delta[0] == delta[0] && delta[1] == delta[1] && delta[2] == delta[2]) {
                    const float th2 = th + delta[0], ph2 = ph + delta[1], D2 = D * (1.0f + delta[2]);
                    const float4 n2 = make_float4(sinf(th2) * cosf(ph2), sinf(th2) * sinf(ph2), -cosf(th2), 0.0f);
                    out = make_float4(n2.x, n2.y, n2.z, distance_to_origin(cam, x, y, D2, n2));
                }
            }
        }
        dst[q] = out;
    }
    count_add(d.counters, kFamRefine, hyps, taps);
}

void launch_gn_step(const Dev &d, int colour, const float4 *src, float4 *dst, hipStream_t s) {
    // the steps use the plain window sums (on the call's quads) in every form
    Dev dp = d;
    dp.prec = 0;
    hipLaunchKernelGGL(k_gn_step<6>, dim3((d.Q + kLanes - 1) / kLanes), dim3(kLanes * kWaves), 0, s, dp, colour, src,
                       dst);
}

void launch_ncc_adjoint(const Dev &d, const float4 *tpd, const float *view_w, float *cost, float4 *grad,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_ncc_adjoint<6>, dim3((d.P + kLanes - 1) / kLanes), dim3(kLanes * kWaves), 0, s, d, tpd,
                       view_w, cost, grad);
}

// An empty launch: HIP loads this file's code object at its first kernel launch
// (tens of ms), so a context does it once when it is created.
__global__ void k_preload_adjoint() {}
void preload_adjoint(hipStream_t s) { hipLaunchKernelGGL(k_preload_adjoint, dim3(1), dim3(1), 0, s); }

}  // namespace dpe
