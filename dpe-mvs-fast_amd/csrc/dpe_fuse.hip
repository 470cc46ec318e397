// dpe_fuse.hip -- depth-map fusion (RunFusion, DPE.cpp:1387-1548) on gfx950.
//
// The reference fuses on one CPU thread: reference views in pair order, pixels in
// raster order, and every accepted point masks the source pixels it used, which
// later pixels then skip (DPE.cpp:1474, 1497, 1528).  Within one reference view a
// pixel's result depends only on the masks set by earlier views and by earlier
// pixels of the same view, i.e. on pixels with a smaller raster index.  The GPU
// evaluates all pixels of a view in parallel and iterates to the fixpoint of
//     state_k = eval_k(masks of earlier views + claims of accepted pixels j < k),
// where a claim is the smallest raster index of an accepted pixel that used a
// source pixel.  Dependencies only point to smaller indices, so the iteration
// reaches the unique fixpoint -- the sequential result -- after at most one more
// round than the longest dependency chain (a few rounds in practice).  Accepted
// pixels are then compacted in raster order with a prefix sum, which reproduces
// the reference's point order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cfloat>

#include "dpe_fuse.h"

namespace dpe {

// Fusion geometry in scalar type T: float is the reference's arithmetic; double
// is the precision form's FP64 reprojection check (camera entries widened, the
// same operations in the same order).
template <typename T> struct FVec3 { T x, y, z; };
template <typename T> struct FVec2 { T x, y; };

// Get3DPointonWorld (host), DPE.cpp:1337-1361: the centre comes from R and t.
template <typename T>
__device__ __forceinline__ FVec3<T> fuse_point_world(int x, int y, T depth, const dpe_camera &cam) {
    const float *R = cam.R, *K = cam.K, *t = cam.t;
    FVec3<T> X, Tr, Cc;
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
__device__ __forceinline__ float3 fuse_point_world(int x, int y, float depth, const dpe_camera &cam) {
    const FVec3<float> X = fuse_point_world<float>(x, y, depth, cam);
    return make_float3(X.x, X.y, X.z);
}

// ProjectCamera, DPE.cpp:1363-1373
template <typename T>
__device__ __forceinline__ void fuse_project(const FVec3<T> &X, const dpe_camera &cam, FVec2<T> &pt, T &depth) {
    const float *R = cam.R, *K = cam.K, *t0 = cam.t;
    FVec3<T> t;
    t.x = (T)R[0] * X.x + (T)R[1] * X.y + (T)R[2] * X.z + (T)t0[0];
    t.y = (T)R[3] * X.x + (T)R[4] * X.y + (T)R[5] * X.z + (T)t0[1];
    t.z = (T)R[6] * X.x + (T)R[7] * X.y + (T)R[8] * X.z + (T)t0[2];
    depth = (T)K[6] * t.x + (T)K[7] * t.y + (T)K[8] * t.z;
    pt.x = ((T)K[0] * t.x + (T)K[1] * t.y + (T)K[2] * t.z) / depth;
    pt.y = ((T)K[3] * t.x + (T)K[4] * t.y + (T)K[5] * t.z) / depth;
}
__device__ __forceinline__ void fuse_project(const float3 &X, const dpe_camera &cam, float2 &pt, float &depth) {
    FVec2<float> p;
    fuse_project<float>(FVec3<float>{X.x, X.y, X.z}, cam, p, depth);
    pt = make_float2(p.x, p.y);
}

// int(v + 0.5f) (DPE.cpp:1491-1492); NaN and huge values map to -1.
template <typename T>
__device__ __forceinline__ int fuse_round(T v) {
    const T r = v + (T)0.5;
    if (!(r > (T)-1.0e9 && r < (T)1.0e9)) return -1;
    return (int)r;
}

// One pixel of one reference view (DPE.cpp:1475-1541 up to the point output);
// T = double for the precision form's FP64 reprojection check.
template <typename T>
__global__ void k_fuse_eval(FuseDev f, int vi, const int *__restrict__ used_old, const uint8_t *__restrict__ acc_old,
                            int *__restrict__ used_new, uint8_t *__restrict__ acc_new, int *__restrict__ changed) {
    const FuseViewDev &rv = f.views[vi];
    const int P = rv.width * rv.height;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int r = p / rv.width, c = p - r * rv.width;
    const long rp = rv.offset + p;
    uint8_t accepted = 0;
    int used[kMaxImages];
    for (int j = 0; j < rv.num_src; ++j) used[j] = -1;
    const float ref_depth = f.depth[rp];
    if (f.mask[rp] != 1 && ref_depth > 0.0f) {
        const float rn0 = f.normal[3 * rp], rn1 = f.normal[3 * rp + 1], rn2 = f.normal[3 * rp + 2];
        const FVec3<T> X = fuse_point_world<T>(c, r, (T)ref_depth, rv.cam);
        int num_consistent = 0;
        float dyn = 0.0f;
        for (int j = 0; j < rv.num_src; ++j) {
            const FuseViewDev &sv = f.views[rv.src[j]];
            FVec2<T> pt;
            T proj_depth;
            fuse_project<T>(X, sv.cam, pt, proj_depth);
            const int src_r = fuse_round<T>(pt.y);
            const int src_c = fuse_round<T>(pt.x);
            if (src_c >= 0 && src_c < sv.width && src_r >= 0 && src_r < sv.height) {
                const long sp = sv.offset + (long)src_r * sv.width + src_c;
                if (f.mask[sp] == 1 || f.claim[sp] < p) continue;
                const float src_depth = f.depth[sp];
                if (src_depth <= 0.0f) continue;
                const FVec3<T> tX = fuse_point_world<T>(src_c, src_r, (T)src_depth, sv.cam);
                FVec2<T> tp;
                fuse_project<T>(tX, rv.cam, tp, proj_depth);
                // sqrt(pow(c - x, 2) + pow(r - y, 2)) evaluates in double (DPE.cpp:1506)
                const T dx = (T)c - tp.x, dy = (T)r - tp.y;
                const float reproj = (float)sqrt((double)dx * (double)dx + (double)dy * (double)dy);
                const float rel = (float)(fabs(proj_depth - (T)ref_depth) / (T)ref_depth);
                const float dot = rn0 * f.normal[3 * sp] + rn1 * f.normal[3 * sp + 1] + rn2 * f.normal[3 * sp + 2];
                float angle = d_acosf(dot);
                if (angle != angle) angle = 0.0f;  // GetAngle, DPE.cpp:1379-1382
                if (reproj < f.max_reproj && rel < f.max_rel && angle < f.max_angle) {
                    used[j] = (int)(sp - sv.offset);
                    const float idx = reproj + 200 * rel + angle * 10;
                    dyn += d_expf(-idx);
                    num_consistent++;
                }
            }
        }
        const float factor = f.states[rp] == DPE_WEAK ? f.weak_factor : f.strong_factor;
        accepted = (num_consistent >= f.min_consistent && num_consistent >= 1 && dyn > factor * num_consistent) ? 1 : 0;
    }
    bool diff = acc_old[p] != accepted;
    acc_new[p] = accepted;
    for (int j = 0; j < rv.num_src; ++j) {
        const int u = accepted ? used[j] : -1;
        diff = diff || used_old[(size_t)j * P + p] != u;
        used_new[(size_t)j * P + p] = u;
    }
    if (diff) atomicOr(changed, 1);
}

__global__ void k_fuse_fill_claims(FuseDev f, int vi) {
    const FuseViewDev &rv = f.views[vi];
    const int k = blockIdx.y;
    const FuseViewDev &sv = f.views[rv.src[k]];
    const int P = sv.width * sv.height;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) f.claim[sv.offset + p] = 0x7FFFFFFF;
}

__global__ void k_fuse_claim(FuseDev f, int vi, const int *__restrict__ used, const uint8_t *__restrict__ acc) {
    const FuseViewDev &rv = f.views[vi];
    const int P = rv.width * rv.height;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !acc[p]) return;
    for (int j = 0; j < rv.num_src; ++j) {
        const int u = used[(size_t)j * P + p];
        if (u >= 0) atomicMin(&f.claim[f.views[rv.src[j]].offset + u], p);
    }
}

// Point output and mask update of the accepted pixels (DPE.cpp:1520-1541).
template <typename T>
__global__ void k_fuse_commit(FuseDev f, int vi, const int *__restrict__ used, const uint8_t *__restrict__ acc,
                              const int *__restrict__ index, dpe_point *__restrict__ out, long base) {
    const FuseViewDev &rv = f.views[vi];
    const int P = rv.width * rv.height;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !acc[p]) return;
    const int r = p / rv.width, c = p - r * rv.width;
    const long rp = rv.offset + p;
    const FVec3<T> X = fuse_point_world<T>(c, r, (T)f.depth[rp], rv.cam);
    float col[3] = {(float)f.color[3 * rp], (float)f.color[3 * rp + 1], (float)f.color[3 * rp + 2]};
    int n = 0;
    for (int j = 0; j < rv.num_src; ++j) {
        const int u = used[(size_t)j * P + p];
        if (u < 0) continue;
        const long sp = f.views[rv.src[j]].offset + u;
        f.mask[sp] = 1;
        col[0] += f.color[3 * sp];
        col[1] += f.color[3 * sp + 1];
        col[2] += f.color[3 * sp + 2];
        n++;
    }
    col[0] /= (n + 1);
    col[1] /= (n + 1);
    col[2] /= (n + 1);
    dpe_point o;
    o.x = (float)X.x;
    o.y = (float)X.y;
    o.z = (float)X.z;
    o.b = (uint8_t)col[0];
    o.g = (uint8_t)col[1];
    o.r = (uint8_t)col[2];
    o.pad = 0;
    out[base + index[p]] = o;
}

__global__ void k_fuse_flags(const uint8_t *__restrict__ acc, int *__restrict__ flags, int P) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P) flags[p] = acc[p];
}

void launch_fuse_eval(const FuseDev &f, int vi, int P, const int *used_old, const uint8_t *acc_old, int *used_new,
                      uint8_t *acc_new, int *changed, hipStream_t s) {
    if (f.fp64)
        hipLaunchKernelGGL(k_fuse_eval<double>, dim3((P + 255) / 256), dim3(256), 0, s, f, vi, used_old, acc_old,
                           used_new, acc_new, changed);
    else
        hipLaunchKernelGGL(k_fuse_eval<float>, dim3((P + 255) / 256), dim3(256), 0, s, f, vi, used_old, acc_old,
                           used_new, acc_new, changed);
}
void launch_fuse_fill_claims(const FuseDev &f, int vi, int num_src, hipStream_t s) {
    if (num_src > 0) hipLaunchKernelGGL(k_fuse_fill_claims, dim3(1024, num_src), dim3(256), 0, s, f, vi);
}
void launch_fuse_claim(const FuseDev &f, int vi, int P, const int *used, const uint8_t *acc, hipStream_t s) {
    hipLaunchKernelGGL(k_fuse_claim, dim3((P + 255) / 256), dim3(256), 0, s, f, vi, used, acc);
}
hipError_t fuse_scan_temp_bytes(int P, size_t *bytes) {
    *bytes = 0;
    return hipcub::DeviceScan::ExclusiveSum(nullptr, *bytes, (int *)nullptr, (int *)nullptr, P);
}
hipError_t launch_fuse_compact(const FuseDev &f, int vi, int P, const int *used, const uint8_t *acc, int *flags,
                               int *index, void *temp, size_t temp_bytes, dpe_point *out, long base, hipStream_t s) {
    hipLaunchKernelGGL(k_fuse_flags, dim3((P + 255) / 256), dim3(256), 0, s, acc, flags, P);
    const hipError_t e = hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, flags, index, P, s);
    if (e != hipSuccess) return e;
    if (f.fp64)
        hipLaunchKernelGGL(k_fuse_commit<double>, dim3((P + 255) / 256), dim3(256), 0, s, f, vi, used, acc, index, out,
                           base);
    else
        hipLaunchKernelGGL(k_fuse_commit<float>, dim3((P + 255) / 256), dim3(256), 0, s, f, vi, used, acc, index, out,
                           base);
    return hipGetLastError();
}

// ------------------------------------------------------------------ TAT fusions
// RunFusion_TAT_Intermediate (DPE.cpp:1550-1718) and RunFusion_TAT_advanced
// (DPE.cpp:1720-1867).  The reference keeps one measurement record per source
// across the pixels of a reference view and overwrites it only where the source
// pixel is in bounds, unmasked and has depth (DPE.cpp:1656-1682): the record a
// pixel p reads is the one taken at the last such pixel q <= p in raster order.
// Within one view only the view's own mask changes and no view is its own
// source, so every validity test reads masks fixed before the view starts.
// The GPU writes per source the key j*P + p where source j is valid at p (-1
// elsewhere), takes one running maximum over the concatenated sources (a key
// below j*P in segment j means "no valid pixel yet", FLT_MAX in the reference),
// recomputes the record at that pixel and decides every pixel in parallel.

// DPE.cpp:1643-1647: the pixel takes part (block >= 128, depth > 0)
__device__ __forceinline__ bool tat_active(const FuseDev &f, long rp) {
    if (f.block && f.block[rp] < 128) return false;
    return f.depth[rp] > 0.0f;
}

// One source record, DPE.cpp:1656-1682; false where the reference leaves it.
__device__ __forceinline__ bool tat_measure(const FuseDev &f, const FuseViewDev &rv, int j, int c, int r,
                                            float ref_depth, long rp, const float3 &X, float &dist, float &rel,
                                            float &angle, long &sp) {
    const FuseViewDev &sv = f.views[rv.src[j]];
    float2 pt;
    float proj_depth;
    fuse_project(X, sv.cam, pt, proj_depth);
    const int src_r = fuse_round(pt.y);
    const int src_c = fuse_round(pt.x);
    if (!(src_c >= 0 && src_c < sv.width && src_r >= 0 && src_r < sv.height)) return false;
    sp = sv.offset + (long)src_r * sv.width + src_c;
    if (f.mask[sp] == 1) return false;
    const float src_depth = f.depth[sp];
    if (src_depth <= 0.0f) return false;
    const float3 tX = fuse_point_world(src_c, src_r, src_depth, sv.cam);
    float2 tp;
    fuse_project(tX, rv.cam, tp, proj_depth);
    const float dx = (float)c - tp.x, dy = (float)r - tp.y;
    dist = (float)sqrt((double)dx * (double)dx + (double)dy * (double)dy);
    rel = fabsf(proj_depth - ref_depth) / ref_depth;
    const float dot = f.normal[3 * rp] * f.normal[3 * sp] + f.normal[3 * rp + 1] * f.normal[3 * sp + 1] +
                      f.normal[3 * rp + 2] * f.normal[3 * sp + 2];
    angle = d_acosf(dot);
    if (angle != angle) angle = 0.0f;  // GetAngle, DPE.cpp:1375-1385
    return true;
}

__global__ void k_tat_keys(FuseDev f, int vi, int *__restrict__ keys) {
    const FuseViewDev &rv = f.views[vi];
    const int P = rv.width * rv.height;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int r = p / rv.width, c = p - r * rv.width;
    const long rp = rv.offset + p;
    const bool active = tat_active(f, rp);
    const float ref_depth = f.depth[rp];
    const float3 X = fuse_point_world(c, r, active ? ref_depth : 1.0f, rv.cam);
    for (int j = 0; j < rv.num_src; ++j) {
        float dist, rel, angle;
        long sp;
        const bool valid = active && tat_measure(f, rv, j, c, r, ref_depth, rp, X, dist, rel, angle, sp);
        keys[(size_t)j * P + p] = valid ? j * P + p : -1;
    }
}

// k-of-n decision, DPE.cpp:1684-1710 (intermediate) and 1836-1855 (advanced).
// Source j passes at k when its record is under all thresholds of k; bit k-2 of
// pass[j] holds that.  The first k with >= k passing sources accepts.
__global__ void k_tat_decide(FuseDev f, int vi, const int *__restrict__ last, uint8_t *__restrict__ acc,
                             uint32_t *__restrict__ colour) {
    const FuseViewDev &rv = f.views[vi];
    const int P = rv.width * rv.height;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const long rp = rv.offset + p;
    uint8_t accepted = 0;
    uint32_t bgr = 0;
    if (tat_active(f, rp)) {
        const int ns = rv.num_src;
        uint32_t pass[kMaxImages];
        uint32_t scol[kMaxImages];
        for (int j = 0; j < ns; ++j) {
            pass[j] = 0;
            scol[j] = 0;
            const int q = last[(size_t)j * P + p];
            if (q < j * P) continue;
            const int qp = q - j * P;
            const int qr = qp / rv.width, qc = qp - qr * rv.width;
            const long qrp = rv.offset + qp;
            const float qd = f.depth[qrp];
            const float3 X = fuse_point_world(qc, qr, qd, rv.cam);
            float dist = FLT_MAX, rel = FLT_MAX, angle = FLT_MAX;
            long sp = 0;
            tat_measure(f, rv, j, qc, qr, qd, qrp, X, dist, rel, angle, sp);  // valid: the key says so
            for (int k = 2; k <= ns; ++k) {
                const bool ok = dist < (float)k * f.dist_base && rel < (float)k * f.depth_base &&
                                (!f.tat_intermediate || angle < ((float)k * f.angle_grad + f.angle_base));
                pass[j] |= (ok ? 1u : 0u) << (k - 2);
            }
            scol[j] = (uint32_t)f.color[3 * sp] | ((uint32_t)f.color[3 * sp + 1] << 8) |
                      ((uint32_t)f.color[3 * sp + 2] << 16);
        }
        for (int k = 2; k <= ns; ++k) {
            int n = 0;
            for (int j = 0; j < ns; ++j) n += (pass[j] >> (k - 2)) & 1u;
            if (n < k) continue;
            float col[3] = {(float)f.color[3 * rp], (float)f.color[3 * rp + 1], (float)f.color[3 * rp + 2]};
            if (f.tat_intermediate) {
                for (int j = 0; j < ns; ++j) {
                    if (!((pass[j] >> (k - 2)) & 1u)) continue;
                    col[0] += (float)(scol[j] & 0xFF);
                    col[1] += (float)((scol[j] >> 8) & 0xFF);
                    col[2] += (float)((scol[j] >> 16) & 0xFF);
                }
                col[0] /= (n + 1.0f);
                col[1] /= (n + 1.0f);
                col[2] /= (n + 1.0f);
            }
            bgr = (uint32_t)(uint8_t)col[0] | ((uint32_t)(uint8_t)col[1] << 8) | ((uint32_t)(uint8_t)col[2] << 16);
            accepted = 1;
            break;
        }
    }
    acc[p] = accepted;
    colour[p] = bgr;
    if (accepted) f.mask[rp] = 1;  // DPE.cpp:1707; read only by later views
}

__global__ void k_tat_commit(FuseDev f, int vi, const uint8_t *__restrict__ acc, const uint32_t *__restrict__ colour,
                             const int *__restrict__ index, dpe_point *__restrict__ out) {
    const FuseViewDev &rv = f.views[vi];
    const int P = rv.width * rv.height;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !acc[p]) return;
    const int r = p / rv.width, c = p - r * rv.width;
    const float3 X = fuse_point_world(c, r, f.depth[rv.offset + p], rv.cam);
    dpe_point o;
    o.x = X.x;
    o.y = X.y;
    o.z = X.z;
    o.b = (uint8_t)(colour[p] & 0xFF);
    o.g = (uint8_t)((colour[p] >> 8) & 0xFF);
    o.r = (uint8_t)((colour[p] >> 16) & 0xFF);
    o.pad = 0;
    out[index[p]] = o;
}

void launch_tat_keys(const FuseDev &f, int vi, int P, int *keys, hipStream_t s) {
    hipLaunchKernelGGL(k_tat_keys, dim3((P + 255) / 256), dim3(256), 0, s, f, vi, keys);
}
hipError_t tat_scan_temp_bytes(int n, size_t *bytes) {
    *bytes = 0;
    return hipcub::DeviceScan::InclusiveScan(nullptr, *bytes, (const int *)nullptr, (int *)nullptr, hipcub::Max(), n);
}
hipError_t launch_tat_last(const int *keys, int *last, int n, void *temp, size_t temp_bytes, hipStream_t s) {
    return hipcub::DeviceScan::InclusiveScan(temp, temp_bytes, keys, last, hipcub::Max(), n, s);
}
void launch_tat_decide(const FuseDev &f, int vi, int P, const int *last, uint8_t *acc, uint32_t *colour,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_tat_decide, dim3((P + 255) / 256), dim3(256), 0, s, f, vi, last, acc, colour);
}
hipError_t launch_tat_compact(const FuseDev &f, int vi, int P, const uint8_t *acc, const uint32_t *colour, int *flags,
                              int *index, void *temp, size_t temp_bytes, dpe_point *out, hipStream_t s) {
    hipLaunchKernelGGL(k_fuse_flags, dim3((P + 255) / 256), dim3(256), 0, s, acc, flags, P);
    const hipError_t e = hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, flags, index, P, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tat_commit, dim3((P + 255) / 256), dim3(256), 0, s, f, vi, acc, colour, index, out);
    return hipGetLastError();
}

// An empty launch: HIP loads this file's code object at its first kernel launch
// (tens of ms), so a context does it once when it is created.
__global__ void k_preload_fuse() {}
void preload_fuse(hipStream_t s) { hipLaunchKernelGGL(k_preload_fuse, dim3(1), dim3(1), 0, s); }

}  // namespace dpe
