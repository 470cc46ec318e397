// dpe_capi.hip -- the C-ABI of libdpe_hip.so (include/dpe_hip.h).
//
// One dpe_ctx = one GPU, one HIP stream, a pool of device buffers that grows to
// the largest problem seen (the reference's GpuWorkspace, DPE.cpp:775-837) and a
// pool of hipEvents for per-kernel timing.  dpe_process restates the launch
// schedule of DPE::RunPatchMatch (DPE.cu:3130-3262) over the kernels of
// dpe_kernels.hip.  Host-side set-up follows DPE::InuputInitialization and
// CudaSpaceInitialization (DPE.cpp:934-1185); every exit() there is a status.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dpe_state.h"
#include "dpe_fuse.h"
#include "dpe_ncc.h"

// Autotuner record of one problem class (image size, source count, window
// slots, precision): launch times of every tile, then the chosen tiles.
struct dpe_tune {
    int W = 0, H = 0, NV = 0, use_edge = 0, prec = 0;
    bool decided = false;
    int full = 0, wtd = 0;                 // chosen tile indices
    std::vector<std::vector<float>> ms[2];  // [0] k_ncc_full, [1] k_ncc_weighted: per tile, per launch
};

struct dpe_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    struct Buf {
        void *ptr = nullptr;
        size_t bytes = 0;
    };
    std::vector<Buf> bufs;
    std::vector<hipEvent_t> events;
    dpe_stats stats{};
    size_t reserved = 0;
    int lds_per_cu = 0, lds_per_block = 0;
    std::vector<dpe_tune> tunes;
    hipGraphExec_t graph = nullptr;  // the last captured schedule (graph form), updated in place when it fits
    int graph_rebuilds = 0, graph_updates = 0;
    // allocation-free calls (dpe_ncc_cost_noalloc): while `arena` is set, buffer()
    // carves every buffer of the call from this device's scratch array instead of
    // the pool, and never allocates
    char *arena = nullptr;
    size_t arena_bytes = 0, arena_used = 0;
};

static thread_local std::string g_create_error;

// Scratch of the allocation-free entry point (dpe_ncc_cost_noalloc): every GPU
// the library runs on holds its own copy of this array from the library's first
// context on that device until the process ends.  Host code never uses the name
// `scratch` as an address (on the host it is only a stand-in for the symbol):
// dpe_create looks up each device's address once, with that device current, and
// the calls use that address.  Calls that overlap on one GPU share the array and
// race; each call rewrites what it reads.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static char *g_scratch_ptr[kMaxScratchDevices];

namespace {

using namespace dpe;

enum BufId {
    B_IMAGES, B_QUADS, B_GDEPTH, B_CAMS, B_VC, B_SPATIAL, B_PLANES, B_COSTS, B_SELECTED, B_VW, B_FIT, B_WEAK,
    B_WEAKREL, B_NEAREST, B_NEIGHBOURS, B_EDGENEIGH, B_COMPLEX, B_LABELBOUND, B_RADIUS, B_NMAP, B_EDGE, B_EDGELOW,
    B_LABEL, B_CANDPOS, B_COSTMAT, B_NSEL, B_RPLANE, B_RCOST, B_NOWPLANE, B_NOWDEPTH, B_NOWCOST, B_WNORM, B_POSTOP,
    B_POSTBL, B_POSTDISP, B_POSTWN, B_POSTFLAGS, B_PCOST, B_COUNTERS, B_GIVEN_PLANES, B_GIVEN_COSTS,
    B_F_VIEWS, B_F_DEPTH, B_F_NORMAL, B_F_STATES, B_F_COLOR, B_F_MASK, B_F_CLAIM, B_F_USED0, B_F_USED1, B_F_ACC0,
    B_F_ACC1, B_F_FLAGS, B_F_INDEX, B_F_TEMP, B_F_POINTS, B_F_CHANGED, B_POSTPLANES, B_WEAKLIST, B_ADJ_WEIGHTS,
    B_ADJ_GRADS, B_QUADS_H, B_FLAG, B_WPRE, B_WSUM, B_VC64, B_WEAKCOUNTS, B_WEAKTEMP, B_NOWROW, B_CURCOST,
    B_CURVALID, B_LASTCAND, B_WEAKCOST, B_WEAKST, B_WEAKTERMS, B_CANDALIAS, B_COUNT
};

struct Fail {
    dpe_status st;
};

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                \
            throw Fail{DPE_ERR_HIP};                                                     \
        }                                                                                \
    } while (0)

void *buffer(dpe_ctx *ctx, int id, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (ctx->arena) {
        const size_t off = (ctx->arena_used + 255) & ~(size_t)255;
        if (off > ctx->arena_bytes || bytes > ctx->arena_bytes - off) {
            ctx->err = "the call needs more than the " + std::to_string(ctx->arena_bytes) +
                       " bytes of the allocation-free scratch";
            throw Fail{DPE_ERR_UNSUPPORTED};
        }
        ctx->arena_used = off + bytes;
        return ctx->arena + off;
    }
    if (ctx->bufs.size() < (size_t)B_COUNT) ctx->bufs.resize(B_COUNT);
    auto &b = ctx->bufs[id];
    if (b.bytes < bytes) {
        if (b.ptr) {
            HIPCHK(hipStreamSynchronize(ctx->stream));
            HIPCHK(hipFree(b.ptr));
            ctx->reserved -= b.bytes;
            b.ptr = nullptr;
            b.bytes = 0;
        }
        HIPCHK(hipMalloc(&b.ptr, bytes));
        b.bytes = bytes;
        ctx->reserved += bytes;
    }
    return b.ptr;
}

dpe_status fail(dpe_ctx *ctx, dpe_status st, const std::string &msg) {
    ctx->err = msg;
    return st;
}

// Plane-independent half of ComputeHomography (DPE.cu:459-485) and the
// camera-centre distance of DepthToWeak (DPE.cu:2644-2649), in host fp32 with
// the reference's operation order.
ViewConst view_const(const dpe_camera &ref, const dpe_camera &src) {
    ViewConst v{};
    float ref_C[3], src_C[3];
    ref_C[0] = -(ref.R[0] * ref.t[0] + ref.R[3] * ref.t[1] + ref.R[6] * ref.t[2]);
    ref_C[1] = -(ref.R[1] * ref.t[0] + ref.R[4] * ref.t[1] + ref.R[7] * ref.t[2]);
    ref_C[2] = -(ref.R[2] * ref.t[0] + ref.R[5] * ref.t[1] + ref.R[8] * ref.t[2]);
    src_C[0] = -(src.R[0] * src.t[0] + src.R[3] * src.t[1] + src.R[6] * src.t[2]);
    src_C[1] = -(src.R[1] * src.t[0] + src.R[4] * src.t[1] + src.R[7] * src.t[2]);
    src_C[2] = -(src.R[2] * src.t[0] + src.R[5] * src.t[1] + src.R[8] * src.t[2]);
    v.Rr[0] = src.R[0] * ref.R[0] + src.R[1] * ref.R[1] + src.R[2] * ref.R[2];
    v.Rr[1] = src.R[0] * ref.R[3] + src.R[1] * ref.R[4] + src.R[2] * ref.R[5];
    v.Rr[2] = src.R[0] * ref.R[6] + src.R[1] * ref.R[7] + src.R[2] * ref.R[8];
    v.Rr[3] = src.R[3] * ref.R[0] + src.R[4] * ref.R[1] + src.R[5] * ref.R[2];
    v.Rr[4] = src.R[3] * ref.R[3] + src.R[4] * ref.R[4] + src.R[5] * ref.R[5];
    v.Rr[5] = src.R[3] * ref.R[6] + src.R[4] * ref.R[7] + src.R[5] * ref.R[8];
    v.Rr[6] = src.R[6] * ref.R[0] + src.R[7] * ref.R[1] + src.R[8] * ref.R[2];
    v.Rr[7] = src.R[6] * ref.R[3] + src.R[7] * ref.R[4] + src.R[8] * ref.R[5];
    v.Rr[8] = src.R[6] * ref.R[6] + src.R[7] * ref.R[7] + src.R[8] * ref.R[8];
    float Cr[3];
    Cr[0] = ref_C[0] - src_C[0];
    Cr[1] = ref_C[1] - src_C[1];
    Cr[2] = ref_C[2] - src_C[2];
    v.tr[0] = src.R[0] * Cr[0] + src.R[1] * Cr[1] + src.R[2] * Cr[2];
    v.tr[1] = src.R[3] * Cr[0] + src.R[4] * Cr[1] + src.R[5] * Cr[2];
    v.tr[2] = src.R[6] * Cr[0] + src.R[7] * Cr[1] + src.R[8] * Cr[2];
    for (int i = 0; i < 9; ++i) v.Ks[i] = src.K[i];
    float dc[3];
    dc[0] = ref.c[0] - src.c[0];
    dc[1] = ref.c[1] - src.c[1];
    dc[2] = ref.c[2] - src.c[2];
    v.base = std::sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
    v.width = src.width;
    v.height = src.height;
    return v;
}

// The precision form's view constants in double from the fp32 cameras:
// A = Ks Rrel Kr^-1, b = Ks trel (Rrel = Rs Rr^T, trel = Rs (Cr - Cs)).
ViewConst64 view_const64(const dpe_camera &ref, const dpe_camera &src) {
    double rC[3], sC[3], Rrel[9], trel[3], KsR[9];
    for (int j = 0; j < 3; ++j) {
        rC[j] = -((double)ref.R[j] * ref.t[0] + (double)ref.R[3 + j] * ref.t[1] + (double)ref.R[6 + j] * ref.t[2]);
        sC[j] = -((double)src.R[j] * src.t[0] + (double)src.R[3 + j] * src.t[1] + (double)src.R[6 + j] * src.t[2]);
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            Rrel[3 * r + c] = (double)src.R[3 * r] * ref.R[3 * c] + (double)src.R[3 * r + 1] * ref.R[3 * c + 1] +
                              (double)src.R[3 * r + 2] * ref.R[3 * c + 2];
    for (int r = 0; r < 3; ++r)
        trel[r] = (double)src.R[3 * r] * (rC[0] - sC[0]) + (double)src.R[3 * r + 1] * (rC[1] - sC[1]) +
                  (double)src.R[3 * r + 2] * (rC[2] - sC[2]);
    ViewConst64 v{};
    for (int r = 0; r < 3; ++r) {
        v.b[r] = (double)src.K[3 * r] * trel[0] + (double)src.K[3 * r + 1] * trel[1] + (double)src.K[3 * r + 2] * trel[2];
        for (int c = 0; c < 3; ++c)
            KsR[3 * r + c] = (double)src.K[3 * r] * Rrel[c] + (double)src.K[3 * r + 1] * Rrel[3 + c] +
                             (double)src.K[3 * r + 2] * Rrel[6 + c];
    }
    // Kr^-1 = [[1/fx, 0, -cx/fx], [0, 1/fy, -cy/fy], [0, 0, 1]]
    const double fx = ref.K[0], cx = ref.K[2], fy = ref.K[4], cy = ref.K[5];
    for (int r = 0; r < 3; ++r) {
        v.A[3 * r + 0] = KsR[3 * r + 0] / fx;
        v.A[3 * r + 1] = KsR[3 * r + 1] / fy;
        v.A[3 * r + 2] = KsR[3 * r + 2] - KsR[3 * r + 0] * cx / fx - KsR[3 * r + 1] * cy / fy;
    }
    v.kinv[0] = 1.0 / fx;
    v.kinv[1] = 1.0 / fy;
    v.kinv[2] = cx / fx;
    v.kinv[3] = cy / fy;
    return v;
}

struct Upload {
    dpe_ctx *ctx;
    bool device_in;
    void to(void *dst, const void *src, size_t bytes) {
        if (bytes == 0) return;
        HIPCHK(hipMemcpyAsync(dst, src, bytes, device_in ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                              ctx->stream));
    }
};

dpe_status validate(dpe_ctx *ctx, const dpe_params *prm, const dpe_inputs *in) {
    if (!ctx) return DPE_ERR_INVALID_ARGUMENT;
    if (!prm || !in) return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "null params or inputs");
    if (in->num_images > DPE_MAX_IMAGES)
        return fail(ctx, DPE_ERR_TOO_MANY_IMAGES, "Can't process so much images: " + std::to_string(in->num_images));
    if (in->num_images < 2) return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "need a reference and at least one source");
    if (in->width <= 0 || in->height <= 0) return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "bad image size");
    // quad offsets are 32-bit byte offsets; tap coordinates stay within 2^17
    if (in->width > 65536 || in->height > 65536 || (long)(in->width + 1) * (in->height + 1) >= (1L << 28))
        return fail(ctx, DPE_ERR_UNSUPPORTED, "image larger than 65536 per side or (W+1)(H+1) >= 2^28");
    if (in->memory != DPE_MEM_HOST && in->memory != DPE_MEM_DEVICE)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "bad memory kind");
    for (int i = 0; i < in->num_images; ++i) {
        if (!in->images[i]) return fail(ctx, DPE_ERR_MISSING_INPUT, "image " + std::to_string(i) + " is null");
        if (in->cameras[i].width != in->width || in->cameras[i].height != in->height)
            return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "camera " + std::to_string(i) + " size differs from the images");
        if (prm->geom_consistency && !in->geom_depths[i])
            return fail(ctx, DPE_ERR_MISSING_INPUT, "geom_consistency needs every view's depth map");
    }
    if (prm->state < DPE_FIRST_INIT || prm->state > DPE_REFINE_ITER)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "bad run state");
    if (prm->state != DPE_FIRST_INIT && (!in->prior_planes || !in->prior_selected))
        return fail(ctx, DPE_ERR_MISSING_INPUT, "REFINE states need the prior planes and selected views");
    if (prm->use_APD && !in->weak_info)
        return fail(ctx, DPE_ERR_MISSING_INPUT, "Can't find weak info (use_APD)");
    if ((prm->use_edge || prm->use_limit) && (!in->edges || !in->edges_low || in->low_width <= 0 || in->low_height <= 0))
        return fail(ctx, DPE_ERR_MISSING_INPUT, "use_edge/use_limit need the edge maps");
    if (prm->use_label && !in->labels) return fail(ctx, DPE_ERR_MISSING_INPUT, "use_label needs the label map");
    if (prm->strong_increment <= 0 || prm->strong_radius < 0 || prm->weak_increment <= 0 || prm->max_iterations < 0)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "bad window parameters");
    if (prm->rotate_time < 1 || prm->rotate_time > 4) return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "rotate_time must be 1..4");
    const int ntap = 2 * prm->strong_radius / prm->strong_increment + 1;
    if (ntap > kMaxTaps) return fail(ctx, DPE_ERR_UNSUPPORTED, "strong window wider than 11 taps");
    return DPE_OK;
}

// Common set-up of images, cameras and constants.  Returns the Dev with the
// image/camera pointers filled.
Dev setup_views(dpe_ctx *ctx, const dpe_params *prm, const dpe_inputs *in, int precision = DPE_PRECISION_FP32) {
    Dev d{};
    d.W = in->width;
    d.H = in->height;
    d.N = in->num_images;
    d.NV = d.N - 1;
    d.P = d.W * d.H;
    d.half_rows = std::min(d.H, 2 * (((d.H / 2) + 15) / 16) * 16);
    d.Wh = (d.W + 1) / 2;
    d.Q = d.half_rows * d.Wh;
    d.prm = *prm;
    d.prm.num_images = d.N;
    d.ntap = 2 * prm->strong_radius / prm->strong_increment + 1;
    d.tap_off0 = -prm->strong_radius;
    d.tap_inc = prm->strong_increment;
    d.qs = (size_t)(d.W + 1) * (d.H + 1);
    const size_t P = d.P;
    Upload up{ctx, in->memory == DPE_MEM_DEVICE};
    // images -> contiguous fp32 (reference) and quads of every view
    float *imgs = (float *)buffer(ctx, B_IMAGES, sizeof(float) * P * d.N);
    for (int i = 0; i < d.N; ++i) up.to(imgs + P * i, in->images[i], sizeof(float) * P);
    d.ref = imgs;
    std::vector<dpe_camera> cams(in->cameras, in->cameras + d.N);
    std::vector<ViewConst> vcs(d.N);
    for (int i = 0; i < d.N; ++i) vcs[i] = view_const(cams[0], cams[i]);
    dpe_camera *dc = (dpe_camera *)buffer(ctx, B_CAMS, sizeof(dpe_camera) * d.N);
    ViewConst *dv = (ViewConst *)buffer(ctx, B_VC, sizeof(ViewConst) * d.N);
    HIPCHK(hipMemcpyAsync(dc, cams.data(), sizeof(dpe_camera) * d.N, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dv, vcs.data(), sizeof(ViewConst) * d.N, hipMemcpyHostToDevice, ctx->stream));
    d.cams = dc;
    d.vc = dv;
    // spatial exponents of the bilateral weight, -sqrt(i^2+j^2) / (2 ss^2) (DPE.cu:554-559)
    std::vector<float> sp(d.ntap * d.ntap);
    const float ss = prm->sigma_spatial;
    for (int ki = 0; ki < d.ntap; ++ki)
        for (int kj = 0; kj < d.ntap; ++kj) {
            const float fi = (float)(d.tap_off0 + ki * d.tap_inc), fj = (float)(d.tap_off0 + kj * d.tap_inc);
            const float spatial = std::sqrt(fi * fi + fj * fj);
            sp[ki * d.ntap + kj] = -spatial / (2.0f * ss * ss);
        }
    float *dsp = (float *)buffer(ctx, B_SPATIAL, sizeof(float) * sp.size());
    HIPCHK(hipMemcpyAsync(dsp, sp.data(), sizeof(float) * sp.size(), hipMemcpyHostToDevice, ctx->stream));
    d.spatial = dsp;
    // copies above are stream-ordered; the host vectors must outlive them
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // quad layout: the precision form's 16-bit format; else fp16 when every texel
    // of the call is fp16-exact (the same samples in half the bytes), else fp32
    int fmt = precision == DPE_PRECISION_FP16 ? QUAD_FP16 : precision == DPE_PRECISION_BF16 ? QUAD_BF16 : QUAD_FP32;
    if (precision == DPE_PRECISION_FP32) {
        int *flag = (int *)buffer(ctx, B_FLAG, sizeof(int));
        HIPCHK(hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
        launch_check_fp16_exact(imgs, P * d.N, flag, ctx->stream);
        int not_exact = 1;
        HIPCHK(hipMemcpyAsync(&not_exact, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        fmt = (not_exact & 1) ? QUAD_FP32 : QUAD_FP16;
        // the clamp-free gather addresses every view's quads with one 32-bit offset
        d.fast_gather = fmt == QUAD_FP16 && !(not_exact & 2) && (size_t)d.N * d.qs * 8 <= 0xFFFFFFFFull;
    }
    d.quad_fmt = fmt;
    d.prec = precision;
    if (fmt == QUAD_FP32) {
        float4 *quads = (float4 *)buffer(ctx, B_QUADS, sizeof(float4) * d.qs * d.N);
        for (int i = 0; i < d.N; ++i) launch_make_quads(d, imgs + P * i, quads + d.qs * i, ctx->stream);
        d.quads = quads;
        d.quads_h = nullptr;
    } else {
        uint2 *quads = (uint2 *)buffer(ctx, B_QUADS_H, sizeof(uint2) * d.qs * d.N);
        for (int i = 0; i < d.N; ++i) launch_make_quads_h(d, imgs + P * i, quads + d.qs * i, fmt, ctx->stream);
        d.quads = nullptr;
        d.quads_h = quads;
    }
    d.vc64 = nullptr;
    if (precision != DPE_PRECISION_FP32) {
        std::vector<ViewConst64> v64(d.N);
        for (int i = 0; i < d.N; ++i) v64[i] = view_const64(cams[0], cams[i]);
        ViewConst64 *dv64 = (ViewConst64 *)buffer(ctx, B_VC64, sizeof(ViewConst64) * d.N);
        HIPCHK(hipMemcpyAsync(dv64, v64.data(), sizeof(ViewConst64) * d.N, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        d.vc64 = dv64;
    }
    HIPCHK(hipGetLastError());
    return d;
}

void copy_out(dpe_ctx *ctx, void *dst, const void *src, size_t bytes, bool device_out) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, device_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
}

struct EventTimer {
    dpe_ctx *ctx;
    int next = 0;
    bool record = true;  // false inside a graph capture: the launches are not timed one by one
    struct Span {
        int a, b, family, tile;  // tile: index of the tile timed for the autotuner, or -1
    };
    std::vector<Span> ncc;
    hipEvent_t get() {
        if (next >= (int)ctx->events.size()) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            ctx->events.push_back(e);
        }
        return ctx->events[next++];
    }
    template <class F>
    void ncc_launch(int family, F f, int tile = -1) {
        if (!record) {
            f();
            HIPCHK(hipGetLastError());
            return;
        }
        const int a = next;
        HIPCHK(hipEventRecord(get(), ctx->stream));
        f();
        HIPCHK(hipGetLastError());
        const int b = next;
        HIPCHK(hipEventRecord(get(), ctx->stream));
        ncc.push_back({a, b, family, tile});
    }
    // launch times of the tuned tiles into the tuner record (stream synchronised)
    void tune_samples(dpe_tune &t) {
        for (const Span &sp : ncc) {
            if (sp.tile < 0) continue;
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ctx->events[sp.a], ctx->events[sp.b]));
            t.ms[sp.family == kFamFull ? 0 : 1][sp.tile].push_back(ms);
        }
    }
    // Per-family launch times and the counters of the NCC kernels into `stt`;
    // gpu_ms is left to the caller.  Needs the stream synchronised.
    // cnt: the kCounterWords striped counters as the kernels left them
    void fill(dpe_stats &stt, const unsigned long long *striped) {
        unsigned long long cnt[2 * kNumFamilies] = {};
        for (int k = 0; k < kCounterStripes; ++k)
            for (int i = 0; i < 2 * kNumFamilies; ++i) cnt[i] += striped[(size_t)k * kCounterStride + i];
        for (const Span &sp : ncc) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ctx->events[sp.a], ctx->events[sp.b]));
            stt.family_ms[sp.family] += ms;
            stt.family_launches[sp.family]++;
            if (sp.family != kFamWeak) {
                stt.ncc_ms += ms;
                stt.ncc_launches++;
            }
        }
        for (int f = 0; f < kNumFamilies; ++f) {
            stt.family_hyps[f] = (int64_t)cnt[2 * f];
            stt.family_taps[f] = (int64_t)cnt[2 * f + 1];
            stt.ncc_hyps += stt.family_hyps[f];
            stt.ncc_taps += stt.family_taps[f];
        }
    }
};

constexpr int kNumCounters = kCounterWords;

// The autotuner's tiles (index = dpe_stats.tile_full / tile_weighted):
//   0: 4 waves            1: 8 waves            2: 2 waves
//   3: 4 waves, <= 3 workgroups per CU          4: 4 waves, <= 2 workgroups per CU
// The caps come from dynamic LDS padding (the CU's LDS shared by fewer
// workgroups); a cap the device cannot express is left out.
constexpr int kTiles = 5;
constexpr int kSharedPatchBytes = (int)sizeof(SharedPatch<6>);
bool tile_config(const dpe_ctx *ctx, int i, NccTile &t) {
    static const int waves[kTiles] = {4, 8, 2, 4, 4};
    static const int cap[kTiles] = {0, 0, 0, 3, 2};
    t.waves = waves[i];
    t.lds_pad = 0;
    if (cap[i] > 0) {
        // LDS per workgroup such that exactly cap[i] workgroups fit in a CU
        const int total = std::min(ctx->lds_per_cu / cap[i], ctx->lds_per_block) / 256 * 256;
        if (total <= kSharedPatchBytes || ctx->lds_per_cu / total != cap[i]) return false;
        t.lds_pad = total - kSharedPatchBytes;
    }
    return true;
}

dpe_tune *tune_record(dpe_ctx *ctx, const Dev &d) {
    for (dpe_tune &t : ctx->tunes)
        if (t.W == d.W && t.H == d.H && t.NV == d.NV && t.use_edge == d.prm.use_edge && t.prec == d.prec) return &t;
    dpe_tune t;
    t.W = d.W;
    t.H = d.H;
    t.NV = d.NV;
    t.use_edge = d.prm.use_edge;
    t.prec = d.prec;
    t.ms[0].resize(kTiles);
    t.ms[1].resize(kTiles);
    ctx->tunes.push_back(std::move(t));
    return &ctx->tunes.back();
}

// Decide once every available tile has 2 launches of each kernel timed: the
// lowest median per kernel.
void tune_decide(const dpe_ctx *ctx, dpe_tune &t) {
    int best[2] = {0, 0};
    float best_ms[2] = {1e30f, 1e30f};
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < kTiles; ++i) {
            NccTile tc;
            if (!tile_config(ctx, i, tc)) continue;
            std::vector<float> v = t.ms[k][i];
            if (v.size() < 2) return;
            std::sort(v.begin(), v.end());
            const float med = v[v.size() / 2];
            if (med < best_ms[k]) {
                best_ms[k] = med;
                best[k] = i;
            }
        }
    t.full = best[0];
    t.wtd = best[1];
    t.decided = true;
}

// Launch a captured schedule: update the context's executable graph in place
// when the topology matches (kernel arguments and grids change), else
// instantiate it anew.
void launch_schedule_graph(dpe_ctx *ctx, hipGraph_t g) {
    bool updated = false;
    if (ctx->graph) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(ctx->graph, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            updated = true;
            ctx->graph_updates++;
        } else {
            (void)hipGetLastError();
            HIPCHK(hipGraphExecDestroy(ctx->graph));
            ctx->graph = nullptr;
        }
    }
    if (!updated) {
        HIPCHK(hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0));
        ctx->graph_rebuilds++;
    }
    HIPCHK(hipGraphLaunch(ctx->graph, ctx->stream));
}

dpe_status process_impl(dpe_ctx *ctx, const dpe_params *prm, const dpe_form *form_in, const dpe_inputs *in,
                        dpe_outputs *out) {
    dpe_status st = validate(ctx, prm, in);
    if (st != DPE_OK) return st;
    const dpe_form form = form_in ? *form_in : dpe_form{};
    if (form.gauss_newton_steps < 0 || form.gauss_newton_steps > 2)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "gauss_newton_steps must be 0, 1 or 2");
    if (form.precision < DPE_PRECISION_FP32 || form.precision > DPE_PRECISION_BF16)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "unknown precision");
    if (form.precision != DPE_PRECISION_FP32 && 2 * prm->strong_radius / prm->strong_increment + 1 != 6)
        return fail(ctx, DPE_ERR_UNSUPPORTED, "the precision form needs the 11x11 step-2 window");
    if (form.autotune < 0 || form.autotune > 1) return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "autotune must be 0 or 1");
    if (form.graph < 0 || form.graph > 1) return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "graph must be 0 or 1");
    for (int r : form.reserved)
        if (r != 0) return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "reserved form fields must be 0");
    if (form.gauss_newton_steps > 0 && 2 * prm->strong_radius / prm->strong_increment + 1 != 6)
        return fail(ctx, DPE_ERR_UNSUPPORTED, "Gauss-Newton steps need the 11x11 step-2 window");
    if (!out || !out->planes || !out->costs || !out->pixel_states || !out->selected_views)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "null output");
    HIPCHK(hipSetDevice(ctx->device));
    EventTimer tm{ctx};
    hipEvent_t ev_start = tm.get();
    Dev d = setup_views(ctx, prm, in, form.precision);
    const size_t P = d.P, Q = d.Q;
    const int N = d.N, NV = d.NV;
    Upload up{ctx, in->memory == DPE_MEM_DEVICE};
    const bool dev_in = in->memory == DPE_MEM_DEVICE;
    // geometric depths
    if (prm->geom_consistency) {
        float *g = (float *)buffer(ctx, B_GDEPTH, sizeof(float) * P * N);
        for (int i = 0; i < N; ++i) up.to(g + P * i, in->geom_depths[i], sizeof(float) * P);
        d.gdepth = g;
    }
    // pixel states and the weak index map (DPE.cpp:989-1030), and the weak
    // pixel lists: all (GenNeighbours, RANSAC) and per colour as the HALF launch
    // covers them (weak propagation, DPE.cu:1868-1902) -- built on the device;
    // only the three list lengths come back to the host (grids and sizes)
    d.weak = (uint8_t *)buffer(ctx, B_WEAK, P);
    int *nmap = (int *)buffer(ctx, B_NMAP, sizeof(int) * P);
    const size_t half_cap = (P + 1) / 2;
    int *dwl = (int *)buffer(ctx, B_WEAKLIST, sizeof(int) * (P + 2 * half_cap + 1));
    int weak_count = 0;
    int cc[2] = {0, 0};
    if (prm->use_APD) {
        up.to(d.weak, in->weak_info, P);
        int *dcnt = (int *)buffer(ctx, B_WEAKCOUNTS, sizeof(int) * 3);
        size_t temp_bytes = 0;
        HIPCHK(weak_lists_temp_bytes(d.W, d.H, &temp_bytes));
        void *temp = buffer(ctx, B_WEAKTEMP, std::max(temp_bytes, (size_t)1));
        HIPCHK(launch_weak_lists(d.weak, d.W, d.H, d.half_rows, nmap, dwl, dwl + P, dwl + P + half_cap, dcnt, temp,
                                 temp_bytes, ctx->stream));
        int cnt_h[3] = {0, 0, 0};
        HIPCHK(hipMemcpyAsync(cnt_h, dcnt, sizeof(cnt_h), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        weak_count = cnt_h[0];
        cc[0] = cnt_h[1];
        cc[1] = cnt_h[2];
    } else {
        HIPCHK(hipMemsetAsync(d.weak, DPE_STRONG, P, ctx->stream));
        HIPCHK(hipMemsetAsync(nmap, 0, sizeof(int) * P, ctx->stream));
    }
    d.weak_list = dwl;
    d.weak_count = weak_count;
    d.weak_colour_list[0] = dwl + P;
    d.weak_colour_count[0] = cc[0];
    d.weak_colour_list[1] = dwl + P + half_cap;
    d.weak_colour_count[1] = cc[1];
    {
        const size_t n = (size_t)cc[0] + cc[1] + 1;
        float *wpre = (float *)buffer(ctx, B_WPRE, sizeof(float) * kWeakPatchTaps * n);
        float *wsum = (float *)buffer(ctx, B_WSUM, sizeof(float) * 3 * kNeighbourNum * n);
        d.weak_wpre[0] = wpre;
        d.weak_wpre[1] = wpre + (size_t)kWeakPatchTaps * cc[0];
        d.weak_wsum[0] = wsum;
        d.weak_wsum[1] = wsum + (size_t)3 * kNeighbourNum * cc[0];
        const size_t nmax = (size_t)std::max(cc[0], cc[1]) + 1;
        d.weak_cost = (float *)buffer(ctx, B_WEAKCOST, sizeof(float) * 8 * (N - 1) * nmax);
        d.weak_st = (WeakState *)buffer(ctx, B_WEAKST, sizeof(WeakState) * nmax);
        d.weak_terms = (float *)buffer(ctx, B_WEAKTERMS, sizeof(float) * kWeakTermSlots * (N - 1) * nmax);
    }
    d.nmap = nmap;
    // state buffers
    d.planes = (float4 *)buffer(ctx, B_PLANES, sizeof(float4) * P);
    d.costs = (float *)buffer(ctx, B_COSTS, sizeof(float) * P);
    d.selected = (uint32_t *)buffer(ctx, B_SELECTED, sizeof(uint32_t) * P);
    d.vw = (uint8_t *)buffer(ctx, B_VW, (size_t)kMaxImages * P);
    d.fit = (float4 *)buffer(ctx, B_FIT, sizeof(float4) * P);
    d.weak_rel = (uint8_t *)buffer(ctx, B_WEAKREL, P);
    d.nearest = (short2 *)buffer(ctx, B_NEAREST, sizeof(short2) * P);
    d.neighbours = (short2 *)buffer(ctx, B_NEIGHBOURS, sizeof(short2) * ((size_t)weak_count * kNeighbourNum + 1));
    d.radius = (int *)buffer(ctx, B_RADIUS, sizeof(int) * P);
    HIPCHK(hipMemsetAsync(d.fit, 0, sizeof(float4) * P, ctx->stream));
    HIPCHK(hipMemsetAsync(d.vw, 0, (size_t)kMaxImages * P, ctx->stream));
    HIPCHK(hipMemsetAsync(d.weak_rel, 0, P, ctx->stream));
    {
        std::vector<int> rad(P, prm->strong_radius);
        HIPCHK(hipMemcpyAsync(d.radius, rad.data(), sizeof(int) * P, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    if (prm->state != DPE_FIRST_INIT) {
        up.to(d.planes, in->prior_planes, sizeof(float4) * P);
        up.to(d.selected, in->prior_selected, sizeof(uint32_t) * P);
    } else {
        HIPCHK(hipMemsetAsync(d.planes, 0, sizeof(float4) * P, ctx->stream));
        HIPCHK(hipMemsetAsync(d.selected, 0, sizeof(uint32_t) * P, ctx->stream));
    }
    if (prm->use_edge || prm->use_limit) {
        uint8_t *e = (uint8_t *)buffer(ctx, B_EDGE, P);
        uint8_t *el = (uint8_t *)buffer(ctx, B_EDGELOW, (size_t)in->low_width * in->low_height);
        up.to(e, in->edges, P);
        up.to(el, in->edges_low, (size_t)in->low_width * in->low_height);
        d.edge = e;
        d.edge_low = el;
        d.low_w = in->low_width;
        d.low_h = in->low_height;
    }
    if (prm->use_edge) {
        d.edge_neigh = (short2 *)buffer(ctx, B_EDGENEIGH, sizeof(short2) * 8 * P);
        d.complexity = (float *)buffer(ctx, B_COMPLEX, sizeof(float) * P);
        HIPCHK(hipMemsetAsync(d.complexity, 0, sizeof(float) * P, ctx->stream));
    }
    if (prm->use_label) {
        int *lab = (int *)buffer(ctx, B_LABEL, sizeof(int) * P);
        up.to(lab, in->labels, sizeof(int) * P);
        d.label = lab;
        d.label_bound = (short2 *)buffer(ctx, B_LABELBOUND, sizeof(short2) * ((size_t)weak_count * 8 + 1));
        HIPCHK(hipMemsetAsync(d.label_bound, 0xFF, sizeof(short2) * ((size_t)weak_count * 8 + 1), ctx->stream));
    }
    // half-iteration and post-pass scratch
    d.cand_pos = (int *)buffer(ctx, B_CANDPOS, sizeof(int) * kStrongSlots * Q);
    d.cand_alias = (uint8_t *)buffer(ctx, B_CANDALIAS, (size_t)kStrongSlots * Q);
    // one cost matrix per colour: a candidate repeating its slot's last plane keeps
    // the row from the colour's previous half-iteration (k_ncc_full)
    float *cost_mat =
        (float *)buffer(ctx, B_COSTMAT, sizeof(float) * std::max((size_t)2 * kStrongSlots * NV * Q, (size_t)NV * P));
    d.cost_mat = cost_mat;
    d.last_cand = (float4 *)buffer(ctx, B_LASTCAND, sizeof(float4) * 2 * kStrongSlots * Q);
    HIPCHK(hipMemsetAsync(d.last_cand, 0xFF, sizeof(float4) * 2 * kStrongSlots * Q, ctx->stream));
    d.nsel = (uint32_t *)buffer(ctx, B_NSEL, sizeof(uint32_t) * 4 * Q);
    d.rplane = (float4 *)buffer(ctx, B_RPLANE, sizeof(float4) * kRefineSlots * Q);
    d.rcost = (float *)buffer(ctx, B_RCOST, sizeof(float) * kRefineSlots * Q);
    d.now_plane = (float4 *)buffer(ctx, B_NOWPLANE, sizeof(float4) * Q);
    d.now_depth = (float *)buffer(ctx, B_NOWDEPTH, sizeof(float) * Q);
    d.now_cost = (float *)buffer(ctx, B_NOWCOST, sizeof(float) * Q);
    d.now_row = (int *)buffer(ctx, B_NOWROW, sizeof(int) * Q);
    d.cur_cost = (float *)buffer(ctx, B_CURCOST, sizeof(float) * NV * P);
    d.cur_valid = (uint32_t *)buffer(ctx, B_CURVALID, sizeof(uint32_t) * P);
    d.wnorm = (float *)buffer(ctx, B_WNORM, sizeof(float) * Q);
    d.post_op = (float4 *)buffer(ctx, B_POSTOP, sizeof(float4) * P);
    d.post_bl = (float *)buffer(ctx, B_POSTBL, sizeof(float) * P);
    d.post_disp = (float *)buffer(ctx, B_POSTDISP, sizeof(float) * P);
    d.post_wn = (float *)buffer(ctx, B_POSTWN, sizeof(float) * P);
    d.post_flags = (uint8_t *)buffer(ctx, B_POSTFLAGS, P);
    d.pcost = (float *)buffer(ctx, B_PCOST, sizeof(float) * kPostSlots * P);
    d.post_planes = (float4 *)buffer(ctx, B_POSTPLANES, sizeof(float4) * kPostSlots * P);
    d.counters = (unsigned long long *)buffer(ctx, B_COUNTERS, sizeof(unsigned long long) * kNumCounters);
    HIPCHK(hipMemsetAsync(d.counters, 0, sizeof(unsigned long long) * kNumCounters, ctx->stream));
    // GenNeighbours constants (DPE.cu:2151-2157): host double math
    {
        const float angle = 45.0f / prm->rotate_time;
        d.gn_cos = (float)std::cos(angle * M_PI / 180.f);
        d.gn_sin = (float)std::sin(angle * M_PI / 180.f);
        d.gn_thr = (float)std::cos((angle / 2.0f) * M_PI / 180.0f);
        d.gn_shift = std::max((int)(std::tan((angle / 2.0f) * M_PI / 180.0f) * 20), 1);
    }
    // tiles: the tuned ones of this problem class, or (while the tuner samples)
    // each half-iteration's candidate and refinement launches in the next tile
    dpe_tune *tune = form.autotune && d.ntap == 6 ? tune_record(ctx, d) : nullptr;
    if (tune && tune->decided) {
        tile_config(ctx, tune->full, d.full_tile);
        tile_config(ctx, tune->wtd, d.wtd_tile);
    }
    std::vector<int> tiles;
    for (int i = 0; i < kTiles; ++i) {
        NccTile t;
        if (tile_config(ctx, i, t)) tiles.push_back(i);
    }
    int half = 0;
    hipStream_t s = ctx->stream;
    HIPCHK(hipStreamSynchronize(s));
    // graph form: the schedule is captured and launched as one graph (not while
    // the tuner times single launches); its launches are not timed one by one
    const bool use_graph = form.graph && !(tune && !tune->decided);
    // ---- RunPatchMatch schedule (DPE.cu:3157-3234)
    HIPCHK(hipEventRecord(ev_start, s));
    // ends a capture that an error left open, so the stream stays usable
    struct CaptureGuard {
        hipStream_t s;
        bool active = false;
        ~CaptureGuard() {
            if (!active) return;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
        }
    } capture{s};
    if (use_graph) {
        tm.record = false;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        capture.active = true;
    }
    launch_gen_edge_inform(d, s);
    launch_find_nearest_strong(d, s);
    launch_gen_neighbours(d, s);
    launch_neighbour_update(d, s);
    launch_random_init_planes(d, s);
    tm.ncc_launch(kFamFull, [&] { launch_ncc_current(d, s); });
    launch_random_init_costs(d, s);
    const NccTile full_tile = d.full_tile, wtd_tile = d.wtd_tile;
    for (int it = 0; it < prm->max_iterations; ++it) {
        for (int c = 0; c < 2; ++c) {
            int tile = -1;
            if (tune && !tune->decided) {
                tile = tiles[half % tiles.size()];
                tile_config(ctx, tile, d.full_tile);
                d.wtd_tile = d.full_tile;
            }
            half++;
            d.cost_mat = cost_mat + (size_t)c * kStrongSlots * NV * Q;
            launch_strong_candidates(d, c, it, s);
            tm.ncc_launch(kFamFull, [&] { launch_ncc_strong(d, c, s); }, tile);
            launch_strong_select(d, c, it, s);
            // Gauss-Newton form: step 1 (from the current plane) replaces the perturbed-normal
            // hypothesis (slot 3), step 2 (from step 1) the random depth + normal one (slot 2)
            if (form.gauss_newton_steps >= 1)
                tm.ncc_launch(kFamRefine, [&] { launch_gn_step(d, c, d.now_plane, d.rplane + 3 * Q, s); });
            if (form.gauss_newton_steps >= 2)
                tm.ncc_launch(kFamRefine, [&] { launch_gn_step(d, c, d.rplane + 3 * Q, d.rplane + 2 * Q, s); });
            tm.ncc_launch(kFamRefine, [&] { launch_ncc_refine(d, c, s); }, tile);
            launch_strong_accept(d, c, s);
        }
        launch_ransac(d, it, s);
        if (weak_count > 0) {
            for (int c = 0; c < 2; ++c)
                if (d.weak_colour_count[c] > 0) tm.ncc_launch(kFamWeak, [&] { launch_weak_half(d, c, it, s); });
        }
    }
    d.full_tile = full_tile;
    d.wtd_tile = wtd_tile;
    launch_depth_and_normal(d, s);
    launch_filter(d, 0, s);
    launch_filter(d, 1, s);
    launch_post_prep(d, s);
    launch_post_planes(d, s);
    tm.ncc_launch(kFamPost, [&] { launch_ncc_post(d, s); });
    launch_post_final(d, s);
    HIPCHK(hipGetLastError());
    if (use_graph) {
        hipGraph_t g = nullptr;
        capture.active = false;
        HIPCHK(hipStreamEndCapture(s, &g));
        struct GraphGuard {
            hipGraph_t g;
            ~GraphGuard() { (void)hipGraphDestroy(g); }
        } graph_guard{g};
        launch_schedule_graph(ctx, g);
    }
    hipEvent_t ev_stop = tm.get();
    HIPCHK(hipEventRecord(ev_stop, s));
    // ---- downloads (DPE.cu:3258-3260, plus the cost map)
    const bool dev_out = dev_in;
    copy_out(ctx, out->planes, d.planes, sizeof(float4) * P, dev_out);
    copy_out(ctx, out->costs, d.costs, sizeof(float) * P, dev_out);
    copy_out(ctx, out->pixel_states, d.weak, P, dev_out);
    copy_out(ctx, out->selected_views, d.selected, sizeof(uint32_t) * P, dev_out);
    std::vector<unsigned long long> cnt(kNumCounters);
    HIPCHK(hipMemcpyAsync(cnt.data(), d.counters, sizeof(unsigned long long) * kNumCounters, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    dpe_stats stt{};
    HIPCHK(hipEventElapsedTime(&stt.gpu_ms, ev_start, ev_stop));
    tm.fill(stt, cnt.data());
    stt.quad_bytes = d.quad_fmt == QUAD_FP32 ? 16 : 8;
    stt.tile_full = stt.tile_weighted = 0;
    stt.graph = use_graph ? 1 : 0;
    if (tune) {
        if (tune->decided) {
            stt.tile_full = tune->full;
            stt.tile_weighted = tune->wtd;
        } else {
            tm.tune_samples(*tune);
            tune_decide(ctx, *tune);
            stt.tile_full = stt.tile_weighted = -1;
        }
    }
    ctx->stats = stt;
    return DPE_OK;
}

dpe_status ncc_adjoint_impl(dpe_ctx *ctx, const dpe_params *prm, const dpe_inputs *in, const float *plane_params,
                            const float *view_weights, float *costs, float *grads) {
    dpe_params p = *prm;
    p.use_APD = p.use_edge = p.use_limit = p.use_label = p.geom_consistency = 0;
    p.state = DPE_FIRST_INIT;
    dpe_status st = validate(ctx, &p, in);
    if (st != DPE_OK) return st;
    if (!plane_params || !costs || !grads) return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "null plane parameters or outputs");
    if (2 * p.strong_radius / p.strong_increment + 1 != 6)
        return fail(ctx, DPE_ERR_UNSUPPORTED, "the adjoint needs the 11x11 step-2 window");
    HIPCHK(hipSetDevice(ctx->device));
    EventTimer tm{ctx};
    Dev d = setup_views(ctx, &p, in);
    const size_t P = d.P;
    const bool dev = in->memory == DPE_MEM_DEVICE;
    const float4 *tpd = (const float4 *)plane_params;
    const float *vw = view_weights;
    float *co = costs;
    float4 *gr = (float4 *)grads;
    if (!dev) {
        float4 *b_tpd = (float4 *)buffer(ctx, B_GIVEN_PLANES, sizeof(float4) * P);
        HIPCHK(hipMemcpyAsync(b_tpd, plane_params, sizeof(float4) * P, hipMemcpyHostToDevice, ctx->stream));
        tpd = b_tpd;
        if (view_weights) {
            float *b_vw = (float *)buffer(ctx, B_ADJ_WEIGHTS, sizeof(float) * P * d.NV);
            HIPCHK(hipMemcpyAsync(b_vw, view_weights, sizeof(float) * P * d.NV, hipMemcpyHostToDevice, ctx->stream));
            vw = b_vw;
        }
        co = (float *)buffer(ctx, B_GIVEN_COSTS, sizeof(float) * P);
        gr = (float4 *)buffer(ctx, B_ADJ_GRADS, sizeof(float4) * P);
    }
    d.counters = (unsigned long long *)buffer(ctx, B_COUNTERS, sizeof(unsigned long long) * kNumCounters);
    HIPCHK(hipMemsetAsync(d.counters, 0, sizeof(unsigned long long) * kNumCounters, ctx->stream));
    tm.ncc_launch(kFamFull, [&] { launch_ncc_adjoint(d, tpd, vw, co, gr, ctx->stream); });
    if (!dev) {
        HIPCHK(hipMemcpyAsync(costs, co, sizeof(float) * P, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(grads, gr, sizeof(float4) * P, hipMemcpyDeviceToHost, ctx->stream));
    }
    std::vector<unsigned long long> cnt(kNumCounters);
    HIPCHK(hipMemcpyAsync(cnt.data(), d.counters, sizeof(unsigned long long) * kNumCounters, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    dpe_stats stt{};
    tm.fill(stt, cnt.data());
    stt.quad_bytes = d.quad_fmt == QUAD_FP32 ? 16 : 8;
    stt.gpu_ms = stt.ncc_ms;
    ctx->stats = stt;
    return DPE_OK;
}

// Upper bound of the scratch bytes one dpe_ncc_cost call with device-resident
// inputs carves (every buffer rounded up to 256 B, quads counted at fp32 size):
// images, quads, cameras, view constants, spatial weights, flag, counters.
size_t ncc_cost_scratch_bytes(const dpe_params *p, const dpe_inputs *in) {
    auto up = [](size_t b) { return (std::max(b, (size_t)16) + 255) & ~(size_t)255; };
    const size_t P = (size_t)in->width * in->height, N = (size_t)in->num_images;
    const size_t qs = (size_t)(in->width + 1) * (in->height + 1);
    const size_t ntap = (size_t)(2 * p->strong_radius / p->strong_increment + 1);
    return up(sizeof(float) * P * N) + up(sizeof(float4) * qs * N) + up(sizeof(dpe_camera) * N) +
           up(sizeof(ViewConst) * N) + up(sizeof(float) * ntap * ntap) + up(sizeof(int)) +
           up(sizeof(unsigned long long) * kNumCounters);
}

dpe_status ncc_cost_impl(dpe_ctx *ctx, const dpe_params *prm, const dpe_form *form, const dpe_inputs *in, int num_hyps,
                         const float *planes, float *costs, bool no_alloc = false) {
    dpe_params p = *prm;
    p.use_APD = p.use_edge = p.use_limit = p.use_label = p.geom_consistency = 0;
    p.state = DPE_FIRST_INIT;
    dpe_status st = validate(ctx, &p, in);
    if (st != DPE_OK) return st;
    if (num_hyps < 1 || !planes || !costs) return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "bad hypotheses");
    const int precision = form ? form->precision : DPE_PRECISION_FP32;
    if (precision < DPE_PRECISION_FP32 || precision > DPE_PRECISION_BF16)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "unknown precision");
    if (precision != DPE_PRECISION_FP32 && 2 * p.strong_radius / p.strong_increment + 1 != 6)
        return fail(ctx, DPE_ERR_UNSUPPORTED, "the precision form needs the 11x11 step-2 window");
    // allocation-free form: device-resident inputs and outputs only, every
    // buffer of the call carved from this device's scratch array, no timing
    // events (their pool could grow)
    struct ArenaGuard {
        dpe_ctx *c;
        ~ArenaGuard() {
            c->arena = nullptr;
            c->arena_bytes = c->arena_used = 0;
        }
    };
    std::unique_ptr<ArenaGuard> arena_guard;
    if (no_alloc) {
        if (in->memory != DPE_MEM_DEVICE)
            return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "the allocation-free call needs device-resident inputs (DPE_MEM_DEVICE)");
        char *base = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_scratch_mu);
            base = ctx->device < kMaxScratchDevices ? g_scratch_ptr[ctx->device] : nullptr;
        }
        if (!base) return fail(ctx, DPE_ERR_HIP, "the device's scratch address was not resolved at dpe_create");
        const size_t need = ncc_cost_scratch_bytes(&p, in);
        if (need > kNoAllocScratchBytes)
            return fail(ctx, DPE_ERR_UNSUPPORTED,
                        "the allocation-free call holds " + std::to_string(kNoAllocScratchBytes) +
                            " bytes of scratch; this problem needs " + std::to_string(need) +
                            " (4 B x pixels x images + 16 B x (W+1)(H+1) x images + 64 KiB)");
        ctx->arena = base;
        ctx->arena_bytes = kNoAllocScratchBytes;
        ctx->arena_used = 0;
        arena_guard.reset(new ArenaGuard{ctx});
    }
    HIPCHK(hipSetDevice(ctx->device));
    EventTimer tm{ctx};
    tm.record = !no_alloc;
    Dev d = setup_views(ctx, &p, in, precision);
    const size_t P = d.P;
    const bool dev = in->memory == DPE_MEM_DEVICE;
    const float4 *pl = (const float4 *)planes;
    float *co = costs;
    if (!dev) {
        float4 *gp = (float4 *)buffer(ctx, B_GIVEN_PLANES, sizeof(float4) * P * num_hyps);
        float *gc = (float *)buffer(ctx, B_GIVEN_COSTS, sizeof(float) * P * num_hyps * d.NV);
        HIPCHK(hipMemcpyAsync(gp, planes, sizeof(float4) * P * num_hyps, hipMemcpyHostToDevice, ctx->stream));
        pl = gp;
        co = gc;
    }
    d.counters = (unsigned long long *)buffer(ctx, B_COUNTERS, sizeof(unsigned long long) * kNumCounters);
    HIPCHK(hipMemsetAsync(d.counters, 0, sizeof(unsigned long long) * kNumCounters, ctx->stream));
    tm.ncc_launch(kFamFull, [&] { launch_ncc_given(d, num_hyps, pl, co, ctx->stream); });
    if (!dev) HIPCHK(hipMemcpyAsync(costs, co, sizeof(float) * P * num_hyps * d.NV, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<unsigned long long> cnt(kNumCounters);
    HIPCHK(hipMemcpyAsync(cnt.data(), d.counters, sizeof(unsigned long long) * kNumCounters, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    dpe_stats stt{};
    tm.fill(stt, cnt.data());
    stt.quad_bytes = d.quad_fmt == QUAD_FP32 ? 16 : 8;
    stt.gpu_ms = stt.ncc_ms;
    ctx->stats = stt;
    return DPE_OK;
}

// Validated views of a fusion call, uploaded back to back (FuseViewDev.offset).
struct FuseUpload {
    std::vector<FuseViewDev> fv;
    long total = 0;
    int maxP = 0;
    FuseDev f{};
};

FuseUpload fuse_upload(dpe_ctx *ctx, const dpe_fuse_view *views, int nv, bool need_states, bool dev) {
    hipStream_t s = ctx->stream;
    FuseUpload u;
    u.fv.resize(nv);
    for (int i = 0; i < nv; ++i) {
        const dpe_fuse_view &v = views[i];
        if (v.width <= 0 || v.height <= 0 || !v.depth || !v.normal || (need_states && !v.pixel_states) || !v.color)
            throw Fail{fail(ctx, DPE_ERR_INVALID_ARGUMENT, "fusion view " + std::to_string(i) + " is incomplete")};
        if (v.num_src < 0 || v.num_src > DPE_MAX_IMAGES)
            throw Fail{fail(ctx, DPE_ERR_INVALID_ARGUMENT, "fusion view " + std::to_string(i) + ": bad source count")};
        FuseViewDev &d = u.fv[i];
        d.width = v.width;
        d.height = v.height;
        d.offset = u.total;
        d.cam = v.camera;
        d.num_src = v.num_src;
        for (int j = 0; j < v.num_src; ++j) {
            if (v.src_index[j] < 0 || v.src_index[j] >= nv || v.src_index[j] == i)
                throw Fail{fail(ctx, DPE_ERR_INVALID_ARGUMENT, "fusion view " + std::to_string(i) + ": bad source index")};
            d.src[j] = v.src_index[j];
        }
        u.total += (long)v.width * v.height;
        u.maxP = std::max(u.maxP, v.width * v.height);
    }
    FuseViewDev *dv = (FuseViewDev *)buffer(ctx, B_F_VIEWS, sizeof(FuseViewDev) * nv);
    float *depth = (float *)buffer(ctx, B_F_DEPTH, sizeof(float) * u.total);
    float *normal = (float *)buffer(ctx, B_F_NORMAL, sizeof(float) * 3 * u.total);
    uint8_t *states = need_states ? (uint8_t *)buffer(ctx, B_F_STATES, u.total) : nullptr;
    uint8_t *color = (uint8_t *)buffer(ctx, B_F_COLOR, 3 * u.total);
    uint8_t *mask = (uint8_t *)buffer(ctx, B_F_MASK, u.total);
    HIPCHK(hipMemcpyAsync(dv, u.fv.data(), sizeof(FuseViewDev) * nv, hipMemcpyHostToDevice, s));
    const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    for (int i = 0; i < nv; ++i) {
        const size_t P = (size_t)u.fv[i].width * u.fv[i].height;
        HIPCHK(hipMemcpyAsync(depth + u.fv[i].offset, views[i].depth, sizeof(float) * P, kind, s));
        HIPCHK(hipMemcpyAsync(normal + 3 * u.fv[i].offset, views[i].normal, sizeof(float) * 3 * P, kind, s));
        if (need_states) HIPCHK(hipMemcpyAsync(states + u.fv[i].offset, views[i].pixel_states, P, kind, s));
        HIPCHK(hipMemcpyAsync(color + 3 * u.fv[i].offset, views[i].color, 3 * P, kind, s));
    }
    HIPCHK(hipMemsetAsync(mask, 0, u.total, s));
    u.f.views = dv;
    u.f.depth = depth;
    u.f.normal = normal;
    u.f.states = states;
    u.f.color = color;
    u.f.mask = mask;
    return u;
}

// Copies n points of one view from the staging buffer to the caller's array.
void fuse_emit(dpe_ctx *ctx, const dpe_point *stage, int64_t n, dpe_point *points, int64_t capacity, int64_t count,
               bool dev) {
    if (points && count < capacity && n > 0) {
        const int64_t m = std::min<int64_t>(n, capacity - count);
        HIPCHK(hipMemcpyAsync(points + count, stage, sizeof(dpe_point) * m,
                              dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
}

// Number of points of a compaction: exclusive sum at P-1 plus the last flag.
int64_t fuse_count(dpe_ctx *ctx, const int *index, const int *flags, int P) {
    int last_index = 0, last_flag = 0;
    HIPCHK(hipMemcpyAsync(&last_index, index + P - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(&last_flag, flags + P - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return (int64_t)last_index + last_flag;
}

dpe_status fuse_impl(dpe_ctx *ctx, const dpe_fuse_view *views, int nv, const dpe_fuse_params *fp, dpe_point *points,
                     int64_t capacity, int64_t *num_points) {
    if (!views || nv <= 0 || !fp || !num_points) return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "bad fusion arguments");
    if (fp->memory != DPE_MEM_HOST && fp->memory != DPE_MEM_DEVICE)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "bad memory kind");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const bool dev = fp->memory == DPE_MEM_DEVICE;
    FuseUpload u = fuse_upload(ctx, views, nv, true, dev);
    const std::vector<FuseViewDev> &fv = u.fv;
    const int maxP = u.maxP;
    FuseDev &f = u.f;
    int *claim = (int *)buffer(ctx, B_F_CLAIM, sizeof(int) * u.total);
    int *used[2] = {(int *)buffer(ctx, B_F_USED0, sizeof(int) * (size_t)maxP * DPE_MAX_IMAGES),
                    (int *)buffer(ctx, B_F_USED1, sizeof(int) * (size_t)maxP * DPE_MAX_IMAGES)};
    uint8_t *acc[2] = {(uint8_t *)buffer(ctx, B_F_ACC0, maxP), (uint8_t *)buffer(ctx, B_F_ACC1, maxP)};
    int *flags = (int *)buffer(ctx, B_F_FLAGS, sizeof(int) * maxP);
    int *index = (int *)buffer(ctx, B_F_INDEX, sizeof(int) * maxP);
    size_t temp_bytes = 0;
    HIPCHK(fuse_scan_temp_bytes(maxP, &temp_bytes));
    void *temp = buffer(ctx, B_F_TEMP, temp_bytes);
    int *changed = (int *)buffer(ctx, B_F_CHANGED, sizeof(int));
    f.claim = claim;
    f.max_reproj = fp->max_reproj_px;
    f.max_rel = fp->max_rel_depth;
    f.max_angle = fp->max_angle_rad;
    f.strong_factor = fp->strong_factor;
    f.weak_factor = fp->weak_factor;
    f.min_consistent = fp->min_consistent;
    if (fp->fp64_reprojection != 0 && fp->fp64_reprojection != 1)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "fp64_reprojection must be 0 or 1");
    f.fp64 = fp->fp64_reprojection;
    // points of one view go to a staging buffer, then to the caller
    dpe_point *stage = (dpe_point *)buffer(ctx, B_F_POINTS, sizeof(dpe_point) * (size_t)maxP);
    int64_t count = 0;
    for (int i = 0; i < nv; ++i) {
        const int P = fv[i].width * fv[i].height;
        const int ns = fv[i].num_src;
        HIPCHK(hipMemsetAsync(acc[0], 0, P, s));
        HIPCHK(hipMemsetAsync(used[0], 0xFF, sizeof(int) * (size_t)P * std::max(ns, 1), s));
        launch_fuse_fill_claims(f, i, ns, s);
        // A pixel reads only claims of smaller raster index, so after round k the first k pixels hold their
        // sequential result and round P + 1 at the latest finds nothing to change; more rounds mean a bug.
        int cur = 0;
        for (int round = 0;; ++round) {
            if (round > P) return fail(ctx, DPE_ERR_INTERNAL, "fusion fixpoint exceeded its P + 1 round bound");
            HIPCHK(hipMemsetAsync(changed, 0, sizeof(int), s));
            launch_fuse_eval(f, i, P, used[cur], acc[cur], used[1 - cur], acc[1 - cur], changed, s);
            cur = 1 - cur;
            launch_fuse_fill_claims(f, i, ns, s);
            launch_fuse_claim(f, i, P, used[cur], acc[cur], s);
            HIPCHK(hipGetLastError());
            int ch = 0;
            HIPCHK(hipMemcpyAsync(&ch, changed, sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (!ch) break;
        }
        HIPCHK(launch_fuse_compact(f, i, P, used[cur], acc[cur], flags, index, temp, temp_bytes, stage, 0, s));
        HIPCHK(hipGetLastError());
        const int64_t n = fuse_count(ctx, index, flags, P);
        fuse_emit(ctx, stage, n, points, capacity, count, dev);
        count += n;
    }
    *num_points = count;
    return DPE_OK;
}

dpe_status fuse_tat_impl(dpe_ctx *ctx, const dpe_fuse_view *views, int nv, const uint8_t *const *blocks,
                         const dpe_fuse_tat_params *tp, dpe_point *points, int64_t capacity, int64_t *num_points) {
    if (!views || nv <= 0 || !tp || !num_points)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "bad fusion arguments");
    if (tp->memory != DPE_MEM_HOST && tp->memory != DPE_MEM_DEVICE)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "bad memory kind");
    if (tp->variant != DPE_TAT_INTERMEDIATE && tp->variant != DPE_TAT_ADVANCED)
        return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "bad TAT fusion variant");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const bool dev = tp->memory == DPE_MEM_DEVICE;
    FuseUpload u = fuse_upload(ctx, views, nv, false, dev);
    const std::vector<FuseViewDev> &fv = u.fv;
    const int maxP = u.maxP;
    FuseDev &f = u.f;
    int maxN = 0;
    for (int i = 0; i < nv; ++i) {
        if ((long)fv[i].num_src * fv[i].width * fv[i].height > 0x7FFFFFFFL)
            return fail(ctx, DPE_ERR_INVALID_ARGUMENT, "TAT fusion view " + std::to_string(i) + " is too large");
        maxN = std::max(maxN, fv[i].num_src * fv[i].width * fv[i].height);
    }
    if (blocks) {
        // block maps packed like the depth maps; a view without one lets every pixel through
        uint8_t *block = (uint8_t *)buffer(ctx, B_F_STATES, u.total);
        const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        for (int i = 0; i < nv; ++i) {
            const size_t P = (size_t)fv[i].width * fv[i].height;
            if (blocks[i]) HIPCHK(hipMemcpyAsync(block + fv[i].offset, blocks[i], P, kind, s));
            else HIPCHK(hipMemsetAsync(block + fv[i].offset, 255, P, s));
        }
        f.block = block;
    }
    f.dist_base = tp->dist_base;
    f.depth_base = tp->depth_base;
    f.angle_base = tp->angle_base;
    f.angle_grad = tp->angle_grad;
    f.tat_intermediate = tp->variant == DPE_TAT_INTERMEDIATE;
    int *keys = (int *)buffer(ctx, B_F_USED0, sizeof(int) * (size_t)std::max(maxN, 1));
    int *last = (int *)buffer(ctx, B_F_USED1, sizeof(int) * (size_t)std::max(maxN, 1));
    uint8_t *acc = (uint8_t *)buffer(ctx, B_F_ACC0, maxP);
    uint32_t *colour = (uint32_t *)buffer(ctx, B_F_CLAIM, sizeof(uint32_t) * maxP);
    int *flags = (int *)buffer(ctx, B_F_FLAGS, sizeof(int) * maxP);
    int *index = (int *)buffer(ctx, B_F_INDEX, sizeof(int) * maxP);
    dpe_point *stage = (dpe_point *)buffer(ctx, B_F_POINTS, sizeof(dpe_point) * (size_t)maxP);
    int64_t count = 0;
    for (int i = 0; i < nv; ++i) {
        const int P = fv[i].width * fv[i].height;
        const int ns = fv[i].num_src;
        if (ns < 2) continue;  // k runs over 2..num_src (DPE.cpp:1684)
        // scan workspace sized for this view's two scans
        size_t scan_bytes = 0, last_bytes = 0;
        HIPCHK(fuse_scan_temp_bytes(P, &scan_bytes));
        HIPCHK(tat_scan_temp_bytes(ns * P, &last_bytes));
        const size_t temp_bytes = std::max(scan_bytes, last_bytes);
        void *temp = buffer(ctx, B_F_TEMP, temp_bytes);
        launch_tat_keys(f, i, P, keys, s);
        HIPCHK(hipGetLastError());
        HIPCHK(launch_tat_last(keys, last, ns * P, temp, temp_bytes, s));
        launch_tat_decide(f, i, P, last, acc, colour, s);
        HIPCHK(hipGetLastError());
        HIPCHK(launch_tat_compact(f, i, P, acc, colour, flags, index, temp, temp_bytes, stage, s));
        const int64_t n = fuse_count(ctx, index, flags, P);
        fuse_emit(ctx, stage, n, points, capacity, count, dev);
        count += n;
    }
    *num_points = count;
    return DPE_OK;
}

}  // namespace

extern "C" {

int dpe_abi_version(void) { return DPE_ABI_VERSION; }

dpe_status dpe_create(int device, dpe_ctx **out) {
    if (!out) {
        g_create_error = "null output pointer";
        return DPE_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        g_create_error = "no HIP device";
        return DPE_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
        g_create_error = "device index out of range";
        return DPE_ERR_NO_DEVICE;
    }
    dpe_ctx *ctx = new dpe_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "cannot create a HIP stream";
        delete ctx;
        return DPE_ERR_HIP;
    }
    if (hipDeviceGetAttribute(&ctx->lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) !=
            hipSuccess ||
        hipDeviceGetAttribute(&ctx->lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) {
        g_create_error = "cannot query the LDS size";
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return DPE_ERR_HIP;
    }
    // load every code object of the library now rather than inside the first call
    preload_kernels(ctx->stream);
    preload_adjoint(ctx->stream);
    preload_fuse(ctx->stream);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
        g_create_error = "cannot load the library's kernels";
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return DPE_ERR_HIP;
    }
    // this device's address of the allocation-free scratch array, once per device
    // (with the device current: each GPU holds its own copy at its own address)
    if (device < kMaxScratchDevices) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[device]) {
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                g_create_error = std::string("cannot resolve the scratch array: ") + hipGetErrorString(e);
                (void)hipStreamDestroy(ctx->stream);
                delete ctx;
                return DPE_ERR_HIP;
            }
            g_scratch_ptr[device] = (char *)p;
        }
    }
    *out = ctx;
    return DPE_OK;
}

void dpe_destroy(dpe_ctx *ctx) {
    if (!ctx) return;
    // teardown: errors cannot be reported any more, so they are ignored
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    for (auto &b : ctx->bufs)
        if (b.ptr) (void)hipFree(b.ptr);
    for (auto e : ctx->events) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *dpe_last_error(const dpe_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

dpe_status dpe_process(dpe_ctx *ctx, const dpe_params *params, const dpe_inputs *in, dpe_outputs *out) {
    return dpe_process_ex(ctx, params, nullptr, in, out);
}

dpe_status dpe_process_ex(dpe_ctx *ctx, const dpe_params *params, const dpe_form *form, const dpe_inputs *in,
                          dpe_outputs *out) {
    if (!ctx) return DPE_ERR_INVALID_ARGUMENT;
    try {
        return process_impl(ctx, params, form, in, out);
    } catch (const Fail &f) {
        return f.st;
    }
}

// ---- batched form: independent problems over concurrent lanes
struct dpe_batch {
    int device = 0;
    std::vector<dpe_ctx *> lanes;
    std::string err;
};

dpe_status dpe_batch_create(int device, int lanes, dpe_batch **out) {
    if (!out) return DPE_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (lanes < 1 || lanes > DPE_MAX_LANES) {
        g_create_error = "lanes must be in [1, DPE_MAX_LANES]";
        return DPE_ERR_INVALID_ARGUMENT;
    }
    dpe_batch *b = new dpe_batch();
    b->device = device;
    for (int i = 0; i < lanes; ++i) {
        dpe_ctx *c = nullptr;
        const dpe_status st = dpe_create(device, &c);
        if (st != DPE_OK) {
            dpe_batch_destroy(b);
            return st;
        }
        b->lanes.push_back(c);
    }
    *out = b;
    return DPE_OK;
}

void dpe_batch_destroy(dpe_batch *b) {
    if (!b) return;
    for (dpe_ctx *c : b->lanes) dpe_destroy(c);
    delete b;
}

const char *dpe_batch_last_error(const dpe_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

dpe_status dpe_batch_process(dpe_batch *b, int n, const dpe_params *params, const dpe_form *form,
                             const dpe_inputs *inputs, dpe_outputs *outputs, dpe_stats *stats, float *wall_ms) {
    if (!b) return DPE_ERR_INVALID_ARGUMENT;
    b->err.clear();
    if (n < 0 || (n > 0 && (!params || !inputs || !outputs))) {
        b->err = "bad batch arrays";
        return DPE_ERR_INVALID_ARGUMENT;
    }
    // longest-processing-time-first assignment on the estimated NCC work
    // (pixels x source views x (iterations + 1)) to the least loaded lane
    std::vector<int> order(n);
    std::vector<double> cost(n);
    for (int i = 0; i < n; ++i) {
        order[i] = i;
        cost[i] = (double)inputs[i].width * inputs[i].height * std::max(inputs[i].num_images - 1, 1) *
                  (std::max(params[i].max_iterations, 0) + 1);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
    const int L = (int)b->lanes.size();
    std::vector<std::vector<int>> work(L);
    std::vector<double> load(L, 0.0);
    for (int i : order) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        work[l].push_back(i);
        load[l] += cost[i];
    }
    // the lanes share what the autotuner has decided
    for (dpe_ctx *c : b->lanes)
        for (const dpe_tune &t : b->lanes[0]->tunes)
            if (t.decided) {
                bool have = false;
                for (dpe_tune &u : c->tunes)
                    if (u.W == t.W && u.H == t.H && u.NV == t.NV && u.use_edge == t.use_edge && u.prec == t.prec) {
                        if (!u.decided) u = t;
                        have = true;
                    }
                if (!have) c->tunes.push_back(t);
            }
    std::vector<dpe_status> status(n, DPE_OK);
    std::vector<std::string> msg(n);
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> threads;
    for (int l = 0; l < L; ++l) {
        if (work[l].empty()) continue;
        threads.emplace_back([&, l] {
            dpe_ctx *c = b->lanes[l];
            for (int i : work[l]) {
                status[i] = dpe_process_ex(c, &params[i], form, &inputs[i], &outputs[i]);
                if (status[i] != DPE_OK) {
                    msg[i] = c->err;
                    break;
                }
                if (stats) stats[i] = c->stats;
            }
        });
    }
    for (std::thread &t : threads) t.join();
    const auto t1 = std::chrono::steady_clock::now();
    if (wall_ms) *wall_ms = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();
    // decided tuner records flow back to lane 0 for the next batch
    for (int l = 1; l < L; ++l)
        for (const dpe_tune &t : b->lanes[l]->tunes)
            if (t.decided) {
                bool have = false;
                for (dpe_tune &u : b->lanes[0]->tunes)
                    if (u.W == t.W && u.H == t.H && u.NV == t.NV && u.use_edge == t.use_edge && u.prec == t.prec) {
                        if (!u.decided) u = t;
                        have = true;
                    }
                if (!have) b->lanes[0]->tunes.push_back(t);
            }
    for (int i = 0; i < n; ++i)
        if (status[i] != DPE_OK) {
            b->err = "problem " + std::to_string(i) + ": " + msg[i];
            return status[i];
        }
    return DPE_OK;
}

dpe_status dpe_last_stats(const dpe_ctx *ctx, dpe_stats *stats) {
    if (!ctx || !stats) return DPE_ERR_INVALID_ARGUMENT;
    *stats = ctx->stats;
    return DPE_OK;
}

dpe_status dpe_ncc_cost(dpe_ctx *ctx, const dpe_params *params, const dpe_inputs *in, int num_hyps,
                        const float *planes, float *costs) {
    return dpe_ncc_cost_ex(ctx, params, nullptr, in, num_hyps, planes, costs);
}

dpe_status dpe_ncc_cost_ex(dpe_ctx *ctx, const dpe_params *params, const dpe_form *form, const dpe_inputs *in,
                           int num_hyps, const float *planes, float *costs) {
    if (!ctx) return DPE_ERR_INVALID_ARGUMENT;
    try {
        return ncc_cost_impl(ctx, params, form, in, num_hyps, planes, costs);
    } catch (const Fail &f) {
        return f.st;
    }
}

dpe_status dpe_ncc_cost_noalloc(dpe_ctx *ctx, const dpe_params *params, const dpe_inputs *in, int num_hyps,
                                const float *planes, float *costs) {
    if (!ctx) return DPE_ERR_INVALID_ARGUMENT;
    try {
        return ncc_cost_impl(ctx, params, nullptr, in, num_hyps, planes, costs, true);
    } catch (const Fail &f) {
        return f.st;
    }
}

dpe_status dpe_ncc_cost_adjoint(dpe_ctx *ctx, const dpe_params *params, const dpe_inputs *in,
                                const float *plane_params, const float *view_weights, float *costs, float *grads) {
    if (!ctx) return DPE_ERR_INVALID_ARGUMENT;
    try {
        return ncc_adjoint_impl(ctx, params, in, plane_params, view_weights, costs, grads);
    } catch (const Fail &f) {
        return f.st;
    }
}

dpe_status dpe_fuse(dpe_ctx *ctx, const dpe_fuse_view *views, int num_views, const dpe_fuse_params *params,
                    dpe_point *points, int64_t capacity, int64_t *num_points) {
    if (!ctx) return DPE_ERR_INVALID_ARGUMENT;
    try {
        return fuse_impl(ctx, views, num_views, params, points, capacity, num_points);
    } catch (const Fail &f) {
        return f.st;
    }
}

dpe_status dpe_fuse_tat(dpe_ctx *ctx, const dpe_fuse_view *views, int num_views, const uint8_t *const *blocks,
                        const dpe_fuse_tat_params *params, dpe_point *points, int64_t capacity,
                        int64_t *num_points) {
    if (!ctx) return DPE_ERR_INVALID_ARGUMENT;
    try {
        return fuse_tat_impl(ctx, views, num_views, blocks, params, points, capacity, num_points);
    } catch (const Fail &f) {
        return f.st;
    }
}

}  // extern "C"
