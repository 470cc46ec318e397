/*
 * dpe_hip.h -- C-ABI of libdpe_hip.so, the MI355X (gfx950) implementation of
 * DPE-MVS's per-reference-image PatchMatch depth/normal estimation and
 * depth-map fusion.
 *
 * The reference has no FFI: its boundary is the C++ class `DPE` (DPE.h:140-243)
 * driven by `ProcessProblem` (main.cpp:241-319) and the free function
 * `RunFusion` (DPE.h:95-96, DPE.cpp:1387-1548).  Every entry point below names
 * the reference interface it replaces.  All structs are plain C; every pointer
 * is host memory unless `memory == DPE_MEM_DEVICE`, in which case it is a
 * device pointer on the context's GPU.  No call exits the process: every
 * failure is a status code plus a message from dpe_last_error().
 */
#ifndef DPE_HIP_H_
#define DPE_HIP_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPE_ABI_VERSION 11
/* main.h:41-44 */
#define DPE_MAX_IMAGES 32
#define DPE_NEIGHBOUR_NUM 9
#define DPE_MAX_SEARCH_RADIUS 4096

typedef enum dpe_status {
    DPE_OK = 0,
    DPE_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size, bad enum */
    DPE_ERR_TOO_MANY_IMAGES = 2,  /* > 32 images per problem: DPE.cpp:944-947 */
    DPE_ERR_MISSING_INPUT = 3,    /* an input the params require is null: DPE.cpp:995-998 */
    DPE_ERR_HIP = 4,              /* HIP runtime error: replaces CudaSafeCall's exit, DPE.cpp:646-679 */
    DPE_ERR_NO_DEVICE = 5,
    DPE_ERR_UNSUPPORTED = 6,
    DPE_ERR_INTERNAL = 7          /* a proven bound of the library was exceeded (e.g. fusion rounds); a bug */
} dpe_status;

/* RunState, main.h:69-73 */
typedef enum dpe_run_state { DPE_FIRST_INIT = 0, DPE_REFINE_INIT = 1, DPE_REFINE_ITER = 2 } dpe_run_state;
/* PixelState, main.h:75-79 */
typedef enum dpe_pixel_state { DPE_WEAK = 0, DPE_STRONG = 1, DPE_UNKNOWN = 2 } dpe_pixel_state;

typedef enum dpe_memory { DPE_MEM_HOST = 0, DPE_MEM_DEVICE = 1 } dpe_memory;

/* Camera, main.h:53-62: 112 bytes, byte-identical.  K and R row-major,
 * R maps world to camera, c = -R^T t. */
typedef struct dpe_camera {
    float K[9];
    float R[9];
    float t[3];
    float c[3];
    int32_t height;
    int32_t width;
    float depth_min;
    float depth_max;
} dpe_camera;

/* PatchMatchParams, main.h:81-109 (bools widened to int32), plus the
 * RNG seed and the index of this process call (both select the counter-based
 * random streams; the reference seeds curand from clock64(), DPE.cu:1036). */
typedef struct dpe_params {
    int32_t max_iterations;   /* 3 */
    int32_t num_images;       /* ref + sources; set from the inputs */
    float sigma_spatial;      /* 5 */
    float sigma_color;        /* 3 */
    int32_t top_k;            /* 4 */
    float depth_min;          /* camera depth_min * 0.6 (DPE.cpp:953) */
    float depth_max;          /* camera depth_max * 1.2 (DPE.cpp:954) */
    int32_t geom_consistency;
    int32_t strong_radius;    /* 5 */
    int32_t strong_increment; /* 2 */
    int32_t weak_radius;      /* 5 */
    int32_t weak_increment;   /* 5 */
    int32_t use_APD;
    int32_t use_edge;
    int32_t use_limit;
    int32_t use_label;
    int32_t use_radius;
    int32_t high_res_img;
    int32_t max_scale_size;
    int32_t scale_size;
    int32_t weak_peak_radius; /* 2 */
    int32_t rotate_time;      /* 4 */
    float ransac_threshold;   /* 0.005 */
    float geom_factor;        /* 0.2 */
    int32_t state;            /* dpe_run_state */
    int32_t call_index;       /* which process call this is (RNG stream) */
    uint64_t seed;            /* RNG key */
} dpe_params;

/* Inputs of one per-reference-image process call: what
 * DPE::InuputInitialization / SupportInitialization gather (DPE.cpp:934-1069,
 * 1187-1214) before CudaSpaceInitialization uploads them (DPE.cpp:1071-1185). */
typedef struct dpe_inputs {
    int32_t width;
    int32_t height;
    int32_t num_images;                     /* ref first, then sources; 2..32 */
    int32_t memory;                         /* dpe_memory of every pointer below */
    const float *images[DPE_MAX_IMAGES];    /* width*height fp32 gray, row-major */
    dpe_camera cameras[DPE_MAX_IMAGES];     /* already scaled to width x height */
    const float *geom_depths[DPE_MAX_IMAGES]; /* per view depth maps; geom_consistency only */
    const uint8_t *weak_info;               /* pixel states; NULL => all STRONG (use_APD off) */
    const float *prior_planes;              /* float4 per pixel: world normal xyz, depth w; not FIRST_INIT */
    const uint32_t *prior_selected;         /* selected-view bits; not FIRST_INIT */
    const uint8_t *edges;                   /* edge map (use_edge or use_limit) */
    const uint8_t *edges_low;               /* edge map at the coarsest scale */
    int32_t low_width;
    int32_t low_height;
    const int32_t *labels;                  /* label map (use_label): 0 edge, -1 small, >0 region */
} dpe_inputs;

/* Outputs: what RunPatchMatch downloads (DPE.cu:3258-3260) plus the cost map. */
typedef struct dpe_outputs {
    float *planes;            /* float4 per pixel: world normal xyz, depth w */
    float *costs;             /* per pixel final cost */
    uint8_t *pixel_states;    /* dpe_pixel_state per pixel */
    uint32_t *selected_views; /* bit i = source view i+1 selected */
} dpe_outputs;

/* NCC kernel families of dpe_stats.family_*. */
#define DPE_NCC_FAMILIES 4
#define DPE_FAMILY_FULL 0   /* initial cost vectors + propagation candidates (DPE.cu:1039-1067, 1218-1670) */
#define DPE_FAMILY_REFINE 1 /* refinement hypotheses (PlaneHypothesisRefinementStrong, DPE.cu:1069-1122) */
#define DPE_FAMILY_POST 2   /* DepthToWeak + LocalRefine samples (DPE.cu:2597-2839) */
#define DPE_FAMILY_WEAK 3   /* weak-pixel propagation (CheckerboardPropagationWeak, DPE.cu:1672-1866) */

/* Timing of the last dpe_process / dpe_ncc_cost call (hipEvent pairs around
 * each launch, milliseconds) and exact counts of the NCC plane hypotheses and
 * taps evaluated on the GPU.  One hypothesis = one (pixel, plane, source view)
 * call of ComputeBilateralNCCOld/New; one tap = one bilinear source sample. */
typedef struct dpe_stats {
    float gpu_ms;             /* whole call: GetLastGpuTimeMs (DPE.cpp:1286) */
    float ncc_ms;             /* launches of families FULL, REFINE and POST */
    int64_t ncc_hyps;         /* all families */
    int64_t ncc_taps;         /* all families */
    int32_t ncc_launches;     /* launches of families FULL, REFINE and POST */
    float family_ms[DPE_NCC_FAMILIES];   /* the WEAK family's time is its whole half-iteration (view selection and
                                            acceptance included); family_launches counts half-iterations */
    int64_t family_hyps[DPE_NCC_FAMILIES];
    int64_t family_taps[DPE_NCC_FAMILIES];
    int32_t family_launches[DPE_NCC_FAMILIES];
    int32_t quad_bytes;       /* bytes of one bilinear source quad (one per tap): 16 fp32, 8 fp16 / bf16 */
    int32_t tile_full;        /* tile of the candidate-cost and refinement/post-pass NCC launches: */
    int32_t tile_weighted;    /* 0 = 4 waves per workgroup (default), 1 = 8 waves, 2 = 2 waves,
                                 3 = 4 waves with <= 3 workgroups per CU, 4 = 4 waves with <= 2;
                                 -1 while the autotuner samples (dpe_form.autotune) */
    int32_t graph;            /* 1: the schedule ran as one captured hipGraph (dpe_form.graph) */
} dpe_stats;

typedef struct dpe_ctx dpe_ctx;

/* Context = one GPU, one HIP stream and a device memory pool (GpuWorkspace,
 * DPE.h:67-84).  One context per host thread.  Replaces cudaSetDevice
 * (main.cpp:343) and the GpuWorkspace constructor.  Loads the library's GPU
 * code on the device (once per process; the first call pays no load). */
dpe_status dpe_create(int device, dpe_ctx **out);
void dpe_destroy(dpe_ctx *ctx);
/* Last error message of this context (or of dpe_create when ctx is NULL). */
const char *dpe_last_error(const dpe_ctx *ctx);
int dpe_abi_version(void);

/* One per-image process call.  Replaces the DPE object's
 * InuputInitialization..RunPatchMatch sequence (main.cpp:249-262, DPE.cu:3130-3262);
 * `params` arrive with depth_min/max already scaled as DPE.cpp:953-954 does. */
dpe_status dpe_process(dpe_ctx *ctx, const dpe_params *params,
                       const dpe_inputs *in, dpe_outputs *out);
dpe_status dpe_last_stats(const dpe_ctx *ctx, dpe_stats *stats);

/* Forms of the process call beyond the plain path (no counterpart in the
 * reference; results are not those of the reference's algorithm and are
 * judged against ground truth, not the oracle). */
#define DPE_PRECISION_FP32 0 /* plain */
#define DPE_PRECISION_FP16 1 /* precision form: fp16 source gathers ... */
#define DPE_PRECISION_BF16 2 /* ... or bf16; both with FP64 homography set-up and shifted (centred) NCC sums */
typedef struct dpe_form {
    int32_t gauss_newton_steps; /* 0: plain.  1 or 2: of the 5 hypotheses of
                                   PlaneHypothesisRefinementStrong (DPE.cu:1080-1104), the perturbed
                                   normal (4th) and then the random depth + normal (3rd) become
                                   Gauss-Newton steps on the per-view costs as residuals (gradients
                                   from the adjoint): from the current plane, then from step 1 */
    int32_t precision;          /* DPE_PRECISION_*.  FP16 / BF16: the source quads are stored in that
                                   format (8 B per bilinear cell instead of 16), the homography is
                                   composed in FP64 and rounded to fp32, and the NCC sums of the
                                   shared-patch kernels (propagation, refinement, DepthToWeak,
                                   LocalRefine) are taken on values centred at the window's centre
                                   (reference texel, source sample at its projection), accumulated
                                   in fp32.  The weak-texture windows read the same 16-bit quads with
                                   the plain sums. */
    int32_t autotune;           /* 1: the shared-patch NCC launches use the tile (waves per workgroup,
                                   workgroups per CU) the context's autotuner chose for this problem
                                   class (image size, source count, window slots, precision).  The
                                   first calls of a class time every tile on their half-iteration
                                   launches, one tile per half-iteration, then the fastest is kept.
                                   Results are bit-identical for every tile. */
    int32_t graph;              /* 1: the call's launch schedule is captured into a hipGraph on the
                                   context's stream and launched as one graph; the context keeps
                                   the executable graph and updates it in place when the next
                                   call's schedule has the same shape.  Per-family launch times
                                   (dpe_stats.family_ms) are not taken in this form.  Calls in
                                   which the autotuner still samples run without the graph. */
    int32_t reserved[4];        /* 0 */
} dpe_form;

/* dpe_process with a form; form == NULL or all zero is dpe_process. */
dpe_status dpe_process_ex(dpe_ctx *ctx, const dpe_params *params, const dpe_form *form,
                          const dpe_inputs *in, dpe_outputs *out);

/* Multi-view bilateral NCC cost vectors, the device function
 * ComputeMultiViewCostVectorOld (DPE.cu:874-883) exposed for tests and
 * measurement: for every pixel p and hypothesis h, planes[(h*H*W + p)*4..+3]
 * is a camera-frame plane (n, d) and costs[(h*(num_images-1) + v)*H*W + p] the
 * cost against source view v+1.  num_hyps >= 1. */
dpe_status dpe_ncc_cost(dpe_ctx *ctx, const dpe_params *params, const dpe_inputs *in,
                        int num_hyps, const float *planes, float *costs);

/* dpe_ncc_cost that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Inputs,
 * planes and costs must be device-resident (in->memory == DPE_MEM_DEVICE).
 * Every buffer of the call (a copy of the images, the bilinear quads, camera
 * constants, counters) comes from a 64 MiB scratch array that each GPU holds
 * from its first dpe_create until the process ends; a problem needs at most
 * 4 B x W*H x num_images + 16 B x (W+1)(H+1) x num_images + 64 KiB of it, and a
 * larger one returns DPE_ERR_UNSUPPORTED before anything runs.  Calls that
 * overlap on one GPU (several contexts or threads) share the array and race.
 * Results equal dpe_ncc_cost's bit for bit; dpe_last_stats reports the counts
 * and no launch times. */
dpe_status dpe_ncc_cost_noalloc(dpe_ctx *ctx, const dpe_params *params, const dpe_inputs *in,
                                int num_hyps, const float *planes, float *costs);

/* dpe_ncc_cost with a form (its precision field; NULL = plain).  No reference counterpart. */
dpe_status dpe_ncc_cost_ex(dpe_ctx *ctx, const dpe_params *params, const dpe_form *form, const dpe_inputs *in,
                           int num_hyps, const float *planes, float *costs);

/* Adjoint form (no counterpart in the reference, which has no derivatives):
 * the aggregated multi-view NCC cost C = sum_v w_v c_v / sum_v w_v of every
 * pixel's plane, c_v = ComputeBilateralNCCOld (DPE.cu:696-782), and its
 * gradient with respect to the plane parameters.
 *   plane_params: float4 per pixel (theta, phi, depth, unused); the
 *     reference-camera normal is n = (sin t cos p, sin t sin p, -cos t) and
 *     depth is the plane's z at the pixel (d = -n . depth K^-1 [x, y, 1]).
 *   view_weights: float [num_images-1][H*W] (w_v per pixel), or NULL for 1.
 *   costs: float [H*W]; grads: float4 per pixel (dC/dtheta, dC/dphi,
 *     dC/ddepth, 0).  Views whose centre projects outside the source, whose
 *     patch variance is under the floor or whose cost is clamped to 0 or 2
 *     contribute their cost and no gradient.
 * Needs the 11x11 step-2 window (strong_radius 5, strong_increment 2). */
dpe_status dpe_ncc_cost_adjoint(dpe_ctx *ctx, const dpe_params *params, const dpe_inputs *in,
                                const float *plane_params, const float *view_weights, float *costs, float *grads);

/* Fusion (RunFusion, DPE.cpp:1387-1548).  One entry per reference image in
 * pair order; src_index lists indices into the same array. */
typedef struct dpe_fuse_view {
    int32_t width;
    int32_t height;
    dpe_camera camera;           /* at width x height */
    const float *depth;          /* 0 = no depth */
    const float *normal;         /* 3 floats per pixel, world frame */
    const uint8_t *pixel_states;
    const uint8_t *color;        /* 3 bytes per pixel, BGR */
    int32_t num_src;
    int32_t src_index[DPE_MAX_IMAGES];
} dpe_fuse_view;

typedef struct dpe_fuse_params {
    float max_reproj_px;      /* 2.0 (DPE.cpp:1508) */
    float max_rel_depth;      /* 0.01 */
    float max_angle_rad;      /* 0.174533 */
    float strong_factor;      /* 0.3 (DPE.cpp:1518) */
    float weak_factor;        /* 0.45 */
    int32_t min_consistent;   /* 1 in the reference (DPE.cpp:1519) */
    int32_t memory;           /* dpe_memory of the view pointers and of `points` */
    int32_t fp64_reprojection; /* 0: the reference's fp32 geometry; 1 (precision form): back-projection,
                                  projection, reprojection error and relative depth in FP64 (points are
                                  emitted rounded to fp32; the angle and the weighting stay fp32) */
} dpe_fuse_params;

/* Points: xyz float + bgr uchar, the PLY vertex of ExportPointCloud
 * (DPE.cpp:532-572).  Call with points==NULL to get the count only. */
typedef struct dpe_point {
    float x, y, z;
    uint8_t b, g, r, pad;
} dpe_point;

dpe_status dpe_fuse(dpe_ctx *ctx, const dpe_fuse_view *views, int num_views,
                    const dpe_fuse_params *params, dpe_point *points, int64_t capacity,
                    int64_t *num_points);

/* Tanks-and-Temples fusions, k-of-n consistency: RunFusion_TAT_Intermediate
 * (DPE.cpp:1550-1718) and RunFusion_TAT_advanced (DPE.cpp:1720-1867).  Views
 * as for dpe_fuse (pixel_states unused, may be NULL).  A pixel of view i with
 * depth > 0 (and block >= 128 when blocks are given) keeps, per source, the
 * last measurement (reprojection error, relative depth, normal angle) taken at
 * this or an earlier pixel of the view in raster order where the source pixel
 * was in bounds, unmasked and had depth > 0; for k = 2..num_src it counts the
 * sources under k*dist_base, k*depth_base (and, intermediate only,
 * k*angle_grad + angle_base) and emits its back-projection at the first k
 * with count >= k, then masks itself in view i.  Colour: intermediate
 * averages the reference and counted source pixels, advanced takes the
 * reference pixel.
 *   blocks: NULL, or num_views pointers to width x height u8 maps (the
 *     reference's blocks/mask_<id>.jpg read at the depth map's size; a NULL
 *     entry lets every pixel of that view through). */
#define DPE_TAT_INTERMEDIATE 0
#define DPE_TAT_ADVANCED 1
typedef struct dpe_fuse_tat_params {
    int32_t variant;          /* DPE_TAT_* */
    float dist_base;          /* 0.25 (DPE.cpp:1555, 1725) */
    float depth_base;         /* 1/3500 intermediate (DPE.cpp:1556), 1/3000 advanced (DPE.cpp:1726) */
    float angle_base;         /* 0.06981317 rad, intermediate only (DPE.cpp:1558) */
    float angle_grad;         /* 0.05235988 rad, intermediate only (DPE.cpp:1559) */
    int32_t memory;           /* dpe_memory of the view pointers, blocks and `points` */
} dpe_fuse_tat_params;

dpe_status dpe_fuse_tat(dpe_ctx *ctx, const dpe_fuse_view *views, int num_views, const uint8_t *const *blocks,
                        const dpe_fuse_tat_params *params, dpe_point *points, int64_t capacity,
                        int64_t *num_points);

/* Batched form (no counterpart in the reference, whose main.cpp processes one
 * problem at a time): many independent problems of uneven image size and
 * source count in one call.  A batch owns `lanes` contexts on one GPU, each
 * with its own stream and buffer pool, driven by one host thread each; the
 * problems go to the lanes longest-first (estimated NCC work), so small
 * problems run concurrently and fill the GPU.  Each problem is one
 * dpe_process_ex call with `form` (typically autotune = 1, graph = 1); the
 * lanes share the autotuner's decisions between batches.  Results equal the
 * problems' single calls bit for bit. */
#define DPE_MAX_LANES 16
typedef struct dpe_batch dpe_batch;
dpe_status dpe_batch_create(int device, int lanes, dpe_batch **out);
void dpe_batch_destroy(dpe_batch *batch);
const char *dpe_batch_last_error(const dpe_batch *batch);
/* params, inputs, outputs: n each; stats: n or NULL; wall_ms: host wall time
 * of the whole batch (all lanes synchronised), or NULL. */
dpe_status dpe_batch_process(dpe_batch *batch, int n, const dpe_params *params, const dpe_form *form,
                             const dpe_inputs *inputs, dpe_outputs *outputs, dpe_stats *stats,
                             float *wall_ms);

#ifdef __cplusplus
}
#endif
#endif /* DPE_HIP_H_ */
