// dpe_state.h -- device buffers of one process call and the launch interface
// between the C-ABI (dpe_capi.hip) and the kernels (dpe_kernels.hip).
//
// HBM layout (P = width*height, Q = pixels of one checkerboard colour as the
// reference's HALF launch covers them, DPE.cu:3148-3155):
//   quads      float4[N][(H+1)(W+1)]   bilinear quads of every view (16 B/tap), or
//              half4[N][(H+1)(W+1)]    the same in fp16 (8 B/tap) when fp16-exact
//   ref        float[P]                reference image (patch texels)
//   gdepth     float[N][P]             geometric-consistency depth maps
//   planes     float4[P]               (n, d) camera frame during iterations
//   costs, selected, view_weight (32 B/px), weak, fit planes, ... as DataPassHelper
//   cand_pos   int[17][Q]              candidate pixel of every hypothesis slot
//   cost_mat   float[2][17][NV][Q]     per-view NCC cost of every slot, per colour
//              (a colour's rows stay for its next half-iteration)
//   last_cand  float4[2][17][Q]        plane of each slot's last evaluated row
//   cur_cost   float[NV][P], cur_valid u32[P]   per-view NCC cost of each pixel's
//              current plane, for the views whose bit is set (known from the
//              evaluation that put the plane there)
//   rplane     float4[5][Q], rcost float[5][Q]   refinement hypotheses
#pragma once
#include "dpe_device.h"

namespace dpe {

constexpr int kStrongSlots = 17;  // 0 = current plane, 1..8 candidates, 9..16 fixed-step candidates
constexpr int kRefineSlots = 5;
constexpr int kPostSlots = 73;
constexpr int kWeakPatchTaps = 36 + 8 * 9;  // centre window + 8 neighbour windows of ComputeBilateralNCCNew    // 0..60 DepthToWeak samples, 61 LocalRefine now, 62..72 LocalRefine samples

// Tile of a shared-patch NCC launch: 64 pixels x `waves` waves per
// workgroup, plus `lds_pad` bytes of dynamic LDS per workgroup that caps the
// workgroups resident per CU (the waves-per-CU knob).  Chosen per problem
// class by the autotuner; the default is 4 waves and no pad.
struct NccTile {
    int waves = 4;
    int lds_pad = 0;
};

// Per-entry state of a weak half-iteration between its kernels (k_weak_select
// -> k_weak_mid -> k_weak_final), one per entry of the colour's weak list.
struct WeakState {
    float4 cand;       // the argmin candidate's plane
    float4 plane_now;  // best plane so far (k_weak_mid)
    float4 hyp[5];     // refinement hypotheses (k_weak_mid)
    float fc_mi;       // the argmin candidate's final cost
    float wn;          // view weight norm
    float cost_now;    // best cost after the candidate and the fitted plane (k_weak_mid)
    uint32_t tsel;     // views selected with the candidates
    uint32_t rng_n;    // draws taken from the entry's random stream
    uint32_t flags;    // kWeakActive | kWeakCand | kWeakFit | kWeakRandom
    uint32_t pad[2];
};
constexpr uint32_t kWeakActive = 1, kWeakCand = 2, kWeakFit = 4, kWeakRandom = 8;
// slots of the weak window terms: the current plane, the fitted plane, the 5 hypotheses
constexpr int kWeakTermSlots = 7;

struct Dev {
    int W, H, N, NV, P;
    int half_rows, Wh, Q;   // colour indexing
    int ntap, tap_off0, tap_inc;  // strong window: taps per dimension, first offset, increment
    size_t qs;              // quad stride per view
    dpe_params prm;
    const dpe_camera *cams;
    const ViewConst *vc;
    const float *spatial;   // [ntap*ntap] spatial weight exponents
    const float *ref;
    const float4 *quads;     // fp32 quads (quad_fmt 0)
    const uint2 *quads_h;    // 16-bit quads, 4 texels per cell: fp16 (quad_fmt 1) or bf16 (2)
    int quad_fmt;
    int prec;                // 0 plain; else the precision form (dpe_form.precision)
    const struct ViewConst64 *vc64;  // double-precision view constants (precision form)
    const float *gdepth;
    float4 *planes;
    float *costs;
    uint32_t *selected;
    uint8_t *vw;
    float4 *fit;
    uint8_t *weak;
    uint8_t *weak_rel;
    short2 *nearest;
    short2 *neighbours;
    short2 *edge_neigh;
    float *complexity;
    short2 *label_bound;
    int *radius;
    const int *nmap;
    const uint8_t *edge;
    const uint8_t *edge_low;
    int low_w, low_h;
    const int *label;
    // strong half-iteration scratch
    int *cand_pos;
    uint8_t *cand_alias;  // [17][Q] cost_mat row holding each slot's plane (k_strong_candidates)
    float *cost_mat;      // the launched colour's [17][NV][Q] rows
    float4 *last_cand;    // [2][17][Q] plane of each slot's last evaluated row (all-ones: none)
    uint32_t *nsel;
    float4 *rplane;
    float *rcost;
    float4 *now_plane;
    float *now_depth;
    float *now_cost;
    int *now_row;         // cost_mat row of now_plane (0 = the current plane)
    float *cur_cost;      // per-view costs of the current planes (layout above)
    uint32_t *cur_valid;  // views of cur_cost that hold the current plane's cost
    float *wnorm;
    // post-pass scratch (DepthToWeak / LocalRefine)
    float4 *post_op;
    float *post_bl;
    float *post_disp;
    float *post_wn;
    uint8_t *post_flags;   // bit0 DepthToWeak evaluates, bit1 LocalRefine evaluates
    float *pcost;          // [kPostSlots][P]
    float4 *post_planes;   // [kPostSlots][P] sample planes (w = NaN: not evaluated)
    // pixels that are WEAK at the start of the call (raster order), all and per colour
    const int *weak_list;
    int weak_count;
    const int *weak_colour_list[2];
    int weak_colour_count[2];
    // reference halves of the weak windows per colour-list entry (k_weak_patch)
    float *weak_wpre[2];    // [kWeakPatchTaps][count] bilateral weight of every tap
    float *weak_wsum[2];    // [3 * kNeighbourNum][count]
    float *weak_cost;       // [8][NV][count] candidate cost rows of the launched colour (k_weak_cand)
    WeakState *weak_st;     // [count] of the launched colour
    float *weak_terms;      // [kWeakTermSlots][NV][count] weighted window terms (k_weak_terms)
    // GenNeighbours constants (DPE.cu:2151-2157, host double math)
    float gn_cos, gn_sin, gn_thr;
    int gn_shift;
    unsigned long long *counters;  // [stripe][2*family] NCC hyps, [..][2*family+1] taps (kCounterStripes)
    // 1: fp16 quads of quarter-grid texels, so windows that project inside the
    // source take the clamp-free gather (sample_fast)
    int fast_gather;
    // tiles of the shared-patch NCC launches: k_ncc_full (candidates, initial
    // costs) and k_ncc_weighted (refinement, post-pass)
    NccTile full_tile, wtd_tile;
};

// Source quad formats (Dev::quad_fmt): fp32 float4 per bilinear cell, or two
// 16-bit texels per 32-bit word (uint2 per cell).
enum QuadFormat { QUAD_FP32 = 0, QUAD_FP16 = 1, QUAD_BF16 = 2 };

// NCC kernel families, each counted and timed on its own (dpe_stats.family_*).
enum NccFamily {
    kFamFull = 0,    // k_ncc_full: initial cost vectors and propagation candidates (8x/16x per pixel)
    kFamRefine = 1,  // k_ncc_weighted<REFINE>: the 5 refinement hypotheses
    kFamPost = 2,    // k_ncc_weighted<POST>: DepthToWeak (61) and LocalRefine (12) samples
    kFamWeak = 3,    // k_weak_update: weak-pixel propagation (ComputeBilateralNCCNew windows)
    kNumFamilies = 4
};
// The NCC counters (per family: hyps, taps) are striped: every wave adds its
// counts to one of kCounterStripes copies, 128 B apart, and the host sums them.
// (One copy took about two million atomic adds per post-pass launch on one
// address, 7 ms of its 46 ms at 1600x1200.)
constexpr int kCounterStripes = 64;
constexpr int kCounterStride = 16;  // 64-bit words per stripe (2 * kNumFamilies used)
constexpr int kCounterWords = kCounterStripes * kCounterStride;

// Launch interface (dpe_kernels.hip); every function enqueues on `stream`.
// Kernels named ncc_* evaluate bilateral NCC windows (the timed hot path).
// one empty launch per source file: loads its code object (dpe_create)
void preload_kernels(hipStream_t s);
void preload_adjoint(hipStream_t s);
void preload_fuse(hipStream_t s);
void launch_make_quads(const Dev &d, const float *img, float4 *quads, hipStream_t s);
void launch_make_quads_h(const Dev &d, const float *img, uint2 *quads, int fmt, hipStream_t s);
// sets *flag (int, zeroed by the caller) when a value of img[0..n) is not fp16-exact
void launch_check_fp16_exact(const float *img, size_t n, int *flag, hipStream_t s);
// The weak-pixel index map and lists of one call from its pixel states
// weak[0..W*H): nmap[i] = rank of i among the WEAK pixels in raster order (0
// for the others); all = the WEAK pixels; col[c] = the WEAK pixels of colour
// (x + y) & 1 == c in rows y < half_rows; the lists in tiled order (tile by
// tile, raster inside a tile); counts = the three list lengths.  col[c] has
// room for (P + 1) / 2 entries.
hipError_t weak_lists_temp_bytes(int W, int H, size_t *bytes);
hipError_t launch_weak_lists(const uint8_t *weak, int W, int H, int half_rows, int *nmap, int *all, int *col0,
                             int *col1, int *counts, void *temp, size_t temp_bytes, hipStream_t s);
void launch_gen_edge_inform(const Dev &d, hipStream_t s);
void launch_find_nearest_strong(const Dev &d, hipStream_t s);
void launch_gen_neighbours(const Dev &d, hipStream_t s);
void launch_neighbour_update(const Dev &d, hipStream_t s);
void launch_random_init_planes(const Dev &d, hipStream_t s);
void launch_ncc_current(const Dev &d, hipStream_t s);          // slot 0 cost vector of every pixel
void launch_random_init_costs(const Dev &d, hipStream_t s);
void launch_strong_candidates(const Dev &d, int colour, int iter, hipStream_t s);
void launch_ncc_strong(const Dev &d, int colour, hipStream_t s);
void launch_strong_select(const Dev &d, int colour, int iter, hipStream_t s);
void launch_ncc_refine(const Dev &d, int colour, hipStream_t s);
void launch_strong_accept(const Dev &d, int colour, hipStream_t s);
void launch_ransac(const Dev &d, int iter, hipStream_t s);
void launch_weak_half(const Dev &d, int colour, int iter, hipStream_t s);
void launch_depth_and_normal(const Dev &d, hipStream_t s);
void launch_filter(const Dev &d, int colour, hipStream_t s);
void launch_post_prep(const Dev &d, hipStream_t s);
void launch_post_planes(const Dev &d, hipStream_t s);
void launch_ncc_post(const Dev &d, hipStream_t s);
void launch_post_final(const Dev &d, hipStream_t s);
void launch_ncc_given(const Dev &d, int num_hyps, const float4 *planes, float *costs, hipStream_t s);
// dpe_adjoint.hip: aggregated cost and its (theta, phi, depth) gradient per pixel
void launch_ncc_adjoint(const Dev &d, const float4 *tpd, const float *view_w, float *cost, float4 *grad,
                        hipStream_t s);
// one Gauss-Newton step of the refinement from src[Q] to dst[Q] (camera-frame planes)
void launch_gn_step(const Dev &d, int colour, const float4 *src, float4 *dst, hipStream_t s);

}  // namespace dpe
