// dpe_oracle.h -- C entry points of the CPU oracle (liboracle_dpe.so).
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg; never by the product.
#ifndef DPE_ORACLE_H_
#define DPE_ORACLE_H_
#include "../include/dpe_hip.h"
#ifdef __cplusplus
extern "C" {
#endif
/* One process call (DPE.cu:3130-3262).  counters (NULL or 12 entries):
 * [0..1] NCC hyps, taps in total; [2+2f], [3+2f] hyps, taps of family f =
 * DPE_FAMILY_* (0..3) and f = 4, evaluations whose results the reference never
 * reads (the GPU library skips them). */
int oracle_process(const dpe_params *params, const dpe_inputs *in, dpe_outputs *out, int nthreads,
                   int64_t *counters);
/* The same call with the multi-view cost (homography, projection, bilinear sample, bilateral weights, window
 * sums, NCC, geometric-consistency cost) evaluated in double from the fp32 inputs and plane hypotheses; states,
 * random streams and decision rules as in oracle_process.  The reference of the precision form's depth error. */
int oracle_process_fp64(const dpe_params *params, const dpe_inputs *in, dpe_outputs *out, int nthreads,
                        int64_t *counters);
/* ComputeMultiViewCostVectorOld for given camera-frame planes (layout of dpe_ncc_cost). */
int oracle_ncc_cost(const dpe_params *params, const dpe_inputs *in, int num_hyps, const float *planes,
                    float *costs, int nthreads);
/* oracle_ncc_cost through the FP64 cost path. */
int oracle_ncc_cost_fp64(const dpe_params *params, const dpe_inputs *in, int num_hyps, const float *planes,
                         float *costs, int nthreads);
/* RunFusion, DPE.cpp:1387-1548. */
int oracle_fuse(const dpe_fuse_view *views, int num_views, const dpe_fuse_params *fp, dpe_point *points,
                int64_t capacity, int64_t *num_points);
/* RunFusion_TAT_Intermediate / RunFusion_TAT_advanced, DPE.cpp:1550-1867 (layout of dpe_fuse_tat). */
int oracle_fuse_tat(const dpe_fuse_view *views, int num_views, const uint8_t *const *blocks,
                    const dpe_fuse_tat_params *tp, dpe_point *points, int64_t capacity, int64_t *num_points);
float oracle_expf(float x);
double oracle_exp(double x);
float oracle_acosf(float x);
void oracle_sincosf(float x, float *s, float *c);
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float oracle_bilinear(const float *im, int W, int H, float u, float v);
void oracle_homography(const dpe_camera *ref, const dpe_camera *src, const float *plane, float *H);
int oracle_point_in_triangle(const int16_t *A, const int16_t *B, const int16_t *C, int px, int py);
/* sort_small, sort_small_weighted, FindMinCostIndex, TransformPDFToCDF (DPE.cu:9-61, 297-311) */
void oracle_sort_small(float *d, int n);
void oracle_sort_small_weighted(int16_t *xy, float *w, int n);
int oracle_find_min_cost_index(const float *costs, int n);
void oracle_pdf_to_cdf(float *probs, int n);
/* FindNearestStrongPoint (DPE.cu:2859-2893) over a pixel-state map; out: (x, y) per pixel */
void oracle_nearest_strong(const uint8_t *weak, int W, int H, int16_t *out);
/* CheckerboardFilterStrong (DPE.cu:1961-2071) of one pixel; returns its filtered depth */
float oracle_filter_strong(const float *depth, const float *costs, const uint8_t *weak, int W, int H, int px,
                           int py);
/* RANSACToGetFitPlane (DPE.cu:2895-3128) of one WEAK pixel on a given state (camera-frame planes, 9 neighbour
 * slots of (x, y)); out: the fitted plane, 0 0 0 0 when none */
void oracle_fit_plane(const dpe_camera *cam, int W, int H, const float *planes, const int16_t *neighbours, int px,
                      int py, uint64_t seed, float *out);
/* GenEdgeInform (DPE.cu:2487-2595) over a whole image; label NULL = use_label off */
void oracle_gen_edge_inform(int W, int H, const uint8_t *edge, const int32_t *label, const uint8_t *weak,
                            int strong_radius, int16_t *edge_neigh, float *complex_map, int16_t *label_bound);
/* DepthToWeak's peak analysis (DPE.cu:2699-2750) over its 61 sampled costs; returns the pixel state */
uint8_t oracle_depth_to_weak_state(const float *costs61, int weak_peak_radius);
/* ComputeGeomConsistencyCost (DPE.cu:919-957) of reference pixel (px, py) for a camera-frame plane against one
 * source camera and its W x H depth map */
float oracle_geom_cost(const dpe_camera *ref, const dpe_camera *src, int W, int H, const float *src_depth, int px,
                       int py, const float *plane);
/* oracle_geom_cost through the FP64 cost path. */
float oracle_geom_cost_fp64(const dpe_camera *ref, const dpe_camera *src, int W, int H, const float *src_depth,
                            int px, int py, const float *plane);
/* the view loop of ComputeMultiViewInitialCost (DPE.cu:832-861) over given costs of source views 1..num_images-1;
 * *selected is the prior mask in, the updated mask out; returns the cost */
float oracle_initial_cost(const float *view_costs, int num_images, uint32_t *selected);
/* GenerateRandomNormal (DPE.cu:365-391) and GeneratePerturbedNormal (DPE.cu:393-428) on the random stream
 * (seed, pixel, phase, call); out: float4 */
void oracle_random_normal(const dpe_camera *cam, int px, int py, float depth, uint64_t seed, uint32_t pixel,
                          uint32_t phase, uint32_t call, float *out);
void oracle_perturbed_normal(const dpe_camera *cam, int px, int py, const float *normal, float perturbation,
                             uint64_t seed, uint32_t pixel, uint32_t phase, uint32_t call, float *out);
/* TransformNormal (to_world = 1) / TransformNormal2RefCam (to_world = 0), DPE.cu:528-546 */
void oracle_transform_normal(const dpe_camera *cam, const float *n, int to_world, float *out);
/* GetDepthandNormal (DPE.cu:1944-1959) of one pixel: camera-frame (n, d) in, (world n, depth) out */
void oracle_depth_and_normal(const dpe_camera *cam, int px, int py, const float *plane, float *out);
/* BresenhamLine (DPE.cu:162-248) between full-resolution pixels A and B over a W x H edge map and a
 * low_w x low_h one; trace receives (up to cap entries) the low-resolution (x, y) of every pixel the two walks
 * test, steps[0] / steps[1] how many each walk tested; returns the result (1 = an edge separates A and B) */
int oracle_bresenham(int W, int H, const uint8_t *edge, int low_w, int low_h, const uint8_t *edge_low, int high_res,
                     int ax, int ay, int bx, int by, int64_t *trace, int cap, int *steps);
/* The ray search of GenNeighbours (DPE.cu:2170-2222) for pixel (px, py) over pixel states and nearest-strong
 * coordinates (x, y per pixel), with the random draws taken cyclically from draws[0..n_draws); points / valid:
 * 64 slots (index 4 * base direction + rotation); returns the number of directions that found a point */
int oracle_direction_search(int W, int H, const uint8_t *weak, const int16_t *nearest, const uint8_t *edge,
                            const uint8_t *edge_low, int low_w, int low_h, int high_res, int rotate_time,
                            int edge_limit, int px, int py, const uint32_t *draws, int n_draws, int16_t *points,
                            uint8_t *valid);
#ifdef __cplusplus
}
#endif
#endif
