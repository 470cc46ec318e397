// dpe_fuse.h -- device layout of the fusion (RunFusion, DPE.cpp:1387-1548).
// All views are packed back to back (offset = first pixel of a view); per pixel:
// depth f32, normal 3 x f32 (world frame), pixel state u8, colour 3 x u8 (BGR),
// mask u8 (the reference's `masks`) and claim i32 (fixpoint bookkeeping).
#pragma once
#include "dpe_device.h"

namespace dpe {

struct FuseViewDev {
    int width, height;
    long offset;
    dpe_camera cam;
    int num_src;
    int src[kMaxImages];
};

struct FuseDev {
    const FuseViewDev *views;
    const float *depth;
    const float *normal;
    const uint8_t *states;
    const uint8_t *color;
    uint8_t *mask;
    int *claim;
    float max_reproj, max_rel, max_angle, strong_factor, weak_factor;
    int min_consistent;
    int fp64;  // RunFusion's reprojection geometry in FP64 (precision form)
    // TAT fusions (DPE.cpp:1550-1867): block map (u8 per pixel, packed like
    // depth; NULL = no blocks) and the k-of-n thresholds
    const uint8_t *block;
    float dist_base, depth_base, angle_base, angle_grad;
    int tat_intermediate;
};

void launch_fuse_eval(const FuseDev &f, int vi, int P, const int *used_old, const uint8_t *acc_old, int *used_new,
                      uint8_t *acc_new, int *changed, hipStream_t s);
void launch_fuse_fill_claims(const FuseDev &f, int vi, int num_src, hipStream_t s);
void launch_fuse_claim(const FuseDev &f, int vi, int P, const int *used, const uint8_t *acc, hipStream_t s);
hipError_t fuse_scan_temp_bytes(int P, size_t *bytes);
hipError_t launch_fuse_compact(const FuseDev &f, int vi, int P, const int *used, const uint8_t *acc, int *flags, int *index,
                         void *temp, size_t temp_bytes, dpe_point *out, long base, hipStream_t s);

// TAT fusions of one reference view: per-source validity keys (j*P + p, or -1),
// their running maximum (the last valid pixel at or before p per source), then
// the k-of-n decision and the compaction in raster order.
void launch_tat_keys(const FuseDev &f, int vi, int P, int *keys, hipStream_t s);
hipError_t tat_scan_temp_bytes(int n, size_t *bytes);
hipError_t launch_tat_last(const int *keys, int *last, int n, void *temp, size_t temp_bytes, hipStream_t s);
void launch_tat_decide(const FuseDev &f, int vi, int P, const int *last, uint8_t *acc, uint32_t *colour,
                       hipStream_t s);
hipError_t launch_tat_compact(const FuseDev &f, int vi, int P, const uint8_t *acc, const uint32_t *colour, int *flags,
                              int *index, void *temp, size_t temp_bytes, dpe_point *out, hipStream_t s);

}  // namespace dpe
