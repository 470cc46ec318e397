/*
 * dpe_host.h -- C entry points of libdpe_host.so, the host side of the C++
 * driver (dpe-mvs-fast_amd/host/): the reference's file formats, image pyramid,
 * state rescaling and the support (edge / label) maps, without OpenCV or Boost.
 * The driver binary `dpe_driver` (main.cpp's CLI and schedule) is built on these
 * and on libdpe_hip.so.  Pure CPU code: nothing here touches a GPU.
 *
 * Matrices follow the reference's .dmb container (ReadBinMat / WriteBinMat,
 * DPE.cpp:293-339): int32 version = 1, rows, cols, OpenCV type code, then the
 * raw row-major data.  Type codes used: 0 = CV_8UC1, 4 = CV_32SC1,
 * 5 = CV_32FC1, 21 = CV_32FC3.
 */
#ifndef DPE_HOST_H_
#define DPE_HOST_H_

#include <stdint.h>
#include <stddef.h>

#include "dpe_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DPE_HOST_CV_8UC1 0
#define DPE_HOST_CV_32SC1 4
#define DPE_HOST_CV_32FC1 5
#define DPE_HOST_CV_32FC3 21

/* Bytes per element of a type code (0 for an unknown code). */
int dpe_host_type_size(int type);

/* WriteBinMat (DPE.cpp:319-339).  Returns 0 on success. */
int dpe_host_write_bin_mat(const char *path, int rows, int cols, int type, const void *data);
/* ReadBinMat (DPE.cpp:293-317): reads the header; with data != NULL also the
 * payload (capacity bytes available).  Returns 0 on success, -1 unreadable,
 * -2 version error, -3 capacity too small. */
int dpe_host_read_bin_mat(const char *path, int *rows, int *cols, int *type, void *data, size_t capacity);

/* ReadCamera (DPE.cpp:341-382): "extrinsic" 4x4, "intrinsic" 3x3, then
 * depth_min interval depth_num depth_max; c = -R^T t in double.  width/height
 * are left 0 (the caller sets the image size).  Returns 0 on success. */
int dpe_host_read_camera(const char *path, dpe_camera *cam);

/* GenerateSampleList (main.cpp:102-145): parses pair.txt.  Fills up to
 * max_problems entries of ref_ids[] and, per problem, src_counts[i] source ids
 * (scores <= 0 dropped) into srcs[i * DPE_MAX_IMAGES ...].  Returns the number
 * of problems, or -1 if the file cannot be read. */
int dpe_host_read_pairs(const char *path, int max_problems, int32_t *ref_ids, int32_t *src_counts, int32_t *srcs);

/* cv::resize(..., INTER_LINEAR) of a float image (SceneCache::GetGrayFloatImage,
 * DPE.cpp:720-737): half-pixel centres, clamped borders.  A factor-2 reduction
 * of an even-sized image is the 2x2 box mean. */
void dpe_host_resize_linear(const float *src, int w, int h, float *dst, int dw, int dh);

/* RescaleMatToTargetSize (DPE.cpp:1313-1335): nearest, o = int(r / s);
 * elem_size bytes per element. */
void dpe_host_rescale_nearest(const void *src, int w, int h, void *dst, int dw, int dh, int elem_size);

/* The support maps the weak-texture path consumes (replaces EdgeSegment,
 * DPE.cpp:136-291, with this project's detector; same semantics as
 * dpe_amd/support.py): edges u8 (255 = edge), labels i32 (0 = edge,
 * -1 = small region, > 0 = region).  Either output may be NULL. */
void dpe_host_support_maps(const uint8_t *image, int w, int h, int scale, uint8_t *edges, int32_t *labels);

/* The reference's own segmentation (EdgeSegment, DPE.cpp:136-291) without
 * OpenCV, for the driver's --support=reference: mode 0 = edges (Canny on the
 * image with median-derived thresholds, use_canny = 1 in the reference's call,
 * main.cpp:187), u8 0/255 at w x h; mode 1 = labels (Roberts + threshold +
 * Connect + HoughLinesP borders of large regions at a quarter of the size,
 * use_canny = 0, main.cpp:207), i32 (0 edge, -1 small region, > 0 region) at
 * round(w / 2^scale) x round(h / 2^scale).  *out_w, *out_h receive the output
 * size.  Returns 0, or -1 for bad arguments.  The OpenCV routines are restated
 * from their published algorithms; agreement with an OpenCV build is not
 * checked here. */
int dpe_host_edge_segment(int scale, const uint8_t *image, int w, int h, int mode, int use_canny, int high_res,
                          uint8_t *edges, int32_t *labels, int *out_w, int *out_h);
/* Its building blocks (exposed for tests): Roberts (DPE.cpp:9-25); Connect
 * (DPE.cpp:27-139, returns the number of labels, counts per label up to
 * capacity); cv::Canny (aperture 3, L2 gradient); cv::HoughLinesP (lines as
 * x0 y0 x1 y1, returns the count). */
void dpe_host_roberts(const uint8_t *src, int w, int h, uint8_t *dst);
int dpe_host_connect(const uint8_t *bin, int w, int h, int32_t *labels, int32_t *counts, int capacity);
void dpe_host_canny(const uint8_t *src, int w, int h, double low, double high, uint8_t *edges);
int dpe_host_hough_lines_p(const uint8_t *img, int w, int h, double rho, double theta, int threshold, int min_len,
                           int max_gap, int32_t *lines, int max_lines);

/* cv::imread of a JPEG (DPE.cpp:713 grayscale, DPE.cpp:1422 colour) without
 * OpenCV: sequential-Huffman 8-bit JPEG (baseline / extended; 1 or 3
 * components, any sampling factors, restart intervals) decoded as libjpeg 6b /
 * libjpeg-turbo does by default: islow integer IDCT; grayscale = the Y
 * component; colour = triangle-filter chroma upsampling (h2v1, h2v2) and the
 * 16-bit fixed-point YCbCr -> RGB tables, written as BGR.  out == NULL: only
 * *width, *height.  Returns 0, -1 (cannot open), -2 (not a supported JPEG),
 * -3 (capacity below width * height * (color ? 3 : 1)). */
int dpe_host_read_jpeg(const char *path, int color, uint8_t *out, size_t capacity, int *width, int *height);

/* ExportPointCloud (DPE.cpp:532-572): binary little-endian PLY, xyz float +
 * bgr uchar per vertex.  Returns 0 on success. */
int dpe_host_write_ply(const char *path, const dpe_point *points, int64_t count);

#ifdef __cplusplus
}
#endif
#endif /* DPE_HOST_H_ */
