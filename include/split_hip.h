/*
 * split_hip.h — C ABI of the MI355X (gfx950) segmentation hot path.
 *
 * Plain pointers and sizes only: every buffer is caller-owned device memory
 * (hipMalloc'd), points are three SoA float planes of n pixels each
 * (channel c of pixel i at points[c*n + i]), centroids are a 3 x k row-major
 * float matrix (channel c of centroid j at centroids[c*k + j]), labels are
 * int32.  `stream` is a hipStream_t (NULL = the default stream).  Every entry
 * point returns SPLIT_HIP_OK (0) or a nonzero status; the thread-local text of
 * the last failure is available from split_hip_last_error().
 *
 * Each entry point names the reference interface it replaces
 * (nitronoid/split, paths relative to the repository root).  The C++
 * drop-in headers under include/split/ wrap these functions with the
 * reference's namespaces, names and argument order.
 */
#ifndef SPLIT_HIP_H
#define SPLIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "split_gen.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* split_hip_stream_t; /* == hipStream_t */

enum split_hip_status {
  SPLIT_HIP_OK = 0,
  SPLIT_HIP_ERR_ARGUMENT = 1,    /* null pointer, bad size, k or dim       */
  SPLIT_HIP_ERR_TEMP = 2,        /* temp_bytes below the *_workspace size  */
  SPLIT_HIP_ERR_HIP = 3,         /* a HIP runtime call or launch failed    */
  SPLIT_HIP_ERR_RANGE = 4,       /* label value range exceeds what temp holds */
  SPLIT_HIP_ERR_UNSUPPORTED = 5  /* a size beyond a documented limit       */
};

/* Largest k accepted by the assignment kernels (centroids held in registers,
 * 16 per lane). */
#define SPLIT_HIP_MAX_K 1024

/* Result of a Lloyd run. */
typedef struct split_hip_cluster_result {
  int32_t iterations; /* Lloyd iterations performed                          */
  float shift;        /* Frobenius norm of the last centroid update           */
} split_hip_cluster_result;

/* Text of the last failure on the calling thread ("" if none). */
const char* split_hip_last_error(void);
/* Text of a status code. */
const char* split_hip_status_string(int status);
/* Library version, e.g. "0.1.0". */
const char* split_hip_version(void);

/* ---- k-means ---------------------------------------------------------- */

/* Replaces split::device::kmeans::label_points_workspace
 * (split/include/split/device/kmeans/label.cuh:39-40,
 *  split/src/split/device/kmeans/label.cu:68-72).  The fused kernel keeps
 * distances in registers, so the temp it needs is 0 bytes. */
size_t split_hip_label_points_workspace(int64_t n, int32_t k);

/* Replaces split::device::kmeans::label_points
 * (split/include/split/device/kmeans/label.cuh:27-31,
 *  split/src/split/device/kmeans/label.cu:74-106).
 * labels[i] = argmin_j ((p0-c0)^2 + (p1-c1)^2) + (p2-c2)^2 over the k
 * centroids, ties to the lowest j.  1 <= k <= SPLIT_HIP_MAX_K. */
int split_hip_label_points(const float* centroids, int32_t k,
                           const float* points, int64_t n, int32_t* labels,
                           void* temp, size_t temp_bytes,
                           split_hip_stream_t stream);

/* Image form of split_hip_label_points (same result): the caller also gives
 * the row-major width x height layout of the n = width*height pixels, so the
 * kernel can walk 16x16 tiles, whose colours are more compact than runs of a
 * row and let more centroids be ruled out per tile.  The reference has no
 * such overload; its callers hold the labels as a height x width array2d
 * (samples/src/main.cu:234-237). */
int split_hip_label_image(const float* centroids, int32_t k,
                          const float* points, int32_t width, int32_t height,
                          int32_t* labels, split_hip_stream_t stream);

/* Replaces split::device::kmeans::calculate_centroids_workspace
 * (split/include/split/device/kmeans/centroids.cuh:39-40,
 *  split/src/split/device/kmeans/centroids.cu:16-20). */
size_t split_hip_calculate_centroids_workspace(int64_t n, int32_t k);

/* Replaces split::device::kmeans::calculate_centroids
 * (split/include/split/device/kmeans/centroids.cuh:26-30,
 *  split/src/split/device/kmeans/centroids.cu:22-101).
 * centroids[c*k+j] = mean of channel c over pixels labelled j; a cluster with
 * no pixels gets the zero centroid (the reference's stated intent,
 * centroids.cu:65-67).  Labels outside [0, k) contribute nothing.
 * Counts are exact integers.  k <= SPLIT_HIP_MAX_K: float sums reduced in a
 * fixed order.  Larger k (segment means, k = number of segments, the
 * caller's use at samples/src/main.cu:255-307): each channel c quantised
 * with its own scale, q = rint(v * 2^(23-e_c)) (max|v_c| < 2^e_c), and
 * accumulated exactly in int64, so the result does not depend on order and
 * a mean moves by at most max|v_c| * 2^-23; the temp needed is about 28
 * bytes per label. */
int split_hip_calculate_centroids(const int32_t* labels, const float* points,
                                  int64_t n, float* centroids, int32_t k,
                                  void* temp, size_t temp_bytes,
                                  split_hip_stream_t stream);

/* No counterpart in the reference.  split_hip_calculate_centroids for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same kernels and
 * results, with their scratch in a static device array of
 * 16 MiB + 256 bytes that each GPU holds until the process ends, instead of
 * a temp block.  The caller passes only inputs and outputs.  Any
 * k <= SPLIT_HIP_MAX_K (bit-identical to split_hip_calculate_centroids with
 * its workspace); larger k up to about 599,000 labels; beyond that it
 * returns SPLIT_HIP_ERR_UNSUPPORTED.  Calls that overlap on one GPU share the
 * array and race: serialise them (one stream, one host thread). */
int split_hip_calculate_centroids_noalloc(const int32_t* labels,
                                          const float* points, int64_t n,
                                          float* centroids, int32_t k,
                                          split_hip_stream_t stream);

/* Replaces split::device::kmeans::propagate_centroids
 * (split/include/split/device/kmeans/propagation.cuh:24-27,
 *  split/src/split/device/kmeans/propagation.cu:9-24): points[c*n+i] =
 * centroids[c*k + labels[i]] for every pixel (3 x n SoA planes).  Pixels
 * whose label is outside [0, k) keep their values (the reference gathers
 * out of range).  Any k. */
int split_hip_propagate_centroids(const int32_t* labels, int64_t n,
                                  const float* centroids, int64_t k,
                                  float* points, split_hip_stream_t stream);

/* Replaces split::device::kmeans::cluster_workspace
 * (split/include/split/device/kmeans/cluster.cuh:46-47,
 *  split/src/split/device/kmeans/cluster.cu:12-25).  dim must be 3. */
size_t split_hip_cluster_workspace(int64_t n, int32_t k, int32_t dim);

/* Replaces split::device::kmeans::cluster
 * (split/include/split/device/kmeans/cluster.cuh:31-37,
 *  split/src/split/device/kmeans/cluster.cu:27-93).
 * Lloyd iterations (assign, then centroid update) starting from the
 * centroids passed in; the final centroids are written back to `centroids`.
 * Iteration stops once the Frobenius norm of a centroid update is
 * <= threshold, or after max_iter iterations (threshold < 0 runs exactly
 * max_iter).  As in the reference, `labels` holds the assignment made in the
 * last iteration, i.e. against the centroids before the last update.
 * Host-synchronous on return (like the reference, which reads the shift back
 * every iteration); `result` may be NULL. */
int split_hip_cluster(const float* points, int64_t n, float* centroids,
                      int32_t k, int32_t* labels, void* temp,
                      size_t temp_bytes, int32_t max_iter, float threshold,
                      split_hip_cluster_result* result,
                      split_hip_stream_t stream);

/* Image form of split_hip_cluster (same results) walking 16x16 tiles of the
 * width x height image, as split_hip_label_image. */
int split_hip_cluster_image(const float* points, int32_t width,
                            int32_t height, float* centroids, int32_t k,
                            int32_t* labels, void* temp, size_t temp_bytes,
                            int32_t max_iter, float threshold,
                            split_hip_cluster_result* result,
                            split_hip_stream_t stream);

/* Replaces split::device::kmeans::uniform_random_initialize
 * (split/include/split/device/kmeans/uniform_random_initialize.cuh:20-22,
 *  split/src/split/device/kmeans/uniform_random_initialize.cu:10-55)
 * with a reproducible form: k distinct pixel indices drawn from SplitMix64
 * seeded by `seed` (index = high 64 bits of draw * n, redrawn on a repeat),
 * and their colours copied as the centroids.  The reference seeds with
 * time(NULL) and can read index n; this form cannot.  k <= n. */
int split_hip_uniform_random_initialize(const float* points, int64_t n,
                                        float* centroids, int32_t k,
                                        uint64_t seed,
                                        split_hip_stream_t stream);

/* ---- connected components -------------------------------------------- */

/* Temp needed by split_hip_connected_components: a linked-root bitmap (one
 * bit per pixel) and the labels and local roots of every 64x64 tile's left
 * and right columns, about 0.375 bytes per pixel plus 256.  The reference
 * ignores its temp argument and allocates 2N+3R ints internally
 * (split/src/split/device/ccl/connected_components.cu:93-94,140-142); its
 * callers pass the clustering temp (at least N ints), which is larger. */
size_t split_hip_connected_components_workspace(int32_t width, int32_t height);

/* Replaces split::device::ccl::connected_components
 * (split/include/split/device/ccl/connected_components.cuh:25-29,
 *  split/src/split/device/ccl/connected_components.cu:80-181).
 * labels: height x width row-major cluster labels.  out[p] = the largest
 * linear pixel index of p's connected component of equal labels — the value
 * the reference's max-propagation reaches at convergence — under 4- or
 * 8-connectivity.  The union-find always runs to completion, so max_iter is
 * accepted for signature parity and ignored.  Passes: per-tile LDS
 * union-find, tile-seam unions, root jumps, a per-tile final resolution. */
int split_hip_connected_components(const int32_t* labels, int32_t width,
                                   int32_t height, void* temp,
                                   size_t temp_bytes, int32_t* out,
                                   int32_t max_iter, int32_t connectivity,
                                   split_hip_stream_t stream);

/* Temp needed by split_hip_compress_labels for n labels whose values span
 * at most `value_span` distinct integers (max - min + 1); pass n for labels
 * produced by split_hip_connected_components. */
size_t split_hip_compress_labels_workspace(int64_t n, int64_t value_span);

/* Replaces split::device::ccl::compress_labels
 * (split/include/split/device/ccl/compress_labels.cuh:21-23,
 *  split/src/split/device/ccl/compress_labels.cu:10-35).
 * In place: labels[i] <- rank of labels[i] among the sorted distinct values.
 * *num_unique receives the number of distinct values (host-synchronous).
 * Fails with SPLIT_HIP_ERR_RANGE if max - min + 1 exceeds what temp holds. */
int split_hip_compress_labels(int32_t* labels, int64_t n, void* temp,
                              size_t temp_bytes, int32_t* num_unique,
                              split_hip_stream_t stream);

/* ---- whole segmentation path ----------------------------------------- */

/* cluster -> connected_components -> compress_labels in one stream-ordered
 * call: the caller sequence of samples/src/main.cu:240-252.  Writes the
 * cluster labels (`labels`, as split_hip_cluster) and the compressed segment
 * labels (`segments`); *num_segments gets the segment count.  `compress` = 0
 * leaves `segments` as component representatives.  temp must hold
 * split_hip_segment_workspace(width, height, k) bytes. */
size_t split_hip_segment_workspace(int32_t width, int32_t height, int32_t k);
int split_hip_segment(const float* points, int32_t width, int32_t height,
                      float* centroids, int32_t k, int32_t* labels,
                      int32_t* segments, void* temp, size_t temp_bytes,
                      int32_t max_iter, float threshold, int32_t connectivity,
                      int32_t compress, split_hip_cluster_result* result,
                      int32_t* num_segments, split_hip_stream_t stream);

/* ---- batched segmentation ----------------------------------------------
 * No counterpart in the reference, whose caller segments one image per
 * cluster -> connected_components -> compress_labels sequence
 * (samples/src/main.cu:240-252).  This runs that sequence for many images of
 * uneven size and k at once; each image's results equal what
 * split_hip_segment gives for it alone (same centroids, labels, segments,
 * iteration count and shift). */
#define SPLIT_HIP_BATCH_MAX_K 64

typedef struct split_hip_image_desc {
  const float* points;  /* 3 SoA planes of width*height floats */
  float* centroids;     /* [3][k], in: initial, out: final */
  int32_t* labels;      /* width*height cluster labels */
  int32_t* segments;    /* width*height segment labels (unused, may be NULL,
                           with connectivity 0) */
  int32_t width, height;
  int32_t k;            /* 1..SPLIT_HIP_BATCH_MAX_K */
  int32_t reserved;
} split_hip_image_desc;

typedef struct split_hip_batch_result {
  int32_t iterations;
  float shift;
  int32_t num_segments; /* -1 when compress == 0 or connectivity == 0 */
  int32_t reserved;
} split_hip_batch_result;

/* Bytes of temp for a batch: a descriptor table plus, per image, centroid
 * buffers, sums and the root bitmap (about 0.2 bytes per pixel).  0 for an
 * invalid batch. */
size_t split_hip_segment_batch_workspace(const split_hip_image_desc* images,
                                         int32_t count);
/* Every image: clustering from its centroids (max_iter / threshold as
 * split_hip_cluster), connected components of its labels (connectivity 4
 * or 8; 0: the clustering alone) and, with `compress`, compressed segment
 * labels.  use_graph != 0 captures the
 * launch sequence into a hipGraph on first use (cached per temp pointer and
 * batch shape) and replays it.  results (count entries) may be NULL.  The
 * call returns after the batch finished on `stream`. */
int split_hip_segment_batch(const split_hip_image_desc* images, int32_t count,
                            void* temp, size_t temp_bytes, int32_t max_iter,
                            float threshold, int32_t connectivity,
                            int32_t compress, int32_t use_graph,
                            split_hip_batch_result* results,
                            split_hip_stream_t stream);

/* ---- seeded synthetic images on the device ------------------------------
 * No counterpart in the reference (whose inputs are image files loaded by
 * samples/src/main.cu:81-132).  The construction of split_gen_image
 * (split_gen.h) run on the GPU for many images at once: same hashes and
 * geometry; noisy pixels may differ from the host image by one float ulp
 * (device log/cos). */
typedef struct split_hip_gen_desc {
  float* planes;  /* out: 3 SoA planes of width*height floats */
  int32_t width, height;
  split_gen_params params;
} split_hip_gen_desc;

size_t split_hip_gen_images_workspace(const split_hip_gen_desc* images,
                                      int32_t count);
/* Returns after the images are written. */
int split_hip_gen_images(const split_hip_gen_desc* images, int32_t count,
                         void* temp, size_t temp_bytes,
                         split_hip_stream_t stream);

/* ---- learned warm start --------------------------------------------------
 * No counterpart in the reference: its cluster() takes the initial
 * centroids as an in/out argument (split/include/split/device/kmeans/
 * cluster.cuh:19-22,31-37) and its caller fills them by random sampling
 * (samples/src/main.cu:230-233).  This network predicts those 16 starting
 * centroids from the image.
 *
 * Network: input (B, 256, 256, 3) bf16 NHWC; four 3x3 stride-2 convolutions
 * 3->16->32->64->64 with ReLU; global average pool; dense 64->128 ReLU;
 * dense 128->48 with sigmoid = 16 centroids in the (3, k) layout.
 * Parameters: split_hip_ws_param_count() floats; layer l holds W_l[n][Kp_l]
 * row-major, columns (ky*3+kx)*Cin + ci (conv) or ci (dense), the bias at
 * column K_l (= 9*Cin or Cin), zeros up to Kp_l = roundup(K_l + 1, 32);
 * layers in order conv1..conv4, dense5, dense6.  Weights and activations
 * are rounded to bf16 for the matrix cores, which accumulate in fp32. */
#define SPLIT_HIP_WS_K 16
#define SPLIT_HIP_WS_SIDE 256

size_t split_hip_ws_param_count(void);
/* Bytes of ws for a batch of `batch` images (activations and gradients). */
size_t split_hip_ws_workspace(int32_t batch);
/* Resamples images (points, width, height of each desc) to the network
 * input: `input` = count x 256 x 256 x 3 bf16 (uint16 bits), area average.
 * temp holds count * 16 bytes.  Returns after the input is written. */
int split_hip_ws_pack(const split_hip_image_desc* images, int32_t count,
                      uint16_t* input, void* temp, size_t temp_bytes,
                      split_hip_stream_t stream);
/* The same area-average resample as float SoA planes: `planes` = count x 3
 * x 256^2 floats, the coarse images on which predicted centroids are
 * refined before the full-size run.  temp holds count * 16 bytes.  Returns
 * after the planes are written. */
int split_hip_ws_coarse(const split_hip_image_desc* images, int32_t count,
                        float* planes, void* temp, size_t temp_bytes,
                        split_hip_stream_t stream);
/* Forward pass over `batch` inputs: input slot b is image index[b] of
 * `input` (index NULL: slot b).  out (batch x 48 floats, may be NULL)
 * receives the sigmoid outputs.  Keeps the activations in ws. */
int split_hip_ws_forward(const float* params, const uint16_t* input,
                         const int32_t* index, int32_t batch, void* ws,
                         size_t ws_bytes, float* out,
                         split_hip_stream_t stream);
/* Loss against targets[index[b]] (48 floats per image) and its gradient
 * w.r.t. every parameter, written to grads (param_count floats).
 * SPLIT_HIP_WS_LOSS_MSE: (1/B) sum (out - target)^2 slot by slot;
 * SPLIT_HIP_WS_LOSS_CHAMFER: (1/B) sum over images of the Chamfer set
 * distance sum_t min_p |t-p|^2 + sum_p min_t |p-t|^2 between the 16
 * predicted and the 16 target colours (batch <= 256).  Uses the activations
 * of the preceding split_hip_ws_forward with the same params, input, index,
 * batch and ws.  loss (device float, may be NULL) receives the loss. */
#define SPLIT_HIP_WS_LOSS_MSE 0
#define SPLIT_HIP_WS_LOSS_CHAMFER 1
int split_hip_ws_backward(const float* params, const uint16_t* input,
                          const float* targets, const int32_t* index,
                          int32_t batch, int32_t loss_kind, void* ws,
                          size_t ws_bytes, float* grads, float* loss,
                          split_hip_stream_t stream);
/* One Adam step (bias-corrected, step >= 1) on params with moments m, v. */
int split_hip_ws_adam(float* params, const float* grads, float* m, float* v,
                      int32_t step, float lr, float beta1, float beta2,
                      float eps, split_hip_stream_t stream);
/* Predicted starting centroids for each image: resample, forward, and write
 * the 16 centroids to images[i].centroids (k must be 16).  ws holds
 * split_hip_ws_workspace(count) bytes.  Returns after they are written. */
int split_hip_ws_predict(const float* params,
                         const split_hip_image_desc* images, int32_t count,
                         void* ws, size_t ws_bytes, split_hip_stream_t stream);
/* Training targets: per image, its 16 centroids (3 x 16 floats) reordered
 * by luminance 0.299 R + 0.587 G + 0.114 B, ties to the lower index. */
int split_hip_ws_targets(const float* centroids, int32_t count, float* targets,
                         split_hip_stream_t stream);

/* ---- erosion of segment labels ------------------------------------------ */

/* Temp needed by split_hip_erode and split_hip_erode_horizontal_sweep:
 * width * height int32. */
size_t split_hip_erode_workspace(int32_t width, int32_t height);

/* Replaces split::device::morph::erode_horizontal_sweep
 * (split/include/split/device/morph/erode.cuh:12-14,
 *  split/src/split/device/morph/erode.cu:57-139): in each row of the
 * height x width labels, a run [s, e) of equal labels keeps the pixels
 * s + n <= p < e - n (n = niterations) and every other pixel becomes -1;
 * i.e. a pixel survives iff the 2n + 1 pixels centred on it in its row lie
 * inside the row and carry its label.  In place; niterations = 0 leaves the
 * labels unchanged (the reference's 0 writes past its buffer and drops every
 * other run). */
int split_hip_erode_horizontal_sweep(int32_t* labels, int32_t width,
                                     int32_t height, int32_t niterations,
                                     void* temp, size_t temp_bytes,
                                     split_hip_stream_t stream);

/* Replaces split::device::morph::erode
 * (split/include/split/device/morph/erode.cuh:16-18,
 *  split/src/split/device/morph/erode.cu:141-159): the sweep over the
 * columns, then over the rows of the result.  Together: a pixel keeps its
 * label iff the (2n + 1) x (2n + 1) square centred on it lies inside the
 * image and carries one label; every other pixel becomes -1.  In place. */
int split_hip_erode(int32_t* labels, int32_t width, int32_t height,
                    int32_t niterations, void* temp, size_t temp_bytes,
                    split_hip_stream_t stream);

/* ---- colour features feeding k-means ------------------------------------ */

/* Conversions of split/include/split/device/color/conversion.cuh:15-62
 * (functors rgb_to_xyz, xyz_to_rgb, xyz_to_lab, lab_to_xyz, rgb_to_lab,
 * lab_to_rgb, rgb_to_ic; formulas in detail/conversion.inl:13-139). */
enum split_hip_color_conversion {
  SPLIT_HIP_RGB_TO_XYZ = 0,
  SPLIT_HIP_XYZ_TO_RGB = 1,
  SPLIT_HIP_XYZ_TO_LAB = 2,
  SPLIT_HIP_LAB_TO_XYZ = 3,
  SPLIT_HIP_RGB_TO_LAB = 4,
  SPLIT_HIP_LAB_TO_RGB = 5,
  SPLIT_HIP_RGB_TO_IC = 6
};

/* Replaces split::device::color::convert_color_space
 * (split/include/split/device/color/conversion.cuh:67-82): out = the
 * conversion of each pixel of `in` (3 x n SoA planes each; out == in is
 * allowed, any other overlap is an argument error).  sRGB companding,
 * sRGB/XYZ matrices and CIE Lab with a D65 white (X, Y, Z scaled to 100). */
int split_hip_convert_color_space(const float* in, float* out, int64_t n,
                                  int32_t conversion,
                                  split_hip_stream_t stream);

/* Temp needed by split_hip_beta_feature and split_hip_lab_beta_feature
 * (4 KiB: copies of the two maxima). */
size_t split_hip_beta_feature_workspace(void);

/* Replaces split::device::color::beta_feature
 * (split/include/split/device/color/beta_feature.cuh:13-16,
 *  split/src/split/device/color/beta_feature.cu:48-78).  lab: 3 x n planes
 * (L, a, b).  With Lmax = max L and lambda = 0.2 * max |a| (the a plane
 * only, as the reference): beta[i] = -mu if |a_i| < lambda, |b_i| < lambda
 * and L_i < 0.2 Lmax; +mu if |a_i|, |b_i| < lambda and L_i > 0.95 Lmax;
 * 0 otherwise.  beta may be one of lab's planes (the sample writes it over
 * L, samples/src/main.cu:224). */
int split_hip_beta_feature(const float* lab, int64_t n, float* beta, float mu,
                           void* temp, size_t temp_bytes,
                           split_hip_stream_t stream);

/* The sample's feature sequence (samples/src/main.cu:216-225) in two
 * passes: rgb -> Lab into `features` (3 x n), L copied to `luminance`
 * (n floats, may be NULL), then beta (as split_hip_beta_feature, mu) written
 * over the L plane.  features == rgb is allowed. */
int split_hip_lab_beta_feature(const float* rgb, int64_t n, float* features,
                               float* luminance, float mu, void* temp,
                               size_t temp_bytes, split_hip_stream_t stream);

/* ---- image ingest and egress -------------------------------------------
 * Host images are pixel-interleaved (c channels per pixel, as
 * split::host::stbi::loadf returns them, split/src/split/host/stbi/
 * stbi_raii.cpp:115-124); the device path uses SoA planes (channel c of
 * pixel i at planes[c*n + i]).  channels = 1..4. */

/* Device temp of the host transfers: n * channels floats. */
size_t split_hip_interleaved_workspace(int64_t n, int32_t channels);

/* Replaces make_device_image (samples/src/main.cu:98-114, one cudaMemcpy2D
 * of 4-byte rows per channel): one contiguous host-to-device copy of the
 * interleaved image into temp, then a device deinterleave into `planes`.
 * Host-synchronous (the host buffer may be reused on return); pinned host
 * memory makes the copy run at full PCIe rate. */
int split_hip_upload_interleaved(const float* host, int64_t n, int32_t channels,
                                 float* planes, void* temp, size_t temp_bytes,
                                 split_hip_stream_t stream);

/* Replaces make_host_image (samples/src/main.cu:116-132): device interleave
 * of `planes` into temp, one contiguous device-to-host copy.
 * Host-synchronous (the host buffer is complete on return). */
int split_hip_download_interleaved(const float* planes, int64_t n,
                                   int32_t channels, float* host, void* temp,
                                   size_t temp_bytes, split_hip_stream_t stream);

/* Device-to-device forms of the two layout changes (stream-ordered):
 * hwc[i*channels + c] <-> planes[c*n + i]. */
int split_hip_deinterleave(const float* hwc, int64_t n, int32_t channels,
                           float* planes, split_hip_stream_t stream);
int split_hip_interleave(const float* planes, int64_t n, int32_t channels,
                         float* hwc, split_hip_stream_t stream);

/* The 8-bit egress of split::host::stbi::quantize
 * (split/src/split/host/stbi/stbi_raii.cpp:135-143) fused with the
 * interleave, on the device: out[i*channels + c] = (uint8)(clamp(v, 0, 1) *
 * 255 + 0.5) with v = planes[c*n + i], the multiply-add rounded once.
 * `out` is device memory of n * channels bytes; stream-ordered. */
int split_hip_quantize_interleaved(const float* planes, int64_t n,
                                   int32_t channels, uint8_t* out,
                                   split_hip_stream_t stream);

/* k-means inertia sum_i min_j |p_i - c_j|^2 (float distances in the
 * assignment's form, summed in double in a fixed order).  No counterpart in
 * the reference, which reports no inertia.  Host-synchronous. */
size_t split_hip_inertia_workspace(void);
int split_hip_inertia(const float* points, int64_t n, const float* centroids,
                      int32_t k, void* temp, size_t temp_bytes,
                      double* inertia, split_hip_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SPLIT_HIP_H */
