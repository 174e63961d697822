/*
 * CPU oracle of the segmentation path.  TEST INFRASTRUCTURE ONLY: linked by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker or the CPU baseline; never part of the product path.
 *
 * Two families:
 *  - oracle_*:  the mathematics, in double precision — FP64 nearest-centroid
 *               labels with the top-2 gap, FP64 means, FP64 Lloyd, textbook
 *               union-find connected components (4/8), dense rank.
 *  - ref_*:     a restatement of nitronoid/split's own algorithm, quirks
 *               included (float dense distances, radix-sorted compacted
 *               reduce_by_key centroids with float cumulative counts, the
 *               aliased shift and its stopping rule, max-propagation CC,
 *               sort/unique/lower_bound compress).
 *
 * Layouts: points are 3 SoA float planes of n (channel c of pixel i at
 * pts[c*n+i]); centroids are 3 x k row-major (c*k + j); labels int32.
 */
#ifndef SPLIT_ORACLE_H
#define SPLIT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mathematics (double) ---- */

/* labels[i] = argmin_j ||p_i - c_j||^2 in double (ties to lowest j).
 * gap[i] (nullable) = (d2 - d1) / d2 for the two smallest distances
 * (0 when d2 == 0; 1 when k == 1). */
void oracle_label_f64(const float* pts, int64_t n, const float* cent, int k,
                      int32_t* labels, double* gap);

/* out[c*k+j] = mean over labels == j (double); 0 for an empty label;
 * labels outside [0,k) ignored.  counts (nullable) receives per-label
 * counts. */
void oracle_centroids_f64(const int32_t* labels, const float* pts, int64_t n,
                          int k, double* out, int64_t* counts);

/* Free-running Lloyd in double for exactly `iters` iterations from `init`
 * (float).  Returns centroids (double 3*k) and the labels of the last
 * assignment (against the centroids before the last update). */
void oracle_lloyd_f64(const float* pts, int64_t n, const float* init, int k,
                      int iters, double* cent_out, int32_t* labels_out);

/* Inertia sum ||p_i - c_{l_i}||^2 (double) for given labels. */
double oracle_inertia(const float* pts, int64_t n, const float* cent, int k,
                      const int32_t* labels);

/* Union-find components of equal labels, 4- or 8-connectivity; out[p] =
 * largest linear index in p's component.  Returns the component count. */
int64_t oracle_cc_uf(const int32_t* labels, int W, int H, int conn,
                     int32_t* out);

/* Dense rank of each value among the sorted distinct values (in place);
 * returns the number of distinct values. */
int64_t oracle_compress(int32_t* labels, int64_t n);

/* SplitMix64 sampler: k distinct indices in [0, n) (index = high 64 bits of
 * draw * n, redrawn on a repeat), as split_hip_uniform_random_initialize. */
int oracle_threads(void);
void oracle_sample_indices(uint64_t seed, int64_t n, int k, int64_t* idx);

/* ---- restatement of the reference (float, quirks included) ---- */

/* label.cu:23-105: float squared distance ((dx^2+dy^2)+dz^2) evaluated with
 * fused multiply-adds (fma(dz,dz,fma(dy,dy,dx*dx))), lexicographic
 * (distance, index) minimum. */
void ref_label_points(const float* cent, int k, const float* pts, int64_t n,
                      int32_t* labels);

/* centroids.cu:22-101: stable sort by label, float cumulative counts from
 * upper_bound, reciprocal of adjacent differences, reduce_by_key sums written
 * compacted by distinct label (sequential float sums in sorted order), then
 * sums x reciprocal count.  Reproduces the empty-cluster slot shift. */
void ref_calculate_centroids(const int32_t* labels, const float* pts,
                             int64_t n, float* cent, int k);

/* cluster.cu:27-93 including the aliasing of the saved centroids with the
 * distance matrix and the reciprocal counts, and the |delta - old_delta| >
 * threshold rule.  Returns the iteration count; *last_delta gets the final
 * delta. */
int ref_cluster(const float* pts, int64_t n, float* cent, int k,
                int32_t* labels, int max_iter, float threshold,
                float* last_delta);

/* connected_components.cu:80-181: max-propagation over horizontal then
 * vertical runs (4-connectivity) until the run maxima repeat or max_iter.
 * Returns the iteration count. */
int ref_connected_components(const int32_t* labels, int W, int H,
                             int32_t* out, int max_iter);

/* compress_labels.cu:10-35: sort, unique, lower_bound; in place. */
int64_t ref_compress_labels(int32_t* labels, int64_t n);

#ifdef __cplusplus
}
#endif

#endif
