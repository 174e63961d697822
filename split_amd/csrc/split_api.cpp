// C++ drop-in entry points (include/split/device/...) over the C ABI.
//
// Each function keeps the reference's namespace, name and argument order and
// forwards to the matching split_hip_* call on the default stream; a nonzero
// status becomes a split::device_error.  Like the reference, which reads a
// scalar back after every call, each wrapper returns with its work finished.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "split/device/ccl/compress_labels.cuh"
#include "split/device/ccl/connected_components.cuh"
#include "split/device/device_error.h"
#include "split/device/kmeans/kmeans.cuh"
#include "split_hip.h"

namespace {

void check(int status, const char* what) {
  if (status != SPLIT_HIP_OK) {
    throw split::device_error(status, std::string(what) + ": " +
                                          split_hip_status_string(status) +
                                          ": " + split_hip_last_error());
  }
}

void sync(const char* what) {
  const hipError_t e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess)
    throw split::device_error(SPLIT_HIP_ERR_HIP,
                              std::string(what) + ": " + hipGetErrorString(e));
}

int rows3(std::int64_t rows, const char* what) {
  if (rows != 3)
    throw split::device_error(SPLIT_HIP_ERR_ARGUMENT,
                              std::string(what) + ": points must have 3 rows");
  return 3;
}

}  // namespace

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace kmeans
{
// split/src/split/device/kmeans/label.cu:68-72
std::size_t label_points_workspace(const int i_npoints, const int i_nclusters)
{
  return std::max<std::size_t>(
      (std::size_t)i_npoints * (std::size_t)i_nclusters * sizeof(real),
      split_hip_label_points_workspace(i_npoints, i_nclusters));
}

// split/src/split/device/kmeans/label.cu:74-106
void label_points(const_view2d<real> di_centroids,
                  const_view2d<real> di_points,
                  view1d<int> do_cluster_labels,
                  device_ptr<void> do_temp)
{
  rows3(di_points.num_rows, "label_points");
  rows3(di_centroids.num_rows, "label_points");
  check(split_hip_label_points(di_centroids.values.data(),
                               (int32_t)di_centroids.num_cols,
                               di_points.values.data(), di_points.num_cols,
                               do_cluster_labels.data(), do_temp.get(), 0,
                               nullptr),
        "label_points");
  sync("label_points");
}

// split/src/split/device/kmeans/centroids.cu:16-20
std::size_t calculate_centroids_workspace(const int i_npoints,
                                          const int i_ncentroids)
{
  return std::max<std::size_t>(
      (std::size_t)i_npoints * 2 * sizeof(int) +
          (std::size_t)i_ncentroids * sizeof(real),
      split_hip_calculate_centroids_workspace(i_npoints, i_ncentroids));
}

// split/src/split/device/kmeans/centroids.cu:22-101
void calculate_centroids(const_view1d<int> di_labels,
                         const_view2d<real> di_points,
                         view2d<real> do_centroids,
                         device_ptr<void> do_temp)
{
  rows3(di_points.num_rows, "calculate_centroids");
  rows3(do_centroids.num_rows, "calculate_centroids");
  const int k = (int)do_centroids.num_cols;
  const std::int64_t n = di_points.num_cols;
  // A reference caller sized temp by the reference's formula (at least
  // N*2*sizeof(int) + K*sizeof(real)); the grid adapts to that size.
  check(split_hip_calculate_centroids(
            di_labels.data(), di_points.values.data(), n,
            do_centroids.values.data(), k, do_temp.get(),
            (std::size_t)n * 2 * sizeof(int) + (std::size_t)k * sizeof(real),
            nullptr),
        "calculate_centroids");
  sync("calculate_centroids");
}

// split/src/split/device/kmeans/propagation.cu:9-24
void propagate_centroids(const_view1d<int> di_cluster_labels,
                         const_view2d<real> di_centroids,
                         view2d<real> do_points)
{
  rows3(di_centroids.num_rows, "propagate_centroids");
  rows3(do_points.num_rows, "propagate_centroids");
  check(split_hip_propagate_centroids(
            di_cluster_labels.data(), (std::int64_t)do_points.num_cols,
            di_centroids.values.data(), (std::int64_t)di_centroids.num_cols,
            do_points.values.data(), nullptr),
        "propagate_centroids");
  sync("propagate_centroids");
}

// split/src/split/device/kmeans/cluster.cu:12-25
std::size_t cluster_workspace(const int i_npoints, const int i_nclusters,
                              const int i_dim)
{
  const std::size_t ref = std::max<std::size_t>(
      (std::size_t)i_nclusters * i_dim * sizeof(real) +
          calculate_centroids_workspace(i_npoints, i_nclusters),
      label_points_workspace(i_npoints, i_nclusters));
  return std::max({ref,
                   split_hip_cluster_workspace(i_npoints, i_nclusters, i_dim),
                   split_hip_compress_labels_workspace(i_npoints, i_npoints)});
}

// split/src/split/device/kmeans/cluster.cu:27-93
cluster_result cluster(const_view2d<real> di_points,
                       view2d<real> dio_centroids,
                       view1d<int> do_cluster_labels,
                       device_ptr<void> do_temp,
                       int i_max_iter,
                       real i_threshold)
{
  rows3(di_points.num_rows, "cluster");
  rows3(dio_centroids.num_rows, "cluster");
  const int k = (int)dio_centroids.num_cols;
  const std::int64_t n = di_points.num_cols;
  split_hip_cluster_result r{0, 0.f};
  check(split_hip_cluster(di_points.values.data(), n,
                          dio_centroids.values.data(), k,
                          do_cluster_labels.data(), do_temp.get(),
                          cluster_workspace((int)n, k, 3), i_max_iter,
                          i_threshold, &r, nullptr),
        "cluster");
  return cluster_result{r.iterations, r.shift};
}

// split/src/split/device/kmeans/uniform_random_initialize.cu:10-55
void uniform_random_initialize(const_view2d<real> di_points,
                               view2d<real> do_centroids, std::uint64_t seed)
{
  rows3(di_points.num_rows, "uniform_random_initialize");
  rows3(do_centroids.num_rows, "uniform_random_initialize");
  check(split_hip_uniform_random_initialize(
            di_points.values.data(), di_points.num_cols,
            do_centroids.values.data(), (int)do_centroids.num_cols, seed,
            nullptr),
        "uniform_random_initialize");
  sync("uniform_random_initialize");
}

void uniform_random_initialize(const_view2d<real> di_points,
                               view2d<real> do_centroids)
{
  uniform_random_initialize(di_points, do_centroids, 0x5EEDull);
}

}  // namespace kmeans

namespace ccl
{
// split/src/split/device/ccl/connected_components.cu:80-181
void connected_components(const_view2d<int> di_labels,
                          device_ptr<void> dio_temp,
                          view1d<int> do_labels,
                          int i_max_iterations,
                          int i_connectivity)
{
  check(split_hip_connected_components(
            di_labels.values.data(), (int32_t)di_labels.num_cols,
            (int32_t)di_labels.num_rows, dio_temp.get(),
            split_hip_connected_components_workspace(
                (int32_t)di_labels.num_cols, (int32_t)di_labels.num_rows),
            do_labels.data(),
            i_max_iterations, i_connectivity, nullptr),
        "connected_components");
  sync("connected_components");
}

std::size_t compress_labels_workspace(const int i_nlabels,
                                      const long long value_span)
{
  return split_hip_compress_labels_workspace(i_nlabels, value_span);
}

// split/src/split/device/ccl/compress_labels.cu:10-35.  The reference's
// contract is a temp block of at least N ints; that is what is assumed here.
int compress_labels(view1d<int> dio_labels, device_ptr<void> dio_temp)
{
  const std::int64_t n = dio_labels.size();
  int32_t count = 0;
  const std::size_t temp_bytes = (std::size_t)n * sizeof(int);
  check(split_hip_compress_labels(dio_labels.data(), n, dio_temp.get(),
                                  temp_bytes, &count, nullptr),
        "compress_labels");
  return count;
}

}  // namespace ccl

SPLIT_DEVICE_NAMESPACE_END
