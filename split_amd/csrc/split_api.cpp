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
#include "split/device/color/beta_feature.cuh"
#include "split/device/color/conversion.cuh"
#include "split/device/ccl/connected_components.cuh"
#include "split/device/device_error.h"
#include "split/device/kmeans/kmeans.cuh"
#include "split/device/morph/erode.cuh"
#include "split/host/image_transfer.h"
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

// Stream-ordered device scratch for the wrappers whose reference
// signatures take no temp argument; freed when the wrapper returns.
class Scratch {
 public:
  Scratch(std::size_t bytes, const char* what) {
    const hipError_t e = hipMallocAsync(&p_, bytes, nullptr);
    if (e != hipSuccess)
      throw split::device_error(SPLIT_HIP_ERR_HIP,
                                std::string(what) + ": " + hipGetErrorString(e));
  }
  ~Scratch() { (void)hipFreeAsync(p_, nullptr); }
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  void* get() const { return p_; }

 private:
  void* p_ = nullptr;
};

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
  // The contract of this entry point: do_temp holds
  // calculate_centroids_workspace(N, K) bytes (a block sized by
  // cluster_workspace(N, ...) holds it for every K <= N, which covers the
  // segment means of main.cu:293-307 in the clustering temp).
  check(split_hip_calculate_centroids(
            di_labels.data(), di_points.values.data(), n,
            do_centroids.values.data(), k, do_temp.get(),
            calculate_centroids_workspace((int)n, k), nullptr),
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
  // The reference's sample reuses this block for connected_components,
  // compress_labels and calculate_centroids with K = number of segments
  // (<= N) (main.cu:240-307): cover all of them.
  return std::max({ref,
                   split_hip_cluster_workspace(i_npoints, i_nclusters, i_dim),
                   split_hip_compress_labels_workspace(i_npoints, i_npoints),
                   calculate_centroids_workspace(i_npoints, i_npoints)});
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

namespace morph
{
namespace
{
void erode_call(view2d<int> labels, int n, bool columns_too, const char* what)
{
  const int32_t w = (int32_t)labels.num_cols, h = (int32_t)labels.num_rows;
  const std::size_t bytes = split_hip_erode_workspace(w, h);
  Scratch temp(std::max<std::size_t>(bytes, 1), what);
  check(columns_too
            ? split_hip_erode(labels.values.data(), w, h, n, temp.get(), bytes,
                              nullptr)
            : split_hip_erode_horizontal_sweep(labels.values.data(), w, h, n,
                                               temp.get(), bytes, nullptr),
        what);
  sync(what);
}
}  // namespace

// split/src/split/device/morph/erode.cu:57-139
void erode_horizontal_sweep(view2d<int> dio_labels, const int niterations)
{
  erode_call(dio_labels, niterations, false, "erode_horizontal_sweep");
}

// split/src/split/device/morph/erode.cu:141-159
void erode(view2d<int> dio_labels, const int niterations)
{
  erode_call(dio_labels, niterations, true, "erode");
}
}  // namespace morph

namespace color
{
// split/include/split/device/color/conversion.cuh:67-82
void convert_color_space(const_view2d<real> di_in, view2d<real> do_out,
                         int conversion)
{
  rows3(di_in.num_rows, "convert_color_space");
  rows3(do_out.num_rows, "convert_color_space");
  if (do_out.num_cols != di_in.num_cols)
    throw split::device_error(SPLIT_HIP_ERR_ARGUMENT,
                              "convert_color_space: in and out sizes differ");
  check(split_hip_convert_color_space(di_in.values.data(), do_out.values.data(),
                                      di_in.num_cols, conversion, nullptr),
        "convert_color_space");
  sync("convert_color_space");
}

// split/src/split/device/color/beta_feature.cu:48-78
void beta_feature(const_view2d<real> di_lab_points, view1d<real> do_beta,
                  const real mu)
{
  rows3(di_lab_points.num_rows, "beta_feature");
  const std::size_t bytes = split_hip_beta_feature_workspace();
  Scratch temp(bytes, "beta_feature");
  check(split_hip_beta_feature(di_lab_points.values.data(), di_lab_points.num_cols,
                               do_beta.data(), mu, temp.get(), bytes, nullptr),
        "beta_feature");
  sync("beta_feature");
}
}  // namespace color

SPLIT_DEVICE_NAMESPACE_END

SPLIT_HOST_NAMESPACE_BEGIN

// samples/src/main.cu:98-114
void make_device_image(const real* h_image, split::device::view2d<real> d_image)
{
  const std::int64_t n = d_image.num_cols;
  const int c = (int)d_image.num_rows;
  const std::size_t bytes = split_hip_interleaved_workspace(n, c);
  Scratch temp(std::max<std::size_t>(bytes, 1), "make_device_image");
  check(split_hip_upload_interleaved(h_image, n, c, d_image.values.data(), temp.get(),
                                     bytes, nullptr),
        "make_device_image");
}

// samples/src/main.cu:116-132
void make_host_image(split::device::view2d<real> d_image, real* h_image)
{
  const std::int64_t n = d_image.num_cols;
  const int c = (int)d_image.num_rows;
  const std::size_t bytes = split_hip_interleaved_workspace(n, c);
  Scratch temp(std::max<std::size_t>(bytes, 1), "make_host_image");
  check(split_hip_download_interleaved(d_image.values.data(), n, c, h_image,
                                       temp.get(), bytes, nullptr),
        "make_host_image");
}

SPLIT_HOST_NAMESPACE_END
