// The reference sample's segmentation sequence (samples/src/main.cu:228-310)
// written against the drop-in headers (include/split/...) and linked with
// libsplit_hip.so, as a reference caller would use it.  The GPU tests run it
// on seeded images and check every output against the CPU oracle.
//
//   split_dropin_sequence IN W H K SEED MAX_ITER THRESHOLD CONN OUT
//
// IN: 3 x (W*H) float32 planes.  Sequence: uniform_random_initialize ->
// cluster_workspace + one temp block -> cluster -> connected_components ->
// compress_labels -> calculate_centroids (cluster labels) ->
// calculate_centroids (segment labels, K = number of segments) ->
// propagate_centroids, all through the one temp block; then, in the same
// temp, the components of a 1-pixel checkerboard (K = N segments at
// 4-connectivity) and their segment means.  Writes OUT.{init,centroids,
// labels,segments,cluster_means,segment_means,propagated,checker,
// checker_means}.bin and OUT.txt (iterations shift nsegments nchecker).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "split/device/ccl/compress_labels.cuh"
#include "split/device/ccl/connected_components.cuh"
#include "split/device/device_error.h"
#include "split/device/kmeans/kmeans.cuh"

namespace {

using split::real;
namespace dev = split::device;

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    std::exit(3);
  }
}

template <class T>
T* dalloc(std::size_t count) {
  void* p = nullptr;
  hip_ok(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
  return static_cast<T*>(p);
}

template <class T>
void save(const std::string& path, const T* d, std::size_t count) {
  std::vector<T> h(count);
  hip_ok(hipMemcpy(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost),
         "hipMemcpy D2H");
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(h.data(), sizeof(T), count, f) != count) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(4);
  }
  std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 10) {
    std::fprintf(stderr,
                 "usage: %s IN W H K SEED MAX_ITER THRESHOLD CONN OUT\n",
                 argv[0]);
    return 2;
  }
  const std::string in = argv[1], out = argv[9];
  const int width = std::atoi(argv[2]), height = std::atoi(argv[3]);
  const int nclusters = std::atoi(argv[4]);
  const std::uint64_t seed = std::strtoull(argv[5], nullptr, 0);
  const int max_iter = std::atoi(argv[6]);
  const real threshold = (real)std::atof(argv[7]);
  const int conn = std::atoi(argv[8]);
  const int npixels = width * height;

  std::vector<real> h_points((std::size_t)3 * npixels);
  FILE* f = std::fopen(in.c_str(), "rb");
  if (!f || std::fread(h_points.data(), sizeof(real), h_points.size(), f) !=
                h_points.size()) {
    std::fprintf(stderr, "cannot read %s\n", in.c_str());
    return 4;
  }
  std::fclose(f);

  try {
    real* d_points = dalloc<real>(h_points.size());
    hip_ok(hipMemcpy(d_points, h_points.data(), h_points.size() * sizeof(real),
                     hipMemcpyHostToDevice),
           "hipMemcpy H2D");
    const dev::const_view2d<real> points(3, npixels, d_points);

    // main.cu:228-232
    real* d_centroids = dalloc<real>((std::size_t)3 * nclusters);
    dev::view2d<real> centroids(3, nclusters, d_centroids);
    dev::kmeans::uniform_random_initialize(points, centroids, seed);
    save(out + ".init.bin", d_centroids, (std::size_t)3 * nclusters);

    // main.cu:234-241: labels, segments and one temp block
    int* d_labels = dalloc<int>(npixels);
    int* d_segments = dalloc<int>(npixels);
    const std::size_t temp_bytes =
        dev::kmeans::cluster_workspace(npixels, nclusters, 3);
    void* d_temp = dalloc<unsigned char>(temp_bytes);
    const dev::device_ptr<void> temp(d_temp);

    // main.cu:244-245
    const dev::kmeans::cluster_result r = dev::kmeans::cluster(
        points, centroids, dev::view1d<int>(d_labels, npixels), temp, max_iter,
        threshold);
    save(out + ".centroids.bin", d_centroids, (std::size_t)3 * nclusters);
    save(out + ".labels.bin", d_labels, npixels);

    // main.cu:248-253
    dev::ccl::connected_components(
        dev::const_view2d<int>(height, width, d_labels), temp,
        dev::view1d<int>(d_segments, npixels), 100, conn);
    const int nsegments = dev::ccl::compress_labels(
        dev::view1d<int>(d_segments, npixels), temp);
    save(out + ".segments.bin", d_segments, npixels);

    // main.cu:259-260: the clusters' means (here of the same points)
    real* d_cmeans = dalloc<real>((std::size_t)3 * nclusters);
    dev::kmeans::calculate_centroids(
        dev::const_view1d<int>(d_labels, npixels), points,
        dev::view2d<real>(3, nclusters, d_cmeans), temp);
    save(out + ".cluster_means.bin", d_cmeans, (std::size_t)3 * nclusters);

    // main.cu:287-297: segment means, then propagated to the pixels
    real* d_smeans = dalloc<real>((std::size_t)3 * nsegments);
    const dev::view2d<real> seg_centroids(3, nsegments, d_smeans);
    dev::kmeans::calculate_centroids(
        dev::const_view1d<int>(d_segments, npixels), points, seg_centroids,
        temp);
    save(out + ".segment_means.bin", d_smeans, (std::size_t)3 * nsegments);
    real* d_prop = dalloc<real>(h_points.size());
    hip_ok(hipMemcpy(d_prop, d_points, h_points.size() * sizeof(real),
                     hipMemcpyDeviceToDevice),
           "hipMemcpy D2D");
    dev::kmeans::propagate_centroids(
        dev::const_view1d<int>(d_segments, npixels), seg_centroids,
        dev::view2d<real>(3, npixels, d_prop));
    save(out + ".propagated.bin", d_prop, h_points.size());

    // K = N segments in the same temp: a 1-pixel checkerboard at
    // 4-connectivity (every pixel its own component), then its means
    std::vector<int> h_checker(npixels);
    for (int y = 0; y < height; ++y)
      for (int x = 0; x < width; ++x)
        h_checker[(std::size_t)y * width + x] = (x + y) & 1;
    hip_ok(hipMemcpy(d_labels, h_checker.data(), npixels * sizeof(int),
                     hipMemcpyHostToDevice),
           "hipMemcpy H2D");
    dev::ccl::connected_components(
        dev::const_view2d<int>(height, width, d_labels), temp,
        dev::view1d<int>(d_segments, npixels), 100, 4);
    const int nchecker = dev::ccl::compress_labels(
        dev::view1d<int>(d_segments, npixels), temp);
    save(out + ".checker.bin", d_segments, npixels);
    real* d_chk = dalloc<real>((std::size_t)3 * nchecker);
    dev::kmeans::calculate_centroids(
        dev::const_view1d<int>(d_segments, npixels), points,
        dev::view2d<real>(3, nchecker, d_chk), temp);
    save(out + ".checker_means.bin", d_chk, (std::size_t)3 * nchecker);

    FILE* t = std::fopen((out + ".txt").c_str(), "w");
    if (!t) return 4;
    std::fprintf(t, "%d %.9g %d %d %zu\n", r.iterations, (double)r.shift,
                 nsegments, nchecker, temp_bytes);
    std::fclose(t);
    for (void* p : {(void*)d_points, (void*)d_centroids, (void*)d_labels,
                    (void*)d_segments, d_temp, (void*)d_cmeans,
                    (void*)d_smeans, (void*)d_prop, (void*)d_chk})
      hip_ok(hipFree(p), "hipFree");
  } catch (const split::device_error& e) {
    std::fprintf(stderr, "split::device_error %d: %s\n", e.status(), e.what());
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
