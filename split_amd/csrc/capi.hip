// C-ABI plumbing (status text, device queries) and the whole segmentation
// path: cluster -> connected components -> compress, stream-ordered, one
// host synchronisation at the end (the caller sequence of
// samples/src/main.cu:240-252).
#include <stdarg.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace splithip {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int fail(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return status;
}

int device_cu_count() {
  static std::mutex mu;
  static std::vector<int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
  if (cache[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    cache[dev] = cus;
  }
  return cache[dev];
}

// Defined in kmeans.hip / ccl.hip.
int cluster_async(const float* pts, int64_t n, int width, int height,
                  float* centroids, int k, int32_t* labels, void* temp,
                  size_t temp_bytes, int max_iter, float threshold,
                  hipStream_t stream, Mailbox** mbox_out);
struct CompressState;
int segments_async(const int32_t* labels, int W, int H, int32_t* segments,
                   int conn, bool compress, void* temp, size_t temp_bytes,
                   hipStream_t stream, CompressState** st_out);
int compress_labels_finish(CompressState* st, int32_t* num_unique,
                           hipStream_t stream);
size_t segments_workspace(int W, int H, bool compress);

}  // namespace splithip

using namespace splithip;

extern "C" {

const char* split_hip_last_error(void) { return g_last_error; }

const char* split_hip_status_string(int status) {
  switch (status) {
    case SPLIT_HIP_OK: return "ok";
    case SPLIT_HIP_ERR_ARGUMENT: return "invalid argument";
    case SPLIT_HIP_ERR_TEMP: return "temp buffer too small";
    case SPLIT_HIP_ERR_HIP: return "HIP runtime error";
    case SPLIT_HIP_ERR_RANGE: return "label value range too large for temp";
    case SPLIT_HIP_ERR_UNSUPPORTED: return "size beyond a documented limit";
    default: return "unknown status";
  }
}

const char* split_hip_version(void) { return "0.1.0"; }

size_t split_hip_segment_workspace(int32_t width, int32_t height, int32_t k) {
  const int64_t n = (int64_t)width * height;
  return std::max(split_hip_cluster_workspace(n, k, 3),
                  segments_workspace(width, height, true));
}

int split_hip_segment(const float* points, int32_t width, int32_t height,
                      float* centroids, int32_t k, int32_t* labels,
                      int32_t* segments, void* temp, size_t temp_bytes,
                      int32_t max_iter, float threshold, int32_t connectivity,
                      int32_t compress, split_hip_cluster_result* result,
                      int32_t* num_segments, split_hip_stream_t stream) {
  if (width < 1 || height < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "image is %d x %d", width, height);
  if (!labels || !segments)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "labels or segments is null");
  const int64_t n = (int64_t)width * height;
  const size_t need = split_hip_segment_workspace(width, height, k);
  if (temp_bytes < need)
    return fail(SPLIT_HIP_ERR_TEMP, "temp holds %zu bytes, segment needs %zu",
                temp_bytes, need);
  Mailbox* mb = nullptr;
  SPLIT_HIP_TRY(cluster_async(points, n, width, height, centroids, k, labels,
                              temp, temp_bytes, max_iter, threshold, stream,
                              &mb));
  CompressState* st = nullptr;
  SPLIT_HIP_TRY(segments_async(labels, width, height, segments, connectivity,
                               compress != 0, temp, temp_bytes, stream, &st));
  if (compress) {
    SPLIT_HIP_TRY(compress_labels_finish(st, num_segments, stream));
  } else {
    SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
    if (num_segments) *num_segments = -1;
  }
  if (result) {
    const volatile Mailbox* v = mb;
    result->iterations = v->iter;
    result->shift = v->shift;
  }
  return SPLIT_HIP_OK;
}

}  // extern "C"
