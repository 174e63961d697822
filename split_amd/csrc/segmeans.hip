// Segment means: calculate_centroids for k beyond what the register-resident
// kernels hold (k = number of connected components, 10^4..10^7), and
// propagate_centroids.
//
// Sums are exact integers: each channel is quantised once with its own
// scale, q = rint(v * s_c) (s_c from the channel's largest |v|, as in the
// tiled Lloyd path), a wave of 256 consecutive pixels reduces its pixels one
// distinct label at a time (labels of connected components come in runs, so
// a wave holds few), and one lane adds the run's sums and count to int64 /
// uint32 global totals with atomics.  Integer atomics make the result
// independent of order; the quantisation moves a mean by at most
// max|v_c| * 2^-23 in channel c.  A final pass divides (double) and writes the centroids;
// labels with no pixel get the zero centroid.
#include <algorithm>

#include "common.hpp"
#include "lloyd_sums.hpp"

namespace splithip {
namespace {

constexpr int kSmThreads = 256;

__global__ void sm_zero(long long* sums, unsigned* cnt, int64_t k,
                        unsigned* maxabs) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * k;
       i += (int64_t)gridDim.x * blockDim.x) {
    sums[i] = 0;
    if (i < k) cnt[i] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x < 3) maxabs[threadIdx.x] = 0u;
}

// Loads pixel indices base + 4*lane + i (vec) or base + lane + 64*i.
__device__ __forceinline__ void sm_load(const float* pts,
                                        const int32_t* labels, int64_t n,
                                        int64_t base, int lane, bool vec,
                                        int (&lab)[4], float (&x)[4],
                                        float (&y)[4], float (&z)[4]) {
  if (vec && base + 256 <= n) {
    const int64_t o = base + 4 * lane;
    const int4 L = *reinterpret_cast<const int4*>(labels + o);
    const float4 X = *reinterpret_cast<const float4*>(pts + o);
    const float4 Y = *reinterpret_cast<const float4*>(pts + n + o);
    const float4 Z = *reinterpret_cast<const float4*>(pts + 2 * n + o);
    lab[0] = L.x; lab[1] = L.y; lab[2] = L.z; lab[3] = L.w;
    x[0] = X.x; x[1] = X.y; x[2] = X.z; x[3] = X.w;
    y[0] = Y.x; y[1] = Y.y; y[2] = Y.z; y[3] = Y.w;
    z[0] = Z.x; z[1] = Z.y; z[2] = Z.z; z[3] = Z.w;
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t p = base + lane + 64 * i;
    const bool v = p < n;
    lab[i] = v ? labels[p] : -1;
    x[i] = v ? pts[p] : 0.f;
    y[i] = v ? pts[n + p] : 0.f;
    z[i] = v ? pts[2 * n + p] : 0.f;
  }
}

__global__ __launch_bounds__(kSmThreads) void sm_accumulate(
    const int32_t* __restrict__ labels, const float* __restrict__ pts,
    int64_t n, int64_t k, int vec, const unsigned* maxabs, long long* sums,
    unsigned* cnt) {
  const int lane = threadIdx.x & (kWave - 1);
  const QScale qs = qscale_of(maxabs);
  const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t s = w0; s * 256 < n; s += nw) {
    int lab[4];
    float x[4], y[4], z[4];
    sm_load(pts, labels, n, s * 256, lane, vec != 0, lab, x, y, z);
    unsigned rem = 0;
    int qx[4], qy[4], qz[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      rem |= (lab[i] >= 0 && (int64_t)lab[i] < k ? 1u : 0u) << i;
      qx[i] = quantise(x[i], qs.x);
      qy[i] = quantise(y[i], qs.y);
      qz[i] = quantise(z[i], qs.z);
    }
    for (;;) {  // one distinct label per round
      const uint64_t act = ballot(rem != 0);
      if (act == 0) break;
      const int first = __builtin_ctzll(act);
      const int mine = (rem & 1u) ? lab[0]
                       : (rem & 2u) ? lab[1]
                       : (rem & 4u) ? lab[2]
                       : (rem & 8u) ? lab[3] : -1;
      const int L = readlane_i(mine, first);
      int sx = 0, sy = 0, sz = 0, c = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bool m = ((rem >> i) & 1u) && lab[i] == L;
        sx += m ? qx[i] : 0;
        sy += m ? qy[i] : 0;
        sz += m ? qz[i] : 0;
        c += m ? 1 : 0;
        rem &= m ? ~(1u << i) : ~0u;
      }
      // |per-lane sum| <= 4 (2^23 - 1), so a wave sum fits int32
      const long long tx = wave_sum_i32(sx), ty = wave_sum_i32(sy),
                      tz = wave_sum_i32(sz);
      const unsigned tc = (unsigned)wave_sum_i32(c);
      if (lane == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&sums[L]),
                  (unsigned long long)tx);
        atomicAdd(reinterpret_cast<unsigned long long*>(&sums[k + L]),
                  (unsigned long long)ty);
        atomicAdd(reinterpret_cast<unsigned long long*>(&sums[2 * k + L]),
                  (unsigned long long)tz);
        atomicAdd(&cnt[L], tc);
      }
    }
  }
}

__global__ void sm_finalize(const long long* sums, const unsigned* cnt,
                            int64_t k, const unsigned* maxabs,
                            float* centroids) {
  const QScale qs = qscale_of(maxabs);
  const float sc[3] = {qs.x, qs.y, qs.z};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < k;
       j += (int64_t)gridDim.x * blockDim.x) {
    const unsigned c = cnt[j];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
      centroids[ch * k + j] = dequantise_mean(sums[ch * k + j], c, sc[ch]);
  }
}

__global__ void propagate_kernel(const int32_t* __restrict__ labels, int64_t n,
                                 const float* __restrict__ cent, int64_t k,
                                 float* points) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int l = labels[i];
    if (l < 0 || (int64_t)l >= k) continue;
    points[i] = cent[l];
    points[n + i] = cent[k + l];
    points[2 * n + i] = cent[2 * k + l];
  }
}

struct SmLayout {
  unsigned* maxabs;
  long long* sums;
  unsigned* cnt;
  size_t bytes;
};

SmLayout sm_layout(void* temp, int64_t k) {
  SmLayout L{};
  char* t = static_cast<char*>(temp);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    char* p = t ? t + o : nullptr;
    o += align_up(bytes, 256);
    return p;
  };
  L.maxabs = reinterpret_cast<unsigned*>(take(3 * sizeof(unsigned)));
  L.sums = reinterpret_cast<long long*>(take(3 * (size_t)k * 8));
  L.cnt = reinterpret_cast<unsigned*>(take((size_t)k * 4));
  L.bytes = o;
  return L;
}

int grid_of(int64_t items, int per_cu) {
  return (int)std::max<int64_t>(
      1, std::min<int64_t>((items + 255) / 256,
                           (int64_t)per_cu * device_cu_count()));
}

}  // namespace

size_t segment_means_workspace(int64_t k) { return sm_layout(nullptr, k).bytes; }

int segment_means(const int32_t* labels, const float* points, int64_t n,
                  float* centroids, int64_t k, void* temp, size_t temp_bytes,
                  hipStream_t stream) {
  const SmLayout L = sm_layout(temp, k);
  if (temp_bytes < L.bytes)
    return fail(SPLIT_HIP_ERR_TEMP,
                "temp holds %zu bytes, calculate_centroids (k = %lld) needs %zu",
                temp_bytes, (long long)k, L.bytes);
  hipLaunchKernelGGL(sm_zero, dim3(grid_of(3 * k, 8)), dim3(256), 0, stream,
                     L.sums, L.cnt, k, L.maxabs);
  SPLIT_HIP_TRY(launch_absmax(points, n, L.maxabs, stream));
  const int vec = (((uintptr_t)points & 15u) == 0) &&
                  (((uintptr_t)labels & 15u) == 0) && (n % 4 == 0);
  hipLaunchKernelGGL(sm_accumulate, dim3(grid_of(n / 4 + 1, 16)),
                     dim3(kSmThreads), 0, stream, labels, points, n, k, vec,
                     L.maxabs, L.sums, L.cnt);
  hipLaunchKernelGGL(sm_finalize, dim3(grid_of(k, 8)), dim3(256), 0, stream,
                     L.sums, L.cnt, k, L.maxabs, centroids);
  SPLIT_HIP_CHECK_LAUNCH("segment means");
  return SPLIT_HIP_OK;
}

}  // namespace splithip

using namespace splithip;

extern "C" int split_hip_propagate_centroids(const int32_t* labels, int64_t n,
                                             const float* centroids, int64_t k,
                                             float* points,
                                             split_hip_stream_t stream) {
  if (!labels || !centroids || !points || n < 1 || k < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "labels, centroids or points missing");
  hipLaunchKernelGGL(propagate_kernel, dim3(grid_of(n, 16)), dim3(256), 0,
                     stream, labels, n, centroids, k, points);
  SPLIT_HIP_CHECK_LAUNCH("propagate_centroids");
  return SPLIT_HIP_OK;
}
