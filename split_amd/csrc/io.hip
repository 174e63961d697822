// Image ingest and egress around the segmentation path: the host image is
// pixel-interleaved (HWC, as stbi::loadf returns it,
// split/src/split/host/stbi/stbi_raii.cpp:115-124) and the device path
// works on SoA planes.  The reference's sample moves each channel with a
// cudaMemcpy2D of one 4-byte element per row (samples/src/main.cu:84-132),
// i.e. n tiny strided DMA rows per channel.  Here the interleaved image
// crosses PCIe as one contiguous copy into device temp and a kernel
// (de)interleaves it in HBM; the 8-bit egress (stbi::quantize,
// stbi_raii.cpp:135-143) is fused into the interleave so that 1 byte per
// channel crosses back instead of 4.
//
// Kernels take quads of 4 pixels per thread: the interleaved side is
// 4 * C consecutive values (float4 / uint32 accesses when aligned), the
// planar side one float4 per plane.
#include <algorithm>

#include "common.hpp"

namespace splithip {
namespace {

constexpr int kIoThreads = 256;

int io_grid(int64_t n) {
  const int64_t quads = (n + 3) / 4;
  const int64_t want = (quads + kIoThreads - 1) / kIoThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(want, 8 * device_cu_count()));
}

// Aligned: both sides allow 16-byte access of whole quads.
template <int C>
__global__ __launch_bounds__(kIoThreads) void deinterleave_kernel(
    const float* hwc, float* planes, int64_t n, int aligned) {
  const int64_t quads = (n + 3) / 4;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < quads;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = 4 * q;
    if (aligned) {
      float v[4 * C];
#pragma unroll
      for (int j = 0; j < C; ++j) {
        const float4 t = reinterpret_cast<const float4*>(hwc + i0 * C)[j];
        v[4 * j] = t.x; v[4 * j + 1] = t.y; v[4 * j + 2] = t.z; v[4 * j + 3] = t.w;
      }
#pragma unroll
      for (int c = 0; c < C; ++c)
        *reinterpret_cast<float4*>(planes + c * n + i0) =
            make_float4(v[c], v[C + c], v[2 * C + c], v[3 * C + c]);
    } else {
      for (int i = 0; i < 4 && i0 + i < n; ++i)
#pragma unroll
        for (int c = 0; c < C; ++c) planes[c * n + i0 + i] = hwc[(i0 + i) * C + c];
    }
  }
}

template <int C>
__global__ __launch_bounds__(kIoThreads) void interleave_kernel(
    const float* planes, float* hwc, int64_t n, int aligned) {
  const int64_t quads = (n + 3) / 4;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < quads;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = 4 * q;
    if (aligned) {
      float v[4 * C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const float4 t = *reinterpret_cast<const float4*>(planes + c * n + i0);
        v[c] = t.x; v[C + c] = t.y; v[2 * C + c] = t.z; v[3 * C + c] = t.w;
      }
#pragma unroll
      for (int j = 0; j < C; ++j)
        reinterpret_cast<float4*>(hwc + i0 * C)[j] =
            make_float4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
    } else {
      for (int i = 0; i < 4 && i0 + i < n; ++i)
#pragma unroll
        for (int c = 0; c < C; ++c) hwc[(i0 + i) * C + c] = planes[c * n + i0 + i];
    }
  }
}

// clamp(v, 0, 1) * 255 + 0.5 truncated to 8 bits, as one fused multiply-add
// (the reference's host expression contracted under its -Ofast -mfma build).
__device__ __forceinline__ uint32_t quantize_u8(float v) {
  const float c = fminf(fmaxf(v, 0.f), 1.f);
  return (uint32_t)fmaf(c, 255.f, 0.5f);
}

template <int C>
__global__ __launch_bounds__(kIoThreads) void quantize_kernel(
    const float* planes, uint8_t* out, int64_t n, int aligned) {
  const int64_t quads = (n + 3) / 4;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < quads;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = 4 * q;
    if (aligned) {
      uint32_t b[4 * C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const float4 t = *reinterpret_cast<const float4*>(planes + c * n + i0);
        b[c] = quantize_u8(t.x);
        b[C + c] = quantize_u8(t.y);
        b[2 * C + c] = quantize_u8(t.z);
        b[3 * C + c] = quantize_u8(t.w);
      }
      // 4C bytes = C little-endian words
#pragma unroll
      for (int j = 0; j < C; ++j)
        reinterpret_cast<uint32_t*>(out + i0 * C)[j] =
            b[4 * j] | (b[4 * j + 1] << 8) | (b[4 * j + 2] << 16) | (b[4 * j + 3] << 24);
    } else {
      for (int i = 0; i < 4 && i0 + i < n; ++i)
#pragma unroll
        for (int c = 0; c < C; ++c)
          out[(i0 + i) * C + c] = (uint8_t)quantize_u8(planes[c * n + i0 + i]);
    }
  }
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

int check_args(const void* a, const void* b, int64_t n, int32_t channels,
               const char* what) {
  if (!a || !b || n < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "%s: buffer missing or n < 1", what);
  if (channels < 1 || channels > 4)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "%s: channels must be 1..4, got %d", what,
                channels);
  return SPLIT_HIP_OK;
}

int launch_deinterleave(const float* hwc, float* planes, int64_t n, int channels,
                        hipStream_t stream) {
  const int al = aligned16(hwc) && aligned16(planes) && n % 4 == 0;
  const dim3 g(io_grid(n)), b(kIoThreads);
  switch (channels) {
    case 1: hipLaunchKernelGGL(deinterleave_kernel<1>, g, b, 0, stream, hwc, planes, n, al); break;
    case 2: hipLaunchKernelGGL(deinterleave_kernel<2>, g, b, 0, stream, hwc, planes, n, al); break;
    case 3: hipLaunchKernelGGL(deinterleave_kernel<3>, g, b, 0, stream, hwc, planes, n, al); break;
    default: hipLaunchKernelGGL(deinterleave_kernel<4>, g, b, 0, stream, hwc, planes, n, al); break;
  }
  SPLIT_HIP_CHECK_LAUNCH("deinterleave");
  return SPLIT_HIP_OK;
}

int launch_interleave(const float* planes, float* hwc, int64_t n, int channels,
                      hipStream_t stream) {
  const int al = aligned16(hwc) && aligned16(planes) && n % 4 == 0;
  const dim3 g(io_grid(n)), b(kIoThreads);
  switch (channels) {
    case 1: hipLaunchKernelGGL(interleave_kernel<1>, g, b, 0, stream, planes, hwc, n, al); break;
    case 2: hipLaunchKernelGGL(interleave_kernel<2>, g, b, 0, stream, planes, hwc, n, al); break;
    case 3: hipLaunchKernelGGL(interleave_kernel<3>, g, b, 0, stream, planes, hwc, n, al); break;
    default: hipLaunchKernelGGL(interleave_kernel<4>, g, b, 0, stream, planes, hwc, n, al); break;
  }
  SPLIT_HIP_CHECK_LAUNCH("interleave");
  return SPLIT_HIP_OK;
}

}  // namespace
}  // namespace splithip

using namespace splithip;

extern "C" {

size_t split_hip_interleaved_workspace(int64_t n, int32_t channels) {
  if (n < 1 || channels < 1 || channels > 4) return 0;
  return align_up((size_t)n * (size_t)channels * sizeof(float), 256);
}

int split_hip_deinterleave(const float* hwc, int64_t n, int32_t channels,
                           float* planes, split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_args(hwc, planes, n, channels, "deinterleave"));
  return launch_deinterleave(hwc, planes, n, channels, stream);
}

int split_hip_interleave(const float* planes, int64_t n, int32_t channels,
                         float* hwc, split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_args(planes, hwc, n, channels, "interleave"));
  return launch_interleave(planes, hwc, n, channels, stream);
}

int split_hip_upload_interleaved(const float* host, int64_t n, int32_t channels,
                                 float* planes, void* temp, size_t temp_bytes,
                                 split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_args(host, planes, n, channels, "upload_interleaved"));
  const size_t need = split_hip_interleaved_workspace(n, channels);
  if (!temp || temp_bytes < need)
    return fail(SPLIT_HIP_ERR_TEMP, "upload_interleaved: temp holds %zu bytes, needs %zu",
                temp_bytes, need);
  float* stage = static_cast<float*>(temp);
  SPLIT_HIP_CHECK(hipMemcpyAsync(stage, host, (size_t)n * channels * sizeof(float),
                                 hipMemcpyHostToDevice, stream));
  SPLIT_HIP_TRY(launch_deinterleave(stage, planes, n, channels, stream));
  SPLIT_HIP_CHECK(hipStreamSynchronize(stream));  // the host buffer is free again
  return SPLIT_HIP_OK;
}

int split_hip_download_interleaved(const float* planes, int64_t n,
                                   int32_t channels, float* host, void* temp,
                                   size_t temp_bytes, split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_args(planes, host, n, channels, "download_interleaved"));
  const size_t need = split_hip_interleaved_workspace(n, channels);
  if (!temp || temp_bytes < need)
    return fail(SPLIT_HIP_ERR_TEMP, "download_interleaved: temp holds %zu bytes, needs %zu",
                temp_bytes, need);
  float* stage = static_cast<float*>(temp);
  SPLIT_HIP_TRY(launch_interleave(planes, stage, n, channels, stream));
  SPLIT_HIP_CHECK(hipMemcpyAsync(host, stage, (size_t)n * channels * sizeof(float),
                                 hipMemcpyDeviceToHost, stream));
  SPLIT_HIP_CHECK(hipStreamSynchronize(stream));  // the host buffer is complete
  return SPLIT_HIP_OK;
}

int split_hip_quantize_interleaved(const float* planes, int64_t n,
                                   int32_t channels, uint8_t* out,
                                   split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_args(planes, out, n, channels, "quantize_interleaved"));
  const int al = aligned16(planes) && ((uintptr_t)out & 3u) == 0 && n % 4 == 0;
  const dim3 g(io_grid(n)), b(kIoThreads);
  switch (channels) {
    case 1: hipLaunchKernelGGL(quantize_kernel<1>, g, b, 0, stream, planes, out, n, al); break;
    case 2: hipLaunchKernelGGL(quantize_kernel<2>, g, b, 0, stream, planes, out, n, al); break;
    case 3: hipLaunchKernelGGL(quantize_kernel<3>, g, b, 0, stream, planes, out, n, al); break;
    default: hipLaunchKernelGGL(quantize_kernel<4>, g, b, 0, stream, planes, out, n, al); break;
  }
  SPLIT_HIP_CHECK_LAUNCH("quantize_interleaved");
  return SPLIT_HIP_OK;
}

}  // extern "C"
