// Colour-space conversion and the beta feature: the per-pixel transforms
// that turn the linear RGB image into the features k-means clusters
// (the caller sequence of samples/src/main.cu:216-225).
//
// Conversions follow the sRGB / CIE definitions with a D65 white point
// (split/include/split/device/color/detail/conversion.inl:13-139): sRGB
// companding, the sRGB <-> XYZ matrices, the CIE Lab companding with the
// 0.008856 / 7.787 approximations, and intensity-chroma.  Every conversion
// is a pointwise map over three SoA float planes, so a kernel streams
// float4 quads of the three planes in and out (in place allowed).
//
// Beta feature (split/src/split/device/color/beta_feature.cu:48-78):
// beta = -mu where |a| and |b| are below lambda and L < 0.2 Lmax, +mu where
// they are below lambda and L > 0.95 Lmax, 0 elsewhere; Lmax = max L and
// lambda = 0.2 * max |a| (the reference takes lambda from the a plane only).
// One pass reduces both maxima (order-preserving integer keys, wave DPP and
// block reductions, atomics spread over 32 copies), a second writes beta.
//
// split_hip_lab_beta_feature fuses the sample's sequence rgb -> Lab, copy
// of L, beta in place of L: the maxima are reduced while Lab is written, so
// the image is read once and the Lab planes once more.
#include <algorithm>

#include "common.hpp"

namespace splithip {
namespace {

constexpr int kColorThreads = 256;

struct C3 {
  float x, y, z;
};

// Powers of positive normal floats as 2^(y log2 x) on the hardware
// transcendental units (v_log_f32, v_exp_f32): the form a fast-math build of
// the reference evaluates pow in.  Its relative error grows with
// |y log2 x| (a few float ulp over the arguments used here, far below the
// conversions' output tolerance).  The branches below only raise normal
// floats >= 0.0031308 to a power.
__device__ __forceinline__ float pow_pos(float x, float y) {
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
// Cube root of a positive normal float: 2^(log2(t) / 3) refined by one
// Newton step on y^3 = t.
__device__ __forceinline__ float cbrt_pos(float t) {
  const float y = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(t) * (1.f / 3.f));
  return y - (y * y * y - t) * __builtin_amdgcn_rcpf(3.f * y * y);
}

// sRGB companded value -> linear, and back.
__device__ __forceinline__ float srgb_to_linear(float v) {
  return v > 0.04045f ? pow_pos((v + 0.055f) * (1.f / 1.055f), 2.4f)
                      : v * (1.f / 12.92f);
}
__device__ __forceinline__ float linear_to_srgb(float v) {
  return v > 0.0031308f ? fmaf(1.055f, pow_pos(v, 1.f / 2.4f), -0.055f)
                        : 12.92f * v;
}
// CIE Lab companding f(t) and its inverse.
__device__ __forceinline__ float lab_f(float t) {
  return t > 0.008856f ? cbrt_pos(t) : fmaf(7.787f, t, 16.f / 116.f);
}
__device__ __forceinline__ float lab_finv(float t) {
  const float t3 = t * t * t;
  return t3 > 0.008856f ? t3 : (t - 16.f / 116.f) * (1.f / 7.787f);
}

__device__ __forceinline__ C3 rgb_to_xyz(C3 c) {
  const float r = 100.f * srgb_to_linear(c.x);
  const float g = 100.f * srgb_to_linear(c.y);
  const float b = 100.f * srgb_to_linear(c.z);
  return {fmaf(b, 0.1804375f, fmaf(g, 0.3575761f, r * 0.4124564f)),
          fmaf(b, 0.0721750f, fmaf(g, 0.7151522f, r * 0.2126729f)),
          fmaf(b, 0.9503041f, fmaf(g, 0.1191920f, r * 0.0193339f))};
}
__device__ __forceinline__ C3 xyz_to_rgb(C3 c) {
  const float x = c.x * 0.01f, y = c.y * 0.01f, z = c.z * 0.01f;
  const float r = fmaf(z, -0.4985314f, fmaf(y, -1.5371385f, x * 3.2404542f));
  const float g = fmaf(z, 0.0415560f, fmaf(y, 1.8760108f, x * -0.9692660f));
  const float b = fmaf(z, 1.0572252f, fmaf(y, -0.2040259f, x * 0.0556434f));
  return {linear_to_srgb(r), linear_to_srgb(g), linear_to_srgb(b)};
}
__device__ __forceinline__ C3 xyz_to_lab(C3 c) {
  const float fx = lab_f(c.x * (1.f / 95.047f));
  const float fy = lab_f(c.y * 0.01f);
  const float fz = lab_f(c.z * (1.f / 108.883f));
  return {fmaf(116.f, fy, -16.f), 500.f * (fx - fy), 200.f * (fy - fz)};
}
__device__ __forceinline__ C3 lab_to_xyz(C3 c) {
  const float fy = (c.x + 16.f) * (1.f / 116.f);
  const float fx = fmaf(c.y, 1.f / 500.f, fy);
  const float fz = fmaf(c.z, -1.f / 200.f, fy);
  return {lab_finv(fx) * 95.047f, lab_finv(fy) * 100.f,
          lab_finv(fz) * 108.883f};
}
__device__ __forceinline__ C3 rgb_to_ic(C3 c) {
  // black pixels give inf/NaN chroma, as in the reference
  const float i = (c.x + c.y + c.z) * (1.f / 3.f);
  const float ri = 1.f / i;
  return {i, c.x * ri, c.y * ri};
}

template <int CONV>
__device__ __forceinline__ C3 convert(C3 c) {
  if constexpr (CONV == SPLIT_HIP_RGB_TO_XYZ) return rgb_to_xyz(c);
  else if constexpr (CONV == SPLIT_HIP_XYZ_TO_RGB) return xyz_to_rgb(c);
  else if constexpr (CONV == SPLIT_HIP_XYZ_TO_LAB) return xyz_to_lab(c);
  else if constexpr (CONV == SPLIT_HIP_LAB_TO_XYZ) return lab_to_xyz(c);
  else if constexpr (CONV == SPLIT_HIP_RGB_TO_LAB) return xyz_to_lab(rgb_to_xyz(c));
  else if constexpr (CONV == SPLIT_HIP_LAB_TO_RGB) return xyz_to_rgb(lab_to_xyz(c));
  else return rgb_to_ic(c);
}

// Order-preserving key of a float (larger float -> larger unsigned key);
// key 0 is below every float's key.
__device__ __forceinline__ unsigned float_key(float f) {
  const unsigned b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float key_float(unsigned k) {
  const unsigned b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  float f;
  memcpy(&f, &b, 4);
  return f;
}

// Quads of four consecutive pixels: quad q covers pixels 4q..4q+3.  `vec`
// (planes 16-byte aligned, n a multiple of 4) selects float4 access.
struct Quad {
  float v[3][4];
};
__device__ __forceinline__ void load_quad(const float* p, int64_t n,
                                          int64_t q, bool vec, Quad& d,
                                          bool (&ok)[4]) {
  const int64_t i0 = 4 * q;
  if (vec) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float4 t = *reinterpret_cast<const float4*>(p + c * n + i0);
      d.v[c][0] = t.x; d.v[c][1] = t.y; d.v[c][2] = t.z; d.v[c][3] = t.w;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) ok[i] = true;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ok[i] = i0 + i < n;
#pragma unroll
      for (int c = 0; c < 3; ++c) d.v[c][i] = ok[i] ? p[c * n + i0 + i] : 0.f;
    }
  }
}
__device__ __forceinline__ void store_plane(float* p, int64_t i0, bool vec,
                                            const float (&v)[4],
                                            const bool (&ok)[4]) {
  if (vec) {
    *reinterpret_cast<float4*>(p + i0) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (ok[i]) p[i0 + i] = v[i];
  }
}

template <int CONV>
__global__ __launch_bounds__(kColorThreads) void convert_kernel(
    const float* in, float* out, int64_t n, int vec) {
  const int64_t quads = (n + 3) / 4;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < quads;
       q += (int64_t)gridDim.x * blockDim.x) {
    Quad d;
    bool ok[4];
    load_quad(in, n, q, vec, d, ok);
    Quad r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const C3 o = convert<CONV>(C3{d.v[0][i], d.v[1][i], d.v[2][i]});
      r.v[0][i] = o.x;
      r.v[1][i] = o.y;
      r.v[2][i] = o.z;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) store_plane(out + c * n, 4 * q, vec, r.v[c], ok);
  }
}

// The two maxima are kept in kKeyCopies copies (keys[c * kKeyStride + 0]:
// L, + 1: |a|), one 128-byte line each; block b adds into copy
// b % kKeyCopies, so that few blocks contend for one address.
constexpr int kKeyCopies = 32;
constexpr int kKeyStride = 32;  // uints

__global__ void beta_init(unsigned* keys) {
  for (int i = threadIdx.x; i < kKeyCopies * kKeyStride; i += blockDim.x)
    keys[i] = 0u;
}

// Block maximum of L and |a| (wave DPP max, then across the waves in LDS),
// added by one thread into the block's copy of the keys.
__device__ __forceinline__ void reduce_maxima(float lmax, float amax,
                                              unsigned* keys) {
  __shared__ float part[2][kColorThreads / kWave];
  lmax = wave_allreduce(lmax, OpMax());
  amax = wave_allreduce(amax, OpMax());
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    part[0][wave] = lmax;
    part[1][wave] = amax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)blockDim.x / kWave; ++w) {
      lmax = fmaxf(lmax, part[0][w]);
      amax = fmaxf(amax, part[1][w]);
    }
    unsigned* k = keys + (blockIdx.x % kKeyCopies) * kKeyStride;
    atomicMax(k, float_key(lmax));
    atomicMax(k + 1, float_key(amax));
  }
}

__global__ __launch_bounds__(kColorThreads) void beta_maxima(
    const float* lab, int64_t n, int vec, unsigned* keys) {
  const int64_t quads = (n + 3) / 4;
  float lmax = -INFINITY, amax = -INFINITY;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < quads;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = 4 * q;
    if (vec) {  // the L and a planes only
      const float4 L = *reinterpret_cast<const float4*>(lab + i0);
      const float4 A = *reinterpret_cast<const float4*>(lab + n + i0);
      lmax = fmaxf(lmax, fmaxf(fmaxf(L.x, L.y), fmaxf(L.z, L.w)));
      amax = fmaxf(amax, fmaxf(fmaxf(fabsf(A.x), fabsf(A.y)),
                               fmaxf(fabsf(A.z), fabsf(A.w))));
    } else {
      for (int i = 0; i < 4 && i0 + i < n; ++i) {
        lmax = fmaxf(lmax, lab[i0 + i]);
        amax = fmaxf(amax, fabsf(lab[n + i0 + i]));
      }
    }
  }
  reduce_maxima(lmax, amax, keys);
}

// rgb -> Lab into out (L, a, b planes), L also into lum (may be null), and
// the maxima of L and |a|.
__global__ __launch_bounds__(kColorThreads) void lab_maxima(
    const float* rgb, float* out, float* lum, int64_t n, int vec,
    unsigned* keys) {
  const int64_t quads = (n + 3) / 4;
  float lmax = -INFINITY, amax = -INFINITY;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < quads;
       q += (int64_t)gridDim.x * blockDim.x) {
    Quad d;
    bool ok[4];
    load_quad(rgb, n, q, vec, d, ok);
    Quad r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const C3 o = xyz_to_lab(rgb_to_xyz(C3{d.v[0][i], d.v[1][i], d.v[2][i]}));
      r.v[0][i] = o.x;
      r.v[1][i] = o.y;
      r.v[2][i] = o.z;
      lmax = ok[i] ? fmaxf(lmax, o.x) : lmax;
      amax = ok[i] ? fmaxf(amax, fabsf(o.y)) : amax;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) store_plane(out + c * n, 4 * q, vec, r.v[c], ok);
    if (lum) store_plane(lum, 4 * q, vec, r.v[0], ok);
  }
  reduce_maxima(lmax, amax, keys);
}

// beta[i] from (L, a, b) of pixel i; beta may alias the L plane.
__global__ __launch_bounds__(kColorThreads) void beta_kernel(
    const float* lab, int64_t n, int vec, const unsigned* keys, float mu,
    float* beta) {
  __shared__ unsigned kmax[2];
  if (threadIdx.x < 2) kmax[threadIdx.x] = 0u;
  __syncthreads();
  if (threadIdx.x < 2 * kKeyCopies)
    atomicMax(&kmax[threadIdx.x & 1], keys[(threadIdx.x >> 1) * kKeyStride + (threadIdx.x & 1)]);
  __syncthreads();
  const float lum_max = key_float(kmax[0]);
  const float lambda = 0.2f * key_float(kmax[1]);
  const float lo = lum_max * 0.20f, hi = lum_max * 0.95f;
  const int64_t quads = (n + 3) / 4;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < quads;
       q += (int64_t)gridDim.x * blockDim.x) {
    Quad d;
    bool ok[4];
    load_quad(lab, n, q, vec, d, ok);
    float o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float L = d.v[0][i];
      const bool grey = fabsf(d.v[1][i]) < lambda && fabsf(d.v[2][i]) < lambda;
      o[i] = !grey ? 0.f : L < lo ? -mu : L > hi ? mu : 0.f;
    }
    store_plane(beta, 4 * q, vec && ((uintptr_t)beta & 15u) == 0, o, ok);
  }
}

int grid_for(int64_t n) {
  const int64_t quads = (n + 3) / 4;
  const int64_t want = (quads + kColorThreads - 1) / kColorThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(want, 8 * device_cu_count()));
}

int plane_vec(const float* p, int64_t n) {
  return ((uintptr_t)p & 15u) == 0 && n % 4 == 0;
}

template <int CONV>
void launch_convert(const float* in, float* out, int64_t n, int vec,
                    hipStream_t stream) {
  hipLaunchKernelGGL((convert_kernel<CONV>), dim3(grid_for(n)),
                     dim3(kColorThreads), 0, stream, in, out, n, vec);
}

// Planes [p, p + 3n) and [q, q + 3n) overlap without being the same.
bool partial_overlap(const float* p, const float* q, int64_t n) {
  if (p == q) return false;
  return p < q + 3 * n && q < p + 3 * n;
}

}  // namespace
}  // namespace splithip

using namespace splithip;

extern "C" {

int split_hip_convert_color_space(const float* in, float* out, int64_t n,
                                  int32_t conversion,
                                  split_hip_stream_t stream) {
  if (!in || !out || n < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "convert_color_space: in/out missing or n < 1");
  if (partial_overlap(in, out, n))
    return fail(SPLIT_HIP_ERR_ARGUMENT,
                "convert_color_space: in and out overlap without being equal");
  const int vec = plane_vec(in, n) && plane_vec(out, n);
  switch (conversion) {
    case SPLIT_HIP_RGB_TO_XYZ: launch_convert<SPLIT_HIP_RGB_TO_XYZ>(in, out, n, vec, stream); break;
    case SPLIT_HIP_XYZ_TO_RGB: launch_convert<SPLIT_HIP_XYZ_TO_RGB>(in, out, n, vec, stream); break;
    case SPLIT_HIP_XYZ_TO_LAB: launch_convert<SPLIT_HIP_XYZ_TO_LAB>(in, out, n, vec, stream); break;
    case SPLIT_HIP_LAB_TO_XYZ: launch_convert<SPLIT_HIP_LAB_TO_XYZ>(in, out, n, vec, stream); break;
    case SPLIT_HIP_RGB_TO_LAB: launch_convert<SPLIT_HIP_RGB_TO_LAB>(in, out, n, vec, stream); break;
    case SPLIT_HIP_LAB_TO_RGB: launch_convert<SPLIT_HIP_LAB_TO_RGB>(in, out, n, vec, stream); break;
    case SPLIT_HIP_RGB_TO_IC: launch_convert<SPLIT_HIP_RGB_TO_IC>(in, out, n, vec, stream); break;
    default:
      return fail(SPLIT_HIP_ERR_ARGUMENT, "convert_color_space: unknown conversion %d",
                  conversion);
  }
  SPLIT_HIP_CHECK_LAUNCH("convert_color_space");
  return SPLIT_HIP_OK;
}

size_t split_hip_beta_feature_workspace(void) {
  return kKeyCopies * kKeyStride * sizeof(unsigned);
}

int split_hip_beta_feature(const float* lab, int64_t n, float* beta, float mu,
                           void* temp, size_t temp_bytes,
                           split_hip_stream_t stream) {
  if (!lab || !beta || n < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "beta_feature: lab/beta missing or n < 1");
  if (!temp || temp_bytes < split_hip_beta_feature_workspace())
    return fail(SPLIT_HIP_ERR_TEMP, "beta_feature: temp holds %zu bytes, needs %zu",
                temp_bytes, split_hip_beta_feature_workspace());
  // beta may be one of lab's planes (the sample writes it over L); any
  // other overlap would change lab before it is read
  if (beta != lab && beta != lab + n && beta != lab + 2 * n &&
      beta < lab + 3 * n && lab < beta + n)
    return fail(SPLIT_HIP_ERR_ARGUMENT,
                "beta_feature: beta overlaps lab without being one of its planes");
  unsigned* keys = static_cast<unsigned*>(temp);
  const int vec = plane_vec(lab, n);
  hipLaunchKernelGGL(beta_init, dim3(1), dim3(256), 0, stream, keys);
  hipLaunchKernelGGL(beta_maxima, dim3(grid_for(n)), dim3(kColorThreads), 0,
                     stream, lab, n, vec, keys);
  hipLaunchKernelGGL(beta_kernel, dim3(grid_for(n)), dim3(kColorThreads), 0,
                     stream, lab, n, vec, keys, mu, beta);
  SPLIT_HIP_CHECK_LAUNCH("beta_feature");
  return SPLIT_HIP_OK;
}

int split_hip_lab_beta_feature(const float* rgb, int64_t n, float* features,
                               float* luminance, float mu, void* temp,
                               size_t temp_bytes, split_hip_stream_t stream) {
  if (!rgb || !features || n < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "lab_beta_feature: rgb/features missing or n < 1");
  if (!temp || temp_bytes < split_hip_beta_feature_workspace())
    return fail(SPLIT_HIP_ERR_TEMP, "lab_beta_feature: temp holds %zu bytes, needs %zu",
                temp_bytes, split_hip_beta_feature_workspace());
  if (partial_overlap(rgb, features, n))
    return fail(SPLIT_HIP_ERR_ARGUMENT,
                "lab_beta_feature: rgb and features overlap without being equal");
  if (luminance && luminance < features + 3 * n && features < luminance + n)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "lab_beta_feature: luminance overlaps features");
  if (luminance && luminance < rgb + 3 * n && rgb < luminance + n)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "lab_beta_feature: luminance overlaps rgb");
  unsigned* keys = static_cast<unsigned*>(temp);
  const int vec = plane_vec(rgb, n) && plane_vec(features, n) &&
                  (!luminance || ((uintptr_t)luminance & 15u) == 0);
  hipLaunchKernelGGL(beta_init, dim3(1), dim3(256), 0, stream, keys);
  hipLaunchKernelGGL(lab_maxima, dim3(grid_for(n)), dim3(kColorThreads), 0,
                     stream, rgb, features, luminance, n, vec, keys);
  hipLaunchKernelGGL(beta_kernel, dim3(grid_for(n)), dim3(kColorThreads), 0,
                     stream, features, n, vec, keys, mu, features);
  SPLIT_HIP_CHECK_LAUNCH("lab_beta_feature");
  return SPLIT_HIP_OK;
}

}  // extern "C"
