// Device copy of the seeded image generator (split_amd/csrc/gen.c): the same
// counter-based hashes, Voronoi colours, value-noise height map, Lambertian
// shading, noise and clamping, for many images per launch.  Used where a
// run needs far more synthetic pixels than the host can make in time (the
// warm-start network's training set, large batches).
//
// Integer hashing, the Voronoi search (brute force over all sites, ties to
// the lower index — the host's grid search finds the same site) and the
// double arithmetic of the height map and shading are IEEE-exact and match
// the host bit for bit (FMA contraction is off in this file).  The noise
// calls log/cos from the device math library, whose last bit may differ from
// glibc's, so a noisy pixel can differ from the host image by one float ulp.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "split_gen.h"

#pragma clang fp contract(off)

namespace splithip {
namespace {

struct GenSite {
  double x, y;
  float r, g, b, pad;
};

struct GenImg {
  float* planes;
  int W, H;
  split_gen_params p;
  GenSite* sites;
  double* h;  // height map (shading images), W*H
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t hash4(uint64_t seed, uint64_t stream,
                                          uint64_t a, uint64_t b) {
  uint64_t h = mix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
  h = mix64(h ^ (a * 0x8CB92BA72F3D8DD7ull));
  h = mix64(h ^ (b * 0xABC98388FB8FAC03ull));
  return h;
}

__device__ __forceinline__ double u01(uint64_t x) {
  return (double)(x >> 11) * 0x1.0p-53;
}

enum { kStreamSite = 1, kStreamLattice = 2, kStreamNoise = 3 };

__global__ void gen_sites(const GenImg* imgs) {
  const GenImg& m = imgs[blockIdx.y];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m.p.sites;
       i += gridDim.x * blockDim.x) {
    const uint64_t s = m.p.seed;
    GenSite t;
    t.x = u01(hash4(s, kStreamSite, (uint64_t)i, 0)) * m.W;
    t.y = u01(hash4(s, kStreamSite, (uint64_t)i, 1)) * m.H;
    t.r = (float)u01(hash4(s, kStreamSite, (uint64_t)i, 2));
    t.g = (float)u01(hash4(s, kStreamSite, (uint64_t)i, 3));
    t.b = (float)u01(hash4(s, kStreamSite, (uint64_t)i, 4));
    t.pad = 0.f;
    m.sites[i] = t;
  }
}

__device__ __forceinline__ double lattice(uint64_t seed, int octave, int64_t ix,
                                          int64_t iy) {
  return u01(hash4(seed, kStreamLattice + 16 * (uint64_t)octave, (uint64_t)ix,
                   (uint64_t)iy));
}

__device__ double value_noise(uint64_t seed, int octave, double x, double y) {
  const double fx = floor(x), fy = floor(y);
  const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
  double tx = x - fx, ty = y - fy;
  tx = tx * tx * (3.0 - 2.0 * tx);
  ty = ty * ty * (3.0 - 2.0 * ty);
  const double a = lattice(seed, octave, ix, iy);
  const double b = lattice(seed, octave, ix + 1, iy);
  const double c = lattice(seed, octave, ix, iy + 1);
  const double d = lattice(seed, octave, ix + 1, iy + 1);
  const double top = a + (b - a) * tx;
  const double bot = c + (d - c) * tx;
  return top + (bot - top) * ty;
}

__global__ void gen_height(const GenImg* imgs) {
  const GenImg& m = imgs[blockIdx.y];
  if (!m.p.shading) return;
  const int64_t n = (int64_t)m.W * m.H;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / m.W), x = (int)(i - (int64_t)y * m.W);
    double h = 0.0, amp = 1.0, cell = m.p.base_cell;
    for (int o = 0; o < m.p.octaves; ++o) {
      h += amp * value_noise(m.p.seed, o, (x + 0.5) / cell, (y + 0.5) / cell);
      amp *= 0.5;
      cell *= 0.5;
    }
    m.h[i] = m.p.height_amp * h;
  }
}

__device__ double gauss(uint64_t seed, uint64_t key) {
  const uint64_t a = hash4(seed, kStreamNoise, key, 0);
  const uint64_t b = hash4(seed, kStreamNoise, key, 1);
  const double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
  const double u2 = u01(b);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__device__ double uniform_pm1(uint64_t seed, uint64_t key) {
  return 2.0 * u01(hash4(seed, kStreamNoise, key, 2)) - 1.0;
}

constexpr int kGenThreads = 256;
constexpr int kSiteChunk = 512;

// One thread per pixel; sites are staged through LDS in chunks.
__global__ __launch_bounds__(kGenThreads) void gen_pixels(const GenImg* imgs) {
  const GenImg& m = imgs[blockIdx.y];
  __shared__ double sx[kSiteChunk], sy[kSiteChunk];
  const int64_t n = (int64_t)m.W * m.H;
  const int W = m.W, H = m.H;
  const double ln = sqrt(0.3 * 0.3 + 0.3 * 0.3 + 0.906 * 0.906);
  const double Lx = 0.3 / ln, Ly = 0.3 / ln, Lz = 0.906 / ln;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    const bool live = i < n;
    const int y = live ? (int)(i / W) : 0;
    const int x = live ? (int)(i - (int64_t)y * W) : 0;
    const double px = x + 0.5, py = y + 0.5;
    double best = INFINITY;
    int bi = 0;
    for (int c0 = 0; c0 < m.p.sites; c0 += kSiteChunk) {
      const int cn = min(kSiteChunk, m.p.sites - c0);
      __syncthreads();
      for (int t = threadIdx.x; t < cn; t += blockDim.x) {
        sx[t] = m.sites[c0 + t].x;
        sy[t] = m.sites[c0 + t].y;
      }
      __syncthreads();
      for (int t = 0; t < cn; ++t) {
        const double dx = px - sx[t], dy = py - sy[t];
        const double d = dx * dx + dy * dy;
        if (d < best) {  // ascending index: ties keep the lower one
          best = d;
          bi = c0 + t;
        }
      }
    }
    if (!live) continue;
    double shade = 1.0;
    if (m.p.shading) {
      const double* h = m.h;
      const int xl = x > 0 ? x - 1 : x, xr = x < W - 1 ? x + 1 : x;
      const int yu = y > 0 ? y - 1 : y, yd = y < H - 1 ? y + 1 : y;
      const double dhdx = (h[(int64_t)y * W + xr] - h[(int64_t)y * W + xl]) /
                          (xr - xl > 0 ? xr - xl : 1);
      const double dhdy = (h[(int64_t)yd * W + x] - h[(int64_t)yu * W + x]) /
                          (yd - yu > 0 ? yd - yu : 1);
      const double nl = sqrt(dhdx * dhdx + dhdy * dhdy + 1.0);
      const double d = (-dhdx * Lx - dhdy * Ly + Lz) / nl;
      shade = d > 0.0 ? d : 0.0;
    }
    const GenSite st = m.sites[bi];
    const float alb[3] = {st.r, st.g, st.b};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = (double)alb[c] * shade;
      const uint64_t key = (uint64_t)i * 3 + (uint64_t)c;
      if (m.p.noise_sigma > 0.0) v += m.p.noise_sigma * gauss(m.p.seed, key);
      if (m.p.noise_uniform > 0.0)
        v += m.p.noise_uniform * uniform_pm1(m.p.seed, key);
      v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
      m.planes[(int64_t)c * n + i] = (float)v;
    }
  }
}

size_t gen_layout(const split_hip_gen_desc* d, int count, void* temp,
                  std::vector<GenImg>* out) {
  char* t = static_cast<char*>(temp);
  size_t o = align_up((size_t)count * sizeof(GenImg), 256);
  if (out) out->resize(count);
  for (int i = 0; i < count; ++i) {
    const int64_t n = (int64_t)d[i].width * d[i].height;
    GenImg m{};
    m.planes = d[i].planes;
    m.W = d[i].width;
    m.H = d[i].height;
    m.p = d[i].params;
    m.sites = reinterpret_cast<GenSite*>(t ? t + o : nullptr);
    o += align_up((size_t)std::max(d[i].params.sites, 1) * sizeof(GenSite), 256);
    if (d[i].params.shading) {
      m.h = reinterpret_cast<double*>(t ? t + o : nullptr);
      o += align_up((size_t)n * sizeof(double), 256);
    }
    if (out) (*out)[i] = m;
  }
  return o;
}

int check_gen(const split_hip_gen_desc* d, int count, bool need_planes) {
  if (!d || count < 1) return fail(SPLIT_HIP_ERR_ARGUMENT, "no images");
  for (int i = 0; i < count; ++i) {
    const split_gen_params& p = d[i].params;
    if (d[i].width < 1 || d[i].height < 1 ||
        (int64_t)d[i].width * d[i].height > INT32_MAX)
      return fail(SPLIT_HIP_ERR_ARGUMENT, "image %d is %d x %d", i, d[i].width,
                  d[i].height);
    if (p.sites < 1 || p.octaves < 0 || (p.shading && !(p.base_cell > 0.0)))
      return fail(SPLIT_HIP_ERR_ARGUMENT,
                  "image %d: sites %d, octaves %d, base_cell %g", i, p.sites,
                  p.octaves, p.base_cell);
    if (need_planes && !d[i].planes)
      return fail(SPLIT_HIP_ERR_ARGUMENT, "image %d has no planes", i);
  }
  return SPLIT_HIP_OK;
}

}  // namespace
}  // namespace splithip

using namespace splithip;

extern "C" {

size_t split_hip_gen_images_workspace(const split_hip_gen_desc* images,
                                      int32_t count) {
  if (check_gen(images, count, false) != SPLIT_HIP_OK) return 0;
  return gen_layout(images, count, nullptr, nullptr);
}

int split_hip_gen_images(const split_hip_gen_desc* images, int32_t count,
                         void* temp, size_t temp_bytes,
                         split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_gen(images, count, true));
  if (!temp) return fail(SPLIT_HIP_ERR_ARGUMENT, "temp is null");
  std::vector<GenImg> imgs;
  const size_t need = gen_layout(images, count, temp, &imgs);
  if (temp_bytes < need)
    return fail(SPLIT_HIP_ERR_TEMP, "temp holds %zu bytes, generator needs %zu",
                temp_bytes, need);
  SPLIT_HIP_CHECK(hipMemcpyAsync(temp, imgs.data(), count * sizeof(GenImg),
                                 hipMemcpyHostToDevice, stream));
  int64_t maxn = 0;
  int maxs = 0;
  for (const auto& m : imgs) {
    maxn = std::max<int64_t>(maxn, (int64_t)m.W * m.H);
    maxs = std::max(maxs, m.p.sites);
  }
  const GenImg* dimg = static_cast<const GenImg*>(temp);
  const int cus = device_cu_count();
  const int per_img = (int)std::max<int64_t>(
      1, std::min<int64_t>((maxn + kGenThreads - 1) / kGenThreads,
                           std::max(1, 8 * cus / count)));
  hipLaunchKernelGGL(gen_sites, dim3((maxs + 255) / 256, count), dim3(256), 0,
                     stream, dimg);
  hipLaunchKernelGGL(gen_height, dim3(per_img, count), dim3(256), 0, stream,
                     dimg);
  hipLaunchKernelGGL(gen_pixels, dim3(per_img, count), dim3(kGenThreads), 0,
                     stream, dimg);
  SPLIT_HIP_CHECK_LAUNCH("gen_images");
  // the descriptor table is read by the kernels: keep the host copy alive
  SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
  return SPLIT_HIP_OK;
}

}  // extern "C"
