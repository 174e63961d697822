// Device-side pieces shared by the Lloyd kernels: traversal geometry,
// segment loading (1-D runs or 16x16 tiles) and the exact pruned
// nearest-centroid assignment.
#pragma once

#include "common.hpp"

namespace splithip {

constexpr int kSegPx = 256;  // pixels per wave segment (64 lanes x 4)

// What a kernel walks: n pixels of three SoA planes, as 1-D runs of 256
// pixels (width == 0) or as 16x16 tiles of a width x height image.
struct Geometry {
  const float* pts;  // plane c at pts + c*n
  int64_t n;
  int64_t nseg;      // segments: ceil(n / kSegPx) or tile count
  int width;         // > 0: 16x16-tile traversal of width x height
  int height;        // rows: ceil(n / width)
  int tiles_x;       // ceil(width / 16)
  // rows that hold `width` pixels: height for an image; n / width for a
  // point set walked as a width-16 image, whose last row may be partial
  int full_rows;
  // tiles are (1 << tw_log) x (256 >> tw_log) pixels, tw_log in 4..6: 16x16,
  // 32x8 or 64x4 (a lane holds 4 consecutive pixels of one tile row)
  int tw_log = 4;
  // optional interleaved copy of the pixels ({x, y, z, 0} per pixel) for
  // scattered reads: one 16-byte load touches one cache line instead of
  // three (one per plane)
  const F4* aos = nullptr;
};

// ------------------------------------------------------------- loading ---

// Pixels in row `row` of a tiled traversal: width, fewer in a partial last
// row (n not a multiple of width), none at or below height.
__device__ __forceinline__ int row_pixels(const Geometry& a, int row) {
  if (row < a.full_rows) return a.width;
  const int64_t left = a.n - (int64_t)row * a.width;
  return left < (int64_t)a.width ? (int)(left > 0 ? left : 0) : a.width;
}

// Traversal modes: 1-D runs of 256 pixels (dword or float4 loads) or 16x16
// tiles (dword or float4 loads per lane row).
enum Mode { kRun = 0, kRunVec = 1, kTile = 2, kTileVec = 3 };
constexpr int kTilePx = 256;

// Tile shape of a Geometry and the lane layout within a tile: lane l holds
// pixels 4 (l mod (w/4)) .. +3 of tile row l / (w/4).
__host__ __device__ __forceinline__ int tile_w(const Geometry& a) {
  return 1 << a.tw_log;
}
__host__ __device__ __forceinline__ int tile_h(const Geometry& a) {
  return kTilePx >> a.tw_log;
}
__device__ __forceinline__ int lane_row(const Geometry& a, int lane) {
  return lane >> (a.tw_log - 2);
}
__device__ __forceinline__ int lane_col(const Geometry& a, int lane) {
  return 4 * (lane & ((1 << (a.tw_log - 2)) - 1));
}
// Tiles across and down a width x rows area.
__host__ __device__ __forceinline__ int64_t tiles_across(int width, int tw_log) {
  return (width + (1 << tw_log) - 1) >> tw_log;
}
__host__ __device__ __forceinline__ int64_t tiles_down(int64_t rows,
                                                       int tw_log) {
  const int th = kTilePx >> tw_log;
  return (rows + th - 1) / th;
}

// Pixel indices of segment s for this lane: idx[i] of pixel i, v[i] whether
// it exists, lvec whether the lane's four pixels are consecutive and exist
// (so per-pixel arrays can be read or written as one dword-aligned 16-byte
// vector).
template <int MODE>
__device__ __forceinline__ void segment_index(const Geometry& a, int64_t s,
                                              int lane, int64_t (&idx)[4],
                                              bool (&v)[4], bool& lvec) {
  if constexpr (MODE == kTile || MODE == kTileVec) {
    const int ty = (int)(s / a.tiles_x);
    const int tx = (int)(s - (int64_t)ty * a.tiles_x);
    const int row = ty * tile_h(a) + lane_row(a, lane);
    const int col = tx * tile_w(a) + lane_col(a, lane);
    const int64_t o = (int64_t)row * a.width + col;
    const int rp = row_pixels(a, row);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = col + i < rp;
      idx[i] = o + i;
    }
    lvec = MODE == kTileVec && col + 3 < rp;
  } else {
    const int64_t base = s * kSegPx;
    lvec = MODE == kRunVec && base + kSegPx <= a.n;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      idx[i] = lvec ? base + 4 * lane + i : base + lane + 64 * i;
      v[i] = idx[i] < a.n;
    }
  }
}

__device__ __forceinline__ void load_colours(const Geometry& a,
                                             const int64_t (&idx)[4],
                                             const bool (&v)[4], bool lvec,
                                             float (&x)[4], float (&y)[4],
                                             float (&z)[4]) {
  const auto p0 = gmem(a.pts);
  const auto p1 = gmem(a.pts + a.n);
  const auto p2 = gmem(a.pts + 2 * a.n);
  if (lvec) {
    const int64_t o = idx[0];
    const F4 X = *gmem(reinterpret_cast<const F4*>(a.pts + o));
    const F4 Y = *gmem(reinterpret_cast<const F4*>(a.pts + a.n + o));
    const F4 Z = *gmem(reinterpret_cast<const F4*>(a.pts + 2 * a.n + o));
    x[0] = X.x; x[1] = X.y; x[2] = X.z; x[3] = X.w;
    y[0] = Y.x; y[1] = Y.y; y[2] = Y.z; y[3] = Y.w;
    z[0] = Z.x; z[1] = Z.y; z[2] = Z.z; z[3] = Z.w;
  } else if (a.aos) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const F4 q = gmem(a.aos)[v[i] ? idx[i] : 0];
      x[i] = q.x;
      y[i] = q.y;
      z[i] = q.z;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t j = v[i] ? idx[i] : 0;
      x[i] = p0[j];
      y[i] = p1[j];
      z[i] = p2[j];
    }
  }
}

// Loads segment s: x/y/z[i] of pixel idx[i] (see segment_index).
template <int MODE>
__device__ __forceinline__ void load_segment(const Geometry& a, int64_t s,
                                             int lane, float (&x)[4],
                                             float (&y)[4], float (&z)[4],
                                             bool (&v)[4], int64_t (&idx)[4],
                                             bool& lvec) {
  segment_index<MODE>(a, s, lane, idx, v, lvec);
  load_colours(a, idx, v, lvec, x, y, z);
}

// ---------------------------------------------------------- assignment ---

__device__ __forceinline__ float sqdist(float px, float py, float pz, float cx,
                                        float cy, float cz) {
  const float dx = px - cx, dy = py - cy, dz = pz - cz;
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// With TOP2, slack[i] also receives a lower bound on the gap between the
// distance from pixel i to its second-nearest centroid and to its nearest
// one (non-squared, rounded down): the second-best evaluated distance, the
// box lower bound of every box-pruned centroid, and for bisector-pruned
// centroids |p-c_j|^2 >= |p-c_g|^2 + 2 (b_j - max_box p.w_j) >= best^2 + 2 m.
// PRUNE = false evaluates every centroid (batches whose pixels come from
// all over an image, whose box keeps nearly all of them anyway), as a
// counted loop unrolled by 4 with COUNTED, else as a walk over a ballot
// mask (per 4096^2 k = 32 iteration 60.0 / 60.7 against 61.2 / 61.5 us
// counted; the batched walker, at 4 waves per SIMD, lost 1.7% with it).
template <int CPL, bool TOP2, bool PRUNE = true, bool COUNTED = false>
__device__ __forceinline__ void assign_segment(
    const float (&cx)[CPL], const float (&cy)[CPL], const float (&cz)[CPL],
    const float (&cn)[CPL], int k, int lane, const float (&x)[4],
    const float (&y)[4], const float (&z)[4], const bool (&v)[4],
    int (&lab)[4], float* slack) {
  float bd[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  float bd2[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  int bk[4] = {0, 0, 0, 0};
  float lbox = INFINITY, mbis = INFINITY;  // per-lane, reduced below
  // One candidate (lane bl of register t) against the four pixels.
  auto eval = [&](int t, int bl) {
    const float ccx = readlane_f(cx[t], bl);
    const float ccy = readlane_f(cy[t], bl);
    const float ccz = readlane_f(cz[t], bl);
    const int j = 64 * t + bl;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float d = sqdist(x[i], y[i], z[i], ccx, ccy, ccz);
      const bool better = d < bd[i];
      // with bd <= bd2 always, the median of (bd, d, bd2) is the new
      // runner-up: one v_med3_f32
      if constexpr (TOP2) bd2[i] = __builtin_amdgcn_fmed3f(bd[i], d, bd2[i]);
      bd[i] = fminf(bd[i], d);
      bk[i] = better ? j : bk[i];
    }
  };
  if constexpr (!PRUNE) {
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
      if constexpr (COUNTED) {
        const int kt = k - 64 * t < 64 ? k - 64 * t : 64;
#pragma unroll 4
        for (int bl = 0; bl < kt; ++bl) eval(t, bl);
      } else {
        uint64_t m = ballot(lane + 64 * t < k);
        while (m) {
          const int bl = __builtin_ctzll(m);
          m &= m - 1;
          eval(t, bl);
        }
      }
    }
  } else {
  // Colour box of the segment's valid pixels.
  float lx = INFINITY, ly = INFINITY, lz = INFINITY;
  float hx = -INFINITY, hy = -INFINITY, hz = -INFINITY;
  if (ballot(!(v[0] && v[1] && v[2] && v[3])) == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      lx = fminf(lx, x[i]);  hx = fmaxf(hx, x[i]);
      ly = fminf(ly, y[i]);  hy = fmaxf(hy, y[i]);
      lz = fminf(lz, z[i]);  hz = fmaxf(hz, z[i]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      lx = v[i] ? fminf(lx, x[i]) : lx;  hx = v[i] ? fmaxf(hx, x[i]) : hx;
      ly = v[i] ? fminf(ly, y[i]) : ly;  hy = v[i] ? fmaxf(hy, y[i]) : hy;
      lz = v[i] ? fminf(lz, z[i]) : lz;  hz = v[i] ? fmaxf(hz, z[i]) : hz;
    }
  }
  lx = wave_allreduce(lx, OpMin()); hx = wave_allreduce(hx, OpMax());
  ly = wave_allreduce(ly, OpMin()); hy = wave_allreduce(hy, OpMax());
  lz = wave_allreduce(lz, OpMin()); hz = wave_allreduce(hz, OpMax());

  // (a) lower / upper bounds of each centroid's squared distance to the box.
  float lower[CPL], upv[CPL];
  float upper = INFINITY;
#pragma unroll
  for (int t = 0; t < CPL; ++t) {
    const bool live = lane + 64 * t < k;
    const float ex = fmaxf(fmaxf(lx - cx[t], cx[t] - hx), 0.f);
    const float ey = fmaxf(fmaxf(ly - cy[t], cy[t] - hy), 0.f);
    const float ez = fmaxf(fmaxf(lz - cz[t], cz[t] - hz), 0.f);
    const float fx = fmaxf(fabsf(cx[t] - lx), fabsf(cx[t] - hx));
    const float fy = fmaxf(fabsf(cy[t] - ly), fabsf(cy[t] - hy));
    const float fz = fmaxf(fabsf(cz[t] - lz), fabsf(cz[t] - hz));
    lower[t] = live ? fmaf(ez, ez, fmaf(ey, ey, ex * ex)) : INFINITY;
    upv[t] = live ? fmaf(fz, fz, fmaf(fy, fy, fx * fx)) : INFINITY;
    upper = fminf(upper, upv[t]);
  }
  upper = wave_allreduce(upper, OpMin());
  const float keep = upper * 1.00001f;

  // (b) bisector test against g, the lowest-index centroid whose farthest
  // corner is nearest.  For any pixel p, |p-c_j|^2 - |p-c_g|^2 =
  // 2 (b_j - p.w_j) with w_j = c_j - c_g and b_j = (|c_j|^2 - |c_g|^2) / 2,
  // so j is dropped when max over the box of p.w_j stays below b_j by the
  // margin.
  int g = 0;
#pragma unroll
  for (int t = CPL - 1; t >= 0; --t) {
    const uint64_t m = ballot(lane + 64 * t < k && upv[t] == upper);
    if (m) g = 64 * t + __builtin_ctzll(m);
  }
  float gx = 0.f, gy = 0.f, gz = 0.f, gn = 0.f;
#pragma unroll
  for (int t = 0; t < CPL; ++t) {
    if ((g >> 6) == t) {
      gx = readlane_f(cx[t], g & 63);
      gy = readlane_f(cy[t], g & 63);
      gz = readlane_f(cz[t], g & 63);
      gn = readlane_f(cn[t], g & 63);
    }
  }
  const float ax = fmaxf(fabsf(lx), fabsf(hx)), ay = fmaxf(fabsf(ly), fabsf(hy)),
              az = fmaxf(fabsf(lz), fabsf(hz));

#pragma unroll
  for (int t = 0; t < CPL; ++t) {
    const float wx = cx[t] - gx, wy = cy[t] - gy, wz = cz[t] - gz;
    const float b = 0.5f * (cn[t] - gn);
    const float mx = fmaf(wz, wz > 0.f ? hz : lz,
                          fmaf(wy, wy > 0.f ? hy : ly, wx * (wx > 0.f ? hx : lx)));
    const float scale = 1.f + fabsf(b) + upper +
                        fmaf(az, fabsf(wz), fmaf(ay, fabsf(wy), ax * fabsf(wx)));
    const bool side = mx >= b - 1e-5f * scale;
    const bool live = lane + 64 * t < k;
    if constexpr (TOP2) {
      lbox = (live && lower[t] > keep) ? fminf(lbox, lower[t]) : lbox;
      mbis = (live && lower[t] <= keep && !side)
                 ? fminf(mbis, (b - mx) - 1e-6f * scale)
                 : mbis;
    }
    uint64_t m = ballot(live && lower[t] <= keep && side);
    while (m) {
      const int bl = __builtin_ctzll(m);
      m &= m - 1;
      eval(t, bl);
    }
  }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) lab[i] = v[i] ? bk[i] : -1;
  if constexpr (TOP2) {
    lbox = wave_allreduce(lbox, OpMin());
    mbis = wave_allreduce(mbis, OpMin());
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      // squared lower bound of the runner-up, then a 2e-5 relative shrink
      // (and growth of the winner) far above the float rounding involved
      // (hardware square root, within 1 ulp: far inside the 2e-5 shrink)
      const float l2 = fminf(fminf(bd2[i], lbox), bd[i] + 2.f * fmaxf(mbis, 0.f));
      const float s = __builtin_amdgcn_sqrtf(l2 * (1.f - 2e-5f)) -
                      __builtin_amdgcn_sqrtf(bd[i] * (1.f + 2e-5f));
      slack[i] = s > 0.f ? s : 0.f;
    }
  }
}


}  // namespace splithip
