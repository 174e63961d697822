// Neighbour-ball re-assignment for the bounded Lloyd iterations at large k.
//
// The box pruning of a 256-pixel batch keeps nearly every centroid when the
// batch mixes pixels from all over the image (the queued pixels) or when
// the colour noise spans several centroid spacings (k = 256): a batch then
// evaluates ~90-240 of 256 candidates.  Here each pixel walks instead the
// list of its previous label's nearest centroids, nearest first, and stops
// as soon as the triangle inequality puts every centroid not yet visited
// beyond the runner-up found so far: a handful of candidates per pixel
// (a batch costs the longest of its 256 walks, ~16 on the 8192^2 k = 256
// image).
//
// Lists: launch s builds, from its centroids c_s, the m nearest other
// centroids of every centroid (one wave per list), and launch s + 1 walks
// them.  An entry holds a lower bound of |c_s[a] - c_s[j]| (the top 20 bits
// of its float: 11 mantissa bits, truncated) over 16 j, the byte offset of
// centroid j in a float4 table, so that entries sort by distance, then
// index, and a walk step gets the centroid's LDS address with one AND.  In
// launch s + 1 the centroids have moved: |c_{s+1}[a] - c_{s+1}[j]| >=
// entry - mv_a - mv_j >= entry - corr[a], corr[a] = mv_a + max_j mv_j.
// For a pixel x of previous label a and every centroid j at or after list
// position p (and every unlisted one): |x - c_j| >= |c_a - c_j| - |x - c_a|
// >= entry_p - corr[a] - |x - c_a|.  The walk starts from c_a, visits the
// list in order and stops when that bound, less a margin for the float
// error, reaches the runner-up; the nearest and the runner-up are then
// exact (ties to the lowest index, as the dense argmin).  A walk that ends
// the list with the nearest not settled is redone over every centroid.
#pragma once

#include "common.hpp"
#include "lloyd_device.hpp"

namespace splithip {

constexpr int kNbMax = 32;  // neighbours listed per centroid (k <= 256)

// Entries per list for k centroids.
__host__ __device__ __forceinline__ int nb_len(int k) {
  return k - 1 < kNbMax ? k - 1 : kNbMax;
}
__device__ __forceinline__ float nb_dist(uint32_t e) {
  return __uint_as_float(e & 0xFFFFF000u);
}
// 16 j: the byte offset of centroid j in a float4 table
__device__ __forceinline__ int nb_off(uint32_t e) { return (int)(e & 0xFF0u); }

__device__ __forceinline__ unsigned wave_min_u32(unsigned x) {
  x = min(x, (unsigned)dpp_i<kDppQuadXor1>((int)x));
  x = min(x, (unsigned)dpp_i<kDppQuadXor2>((int)x));
  x = min(x, (unsigned)dpp_i<kDppRowHalfMirror>((int)x));
  x = min(x, (unsigned)dpp_i<kDppRowMirror>((int)x));
  x = min(x, (unsigned)dpp_i<kDppRowBcast15, 0xA>((int)x, -1));
  x = min(x, (unsigned)dpp_i<kDppRowBcast31, 0xC>((int)x, -1));
  return (unsigned)readlane_i((int)x, 63);
}

// Centroid readers for build_neighbours: SoA planes (x[k], y[k], z[k]) or
// {x, y, z, -} vectors.
struct CentSoA {
  const float* c;
  int k;
  __device__ float3 operator()(int j) const {
    return make_float3(c[j], c[k + j], c[2 * k + j]);
  }
};
struct CentAoS {
  const float4* c;
  __device__ float3 operator()(int j) const {
    const float4 q = c[j];
    return make_float3(q.x, q.y, q.z);
  }
};

// One wave: the list of centroid a (m entries, nearest first) into out,
// k <= 256.
template <class Cent>
__device__ __forceinline__ void build_neighbours(Cent c, int k, int a, int m,
                                                 uint32_t* out, int lane) {
  const float3 ca = c(a);
  uint32_t key[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int j = lane + 64 * t;
    key[t] = 0xFFFFFFFFu;
    if (j < k && j != a) {
      // the float distance is within ~3e-7 of |c_a - c_j|: shrunk by 1e-6
      // and truncated, a lower bound
      const float3 cj = c(j);
      const float d =
          __builtin_amdgcn_sqrtf(sqdist(ca.x, ca.y, ca.z, cj.x, cj.y, cj.z)) *
          (1.f - 1e-6f);
      key[t] = (__float_as_uint(d) & 0xFFFFF000u) | ((uint32_t)j << 4);
    }
  }
  for (int r = 0; r < m; ++r) {
    const unsigned mn =
        wave_min_u32(min(min(key[0], key[1]), min(key[2], key[3])));
    if (lane == 0) out[r] = mn;
#pragma unroll
    for (int t = 0; t < 4; ++t) key[t] = key[t] == mn ? 0xFFFFFFFFu : key[t];
  }
}

// Nearest centroid (ties to the lowest index) and the runner-up distance
// bound of up to 256 pixels, 4 per lane, with previous labels old[] (valid
// where v[]).  cent: {x, y, z, -} per centroid; nb: k lists of m entries;
// corr: the per-label list correction; margin: absolute slack for the float
// error of the distances.  lab = -1 where !v; slack as assign_segment's.
__device__ __forceinline__ void assign_ball(
    const float4* cent, const uint32_t* nb, int m, const float* corr,
    float margin, int k, const float (&x)[4], const float (&y)[4],
    const float (&z)[4], const bool (&v)[4], const int (&old)[4],
    int (&lab)[4], float (&slack)[4]) {
  // bk: 16 x the index of the best centroid (its byte offset in cent; the
  // offsets order like the indices)
  const char* cb = reinterpret_cast<const char*>(cent);
  float best[4], second[4], lim[4];
  int bk[4], base[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int a = v[i] ? old[i] : 0;
    const float4 c = cent[a];
    best[i] = sqdist(x[i], y[i], z[i], c.x, c.y, c.z);
    second[i] = INFINITY;
    bk[i] = 16 * a;
    base[i] = a * m;
    // |x - c_a| <= the float distance (1 + 1e-6); margin covers the rest.
    // Invalid pixels: -inf, so that their walk stops at once.
    lim[i] = v[i] ? corr[a] + __builtin_amdgcn_sqrtf(best[i]) * (1.f + 1e-6f) +
                        margin
                  : -INFINITY;
  }
  // A walk stops at the first entry whose bound puts every centroid not yet
  // visited at or beyond the runner-up: (lb > 0) & (lb^2 >= second).  The
  // test is sticky without a flag: entries ascend and a stopped walk's
  // runner-up no longer changes.  Branch-free steps: the four slots' list
  // entries are read together (and the next step's one step ahead), then
  // their four centroids, so that a step waits for about one LDS round trip
  // instead of two per slot (per-slot branches serialised eight of them);
  // stopped slots read in-bounds entries (a = 0 for invalid pixels) and
  // count their centroid as infinitely far, which leaves best, second and
  // bk unchanged.
  auto stops = [&](uint32_t e, int i) {
    const float lb = nb_dist(e) - lim[i];
    return (lb > 0.f) & (lb * lb >= second[i]);
  };
  uint32_t e[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) e[i] = nb[base[i]];
  for (int p = 0; p < m; ++p) {
    bool stop[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) stop[i] = stops(e[i], i);
    if (ballot(!(stop[0] & stop[1] & stop[2] & stop[3])) == 0) break;
    const int pn = p + 1 < m ? p + 1 : p;
    uint32_t en[4];
    float4 c[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      en[i] = nb[base[i] + pn];
      c[i] = *reinterpret_cast<const float4*>(cb + nb_off(e[i]));
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int off = nb_off(e[i]);
      const float dd = sqdist(x[i], y[i], z[i], c[i].x, c[i].y, c[i].z);
      const float d = stop[i] ? INFINITY : dd;
      // bitwise, not short-circuit: no branches
      const bool take = (d < best[i]) | ((d == best[i]) & (off < bk[i]));
      // best <= second: the median is the new runner-up
      second[i] = __builtin_amdgcn_fmed3f(best[i], d, second[i]);
      bk[i] = take ? off : bk[i];
      best[i] = fminf(best[i], d);
      e[i] = en[i];
    }
  }
  // The unlisted centroids (m < k - 1) are at least the last entry's bound
  // away: a walk whose stop test holds at the last entry is settled (every
  // walk that stopped early, by stickiness); else the bound still caps the
  // runner-up if it exceeds the nearest, or the walk is redone over all k.
  bool redo[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    redo[i] = false;
    if (m < k - 1) {
      const uint32_t last = nb[base[i] + m - 1];
      const float lb = nb_dist(last) - lim[i];
      if (!stops(last, i)) {
        if (lb > 0.f && lb * lb > best[i])
          second[i] = fminf(second[i], lb * lb);
        else
          redo[i] = true;
      }
    }
  }
  if (ballot(redo[0] || redo[1] || redo[2] || redo[3])) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (!redo[i]) continue;
      best[i] = INFINITY;
      second[i] = INFINITY;
      bk[i] = 0;
    }
    for (int j = 0; j < k; ++j) {
      const float4 c = cent[j];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (!redo[i]) continue;
        const float d = sqdist(x[i], y[i], z[i], c.x, c.y, c.z);
        second[i] = __builtin_amdgcn_fmed3f(best[i], d, second[i]);
        bk[i] = d < best[i] ? 16 * j : bk[i];
        best[i] = fminf(best[i], d);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    lab[i] = v[i] ? bk[i] >> 4 : -1;
    // the same 2e-5 relative shrink as assign_segment's slack
    const float s = __builtin_amdgcn_sqrtf(second[i] * (1.f - 2e-5f)) -
                    __builtin_amdgcn_sqrtf(best[i] * (1.f + 2e-5f));
    slack[i] = s > 0.f ? s : 0.f;
  }
}

}  // namespace splithip
