// Exact integer accumulation of quantised pixel sums, shared by the
// single-image and the batched Lloyd kernels.
//
// Each channel c is quantised once per image with its own power-of-two
// scale, q = rint(v * s_c) with |q| <= 2^23 - 1 (s_c from the channel's
// largest |v|, see scale_of), and summed as integers: per-lane int32
// registers, exact wave sums, int64 LDS slices and int64 totals.  Integer
// sums do not depend on order, so every traversal of an image gives the same
// totals.  Rounding to nearest moves each pixel by at most half a quantum,
// so a mean moves by at most 1 / (2 s_c) <= max|v_c| * 2^-23, whatever the
// scales of the other channels.
#pragma once

#include "common.hpp"

namespace splithip {

constexpr int kRunFlushTiles = 64;

__device__ __forceinline__ int wave_sum_i32(int x) {
  return (int)wave_sum_u32((unsigned)x);
}

// Exact wave sum of int32 values whose total may exceed 32 bits.
__device__ __forceinline__ long long wave_sum_wide(int s) {
  const int lo = s & 0xFFFF;  // 0..65535
  const int hi = s >> 16;     // arithmetic shift: s = hi * 65536 + lo
  return ((long long)wave_sum_i32(hi) << 16) + (long long)wave_sum_i32(lo);
}

// SHARED: the slice is shared by the waves of a block (LDS atomics); else it
// is the wave's own.
template <bool SHARED = false>
__device__ __forceinline__ void slice_add64(long long* slice, int k, int label,
                                            long long sx, long long sy,
                                            long long sz, long long n,
                                            int lane) {
  if (lane == 0) {
    if constexpr (SHARED) {
      auto* u = reinterpret_cast<unsigned long long*>(slice);
      atomicAdd(u + label, (unsigned long long)sx);
      atomicAdd(u + k + label, (unsigned long long)sy);
      atomicAdd(u + 2 * k + label, (unsigned long long)sz);
      atomicAdd(u + 3 * k + label, (unsigned long long)n);
    } else {
      slice[label] += sx;
      slice[k + label] += sy;
      slice[2 * k + label] += sz;
      slice[3 * k + label] += n;
    }
  }
}

// Register run of a wave-uniform label: per-lane int32 sums of quantised
// pixels, flushed (exact wide wave sums) on a label change or every
// kRunFlushTiles tiles, which bounds a lane's sum by 256 (2^23 - 1) < 2^31.
struct IRun {
  int x, y, z, n;
  int label;
  int tiles;
};

__device__ __forceinline__ void irun_flush(IRun& r, long long* slice, int k,
                                           int lane) {
  if (r.label >= 0) {
    const long long sx = wave_sum_wide(r.x);
    const long long sy = wave_sum_wide(r.y);
    const long long sz = wave_sum_wide(r.z);
    const long long sn = wave_sum_i32(r.n);
    slice_add64(slice, k, r.label, sx, sy, sz, sn, lane);
  }
  r.x = r.y = r.z = r.n = 0;
  r.label = -1;
  r.tiles = 0;
}

// Adds sign * (q, 1) of each pixel with lab[i] >= 0 into its label's slice
// entries, one distinct label at a time (|per-lane sum| <= 4 (2^23 - 1), so
// a wave sum stays below 2^31).
template <bool SHARED = false>
__device__ __forceinline__ void add_by_label(long long* slice, int k, int lane,
                                             const int (&lab)[4],
                                             const int (&qx)[4],
                                             const int (&qy)[4],
                                             const int (&qz)[4], int sign) {
  unsigned rem = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) rem |= (lab[i] >= 0 ? 1u : 0u) << i;
  for (;;) {
    const uint64_t act = ballot(rem != 0);
    if (act == 0) break;
    const int first_lane = __builtin_ctzll(act);
    const int mine = (rem & 1u) ? lab[0]
                     : (rem & 2u) ? lab[1]
                     : (rem & 4u) ? lab[2]
                     : (rem & 8u) ? lab[3] : -1;
    const int L = readlane_i(mine, first_lane);
    int sx = 0, sy = 0, sz = 0, cnt = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool m = ((rem >> i) & 1u) && lab[i] == L;
      sx += m ? qx[i] : 0;
      sy += m ? qy[i] : 0;
      sz += m ? qz[i] : 0;
      cnt += m ? 1 : 0;
      rem &= m ? ~(1u << i) : ~0u;
    }
    const long long tx = wave_sum_i32(sx), ty = wave_sum_i32(sy),
                    tz = wave_sum_i32(sz), tn = wave_sum_i32(cnt);
    slice_add64<SHARED>(slice, k, L, sign * tx, sign * ty, sign * tz,
                        sign * tn, lane);
  }
}

// Quantisation scale of a channel whose largest |v| has the float bits
// maxabs_bits: s = 2^(23 - e) with max|v| < 2^e, halved in the one case
// where rint(max|v| * s) would reach 2^23, so |rint(v * s)| <= 2^23 - 1 for
// every v of the channel.  v * s is exact (a power-of-two scaling).  e is
// held at -100 or above so that s stays finite for channels of tiny or zero
// range (an all-zero channel quantises to 0).
__device__ __forceinline__ float scale_of(unsigned maxabs_bits) {
  const float m = __uint_as_float(maxabs_bits);
  int e = 0;
  frexpf(m, &e);  // m < 2^e (m = 0 -> e = 0)
  e = e < -100 ? -100 : e;
  float s = ldexpf(1.f, 23 - e);
  if (m * s > 8388607.f) s *= 0.5f;
  return s;
}

// Per-channel scales of an image (maxabs: the three channels' largest |v|
// as float bits).
struct QScale {
  float x, y, z;
};
__device__ __forceinline__ QScale qscale_of(const unsigned* maxabs) {
  return QScale{scale_of(maxabs[0]), scale_of(maxabs[1]), scale_of(maxabs[2])};
}
__device__ __forceinline__ float qmax_of(const unsigned* maxabs) {
  return fmaxf(__uint_as_float(maxabs[0]),
               fmaxf(__uint_as_float(maxabs[1]), __uint_as_float(maxabs[2])));
}
__device__ __forceinline__ int quantise(float v, float s) {
  return (int)rintf(v * s);
}
// Mean of a channel from its integer sum: sum / (count * s), in double.
__device__ __forceinline__ float dequantise_mean(long long sum, long long cnt,
                                                 float s) {
  return cnt > 0 ? (float)((double)sum / ((double)cnt * (double)s)) : 0.f;
}

}  // namespace splithip
