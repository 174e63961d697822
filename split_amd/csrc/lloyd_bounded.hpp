// Bounded Lloyd iterations over 16x16 tiles: per-pixel state, the skip
// test and the re-assignment of failing pixels, shared by the single-image
// kernels (lloyd_tiles.hip) and the batched ones (batch.hip).  The method is
// described at the top of lloyd_tiles.hip.
#pragma once

#include <algorithm>

#include "common.hpp"
#include "lloyd_device.hpp"
#include "lloyd_sums.hpp"

namespace splithip {

constexpr int kMaxTileK = 256;  // labels are 8 bits of the per-pixel state

struct TileArgs : Geometry {
  int k;
  float* cbuf;                // two 3k centroid buffers
  long long* totals;          // [4][k]: quantised channel sums, counts
  long long* partials;        // step accumulator copies, [copies][4k]
  float* fcum;                // per-centroid bound consumption F[j], k
  uint32_t* pstate;           // per-pixel state (see pack_state)
  const unsigned* maxabs;     // bits of max |v| of each channel (3)
  LloydState* state;
  Mailbox* mbox;
  // bounded iterations of the single-image path: units (pairs of adjacent
  // tiles) and the multiplier of their spread order
  int64_t nunits, unit_mul;
  // a tile with more failing pixels than this is re-assigned whole
  int whole_fails;
  // neighbour lists of the ball re-assignment, two buffers of k lists
  // (launch s writes buffer s % 2); null: box-pruned re-assignment
  uint32_t* nb = nullptr;
};

// Per-pixel state: label | budget, where budget = slack + F[label] at the
// iteration the pixel was assigned, rounded down, and slack is a lower
// bound on (distance to the second-nearest centroid - distance to the
// nearest).  F[j] sums, over the centroid updates so far, the move of
// centroid j plus the largest move of any other centroid (rounded up).  If
// pixel p took label a when F[a] was F0, then after further updates its
// distance to c_a grew by at most the moves of a and its distance to every
// other centroid shrank by at most the largest other move, so the label is
// provably unchanged while budget - F[a] > margin (Hamerly's bound, with
// the history folded into one value per pixel).  The budget keeps the top
// 24 bits of its float (sign, exponent, 15 mantissa bits); dropping the low
// bits rounds a non-negative value down.
// A float <= a + b (both >= 0): the rounded sum, one ulp lower (rounding
// to nearest is off by at most half an ulp); x >= 0 to float rounded up.
__device__ __forceinline__ float add_down(float a, float b) {
  const float s = a + b;
  return s > 0.f ? __uint_as_float(__float_as_uint(s) - 1u) : 0.f;
}
__device__ __forceinline__ float to_float_up(double x) {
  const float f = (float)x;
  return (double)f < x ? __uint_as_float(__float_as_uint(f) + 1u) : f;
}

__device__ __forceinline__ uint32_t pack_state(int label, float budget) {
  return (__float_as_uint(budget) & 0xFFFFFF00u) | (uint32_t)(label & 0xFF);
}
__device__ __forceinline__ float state_budget(uint32_t s) {
  return __uint_as_float(s & 0xFFFFFF00u);
}

// F[j] after one more centroid update: + the move of centroid j + the
// largest move of any other centroid, rounded up.  d2j = squared move of j;
// m1, m2 = the largest and second-largest squared moves, arg = index of m1.
__device__ __forceinline__ float bound_step(float f, double d2j, int j,
                                            double m1, double m2, int arg) {
  const double other = sqrt(j == arg ? m2 : m1);
  return to_float_up((double)f + (sqrt(d2j) + other) * (1.0 + 1e-5) + 1e-7);
}

// The second half of reassign: given the pixels' colours, new labels and
// slack, writes their state and accumulates (SHARED: into a slice that the
// block's waves share).
template <bool FULL, bool SHARED = false>
__device__ __forceinline__ void commit_assignment(
    const TileArgs& a, const float* fl, QScale qs, int lane,
    const int64_t (&idx)[4], const bool (&v)[4], bool lvec,
    const int (&old)[4], const float (&x)[4], const float (&y)[4],
    const float (&z)[4], const int (&lab)[4], const float (&slack)[4],
    long long* slice, IRun& run) {
  static_assert(!(FULL && SHARED), "the full iteration uses per-wave slices");
  const int k = a.k;
  uint32_t ns[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    ns[i] = pack_state(lab[i],
                       add_down(v[i] ? slack[i] : 0.f, fl[v[i] ? lab[i] : 0]));
  if (lvec) {
    *gmem(reinterpret_cast<U4*>(a.pstate + idx[0])) =
        U4{ns[0], ns[1], ns[2], ns[3]};
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (v[i]) gmem(a.pstate)[idx[i]] = ns[i];
  }
  int qx[4], qy[4], qz[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    qx[i] = quantise(x[i], qs.x);
    qy[i] = quantise(y[i], qs.y);
    qz[i] = quantise(z[i], qs.z);
  }
  if constexpr (FULL) {
    const int l0 = readfirstlane_i(lab[0]);
    const bool same = lab[0] == l0 && lab[1] == l0 && lab[2] == l0 &&
                      lab[3] == l0;
    if (l0 >= 0 && ballot(!same) == 0) {
      if (l0 != run.label || run.tiles >= kRunFlushTiles) {
        irun_flush(run, slice, k, lane);
        run.label = l0;
      }
      run.x += (qx[0] + qx[1]) + (qx[2] + qx[3]);
      run.y += (qy[0] + qy[1]) + (qy[2] + qy[3]);
      run.z += (qz[0] + qz[1]) + (qz[2] + qz[3]);
      run.n += 4;
      run.tiles += 1;
    } else {
      add_by_label(slice, k, lane, lab, qx, qy, qz, 1);
    }
  } else {
    // label changes of a bounded iteration: each changed pixel adds +q to
    // its new and -q to its old label's slice entries with LDS atomics
    // (wave sums per distinct label, used above 16 changes per batch until
    // measured against: 4096^2 k = 32 iteration 64.8 -> 63.3 us, 8192^2
    // k = 256 521 -> 493 us, mostly in the early iterations where many
    // labels change)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (!(v[i] && lab[i] != old[i])) continue;
      unsigned long long* pt =
          reinterpret_cast<unsigned long long*>(slice + lab[i]);
      unsigned long long* pf =
          reinterpret_cast<unsigned long long*>(slice + old[i]);
      atomicAdd(pt, (unsigned long long)(long long)qx[i]);
      atomicAdd(pt + k, (unsigned long long)(long long)qy[i]);
      atomicAdd(pt + 2 * k, (unsigned long long)(long long)qz[i]);
      atomicAdd(pt + 3 * k, 1ull);
      atomicAdd(pf, (unsigned long long)(-(long long)qx[i]));
      atomicAdd(pf + k, (unsigned long long)(-(long long)qy[i]));
      atomicAdd(pf + 2 * k, (unsigned long long)(-(long long)qz[i]));
      atomicAdd(pf + 3 * k, ~0ull);
    }
  }
}

// Re-assigns up to 256 pixels (4 per lane: idx/v) whose previous labels are
// old[] (-1: none), writes their new state, and accumulates: every pixel
// (FULL) or the +/- changes of label (bounded iterations).
template <int CPL, bool FULL, bool PRUNE = true, bool COUNTED = false>
__device__ __forceinline__ void reassign(
    const TileArgs& a, const float (&cx)[CPL], const float (&cy)[CPL],
    const float (&cz)[CPL], const float (&cn)[CPL], const float* fl,
    QScale qs, int lane, const int64_t (&idx)[4], const bool (&v)[4],
    bool lvec, const int (&old)[4], long long* slice, IRun& run) {
  float x[4], y[4], z[4];
  load_colours(a, idx, v, lvec, x, y, z);
  int lab[4];
  float slack[4];
  assign_segment<CPL, true, PRUNE, COUNTED>(cx, cy, cz, cn, a.k, lane, x, y,
                                            z, v, lab, slack);
  commit_assignment<FULL>(a, fl, qs, lane, idx, v, lvec, old, x, y, z,
                          lab, slack, slice, run);
}

// Tile source for walk_tiles: the tiles [s, end) in order.
struct TileRange {
  int64_t s, end;
  __device__ bool operator()(int64_t& t) {
    if (s >= end) return false;
    t = s++;
    return true;
  }
};

constexpr int kQueueCap = 512;      // per-wave queue of failing pixels
// Entries after an image's last pixel state that the walkers' 16-byte lane
// loads may read (and ignore); state buffers are allocated with them.
constexpr int kStatePad = 64;
// More failing pixels: re-assign the tile whole (its compact colour box keeps
// ~7 candidates) rather than queue them (a queued batch evaluates nearly all
// k).  On the 4096² k = 32 image and the 256-image batch: 64 / 96 / 128 /
// 160 / 224 / never -> 71.5 / 70.5 / 71.6 / 69.6 / 70.7 / 70.9 us per
// iteration and 202 / 194 / 193 / 191 / 191 / 196 ms.
constexpr int kWholeTileFails = 160;

// Processes, for one wave, the tiles that next(t) hands out until it
// returns false: FULL re-assigns every pixel (run accumulation), otherwise
// tiles whose pixels all pass the skip test are left alone, tiles with many
// failing pixels are re-assigned whole and the other failing pixels are
// queued and re-assigned in batches of 256.  The state of the next tile is
// loaded while the current one is tested.  On return every change is in
// `slice` (the queue is drained, the run flushed).
template <int CPL, int MODE, bool FULL, class NextTile>
__device__ __forceinline__ void walk_tiles(
    const TileArgs& a, NextTile next, const float (&cx)[CPL],
    const float (&cy)[CPL], const float (&cz)[CPL], const float (&cn)[CPL],
    const float* fl, QScale qs, float margin, long long* slice, int* qidx,
    int* qold, int lane) {
  const int k = a.k;
  IRun run{0, 0, 0, 0, -1, 0};
  int qlen = 0;  // wave-uniform
  // Re-assigns the first min(qlen, 256) queued pixels, then moves the rest
  // to the front.
  auto drain = [&]() {
    const int c = qlen < 256 ? qlen : 256;
    int64_t idx[4];
    bool v[4];
    int old[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = lane + 64 * i;
      v[i] = e < c;
      idx[i] = qidx[v[i] ? e : 0];
      old[i] = v[i] ? qold[e] : -1;
    }
    // queued pixels come from all over the chunk: no box pruning
    reassign<CPL, false, false>(a, cx, cy, cz, cn, fl, qs, lane, idx, v,
                                false, old, slice, run);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = 256 + lane + 64 * i;
      if (e < qlen) {
        qidx[e - 256] = qidx[e];
        qold[e - 256] = qold[e];
      }
    }
    qlen -= c;
  };

  if constexpr (FULL) {
    int64_t s;
    while (next(s)) {
      int64_t idx[4];
      bool v[4], lvec;
      segment_index<MODE>(a, s, lane, idx, v, lvec);
      const int old[4] = {-1, -1, -1, -1};
      reassign<CPL, true>(a, cx, cy, cz, cn, fl, qs, lane, idx, v, lvec,
                          old, slice, run);
    }
  } else {
    static_assert(MODE == kTile || MODE == kTileVec, "bounded walks cover tiles");
    // Tile position (column tx, row ty) of tile s, wave-uniform: advanced by
    // one per tile (the walks take consecutive tiles), divided out only when
    // a walk jumps.  Tile indices of one image fit in int.
    struct Pos {
      int64_t s;
      int tx, ty;
    };
    auto locate = [&](int64_t t, Pos& p) {
      if (t == p.s + 1) {
        p.s = t;
        if (++p.tx == a.tiles_x) {
          p.tx = 0;
          ++p.ty;
        }
      } else {
        p.s = t;
        p.ty = (int)t / a.tiles_x;
        p.tx = (int)t - p.ty * a.tiles_x;
      }
    };
    // One 16-byte load per lane on every path, so that the wait for a
    // tile's state (issued a tile earlier) never covers the next tile's
    // load: lanes below the image read the tile's first row, lanes past its
    // right edge read on into the next row or the kStatePad entries after
    // the last pixel; visit() masks such pixels.
    auto load_state = [&](const Pos& p, bool have, uint32_t (&ps)[4]) {
      const int row = p.ty * tile_h(a) + lane_row(a, lane);
      const int64_t o =
          have ? (int64_t)(row < a.height ? row : p.ty * tile_h(a)) * a.width +
                     p.tx * tile_w(a) + lane_col(a, lane)
               : 0;
      const U4 q = *gmem(reinterpret_cast<const U4*>(a.pstate + o));
      ps[0] = q.x; ps[1] = q.y; ps[2] = q.z; ps[3] = q.w;
    };
    auto visit = [&](const Pos& p, const uint32_t (&ps)[4]) {
      int64_t idx[4];
      bool v[4], lvec;
      int old[4];
      bool fail[4];
      uint64_t fb[4];
      int nfail = 0;
      const int row = p.ty * tile_h(a) + lane_row(a, lane);
      const int col = p.tx * tile_w(a) + lane_col(a, lane);
      const int64_t o = (int64_t)row * a.width + col;
      if ((p.ty + 1) * tile_h(a) <= a.full_rows &&
          (p.tx + 1) * tile_w(a) <= a.width) {
        // every pixel exists: no per-lane guards
        lvec = MODE == kTileVec;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          idx[i] = o + i;
          v[i] = true;
          old[i] = (int)(ps[i] & 0xFFu);
          fail[i] = !(state_budget(ps[i]) - fl[old[i]] > margin);
          fb[i] = ballot(fail[i]);
          nfail += __builtin_popcountll(fb[i]);
        }
      } else {
        const int rp = row_pixels(a, row);
        lvec = MODE == kTileVec && col + 3 < rp;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          idx[i] = o + i;
          v[i] = col + i < rp;
          old[i] = v[i] ? (int)(ps[i] & 0xFFu) : -1;
          fail[i] = v[i] &&
                    !(state_budget(ps[i]) - fl[v[i] ? old[i] : 0] > margin);
          fb[i] = ballot(fail[i]);
          nfail += __builtin_popcountll(fb[i]);
        }
      }
      if (nfail == 0) return;  // every pixel provably unchanged
      if (nfail > kWholeTileFails) {
        reassign<CPL, false>(a, cx, cy, cz, cn, fl, qs, lane, idx, v,
                             lvec, old, slice, run);
        return;
      }
      // queue the failing pixels (in lane order within each slot)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(fb[i] >> 32),
            __builtin_amdgcn_mbcnt_lo((uint32_t)fb[i], 0u));
        if (fail[i]) {
          qidx[qlen + below] = (int)idx[i];
          qold[qlen + below] = old[i];
        }
        qlen += __builtin_popcountll(fb[i]);
      }
      if (qlen >= 256) drain();
    };
    Pos pa{-2, 0, 0};
    int64_t s = 0;
    bool have = next(s);
    if (have) locate(s, pa);
    uint32_t ps[4];
    load_state(pa, have, ps);
    while (have) {
      int64_t s2 = 0;
      const bool have2 = next(s2);
      Pos pb = pa;
      if (have2) locate(s2, pb);
      uint32_t ps2[4];
      load_state(pb, have2, ps2);
      visit(pa, ps);
      pa = pb;
      have = have2;
#pragma unroll
      for (int i = 0; i < 4; ++i) ps[i] = ps2[i];
    }
  }
  if constexpr (!FULL) {
    while (qlen > 0) drain();
  }
  irun_flush(run, slice, k, lane);
}

}  // namespace splithip
