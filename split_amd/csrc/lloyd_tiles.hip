// Bounded Lloyd iterations over 16x16 tiles with exact integer sums.
//
// Used by cluster_image / segment when k <= 256 (any iteration count).
// Results equal plain Lloyd: every label is the dense float argmin (ties to
// the lowest index), and the centroids are exact functions of cluster
// membership.
//
// Per-pixel state (4 bytes, in temp): the label and a "budget" (see
// pack_state in lloyd_bounded.hpp): a pixel provably keeps its label while
// budget - F[label] > margin, where F[j] accumulates, over the centroid
// updates, the move of centroid j plus the largest move of any other
// centroid (Hamerly's bound; triangle inequality on both distances).  margin
// (1e-5 of the data scale) exceeds the float error of the distances, so the
// float argmin is unchanged too.
//
// Iteration 1 assigns every tile.  Later iterations read the state, skip
// tiles whose every pixel passes the test without touching their colours,
// re-assign tiles with many failing pixels whole, and queue the other
// failing pixels for re-assignment in batches of 256 (writing fresh state).
//
// Sums: each channel is quantised once with its own power-of-two scale,
// q = rint(v * s_c) with |q| <= 2^23 - 1 (s_c from the channel's largest
// |v|), and summed as integers into int64 totals in temp: iteration 1 adds
// every pixel, later iterations add +q to the new and -q to the old cluster
// of each pixel whose label changed.  Integer sums are exact, so totals are
// independent of order and of which pixels were skipped, and identical run
// to run.  The quantisation moves a mean by at most max|v_c| * 2^-23 in
// channel c.
#include <algorithm>

#include "common.hpp"
#include "lloyd_device.hpp"
#include "lloyd_sums.hpp"
#include "lloyd_bounded.hpp"
#include "lloyd_ball.hpp"

namespace splithip {

// Buffers across launches.  Launch s (1-based) assigns with centroids c_s
// and bound consumption F_s, and adds its integer sums into accumulator
// A[s % 3] (kAccCopies copies, block b -> copy b % kAccCopies).  Launch
// s >= 2 first finalises step s - 1 in every block: totals_{s-1} =
// totals_{s-2} (T[s % 2]) + A[(s - 1) % 3], c_s from the totals, the shift
// |c_s - c_{s-1}|, F_s and the stop decision; block 0 also publishes them
// (T[(s - 1) % 2], cbuf[s % 2], fcum[s % 2], state, mailbox) and clears
// A[(s + 1) % 3] for launch s + 1.  No block of a launch writes what
// another block of the same launch reads, so one launch per iteration
// needs no fence or second kernel.
constexpr int kAccCopies = 8;

__device__ __forceinline__ long long* totals_buf(const TileArgs& a, int j) {
  return a.totals + (size_t)(j & 1) * 4 * a.k;
}
__device__ __forceinline__ long long* acc_buf(const TileArgs& a, int j) {
  return a.partials + (size_t)(j % 3) * kAccCopies * 4 * a.k;
}

// Finalisation of step s - 1 at the start of launch s >= 2 (see above).
// LDS scratch: tot (4k int64), d2 (k doubles), cs (3k floats: c_s on
// return); fs (k floats) receives F_s, and corr (if not null) the
// neighbour-list correction |c_s[j] - c_{s-1}[j]| + max_i |c_s[i] -
// c_{s-1}[i]|, rounded up.  Returns the stop decision (the same
// in every block), also when an earlier launch already stopped (done_sh:
// an LDS flag).  Every global load of the finalisation is issued before the
// first barrier, so their latencies overlap (one round trip instead of
// four: the fixed cost of a launch, timed without the walk, 6.6 -> 5.9 us;
// the 4096^2 k = 32 iteration 70.6 -> 69.0 us).
__device__ __forceinline__ bool tiles_prologue(const TileArgs& a, int s,
                                               long long* tot, double* d2,
                                               float* cs, float* fs,
                                               int* done_sh, QScale& qs,
                                               float& maxabs,
                                               float* corr = nullptr) {
  const int k = a.k, V = 4 * k;
  const int tid = threadIdx.x;
  LloydState* st = a.state;
  const float* cold = a.cbuf + (size_t)((s - 1) & 1) * 3 * k;
  const float* fold = a.fcum + (size_t)((s - 1) & 1) * k;
  float c_old[3] = {0.f, 0.f, 0.f}, f_old = 0.f;
  if (tid < k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) c_old[c] = cold[c * k + tid];
    f_old = fold[tid];
  }
  const float threshold = st->threshold;
  const int max_iter = st->max_iter;
  const unsigned mb[3] = {a.maxabs[0], a.maxabs[1], a.maxabs[2]};
  if (tid == 0) *done_sh = st->done;
  {
    const long long* tp = totals_buf(a, s - 2);
    const long long* acc = acc_buf(a, s - 1);
    for (int v = tid; v < V; v += blockDim.x) {
      long long x[kAccCopies];
#pragma unroll
      for (int c = 0; c < kAccCopies; ++c) x[c] = acc[(size_t)c * V + v];
      long long t = tp[v];
#pragma unroll
      for (int c = 0; c < kAccCopies; ++c) t += x[c];
      tot[v] = t;
    }
  }
  __syncthreads();
  if (*done_sh) return true;
  qs = QScale{scale_of(mb[0]), scale_of(mb[1]), scale_of(mb[2])};
  maxabs = fmaxf(__uint_as_float(mb[0]),
                 fmaxf(__uint_as_float(mb[1]), __uint_as_float(mb[2])));
  const float sc[3] = {qs.x, qs.y, qs.z};
  if (tid < k) {
    const long long cnt = tot[3 * k + tid];
    double sq = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float nc = dequantise_mean(tot[c * k + tid], cnt, sc[c]);
      const double d = (double)nc - (double)c_old[c];
      sq += d * d;
      cs[c * k + tid] = nc;
    }
    d2[tid] = sq;
  }
  __syncthreads();
  __shared__ double top[2];
  __shared__ int top_arg;
  __shared__ float shift_sh;
  if (tid < kWave) {
    // shift: lane l sums d2[l], d2[l + 64], ...; the 64 partial sums are
    // then added in lane order
    double part = 0.0;
    for (int j = tid; j < k; j += kWave) part += d2[j];
    const uint64_t pb = (uint64_t)__double_as_longlong(part);
    const uint32_t plo = (uint32_t)pb, phi = (uint32_t)(pb >> 32);
    double sum = 0.0;
#pragma unroll
    for (int l = 0; l < kWave; ++l)
      sum += __longlong_as_double(
          (long long)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)phi, l)
                       << 32) |
                      (uint32_t)__builtin_amdgcn_readlane((int)plo, l)));
    // largest and second-largest squared move; arg = lowest index of the
    // largest, m2 = largest over the other indices
    double m1 = -1.0;  // moves are >= 0: the lane's first index is taken
    int arg = 0x7fffffff;
    for (int j = tid; j < k; j += kWave)
      if (d2[j] > m1) {
        m1 = d2[j];
        arg = j;
      }
    double g1 = m1;
    for (int o = kWave / 2; o > 0; o >>= 1) g1 = fmax(g1, __shfl_xor(g1, o));
    int ga = m1 == g1 ? arg : 0x7fffffff;
    for (int o = kWave / 2; o > 0; o >>= 1) ga = min(ga, __shfl_xor(ga, o));
    double m2 = 0.0;
    for (int j = tid; j < k; j += kWave)
      if (j != ga) m2 = fmax(m2, d2[j]);
    for (int o = kWave / 2; o > 0; o >>= 1) m2 = fmax(m2, __shfl_xor(m2, o));
    if (tid == 0) {
      shift_sh = (float)sqrt(sum);
      top[0] = g1;
      top[1] = m2;
      top_arg = ga;
    }
  }
  __syncthreads();
  if (tid < k) {
    fs[tid] = bound_step(f_old, d2[tid], tid, top[0], top[1], top_arg);
    if (corr)
      corr[tid] = to_float_up((sqrt(d2[tid]) + sqrt(top[0])) * (1.0 + 1e-9));
  }
  const float shift = shift_sh;
  const int it = s - 1;
  const bool done = !(shift > threshold) || it >= max_iter;
  if (blockIdx.x == 0) {
    long long* tn = totals_buf(a, s - 1);
    for (int v = tid; v < V; v += blockDim.x) tn[v] = tot[v];
    long long* an = acc_buf(a, s + 1);
    for (int v = tid; v < kAccCopies * V; v += blockDim.x) an[v] = 0;
    float* cn = a.cbuf + (size_t)(s & 1) * 3 * k;
    for (int v = tid; v < 3 * k; v += blockDim.x) cn[v] = cs[v];
    if (tid < k) a.fcum[(size_t)(s & 1) * k + tid] = fs[tid];
    if (tid == 0) {
      st->iter = it;
      st->shift = shift;
      st->cur = s & 1;
      st->done = done;
      if (a.mbox)  // one 16-byte store; the host polls it between chunks
        *reinterpret_cast<int4*>(a.mbox) =
            make_int4(it, done, __float_as_int(shift), 0);
    }
  }
  return done;
}

// ------------------------------------------------ bounded iterations ---
//
// Units are pairs of horizontally adjacent tiles, ranked in a spread order:
// rank r is unit r * mul mod nunits (units row-major, mul coprime to
// nunits), so that consecutive ranks land far apart in the image.  Tiles
// that fail the skip test come in large patches; spreading every block's
// units over the whole image evens out the work per block.  A block owns
// the ranks [r0, r1); per round of up to kBlockUnits units it lists them in
// LDS as {offset of the first pixel, first row} and its waves take them
// through an LDS counter.  Pixel offsets are int32 (the tiled path
// requires n < 2^31).
constexpr int kBlockUnits = 2048;
constexpr int kMaxQueueWaves = 16;  // waves per block (1024 threads)

// One tile: offset of its first pixel, first row and column.
struct TileGeo {
  int off, row0, col0;
};

// Fills list[0, cnt) with the units of ranks q0 .. q0 + cnt - 1.
__device__ __forceinline__ void list_units(const TileArgs& a, int64_t q0,
                                           int cnt, int2* list) {
  const int upr = (a.tiles_x + 1) / 2;
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
    const int64_t u = (q0 + i) * a.unit_mul % a.nunits;
    const int ty = (int)(u / upr);
    const int c = (int)(u - (int64_t)ty * upr);
    const int row0 = ty * tile_h(a);
    list[i] = make_int2(row0 * a.width + 2 * c * tile_w(a), row0);
  }
}

// BALL: failing pixels are re-assigned by neighbour-ball walks
// (assign_ball) instead of the box-pruned assign_segment, and the waves
// share one slice.
template <int CPL, int MODE, bool BALL = false>
struct BoundedWalker {
  static constexpr bool kBall = BALL;
  const TileArgs& a;
  const float (&cx)[CPL];
  const float (&cy)[CPL];
  const float (&cz)[CPL];
  const float (&cn)[CPL];
  const float* fl;
  QScale qs;
  float margin;
  long long* slice;
  int* qidx;
  int* qold;
  int lane, lane_off;  // lane l: row lane_row, columns lane_col..+3 of a tile
  IRun run;
  int qlen;            // wave-uniform
  // BALL: centroids {x, y, z, -}, neighbour lists (nbm entries each) and
  // their correction, all in LDS
  const float4* cent4 = nullptr;
  const uint32_t* nbl = nullptr;
  int nbm = 0;
  const float* corr = nullptr;

  // Re-assigns up to 256 pixels (idx/v, previous labels old); PRUNE: the
  // box-pruned assignment (whole tiles), else every centroid (queued
  // batches, whose pixels come from all over the image).
  template <bool PRUNE = true>
  __device__ __forceinline__ void assign(const int64_t (&idx)[4],
                                         const bool (&v)[4], bool lvec,
                                         const int (&old)[4]) {
    if constexpr (BALL) {
      float x[4], y[4], z[4];
      load_colours(a, idx, v, lvec, x, y, z);
      int lab[4];
      float slack[4];
      assign_ball(cent4, nbl, nbm, corr, margin, a.k, x, y, z, v, old, lab,
                  slack);
      commit_assignment<false, true>(a, fl, qs, lane, idx, v, lvec, old, x,
                                     y, z, lab, slack, slice, run);
    } else {
      reassign<CPL, false, PRUNE, true>(a, cx, cy, cz, cn, fl, qs, lane, idx,
                                        v, lvec, old, slice, run);
    }
  }

  __device__ __forceinline__ void tile_pixels(const TileGeo& g,
                                              int64_t (&idx)[4], bool (&v)[4],
                                              bool& lvec) const {
    const int col = g.col0 + lane_col(a, lane);
    const int rp = row_pixels(a, g.row0 + lane_row(a, lane));
    const int o = g.off + lane_off;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = col + i < rp;
      idx[i] = o + i;
    }
    lvec = MODE == kTileVec && col + 3 < rp;
  }

  // Tiles wholly inside the image (wave-uniform: the geometry is scalar).
  __device__ __forceinline__ bool inner(const TileGeo& g) const {
    return g.row0 + tile_h(a) <= a.full_rows && g.col0 + tile_w(a) <= a.width;
  }

  // One 16-byte load per lane on every path, so that the wait for this
  // tile's state (issued a tile earlier) never has to cover the next tile's
  // load as well.  Lanes below the image read the tile's first row, lanes
  // past its right edge read on into the next row or the padding after the
  // last pixel (fewer than a tile width, kStatePad); visit() masks such
  // pixels.
  __device__ __forceinline__ void load_state(const TileGeo& g, bool have,
                                             uint32_t (&ps)[4]) const {
    const bool rowok = g.row0 + lane_row(a, lane) < a.height;
    const int o = have ? g.off + (rowok ? lane_off : lane_col(a, lane)) : 0;
    const U4 q = *reinterpret_cast<const U4*>(a.pstate + o);
    ps[0] = q.x; ps[1] = q.y; ps[2] = q.z; ps[3] = q.w;
  }

  // Re-assigns the first min(qlen, 256) queued pixels, then moves the rest
  // to the front.
  __device__ __forceinline__ void drain() {
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
    assign<false>(idx, v, false, old);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = 256 + lane + 64 * i;
      if (e < qlen) {
        qidx[e - 256] = qidx[e];
        qold[e - 256] = qold[e];
      }
    }
    qlen -= c;
  }

  // Skip test of one tile; re-assigns it whole or queues its failing pixels.
  // The geometry is re-read as wave-uniform values, so that the interior
  // test is a scalar branch (the compiler kept it in vector registers).
  __device__ __forceinline__ void visit(const TileGeo& gv,
                                        const uint32_t (&ps)[4]) {
    const TileGeo g{readfirstlane_i(gv.off), readfirstlane_i(gv.row0),
                    readfirstlane_i(gv.col0)};
    const int col = g.col0 + lane_col(a, lane);
    int old[4];
    bool fail[4];
    uint64_t fb[4];
    int nfail = 0;
    if (inner(g)) {  // every pixel exists: no per-lane guards
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        old[i] = (int)(ps[i] & 0xFFu);
        fail[i] = !(state_budget(ps[i]) - fl[old[i]] > margin);
        fb[i] = ballot(fail[i]);
        nfail += __builtin_popcountll(fb[i]);
      }
    } else {
      const int rp = row_pixels(a, g.row0 + lane_row(a, lane));
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bool v = col + i < rp;
        old[i] = (int)(ps[i] & 0xFFu);
        fail[i] = v && !(state_budget(ps[i]) - fl[v ? old[i] : 0] > margin);
        fb[i] = ballot(fail[i]);
        nfail += __builtin_popcountll(fb[i]);
      }
    }
    if (nfail == 0) return;  // every pixel provably unchanged
    if (nfail > a.whole_fails) {
      int64_t idx[4];
      bool v[4], lvec;
      tile_pixels(g, idx, v, lvec);
#pragma unroll
      for (int i = 0; i < 4; ++i) old[i] = v[i] ? old[i] : -1;
      assign(idx, v, lvec, old);
      return;
    }
    // queue the failing pixels (in lane order within each slot)
    const int o = g.off + lane_off;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t below = __builtin_amdgcn_mbcnt_hi(
          (uint32_t)(fb[i] >> 32),
          __builtin_amdgcn_mbcnt_lo((uint32_t)fb[i], 0u));
      if (fail[i]) {
        qidx[qlen + below] = o + i;
        qold[qlen + below] = old[i];
      }
      qlen += __builtin_popcountll(fb[i]);
    }
    if (qlen >= 256) drain();
  }

  // Walks the units list[0, cnt), taken through *counter (zero on entry),
  // loading the state of the next tile while the current one is tested.
  __device__ __forceinline__ void walk(const int2* list, int cnt,
                                       unsigned* counter) {
    int pend_off = 0, pend_row0 = 0, pend_col0 = 0;
    bool pending = false;
    auto next = [&](TileGeo& g) {
      if (pending) {
        pending = false;
        g = TileGeo{pend_off, pend_row0, pend_col0};
        return true;
      }
      unsigned i = 0u;
      if (lane == 0) i = atomicAdd(counter, 1u);
      const int u = readfirstlane_i((int)i);
      if (u >= cnt) return false;
      const int2 e = list[u];
      // the entry is the same for every lane: keep the geometry scalar
      const int off = readfirstlane_i(e.x), row0 = readfirstlane_i(e.y);
      const int col0 = off - row0 * a.width;
      g = TileGeo{off, row0, col0};
      if (col0 + tile_w(a) < a.width) {
        pending = true;
        pend_off = off + tile_w(a);
        pend_row0 = row0;
        pend_col0 = col0 + tile_w(a);
      }
      return true;
    };
    TileGeo g{0, 0, 0};
    bool have = next(g);
    uint32_t ps[4];
    load_state(g, have, ps);
    while (have) {
      TileGeo g2{0, 0, 0};
      const bool have2 = next(g2);
      uint32_t ps2[4];
      load_state(g2, have2, ps2);
      visit(g, ps);
      g = g2;
      have = have2;
#pragma unroll
      for (int i = 0; i < 4; ++i) ps[i] = ps2[i];
    }
  }

  // After the walk: the waves' queues (fewer than 256 pixels each) are
  // re-assigned as one list, their concatenation in wave order, in batches
  // of 256 handed out through an LDS counter.  queues = the first wave's
  // queue (each wave's spans 2 * kQueueCap ints: indices, then labels).
  __device__ __forceinline__ void finish_block(int* queues, int wave,
                                               int nwaves) {
    __shared__ int start[kMaxQueueWaves + 1];
    __shared__ unsigned next_batch;
    if (lane == 0) start[wave + 1] = qlen;
    if (threadIdx.x == 0) {
      start[0] = 0;
      next_batch = 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 1; w <= nwaves; ++w) start[w] += start[w - 1];
    __syncthreads();
    const int total = start[nwaves];
    for (;;) {
      unsigned b = 0u;
      if (lane == 0) b = atomicAdd(&next_batch, 1u);
      const int e0 = 256 * readfirstlane_i((int)b);
      if (e0 >= total) break;
      int64_t idx[4];
      bool v[4];
      int old[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int pos = e0 + lane + 64 * i;
        v[i] = pos < total;
        int w = 0;
        for (int j = 1; j < nwaves; ++j) w += pos >= start[j] ? 1 : 0;
        const int* q = queues + w * 2 * kQueueCap;
        const int e = v[i] ? pos - start[w] : 0;
        idx[i] = v[i] ? q[e] : 0;
        old[i] = v[i] ? q[kQueueCap + e] : -1;
      }
      assign(idx, v, false, old);
    }
    qlen = 0;
  }
};


// The bounded walk of one block: its ranks [r0, r1) of the spread unit
// order, listed kBlockUnits at a time, then the waves' queue remainders.
template <class Walker>
__device__ __forceinline__ void walk_block(const TileArgs& a, Walker& w,
                                           int2* list, int* queues, int wave,
                                           int nwaves) {
  __shared__ unsigned unit_next;
  const int64_t G = gridDim.x;
  const int64_t r0 = a.nunits * blockIdx.x / G;
  const int64_t r1 = a.nunits * (blockIdx.x + 1) / G;
  for (int64_t q0 = r0; q0 < r1; q0 += kBlockUnits) {
    const int cnt = (int)std::min<int64_t>(kBlockUnits, r1 - q0);
    list_units(a, q0, cnt, list);
    if (threadIdx.x == 0) unit_next = 0u;
    __syncthreads();
    w.walk(list, cnt, &unit_next);
    __syncthreads();  // the list is refilled by the next round
  }
  // The queue remainders: box-pruned waves drain their own as soon as their
  // walk ends (no barrier: 67.8 -> 65.5 us per iteration on the 4096^2
  // k = 32 image, where the pooled remainders waited for the block's
  // slowest walker); the ball walks pool them (a partial batch costs its
  // longest walk: 521 -> 581 us per iteration on the 8192^2 k = 256 image
  // with their own drains).
  if constexpr (Walker::kBall) {
    w.finish_block(queues, wave, nwaves);
  } else {
    while (w.qlen > 0) w.drain();
  }
}

// The block's share of step s, summed over its nslices slices, is added
// (device atomics) into copy blockIdx % kAccCopies of accumulator A[s % 3];
// the next launch's prologue adds the copies into the totals.
__device__ __forceinline__ void add_block_sums(const TileArgs& a, int s,
                                               const long long* slices,
                                               int nslices) {
  const int V = 4 * a.k;
  __syncthreads();
  long long* acc = acc_buf(a, s) + (size_t)(blockIdx.x % kAccCopies) * V;
  for (int vv = threadIdx.x; vv < V; vv += blockDim.x) {
    long long tot = 0;
    for (int w = 0; w < nslices; ++w) tot += slices[(size_t)w * V + vv];
    if (tot != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(acc + vv),
                (unsigned long long)tot);
  }
}

// Neighbour lists of centroids c (of launch s) for launch s + 1, when the
// run uses them: list j by wave j / G of block j % G.
template <class Cent>
__device__ __forceinline__ void build_lists(const TileArgs& a, int s, Cent c,
                                            int wave, int nwaves, int lane) {
  if (a.nb == nullptr) return;
  const int k = a.k, m = nb_len(k);
  uint32_t* out = a.nb + (size_t)(s & 1) * k * m;
  for (int j = wave * (int)gridDim.x + (int)blockIdx.x; j < k;
       j += nwaves * (int)gridDim.x)
    build_neighbours(c, k, j, m, out + (size_t)j * m, lane);
}

// Which of the two kernels enqueued for a bounded launch s of a run with
// neighbour lists runs the step (the other returns after the
// finalisation): the ball walk when the largest list correction is below
// 0.85 of the mean nearest-neighbour distance in the lists, else the
// box-pruned step.  A walk visits its label's list until the entries are
// farther than the correction plus the pixel's distance to its centroid;
// with corrections near the spacing of the centroids it reaches the end of
// the list and rescans all k.  On the 8192^2 k = 256 image the ratio is
// 1.74 / 0.80 / 0.46 / 0.34 in launches 2 / 3 / 4 / 5, where the box step
// takes 2.9 / 2.8 / 2.6 / 2.6 ms and the walk 8.8 / 4.0 / 1.9 / 1.6 ms; with
// the branch-free walk steps launch 3 walks too (0.72 ms less per run than
// at a 0.6 threshold).
// Every block of both kernels computes it from the same data (corr from the
// finalisation, the lists of launch s - 1) in the same order of operations
// (wave 0: per-lane sums in index order, then a fixed DPP reduction; no
// atomics, whose order would vary), so every block reaches the same
// decision and exactly one kernel runs the step.  Block-uniform; red: 2
// floats of LDS.
__device__ __forceinline__ bool prefer_ball(const TileArgs& a, int s,
                                            const float* corr, float* red) {
  const int k = a.k, m = nb_len(k);
  const uint32_t* nb = a.nb + (size_t)((s - 1) & 1) * k * m;
  if (threadIdx.x < kWave) {
    float mx = 0.f, sum = 0.f;
    for (int j = threadIdx.x; j < k; j += kWave) {
      mx = fmaxf(mx, corr[j]);
      sum += nb_dist(nb[(size_t)j * m]);
    }
    mx = wave_allreduce(mx, OpMax());
    sum = wave_allreduce(sum, OpAdd());
    if (threadIdx.x == 0) {
      red[0] = mx;
      red[1] = sum;
    }
  }
  __syncthreads();
  const bool ball = red[0] < 0.85f * (red[1] / (float)k);
  __syncthreads();  // red is free again
  return ball;
}

// Launch s of a tiled run (FULL: s == 1, every pixel assigned).
// k <= 64: 1024-thread blocks at 4 waves per SIMD (<= 128 VGPRs).  Larger
// k: 256-thread blocks at 2 waves per SIMD, whose 2 or 4 centroids per lane
// need up to 186 VGPRs (capped at 128 they spilled 29 / 59 VGPRs to scratch;
// on the 8192^2 k = 256 image 1,756 -> 1,723 us per iteration without).
template <int CPL, int MODE, bool FULL>
__global__ __launch_bounds__(CPL >= 2 ? 256 : 1024, CPL >= 2 ? 2 : 1) void tiles_step(TileArgs a, int s) {
  // 16-byte aligned: the float4 tables in it are read with ds_read_b128
  // (the static LDS variables before it had left it 16-aligned by chance;
  // four more bytes of them made every centroid read of the walks
  // unaligned: 410 -> 762 us per 8192^2 k = 256 iteration)
  extern __shared__ __align__(16) long long lds64[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  const int k = a.k;
  // done only ever goes 0 -> 1, and a block that sees it set by block 0 of
  // this launch would have reached the same decision; read once per block so
  // that the whole block returns (or not) together
  __shared__ int done_sh;
  if (s == 1) {
    if (threadIdx.x == 0) done_sh = a.state->done;
    __syncthreads();
    if (done_sh) return;
  }

  const int V = 4 * k;
  long long* slices = lds64;                                   // nwaves x V
  float* fl = reinterpret_cast<float*>(lds64 + nwaves * V);    // F[0..k)
  int* queues = reinterpret_cast<int*>(fl + kMaxTileK);
  int* qidx = queues + wave * 2 * kQueueCap;
  int* qold = qidx + kQueueCap;
  int2* list = reinterpret_cast<int2*>(queues + nwaves * 2 * kQueueCap);
  // this launch's centroids c_s (and F_s in fl): the initial ones for s = 1,
  // else finalised here from step s - 1 (scratch in the unit-list area,
  // 8 KiB + 2 KiB + 3 KiB at k = 256 < 16 KiB)
  const float* cent;
  QScale qs;
  float maxabs;
  if (s == 1) {
    maxabs = qmax_of(a.maxabs);
    qs = qscale_of(a.maxabs);
    cent = a.cbuf + 3 * k;  // cbuf[1] holds c_1, fcum[1] F_1 = 0
    for (int i = threadIdx.x; i < k; i += blockDim.x) fl[i] = a.fcum[k + i];
  } else {
    long long* tot = reinterpret_cast<long long*>(list);
    double* d2 = reinterpret_cast<double*>(tot + 4 * kMaxTileK);
    float* cs = reinterpret_cast<float*>(d2 + kMaxTileK);
    float* corr = cs + 3 * kMaxTileK;  // 14 KiB of the 16 KiB list area
    if (tiles_prologue(a, s, tot, d2, cs, fl, &done_sh, qs, maxabs,
                       a.nb ? corr : nullptr))
      return;
    if (a.nb) {
      __shared__ float red[2];
      __syncthreads();  // corr is written
      if (prefer_ball(a, s, corr, red)) return;  // tiles_step_ball runs it
    }
    cent = cs;
  }
  const float margin = 1e-5f * (1.f + 4.f * maxabs);
  for (int i = threadIdx.x; i < nwaves * V; i += blockDim.x) slices[i] = 0;

  float cx[CPL], cy[CPL], cz[CPL], cn[CPL];
#pragma unroll
  for (int t = 0; t < CPL; ++t) {
    const int j = lane + 64 * t;
    const bool live = j < k;
    cx[t] = live ? cent[j] : 0.f;
    cy[t] = live ? cent[k + j] : 0.f;
    cz[t] = live ? cent[2 * k + j] : 0.f;
    cn[t] = fmaf(cz[t], cz[t], fmaf(cy[t], cy[t], cx[t] * cx[t]));
  }
  // lists for launch s + 1 (a run with neighbour lists)
  build_lists(a, s, CentSoA{cent, k}, wave, nwaves, lane);
  __syncthreads();  // cent (LDS for s >= 2) is read; the list area is free
  long long* slice = slices + (size_t)wave * V;

  if constexpr (FULL) {
    // every tile costs the same: contiguous tiles per wave, so that runs of
    // one label accumulate in registers
    const int64_t gw = (int64_t)blockIdx.x * nwaves + wave;
    const int64_t W = (int64_t)gridDim.x * nwaves;
    walk_tiles<CPL, MODE, FULL>(
        a, TileRange{a.nseg * gw / W, a.nseg * (gw + 1) / W}, cx, cy, cz, cn,
        fl, qs, margin, slice, qidx, qold, lane);
  } else {
    BoundedWalker<CPL, MODE> w{a,     cx,    cy,   cz,   cn,
                               fl,    qs,    margin, slice, qidx,
                               qold,  lane,  lane_row(a, lane) * a.width + lane_col(a, lane),
                               IRun{0, 0, 0, 0, -1, 0}, 0};
    walk_block(a, w, list, queues, wave, nwaves);
  }
  add_block_sums(a, s, slices, nwaves);
}

// Bounded launch s >= 2 of a run with neighbour lists (k > 64): failing
// pixels re-assigned by neighbour-ball walks.  1024-thread blocks, one per
// CU; no centroids in registers, so the waves share one slice of the step
// sums (LDS atomics) and the lists fit: at k = 256, 4 KiB of centroids,
// 8 KiB slice, 32 KiB of lists, 64 KiB of queues, 16 KiB of units.
template <int MODE>
__global__ __launch_bounds__(1024, 1) void tiles_step_ball(TileArgs a, int s) {
  // Failing pixels are gathered from the three planes, not the interleaved
  // copy: ~40% of a tile's pixels fail here, so the planes' lines are read
  // anyway (8192^2 k = 256 iteration 405.8 -> 399.1 us; the box-pruned
  // walker's sparser failures stay on the interleaved copy, one line per
  // pixel: 4096^2 k = 32 63.4-64.1 against 64.6-65.2 us on the planes)
  a.aos = nullptr;
  // 16-byte aligned: the float4 tables in it are read with ds_read_b128
  // (the static LDS variables before it had left it 16-aligned by chance;
  // four more bytes of them made every centroid read of the walks
  // unaligned: 410 -> 762 us per 8192^2 k = 256 iteration)
  extern __shared__ __align__(16) long long lds64[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  const int k = a.k, V = 4 * k, m = nb_len(k);
  float4* cent4 = reinterpret_cast<float4*>(lds64);                    // kMaxTileK
  long long* slice = reinterpret_cast<long long*>(cent4 + kMaxTileK);  // V
  float* fl = reinterpret_cast<float*>(slice + V);                     // kMaxTileK
  float* corr = fl + kMaxTileK;                                        // kMaxTileK
  uint32_t* nbl = reinterpret_cast<uint32_t*>(corr + kMaxTileK);       // k m, even
  int* queues = reinterpret_cast<int*>(nbl + ((k * m + 1) & ~1));
  int* qidx = queues + wave * 2 * kQueueCap;
  int* qold = qidx + kQueueCap;
  int2* list = reinterpret_cast<int2*>(queues + nwaves * 2 * kQueueCap);
  __shared__ int done_sh;
  QScale qs;
  float maxabs;
  {
    long long* tot = reinterpret_cast<long long*>(list);
    double* d2 = reinterpret_cast<double*>(tot + 4 * kMaxTileK);
    float* cs = reinterpret_cast<float*>(d2 + kMaxTileK);
    if (tiles_prologue(a, s, tot, d2, cs, fl, &done_sh, qs, maxabs, corr))
      return;
    __shared__ float red[2];
    __syncthreads();  // corr is written
    if (!prefer_ball(a, s, corr, red)) return;  // tiles_step ran it
    for (int j = threadIdx.x; j < k; j += blockDim.x)
      cent4[j] = make_float4(cs[j], cs[k + j], cs[2 * k + j], 0.f);
  }
  const uint32_t* nb_in = a.nb + (size_t)((s - 1) & 1) * k * m;
  for (int i = threadIdx.x; i < k * m; i += blockDim.x) nbl[i] = nb_in[i];
  for (int i = threadIdx.x; i < V; i += blockDim.x) slice[i] = 0;
  __syncthreads();  // centroids, lists, correction and slice are set
  build_lists(a, s, CentAoS{cent4}, wave, nwaves, lane);

  const float margin = 1e-5f * (1.f + 4.f * maxabs);
  const float none[1] = {0.f};
  BoundedWalker<1, MODE, true> w{a,     none,  none, none,   none,
                                 fl,    qs,    margin, slice, qidx,
                                 qold,  lane,  lane_row(a, lane) * a.width + lane_col(a, lane),
                                 IRun{0, 0, 0, 0, -1, 0}, 0};
  w.cent4 = cent4;
  w.nbl = nbl;
  w.nbm = m;
  w.corr = corr;
  walk_block(a, w, list, queues, wave, nwaves);
  add_block_sums(a, s, slice, 1);
}

// c_1 (the initial centroids) into cbuf[1], zero totals, accumulators and
// F (both buffers).
__global__ void tiles_init(LloydState* st, float* cbuf, const float* cent,
                           int k, int max_iter, float threshold,
                           long long* totals, float* fcum, unsigned* maxabs) {
  for (int i = threadIdx.x; i < 3 * k; i += blockDim.x)
    cbuf[3 * k + i] = cent[i];
  for (int i = threadIdx.x; i < 4 * k * (2 + 3 * kAccCopies); i += blockDim.x)
    totals[i] = 0;  // two totals buffers, then the three accumulators
  for (int i = threadIdx.x; i < 2 * k; i += blockDim.x) fcum[i] = 0.f;
  if (threadIdx.x == 0) {
    st->cur = 1;
    st->iter = 0;
    st->done = max_iter <= 0;
    st->max_iter = max_iter;
    st->threshold = threshold;
    st->shift = 0.f;
    st->next = 0u;
  }
  if (threadIdx.x < 3) maxabs[threadIdx.x] = 0u;
}

// Largest |v| of each of the three planes of n floats (plane c at pts +
// c n) as float bits into out[c] (non-negative floats order like their bit
// patterns).  blockIdx.y is the plane; a plane's 16-byte-aligned middle is
// read as float4, four loads in flight per thread.
__global__ __launch_bounds__(256) void planes_absmax(const float* pts,
                                                     int64_t n,
                                                     unsigned* out) {
  const float* p = pts + (int64_t)blockIdx.y * n;
  // [0, head): elements before the first 16-byte boundary (all of them
  // when p is not 4-byte aligned), then whole float4s, then the tail
  // [tail0, n)
  const bool al = ((uintptr_t)p & 3u) == 0;
  const int64_t head =
      al ? std::min<int64_t>(
               n, (int64_t)(((16u - ((uintptr_t)p & 15u)) & 15u) / 4u))
         : n;
  const float4* q4 = reinterpret_cast<const float4*>(p + head);
  const int64_t chunks = (n - head) / 4;
  const int64_t tail0 = head + 4 * chunks;
  float m = 0.f;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t c = t;
  for (; c + 3 * stride < chunks; c += 4 * stride) {
    float4 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = q4[c + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      m = fmaxf(m, fmaxf(fmaxf(fabsf(q[u].x), fabsf(q[u].y)),
                         fmaxf(fabsf(q[u].z), fabsf(q[u].w))));
  }
  for (; c < chunks; c += stride) {
    const float4 q = q4[c];
    m = fmaxf(m, fmaxf(fmaxf(fabsf(q.x), fabsf(q.y)),
                       fmaxf(fabsf(q.z), fabsf(q.w))));
  }
  for (int64_t i = t; i < head; i += stride) m = fmaxf(m, fabsf(p[i]));
  for (int64_t i = tail0 + t; i < n; i += stride) m = fmaxf(m, fabsf(p[i]));
  // one atomic per block (many blocks updating one address serialise)
  __shared__ float part[256 / kWave];
  m = wave_allreduce(m, OpMax());
  if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)blockDim.x / kWave; ++w) m = fmaxf(m, part[w]);
    atomicMax(out + blockIdx.y, __float_as_uint(m));
  }
}

// planes_absmax plus the interleaved copy of the pixels ({x, y, z, 0}) that
// the queued pixels are read from.
__global__ __launch_bounds__(256) void tiles_absmax_pack(const float* pts,
                                                         int64_t n, F4* aos,
                                                         unsigned* out) {
  // one pixel per lane and step: each plane load reads 256 contiguous bytes
  // and each store writes 1 KiB, four steps in flight
  float mx = 0.f, my = 0.f, mz = 0.f;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    float x[4], y[4], z[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t j = i + u * stride;
      x[u] = pts[j];
      y[u] = pts[n + j];
      z[u] = pts[2 * n + j];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      aos[i + u * stride] = F4{x[u], y[u], z[u], 0.f};
      mx = fmaxf(mx, fabsf(x[u]));
      my = fmaxf(my, fabsf(y[u]));
      mz = fmaxf(mz, fabsf(z[u]));
    }
  }
  for (; i < n; i += stride) {
    const float x = pts[i], y = pts[n + i], z = pts[2 * n + i];
    aos[i] = F4{x, y, z, 0.f};
    mx = fmaxf(mx, fabsf(x));
    my = fmaxf(my, fabsf(y));
    mz = fmaxf(mz, fabsf(z));
  }
  __shared__ float part[3][256 / kWave];
  mx = wave_allreduce(mx, OpMax());
  my = wave_allreduce(my, OpMax());
  mz = wave_allreduce(mz, OpMax());
  if ((threadIdx.x & (kWave - 1)) == 0) {
    part[0][threadIdx.x / kWave] = mx;
    part[1][threadIdx.x / kWave] = my;
    part[2][threadIdx.x / kWave] = mz;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    float m = part[threadIdx.x][0];
    for (int w = 1; w < (int)blockDim.x / kWave; ++w)
      m = fmaxf(m, part[threadIdx.x][w]);
    atomicMax(out + threadIdx.x, __float_as_uint(m));
  }
}

__global__ void tiles_labels(const LloydState* st, const uint32_t* pstate,
                             int32_t* labels, int64_t n) {
  if (st->iter == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    labels[i] = (int32_t)(pstate[i] & 0xFFu);
}

__global__ void tiles_finish(const LloydState* st, const float* cbuf,
                             float* cent, int k) {
  const float* src = cbuf + (size_t)st->cur * 3 * k;
  for (int i = threadIdx.x; i < 3 * k; i += blockDim.x) cent[i] = src[i];
}

// ------------------------------------------------------------ launchers ---

// One block per CU where the LDS allows (k <= 128: up to 129 KB): the
// waves of a CU then share one pool of tiles in the bounded iterations and
// finish together.
static int tiles_threads(int k) { return k <= 64 ? 1024 : 256; }
// Bounded iterations re-assign by neighbour-ball walks above this k (on
// the 8192^2 k = 256 image a queued batch's colour box keeps ~230 of 256
// centroids, its longest walk ~16).
constexpr int kBallMinK = 65;
// Step launches use at most this many blocks.
static int tiles_max_blocks(int k) { return k <= 64 ? 512 : 1024; }
// Units (pairs of adjacent tiles) of n pixels in rows of width.
static int64_t tile_units(int64_t n, int width, int tw_log) {
  if (width <= 0) return 0;
  const int64_t tiles_x = tiles_across(width, tw_log);
  const int64_t rows = (n + width - 1) / width;
  return (tiles_x + 1) / 2 * tiles_down(rows, tw_log);
}

struct TileLayout {
  LloydState* st;
  float* cbuf;
  long long* totals;
  long long* partials;
  float* fcum;
  unsigned* maxabs;
  uint32_t* pstate;
  F4* aos;
  uint32_t* nb;
  size_t bytes;
};

static TileLayout tiles_layout(void* temp, int64_t n, int k) {
  TileLayout L{};
  size_t o = 0;
  char* t = static_cast<char*>(temp);
  auto take = [&](size_t bytes) {
    char* p = t ? t + o : nullptr;
    o += align_up(bytes, 256);
    return p;
  };
  L.st = reinterpret_cast<LloydState*>(take(sizeof(LloydState)));
  L.cbuf = reinterpret_cast<float*>(take(6 * (size_t)k * sizeof(float)));
  // two totals buffers and three accumulators of kAccCopies copies each,
  // contiguous
  L.totals = reinterpret_cast<long long*>(
      take((size_t)(2 + 3 * kAccCopies) * 4 * k * 8));
  L.partials = L.totals ? L.totals + 2 * 4 * (size_t)k : nullptr;
  L.fcum = reinterpret_cast<float*>(take(2 * (size_t)k * sizeof(float)));
  L.maxabs = reinterpret_cast<unsigned*>(take(3 * sizeof(unsigned)));
  L.nb = reinterpret_cast<uint32_t*>(
      take(2 * (size_t)k * nb_len(k) * sizeof(uint32_t)));
  L.pstate = reinterpret_cast<uint32_t*>(
      take(((size_t)n + kStatePad) * sizeof(uint32_t)));
  L.aos = reinterpret_cast<F4*>(take((size_t)n * sizeof(F4)));
  L.bytes = o;
  return L;
}

// A multiplier m coprime to n near n * (golden ratio - 1): r -> r * m mod n
// is a permutation of [0, n) that sends consecutive r far apart.
static int64_t spread_multiplier(int64_t n) {
  if (n <= 2) return 1;
  int64_t m = std::max<int64_t>(1, (int64_t)((double)n * 0.6180339887));
  auto gcd = [](int64_t x, int64_t y) {
    while (y) {
      const int64_t t = x % y;
      x = y;
      y = t;
    }
    return x;
  };
  while (gcd(m, n) != 1) ++m;
  return m;
}

int launch_absmax(const float* pts, int64_t n, unsigned* out,
                  hipStream_t stream) {
  const int64_t per_plane = std::max<int64_t>(
      1, std::min<int64_t>((n / 4 + 1023) / 1024, 3 * device_cu_count()));
  hipLaunchKernelGGL(planes_absmax, dim3((unsigned)per_plane, 3), dim3(256), 0,
                     stream, pts, n, out);
  SPLIT_HIP_CHECK_LAUNCH("planes_absmax");
  return SPLIT_HIP_OK;
}

size_t tiles_workspace(int64_t n, int k) {
  return tiles_layout(nullptr, n, k).bytes;
}

bool tiles_supported(int64_t n, int width, int k) {
  // pixel indices are listed as int32
  return width > 0 && k >= 1 && k <= kMaxTileK && n <= 0x7fffffff;
}

template <bool FULL, int MODE>
static void tiles_launch_mode(const TileArgs& a, int s, int grid,
                              int threads, size_t lds, hipStream_t stream,
                              bool ball = false) {
  if (ball) {
    hipLaunchKernelGGL(tiles_step_ball<MODE>, dim3(grid), dim3(threads), lds,
                       stream, a, s);
    return;
  }
  switch (a.k <= 64 ? 1 : a.k <= 128 ? 2 : 4) {
    case 1:
      hipLaunchKernelGGL((tiles_step<1, MODE, FULL>), dim3(grid),
                         dim3(threads), lds, stream, a, s);
      break;
    case 2:
      hipLaunchKernelGGL((tiles_step<2, MODE, FULL>), dim3(grid),
                         dim3(threads), lds, stream, a, s);
      break;
    default:
      hipLaunchKernelGGL((tiles_step<4, MODE, FULL>), dim3(grid),
                         dim3(threads), lds, stream, a, s);
      break;
  }
}

int cluster_tiles_async(const float* pts, int64_t n, int width,
                        float* centroids, int k, int32_t* labels, void* temp,
                        size_t temp_bytes, int max_iter, float threshold,
                        hipStream_t stream, Mailbox** mbox_out) {
  // rows of width pixels; the last one may be partial (a point set walked
  // as a width-16 image)
  const int height = (int)((n + width - 1) / width);
  const TileLayout L = tiles_layout(temp, n, k);
  if (temp_bytes < L.bytes)
    return fail(SPLIT_HIP_ERR_TEMP,
                "temp holds %zu bytes, tiled cluster needs %zu", temp_bytes,
                L.bytes);
  DriverRes* res = nullptr;
  SPLIT_HIP_TRY(driver_res(&res));
  res->host->iter = 0;
  res->host->done = max_iter <= 0;
  res->host->shift = 0.f;

  hipLaunchKernelGGL(tiles_init, dim3(1), dim3(256), 0, stream, L.st, L.cbuf,
                     centroids, k, max_iter, threshold, L.totals, L.fcum,
                     L.maxabs);
  // at least 4096 pixels per block: each block ends with three device
  // atomics on the same three words, which serialise (a 512^2 image took
  // 26.5 us with 2048 blocks of 128 pixels)
  const int64_t pack_blocks = std::max<int64_t>(
      1, std::min<int64_t>((n + 4095) / 4096, 8 * (int64_t)device_cu_count()));
  hipLaunchKernelGGL(tiles_absmax_pack, dim3((unsigned)pack_blocks), dim3(256),
                     0, stream, pts, n, L.aos, L.maxabs);
  SPLIT_HIP_CHECK_LAUNCH("tiles_init/tiles_absmax_pack");

  TileArgs a{};
  a.pts = pts;
  a.n = n;
  a.width = width;
  a.height = height;
  a.full_rows = (int)(n / width);
  // 32x8 tiles for k <= 64 on images at least 32 wide, else 16x16: on the
  // 4096^2 k = 32 image 69.3 / 68.4 / 70.2 us per iteration with 16x16 /
  // 32x8 / 64x4 tiles, on the 8192^2 k = 256 one 1.72 / 1.80 / 1.93 ms
  // (the wider tiles' colour boxes keep more of 256 candidates)
  a.tw_log = k <= 64 && width >= 32 ? 5 : 4;
  a.tiles_x = (int)tiles_across(width, a.tw_log);
  a.nseg = (int64_t)a.tiles_x * tiles_down(height, a.tw_log);
  a.k = k;
  a.cbuf = L.cbuf;
  a.totals = L.totals;
  a.partials = L.partials;
  a.fcum = L.fcum;
  a.pstate = L.pstate;
  a.aos = L.aos;
  a.maxabs = L.maxabs;
  a.state = L.st;
  a.mbox = res->dev;
  a.nunits = tile_units(n, width, a.tw_log);
  // measured on the 8192^2 k=256 image (every failing tile whole: 2.49 ms
  // per iteration; at 16 / 64 / 160 failing pixels: 2.34 / 1.83 / 1.76 ms)
  // and the 4096^2 k=32 one (77.7 / 73.4 / 71.0 / 70.1 us)
  a.whole_fails = kWholeTileFails;
  a.unit_mul = spread_multiplier(a.nunits);
  a.nb = k >= kBallMinK ? L.nb : nullptr;

  const int threads = tiles_threads(k);
  const int nwaves = threads / kWave;
  const int64_t want = (a.nseg + nwaves - 1) / nwaves;
  const size_t lds = (size_t)nwaves * 4 * k * 8 + kMaxTileK * sizeof(float) +
                     (size_t)nwaves * 2 * kQueueCap * sizeof(int) +
                     kBlockUnits * sizeof(int2);
  // resident blocks per CU of the bounded kernel (its register use is the
  // larger of the two), so that the grid is one wave of blocks
  int per_cu = 1;
  {
    const void* fn = k <= 64    ? (const void*)tiles_step<1, kTileVec, false>
                     : k <= 128 ? (const void*)tiles_step<2, kTileVec, false>
                                : (const void*)tiles_step<4, kTileVec, false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = 1;
  }
  const int grid = (int)std::max<int64_t>(
      1, std::min<int64_t>({want, (int64_t)device_cu_count() * per_cu,
                            (int64_t)tiles_max_blocks(k)}));
  // the ball kernel of the bounded launches: 1024 threads, one block per CU
  const int ball_threads = 1024, ball_waves = ball_threads / kWave;
  const size_t ball_lds =
      kMaxTileK * sizeof(float4) + (size_t)4 * k * 8 +
      2 * kMaxTileK * sizeof(float) +
      (size_t)((k * nb_len(k) + 1) & ~1) * sizeof(uint32_t) +
      (size_t)ball_waves * 2 * kQueueCap * sizeof(int) +
      kBlockUnits * sizeof(int2);
  const int ball_grid = (int)std::max<int64_t>(
      1, std::min<int64_t>((a.nseg + ball_waves - 1) / ball_waves,
                           (int64_t)device_cu_count()));
  // a lane's four pixels move as one dword-aligned 16-byte vector (any width)
  const bool vecmode = aligned4(pts) && aligned4(L.pstate);
  // launch s = 1 .. max_iter + 1: launch s finalises step s - 1 and runs
  // step s; the last one only finalises (it stops at the iteration limit)
  auto launch = [&](int s) {
    if (s == 1 && vecmode)
      tiles_launch_mode<true, kTileVec>(a, s, grid, threads, lds, stream);
    else if (s == 1)
      tiles_launch_mode<true, kTile>(a, s, grid, threads, lds, stream);
    else if (vecmode)
      tiles_launch_mode<false, kTileVec>(a, s, grid, threads, lds, stream);
    else
      tiles_launch_mode<false, kTile>(a, s, grid, threads, lds, stream);
    // with neighbour lists a bounded step is enqueued twice, box-pruned then
    // ball walk; prefer_ball lets exactly one of them run it
    if (s > 1 && a.nb && vecmode)
      tiles_launch_mode<false, kTileVec>(a, s, ball_grid, ball_threads,
                                         ball_lds, stream, true);
    else if (s > 1 && a.nb)
      tiles_launch_mode<false, kTile>(a, s, ball_grid, ball_threads, ball_lds,
                                      stream, true);
  };

  // Launches are enqueued in chunks, two chunks in flight, and the host
  // stops enqueueing once the mailbox says done.  A run that cannot stop
  // early (threshold < 0) is enqueued at once, small images in chunks of
  // 16 (their launches take ~15 us); launches after the stop return at
  // once.
  const int total = max_iter > 0 ? max_iter + 1 : 0;
  const int kChunk = threshold < 0.f ? std::max(total, 1)
                     : n < ((int64_t)1 << 22) ? 16 : 4;
  int launched = 0, chunk = 0;
  while (launched < total) {
    const int cnt = std::min(kChunk, total - launched);
    for (int i = 0; i < cnt; ++i) launch(launched + i + 1);
    SPLIT_HIP_CHECK_LAUNCH("tiles_step");
    launched += cnt;
    SPLIT_HIP_CHECK(hipEventRecord(res->ev[chunk & 1], stream));
    if (chunk > 0) {
      SPLIT_HIP_CHECK(hipEventSynchronize(res->ev[(chunk - 1) & 1]));
      if (reinterpret_cast<volatile Mailbox*>(res->host)->done) break;
    }
    ++chunk;
  }
  if (labels) {
    hipLaunchKernelGGL(tiles_labels, dim3(4 * device_cu_count()), dim3(256),
                       0, stream, L.st, L.pstate, labels, n);
  }
  hipLaunchKernelGGL(tiles_finish, dim3(1), dim3(256), 0, stream, L.st,
                     L.cbuf, centroids, k);
  SPLIT_HIP_CHECK_LAUNCH("tiles_labels/tiles_finish");
  *mbox_out = res->host;
  return SPLIT_HIP_OK;
}

}  // namespace splithip
