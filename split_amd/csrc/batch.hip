// Batched segmentation: many images of uneven size and k (<= 64) through
// one captured hipGraph of persistent launches.
//
// The images' 16x16 Lloyd tiles are numbered consecutively (image i owns
// global tiles [tile0_i, tile0_{i+1})).  Images are split into two groups
// by k (kBoundedMinK; by default every image is bounded), each walked by its
// own persistent kernel once per Lloyd iteration:
//  * plain images: every iteration assigns every pixel of every tile
//    (box-pruned exact assignment) and accumulates; nothing is stored per
//    pixel, the labels come from one assignment pass at the end;
//  * bounded images: the single-image bounded iteration
//    (lloyd_bounded.hpp) — the first iteration records each pixel's label
//    and slack, later ones skip the pixels whose bound proves the label
//    unchanged and re-assign the rest; the labels come from that state.
// A wave walks a contiguous run of its group's tiles, switching image
// (centroids, k, scale, bound table) where the run crosses an image
// boundary.  Sums are
// exact integers (quantised as in the single-image tiled path), so every
// image gets exactly the centroids, labels and stopping iteration it gets on
// its own.  The wave that completes an image's last tile of the iteration
// (per-image tile counter) finalises that image: new centroids, shift, bound
// table, stop flag.  Images finish independently; later launches skip
// their tiles.
//
// After the iterations: the labels (from the per-pixel state) and
// connected components with the single-image passes (ccl_device.hpp) over
// every image's 64x64 tiles in one grid: tile pass, seam merge, root jumps,
// a per-image scan of the root bitmap, and the final per-tile resolution
// that writes each pixel's root or dense rank.
//
// The whole sequence is captured once into a hipGraph per (temp, batch
// shape, settings) and replayed; per-image buffer pointers live in a
// descriptor table in temp that is uploaded before each replay.
// An autotuner picks the Lloyd tile shape (16x16, 32x8 or 64x4) and the
// step kernels' block size and blocks per CU by timing iterations of each
// candidate on the batch itself, cached per batch signature.
#include <algorithm>
#include <cstddef>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "ccl_device.hpp"
#include "common.hpp"
#include "lloyd_device.hpp"
#include "lloyd_sums.hpp"
#include "lloyd_bounded.hpp"

namespace splithip {

constexpr int kBatchMaxK = 64;

struct BatchImg {
  // set by the host
  const float* pts;
  int32_t* labels;
  int32_t* segments;
  const float* centroids_in;
  float* centroids_out;
  int W, H, k, tiles_x;  // 16x16 Lloyd tiles
  int ctiles_x, pad0;    // 64x64 CC tiles
  int64_t tile0;         // first global Lloyd tile
  int64_t cost0;         // the group's tile cost before this image
  int64_t ctile0;        // first global CC tile
  int64_t px0;           // first global pixel (multiple of 64)
  float* cbuf;           // 2 x 3k
  long long* totals;     // 4k
  unsigned* bitmap;      // 2*ceil(n/64) words: local roots, then roots
  unsigned* linked;      // 2*ceil(n/64) words: roots linked under another
  unsigned* prefix;      // ceil(n/32) words
  int32_t* edge;         // 4 x (CC tiles x 64): the tiles' edge columns
  uint32_t* pstate;      // per-pixel Lloyd state (lloyd_bounded.hpp), n
  float* fcum;           // per-centroid bound consumption, k
  // device state
  int cur, iter, done, nseg;
  float shift;
  unsigned maxabs[3];  // largest |v| of each channel (float bits)
  int tiles_done;
  int bounded;  // iterations after the first are bounded (keep the totals)
  int tw_log;   // Lloyd tile shape (Geometry::tw_log), the same in a batch
};

// One chunk of a bounded iteration's work: tiles [s, s + tiles) (global
// tile numbers) of image img of the group.
struct ChunkEnt {
  int img;
  int tiles;
  int64_t s;
};
struct ChunkTable {
  int n;     // entries this iteration
  int pad[63];
  ChunkEnt e[1];  // kMaxChunks entries follow the header
};

struct BatchArgs {
  BatchImg* img;
  ChunkTable* chunks;  // per bounded iteration, built by batch_chunks
  int count;
  int64_t tiles;   // total Lloyd tiles
  int64_t ctiles;  // total CC tiles
  int64_t px;      // total (padded) pixels
  int max_iter;
  float threshold;
};

// The images of one Lloyd kernel: img[0..count) of the table, whose tiles
// are the global tiles [t0, t1).
struct Group {
  BatchImg* img;
  int count;
  int64_t t0, t1;
  int64_t c0 = 0, c1 = 0;  // tile cost of the group: [c0, c1)
};

// Estimated cost of a Lloyd tile of an image with k centroids in the full
// first iteration (the box keeps more candidates at larger k), in units of
// ~7 VALU instructions: the first iteration splits the group's cost, not its
// tiles, evenly over the waves.
__host__ __device__ __forceinline__ int64_t tile_cost(int k) {
  return 32 + k;
}

// Images with k >= kBoundedMinK take bounded iterations.  They pay for
// reading and writing the per-pixel state and for queueing, and win at every
// k on the shaded workloads (batch_256: 253 ms with every image bounded,
// 270 / 276 / 287 ms with plain iterations below k = 8 / 16 / 32; the
// results are identical), so every image is bounded; the plain group stays
// for a threshold above 1.
constexpr int kBoundedMinK = 1;

// Image owning global index g of the field at byte offset OFF (tile0,
// ctile0 or px0): the last image whose start is <= g.
template <int64_t BatchImg::*F>
__device__ __forceinline__ int find_image(const BatchImg* img, int count,
                                          int64_t g) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (img[mid].*F <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// The fields of one image the pixel-space kernels use, held in registers
// while a wave walks that image.
struct PxImg {
  int32_t* labels;
  int32_t* segments;
  unsigned* bitmap;
  unsigned* linked;
  unsigned* prefix;
  const uint32_t* pstate;
  int64_t n, px0;
  int W, H, iter, bounded;
};

__device__ __forceinline__ PxImg px_image(const BatchImg& m) {
  return PxImg{m.labels, m.segments, m.bitmap, m.linked, m.prefix, m.pstate,
               (int64_t)m.W * m.H, m.px0, m.W, m.H, m.iter, m.bounded};
}

// A wave's contiguous share of the padded global pixel space in 64-pixel
// units (each image starts at a multiple of 64, so a unit lies in one
// image): calls f(m, q0) with q0 the unit's first pixel in image m (lane l
// owns pixel q0 + l, which exists while q0 + l < m.n).  The image is found
// by one binary search, then by walking forward; a grid-stride loop paid a
// binary search (8 dependent table loads) per 64 pixels.
template <class F>
__device__ __forceinline__ void for_each_unit(const BatchArgs& a, F&& f) {
  const int64_t units = a.px >> 6;
  const int wpb = (int)(blockDim.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * wpb;
  const int64_t gw = (int64_t)blockIdx.x * wpb +
                     __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t u0 = units * gw / nw, u1 = units * (gw + 1) / nw;
  if (u0 >= u1) return;
  int i = find_image<&BatchImg::px0>(a.img, a.count, u0 << 6);
  PxImg m = px_image(a.img[i]);
  int64_t next = i + 1 < a.count ? a.img[i + 1].px0 : INT64_MAX;
  for (int64_t u = u0; u < u1; ++u) {
    const int64_t g0 = u << 6;
    if (g0 >= next) {
      while (g0 >= next) {
        ++i;
        next = i + 1 < a.count ? a.img[i + 1].px0 : INT64_MAX;
      }
      m = px_image(a.img[i]);
    }
    f(m, g0 - m.px0);
  }
}

__device__ __forceinline__ Geometry image_geometry(const BatchImg& m) {
  Geometry g{};
  g.pts = m.pts;
  g.n = (int64_t)m.W * m.H;
  g.width = m.W;
  g.height = m.H;
  g.full_rows = m.H;
  g.tw_log = m.tw_log;
  g.tiles_x = m.tiles_x;
  g.nseg = (int64_t)m.tiles_x * tiles_down(m.H, g.tw_log);
  return g;
}

__global__ void batch_init(BatchArgs a) {
  BatchImg& m = a.img[blockIdx.x];
  const int k = m.k;
  for (int i = threadIdx.x; i < 3 * k; i += blockDim.x) m.cbuf[i] = m.centroids_in[i];
  for (int i = threadIdx.x; i < 4 * k; i += blockDim.x) m.totals[i] = 0;
  if (threadIdx.x == 0) {
    m.cur = 0;
    m.iter = 0;
    m.done = a.max_iter <= 0;
    m.nseg = -1;
    m.shift = 0.f;
    m.tiles_done = 0;
  }
  if (threadIdx.x < 3) m.maxabs[threadIdx.x] = 0u;
  for (int i = threadIdx.x; i < k; i += blockDim.x) m.fcum[i] = 0.f;
}

// Largest |v| of each channel per image.  A wave takes a contiguous run of
// the batch's Lloyd tiles and keeps running maxima while the run stays in
// one image, so an image receives one atomic per channel and wave that
// touches it.
__global__ __launch_bounds__(256) void batch_absmax(BatchArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / kWave;
  const int64_t s1 = a.tiles * (gw + 1) / nw;
  int64_t s = a.tiles * gw / nw;
  while (s < s1) {
    const int i = find_image<&BatchImg::tile0>(a.img, a.count, s);
    const BatchImg& m = a.img[i];
    const Geometry g = image_geometry(m);
    const int64_t end = std::min<int64_t>(s1, m.tile0 + g.nseg);
    const bool vec = aligned4(g.pts);  // dword-aligned 16-byte loads
    float mx = 0.f, my = 0.f, mz = 0.f;
    for (; s < end; ++s) {
      int64_t idx[4];
      bool v[4], lvec;
      segment_index<kTileVec>(g, s - m.tile0, lane, idx, v, lvec);
      lvec = lvec && vec;
      float x[4], y[4], z[4];
      load_colours(g, idx, v, lvec, x, y, z);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        mx = v[j] ? fmaxf(mx, fabsf(x[j])) : mx;
        my = v[j] ? fmaxf(my, fabsf(y[j])) : my;
        mz = v[j] ? fmaxf(mz, fabsf(z[j])) : mz;
      }
    }
    mx = wave_allreduce(mx, OpMax());
    my = wave_allreduce(my, OpMax());
    mz = wave_allreduce(mz, OpMax());
    if (lane == 0) {
      atomicMax(&a.img[i].maxabs[0], __float_as_uint(mx));
      atomicMax(&a.img[i].maxabs[1], __float_as_uint(my));
      atomicMax(&a.img[i].maxabs[2], __float_as_uint(mz));
    }
  }
}

// Per-wave LDS: accumulation slice [4][kBatchMaxK] of int64 and 64 doubles
// of finalisation scratch; the bounded kernel adds the bound table and the
// queue of failing pixels.
struct WaveLdsPlain {
  long long slice[4 * kBatchMaxK];
  double scratch[kWave];
};
struct WaveLds {
  long long slice[4 * kBatchMaxK];
  double scratch[kWave];
  float fl[kBatchMaxK];
  int qidx[kQueueCap];
  int qold[kQueueCap];
};

// Flushes the wave's slice for image m into its totals, then counts the
// wave's tiles; the wave that completes the image's tiles finalises it.
template <class L>
__device__ void batch_finish_segment(const BatchArgs& a, BatchImg& m,
                                     int ntiles, L& w, int lane) {
  const int k = m.k;
  for (int v = lane; v < 4 * k; v += kWave) {
    const long long t = w.slice[v];
    if (t != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(&m.totals[v]),
                (unsigned long long)t);
    w.slice[v] = 0;
  }
  __threadfence();
  int last = 0;
  if (lane == 0) {
    const int img_tiles = (int)(image_geometry(m).nseg);
    const int before = atomicAdd(&m.tiles_done, ntiles);
    last = before + ntiles == img_tiles;
  }
  last = readfirstlane_i(readlane_i(last, 0));
  if (!last) return;
  __threadfence();
  // this wave finalises image i for this iteration
  const QScale qs = qscale_of(m.maxabs);
  const float sc[3] = {qs.x, qs.y, qs.z};
  const float* cent = m.cbuf + (size_t)m.cur * 3 * k;
  float* cent_out = m.cbuf + (size_t)(1 - m.cur) * 3 * k;
  double d2 = 0.0;
  if (lane < k) {
    const long long cnt = __hip_atomic_load(&m.totals[3 * k + lane],
                                            __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long sum = __hip_atomic_load(&m.totals[c * k + lane],
                                              __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
      const float nc = dequantise_mean(sum, cnt, sc[c]);
      const double d = (double)nc - (double)cent[c * k + lane];
      d2 += d * d;
      cent_out[c * k + lane] = nc;
    }
  }
  w.scratch[lane] = d2;
  // bounded iterations update the totals by label changes; iterations that
  // re-assign every pixel start from zero
  if (!m.bounded)
    for (int v = lane; v < 4 * k; v += kWave) m.totals[v] = 0;
  __builtin_amdgcn_wave_barrier();
  double m1 = 0.0, m2 = 0.0;  // largest and second-largest squared moves
  int arg = -1;
  for (int j = 0; j < k; ++j) {
    const double d = w.scratch[j];
    if (d > m1) {
      m2 = m1;
      m1 = d;
      arg = j;
    } else if (d > m2) {
      m2 = d;
    }
  }
  if (m.bounded && lane < k)
    m.fcum[lane] = bound_step(m.fcum[lane], d2, lane, m1, m2, arg);
  __threadfence();
  if (lane == 0) {
    double s = 0.0;
    for (int j = 0; j < kWave; ++j) s += w.scratch[j];
    const float shift = (float)sqrt(s);
    const int it = m.iter + 1;
    m.shift = shift;
    m.iter = it;
    m.cur = 1 - m.cur;
    m.done = !(shift > a.threshold) || it >= a.max_iter;
    m.tiles_done = 0;
  }
}

// One Lloyd iteration of the plain images: assign every pixel (exact
// box-pruned argmin), accumulate quantised sums.
__global__ __launch_bounds__(1024) void batch_step_plain(BatchArgs a, Group g) {
  extern __shared__ WaveLdsPlain plds[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  WaveLdsPlain& w = plds[wave];
  for (int v = lane; v < 4 * kBatchMaxK; v += kWave) w.slice[v] = 0;

  const int64_t gw = (int64_t)blockIdx.x * nwaves + wave;
  const int64_t W = (int64_t)gridDim.x * nwaves;
  const int64_t n = g.t1 - g.t0;
  const int64_t s0 = g.t0 + n * gw / W;
  const int64_t s1 = g.t0 + n * (gw + 1) / W;

  int64_t s = s0;
  while (s < s1) {
    const int i = find_image<&BatchImg::tile0>(g.img, g.count, s);
    BatchImg& m = g.img[i];
    const Geometry geo = image_geometry(m);
    const int64_t end = std::min<int64_t>(s1, m.tile0 + geo.nseg);
    if (m.done) {
      s = end;
      continue;
    }
    const int k = m.k;
    const float* cent = m.cbuf + (size_t)m.cur * 3 * k;
    float cx[1], cy[1], cz[1], cn[1];
    {
      const bool live = lane < k;
      cx[0] = live ? cent[lane] : 0.f;
      cy[0] = live ? cent[k + lane] : 0.f;
      cz[0] = live ? cent[2 * k + lane] : 0.f;
      cn[0] = fmaf(cz[0], cz[0], fmaf(cy[0], cy[0], cx[0] * cx[0]));
    }
    const QScale qs = qscale_of(m.maxabs);
    const bool vec = aligned4(geo.pts);  // dword-aligned 16-byte loads
    IRun run{0, 0, 0, 0, -1, 0};
    for (int64_t t = s; t < end; ++t) {
      int64_t idx[4];
      bool v[4], lvec;
      segment_index<kTileVec>(geo, t - m.tile0, lane, idx, v, lvec);
      lvec = lvec && vec;
      float x[4], y[4], z[4];
      load_colours(geo, idx, v, lvec, x, y, z);
      int lab[4];
      assign_segment<1, false>(cx, cy, cz, cn, k, lane, x, y, z, v, lab,
                               nullptr);
      int qx[4], qy[4], qz[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        qx[j] = quantise(x[j], qs.x);
        qy[j] = quantise(y[j], qs.y);
        qz[j] = quantise(z[j], qs.z);
      }
      const int l0 = readfirstlane_i(lab[0]);
      const bool same = lab[0] == l0 && lab[1] == l0 && lab[2] == l0 &&
                        lab[3] == l0;
      if (l0 >= 0 && ballot(!same) == 0) {
        if (l0 != run.label || run.tiles >= kRunFlushTiles) {
          irun_flush(run, w.slice, k, lane);
          run.label = l0;
        }
        run.x += (qx[0] + qx[1]) + (qx[2] + qx[3]);
        run.y += (qy[0] + qy[1]) + (qy[2] + qy[3]);
        run.z += (qz[0] + qz[1]) + (qz[2] + qz[3]);
        run.n += 4;
        run.tiles += 1;
      } else {
        add_by_label(w.slice, k, lane, lab, qx, qy, qz, 1);
      }
    }
    irun_flush(run, w.slice, k, lane);
    batch_finish_segment(a, m, (int)(end - s), w, lane);
    s = end;
  }
}

// Tiles per chunk of the bounded iterations' round-robin deal: smaller
// chunks balance the waves better, each chunk end costs a partial queue
// drain and an image flush (batch_256: 224 / 218 / 202 / 207 / 233 ms and
// the warm-start set: 121 / 109 / 107-111 / 108 / 111 ms at 64 / 96 / 128 /
// 192 / 256 tiles while every chunk was 128 tiles).  With few running
// images (the tail of a batch, or a batch of small images) 128-tile chunks
// leave most waves idle, so each iteration's chunks are sized to the tiles
// still running: about one chunk per wave, between kMinChunkTiles and
// kBalanceTiles tiles.
constexpr int64_t kBalanceTiles = 128;
constexpr int64_t kMinChunkTiles = 16;
constexpr int kMaxChunks = 32768;  // chunk table entries

// The tiles [s, s1) of a group (possibly spanning images) for one wave.
template <bool FIRST>
__device__ __forceinline__ void bounded_run(const BatchArgs& a, const Group& g, WaveLds& w,
                            int lane, int64_t s, int64_t s1, int known = -1) {
  while (s < s1) {
    const int i =
        known >= 0 ? known : find_image<&BatchImg::tile0>(g.img, g.count, s);
    known = -1;
    BatchImg& m = g.img[i];
    TileArgs ta{};
    static_cast<Geometry&>(ta) = image_geometry(m);
    const int64_t end = std::min<int64_t>(s1, m.tile0 + ta.nseg);
    if (m.done) {
      s = end;
      continue;
    }
    const int k = m.k;
    ta.k = k;
    ta.pstate = m.pstate;
    const float* cent = m.cbuf + (size_t)m.cur * 3 * k;
    float cx[1], cy[1], cz[1], cn[1];
    {
      const bool live = lane < k;
      cx[0] = live ? cent[lane] : 0.f;
      cy[0] = live ? cent[k + lane] : 0.f;
      cz[0] = live ? cent[2 * k + lane] : 0.f;
      cn[0] = fmaf(cz[0], cz[0], fmaf(cy[0], cy[0], cx[0] * cx[0]));
    }
    const float maxabs = qmax_of(m.maxabs);
    const QScale qs = qscale_of(m.maxabs);
    const float margin = 1e-5f * (1.f + 4.f * maxabs);
    if (lane < k) w.fl[lane] = m.fcum[lane];
    __builtin_amdgcn_wave_barrier();
    // dword-aligned 16-byte loads and stores: any width
    const bool vec = aligned4(ta.pts) && aligned4(ta.pstate);
    const int64_t t0 = s - m.tile0, t1 = end - m.tile0;
    if (vec)
      walk_tiles<1, kTileVec, FIRST>(ta, TileRange{t0, t1}, cx, cy, cz,
                                     cn, w.fl, qs, margin, w.slice,
                                     w.qidx, w.qold, lane);
    else
      walk_tiles<1, kTile, FIRST>(ta, TileRange{t0, t1}, cx, cy, cz, cn,
                                  w.fl, qs, margin, w.slice, w.qidx,
                                  w.qold, lane);
    batch_finish_segment(a, m, (int)(end - s), w, lane);
    s = end;
  }
}

// One Lloyd iteration of the bounded images.  FIRST: assign every pixel and
// record its state; otherwise the bounded iteration.  At least 4 waves per
// SIMD (<= 128 VGPRs): the bounded iteration otherwise took 144 and two
// 512-thread blocks no longer fitted on a CU (C4 set: 97 -> 93 ms).
template <bool FIRST>
__global__ __launch_bounds__(512, 4) void batch_step_bounded(BatchArgs a, Group g) {
  extern __shared__ WaveLds wlds[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  WaveLds& w = wlds[wave];
  for (int v = lane; v < 4 * kBatchMaxK; v += kWave) w.slice[v] = 0;

  const int64_t gw = (int64_t)blockIdx.x * nwaves + wave;
  const int64_t W = (int64_t)gridDim.x * nwaves;
  if constexpr (FIRST) {
    // one contiguous run per wave, of equal estimated cost: wave gw takes
    // the tiles whose cost interval starts in its share of [c0, c1) (with
    // equal tile counts the waves on k = 64 images set the launch's time)
    const int64_t C = g.c1 - g.c0;
    auto tile_at = [&](int64_t c) {
      const int i = find_image<&BatchImg::cost0>(g.img, g.count, c);
      const int64_t wt = tile_cost(g.img[i].k);
      return g.img[i].tile0 + (c - g.img[i].cost0 + wt - 1) / wt;
    };
    bounded_run<FIRST>(a, g, w, lane, tile_at(g.c0 + C * gw / W),
                       tile_at(g.c0 + C * (gw + 1) / W));
  } else {
    // failing tiles come in patches: the running images' chunks (the table
    // batch_chunks built for this iteration) are dealt out round-robin, so
    // that every wave sees every part of the images
    const int nc = a.chunks->n;
    if (nc >= 0) {
      for (int64_t c = gw; c < nc; c += W) {
        const ChunkEnt e = a.chunks->e[c];
        bounded_run<FIRST>(a, g, w, lane, e.s, e.s + e.tiles, e.img);
      }
    } else {  // no table (too many images): fixed chunks over every tile
      for (int64_t c = g.t0 + gw * kBalanceTiles; c < g.t1;
           c += W * kBalanceTiles)
        bounded_run<FIRST>(a, g, w, lane, c,
                           std::min<int64_t>(g.t1, c + kBalanceTiles));
    }
  }
}

// The chunk table of one bounded iteration (one block): the group's running
// images (not done), their tiles cut into chunks of
// clamp(ceil(running tiles / target), kMinChunkTiles, kBalanceTiles) tiles
// (more if the table would overflow), entries in image order.
__global__ __launch_bounds__(1024) void batch_chunks(BatchArgs a, Group g,
                                                     int target) {
  __shared__ long long wsum[1024 / kWave];
  __shared__ long long cs_sh;
  __shared__ long long run_sh;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & (kWave - 1), wv = tid / kWave, nw = nt / kWave;
  if (g.count > kMaxChunks / 2) {  // the table could overflow: no table
    if (tid == 0) a.chunks->n = -1;
    return;
  }
  auto img_tiles = [&](int i) -> int64_t {
    const BatchImg& m = g.img[i];
    return m.done ? 0 : (int64_t)m.tiles_x * tiles_down(m.H, m.tw_log);
  };
  // inclusive wave scan by shuffles
  auto wave_scan = [&](long long x) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const long long y = __shfl_up(x, o);
      x += lane >= o ? y : 0;
    }
    return x;
  };
  long long tot = 0;
  for (int i = tid; i < g.count; i += nt) tot += img_tiles(i);
  tot = wave_scan(tot);
  if (lane == kWave - 1) wsum[wv] = tot;
  __syncthreads();
  if (tid == 0) {
    long long T = 0;
    for (int w = 0; w < nw; ++w) T += wsum[w];
    long long cs = (T + target - 1) / target;
    cs = cs < kMinChunkTiles ? kMinChunkTiles : cs > kBalanceTiles ? kBalanceTiles : cs;
    // entries <= T / cs + images <= kMaxChunks (images <= kMaxChunks / 2)
    const long long room = kMaxChunks - g.count;
    if (T / cs > room) cs = (T + room - 1) / room;
    cs_sh = cs;
    run_sh = 0;
  }
  __syncthreads();
  const long long cs = cs_sh;
  // images in rounds of nt: chunk counts, their exclusive prefix, entries
  for (int base = 0; base < g.count; base += nt) {
    const int i = base + tid;
    const int64_t ti = i < g.count ? img_tiles(i) : 0;
    const long long ci = (ti + cs - 1) / cs;
    const long long inc = wave_scan(ci);
    if (lane == kWave - 1) wsum[wv] = inc;
    __syncthreads();
    long long before = run_sh;
    for (int w = 0; w < wv; ++w) before += wsum[w];
    const long long first = before + inc - ci;
    if (ti > 0) {
      const int64_t t0 = g.img[i].tile0;
      for (long long j = 0; j < ci; ++j) {
        ChunkEnt e;
        e.img = i;
        e.s = t0 + j * cs;
        e.tiles = (int)std::min<int64_t>(cs, ti - j * cs);
        a.chunks->e[first + j] = e;
      }
    }
    __syncthreads();
    if (tid == 0)
      for (int w = 0; w < nw; ++w) run_sh += wsum[w];
    __syncthreads();
  }
  if (tid == 0) a.chunks->n = (int)run_sh;
}

// Labels of the plain images: the assignment of the last iteration, i.e.
// against the previous centroid buffer (images that ran an iteration).
__global__ __launch_bounds__(256) void batch_labels_plain(Group g) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t s = g.t0 + gw; s < g.t1; s += nw) {
    const int i = find_image<&BatchImg::tile0>(g.img, g.count, s);
    const BatchImg& m = g.img[i];
    if (m.iter == 0) continue;
    const Geometry geo = image_geometry(m);
    const int k = m.k;
    const float* cent = m.cbuf + (size_t)(1 - m.cur) * 3 * k;
    float cx[1], cy[1], cz[1], cn[1];
    const bool live = lane < k;
    cx[0] = live ? cent[lane] : 0.f;
    cy[0] = live ? cent[k + lane] : 0.f;
    cz[0] = live ? cent[2 * k + lane] : 0.f;
    cn[0] = fmaf(cz[0], cz[0], fmaf(cy[0], cy[0], cx[0] * cx[0]));
    int64_t idx[4];
    bool v[4], lvec;
    segment_index<kTileVec>(geo, s - m.tile0, lane, idx, v, lvec);
    lvec = lvec && aligned4(geo.pts) && aligned4(m.labels);
    float x[4], y[4], z[4];
    load_colours(geo, idx, v, lvec, x, y, z);
    int lab[4];
    assign_segment<1, false>(cx, cy, cz, cn, k, lane, x, y, z, v, lab, nullptr);
    if (lvec) {
      *reinterpret_cast<I4*>(m.labels + idx[0]) = I4{lab[0], lab[1], lab[2], lab[3]};
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (v[j]) m.labels[idx[j]] = lab[j];
    }
  }
}

// Labels of the bounded images, from the per-pixel state (images that ran
// an iteration).
__global__ __launch_bounds__(256) void batch_labels_state(BatchArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  for_each_unit(a, [&](const PxImg& m, int64_t q0) {
    const int64_t q = q0 + lane;
    if (m.bounded && m.iter > 0 && q < m.n)
      m.labels[q] = (int32_t)(m.pstate[q] & 0xFFu);
  });
}

__global__ void batch_finish(BatchArgs a) {
  BatchImg& m = a.img[blockIdx.x];
  const float* src = m.cbuf + (size_t)m.cur * 3 * m.k;
  for (int i = threadIdx.x; i < 3 * m.k; i += blockDim.x) m.centroids_out[i] = src[i];
}

// The batched connected components: the single-image passes
// (ccl_device.hpp) over every image's 64x64 tiles and bitmap words.
constexpr int kCcSide = 64;
using BatchCcTile = CcTile<kCcSide>;

__device__ __forceinline__ CcArgs image_cc_args(const BatchImg& m,
                                                bool compress) {
  const int64_t nt = (int64_t)m.ctiles_x * ((m.H + kCcSide - 1) / kCcSide);
  CcArgs c{};
  c.labels = m.labels;
  c.out = m.segments;
  c.W = m.W;
  c.H = m.H;
  c.ntx = m.ctiles_x;
  c.linked = m.linked;
  c.local_root = compress ? m.bitmap : nullptr;
  c.llab = m.edge;
  c.rlab = m.edge + nt * kCcSide;
  c.lroot = m.edge + 2 * nt * kCcSide;
  c.rroot = m.edge + 3 * nt * kCcSide;
  return c;
}

// Calls f(m, tx, ty) for the block's contiguous share of the global CC
// tiles: one binary search for its first tile's image, then a walk forward
// (a search per tile cost ~8 dependent loads of latency each).
template <class F>
__device__ __forceinline__ void for_each_cc_tile(const BatchArgs& a, F&& f) {
  const int64_t c0 = a.ctiles * blockIdx.x / gridDim.x;
  const int64_t c1 = a.ctiles * (blockIdx.x + 1) / gridDim.x;
  if (c0 >= c1) return;
  int i = find_image<&BatchImg::ctile0>(a.img, a.count, c0);
  int64_t next = i + 1 < a.count ? a.img[i + 1].ctile0 : INT64_MAX;
  for (int64_t c = c0; c < c1; ++c) {
    while (c >= next) {
      ++i;
      next = i + 1 < a.count ? a.img[i + 1].ctile0 : INT64_MAX;
    }
    const BatchImg& m = a.img[i];
    const int64_t local = c - m.ctile0;
    const int ty = (int)(local / m.ctiles_x);
    const int tx = (int)(local - (int64_t)ty * m.ctiles_x);
    f(m, tx, ty);
  }
}

// Zeroes every image's linked-root and local-root bitmaps (images strided
// over blockIdx.y; 8-byte stores: the allocations are 256-byte aligned and
// hold 2 ceil(n/64) words).
__global__ __launch_bounds__(256) void batch_cc_clear(BatchArgs a) {
  for (int i = blockIdx.y; i < a.count; i += gridDim.y) {
    const BatchImg& m = a.img[i];
    const int64_t pairs = ((int64_t)m.W * m.H + 63) / 64;  // uint2 per bitmap
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
         q < pairs; q += (int64_t)gridDim.x * blockDim.x) {
      reinterpret_cast<uint2*>(m.linked)[q] = make_uint2(0u, 0u);
      reinterpret_cast<uint2*>(m.bitmap)[q] = make_uint2(0u, 0u);
    }
  }
}

template <int CONN>
__global__ __launch_bounds__(BatchCcTile::kThreads) void batch_cc_tile(
    BatchArgs a, int compress) {
  __shared__ int lab[BatchCcTile::kPx];
  __shared__ int par[BatchCcTile::kPx];
  for_each_cc_tile(a, [&](const BatchImg& m, int tx, int ty) {
    cc_tile_body<CONN, kCcSide>(image_cc_args(m, compress != 0), tx, ty, lab,
                                par);
    __syncthreads();  // lab / par are reused by the next tile
  });
}

template <int CONN>
__global__ __launch_bounds__(3 * kCcSide) void batch_cc_merge(BatchArgs a) {
  for_each_cc_tile(a, [&](const BatchImg& m, int tx, int ty) {
    cc_merge_body<CONN, kCcSide>(image_cc_args(m, false), tx, ty);
  });
}

// Root jumps over every image's linked-root words, in decreasing order of
// the padded global pixel space (each image's ancestors first): wave w of
// W takes the w-th share from the top and walks it downwards.
__global__ __launch_bounds__(256) void batch_cc_jump(BatchArgs a) {
  const int64_t units = a.px >> 6;
  const int wpb = (int)(blockDim.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * wpb;
  const int64_t gw = nw - 1 - ((int64_t)blockIdx.x * wpb +
                               __builtin_amdgcn_readfirstlane(
                                   (int)(threadIdx.x >> 6)));
  const int64_t u0 = units * gw / nw, u1 = units * (gw + 1) / nw;
  if (u0 >= u1) return;
  const int lane = threadIdx.x & (kWave - 1);
  int i = find_image<&BatchImg::px0>(a.img, a.count, (u1 - 1) << 6);
  // two lanes per 64-pixel unit (one bitmap word each), 32 units per pass;
  // i: the image of the pass's top unit
  for (int64_t ub = u1 - 1; ub >= u0; ub -= 32) {
    const int64_t u = ub - (lane >> 1);
    const bool ok = u >= u0;
    const int64_t g0 = (ok ? u : u0) << 6;
    int j = i;
    while (j > 0 && a.img[j].px0 > g0) --j;
    if (ok) {
      const BatchImg& m = a.img[j];
      const int64_t w = ((g0 - m.px0) >> 5) + 1 - (lane & 1);
      if (w * 32 < (int64_t)m.W * m.H) cc_jump_word(m.segments, m.linked, w);
    }
    i = __builtin_amdgcn_readlane(j, kWave - 1);  // the pass's lowest unit
  }
}

// Per image (one block): roots = local roots not linked, written back into
// the bitmap, and the exclusive popcount prefix of its words.
__global__ __launch_bounds__(1024) void batch_scan(BatchArgs a) {
  BatchImg& m = a.img[blockIdx.x];
  const int64_t n = (int64_t)m.W * m.H;
  const int64_t words = (n + 31) / 32;
  const int nt = blockDim.x;
  const int64_t per = (words + nt - 1) / nt;
  const int64_t w0 = (int64_t)threadIdx.x * per;
  unsigned sum = 0;
  for (int64_t w = w0; w < w0 + per && w < words; ++w) {
    const unsigned r = m.bitmap[w] & ~m.linked[w];
    m.bitmap[w] = r;
    sum += __popc(r);
  }
  __shared__ unsigned s[1024];
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {
    const unsigned add = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  unsigned run = s[threadIdx.x] - sum;
  for (int64_t w = w0; w < w0 + per && w < words; ++w) {
    m.prefix[w] = run;
    run += __popc(m.bitmap[w]);
  }
  if (threadIdx.x == nt - 1) m.nseg = (int)s[nt - 1];
}

// Every pixel's root (or, with compress, its dense rank), per tile.
__global__ __launch_bounds__(BatchCcTile::kThreads) void batch_cc_final(
    BatchArgs a, int compress) {
  __shared__ int sv[BatchCcTile::kPx];
  for_each_cc_tile(a, [&](const BatchImg& m, int tx, int ty) {
    cc_final_body<kCcSide>(image_cc_args(m, false), tx, ty,
                           compress ? m.bitmap : nullptr, m.prefix, nullptr,
                           sv);
    __syncthreads();  // sv is reused by the next tile
  });
}

// ------------------------------------------------------------------ host ---

struct BatchPlan {
  size_t chunks_off = 0;  // chunk table offset in temp
  std::vector<BatchImg> img;  // table order: plain images, then bounded ones
  std::vector<int> order;     // table entry -> caller's index
  int nplain = 0;
  int64_t tiles_plain = 0;
  size_t total_bytes = 0;
  int64_t tiles = 0, ctiles = 0, px = 0;
  int64_t cost_bounded = 0;  // tile cost of the bounded group
};

static int make_plan(const split_hip_image_desc* d, int count, void* temp,
                     bool need_segments, int tw_log, BatchPlan* P) {
  for (int i = 0; i < count; ++i) {
    const split_hip_image_desc& e = d[i];
    if (e.width < 1 || e.height < 1 || (int64_t)e.width * e.height > INT32_MAX)
      return fail(SPLIT_HIP_ERR_ARGUMENT, "image %d is %d x %d", i, e.width,
                  e.height);
    if (e.k < 1 || e.k > kBatchMaxK)
      return fail(SPLIT_HIP_ERR_UNSUPPORTED, "image %d: k = %d (batch k <= %d)",
                  i, e.k, kBatchMaxK);
    if (temp && (!e.points || !e.centroids || !e.labels ||
                 (need_segments && !e.segments)))
      return fail(SPLIT_HIP_ERR_ARGUMENT, "image %d has a null buffer", i);
  }
  auto bounded = [&](int i) { return d[i].k >= kBoundedMinK; };
  P->order.clear();
  for (int i = 0; i < count; ++i)
    if (!bounded(i)) P->order.push_back(i);
  P->nplain = (int)P->order.size();
  for (int i = 0; i < count; ++i)
    if (bounded(i)) P->order.push_back(i);

  P->img.assign(count, BatchImg{});
  char* t = static_cast<char*>(temp);
  size_t o = align_up((size_t)count * sizeof(BatchImg), 256);
  P->chunks_off = o;
  o += align_up(offsetof(ChunkTable, e) + (size_t)kMaxChunks * sizeof(ChunkEnt), 256);
  int64_t tiles = 0, ctiles = 0, px = 0, cost = 0;
  for (int j = 0; j < count; ++j) {
    BatchImg& m = P->img[j];
    if (j == P->nplain) cost = 0;  // each group's costs count from 0
    const split_hip_image_desc& e = d[P->order[j]];
    const int64_t n = (int64_t)e.width * e.height;
    if (j == P->nplain) P->tiles_plain = tiles;
    m.pts = e.points;
    m.labels = e.labels;
    m.segments = e.segments;
    m.centroids_in = e.centroids;
    m.centroids_out = e.centroids;
    m.W = e.width;
    m.H = e.height;
    m.k = e.k;
    m.bounded = j >= P->nplain;
    m.tw_log = tw_log;
    m.tiles_x = (int)tiles_across(e.width, tw_log);
    m.ctiles_x = (e.width + kCcSide - 1) / kCcSide;
    m.tile0 = tiles;
    m.cost0 = cost;
    m.ctile0 = ctiles;
    m.px0 = px;
    const int64_t img_tiles = (int64_t)m.tiles_x * tiles_down(e.height, tw_log);
    tiles += img_tiles;
    cost += img_tiles * tile_cost(e.k);
    const int64_t img_ctiles =
        (int64_t)m.ctiles_x * ((e.height + kCcSide - 1) / kCcSide);
    ctiles += img_ctiles;
    px += (n + 63) / 64 * 64;
    auto take = [&](size_t bytes) {
      char* p = t ? t + o : nullptr;
      o += align_up(bytes, 256);
      return p;
    };
    m.cbuf = reinterpret_cast<float*>(take(6 * (size_t)e.k * sizeof(float)));
    m.totals = reinterpret_cast<long long*>(take(4 * (size_t)e.k * 8));
    m.bitmap = reinterpret_cast<unsigned*>(take(2 * (size_t)((n + 63) / 64) * 4));
    m.linked = reinterpret_cast<unsigned*>(take(2 * (size_t)((n + 63) / 64) * 4));
    m.edge = reinterpret_cast<int32_t*>(take(4 * (size_t)img_ctiles * kCcSide * 4));
    m.prefix = reinterpret_cast<unsigned*>(take((size_t)((n + 31) / 32) * 4));
    m.pstate = reinterpret_cast<uint32_t*>(take(((size_t)n + kStatePad) * 4));
    m.fcum = reinterpret_cast<float*>(take((size_t)e.k * sizeof(float)));
  }
  if (P->nplain == count) P->tiles_plain = tiles;
  P->tiles = tiles;
  P->cost_bounded = P->nplain < count ? cost : 0;
  P->ctiles = ctiles;
  P->px = px;
  P->total_bytes = o;
  return SPLIT_HIP_OK;
}

struct StepTune {
  int threads = 512;
  int per_cu = 1;
};

static size_t step_lds(bool bounded, int threads) {
  return (size_t)(threads / kWave) *
         (bounded ? sizeof(WaveLds) : sizeof(WaveLdsPlain));
}

static const void* step_fn(bool bounded, bool first) {
  if (!bounded) return (const void*)batch_step_plain;
  return first ? (const void*)batch_step_bounded<true>
               : (const void*)batch_step_bounded<false>;
}

static void launch_step(const BatchArgs& a, const Group& g, bool bounded,
                        bool first, const StepTune& t, hipStream_t stream) {
  const int nwaves = t.threads / kWave;
  const int64_t want = (g.t1 - g.t0 + nwaves - 1) / nwaves;
  const int grid = (int)std::max<int64_t>(
      1, std::min<int64_t>(want, (int64_t)device_cu_count() * t.per_cu));
  const size_t lds = step_lds(bounded, t.threads);
  if (!bounded)
    hipLaunchKernelGGL(batch_step_plain, dim3(grid), dim3(t.threads), lds,
                       stream, a, g);
  else if (first)
    hipLaunchKernelGGL(batch_step_bounded<true>, dim3(grid), dim3(t.threads),
                       lds, stream, a, g);
  else {
    // about one chunk per wave of the step grid (never below
    // kMinChunkTiles, never above kBalanceTiles tiles)
    hipLaunchKernelGGL(batch_chunks, dim3(1), dim3(1024), 0, stream, a, g,
                       grid * nwaves);
    hipLaunchKernelGGL(batch_step_bounded<false>, dim3(grid), dim3(t.threads),
                       lds, stream, a, g);
  }
}

// Autotuning of the Lloyd tile shape and the step kernels' block shapes,
// cached per batch signature (device, images, padded pixels, plain images,
// largest k of each group).  Every tile shape gives the same results (exact
// argmin, integer sums); for each shape the tables are re-planned and each
// group's block size x blocks-per-CU candidates are timed over one full
// and three bounded iterations of the batch itself (after one untimed full
// iteration); the shape with the smallest sum of its groups' best times
// wins.  batch_256: 16x16 / 32x8 / 64x4 tiles took 178.8 / 171.8 / 176.9 ms
// per batch (bounded iteration 2.40 / 2.20 / 2.23 ms).
struct BatchTune {
  int tw_log = 4;
  StepTune tp, tb;
};
using BatchTuneKey = std::tuple<int, int, int64_t, int, int, int>;
static std::mutex g_tune_mu;
static std::map<BatchTuneKey, BatchTune> g_tune;

constexpr int kTileShapes[] = {4, 5, 6};  // tw_log: 16x16, 32x8, 64x4

static int time_group(const BatchArgs& a, const Group& g, bool bounded,
                      const StepTune& cand, hipEvent_t e0, hipEvent_t e1,
                      hipStream_t stream, float* ms) {
  BatchArgs ta = a;
  ta.max_iter = 1 << 30;  // never stop while timing
  ta.threshold = -1.f;
  hipLaunchKernelGGL(batch_init, dim3(a.count), dim3(256), 0, stream, ta);
  hipLaunchKernelGGL(batch_absmax, dim3(8 * device_cu_count()), dim3(256), 0,
                     stream, ta);
  launch_step(ta, g, bounded, true, cand, stream);  // untimed
  SPLIT_HIP_CHECK(hipEventRecord(e0, stream));
  launch_step(ta, g, bounded, true, cand, stream);
  for (int r = 0; r < 3; ++r) {
    // a full step re-assigns (and records) every pixel: the bounded steps
    // after it run from a consistent state
    launch_step(ta, g, bounded, false, cand, stream);
  }
  SPLIT_HIP_CHECK(hipEventRecord(e1, stream));
  SPLIT_HIP_CHECK(hipEventSynchronize(e1));
  SPLIT_HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
  SPLIT_HIP_CHECK_LAUNCH("batch autotune");
  return SPLIT_HIP_OK;
}

// Best block shape of one group on the uploaded tables, and its time.
static int tune_group(const BatchArgs& a, const Group& g, bool bounded,
                      hipEvent_t e0, hipEvent_t e1, hipStream_t stream,
                      StepTune* out, float* out_ms) {
  StepTune best;
  float best_ms = 1e30f;
  const int cands_plain[] = {256, 512, 1024};
  const int cands_bounded[] = {256, 512};
  const int* cands = bounded ? cands_bounded : cands_plain;
  const int ncand = bounded ? 2 : 3;
  for (int c = 0; c < ncand; ++c) {
    const int threads = cands[c];
    int occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, step_fn(bounded, true), threads,
            step_lds(bounded, threads)) != hipSuccess ||
        occ < 1)
      occ = 1;
    for (int per_cu = 1; per_cu <= occ; per_cu *= 2) {
      const StepTune cand{threads, per_cu};
      float ms = 0.f;
      SPLIT_HIP_TRY(time_group(a, g, bounded, cand, e0, e1, stream, &ms));
      if (ms < best_ms) {
        best_ms = ms;
        best = cand;
      }
    }
  }
  *out = best;
  *out_ms = best_ms;
  return SPLIT_HIP_OK;
}


struct Groups {
  Group plain, bounded;
  StepTune tp, tb;
};

// Enqueues the whole batched segmentation on `stream`.
static void enqueue_batch(const BatchArgs& a, const Groups& G, int conn,
                          int compress, hipStream_t stream) {
  const int cus = device_cu_count();
  hipLaunchKernelGGL(batch_init, dim3(a.count), dim3(256), 0, stream, a);
  hipLaunchKernelGGL(batch_absmax, dim3(8 * cus), dim3(256), 0, stream, a);
  for (int it = 0; it < a.max_iter; ++it) {
    if (G.plain.count) launch_step(a, G.plain, false, it == 0, G.tp, stream);
    if (G.bounded.count) launch_step(a, G.bounded, true, it == 0, G.tb, stream);
  }
  if (G.plain.count)
    hipLaunchKernelGGL(batch_labels_plain, dim3(8 * cus), dim3(256), 0, stream,
                       G.plain);
  if (G.bounded.count)
    hipLaunchKernelGGL(batch_labels_state, dim3(8 * cus), dim3(256), 0, stream,
                       a);
  hipLaunchKernelGGL(batch_finish, dim3(a.count), dim3(256), 0, stream, a);
  if (conn == 0) return;  // clustering only
  const int cgrid = (int)std::min<int64_t>(a.ctiles, (int64_t)cus * 8);
  hipLaunchKernelGGL(batch_cc_clear, dim3(16, std::min(a.count, 65535)),
                     dim3(256), 0, stream, a);
  if (conn == 8) {
    hipLaunchKernelGGL(batch_cc_tile<8>, dim3(cgrid),
                       dim3(BatchCcTile::kThreads), 0, stream, a, compress);
    hipLaunchKernelGGL(batch_cc_merge<8>, dim3(cgrid), dim3(3 * kCcSide), 0,
                       stream, a);
  } else {
    hipLaunchKernelGGL(batch_cc_tile<4>, dim3(cgrid),
                       dim3(BatchCcTile::kThreads), 0, stream, a, compress);
    hipLaunchKernelGGL(batch_cc_merge<4>, dim3(cgrid), dim3(3 * kCcSide), 0,
                       stream, a);
  }
  hipLaunchKernelGGL(batch_cc_jump, dim3(32 * cus), dim3(256), 0, stream, a);
  if (compress)
    hipLaunchKernelGGL(batch_scan, dim3(a.count), dim3(1024), 0, stream, a);
  hipLaunchKernelGGL(batch_cc_final, dim3(cgrid), dim3(BatchCcTile::kThreads),
                     0, stream, a, compress);
}

struct GraphEntry {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};
// Everything a captured graph bakes in: (device, temp, images, plain
// images, plain tiles, Lloyd tiles, CC tiles, pixels, max_iter, threshold,
// connectivity, compress, tuned plain / bounded block shapes, tile shape).
// Per-image sizes,
// k and buffers are read from the descriptor table at run time; the totals
// are launch bounds and grid sizes.
using GraphKey = std::tuple<int, void*, int, int, int64_t, int64_t, int64_t,
                            int64_t, int, float, int, int, int, int, int,
                            int, int>;
static std::mutex g_graph_mu;
static std::map<GraphKey, GraphEntry> g_graphs;

}  // namespace splithip

using namespace splithip;

extern "C" {

size_t split_hip_segment_batch_workspace(const split_hip_image_desc* images,
                                         int32_t count) {
  if (!images || count < 1) return 0;
  BatchPlan P;
  if (make_plan(images, count, nullptr, false, 4, &P) != SPLIT_HIP_OK) return 0;
  return P.total_bytes;
}

int split_hip_segment_batch(const split_hip_image_desc* images, int32_t count,
                            void* temp, size_t temp_bytes, int32_t max_iter,
                            float threshold, int32_t connectivity,
                            int32_t compress, int32_t use_graph,
                            split_hip_batch_result* results,
                            split_hip_stream_t stream) {
  if (!images || count < 1 || !temp)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "images, count or temp missing");
  // connectivity 0: the clustering alone (centroids and labels; segments
  // may be null and are not written)
  if (connectivity != 0 && connectivity != 4 && connectivity != 8)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "connectivity %d", connectivity);
  BatchPlan P;
  SPLIT_HIP_TRY(make_plan(images, count, temp, connectivity != 0, 4, &P));
  if (temp_bytes < P.total_bytes)
    return fail(SPLIT_HIP_ERR_TEMP, "temp holds %zu bytes, batch needs %zu",
                temp_bytes, P.total_bytes);
  BatchArgs a;
  a.img = static_cast<BatchImg*>(temp);
  a.chunks = reinterpret_cast<ChunkTable*>(static_cast<char*>(temp) + P.chunks_off);
  a.count = count;
  a.max_iter = std::max(0, max_iter);
  a.threshold = threshold;
  int device = 0;
  SPLIT_HIP_CHECK(hipGetDevice(&device));
  int kp = 0, kb = 0;
  for (int j = 0; j < count; ++j)
    (j < P.nplain ? kp : kb) = std::max(j < P.nplain ? kp : kb, P.img[j].k);
  Groups G;
  // tables for tile shape tw (re-planned unless they have it), uploaded
  auto upload = [&](int tw) -> int {
    if (P.img[0].tw_log != tw)
      SPLIT_HIP_TRY(make_plan(images, count, temp, connectivity != 0, tw, &P));
    a.tiles = P.tiles;
    a.ctiles = P.ctiles;
    a.px = P.px;
    SPLIT_HIP_CHECK(hipMemcpyAsync(temp, P.img.data(),
                                   (size_t)count * sizeof(BatchImg),
                                   hipMemcpyHostToDevice, stream));
    G.plain = Group{a.img, P.nplain, 0, P.tiles_plain};
    G.bounded = Group{a.img + P.nplain, count - P.nplain, P.tiles_plain, P.tiles,
                      0, P.cost_bounded};
    return SPLIT_HIP_OK;
  };
  const BatchTuneKey tkey{device, count, P.px, P.nplain, kp, kb};
  BatchTune tune;
  bool tuned = false;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune.find(tkey);
    if (it != g_tune.end()) {
      tune = it->second;
      tuned = true;
    }
  }
  if (!tuned) {
    hipEvent_t e0, e1;
    SPLIT_HIP_CHECK(hipEventCreate(&e0));
    SPLIT_HIP_CHECK(hipEventCreate(&e1));
    float best_total = 1e30f;
    for (const int tw : kTileShapes) {
      SPLIT_HIP_TRY(upload(tw));
      BatchTune cand;
      cand.tw_log = tw;
      float total = 0.f, ms = 0.f;
      if (G.plain.count) {
        SPLIT_HIP_TRY(tune_group(a, G.plain, false, e0, e1, stream, &cand.tp, &ms));
        total += ms;
      }
      if (G.bounded.count) {
        SPLIT_HIP_TRY(tune_group(a, G.bounded, true, e0, e1, stream, &cand.tb, &ms));
        total += ms;
      }
      if (total < best_total) {
        best_total = total;
        tune = cand;
      }
    }
    SPLIT_HIP_CHECK(hipEventDestroy(e0));
    SPLIT_HIP_CHECK(hipEventDestroy(e1));
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[tkey] = tune;
  }
  SPLIT_HIP_TRY(upload(tune.tw_log));
  G.tp = tune.tp;
  G.tb = tune.tb;

  if (use_graph) {
    const GraphKey key{device,      temp,          count,      P.nplain,
                       P.tiles_plain, P.tiles,     P.ctiles,   P.px,
                       a.max_iter,  threshold,     connectivity, compress != 0,
                       G.tp.threads, G.tp.per_cu,  G.tb.threads, G.tb.per_cu,
                       tune.tw_log};
    hipGraphExec_t exec = nullptr;
    {
      std::lock_guard<std::mutex> lock(g_graph_mu);
      auto it = g_graphs.find(key);
      if (it != g_graphs.end()) exec = it->second.exec;
    }
    if (!exec) {
      hipStream_t cap = nullptr;
      SPLIT_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      GraphEntry e;
      SPLIT_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
      enqueue_batch(a, G, connectivity, compress != 0, cap);
      const hipError_t le = hipGetLastError();
      SPLIT_HIP_CHECK(hipStreamEndCapture(cap, &e.graph));
      if (le != hipSuccess)
        return fail(SPLIT_HIP_ERR_HIP, "capture of the batch failed: %s",
                    hipGetErrorString(le));
      SPLIT_HIP_CHECK(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
      SPLIT_HIP_CHECK(hipStreamDestroy(cap));
      std::lock_guard<std::mutex> lock(g_graph_mu);
      g_graphs[key] = e;
      exec = e.exec;
    }
    SPLIT_HIP_CHECK(hipGraphLaunch(exec, stream));
  } else {
    enqueue_batch(a, G, connectivity, compress != 0, stream);
    SPLIT_HIP_CHECK_LAUNCH("batch kernels");
  }
  if (results) {
    std::vector<BatchImg> back(count);
    SPLIT_HIP_CHECK(hipMemcpyAsync(back.data(), temp,
                                   (size_t)count * sizeof(BatchImg),
                                   hipMemcpyDeviceToHost, stream));
    SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
    for (int j = 0; j < count; ++j) {
      split_hip_batch_result& r = results[P.order[j]];
      r.iterations = back[j].iter;
      r.shift = back[j].shift;
      r.num_segments = compress && connectivity != 0 ? back[j].nseg : -1;
    }
  } else {
    SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
  }
  return SPLIT_HIP_OK;
}

}  // extern "C"
