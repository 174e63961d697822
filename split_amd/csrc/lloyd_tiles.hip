// Bounded Lloyd iterations over 16x16 tiles with exact integer sums.
//
// Used by cluster_image / segment when k <= 256.  Results equal plain Lloyd:
// every label is the dense float argmin (ties to the lowest index), and the
// centroids are exact functions of cluster membership.
//
// Per-pixel state (4 bytes, in temp): label | t0 << 8 | fp16 slack << 16,
// where `slack` is a lower bound on (distance to the second-nearest centroid
// - distance to the nearest one) for the centroid set number t0, rounded
// down.  If centroid j moves by delta_j between sets, distances move by at
// most delta_j (triangle inequality), so a pixel provably keeps its label
// while
//     slack - 2 * (DM[t] - DM[t0]) > margin,
// DM[t] being the cumulative sum of max_j delta_j up to set t (Hamerly's
// bound with the largest drift on both sides).  margin (1e-5 of the data
// scale) exceeds the float error of the distances, so the float argmin is
// unchanged too.
//
// Iteration 1 assigns every tile.  Later iterations read the state, skip
// tiles whose every pixel passes the test without touching their colours,
// and re-assign the others whole (writing fresh state).
//
// Sums: channels are quantised once, q = trunc(v * 2^S) with |q| <= 2^22
// (S from the largest |v| of the image), and summed as integers into int64
// totals in temp: iteration 1 adds every pixel, later iterations add +q to
// the new and -q to the old cluster of each pixel whose label changed.
// Integer sums are exact, so totals are independent of order and of which
// pixels were skipped, and identical run to run.  The quantisation moves a
// mean by at most max|v| * 2^-22.
#include <algorithm>

#include "common.hpp"
#include "lloyd_device.hpp"
#include "lloyd_sums.hpp"

namespace splithip {

constexpr int kDmSlots = 256;    // drift history entries; t0 is 8 bits

struct TileArgs : Geometry {
  int k;
  float* cbuf;                // two 3k centroid buffers
  long long* totals;          // [4][k]: quantised channel sums, counts
  float* dm;                  // cumulative max drift, kDmSlots entries
  uint32_t* pstate;           // per-pixel state (see above)
  const unsigned* maxabs;     // bits of max |v| (non-negative float)
  LloydState* state;
  Mailbox* mbox;
};

// fp16 bits of s >= 0 rounded down (to the next smaller half when the
// nearest one is above s; a value past the fp16 range becomes 65504).
__device__ __forceinline__ unsigned short half_bits_down(float s) {
  const _Float16 h = (_Float16)s;
  unsigned short b = __builtin_bit_cast(unsigned short, h);
  if ((float)h > s && b > 0) b -= 1;
  return b;
}
__device__ __forceinline__ float half_bits_to_float(unsigned short b) {
  return (float)__builtin_bit_cast(_Float16, b);
}

__device__ __forceinline__ uint32_t pack_state(int label, int t0, float slack) {
  const unsigned short h = half_bits_down(slack);
  return (uint32_t)(label & 0xFF) | ((uint32_t)(t0 & 0xFF) << 8) |
         ((uint32_t)h << 16);
}

// Last block: totals -> centroids, shift, drift history, stop decision.
__device__ void tiles_finalize(const TileArgs& a, const float* cent,
                               float* cent_out, float fscale, double* scratch) {
  const int k = a.k;
  const int tid = threadIdx.x, nt = blockDim.x;
  double* d2 = scratch;             // k entries: |new - old|^2
  double* part = scratch + k;       // kWave partial sums
  double* pmax = part + kWave;      // kWave partial maxima
  const double inv = 1.0 / (double)fscale;
  for (int j = tid; j < k; j += nt) {
    const long long cnt = __hip_atomic_load(&a.totals[3 * k + j],
                                            __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long sum = __hip_atomic_load(&a.totals[c * k + j],
                                              __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
      const float nc = cnt > 0 ? (float)((double)sum * inv / (double)cnt) : 0.f;
      const double d = (double)nc - (double)cent[c * k + j];
      s += d * d;
      cent_out[c * k + j] = nc;
    }
    d2[j] = s;
  }
  __syncthreads();
  if (tid < kWave) {
    double s = 0.0, m = 0.0;
    for (int j = tid; j < k; j += kWave) {
      s += d2[j];
      m = fmax(m, d2[j]);
    }
    part[tid] = s;
    pmax[tid] = m;
  }
  __syncthreads();
  if (tid == 0) {
    double s = 0.0, m = 0.0;
    for (int j = 0; j < kWave; ++j) {
      s += part[j];
      m = fmax(m, pmax[j]);
    }
    LloydState* st = a.state;
    const float shift = (float)sqrt(s);
    const int it = st->iter + 1;
    // rounded up: the skip test must never under-estimate the drift
    a.dm[it] = (float)((double)a.dm[it - 1] + sqrt(m) * (1.0 + 1e-5) + 1e-7);
    const int done = !(shift > st->threshold) || it >= st->max_iter;
    st->iter = it;
    st->shift = shift;
    st->cur = 1 - st->cur;
    st->done = done;
    if (a.mbox) {
      volatile Mailbox* mb = a.mbox;
      mb->iter = it;
      mb->shift = shift;
      mb->done = done;
      __threadfence_system();
    }
    st->arrive = 0u;
  }
}

// Re-assigns up to 256 pixels (4 per lane: idx/v) whose previous labels are
// old[] (-1: none), writes their new state, and accumulates: every pixel
// (FULL) or the +/- changes of label (bounded iterations).
template <int CPL, bool FULL>
__device__ __forceinline__ void reassign(
    const TileArgs& a, const float (&cx)[CPL], const float (&cy)[CPL],
    const float (&cz)[CPL], const float (&cn)[CPL], int it, float fscale,
    int lane, const int64_t (&idx)[4], const bool (&v)[4], bool lvec,
    const int (&old)[4], long long* slice, IRun& run) {
  const int k = a.k;
  float x[4], y[4], z[4];
  load_colours(a, idx, v, lvec, x, y, z);
  int lab[4];
  float slack[4];
  assign_segment<CPL, true>(cx, cy, cz, cn, k, lane, x, y, z, v, lab, slack);
  uint32_t ns[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) ns[i] = pack_state(lab[i], it, slack[i]);
  if (lvec) {
    *reinterpret_cast<uint4*>(a.pstate + idx[0]) =
        make_uint4(ns[0], ns[1], ns[2], ns[3]);
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (v[i]) a.pstate[idx[i]] = ns[i];
  }
  int qx[4], qy[4], qz[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    qx[i] = (int)(x[i] * fscale);
    qy[i] = (int)(y[i] * fscale);
    qz[i] = (int)(z[i] * fscale);
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
    bool any = false;
    int to[4], from[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool chg = v[i] && lab[i] != old[i];
      to[i] = chg ? lab[i] : -1;
      from[i] = chg ? old[i] : -1;
      any = any || chg;
    }
    if (ballot(any) != 0) {
      add_by_label(slice, k, lane, to, qx, qy, qz, 1);
      add_by_label(slice, k, lane, from, qx, qy, qz, -1);
    }
  }
}

constexpr int kQueueCap = 512;      // per-wave queue of failing pixels
constexpr int kWholeTileFails = 96;  // more failing pixels: re-assign the tile

template <int CPL, int MODE, bool FULL>
__global__ __launch_bounds__(1024) void tiles_step(TileArgs a) {
  extern __shared__ long long lds64[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  const int k = a.k;
  const LloydState* st = a.state;
  if (st->done) return;
  const int cur = st->cur;
  const int it = st->iter;  // index of the centroid set assigned against
  const float* cent = a.cbuf + (size_t)cur * 3 * k;
  float* cent_out = a.cbuf + (size_t)(1 - cur) * 3 * k;

  const int V = 4 * k;
  long long* slices = lds64;  // nwaves x V
  float* dmt = reinterpret_cast<float*>(
      lds64 + std::max(nwaves * V, k + 2 * kWave));
  int* qidx = reinterpret_cast<int*>(dmt + kDmSlots) + wave * 2 * kQueueCap;
  int* qold = qidx + kQueueCap;
  for (int i = threadIdx.x; i < nwaves * V; i += blockDim.x) slices[i] = 0;
  if constexpr (!FULL)
    for (int i = threadIdx.x; i <= it; i += blockDim.x) dmt[i] = a.dm[i];

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
  const float maxabs = __uint_as_float(*a.maxabs);
  const float fscale = scale_of(*a.maxabs);
  const float margin = 1e-5f * (1.f + 4.f * maxabs);
  __syncthreads();
  const float drift_now = FULL ? 0.f : dmt[it];
  long long* slice = slices + (size_t)wave * V;

  const int64_t gw = (int64_t)blockIdx.x * nwaves + wave;
  const int64_t W = (int64_t)gridDim.x * nwaves;
  const int64_t s0 = a.nseg * gw / W;
  const int64_t s1 = a.nseg * (gw + 1) / W;

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
    reassign<CPL, false>(a, cx, cy, cz, cn, it, fscale, lane, idx, v, false,
                         old, slice, run);
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

  for (int64_t s = s0; s < s1; ++s) {
    int64_t idx[4];
    bool v[4], lvec;
    segment_index<MODE>(a, s, lane, idx, v, lvec);
    if constexpr (FULL) {
      const int old[4] = {-1, -1, -1, -1};
      reassign<CPL, true>(a, cx, cy, cz, cn, it, fscale, lane, idx, v, lvec,
                          old, slice, run);
    } else {
      uint32_t ps[4];
      if (lvec) {
        const uint4 q = *reinterpret_cast<const uint4*>(a.pstate + idx[0]);
        ps[0] = q.x; ps[1] = q.y; ps[2] = q.z; ps[3] = q.w;
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) ps[i] = v[i] ? a.pstate[idx[i]] : 0u;
      }
      int old[4];
      bool fail[4];
      uint64_t fb[4];
      int nfail = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        old[i] = v[i] ? (int)(ps[i] & 0xFFu) : -1;
        const int t0 = (int)((ps[i] >> 8) & 0xFFu);
        const float sl = half_bits_to_float((unsigned short)(ps[i] >> 16));
        fail[i] = v[i] && !(sl - 2.f * (drift_now - dmt[t0]) > margin);
        fb[i] = ballot(fail[i]);
        nfail += __builtin_popcountll(fb[i]);
      }
      if (nfail == 0) continue;  // every pixel provably unchanged
      if (nfail > kWholeTileFails) {
        reassign<CPL, false>(a, cx, cy, cz, cn, it, fscale, lane, idx, v, lvec,
                             old, slice, run);
        continue;
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
    }
  }
  if constexpr (!FULL) {
    while (qlen > 0) drain();
  }
  if constexpr (FULL) irun_flush(run, slice, k, lane);
  __syncthreads();
  for (int vv = threadIdx.x; vv < V; vv += blockDim.x) {
    long long tot = 0;
    for (int w = 0; w < nwaves; ++w) tot += slices[(size_t)w * V + vv];
    if (tot != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(&a.totals[vv]),
                (unsigned long long)tot);
  }
  __threadfence();
  __shared__ unsigned ticket;
  __syncthreads();
  if (threadIdx.x == 0) ticket = atomicAdd(&a.state->arrive, 1u);
  __syncthreads();
  if (ticket != gridDim.x - 1) return;
  __threadfence();
  tiles_finalize(a, cent, cent_out, fscale,
                 reinterpret_cast<double*>(lds64));
}

__global__ void tiles_init(LloydState* st, float* cbuf, const float* cent,
                           int k, int max_iter, float threshold,
                           long long* totals, float* dm, unsigned* maxabs) {
  for (int i = threadIdx.x; i < 3 * k; i += blockDim.x) cbuf[i] = cent[i];
  for (int i = threadIdx.x; i < 4 * k; i += blockDim.x) totals[i] = 0;
  if (threadIdx.x == 0) {
    st->arrive = 0u;
    st->cur = 0;
    st->iter = 0;
    st->done = max_iter <= 0;
    st->max_iter = max_iter;
    st->threshold = threshold;
    st->shift = 0.f;
    st->reserved = 0;
    dm[0] = 0.f;
    *maxabs = 0u;
  }
}

// Largest |v| over the three planes (as float bits; non-negative floats
// order like their bit patterns).
__global__ __launch_bounds__(256) void tiles_absmax(const float* pts,
                                                    int64_t count, int vec,
                                                    unsigned* out) {
  float m = 0.f;
  const int64_t chunks = vec ? count / 4 : 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // four independent loads in flight per thread
  for (; c + 3 * stride < chunks; c += 4 * stride) {
    float4 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      q[u] = reinterpret_cast<const float4*>(pts)[c + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      m = fmaxf(m, fmaxf(fmaxf(fabsf(q[u].x), fabsf(q[u].y)),
                         fmaxf(fabsf(q[u].z), fabsf(q[u].w))));
  }
  for (; c < chunks; c += stride) {
    const float4 q = reinterpret_cast<const float4*>(pts)[c];
    m = fmaxf(m, fmaxf(fmaxf(fabsf(q.x), fabsf(q.y)),
                       fmaxf(fabsf(q.z), fabsf(q.w))));
  }
  for (int64_t i = 4 * chunks + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       i < count; i += stride)
    m = fmaxf(m, fabsf(pts[i]));
  m = wave_allreduce(m, OpMax());
  if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(out, __float_as_uint(m));
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

static int tiles_threads(int k) { return k <= 64 ? 512 : 256; }

struct TileLayout {
  LloydState* st;
  float* cbuf;
  long long* totals;
  float* dm;
  unsigned* maxabs;
  uint32_t* pstate;
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
  L.totals = reinterpret_cast<long long*>(take(4 * (size_t)k * 8));
  L.dm = reinterpret_cast<float*>(take(kDmSlots * sizeof(float)));
  L.maxabs = reinterpret_cast<unsigned*>(take(sizeof(unsigned)));
  L.pstate = reinterpret_cast<uint32_t*>(take((size_t)n * sizeof(uint32_t)));
  L.bytes = o;
  return L;
}

size_t tiles_workspace(int64_t n, int k) {
  return tiles_layout(nullptr, n, k).bytes;
}

bool tiles_supported(int width, int k, int max_iter) {
  return width > 0 && k >= 1 && k <= 256 && max_iter <= kDmSlots - 2;
}

template <bool FULL, int MODE>
static void tiles_launch_mode(const TileArgs& a, int grid, int threads,
                              size_t lds, hipStream_t stream) {
  switch (a.k <= 64 ? 1 : a.k <= 128 ? 2 : 4) {
    case 1:
      hipLaunchKernelGGL((tiles_step<1, MODE, FULL>), dim3(grid),
                         dim3(threads), lds, stream, a);
      break;
    case 2:
      hipLaunchKernelGGL((tiles_step<2, MODE, FULL>), dim3(grid),
                         dim3(threads), lds, stream, a);
      break;
    default:
      hipLaunchKernelGGL((tiles_step<4, MODE, FULL>), dim3(grid),
                         dim3(threads), lds, stream, a);
      break;
  }
}

int cluster_tiles_async(const float* pts, int width, int height,
                        float* centroids, int k, int32_t* labels, void* temp,
                        size_t temp_bytes, int max_iter, float threshold,
                        hipStream_t stream, Mailbox** mbox_out) {
  const int64_t n = (int64_t)width * height;
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
                     centroids, k, max_iter, threshold, L.totals, L.dm,
                     L.maxabs);
  const int vec = ((uintptr_t)pts & 15u) == 0;
  hipLaunchKernelGGL(tiles_absmax, dim3(8 * device_cu_count()), dim3(256), 0,
                     stream, pts, 3 * n, vec, L.maxabs);
  SPLIT_HIP_CHECK_LAUNCH("tiles_init/tiles_absmax");

  TileArgs a{};
  a.pts = pts;
  a.n = n;
  a.width = width;
  a.height = height;
  a.tiles_x = (width + kTileSide - 1) / kTileSide;
  a.nseg = (int64_t)a.tiles_x * ((height + kTileSide - 1) / kTileSide);
  a.k = k;
  a.cbuf = L.cbuf;
  a.totals = L.totals;
  a.dm = L.dm;
  a.pstate = L.pstate;
  a.maxabs = L.maxabs;
  a.state = L.st;
  a.mbox = res->dev;

  const int threads = tiles_threads(k);
  const int nwaves = threads / kWave;
  const int64_t want = (a.nseg + nwaves - 1) / nwaves;
  const size_t lds =
      std::max<size_t>((size_t)nwaves * 4 * k, (size_t)k + 2 * kWave) * 8 +
      kDmSlots * sizeof(float) +
      (size_t)nwaves * 2 * kQueueCap * sizeof(int);
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
      1, std::min<int64_t>(want, (int64_t)device_cu_count() * per_cu));
  const bool vecmode = vec && (width % 4 == 0) &&
                       (((uintptr_t)L.pstate & 15u) == 0);
  auto launch = [&](bool full) {
    if (full) {
      if (vecmode) tiles_launch_mode<true, kTileVec>(a, grid, threads, lds, stream);
      else tiles_launch_mode<true, kTile>(a, grid, threads, lds, stream);
    } else {
      if (vecmode) tiles_launch_mode<false, kTileVec>(a, grid, threads, lds, stream);
      else tiles_launch_mode<false, kTile>(a, grid, threads, lds, stream);
    }
  };

  constexpr int kChunk = 4;
  int launched = 0, chunk = 0;
  while (launched < max_iter) {
    const int cnt = std::min(kChunk, max_iter - launched);
    for (int i = 0; i < cnt; ++i) launch(launched + i == 0);
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
