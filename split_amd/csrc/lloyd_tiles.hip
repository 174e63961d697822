// Bounded Lloyd iterations over 16x16 tiles with exact integer sums.
//
// Used by cluster_image / segment when k <= 256 (any iteration count).  Results equal plain Lloyd:
// every label is the dense float argmin (ties to the lowest index), and the
// centroids are exact functions of cluster membership.
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
#include "lloyd_bounded.hpp"

namespace splithip {

// totals (tot, [4][k]) -> centroids, shift, bound consumption, stop
// decision; one block.
__device__ void tiles_finalize(const TileArgs& a, const long long* tot,
                               const float* cent, float* cent_out,
                               float fscale, double* scratch) {
  const int k = a.k;
  const int tid = threadIdx.x, nt = blockDim.x;
  double* d2 = scratch;             // k entries: |new - old|^2
  double* part = scratch + k;       // kWave partial sums
  const double inv = 1.0 / (double)fscale;
  for (int j = tid; j < k; j += nt) {
    const long long cnt = tot[3 * k + j];
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long sum = tot[c * k + j];
      const float nc = cnt > 0 ? (float)((double)sum * inv / (double)cnt) : 0.f;
      const double d = (double)nc - (double)cent[c * k + j];
      s += d * d;
      cent_out[c * k + j] = nc;
    }
    d2[j] = s;
  }
  __syncthreads();
  if (tid < kWave) {
    double s = 0.0;
    for (int j = tid; j < k; j += kWave) s += d2[j];
    part[tid] = s;
  }
  __syncthreads();
  __shared__ double top[2];
  __shared__ int top_arg;
  __shared__ float shift_sh;
  if (tid == 0) {
    double s = 0.0;
    for (int j = 0; j < kWave; ++j) s += part[j];
    shift_sh = (float)sqrt(s);
    double m1 = 0.0, m2 = 0.0;
    int arg = -1;
    for (int j = 0; j < k; ++j) {
      const double d = d2[j];
      if (d > m1) {
        m2 = m1;
        m1 = d;
        arg = j;
      } else if (d > m2) {
        m2 = d;
      }
    }
    top[0] = m1;
    top[1] = m2;
    top_arg = arg;
  }
  __syncthreads();
  for (int j = tid; j < k; j += nt)
    a.fcum[j] = bound_step(a.fcum[j], d2[j], j, top[0], top[1], top_arg);
  if (tid == 0) {
    LloydState* st = a.state;
    const float shift = shift_sh;
    const int it = st->iter + 1;
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
  }
}

template <int CPL, int MODE, bool FULL>
__global__ __launch_bounds__(1024) void tiles_step(TileArgs a) {
  extern __shared__ long long lds64[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  const int k = a.k;
  const LloydState* st = a.state;
  if (st->done) return;
  const float* cent = a.cbuf + (size_t)st->cur * 3 * k;

  const int V = 4 * k;
  long long* slices = lds64;                                   // nwaves x V
  float* fl = reinterpret_cast<float*>(lds64 + nwaves * V);    // F[0..k)
  int* qidx = reinterpret_cast<int*>(fl + kMaxTileK) + wave * 2 * kQueueCap;
  int* qold = qidx + kQueueCap;
  for (int i = threadIdx.x; i < nwaves * V; i += blockDim.x) slices[i] = 0;
  for (int i = threadIdx.x; i < k; i += blockDim.x) fl[i] = a.fcum[i];
  __shared__ unsigned unit_next;
  if (threadIdx.x == 0) unit_next = 0u;

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
  long long* slice = slices + (size_t)wave * V;

  if constexpr (FULL) {
    // every tile costs the same: contiguous tiles per wave, so that runs of
    // one label accumulate in registers
    const int64_t gw = (int64_t)blockIdx.x * nwaves + wave;
    const int64_t W = (int64_t)gridDim.x * nwaves;
    walk_tiles<CPL, MODE, FULL>(
        a, TileRange{a.nseg * gw / W, a.nseg * (gw + 1) / W}, cx, cy, cz, cn,
        fl, fscale, margin, slice, qidx, qold, lane);
  } else {
    const int64_t G = gridDim.x;
    walk_tiles<CPL, MODE, FULL>(
        a,
        SpreadUnits{a.nunits * blockIdx.x / G, a.nunits * (blockIdx.x + 1) / G,
                    a.nunits, a.unit_mul, a.units_per_row, a.tiles_x,
                    &unit_next, lane, -1},
        cx, cy, cz, cn, fl, fscale, margin, slice, qidx, qold, lane);
  }
  // the block's share of the step, summed over its waves; tiles_reduce adds
  // the shares of all blocks into the totals after this launch ends
  __syncthreads();
  long long* part = a.partials + (size_t)blockIdx.x * V;
  for (int vv = threadIdx.x; vv < V; vv += blockDim.x) {
    long long tot = 0;
    for (int w = 0; w < nwaves; ++w) tot += slices[(size_t)w * V + vv];
    part[vv] = tot;
  }
}

// One block after each step launch: totals += the shares of the step's
// `blocks` blocks (fixed order; integer sums), then tiles_finalize.
constexpr int kReduceThreads = 1024;

__global__ __launch_bounds__(kReduceThreads) void tiles_reduce(TileArgs a,
                                                                int blocks) {
  __shared__ long long red[kReduceThreads];
  __shared__ long long tot[4 * kMaxTileK];
  __shared__ double scratch[kMaxTileK + kWave];
  LloydState* st = a.state;
  if (st->done) return;
  const int k = a.k, V = 4 * k;
  const int tid = threadIdx.x;
  static_assert(4 * kMaxTileK <= kReduceThreads, "one pass over the entries");
  // thread t sums entry t % V over blocks t / V, t / V + per, ...
  const int per = kReduceThreads / V;
  const int v = tid % V, q = tid / V;
  long long sum = 0;
  if (q < per)
    for (int g = q; g < blocks; g += per) sum += a.partials[(size_t)g * V + v];
  red[tid] = sum;
  __syncthreads();
  if (tid < V) {
    long long t = a.totals[tid];
    for (int r = 0; r < per; ++r) t += red[r * V + tid];
    a.totals[tid] = t;
    tot[tid] = t;
  }
  __syncthreads();
  const int cur = st->cur;
  const float* cent = a.cbuf + (size_t)cur * 3 * k;
  float* cent_out = a.cbuf + (size_t)(1 - cur) * 3 * k;
  tiles_finalize(a, tot, cent, cent_out, scale_of(*a.maxabs), scratch);
}

__global__ void tiles_init(LloydState* st, float* cbuf, const float* cent,
                           int k, int max_iter, float threshold,
                           long long* totals, float* fcum, unsigned* maxabs) {
  for (int i = threadIdx.x; i < 3 * k; i += blockDim.x) cbuf[i] = cent[i];
  for (int i = threadIdx.x; i < 4 * k; i += blockDim.x) totals[i] = 0;
  for (int i = threadIdx.x; i < k; i += blockDim.x) fcum[i] = 0.f;
  if (threadIdx.x == 0) {
    st->cur = 0;
    st->iter = 0;
    st->done = max_iter <= 0;
    st->max_iter = max_iter;
    st->threshold = threshold;
    st->shift = 0.f;
    st->next = 0u;
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

// One block per CU where the LDS allows (k <= 128: up to 129 KB): the
// waves of a CU then share one pool of tiles in the bounded iterations and
// finish together.
static int tiles_threads(int k) { return k <= 128 ? 1024 : 256; }
// Step launches use at most this many blocks (partials in temp are sized
// for it).
static int tiles_max_blocks(int k) { return k <= 128 ? 512 : 1024; }

struct TileLayout {
  LloydState* st;
  float* cbuf;
  long long* totals;
  long long* partials;
  float* fcum;
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
  L.partials = reinterpret_cast<long long*>(
      take((size_t)tiles_max_blocks(k) * 4 * k * 8));
  L.fcum = reinterpret_cast<float*>(take((size_t)k * sizeof(float)));
  L.maxabs = reinterpret_cast<unsigned*>(take(sizeof(unsigned)));
  L.pstate = reinterpret_cast<uint32_t*>(take((size_t)n * sizeof(uint32_t)));
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

int launch_absmax(const float* pts, int64_t count, unsigned* out,
                  hipStream_t stream) {
  const int vec = ((uintptr_t)pts & 15u) == 0;
  hipLaunchKernelGGL(tiles_absmax, dim3(8 * device_cu_count()), dim3(256), 0,
                     stream, pts, count, vec, out);
  SPLIT_HIP_CHECK_LAUNCH("absmax");
  return SPLIT_HIP_OK;
}

size_t tiles_workspace(int64_t n, int k) {
  return tiles_layout(nullptr, n, k).bytes;
}

bool tiles_supported(int width, int k, int max_iter) {
  (void)max_iter;
  return width > 0 && k >= 1 && k <= kMaxTileK;
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
                     centroids, k, max_iter, threshold, L.totals, L.fcum,
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
  a.partials = L.partials;
  a.fcum = L.fcum;
  a.pstate = L.pstate;
  a.maxabs = L.maxabs;
  a.state = L.st;
  a.mbox = res->dev;
  a.units_per_row = (a.tiles_x + 1) / 2;
  a.nunits = (int64_t)a.units_per_row * ((height + kTileSide - 1) / kTileSide);
  a.unit_mul = spread_multiplier(a.nunits);

  const int threads = tiles_threads(k);
  const int nwaves = threads / kWave;
  const int64_t want = (a.nseg + nwaves - 1) / nwaves;
  const size_t lds = (size_t)nwaves * 4 * k * 8 + kMaxTileK * sizeof(float) +
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
      1, std::min<int64_t>({want, (int64_t)device_cu_count() * per_cu,
                            (int64_t)tiles_max_blocks(k)}));
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
    hipLaunchKernelGGL(tiles_reduce, dim3(1), dim3(kReduceThreads), 0, stream,
                       a, grid);
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
