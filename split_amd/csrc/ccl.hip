// Connected-component labelling and dense label compression on gfx950.
//
// connected_components replaces the reference's iterated max-propagation
// over horizontal then vertical same-label runs
// (split/src/split/device/ccl/connected_components.cu:80-181), which needs
// about one iteration per turn of a spiral and a device->host read per
// iteration.  Here a three-kernel union-find produces the same fixed point —
// every pixel labelled with the LARGEST linear index in its component — in a
// bounded number of passes:
//
//  1. ccl_tile:   one 256-thread block per 32x32 tile; labels staged in LDS,
//                 lock-free union-find in LDS (atomicMax links a smaller root
//                 under a larger one, so a root is its tree's largest index),
//                 then each pixel's global root index is written to `out`.
//  2. ccl_merge:  pixels on tile edges union with equal-label neighbours in
//                 other tiles, directly on the `out` forest (global atomicMax
//                 with retry, as in Playne & Hawick's block-merge scheme).
//  3. ccl_flatten: every pixel follows its chain to the root.
//
// The parent array is the output buffer itself, so no temp is needed.
//
// compress_labels replaces copy/sort/unique/lower_bound
// (split/src/split/device/ccl/compress_labels.cu:10-35) with a presence
// bitmap over [min, max], a popcount prefix scan and a rank lookup: the same
// dense rank, computed without sorting.
#include <limits.h>

#include "common.hpp"

namespace splithip {

constexpr int kTile = 32;
constexpr int kTilePx = kTile * kTile;
constexpr int kTileThreads = 256;

__device__ __forceinline__ int ld_rlx(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------- LDS forest ---

__device__ __forceinline__ int lds_find(const volatile int* par, int i) {
  for (;;) {
    const int p = par[i];
    if (p == i) return i;
    i = p;
  }
}

__device__ __forceinline__ void lds_union(int* par, int a, int b) {
  for (;;) {
    a = lds_find(par, a);
    b = lds_find(par, b);
    if (a == b) return;
    if (a > b) { const int t = a; a = b; b = t; }
    const int old = atomicMax(&par[a], b);
    if (old == a) return;
    a = old;
  }
}

template <int CONN>
__global__ __launch_bounds__(kTileThreads) void ccl_tile(
    const int32_t* __restrict__ labels, int W, int H, int32_t* out) {
  __shared__ int lab[kTilePx];
  __shared__ int par[kTilePx];
  const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
#pragma unroll
  for (int j = 0; j < kTilePx / kTileThreads; ++j) {
    const int li = threadIdx.x + kTileThreads * j;
    const int gx = x0 + (li & (kTile - 1)), gy = y0 + li / kTile;
    const bool in = gx < W && gy < H;
    lab[li] = in ? labels[(int64_t)gy * W + gx] : 0;
    par[li] = li;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kTilePx / kTileThreads; ++j) {
    const int li = threadIdx.x + kTileThreads * j;
    const int lx = li & (kTile - 1), ly = li / kTile;
    const int gx = x0 + lx, gy = y0 + ly;
    if (gx >= W || gy >= H) continue;
    const int l = lab[li];
    if (lx > 0 && lab[li - 1] == l) lds_union(par, li, li - 1);
    if (ly > 0) {
      if (lab[li - kTile] == l) lds_union(par, li, li - kTile);
      if (CONN == 8) {
        if (lx > 0 && lab[li - kTile - 1] == l)
          lds_union(par, li, li - kTile - 1);
        if (lx < kTile - 1 && gx + 1 < W && lab[li - kTile + 1] == l)
          lds_union(par, li, li - kTile + 1);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kTilePx / kTileThreads; ++j) {
    const int li = threadIdx.x + kTileThreads * j;
    const int gx = x0 + (li & (kTile - 1)), gy = y0 + li / kTile;
    if (gx >= W || gy >= H) continue;
    const int r = lds_find(par, li);
    out[(int64_t)gy * W + gx] =
        (y0 + r / kTile) * W + x0 + (r & (kTile - 1));
  }
}

// ------------------------------------------------------- global forest ---

__device__ __forceinline__ int g_find(const int* par, int i) {
  for (;;) {
    const int p = ld_rlx(par + i);
    if (p == i) return i;
    i = p;
  }
}

__device__ __forceinline__ void g_union(int* par, int a, int b) {
  for (;;) {
    a = g_find(par, a);
    b = g_find(par, b);
    if (a == b) return;
    if (a > b) { const int t = a; a = b; b = t; }
    const int old = atomicMax(par + a, b);
    if (old == a) return;
    a = old;
  }
}

// One block per tile; thread t examines one pixel on the tile's top row,
// left column or (8-connectivity) right column and unions it with each
// equal-label neighbour that lies in another tile.
template <int CONN>
__global__ __launch_bounds__(3 * kTile) void ccl_merge(
    const int32_t* __restrict__ labels, int W, int H, int32_t* out) {
  const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
  const int t = threadIdx.x;
  int gx, gy;
  if (t < kTile) {            // top row
    gx = x0 + t; gy = y0;
  } else if (t < 2 * kTile) { // left column
    gx = x0; gy = y0 + t - kTile;
  } else {                    // right column (up-right neighbours)
    if (CONN != 8) return;
    gx = x0 + kTile - 1; gy = y0 + t - 2 * kTile;
  }
  if (gx >= W || gy >= H) return;
  const int p = gy * W + gx;
  const int l = labels[p];
  const int tx = gx / kTile, ty = gy / kTile;
  // "before" neighbours: left, up, and for 8-connectivity up-left, up-right
  const int ndx[4] = {-1, 0, -1, 1};
  const int ndy[4] = {0, -1, -1, -1};
  const int nn = CONN == 8 ? 4 : 2;
  for (int q = 0; q < nn; ++q) {
    const int nx = gx + ndx[q], ny = gy + ndy[q];
    if (nx < 0 || ny < 0 || nx >= W) continue;
    if (nx / kTile == tx && ny / kTile == ty) continue;
    const int np = ny * W + nx;
    if (labels[np] == l) g_union(out, p, np);
  }
}

__global__ void ccl_flatten(int32_t* out, int64_t n) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int r0 = out[p];
    int r = r0;
    for (;;) {
      const int q = ld_rlx(out + r);
      if (q == r) break;
      r = q;
    }
    if (r != r0) out[p] = r;
  }
}

// ------------------------------------------------------------ compress ---

struct CompressState {
  int lo;
  int hi;
  unsigned arrive;
  int err;          // 1: value span exceeds the bitmap capacity of temp
  int64_t words;    // bitmap words in use
  int64_t total;    // distinct values
};

constexpr size_t kCompressHead = 64;
constexpr int kScanThreads = 256;
constexpr int kScanWordsPerThread = 16;
constexpr int kScanChunk = kScanThreads * kScanWordsPerThread;  // 4096 words

__global__ void compress_init(CompressState* st) {
  st->lo = INT_MAX;
  st->hi = INT_MIN;
  st->arrive = 0u;
  st->err = 0;
  st->words = 0;
  st->total = 0;
}

__device__ __forceinline__ int wave_min_i(int x) {
  x = min(x, dpp_i<kDppQuadXor1>(x));
  x = min(x, dpp_i<kDppQuadXor2>(x));
  x = min(x, dpp_i<kDppRowHalfMirror>(x));
  x = min(x, dpp_i<kDppRowMirror>(x));
  x = min(x, dpp_i<kDppRowBcast15, 0xA>(x, INT_MAX));
  x = min(x, dpp_i<kDppRowBcast31, 0xC>(x, INT_MAX));
  return readlane_i(x, 63);
}
__device__ __forceinline__ int wave_max_i(int x) {
  x = max(x, dpp_i<kDppQuadXor1>(x));
  x = max(x, dpp_i<kDppQuadXor2>(x));
  x = max(x, dpp_i<kDppRowHalfMirror>(x));
  x = max(x, dpp_i<kDppRowMirror>(x));
  x = max(x, dpp_i<kDppRowBcast15, 0xA>(x, INT_MIN));
  x = max(x, dpp_i<kDppRowBcast31, 0xC>(x, INT_MIN));
  return readlane_i(x, 63);
}

__global__ __launch_bounds__(256) void compress_minmax(const int32_t* v,
                                                       int64_t n,
                                                       CompressState* st) {
  int lo = INT_MAX, hi = INT_MIN;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int x = v[i];
    lo = min(lo, x);
    hi = max(hi, x);
  }
  lo = wave_min_i(lo);
  hi = wave_max_i(hi);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicMin(&st->lo, lo);
    atomicMax(&st->hi, hi);
  }
}

__global__ void compress_clear(CompressState* st, unsigned* bm,
                               int64_t cap_words) {
  const int64_t words = (((int64_t)st->hi - st->lo + 1) + 31) / 32;
  if (words > cap_words) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->err = 1;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->words = words;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (int64_t)gridDim.x * blockDim.x)
    bm[i] = 0u;
}

__global__ void compress_mark(const int32_t* v, int64_t n,
                              const CompressState* st, unsigned* bm) {
  if (st->err) return;
  const int lo = st->lo;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    const int x = i < n ? v[i] : lo;
    const int prev = __shfl_up(x, 1);
    const bool lead = (threadIdx.x & (kWave - 1)) == 0 || prev != x;
    if (i < n && lead) {
      const uint32_t d = (uint32_t)((int64_t)x - lo);
      atomicOr(&bm[d >> 5], 1u << (d & 31));
    }
  }
}

// Per 4096-word chunk: exclusive popcount prefix of each word inside the
// chunk, the chunk total into csum[]; the last block scans csum[] in place.
__global__ __launch_bounds__(kScanThreads) void compress_scan(
    CompressState* st, const unsigned* bm, unsigned* prefix, unsigned* csum) {
  if (st->err) return;
  const int64_t words = st->words;
  const int64_t nchunks = (words + kScanChunk - 1) / kScanChunk;
  __shared__ unsigned s[kScanThreads];
  __shared__ unsigned ticket;
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int64_t w0 = c * kScanChunk + (int64_t)threadIdx.x * kScanWordsPerThread;
    unsigned cnt[kScanWordsPerThread];
    unsigned tsum = 0;
#pragma unroll
    for (int j = 0; j < kScanWordsPerThread; ++j) {
      const int64_t w = w0 + j;
      cnt[j] = w < words ? __popc(bm[w]) : 0u;
      tsum += cnt[j];
    }
    s[threadIdx.x] = tsum;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
      const unsigned add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    unsigned run = s[threadIdx.x] - tsum;  // exclusive
#pragma unroll
    for (int j = 0; j < kScanWordsPerThread; ++j) {
      const int64_t w = w0 + j;
      if (w < words) prefix[w] = run;
      run += cnt[j];
    }
    if (threadIdx.x == kScanThreads - 1) csum[c] = s[threadIdx.x];
    __syncthreads();
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) ticket = atomicAdd(&st->arrive, 1u);
  __syncthreads();
  if (ticket != gridDim.x - 1) return;
  __threadfence();
  // Last block: exclusive scan of the chunk totals, in place.
  const int64_t per = (nchunks + kScanThreads - 1) / kScanThreads;
  const int64_t b0 = (int64_t)threadIdx.x * per;
  uint64_t tsum = 0;
  for (int64_t j = b0; j < b0 + per && j < nchunks; ++j) tsum += csum[j];
  __shared__ uint64_t s64[kScanThreads];
  s64[threadIdx.x] = tsum;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint64_t add = threadIdx.x >= off ? s64[threadIdx.x - off] : 0u;
    __syncthreads();
    s64[threadIdx.x] += add;
    __syncthreads();
  }
  uint64_t run = s64[threadIdx.x] - tsum;
  for (int64_t j = b0; j < b0 + per && j < nchunks; ++j) {
    const unsigned c = csum[j];
    csum[j] = (unsigned)run;
    run += c;
  }
  if (threadIdx.x == kScanThreads - 1) {
    st->total = (int64_t)s64[threadIdx.x];
    st->arrive = 0u;
  }
}

__global__ void compress_map(int32_t* v, int64_t n, const CompressState* st,
                             const unsigned* bm, const unsigned* prefix,
                             const unsigned* csum) {
  if (st->err) return;
  const int lo = st->lo;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t d = (uint32_t)((int64_t)v[i] - lo);
    const uint32_t w = d >> 5;
    const unsigned below = bm[w] & ((1u << (d & 31)) - 1u);
    v[i] = (int32_t)(csum[w / kScanChunk] + prefix[w] + __popc(below));
  }
}

static int grid_for(int64_t n, int threads, int per_cu) {
  const int64_t want = (n + threads - 1) / threads;
  const int64_t cap = (int64_t)device_cu_count() * per_cu;
  return (int)std::max<int64_t>(1, std::min(want, cap));
}

static size_t compress_csum_bytes(int64_t words) {
  return align_up(((words + kScanChunk - 1) / kScanChunk + 1) * 4, 16);
}

int connected_components_impl(const int32_t* labels, int W, int H,
                              int32_t* out, int conn, hipStream_t stream) {
  if (!labels || !out)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "labels or out is null");
  if (W < 1 || H < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "image is %d x %d", W, H);
  if ((int64_t)W * H > INT_MAX)
    return fail(SPLIT_HIP_ERR_UNSUPPORTED,
                "%d x %d pixels exceed the int32 label range", W, H);
  if (conn != 4 && conn != 8)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "connectivity %d is not 4 or 8",
                conn);
  const int ntx = (W + kTile - 1) / kTile, nty = (H + kTile - 1) / kTile;
  if (nty > 65535)
    return fail(SPLIT_HIP_ERR_UNSUPPORTED, "height %d too large", H);
  const dim3 tiles(ntx, nty);
  if (conn == 8) {
    hipLaunchKernelGGL(ccl_tile<8>, tiles, dim3(kTileThreads), 0, stream,
                       labels, W, H, out);
    hipLaunchKernelGGL(ccl_merge<8>, tiles, dim3(3 * kTile), 0, stream,
                       labels, W, H, out);
  } else {
    hipLaunchKernelGGL(ccl_tile<4>, tiles, dim3(kTileThreads), 0, stream,
                       labels, W, H, out);
    hipLaunchKernelGGL(ccl_merge<4>, tiles, dim3(3 * kTile), 0, stream,
                       labels, W, H, out);
  }
  SPLIT_HIP_CHECK_LAUNCH("ccl_tile/ccl_merge");
  const int64_t n = (int64_t)W * H;
  hipLaunchKernelGGL(ccl_flatten, dim3(grid_for(n, 256, 8)), dim3(256), 0,
                     stream, out, n);
  SPLIT_HIP_CHECK_LAUNCH("ccl_flatten");
  return SPLIT_HIP_OK;
}

int compress_labels_async(int32_t* labels, int64_t n, void* temp,
                          size_t temp_bytes, hipStream_t stream,
                          CompressState** st_out) {
  if (!labels || !temp)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "labels or temp is null");
  if (n < 1) return fail(SPLIT_HIP_ERR_ARGUMENT, "n = %lld < 1", (long long)n);
  const size_t min_need = split_hip_compress_labels_workspace(n, 1);
  if (temp_bytes < min_need)
    return fail(SPLIT_HIP_ERR_TEMP, "temp holds %zu bytes, needs >= %zu",
                temp_bytes, min_need);
  // Largest bitmap (in 32-bit words) the temp block holds next to its
  // prefix array and chunk sums.
  auto fits = [&](int64_t w) {
    return kCompressHead + compress_csum_bytes(w) +
               2 * align_up((size_t)w * 4, 16) <=
           temp_bytes;
  };
  int64_t lo = 1, hi = (int64_t)(temp_bytes / 8) + 1;  // fits(lo), !fits(hi)
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (fits(mid))
      lo = mid;
    else
      hi = mid;
  }
  const int64_t cap = lo;
  char* t = static_cast<char*>(temp);
  CompressState* st = reinterpret_cast<CompressState*>(t);
  unsigned* csum = reinterpret_cast<unsigned*>(t + kCompressHead);
  unsigned* bm = reinterpret_cast<unsigned*>(t + kCompressHead +
                                             compress_csum_bytes(cap));
  unsigned* prefix = bm + align_up((size_t)cap * 4, 16) / 4;

  hipLaunchKernelGGL(compress_init, dim3(1), dim3(1), 0, stream, st);
  hipLaunchKernelGGL(compress_minmax, dim3(grid_for(n, 256, 8)), dim3(256), 0,
                     stream, labels, n, st);
  hipLaunchKernelGGL(compress_clear, dim3(grid_for(cap, 256, 8)), dim3(256), 0,
                     stream, st, bm, cap);
  hipLaunchKernelGGL(compress_mark, dim3(grid_for(n, 256, 8)), dim3(256), 0,
                     stream, labels, n, st, bm);
  hipLaunchKernelGGL(compress_scan,
                     dim3(grid_for(cap, kScanChunk, 1)),
                     dim3(kScanThreads), 0, stream, st, bm, prefix, csum);
  hipLaunchKernelGGL(compress_map, dim3(grid_for(n, 256, 8)), dim3(256), 0,
                     stream, labels, n, st, bm, prefix, csum);
  SPLIT_HIP_CHECK_LAUNCH("compress kernels");
  *st_out = st;
  return SPLIT_HIP_OK;
}

int compress_labels_finish(CompressState* st, int32_t* num_unique,
                           hipStream_t stream) {
  CompressState h{};
  SPLIT_HIP_CHECK(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost,
                                 stream));
  SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
  if (h.err)
    return fail(SPLIT_HIP_ERR_RANGE,
                "label values span [%d, %d]; temp holds a bitmap for fewer",
                h.lo, h.hi);
  if (num_unique) *num_unique = (int32_t)h.total;
  return SPLIT_HIP_OK;
}

}  // namespace splithip

using namespace splithip;

extern "C" {

size_t split_hip_connected_components_workspace(int32_t, int32_t) {
  return 0;
}

int split_hip_connected_components(const int32_t* labels, int32_t width,
                                   int32_t height, void*, size_t,
                                   int32_t* out, int32_t, int32_t connectivity,
                                   split_hip_stream_t stream) {
  return connected_components_impl(labels, width, height, out, connectivity,
                                   stream);
}

size_t split_hip_compress_labels_workspace(int64_t n, int64_t value_span) {
  if (value_span < 1) value_span = 1;
  (void)n;
  const int64_t words = (value_span + 31) / 32;
  return kCompressHead + compress_csum_bytes(words) +
         2 * align_up((size_t)words * 4, 16);
}

int split_hip_compress_labels(int32_t* labels, int64_t n, void* temp,
                              size_t temp_bytes, int32_t* num_unique,
                              split_hip_stream_t stream) {
  CompressState* st = nullptr;
  SPLIT_HIP_TRY(compress_labels_async(labels, n, temp, temp_bytes, stream, &st));
  return compress_labels_finish(st, num_unique, stream);
}

}  // extern "C"
