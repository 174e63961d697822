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
#include "ccl_device.hpp"

namespace splithip {


template <int CONN>
__global__ __launch_bounds__(kCcThreads) void ccl_tile(
    const int32_t* __restrict__ labels, int W, int H, int32_t* out,
    unsigned* flag) {
  __shared__ int lab[kCcTilePx];
  __shared__ int par[kCcTilePx];
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *flag = 0u;
  ccl_tile_body<CONN>(labels, W, H, out, blockIdx.x, blockIdx.y, lab, par);
}

template <int CONN>
__global__ __launch_bounds__(3 * kCcTile) void ccl_merge(
    const int32_t* __restrict__ labels, int W, int H, int32_t* out) {
  ccl_merge_body<CONN>(labels, W, H, out, blockIdx.x, blockIdx.y, threadIdx.x);
}

// Every pixel follows its chain to the root.  With `roots` non-null, wave w
// also writes the presence bitmap of roots for its 64 pixels (bit p set iff
// pixel p is a root), which is the value set compress_labels needs.
__global__ void ccl_jump(int32_t* out, int64_t n, int W) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x)
    ccl_jump_one(out, p, W);
}

template <int CONN>
__global__ void ccl_check(const int32_t* __restrict__ labels, int W, int H,
                          int32_t* out, unsigned* flag) {
  const int64_t n = (int64_t)W * H;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x)
    ccl_check_one<CONN>(labels, W, H, out, p, flag);
}

// flag == null: always; otherwise only when the check pass found a failure.
__global__ void ccl_flatten(int32_t* out, int64_t n, uint32_t* roots,
                            const unsigned* flag) {
  if (flag && *flag == 0u) return;
  const int lane = threadIdx.x & (kWave - 1);
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = base + threadIdx.x;
    const bool root = p < n && ccl_flatten_one(out, p) == p;
    const uint64_t m = ballot(root);
    if (roots && lane == 0 && p < n) {
      const int64_t w = p >> 5;  // p (lane 0) is a multiple of 64
      *reinterpret_cast<uint2*>(roots + w) =
          make_uint2((uint32_t)m, (uint32_t)(m >> 32));
    }
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
  unsigned ccl_flag;  // connected-component check failures (segment path)
};

constexpr size_t kCompressHead = 64;
constexpr size_t kCcTempBytes = 256;  // the check pass's failure flag
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

// Labels are read four at a time (int4) when the array is 16-byte aligned;
// `vec` says so.  Element i of chunk c is labels[4c + i].
__device__ __forceinline__ void load4(const int32_t* v, int64_t n, int64_t c,
                                      bool vec, int (&x)[4], bool (&ok)[4]) {
  const int64_t i0 = 4 * c;
  if (vec && i0 + 3 < n) {
    const int4 q = *reinterpret_cast<const int4*>(v + i0);
    x[0] = q.x; x[1] = q.y; x[2] = q.z; x[3] = q.w;
#pragma unroll
    for (int i = 0; i < 4; ++i) ok[i] = true;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ok[i] = i0 + i < n;
      x[i] = ok[i] ? v[i0 + i] : 0;
    }
  }
}

__global__ __launch_bounds__(256) void compress_minmax(const int32_t* v,
                                                       int64_t n, int vec,
                                                       CompressState* st) {
  int lo = INT_MAX, hi = INT_MIN;
  const int64_t chunks = (n + 3) / 4;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < chunks;
       c += (int64_t)gridDim.x * blockDim.x) {
    int x[4];
    bool ok[4];
    load4(v, n, c, vec, x, ok);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      lo = ok[i] ? min(lo, x[i]) : lo;
      hi = ok[i] ? max(hi, x[i]) : hi;
    }
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

__global__ void compress_mark(const int32_t* v, int64_t n, int vec,
                              const CompressState* st, unsigned* bm) {
  if (st->err) return;
  const int lo = st->lo;
  const int64_t chunks = (n + 3) / 4;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < chunks;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = base + threadIdx.x;
    int x[4];
    bool ok[4];
    load4(v, n, c < chunks ? c : chunks - 1, vec, x, ok);
    // skip a value equal to the one before it (runs are common)
    const int prev = __shfl_up(x[3], 1);
    const bool first_lane = (threadIdx.x & (kWave - 1)) == 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int before = i == 0 ? prev : x[i - 1];
      const bool lead = (i == 0 && first_lane) || before != x[i];
      if (c < chunks && ok[i] && lead) {
        const uint32_t d = (uint32_t)((int64_t)x[i] - lo);
        // test first: values repeat, and an atomic on a hot word serialises
        const unsigned bit = 1u << (d & 31);
        if (!(__hip_atomic_load(&bm[d >> 5], __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT) & bit))
          atomicOr(&bm[d >> 5], bit);
      }
    }
  }
}

// Value set known to be the roots of a connected-component labelling of n
// pixels: values lie in [0, n) and the bitmap was written by ccl_flatten.
__global__ void compress_init_roots(CompressState* st, int64_t n) {
  st->lo = 0;
  st->hi = (int)(n - 1);
  st->arrive = 0u;
  st->err = 0;
  st->words = (n + 31) / 32;
  st->total = 0;
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
                              int32_t* out, int conn, hipStream_t stream,
                              uint32_t* roots, unsigned* flag) {
  if (!labels || !out || !flag)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "labels, out or temp is null");
  if (W < 1 || H < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "image is %d x %d", W, H);
  if ((int64_t)W * H > INT_MAX)
    return fail(SPLIT_HIP_ERR_UNSUPPORTED,
                "%d x %d pixels exceed the int32 label range", W, H);
  if (conn != 4 && conn != 8)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "connectivity %d is not 4 or 8",
                conn);
  const int ntx = (W + kCcTile - 1) / kCcTile, nty = (H + kCcTile - 1) / kCcTile;
  if (nty > 65535)
    return fail(SPLIT_HIP_ERR_UNSUPPORTED, "height %d too large", H);
  const dim3 tiles(ntx, nty);
  const int64_t n = (int64_t)W * H;
  const dim3 pgrid(grid_for(n, 256, 8));
  if (conn == 8) {
    hipLaunchKernelGGL(ccl_tile<8>, tiles, dim3(kCcThreads), 0, stream,
                       labels, W, H, out, flag);
    hipLaunchKernelGGL(ccl_merge<8>, tiles, dim3(3 * kCcTile), 0, stream,
                       labels, W, H, out);
  } else {
    hipLaunchKernelGGL(ccl_tile<4>, tiles, dim3(kCcThreads), 0, stream,
                       labels, W, H, out, flag);
    hipLaunchKernelGGL(ccl_merge<4>, tiles, dim3(3 * kCcTile), 0, stream,
                       labels, W, H, out);
  }
  SPLIT_HIP_CHECK_LAUNCH("ccl_tile/ccl_merge");
  hipLaunchKernelGGL(ccl_jump, pgrid, dim3(256), 0, stream, out, n, W);
  hipLaunchKernelGGL(ccl_flatten, pgrid, dim3(256), 0, stream, out, n, roots,
                     (const unsigned*)nullptr);
  if (conn == 8)
    hipLaunchKernelGGL(ccl_check<8>, pgrid, dim3(256), 0, stream, labels, W,
                       H, out, flag);
  else
    hipLaunchKernelGGL(ccl_check<4>, pgrid, dim3(256), 0, stream, labels, W,
                       H, out, flag);
  hipLaunchKernelGGL(ccl_flatten, pgrid, dim3(256), 0, stream, out, n, roots,
                     (const unsigned*)flag);
  SPLIT_HIP_CHECK_LAUNCH("ccl_flatten/ccl_check");
  return SPLIT_HIP_OK;
}

// Temp block of compress_labels: [state][chunk sums][bitmap][word prefix].
struct CompressLayout {
  CompressState* st;
  unsigned* csum;
  unsigned* bm;
  unsigned* prefix;
  int64_t cap;  // bitmap words the block holds
};

static int compress_layout(void* temp, size_t temp_bytes, int64_t n,
                           CompressLayout* L) {
  if (!temp) return fail(SPLIT_HIP_ERR_ARGUMENT, "temp is null");
  const size_t min_need = split_hip_compress_labels_workspace(n, 1);
  if (temp_bytes < min_need)
    return fail(SPLIT_HIP_ERR_TEMP, "temp holds %zu bytes, needs >= %zu",
                temp_bytes, min_need);
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
  char* t = static_cast<char*>(temp);
  L->cap = lo;
  L->st = reinterpret_cast<CompressState*>(t);
  L->csum = reinterpret_cast<unsigned*>(t + kCompressHead);
  L->bm = reinterpret_cast<unsigned*>(t + kCompressHead +
                                      compress_csum_bytes(lo));
  L->prefix = L->bm + align_up((size_t)lo * 4, 16) / 4;
  return SPLIT_HIP_OK;
}

static int compress_rank(const CompressLayout& L, int32_t* labels, int64_t n,
                         hipStream_t stream) {
  hipLaunchKernelGGL(compress_scan, dim3(grid_for(L.cap, kScanChunk, 1)),
                     dim3(kScanThreads), 0, stream, L.st, L.bm, L.prefix,
                     L.csum);
  hipLaunchKernelGGL(compress_map, dim3(grid_for(n, 256, 8)), dim3(256), 0,
                     stream, labels, n, L.st, L.bm, L.prefix, L.csum);
  SPLIT_HIP_CHECK_LAUNCH("compress scan/map");
  return SPLIT_HIP_OK;
}

int compress_labels_async(int32_t* labels, int64_t n, void* temp,
                          size_t temp_bytes, hipStream_t stream,
                          CompressState** st_out) {
  if (!labels) return fail(SPLIT_HIP_ERR_ARGUMENT, "labels is null");
  if (n < 1) return fail(SPLIT_HIP_ERR_ARGUMENT, "n = %lld < 1", (long long)n);
  CompressLayout L;
  SPLIT_HIP_TRY(compress_layout(temp, temp_bytes, n, &L));
  const int vec = ((uintptr_t)labels & 15u) == 0;
  const int64_t chunks = (n + 3) / 4;
  hipLaunchKernelGGL(compress_init, dim3(1), dim3(1), 0, stream, L.st);
  hipLaunchKernelGGL(compress_minmax, dim3(grid_for(chunks, 256, 8)),
                     dim3(256), 0, stream, labels, n, vec, L.st);
  hipLaunchKernelGGL(compress_clear, dim3(grid_for(L.cap, 256, 8)), dim3(256),
                     0, stream, L.st, L.bm, L.cap);
  hipLaunchKernelGGL(compress_mark, dim3(grid_for(chunks, 256, 8)), dim3(256),
                     0, stream, labels, n, vec, L.st, L.bm);
  SPLIT_HIP_CHECK_LAUNCH("compress kernels");
  SPLIT_HIP_TRY(compress_rank(L, labels, n, stream));
  *st_out = L.st;
  return SPLIT_HIP_OK;
}

// Connected components of `labels` into `segments`, then their dense ranks,
// with the root bitmap written by the flatten pass (no min/max, clear or
// mark pass).
int segments_async(const int32_t* labels, int W, int H, int32_t* segments,
                   int conn, bool compress, void* temp, size_t temp_bytes,
                   hipStream_t stream, CompressState** st_out) {
  const int64_t n = (int64_t)W * H;
  if (!compress) {
    *st_out = nullptr;
    return connected_components_impl(labels, W, H, segments, conn, stream,
                                     nullptr, static_cast<unsigned*>(temp));
  }
  CompressLayout L;
  SPLIT_HIP_TRY(compress_layout(temp, temp_bytes, n, &L));
  const int64_t words = 2 * ((n + 63) / 64);
  if (L.cap < words)
    return fail(SPLIT_HIP_ERR_TEMP, "temp holds a bitmap of %lld words, %lld"
                " needed", (long long)L.cap, (long long)words);
  SPLIT_HIP_TRY(connected_components_impl(labels, W, H, segments, conn,
                                          stream, L.bm, &L.st->ccl_flag));
  hipLaunchKernelGGL(compress_init_roots, dim3(1), dim3(1), 0, stream, L.st,
                     n);
  SPLIT_HIP_CHECK_LAUNCH("compress_init_roots");
  SPLIT_HIP_TRY(compress_rank(L, segments, n, stream));
  *st_out = L.st;
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
  return kCcTempBytes;
}

int split_hip_connected_components(const int32_t* labels, int32_t width,
                                   int32_t height, void* temp,
                                   size_t temp_bytes, int32_t* out, int32_t,
                                   int32_t connectivity,
                                   split_hip_stream_t stream) {
  if (!labels || !out)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "labels or out is null");
  if (!temp || temp_bytes < kCcTempBytes)
    return fail(SPLIT_HIP_ERR_TEMP,
                "temp holds %zu bytes, connected_components needs %zu",
                temp ? temp_bytes : (size_t)0, kCcTempBytes);
  return connected_components_impl(labels, width, height, out, connectivity,
                                   stream, nullptr,
                                   static_cast<unsigned*>(temp));
}

size_t split_hip_compress_labels_workspace(int64_t n, int64_t value_span) {
  if (value_span < 1) value_span = 1;
  (void)n;
  // whole 64-bit words, as the flatten pass writes the root bitmap in pairs
  const int64_t words = 2 * ((value_span + 63) / 64);
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
