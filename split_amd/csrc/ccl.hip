// Connected-component labelling and dense label compression on gfx950.
//
// connected_components replaces the reference's iterated max-propagation
// over horizontal then vertical same-label runs
// (split/src/split/device/ccl/connected_components.cu:80-181), which needs
// about one iteration per turn of a spiral and a device->host read per
// iteration.  Here a union-find produces the same fixed point — every pixel
// labelled with the LARGEST linear index in its component — in a fixed
// number of passes over 64x64 tiles:
//
//  1. cc_tile:  one 1024-thread block per tile (wave w: rows w, w+16, w+32,
//               w+48; lane = column).  Each pixel links to the end of its
//               row run (ballot), vertical/diagonal unions between runs in
//               an LDS forest (atomicMax puts the smaller root under the
//               larger, so a root is the largest index of its tree), one
//               find per run; every pixel writes its local root's global
//               index into `out`.  The tile's left and right columns
//               (labels and local roots) go to compact edge arrays, so the
//               merge pass reads no image columns.
//  2. cc_merge: the tile seams: top row against the row above, left column
//               against the left tile's right column, and (8-connectivity)
//               the up-right diagonal of the right column; unions on the
//               global forest, deduplicated along runs.  A root that a union
//               links under another is marked in the linked-root bitmap.
//  3. cc_jump:  every marked root points at its root by pointer jumping
//               (its word follows its parent's current word), ancestors
//               first.  Only a root's own thread writes its word.
//  4. cc_scan:  (with compression) roots = local roots not linked; popcount
//               prefix of the root bitmap.
//  5. cc_final: per tile, from the tile's own forest words staged in LDS: a
//               pixel's root is its word (a root or a jumped root), or the
//               word of its in-tile local root; writes the root index, or its
//               rank among the roots (dense labels, compress_labels).
//
// compress_labels replaces copy/sort/unique/lower_bound
// (split/src/split/device/ccl/compress_labels.cu:10-35) with a presence
// bitmap over [min, max], a popcount prefix scan and a rank lookup: the same
// dense rank, computed without sorting.
#include <limits.h>

#include <algorithm>

#include "common.hpp"
#include "ccl_device.hpp"

namespace splithip {

template <int CONN, int S>
__global__ __launch_bounds__(CcTile<S>::kThreads) void cc_tile(CcArgs a) {
  __shared__ int lab[CcTile<S>::kPx];
  __shared__ int par[CcTile<S>::kPx];
  cc_tile_body<CONN, S>(a, blockIdx.x, blockIdx.y, lab, par);
}

template <int CONN, int S>
__global__ __launch_bounds__(3 * S) void cc_merge(CcArgs a) {
  cc_merge_body<CONN, S>(a, blockIdx.x, blockIdx.y);
}

// Words in decreasing order.  Only a root's own thread writes its word; on
// the 8192^2 spiral image (808k linked roots, chains of up to ~7,000
// roots): 653 us; a path-halving walk that writes the root at its end took
// 3,457 us in increasing order and 1,111 us in decreasing order, pointer
// jumping in increasing order 5,270 us.  On the 4096^2 k = 32 segment
// labels: 20 us (walk: 28 us).
__global__ __launch_bounds__(256) void cc_jump(int32_t* out,
                                               const unsigned* linked,
                                               int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (int64_t)gridDim.x * blockDim.x)
    cc_jump_word(out, linked, words - 1 - i);
}

template <int S>
__global__ __launch_bounds__(CcTile<S>::kThreads) void cc_final(
    CcArgs a, const unsigned* rb, const unsigned* prefix,
    const unsigned* csum) {
  __shared__ int sv[CcTile<S>::kPx];
  cc_final_body<S>(a, blockIdx.x, blockIdx.y, rb, prefix, csum, sv);
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
constexpr int kScanChunk = kCcScanChunk;  // 4 words per thread

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

// One 1024-thread block per CU, four 16-byte loads in flight per thread,
// one atomic pair per block: with one pair per wave the 16k atomics on the
// two words serialised (216 us for 67M labels).
__global__ __launch_bounds__(1024) void compress_minmax(const int32_t* v,
                                                        int64_t n, int vec,
                                                        CompressState* st) {
  int lo = INT_MAX, hi = INT_MIN;
  const int64_t chunks = (n + 3) / 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c0 < chunks;
       c0 += 4 * stride) {
    int x[4][4];
    bool ok[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t c = c0 + u * stride;
      if (c < chunks) {
        load4(v, n, c, vec, x[u], ok[u]);
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) ok[u][i] = false;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        lo = ok[u][i] ? min(lo, x[u][i]) : lo;
        hi = ok[u][i] ? max(hi, x[u][i]) : hi;
      }
  }
  lo = wave_min_i(lo);
  hi = wave_max_i(hi);
  __shared__ int slo[16], shi[16];
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    slo[wave] = lo;
    shi[wave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x / kWave); ++w) {
      lo = min(lo, slo[w]);
      hi = max(hi, shi[w]);
    }
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

// Sets the bit of every value.  A wave first takes up to four distinct
// values of its 256 one at a time (ballot of equality: labels come in runs,
// and the stress images' few components alternate pixel by pixel), one lane
// marking each; values left after that are marked per lane, skipping a
// value equal to the one before it.  A mark tests the word with a plain
// load before the atomicOr (a stale clear bit only repeats the OR): with
// every pixel loading its word at agent scope, the few hot words of an
// image with 4 components took 496 us for 67M labels.
__device__ __forceinline__ void mark_value(unsigned* bm, int lo, int x) {
  const uint32_t d = (uint32_t)((int64_t)x - lo);
  const unsigned bit = 1u << (d & 31);
  if (!(bm[d >> 5] & bit)) atomicOr(&bm[d >> 5], bit);
}

__global__ void compress_mark(const int32_t* v, int64_t n, int vec,
                              const CompressState* st, unsigned* bm) {
  if (st->err) return;
  const int lo = st->lo;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t chunks = (n + 3) / 4;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < chunks;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = base + threadIdx.x;
    int x[4];
    bool ok[4];
    load4(v, n, c < chunks ? c : chunks - 1, vec, x, ok);
    unsigned pend = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) pend |= (c < chunks && ok[i] ? 1u : 0u) << i;
    for (int r = 0; r < 4; ++r) {
      const uint64_t act = ballot(pend != 0);
      if (act == 0) break;
      const int first = __builtin_ctzll(act);
      const int mine = (pend & 1u)   ? x[0]
                       : (pend & 2u) ? x[1]
                       : (pend & 4u) ? x[2]
                       : x[3];
      const int V = __shfl(mine, first);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        pend &= x[i] == V ? ~(1u << i) : ~0u;
      if (lane == first) mark_value(bm, lo, V);
    }
    if (ballot(pend != 0) == 0) continue;
    const int prev = __shfl_up(x[3], 1);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int before = i == 0 ? prev : x[i - 1];
      const bool lead = (i == 0 && lane == 0) || before != x[i];
      if (((pend >> i) & 1u) && lead) mark_value(bm, lo, x[i]);
    }
  }
}

// Exclusive popcount prefix of a bitmap of st->words words: per chunk of
// kScanChunk words (256 threads x 4 consecutive words), each word's prefix
// inside its chunk into prefix[], the chunk total into csum[] (chunk_scan
// then scans csum[] in place and stores the total).  With `mask`,
// the bitmap is first replaced by bm & ~mask (the roots: local roots not
// linked).
__global__ __launch_bounds__(kScanThreads) void bitmap_scan(
    CompressState* st, unsigned* bm, const unsigned* mask, unsigned* prefix,
    unsigned* csum) {
  if (st->err) return;
  const int64_t words = st->words;
  const int64_t nchunks = (words + kScanChunk - 1) / kScanChunk;
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  __shared__ unsigned wsum[kScanThreads / kWave];
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int64_t w0 = c * kScanChunk + 4 * (int64_t)threadIdx.x;
    unsigned v[4], cnt[4], tsum = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t w = w0 + i;
      v[i] = w < words ? bm[w] : 0u;
      if (mask && w < words) {
        v[i] &= ~mask[w];
        bm[w] = v[i];
      }
      cnt[i] = __popc(v[i]);
      tsum += cnt[i];
    }
    // inclusive wave scan of the thread sums, then the waves' totals
    unsigned inc = tsum;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const unsigned y = __shfl_up(inc, o);
      inc += lane >= o ? y : 0u;
    }
    if (lane == kWave - 1) wsum[wv] = inc;
    __syncthreads();
    unsigned before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < kScanThreads / kWave; ++k) {
      before += k < wv ? wsum[k] : 0u;
      total += wsum[k];
    }
    unsigned run = before + inc - tsum;  // exclusive
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (w0 + i < words) prefix[w0 + i] = run;
      run += cnt[i];
    }
    if (threadIdx.x == 0) csum[c] = total;
    __syncthreads();
  }
}

// Exclusive scan of the chunk totals of bitmap_scan, in place (one block);
// the total into st.
__global__ __launch_bounds__(kScanThreads) void chunk_scan(CompressState* st,
                                                           unsigned* csum) {
  if (st->err) return;
  const int64_t nchunks = (st->words + kScanChunk - 1) / kScanChunk;
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
  if (threadIdx.x == kScanThreads - 1) st->total = (int64_t)s64[threadIdx.x];
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

// Zeroes the linked-root (and local-root) bitmaps; with st, sets up the
// state for the root scan (values are pixel indices in [0, n)).
__global__ __launch_bounds__(256) void cc_clear(unsigned* bm, int64_t words,
                                                CompressState* st, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (int64_t)gridDim.x * blockDim.x)
    bm[i] = 0u;
  if (st && blockIdx.x == 0 && threadIdx.x == 0) {
    st->lo = 0;
    st->hi = (int)(n - 1);
    st->arrive = 0u;
    st->err = 0;
    st->words = (n + 31) / 32;
    st->total = 0;
  }
}

// Temp block of connected_components: [state][linked-root bitmap][local-
// root bitmap, word prefix, chunk sums: with compression][edge arrays].
// Bitmaps carry 2 words after the last pixel's (or_bits64 may touch them).
struct CcLayout {
  CompressState* st;
  unsigned* linked;
  unsigned* local_root;
  unsigned* prefix;
  unsigned* csum;
  int32_t* edge[4];
  int64_t words;  // bitmap words incl. padding
  size_t bytes;
};

// Tile side of the single-image CC passes.  Measured on the 4096^2 k=32
// segment labels / the 8192^2 spiral: 64 -> 256 / 5,330 us for CC +
// compress, 32 -> 280 / 9,840 us (four times the seams; longer root
// chains for the jump pass).
constexpr int kCcSide = 64;
static int cc_side() { return kCcSide; }

static CcLayout cc_layout(void* temp, int W, int H, bool compress) {
  CcLayout L{};
  const int S = cc_side();
  const int64_t n = (int64_t)W * H;
  const int64_t ntiles = (int64_t)((W + S - 1) / S) * ((H + S - 1) / S);
  L.words = (n + 31) / 32 + 2;
  char* t = static_cast<char*>(temp);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    char* q = t ? t + o : nullptr;
    o += align_up(bytes, 256);
    return q;
  };
  L.st = reinterpret_cast<CompressState*>(take(sizeof(CompressState)));
  // linked and local-root bitmaps adjacent: one clear covers both
  L.linked = reinterpret_cast<unsigned*>(take((size_t)L.words * 4));
  if (compress) {
    L.local_root = L.linked ? L.linked + align_up((size_t)L.words * 4, 256) / 4
                            : nullptr;
    take((size_t)L.words * 4);
    L.prefix = reinterpret_cast<unsigned*>(take((size_t)L.words * 4));
    L.csum = reinterpret_cast<unsigned*>(take(compress_csum_bytes(L.words)));
  }
  for (int e = 0; e < 4; ++e)
    L.edge[e] = reinterpret_cast<int32_t*>(take((size_t)ntiles * S * 4));
  L.bytes = o;
  return L;
}

static int cc_check_args(const int32_t* labels, int W, int H, int32_t* out,
                         int conn) {
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
  if ((H + cc_side() - 1) / cc_side() > 65535)
    return fail(SPLIT_HIP_ERR_UNSUPPORTED, "height %d too large", H);
  return SPLIT_HIP_OK;
}

// The pipeline of the header comment; with compress, out receives dense
// ranks and L.st the count.
static int cc_run(const int32_t* labels, int W, int H, int32_t* out, int conn,
                  bool compress, const CcLayout& L, hipStream_t stream) {
  const int64_t n = (int64_t)W * H;
  const int S = cc_side();
  const int ntx = (W + S - 1) / S, nty = (H + S - 1) / S;
  const dim3 tiles(ntx, nty);
  CcArgs a{};
  a.labels = labels;
  a.out = out;
  a.W = W;
  a.H = H;
  a.ntx = ntx;
  a.linked = L.linked;
  a.local_root = compress ? L.local_root : nullptr;
  a.llab = L.edge[0];
  a.rlab = L.edge[1];
  a.lroot = L.edge[2];
  a.rroot = L.edge[3];
  const int64_t clear_words =
      compress ? (int64_t)(L.local_root - L.linked) + L.words : L.words;
  hipLaunchKernelGGL(cc_clear, dim3(grid_for(clear_words, 256, 4)), dim3(256),
                     0, stream, L.linked, clear_words,
                     compress ? L.st : nullptr, n);
  constexpr int kS = kCcSide;
  if (conn == 8) {
    hipLaunchKernelGGL((cc_tile<8, kS>), tiles, dim3(CcTile<kS>::kThreads), 0,
                       stream, a);
    hipLaunchKernelGGL((cc_merge<8, kS>), tiles, dim3(3 * kS), 0, stream, a);
  } else {
    hipLaunchKernelGGL((cc_tile<4, kS>), tiles, dim3(CcTile<kS>::kThreads), 0,
                       stream, a);
    hipLaunchKernelGGL((cc_merge<4, kS>), tiles, dim3(3 * kS), 0, stream, a);
  }
  SPLIT_HIP_CHECK_LAUNCH("cc_tile/cc_merge");
  const int64_t bits_words = (n + 31) / 32;
  // 32 blocks per CU: about one bitmap word per thread on the 8192^2 image
  // (8 per CU: 797 us there)
  hipLaunchKernelGGL(cc_jump, dim3(grid_for(bits_words, 256, 32)), dim3(256),
                     0, stream, out, L.linked, bits_words);
  if (compress) {
    hipLaunchKernelGGL(bitmap_scan,
                       dim3(grid_for((bits_words + kScanChunk - 1) / kScanChunk,
                                     1, 8)),
                       dim3(kScanThreads), 0, stream, L.st, L.local_root,
                       L.linked, L.prefix, L.csum);
    hipLaunchKernelGGL(chunk_scan, dim3(1), dim3(kScanThreads), 0, stream,
                       L.st, L.csum);
  }
  hipLaunchKernelGGL(cc_final<kS>, tiles, dim3(CcTile<kS>::kThreads), 0,
                     stream, a, compress ? L.local_root : nullptr, L.prefix,
                     L.csum);
  SPLIT_HIP_CHECK_LAUNCH("cc_jump/cc_final");
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
  hipLaunchKernelGGL(bitmap_scan, dim3(grid_for(L.cap, kScanChunk, 8)),
                     dim3(kScanThreads), 0, stream, L.st, L.bm,
                     (const unsigned*)nullptr, L.prefix, L.csum);
  hipLaunchKernelGGL(chunk_scan, dim3(1), dim3(kScanThreads), 0, stream, L.st,
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
  hipLaunchKernelGGL(compress_minmax, dim3(grid_for(chunks, 1024, 1)),
                     dim3(1024), 0, stream, labels, n, vec, L.st);
  hipLaunchKernelGGL(compress_clear, dim3(grid_for(L.cap, 256, 8)), dim3(256),
                     0, stream, L.st, L.bm, L.cap);
  hipLaunchKernelGGL(compress_mark, dim3(grid_for(chunks, 256, 8)), dim3(256),
                     0, stream, labels, n, vec, L.st, L.bm);
  SPLIT_HIP_CHECK_LAUNCH("compress kernels");
  SPLIT_HIP_TRY(compress_rank(L, labels, n, stream));
  *st_out = L.st;
  return SPLIT_HIP_OK;
}

// Connected components of `labels` into `segments`, with compress their
// dense ranks (the root bitmap comes from the CC passes themselves: no
// min/max, clear or mark pass).
int segments_async(const int32_t* labels, int W, int H, int32_t* segments,
                   int conn, bool compress, void* temp, size_t temp_bytes,
                   hipStream_t stream, CompressState** st_out) {
  SPLIT_HIP_TRY(cc_check_args(labels, W, H, segments, conn));
  if (!temp) return fail(SPLIT_HIP_ERR_ARGUMENT, "temp is null");
  const CcLayout L = cc_layout(temp, W, H, compress);
  if (temp_bytes < L.bytes)
    return fail(SPLIT_HIP_ERR_TEMP,
                "temp holds %zu bytes, connected components need %zu",
                temp_bytes, L.bytes);
  SPLIT_HIP_TRY(cc_run(labels, W, H, segments, conn, compress, L, stream));
  *st_out = compress ? L.st : nullptr;
  return SPLIT_HIP_OK;
}

size_t segments_workspace(int W, int H, bool compress) {
  return cc_layout(nullptr, W, H, compress).bytes;
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

size_t split_hip_connected_components_workspace(int32_t width,
                                                int32_t height) {
  if (width < 1 || height < 1) return 256;
  return cc_layout(nullptr, width, height, false).bytes;
}

int split_hip_connected_components(const int32_t* labels, int32_t width,
                                   int32_t height, void* temp,
                                   size_t temp_bytes, int32_t* out, int32_t,
                                   int32_t connectivity,
                                   split_hip_stream_t stream) {
  SPLIT_HIP_TRY(cc_check_args(labels, width, height, out, connectivity));
  const CcLayout L = cc_layout(temp, width, height, false);
  if (!temp || temp_bytes < L.bytes)
    return fail(SPLIT_HIP_ERR_TEMP,
                "temp holds %zu bytes, connected_components needs %zu",
                temp ? temp_bytes : (size_t)0, L.bytes);
  return cc_run(labels, width, height, out, connectivity, false, L, stream);
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
