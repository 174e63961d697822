// k-means on gfx950: fused nearest-centroid assignment + per-cluster
// sum/count reduction + centroid update, one launch per Lloyd iteration.
//
// Replaces the reference's Thrust pipeline (split/src/split/device/kmeans/
// label.cu:35-106, centroids.cu:22-101, cluster.cu:27-93): there the N x K
// distance matrix is written to and re-read from memory, pixels are radix
// sorted by label every iteration and the host reads the shift back each
// iteration.  Here one pass over the three SoA float planes does all of it:
//
//  * Each wave owns a contiguous run of 256-pixel segments (64 lanes x 4
//    pixels, float4 loads per plane when aligned).  When the image width is
//    known a segment is a 16x16 tile (lane l: row l/4, columns 4*(l%4)..+3),
//    whose colours are far more compact than a 1x256 run of a row.
//  * Per segment, the wave reduces the colour box of its pixels and drops
//    every centroid that provably cannot be the nearest for any pixel in it:
//    (a) its squared distance to the box exceeds the smallest farthest-corner
//    distance, or (b) the whole box lies on the far side of its bisector with
//    the best-guess centroid g (the one with that smallest corner distance).
//    Both tests keep a 1e-5 relative margin, far above the few-ulp error of
//    the float bounds and distances, so a dropped centroid's float distance
//    is strictly larger than g's for every pixel in the box: the labels equal
//    the dense float argmin over all k (ties to the lowest index).
//  * Centroids live in VGPRs (lane j holds centroid j + 64*t); survivors are
//    broadcast with v_readlane in increasing index order.
//  * Sums/counts: while consecutive segments are label-uniform each lane adds
//    its pixels into registers; on a label change the run is reduced across
//    the wave (DPP) into a per-wave LDS slice.  Mixed segments are reduced one
//    distinct label at a time.  No atomics touch the sums, so the result is
//    bit-identical run to run.
//  * Each block folds its wave slices into a per-block partial column; the
//    last block to finish (arrival counter) reduces the columns in a fixed
//    order in double, writes the new centroids, the shift and the stop flag.
#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "lloyd_device.hpp"

namespace splithip {

constexpr int kMaxGrid = 1024;        // upper bound on step-kernel blocks
constexpr size_t kStateBytes = 256;   // LloydState slot at the head of temp

struct StepArgs : Geometry {
  int k;
  int use_prev;              // Lloyd label pass: assign against previous
  const float* cent;         // centroids to assign against (standalone)
  float* cent_out;           // new centroids (standalone calculate)
  float* cbuf;               // Lloyd: two 3k centroid buffers
  LloydState* state;         // Lloyd state, or null when standalone
  Mailbox* mbox;             // host-mapped mailbox (Lloyd), may be null
  unsigned* arrive;          // arrival counter (standalone calculate)
  int32_t* labels;           // labels written by assignment, may be null
  const int32_t* labels_in;  // labels read instead of assigning
  float* partial;            // [4k][gridDim.x] per-block partial columns
};

// --------------------------------------------------------- accumulation ---

struct RunAcc {
  float x, y, z;
  unsigned n;
  int label;  // wave-uniform label of the open run, -1 if none
};

__device__ __forceinline__ void slice_add(float* slice, int k, int label,
                                          float sx, float sy, float sz,
                                          unsigned cnt, int lane) {
  if (lane == 0) {
    slice[label] += sx;
    slice[k + label] += sy;
    slice[2 * k + label] += sz;
    reinterpret_cast<unsigned*>(slice)[3 * k + label] += cnt;
  }
}

__device__ __forceinline__ void flush_run(RunAcc& r, float* slice, int k,
                                          int lane) {
  if (r.label >= 0) {
    const float sx = wave_allreduce(r.x, OpAdd());
    const float sy = wave_allreduce(r.y, OpAdd());
    const float sz = wave_allreduce(r.z, OpAdd());
    const unsigned sn = wave_sum_u32(r.n);
    slice_add(slice, k, r.label, sx, sy, sz, sn, lane);
  }
  r.x = r.y = r.z = 0.f;
  r.n = 0;
  r.label = -1;
}

__device__ __forceinline__ void accumulate_segment(
    RunAcc& r, float* slice, int k, int lane, const float (&x)[4],
    const float (&y)[4], const float (&z)[4], const int (&lab)[4]) {
  const int l0 = readfirstlane_i(lab[0]);
  const bool same = lab[0] == l0 && lab[1] == l0 && lab[2] == l0 &&
                    lab[3] == l0;
  if (l0 >= 0 && ballot(!same) == 0) {
    // Whole segment carries one label: extend (or open) the register run.
    if (l0 != r.label) {
      flush_run(r, slice, k, lane);
      r.label = l0;
    }
    r.x += (x[0] + x[1]) + (x[2] + x[3]);
    r.y += (y[0] + y[1]) + (y[2] + y[3]);
    r.z += (z[0] + z[1]) + (z[2] + z[3]);
    r.n += 4;
    return;
  }
  // Mixed segment: reduce one distinct label at a time.
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
    float sx = 0.f, sy = 0.f, sz = 0.f;
    unsigned cnt = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool m = ((rem >> i) & 1u) && lab[i] == L;
      sx += m ? x[i] : 0.f;
      sy += m ? y[i] : 0.f;
      sz += m ? z[i] : 0.f;
      cnt += m ? 1u : 0u;
      rem &= m ? ~(1u << i) : ~0u;
    }
    sx = wave_allreduce(sx, OpAdd());
    sy = wave_allreduce(sy, OpAdd());
    sz = wave_allreduce(sz, OpAdd());
    cnt = wave_sum_u32(cnt);
    slice_add(slice, k, L, sx, sy, sz, cnt, lane);
  }
}

// ------------------------------------------------------- final reduction ---

// Run by the last block of a launch: reduce the per-block partial columns
// in a fixed order (double sums, exact integer counts), write the new
// centroids and, in Lloyd mode, the shift and the stop decision.
__device__ void finalize_centroids(const StepArgs& a, const float* cent,
                                   float* cent_out, unsigned* arrive,
                                   float* lds) {
  const int k = a.k;
  const int V = 4 * k;
  const int G = gridDim.x;
  const int nt = blockDim.x;
  const int tid = threadIdx.x;
  const int T = nt / V > 1 ? nt / V : 1;  // threads per value
  // LDS: tot[V] | (T > 1: red[nt]) ... later reused as shv[k] | part[64]
  double* tot = reinterpret_cast<double*>(lds);
  double* red = tot + V;
  const unsigned* upart = reinterpret_cast<const unsigned*>(a.partial);

  for (int v = tid / T; v < V; v += nt / T) {
    const int sub = tid % T;
    double acc = 0.0;
    const size_t row = (size_t)v * G;
    if (v < 3 * k) {
      for (int g = sub; g < G; g += T) acc += (double)a.partial[row + g];
    } else {
      uint64_t c = 0;
      for (int g = sub; g < G; g += T) c += upart[row + g];
      acc = (double)c;
    }
    if (T > 1)
      red[v * T + sub] = acc;
    else
      tot[v] = acc;
  }
  __syncthreads();
  if (T > 1) {
    for (int v = tid; v < V; v += nt) {
      double s = 0.0;
      for (int t = 0; t < T; ++t) s += red[v * T + t];
      tot[v] = s;
    }
    __syncthreads();
  }
  double* shv = tot + V;      // k entries (red is consumed)
  double* part = shv + k;     // kWave entries
  for (int j = tid; j < k; j += nt) {
    const double cnt = tot[3 * k + j];
    double sh = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float nc = cnt > 0.0 ? (float)(tot[c * k + j] / cnt) : 0.f;
      if (cent) {
        const double d = (double)nc - (double)cent[c * k + j];
        sh += d * d;
      }
      cent_out[c * k + j] = nc;
    }
    shv[j] = sh;
  }
  __syncthreads();
  if (tid < kWave) {
    double s = 0.0;
    for (int j = tid; j < k; j += kWave) s += shv[j];
    part[tid] = s;
  }
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int j = 0; j < kWave; ++j) s += part[j];
    const float shift = (float)sqrt(s);
    if (a.state) {
      LloydState* st = a.state;
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
    *arrive = 0u;
  }
}

// -------------------------------------------------------------- kernel ---

template <int CPL, int MODE, bool ASSIGN, bool ACCUM>
__global__ __launch_bounds__(1024) void lloyd_step(StepArgs a) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  const int k = a.k;

  const float* cent = a.cent;
  float* cent_out = a.cent_out;
  unsigned* arrive = a.arrive;
  if (a.state) {
    const LloydState* st = a.state;
    if (a.use_prev ? st->iter == 0 : st->done) return;
    const int cur = st->cur;
    cent = a.cbuf + (size_t)(a.use_prev ? 1 - cur : cur) * 3 * k;
    cent_out = a.cbuf + (size_t)(1 - cur) * 3 * k;
    arrive = &a.state->arrive;
  }

  float cx[CPL], cy[CPL], cz[CPL], cn[CPL];
  if constexpr (ASSIGN) {
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
      const int j = lane + 64 * t;
      const bool live = j < k;
      cx[t] = live ? cent[j] : 0.f;
      cy[t] = live ? cent[k + j] : 0.f;
      cz[t] = live ? cent[2 * k + j] : 0.f;
      cn[t] = fmaf(cz[t], cz[t], fmaf(cy[t], cy[t], cx[t] * cx[t]));
    }
  }

  float* slice = lds + (size_t)wave * 4 * k;
  if constexpr (ACCUM) {
    for (int i = threadIdx.x; i < nwaves * 4 * k; i += blockDim.x) lds[i] = 0.f;
    __syncthreads();
  }

  const int64_t gw = (int64_t)blockIdx.x * nwaves + wave;
  const int64_t W = (int64_t)gridDim.x * nwaves;
  const int64_t s0 = a.nseg * gw / W;
  const int64_t s1 = a.nseg * (gw + 1) / W;

  RunAcc run{0.f, 0.f, 0.f, 0u, -1};
  for (int64_t s = s0; s < s1; ++s) {
    float x[4], y[4], z[4];
    bool v[4], lvec;
    int64_t idx[4];
    load_segment<MODE>(a, s, lane, x, y, z, v, idx, lvec);
    int lab[4];
    if constexpr (ASSIGN) {
      assign_segment<CPL, false>(cx, cy, cz, cn, k, lane, x, y, z, v, lab, nullptr);
      if (a.labels) {
        if (lvec) {
          *reinterpret_cast<I4*>(a.labels + idx[0]) =
              I4{lab[0], lab[1], lab[2], lab[3]};
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (v[i]) a.labels[idx[i]] = lab[i];
        }
      }
    } else {
      if (lvec) {
        const I4 L = *reinterpret_cast<const I4*>(a.labels_in + idx[0]);
        lab[0] = L.x; lab[1] = L.y; lab[2] = L.z; lab[3] = L.w;
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) lab[i] = v[i] ? a.labels_in[idx[i]] : -1;
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
        lab[i] = (lab[i] >= 0 && lab[i] < k) ? lab[i] : -1;
    }
    if constexpr (ACCUM) accumulate_segment(run, slice, k, lane, x, y, z, lab);
  }

  if constexpr (ACCUM) {
    flush_run(run, slice, k, lane);
    __syncthreads();
    const int V = 4 * k;
    const int G = gridDim.x;
    const unsigned* ulds = reinterpret_cast<const unsigned*>(lds);
    for (int v = threadIdx.x; v < V; v += blockDim.x) {
      if (v < 3 * k) {
        float s = 0.f;
        for (int w = 0; w < nwaves; ++w) s += lds[(size_t)w * V + v];
        a.partial[(size_t)v * G + blockIdx.x] = s;
      } else {
        unsigned s = 0;
        for (int w = 0; w < nwaves; ++w) s += ulds[(size_t)w * V + v];
        reinterpret_cast<unsigned*>(a.partial)[(size_t)v * G + blockIdx.x] = s;
      }
    }
    __threadfence();
    __shared__ unsigned ticket;
    __syncthreads();
    if (threadIdx.x == 0) ticket = atomicAdd(arrive, 1u);
    __syncthreads();
    if (ticket != (unsigned)G - 1) return;
    __threadfence();
    finalize_centroids(a, a.state ? cent : nullptr, cent_out, arrive, lds);
  }
}

// Lloyd set-up: state fields and the first centroid buffer.
__global__ void lloyd_init(LloydState* st, float* cbuf, const float* cent,
                           int k, int max_iter, float threshold) {
  for (int i = threadIdx.x; i < 3 * k; i += blockDim.x) cbuf[i] = cent[i];
  if (threadIdx.x == 0) {
    st->arrive = 0u;
    st->cur = 0;
    st->iter = 0;
    st->done = max_iter <= 0;
    st->max_iter = max_iter;
    st->threshold = threshold;
    st->shift = 0.f;
    st->next = 0u;
  }
}

// Lloyd wrap-up: copy the current centroid buffer out.
__global__ void lloyd_finish(const LloydState* st, const float* cbuf,
                             float* cent, int k) {
  const float* src = cbuf + (size_t)st->cur * 3 * k;
  for (int i = threadIdx.x; i < 3 * k; i += blockDim.x) cent[i] = src[i];
}

__global__ void zero_u32(unsigned* p, int count) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0u;
}

// Seeded sampler: k distinct pixel indices from SplitMix64, colours copied.
__device__ __forceinline__ uint64_t splitmix64_next(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void sample_init(const float* pts, int64_t n,
                                                   float* cent, int k,
                                                   uint64_t seed) {
  __shared__ int64_t chosen[SPLIT_HIP_MAX_K];
  __shared__ int64_t cand;
  __shared__ int dup;
  uint64_t s = seed;  // only thread 0's copy advances
  for (int j = 0; j < k;) {
    if (threadIdx.x == 0) {
      cand = (int64_t)__umul64hi(splitmix64_next(s), (uint64_t)n);
      dup = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < j; i += blockDim.x)
      if (chosen[i] == cand) dup = 1;
    __syncthreads();
    const bool accept = !dup;
    if (accept && threadIdx.x == 0) chosen[j] = cand;
    __syncthreads();
    j += accept ? 1 : 0;
  }
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    const int64_t i = chosen[j];
    cent[j] = pts[i];
    cent[k + j] = pts[n + i];
    cent[2 * k + j] = pts[2 * n + i];
  }
}

// ------------------------------------------------------------ launchers ---

static int cpl_for(int k) {
  return k <= 64 ? 1 : k <= 128 ? 2 : k <= 256 ? 4 : k <= 512 ? 8 : 16;
}

// Block size by k keeps the per-wave LDS slices (16*k bytes each) at 32 KiB.
static int threads_for(int k) {
  return k <= 128 ? 1024 : k <= 256 ? 512 : k <= 512 ? 256 : 128;
}

struct StepCfg {
  int grid;
  int threads;
  size_t lds;
  int mode;
};

// partial_cap: bytes available for the per-block partial columns (16*k
// bytes per block); the grid shrinks to fit, so any temp that holds one
// column works.
static StepCfg step_cfg(const StepArgs& a, bool accum, size_t partial_cap = 0) {
  StepCfg c;
  c.threads = threads_for(a.k);
  const int nwaves = c.threads / kWave;
  const int64_t want = (a.nseg + nwaves - 1) / nwaves;
  int64_t cap =
      accum ? std::min(device_cu_count() * (1024 / c.threads), kMaxGrid)
            : 4 * device_cu_count() * (1024 / c.threads);
  if (accum)
    cap = std::min<int64_t>(cap, (int64_t)(partial_cap / (16 * (size_t)a.k)));
  c.grid = (int)std::max<int64_t>(1, std::min<int64_t>(want, cap));
  if (accum) {
    const size_t V = 4 * (size_t)a.k;
    const size_t slices = (size_t)nwaves * V * sizeof(float);
    const size_t T = (size_t)c.threads / V > 1 ? (size_t)c.threads / V : 1;
    const size_t fin =
        (V + std::max(T > 1 ? (size_t)c.threads : 0, (size_t)a.k + kWave)) *
        sizeof(double);
    c.lds = std::max(slices, fin);
  } else {
    c.lds = 0;
  }
  // a lane's four consecutive pixels move as one dword-aligned 16-byte
  // vector, whatever n, the width or the plane offsets
  const bool vec = aligned4(a.pts) && (!a.labels || aligned4(a.labels)) &&
                   (!a.labels_in || aligned4(a.labels_in));
  if (a.width > 0)
    c.mode = vec ? kTileVec : kTile;
  else
    c.mode = vec ? kRunVec : kRun;
  return c;
}

template <bool ASSIGN, bool ACCUM, int MODE>
static void launch_mode(const StepArgs& a, const StepCfg& c,
                        hipStream_t stream) {
  const dim3 g(c.grid), b(c.threads);
  switch (ASSIGN ? cpl_for(a.k) : 1) {
#define SPLIT_LAUNCH(CPL)                                                    \
  case CPL:                                                                  \
    hipLaunchKernelGGL((lloyd_step<CPL, MODE, ASSIGN, ACCUM>), g, b, c.lds,  \
                       stream, a);                                           \
    break;
    SPLIT_LAUNCH(1)
    SPLIT_LAUNCH(2)
    SPLIT_LAUNCH(4)
    SPLIT_LAUNCH(8)
    SPLIT_LAUNCH(16)
#undef SPLIT_LAUNCH
  }
}

template <bool ASSIGN, bool ACCUM>
static int launch_step(const StepArgs& a, const StepCfg& c,
                       hipStream_t stream) {
  switch (c.mode) {
    case kRun: launch_mode<ASSIGN, ACCUM, kRun>(a, c, stream); break;
    case kRunVec: launch_mode<ASSIGN, ACCUM, kRunVec>(a, c, stream); break;
    case kTile:
      if constexpr (ASSIGN) launch_mode<ASSIGN, ACCUM, kTile>(a, c, stream);
      break;
    default:
      if constexpr (ASSIGN) launch_mode<ASSIGN, ACCUM, kTileVec>(a, c, stream);
      break;
  }
  SPLIT_HIP_CHECK_LAUNCH("lloyd_step");
  return SPLIT_HIP_OK;
}

// Geometry of a StepArgs: 1-D runs (width == 0) or 16x16 tiles.
static void set_geometry(StepArgs& a, int64_t n, int width, int height) {
  a.n = n;
  if (width > 0) {
    a.width = width;
    a.height = height;
    a.full_rows = height;
    a.tw_log = 4;
    a.tiles_x = (int)tiles_across(width, a.tw_log);
    a.nseg = (int64_t)a.tiles_x * tiles_down(height, a.tw_log);
  } else {
    a.width = a.height = a.tiles_x = a.full_rows = 0;
    a.nseg = (n + kSegPx - 1) / kSegPx;
  }
}

static int check_image(int width, int height) {
  if (width < 1 || height < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "image is %d x %d", width, height);
  return SPLIT_HIP_OK;
}

// Partial columns for the full grid, and for a single block (the least any
// accumulating entry point accepts).
static size_t partial_bytes(int k) {
  return align_up(4 * (size_t)k * kMaxGrid * sizeof(float), 256);
}
static size_t partial_bytes_min(int k) { return 16 * (size_t)k; }

int driver_res(DriverRes** out) {
  static thread_local DriverRes r;
  int dev = 0;
  SPLIT_HIP_CHECK(hipGetDevice(&dev));
  if (r.host == nullptr) {
    SPLIT_HIP_CHECK(hipHostMalloc((void**)&r.host, sizeof(Mailbox),
                                  hipHostMallocMapped));
  }
  if (r.device != dev) {
    SPLIT_HIP_CHECK(hipHostGetDevicePointer((void**)&r.dev, r.host, 0));
    // The events only pace the host (it waits for chunk c - 1 while chunk c
    // runs, then reads the mailbox): without the system-scope fence an event
    // costs no cache writeback and invalidation between two Lloyd launches
    // (~6 us each).  A mailbox read that lags only enqueues launches that
    // return at once; results are read after a full stream synchronisation.
    for (int i = 0; i < 2; ++i)
      SPLIT_HIP_CHECK(hipEventCreateWithFlags(
          &r.ev[i], hipEventDisableTiming | hipEventDisableSystemFence));
    r.device = dev;
  }
  *out = &r;
  return SPLIT_HIP_OK;
}

static int check_points(const float* pts, int64_t n, int k) {
  if (!pts) return fail(SPLIT_HIP_ERR_ARGUMENT, "points is null");
  if (n < 1) return fail(SPLIT_HIP_ERR_ARGUMENT, "n = %lld < 1", (long long)n);
  if (k < 1) return fail(SPLIT_HIP_ERR_ARGUMENT, "k = %d < 1", k);
  if (k > SPLIT_HIP_MAX_K)
    return fail(SPLIT_HIP_ERR_UNSUPPORTED, "k = %d exceeds %d", k,
                SPLIT_HIP_MAX_K);
  return SPLIT_HIP_OK;
}

// Width of the rows a point set without image geometry is walked in by the
// bounded tiled path: a 16x16 tile is then 256 consecutive points.
constexpr int kPointsWidth = 16;

// Internal entry used by the pipeline: Lloyd loop without the final host
// synchronisation (the mailbox holds the result once the stream drains).
// width == 0: a point set without image geometry.  The bounded tiled path
// (k <= 256) takes it as rows of kPointsWidth points; the 1-D runs path
// (float sums) serves larger k and temp blocks too small for the tiled
// path's per-pixel state.
int cluster_async(const float* pts, int64_t n, int width, int height,
                  float* centroids, int k, int32_t* labels, void* temp,
                  size_t temp_bytes, int max_iter, float threshold,
                  hipStream_t stream, Mailbox** mbox_out) {
  SPLIT_HIP_TRY(check_points(pts, n, k));
  if (!centroids) return fail(SPLIT_HIP_ERR_ARGUMENT, "centroids is null");
  if (!temp) return fail(SPLIT_HIP_ERR_ARGUMENT, "temp is null");
  const int tw = width > 0 ? width : kPointsWidth;
  if (tiles_supported(n, tw, k) && temp_bytes >= tiles_workspace(n, k))
    return cluster_tiles_async(pts, n, tw, centroids, k, labels, temp,
                               temp_bytes, max_iter, threshold, stream,
                               mbox_out);
  const size_t head = kStateBytes + align_up(6 * (size_t)k * sizeof(float), 256);
  if (temp_bytes < head + partial_bytes_min(k))
    return fail(SPLIT_HIP_ERR_TEMP, "temp holds %zu bytes, cluster needs >= %zu",
                temp_bytes, head + partial_bytes_min(k));
  DriverRes* res = nullptr;
  SPLIT_HIP_TRY(driver_res(&res));

  char* t = static_cast<char*>(temp);
  LloydState* st = reinterpret_cast<LloydState*>(t);
  float* cbuf = reinterpret_cast<float*>(t + kStateBytes);
  float* partial = reinterpret_cast<float*>(t + head);

  res->host->iter = 0;
  res->host->done = max_iter <= 0;
  res->host->shift = 0.f;
  hipLaunchKernelGGL(lloyd_init, dim3(1), dim3(256), 0, stream, st, cbuf,
                     centroids, k, max_iter, threshold);
  SPLIT_HIP_CHECK_LAUNCH("lloyd_init");

  StepArgs a{};
  a.pts = pts;
  set_geometry(a, n, width, height);
  a.k = k;
  a.cbuf = cbuf;
  a.state = st;
  a.mbox = res->dev;
  a.partial = partial;
  const StepCfg cfg = step_cfg(a, true, temp_bytes - head);

  // Enqueue iterations in chunks; after queuing chunk c, wait for chunk c-1
  // and stop once its last block has reported convergence.  The GPU always
  // has the next chunk queued, and kernels past convergence exit at once.
  // A run that cannot stop early (threshold < 0) is enqueued at once, small
  // point sets in chunks of 16 (as the tiled path).
  const int kChunk = threshold < 0.f ? std::max(max_iter, 1)
                     : n < ((int64_t)1 << 22) ? 16 : 4;
  int launched = 0, chunk = 0;
  while (launched < max_iter) {
    const int cnt = std::min(kChunk, max_iter - launched);
    for (int i = 0; i < cnt; ++i)
      SPLIT_HIP_TRY((launch_step<true, true>(a, cfg, stream)));
    launched += cnt;
    SPLIT_HIP_CHECK(hipEventRecord(res->ev[chunk & 1], stream));
    if (chunk > 0) {
      SPLIT_HIP_CHECK(hipEventSynchronize(res->ev[(chunk - 1) & 1]));
      if (reinterpret_cast<volatile Mailbox*>(res->host)->done) break;
    }
    ++chunk;
  }

  if (labels) {
    StepArgs b = a;
    b.use_prev = 1;
    b.labels = labels;
    b.partial = nullptr;
    b.mbox = nullptr;
    const StepCfg lc = step_cfg(b, false);
    SPLIT_HIP_TRY((launch_step<true, false>(b, lc, stream)));
  }
  hipLaunchKernelGGL(lloyd_finish, dim3(1), dim3(256), 0, stream, st, cbuf,
                     centroids, k);
  SPLIT_HIP_CHECK_LAUNCH("lloyd_finish");
  *mbox_out = res->host;
  return SPLIT_HIP_OK;
}

int calculate_centroids_impl(const int32_t* labels, const float* points,
                             int64_t n, float* centroids, int32_t k,
                             void* temp, size_t temp_bytes,
                             hipStream_t stream) {
  if (k > SPLIT_HIP_MAX_K) {  // segment means: k beyond the register kernels
    if (!points || !labels || !centroids || !temp || n < 1)
      return fail(SPLIT_HIP_ERR_ARGUMENT,
                  "points, labels, centroids or temp is null, or n < 1");
    return segment_means(labels, points, n, centroids, k, temp, temp_bytes,
                         stream);
  }
  SPLIT_HIP_TRY(check_points(points, n, k));
  if (!labels || !centroids || !temp)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "labels, centroids or temp is null");
  if (temp_bytes < kStateBytes + partial_bytes_min(k))
    return fail(SPLIT_HIP_ERR_TEMP,
                "temp holds %zu bytes, calculate_centroids needs >= %zu",
                temp_bytes, kStateBytes + partial_bytes_min(k));
  char* t = static_cast<char*>(temp);
  StepArgs a{};
  a.pts = points;
  set_geometry(a, n, 0, 0);
  a.k = k;
  a.labels_in = labels;
  a.cent_out = centroids;
  a.arrive = reinterpret_cast<unsigned*>(t);
  a.partial = reinterpret_cast<float*>(t + kStateBytes);
  hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, stream, a.arrive, 1);
  SPLIT_HIP_CHECK_LAUNCH("zero_u32");
  const StepCfg c = step_cfg(a, true, temp_bytes - kStateBytes);
  return launch_step<false, true>(a, c, stream);
}

// ------------------------------------------- calculate_centroids_noalloc ---
//
// Scratch of the allocation-free calculate_centroids: a file-scope device
// array that each GPU holds for the life of the process.  Its size is the
// workspace calculate_centroids asks for at k = SPLIT_HIP_MAX_K (arrival
// counter + partial columns for the full grid, 16 MiB + 256 B), so every
// k <= SPLIT_HIP_MAX_K runs the same grid as the allocating entry point;
// larger k (segment means, ~28 B per label) fits up to kNoAllocMaxSegments.
// Every call on a GPU uses that GPU's one array: the kernels clear the
// counter and sums they need at the start of each call, and calls that
// overlap on one GPU (two streams, two host threads) race.
constexpr size_t kNoAllocScratchBytes =
    kStateBytes + (size_t)4 * SPLIT_HIP_MAX_K * kMaxGrid * sizeof(float);
constexpr int64_t kNoAllocMaxSegments =
    (int64_t)(kNoAllocScratchBytes - 1024) / 28;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kNoAllocMaxDevices = 64;

// The device address of this GPU's scratch, looked up once per device (with
// that device current) and cached.  The array's name in host code is only a
// stand-in for the symbol: kernels and copies get the address from here.
static int noalloc_scratch(void** out) {
  static std::mutex mu;
  static void* ptrs[kNoAllocMaxDevices] = {};
  int dev = 0;
  SPLIT_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices)
    return fail(SPLIT_HIP_ERR_UNSUPPORTED, "device %d beyond %d", dev,
                kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    SPLIT_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p)
      return fail(SPLIT_HIP_ERR_HIP, "hipGetSymbolAddress gave no address");
    ptrs[dev] = p;
  }
  *out = ptrs[dev];
  return SPLIT_HIP_OK;
}

}  // namespace splithip

using namespace splithip;

extern "C" {

size_t split_hip_label_points_workspace(int64_t, int32_t) { return 0; }

int split_hip_label_points(const float* centroids, int32_t k,
                           const float* points, int64_t n, int32_t* labels,
                           void*, size_t, split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_points(points, n, k));
  if (!centroids || !labels)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "centroids or labels is null");
  StepArgs a{};
  a.pts = points;
  set_geometry(a, n, 0, 0);
  a.k = k;
  a.cent = centroids;
  a.labels = labels;
  const StepCfg c = step_cfg(a, false);
  return launch_step<true, false>(a, c, stream);
}

int split_hip_label_image(const float* centroids, int32_t k,
                          const float* points, int32_t width, int32_t height,
                          int32_t* labels, split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_image(width, height));
  const int64_t n = (int64_t)width * height;
  SPLIT_HIP_TRY(check_points(points, n, k));
  if (!centroids || !labels)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "centroids or labels is null");
  StepArgs a{};
  a.pts = points;
  set_geometry(a, n, width, height);
  a.k = k;
  a.cent = centroids;
  a.labels = labels;
  const StepCfg c = step_cfg(a, false);
  return launch_step<true, false>(a, c, stream);
}

size_t split_hip_calculate_centroids_workspace(int64_t, int32_t k) {
  if (k < 1) k = 1;
  if (k > SPLIT_HIP_MAX_K) return segment_means_workspace(k);
  return kStateBytes + partial_bytes(k);
}

int split_hip_calculate_centroids(const int32_t* labels, const float* points,
                                  int64_t n, float* centroids, int32_t k,
                                  void* temp, size_t temp_bytes,
                                  split_hip_stream_t stream) {
  return calculate_centroids_impl(labels, points, n, centroids, k, temp,
                                  temp_bytes, stream);
}

int split_hip_calculate_centroids_noalloc(const int32_t* labels,
                                          const float* points, int64_t n,
                                          float* centroids, int32_t k,
                                          split_hip_stream_t stream) {
  if (k > SPLIT_HIP_MAX_K && segment_means_workspace(k) > kNoAllocScratchBytes)
    return fail(SPLIT_HIP_ERR_UNSUPPORTED,
                "calculate_centroids_noalloc: k = %d needs %zu bytes of "
                "scratch, the static scratch holds %zu (k <= %lld)",
                k, segment_means_workspace(k), kNoAllocScratchBytes,
                (long long)kNoAllocMaxSegments);
  void* scratch_ptr = nullptr;
  SPLIT_HIP_TRY(noalloc_scratch(&scratch_ptr));
  return calculate_centroids_impl(labels, points, n, centroids, k, scratch_ptr,
                                  kNoAllocScratchBytes, stream);
}

size_t split_hip_cluster_workspace(int64_t n, int32_t k, int32_t) {
  if (k < 1) k = 1;
  const size_t runs = kStateBytes + align_up(6 * (size_t)k * sizeof(float), 256) +
                      partial_bytes(k);
  return std::max(runs, tiles_workspace(n, k));
}

int split_hip_cluster(const float* points, int64_t n, float* centroids,
                      int32_t k, int32_t* labels, void* temp,
                      size_t temp_bytes, int32_t max_iter, float threshold,
                      split_hip_cluster_result* result,
                      split_hip_stream_t stream) {
  Mailbox* mb = nullptr;
  SPLIT_HIP_TRY(cluster_async(points, n, 0, 0, centroids, k, labels, temp,
                              temp_bytes, max_iter, threshold, stream, &mb));
  SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
  if (result) {
    const volatile Mailbox* v = mb;
    result->iterations = v->iter;
    result->shift = v->shift;
  }
  return SPLIT_HIP_OK;
}

int split_hip_cluster_image(const float* points, int32_t width,
                            int32_t height, float* centroids, int32_t k,
                            int32_t* labels, void* temp, size_t temp_bytes,
                            int32_t max_iter, float threshold,
                            split_hip_cluster_result* result,
                            split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_image(width, height));
  Mailbox* mb = nullptr;
  SPLIT_HIP_TRY(cluster_async(points, (int64_t)width * height, width, height,
                              centroids, k, labels, temp, temp_bytes,
                              max_iter, threshold, stream, &mb));
  SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
  if (result) {
    const volatile Mailbox* v = mb;
    result->iterations = v->iter;
    result->shift = v->shift;
  }
  return SPLIT_HIP_OK;
}

int split_hip_uniform_random_initialize(const float* points, int64_t n,
                                        float* centroids, int32_t k,
                                        uint64_t seed,
                                        split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_points(points, n, k));
  if (!centroids) return fail(SPLIT_HIP_ERR_ARGUMENT, "centroids is null");
  if ((int64_t)k > n)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "k = %d exceeds n = %lld", k,
                (long long)n);
  hipLaunchKernelGGL(sample_init, dim3(1), dim3(256), 0, stream, points, n,
                     centroids, k, seed);
  SPLIT_HIP_CHECK_LAUNCH("sample_init");
  return SPLIT_HIP_OK;
}

}  // extern "C"
