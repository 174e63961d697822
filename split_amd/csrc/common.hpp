// Shared helpers of the gfx950 kernels and their C-ABI launchers: status
// handling, wave64 cross-lane reductions built on DPP row operations, and
// the device-side bookkeeping structs that live in the caller's temp block.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "split_hip.h"

namespace splithip {

// ---------------------------------------------------------------- status ---

void set_error(const char* fmt, ...);
int fail(int status, const char* fmt, ...);

#define SPLIT_HIP_CHECK(expr)                                                \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return ::splithip::fail(SPLIT_HIP_ERR_HIP, "%s failed: %s (%s:%d)",    \
                              #expr, hipGetErrorString(e_), __FILE__,        \
                              __LINE__);                                     \
  } while (0)

#define SPLIT_HIP_CHECK_LAUNCH(name)                                         \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess)                                                    \
      return ::splithip::fail(SPLIT_HIP_ERR_HIP, "launch of %s failed: %s",  \
                              name, hipGetErrorString(e_));                  \
  } while (0)

#define SPLIT_HIP_TRY(expr)                                                  \
  do {                                                                       \
    int s_ = (expr);                                                         \
    if (s_ != SPLIT_HIP_OK) return s_;                                       \
  } while (0)

// Number of compute units of the current device (cached per device).
int device_cu_count();

constexpr int kWave = 64;

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ------------------------------------------------------ wave reductions ---
//
// Reductions run on all 64 lanes (callers keep EXEC full).  Inside each row
// of 16 lanes: quad xor-1, quad xor-2, half-row mirror, row mirror — after
// which every lane of a row holds the row result.  Then row_bcast:15 folds
// row 0 into row 1 and row 2 into row 3, row_bcast:31 folds rows 0-1 into
// rows 2-3, and lane 63 holds the wave result, read with v_readlane.  The
// combination order is fixed, so float sums are bit-identical run to run.

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp_i(int x, int old = 0) {
  return __builtin_amdgcn_update_dpp(old, x, CTRL, ROW_MASK, 0xF, false);
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_f(float x, float old = 0.f) {
  return __int_as_float(
      dpp_i<CTRL, ROW_MASK>(__float_as_int(x), __float_as_int(old)));
}

__device__ __forceinline__ float readlane_f(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}
__device__ __forceinline__ int readlane_i(int x, int lane) {
  return __builtin_amdgcn_readlane(x, lane);
}
__device__ __forceinline__ int readfirstlane_i(int x) {
  return __builtin_amdgcn_readfirstlane(x);
}

constexpr int kDppQuadXor1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

struct OpAdd {
  static constexpr float identity = 0.f;
  __device__ float operator()(float a, float b) const { return a + b; }
};
struct OpMin {
  static constexpr float identity = INFINITY;
  __device__ float operator()(float a, float b) const { return fminf(a, b); }
};
struct OpMax {
  static constexpr float identity = -INFINITY;
  __device__ float operator()(float a, float b) const { return fmaxf(a, b); }
};

template <class Op>
__device__ __forceinline__ float wave_allreduce(float x, Op op) {
  x = op(x, dpp_f<kDppQuadXor1>(x));
  x = op(x, dpp_f<kDppQuadXor2>(x));
  x = op(x, dpp_f<kDppRowHalfMirror>(x));
  x = op(x, dpp_f<kDppRowMirror>(x));
  x = op(x, dpp_f<kDppRowBcast15, 0xA>(x, Op::identity));
  x = op(x, dpp_f<kDppRowBcast31, 0xC>(x, Op::identity));
  return readlane_f(x, 63);
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned x) {
  x += (unsigned)dpp_i<kDppQuadXor1>((int)x);
  x += (unsigned)dpp_i<kDppQuadXor2>((int)x);
  x += (unsigned)dpp_i<kDppRowHalfMirror>((int)x);
  x += (unsigned)dpp_i<kDppRowMirror>((int)x);
  x += (unsigned)dpp_i<kDppRowBcast15, 0xA>((int)x, 0);
  x += (unsigned)dpp_i<kDppRowBcast31, 0xC>((int)x, 0);
  return (unsigned)readlane_i((int)x, 63);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// 16-byte vectors that need only 4-byte alignment.  gfx950 global memory
// takes dword-aligned dwordx4 accesses, so a lane's four consecutive pixels
// (or labels, or states) move with one instruction whatever the image
// width or plane offset.
struct __attribute__((aligned(4))) F4 {
  float x, y, z, w;
};
struct __attribute__((aligned(4))) U4 {
  uint32_t x, y, z, w;
};
struct __attribute__((aligned(4))) I4 {
  int32_t x, y, z, w;
};
__host__ __device__ inline bool aligned4(const void* p) {
  return ((uintptr_t)p & 3u) == 0;
}

// Global-memory view of a generic pointer.  Pointers that reach a kernel
// through memory (the batched path's per-image records) are generic to the
// compiler, and its flat accesses count against both the vector-memory and
// the LDS wait counters, so every LDS wait would also wait for outstanding
// loads; accesses through gmem() compile to global_* instructions.  Only for
// pointers into device memory.  (The host pass, which only parses device
// code, sees the identity.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
__device__ __forceinline__ T __attribute__((address_space(1)))* gmem(T* p) {
  return (T __attribute__((address_space(1)))*)p;
}
#else
template <class T>
__device__ __forceinline__ T* gmem(T* p) {
  return p;
}
#endif

// -------------------------------------------------- temp-block structs ---

// Lloyd loop state, kept in the caller's temp block.  Written by the last
// block of each step launch (1-D runs path) or by block 0 of each launch's
// finalisation (tiled path); read by the blocks of later launches.
struct LloydState {
  unsigned arrive;  // blocks of the current launch that have finished
  int cur;          // which of the two centroid buffers is current
  int iter;         // Lloyd iterations completed
  int done;         // set once converged or max_iter reached
  int max_iter;
  float threshold;
  float shift;      // Frobenius norm of the last centroid update
  unsigned next;    // tile chunks handed out in the current launch
};

// Host-mapped mailbox the finalising block of a launch writes, so the host
// can stop enqueuing iterations without a device->host copy.
struct alignas(16) Mailbox {
  int iter;
  int done;
  float shift;
  int reserved;
};
static_assert(sizeof(Mailbox) == 16, "written as one 16-byte store");

// Per-thread resources of the Lloyd drivers: a host-mapped mailbox and two
// events, created on first use and kept for the life of the thread.
struct DriverRes {
  Mailbox* host = nullptr;
  Mailbox* dev = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  int device = -1;
};
int driver_res(DriverRes** out);

// Bounded tiled Lloyd (lloyd_tiles.hip).
// Largest |v| of each of the three planes of n floats as float bits into
// out[0..3) (which must start at 0), stream-ordered.
int launch_absmax(const float* pts, int64_t n, unsigned* out,
                  hipStream_t stream);
// calculate_centroids for any k (segment means; exact integer sums).
size_t segment_means_workspace(int64_t k);
int segment_means(const int32_t* labels, const float* points, int64_t n,
                  float* centroids, int64_t k, void* temp, size_t temp_bytes,
                  hipStream_t stream);
size_t tiles_workspace(int64_t n, int k);
bool tiles_supported(int64_t n, int width, int k);
// Bounded Lloyd over the n pixels as rows of `width` (the last row may be
// partial).
int cluster_tiles_async(const float* pts, int64_t n, int width,
                        float* centroids, int k, int32_t* labels, void* temp,
                        size_t temp_bytes, int max_iter, float threshold,
                        hipStream_t stream, Mailbox** mbox_out);

}  // namespace splithip
