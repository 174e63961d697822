// Erosion of a label image (split::device::morph::erode,
// split/src/split/device/morph/erode.cu:57-159), the step after connected
// components in the reference's sample (samples/src/main.cu:302-304).
//
// The reference erodes each row's runs of equal labels (erode.cu:57-139):
// a run [s, e) keeps the pixels p with s + n <= p < e - n and every other
// pixel becomes -1; erode() applies this to the columns (through a
// transpose) and then to the rows of the result.  Per pixel that is a
// window test: p survives a row sweep iff the 2n + 1 pixels centred on it
// lie inside the row and carry its label.
//
// Here each sweep is one pass with a run counter along the line:
//  * columns: a wave takes 64 adjacent columns (coalesced row segments) and
//    a chunk of kChunk rows, scans down from n rows above the chunk
//    (length of the equal-label run above each pixel, capped at n) into a
//    bit mask, then up from n rows below it, writing label or -1;
//  * rows, n <= 64: lanes along the row and ballots of run starts (see
//    erode_rows_bits); larger n: the column scan with one lane per row
//    over a chunk of kChunk columns.
// Every pixel is read about 2 (1 + n / kChunk) times (columns) or twice
// (rows) and written once per sweep.
#include <algorithm>

#include "common.hpp"

namespace splithip {
namespace {

constexpr int kChunk = 64;  // pixels per lane and sweep (one mask bit each)

// One sweep along `lines` lines of `len` pixels: pixel i of line l at
// in[l * lstride + i * pstride].  Lane = line (blockIdx.x * 64 + lane),
// blockIdx.y = chunk of kChunk pixels along the line.
__global__ __launch_bounds__(kWave) void erode_sweep(const int32_t* in,
                                                     int32_t* out, int lines,
                                                     int len, int64_t lstride,
                                                     int64_t pstride, int n) {
  const int line = blockIdx.x * kWave + threadIdx.x;
  if (line >= lines) return;
  const int p0 = blockIdx.y * kChunk;
  const int p1 = std::min(len, p0 + kChunk);
  const int32_t* src = in + (int64_t)line * lstride;
  int32_t* dst = out + (int64_t)line * lstride;
  // run of equal labels ending at p (looking back), capped at n
  uint64_t above = 0;
  {
    const int start = std::max(0, p0 - n);
    int prev = src[(int64_t)start * pstride];
    int run = 0;
    for (int p = start + 1; p < p1; ++p) {
      const int cur = src[(int64_t)p * pstride];
      run = cur == prev ? std::min(run + 1, n) : 0;
      prev = cur;
      if (p >= p0 && run >= n) above |= 1ull << (p - p0);
    }
  }
  // run looking forward, then the result
  {
    const int end = std::min(len - 1, p1 - 1 + n);
    int prev = src[(int64_t)end * pstride];
    int run = 0;
    if (end < p1) dst[(int64_t)end * pstride] = -1;  // the line's last pixel
    for (int p = end - 1; p >= p0; --p) {
      const int cur = src[(int64_t)p * pstride];
      run = cur == prev ? std::min(run + 1, n) : 0;
      prev = cur;
      if (p < p1) {
        const bool keep = run >= n && ((above >> (p - p0)) & 1ull);
        dst[(int64_t)p * pstride] = keep ? cur : -1;
      }
    }
  }
}

// Row sweep for n <= 64 with lanes along the row (coalesced): a wave takes
// kSpan consecutive 64-pixel segments of one row.  Bit j of a segment's
// change mask says that column seg * 64 + j starts a run (its label differs
// from the one on its left; columns <= 0 and >= width count as starts).
// With the masks of the segment and its two neighbours, each lane finds the
// last run start at or before its column and the next one after it, i.e.
// how many pixels of its run lie before and after it (at least 64 + 1 when
// the neighbouring mask has no start, which is enough for n <= 64).
constexpr int kSpan = 8;
constexpr int kMaxRowBitsN = 64;

__global__ __launch_bounds__(256) void erode_rows_bits(const int32_t* in,
                                                       int32_t* out, int width,
                                                       int height, int n) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int segs = (width + kWave - 1) / kWave;
  const int spans = (segs + kSpan - 1) / kSpan;
  if (wave >= (int64_t)height * spans) return;  // wave-uniform
  const int row = (int)(wave / spans);
  const int seg0 = (int)(wave - (int64_t)row * spans) * kSpan;
  const int32_t* src = in + (int64_t)row * width;
  int32_t* dst = out + (int64_t)row * width;
  int lab_next = 0;
  auto starts = [&](int seg, int& lab) {
    const int c = seg * kWave + lane;
    bool st = true;
    lab = 0;
    if (c > 0 && c < width) {
      lab = src[c];
      st = lab != src[c - 1];
    } else if (c == 0) {
      lab = src[0];
    }
    return (uint64_t)ballot(st);
  };
  int lab_prev, lab_cur;
  uint64_t m_prev = starts(seg0 - 1, lab_prev);
  uint64_t m_cur = starts(seg0, lab_cur);
  uint64_t m_next = starts(seg0 + 1, lab_next);
  const uint64_t upto = lane == 63 ? ~0ull : (2ull << lane) - 1;  // bits 0..lane
  const int seg1 = std::min(segs, seg0 + kSpan);
  for (int seg = seg0; seg < seg1; ++seg) {
    const uint64_t below = m_cur & upto;
    const uint64_t after_bits = m_cur & ~upto;
    int before, after;
    if (below)
      before = lane - (63 - __builtin_clzll(below));
    else if (m_prev)
      before = lane + 64 - (63 - __builtin_clzll(m_prev));
    else
      before = kMaxRowBitsN + 1;
    if (after_bits)
      after = __builtin_ctzll(after_bits) - 1 - lane;
    else if (m_next)
      after = 64 + __builtin_ctzll(m_next) - 1 - lane;
    else
      after = kMaxRowBitsN + 1;
    const int c = seg * kWave + lane;
    if (c < width) dst[c] = before >= n && after >= n ? lab_cur : -1;
    m_prev = m_cur;
    m_cur = m_next;
    lab_cur = lab_next;
    m_next = starts(seg + 2, lab_next);
  }
}

__global__ void copy_labels(const int32_t* in, int32_t* out, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

int launch_sweep(const int32_t* in, int32_t* out, int width, int height,
                 bool columns, int n, hipStream_t stream) {
  if (!columns && n <= kMaxRowBitsN) {
    const int64_t segs = (width + kWave - 1) / kWave;
    const int64_t waves = (int64_t)height * ((segs + kSpan - 1) / kSpan);
    const int64_t blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(erode_rows_bits, dim3((unsigned)blocks), dim3(256), 0,
                       stream, in, out, width, height, n);
    SPLIT_HIP_CHECK_LAUNCH("erode_rows_bits");
    return SPLIT_HIP_OK;
  }
  // columns: lines = columns (adjacent lanes read adjacent words)
  const int lines = columns ? width : height;
  const int len = columns ? height : width;
  const int64_t lstride = columns ? 1 : width;
  const int64_t pstride = columns ? width : 1;
  const dim3 grid((lines + kWave - 1) / kWave, (len + kChunk - 1) / kChunk);
  hipLaunchKernelGGL(erode_sweep, grid, dim3(kWave), 0, stream, in, out, lines,
                     len, lstride, pstride, n);
  SPLIT_HIP_CHECK_LAUNCH("erode_sweep");
  return SPLIT_HIP_OK;
}

int check_erode(const int32_t* labels, int width, int height, int n,
                void* temp, size_t temp_bytes, const char* what) {
  if (!labels || width < 1 || height < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "%s: labels missing or empty image", what);
  if (n < 0) return fail(SPLIT_HIP_ERR_ARGUMENT, "%s: niterations < 0", what);
  if ((int64_t)width * height > 0x7fffffffLL)
    return fail(SPLIT_HIP_ERR_UNSUPPORTED, "%s: more than 2^31 - 1 pixels", what);
  const size_t need = (size_t)width * height * sizeof(int32_t);
  if (n > 0 && (!temp || temp_bytes < need))
    return fail(SPLIT_HIP_ERR_TEMP, "%s: temp holds %zu bytes, needs %zu", what,
                temp_bytes, need);
  return SPLIT_HIP_OK;
}

}  // namespace
}  // namespace splithip

using namespace splithip;

extern "C" {

size_t split_hip_erode_workspace(int32_t width, int32_t height) {
  if (width < 1 || height < 1) return 0;
  return align_up((size_t)width * height * sizeof(int32_t), 256);
}

int split_hip_erode_horizontal_sweep(int32_t* labels, int32_t width,
                                     int32_t height, int32_t niterations,
                                     void* temp, size_t temp_bytes,
                                     split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_erode(labels, width, height, niterations, temp, temp_bytes,
                            "erode_horizontal_sweep"));
  if (niterations == 0) return SPLIT_HIP_OK;
  const int64_t count = (int64_t)width * height;
  int32_t* copy = static_cast<int32_t*>(temp);
  hipLaunchKernelGGL(copy_labels, dim3(4 * device_cu_count()), dim3(256), 0,
                     stream, labels, copy, count);
  SPLIT_HIP_CHECK_LAUNCH("erode copy");
  return launch_sweep(copy, labels, width, height, false, niterations, stream);
}

int split_hip_erode(int32_t* labels, int32_t width, int32_t height,
                    int32_t niterations, void* temp, size_t temp_bytes,
                    split_hip_stream_t stream) {
  SPLIT_HIP_TRY(check_erode(labels, width, height, niterations, temp, temp_bytes,
                            "erode"));
  if (niterations == 0) return SPLIT_HIP_OK;
  int32_t* mid = static_cast<int32_t*>(temp);
  SPLIT_HIP_TRY(launch_sweep(labels, mid, width, height, true, niterations, stream));
  return launch_sweep(mid, labels, width, height, false, niterations, stream);
}

}  // extern "C"
