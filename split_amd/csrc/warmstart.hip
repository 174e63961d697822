// Warm-start network: a small CNN that predicts 16 starting centroids from
// an image, with its inference and training kernels.
//
//   input (B, 256, 256, 3) bf16 NHWC (area-average resample of the image)
//   conv1..conv4: 3x3 stride 2 pad 1, 3->16->32->64->64, ReLU  (bf16 out)
//   global average pool -> 64 (fp32)
//   fc5 64->128 ReLU (bf16), fc6 128->48 (fp32 logits), sigmoid
//   loss (1/B) sum (sigmoid - target)^2, or the Chamfer set distance
//   between the 16 predicted and 16 target colours
//
// Every layer — forward, weight gradient and input gradient — is one
// implicit GEMM on the matrix cores (v_mfma_f32_16x16x32_bf16: bf16
// operands, fp32 accumulation).  The GEMM kernel is generic over three
// functors: the A and B loaders gather their operand tiles straight from the
// activation / weight / gradient tensors (im2col and its transposes are
// never materialised) and round them to bf16 into LDS; the epilogue stores
// (ReLU + bf16, fp32, or per-split partials for long reductions, which a
// second pass sums in a fixed order).  Biases are an extra weight column
// multiplied by a constant-1 input column, so the bias gradient falls out
// of the weight-gradient GEMM.
//
// Parameter layout (float32 master copy, bf16 compute copy): layer l holds
// W_l[n][0:Kp_l] with columns (ky*3+kx)*Cin + ci (conv) or ci (dense), the
// bias at column K_l, zeros up to Kp_l = roundup(K_l + 1, 32).
#include <algorithm>
#include <vector>

#include "common.hpp"

namespace splithip {
namespace ws {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kSide = 256;
constexpr int kOut = 3 * SPLIT_HIP_WS_K;  // 48 outputs
constexpr int kBM = 64, kBK = 32, kThreads = 256;
// shorts of padding per LDS row: rows stay 16-byte aligned for the MFMA
// operands' 16-byte LDS reads (a row is 80 bytes)
constexpr int kLdsPad = 8;
constexpr int kMaxSplits = 256;

struct LayerDim {
  int conv, cin, cout, K, Kp, off;
  int hin, hout;  // spatial side in / out (conv)
};

__host__ __device__ constexpr int kp_of(int K) { return (K + 1 + 31) / 32 * 32; }

// conv1..4, fc5, fc6
constexpr int kCin[6] = {3, 16, 32, 64, 64, 128};
constexpr int kCout[6] = {16, 32, 64, 64, 128, kOut};

static std::vector<LayerDim> layer_dims() {
  std::vector<LayerDim> d;
  int off = 0, side = kSide;
  for (int l = 0; l < 6; ++l) {
    LayerDim L{};
    L.conv = l < 4;
    L.cin = kCin[l];
    L.cout = kCout[l];
    L.K = L.conv ? 9 * L.cin : L.cin;
    L.Kp = kp_of(L.K);
    L.off = off;
    off += L.cout * L.Kp;
    L.hin = side;
    L.hout = L.conv ? (side + 1) / 2 : 1;
    if (L.conv) side = L.hout;
    d.push_back(L);
  }
  return d;
}

static size_t param_count() {
  const auto d = layer_dims();
  return (size_t)d.back().off + (size_t)d.back().cout * d.back().Kp;
}

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}
__device__ __forceinline__ void unpack8(const uint4 q, float (&v)[8]) {
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[2 * i] = __uint_as_float(w[i] << 16);
    v[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
  }
}

// ------------------------------------------------------------- loaders ---
// load8(i, r, v): v[e] = operand(i, r + e), e = 0..7, r a multiple of 8;
// zero where the operand is zero (padding, outside the reduction range).

// Forward A of a conv layer: im2col of the bf16 NHWC input, rows = output
// pixels (b, oy, ox), columns k = (ky*3+kx)*Cin + ci, column K = 1.
struct ConvA {
  const uint16_t* X;
  const int32_t* index;  // batch slot -> image (input of conv1 only)
  int64_t img;           // elements per image of X
  int Hi, Cin, Ho, K;
  __device__ void load8(int64_t i, int r, float (&v)[8]) const {
    const int hw = Ho * Ho;
    const int b = (int)(i / hw);
    const int rem = (int)(i - (int64_t)b * hw);
    const int oy = rem / Ho, ox = rem - oy * Ho;
    const uint16_t* base = X + (int64_t)(index ? index[b] : b) * img;
    if (Cin % 8 == 0 && r + 8 <= K) {
      const int tap = r / Cin, ci = r - tap * Cin;
      const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
      if (iy >= 0 && iy < Hi && ix >= 0 && ix < Hi) {
        unpack8(*reinterpret_cast<const uint4*>(
                    base + ((int64_t)iy * Hi + ix) * Cin + ci), v);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = 0.f;
      }
      return;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = r + e;
      float x = 0.f;
      if (k < K) {
        const int tap = k / Cin, ci = k - tap * Cin;
        const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
        if (iy >= 0 && iy < Hi && ix >= 0 && ix < Hi)
          x = bf2f(base[((int64_t)iy * Hi + ix) * Cin + ci]);
      } else if (k == K) {
        x = 1.f;
      }
      v[e] = x;
    }
  }
};

// Weight-gradient A of a conv layer: the transpose of ConvA (rows k,
// reduction over output pixels m < M).
struct ConvAT {
  ConvA a;
  int64_t M;
  __device__ void load8(int64_t k, int64_t r, float (&v)[8]) const {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float x = 0.f;
      const int64_t m = r + e;
      if (m < M && k <= a.K) {
        if (k == a.K) {
          x = 1.f;
        } else {
          const int hw = a.Ho * a.Ho;
          const int b = (int)(m / hw);
          const int rem = (int)(m - (int64_t)b * hw);
          const int oy = rem / a.Ho, ox = rem - oy * a.Ho;
          const int tap = (int)k / a.Cin, ci = (int)k - tap * a.Cin;
          const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
          if (iy >= 0 && iy < a.Hi && ix >= 0 && ix < a.Hi)
            x = bf2f(a.X[(int64_t)(a.index ? a.index[b] : b) * a.img +
                         ((int64_t)iy * a.Hi + ix) * a.Cin + ci]);
        }
      }
      v[e] = x;
    }
  }
};

// Dense A: rows b, columns k < Cin from a row-major activation (fp32 or
// bf16), column Cin = 1.
template <class T>
struct DenseA {
  const T* F;
  int Cin;
  __device__ float at(int64_t b, int k) const {
    if constexpr (sizeof(T) == 2)
      return bf2f(F[b * Cin + k]);
    else
      return F[b * Cin + k];
  }
  __device__ void load8(int64_t b, int r, float (&v)[8]) const {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = r + e;
      v[e] = k < Cin ? at(b, k) : (k == Cin ? 1.f : 0.f);
    }
  }
};

// Its transpose for the weight gradient (rows k, reduction over b < M).
template <class T>
struct DenseAT {
  DenseA<T> a;
  int64_t M;
  __device__ void load8(int64_t k, int64_t r, float (&v)[8]) const {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t b = r + e;
      v[e] = (b < M && k <= a.Cin) ? (k < a.Cin ? a.at(b, (int)k) : 1.f) : 0.f;
    }
  }
};

// Forward B: weight rows W[n][k] (bf16, stride Kp), as B(k, n).
struct WeightB {
  const uint16_t* W;
  int Kp;
  __device__ void load8(int64_t r, int64_t n, float (&v)[8]) const {
    if (r + 8 <= Kp) {
      unpack8(*reinterpret_cast<const uint4*>(W + n * Kp + r), v);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = 0.f;
    }
  }
};

// Gradient at a layer output through its ReLU: G(m, n) = Y > 0 ? dY : 0,
// Y bf16 [M][N], dY fp32 [M][N] (Y == null: no ReLU), as the B operand of
// the weight gradient (reduction over m).
struct GradB {
  const float* dY;
  const uint16_t* Y;
  int N;
  int64_t M;
  __device__ float at(int64_t m, int64_t n) const {
    const float g = dY[m * N + n];
    return (!Y || bf2f(Y[m * N + n]) > 0.f) ? g : 0.f;
  }
  __device__ void load8(int64_t r, int64_t n, float (&v)[8]) const {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = r + e < M ? at(r + e, n) : 0.f;
  }
};

// The same gradient as the A operand of a dense input gradient (rows b,
// reduction over n < N).
struct GradA {
  GradB g;
  __device__ void load8(int64_t b, int r, float (&v)[8]) const {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = r + e < g.N ? g.at(b, r + e) : 0.f;
  }
};

// Input gradient of a conv layer, A operand: rows = input pixels
// (b, iy, ix), reduction r = tap*Cout + co; the output pixel a tap of
// (iy, ix) feeds is (oy, ox) = ((iy+1-ky)/2, (ix+1-kx)/2) when both are
// whole and inside the output.
struct ConvGradA {
  GradB g;  // at the layer output, N = Cout
  int Hi, Ho;
  __device__ void load8(int64_t p, int r, float (&v)[8]) const {
    const int Cout = g.N;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = 0.f;
    if (r >= 9 * Cout) return;
    const int hw = Hi * Hi;
    const int b = (int)(p / hw);
    const int rem = (int)(p - (int64_t)b * hw);
    const int iy = rem / Hi, ix = rem - iy * Hi;
    const int tap = r / Cout, co = r - tap * Cout;  // Cout % 8 == 0
    const int ty = iy + 1 - tap / 3, tx = ix + 1 - tap % 3;
    if (ty < 0 || tx < 0 || (ty & 1) || (tx & 1)) return;
    const int oy = ty >> 1, ox = tx >> 1;
    if (oy >= Ho || ox >= Ho) return;
    const int64_t m = ((int64_t)b * Ho + oy) * Ho + ox;
    const float4 d0 = *reinterpret_cast<const float4*>(g.dY + m * Cout + co);
    const float4 d1 = *reinterpret_cast<const float4*>(g.dY + m * Cout + co + 4);
    float y[8];
    unpack8(*reinterpret_cast<const uint4*>(g.Y + m * Cout + co), y);
    const float d[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = y[e] > 0.f ? d[e] : 0.f;
  }
};

// Input-gradient B of a conv layer: B(r = tap*Cout + co, ci) =
// W[co][tap*Cin + ci].
struct ConvGradB {
  const uint16_t* W;
  int Kp, Cin, Cout;
  __device__ void load8(int64_t r, int64_t ci, float (&v)[8]) const {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int rr = (int)r + e;
      float x = 0.f;
      if (rr < 9 * Cout) {
        const int tap = rr / Cout, co = rr - tap * Cout;
        x = bf2f(W[(int64_t)co * Kp + tap * Cin + ci]);
      }
      v[e] = x;
    }
  }
};

// Input-gradient B of a dense layer: B(n, ci) = W[n][ci], n < N.
struct DenseGradB {
  const uint16_t* W;
  int Kp, N;
  __device__ void load8(int64_t r, int64_t ci, float (&v)[8]) const {
#pragma unroll
    for (int e = 0; e < 8; ++e)
      v[e] = r + e < N ? bf2f(W[(r + e) * Kp + ci]) : 0.f;
  }
};

// ----------------------------------------------------------- epilogues ---

struct EpReluBf16 {
  uint16_t* out;
  int ld;
  __device__ void operator()(int, int64_t i, int64_t n, float v) const {
    out[i * ld + n] = f2bf(fmaxf(v, 0.f));
  }
};
struct EpF32 {
  float* out;
  int ld;
  __device__ void operator()(int, int64_t i, int64_t n, float v) const {
    out[i * ld + n] = v;
  }
};
struct EpPartial {
  float* part;
  int64_t M;
  int N;
  __device__ void operator()(int z, int64_t i, int64_t n, float v) const {
    part[((int64_t)z * M + i) * N + n] = v;
  }
};

// ---------------------------------------------------------------- GEMM ---
//
// C[i][n] = sum_r A(i, r) B(r, n) for i < M, n < N, r in this block's split
// of [0, R).  Block: 256 threads = 4 waves; tile 64 x BN; wave w owns rows
// 16w..16w+15 and every 16-column subtile.  Per 32-deep step each thread
// loads 8 A values (row t/4, columns 8*(t%4)..) and up to 8 B values, rounds
// them to bf16 into LDS, then each wave issues one gfx950
// v_mfma_f32_16x16x32_bf16 per subtile.
//
// Operands: lane l holds row (A) or column (B) l%16 and the 8 k values
// 8*(l/16) .. 8*(l/16)+7 of the 32-deep step (A and B use the same k
// assignment, so each product pairs matching k); D[i][j] in lane
// j + 16*(i/4), element i%4.
template <int BN, class LA, class LB, class EP>
__global__ __launch_bounds__(kThreads) void gemm_kernel(int64_t M, int64_t N,
                                                        int64_t R,
                                                        int64_t rchunk, LA la,
                                                        LB lb, EP ep) {
  __shared__ __align__(16) uint16_t As[kBM][kBK + kLdsPad];
  __shared__ __align__(16) uint16_t Bs[BN][kBK + kLdsPad];
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t n0 = (int64_t)blockIdx.y * BN;
  const int64_t rbeg = (int64_t)blockIdx.z * rchunk;
  const int64_t rend = std::min<int64_t>(R, rbeg + rchunk);
  f4 acc[BN / 16];
#pragma unroll
  for (int j = 0; j < BN / 16; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
  const int lrow = t >> 2, lseg = (t & 3) * 8;

  for (int64_t r0 = rbeg; r0 < rend; r0 += kBK) {
    {
      float v[8];
      if (m0 + lrow < M) {
        la.load8(m0 + lrow, r0 + lseg, v);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = 0.f;
      }
      uint32_t pk[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float a0 = r0 + lseg + 2 * e < rend ? v[2 * e] : 0.f;
        const float a1 = r0 + lseg + 2 * e + 1 < rend ? v[2 * e + 1] : 0.f;
        pk[e] = (uint32_t)f2bf(a0) | ((uint32_t)f2bf(a1) << 16);
      }
      uint32_t* dst = reinterpret_cast<uint32_t*>(&As[lrow][lseg]);
      dst[0] = pk[0]; dst[1] = pk[1]; dst[2] = pk[2]; dst[3] = pk[3];
    }
    if (lrow < BN) {
      float v[8];
      if (n0 + lrow < N) {
        lb.load8(r0 + lseg, n0 + lrow, v);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = 0.f;
      }
      uint32_t pk[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float b0 = r0 + lseg + 2 * e < rend ? v[2 * e] : 0.f;
        const float b1 = r0 + lseg + 2 * e + 1 < rend ? v[2 * e + 1] : 0.f;
        pk[e] = (uint32_t)f2bf(b0) | ((uint32_t)f2bf(b1) << 16);
      }
      uint32_t* dst = reinterpret_cast<uint32_t*>(&Bs[lrow][lseg]);
      dst[0] = pk[0]; dst[1] = pk[1]; dst[2] = pk[2]; dst[3] = pk[3];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kBK; ks += 32) {
      const bf8 a = *reinterpret_cast<const bf8*>(
          &As[16 * w + (lane & 15)][ks + 8 * (lane >> 4)]);
#pragma unroll
      for (int j = 0; j < BN / 16; ++j) {
        const bf8 b = *reinterpret_cast<const bf8*>(
            &Bs[16 * j + (lane & 15)][ks + 8 * (lane >> 4)]);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < BN / 16; ++j) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t i = m0 + 16 * w + 4 * (lane >> 4) + e;
      const int64_t n = n0 + 16 * j + (lane & 15);
      if (i < M && n < N) ep((int)blockIdx.z, i, n, acc[j][e]);
    }
  }
}

template <int BN, class LA, class LB, class EP>
static void gemm_bn(int64_t M, int64_t N, int64_t R, int splits, LA la, LB lb,
                    EP ep, hipStream_t s) {
  const int64_t rchunk = ((R + splits - 1) / splits + kBK - 1) / kBK * kBK;
  const dim3 grid((unsigned)((M + kBM - 1) / kBM), (unsigned)((N + BN - 1) / BN),
                  (unsigned)splits);
  hipLaunchKernelGGL((gemm_kernel<BN, LA, LB, EP>), grid, dim3(kThreads), 0, s,
                     M, N, R, rchunk, la, lb, ep);
}

template <class LA, class LB, class EP>
static void gemm(int64_t M, int64_t N, int64_t R, int splits, LA la, LB lb,
                 EP ep, hipStream_t s) {
  if (N <= 16)
    gemm_bn<16>(M, N, R, splits, la, lb, ep, s);
  else if (N <= 32)
    gemm_bn<32>(M, N, R, splits, la, lb, ep, s);
  else
    gemm_bn<64>(M, N, R, splits, la, lb, ep, s);
}

// --------------------------------------------------------- small kernels ---

__global__ void cast_bf16(const float* p, uint16_t* q, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    q[i] = f2bf(p[i]);
}

// grad[n*Kp + k] = sum over z of part[z][k][n] (fixed order).
__global__ void reduce_partials(const float* part, int splits, int Kp, int N,
                                float* grad) {
  const int64_t total = (int64_t)Kp * N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(e / N), n = (int)(e - (int64_t)k * N);
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += part[((int64_t)z * Kp + k) * N + n];
    grad[(int64_t)n * Kp + k] = s;
  }
}

// Global average pool of (B, P, C) bf16 -> (B, C) fp32, positions in order.
__global__ void gap_forward(const uint16_t* Y, int P, int C, int B, float* F) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= B * C) return;
  const int b = e / C, c = e - b * C;
  float s = 0.f;
  for (int p = 0; p < P; ++p) s += bf2f(Y[((int64_t)b * P + p) * C + c]);
  F[e] = s * (1.f / (float)P);
}

__global__ void gap_backward(const float* dF, int P, int C, int B, float* dY) {
  const int64_t total = (int64_t)B * P * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t b = e / ((int64_t)P * C);
    dY[e] = dF[b * C + c] * (1.f / (float)P);
  }
}

// Sigmoid outputs, loss and its logit gradient; one block, fixed order.
__global__ __launch_bounds__(256) void sigmoid_loss(const float* Z,
                                                    const float* T,
                                                    const int32_t* index,
                                                    int B, float* P, float* dZ,
                                                    float* loss) {
  __shared__ float part[256];
  float s = 0.f;
  const int total = B * kOut;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int b = e / kOut, j = e - b * kOut;
    const float p = 1.f / (1.f + expf(-Z[e]));
    P[e] = p;
    if (T) {
      const float d = p - T[(int64_t)(index ? index[b] : b) * kOut + j];
      s += d * d;
      dZ[e] = 2.f * d * p * (1.f - p) / (float)B;
    }
  }
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0 && loss) {
    float tot = 0.f;
    for (int i = 0; i < (int)blockDim.x; ++i) tot += part[i];
    *loss = tot / (float)B;
  }
}

// Chamfer set distance between each image's 16 predicted and 16 target
// colours, (1/B) sum_b [sum_t min_p |t-p|^2 + sum_p min_t |p-t|^2], and its
// logit gradient (ties to the lower index); one thread per image, the loss
// summed in image order by one thread.
__global__ __launch_bounds__(256) void chamfer_loss(const float* Z,
                                                    const float* T,
                                                    const int32_t* index,
                                                    int B, float* P, float* dZ,
                                                    float* loss) {
  constexpr int k = SPLIT_HIP_WS_K;
  __shared__ float part[256];
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    const float* z = Z + (int64_t)b * kOut;
    const float* t = T + (int64_t)(index ? index[b] : b) * kOut;
    float p[kOut], g[kOut];
#pragma unroll
    for (int e = 0; e < kOut; ++e) {
      p[e] = 1.f / (1.f + expf(-z[e]));
      g[e] = 0.f;
    }
    float L = 0.f;
    for (int i = 0; i < k; ++i) {  // each target pulls its nearest prediction
      float best = INFINITY;
      int bj = 0;
      for (int j = 0; j < k; ++j) {
        const float dx = p[j] - t[i], dy = p[k + j] - t[k + i],
                    dz = p[2 * k + j] - t[2 * k + i];
        const float d = dx * dx + dy * dy + dz * dz;
        if (d < best) { best = d; bj = j; }
      }
      L += best;
#pragma unroll
      for (int c = 0; c < 3; ++c) g[c * k + bj] += 2.f * (p[c * k + bj] - t[c * k + i]);
    }
    for (int j = 0; j < k; ++j) {  // each prediction moves to its nearest target
      float best = INFINITY;
      int bi = 0;
      for (int i = 0; i < k; ++i) {
        const float dx = p[j] - t[i], dy = p[k + j] - t[k + i],
                    dz = p[2 * k + j] - t[2 * k + i];
        const float d = dx * dx + dy * dy + dz * dz;
        if (d < best) { best = d; bi = i; }
      }
      L += best;
#pragma unroll
      for (int c = 0; c < 3; ++c) g[c * k + j] += 2.f * (p[c * k + j] - t[c * k + bi]);
    }
#pragma unroll
    for (int e = 0; e < kOut; ++e) {
      P[(int64_t)b * kOut + e] = p[e];
      dZ[(int64_t)b * kOut + e] = g[e] / (float)B * p[e] * (1.f - p[e]);
    }
    part[b] = L;  // B <= 256 per call of this kernel (checked by the host)
  }
  __syncthreads();
  if (threadIdx.x == 0 && loss) {
    float tot = 0.f;
    for (int b = 0; b < B; ++b) tot += part[b];
    *loss = tot / (float)B;
  }
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v,
                            int64_t n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

struct PackImg {
  const float* pts;
  int W, H;
};

// Area-average resample of SoA float images to (256, 256, 3) bf16: output
// cell X averages source columns [X*W/256, max((X+1)*W/256, that + 1)),
// rows likewise; summed in float (x inner), divided, rounded to bf16.
__global__ void pack_kernel(const PackImg* imgs, uint16_t* out) {
  const PackImg m = imgs[blockIdx.y];
  const int64_t n = (int64_t)m.W * m.H;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < kSide * kSide;
       e += gridDim.x * blockDim.x) {
    const int Y = e / kSide, X = e - Y * kSide;
    const int x0 = (int)((int64_t)X * m.W / kSide);
    const int x1 = std::max((int)((int64_t)(X + 1) * m.W / kSide), x0 + 1);
    const int y0 = (int)((int64_t)Y * m.H / kSide);
    const int y1 = std::max((int)((int64_t)(Y + 1) * m.H / kSide), y0 + 1);
    const float cnt = (float)((x1 - x0) * (y1 - y0));
    uint16_t* o = out + ((int64_t)blockIdx.y * kSide * kSide + e) * 3;
    for (int c = 0; c < 3; ++c) {
      const float* pl = m.pts + c * n;
      float s = 0.f;
      for (int yy = y0; yy < y1; ++yy)
        for (int xx = x0; xx < x1; ++xx) s += pl[(int64_t)yy * m.W + xx];
      o[c] = f2bf(__fdiv_rn(s, cnt));
    }
  }
}

// The same area-average resample to float SoA planes (count x 3 x 256^2):
// the coarse image on which the predicted centroids are refined.
__global__ void coarse_kernel(const PackImg* imgs, float* out) {
  const PackImg m = imgs[blockIdx.y];
  const int64_t n = (int64_t)m.W * m.H;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < kSide * kSide;
       e += gridDim.x * blockDim.x) {
    const int Y = e / kSide, X = e - Y * kSide;
    const int x0 = (int)((int64_t)X * m.W / kSide);
    const int x1 = std::max((int)((int64_t)(X + 1) * m.W / kSide), x0 + 1);
    const int y0 = (int)((int64_t)Y * m.H / kSide);
    const int y1 = std::max((int)((int64_t)(Y + 1) * m.H / kSide), y0 + 1);
    const float cnt = (float)((x1 - x0) * (y1 - y0));
    float* o = out + (int64_t)blockIdx.y * 3 * kSide * kSide + e;
    for (int c = 0; c < 3; ++c) {
      const float* pl = m.pts + c * n;
      float s = 0.f;
      for (int yy = y0; yy < y1; ++yy)
        for (int xx = x0; xx < x1; ++xx) s += pl[(int64_t)yy * m.W + xx];
      o[(int64_t)c * kSide * kSide] = __fdiv_rn(s, cnt);
    }
  }
}

__global__ void scatter_centroids(const float* P, float* const* dst, int B) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= B * kOut) return;
  const int b = e / kOut, j = e - b * kOut;
  dst[b][j] = P[e];
}

// Per image: 16 centroids (3, 16) -> ordered by luminance
// Y = (0.299 R + 0.587 G) + 0.114 B (each operation rounded), ties to the
// lower index.
__global__ void targets_kernel(const float* C, int count, float* T) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= count) return;
  constexpr int k = SPLIT_HIP_WS_K;
  const float* c = C + (int64_t)b * kOut;
  float y[k];
  int order[k];
  for (int j = 0; j < k; ++j) {
    y[j] = __fadd_rn(__fadd_rn(__fmul_rn(0.299f, c[j]), __fmul_rn(0.587f, c[k + j])),
                     __fmul_rn(0.114f, c[2 * k + j]));
    order[j] = j;
  }
  for (int i = 1; i < k; ++i) {  // insertion sort, stable
    const int o = order[i];
    int j = i - 1;
    while (j >= 0 && y[order[j]] > y[o]) {
      order[j + 1] = order[j];
      --j;
    }
    order[j + 1] = o;
  }
  float* t = T + (int64_t)b * kOut;
  for (int j = 0; j < k; ++j)
    for (int ch = 0; ch < 3; ++ch) t[ch * k + j] = c[ch * k + order[j]];
}

// Squared distance of each pixel to its nearest centroid (the assignment's
// fmaf form), summed in double: per-block partials, then one block in order.
__global__ __launch_bounds__(256) void inertia_partial(const float* pts,
                                                       int64_t n,
                                                       const float* cent,
                                                       int k, double* part) {
  __shared__ float cs[3 * SPLIT_HIP_MAX_K];
  __shared__ double red[256];
  for (int i = threadIdx.x; i < 3 * k; i += blockDim.x) cs[i] = cent[i];
  __syncthreads();
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float x = pts[i], y = pts[n + i], z = pts[2 * n + i];
    float best = INFINITY;
    for (int j = 0; j < k; ++j) {
      const float dx = x - cs[j], dy = y - cs[k + j], dz = z - cs[2 * k + j];
      best = fminf(best, fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
    }
    s += (double)best;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < (int)blockDim.x; ++i) t += red[i];
    part[blockIdx.x] = t;
  }
}

__global__ void inertia_final(const double* part, int nb, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < nb; ++i) t += part[i];
    *out = t;
  }
}

constexpr int kInertiaBlocks = 1024;

// ------------------------------------------------------------ workspace ---

struct Ws {
  uint16_t* wbf;                          // bf16 weights
  uint16_t *Y1, *Y2, *Y3, *Y4, *H;        // activations (bf16)
  float *F, *Z6, *P;                      // fp32
  float *dY1, *dY2, *dY3, *dY4, *dF, *dH, *dZ6;
  float* part;                            // split partials
  float* loss;
  uint16_t* input;                        // packed input (predict)
  PackImg* pack;                          // descriptor tables (predict)
  float** dst;
  size_t bytes;
};

static size_t part_floats() {
  size_t mx = 0;
  for (const auto& L : layer_dims())
    mx = std::max(mx, (size_t)kMaxSplits * L.Kp * L.cout);
  return mx;
}

static Ws ws_layout(void* base, int B) {
  Ws w{};
  char* t = static_cast<char*>(base);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    char* p = t ? t + o : nullptr;
    o += align_up(bytes, 256);
    return p;
  };
  const auto d = layer_dims();
  const size_t b = (size_t)B;
  w.wbf = reinterpret_cast<uint16_t*>(take(param_count() * 2));
  w.Y1 = reinterpret_cast<uint16_t*>(take(b * 128 * 128 * 16 * 2));
  w.Y2 = reinterpret_cast<uint16_t*>(take(b * 64 * 64 * 32 * 2));
  w.Y3 = reinterpret_cast<uint16_t*>(take(b * 32 * 32 * 64 * 2));
  w.Y4 = reinterpret_cast<uint16_t*>(take(b * 16 * 16 * 64 * 2));
  w.H = reinterpret_cast<uint16_t*>(take(b * 128 * 2));
  w.F = reinterpret_cast<float*>(take(b * 64 * 4));
  w.Z6 = reinterpret_cast<float*>(take(b * kOut * 4));
  w.P = reinterpret_cast<float*>(take(b * kOut * 4));
  w.dY1 = reinterpret_cast<float*>(take(b * 128 * 128 * 16 * 4));
  w.dY2 = reinterpret_cast<float*>(take(b * 64 * 64 * 32 * 4));
  w.dY3 = reinterpret_cast<float*>(take(b * 32 * 32 * 64 * 4));
  w.dY4 = reinterpret_cast<float*>(take(b * 16 * 16 * 64 * 4));
  w.dF = reinterpret_cast<float*>(take(b * 64 * 4));
  w.dH = reinterpret_cast<float*>(take(b * 128 * 4));
  w.dZ6 = reinterpret_cast<float*>(take(b * kOut * 4));
  w.part = reinterpret_cast<float*>(take(part_floats() * 4));
  w.loss = reinterpret_cast<float*>(take(4));
  w.input = reinterpret_cast<uint16_t*>(take(b * kSide * kSide * 3 * 2));
  w.pack = reinterpret_cast<PackImg*>(take(b * sizeof(PackImg)));
  w.dst = reinterpret_cast<float**>(take(b * sizeof(float*)));
  w.bytes = o;
  (void)d;
  return w;
}

static int grid_for(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256,
                                                     8 * device_cu_count()));
}

// Forward pass over `batch` images of `input` (slots via index).
static void forward(const float* params, const uint16_t* input,
                    const int32_t* index, int B, const Ws& w, hipStream_t s) {
  const auto d = layer_dims();
  const int64_t pc = (int64_t)param_count();
  hipLaunchKernelGGL(cast_bf16, dim3(grid_for(pc)), dim3(256), 0, s, params,
                     w.wbf, pc);
  const uint16_t* X = input;
  uint16_t* outs[4] = {w.Y1, w.Y2, w.Y3, w.Y4};
  for (int l = 0; l < 4; ++l) {
    const LayerDim& L = d[l];
    ConvA a{X, l == 0 ? index : nullptr,
            (int64_t)L.hin * L.hin * L.cin, L.hin, L.cin, L.hout, L.K};
    const int64_t M = (int64_t)B * L.hout * L.hout;
    gemm(M, L.cout, L.Kp, 1, a, WeightB{w.wbf + L.off, L.Kp},
         EpReluBf16{outs[l], L.cout}, s);
    X = outs[l];
  }
  hipLaunchKernelGGL(gap_forward, dim3((B * 64 + 255) / 256), dim3(256), 0, s,
                     w.Y4, 16 * 16, 64, B, w.F);
  gemm(B, 128, d[4].Kp, 1, DenseA<float>{w.F, 64},
       WeightB{w.wbf + d[4].off, d[4].Kp}, EpReluBf16{w.H, 128}, s);
  gemm(B, kOut, d[5].Kp, 1, DenseA<uint16_t>{w.H, 128},
       WeightB{w.wbf + d[5].off, d[5].Kp}, EpF32{w.Z6, kOut}, s);
}

static int splits_for(int64_t R) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kMaxSplits, R / 2048));
}

template <class LA>
static void wgrad(const LayerDim& L, int64_t M, LA at, GradB g, float* grad,
                  float* part, hipStream_t s) {
  const int splits = splits_for(M);
  gemm(L.Kp, L.cout, M, splits, at, g, EpPartial{part, L.Kp, L.cout}, s);
  hipLaunchKernelGGL(reduce_partials, dim3(grid_for((int64_t)L.Kp * L.cout)),
                     dim3(256), 0, s, part, splits, L.Kp, L.cout,
                     grad + L.off);
}

static void backward(const uint16_t* input, const int32_t* index, int B,
                     const Ws& w, float* grads, hipStream_t s) {
  const auto d = layer_dims();
  // fc6 (dZ6 from sigmoid_loss; no ReLU)
  {
    const LayerDim& L = d[5];
    GradB g{w.dZ6, nullptr, kOut, B};
    wgrad(L, B, DenseAT<uint16_t>{{w.H, 128}, B}, g, grads, w.part, s);
    gemm(B, 128, kOut, 1, GradA{g}, DenseGradB{w.wbf + L.off, L.Kp, kOut},
         EpF32{w.dH, 128}, s);
  }
  // fc5 (through its ReLU)
  {
    const LayerDim& L = d[4];
    GradB g{w.dH, w.H, 128, B};
    wgrad(L, B, DenseAT<float>{{w.F, 64}, B}, g, grads, w.part, s);
    gemm(B, 64, 128, 1, GradA{g}, DenseGradB{w.wbf + L.off, L.Kp, 128},
         EpF32{w.dF, 64}, s);
  }
  hipLaunchKernelGGL(gap_backward, dim3(grid_for((int64_t)B * 256 * 64)),
                     dim3(256), 0, s, w.dF, 256, 64, B, w.dY4);
  const uint16_t* ins[4] = {input, w.Y1, w.Y2, w.Y3};
  uint16_t* outs[4] = {w.Y1, w.Y2, w.Y3, w.Y4};
  float* douts[4] = {w.dY1, w.dY2, w.dY3, w.dY4};
  for (int l = 3; l >= 0; --l) {
    const LayerDim& L = d[l];
    const int64_t M = (int64_t)B * L.hout * L.hout;
    GradB g{douts[l], outs[l], L.cout, M};
    ConvA a{ins[l], l == 0 ? index : nullptr, (int64_t)L.hin * L.hin * L.cin,
            L.hin, L.cin, L.hout, L.K};
    wgrad(L, M, ConvAT{a, M}, g, grads, w.part, s);
    if (l > 0) {
      const int64_t Mi = (int64_t)B * L.hin * L.hin;
      gemm(Mi, L.cin, 9 * L.cout, 1, ConvGradA{g, L.hin, L.hout},
           ConvGradB{w.wbf + L.off, L.Kp, L.cin, L.cout},
           EpF32{douts[l - 1], L.cin}, s);
    }
  }
}

}  // namespace ws
}  // namespace splithip

using namespace splithip;
using namespace splithip::ws;

extern "C" {

size_t split_hip_ws_param_count(void) { return param_count(); }

size_t split_hip_ws_workspace(int32_t batch) {
  if (batch < 1) return 0;
  return ws_layout(nullptr, batch).bytes;
}

int split_hip_ws_pack(const split_hip_image_desc* images, int32_t count,
                      uint16_t* input, void* temp, size_t temp_bytes,
                      split_hip_stream_t stream) {
  if (!images || count < 1 || !input || !temp)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "images, input or temp missing");
  if (temp_bytes < (size_t)count * sizeof(PackImg))
    return fail(SPLIT_HIP_ERR_TEMP, "temp holds %zu bytes, pack needs %zu",
                temp_bytes, (size_t)count * sizeof(PackImg));
  std::vector<PackImg> h(count);
  for (int i = 0; i < count; ++i) {
    if (!images[i].points || images[i].width < 1 || images[i].height < 1)
      return fail(SPLIT_HIP_ERR_ARGUMENT, "image %d: no points or empty", i);
    h[i] = PackImg{images[i].points, images[i].width, images[i].height};
  }
  SPLIT_HIP_CHECK(hipMemcpyAsync(temp, h.data(), count * sizeof(PackImg),
                                 hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(pack_kernel, dim3(64, count), dim3(256), 0, stream,
                     static_cast<const PackImg*>(temp), input);
  SPLIT_HIP_CHECK_LAUNCH("ws pack");
  SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
  return SPLIT_HIP_OK;
}

int split_hip_ws_coarse(const split_hip_image_desc* images, int32_t count,
                        float* planes, void* temp, size_t temp_bytes,
                        split_hip_stream_t stream) {
  if (!images || count < 1 || !planes || !temp)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "images, planes or temp missing");
  if (temp_bytes < (size_t)count * sizeof(PackImg))
    return fail(SPLIT_HIP_ERR_TEMP, "temp holds %zu bytes, coarse needs %zu",
                temp_bytes, (size_t)count * sizeof(PackImg));
  std::vector<PackImg> h(count);
  for (int i = 0; i < count; ++i) {
    if (!images[i].points || images[i].width < 1 || images[i].height < 1)
      return fail(SPLIT_HIP_ERR_ARGUMENT, "image %d: no points or empty", i);
    h[i] = PackImg{images[i].points, images[i].width, images[i].height};
  }
  SPLIT_HIP_CHECK(hipMemcpyAsync(temp, h.data(), count * sizeof(PackImg),
                                 hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(coarse_kernel, dim3(64, count), dim3(256), 0, stream,
                     static_cast<const PackImg*>(temp), planes);
  SPLIT_HIP_CHECK_LAUNCH("ws coarse");
  SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
  return SPLIT_HIP_OK;
}

int split_hip_ws_forward(const float* params, const uint16_t* input,
                         const int32_t* index, int32_t batch, void* ws,
                         size_t ws_bytes, float* out,
                         split_hip_stream_t stream) {
  if (!params || !input || !ws || batch < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "params, input, ws or batch missing");
  const Ws w = ws_layout(ws, batch);
  if (ws_bytes < w.bytes)
    return fail(SPLIT_HIP_ERR_TEMP, "ws holds %zu bytes, batch %d needs %zu",
                ws_bytes, batch, w.bytes);
  forward(params, input, index, batch, w, stream);
  hipLaunchKernelGGL(sigmoid_loss, dim3(1), dim3(256), 0, stream, w.Z6,
                     (const float*)nullptr, index, batch, w.P, w.dZ6,
                     (float*)nullptr);
  if (out)
    SPLIT_HIP_CHECK(hipMemcpyAsync(out, w.P, (size_t)batch * kOut * 4,
                                   hipMemcpyDeviceToDevice, stream));
  SPLIT_HIP_CHECK_LAUNCH("ws forward");
  return SPLIT_HIP_OK;
}

int split_hip_ws_backward(const float* params, const uint16_t* input,
                          const float* targets, const int32_t* index,
                          int32_t batch, int32_t loss_kind, void* ws,
                          size_t ws_bytes, float* grads, float* loss,
                          split_hip_stream_t stream) {
  if (!params || !input || !targets || !ws || !grads || batch < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "missing argument to ws backward");
  if (loss_kind != SPLIT_HIP_WS_LOSS_MSE && loss_kind != SPLIT_HIP_WS_LOSS_CHAMFER)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "loss kind %d", loss_kind);
  if (loss_kind == SPLIT_HIP_WS_LOSS_CHAMFER && batch > 256)
    return fail(SPLIT_HIP_ERR_UNSUPPORTED, "chamfer loss: batch %d > 256", batch);
  const Ws w = ws_layout(ws, batch);
  if (ws_bytes < w.bytes)
    return fail(SPLIT_HIP_ERR_TEMP, "ws holds %zu bytes, batch %d needs %zu",
                ws_bytes, batch, w.bytes);
  if (loss_kind == SPLIT_HIP_WS_LOSS_MSE)
    hipLaunchKernelGGL(sigmoid_loss, dim3(1), dim3(256), 0, stream, w.Z6,
                       targets, index, batch, w.P, w.dZ6, w.loss);
  else
    hipLaunchKernelGGL(chamfer_loss, dim3(1), dim3(256), 0, stream, w.Z6,
                       targets, index, batch, w.P, w.dZ6, w.loss);
  backward(input, index, batch, w, grads, stream);
  if (loss)
    SPLIT_HIP_CHECK(hipMemcpyAsync(loss, w.loss, 4, hipMemcpyDeviceToDevice,
                                   stream));
  SPLIT_HIP_CHECK_LAUNCH("ws backward");
  return SPLIT_HIP_OK;
}

int split_hip_ws_adam(float* params, const float* grads, float* m, float* v,
                      int32_t step, float lr, float beta1, float beta2,
                      float eps, split_hip_stream_t stream) {
  if (!params || !grads || !m || !v || step < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "missing argument to ws adam");
  const int64_t n = (int64_t)param_count();
  const float c1 = 1.f - powf(beta1, (float)step);
  const float c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, stream,
                     params, grads, m, v, n, lr, beta1, beta2, eps, c1, c2);
  SPLIT_HIP_CHECK_LAUNCH("ws adam");
  return SPLIT_HIP_OK;
}

int split_hip_ws_predict(const float* params,
                         const split_hip_image_desc* images, int32_t count,
                         void* ws, size_t ws_bytes,
                         split_hip_stream_t stream) {
  if (!params || !images || count < 1 || !ws)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "missing argument to ws predict");
  const Ws w = ws_layout(ws, count);
  if (ws_bytes < w.bytes)
    return fail(SPLIT_HIP_ERR_TEMP, "ws holds %zu bytes, %d images need %zu",
                ws_bytes, count, w.bytes);
  std::vector<PackImg> h(count);
  std::vector<float*> dst(count);
  for (int i = 0; i < count; ++i) {
    if (images[i].k != SPLIT_HIP_WS_K || !images[i].centroids ||
        !images[i].points || images[i].width < 1 || images[i].height < 1)
      return fail(SPLIT_HIP_ERR_ARGUMENT,
                  "image %d: needs points, a size and k = %d centroids", i,
                  SPLIT_HIP_WS_K);
    h[i] = PackImg{images[i].points, images[i].width, images[i].height};
    dst[i] = images[i].centroids;
  }
  SPLIT_HIP_CHECK(hipMemcpyAsync(w.pack, h.data(), count * sizeof(PackImg),
                                 hipMemcpyHostToDevice, stream));
  SPLIT_HIP_CHECK(hipMemcpyAsync(w.dst, dst.data(), count * sizeof(float*),
                                 hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(pack_kernel, dim3(64, count), dim3(256), 0, stream,
                     w.pack, w.input);
  forward(params, w.input, nullptr, count, w, stream);
  hipLaunchKernelGGL(sigmoid_loss, dim3(1), dim3(256), 0, stream, w.Z6,
                     (const float*)nullptr, (const int32_t*)nullptr, count,
                     w.P, w.dZ6, (float*)nullptr);
  hipLaunchKernelGGL(scatter_centroids, dim3((count * kOut + 255) / 256),
                     dim3(256), 0, stream, w.P, w.dst, count);
  SPLIT_HIP_CHECK_LAUNCH("ws predict");
  SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
  return SPLIT_HIP_OK;
}

int split_hip_ws_targets(const float* centroids, int32_t count, float* targets,
                         split_hip_stream_t stream) {
  if (!centroids || !targets || count < 1)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "missing argument to ws targets");
  hipLaunchKernelGGL(targets_kernel, dim3((count + 63) / 64), dim3(64), 0,
                     stream, centroids, count, targets);
  SPLIT_HIP_CHECK_LAUNCH("ws targets");
  return SPLIT_HIP_OK;
}

size_t split_hip_inertia_workspace(void) {
  return align_up(kInertiaBlocks * sizeof(double), 256) + 256;
}

int split_hip_inertia(const float* points, int64_t n, const float* centroids,
                      int32_t k, void* temp, size_t temp_bytes,
                      double* inertia, split_hip_stream_t stream) {
  if (!points || !centroids || !temp || !inertia || n < 1 || k < 1 ||
      k > SPLIT_HIP_MAX_K)
    return fail(SPLIT_HIP_ERR_ARGUMENT, "missing argument to inertia");
  if (temp_bytes < split_hip_inertia_workspace())
    return fail(SPLIT_HIP_ERR_TEMP, "temp holds %zu bytes, inertia needs %zu",
                temp_bytes, split_hip_inertia_workspace());
  double* part = static_cast<double*>(temp);
  double* res = reinterpret_cast<double*>(
      static_cast<char*>(temp) + align_up(kInertiaBlocks * sizeof(double), 256));
  hipLaunchKernelGGL(inertia_partial, dim3(kInertiaBlocks), dim3(256), 0,
                     stream, points, n, centroids, k, part);
  hipLaunchKernelGGL(inertia_final, dim3(1), dim3(64), 0, stream, part,
                     kInertiaBlocks, res);
  SPLIT_HIP_CHECK_LAUNCH("inertia");
  SPLIT_HIP_CHECK(hipMemcpyAsync(inertia, res, sizeof(double),
                                 hipMemcpyDeviceToHost, stream));
  SPLIT_HIP_CHECK(hipStreamSynchronize(stream));
  return SPLIT_HIP_OK;
}

}  // extern "C"
