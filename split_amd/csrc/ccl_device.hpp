// Device bodies of the connected-component passes, shared by the
// single-image and the batched kernels.
#pragma once

#include "common.hpp"

namespace splithip {

constexpr int kCcTile = 32;
constexpr int kCcTilePx = kCcTile * kCcTile;
constexpr int kCcThreads = 256;

// Accesses to the global forest (the output array).  Loads are plain:
// every value a forest word ever holds is an ancestor of its pixel (links
// only point to larger indices and only ever move up the tree), so a load
// that returns an older value still leads to the same root.  Stores are
// plain where one thread per kernel writes the word (the tile pass, the
// final pass, flatten: a pixel's own entry).  Where two threads may store
// to one word — union links, and the batched path's walk (ccl_jump_one),
// whose path-halving store into a root can race with that root's own final
// store — the stores are device atomics (g_raise), performed at the
// coherence point and monotone, so the larger value, the one nearer the
// root, is what every later kernel reads.  With plain racing stores the
// 2048^2 spiral ended with tile roots holding the halving store's value
// after later kernels had written their root.
__device__ __forceinline__ int g_load(const int32_t* par, int64_t i) {
  return par[i];
}
__device__ __forceinline__ void g_store(int32_t* par, int64_t i, int v) {
  par[i] = v;
}
__device__ __forceinline__ void g_raise(int32_t* par, int64_t i, int v) {
  atomicMax(par + i, v);
}

// ---------------------------------------------------------- LDS forest ---
//
// Parents only ever grow (a link puts the smaller root under the larger),
// so a root is the largest index of its tree.  find() halves paths with
// plain stores: a non-root's parent is replaced by its grandparent, which is
// still an ancestor, and unions only ever write roots (atomicMax, retried
// when the root moved), so both kinds of write are safe together.

// The forest lives in LDS; callers pass the __shared__ array, which these
// helpers address as LDS explicitly (a generic pointer would compile to
// flat instructions, which also wait on the vector-memory counter).
#if defined(__HIP_DEVICE_COMPILE__)
typedef volatile __attribute__((address_space(3))) int lds_int;
#else
typedef volatile int lds_int;
#endif

__device__ __forceinline__ int lds_find(int* par_generic, int x) {
  lds_int* par = (lds_int*)par_generic;
  for (;;) {
    const int p = par[x];
    if (p == x) return x;
    const int gp = par[p];
    if (gp == p) return p;
    par[x] = gp;
    x = gp;
  }
}

__device__ __forceinline__ void lds_union(int* par, int a, int b) {
  for (;;) {
    a = lds_find(par, a);
    b = lds_find(par, b);
    if (a == b) return;
    if (a > b) { const int t = a; a = b; b = t; }
    const int old = __hip_atomic_fetch_max((__attribute__((address_space(3))) int*)par + a, b,
                                           __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old == a) return;
    a = old;
  }
}

// One 256-thread block per 32x32 tile.  Thread t handles tile pixels
// t + 256*j: a wave covers two tile rows per j, one pixel per lane.
//  1. Horizontal runs: each pixel's parent is the last pixel of its run in
//     the row (ballot of "same label as right neighbour", then ctz).
//  2. Vertical (and, for 8-connectivity, diagonal) unions between runs,
//     skipping a union when the pixel to the left already made it (same run
//     below, same run above).
//  3. Each pixel writes the global index of its root.
template <int CONN>
__device__ __forceinline__ void ccl_tile_body(const int32_t* __restrict__ labels,
                                              int W, int H, int32_t* out,
                                              int tx, int ty, int* lab,
                                              int* par) {
  const int x0 = tx * kCcTile, y0 = ty * kCcTile;
  const int tw = min(kCcTile, W - x0);  // tile width inside the image
  const int th = min(kCcTile, H - y0);
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int j = 0; j < kCcTilePx / kCcThreads; ++j) {
    const int li = threadIdx.x + kCcThreads * j;
    const int lx = li & (kCcTile - 1), ly = li / kCcTile;
    const bool in = lx < tw && ly < th;
    lab[li] = in ? labels[(int64_t)(y0 + ly) * W + x0 + lx] : 0;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kCcTilePx / kCcThreads; ++j) {
    const int li = threadIdx.x + kCcThreads * j;
    const int lx = li & (kCcTile - 1), ly = li / kCcTile;
    const bool in = lx < tw && ly < th;
    const bool right = in && lx + 1 < tw && lab[li + 1] == lab[li];
    const uint64_t m = ballot(right);
    const uint32_t row = lane < 32 ? (uint32_t)m : (uint32_t)(m >> 32);
    // bit tw-1 (and above) is clear, so the shifted complement is nonzero
    const int end = lx + __builtin_ctz(~(row >> lx));
    par[li] = in ? li - lx + end : li;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kCcTilePx / kCcThreads; ++j) {
    const int li = threadIdx.x + kCcThreads * j;
    const int lx = li & (kCcTile - 1), ly = li / kCcTile;
    if (lx >= tw || ly >= th || ly == 0) continue;
    const int l = lab[li];
    const int up = li - kCcTile;
    const bool left_same = lx > 0 && lab[li - 1] == l;
    const bool up_same = lab[up] == l;
    if (up_same && !(left_same && lab[up - 1] == l)) lds_union(par, li, up);
    if (CONN == 8) {
      if (lx > 0 && !left_same && !up_same && lab[up - 1] == l)
        lds_union(par, li, up - 1);
      if (lx + 1 < tw && !up_same && lab[up + 1] == l)
        lds_union(par, li, up + 1);
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kCcTilePx / kCcThreads; ++j) {
    const int li = threadIdx.x + kCcThreads * j;
    const int lx = li & (kCcTile - 1), ly = li / kCcTile;
    if (lx >= tw || ly >= th) continue;
    const int r = lds_find(par, li);
    g_store(out, (int64_t)(y0 + ly) * W + x0 + lx,
            (y0 + r / kCcTile) * W + x0 + (r & (kCcTile - 1)));
  }
}

// ------------------------------------------------------- global forest ---

// Finds do not shorten the paths they walk: halving them (atomicMax into
// the linked roots' words) made the 8192^2 spiral image's merge 704 us
// instead of 362 for a jump pass 125 us shorter.
__device__ __forceinline__ int g_find(const int* par, int i) {
  for (;;) {
    const int p = g_load(par, i);
    if (p == i) return i;
    i = p;
  }
}

// Joins the trees of a and b: the smaller root goes under the larger.  A
// find may stop at a word's older value (another XCD's link not yet seen);
// atomicMax returns the current value, and a root that turns out to be
// linked already is retried from its current parent.  With `linked`, the
// root that a successful link puts under another gets its bit set there
// (the one thread whose atomicMax saw the root unlinked).
__device__ __forceinline__ void g_union(int* par, int a, int b,
                                        unsigned* linked = nullptr) {
  for (;;) {
    a = g_find(par, a);
    b = g_find(par, b);
    if (a == b) return;
    if (a > b) { const int t = a; a = b; b = t; }
    const int old = atomicMax(par + a, b);
    if (old == a) {
      if (linked) atomicOr(linked + (a >> 5), 1u << (a & 31));
      return;
    }
    a = old;
  }
}

// One block per tile; thread t examines one pixel on the tile's top row,
// left column or (8-connectivity) right column and unions it with each
// equal-label neighbour that lies in another tile.  A union is skipped when
// a neighbouring edge pixel already joins the same two runs (equal labels
// along the edge on both sides), which leaves about one union per pair of
// touching components instead of one per edge pixel.
template <int CONN>
__device__ __forceinline__ void ccl_merge_body(const int32_t* __restrict__ labels,
                                               int W, int H, int32_t* out,
                                               int tx, int ty, int t) {
  const int x0 = tx * kCcTile, y0 = ty * kCcTile;
  int gx, gy, side;
  if (t < kCcTile) {             // top row: up, up-left, up-right
    gx = x0 + t; gy = y0; side = 0;
  } else if (t < 2 * kCcTile) {  // left column: left, up-left
    gx = x0; gy = y0 + t - kCcTile; side = 1;
  } else {                     // right column: up-right into the next tile
    if (CONN != 8) return;
    gx = x0 + kCcTile - 1; gy = y0 + t - 2 * kCcTile; side = 2;
  }
  if (gx >= W || gy >= H) return;
  const int p = gy * W + gx;
  const int l = labels[p];
  if (side == 0) {
    if (gy == 0) return;
    const int u = p - W;
    const bool u_same = labels[u] == l;
    const bool left_in = gx > x0;
    const bool left_same = left_in && labels[p - 1] == l;
    if (u_same && !(left_same && labels[u - 1] == l)) g_union(out, p, u);
    if (CONN == 8) {
      if (gx > 0 && !u_same && !left_same && labels[u - 1] == l)
        g_union(out, p, u - 1);
      if (gx + 1 < W && !u_same && labels[u + 1] == l) g_union(out, p, u + 1);
    }
  } else if (side == 1) {
    if (gx == 0) return;
    const bool up_in = gy > y0;
    const bool lf_same = labels[p - 1] == l;
    const bool up_same = up_in && labels[p - W] == l;
    if (lf_same && !(up_same && labels[p - W - 1] == l)) g_union(out, p, p - 1);
    if (CONN == 8 && up_in && !lf_same && !up_same && labels[p - W - 1] == l)
      g_union(out, p, p - W - 1);
  } else {
    if (gy == y0 || gx + 1 >= W) return;
    const int ur = p - W + 1;
    if (labels[ur] == l && labels[p - W] != l && labels[p + 1] != l)
      g_union(out, p, ur);
  }
}


// Shortens the global forest before the flatten pass.  After ccl_tile and
// ccl_merge a pixel's parent is its tile's local root (inside the tile)
// unless the pixel is such a root, whose parent chain crosses tiles and can
// be as long as the component is winding.  Pixels whose parent lies in
// another tile, and roots, walk their chain with path halving (every write
// replaces a parent by an ancestor, so the forest stays valid under
// concurrent walks) and then point straight at the root they found.
__device__ __forceinline__ void ccl_jump_one(int32_t* out, int64_t p, int W) {
  const int q = g_load(out, p);
  if (q == p) return;
  const int py = (int)(p / W), px = (int)(p - (int64_t)py * W);
  const int qy = q / W, qx = q - qy * W;
  if ((py / kCcTile) == (qy / kCcTile) && (px / kCcTile) == (qx / kCcTile))
    return;  // an in-tile parent: its local root does the walk
  int x = q;
  for (;;) {
    const int px2 = g_load(out, x);
    if (px2 == x) break;
    const int gp = g_load(out, px2);
    if (gp == px2) {
      x = px2;
      break;
    }
    g_raise(out, x, gp);
    x = gp;
  }
  g_raise(out, p, x);
}

// Follows the chain of pixel p to its root and writes it back.
__device__ __forceinline__ int ccl_flatten_one(int32_t* out, int64_t p) {
  const int r0 = g_load(out, p);
  int r = r0;
  for (;;) {
    const int q = g_load(out, r);
    if (q == r) break;
    r = q;
  }
  if (r != r0) g_store(out, p, r);
  return r;
}

// Check of the flattened forest at pixel p: p's value must be a root, and
// every equal-label neighbour to the right and below (and on the two lower
// diagonals with 8-connectivity) must carry the same value.  Together these
// mean each component maps to one root, the largest index of its tree.  A
// neighbour pair that disagrees is united (the roots are joined as in the
// merge pass); any failure sets *flag, after which a second flatten pass
// runs.
template <int CONN>
__device__ __forceinline__ void ccl_check_one(const int32_t* __restrict__ labels,
                                              int W, int H, int32_t* out,
                                              int64_t p, unsigned* flag) {
  const int y = (int)(p / W), x = (int)(p - (int64_t)y * W);
  const int l = labels[p];
  const int r = g_load(out, p);
  bool bad = g_load(out, r) != r;
  auto pair = [&](int64_t q) {
    if (labels[q] != l) return;
    const int rq = g_load(out, q);
    if (rq != r) {
      g_union(out, r, rq);
      bad = true;
    }
  };
  if (x + 1 < W) pair(p + 1);
  if (y + 1 < H) {
    pair(p + W);
    if (CONN == 8) {
      if (x > 0) pair(p + W - 1);
      if (x + 1 < W) pair(p + W + 1);
    }
  }
  if (bad) atomicOr(flag, 1u);
}

}  // namespace splithip
