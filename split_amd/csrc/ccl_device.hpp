// Device bodies of the connected-component passes (64x64 tiles), shared by
// the single-image kernels (ccl.hip) and the batched ones (batch.hip).
#pragma once

#include <limits.h>

#include "common.hpp"

namespace splithip {

// Accesses to the global forest (the output array).  Loads are plain:
// every value a forest word ever holds is an ancestor of its pixel (links
// only point to larger indices and only ever move up the tree), so a load
// that returns an older value still leads to the same root.  Stores are
// plain where one thread per kernel writes the word (the tile pass, the
// final pass).  Union links, where two threads may store to one root's
// word, are device atomics (g_raise, atomicMax): performed at the coherence
// point and monotone, so the larger value, the one nearer the root, wins
// whatever the order.  The root-jump pass writes each linked root's word
// from that root's thread only, also with g_raise, for the visibility of
// its progress to the other XCDs within the pass.  (An earlier walk whose
// path-halving store into a root could land after that root's own final
// store left, with plain stores, tile roots of the 2048^2 spiral holding
// their chain grandparent.)
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

// ------------------------------------------------- 64x64-tile passes ---
//
// Tile side S of the CC passes (64).  Thread t of a block handles tile
// pixels t + (S*S/4) j, j = 0..3: with S = 64 a wave holds one tile row
// per j, with S = 32 two (lanes 0-31, 32-63).  More pixels per thread (and
// more blocks per CU) was slower: the 4096^2 k = 32 labels' tile pass took
// 154 / 162 / 265 us at 4 / 8 / 16 pixels per thread.
template <int S>
struct CcTile {
  static constexpr int kPx = S * S;
  static constexpr int kPer = 4;  // pixels per thread
  static constexpr int kThreads = kPx / kPer;
  static constexpr int kRowsPerPass = kThreads / S;
};

// Inputs, outputs and scratch of one connected-components run.
struct CcArgs {
  const int32_t* labels;
  int32_t* out;   // the forest, then the result
  int W, H, ntx;
  unsigned* linked;      // bit p: root p was linked under another (N bits)
  unsigned* local_root;  // bit p: p is its tile's local root (compression)
  // per tile t, rows 0..S-1: labels and local roots of the left (x0) and
  // right (x0 + S - 1) columns
  int32_t* llab;
  int32_t* rlab;
  int32_t* lroot;
  int32_t* rroot;
};

// ORs `bits` (the first `count` <= 64 bits used) into the bitmap at bit
// offset p0 (any alignment); lanes k0, k0 + 1, k0 + 2 write one word each.
__device__ __forceinline__ void or_bits(unsigned* bm, int64_t p0,
                                        uint64_t bits, int lane, int k0) {
  const int k = lane - k0;
  if (bits == 0 || k < 0 || k > 2) return;
  const int s = (int)(p0 & 31);
  // word w: bits 0..31-s, word w+1: the next 32, word w+2: the last s
  const unsigned v = k == 0   ? (unsigned)(bits << s)
                     : k == 1 ? (unsigned)(bits >> (32 - s))
                              : (unsigned)((bits >> (32 - s)) >> 32);
  if (v) atomicOr(bm + (p0 >> 5) + k, v);
}

// The tile pass for tile (tx, ty), by a block of CcTile<S>::kThreads
// threads; lab and par are the block's LDS arrays of S*S ints.
template <int CONN, int S>
__device__ __forceinline__ void cc_tile_body(const CcArgs& a, int tx, int ty,
                                             int* lab, int* par) {
  using G = CcTile<S>;
  const int x0 = tx * S, y0 = ty * S;
  const int tw = min(S, a.W - x0), th = min(S, a.H - y0);
  const int lane = threadIdx.x & (kWave - 1);
  const int col = threadIdx.x & (S - 1);
  const int rowbase = lane - col;  // first lane of this lane's row
  const int64_t T = (int64_t)ty * a.ntx + tx;
  int l[G::kPer], end[G::kPer];
#pragma unroll
  for (int j = 0; j < G::kPer; ++j) {
    const int r = threadIdx.x / S + G::kRowsPerPass * j;
    const bool in = col < tw && r < th;
    l[j] = in ? a.labels[(int64_t)(y0 + r) * a.W + x0 + col] : INT_MIN;
    lab[r * S + col] = l[j];
  }
  // each pixel's parent: the last pixel of its row run
#pragma unroll
  for (int j = 0; j < G::kPer; ++j) {
    const int r = threadIdx.x / S + G::kRowsPerPass * j;
    const bool in = col < tw && r < th;
    const int right = __shfl_down(l[j], 1);
    const uint64_t m = ballot(in && col + 1 < tw && right == l[j]);
    // bit S-1 of the row's bits is clear, so the shifted complement is
    // nonzero below it
    end[j] = col + (int)__builtin_ctzll(~(m >> lane));
    par[r * S + col] = r * S + (in ? end[j] : col);
  }
  __syncthreads();
  // unions between runs of consecutive rows, once per run overlap
#pragma unroll
  for (int j = 0; j < G::kPer; ++j) {
    const int r = threadIdx.x / S + G::kRowsPerPass * j;
    const int left = __shfl_up(l[j], 1);
    if (r == 0 || r >= th || col >= tw) continue;
    const int li = r * S + col, c = l[j], up = li - S;
    const bool left_same = col > 0 && left == c;
    const int lu = lab[up];
    const int lul = col > 0 ? lab[up - 1] : INT_MIN;
    const bool up_same = lu == c;
    if (up_same && !(left_same && lul == c)) lds_union(par, li, up);
    if (CONN == 8) {
      if (col > 0 && !left_same && !up_same && lul == c)
        lds_union(par, li, up - 1);
      if (col + 1 < tw && !up_same && lab[up + 1] == c)
        lds_union(par, li, up + 1);
    }
  }
  __syncthreads();
  // roots: one find per run (at its end), shared along the run
#pragma unroll
  for (int j = 0; j < G::kPer; ++j) {
    const int r = threadIdx.x / S + G::kRowsPerPass * j;
    const bool in = col < tw && r < th;
    const int li = r * S + col;
    const int found = (in && col == end[j]) ? lds_find(par, li) : 0;
    const int rl = __shfl(found, rowbase + (in ? end[j] : col));
    const int g = (y0 + rl / S) * a.W + x0 + (rl & (S - 1));
    const int64_t p = (int64_t)(y0 + r) * a.W + x0 + col;
    if (in) g_store(a.out, p, g);
    if (a.local_root) {
      const uint64_t roots = ballot(in && rl == li);
      if (r < th)
        or_bits(a.local_root, (int64_t)(y0 + r) * a.W + x0,
                S == 64 ? roots : (roots >> rowbase) & 0xFFFFFFFFull, lane,
                rowbase);
    }
    if (in && col == 0) {
      a.llab[T * S + r] = l[j];
      a.lroot[T * S + r] = g;
    }
    if (in && col == S - 1) {
      a.rlab[T * S + r] = l[j];
      a.rroot[T * S + r] = g;
    }
  }
}

// One block of three groups of S threads per tile: group 0 the top row (up,
// up-left, up-right neighbours in the rows above), group 1 the left column
// (left and up-left neighbours in the left tile, from the edge arrays),
// group 2 the right column's up-right neighbour in the right tile
// (8-connectivity).  A union is skipped where the neighbouring seam pixel
// already joins the same two runs, which leaves about one union per pair of
// touching runs.
template <int CONN, int S>
__device__ __forceinline__ void cc_merge_body(const CcArgs& a, int tx, int ty) {
  const int x0 = tx * S, y0 = ty * S;
  const int tw = min(S, a.W - x0), th = min(S, a.H - y0);
  const int t = threadIdx.x % S;
  const int grp = threadIdx.x / S;
  const int64_t T = (int64_t)ty * a.ntx + tx;
  const int W = a.W;
  if (grp == 0) {
    if (y0 == 0 || t >= tw) return;
    const int x = x0 + t;
    const int p = y0 * W + x, u = p - W;
    const int c = a.labels[p];
    const int lu = a.labels[u];
    const bool left_same = t > 0 && a.labels[p - 1] == c;
    const int lul = x > 0 ? a.labels[u - 1] : INT_MIN;
    if (lu == c && !(left_same && lul == c)) g_union(a.out, p, u, a.linked);
    if (CONN == 8) {
      if (x > 0 && lu != c && !left_same && lul == c)
        g_union(a.out, p, u - 1, a.linked);
      if (x + 1 < W && lu != c && a.labels[u + 1] == c)
        g_union(a.out, p, u + 1, a.linked);
    }
  } else if (grp == 1) {
    if (x0 == 0 || t >= th) return;
    const int64_t e = T * S + t, el = (T - 1) * S + t;
    const int c = a.llab[e];
    const int lq = a.rlab[el];
    const bool up_in = t > 0;
    const bool up_same = up_in && a.llab[e - 1] == c;
    const int ul = up_in ? a.rlab[el - 1] : INT_MIN;
    if (lq == c && !(up_same && ul == c))
      g_union(a.out, a.lroot[e], a.rroot[el], a.linked);
    if (CONN == 8 && up_in && lq != c && !up_same && ul == c)
      g_union(a.out, a.lroot[e], a.rroot[el - 1], a.linked);
  } else if (CONN == 8) {
    if (tw < S || x0 + S >= W || t == 0 || t >= th) return;
    const int64_t e = T * S + t, er = (T + 1) * S + t;
    const int c = a.rlab[e];
    if (a.llab[er - 1] == c && a.rlab[e - 1] != c && a.llab[er] != c)
      g_union(a.out, a.rroot[e], a.lroot[er - 1], a.linked);
  }
}

// Bitmap words per chunk of the rank prefix (chunk sums, then word
// prefixes within a chunk).
constexpr int kCcScanChunk = 1024;

// Dense rank of a root: chunk sum (csum may be null: prefix is then a
// prefix over all words) + word prefix + the lower bits of its word.
__device__ __forceinline__ int32_t root_rank(int root, const unsigned* rb,
                                             const unsigned* prefix,
                                             const unsigned* csum) {
  const uint32_t w = (uint32_t)root >> 5;
  const unsigned below = rb[w] & ((1u << (root & 31)) - 1u);
  return (int32_t)((csum ? csum[w / kCcScanChunk] : 0u) + prefix[w] +
                   __popc(below));
}

// Per tile: every pixel's root, from the tile's forest words staged in LDS.
// After the jump pass a pixel's word holds itself (an unlinked local root:
// a root), the root (a linked local root, jumped), or its local root in the
// same tile (any other pixel; the tile pass wrote it and nothing rewrote it),
// whose own word is then the root.  So the root is the word of the pixel's
// word when that lies in the tile, else the pixel's word.  Writes the root
// index (rb == null) or its dense rank among the roots.
template <int S>
__device__ __forceinline__ void cc_final_body(const CcArgs& a, int tx, int ty,
                                              const unsigned* rb,
                                              const unsigned* prefix,
                                              const unsigned* csum, int* sv) {
  using G = CcTile<S>;
  const int x0 = tx * S, y0 = ty * S;
  const int tw = min(S, a.W - x0), th = min(S, a.H - y0);
  const int col = threadIdx.x & (S - 1);
  const double inv_w = 1.0 / (double)a.W;
  int v[G::kPer];
#pragma unroll
  for (int j = 0; j < G::kPer; ++j) {
    const int r = threadIdx.x / S + G::kRowsPerPass * j;
    const bool in = col < tw && r < th;
    v[j] = in ? g_load(a.out, (int64_t)(y0 + r) * a.W + x0 + col) : 0;
    sv[r * S + col] = v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < G::kPer; ++j) {
    const int r = threadIdx.x / S + G::kRowsPerPass * j;
    if (col >= tw || r >= th) continue;
    const int64_t p = (int64_t)(y0 + r) * a.W + x0 + col;
    const int v1 = v[j];
    // row and column of v1 (double division, corrected: exact for v1 < 2^31)
    int vy = (int)((double)v1 * inv_w);
    if ((int64_t)vy * a.W > v1) --vy;
    else if ((int64_t)(vy + 1) * a.W <= v1) ++vy;
    const int vx = v1 - vy * a.W;
    const bool in_tile = vy >= y0 && vy < y0 + th && vx >= x0 && vx < x0 + tw;
    const int root = in_tile ? sv[(vy - y0) * S + (vx - x0)] : v1;
    g_store(a.out, p, rb ? root_rank(root, rb, prefix, csum) : root);
  }
}

// The root-jump pass for the linked roots of bitmap word w (pixels
// 32w..32w+31 of the forest `out`): each points at its tree's root by
// pointer jumping — its word moves to its parent's current word and is
// published after every step, so the roots below it advance through it by
// the distance it has covered.  Roots in decreasing order (the callers also
// take words in decreasing order: ancestors, which have larger indices,
// start first and descendants mostly find them done).
__device__ __forceinline__ void cc_jump_word(int32_t* out,
                                             const unsigned* linked,
                                             int64_t w) {
  unsigned m = linked[w];
  while (m) {
    const int b = 31 - __builtin_clz(m);
    m &= ~(1u << b);
    const int a = (int)(w * 32 + b);
    int x = g_load(out, a);
    for (;;) {
      const int px = g_load(out, x);
      if (px == x) break;
      x = px;
      g_raise(out, a, x);
    }
  }
}

}  // namespace splithip
