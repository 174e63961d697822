/*
 * CPU oracle of the segmentation path — TEST INFRASTRUCTURE ONLY (see
 * oracle.h).  Plain C11, single-threaded, no fast-math.
 *
 * Parity pins (tests/test_oracle.py): every function here is checked against
 * closed forms, brute force on tiny inputs or an independent implementation
 * (numpy / pure-Python BFS) — see the test file for the list.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads the OpenMP build's parallel loops use (1 in the plain build). */
int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------ FP64 ---- */

void oracle_label_f64(const float* pts, int64_t n, const float* cent, int k,
                      int32_t* labels, double* gap) {
  for (int64_t i = 0; i < n; ++i) {
    const double px = pts[i], py = pts[n + i], pz = pts[2 * n + i];
    double best = INFINITY, second = INFINITY;
    int bj = 0;
    for (int j = 0; j < k; ++j) {
      const double dx = px - (double)cent[j];
      const double dy = py - (double)cent[k + j];
      const double dz = pz - (double)cent[2 * k + j];
      const double d = dx * dx + dy * dy + dz * dz;
      if (d < best) {
        second = best;
        best = d;
        bj = j;
      } else if (d < second) {
        second = d;
      }
    }
    labels[i] = bj;
    if (gap) {
      if (k == 1)
        gap[i] = 1.0;
      else
        gap[i] = second > 0.0 ? (second - best) / second : 0.0;
    }
  }
}

void oracle_centroids_f64(const int32_t* labels, const float* pts, int64_t n,
                          int k, double* out, int64_t* counts) {
  double* s = (double*)calloc((size_t)3 * k, sizeof(double));
  int64_t* c = (int64_t*)calloc((size_t)k, sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) {
    const int l = labels[i];
    if (l < 0 || l >= k) continue;
    s[l] += pts[i];
    s[k + l] += pts[n + i];
    s[2 * k + l] += pts[2 * n + i];
    c[l] += 1;
  }
  for (int j = 0; j < k; ++j)
    for (int ch = 0; ch < 3; ++ch)
      out[ch * k + j] = c[j] ? s[ch * k + j] / (double)c[j] : 0.0;
  if (counts) memcpy(counts, c, (size_t)k * sizeof(int64_t));
  free(s);
  free(c);
}

static void label_f64_dcent(const float* pts, int64_t n, const double* cent,
                            int k, int32_t* labels) {
  for (int64_t i = 0; i < n; ++i) {
    const double px = pts[i], py = pts[n + i], pz = pts[2 * n + i];
    double best = INFINITY;
    int bj = 0;
    for (int j = 0; j < k; ++j) {
      const double dx = px - cent[j], dy = py - cent[k + j],
                   dz = pz - cent[2 * k + j];
      const double d = dx * dx + dy * dy + dz * dz;
      if (d < best) {
        best = d;
        bj = j;
      }
    }
    labels[i] = bj;
  }
}

void oracle_lloyd_f64(const float* pts, int64_t n, const float* init, int k,
                      int iters, double* cent_out, int32_t* labels_out) {
  for (int j = 0; j < 3 * k; ++j) cent_out[j] = init[j];
  for (int it = 0; it < iters; ++it) {
    label_f64_dcent(pts, n, cent_out, k, labels_out);
    oracle_centroids_f64(labels_out, pts, n, k, cent_out, NULL);
  }
}

double oracle_inertia(const float* pts, int64_t n, const float* cent, int k,
                      const int32_t* labels) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const int l = labels[i];
    if (l < 0 || l >= k) continue;
    const double dx = (double)pts[i] - cent[l];
    const double dy = (double)pts[n + i] - cent[k + l];
    const double dz = (double)pts[2 * n + i] - cent[2 * k + l];
    s += dx * dx + dy * dy + dz * dz;
  }
  return s;
}

/* Union-find over pixel indices, linking the smaller root under the larger
 * so that a root is the largest index of its tree. */
static int64_t uf_find(int64_t* par, int64_t i) {
  int64_t r = i;
  while (par[r] != r) r = par[r];
  while (par[i] != r) {
    const int64_t nx = par[i];
    par[i] = r;
    i = nx;
  }
  return r;
}

static void uf_union(int64_t* par, int64_t a, int64_t b) {
  a = uf_find(par, a);
  b = uf_find(par, b);
  if (a == b) return;
  if (a < b)
    par[a] = b;
  else
    par[b] = a;
}

int64_t oracle_cc_uf(const int32_t* labels, int W, int H, int conn,
                     int32_t* out) {
  const int64_t n = (int64_t)W * H;
  int64_t* par = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) par[i] = i;
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      const int64_t p = (int64_t)y * W + x;
      const int32_t l = labels[p];
      if (x > 0 && labels[p - 1] == l) uf_union(par, p, p - 1);
      if (y > 0) {
        if (labels[p - W] == l) uf_union(par, p, p - W);
        if (conn == 8) {
          if (x > 0 && labels[p - W - 1] == l) uf_union(par, p, p - W - 1);
          if (x + 1 < W && labels[p - W + 1] == l) uf_union(par, p, p - W + 1);
        }
      }
    }
  }
  int64_t roots = 0;
  for (int64_t p = 0; p < n; ++p) {
    const int64_t r = uf_find(par, p);
    out[p] = (int32_t)r;
    roots += (r == p);
  }
  free(par);
  return roots;
}

static int cmp_i32(const void* a, const void* b) {
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* index of the first element >= v in sorted u[0..m) */
static int64_t lower_bound_i32(const int32_t* u, int64_t m, int32_t v) {
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (u[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

int64_t oracle_compress(int32_t* labels, int64_t n) {
  int32_t* u = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
  memcpy(u, labels, (size_t)n * sizeof(int32_t));
  qsort(u, (size_t)n, sizeof(int32_t), cmp_i32);
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i)
    if (m == 0 || u[m - 1] != u[i]) u[m++] = u[i];
  for (int64_t i = 0; i < n; ++i)
    labels[i] = (int32_t)lower_bound_i32(u, m, labels[i]);
  free(u);
  return m;
}

static uint64_t splitmix64_next(uint64_t* s) {
  *s += 0x9E3779B97F4A7C15ull;
  uint64_t z = *s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void oracle_sample_indices(uint64_t seed, int64_t n, int k, int64_t* idx) {
  uint64_t s = seed;
  int j = 0;
  while (j < k) {
    const uint64_t z = splitmix64_next(&s);
    const int64_t cand =
        (int64_t)(((unsigned __int128)z * (unsigned __int128)(uint64_t)n) >> 64);
    int dup = 0;
    for (int i = 0; i < j; ++i)
      if (idx[i] == cand) dup = 1;
    if (!dup) idx[j++] = cand;
  }
}

/* ----------------------------------------------- reference restatement -- */

static float ref_dist(const float* pts, int64_t n, int64_t i, const float* cent,
                      int k, int j) {
  const float dx = pts[i] - cent[j];
  const float dy = pts[n + i] - cent[k + j];
  const float dz = pts[2 * n + i] - cent[2 * k + j];
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

void ref_label_points(const float* cent, int k, const float* pts, int64_t n,
                      int32_t* labels) {
  /* pixels are independent: the OpenMP build (timing only) splits them */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    float best = INFINITY;
    int bj = 0;
    for (int j = 0; j < k; ++j) {
      const float d = ref_dist(pts, n, i, cent, k, j);
      if (d < best || (d == best && j < bj)) {
        best = d;
        bj = j;
      }
    }
    labels[i] = bj;
  }
}

typedef struct {
  int32_t label;
  int64_t index;
} keyed_t;

static int cmp_keyed(const void* a, const void* b) {
  const keyed_t* x = (const keyed_t*)a;
  const keyed_t* y = (const keyed_t*)b;
  if (x->label != y->label) return (x->label > y->label) - (x->label < y->label);
  return (x->index > y->index) - (x->index < y->index); /* stable */
}

void ref_calculate_centroids(const int32_t* labels, const float* pts,
                             int64_t n, float* cent, int k) {
  keyed_t* s = (keyed_t*)malloc((size_t)n * sizeof(keyed_t));
  /* sort_by_key of (label, index) pairs (centroids.cu:48): labels in
   * [0, k) take a stable counting sort (the radix sort's order: by label,
   * then index), anything else the comparison sort */
  int in_range = 1;
  for (int64_t i = 0; i < n; ++i)
    if (labels[i] < 0 || labels[i] >= k) in_range = 0;
  if (in_range) {
    int64_t* start = (int64_t*)calloc((size_t)k + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) ++start[labels[i] + 1];
    for (int j = 0; j < k; ++j) start[j + 1] += start[j];
    for (int64_t i = 0; i < n; ++i) {
      keyed_t* e = &s[start[labels[i]]++];
      e->label = labels[i];
      e->index = i;
    }
    free(start);
  } else {
    for (int64_t i = 0; i < n; ++i) {
      s[i].label = labels[i];
      s[i].index = i;
    }
    qsort(s, (size_t)n, sizeof(keyed_t), cmp_keyed);
  }
  /* upper_bound cumulative counts, stored as float (centroids.cu:53-57) */
  float* rval = (float*)malloc((size_t)k * sizeof(float));
  int64_t pos = 0;
  for (int j = 0; j < k; ++j) {
    while (pos < n && s[pos].label <= j) ++pos;
    rval[j] = (float)pos;
  }
  /* adjacent_difference -> reciprocal, in place (centroids.cu:60-63) */
  for (int j = k - 1; j >= 1; --j) rval[j] = 1.0f / (rval[j] - rval[j - 1]);
  if (k > 0) rval[0] = 1.0f / rval[0];
  for (int j = 0; j < 3 * k; ++j) cent[j] = 0.0f; /* :67 */
  /* reduce_by_key, compacted by distinct key (centroids.cu:78-90) */
  int slot = 0;
  for (int64_t a = 0; a < n;) {
    int64_t b = a;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    while (b < n && s[b].label == s[a].label) {
      const int64_t i = s[b].index;
      sx += pts[i];
      sy += pts[n + i];
      sz += pts[2 * n + i];
      ++b;
    }
    if (slot < k) {
      cent[slot] = sx;
      cent[k + slot] = sy;
      cent[2 * k + slot] = sz;
    }
    ++slot;
    a = b;
  }
  /* multiply by the reciprocal counts, cycling over k (:93-100) */
  for (int ch = 0; ch < 3; ++ch)
    for (int j = 0; j < k; ++j) cent[ch * k + j] *= rval[j];
  free(rval);
  free(s);
}

int ref_cluster(const float* pts, int64_t n, float* cent, int k,
                int32_t* labels, int max_iter, float threshold,
                float* last_delta) {
  float* old = (float*)malloc((size_t)3 * k * sizeof(float));
  float* rv = (float*)malloc((size_t)k * sizeof(float));
  float delta = 0.f;
  float old_delta = delta + threshold + 1.f;
  int iter = 0;
  while (iter < max_iter && fabsf(delta - old_delta) > threshold) {
    ++iter;
    /* saved centroids live at temp + 2N floats (cluster.cu:40-50,75-77) */
    memcpy(old, cent, (size_t)3 * k * sizeof(float));
    /* label_points writes N*K distances from temp[0] (label.cu:85-89),
     * which overwrite the saved copy where 2N + j < N*K */
    for (int j = 0; j < 3 * k; ++j) {
      const int64_t flat = 2 * n + j;
      if (flat < n * (int64_t)k)
        old[j] = ref_dist(pts, n, flat / k, cent, k, (int)(flat % k));
    }
    ref_label_points(cent, k, pts, n, labels);
    ref_calculate_centroids(labels, pts, n, cent, k);
    /* calculate_centroids' reciprocal counts sit at temp + 2N floats
     * (centroids.cu:44-45), i.e. over the first k saved values */
    {
      /* recompute them exactly as ref_calculate_centroids does */
      int64_t* cnt = (int64_t*)calloc((size_t)k + 1, sizeof(int64_t));
      for (int64_t i = 0; i < n; ++i)
        if (labels[i] >= 0 && labels[i] < k) cnt[labels[i]]++;
      float cum_prev = 0.f;
      int64_t cum = 0;
      for (int j = 0; j < k; ++j) {
        cum += cnt[j];
        const float cf = (float)cum;
        rv[j] = j == 0 ? 1.0f / cf : 1.0f / (cf - cum_prev);
        cum_prev = cf;
      }
      free(cnt);
      for (int j = 0; j < k; ++j) old[j] = rv[j];
    }
    old_delta = delta;
    float ss = 0.f;
    for (int j = 0; j < 3 * k; ++j) {
      const float d = cent[j] - old[j];
      ss += d * d;
    }
    delta = sqrtf(ss);
  }
  if (last_delta) *last_delta = delta;
  free(rv);
  free(old);
  return iter;
}

int ref_connected_components(const int32_t* labels, int W, int H,
                             int32_t* out, int max_iter) {
  const int64_t n = (int64_t)W * H;
  int32_t* row_keys = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  int32_t* col_keys = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  /* horizontal runs in row-major order (connected_components.cu:36-55,120) */
  int32_t r = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (i > 0 && (i / W != (i - 1) / W || labels[i] != labels[i - 1])) ++r;
    row_keys[i] = r;
  }
  /* vertical runs in column-major order, transposed back (:110-127) */
  int32_t v = 0;
  int prev_x = -1, prev_l = 0;
  for (int64_t t = 0; t < n; ++t) {
    const int x = (int)(t / H), y = (int)(t % H);
    const int32_t l = labels[(int64_t)y * W + x];
    if (t > 0 && (x != prev_x || l != prev_l)) ++v;
    col_keys[(int64_t)y * W + x] = v;
    prev_x = x;
    prev_l = l;
  }
  for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)i; /* :131 */
  const int64_t nmax = (r > v ? r : v) + 1;            /* :133 */
  int32_t* maxima = (int32_t*)calloc((size_t)nmax, sizeof(int32_t));
  int32_t* buffer = (int32_t*)calloc((size_t)nmax, sizeof(int32_t));
  int32_t* old = (int32_t*)calloc((size_t)nmax, sizeof(int32_t));
  old[0] = -1; /* :143 */
  int iter;
  for (iter = 0; iter < max_iter; ++iter) {
    if (memcmp(maxima, old, (size_t)nmax * sizeof(int32_t)) == 0) break;
    memcpy(old, maxima, (size_t)nmax * sizeof(int32_t));
    /* max over each horizontal run (:156-162) */
    for (int64_t i = 0; i < n; ++i) {
      const int32_t k = row_keys[i];
      if (i == 0 || row_keys[i - 1] != k)
        buffer[k] = out[i];
      else if (out[i] > buffer[k])
        buffer[k] = out[i];
    }
    /* max over each vertical run of the run maxima (:166-175) */
    int32_t last = -1;
    for (int64_t t = 0; t < n; ++t) {
      const int x = (int)(t / H), y = (int)(t % H);
      const int64_t p = (int64_t)y * W + x;
      const int32_t k = col_keys[p];
      const int32_t val = buffer[row_keys[p]];
      if (k != last)
        maxima[k] = val;
      else if (val > maxima[k])
        maxima[k] = val;
      last = k;
    }
    /* gather (:177-178) */
    for (int64_t i = 0; i < n; ++i) out[i] = maxima[col_keys[i]];
  }
  free(old);
  free(buffer);
  free(maxima);
  free(col_keys);
  free(row_keys);
  return iter;
}

int64_t ref_compress_labels(int32_t* labels, int64_t n) {
  return oracle_compress(labels, n);
}
