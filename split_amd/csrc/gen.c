/*
 * Seeded synthetic inputs for the segmentation path: Voronoi colour images
 * with optional Lambertian shading and noise, and the connected-component
 * stress image of square spirals plus a checkerboard band.
 *
 * Every random number comes from counter-based SplitMix64 hashes of
 * (seed, purpose, index), so an image is a pure function of its parameters
 * whatever the thread count.  Arithmetic is IEEE double with FMA contraction
 * disabled at build time (-ffp-contract=off), and the transcendental calls
 * (log, cos, sqrt) are glibc's, so the same image is produced on any host
 * that runs this generator.
 *
 * Layout of an image: three float planes of W*H pixels (SoA, row-major).
 */
#include "split_gen.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Hash of (seed, stream, a, b) -> 64 random bits. */
static uint64_t hash4(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b) {
  uint64_t h = mix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
  h = mix64(h ^ (a * 0x8CB92BA72F3D8DD7ull));
  h = mix64(h ^ (b * 0xABC98388FB8FAC03ull));
  return h;
}

/* [0, 1) with 53 random bits. */
static double u01(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

enum { STREAM_SITE = 1, STREAM_LATTICE = 2, STREAM_NOISE = 3 };

typedef struct {
  double x, y;
  float r, g, b;
} site_t;

/* Nearest site by squared distance to the pixel centre, ties to the lower
 * site index, using a uniform bucket grid and ring search. */
typedef struct {
  int gw, gh;
  double cell;
  int* start; /* gw*gh+1 offsets into idx */
  int* idx;
} site_grid_t;

static int grid_build(site_grid_t* g, const site_t* s, int n, int W, int H) {
  g->cell = sqrt((double)W * (double)H / (double)n);
  if (g->cell < 1.0) g->cell = 1.0;
  g->gw = (int)ceil(W / g->cell);
  g->gh = (int)ceil(H / g->cell);
  const int nc = g->gw * g->gh;
  g->start = (int*)calloc((size_t)nc + 1, sizeof(int));
  g->idx = (int*)malloc((size_t)n * sizeof(int));
  int* fill = (int*)calloc((size_t)nc, sizeof(int));
  if (!g->start || !g->idx || !fill) {
    free(fill);
    return -1;
  }
  for (int i = 0; i < n; ++i) {
    int cx = (int)(s[i].x / g->cell), cy = (int)(s[i].y / g->cell);
    if (cx >= g->gw) cx = g->gw - 1;
    if (cy >= g->gh) cy = g->gh - 1;
    g->start[cy * g->gw + cx + 1]++;
  }
  for (int c = 0; c < nc; ++c) g->start[c + 1] += g->start[c];
  for (int i = 0; i < n; ++i) {
    int cx = (int)(s[i].x / g->cell), cy = (int)(s[i].y / g->cell);
    if (cx >= g->gw) cx = g->gw - 1;
    if (cy >= g->gh) cy = g->gh - 1;
    const int c = cy * g->gw + cx;
    g->idx[g->start[c] + fill[c]++] = i; /* ascending site index per cell */
  }
  free(fill);
  return 0;
}

static void grid_free(site_grid_t* g) {
  free(g->start);
  free(g->idx);
}

static int nearest_site(const site_grid_t* g, const site_t* s, double px,
                        double py) {
  int cx = (int)(px / g->cell), cy = (int)(py / g->cell);
  if (cx >= g->gw) cx = g->gw - 1;
  if (cy >= g->gh) cy = g->gh - 1;
  double best = INFINITY;
  int bi = -1;
  const int rmax = g->gw > g->gh ? g->gw : g->gh;
  for (int r = 0; r <= rmax; ++r) {
    for (int yy = cy - r; yy <= cy + r; ++yy) {
      if (yy < 0 || yy >= g->gh) continue;
      const int ring_row = (yy == cy - r || yy == cy + r);
      for (int xx = cx - r; xx <= cx + r; xx += (ring_row ? 1 : 2 * r)) {
        if (xx >= 0 && xx < g->gw) {
          const int c = yy * g->gw + xx;
          for (int t = g->start[c]; t < g->start[c + 1]; ++t) {
            const int i = g->idx[t];
            const double dx = px - s[i].x, dy = py - s[i].y;
            const double d = dx * dx + dy * dy;
            if (d < best || (d == best && i < bi)) {
              best = d;
              bi = i;
            }
          }
        }
        if (r == 0) break;
      }
    }
    /* Every cell of ring r+1 is at least r full cells away. */
    const double lb = (double)r * g->cell;
    if (bi >= 0 && best < lb * lb) break;
  }
  return bi;
}

/* Value noise: lattice of hashed values, smoothstep-weighted bilinear. */
static double lattice(uint64_t seed, int octave, int64_t ix, int64_t iy) {
  return u01(hash4(seed, STREAM_LATTICE + 16 * (uint64_t)octave,
                   (uint64_t)ix, (uint64_t)iy));
}

static double value_noise(uint64_t seed, int octave, double x, double y) {
  const double fx = floor(x), fy = floor(y);
  const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
  double tx = x - fx, ty = y - fy;
  tx = tx * tx * (3.0 - 2.0 * tx);
  ty = ty * ty * (3.0 - 2.0 * ty);
  const double a = lattice(seed, octave, ix, iy);
  const double b = lattice(seed, octave, ix + 1, iy);
  const double c = lattice(seed, octave, ix, iy + 1);
  const double d = lattice(seed, octave, ix + 1, iy + 1);
  const double top = a + (b - a) * tx;
  const double bot = c + (d - c) * tx;
  return top + (bot - top) * ty;
}

static double height_at(const split_gen_params* p, int x, int y) {
  double h = 0.0, amp = 1.0, cell = p->base_cell;
  for (int o = 0; o < p->octaves; ++o) {
    h += amp * value_noise(p->seed, o, (x + 0.5) / cell, (y + 0.5) / cell);
    amp *= 0.5;
    cell *= 0.5;
  }
  return p->height_amp * h;
}

static double gauss(uint64_t seed, uint64_t key) {
  const uint64_t a = hash4(seed, STREAM_NOISE, key, 0);
  const uint64_t b = hash4(seed, STREAM_NOISE, key, 1);
  const double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53; /* (0, 1] */
  const double u2 = u01(b);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static double uniform_pm1(uint64_t seed, uint64_t key) {
  return 2.0 * u01(hash4(seed, STREAM_NOISE, key, 2)) - 1.0;
}

int split_gen_image(float* planes, int W, int H, const split_gen_params* p) {
  if (!planes || !p || W < 1 || H < 1 || p->sites < 1) return -1;
  const int64_t n = (int64_t)W * H;
  site_t* s = (site_t*)malloc((size_t)p->sites * sizeof(site_t));
  if (!s) return -1;
  for (int i = 0; i < p->sites; ++i) {
    s[i].x = u01(hash4(p->seed, STREAM_SITE, (uint64_t)i, 0)) * W;
    s[i].y = u01(hash4(p->seed, STREAM_SITE, (uint64_t)i, 1)) * H;
    s[i].r = (float)u01(hash4(p->seed, STREAM_SITE, (uint64_t)i, 2));
    s[i].g = (float)u01(hash4(p->seed, STREAM_SITE, (uint64_t)i, 3));
    s[i].b = (float)u01(hash4(p->seed, STREAM_SITE, (uint64_t)i, 4));
  }
  site_grid_t g;
  if (grid_build(&g, s, p->sites, W, H) != 0) {
    free(s);
    return -1;
  }
  double* h = NULL;
  if (p->shading) {
    h = (double*)malloc((size_t)n * sizeof(double));
    if (!h) {
      grid_free(&g);
      free(s);
      return -1;
    }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) h[(int64_t)y * W + x] = height_at(p, x, y);
  }
  const double ln = sqrt(0.3 * 0.3 + 0.3 * 0.3 + 0.906 * 0.906);
  const double Lx = 0.3 / ln, Ly = 0.3 / ln, Lz = 0.906 / ln;
#pragma omp parallel for schedule(dynamic, 16)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      const int64_t i = (int64_t)y * W + x;
      const site_t* st = &s[nearest_site(&g, s, x + 0.5, y + 0.5)];
      double shade = 1.0;
      if (h) {
        const int xl = x > 0 ? x - 1 : x, xr = x < W - 1 ? x + 1 : x;
        const int yu = y > 0 ? y - 1 : y, yd = y < H - 1 ? y + 1 : y;
        const double dhdx =
            (h[(int64_t)y * W + xr] - h[(int64_t)y * W + xl]) / (xr - xl > 0 ? xr - xl : 1);
        const double dhdy =
            (h[(int64_t)yd * W + x] - h[(int64_t)yu * W + x]) / (yd - yu > 0 ? yd - yu : 1);
        const double nl = sqrt(dhdx * dhdx + dhdy * dhdy + 1.0);
        const double d = (-dhdx * Lx - dhdy * Ly + Lz) / nl;
        shade = d > 0.0 ? d : 0.0;
      }
      const float alb[3] = {st->r, st->g, st->b};
      for (int c = 0; c < 3; ++c) {
        double v = (double)alb[c] * shade;
        const uint64_t key = (uint64_t)i * 3 + (uint64_t)c;
        if (p->noise_sigma > 0.0) v += p->noise_sigma * gauss(p->seed, key);
        if (p->noise_uniform > 0.0)
          v += p->noise_uniform * uniform_pm1(p->seed, key);
        v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
        planes[(int64_t)c * n + i] = (float)v;
      }
    }
  }
  free(h);
  grid_free(&g);
  free(s);
  return 0;
}

int split_gen_site_count(int W, int H, const split_gen_params* p) {
  (void)W;
  (void)H;
  return p ? p->sites : -1;
}

/* Square spiral wall of 1-px width with a 1-px corridor, drawn as a turtle
 * walk inside the side x side square at (ox, oy): segment lengths
 * side-1 (x3), then side-3 (x2), side-5 (x2), ... turning clockwise. */
static void draw_spiral(int32_t* lab, int W, int ox, int oy, int side,
                        int32_t wall) {
  int x = 0, y = 0, dir = 0; /* 0 right, 1 down, 2 left, 3 up */
  static const int dx[4] = {1, 0, -1, 0}, dy[4] = {0, 1, 0, -1};
  lab[(int64_t)(oy + y) * W + ox + x] = wall;
  int len = side - 1, seg = 0;
  while (len > 0) {
    for (int t = 0; t < len; ++t) {
      x += dx[dir];
      y += dy[dir];
      lab[(int64_t)(oy + y) * W + ox + x] = wall;
    }
    dir = (dir + 1) & 3;
    ++seg;
    if (seg >= 3 && (seg - 3) % 2 == 0) len -= 2;
  }
}

int split_gen_cc_stress(int32_t* labels, int W, int H) {
  if (!labels || W < 4 || H < 8) return -1;
  const int band = H / 8;
  const int top = H - band;
  const int64_t n = (int64_t)W * H;
  for (int64_t i = 0; i < n; ++i) labels[i] = 0;
  /* Tile the top region with square spirals, left to right; the leftover
   * strip keeps label 2. */
  const int side = top < W ? top : W;
  int ox = 0;
  while (ox + side <= W) {
    draw_spiral(labels, W, ox, 0, side, 1);
    ox += side;
  }
  for (int y = 0; y < top; ++y)
    for (int x = ox; x < W; ++x) labels[(int64_t)y * W + x] = 2;
  /* 1-px checkerboard band at the bottom. */
  for (int y = top; y < H; ++y)
    for (int x = 0; x < W; ++x) labels[(int64_t)y * W + x] = (x + y) & 1;
  return 0;
}
