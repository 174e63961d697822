/* Seeded synthetic inputs for the segmentation path (host C). */
#ifndef SPLIT_GEN_H
#define SPLIT_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct split_gen_params {
  uint64_t seed;
  int32_t sites;        /* Voronoi sites, uniform in [0,W)x[0,H)            */
  int32_t shading;      /* 1: multiply by max(0, n.L) of a value-noise height */
  int32_t octaves;      /* value-noise octaves (frequency x2, amplitude /2)   */
  int32_t reserved;
  double base_cell;     /* lattice cell of the first octave, in pixels        */
  double height_amp;    /* height-map amplitude, in pixels                    */
  double noise_sigma;   /* iid Gaussian noise per channel (0 = none)          */
  double noise_uniform; /* iid uniform noise in [-a, a] per channel (0 = none)*/
} split_gen_params;

/* Three float planes of W*H pixels; values clamped to [0, 1]. */
int split_gen_image(float* planes, int W, int H, const split_gen_params* p);

/* Connected-component stress labels: square 1-px spirals with 1-px
 * corridors tiling the top 7/8 (label 1 wall, 0 corridor, 2 leftover strip),
 * and a 1-px checkerboard band (labels 0/1) in the bottom H/8 rows. */
int split_gen_cc_stress(int32_t* labels, int W, int H);

#ifdef __cplusplus
}
#endif

#endif
