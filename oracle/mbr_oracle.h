/*
 * mbr_oracle.h — CPU FP64 restatement of the Mandelbulber 1.21 render core.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the timed CPU baseline.  The product (mandelbulber_amd) never
 * links, loads or calls it.
 */
#ifndef MBR_ORACLE_H_
#define MBR_ORACLE_H_

#include <stdint.h>
#include "mbr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Jitter source of the oracle.  NONE and HASH match enum mbr_jitter;
 * GLIBC replays the reference exactly: srand(seed) then rand()%1001 per
 * march step in pixel order (single thread only). */
enum oracle_jitter { ORACLE_JITTER_NONE = 0, ORACLE_JITTER_HASH = 1, ORACLE_JITTER_GLIBC = 2 };

typedef struct oracle_stats {
  uint64_t rays, hits;
  uint64_t march_de_calls, shade_de_calls, iterations, missed_de;
  uint64_t max_steps;        /* largest step record of one ray */
  uint64_t total_steps;      /* sum of step-record lengths */
} oracle_stats;

/* One primary ray's march (RayMarching, src/Render3D.cpp:79-199). */
typedef struct oracle_march_out {
  double point[3];
  double depth;              /* final scan */
  double dist_thresh;        /* distThresh after the bisection */
  double last_dist;
  int found;
  int step_count;            /* buffCount */
  int object;                /* objectOut of the last DE call */
  int de_calls;
} oracle_march_out;

uint32_t oracle_jitter_draw(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t step);

/* Compute<Mode> (src/fractal.cpp:40-1241); mode 0 = normal, 1 = colouring.
 * n_iter = N used by the call (N, N*5 for normals, N*10 for colouring). */
double oracle_compute(const mbr_scene *s, int mode, const double p[3], int n_iter, int *iters);

/* CalculateDistance (src/fractal.cpp:1245-1448). */
double oracle_calculate_distance(const mbr_scene *s, const double p[3], double detail_size,
                                 int n_iter, int min_n, int normal_mode,
                                 int *max_iter, int *iters, int *object);

/* Camera of MainThread (src/Render3D.cpp:236-256, :325-333): rotation
 * matrix (row-major m11..m33), ray start and main light vector. */
void oracle_camera(const mbr_scene *s, double R[9], double start[3], double light[3]);

/* Per-pixel view vector (src/Render3D.cpp:389-438); returns 1 when the
 * fish-eye cut removes the pixel. */
int oracle_view_vector(const mbr_scene *s, int x, int y, double dir[3]);

/* AO direction table (src/Render3D.cpp:271-313); returns the count.
 * dirs: max*3 doubles, rgb: max*3 ints. lightmap NULL = white fallback. */
int oracle_ao_table(int quality, const mbr_texture *lightmap, double *dirs, int *rgb, int max);

/* RayMarching with the given start and direction; jitter from the hash. */
int oracle_march(const mbr_scene *s, const double start[3], const double dir[3],
                 uint32_t pixel, oracle_march_out *out);

/* Render rows [row_begin,row_end) of the scene into host planes (planes
 * sized for the whole image).  nthreads <= 0 uses the host core count. */
int oracle_render(const mbr_scene *s, const mbr_texture *lightmap, int row_begin, int row_end,
                  int nthreads, int jitter, uint32_t glibc_seed, mbr_framebuffer *fb,
                  oracle_stats *st);

#ifdef __cplusplus
}
#endif
#endif
