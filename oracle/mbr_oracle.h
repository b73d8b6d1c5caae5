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

/* CRotationMatrix::SetRotation(angles[3]) (src/algebra.cpp:145-151), row-major */
void oracle_rotation3(const double angles[3], double m[9]);

/* VolumetricShader (src/shaders.cpp:185-478) over a synthetic step record:
 * steps[i*3 + {0,1,2}] = {step, distance, distThresh} for i < n.  pixel and
 * out are {R, G, B, alpha}; *opacity receives opacityOut.R (0 when no step
 * group was emitted). */
void oracle_volumetric(const mbr_scene *s, double depth, const double *steps, int n,
                       const double pixel[4], double out[4], double *opacity);

/* IndexToColour (src/shaders.cpp:1273-1305) with the scene's palette */
void oracle_index_to_colour(const mbr_scene *s, int index, int rgb[3]);

/* BackgroundShader gradient (src/shaders.cpp:159-181) for a view vector */
void oracle_background(const mbr_scene *s, const double view[3], double rgb[3]);

/* ObjectShader terms at a hit point (src/shaders.cpp:21-112): normal[3],
 * shade, shadow, specular, ao[3], colour[3] and the shader output rgb[3]
 * (without the specular term, which is spec_rgb[3]). */
typedef struct oracle_shade_out {
  double normal[3];
  double shade, shadow, specular;
  double ao[3], colour[3], rgb[3], spec_rgb[3];
} oracle_shade_out;
int oracle_object_shader(const mbr_scene *s, const double point[3], const double view[3],
                         double dist_thresh, double depth, double last_dist, int object,
                         oracle_shade_out *out);

/* Render rows [row_begin,row_end) of the scene into host planes (planes
 * sized for the whole image).  nthreads <= 0 uses the host core count. */
/* PlaceRandomLights(params, false) (src/shaders.cpp:1134-1212) with the C
 * library's srand/rand: aux_light_number + 4 records; -1 when capacity is
 * short, -2 after 20 million draws without a placement.  Not thread safe
 * (rand). */
int oracle_place_lights(const mbr_scene *s, mbr_light *lights, int capacity);
int oracle_scene_has_lights(const mbr_scene *s);

int oracle_render(const mbr_scene *s, const mbr_texture *lightmap, int row_begin, int row_end,
                  int nthreads, int jitter, uint32_t glibc_seed, mbr_framebuffer *fb,
                  oracle_stats *st);

/* Screen-space ambient occlusion (PostRendering_SSAO / ThreadSSAO,
 * src/image.cpp:249-415) on host planes with the scene's ssao_quality,
 * persp, perspective_type and global_ilum; image16 (interleaved RGB16) in
 * place. */
int oracle_ssao(const mbr_scene *s, int width, int height, const float *z, const uint16_t *opacity,
                const uint8_t *const colour[3], uint16_t *image16);

/* Depth of field (PostRendering_DOF, src/image.cpp:57-247) with deep =
 * dof_radius x width / 1000 and focus 10^(dof_focus / 10 - 16)
 * (src/Render3D.cpp:1101-1105).  The z order and its randomisation draw
 * glibc rand() after srand(seed).  oracle_dof_order returns the pixels in
 * processing order; oracle_dof blurs image16 (interleaved RGB16) and alpha
 * (may be NULL) in place. */
int oracle_dof_order(const mbr_scene *s, int width, int height, const float *z, uint32_t seed, int *order);
int oracle_dof(const mbr_scene *s, int width, int height, const float *z, uint32_t seed, uint16_t *image16,
               uint16_t *alpha);

#ifdef __cplusplus
}
#endif
#endif
