// mbr_kernels.hpp — launch interface between the host API and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mbr_scene_dev.hpp"

namespace mbr {

// Counters accumulated on the device (one atomic add per workgroup).
enum StatIdx {
  ST_MARCH_DE = 0, ST_SHADE_DE, ST_MARCH_ITERS, ST_SHADE_ITERS, ST_HITS, ST_MISSED, ST_REDO,
  ST_OVF_RAYS,   // rays whose step record outgrew the slab
  ST_POOL_MISS,  // records dropped because the overflow pool was full
  ST_OVF_CHUNKS, // overflow chunks taken
  ST_AO_DE, ST_AO_ITERS, ST_AO_STEPS,
  ST_COUNT = 14
};

// Band counters (BandBufs::counters), zeroed on the device at band start:
// hits appended, rays handed to the FP64 continuation, primary-ray queue,
// overflow runs handed out, AO work queue, rays listed for the FP64 redo,
// overflow chunks.
enum CounterIdx {
  CNT_HITS = 0, CNT_HAND = 1, CNT_RAYQ = 2, CNT_DIRS = 3, CNT_AOQ = 4, CNT_REDO = 5, CNT_CHUNKS = 6, CNT_COUNT = 8
};

// Formulas of the render core; the generic ones (no kernels of their own)
// run in FP64 through the F = 0 kernels.
inline bool formula_supported(int f) { return (f >= 1 && f <= 8) || (f >= 10 && f <= 37); }
// Kernel family of a scene: its formula, or kGenericFamily when the scene
// turns on the IFS folding pre-fold, which only the generic kernels run.
constexpr int kGenericFamily = 1000;
inline bool formula_generic(int f) {
  return f == kGenericFamily || (formula_supported(f) && f != 2 && f != 7 && f != 8);
}

// Maximum step records of one ray (the loop bound of src/Render3D.cpp:95).
constexpr int kMaxSteps = 10000;
// Overflow records come in chunks of kChunk; a ray's directory lists the
// chunks of its records slab, slab + kChunk, ... (at most kDirSlots).
constexpr int kChunk = 256;
constexpr int kDirSlots = 40;  // >= (kMaxSteps - slab) / kChunk for any slab >= 0

// Per-band device buffers.  A band is a run of consecutive pixels of the
// image (row-major), at most `cap` rays.  Step records are step-major
// structure-of-arrays: record i < slab of ray r lives at [i * cap + r];
// records from `slab` on go to chunks of kChunk records of the overflow
// pool, listed in a per-ray directory, both taken when first needed.
struct BandBufs {
  int cap;          // rays per band
  int slab;         // step records kept per ray in the slab
  int rec_bytes;    // 4 (FP32 records) or 8 (FP64 records)
  // per ray
  uint8_t *found, *object;
  int *count;       // step record length (buffCount)
  double *depth, *thr, *last, *px, *py, *pz;
  void *rec_step, *rec_dist, *rec_thr;   // slab planes [slab][cap]
  int *rec_it;                           // iteration counts [slab][cap] (iteration fog only, else null)
  long long *ovf_off;                    // per ray: its directory, -1 none, -2 pool full
  // compaction lists: hits, rays to re-march in FP64
  int *hit_list, *redo_list;
  unsigned long long *counters;          // CNT_* below
  // rays handed from the march to its FP64 continuation
  int *hand_ray, *hand_i;
  double *hand_scan, *hand_step, *hand_drift, *hand_dist;
  // overflow pool: ovf_dirs directories of kDirSlots chunk indices,
  // ovf_chunks chunks of kChunk records
  int *ovf_dir;
  void *ovf_step, *ovf_dist, *ovf_thr;
  int *ovf_it;                           // with rec_it
  long long ovf_dirs, ovf_chunks;
  // shade output per ray
  void *sh_r, *sh_g, *sh_b;
  uint8_t *col_r, *col_g, *col_b;
  // raytraced reflections: the specular part of the shade, kept apart
  // (sh_* then holds the object part); per ray of a reflection level its
  // origin, direction and initial threshold, and whether it exists
  void *sp_r, *sp_g, *sp_b;
  double *ray_o, *ray_d, *ray_t0;
  uint8_t *active;
  // the context's cancel flag (mbr_cancel): pinned host memory the march
  // and AO kernels poll while handing out work; null: never polled
  const volatile int *cancel;
};

struct FrameOut {
  float *rgb[3];
  float *z;
  uint8_t *colour[3];
  uint16_t *alpha, *opacity;
};

struct AODev {
  const float *dir_f;
  const double *dir_d;
  const int *rgb;
  int count;
};

// precision: 0 = FP64; 1 = FP64 march, FP32 records and shading; 2 = FP32
// march with FP64 continuation, FP32 records and shading (mbr_precision)
// All launches of a band read their work counts (hits, hand-overs) from
// b.counters on the device: a band is enqueued without host round trips.
// persist_blocks: workgroups of the persistent launches (a few per CU).
hipError_t launch_band_reset(const BandBufs &b, hipStream_t s);
hipError_t launch_zero(unsigned long long *p, int n, hipStream_t s);
// ip8: the scene is a Mandelbulb of power 8 (kernels with the power fixed
// at compile time for the mixed-precision modes)
hipError_t launch_march(int precision, const DevScene *sc, const BandBufs &b, int band_first, int band_rays,
                        int formula, int plain, int persist_blocks, unsigned long long *stats, hipStream_t s,
                        bool ip8 = false);
hipError_t launch_march_fix(int precision, const DevScene *sc, const BandBufs &b, int band_first, int formula,
                            int persist_blocks, unsigned long long *stats, hipStream_t s, bool ip8 = false);
// Buffers of the shading launches (element type T = the path's precision).
struct ShadeBufs {
  void *base, *spec, *col, *fast_ao, *ao;
  void *aux;                    // [6][n]: aux light diffuse and specular terms
  long long n;                  // hit capacity (plane stride)
  int slow_ao;                  // 1: run the AmbientOcclusion rays
  int deep;                     // 1: FP32 shading starts distance estimates in FP64 where needed
  int plain;                    // 1: no primitives, limits, interior, iteration threshold, pre-folds, Julia
  int formula;
  int lights;                   // 1: aux lights take part (k_aux runs after k_surface)
  int ip8;                      // 1: Mandelbulb of power 8 (AO kernel with the power fixed)
};
// Shading in three launches: surface (normals, shadow, specular, colour,
// fast AO), AO rays, composition; ev_ao (when not null) is recorded around
// the AO launch.
hipError_t launch_shade(int precision, const DevScene *sc, const BandBufs &b, int band_first, AODev ao,
                        ShadeBufs sbuf, int persist_blocks, unsigned long long *stats, hipStream_t s,
                        hipEvent_t ev_ao0, hipEvent_t ev_ao1);
// Scene parameters the adjoint differentiates (radians), and zoom/persp.
struct AdjParams {
  double power, vp[3], alpha, beta, gamma, light_alpha, light_beta;
  double zoom_persp;  // the ray start sits zoom/persp behind vp
};
// dL/dtheta of the band's pixels added into grad[9] (device doubles); g* are
// the image adjoint planes dL/dC (full image, row-major).  With jac (27
// planes of the full image, plane i*3+ch = dC_ch/dtheta_i) every pixel is
// visited and its tangents written; g* and grad may then be null.
hipError_t launch_adjoint(const DevScene *sc, const BandBufs &b, int band_first, int band_rays, AdjParams p,
                          const float *gR, const float *gG, const float *gB, double *grad, uint8_t *skipped,
                          float *jac, int blocks, hipStream_t s);
// Background, VolumetricShader and framebuffer planes.  fx: the scene has
// volumetric lights or iteration fog, whose shadow marches (and AO rays)
// run in a grid-stride kernel of persist_blocks workgroups templated on the
// formula; vis: visible aux lights glow along the rays.
hipError_t launch_finish(int precision, const DevScene *sc, const BandBufs &b, int band_first, int band_rays,
                         FrameOut out, hipStream_t s, int formula, int fx, int vis, int persist_blocks,
                         unsigned long long *stats, AODev ao);

// Raytraced reflections (MainThread, src/Render3D.cpp:451-560), FP64.
// launch_refl_setup: for the hits of level `prev` whose surface reflects,
// the mirrored ray of level `next` (origin, direction, initial threshold)
// and its entry in next's ray list; launch_refl_march: the FP64 march of
// next's rays; launch_refl_finish: per pixel, the levels shaded from the
// last to the first (levels: device array of n_levels BandBufs).
hipError_t launch_refl_clear(const BandBufs &next, int band_rays, hipStream_t s);
hipError_t launch_refl_setup(const DevScene *sc, const BandBufs &prev, const BandBufs &next, int band_first,
                             int formula, int blocks, unsigned long long *stats, hipStream_t s);
hipError_t launch_refl_march(const DevScene *sc, const BandBufs &next, int band_first, int formula, int blocks,
                             unsigned long long *stats, hipStream_t s);
hipError_t launch_refl_finish(const DevScene *sc, const BandBufs *levels, int n_levels, int band_first,
                              int band_rays, FrameOut out, int vis, int fx, int formula, AODev ao,
                              unsigned long long *stats, hipStream_t s);

// Screen-space ambient occlusion of one compiled frame (mbr_post.hip).
// cs/sn: device tables of cos/sin(2 pi i / quality), i < quality.
struct SsaoParams {
  int width, height, quality, type;  // type: enum mbr_perspective
  double scale, aspect, fov, intensity;
  const double *cs, *sn;
};
hipError_t launch_ssao(const SsaoParams &p, const float *z, const uint16_t *opacity, const uint8_t *const colour[3],
                       uint16_t *image16, hipStream_t s);

// CompileImage on the device (mbr_post.hip): gamma = 65536-entry table.
struct CompileParams {
  int n, hdr;
  double brightness, contrast;
};
hipError_t launch_compile(const CompileParams &p, const float *r, const float *g, const float *b,
                          const uint16_t *gamma, uint16_t *image16, hipStream_t s);

// Depth of field (mbr_post.hip, mbr_post_order.cpp).  tiles_x/tiles_y:
// 16x16 pixel tiles; rank_bits: bits of a splat rank in the sort key.
struct DofParams {
  int width, height, tiles_x, tiles_y, rank_bits;
  double neutral, deep;
};
// host: rank[pixel] = its position in the splat order
void dof_order(const float *z, int width, int height, double neutral, double deep, uint32_t seed, uint32_t *rank);
hipError_t dof_count(const DofParams &p, const float *z, unsigned long long *cnt, hipStream_t s);
hipError_t dof_scan(void *tmp, size_t &tmp_bytes, const unsigned long long *cnt, unsigned long long *off, int n,
                    hipStream_t s);
hipError_t dof_fill(const DofParams &p, const float *z, const uint32_t *rank, const unsigned long long *off,
                    unsigned long long *keys, uint32_t *vals, hipStream_t s);
hipError_t dof_sort(void *tmp, size_t &tmp_bytes, const unsigned long long *kin, unsigned long long *kout,
                    const uint32_t *vin, uint32_t *vout, int n, int bits, hipStream_t s);
// seg: tiles + 1 entries (filled here); keys/vals: the sorted lists
hipError_t dof_fold(const DofParams &p, const float *z, const unsigned long long *keys, long long n,
                    unsigned long long *seg, const uint32_t *vals, const uint16_t *img_in, const uint16_t *alpha_in,
                    uint16_t *img_out, uint16_t *alpha_out, hipStream_t s);

}  // namespace mbr
