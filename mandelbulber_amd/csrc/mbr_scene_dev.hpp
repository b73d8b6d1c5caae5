// mbr_scene_dev.hpp — the device-resident scene block.
//
// The host derives it once per mbr_set_scene from the FP64 mbr_scene POD:
// formula constants, rotation matrices, camera, light vector and shading
// constants, each in FP64 and FP32.  Kernels read it through a
// `const DevScene* __restrict__` whose every access is wave-uniform, so the
// compiler fetches it with scalar loads into SGPRs: the formula constants
// never occupy vector registers or LDS.
#pragma once

#include <stdint.h>

namespace mbr {

// Formula and CalculateDistance constants (src/fractal.h:102-293).
template <class T>
struct FractalK {
  int formula, N, minN;
  int julia_mode, tglad_fold, sph_fold, mb_rot_on, mb_main_rot;
  int interior, iter_thresh, limits, only_plane;
  int prim_plane, prim_box, prim_ibox, prim_sphere, prim_isphere, prim_water;
  int water_iterations, frame_no, water_rotated;
  int ipow;  // power as an integer in [2, 32] when it is one, else 0
  int linear_de;  // linear_DE_mode: delta-DE distance 0.5 r / dr instead of 0.5 r ln r / dr
  T mb_sharp, fip_factor, fip_zfactor, c_add;  // smoothMandelbox, foldingIntPow2, aexion
  // kaleidoscopic IFS (RecalculateIFSParams, src/callbacks.cpp:1538-1546):
  // per enabled vector its rotation (identity flag), unit direction,
  // distance and intensity; the main rotation about the offset
  int ifs_abs[3], ifs_menger, ifs_en[9], ifs_rot_on[9], ifs_main_on;
  T ifs_scale, ifs_offset[3], ifs_edge[3];
  T ifs_rot[9][9], ifs_dir[9][3], ifs_dist[9], ifs_int[9], ifs_main[9];
  int ifs_fold;  // IFS folding mode: the same vectors as a pre-fold of every iteration
  // fake lights' orbit trap: iterations summed and the trap point
  int fl_min, fl_max;
  T fl_trap[3];
  // generalizedFoldBox: face normals of the fold polyhedron (genFoldBox.type),
  // the fold plane offset (solid) and the melt layer
  int gf_sides;
  T gf_nv[20][3], gf_solid, gf_melt;
  // mandelboxVaryScale4D: scale variation, fold, minimum radius, radius power, w addend
  T v4_scale_vary, v4_fold, v4_minr, v4_rpower, v4_wadd;
  // hybrid sequence: formulas, iterations and powers of the five slots,
  // cyclic flag and the length of one cycle
  int hyb_f[5], hyb_it[5], hyb_cyclic, hyb_period;
  T hyb_p[5];
  T power, bailout, constant_factor;
  T julia[3];
  T fold_limit, fold_value, sph_fR2, sph_mR2, sph_factor1;
  T mb_scale, mb_abs_scale, mb_limit, mb_value, mb_mR2, mb_fR2, mb_factor1;
  T mb_offset[3];
  T mb_col_R, mb_col_X, mb_col_Y, mb_col_Z, mb_col_Sp1, mb_col_Sp2;
  T mb_main[9];
  T mb_rot[2][3][9];
  T mb_rotinv[2][3][9];
  T limit_min[3], limit_max[3];
  T plane_c[3], plane_n[3];
  T box_c1[3], box_c2[3], ibox_c1[3], ibox_c2[3];
  T sph_c[3], sph_r, isph_c[3], isph_r;
  T water_height, water_amplitude, water_length, water_rot[9], water_phase;
};

// Shading constants (sParamRender fields used by src/shaders.cpp).
template <class T>
struct ShadeK {
  T quality, DE_factor, smoothness, resolution, persp, view_distance_max;
  T light[3];           // shadow vector, MainThread src/Render3D.cpp:252-256
  T main_light[3];      // mainLightIntensity * colour / 65536 (ObjectShader :28-30)
  T direct_light, shading, specular, global_ilum, fast_ao_tune;
  T soft_range;         // tan(shadowConeAngle)
  T coloring_speed, palette_offset;
  T fog_visibility;     // pow(10, post_fog_visibility/10 - 16)
  T fog_density, fog_c1_dist, fog_c2_dist, fog_reduce;
  T glow_intensity;
  T aux_intensity, aux_visibility;  // auxLightIntensity, auxLightVisibility
};

struct DevScene {
  // image / camera (FP64: ray origins and scan stay FP64 in every path)
  int width, height, tiles, tile_index;
  int persp_type, fish_eye_cut, frame_no, jitter_mode, hit_refine;
  int deep;  // the pixel footprint can fall below FP32 resolution of the coordinates
  int plain; // no primitives, limits, interior mode, iteration threshold, pre-folds or Julia constant
  uint32_t jitter_seed, pixel_base;  // pixel index offset of this tile
  double aspect, fov, tile_x, tile_y;
  double R[9];
  int R_identity, pad1;
  double start[3];
  double view_distance_min, view_distance_max;

  // switches
  int constant_thresh, interior, shadow, penetrating;
  int ao, fast_ao, ao_quality, ao_count;
  int slow_shading, coloring, bsoft, post_fog;
  int N, minN;

  // auxiliary lights that take part, in the order of the reference's Lights
  // array: 8 values per light (x, y, z, intensity, colour R, G, B, unused),
  // in FP64 and FP32.  light_div is lightsPlaced (at least 4), the divisor
  // of every light's intensity (src/shaders.cpp:1053-1062).
  const double *lights_d;
  const float *lights_f;
  int n_lights, light_div;
  // fake lights (FakeLights and their glow): on, intensity, visibility and
  // visibility size
  int fake_lights;
  double fl_intensity, fl_visibility, fl_visibility_size;
  // textured background (BackgroundShader's texture branches): the
  // context's background texture (8-bit RGB rows), dome layout
  int textured_bg, fulldome, bg_w, bg_h;
  const uint8_t *bg_tex;
  // environment mapping (EnvMapping, reflect > 0 without raytraced
  // reflections): the context's envmap texture and the reflect factor
  int env_map, env_w, env_h;
  const uint8_t *env_tex;
  double reflect;

  // volumetric lights (VolumetricShader, src/shaders.cpp:314-341): vol_main
  // for the main light; vol_aux_on[k] when aux light slot k is enabled and
  // its volumetric light switched on, vol_aux[k] its position, intensity and
  // colour as in the light table; vol_intensity[0..4] their intensities.
  // shadow_fx: volumetric lights or iteration fog are on, so the shadow
  // marches take unit DE steps and a threshold from the distance to the eye.
  int shadow_fx, vol_main, vol_any;
  int vol_aux_on[4];
  double vol_aux[4][8];
  double vol_intensity[5];
  // iteration fog (IterOpacity, src/shaders.cpp:1214-1222, and its
  // VolumetricShader block :385-465); aux_slot_on[k]: aux light slot k is
  // enabled (the fog lights every enabled slot 0-3)
  int iter_fog;
  int aux_slot_on[4];
  double iter_fog_opacity, iter_fog_trim;
  double main_light_intensity;  // mainLightIntensity (the fog's light)
  // raytraced reflections: bounces (0 = off) and the reflection coefficient
  // of each object type (OBJ_*: fractal, plane, water, sphere, inverted
  // sphere, box, inverted box)
  int refl_max;
  double refl[7];

  // integer colours (sRGB, 16-bit channel values)
  int main_light_colour[3];
  int glow1[3], glow2[3];
  int bg1[3], bg2[3], bg3[3];
  int post_fog_colour[3];
  int fog1[3], fog2[3], fog3[3];
  int plane_col[3], box_col[3], ibox_col[3], sphere_col[3], isphere_col[3], water_col[3];
  int palette[256][3];

  FractalK<double> fd;
  FractalK<float> ff;
  ShadeK<double> sd;
  ShadeK<float> sf;

  template <class T>
  __host__ __device__ const FractalK<T> &frac() const;
  template <class T>
  __host__ __device__ const ShadeK<T> &shade() const;
  template <class T>
  __host__ __device__ const T *lights() const;
};

template <>
__host__ __device__ inline const FractalK<double> &DevScene::frac<double>() const { return fd; }
template <>
__host__ __device__ inline const FractalK<float> &DevScene::frac<float>() const { return ff; }
template <>
__host__ __device__ inline const ShadeK<double> &DevScene::shade<double>() const { return sd; }
template <>
__host__ __device__ inline const ShadeK<float> &DevScene::shade<float>() const { return sf; }
template <>
__host__ __device__ inline const double *DevScene::lights<double>() const { return lights_d; }
template <>
__host__ __device__ inline const float *DevScene::lights<float>() const { return lights_f; }

}  // namespace mbr
