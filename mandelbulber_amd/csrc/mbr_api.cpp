// mbr_api.cpp — host side of the C ABI (include/mbr.h).
//
// mbr_set_scene derives everything MainThread computes once per frame
// (src/Render3D.cpp:236-333) and everything ReadInterface derives
// (src/interface.cpp:688-730, src/callbacks.cpp:1538-1555) into the device
// scene block, after the parameter clamps of CheckPrameters
// (src/interface.cpp:3587-3602).  mbr_render[_device] runs the frame as
// bands of at most kBandRays rays: reset -> march -> FP64 continuation ->
// shade over compacted hits -> finish, all enqueued on one stream with the
// work counts passed between launches in device memory; the host waits once
// per frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mbr.h"
#include "mbr_core.hpp"
#include "mbr_fractal.hpp"
#include "mbr_kernels.hpp"
#include "mbr_rand.hpp"
#include "mbr_scene_dev.hpp"

using namespace mbr;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kBandRays = 1 << 21;   // rays per band (one 1080p frame)
constexpr int kSlab = 96;            // step records per ray kept in the slab
constexpr size_t kAOBudget = size_t(8) << 30;  // bytes of per-(hit, direction) AO intensities per band

// Working memory of mbr_render_device_noalloc: one file-scope device array,
// which every GPU the library runs on holds (its own copy) until the process
// ends.  Host code never names it after set-up: mbr_init takes its device
// address once per device (scratch_ptr) and everything else uses that.
constexpr size_t kScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
// fewest rays per band the scratch must hold for a scene
constexpr int kNoallocMinBand = 16384;

// ---------------------------------------------------- rotation matrices
// CRotationMatrix (src/algebra.cpp:80-218): row-major, post-multiplied
// rotations, each skipped when its angle is 0; `identity` mirrors the
// reference's `zero` flag (RotateVector returns its argument unchanged).
struct Rot {
  double m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  bool identity = true;
  void mul(const double *b) {
    double c[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) c[i * 3 + j] = m[i * 3] * b[j] + m[i * 3 + 1] * b[3 + j] + m[i * 3 + 2] * b[6 + j];
    std::memcpy(m, c, sizeof c);
    identity = false;
  }
  void rx(double a) {
    if (a == 0.0) return;
    double s = std::sin(a), c = std::cos(a), t[9] = {1, 0, 0, 0, c, -s, 0, s, c};
    mul(t);
  }
  void ry(double a) {
    if (a == 0.0) return;
    double s = std::sin(a), c = std::cos(a), t[9] = {c, 0, s, 0, 1, 0, -s, 0, c};
    mul(t);
  }
  void rz(double a) {
    if (a == 0.0) return;
    double s = std::sin(a), c = std::cos(a), t[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
    mul(t);
  }
  void apply(const double v[3], double o[3]) const {
    if (identity) { o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; return; }
    for (int i = 0; i < 3; i++) o[i] = m[i * 3] * v[0] + m[i * 3 + 1] * v[1] + m[i * 3 + 2] * v[2];
  }
};

// SetRotation(angles[3]) (src/algebra.cpp:145-151)
Rot set_rotation3(const double a[3]) {
  Rot r;
  r.rz(a[2]);
  r.ry(a[1]);
  r.rx(a[0]);
  return r;
}

// Face normals of generalizedFoldBox's fold polyhedra in the order
// RecalculateIFSParams lists them (src/callbacks.cpp:1557-1664), selected by
// genFoldBox.type (src/fractal.cpp:538-576; enumGeneralizedFoldBoxType):
// 0 tetrahedron, 1 cube, 2 octahedron, 3 dodecahedron, 4 octahedron + cube,
// 5 icosahedron, 6 hexagonal prism, 7 pentagonal prism; any other type keeps
// the tetrahedron.  Returns the number of faces.
int fold_polyhedron(int type, double nv[20][3]) {
  int n = 0;
  auto put = [&](double x, double y, double z) {
    nv[n][0] = x;
    nv[n][1] = y;
    nv[n][2] = z;
    n++;
  };
  const double t = 1.0 / std::sqrt(3.0);
  // the tetrahedron's normals carry an odd number of minus signs, the four
  // the octahedron adds an even number
  static const int odd[4][3] = {{1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {-1, -1, -1}};
  static const int even[4][3] = {{1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {1, -1, -1}};
  auto diagonals = [&](const int (*sg)[3]) {
    for (int i = 0; i < 4; i++) put(sg[i][0] * t, sg[i][1] * t, sg[i][2] * t);
  };
  auto axes = [&]() {
    for (int a = 0; a < 3; a++)
      for (int sgn = 1; sgn >= -1; sgn -= 2) put(a == 0 ? sgn : 0, a == 1 ? sgn : 0, a == 2 ? sgn : 0);
  };
  const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  switch (type) {
    case 1:
      axes();
      break;
    case 2:
    case 4:
      diagonals(odd);
      diagonals(even);
      if (type == 4) axes();
      break;
    case 3: {
      const double b = 1.0 / std::sqrt(phi * phi + 1.0), ab = phi * b;
      for (int k = 0; k < 4; k++) put(0, (k < 2 ? b : -b), (k % 2 ? -ab : ab));
      for (int k = 0; k < 4; k++) put((k < 2 ? b : -b), (k % 2 ? -ab : ab), 0);
      for (int k = 0; k < 4; k++) put((k % 2 ? -ab : ab), 0, (k < 2 ? b : -b));
      break;
    }
    case 5: {
      const double f = std::sqrt(phi * phi + 1.0 / (phi * phi));
      const double cc = phi / f, dd = 1.0 / phi / f;
      diagonals(odd);
      diagonals(even);
      for (int k = 0; k < 4; k++) put(0, (k < 2 ? dd : -dd), (k % 2 ? -cc : cc));
      for (int k = 0; k < 4; k++) put((k < 2 ? cc : -cc), 0, (k % 2 ? -dd : dd));
      for (int k = 0; k < 4; k++) put((k < 2 ? dd : -dd), (k % 2 ? -cc : cc), 0);
      break;
    }
    case 6:
    case 7: {
      // the prisms' angle step is the reference's 3.14159 * 2 / k
      const int k = type == 6 ? 6 : 5;
      const double step = 3.14159 * 2.0 / k;
      put(0, 0, 1);
      put(0, 0, -1);
      for (int i = 0; i < k; i++) put(std::cos(step * i), std::sin(step * i), 0);
      break;
    }
    default:
      diagonals(odd);
      break;
  }
  return n;
}

template <class T>
void fill_fractal(FractalK<T> &f, const mbr_scene &s) {
  std::memset(&f, 0, sizeof f);
  f.formula = s.formula;
  f.N = s.N;
  f.minN = s.minN;
  f.julia_mode = s.julia_mode;
  f.tglad_fold = s.tglad_folding_mode;
  f.sph_fold = s.spherical_folding_mode;
  f.mb_rot_on = s.mb_rotations_enabled;
  f.interior = s.interior_mode;
  f.iter_thresh = s.iter_thresh;
  f.limits = s.limits_enabled;
  f.only_plane = s.primitive_only_plane;
  f.prim_plane = s.plane_enabled;
  f.prim_box = s.box_enabled;
  f.prim_ibox = s.inv_box_enabled;
  f.prim_sphere = s.sphere_enabled;
  f.prim_isphere = s.inv_sphere_enabled;
  f.prim_water = s.water_enabled;
  f.water_iterations = s.water_iterations;
  f.frame_no = s.frame_no;
  f.power = (T)s.power;
  f.ipow = (s.power == std::floor(s.power) && s.power >= 2 && s.power <= 32) ? (int)s.power : 0;
  f.linear_de = s.linear_DE_mode;
  f.mb_sharp = (T)s.mb_sharpness;
  f.fip_factor = (T)s.folding_int_pow_factor;
  f.fip_zfactor = (T)s.folding_int_pow_z_factor;
  f.c_add = (T)s.c_add;
  {
    // RecalculateIFSParams: SetRotation(alfa, beta, gamma) = Z(alfa) Y(beta)
    // X(gamma) (src/algebra.cpp:153-159); CVector3::Normalize (x 1/length)
    Rot mr;
    mr.rz(s.ifs_rotation[0]);
    mr.ry(s.ifs_rotation[1]);
    mr.rx(s.ifs_rotation[2]);
    f.ifs_main_on = !mr.identity;
    for (int k = 0; k < 9; k++) f.ifs_main[k] = (T)mr.m[k];
    for (int i = 0; i < 9; i++) {
      Rot r;
      r.rz(s.ifs_angles[i][0]);
      r.ry(s.ifs_angles[i][1]);
      r.rx(s.ifs_angles[i][2]);
      f.ifs_rot_on[i] = !r.identity;
      for (int k = 0; k < 9; k++) f.ifs_rot[i][k] = (T)r.m[k];
      const double *dv = s.ifs_direction[i];
      const double n = 1.0 / std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
      for (int c = 0; c < 3; c++) f.ifs_dir[i][c] = (T)(dv[c] * n);
      f.ifs_dist[i] = (T)s.ifs_distance[i];
      f.ifs_int[i] = (T)s.ifs_intensity[i];
      f.ifs_en[i] = s.ifs_enabled[i];
    }
    for (int c = 0; c < 3; c++) {
      f.ifs_abs[c] = s.ifs_abs[c];
      f.ifs_offset[c] = (T)s.ifs_offset[c];
      f.ifs_edge[c] = (T)s.ifs_edge[c];
    }
    f.ifs_menger = s.ifs_menger_sponge_mode;
    f.ifs_fold = s.ifs_folding_mode;
    f.fl_min = s.fake_lights_min_iter;
    f.fl_max = s.fake_lights_max_iter;
    for (int c = 0; c < 3; c++) f.fl_trap[c] = (T)s.fake_lights_orbit_trap[c];
    int period = 0;
    for (int k = 0; k < 5; k++) {
      f.hyb_f[k] = s.hybrid_formula[k];
      f.hyb_it[k] = s.hybrid_iterations[k];
      f.hyb_p[k] = (T)s.hybrid_power[k];
      if (s.hybrid_formula[k] != 0 && (k < 4 || s.hybrid_cyclic)) period += s.hybrid_iterations[k];
    }
    f.hyb_cyclic = s.hybrid_cyclic;
    f.hyb_period = period;
    f.ifs_scale = (T)s.ifs_scale;
  }
  f.bailout = (T)s.bailout;
  f.constant_factor = (T)s.constant_factor;
  for (int i = 0; i < 3; i++) f.julia[i] = (T)s.julia[i];
  f.fold_limit = (T)s.folding_limit;
  f.fold_value = (T)s.folding_value;
  const double sf2 = s.spherical_folding_fixed * s.spherical_folding_fixed;
  const double sm2 = s.spherical_folding_min * s.spherical_folding_min;
  f.sph_fR2 = (T)sf2;
  f.sph_mR2 = (T)sm2;
  f.sph_factor1 = (T)(sf2 / sm2);
  f.mb_scale = (T)s.mb_scale;
  f.mb_abs_scale = (T)std::fabs(s.mb_scale);
  f.mb_limit = (T)s.mb_folding_limit;
  f.mb_value = (T)s.mb_folding_value;
  const double fR2 = s.mb_fixed_radius * s.mb_fixed_radius;
  const double mR2 = s.mb_min_radius * s.mb_min_radius;
  f.mb_fR2 = (T)fR2;
  f.mb_mR2 = (T)mR2;
  f.mb_factor1 = (T)(fR2 / mR2);
  for (int i = 0; i < 3; i++) f.mb_offset[i] = (T)s.mb_offset[i];
  f.mb_col_R = (T)s.mb_color_R;
  f.mb_col_X = (T)s.mb_color_X;
  f.mb_col_Y = (T)s.mb_color_Y;
  f.mb_col_Z = (T)s.mb_color_Z;
  f.mb_col_Sp1 = (T)s.mb_color_Sp1;
  f.mb_col_Sp2 = (T)s.mb_color_Sp2;
  {
    double nv[20][3];
    f.gf_sides = fold_polyhedron(s.mb_fold_mode, nv);
    for (int i = 0; i < f.gf_sides; i++)
      for (int c = 0; c < 3; c++) f.gf_nv[i][c] = (T)nv[i][c];
  }
  f.gf_solid = (T)s.mb_solid;
  f.gf_melt = (T)s.mb_melt;
  f.v4_scale_vary = (T)s.mb_vary_scale_vary;
  f.v4_fold = (T)s.mb_vary_fold;
  f.v4_minr = (T)s.mb_vary_minr;
  f.v4_rpower = (T)s.mb_vary_rpower;
  f.v4_wadd = (T)s.mb_vary_wadd;
  // RecalculateIFSParams, Mandelbox part (src/callbacks.cpp:1548-1555)
  Rot main = set_rotation3(s.mb_rotation_main);
  f.mb_main_rot = !main.identity;
  for (int i = 0; i < 9; i++) f.mb_main[i] = (T)main.m[i];
  for (int fold = 0; fold < 2; fold++)
    for (int axis = 0; axis < 3; axis++) {
      Rot r = set_rotation3(s.mb_rotation[fold][axis]);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          f.mb_rot[fold][axis][i * 3 + j] = (T)r.m[i * 3 + j];
          f.mb_rotinv[fold][axis][i * 3 + j] = (T)r.m[j * 3 + i];
        }
    }
  for (int i = 0; i < 3; i++) {
    f.limit_min[i] = (T)s.limit_min[i];
    f.limit_max[i] = (T)s.limit_max[i];
  }
  // PrimitivePlane normalises its normal on every call (src/primitives.cpp:17-18)
  {
    const double *n = s.plane_normal;
    double inv = 1.0 / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int i = 0; i < 3; i++) {
      f.plane_n[i] = (T)(n[i] * inv);
      f.plane_c[i] = (T)s.plane_centre[i];
    }
  }
  for (int i = 0; i < 3; i++) {
    f.box_c1[i] = (T)(s.box_centre[i] - 0.5 * s.box_size[i]);
    f.box_c2[i] = (T)(s.box_centre[i] + 0.5 * s.box_size[i]);
    f.ibox_c1[i] = (T)(s.inv_box_centre[i] - 0.5 * s.inv_box_size[i]);
    f.ibox_c2[i] = (T)(s.inv_box_centre[i] + 0.5 * s.inv_box_size[i]);
    f.sph_c[i] = (T)s.sphere_centre[i];
    f.isph_c[i] = (T)s.inv_sphere_centre[i];
  }
  f.sph_r = (T)s.sphere_radius;
  f.isph_r = (T)s.inv_sphere_radius;
  f.water_height = (T)s.water_height;
  f.water_amplitude = (T)s.water_amplitude;
  f.water_length = (T)s.water_length;
  Rot wr;
  wr.rz(s.water_rotation / 180 * kPi);
  f.water_rotated = !wr.identity;
  for (int i = 0; i < 9; i++) f.water_rot[i] = (T)wr.m[i];
  f.water_phase = (T)(s.water_anim_speed * s.frame_no);
}

template <class T>
void fill_shade(ShadeK<T> &k, const mbr_scene &s, const double light[3], double resolution) {
  std::memset(&k, 0, sizeof k);
  k.quality = (T)s.quality;
  k.DE_factor = (T)s.DE_factor;
  k.smoothness = (T)s.smoothness;
  k.resolution = (T)resolution;
  k.persp = (T)s.persp;
  k.view_distance_max = (T)s.view_distance_max;
  for (int i = 0; i < 3; i++) k.light[i] = (T)light[i];
  k.main_light[0] = (T)(s.main_light_intensity * s.main_light_colour.R / 65536.0);
  k.main_light[1] = (T)(s.main_light_intensity * s.main_light_colour.G / 65536.0);
  k.main_light[2] = (T)(s.main_light_intensity * s.main_light_colour.B / 65536.0);
  k.direct_light = (T)s.direct_light;
  k.shading = (T)s.shading;
  k.specular = (T)s.specular;
  k.global_ilum = (T)s.global_ilum;
  k.fast_ao_tune = (T)s.fast_ao_tune;
  k.soft_range = (T)std::tan(s.shadow_cone_angle / 180.0 * kPi);
  k.coloring_speed = (T)s.coloring_speed;
  k.palette_offset = (T)s.palette_offset;
  k.fog_visibility = (T)std::pow(10.0, s.fog_visibility / 10 - 16.0);
  k.fog_density = (T)s.fog_density;
  k.fog_c1_dist = (T)s.fog_colour1_distance;
  k.fog_c2_dist = (T)s.fog_colour2_distance;
  k.fog_reduce = (T)s.fog_distance_factor;
  k.glow_intensity = (T)s.glow_intensity;
  k.aux_intensity = (T)s.aux_light_intensity;
  k.aux_visibility = (T)s.aux_light_visibility;
}

void rgb3(int *o, const mbr_rgb &c) { o[0] = c.R; o[1] = c.G; o[2] = c.B; }

// AO direction table (src/Render3D.cpp:271-313); NULL lightmap = the white
// 100x100 fallback texture (src/texture.cpp:58-74).
int ao_table(int AO_quality, const mbr_texture *lm, std::vector<double> &dirs, std::vector<int> &rgb) {
  dirs.clear();
  rgb.clear();
  const int W = lm ? lm->width : 100, H = lm ? lm->height : 100;
  int counter = 0;
  for (double b = -0.49 * kPi; b < 0.49 * kPi; b += 1.0 / AO_quality) {
    for (double a = 0.0; a < 2.0 * kPi; a += ((2.0 / AO_quality) / std::cos(b))) {
      int X = (int)((a + b) / (2.0 * kPi) * W + W * 8.5) % W;
      int Y = (int)(b / (kPi)*H + H * 8.5) % H;
      int R = 255, G = 255, B = 255;
      if (lm) {
        const uint8_t *px = lm->rgb + 3 * ((size_t)X + (size_t)Y * W);
        R = px[0]; G = px[1]; B = px[2];
      }
      if (R > 10 || G > 10 || B > 10) {
        dirs.push_back(std::cos(a + b) * std::cos(b));
        dirs.push_back(std::sin(a + b) * std::cos(b));
        dirs.push_back(std::sin(b));
        dirs.push_back(0.0);
        rgb.push_back(R); rgb.push_back(G); rgb.push_back(B); rgb.push_back(0);
        counter++;
      }
      if (counter >= 10000) break;
    }
    if (counter >= 10000) break;
  }
  if (counter == 0) {
    counter = 1;
    dirs.assign(4, 0.0);
    rgb.assign(4, 0);
  }
  return counter;
}

}  // namespace

// Device buffers of one chain of band launches: the single-frame path uses
// slot 0; a batch runs up to kMaxSlots frame chains concurrently.
struct Slot {
  BandBufs band{};
  int prec = -1;
  // shading buffers (per hit; element type of the band's precision)
  void *sh_base = nullptr, *sh_spec = nullptr, *sh_col = nullptr, *sh_fast = nullptr, *sh_ao = nullptr;
  void *sh_aux = nullptr;
  size_t sh_ao_bytes = 0;
  // raytraced reflections: specular planes (every level), mirrored rays
  // (levels after the first)
  void *rf_sp = nullptr;
  double *rf_o = nullptr, *rf_d = nullptr, *rf_t0 = nullptr;
  uint8_t *rf_active = nullptr;
};

constexpr int kMaxSlots = 4;

struct mbr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool has_scene = false;
  mbr_scene scene;
  DevScene host_sc;
  DevScene *d_sc = nullptr;
  float *d_ao_f = nullptr;
  double *d_ao_d = nullptr;
  int *d_ao_rgb = nullptr;
  int ao_count = 0;
  // textures of mbr_set_textures (device copies, 8-bit RGB rows)
  uint8_t *d_bg = nullptr, *d_env = nullptr;
  int bg_w = 0, bg_h = 0, env_w = 0, env_h = 0;
  Slot slots[kMaxSlots];
  unsigned long long *d_stats = nullptr;
  int cu_count = 0;
  std::vector<hipEvent_t> ev;  // 6 per band of the last frame
  hipEvent_t split_fork = nullptr, split_join = nullptr;  // the two-stream frame's fork and join
  // mbr_cancel's flag: coherent pinned host memory (h_cancel) that the
  // kernels read through d_cancel
  int *h_cancel = nullptr;
  int *d_cancel = nullptr;
  // scratch planes of mbr_render (host-buffer entry point)
  void *d_planes = nullptr;
  size_t d_planes_bytes = 0;
  // batch: per-frame scene blocks and AO tables, streams, tuned launch shape
  DevScene *d_batch_sc = nullptr;
  int batch_sc_cap = 0;
  float *d_batch_ao_f = nullptr;
  double *d_batch_ao_d = nullptr;
  int *d_batch_ao_rgb = nullptr;
  size_t batch_ao_cap = 0;
  hipStream_t side[kMaxSlots] = {};
  double *d_grad = nullptr;  // mbr_adjoint accumulator
  int tuned_slots = 0, tuned_bpc = 0;
  // post effects: cos/sin table of the last SSAO quality, staging planes of
  // the host-memory calls
  double *d_ssao_tab = nullptr;
  int ssao_tab_q = 0;
  uint16_t *d_gamma = nullptr;  // gamma table of the last compile
  double gamma_of_table = 0.0;
  void *d_post = nullptr;
  size_t d_post_bytes = 0;
  // DOF: per-pixel buffers and the (tile, splat) lists
  void *d_dof = nullptr;
  size_t d_dof_bytes = 0;
  void *d_dof_pairs = nullptr;
  size_t d_dof_pairs_bytes = 0;
  // aux lights: the Lights array (placed by mbr_place_lights or given by
  // mbr_set_lights), the inputs of the automatic placement, and the device
  // list of the lights of the current scene
  std::vector<mbr_light> lights;
  bool lights_user = false, lights_placed = false;
  double lights_key[7] = {};
  double *d_lights_d = nullptr;
  float *d_lights_f = nullptr;
  int d_lights_cap = 0;
  // raytraced reflections: the slots of the levels after the first, and the
  // device copy of every level's band buffers
  std::vector<Slot> refl_slots;
  BandBufs *d_levels = nullptr;
  std::vector<BandBufs> levels_host;

  int fail(int rc, const std::string &msg) {
    err = msg;
    return rc;
  }
  int hip_fail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? MBR_ERR_OUT_OF_MEMORY : MBR_ERR_HIP;
  }
};

#define HIPCHECK(ctx, expr)                                 \
  do {                                                      \
    hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) return (ctx)->hip_fail(e_, #expr); \
  } while (0)

namespace {

void free_band(Slot &sl) {
  BandBufs &b = sl.band;
  void *ptrs[] = {b.found, b.object, b.count, b.depth, b.thr, b.last, b.px, b.py, b.pz, b.rec_step, b.rec_dist,
                  b.rec_thr, b.rec_it, b.ovf_it, b.ovf_off, b.hit_list, b.redo_list, b.counters, b.ovf_dir,
                  b.ovf_step, b.ovf_dist, b.ovf_thr,
                  b.hand_ray, b.hand_i, b.hand_scan, b.hand_step, b.hand_drift, b.hand_dist,
                  b.sh_r, b.sh_g, b.sh_b, b.col_r, b.col_g, b.col_b};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  b = BandBufs{};
  void *sp[] = {sl.sh_base, sl.sh_spec, sl.sh_col, sl.sh_fast, sl.sh_ao, sl.sh_aux};
  for (void *p : sp)
    if (p) (void)hipFree(p);
  sl.sh_base = sl.sh_spec = sl.sh_col = sl.sh_fast = sl.sh_ao = sl.sh_aux = nullptr;
  void *rp[] = {sl.rf_sp, sl.rf_o, sl.rf_d, sl.rf_t0, sl.rf_active};
  for (void *p : rp)
    if (p) (void)hipFree(p);
  sl.rf_sp = nullptr;
  sl.rf_o = sl.rf_d = sl.rf_t0 = nullptr;
  sl.rf_active = nullptr;
  sl.sh_ao_bytes = 0;
  sl.prec = -1;
}

// Overflow pool of `dirs` directories and `chunks` chunks (freeing the
// previous one).
int alloc_pool(mbr_ctx *c, Slot &sl, long long dirs, long long chunks) {
  BandBufs &b = sl.band;
  void *old[] = {b.ovf_dir, b.ovf_step, b.ovf_dist, b.ovf_thr, b.ovf_it};
  for (void *p : old)
    if (p) (void)hipFree(p);
  b.ovf_dir = nullptr;
  b.ovf_step = b.ovf_dist = b.ovf_thr = nullptr;
  b.ovf_it = nullptr;
  b.ovf_dirs = b.ovf_chunks = 0;
  const size_t bytes = (size_t)chunks * kChunk * b.rec_bytes;
  HIPCHECK(c, hipMalloc((void **)&b.ovf_dir, (size_t)dirs * kDirSlots * sizeof(int)));
  HIPCHECK(c, hipMalloc(&b.ovf_step, bytes));
  HIPCHECK(c, hipMalloc(&b.ovf_dist, bytes));
  HIPCHECK(c, hipMalloc(&b.ovf_thr, bytes));
  if (b.rec_it) HIPCHECK(c, hipMalloc((void **)&b.ovf_it, (size_t)chunks * kChunk * sizeof(int)));
  b.ovf_dirs = dirs;
  b.ovf_chunks = chunks;
  return MBR_OK;
}

// Grow the pool after a render that found it full and asked for `dirs`
// directories and `chunks` chunks: at least double it, and at least 1.25 x
// what was asked for (rays that found no directory asked for no chunk, so a
// second render may ask for more).
int grow_pool(mbr_ctx *c, Slot &sl, unsigned long long dirs, unsigned long long chunks) {
  const BandBufs &b = sl.band;
  return alloc_pool(c, sl, std::max<long long>(2 * b.ovf_dirs, (long long)(dirs + dirs / 4 + 64)),
                    std::max<long long>(2 * b.ovf_chunks, (long long)(chunks + chunks / 4 + 64)));
}

constexpr int kPoolAttempts = 8;  // renders of a frame while its overflow pool grows

// Band buffers for `cap` rays and AO intensities for `ao_count` directions
// per hit; kept while precision and cap stay the same.
int alloc_slot(mbr_ctx *c, Slot &sl, int precision, int cap, int ao_count, bool iters = false) {
  if (sl.prec != precision || sl.band.cap != cap) {
    free_band(sl);
    BandBufs &b = sl.band;
    b.cap = cap;
    b.slab = kSlab;
    b.rec_bytes = precision == MBR_PREC_FP64 ? 8 : 4;
    const size_t n = (size_t)b.cap;
    HIPCHECK(c, hipMalloc((void **)&b.found, n));
    HIPCHECK(c, hipMalloc((void **)&b.object, n));
    HIPCHECK(c, hipMalloc((void **)&b.count, n * sizeof(int)));
    HIPCHECK(c, hipMalloc((void **)&b.depth, n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&b.thr, n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&b.last, n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&b.px, n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&b.py, n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&b.pz, n * sizeof(double)));
    const size_t rec = n * (size_t)b.slab * b.rec_bytes;
    HIPCHECK(c, hipMalloc(&b.rec_step, rec));
    HIPCHECK(c, hipMalloc(&b.rec_dist, rec));
    HIPCHECK(c, hipMalloc(&b.rec_thr, rec));
    HIPCHECK(c, hipMalloc((void **)&b.ovf_off, n * sizeof(long long)));
    HIPCHECK(c, hipMalloc((void **)&b.hit_list, n * sizeof(int)));
    HIPCHECK(c, hipMalloc((void **)&b.redo_list, n * sizeof(int)));
    HIPCHECK(c, hipMalloc((void **)&b.counters, CNT_COUNT * sizeof(unsigned long long)));
    HIPCHECK(c, hipMalloc((void **)&b.hand_ray, n * sizeof(int)));
    HIPCHECK(c, hipMalloc((void **)&b.hand_i, n * sizeof(int)));
    HIPCHECK(c, hipMalloc((void **)&b.hand_scan, n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&b.hand_step, n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&b.hand_drift, n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&b.hand_dist, n * sizeof(double)));
    HIPCHECK(c, hipMalloc(&b.sh_r, n * b.rec_bytes));
    HIPCHECK(c, hipMalloc(&b.sh_g, n * b.rec_bytes));
    HIPCHECK(c, hipMalloc(&b.sh_b, n * b.rec_bytes));
    HIPCHECK(c, hipMalloc((void **)&b.col_r, n));
    HIPCHECK(c, hipMalloc((void **)&b.col_g, n));
    HIPCHECK(c, hipMalloc((void **)&b.col_b, n));
    HIPCHECK(c, hipMalloc(&sl.sh_base, 3 * n * b.rec_bytes));
    HIPCHECK(c, hipMalloc(&sl.sh_spec, 3 * n * b.rec_bytes));
    HIPCHECK(c, hipMalloc(&sl.sh_col, 3 * n * b.rec_bytes));
    HIPCHECK(c, hipMalloc(&sl.sh_fast, n * b.rec_bytes));
    // aux lights: normal and detail, then diffuse and specular sums (6);
    // environment mapping: its reflect-scaled colour (3)
    HIPCHECK(c, hipMalloc(&sl.sh_aux, 9 * n * b.rec_bytes));
    // one directory and one chunk per 64 rays to begin with; grown when a
    // frame needs more
    int rc = alloc_pool(c, sl, std::max<long long>(256, (long long)n / 64), std::max<long long>(256, (long long)n / 64));
    if (rc != MBR_OK) return rc;
    sl.prec = precision;
  }
  const size_t need = (size_t)sl.band.cap * (size_t)std::max(1, ao_count) * sl.band.rec_bytes;
  if (ao_count > 0 && need > sl.sh_ao_bytes) {
    if (sl.sh_ao) (void)hipFree(sl.sh_ao);
    sl.sh_ao = nullptr;
    sl.sh_ao_bytes = 0;
    HIPCHECK(c, hipMalloc(&sl.sh_ao, need));
    sl.sh_ao_bytes = need;
  }
  // iteration counts of the step records (iteration fog), kept once made
  BandBufs &b = sl.band;
  if (iters && !b.rec_it) {
    HIPCHECK(c, hipMalloc((void **)&b.rec_it, (size_t)b.cap * (size_t)b.slab * sizeof(int)));
    HIPCHECK(c, hipMalloc((void **)&b.ovf_it, (size_t)b.ovf_chunks * kChunk * sizeof(int)));
  }
  b.cancel = c->d_cancel;
  return MBR_OK;
}

// Rays per band: every ray of a band may hit, and the AO pass keeps one
// intensity per (hit, direction).
int band_cap(int precision, int ao_count) {
  if (ao_count <= 0) return kBandRays;
  const size_t per_ray = (size_t)ao_count * (precision == MBR_PREC_FP64 ? 8 : 4);
  return (int)std::min<size_t>((size_t)kBandRays, std::max<size_t>(4096, kAOBudget / per_ray));
}

int kernel_family(const mbr_scene &s) { return s.ifs_folding_mode ? kGenericFamily : s.formula; }

// The precision a scene renders in.  The generic formulas run in FP64,
// except the analytic ones (trig_DE, kaleidoscopic IFS, smoothMandelbox,
// mandelboxVaryScale4D, generalizedFoldBox, and any analytic formula under
// the IFS folding pre-fold) without aux lights, volumetric lights or
// iteration fog: their FP32 modes march in FP64 and shade in FP32
// (FP32_MARCH runs as FP32_FIXUP for them).
// h: the scene's derived block, when made (batches reject those features).
int scene_precision(const mbr_scene &s, const DevScene *h, int precision) {
  if (precision == MBR_PREC_FP64) return precision;
  // fake lights and environment mapping render in FP64
  if (s.fake_lights_enabled || (s.reflect > 0.0 && !s.raytraced_reflections)) return MBR_PREC_FP64;
  const int fam = kernel_family(s);
  if (!formula_generic(fam)) return precision;
  // the IFS folding pre-fold keeps the FP32 mode of an analytic formula
  const int f = fam == kGenericFamily ? s.formula : fam;
  const bool analytic = f == 1 || f == 2 || f == 7 || f == 8 || f == 10 || f == 17 || f == 18 || f == 37;
  if (!analytic || (h && (h->n_lights > 0 || h->shadow_fx))) return MBR_PREC_FP64;
  return MBR_PREC_FP32_FIXUP;
}

int check_supported(mbr_ctx *c, const mbr_scene &s) {
  if (s.abi_version != MBR_ABI_VERSION) return c->fail(MBR_ERR_INVALID_ARG, "scene abi_version mismatch");
  if (s.formula == MBR_FORMULA_HYBRID) {
    int period = 0;
    for (int k = 0; k < 5; k++) {
      const int hf = s.hybrid_formula[k];
      if (hf != 0 && (hf == MBR_FORMULA_HYBRID || !formula_supported(hf)))
        return c->fail(MBR_ERR_UNSUPPORTED, "hybrid slot " + std::to_string(k + 1) + ": formula " + std::to_string(hf) +
                                                " is not in the render core");
      if (hf != 0 && s.hybrid_iterations[k] < 0) return c->fail(MBR_ERR_INVALID_ARG, "negative hybrid iterations");
      if (hf != 0 && (k < 4 || s.hybrid_cyclic)) period += s.hybrid_iterations[k];
    }
    // a cyclic sequence of no iterations never ends in the reference
    if (s.hybrid_cyclic && period == 0) return c->fail(MBR_ERR_INVALID_ARG, "cyclic hybrid of zero iterations");
  }
  if (!formula_supported(s.formula))
    return c->fail(MBR_ERR_UNSUPPORTED, "formula " + std::to_string(s.formula) +
                                            " is not in the render core (supported: 1-8, 10-37)");
  if (s.reflect > 0.0 && !s.raytraced_reflections && !c->d_env)
    return c->fail(MBR_ERR_UNSUPPORTED, "environment mapping (reflect > 0): no envmap texture (mbr_set_textures)");
  if (s.raytraced_reflections && s.reflections_max > MBR_MAX_REFLECTIONS)
    return c->fail(MBR_ERR_INVALID_ARG, "reflections_max above " + std::to_string(MBR_MAX_REFLECTIONS));
  if (s.aux_light_number < 0) return c->fail(MBR_ERR_INVALID_ARG, "negative aux_light_number");
  if (s.textured_background && !c->d_bg)
    return c->fail(MBR_ERR_UNSUPPORTED, "textured background: no background texture (mbr_set_textures)");
  if (s.image_width <= 0 || s.image_height <= 0 || s.N <= 0 || s.N > 100000)
    return c->fail(MBR_ERR_INVALID_ARG, "bad image size or N");
  return MBR_OK;
}

}  // namespace

extern "C" {

int mbr_version(void) { return MBR_ABI_VERSION; }

size_t mbr_scene_layout(size_t offsets[8]) {
  if (offsets) {
    offsets[0] = offsetof(mbr_scene, abi_version);
    offsets[1] = offsetof(mbr_scene, vp);
    offsets[2] = offsetof(mbr_scene, formula);
    offsets[3] = offsetof(mbr_scene, mb_rotation);
    offsets[4] = offsetof(mbr_scene, palette);
    offsets[5] = offsetof(mbr_scene, fog_colour3);
    offsets[6] = offsetof(mbr_scene, reflections_max);
    offsets[7] = offsetof(mbr_scene, jitter_seed);
  }
  return sizeof(mbr_scene);
}

uint32_t mbr_jitter_draw(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t step) {
  return mbr::jitter_draw(seed, frame, pixel, step);
}

int mbr_init(const mbr_device_opts *opts, mbr_ctx **out) {
  if (!out) return MBR_ERR_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MBR_ERR_NO_DEVICE;
  mbr_ctx *c = new mbr_ctx();
  c->device = opts ? opts->device : 0;
  if (c->device < 0 || c->device >= ndev) {
    delete c;
    return MBR_ERR_NO_DEVICE;
  }
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void **)&c->d_sc, sizeof(DevScene)) != hipSuccess ||
      hipMalloc((void **)&c->d_stats, ST_COUNT * sizeof(unsigned long long)) != hipSuccess) {
    delete c;
    return MBR_ERR_HIP;
  }
  if (hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess)
    c->cu_count = 0;
  if (hipHostMalloc((void **)&c->h_cancel, sizeof(int), hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void **)&c->d_cancel, c->h_cancel, 0) != hipSuccess) {
    mbr_destroy(c);
    return MBR_ERR_HIP;
  }
  __atomic_store_n(c->h_cancel, 0, __ATOMIC_SEQ_CST);
  if (c->device < kMaxDevices) {
    // the device address of this GPU's copy of `scratch` (the device is
    // current), once per device
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
      void *p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        mbr_destroy(c);
        return MBR_ERR_HIP;
      }
      g_scratch_ptr[c->device] = p;
    }
  }
  *out = c;
  return MBR_OK;
}

void mbr_destroy(mbr_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (Slot &sl : c->slots) free_band(sl);
  for (hipStream_t st : c->side)
    if (st) (void)hipStreamDestroy(st);
  void *bp[] = {c->d_batch_sc, c->d_batch_ao_f, c->d_batch_ao_d, c->d_batch_ao_rgb, c->d_grad, c->d_ssao_tab,
                c->d_post, c->d_dof, c->d_dof_pairs, c->d_gamma, c->d_lights_d, c->d_lights_f, c->d_bg, c->d_env};
  for (void *p : bp)
    if (p) (void)hipFree(p);
  for (Slot &sl : c->refl_slots) free_band(sl);
  if (c->d_levels) (void)hipFree(c->d_levels);
  void *ptrs[] = {c->d_sc, c->d_ao_f, c->d_ao_d, c->d_ao_rgb, c->d_stats, c->d_planes};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->h_cancel) (void)hipHostFree(c->h_cancel);
  if (c->split_fork) (void)hipEventDestroy(c->split_fork);
  if (c->split_join) (void)hipEventDestroy(c->split_join);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *mbr_last_error(const mbr_ctx *c) { return c ? c->err.c_str() : "null context"; }

int mbr_cancel(mbr_ctx *c) {
  if (!c || !c->h_cancel) return MBR_ERR_INVALID_ARG;
  __atomic_store_n(c->h_cancel, 1, __ATOMIC_SEQ_CST);
  return MBR_OK;
}

}  // extern "C"

namespace {

// A scene after CheckPrameters and derivation: the device scene block and
// the AO direction table (directions as x, y, z, 0; lightmap colours).
struct Prepared {
  mbr_scene s;
  DevScene d;
  std::vector<double> dirs;
  std::vector<int> rgb;
};

int prepare_scene(mbr_ctx *c, const mbr_scene *scene_in, const mbr_texture *lightmap, Prepared &p) {
  mbr_scene &s = p.s;
  s = *scene_in;
  int rc = check_supported(c, s);
  if (rc != MBR_OK) return rc;
  // CheckPrameters (src/interface.cpp:3587-3602)
  if (s.zoom <= 1e-15) s.zoom = 1e-15;
  if (s.DE_factor <= 0.0001) s.DE_factor = 0.0001;
  if (s.smoothness <= 0.00001) s.smoothness = 0.00001;
  if (s.persp <= 0.001) s.persp = 0.001;
  if (s.ao_quality >= 30) s.ao_quality = 30;
  if (s.ao_quality <= 1) s.ao_quality = 1;
  if (s.tiles < 1) s.tiles = 1;
  if (s.tiles > 100) s.tiles = 100;
  if (s.fog_density < 0.0) s.fog_density = 0.0;
  if (lightmap && (!lightmap->rgb || lightmap->width <= 0 || lightmap->height <= 0))
    return c->fail(MBR_ERR_INVALID_ARG, "bad lightmap texture");

  DevScene &d = p.d;
  std::memset(&d, 0, sizeof d);
  d.width = s.image_width;
  d.height = s.image_height;
  d.tiles = s.tiles;
  d.tile_index = s.tile_index;
  d.persp_type = s.perspective_type;
  d.fish_eye_cut = s.fish_eye_cut;
  d.frame_no = s.frame_no;
  d.jitter_mode = s.jitter_mode;
  d.jitter_seed = s.jitter_seed;
  d.hit_refine = s.hit_refine;
  d.pixel_base = (uint32_t)((long long)s.tile_index * s.image_width * s.image_height);
  d.aspect = (double)s.image_width / s.image_height;
  d.fov = s.persp;
  d.tile_x = (double)(s.tile_index % s.tiles);
  d.tile_y = (double)(s.tile_index / s.tiles);
  // camera (src/Render3D.cpp:236-256, :325-333)
  Rot R;
  R.rz(s.alpha);
  R.rx(s.beta);
  R.ry(s.gamma);
  std::memcpy(d.R, R.m, sizeof d.R);
  d.R_identity = R.identity;
  const double baseY0[3] = {0.0, 1.0, 0.0};
  double baseY[3];
  R.apply(baseY0, baseY);
  const double lv[3] = {std::cos(s.main_light_alpha - 0.5 * kPi) * std::cos(-s.main_light_beta),
                        std::sin(s.main_light_alpha - 0.5 * kPi) * std::cos(-s.main_light_beta),
                        std::sin(-s.main_light_beta)};
  double light[3];
  R.apply(lv, light);
  if (s.perspective_type == MBR_PERSP_FISH_EYE || s.perspective_type == MBR_PERSP_EQUIRECTANGULAR) {
    for (int i = 0; i < 3; i++) d.start[i] = s.vp[i];
  } else {
    const double f = 1.0 / s.persp * s.zoom;
    for (int i = 0; i < 3; i++) d.start[i] = s.vp[i] - baseY[i] * f;
  }
  d.view_distance_min = s.view_distance_min;
  d.view_distance_max = s.view_distance_max;
  const double resolution = 1.0 / s.image_width / s.tiles;  // src/interface.cpp:695

  d.constant_thresh = s.constant_DE_threshold;
  d.interior = s.interior_mode;
  d.shadow = s.shadow;
  d.penetrating = s.penetrating_lights;
  d.ao = s.global_illumination;
  d.fast_ao = s.fast_global_illumination;
  d.ao_quality = s.ao_quality;
  d.slow_shading = s.slow_shading;
  d.coloring = s.coloring_enabled;
  d.bsoft = (!s.iter_fog_enabled && !s.limits_enabled && !s.iter_thresh) &&
            std::tan(s.shadow_cone_angle / 180.0 * kPi) > 0.0;
  // volumetric lights (src/interface.cpp:697-702) and iteration fog change
  // the shadow marches; the aux lights' volumetric switches are resolved
  // against the light table in setup_lights
  d.vol_main = s.volumetric_light_enabled[0];
  d.iter_fog = s.iter_fog_enabled;
  d.iter_fog_opacity = s.iter_fog_opacity;
  d.iter_fog_trim = s.iter_fog_opacity_trim;
  d.main_light_intensity = s.main_light_intensity;
  d.refl_max = s.raytraced_reflections ? std::max(0, s.reflections_max) : 0;
  {
    const double rf[7] = {s.reflect, s.plane_reflect, s.water_reflect, s.sphere_reflect, s.inv_sphere_reflect,
                          s.box_reflect, s.inv_box_reflect};  // OBJ_* order
    std::memcpy(d.refl, rf, sizeof rf);
  }
  d.shadow_fx = s.iter_fog_enabled;
  for (int i = 0; i < 5; i++) {
    if (s.volumetric_light_enabled[i]) d.shadow_fx = 1;
    d.vol_intensity[i] = s.volumetric_light_intensity[i];
  }
  d.post_fog = s.post_fog_enabled;
  d.textured_bg = s.textured_background;
  d.fulldome = s.background_as_fulldome;
  d.bg_tex = c->d_bg;
  d.bg_w = c->bg_w;
  d.bg_h = c->bg_h;
  d.env_map = s.reflect > 0.0 && !s.raytraced_reflections && c->d_env;
  d.env_tex = c->d_env;
  d.env_w = c->env_w;
  d.env_h = c->env_h;
  d.reflect = s.reflect;
  d.fake_lights = s.fake_lights_enabled;
  d.fl_intensity = s.fake_lights_intensity;
  d.fl_visibility = s.fake_lights_visibility;
  d.fl_visibility_size = s.fake_lights_visibility_size;
  d.N = s.N;
  d.minN = s.minN;
  rgb3(d.main_light_colour, s.main_light_colour);
  rgb3(d.glow1, s.glow_color1);
  rgb3(d.glow2, s.glow_color2);
  rgb3(d.bg1, s.background_color1);
  rgb3(d.bg2, s.background_color2);
  rgb3(d.bg3, s.background_color3);
  rgb3(d.post_fog_colour, s.post_fog_color);
  rgb3(d.fog1, s.fog_colour1);
  rgb3(d.fog2, s.fog_colour2);
  rgb3(d.fog3, s.fog_colour3);
  rgb3(d.plane_col, s.plane_colour);
  rgb3(d.box_col, s.box_colour);
  rgb3(d.ibox_col, s.inv_box_colour);
  rgb3(d.sphere_col, s.sphere_colour);
  rgb3(d.isphere_col, s.inv_sphere_colour);
  rgb3(d.water_col, s.water_colour);
  for (int i = 0; i < 256; i++) rgb3(d.palette[i], s.palette[i]);
  fill_fractal(d.fd, s);
  fill_fractal(d.ff, s);
  fill_shade(d.sd, s, light, resolution);
  fill_shade(d.sf, s, light, resolution);

  d.ao_count = ao_table(s.ao_quality, lightmap, p.dirs, p.rgb);
  // Deep zoom: every hit lies at least the distance estimate at the ray
  // start away, so its threshold is at least thr_min; when that is below
  // what FP32 resolves at the hit's coordinates (|start| + a few units of
  // march), FP32 shading must start its estimates in FP64.
  {
    const V3<double> st = v3<double>(d.start[0], d.start[1], d.start[2]);
    DEInfo info;
    const double de0 = calc_distance<double>(d.fd, st, 0.0, s.N, s.minN, false, info);
    const double near = std::max(de0, s.view_distance_min);
    const double thr_min = s.constant_DE_threshold ? s.quality : near * resolution * s.persp / s.quality;
    d.deep = thr_min < 2.0 * kMixedSwitch * (length(st) + 3.0);
  }
  // Plain scenes run kernel variants compiled without the branches (and the
  // live scene constants) of the optional distance-estimate terms.
  d.plain = d.fd.limits == 0 && d.fd.only_plane == 0 && d.fd.interior == 0 && d.fd.iter_thresh == 0 &&
            !d.fd.prim_plane && !d.fd.prim_box && !d.fd.prim_ibox && !d.fd.prim_sphere && !d.fd.prim_isphere &&
            !d.fd.prim_water && !d.fd.tglad_fold && !d.fd.sph_fold && !d.fd.julia_mode && !d.fd.ifs_fold;
  return MBR_OK;
}

bool scene_has_lights(const mbr_scene &s) {
  if (s.aux_light_number > 0) return true;
  for (int i = 0; i < 4; i++)
    if (s.aux_light_pre_enabled[i]) return true;
  return false;
}

constexpr long long kMaxLightDraws = 20000000;  // placement trials before giving up

// PlaceRandomLights(params, false) (src/shaders.cpp:1134-1212) on the host,
// in FP64, with the generic distance estimator (the reference's operation
// order, as in the oracle) and a private copy of the
// C library's rand() sequence.  The reference draws the two terms of each
// coordinate left to right and the three colour channels in order; an
// accepted light gets its colour normalised to a maximum channel of 65536.
int place_lights(const mbr_scene &s, std::vector<mbr_light> &out) {
  const int n = std::max(0, s.aux_light_number);
  out.assign((size_t)n + 4, mbr_light{});
  FractalK<double> fd;
  std::memset(&fd, 0, sizeof fd);
  fill_fractal(fd, s);
  GlibcRand rnd((uint32_t)s.aux_light_random_seed);
  auto random = [&](int max) { return rnd.next() % (max + 1); };  // Random, src/common_math.cpp:25-32
  int trial_number = 0;
  double radius_multiplier = 1.0;
  long long draws = 0;
  for (int i = 0; i < n; i++) {
    trial_number++;
    if (++draws > kMaxLightDraws) return MBR_ERR_INVALID_ARG;
    double pos[3];
    for (int k = 0; k < 3; k++) {
      const double a = random(2000000) / 1000000.0 - 1.0;
      const double b = random(1000000) / 1.0e12;
      pos[k] = s.aux_light_random_center[k] + (a + b) * s.aux_light_distribution_radius * radius_multiplier;
    }
    DEInfo info;
    const double distance =
        calc_distance_generic<double>(fd, v3<double>(pos[0], pos[1], pos[2]), 0.0, s.N, s.minN, false, info);
    if (trial_number > 1000) {
      radius_multiplier *= 1.1;
      trial_number = 0;
    }
    if (distance > 0 && !info.max_iter && distance < s.aux_light_max_dist * radius_multiplier) {
      radius_multiplier = 1.0;
      mbr_light &L = out[(size_t)i];
      for (int k = 0; k < 3; k++) L.position[k] = pos[k];
      int col[3];
      for (int k = 0; k < 3; k++) col[k] = 20000 + random(80000);
      const double maxColour = std::max(std::max((double)col[0], (double)col[1]), (double)col[2]);
      for (int k = 0; k < 3; k++) L.colour[k] = (int)(col[k] * 65536.0 / maxColour);
      L.intensity = distance * distance / s.aux_light_max_dist;
      L.enabled = 1;
    } else {
      i--;
    }
  }
  return MBR_OK;
}

// PlaceRandomLights(params, true): the scene's predefined lights over slots
// 0-3, the others of those slots marked disabled.
void apply_predefined(const mbr_scene &s, std::vector<mbr_light> &t) {
  for (int i = 0; i < 4 && i < (int)t.size(); i++) {
    mbr_light &L = t[(size_t)i];
    if (s.aux_light_pre_enabled[i]) {
      for (int k = 0; k < 3; k++) L.position[k] = s.aux_light_pre[i][k];
      L.intensity = s.aux_light_pre_intensity[i];
      L.colour[0] = s.aux_light_pre_colour[i].R;
      L.colour[1] = s.aux_light_pre_colour[i].G;
      L.colour[2] = s.aux_light_pre_colour[i].B;
      L.enabled = 1;
    } else {
      L.enabled = 0;
    }
  }
}

// The lights of scene s in the device scene block d: place them when the
// context has no placement for these inputs (MainRender places them once,
// src/Render3D.cpp:1990-1993), write the predefined ones over slots 0-3
// (:2305), and upload the list of the lights that take part.
int setup_lights(mbr_ctx *c, const mbr_scene &s, DevScene &d) {
  d.n_lights = 0;
  d.light_div = 0;
  d.lights_d = nullptr;
  d.lights_f = nullptr;
  if (!scene_has_lights(s)) return MBR_OK;
  const double key[7] = {(double)s.aux_light_random_seed, (double)s.aux_light_number, s.aux_light_distribution_radius,
                         s.aux_light_max_dist, s.aux_light_random_center[0], s.aux_light_random_center[1],
                         s.aux_light_random_center[2]};
  if (!c->lights_user && (!c->lights_placed || std::memcmp(key, c->lights_key, sizeof key) != 0)) {
    std::vector<mbr_light> t;
    if (place_lights(s, t) != MBR_OK)
      return c->fail(MBR_ERR_INVALID_ARG, "aux light placement: no position accepted in 20 million draws");
    c->lights.swap(t);
    std::memcpy(c->lights_key, key, sizeof key);
    c->lights_placed = true;
  }
  const int number = s.aux_light_number;
  const int count = std::max(number, 4);  // lightsPlaced
  if ((int)c->lights.size() < count)
    return c->fail(MBR_ERR_INVALID_ARG, "light table has " + std::to_string(c->lights.size()) +
                                            " entries, the scene needs " + std::to_string(count));
  apply_predefined(s, c->lights);
  // volumetric aux lights: slots 0-3 when enabled (src/shaders.cpp:326-328)
  for (int k = 0; k < 4; k++) {
    const mbr_light &L = c->lights[(size_t)k];
    d.vol_aux_on[k] = s.volumetric_light_enabled[k + 1] && L.enabled;
    d.aux_slot_on[k] = L.enabled;
    const double v[8] = {L.position[0], L.position[1], L.position[2], L.intensity,
                         (double)L.colour[0], (double)L.colour[1], (double)L.colour[2], 0.0};
    std::memcpy(d.vol_aux[k], v, sizeof v);
  }
  std::vector<double> ld;
  for (int i = 0; i < count; i++) {
    const mbr_light &L = c->lights[(size_t)i];
    if (!(i < number || L.enabled)) continue;
    const double v[8] = {L.position[0], L.position[1], L.position[2], L.intensity,
                         (double)L.colour[0], (double)L.colour[1], (double)L.colour[2], 0.0};
    ld.insert(ld.end(), v, v + 8);
  }
  const int n_used = (int)(ld.size() / 8);
  d.light_div = count;
  if (n_used == 0) return MBR_OK;
  // renders of the previous scene may still read the table
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (n_used > c->d_lights_cap) {
    if (c->d_lights_d) (void)hipFree(c->d_lights_d);
    if (c->d_lights_f) (void)hipFree(c->d_lights_f);
    c->d_lights_d = nullptr;
    c->d_lights_f = nullptr;
    c->d_lights_cap = 0;
    HIPCHECK(c, hipMalloc((void **)&c->d_lights_d, (size_t)n_used * 8 * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&c->d_lights_f, (size_t)n_used * 8 * sizeof(float)));
    c->d_lights_cap = n_used;
  }
  std::vector<float> lf(ld.begin(), ld.end());
  HIPCHECK(c, hipMemcpy(c->d_lights_d, ld.data(), ld.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(c->d_lights_f, lf.data(), lf.size() * sizeof(float), hipMemcpyHostToDevice));
  d.lights_d = c->d_lights_d;
  d.lights_f = c->d_lights_f;
  d.n_lights = n_used;
  return MBR_OK;
}

}  // namespace

extern "C" {

int mbr_place_lights(const mbr_scene *scene, mbr_light *lights, int capacity, int *count) {
  if (!scene || !count) return MBR_ERR_INVALID_ARG;
  mbr_ctx tmp;
  int rc = check_supported(&tmp, *scene);
  if (rc != MBR_OK) return rc;
  const int n = std::max(0, scene->aux_light_number) + 4;
  *count = n;
  if (!lights || capacity < n) return MBR_ERR_INVALID_ARG;
  std::vector<mbr_light> t;
  rc = place_lights(*scene, t);
  if (rc != MBR_OK) return rc;
  apply_predefined(*scene, t);
  std::memcpy(lights, t.data(), t.size() * sizeof(mbr_light));
  return MBR_OK;
}

int mbr_set_lights(mbr_ctx *c, const mbr_light *lights, int count) {
  if (!c || count < 0 || (count > 0 && !lights)) return MBR_ERR_INVALID_ARG;
  if (!lights || count == 0) {
    c->lights.clear();
    c->lights_user = false;
    c->lights_placed = false;
    return MBR_OK;
  }
  c->lights.assign(lights, lights + count);
  c->lights_user = true;
  return MBR_OK;
}

int mbr_get_lights(mbr_ctx *c, mbr_light *lights, int capacity, int *count) {
  if (!c || !count) return MBR_ERR_INVALID_ARG;
  *count = (int)c->lights.size();
  if (!lights) return MBR_OK;
  if (capacity < *count) return c->fail(MBR_ERR_INVALID_ARG, "capacity below the light count");
  std::memcpy(lights, c->lights.data(), c->lights.size() * sizeof(mbr_light));
  return MBR_OK;
}

int mbr_set_textures(mbr_ctx *c, const mbr_texture *background, const mbr_texture *envmap) {
  if (!c) return MBR_ERR_INVALID_ARG;
  for (const mbr_texture *t : {background, envmap})
    if (t && (!t->rgb || t->width <= 0 || t->height <= 0 || (long long)t->width * t->height > (1ll << 28)))
      return c->fail(MBR_ERR_INVALID_ARG, "bad texture");
  HIPCHECK(c, hipSetDevice(c->device));
  // the previous frame may still read the old copies
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  auto put = [&](const mbr_texture *t, uint8_t *&d, int &w, int &h) -> int {
    if (d) (void)hipFree(d);
    d = nullptr;
    w = h = 0;
    if (!t) return MBR_OK;
    const size_t bytes = (size_t)t->width * t->height * 3;
    HIPCHECK(c, hipMalloc((void **)&d, bytes));
    HIPCHECK(c, hipMemcpy(d, t->rgb, bytes, hipMemcpyHostToDevice));
    w = t->width;
    h = t->height;
    return MBR_OK;
  };
  int rc = put(background, c->d_bg, c->bg_w, c->bg_h);
  if (rc == MBR_OK) rc = put(envmap, c->d_env, c->env_w, c->env_h);
  return rc;
}

int mbr_set_scene(mbr_ctx *c, const mbr_scene *scene_in, const mbr_texture *lightmap) {
  if (!c || !scene_in) return MBR_ERR_INVALID_ARG;
  Prepared p;
  int rc = prepare_scene(c, scene_in, lightmap, p);
  if (rc != MBR_OK) return rc;
  HIPCHECK(c, hipSetDevice(c->device));
  rc = setup_lights(c, p.s, p.d);
  if (rc != MBR_OK) return rc;
  p.d.vol_any = p.d.vol_main || p.d.vol_aux_on[0] || p.d.vol_aux_on[1] || p.d.vol_aux_on[2] || p.d.vol_aux_on[3];
  const DevScene &d = p.d;
  const std::vector<double> &dirs = p.dirs;
  const std::vector<int> &rgb = p.rgb;
  const int count = d.ao_count;
  HIPCHECK(c, hipSetDevice(c->device));
  if (count > c->ao_count) {
    if (c->d_ao_f) (void)hipFree(c->d_ao_f);
    if (c->d_ao_d) (void)hipFree(c->d_ao_d);
    if (c->d_ao_rgb) (void)hipFree(c->d_ao_rgb);
    c->d_ao_f = nullptr; c->d_ao_d = nullptr; c->d_ao_rgb = nullptr;
    HIPCHECK(c, hipMalloc((void **)&c->d_ao_f, (size_t)count * 4 * sizeof(float)));
    HIPCHECK(c, hipMalloc((void **)&c->d_ao_d, (size_t)count * 4 * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&c->d_ao_rgb, (size_t)count * 4 * sizeof(int)));
  }
  c->ao_count = std::max(c->ao_count, count);
  std::vector<float> dirs_f(dirs.begin(), dirs.end());
  HIPCHECK(c, hipMemcpyAsync(c->d_ao_f, dirs_f.data(), dirs_f.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_ao_d, dirs.data(), dirs.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_ao_rgb, rgb.data(), rgb.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_sc, &d, sizeof d, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->host_sc = d;
  c->scene = p.s;
  c->has_scene = true;
  return MBR_OK;
}

}  // extern "C"

namespace {

// The shading buffers of slot sl for a frame of the scene `host` (no
// reflection levels: the shade keeps specular and object parts together).
// The power-8 Mandelbulb runs its trig-free kernels (ip8) in the FP32 modes
// and in FP64_DEMOIVRE (trig_free); plain FP64 keeps the reference's trig.
ShadeBufs frame_shade_bufs(Slot &sl, const DevScene &host, int formula, bool trig_free) {
  BandBufs &b = sl.band;
  b.sp_r = b.sp_g = b.sp_b = nullptr;
  b.ray_o = b.ray_d = b.ray_t0 = nullptr;
  b.active = nullptr;
  ShadeBufs sbuf;
  sbuf.base = sl.sh_base;
  sbuf.spec = sl.sh_spec;
  sbuf.col = sl.sh_col;
  sbuf.fast_ao = sl.sh_fast;
  sbuf.ao = sl.sh_ao;
  sbuf.aux = sl.sh_aux;
  sbuf.n = b.cap;
  sbuf.slow_ao = host.ao && !host.fast_ao;
  sbuf.deep = host.deep;
  sbuf.plain = host.plain;
  sbuf.formula = formula;
  sbuf.lights = host.n_lights > 0 || host.fake_lights;
  sbuf.ip8 = trig_free && formula == 2 && host.fd.ipow == 8;
  return sbuf;
}

// One band, rays [band_first, band_first + band_rays), through the buffers
// of slot sl on stream st: reset -> march -> FP64 continuation -> shade ->
// finish.  With e (6 events) the launches are bracketed for timing.
int enqueue_band(mbr_ctx *c, Slot &sl, const ShadeBufs &sbuf, int precision, const DevScene *d_sc,
                 const DevScene &host, int formula, AODev ao, long long band_first, int band_rays, const FrameOut &out,
                 int persist, hipStream_t st, hipEvent_t *e, unsigned long long *d_stats) {
  const BandBufs &b = sl.band;
  HIPCHECK(c, launch_band_reset(b, st));
  if (e) HIPCHECK(c, hipEventRecord(e[0], st));
  HIPCHECK(c, launch_march(precision, d_sc, b, (int)band_first, band_rays, formula, host.plain, persist, d_stats, st,
                           sbuf.ip8 != 0));
  HIPCHECK(c, launch_march_fix(precision, d_sc, b, (int)band_first, formula, persist / 2, d_stats, st,
                               sbuf.ip8 != 0));
  if (e) HIPCHECK(c, hipEventRecord(e[1], st));
  HIPCHECK(c, launch_shade(precision, d_sc, b, (int)band_first, ao, sbuf, persist, d_stats, st,
                           e ? e[4] : nullptr, e ? e[5] : nullptr));
  if (e) HIPCHECK(c, hipEventRecord(e[2], st));
  HIPCHECK(c, launch_finish(precision, d_sc, b, (int)band_first, band_rays, out, st, formula,
                            host.vol_any || host.iter_fog || host.fake_lights,
                            host.n_lights > 0 && host.sd.aux_visibility > 0, persist,
                            d_stats, ao));
  if (e) HIPCHECK(c, hipEventRecord(e[3], st));
  return MBR_OK;
}

// Enqueue every band of one frame (rays [first, last) of the image) on
// stream st, with the scene block d_sc, through the buffers of slot sl.
// `persist` is the workgroup count of the persistent launches.  With ev
// (6 events per band) the launches are bracketed for timing; a captured
// batch passes none.
int enqueue_frame(mbr_ctx *c, Slot &sl, int precision, const DevScene *d_sc, const DevScene &host, int formula,
                  AODev ao, long long first, long long last, const FrameOut &out, int persist, hipStream_t st,
                  hipEvent_t *ev, volatile const int *cancel, unsigned long long *stats = nullptr,
                  bool demoivre = false) {
  unsigned long long *const d_stats = stats ? stats : c->d_stats;
  const ShadeBufs sbuf = frame_shade_bufs(sl, host, formula, precision != MBR_PREC_FP64 || demoivre);
  const long long cap = sl.band.cap;
  int band = 0;
  for (long long band_first = first; band_first < last; band_first += cap, band++) {
    if (cancel && *cancel) return c->fail(MBR_ERR_CANCELLED, "render cancelled");
    const int band_rays = (int)std::min<long long>(cap, last - band_first);
    const int rc = enqueue_band(c, sl, sbuf, precision, d_sc, host, formula, ao, band_first, band_rays, out, persist,
                                st, ev ? ev + 6 * band : nullptr, d_stats);
    if (rc != MBR_OK) return rc;
  }
  return MBR_OK;
}

// An FP32 frame of at least kSplitRays rays renders as n_bands equal bands
// that alternate between two slots and two streams (st, and st1 forked from
// it and joined back at the end): while one band's march drains its longest
// rays, or its continuation, surface and composition kernels leave the
// vector units idle, the other band's kernels fill the CUs.  Each band's
// results depend only on its own rays, so the frame equals the one-stream
// frame bit for bit.  Measured on the 1080p power-8 bulb (MI355X): two
// bands 13.5 ms against 14.4 ms on one stream; four or more bands, a
// staggered start of the second stream and smaller march or AO grids were
// all slower.  FP64 frames stay on one stream (two FP64 bands at once took
// 133-143 ms against 127 ms).
constexpr long long kSplitRays = 1 << 18;
constexpr int kSplitBands = 2;

int enqueue_frame_split(mbr_ctx *c, Slot *sl[2], int precision, const DevScene *d_sc, const DevScene &host,
                        int formula, AODev ao, long long first, long long last, int n_bands, const FrameOut &out,
                        int persist, hipStream_t st, hipStream_t st1, hipEvent_t *ev, volatile const int *cancel) {
  // FP32 modes only: the trig-free power-8 kernels where they apply
  const ShadeBufs sbuf[2] = {frame_shade_bufs(*sl[0], host, formula, true),
                             frame_shade_bufs(*sl[1], host, formula, true)};
  const long long size = (last - first + n_bands - 1) / n_bands;
  if (size > sl[0]->band.cap || size > sl[1]->band.cap) return c->fail(MBR_ERR_INVALID_ARG, "split band above slot");
  HIPCHECK(c, hipEventRecord(c->split_fork, st));
  HIPCHECK(c, hipStreamWaitEvent(st1, c->split_fork, 0));
  int rc = MBR_OK, band = 0;
  for (long long band_first = first; band_first < last; band_first += size, band++) {
    if (cancel && *cancel) {
      rc = c->fail(MBR_ERR_CANCELLED, "render cancelled");
      break;
    }
    const int k = band & 1;
    const int band_rays = (int)std::min<long long>(size, last - band_first);
    rc = enqueue_band(c, *sl[k], sbuf[k], precision, d_sc, host, formula, ao, band_first, band_rays, out, persist,
                      k ? st1 : st, ev ? ev + 6 * band : nullptr, c->d_stats);
    if (rc != MBR_OK) break;
  }
  // st continues only after the other stream's bands, also when a band
  // could not be enqueued
  HIPCHECK(c, hipEventRecord(c->split_join, st1));
  HIPCHECK(c, hipStreamWaitEvent(st, c->split_join, 0));
  return rc;
}

ShadeBufs shade_bufs(Slot &sl, const DevScene &host, int formula) {
  ShadeBufs sbuf;
  sbuf.base = sl.sh_base;
  sbuf.spec = sl.sh_spec;
  sbuf.col = sl.sh_col;
  sbuf.fast_ao = sl.sh_fast;
  sbuf.ao = sl.sh_ao;
  sbuf.aux = sl.sh_aux;
  sbuf.n = sl.band.cap;
  sbuf.slow_ao = host.ao && !host.fast_ao;
  sbuf.deep = 0;
  sbuf.plain = host.plain;
  sbuf.formula = formula;
  sbuf.lights = host.n_lights > 0 || host.fake_lights;
  sbuf.ip8 = 0;  // reflections render in FP64 with the reference's trig
  return sbuf;
}

// The reflection planes of a level's slot (level 0: the specular planes
// only), made once and kept with the slot.
int alloc_refl(mbr_ctx *c, Slot &sl, bool mirrored) {
  const size_t n = (size_t)sl.band.cap;
  if (!sl.rf_sp) HIPCHECK(c, hipMalloc(&sl.rf_sp, 3 * n * sizeof(double)));
  if (mirrored && !sl.rf_o) {
    HIPCHECK(c, hipMalloc((void **)&sl.rf_o, 3 * n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&sl.rf_d, 3 * n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&sl.rf_t0, n * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&sl.rf_active, n));
  }
  BandBufs &b = sl.band;
  b.sp_r = sl.rf_sp;
  b.sp_g = (double *)sl.rf_sp + n;
  b.sp_b = (double *)sl.rf_sp + 2 * n;
  b.ray_o = mirrored ? sl.rf_o : nullptr;
  b.ray_d = mirrored ? sl.rf_d : nullptr;
  b.ray_t0 = mirrored ? sl.rf_t0 : nullptr;
  b.active = mirrored ? sl.rf_active : nullptr;
  return MBR_OK;
}

// One frame with raytraced reflections (FP64): per band the primary march,
// then per further level the mirrored rays of the previous level's
// reflecting hits and their march, the shading of every level, and the
// back-to-front composition (MainThread, src/Render3D.cpp:440-589).
int enqueue_frame_refl(mbr_ctx *c, std::vector<Slot *> &lv, const DevScene *d_sc, const DevScene &host, int formula,
                       AODev ao, long long first, long long last, const FrameOut &out, int persist, hipStream_t st,
                       hipEvent_t *ev, volatile const int *cancel) {
  const int n_levels = (int)lv.size();
  // the device copy of the levels' band buffers (the previous frame's
  // finish may still read it)
  HIPCHECK(c, hipStreamSynchronize(st));
  c->levels_host.resize((size_t)n_levels);
  for (int k = 0; k < n_levels; k++) c->levels_host[(size_t)k] = lv[(size_t)k]->band;
  if (!c->d_levels) HIPCHECK(c, hipMalloc((void **)&c->d_levels, (MBR_MAX_REFLECTIONS + 1) * sizeof(BandBufs)));
  HIPCHECK(c, hipMemcpy(c->d_levels, c->levels_host.data(), (size_t)n_levels * sizeof(BandBufs),
                        hipMemcpyHostToDevice));
  const long long cap = lv[0]->band.cap;
  int band = 0;
  for (long long band_first = first; band_first < last; band_first += cap, band++) {
    if (cancel && *cancel) return c->fail(MBR_ERR_CANCELLED, "render cancelled");
    const int band_rays = (int)std::min<long long>(cap, last - band_first);
    hipEvent_t *e = ev ? ev + 6 * band : nullptr;
    BandBufs &b0 = lv[0]->band;
    HIPCHECK(c, launch_band_reset(b0, st));
    if (e) HIPCHECK(c, hipEventRecord(e[0], st));
    HIPCHECK(c, launch_march(MBR_PREC_FP64, d_sc, b0, (int)band_first, band_rays, formula, host.plain, persist,
                             c->d_stats, st));
    HIPCHECK(c, launch_march_fix(MBR_PREC_FP64, d_sc, b0, (int)band_first, formula, persist / 2, c->d_stats, st));
    for (int k = 1; k < n_levels; k++) {
      BandBufs &bk = lv[(size_t)k]->band;
      HIPCHECK(c, launch_band_reset(bk, st));
      HIPCHECK(c, launch_refl_clear(bk, band_rays, st));
      HIPCHECK(c, launch_refl_setup(d_sc, lv[(size_t)k - 1]->band, bk, (int)band_first, formula, persist, c->d_stats,
                                    st));
      HIPCHECK(c, launch_refl_march(d_sc, bk, (int)band_first, formula, persist, c->d_stats, st));
    }
    if (e) HIPCHECK(c, hipEventRecord(e[1], st));
    for (int k = 0; k < n_levels; k++) {
      const ShadeBufs sbuf = shade_bufs(*lv[(size_t)k], host, formula);
      HIPCHECK(c, launch_shade(MBR_PREC_FP64, d_sc, lv[(size_t)k]->band, (int)band_first, ao, sbuf, persist,
                               c->d_stats, st, (e && k == 0) ? e[4] : nullptr, (e && k == 0) ? e[5] : nullptr));
    }
    if (e) HIPCHECK(c, hipEventRecord(e[2], st));
    HIPCHECK(c, launch_refl_finish(d_sc, c->d_levels, n_levels, (int)band_first, band_rays, out,
                                   host.n_lights > 0 && host.sd.aux_visibility > 0,
                                   host.shadow_fx || host.fake_lights, formula, ao,
                                   c->d_stats, st));
    if (e) HIPCHECK(c, hipEventRecord(e[3], st));
  }
  return MBR_OK;
}

FrameOut frame_out(const mbr_framebuffer *fb) {
  FrameOut out;
  for (int i = 0; i < 3; i++) {
    out.rgb[i] = fb->rgb[i];
    out.colour[i] = fb->colour[i];
  }
  out.z = fb->z;
  out.alpha = fb->alpha;
  out.opacity = fb->opacity;
  return out;
}

void fill_stats(mbr_stats *stats, const unsigned long long *v, long long rays, int launches) {
  std::memset(stats, 0, sizeof *stats);
  stats->rays = (uint64_t)rays;
  stats->hits = v[ST_HITS];
  stats->march_de_calls = v[ST_MARCH_DE];
  stats->shade_de_calls = v[ST_SHADE_DE];
  stats->iterations = v[ST_MARCH_ITERS] + v[ST_SHADE_ITERS];
  stats->march_iterations = v[ST_MARCH_ITERS];
  stats->shade_iterations = v[ST_SHADE_ITERS];
  stats->march_launches = launches;
  stats->shade_launches = launches;
  stats->finish_launches = launches;
  stats->missed_de = v[ST_MISSED];
  stats->step_overflow_rays = v[ST_OVF_RAYS];
  stats->fp64_redo_rays = v[ST_REDO];
  stats->ao_de_calls = v[ST_AO_DE];
  stats->ao_iterations = v[ST_AO_ITERS];
  stats->ao_steps = v[ST_AO_STEPS];
}

// After a render has finished on the device: MBR_ERR_CANCELLED when
// mbr_cancel was called since the last cancelled render (the flag is then
// cleared; the frame is incomplete), else MBR_OK.
int consume_cancel(mbr_ctx *c) {
  if (c->h_cancel && __atomic_exchange_n(c->h_cancel, 0, __ATOMIC_SEQ_CST) != 0)
    return c->fail(MBR_ERR_CANCELLED, "render cancelled (mbr_cancel)");
  return MBR_OK;
}

// Checks the requested precision; MBR_PREC_FP64_DEMOIVRE renders as FP64
// with the trig-free power-8 iteration (demoivre set).
int check_precision(mbr_ctx *c, int &precision, bool &demoivre) {
  if (precision != MBR_PREC_FP64 && precision != MBR_PREC_FP32_FIXUP && precision != MBR_PREC_FP32_MARCH &&
      precision != MBR_PREC_FP64_DEMOIVRE)
    return c->fail(MBR_ERR_INVALID_ARG, "unknown precision");
  demoivre = precision == MBR_PREC_FP64_DEMOIVRE;
  if (demoivre) precision = MBR_PREC_FP64;
  return MBR_OK;
}

}  // namespace

extern "C" {

int mbr_render_device(mbr_ctx *c, const mbr_render_opts *opts, mbr_framebuffer *fb, mbr_stats *stats) {
  if (!c || !fb) return MBR_ERR_INVALID_ARG;
  if (!c->has_scene) return c->fail(MBR_ERR_NO_SCENE, "mbr_set_scene was not called");
  const mbr_scene &s = c->scene;
  if (fb->width != s.image_width || fb->height != s.image_height)
    return c->fail(MBR_ERR_INVALID_ARG, "framebuffer size differs from the scene's image size");
  int precision = opts ? opts->precision : MBR_PREC_FP32_FIXUP;
  bool demoivre = false;
  int rc = check_precision(c, precision, demoivre);
  if (rc != MBR_OK) return rc;
  // formulas without kernels of their own (delta DE, trig_DE) and
  // reflections render in FP64
  precision = c->host_sc.refl_max > 0 ? MBR_PREC_FP64 : scene_precision(s, &c->host_sc, precision);
  int row0 = opts ? opts->row_begin : 0, row1 = opts ? opts->row_end : 0;
  if (row1 <= 0 || row1 > s.image_height) row1 = s.image_height;
  if (row0 < 0) row0 = 0;
  if (row0 >= row1) return c->fail(MBR_ERR_INVALID_ARG, "empty row range");
  hipStream_t st = (opts && opts->stream) ? (hipStream_t)opts->stream : c->stream;
  HIPCHECK(c, hipSetDevice(c->device));
  const bool slow_ao = c->host_sc.ao && !c->host_sc.fast_ao;
  const int ao_count = slow_ao ? c->host_sc.ao_count : 0;
  Slot &sl = c->slots[0];
  rc = alloc_slot(c, sl, precision, band_cap(precision, ao_count), ao_count, c->host_sc.iter_fog != 0);
  if (rc != MBR_OK) return rc;
  // reflection levels: a slot each after the first
  const int refl = c->host_sc.refl_max;
  std::vector<Slot *> levels;
  if (refl > 0) {
    if ((int)c->refl_slots.size() < refl) c->refl_slots.resize((size_t)refl);
    levels.push_back(&sl);
    rc = alloc_refl(c, sl, false);
    if (rc != MBR_OK) return rc;
    for (int k = 0; k < refl; k++) {
      Slot &lk = c->refl_slots[(size_t)k];
      rc = alloc_slot(c, lk, precision, sl.band.cap, ao_count, c->host_sc.iter_fog != 0);
      if (rc == MBR_OK) rc = alloc_refl(c, lk, true);
      if (rc != MBR_OK) return rc;
      levels.push_back(&lk);
    }
  }
  const FrameOut out = frame_out(fb);
  const AODev ao{c->d_ao_f, c->d_ao_d, c->d_ao_rgb, c->host_sc.ao_count};
  const long long first = (long long)row0 * s.image_width;
  const long long last = (long long)row1 * s.image_width;
  // large frames without reflections: bands over two slots and streams
  const bool split = refl == 0 && precision != MBR_PREC_FP64 && last - first >= kSplitRays;
  Slot *pair[2] = {&sl, &c->slots[1]};
  int n_bands = (int)((last - first + sl.band.cap - 1) / sl.band.cap);
  if (split) {
    rc = alloc_slot(c, c->slots[1], precision, sl.band.cap, ao_count, c->host_sc.iter_fog != 0);
    if (rc != MBR_OK) return rc;
    if (!c->side[1]) HIPCHECK(c, hipStreamCreateWithFlags(&c->side[1], hipStreamNonBlocking));
    if (!c->split_fork) HIPCHECK(c, hipEventCreateWithFlags(&c->split_fork, hipEventDisableTiming));
    if (!c->split_join) HIPCHECK(c, hipEventCreateWithFlags(&c->split_join, hipEventDisableTiming));
    n_bands = std::max(kSplitBands, n_bands + (n_bands & 1));
  }
  // 6 events per band, then the frame's first and last
  while ((int)c->ev.size() < 6 * n_bands + 2) {
    hipEvent_t e;
    HIPCHECK(c, hipEventCreate(&e));
    c->ev.push_back(e);
  }
  hipEvent_t *frame_ev = &c->ev[6 * (size_t)n_bands];
  const int persist = c->cu_count > 0 ? c->cu_count * 8 : 2048;
  unsigned long long v[ST_COUNT];
  for (int attempt = 0;; attempt++) {
    HIPCHECK(c, launch_zero(c->d_stats, ST_COUNT, st));
    HIPCHECK(c, hipEventRecord(frame_ev[0], st));
    if (refl > 0)
      rc = enqueue_frame_refl(c, levels, c->d_sc, c->host_sc, kernel_family(s), ao, first, last, out, persist, st,
                              c->ev.data(), opts ? opts->cancel : nullptr);
    else if (split)
      rc = enqueue_frame_split(c, pair, precision, c->d_sc, c->host_sc, kernel_family(s), ao, first, last, n_bands, out,
                               persist, st, c->side[1], c->ev.data(), opts ? opts->cancel : nullptr);
    else
      rc = enqueue_frame(c, sl, precision, c->d_sc, c->host_sc, kernel_family(s), ao, first, last, out, persist, st,
                         c->ev.data(), opts ? opts->cancel : nullptr, nullptr, demoivre);
    if (rc != MBR_OK) return rc;
    HIPCHECK(c, hipEventRecord(frame_ev[1], st));
    HIPCHECK(c, hipMemcpyAsync(v, c->d_stats, sizeof v, hipMemcpyDeviceToHost, st));
    HIPCHECK(c, hipStreamSynchronize(st));
    rc = consume_cancel(c);
    if (rc != MBR_OK) return rc;
    if (v[ST_POOL_MISS] == 0) break;
    // some ray found the overflow pool full: grow it to hold every
    // overflowing ray of the frame and render again
    if (attempt + 1 >= kPoolAttempts) return c->fail(MBR_ERR_HIP, "overflow pool still full after growing it");
    rc = grow_pool(c, sl, v[ST_OVF_RAYS], v[ST_OVF_CHUNKS]);
    if (rc == MBR_OK && split) rc = grow_pool(c, c->slots[1], v[ST_OVF_RAYS], v[ST_OVF_CHUNKS]);
    if (rc != MBR_OK) return rc;
    for (int k = 0; k < refl; k++) {
      rc = grow_pool(c, c->refl_slots[(size_t)k], v[ST_OVF_RAYS], v[ST_OVF_CHUNKS]);
      if (rc != MBR_OK) return rc;
    }
  }
  double ms_march = 0, ms_shade = 0, ms_finish = 0, ms_ao = 0;
  for (int band = 0; band < n_bands; band++) {
    const hipEvent_t *e = &c->ev[6 * band];
    float t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    HIPCHECK(c, hipEventElapsedTime(&t0, e[0], e[1]));
    HIPCHECK(c, hipEventElapsedTime(&t1, e[1], e[2]));
    HIPCHECK(c, hipEventElapsedTime(&t2, e[2], e[3]));
    HIPCHECK(c, hipEventElapsedTime(&t3, e[4], e[5]));
    ms_march += t0;
    ms_shade += t1;
    ms_finish += t2;
    ms_ao += t3;
  }
  float ms_frame = 0;
  HIPCHECK(c, hipEventElapsedTime(&ms_frame, frame_ev[0], frame_ev[1]));
  // the AO phase of the frame: first AO start to last AO end, over bands
  // that may have run on two streams (event times relative to the frame's
  // first event, recorded on st before any band)
  double ao_first = 1e300, ao_last = -1e300;
  for (int band = 0; band < n_bands; band++) {
    const hipEvent_t *e = &c->ev[6 * band];
    float a = 0, b = 0;
    HIPCHECK(c, hipEventElapsedTime(&a, frame_ev[0], e[4]));
    HIPCHECK(c, hipEventElapsedTime(&b, frame_ev[0], e[5]));
    ao_first = std::min(ao_first, (double)a);
    ao_last = std::max(ao_last, (double)b);
  }
  if (stats) {
    fill_stats(stats, v, last - first, n_bands);
    // the parts are sums over bands; with two streams they overlap, and the
    // total is the frame's own span
    stats->ms_march = ms_march;
    stats->ms_shade = ms_shade;
    stats->ms_finish = ms_finish;
    stats->ms_ao = ms_ao;
    stats->ms_ao_span = n_bands > 0 ? ao_last - ao_first : 0.0;
    stats->ms_total = ms_frame;
  }
  return MBR_OK;
}

}  // extern "C"

namespace {

// Band buffers of one slot carved out of `base` (kScratchBytes): the counters
// of the render, the per-ray planes of alloc_slot for `cap` rays, and an
// overflow pool of one directory and one chunk per 64 rays.  Returns the
// bytes used; with base == nullptr only sizes the layout.
size_t carve_scratch(unsigned char *base, int precision, int cap, int ao_count, bool iters, Slot &sl,
                     unsigned long long **stats) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> void * {
    off = (off + 255) & ~size_t(255);
    void *p = base ? base + off : nullptr;
    off += bytes;
    return p;
  };
  const size_t n = (size_t)cap;
  const size_t rec = precision == MBR_PREC_FP64 ? 8 : 4;
  sl = Slot{};
  BandBufs &b = sl.band;
  b.cap = cap;
  b.slab = kSlab;
  b.rec_bytes = (int)rec;
  *stats = (unsigned long long *)take(ST_COUNT * sizeof(unsigned long long));
  b.counters = (unsigned long long *)take(CNT_COUNT * sizeof(unsigned long long));
  b.found = (uint8_t *)take(n);
  b.object = (uint8_t *)take(n);
  b.count = (int *)take(n * sizeof(int));
  b.depth = (double *)take(n * sizeof(double));
  b.thr = (double *)take(n * sizeof(double));
  b.last = (double *)take(n * sizeof(double));
  b.px = (double *)take(n * sizeof(double));
  b.py = (double *)take(n * sizeof(double));
  b.pz = (double *)take(n * sizeof(double));
  b.rec_step = take(n * kSlab * rec);
  b.rec_dist = take(n * kSlab * rec);
  b.rec_thr = take(n * kSlab * rec);
  b.ovf_off = (long long *)take(n * sizeof(long long));
  b.hit_list = (int *)take(n * sizeof(int));
  b.redo_list = (int *)take(n * sizeof(int));
  b.hand_ray = (int *)take(n * sizeof(int));
  b.hand_i = (int *)take(n * sizeof(int));
  b.hand_scan = (double *)take(n * sizeof(double));
  b.hand_step = (double *)take(n * sizeof(double));
  b.hand_drift = (double *)take(n * sizeof(double));
  b.hand_dist = (double *)take(n * sizeof(double));
  b.sh_r = take(n * rec);
  b.sh_g = take(n * rec);
  b.sh_b = take(n * rec);
  b.col_r = (uint8_t *)take(n);
  b.col_g = (uint8_t *)take(n);
  b.col_b = (uint8_t *)take(n);
  sl.sh_base = take(3 * n * rec);
  sl.sh_spec = take(3 * n * rec);
  sl.sh_col = take(3 * n * rec);
  sl.sh_fast = take(n * rec);
  sl.sh_aux = take(9 * n * rec);
  if (ao_count > 0) {
    sl.sh_ao_bytes = n * (size_t)ao_count * rec;
    sl.sh_ao = take(sl.sh_ao_bytes);
  }
  const long long pool = std::max<long long>(256, (long long)n / 64);
  b.ovf_dirs = pool;
  b.ovf_chunks = pool;
  b.ovf_dir = (int *)take((size_t)pool * kDirSlots * sizeof(int));
  b.ovf_step = take((size_t)pool * kChunk * rec);
  b.ovf_dist = take((size_t)pool * kChunk * rec);
  b.ovf_thr = take((size_t)pool * kChunk * rec);
  if (iters) {
    b.rec_it = (int *)take(n * kSlab * sizeof(int));
    b.ovf_it = (int *)take((size_t)pool * kChunk * sizeof(int));
  }
  sl.prec = precision;
  return off;
}

}  // namespace

extern "C" {

int mbr_render_device_noalloc(mbr_ctx *c, const mbr_render_opts *opts, mbr_framebuffer *fb, mbr_stats *stats) {
  if (!c || !fb) return MBR_ERR_INVALID_ARG;
  if (!c->has_scene) return c->fail(MBR_ERR_NO_SCENE, "mbr_set_scene was not called");
  const mbr_scene &s = c->scene;
  if (fb->width != s.image_width || fb->height != s.image_height)
    return c->fail(MBR_ERR_INVALID_ARG, "framebuffer size differs from the scene's image size");
  int precision = opts ? opts->precision : MBR_PREC_FP32_FIXUP;
  bool demoivre = false;
  int rc = check_precision(c, precision, demoivre);
  if (rc != MBR_OK) return rc;
  if (c->host_sc.refl_max > 0)
    return c->fail(MBR_ERR_UNSUPPORTED, "mbr_render_device_noalloc: raytraced reflections need a band slot per "
                                        "level; render them with mbr_render_device");
  precision = scene_precision(s, &c->host_sc, precision);
  unsigned char *base = c->device < kMaxDevices ? (unsigned char *)g_scratch_ptr[c->device] : nullptr;
  if (!base) return c->fail(MBR_ERR_HIP, "mbr_render_device_noalloc: no scratch address for this device");
  int row0 = opts ? opts->row_begin : 0, row1 = opts ? opts->row_end : 0;
  if (row1 <= 0 || row1 > s.image_height) row1 = s.image_height;
  if (row0 < 0) row0 = 0;
  if (row0 >= row1) return c->fail(MBR_ERR_INVALID_ARG, "empty row range");
  const bool slow_ao = c->host_sc.ao && !c->host_sc.fast_ao;
  const int ao_count = slow_ao ? c->host_sc.ao_count : 0;
  const bool iters = c->host_sc.iter_fog != 0;
  // the largest band whose buffers fit the scratch (multiple of 256 rays)
  Slot sl;
  unsigned long long *d_stats = nullptr;
  const size_t per_256 = carve_scratch(nullptr, precision, 512, ao_count, iters, sl, &d_stats) -
                         carve_scratch(nullptr, precision, 256, ao_count, iters, sl, &d_stats);
  const size_t fixed = carve_scratch(nullptr, precision, 256, ao_count, iters, sl, &d_stats) - per_256;
  long long cap = kScratchBytes > fixed ? (long long)((kScratchBytes - fixed) / per_256) * 256 : 0;
  cap = std::min<long long>(cap, kBandRays);
  while (cap >= 256 && carve_scratch(nullptr, precision, (int)cap, ao_count, iters, sl, &d_stats) > kScratchBytes)
    cap -= 256;
  if (cap < kNoallocMinBand)
    return c->fail(MBR_ERR_OUT_OF_MEMORY,
                   "mbr_render_device_noalloc: this scene needs " + std::to_string(per_256 / 256) +
                       " bytes per ray (AO directions x record size), so the " + std::to_string(kScratchBytes >> 20) +
                       " MiB scratch holds fewer than " + std::to_string(kNoallocMinBand) +
                       " rays per band; render it with mbr_render_device");
  carve_scratch(base, precision, (int)cap, ao_count, iters, sl, &d_stats);
  sl.band.cancel = c->d_cancel;
  hipStream_t st = (opts && opts->stream) ? (hipStream_t)opts->stream : c->stream;
  HIPCHECK(c, hipSetDevice(c->device));
  const FrameOut out = frame_out(fb);
  const AODev ao{c->d_ao_f, c->d_ao_d, c->d_ao_rgb, c->host_sc.ao_count};
  const long long first = (long long)row0 * s.image_width;
  const long long last = (long long)row1 * s.image_width;
  const int persist = c->cu_count > 0 ? c->cu_count * 8 : 2048;
  HIPCHECK(c, launch_zero(d_stats, ST_COUNT, st));
  rc = enqueue_frame(c, sl, precision, c->d_sc, c->host_sc, kernel_family(s), ao, first, last, out, persist, st, nullptr,
                     opts ? opts->cancel : nullptr, d_stats, demoivre);
  if (rc != MBR_OK) return rc;
  unsigned long long v[ST_COUNT];
  HIPCHECK(c, hipMemcpyAsync(v, d_stats, sizeof v, hipMemcpyDeviceToHost, st));
  HIPCHECK(c, hipStreamSynchronize(st));
  rc = consume_cancel(c);
  if (rc != MBR_OK) return rc;
  if (v[ST_POOL_MISS] != 0)
    return c->fail(MBR_ERR_OUT_OF_MEMORY, "mbr_render_device_noalloc: rays of more than " + std::to_string(kSlab) +
                                              " steps outgrew the scratch's overflow pool (" +
                                              std::to_string(sl.band.ovf_chunks) +
                                              " chunks per band); render this frame with mbr_render_device");
  if (stats) {
    fill_stats(stats, v, last - first, (int)((last - first + cap - 1) / cap));
    stats->ms_march = stats->ms_shade = stats->ms_finish = stats->ms_ao = stats->ms_total = 0.0;
  }
  return MBR_OK;
}

int mbr_compile_image_device(mbr_ctx *c, const mbr_scene *s, int w, int h, const float *const rgb[3], uint16_t *image16,
                             void *stream) {
  if (!c || !s || !rgb || !rgb[0] || !rgb[1] || !rgb[2] || !image16 || w <= 0 || h <= 0)
    return MBR_ERR_INVALID_ARG;
  HIPCHECK(c, hipSetDevice(c->device));
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  if (!c->d_gamma || c->gamma_of_table != s->image_gamma) {
    // CalculateGammaTable (src/cimage.cpp:176-182) with the host libm
    std::vector<uint16_t> gamma(65536);
    for (int i = 0; i < 65536; i++) gamma[i] = (uint16_t)(int)(std::pow(i / 65536.0, 1.0 / s->image_gamma) * 65535);
    if (!c->d_gamma) HIPCHECK(c, hipMalloc(&c->d_gamma, 65536 * sizeof(uint16_t)));
    else HIPCHECK(c, hipDeviceSynchronize());  // an earlier compile may still read the old table
    HIPCHECK(c, hipMemcpyAsync(c->d_gamma, gamma.data(), 65536 * sizeof(uint16_t), hipMemcpyHostToDevice, st));
    HIPCHECK(c, hipStreamSynchronize(st));
    c->gamma_of_table = s->image_gamma;
  }
  CompileParams p;
  p.n = w * h;
  p.hdr = s->hdr;
  p.brightness = s->brightness;
  p.contrast = s->contrast;
  HIPCHECK(c, launch_compile(p, rgb[0], rgb[1], rgb[2], c->d_gamma, image16, st));
  return MBR_OK;
}

int mbr_post_ssao(mbr_ctx *c, const mbr_scene *s, const mbr_post_planes *pl, int on_device, void *stream) {
  if (!c || !s || !pl) return MBR_ERR_INVALID_ARG;
  const int w = pl->width, h = pl->height, q = s->ssao_quality;
  if (w <= 0 || h <= 0 || !pl->z || !pl->opacity || !pl->colour[0] || !pl->colour[1] || !pl->colour[2] ||
      !pl->image16)
    return c->fail(MBR_ERR_INVALID_ARG, "mbr_post_ssao: planes missing");
  if (q < 1 || q > 1024) return c->fail(MBR_ERR_INVALID_ARG, "mbr_post_ssao: ssao_quality must be in [1, 1024]");
  if (s->perspective_type < MBR_PERSP_THREE_POINT || s->perspective_type > MBR_PERSP_EQUIRECTANGULAR)
    return c->fail(MBR_ERR_INVALID_ARG, "mbr_post_ssao: unknown perspective_type");
  HIPCHECK(c, hipSetDevice(c->device));
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  if (q != c->ssao_tab_q) {
    // the direction table of src/image.cpp:266-271, from the host libm
    std::vector<double> tab(2 * (size_t)q);
    for (int i = 0; i < q; i++) {
      tab[i] = cos((double)i / q * 2.0 * M_PI);
      tab[q + i] = sin((double)i / q * 2.0 * M_PI);
    }
    // an earlier device-plane call may still read the old table on another stream
    if (c->d_ssao_tab) HIPCHECK(c, hipDeviceSynchronize());
    if (c->d_ssao_tab) HIPCHECK(c, hipFree(c->d_ssao_tab));
    c->d_ssao_tab = nullptr;
    c->ssao_tab_q = 0;
    HIPCHECK(c, hipMalloc(&c->d_ssao_tab, tab.size() * sizeof(double)));
    HIPCHECK(c, hipMemcpyAsync(c->d_ssao_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHECK(c, hipStreamSynchronize(st));
    c->ssao_tab_q = q;
  }
  SsaoParams p;
  p.width = w;
  p.height = h;
  p.quality = q;
  p.type = s->perspective_type;
  p.scale = (double)w / (q * q) / 2.0;
  p.aspect = (double)w / h;
  p.fov = s->persp;
  p.intensity = s->global_ilum;
  p.cs = c->d_ssao_tab;
  p.sn = c->d_ssao_tab + q;
  if (on_device) {
    HIPCHECK(c, launch_ssao(p, pl->z, pl->opacity, pl->colour, pl->image16, st));
    return MBR_OK;
  }
  // host planes: stage z, opacity, colour and image16 on the device
  const size_t n = (size_t)w * h;
  const size_t bytes = n * (4 + 2 + 3 + 6) + 64 * 6;
  if (bytes > c->d_post_bytes) {
    if (c->d_post) HIPCHECK(c, hipFree(c->d_post));
    c->d_post = nullptr;
    c->d_post_bytes = 0;
    HIPCHECK(c, hipMalloc(&c->d_post, bytes));
    c->d_post_bytes = bytes;
  }
  char *base = (char *)c->d_post;
  size_t off = 0;
  auto take = [&](size_t sz) { void *ptr = base + off; off += (sz + 63) / 64 * 64; return ptr; };
  float *dz = (float *)take(n * 4);
  uint16_t *dop = (uint16_t *)take(n * 2);
  uint8_t *dcol[3] = {(uint8_t *)take(n), (uint8_t *)take(n), (uint8_t *)take(n)};
  uint16_t *dimg = (uint16_t *)take(n * 6);
  HIPCHECK(c, hipMemcpyAsync(dz, pl->z, n * 4, hipMemcpyHostToDevice, st));
  HIPCHECK(c, hipMemcpyAsync(dop, pl->opacity, n * 2, hipMemcpyHostToDevice, st));
  for (int i = 0; i < 3; i++) HIPCHECK(c, hipMemcpyAsync(dcol[i], pl->colour[i], n, hipMemcpyHostToDevice, st));
  HIPCHECK(c, hipMemcpyAsync(dimg, pl->image16, n * 6, hipMemcpyHostToDevice, st));
  HIPCHECK(c, launch_ssao(p, dz, dop, dcol, dimg, st));
  HIPCHECK(c, hipMemcpyAsync(pl->image16, dimg, n * 6, hipMemcpyDeviceToHost, st));
  HIPCHECK(c, hipStreamSynchronize(st));
  return MBR_OK;
}

int mbr_post_dof(mbr_ctx *c, const mbr_scene *s, const mbr_post_planes *pl, int on_device, void *stream) {
  if (!c || !s || !pl) return MBR_ERR_INVALID_ARG;
  const int w = pl->width, h = pl->height;
  if (w <= 0 || h <= 0 || !pl->z || !pl->image16) return c->fail(MBR_ERR_INVALID_ARG, "mbr_post_dof: planes missing");
  if ((long long)w * h > (1ll << 28)) return c->fail(MBR_ERR_INVALID_ARG, "mbr_post_dof: image too large");
  HIPCHECK(c, hipSetDevice(c->device));
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  const int n = w * h;
  DofParams p;
  p.width = w;
  p.height = h;
  p.tiles_x = (w + 15) / 16;
  p.tiles_y = (h + 15) / 16;
  p.rank_bits = 1;
  while ((1ll << p.rank_bits) < n) p.rank_bits++;
  int tile_bits = 1;
  while ((1ll << tile_bits) < (long long)p.tiles_x * p.tiles_y) tile_bits++;
  p.deep = s->dof_radius * w / 1000.0;  // src/Render3D.cpp:1103-1104
  p.neutral = pow(10, s->dof_focus / 10.0 - 16.0);
  const int tiles = p.tiles_x * p.tiles_y;

  // splat order on the host (sequential by definition: sort + shuffle)
  std::vector<float> hz;
  const float *zh = pl->z;
  if (on_device) {
    hz.resize(n);
    HIPCHECK(c, hipMemcpyAsync(hz.data(), pl->z, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(c, hipStreamSynchronize(st));
    zh = hz.data();
  }
  std::vector<uint32_t> rank(n);
  dof_order(zh, w, h, p.neutral, p.deep, s->jitter_seed, rank.data());

  // per-pixel device buffers: rank, tile counts, offsets, segment starts,
  // copies of the input image and alpha (and z, output for host planes)
  size_t scan_tmp = 0;
  HIPCHECK(c, dof_scan(nullptr, scan_tmp, nullptr, nullptr, n, st));
  const size_t bytes = (size_t)n * (4 + 8 + 8 + 4 + 6 + 2 + 6 + 2) + ((size_t)tiles + 1) * 8 + scan_tmp + 64 * 12;
  if (bytes > c->d_dof_bytes) {
    if (c->d_dof) HIPCHECK(c, hipFree(c->d_dof));
    c->d_dof = nullptr;
    c->d_dof_bytes = 0;
    HIPCHECK(c, hipMalloc(&c->d_dof, bytes));
    c->d_dof_bytes = bytes;
  }
  char *base = (char *)c->d_dof;
  size_t off = 0;
  auto take = [&](size_t sz) { void *ptr = base + off; off += (sz + 63) / 64 * 64; return ptr; };
  uint32_t *d_rank = (uint32_t *)take((size_t)n * 4);
  unsigned long long *d_cnt = (unsigned long long *)take((size_t)n * 8);
  unsigned long long *d_off = (unsigned long long *)take((size_t)n * 8);
  float *d_z = (float *)take((size_t)n * 4);
  uint16_t *d_img_in = (uint16_t *)take((size_t)n * 6);
  uint16_t *d_alpha_in = (uint16_t *)take((size_t)n * 2);
  uint16_t *d_img_out = (uint16_t *)take((size_t)n * 6);
  uint16_t *d_alpha_out = (uint16_t *)take((size_t)n * 2);
  unsigned long long *d_seg = (unsigned long long *)take(((size_t)tiles + 1) * 8);
  void *d_scan_tmp = take(scan_tmp);
  const hipMemcpyKind in_kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const float *z = on_device ? pl->z : d_z;
  HIPCHECK(c, hipMemcpyAsync(d_rank, rank.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
  if (!on_device) HIPCHECK(c, hipMemcpyAsync(d_z, pl->z, (size_t)n * 4, hipMemcpyHostToDevice, st));
  HIPCHECK(c, hipMemcpyAsync(d_img_in, pl->image16, (size_t)n * 6, in_kind, st));
  if (pl->alpha) HIPCHECK(c, hipMemcpyAsync(d_alpha_in, pl->alpha, (size_t)n * 2, in_kind, st));
  HIPCHECK(c, dof_count(p, z, d_cnt, st));
  HIPCHECK(c, dof_scan(d_scan_tmp, scan_tmp, d_cnt, d_off, n, st));
  unsigned long long last[2] = {0, 0};
  HIPCHECK(c, hipMemcpyAsync(&last[0], d_off + n - 1, 8, hipMemcpyDeviceToHost, st));
  HIPCHECK(c, hipMemcpyAsync(&last[1], d_cnt + n - 1, 8, hipMemcpyDeviceToHost, st));
  HIPCHECK(c, hipStreamSynchronize(st));
  const unsigned long long pairs = last[0] + last[1];
  if (pairs > 0x7fffffffull) return c->fail(MBR_ERR_INVALID_ARG, "mbr_post_dof: too many splat-tile pairs");
  uint16_t *img_out = on_device ? pl->image16 : d_img_out;
  uint16_t *alpha_out = pl->alpha ? (on_device ? pl->alpha : d_alpha_out) : nullptr;
  const uint16_t *alpha_in = pl->alpha ? d_alpha_in : nullptr;
  // (tile, rank) keys and source values, double-buffered for the sort
  const int np = (int)pairs;
  size_t sort_tmp = 0;
  HIPCHECK(c, dof_sort(nullptr, sort_tmp, nullptr, nullptr, nullptr, nullptr, np, p.rank_bits + tile_bits, st));
  const size_t pbytes = (size_t)np * (8 + 8 + 4 + 4) + sort_tmp + 64 * 6;
  if (pbytes > c->d_dof_pairs_bytes) {
    if (c->d_dof_pairs) HIPCHECK(c, hipFree(c->d_dof_pairs));
    c->d_dof_pairs = nullptr;
    c->d_dof_pairs_bytes = 0;
    HIPCHECK(c, hipMalloc(&c->d_dof_pairs, pbytes));
    c->d_dof_pairs_bytes = pbytes;
  }
  base = (char *)c->d_dof_pairs;
  off = 0;
  unsigned long long *k0 = (unsigned long long *)take((size_t)np * 8);
  unsigned long long *k1 = (unsigned long long *)take((size_t)np * 8);
  uint32_t *v0 = (uint32_t *)take((size_t)np * 4);
  uint32_t *v1 = (uint32_t *)take((size_t)np * 4);
  void *d_sort_tmp = take(sort_tmp);
  if (np > 0) {
    HIPCHECK(c, dof_fill(p, z, d_rank, d_off, k0, v0, st));
    HIPCHECK(c, dof_sort(d_sort_tmp, sort_tmp, k0, k1, v0, v1, np, p.rank_bits + tile_bits, st));
  }
  HIPCHECK(c, dof_fold(p, z, k1, np, d_seg, v1, d_img_in, alpha_in, img_out, alpha_out, st));
  if (!on_device) {
    HIPCHECK(c, hipMemcpyAsync(pl->image16, d_img_out, (size_t)n * 6, hipMemcpyDeviceToHost, st));
    if (pl->alpha) HIPCHECK(c, hipMemcpyAsync(pl->alpha, d_alpha_out, (size_t)n * 2, hipMemcpyDeviceToHost, st));
    HIPCHECK(c, hipStreamSynchronize(st));
  }
  return MBR_OK;
}

int mbr_render(mbr_ctx *c, const mbr_render_opts *opts, mbr_framebuffer *fb, mbr_stats *stats) {
  if (!c || !fb) return MBR_ERR_INVALID_ARG;
  if (!c->has_scene) return c->fail(MBR_ERR_NO_SCENE, "mbr_set_scene was not called");
  const size_t n = (size_t)fb->width * fb->height;
  // planes: rgb 3 x f32, z f32, colour 3 x u8, alpha u16, opacity u16
  const size_t bytes = n * (4 * 4 + 3 + 2 * 2) + 64 * 9;
  HIPCHECK(c, hipSetDevice(c->device));
  if (bytes > c->d_planes_bytes) {
    if (c->d_planes) (void)hipFree(c->d_planes);
    c->d_planes = nullptr;
    HIPCHECK(c, hipMalloc(&c->d_planes, bytes));
    c->d_planes_bytes = bytes;
  }
  char *base = (char *)c->d_planes;
  size_t off = 0;
  auto take = [&](size_t sz) { void *p = base + off; off += (sz + 63) / 64 * 64; return p; };
  mbr_framebuffer dfb;
  dfb.width = fb->width;
  dfb.height = fb->height;
  for (int i = 0; i < 3; i++) dfb.rgb[i] = fb->rgb[i] ? (float *)take(n * 4) : nullptr;
  dfb.z = fb->z ? (float *)take(n * 4) : nullptr;
  for (int i = 0; i < 3; i++) dfb.colour[i] = fb->colour[i] ? (uint8_t *)take(n) : nullptr;
  dfb.alpha = fb->alpha ? (uint16_t *)take(n * 2) : nullptr;
  dfb.opacity = fb->opacity ? (uint16_t *)take(n * 2) : nullptr;
  int rc = mbr_render_device(c, opts, &dfb, stats);
  if (rc != MBR_OK) return rc;
  hipStream_t st = (opts && opts->stream) ? (hipStream_t)opts->stream : c->stream;
  int row0 = opts ? opts->row_begin : 0, row1 = opts ? opts->row_end : 0;
  if (row1 <= 0 || row1 > fb->height) row1 = fb->height;
  if (row0 < 0) row0 = 0;
  const size_t o0 = (size_t)row0 * fb->width, cnt = (size_t)(row1 - row0) * fb->width;
  for (int i = 0; i < 3; i++) {
    if (fb->rgb[i]) HIPCHECK(c, hipMemcpyAsync(fb->rgb[i] + o0, dfb.rgb[i] + o0, cnt * 4, hipMemcpyDeviceToHost, st));
    if (fb->colour[i])
      HIPCHECK(c, hipMemcpyAsync(fb->colour[i] + o0, dfb.colour[i] + o0, cnt, hipMemcpyDeviceToHost, st));
  }
  if (fb->z) HIPCHECK(c, hipMemcpyAsync(fb->z + o0, dfb.z + o0, cnt * 4, hipMemcpyDeviceToHost, st));
  if (fb->alpha) HIPCHECK(c, hipMemcpyAsync(fb->alpha + o0, dfb.alpha + o0, cnt * 2, hipMemcpyDeviceToHost, st));
  if (fb->opacity) HIPCHECK(c, hipMemcpyAsync(fb->opacity + o0, dfb.opacity + o0, cnt * 2, hipMemcpyDeviceToHost, st));
  HIPCHECK(c, hipStreamSynchronize(st));
  return MBR_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batch
namespace {

struct BatchFrame {
  Prepared p;
  long long ao_off = 0;   // first direction of the frame's AO table in the batch table
  long long rays = 0;
};

// Capture every frame's band launches into one graph: frames go to the
// `slots` chains greedily by ray count (largest first, to the least loaded
// chain), chains fork from and join back to stream s0.
int build_batch_graph(mbr_ctx *c, std::vector<BatchFrame> &frames, const mbr_framebuffer *fbs, int precision,
                      bool demoivre, int slots, int bpc, hipStream_t s0, hipGraphExec_t *exec, int *nodes) {
  const int persist = (c->cu_count > 0 ? c->cu_count : 256) * bpc;
  std::vector<int> order(frames.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return frames[a].rays > frames[b].rays; });
  std::vector<long long> load(slots, 0);
  std::vector<int> chain(frames.size(), 0);
  for (int f : order) {
    int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    chain[f] = k;
    load[k] += frames[f].rays;
  }
  hipEvent_t fork = nullptr, join[kMaxSlots] = {};
  HIPCHECK(c, hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  for (int k = 1; k < slots; k++) HIPCHECK(c, hipEventCreateWithFlags(&join[k], hipEventDisableTiming));
  hipGraph_t graph = nullptr;
  int rc = MBR_OK;
  HIPCHECK(c, hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
  int forked = 1;  // chains 1..forked-1 wait on the fork
  auto body = [&]() -> int {
    HIPCHECK(c, launch_zero(c->d_stats, ST_COUNT, s0));
    HIPCHECK(c, hipEventRecord(fork, s0));
    for (int k = 1; k < slots; k++) {
      HIPCHECK(c, hipStreamWaitEvent(c->side[k], fork, 0));
      forked = k + 1;
    }
    for (int k = 0; k < slots; k++) {
      hipStream_t st = k == 0 ? s0 : c->side[k];
      for (int f : order) {
        if (chain[f] != k) continue;
        const BatchFrame &fr = frames[f];
        const AODev ao{c->d_batch_ao_f + 4 * fr.ao_off, c->d_batch_ao_d + 4 * fr.ao_off,
                       c->d_batch_ao_rgb + 4 * fr.ao_off, fr.p.d.ao_count};
        int r = enqueue_frame(c, c->slots[k], precision, c->d_batch_sc + f, fr.p.d, kernel_family(fr.p.s), ao, 0, fr.rays,
                              frame_out(&fbs[f]), persist, st, nullptr, nullptr, nullptr, demoivre);
        if (r != MBR_OK) return r;
      }
    }
    for (int k = 1; k < slots; k++) {
      HIPCHECK(c, hipEventRecord(join[k], c->side[k]));
      HIPCHECK(c, hipStreamWaitEvent(s0, join[k], 0));
    }
    return MBR_OK;
  };
  rc = body();
  if (rc != MBR_OK) {
    // a chain left forked would keep its stream in the (failed) capture:
    // join every forked chain back before ending it (joining one twice is
    // harmless)
    for (int k = 1; k < forked; k++) {
      (void)hipEventRecord(join[k], c->side[k]);
      (void)hipStreamWaitEvent(s0, join[k], 0);
    }
  }
  hipError_t ec = hipStreamEndCapture(s0, &graph);
  if (rc == MBR_OK && ec != hipSuccess) rc = c->hip_fail(ec, "hipStreamEndCapture");
  if (rc == MBR_OK) {
    hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) rc = c->hip_fail(ei, "hipGraphInstantiate");
    size_t nn = 0;
    if (rc == MBR_OK && hipGraphGetNodes(graph, nullptr, &nn) == hipSuccess) *nodes = (int)nn;
  }
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(fork);
  for (int k = 1; k < slots; k++) (void)hipEventDestroy(join[k]);
  return rc;
}

// Launch the graph `replays` times; mean GPU ms per launch.
int time_graph(mbr_ctx *c, hipGraphExec_t exec, hipStream_t s0, int replays, double *ms) {
  hipEvent_t e0, e1;
  HIPCHECK(c, hipEventCreate(&e0));
  HIPCHECK(c, hipEventCreate(&e1));
  HIPCHECK(c, hipEventRecord(e0, s0));
  for (int i = 0; i < replays; i++) HIPCHECK(c, hipGraphLaunch(exec, s0));
  HIPCHECK(c, hipEventRecord(e1, s0));
  HIPCHECK(c, hipEventSynchronize(e1));
  float t = 0;
  HIPCHECK(c, hipEventElapsedTime(&t, e0, e1));
  *ms = t / replays;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return MBR_OK;
}

}  // namespace

extern "C" {

int mbr_render_batch(mbr_ctx *c, const mbr_scene *scenes, int n, const mbr_texture *lightmap, mbr_framebuffer *fbs,
                     const mbr_batch_opts *opts, mbr_stats *stats, mbr_batch_result *result) {
  if (!c || !scenes || !fbs || n <= 0) return MBR_ERR_INVALID_ARG;
  int precision = opts ? opts->precision : MBR_PREC_FP32_FIXUP;
  bool demoivre = false;
  int rc = check_precision(c, precision, demoivre);
  if (rc != MBR_OK) return rc;
  // the batch renders in the precision every frame can take (a generic
  // formula's frame that renders in FP64 only makes the whole batch FP64)
  for (int i = 0; i < n; i++) {
    const int p = scene_precision(scenes[i], nullptr, precision);
    if (p == MBR_PREC_FP64 || (p == MBR_PREC_FP32_FIXUP && precision == MBR_PREC_FP32_MARCH)) precision = p;
  }
  HIPCHECK(c, hipSetDevice(c->device));
  // derive every scene on the host
  std::vector<BatchFrame> frames(n);
  long long ao_total = 0;
  int max_ao = 0;
  for (int i = 0; i < n; i++) {
    rc = prepare_scene(c, &scenes[i], lightmap, frames[i].p);
    if (rc != MBR_OK) return c->fail(rc, "scene " + std::to_string(i) + ": " + c->err);
    if (scene_has_lights(frames[i].p.s) || frames[i].p.s.fake_lights_enabled || frames[i].p.d.shadow_fx ||
        frames[i].p.d.refl_max > 0)
      return c->fail(MBR_ERR_UNSUPPORTED, "scene " + std::to_string(i) +
                                              ": aux lights, fake lights, volumetric lights, iteration fog and reflections "
                                              "render through mbr_set_scene, not in a batch");
    const mbr_scene &s = frames[i].p.s;
    if (fbs[i].width != s.image_width || fbs[i].height != s.image_height)
      return c->fail(MBR_ERR_INVALID_ARG, "framebuffer " + std::to_string(i) + " differs from its scene's size");
    frames[i].rays = (long long)s.image_width * s.image_height;
    frames[i].ao_off = ao_total;
    ao_total += frames[i].p.d.ao_count;
    const DevScene &d = frames[i].p.d;
    if (d.ao && !d.fast_ao) max_ao = std::max(max_ao, d.ao_count);
  }
  // upload scene blocks and AO tables
  if (n > c->batch_sc_cap) {
    if (c->d_batch_sc) (void)hipFree(c->d_batch_sc);
    c->d_batch_sc = nullptr;
    c->batch_sc_cap = 0;
    HIPCHECK(c, hipMalloc((void **)&c->d_batch_sc, (size_t)n * sizeof(DevScene)));
    c->batch_sc_cap = n;
  }
  if ((size_t)ao_total > c->batch_ao_cap) {
    void *p[] = {c->d_batch_ao_f, c->d_batch_ao_d, c->d_batch_ao_rgb};
    for (void *q : p)
      if (q) (void)hipFree(q);
    c->d_batch_ao_f = nullptr; c->d_batch_ao_d = nullptr; c->d_batch_ao_rgb = nullptr;
    c->batch_ao_cap = 0;
    HIPCHECK(c, hipMalloc((void **)&c->d_batch_ao_f, (size_t)ao_total * 4 * sizeof(float)));
    HIPCHECK(c, hipMalloc((void **)&c->d_batch_ao_d, (size_t)ao_total * 4 * sizeof(double)));
    HIPCHECK(c, hipMalloc((void **)&c->d_batch_ao_rgb, (size_t)ao_total * 4 * sizeof(int)));
    c->batch_ao_cap = (size_t)ao_total;
  }
  {
    std::vector<DevScene> ds(n);
    std::vector<float> af((size_t)ao_total * 4);
    std::vector<double> ad((size_t)ao_total * 4);
    std::vector<int> ar((size_t)ao_total * 4);
    for (int i = 0; i < n; i++) {
      ds[i] = frames[i].p.d;
      const size_t o = (size_t)frames[i].ao_off * 4;
      for (size_t j = 0; j < frames[i].p.dirs.size(); j++) {
        ad[o + j] = frames[i].p.dirs[j];
        af[o + j] = (float)frames[i].p.dirs[j];
      }
      std::copy(frames[i].p.rgb.begin(), frames[i].p.rgb.end(), ar.begin() + o);
    }
    HIPCHECK(c, hipMemcpy(c->d_batch_sc, ds.data(), (size_t)n * sizeof(DevScene), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_batch_ao_f, af.data(), af.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_batch_ao_d, ad.data(), ad.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_batch_ao_rgb, ar.data(), ar.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  hipStream_t s0 = (opts && opts->stream) ? (hipStream_t)opts->stream : c->stream;
  for (int k = 1; k < kMaxSlots; k++)
    if (!c->side[k]) HIPCHECK(c, hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking));
  const int cap = band_cap(precision, max_ao);

  // launch shapes to time (autotune) or the one to use
  std::vector<std::pair<int, int>> cands;
  int slots = opts && opts->slots ? opts->slots : (c->tuned_slots ? c->tuned_slots : 2);
  int bpc = opts && opts->blocks_per_cu ? opts->blocks_per_cu : (c->tuned_bpc ? c->tuned_bpc : 8);
  if (opts && opts->autotune) {
    for (int sl : {1, 2, 4})
      for (int b : {4, 8, 16}) cands.push_back({sl, b});
  }
  if (slots < 1 || slots > kMaxSlots || bpc < 1 || bpc > 32)
    return c->fail(MBR_ERR_INVALID_ARG, "slots must be 1..4 and blocks_per_cu 1..32");
  const int n_slots = cands.empty() ? slots : kMaxSlots;
  for (int k = 0; k < n_slots; k++) {
    rc = alloc_slot(c, c->slots[k], precision, cap, max_ao);
    if (rc != MBR_OK) return rc;
  }
  mbr_batch_result res;
  std::memset(&res, 0, sizeof res);
  unsigned long long v[ST_COUNT];
  // Build the graph of one launch shape and launch it `replays` times; when
  // a record found the overflow pool full, grow every slot's pool to what the
  // batch asked for and do it again.
  auto run_shape = [&](int sl_n, int b_n, int reps, bool warm, double *ms) -> int {
    for (int attempt = 0;; attempt++) {
      const auto t0 = std::chrono::steady_clock::now();
      hipGraphExec_t ex = nullptr;
      int nodes = 0;
      int r2 = build_batch_graph(c, frames, fbs, precision, demoivre, sl_n, b_n, s0, &ex, &nodes);
      if (r2 != MBR_OK) return r2;
      res.ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      res.graph_nodes = nodes;
      hipError_t e = warm ? hipGraphLaunch(ex, s0) : hipSuccess;
      if (e == hipSuccess) r2 = time_graph(c, ex, s0, reps, ms);
      else r2 = c->hip_fail(e, "hipGraphLaunch");
      (void)hipGraphExecDestroy(ex);
      if (r2 != MBR_OK) return r2;
      HIPCHECK(c, hipMemcpy(v, c->d_stats, sizeof v, hipMemcpyDeviceToHost));
      r2 = consume_cancel(c);
      if (r2 != MBR_OK) return r2;
      if (v[ST_POOL_MISS] == 0) return MBR_OK;
      if (attempt + 1 >= kPoolAttempts) return c->fail(MBR_ERR_HIP, "overflow pool still full after growing it");
      for (int k = 0; k < kMaxSlots; k++) {
        if (c->slots[k].prec < 0) continue;
        r2 = grow_pool(c, c->slots[k], v[ST_OVF_RAYS], v[ST_OVF_CHUNKS]);
        if (r2 != MBR_OK) return r2;
      }
    }
  };
  if (!cands.empty()) {
    double best = 1e300;
    for (size_t i = 0; i < cands.size(); i++) {
      double ms = 0;
      rc = run_shape(cands[i].first, cands[i].second, 1, true, &ms);
      if (rc != MBR_OK) return rc;
      if (i < 16) res.tuned_ms[i] = ms;
      if (ms < best) {
        best = ms;
        slots = cands[i].first;
        bpc = cands[i].second;
      }
    }
    res.tuned_candidates = (int)cands.size();
    c->tuned_slots = slots;
    c->tuned_bpc = bpc;
  }
  const int replays = opts && opts->replays > 0 ? opts->replays : 1;
  rc = run_shape(slots, bpc, replays, false, &res.ms_per_launch);
  if (rc != MBR_OK) return rc;
  res.slots = slots;
  res.blocks_per_cu = bpc;
  if (result) *result = res;
  if (stats) {
    long long rays = 0;
    for (auto &f : frames) rays += f.rays;
    fill_stats(stats, v, rays, n);
    stats->ms_total = res.ms_per_launch;
  }
  return MBR_OK;
}

}  // extern "C"

namespace {

// The adjoint pass of mbr_adjoint (grad_rgb, grad) and mbr_adjoint_jacobian
// (jac): the FP64 march of the current scene with refined hits, then the
// adjoint kernel per band.
int run_adjoint(mbr_ctx *c, const mbr_render_opts *opts, const float *const grad_rgb[3], double grad[9],
                uint8_t *skipped, float *jac, mbr_stats *stats) {
  if (!c->has_scene) return c->fail(MBR_ERR_NO_SCENE, "mbr_set_scene was not called");
  const mbr_scene &s = c->scene;
  const DevScene &h = c->host_sc;
  if (s.formula != MBR_FORMULA_TRIG_OPTIM) return c->fail(MBR_ERR_UNSUPPORTED, "mbr_adjoint: formula must be 2");
  if (h.n_lights > 0) return c->fail(MBR_ERR_UNSUPPORTED, "mbr_adjoint: aux lights have no gradient here");
  if (!s.hit_refine) return c->fail(MBR_ERR_UNSUPPORTED, "mbr_adjoint: the scene needs hit_refine = 1");
  if (h.shadow || h.ao || h.shadow_fx || h.refl_max > 0 || s.glow_intensity != 0.0 || s.fog_density != 0.0 || h.post_fog ||
      s.perspective_type != MBR_PERSP_THREE_POINT || s.tiles != 1 || s.constant_DE_threshold || s.interior_mode ||
      s.limits_enabled || s.iter_thresh || h.fd.prim_plane || h.fd.prim_box || h.fd.prim_ibox ||
      h.fd.prim_sphere || h.fd.prim_isphere || h.fd.prim_water || h.fd.tglad_fold || h.fd.sph_fold || h.fd.ifs_fold)
    return c->fail(MBR_ERR_UNSUPPORTED, "mbr_adjoint: shadows, AO, volumetric lights, glow, fog, post fog, fish-eye, tiles, "
                                        "constant threshold, interior mode, limits, iteration threshold, "
                                        "primitives and pre-folds have no gradient here");
  hipStream_t st = (opts && opts->stream) ? (hipStream_t)opts->stream : c->stream;
  HIPCHECK(c, hipSetDevice(c->device));
  Slot &sl = c->slots[0];
  int rc = alloc_slot(c, sl, MBR_PREC_FP64, kBandRays, 0);
  if (rc != MBR_OK) return rc;
  if (!c->d_grad) HIPCHECK(c, hipMalloc((void **)&c->d_grad, 16 * sizeof(double)));
  AdjParams p;
  p.power = s.power;
  for (int i = 0; i < 3; i++) p.vp[i] = s.vp[i];
  p.alpha = s.alpha;
  p.beta = s.beta;
  p.gamma = s.gamma;
  p.light_alpha = s.main_light_alpha;
  p.light_beta = s.main_light_beta;
  p.zoom_persp = s.zoom / s.persp;
  const int persist = c->cu_count > 0 ? c->cu_count * 8 : 2048;
  const long long last = (long long)s.image_width * s.image_height;
  const float *gR = grad_rgb ? grad_rgb[0] : nullptr, *gG = grad_rgb ? grad_rgb[1] : nullptr,
              *gB = grad_rgb ? grad_rgb[2] : nullptr;
  unsigned long long v[ST_COUNT];
  for (int attempt = 0;; attempt++) {
    HIPCHECK(c, launch_zero(c->d_stats, ST_COUNT, st));
    HIPCHECK(c, launch_zero((unsigned long long *)c->d_grad, 16, st));
    for (long long band_first = 0; band_first < last; band_first += sl.band.cap) {
      const int band_rays = (int)std::min<long long>(sl.band.cap, last - band_first);
      HIPCHECK(c, launch_band_reset(sl.band, st));
      HIPCHECK(c, launch_march(MBR_PREC_FP64, c->d_sc, sl.band, (int)band_first, band_rays, s.formula,
                               c->host_sc.plain, persist, c->d_stats, st));
      HIPCHECK(c, launch_march_fix(MBR_PREC_FP64, c->d_sc, sl.band, (int)band_first, s.formula, persist / 2,
                                   c->d_stats, st));
      HIPCHECK(c, launch_adjoint(c->d_sc, sl.band, (int)band_first, band_rays, p, gR, gG, gB,
                                 jac ? nullptr : c->d_grad, skipped, jac, persist, st));
    }
    HIPCHECK(c, hipMemcpyAsync(v, c->d_stats, sizeof v, hipMemcpyDeviceToHost, st));
    HIPCHECK(c, hipStreamSynchronize(st));
    rc = consume_cancel(c);
    if (rc != MBR_OK) return rc;
    if (v[ST_POOL_MISS] == 0) break;
    if (attempt + 1 >= kPoolAttempts) return c->fail(MBR_ERR_HIP, "overflow pool still full after growing it");
    rc = grow_pool(c, sl, v[ST_OVF_RAYS], v[ST_OVF_CHUNKS]);
    if (rc != MBR_OK) return rc;
  }
  if (grad) HIPCHECK(c, hipMemcpy(grad, c->d_grad, 9 * sizeof(double), hipMemcpyDeviceToHost));
  if (stats) fill_stats(stats, v, last, (int)((last + sl.band.cap - 1) / sl.band.cap));
  return MBR_OK;
}

}  // namespace

extern "C" {

int mbr_adjoint(mbr_ctx *c, const mbr_render_opts *opts, const float *const grad_rgb[3], double grad[9],
                uint8_t *skipped, mbr_stats *stats) {
  if (!c || !grad_rgb || !grad || !grad_rgb[0] || !grad_rgb[1] || !grad_rgb[2]) return MBR_ERR_INVALID_ARG;
  return run_adjoint(c, opts, grad_rgb, grad, skipped, nullptr, stats);
}

int mbr_adjoint_jacobian(mbr_ctx *c, const mbr_render_opts *opts, float *jac, uint8_t *skipped, mbr_stats *stats) {
  if (!c || !jac) return MBR_ERR_INVALID_ARG;
  return run_adjoint(c, opts, nullptr, nullptr, skipped, jac, stats);
}

}  // extern "C"
