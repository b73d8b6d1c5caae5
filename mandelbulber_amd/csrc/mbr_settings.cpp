// mbr_settings.cpp — .fract settings loader of the render core.
//
// Mirrors the reference's settings model: LoadSettings / LoadSettings2 /
// LoadOneSetting / LoadSettingsPost (src/settings.cpp:611-1195) and
// GetPaletteFromString (:94-119), NewPalette (src/image.cpp:28-55).  The
// default values are those of the reference defaults file share/defaults,
// which LoadSettings reads before every scene (src/settings.cpp:611-626);
// keys absent from that file start at zero, as the reference's static
// parameter struct does.  Keys that configure parts of the application
// outside the render core (post effects, file names, GUI state) are accepted
// and ignored so that reference scenes load without warnings.
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mbr.h"

namespace {

enum KeyType { KT_DOUBLE, KT_INT, KT_UINT, KT_IGNORE, KT_PALETTE, KT_STRING };

struct KeyDef {
  std::string name;
  KeyType type;
  size_t offset;
  double scale;  // applied to doubles (degrees -> radians)
};

constexpr double kPi = 3.14159265358979323846;
constexpr double kDeg = kPi / 180.0;

#define OFF(field) offsetof(mbr_scene, field)
#define OFF_RGB(field, c) (offsetof(mbr_scene, field) + offsetof(mbr_rgb, c))
#define OFF_A(field, i, T) (offsetof(mbr_scene, field) + sizeof(T) * (i))

std::vector<KeyDef> build_table() {
  std::vector<KeyDef> t;
  auto D = [&](const std::string &n, size_t o, double s = 1.0) { t.push_back({n, KT_DOUBLE, o, s}); };
  auto I = [&](const std::string &n, size_t o) { t.push_back({n, KT_INT, o, 1.0}); };
  auto G = [&](const std::string &n) { t.push_back({n, KT_IGNORE, 0, 1.0}); };
  auto RGB = [&](const std::string &prefix, size_t r, size_t g, size_t b) {
    I(prefix + "R", r); I(prefix + "G", g); I(prefix + "B", b);
  };

  G("Mandelbulber");
  I("image_width", OFF(image_width)); I("image_height", OFF(image_height));
  I("tiles", OFF(tiles)); I("tile_no", OFF(tile_index)); I("frame_no", OFF(frame_no));
  D("x_min", OFF_A(limit_min, 0, double)); D("x_max", OFF_A(limit_max, 0, double));
  D("y_min", OFF_A(limit_min, 1, double)); D("y_max", OFF_A(limit_max, 1, double));
  D("z_min", OFF_A(limit_min, 2, double)); D("z_max", OFF_A(limit_max, 2, double));
  D("view_point_x", OFF_A(vp, 0, double)); D("view_point_y", OFF_A(vp, 1, double));
  D("view_point_z", OFF_A(vp, 2, double));
  D("angle_alfa", OFF(alpha), kDeg); D("angle_beta", OFF(beta), kDeg); D("angle_gamma", OFF(gamma), kDeg);
  D("zoom", OFF(zoom)); D("perspective", OFF(persp));
  I("formula", OFF(formula)); D("power", OFF(power)); I("N", OFF(N)); I("minN", OFF(minN));
  D("fractal_constant_factor", OFF(constant_factor));
  D("quality", OFF(quality)); D("smoothness", OFF(smoothness));
  I("julia_mode", OFF(julia_mode));
  D("julia_a", OFF_A(julia, 0, double)); D("julia_b", OFF_A(julia, 1, double)); D("julia_c", OFF_A(julia, 2, double));
  I("tglad_folding_mode", OFF(tglad_folding_mode));
  D("folding_limit", OFF(folding_limit)); D("folding_value", OFF(folding_value));
  I("spherical_folding_mode", OFF(spherical_folding_mode));
  D("spherical_folding_fixed", OFF(spherical_folding_fixed)); D("spherical_folding_min", OFF(spherical_folding_min));
  I("IFS_folding_mode", OFF(ifs_folding_mode));
  I("iteration_threshold_mode", OFF(iter_thresh));
  G("analityc_DE_mode");  // derived from the formula (src/interface.cpp:690-693)
  D("DE_factor", OFF(DE_factor));
  D("brightness", OFF(brightness)); D("contrast", OFF(contrast)); D("gamma", OFF(image_gamma)); I("hdr", OFF(hdr));
  G("ambient");
  D("reflect", OFF(reflect));
  D("shadows_intensity", OFF(direct_light)); D("shadows_cone_angle", OFF(shadow_cone_angle));
  D("ambient_occlusion", OFF(global_ilum)); I("ambient_occlusion_quality", OFF(ao_quality));
  D("ambient_occlusion_fast_tune", OFF(fast_ao_tune));
  D("shading", OFF(shading)); D("specular", OFF(specular)); D("glow_intensity", OFF(glow_intensity));
  RGB("glow_color_1_", OFF_RGB(glow_color1, R), OFF_RGB(glow_color1, G), OFF_RGB(glow_color1, B));
  RGB("glow_color_2_", OFF_RGB(glow_color2, R), OFF_RGB(glow_color2, G), OFF_RGB(glow_color2, B));
  RGB("background_color_1_", OFF_RGB(background_color1, R), OFF_RGB(background_color1, G), OFF_RGB(background_color1, B));
  RGB("background_color_2_", OFF_RGB(background_color2, R), OFF_RGB(background_color2, G), OFF_RGB(background_color2, B));
  RGB("background_color_3_", OFF_RGB(background_color3, R), OFF_RGB(background_color3, G), OFF_RGB(background_color3, B));
  RGB("fog_colour_1_", OFF_RGB(fog_colour1, R), OFF_RGB(fog_colour1, G), OFF_RGB(fog_colour1, B));
  RGB("fog_colour_2_", OFF_RGB(fog_colour2, R), OFF_RGB(fog_colour2, G), OFF_RGB(fog_colour2, B));
  RGB("fog_colour_3_", OFF_RGB(fog_colour3, R), OFF_RGB(fog_colour3, G), OFF_RGB(fog_colour3, B));
  I("textured_background", OFF(textured_background)); I("background_as_fuldome", OFF(background_as_fulldome));
  I("shadows_enabled", OFF(shadow)); I("ambient_occlusion_enabled", OFF(global_illumination));
  I("fast_ambient_occlusion_mode", OFF(fast_global_illumination));
  I("fractal_color", OFF(coloring_enabled));
  G("coloring_random_seed"); G("coloring_saturation");  // consumed by LoadSettingsPost
  D("coloring_speed", OFF(coloring_speed)); D("coloring_palette_offset", OFF(palette_offset));
  I("slow_shading", OFF(slow_shading)); I("limits_enabled", OFF(limits_enabled));
  I("post_fog_enabled", OFF(post_fog_enabled)); D("post_fog_visibility", OFF(fog_visibility));
  G("post_fog_visibility_front");
  RGB("post_fog_color_", OFF_RGB(post_fog_color, R), OFF_RGB(post_fog_color, G), OFF_RGB(post_fog_color, B));
  I("post_SSAO_enabled", OFF(ssao_enabled)); I("post_SSAO_quality", OFF(ssao_quality));
  I("post_DOF_enabled", OFF(dof_enabled)); D("post_DOF_focus", OFF(dof_focus)); D("post_DOF_radius", OFF(dof_radius));
  D("main_light_intensity", OFF(main_light_intensity));
  D("main_light_alfa", OFF(main_light_alpha)); D("main_light_beta", OFF(main_light_beta));
  RGB("main_light_colour_", OFF_RGB(main_light_colour, R), OFF_RGB(main_light_colour, G), OFF_RGB(main_light_colour, B));
  D("aux_light_intensity", OFF(aux_light_intensity)); I("aux_light_random_seed", OFF(aux_light_random_seed));
  I("aux_light_number", OFF(aux_light_number));
  D("aux_light_max_dist", OFF(aux_light_max_dist));
  D("aux_light_distribution_radius", OFF(aux_light_distribution_radius));
  D("aux_light_visibility", OFF(aux_light_visibility));
  D("aux_light_random_center_X", OFF_A(aux_light_random_center, 0, double));
  D("aux_light_random_center_Y", OFF_A(aux_light_random_center, 1, double));
  D("aux_light_random_center_Z", OFF_A(aux_light_random_center, 2, double));
  for (int i = 1; i <= 4; i++) {
    std::string p = "aux_light_predefined_" + std::to_string(i) + "_";
    const size_t pos = OFF(aux_light_pre) + sizeof(double) * 3 * (i - 1);
    D(p + "x", pos); D(p + "y", pos + sizeof(double)); D(p + "z", pos + 2 * sizeof(double));
    D(p + "intensity", OFF_A(aux_light_pre_intensity, i - 1, double));
    const size_t col = OFF(aux_light_pre_colour) + sizeof(mbr_rgb) * (i - 1);
    RGB(p + "colour_", col + offsetof(mbr_rgb, R), col + offsetof(mbr_rgb, G), col + offsetof(mbr_rgb, B));
    I(p + "enabled", OFF_A(aux_light_pre_enabled, i - 1, int));
  }
  D("IFS_scale", OFF(ifs_scale));
  D("IFS_rot_alfa", OFF_A(ifs_rotation, 0, double)); D("IFS_rot_beta", OFF_A(ifs_rotation, 1, double));
  D("IFS_rot_gamma", OFF_A(ifs_rotation, 2, double));
  D("IFS_offsetX", OFF_A(ifs_offset, 0, double)); D("IFS_offsetY", OFF_A(ifs_offset, 1, double));
  D("IFS_offsetZ", OFF_A(ifs_offset, 2, double));
  D("IFS_edgeX", OFF_A(ifs_edge, 0, double)); D("IFS_edgeY", OFF_A(ifs_edge, 1, double));
  D("IFS_edgeZ", OFF_A(ifs_edge, 2, double));
  I("IFS_absX", OFF_A(ifs_abs, 0, int)); I("IFS_absY", OFF_A(ifs_abs, 1, int)); I("IFS_absZ", OFF_A(ifs_abs, 2, int));
  I("IFS_menger_sponge_mode", OFF(ifs_menger_sponge_mode));
  for (int i = 0; i < 9; i++) {
    const std::string p = "IFS_" + std::to_string(i) + "_";
    for (int c = 0; c < 3; c++) {
      D(p + "xyz"[c], OFF(ifs_direction) + sizeof(double) * (3 * i + c));
      D(p + (c == 0 ? "alfa" : c == 1 ? "beta" : "gamma"), OFF(ifs_angles) + sizeof(double) * (3 * i + c));
    }
    D(p + "distance", OFF_A(ifs_distance, i, double));
    D(p + "intensity", OFF_A(ifs_intensity, i, double));
    I(p + "enabled", OFF_A(ifs_enabled, i, int));
  }
  I("start_frame", OFF(start_frame)); I("end_frame", OFF(end_frame));
  I("frames_per_keyframe", OFF(frames_per_keyframe)); I("hybrid_cyclic", OFF(hybrid_cyclic));
  for (int i = 1; i <= 5; i++) {
    I("hybrid_formula_" + std::to_string(i), OFF_A(hybrid_formula, i - 1, int));
    I("hybrid_iterations_" + std::to_string(i), OFF_A(hybrid_iterations, i - 1, int));
    D("hybrid_power_" + std::to_string(i), OFF_A(hybrid_power, i - 1, double));
  }
  I("fish_eye", OFF(perspective_type)); I("fish_eye_180cut", OFF(fish_eye_cut));
  G("stereo_enabled"); G("stereo_eye_distance");
  D("mandelbox_scale", OFF(mb_scale)); D("mandelbox_folding_limit", OFF(mb_folding_limit));
  D("mandelbox_folding_value", OFF(mb_folding_value)); D("mandelbox_folding_min_radius", OFF(mb_min_radius));
  D("mandelbox_folding_fixed_radius", OFF(mb_fixed_radius)); D("mandelbox_sharpness", OFF(mb_sharpness));
  D("mandelbox_offset_X", OFF_A(mb_offset, 0, double)); D("mandelbox_offset_Y", OFF_A(mb_offset, 1, double));
  D("mandelbox_offset_Z", OFF_A(mb_offset, 2, double));
  D("mandelbox_color_R", OFF(mb_color_R)); D("mandelbox_color_X", OFF(mb_color_X));
  D("mandelbox_color_Y", OFF(mb_color_Y)); D("mandelbox_color_Z", OFF(mb_color_Z));
  D("mandelbox_color_Sp1", OFF(mb_color_Sp1)); D("mandelbox_color_Sp2", OFF(mb_color_Sp2));
  I("mandelbox_rotation_enabled", OFF(mb_rotations_enabled));
  I("mandelbox_fold_mode", OFF(mb_fold_mode)); D("mandelbox_solid", OFF(mb_solid)); D("mandelbox_melt", OFF(mb_melt));
  const char *comp[3] = {"alfa", "beta", "gamma"};
  const char *axis[3] = {"X", "Y", "Z"};
  for (int c = 0; c < 3; c++)
    D(std::string("mandelbox_rotation_main_") + comp[c], OFF_A(mb_rotation_main, c, double), kDeg);
  for (int fold = 0; fold < 2; fold++)
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++)
        D(std::string("mandelbox_rotation_") + axis[a] + std::to_string(fold + 1) + "_" + comp[c],
          offsetof(mbr_scene, mb_rotation) + sizeof(double) * (fold * 9 + a * 3 + c), kDeg);
  D("view_distance_max", OFF(view_distance_max)); D("view_distance_min", OFF(view_distance_min));
  I("interior_mode", OFF(interior_mode)); I("linear_DE_mode", OFF(linear_DE_mode));
  I("constant_DE_threshold", OFF(constant_DE_threshold));
  D("FoldingIntPow_folding_factor", OFF(folding_int_pow_factor)); D("FoldingIntPow_z_factor", OFF(folding_int_pow_z_factor)); G("dynamic_DE_correction");
  I("penetrating_lights", OFF(penetrating_lights));
  I("raytraced_reflections", OFF(raytraced_reflections)); I("reflections_max", OFF(reflections_max));
  D("mandelbox_vary_scale_vary", OFF(mb_vary_scale_vary)); D("mandelbox_vary_fold", OFF(mb_vary_fold));
  D("mandelbox_vary_minr", OFF(mb_vary_minr)); D("mandelbox_vary_rpower", OFF(mb_vary_rpower));
  D("mandelbox_vary_wadd", OFF(mb_vary_wadd)); D("c_add", OFF(c_add));
  D("volumetric_fog_density", OFF(fog_density));
  D("volumetric_fog_colour_1_distance", OFF(fog_colour1_distance));
  D("volumetric_fog_colour_2_distance", OFF(fog_colour2_distance));
  D("volumetric_fog_distance_factor", OFF(fog_distance_factor));
  I("iteration_fog_enable", OFF(iter_fog_enabled)); D("iteration_fog_opacity", OFF(iter_fog_opacity));
  D("iteration_fog_opacity_trim", OFF(iter_fog_opacity_trim));
  G("volumetric_light_intensity"); G("volumetric_light_quality");
  for (int i = 0; i < 5; i++) {
    I("volumetric_light_enabled_" + std::to_string(i), OFF_A(volumetric_light_enabled, i, int));
    D("volumetric_light_intensity_" + std::to_string(i), OFF_A(volumetric_light_intensity, i, double));
  }
  I("primitive_only_plane", OFF(primitive_only_plane));
  I("primitive_plane_enabled", OFF(plane_enabled));
  D("primitive_plane_centre_X", OFF_A(plane_centre, 0, double)); D("primitive_plane_centre_Y", OFF_A(plane_centre, 1, double));
  D("primitive_plane_centre_Z", OFF_A(plane_centre, 2, double));
  D("primitive_plane_normal_X", OFF_A(plane_normal, 0, double)); D("primitive_plane_normal_Y", OFF_A(plane_normal, 1, double));
  D("primitive_plane_normal_Z", OFF_A(plane_normal, 2, double));
  RGB("primitive_plane_colour_", OFF_RGB(plane_colour, R), OFF_RGB(plane_colour, G), OFF_RGB(plane_colour, B));
  I("primitive_box_enabled", OFF(box_enabled));
  D("primitive_box_centre_X", OFF_A(box_centre, 0, double)); D("primitive_box_centre_Y", OFF_A(box_centre, 1, double));
  D("primitive_box_centre_Z", OFF_A(box_centre, 2, double));
  D("primitive_box_size_X", OFF_A(box_size, 0, double)); D("primitive_box_size_Y", OFF_A(box_size, 1, double));
  D("primitive_box_size_Z", OFF_A(box_size, 2, double));
  RGB("primitive_box_colour_", OFF_RGB(box_colour, R), OFF_RGB(box_colour, G), OFF_RGB(box_colour, B));
  I("primitive_invertedBox_enabled", OFF(inv_box_enabled));
  D("primitive_invertedBox_centre_X", OFF_A(inv_box_centre, 0, double));
  D("primitive_invertedBox_centre_Y", OFF_A(inv_box_centre, 1, double));
  D("primitive_invertedBox_centre_Z", OFF_A(inv_box_centre, 2, double));
  D("primitive_invertedBox_size_X", OFF_A(inv_box_size, 0, double));
  D("primitive_invertedBox_size_Y", OFF_A(inv_box_size, 1, double));
  D("primitive_invertedBox_size_Z", OFF_A(inv_box_size, 2, double));
  RGB("primitive_invertedBox_colour_", OFF_RGB(inv_box_colour, R), OFF_RGB(inv_box_colour, G), OFF_RGB(inv_box_colour, B));
  I("primitive_sphere_enabled", OFF(sphere_enabled));
  D("primitive_sphere_centre_X", OFF_A(sphere_centre, 0, double)); D("primitive_sphere_centre_Y", OFF_A(sphere_centre, 1, double));
  D("primitive_sphere_centre_Z", OFF_A(sphere_centre, 2, double));
  D("primitive_sphere_radius", OFF(sphere_radius));
  RGB("primitive_sphere_colour_", OFF_RGB(sphere_colour, R), OFF_RGB(sphere_colour, G), OFF_RGB(sphere_colour, B));
  I("primitive_invertedSphere_enabled", OFF(inv_sphere_enabled));
  D("primitive_invertedSphere_centre_X", OFF_A(inv_sphere_centre, 0, double));
  D("primitive_invertedSphere_centre_Y", OFF_A(inv_sphere_centre, 1, double));
  D("primitive_invertedSphere_centre_Z", OFF_A(inv_sphere_centre, 2, double));
  D("primitive_invertedSphere_radius", OFF(inv_sphere_radius));
  RGB("primitive_invertedSphere_colour_", OFF_RGB(inv_sphere_colour, R), OFF_RGB(inv_sphere_colour, G),
      OFF_RGB(inv_sphere_colour, B));
  I("primitive_water_enabled", OFF(water_enabled));
  D("primitive_water_level", OFF(water_height)); D("primitive_water_amplitude", OFF(water_amplitude));
  D("primitive_water_length", OFF(water_length)); I("primitive_water_iterations", OFF(water_iterations));
  D("primitive_water_rotation", OFF(water_rotation)); D("primitive_water_anim_speed", OFF(water_anim_speed));
  RGB("primitive_water_colour_", OFF_RGB(water_colour, R), OFF_RGB(water_colour, G), OFF_RGB(water_colour, B));
  D("primitive_plane_reflect", OFF(plane_reflect)); D("primitive_box_reflect", OFF(box_reflect));
  D("primitive_invertedBox_reflect", OFF(inv_box_reflect)); D("primitive_sphere_reflect", OFF(sphere_reflect));
  D("primitive_invertedSphere_reflect", OFF(inv_sphere_reflect)); D("primitive_water_reflect", OFF(water_reflect));
  I("fake_lights_enabled", OFF(fake_lights_enabled));
  I("fake_lights_min_iter", OFF(fake_lights_min_iter)); I("fake_lights_max_iter", OFF(fake_lights_max_iter));
  D("fake_lights_visibility", OFF(fake_lights_visibility));
  D("fake_lights_visibility_size", OFF(fake_lights_visibility_size));
  D("fake_lights_intensity", OFF(fake_lights_intensity));
  D("fake_lights_orbit_trap_X", OFF_A(fake_lights_orbit_trap, 0, double));
  D("fake_lights_orbit_trap_Y", OFF_A(fake_lights_orbit_trap, 1, double));
  D("fake_lights_orbit_trap_Z", OFF_A(fake_lights_orbit_trap, 2, double));
  G("ocl_custom_DE_mode"); G("ocl_custom_formula_name"); G("ocl_delta_DE_step"); G("ocl_DOF_method");
  for (int i = 0; i < 15; i++) G("ocl_custom_par_" + std::to_string(i));
  t.push_back({"file_destination", KT_STRING, OFF(file_destination), 1.0});
  t.push_back({"file_animation_path", KT_STRING, OFF(file_animation_path), 1.0});
  t.push_back({"file_keyframes", KT_STRING, OFF(file_keyframes), 1.0});
  for (const char *k : {"file_background", "file_envmap", "file_lightmap"}) G(k);
  t.push_back({"palette", KT_PALETTE, 0, 1.0});
  // render-core settings with no reference key
  D("bailout", OFF(bailout));
  I("jitter_mode", OFF(jitter_mode));
  t.push_back({"jitter_seed", KT_UINT, OFF(jitter_seed), 1.0});
  I("hit_refine", OFF(hit_refine));
  return t;
}

const std::vector<KeyDef> &table() {
  static const std::vector<KeyDef> t = build_table();
  return t;
}

// atof2 (src/settings.cpp:66-79): locale-independent strtod
double parse_double(const char *s) {
  return std::strtod(s, nullptr);
}

// GetPaletteFromString (src/settings.cpp:94-119): 256 hex RRGGBB words
bool parse_palette(mbr_rgb *palette, const char *str) {
  int index = 0;
  size_t len = std::strlen(str);
  size_t i = 0;
  while (index < 256) {
    while (i < len && str[i] == ' ') i++;
    if (i >= len) break;
    unsigned colour = 0;
    std::sscanf(str + i, "%x", &colour);
    palette[index].R = static_cast<int>(colour / 65536);
    palette[index].G = static_cast<int>((colour / 256) % 256);
    palette[index].B = static_cast<int>(colour % 256);
    index++;
    while (i < len && str[i] != ' ') i++;
  }
  return index == 256;
}

// Random(max) with the reference's generator (src/common_math.cpp:25-32)
int random_max(int max) { return std::rand() % (max + 1); }

// NewPalette (src/image.cpp:28-55)
void new_palette(mbr_rgb *p, double saturation) {
  for (int i = 0; i < 255; i++) {
    int Y = static_cast<int>((random_max(255) - 128) / (1.0 + saturation));
    int R = static_cast<int>(Y + 128 + (random_max(255) - 128) * saturation);
    int G = static_cast<int>(Y + 128 + (random_max(255) - 128) * saturation);
    int B = static_cast<int>(Y + 128 + (random_max(255) - 128) * saturation);
    p[i].R = R < 0 ? 0 : (R > 255 ? 255 : R);
    p[i].G = G < 0 ? 0 : (G > 255 ? 255 : G);
    p[i].B = B < 0 ? 0 : (B > 255 ? 255 : B);
  }
  p[255] = {255, 255, 255};
  p[254] = {0, 0, 0};
  p[253] = {0, 20, 40};
}

struct LoadState {
  bool palette_loaded = false;
  int coloring_seed = 123456;
  double colour_saturation = 1.0;
};

int set_one(mbr_scene *s, const char *key, const char *value, LoadState *st) {
  if (!std::strcmp(key, "coloring_random_seed")) { if (st) st->coloring_seed = std::atoi(value); return MBR_OK; }
  if (!std::strcmp(key, "coloring_saturation")) { if (st) st->colour_saturation = parse_double(value); return MBR_OK; }
  for (const KeyDef &k : table()) {
    if (k.name != key) continue;
    char *base = reinterpret_cast<char *>(s);
    switch (k.type) {
      case KT_DOUBLE: *reinterpret_cast<double *>(base + k.offset) = parse_double(value) * k.scale; break;
      case KT_INT: *reinterpret_cast<int *>(base + k.offset) = std::atoi(value); break;
      case KT_UINT: *reinterpret_cast<uint32_t *>(base + k.offset) = static_cast<uint32_t>(std::strtoul(value, nullptr, 10)); break;
      case KT_PALETTE: {
        bool ok = parse_palette(s->palette, value);
        if (st && ok) st->palette_loaded = true;
        break;
      }
      case KT_STRING: {
        char *dst = base + k.offset;
        std::strncpy(dst, value, sizeof(s->file_destination) - 1);
        dst[sizeof(s->file_destination) - 1] = 0;
        break;
      }
      case KT_IGNORE: break;
    }
    return MBR_OK;
  }
  return MBR_ERR_INVALID_ARG;
}

// one settings file in the reference's "key value;" format (LoadSettings2,
// src/settings.cpp:628-657): key = first word, value = text up to ';'
int load_file(const char *path, mbr_scene *s, LoadState *st) {
  FILE *f = std::fopen(path, "r");
  if (!f) return MBR_ERR_IO;
  char line[8192];
  while (std::fgets(line, sizeof line, f)) {
    char *p = line;
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') p++;
    if (!*p) continue;
    char key[128];
    int n = 0;
    while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r' && n < 127) key[n++] = *p++;
    key[n] = 0;
    while (*p == ' ') p++;
    char *end = std::strchr(p, ';');
    std::string value = end ? std::string(p, end) : std::string(p);
    while (!value.empty() && (value.back() == '\n' || value.back() == '\r')) value.pop_back();
    if (set_one(s, key, value.c_str(), st) != MBR_OK)
      std::fprintf(stderr, "Warning! Unknown parameter: %s %s\n", key, value.c_str());
  }
  std::fclose(f);
  return MBR_OK;
}

}  // namespace

extern "C" int mbr_scene_defaults(mbr_scene *s) {
  if (!s) return MBR_ERR_INVALID_ARG;
  std::memset(s, 0, sizeof *s);
  s->abi_version = MBR_ABI_VERSION;
  s->image_width = 800; s->image_height = 600; s->tiles = 1;
  for (int i = 0; i < 3; i++) { s->limit_min[i] = -10; s->limit_max[i] = 10; }
  s->alpha = -20 * kDeg; s->beta = 30 * kDeg; s->gamma = 0;
  s->zoom = 2.5; s->persp = 0.5;
  s->formula = MBR_FORMULA_TRIG_OPTIM; s->power = 9; s->N = 250; s->minN = 1;
  s->constant_factor = 1; s->quality = 1; s->smoothness = 1;
  s->folding_limit = 1; s->folding_value = 2;
  s->spherical_folding_fixed = 1; s->spherical_folding_min = 0.5;
  s->DE_factor = 1;
  s->reflect = 0; s->direct_light = 0.69999999999999996; s->global_ilum = 1;
  s->ao_quality = 4; s->fast_ao_tune = 1; s->shading = 1; s->specular = 1; s->glow_intensity = 1;
  s->glow_color1 = {40984, 44713, 49490}; s->glow_color2 = {57192, 60888, 62408};
  s->background_color1 = {0, 38306, 65535}; s->background_color2 = {65535, 65535, 65535};
  s->background_color3 = {0, 10000, 500};
  s->fog_colour1 = {0, 0, 0}; s->fog_colour2 = {0, 30000, 65535}; s->fog_colour3 = {65535, 65535, 65535};
  s->shadow = 1; s->global_illumination = 0; s->fast_global_illumination = 0;
  s->coloring_enabled = 1; s->coloring_speed = 1; s->palette_offset = 0;
  s->fog_visibility = 20; s->post_fog_color = {59399, 61202, 65535};
  s->main_light_intensity = 1;
  s->main_light_alpha = -0.78539816339744828; s->main_light_beta = 0.78539816339744828;
  s->main_light_colour = {65535, 65535, 65535};
  s->mb_scale = 2; s->mb_folding_limit = 1; s->mb_folding_value = 2; s->mb_min_radius = 0.5;
  s->mb_fixed_radius = 1;
  s->mb_color_R = 0; s->mb_color_X = 0.029999999999999999; s->mb_color_Y = 0.050000000000000003;
  s->mb_color_Z = 0.070000000000000007; s->mb_color_Sp1 = 0.20000000000000001;
  s->mb_color_Sp2 = 0.20000000000000001;
  s->view_distance_max = 50; s->view_distance_min = 1.0000000000000001e-15;
  s->penetrating_lights = 1; s->reflections_max = 5;
  s->fog_density = 0; s->fog_colour1_distance = 1; s->fog_colour2_distance = 2; s->fog_distance_factor = 1;
  s->plane_normal[2] = -1;
  s->plane_colour = s->box_colour = s->inv_box_colour = s->sphere_colour = s->inv_sphere_colour = {20000, 20000, 20000};
  for (int i = 0; i < 3; i++) { s->box_size[i] = 2; s->inv_box_size[i] = 10; }
  s->sphere_radius = 1.5; s->inv_sphere_radius = 5;
  s->water_amplitude = 0.02; s->water_length = 0.20000000000000001; s->water_iterations = 5;
  s->water_colour = {0, 5000, 10000};
  // palette: the defaults file holds 256 zero entries
  s->bailout = 100.0;                // src/fractal.cpp:1153
  s->jitter_mode = MBR_JITTER_HASH;
  s->jitter_seed = 1234;             // aux_light_random_seed, the reference's srand seed
  s->brightness = 1.0;
  s->contrast = 1.0;
  s->image_gamma = 1.0;
  s->hdr = 0;
  std::strncpy(s->file_destination, "images/image", sizeof(s->file_destination) - 1);
  s->start_frame = 0;                // share/defaults
  s->end_frame = 1000;
  s->frames_per_keyframe = 100;
  std::strncpy(s->file_animation_path, "paths/path.txt", sizeof(s->file_animation_path) - 1);
  std::strncpy(s->file_keyframes, "keyframes/keyframe", sizeof(s->file_keyframes) - 1);
  for (int i = 0; i < 5; i++) {       // share/defaults: formula 0 0 0 0 2, iterations 1, power 2
    s->hybrid_formula[i] = i == 4 ? 2 : 0;
    s->hybrid_iterations[i] = 1;
    s->hybrid_power[i] = 2.0;
  }
  s->ifs_scale = 2.0;                // share/defaults: IFS_scale 2, IFS_offsetX 1, IFS_%d_x 1, intensity 1
  s->ifs_offset[0] = 1.0;
  for (int i = 0; i < 9; i++) {
    s->ifs_direction[i][0] = 1.0;
    s->ifs_intensity[i] = 1.0;
  }
  s->mb_sharpness = 3.0;
  s->folding_int_pow_factor = 2.0;
  s->folding_int_pow_z_factor = 5.0;
  // share/defaults:300-302, 326-330
  s->mb_fold_mode = 1;
  s->mb_solid = 1.0;
  s->mb_melt = 0.0;
  s->mb_vary_scale_vary = 0.1;
  s->mb_vary_fold = 1.0;
  s->mb_vary_minr = 0.5;
  s->mb_vary_rpower = 1.0;
  s->mb_vary_wadd = 0.0;
  // fake lights (src/settings.cpp:520-528)
  s->fake_lights_intensity = 1.0;
  s->fake_lights_visibility = 1.0;
  s->fake_lights_visibility_size = 5.0;
  s->fake_lights_orbit_trap[0] = 2.0;
  s->fake_lights_min_iter = 1;
  s->fake_lights_max_iter = 2;
  s->c_add = -1.3;
  s->ssao_enabled = 1;               // share/defaults: post_SSAO_enabled 1, quality 20
  s->ssao_quality = 20;
  s->dof_enabled = 0;
  s->dof_focus = 21.699999999999999;
  s->dof_radius = 10.0;
  // primitive reflection coefficients: the reference writes them only when
  // they differ from 0 (water: 0.7), so a file without them means these
  s->water_reflect = 0.7;
  // share/defaults: volumetric lights and iteration fog
  for (int i = 0; i < 5; i++) s->volumetric_light_intensity[i] = 100.0;
  s->iter_fog_opacity = 1000.0;
  s->iter_fog_opacity_trim = 4.0;
  // share/defaults: aux light block
  s->aux_light_intensity = 1.0;
  s->aux_light_random_seed = 1234;
  s->aux_light_max_dist = 0.10000000000000001;
  s->aux_light_distribution_radius = 3.0;
  s->aux_light_visibility = 1.0;
  const double pre[4][4] = {{3, -3, -3, 1.3}, {-3, -3, 0, 1}, {-3, 3, -1, 3}, {0, -1, 3, 2}};
  const int pre_col[4][3] = {{45761, 53633, 59498}, {62875, 55818, 50083}, {64884, 64928, 48848}, {52704, 62492, 45654}};
  for (int i = 0; i < 4; i++) {
    for (int k = 0; k < 3; k++) s->aux_light_pre[i][k] = pre[i][k];
    s->aux_light_pre_intensity[i] = pre[i][3];
    s->aux_light_pre_colour[i] = mbr_rgb{pre_col[i][0], pre_col[i][1], pre_col[i][2]};
  }
  return MBR_OK;
}

namespace {

// CMorph::CatmullRom for one value (src/morph.cpp:49-109)
double catmull_rom(double v1, double v2, double v3, double v4, double factor) {
  const double factor2 = factor * factor, factor3 = factor2 * factor;
  bool logarithmic = false, negative = false;
  if ((v1 > 0 && v2 > 0 && v3 > 0 && v4 > 0) || (v1 < 0 && v2 < 0 && v3 < 0 && v4 < 0)) {
    if (v1 < 0) negative = true;
    const double average = (v1 + v2 + v3 + v4) / 4.0;
    if (average > 0) {
      const double deviation = (std::fabs(v2 - v1) + std::fabs(v3 - v2) + std::fabs(v4 - v3)) / average;
      if (deviation > 0.1) {
        v1 = std::log(std::fabs(v1));
        v2 = std::log(std::fabs(v2));
        v3 = std::log(std::fabs(v3));
        v4 = std::log(std::fabs(v4));
        logarithmic = true;
      }
    }
  }
  double value = 0.5 * ((2 * v2) + (-v1 + v3) * factor + (2 * v1 - 5 * v2 + 4 * v3 - v4) * factor2 +
                        (-v1 + 3 * v2 - 3 * v3 + v4) * factor3);
  if (logarithmic) value = negative ? -std::exp(value) : std::exp(value);
  if (value > 1e20) value = 1e20;
  if (value < -1e20) value = 1e20;  // (the reference clamps negative overflow to +1e20 too)
  if (std::fabs(value) < 1e-20) value = 0.0;
  return value;
}

}  // namespace

extern "C" int mbr_keyframe_scene(const mbr_scene *keys, int n_keys, int frames_per_key, int frame, mbr_scene *s) {
  if (!keys || !s || n_keys < 1 || frames_per_key < 1 || frame < 0) return MBR_ERR_INVALID_ARG;
  const int key = frame / frames_per_key;
  const double factor = (frame % frames_per_key) / (double)frames_per_key;
  auto at = [&](int k) -> const mbr_scene & { return keys[k]; };
  // support values as CMorph picks them (clamped at both ends)
  const int k1 = key >= 1 ? key - 1 : key;
  const int k2 = key < n_keys ? key : n_keys - 1;
  const int k3 = key < n_keys - 1 ? key + 1 : n_keys - 1;
  const int k4 = key < n_keys - 2 ? key + 2 : n_keys - 1;
  if (k1 >= n_keys) return MBR_ERR_INVALID_ARG;
  auto spline = [&](size_t offset) {
    auto v = [&](int k) { return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(&at(k)) + offset); };
    *reinterpret_cast<double *>(reinterpret_cast<char *>(s) + offset) = catmull_rom(v(k1), v(k2), v(k3), v(k4), factor);
  };
  for (const KeyDef &k : table())
    if (k.type == KT_DOUBLE) spline(k.offset);
  // N is a double of sFractalD in the reference, read back as int
  const double n = catmull_rom(at(k1).N, at(k2).N, at(k3).N, at(k4).N, factor);
  s->N = (int)n;
  if (s->zoom < 1e-15) s->zoom = 1e-15;  // src/Render3D.cpp:2287
  return MBR_OK;
}

extern "C" int mbr_set_setting(mbr_scene *s, const char *key, const char *value) {
  if (!s || !key || !value) return MBR_ERR_INVALID_ARG;
  return set_one(s, key, value, nullptr);
}

extern "C" int mbr_load_settings(const char *path, const char *const *overrides, int n_overrides, mbr_scene *s) {
  if (!s || !path) return MBR_ERR_INVALID_ARG;
  mbr_scene_defaults(s);
  LoadState st;
  int rc = load_file(path, s, &st);
  if (rc != MBR_OK) return rc;
  // -o key=value overrides (src/settings.cpp:659-676)
  for (int i = 0; i < n_overrides; i++) {
    if (!overrides || !overrides[i]) continue;
    const char *o = overrides[i];
    const char *eq = std::strpbrk(o, "= \t");
    if (!eq) { std::fprintf(stderr, "Warning! Bad override string: %s\n", o); continue; }
    std::string key(o, eq);
    std::string value(eq + 1);
    if (set_one(s, key.c_str(), value.c_str(), &st) != MBR_OK)
      std::fprintf(stderr, "Warning! Unknown parameter: %s %s\n", key.c_str(), value.c_str());
  }
  // LoadSettingsPost (src/settings.cpp:1185-1195)
  if (!st.palette_loaded) {
    std::srand(static_cast<unsigned>(st.coloring_seed));
    new_palette(s->palette, st.colour_saturation);
  }
  return MBR_OK;
}
