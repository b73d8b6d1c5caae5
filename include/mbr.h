/*
 * mbr.h — C ABI of the MI355X render core for Mandelbulber 1.21 scenes.
 *
 * This boundary replaces the reference's OpenCL render seam, class CclSupport
 * (src/cl_support.hpp:59-156, dummy fallback :158-170), which Render() calls
 * at src/Render3D.cpp:719 and :1160-1251, and the CPU MainThread pool
 * (src/Render3D.cpp:203-706) that Render() falls back to.  Parameters cross
 * it as one FP64 POD (mbr_scene) that carries the fields of sParamRender
 * (src/fractparams.h:18-146) and sFractal (src/fractal.h:251-293) used by the
 * render core.  Output is the cImage buffer set (src/cimage.hpp:82-89,
 * :157-162) as caller-owned planes.
 *
 * Plain C: no C++ or torch types cross this boundary.  Every entry point
 * returns an int status (MBR_OK = 0, negative = error); mbr_last_error()
 * gives the message.  A context is not thread-safe; it owns one HIP stream
 * unless the caller supplies one per call.
 */
#ifndef MBR_H_
#define MBR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBR_ABI_VERSION 3

/* largest reflections_max the render core accepts (mirrored rays per pixel) */
#define MBR_MAX_REFLECTIONS 16

/* ---- status codes (reference analogue: checkErr -> bool + CPU fallback,
 *      src/cl_support.cpp:109-135, src/Render3D.cpp:719) ---- */
enum mbr_status {
  MBR_OK = 0,
  MBR_ERR_INVALID_ARG = -1,
  MBR_ERR_NO_DEVICE = -2,
  MBR_ERR_HIP = -3,
  MBR_ERR_UNSUPPORTED = -4, /* scene uses a feature outside the render core */
  MBR_ERR_NO_SCENE = -5,
  MBR_ERR_CANCELLED = -6,
  MBR_ERR_IO = -7,
  MBR_ERR_OUT_OF_MEMORY = -8
};

/* ---- formula enum: values of enumFractalFormula (src/fractal.h:25-65) ---- */
enum mbr_formula {
  MBR_FORMULA_NONE = 0,
  MBR_FORMULA_TRIG_DE = 1,       /* Mandelbulb with the analytic derivative vector */
  MBR_FORMULA_TRIG_OPTIM = 2,    /* Mandelbulb, "trigonometric" in the GUI */
  MBR_FORMULA_FAST_TRIG = 3,
  MBR_FORMULA_HYPERCOMPLEX = 4,
  MBR_FORMULA_QUATERNION = 5,
  MBR_FORMULA_MINUS_FAST_TRIG = 6,
  MBR_FORMULA_MENGER_SPONGE = 7,
  MBR_FORMULA_TGLAD = 8,         /* Mandelbox */
  MBR_FORMULA_KALEIDOSCOPIC = 10, /* kaleidoscopic IFS */
  MBR_FORMULA_HYBRID = 12,       /* hybrid_formula_1..5 sequence */
  MBR_FORMULA_XENODREAMBUIE = 11,
  MBR_FORMULA_MANDELBULB2 = 13,
  MBR_FORMULA_MANDELBULB3 = 14,
  MBR_FORMULA_MANDELBULB4 = 15,
  MBR_FORMULA_FOLDING_INT_POW2 = 16,
  MBR_FORMULA_SMOOTH_MANDELBOX = 17,
  MBR_FORMULA_MANDELBOX_VARY_SCALE_4D = 18,
  MBR_FORMULA_AEXION = 19,
  MBR_FORMULA_BENESI = 20,
  MBR_FORMULA_BRISTORBROT = 21,
  /* hybrid building blocks (src/fractal.cpp:984-1103) */
  MBR_FORMULA_INVERT_X = 22, MBR_FORMULA_INVERT_Y = 23, MBR_FORMULA_INVERT_Z = 24, MBR_FORMULA_INVERT_R = 25,
  MBR_FORMULA_SPHERICAL_FOLD = 26, MBR_FORMULA_POW_XYZ = 27,
  MBR_FORMULA_SCALE_X = 28, MBR_FORMULA_SCALE_Y = 29, MBR_FORMULA_SCALE_Z = 30,
  MBR_FORMULA_OFFSET_X = 31, MBR_FORMULA_OFFSET_Y = 32, MBR_FORMULA_OFFSET_Z = 33,
  MBR_FORMULA_ANGLE_MULTIPLY_X = 34, MBR_FORMULA_ANGLE_MULTIPLY_Y = 35, MBR_FORMULA_ANGLE_MULTIPLY_Z = 36,
  MBR_FORMULA_GENERALIZED_FOLD_BOX = 37
  /* all but 2, 7 and 8 render through the generic FP64 path (analytic DE for
   * 1, 17, 18 and 37, delta DE for the rest, as the reference's analitycDE):
   * every precision mode renders them in FP64 */
};

/* enumPerspectiveType (src/common_math.h) */
enum mbr_perspective {
  MBR_PERSP_THREE_POINT = 0,
  MBR_PERSP_FISH_EYE = 1,
  MBR_PERSP_EQUIRECTANGULAR = 2
};

/* enumObjectType (src/fractal.h:82-91) */
enum mbr_object {
  MBR_OBJ_FRACTAL = 0, MBR_OBJ_PLANE = 1, MBR_OBJ_WATER = 2, MBR_OBJ_SPHERE = 3,
  MBR_OBJ_SPHERE_INV = 4, MBR_OBJ_BOX = 5, MBR_OBJ_BOX_INV = 6
};

/* March step jitter.  The reference multiplies every step by
 * 1 - Random(1000)/10000 with glibc rand() shared by all threads
 * (src/Render3D.cpp:129-134, src/common_math.cpp:25-32).  A GPU cannot
 * replay that sequential stream, so the render core draws the same integer
 * in [0,1000] from a counter-based hash of (seed, frame, pixel, step). */
enum mbr_jitter {
  MBR_JITTER_NONE = 0,
  MBR_JITTER_HASH = 1
};

/* Arithmetic of the render.
 *   FP64        FP64 everywhere: the reference's CPU arithmetic.
 *   FP32_FIXUP  primary-ray march (ray origins, scan, every hit decision and
 *               the binary search) in FP64, so hits and depths are those of
 *               the FP64 march; normals, shadow, AO, colouring, fog in FP32.
 *   FP32_MARCH  as FP32_FIXUP, but the march iterates in FP32 until FP32
 *               error could change a hit decision, continues in FP64 from
 *               there, and re-marches in FP64 the rays whose result the FP32
 *               part could have changed.
 *   FP64_DEMOIVRE  FP64 everywhere, but the march, its FP64 continuation
 *               and the AO rays of the power-8 Mandelbulb iterate with the
 *               trig-free integer-power form of the FP32 modes (de Moivre,
 *               hardware rsq) instead of the reference's asin/atan2/pow/
 *               sin/cos; normals, shadow and colouring keep the trig form.
 *               Not bit-exact with the reference; it separates what the
 *               iteration form gains from what FP32 arithmetic gains.  Other
 *               scenes render as FP64. */
enum mbr_precision {
  MBR_PREC_FP64 = 0,
  MBR_PREC_FP32_FIXUP = 1,
  MBR_PREC_FP32_MARCH = 2,
  MBR_PREC_FP64_DEMOIVRE = 3
};

typedef struct mbr_rgb { int R, G, B; } mbr_rgb; /* sRGB (src/cimage.hpp:39-44) */

/*
 * Scene POD.  Field names follow sParamRender / sFractal; each comment names
 * the .fract key that sets it (LoadOneSetting, src/settings.cpp:700-1183).
 * Angles are radians (the loader converts the degree keys, settings.cpp:787).
 */
typedef struct mbr_scene {
  int abi_version;                 /* = MBR_ABI_VERSION */

  /* image */
  int image_width, image_height;   /* image_width, image_height */
  int tiles, tile_index;           /* tiles, tile_no */

  /* camera */
  double vp[3];                    /* view_point_x/y/z */
  double alpha, beta, gamma;       /* angle_alfa/beta/gamma (radians) */
  double zoom;                     /* zoom */
  double persp;                    /* perspective (fov factor) */
  int perspective_type;            /* fish_eye: enum mbr_perspective */
  int fish_eye_cut;                /* fish_eye_180cut */
  double view_distance_min;        /* view_distance_min */
  double view_distance_max;        /* view_distance_max */

  /* ray marching */
  double quality;                  /* quality (DE threshold factor) */
  double DE_factor;                /* DE_factor */
  double smoothness;               /* smoothness */

  /* fractal */
  int formula;                     /* formula: enum mbr_formula */
  int N;                           /* N */
  int minN;                        /* minN */
  double power;                    /* power */
  double bailout;                  /* escape radius of trig formulas; the
                                      reference hard-codes 1e2
                                      (src/fractal.cpp:1153) */
  double constant_factor;          /* fractal_constant_factor */
  int julia_mode;                  /* julia_mode */
  double julia[3];                 /* julia_a/b/c */
  int interior_mode;               /* interior_mode */
  int iter_thresh;                 /* iteration_threshold_mode */
  int constant_DE_threshold;       /* constant_DE_threshold */
  int limits_enabled;              /* limits_enabled */
  double limit_min[3];             /* x_min, y_min, z_min */
  double limit_max[3];             /* x_max, y_max, z_max */
  int tglad_folding_mode;          /* tglad_folding_mode (pre-fold) */
  double folding_limit;            /* folding_limit */
  double folding_value;            /* folding_value */
  int spherical_folding_mode;      /* spherical_folding_mode (pre-fold) */
  double spherical_folding_fixed;  /* spherical_folding_fixed */
  double spherical_folding_min;    /* spherical_folding_min */
  int ifs_folding_mode;            /* IFS_folding_mode: the IFS vectors fold z before every iteration */

  /* Mandelbox (sFractalMandelbox, src/fractal.h:172-179) */
  double mb_scale;                 /* mandelbox_scale */
  double mb_folding_limit;         /* mandelbox_folding_limit */
  double mb_folding_value;         /* mandelbox_folding_value */
  double mb_min_radius;            /* mandelbox_folding_min_radius */
  double mb_fixed_radius;          /* mandelbox_folding_fixed_radius */
  double mb_offset[3];             /* mandelbox_offset_X/Y/Z */
  double mb_color_R, mb_color_X, mb_color_Y, mb_color_Z;  /* mandelbox_color_* */
  double mb_color_Sp1, mb_color_Sp2;
  int mb_rotations_enabled;        /* mandelbox_rotation_enabled */
  double mb_rotation_main[3];      /* mandelbox_rotation_main_alfa/beta/gamma */
  double mb_rotation[2][3][3];     /* mandelbox_rotation_{X,Y,Z}{1,2}_{alfa,beta,gamma} */

  /* primitives (sFractalPrimitives[D], src/fractal.h:181-215) */
  int primitive_only_plane;        /* primitive_only_plane */
  int plane_enabled;   double plane_centre[3], plane_normal[3];
  int box_enabled;     double box_centre[3], box_size[3];
  int inv_box_enabled; double inv_box_centre[3], inv_box_size[3];
  int sphere_enabled;  double sphere_centre[3], sphere_radius;
  int inv_sphere_enabled; double inv_sphere_centre[3], inv_sphere_radius;
  int water_enabled;   double water_height, water_amplitude, water_length,
                              water_rotation, water_anim_speed;
  int water_iterations;
  int frame_no;                    /* frame_no (water animation phase) */
  mbr_rgb plane_colour, box_colour, inv_box_colour, sphere_colour,
          inv_sphere_colour, water_colour;

  /* lighting and shading (sImageAdjustments, src/cimage.hpp:46-62) */
  int shadow;                      /* shadows_enabled */
  int global_illumination;         /* ambient_occlusion_enabled */
  int fast_global_illumination;    /* fast_ambient_occlusion_mode */
  int ao_quality;                  /* ambient_occlusion_quality */
  int slow_shading;                /* slow_shading */
  int penetrating_lights;          /* penetrating_lights */
  double main_light_alpha;         /* main_light_alfa (radians) */
  double main_light_beta;          /* main_light_beta (radians) */
  double shadow_cone_angle;        /* shadows_cone_angle (degrees) */
  double main_light_intensity;     /* main_light_intensity */
  double direct_light;             /* shadows_intensity */
  double shading;                  /* shading */
  double specular;                 /* specular */
  double global_ilum;              /* ambient_occlusion */
  double fast_ao_tune;             /* ambient_occlusion_fast_tune */
  double reflect;                  /* reflect (with raytraced_reflections: the fractal's reflection
                                      coefficient; without: environment mapping, not supported) */
  double glow_intensity;           /* glow_intensity */
  mbr_rgb main_light_colour;       /* main_light_colour_R/G/B */
  mbr_rgb glow_color1, glow_color2;/* glow_color_1_*, glow_color_2_* */
  int coloring_enabled;            /* fractal_color */
  double coloring_speed;           /* coloring_speed */
  double palette_offset;           /* coloring_palette_offset */
  mbr_rgb palette[256];            /* palette */

  /* background (BackgroundShader, src/shaders.cpp:114-183) */
  int textured_background;         /* textured_background (must be 0) */
  mbr_rgb background_color1, background_color2, background_color3;

  /* fog (VolumetricShader, src/shaders.cpp:185-478) */
  int post_fog_enabled;            /* post_fog_enabled */
  double fog_visibility;           /* post_fog_visibility */
  mbr_rgb post_fog_color;          /* post_fog_color_R/G/B */
  double fog_density;              /* volumetric_fog_density */
  double fog_colour1_distance;     /* volumetric_fog_colour_1_distance */
  double fog_colour2_distance;     /* volumetric_fog_colour_2_distance */
  double fog_distance_factor;      /* volumetric_fog_distance_factor */
  mbr_rgb fog_colour1, fog_colour2, fog_colour3; /* fog_colour_{1,2,3}_* */

  /* auxiliary lights: random ones, predefined ones (parameters below) */
  int aux_light_number;            /* aux_light_number */
  int aux_light_pre_enabled[4];    /* aux_light_predefined_N_enabled */
  /* volumetric lights (main light, aux lights 1-4) and iteration fog:
   * intensities and opacities below */
  int volumetric_light_enabled[5]; /* volumetric_light_enabled_N */
  int iter_fog_enabled;            /* iteration_fog_enable */
  /* features outside the render core: a scene that enables any of them is
   * rejected with MBR_ERR_UNSUPPORTED */
  int fake_lights_enabled;         /* fake_lights_enabled */
  int raytraced_reflections;       /* raytraced_reflections */
  int reflections_max;             /* reflections_max */

  /* settings of the render core that have no .fract key */
  int jitter_mode;                 /* enum mbr_jitter */
  uint32_t jitter_seed;
  int hit_refine;                  /* 1: after the binary search, Newton steps move
                                      the hit to the middle of the refinement band
                                      (distance = (1 - 0.005 quality) x threshold),
                                      so depth is a smooth function of the scene
                                      parameters (the differentiable renders);
                                      0: the reference's binary-search result */

  /* image output (sImageAdjustments, src/fractparams.h; CompileImage,
   * src/cimage.cpp:132-213) */
  double brightness;               /* brightness */
  double contrast;                 /* contrast */
  double image_gamma;              /* gamma */
  int hdr;                         /* hdr */
  char file_destination[512];      /* file_destination (path prefix of the images) */

  /* animation (src/Render3D.cpp:1995-2310) */
  int start_frame;                 /* start_frame */
  int end_frame;                   /* end_frame */
  int frames_per_keyframe;         /* frames_per_keyframe */
  char file_animation_path[512];   /* file_animation_path (flight path: vp x y z, alfa beta gamma per line) */
  char file_keyframes[512];        /* file_keyframes (prefix of the keyframe .fract files) */

  /* formula parameters of the delta-DE and fold formulas */
  int linear_DE_mode;              /* linear_DE_mode (delta-DE formulas: 0.5 r / dr) */
  double mb_sharpness;             /* mandelbox_sharpness (smoothMandelbox) */
  double folding_int_pow_factor;   /* FoldingIntPow_folding_factor */
  double folding_int_pow_z_factor; /* FoldingIntPow_z_factor */
  double c_add;                    /* c_add (aexion) */

  /* kaleidoscopic IFS (sFractalIFS, src/fractal.h; angles in radians, as
   * the settings files store them) */
  double ifs_scale;                /* IFS_scale */
  double ifs_rotation[3];          /* IFS_rot_alfa / beta / gamma */
  double ifs_offset[3];            /* IFS_offsetX/Y/Z */
  double ifs_edge[3];              /* IFS_edgeX/Y/Z */
  int ifs_abs[3];                  /* IFS_absX/Y/Z */
  int ifs_menger_sponge_mode;      /* IFS_menger_sponge_mode */
  double ifs_direction[9][3];      /* IFS_%d_x/y/z (normalised by mbr_set_scene) */
  double ifs_angles[9][3];         /* IFS_%d_alfa/beta/gamma */
  double ifs_distance[9];          /* IFS_%d_distance */
  double ifs_intensity[9];         /* IFS_%d_intensity */
  int ifs_enabled[9];              /* IFS_%d_enabled */

  /* hybrid sequence (CreateFormulaSequence, src/callbacks.cpp:1666-1703) */
  int hybrid_formula[5];           /* hybrid_formula_1..5 (0 = none) */
  int hybrid_iterations[5];        /* hybrid_iterations_1..5 */
  double hybrid_power[5];          /* hybrid_power_1..5 */
  int hybrid_cyclic;               /* hybrid_cyclic */

  /* post effects on the compiled image (src/image.cpp; order in the
   * -nogui render: CompileImage, SSAO, DOF, src/Render3D.cpp:1094-1105) */
  int ssao_enabled;                /* post_SSAO_enabled */
  int ssao_quality;                /* post_SSAO_quality */
  int dof_enabled;                 /* post_DOF_enabled */
  double dof_focus;                /* post_DOF_focus */
  double dof_radius;               /* post_DOF_radius */

  /* auxiliary lights: placement (PlaceRandomLights, src/shaders.cpp:1134-1212),
   * surface shading (AuxLightsShader / LightShading / AuxShadow, :997-1132)
   * and their glow in VolumetricShader (:241-300).  aux_light_number and
   * aux_light_pre_enabled sit with the switches above. */
  double aux_light_intensity;           /* aux_light_intensity */
  int aux_light_random_seed;            /* aux_light_random_seed (srand seed of the placement) */
  double aux_light_max_dist;            /* aux_light_max_dist */
  double aux_light_distribution_radius; /* aux_light_distribution_radius */
  double aux_light_visibility;          /* aux_light_visibility */
  double aux_light_random_center[3];    /* aux_light_random_center_X/Y/Z */
  double aux_light_pre[4][3];           /* aux_light_predefined_N_x/y/z */
  double aux_light_pre_intensity[4];    /* aux_light_predefined_N_intensity */
  mbr_rgb aux_light_pre_colour[4];      /* aux_light_predefined_N_colour_R/G/B */

  /* volumetric lights and iteration fog (VolumetricShader, src/shaders.cpp:
   * 314-341, 385-465; their effect on MainShadow / AuxShadow, :494-548,
   * :1084-1117); the switches sit with volumetric_light_enabled above */
  double volumetric_light_intensity[5]; /* volumetric_light_intensity_0..4 */
  double iter_fog_opacity;              /* iteration_fog_opacity */
  double iter_fog_opacity_trim;         /* iteration_fog_opacity_trim */

  /* raytraced reflections (MainThread's reflection loop, src/Render3D.cpp:
   * 451-560): reflections_max bounces when raytraced_reflections is on
   * (both with the switches above); the fractal reflects `reflect`, each
   * primitive its own coefficient */
  double plane_reflect;                 /* primitive_plane_reflect */
  double box_reflect;                   /* primitive_box_reflect */
  double inv_box_reflect;               /* primitive_invertedBox_reflect */
  double sphere_reflect;                /* primitive_sphere_reflect */
  double inv_sphere_reflect;            /* primitive_invertedSphere_reflect */
  double water_reflect;                 /* primitive_water_reflect */

  /* generalizedFoldBox (formula 37, src/fractal.cpp:531-749) and
   * mandelboxVaryScale4D (formula 18, :887-915); both also use the Mandelbox
   * sphere fold, scale, offset, rotation and colour factors above */
  int mb_fold_mode;                     /* mandelbox_fold_mode: fold polyhedron, 0 tetrahedron,
                                           1 cube, 2 octahedron, 3 dodecahedron, 4 octahedron +
                                           cube, 5 icosahedron, 6 hexagonal prism, 7 pentagonal
                                           prism (enumGeneralizedFoldBoxType, src/fractal.h:72-82) */
  double mb_solid;                      /* mandelbox_solid: distance of the fold planes */
  double mb_melt;                       /* mandelbox_melt */
  double mb_vary_scale_vary;            /* mandelbox_vary_scale_vary */
  double mb_vary_fold;                  /* mandelbox_vary_fold */
  double mb_vary_minr;                  /* mandelbox_vary_minr */
  double mb_vary_rpower;                /* mandelbox_vary_rpower */
  double mb_vary_wadd;                  /* mandelbox_vary_wadd */

  /* fake lights (orbit-trap lights, FakeLights src/shaders.cpp:1225-1270, and
   * their glow in VolumetricShader :302-312; fake_lights_enabled above):
   * the orbit's inverse squared distances to the trap point, summed over the
   * iterations min_iter..max_iter, light the surface and the air */
  int fake_lights_min_iter;             /* fake_lights_min_iter */
  int fake_lights_max_iter;             /* fake_lights_max_iter */
  double fake_lights_intensity;         /* fake_lights_intensity */
  double fake_lights_visibility;        /* fake_lights_visibility */
  double fake_lights_visibility_size;   /* fake_lights_visibility_size */
  double fake_lights_orbit_trap[3];     /* fake_lights_orbit_trap_X, _Y, _Z */
  int background_as_fulldome;           /* background_as_fuldome: the background texture is a
                                           fisheye dome (two hemispheres side by side) */
} mbr_scene;

/* Optional AO lightmap (file_lightmap).  The reference loads it as JPEG and
 * substitutes a white 100x100 texture when loading fails
 * (src/texture.cpp:58-74); pass NULL for that fallback. */
typedef struct mbr_texture {
  int width, height;
  const uint8_t *rgb;               /* width*height*3 bytes, row-major */
} mbr_texture;

/* Framebuffer planes (cImage layout as structure-of-arrays).  Any pointer
 * may be NULL to skip that plane.  Index = x + y*width. */
typedef struct mbr_framebuffer {
  int width, height;
  float *rgb[3];       /* imageFloat R, G, B (sRGBfloat) */
  float *z;            /* zBuffer: depth of ray 0, 1e20 on miss */
  uint8_t *colour[3];  /* colourBuffer R, G, B (sRGB8) */
  uint16_t *alpha;     /* alphaBuffer */
  uint16_t *opacity;   /* opacityBuffer */
} mbr_framebuffer;

typedef struct mbr_render_opts {
  int precision;                   /* enum mbr_precision */
  void *stream;                    /* hipStream_t; NULL = the context's stream */
  volatile const int *cancel;      /* cooperative stop flag (programClosed,
                                      src/interface.h:924) read on the host
                                      between bands; may be NULL.  To stop
                                      a frame already on the GPU, call
                                      mbr_cancel from another thread. */
  int row_begin, row_end;          /* render rows [row_begin,row_end); 0,0 = all */
} mbr_render_opts;

/* Counters of one render (the reference prints iters/s, average N, average
 * DE steps and missed-DE %, src/Render3D.cpp:638-652). */
typedef struct mbr_stats {
  double ms_total;                 /* GPU time of the render, HIP events */
  double ms_march, ms_shade, ms_finish;  /* summed over bands: frames of 2^18 or more
                                            rays run two bands at a time, so these may
                                            add up to more than ms_total */
  uint64_t rays, hits;
  uint64_t march_de_calls;         /* DE calls of primary march + bisection */
  uint64_t shade_de_calls;         /* DE calls of normals, shadow, AO, colour */
  uint64_t iterations;             /* formula iterations over all DE calls */
  uint64_t march_iterations;       /* ... of which in the march kernel */
  uint64_t shade_iterations;       /* ... of which in the shade kernel */
  uint64_t march_launches, shade_launches, finish_launches;
  uint64_t missed_de;              /* hits with dist < 0.1*thresh */
  uint64_t step_overflow_rays;     /* rays whose step record outgrew the slab
                                      (their later records use the overflow pool) */
  uint64_t fp64_redo_rays;         /* FP32_MARCH rays re-marched in FP64 because a
                                      hit or bisection decision fell inside the
                                      FP32 error bound */
  double ms_ao;                    /* part of ms_shade spent in the AO kernel */
  uint64_t ao_de_calls;            /* DE calls of the AO rays (in shade_de_calls) */
  uint64_t ao_iterations;          /* formula iterations of those calls */
  uint64_t ao_steps;               /* AO ray steps taken */
  double ms_ao_span;               /* from the first band's AO launch start to the
                                      last band's AO end (equals ms_ao on one stream) */
} mbr_stats;

typedef struct mbr_device_opts {
  int device;                      /* HIP device ordinal */
} mbr_device_opts;

typedef struct mbr_ctx mbr_ctx;

/* Library version (MBR_ABI_VERSION). */
int mbr_version(void);

/* Fill *s with the values of the reference defaults file (share/defaults),
 * which LoadSettings reads before every scene (src/settings.cpp:611-626). */
int mbr_scene_defaults(mbr_scene *s);

/* LoadSettings (src/settings.cpp:611-698): defaults, then the .fract file,
 * then "key=value" overrides (-o, src/settings.cpp:659-676), then
 * LoadSettingsPost (random palette when the file has none, :1185-1195).
 * Unknown keys are reported to stderr, as the reference does. */
int mbr_load_settings(const char *path, const char *const *overrides,
                      int n_overrides, mbr_scene *s);

/* Parse one "key value" pair into *s (LoadOneSetting, :700). */
int mbr_set_setting(mbr_scene *s, const char *key, const char *value);

/* Replaces CclSupport::InitDevice/InitFractal (cl_support.cpp:137, :231). */
int mbr_init(const mbr_device_opts *opts, mbr_ctx **out);
void mbr_destroy(mbr_ctx *ctx);
const char *mbr_last_error(const mbr_ctx *ctx);

/* Stops the context's render in progress, from any thread: the reference
 * polls programClosed before every pixel (src/Render3D.cpp:369-378) and
 * returns.  Here the flag lives in pinned host memory that the march and AO
 * kernels read while handing out work, so a frame on the GPU stops within
 * a few hundred microseconds: rays not yet marched are left empty and AO
 * rays not yet started are skipped.  The render then returns
 * MBR_ERR_CANCELLED with an incomplete frame, and the flag is cleared.  A
 * cancel made while no render runs stops the next render at once. */
int mbr_cancel(mbr_ctx *ctx);

/* Replaces Params2Cl + CclSupport::SetParams (interface.cpp:3634,
 * cl_support.cpp:484).  Copies the scene; derives the rotation matrix, ray
 * start, light vector, resolution and AO direction table as MainThread does
 * (src/Render3D.cpp:236-333) and the formula constants as
 * RecalculateIFSParams does (src/callbacks.cpp:1538-1555). */
int mbr_set_scene(mbr_ctx *ctx, const mbr_scene *scene,
                  const mbr_texture *lightmap);

/* The textures of textured_background (file_background) and of environment
 * mapping (file_envmap), decoded by the caller into 8-bit RGB as the
 * reference's cTexture holds them (src/texture.cpp).  The context keeps
 * device copies until replaced; NULL clears one.  Call before
 * mbr_set_scene: a scene with textured_background set and no background
 * texture returns MBR_ERR_UNSUPPORTED.  Replaces the reference's
 * backgroundTexture / envmapTexture loading (src/Render3D.cpp:1845-1862). */
int mbr_set_textures(mbr_ctx *ctx, const mbr_texture *background, const mbr_texture *envmap);

/* Replaces CclSupport::Render (cl_support.cpp:527-742) and the CPU
 * MainThread pool: renders the scene into caller-owned HOST planes. */
int mbr_render(mbr_ctx *ctx, const mbr_render_opts *opts,
               mbr_framebuffer *fb, mbr_stats *stats);

/* Same, with the planes in DEVICE memory (no host copies); the render is
 * enqueued on the stream and complete when the call returns. */
int mbr_render_device(mbr_ctx *ctx, const mbr_render_opts *opts,
                      mbr_framebuffer *fb, mbr_stats *stats);

/* mbr_render_device without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code
 * under HIP graph capture or a serving loop that forbids hipMalloc.  Every
 * working buffer (counters, step records, hit lists, shading planes, AO
 * intensities, overflow pool) is carved from one static 256 MiB device
 * array per GPU, whose address mbr_init takes once per device; the call
 * makes no hipMalloc or hipFree.  The caller passes the scene (mbr_set_scene,
 * which may allocate) and the DEVICE planes, as for mbr_render_device; the
 * results are the same.  The frame is rendered in bands as large as the
 * array holds for this scene; a scene needing more than 256 MiB / 16384
 * bytes per ray (very high AO quality in FP64), raytraced reflections, or a
 * frame whose long rays outgrow the array's overflow pool is rejected
 * (MBR_ERR_OUT_OF_MEMORY / MBR_ERR_UNSUPPORTED, with mbr_last_error saying
 * which).  All calls on one GPU share the array: calls that overlap on one
 * GPU (from any context or thread) race and must be serialised by the
 * caller.  No kernel timings are taken (the ms fields of stats are 0). */
int mbr_render_device_noalloc(mbr_ctx *ctx, const mbr_render_opts *opts,
                              mbr_framebuffer *fb, mbr_stats *stats);

/* Gradient of a pixel loss (no counterpart in the reference, which renders
 * only forward): given the image adjoint dL/dC as three DEVICE float planes
 * of the scene's size (dL/dR, dL/dG, dL/dB of imageFloat), writes dL/dtheta
 * to grad[9] (host) for theta = (power, vp[0], vp[1], vp[2], alpha, beta,
 * gamma, main_light_alpha, main_light_beta), angles in radians.  The image
 * is the FP64 render of the current scene, which must be a trig_optim
 * Mandelbulb with hit_refine = 1 and no shadows, AO, glow, fog, post fog,
 * primitives, limits, interior or constant-threshold mode (their
 * contributions are steps in theta); the surface colour index counts as a
 * constant.  Hit pixels whose ray meets the surface within about 3 degrees
 * of grazing (where the depth derivative diverges), whose refined hit is not
 * a root of the band-centre equation (the distance estimate jumps across
 * the band there), or whose colour tangent exceeds 1e5 per unit parameter
 * (a degenerate normal), contribute nothing; when
 * `skipped` (a DEVICE uint8 plane of the image's size, may be NULL) is
 * given, it is set to 1 on those pixels and 0 on the others the call
 * visits (pixels whose adjoint is zero are not visited).  Rays are marched
 * again (FP64) inside the call. */
int mbr_adjoint(mbr_ctx *ctx, const mbr_render_opts *opts,
                const float *const grad_rgb[3], double grad[9], uint8_t *skipped,
                mbr_stats *stats);

/* The tangents behind mbr_adjoint, pixel by pixel (no counterpart in the
 * reference): for every pixel of the current scene's FP64 image, jac (a
 * DEVICE float array of 27 * width * height) receives dC_ch/dtheta_i at
 * index (i * 3 + ch) * width * height + pixel, theta as for mbr_adjoint and
 * ch = R, G, B.  Pixels that mbr_adjoint leaves out get zeros (and 1 in
 * `skipped` when given).  mbr_adjoint's result equals the sum over pixels
 * and channels of grad_rgb times these tangents; this form serves
 * pixel-level gradient checks and Jacobian-vector products. */
int mbr_adjoint_jacobian(mbr_ctx *ctx, const mbr_render_opts *opts, float *jac,
                         uint8_t *skipped, mbr_stats *stats);

/* Batched rendering of many frames (animation frames, or unrelated scenes of
 * mixed formulas and sizes) as ONE launch of a HIP graph: the frames' band
 * launches are captured once, spread over `slots` concurrent chains, and
 * the graph is launched `replays` times.  The reference renders a flight
 * with one Render() per frame (src/Render3D.cpp:2305-2521, keyframe loop);
 * mbr_render_batch(scenes, n) is n such renders. */
typedef struct mbr_batch_opts {
  int precision;                   /* enum mbr_precision */
  void *stream;                    /* hipStream_t the graph is launched on; NULL = the context's */
  int slots;                       /* concurrent frame chains 1..4; 0 = tuned or 2 */
  int blocks_per_cu;               /* workgroups per CU of the persistent kernels; 0 = tuned or 8 */
  int autotune;                    /* 1: time every (slots, blocks_per_cu) candidate on this batch,
                                      keep the fastest in the context for later batches */
  int replays;                     /* graph launches (>= 1); the last one's frames are returned */
} mbr_batch_opts;

typedef struct mbr_batch_result {
  double ms_per_launch;            /* GPU time of one graph launch (HIP events), mean over replays */
  double ms_build;                 /* host time of capture + instantiate */
  int slots, blocks_per_cu;        /* launch shape used */
  int graph_nodes;
  int tuned_candidates;            /* candidates timed when autotune was set */
  double tuned_ms[16];             /* their launch times, in candidate order */
} mbr_batch_result;

/* Render n scenes into n caller-owned DEVICE framebuffers (fbs[i] sized as
 * scenes[i]).  Stats sum over the frames of the last replay.  Frames are
 * independent: results equal n calls of mbr_set_scene + mbr_render_device. */
int mbr_render_batch(mbr_ctx *ctx, const mbr_scene *scenes, int n,
                     const mbr_texture *lightmap, mbr_framebuffer *fbs,
                     const mbr_batch_opts *opts, mbr_stats *stats,
                     mbr_batch_result *result);


/* CompileImage (src/cimage.cpp:132-213): tone-map the float planes of a
 * render (HOST memory) with the scene's brightness, contrast, hdr and gamma
 * into interleaved 16-bit RGB (image16: 3 * width * height). */
int mbr_compile_image(const mbr_scene *scene, int width, int height,
                      const float *const rgb[3], uint16_t *image16);

/* mbr_compile_image on the GPU: the float planes and image16 are DEVICE
 * memory; queued on `stream` (NULL = the context's) without a host wait.
 * The same tone map and gamma table (computed on the host with the host
 * libm, as the reference's), so the result equals mbr_compile_image's
 * (hdr = 1 uses the device tanhf: equal up to its last bit). */
int mbr_compile_image_device(mbr_ctx *ctx, const mbr_scene *scene, int width, int height,
                             const float *const rgb[3], uint16_t *image16, void *stream);

/* SavePNG / SavePNG16 / SavePNG16Alpha (src/files.cpp): write image16 (and
 * the alpha plane when not NULL) as an 8- or 16-bit PNG. */
int mbr_write_png(const char *path, int width, int height, const uint16_t *image16,
                  const uint16_t *alpha, int bits);

/* Planes of one rendered frame for the post effects.  image16 is the
 * compiled image (mbr_compile_image), interleaved RGB, 3*width*height. */
typedef struct mbr_post_planes {
  int width, height;
  const float *z;                  /* zBuffer (1e20 on a miss) */
  const uint16_t *opacity;         /* opacity plane */
  const uint8_t *colour[3];        /* colour planes (palette colour of the hit) */
  uint16_t *image16;               /* in/out */
  uint16_t *alpha;                 /* in/out for DOF; unused by SSAO (may be NULL) */
} mbr_post_planes;

/* Screen-space ambient occlusion (PostRendering_SSAO / ThreadSSAO,
 * src/image.cpp:249-415; called from src/Render3D.cpp:1096-1099) with the
 * scene's ssao_quality, persp, perspective_type and global_ilum: for every
 * hit pixel, the horizon angle over ssao_quality screen directions of the
 * zBuffer, added to image16 as colour x AO x intensity x (1 - opacity).
 * on_device = 1: the planes are device memory and the kernel is queued on
 * `stream` (NULL = the context's) without a host wait; 0: host memory,
 * copied in and out, returns when done.  Needs no mbr_set_scene. */
int mbr_post_ssao(mbr_ctx *ctx, const mbr_scene *scene, const mbr_post_planes *planes,
                  int on_device, void *stream);

/* Depth of field (PostRendering_DOF, src/image.cpp:57-247; called from
 * src/Render3D.cpp:1101-1105) with focus 10^(dof_focus / 10 - 16) and
 * radius dof_radius x width / 1000: every pixel is splatted as a disc of
 * radius |z - focus| / z x radius (at most 100 pixels) onto image16 and
 * alpha (may be NULL), in the reference's order (z sorted, then shuffled
 * among pixels of similar blur with glibc-rand draws after
 * srand(jitter_seed)).  The order is computed on the host (the device z
 * plane is read back first); the splats run on the GPU and give the
 * reference's result exactly.  on_device / stream as for mbr_post_ssao;
 * the call waits for the device once (the size of the splat lists). */
int mbr_post_dof(mbr_ctx *ctx, const mbr_scene *scene, const mbr_post_planes *planes,
                 int on_device, void *stream);

/* Keyframe animation (CMorph::CatmullRom, src/morph.cpp:49-109, applied as
 * src/Render3D.cpp:2278-2295 does): every double parameter of `scene` (the
 * double keys of the settings table, and N, which the reference keeps as a
 * double) becomes the Catmull-Rom spline through the n_keys keyframe values
 * at `frame` (frames_per_key frames per key; logarithmic where all four
 * support values share a sign and vary by more than 10% of their mean);
 * the other fields of `scene` are left as they are. */
int mbr_keyframe_scene(const mbr_scene *keys, int n_keys, int frames_per_key, int frame, mbr_scene *scene);

/* One auxiliary light (sLight, src/shaders.h; the Lights array of
 * src/shaders.cpp:17). */
typedef struct mbr_light {
  double position[3];
  double intensity;
  int colour[3];                   /* sRGB: 16-bit channel values */
  int enabled;
} mbr_light;

/* Replaces PlaceRandomLights(params, false) (src/shaders.cpp:1134-1212), on
 * the host, in FP64: aux_light_number random lights drawn with the C
 * library's rand() sequence of srand(aux_light_random_seed) (glibc's
 * generator, kept private: the process's own rand() state is untouched) and
 * accepted where the distance estimate lies in (0, aux_light_max_dist), then
 * the predefined lights written over slots 0-3.  Writes aux_light_number + 4
 * records (`capacity` must allow that; *count gets the number).  Returns
 * MBR_ERR_INVALID_ARG when 20 million draws place no further light (the
 * reference would loop forever). */
int mbr_place_lights(const mbr_scene *scene, mbr_light *lights, int capacity, int *count);

/* The context's light table, which mbr_set_scene uses.  Without a call,
 * mbr_set_scene places the lights itself (mbr_place_lights) on its first
 * scene with aux lights and again whenever the placement inputs (seed,
 * number, radius, centre, maximum distance) change; the random lights then
 * stay where they are while fractal parameters change, as in the
 * reference's animation loop (src/Render3D.cpp:1990-1993, :2305).  Either
 * way each mbr_set_scene writes the scene's predefined lights over slots 0-3
 * (PlaceRandomLights(params, true)).  lights == NULL (or count 0) returns
 * to automatic placement. */
int mbr_set_lights(mbr_ctx *ctx, const mbr_light *lights, int count);
int mbr_get_lights(mbr_ctx *ctx, mbr_light *lights, int capacity, int *count);

/* Layout probe for bindings: sizeof(mbr_scene) and the offsets of
 * abi_version, vp, formula, mb_rotation, palette, fog_colour3,
 * reflections_max, jitter_seed (in that order). */
size_t mbr_scene_layout(size_t offsets[8]);

/* The jitter draw of MBR_JITTER_HASH: an integer in [0,1000] standing in for
 * Random(1000) at march step `step` of pixel `pixel` in frame `frame`. */
uint32_t mbr_jitter_draw(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t step);

#ifdef __cplusplus
}
#endif

#endif /* MBR_H_ */
