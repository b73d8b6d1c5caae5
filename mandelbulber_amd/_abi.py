"""ctypes mirror of include/mbr.h (the C ABI of the render core).

The layout is checked against the compiled library by ``mbr_scene_layout``
when the library is loaded (see ``mandelbulber_amd._lib``).
"""
import ctypes as C

MBR_ABI_VERSION = 3

MBR_OK = 0
STATUS = {
    0: "MBR_OK", -1: "MBR_ERR_INVALID_ARG", -2: "MBR_ERR_NO_DEVICE", -3: "MBR_ERR_HIP",
    -4: "MBR_ERR_UNSUPPORTED", -5: "MBR_ERR_NO_SCENE", -6: "MBR_ERR_CANCELLED",
    -7: "MBR_ERR_IO", -8: "MBR_ERR_OUT_OF_MEMORY",
}

FORMULA_NONE, FORMULA_TRIG_OPTIM, FORMULA_MENGER_SPONGE, FORMULA_TGLAD = 0, 2, 7, 8
PERSP_THREE_POINT, PERSP_FISH_EYE, PERSP_EQUIRECTANGULAR = 0, 1, 2
JITTER_NONE, JITTER_HASH = 0, 1
PREC_FP64, PREC_FP32_FIXUP, PREC_FP32_MARCH, PREC_FP64_DEMOIVRE = 0, 1, 2, 3


class RGB(C.Structure):
    _fields_ = [("R", C.c_int), ("G", C.c_int), ("B", C.c_int)]


D3 = C.c_double * 3


class Scene(C.Structure):
    """mbr_scene: field names follow include/mbr.h."""
    _fields_ = [
        ("abi_version", C.c_int),
        ("image_width", C.c_int), ("image_height", C.c_int),
        ("tiles", C.c_int), ("tile_index", C.c_int),
        ("vp", D3),
        ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double),
        ("zoom", C.c_double), ("persp", C.c_double),
        ("perspective_type", C.c_int), ("fish_eye_cut", C.c_int),
        ("view_distance_min", C.c_double), ("view_distance_max", C.c_double),
        ("quality", C.c_double), ("DE_factor", C.c_double), ("smoothness", C.c_double),
        ("formula", C.c_int), ("N", C.c_int), ("minN", C.c_int),
        ("power", C.c_double), ("bailout", C.c_double), ("constant_factor", C.c_double),
        ("julia_mode", C.c_int), ("julia", D3),
        ("interior_mode", C.c_int), ("iter_thresh", C.c_int),
        ("constant_DE_threshold", C.c_int), ("limits_enabled", C.c_int),
        ("limit_min", D3), ("limit_max", D3),
        ("tglad_folding_mode", C.c_int), ("folding_limit", C.c_double), ("folding_value", C.c_double),
        ("spherical_folding_mode", C.c_int), ("spherical_folding_fixed", C.c_double),
        ("spherical_folding_min", C.c_double),
        ("ifs_folding_mode", C.c_int),
        ("mb_scale", C.c_double), ("mb_folding_limit", C.c_double), ("mb_folding_value", C.c_double),
        ("mb_min_radius", C.c_double), ("mb_fixed_radius", C.c_double), ("mb_offset", D3),
        ("mb_color_R", C.c_double), ("mb_color_X", C.c_double), ("mb_color_Y", C.c_double),
        ("mb_color_Z", C.c_double), ("mb_color_Sp1", C.c_double), ("mb_color_Sp2", C.c_double),
        ("mb_rotations_enabled", C.c_int), ("mb_rotation_main", D3),
        ("mb_rotation", ((C.c_double * 3) * 3) * 2),
        ("primitive_only_plane", C.c_int),
        ("plane_enabled", C.c_int), ("plane_centre", D3), ("plane_normal", D3),
        ("box_enabled", C.c_int), ("box_centre", D3), ("box_size", D3),
        ("inv_box_enabled", C.c_int), ("inv_box_centre", D3), ("inv_box_size", D3),
        ("sphere_enabled", C.c_int), ("sphere_centre", D3), ("sphere_radius", C.c_double),
        ("inv_sphere_enabled", C.c_int), ("inv_sphere_centre", D3), ("inv_sphere_radius", C.c_double),
        ("water_enabled", C.c_int), ("water_height", C.c_double), ("water_amplitude", C.c_double),
        ("water_length", C.c_double), ("water_rotation", C.c_double), ("water_anim_speed", C.c_double),
        ("water_iterations", C.c_int), ("frame_no", C.c_int),
        ("plane_colour", RGB), ("box_colour", RGB), ("inv_box_colour", RGB), ("sphere_colour", RGB),
        ("inv_sphere_colour", RGB), ("water_colour", RGB),
        ("shadow", C.c_int), ("global_illumination", C.c_int), ("fast_global_illumination", C.c_int),
        ("ao_quality", C.c_int), ("slow_shading", C.c_int), ("penetrating_lights", C.c_int),
        ("main_light_alpha", C.c_double), ("main_light_beta", C.c_double),
        ("shadow_cone_angle", C.c_double), ("main_light_intensity", C.c_double),
        ("direct_light", C.c_double), ("shading", C.c_double), ("specular", C.c_double),
        ("global_ilum", C.c_double), ("fast_ao_tune", C.c_double), ("reflect", C.c_double),
        ("glow_intensity", C.c_double),
        ("main_light_colour", RGB), ("glow_color1", RGB), ("glow_color2", RGB),
        ("coloring_enabled", C.c_int), ("coloring_speed", C.c_double), ("palette_offset", C.c_double),
        ("palette", RGB * 256),
        ("textured_background", C.c_int),
        ("background_color1", RGB), ("background_color2", RGB), ("background_color3", RGB),
        ("post_fog_enabled", C.c_int), ("fog_visibility", C.c_double), ("post_fog_color", RGB),
        ("fog_density", C.c_double), ("fog_colour1_distance", C.c_double),
        ("fog_colour2_distance", C.c_double), ("fog_distance_factor", C.c_double),
        ("fog_colour1", RGB), ("fog_colour2", RGB), ("fog_colour3", RGB),
        ("aux_light_number", C.c_int), ("aux_light_pre_enabled", C.c_int * 4),
        ("volumetric_light_enabled", C.c_int * 5), ("iter_fog_enabled", C.c_int),
        ("fake_lights_enabled", C.c_int), ("raytraced_reflections", C.c_int),
        ("reflections_max", C.c_int),
        ("jitter_mode", C.c_int), ("jitter_seed", C.c_uint32), ("hit_refine", C.c_int),
        ("brightness", C.c_double), ("contrast", C.c_double), ("image_gamma", C.c_double), ("hdr", C.c_int),
        ("file_destination", C.c_char * 512),
        ("start_frame", C.c_int), ("end_frame", C.c_int), ("frames_per_keyframe", C.c_int),
        ("file_animation_path", C.c_char * 512), ("file_keyframes", C.c_char * 512),
        ("linear_DE_mode", C.c_int), ("mb_sharpness", C.c_double), ("folding_int_pow_factor", C.c_double),
        ("folding_int_pow_z_factor", C.c_double), ("c_add", C.c_double),
        ("ifs_scale", C.c_double), ("ifs_rotation", C.c_double * 3), ("ifs_offset", C.c_double * 3),
        ("ifs_edge", C.c_double * 3), ("ifs_abs", C.c_int * 3), ("ifs_menger_sponge_mode", C.c_int),
        ("ifs_direction", (C.c_double * 3) * 9), ("ifs_angles", (C.c_double * 3) * 9),
        ("ifs_distance", C.c_double * 9), ("ifs_intensity", C.c_double * 9), ("ifs_enabled", C.c_int * 9),
        ("hybrid_formula", C.c_int * 5), ("hybrid_iterations", C.c_int * 5), ("hybrid_power", C.c_double * 5),
        ("hybrid_cyclic", C.c_int), ("ssao_enabled", C.c_int), ("ssao_quality", C.c_int), ("dof_enabled", C.c_int),
        ("dof_focus", C.c_double), ("dof_radius", C.c_double),
        ("aux_light_intensity", C.c_double), ("aux_light_random_seed", C.c_int),
        ("aux_light_max_dist", C.c_double), ("aux_light_distribution_radius", C.c_double),
        ("aux_light_visibility", C.c_double), ("aux_light_random_center", C.c_double * 3),
        ("aux_light_pre", (C.c_double * 3) * 4), ("aux_light_pre_intensity", C.c_double * 4),
        ("aux_light_pre_colour", RGB * 4),
        ("volumetric_light_intensity", C.c_double * 5), ("iter_fog_opacity", C.c_double),
        ("iter_fog_opacity_trim", C.c_double),
        ("plane_reflect", C.c_double), ("box_reflect", C.c_double), ("inv_box_reflect", C.c_double),
        ("sphere_reflect", C.c_double), ("inv_sphere_reflect", C.c_double), ("water_reflect", C.c_double),
        ("mb_fold_mode", C.c_int), ("mb_solid", C.c_double), ("mb_melt", C.c_double),
        ("mb_vary_scale_vary", C.c_double), ("mb_vary_fold", C.c_double), ("mb_vary_minr", C.c_double),
        ("mb_vary_rpower", C.c_double), ("mb_vary_wadd", C.c_double),
        ("fake_lights_min_iter", C.c_int), ("fake_lights_max_iter", C.c_int),
        ("fake_lights_intensity", C.c_double), ("fake_lights_visibility", C.c_double),
        ("fake_lights_visibility_size", C.c_double), ("fake_lights_orbit_trap", C.c_double * 3),
        ("background_as_fulldome", C.c_int),
    ]


class Light(C.Structure):
    """One auxiliary light (mbr_light)."""
    _fields_ = [("position", C.c_double * 3), ("intensity", C.c_double), ("colour", C.c_int * 3),
                ("enabled", C.c_int)]


LAYOUT_FIELDS = ("abi_version", "vp", "formula", "mb_rotation", "palette", "fog_colour3",
                 "reflections_max", "jitter_seed")


class Texture(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("rgb", C.POINTER(C.c_uint8))]


class Framebuffer(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int),
        ("rgb", C.c_void_p * 3),
        ("z", C.c_void_p),
        ("colour", C.c_void_p * 3),
        ("alpha", C.c_void_p),
        ("opacity", C.c_void_p),
    ]


class PostPlanes(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int),
        ("z", C.c_void_p), ("opacity", C.c_void_p), ("colour", C.c_void_p * 3),
        ("image16", C.c_void_p), ("alpha", C.c_void_p),
    ]


class BatchOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("stream", C.c_void_p), ("slots", C.c_int), ("blocks_per_cu", C.c_int),
        ("autotune", C.c_int), ("replays", C.c_int),
    ]


class BatchResult(C.Structure):
    _fields_ = [
        ("ms_per_launch", C.c_double), ("ms_build", C.c_double), ("slots", C.c_int), ("blocks_per_cu", C.c_int),
        ("graph_nodes", C.c_int), ("tuned_candidates", C.c_int), ("tuned_ms", C.c_double * 16),
    ]


class RenderOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("stream", C.c_void_p),
        ("cancel", C.POINTER(C.c_int)),
        ("row_begin", C.c_int), ("row_end", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double), ("ms_march", C.c_double), ("ms_shade", C.c_double),
        ("ms_finish", C.c_double),
        ("rays", C.c_uint64), ("hits", C.c_uint64),
        ("march_de_calls", C.c_uint64), ("shade_de_calls", C.c_uint64),
        ("iterations", C.c_uint64), ("march_iterations", C.c_uint64), ("shade_iterations", C.c_uint64),
        ("march_launches", C.c_uint64), ("shade_launches", C.c_uint64), ("finish_launches", C.c_uint64),
        ("missed_de", C.c_uint64),
        ("step_overflow_rays", C.c_uint64), ("fp64_redo_rays", C.c_uint64),
        ("ms_ao", C.c_double), ("ao_de_calls", C.c_uint64), ("ao_iterations", C.c_uint64),
        ("ao_steps", C.c_uint64), ("ms_ao_span", C.c_double),
    ]


class DeviceOpts(C.Structure):
    _fields_ = [("device", C.c_int)]


def scene_layout_py():
    """(sizeof, offsets) of the ctypes mirror, comparable with mbr_scene_layout."""
    return C.sizeof(Scene), [getattr(Scene, f).offset for f in LAYOUT_FIELDS]
