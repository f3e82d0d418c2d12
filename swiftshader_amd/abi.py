"""ctypes mirror of include/swr.h (the C ABI of the draw path).

The structures here are byte-for-byte the C structs of ``include/swr.h``;
the constants are the Vulkan enum values the descriptor uses.  The same
descriptor layout is what ``sw::Renderer::draw`` consumes through DrawCall /
DrawData in the reference (src/Device/Renderer.hpp:58-113, 211-213).
"""
import ctypes as C

ABI_VERSION = 12

OK = 0
ERROR_INVALID_ARGUMENT = -1
ERROR_UNSUPPORTED = -2
ERROR_OUT_OF_MEMORY = -3
ERROR_DEVICE = -4

FORMAT_UNDEFINED = 0
FORMAT_R8G8B8A8_UNORM = 37
FORMAT_R8G8B8A8_SRGB = 43
FORMAT_B8G8R8A8_UNORM = 44
FORMAT_B8G8R8A8_SRGB = 50
FORMAT_R32_UINT = 98
FORMAT_R32G32B32A32_SFLOAT = 109
FORMAT_D16_UNORM = 124
FORMAT_D32_SFLOAT = 126
FORMAT_S8_UINT = 127

COMPARE_NEVER, COMPARE_LESS, COMPARE_EQUAL, COMPARE_LESS_OR_EQUAL = 0, 1, 2, 3
COMPARE_GREATER, COMPARE_NOT_EQUAL, COMPARE_GREATER_OR_EQUAL, COMPARE_ALWAYS = 4, 5, 6, 7

STENCIL_OP_KEEP, STENCIL_OP_ZERO, STENCIL_OP_REPLACE = 0, 1, 2
STENCIL_OP_INCREMENT_AND_CLAMP, STENCIL_OP_DECREMENT_AND_CLAMP = 3, 4
STENCIL_OP_INVERT, STENCIL_OP_INCREMENT_AND_WRAP, STENCIL_OP_DECREMENT_AND_WRAP = 5, 6, 7

BLEND_FACTOR_ZERO, BLEND_FACTOR_ONE = 0, 1
BLEND_FACTOR_SRC_COLOR, BLEND_FACTOR_ONE_MINUS_SRC_COLOR = 2, 3
BLEND_FACTOR_DST_COLOR, BLEND_FACTOR_ONE_MINUS_DST_COLOR = 4, 5
BLEND_FACTOR_SRC_ALPHA, BLEND_FACTOR_ONE_MINUS_SRC_ALPHA = 6, 7
BLEND_FACTOR_DST_ALPHA, BLEND_FACTOR_ONE_MINUS_DST_ALPHA = 8, 9
BLEND_FACTOR_CONSTANT_COLOR, BLEND_FACTOR_ONE_MINUS_CONSTANT_COLOR = 10, 11
BLEND_FACTOR_CONSTANT_ALPHA, BLEND_FACTOR_ONE_MINUS_CONSTANT_ALPHA = 12, 13
BLEND_FACTOR_SRC_ALPHA_SATURATE = 14

BLEND_OP_ADD, BLEND_OP_SUBTRACT, BLEND_OP_REVERSE_SUBTRACT, BLEND_OP_MIN, BLEND_OP_MAX = 0, 1, 2, 3, 4

TOPOLOGY_POINT_LIST, TOPOLOGY_LINE_LIST, TOPOLOGY_LINE_STRIP = 0, 1, 2
TOPOLOGY_TRIANGLE_LIST, TOPOLOGY_TRIANGLE_STRIP, TOPOLOGY_TRIANGLE_FAN = 3, 4, 5
BLEND_OP_MULTIPLY_EXT, BLEND_OP_SCREEN_EXT, BLEND_OP_OVERLAY_EXT, BLEND_OP_DARKEN_EXT = (1000148012, 1000148013,
                                                                                        1000148014, 1000148015)
BLEND_OP_LIGHTEN_EXT, BLEND_OP_COLORDODGE_EXT, BLEND_OP_COLORBURN_EXT = 1000148016, 1000148017, 1000148018
BLEND_OP_HARDLIGHT_EXT, BLEND_OP_SOFTLIGHT_EXT, BLEND_OP_DIFFERENCE_EXT = 1000148019, 1000148020, 1000148021
BLEND_OP_EXCLUSION_EXT = 1000148022
BLEND_OP_HSL_HUE_EXT, BLEND_OP_HSL_SATURATION_EXT = 1000148031, 1000148032
BLEND_OP_HSL_COLOR_EXT, BLEND_OP_HSL_LUMINOSITY_EXT = 1000148033, 1000148034
ADVANCED_BLEND_OPS = tuple(range(1000148012, 1000148023)) + tuple(range(1000148031, 1000148035))
LINE_RASTERIZATION_DEFAULT, LINE_RASTERIZATION_RECTANGULAR = 0, 1
LINE_RASTERIZATION_BRESENHAM, LINE_RASTERIZATION_RECTANGULAR_SMOOTH = 2, 3
PROVOKING_VERTEX_FIRST, PROVOKING_VERTEX_LAST = 0, 1
INDEX_TYPE_UINT16, INDEX_TYPE_UINT32, INDEX_TYPE_UINT8 = 0, 1, 1000265000

CULL_NONE, CULL_FRONT, CULL_BACK, CULL_FRONT_AND_BACK = 0, 1, 2, 3
FRONT_FACE_COUNTER_CLOCKWISE, FRONT_FACE_CLOCKWISE = 0, 1

FILTER_NEAREST, FILTER_LINEAR = 0, 1
MIPMAP_MODE_NEAREST, MIPMAP_MODE_LINEAR = 0, 1
ADDRESS_REPEAT, ADDRESS_MIRRORED_REPEAT, ADDRESS_CLAMP_TO_EDGE, ADDRESS_MIRROR_CLAMP_TO_EDGE = 0, 1, 2, 4

SHADER_INTERPOLATE, SHADER_SAMPLE, SHADER_MODULATE = 0, 1, 2

MAX_MIP_LEVELS = 15


class Viewport(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("x", "y", "width", "height", "min_depth", "max_depth")]


class Rect(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("width", C.c_uint32), ("height", C.c_uint32)]


class StencilOpState(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("fail_op", "pass_op", "depth_fail_op", "compare_op",
                                            "compare_mask", "write_mask", "reference")]


class BlendState(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("blend_enable", "src_color_blend_factor", "dst_color_blend_factor",
                                            "color_blend_op", "src_alpha_blend_factor", "dst_alpha_blend_factor",
                                            "alpha_blend_op", "color_write_mask")]


class SamplerState(C.Structure):
    _fields_ = [("mag_filter", C.c_uint32), ("min_filter", C.c_uint32), ("mipmap_mode", C.c_uint32),
                ("address_mode_u", C.c_uint32), ("address_mode_v", C.c_uint32),
                ("mip_lod_bias", C.c_float), ("anisotropy_enable", C.c_uint32), ("max_anisotropy", C.c_float),
                ("min_lod", C.c_float), ("max_lod", C.c_float)]


class ImageLevel(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("row_pitch_bytes", C.c_uint32), ("reserved", C.c_uint32)]


class Texture(C.Structure):
    _fields_ = [("format", C.c_uint32), ("level_count", C.c_uint32), ("levels", ImageLevel * MAX_MIP_LEVELS)]


class Attachment(C.Structure):
    _fields_ = [("data", C.c_void_p), ("format", C.c_uint32), ("row_pitch_bytes", C.c_uint32),
                ("sample_pitch_bytes", C.c_uint64)]


class DrawDesc(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p), ("colors", C.c_void_p), ("texcoords", C.c_void_p),
        ("vertex_count", C.c_uint32), ("triangle_count", C.c_uint32), ("indices", C.c_void_p),
        ("viewport", Viewport), ("scissor", Rect), ("render_area", Rect),
        ("cull_mode", C.c_uint32), ("front_face", C.c_uint32),
        ("samples", C.c_uint32), ("sample_mask", C.c_uint32),
        ("depth_test_enable", C.c_uint32), ("depth_write_enable", C.c_uint32), ("depth_compare_op", C.c_uint32),
        ("stencil_test_enable", C.c_uint32), ("front", StencilOpState), ("back", StencilOpState),
        ("blend", BlendState), ("blend_constants", C.c_float * 4),
        ("shader", C.c_uint32), ("sampler", SamplerState), ("texture", Texture),
        ("framebuffer_width", C.c_uint32), ("framebuffer_height", C.c_uint32),
        ("color", Attachment), ("depth", Attachment), ("stencil", Attachment),
        ("occlusion_count", C.c_void_p),
        ("primitive_id", Attachment),
        ("topology", C.c_uint32), ("provoking_vertex", C.c_uint32),
        ("depth_bias_enable", C.c_uint32), ("depth_bias_constant_factor", C.c_float),
        ("depth_bias_clamp", C.c_float), ("depth_bias_slope_factor", C.c_float),
        ("line_width", C.c_float), ("line_rasterization_mode", C.c_uint32),
        ("point_sizes", C.c_void_p), ("index_type", C.c_uint32),
        ("depth_clamp_enable", C.c_uint32), ("depth_clip_control", C.c_uint32),
        ("depth_bounds_test_enable", C.c_uint32), ("min_depth_bounds", C.c_float), ("max_depth_bounds", C.c_float),
        ("alpha_to_coverage_enable", C.c_uint32),
    ]


class BCDesc(C.Structure):
    _fields_ = [("src", C.c_void_p), ("dst", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("dst_row_pitch_bytes", C.c_uint32), ("dst_bytes_per_pixel", C.c_uint32), ("n", C.c_uint32),
                ("no_alpha_or_unsigned", C.c_uint32)]


class ETCDesc(C.Structure):
    _fields_ = [("src", C.c_void_p), ("dst", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("dst_row_pitch_bytes", C.c_uint32), ("dst_bytes_per_pixel", C.c_uint32),
                ("input_type", C.c_uint32)]


ETC_R_SIGNED, ETC_R_UNSIGNED, ETC_RG_SIGNED, ETC_RG_UNSIGNED = 0, 1, 2, 3
ETC_RGB, ETC_RGB_PUNCHTHROUGH_ALPHA, ETC_RGBA = 4, 5, 6


class ResolveDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32),
                ("src", C.c_void_p), ("src_row_pitch_bytes", C.c_uint32), ("src_sample_pitch_bytes", C.c_uint64),
                ("src_samples", C.c_uint32), ("dst", C.c_void_p), ("dst_row_pitch_bytes", C.c_uint32)]


class _DepthStencilValue(C.Structure):
    _fields_ = [("depth", C.c_float), ("stencil", C.c_uint32)]


class ClearValue(C.Union):
    _fields_ = [("color", C.c_float * 4), ("depth_stencil", _DepthStencilValue)]


class ClearDesc(C.Structure):
    _fields_ = [("target", Attachment), ("width", C.c_uint32), ("height", C.c_uint32), ("samples", C.c_uint32),
                ("use_area", C.c_uint32), ("area", Rect), ("value", ClearValue)]


class GradDesc(C.Structure):
    _fields_ = [("d_color", C.c_void_p), ("d_color_row_pitch_bytes", C.c_uint32),
                ("d_color_sample_pitch_bytes", C.c_uint64),
                ("d_positions", C.c_void_p), ("d_colors", C.c_void_p), ("d_texcoords", C.c_void_p),
                ("d_texture", C.c_void_p * MAX_MIP_LEVELS),
                ("sample_state", C.c_void_p), ("sample_state_row_pitch_bytes", C.c_uint32),
                ("sample_state_sample_pitch_bytes", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("tile_width", C.c_uint32), ("tile_height", C.c_uint32), ("raster_workgroups_per_cu", C.c_uint32),
                ("reserved", C.c_uint32 * 5)]


class Profile(C.Structure):
    _fields_ = [("setup_ms", C.c_double), ("bin_ms", C.c_double), ("raster_ms", C.c_double),
                ("total_ms", C.c_double), ("draws", C.c_uint64), ("pairs", C.c_uint64),
                ("block_tests", C.c_uint64), ("block_covers", C.c_uint64), ("shaded", C.c_uint64)]


# Symbols declared by include/swr.h, in declaration order.
EXPORTED_SYMBOLS = (
    "swr_abi_version", "swr_last_error", "swr_init", "swr_release", "swr_draw", "swr_draw_noalloc",
    "swr_draw_batch", "swr_resolve", "swr_clear", "swr_clear_batch", "swr_decode_bc", "swr_decode_etc", "swr_batch_plan_create", "swr_batch_plan_run", "swr_batch_plan_check",
    "swr_batch_plan_autotune", "swr_batch_plan_set_grads", "swr_batch_plan_backward", "swr_batch_plan_destroy", "swr_draw_backward", "swr_generate_mips", "swr_mip_backward",
    "swr_profile_enable", "swr_profile_read", "swr_set_tuning", "swr_get_tuning",
)


def bind(lib):
    """Attach argtypes/restype to the entry points of a loaded libswr."""
    lib.swr_abi_version.restype = C.c_int
    lib.swr_last_error.restype = C.c_char_p
    lib.swr_init.argtypes = [C.c_int]
    lib.swr_init.restype = C.c_int
    lib.swr_release.restype = C.c_int
    lib.swr_draw.argtypes = [C.POINTER(DrawDesc), C.c_void_p]
    lib.swr_draw.restype = C.c_int
    lib.swr_draw_noalloc.argtypes = [C.POINTER(DrawDesc), C.c_void_p]
    lib.swr_draw_noalloc.restype = C.c_int
    lib.swr_draw_batch.argtypes = [C.POINTER(DrawDesc), C.c_uint32, C.c_void_p]
    lib.swr_draw_batch.restype = C.c_int
    lib.swr_resolve.argtypes = [C.POINTER(ResolveDesc), C.c_void_p]
    lib.swr_resolve.restype = C.c_int
    lib.swr_clear.argtypes = [C.POINTER(ClearDesc), C.c_void_p]
    lib.swr_clear.restype = C.c_int
    lib.swr_clear_batch.argtypes = [C.POINTER(ClearDesc), C.c_uint32, C.c_void_p]
    lib.swr_clear_batch.restype = C.c_int
    lib.swr_decode_bc.argtypes = [C.POINTER(BCDesc), C.c_void_p]
    lib.swr_decode_bc.restype = C.c_int
    lib.swr_decode_etc.argtypes = [C.POINTER(ETCDesc), C.c_void_p]
    lib.swr_decode_etc.restype = C.c_int
    lib.swr_batch_plan_create.argtypes = [C.POINTER(DrawDesc), C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p),
                                          C.c_void_p]
    lib.swr_batch_plan_create.restype = C.c_int
    lib.swr_batch_plan_run.argtypes = [C.c_void_p, C.c_void_p]
    lib.swr_batch_plan_run.restype = C.c_int
    lib.swr_batch_plan_check.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
    lib.swr_batch_plan_check.restype = C.c_int
    lib.swr_batch_plan_autotune.argtypes = [C.c_void_p, C.POINTER(Tuning), C.c_void_p]
    lib.swr_batch_plan_autotune.restype = C.c_int
    lib.swr_batch_plan_set_grads.argtypes = [C.c_void_p, C.POINTER(GradDesc), C.c_uint32, C.c_void_p]
    lib.swr_batch_plan_set_grads.restype = C.c_int
    lib.swr_batch_plan_backward.argtypes = [C.c_void_p, C.c_void_p]
    lib.swr_batch_plan_backward.restype = C.c_int
    lib.swr_batch_plan_destroy.argtypes = [C.c_void_p]
    lib.swr_batch_plan_destroy.restype = C.c_int
    lib.swr_draw_backward.argtypes = [C.POINTER(DrawDesc), C.POINTER(GradDesc), C.c_void_p]
    lib.swr_draw_backward.restype = C.c_int
    lib.swr_generate_mips.argtypes = [C.POINTER(Texture), C.c_void_p]
    lib.swr_generate_mips.restype = C.c_int
    lib.swr_mip_backward.argtypes = [C.POINTER(Texture), C.c_void_p]
    lib.swr_mip_backward.restype = C.c_int
    lib.swr_profile_enable.argtypes = [C.c_int]
    lib.swr_profile_enable.restype = C.c_int
    lib.swr_profile_read.argtypes = [C.POINTER(Profile)]
    lib.swr_profile_read.restype = C.c_int
    lib.swr_set_tuning.argtypes = [C.POINTER(Tuning)]
    lib.swr_set_tuning.restype = C.c_int
    lib.swr_get_tuning.argtypes = [C.POINTER(Tuning)]
    lib.swr_get_tuning.restype = C.c_int
    return lib
