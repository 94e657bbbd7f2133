"""ctypes mirror of include/rvpf.h (the C ABI of the post-filter chain).

The structs here must match the C layout byte for byte; tests compare
``ctypes.sizeof`` with the sizes the compiled libraries report.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))

RVPF_OK = 0
RVPF_E_INVALID = -1
RVPF_E_UNSUPPORTED = -2
RVPF_E_DEVICE = -3
RVPF_E_NOMEM = -4

RESTORE_NONE = 0
RESTORE_WIENER = 1
RESTORE_SGRPROJ = 2
RESTORE_SWITCHABLE = 3


class RvpfMi(C.Structure):
    _fields_ = [
        ("mi_size", C.c_uint8),
        ("y_mode", C.c_uint8),
        ("ref_frame0", C.c_int8),
        ("seg_skip", C.c_uint8),
        ("delta_lf", C.c_int8 * 4),
    ]


class RvpfLrUnit(C.Structure):
    _fields_ = [
        ("type", C.c_uint8),
        ("sgr_set", C.c_uint8),
        ("sgr_xqd", C.c_int8 * 2),
        ("wiener", (C.c_int8 * 3) * 2),
        ("pad", C.c_uint8 * 2),
    ]


class RvpfGrainParams(C.Structure):
    _fields_ = [
        ("apply_grain", C.c_int32),
        ("grain_seed", C.c_uint32),
        ("num_y_points", C.c_int32),
        ("point_y_value", C.c_int32 * 16),
        ("point_y_scaling", C.c_int32 * 16),
        ("chroma_scaling_from_luma", C.c_int32),
        ("num_cb_points", C.c_int32),
        ("point_cb_value", C.c_int32 * 16),
        ("point_cb_scaling", C.c_int32 * 16),
        ("num_cr_points", C.c_int32),
        ("point_cr_value", C.c_int32 * 16),
        ("point_cr_scaling", C.c_int32 * 16),
        ("grain_scaling_minus_8", C.c_int32),
        ("ar_coeff_lag", C.c_int32),
        ("ar_coeffs_y_plus_128", C.c_int32 * 24),
        ("ar_coeffs_cb_plus_128", C.c_int32 * 25),
        ("ar_coeffs_cr_plus_128", C.c_int32 * 25),
        ("ar_coeff_shift_minus_6", C.c_int32),
        ("grain_scale_shift", C.c_int32),
        ("cb_mult", C.c_int32),
        ("cb_luma_mult", C.c_int32),
        ("cb_offset", C.c_int32),
        ("cr_mult", C.c_int32),
        ("cr_luma_mult", C.c_int32),
        ("cr_offset", C.c_int32),
        ("overlap_flag", C.c_int32),
        ("clip_to_restricted_range", C.c_int32),
    ]


class RvpfFrameParams(C.Structure):
    _fields_ = [
        ("bit_depth", C.c_int32),
        ("ssx", C.c_int32),
        ("ssy", C.c_int32),
        ("mono_chrome", C.c_int32),
        ("matrix_coefficients", C.c_int32),
        ("enable_cdef", C.c_int32),
        ("film_grain_params_present", C.c_int32),
        ("frame_width", C.c_int32),
        ("frame_height", C.c_int32),
        ("upscaled_width", C.c_int32),
        ("use_superres", C.c_int32),
        ("mi_rows", C.c_int32),
        ("mi_cols", C.c_int32),
        ("coded_lossless", C.c_int32),
        ("allow_intrabc", C.c_int32),
        ("loop_filter_level", C.c_int32 * 4),
        ("loop_filter_sharpness", C.c_int32),
        ("loop_filter_delta_enabled", C.c_int32),
        ("loop_filter_ref_deltas", C.c_int32 * 8),
        ("loop_filter_mode_deltas", C.c_int32 * 2),
        ("delta_lf_multi", C.c_int32),
        ("segmentation_enabled", C.c_int32),
        ("seg_feature_enabled", (C.c_int32 * 8) * 8),
        ("seg_feature_data", (C.c_int32 * 8) * 8),
        ("cdef_damping", C.c_int32),
        ("cdef_y_pri_strength", C.c_int32 * 8),
        ("cdef_y_sec_strength", C.c_int32 * 8),
        ("cdef_uv_pri_strength", C.c_int32 * 8),
        ("cdef_uv_sec_strength", C.c_int32 * 8),
        ("frame_restoration_type", C.c_int32 * 3),
        ("loop_restoration_size", C.c_int32 * 3),
        ("lr_unit_rows", C.c_int32 * 3),
        ("lr_unit_cols", C.c_int32 * 3),
        ("grain", RvpfGrainParams),
    ]


class RvpfSideInfo(C.Structure):
    _fields_ = [
        ("mi", C.c_void_p),
        ("mi_stride", C.c_int32),
        ("tx_size", C.c_void_p * 3),
        ("cdef_idx", C.c_void_p),
        ("cdef_stride", C.c_int32),
        ("lr_units", C.c_void_p * 3),
    ]


class RvpfPlanes(C.Structure):
    _fields_ = [
        ("data", C.c_void_p * 3),
        ("stride", C.c_int64 * 3),
        ("width", C.c_int32 * 3),
        ("height", C.c_int32 * 3),
        ("bytes_per_sample", C.c_int32),
    ]


class SynthCfg(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("bit_depth", C.c_int32),
        ("ssx", C.c_int32),
        ("ssy", C.c_int32),
        ("mono", C.c_int32),
        ("sb_size", C.c_int32),
        ("qindex", C.c_int32),
        ("seed", C.c_uint64),
        ("frame_index", C.c_int32),
        ("pan_px", C.c_double),
        ("lf", C.c_int32),
        ("lf_extras", C.c_int32),
        ("cdef", C.c_int32),
        ("cdef_bits", C.c_int32),
        ("lr", C.c_int32),
        ("lr_unit_size", C.c_int32),
        ("grain", C.c_int32),
        ("grain_lag", C.c_int32),
        ("grain_y_points", C.c_int32),
        ("mc_identity", C.c_int32),
        ("superres_denom", C.c_int32),
    ]


class SynthGeom(C.Structure):
    _fields_ = [
        ("mi_rows", C.c_int32),
        ("mi_cols", C.c_int32),
        ("mi_stride", C.c_int32),
        ("mi_alloc_rows", C.c_int32),
        ("cdef_stride", C.c_int32),
        ("cdef_rows", C.c_int32),
        ("plane_w", C.c_int32 * 3),
        ("plane_h", C.c_int32 * 3),
        ("crop_w", C.c_int32 * 3),
        ("crop_h", C.c_int32 * 3),
        ("lr_unit_rows", C.c_int32 * 3),
        ("lr_unit_cols", C.c_int32 * 3),
        ("num_planes", C.c_int32),
        ("coded_width", C.c_int32),
        ("ref_w", C.c_int32 * 3),
    ]


class RvpfBatchFrame(C.Structure):
    _fields_ = [
        ("params", C.c_void_p),
        ("side", RvpfSideInfo),
        ("recon", RvpfPlanes),
        ("ref_out", RvpfPlanes),
        ("display", RvpfPlanes),
    ]


class RvpfTileConfig(C.Structure):
    _fields_ = [
        ("shape", C.c_int32),
        ("workgroups_per_cu", C.c_int32),
        ("chunk", C.c_int32),
        ("us", C.c_float),
    ]


class RvpfLrdItem(C.Structure):
    _fields_ = [
        ("plane", C.c_int32),
        ("x0", C.c_int32),
        ("y0", C.c_int32),
        ("w", C.c_int32),
        ("h", C.c_int32),
        ("stripe_start", C.c_int32),
        ("stripe_end", C.c_int32),
        ("unit", C.c_int32),
    ]


class RvpfPackedLayout(C.Structure):
    _fields_ = [
        ("offset", C.c_int64 * 3),
        ("width", C.c_int32 * 3),
        ("height", C.c_int32 * 3),
        ("num_planes", C.c_int32),
        ("bytes_per_sample", C.c_int32),
        ("bytes", C.c_int64),
    ]


class RvpfCdefSearchResult(C.Structure):
    _fields_ = [
        ("damping_minus_3", C.c_int32),
        ("bits", C.c_int32),
        ("y_pri", C.c_int32 * 8),
        ("y_sec", C.c_int32 * 8),
        ("uv_pri", C.c_int32 * 8),
        ("uv_sec", C.c_int32 * 8),
    ]

    def as_tuple(self):
        n = 1 << self.bits
        return (self.damping_minus_3, self.bits, [(self.y_pri[i], self.y_sec[i]) for i in range(n)],
                [(self.uv_pri[i], self.uv_sec[i]) for i in range(n)])


TUNE_STAGES = ["deblock", "cdef", "lr", "grain_prepare", "grain_apply", "copy"]
TUNE_SHAPES = [3, 1, 1, 1, 1, 1]

STAGES = ["deblock", "cdef", "lr", "grain_prepare", "grain_apply", "copy", "upscale"]


class RvpfProfile(C.Structure):
    _fields_ = [("ms", C.c_float * 7), ("launches", C.c_int32 * 7)]


STRUCT_SIZES = [
    C.sizeof(RvpfFrameParams),
    C.sizeof(RvpfSideInfo),
    C.sizeof(RvpfMi),
    C.sizeof(RvpfLrUnit),
    C.sizeof(RvpfGrainParams),
    C.sizeof(RvpfPlanes),
    C.sizeof(RvpfBatchFrame),
    C.sizeof(RvpfTileConfig),
    C.sizeof(RvpfLrdItem),
    C.sizeof(RvpfPackedLayout),
    C.sizeof(RvpfCdefSearchResult),
]


def lib_path(name):
    return os.path.join(HERE, name)


# Functions the product library exports (every declaration of include/rvpf.h).
RVPF_EXPORTS = [
    "rvpf_ctx_create",
    "rvpf_ctx_destroy",
    "rvpf_ctx_sync",
    "rvpf_ctx_last_error",
    "rvpf_ctx_set_profiling",
    "rvpf_ctx_read_profile",
    "rvpf_validate",
    "rvpf_deblock",
    "rvpf_cdef",
    "rvpf_upscale",
    "rvpf_lr",
    "rvpf_grain",
    "rvpf_grain_noalloc",
    "rvpf_packed_layout_of",
    "rvpf_output_packed",
    "rvpf_postfilter_frame",
    "rvpf_lf_search",
    "rvpf_cdef_search",
    "rvpf_batch_create",
    "rvpf_batch_destroy",
    "rvpf_batch_autotune",
    "rvpf_batch_get_tiles",
    "rvpf_batch_set_tiles",
    "rvpf_batch_run",
    "rvpf_lrd_gather",
    "rvpf_lrd_wiener",
    "rvpf_lrd_sgr_filter",
    "rvpf_lrd_sgr",
    "rvpf_lrd_rect_sse",
    "rvpf_ac_q",
]
