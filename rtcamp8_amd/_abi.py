"""ctypes mirror of include/rtc8x.h (structs, enums, function prototypes).

Pointers inside the structs are plain integers (``c_void_p``): device addresses for the library,
host addresses when the same structs are handed to the CPU oracle in the test-suite.
"""
from __future__ import annotations

import ctypes as C

c_f3 = C.c_float * 3
c_i3 = C.c_int32 * 3

RTC8X_OK = 0
RTC8X_ERR_INVALID_ARGUMENT = 1
RTC8X_ERR_HIP = 2
RTC8X_ERR_OUT_OF_MEMORY = 3
RTC8X_ERR_NOT_SUPPORTED = 4

UPPER_TABLE_SIZE = 32768
LOWER_TABLE_SIZE = 4096
LEAF_SIZE = 512
TILE_BIT = 0x80000000

ENV_NONE, ENV_CONSTANT, ENV_LATLONG = 0, 1, 2
ENV_SAMPLER_HIERARCHICAL, ENV_SAMPLER_ALIAS = 0, 1
PHASE_SCHLICK, PHASE_HENYEY_GREENSTEIN = 0, 1
TRACKING_GLOBAL_MAJORANT, TRACKING_LOCAL_MAJORANT = 0, 1
ENTRY_PATH_TRACE, ENTRY_GENERATE_TRAINING_DATA, ENTRY_PATH_TRACE_WITH_NRC = 0, 1, 2
STAT_PATHS, STAT_VERTICES, STAT_LOOKUPS, STAT_SCATTER, STAT_COUNT = 0, 1, 2, 3, 8

INVALID_VERTEX_INDEX = 0x003FFFFF
NUM_TRAINING_DATA_PER_FRAME = 65536
TRAIN_BUFFER_SIZE = 2 * 65536
MAX_REGIONS = 32768
DEFAULT_REGION_CELLS = 4


class RootTile(C.Structure):
    _fields_ = [("key", C.c_uint64), ("entry", C.c_uint32), ("state", C.c_uint32)]


class Grid(C.Structure):
    _fields_ = [
        ("inv_mat", C.c_float * 9),
        ("vec", c_f3),
        ("world_bbox_min", c_f3),
        ("world_bbox_max", c_f3),
        ("index_bbox_min", c_i3),
        ("index_bbox_max", c_i3),
        ("max_value", C.c_float),
        ("background", C.c_float),
        ("num_root_tiles", C.c_uint32),
        ("num_upper", C.c_uint32),
        ("num_lower", C.c_uint32),
        ("num_leaves", C.c_uint32),
        ("root_tiles", C.c_void_p),
        ("upper_tables", C.c_void_p),
        ("lower_tables", C.c_void_p),
        ("leaf_values", C.c_void_p),
        ("tile_values", C.c_void_p),
        ("leaf_origins", C.c_void_p),
        ("cell_origin", c_i3),
        ("cell_dims", c_i3),
        ("cell_majorants", C.c_void_p),
        ("leaf_map_origin", c_i3),
        ("leaf_map_dims", c_i3),
        ("leaf_map", C.c_void_p),
        ("leaf_masks", C.c_void_p),
        ("lower_masks", C.c_void_p),
        ("upper_masks", C.c_void_p),
        ("map_mat_d", C.c_double * 9),
        ("map_inv_mat_d", C.c_double * 9),
        ("map_vec_d", C.c_double * 3),
        ("region_dims", c_i3),
        ("region_cells", C.c_uint32),
        ("region_unit", C.c_float),
        ("region_majorants", C.c_void_p),
    ]


NVDB_CODEC_NONE, NVDB_CODEC_ZIP, NVDB_CODEC_BLOSC = 0, 1, 2


class NvdbGridInfo(C.Structure):
    _fields_ = [
        ("grid_size", C.c_uint64),
        ("file_size", C.c_uint64),
        ("voxel_count", C.c_uint64),
        ("file_offset", C.c_uint64),
        ("grid_type", C.c_uint32),
        ("grid_class", C.c_uint32),
        ("codec", C.c_uint32),
        ("version", C.c_uint32),
        ("node_count", C.c_uint32 * 4),
        ("tile_count", C.c_uint32 * 3),
        ("segment", C.c_uint32),
        ("world_bbox", C.c_double * 6),
        ("index_bbox", C.c_int32 * 6),
        ("voxel_size", C.c_double * 3),
        ("name", C.c_char * 256),
    ]


class SdrConfig(C.Structure):
    _fields_ = [
        ("alpha_override", C.c_float),
        ("brightness_scale", C.c_float),
        ("apply_tone_map", C.c_int32),
        ("apply_srgb_gamma", C.c_int32),
        ("flip_y", C.c_int32),
    ]


class ToneMapParams(C.Structure):
    _fields_ = [
        ("accum", C.c_void_p),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("samples", C.c_uint32),
        ("exponent", C.c_float),
        ("brightness", C.c_float),
        ("out_rgba32f", C.c_void_p),
        ("out_rgba8", C.c_void_p),
        ("sdr", SdrConfig),
    ]


class FbmDesc(C.Structure):
    _fields_ = [
        ("resolution", C.c_uint32),
        ("octaves", C.c_uint32),
        ("seed", C.c_uint64),
        ("base_frequency", C.c_float),
        ("lacunarity", C.c_float),
        ("gain", C.c_float),
        ("sparse", C.c_int32),
        ("active_leaf_fraction", C.c_float),
        ("contrast", C.c_float),
    ]


class AliasEntry(C.Structure):
    _fields_ = [("second_index", C.c_uint32), ("prob_to_pick_first", C.c_float)]


class EnvLight(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("sampler", C.c_int32),
        ("power_coeff", C.c_float),
        ("rotation", C.c_float),
        ("constant_radiance", c_f3),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("texels", C.c_void_p),
        ("map_dim_x", C.c_uint32),
        ("map_dim_y", C.c_uint32),
        ("num_mip_levels", C.c_uint32),
        ("reserved0", C.c_uint32),
        ("mips", C.c_void_p),
        ("map_integral", C.c_float),
        ("alias_integral", C.c_float),
        ("alias_rows", C.c_void_p),
        ("alias_cols", C.c_void_p),
        ("alias_weights", C.c_void_p),
    ]


class SunLight(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("direction", c_f3),
        ("irradiance", c_f3),
        ("select_prob", C.c_float),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("position", c_f3),
        ("orientation", C.c_float * 9),
        ("vh", C.c_float),
        ("vw", C.c_float),
    ]


class RadianceQuery(C.Structure):
    _fields_ = [
        ("position", c_f3),
        ("normal_phi", C.c_float),
        ("normal_theta", C.c_float),
        ("vout_phi", C.c_float),
        ("vout_theta", C.c_float),
        ("roughness", C.c_float),
        ("diffuse_reflectance", c_f3),
        ("specular_reflectance", c_f3),
    ]


class TerminalInfo(C.Structure):
    _fields_ = [("throughput", c_f3), ("has_query", C.c_uint32)]


class TrainingVertexInfo(C.Structure):
    _fields_ = [
        ("local_throughput", c_f3),
        ("prev_vertex_index", C.c_uint32),
        ("path_length", C.c_uint32),
        ("is_unbiased", C.c_uint32),
    ]


class TrainingSuffixTerminal(C.Structure):
    _fields_ = [
        ("prev_vertex_index", C.c_uint32),
        ("has_query", C.c_uint32),
        ("path_length", C.c_uint32),
        ("is_unbiased", C.c_uint32),
    ]


class NrcBuffers(C.Structure):
    _fields_ = [
        ("max_num_training_suffixes", C.c_uint32),
        ("nrc_buffer_index", C.c_uint32),
        ("offset_to_select_unbiased_path", C.c_uint32),
        ("radiance_scale", C.c_float),
        ("path_termination_factor", C.c_float),
        ("num_samples", C.c_uint32),
        ("num_training_data", C.c_void_p),
        ("train_image_size", C.c_void_p),
        ("target_min_max", C.c_void_p),
        ("target_avg", C.c_void_p),
        ("inference_queries", C.c_void_p),
        ("inference_terminals", C.c_void_p),
        ("train_suffix_terminals", C.c_void_p),
        ("inferred_radiance", C.c_void_p),
        ("per_frame_contribution", C.c_void_p),
        ("train_queries", C.c_void_p * 2),
        ("train_targets", C.c_void_p * 2),
        ("train_vertex_info", C.c_void_p),
        ("data_shufflers", C.c_void_p),
        ("scratch", C.c_void_p),
    ]


class FrameParams(C.Structure):
    _fields_ = [
        ("image_width", C.c_uint32),
        ("image_height", C.c_uint32),
        ("camera", Camera),
        ("grid", Grid),
        ("density_coeff", C.c_float),
        ("scattering_albedo", C.c_float),
        ("scattering_forwardness", C.c_float),
        ("phase", C.c_int32),
        ("env", EnvLight),
        ("sun", SunLight),
        ("world_aabb_min", c_f3),
        ("world_aabb_max", c_f3),
        ("world_radius", C.c_float),
        ("rr_after_path_length", C.c_uint32),
        ("max_path_length", C.c_uint32),
        ("tracking", C.c_int32),
        ("rng_states", C.c_void_p),
        ("accum", C.c_void_p),
        ("path_vertices", C.c_void_p),
        ("stats", C.c_void_p),
        ("pixel_cost", C.c_void_p),
        ("work_order", C.c_void_p),
        ("nrc", NrcBuffers),
        ("event_eighths", C.c_uint32),
        ("tile_shape", C.c_uint32),
    ]


class HosekBand(C.Structure):
    _fields_ = [
        ("configs", (C.c_float * 9) * 11),
        ("radiances", C.c_float * 11),
        ("turbidity", C.c_float),
        ("solar_radius", C.c_float),
        ("emission_correction_factor_sky", C.c_float * 11),
        ("emission_correction_factor_sun", C.c_float * 11),
        ("albedo", C.c_float),
        ("elevation", C.c_float),
    ]


SKY_BANDS = 16

NRC_FP16, NRC_BF16, NRC_BF16_FP8, NRC_FP32 = 0, 1, 2, 3
NRC_TRIANGLE_WAVE, NRC_HASH_GRID = 0, 1


class NrcDesc(C.Structure):
    _fields_ = [
        ("num_hidden_layers", C.c_uint32),
        ("learning_rate", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("epsilon", C.c_float),
        ("l2_reg", C.c_float),
        ("ema_decay", C.c_float),
        ("precision", C.c_int32),
        ("seed", C.c_uint64),
        ("encoding", C.c_int32),
        ("loss_scale", C.c_float),
    ]


class HosekSky(C.Structure):
    _fields_ = [
        ("bands", C.c_void_p),
        ("solar_rows", C.c_void_p),
        ("limb_darkening", C.c_void_p),
        ("solar_turbidity_base", C.c_int32),
        ("cmf_x", C.c_float * SKY_BANDS),
        ("cmf_y", C.c_float * SKY_BANDS),
        ("cmf_z", C.c_float * SKY_BANDS),
        ("center_wavelengths", C.c_float * SKY_BANDS),
        ("integral_cmf", C.c_float),
    ]


STRUCTS = {
    "rtc8x_nrc_desc": NrcDesc,
    "rtc8x_hosek_band": HosekBand,
    "rtc8x_hosek_sky": HosekSky,
    "rtc8x_root_tile": RootTile,
    "rtc8x_grid": Grid,
    "rtc8x_fbm_desc": FbmDesc,
    "rtc8x_alias_entry": AliasEntry,
    "rtc8x_env_light": EnvLight,
    "rtc8x_sun_light": SunLight,
    "rtc8x_camera": Camera,
    "rtc8x_radiance_query": RadianceQuery,
    "rtc8x_terminal_info": TerminalInfo,
    "rtc8x_training_vertex_info": TrainingVertexInfo,
    "rtc8x_training_suffix_terminal": TrainingSuffixTerminal,
    "rtc8x_nrc_buffers": NrcBuffers,
    "rtc8x_frame_params": FrameParams,
    "rtc8x_nvdb_grid_info": NvdbGridInfo,
    "rtc8x_sdr_config": SdrConfig,
    "rtc8x_tone_map_params": ToneMapParams,
}

_P = C.c_void_p
_S = C.c_int  # rtc8x_status

# name -> (restype, argtypes); every symbol declared in include/rtc8x.h
PROTOTYPES = {
    "rtc8x_abi_version": (C.c_int32, []),
    "rtc8x_last_error": (C.c_char_p, []),
    "rtc8x_sizeof": (C.c_size_t, [C.c_char_p]),
    "rtc8x_device_count": (_S, [C.POINTER(C.c_int32)]),
    "rtc8x_set_device": (_S, [C.c_int32]),
    "rtc8x_device_name": (_S, [C.c_char_p, C.c_size_t]),
    "rtc8x_malloc": (_S, [C.POINTER(C.c_void_p), C.c_size_t]),
    "rtc8x_free": (_S, [_P]),
    "rtc8x_memcpy_htod": (_S, [_P, _P, C.c_size_t]),
    "rtc8x_memcpy_dtoh": (_S, [_P, _P, C.c_size_t]),
    "rtc8x_memcpy_dtod": (_S, [_P, _P, C.c_size_t]),
    "rtc8x_memset": (_S, [_P, C.c_int32, C.c_size_t]),
    "rtc8x_copy_async": (_S, [_P, _P, C.c_size_t, _P]),
    "rtc8x_gather_probe": (_S, [_P, C.POINTER(Grid), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "rtc8x_march_probe": (_S, [_P, C.POINTER(FrameParams), _P, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "rtc8x_device_synchronize": (_S, []),
    "rtc8x_stream_create": (_S, [C.POINTER(C.c_void_p)]),
    "rtc8x_stream_destroy": (_S, [_P]),
    "rtc8x_stream_synchronize": (_S, [_P]),
    "rtc8x_event_create": (_S, [C.POINTER(C.c_void_p)]),
    "rtc8x_event_destroy": (_S, [_P]),
    "rtc8x_event_record": (_S, [_P, _P]),
    "rtc8x_event_elapsed_ms": (_S, [C.POINTER(C.c_float), _P, _P]),
    "rtc8x_stream_wait_event": (_S, [_P, _P]),
    "rtc8x_grid_build_fbm": (_S, [C.POINTER(FbmDesc), C.POINTER(Grid), _P]),
    "rtc8x_grid_upload": (_S, [C.POINTER(Grid), C.POINTER(Grid)]),
    "rtc8x_grid_destroy": (_S, [C.POINTER(Grid)]),
    "rtc8x_grid_build_regions": (_S, [C.POINTER(Grid), C.c_uint32]),
    "rtc8x_grid_probe": (_S, [C.POINTER(Grid), C.c_uint32, _P, _P, _P, _P, _P]),
    "rtc8x_nvdb_list": (_S, [_P, C.c_size_t, C.POINTER(NvdbGridInfo), C.c_uint32, C.POINTER(C.c_uint32)]),
    "rtc8x_nanovdb_decode": (_S, [_P, C.c_size_t, C.POINTER(Grid)]),
    "rtc8x_nvdb_read_grid": (_S, [_P, C.c_size_t, C.c_uint32, C.POINTER(Grid)]),
    "rtc8x_grid_free_host": (_S, [C.POINTER(Grid)]),
    "rtc8x_grid_upload_nanovdb": (_S, [_P, C.c_size_t, C.POINTER(Grid)]),
    "rtc8x_grid_load_nvdb": (_S, [_P, C.c_size_t, C.c_uint32, C.POINTER(Grid)]),
    "rtc8x_grid_download": (_S, [C.POINTER(Grid), C.POINTER(Grid)]),
    "rtc8x_nanovdb_encode": (_S, [C.POINTER(Grid), C.c_char_p, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "rtc8x_hosek_model_create": (_S, [_P, C.c_size_t, C.POINTER(C.c_void_p)]),
    "rtc8x_hosek_model_destroy": (_S, [_P]),
    "rtc8x_hosek_bands": (_S, [_P, C.c_float, C.c_float, _P]),
    "rtc8x_hosek_model_state": (_S, [_P, C.c_float, C.c_float, C.POINTER(HosekSky), _P]),
    "rtc8x_tone_map": (_S, [_P, C.POINTER(ToneMapParams)]),
    "rtc8x_sdr_convert": (_S, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(SdrConfig), _P]),
    "rtc8x_png_encode": (_S, [_P, C.c_uint32, C.c_uint32, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "rtc8x_bmp_encode": (_S, [_P, C.c_uint32, C.c_uint32, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "rtc8x_exr_encode": (_S, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_int32, _P, C.c_size_t,
                              C.POINTER(C.c_size_t)]),
    "rtc8x_nvdb_write": (_S, [C.POINTER(C.c_void_p), C.POINTER(C.c_size_t), C.c_uint32, C.c_uint32, _P, C.c_size_t,
                              C.POINTER(C.c_size_t)]),
    "rtc8x_env_build_sampler": (_S, [C.POINTER(EnvLight), _P]),
    "rtc8x_env_release": (_S, [C.POINTER(EnvLight)]),
    "rtc8x_camera_setup": (_S, [C.POINTER(Camera), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.c_float]),
    "rtc8x_trace": (_S, [_P, C.POINTER(FrameParams), C.c_int32, C.c_uint32]),
    "rtc8x_trace_batch": (_S, [_P, _P, _P, _P, C.c_uint32, C.c_int32]),
    "rtc8x_trace_batch_static": (_S, [_P, _P, _P, _P, C.c_uint32, C.c_int32]),
    "rtc8x_order_by_cost": (_S, [_P, _P, C.c_uint32, _P, _P]),
    "rtc8x_order_scratch_bytes": (C.c_size_t, [C.c_uint32]),
    "rtc8x_order_by_tile_cost": (_S, [_P, _P, C.c_uint32, C.c_uint32, _P, _P]),
    "rtc8x_order_tiles_scratch_bytes": (C.c_size_t, [C.c_uint32, C.c_uint32]),
    "rtc8x_seed_rng": (_S, [_P, C.c_uint32, C.c_uint32, C.c_uint64]),
    "rtc8x_sky_bake": (_S, [_P, C.POINTER(HosekSky), C.POINTER(C.c_float), C.c_uint32, C.c_uint32, _P]),
    "rtc8x_sky_eval": (_S, [_P, C.POINTER(HosekSky), C.POINTER(C.c_float), C.c_uint32, _P, _P, _P]),
    "rtc8x_nrc_prepare": (_S, [_P, C.POINTER(FrameParams), C.c_int32]),
    "rtc8x_nrc_propagate": (_S, [_P, C.POINTER(FrameParams)]),
    "rtc8x_nrc_shuffle": (_S, [_P, C.POINTER(FrameParams)]),
    "rtc8x_nrc_accumulate": (_S, [_P, C.POINTER(FrameParams)]),
    "rtc8x_nrc_scratch_bytes": (C.c_size_t, [C.c_uint32, C.c_uint32]),
    "rtc8x_nrc_create": (_S, [C.POINTER(NrcDesc), C.POINTER(C.c_void_p)]),
    "rtc8x_nrc_destroy": (_S, [_P]),
    "rtc8x_nrc_infer": (_S, [_P, _P, _P, C.c_uint32, _P]),
    "rtc8x_nrc_infer_training": (_S, [_P, _P, C.POINTER(FrameParams)]),
    "rtc8x_nrc_infer_render": (_S, [_P, _P, C.POINTER(FrameParams)]),
    "rtc8x_nrc_train": (_S, [_P, _P, _P, _P, C.c_uint32, C.POINTER(C.c_float)]),
    "rtc8x_nrc_get_loss_scale": (_S, [C.c_void_p, C.POINTER(C.c_float)]),
    "rtc8x_nrc_num_params": (C.c_uint32, [_P]),
    "rtc8x_nrc_get_weights": (_S, [_P, C.c_int32, _P, C.c_size_t]),
    "rtc8x_nrc_set_weights": (_S, [_P, _P, C.c_size_t]),
    "rtc8x_nrc_encode": (_S, [_P, _P, C.c_uint32, _P]),
    "rtc8x_nrc_encode_net": (_S, [_P, _P, _P, C.c_uint32, _P, C.c_int32]),
    "rtc8x_nrc_gradient": (_S, [_P, _P, _P, _P, C.c_uint32, _P]),
}
