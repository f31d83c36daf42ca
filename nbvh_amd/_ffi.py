"""ctypes binding of libnbvh_hip.so (the C ABI declared in include/nbvh.h).

The library is loaded from this package directory only (it is built in-tree by
nbvh_amd.build); a missing library raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnbvh_hip.so")

# ---- status codes (nbvh_status) ----
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_HIP = 2
ERR_NO_DEVICE = 3
ERR_OUT_OF_MEMORY = 4
ERR_FORMAT = 5
ERR_STATE = 6
ERR_UNSUPPORTED = 7
ERR_TIMEOUT = 8
ERR_CAPACITY = 9

# ---- enums ----
PREC_FP32, PREC_FP16, PREC_BF16, PREC_FP8 = 0, 1, 2, 3
QUERY_SORT_RAYS = 1
QUERY_WAVEFRONT = 2
QUERY_ARRAY_ORDER = 4
NOALLOC_MAX_SORTED_RAYS = 524288  # NBVH_NOALLOC_MAX_SORTED_RAYS
BACKGROUND_CONSTANT, BACKGROUND_SKY = 0, 1
TRAIN_GRADIENTS_ONLY = 1
TRAIN_TABLES, TRAIN_MLP, TRAIN_GRAD_TABLES, TRAIN_GRAD_MLP = 0, 1, 2, 3
QUERY_NEURAL_CLOSEST, QUERY_NEURAL_ANY, QUERY_TRIANGLE_CLOSEST, QUERY_TRIANGLE_ANY = 0, 1, 2, 3
HIT_NONE, HIT_NEURAL, HIT_TRIANGLE = 0, 1, 2
FLAG_STACK_OVERFLOW, FLAG_STEP_CAP = 1, 2
SEGMENT_EQUIDISTANT_INTERIOR, SEGMENT_EQUIDISTANT_EXTERIOR, SEGMENT_UNIFORM = 0, 1, 2
DEPTH_LOCAL_T, DEPTH_WORLD_T = 0, 1
NO_NEURAL_CHILD = 0xFFFFFFFF

# ---- numpy record layouts (identical to the C structs) ----
NODE_DTYPE = np.dtype([("aabb_min", "<f4", 3), ("prim_count", "<u4"), ("aabb_max", "<f4", 3),
                       ("left_or_prim_offset", "<u4")])
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim_id", "<u4"), ("node_id", "<u4"), ("info", "<u4"),
                      ("normal_or_uv", "<f4", 3), ("logit", "<f4"), ("albedo", "<f4", 3), ("reserved", "<f4")])
assert NODE_DTYPE.itemsize == 32 and RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 48


class GridConfig(ctypes.Structure):
    _fields_ = [("n_levels", ctypes.c_uint32), ("n_features_per_level", ctypes.c_uint32),
                ("log2_hashmap_size", ctypes.c_uint32), ("base_resolution", ctypes.c_uint32),
                ("per_level_scale", ctypes.c_float)]


class NetworkConfig(ctypes.Structure):
    _fields_ = [("grid", GridConfig), ("n_points", ctypes.c_uint32), ("segment_distribution", ctypes.c_uint32),
                ("n_neurons", ctypes.c_uint32), ("n_hidden_layers", ctypes.c_uint32),
                ("n_outputs", ctypes.c_uint32), ("depth_mode", ctypes.c_uint32),
                ("depth_linear_activation", ctypes.c_uint32), ("interpolate_depth", ctypes.c_uint32),
                ("depth_interpolation_eps", ctypes.c_float)]


class QueryParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("precision", ctypes.c_uint32), ("lod", ctypes.c_uint32),
                ("max_traversal_steps", ctypes.c_uint32), ("visibility_threshold", ctypes.c_float),
                ("image_width", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("segment_seed", ctypes.c_uint32)]


class Camera(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("du", ctypes.c_float * 3), ("dv", ctypes.c_float * 3),
                ("dir_top_left", ctypes.c_float * 3)]


class CreateInfo(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("reserved0", ctypes.c_uint32), ("hbm_budget_bytes", ctypes.c_uint64)]


class QueryPart(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_void_p), ("hits", ctypes.c_void_p), ("n_rays", ctypes.c_uint32),
                ("any_hit", ctypes.c_uint32)]


class PathParams(ctypes.Structure):
    _fields_ = [("query", QueryParams), ("bounce_query_flags", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("rr_start_depth", ctypes.c_uint32), ("sample", ctypes.c_uint32), ("background", ctypes.c_uint32),
                ("background_color", ctypes.c_float * 3), ("environment_albedo", ctypes.c_float * 3),
                ("normal_offset", ctypes.c_float), ("path_lod_switch", ctypes.c_uint32),
                ("lod_switch", ctypes.c_uint32), ("secondary_visibility_threshold", ctypes.c_float),
                ("reserved", ctypes.c_uint32), ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_uint64)]


class PathStats(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_uint32 * 16), ("reserved", ctypes.c_uint32 * 16)]


class TrainConfig(ctypes.Structure):
    _fields_ = [("learning_rate", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("epsilon", ctypes.c_float), ("l2_reg", ctypes.c_float), ("max_batch", ctypes.c_uint32),
                ("decay_start", ctypes.c_uint32), ("decay_interval", ctypes.c_uint32), ("decay_base", ctypes.c_float),
                ("reserved", ctypes.c_uint32)]


class LearnerConfig(ctypes.Structure):
    _fields_ = [("batch_size", ctypes.c_uint32), ("log2_bvh_max_traversals", ctypes.c_uint32),
                ("sampling_bbox_scaling", ctypes.c_float), ("fixed_node_dilation", ctypes.c_float),
                ("node_extent_dilation", ctypes.c_float), ("learning_mint_epsilon", ctypes.c_float),
                ("learning_maxt_epsilon", ctypes.c_float), ("max_training_bounces", ctypes.c_uint32),
                ("training_node_survival", ctypes.c_uint32), ("survival_probability_scale", ctypes.c_float),
                ("min_rr_survival_prob", ctypes.c_float), ("tree_cut_learning", ctypes.c_uint32),
                ("neural_bvh_min_depth", ctypes.c_uint32), ("num_lods_learned", ctypes.c_uint32),
                ("max_training_steps", ctypes.c_uint32), ("lod_decision", ctypes.c_uint32),
                ("splitting_interval", ctypes.c_int32), ("last_split_iteration", ctypes.c_int32),
                ("start_split_num", ctypes.c_int32), ("split_scaling", ctypes.c_float),
                ("split_interval_scaling", ctypes.c_float), ("reserved", ctypes.c_uint32 * 4)]


class LearnerState(ctypes.Structure):
    _fields_ = [("step", ctypes.c_uint32), ("n_nodes", ctypes.c_uint32), ("current_learned_lod", ctypes.c_uint32),
                ("all_lods_assigned", ctypes.c_uint32), ("tree_cut_learning", ctypes.c_uint32),
                ("current_num_splits", ctypes.c_uint32), ("total_splits", ctypes.c_uint32),
                ("sample_failures", ctypes.c_uint32), ("max_node_loss", ctypes.c_float), ("last_loss", ctypes.c_float),
                ("expected_total_splits", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 5)]


TREE_CUT_DTYPE = np.dtype([("loss", "<f4"), ("iter_alive", "<u4"), ("depth", "<u4"), ("samples", "<u4")])
LEARN_INPUTS, LEARN_TARGETS, LEARN_NODES, LEARN_TREE_CUT, LEARN_SPLITS, LEARN_NBVH = range(6)


class LaunchShape(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_uint32), ("queue_bands", ctypes.c_uint32), ("ray_order", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("n_streams", ctypes.c_uint32), ("max_batch_rays", ctypes.c_uint32),
                ("staging_bytes", ctypes.c_uint64), ("reserved", ctypes.c_uint32 * 4)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("n_launches", ctypes.c_uint32), ("batch_parts", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("submit_ns", ctypes.c_uint64), ("launch_ns", ctypes.c_uint64),
                ("complete_ns", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("n_submitted", ctypes.c_uint64), ("n_completed", ctypes.c_uint64), ("n_failed", ctypes.c_uint64),
                ("n_launches", ctypes.c_uint64), ("n_batched_requests", ctypes.c_uint64),
                ("n_rays", ctypes.c_uint64), ("n_pending", ctypes.c_uint32), ("n_inflight", ctypes.c_uint32)]


REQUEST_DEVICE_BUFFERS = 1


class NbvhError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"nbvh status {status}: {message}")
        self.status = status


_lib = None

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
SZ = ctypes.c_size_t
F32 = ctypes.c_float
PU32 = ctypes.POINTER(ctypes.c_uint32)
PU64 = ctypes.POINTER(ctypes.c_uint64)
PSZ = ctypes.POINTER(ctypes.c_size_t)

# name -> argtypes (restype int32 unless listed in _RESTYPES)
SIGNATURES = {
    "nbvh_abi_version": [],
    "nbvh_status_string": [ctypes.c_int32],
    "nbvh_last_error": [P],
    "nbvh_create": [ctypes.POINTER(CreateInfo), ctypes.POINTER(P)],
    "nbvh_destroy": [P],
    "nbvh_memory_usage": [P, PU64, PU64],
    "nbvh_mesh_icosphere": [U32, U64, F32, P, F32, P, PU32, P, PU32],
    "nbvh_mesh_geosphere": [U32, U64, F32, P, F32, P, PU32, P, PU32],
    "nbvh_bvh_build": [P, U32, P, U32, U32, P, U32, PU32, P],
    "nbvh_nbvh_cut": [P, U32, U32, P, U32, PU32],
    "nbvh_nbvh_parse": [P, SZ, P, U32, PU32],
    "nbvh_nbvh_serialize": [P, U32, P, SZ, PSZ],
    "nbvh_network_param_counts": [ctypes.POINTER(NetworkConfig), PU64, PU64],
    "nbvh_network_init_params": [ctypes.POINTER(NetworkConfig), U64, F32, P, P],
    "nbvh_tcnn_params_to_network": [ctypes.POINTER(NetworkConfig), P, U64, P, P],
    "nbvh_network_to_tcnn_params": [ctypes.POINTER(NetworkConfig), P, P, P, U64, PU64],
    "nbvh_camera_look_at": [P, P, P, F32, U32, U32, ctypes.POINTER(Camera)],
    "nbvh_scene_create": [P, P, U32, P, U32, ctypes.POINTER(P)],
    "nbvh_scene_create_empty": [P, ctypes.POINTER(P)],
    "nbvh_scene_destroy": [P],
    "nbvh_scene_get_bvh": [P, P, U32, PU32, P, U32, PU32],
    "nbvh_scene_build_nbvh": [P, U32],
    "nbvh_scene_load_nbvh": [P, P, SZ],
    "nbvh_scene_set_nbvh": [P, P, U32],
    "nbvh_scene_get_nbvh": [P, P, U32, PU32],
    "nbvh_scene_set_bbox": [P, P, P],
    "nbvh_scene_get_bbox": [P, P, P],
    "nbvh_scene_load_network": [P, ctypes.POINTER(NetworkConfig), P, U64, P, U64],
    "nbvh_scene_init_network": [P, ctypes.POINTER(NetworkConfig), U64, F32],
    "nbvh_scene_get_network": [P, ctypes.POINTER(NetworkConfig), P, U64, P, U64],
    "nbvh_raygen": [P, ctypes.POINTER(Camera), U32, U32, U32, U32, P, P],
    "nbvh_query": [P, P, ctypes.POINTER(QueryParams), P, U32, P, P],
    "nbvh_query_noalloc": [P, P, ctypes.POINTER(QueryParams), P, U32, P, P],
    "nbvh_scatter_diffuse": [P, P, P, P, U32, U32, U32, P, P],
    "nbvh_segment_labels": [P, P, P, P, P, U32, P, P],
    "nbvh_query_host": [P, P, ctypes.POINTER(QueryParams), P, U32, P],
    "nbvh_query_parts": [P, P, ctypes.POINTER(QueryParams), ctypes.POINTER(QueryPart), U32, P],
    "nbvh_query_hybrid": [P, P, P, ctypes.POINTER(QueryParams), P, U32, P, P],
    "nbvh_trainer_create": [P, P, ctypes.POINTER(TrainConfig), ctypes.POINTER(P)],
    "nbvh_trainer_destroy": [P],
    "nbvh_trainer_step": [P, P, P, U32, U32, ctypes.POINTER(F32), P],
    "nbvh_trainer_get": [P, U32, P, U64, PU64],
    "nbvh_trainer_commit": [P],
    "nbvh_flip": [P, P, P, U32, U32, F32, P, ctypes.POINTER(F32), P],
    "nbvh_learner_config_defaults": [ctypes.POINTER(LearnerConfig)],
    "nbvh_learner_create": [P, P, ctypes.POINTER(LearnerConfig), ctypes.POINTER(P)],
    "nbvh_learner_destroy": [P],
    "nbvh_learner_sample": [P, U32, P],
    "nbvh_learner_update": [P, P, P],
    "nbvh_learner_step": [P, P, ctypes.POINTER(F32), P],
    "nbvh_learner_get_state": [P, ctypes.POINTER(LearnerState)],
    "nbvh_learner_get": [P, U32, P, U64, PU64],
    "nbvh_path_trace": [P, P, P, ctypes.POINTER(PathParams), ctypes.POINTER(Camera), U32, U32, P,
                        ctypes.POINTER(PathStats), P],
    "nbvh_path_trace_workspace_size": [P, P, P, ctypes.POINTER(PathParams), U32, U32, PU64],
    "nbvh_scene_set_launch_shape": [P, U32, ctypes.POINTER(LaunchShape)],
    "nbvh_scene_get_launch_shape": [P, U32, ctypes.POINTER(LaunchShape)],
    "nbvh_autotune": [P, P, ctypes.POINTER(QueryParams), P, U32, P, P, ctypes.POINTER(LaunchShape),
                      ctypes.POINTER(ctypes.c_float)],
    "nbvh_autotune_parts": [P, P, ctypes.POINTER(QueryParams), ctypes.POINTER(QueryPart), U32, P,
                            ctypes.POINTER(LaunchShape), ctypes.POINTER(ctypes.c_float)],
    "nbvh_network_inference": [P, P, U32, P, U32, P, P, P],
    "nbvh_network_inference_host": [P, P, U32, P, U32, P, P],
    "nbvh_server_create": [P, ctypes.POINTER(ServerConfig), ctypes.POINTER(P)],
    "nbvh_server_destroy": [P],
    "nbvh_request_submit": [P, P, ctypes.POINTER(QueryParams), P, U32, P, U32, PU64],
    "nbvh_request_wait": [P, U64, ctypes.c_int64, ctypes.POINTER(RequestInfo)],
    "nbvh_server_get_stats": [P, ctypes.POINTER(ServerStats)],
}
_RESTYPES = {"nbvh_abi_version": ctypes.c_uint32, "nbvh_status_string": ctypes.c_char_p,
             "nbvh_last_error": ctypes.c_char_p}


def lib():
    """Load libnbvh_hip.so from the package directory (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not found: build it with `python -m nbvh_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, ctypes.c_int32)
        _lib = L
    return _lib


def check(status, ctx=None):
    if status != OK:
        msg = lib().nbvh_last_error(ctx)
        raise NbvhError(status, msg.decode() if msg else lib().nbvh_status_string(status).decode())


def ptr(a):
    """ctypes void pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)
