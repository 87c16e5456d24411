"""ctypes binding of libplt_amd.so (the C ABI declared in include/plt_abi.h).

The library is the product: if it is missing this module raises instead of falling back to any
host implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libplt_amd.so"

# Plane indices and material types (include/plt_abi.h).
Q_WI_X, Q_WI_Y, Q_WI_Z, Q_WO_X, Q_WO_Y, Q_WO_Z, Q_LAMBDA, Q_MAT, Q_PARAM, Q_ROUGH_Q, Q_ROUGH_L, Q_COH_AREA, \
    Q_S0, Q_S1, Q_S2, Q_S3, Q_PLANES = range(17)
R_S0, R_S1, R_S2, R_S3, R_PDF, R_PLANES = range(6)
MAT_CONDUCTOR, MAT_DIELECTRIC = 3, 4  # MaterialType values (MaterialTypes.slang:36-53)

PLT_OK = 0
PLT_WARN_UNKNOWN_KEY = 1
PLT_EINVAL, PLT_ENOMEM, PLT_EHIP, PLT_ENOTIMPL, PLT_ESTATE = -1, -2, -3, -4, -5
STATUS_NAMES = {0: "PLT_OK", 1: "PLT_WARN_UNKNOWN_KEY", -1: "PLT_EINVAL", -2: "PLT_ENOMEM", -3: "PLT_EHIP",
                -4: "PLT_ENOTIMPL", -5: "PLT_ESTATE"}


MQ_WI, MQ_WO, MQ_LOBE, MQ_MAT, MQ_LAMBDA, MQ_S, MQ_AREA, MQ_PLANES = 0, 3, 6, 7, 8, 12, 28, 29
MR_S, MR_PDF, MR_LAMBDA_PDF, MR_PLANES = 0, 16, 17, 21
MSQ_WI, MSQ_MAT, MSQ_LAMBDA, MSQ_PLANES = 0, 3, 4, 5
MSR_WO, MSR_PDF, MSR_WEIGHT, MSR_LOBE, MSR_VALID, MSR_PLANES = 0, 3, 4, 5, 6, 7


class PLTError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class SpectralProfile(C.Structure):
    """plt_spectral_profile (Scene/Material/BasicMaterialData.slang:43-54)."""
    _fields_ = [("rgb", C.c_float * 3), ("intensity", C.c_float), ("bins", C.c_uint32),
                ("min_wavelength", C.c_float), ("max_wavelength", C.c_float),
                ("pdf", C.c_float * 128), ("icdf", C.c_float * 128)]


class Material(C.Structure):
    """plt_material: the reference's BasicMaterialData parameter block (binary16 fields as bits)."""
    _fields_ = [("type", C.c_uint32), ("base_color", C.c_uint16 * 4), ("specular", C.c_uint16 * 4),
                ("data1", C.c_uint16 * 4), ("data2", C.c_uint16 * 4), ("transmission", C.c_uint16 * 3),
                ("ior", C.c_uint16), ("emission_profile", C.c_uint32), ("ior_n_profile", C.c_uint32),
                ("ior_k_profile", C.c_uint32)]


class LaunchConfig(C.Structure):
    """plt_launch_config: sample-pass schedule, megakernel block size, wavefront blocks per CU, graph
    replay, solve-pass schedule and its event-area sizes."""
    _fields_ = [("schedule", C.c_uint32), ("sample_block", C.c_uint32), ("wf_blocks_per_cu", C.c_uint32),
                ("graph", C.c_uint32), ("solve_schedule", C.c_uint32), ("solve_events_per_path", C.c_uint32),
                ("solve_hits_per_path", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3), ("fov_y", C.c_float),
                ("aspect", C.c_float), ("focal_distance", C.c_float), ("jitter", C.c_float * 2),
                ("polarizer", C.c_float), ("polarizer_theta", C.c_float)]


class ProceduralSceneDesc(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("sphere_count", C.c_uint32), ("subdivision", C.c_uint32),
                ("material_kinds", C.c_uint32), ("ground_size", C.c_float), ("aspect", C.c_float)]


class SceneArrays(C.Structure):
    _fields_ = [("n_vertices", C.c_uint32), ("n_triangles", C.c_uint32), ("n_meshes", C.c_uint32),
                ("n_materials", C.c_uint32), ("n_profiles", C.c_uint32), ("n_bvh_nodes", C.c_uint32),
                ("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32)), ("tri_mesh", C.POINTER(C.c_uint32)),
                ("mesh_first_tri", C.POINTER(C.c_uint32)), ("mesh_material", C.POINTER(C.c_uint32)),
                ("materials", C.POINTER(Material)), ("profiles", C.POINTER(SpectralProfile)),
                ("camera", Camera), ("bvh_max_depth", C.c_uint32)]


class BvhInfo(C.Structure):
    """plt_bvh_info: the builder whose tree a scene holds, its size, depth and build time."""
    _fields_ = [("builder", C.c_uint32), ("n_nodes", C.c_uint32), ("max_depth", C.c_uint32),
                ("iterations", C.c_uint32), ("build_ms", C.c_float), ("pad", C.c_uint32 * 3)]


BVH_HOST_SAH, BVH_DEVICE_PLOC = 0, 1


class Light(C.Structure):
    """plt_light: the reference's LightData (Scene/Lights/LightData.slang:47-66)."""
    _fields_ = [("pos", C.c_float * 3), ("type", C.c_uint32), ("dir", C.c_float * 3), ("opening_angle", C.c_float),
                ("cos_opening_angle", C.c_float), ("cos_subtended_angle", C.c_float), ("penumbra_angle", C.c_float),
                ("intensity_profile", C.c_uint32), ("surface_area", C.c_float), ("pad", C.c_float * 3),
                ("trans", C.c_float * 16), ("trans_it", C.c_float * 16)]


class SceneDesc(C.Structure):
    _fields_ = [("n_vertices", C.c_uint32), ("n_triangles", C.c_uint32), ("n_meshes", C.c_uint32),
                ("n_materials", C.c_uint32), ("n_profiles", C.c_uint32),
                ("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32)), ("tri_mesh", C.POINTER(C.c_uint32)),
                ("mesh_material", C.POINTER(C.c_uint32)), ("materials", C.POINTER(Material)),
                ("profiles", C.POINTER(SpectralProfile)), ("camera", Camera)]


# (name, restype, argtypes) for every exported symbol of include/plt_abi.h.
SIGNATURES = [
    ("plt_last_error", C.c_char_p, []),
    ("plt_abi_version", C.c_int, []),
    ("plt_context_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("plt_context_destroy", C.c_int, [C.c_void_p]),
    ("plt_spectral_profile_build", C.c_int,
     [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.POINTER(SpectralProfile)]),
    ("plt_context_set_spectral_profiles", C.c_int, [C.c_void_p, C.POINTER(SpectralProfile), C.c_uint32]),
    ("plt_sample_generator_stream", C.c_int,
     [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("plt_sample_generator_stream_typed", C.c_int,
     [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("plt_debug_fail_upload", C.c_int, [C.c_int]),
    ("plt_bsdf_generate_queries", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("plt_bsdf_eval_queries", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("plt_scene_create_procedural", C.c_int, [C.c_void_p, C.POINTER(ProceduralSceneDesc), C.POINTER(C.c_void_p)]),
    ("plt_material_eval_batch", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
    ("plt_pass_set_schedule", C.c_int, [C.c_void_p, C.c_int]),
    ("plt_pass_set_graph", C.c_int, [C.c_void_p, C.c_int]),
    ("plt_pass_set_launch_config", C.c_int, [C.c_void_p, C.POINTER(LaunchConfig)]),
    ("plt_pass_get_launch_config", C.c_int, [C.c_void_p, C.POINTER(LaunchConfig)]),
    ("plt_pass_autotune", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(LaunchConfig),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("plt_surrogate_create", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("plt_surrogate_destroy", None, [C.c_void_p]),
    ("plt_surrogate_param_count", C.c_uint32, []),
    ("plt_surrogate_params", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int]),
    ("plt_surrogate_forward", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("plt_surrogate_gradients", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.POINTER(C.c_float),
                                          C.POINTER(C.c_float), C.c_void_p]),
    ("plt_surrogate_train_step", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_float,
                                           C.POINTER(C.c_float), C.c_void_p]),
    ("plt_surrogate_set_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                          C.POINTER(C.c_float)]),
    ("plt_surrogate_prepare", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.POINTER(C.c_uint32), C.c_void_p]),
    ("plt_surrogate_shuffle", C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("plt_surrogate_get_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                          C.POINTER(C.c_float)]),
    ("plt_surrogate_train", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                      C.c_float, C.c_uint32, C.c_int, C.POINTER(C.c_float), C.c_void_p]),
    ("plt_surrogate_generate", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]),
    ("plt_surrogate_eval", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_int, C.POINTER(C.c_uint32), C.c_void_p]),
    ("plt_mlp_gemm_nt", C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("plt_material_sample_batch", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                            C.c_void_p, C.c_void_p]),
    ("plt_scene_create", C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    ("plt_scene_create_ex", C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.c_int, C.POINTER(C.c_void_p),
                                      C.POINTER(BvhInfo)]),
    ("plt_scene_build_bvh", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(BvhInfo)]),
    ("plt_scene_get_bvh", C.c_int, [C.c_void_p, C.POINTER(BvhInfo), C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]),
    ("plt_scene_destroy", C.c_int, [C.c_void_p]),
    ("plt_scene_get_arrays", C.c_int, [C.c_void_p, C.POINTER(SceneArrays)]),
    ("plt_scene_trace", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    ("plt_scene_set_lights", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("plt_scene_set_env_map", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_float,
                                        C.c_void_p]),
    ("plt_pass_create", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    ("plt_pass_destroy", C.c_int, [C.c_void_p]),
    ("plt_pass_set_option", C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    ("plt_pass_get_option", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    ("plt_pass_option_name", C.c_char_p, [C.c_uint32]),
    ("plt_pass_set_scene", C.c_int, [C.c_void_p, C.c_void_p]),
    ("plt_pass_set_frame", C.c_int, [C.c_void_p, C.c_uint32]),
    ("plt_pass_execute", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("plt_pass_copy_bounce_records", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("plt_pass_set_profiling", C.c_int, [C.c_void_p, C.c_int]),
    ("plt_pass_get_kernel_times", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.POINTER(C.c_uint32)]),
    ("plt_pass_get_kernel_times_split", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                  C.POINTER(C.c_uint32)]),
    ("plt_pass_get_solve_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("plt_pass_set_surrogate", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("plt_scene_set_camera", C.c_int, [C.c_void_p, C.POINTER(Camera)]),
    ("plt_pass_copy_surrogate_queries", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("plt_pass_get_surrogate_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("plt_spectrum_load_spd", C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                        C.POINTER(C.c_uint32)]),
    ("plt_spectral_profile_constant", C.c_int, [C.c_float, C.POINTER(SpectralProfile)]),
    ("plt_spectral_profile_build_sampled", C.c_int, [C.c_uint32, C.c_float, C.c_float, C.POINTER(C.c_float), C.c_int,
                                                     C.POINTER(SpectralProfile)]),
    ("plt_spectral_profiles_from_material", C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(SpectralProfile),
                                                      C.POINTER(SpectralProfile)]),
    ("plt_spectral_profile_from_emitter", C.c_int, [C.c_char_p, C.c_char_p, C.c_float, C.POINTER(SpectralProfile)]),
    ("plt_pass_set_sample_pattern", C.c_int, [C.c_void_p, C.c_int, C.c_uint32]),
    ("plt_pass_get_jitter", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("plt_sample_pattern_generate", C.c_int, [C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]),
    ("plt_accumulate_create", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("plt_accumulate_destroy", C.c_int, [C.c_void_p]),
    ("plt_accumulate_set_option", C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    ("plt_accumulate_get_option", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    ("plt_accumulate_option_name", C.c_char_p, [C.c_uint32]),
    ("plt_accumulate_set_enabled", C.c_int, [C.c_void_p, C.c_int]),
    ("plt_accumulate_reset", C.c_int, [C.c_void_p]),
    ("plt_accumulate_get_frame_count", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    ("plt_accumulate_execute", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p,
                                         C.c_uint32, C.c_void_p]),
    ("plt_tonemapper_create", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("plt_tonemapper_destroy", C.c_int, [C.c_void_p]),
    ("plt_tonemapper_set_option", C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    ("plt_tonemapper_get_option", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    ("plt_tonemapper_option_name", C.c_char_p, [C.c_uint32]),
    ("plt_tonemapper_set_scene_metadata", C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
    ("plt_tonemapper_get_color_transform", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("plt_tonemapper_execute", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                         C.c_void_p]),
    ("plt_tonemapper_execute_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                                 C.c_void_p, C.c_void_p]),
]

_lib = None


def load() -> C.CDLL:
    """Loads libplt_amd.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pltfalcor_amd.build` "
                           "(the HIP library is required; there is no host fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def reload_signatures() -> C.CDLL:
    """Applies signatures registered after the first load (modules that extend SIGNATURES)."""
    lib = load()
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def check(status: int) -> int:
    """Raises PLTError for negative statuses; returns warnings (positive) unchanged."""
    if status < 0:
        msg = load().plt_last_error().decode(errors="replace")
        raise PLTError(status, msg)
    return status
