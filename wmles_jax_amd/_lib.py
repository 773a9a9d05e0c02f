"""ctypes binding of libwmles_amd.so (C ABI declared in include/wmles_amd.h).

The library is required: importing the package on a machine where it is missing or
fails to load raises immediately.  There is no CPU fallback.
"""
import ctypes as C
import os

# torch first: its HIP runtime (libamdhip64.so.7) must be the one in the process before
# the library is loaded, so that the library binds to that same runtime.  Loading the
# library first would bring in a second copy of the runtime from /opt/rocm, and two
# runtimes in one process do not share the device.
import torch  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WMX_LIB", os.path.join(_HERE, "libwmles_amd.so"))

WMX_MAX_BLOCKS = 64
FACE_INTERFACE, BC_SLIP, BC_NOSLIP, BC_DIRICHLET, BC_FARFIELD, BC_EXTRAPOLATE, BC_WALLMODEL, \
    BC_INFLOW, BC_EXTERNAL = range(9)
SCHEME_KEP, SCHEME_KEEP, SCHEME_MUSCLHLLC, SCHEME_HYBRID = range(4)
LIMITER_ALBADA, LIMITER_MINMOD = range(2)
SGS_NONE, SGS_SMAGORINSKY, SGS_WALE, SGS_VREMAN = range(4)
PREC_FP64, PREC_FP32_BULK = range(2)


class FaceDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nb", C.c_int32), ("src_axis", C.c_int32),
                ("src_side", C.c_int32), ("mapping", C.c_int32 * 3), ("flip", C.c_int32 * 3),
                ("shift", C.c_double * 3), ("prim", C.c_double * 5), ("T_wall", C.c_double),
                ("n_match", C.c_int32), ("inflow_id", C.c_int32)]


class BlockDesc(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("case_id", C.c_int32), ("faces", FaceDesc * 6),
                ("vol", C.c_void_p), ("xc", C.c_void_p), ("S", C.c_void_p * 3),
                ("xf", C.c_void_p * 3)]


class NumericsDesc(C.Structure):
    _fields_ = [("gamma", C.c_double), ("R", C.c_double), ("Pr", C.c_double), ("Pr_t", C.c_double),
                ("mu_ref", C.c_double), ("T_ref", C.c_double), ("S_suth", C.c_double),
                ("scheme", C.c_int32), ("central", C.c_int32), ("order", C.c_int32),
                ("limiter", C.c_int32), ("th0", C.c_double), ("th1", C.c_double),
                ("sgs", C.c_int32), ("sgs_c", C.c_double),
                ("kappa", C.c_double), ("B", C.c_double), ("n_iter", C.c_int32),
                ("wm_ekb", C.c_double), ("wm_pr13", C.c_double), ("wm_pr3", C.c_double),
                ("wm_beta", C.c_double),
                ("bc_passes", C.c_int32), ("cfl", C.c_double), ("precision", C.c_int32),
                ("n_sponge", C.c_int32), ("sponge", (C.c_double * 5) * 4),
                ("n_cases", C.c_int32), ("mass_flux_dir", C.c_int32 * 16),
                ("mass_flux_target", C.c_double * 16)]


class InflowDesc(C.Structure):
    _fields_ = [("n_tab", C.c_int32), ("table", C.c_void_p), ("n_modes", C.c_int32),
                ("modes", C.c_void_p), ("U_inf", C.c_double), ("L", C.c_double),
                ("rho_inf", C.c_double), ("p_inf", C.c_double)]


class WallNetDesc(C.Structure):
    _fields_ = [("w_bf16", C.c_void_p), ("w_fp32", C.c_void_p),
                ("x_center", C.c_float * 3), ("x_scale", C.c_float * 3),
                ("y_center", C.c_float * 2), ("y_scale", C.c_float * 2),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("n_polish", C.c_int32), ("tol", C.c_double), ("fallback_count", C.c_void_p)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"wmles_jax_amd: HIP library {LIB_PATH} is missing; build it with "
            "`python wmles_jax_amd/build.py` (hipcc, gfx950). There is no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    PP = C.POINTER(C.c_void_p)
    sig = {
        "wmx_last_error": ([], C.c_char_p),
        "wmx_build_info": ([], C.c_char_p),
        "wmx_plan_create": ([C.c_int32, C.POINTER(BlockDesc), C.POINTER(NumericsDesc), C.c_int32,
                             C.POINTER(InflowDesc), C.POINTER(P)], C.c_int),
        "wmx_plan_destroy": ([P], C.c_int),
        "wmx_plan_workspace_bytes": ([P, C.POINTER(C.c_size_t)], C.c_int),
        "wmx_plan_bind_workspace": ([P, P, C.c_size_t], C.c_int),
        "wmx_plan_wall_faces": ([P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "wmx_plan_fill_phases": ([P, C.POINTER(C.c_int32)], C.c_int),
        "wmx_plan_fill_passes": ([P, C.POINTER(C.c_int32)], C.c_int),
        "wmx_plan_profile": ([P, C.c_int32], C.c_int),
        "wmx_plan_set_gauge": ([P, C.c_double, C.c_double], C.c_int),
        "wmx_plan_set_case_gauge": ([P, C.c_int32, C.c_double, C.c_double], C.c_int),
        "wmx_plan_profile_read": ([P, C.POINTER(C.c_double), C.POINTER(C.c_int32)], C.c_int),
        "wmx_fill_nodes": ([C.c_int32, C.POINTER(BlockDesc), PP, C.c_int32, C.c_int32, P], C.c_int),
        "wmx_compute_geometry": ([P, C.POINTER(C.c_int32), P, P, P, P, P, P, P, P, P], C.c_int),
        "wmx_apply_bcs": ([P, PP, C.c_double, P], C.c_int),
        "wmx_rhs": ([P, PP, PP, C.c_double, P], C.c_int),
        "wmx_advance": ([P, PP, PP, C.c_double, P, C.c_double, P], C.c_int),
        "wmx_max_dt": ([P, PP, P, P], C.c_int),
        "wmx_max_dt_noalloc": ([P, PP, P, P], C.c_int),
        "wmx_plan_sweep_phases": ([P, C.POINTER(C.c_int32), P], C.c_int),
        "wmx_apply_bcs_phases": ([P, PP, C.c_int32, C.c_int32, C.c_double, P], C.c_int),
        "wmx_rhs_filled": ([P, PP, PP, P], C.c_int),
        "wmx_scheme_flux": ([C.POINTER(NumericsDesc), C.POINTER(C.c_int32), C.c_int32, P, P, P, P, P], C.c_int),
        "wmx_sgs_eval": ([C.POINTER(NumericsDesc), C.c_int64, P, P, P, P, P], C.c_int),
        "wmx_wall_flux": ([C.POINTER(BlockDesc), C.c_int32, C.c_int32, C.POINTER(NumericsDesc), P, P, P, P, P, P],
                          C.c_int),
        "wmx_step": ([P, PP, PP, P, C.c_double, C.c_int32, P], C.c_int),
        "wmx_plan_set_tiling": ([P, C.c_int32, C.c_int32], C.c_int),
        "wmx_plan_set_persistent": ([P, C.c_int32], C.c_int),
        "wmx_plan_set_fused_gradients": ([P, C.c_int32], C.c_int),
        "wmx_tile_model": ([C.c_int32, P, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "wmx_plan_tiling": ([P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "wmx_mass_flux": ([P, PP, P], C.c_int),
        "wmx_wall_stats": ([P, PP, P, P, P], C.c_int),
        "wmx_plane_stats": ([P, PP, C.c_int32, P, P], C.c_int),
        "wmx_fill_face": ([C.POINTER(BlockDesc), C.c_int32, C.c_int32, C.POINTER(NumericsDesc), P,
                           C.c_double, P], C.c_int),
        "wmx_wall_model_eval": ([C.POINTER(NumericsDesc), C.c_int32, C.c_int64, P, P, P, P,
                                 C.c_double, P, P, P], C.c_int),
        "wmx_mlp_sizes": ([C.POINTER(C.c_int32)] * 3, C.c_int),
        "wmx_wm_dataset": ([C.POINTER(NumericsDesc), C.POINTER(WallNetDesc), C.c_int64, C.c_uint64,
                            P, P, P, P], C.c_int),
        "wmx_mlp_train_step": ([P, P, P, C.c_int64, C.c_int64, C.c_int32, P, P, P, P, P, C.c_int32,
                                C.c_float, C.c_int32, P], C.c_int),
        "wmx_mlp_pack": ([P, P, P, P], C.c_int),
        "wmx_mlp_forward": ([P, P, C.c_int64, P, P], C.c_int),
        "wmx_wall_net_eval": ([C.POINTER(NumericsDesc), C.POINTER(WallNetDesc), C.c_int32, C.c_int64,
                               P, P, P, P, P, C.c_double, P, P, P, P], C.c_int),
        "wmx_plan_set_wall_net": ([P, C.POINTER(WallNetDesc)], C.c_int),
        "wmx_mfma_bf16_probe": ([P, P, P, P], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()


class WmxError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise WmxError(lib.wmx_last_error().decode())


def ptr_array(tensors):
    arr = (C.c_void_p * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = t.data_ptr()
    return arr


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


__all__ = ["lib", "check", "WmxError", "FaceDesc", "BlockDesc", "NumericsDesc", "InflowDesc", "WallNetDesc",
           "ptr_array", "stream_ptr", "LIB_PATH"]
