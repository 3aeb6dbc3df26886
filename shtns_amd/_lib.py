"""ctypes binding of libshtns_amd.so, the C-ABI library declared in include/shtns.h and
include/shtns_hip.h.  The library is loaded from this package directory; when it is
missing, importing fails loudly (there is no fallback implementation)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshtns_amd.so")

# Every symbol the headers declare, used by the ABI test and by load().
EXPORTED_SYMBOLS = [
    # shtns.h
    "nlm_calc", "nlm_cplx_calc", "shtns_verbose", "shtns_print_version", "shtns_get_build_info",
    "shtns_print_cfg", "shtns_init", "shtns_create", "shtns_set_grid", "shtns_set_grid_auto",
    "shtns_use_threads", "shtns_use_gpu", "shtns_reset", "shtns_destroy", "shtns_unset_grid",
    "shtns_malloc", "shtns_free", "sh00_1", "sh10_ct", "sh11_st", "shlm_e1", "shtns_gauss_wts",
    "legendre_sphPlm_array", "spat_to_SH", "SH_to_spat", "spat_to_SHsphtor", "SHsphtor_to_spat",
    "SHsph_to_spat", "SHtor_to_spat", "spat_to_SH_l", "SH_to_spat_l", "SHsphtor_to_spat_l",
    "SHsph_to_spat_l", "SHtor_to_spat_l", "spat_to_SHsphtor_l", "SH_to_point",
    "spat_to_SHqst", "SHqst_to_spat", "spat_to_SHqst_l", "SHqst_to_spat_l",
    "SH_to_spat_cplx", "spat_cplx_to_SH", "SHsphtor_to_spat_cplx", "spat_cplx_to_SHsphtor",
    "SHqst_to_spat_cplx", "spat_cplx_to_SHqst", "SH_to_spat_ml", "spat_to_SH_ml",
    "SHsphtor_to_spat_ml", "spat_to_SHsphtor_ml", "SHsph_to_spat_ml", "SHtor_to_spat_ml",
    "SHqst_to_spat_ml", "spat_to_SHqst_ml",
    "SH_to_point_cplx", "SH_to_grad_point", "SHqst_to_point", "SH_to_lat", "SHqst_to_lat", "SH_Zrotate",
    "legendre_sphPlm_deriv_array", "mul_ct_matrix", "st_dt_matrix", "SH_mul_mx",
    "SH_Yrotate", "SH_Yrotate90", "SH_Xrotate90", "shtns_rotation_create", "shtns_rotation_destroy",
    "shtns_rotation_set_angles_ZYZ", "shtns_rotation_set_angles_ZXZ", "shtns_rotation_set_angle_axis",
    "shtns_rotation_wigner_d_matrix", "shtns_rotation_apply_cplx", "shtns_rotation_apply_real",
    "shtns_last_error", "shtns_last_error_string",
    # shtns_hip.h
    "cu_spat_to_SH", "cu_SH_to_spat", "cu_spat_to_SHsphtor", "cu_SHsphtor_to_spat",
    "cu_spat_to_SHqst", "cu_SHqst_to_spat", "cu_SH_to_spat_cplx", "cu_spat_cplx_to_SH",
    "cu_SHsphtor_to_spat_cplx", "cu_spat_cplx_to_SHsphtor", "cu_SH_to_spat_cplx_batch", "cu_spat_cplx_to_SH_batch",
    "cu_SHsphtor_to_spat_cplx_batch", "cu_spat_cplx_to_SHsphtor_batch",
    "cushtns_clone", "cushtns_init_gpu", "cushtns_set_streams", "cushtns_release_gpu",
    "cu_SH_to_spat_batch", "cu_spat_to_SH_batch", "cu_SHsphtor_to_spat_batch",
    "cu_spat_to_SHsphtor_batch", "cu_SH_to_spat_adjoint_batch", "cu_spat_to_SH_adjoint_batch",
    "shtns_profile", "shtns_profile_get",
    "cu_spat_to_SH_float", "cu_SH_to_spat_float", "cu_SH_to_spat_float_batch", "cu_spat_to_SH_float_batch",
    "cushtns_set_precision", "cushtns_tune", "cushtns_tuning", "cushtns_set_tuning", "cushtns_batch_create", "cushtns_batch_arrays",
    "cushtns_batch_SH_to_spat", "cushtns_batch_spat_to_SH", "cushtns_batch_destroy",
    "cu_SH_to_spat_batch_noalloc",
]

class Cplx(ctypes.Structure):
    """C99 double complex (returned in two SSE registers, as a struct of two doubles)."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


# public prefix of struct shtns_info (shtns.h:79-94)
class ShtnsInfo(ctypes.Structure):
    _fields_ = [
        ("nlm", ctypes.c_uint),
        ("lmax", ctypes.c_ushort),
        ("mmax", ctypes.c_ushort),
        ("mres", ctypes.c_ushort),
        ("nlat_2", ctypes.c_ushort),
        ("nlat", ctypes.c_uint),
        ("nphi", ctypes.c_uint),
        ("nspat", ctypes.c_uint),
        ("li", ctypes.POINTER(ctypes.c_ushort)),
        ("mi", ctypes.POINTER(ctypes.c_ushort)),
        ("ct", ctypes.POINTER(ctypes.c_double)),
        ("st", ctypes.POINTER(ctypes.c_double)),
        ("nlat_padded", ctypes.c_uint),
        ("nlm_cplx", ctypes.c_uint),
    ]


_lib = None


def load():
    """Load libshtns_amd.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(
            f"shtns_amd: {LIB_PATH} is missing; build it with `make -C shtns_amd/csrc` "
            "(or __graft_entry__.build()).  There is no fallback implementation.")
    # torch ships its own HIP runtime.  With both in one process, torch's must be loaded
    # first: if this library initializes the GPU before torch is imported, torch no longer
    # sees the GPU afterwards (measured).  Importing torch (not initializing it) suffices.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    P, I, D, L = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
    Cplx_ = Cplx
    CFG = ctypes.POINTER(ShtnsInfo)
    sig = {
        "nlm_calc": (L, [L, L, L]),
        "nlm_cplx_calc": (L, [L, L, L]),
        "shtns_verbose": (None, [I]),
        "shtns_get_build_info": (ctypes.c_char_p, []),
        "shtns_print_cfg": (None, [CFG]),
        "shtns_create": (CFG, [I, I, I, I]),
        "shtns_set_grid": (I, [CFG, I, D, I, I]),
        "shtns_set_grid_auto": (I, [CFG, I, D, I, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "shtns_init": (CFG, [I, I, I, I, I, I]),
        "shtns_use_threads": (I, [I]),
        "shtns_use_gpu": (I, [I]),
        "shtns_destroy": (None, [CFG]),
        "shtns_reset": (None, []),
        "shtns_unset_grid": (None, [CFG]),
        "shtns_malloc": (P, [ctypes.c_size_t]),
        "shtns_free": (None, [P]),
        "sh00_1": (D, [CFG]),
        "sh10_ct": (D, [CFG]),
        "sh11_st": (D, [CFG]),
        "shlm_e1": (D, [CFG, I, I]),
        "shtns_gauss_wts": (I, [CFG, P]),
        "legendre_sphPlm_array": (I, [CFG, I, I, D, P]),
        "SH_to_point": (D, [CFG, P, D, D]),
        "SH_to_point_cplx": (Cplx_, [CFG, P, D, D]),
        "SH_to_grad_point": (None, [CFG, P, P, D, D, P, P, P]),
        "SHqst_to_point": (None, [CFG, P, P, P, D, D, P, P, P]),
        "SH_to_lat": (None, [CFG, P, D, P, I, I, I]),
        "SH_Zrotate": (None, [CFG, P, D, P]),
        "SH_Yrotate": (None, [CFG, P, D, P]),
        "SH_Yrotate90": (None, [CFG, P, P]),
        "SH_Xrotate90": (None, [CFG, P, P]),
        "shtns_rotation_create": (P, [I, I, I]),
        "shtns_rotation_destroy": (None, [P]),
        "shtns_rotation_set_angles_ZYZ": (None, [P, D, D, D]),
        "shtns_rotation_set_angles_ZXZ": (None, [P, D, D, D]),
        "shtns_rotation_set_angle_axis": (None, [P, D, D, D, D]),
        "shtns_rotation_wigner_d_matrix": (I, [P, I, P]),
        "shtns_rotation_apply_cplx": (None, [P, P, P]),
        "shtns_rotation_apply_real": (None, [P, P, P]),
        "legendre_sphPlm_deriv_array": (I, [CFG, I, I, D, D, P, P]),
        "mul_ct_matrix": (None, [CFG, P]),
        "st_dt_matrix": (None, [CFG, P]),
        "SH_mul_mx": (None, [CFG, P, P, P]),
        "SH_to_spat_ml": (None, [CFG, I, P, P, I]),
        "spat_to_SH_ml": (None, [CFG, I, P, P, I]),
        "SHsphtor_to_spat_ml": (None, [CFG, I, P, P, P, P, I]),
        "spat_to_SHsphtor_ml": (None, [CFG, I, P, P, P, P, I]),
        "SHsph_to_spat_ml": (None, [CFG, I, P, P, P, I]),
        "SHtor_to_spat_ml": (None, [CFG, I, P, P, P, I]),
        "SHqst_to_spat_ml": (None, [CFG, I, P, P, P, P, P, P, I]),
        "spat_to_SHqst_ml": (None, [CFG, I, P, P, P, P, P, P, I]),
        "SHqst_to_lat": (None, [CFG, P, P, P, D, P, P, P, I, I, I]),
        "spat_to_SH": (None, [CFG, P, P]),
        "SH_to_spat": (None, [CFG, P, P]),
        "spat_to_SH_l": (None, [CFG, P, P, I]),
        "SH_to_spat_l": (None, [CFG, P, P, I]),
        "SHsphtor_to_spat": (None, [CFG, P, P, P, P]),
        "spat_to_SHsphtor": (None, [CFG, P, P, P, P]),
        "SHsphtor_to_spat_l": (None, [CFG, P, P, P, P, I]),
        "spat_to_SHsphtor_l": (None, [CFG, P, P, P, P, I]),
        "SHsph_to_spat": (None, [CFG, P, P, P]),
        "SHtor_to_spat": (None, [CFG, P, P, P]),
        "SHqst_to_spat": (None, [CFG, P, P, P, P, P, P]),
        "spat_to_SHqst": (None, [CFG, P, P, P, P, P, P]),
        "SHqst_to_spat_l": (None, [CFG, P, P, P, P, P, P, I]),
        "spat_to_SHqst_l": (None, [CFG, P, P, P, P, P, P, I]),
        "cu_SHqst_to_spat": (None, [CFG, P, P, P, P, P, P, I]),
        "cu_spat_to_SHqst": (None, [CFG, P, P, P, P, P, P, I]),
        "SH_to_spat_cplx": (None, [CFG, P, P]),
        "spat_cplx_to_SH": (None, [CFG, P, P]),
        "cu_SH_to_spat_cplx": (None, [CFG, P, P]),
        "cu_spat_cplx_to_SH": (None, [CFG, P, P]),
        "SHsphtor_to_spat_cplx": (None, [CFG, P, P, P, P]),
        "SHqst_to_spat_cplx": (None, [CFG, P, P, P, P, P, P]),
        "spat_cplx_to_SHqst": (None, [CFG, P, P, P, P, P, P]),
        "spat_cplx_to_SHsphtor": (None, [CFG, P, P, P, P]),
        "cu_SHsphtor_to_spat_cplx": (None, [CFG, P, P, P, P]),
        "cu_spat_cplx_to_SHsphtor": (None, [CFG, P, P, P, P]),
        "cu_SH_to_spat_cplx_batch": (None, [CFG, P, P, I, L, L]),
        "cu_spat_cplx_to_SH_batch": (None, [CFG, P, P, I, L, L]),
        "cu_SHsphtor_to_spat_cplx_batch": (None, [CFG, P, P, P, P, I, L, L]),
        "cu_spat_cplx_to_SHsphtor_batch": (None, [CFG, P, P, P, P, I, L, L]),
        "shtns_last_error": (I, []),
        "shtns_last_error_string": (ctypes.c_char_p, []),
        "cu_SH_to_spat": (None, [CFG, P, P, I]),
        "cu_spat_to_SH": (None, [CFG, P, P, I]),
        "cu_SHsphtor_to_spat": (None, [CFG, P, P, P, P, I]),
        "cu_spat_to_SHsphtor": (None, [CFG, P, P, P, P, I]),
        "cushtns_init_gpu": (I, [CFG]),
        "cushtns_set_streams": (None, [CFG, P, P]),
        "cushtns_clone": (CFG, [CFG, P, P]),
        "cushtns_release_gpu": (None, [CFG]),
        "cu_SH_to_spat_batch": (None, [CFG, P, P, I, I, L, L]),
        "cu_SH_to_spat_batch_noalloc": (None, [CFG, P, P, I, I, L, L]),
        "cu_spat_to_SH_batch": (None, [CFG, P, P, I, I, L, L]),
        "cu_SHsphtor_to_spat_batch": (None, [CFG, P, P, P, P, I, I, L, L]),
        "cu_spat_to_SHsphtor_batch": (None, [CFG, P, P, P, P, I, I, L, L]),
        "cu_SH_to_spat_adjoint_batch": (None, [CFG, P, P, I, I, L, L]),
        "cu_spat_to_SH_adjoint_batch": (None, [CFG, P, P, I, I, L, L]),
        "shtns_profile": (None, [CFG, I]),
        "shtns_profile_get": (I, [CFG, P, P]),
        "cu_SH_to_spat_float": (None, [CFG, P, P, I]),
        "cu_spat_to_SH_float": (None, [CFG, P, P, I]),
        "cu_SH_to_spat_float_batch": (None, [CFG, P, P, I, I, L, L]),
        "cu_spat_to_SH_float_batch": (None, [CFG, P, P, I, I, L, L]),
        "cushtns_set_precision": (I, [CFG, I]),
        "cushtns_tune": (I, [CFG, I]),
        "cushtns_tuning": (I, [CFG, P]),
        "cushtns_set_tuning": (I, [CFG, P]),
        "cushtns_batch_create": (P, [I, P, I, I]),
        "cushtns_batch_arrays": (I, [P, I, P, P]),
        "cushtns_batch_SH_to_spat": (I, [P, P]),
        "cushtns_batch_spat_to_SH": (I, [P, P]),
        "cushtns_batch_destroy": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check_error(what=""):
    lib = load()
    code = lib.shtns_last_error()
    if code:
        msg = lib.shtns_last_error_string().decode(errors="replace")
        raise RuntimeError(f"shtns_amd{(' ' + what) if what else ''}: {msg} (code {code})")
