"""ctypes binding of libgila_amd.so (include/gila_amd.h).

The library is the only compute path: importing this module raises if the
built library is missing, so nothing silently falls back to a CPU path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# GILA_AMD_LIB selects another build of the library (kernel-variant timing)
LIB_PATH = os.environ.get("GILA_AMD_LIB") or os.path.join(_HERE, "libgila_amd.so")

GLA_OK = 0
GLA_ERR_SIZE = 1
GLA_ERR_NONFINITE = 2
GLA_ERR_NO_DEVICE = 3
GLA_ERR_OVERLAP = 4
GLA_ERR_SCALE = 5
GLA_ERR_TYPE = 6
GLA_ERR_ARG = 7
GLA_ERR_UNSUPPORTED = 8
GLA_ERR_HIP = 9

GLA_C64 = 1
GLA_C128 = 2

GLA_NOALLOC_SCRATCH_BYTES = 256 << 20

EXPORTED = (
    "gla_last_error", "gla_init", "gla_opr_dims_for", "gla_vol_even_gen", "gla_opr_create_from_fur", "gla_opr_create",
    "gla_opr_export_fur", "gla_opr_get_dims", "gla_opr_adjoint", "gla_opr_destroy", "gla_opr_apply",
    "gla_opr_apply_axpby", "gla_opr_apply_host", "gla_opr_apply_noalloc", "gla_opr_reserve", "gla_batch_create", "gla_batch_launch",
    "gla_batch_destroy", "gla_opr_workspace_bytes", "gla_opr_profile",
    "gla_opr_profile_seq",
    "gla_ls_apply", "gla_krylov_dots_workspace", "gla_krylov_dots", "gla_krylov_axpy",
    "gla_krylov_rows", "gla_krylov_next", "gla_krylov_givens", "gla_opr_get_vols", "gla_opr_get_opts", "gla_opr_save", "gla_opr_load",
)


class gla_vol(ctypes.Structure):
    _fields_ = [("cel", ctypes.c_int64 * 3),
                ("scl_num", ctypes.c_int64 * 3), ("scl_den", ctypes.c_int64 * 3),
                ("org_num", ctypes.c_int64 * 3), ("org_den", ctypes.c_int64 * 3)]


class gla_opts(ctypes.Structure):
    _fields_ = [("frq_phz_re", ctypes.c_double), ("frq_phz_im", ctypes.c_double),
                ("int_ord", ctypes.c_int32), ("adj_mod", ctypes.c_int32),
                ("dtype", ctypes.c_int32), ("device", ctypes.c_int32)]


class gla_dims(ctypes.Structure):
    _fields_ = [("src_cel", ctypes.c_int64 * 3), ("trg_cel", ctypes.c_int64 * 3),
                ("brn", ctypes.c_int64 * 3), ("fur_ext", ctypes.c_int64 * 3),
                ("sym", ctypes.c_int32 * 3), ("dtype", ctypes.c_int32),
                ("adj_mod", ctypes.c_int32), ("self_opr", ctypes.c_int32),
                ("dev_brn", ctypes.c_int64 * 3),
                ("src_div", ctypes.c_int64 * 3), ("trg_div", ctypes.c_int64 * 3)]


class GlaError(RuntimeError):
    """An error raised by the library; `status` is the gla_status code."""

    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


def load(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(
            f"libgila_amd.so not found at {path}: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (or make -C gilaelectromagnetics.jl_amd/csrc)")
    lib = ctypes.CDLL(path)
    vp, i64, i32, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    P = ctypes.POINTER
    lib.gla_last_error.restype = ctypes.c_char_p
    lib.gla_last_error.argtypes = []
    sigs = {
        "gla_init": [i32],
        "gla_opr_dims_for": [P(gla_vol), P(gla_vol), P(gla_dims)],
        "gla_vol_even_gen": [P(gla_vol), P(gla_vol)],
        "gla_opr_create_from_fur": [P(gla_opts), P(gla_vol), P(gla_vol), vp, i32, i64, P(vp)],
        "gla_opr_create": [P(gla_opts), P(gla_vol), P(gla_vol), P(vp)],
        "gla_opr_export_fur": [vp, vp, i64],
        "gla_opr_get_dims": [vp, P(gla_dims)],
        "gla_opr_adjoint": [vp, P(vp)],
        "gla_opr_destroy": [vp],
        "gla_opr_apply": [vp, vp, i64, vp, i64, i64, vp],
        "gla_opr_apply_axpby": [vp, vp, i64, vp, i64, i64, d, d, d, d, vp],
        "gla_opr_apply_noalloc": [vp, vp, i64, vp, i64, i64, vp],
        "gla_opr_apply_host": [vp, vp, i64, vp, i64, i64],
        "gla_opr_reserve": [vp, i64],
        "gla_batch_create": [vp, vp, vp, vp, i32, i32, P(vp)],
        "gla_batch_launch": [vp, vp],
        "gla_batch_destroy": [vp],
        "gla_opr_profile": [vp, vp, i64, vp, i64, i64, vp, i32, P(d), P(i64)],
        "gla_opr_profile_seq": [vp, vp, i64, vp, i64, i64, vp, i32, P(d), P(i64), P(d)],
        "gla_ls_apply": [vp, vp, i64, i64, vp, i64, vp, i64, i64, vp],
        "gla_krylov_dots": [i32, i64, i64, i64, vp, i64, i64, vp, i64, vp, i64, vp, vp, vp, vp],
        "gla_krylov_axpy": [i32, i64, i64, i64, vp, i64, i64, vp, i64, vp, i64, d, vp, i64, vp, vp, vp],
        "gla_krylov_rows": [i32, i64, i64, vp, i64, i64, vp, vp, i64, vp, vp],
        "gla_krylov_next": [i32, i64, i64, vp, i64, i64, vp, vp, i64, vp, vp, i64, vp],
        "gla_krylov_givens": [i64, i64, vp, vp, i64, vp, vp, vp, vp, vp],
        "gla_opr_get_vols": [vp, P(gla_vol), P(gla_vol)],
        "gla_opr_get_opts": [vp, P(gla_opts)],
        "gla_opr_save": [vp, ctypes.c_char_p],
        "gla_opr_load": [ctypes.c_char_p, i32, i32, P(vp)],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.gla_opr_workspace_bytes.argtypes = [vp, i64]
    lib.gla_opr_workspace_bytes.restype = i64
    lib.gla_krylov_dots_workspace.argtypes = [i64, i64, i64]
    lib.gla_krylov_dots_workspace.restype = i64
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def check(status):
    if status != GLA_OK:
        raise GlaError(status, lib().gla_last_error().decode())
