"""ctypes binding of libpicamd.so (C ABI declared in include/pic_capi.h).

The library is loaded from ``msc-thesis_amd/lib/libpicamd.so`` next to this package.  There
is no fallback: if the library is missing the import raises, so nothing can silently run
on another code path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libpicamd.so")

PIC_OK = 0
PIC_ERR_INVALID_CONFIG = -1
PIC_ERR_OOM = -2
PIC_ERR_HIP = -3
PIC_ERR_OVERFLOW = -4
PIC_ERR_NONFINITE = -5
PIC_ERR_STATE = -6
PIC_ERR_UNSUPPORTED = -7
PIC_ERR_ARG = -8

PIC_FP32 = 0
PIC_FP64 = 1
PIC_COLLIDE_NONE = 0
PIC_COLLIDE_MCC = 1
PIC_COLLIDE_DSMC = 2
PIC_GRID_RHO, PIC_GRID_PHI, PIC_GRID_EX, PIC_GRID_EY = 0, 1, 2, 3
PIC_DIAG_COLS = 8
(PIC_DIAG_FIELD_ENERGY, PIC_DIAG_KINETIC, PIC_DIAG_PX, PIC_DIAG_PY, PIC_DIAG_MODE1,
 PIC_DIAG_NPART, PIC_DIAG_NCOLL) = range(7)
DIAG_NAMES = ("field_energy", "kinetic", "px", "py", "mode1", "npart", "ncoll", "reserved")
STAGE_NAMES = ("push_deposit", "fft_rows", "fft_cols", "fft_field", "sort", "collide")

_ERR_NAMES = {
    PIC_ERR_INVALID_CONFIG: "PIC_ERR_INVALID_CONFIG",
    PIC_ERR_OOM: "PIC_ERR_OOM",
    PIC_ERR_HIP: "PIC_ERR_HIP",
    PIC_ERR_OVERFLOW: "PIC_ERR_OVERFLOW",
    PIC_ERR_NONFINITE: "PIC_ERR_NONFINITE",
    PIC_ERR_STATE: "PIC_ERR_STATE",
    PIC_ERR_UNSUPPORTED: "PIC_ERR_UNSUPPORTED",
    PIC_ERR_ARG: "PIC_ERR_ARG",
}


class PicError(RuntimeError):
    """A C-ABI call returned a negative status; ``code`` holds it."""

    def __init__(self, code, msg):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class pic_config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double),
        ("dt", ctypes.c_double), ("qm", ctypes.c_double), ("qp", ctypes.c_double),
        ("capacity", ctypes.c_int64),
        ("precision", ctypes.c_int32), ("device", ctypes.c_int32),
        ("sort_every", ctypes.c_int32), ("tile", ctypes.c_int32),
        ("halo", ctypes.c_int32), ("splits", ctypes.c_int32),
        ("collide", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("mcc_ng", ctypes.c_double), ("mcc_v_unit", ctypes.c_double),
        ("mcc_dt_phys", ctypes.c_double), ("mcc_neutral_mass", ctypes.c_double),
        ("dsmc_coeff", ctypes.c_double), ("dsmc_gmax0", ctypes.c_double),
    ]


class pic_grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("lx", ctypes.c_double), ("ly", ctypes.c_double)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_CFG = ctypes.POINTER(pic_config)
_GRID = ctypes.POINTER(pic_grid)

# name -> (restype, argtypes)
SIGNATURES = {
    "pic_last_error": (ctypes.c_char_p, []),
    "pic_version": (_I, []),
    "pic_config_default": (None, [_CFG]),
    "pic_mcc_preset": (_I, [_CFG, _D, _D, _D]),
    "pic_mcc_nu_max": (_D, [_CFG]),
    "pic_create": (_I, [_CFG, ctypes.POINTER(_P)]),
    "pic_batch_create": (_I, [_CFG, _I, ctypes.POINTER(_P)]),
    "pic_destroy": (_I, [_P]),
    "pic_num_runs": (_I, [_P, ctypes.POINTER(_I)]),
    "pic_set_particles": (_I, [_P, _I, _DP, _DP, _DP, _DP, _I64]),
    "pic_set_particles_device": (_I, [_P, _I, _P, _P, _P, _P, _I64]),
    "pic_get_particles": (_I, [_P, _I, _DP, _DP, _DP, _DP, _I64, _I]),
    "pic_get_particles_device": (_I, [_P, _I, _P, _P, _P, _P, _I64, _I]),
    "pic_num_particles": (_I, [_P, _I, ctypes.POINTER(_I64)]),
    "pic_step": (_I, [_P, _I]),
    "pic_sync": (_I, [_P]),
    "pic_check_finite": (_I, [_P]),
    "pic_collision_counts": (_I, [_P, _I, ctypes.POINTER(_I64)]),
    "pic_get_grid": (_I, [_P, _I, _I, _DP]),
    "pic_get_diagnostics": (_I, [_P, _I, _I64, _I, _DP]),
    "pic_current_step": (_I, [_P, _I, ctypes.POINTER(_I64)]),
    "pic_stream": (_I, [_P, ctypes.POINTER(_P)]),
    "pic_use_graphs": (_I, [_P, _I]),
    "pic_get_tuning": (_I, [_P] + [ctypes.POINTER(_I)] * 4),
    "pic_set_tuning": (_I, [_P, _I, _I, _I, _I]),
    "pic_sort_fused": (_I, [_P, ctypes.POINTER(_I)]),
    "pic_autotune": (_I, [_P, _I, _DP]),
    "pic_stage_times": (_I, [_P, _I, _DP]),
    "pic_run_compat": (_I, [_DP, _DP, _DP, _DP, _DP, _DP, _I64, _I, _CFG, _DP]),
    "pic_lattice_count": (_I64, [_I, _I, _I]),
    "pic_load_landau": (_I, [_I, _I, _D, _D, _I, _D, _D, _I, ctypes.c_uint64,
                             _DP, _DP, _DP, _DP]),
    "pic_load_two_stream": (_I, [_I, _I, _D, _D, _I, _D, _D, _D, _I, ctypes.c_uint64,
                                 _DP, _DP, _DP, _DP]),
    "pic_load_reference": (_I, [_I, _I, _I, _I, _I, _I, _D, _D, _D, _D, _DP, _DP, _DP, _DP]),
    "pic_op_cell_keys": (_I, [_P, _P, _I64, _I, _GRID, _P, _P]),
    "pic_op_sort": (_I, [_P, _I64, _I, _P, _P, _P, _P]),
    "pic_op_deposit": (_I, [_P, _P, _I64, _I, _GRID, _D, _P, _P]),
    "pic_op_deposit_noalloc": (_I, [_P, _P, _I64, _I, _GRID, _D, _P, _P]),
    "pic_op_mcc_reference": (_I, [_P, _P, _I64, _D, _P, _I64, ctypes.POINTER(_I64), _P]),
    "pic_op_dsmc_reference": (_I, [_P, _P, _P, _P, _I64, _GRID, _D, _P, _I64,
                                   ctypes.POINTER(_I64), ctypes.POINTER(_I64), _P]),
    "pic_op_poisson": (_I, [_P, _P, _GRID, _P]),
    "pic_op_efield": (_I, [_P, _P, _P, _GRID, _P]),
    "pic_op_gather_push": (_I, [_P, _P, _P, _P, _I64, _I, _P, _P, _GRID, _D, _D, _D, _P, _P]),
    "pic_op_mcc": (_I, [_P, _P, _P, _P, _P, ctypes.POINTER(_I64), _I64, _CFG, _I64, _P, _P]),
    "pic_op_dsmc": (_I, [_P, _P, _P, _P, _I, _P, _D, ctypes.c_uint64, _I64, _P, _P]),
    "pic_save_checkpoint": (_I, [_P, _I, ctypes.c_char_p]),
    "pic_checkpoint_info": (_I, [ctypes.c_char_p, _CFG, ctypes.POINTER(_I64),
                                 ctypes.POINTER(_I64)]),
    "pic_load_checkpoint": (_I, [_P, _I, ctypes.c_char_p]),
    "pic_adjoint_create": (_I, [_CFG, _DP, _DP, _I64, _I, ctypes.POINTER(ctypes.c_int32),
                                ctypes.POINTER(ctypes.c_int32), _I, _I, ctypes.POINTER(_P)]),
    "pic_adjoint_destroy": (_I, [_P]),
    "pic_adjoint_forward": (_I, [_P, _P, _P, _P, _DP]),
    "pic_adjoint_backward": (_I, [_P, _D, _P, _P, _P]),
    "pic_adjoint_stream": (_I, [_P, ctypes.POINTER(_P)]),
}


def _share_torch_hip_runtime():
    """Load torch's HIP runtime first when torch is installed.

    torch ships its own libamdhip64 / libhsa-runtime64 with the same sonames as the system
    ROCm ones.  Loaded first, they satisfy libpicamd.so's dependencies too, so the process
    has one HIP runtime and torch tensors and engine buffers live in one device context.
    Loaded second, torch would bring a second runtime into the process beside the one
    libpicamd.so pulled in, and the second runtime can fail to open the device.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load():
    _share_torch_hip_runtime()
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libpicamd.so not found at {LIB_PATH}; build it with __graft_entry__.build() "
            "(the engine has no other code path)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()


def check(rc):
    """Raise PicError for a negative status from the C ABI."""
    if rc != PIC_OK:
        msg = lib.pic_last_error()
        raise PicError(rc, msg.decode() if msg else "")
    return rc


def default_config(**kw):
    c = pic_config()
    lib.pic_config_default(ctypes.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(f"pic_config has no field {k}")
        setattr(c, k, v)
    return c
