"""ctypes binding of libjellyfysh_amd.so (the C ABI in include/jellyfysh_amd.h).

The library is the only compute path: if it is missing, or no HIP device is visible when a
device call is made, the functions here raise instead of computing anything elsewhere.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjellyfysh_amd.so")

MAX_TAGGERS = 16
MAX_LEAVES = 4
MAX_FACTOR_MAP = 8
SYSTEMS_NOALLOC_MAX = 1024   # ECMC_SYSTEMS_NOALLOC_MAX
(TAG_PAIR_CELLS, TAG_PAIR_ALL, TAG_CELL_BOUNDARY, TAG_SAMPLING, TAG_END_OF_CHAIN, TAG_END_OF_RUN,
 TAG_START_OF_RUN, TAG_SURPLUS, TAG_CELL_BOUNDING, TAG_CELL_VETO, TAG_FACTOR_MAP) = range(11)
EH_TWO_LEAF, EH_TWO_LEAF_BOUNDING, EH_CELL_BOUNDING, EH_CELL_VETO = 0, 1, 2, 3
(POT_HARD_SPHERE, POT_INVERSE_POWER, POT_LENNARD_JONES, POT_INVERSE_POWER_TRUNC, POT_MERGED_IMAGE_COULOMB,
 POT_INVERSE_POWER_COULOMB_BOUND, POT_DISPLACED_EVEN_POWER, POT_CELL_BOUNDING, POT_HARD_DIPOLE) = range(9)
SCHED_LIST, SCHED_HEAP = 0, 1

STATUS = {0: "ok", -1: "argument", -2: "hip", -3: "configuration", -4: "scheduler", -5: "tag activator",
          -6: "state", -7: "numeric"}


class EcmcError(RuntimeError):
    """A C-ABI call returned a non-zero ecmc_status."""

    def __init__(self, status, message):
        super().__init__("%s error (%d): %s" % (STATUS.get(status, "unknown"), status, message))
        self.status = status


class TaggerDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("event_handler", ctypes.c_int32), ("potential", ctypes.c_int32),
                ("bounding_potential", ctypes.c_int32), ("par", ctypes.c_double * 4), ("bpar", ctypes.c_double * 4),
                ("use_charge", ctypes.c_int32), ("n_create", ctypes.c_int32), ("create", ctypes.c_int32 * MAX_TAGGERS),
                ("n_trash", ctypes.c_int32), ("trash", ctypes.c_int32 * MAX_TAGGERS), ("n_handlers", ctypes.c_int32),
                ("value", ctypes.c_double), ("ivalue", ctypes.c_int32), ("ivalue2", ctypes.c_int32),
                ("fmap_local", ctypes.c_int32), ("fmap_n", ctypes.c_int32 * MAX_LEAVES),
                ("fmap_target", (ctypes.c_int32 * MAX_FACTOR_MAP) * MAX_LEAVES)]


class SystemDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("n", ctypes.c_int32), ("system_length", ctypes.c_double),
                ("beta", ctypes.c_double), ("cells_per_side", ctypes.c_int32 * 3), ("neighbor_layers", ctypes.c_int32),
                ("max_occupants", ctypes.c_int32), ("scheduler", ctypes.c_int32), ("rng_skip", ctypes.c_int32),
                ("leaves_per_root", ctypes.c_int32), ("seed", ctypes.c_uint64), ("n_taggers", ctypes.c_int32), ("taggers", TaggerDesc * MAX_TAGGERS)]


SCOPE_ALL, SCOPE_INTER, SCOPE_INTRA = 0, 1, 2


class EnergyTerm(ctypes.Structure):
    _fields_ = [("potential", ctypes.c_int32), ("scope", ctypes.c_int32), ("use_charge", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("par", ctypes.c_double * 4)]


EVENT_DTYPE = np.dtype([("tagger", np.int32), ("flags", np.int32), ("active", np.int32), ("target", np.int32),
                        ("tq", np.float64), ("tr", np.float64), ("pos", np.float64, (3,)), ("ncand", np.int32),
                        ("chash", np.uint32)])

# Exported symbols (kept in step with include/jellyfysh_amd.h; checked by the tests).
_SIGNATURES = None
_lib = None


def _signatures():
    d, i32, i64, u64, vp = ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_void_p
    pd, pi32, pi64 = ctypes.POINTER(d), ctypes.POINTER(i32), ctypes.POINTER(i64)
    return {
        "ecmc_last_error": (ctypes.c_char_p, []),
        "ecmc_device_count": (i32, []),
        "ecmc_system_create": (i32, [vp, pd, pd, ctypes.POINTER(vp)]),
        "ecmc_system_destroy": (None, [vp]),
        "ecmc_system_run": (i32, [vp, i64, vp, pi64]),
        "ecmc_system_set_stop_mask": (i32, [vp, i32]),
        "ecmc_system_positions": (i32, [vp, pd]),
        "ecmc_system_active": (i32, [vp, pi32, pi32, pd, pd, pi32]),
        "ecmc_system_estimators": (i32, [vp, pd]),
        "ecmc_system_rng_words": (i32, [vp, pi64]),
        "ecmc_system_last_kernel_ms": (i32, [vp, pd]),
        "ecmc_system_launch_lds": (i32, [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int32)]),
        "ecmc_system_kernel_variant": (i32, [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
        "ecmc_desc_kernel_variant": (i32, [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
        "ecmc_systems_run": (i32, [vp, i32, i64, pi64]),
        "ecmc_systems_run_noalloc": (i32, [vp, i32, i64, pi64]),
        "ecmc_system_begin_leg": (i32, [vp]),
        "ecmc_system_send_event_time": (i32, [vp, i32, pd, pd, pi32, pi32]),
        "ecmc_system_send_out_state": (i32, [vp, i32, vp]),
        "ecmc_pair_displacement": (i32, [i32, pd, i32, i64, vp, vp, vp, vp, vp, vp, d, vp, vp]),
        "ecmc_pair_derivative": (i32, [i32, pd, i32, i64, vp, vp, vp, vp, vp, d, vp, vp]),
        "ecmc_pair_time_bounds": (i32, [i32, pd, i32, i64, vp, vp, vp, vp, vp, vp, d, vp, vp, vp]),
        "ecmc_math_log": (i32, [i64, vp, vp, vp]),
        "ecmc_math_pow": (i32, [i64, vp, vp, vp, vp]),
        "ecmc_mt_words": (i32, [u64, i64, vp, vp]),
        "ecmc_fp32_math_error": (i32, [i32, ctypes.c_uint32, ctypes.c_uint32, vp, vp]),
        "ecmc_copy_bytes": (i32, [vp, vp, i64, vp]),
        "ecmc_ewald_create": (i32, [i32, i32, d, d, ctypes.POINTER(vp)]),
        "ecmc_ewald_copy": (i32, [vp, ctypes.POINTER(vp)]),
        "ecmc_ewald_destroy": (None, [vp]),
        "ecmc_ewald_size": (i32, [vp, pi64]),
        "ecmc_ewald_derivative": (i32, [vp, i64, vp, vp, vp]),
        "ecmc_batch_create": (i32, [vp, vp, pd, pd, ctypes.POINTER(vp)]),
        "ecmc_batch_destroy": (None, [vp]),
        "ecmc_batch_set_classes": (i32, [vp, i32, vp, vp]),
        "ecmc_batch_set_mode": (i32, [vp, i32]),
        "ecmc_batch_set_log": (i32, [vp, i32]),
        "ecmc_batch_run": (i32, [vp, i32]),
        "ecmc_batch_run_graph": (i32, [vp, i32, i32]),
        "ecmc_batch_last_ms": (i32, [vp, pd]),
        "ecmc_batch_set_max_blocks": (i32, [vp, i32]),
        "ecmc_separation_norms": (i32, [i32, i64, vp, d, vp, vp]),
        "ecmc_batch_set_class_timing": (i32, [vp, i32]),
        "ecmc_batch_class_ms": (i32, [vp, i32, pd, ctypes.POINTER(i32)]),
        "ecmc_batch_states": (i32, [vp, vp, vp, vp, vp, vp, vp]),
        "ecmc_batch_positions": (i32, [vp, pd]),
        "ecmc_batch_log": (i32, [vp, vp]),
        "ecmc_batch_audit": (i32, [vp, ctypes.POINTER(u64)]),
        "ecmc_batch_snapshot": (i32, [vp, vp, pi64]),
        "ecmc_batch_restore": (i32, [vp, vp, i64]),
        "ecmc_energy_workspace_bytes": (i32, [vp, i32, pi64]),
        "ecmc_energy_forward": (i32, [vp, i32, vp, vp, vp, d, vp, vp, vp]),
        "ecmc_energy_backward": (i32, [vp, i32, vp, vp, vp, d, vp, vp, vp, vp, vp]),
    }


def exported_symbols():
    return sorted(_signatures())


def lib():
    """Load the library (raises if it was not built)."""
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7); loading
        # torch first makes this library bind to that runtime, so one process holds a single
        # HIP runtime and torch tensors' device pointers are valid here.
        import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libjellyfysh_amd.so is not built (run __graft_entry__.build() or "
                               "python -m jellyfysh_amd.csrc.build); there is no other compute path")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _signatures().items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        raise EcmcError(status, lib().ecmc_last_error().decode(errors="replace"))


def require_device():
    n = lib().ecmc_device_count()
    if n < 1:
        raise RuntimeError("no HIP device is visible: the ECMC kernels need an MI355X (gfx950)")
    return n


def dptr(t):
    """Raw device pointer of a torch tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())
