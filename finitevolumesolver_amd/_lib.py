"""ctypes binding of libfvs_hip.so (the C ABI declared in include/fvs.h).

The HIP library is the only compute path: importing a compute entry point
without the built library raises ImportError (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("FVS_HIP_LIB", _PKG / "lib" / "libfvs_hip.so"))

FVS_OK = 0
FVS_TIME_STEP_TOO_LARGE = 1
FVS_NEWTON_NOT_CONVERGED = 2
FVS_BAD_ARGUMENT = 3
FVS_HIP_ERROR = 4

BASE_HLLEM_LARROUTUROU = 0
BASE_HLLEM = 1
BASE_GODUNOV_EXACT = 2
BASE_HLL_EINFELDT = 3
RECON_NONE = 0
RECON_MUSCL_HANCOCK = 1
RECON_MUSCL_HANCOCK_PRIM = 2
RECON_MUSCL_HANCOCK_CHARACTERISTIC = 3
LIMITER_MINMOD = 0
LIMITER_VANLEER = 1
LIMITER_NONE = 2
LIMITER_NO_GRADIENT = 3
FP64 = 0
FP32_FLUX = 1
FP64_RCP = 2
LEVEL_TIMERS = 4
TIMER_NAMES = ("sweep_x", "sweep_y", "sweep_z", "cfl")
GUESS_MEAN_PRESSURE = 0
GUESS_LEARNED = 1
SPLIT_GODUNOV = 0
SPLIT_STRANG = 1
BOUNDARY_EXACT = 0
BOUNDARY_HLLEM = 1


class MethodDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("gamma", C.c_double),
        ("base", C.c_int),
        ("reconstruction", C.c_int),
        ("limiter", C.c_int),
        ("precision", C.c_int),
        ("quirks", C.c_int),
        ("newton_guess", C.c_int),
        ("max_newton_iterations", C.c_int),
        ("newton_tolerance", C.c_double),
        ("boundary_solver", C.c_int),
    ]


class View(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("ncomp", C.c_int),
        ("lo", C.c_int64 * 3),
        ("hi", C.c_int64 * 3),
        ("stride", C.c_int64 * 3),
        ("comp_stride", C.c_int64),
    ]


class CutCellData(C.Structure):
    """fvs_cutcell_data: CutCellData<Rank> as views (include/fvs.h)."""
    _fields_ = [
        ("volume_fractions", View),
        ("face_fractions", View * 3),
        ("boundary_normals", View),
        ("boundary_centroids", View),
        ("unshielded", View * 3),
        ("shielded_left", View * 3),
        ("shielded_right", View * 3),
        ("doubly_shielded", View * 3),
    ]


class PatchDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int64 * 3),
        ("dx", C.c_double * 3),
        ("splitting", C.c_int),
        ("cfl", C.c_double),
        ("smallest_dt", C.c_double),
    ]


# Every entry point of include/fvs.h with its ctypes signature.
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PI64 = C.POINTER(C.c_int64)
_PD = C.POINTER(C.c_double)
SIGNATURES = {
    "fvs_create": (_I, [C.POINTER(MethodDesc), _P, C.POINTER(_P)]),
    "fvs_destroy": (_I, [_P]),
    "fvs_last_error": (C.c_char_p, []),
    "fvs_version": (C.c_char_p, []),
    "fvs_stable_dt": (_I, [_P, C.POINTER(View), _D, _I, _P]),
    "fvs_numeric_fluxes": (_I, [_P, C.POINTER(View), C.POINTER(View), _D, _D, _I]),
    "fvs_update_conservatively": (_I, [_P, C.POINTER(View), C.POINTER(View), C.POINTER(View), _D, _D, _I]),
    "fvs_complete_from_cons": (_I, [_P, C.POINTER(View), C.POINTER(View)]),
    "fvs_transmissive_fill": (_I, [_P, C.POINTER(View), _PI64, _PI64, _I]),
    "fvs_level_create": (_I, [_P, C.POINTER(PatchDesc), C.POINTER(_P)]),
    "fvs_level_destroy": (_I, [_P]),
    "fvs_level_state": (_P, [_P, _I]),
    "fvs_level_set_state": (_I, [_P, C.POINTER(_P), _D]),
    "fvs_level_get_state": (_I, [_P, C.POINTER(_P)]),
    "fvs_level_advance": (_I, [_P, _I, _D, _I, _PD, C.POINTER(_I), C.POINTER(_I)]),
    "fvs_level_time": (_I, [_P, _PD, _PI64]),
    "fvs_level_hint_max_speed": (_I, [_P, _PD]),
    "fvs_level_cfl_redos": (_I, [_P, _PI64]),
    "fvs_level_set_timing": (_I, [_P, _I]),
    "fvs_level_set_graphs": (_I, [_P, _I]),
    "fvs_level_kernel_times": (_I, [_P, _PD, _PI64]),
    "fvs_split_sweep": (_I, [_P, C.POINTER(_P), C.POINTER(_P), _PI64, _D, _D, _I, _P]),
    "fvs_patch_stable_dt": (_I, [_P, C.POINTER(_P), _PI64, _PD, _P]),
    "fvs_patch_stable_dt_noalloc": (_I, [_P, C.POINTER(_P), _PI64, _PD, _P]),
    "fvs_stream_copy": (_I, [_P, C.POINTER(_P), C.POINTER(_P), _I, C.c_int64]),
    "fvs_godunov_sweep": (_I, [_P, C.POINTER(_P), C.POINTER(_P), _PI64, _D, _D, _I, _P, _P]),
    "fvs_newton_stats": (_I, [_P, C.POINTER(C.c_uint64), _I]),
    "fvs_exact_middle_state": (_I, [_P, _I, _P, _P, C.c_int64, _P, _P, _P, _P, _P]),
    "fvs_mlp_create": (_I, [_P, C.c_uint64, C.POINTER(_P)]),
    "fvs_mlp_destroy": (_I, [_P]),
    "fvs_mlp_num_params": (_I, []),
    "fvs_mlp_get_params": (_I, [_P, _P]),
    "fvs_mlp_set_params": (_I, [_P, _P]),
    "fvs_mlp_generate": (_I, [_P, C.c_int64, C.c_uint64, _D, _PI64]),
    "fvs_mlp_dataset": (_I, [_P, _PI64, _P, _P, _P, _P, _P]),
    "fvs_mlp_forward": (_I, [_P, _P, C.c_int64, _P]),
    "fvs_mlp_gradients": (_I, [_P, _P, _P, _P, C.c_int64, C.c_float, _P, _P]),
    "fvs_mlp_train": (_I, [_P, _I, C.c_int64, C.c_float, C.c_float, C.c_uint64, _P]),
    "fvs_mlp_infer_pairs": (_I, [_P, _I, _I, _P, _P, C.c_int64, _P, _P]),
    "fvs_mlp_face_guess": (_I, [_P, _I, _D, C.POINTER(_P), _PI64, _I, _P]),
    "fvs_set_mlp": (_I, [_P, _P]),
    "fvs_kbn_fill_cutcell_data": (_I, [_P, _P, _I, _PI64, _PI64]),
    "fvs_kbn_is_cut_cell": (_I, [_P, _P, _PI64, _PI64, _P]),
    "fvs_kbn_reference_states": (_I, [_P, C.POINTER(View), C.POINTER(View), _P]),
    "fvs_kbn_boundary_fluxes": (_I, [_P, C.POINTER(View), C.POINTER(View), C.POINTER(View), _P,
                                     _D, _D, _I]),
    "fvs_kbn_regular_fluxes": (_I, [_P, C.POINTER(View), C.POINTER(View), _P, _D, _D, _I]),
    "fvs_kbn_stable_dt": (_I, [_P, C.POINTER(View), _P, _D, _I, _P]),
    "fvs_kbn_cutcell_fluxes": (_I, [_P, C.POINTER(View), C.POINTER(View), C.POINTER(View),
                                    C.POINTER(View), C.POINTER(View), _P, _D, _D, _I]),
    "fvs_kbn_update_conservatively": (_I, [_P, C.POINTER(View), C.POINTER(View), C.POINTER(View),
                                           C.POINTER(View), C.POINTER(View), C.POINTER(View), _P,
                                           _D, _D, _I]),
    "fvs_kbn_update_conservatively2": (_I, [_P, C.POINTER(View), C.POINTER(View), C.POINTER(View),
                                            C.POINTER(View), _P, _D, _D, _I]),
    "fvs_kbn_complete_from_cons": (_I, [_P, C.POINTER(View), C.POINTER(View), _P]),
    "fvs_kbn_level_create": (_I, [_P, C.POINTER(PatchDesc), _P, _I, C.POINTER(_P)]),
    "fvs_kbn_level_destroy": (_I, [_P]),
    "fvs_kbn_level_ncut": (_I, [_P]),
    "fvs_kbn_level_set_state": (_I, [_P, C.POINTER(_P), _D]),
    "fvs_kbn_level_get_state": (_I, [_P, C.POINTER(_P)]),
    "fvs_kbn_level_time": (_I, [_P, _PD, _PI64]),
    "fvs_kbn_level_advance": (_I, [_P, _I, _D, _I, _PD, C.POINTER(_I), C.POINTER(_I)]),
    "fvs_batch_create": (_I, [_P, C.c_int64, _P, _P, C.POINTER(_P)]),
    "fvs_batch_destroy": (_I, [_P]),
    "fvs_batch_pool": (_P, [_P, _I]),
    "fvs_batch_pool_size": (C.c_int64, [_P]),
    "fvs_batch_offset": (C.c_int64, [_P, C.c_int64]),
    "fvs_batch_upload": (_I, [_P, _I, _P]),
    "fvs_batch_download": (_I, [_P, _I, _P]),
    "fvs_batch_set_dt": (_I, [_P, _D]),
    "fvs_batch_fill_ghosts": (_I, [_P, _I]),
    "fvs_batch_sweep": (_I, [_P, _I, _I, _P]),
    "fvs_batch_tune": (_I, [_P, _I, _I, C.POINTER(_I), C.POINTER(C.c_float)]),
    "fvs_batch_tile_shape": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "fvs_batch_capture": (_I, [_P]),
    "fvs_batch_replay": (_I, [_P, _I]),
    "fvs_batch_cfl": (_P, [_P]),
    "fvs_batch_read_cfl": (_I, [_P, _P]),
    "fvs_batch_set_shape": (_I, [_P, _I, _I]),
    "fvs_amr_create": (_I, [_P, _I, _P, _P, _P, _I, C.POINTER(_P)]),
    "fvs_amr_destroy": (_I, [_P]),
    "fvs_amr_fill_ghosts": (_I, [_P, _I, _I]),
    "fvs_amr_average_down": (_I, [_P, _I]),
    "fvs_amr_advance": (_I, [_P, _D]),
    "fvs_amr_level_cfl": (_I, [_P, _P]),
    "fvs_amr_stable_dt": (_I, [_P, _PD]),
    "fvs_amr_set_graphs": (_I, [_P, _I]),
    "fvs_amr_set_sequencing": (_I, [_P, _I]),
    "fvs_amr_cycle": (_I, [_P, C.POINTER(C.c_int64)]),
    "fvs_amr_set_cycle": (_I, [_P, C.c_int64]),
    "fvs_amr_tag_cells": (_I, [_P, _I, _D, _D, _P]),
    "fvs_amr_transfer": (_I, [_P, _P]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libfvs_hip.so once; raises ImportError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python __graft_entry__.py build` "
            "(hipcc --offload-arch=gfx950). There is no CPU fallback.")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FvsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fvs error {code}: {msg}")
        self.code = code


class TimeStepTooLarge(FvsError):
    """Mirror of fub::TimeStepTooLarge (include/fub/TimeStepError.hpp:71-74)."""


class NewtonNotConverged(FvsError):
    """The exception NewtonIteration throws at its iteration cap
    (include/fub/NewtonIteration.hpp:44-50)."""


def check(code: int) -> int:
    if code == FVS_OK:
        return code
    msg = load().fvs_last_error().decode()
    if code == FVS_TIME_STEP_TOO_LARGE:
        raise TimeStepTooLarge(code, msg)
    if code == FVS_NEWTON_NOT_CONVERGED:
        raise NewtonNotConverged(code, msg)
    raise FvsError(code, msg)


def ptr_array(ptrs) -> C.Array:
    arr = (C.c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


def i64x3(vals) -> C.Array:
    v = list(vals) + [0] * (3 - len(vals))
    return (C.c_int64 * 3)(*v[:3])
