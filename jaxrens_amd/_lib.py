"""ctypes binding of ``libjrhip.so`` (the C ABI declared in include/jaxrens_amd.h).

The library is the product path: importing this module raises if the compiled
HIP library is missing, and no call here falls back to host arithmetic.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libjrhip.so"

JR_MOVE_GMC = 0
JR_MOVE_VOLUME = 1
JR_MOVE_HMC = 2
JR_MOVE_MORPH = 3
JR_MOVE_SHEAR = 4
JR_MOVE_STRETCH = 5
JR_MOVE_SWAP = 6
JR_MAX_IMAGES = 4
JR_MAX_MOVES = 8
JR_MAX_SPECIES = 8
JR_NCOUNT = 7
JR_PREC_FP64 = 0
JR_PREC_FP32_PAIRS = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class JrPot(C.Structure):
    _fields_ = [
        ("n_species", C.c_int32),
        ("precision", C.c_int32),
        ("rc2", C.c_double),
        ("eps4", _dp),
        ("eps24", _dp),
        ("sig2", _dp),
        ("images", C.c_int32 * 3),
        ("eval_chunks", C.c_int32),
    ]


class JrMove(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("n_reflect", C.c_int32),
        ("n_leapfrog", C.c_int32),
        ("flat_v_prior", C.c_int32),
        ("n_species", C.c_int32),
        ("species_mask", C.c_int32),
        ("perturb_angle", C.c_double),
        ("max_vpa", C.c_double),
        ("min_vpa", C.c_double),
        ("min_aspect", C.c_double),
        ("refresh_direction", C.c_int32),
        ("dir_slot", C.c_int32),
    ]


_fp = C.POINTER(C.c_float)

JR_LEARNED_PLAIN = 0
JR_LEARNED_AUTO = 1
JR_LEARNED_SCREENED = 2


class JrMPNN(C.Structure):
    _fields_ = [
        ("n_species", C.c_int32),
        ("channels", C.c_int32),
        ("n_basis", C.c_int32),
        ("pad_", C.c_int32),
        ("rc", C.c_double),
        ("core_eps", C.c_double),
        ("core_sig", C.c_double),
        ("core_rc", C.c_double),
        ("emb", _fp), ("W1", _fp), ("U1", _fp), ("c1", _fp),
        ("W2", _fp), ("U2", _fp), ("c2", _fp), ("wo", _fp),
    ]


class JrLearned(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("ring_capacity", C.c_int32),
        ("train_batch", C.c_int32),
        ("train_steps", C.c_int32),
        ("min_train", C.c_int32),
        ("min_heldout", C.c_int32),
        ("lr", C.c_double),
        ("rmse_bound", C.c_double),
        ("delta_k", C.c_double),
        ("train_max_epa", C.c_double),
        ("base_eps", C.c_double),
        ("base_sig", C.c_double),
        ("base_rc", C.c_double),
        ("sur_params", _fp),
    ]


class JrLearnedStats(C.Structure):
    _fields_ = [
        ("iterations_screened", C.c_int64),
        ("iterations_plain", C.c_int64),
        ("proposals", C.c_int64),
        ("stage1_pass", C.c_int64),
        ("accepted", C.c_int64),
        ("exact_force_evals", C.c_int64),
        ("exact_energy_evals", C.c_int64),
        ("surrogate_evals", C.c_int64),
        ("train_steps", C.c_int64),
        ("stored", C.c_int64),
        ("heldout_rmse", C.c_double),
        ("train_loss", C.c_double),
        ("walk_ms", C.c_double),
        ("train_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class JrConfig(C.Structure):
    _fields_ = [
        ("n_replicas", C.c_int32),
        ("n_walkers", C.c_int32),
        ("n_atoms_max", C.c_int32),
        ("n_moves", C.c_int32),
        ("n_mcmc_steps", C.c_int32),
        ("n_extra", C.c_int32),
        ("n_cull", C.c_int32),
        ("n_swap_cycles", C.c_int32),
        ("re_interval", C.c_int32),
        ("device", C.c_int32),
        ("waves_per_block", C.c_int32),
        ("pad_", C.c_int32),
        ("seed", C.c_uint64),
        ("pot", JrPot),
        ("moves", C.POINTER(JrMove)),
        ("move_weights", _dp),
        ("mpnn", C.POINTER(JrMPNN)),
        ("learned", C.POINTER(JrLearned)),
        ("min_distance", _dp),
        ("min_distance_types", _dp),
    ]


class JrTape(C.Structure):
    _fields_ = [
        ("sel_u", _dp),
        ("clone_u", _dp),
        ("extra_u", _dp),
        ("schedule", _ip),
        ("normals", _dp),
        ("uniforms", _dp),
        ("re_u", _dp),
    ]


class JrIterOut(C.Structure):
    _fields_ = [
        ("emax", _dp),
        ("worst", _ip),
        ("clone", _ip),
        ("dead_volume", _dp),
        ("dead_pos", _dp),
        ("dead_types", _ip),
        ("dead_cell", _dp),
        ("counters", _ip),
        ("re_acc", _ip),
        ("re_att", _ip),
        ("walk", _ip),
        ("dead_energy", _dp),
    ]



# (name, restype, argtypes) of every exported symbol; the header declares
# exactly these.
SIGNATURES = [
    ("jr_global_error", C.c_char_p, []),
    ("jr_last_error", C.c_char_p, [C.c_void_p]),
    ("jr_eval_energy_forces", C.c_int,
     [C.POINTER(JrPot), C.c_int32, C.c_int32, C.c_int32, _dp, _ip, _ip, _dp,
      _dp, _dp, _dp, _dp, _dp]),
    ("jr_eval_energy_forces_noalloc", C.c_int,
     [C.POINTER(JrPot), C.c_int32, C.c_int32, C.c_int32, _dp, _ip, _ip, _dp,
      _dp, _dp, _dp, _dp, _dp]),
    ("jr_create", C.c_int, [C.POINTER(JrConfig), C.POINTER(C.c_void_p)]),
    ("jr_destroy", None, [C.c_void_p]),
    ("jr_set_population", C.c_int, [C.c_void_p, _dp, _ip, _ip, _dp, _dp]),
    ("jr_get_population", C.c_int,
     [C.c_void_p, _dp, _ip, _ip, _dp, _dp, _dp]),
    ("jr_get_direction_slot", C.c_int, [C.c_void_p, C.c_int32, _dp]),
    ("jr_set_direction_slot", C.c_int, [C.c_void_p, C.c_int32, _dp]),
    ("jr_set_run_params", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("jr_get_step_sizes", C.c_int, [C.c_void_p, _dp]),
    ("jr_set_pair_valid", C.c_int, [C.c_void_p, _ip]),
    ("jr_set_iteration", C.c_int, [C.c_void_p, C.c_int64]),
    ("jr_get_rng_state", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("jr_set_rng_state", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("jr_get_emax", C.c_int, [C.c_void_p, _dp]),
    ("jr_set_emax", C.c_int, [C.c_void_p, _dp]),
    ("jr_ns_iterations", C.c_int,
     [C.c_void_p, C.c_int32, C.POINTER(JrTape), C.POINTER(JrIterOut)]),
    ("jr_ns_iterations_graph", C.c_int,
     [C.c_void_p, C.c_int32, C.POINTER(JrIterOut)]),
    ("jr_get_last_tape", C.c_int,
     [C.c_void_p, _dp, _dp, _dp, _ip, _dp, _dp, _dp]),
    ("jr_replica_exchange", C.c_int, [C.c_void_p, _dp, _ip, _ip]),
    ("jr_trial_moves", C.c_int,
     [C.c_void_p, C.c_int32, _dp, _dp, C.c_int32, _ip, _dp, _dp, _ip, _ip]),
    ("jr_initial_walk", C.c_int,
     [C.c_void_p, C.c_int32, _dp, _ip, _dp, _dp, C.c_int32, _ip]),
    ("jr_autotune", C.c_int,
     [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
      C.POINTER(C.c_double)]),
    ("jr_set_launch", C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    ("jr_set_walk_block_cap", C.c_int, [C.c_void_p, C.c_int32]),
    ("jr_walk_timing", C.c_int,
     [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int32]),
    ("jr_get_info", C.c_int,
     [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
      C.POINTER(C.c_int32)]),
    ("jr_mpnn_eval", C.c_int,
     [C.POINTER(JrMPNN), C.c_int32, C.c_int32, C.c_int32, _dp, _ip, _ip, _dp,
      _dp, _dp, _dp, _dp]),
    ("jr_sur_eval", C.c_int,
     [_fp, _dp, C.c_int32, C.c_int32, C.c_int32, _dp, _ip, _ip, _dp, C.c_int32, _dp, _dp,
      C.POINTER(C.c_uint16)]),
    ("jr_sur_grad", C.c_int,
     [_fp, C.c_int32, C.c_int32, C.POINTER(C.c_uint16), _fp, _fp, _fp]),
    ("jr_sur_adam", C.c_int,
     [C.c_int32, _fp, _fp, _fp, _fp, C.c_double, C.c_int32]),
    ("jr_learned_get_stats", C.c_int,
     [C.c_void_p, C.POINTER(JrLearnedStats), C.c_int32]),
    ("jr_learned_get_params", C.c_int, [C.c_void_p, _fp]),
    ("jr_learned_set_params", C.c_int, [C.c_void_p, _fp, C.c_int32]),
    ("jr_learned_set_mode", C.c_int, [C.c_void_p, C.c_int32]),
]

_LIB = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libjrhip.so and declare every symbol; raises if it is missing."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"{p} not found: the HIP library is required (no host fallback). "
            "Build it with `python -m jaxrens_amd.build`."
        )
    lib = C.CDLL(str(p))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


class JrError(RuntimeError):
    pass


def check(rc: int, handle=None) -> None:
    if rc != 0:
        lib = load_library()
        msg = (lib.jr_last_error(handle) if handle is not None
               else lib.jr_global_error())
        raise JrError(f"libjrhip error {rc}: {msg.decode() if msg else ''}")


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))
