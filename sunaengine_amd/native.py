"""ctypes binding of the C ABI in include/sunaengine.h.

The shared library is the product: there is no CPU fallback.  If
libsunaengine_amd.so is missing or cannot be loaded, every entry point
raises NativeMissing with the reason.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsunaengine_amd.so")

# error codes (include/sunaengine.h)
SN_OK, SN_E_ARG, SN_E_CAPACITY, SN_E_DEVICE, SN_E_UNSUPPORTED, SN_E_STATE = 0, -1, -2, -3, -4, -5
SN_SPAWN_NO_WELD, SN_SPAWN_OWN_GROUP = 1, 2
#: largest particle_count sn_digest_state_noalloc accepts (its fixed device array)
SN_DIGEST_NOALLOC_MAX_PARTICLES = 8388608

#: every symbol the header declares; tests check the library exports them all
EXPORTS = [
    "sn_create", "sn_destroy", "sn_last_error", "sn_load_scene", "sn_load_scenes", "sn_step",
    "sn_sync", "sn_read_hash", "sn_read_hash_stream", "sn_reset_stream", "sn_read_chain",
    "sn_read_state", "sn_read_derived", "sn_read_bond", "sn_autotune", "sn_write_state_derived", "sn_debug_counters", "sn_substep_count",
    "sn_apply_impulse", "sn_spawn_fluid", "sn_set_events", "sn_set_active", "sn_profile_step",
    "sn_reserve_slots", "sn_load_slot", "sn_query", "sn3_query",
    "sn3_create", "sn3_destroy", "sn3_last_error", "sn3_load", "sn3_set_inputs", "sn3_step",
    "sn3_sync", "sn3_read_hash", "sn3_read_stream", "sn3_reset_stream", "sn3_read_chain",
    "sn3_read_state", "sn3_debug_counters", "sn3_fnv_digest", "sn3_set_events", "sn3_set_active",
    "sn3_profile_step", "sn3_reserve_slots", "sn3_load_slot",
    "sn_digest_last_error", "sn_digest_state", "sn_digest_state_noalloc", "sn_read_push_hits",
]


class NativeMissing(RuntimeError):
    pass


class SnError(RuntimeError):
    """A failed sn_* call; `code` is the SN_E_* value, the message is sn_last_error()."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class sn_tune_trial(C.Structure):
    _fields_ = [("tile_cells", C.c_uint32), ("graph_substeps", C.c_uint32), ("ms_per_substep", C.c_float)]


class sn_opts(C.Structure):
    _fields_ = [("max_particles", C.c_uint32), ("max_cells", C.c_uint32),
                ("max_scenes", C.c_uint32), ("iters", C.c_uint32),
                ("stream_cap", C.c_uint32), ("device", C.c_uint32), ("no_debug", C.c_uint32),
                ("materials", C.c_uint32)]


class sn_scene(C.Structure):
    _fields_ = [("n", C.c_uint32), ("n_fluid", C.c_uint32), ("state8", C.c_void_p),
                ("dom_w", C.c_int32), ("dom_h", C.c_int32), ("cell_shift", C.c_uint32),
                ("st_acc", C.c_void_p)]


class sn3_scene(C.Structure):
    _fields_ = [("n", C.c_uint32), ("parts16", C.c_void_p), ("gx", C.c_uint32),
                ("gy", C.c_uint32), ("gz", C.c_uint32)]


class sn3_event(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("substep", C.c_uint32), ("cx", C.c_int32),
                ("cy", C.c_int32), ("cz", C.c_int32), ("r", C.c_int32), ("ix", C.c_int32),
                ("iy", C.c_int32), ("iz", C.c_int32), ("limit", C.c_uint32)]


class sn_event(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("substep", C.c_uint32), ("cx", C.c_int32),
                ("cy", C.c_int32), ("r", C.c_int32), ("ix", C.c_int32), ("iy", C.c_int32),
                ("limit", C.c_uint32)]


_LIB = None


def lib():
    """Load the library once; raise NativeMissing if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise NativeMissing(f"{LIB_PATH} is not built; run `python -m sunaengine_amd.build`")
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:
        raise NativeMissing(f"cannot load {LIB_PATH}: {e}") from e
    P, u32, i32, sz = C.c_void_p, C.c_uint32, C.c_int32, C.c_size_t
    sig = {
        "sn_create": (C.c_int, [C.POINTER(sn_opts), P, P, C.POINTER(P)]),
        "sn_destroy": (None, [P]),
        "sn_last_error": (C.c_char_p, [P]),
        "sn_load_scene": (C.c_int, [P, C.POINTER(sn_scene)]),
        "sn_load_scenes": (C.c_int, [P, P, u32]),
        "sn_step": (C.c_int, [P, u32]),
        "sn_sync": (C.c_int, [P]),
        "sn_read_hash": (C.c_int, [P, u32, P]),
        "sn_read_hash_stream": (C.c_int, [P, u32, P, sz, C.POINTER(sz)]),
        "sn_reset_stream": (C.c_int, [P]),
        "sn_read_chain": (C.c_int, [P, u32, P]),
        "sn_read_state": (C.c_int, [P, u32, P, sz]),
        "sn_read_derived": (C.c_int, [P, u32, P, sz]),
        "sn_read_bond": (C.c_int, [P, u32, P, sz]),
        "sn_autotune": (C.c_int, [P, u32, u32, P, u32, P, P]),
        "sn_write_state_derived": (C.c_int, [P, u32, P, sz, P, sz]),
        "sn_debug_counters": (C.c_int, [P, u32, P]),
        "sn_substep_count": (C.c_int, [P, u32, C.POINTER(u32)]),
        "sn_apply_impulse": (C.c_int, [P, u32, i32, i32, i32, i32, i32, u32, C.POINTER(u32)]),
        "sn_spawn_fluid": (C.c_int, [P, P, u32, u32, C.POINTER(u32)]),
        "sn_set_events": (C.c_int, [P, P, u32]),
        "sn_read_push_hits": (C.c_int, [P, u32, C.POINTER(u32)]),
        "sn_set_active": (C.c_int, [P, P, u32]),
        "sn_reserve_slots": (C.c_int, [P, u32, P, P]),
        "sn_query": (C.c_int, [P]),
        "sn3_query": (C.c_int, [P]),
        "sn_load_slot": (C.c_int, [P, u32, P]),
        "sn_profile_step": (C.c_int, [P, u32, u32, P, P]),
        "sn3_create": (C.c_int, [C.POINTER(sn_opts), P, u32, C.POINTER(P)]),
        "sn3_destroy": (None, [P]),
        "sn3_last_error": (C.c_char_p, [P]),
        "sn3_load": (C.c_int, [P, P, u32]),
        "sn3_set_inputs": (C.c_int, [P, i32, i32, u32, P, i32]),
        "sn3_step": (C.c_int, [P, u32]),
        "sn3_sync": (C.c_int, [P]),
        "sn3_read_hash": (C.c_int, [P, u32, P]),
        "sn3_read_stream": (C.c_int, [P, u32, P, sz, C.POINTER(sz)]),
        "sn3_reset_stream": (C.c_int, [P]),
        "sn3_read_chain": (C.c_int, [P, u32, P]),
        "sn3_read_state": (C.c_int, [P, u32, P, sz]),
        "sn3_debug_counters": (C.c_int, [P, u32, P]),
        "sn3_fnv_digest": (C.c_int, [P, u32, P]),
        "sn3_set_events": (C.c_int, [P, P, u32]),
        "sn3_set_active": (C.c_int, [P, P, u32]),
        "sn3_reserve_slots": (C.c_int, [P, u32, P, P]),
        "sn3_load_slot": (C.c_int, [P, u32, P]),
        "sn3_profile_step": (C.c_int, [P, u32, P]),
        "sn_digest_last_error": (C.c_char_p, []),
        "sn_digest_state": (C.c_int, [P, sz, u32, u32, u32, u32, P, u32, P, P]),
        "sn_digest_state_noalloc": (C.c_int, [P, sz, u32, u32, u32, u32, P, u32, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def digest_state(state_ptr: int, state_words: int, particle_count: int, stride_words: int,
                 hashed_words: int, n_fluid=None, bond_ptr=None, bond_cap: int = 0, stream=None,
                 entry: str = "sn_digest_state"):
    """digestState (src/hash/statehash.js:80-159) of device-resident words.

    `state_ptr` / `bond_ptr` are device addresses; returns the 8 accumulator
    words as a numpy uint32 array.  `entry` names the C entry point to call
    (each takes the same arguments)."""
    import numpy as np
    out = np.zeros(8, np.uint32)
    nf = particle_count if n_fluid is None else n_fluid
    rc = getattr(lib(), entry)(state_ptr, state_words, particle_count, stride_words, hashed_words, nf,
                               bond_ptr, bond_cap, out.ctypes.data, stream)
    if rc != SN_OK:
        raise SnError(rc, (lib().sn_digest_last_error() or b"").decode() or f"sn error {rc}")
    return out


def check(handle, rc: int, three_d: bool = False) -> None:
    if rc != SN_OK:
        msg = (lib().sn3_last_error if three_d else lib().sn_last_error)(handle)
        raise SnError(rc, (msg or b"").decode() or f"sn error {rc}")
