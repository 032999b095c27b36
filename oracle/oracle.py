"""ORACLE — test infrastructure only.  ctypes binding of oracle/_build/liboracle.so."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib_path() -> str:
    return os.path.join(HERE, "_build", "liboracle.so")


def build() -> str:
    """Compile the oracle with make (g++ on its own sources)."""
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return lib_path()


def _lib():
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not os.path.exists(p):
            build()
        L = C.CDLL(p)
        P = C.c_void_p
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        L.or_create.restype = P
        L.or_create.argtypes = [C.c_uint32, C.c_uint32, i32p, i32p, C.c_int]
        L.or_destroy.argtypes = [P]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_load.restype = C.c_int
        L.or_load.argtypes = [P, C.c_uint32, C.c_uint32, i32p, C.c_int32, C.c_int32,
                              C.c_uint32, C.c_uint32, C.c_uint32]
        L.or_step.argtypes = [P, C.c_uint32, C.c_void_p]
        L.or_digest.argtypes = [P, u32p]
        L.or_chain.argtypes = [P, u32p]
        L.or_read_state.argtypes = [P, i32p]
        L.or_read_derived.argtypes = [P, i32p]
        L.or_read_bond.argtypes = [P, u32p]
        L.or_write_state_derived.argtypes = [P, i32p, C.c_void_p]
        L.or_debug_counters.argtypes = [P, u32p]
        L.or_substep_index.restype = C.c_uint32
        L.or_substep_index.argtypes = [P]
        L.or_apply_impulse.restype = C.c_uint32
        L.or_apply_impulse.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32, C.c_uint32]
        L.or_spawn.restype = C.c_uint32
        L.or_spawn.argtypes = [P, C.c_uint32, i32p, i32p]
        L.or_friction_probe.argtypes = [P, np.ctypeslib.ndpointer(np.int64, flags="C")]
        L.or_nbr_probe.argtypes = [P, i32p, u32p, u32p, C.POINTER(C.c_uint32)]
        L.or_friction_dp.argtypes = [P, i32p, i32p]
        L.or_contain_probe.argtypes = [P, i32p, i32p]
        _LIB = L
    return _LIB


def set_threads(n: int) -> None:
    _lib().or_set_threads(int(n))


class OracleWorld:
    """One 2D world on the CPU oracle, with the reference engine's host API."""

    def __init__(self, image, luts, mats, iters: int = 4, max_particles=None, max_cells=None):
        L = _lib()
        self.max_particles = int(max_particles or image.n)
        self.max_cells = int(max_cells or image.cellTotal)
        self._h = L.or_create(self.max_particles, self.max_cells,
                              np.ascontiguousarray(luts, np.int32),
                              np.ascontiguousarray(mats, np.int32).reshape(-1), int(iters))
        self.load(image)

    def load(self, image):
        rc = _lib().or_load(self._h, image.n, image.nFluid,
                            np.ascontiguousarray(image.state, np.int32).reshape(-1),
                            image.domW, image.domH, image.cellShift, image.gridW, image.gridH)
        if rc != 0:
            raise ValueError("oracle: scene exceeds capacity")
        self.n, self.nFluid = image.n, image.nFluid

    def close(self):
        if self._h:
            _lib().or_destroy(self._h)
            self._h = None

    __del__ = close

    def step(self, count: int, stream: bool = False):
        """Advance `count` substeps; with stream=True return the per-substep
        8-word digests as uint32 [count, 8]."""
        if stream:
            out = np.zeros((count, 8), np.uint32)
            _lib().or_step(self._h, count, out.ctypes.data)
            return out
        _lib().or_step(self._h, count, None)
        return None

    def digest(self) -> np.ndarray:
        a = np.zeros(8, np.uint32)
        _lib().or_digest(self._h, a)
        return a

    def chain(self) -> np.ndarray:
        a = np.zeros(8, np.uint32)
        _lib().or_chain(self._h, a)
        return a

    def read_state(self) -> np.ndarray:
        a = np.zeros(self.n * 8, np.int32)
        _lib().or_read_state(self._h, a)
        return a.reshape(-1, 8)

    def read_derived(self) -> np.ndarray:
        a = np.zeros(self.n * 16, np.int32)
        _lib().or_read_derived(self._h, a)
        return a.reshape(-1, 16)

    def read_bond(self) -> np.ndarray:
        a = np.zeros(self.n * 48, np.uint32)
        _lib().or_read_bond(self._h, a)
        return a.reshape(-1, 48)

    def write_state_derived(self, state, derived=None):
        st = np.ascontiguousarray(state, np.int32).reshape(-1)
        dv = None if derived is None else np.ascontiguousarray(derived, np.int32).reshape(-1)
        _lib().or_write_state_derived(self._h, st, None if dv is None else dv.ctypes.data)

    def debug_counters(self) -> np.ndarray:
        a = np.zeros(16, np.uint32)
        _lib().or_debug_counters(self._h, a)
        return a

    @property
    def substep(self) -> int:
        return int(_lib().or_substep_index(self._h))

    def apply_impulse(self, cx, cy, r, ix, iy, limit=3000) -> int:
        return int(_lib().or_apply_impulse(self._h, cx, cy, r, ix, iy, limit))

    def spawn(self, fluid8, boundary8) -> int:
        f = np.ascontiguousarray(fluid8, np.int32).reshape(-1)
        b = np.ascontiguousarray(boundary8, np.int32).reshape(-1)
        if b.size == 0:
            b = np.zeros(1, np.int32)
        n = int(_lib().or_spawn(self._h, f.size // 8, f, b))
        if n != self.n:
            self.nFluid += f.size // 8
            self.n = n
        return n


def _lib3():
    L = _lib()
    if not hasattr(L, "_or3_ready"):
        P = C.c_void_p
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        L.or3_create.restype = P
        L.or3_create.argtypes = [C.c_uint32, C.c_uint32, i32p, C.c_uint32, C.c_int]
        L.or3_destroy.argtypes = [P]
        L.or3_load.restype = C.c_int
        L.or3_load.argtypes = [P, C.c_uint32, i32p, C.c_uint32, C.c_uint32, C.c_uint32]
        L.or3_set_inputs.argtypes = [P, C.c_int32, C.c_int32, C.c_uint32, C.c_void_p, C.c_int32]
        L.or3_step.argtypes = [P, C.c_uint32, C.c_void_p]
        L.or3_digest.argtypes = [P, u32p]
        L.or3_chain.argtypes = [P, u32p]
        L.or3_read.argtypes = [P, i32p]
        L.or3_debug_counters.argtypes = [P, u32p]
        L.or3_fnv_digest.argtypes = [P, u32p]
        L.or3_grid_probe.argtypes = [P, u32p, u32p, u32p, u32p, C.POINTER(C.c_uint32)]
        L.or3_dp_probe.argtypes = [P]
        L.or3_apply_impulse.restype = C.c_uint32
        L.or3_apply_impulse.argtypes = [P] + [C.c_int32] * 7 + [C.c_uint32]
        L._or3_ready = True
    return L


class OracleWorld3:
    """One 3D world on the CPU oracle (3d/sim3d.wgsl restated)."""

    def __init__(self, scene3, luts3, iters: int = 3, max_particles=None, max_cells=None):
        L = _lib3()
        self.n = scene3.n
        cells = scene3.gx * scene3.gy * scene3.gz
        lut = np.ascontiguousarray(luts3, np.int32)
        self._h = L.or3_create(int(max_particles or self.n), int(max_cells or cells), lut, lut.size,
                               int(iters))
        self._cells = cells
        rc = L.or3_load(self._h, self.n, np.ascontiguousarray(scene3.parts, np.int32).reshape(-1),
                        scene3.gx, scene3.gy, scene3.gz)
        if rc:
            raise ValueError("oracle3: scene exceeds capacity")

    def close(self):
        if self._h:
            _lib3().or3_destroy(self._h)
            self._h = None

    __del__ = close

    def set_inputs(self, stir_amp=0, stir_phase=0, ball_on=0, ball=(0, 0, 0), ball_reach=0):
        b = (C.c_int32 * 3)(*ball)
        _lib3().or3_set_inputs(self._h, stir_amp, stir_phase, ball_on, C.cast(b, C.c_void_p), ball_reach)

    def step(self, count: int, stream: bool = False):
        if stream:
            out = np.zeros((count, 8), np.uint32)
            _lib3().or3_step(self._h, count, out.ctypes.data)
            return out
        _lib3().or3_step(self._h, count, None)
        return None

    def digest(self):
        a = np.zeros(8, np.uint32)
        _lib3().or3_digest(self._h, a)
        return a

    def chain(self):
        a = np.zeros(8, np.uint32)
        _lib3().or3_chain(self._h, a)
        return a

    def read(self):
        a = np.zeros(self.n * 16, np.int32)
        _lib3().or3_read(self._h, a)
        return a.reshape(-1, 16)

    def debug_counters(self):
        a = np.zeros(16, np.uint32)
        _lib3().or3_debug_counters(self._h, a)
        return a

    def fnv_digest(self):
        a = np.zeros(8, np.uint32)
        _lib3().or3_fnv_digest(self._h, a)
        return a

    def apply_impulse(self, cx, cy, cz, r, ix, iy, iz, limit=3000) -> int:
        """The 3D stir event on the current state (slots in ascending order
        within r, at most `limit`); returns the hit count."""
        return int(_lib3().or3_apply_impulse(self._h, cx, cy, cz, r, ix, iy, iz, limit))

    def grid_probe(self):
        """Run the next substep through buildNbr only (the world is left
        mid-substep): returns cellStart, sortedIds, nbr [n, 48], nbrN and the
        pairs truncation dropped; read() then holds the predicted positions."""
        cells = self._cells
        cs = np.zeros(cells + 1, np.uint32)
        ids = np.zeros(self.n, np.uint32)
        nbr = np.zeros(self.n * 48, np.uint32)
        nn = np.zeros(self.n, np.uint32)
        dropped = C.c_uint32()
        _lib3().or3_grid_probe(self._h, cs, ids, nbr, nn, C.byref(dropped))
        return {"cellStart": cs, "sortedIds": ids, "nbr": nbr.reshape(-1, 48), "nbrN": nn,
                "dropped": int(dropped.value)}

    def dp_probe(self):
        """Run the next substep through iteration 0's solveDp (left
        mid-substep) and return the records: pr, lambda, rho and dp."""
        _lib3().or3_dp_probe(self._h)
        return self.read()


def nbr_probe(world: "OracleWorld") -> dict:
    """Run the next substep of `world` through buildNbr only and return its
    lists: pred [n, 2], nbr [n, 48] (list order), nbrN [n] (count plus the
    seal / ext-seal / congeal bits) and the pairs dropped by truncation
    (the world is left mid-substep)."""
    n = world.n
    pred = np.zeros(n * 2, np.int32)
    nbr = np.zeros(n * 48, np.uint32)
    nbrn = np.zeros(n, np.uint32)
    dropped = C.c_uint32()
    _lib().or_nbr_probe(world._h, pred, nbr, nbrn, C.byref(dropped))
    return {"pred": pred.reshape(-1, 2), "nbr": nbr.reshape(-1, 48), "nbrN": nbrn, "dropped": int(dropped.value)}


def friction_dp(world: "OracleWorld", n_fluid: int) -> dict:
    """Run the next substep of `world` up to its first friction pass and
    return each fluid slot's friction-pass dp (ticks << ACC, friction plus
    the bond gather, before applyFric) and pred, both (n_fluid, 2).  The
    world is left mid-substep."""
    dp = np.zeros(2 * n_fluid, np.int32)
    pred = np.zeros(2 * n_fluid, np.int32)
    _lib().or_friction_dp(world._h, dp, pred)
    return {"dp": dp.reshape(n_fluid, 2), "pred": pred.reshape(n_fluid, 2)}


def contain_probe(world: "OracleWorld", n_fluid: int) -> dict:
    """Run the next substep of `world` through its solver iterations and
    return each fluid slot's pred just before and just after the containment
    passes, both (n_fluid, 2).  The world is left mid-substep."""
    before = np.zeros(2 * n_fluid, np.int32)
    after = np.zeros(2 * n_fluid, np.int32)
    _lib().or_contain_probe(world._h, before, after)
    return {"before": before.reshape(n_fluid, 2), "after": after.reshape(n_fluid, 2)}


def friction_probe(world: "OracleWorld", n_fluid: int) -> dict:
    """Run the next substep of `world` up to its first friction pass and sum
    the friction dp over the fluid slots (the world is left mid-substep)."""
    out = np.zeros(5, np.int64)
    _lib().or_friction_probe(world._h, out)
    return {"dp": (int(out[0]), int(out[1])), "pairs": int(out[2]), "boundary": int(out[3]),
            "full": int(out[4])}
