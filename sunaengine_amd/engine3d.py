"""The 3D engine: the reference 3D host's App.runSubsteps / App.chain over the
C ABI (3d/app.js:606-743), for one or many 3D scenes resident on one GPU."""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import native
from . import scenes3d as s3
from . import tables3d


class Engine3D:
    def __init__(self, max_particles: int, max_cells: int, iters: int = 3, max_scenes: int = 1,
                 stream_cap: int = 4096, device: int = 0, ball_r: float = tables3d.BALL_R_DEFAULT,
                 debug: bool = True):
        if not (isinstance(iters, int) and 1 <= iters <= 16):
            raise s3.Scene3Error("iters must be an integer in 1..16")
        self._luts = np.ascontiguousarray(tables3d.luts3d(ball_r), np.int32)
        L = native.lib()
        o = native.sn_opts(max_particles, max_cells, max_scenes, iters, stream_cap, device, 0 if debug else 1)
        h = C.c_void_p()
        rc = L.sn3_create(C.byref(o), self._luts.ctypes.data, self._luts.size, C.byref(h))
        self._h = h
        if rc != native.SN_OK:
            msg = L.sn3_last_error(h).decode() if h else f"sn3_create failed ({rc})"
            if h:
                L.sn3_destroy(h)
            self._h = None
            raise native.SnError(rc, msg)
        self.scenes = []
        self.iters = iters

    def destroy(self):
        if getattr(self, "_h", None):
            native.lib().sn3_destroy(self._h)
            self._h = None

    __del__ = destroy

    def _check(self, rc):
        native.check(self._h, rc, three_d=True)

    def load(self, scenes: Sequence[s3.Scene3]):
        arr = (native.sn3_scene * len(scenes))()
        keep = []
        for k, sc in enumerate(scenes):
            s3.validate3(sc)
            a = np.ascontiguousarray(sc.parts, np.int32)
            keep.append(a)
            arr[k] = native.sn3_scene(sc.n, a.ctypes.data, sc.gx, sc.gy, sc.gz)
        self._check(native.lib().sn3_load(self._h, arr, len(scenes)))
        self.scenes = list(scenes)
        return self

    def reserve_slots(self, cap_particles: Sequence[int], cap_cells: Sequence[int]):
        """Replace the world with empty, inactive partitions (continuous batching)."""
        cp = np.ascontiguousarray(cap_particles, np.uint32)
        cc = np.ascontiguousarray(cap_cells, np.uint32)
        assert cp.size == cc.size
        self._check(native.lib().sn3_reserve_slots(self._h, cp.size, cp.ctypes.data, cc.ctypes.data))
        self.scenes = [None] * cp.size
        return self

    def load_slot(self, slot: int, sc: s3.Scene3):
        """Place `sc` into free partition `slot` and activate it."""
        s3.validate3(sc)
        a = np.ascontiguousarray(sc.parts, np.int32)
        one = native.sn3_scene(sc.n, a.ctypes.data, sc.gx, sc.gy, sc.gz)
        self._check(native.lib().sn3_load_slot(self._h, int(slot), C.byref(one)))
        self.scenes[slot] = sc
        return self

    def set_inputs(self, stir_amp=0, stir_phase=0, ball_on=0, ball=(0, 0, 0), ball_reach=0):
        b = (C.c_int32 * 3)(*ball)
        self._check(native.lib().sn3_set_inputs(self._h, stir_amp, stir_phase, ball_on,
                                                C.cast(b, C.c_void_p), ball_reach))

    def step(self, count: int = 1):
        self._check(native.lib().sn3_step(self._h, int(count)))

    def busy(self) -> bool:
        """True while enqueued work is still running (does not wait)."""
        rc = native.lib().sn3_query(self._h)
        if rc == 1:
            return True
        self._check(rc)
        return False

    def sync(self):
        self._check(native.lib().sn3_sync(self._h))

    def digest(self, scene: int = 0) -> np.ndarray:
        a = np.zeros(8, np.uint32)
        self._check(native.lib().sn3_read_hash(self._h, scene, a.ctypes.data))
        return a

    def chain(self, scene: int = 0) -> np.ndarray:
        a = np.zeros(8, np.uint32)
        self._check(native.lib().sn3_read_chain(self._h, scene, a.ctypes.data))
        return a

    def hash_stream(self, scene: int = 0, max_substeps: int = 1 << 16) -> np.ndarray:
        out = np.zeros((max_substeps, 8), np.uint32)
        n = C.c_size_t()
        self._check(native.lib().sn3_read_stream(self._h, scene, out.ctypes.data, max_substeps,
                                                 C.byref(n)))
        return out[: n.value]

    def reset_stream(self):
        self._check(native.lib().sn3_reset_stream(self._h))

    def read_state(self, scene: int = 0) -> np.ndarray:
        n = self.scenes[scene].n
        out = np.zeros((max(n, 1), 16), np.int32)
        self._check(native.lib().sn3_read_state(self._h, scene, out.ctypes.data, out.size))
        return out[:n]

    def debug_counters(self, scene: int = 0) -> np.ndarray:
        a = np.zeros(16, np.uint32)
        self._check(native.lib().sn3_debug_counters(self._h, scene, a.ctypes.data))
        return a

    def fnv_digest(self, scene: int = 0) -> np.ndarray:
        a = np.zeros(8, np.uint32)
        self._check(native.lib().sn3_fnv_digest(self._h, scene, a.ctypes.data))
        return a

    def fnv_hex(self, scene: int = 0) -> str:
        return "".join("%08x" % int(x) for x in self.fnv_digest(scene))

    def set_events(self, events):
        """(scene, substep, cx, cy, cz, r, ix, iy, iz[, limit]) tuples."""
        evs = sorted(((e[0], e[1], k, e) for k, e in enumerate(events)))
        arr = (native.sn3_event * max(len(evs), 1))()
        for j, (_, _, _, e) in enumerate(evs):
            arr[j] = native.sn3_event(*e[:9], e[9] if len(e) > 9 else 3000)
        self._check(native.lib().sn3_set_events(self._h, arr, len(evs)))

    def set_active(self, mask):
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8))
        self._check(native.lib().sn3_set_active(self._h, m.ctypes.data, m.size))

    def profile_step(self, count: int = 1) -> dict:
        out = np.zeros(5, np.float32)
        self._check(native.lib().sn3_profile_step(self._h, int(count), out.ctypes.data))
        return dict(zip(("total", "predict_sort", "nbr_density", "solve", "hash_finish"),
                        (float(x) for x in out)))
