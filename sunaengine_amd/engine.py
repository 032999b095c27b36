"""The engine object: the reference's createEngine()/Engine API over the C ABI.

Mirrors ``createEngine`` / ``Engine`` of ``src/engine.js`` (``:2786-3518``,
``:3704-4480``, ``:5901-6098``): same method names in snake_case, same
argument meaning, same refusal rules (``SceneError`` for what loadScene
asserts, ``SnError`` for engine-side failures).  All compute runs in the HIP
library; this module only validates, packs arrays and reads results.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import materials as mt
from . import native
from . import scenes as sc
from . import tables

ITERS = 4
MAXNBR = 48
PARTICLE_WORDS = 8
DERIVED_WORDS = 16
CHAIN_INIT = [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344,
              0xA4093822, 0x299F31D0, 0x082EFA98, 0xEC4E6C89]


def hex8(words) -> str:
    return "".join("%08x" % (int(w) & 0xFFFFFFFF) for w in words)


def digest_dict(acc) -> dict:
    acc = np.asarray(acc, dtype=np.uint32)
    return {"acc": acc, "hex": hex8(acc), "setHex": hex8(acc[:4]), "slotHex": hex8(acc[4:])}


class Engine:
    """One world of one or more independent 2D scenes resident on one GPU."""

    def __init__(self, max_particles: int = 20000, max_cells: int = 65536, iters: int = ITERS,
                 max_scenes: int = 1, stream_cap: int = 4096, device: int = 0, mats=None, luts=None,
                 debug: bool = True, materials: bool = False):
        if not (isinstance(iters, int) and 1 <= iters <= 16):
            raise sc.SceneError(f"iters = {iters} must be an integer in 1..16 (shipped default {ITERS})")
        self.mats_rows = mt.MATS if mats is None else mats
        self._mats_img = mt.mats_image(self.mats_rows)
        self._luts = tables.load_luts() if luts is None else np.ascontiguousarray(luts, np.int32)
        self.max_particles, self.max_cells, self.iters = max_particles, max_cells, iters
        self.max_scenes = max_scenes
        L = native.lib()
        # debug=False is the reference's stripped build (src/engine.js:2884-2894):
        # debug_counters() reads zeros; hashes and state are unchanged
        # materials=True runs the friction, containment and bond passes from the
        # start; otherwise they switch on when a loaded scene needs them
        o = native.sn_opts(max_particles, max_cells, max_scenes, iters, stream_cap, device, 0 if debug else 1,
                           1 if materials else 0)
        h = C.c_void_p()
        rc = L.sn_create(C.byref(o), self._luts.ctypes.data, self._mats_img.ctypes.data, C.byref(h))
        self._h = h
        if rc != native.SN_OK:
            msg = L.sn_last_error(h).decode() if h else f"sn_create failed ({rc})"
            if h:
                L.sn_destroy(h)
            self._h = None
            raise native.SnError(rc, msg)
        self.scenes: List[sc.Scene] = []
        self.images: List[sc.LoadImage] = []

    # ------------------------------------------------------------ lifecycle
    def destroy(self):
        if getattr(self, "_h", None):
            native.lib().sn_destroy(self._h)
            self._h = None

    __del__ = destroy

    def _check(self, rc):
        native.check(self._h, rc)

    # ----------------------------------------------------------------- load
    def load_scene(self, scene: sc.Scene, st_acc: Optional[np.ndarray] = None):
        """loadScene (src/engine.js:3704): validate, build the state words, upload."""
        return self.load_scenes([scene], None if st_acc is None else [st_acc])

    def load_scenes(self, scenes: Sequence[sc.Scene], st_accs=None):
        imgs = [sc.load_image(s, self.mats_rows, max_particles=self.max_particles,
                              max_cells=self.max_cells) for s in scenes]
        return self._upload(list(scenes), imgs, st_accs)

    def _upload(self, scenes, imgs, st_accs=None):
        arr = (native.sn_scene * len(imgs))()
        keep = []
        for k, im in enumerate(imgs):
            st = np.ascontiguousarray(im.state, np.int32)
            keep.append(st)
            acc_ptr = None
            if st_accs is not None and st_accs[k] is not None:
                a = np.ascontiguousarray(st_accs[k], np.int32).reshape(-1)
                keep.append(a)
                acc_ptr = a.ctypes.data
            arr[k] = native.sn_scene(im.n, im.nFluid, st.ctypes.data, im.domW, im.domH,
                                     im.cellShift, acc_ptr)
        self._check(native.lib().sn_load_scenes(self._h, arr, len(imgs)))
        self.scenes, self.images = scenes, imgs
        return self

    # ------------------------------------------------------- slotted worlds
    def reserve_slots(self, cap_particles: Sequence[int], cap_cells: Sequence[int]):
        """Replace the world with empty, inactive partitions (continuous batching)."""
        cp = np.ascontiguousarray(cap_particles, np.uint32)
        cc = np.ascontiguousarray(cap_cells, np.uint32)
        assert cp.size == cc.size
        self._check(native.lib().sn_reserve_slots(self._h, cp.size, cp.ctypes.data, cc.ctypes.data))
        self.scenes = [None] * cp.size
        self.images = [None] * cp.size
        return self

    def load_slot(self, slot: int, scene: sc.Scene):
        """Place `scene` into free partition `slot` and activate it; the other
        partitions keep running state."""
        im = sc.load_image(scene, self.mats_rows, max_particles=self.max_particles, max_cells=self.max_cells)
        st = np.ascontiguousarray(im.state, np.int32)
        one = native.sn_scene(im.n, im.nFluid, st.ctypes.data, im.domW, im.domH, im.cellShift, None)
        self._check(native.lib().sn_load_slot(self._h, int(slot), C.byref(one)))
        self.scenes[slot], self.images[slot] = scene, im
        return self

    @property
    def n(self) -> int:
        return self.images[0].n

    @property
    def n_fluid(self) -> int:
        return self.images[0].nFluid

    # ----------------------------------------------------------------- step
    def step(self, count: int = 1) -> int:
        """step(count) (src/engine.js:3963): exactly `count` substeps, enqueued."""
        if not isinstance(count, (int, np.integer)) or count < 0:
            raise sc.SceneError("step(count) needs a non-negative integer")
        if not self.images:
            raise sc.SceneError("step() called before loadScene()")
        self._check(native.lib().sn_step(self._h, int(count)))
        return self.substep_count()

    def busy(self) -> bool:
        """True while enqueued work is still running (does not wait)."""
        rc = native.lib().sn_query(self._h)
        if rc == 1:
            return True
        self._check(rc)
        return False

    def sync(self):
        self._check(native.lib().sn_sync(self._h))

    def check_errors(self):
        self.sync()

    def substep_count(self, scene: int = 0) -> int:
        v = C.c_uint32()
        self._check(native.lib().sn_substep_count(self._h, scene, C.byref(v)))
        return int(v.value)

    # ------------------------------------------------------------- readback
    def digest(self, scene: int = 0) -> dict:
        """digest() (src/engine.js:4453): H_set ‖ H_slot of the current state."""
        a = np.zeros(8, np.uint32)
        self._check(native.lib().sn_read_hash(self._h, scene, a.ctypes.data))
        return digest_dict(a)

    def chain(self, scene: int = 0) -> np.ndarray:
        a = np.zeros(8, np.uint32)
        self._check(native.lib().sn_read_chain(self._h, scene, a.ctypes.data))
        return a

    def chain_hex(self, scene: int = 0) -> str:
        return hex8(self.chain(scene))

    def hash_stream(self, scene: int = 0) -> np.ndarray:
        """Per-substep 8-word digests since the last reset_stream(), uint32 [k, 8]."""
        cap = self.substep_count(scene) + 1
        out = np.zeros((max(cap, 1), 8), np.uint32)
        n = C.c_size_t()
        self._check(native.lib().sn_read_hash_stream(self._h, scene, out.ctypes.data, cap,
                                                     C.byref(n)))
        return out[: n.value]

    def reset_stream(self):
        self._check(native.lib().sn_reset_stream(self._h))

    def read_state(self, scene: int = 0) -> np.ndarray:
        n = self.images[scene].n if scene < len(self.images) else 0
        out = np.zeros((max(n, 1), 8), np.int32)
        self._check(native.lib().sn_read_state(self._h, scene, out.ctypes.data, out.size))
        return out[:n]

    def read_derived(self, scene: int = 0) -> np.ndarray:
        n = self.images[scene].n if scene < len(self.images) else 0
        out = np.zeros((max(n, 1), 16), np.int32)
        self._check(native.lib().sn_read_derived(self._h, scene, out.ctypes.data, out.size))
        return out[:n]

    def autotune(self, trial_substeps: int = 16, lookahead: int = 0) -> dict:
        """Tune the world's launch shape (neighbour tile width, hipGraph
        length) on the world itself, timing every trial from the world
        advanced `lookahead` substeps; state and streams are left as they
        were.  Returns {"trials": [...], "chosen": {...}}."""
        trials = (native.sn_tune_trial * 16)()
        n = C.c_uint32()
        best = native.sn_tune_trial()
        self._check(native.lib().sn_autotune(self._h, int(trial_substeps), int(lookahead), trials, 16,
                                             C.byref(n), C.byref(best)))
        as_dict = lambda t: {"tile_cells": int(t.tile_cells), "graph_substeps": int(t.graph_substeps),
                             "ms_per_substep": round(float(t.ms_per_substep), 4)}
        return {"trials": [as_dict(trials[i]) for i in range(min(n.value, 16))], "chosen": as_dict(best)}

    def write_state_derived(self, state, derived=None, scene: int = 0):
        """The buffer writes of restore(snap) (src/demo_boot.js:308-321): state
        (n, 8) and optionally derived (n, 16) words, slot order."""
        st = np.ascontiguousarray(state, np.int32).reshape(-1)
        dv = None if derived is None else np.ascontiguousarray(derived, np.int32).reshape(-1)
        self._check(native.lib().sn_write_state_derived(self._h, scene, st.ctypes.data, st.size,
                                                        None if dv is None else dv.ctypes.data,
                                                        0 if dv is None else dv.size))

    def read_bond(self, scene: int = 0) -> np.ndarray:
        """readBond (src/engine.js:4295): (n, 48) packed bond words in slot order."""
        n = self.images[scene].n if scene < len(self.images) else 0
        out = np.zeros((max(n, 1), 48), np.uint32)
        self._check(native.lib().sn_read_bond(self._h, scene, out.ctypes.data, out.size))
        return out[:n]

    def debug_counters(self, scene: int = 0) -> np.ndarray:
        a = np.zeros(16, np.uint32)
        self._check(native.lib().sn_debug_counters(self._h, scene, a.ctypes.data))
        return a

    def assert_spec_valid_regime(self, where: str = "run", scene: int = 0) -> bool:
        """assertSpecValidRegime (src/engine.js:4427-4450)."""
        names = ["mul/shift saturation", "divide saturation or bad shift", "divide by zero",
                 "sqrt of negative", "rsqrt of non-positive", "add/sub/neg saturation",
                 "neighbour-list truncation", "cell/scan capacity overflow", "bond row saturation"]
        dbg = self.debug_counters(scene)
        hit = [f"{names[i] if i < len(names) else f'RESERVED dbg[{i}]'} (dbg[{i}]={int(v)})"
               for i, v in enumerate(dbg) if v]
        if hit:
            raise sc.SceneError(f"[aether] {where}: simulation left the spec-valid regime — "
                                + "; ".join(hit))
        return True

    # --------------------------------------------------------------- inputs
    def apply_impulse(self, cx, cy, radius, ix, iy, limit: int = 3000, scene: int = 0) -> int:
        """applyImpulse (src/engine.js:6077) on the exact current state."""
        hits = C.c_uint32()
        self._check(native.lib().sn_apply_impulse(self._h, scene, int(cx), int(cy), int(radius),
                                                  int(ix), int(iy), int(limit), C.byref(hits)))
        return int(hits.value)

    def spawn_fluid(self, particles: Iterable[dict], no_weld: bool = False, own_group: bool = False) -> int:
        """spawnFluid(list, {noWeld, ownGroup}) (src/engine.js:5901); returns
        the new n.  The engine shifts its own boundary block and assigns the
        spawn's weld group."""
        plist = list(particles)
        if not plist:
            return self.n
        scene = self.scenes[0]
        fl, _ = sc.spawn_records(scene, plist, self.mats_rows)
        n_out = C.c_uint32()
        fl = np.ascontiguousarray(fl, np.int32)
        flags = (native.SN_SPAWN_NO_WELD if no_weld else 0) | (native.SN_SPAWN_OWN_GROUP if own_group else 0)
        self._check(native.lib().sn_spawn_fluid(self._h, fl.ctypes.data, len(plist), flags, C.byref(n_out)))
        new_n = int(n_out.value)
        if new_n != self.images[0].n:
            im = self.images[0]
            im.n, im.nFluid = new_n, im.nFluid + len(plist)
        return new_n

    def push_hits(self, scene: int = 0) -> int:
        """Particles moved by stir events and apply_impulse since the load."""
        out = C.c_uint32()
        self._check(native.lib().sn_read_push_hits(self._h, scene, C.byref(out)))
        return int(out.value)

    def set_events(self, events: Sequence[tuple]):
        """Device-resident stir schedule: (scene, substep, cx, cy, r, ix, iy[, limit])."""
        evs = sorted(((e[0], e[1], k, e) for k, e in enumerate(events)))
        arr = (native.sn_event * max(len(evs), 1))()
        for j, (_, _, _, e) in enumerate(evs):
            lim = e[7] if len(e) > 7 else 3000
            arr[j] = native.sn_event(e[0], e[1], e[2], e[3], e[4], e[5], e[6], lim)
        self._check(native.lib().sn_set_events(self._h, arr, len(evs)))

    def set_active(self, mask: Sequence[bool]):
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8))
        self._check(native.lib().sn_set_active(self._h, m.ctypes.data, m.size))

    # kernel classes of sn_profile_step, slot 0 = the whole span
    PROFILE_KEYS = ("total", "predict", "scan", "scatter", "canon_gather", "nbr_solveA", "solveA",
                    "colour", "xnc", "st_thermal_hash", "hash_finish", "impulse", "lists", "apply",
                    "friction", "materials")

    def profile_step(self, count: int = 1, classes: Optional[Iterable[str]] = None) -> dict:
        """Run `count` substeps eagerly with an event pair around every launch
        of the kernel classes named in `classes` (all when None).

        Returns {class: {"ms": summed device ms, "calls": launches}}; "total"
        is the span from the first timed launch's start to the last one's end.
        """
        mask = 0xFFFFFFFF
        if classes is not None:
            mask = 0
            for c in classes:
                mask |= 1 << (self.PROFILE_KEYS.index(c) - 1)
        ms = np.zeros(16, np.float32)
        calls = np.zeros(16, np.uint32)
        self._check(native.lib().sn_profile_step(self._h, int(count), mask, ms.ctypes.data, calls.ctypes.data))
        return {k: {"ms": float(ms[i]), "calls": int(calls[i])} for i, k in enumerate(self.PROFILE_KEYS)}


def create_engine(**opts) -> Engine:
    """createEngine(opts) with the reference's option names where they exist."""
    m = {"maxParticles": "max_particles", "maxCells": "max_cells", "maxScenes": "max_scenes"}
    return Engine(**{m.get(k, k): v for k, v in opts.items()})
