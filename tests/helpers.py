"""Shared helpers for the parity tests (oracle runs, GPU engine runs)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def oracle_world(scene, iters, max_particles=None, max_cells=None):
    from oracle import OracleWorld
    from sunaengine_amd import materials, scenes, tables
    img = scenes.load_image(scene)
    return OracleWorld(img, tables.load_luts(), materials.mats_image(), iters=iters,
                       max_particles=max_particles or img.n, max_cells=max_cells or img.cellTotal)


def gpu_engine(scene, iters, max_particles=None, max_cells=None, stream_cap=4096):
    from sunaengine_amd.engine import Engine
    from sunaengine_amd import scenes
    img = scenes.load_image(scene)
    e = Engine(max_particles=max_particles or img.n, max_cells=max_cells or img.cellTotal,
               iters=iters, stream_cap=stream_cap)
    e.load_scene(scene)
    return e


def first_mismatch(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} != {b.shape}"
    bad = np.nonzero((a != b).reshape(a.shape[0], -1).any(axis=1))[0]
    return None if bad.size == 0 else int(bad[0])


class DeviceWords:
    """A device copy of a host array, allocated through the same HIP runtime
    the library links (/opt/rocm's libamdhip64): test plumbing for entry
    points that take device pointers.  torch is not used for this because it
    loads a second, bundled HIP runtime into the process."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes as C
            from sunaengine_amd import native
            native.lib()   # the library's runtime is loaded first; the same soname resolves to it
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, a):
        import ctypes as C
        host = np.ascontiguousarray(a).view(np.uint32).reshape(-1)
        self.words = host.size
        p = C.c_void_p()
        h = self.hip()
        assert h.hipMalloc(C.byref(p), max(4, host.nbytes)) == 0, "hipMalloc failed"
        self.ptr = p.value
        if host.nbytes:
            assert h.hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, 1) == 0, "hipMemcpy failed"
        assert h.hipDeviceSynchronize() == 0

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None

    __del__ = free


def oracle_stream_2d(scene, iters, steps, events):
    """The oracle's per-substep stream of a 2D scene with stir events
    (substep, cx, cy, r, ix, iy[, limit]) applied immediately before their
    substep, in the order given (arming order, src/demo_boot.js:244-272)."""
    o = oracle_world(scene, iters)
    rows = []
    for s in range(steps):
        for ev in events:
            if ev[0] == s:
                o.apply_impulse(*ev[1:6], ev[6] if len(ev) > 6 else 3000)
        rows.append(o.step(1, stream=True)[0])
    return o, np.array(rows).reshape(-1, 8)


def oracle_stream_3d(scene, iters, steps, events):
    """As oracle_stream_2d for a 3D scene: events (substep, cx, cy, cz, r,
    ix, iy, iz[, limit])."""
    from oracle import OracleWorld3
    from sunaengine_amd import tables3d
    o = OracleWorld3(scene, tables3d.luts3d(), iters)
    rows = []
    for s in range(steps):
        for ev in events:
            if ev[0] == s:
                o.apply_impulse(*ev[1:8], ev[8] if len(ev) > 8 else 3000)
        rows.append(o.step(1, stream=True)[0])
    return o, np.array(rows).reshape(-1, 8)
