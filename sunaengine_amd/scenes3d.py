"""3D scenes: the reference's seeded initial column (3d/app.js:232-292) and
the 2M-particle benchmark block.

A 3D record is 16 int32 words {p.xyzw, v.xyzw, pr.xyzw, dp.xyzw}
(3d/sim3d.wgsl:46-51) with the material index in p.w.  The world is
gx*2 x gy*2 x gz*2 world units (cells are one kernel radius, 2 wu).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .scenes import splitmix32, xorshift32

ONE = 65536
H3 = 131072
MAX_CELLS_3D = 1 << 24


class Scene3Error(ValueError):
    pass


@dataclass
class Scene3:
    n: int
    parts: np.ndarray      # int32 [n, 16]
    gx: int
    gy: int
    gz: int
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def cell_total(self) -> int:
        return self.gx * self.gy * self.gz


def _js_round(x: float) -> int:
    import math
    f = math.floor(x)
    return int(f + 1) if (x - f) >= 0.5 else int(f)


def column_scene(n: int, gx: int = 28, gy: int = 30, gz: int = 28, spacing: float = 1.0,
                 frac: float = 0.55, layers: bool = False) -> Scene3:
    """App.initialState(n, 'pool' | 'layers') (3d/app.js:247-292): a column in
    one corner with splitmix32 jitter in [-4096, 4096) per axis."""
    world = [gx * 2.0, gy * 2.0, gz * 2.0]
    bx = int((world[0] * frac) // spacing)
    bz = int((world[2] * frac) // spacing)
    by_max = int((world[1] * 0.92) // spacing)
    by = min(by_max, -(-n // (bx * bz)))
    a = np.zeros((n, 16), np.int64)
    c = 0
    for j in range(by):
        for k in range(bz):
            for i in range(bx):
                if c >= n:
                    break
                jx = (splitmix32(c * 3 + 0) % 8192) - 4096
                jy = (splitmix32(c * 3 + 1) % 8192) - 4096
                jz = (splitmix32(c * 3 + 2) % 8192) - 4096
                a[c, 0] = _js_round((1.5 + i * spacing) * ONE) + jx
                a[c, 1] = _js_round((1.5 + j * spacing) * ONE) + jy
                a[c, 2] = _js_round((1.5 + k * spacing) * ONE) + jz
                if layers:
                    f = j / by
                    a[c, 3] = 1 if f < 0.34 else (0 if f < 0.67 else 2)
                c += 1
    a = a[:c]
    return Scene3(n=c, parts=a.astype(np.int32), gx=gx, gy=gy, gz=gz, name="column")


def block_scene(side: int = 128, gx: int = 128, gy: int = 64, gz: int = 128, seed: int = 0x3D3D3D,
                pitch: int = ONE >> 1, jitter: int = 1024, mat: int = 0, origin_wu: float = 1.5) -> Scene3:
    """side^3 block at `pitch` ticks from (1.5, 1.5, 1.5) wu, jittered by
    xorshift32(seed) integers in [-jitter, jitter) per axis (x fastest)."""
    n = side ** 3
    o = int(origin_wu * ONE)
    rnd = xorshift32(seed)
    span = 2 * jitter
    jit = np.empty((n, 3), np.int64)
    for c in range(n):
        jit[c, 0] = (rnd() % span) - jitter
        jit[c, 1] = (rnd() % span) - jitter
        jit[c, 2] = (rnd() % span) - jitter
    idx = np.arange(n, dtype=np.int64)
    i = idx % side
    k = (idx // side) % side
    j = idx // (side * side)
    a = np.zeros((n, 16), np.int64)
    a[:, 0] = o + i * pitch + jit[:, 0]
    a[:, 1] = o + j * pitch + jit[:, 1]
    a[:, 2] = o + k * pitch + jit[:, 2]
    a[:, 3] = mat
    s = Scene3(n=n, parts=a.astype(np.int32), gx=gx, gy=gy, gz=gz, name=f"block3d_{n}",
               meta=dict(seed=seed, pitch=pitch, jitter=jitter))
    validate3(s)
    return s


def validate3(s: Scene3) -> None:
    if s.gx < 1 or s.gy < 1 or s.gz < 1 or s.cell_total > MAX_CELLS_3D:
        raise Scene3Error("3D grid out of range")
    p = s.parts
    for ax, g in ((0, s.gx), (1, s.gy), (2, s.gz)):
        if s.n and (p[:, ax].min() < 0 or p[:, ax].max() >= g * H3):
            raise Scene3Error("3D position outside the world")
    if s.n and (p[:, 3].min() < 0 or p[:, 3].max() >= 6):
        raise Scene3Error("3D material outside the table")
