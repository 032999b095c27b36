"""Seeded integer scene construction and the load-time state image.

Restates the reference's scene builders (``makeScene``, ``boundaryRing``,
``xorshift32``, ``splitmix32``, ``permuteScene``: ``src/engine.js:2153-2169,
2396-2679``) and the validation and 8-word state construction of
``Engine.loadScene`` (``src/engine.js:3704-3951``).  Nothing here touches a
device; the GPU engine and the CPU oracle both start from ``load_image``.

Also defines the benchmark workloads built on the same primitives (a walled
dam break at 0.5 wu pitch, the 1M-particle thermal scene, the multi-scene
batch), each fully determined by its seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import materials as mt

ONE = 65536
H = 163840
CELL_SHIFT = 18
MAX_CELL_TOTAL = 1048576
MAX_POS = 134217728
VMAX = 65536
PARTICLE_WORDS = 8
DERIVED_WORDS = 16
FLAG_BURNING, FLAG_DEAD, FLAG_ANCHOR = 1, 2, 4
FLAG_SEAL = 128
FLAG_MASK = 7
SPAWN_FLAG_MASK = FLAG_BURNING | FLAG_ANCHOR
WELD_TAG_SHIFT = 8
WELD_TAG_MASK = 0x000FFF00
MAXNBR = 48
BOND_CAP = 48


class SceneError(ValueError):
    """A scene the reference's loadScene would refuse ('[aether] ...' throws)."""


def xorshift32(seed: int):
    """src/engine.js:2153-2161"""
    s = (seed & 0xFFFFFFFF) or 0x2545F491

    def rnd():
        nonlocal s
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        return s

    return rnd


def splitmix32(a: int) -> int:
    """src/engine.js:2163-2169"""
    a = (a + 0x9E3779B9) & 0xFFFFFFFF
    t = a
    t = ((t ^ (t >> 16)) * 0x21F0AAAD) & 0xFFFFFFFF
    t = ((t ^ (t >> 15)) * 0x735A2D97) & 0xFFFFFFFF
    return (t ^ (t >> 15)) & 0xFFFFFFFF


@dataclass
class Scene:
    n: int
    nFluid: int
    pos: np.ndarray          # int32 [n*2]
    vel: np.ndarray          # int32 [n*2]
    matId: np.ndarray        # uint32 [n]
    domW: int
    domH: int
    cellShift: int = CELL_SHIFT
    flags: Optional[np.ndarray] = None
    temp: Optional[np.ndarray] = None
    fuel: Optional[np.ndarray] = None
    name: str = ""
    meta: dict = field(default_factory=dict)


def empty_scene(n, n_fluid, dom_w_wu, dom_h_wu) -> Scene:
    return Scene(n=n, nFluid=n_fluid, pos=np.zeros(2 * n, np.int32), vel=np.zeros(2 * n, np.int32),
                 matId=np.zeros(n, np.uint32), domW=int(dom_w_wu * ONE), domH=int(dom_h_wu * ONE))


def boundary_ring(dom_w_wu: int, dom_h_wu: int, wall_rows: int):
    """Two floor rows and two wall columns per side (src/engine.js:2417-2433)."""
    out = []
    half = ONE >> 1
    for row in range(2):
        y = dom_h_wu * ONE - half - row * ONE
        for c in range(dom_w_wu):
            out.append((half + c * ONE, y))
    for x_base in (half, dom_w_wu * ONE - half):
        for col in range(2):
            x = half + col * ONE if x_base == half else x_base - col * ONE
            for r in range(wall_rows):
                out.append((x, dom_h_wu * ONE - half - (2 + r) * ONE))
    return out


def _assemble(fluid, bnd, dom_w_wu, dom_h_wu, name, fluid_vel=None) -> Scene:
    s = empty_scene(len(fluid) + len(bnd), len(fluid), dom_w_wu, dom_h_wu)
    for i, f in enumerate(fluid):
        s.pos[2 * i] = f[0]
        s.pos[2 * i + 1] = f[1]
        s.matId[i] = f[2]
        if fluid_vel is not None:
            s.vel[2 * i] = fluid_vel[i][0]
            s.vel[2 * i + 1] = fluid_vel[i][1]
    for k, b in enumerate(bnd):
        i = len(fluid) + k
        s.pos[2 * i] = b[0]
        s.pos[2 * i + 1] = b[1]
        s.matId[i] = mt.WATER
    s.name = name
    return s


def make_scene(name: str, **opts) -> Scene:
    """Canonical scenes (src/engine.js:2469-2649).  dambreak_v1 is frozen."""
    DOM_W, DOM_H = 128, 72
    if name == "dambreak_v1":
        rnd = xorshift32(0x1234567)
        fluid = []
        for gy in range(60):
            for gx in range(60):
                jx = (rnd() % 4097) - 2048
                jy = (rnd() % 4097) - 2048
                fluid.append((3 * ONE + gx * ONE + jx, (DOM_H - 3) * ONE - gy * ONE + jy, mt.WATER))
        return _assemble(fluid, boundary_ring(DOM_W, DOM_H, 60), DOM_W, DOM_H, name)
    if name == "overdense_v1":
        side, sp = 32, ONE >> 1
        cx, cy = 64 * ONE, 40 * ONE
        fluid = [(cx + (gx - side // 2) * sp, cy + (gy - side // 2) * sp, mt.WATER)
                 for gy in range(side) for gx in range(side)]
        return _assemble(fluid, boundary_ring(DOM_W, DOM_H, 60), DOM_W, DOM_H, name)
    if name == "pool_v1":
        cw, ch = opts.get("w", 100), opts.get("h", 24)
        mat = opts.get("mat", mt.WATER)
        if not (13 + cw <= DOM_W - 2 and ch <= DOM_H - 5):
            raise SceneError(f"pool_v1 w={cw} h={ch} does not fit the domain")
        rnd = xorshift32(0x0BE7A1)
        fluid = []
        for gy in range(ch):
            for gx in range(cw):
                jx = (rnd() % 4097) - 2048
                jy = (rnd() % 4097) - 2048
                fluid.append((13 * ONE + gx * ONE + jx, (DOM_H - 3) * ONE - gy * ONE + jy, mat))
        return _assemble(fluid, boundary_ring(DOM_W, DOM_H, 60), DOM_W, DOM_H, name)
    if name == "droplet_v1":
        mat = opts.get("mat", mt.WATER)
        rx, ry = opts.get("rx", 13), opts.get("ry", 2)
        cx, cy = 64 * ONE, 22 * ONE
        fluid = []
        for gy in range(-ry, ry + 1):
            for gx in range(-rx, rx + 1):
                nx, ny = gx / rx, gy / ry
                if nx * nx + ny * ny > 1.0:
                    continue
                fluid.append((cx + gx * ONE, cy + gy * ONE, mat))
        return _assemble(fluid, boundary_ring(DOM_W, DOM_H, 60), DOM_W, DOM_H, name)
    if name == "empty_v1":
        w = max(8, int(opts.get("domWwu", DOM_W)))
        h = max(16, int(opts.get("domHwu", DOM_H)))
        return _assemble([], boundary_ring(w, h, h - 12), w, h, name)
    raise SceneError(f'unknown scene "{name}"')


def permute_scene(scene: Scene, seed: int) -> Scene:
    """Deterministic permutation of the fluid block (src/engine.js:2653-2679)."""
    rnd = xorshift32(seed)
    perm = list(range(scene.nFluid))
    for i in range(scene.nFluid - 1, 0, -1):
        j = rnd() % (i + 1)
        perm[i], perm[j] = perm[j], perm[i]
    out = Scene(n=scene.n, nFluid=scene.nFluid, pos=scene.pos.copy(), vel=scene.vel.copy(),
                matId=scene.matId.copy(), domW=scene.domW, domH=scene.domH,
                cellShift=scene.cellShift,
                flags=None if scene.flags is None else scene.flags.copy(),
                temp=None if scene.temp is None else scene.temp.copy(),
                fuel=None if scene.fuel is None else scene.fuel.copy(), name=scene.name + "+perm")
    p = np.asarray(perm, dtype=np.int64)
    nf = scene.nFluid
    out.pos[0:2 * nf:2] = scene.pos[2 * p]
    out.pos[1:2 * nf:2] = scene.pos[2 * p + 1]
    out.vel[0:2 * nf:2] = scene.vel[2 * p]
    out.vel[1:2 * nf:2] = scene.vel[2 * p + 1]
    out.matId[:nf] = scene.matId[p]
    for attr in ("temp", "fuel", "flags"):
        a = getattr(scene, attr)
        if a is not None:
            getattr(out, attr)[:nf] = a[p]
    return out


# ------------------------------------------------------------------- loading
@dataclass
class LoadImage:
    """What loadScene writes to the device: the 8-word records and the grid."""
    n: int
    nFluid: int
    state: np.ndarray       # int32 [n, 8]
    domW: int
    domH: int
    cellShift: int
    gridW: int
    gridH: int

    @property
    def cellTotal(self) -> int:
        return self.gridW * self.gridH


def grid_dims(dom_w: int, dom_h: int, cell_shift: int):
    cell = 2 ** cell_shift
    return max(1, -(-dom_w // cell)), max(1, -(-dom_h // cell))


def load_image(scene: Scene, mats=None, max_particles: Optional[int] = None,
               max_cells: Optional[int] = None) -> LoadImage:
    """Validate like loadScene and build its state words (src/engine.js:3704-3937).

    Differences from the reference host, by design: the reference refuses
    maxParticles > 131071 unconditionally (a 17-bit bond partner field,
    src/engine.js:3331-3334); here the native loader applies that cap per
    scene, and only in a world that runs the material passes (a fluid row
    with friction or bonds), because the partner field is only ever written
    for bondable rows.
    """
    rows = mt.MATS if mats is None else mats
    mat_count = len(rows)
    n, nf = int(scene.n), int(scene.nFluid)
    if n < 0:
        raise SceneError("scene.n must be a non-negative integer")
    if max_particles is not None and n > max_particles:
        raise SceneError(f"scene.n = {n} exceeds maxParticles = {max_particles}")
    if not (0 <= nf <= n):
        raise SceneError("scene.nFluid out of range")
    pos = np.asarray(scene.pos, dtype=np.int64)
    vel = np.asarray(scene.vel, dtype=np.int64)
    mid = np.asarray(scene.matId, dtype=np.int64)
    if pos.size != 2 * n or vel.size != 2 * n or mid.size != n:
        raise SceneError("scene array lengths must be n*2 / n*2 / n")
    if n and mid.max() >= mat_count:
        raise SceneError(f"matId >= matCount {mat_count}")
    px, py = pos[0::2], pos[1::2]
    lo = np.where(np.arange(n) < nf, 0, -MAX_POS)
    bad = (px < lo) | (px > MAX_POS) | (py < lo) | (py > MAX_POS)
    if bad.any():
        i = int(np.argmax(bad))
        raise SceneError(f"pos[{i}] outside [{int(lo[i])}, 2^27]")
    if scene.flags is not None:
        f = np.asarray(scene.flags, dtype=np.int64)
        if ((f & ~FLAG_MASK) != 0).any():
            raise SceneError("flags set a reserved bit (legal mask 0x7)")
    if scene.temp is not None:
        t = np.asarray(scene.temp, dtype=np.int64)
        if ((t < mt.TEMP_MIN) | (t > mt.TEMP_MAX)).any():
            raise SceneError("temp outside [TEMP_MIN, TEMP_MAX]")
    if scene.fuel is not None and (np.asarray(scene.fuel, dtype=np.int64) < 0).any():
        raise SceneError("fuel < 0")
    cs = int(scene.cellShift)
    if not (1 <= cs <= 27):
        raise SceneError(f"scene.cellShift = {cs} outside the 1..27 contract")
    if 2 ** cs < H:
        raise SceneError(f"scene.cellShift = {cs} gives cells smaller than the kernel radius")
    gw, gh = grid_dims(int(scene.domW), int(scene.domH), cs)
    if gw * gh > MAX_CELL_TOTAL:
        raise SceneError(f"cellTotal {gw * gh} exceeds the scan limit {MAX_CELL_TOTAL}")
    if max_cells is not None and gw * gh > max_cells:
        raise SceneError(f"cellTotal {gw * gh} exceeds allocated maxCells {max_cells}")

    F = mt.F
    phantom = np.array([r[F["phantom"]] != 0 for r in rows])
    seal = np.array([mt.seal_row(r) for r in rows])
    welds = np.array([mt.spawn_welds(r) for r in rows])
    fuel0 = np.array([r[F["fuel0"]] for r in rows], dtype=np.int64)
    given = (np.zeros(n, np.int64) if scene.flags is None
             else np.asarray(scene.flags, dtype=np.int64))
    want = np.where(phantom[mid], FLAG_DEAD, 0) if n else np.zeros(0, np.int64)
    if n and (((given & FLAG_DEAD) != 0) & (want == 0)).any():
        raise SceneError("scene.flags sets FLAG_DEAD on a non-phantom material")
    tag = np.where((np.arange(n) < nf) & welds[mid], 1 << WELD_TAG_SHIFT, 0) if n else np.zeros(0)
    st = np.zeros((n, PARTICLE_WORDS), dtype=np.int64)
    if n:
        st[:, 0] = px
        st[:, 1] = py
        st[:, 2] = vel[0::2]
        st[:, 3] = vel[1::2]
        st[:, 4] = mid
        st[:, 5] = given | want | np.where(seal[mid], FLAG_SEAL, 0) | tag
        st[:, 6] = (np.asarray(scene.temp, dtype=np.int64) if scene.temp is not None
                    else mt.AMBIENT)
        st[:, 7] = (np.asarray(scene.fuel, dtype=np.int64) if scene.fuel is not None
                    else fuel0[mid])
    st = (((st + 2 ** 31) % 2 ** 32) - 2 ** 31).astype(np.int32)
    return LoadImage(n=n, nFluid=nf, state=st, domW=int(scene.domW), domH=int(scene.domH),
                     cellShift=cs, gridW=gw, gridH=gh)


def spawn_records(scene: Scene, particles, mats=None):
    """The k fluid records and the rebuilt boundary block of spawnFluid
    (src/engine.js:5901-5971); both int32 [*, 8].  The fluid records carry
    the caller's FLAG_BURNING/FLAG_ANCHOR and the material-derived FLAG_DEAD
    and FLAG_SEAL but no weld tag (the engine adds the spawn's group id);
    the boundary block is the loaded scene's, velocity and flags zero, as
    the reference rewrites it (the engine takes it from its own state; the
    oracle is given it)."""
    rows = mt.MATS if mats is None else mats
    F = mt.F
    mat_count = len(rows)
    k = len(particles)
    fl = np.zeros((k, PARTICLE_WORDS), dtype=np.int64)
    for i, p in enumerate(particles):
        m = min(mat_count - 1, max(0, int(p.get("mat", mt.WATER))))
        row = rows[m]
        fl[i, 0] = int(p["x"])
        fl[i, 1] = int(p["y"])
        fl[i, 2] = int(p.get("vx", 0))
        fl[i, 3] = int(p.get("vy", 0))
        fl[i, 4] = m
        fl[i, 5] = ((int(p.get("flags", 0)) & SPAWN_FLAG_MASK) |
                    (FLAG_DEAD if row[F["phantom"]] != 0 else 0) |
                    (FLAG_SEAL if mt.seal_row(row) else 0))
        t = p.get("temp", row[F["spawnTemp"]])
        fl[i, 6] = max(mt.TEMP_MIN, min(mt.TEMP_MAX, int(t)))
        fl[i, 7] = max(0, int(p.get("fuel", row[F["fuel0"]])))
    nb = scene.n - scene.nFluid
    bd = np.zeros((nb, PARTICLE_WORDS), dtype=np.int64)
    for b in range(nb):
        src = scene.nFluid + b
        bm = int(scene.matId[src])
        bd[b, 0] = scene.pos[2 * src]
        bd[b, 1] = scene.pos[2 * src + 1]
        bd[b, 4] = bm
        bd[b, 6] = scene.temp[src] if scene.temp is not None else mt.AMBIENT
        bd[b, 7] = scene.fuel[src] if scene.fuel is not None else rows[bm][F["fuel0"]]
    wrap = lambda a: (((a + 2 ** 31) % 2 ** 32) - 2 ** 31).astype(np.int32)
    return wrap(fl), wrap(bd)


# ------------------------------------------------------------ benchmark scenes
def walled_block_scene(nx: int, ny: int, dom_w_wu: int, dom_h_wu: int, seed: int,
                       pitch: int = ONE >> 1, jitter: int = 1024, x0: int = 3 * ONE,
                       name: str = "block") -> Scene:
    """A ny x nx fluid block at `pitch` ticks standing on the floor at x0,
    jittered by xorshift32(seed) integers in [-jitter, jitter), inside the
    walled box boundary_ring(dom_w, dom_h, dom_h - 12)."""
    rnd = xorshift32(seed)
    span = 2 * jitter
    ys = (dom_h_wu - 3) * ONE
    fluid = []
    for gy in range(ny):
        for gx in range(nx):
            jx = (rnd() % span) - jitter
            jy = (rnd() % span) - jitter
            fluid.append((x0 + gx * pitch + jx, ys - gy * pitch + jy, mt.WATER))
    s = _assemble(fluid, boundary_ring(dom_w_wu, dom_h_wu, dom_h_wu - 12), dom_w_wu, dom_h_wu, name)
    s.meta = dict(seed=seed, pitch=pitch, jitter=jitter)
    return s


def dam_break_4096(seed: int = 0xC0FFEE) -> Scene:
    """Dam break: 64x64 water at 0.5 wu pitch, jitter in [-1/64, 1/64) wu, in a
    walled 256 x 128 wu box (976 boundary particles, 64 x 32 cells)."""
    return walled_block_scene(64, 64, 256, 128, seed, name="dambreak_4096")


def thermal_1m(seed: int = 0x5EED01, temp_seed: int = 0x7E3317, side: int = 1024) -> Scene:
    """Single 1,048,576-particle scene: a 1024 x 1024 water block at 0.5 wu
    pitch in a walled 2048 x 1024 wu box (8,144 boundary particles, 512 x 256
    cells), temperatures uniform integers in [0, 100) degrees from
    xorshift32(temp_seed).  `side` shrinks the block for parity-sized runs."""
    dom_w, dom_h = (2048, 1024) if side == 1024 else (max(64, side), max(64, side))
    s = walled_block_scene(side, side, dom_w, dom_h, seed, name=f"thermal_{side * side}")
    rnd = xorshift32(temp_seed)
    t = np.zeros(s.n, np.int32)
    for i in range(s.nFluid):
        t[i] = rnd() % (100 * ONE)
    t[s.nFluid:] = mt.AMBIENT
    s.temp = t
    s.meta.update(temp_seed=temp_seed, zero_temps=int((t[:s.nFluid] == 0).sum()))
    return s


def fluid_block_scene(nf: int, seed: int, x_pick: int = 0, name: str = "block") -> Scene:
    """Exactly `nf` water particles as a 0.5 wu block about 2:1 wide, standing
    on the floor of a walled 256 x 128 wu box (scaled up by a power of two
    until the block fits), at x position `x_pick` (in half-wu steps, taken
    modulo the free width), jittered by its own seed."""
    bw = int(math.ceil(math.sqrt(2 * nf)))
    bh = -(-nf // bw)
    k = 1
    while 256 * k < bw * 0.5 + 16 or 128 * k < bh * 0.5 + 24:
        k *= 2
    dom_w, dom_h = 256 * k, 128 * k
    width_wu = (bw * (ONE >> 1)) / ONE
    max_x0 = int((dom_w - 6 - width_wu) * 2)       # half-wu steps
    x0 = (3 * ONE) + (x_pick % max(1, max_x0)) * (ONE >> 1)
    s = walled_block_scene(bw, bh, dom_w, dom_h, seed, x0=x0, name=name)
    # trim the partial last row down to exactly nf fluid particles
    drop = bw * bh - nf
    if drop:
        keep = np.ones(s.n, bool)
        keep[nf:bw * bh] = False
        idx = np.nonzero(keep)[0]
        s = Scene(n=s.n - drop, nFluid=nf, pos=s.pos.reshape(-1, 2)[idx].reshape(-1).copy(),
                  vel=s.vel.reshape(-1, 2)[idx].reshape(-1).copy(), matId=s.matId[idx].copy(),
                  domW=s.domW, domH=s.domH, name=s.name, meta=s.meta)
    s.meta.update(block=(bw, bh), x0=int(x0), box=(dom_w, dom_h))
    return s


def batch_scenes(count: int = 512, seed: int = 0xBA7C4, lo: int = 1024, hi: int = 65536):
    """`count` independent scenes in the 256 x 128 wu walled box, fluid counts
    log-uniform in [lo, hi] (rounded), each a rectangular 0.5 wu block (about
    2:1 wide) at a seeded position on the floor, jittered by its own seed."""
    rnd = xorshift32(seed)
    out = []
    for k in range(count):
        u = (rnd() & 0xFFFFFF) / float(1 << 24)
        nf = int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))
        bw = int(math.ceil(math.sqrt(2 * nf)))
        bh = -(-nf // bw)
        width_wu = (bw * (ONE >> 1)) / ONE
        max_x0 = int((250 - width_wu) * 2)       # half-wu steps
        x0 = (3 * ONE) + (rnd() % max(1, max_x0)) * (ONE >> 1)
        s = walled_block_scene(bw, bh, 256, 128, splitmix32(seed + k), x0=x0, name=f"batch{k}")
        # trim the partial last row down to exactly nf fluid particles
        drop = bw * bh - nf
        if drop:
            keep = np.ones(s.n, bool)
            keep[nf:bw * bh] = False
            idx = np.nonzero(keep)[0]
            s = Scene(n=s.n - drop, nFluid=nf, pos=s.pos.reshape(-1, 2)[idx].reshape(-1).copy(),
                      vel=s.vel.reshape(-1, 2)[idx].reshape(-1).copy(), matId=s.matId[idx].copy(),
                      domW=s.domW, domH=s.domH, name=s.name, meta=s.meta)
        s.meta.update(block=(bw, bh), x0=int(x0))
        out.append(s)
    return out


def materials_mix(seed: int = 0x3A7E11) -> Scene:
    """A mixed-material 2D scene at the bonded-scene size limit, every block
    at rest spacing (1 wu, as dambreak_v1): a SAND heap, a METAL bar, a
    weld-tagged STONE slab with its bottom row held, RICE grains spread over
    the bond ramp's temperatures resting on the slab, and a water column, in
    a walled 1024 x 256 wu box.  109,928 particles (106,880 fluid), under
    the 131,071 of MAX_PARTICLES_BONDED; every material pass has work in it
    (friction, bonds and their breaks, containment, held-wall bursts,
    formation)."""
    rnd = xorshift32(seed)
    floor = (256 - 3) * ONE
    parts, flags, temps = [], [], []

    def block(mat, nx, ny, x0_wu, y_floor, jitter, temp=None, anchor_rows=0):
        for gy in range(ny):
            for gx in range(nx):
                jx = (rnd() % (2 * jitter)) - jitter
                jy = (rnd() % (2 * jitter)) - jitter
                parts.append((x0_wu * ONE + gx * ONE + jx, y_floor - gy * ONE + jy, mat))
                flags.append(FLAG_ANCHOR if gy < anchor_rows else 0)
                temps.append(mt.AMBIENT if temp is None else temp(gx, gy))

    block(mt.SAND, 200, 120, 10, floor, 2048)
    block(mt.METAL, 80, 6, 215, floor, 512)
    block(mt.STONE, 400, 6, 300, floor, 512, anchor_rows=1)
    block(mt.RICE, 200, 100, 320, floor - 7 * ONE, 2048, temp=lambda gx, gy: mt.DEG(40 + (gx * 7 + gy * 13) % 60))
    block(mt.WATER, 300, 200, 710, floor, 2048)
    s = _assemble(parts, boundary_ring(1024, 256, 256 - 6), 1024, 256, "materials_mix")
    f = np.zeros(s.n, np.uint32)
    f[: len(flags)] = flags
    t = np.full(s.n, mt.AMBIENT, np.int64)
    t[: len(temps)] = temps
    s.flags, s.temp = f, t.astype(np.int32)
    s.meta = dict(seed=seed)
    return s
