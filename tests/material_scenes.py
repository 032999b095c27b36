"""Seeded 2D scenes that drive the material passes (friction, bonds,
containment, burst, freezing).  Every builder returns a scenes.Scene; each
names what it is meant to reach, and the tests check on the oracle that it
does reach it.
"""
import numpy as np

from sunaengine_amd import materials as mt
from sunaengine_amd import scenes

ONE = 65536
DEG = mt.DEG


def _block(rnd, mat, nx, ny, x0, y_floor, pitch, jitter=512):
    """ny x nx particles of `mat` at `pitch` ticks, bottom row at y_floor,
    jittered by [-jitter, jitter) ticks."""
    out = []
    for gy in range(ny):
        for gx in range(nx):
            jx = (rnd() % (2 * jitter)) - jitter
            jy = (rnd() % (2 * jitter)) - jitter
            out.append((x0 + gx * pitch + jx, y_floor - gy * pitch + jy, mat))
    return out


def _scene(parts, dom_w_wu, dom_h_wu, name, temps=None, flags=None):
    s = scenes._assemble(parts, scenes.boundary_ring(dom_w_wu, dom_h_wu, dom_h_wu - 6), dom_w_wu, dom_h_wu, name)
    if temps is not None:
        t = np.full(s.n, mt.AMBIENT, np.int64)
        t[: len(temps)] = temps
        s.temp = t.astype(np.int32)
    if flags is not None:
        f = np.zeros(s.n, np.uint32)
        f[: len(flags)] = flags
        s.flags = f
    return s


def sand_and_water(seed=0x5A4D):
    """A SAND heap beside a water column: friction between grains, grains
    and water (mat_mix), grains and the walls."""
    rnd = scenes.xorshift32(seed)
    floor = (40 - 3) * ONE
    parts = _block(rnd, mt.SAND, 28, 20, 6 * ONE, floor, ONE >> 1)
    parts += _block(rnd, mt.WATER, 20, 24, 24 * ONE, floor, ONE >> 1)
    return _scene(parts, 56, 40, "sand_and_water")


def stone_in_water(seed=0x57011E):
    """A weld-tagged STONE slab (seal class: bonds form at substep 0, water
    beside it is contained) with its bottom row held (FLAG_ANCHOR: the burst
    test of containStress applies), under a water column."""
    rnd = scenes.xorshift32(seed)
    floor = (48 - 3) * ONE
    stone = _block(rnd, mt.STONE, 16, 5, 10 * ONE, floor, ONE, jitter=256)
    water = _block(rnd, mt.WATER, 36, 30, 5 * ONE, floor - 6 * ONE, ONE >> 1)
    flags = [scenes.FLAG_ANCHOR if k < 16 else 0 for k in range(len(stone))] + [0] * len(water)
    return _scene(stone + water, 48, 48, "stone_in_water", flags=flags)


def rice_on_rubber(seed=0x41CE):
    """RICE grains whose temperatures span the bond ramp (45..95 C: bond_k
    varies per particle), falling onto a RUBBER slab: bonds reform under the
    velocity and pressure gates and break past bondBreak."""
    rnd = scenes.xorshift32(seed)
    floor = (44 - 3) * ONE
    rubber = _block(rnd, mt.RUBBER, 24, 3, 6 * ONE, floor, ONE, jitter=256)
    rice = _block(rnd, mt.RICE, 26, 16, 5 * ONE, floor - 6 * ONE, (ONE * 3) >> 2)
    temps = [mt.AMBIENT] * len(rubber) + [DEG(40 + (k * 37) % 60) for k in range(len(rice))]
    return _scene(rubber + rice, 40, 44, "rice_on_rubber", temps=temps)


def freezing_pool(seed=0xF2EE):
    """Water, the left half below 0 C: it freezes to ICE on the first
    substep (FLAG_CONGEAL, congeal bonds with same-row neighbours), then
    warms and melts back.  Every row at load is frictionless and bondless,
    so the world must be created with materials on."""
    rnd = scenes.xorshift32(seed)
    floor = (36 - 3) * ONE
    water = _block(rnd, mt.WATER, 40, 20, 4 * ONE, floor, ONE >> 1)
    temps = [DEG(-12) if (k % 40) < 20 else DEG(25) for k in range(len(water))]
    return _scene(water, 32, 36, "freezing_pool", temps=temps)


def packed_against_stone(seed=0xD0C7):
    """Water packed at 0.25 wu against a held STONE wall: neighbour lists
    fill (containOverflow gathers the seal slots itself) and the crush test
    of containStress bursts wall particles."""
    rnd = scenes.xorshift32(seed)
    floor = (40 - 3) * ONE
    wall = _block(rnd, mt.STONE, 3, 20, 20 * ONE, floor, ONE >> 1, jitter=128)
    water = _block(rnd, mt.WATER, 44, 30, 8 * ONE + (ONE >> 3), floor, ONE >> 2, jitter=256)
    flags = [scenes.FLAG_ANCHOR] * len(wall) + [0] * len(water)
    return _scene(wall + water, 40, 40, "packed_against_stone", flags=flags)


def hot_metal(seed=0x3E7A1):
    """A METAL bar whose right half is past its melting point: melted
    partners leave the bond rows (bondK of the new row is 0, the bond is
    tombstoned), the ramp weakens the rest."""
    rnd = scenes.xorshift32(seed)
    floor = (30 - 3) * ONE
    bar = _block(rnd, mt.METAL, 24, 4, 6 * ONE, floor, ONE, jitter=256)
    temps = [DEG(1150) if (k % 24) < 12 else DEG(1250) for k in range(len(bar))]
    return _scene(bar, 36, 30, "hot_metal", temps=temps)


def seal_crossing(seed=0xC505):
    """A held STONE line (seal class, anchored) and water probes just above
    it, thrown down across it at 0.6 wu per substep, between the stones:
    each probe's straight move crosses a seal segment, so contain reverts it
    to where it started (src/sim.wgsl:2104-2135).  A second row of probes
    is thrown straight at stones (the keep-out push)."""
    rnd = scenes.xorshift32(seed)
    y_line = 20 * ONE
    stones = [(10 * ONE + k * ONE + (rnd() % 64), y_line, mt.STONE) for k in range(16)]
    probes = [(10 * ONE + k * ONE + (ONE >> 1), y_line - (ONE // 5), mt.WATER) for k in range(0, 15, 2)]
    probes += [(10 * ONE + k * ONE, y_line - (ONE // 3), mt.WATER) for k in range(1, 15, 2)]
    s = _scene(stones + probes, 40, 40, "seal_crossing",
               flags=[scenes.FLAG_ANCHOR] * len(stones) + [0] * len(probes))
    for i in range(len(stones), len(stones) + len(probes)):
        s.vel[2 * i + 1] = (ONE * 3) // 5
    s.meta = dict(n_stones=len(stones), n_cross=8)
    return s


def dense_crossing(seed=0xDC05):
    """Water packed at 0.25 wu (every list full, so containOverflow gathers
    the seal slots itself) thrown down at 0.6 wu per substep onto a held
    STONE line wider than the water: no water particle may end a substep
    below the line."""
    rnd = scenes.xorshift32(seed)
    y_line = 24 * ONE
    stones = [(8 * ONE + k * ONE + (rnd() % 64), y_line, mt.STONE) for k in range(24)]
    water = _block(rnd, mt.WATER, 40, 8, 12 * ONE, y_line - (ONE // 4), ONE >> 2, jitter=128)
    s = _scene(stones + water, 40, 40, "dense_crossing",
               flags=[scenes.FLAG_ANCHOR] * len(stones) + [0] * len(water))
    for i in range(len(stones), len(stones) + len(water)):
        s.vel[2 * i + 1] = (ONE * 3) // 5
    s.meta = dict(n_stones=len(stones), y_line=y_line)
    return s


# name -> (builder, substeps, create the world with materials on)
CASES = {
    "sand_and_water": (sand_and_water, 48, False),
    "stone_in_water": (stone_in_water, 48, False),
    "rice_on_rubber": (rice_on_rubber, 48, False),
    "freezing_pool": (freezing_pool, 40, True),
    "packed_against_stone": (packed_against_stone, 32, False),
    "hot_metal": (hot_metal, 32, False),
    "seal_crossing": (seal_crossing, 16, False),
    "dense_crossing": (dense_crossing, 16, False),
}
