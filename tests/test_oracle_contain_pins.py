"""The oracle's containment geometry against checks that do not restate it
(src/sim.wgsl:1970-2235).  Each case reads pred just before and just after
the containment passes (the oracle's contain probe), so the solver's own
moves are taken as given and only contain is judged.

* The crossing revert is the textbook segment-intersection test.  A held
  line of seal-class STONE with gaps of 0.3 to 1.6 wu, and water probes
  thrown across and along it from random starts: a probe's move is reverted
  to its start exactly when its straight move pos -> pred meets a seal
  segment, computed here by exact rational parametric intersection
  (pos + t m = p_j + u e, 0 <= t, u <= 1) over every pair of seal slots
  whose distance, at the 1/2048-wu resolution the pass works in, lies in
  [0.40, 1.45] wu.  Random starts make touching and collinear cases
  improbable; each case asserts none occurs.
* The keep-out push: a probe whose pred ends closer than 0.35 wu to a seal
  slot is moved radially out to 0.35 wu (within the pass's 1/2048-wu
  rounding and its floored-root normalisation); a non-seal solid (SAND) at
  the same place pushes nothing.
* A move that stops short of a slanted segment, inside its bounding box
  and with its line crossing the segment, is kept.
* Liquids pass a wall of sealMode rows: a seal_liquid probe (bondK = fric =
  0, grav > 0) crossing a segment whose two endpoints both have sealMode
  != 0 keeps its move; a frictionless non-liquid (water without gravity) on
  the same path, or water crossing a segment with one ordinary endpoint, is
  reverted.
"""
from fractions import Fraction

import numpy as np

from sunaengine_amd import materials as mt
from sunaengine_amd import scenes, tables
from tests import material_scenes as ms
from tests.test_oracle_bond_pins import _table

ONE = 65536
F = mt.F
FLAG_ANCHOR = 4
KEEP = 717 * 32                         # 0.35 wu in ticks (SEAL_KEEP4 << BPRE)
SEG_MIN2, SEG_MAX2 = 670761, 8820900    # (0.40 wu)^2, (1.45 wu)^2 at 2048 per wu
Y0 = 20 * ONE
DOM = 48


def _rne(x, s):
    q, r = divmod(int(x), 1 << s)
    half = 1 << (s - 1)
    if r > half or (r == half and q & 1):
        q += 1
    return q


def _world(parts, rows, flags, vel):
    from oracle import OracleWorld
    s = ms._scene(parts, DOM, DOM, "contain", flags=flags)
    for i, (vx, vy) in enumerate(vel):
        s.vel[2 * i], s.vel[2 * i + 1] = vx, vy
    im = scenes.load_image(s, mats=rows)
    return OracleWorld(im, tables.load_luts(), mt.mats_image(rows), iters=4), s


def _probe(parts, rows, flags, vel):
    from oracle.oracle import contain_probe
    o, s = _world(parts, rows, flags, vel)
    pos = o.read_state()[: s.nFluid, :2].astype(np.int64)
    r = contain_probe(o, s.nFluid)
    o.close()
    return pos, r["before"].astype(np.int64), r["after"].astype(np.int64)


def _seal_segments(seals):
    """Every pair of seal slots whose 1/2048-wu distance is in the
    segment range."""
    out = []
    for a in range(len(seals)):
        for b in range(a + 1, len(seals)):
            ex4 = _rne(seals[b][0] - seals[a][0], 5)
            ey4 = _rne(seals[b][1] - seals[a][1], 5)
            e2 = ex4 * ex4 + ey4 * ey4
            if SEG_MIN2 <= e2 <= SEG_MAX2:
                out.append((seals[a], seals[b]))
    return out


def _meets(p0, p1, a, b):
    """Exact parametric intersection of segments p0-p1 and a-b; None when a
    degenerate (parallel, collinear or touching) case would need the
    pass's tie rules."""
    mx, my = p1[0] - p0[0], p1[1] - p0[1]
    ex, ey = b[0] - a[0], b[1] - a[1]
    den = mx * ey - my * ex
    if den == 0:
        return None
    t = Fraction((a[0] - p0[0]) * ey - (a[1] - p0[1]) * ex, den)
    u = Fraction((a[0] - p0[0]) * my - (a[1] - p0[1]) * mx, den)
    if t in (0, 1) or u in (0, 1):
        return None
    return 0 < t < 1 and 0 < u < 1


def _line(rnd, gaps_wu):
    xs = [8 * ONE]
    for gw in gaps_wu:
        xs.append(xs[-1] + int(gw * ONE) + (rnd() % 2048) - 1024)
    # slanted segments: a move that stops short of one can still lie inside
    # its bounding box, where only the same-side test keeps it
    return [(x, Y0 + (rnd() % 16384) - 8192) for x in xs]


def test_crossing_revert_is_segment_intersection():
    rnd = scenes.xorshift32(0xC0217A)
    checked = reverted = kept = 0
    for trial in range(6):
        gaps = [(0.3, 0.6, 1.0, 1.3, 1.6, 0.45, 1.42)[(rnd() % 7)] for _ in range(18)]
        stones = _line(rnd, gaps)
        x_lo, x_hi = stones[0][0], stones[-1][0]
        probes, vel = [], []
        for _ in range(48):
            x = x_lo + rnd() % (x_hi - x_lo)
            y = Y0 - int(0.15 * ONE) - rnd() % int(0.45 * ONE)
            probes.append((x, y, mt.WATER))
            vel.append((int((rnd() % 24000) - 12000), int(0.05 * ONE) + rnd() % int(0.7 * ONE)))
        parts = [(x, y, mt.STONE) for x, y in stones] + probes
        flags = [FLAG_ANCHOR] * len(stones) + [0] * len(probes)
        pos, before, after = _probe(parts, None, flags, [(0, 0)] * len(stones) + vel)
        ns = len(stones)
        seals = [tuple(before[j]) for j in range(ns)]
        assert all(tuple(pos[j]) == seals[j] for j in range(ns)), "held seals moved"
        segs = _seal_segments(seals)
        assert segs
        for i in range(ns, len(parts)):
            p0, p1 = tuple(pos[i]), tuple(before[i])
            near = min((p1[0] - s[0]) ** 2 + (p1[1] - s[1]) ** 2 for s in seals)
            if near < (KEEP + 64) ** 2:
                continue                   # the keep-out push moves it first: its own test
            hits = [_meets(p0, p1, a, b) for a, b in segs]
            assert None not in hits, (trial, i)
            want = p0 if any(hits) else p1
            assert tuple(after[i]) == want, (trial, i, p0, p1, tuple(after[i]))
            checked += 1
            reverted += 1 if any(hits) else 0
            kept += 0 if any(hits) else 1
    assert checked > 150 and reverted > 40 and kept > 40, (checked, reverted, kept)


def test_keep_out_push_is_radial_to_the_keep_distance():
    rnd = scenes.xorshift32(0x4EE9)
    pushed = 0
    for trial in range(30):
        stone = (24 * ONE, Y0)
        ang = (rnd() % 6283) / 1000.0
        r = int(0.05 * ONE) + rnd() % int(0.2 * ONE)
        probe = (stone[0] + int(r * np.cos(ang)), stone[1] + int(r * np.sin(ang)))
        for solid, pushes in ((mt.STONE, True), (mt.SAND, False)):
            parts = [(stone[0], stone[1], solid), (probe[0], probe[1], mt.WATER)]
            pos, before, after = _probe(parts, None, [FLAG_ANCHOR, 0], [(0, 0), (0, 0)])
            d0 = before[1] - before[0]
            r0 = float(np.hypot(*d0))
            if not pushes:
                assert tuple(after[1]) == tuple(before[1])
                continue
            if r0 >= KEEP - 64 or r0 < 2048:
                continue
            d1 = after[1] - before[0]
            r1 = float(np.hypot(*d1))
            # the unit vector divides by the floored integer root d4 of the
            # 1/2048-wu distance, so the push can overshoot by (KEEP - d4) / d4
            # of those units; and each 1/2048-wu rounding is half a unit
            d4 = r0 / 32.0
            over = (KEEP / 32.0 - d4) / max(d4 - 1.0, 1.0) * 32.0
            assert KEEP - 64 <= r1 <= KEEP + over + 64, (trial, r0, r1, over)
            cross = float(d0[0]) * d1[1] - float(d0[1]) * d1[0]
            assert abs(cross) <= 64 * r1 and float(np.dot(d0, d1)) > 0, (trial, d0, d1)
            pushed += 1
    assert pushed >= 15, pushed


def test_liquids_pass_sealmode_walls():
    # W0: water without gravity, so not a seal liquid (grav > 0 is part of
    # the definition), and frictionless, so its move is the same as water's
    rows, ids = _table(SM=(mt.STONE, {"sealMode": 1}), W0=(mt.WATER, {"grav": 0}))
    assert mt.MATS[mt.WATER][F["fric"]] == 0 and mt.MATS[mt.WATER][F["grav"]] > 0
    stones = [(20 * ONE, Y0), (21 * ONE, Y0)]
    start = (int(20.5 * ONE), Y0 - int(0.25 * ONE))
    for walls, probe_mat, reverted in (((ids["SM"], ids["SM"]), mt.WATER, False),
                                       ((ids["SM"], ids["SM"]), ids["W0"], True),
                                       ((ids["SM"], mt.STONE), mt.WATER, True),
                                       ((mt.STONE, mt.STONE), mt.WATER, True)):
        parts = [(stones[0][0], stones[0][1], walls[0]), (stones[1][0], stones[1][1], walls[1]),
                 (start[0], start[1], probe_mat)]
        pos, before, after = _probe(parts, rows, [FLAG_ANCHOR, FLAG_ANCHOR, 0],
                                    [(0, 0), (0, 0), (0, ONE)])
        assert before[2][1] > Y0 + KEEP, "the probe's move does not cross the line"
        assert _meets(tuple(pos[2]), tuple(before[2]), tuple(before[0]), tuple(before[1]))
        want = tuple(pos[2]) if reverted else tuple(before[2])
        assert tuple(after[2]) == want, (walls, probe_mat)


def test_move_stopping_short_of_a_slanted_segment_is_kept():
    """A move inside a slanted segment's bounding box, whose line crosses
    the segment but which ends before it, is kept; the same move carried on
    through the segment is reverted."""
    a = (20 * ONE, Y0 - int(0.3 * ONE))
    b = (int(21.2 * ONE), Y0 + int(0.3 * ONE))
    start = (int(20.6 * ONE), Y0 - int(0.4 * ONE))
    for vy, crosses in ((int(0.3 * ONE), False), (int(0.7 * ONE), True)):
        parts = [(a[0], a[1], mt.STONE), (b[0], b[1], mt.STONE), (start[0], start[1], mt.WATER)]
        pos, before, after = _probe(parts, None, [FLAG_ANCHOR, FLAG_ANCHOR, 0], [(0, 0), (0, 0), (0, vy)])
        p0, p1, pa, pb = tuple(pos[2]), tuple(before[2]), tuple(before[0]), tuple(before[1])
        assert min((p1[0] - s[0]) ** 2 + (p1[1] - s[1]) ** 2 for s in (pa, pb)) > (KEEP + 64) ** 2
        if not crosses:
            assert min(pa[1], pb[1]) <= p1[1] <= max(pa[1], pb[1])     # inside the bounding box
        assert bool(_meets(p0, p1, pa, pb)) == crosses
        assert tuple(after[2]) == (p0 if crosses else p1), (vy, p0, p1, tuple(after[2]))
