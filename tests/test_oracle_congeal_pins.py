"""The oracle's congeal path beyond a pool freezing on its own, by brute
force (src/sim.wgsl:2612-2730, 2854-2960).

LAVA below its freezing point beside a held STONE slab and a held METAL
slab, none weld-tagged.  On substep 0 the lava freezes to STONE with
FLAG_CONGEAL; substep 0 is a formation substep.  markCongealNbr wakes the
settled STONE slab (its neighbour rows now hold a congealed particle), so
its slots propose too, without the velocity and pressure gates.  The bond
set after the substep must equal the O(n^2) search over the final positions
for pairs with

* the same material row after the substep (no frozen-lava <-> METAL bond),
* at least one end that froze this substep (the settled slab does not bond
  to itself: STONE does not reform and carries no weld tag), and
* D2_FLOOR <= d2 <= BOND_R2 at 1/32-tick resolution, rest length isqrt(d2),

and every such bond is mutual.  FLAG_CONGEAL is cleared by the same
formation pass, and the held slabs keep FLAG_ANCHOR.

A freeze on a non-formation substep leaves FLAG_CONGEAL visible: it is set
when the product is a bonded row that holds shape (ICE) and not otherwise
(SAND).
"""
import math

import numpy as np

from sunaengine_amd import materials as mt
from sunaengine_amd import scenes
from tests import material_scenes as ms
from tests.helpers import oracle_world

ONE = 65536
BOND_R2 = 17623204
D2_FLOOR = 262144
FLAG_ANCHOR, FLAG_CONGEAL = 4, 32


def _rne(x, s):
    q, r = divmod(int(x), 1 << s)
    half = 1 << (s - 1)
    if r > half or (r == half and q & 1):
        q += 1
    return q


def test_frozen_lava_congeals_onto_a_settled_slab_only_of_its_own_row():
    F = mt.F
    assert mt.MATS[mt.LAVA][F["freezeTo"]] == mt.STONE
    assert mt.MATS[mt.STONE][F["bondReform"]] == 0
    rnd = scenes.xorshift32(0xC06E)
    floor = (30 - 3) * ONE
    stone = ms._block(rnd, mt.STONE, 8, 2, 6 * ONE, floor, ONE, jitter=256)
    metal = ms._block(rnd, mt.METAL, 6, 2, 15 * ONE, floor, ONE, jitter=256)
    lava = ms._block(rnd, mt.LAVA, 11, 2, 10 * ONE, floor - 2 * ONE, ONE, jitter=256)
    held = len(stone) + len(metal)
    cold = mt.MATS[mt.LAVA][F["freezePt"]] - mt.DEG(60)
    temps = [mt.DEG(20)] * held + [cold] * len(lava)
    s = ms._scene(stone + metal + lava, 30, 30, "lava_on_slabs", temps=temps,
                  flags=[FLAG_ANCHOR] * held + [0] * len(lava))
    o = oracle_world(s, 4)
    o.step(1)
    st = o.read_state()
    nf = s.nFluid
    assert o.debug_counters()[6] == 0                       # no list was truncated
    froze = np.zeros(nf, bool)
    froze[held:] = True
    assert (st[held:nf, 4] == mt.STONE).all()              # every lava slot froze
    assert not (st[:nf, 5] & FLAG_CONGEAL).any()
    assert (st[:held, 5] & FLAG_ANCHOR).all()
    bond = o.read_bond()
    got = set()
    for i in range(nf):
        for w in bond[i]:
            if w == 0:
                break
            got.add((i, int(w) & 0x1FFFF, (int(w) >> 17) & 0x1FFF))
    want = set()
    for i in range(nf):
        for j in range(nf):
            if i == j or st[i, 4] != st[j, 4] or not (froze[i] or froze[j]):
                continue
            d4x = _rne(int(st[i, 0]) - int(st[j, 0]), 5)
            d4y = _rne(int(st[i, 1]) - int(st[j, 1]), 5)
            if abs(d4x) > 5120 or abs(d4y) > 5120:
                continue
            d2 = d4x * d4x + d4y * d4y
            if D2_FLOOR <= d2 <= BOND_R2:
                want.add((i, j, math.isqrt(d2)))
    cross = {(i, j) for i, j, _ in want if froze[i] != froze[j]}
    assert len(cross) >= 10 and len(want) > 2 * len(lava), (len(cross), len(want))
    assert got == want
    assert all((j, i, L) in got for i, j, L in got)          # mutual
    # the settled slab did not bond to itself, nor anything to METAL
    assert not any(not froze[i] and not froze[j] for i, j, _ in got)
    assert not any(st[i, 4] == mt.METAL for i, _, _ in got)


def test_only_a_freeze_into_a_bonded_row_is_flagged():
    """A freeze marks FLAG_CONGEAL only when its product is a bonded row
    that holds shape (fuse_row): a row driven from 30 C toward -100 C
    (tempDrive 0.1, no ambient cooling) freezes on substep 1, which is not a
    formation substep, so the flag is still there after it: set for a
    product of ICE, not for one of SAND (src/sim.wgsl:2676-2693)."""
    from tests.test_oracle_bond_pins import _table
    F = mt.F
    assert mt.MATS[mt.SAND][F["bondK"]] == 0 and mt.MATS[mt.ICE][F["bondK"]] != 0
    for product, flagged in ((mt.ICE, True), (mt.SAND, False)):
        rows, ids = _table(C=(mt.WATER, {"freezePt": mt.DEG(10), "freezeTo": product, "coolRate": 0,
                                         "tempTarget": mt.DEG(-100), "tempDrive": 6554}))
        rnd = scenes.xorshift32(0xF1A6)
        pool = ms._block(rnd, ids["C"], 6, 3, 10 * ONE, 27 * ONE, ONE, jitter=256)
        s = ms._scene(pool, 30, 30, "drive_freeze", temps=[mt.DEG(30)] * len(pool))
        from oracle import OracleWorld
        from sunaengine_amd import tables
        o = OracleWorld(scenes.load_image(s, mats=rows), tables.load_luts(), mt.mats_image(rows), iters=4)
        nf = s.nFluid
        o.step(1)
        st = o.read_state()
        assert (st[:nf, 4] == ids["C"]).all() and (st[:nf, 6] > mt.DEG(10)).all()
        o.step(1)
        st = o.read_state()
        o.close()
        assert (st[:nf, 4] == product).all()
        assert bool((st[:nf, 5] & FLAG_CONGEAL).all()) == flagged and (flagged or not (st[:nf, 5] & FLAG_CONGEAL).any())
