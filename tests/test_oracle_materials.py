"""The oracle's material passes against checks that do not restate them.

* Freezing: water poured below 0 C becomes ICE on substep 0 and forms
  congeal bonds to exactly the brute-force set of same-row pairs in range;
  water at 100 C becomes its boil row.
* Weld bonds by brute force: a weld-tagged STONE slab at rest forms, on
  substep 0, exactly the bonds an O(n^2) search over its final positions
  gives (every same-tag pair with D2_FLOOR <= d2 <= BOND_R2 at 1/32-tick
  resolution, rest length isqrt(d2)), src/sim.wgsl:2794-3097.
* bondForm2 leaves only mutual bonds: right after a formation substep every
  word has its mirror in the partner's row with the same rest length, and no
  tombstone survives.
* The bond fold of the state hash equals the vectorised digestState twin
  (sunaengine_amd.tables.digest_words, after src/hash/statehash.js:80-159).
* A move across a held seal segment is reverted: the particle ends the
  substep where it started (src/sim.wgsl:2104-2135); with every list full
  (containOverflow) packed water stays above a held line.
* Burning spends fuel at exactly burnRate per substep; ICE melts past its
  melting point; applyImpulse and spawnFluid's slot layout match their
  closed forms.
* Friction is antisymmetric: with no boundary neighbour, no truncation and
  no saturating term, the friction terms of a SAND heap cancel pairwise, so
  one application of friction moves the heap's summed dp by zero
  (src/sim.wgsl:1727-1781).
"""
import math

import numpy as np

from sunaengine_amd import materials as mt
from sunaengine_amd import scenes, tables
from tests import material_scenes as ms
from tests.helpers import oracle_world

ONE = 65536
BOND_R2 = 17623204
D2_FLOOR = 262144
TOMB = 0x40000000


def _shr_rne(x, s):
    """RNE arithmetic shift, exact integer arithmetic."""
    q, r = divmod(int(x), 1 << s)
    half = 1 << (s - 1)
    if r > half or (r == half and q & 1):
        q += 1
    return q


def _rows(bond, nf):
    out = []
    for i in range(nf):
        row = []
        for w in bond[i]:
            if w == 0:
                break
            row.append(int(w))
        out.append(row)
    return out


def test_weld_bonds_match_brute_force():
    rnd = scenes.xorshift32(0xB0D5)
    floor = (30 - 3) * ONE
    slab = ms._block(rnd, mt.STONE, 12, 4, 8 * ONE, floor, ONE, jitter=256)
    s = ms._scene(slab, 30, 30, "stone_slab")
    o = oracle_world(s, 4)
    o.step(1)
    st = o.read_state()
    bond = o.read_bond()
    nf = s.nFluid
    got = {(i, w & 0x1FFFF, (w >> 17) & 0x1FFF) for i, row in enumerate(_rows(bond, nf)) for w in row}
    want = set()
    for i in range(nf):
        for j in range(nf):
            if i == j:
                continue
            d4x = _shr_rne(int(st[i, 0]) - int(st[j, 0]), 5)
            d4y = _shr_rne(int(st[i, 1]) - int(st[j, 1]), 5)
            if abs(d4x) > 5120 or abs(d4y) > 5120:
                continue
            d2 = d4x * d4x + d4y * d4y
            if D2_FLOOR <= d2 <= BOND_R2:
                want.add((i, j, math.isqrt(d2)))
    assert len(want) > 4 * nf
    assert got == want
    # the weld tags are cleared by the same formation pass
    assert not (st[:nf, 5] & 0x000FFF00).any()


def test_freezing_forms_congeal_bonds_by_brute_force():
    """Water poured at -10 C freezes to its freezeTo row (ICE) on substep 0's
    thermal pass, with FLAG_CONGEAL and FLAG_SEAL; substep 0 is a formation
    substep, so bondForm1 bonds every same-row pair in range without the
    velocity and pressure gates, bondForm2 keeps the mutual ones and clears
    FLAG_CONGEAL (src/sim.wgsl:2612-2707, 2794-3097).  After one substep the
    bond set equals the O(n^2) pair search over the final positions."""
    rnd = scenes.xorshift32(0xF0E5)
    floor = (30 - 3) * ONE
    pool = ms._block(rnd, mt.WATER, 12, 6, 8 * ONE, floor, ONE, jitter=256)
    s = ms._scene(pool, 30, 30, "cold_pool", temps=[mt.DEG(-10)] * len(pool))
    o = oracle_world(s, 4)
    o.step(1)
    st = o.read_state()
    nf = s.nFluid
    ice = mt.MATS[mt.WATER][mt.F["freezeTo"]]
    assert (st[:nf, 4] == ice).all()
    assert ((st[:nf, 5] & 128) != 0).all() and not (st[:nf, 5] & 32).any()   # SEAL set, CONGEAL cleared
    bond = o.read_bond()
    got = {(i, w & 0x1FFFF, (w >> 17) & 0x1FFF) for i, row in enumerate(_rows(bond, nf)) for w in row}
    want = set()
    for i in range(nf):
        for j in range(nf):
            if i == j:
                continue
            d4x = _shr_rne(int(st[i, 0]) - int(st[j, 0]), 5)
            d4y = _shr_rne(int(st[i, 1]) - int(st[j, 1]), 5)
            if abs(d4x) > 5120 or abs(d4y) > 5120:
                continue
            d2 = d4x * d4x + d4y * d4y
            if D2_FLOOR <= d2 <= BOND_R2:
                want.add((i, j, math.isqrt(d2)))
    assert len(want) > 4 * nf
    assert got == want


def test_boiling_water_turns_to_its_boil_row():
    """Water held uniformly at 101 C (no conduction; cooling toward ambient
    takes off 0.05 C) is past boilPt on substep 0's thermal pass and becomes
    its boilTo row with that row's fuel0; at 99 C it stays water
    (src/sim.wgsl:2612-2707)."""
    steam = mt.MATS[mt.WATER][mt.F["boilTo"]]
    for deg, want in ((101, steam), (99, mt.WATER)):
        rnd = scenes.xorshift32(0xB011)
        floor = (30 - 3) * ONE
        pool = ms._block(rnd, mt.WATER, 12, 6, 8 * ONE, floor, ONE, jitter=256)
        s = ms._scene(pool, 30, 30, "hot_pool", temps=[mt.DEG(deg)] * len(pool))
        o = oracle_world(s, 4)
        o.step(1)
        st = o.read_state()
        assert (st[: s.nFluid, 4] == want).all(), deg
        assert (st[: s.nFluid, 7] == mt.MATS[want][mt.F["fuel0"]]).all()


def test_bonds_are_mutual_after_formation():
    s = ms.rice_on_rubber()
    o = oracle_world(s, 4)
    done = 0
    for k in (1, 5, 21, 41):
        o.step(k - done)
        done = k
        bond = o.read_bond()
        assert not (bond & TOMB).any()
        rows = _rows(bond, s.nFluid)
        pairs = {(i, w & 0x1FFFF): (w >> 17) & 0x1FFF for i, row in enumerate(rows) for w in row}
        assert pairs, "no bonds formed"
        for (i, j), l0 in pairs.items():
            assert pairs.get((j, i)) == l0, f"bond {i}->{j} has no mirror after substep {k}"


def test_bond_fold_matches_digest_twin():
    for build, k in ((ms.stone_in_water, 6), (ms.rice_on_rubber, 10), (ms.freezing_pool, 7)):
        s = build()
        o = oracle_world(s, 4)
        o.step(k)
        st = o.read_state()
        bond = o.read_bond()
        assert (bond != 0).any()
        want = tables.digest_words(st.reshape(-1), 8, 8, bond=bond, n_fluid=s.nFluid)
        assert want == [int(x) for x in o.digest()], s.name
        # and the fold matters: without it the digest differs
        assert tables.digest_words(st.reshape(-1), 8, 8) != want


def test_crossing_a_seal_segment_reverts_the_move():
    """Water thrown across a held STONE line (each probe's straight move
    crosses the segment between two stones) ends the substep where it
    started, with zero velocity; the held stones do not move."""
    s = ms.seal_crossing()
    o = oracle_world(s, 4)
    st0 = o.read_state().copy()
    o.step(1)
    st = o.read_state()
    ns = s.meta["n_stones"]
    probes = slice(ns, s.nFluid)
    assert np.array_equal(st[probes, :2], st0[probes, :2])
    assert not st[probes, 2:4].any()
    assert np.array_equal(st[:ns, :2], st0[:ns, :2])
    # without the line, the same probes fall through: the revert is contain's doing
    s2 = ms.seal_crossing()
    s2.matId[:ns] = mt.WATER
    s2.flags[:ns] = 0
    o2 = oracle_world(s2, 4)
    o2.step(1)
    assert (o2.read_state()[probes, 1] > st0[probes, 1]).all()


def test_dense_crossing_is_contained_with_full_lists():
    """Water packed at 0.25 wu (every list full: containOverflow gathers the
    seal slots from the cells itself) thrown across a held STONE line: after
    one substep no water particle is below the line, and without the line
    they fall through.  (From the next substep on the crushed wall bursts,
    src/sim.wgsl:2742-2771, and the water breaks through.)"""
    s = ms.dense_crossing()
    ns, yl = s.meta["n_stones"], s.meta["y_line"]
    o = oracle_world(s, 4)
    o.step(1)
    st = o.read_state()
    assert o.debug_counters()[6] > 0
    assert st[ns:s.nFluid, 1].max() <= yl
    s2 = ms.dense_crossing()
    s2.matId[:ns] = mt.WATER
    s2.flags[:ns] = 0
    o2 = oracle_world(s2, 4)
    o2.step(1)
    assert o2.read_state()[ns:s2.nFluid, 1].max() > yl


def test_friction_terms_cancel_pairwise():
    from oracle import oracle as orc
    rnd = scenes.xorshift32(0xF71C)
    # a loose SAND cloud in mid-air, far from every wall: no boundary neighbour
    parts = ms._block(rnd, mt.SAND, 14, 10, 20 * ONE, 30 * ONE, (ONE * 3) >> 2, jitter=4096)
    s = ms._scene(parts, 60, 60, "sand_cloud")
    vel = np.zeros(2 * s.n, np.int32)
    r2 = scenes.xorshift32(0x1234)
    for i in range(s.nFluid):
        vel[2 * i] = (r2() % 8192) - 4096
        vel[2 * i + 1] = (r2() % 8192) - 4096
    s.vel = vel
    w = oracle_world(s, 1)
    sums = orc.friction_probe(w, s.nFluid)
    assert sums["pairs"] > 500 and sums["boundary"] == 0 and sums["full"] == 0
    assert sums["dp"] == (0, 0)


def _pool(mat, deg, w=12, h=6, seed=0x7007):
    rnd = scenes.xorshift32(seed)
    floor = (30 - 3) * ONE
    pool = ms._block(rnd, mat, w, h, 8 * ONE, floor, ONE, jitter=256)
    return ms._scene(pool, 30, 30, f"pool_{mat}", temps=[mt.DEG(deg)] * len(pool))


def test_burning_spends_fuel_at_the_burn_rate():
    """A row with fuel, held past its ignition point, burns from substep 0:
    FLAG_BURNING is set and fuel falls by exactly burnRate per substep while
    it lasts, whatever the temperature does afterwards (burning is sticky,
    src/sim.wgsl:2612-2707)."""
    row = 4
    R = mt.MATS[row]
    F = mt.F
    s = _pool(row, 250)
    o = oracle_world(s, 4)
    for k in (1, 5):
        o.step(1 if k == 1 else 4)
        st = o.read_state()
        assert (st[: s.nFluid, 4] == row).all()
        assert ((st[: s.nFluid, 5] & 1) != 0).all()
        assert (st[: s.nFluid, 7] == R[F["fuel0"]] - k * R[F["burnRate"]]).all(), k


def test_ice_melts_past_its_melting_point():
    """ICE held at 5 C melts to its meltTo row on substep 0; at -5 C it stays."""
    ice = mt.MATS[mt.WATER][mt.F["freezeTo"]]
    for deg, want in ((5, mt.MATS[ice][mt.F["meltTo"]]), (-5, ice)):
        s = _pool(ice, deg)
        o = oracle_world(s, 4)
        o.step(1)
        assert (o.read_state()[: s.nFluid, 4] == want).all(), deg


def test_steam_condenses_below_its_condensation_point():
    """The boil row of water held at 30 C (below its condPt) becomes its
    condTo row on substep 0; at 90 C it stays."""
    steam = mt.MATS[mt.WATER][mt.F["boilTo"]]
    R = mt.MATS[steam]
    assert R[mt.F["condPt"]] > mt.DEG(30) and R[mt.F["condPt"]] < mt.DEG(90)
    for deg, want in ((30, R[mt.F["condTo"]]), (90, steam)):
        s = _pool(steam, deg)
        o = oracle_world(s, 4)
        o.step(1)
        assert (o.read_state()[: s.nFluid, 4] == want).all(), deg


def test_apply_impulse_closed_form():
    """applyImpulse (src/engine.js:6077-6098) on the current state: fluid
    slots in ascending order whose centre is within r (the test in IEEE
    doubles, as JavaScript computes it), at most `limit` of them, get (ix, iy)
    added to their velocity, clamped to +-VMAX; nothing else changes."""
    s = scenes.make_scene("pool_v1", w=40, h=12)
    o = oracle_world(s, 4)
    o.step(3)
    st0 = o.read_state().copy()
    cx, cy, r, ix, iy, limit = 30 * ONE, 60 * ONE, 6 * ONE, 40000, -50000, 25
    dx = st0[: s.nFluid, 0].astype(np.float64) - float(cx)
    dy = st0[: s.nFluid, 1].astype(np.float64) - float(cy)
    inside = np.nonzero(dx * dx + dy * dy <= float(r) * float(r))[0]
    assert inside.size > limit
    hit = inside[:limit]
    want = st0.copy()
    want[hit, 2] = np.clip(want[hit, 2].astype(np.int64) + ix, -65536, 65536)
    want[hit, 3] = np.clip(want[hit, 3].astype(np.int64) + iy, -65536, 65536)
    assert o.apply_impulse(cx, cy, r, ix, iy, limit) == limit
    assert np.array_equal(o.read_state(), want)


def test_spawn_fluid_slot_layout():
    """spawnFluid (src/engine.js:5901-6020): the new fluid records take slots
    nFluid..nFluid+k-1, the boundary block moves up by k unchanged, existing
    fluid slots keep their words, and the new slots' bond rows are empty."""
    s = ms.stone_in_water()
    o = oracle_world(s, 4, max_particles=s.n + 16)
    o.step(5)
    st0 = o.read_state().copy()
    nf, nb = s.nFluid, s.n - s.nFluid
    drop = [dict(x=(14 + dx) * ONE, y=12 * ONE, vx=0, vy=2000, mat=mt.SAND) for dx in range(5)]
    fl, bd = scenes.spawn_records(s, drop)
    assert o.spawn(fl, bd) == s.n + 5
    st = o.read_state()
    assert np.array_equal(st[:nf], st0[:nf])
    assert np.array_equal(st[nf:nf + 5], fl)
    assert np.array_equal(st[nf + 5:nf + 5 + nb], bd)
    assert not o.read_bond()[nf:nf + 5].any()
