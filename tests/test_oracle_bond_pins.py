"""The oracle's bond gather break rule, bondForm1's reform gates and
containStress, pinned by closed forms and brute force
(src/sim.wgsl:1790-1900, 2742-2771, 2794-3097).

Most cases hold their particles still: a particle with FLAG_ANCHOR keeps
pred == pos through predict and the solver (applyDp skips it) and contain
skips it, so the distances these passes see are exactly the ones the test
wrote.  A bond under test is formed on substep 0 (rows with a weld tag at
load bond without gates), then the pair is written back anchored at a
chosen offset and one more substep runs.  Substep 1 is not a formation
substep, so a bond the gather broke still shows its tombstone bit.

* Break rule, no pressure term: a bond of rest length L0 breaks iff
  d2 > T, T = L0^2 * min(bondBreak_i, bondBreak_j) / 2^16 rounded to
  nearest-even (the fixed-point core's fp_mul_s; d in 1/32-tick units).
  The formation gap is chosen so that T is a sum of two squares: the pair
  survives at d2 == T exactly and breaks at the next sum of two squares
  above it, on a diagonal and on an axis (isqrt(T) and one more), for
  equal and for mixed rows.
* bondBreak == 0 (STONE) survives every stretch below the kernel radius and
  breaks at d2 >= 5120^2.
* A partner with FLAG_BURST or on a bondK == 0 row loses the bond from the
  other side in the same substep; a partner whose temperature ramp has
  taken kEff to 0 keeps it.
* Mohr-Coulomb: with bondMu != 0 on both rows the bound becomes
  L0^2 * (bondBreak + Pq * bondMu / 2^16) / 2^16 with Pq = -(lam_i + lam_j)
  / 16 (Q20 -> Q16) when positive.  lam is read back from the substep (all
  particles anchored: every iteration computes the same lam).  A pair that
  survives only because of the pressure term breaks when either row has
  bondMu == 0.
* Reform gates by brute force: RICE at mixed temperatures and velocities,
  one substep from load (a formation substep; RICE rows carry no weld tag,
  so every bond comes from the reform path).  The bond set equals an
  O(n * list) search over the substep's own neighbour lists and final
  state: both kEff != 0, same bond group, range, the relative-velocity box
  and disc at 1/8-tick resolution, both -lam >= bondFormP, L0 range, and
  mutual proposal with equal L0.
* containStress: an anchored STONE with anchored WATER around it bursts
  iff 32 * sum(2560 - isqrt(d2)) > burstP over non-seal, non-phantom
  neighbours closer than 2560; burst clears FLAG_ANCHOR and the weld tag.
  Cases sit on either side of the bound, with one where a rounded square
  root would cross it, and with seal-class and phantom neighbours that
  must not count.
"""
import math

import numpy as np

from sunaengine_amd import materials as mt
from sunaengine_amd import scenes, tables
from tests import material_scenes as ms

ONE = 65536
F = mt.F
TOMB = 0x40000000
FLAG_DEAD, FLAG_ANCHOR, FLAG_BURST = 2, 4, 64
WELD_TAG_MASK = 0x000FFF00
HARD_BREAK2 = 5120 * 5120
BOND_R2 = 17623204
D2_FLOOR = 262144
DOM = 40
CENTER = (15 * ONE, 15 * ONE)


def _shr_rne(x, s):
    q, r = divmod(int(x), 1 << s)
    half = 1 << (s - 1)
    if r > half or (r == half and q & 1):
        q += 1
    return q


def _table(**extra):
    """The shipped rows plus named copies: name -> (base row, {field: value})."""
    rows = [list(r) for r in mt.MATS]
    ids = {}
    for name, (base, fields) in extra.items():
        r = list(mt.MATS[base])
        for k, v in fields.items():
            r[F[k]] = v
        rows.append(r)
        ids[name] = len(rows) - 1
    return rows, ids


def _world(parts, rows=None, flags=None, temps=None, vel=None, iters=4):
    from oracle import OracleWorld
    s = ms._scene(parts, DOM, DOM, "pin", temps=temps, flags=flags)
    if vel is not None:
        for i, (vx, vy) in enumerate(vel):
            s.vel[2 * i], s.vel[2 * i + 1] = vx, vy
    im = scenes.load_image(s, mats=rows)
    return OracleWorld(im, tables.load_luts(), mt.mats_image(rows), iters=iters), s


def _bonded(mat_i, mat_j, rows=None, extra=(), gap=(ONE * 3) // 5):
    """Two particles welded on substep 0 (`gap` ticks apart, away from
    `extra`, which are anchored).  Returns the world, its state and the rest
    length."""
    parts = [(30 * ONE, 12 * ONE, mat_i), (30 * ONE + gap, 12 * ONE, mat_j)] + list(extra)
    flags = [0, 0] + [FLAG_ANCHOR] * len(extra)
    o, s = _world(parts, rows, flags=flags)
    o.step(1)
    b = o.read_bond()
    assert b[0, 0] & 0x1FFFF == 1 and b[1, 0] & 0x1FFFF == 0 and b[0, 1] == 0, "pair did not weld"
    L0 = (int(b[0, 0]) >> 17) & 0x1FFF
    assert (int(b[1, 0]) >> 17) & 0x1FFF == L0
    return o, o.read_state(), L0


def _stretch(o, st, d4, at=CENTER, edit=None):
    """Write the pair anchored at `at` and `at` + 32 * d4, run one substep,
    return (row i word, row j word, derived)."""
    st2 = st.copy()
    st2[0, 0], st2[0, 1] = at
    st2[1, 0], st2[1, 1] = at[0] + 32 * d4[0], at[1] + 32 * d4[1]
    st2[:2, 2:4] = 0
    st2[:2, 5] |= FLAG_ANCHOR
    if edit is not None:
        edit(st2)
    o.write_state_derived(st2)
    o.step(1)
    b = o.read_bond()
    return int(b[0, 0]), int(b[1, 0]), o.read_derived()


def _two_squares(n):
    """(a, b) with a^2 + b^2 == n and b <= a as large as possible (the most
    diagonal offset), or None."""
    for b in range(math.isqrt(n // 2), -1, -1):
        r = n - b * b
        a = math.isqrt(r)
        if a * a == r:
            return (a, b)
    return None


def _broken(w):
    return (w & TOMB) != 0


def test_ratio_break_at_the_closed_form_bound():
    rows, ids = _table(R147=(mt.RUBBER, {"bondBreak": 147456}), R102=(mt.RUBBER, {"bondBreak": 102400}))
    cases = [(mt.RUBBER, mt.RUBBER), (ids["R147"], ids["R147"]), (mt.RUBBER, ids["R147"]),
             (ids["R147"], ids["R102"])]
    for mi, mj in cases:
        mb = min(rows[mi][F["bondBreak"]], rows[mj][F["bondBreak"]])
        # a formation gap whose bound is itself a sum of two squares, so the
        # surviving pair sits exactly on it
        for k in range(24):
            gap = (ONE * 3) // 5 + 32 * k
            o, st, L0 = _bonded(mi, mj, rows, gap=gap)
            T = _shr_rne(L0 * L0 * mb, 16)        # fp_mul_s(L0^2, mb, 16)
            if _two_squares(T) is not None:
                break
        else:
            raise AssertionError("no formation gap gives a representable bound")
        assert T + 2 < HARD_BREAK2
        above = next(m for m in range(T + 1, T + 400) if _two_squares(m) is not None)
        a = math.isqrt(T)
        keeps = [_two_squares(T), (a, 0)]
        breaks = [_two_squares(above), (a + 1, 0)]
        assert keeps[0][1] > 0 or breaks[0][1] > 0        # one diagonal offset at least
        for d4 in keeps:                                   # one world: the bond survives
            wi, wj, _ = _stretch(o, st, d4)
            assert not _broken(wi) and not _broken(wj), (mi, mj, L0, T, d4)
            st = o.read_state()
        for d4 in breaks:
            o, st, L0b = _bonded(mi, mj, rows, gap=gap)
            assert L0b == L0
            wi, wj, _ = _stretch(o, st, d4)
            assert _broken(wi) and _broken(wj), (mi, mj, L0, T, d4)
            # the tombstone keeps the word: partner and rest length
            assert wi & ~TOMB == (1 | (L0 << 17))


def test_unbreakable_row_breaks_at_the_kernel_radius():
    assert mt.MATS[mt.STONE][F["bondBreak"]] == 0
    for d4, broken in [((4095, 3072), False), ((5119, 0), False), ((3072, 4096), True),
                       ((5120, 0), True), ((5121, 0), True), ((0, -5121), True)]:
        o, st, L0 = _bonded(mt.STONE, mt.STONE)
        assert d4[0] ** 2 + d4[1] ** 2 > 16 * L0 * L0 or broken
        wi, wj, _ = _stretch(o, st, d4)
        assert _broken(wi) == broken and _broken(wj) == broken, (d4, L0)


def test_partner_state_vetoes():
    rows, ids = _table(COLD=(mt.RUBBER, {"rampLo": mt.DEG(30), "rampHi": mt.DEG(40), "rampSign": -1}))
    # FLAG_BURST on j: i's word is tombstoned, j's row is not gathered
    o, st, L0 = _bonded(mt.RUBBER, mt.RUBBER, rows)

    def burst(s):
        s[1, 5] |= FLAG_BURST
    wi, wj, _ = _stretch(o, st, (L0, 0), edit=burst)
    assert _broken(wi) and not _broken(wj)
    # j moved to a bondK == 0 row (WATER): the same
    o, st, L0 = _bonded(mt.RUBBER, mt.RUBBER, rows)

    def to_water(s):
        s[1, 4] = mt.WATER
    wi, wj, _ = _stretch(o, st, (L0, 0), edit=to_water)
    assert _broken(wi) and not _broken(wj)
    # a partner whose ramp gives kEff == 0 (COLD at 60 C) keeps the bond,
    # also stretched (no spring force, still inside the bound)
    o, st, L0 = _bonded(ids["COLD"], ids["COLD"], rows)

    def heat(s):
        s[1, 6] = mt.DEG(60)
    wi, wj, _ = _stretch(o, st, ((L0 * 3) // 2, 0), edit=heat)
    assert not _broken(wi) and not _broken(wj)
    assert o.read_state()[1, 6] > mt.DEG(45)


def _pressure_cluster(pitch=0.35):
    """Anchored WATER on a lattice around CENTER, clear of the pair's
    positions (CENTER and up to 1.7 wu to its right)."""
    p = int(pitch * ONE)
    out = []
    for gy in range(-7, 8):
        for gx in range(-7, 12):
            x, y = CENTER[0] + gx * p + p // 2, CENTER[1] + gy * p + p // 3
            if abs(y - CENTER[1]) < int(0.4 * ONE) and CENTER[0] - int(0.4 * ONE) < x < CENTER[0] + int(2.1 * ONE):
                continue
            out.append((x, y, mt.WATER))
    return out


def test_pressure_raises_the_break_bound():
    rows, ids = _table(MU=(mt.RUBBER, {"bondBreak": 147456, "bondMu": 2 * ONE}),
                       M0=(mt.RUBBER, {"bondBreak": 147456, "bondMu": 0}))
    cluster = _pressure_cluster()
    mb = 147456

    def bound(L0, lam_i, lam_j, mub):
        m = mb
        if mub:
            lp = -(lam_i + lam_j)
            pq = _shr_rne(lp, 4) if lp > 0 else 0
            m = mb + _shr_rne(pq * mub, 16)
        return (L0 * L0 * m) >> 16

    o, st, L0 = _bonded(ids["MU"], ids["MU"], rows, cluster)
    t0 = (L0 * L0 * mb) >> 16
    # a first pass at the no-pressure bound measures lam at this spot
    _, _, dv = _stretch(o, st, (math.isqrt(t0), 0))
    lam_i, lam_j = int(dv[0, 11]), int(dv[1, 11])
    assert lam_i < -20000 and lam_j < -20000, (lam_i, lam_j)
    t1 = bound(L0, lam_i, lam_j, 2 * ONE)
    assert t1 > t0 + t0 // 20
    # lam moves with j's position, so each run is checked against the bound
    # of its own lam; the break case steps past the bound it last saw
    survived_by_pressure = broke_past_it = 0
    d = math.isqrt((t0 + t1) // 2)
    for _ in range(6):
        o, st, _ = _bonded(ids["MU"], ids["MU"], rows, cluster)
        wi, wj, dv = _stretch(o, st, (d, 0))
        li, lj = int(dv[0, 11]), int(dv[1, 11])
        tb = bound(L0, li, lj, 2 * ONE)
        assert d * d <= tb or d * d >= tb + 2          # off the rounding unit
        assert _broken(wi) == (d * d > tb) and _broken(wj) == (d * d > tb), (d, tb, li, lj)
        survived_by_pressure += int(t0 < d * d <= tb)
        broke_past_it += int(d * d > tb)
        # the same spot with bondMu == 0 on one row: the bound is t0
        o, st, _ = _bonded(ids["MU"], ids["M0"], rows, cluster)
        wi, wj, dv = _stretch(o, st, (d, 0))
        assert _broken(wi) and _broken(wj), (d, t0)
        if broke_past_it:
            break
        d = math.isqrt(tb) + 2
    assert survived_by_pressure >= 1 and broke_past_it == 1


def _rice_scene(seed=0x41CE, vj=1500):
    """RICE packed at 0.45 wu beside a block at 1 wu, temperatures 20 / 60 /
    120 / 75 C (20 C is below the ramp: kEff 0) and random velocities."""
    rnd = scenes.xorshift32(seed)
    pitch = int(0.45 * ONE)
    parts = ms._block(rnd, mt.RICE, 14, 14, 6 * ONE, 33 * ONE, pitch, jitter=int(0.1 * ONE))
    parts += ms._block(rnd, mt.RICE, 9, 9, 6 * ONE + 14 * pitch, 33 * ONE, ONE, jitter=int(0.1 * ONE))
    temps = [mt.DEG([20, 60, 120, 75][k % 4]) for k in range(len(parts))]
    vel = [((rnd() % (2 * vj)) - vj, (rnd() % (2 * vj)) - vj) for _ in parts]
    return parts, temps, vel


def test_reform_gates_match_brute_force():
    from oracle.oracle import nbr_probe
    parts, temps, vel = _rice_scene()
    probe_world, s = _world(parts, temps=temps, vel=vel)
    pr = nbr_probe(probe_world)
    o, _ = _world(parts, temps=temps, vel=vel)
    o.step(1)
    st, dv, bond = o.read_state(), o.read_derived(), o.read_bond()
    nf = s.nFluid
    R = mt.MATS[mt.RICE]
    assert R[F["bondReform"]] != 0 and not (st[:nf, 5] & WELD_TAG_MASK).any()
    lo = R[F["rampLo"]]
    assert (np.abs(st[:nf, 6].astype(np.int64) - lo) > ONE).all()     # clear of the ramp's low end
    k_on = st[:nf, 6] > lo
    vl4 = _shr_rne(R[F["bondFormV"]], 3)
    pl = R[F["bondFormP"]]
    rejected = dict(kEff=0, velocity=0, pressure=0, range=0)
    cand = []
    for i in range(nf):
        c = {}
        if k_on[i]:
            cnt = int(pr["nbrN"][i]) & 0x1FFFFFFF
            for j in (int(x) for x in pr["nbr"][i, :cnt]):
                if j >= nf:
                    continue
                d4x = _shr_rne(int(st[i, 0]) - int(st[j, 0]), 5)
                d4y = _shr_rne(int(st[i, 1]) - int(st[j, 1]), 5)
                d2 = d4x * d4x + d4y * d4y
                if abs(d4x) > 5120 or abs(d4y) > 5120 or not D2_FLOOR <= d2 <= BOND_R2:
                    rejected["range"] += 1
                    continue
                if not k_on[j]:
                    rejected["kEff"] += 1
                    continue
                r4x = _shr_rne(int(st[i, 2]) - int(st[j, 2]), 3)
                r4y = _shr_rne(int(st[i, 3]) - int(st[j, 3]), 3)
                if abs(r4x) > vl4 or abs(r4y) > vl4 or r4x * r4x + r4y * r4y > vl4 * vl4:
                    rejected["velocity"] += 1
                    continue
                if -int(dv[i, 11]) < pl or -int(dv[j, 11]) < pl:
                    rejected["pressure"] += 1
                    continue
                L0 = math.isqrt(d2)
                if 512 <= L0 <= 5120:
                    c[j] = L0
        assert len(c) < 48
        cand.append(c)
    want = {(i, j, L0) for i in range(nf) for j, L0 in cand[i].items() if cand[j].get(i) == L0}
    got = set()
    for i in range(nf):
        for w in bond[i]:
            if w == 0:
                break
            got.add((i, int(w) & 0x1FFFF, (int(w) >> 17) & 0x1FFF))
    assert all(v >= 40 for v in rejected.values()), rejected
    assert len(want) >= 40
    assert got == want


def _stress_case(offsets, extra=(), stone_anchored=True):
    """An anchored STONE at CENTER, anchored WATER at CENTER + 32 * offset,
    and `extra` (dx4, dy4, mat) also anchored.  A phantom row is written
    back without the FLAG_DEAD the loader gives it, so that buildNbr lists it
    and containStress itself has to skip it.  Returns the stone's flags
    before and after one substep."""
    parts = [(CENTER[0], CENTER[1], mt.STONE)]
    parts += [(CENTER[0] + 32 * a, CENTER[1] + 32 * b, mt.WATER) for a, b in offsets]
    parts += [(CENTER[0] + 32 * a, CENTER[1] + 32 * b, m) for a, b, m in extra]
    flags = [FLAG_ANCHOR if stone_anchored else 0] + [FLAG_ANCHOR] * (len(parts) - 1)
    o, s = _world(parts, flags=flags)
    st0 = o.read_state()
    st0[:, 5] &= ~FLAG_DEAD
    o.write_state_derived(st0)
    before = int(st0[0, 5])
    o.step(1)
    st = o.read_state()
    moved = st[1:len(parts), :2] - np.array([[p[0], p[1]] for p in parts[1:]])
    assert not moved.any()
    return before, int(st[0, 5])


def test_contain_stress_closed_form():
    burst_p = mt.MATS[mt.STONE][F["burstP"]]
    assert burst_p == 163840          # stress > 5120 bursts

    def stress(offsets):
        return sum(2560 - math.isqrt(a * a + b * b) for a, b in offsets if a * a + b * b < 2560 * 2560)

    four = lambda r: [(r, 0), (-r, 0), (0, r), (0, -r)]
    # 2 * 907^2 = 1,645,298: sqrt 1282.69, isqrt 1282; three at 1279 add
    # 3 * 1281, so the floor root gives 5121 (bursts) where a rounded root
    # would give 5120 (does not)
    diag = [(1279, 0), (-1279, 0), (0, 1279), (907, -907)]
    cases = [
        (four(1280), (), True, False),           # 5120 exactly: not above
        (four(1279), (), True, True),            # 5124
        (diag, (), True, True),                  # 5121 with the floor root
        ([(1280, 0), (-1280, 0), (0, 1280), (900, -903)], (), True, True),     # 1280*3 + 1286 = 5126
        ([(1281, 0), (-1281, 0), (0, 1281), (903, -903)], (), True, False),   # 1279*3 + 1283 = 5120
        (four(1279), (), False, False),          # not anchored: never bursts
        # a seal-class neighbour (STONE) and a phantom row do not count
        (four(1280), ((0, 900, mt.STONE), (700, 700, 24)), True, False),
        # a boundary-free non-seal row (SAND) counts like WATER
        (four(1280), ((-700, -700, mt.SAND),), True, True),
    ]
    assert mt.MATS[24][F["phantom"]] != 0 and not mt.seal_row(mt.MATS[mt.SAND])
    for offsets, extra, anchored, want in cases:
        counted = list(offsets) + [(a, b) for a, b, m in extra if m == mt.SAND]
        s = stress(counted)
        assert (32 * s > burst_p) == want or not anchored, (offsets, s)
        before, after = _stress_case(offsets, extra, anchored)
        assert bool(after & FLAG_BURST) == want, (offsets, extra, anchored, s)
        if want:
            assert not after & (FLAG_ANCHOR | WELD_TAG_MASK) and before & FLAG_ANCHOR
        elif anchored:
            assert after & FLAG_ANCHOR
