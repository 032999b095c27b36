"""The oracle's bond spring force, corrosion dose and split products, pinned
by closed forms (src/sim.wgsl:1663-1671, 1795-1934, 2530-2707, 573-593).

* Bond spring.  The bond gather is a position-based distance constraint:
  with C = |d|^2 - L0^2 its projection moves particle i by
  -k * C / (4 |d|^2) * d (half of the Newton step along grad C = 2d, the
  other half going to j), in ticks, with d = pred_i - pred_j.  A welded
  RUBBER pair (fric = 0, so the friction pass carries the bond term alone)
  is written back anchored (pred == pos through the colour passes) at a
  chosen offset, and the friction pass's dp (ticks << ACC) is read before it
  is applied.  dp / 2^ACC must equal that real-number projection to within
  the fixed-point rounding (2^-12 relative plus a few ticks): a wrong shift
  is a factor of two or more, a wrong sign points the wrong way.  Also:
  dp_j == -dp_i exactly (every step is sign-symmetric), the stiffness of a
  mixed pair is min(k_i, k_j) (an average would be 2.5x here), pairs closer
  than the D2 floor carry no force, and the temperature ramp scales k
  linearly between rampLo and rampHi (0 below, the table value above) while
  the bond survives, with the cooler end deciding from either side.  (The
  clamp of the ramp fraction to [0, 1] is inert for every loadable row: the
  table bake derives rampInv from the span.)
* Corrosion.  An anchored ACID beside an anchored target at offset d gives
  the target the dose RNE(clamp(RNE(W(d) * str / 2^12), 0, 2^24) / 2^8),
  str = ACID.corrode (+ the target's own corrode when ACID can eat it), and
  the target becomes corrTo on substep 0 iff dose >= corrPt: the test sets
  corrPt to the dose and to the dose + 1.  The attacker spends its own
  strength only against a target whose corrPt is reachable; same-material
  pairs take no dose; corrosion outranks boiling in the transition chain,
  and a transmuted anchored slot loses FLAG_ANCHOR.
* Split products.  pick_product takes splitTo for slot i iff
  (h(i) & 255) < splitPct, where h is the 2-round multiply-xorshift hash
  (shifts 16/15/15, multipliers 0x21f0aaad, 0x735a2d97; C. Wellons'
  hash-prospector) of i + 0x9e3779b9.  Closed forms: splitPct 0 never
  splits, 256 always does; the split set grows with splitPct (nested); and
  a 2,000-slot block transmuting at once splits exactly at the slots the
  published hash picks, at a rate within 4 sigma of splitPct / 256.
"""
import math

import numpy as np

from sunaengine_amd import materials as mt
from sunaengine_amd import scenes, tables
from tests import material_scenes as ms
from tests.test_oracle_bond_pins import CENTER, FLAG_ANCHOR, ONE, _bonded, _table, _world

F = mt.F
ACC = 8
D2_FLOOR = 262144
TN = 3202
HS, H2S = 10240, 104857600
CORR_NEVER = 2147483647


def _rne(num, s):
    q, r = divmod(int(num), 1 << s)
    half = 1 << (s - 1)
    if r > half or (r == half and q & 1):
        q += 1
    return q


def _spring(mi, mj, rows, d4, temps=None):
    """The friction-pass dp of a welded (mi, mj) pair written back anchored
    at CENTER and CENTER + 32 * d4 (d4 in 1/32-tick units)."""
    from oracle.oracle import friction_dp
    o, st, L0 = _bonded(mi, mj, rows)
    st2 = st.copy()
    st2[0, 0], st2[0, 1] = CENTER
    st2[1, 0], st2[1, 1] = CENTER[0] + 32 * d4[0], CENTER[1] + 32 * d4[1]
    st2[:2, 2:4] = 0
    st2[:2, 5] |= FLAG_ANCHOR
    if temps is not None:
        st2[0, 6], st2[1, 6] = temps
    o.write_state_derived(st2)
    r = friction_dp(o, 2)
    d = r["pred"][0].astype(np.int64) - r["pred"][1].astype(np.int64)
    assert tuple(d) == (-32 * d4[0], -32 * d4[1]), "anchored preds moved"
    o.close()
    return r["dp"].astype(np.int64), L0


def _projection(d4, L0, k):
    """-k C / (4 |d|^2) d in ticks, C = |d|^2 - L0^2 (d, L0 in 1/32 ticks),
    for particle i at d = pred_i - pred_j = -d4."""
    dx, dy = -d4[0], -d4[1]
    D2 = dx * dx + dy * dy
    s = -(k / 65536.0) * (D2 - L0 * L0) / (4.0 * D2)
    return np.array([s * dx * 32.0, s * dy * 32.0])


def _close(dp_i, want):
    got = dp_i / float(1 << ACC)
    tol = np.abs(want) * 2.0 ** -12 + 8.0
    return bool(np.all(np.abs(got - want) <= tol)), got


def test_bond_spring_is_the_distance_projection():
    rows, ids = _table(STIFF=(mt.RUBBER, {"bondK": 65536}))
    k = mt.MATS[mt.RUBBER][F["bondK"]]
    assert k == 16384 and mt.MATS[mt.RUBBER][F["fric"]] == 0 and mt.MATS[mt.RUBBER][F["rampSign"]] == 0
    for stretch in (1.5, 1.1, 0.8, 0.6):
        for direction in ((1.0, 0.0), (0.6, 0.8), (-0.28, 0.96)):
            o, st, L0 = _bonded(mt.RUBBER, mt.RUBBER, rows)
            o.close()
            r = stretch * L0
            d4 = (int(round(r * direction[0])), int(round(r * direction[1])))
            assert d4[0] ** 2 + d4[1] ** 2 >= D2_FLOOR
            dp, L0b = _spring(mt.RUBBER, mt.RUBBER, rows, d4)
            assert L0b == L0
            want = _projection(d4, L0, k)
            ok, got = _close(dp[0], want)
            assert ok, (stretch, direction, got, want)
            assert np.abs(want).max() > 200.0          # far above the tolerance
            assert dp[1][0] == -dp[0][0] and dp[1][1] == -dp[0][1]
            # stretched pairs pull i toward j (d = pred_i - pred_j = -d4), compressed ones push
            dot = float(dp[0][0]) * d4[0] + float(dp[0][1]) * d4[1]
            assert (dot > 0) == (stretch > 1.0)
    # a mixed pair takes the smaller stiffness
    o, st, L0 = _bonded(mt.RUBBER, ids["STIFF"], rows)
    o.close()
    d4 = (int(1.4 * L0), 0)
    dp, _ = _spring(mt.RUBBER, ids["STIFF"], rows, d4)
    ok, got = _close(dp[0], _projection(d4, L0, k))
    assert ok, got
    dp2, _ = _spring(ids["STIFF"], ids["STIFF"], rows, d4)
    ok, got = _close(dp2[0], _projection(d4, L0, 65536))
    assert ok, got


def test_bond_spring_floor_and_ramp():
    span = mt.DEG(10)
    rows, ids = _table(WARM=(mt.RUBBER, {"rampLo": mt.DEG(30), "rampHi": mt.DEG(40), "rampSign": 1,
                                         "rampInv": (1 << 32) // span}))
    k = rows[ids["WARM"]][F["bondK"]]
    # closer than the D2 floor: no force, bond kept
    o, st, L0 = _bonded(mt.RUBBER, mt.RUBBER, rows)
    o.close()
    dp, _ = _spring(mt.RUBBER, mt.RUBBER, rows, (500, 100))
    assert 500 ** 2 + 100 ** 2 < D2_FLOOR and not dp.any()
    # the ramp: 0 below rampLo, linear inside, the table value above
    o, st, L0 = _bonded(ids["WARM"], ids["WARM"], rows)
    o.close()
    d4 = (int(1.4 * L0), int(0.3 * L0))
    for T, frac in ((mt.DEG(20), 0.0), (mt.DEG(30), 0.0), (mt.DEG(33), 0.3), (mt.DEG(35), 0.5),
                    (mt.DEG(38), 0.8), (mt.DEG(40), 1.0), (mt.DEG(60), 1.0)):
        dp, _ = _spring(ids["WARM"], ids["WARM"], rows, d4, temps=(T, T))
        want = _projection(d4, L0, k * frac)
        ok, got = _close(dp[0], want)
        assert ok, (T, frac, got, want)
        if frac == 0.0:
            assert not dp.any()
    # different temperatures: the cooler end decides (min of the two kEff),
    # from either side
    for temps in ((mt.DEG(33), mt.DEG(39)), (mt.DEG(39), mt.DEG(33))):
        dp, _ = _spring(ids["WARM"], ids["WARM"], rows, d4, temps=temps)
        ok, got = _close(dp[0], _projection(d4, L0, k * 0.3))
        assert ok, (temps, got)


# ---------------------------------------------------------------- corrosion
def _w_at(dx, dy):
    """The frozen W table at the lut_index of offset (dx, dy) ticks."""
    ax, ay = (abs(dx) + 8) >> 4, (abs(dy) + 8) >> 4
    r2 = ax * ax + ay * ay
    assert ax <= HS and ay <= HS and r2 < H2S
    return int(tables.load_luts()[r2 >> 15])


def _dose(W, strength):
    c = min(max(_rne(W * strength, 20 - ACC), 0), 1 << 24)
    return _rne(c, ACC)


def _pair_after(mat_a, mat_t, rows, dx, temps=None):
    """One substep of an anchored (attacker, target) pair dx ticks apart;
    returns the state."""
    parts = [(CENTER[0], CENTER[1], mat_a), (CENTER[0] + dx, CENTER[1], mat_t)]
    o, s = _world(parts, rows, flags=[FLAG_ANCHOR, FLAG_ANCHOR], temps=temps)
    o.step(1)
    st = o.read_state()
    o.close()
    return st


def test_corrosion_threshold_is_the_closed_form():
    acid = mt.MATS[mt.ACID]
    assert acid[F["corrode"]] == 65536
    for dx in (int(0.45 * ONE), int(0.9 * ONE), int(1.7 * ONE)):
        W = _w_at(dx, 0)
        dose = _dose(W, acid[F["corrode"]])        # the target is not corrosive: no spend term
        assert 50 < dose < 20000
        for pt, eaten in ((dose, True), (dose + 1, False), (dose - 1, True)):
            rows, ids = _table(T=(mt.STONE, {"corrPt": pt, "corrTo": mt.SAND}))
            st = _pair_after(mt.ACID, ids["T"], rows, dx)
            assert (st[1, 4] == mt.SAND) == eaten, (dx, W, dose, pt)
            if eaten:
                assert st[1, 5] & FLAG_ANCHOR == 0    # transmuted matter is no longer held
            else:
                assert st[1, 4] == ids["T"] and st[1, 5] & FLAG_ANCHOR


def test_corrosion_attack_spend_and_same_material():
    dx = int(0.8 * ONE)
    W = _w_at(dx, 0)
    acid = mt.MATS[mt.ACID]
    # the acid spends its own strength against something it can eat
    own = _dose(W, acid[F["corrode"]])
    rows, ids = _table(A=(mt.ACID, {"corrPt": own, "corrTo": mt.WATER}),
                       T=(mt.STONE, {"corrPt": 1 << 30}),
                       IMM=(mt.STONE, {"corrPt": CORR_NEVER}))
    st = _pair_after(ids["A"], ids["T"], rows, dx)
    assert st[0, 4] == mt.WATER and st[1, 4] == ids["T"]
    # ... and nothing against an immune target
    st = _pair_after(ids["A"], ids["IMM"], rows, dx)
    assert st[0, 4] == ids["A"]
    # a corrosive target adds its own strength to the attack (str = attack + spend)
    rows, ids = _table(A=(mt.ACID, {"corrPt": 1 << 30}),
                       T=(mt.STONE, {"corrode": 30000, "corrPt": _dose(W, 65536 + 30000), "corrTo": mt.SAND}),
                       T1=(mt.STONE, {"corrode": 30000, "corrPt": _dose(W, 65536 + 30000) + 1, "corrTo": mt.SAND}))
    assert _dose(W, 65536 + 30000) > _dose(W, 65536)
    assert _pair_after(ids["A"], ids["T"], rows, dx)[1, 4] == mt.SAND
    assert _pair_after(ids["A"], ids["T1"], rows, dx)[1, 4] == ids["T1"]
    # two slots of one material take no dose from each other
    rows, ids = _table(A=(mt.ACID, {"corrPt": 0, "corrTo": mt.WATER}))
    st = _pair_after(ids["A"], ids["A"], rows, dx)
    # corrPt = 0 is reached by a zero dose: both transmute with no pair term at all
    assert st[0, 4] == mt.WATER and st[1, 4] == mt.WATER
    rows, ids = _table(A=(mt.ACID, {"corrPt": 1}))
    st = _pair_after(ids["A"], ids["A"], rows, dx)
    assert st[0, 4] == ids["A"] and st[1, 4] == ids["A"]


def test_corrosion_outranks_boiling():
    dx = int(0.6 * ONE)
    dose = _dose(_w_at(dx, 0), 65536)
    rows, ids = _table(T=(mt.WATER, {"corrPt": dose, "corrTo": mt.SAND}))
    hot = mt.DEG(150)
    st = _pair_after(mt.ACID, ids["T"], rows, dx, temps=[mt.DEG(20), hot])
    assert st[1, 4] == mt.SAND                       # not STEAM: corrode comes first
    rows, ids = _table(T=(mt.WATER, {"corrPt": dose + 1, "corrTo": mt.SAND}))
    st = _pair_after(mt.ACID, ids["T"], rows, dx, temps=[mt.DEG(20), hot])
    assert st[1, 4] == mt.MATS[mt.WATER][F["boilTo"]]


# ------------------------------------------------------------ split products
def _published_hash(i):
    """C. Wellons' 2-round multiply-xorshift (hash-prospector: 16, 0x21f0aaad,
    15, 0x735a2d97, 15) of i + 0x9e3779b9, over u32."""
    M = 0xFFFFFFFF
    x = (i + 0x9E3779B9) & M
    x = ((x ^ (x >> 16)) * 0x21F0AAAD) & M
    x = ((x ^ (x >> 15)) * 0x735A2D97) & M
    return x ^ (x >> 15)


def _split_block(pct, n_side=40):
    """n_side^2 slots of a row with corrPt = 0 (transmutes on substep 0 with
    no neighbour) into SAND, splitting into ASH at splitPct = pct; returns
    the slots that became ASH."""
    from oracle import OracleWorld
    rows, ids = _table(T=(mt.STONE, {"corrPt": 0, "corrTo": mt.SAND, "splitTo": mt.ASH, "splitPct": pct}))
    rnd = scenes.xorshift32(0x5B117)
    parts = ms._block(rnd, ids["T"], n_side, n_side, 4 * ONE, 40 * ONE, ONE, jitter=256)
    s = ms._scene(parts, 48, 48, "split")
    im = scenes.load_image(s, mats=rows)
    o = OracleWorld(im, tables.load_luts(), mt.mats_image(rows), iters=1)
    o.step(1)
    st = o.read_state()[: n_side * n_side]
    o.close()
    assert set(np.unique(st[:, 4])) <= {mt.SAND, mt.ASH}
    return set(np.nonzero(st[:, 4] == mt.ASH)[0].tolist())


def test_split_products_follow_the_published_hash():
    n = 40 * 40
    assert _split_block(0) == set()
    assert _split_block(256) == set(range(n))
    prev = set()
    for pct in (26, 100, 200):
        got = _split_block(pct)
        want = {i for i in range(n) if (_published_hash(i) & 255) < pct}
        assert got == want, pct
        assert prev <= got
        prev = got
        p = pct / 256.0
        assert abs(len(got) - n * p) <= 4 * math.sqrt(n * p * (1 - p))
