"""GPU parity of the material passes: friction and the bond gather,
containment (contain / containOverflow), containStress, markCongealNbr and
bond formation (src/sim.wgsl:1673-2235, 2708-3097), with the state hash
folding the bond rows in.

Every scene of tests/material_scenes.py runs on the GPU and on the oracle:
the per-substep hash stream, the chain, the state, the bond rows and the
debug counters must be identical.  The same runs check on the oracle's
output that the scene reaches the path it is there for.
"""
import numpy as np
import pytest

from tests import material_scenes as ms
from tests.helpers import first_mismatch, golden, oracle_world

pytestmark = pytest.mark.gpu

CHUNKS = 4
TOMB = 0x40000000


def _engine(scene, materials, debug=True):
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    im = scenes.load_image(scene)
    e = Engine(max_particles=im.n, max_cells=im.cellTotal, iters=4, materials=materials, debug=debug)
    e.load_scene(scene)
    return e


@pytest.mark.parametrize("name", list(ms.CASES))
def test_material_scene_matches_oracle(name):
    build, steps, materials = ms.CASES[name]
    s = build()
    print(f"{name}: n={s.n} nFluid={s.nFluid} x{steps} materials={materials}", flush=True)
    o = oracle_world(s, 4)
    e = _engine(s, materials)
    seen = dict(bonds=0, tombs=0, burst=0, congeal=0)
    per = steps // CHUNKS
    for c in range(CHUNKS):
        ref = o.step(per, stream=True)
        e.step(per)
        got = e.hash_stream()[-per:]
        bad = first_mismatch(got, ref)
        assert bad is None, f"{name}: first differing substep {c * per + bad}"
        ob = o.read_bond()
        assert np.array_equal(e.read_bond(), ob), f"{name}: bond rows differ after {(c + 1) * per} substeps"
        st = o.read_state()
        assert np.array_equal(e.read_state(), st), f"{name}: state differs after {(c + 1) * per} substeps"
        seen["bonds"] = max(seen["bonds"], int((ob != 0).sum()))
        seen["tombs"] += int(((ob & TOMB) != 0).sum())
        seen["burst"] += int(((st[:, 5] & 64) != 0).sum())
        seen["congeal"] += int(((st[:, 5] & 32) != 0).sum())
    assert np.array_equal(e.chain(), o.chain())
    assert np.array_equal(e.debug_counters(), o.debug_counters())
    print(f"{name}: {seen} dbg={list(o.debug_counters())}")
    # the path each scene is for
    if name in ("stone_in_water", "rice_on_rubber", "hot_metal", "freezing_pool"):
        assert seen["bonds"] > 0
    if name in ("stone_in_water", "rice_on_rubber", "freezing_pool"):
        assert seen["tombs"] > 0          # bonds broke (stretch, melted partner) between formations
    if name == "freezing_pool":
        assert seen["congeal"] > 0
    if name in ("stone_in_water", "packed_against_stone"):
        assert seen["burst"] > 0
    if name == "packed_against_stone":
        assert o.debug_counters()[6] > 0  # full lists: containOverflow gathers the seals itself


@pytest.mark.parametrize("name", ["stone_in_water", "seal_crossing"])
def test_material_scene_in_an_unbounded_domain(name):
    """The same scenes in a domain 2^29 ticks wide: friction and the bond
    gather take the saturating pair forms, and contain runs the reference's
    full pair loop (its bounded-scene box filter does not apply)."""
    build, steps, materials = ms.CASES[name]
    s = build()
    s.domW = 1 << 29
    o = oracle_world(s, 4)
    e = _engine(s, materials)
    ref = o.step(steps, stream=True)
    e.step(steps)
    bad = first_mismatch(e.hash_stream(), ref)
    assert bad is None, f"{name} (unbounded): first differing substep {bad}"
    assert np.array_equal(e.read_state(), o.read_state())
    assert np.array_equal(e.read_bond(), o.read_bond())
    assert np.array_equal(e.debug_counters(), o.debug_counters())
    e.destroy()


def test_spawn_in_a_material_world():
    """spawnFluid in a bonded world: the existing fluid slots keep their bond
    rows across the reload, the new slots start empty, and formation runs
    the next substep (src/engine.js:5901-6020)."""
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    from sunaengine_amd import materials as mt
    s = ms.stone_in_water()
    im = scenes.load_image(s)
    cap = im.n + 64
    o = oracle_world(s, 4, max_particles=cap)
    e = Engine(max_particles=cap, max_cells=im.cellTotal, iters=4)
    e.load_scene(s)
    o.step(9)
    e.step(9)
    assert (e.read_bond() != 0).any()
    drop = [dict(x=(14 + dx) * 65536, y=(12 + dy) * 65536, vx=0, vy=2000, mat=mt.SAND if dx < 3 else mt.WATER)
            for dy in range(2) for dx in range(6)]
    fl, bd = scenes.spawn_records(s, drop)
    assert o.spawn(fl, bd) == e.spawn_fluid(drop)
    done = 0
    for k in (1, 5, 20):
        o.step(k - done)
        e.step(k - done)
        done = k
        assert np.array_equal(e.read_state(), o.read_state()), f"state after {k}"
        assert np.array_equal(e.read_bond(), o.read_bond()), f"bonds after {k}"
        assert np.array_equal(e.digest()["acc"], o.digest()), f"digest after {k}"
    assert np.array_equal(e.chain(), o.chain())


def test_material_scenes_batched_equal_alone():
    """Three material scenes in one world (their bond rows and slot maps side
    by side): each scene's per-substep stream and bond rows equal its own
    oracle run."""
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    ss = [ms.stone_in_water(), ms.rice_on_rubber(), ms.freezing_pool()]
    ims = [scenes.load_image(s) for s in ss]
    e = Engine(max_particles=sum(i.n for i in ims), max_cells=sum(i.cellTotal for i in ims), iters=4,
               max_scenes=len(ss))
    e.load_scenes(ss)
    e.step(24)
    for k, s in enumerate(ss):
        o = oracle_world(s, 4)
        ref = o.step(24, stream=True)
        bad = first_mismatch(e.hash_stream(k), ref)
        assert bad is None, f"{s.name}: first differing substep {bad}"
        assert np.array_equal(e.read_bond(k), o.read_bond()), s.name
    e.destroy()


def test_slotted_world_switches_to_material_passes():
    """A slotted world running a water job takes a bonded job into a free
    partition: the world switches to the material passes mid-run, and both
    jobs' streams equal their own oracle runs (the water job's included:
    the passes are bit-inert for it)."""
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    water = scenes.batch_scenes(count=1, seed=93, lo=900, hi=900)[0]
    stone = ms.stone_in_water()
    caps_n, caps_c = [2000, 2000], [2048, 2048]
    e = Engine(max_particles=sum(caps_n), max_cells=sum(caps_c), iters=4, max_scenes=2)
    e.reserve_slots(caps_n, caps_c)
    e.load_slot(0, water)
    e.step(10)
    e.load_slot(1, stone)
    e.step(20)
    for slot, s, steps in ((0, water, 30), (1, stone, 20)):
        o = oracle_world(s, 4)
        ref = o.step(steps, stream=True)
        bad = first_mismatch(e.hash_stream(slot), ref)
        assert bad is None, f"slot {slot}: first differing substep {bad}"
    assert np.array_equal(e.read_bond(1), o.read_bond())   # o: the stone job's oracle, 20 substeps
    e.destroy()


def test_materials_mix_at_scale_first_substeps():
    """The measured mixed-material scene itself (109,928 particles: SAND,
    METAL, held STONE, RICE over the bond ramp, water) against the oracle
    for its first 8 substeps: streams, state and bond rows."""
    from sunaengine_amd import scenes
    s = scenes.materials_mix()
    print(f"materials_mix n={s.n} x8", flush=True)
    o = oracle_world(s, 4)
    e = _engine(s, False)
    ref = o.step(8, stream=True)
    e.step(8)
    assert first_mismatch(e.hash_stream(), ref) is None
    assert np.array_equal(e.read_state(), o.read_state())
    assert np.array_equal(e.read_bond(), o.read_bond())
    e.destroy()


def test_stripped_build_on_material_scene():
    """debug:false runs the same material passes: identical stream and bonds."""
    s = ms.stone_in_water()
    a = _engine(s, False, debug=True)
    b = _engine(s, False, debug=False)
    a.step(24)
    b.step(24)
    assert np.array_equal(a.hash_stream(), b.hash_stream())
    assert np.array_equal(a.read_bond(), b.read_bond())
    assert not b.debug_counters().any()


def test_water_is_bit_inert_under_the_material_passes():
    """The frozen dambreak_v1 goldens hold in a world that runs every
    material pass (the passes return before writing for WATER,
    src/sim.wgsl:1698-1702,2062,2147)."""
    from sunaengine_amd import scenes
    g = golden("dambreak_v1.json")
    s = scenes.make_scene("dambreak_v1")
    e = _engine(s, True)
    done = 0
    for k in (1, 10, 100):
        e.step(k - done)
        done = k
        assert e.digest()["hex"] == g["checkpoints"][str(k)], f"substep {k}"


def test_material_mode_refusals():
    """What a world refuses, and how (the reference's assert and fault
    conventions): a rigid-body flag at load; a bonded scene above
    MAX_PARTICLES_BONDED (src/engine.js:261,3331); a water world created
    without materials that freezes into ICE reports SN_E_UNSUPPORTED (at the
    next call that checks the scene's fault word) instead of a wrong hash;
    created with materials it runs."""
    from sunaengine_amd import native, scenes
    from sunaengine_amd import materials as mt
    from sunaengine_amd.engine import Engine
    # FLAG_RIGID cannot come from a scene's flags (legal mask 0x7), so it is
    # put into the state image directly
    s = ms.stone_in_water()
    im = scenes.load_image(s)
    e = Engine(max_particles=im.n, max_cells=im.cellTotal, iters=4)
    im.state[0, 5] |= 16
    arr = (native.sn_scene * 1)()
    st = np.ascontiguousarray(im.state, np.int32)
    arr[0] = native.sn_scene(im.n, im.nFluid, st.ctypes.data, im.domW, im.domH, im.cellShift, None)
    rc = native.lib().sn_load_scenes(e._h, arr, 1)
    assert rc == native.SN_E_UNSUPPORTED and b"FLAG_RIGID" in native.lib().sn_last_error(e._h)
    e.destroy()
    # a bonded scene of 131,072 slots
    big = scenes.walled_block_scene(400, 326, 512, 256, 0xB16, pitch=65536 >> 1)
    big.matId[: big.nFluid] = mt.STONE
    bim = scenes.load_image(big)
    assert bim.n > 131071
    e = Engine(max_particles=bim.n, max_cells=bim.cellTotal, iters=4)
    with pytest.raises(native.SnError) as ex:
        e.load_scene(big)
    assert ex.value.code == native.SN_E_CAPACITY and "MAX_PARTICLES_BONDED" in str(ex.value)
    e.destroy()
    # freezing without and with the material passes
    cold = ms.freezing_pool()
    plain = _engine(cold, False)
    with pytest.raises(native.SnError) as ex:   # raised by step() or by the sync after it
        plain.step(4)
        plain.sync()
    assert ex.value.code == native.SN_E_UNSUPPORTED
    plain.destroy()
    full = _engine(cold, True)
    full.step(4)
    full.sync()
    full.destroy()



def test_spawn_weld_groups_and_boundary_rewrite():
    """spawnFluid's weld tags (src/engine.js:4570-4580,5933-5962): after a
    load the first spawn is group 2, the next group 3, each tag written only
    on rows that weld on spawn (STONE; not WATER) and not at all with
    noWeld; the boundary block moves up with its flags cleared.  The oracle
    is handed the records with the tags the reference assigns and the
    boundary block rebuilt from the scene; the engine makes both itself."""
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    from sunaengine_amd import materials as mt
    s = ms.stone_in_water()
    im = scenes.load_image(s)
    cap = im.n + 64
    o = oracle_world(s, 4, max_particles=cap)
    e = Engine(max_particles=cap, max_cells=im.cellTotal, iters=4)
    e.load_scene(s)
    o.step(5)
    e.step(5)
    tags = []
    for group, no_weld in ((2, False), (3, False), (None, True)):
        drop = [dict(x=(10 + dx) * 65536, y=(8 + dy) * 65536, vy=1500, mat=mt.STONE if dx < 4 else mt.WATER)
                for dy in range(2) for dx in range(6)]
        fl, bd = scenes.spawn_records(s, drop)
        for i, p in enumerate(drop):
            if group is not None and p["mat"] == mt.STONE:
                fl[i, 5] |= group << 8
        assert o.spawn(fl, bd) == e.spawn_fluid(drop, no_weld=no_weld)
        st = e.read_state()
        nf = o.nFluid
        tags.append(sorted({int(f) >> 8 & 0xFFF for f in st[nf - len(drop):nf, 5]}))
        assert not st[nf:, 5].any(), "boundary flags are cleared by the rewrite"
        assert np.array_equal(st, o.read_state())
        o.step(3)
        e.step(3)
    assert tags == [[0, 2], [0, 3], [0]]
    o.step(12)
    e.step(12)
    assert np.array_equal(e.read_state(), o.read_state())
    assert np.array_equal(e.read_bond(), o.read_bond())
    assert (e.read_bond() != 0).any()
    assert np.array_equal(e.chain(), o.chain())
    e.destroy()
