"""GPU parity: the HIP engine against the frozen goldens and the CPU oracle.

Every comparison is bit-exact (integer path).  Per-substep checks compare
the full 8-word digest (H_set ‖ H_slot) of every substep.
"""
import numpy as np
import pytest

from tests.helpers import first_mismatch, golden, gpu_engine, oracle_world

pytestmark = pytest.mark.gpu


def _print(msg):
    print(msg, flush=True)


def test_dambreak_v1_goldens():
    """The reference's frozen gate: dambreak_v1, iters=4, 1000 substeps."""
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    g = golden("dambreak_v1.json")
    e = Engine(max_particles=8192, max_cells=65536, iters=4)
    e.load_scene(scenes.make_scene("dambreak_v1"))
    done = 0
    for c in (1, 10, 100, 1000):
        e.step(c - done)
        done = c
        _print(f"dambreak_v1 checkpoint {c}")
        assert e.digest()["hex"] == g["checkpoints"][str(c)], f"checkpoint {c}"
    assert e.chain_hex() == g["chain_1000"]
    assert list(e.debug_counters()) == g["dbg_1000"]
    # the per-substep stream's last entry is the digest of the current state
    hs = e.hash_stream()
    assert hs.shape == (1000, 8)
    assert "".join("%08x" % w for w in hs[-1]) == g["checkpoints"]["1000"]
    e.destroy()


def _stream_parity(scene, iters, steps, **kw):
    o = oracle_world(scene, iters)
    ref = o.step(steps, stream=True)
    e = gpu_engine(scene, iters, **kw)
    e.step(steps)
    got = e.hash_stream()
    bad = first_mismatch(got, ref)
    assert bad is None, f"{scene.name}: first differing substep {bad}"
    assert np.array_equal(e.chain(), o.chain())
    assert np.array_equal(e.debug_counters(), o.debug_counters())
    assert np.array_equal(e.read_state(), o.read_state())
    return e, o


def test_dam_break_4096_iters3_stream():
    """The 4096-particle dam break at 0.5 wu pitch, 3 iterations."""
    from sunaengine_amd import scenes
    _print("dam_break_4096 x300")
    e, o = _stream_parity(scenes.dam_break_4096(), 3, 300)
    e.destroy()


def test_overdense_truncation_stream():
    """overdense_v1: 4x rest density, neighbour truncation on every substep."""
    from sunaengine_amd import scenes
    _print("overdense_v1 x120")
    e, o = _stream_parity(scenes.make_scene("overdense_v1"), 4, 120)
    assert o.debug_counters()[6] > 0
    e.destroy()


def test_unbounded_domain_stream():
    """A domain wider than 2^28 ticks: the kernels take the saturating pair
    geometry (the bounded fast path is off for this scene) and must still
    match the oracle bit for bit."""
    from sunaengine_amd import scenes
    wide = scenes.dam_break_4096()
    wide.domW = 1 << 29
    wide.name = "dam_break_4096_wide"
    _print("dam_break_4096 in a 2^29-tick wide domain x120")
    e, o = _stream_parity(wide, 3, 120)
    e.destroy()


def test_thermal_small_stream():
    """Non-uniform temperatures: conduction and ambient relaxation."""
    from sunaengine_amd import scenes
    s = scenes.thermal_1m(side=48)
    _print("thermal_2304 x150")
    e, o = _stream_parity(s, 4, 150)
    e.destroy()


def test_pool_with_impulses():
    """applyImpulse between substeps (ascending slots, limit, VMAX clamp)."""
    from sunaengine_amd import scenes
    s = scenes.make_scene("pool_v1", w=60, h=20)
    o = oracle_world(s, 4)
    e = gpu_engine(s, 4)
    for k, (cx, cy, r, ix, iy, lim) in enumerate([
            (30 * 65536, 60 * 65536, 8 * 65536, 40000, -20000, 3000),
            (40 * 65536, 62 * 65536, 6 * 65536, -131072, 90000, 25),
            (20 * 65536, 64 * 65536, 10 * 65536, 65536, 65536, 3000)]):
        o.step(17)
        e.step(17)
        h_o = o.apply_impulse(cx, cy, r, ix, iy, lim)
        h_e = e.apply_impulse(cx, cy, r, ix, iy, lim)
        _print(f"impulse {k}: hits {h_o}")
        assert h_o == h_e
        assert np.array_equal(e.read_state(), o.read_state())
    o.step(40)
    e.step(40)
    assert e.digest()["hex"] == "".join("%08x" % w for w in o.digest())
    assert np.array_equal(e.chain(), o.chain())
    e.destroy()


def test_spawn_fluid_matches_oracle():
    """spawnFluid: boundary block shifts up, formation runs the next substep."""
    from sunaengine_amd import scenes
    s = scenes.make_scene("pool_v1", w=40, h=12)
    o = oracle_world(s, 4, max_particles=2000)
    e = gpu_engine(s, 4, max_particles=2000)
    o.step(9)
    e.step(9)
    drop = [dict(x=(60 + dx) * 65536, y=20 * 65536 + dy * 65536, vx=0, vy=3000)
            for dy in range(3) for dx in range(3)]
    fl, bd = scenes.spawn_records(s, drop)
    assert o.spawn(fl, bd) == e.spawn_fluid(drop)
    o.step(30)
    e.step(30)
    assert np.array_equal(e.read_state(), o.read_state())
    assert np.array_equal(e.chain(), o.chain())
    e.destroy()


def test_batched_scenes_equal_solo():
    """Independent scenes advanced together: each equals its own oracle run."""
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    ss = scenes.batch_scenes(count=6, seed=77, lo=300, hi=2000)
    tot_n = sum(s.n for s in ss)
    tot_c = sum(scenes.load_image(s).cellTotal for s in ss)
    e = Engine(max_particles=tot_n, max_cells=tot_c, iters=3, max_scenes=len(ss))
    e.load_scenes(ss)
    e.step(80)
    for k, s in enumerate(ss):
        o = oracle_world(s, 3)
        ref = o.step(80, stream=True)
        got = e.hash_stream(k)
        _print(f"batch scene {k}: n={s.n}")
        assert first_mismatch(got, ref) is None, f"scene {k}"
        assert np.array_equal(e.chain(k), o.chain())
    e.destroy()


def test_slotted_world_continuous_batching():
    """Scenes placed into and retired from partitions of a running world:
    each scene's per-substep stream equals its own oracle run, whatever ran
    beside it and however the partitions were reused."""
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    ss = scenes.batch_scenes(count=3, seed=91, lo=400, hi=1500)
    caps_n = [2000, 2600, 2600]
    caps_c = [2048, 4096, 2048]
    e = Engine(max_particles=sum(caps_n), max_cells=sum(caps_c), iters=3, max_scenes=3)
    e.reserve_slots(caps_n, caps_c)
    ran = {}
    e.load_slot(1, ss[0]); e.step(20)                    # A alone in slot 1
    e.load_slot(2, ss[1]); e.step(30)                    # A and B
    ran[0] = (1, 50, e.hash_stream(1))
    e.set_active([0, 0, 1]); e.step(10)                  # A retired, B alone
    e.load_slot(1, ss[2]); e.step(15)                    # C reuses slot 1 beside B
    ran[1] = (2, 55, e.hash_stream(2))
    ran[2] = (1, 15, e.hash_stream(1))
    for k, (slot, steps, got) in ran.items():
        o = oracle_world(ss[k], 3)
        ref = o.step(steps, stream=True)
        _print(f"slotted scene {k}: slot {slot}, {steps} substeps, n={ss[k].n}")
        assert got.shape == ref.shape
        assert first_mismatch(got, ref) is None, f"scene {k}"
    e.destroy()


@pytest.mark.parametrize("wide", [False, True], ids=["bounded", "unbounded"])
def test_full_scale_1m_first_substeps(wide):
    """The 1,048,576-particle scene itself (the measured workload), first
    substeps against the oracle on every substep; `unbounded` widens the
    domain past 2^28 ticks so the saturating pair geometry runs at scale."""
    from sunaengine_amd import scenes
    from oracle.oracle import set_threads
    set_threads(16)
    s = scenes.thermal_1m()
    if wide:
        s.domW = 1 << 29
        s.name += "_wide"
    _print(f"{s.name}: n={s.n}, 4 substeps")
    e, o = _stream_parity(s, 4, 4)
    e.destroy()


@pytest.mark.parametrize("name", ["pool", "stone_in_water"])
def test_restore_state_derived_matches_oracle(name):
    """restore(snap) with derived (src/demo_boot.js:308-321): a mid-run
    snapshot written into a freshly loaded world continues exactly as the
    oracle's restored world does; a water scene also rejoins its own
    timeline (its formation passes are bit-inert)."""
    from sunaengine_amd import scenes
    from tests import material_scenes as ms
    s = scenes.make_scene("pool_v1", w=40, h=12) if name == "pool" else ms.stone_in_water()
    donor = oracle_world(s, 4)
    donor.step(9)
    st, dv = donor.read_state(), donor.read_derived()
    o = oracle_world(s, 4)
    o.write_state_derived(st, dv)
    e = gpu_engine(s, 4)
    e.write_state_derived(st, dv)
    ref = o.step(20, stream=True)
    e.step(20)
    assert first_mismatch(e.hash_stream(), ref) is None
    assert np.array_equal(e.read_state(), o.read_state())
    if name == "pool":
        assert first_mismatch(ref, donor.step(20, stream=True)) is None
    e.destroy()

