"""GPU parity of device-resident stir schedules (sn_set_events) against the
oracle applying the same impulses from the host.

The reference applies a push immediately before substep s, on the exact
current state, in arming order (src/demo_boot.js:244-272), through
applyImpulse (src/engine.js:6077-6098).  The engine applies a scene's
scheduled events inside its captured substeps; the oracle steps one substep
at a time and calls apply_impulse for each event of that substep in the
order it was armed.  Every substep's H_set || H_slot must be identical.
"""
import numpy as np
import pytest

from tests.helpers import first_mismatch, oracle_world

pytestmark = pytest.mark.gpu

ONE = 65536


def oracle_stream(scene, iters, steps, events):
    """events: (substep, cx, cy, r, ix, iy, limit) in arming order."""
    o = oracle_world(scene, iters)
    rows = []
    for s in range(steps):
        for ev in events:
            if ev[0] == s:
                o.apply_impulse(*ev[1:])
        rows.append(o.step(1, stream=True)[0])
    return o, np.array(rows)


def _timing_schedule(scene):
    """Events on substep 0, two and three on one substep (overlapping, large
    enough that VMAX clamps, so their order shows), the substeps either side
    of the 8-substep graph boundaries, a clipped limit, a limit of 0 and an
    event past the end of the run."""
    p = scene.pos.reshape(-1, 2)[:scene.nFluid]
    cx, cy = int(p[:, 0].mean()), int(p[:, 1].mean())
    return [
        (0, cx, cy, 6 * ONE, 40000, -30000, 3000),
        (3, cx, cy, 5 * ONE, 60000, 0, 3000),
        (3, cx + 2 * ONE, cy, 5 * ONE, -65536, 20000, 3000),
        (3, cx + ONE, cy + ONE, 3 * ONE, 30000, 65536, 3000),
        (7, cx, cy, 4 * ONE, 0, -50000, 3000),
        (8, cx + 3 * ONE, cy, 4 * ONE, -20000, 10000, 3000),
        (11, cx, cy, 30 * ONE, 12345, -54321, 5),          # many hits, only the first 5 slots
        (15, cx, cy, 8 * ONE, 65536, 65536, 0),              # limit 0: nothing moves
        (16, cx - 2 * ONE, cy, 8 * ONE, -30000, -30000, 3000),
        (16, cx - 2 * ONE, cy, 8 * ONE, 30000, 30000, 3000),
        (23, cx, cy - 4 * ONE, 10 * ONE, 1, -1, 3000),
        (500, cx, cy, 8 * ONE, 65536, 0, 3000),              # never reached
    ]


@pytest.mark.parametrize("pieces", [(30,), (1, 2, 5, 9, 13)])
def test_schedule_timing_matches_oracle(pieces):
    from sunaengine_amd import scenes
    from tests.helpers import gpu_engine
    s = scenes.fluid_block_scene(3000, seed=0xE1E1, x_pick=40)
    evs = _timing_schedule(s)
    steps = sum(pieces)
    print(f"events: block n={s.n} pieces={pieces}", flush=True)
    o, ref = oracle_stream(s, 4, steps, evs)
    e = gpu_engine(s, 4)
    e.set_events([(0,) + ev for ev in evs])
    for p in pieces:
        e.step(p)
    got = e.hash_stream()
    bad = first_mismatch(got, ref)
    assert bad is None, f"first differing substep {bad}"
    assert np.array_equal(e.read_state(), o.read_state())
    assert np.array_equal(e.chain(), o.chain())
    e.destroy()
    # the schedule must matter: the same scene without it differs
    o2 = oracle_world(s, 4)
    assert not np.array_equal(o2.step(steps, stream=True), ref)


def test_batched_per_scene_schedules_match_oracle():
    """The batched workload's input path: independent scenes of uneven sizes,
    each with its own stir schedule (radius 8 wu, magnitudes up to 2 wu per
    substep, clamped to VMAX), advanced together in one world."""
    from sunaengine_amd import scenes, serve
    from sunaengine_amd.engine import Engine
    ss = scenes.batch_scenes(count=4, seed=0x5E7E, lo=1024, hi=6000)
    steps = 48
    imgs = [scenes.load_image(x) for x in ss]
    evs = []
    for k, x in enumerate(ss):
        box = (8 * ONE, x.domH // 2, x.domW - 8 * ONE, x.domH - 8 * ONE)
        evs.append(serve.stir_events_2d(scenes.splitmix32(0xA11 + k), steps, box, mean_gap=4))
    assert sum(len(v) for v in evs) > 20
    e = Engine(max_particles=sum(im.n for im in imgs), max_cells=sum(im.cellTotal for im in imgs), iters=4,
               max_scenes=len(ss))
    e.load_scenes(ss)
    e.set_events([(k,) + tuple(ev) for k, v in enumerate(evs) for ev in v])
    e.step(steps)
    for k, x in enumerate(ss):
        print(f"events: batched scene {k} n={x.n} events={len(evs[k])}", flush=True)
        o, ref = oracle_stream(x, 4, steps, evs[k])
        bad = first_mismatch(e.hash_stream(k), ref)
        assert bad is None, f"scene {k}: first differing substep {bad}"
        assert np.array_equal(e.read_state(k), o.read_state())
    e.destroy()


def test_apply_impulse_keeps_schedule_cursors():
    """A host applyImpulse between steps leaves the device schedule's
    cursors where they were: the scheduled events still fire on time."""
    from sunaengine_amd import scenes
    from tests.helpers import gpu_engine
    s = scenes.fluid_block_scene(2000, seed=0xE2E2, x_pick=10)
    p = s.pos.reshape(-1, 2)[:s.nFluid]
    cx, cy = int(p[:, 0].mean()), int(p[:, 1].mean())
    sched = [(2, cx, cy, 6 * ONE, 30000, 0, 3000), (6, cx, cy, 6 * ONE, -30000, 9000, 3000)]
    host = (4, cx + ONE, cy + 2 * ONE, 5 * ONE, 0, -40000, 100)
    o = oracle_world(s, 4)
    rows = []
    for t in range(10):
        for ev in sched:
            if ev[0] == t:
                o.apply_impulse(*ev[1:])
        if t == host[0]:
            hits_ref = o.apply_impulse(*host[1:])
        rows.append(o.step(1, stream=True)[0])
    e = gpu_engine(s, 4)
    e.set_events([(0,) + ev for ev in sched])
    e.step(4)
    print("events: host impulse between scheduled ones", flush=True)
    assert e.apply_impulse(*host[1:]) == hits_ref
    e.step(6)
    assert first_mismatch(e.hash_stream(), np.array(rows)) is None
    e.destroy()


def test_3d_schedule_matches_oracle():
    """The 3D stir events (sn3_set_events) against the oracle applying them
    between substeps: substep 0, three on one substep with clamping overlap,
    a clipped limit, graph boundaries, stepping in pieces."""
    from sunaengine_amd import scenes3d
    from sunaengine_amd.engine3d import Engine3D
    from tests.helpers import oracle_stream_3d
    s = scenes3d.column_scene(3000, gx=12, gy=14, gz=12)
    p = s.parts[:, :3].astype(np.int64)
    c = [int(x) for x in p.mean(axis=0)]
    evs = [(0, c[0], c[1], c[2], 4 * ONE, 40000, 0, -30000, 3000),
           (3, c[0], c[1], c[2], 3 * ONE, 60000, 20000, 0, 3000),
           (3, c[0] + ONE, c[1], c[2], 3 * ONE, -65536, 0, 65536, 3000),
           (3, c[0], c[1] - ONE, c[2] + ONE, 2 * ONE, 30000, -65536, 30000, 3000),
           (8, c[0], c[1], c[2], 20 * ONE, 1234, -4321, 999, 7),
           (15, c[0], c[1], c[2], 5 * ONE, -20000, 50000, 0, 3000),
           (16, c[0], c[1], c[2], 5 * ONE, 20000, -50000, 0, 0)]
    steps = 26
    o, ref = oracle_stream_3d(s, 3, steps, evs)
    e = Engine3D(max_particles=s.n, max_cells=s.cell_total, iters=3)
    e.load([s])
    e.set_events([(0,) + ev for ev in evs])
    for piece in (2, 7, 17):
        e.step(piece)
    print(f"events: 3d column n={s.n}", flush=True)
    bad = first_mismatch(e.hash_stream(0, max_substeps=steps), ref)
    assert bad is None, f"first differing substep {bad}"
    assert np.array_equal(e.read_state()[:, :8], o.read()[:, :8])
    e.destroy()
    o2, plain = oracle_stream_3d(s, 3, steps, [])
    assert not np.array_equal(plain, ref)


def test_push_hits_count_scheduled_and_host_pushes():
    """sn_read_push_hits: the particles moved by scheduled events and by
    host applyImpulse calls since the load, each event counting its hits
    (at most its limit), equal to the oracle's hit counts."""
    from sunaengine_amd import scenes
    from tests.helpers import gpu_engine
    s = scenes.fluid_block_scene(2000, seed=0xE3E3, x_pick=20)
    p = s.pos.reshape(-1, 2)[:s.nFluid]
    cx, cy = int(p[:, 0].mean()), int(p[:, 1].mean())
    sched = [(1, cx, cy, 5 * ONE, 20000, 0, 3000), (3, cx, cy, 30 * ONE, 0, -20000, 17)]
    o = oracle_world(s, 4)
    want = 0
    for t in range(6):
        for ev in sched:
            if ev[0] == t:
                want += o.apply_impulse(*ev[1:])
        if t == 4:
            want += o.apply_impulse(cx, cy + ONE, 4 * ONE, 0, 30000, 3000)
        o.step(1)
    e = gpu_engine(s, 4)
    e.set_events([(0,) + ev for ev in sched])
    e.step(4)
    e.apply_impulse(cx, cy + ONE, 4 * ONE, 0, 30000, 3000)
    e.step(2)
    print(f"events: push hits {want}", flush=True)
    assert e.push_hits() == want and want > 17
    e.destroy()
