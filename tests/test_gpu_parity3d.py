"""GPU parity of the 3D solver against the 3D CPU oracle (bit-exact)."""
import numpy as np
import pytest

from tests.helpers import first_mismatch

pytestmark = pytest.mark.gpu


def _run(scene, iters, steps, inputs=None):
    from oracle import OracleWorld3
    from sunaengine_amd import tables3d
    from sunaengine_amd.engine3d import Engine3D
    o = OracleWorld3(scene, tables3d.luts3d(), iters)
    e = Engine3D(max_particles=scene.n, max_cells=scene.cell_total, iters=iters)
    e.load([scene])
    if inputs:
        o.set_inputs(**inputs)
        e.set_inputs(**inputs)
    ref = o.step(steps, stream=True)
    e.step(steps)
    got = e.hash_stream()
    return o, e, ref, got


def test_column_pool_stream():
    """The reference 3D page's own initial column (rest spacing)."""
    from sunaengine_amd import scenes3d
    print("3d column 4000 x60", flush=True)
    o, e, ref, got = _run(scenes3d.column_scene(4000), 3, 60)
    assert first_mismatch(got, ref) is None
    assert np.array_equal(e.chain(), o.chain())
    assert np.array_equal(e.read_state()[:, :8], o.read()[:, :8])
    assert np.array_equal(e.fnv_digest(), o.fnv_digest())
    assert np.array_equal(e.debug_counters(), o.debug_counters())
    e.destroy()


def test_overdense_block_truncation_stream():
    """0.5 wu pitch (8x rest density): neighbour truncation and saturating
    position sums every substep."""
    from sunaengine_amd import scenes3d
    s = scenes3d.block_scene(side=20, gx=16, gy=16, gz=16, seed=11)
    print("3d block 8000 x40", flush=True)
    o, e, ref, got = _run(s, 3, 40)
    assert o.debug_counters()[6] > 0
    assert first_mismatch(got, ref) is None
    assert np.array_equal(e.debug_counters(), o.debug_counters())
    assert np.array_equal(e.read_state(), o.read())
    e.destroy()


def test_layers_with_stir_and_ball():
    """Three materials out of order, the stir triangle wave and the cursor ball."""
    from sunaengine_amd import scenes3d
    s = scenes3d.column_scene(3000, layers=True)
    inputs = dict(stir_amp=6000, stir_phase=384 * 17, ball_on=1,
                  ball=(20 * 65536, 8 * 65536, 20 * 65536), ball_reach=int((3.4 + 2.0) * 65536))
    print("3d layers+ball 3000 x50", flush=True)
    o, e, ref, got = _run(s, 2, 50, inputs)
    assert first_mismatch(got, ref) is None
    assert np.array_equal(e.read_state()[:, :8], o.read()[:, :8])
    e.destroy()


def test_slotted_3d_world():
    """3D scenes placed into and retired from partitions of a running world
    each equal their own oracle run."""
    from oracle import OracleWorld3
    from sunaengine_amd import scenes3d, tables3d
    from sunaengine_amd.engine3d import Engine3D
    a = scenes3d.block_scene(side=10, gx=12, gy=12, gz=12, seed=5)
    b = scenes3d.column_scene(2000)
    c = scenes3d.block_scene(side=12, gx=12, gy=14, gz=12, seed=8)
    caps_n = [3000, 2500]
    caps_c = [max(a.cell_total, c.cell_total), b.cell_total]
    e = Engine3D(max_particles=sum(caps_n), max_cells=sum(caps_c), iters=2, max_scenes=2)
    e.reserve_slots(caps_n, caps_c)
    e.load_slot(0, a); e.step(12)
    e.load_slot(1, b); e.step(10)
    got_a = e.hash_stream(0)                  # 22 substeps
    e.set_active([0, 1]); e.step(5)
    e.load_slot(0, c); e.step(9)
    got_b, got_c = e.hash_stream(1), e.hash_stream(0)   # 24 and 9 substeps
    for sc, steps, got in ((a, 22, got_a), (b, 24, got_b), (c, 9, got_c)):
        o = OracleWorld3(sc, tables3d.luts3d(), 2)
        ref = o.step(steps, stream=True)
        print(f"3d slotted n={sc.n} x{steps}", flush=True)
        assert got.shape == ref.shape
        assert first_mismatch(got, ref) is None
    e.destroy()


def test_large_block_first_substeps():
    """A 64^3 block (262,144 particles, 8x rest density, cells packed far
    past one wave) in a 48 x 40 x 48 grid: the tiled neighbour sweep at scale
    against the oracle on every substep."""
    from oracle.oracle import set_threads
    from sunaengine_amd import scenes3d
    set_threads(16)
    s = scenes3d.block_scene(side=64, gx=48, gy=40, gz=48, seed=0x3D64)
    print(f"3d block {s.n} x3", flush=True)
    o, e, ref, got = _run(s, 3, 3)
    assert first_mismatch(got, ref) is None
    assert np.array_equal(e.debug_counters(), o.debug_counters())
    assert np.array_equal(e.read_state(), o.read())
    e.destroy()
