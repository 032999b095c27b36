"""sn_autotune: the launch shape is tuned on the live world and leaves no
trace.  State, bond rows, debug counters and the hash stream are as before
the call, and the following substeps still equal the oracle's."""
import numpy as np
import pytest

from tests import material_scenes as ms
from tests.helpers import first_mismatch, oracle_world

pytestmark = pytest.mark.gpu


def test_autotune_batched_world_leaves_no_trace():
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    ss = scenes.batch_scenes(count=6, seed=77, lo=300, hi=2000)
    tot_n = sum(s.n for s in ss)
    tot_c = sum(scenes.load_image(s).cellTotal for s in ss)
    e = Engine(max_particles=tot_n, max_cells=tot_c, iters=3, max_scenes=len(ss))
    e.load_scenes(ss)
    e.step(10)
    before = [(e.read_state(k), e.hash_stream(k), e.debug_counters(k), e.substep_count(k)) for k in range(len(ss))]
    rep = e.autotune(trial_substeps=8, lookahead=12)
    print(rep, flush=True)
    assert len(rep["trials"]) == 12
    best = min(rep["trials"], key=lambda t: t["ms_per_substep"])
    assert (rep["chosen"]["tile_cells"], rep["chosen"]["graph_substeps"]) == (best["tile_cells"], best["graph_substeps"])
    for k, (st, hs, dbg, sc) in enumerate(before):
        assert np.array_equal(e.read_state(k), st) and np.array_equal(e.hash_stream(k), hs)
        assert np.array_equal(e.debug_counters(k), dbg) and e.substep_count(k) == sc
    e.step(30)
    for k, s in enumerate(ss):
        ref = oracle_world(s, 3).step(40, stream=True)
        assert first_mismatch(e.hash_stream(k), ref) is None, f"scene {k}"
    e.destroy()


def test_autotune_material_world_keeps_bonds():
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    s = ms.stone_in_water()
    im = scenes.load_image(s)
    e = Engine(max_particles=im.n, max_cells=im.cellTotal, iters=4)
    e.load_scene(s)
    e.step(6)
    bond = e.read_bond()
    assert (bond != 0).any()
    e.autotune(trial_substeps=8)
    assert np.array_equal(e.read_bond(), bond)
    e.step(10)
    o = oracle_world(s, 4)
    ref = o.step(16, stream=True)
    assert first_mismatch(e.hash_stream(), ref) is None
    assert np.array_equal(e.read_bond(), o.read_bond())
    e.destroy()


def test_new_entry_points_refuse_bad_calls():
    """autotune before any load is a state error; readBond with a short
    buffer, restore into a scene that does not exist or with a matId outside
    the table are argument errors; none of them changes the world."""
    import ctypes as C
    from sunaengine_amd import native, scenes
    from sunaengine_amd.engine import Engine
    s = scenes.make_scene("pool_v1", w=20, h=8)
    im = scenes.load_image(s)
    e = Engine(max_particles=im.n, max_cells=im.cellTotal, iters=4)
    with pytest.raises(native.SnError) as ex:
        e.autotune()
    assert ex.value.code == native.SN_E_STATE
    e.load_scene(s)
    e.step(3)
    st = e.read_state()
    short = np.zeros(8, np.uint32)
    assert native.lib().sn_read_bond(e._h, 0, short.ctypes.data, short.size) == native.SN_E_ARG
    with pytest.raises(native.SnError) as ex:
        e.write_state_derived(st, scene=1)
    assert ex.value.code == native.SN_E_ARG
    bad = st.copy()
    bad[0, 4] = 64
    with pytest.raises(native.SnError) as ex:
        e.write_state_derived(bad)
    assert ex.value.code == native.SN_E_ARG
    assert np.array_equal(e.read_state(), st)
    e.step(2)
    ref = oracle_world(s, 4).step(5, stream=True)
    assert first_mismatch(e.hash_stream(), ref) is None
    e.destroy()
