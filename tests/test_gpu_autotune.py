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
