"""Scene construction and loadScene's validation (CPU)."""
import numpy as np
import pytest

from sunaengine_amd import materials as mt
from sunaengine_amd import scenes


def test_dambreak_v1_shape():
    s = scenes.make_scene("dambreak_v1")
    assert (s.n, s.nFluid) == (4096, 3600)
    assert s.domW == 128 * 65536 and s.domH == 72 * 65536
    im = scenes.load_image(s)
    assert (im.gridW, im.gridH) == (32, 18)
    assert (im.state[:, 5] == 0).all()              # water: no flags, no weld tag
    assert (im.state[:, 6] == mt.AMBIENT).all()     # absent temp means AMBIENT


def test_boundary_ring_counts():
    assert len(scenes.boundary_ring(128, 72, 60)) == 2 * 128 + 4 * 60
    assert len(scenes.boundary_ring(256, 128, 116)) == 976
    assert len(scenes.boundary_ring(2048, 1024, 1012)) == 8144


def test_benchmark_scenes():
    d = scenes.dam_break_4096()
    assert (d.n, d.nFluid) == (4096 + 976, 4096)
    im = scenes.load_image(d)
    assert (im.gridW, im.gridH) == (64, 32)
    small = scenes.thermal_1m(side=32)
    assert small.nFluid == 1024 and small.temp is not None
    assert (small.temp[: small.nFluid] < 100 * 65536).all()
    b = scenes.batch_scenes(count=16, seed=3)
    assert all(1024 <= s.nFluid <= 65536 for s in b)
    for s in b:
        scenes.load_image(s)                       # every batch scene is loadable


@pytest.mark.parametrize("mutate,msg", [
    (lambda s: setattr(s, "cellShift", 17), "kernel radius"),
    (lambda s: setattr(s, "cellShift", 28), "1..27"),
    (lambda s: s.pos.__setitem__(0, -5), "outside"),
    (lambda s: s.matId.__setitem__(0, 99), "matCount"),
    (lambda s: setattr(s, "flags", np.full(s.n, 8, np.uint32)), "reserved bit"),
    (lambda s: setattr(s, "nFluid", s.n + 1), "nFluid"),
])
def test_load_scene_refusals(mutate, msg):
    s = scenes.make_scene("pool_v1", w=10, h=5)
    mutate(s)
    with pytest.raises(scenes.SceneError, match=msg):
        scenes.load_image(s)


def test_dead_flag_is_derived_from_the_row():
    s = scenes.make_scene("pool_v1", w=10, h=5)
    s.matId[0] = mt.VOID
    im = scenes.load_image(s)
    assert im.state[0, 5] & scenes.FLAG_DEAD
    s2 = scenes.make_scene("pool_v1", w=10, h=5)
    s2.flags = np.zeros(s2.n, np.uint32)
    s2.flags[0] = scenes.FLAG_DEAD
    with pytest.raises(scenes.SceneError, match="FLAG_DEAD"):
        scenes.load_image(s2)


def test_weld_tag_only_on_welding_rows():
    s = scenes.make_scene("pool_v1", w=10, h=5)
    s.matId[: s.nFluid] = mt.STONE
    im = scenes.load_image(s)
    assert ((im.state[: s.nFluid, 5] & scenes.WELD_TAG_MASK) == 1 << 8).all()
    assert ((im.state[: s.nFluid, 5] & scenes.FLAG_SEAL) != 0).all()


def test_spawn_records_follow_spawnfluid():
    s = scenes.make_scene("pool_v1", w=10, h=5)
    fl, bd = scenes.spawn_records(s, [dict(x=100, y=200, vx=1, vy=2)])
    assert list(fl[0]) == [100, 200, 1, 2, mt.WATER, 0, mt.DEG(20), 0]
    assert bd.shape == (s.n - s.nFluid, 8)
    assert (bd[:, 2:4] == 0).all() and (bd[:, 5] == 0).all()
