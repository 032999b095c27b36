"""GPU parity on cells packed with thousands of particles.

Walls and collisions can pile thousands of particles into one grid cell.
Such cells take their own paths in 2D: the canonical rank of a cell above
256 particles runs in k_canon_heavy_rank / _place, and a tile whose three
candidate rows hold more than 32,768 slots takes the unstaged sweep.  In 3D
they run the ordinary kernels, whose per-cell work then grows with the cell.
These scenes start with 3,600 / 10,000 / 36,100 particles (2D) or 4,096 (3D)
inside one cell and keep cells of several hundred to thousands for every
substep checked; the largest starts on the unstaged sweep.

The same scenes also pin the debug:false build (src/engine.js:2884-2894):
identical hashes, chain and state, and debug counters that stay zero.  In
that build a sweep stops testing candidates once every list of its wave is
full, and (2D) finishes the last few unfilled lists of a wave one particle
at a time; these scenes reach both on every substep.
"""
import numpy as np
import pytest

from tests.helpers import first_mismatch, oracle_world

pytestmark = pytest.mark.gpu

ONE = 65536


def _packed2d(side, pitch_shift):
    from sunaengine_amd import scenes
    # x0 = 101 wu and the top row at 127 wu: the block starts inside the one
    # cell [100, 104) x [124, 128) of 4 wu cells
    return scenes.walled_block_scene(side, side, 256, 130, 0xFACE + side, pitch=ONE >> pitch_shift,
                                     x0=101 * ONE, name=f"packed_{side * side}")


def _gpu2d(scene, iters, steps, debug):
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    im = scenes.load_image(scene)
    e = Engine(max_particles=im.n, max_cells=im.cellTotal, iters=iters, debug=debug)
    e.load_scene(scene)
    e.step(steps)
    return e


@pytest.mark.parametrize("side,pitch_shift,steps", [(60, 5, 12), (100, 6, 12), (190, 7, 6)])
def test_packed_cell_2d(side, pitch_shift, steps):
    s = _packed2d(side, pitch_shift)
    print(f"{s.name} x{steps} (debug on and off)", flush=True)
    o = oracle_world(s, 4)
    ref = o.step(steps, stream=True)
    assert o.debug_counters()[6] > 0
    for debug in (True, False):
        e = _gpu2d(s, 4, steps, debug)
        got = e.hash_stream()
        bad = first_mismatch(got, ref)
        assert bad is None, f"{s.name} debug={debug}: first differing substep {bad}"
        assert np.array_equal(e.chain(), o.chain())
        assert np.array_equal(e.read_state(), o.read_state())
        if debug:
            assert np.array_equal(e.debug_counters(), o.debug_counters())
        else:
            assert not e.debug_counters().any()
        e.destroy()


def test_debug_off_matches_on_dam_break():
    """An ordinary scene: debug:false changes no hash, zeroes the counters."""
    from sunaengine_amd import scenes
    s = scenes.make_scene("overdense_v1")
    print("overdense_v1 x60 debug off", flush=True)
    o = oracle_world(s, 4)
    ref = o.step(60, stream=True)
    e = _gpu2d(s, 4, 60, False)
    assert first_mismatch(e.hash_stream(), ref) is None
    assert np.array_equal(e.read_state(), o.read_state())
    assert not e.debug_counters().any()
    e.destroy()


def test_packed_cell_3d():
    from oracle import OracleWorld3
    from sunaengine_amd import scenes3d, tables3d
    from sunaengine_amd.engine3d import Engine3D
    # 16^3 at 1/16 wu from 2.5 wu: every particle inside cell (1, 1, 1) of 2 wu cells
    s = scenes3d.block_scene(side=16, gx=8, gy=8, gz=8, pitch=ONE >> 4, origin_wu=2.5, seed=0x3EA7)
    assert len({tuple(r) for r in (s.parts[:, :3] >> 17).tolist()}) == 1
    print("3d packed 4096 x10 (debug on and off)", flush=True)
    o = OracleWorld3(s, tables3d.luts3d(), 3)
    ref = o.step(10, stream=True)
    assert o.debug_counters()[6] > 0
    for debug in (True, False):
        e = Engine3D(max_particles=s.n, max_cells=s.cell_total, iters=3, debug=debug)
        e.load([s])
        e.step(10)
        bad = first_mismatch(e.hash_stream(), ref)
        assert bad is None, f"3d packed debug={debug}: first differing substep {bad}"
        assert np.array_equal(e.chain(), o.chain())
        assert np.array_equal(e.read_state()[:, :8], o.read()[:, :8])
        if debug:
            assert np.array_equal(e.debug_counters(), o.debug_counters())
        else:
            assert not e.debug_counters().any()
        e.destroy()


def test_corner_pileup_debug_off_long_run():
    """A 76,175-particle dam break in the 256 x 128 wu box piles about 26,000
    particles into its floor corner cell within a few hundred substeps, nearly
    all on the one corner point; the chunks of that cell that hold a particle
    off the point, with fewer than 48 neighbours in range, sweep every block of
    the pile, position by position.  600 substeps: the stripped build's stream equals the
    instrumented build's (itself pinned to the oracle above and, here, over
    the first 20 substeps)."""
    from sunaengine_amd import scenes, serve
    from sunaengine_amd.engine import Engine
    job = serve.make_job(6, 0xC0A1, "2d", 76175, 600)
    s = job.scene
    im = scenes.load_image(s)
    print("corner pile-up 76175 x600 (debug on vs off) + oracle x20", flush=True)
    o = oracle_world(s, 4)
    ref20 = o.step(20, stream=True)
    streams = []
    for debug in (True, False):
        e = Engine(max_particles=im.n, max_cells=im.cellTotal, iters=4, debug=debug)
        e.load_scene(s)
        e.step(600)
        streams.append((e.hash_stream(), e.chain(), e.read_state()))
        e.destroy()
    on, off = streams
    assert first_mismatch(on[0][:20], ref20) is None
    bad = first_mismatch(off[0], on[0])
    assert bad is None, f"debug off differs from debug on at substep {bad}"
    assert np.array_equal(off[1], on[1]) and np.array_equal(off[2], on[2])
    st = off[2]
    cells = (st[:, 1].astype(np.int64) >> 18) * 64 + (st[:, 0].astype(np.int64) >> 18)
    assert np.unique(cells, return_counts=True)[1].max() > 8192   # the pile-up formed


def _corner_point_pile(n_pile, n_spread, W=48, Hh=24, seed=0xC0C0):
    """n_pile water particles on the exact corner point the wall clamp
    holds them at (the bottom-right one, where a floor corner pile-up ends),
    with n_spread others scattered over the 3 wu beside it, every
    (n_pile + n_spread) / n_spread-th in slot order.  A cell's slots are
    ordered by id, so nearly every 64-slot block of the cell's candidates is
    one exact position."""
    from sunaengine_amd import materials as mt
    from sunaengine_amd import scenes
    from tests import material_scenes as ms
    rnd = scenes.xorshift32(seed)
    wall = ONE >> 1
    cx, cy = W * ONE - wall, Hh * ONE - wall
    every = (n_pile + n_spread) // n_spread
    parts = []
    for i in range(n_pile + n_spread):
        if i % every == 0:
            parts.append((cx - rnd() % (3 * ONE), cy - rnd() % (3 * ONE), mt.WATER))
        else:
            parts.append((cx, cy, mt.WATER))
    return ms._scene(parts, W, Hh, "corner_point_pile")


@pytest.mark.parametrize("n_pile,n_spread,steps", [(3000, 300, 12), (34000, 600, 4)])
def test_corner_point_pile(n_pile, n_spread, steps):
    """A pile-up stacked on one point: most candidate blocks are a single
    position, which the sweep takes per position instead of per slot once
    every lane that reaches it holds a full list.  3,000 on the point keep
    the staged sweep; 34,000 exceed its 32,768 candidates and take the
    unstaged one.  Both builds against the oracle: stream, chain, state, and
    (debug) the truncation count, which the grouped blocks add up by
    position."""
    s = _corner_point_pile(n_pile, n_spread)
    print(f"{s.name} {s.n} x{steps} (debug on and off)", flush=True)
    o = oracle_world(s, 4)
    ref = o.step(steps, stream=True)
    st = o.read_state()
    assert np.unique(st[:, :2], axis=0, return_counts=True)[1].max() > n_pile * 5 // 6   # still one stack
    assert o.debug_counters()[6] > 0
    for debug in (True, False):
        e = _gpu2d(s, 4, steps, debug)
        bad = first_mismatch(e.hash_stream(), ref)
        assert bad is None, f"{s.name} debug={debug}: first differing substep {bad}"
        assert np.array_equal(e.chain(), o.chain())
        assert np.array_equal(e.read_state(), st)
        if debug:
            assert np.array_equal(e.debug_counters(), o.debug_counters())
        else:
            assert not e.debug_counters().any()
        e.destroy()
