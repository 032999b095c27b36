"""GPU parity in the packed regime, against oracle fixtures made offline
(tests/golden/make_packed_fixtures.py).

* pileup_76k: the 76,175-particle dam break in the 256 x 128 wu box piles
  more than 16,384 particles into its floor corner cell, so the canonical
  rank of that cell runs in k_canon_heavy_rank (the pairwise count over id
  ranges) as well as the sorted path for cells of 257..16,384; its tiles
  take the heavy queue and the spatial chunks.  Both builds (debug counters
  on and off) are compared with the oracle at every substep.
* thermal_1m: the measured 1M-particle scene, every substep from load up to
  the fixture's length, far into the collapse where floor cells hold
  thousands of particles.

Each fixture also records the largest cell occupancy per substep; the tests
check that the regime was reached.
"""
import os

import numpy as np
import pytest

from tests.helpers import GOLDEN, first_mismatch

pytestmark = pytest.mark.gpu


def _fixture(name):
    p = os.path.join(GOLDEN, f"{name}_stream.npz")
    if not os.path.exists(p):
        pytest.skip(f"{p} not generated")
    return np.load(p)


def _run(scene, steps, debug):
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    im = scenes.load_image(scene)
    e = Engine(max_particles=im.n, max_cells=im.cellTotal, iters=4, debug=debug,
               stream_cap=max(4096, steps + 8))
    e.load_scene(scene)
    e.step(steps)
    return e


@pytest.mark.parametrize("debug", [True, False])
def test_pileup_76k_matches_oracle_fixture(debug):
    from sunaengine_amd import serve
    fx = _fixture("pileup_76k")
    ref = fx["stream"]
    assert int(fx["maxcell"].max()) > 16384, "the fixture must reach a cell above 16,384 slots"
    s = serve.make_job(6, 0xC0A1, "2d", 76175, 600).scene
    print(f"pileup_76k x{ref.shape[0]} debug={debug}", flush=True)
    e = _run(s, ref.shape[0], debug)
    bad = first_mismatch(e.hash_stream(), ref)
    assert bad is None, f"first differing substep {bad} (largest cell there {fx['maxcell'][bad]})"
    assert np.array_equal(e.chain(), fx["chain"])
    if debug:
        assert np.array_equal(e.debug_counters(), fx["dbg"])
    e.destroy()


@pytest.mark.parametrize("debug", [True, False])
def test_thermal_1m_matches_oracle_fixture(debug):
    from sunaengine_amd import scenes
    fx = _fixture("thermal_1m")
    ref = fx["stream"]
    assert int(fx["maxcell"].max()) > 4000, "the fixture must reach cells of thousands"
    print(f"thermal_1m x{ref.shape[0]} debug={debug}", flush=True)
    e = _run(scenes.thermal_1m(), ref.shape[0], debug)
    bad = first_mismatch(e.hash_stream(), ref)
    assert bad is None, f"first differing substep {bad} (largest cell there {fx['maxcell'][bad]})"
    assert np.array_equal(e.chain(), fx["chain"])
    if debug:
        assert np.array_equal(e.debug_counters(), fx["dbg"])
    else:
        assert not e.debug_counters().any()
    e.destroy()
