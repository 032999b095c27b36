"""Write the packed-regime fixtures: the oracle's per-substep H_set || H_slot
for scenes whose cells pack far past 256 particles, where the engine's
heavy-cell paths run (the canonical rank of cells above 256 and above 16,384
slots, heavy neighbour tiles, spatial chunks, the stripped build's tail).

    python tests/golden/make_packed_fixtures.py [thermal_1m|pileup_76k] [substeps]

The oracle (test infrastructure, oracle/) runs on this host's cores; each
fixture holds the stream (uint32 [substeps, 8]), the chain and debug
counters after the last substep, and the largest cell occupancy after each
substep (so a test can check which regime the run reached).  Outputs:
tests/golden/<name>_stream.npz.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

DEFAULT_STEPS = {"thermal_1m": 200, "pileup_76k": 160}


def scene_of(name):
    from sunaengine_amd import scenes, serve
    if name == "thermal_1m":
        return scenes.thermal_1m()
    if name == "pileup_76k":
        return serve.make_job(6, 0xC0A1, "2d", 76175, 600).scene
    raise SystemExit(f"unknown fixture {name}")


def make(name, steps):
    from oracle import OracleWorld
    from sunaengine_amd import materials, scenes, tables
    s = scene_of(name)
    im = scenes.load_image(s)
    o = OracleWorld(im, tables.load_luts(), materials.mats_image(), iters=4)
    stream = np.zeros((steps, 8), np.uint32)
    maxcell = np.zeros(steps, np.int64)
    t0 = time.time()
    for k in range(steps):
        stream[k] = o.step(1, stream=True)[0]
        st = o.read_state()
        cells = (st[:, 1].astype(np.int64) >> im.cellShift) * im.gridW + (st[:, 0].astype(np.int64) >> im.cellShift)
        maxcell[k] = np.bincount(cells).max()
        if k % 10 == 9 or k == steps - 1:
            # every 10 substeps the fixture so far is written, so a long run
            # that is stopped still leaves a usable prefix
            print(f"{name}: substep {k + 1}/{steps}, largest cell {maxcell[k]}, {time.time() - t0:.0f} s",
                  flush=True)
            save(name, stream[:k + 1], o, maxcell[:k + 1], im)


def save(name, stream, o, maxcell, im):
    out = os.path.join(HERE, f"{name}_stream.npz")
    np.savez_compressed(out, stream=stream, chain=o.chain(), dbg=o.debug_counters(), maxcell=maxcell,
                        n=np.int64(im.n), n_fluid=np.int64(im.nFluid))


if __name__ == "__main__":
    names = [sys.argv[1]] if len(sys.argv) > 1 else list(DEFAULT_STEPS)
    for nm in names:
        make(nm, int(sys.argv[2]) if len(sys.argv) > 2 else DEFAULT_STEPS[nm])
