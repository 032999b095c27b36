"""Shared helpers for the parity tests (oracle runs, GPU engine runs)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def oracle_world(scene, iters, max_particles=None, max_cells=None):
    from oracle import OracleWorld
    from sunaengine_amd import materials, scenes, tables
    img = scenes.load_image(scene)
    return OracleWorld(img, tables.load_luts(), materials.mats_image(), iters=iters,
                       max_particles=max_particles or img.n, max_cells=max_cells or img.cellTotal)


def gpu_engine(scene, iters, max_particles=None, max_cells=None, stream_cap=4096):
    from sunaengine_amd.engine import Engine
    from sunaengine_amd import scenes
    img = scenes.load_image(scene)
    e = Engine(max_particles=max_particles or img.n, max_cells=max_cells or img.cellTotal,
               iters=iters, stream_cap=stream_cap)
    e.load_scene(scene)
    return e


def first_mismatch(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} != {b.shape}"
    bad = np.nonzero((a != b).reshape(a.shape[0], -1).any(axis=1))[0]
    return None if bad.size == 0 else int(bad[0])
