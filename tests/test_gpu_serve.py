"""Serving loop on the GPU: every job's per-substep stream is the same when it
runs in the continuous-batching mix (sharing worlds, admitted and retired
between rounds, partitions reused) as when it runs alone, and both equal the
oracle's stream of the job (its scene with its stir events applied between
substeps in arming order)."""
import numpy as np
import pytest

from tests.helpers import first_mismatch, oracle_stream_2d, oracle_stream_3d

pytestmark = pytest.mark.gpu


def _solo(job):
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    from sunaengine_amd.engine3d import Engine3D
    if job.kind == "2d":
        im = scenes.load_image(job.scene)
        e = Engine(max_particles=im.n, max_cells=im.cellTotal, iters=4)
        e.load_scene(job.scene)
        e.set_events([(0,) + tuple(ev) for ev in job.events])
        e.step(job.substeps)
        out = e.hash_stream()
    else:
        e = Engine3D(max_particles=job.scene.n, max_cells=job.scene.cell_total, iters=3)
        e.load([job.scene])
        e.set_events([(0,) + tuple(ev) for ev in job.events])
        e.step(job.substeps)
        out = e.hash_stream(0, max_substeps=job.substeps)
    e.destroy()
    return out


def test_serving_mix_equals_alone():
    from sunaengine_amd import serve
    specs = [("2d", 3000, 70), ("3d", 1500, 45), ("2d", 6000, 33), ("2d", 2500, 90),
             ("3d", 2700, 60), ("2d", 4200, 21), ("3d", 1000, 38)]
    jobs = [serve.make_job(k, 1000 + 17 * k, kind, n, steps) for k, (kind, n, steps) in enumerate(specs)]
    one = 65536
    for j in jobs:   # a strong stir at substep 3 of every job, beside its seeded ones
        if j.kind == "2d":
            extra = (3, j.scene.domW // 2, j.scene.domH * 3 // 4, 8 * one, one, -30000, 3000)
        else:
            w = (j.scene.gx * 2 * one, j.scene.gy * 2 * one, j.scene.gz * 2 * one)
            extra = (3, w[0] // 2, w[1] // 4, w[2] // 2, 8 * one, 40000, 20000, -10000, 3000)
        j.events = sorted(list(j.events) + [extra], key=lambda ev: ev[0])
    srv = serve.Server(classes2d=(serve.SlotClass(7000, 2048, 2),),
                       classes3d=(serve.SlotClass(3000, 4096, 2),), chunk=16, stream_cap=256)
    srv.run(jobs, realtime=False)
    srv.close()
    for j in jobs:
        print(f"job {j.jid} {j.kind} n={j.n} x{j.substeps} events={len(j.events)}", flush=True)
        assert j.stream is not None and j.stream.shape == (j.substeps, 8)
        ref = _solo(j)
        assert first_mismatch(j.stream, ref) is None, f"job {j.jid}"
        if j.kind == "2d":
            _, orc = oracle_stream_2d(j.scene, 4, j.substeps, j.events)
        else:
            _, orc = oracle_stream_3d(j.scene, 3, j.substeps, j.events)
        assert first_mismatch(j.stream, orc) is None, f"job {j.jid}: differs from the oracle"
