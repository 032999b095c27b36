"""Serving loop: replay/verify jobs of mixed sizes with continuous batching.

A job is a seeded initial state (2D or 3D), a substep count and a stir
schedule; its result is the per-substep H_set ‖ H_slot stream.  Jobs arrive
over time (a Poisson stream in the benchmark), wait in a queue, and are
admitted into a free partition of a resident world as soon as one that fits
them is free; admission is by device memory, each partition being a fixed
slice of particle slots and grid cells.  The 2D and 3D worlds each keep a
round in flight on their own stream (overlap=False: they take turns, one
round on the device at a time); one host thread enqueues the rounds, polls
them without blocking, and admits and retires jobs in between.  A world
advances all running jobs by rounds of `chunk` substeps (fewer when every running job needs
fewer); a job that ends inside a round simply runs on to the round's end,
and its result is the first `substeps` entries of its stream, which the
extra substeps do not touch.  A finished job's stream is read back and its
partition retired.  A scene's hashes depend only on its own slots, so a
job's stream is the same alone and in any mix.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import scenes as sc2
from . import scenes3d as sc3
from .engine import Engine
from .engine3d import Engine3D

ONE = 65536


@dataclass
class Job:
    jid: int
    kind: str                   # "2d" or "3d"
    scene: object               # scenes.Scene or scenes3d.Scene3
    substeps: int
    events: list                # job-relative stir events (see stir_events_2d / _3d)
    arrival: float = 0.0        # seconds after the run starts
    admitted: Optional[float] = None
    finished: Optional[float] = None
    slot: Optional[int] = None
    done: int = 0
    stream: Optional[np.ndarray] = None

    @property
    def n(self) -> int:
        return int(self.scene.n)

    @property
    def work(self) -> int:
        """Particle-substeps of the job."""
        return self.n * self.substeps


# --------------------------------------------------------------- workloads
def stir_events_2d(seed: int, substeps: int, box, mean_gap: int = 50, radius_wu: int = 8,
                   max_wu: float = 2.0, limit: int = 3000):
    """Stir impulses at Bernoulli(1/mean_gap) substep indices (a Poisson
    stream in discrete time): radius 8 wu, magnitude uniform in [0, 2] wu per
    substep in a uniform direction, centred in the box's lower half."""
    rnd = sc2.xorshift32(seed)
    (x0, y0, x1, y1) = box
    out = []
    for t in range(substeps):
        if rnd() % mean_gap:
            continue
        cx = x0 + rnd() % max(1, x1 - x0)
        cy = y0 + rnd() % max(1, y1 - y0)
        mag = (rnd() % (int(max_wu * ONE) + 1))
        ang = (rnd() % 3600) * (math.pi / 1800.0)
        out.append((t, int(cx), int(cy), radius_wu * ONE, int(round(mag * math.cos(ang))),
                    int(round(mag * math.sin(ang))), limit))
    return out


def stir_events_3d(seed: int, substeps: int, extent, mean_gap: int = 50, radius_wu: int = 8,
                   max_wu: float = 2.0, limit: int = 3000):
    """The 3D form: a uniform direction on the sphere."""
    rnd = sc2.xorshift32(seed)
    wx, wy, wz = extent
    out = []
    for t in range(substeps):
        if rnd() % mean_gap:
            continue
        c = [rnd() % max(1, wx), rnd() % max(1, wy // 2), rnd() % max(1, wz)]
        mag = rnd() % (int(max_wu * ONE) + 1)
        u = (rnd() % 20001) / 10000.0 - 1.0
        ph = (rnd() % 3600) * (math.pi / 1800.0)
        s = math.sqrt(max(0.0, 1.0 - u * u))
        d = (s * math.cos(ph), u, s * math.sin(ph))
        out.append((t, int(c[0]), int(c[1]), int(c[2]), radius_wu * ONE,
                    int(round(mag * d[0])), int(round(mag * d[1])), int(round(mag * d[2])), limit))
    return out


def make_job(jid: int, seed: int, kind: str, n: int, substeps: int, arrival: float = 0.0) -> Job:
    """A seeded job: a 0.5 wu water block of about n particles (2D: in a
    walled box; 3D: a cube in a cell grid with margin) and its stir schedule."""
    if kind == "2d":
        s = sc2.fluid_block_scene(n, seed, x_pick=sc2.splitmix32(seed) & 0xFFFF, name=f"job{jid}")
        box = (8 * ONE, s.domH // 2, s.domW - 8 * ONE, s.domH - 8 * ONE)
        ev = stir_events_2d(seed ^ 0x5A5A, substeps, box)
    else:
        side = max(4, int(round(n ** (1.0 / 3.0))))
        g = max(8, (side + 1) // 4 + 6)
        s = sc3.block_scene(side=side, gx=g, gy=g, gz=g, seed=seed)
        ev = stir_events_3d(seed ^ 0x5A5A, substeps, (s.gx * 2 * ONE, s.gy * 2 * ONE, s.gz * 2 * ONE))
    return Job(jid, kind, s, substeps, ev, arrival)


def poisson_jobs(count: int, rate: float, seed: int = 0x5E7E, n_lo: int = 4096, n_hi: int = 262144,
                 steps_lo: int = 100, steps_hi: int = 5000, frac_3d: float = 0.3) -> List[Job]:
    """`count` jobs with exponential inter-arrival times at `rate` jobs/s:
    70% 2D / 30% 3D, particle counts log-uniform in [n_lo, n_hi], substeps
    uniform in [steps_lo, steps_hi]."""
    rnd = sc2.xorshift32(seed)
    t = 0.0
    jobs = []
    for j in range(count):
        u = ((rnd() & 0xFFFFFF) + 0.5) / float(1 << 24)
        t += -math.log(u) / rate
        kind = "3d" if (rnd() % 1000) < int(frac_3d * 1000) else "2d"
        v = (rnd() & 0xFFFFFF) / float(1 << 24)
        n = int(round(math.exp(math.log(n_lo) + v * (math.log(n_hi) - math.log(n_lo)))))
        steps = steps_lo + rnd() % (steps_hi - steps_lo + 1)
        jobs.append(make_job(j, sc2.splitmix32(seed + 7919 * j), kind, n, steps, t))
    return jobs


# -------------------------------------------------------------- partitions
@dataclass
class SlotClass:
    particles: int
    cells: int
    count: int


DEFAULT_2D = (SlotClass(20000, 2048, 24), SlotClass(70000, 8192, 16), SlotClass(270000, 8192, 8))
DEFAULT_3D = (SlotClass(20000, 8192, 12), SlotClass(70000, 16384, 8), SlotClass(270000, 65536, 4))


def _footprint(job: Job):
    if job.kind == "2d":
        im = sc2.load_image(job.scene, max_particles=1 << 30, max_cells=1 << 30)
        return im.n, im.cellTotal
    return job.scene.n, job.scene.cell_total


class Server:
    """Two resident slotted worlds (2D, 3D) and the admission / round loop."""

    def __init__(self, classes2d: Sequence[SlotClass] = DEFAULT_2D, classes3d: Sequence[SlotClass] = DEFAULT_3D,
                 iters2d: int = 4, iters3d: int = 3, chunk: int = 64, stream_cap: int = 8192,
                 debug: bool = False):
        """`debug=False` (the default here) is the reference's stripped build:
        the same hashes, no debug counters, and neighbour sweeps that skip
        work only the counters need, which keeps cells that walls have
        packed with thousands of particles cheap."""
        self.chunk = chunk
        self.caps = {}
        self.engines = {}
        for kind, classes, iters in (("2d", classes2d, iters2d), ("3d", classes3d, iters3d)):
            cp, cc = [], []
            for c in classes:
                cp += [c.particles] * c.count
                cc += [c.cells] * c.count
            if not cp:
                continue
            if kind == "2d":
                e = Engine(max_particles=sum(cp), max_cells=sum(cc), iters=iters, max_scenes=len(cp),
                           stream_cap=stream_cap, debug=debug)
            else:
                e = Engine3D(max_particles=sum(cp), max_cells=sum(cc), iters=iters, max_scenes=len(cp),
                             stream_cap=stream_cap, debug=debug)
            e.reserve_slots(cp, cc)
            self.engines[kind] = e
            self.caps[kind] = list(zip(cp, cc))
        self.busy = {k: [None] * len(v) for k, v in self.caps.items()}
        self.stream_cap = stream_cap

    def close(self):
        for e in self.engines.values():
            e.destroy()
        self.engines = {}

    # ---- admission by memory: the smallest free partition the job fits
    def _find_slot(self, job: Job, fp) -> Optional[int]:
        best = None
        for s, (cp, cc) in enumerate(self.caps.get(job.kind, [])):
            if self.busy[job.kind][s] is None and fp[0] <= cp and fp[1] <= cc:
                if best is None or cp < self.caps[job.kind][best][0]:
                    best = s
        return best

    def fits_anywhere(self, job: Job) -> bool:
        fp = _footprint(job)
        return any(fp[0] <= cp and fp[1] <= cc for cp, cc in self.caps.get(job.kind, []))

    def _set_events(self, kind: str):
        evs = []
        for s, job in enumerate(self.busy[kind]):
            if job is None:
                continue
            for e in job.events:
                evs.append((s,) + tuple(e))
        self.engines[kind].set_events(evs)

    def _mask(self, kind: str):
        return [1 if j is not None else 0 for j in self.busy[kind]]

    def run(self, jobs: Sequence[Job], clock: Callable[[], float] = time.perf_counter,
            on_finish: Optional[Callable[[Job], None]] = None, realtime: bool = True,
            poll_s: float = 0.0002, trace: Optional[list] = None, overlap: bool = True) -> dict:
        """Serve `jobs` (arrival times in seconds from the start).  With
        realtime=False every job is treated as already arrived.  `trace`, if
        given, receives one tuple per round: (world, launch time, substeps,
        jobs running, their particles, seconds from launch to read-back,
        jobs queued for that world).

        One host thread drives both worlds.  With overlap (the default) each
        world keeps a round in flight on its own stream, so the device runs a
        2D and a 3D round side by side (at 10 jobs/s, 100 jobs: 2.23e8
        particle-substeps/s against 2.04e8 taking turns, p50 latency 18 s
        against 27 s); without, one round is on the device at a time,
        alternating between the worlds that have running jobs.  Each round is
        enqueued, polled without blocking, read back, and jobs are admitted
        in between."""
        for j in jobs:
            if j.substeps + self.chunk > self.stream_cap:
                raise ValueError(f"job {j.jid}: {j.substeps} substeps exceed the H ring ({self.stream_cap})")
            if not self.fits_anywhere(j):
                raise ValueError(f"job {j.jid}: no partition can hold it")
        fps = {j.jid: _footprint(j) for j in jobs}
        pending = {k: deque(sorted((j for j in jobs if j.kind == k), key=lambda j: j.arrival))
                   for k in self.engines}
        queue = {k: [] for k in self.engines}
        inflight = {k: 0 for k in self.engines}      # substeps of the round on the device, 0 = idle
        rounds = {k: 0 for k in self.engines}
        busy_s = {k: 0.0 for k in self.engines}       # launch to read-back of each world's rounds
        work = {k: 0 for k in self.engines}           # particle-substeps its rounds advanced (running jobs)
        launched = {k: 0.0 for k in self.engines}
        t0 = clock()

        def finish_round(kind: str):
            eng = self.engines[kind]
            eng.sync()
            k = inflight[kind]
            inflight[kind] = 0
            rounds[kind] += 1
            busy_s[kind] += clock() - launched[kind]
            if trace is not None:
                running = [j for j in self.busy[kind] if j is not None]
                trace.append((kind, launched[kind] - t0, k, len(running), sum(j.n for j in running),
                              clock() - launched[kind], len(queue[kind])))
            retired = False
            for j in self.busy[kind]:
                if j is None:
                    continue
                j.done += k
                if j.done >= j.substeps:
                    st = eng.hash_stream(j.slot) if kind == "2d" else \
                        eng.hash_stream(j.slot, max_substeps=j.done)
                    j.stream = st[: j.substeps]
                    j.finished = clock() - t0
                    self.busy[kind][j.slot] = None
                    retired = True
                    if on_finish:
                        on_finish(j)
            if retired:
                eng.set_active(self._mask(kind))

        def admit_and_launch(kind: str):
            eng = self.engines[kind]
            admitted = False
            for job in list(queue[kind]):
                sl = self._find_slot(job, fps[job.jid])
                if sl is None:
                    continue
                eng.load_slot(sl, job.scene)
                self.busy[kind][sl] = job
                job.slot, job.admitted, job.done = sl, clock() - t0, 0
                queue[kind].remove(job)
                admitted = True
            if admitted:
                self._set_events(kind)
            running = [j for j in self.busy[kind] if j is not None]
            if running:
                k = min(self.chunk, max(j.substeps - j.done for j in running))
                launched[kind] = clock()
                eng.step(k)                 # enqueued; finish_round reads it back
                inflight[kind] = k
                work[kind] += k * sum(j.n for j in running)

        if overlap:
            # every world keeps a round in flight on its own stream
            while True:
                now = clock() - t0
                for kind in self.engines:
                    pend = pending[kind]
                    while pend and (not realtime or pend[0].arrival <= now):
                        queue[kind].append(pend.popleft())
                for kind in self.engines:
                    if inflight[kind] and not self.engines[kind].busy():
                        finish_round(kind)
                for kind in self.engines:
                    if not inflight[kind]:
                        admit_and_launch(kind)
                if any(inflight.values()):
                    time.sleep(poll_s)
                    continue
                if any(queue.values()):   # every partition is free here, and each job fits one
                    raise RuntimeError("queued jobs fit no free partition of an idle world")
                if not any(pending.values()):
                    break
                nxt = min(p[0].arrival for p in pending.values() if p)
                time.sleep(max(0.0, min(nxt - (clock() - t0), 0.005)))
            return {"wall_s": clock() - t0, "rounds": sum(rounds.values()), "rounds_by_world": rounds,
                    "busy_s_by_world": busy_s, "round_work_by_world": work}
        order = list(self.engines)
        turn = 0              # the world whose round is tried first next time
        current = None        # the world with a round on the device
        while True:
            now = clock() - t0
            for kind in self.engines:
                pend = pending[kind]
                while pend and (not realtime or pend[0].arrival <= now):
                    queue[kind].append(pend.popleft())
            if current is not None:
                if self.engines[current].busy():
                    time.sleep(poll_s)
                    continue
                finish_round(current)
                current = None
            for i in range(len(order)):
                kind = order[(turn + i) % len(order)]
                admit_and_launch(kind)
                if inflight[kind]:
                    current = kind
                    turn = (order.index(kind) + 1) % len(order)
                    break
            if current is None:
                if any(queue.values()):   # every partition is free here, and each job fits one
                    raise RuntimeError("queued jobs fit no free partition of an idle world")
                if not any(pending.values()):
                    break
                nxt = min(p[0].arrival for p in pending.values() if p)
                time.sleep(max(0.0, min(nxt - (clock() - t0), 0.005)))
        return {"wall_s": clock() - t0, "rounds": sum(rounds.values()), "rounds_by_world": rounds,
                "busy_s_by_world": busy_s, "round_work_by_world": work}


def summarize(jobs: Sequence[Job], wall: float) -> dict:
    lat = np.array([j.finished - j.arrival for j in jobs])
    wait = np.array([j.admitted - j.arrival for j in jobs])
    work = float(sum(j.work for j in jobs))
    span = max(j.finished for j in jobs) - min(j.arrival for j in jobs)
    return {
        "jobs": len(jobs),
        "jobs_2d": sum(1 for j in jobs if j.kind == "2d"),
        "jobs_3d": sum(1 for j in jobs if j.kind == "3d"),
        "particle_substeps": work,
        "throughput": work / span if span > 0 else 0.0,
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "p50_wait_s": float(np.percentile(wait, 50)),
        "p99_wait_s": float(np.percentile(wait, 99)),
        "wall_s": wall,
    }
