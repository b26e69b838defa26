"""Many independent KMC trajectories on one GPU: batched execution and a
request-serving loop with continuous batching.

The reference runs one trajectory per process (main.cpp:380-412).  Here
independent trajectories of one grid shape advance together through a device
batch (SimulationBatch: one launch sequence per event step for all of them),
and a server admits arriving requests by device-memory footprint:

* requests wait in arrival order until their estimated footprint
  (mgk_estimate_bytes) fits the HBM budget;
* an admitted request goes to the worker of its grid shape and precision; a
  set-up thread of that worker builds its simulation and runs the first
  (cold) LS solve on the simulation's own stream, beside the batch;
* the worker's batch has rows (slots).  A finished request's row is parked:
  inactive (no event cap), its simulation kept until a joining request takes
  the row (mgk_sim_batch_set_member), so membership changes keep the batch
  and its captured HIP graphs.  The batch is rebuilt only to grow, to drop
  parked rows when the admission loop waits for device memory, or when parked
  rows outnumber the running ones;
* workers of different shapes run in their own threads (the native calls
  release the interpreter lock), so their launches overlap on the GPU;
* a request completes when its trajectory has made its events (or has no
  rate left); its latency is completion time minus arrival time.
"""
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import abi, configs
from .solver import Simulation, SimulationBatch, estimate_bytes


@dataclass
class Request:
    rid: int
    shape: Tuple[int, ...]
    events: int
    seed: int
    arrival: float = 0.0               # seconds after the start of the run
    precision: int = abi.MGK_FP32_REFINE
    instability_kmc: int = 0           # instability_mode "kmc" (1) or "cascade" (0)
    device_inputs: bool = False        # E field and constructor draws generated on the GPU
    # filled by the server
    admitted: Optional[float] = None
    started: Optional[float] = None
    finished: Optional[float] = None
    events_done: int = 0
    footprint: int = 0
    stopped: bool = False              # ended short of its events: an attempt without rate (R not > 0)

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def request_simulation(req):
    """The trajectory of a request: the benchmark workloads' material and
    catalog (correlation length 3 voxels in 2D, 4 in 3D) with its own seed;
    with device_inputs the field and the constructor draws come from the GPU
    (the reference's distributions, Philox streams)."""
    corr = 3.0 if len(req.shape) == 2 else 4.0
    cfg = configs.sim_config(req.shape, precision=req.precision, seed=req.seed, instability_kmc=req.instability_kmc,
                             device_draws=int(req.device_inputs))
    E, nu = configs.material(req.shape, corr, req.seed, device=req.device_inputs)
    return Simulation(cfg, E, nu)


def request_footprint(req):
    M = 6 if len(req.shape) == 2 else 12
    return estimate_bytes(req.shape, M, per_voxel_catalog=False)


class _Slot:
    """One row of a worker's batch: a running request, or a parked one whose
    simulation stays until the row is reused or dropped."""
    __slots__ = ("req", "sim", "parked")

    def __init__(self, req, sim):
        self.req, self.sim, self.parked = req, sim, False


class _Worker(threading.Thread):
    """The trajectories of one grid shape and precision."""

    def __init__(self, server, key):
        super().__init__(daemon=True)
        self.server = server
        self.key = key
        self.inbox: List[Request] = []                        # admitted, not yet set up
        self.ready: List[Tuple[Request, Simulation]] = []     # set up, waiting for a row
        self.slots: List[_Slot] = []
        self.batch: Optional[SimulationBatch] = None
        self.cv = threading.Condition()
        self.error: Optional[BaseException] = None
        self.setting_up = 0
        self.rebuilds = 0       # batches built (growth, compaction)
        self.reuses = 0         # joins into a parked row
        self._setup = threading.Thread(target=self._setup_run, daemon=True)

    def submit(self, req):
        with self.cv:
            self.inbox.append(req)
            self.cv.notify_all()

    def start(self):
        super().start()
        self._setup.start()

    # -- set-up thread: simulation and first solve, off the batch's path ----
    def _setup_run(self):
        try:
            srv = self.server
            while True:
                with self.cv:
                    while not self.inbox and not srv._closing:
                        self.cv.wait(0.05)
                    if not self.inbox:
                        return
                    req = self.inbox.pop(0)
                    self.setting_up += 1
                try:
                    sim = request_simulation(req)
                    sim.elastic_run()
                except BaseException:
                    with self.cv:
                        self.setting_up -= 1
                    raise
                # one critical section: the worker never sees the request in
                # neither list
                with self.cv:
                    self.setting_up -= 1
                    self.ready.append((req, sim))
                    self.cv.notify_all()
        except BaseException as e:   # reported by Server.run
            self.error = e
            self.server._wake()

    def run(self):
        try:
            self._loop()
        except BaseException as e:   # reported by Server.run
            self.error = e
            self.server._wake()

    def _active(self):
        return [sl for sl in self.slots if not sl.parked]

    def _drop(self, slot):
        """Releases a parked row's simulation and its footprint."""
        slot.sim.close()
        self.server._release(slot.req)

    def _rebuild(self, joining):
        """A batch of the running rows plus `joining`; parked rows go."""
        for sl in self.slots:
            if sl.parked:
                self._drop(sl)
        self.slots = [sl for sl in self.slots if not sl.parked] + [_Slot(r, s) for r, s in joining]
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        if self.slots:
            self.batch = SimulationBatch([sl.sim for sl in self.slots])
            self.rebuilds += 1

    def _loop(self):
        srv = self.server
        while True:
            with self.cv:
                while (not self.ready and not self._active() and not (srv._closing and not self.inbox
                                                                        and not self.setting_up)):
                    if self.slots and srv.pressure:
                        break
                    self.cv.wait(0.05)
                new, self.ready = self.ready, []
                finished = srv._closing and not self.inbox and not self.setting_up and not new
            t = srv.now()
            for req, _ in new:
                req.started = t
            # joining requests take parked rows first; the rest grow the batch
            grow = []
            for req, sim in new:
                park = next((sl for sl in self.slots if sl.parked), None)
                if park is not None and self.batch is not None:
                    self.batch.set_member(self.slots.index(park), sim)
                    self._drop(park)
                    park.req, park.sim, park.parked = req, sim, False
                    self.reuses += 1
                else:
                    grow.append((req, sim))
            # rebuild to grow, to give parked rows' memory to a waiting
            # admission, or when parked rows outnumber the running ones (every
            # row's blocks are launched each step, inactive ones included)
            n_parked = sum(1 for sl in self.slots if sl.parked)
            if grow or (n_parked and (srv.pressure or n_parked > len(self.slots) - n_parked)):
                self._rebuild(grow)
            if not self._active():
                # nothing runs: parked rows give their memory back
                if self.slots:
                    self._rebuild([])
                if finished:
                    return
                continue
            caps = [0 if sl.parked else sl.req.events - sl.req.events_done for sl in self.slots]
            stats = self.batch.event_steps(srv.step_events, caps=caps)
            t = srv.now()
            for sl, st in zip(self.slots, stats):
                if sl.parked:
                    continue
                made = int(st.kmc_events)
                sl.req.events_done += made
                # every attempt makes an event, runs a cascade or finds no
                # rate: a step with neither events nor cascades has stopped
                stopped = made == 0 and int(st.cascades) == 0
                if sl.req.events_done >= sl.req.events or stopped:
                    sl.req.finished = t
                    sl.req.stopped = sl.req.events_done < sl.req.events
                    if srv.on_complete is not None:
                        srv.on_complete(sl.req, sl.sim)
                    sl.parked = True


class Server:
    """Admits requests by device-memory footprint and runs them batched by
    grid shape (see the module documentation)."""

    def __init__(self, hbm_budget_bytes, step_events=8, on_complete=None, progress=None, progress_s=5.0):
        """on_complete(request, simulation): called by the worker thread when
        a request finishes, before its simulation is released.
        progress(text): called every progress_s seconds while serving."""
        self.progress = progress
        self.progress_s = float(progress_s)
        self._last_progress = 0.0
        self._requests = []
        self.budget = int(hbm_budget_bytes)
        self.step_events = int(step_events)
        self.on_complete = on_complete
        self.peak_used = 0
        self.used = 0
        self._lock = threading.Lock()
        self._wake_cv = threading.Condition(self._lock)
        self._closing = False
        self._t0 = None
        self.workers = {}
        self.pressure = False   # the admission loop waits for device memory: parked rows go

    def now(self):
        return time.perf_counter() - self._t0

    def _wake(self):
        with self._wake_cv:
            self._wake_cv.notify_all()

    def _release(self, req):
        with self._wake_cv:
            self.used -= req.footprint
            self._wake_cv.notify_all()

    def _worker(self, req):
        key = (tuple(req.shape), req.precision)
        w = self.workers.get(key)
        if w is None:
            w = _Worker(self, key)
            self.workers[key] = w
            w.start()
        return w

    def run(self, requests: List[Request]):
        """Serves `requests` (arrival times relative to the start) and
        returns them with admission, start and completion times filled."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        self._requests = list(requests)
        for r in pending:
            r.footprint = request_footprint(r)
            if r.footprint > self.budget:
                raise ValueError(f"request {r.rid} ({r.shape}) needs {r.footprint} bytes, over the budget")
        self._t0 = time.perf_counter()
        self._closing = False
        i = 0
        while i < len(pending):
            r = pending[i]
            wait = r.arrival - self.now()
            if wait > 0:
                time.sleep(min(wait, 0.01))
                self._check()
                continue
            with self._wake_cv:
                while self.used + r.footprint > self.budget:
                    self.pressure = True
                    self._wake_cv.wait(0.01)
                    self._check()
                self.pressure = False
                self.used += r.footprint
                self.peak_used = max(self.peak_used, self.used)
            r.admitted = self.now()
            self._worker(r).submit(r)
            i += 1
        self._closing = True
        for w in self.workers.values():
            with w.cv:
                w.cv.notify()
        for w in self.workers.values():
            while w.is_alive() or w._setup.is_alive():
                w.join(timeout=0.2)
                self._check()
        self._check()
        return requests

    def stats(self):
        """Batches built and joins into parked rows, over all workers."""
        return dict(rebuilds=sum(w.rebuilds for w in self.workers.values()),
                    reuses=sum(w.reuses for w in self.workers.values()))

    def _report(self):
        if self.progress is None or self.now() - self._last_progress < self.progress_s:
            return
        self._last_progress = self.now()
        done = sum(1 for r in self._requests if r.finished is not None)
        ev = sum(r.events_done for r in self._requests)
        self.progress(f"t={self.now():.1f}s admitted={sum(1 for r in self._requests if r.admitted is not None)} "
                      f"done={done}/{len(self._requests)} events={ev} used={self.used / 2**30:.2f}GiB")

    def _check(self):
        self._report()
        for w in self.workers.values():
            if w.error is not None:
                self._closing = True
                raise RuntimeError(f"worker {w.key} failed") from w.error


def summary(requests, wall_s=None):
    """p50/p99 completion latency (arrival to result) and aggregate events/s."""
    lat = np.array([r.latency for r in requests if r.latency is not None])
    ev = sum(r.events_done for r in requests)
    t_end = max((r.finished for r in requests if r.finished is not None), default=0.0)
    t_start = min((r.arrival for r in requests), default=0.0)
    span = wall_s if wall_s is not None else (t_end - t_start)
    return dict(requests=len(requests), completed=int(lat.size), events=int(ev),
                stopped_short=sum(1 for r in requests if r.stopped),
                p50_latency_s=float(np.percentile(lat, 50)) if lat.size else None,
                p99_latency_s=float(np.percentile(lat, 99)) if lat.size else None,
                events_per_s=ev / span if span > 0 else None, span_s=span)
