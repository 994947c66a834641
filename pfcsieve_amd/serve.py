"""Serving loop: continuous batching of sieve jobs on one MI355X.

Requests arrive over time (``submit``).  Between rounds of the device batcher
the loop admits the requests that have arrived, in arrival order, each one
whose HBM footprint (pfc_job_footprint: prime state + tables + factor
buffer) still fits the budget beside the resident jobs (a request that does
not fit yet waits; smaller ones behind it may go first); admitted jobs keep all their state on
the device and share every persistent-kernel launch with the jobs already
running.  A round walks at most ``max_steps`` multiplier steps per segment,
or, with ``round_updates``, about that many p*n updates over all resident
segments, an equal share per segment (pfc_batcher_round_work), which keeps
rounds short while large jobs are resident and lets a small job finish in a
round or a few beside them; either way newly admitted work starts within
one round.  Finished jobs are harvested on the host (verify,
isPrime filter, checksum) and their latency recorded.

Harvest (the CPU verification of a finished job's factor records, its
result, the release of its device memory) runs on worker threads while the
device works through the next rounds (the library releases the interpreter
in every call, and its device pool and errors are locked or per thread), so
that it no longer stalls every other job's round.  Job creation stays
between rounds, on the loop's own thread: its uploads and table check are
device work that would wait behind a running round (the persistent phase
kernels hold every wavefront slot), which on the worker threads delayed
each new job by about a round.
"""
from __future__ import annotations

import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field
from typing import Dict, List, Optional

from .sieve import Batcher, Device, Job, SieveResult, footprint, job_result


@dataclass
class Request:
    rid: int
    desc: tuple
    arrival: float                 # seconds since the server's start
    bytes: int = 0
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[SieveResult] = None


@dataclass
class ServeStats:
    jobs: int = 0
    wall_s: float = 0.0
    rounds: int = 0
    latencies: List[float] = field(default_factory=list)
    peak_resident_bytes: int = 0
    admit_s: float = 0.0           # job creation (tables, device buffers, table check)
    round_s: float = 0.0           # batcher rounds (device phases + the round's host work)
    harvest_s: float = 0.0         # verification and release of finished jobs, summed
    idle_s: float = 0.0            # the loop waiting with nothing to run
    overlap: bool = True           # harvest ran beside the rounds

    def pct(self, q: float) -> float:
        if not self.latencies:
            return float("nan")
        xs = sorted(self.latencies)
        k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
        return xs[k]


class Server:
    def __init__(self, dev: Device, hbm_budget: int, max_steps: int = 20000, verify_threads: int = 2,
                 progress=None, round_updates: int = 0, overlap: bool = True):
        self.progress = progress    # optional callable receiving one line per finished job
        self.dev = dev
        self.budget = hbm_budget
        self.max_steps = max_steps
        self.round_updates = round_updates
        self.verify_threads = verify_threads
        self.overlap = overlap      # False: harvest on the loop's thread, between rounds
        self.batcher = Batcher(dev)
        self.waiting: List[Request] = []
        self.active: Dict[int, tuple] = {}      # job handle -> (request, job)
        self.harvesting: List[tuple] = []       # (request, future of (result, finished, seconds))
        self.resident = 0
        self.done: List[Request] = []
        self.t0 = None

    def submit(self, req: Request):
        req.bytes = footprint(self.dev, *req.desc)
        if req.bytes > self.budget:
            raise ValueError(f"request {req.rid} needs {req.bytes} bytes, over the budget")
        self.waiting.append(req)

    def _now(self) -> float:
        return time.perf_counter() - self.t0

    def _harvest(self, j: Job):
        """Result of a finished job (ready at `finished`), then its release;
        the release can wait for a running round (a device free
        synchronises), so it comes after the result's time."""
        t = time.perf_counter()
        try:
            res = job_result(j, self.verify_threads)
            finished = self._now()
        finally:
            j.destroy()
        return res, finished, time.perf_counter() - t

    def _admit(self, st: ServeStats):
        t = self._now()
        keep = []
        for r in self.waiting:
            if r.arrival <= t and self.resident + r.bytes <= self.budget:
                job = Job(self.dev, *r.desc)
                self.batcher.add(job)
                self.active[job.handle.value] = (r, job)
                self.resident += r.bytes
                r.admitted = t
            else:
                keep.append(r)
        self.waiting = keep
        st.admit_s += self._now() - t

    def _finish(self, st: ServeStats, r: Request, harvested):
        r.result, r.finished, dt = harvested
        st.harvest_s += dt
        self.resident -= r.bytes
        self.done.append(r)
        st.latencies.append(r.finished - r.arrival)
        st.jobs += 1
        if self.progress:
            self.progress(f"served job {r.rid} ({st.jobs} done) at {r.finished:.1f} s, "
                          f"latency {r.finished - r.arrival:.2f} s")

    def _collect(self, st: ServeStats):
        still = []
        for r, f in self.harvesting:
            if f.done():
                self._finish(st, r, f.result())
            else:
                still.append((r, f))
        self.harvesting = still

    def run(self, stats: Optional[ServeStats] = None) -> ServeStats:
        st = stats or ServeStats()
        st.overlap = self.overlap
        self.t0 = time.perf_counter()
        self.waiting.sort(key=lambda r: r.arrival)
        self.pool = ThreadPoolExecutor(max_workers=2) if self.overlap else None
        try:
            while self.waiting or self.active or self.harvesting:
                self._admit(st)
                self._collect(st)
                st.peak_resident_bytes = max(st.peak_resident_bytes, self.resident)
                if not self.active:
                    # nothing to run: wait for a harvest (it frees budget)
                    # or for the next arrival
                    t = self._now()
                    futs = [f for _, f in self.harvesting]
                    nap = max(0.0, min(0.05, self.waiting[0].arrival - t)) if self.waiting else 0.05
                    if futs:
                        wait(futs, timeout=nap, return_when=FIRST_COMPLETED)
                    elif self.waiting:
                        time.sleep(nap)
                    st.idle_s += self._now() - t
                    continue
                t = self._now()
                if self.round_updates:
                    finished = self.batcher.round_work(self.round_updates)
                else:
                    finished = self.batcher.round(self.max_steps)
                st.rounds += 1
                st.round_s += self._now() - t
                for job in finished:
                    r, j = self.active.pop(job.handle.value)
                    if self.overlap:
                        self.harvesting.append((r, self.pool.submit(self._harvest, j)))
                    else:
                        self._finish(st, r, self._harvest(j))
        finally:
            if self.pool is not None:
                self.pool.shutdown(wait=True)
        st.wall_s = self._now()
        self.batcher.destroy()
        return st
