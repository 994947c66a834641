"""Serving loop: continuous batching of sieve jobs on one MI355X.

Requests arrive over time (``submit``).  Between rounds of the device batcher
the loop admits the requests that have arrived, in arrival order, each one
whose HBM footprint (pfc_job_footprint: prime state + tables + factor
buffer) still fits the budget beside the resident jobs (a request that does
not fit yet waits; smaller ones behind it may go first); admitted jobs keep all their state on
the device and share every persistent-kernel launch with the jobs already
running.  A round walks at most ``max_steps`` multiplier steps per segment,
or, with ``round_updates``, the step count that gives about that many p*n
updates over all resident segments (pfc_batcher_round_work), which keeps
rounds short while large jobs are resident; either way newly admitted work
starts within one round.  Finished jobs are harvested on the host (verify,
isPrime filter, checksum) and their latency recorded.
"""
from __future__ import annotations

import time
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
    harvest_s: float = 0.0         # verification and release of finished jobs
    idle_s: float = 0.0            # waiting for arrivals with nothing resident

    def pct(self, q: float) -> float:
        if not self.latencies:
            return float("nan")
        xs = sorted(self.latencies)
        k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
        return xs[k]


class Server:
    def __init__(self, dev: Device, hbm_budget: int, max_steps: int = 20000, verify_threads: int = 2,
                 progress=None, round_updates: int = 0):
        self.progress = progress    # optional callable receiving one line per finished job
        self.dev = dev
        self.budget = hbm_budget
        self.max_steps = max_steps
        self.round_updates = round_updates
        self.verify_threads = verify_threads
        self.batcher = Batcher(dev)
        self.waiting: List[Request] = []
        self.active: Dict[int, tuple] = {}      # job handle -> (request, job)
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

    def _admit(self):
        now = self._now()
        keep = []
        for r in self.waiting:
            if r.arrival <= now and self.resident + r.bytes <= self.budget:
                job = Job(self.dev, *r.desc)
                self.batcher.add(job)
                self.active[job.handle.value] = (r, job)
                self.resident += r.bytes
                r.admitted = now
            else:
                keep.append(r)
        self.waiting = keep

    def run(self, stats: Optional[ServeStats] = None) -> ServeStats:
        st = stats or ServeStats()
        self.t0 = time.perf_counter()
        self.waiting.sort(key=lambda r: r.arrival)
        while self.waiting or self.active:
            t = self._now()
            self._admit()
            now = self._now()
            st.admit_s += now - t
            st.peak_resident_bytes = max(st.peak_resident_bytes, self.resident)
            if not self.active:
                nxt = self.waiting[0].arrival
                time.sleep(max(0.0, min(0.05, nxt - now)))
                st.idle_s += self._now() - now
                continue
            if self.round_updates:
                finished = self.batcher.round_work(self.round_updates)
            else:
                finished = self.batcher.round(self.max_steps)
            st.rounds += 1
            t, now = now, self._now()
            st.round_s += now - t
            for job in finished:
                r, j = self.active.pop(job.handle.value)
                r.result = job_result(j, self.verify_threads)
                r.finished = now
                self.resident -= r.bytes
                j.destroy()
                self.done.append(r)
                st.latencies.append(r.finished - r.arrival)
                st.jobs += 1
                if self.progress:
                    self.progress(f"served job {r.rid} ({st.jobs} done) at {now:.1f} s, "
                                  f"latency {r.finished - r.arrival:.2f} s")
            st.harvest_s += self._now() - now
        st.wall_s = self._now()
        self.batcher.destroy()
        return st
