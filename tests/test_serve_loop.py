"""The serving loop's host logic on the CPU: admission in arrival order within
the HBM budget, rounds until every job is done, harvest between rounds or on
worker threads beside them, latency bookkeeping.  The device side (jobs, the
batcher, the harvest) is replaced by stand-ins that keep the same calling
contract; the GPU tests run the loop over the real library."""
import threading
import time

import pytest

import pfcsieve_amd.serve as S


class _Handle:
    def __init__(self, v):
        self.value = v


class FakeJob:
    made = []
    lock = threading.Lock()

    def __init__(self, dev, rounds, nbytes):
        self.rounds_left = rounds
        self.nbytes = nbytes
        with FakeJob.lock:
            FakeJob.made.append(self)
            self.handle = _Handle(len(FakeJob.made))
        self.destroyed = False
        self.harvest_thread = None

    def destroy(self):
        assert not self.destroyed
        self.destroyed = True


class FakeBatcher:
    def __init__(self, dev):
        self.jobs = []
        self.calls = []

    def add(self, job):
        self.jobs.append(job)

    def _round(self):
        time.sleep(0.002)                  # the device's share of a round
        done = []
        for j in self.jobs:
            j.rounds_left -= 1
            if j.rounds_left <= 0:
                done.append(j)
        self.jobs = [j for j in self.jobs if j not in done]
        return done

    def round(self, max_steps):
        self.calls.append(("steps", max_steps))
        return self._round()

    def round_work(self, max_updates):
        self.calls.append(("work", max_updates))
        return self._round()

    def destroy(self):
        assert not self.jobs


def _fake_result(job, verify_threads):
    job.harvest_thread = threading.current_thread().name
    time.sleep(0.003)
    return ("result", job.handle.value, verify_threads)


@pytest.fixture
def fakes(monkeypatch):
    FakeJob.made = []
    monkeypatch.setattr(S, "Job", FakeJob)
    monkeypatch.setattr(S, "Batcher", FakeBatcher)
    monkeypatch.setattr(S, "footprint", lambda dev, rounds, nbytes: nbytes)
    monkeypatch.setattr(S, "job_result", _fake_result)


# (rounds of work, bytes, arrival s)
REQS = [(3, 40, 0.0), (1, 30, 0.0), (5, 50, 0.01), (2, 20, 0.02), (1, 60, 0.03), (4, 10, 0.03),
        (2, 30, 0.05), (1, 5, 0.06)]


@pytest.mark.parametrize("overlap", [True, False])
@pytest.mark.parametrize("round_updates", [0, 10**9])
def test_every_request_served_within_budget(fakes, overlap, round_updates):
    srv = S.Server(None, hbm_budget=100, max_steps=123, verify_threads=3, round_updates=round_updates,
                   overlap=overlap)
    for k, (rounds, nbytes, t) in enumerate(REQS):
        srv.submit(S.Request(k, (rounds, nbytes), t))
    st = srv.run()
    assert st.jobs == len(REQS) and sorted(r.rid for r in srv.done) == list(range(len(REQS)))
    assert st.peak_resident_bytes <= 100 and srv.resident == 0
    assert all(j.destroyed for j in FakeJob.made) and len(FakeJob.made) == len(REQS)
    for r in srv.done:
        assert r.result == ("result", r.result[1], 3)
        assert r.arrival - 1e-9 <= r.admitted <= r.finished
        assert r.finished - r.arrival in st.latencies
    kinds = {c[0] for c in srv.batcher.calls}
    assert kinds == ({"work"} if round_updates else {"steps"})
    assert all(c[1] == (round_updates or 123) for c in srv.batcher.calls)
    threads = {j.harvest_thread for j in FakeJob.made}
    main = threading.main_thread().name
    assert (main not in threads) if overlap else (threads == {main})
    assert st.overlap == overlap
    assert st.round_s > 0 and st.rounds == len(srv.batcher.calls)


def test_request_over_budget_rejected(fakes):
    srv = S.Server(None, hbm_budget=100)
    with pytest.raises(ValueError):
        srv.submit(S.Request(0, (1, 101), 0.0))


def test_admission_waits_for_harvested_bytes(fakes):
    """With room for one job at a time, each job is admitted only after the
    previous one's harvest has returned its bytes."""
    srv = S.Server(None, hbm_budget=60, verify_threads=1)
    for k in range(4):
        srv.submit(S.Request(k, (2, 60), 0.0))
    st = srv.run()
    assert st.peak_resident_bytes == 60
    done = sorted(srv.done, key=lambda r: r.rid)
    for a, b in zip(done, done[1:]):
        assert b.admitted >= a.finished
