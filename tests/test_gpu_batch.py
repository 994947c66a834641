"""Batched and continuously batched execution: every job's result equals the
oracle's (and therefore running it alone), whatever it shares launches with."""
import pytest

pytestmark = pytest.mark.gpu

F, P, Cm = 1, 2, 4
R = 1 << 64

MIX = [
    (F, 10**9, 10**9 + 3 * 10**5, 101, 3000),
    (P, 10**12, 10**12 + 2 * 10**5, 2000, 9000),
    (Cm, 101, 40000, 101, 2500),
    (F | Cm, (1 << 62) - 10**5, 1 << 62, 500, 1800),
    (F, R - 10**5, R - 1, 101, 700),
    (P, 101, 20000, 101, 4000),
    (Cm, 10**7, 10**7 + 10**5, 5000, 7000),
    (F, 3, 500, 101, 400),
    (F | Cm, 140000, 400000, 65000, 66100),   # n across 65,536: product steps, then pairs
]


def _ref(orc, job):
    r = orc.run(*job)
    return r.checksum, r.primecount, [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in r.factors]


@pytest.mark.parametrize("max_steps", [0, 700])
def test_batch_matches_oracle(orc, dev, max_steps):
    import pfcsieve_amd as PK
    got = PK.run_batch(MIX, device=dev, max_steps=max_steps, verify_threads=8)
    for job, g in zip(MIX, got):
        cs, pc, fac = _ref(orc, job)
        assert (g.checksum, g.primecount, g.factors) == (cs, pc, fac), job


def test_batch_admits_by_memory_budget(orc, dev):
    """A job list larger than the memory budget runs in one pass: with room
    for about two jobs at a time, jobs are admitted as others finish, and
    every result (returned in input order) is the oracle's."""
    import pfcsieve_amd as PK
    sizes = [PK.footprint(dev, *d) for d in MIX]
    budget = max(sizes) + min(sizes)
    seen = []
    got = PK.run_batch(MIX, device=dev, max_steps=900, verify_threads=8, hbm_budget=budget,
                       progress=seen.append)
    assert len(seen) > len(MIX) // 2
    for job, g in zip(MIX, got):
        cs, pc, fac = _ref(orc, job)
        assert (g.checksum, g.primecount, g.factors) == (cs, pc, fac), job
    with pytest.raises(ValueError):
        PK.run_batch(MIX[:1], device=dev, hbm_budget=sizes[0] - 1)


@pytest.mark.parametrize("ppl", [2, 4])
def test_batch_wider_lanes(orc, dev, ppl):
    """The batch kernel with 2 and 4 candidates per lane (k_batch<2>, <4>)
    gives every job the oracle's result."""
    import pfcsieve_amd as PK
    dev.set_tuning(primes_per_lane=ppl)
    try:
        got = PK.run_batch(MIX, device=dev, max_steps=1100, verify_threads=8)
    finally:
        dev.set_tuning()
    for job, g in zip(MIX, got):
        cs, pc, fac = _ref(orc, job)
        assert (g.checksum, g.primecount, g.factors) == (cs, pc, fac), (ppl, job)


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("blocks,ppl", [(3, 1), (64, 1), (5, 2), (7, 4)])
def test_batch_multi_block_walk_items(orc, dev, graph, blocks, ppl):
    """Walk items of several blocks of 64 * K candidates (tuning.
    batch_walk_blocks; short rounds of a serving loop take them by
    themselves): every job's result is the oracle's, with ragged last items
    and blocks past a segment's last candidate."""
    import pfcsieve_amd as PK
    with pytest.raises(PK.PfcError):
        dev.set_tuning(batch_walk_blocks=65)
    dev.set_tuning(batch_walk_blocks=blocks, use_graph=graph, primes_per_lane=ppl)
    try:
        got = PK.run_batch(MIX, device=dev, max_steps=300, verify_threads=8)
    finally:
        dev.set_tuning()
    for job, g in zip(MIX, got):
        cs, pc, fac = _ref(orc, job)
        assert (g.checksum, g.primecount, g.factors) == (cs, pc, fac), (blocks, ppl, graph, job)


def test_batch_multi_segment_jobs(orc, dev):
    import pfcsieve_amd as PK
    dev.set_tuning(seg_candidates=1000)
    try:
        got = PK.run_batch(MIX[:4], device=dev, max_steps=1500, verify_threads=8)
    finally:
        dev.set_tuning()
    for job, g in zip(MIX[:4], got):
        cs, pc, fac = _ref(orc, job)
        assert (g.checksum, g.primecount, g.factors) == (cs, pc, fac), job
        assert g.segments > 1


def test_continuous_admission(orc, dev):
    """Jobs added while others are mid-walk finish with their standalone result."""
    import pfcsieve_amd as PK
    b = PK.Batcher(dev)
    jobs = [PK.Job(dev, *d) for d in MIX]
    done = []
    try:
        for k, j in enumerate(jobs):
            b.add(j)
            done += b.round(max_steps=500)      # admit one job per round
        while b.pending():
            done += b.round(max_steps=500)
        assert len(done) == len(jobs)
        for d, j in zip(MIX, jobs):
            g = PK.job_result(j, verify_threads=8)
            cs, pc, fac = _ref(orc, d)
            assert (g.checksum, g.primecount, g.factors) == (cs, pc, fac), d
    finally:
        b.destroy()
        for j in jobs:
            j.destroy()


def test_round_work_matches_oracle_and_rejects_zero(orc, dev):
    """Rounds sized by work (pfc_batcher_round_work: steps = budget / resident
    candidates, at least 256) give every job the oracle's result; a zero
    budget is an invalid argument."""
    import pfcsieve_amd as PK
    from pfcsieve_amd._native import PfcError
    b = PK.Batcher(dev)
    jobs = [PK.Job(dev, *d) for d in MIX]
    try:
        with pytest.raises(PfcError, match="max_updates"):
            b.round_work(0)
        for j in jobs:
            b.add(j)
        done, rounds = [], 0
        while b.pending():
            done += b.round_work(2 * 10**7)
            rounds += 1
        assert len(done) == len(jobs) and rounds > 2
        for d, j in zip(MIX, jobs):
            g = PK.job_result(j, verify_threads=8)
            cs, pc, fac = _ref(orc, d)
            assert (g.checksum, g.primecount, g.factors) == (cs, pc, fac), d
    finally:
        b.destroy()
        for j in jobs:
            j.destroy()


def test_batch_returns_jobs_with_nothing_left(orc, dev):
    """A job whose position is already at pmax (an empty p-range, or a job
    that has been run to the end) is handed back by the next round, alone or
    beside running jobs."""
    import pfcsieve_amd as PK
    empty = (F, 10**9, 10**9, 101, 500)
    got = PK.run_batch([empty], device=dev, verify_threads=2)
    assert (got[0].checksum, got[0].primecount, got[0].factorcount) == (0, 0, 0)
    done_job = PK.Job(dev, *MIX[0])
    done_job.run()
    b = PK.Batcher(dev)
    try:
        b.add(done_job)
        assert b.round() == [done_job] and b.pending() == 0
    finally:
        b.destroy()
        done_job.destroy()
    got = PK.run_batch([empty, MIX[2]], device=dev, verify_threads=8)
    cs, pc, fac = _ref(orc, MIX[2])
    assert (got[1].checksum, got[1].primecount, got[1].factors) == (cs, pc, fac)


def test_wide_composite_range_walks_bitmap(orc, dev):
    """A C job whose n-range is wider than 2^27 carries no composite list
    (only jobs up to that width get one): its walk scans the composite bitmap,
    alone and in the batch kernel, and both give the oracle's result."""
    import pfcsieve_amd as PK
    job = (Cm, 10**12, 10**12 + 600, 101, 101 + (1 << 27) + 4000)
    ref = orc.run(*job)
    want = (ref.checksum, ref.primecount, [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in ref.factors])
    assert ref.primecount > 10
    plain = PK.run_sieve(*job, device=dev, verify_threads=8)
    assert (plain.checksum, plain.primecount, plain.factors) == want
    [bat] = PK.run_batch([job], device=dev, verify_threads=8)
    assert (bat.checksum, bat.primecount, bat.factors) == want


@pytest.mark.parametrize("ppl", [1, 2])
def test_graph_captured_rounds_match_eager_and_oracle(orc, dev, ppl):
    """Batcher rounds replayed from one captured graph (tuning.use_graph):
    many short rounds (each a replay, with the rounds' own eager generation
    launches and host synchronisations between replays), jobs added between
    rounds (the slot count changes, so the graph is captured again), and a
    second batch through the same device: every job's result equals the eager
    batcher's and the oracle's."""
    import pfcsieve_amd as PK
    results = {}
    for graph in (0, 1):
        dev.set_tuning(use_graph=graph, primes_per_lane=ppl)
        try:
            b = PK.Batcher(dev)
            jobs = [PK.Job(dev, *d) for d in MIX]
            try:
                rounds = 0
                for k, j in enumerate(jobs):
                    b.add(j)
                    if k % 3 == 2:
                        b.round(max_steps=150)
                        rounds += 1
                while b.pending():
                    b.round(max_steps=150)
                    rounds += 1
                assert rounds > 10
                results[graph] = [PK.job_result(j, verify_threads=8) for j in jobs]
            finally:
                b.destroy()
                for j in jobs:
                    j.destroy()
        finally:
            dev.set_tuning()
    for d, e, g in zip(MIX, results[0], results[1]):
        assert (g.checksum, g.primecount, g.factors) == (e.checksum, e.primecount, e.factors), d
        cs, pc, fac = _ref(orc, d)
        assert (g.checksum, g.primecount, g.factors) == (cs, pc, fac), d


def test_round_work_small_job_finishes_beside_a_large_one(orc, dev):
    """pfc_batcher_round_work shares a round's updates equally among the
    segments: beside a job of ~1.45e6 candidates that needs three rounds of
    1.2e10 updates, a job of ~360 candidates takes all of its 19,899 steps in
    the first round (dividing the round by candidates would have given it
    ~8,300 steps a round, three rounds).  Both results are the oracle's."""
    import pfcsieve_amd as PK
    from pfcsieve_amd.sieve import Batcher, Job, job_result
    big = (F, 10**9, 10**9 + 3 * 10**7, 101, 20000)
    small = (F, 10**12, 10**12 + 10**4, 101, 20000)
    jb, js = Job(dev, *big), Job(dev, *small)
    b = Batcher(dev)
    try:
        b.add(jb)
        b.add(js)
        first = b.round_work(12 * 10**9)
        assert [j.handle.value for j in first] == [js.handle.value]
        rounds = 1
        while not b.round_work(12 * 10**9):
            rounds += 1
            assert rounds < 10
        assert rounds + 1 >= 3
        for job, j in ((small, js), (big, jb)):
            r = job_result(j, 8)
            ref = orc.run(*job)
            assert (r.checksum, r.primecount, r.factors) == \
                (ref.checksum, ref.primecount, [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in ref.factors]), job
    finally:
        b.destroy()
        jb.destroy()
        js.destroy()
