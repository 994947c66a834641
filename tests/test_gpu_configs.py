"""The benchmark configurations themselves on the GPU.

* the reference's CPU-runnable case, whole job: -! n [101, 5000),
  p [1e9, 1.001e9), against the oracle;
* the seeded 256-unit uneven mix (splitmix64 seed 12345): every unit run
  plain (one job after another) and batched (one batcher, rounds sized by
  work) with identical checksums, candidate counts and raw factor records,
  and every unit small enough for the oracle against it;
* a prefix of the seeded served stream (splitmix64 seed 2024) through the
  serving loop, each result equal to the same job run alone.
"""
import math

import pytest

pytestmark = pytest.mark.gpu


def _oracle(orc, desc):
    r = orc.run(*desc)
    return r.checksum, r.primecount, [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in r.factors]


def test_factorial_n5000_p1e9_whole_job(orc, dev):
    import pfcsieve_amd as PK
    desc = (1, 10**9, 10**9 + 10**6, 101, 5000)
    got = PK.run_sieve(*desc, device=dev, verify_threads=8)
    assert got.primecount == 48155
    assert (got.checksum, got.primecount, got.factors) == _oracle(orc, desc)
    assert got.factors_txt().endswith("%016X\n" % got.checksum)


def _updates(u):
    from pfcsieve_amd import workload as W
    return (u.pmax - u.pmin) / math.log(u.pmin) * W.walk_steps(u)


@pytest.fixture(scope="module")
def mix256(dev):
    """All 256 seeded units, plain and batched, raw results per unit."""
    import pfcsieve_amd as PK
    from pfcsieve_amd import workload as W
    units = W.mix(256, seed=12345)
    dev.set_tuning(seg_candidates=1 << 26)       # every unit one p-segment, as the benchmark runs it
    try:
        plain = []
        for u in units:
            j = PK.Job(dev, *u.desc())
            try:
                j.run()
                plain.append(j.result())
            finally:
                j.destroy()
        print(f"mix256: plain done ({len(plain)} units)", flush=True)
        jobs = [PK.Job(dev, *u.desc()) for u in units]
        b = PK.Batcher(dev)
        try:
            for j in jobs:
                b.add(j)
            left, rounds, K = sum(_updates(u) for u in units), 0, 6
            while b.pending():
                if rounds >= K - 1:
                    b.round(0)
                else:
                    budget = left / (K - rounds)
                    b.round_work(int(budget))
                    left -= budget
                rounds += 1
            batched = [j.result() for j in jobs]
        finally:
            b.destroy()
            for j in jobs:
                j.destroy()
        print(f"mix256: batched done in {rounds} rounds", flush=True)
    finally:
        dev.set_tuning()
    return units, plain, batched, rounds


def test_mix256_batched_equals_plain(mix256):
    units, plain, batched, rounds = mix256
    assert len(plain) == len(batched) == 256 and rounds == 6
    for u, a, b in zip(units, plain, batched):
        assert a == b, u
    assert sum(r[1] for r in plain) > 10**9          # ~1.35e9 candidates over the mix


def test_mix256_small_units_match_oracle(orc, mix256):
    """Every unit of the mix whose walk the oracle finishes in a few seconds
    (p-width / ln p x steps plus twice the set-up products below 3e9)."""
    import pfcsieve_amd as PK
    units, plain, _, _ = mix256
    checked = 0
    for u, got in zip(units, plain):
        cost = _updates(u) + 2 * (u.pmax - u.pmin) / math.log(u.pmin) * PK.setup_table_cost(
            {1: 0, 2: 1, 4: 2}[u.modes], u.nmin - 1)
        if cost > 3e9:
            continue
        assert got == _oracle(orc, u.desc()), u
        checked += 1
    assert checked >= 18


def test_served_stream_prefix_equals_standalone(dev):
    """The first 16 jobs of the seeded 10,000-job stream arrive as a Poisson
    process faster than the device finishes them, are admitted by HBM
    footprint and share batcher rounds; each result equals the job run alone."""
    import pfcsieve_amd as PK
    import pfcsieve_amd.serve as S
    from pfcsieve_amd import workload as W
    units = W.mix(10_000, seed=2024)[:16]
    arrivals = W.poisson_arrivals(len(units), rate=20.0, seed=777)
    srv = S.Server(dev, hbm_budget=int(dev.info()["hbm_bytes"] * 0.8), verify_threads=8,
                   round_updates=2 * 10**11)
    for k, (u, a) in enumerate(zip(units, arrivals)):
        srv.submit(S.Request(k, u.desc(), a))
    st = srv.run()
    assert st.jobs == len(units) and st.rounds > 1
    for r in sorted(srv.done, key=lambda r: r.rid):
        alone = PK.run_sieve(*r.desc, device=dev, verify_threads=8)
        got = r.result
        assert (got.checksum, got.primecount, got.factors) == (alone.checksum, alone.primecount, alone.factors), r.desc


@pytest.mark.parametrize("desc", [
    (5, 10**12, 10**12 + 300_000, 101, 10**6),                 # -! -c, the headline's n-range and p
    (2, 2**62 - 4 * 10**9, 2**62 - 4 * 10**9 + 400_000, 101, 10**7),   # -#, the primorial job's n-range and p
])
def test_full_n_range_job_every_candidate(orc, dev, desc):
    """The benchmark configurations' full n-ranges through the job path the
    benchmark times (composite list, product steps below n = 65,536, paired
    and, near 2^62, triple steps; launches of 2^18 steps) and through the batch
    kernel, on p-windows of about 10^4 candidates: checksum, candidate count
    and factor records of every candidate against the oracle."""
    import pfcsieve_amd as PK
    want = _oracle(orc, desc)
    assert want[1] > 8000
    got = PK.run_sieve(*desc, device=dev, verify_threads=8)
    assert (got.checksum, got.primecount, got.factors) == want
    [bat] = PK.run_batch([desc], device=dev, max_steps=300_000, verify_threads=8)
    assert (bat.checksum, bat.primecount, bat.factors) == want
