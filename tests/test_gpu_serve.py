"""Serving loop: requests arriving while others run are admitted by HBM
footprint, share launches, and each finishes with its standalone result."""
import pytest

pytestmark = pytest.mark.gpu

F, P, Cm = 1, 2, 4

REQS = [
    ((F, 10**9, 10**9 + 2 * 10**5, 101, 2500), 0.0),
    ((Cm, 101, 30000, 101, 2000), 0.0),
    ((P, 10**11, 10**11 + 10**5, 3000, 8000), 0.05),
    ((F | Cm, 10**12, 10**12 + 10**5, 101, 1500), 0.1),
    ((F, 3, 3000, 101, 2000), 0.1),
    ((P, 101, 10000, 101, 3000), 0.2),
    ((Cm, 10**13, 10**13 + 10**5, 40000, 42000), 0.25),
    ((F, 2**63 - 10**5, 2**63, 101, 900), 0.3),
]


@pytest.mark.parametrize("overlap", [True, False])
@pytest.mark.parametrize("round_updates", [0, 3 * 10**6])
def test_served_results_equal_standalone(orc, dev, round_updates, overlap):
    """Rounds of at most 400 steps per segment, or of about 3e6 updates
    (pfc_batcher_round_work: an equal share per segment); finished jobs
    harvested on worker threads beside the rounds, or between them."""
    import pfcsieve_amd.serve as S
    srv = S.Server(dev, hbm_budget=1 << 40, max_steps=400, verify_threads=8, round_updates=round_updates,
                   overlap=overlap)
    for k, (d, t) in enumerate(REQS):
        srv.submit(S.Request(k, d, t))
    st = srv.run()
    assert st.jobs == len(REQS) and len(st.latencies) == len(REQS)
    assert st.rounds > 1
    for r in sorted(srv.done, key=lambda r: r.rid):
        ref = orc.run(*r.desc)
        got = r.result
        assert (got.checksum, got.primecount) == (ref.checksum, ref.primecount), r.desc
        assert got.factors == [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in ref.factors]
        assert r.finished >= r.arrival and r.admitted >= r.arrival - 1e-9


def test_admission_respects_budget(orc, dev):
    """With a budget that fits one job at a time, jobs run one after another."""
    import pfcsieve_amd as PK
    import pfcsieve_amd.serve as S
    sizes = [PK.footprint(dev, *d) for d, _ in REQS[:4]]
    srv = S.Server(dev, hbm_budget=max(sizes), max_steps=1000, verify_threads=8)
    for k, (d, t) in enumerate(REQS[:4]):
        srv.submit(S.Request(k, d, 0.0))
    st = srv.run()
    assert st.peak_resident_bytes <= max(sizes)
    for r in srv.done:
        ref = orc.run(*r.desc)
        assert (r.result.checksum, r.result.primecount) == (ref.checksum, ref.primecount)
