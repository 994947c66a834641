"""Walk-completion guard (the reference's check.cl:62-73 test, reported as
cl_sieve.cpp:476-480 does): every wavefront counts the multiplier steps it
takes and the drains compare the total with candidates x planned steps.  A
walk cut short on the host (pfc_job_debug_shorten_walk: the last launch of
each segment ends early, no fault involved) must fail the job with
PFC_ERR_VALIDATION in every execution form; the same job without the cut
passes and matches the oracle."""
import pytest

pytestmark = pytest.mark.gpu

F, P, Cm = 1, 2, 4
JOBS = [
    (F, 10**9, 10**9 + 10**5, 101, 3000),
    (P, 10**12, 10**12 + 10**5, 2000, 9000),          # 9000 is composite
    (Cm, 10**7, 10**7 + 10**5, 5000, 7000),
    (F | Cm, 140000, 400000, 65000, 66100),           # product steps, then pairs
]


def _run(dev, desc, form, short):
    import pfcsieve_amd as PK
    from pfcsieve_amd._native import check, lib
    dev.set_tuning(use_graph=1 if form == "graph" else 0, walk_chunk=700)
    try:
        j = PK.Job(dev, *desc)
        try:
            check(lib().pfc_job_debug_shorten_walk(j.handle, short))
            if form == "batched":
                b = PK.Batcher(dev)
                try:
                    b.add(j)
                    while b.pending():
                        b.round(max_steps=900)
                finally:
                    b.destroy()
            else:
                j.run()
            return PK.job_result(j, verify_threads=8)
        finally:
            j.destroy()
    finally:
        dev.set_tuning()


@pytest.mark.parametrize("form", ["plain", "graph", "batched"])
@pytest.mark.parametrize("desc", JOBS)
def test_shortened_walk_fails_validation(orc, dev, desc, form):
    from pfcsieve_amd._native import PfcError
    with pytest.raises(PfcError, match="gpu validation failure") as ei:
        _run(dev, desc, form, 1)
    assert ei.value.code == -5                         # PFC_ERR_VALIDATION
    got = _run(dev, desc, form, 0)
    ref = orc.run(*desc)
    want = (ref.checksum, ref.primecount, [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in ref.factors])
    assert (got.checksum, got.primecount, got.factors) == want
