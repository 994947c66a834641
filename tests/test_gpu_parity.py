"""GPU parity: the HIP path through the C ABI against the CPU oracle and the
reference's golden self-test vectors.  Bit-exact: candidate sets, factor
records, primecounts and checksums must be identical."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F, P, Cm = 1, 2, 4
TWO64 = 1 << 64


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def gen_on_gpu(lo, hi, cap=None):
    torch = _torch()
    import pfcsieve_amd._native as N
    span = hi - lo
    if cap is None:
        cap = int(2.2 * span / max(np.log(max(span, 3)), 1.0)) + 4096
    d_p = torch.zeros(cap, dtype=torch.int64, device="cuda")
    d_ctrl = torch.zeros(4, dtype=torch.int64, device="cuda")
    N.check(N.lib().pfc_stage_generate(lo, hi, _ptr(d_p), cap, _ptr(d_ctrl), None))
    torch.cuda.synchronize()
    ctrl = d_ctrl.cpu().numpy().view(np.uint32)
    n = int(ctrl[0])
    assert ctrl[2] & 1 == 0, "candidate overflow"
    return np.sort(d_p[:n].cpu().numpy().view(np.uint64)), d_p, d_ctrl, cap


@pytest.mark.parametrize("lo,hi", [
    (3, 100000),                       # primes below 114 plus the sieve
    (101, 114),                        # only addsmallprimes territory
    (114, 120),                        # empty-ish window
    (10**9, 10**9 + 10**6),            # cfg1 window: 48,155 candidates
    ((1 << 32) - 300000, (1 << 32) + 300000),
    ((1 << 62) - 4 * 10**6, 1 << 62),  # cfg3 carry edge, scaled down
    (TWO64 - 2 * 10**6, TWO64 - 1),    # top of the 64-bit range
])
def test_generate_matches_oracle(orc, lo, hi):
    got, *_ = gen_on_gpu(lo, hi)
    ref = orc.candidates(lo, hi)
    assert len(got) == len(ref)
    assert np.array_equal(got, np.sort(ref))


def test_generate_cfg1_count(orc):
    got, *_ = gen_on_gpu(10**9, 10**9 + 10**6)
    assert len(got) == 48155


SMALL_JOBS = [
    (F, 101, 30000, 101, 4000),
    (P, 101, 30000, 101, 4000),            # 4000 is composite
    (Cm, 101, 30000, 101, 4000),
    (F | Cm, 101, 30000, 101, 4000),
    (F, 3, 2000, 101, 3000),               # p below nmin: residues hit 0
    (F, 10**9, 10**9 + 10**5, 101, 5000),  # cfg1 slice
    (Cm, 10**12, 10**12 + 2 * 10**5, 101, 3000),
    (P, (1 << 62) - 10**5, 1 << 62, 101, 5000),
    (F | Cm, TWO64 - 10**5, TWO64 - 1, 101, 2000),
    (F, 10**6, 10**6 + 10**5, 2000, 9000),  # larger nmin: non-trivial setup
    (P, 10**7, 10**7 + 10**5, 5000, 12000),
    (Cm, 10**7, 10**7 + 10**5, 5000, 12000),
    (F | Cm, 10**7, 10**7 + 10**5, 5000, 9000),
]


@pytest.mark.parametrize("job", SMALL_JOBS)
def test_job_matches_oracle(orc, dev, job):
    import pfcsieve_amd as PK
    modes, pmin, pmax, nmin, nmax = job
    ref = orc.run(modes, pmin, pmax, nmin, nmax)
    got = PK.run_sieve(modes, pmin, pmax, nmin, nmax, device=dev, verify_threads=8)
    assert got.primecount == ref.primecount
    assert got.factorcount == ref.factorcount
    assert got.checksum == ref.checksum, f"{got.checksum:016X} != {ref.checksum:016X}"
    ref_f = [(int(r["p"]), int(r["nc"]), int(r["type"])) for r in ref.factors]
    assert got.factors == ref_f


def test_job_many_segments_same_result(orc, dev):
    """Splitting the p-range into many segments changes nothing."""
    import pfcsieve_amd as PK
    job = (F | Cm, 101, 200000, 101, 3000)
    ref = orc.run(*job)
    dev.set_tuning(seg_candidates=1000, walk_chunk=257, primes_per_lane=1)
    try:
        got = PK.run_sieve(*job, device=dev, verify_threads=8)
    finally:
        dev.set_tuning()
    assert got.segments > 10
    assert (got.checksum, got.primecount, got.factorcount) == (ref.checksum, ref.primecount, ref.factorcount)


@pytest.mark.parametrize("ppl", [1, 2, 4])
@pytest.mark.parametrize("block", [64, 256])
def test_launch_shapes_identical(orc, dev, ppl, block):
    import pfcsieve_amd as PK
    job = (F | Cm, 10**9, 10**9 + 3 * 10**5, 101, 2500)
    ref = orc.run(*job)
    dev.set_tuning(primes_per_lane=ppl, block_size=block, walk_chunk=1000)
    try:
        got = PK.run_sieve(*job, device=dev, verify_threads=8)
    finally:
        dev.set_tuning()
    assert (got.checksum, got.primecount, got.factorcount) == (ref.checksum, ref.primecount, ref.factorcount)


def test_primorial_prime_nmax_rejected(dev):
    import pfcsieve_amd as PK
    with pytest.raises(PK.PfcError) as e:
        PK.run_sieve(P, 10**6, 10**6 + 1000, 101, 4001, device=dev)   # 4001 is prime
    assert "CPU sieve failure" in str(e.value)


@pytest.mark.parametrize("case", list(range(1, 17)))
def test_reference_selftest_vector(dev, golden, case):
    """The reference's -s self-test (cl_sieve.cpp:1777-2095), all 16 ranges."""
    import pfcsieve_amd as PK
    g = golden[case - 1]
    print(f"selftest case {case}: modes={g['modes']} p=[{g['pmin']},{g['pmax']}) n=[{g['nmin']},{g['nmax']})",
          flush=True)
    got = PK.run_sieve(g["modes"], g["pmin"], g["pmax"], g["nmin"], g["nmax"], device=dev,
                       verify_threads=16)
    assert got.factorcount == g["factorcount"]
    assert got.primecount == g["primecount"]
    assert f"{got.checksum:016X}" == g["checksum"]
