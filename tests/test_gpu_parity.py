"""GPU parity: the HIP path through the C ABI against the CPU oracle and the
reference's golden self-test vectors.  Bit-exact: candidate sets, factor
records, primecounts and checksums must be identical."""
import ctypes as C

import numpy as np
import pytest

from pfcsieve_amd._native import CTRL_BYTES

pytestmark = pytest.mark.gpu

F, P, Cm = 1, 2, 4
TWO64 = 1 << 64


def gen_on_gpu(dev, lo, hi, cap=None):
    """Stage (a) alone through the C ABI, on library-allocated device memory."""
    import pfcsieve_amd._native as N
    from pfcsieve_amd.sieve import DeviceBuffer
    span = hi - lo
    if cap is None:
        cap = int(2.2 * span / max(np.log(max(span, 3)), 1.0)) + 4096
    d_p = DeviceBuffer(dev, cap * 8)
    d_ctrl = DeviceBuffer(dev, CTRL_BYTES)
    d_ctrl.zero()
    N.check(N.lib().pfc_stage_generate(lo, hi, d_p.ptr, cap, d_ctrl.ptr, None))
    N.check(N.lib().pfc_device_synchronize(dev.handle))
    ctrl = np.frombuffer(d_ctrl.download(), dtype=np.uint32)
    n = int(ctrl[0])
    assert ctrl[2] & 1 == 0, "candidate overflow"
    got = np.sort(np.frombuffer(d_p.download(8 * n), dtype=np.uint64)) if n else np.zeros(0, np.uint64)
    d_p.free()
    d_ctrl.free()
    return got


@pytest.mark.parametrize("lo,hi", [
    (3, 100000),                       # primes below 114 plus the sieve
    (101, 114),                        # only addsmallprimes territory
    (114, 120),                        # empty-ish window
    (10**9, 10**9 + 10**6),            # the reference's CPU-runnable window: 48,155 candidates
    ((1 << 32) - 300000, (1 << 32) + 300000),
    ((1 << 62) - 4 * 10**6, 1 << 62),  # just below 2^62
    (TWO64 - 2 * 10**6, TWO64 - 1),    # top of the 64-bit range
])
def test_generate_matches_oracle(orc, dev, lo, hi):
    got = gen_on_gpu(dev, lo, hi)
    ref = orc.candidates(lo, hi)
    assert len(got) == len(ref)
    assert np.array_equal(got, np.sort(ref))


def test_generate_factorial_window_count(orc, dev):
    got = gen_on_gpu(dev, 10**9, 10**9 + 10**6)
    assert len(got) == 48155


SMALL_JOBS = [
    (F, 101, 30000, 101, 4000),
    (P, 101, 30000, 101, 4000),            # 4000 is composite
    (Cm, 101, 30000, 101, 4000),
    (F | Cm, 101, 30000, 101, 4000),
    (F, 3, 2000, 101, 3000),               # p below nmin: residues hit 0
    (F, 10**9, 10**9 + 10**5, 101, 5000),  # slice of the -! n [101,5000) p [1e9,1.001e9) job
    (Cm, 10**12, 10**12 + 2 * 10**5, 101, 3000),
    (P, (1 << 62) - 10**5, 1 << 62, 101, 5000),
    (F | Cm, TWO64 - 10**5, TWO64 - 1, 101, 2000),
    (F, 10**6, 10**6 + 10**5, 2000, 9000),  # larger nmin: non-trivial setup
    (P, 10**7, 10**7 + 10**5, 5000, 12000),
    (Cm, 10**7, 10**7 + 10**5, 5000, 12000),
    (F | Cm, 10**7, 10**7 + 10**5, 5000, 9000),
    # n across 65,536, where pairs stop being product steps (p above twice every n)
    (F | Cm, 140000, 400000, 65000, 66100),
    (Cm, 140000, 400000, 65000, 66100),
    (P, 140000, 400000, 64000, 67000),    # 67000 is composite
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


@pytest.mark.parametrize("job", [(F | Cm, 101, 200000, 101, 3000), (P, 10**9, 10**9 + 4 * 10**5, 2000, 6000),
                                 (Cm, 10**12, 10**12 + 3 * 10**5, 101, 2500)])
def test_graph_replay_matches_eager(orc, dev, job):
    """The captured segment pipeline, replayed once per segment (many
    replays, with eager jobs run between them), gives the oracle's result."""
    import pfcsieve_amd as PK
    ref = orc.run(*job)
    dev.set_tuning(seg_candidates=2000, walk_chunk=900, use_graph=1)
    try:
        g = PK.Job(dev, *job)
        k = 0
        while g.position() < job[2]:
            g.run(1)
            k += 1
            if k % 3 == 0:   # eager work between replays
                dev.set_tuning(seg_candidates=2000, walk_chunk=900, use_graph=0)
                e = PK.run_sieve(F, 10**6, 10**6 + 5000, 101, 600, device=dev)
                r2 = orc.run(F, 10**6, 10**6 + 5000, 101, 600)
                assert e.checksum == r2.checksum
                dev.set_tuning(seg_candidates=2000, walk_chunk=900, use_graph=1)
        got = PK.job_result(g, verify_threads=8)
        g.destroy()
    finally:
        dev.set_tuning()
    assert k > 3
    assert (got.checksum, got.primecount, got.factorcount) == (ref.checksum, ref.primecount, ref.factorcount)


def _table(dev, kind, N, corrupt=False):
    """pfc_build_setup_table on the host, uploaded to device buffers
    (corrupt: the first chunk is changed before the upload)."""
    import pfcsieve_amd._native as Nt
    from pfcsieve_amd.sieve import DeviceBuffer, TableLayout
    t = TableLayout(kind, N)
    ch = (C.c_uint64 * max(len(t.chunks), 1))(*t.chunks)
    gr = (C.c_uint32 * len(t.groups))(*t.groups)
    if corrupt:
        assert t.chunks
        ch[0] = ch[0] + 2 if ch[0] < 2**62 else ch[0] - 2
    dc = DeviceBuffer(dev, 8 * max(len(t.chunks), 1))
    dc.upload(bytes(ch))
    dg = DeviceBuffer(dev, 4 * len(t.groups))
    dg.upload(bytes(gr))
    return Nt.SetupTable(dc.ptr.value, dg.ptr.value, t.npos, t.digit_bits), (dc, dg)


@pytest.mark.parametrize("modes,nmin,nmax", [(F, 2000, 2600), (Cm, 2000, 2600), (F | Cm, 2000, 2600)])
def test_stages_individually(orc, dev, modes, nmin, nmax):
    """generate -> setup -> walk (two launches) -> check, each called through
    the stage ABI; residues after setup and the checksum against the oracle."""
    import pfcsieve_amd._native as Nt
    from pfcsieve_amd.sieve import DeviceBuffer
    L = Nt.lib()
    lo, hi = 10**10, 10**10 + 50000
    cap = 8192
    bufs = {k: DeviceBuffer(dev, cap * w) for k, w in
            [("p", 8), ("v", 4), ("one", 8), ("ra", 8), ("rb", 8)]}
    ctrl = DeviceBuffer(dev, CTRL_BYTES)
    ctrl.zero()
    fac = DeviceBuffer(dev, 16 * 1024)
    Nt.check(L.pfc_stage_generate(lo, hi, bufs["p"].ptr, cap, ctrl.ptr, None))
    two = modes == (F | Cm)
    ta, keep_a = _table(dev, 0 if modes & F else 2, nmin - 1)
    tb, keep_b = _table(dev, 2, nmin - 1)
    Nt.check(L.pfc_stage_setup(bufs["p"].ptr, ctrl.ptr, bufs["v"].ptr, bufs["one"].ptr, bufs["ra"].ptr,
                               bufs["rb"].ptr if two else None, C.byref(ta), C.byref(tb) if two else None,
                               cap, None))
    Nt.check(L.pfc_device_synchronize(dev.handle))
    n = int(np.frombuffer(ctrl.download(), np.uint32)[0])
    p = np.frombuffer(bufs["p"].download(8 * n), np.uint64)
    ra = np.frombuffer(bufs["ra"].download(8 * n), np.uint64)
    for i in range(0, n, max(1, n // 50)):
        pi = int(p[i])
        rec, _ = orc.candidate_record(modes, pi, nmin, nmin)    # state after setup, no walk
        s = 64 - pi.bit_length()
        # plain residue X << s on the device, Mont(X) in the oracle's record
        assert (int(ra[i]) >> s) * 2**64 % pi == rec[6]
    # composite bitmap over [nmin, nmax)
    comp = np.zeros((nmax - nmin) // 32 + 2, np.uint32)
    isp = set(int(x) for x in orc.primes(nmin, nmax - 1))
    for k in range(nmax - nmin):
        if nmin + k not in isp:
            comp[k >> 5] |= np.uint32(1 << (k & 31))
    dcomp = DeviceBuffer(dev, comp.nbytes)
    dcomp.upload(comp.tobytes())
    mid = (nmin + nmax) // 2
    for j0, j1 in [(nmin, mid), (mid, nmax)]:
        Nt.check(L.pfc_stage_walk(modes, j0, j1, None, dcomp.ptr if modes & Cm else None, nmin, bufs["p"].ptr,
                                  bufs["v"].ptr, bufs["ra"].ptr, bufs["rb"].ptr if two else None,
                                  ctrl.ptr, fac.ptr, 1024, cap, None))
    Nt.check(L.pfc_stage_check(modes, nmax, bufs["p"].ptr, bufs["v"].ptr, bufs["one"].ptr, bufs["ra"].ptr,
                               bufs["rb"].ptr if two else None, ctrl.ptr, cap, None))
    Nt.check(L.pfc_device_synchronize(dev.handle))
    c = np.frombuffer(ctrl.download(), np.uint64)
    ref = orc.run(modes, lo, hi, nmin, nmax)
    assert int(c[2]) == ref.dev_checksum and int(c[3]) == ref.primecount == n
    # walk-completion count: every candidate took every multiplier step
    steps = (nmax - nmin if modes & F else 0) + ((nmax - nmin) - len(isp) if modes & Cm else 0)
    assert int(c[4]) == n * steps


@pytest.mark.gpu
@pytest.mark.parametrize("offset", [0, 1])
def test_primorial_walk_stage(orc, dev, offset):
    """Primorial walk through the stage ABI with walk boundaries that are not
    multiples of 16 steps, once with the prime list at a 64-byte aligned
    device address (16-prime block loads) and once 4 bytes past it (per-prime
    loads); residues are folded by the check stage and compared with the
    oracle's checksum."""
    import pfcsieve_amd._native as Nt
    from pfcsieve_amd.sieve import DeviceBuffer
    L = Nt.lib()
    nmin, nmax = 2000, 2600
    lo, hi = 10**10, 10**10 + 50000
    cap = 8192
    bufs = {k: DeviceBuffer(dev, cap * w) for k, w in [("p", 8), ("v", 4), ("one", 8), ("ra", 8)]}
    ctrl = DeviceBuffer(dev, CTRL_BYTES)
    ctrl.zero()
    fac = DeviceBuffer(dev, 16 * 1024)
    Nt.check(L.pfc_stage_generate(lo, hi, bufs["p"].ptr, cap, ctrl.ptr, None))
    ta, keep_a = _table(dev, 1, nmin - 1)
    Nt.check(L.pfc_stage_setup(bufs["p"].ptr, ctrl.ptr, bufs["v"].ptr, bufs["one"].ptr, bufs["ra"].ptr,
                               None, C.byref(ta), None, cap, None))
    primes = np.array([int(x) for x in orc.primes(nmin, nmax - 1)], np.uint32)
    assert len(primes) > 48
    host = np.concatenate([np.zeros(offset, np.uint32), primes])
    dlist = DeviceBuffer(dev, host.nbytes)
    dlist.upload(host.tobytes())
    base = dlist.ptr.value + 4 * offset
    assert (base % 64 == 0) == (offset == 0)
    for j0, j1 in [(0, 37), (37, 38), (38, len(primes))]:
        Nt.check(L.pfc_stage_walk(P, j0, j1, C.c_void_p(base), None, nmin, bufs["p"].ptr, bufs["v"].ptr,
                                  bufs["ra"].ptr, None, ctrl.ptr, fac.ptr, 1024, cap, None))
    Nt.check(L.pfc_stage_check(P, nmax, bufs["p"].ptr, bufs["v"].ptr, bufs["one"].ptr, bufs["ra"].ptr,
                               None, ctrl.ptr, cap, None))
    Nt.check(L.pfc_device_synchronize(dev.handle))
    c = np.frombuffer(ctrl.download(), np.uint64)
    u = np.frombuffer(ctrl.download(), np.uint32)
    ref = orc.run(P, lo, hi, nmin, nmax)
    assert int(c[2]) == ref.dev_checksum and int(c[3]) == ref.primecount == int(u[0])
    assert int(u[1]) == ref.raw_factors
    assert int(c[4]) == int(u[0]) * len(primes)        # walk-completion count


def _x_of(orc, kind, N, p):
    """X(N) mod p from its definition: N!, N# or N!/N#."""
    primes = set(int(q) for q in orc.primes(2, N)) if N >= 2 else set()
    x = 1
    for n in range(2, N + 1):
        if kind == 0 or (kind == 1) == (n in primes):
            x = x * n % p
    return x


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("N", [1, 100, 5000, 123457])
def test_table_self_check(orc, dev, kind, N):
    """The device self-check of a start table: both sides equal Mont(X(N))
    modulo 2^64-59 computed from the definition, and a table with one
    changed chunk is caught."""
    import pfcsieve_amd._native as Nt
    from pfcsieve_amd.sieve import DeviceBuffer
    L = Nt.lib()
    pt = 2**64 - 59
    want = _x_of(orc, kind, N, pt) * 2**64 % pt
    work = DeviceBuffer(dev, 8 * Nt.VERIFY_WORDS)
    t, keep = _table(dev, kind, N)
    Nt.check(L.pfc_stage_verify_table(kind, N, C.byref(t), work.ptr, None))
    Nt.check(L.pfc_device_synchronize(dev.handle))
    got = np.frombuffer(work.download(16), np.uint64)
    assert int(got[0]) == want and int(got[1]) == want
    if t.npos:
        tb, keep_b = _table(dev, kind, N, corrupt=True)
        Nt.check(L.pfc_stage_verify_table(kind, N, C.byref(tb), work.ptr, None))
        Nt.check(L.pfc_device_synchronize(dev.handle))
        bad = np.frombuffer(work.download(16), np.uint64)
        assert int(bad[0]) == want and int(bad[1]) != want


@pytest.mark.gpu
def test_job_create_runs_table_check(dev):
    """Jobs of every mode are created with the self-check on (the default) and
    with it skipped, and give the same results."""
    import pfcsieve_amd as PK
    res = []
    for skip in (0, 1):
        dev.set_tuning(skip_table_check=skip)
        try:
            for modes, nmax in ((F, 3000), (P, 3000), (Cm, 3000), (F | Cm, 3000)):
                j = PK.Job(dev, modes, 10**9, 10**9 + 20000, 2500, nmax)
                j.run()
                st = j.stats()
                res.append((skip, modes, st.dev_checksum, st.primecount))
                j.destroy()
        finally:
            dev.set_tuning()
    assert [r[1:] for r in res[:4]] == [r[1:] for r in res[4:]]


@pytest.mark.gpu
def test_autotune_then_run_matches_oracle(orc, dev):
    """pfc_autotune picks a valid launch shape and n-chunk length, and a job
    run with the tuned device gives the oracle's results."""
    import pfcsieve_amd as PK
    try:
        t = dev.autotune(F | Cm, 10**9, 10**9 + 10**6, 101, 20000)
        assert t["primes_per_lane"] in (1, 2, 4) and t["block_size"] in (64, 128, 256)
        assert t["walk_chunk"] in (1 << 14, 1 << 16, 1 << 18)
        assert t["lds_segment"] in (15360, 30720, 61440)
        assert dev.tuning() == t
        r = PK.run_sieve(F | Cm, 10**9, 10**9 + 300000, 101, 20000, device=dev, verify_threads=4)
        ref = orc.run(F | Cm, 10**9, 10**9 + 300000, 101, 20000)
        assert (r.checksum, r.primecount, r.factorcount) == (ref.checksum, ref.primecount, ref.factorcount)
    finally:
        dev.set_tuning()


@pytest.mark.gpu
def test_wilson_and_prime_values_on_device(orc, dev):
    """Properties fixed by the mathematics, checked on the device's output:
    Wilson's theorem gives p | (p-1)!+1 and p | (p-2)!-1 for every prime p in
    range; n!+1 (n = 116, 154, 320), n!-1 (n = 166, 324, 379), n#+1 (n = 379,
    1019, 1021, 2657) and n#-1 (n = 317, 337, 991, 1873, 2053, 2377) are prime,
    so no factor of them may be reported."""
    import pfcsieve_amd as PK
    r = PK.run_sieve(F, 3, 300000, 101, 2000, device=dev, verify_threads=8)
    got = set((p, nc) for p, nc, t in r.factors)
    for p in (int(x) for x in orc.primes(103, 2001)):
        assert (p, p - 1) in got, p
    for p in (int(x) for x in orc.primes(103, 2002)):
        assert (p, -(p - 2)) in got, p
    ncs = set(nc for _, nc, _ in r.factors)
    assert not ncs & {116, 154, 320, -166, -324, -379}
    rp = PK.run_sieve(P, 3, 300000, 101, 3000, device=dev, verify_threads=8)
    ncs = set(nc for _, nc, _ in rp.factors)
    assert rp.factorcount > 0
    assert not ncs & {379, 1019, 1021, 2657, -317, -337, -991, -1873, -2053, -2377}
    ref = orc.run(P, 3, 300000, 101, 3000)
    assert (rp.checksum, rp.primecount, rp.factorcount) == (ref.checksum, ref.primecount, ref.factorcount)


@pytest.mark.gpu
@pytest.mark.parametrize("modes,lo,width,nmin,nmax", [
    (F | Cm, 10**12, 7_200_000, 101, 10**6),             # the headline shape: -! -c, n up to 1e6
    (P, 2**62 - 4 * 10**9, 11_300_000, 101, 10**7),      # the primorial shape: n up to 1e7, p near 2^62
])
def test_full_size_walk_sampled_against_oracle(orc, dev, modes, lo, width, nmin, nmax):
    """The full n-range of the benchmark configurations on a p-window of about
    2^18 candidates, through the stage ABI (generate, set-up, walk in 65,536-step
    launches); the final residues of 24 sampled candidates against the oracle's
    own walk of each candidate over the same n-range."""
    import random
    import pfcsieve_amd._native as Nt
    from pfcsieve_amd.sieve import DeviceBuffer
    L = Nt.lib()
    cap = 1 << 19
    two = modes == (F | Cm)
    bufs = {k: DeviceBuffer(dev, cap * w) for k, w in [("p", 8), ("v", 4), ("one", 8), ("ra", 8), ("rb", 8)]}
    ctrl = DeviceBuffer(dev, CTRL_BYTES)
    ctrl.zero()
    fac = DeviceBuffer(dev, 16 * 4096)
    Nt.check(L.pfc_stage_generate(lo, lo + width, bufs["p"].ptr, cap, ctrl.ptr, None))
    ta, keep_a = _table(dev, 1 if modes == P else 0, nmin - 1)
    tb, keep_b = _table(dev, 2, nmin - 1)
    Nt.check(L.pfc_stage_setup(bufs["p"].ptr, ctrl.ptr, bufs["v"].ptr, bufs["one"].ptr, bufs["ra"].ptr,
                               bufs["rb"].ptr if two else None, C.byref(ta), C.byref(tb) if two else None,
                               cap, None))
    if modes == P:
        steps = np.array([int(q) for q in orc.primes(nmin, nmax - 1)], np.uint32)
        dlist = DeviceBuffer(dev, steps.nbytes)
        dlist.upload(steps.tobytes())
        j0, j1, lst, comp = 0, len(steps), dlist.ptr, None
    else:
        sv = np.ones(nmax, dtype=bool)
        sv[:2] = False
        for i in range(2, int(nmax ** 0.5) + 1):
            if sv[i]:
                sv[i * i::i] = False
        bits = ~sv[nmin:nmax]
        words = np.zeros((nmax - nmin) // 32 + 2, np.uint32)
        idx = np.nonzero(bits)[0]
        np.bitwise_or.at(words, idx >> 5, (np.uint32(1) << (idx & 31).astype(np.uint32)))
        dcomp = DeviceBuffer(dev, words.nbytes)
        dcomp.upload(words.tobytes())
        j0, j1, lst, comp = nmin, nmax, None, dcomp.ptr
    for a in range(j0, j1, 65536):
        Nt.check(L.pfc_stage_walk(modes, a, min(j1, a + 65536), lst, comp, nmin, bufs["p"].ptr, bufs["v"].ptr,
                                  bufs["ra"].ptr, bufs["rb"].ptr if two else None, ctrl.ptr,
                                  fac.ptr, 4096, cap, None))
    Nt.check(L.pfc_device_synchronize(dev.handle))
    u = np.frombuffer(ctrl.download(), np.uint32)
    n, flags = int(u[0]), int(u[2])
    assert flags & 1 == 0 and 200_000 < n <= cap
    p = np.frombuffer(bufs["p"].download(8 * n), np.uint64)
    ra = np.frombuffer(bufs["ra"].download(8 * n), np.uint64)
    rb = np.frombuffer(bufs["rb"].download(8 * n), np.uint64) if two else None
    rng = random.Random(modes * 7919 + nmax)
    for i in rng.sample(range(n), 24):
        pi = int(p[i])
        s = 64 - pi.bit_length()
        rec, _ = orc.candidate_record(modes, pi, nmin, nmax)
        # The walk keeps plain residues X << s; the oracle reports Mont(X).
        assert (int(ra[i]) >> s) * 2**64 % pi == rec[6], (pi, "res_a")
        if two:
            assert (int(rb[i]) >> s) * 2**64 % pi == rec[4], (pi, "res_b")


@pytest.mark.gpu
def test_job_result_and_sync(orc, dev):
    """pfc_job_result returns the device checksum part, the candidate count
    and the sorted raw records in one call, equal to the oracle's."""
    import pfcsieve_amd as PK
    job = (F | Cm, 10**6, 10**6 + 300000, 101, 3000)
    j = PK.Job(dev, *job)
    try:
        j.sync()
        j.run()
        j.sync()
        cs, pc, recs = j.result()
        st = j.stats()
        assert (cs, pc, len(recs)) == (st.dev_checksum, st.primecount, st.raw_factors)
        ref = orc.run(*job)
        assert (cs, pc, len(recs)) == (ref.dev_checksum, ref.primecount, ref.raw_factors)
        assert recs == sorted(recs, key=lambda r: (r[0], abs(r[1]), r[2]))
    finally:
        j.destroy()


@pytest.mark.gpu
@pytest.mark.parametrize("lds", [30720, 61440])
@pytest.mark.parametrize("job", [(F, 101, 200000, 101, 600), (Cm, 10**12, 10**12 + 10**6, 101, 500),
                                 (P, TWO64 - 3 * 10**6, TWO64 - 1, 101, 400)])
def test_generator_lds_segments(orc, dev, lds, job):
    """Jobs run with the larger generator LDS segments give the oracle's
    results (candidate set, sums and factors)."""
    import pfcsieve_amd as PK
    dev.set_tuning(lds_segment=lds)
    try:
        got = PK.run_sieve(*job, device=dev, verify_threads=8)
    finally:
        dev.set_tuning()
    ref = orc.run(*job)
    assert (got.checksum, got.primecount, got.factorcount) == (ref.checksum, ref.primecount, ref.factorcount)


def _random_jobs(seed, k):
    """Seeded small jobs over every mode, p ranges from below 114 to just under
    2^64, and n ranges from 101 up to a few thousand wide."""
    import random
    rng = random.Random(seed)
    jobs = []
    while len(jobs) < k:
        modes = rng.choice([F, P, Cm, F | Cm])
        nmin = 101 + rng.randrange(3000)
        nmax = nmin + 1 + rng.randrange(2500)
        if modes == P and all(nmax % q for q in range(2, int(nmax ** 0.5) + 1)):
            continue                                  # primorial needs a composite nmax
        where = rng.randrange(4)
        if where == 0:
            pmin = 3 + rng.randrange(20000)
        elif where == 1:
            pmin = rng.randrange(10**9, 10**13)
        elif where == 2:
            pmin = (1 << 62) - rng.randrange(1, 10**9)
        else:
            pmin = TWO64 - 1 - rng.randrange(10**5, 10**9)
        pmax = min(TWO64 - 1, pmin + 1 + rng.randrange(2 * 10**5))
        jobs.append((modes, pmin, pmax, nmin, nmax))
    return jobs


@pytest.mark.gpu
def test_random_jobs_match_oracle(orc, dev):
    """40 seeded random jobs (every mode, p from the small primes to 2^64) give
    the oracle's checksum, primecount and factor list."""
    import pfcsieve_amd as PK
    for job in _random_jobs(20251203, 40):
        got = PK.run_sieve(*job, device=dev, verify_threads=8)
        ref = orc.run(*job)
        want = [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in ref.factors]
        assert (got.checksum, got.primecount, got.factors) == (ref.checksum, ref.primecount, want), job


@pytest.mark.gpu
def test_factor_buffer_overflow_splits_segments(orc, dev):
    """A segment whose factor records exceed the buffer is walked again in
    halves (plain and batched), giving the oracle's results; a single prime
    that alone overflows it fails with PFC_ERR_FACTOR_OVERFLOW."""
    import pfcsieve_amd as PK
    job = (F, 101, 30000, 101, 3000)
    ref = orc.run(*job)
    assert ref.raw_factors > 1000
    got = PK.run_sieve(*job, device=dev, verify_threads=8, factor_capacity=300)
    want = [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in ref.factors]
    assert (got.checksum, got.primecount, got.factors) == (ref.checksum, ref.primecount, want)
    assert got.segments > 4
    j = PK.Job(dev, *job, factor_capacity=300)
    b = PK.Batcher(dev)
    try:
        b.add(j)
        while b.pending():
            b.round(0)
        g = PK.job_result(j, verify_threads=8)
        assert (g.checksum, g.primecount, g.factors) == (ref.checksum, ref.primecount, want)
    finally:
        b.destroy()
        j.destroy()
    with pytest.raises(PK.PfcError) as e:
        PK.run_sieve(F, 103, 104, 101, 103, device=dev, factor_capacity=1)   # 101!-1 and 102!+1
    assert e.value.code == -4


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "batched"])
@pytest.mark.parametrize("chunk", [7, 33, 901])
@pytest.mark.parametrize("modes", [Cm, F | Cm])
def test_composite_list_short_launches(orc, dev, modes, chunk, form):
    """Jobs read their composite list 16 entries per block, with single steps
    before the first whole block and after the last one of each launch (or
    batch round).  Launches of 7, 33 and 901 n (rounds of as many steps) put
    those ends all over the list, with pairs and product steps in between
    (p in [1e5, 3e5) exceeds twice every multiplier below 3001); factor
    records and checksum against the oracle."""
    import pfcsieve_amd as PK
    job = (modes, 10**5, 3 * 10**5, 101, 3001)
    ref = orc.run(*job)
    want = (ref.checksum, ref.primecount, [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in ref.factors])
    assert len(want[2]) > 100
    if form == "plain":
        dev.set_tuning(walk_chunk=chunk)
        try:
            got = PK.run_sieve(*job, device=dev, verify_threads=8)
        finally:
            dev.set_tuning()
    else:
        [got] = PK.run_batch([job], device=dev, max_steps=chunk, verify_threads=8)
    assert (got.checksum, got.primecount, got.factors) == want


@pytest.mark.gpu
@pytest.mark.parametrize("modes", [F, Cm, F | Cm])
def test_walk_launches_end_at_hits(orc, dev, modes):
    """The walk tests every second step of a chain when all p exceed twice the
    largest multiplier, so a launch may end on an untested step.  Here every
    launch ends at a hit step or right after one (p in [1e5, 3e5), n below
    3001, so pairs are taken); factor records and checksum against the
    oracle."""
    import sympy
    import pfcsieve_amd._native as Nt
    from pfcsieve_amd.sieve import DeviceBuffer
    L = Nt.lib()
    nmin, nmax = 101, 3001
    lo, hi = 10**5, 3 * 10**5
    ref = orc.run(modes, lo, hi, nmin, nmax)
    hits = sorted(set(abs(int(r["nc"])) for r in ref.factors))
    assert len(hits) > 100
    cuts = sorted({nmin, nmax} | {n for h in hits for n in (h, h + 1) if nmin < n < nmax})
    cap = 1 << 15
    two = modes == (F | Cm)
    bufs = {k: DeviceBuffer(dev, cap * w) for k, w in [("p", 8), ("v", 4), ("one", 8), ("ra", 8), ("rb", 8)]}
    ctrl = DeviceBuffer(dev, CTRL_BYTES)
    ctrl.zero()
    fcap = 1 << 14
    fac = DeviceBuffer(dev, 16 * fcap)
    Nt.check(L.pfc_stage_generate(lo, hi, bufs["p"].ptr, cap, ctrl.ptr, None))
    ta, keep_a = _table(dev, 0 if modes & F else 2, nmin - 1)
    tb, keep_b = _table(dev, 2, nmin - 1)
    Nt.check(L.pfc_stage_setup(bufs["p"].ptr, ctrl.ptr, bufs["v"].ptr, bufs["one"].ptr, bufs["ra"].ptr,
                               bufs["rb"].ptr if two else None, C.byref(ta), C.byref(tb) if two else None,
                               cap, None))
    comp = np.zeros((nmax - nmin) // 32 + 2, np.uint32)
    isp = set(int(x) for x in orc.primes(nmin, nmax - 1))
    for k in range(nmax - nmin):
        if nmin + k not in isp:
            comp[k >> 5] |= np.uint32(1 << (k & 31))
    dcomp = DeviceBuffer(dev, comp.nbytes)
    dcomp.upload(comp.tobytes())
    for j0, j1 in zip(cuts[:-1], cuts[1:]):
        Nt.check(L.pfc_stage_walk(modes, j0, j1, None, dcomp.ptr if modes & Cm else None, nmin, bufs["p"].ptr,
                                  bufs["v"].ptr, bufs["ra"].ptr, bufs["rb"].ptr if two else None,
                                  ctrl.ptr, fac.ptr, fcap, cap, None))
    Nt.check(L.pfc_stage_check(modes, nmax, bufs["p"].ptr, bufs["v"].ptr, bufs["one"].ptr, bufs["ra"].ptr,
                               bufs["rb"].ptr if two else None, ctrl.ptr, cap, None))
    Nt.check(L.pfc_device_synchronize(dev.handle))
    c = np.frombuffer(ctrl.download(), np.uint64)
    u = np.frombuffer(ctrl.download(), np.uint32)
    assert int(c[2]) == ref.dev_checksum and int(c[3]) == ref.primecount == int(u[0])
    assert int(u[1]) == ref.raw_factors and ref.raw_factors <= fcap
    steps = (nmax - nmin if modes & F else 0) + ((nmax - nmin) - len(isp) if modes & Cm else 0)
    assert int(c[4]) == int(u[0]) * steps                # walk-completion count over all launches
    raw = np.frombuffer(fac.download(16 * int(u[1])), dtype=orc.FACTOR_DTYPE)
    got = sorted((int(r["p"]), int(r["nc"]), int(r["type"])) for r in raw if sympy.isprime(int(r["p"])))
    assert got == sorted((int(r["p"]), int(r["nc"]), int(r["type"])) for r in ref.factors)


@pytest.mark.gpu
def test_primorial_triples_find_planted_factors(orc, dev):
    """P-mode walks over the primes in [101, 400) with every p above
    2^14 * 397^2 test their steps in triples; p-windows around known factors
    of n# +- 1 (tests/golden/primorial_factors.json, hits at all three
    positions of a triple) give the oracle's results, with the planted record
    among them, run alone and batched."""
    import json
    import os
    import pfcsieve_amd as PK
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "primorial_factors.json")
    with open(path) as f:
        vecs = json.load(f)["vectors"]
    descs = [(P, v["p"] - 3000, v["p"] + 3000, 101, 400) for v in vecs]
    refs = []
    for desc in descs:
        r = orc.run(*desc)
        refs.append((r.checksum, r.primecount, [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in r.factors]))
    for desc, v, ref in zip(descs, vecs, refs):
        got = PK.run_sieve(*desc, device=dev, verify_threads=8)
        assert (got.checksum, got.primecount, got.factors) == ref, v
        assert (v["p"], v["c"] * v["n"], 1) in got.factors, v
    for g, ref in zip(PK.run_batch(descs, device=dev, verify_threads=8), refs):
        assert (g.checksum, g.primecount, g.factors) == ref


@pytest.fixture(scope="module")
def planted_narrow(orc):
    """Jobs around the planted factors of tests/golden/narrow_factors.json
    (p in (2^34, 2^52), n < 200) with the oracle's results: every mode that
    walks the factor's chain (F or C, and F|C), from nmin = 101 and 104
    (which puts the hit at the tested and at the untested step of a pair in
    both chains), n up to 300."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "narrow_factors.json")
    with open(path) as f:
        vecs = json.load(f)["vectors"]
    descs, planted = [], []
    for v in vecs:
        chain = F if v["kind"] == "F" else Cm
        for modes in (chain, F | Cm):
            for nmin in (101, 104):
                if v["n"] >= nmin:
                    descs.append((modes, v["p"] - 2000, v["p"] + 2000, nmin, 300))
                    planted.append((v["p"], v["c"] * v["n"], 0 if v["kind"] == "F" else 2))
    refs = []
    for desc in descs:
        r = orc.run(*desc)
        refs.append((r.checksum, r.primecount, [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in r.factors]))
    return descs, planted, refs


@pytest.mark.parametrize("forms_off,ppl", [(0, 1), (0, 2), (0, 4), (1, 1), (1, 4), (3, 1)])
def test_walk_forms_find_planted_factors(dev, planted_narrow, forms_off, ppl):
    """The planted jobs give the oracle's results, with the planted record
    among them, in each walk form: narrow (plain residues below 2p, the form
    these moduli take), medium (narrow form off: residues at half the shifted
    scale) and shifted (both off), with 1, 2 and 4 candidates per lane, run
    alone and batched."""
    import pfcsieve_amd as PK
    descs, planted, refs = planted_narrow
    dev.set_tuning(primes_per_lane=ppl, walk_forms_off=forms_off)
    try:
        for desc, rec, ref in zip(descs, planted, refs):
            got = PK.run_sieve(*desc, device=dev, verify_threads=8)
            assert (got.checksum, got.primecount, got.factors) == ref, desc
            assert rec in got.factors, (desc, rec)
        for desc, g, ref in zip(descs, PK.run_batch(descs, device=dev, verify_threads=8), refs):
            assert (g.checksum, g.primecount, g.factors) == ref, desc
    finally:
        dev.set_tuning()


NARROW_TOP = [
    (F | Cm, 3 * 10**12, 3 * 10**12 + 2 * 10**5, 900_000, 1_300_000),        # p * n up to 2^61.8
    (F, 4_398_000_000_000, 4_398_000_000_000 + 2 * 10**5, 950_000, 1_060_000),  # crosses 2^62 at n ~ 1.0486e6
    (P, 2**41, 2**41 + 4 * 10**5, 101, 2_000_000),                          # primes up to 2e6: p * q < 2^62
]


@pytest.mark.parametrize("forms_off", [0, 1])
def test_narrow_form_near_its_bound(orc, dev, forms_off):
    """Moduli and multipliers with p * n in [2^61, 2^62), the top of the narrow
    form's range (t = X * n up to 2^63), and a job whose n-range crosses
    2^62 / p so that its launches change form: the oracle's results with the
    narrow form on (0) and off (1, the medium form), alone and batched."""
    import pfcsieve_amd as PK
    refs = []
    for desc in NARROW_TOP:
        r = orc.run(*desc)
        refs.append((r.checksum, r.primecount, [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in r.factors]))
    dev.set_tuning(walk_forms_off=forms_off, walk_chunk=50_000)
    try:
        for desc, ref in zip(NARROW_TOP, refs):
            got = PK.run_sieve(*desc, device=dev, verify_threads=8)
            assert (got.checksum, got.primecount, got.factors) == ref, desc
        for desc, g, ref in zip(NARROW_TOP, PK.run_batch(NARROW_TOP, device=dev, max_steps=40_000,
                                                          verify_threads=8), refs):
            assert (g.checksum, g.primecount, g.factors) == ref, desc
    finally:
        dev.set_tuning()


@pytest.mark.gpu
@pytest.mark.parametrize("modes,lo,width,nmin,nmax,cut", [
    (F | Cm, 10**12, 2 * 10**8, 101, 10**6, 73_212_347),            # the -! -c benchmark shape
    (P, 2**62 - 4 * 10**8, 4 * 10**8, 101, 10**7, 151_515_151),      # the primorial shape near 2^62
])
def test_full_n_range_additive_over_p(dev, modes, lo, width, nmin, nmax, cut):
    """At the benchmark configurations' full n-ranges, on p-windows too long
    for the oracle (1.4e13 and 6.2e12 updates), results add over a p-split
    that falls inside a segment, with the two parts cut into segments of
    another size: primecounts and checksums (mod 2^64) of [lo, lo + cut) and
    [lo + cut, lo + width) sum to the whole window's, and their factor lists
    concatenate to its list."""
    import pfcsieve_amd as PK
    whole = PK.run_sieve(modes, lo, lo + width, nmin, nmax, device=dev, verify_threads=8)
    dev.set_tuning(seg_candidates=1 << 21)
    try:
        a = PK.run_sieve(modes, lo, lo + cut, nmin, nmax, device=dev, verify_threads=8)
        b = PK.run_sieve(modes, lo + cut, lo + width, nmin, nmax, device=dev, verify_threads=8)
    finally:
        dev.set_tuning()
    assert whole.primecount > 5 * 10**6 and a.segments > 1 and b.segments > 1
    assert a.primecount + b.primecount == whole.primecount
    assert (a.checksum + b.checksum) % 2**64 == whole.checksum
    assert sorted(a.factors + b.factors) == whole.factors
