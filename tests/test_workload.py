"""Seeded job mixes (CPU)."""
import math

from pfcsieve_amd import workload as W


def test_splitmix_known_values():
    # splitmix64 reference outputs for seed 0 (Vigna's published generator)
    r = W.SplitMix64(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_mix_ranges_and_determinism():
    a = W.mix(256, seed=12345)
    b = W.mix(256, seed=12345)
    assert [u.desc() for u in a] == [u.desc() for u in b]
    assert {u.modes for u in a} == {1, 2, 4}
    for u in a:
        assert 101 <= u.nmin <= 101 + 10**6
        assert 10**3 <= u.nmax - u.nmin <= 10**6 + 64
        assert 10**10 <= u.pmin <= 2**63
        assert 10**6 <= u.pmax - u.pmin <= 10**9
        if u.modes == 2:
            assert not W._is_prime_small(u.nmax)


def test_poisson_arrivals():
    t = W.poisson_arrivals(20000, rate=4.0, seed=1)
    assert all(b > a for a, b in zip(t, t[1:]))
    assert abs(t[-1] / len(t) - 0.25) < 0.01


def test_walk_steps():
    u = W.Unit(1, 10**10, 10**10 + 10**6, 101, 1101)
    assert W.walk_steps(u) == 1000
    u = W.Unit(2, 10**10, 10**10 + 10**6, 101, 1101)
    assert W.walk_steps(u) == sum(1 for n in range(101, 1101) if W._is_prime_small(n))
    u = W.Unit(4, 10**10, 10**10 + 10**6, 101, 1101)
    assert W.walk_steps(u) == 1000 - sum(1 for n in range(101, 1101) if W._is_prime_small(n))


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N with N > 1 exits with status 2 and a JSON line that
    says the benchmark runs on one GPU (before touching a device)."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
