"""Benchmark of the sieve hot path on one MI355X.

Workload (default): plain factorial + compositorial sieve, -! -c,
n in [101, 1,000,000), p from 1e12 (the 2e9-wide p-range of the headline
configuration, processed one p-segment at a time).  One "step" is one
p-segment of about 2^24 candidates run through generate -> setup -> walk ->
check, with its factor records drained; inputs are generated on the device.

Prints ONE JSON line.  `value` is p*n residue updates per second over the
timed steps (updates = candidates x multiplier steps, F and C walks counted
separately).  `roofline` prices the updates at 19 INT32 multiplies each (the
reference's 64-bit Montgomery product, 11 32x32->64 partial products) against
the measured INT32 multiply lane-op rate of the device.
`cpu_baseline` times the CPU oracle (OpenMP, all host cores) on a bounded
p-sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W_MUL = 19  # INT32 multiply instructions per update in the reference formulation
# HBM bytes per candidate per walk launch, from the FETCH_SIZE (x2.000 copy-kernel
# calibration) and WRITE_SIZE counters of the walk on this workload, 2^24-candidate
# segments (profiles/k_walk_fc_1e12_pmc.md): p, v and both residues read, both
# residues written, plus the composite-list reads.
WALK_FETCH_PER_CAND = 28.90
WALK_WRITE_PER_CAND = 16.00


def composites_in(nmin, nmax):
    import numpy as np
    sieve = np.ones(nmax, dtype=bool)
    sieve[:2] = False
    for i in range(2, int(nmax ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = False
    return int((~sieve[nmin:nmax]).sum())


PRODUCT_LIMIT = 65536  # pairs with both multipliers up to this are one step by their product


def mads_per_update(nmin, nmax):
    """v_mad_u64_u32 instructions the F|C walk of n in [nmin, nmax) issues per
    update: five per step, and five per pair where a pair is one step by its
    product (factorial trips of four whose last n is at most PRODUCT_LIMIT;
    composite-list blocks of 16 whose last composite is).  Assumes the first
    walk launch starts at nmin and reaches past PRODUCT_LIMIT."""
    f_prod = max(0, (min(nmax, PRODUCT_LIMIT + 1) - nmin) // 4 * 4)
    c_prod = composites_in(nmin, min(nmax, PRODUCT_LIMIT + 1)) // 16 * 16
    steps = (nmax - nmin) + composites_in(nmin, nmax)
    return 5.0 * (steps - (f_prod + c_prod) / 2) / steps


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(modes, p0, width, nmin, nmax, steps_per_cand, width3=600_000):
    """The oracle (C, OpenMP over candidates, all host threads) on a bounded
    p-sample of the headline configuration, plus the reference's CPU-runnable
    case in full and a p-sample of the primorial configuration."""
    import oracle as O
    from pfcsieve_amd import workload as W
    O.build()
    t = time.perf_counter()
    r = O.run(modes, p0, p0 + width, nmin, nmax)
    dt = time.perf_counter() - t
    upd = r.primecount * steps_per_cand
    out = {"value": upd / dt, "unit": "updates/s", "cores": O.num_threads(), "kind": "port",
           "sample": f"-! -c p in [{p0}, {p0 + width}) ({r.primecount} candidates), n in [{nmin}, {nmax}); "
                     f"{dt:.1f} s",
           "cpu_model": _cpu_model(), "nproc": os.cpu_count()}
    # the reference's CPU-runnable case in full: -! -n 101 -N 5000 -p 1e9 -P 1.001e9
    t = time.perf_counter()
    r1 = O.run(1, 10**9, 10**9 + 10**6, 101, 5000)
    dt1 = time.perf_counter() - t
    out["factorial_n5000_p1e9_full"] = {"candidates": r1.primecount, "s": dt1, "updates_per_s": r1.primecount * 4899 / dt1,
                        "checksum": "%016X" % r1.checksum}
    # primorial p-sample: -# -n 101 -N 1e7 from p = 2^62 - 4e9
    p3 = 2**62 - 4 * 10**9
    steps3 = W.walk_steps(W.Unit(2, 0, 0, 101, 10**7))
    t = time.perf_counter()
    r3 = O.run(2, p3, p3 + width3, 101, 10**7)
    dt3 = time.perf_counter() - t
    out["primorial_n1e7_p2e62_sample"] = {"sample": f"-# p in [{p3}, {p3 + width3}), n in [101, 1e7)",
                          "candidates": r3.primecount, "s": dt3,
                          "updates_per_s": r3.primecount * steps3 / dt3}
    return out


def _progress(on, msg):
    if on:
        print(msg, file=sys.stderr, flush=True)


def mix_compare(dev, units, verify_threads=16, progress=False):
    """Plain (one job at a time, one range per launch sequence) against batched
    (all units through one batcher, shared persistent launches) on the same
    seeded unit mix.  Returns updates/s of both and the ratio."""
    import pfcsieve_amd as P
    from pfcsieve_amd import workload as W
    steps = [W.walk_steps(u) for u in units]
    # plain
    t = time.perf_counter()
    cand_plain, sig_plain = [], []
    for u in units:
        j = P.Job(dev, *u.desc())
        j.run()
        st = j.stats()
        cand_plain.append(st.primecount)
        sig_plain.append((st.dev_checksum, st.primecount, st.raw_factors))
        j.destroy()
        _progress(progress, f"plain: unit {len(cand_plain)}/{len(units)} done, {time.perf_counter() - t:.1f} s")
    t_plain = time.perf_counter() - t
    # batched (job creation -- host tables, device state -- inside the timed
    # region, as it is for the plain form)
    t = time.perf_counter()
    jobs = [P.Job(dev, *u.desc()) for u in units]
    b = P.Batcher(dev)
    for j in jobs:
        b.add(j)
    rounds = 0
    while b.pending():
        b.round(0)
        rounds += 1
        _progress(progress, f"batched: round {rounds} done, {b.pending()} jobs pending, {time.perf_counter() - t:.1f} s")
    cand_batch = [j.stats().primecount for j in jobs]
    sig_batch = [(j.stats().dev_checksum, j.stats().primecount, j.stats().raw_factors) for j in jobs]
    for j in jobs:
        j.destroy()
    b.destroy()
    t_batch = time.perf_counter() - t
    assert cand_plain == cand_batch
    upd = sum(c * s for c, s in zip(cand_plain, steps))
    return {"units": len(units), "candidates": sum(cand_plain), "walk_updates": upd,
            "plain_s": t_plain, "batched_s": t_batch, "plain_updates_per_s": upd / t_plain,
            "batched_updates_per_s": upd / t_batch, "batched_over_plain": t_plain / t_batch,
            "rounds": rounds, "results_identical": sig_plain == sig_batch}


def serve_run(dev, units, max_steps=50_000, round_updates=0, progress=False):
    """Served stream: the units arrive as a seeded Poisson process at 80% of
    the batch capacity measured first on the same units; admission by HBM
    footprint; rounds of at most max_steps steps per segment, or of about
    round_updates p*n updates when that is set; reports throughput and p50/p99
    job latency."""
    import pfcsieve_amd as P
    import pfcsieve_amd.serve as S
    from pfcsieve_amd import workload as W
    steps = [W.walk_steps(u) for u in units]
    t = time.perf_counter()
    say = (lambda m: _progress(True, m)) if progress else None
    res = P.run_batch([u.desc() for u in units], device=dev, verify_threads=16, progress=say)
    cap_s = time.perf_counter() - t
    rate = 0.8 * len(units) / cap_s
    arrivals = W.poisson_arrivals(len(units), rate, seed=777)
    info = dev.info()
    srv = S.Server(dev, hbm_budget=int(info["hbm_bytes"] * 0.8), max_steps=max_steps, verify_threads=16,
                   progress=say, round_updates=round_updates)
    for k, (u, a) in enumerate(zip(units, arrivals)):
        srv.submit(S.Request(k, u.desc(), a))
    st = srv.run()
    by_id = {r.rid: r for r in srv.done}
    same = all((by_id[k].result.checksum, by_id[k].result.primecount) == (res[k].checksum, res[k].primecount)
               for k in range(len(units)))
    upd = sum(r.primecount * s for r, s in zip(res, steps))
    return {"jobs": st.jobs, "arrival_rate_per_s": rate, "wall_s": st.wall_s, "rounds": st.rounds,
            "updates_per_s": upd / st.wall_s, "p50_latency_s": st.pct(50), "p99_latency_s": st.pct(99),
            "peak_resident_bytes": st.peak_resident_bytes, "results_equal_batch": same,
            "round_policy": (f"about {round_updates:.3g} updates per round" if round_updates
                             else f"at most {max_steps} steps per segment per round")}


def verify_baseline(dev):
    """The CPU factor verification (getResults' OpenMP verify, cl_sieve.cpp:509-540)
    on the factor records of self-test case 3 (-! p in [101, 1e5), n in [101, 1e6),
    42,821 factors), at 2 threads (the -v default) and at all host threads."""
    import ctypes as C
    import pfcsieve_amd as P
    from pfcsieve_amd._native import check, lib
    j = P.Job(dev, 1, 101, 100000, 101, 1000000)
    j.run()
    arr, n = j.raw_factors()
    j.destroy()
    out = {"factors": n, "sample": "self-test case 3 factor records"}
    for threads in (2, os.cpu_count() or 2):
        cp = (type(arr[0]) * n)(*arr[:n])
        cs, kept = C.c_uint64(0), C.c_size_t()
        t = time.perf_counter()
        check(lib().pfc_harvest(cp, n, threads, C.byref(cs), C.byref(kept)))
        out[f"threads_{threads}_s"] = time.perf_counter() - t
    # The reference's own verify() (verifyprime.cpp:149, compiled from its
    # source into oracle/_ref when the reference tree was present at build
    # time) on the first records, one call per record on all host threads as
    # getResults' OpenMP loop does, beside the library's harvest of the same
    # records (which also sorts, filters by isPrime and sums the checksum).
    import oracle as orc
    if orc.ref_available():
        from concurrent.futures import ThreadPoolExecutor
        m = min(n, 4000)
        threads = os.cpu_count() or 2
        orc.ref_is_prime(3)                     # loads the library before the threads start
        recs = [(int(f.p), abs(int(f.nc)), 1 if f.nc > 0 else -1, int(f.type)) for f in arr[:m]]
        t = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            ok = list(ex.map(lambda r: orc.ref_verify(r[0], r[1], r[2], r[3], 1000000), recs))
        t_ref = time.perf_counter() - t
        cp = (type(arr[0]) * m)(*arr[:m])
        cs, kept = C.c_uint64(0), C.c_size_t()
        t = time.perf_counter()
        check(lib().pfc_harvest(cp, m, threads, C.byref(cs), C.byref(kept)))
        out["reference_verify"] = {"records": m, "threads": threads, "all_verified": all(ok),
                                   "reference_s": t_ref, "library_harvest_s": time.perf_counter() - t}
    return out


def other_configs(dev):
    """Throughput on the other single-GPU workloads: the reference's CPU-runnable
    case (-! n [101,5000) p [1e9,1.001e9), whole job), the primorial walk near
    2^62 (-# n [101,1e7), one p-segment), single-mode walks and large-nmin set-up."""
    import pfcsieve_amd as P
    out = {}
    # -! -n 101 -N 5000 -p 1e9 -P 1.001e9 (48,155 candidates), whole job
    t = time.perf_counter()
    r = P.run_sieve(1, 10**9, 10**9 + 10**6, 101, 5000, device=dev, verify_threads=16)
    dt = time.perf_counter() - t
    out["factorial_n5000_p1e9"] = {"workload": "-! n [101,5000) p [1e9,1.001e9)", "candidates": r.primecount,
                   "updates": r.primecount * 4899, "wall_s": dt, "updates_per_s": r.primecount * 4899 / dt,
                   "factors": r.factorcount, "checksum": "%016X" % r.checksum}
    # -# -n 101 -N 1e7 -p 2^62-4e9 -P 2^62, one segment
    from pfcsieve_amd import workload as W
    steps = W.walk_steps(W.Unit(2, 0, 0, 101, 10**7))
    j = P.Job(dev, 2, 2**62 - 4 * 10**9, 2**62, 101, 10**7)
    t = time.perf_counter()
    j.run(1)
    dt = time.perf_counter() - t
    s = j.stats()
    j.destroy()
    out["primorial_n1e7_p2e62_segment"] = {"workload": "-# n [101,1e7) p from 2^62-4e9, one p-segment", "candidates": s.primecount,
                           "steps_per_candidate": steps, "wall_s": dt,
                           "updates_per_s": s.primecount * steps / dt,
                           "walk_updates_per_s": s.walk_updates / (s.walk_ms * 1e-3)}
    # The same segment with the launch shape and n-chunk length the autotuner
    # picks for this workload (it leaves them in the device tuning).
    saved = dev.tuning()
    best = dev.autotune(2, 2**62 - 4 * 10**9, 2**62, 101, 10**7)
    j = P.Job(dev, 2, 2**62 - 4 * 10**9, 2**62, 101, 10**7)
    j.run(1)
    s = j.stats()
    j.destroy()
    dev.set_tuning(**{k: saved[k] for k in ("primes_per_lane", "block_size", "walk_chunk", "seg_candidates",
                                            "use_graph", "skip_table_check", "lds_segment")})
    out["primorial_n1e7_p2e62_segment"]["autotuned"] = {
        "tuning": {k: best[k] for k in ("primes_per_lane", "block_size", "walk_chunk", "lds_segment")},
        "walk_updates_per_s": s.walk_updates / (s.walk_ms * 1e-3)}
    # Single-mode walks of the uneven mixes at the headline shape: -! and -c
    # alone, n in [101, 1e6), p from 1e12, one p-segment each.
    for name, modes in (("f_segment", 1), ("c_segment", 4)):
        steps = W.walk_steps(W.Unit(modes, 0, 0, 101, 10**6))
        j = P.Job(dev, modes, 10**12, 10**12 + 2 * 10**9, 101, 10**6)
        j.run(1)
        s = j.stats()
        j.destroy()
        out[name] = {"workload": ("-!" if modes == 1 else "-c") + " n [101,1e6) p from 1e12, one p-segment",
                     "candidates": s.primecount, "steps_per_candidate": steps,
                     "walk_updates_per_s": s.walk_updates / (s.walk_ms * 1e-3)}
    # Start-residue set-up at large nmin (the dominant cost of the uneven mixes):
    # -! n [1e6, 1e6+256) from p = 1e15, one p-segment.  Set-up time is the
    # segment's device time minus its walk launches (generate and check are
    # ~3 ms of it).
    nchunks, nbits = P.setup_table_size(0, 10**6 - 1)
    products = nchunks + nbits - 1               # chunk products and squarings
    j = P.Job(dev, 1, 10**15, 10**15 + 2 * 10**9, 10**6, 10**6 + 256)
    j.run(1)
    s = j.stats()
    j.destroy()
    setup_s = (s.device_ms - s.walk_ms) * 1e-3
    out["setup_nmin_1e6"] = {"workload": "-! n [1e6,1e6+256) p from 1e15, one p-segment: start residues "
                                         "999999! mod p per candidate",
                             "candidates": s.primecount, "modular_products_per_candidate": products,
                             "setup_s": setup_s, "candidates_per_s": s.primecount / setup_s,
                             "products_per_s": s.primecount * products / setup_s}
    return out


def shape_sweep(dev, modes, pmin, nmin, nmax):
    """Walk rate of one 2^21-candidate headline segment at every launch shape
    the autotuner considers (primes per lane x workgroup size)."""
    import pfcsieve_amd as P
    out = []
    for ppl in (1, 2, 4):
        for blk in (64, 128, 256):
            dev.set_tuning(primes_per_lane=ppl, block_size=blk, seg_candidates=1 << 21)
            j = P.Job(dev, modes, pmin, pmin + 10**9, nmin, nmax)
            j.run(1)
            s = j.stats()
            j.destroy()
            out.append({"primes_per_lane": ppl, "block_size": blk,
                        "walk_updates_per_s": s.walk_updates / (s.walk_ms * 1e-3)})
            _progress(True, f"sweep ppl={ppl} block={blk}: {out[-1]['walk_updates_per_s']:.4e} updates/s")
    dev.set_tuning()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seg-candidates", type=int, default=1 << 24)
    ap.add_argument("--ppl", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--walk-chunk", type=int, default=0, help="multiplier steps per walk launch (0 = default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-width", type=int, default=1_000_000)
    ap.add_argument("--mix-units", type=int, default=24,
                    help="units of the seeded 256-unit batched mix to compare plain vs batched (0 = skip)")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--configs", action="store_true", help="also measure the other single-GPU workloads (whole small job, primorial segment, single modes, set-up)")
    ap.add_argument("--serve-round-updates", type=float, default=4e11,
                    help="serve with rounds of about this many p*n updates, ~0.1 s of walk "
                         "(0: at most 50,000 steps per segment per round)")
    ap.add_argument("--serve-jobs", type=int, default=0,
                    help="jobs of the seeded served stream to run through the serving loop (0 = skip)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="run one plain copy of this many bytes first (HBM counter calibration)")
    ap.add_argument("--progress", action="store_true", help="progress lines on stderr during long mixes")
    ap.add_argument("--sweep", action="store_true", help="walk rate at every (primes per lane, block) shape")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; multi-GPU sharding over p is "
                                   "a design only, not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import pfcsieve_amd as P
    modes, nmin, nmax = P.MODE_FACTORIAL | P.MODE_COMPOSITORIAL, 101, 1_000_000
    pmin = 10**12
    ncomp = composites_in(nmin, nmax)
    steps_per_cand = (nmax - nmin) + ncomp

    dev = P.Device(0)
    if args.copy_probe:
        dev.probe_copy(args.copy_probe)
    rates = dev.probe_mul_rates()
    p_mul = max(rates.values())
    dev.set_tuning(seg_candidates=args.seg_candidates, primes_per_lane=args.ppl, block_size=args.block,
                   walk_chunk=args.walk_chunk)
    if args.autotune:
        dev.autotune(modes, pmin, pmin + 2 * 10**9, nmin, nmax)
    total = args.steps + args.warmup
    # p-range: the headline 2e9 window, or wider if more segments are asked for
    job = P.Job(dev, modes, pmin, pmin + 2 * 10**9 * max(1, (total + 16) // 17), nmin, nmax)
    for _ in range(args.warmup):
        job.run(1)
    s0 = job.stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        job.run(1)
    t1 = time.perf_counter()
    s1 = job.stats()
    cand = s1.primecount - s0.primecount
    updates = cand * steps_per_cand
    wall = t1 - t0
    dev_ms = s1.device_ms - s0.device_ms
    ups = updates / wall
    job.destroy()

    # Dominant kernel: the walk.  Its launches are bracketed by HIP events on
    # the job's stream inside the library; achieved = algorithmic multiplies
    # of those launches / their summed duration.
    walk_ms = s1.walk_ms - s0.walk_ms
    walk_launches = s1.walk_launches - s0.walk_launches
    walk_updates = s1.walk_updates - s0.walk_updates
    achieved = walk_updates * W_MUL / (walk_ms * 1e-3) / 1e12
    # The same launches counted by the multiplies they issue, against the measured
    # v_mad_u64_u32 lane rate.
    mpu = mads_per_update(nmin, nmax)
    issued = {"mad_u64_u32_per_update": mpu, "achieved": walk_updates * mpu / (walk_ms * 1e-3) / 1e12,
              "peak": rates["mad_u64_u32"] / 1e12, "unit": "Tops/s"}
    issued["frac"] = issued["achieved"] / issued["peak"]
    line = {
        "metric": "p\u00b7n residue updates/sec on 1 MI355X, batched vs plain; % INT32 VALU peak",
        "value": ups, "unit": "updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64 (exact modular integer)", "data": "synthetic (device-generated primes)",
        "config": {"workload": "-! -c, n in [101, 1e6), p from 1e12 (headline config, per p-segment)",
                   "form": "plain: one range per launch sequence, as the headline configuration is defined "
                           "(the batched form is under 'batched', the INT32 fraction under 'roofline')",
                   "candidates_timed": cand, "steps_per_candidate": steps_per_cand,
                   "seg_candidates": args.seg_candidates, "tuning": dev.tuning(), "device_ms": dev_ms,
                   "walk_launches": walk_launches, "walk_ms_per_launch": walk_ms / max(walk_launches, 1),
                   "walk_updates_per_s": walk_updates / (walk_ms * 1e-3)},
        "roofline": {"bound": "valu-int32-mul", "achieved": achieved, "peak": p_mul / 1e12,
                     "unit": "Tops/s", "frac": achieved / (p_mul / 1e12),
                     "traffic": (WALK_FETCH_PER_CAND + WALK_WRITE_PER_CAND) * cand / max(args.steps, 1),
                     "issued": issued,
                     "note": f"walk kernel: {W_MUL} INT32 multiplies per update x updates per launch / "
                             f"average launch duration (HIP events); peak = measured lane-op rate "
                             f"(mad_u64_u32 {rates['mad_u64_u32']:.3e}, mul_lo {rates['mul_lo_u32']:.3e}, "
                             f"mul_hi {rates['mul_hi_u32']:.3e}); traffic = HBM bytes per walk launch "
                             f"(PMC bytes per candidate x candidates per segment)"},
    }
    if args.configs:
        line["other_configs"] = other_configs(dev)
    if args.sweep:
        line["shape_sweep"] = shape_sweep(dev, modes, pmin, nmin, nmax)
        dev.set_tuning(seg_candidates=args.seg_candidates, primes_per_lane=args.ppl, block_size=args.block,
                       walk_chunk=args.walk_chunk)
    if args.mix_units:
        from pfcsieve_amd import workload as W
        units = W.mix(256, seed=12345)[: args.mix_units]
        line["batched"] = mix_compare(dev, units, progress=args.progress)
        line["batched"]["mix"] = ("first %d of 256 seeded units (splitmix64 seed 12345): modes F/P/C equal odds, "
                                  "n-width log-U[1e3,1e6] from 101+U[0,1e6], p-width log-U[1e6,1e9] from "
                                  "log-U[1e10,2^63]" % args.mix_units)
    if args.serve_jobs:
        from pfcsieve_amd import workload as W
        units = W.mix(10_000, seed=2024)[: args.serve_jobs]
        line["served"] = serve_run(dev, units, round_updates=int(args.serve_round_updates), progress=args.progress)
        line["served"]["stream"] = ("first %d of 10,000 seeded jobs (splitmix64 seed 2024, same unit distribution "
                                    "as the batched mix), Poisson arrivals at 0.8x measured batch capacity"
                                    % args.serve_jobs)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(modes, pmin, args.cpu_width, nmin, nmax, steps_per_cand)
        line["cpu_baseline"]["gpu_over_cpu"] = ups / line["cpu_baseline"]["value"]
        line["cpu_baseline"]["verify"] = verify_baseline(dev)
    dev.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
