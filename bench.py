"""Benchmark of the sieve hot path on one MI355X.

Prints ONE JSON line.  The metric is p*n residue updates per second (updates
= candidates x multiplier steps; F and C walks of a -! -c job counted
separately), batched against plain, plus the walk's fraction of the INT32
VALU multiply peak.

Workloads (synthetic: every candidate prime is generated on the device):

* ``value`` -- the batched uneven mix: all 256 units of the seeded mix
  (splitmix64 seed 12345; modes F/P/C with equal odds, n-width log-U[1e3,1e6]
  from 101+U[0,1e6], p-width log-U[1e6,1e9] from log-U[1e10,2^63]), every
  unit one p-segment.  Batched: all units in one batcher, timed from job
  creation to job destruction; a step is one batcher round (its device
  phase -- work-item upload, queue clear, the persistent set-up, per-mode
  walk and fold launches, control-block readback -- replayed from one
  captured HIP graph), and the K
  rounds are sized by work (round r walks about
  the remaining updates / (K - r); the last walks everything left), so every
  step does work and the K steps cover the mix.  Plain (``mix.plain``): the
  same units one job after another, one range per launch sequence.
  Warmup: W batcher rounds over four small units of another seed.
* ``headline`` -- the plain -! -c job over exactly p in [1e12, 1e12+2e9),
  n in [101, 1e6): one p-segment holding all its ~72.38M candidates, walked
  in K launches of ceil(999,899 / K) multiplier steps (a step is one walk
  launch; generate and set-up run before the first, the checksum fold after
  the last); the whole job is timed.  Warmup: the same n-range on
  p in [1e12 - 1e8, 1e12), walked in W launches.
* ``roofline`` -- the walk kernel (k_walk) of the headline job: INT32
  multiply instructions it issues (counted from its instruction sequence per
  launch: five per update in the narrow form the headline takes) per second
  of launch time (HIP events on the job's stream) against 256 CUs x 64 lanes
  x 2.4 GHz.
* ``served`` -- a prefix of the seeded served stream (splitmix64 seed 2024,
  units drawn as in the mix), Poisson arrivals at 0.8x the capacity that the
  prefix's own device-work bound gives at the kernels' measured rates (walk,
  set-up, generation), admitted by HBM footprint, rounds of about 4e11 p*n
  updates shared equally among the resident jobs, finished jobs harvested
  beside the next rounds; throughput and p50/p99 job latency.
* ``cpu_baseline`` -- the CPU oracle (OpenMP over candidates) on a bounded
  p-sample of the headline job, plus the reference's CPU-runnable case in
  full.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_LANE_OPS = 256 * 64 * 2.4e9   # MI355X: 256 CUs x 64 lanes x 2.4 GHz peak engine clock
W_MUL_REF = 19                     # INT32 multiplies per update in the reference's 64-bit Montgomery product
PRODUCT_LIMIT = 65536              # pairs with both multipliers up to this are one step by their product
# HBM bytes per candidate per walk launch of the headline job (FETCH_SIZE x
# the copy-kernel calibration, and WRITE_SIZE, of k_walk: p, v and both
# residues read, both residues written, plus the launch's composite-list
# reads amortised over its candidates).
WALK_FETCH_PER_CAND = 28.19
WALK_WRITE_PER_CAND = 16.50

HEAD_MODES, HEAD_PMIN, HEAD_PMAX, HEAD_NMIN, HEAD_NMAX = 5, 10**12, 10**12 + 2 * 10**9, 101, 10**6


def _progress(on, msg):
    if on:
        print(msg, file=sys.stderr, flush=True)


def prime_flags(n):
    import numpy as np
    s = np.ones(n + 1, dtype=bool)
    s[:2] = False
    for i in range(2, int(n ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    return s


def narrow_fits(p, bmax):
    """The walk's narrow form (plain residues below 2p, five multiplies per
    step, pfc::narrow_fits): 2^34 < p < 2^55, 1 <= bmax < 2^22, p * bmax < 2^62."""
    return 2**34 < p < 2**55 and 1 <= bmax < 2**22 and p * bmax < 2**62


def walk_muls(nmin, nmax, chunk, pmax):
    """INT32 multiply instructions per candidate that the F|C walk of
    n in [nmin, nmax) issues in launches of `chunk` steps, with every step
    paired (all p above twice the largest n), for candidates up to pmax.
    A launch whose largest multiplier fits the narrow form with pmax takes
    five per update (v_mul_lo_u32, v_mad_u64_u32, v_mul_hi_u32,
    v_mad_i32_i24, v_mad_u64_u32).  Otherwise, for pmax below 2^63, the
    medium form takes six per multiplier step (v_mul_lo_u32, v_mad_u64_u32,
    two v_mul_hi_u32, v_mul_lo_u32, v_mad_u64_u32), a pair with product
    below 2^28 being one step; above it the shifted 3-by-2 step issues five
    v_mad_u64_u32 per multiplier step and per product step of two
    multipliers up to PRODUCT_LIMIT.  Pairs are the factorial trips of four
    and the composite-list blocks of 16, aligned to the list.  Returns
    (multiplies, updates, product updates, narrow updates) per candidate."""
    import numpy as np
    isp = prime_flags(nmax)
    comp = np.nonzero(~isp[nmin:nmax])[0] + nmin          # composites in [nmin, nmax), the C list
    medium = pmax < 2**63
    per_step = 6.0 if medium else 5.0
    muls = prod = narrow = 0.0
    for a in range(nmin, nmax, chunk):
        b = min(nmax, a + chunk)
        c0 = int(np.searchsorted(comp, a))
        c1 = int(np.searchsorted(comp, b))
        if narrow_fits(pmax, b - 1):
            narrow += (b - a) + (c1 - c0)
            muls += 5.0 * ((b - a) + (c1 - c0))
            continue
        lim = 16383 if medium else PRODUCT_LIMIT           # medium: m * n < 2^28
        pp = 0
        jp = min(b, lim + 1)
        if jp > a:
            pp += (jp - a) // 4 * 4
        j = min(c1, (c0 + 15) & ~15)
        while j + 16 <= c1:
            if comp[j + 15] <= lim:
                pp += 16
            j += 16
        prod += pp
        muls += per_step * ((b - a) + (c1 - c0)) - per_step / 2 * pp
    upd = (nmax - nmin) + len(comp)
    return muls, upd, prod, narrow


def measure_headline(dev, K, W, progress):
    """The plain -! -c job over exactly [1e12, 1e12 + 2e9): one p-segment,
    K walk launches; warmup on [1e12 - 1e8, 1e12) in W launches."""
    import pfcsieve_amd as P
    nsteps = HEAD_NMAX - HEAD_NMIN
    chunk = -(-nsteps // K)
    try:
        if W:
            dev.set_tuning(seg_candidates=1 << 27, walk_chunk=-(-nsteps // W))
            wj = P.Job(dev, HEAD_MODES, HEAD_PMIN - 10**8, HEAD_PMIN, HEAD_NMIN, HEAD_NMAX)
            wj.run()
            ws = wj.stats()
            wj.destroy()
            _progress(progress, f"headline warmup: {ws.walk_launches} launches over {ws.primecount} candidates")
        dev.set_tuning(seg_candidates=1 << 27, walk_chunk=chunk)
        job = P.Job(dev, HEAD_MODES, HEAD_PMIN, HEAD_PMAX, HEAD_NMIN, HEAD_NMAX)
        t0 = time.perf_counter()
        job.run()
        wall = time.perf_counter() - t0
        s = job.stats()
        job.destroy()
    finally:
        dev.set_tuning()
    assert s.segments == 1 and s.walk_launches == K, (s.segments, s.walk_launches)
    _progress(progress, f"headline: {s.primecount} candidates, {wall:.1f} s")
    return {"workload": "-! -c n in [101, 1e6) p in [1e12, 1e12+2e9): plain, one range per launch sequence",
            "pmin": HEAD_PMIN, "pmax": HEAD_PMAX, "nmin": HEAD_NMIN, "nmax": HEAD_NMAX,
            "steps": K, "warmup": W,
            "step": f"one k_walk launch of {chunk} multiplier steps over every candidate (generate and set-up "
                    f"before the first, the checksum fold after the last)",
            "candidates_timed": s.primecount, "steps_per_candidate": s.walk_updates // max(s.primecount, 1),
            "segments": s.segments, "updates": s.walk_updates, "wall_s": wall, "ms_per_step": wall * 1e3 / K,
            "updates_per_s": s.walk_updates / wall, "device_ms": s.device_ms,
            "walk_launches": s.walk_launches, "walk_ms": s.walk_ms,
            "walk_ms_per_launch": s.walk_ms / s.walk_launches,
            "walk_updates_per_s": s.walk_updates / (s.walk_ms * 1e-3), "walk_chunk": chunk}


def roofline(head, mul_rates):
    """Roofline figures of k_walk on the headline job."""
    muls, upd, prod, narrow = walk_muls(HEAD_NMIN, HEAD_NMAX, head["walk_chunk"], HEAD_PMAX - 1)
    launches = head["walk_launches"]
    avg_s = head["walk_ms"] * 1e-3 / launches
    cand = head["candidates_timed"]
    achieved = muls * cand / launches / avg_s
    ups = head["walk_updates_per_s"]
    mad_rate = mul_rates["mad_u64_u32"]
    return {
        "bound": "valu-int32", "kernel": "k_walk on the headline job",
        "achieved": achieved / 1e12, "peak": PEAK_LANE_OPS / 1e12, "unit": "Tops/s",
        "frac": achieved / PEAK_LANE_OPS,
        "traffic": (WALK_FETCH_PER_CAND + WALK_WRITE_PER_CAND) * cand,
        "mul_per_update": muls / upd, "narrow_share": narrow / upd, "updates_per_launch": upd * cand / launches,
        "avg_launch_ms": avg_s * 1e3,
        "frac_of_measured_mad_rate": achieved / mad_rate,
        "issue": "the walk is bound by VALU issue: in the narrow form (p * n < 2^62) each update is five "
                 "32-bit multiply instructions, and every pair of updates adds a compare and a masked add "
                 "(the reduction below p) and a 64-bit add and compare (the band test); rocprof VALU and "
                 "SALU counts per update in profiles/",
        "reference_equivalent": {"mul_per_update": W_MUL_REF, "achieved": ups * W_MUL_REF / 1e12,
                                 "frac": ups * W_MUL_REF / PEAK_LANE_OPS,
                                 "note": "updates/s priced at the reference's 19 INT32 multiplies per update "
                                         "(its 64-bit Montgomery product); not instructions this walk issues"},
        "measured_lane_rates": {k: v / 1e12 for k, v in mul_rates.items()},
        "note": "achieved = INT32 multiply instructions issued per launch (walk_muls: 5 per update in the "
                "narrow form) / average k_walk launch duration (HIP events on the job's stream); peak = "
                "256 CU x 64 lanes x 2.4 GHz; traffic = HBM bytes per launch (rocprof FETCH_SIZE x copy "
                "calibration + WRITE_SIZE, per candidate, x candidates per launch)",
    }


def _unit_updates(u):
    """Walk updates of a unit from its p-width (prime density 1/ln p)."""
    from pfcsieve_amd import workload as Wl
    return (u.pmax - u.pmin) / math.log(u.pmin) * Wl.walk_steps(u)


def batched_rounds(dev, units, K, progress, tag="batched"):
    """Runs units through one batcher in K rounds sized by work: round r
    walks about the estimated remaining updates / (K - r) (pfc_batcher_round_work),
    the last one everything left (every unit is one p-segment, so that ends
    the batch).  Returns (wall seconds, rounds, per-job (checksum, primecount,
    factors), per-job stats)."""
    import pfcsieve_amd as P
    t = time.perf_counter()
    jobs = [P.Job(dev, *u.desc()) for u in units]
    b = P.Batcher(dev)
    for j in jobs:
        b.add(j)
    left_est = sum(_unit_updates(u) for u in units)
    rounds = 0
    while b.pending():
        if K - rounds <= 1:
            b.round(0)
        else:
            budget = left_est / (K - rounds)
            b.round_work(max(1, int(budget)))
            left_est -= budget
        rounds += 1
        _progress(progress, f"{tag}: round {rounds} done, {b.pending()} jobs pending, {time.perf_counter() - t:.1f} s")
    stats = [j.stats() for j in jobs]
    sigs = [j.result() for j in jobs]
    for j in jobs:
        j.destroy()
    b.destroy()
    return time.perf_counter() - t, rounds, sigs, stats


def measure_mix(dev, K, W, progress, nunits=256):
    """All units of the seeded 256-unit mix, plain (one job after another,
    from job creation to destruction) and batched (one batcher, K rounds);
    every unit's checksum, candidate count and raw factor records must agree."""
    import pfcsieve_amd as P
    from pfcsieve_amd import workload as Wl
    units = Wl.mix(256, seed=12345)[:nunits]
    steps = [Wl.walk_steps(u) for u in units]
    dev.set_tuning(seg_candidates=1 << 26)        # every unit (p-width <= 1e9) is one p-segment
    try:
        if W:
            warm = sorted(Wl.mix(64, seed=4242), key=_unit_updates)[:4]
            batched_rounds(dev, warm, W, False, "warmup")
        t = time.perf_counter()
        plain = []
        for k, u in enumerate(units):
            j = P.Job(dev, *u.desc())
            j.run()
            plain.append(j.result())
            j.destroy()
            _progress(progress, f"plain: unit {k + 1}/{len(units)} done, {time.perf_counter() - t:.1f} s")
        t_plain = time.perf_counter() - t
        # batched: every round's device phase replayed from one captured HIP graph
        dev.set_tuning(seg_candidates=1 << 26, use_graph=1)
        t_batch, rounds, batched, stats = batched_rounds(dev, units, K, progress)
    finally:
        dev.set_tuning()
    cand = [r[1] for r in batched]
    upd = sum(c * s for c, s in zip(cand, steps))
    upd_narrow = sum(c * narrow_steps(u) for c, u in zip(cand, units))
    products = sum(c * setup_products(u) for c, u in zip(cand, units))
    return {"units": len(units), "candidates": sum(cand), "walk_updates": upd, "walk_updates_narrow": upd_narrow,
            "setup_products": products,
            "segments": sum(s.segments for s in stats), "plain_s": t_plain, "batched_s": t_batch,
            "rounds": rounds, "plain_updates_per_s": upd / t_plain, "batched_updates_per_s": upd / t_batch,
            "batched_over_plain": t_plain / t_batch, "results_identical": plain == batched}


_TABLE_COST = {}


def setup_products(u):
    """Modular products per candidate for the start residue X(nmin - 1)."""
    import pfcsieve_amd as P
    key = ({1: 0, 2: 1, 4: 2}[u.modes], u.nmin - 1)
    if key not in _TABLE_COST:
        _TABLE_COST[key] = P.setup_table_cost(*key)
    return _TABLE_COST[key]


def narrow_steps(u):
    """Walk steps per candidate of unit u that the narrow form takes for
    every p of the unit: those whose multiplier n satisfies
    narrow_fits(pmax - 1, n) (launch boundaries ignored)."""
    from pfcsieve_amd import workload as Wl
    pmax = u.pmax - 1
    if not (2**34 < u.pmin and pmax < 2**55):
        return 0
    nlim = min(2**22 - 1, (2**62 - 1) // pmax)
    hi = min(u.nmax, nlim + 1)
    return Wl.walk_steps(Wl.Unit(u.modes, u.pmin, u.pmax, u.nmin, hi)) if hi > u.nmin else 0


def stream_bound(units, rates):
    """Device-work bound of a list of units at the kernels' measured rates:
    walk updates at rates["walk_narrow"] where the narrow form applies and
    rates["walk"] (the medium form) elsewhere, start-residue products at rates["setup"],
    candidates at rates["generate"] (candidate counts from the prime density
    1/ln p)."""
    from pfcsieve_amd import workload as Wl
    t = 0.0
    for u in units:
        c = (u.pmax - u.pmin) / math.log(u.pmin)
        t += c * setup_products(u) / rates["setup"] + c / rates["generate"]
        steps, nsteps = Wl.walk_steps(u), narrow_steps(u)
        t += c * (steps - nsteps) / rates["walk"] + c * nsteps / rates["walk_narrow"]
    return t


def serve_prefix(dev, njobs, rates, round_updates, progress, start=0, full_rate=False, latencies_path=None,
                 rate=0.0, overlap=True):
    """Jobs [start, start + njobs) of the seeded 10,000-job stream through the
    serving loop.  Poisson arrivals at 0.8 x the capacity that a device-work
    bound gives (stream_bound at the measured kernel rates), instead of a
    separate capacity run: the bound of the jobs served, or with full_rate
    (and for any slice not starting at job 0) that of the whole stream, so
    that slices of it run at the whole stream's rate.  A slice keeps the
    stream's own arrival times, shifted so that its first job arrives one
    inter-arrival gap after the clock starts; it starts with nothing
    resident.  latencies_path: write every job's (index, arrival, latency)
    as JSON there.  rate > 0 fixes the arrival rate (jobs/s) instead.
    overlap=False harvests finished jobs between rounds instead of on worker
    threads beside them."""
    import pfcsieve_amd.serve as S
    from pfcsieve_amd import workload as Wl
    stream = Wl.mix(10_000, seed=2024)
    units = stream[start:start + njobs]
    full = full_rate or start > 0
    bound_all = stream_bound(stream, rates) if full else None
    bound = stream_bound(units, rates)
    lam = rate if rate > 0 else 0.8 * len(stream) / bound_all if full else 0.8 * njobs / bound
    arrivals = Wl.poisson_arrivals(start + njobs, lam, seed=777)
    shift = arrivals[start - 1] if start > 0 else 0.0
    arrivals = [a - shift for a in arrivals[start:start + njobs]]
    info = dev.info()
    say = (lambda m: _progress(True, m)) if progress else None
    srv = S.Server(dev, hbm_budget=int(info["hbm_bytes"] * 0.8), verify_threads=16, progress=say,
                   round_updates=round_updates, overlap=overlap)
    for k, (u, a) in enumerate(zip(units, arrivals)):
        srv.submit(S.Request(start + k, u.desc(), a))
    st = srv.run()
    upd = sum(r.result.primecount * Wl.walk_steps(stream[r.rid]) for r in srv.done)
    if latencies_path:
        with open(latencies_path, "w") as f:
            json.dump([[r.rid, r.arrival, r.finished - r.arrival] for r in srv.done], f)
    which = f"first {njobs}" if start == 0 else f"jobs [{start}, {start + njobs})"
    return {"jobs": st.jobs, "stream": f"{which} of 10,000 seeded jobs (splitmix64 seed 2024, units as in "
                                       f"the mix), Poisson arrivals (seed 777)"
                                       + (" at the whole stream's rate" if full else ""),
            "device_work_bound_s": bound, "device_work_bound_stream_s": bound_all,
            "arrival_rate_per_s": lam, "last_arrival_s": arrivals[-1],
            "wall_s": st.wall_s, "rounds": st.rounds, "updates_per_s": upd / st.wall_s,
            "jobs_per_s": st.jobs / st.wall_s, "p50_latency_s": st.pct(50), "p99_latency_s": st.pct(99),
            "max_latency_s": max(st.latencies) if st.latencies else None,
            "host_split_s": {"admit": st.admit_s, "rounds": st.round_s, "harvest": st.harvest_s,
                             "idle": st.idle_s, "harvest_beside_rounds": st.overlap},
            "peak_resident_bytes": st.peak_resident_bytes, "round_policy": f"about {round_updates:.3g} p*n "
                                                                           f"updates per round"}


def kernel_rates(dev, reps=3):
    """Standalone rates of the kernels around the walk: candidate generation
    (k_generate through pfc_stage_generate, the whole p-window of the
    headline job and of the primorial job near 2^62, best of `reps`), and
    start-residue set-up (k_setup) of 999999! per candidate on one
    p-segment from p = 1e15."""
    import ctypes as C
    import pfcsieve_amd as P
    import pfcsieve_amd._native as N
    from pfcsieve_amd.sieve import DeviceBuffer
    out = {}
    for name, lo, hi in (("generate_p1e12_2e9", 10**12, 10**12 + 2 * 10**9),
                         ("generate_p2e62_4e9", 2**62 - 4 * 10**9, 2**62)):
        span = hi - lo
        cap = int(2.2 * span / math.log(span)) + 4096
        d_p = DeviceBuffer(dev, 8 * cap)
        ctrl = DeviceBuffer(dev, N.CTRL_BYTES)
        best = None
        for _ in range(reps):
            ctrl.zero()
            N.check(N.lib().pfc_device_synchronize(dev.handle))
            t = time.perf_counter()
            N.check(N.lib().pfc_stage_generate(lo, hi, d_p.ptr, cap, ctrl.ptr, None))
            N.check(N.lib().pfc_device_synchronize(dev.handle))
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        n = int.from_bytes(ctrl.download(4), "little")
        out[name] = {"window": f"p in [{lo}, {hi})", "candidates": n, "s": best, "candidates_per_s": n / best,
                     "integers_per_s": span / best}
        d_p.free()
        ctrl.free()
    # The medium walk (p * n above 2^62, p below 2^63): -! -c, n in [9e5, 1e6),
    # p from 2^61.
    j = P.Job(dev, 5, 2**61, 2**61 + 5 * 10**8, 900_000, 10**6)
    j.run()
    s = j.stats()
    j.destroy()
    out["walk_medium_p2e61"] = {"workload": "-! -c n [9e5, 1e6) p in [2^61, 2^61 + 5e8): the medium walk "
                                            "(p * n > 2^62, p < 2^63), walk launches timed by HIP events",
                                 "candidates": s.primecount, "walk_updates": s.walk_updates,
                                 "walk_ms": s.walk_ms, "walk_updates_per_s": s.walk_updates / (s.walk_ms * 1e-3)}
    products = P.setup_table_cost(0, 10**6 - 1)
    j = P.Job(dev, 1, 10**15, 10**15 + 2 * 10**9, 10**6, 10**6 + 256)
    j.run(1)
    s = j.stats()
    j.destroy()
    setup_s = (s.device_ms - s.walk_ms) * 1e-3
    out["setup_nmin_1e6"] = {"workload": "-! n [1e6, 1e6+256) p from 1e15, one p-segment: start residue "
                                         "999999! mod p per candidate (generate and check included, ~1%)",
                             "candidates": s.primecount, "modular_products_per_candidate": products,
                             "s": setup_s, "candidates_per_s": s.primecount / setup_s,
                             "products_per_s": s.primecount * products / setup_s}
    return out


def walk_forms(dev):
    """The walk of one p-segment (p-width 1e8) in each arithmetic form its
    moduli allow (pfc_tuning.walk_forms_off): -! -c at n in [101, 2e5) and
    [9e5, 1e6) from p = 1e12 and at n in [9e5, 1.3e6) from 3e12 (p * n
    between 2^61 and 2^62: the top of the narrow range) (narrow, medium,
    shifted), from p = 2^61
    (medium, shifted), -! at n in [9e5, 1e6) from 2^61, -# at n in
    [101, 1e6) from p = 2^62 - 1e8 (medium, shifted), and -! -c and -! at
    n in [101, 65536) (every pair a product step in the shifted form) from
    1e12 and 2^61; walk launches timed by HIP events, the same candidates and
    multipliers in every form."""
    import pfcsieve_amd as P
    from pfcsieve_amd import workload as Wl
    out = {}
    for name, modes, pmin, nmin, nmax, forms in (
            ("fc_p1e12", 5, 10**12, 101, 200_000, (("narrow", 0), ("medium", 1), ("shifted", 3))),
            ("fc_p1e12_n9e5", 5, 10**12, 900_000, 1_000_000, (("narrow", 0), ("medium", 1), ("shifted", 3))),
            ("fc_p3e12_n9e5", 5, 3 * 10**12, 900_000, 1_300_000, (("narrow", 0), ("medium", 1), ("shifted", 3))),
            ("fc_p2e61", 5, 2**61, 101, 200_000, (("medium", 0), ("shifted", 2))),
            ("fc_p2e61_n9e5", 5, 2**61, 900_000, 1_000_000, (("medium", 0), ("shifted", 2))),
            ("f_p2e61_n9e5", 1, 2**61, 900_000, 1_000_000, (("medium", 0), ("shifted", 2))),
            ("p_p2e62", 2, 2**62 - 10**8, 101, 1_000_000, (("medium", 0), ("shifted", 2))),
            ("fc_p1e12_n64k", 5, 10**12, 101, 65_536, (("narrow", 0), ("medium", 1), ("shifted", 3))),
            ("fc_p2e61_n64k", 5, 2**61, 101, 65_536, (("medium", 0), ("shifted", 2))),
            ("f_p2e61_n64k", 1, 2**61, 101, 65_536, (("medium", 0), ("shifted", 2)))):
        rows = {}
        for form, off in forms:
            dev.set_tuning(walk_forms_off=off, seg_candidates=1 << 24)
            try:
                j = P.Job(dev, modes, pmin, pmin + 10**8, nmin, nmax)
                j.run()
                s = j.stats()
                res = j.result()
                j.destroy()
            finally:
                dev.set_tuning()
            rows[form] = {"walk_launches": s.walk_launches, "walk_ms": s.walk_ms,
                          "walk_updates_per_s": s.walk_updates / (s.walk_ms * 1e-3),
                          "checksum": f"{res[0]:016X}"}
        assert len({r["checksum"] for r in rows.values()}) == 1, rows
        out[name] = {"workload": f"modes {modes}, n in [{nmin}, {nmax}), p in [{pmin}, {pmin + 10**8})",
                     "candidates": s.primecount,
                     "steps_per_candidate": (Wl.walk_steps(Wl.Unit(1, 0, 0, nmin, nmax))
                                             + Wl.walk_steps(Wl.Unit(4, 0, 0, nmin, nmax))) if modes == 5
                     else Wl.walk_steps(Wl.Unit(modes, 0, 0, nmin, nmax)),
                     "forms": rows}
    return out


def walk_modes(dev):
    """One 2^24-candidate p-segment of each walk mode at the shapes the mixes
    use: -! and -c alone at n in [101, 1e6) from p = 1e12, and -# at
    n in [101, 1e7) from p = 2^62 - 4e9 (walk launches timed by HIP events)."""
    import pfcsieve_amd as P
    from pfcsieve_amd import workload as Wl
    out = {}
    for name, modes, pmin, nmin, nmax in (("factorial", 1, 10**12, 101, 10**6), ("compositorial", 4, 10**12, 101, 10**6),
                                          ("primorial", 2, 2**62 - 4 * 10**9, 101, 10**7)):
        j = P.Job(dev, modes, pmin, pmin + 4 * 10**9, nmin, nmax)
        j.run(1)
        s = j.stats()
        j.destroy()
        out[name] = {"workload": f"modes {modes}, n in [{nmin}, {nmax}), one p-segment from {pmin}",
                     "candidates": s.primecount, "steps_per_candidate": Wl.walk_steps(Wl.Unit(modes, 0, 0, nmin, nmax)),
                     "walk_launches": s.walk_launches, "walk_ms": s.walk_ms,
                     "walk_updates_per_s": s.walk_updates / (s.walk_ms * 1e-3)}
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(width):
    """The oracle (C, OpenMP over candidates, all host threads) on a bounded
    p-sample of the headline job, and the reference's CPU-runnable case
    (-! n [101, 5000) p [1e9, 1.001e9)) in full."""
    import oracle as O
    O.build()
    steps = walk_muls(HEAD_NMIN, HEAD_NMAX, HEAD_NMAX, HEAD_PMAX - 1)[1]
    t = time.perf_counter()
    r = O.run(HEAD_MODES, HEAD_PMIN, HEAD_PMIN + width, HEAD_NMIN, HEAD_NMAX)
    dt = time.perf_counter() - t
    out = {"value": r.primecount * steps / dt, "unit": "updates/s", "cores": O.num_threads(), "kind": "port",
           "sample": f"-! -c p in [{HEAD_PMIN}, {HEAD_PMIN + width}) ({r.primecount} candidates), "
                     f"n in [{HEAD_NMIN}, {HEAD_NMAX}); {dt:.1f} s",
           "algorithm": "the oracle's walk: one 64-bit Montgomery product per update (the reference's "
                        "formulation, iterate.cl:156-198), where the GPU walk divides by 32-bit multipliers "
                        "(3-by-2 division, paired tests); a GPU/CPU ratio mixes that change with the hardware",
           "cpu_model": _cpu_model(), "nproc": os.cpu_count()}
    t = time.perf_counter()
    r1 = O.run(1, 10**9, 10**9 + 10**6, 101, 5000)
    dt1 = time.perf_counter() - t
    out["factorial_n5000_p1e9_full"] = {"candidates": r1.primecount, "s": dt1,
                                        "updates_per_s": r1.primecount * 4899 / dt1,
                                        "checksum": "%016X" % r1.checksum}
    return out


def small_job(dev):
    """The reference's CPU-runnable case on the GPU, whole job including
    table self-check and harvest: -! n [101, 5000) p [1e9, 1.001e9)."""
    import pfcsieve_amd as P
    t = time.perf_counter()
    r = P.run_sieve(1, 10**9, 10**9 + 10**6, 101, 5000, device=dev, verify_threads=16)
    dt = time.perf_counter() - t
    return {"workload": "-! n [101,5000) p [1e9,1.001e9)", "candidates": r.primecount, "wall_s": dt,
            "updates_per_s": r.primecount * 4899 / dt, "factors": r.factorcount, "checksum": "%016X" % r.checksum}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="K: batcher rounds of the mix, walk launches of the headline")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mix-units", type=int, default=256, help="units of the seeded mix (256 = the whole mix)")
    ap.add_argument("--serve-jobs", type=int, default=200, help="jobs of the seeded served stream (0 = skip)")
    ap.add_argument("--serve-from", type=int, default=0, help="first job of the served slice of the stream")
    ap.add_argument("--serve-full-rate", action="store_true",
                    help="arrival rate from the whole 10,000-job stream's device-work bound (slices)")
    ap.add_argument("--serve-latencies", default="", help="write every served job's latency to this JSON file")
    ap.add_argument("--serve-rate", type=float, default=0.0, help="fixed arrival rate, jobs/s (0: from the bound)")
    ap.add_argument("--serve-serial", action="store_true",
                    help="harvest served jobs between rounds, not on worker threads beside them")
    ap.add_argument("--serve-round-updates", type=float, default=4e11)
    ap.add_argument("--no-headline", action="store_true")
    ap.add_argument("--kernel-rates", action="store_true", help="also time k_generate and k_setup standalone")
    ap.add_argument("--modes", action="store_true", help="also time one p-segment of each walk mode")
    ap.add_argument("--forms", action="store_true", help="also time one p-segment in each walk arithmetic form")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="run one plain device copy of this many bytes first (HBM counter calibration)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-width", type=int, default=1_000_000)
    ap.add_argument("--progress", action="store_true", help="progress lines on stderr")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on exactly one GPU; sharding over p across GPUs is "
                                   "designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)
    K, W = max(1, args.steps), max(0, args.warmup)

    import pfcsieve_amd as P
    dev = P.Device(0)
    if args.copy_probe:
        dev.probe_copy(args.copy_probe)
    rates = dev.probe_mul_rates()
    head = None if args.no_headline else measure_headline(dev, K, W, args.progress)
    mix = measure_mix(dev, K, W, args.progress, args.mix_units) if args.mix_units else None
    line = {
        "metric": "p·n residue updates/sec on 1 MI355X, batched vs plain; % INT32 VALU peak",
        "value": mix["batched_updates_per_s"] if mix else head["updates_per_s"] if head else None,
        "unit": "updates/s", "n_gpus": 1, "steps": mix["rounds"] if mix else K, "warmup": W,
        "ms_per_step": (mix["batched_s"] / mix["rounds"] * 1e3 if mix else head["wall_s"] / K * 1e3 if head
                        else None),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 (exact modular integer)", "data": "synthetic (seeded job mix; device-generated primes)",
        "config": {"workload": (f"batched uneven mix: {mix['units']} seeded units (splitmix64 seed 12345)"
                                if mix else "headline -! -c job" if head else "kernel rates only"),
                   "form": "batched: all units through one batcher, each round's persistent set-up, per-mode "
                           "walk and fold launches replayed from a captured HIP graph; a step is one batcher "
                           "round, the K rounds sized by work; plain (one job after another, eager) beside it "
                           "under 'mix'",
                   "mix": mix, "headline": head, "tuning": dev.tuning()},
    }
    if head:
        line["roofline"] = roofline(head, rates)
    if args.kernel_rates or (head and mix):
        line["kernel_rates"] = kernel_rates(dev)
    if head and mix:
        # The mix's device work at the kernels' own measured rates: walk updates
        # at the headline walk rate (the narrow form) where the narrow form
        # applies and at the medium walk's rate elsewhere (every p of the mix
        # is below 2^63), start-residue
        # products at the set-up rate, candidates at the generator's rate on
        # the 1e12 window.
        kr = line["kernel_rates"]
        medium = kr["walk_medium_p2e61"]["walk_updates_per_s"]
        bound = (mix["walk_updates_narrow"] / head["walk_updates_per_s"]
                 + (mix["walk_updates"] - mix["walk_updates_narrow"]) / medium
                 + mix["setup_products"] / kr["setup_nmin_1e6"]["products_per_s"]
                 + mix["candidates"] / kr["generate_p1e12_2e9"]["candidates_per_s"])
        mix["device_work_bound_s"] = bound
        mix["batched_frac_of_bound"] = bound / mix["batched_s"]
        mix["plain_frac_of_bound"] = bound / mix["plain_s"]
    if args.serve_jobs:
        if "kernel_rates" not in line:
            line["kernel_rates"] = kernel_rates(dev)
        kr = line["kernel_rates"]
        if head:
            walk_narrow = head["walk_updates_per_s"]
        else:
            line["walk_modes"] = line.get("walk_modes") or walk_modes(dev)
            walk_narrow = line["walk_modes"]["factorial"]["walk_updates_per_s"]
        rates = {"walk": kr["walk_medium_p2e61"]["walk_updates_per_s"], "walk_narrow": walk_narrow,
                 "setup": kr["setup_nmin_1e6"]["products_per_s"],
                 "generate": kr["generate_p1e12_2e9"]["candidates_per_s"]}
        line["served"] = serve_prefix(dev, args.serve_jobs, rates, int(args.serve_round_updates), args.progress,
                                      start=args.serve_from, full_rate=args.serve_full_rate,
                                      latencies_path=args.serve_latencies or None, rate=args.serve_rate,
                                      overlap=not args.serve_serial)
        line["served"]["rates"] = rates
    line["config"]["small_job"] = small_job(dev)
    if args.kernel_rates and "kernel_rates" not in line:
        line["kernel_rates"] = kernel_rates(dev)
    if args.modes and "walk_modes" not in line:
        line["walk_modes"] = walk_modes(dev)
    if args.forms:
        line["walk_forms"] = walk_forms(dev)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_width)
        if head:
            line["cpu_baseline"]["gpu_over_cpu"] = head["updates_per_s"] / line["cpu_baseline"]["value"]
    dev.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
