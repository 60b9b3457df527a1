"""Benchmark: plain GPU radix join at full scale on one MI355X.

Workload (the reference's headline configuration for this path):
  R = S = 2^27 tuples of (int64 key, int64 payload); R's keys are a seeded
  permutation of 1..2^27 with payload = key; S's keys are uniform over R's
  keys with payload = row index; two-pass radix partitioning with 9 + 9 bits,
  then LDS hash build/probe; output {match count, SUM(R.value), SUM(S.value)}.

A step is one complete join (pass-1 histograms, pass-1 scatter, pass-2,
build/probe, aggregate read-back) over inputs already resident in HBM.
Throughput = (|R| + |S|) / step time, as the reference defines it
(guides/benchmarking.md:45-56).  Phase and kernel times come from HIP
events recorded inside the timed steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--log2n 27]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "fast-interconnects_amd"))

METRIC = "joined tuples/s + % HBM roofline; batched/served vs plain launch; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
SEED = 42


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--bits1", type=int, default=9)
    ap.add_argument("--bits2", type=int, default=9)
    ap.add_argument("--chunks", type=int, default=0, help="pass-1 chunks (0 = automatic)")
    ap.add_argument("--block", type=int, default=0, help="pass-1 threads per block (0 = automatic)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-log2n", type=int, default=0,
                    help="CPU baseline size (log2 of |R| = |S|; 0 = the GPU workload's size)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0 = the host share)")
    ap.add_argument("--workload", choices=["plain", "skew", "batched", "served", "ooc", "np", "operator"],
                    default="plain",
                    help="plain: R = S = 2^27 int64, 9+9 bits (the headline line); skew: R = 2^27 int32 unique, "
                         "S = 2^30 int32 Zipf over R, 8+8 bits; batched: 4096 int32 joins of log-uniform sizes")
    ap.add_argument("--zipf", type=float, default=1.0, help="skew: Zipf exponent")
    ap.add_argument("--skew-log2s", type=int, default=30, help="skew: log2 |S|")
    ap.add_argument("--joins", type=int, default=4096, help="batched: number of joins")
    ap.add_argument("--wave-gib", type=float, default=24.0, help="batched: input bytes per submission wave")
    ap.add_argument("--builds", type=int, default=256, help="served: resident build sides")
    ap.add_argument("--probe-inputs", type=int, default=48, help="served: distinct probe sides in device memory")
    ap.add_argument("--requests", type=int, default=2000, help="served: requests in the Poisson stream")
    ap.add_argument("--load", type=float, default=0.8, help="served: offered load as a fraction of capacity")
    ap.add_argument("--hbm-gb", type=float, default=200.0, help="served: HBM budget (GB)")
    ap.add_argument("--table-build-max", type=int, default=1 << 18,
                    help="served: builds of at most this many tuples kept as resident hash tables (0 = none)")
    ap.add_argument("--cache-gb", type=float, default=2.0, help="ooc: device cache for pass-1 partitions (GB)")
    ap.add_argument("--pass1-cpu", action="store_true", help="ooc: pass 1 on host threads")
    ap.add_argument("--probe-slices", type=int, default=0,
                    help="ooc with --pass1-cpu: probe-side slices partitioned while the GPU joins the previous "
                         "one (0 = automatic)")
    ap.add_argument("--lanes", type=int, default=4, help="batched: streams the size classes are spread over")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched: skip the per-join comparators (a profile of the batched form alone)")
    ap.add_argument("--plain-sample", type=int, default=0,
                    help="batched: time the plain per-join launch on every k-th join only (0 = all)")
    return ap.parse_args()


def log(msg):
    """Progress of a long workload (stderr, so the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# Host threads a job on one GPU of the box may use: the box gives each GPU a
# share of 16 cores, whatever nproc reports for the whole machine.
HOST_CPU_SHARE = 16


def host_cpu():
    """(cores this process may run on, CPU model name)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model


def cpu_baseline(log2n, bits1, bits2, threads):
    """The oracle's multi-threaded radix join (C restatement of the reference
    CPU partitioner + per-partition hash join) on the same workload, built by
    the same seeded generators as the device inputs.  Returns the baseline
    line and the oracle's result for these inputs."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as orc

    n = 1 << log2n
    rk = orc.gen_perm(n, SEED, np.int64, 1)
    rp = rk.copy()
    sk = orc.gen_uniform(n, SEED, 1, 1, n, np.int64)
    sp = orc.gen_iota(n, 0, np.int64)
    log(f"cpu baseline: inputs of 2 x 2^{log2n} tuples generated, {threads} threads")
    orc.radix_join(rk, rp, sk, sp, bits1, bits2, threads)  # warm-up
    times = []
    t_end = time.time() + 20.0
    while len(times) < 3 or (time.time() < t_end and len(times) < 10):
        t0 = time.perf_counter()
        res = orc.radix_join(rk, rp, sk, sp, bits1, bits2, threads)
        times.append(time.perf_counter() - t0)
    log(f"cpu baseline: {len(times)} runs, median {np.median(times):.3f} s")
    med = float(np.median(times))
    nproc, model = host_cpu()
    return {
        "value": 2 * n / med,
        "unit": "tuples/s",
        "cores": threads,
        "kind": "port",
        "host_nproc": nproc,
        "cpu_model": model,
        "sample": f"R = S = 2^{log2n} int64 (the same seeded generators as the GPU inputs), {bits1}+{bits2} "
                  f"bits, {threads} threads (one per core of the GPU's host share), 1 warm-up, median of "
                  f"{len(times)} runs",
    }, tuple(res)


def measured_traffic(kernel, log2n, bits1, bits2):
    """HBM bytes per launch of `kernel` at this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counter values scaled by what
    the same counters read for the plain copy kernel); None when the summary
    does not cover this workload or kernel."""
    path = os.path.join(REPO, "profiles", "radix_join_2e27_i64_9+9_hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    w = t.get("workload", {})
    if (w.get("log2n"), w.get("bits1"), w.get("bits2")) != (log2n, bits1, bits2):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k["fetch_bytes"] + k["write_bytes"]


def stream_copy_gbs(fij, nbytes, reps=5):
    """HBM bandwidth of the library's plain copy kernel (bytes read + written
    per second): what a streaming kernel reaches on this GPU."""
    src = fij.DeviceArray(nbytes // 8, np.uint64)
    dst = fij.DeviceArray(nbytes // 8, np.uint64)
    src.fill(1)
    fij.copy_kernel(dst, src)
    e0, e1 = fij.Event(), fij.Event()
    e0.record()
    for _ in range(reps):
        fij.copy_kernel(dst, src)
    e1.record()
    ms = e1.elapsed_ms(e0) / reps
    return 2 * nbytes / (ms / 1000.0) / 1e9


def access_pattern_ceilings(fij, n, bits1, bits2, chunks, scatter_ms, join_ms, reps=5):
    """Times of the scatter's and the join's memory patterns alone (the
    library's calibration kernels, HIP events, median of `reps`), beside the
    kernels' own times: how close each kernel is to what its access pattern
    costs on this GPU.  The line-scatter copy moves the scatter's bytes (n
    16-byte words read, n written as whole 128-byte lines to 2^bits1
    partitions x `chunks` chunk pieces); the run gathers read 2n words as the
    join's items (2^(bits1+bits2) per relation) in runs of one pass-2 tile
    each, 128-byte aligned or shifted by one 16-byte word."""
    def timed(fn):
        e0, e1 = fij.Event(), fij.Event()
        ms = []
        fn()
        for _ in range(reps):
            e0.record()
            fn()
            e1.record()
            ms.append(e1.elapsed_ms(e0))
        return float(np.median(ms))

    src = fij.DeviceArray(4 * n, np.int64)  # 2n 16-byte words
    dst = fij.DeviceArray(2 * n, np.int64)
    src.fill(3)
    f1, f2 = 1 << bits1, 1 << bits2
    line_ms = timed(lambda: fij.calib_line_scatter(dst, src, f1, chunks, nbytes=16 * n))
    del dst
    ck = fij.DeviceArray.zeros(1, np.uint32)
    part = n // f1
    run = 8192 // f2  # the plan's pass-2 tiles hold 8192 tuples: one run per digit and tile
    gather = {}
    for name, r, shift in ((f"runs{run}_aligned", run, 0), (f"runs{run}_shifted", run, 1),
                           (f"runs{2 * run}_shifted", 2 * run, 1)):
        gather[name] = timed(lambda: fij.calib_run_gather(src, part, f2, r, shift, ck))
    return {
        "line_scatter_copy_ms": line_ms,
        "scatter_over_line_scatter_copy": scatter_ms / line_ms,
        "run_gather_ms": gather,
        "join_over_shifted_run_gather": join_ms / gather[f"runs{run}_shifted"],
    }


def skew_main(a):
    """Skewed join: R = 2^27 int32 tuples with unique permuted keys 1..2^27,
    S = 2^skew_log2s int32 tuples with Zipf(--zipf) keys over R's keys
    (payload = row index), 8 + 8 radix bits.  Heavy (pass-1, pass-2)
    partitions are split into work units joined by many waves."""
    import fij

    n_r, n_s = 1 << 27, 1 << a.skew_log2s
    kt = fij.KeyType.I32
    rk = fij.DeviceArray(n_r, np.int32)
    rp = fij.DeviceArray(n_r, np.int32)
    sk = fij.DeviceArray(n_s, np.int32)
    sp = fij.DeviceArray(n_s, np.int32)
    fij.gen_permutation(rk, SEED, 1)
    fij.gen_iota(rp, 1)
    fij.gen_zipf(sk, n_r, a.zipf, SEED, 1, SEED, 1)
    fij.gen_iota(sp, 0)
    fij.synchronize()
    plan = fij.JoinPlan(kt, n_r, n_s, 8, 8)
    for _ in range(a.warmup):
        res, _ = plan.execute(rk, rp, sk, sp)
    times, phases = [], []
    fij.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res, t = plan.execute(rk, rp, sk, sp)
        phases.append(t)
    wall = time.perf_counter() - t0
    ms = wall * 1000.0 / a.steps
    assert res[0] == n_s, f"match count {res[0]} != {n_s}"  # every S key is one of R's keys
    # closed forms: R.value = index + 1 of the key's row; S.value = row index
    assert res[2] == (n_s * (n_s - 1) // 2) % (1 << 64)
    gpu_ms = float(np.median([t["total_ms"] for t in phases]))
    b_alg = (n_r + n_s) * (4 + 5 * 8)
    out = {
        "metric": METRIC, "value": (n_r + n_s) / (ms / 1000.0), "unit": "tuples/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded on-device generators, Zipf probe keys)",
        "config": {"workload": f"skewed radix join, R = 2^27 int32 unique keys, S = 2^{a.skew_log2s} int32 "
                               f"Zipf({a.zipf}) over R's keys, 8+8 radix bits, count + sums",
                   "build_tuples": n_r, "probe_tuples": n_s, "radix_bits": [8, 8], "zipf": a.zipf,
                   "parallelism": "single GPU"},
        "gpu_ms_per_join": gpu_ms,
        "phases_ms": {k: float(np.median([t[k] for t in phases])) for k in phases[0]},
        "pipeline_roofline_frac": b_alg / (gpu_ms / 1000.0) / (HBM_PEAK_GBS * 1e9),
        "result": {"count": res[0], "sum_build": res[1], "sum_probe": res[2]},
    }
    print(json.dumps(out))


def batched_main(a):
    """Batched joins: a.joins independent int32 joins; |R| log-uniform in
    2^10..2^22 with unique permuted keys, |S| = |R| x U{4..16} with keys
    uniform over that join's R (payloads: row indices).  The stream is
    generated and joined in waves of at most --wave-gib of input.  Measured:
    the batched form (one submission per wave) against the plain per-join
    launch: one join at a time through a plan of its size class, plans
    allocated beforehand and reused, so that no allocation is timed (the
    reference's throughput excludes allocation, guides/benchmarking.md:45-56).
    Reported beside it: the same launch with a plan created and destroyed per
    join (one run of the reference's radix-join binary per pair), sampled on
    every --plain-sample-th join."""
    import fij

    rng = np.random.default_rng(SEED)
    n_r = np.floor(2.0 ** rng.uniform(10, 22, a.joins)).astype(np.int64)
    n_s = n_r * rng.integers(4, 17, a.joins)
    total = int((n_r + n_s).sum())
    per_join_bytes = (n_r + n_s) * 8
    waves, cur, acc = [], [], 0
    for i in range(a.joins):
        if cur and acc + per_join_bytes[i] > a.wave_gib * 2**30:
            waves.append(cur)
            cur, acc = [], 0
        cur.append(i)
        acc += per_join_bytes[i]
    waves.append(cur)
    batch = fij.JoinBatch(np.int32, max(len(w) for w in waves), use_graph=True, lanes=a.lanes)
    res_dev = fij.DeviceArray(3, np.uint64)
    # per-join algorithmic bytes: a small join (one block, LDS table) reads
    # each tuple's key and payload once (t); a partitioned join with P passes
    # moves k + (2P + 1) t per tuple (key histogram, P scatters, join read)
    classes = [fij.autotune(np.int32, int(r), int(s_)) for r, s_ in zip(n_r, n_s)]
    b_alg = sum((int(r) + int(s_)) * (8 if small else 4 + (2 * (2 if b2 else 1) + 1) * 8)
                for (small, _, b2), r, s_ in zip(classes, n_r, n_s))
    # the plain per-join launch: each join through a plan of the bits a
    # join run alone takes (fij_join_plan_bits: one pass where the build
    # side's partitions fit block tables, else its size class's two passes),
    # one plan per bits, sized for its largest join, allocated before any
    # timing
    plain_bits = [fij.plan_bits(np.int32, int(r), int(s_)) for r, s_ in zip(n_r, n_s)]
    pooled = {}
    for (b1, b2), r, s_ in zip(plain_bits, n_r, n_s):
        cap = pooled.get((b1, b2), (0, 0))
        pooled[(b1, b2)] = (max(cap[0], int(r)), max(cap[1], int(s_)))
    pooled = {key: fij.JoinPlan(fij.KeyType.I32, r, s_, *key) for key, (r, s_) in pooled.items()}
    plain_step = a.plain_sample or 16
    t_alloc = 0.0
    n_alloc = 0
    t_batch = t_reused = t_cold = 0.0
    gpu_batch = 0.0
    stats_acc = {"small_joins": 0, "partitioned_joins": 0, "prepare_ms": 0.0, "grow_ms": 0.0, "record_ms": 0.0}
    checked = 0
    for w in waves:
        # inputs of the wave in one arena
        lens = [(int(n_r[i]), int(n_s[i])) for i in w]
        words = sum(2 * (r + s) for r, s in lens)
        arena = fij.DeviceArray(words, np.int32)
        joins, off = [], 0
        for i, (r, s_) in zip(w, lens):
            rk = arena.view(off, off + r); off += r
            rp = arena.view(off, off + r); off += r
            sk = arena.view(off, off + s_); off += s_
            sp = arena.view(off, off + s_); off += s_
            fij.gen_permutation(rk, SEED + i, 1)
            fij.gen_iota(rp, 1)
            fij.gen_uniform(sk, SEED + i, 1, 1, r)
            fij.gen_iota(sp, 0)
            joins.append((rk, rp, sk, sp))
        fij.synchronize()
        # the first submission of a wave grows the size classes' buffers
        # (device allocation); allocation is not part of the reference's
        # throughput, so the wave is timed on its second submission and the
        # first one is reported beside it
        t0 = time.perf_counter()
        cold, _ = batch.execute(joins)
        t_cold += time.perf_counter() - t0
        t0 = time.perf_counter()
        got, st = batch.execute(joins)
        t_batch += time.perf_counter() - t0
        gpu_batch += st["gpu_ms"] / 1000.0
        for k in stats_acc:
            stats_acc[k] += st[k]
        assert cold == got, "first and second submission of a wave differ"
        for (r, s_), g in zip(lens, got):
            assert g[0] == s_ and g[2] == s_ * (s_ - 1) // 2, "batched join result"
        # the plain per-join launch, one join at a time, every join
        for j in range(0 if a.no_plain else len(joins)):
            b1, b2 = plain_bits[w[j]]
            fij.synchronize()
            t0 = time.perf_counter()
            pooled[(b1, b2)].enqueue_sized(*joins[j], res_dev)
            fij.synchronize()
            t_reused += time.perf_counter() - t0
            g = tuple(int(x) for x in res_dev.to_host())
            assert g == got[j], "plain and batched results differ"
            checked += 1
        # with a plan allocated per join, on a sample
        for j in range(0, 0 if a.no_plain else len(joins), plain_step):
            r, s_ = lens[j]
            b1, b2 = plain_bits[w[j]]
            t0 = time.perf_counter()
            plan = fij.JoinPlan(fij.KeyType.I32, r, s_, b1, b2)
            g, _ = plan.execute(*joins[j])
            del plan
            fij.synchronize()
            t_alloc += time.perf_counter() - t0
            n_alloc += 1
            assert g == got[j], "allocating plain launch and batched results differ"
        del joins, arena
        log(f"wave of {len(w)} joins done ({checked} checked against the plain launch)")
    value = total / t_batch
    out = {
        "metric": METRIC, "value": value, "unit": "tuples/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": t_batch * 1000.0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic (seeded on-device generators)",
        "config": {"workload": f"batched joins: {a.joins} int32 joins, |R| log-uniform 2^10..2^22, "
                               "|S| = |R| x U{4..16}, count + sums per join", "joins": a.joins,
                   "tuples": total, "waves": len(waves), "lanes": a.lanes, "parallelism": "single GPU"},
        "batched": {"seconds": t_batch, "gpu_seconds": gpu_batch, "tuples_per_s": value, **stats_acc,
                    "first_submission_seconds": t_cold,
                    "note": "each wave timed on its second submission; the first also grows the size classes' "
                            "buffers (device allocation, excluded as in the reference's throughput)"},
        "algorithmic_bytes": b_alg,
        "pipeline_roofline_frac": b_alg / gpu_batch / (HBM_PEAK_GBS * 1e9),
        "wall_roofline_frac": b_alg / t_batch / (HBM_PEAK_GBS * 1e9),
        "plain_per_join": {"seconds": t_reused, "tuples_per_s": total / t_reused if t_reused else None,
                           "note": "one join at a time through a preallocated plan of its size class"},
        "speedup_vs_plain": t_reused / t_batch if t_reused else None,
        "plain_per_join_allocating": {
            "seconds_per_join": t_alloc / max(n_alloc, 1), "joins_timed": n_alloc,
            "note": "a plan created, executed and destroyed per join (allocation timed), every "
                    f"{plain_step}th join"},
    }
    print(json.dumps(out))


def latency_by_slot_class(lat, probe_len, slot_classes, alone_s):
    """Latency percentiles of the requests grouped by the smallest slot class
    that holds their probe side (the server may run one in a larger free
    slot), beside the median time of the same requests joined alone by the
    plain launch: what a request waits for beyond its own work."""
    out = {}
    lo = 0
    for _, cap in slot_classes:
        sel = (probe_len > lo) & (probe_len <= cap)
        if sel.any():
            out[f"probe <= {cap}"] = {"requests": int(sel.sum()), "p50_ms": float(np.percentile(lat[sel], 50)),
                                      "p99_ms": float(np.percentile(lat[sel], 99)),
                                      "alone_p50_ms": float(np.median(alone_s[sel]) * 1000.0)}
        lo = cap
    return out


def served_main(a):
    """Served joins: a Poisson stream of requests against a pool of a.builds
    int64 build sides (|R| log-uniform 2^12..2^26, keys a permutation of
    1..|R|, payload = key), kept resident: builds of at most
    a.table_build_max tuples as open-addressing hash tables, larger ones
    partitioned once at registration.  Each request
    brings one of a.probe_inputs probe sides (|S| log-uniform 2^14..2^28,
    keys uniform over 1..1.25|R|: 80% join, 20% miss; payload = row index)
    held in device memory.  Slots and resident builds are charged against
    the HBM budget.  Measured: capacity (all requests queued at once), then
    latency percentiles and throughput at a.load x capacity, and the plain
    per-join launch over the same stream (both sides partitioned and joined,
    one request at a time, through preallocated plans; beside it with a plan
    allocated per request).  Every result equals the same request joined
    alone."""
    import threading

    import fij

    rng = np.random.default_rng(SEED)
    n_b = np.floor(2.0 ** rng.uniform(12, 26, a.builds)).astype(np.int64)
    n_p = np.floor(2.0 ** rng.uniform(14, 28, a.probe_inputs)).astype(np.int64)
    which = rng.integers(0, a.builds, a.probe_inputs)
    slot_classes = [(8, 1 << 20), (4, 1 << 24), (2, 1 << 28)]
    srv = fij.JoinServer(np.int64, slot_classes, max_probe_len=1 << 28, hbm_budget=int(a.hbm_gb * 1e9),
                         table_build_max=a.table_build_max)
    scratch_k = fij.DeviceArray(int(n_b.max()), np.int64)
    scratch_p = fij.DeviceArray(int(n_b.max()), np.int64)
    t0 = time.perf_counter()
    ids = []
    for i, n in enumerate(n_b):
        k, p = scratch_k.view(0, int(n)), scratch_p.view(0, int(n))
        fij.gen_permutation(k, SEED + i, 1)
        fij.check(fij.lib.fij_memcpy(fij.C.c_void_p(p.ptr), fij.C.c_void_p(k.ptr), k.nbytes(), None))
        ids.append(srv.add_build(k, p))
    t_register = time.perf_counter() - t0
    log(f"{a.builds} build sides registered in {t_register:.2f} s")
    del scratch_k, scratch_p
    probes = []
    for j in range(a.probe_inputs):
        b = int(which[j])
        sk = fij.DeviceArray(int(n_p[j]), np.int64)
        sp = fij.DeviceArray(int(n_p[j]), np.int64)
        fij.gen_uniform(sk, SEED + 7 * j, 3, 1, int(n_b[b]) * 5 // 4)
        fij.gen_iota(sp, 0)
        probes.append((b, sk, sp))
    fij.synchronize()
    used, _ = srv.memory()
    # each probe input joined alone, and by the plain per-join launch: both
    # sides partitioned and joined through a plan of the pair's size class,
    # allocated before timing (no allocation timed); beside it, with a plan
    # created and destroyed per join
    alone, t_plain, t_alloc = [], [], []
    # the plain launch's bits for a pair run alone (as the server chooses
    # them: one pass where the build side's partitions fit block tables)
    classes = [(False,) + fij.plan_bits(np.int64, int(n_b[b]), int(n_p[j])) for j, (b, _, _) in enumerate(probes)]
    caps = {}
    for (_, b1, b2), (b, _, _), j in zip(classes, probes, range(len(probes))):
        c = caps.get((b1, b2), (0, 0))
        caps[(b1, b2)] = (max(c[0], int(n_b[b])), max(c[1], int(n_p[j])))
    pooled = {key: fij.JoinPlan(fij.KeyType.I64, r, s_, *key) for key, (r, s_) in caps.items()}
    res_dev = fij.DeviceArray(3, np.uint64)
    bk = fij.DeviceArray(int(n_b.max()), np.int64)
    for j, (b, sk, sp) in enumerate(probes):
        alone.append(srv.wait(srv.submit(ids[b], sk, sp))[0])
        n = int(n_b[b])
        k, p = bk.view(0, n), bk.view(0, n)  # payload = key
        fij.gen_permutation(k, SEED + b, 1)
        fij.synchronize()
        _, b1, b2 = classes[j]
        t0 = time.perf_counter()
        pooled[(b1, b2)].enqueue_sized(k, p, sk, sp, res_dev)
        fij.synchronize()
        t_plain.append(time.perf_counter() - t0)
        g = tuple(int(x) for x in res_dev.to_host())
        assert g == alone[j], f"probe input {j}: served {alone[j]} != plain {g}"
        t0 = time.perf_counter()
        plan = fij.JoinPlan(fij.KeyType.I64, n, int(n_p[j]), b1, b2)
        g, _ = plan.execute(k, p, sk, sp)
        del plan
        fij.synchronize()
        t_alloc.append(time.perf_counter() - t0)
        assert g == alone[j], f"probe input {j}: served {alone[j]} != allocating plain {g}"
    del bk, pooled
    log(f"{a.probe_inputs} probe inputs joined alone and by the plain launch")
    stream = rng.integers(0, a.probe_inputs, a.requests)
    tuples = np.array([int(n_b[probes[j][0]]) + int(n_p[j]) for j in range(a.probe_inputs)], dtype=np.float64)

    def run(arrivals):
        rids = [None] * len(stream)
        lat = np.zeros(len(stream))
        done = threading.Event()
        errors = []

        def waiter():
            try:
                for i in range(len(stream)):
                    while rids[i] is None:
                        time.sleep(20e-6)
                    got, t = srv.wait(rids[i])
                    lat[i] = t["latency_ms"]
                    if got != alone[stream[i]]:
                        errors.append(i)
            finally:
                done.set()

        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        for i, j in enumerate(stream):
            if arrivals is not None:
                dt = arrivals[i] - (time.perf_counter() - start)
                if dt > 300e-6:
                    time.sleep(dt - 200e-6)
                while time.perf_counter() - start < arrivals[i]:
                    pass
            rids[i] = srv.submit(ids[probes[j][0]], probes[j][1], probes[j][2])
        th.join()
        wall = time.perf_counter() - start
        assert not errors, f"requests {errors[:5]} differ from the same request alone"
        return wall, lat

    run(None)  # warm-up
    wall_sat, lat_sat = run(None)
    capacity = len(stream) / wall_sat
    log(f"saturated: {len(stream)} requests in {wall_sat:.2f} s")
    rate = a.load * capacity
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(stream)))
    wall, lat = run(arrivals)
    total = float(tuples[stream].sum())
    # per request, by how its build side is resident: a hash table -- the
    # probe side read once and one table entry per probe tuple (2t); one pass
    # -- the probe side's histogram, scatter and join read (k + 3t) plus the
    # build side's read (t); two passes -- k + 5t per probe tuple plus t per
    # build tuple
    def request_bytes(j):
        nb, np_ = int(n_b[probes[j][0]]), int(n_p[j])
        if a.table_build_max and nb <= a.table_build_max:
            return np_ * 32
        if fij.plan_bits(np.int64, nb, 1 << 28)[1] == 0:
            return np_ * 56 + nb * 16
        return np_ * 88 + nb * 16
    b_alg = float(sum(request_bytes(j) for j in stream))
    plain_s = float(np.array(t_plain)[stream].sum())
    alloc_s = float(np.array(t_alloc)[stream].sum())
    value = total / wall
    out = {
        "metric": METRIC, "value": value, "unit": "tuples/s", "n_gpus": 1, "steps": len(stream), "warmup": 1,
        "ms_per_step": wall * 1000.0 / len(stream), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64", "data": "synthetic (seeded on-device generators)",
        "config": {"workload": f"served joins: Poisson stream of {len(stream)} requests at {a.load} x capacity, "
                               f"{a.builds} resident int64 builds |R| log-uniform 2^12..2^26, probes |S| "
                               "log-uniform 2^14..2^28 (80% matching keys), count + sums per request",
                   "builds": a.builds, "probe_inputs": a.probe_inputs, "requests": len(stream),
                   "slot_classes": slot_classes, "hbm_budget_gb": a.hbm_gb,
                   "table_build_max": a.table_build_max, "parallelism": "single GPU"},
        "tuples_note": "value counts |R| + |S| per request (the joined relations); the resident |R| is "
                       "partitioned once at registration",
        "probe_tuples_per_s": float(n_p[stream].sum()) / wall,
        "pipeline_roofline_frac": b_alg / wall / (HBM_PEAK_GBS * 1e9),
        "offered_rate_per_s": rate, "capacity_per_s": capacity,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "latency_by_slot_class": latency_by_slot_class(lat, n_p[stream], slot_classes,
                                                       np.array(t_plain)[stream]),
        "saturated": {"tuples_per_s": total / wall_sat, "p50_ms": float(np.percentile(lat_sat, 50)),
                      "p99_ms": float(np.percentile(lat_sat, 99))},
        "plain_per_join": {"seconds": plain_s, "tuples_per_s": total / plain_s,
                           "note": "each request's pair joined alone, both sides partitioned, through a "
                                   "preallocated plan of its size class"},
        "speedup_vs_plain": (total / wall) / (total / plain_s),
        "speedup_vs_plain_saturated": plain_s / wall_sat,
        "plain_per_join_allocating": {"seconds": alloc_s, "tuples_per_s": total / alloc_s,
                                      "note": "a plan created, executed and destroyed per request"},
        "registration_s": t_register, "server_device_bytes": used,
    }
    print(json.dumps(out))


def ooc_main(a):
    """Out-of-core join: R = S = 2^log2n int64 tuples in pinned host memory
    (the plain join's generators), pass-1 partitions cached in device memory up to
    --cache-gb and spilled to pinned host memory beyond it, pass 2 and the
    join streamed over partition groups.  Pass 1 on the GPU (reading the
    relations over the host link) or, with --pass1-cpu, on host threads:
    then R's cached prefix is copied while the probe side is partitioned in
    slices, each slice joined on the GPU while the next is partitioned."""
    import fij

    n = 1 << a.log2n
    cols = []
    tmp = fij.DeviceArray(n, np.int64)
    for gen in ("rk", "rp", "sk", "sp"):
        if gen in ("rk", "rp"):
            fij.gen_permutation(tmp, SEED, 1)  # payload = key
        elif gen == "sk":
            fij.gen_uniform(tmp, SEED, 1, 1, n)
        else:
            fij.gen_iota(tmp, 0)
        h = fij.HostArray(n, np.int64)
        fij.check(fij.lib.fij_memcpy(fij.C.c_void_p(h.ptr), fij.C.c_void_p(tmp.ptr), tmp.nbytes(), None))
        cols.append(h)
    del tmp
    fij.synchronize()
    cache = int(a.cache_gb * 1e9)
    threads = a.cpu_threads or min(16, os.cpu_count() or 1)  # the box's CPU share
    kw = dict(pass1_cpu=a.pass1_cpu, cache_bytes=cache, staging_tuples=1 << 24, cpu_threads=threads,
              probe_slices=a.probe_slices)
    for _ in range(a.warmup):
        fij.ooc_join(np.int64, *cols, a.bits1, a.bits2, **kw)
    times, walls = [], []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        res, t = fij.ooc_join(np.int64, *cols, a.bits1, a.bits2, **kw)
        walls.append(time.perf_counter() - t0)
        times.append(t)
        log(f"ooc join: {t['total_ms']:.1f} ms")
    assert res[0] == n and res[2] == (n * (n - 1) // 2) % (1 << 64), "ooc join result"
    total_ms = float(np.median([t["total_ms"] for t in times]))
    out = {
        "metric": METRIC, "value": 2 * n / (total_ms / 1000.0), "unit": "tuples/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": total_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64", "data": "synthetic (seeded on-device generators, copied to pinned host memory)",
        "config": {"workload": f"out-of-core radix join, R = S = 2^{a.log2n} int64 in pinned host memory, "
                               f"{a.bits1}+{a.bits2} bits, {a.cache_gb} GB device cache, pass 1 on "
                               f"{'host threads' if a.pass1_cpu else 'the GPU'}",
                   "cache_gb": a.cache_gb, "pass1": "cpu" if a.pass1_cpu else "gpu",
                   "cpu_threads": threads if a.pass1_cpu else 0, "probe_slices": a.probe_slices,
                   "parallelism": "single GPU"},
        "phases_ms": {k: float(np.median([t[k] for t in times])) for k in ("prefix_sum_ms", "partition_ms", "join_ms")},
        "cached_tuples": [times[0]["cached_build_tuples"], times[0]["cached_probe_tuples"]],
        "groups": times[0]["groups"],
        "wall_ms_incl_allocation": float(np.median(walls)) * 1000.0,
        "result": {"count": res[0], "sum_build": res[1], "sum_probe": res[2]},
    }
    print(json.dumps(out))


def np_main(a):
    """The no-partitioning join (the hashjoin binary's comparison point):
    R = S = 2^log2n int64 tuples with the plain join's generators, a global
    linear-probing table of 2 * next_pow2(|R|) entries, grid 2 x CUs of 256
    threads (hashjoin/src/main.rs:383-390); and the same probe side against
    a build side of 2^22 tuples, whose table (128 MB) fits the GPU's last
    level cache."""
    import fij

    n = 1 << a.log2n
    grid = 2 * fij.device_limits()["compute_units"]
    sk = fij.DeviceArray(n, np.int64)
    sp = fij.DeviceArray(n, np.int64)
    fij.gen_iota(sp, 0)
    lines = {}
    for log2r in (a.log2n, 22):
        nr = 1 << log2r
        rk = fij.DeviceArray(nr, np.int64)
        fij.gen_permutation(rk, SEED, 1)
        fij.gen_uniform(sk, SEED, 1, 1, nr)
        fij.synchronize()
        entries = 2 * nr
        ht = fij.HashTable(entries, fij.KeyType.I64)
        hj = fij.HipHashJoin(fij.HashingScheme.LinearProbing, ht, (grid, 256), (grid, 256))
        agg = fij.aggregates_array(grid)
        e = [fij.Event() for _ in range(3)]
        tb, tp = [], []
        for step in range(a.warmup + a.steps):
            fij.check(fij.lib.fij_np_table_init(int(fij.KeyType.I64), fij._ptr(ht.mem), entries, None))
            agg.fill(0)
            e[0].record()
            hj.build(rk, rk)  # payload = key
            e[1].record()
            hj.probe_sum(sk, sp, None, agg)
            e[2].record()
            fij.synchronize()
            if step >= a.warmup:
                tb.append(e[1].elapsed_ms(e[0]))
                tp.append(e[2].elapsed_ms(e[1]))
        res = fij.sum_aggregates(agg)
        assert res[0] == n and res[2] == (n * (n - 1) // 2) % (1 << 64), "NP join result"
        b, p_ = float(np.median(tb)), float(np.median(tp))
        lines[f"R=2^{log2r}"] = {"build_ms": b, "probe_ms": p_, "tuples_per_s": (nr + n) / ((b + p_) / 1000.0),
                                  "table_bytes": entries * 16}
        del rk, ht, hj
    first = lines[f"R=2^{a.log2n}"]
    out = {
        "metric": METRIC, "value": first["tuples_per_s"], "unit": "tuples/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": first["build_ms"] + first["probe_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic (seeded on-device generators)",
        "config": {"workload": f"no-partitioning linear-probing join, R = S = 2^{a.log2n} int64, grid {grid} x 256",
                   "parallelism": "single GPU"},
        "lines": lines,
    }
    print(json.dumps(out))


def operator_main(a):
    """The operator API the reference's radix-join binary drives, one op at a
    time on one stream, timed by HIP events: chunked and contiguous prefix
    sums, prefix sum with payload copy, the five scatter variants, task
    assignment, and the LDS perfect and bucket-chaining joins.  Shape: R =
    permutation of 1..2^20 (payload = key), S = 2^24 keys uniform over R's
    keys (payload = row index), int32, one pass of 8 bits; partitioner grid
    = one chunk per CU, join grid = two blocks per CU (the greedy task rule
    then gives every partition a task of its own; with one block per CU a
    partition a little below the average size pulls in the next one).  `value`: R and S
    tuples per second of contiguous prefix sums + scatters (SSWWCv2) +
    chaining join (the join takes single-chunk relations)."""
    import fij

    cus = fij.device_limits()["compute_units"]
    n_r, n_s, bits = 1 << 20, 1 << 24, 8
    dt, kt, k = np.int32, fij.KeyType.I32, 4
    t = 2 * k
    rk, rp = fij.DeviceArray(n_r, dt), fij.DeviceArray(n_r, dt)
    sk, sp = fij.DeviceArray(n_s, dt), fij.DeviceArray(n_s, dt)
    fij.gen_permutation(rk, SEED, 1)
    fij.check(fij.lib.fij_memcpy(fij.C.c_void_p(rp.ptr), fij.C.c_void_p(rk.ptr), rk.nbytes(), None))
    fij.gen_uniform(sk, SEED, 1, 1, n_r)
    fij.gen_iota(sp, 0)
    fij.synchronize()
    st = fij.Stream()
    e0, e1 = fij.Event(), fij.Event()
    CH, CO = fij.HistogramAlgorithm.Chunked, fij.HistogramAlgorithm.Contiguous
    FIRST = fij.RadixPass.First

    def timed(op, before=None):
        ts = []
        for i in range(a.warmup + a.steps):
            if before:
                before()
            e0.record(st)
            op()
            e1.record(st)
            st.synchronize()
            if i >= a.warmup:
                ts.append(e1.elapsed_ms(e0))
        return float(np.median(ts))

    def line(ms, nbytes):
        gbs = nbytes / (ms / 1000.0) / 1e9
        return {"ms": ms, "bytes": nbytes, "gbs": gbs, "frac": gbs / 8000.0}

    ops = {}
    for name, alg in (("prefix_sum_chunked", CH), ("prefix_sum_contiguous", CO)):
        part = fij.GpuRadixPartitioner(alg, fij.RadixPartitionAlgorithm.SSWWCv2, fij.RadixBits(bits), cus, 1024)
        offs = fij.PartitionOffsets(alg, cus, bits, kt)
        ops[name] = line(timed(lambda: part.prefix_sum(FIRST, sk, offs, stream=st)), n_s * k)
    part = fij.GpuRadixPartitioner(CO, fij.RadixPartitionAlgorithm.SSWWCv2, fij.RadixBits(bits), cus, 1024)
    offs = fij.PartitionOffsets(CO, cus, bits, kt)
    dk, dp = fij.DeviceArray(n_s, dt), fij.DeviceArray(n_s, dt)
    ops["prefix_sum_and_copy_with_payload"] = line(
        timed(lambda: part.prefix_sum_and_copy_with_payload(FIRST, sk, sp, dk, dp, offs, stream=st)), n_s * 4 * k)
    del dk, dp

    def partitioned(keys, pays, alg, block, hist=CH):
        """(partitioner, offsets, relation, timed scatter ms): every timed
        scatter follows its own untimed prefix sum (a scatter moves the
        offsets into the relation)."""
        pt = fij.GpuRadixPartitioner(hist, alg, fij.RadixBits(bits), cus, block)
        of = fij.PartitionOffsets(hist, cus, bits, kt)
        rel = fij.PartitionedRelation(len(keys), hist, bits, cus, kt)
        ms = timed(lambda: pt.partition(FIRST, keys, pays, of, rel, stream=st),
                   before=lambda: pt.prefix_sum(FIRST, keys, of, stream=st))
        return pt, of, rel, ms

    for alg in fij.RadixPartitionAlgorithm:
        block = 512 if alg == fij.RadixPartitionAlgorithm.HSSWWCv4 else 1024
        ms = partitioned(sk, sp, alg, block)[3]
        ops[f"partition_{alg.name}"] = line(ms, n_s * (2 * k + t))
    # high fanouts: line rings no longer fit LDS, the buffered variants sort
    # each tile by digit; the direct scatter (NC) beside them
    for hb in (10, 12, 14):
        for alg in (fij.RadixPartitionAlgorithm.SSWWCv2, fij.RadixPartitionAlgorithm.NC):
            pt = fij.GpuRadixPartitioner(CH, alg, fij.RadixBits(hb), cus, 1024)
            of = fij.PartitionOffsets(CH, cus, hb, kt)
            rel = fij.PartitionedRelation(n_s, CH, hb, cus, kt)
            ms = timed(lambda: pt.partition(FIRST, sk, sp, of, rel, stream=st),
                       before=lambda: pt.prefix_sum(FIRST, sk, of, stream=st))
            ops[f"partition_{alg.name}_{hb}bits"] = line(ms, n_s * (2 * k + t))
            del pt, of, rel
    # the join's inputs: contiguous layout (one range per partition), as
    # the join and its task assignment take single-chunk relations
    psum_r = timed(lambda: part.prefix_sum(FIRST, rk, fij.PartitionOffsets(CO, cus, bits, kt), stream=st))
    _, _, brel, part_r = partitioned(rk, rp, fij.RadixPartitionAlgorithm.SSWWCv2, 1024, CO)
    _, _, prel, part_s = partitioned(sk, sp, fij.RadixPartitionAlgorithm.SSWWCv2, 1024, CO)
    jgrid = 2 * cus
    tasks = fij.DeviceArray.zeros(jgrid + 1, np.uint32)
    ops["assign_tasks"] = line(timed(lambda: fij.assign_tasks(prel, bits, jgrid, tasks, stream=st)), 8 * (1 << bits))
    want_build = int(sk.to_host().astype(np.int64).sum())
    joins = {}
    for scheme in (fij.HashingScheme.Perfect, fij.HashingScheme.BucketChaining):
        j = fij.HipRadixJoin(FIRST, fij.RadixBits(bits), scheme, jgrid, 1024)
        agg = fij.aggregates_array(jgrid)
        ms = timed(lambda: j.join(brel, prel, None, tasks, agg, stream=st), before=lambda: agg.fill(0, stream=st))
        got = fij.sum_aggregates(agg)
        assert got == (n_s, want_build, n_s * (n_s - 1) // 2), f"{scheme.name} join result {got}"
        ops[f"join_{scheme.name}"] = line(ms, (n_r + n_s) * t)
        joins[scheme.name] = ms
    total = ops["prefix_sum_contiguous"]["ms"] + psum_r + part_r + part_s + joins["BucketChaining"]
    out = {
        "metric": METRIC, "value": (n_r + n_s) / (total / 1000.0), "unit": "tuples/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": total, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded on-device generators)",
        "config": {"workload": "operator API, one pass of 8 bits: R = permutation of 1..2^20 int32 (payload = key), "
                               "S = 2^24 uniform over R's keys; contiguous prefix sums + SSWWCv2 scatters + "
                               "chaining join",
                   "partitioner_grid": cus, "join_grid": jgrid, "parallelism": "single GPU"},
        "ops": ops,
        "note": "bytes per op: prefix sums N k, copy with payload N 4k, scatters N (2k + t), joins (R + S) t, "
                "assign_tasks 8 B per partition; frac = GB/s / 8000",
    }
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              "(the radix join runs on one MI355X, multi-GPU sharding is design only)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "skew":
        return skew_main(a)
    if a.workload == "batched":
        return batched_main(a)
    if a.workload == "served":
        return served_main(a)
    if a.workload == "ooc":
        return ooc_main(a)
    if a.workload == "np":
        return np_main(a)
    if a.workload == "operator":
        return operator_main(a)
    import fij

    n = 1 << a.log2n
    kt = fij.KeyType.I64
    rk = fij.DeviceArray(n, np.int64)
    rp = fij.DeviceArray(n, np.int64)
    sk = fij.DeviceArray(n, np.int64)
    sp = fij.DeviceArray(n, np.int64)
    fij.gen_permutation(rk, SEED, 1)
    fij.check(fij.lib.fij_memcpy(fij.C.c_void_p(rp.ptr), fij.C.c_void_p(rk.ptr), rk.nbytes(), None))
    fij.gen_uniform(sk, SEED, 1, 1, n)
    fij.gen_iota(sp, 0)
    fij.synchronize()

    plan = fij.JoinPlan(kt, n, n, a.bits1, a.bits2, pass1_chunks=a.chunks, pass1_block=a.block)
    for _ in range(a.warmup):
        res, _ = plan.execute(rk, rp, sk, sp)
    phases = {"prefix_sum_ms": [], "partition_ms": [], "join_ms": [], "total_ms": [], "pass2_ms": [],
              "build_probe_ms": [], "deferred_ms": []}
    fij.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res, t = plan.execute(rk, rp, sk, sp)
        for k in phases:
            phases[k].append(t[k])
    fij.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / a.steps
    gpu_ms = float(np.median(phases["total_ms"]))
    # closed forms: every S key matches exactly one R key, whose payload is
    # the key, so SUM(R.value) = SUM(S.key); S's payload is the row index
    sum_sk = int(sk.to_host().view(np.uint64).sum(dtype=np.uint64))
    want = (n, sum_sk, (n * (n - 1) // 2) % (1 << 64))
    assert tuple(res) == want, f"join result {tuple(res)} != closed form {want}"

    value = 2 * n / (ms_per_step / 1000.0)
    t_bytes = 16
    b_alg = 2 * n * (8 + (2 * 2 + 1) * t_bytes)  # (|R|+|S|) * (k + (2P+1) t), P = 2
    med = {k: float(np.median(v)) for k, v in phases.items()}
    # Per-kernel launch times from the HIP events around each phase, and the
    # algorithmic HBM bytes of one launch (each tuple's bytes read and written
    # once; relation-sized kernels run once per relation).
    # (hist includes the small offsets scan; sort2 the small plan kernel; the
    # deferred-item join moves no bytes of its own at this shape)
    kernels = {
        "hist_kernel": {"launches": 2, "ms": med["prefix_sum_ms"] / 2, "bytes": n * 8},
        "scatter_swwc_kernel": {"launches": 2, "ms": med["partition_ms"] / 2, "bytes": n * (8 + 8 + t_bytes)},
        "sort2_kernel": {"launches": 1, "ms": med["pass2_ms"], "bytes": 2 * n * 2 * t_bytes},
        "join_kernel": {"launches": 1, "ms": med["build_probe_ms"], "bytes": 2 * n * t_bytes},
    }
    for k in kernels.values():
        k["gbs"] = k["bytes"] / (k["ms"] / 1000.0) / 1e9
        k["frac"] = k["gbs"] / HBM_PEAK_GBS
    dominant = max(kernels, key=lambda k: kernels[k]["ms"] * kernels[k]["launches"])
    achieved = kernels[dominant]["gbs"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded on-device generators)",
        "config": {
            "workload": f"plain radix join, R = S = 2^{a.log2n} (int64 key, int64 payload), "
                        f"{a.bits1}+{a.bits2} radix bits, LDS hash build/probe, count + sums",
            "build_tuples": n,
            "probe_tuples": n,
            "radix_bits": [a.bits1, a.bits2],
            "parallelism": "single GPU",
        },
        "gpu_ms_per_join": gpu_ms,
        "phases_ms": med,
        "kernels": kernels,
        "pipeline_roofline_frac": b_alg / (gpu_ms / 1000.0) / (HBM_PEAK_GBS * 1e9),
        "stream_copy_gbs": stream_copy_gbs(fij, 2 << 30),
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dominant, a.log2n, a.bits1, a.bits2),
        },
        "result": {"count": res[0], "sum_build": res[1], "sum_probe": res[2]},
    }
    out["result_check"] = "count, SUM(R.value) and SUM(S.value) equal their closed forms"
    del rk, rp, sk, sp, plan
    out["pattern_ceilings"] = access_pattern_ceilings(
        fij, n, a.bits1, a.bits2, a.chunks or fij.device_limits()["compute_units"],
        kernels["scatter_swwc_kernel"]["ms"], kernels["join_kernel"]["ms"])
    if not a.no_cpu:
        threads = a.cpu_threads or min(HOST_CPU_SHARE, host_cpu()[0])
        cpu_log2n = a.cpu_log2n or a.log2n
        out["cpu_baseline"], cpu_res = cpu_baseline(cpu_log2n, a.bits1, a.bits2, threads)
        if cpu_log2n == a.log2n:
            # the oracle joined the same inputs: its result must be the GPU's
            assert cpu_res == tuple(res), f"oracle {cpu_res} != GPU {tuple(res)}"
            out["cpu_baseline"]["same_result_as_gpu"] = True
    print(json.dumps(out))


if __name__ == "__main__":
    main()
