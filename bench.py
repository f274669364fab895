#!/usr/bin/env python3
"""Benchmark: Gbases/s counted on one MI355X.

Headline workload (N=1): reads from a repetitive, low-complexity genome.  A
100 Mbp genome (seed 3, GC 45%) of which 35% is interspersed copies of 50
random 300 bp repeat families (5% divergence) and 10% tandem arrays (units of
2-12 bp, arrays of 200-2000 bp); 150 bp reads at 20x (2 Gbp, 13.3 M reads)
with 0.3% substitutions; counted with k=31, m=23, ci=1, n=4096 partitions.
This is the largest single-GPU configuration of the benchmark set.  The FASTQ
text is generated and parsed on the host once; the ACGT chunks and offsets are
uploaded to HBM before the timed region.  One step = one whole counting job on
the device (pack, superkmers, routing, counting, drain).

Beside it: the short-read workload (250 Mbp of 150 bp reads at 50x, k=31 m=21
ci=2 n=256) and a k = 127 workload (250 Mbp of 250 bp reads, m=21 ci=2)
measured the same way, the end-to-end CLI rate (FASTQ file to TSV file) of the
first two, the many-sample batched leg, the serving leg and the CPU baseline.

Prints one JSON line.  --gpus N > 1 is not supported: this engine runs on one
GPU.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gbases/s counted on 1 MI355X (% HBM roofline); batched jobs/s + p99 latency"
HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "repeat_reads": {
        "desc": ("repetitive genome reads: 100 Mbp genome seed 3 (GC 45%, 35% interspersed copies of 50 "
                 "300 bp families at 5% divergence, 10% tandem arrays), 150 bp reads at 20x (2 Gbp), "
                 "0.3% substitutions, k=31 m=23 ci=1 n=4096"),
        "spec": ("kind=repeat_fq genome=100000000 seed=3 gc=0.45 cov=20 rl=150 sub=0.003 "
                 "fam=50 flen=300 div=0.05 rep=0.35 tan=0.10"),
        "k": 31, "m": 23, "ci": 1, "n": 4096,
        # the batch parameters tg_autotune_many picks for this job shape
        "config": "kmers_per_bucket=16384,max_tiles=1024,load_factor=0.75",
    },
    "short_reads": {
        "desc": ("short-read FASTQ: 5 Mbp genome seed 2, 150 bp reads at 50x (1,666,667 reads, 250 Mbp), "
                 "0.5% substitutions, N 0.01%, k=31 m=21 ci=2 n=256"),
        "spec": "kind=reads_fq genome=5000000 seed=2 gc=0.5 reads=1666667 rl=150 sub=0.005 nrate=0.0001",
        "k": 31, "m": 21, "ci": 2, "n": 256,
        "config": "kmers_per_bucket=32768,max_tiles=1024,load_factor=0.9",
    },
}
WORKLOADS["long_k"] = {
    "desc": ("k = 127 on 250 bp reads: 5 Mbp genome seed 2, 1,000,000 reads (250 Mbp), 0.5% substitutions, "
             "N 0.01%, k=127 m=21 ci=2"),
    "spec": "kind=reads_fq genome=5000000 seed=2 gc=0.5 reads=1000000 rl=250 sub=0.005 nrate=0.0001",
    "k": 127, "m": 21, "ci": 2, "n": 16,
    "config": None,  # k > 64 does not run through the batched engine
}
HEADLINE = "repeat_reads"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(n_bases, n_sk, sum_len, n_unique, n_out, k):
    """Algorithmic bytes of one job (whole path) and of the counting kernel.

    Job: N (read ASCII chunks) + N/4 (write packed) + N/4 (read packed)
       + 2P (write + read partition records, P = sum over superkmers of
         2 + ceil(len/4)) + 2 U s_slot (insert + drain, s_slot = 8 ceil(k/32) + 4)
       + U_out (8 ceil(k/32) + 4) (compacted output).
    Counting kernel: P (read its records) + 2 U s_slot + U_out (8 ceil(k/32) + 4).
    """
    nw = (k + 31) // 32
    s_slot = 8 * nw + 4
    P = 2 * n_sk + (sum_len + 3 * n_sk) // 4
    job = n_bases + n_bases // 4 + n_bases // 4 + 2 * P + 2 * n_unique * s_slot + n_out * s_slot
    count = P + 2 * n_unique * s_slot + n_out * s_slot
    superkmers = n_bases // 4 + P
    return job, count, superkmers, P


def alg_bytes_wide(n_bases, n_unique, n_out, k):
    """Algorithmic bytes of a k > 64 job (no partition records: every k-mer
    goes from the packed stream into the table).

    Job: N (read ASCII) + N/4 (write packed) + 2 N/4 (superkmer pass and
       table insert read the packed stream) + 2 U s_slot (insert + drain)
       + U_out s_slot (output), s_slot = 8 ceil(k/32) + 4.
    Insert ("count"): N/4 + 2 U s_slot + U_out s_slot, as the counting kernel
    of k <= 64 with the packed stream in place of the records.
    """
    s_slot = 8 * ((k + 31) // 32) + 4
    tab = 2 * n_unique * s_slot + n_out * s_slot
    job = n_bases + 3 * (n_bases // 4) + tab
    return job, n_bases // 4 + tab, n_bases // 4


def host_threads():
    """Threads for host work: the CPUs this process may run on, at most 16 (a
    shared multi-GPU host grants one GPU's job a 16-CPU share, whatever the
    CPU count it reports)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return n, max(1, min(16, n))


def cpu_baseline(bases, off, wl, threads, nproc, frac=8):
    """The oracle's multithreaded two-phase count (oracle/cpu_count.c: the
    reference pipeline restated in C, superkmers routed to n partitions, then
    one table per partition), pinned to the oracle's brute-force count by
    tests/test_oracle.py, on the first 1/frac of the workload's reads.  The
    reference itself is not on this machine, so this is the fallback the
    baseline plan names."""
    import oracle as O
    nch = (len(off) - 1) // frac
    nb = int(off[nch])
    O.count_parallel(bases[:min(nb, 1 << 22)], off[:1 + int(np.searchsorted(off, min(nb, 1 << 22), "right")) - 1],
                     wl["k"], wl["m"], wl["n"], wl["ci"], 2**64 - 1, threads, keep=False)  # warm the pages
    t0 = time.perf_counter()
    _, _, info = O.count_parallel(bases[:nb], off[:nch + 1], wl["k"], wl["m"], wl["n"], wl["ci"], 2**64 - 1,
                                  threads, keep=False)
    wall = time.perf_counter() - t0
    return {"value": round(nb / wall / 1e9, 5), "unit": "Gbases/s", "cores": threads, "kind": "port",
            "nproc": nproc,
            "sample": f"first {nch} of {len(off) - 1} reads ({nb} bases) of the headline workload; "
                      f"oracle/cpu_count.c (two-phase restatement of the reference pipeline, "
                      f"n={wl['n']} partitions) on {threads} threads of {nproc} CPUs, wall {wall:.2f} s "
                      f"(phase1 {info['phase1_s']:.2f} s, phase2 {info['phase2_s']:.2f} s), "
                      f"{info['out']} k-mers reported"}


# ------------------------------------------------------- many samples ----
C4_SPEC = ("kind=samples_fq samples=2000 seed=4 index={i} median=300000 sigma=1.2 lo=10000 "
           "hi=8000000 cov=10 rl=150 sub=0.005")
C4_K, C4_M = 21, 15


def gen_samples(T, n, threads):
    """The first n samples of the many-sample workload, generated in parallel."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(T.synth_parsed, [C4_SPEC.format(i=i) for i in range(n)]))


def bench_batched(T, eng, n_samples, threads, use_graph=True):
    """Run A: the plain per-job sequence (tg_count_device) once per sample.
    Run B: all samples through tg_count_many_device (batched launch sequences,
    one captured hipGraph replayed per batch).  Inputs are resident in HBM
    before either timed region; results stay in HBM (A leaves each job's
    result in its buffers; B moves every job's result to its result store)."""
    t0 = time.perf_counter()
    parts = gen_samples(T, n_samples, threads)
    sizes = np.array([len(b) for b, _ in parts], dtype=np.uint64)
    nchs = np.array([len(o) - 1 for _, o in parts], dtype=np.uint64)
    job_base = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    job_chunk = np.concatenate([[0], np.cumsum(nchs)]).astype(np.uint64)
    bases = np.concatenate([b for b, _ in parts])
    off = np.concatenate([[0]] + [o[1:] + job_base[j] for j, (_, o) in enumerate(parts)]).astype(np.uint64)
    # per-job offsets relative to the job (what tg_count_device takes)
    roff = np.concatenate([o for _, o in parts]).astype(np.uint64)
    roff_start = np.concatenate([[0], np.cumsum(nchs + 1)]).astype(np.uint64)
    del parts
    N = int(job_base[-1])
    log(f"many samples: {n_samples} jobs, {N} bases, {int(job_chunk[-1])} chunks "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    d_b = eng.dev_upload(bases)
    d_o = eng.dev_upload(off)
    d_r = eng.dev_upload(roff)
    del bases
    base_ptr, off_ptr = d_b.value, d_r.value
    eng.set_batch_config(use_graph=1 if use_graph else 0)

    def run_a():
        st = []
        for j in range(n_samples):
            st.append(eng.count_device(base_ptr + int(job_base[j]), int(sizes[j]),
                                       off_ptr + 8 * int(roff_start[j]), int(nchs[j]), C4_K, C4_M, 16, 1))
        eng.sync()
        return st

    def run_b():
        st = eng.count_many_device(d_b, N, d_o, int(job_chunk[-1]), job_chunk, job_base, C4_K, C4_M, 1)
        eng.sync()
        return st

    # autotune the batched path on the first quarter of the jobs, then warm both
    # paths (buffer pools, graph capture)
    t0 = time.perf_counter()
    nt = max(1, n_samples // 4)
    tuned, tuned_ms = eng.autotune_many(d_b, int(job_base[nt]), d_o, int(job_chunk[nt]), job_chunk[:nt + 1],
                                        job_base[:nt + 1], C4_K, C4_M, 1)
    t_tune = time.perf_counter() - t0
    log(f"autotuned on {nt} jobs in {t_tune:.1f} s: {tuned}")
    big = int(np.argmax(sizes))
    eng.count_device(base_ptr + int(job_base[big]), int(sizes[big]), off_ptr + 8 * int(roff_start[big]),
                     int(nchs[big]), C4_K, C4_M, 16, 1)
    run_b()
    t0 = time.perf_counter()
    st_b = run_b()
    t_b = time.perf_counter() - t0
    mb = eng.last_many.as_dict()
    # run A: the plain per-job launch sequence (host-driven between its steps)
    eng.set_plain_sequence(True)
    eng.count_device(base_ptr + int(job_base[big]), int(sizes[big]), off_ptr + 8 * int(roff_start[big]),
                     int(nchs[big]), C4_K, C4_M, 16, 1)
    t0 = time.perf_counter()
    st_a = run_a()
    t_a = time.perf_counter() - t0
    eng.set_plain_sequence(False)
    # for reference: the batched engine called once per job (graph per job)
    run_a()
    t0 = time.perf_counter()
    st_g = run_a()
    t_g = time.perf_counter() - t0
    same = all(a.n_kmers == b["n_kmers"] and a.n_superkmers == b["n_superkmers"] and
               a.n_unique == b["n_unique"] and a.n_out == b["n_out"] for a, b in zip(st_a, st_b))
    same = same and all(a.n_out == g.n_out and a.n_unique == g.n_unique for a, g in zip(st_a, st_g))
    # full key/count equality on a spread of jobs
    checked = 0
    for j in sorted(set([0, n_samples - 1, big] + list(range(0, n_samples, max(1, n_samples // 8))))):
        eng.count_device(base_ptr + int(job_base[j]), int(sizes[j]), off_ptr + 8 * int(roff_start[j]),
                         int(nchs[j]), C4_K, C4_M, 16, 1)
        ak, ac = eng.result(C4_K)
        bk, bc = eng.many_result(j, C4_K)
        oa, ob = np.argsort(ak[:, 0], kind="stable"), np.argsort(bk[:, 0], kind="stable")
        same = same and np.array_equal(ak[oa], bk[ob]) and np.array_equal(ac[oa], bc[ob])
        checked += 1
    eng.dev_free(d_b)
    eng.dev_free(d_o)
    eng.dev_free(d_r)
    return {
        "workload": f"many samples: first {n_samples} of the 2000-sample family (seed 4, log-normal genome "
                    f"median 300 kbp sigma 1.2 clipped 10 kbp-8 Mbp, 150 bp reads at 10x, 0.5% subs), "
                    f"k={C4_K} m={C4_M}",
        "jobs": n_samples, "bases": N,
        "run_a_per_job": {"s": round(t_a, 4), "jobs_per_s": round(n_samples / t_a, 2),
                          "gbases_per_s": round(N / t_a / 1e9, 3),
                          "sequence": "plain launch sequence per job (tg_set_plain_sequence)"},
        "graph_per_job": {"s": round(t_g, 4), "jobs_per_s": round(n_samples / t_g, 2),
                          "gbases_per_s": round(N / t_g / 1e9, 3),
                          "sequence": "batched engine called once per job (one graph launch each)"},
        "run_b_batched": {"s": round(t_b, 4), "jobs_per_s": round(n_samples / t_b, 2),
                          "gbases_per_s": round(N / t_b / 1e9, 3), "batches": mb["n_batches"],
                          "graph_launches": mb["n_graph_launches"], "captures": mb["n_captures"]},
        "speedup_b_over_a": round(t_a / t_b, 3),
        "a_equals_b": bool(same), "jobs_fully_compared": checked,
        "batch_config": eng.batch_config().as_dict(),
        "autotune": {"jobs": nt, "s": round(t_tune, 2), "best_ms_on_tuning_jobs": round(tuned_ms, 3)},
    }


# ------------------------------------------------------------ serving ----
C5_SPEC = ("kind=samples_fq seed=5 index={i} median=1000000 sigma=1.0 lo=10000 hi=20000000 "
           "cov=15 rl=150 sub=0")
C5_K, C5_M, C5_CI = 63, 21, 2


def bench_serving(T, n_req, lambdas, threads, lanes=2):
    """Poisson arrivals (seed 5) of counting requests at each rate in `lambdas`
    (jobs/s); latency = submit to result in host memory, per request."""
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        reqs = list(ex.map(T.synth_parsed, [C5_SPEC.format(i=i) for i in range(n_req)]))
    reqs = [(T.pinned_copy(b), T.pinned_copy(o)) for b, o in reqs]
    sizes = np.array([len(b) for b, _ in reqs], dtype=np.float64)
    log(f"serving: {n_req} requests, {int(sizes.sum())} bases (mean {sizes.mean() / 1e6:.1f} M, "
        f"max {sizes.max() / 1e6:.1f} M), generated in {time.perf_counter() - t0:.1f} s")
    srv = T.Server(C5_K, C5_M, C5_CI, lanes=lanes, max_batch_jobs=64, max_batch_bases=1 << 29)

    lens = [np.diff(o.astype(np.int64)) for _, o in reqs]
    closed = [int(np.maximum(L - C5_K + 1, 0).sum()) for L in lens]
    bad = []

    def run(arrivals):
        ids = []
        t_start = time.perf_counter()
        for i, a in enumerate(arrivals):
            dt = t_start + a - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append((srv.submit(*reqs[i % n_req]), time.perf_counter() - t_start))
        lat, end, bj = [], 0.0, []
        for i, (rid, ts) in enumerate(ids):
            if not srv.wait(rid, 300.0):
                raise RuntimeError(f"request {i} not done in 300 s")
            _, _, st, tm = srv.result(rid, copy=False)
            # every result: its k-mer total against the closed form
            if st["n_kmers"] != closed[i % n_req]:
                bad.append(i % n_req)
            lat.append(tm["submit_to_done_s"])
            bj.append(tm["batch_jobs"])
            end = max(end, ts + tm["submit_to_done_s"])
            srv.release(rid)
        lat = np.array(lat) * 1e3
        n = len(arrivals)
        tot = float(sum(sizes[i % n_req] for i in range(n)))
        return {"p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "max_ms": round(float(lat.max()), 3), "mean_ms": round(float(lat.mean()), 3),
                "jobs_per_s": round(n / end, 2), "gbases_per_s": round(tot / end / 1e9, 3),
                "mean_batch_jobs": round(float(np.mean(bj)), 2), "requests": n}

    run(np.zeros(min(8, n_req)))  # warm lanes (buffers, graph capture, pinned pool)
    sweep = []
    for lam in lambdas:
        # about 6 s of arrivals at low rates, all requests from 20 jobs/s up
        n_at = min(n_req, max(24, int(lam * 6)))
        rng = np.random.default_rng(5)
        arr = np.cumsum(rng.exponential(1.0 / lam, n_at))
        arr -= arr[0]
        r = run(arr)
        r["lambda"] = lam
        sweep.append(r)
        log(f"  lambda {lam}/s: p50 {r['p50_ms']} ms p99 {r['p99_ms']} ms, {r['jobs_per_s']} jobs/s, "
            f"batch {r['mean_batch_jobs']}")
    sat = run(np.zeros(n_req))
    # full results of four requests (the largest among them) against the
    # oracle's two-phase CPU count, which tests pin to its brute-force count
    import oracle as O
    picks = sorted(set([int(np.argmax(sizes)), 0, n_req // 3, 2 * n_req // 3]))
    full_ok = True
    for i in picks:
        rid = srv.submit(*reqs[i])
        srv.wait(rid, 300.0)
        gk, gc, st, _ = srv.result(rid)
        srv.release(rid)
        ok, oc, info_o = O.count_parallel(np.asarray(reqs[i][0]), np.asarray(reqs[i][1]), C5_K, C5_M, 64,
                                          C5_CI, 2**64 - 1, threads)
        same = len(gc) == len(oc)
        if same:
            a = np.lexsort((gk[:, 0], gk[:, 1]))
            b = np.lexsort((ok[:, 0], ok[:, 1]))
            same = bool(np.array_equal(gk[a], ok[b]) and np.array_equal(gc[a], oc[b]))
        full_ok = full_ok and same and st["n_unique"] == info_o["unique"]
    info = srv.stats()
    srv.close()
    head = next((r for r in sweep if r["lambda"] == 100), sweep[-1] if sweep else sat)
    return {
        "workload": f"serving: {n_req} requests of the seed-5 family (own random genome each, log-normal "
                    f"length median 1 Mbp sigma 1.0 max 20 Mbp, 150 bp reads at 15x, no errors), "
                    f"k={C5_K} m={C5_M} ci={C5_CI}, Poisson arrivals seed 5, {lanes} lanes",
        "headline": {"lambda": head.get("lambda"), "p50_ms": head["p50_ms"], "p99_ms": head["p99_ms"],
                     "jobs_per_s": head["jobs_per_s"]},
        "sweep": sweep, "saturation": sat,
        "checks": {"kmers_closed_form_all_requests": not bad, "mismatching_requests": sorted(set(bad))[:10],
                   "full_results_vs_oracle": {"requests": picks, "bases": [int(sizes[i]) for i in picks],
                                              "identical": bool(full_ok)}},
        "server": {k_: info[k_] for k_ in ("lanes", "lane_bases", "hbm_budget", "batches",
                                            "pinned_pool_bytes")},
    }


def bench_output(T, eng, k):
    """TSV output of the last result: device formatting (alone, then with the
    copy into host memory) against the single-threaded host formatter."""
    lib, h = T.lib(), eng._h
    n = int(lib.tg_result_size(h))
    size = int(lib.tg_result_tsv(h, k, 0, n, None, 0))
    buf = np.empty(max(size, 1), dtype=np.uint8)
    ptr = buf.ctypes.data_as(T.ctypes.c_char_p)

    def timed(fn, reps):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2]

    t_fmt = timed(lambda: lib.tg_result_tsv(h, k, 0, n, None, 0), 5)
    t_dev = timed(lambda: lib.tg_result_tsv(h, k, 0, n, ptr, size), 5)
    keys, cnt = eng.result(k)
    ka = np.ascontiguousarray(keys.reshape(-1))
    hbuf = np.empty(max(n, 1) * (k + 12), dtype=np.uint8)
    hres = []
    t_host = timed(lambda: hres.append(lib.tg_format_tsv(k, ka.ctypes.data_as(T._u64p),
                                                          cnt.ctypes.data_as(T._u32p), n,
                                                          hbuf.ctypes.data_as(T.ctypes.c_char_p),
                                                          len(hbuf))), 3)
    same = hres[-1] == size and bytes(buf[:size]) == bytes(hbuf[:size])
    return {"lines": n, "bytes": size, "identical_to_host": bool(same),
            "device_format_ms": round(t_fmt * 1e3, 3), "device_format_gb_s": round(size / t_fmt / 1e9, 2),
            "device_format_and_copy_ms": round(t_dev * 1e3, 3),
            "device_format_and_copy_gb_s": round(size / t_dev / 1e9, 2),
            "host_format_ms": round(t_host * 1e3, 3), "host_format_gb_s": round(size / t_host / 1e9, 2),
            "note": "TSV of the job's result; host = one thread; copy = pageable host memory"}


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on a workload, from the committed PMC
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a copy kernel)."""
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.md")
    if not os.path.exists(p):
        return None
    out = {}
    with open(p) as f:
        for line in f:
            for key in ("hbm_bytes_per_launch", "fetch_bytes_per_launch", "write_bytes_per_launch"):
                if line.startswith(key + ":"):
                    out[key] = int(line.split(":", 1)[1].strip())
    return out or None


def bench_single(T, eng, name, bases, off, steps, warmup, config=None):
    """One workload, inputs resident in HBM: batch parameters (the ones
    given, else autotuned), warm-up, `steps` timed jobs, per-kernel HIP-event
    times of 3 more jobs, and the algorithmic-bytes roofline of the job and of
    its dominant kernel."""
    wl = WORKLOADS[name]
    K, M, CI, NPART = wl["k"], wl["m"], wl["ci"], wl["n"]
    d_b = eng.dev_upload(bases)
    d_o = eng.dev_upload(off)
    N, nch = len(bases), len(off) - 1
    # the autotuner picks the engine's batch parameters for this job shape
    # (bucket size, load factor, routing tiles), checking every candidate
    # reproduces the same totals
    t_tune = time.perf_counter()
    jc = np.array([0, nch], dtype=np.uint64)
    jb = np.array([0, N], dtype=np.uint64)
    if K > 64:
        tuned = {"path": "k > 64: HBM table of positions"}
    elif config:
        kw = {}
        for item in config.split(","):
            key, v = item.split("=")
            kw[key] = float(v) if key in ("load_factor", "rho") else int(v)
        tuned = eng.set_batch_config(**kw).as_dict()
    else:
        tuned, _ = eng.autotune_many(d_b, N, d_o, nch, jc, jb, K, M, CI)
    t_tune = time.perf_counter() - t_tune
    log(f"{name}: {'configured' if config or K > 64 else 'autotuned'} in {t_tune:.1f} s: {tuned}")
    for _ in range(warmup):
        eng.count_device(d_b, N, d_o, nch, K, M, NPART, CI)
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = eng.count_device(d_b, N, d_o, nch, K, M, NPART, CI)
    eng.sync()
    el = time.perf_counter() - t0
    ms_step = el * 1e3 / steps
    gbases = N * steps / el / 1e9
    log(f"{name}: {ms_step:.3f} ms per job, {gbases:.2f} Gbases/s")
    # per-kernel times: the same job with HIP events around every kernel on
    # the context stream (plain launches of the same kernels)
    eng.reset_kernel_times()
    eng.set_profiling(True)
    prof_steps = 3
    for _ in range(prof_steps):
        eng.count_device(d_b, N, d_o, nch, K, M, NPART, CI)
    eng.sync()
    eng.set_profiling(False)
    kt = eng.kernel_times()
    # exact counters of the job for the algorithmic-bytes figures
    # a superkmer of j k-mers has j + k - 1 bases
    sum_len = st.n_kmers + st.n_superkmers * (K - 1)
    if K > 64:
        job_b, count_b, sk_b = alg_bytes_wide(N, st.n_unique, st.n_out, K)
        P = 0
    else:
        job_b, count_b, sk_b, P = alg_bytes(N, st.n_superkmers, sum_len, st.n_unique, st.n_out, K)
    job_roof_gbases = HBM_PEAK_GBS * 1e9 * N / job_b / 1e9
    kern = {k_: (ms / max(n, 1), n) for k_, (ms, n) in kt.items()}
    per_launch = {"count": count_b, "superkmers": sk_b}
    dom = max((x for x in kern if x in per_launch), key=lambda x: kern[x][0])
    dom_bytes = per_launch[dom]
    avg_ms = kern[dom][0]
    achieved = dom_bytes / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
            "avg_launch_ms": round(avg_ms, 4), "alg_bytes_per_launch": int(dom_bytes)}
    tr = load_traffic(dom, name)
    if tr:
        roof["traffic"] = tr.get("hbm_bytes_per_launch")
    out = {
        "workload": wl["desc"], "value": round(gbases, 4), "unit": "Gbases/s", "ms_per_step": round(ms_step, 3),
        "steps": steps, "warmup": warmup,
        "config": {"workload": wl["desc"], "k": K, "m": M, "ci": CI, "n_partitions": NPART, "bases": N,
                   "chunks": nch},
        "kmers_per_s": round(st.n_kmers * steps / el, 1),
        "job_roofline": {"alg_bytes": int(job_b), "roofline_gbases_s": round(job_roof_gbases, 2),
                         "frac": round(gbases / job_roof_gbases, 4)},
        "stats": {"kmers": st.n_kmers, "superkmers": st.n_superkmers, "unique": st.n_unique,
                  "unique_out": st.n_out, "subparts": st.n_subparts, "work_items": st.n_lds_groups,
                  "spilled_kmers": st.n_hbm_parts, "partition_bytes": int(P)},
        "kernels_ms": {k_: round(v[0], 4) for k_, v in kern.items()},
        "kernel_timing": f"HIP events around each kernel, {prof_steps} profiled jobs after the timed region",
        "engine_config": dict(tuned, autotune_s=round(t_tune, 2)),
        "roofline": roof,
    }
    eng.dev_free(d_b)
    eng.dev_free(d_o)
    return out


def _bgzf(text, threads, block=65280, level=1):
    """BGZF members (SAM/BAM specification 4.1) of `text`, compressed on
    `threads` threads, then the empty end-of-file member."""
    import struct
    import zlib
    from concurrent.futures import ThreadPoolExecutor
    mv = memoryview(text)

    def one(i):
        blk = mv[i:i + block]
        c = zlib.compressobj(level, zlib.DEFLATED, -15)
        body = c.compress(blk) + c.flush()
        head = b"\x1f\x8b\x08\x04\x00\x00\x00\x00\x00\xff" + struct.pack("<H", 6) + b"BC" + \
            struct.pack("<HH", 2, 18 + len(body) + 8 - 1)
        return head + body + struct.pack("<II", zlib.crc32(blk) & 0xFFFFFFFF, len(blk))

    with ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(one, range(0, len(text), block)))
    parts.append(bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000"))
    return b"".join(parts)


def bench_end_to_end(name, text, threads, formats=("plain",)):
    """The tuna CLI on the workload's FASTQ file, per input format: plain
    (mapped, parsed on the host threads), gzip (inflated on one thread while
    the parse and the upload follow it, as the reference reads gzip input on
    one decompressor thread) and BGZF (members inflated on all host threads);
    then count, format TSV on the device and write it.  Wall time of the whole
    process, best of two runs (page cache warm)."""
    import gzip as gz
    wl = WORKLOADS[name]
    cli = os.path.join(ROOT, "tuna_amd", "tuna")
    need = 3 * len(text) + (1 << 30)
    base = "/dev/shm" if os.path.isdir("/dev/shm") and shutil.disk_usage("/dev/shm").free > 4 * need else None
    d = tempfile.mkdtemp(prefix="tg_e2e_", dir=base)
    res = {}
    try:
        for fmt in formats:
            fq = os.path.join(d, {"plain": "reads.fq", "gzip": "reads.fq.gz", "bgzf": "reads.bgzf.fq.gz"}[fmt])
            t0 = time.perf_counter()
            with open(fq, "wb") as f:
                if fmt == "plain":
                    f.write(memoryview(text))
                elif fmt == "gzip":
                    f.write(gz.compress(memoryview(text), 1))
                else:
                    f.write(_bgzf(text, threads))
            log(f"{name}: wrote {fmt} input, {os.path.getsize(fq)} bytes ({time.perf_counter() - t0:.1f} s)")
            out = os.path.join(d, "out.tsv")
            best = None
            for _ in range(2):
                t0 = time.perf_counter()
                r = subprocess.run([cli, "-k", str(wl["k"]), "-m", str(wl["m"]), "-n", str(wl["n"]), "-ci",
                                    str(wl["ci"]), "-t", str(threads), "-hp", fq, out],
                                   capture_output=True, text=True, timeout=600)
                wall = time.perf_counter() - t0
                if r.returncode != 0:
                    best = {"error": r.stderr[-300:]}
                    break
                st = dict(l.split(": ", 1) for l in r.stderr.splitlines() if ": " in l and not l.startswith("tuna"))
                size = os.path.getsize(out)
                if best is None or wall < best["wall_s"]:
                    best = {"wall_s": round(wall, 3), "phase1_s": float(st.get("phase1", "0").rstrip("s")),
                            "phase2_s": float(st.get("phase2", "0").rstrip("s")),
                            "unique_kmers": int(st.get("unique_kmers", 0)), "tsv_bytes": size}
                os.remove(out)
            res[fmt] = dict(best, input_bytes=os.path.getsize(fq),
                            command=f"tuna -k {wl['k']} -m {wl['m']} -n {wl['n']} -ci {wl['ci']} -t {threads} -hp "
                                    f"{os.path.basename(fq)} out.tsv")
            os.remove(fq)
            log(f"{name}: end to end, {fmt}: {res[fmt]}")
        return dict(res, fastq_bytes=len(text), threads=threads, output_dir=("tmpfs" if base else "tempdir"))
    finally:
        shutil.rmtree(d, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end CLI runs")
    ap.add_argument("--samples", type=int, default=2000,
                    help="jobs of the many-sample leg (0 skips it)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched leg: plain launches instead of graph replay (same kernels)")
    ap.add_argument("--requests", type=int, default=128,
                    help="requests of the serving leg (0 skips it)")
    ap.add_argument("--lambdas", default="5,20,50,100,200", help="arrival rates (jobs/s) of the serving leg")
    ap.add_argument("--config", default=None,
                    help="single legs: batch parameters as key=value,... instead of autotuning them "
                         "(e.g. kmers_per_bucket=4096,max_tiles=1024,load_factor=0.9)")
    ap.add_argument("--autotune", action="store_true",
                    help="single legs: autotune the batch parameters instead of using the workload's "
                         "(the ones the autotuner picked for it, fixed so that profiles match the run)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default=None,
                    help="with --only single: the one workload to run (default: both)")
    ap.add_argument("--only", choices=["all", "single", "batched", "serving", "copy"], default="all",
                    help="run one leg; 'copy' runs a plain 256 MiB device copy kernel 5 times "
                         "(the reference size for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this engine runs on one GPU "
              f"(value is the one GPU's throughput)", file=sys.stderr)
        sys.exit(2)
    import tuna_amd as T

    nproc, threads = host_threads()
    if args.only == "copy":
        import torch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": src.numel() * 4,
                          "launches": 5}))
        return
    lambdas = [float(x) for x in args.lambdas.split(",") if x]
    if args.only == "batched":
        eng = T.Engine(0)
        print(json.dumps({"batched": bench_batched(T, eng, args.samples, threads, not args.no_graph)}))
        eng.close()
        return
    if args.only == "serving":
        print(json.dumps({"serving": bench_serving(T, args.requests, lambdas, threads)}))
        return
    names = [HEADLINE, "short_reads", "long_k"]
    if args.only == "single" and args.workload:
        names = [args.workload]
    eng = T.Engine(0)
    legs, texts = {}, {}
    for name in names:
        t0 = time.perf_counter()
        text = T.synth(WORKLOADS[name]["spec"])
        bases, off = T.parse_buffer(text, threads=threads)
        log(f"{name}: generated {len(text)} bytes FASTQ -> {len(bases)} bases in {len(off) - 1} chunks "
            f"({time.perf_counter() - t0:.1f} s)")
        cfg = args.config or (None if args.autotune else WORKLOADS[name].get("config"))
        legs[name] = bench_single(T, eng, name, bases, off, args.steps, args.warmup, cfg)
        if name == "short_reads" and args.only == "all":
            legs[name]["output"] = bench_output(T, eng, WORKLOADS[name]["k"])
        if name == HEADLINE and args.only == "all":
            if not args.no_cpu:
                try:
                    legs[name]["cpu_baseline"] = cpu_baseline(bases, off, WORKLOADS[name], threads, nproc)
                except Exception as e:  # reported, not fatal
                    legs[name]["cpu_baseline"] = {"value": None, "error": str(e)[:200]}
        del bases, off
        if args.only == "all" and not args.no_e2e:
            texts[name] = text
        del text
    head = legs[names[0]]
    out = {
        "metric": METRIC, "value": head["value"], "unit": "Gbases/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": head["config"],
    }
    for key in ("kmers_per_s", "job_roofline", "stats", "kernels_ms", "kernel_timing", "engine_config",
                "roofline"):
        if key in head:
            out[key] = head[key]
    if "cpu_baseline" in head:
        out["cpu_baseline"] = head["cpu_baseline"]
    for name in names[1:]:
        out[name] = legs[name]
    if texts:
        e2e = {}
        for name, text in texts.items():
            fmts = ("plain", "gzip", "bgzf") if name == "short_reads" else ("plain",)
            r = bench_end_to_end(name, text, threads, fmts)
            N = legs[name]["config"]["bases"]
            for fmt in fmts:
                if "wall_s" in r.get(fmt, {}):
                    r[fmt]["gbases_per_s"] = round(N / r[fmt]["wall_s"] / 1e9, 3)
            e2e[name] = r
        out["end_to_end"] = e2e
        texts.clear()
    if args.samples > 0 and args.only == "all":
        out["batched"] = bench_batched(T, eng, args.samples, threads, not args.no_graph)
    if args.requests > 0 and args.only == "all":
        out["serving"] = bench_serving(T, args.requests, lambdas, threads)
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
