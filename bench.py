#!/usr/bin/env python3
"""Benchmark: Gbases/s counted on one MI355X for the short-read FASTQ workload.

Workload (N=1): a 5 Mbp uniform random genome (seed 2, GC 50%), 150 bp reads
at uniform positions and strands to 50x (1,666,667 reads, 250 Mbp), 0.5%
i.i.d. substitutions, N at 0.01% per base, Phred 30 qualities; counted with
k=31, m=21, ci=2, n=256 partitions.  The FASTQ text is generated and parsed on
the host once; the ACGT chunks and offsets are uploaded to HBM before the
timed region.  One step = one whole counting job on the device (pack,
superkmers, routing, counting, drain), including the job's own host syncs.

Prints one JSON line (see the fields below).  --gpus N > 1 is not supported:
this engine runs on one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gbases/s counted on 1 MI355X (% HBM roofline); batched jobs/s + p99 latency"
HBM_PEAK_GBS = 8000.0
WORKLOAD = ("short-read FASTQ: 5 Mbp genome seed 2, 150 bp reads at 50x "
            "(1,666,667 reads, 250 Mbp), 0.5% substitutions, N 0.01%, k=31 m=21 ci=2 n=256")
SPEC = "kind=reads_fq genome=5000000 seed=2 gc=0.5 reads=1666667 rl=150 sub=0.005 nrate=0.0001"
K, M, CI, NPART = 31, 21, 2, 256


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


def cpu_baseline(text, n_bases_total, threads):
    """Reference tuna (our driver over the reference's own Phase-1 harness and
    kache table, oracle/_ref) on a bounded sample of the same reads, split
    into `threads` files so both phases run on `threads` cores."""
    import oracle as O
    reads_sample = 400000  # 60 Mbp: ~10-30 s of CPU work
    lines = text.split(b"\n")
    nrec = min(reads_sample, len(lines) // 4)
    tmp = tempfile.mkdtemp(prefix="tg_cpu_")
    files = []
    per = (nrec + threads - 1) // threads
    for t in range(threads):
        a, b = t * per, min((t + 1) * per, nrec)
        if a >= b:
            break
        p = os.path.join(tmp, f"s{t}.fq")
        with open(p, "wb") as f:
            f.write(b"\n".join(lines[4 * a:4 * b]) + b"\n")
        files.append(p)
    sample_bases = sum(len(lines[4 * i + 1]) for i in range(nrec))
    if O.ref_available():
        t0 = time.perf_counter()
        r = subprocess.run([O.REF_BIN, "count", str(K), str(M), str(NPART), str(threads), str(CI),
                            str(2**64 - 1), "-"] + files, capture_output=True, text=True)
        wall = time.perf_counter() - t0
        if r.returncode == 0:
            st = dict(l.split(": ") for l in r.stderr.strip().split("\n") if ": " in l)
            return {"value": round(sample_bases / wall / 1e9, 5), "unit": "Gbases/s", "cores": threads,
                    "kind": "reference",
                    "sample": f"first {nrec} reads ({sample_bases} bases) of the workload in {len(files)} "
                              f"FASTQ files; reference Phase 1 (partition_kmers_mem) + kache table via "
                              f"oracle/_ref, wall {wall:.2f} s (phase1 {float(st.get('phase1', 0)):.2f} s, "
                              f"phase2 {float(st.get('phase2', 0)):.2f} s)"}
        log("reference baseline failed:", r.stderr[-400:])
    # fallback: the single-thread oracle port on a smaller sample
    nrec = min(nrec, 100000)
    bases, off = O.parse(b"\n".join(lines[:4 * nrec]) + b"\n")
    t0 = time.perf_counter()
    O.count(bases, off, K, CI)
    wall = time.perf_counter() - t0
    return {"value": round(len(bases) / wall / 1e9, 5), "unit": "Gbases/s", "cores": 1, "kind": "port",
            "sample": f"first {nrec} reads ({len(bases)} bases), oracle brute-force count"}


def load_traffic(kernel):
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel}_short_reads_c2.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this engine runs on one GPU "
              f"(value is the one GPU's throughput)", file=sys.stderr)
        sys.exit(2)
    import tuna_amd as T

    t0 = time.perf_counter()
    text = T.synth(SPEC)
    bases, off = T.parse_buffer(text)
    log(f"generated {len(text)} bytes FASTQ -> {len(bases)} bases in {len(off) - 1} chunks "
        f"({time.perf_counter() - t0:.1f} s)")
    eng = T.Engine(0)
    d_b = eng.dev_upload(bases)
    d_o = eng.dev_upload(off)
    N, nch = len(bases), len(off) - 1

    for _ in range(args.warmup):
        eng.count_device(d_b, N, d_o, nch, K, M, NPART, CI)
    eng.sync()
    eng.reset_kernel_times()
    eng.set_profiling(True)
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = eng.count_device(d_b, N, d_o, nch, K, M, NPART, CI)
    eng.sync()
    el = time.perf_counter() - t0
    eng.set_profiling(False)
    kt = eng.kernel_times()
    ms_step = el * 1e3 / args.steps
    gbases = N * args.steps / el / 1e9

    # exact counters of the job for the algorithmic-bytes figures
    sk = eng.superkmers(bases, off, K, M, NPART)
    sum_len = int(sk[1].astype(np.int64).sum())
    job_b, count_b, sk_b, P = alg_bytes(N, st.n_superkmers, sum_len, st.n_unique, st.n_out, K)
    job_roof_gbases = HBM_PEAK_GBS * 1e9 * N / job_b / 1e9
    kern = {name: (ms / max(n, 1), n) for name, (ms, n) in kt.items()}
    dom = max(kern, key=lambda x: kern[x][0] * kern[x][1])
    per_launch = {"count": count_b, "superkmers": sk_b}
    dom_bytes = per_launch.get(dom)
    if dom_bytes is None:
        dom = "count"
        dom_bytes = count_b
    avg_ms = kern[dom][0]
    achieved = dom_bytes / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
            "avg_launch_ms": round(avg_ms, 4), "alg_bytes_per_launch": int(dom_bytes)}
    tr = load_traffic(dom)
    if tr:
        roof["traffic"] = tr.get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC, "value": round(gbases, 4), "unit": "Gbases/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "k": K, "m": M, "ci": CI, "n_partitions": NPART,
                   "bases": N, "chunks": nch},
        "kmers_per_s": round(st.n_kmers * args.steps / el, 1),
        "job_roofline": {"alg_bytes": int(job_b), "roofline_gbases_s": round(job_roof_gbases, 2),
                         "frac": round(gbases / job_roof_gbases, 4)},
        "stats": {"kmers": st.n_kmers, "superkmers": st.n_superkmers, "unique": st.n_unique,
                  "unique_out": st.n_out, "subparts": st.n_subparts, "work_items": st.n_lds_groups,
                  "spilled_kmers": st.n_hbm_parts},
        "kernels_ms": {k_: round(v[0], 4) for k_, v in kern.items()},
        "roofline": roof,
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        try:
            out["cpu_baseline"] = cpu_baseline(bytes(text), N, threads)
        except Exception as e:  # reported, not fatal
            out["cpu_baseline"] = {"value": None, "error": str(e)[:200]}
    eng.dev_free(d_b)
    eng.dev_free(d_o)
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
