"""Batched path (many independent jobs per launch sequence): every job's
result must equal counting it alone (oracle and single-job path), bit-exact,
whatever the batching, graph replay or overflow handling."""
import random
import sys
import time

import numpy as np
import pytest

import oracle as O
import tuna_amd as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def eng():
    e = T.Engine(0)
    yield e
    e.close()


def _sorted(keys, cnt):
    if keys.shape[1] == 1:
        order = np.lexsort((keys[:, 0],))
    else:
        order = np.lexsort((keys[:, 0], keys[:, 1]))
    return keys[order], cnt[order]


def concat(parts):
    """parts: list of (bases u8, off u64) -> bases, off, job_chunk."""
    bases, offs, jc = [], [np.zeros(1, dtype=np.uint64)], [0]
    base = 0
    nch = 0
    for b, o in parts:
        bases.append(np.asarray(b, dtype=np.uint8))
        offs.append(np.asarray(o[1:], dtype=np.uint64) + base)
        base += len(b)
        nch += len(o) - 1
        jc.append(nch)
    allb = np.concatenate(bases) if bases else np.zeros(0, dtype=np.uint8)
    return allb, np.concatenate(offs), np.array(jc, dtype=np.uint64)


def rand_fa(rng, n, name="g"):
    return (">" + name + "\n" + "".join(rng.choice("ACGT") for _ in range(n)) + "\n").encode()


def job_inputs(seed):
    rng = random.Random(seed)
    texts = [
        rand_fa(rng, 5000),
        T.synth("kind=reads_fq genome=40000 seed=11 cov=8 rl=150 sub=0.005 nrate=0.001").tobytes(),
        b">empty\n",                                 # no chunks at all
        b">short\nACGTN" + b"ACG" * 3 + b"\n",       # only chunks shorter than k
        rand_fa(rng, 21, "exact"),                   # one k-mer
        T.synth("kind=samples_fq samples=2000 seed=4 index=3 cov=5 rl=150 sub=0.005").tobytes(),
        rand_fa(rng, 120000),
        b">lc\n" + b"A" * 9000 + b"CG" * 3000 + b"ACGTTGCA" * 900 + b"\n",  # low complexity
        T.synth("kind=reads_fq genome=25000 seed=12 cov=12 rl=100 sub=0.01 nrate=0.0").tobytes(),
    ]
    return [T.parse_buffer(t) for t in texts]


def check_jobs(eng, parts, k, m, ci=1, cx=2**64 - 1, label=""):
    bases, off, jc = concat(parts)
    say(f"count_many {label} k={k} m={m} jobs={len(parts)} N={len(bases)} chunks={len(off) - 1}")
    st = eng.count_many(bases, off, jc, k, m, ci, cx)
    for j, (b, o) in enumerate(parts):
        gk, gc = eng.many_result(j, k)
        ok, oc, tot, nu = O.count(b, o, k, ci, cx)
        assert st[j]["n_bases"] == len(b), j
        assert st[j]["n_kmers"] == tot, j
        assert st[j]["n_unique"] == nu, j
        assert st[j]["n_out"] == len(oc) == len(gc), j
        if len(o) > 1:
            nsk = len(O.superkmers(b, o, k, m)[0])
            assert st[j]["n_superkmers"] == nsk, j
        gk, gc = _sorted(gk, gc)
        ok, oc = _sorted(ok, oc)
        assert np.array_equal(gk, ok), j
        assert np.array_equal(gc, oc), j
    return st


def test_many_jobs_match_oracle(eng):
    check_jobs(eng, job_inputs(1), 21, 15, label="mixed")


def test_many_jobs_filters(eng):
    check_jobs(eng, job_inputs(2), 31, 21, ci=2, cx=40, label="ci2 cx40")


def test_many_jobs_long_k(eng):
    check_jobs(eng, job_inputs(3), 63, 21, ci=2, label="k63")
    check_jobs(eng, job_inputs(4), 45, 9, label="k45")


def test_many_odd_k(eng):
    check_jobs(eng, job_inputs(5), 32, 16, label="k32")
    check_jobs(eng, job_inputs(6), 5, 3, label="k5")


def test_graph_replay_equals_plain_launches(eng):
    parts = job_inputs(7)
    bases, off, jc = concat(parts)
    eng.set_batch_config(use_graph=0)
    try:
        say("plain launches")
        eng.count_many(bases, off, jc, 25, 11)
        assert eng.last_many.n_graph_launches == 0
        plain = [_sorted(*eng.many_result(j, 25)) for j in range(len(parts))]
    finally:
        eng.set_batch_config(use_graph=1)
    say("graph replay x3")
    for rep in range(3):
        eng.count_many(bases, off, jc, 25, 11)
        ms = eng.last_many
        assert ms.n_graph_launches == ms.n_batches >= 1
        for j in range(len(parts)):
            gk, gc = _sorted(*eng.many_result(j, 25))
            assert np.array_equal(gk, plain[j][0]) and np.array_equal(gc, plain[j][1]), (rep, j)
    # the class and its graph are reused by the later calls
    assert eng.last_many.n_captures == 0


def test_many_batches_in_one_call(eng):
    parts = job_inputs(8)
    old = eng.batch_config()
    eng.set_batch_config(batch_bases=1 << 16)
    try:
        check_jobs(eng, parts, 21, 15, label="small batches")
        assert eng.last_many.n_batches > 3
    finally:
        eng.set_batch_config(batch_bases=old.batch_bases)


def test_spill_and_retry_path(eng):
    # one bucket per job and a tiny distinct estimate: work items far larger
    # than an LDS table, so most keys go through the tagged spill table, which
    # overflows and is regrown
    rng = random.Random(9)
    texts = [rand_fa(rng, 300000), rand_fa(rng, 80000), rand_fa(rng, 3000)]
    parts = [T.parse_buffer(t) for t in texts]
    old = eng.batch_config()
    eng.set_batch_config(kmers_per_bucket=1 << 20, rho=0.02)
    try:
        check_jobs(eng, parts, 21, 15, label="spill")
        check_jobs(eng, parts, 63, 21, label="spill k63")
    finally:
        eng.set_batch_config(kmers_per_bucket=old.kmers_per_bucket, rho=0.0)


def test_ratio_learned_on_other_input(eng):
    # the distinct/k-mer ratio learned on high-coverage reads (about 0.24)
    # sizes the work items of a genome whose k-mers are all distinct: the
    # items then hold about three tables' worth of keys each and overflow the
    # spill table; the overflowed attempt must end quickly and the retry must
    # re-split the items
    reads = T.parse_buffer(T.synth("kind=reads_fq genome=200000 seed=2 cov=20 rl=150 sub=0.005 nrate=0.0001"))
    genome = T.parse_buffer(T.synth("kind=genome_fa len=4600000 seed=1"))
    assert eng.batch_config().rho == 0.0
    check_jobs(eng, [reads], 31, 21, label="reads (learns the ratio)")
    t0 = time.time()
    check_jobs(eng, [genome], 31, 21, label="genome (all k-mers distinct)")
    ms = eng.last_many
    say(f"  retries {ms.n_retries}, {time.time() - t0:.1f} s with the oracle")
    assert ms.n_retries <= 1
    check_jobs(eng, [genome], 31, 21, label="genome again (ratio learned)")
    assert eng.last_many.n_retries == 0


def test_many_samples_workload(eng):
    # jobs of the many-sample workload (log-normal genome sizes, 10x reads)
    parts = []
    for i in range(6):
        t = T.synth(f"kind=samples_fq samples=2000 seed=4 index={i} cov=10 rl=150 sub=0.005 hi=400000")
        parts.append(T.parse_buffer(t.tobytes()))
    check_jobs(eng, parts, 21, 15, label="samples")


def test_results_equal_single_job_path(eng):
    parts = job_inputs(10)
    bases, off, jc = concat(parts)
    st = eng.count_many(bases, off, jc, 31, 21)
    for j, (b, o) in enumerate(parts):
        if len(o) < 2:
            continue
        sk, sc = _sorted(*eng.count(b, o, 31, 21))
        s1 = eng.last_stats
        assert st[j]["n_superkmers"] == s1.n_superkmers
        mk, mc = _sorted(*eng.many_result(j, 31))
        assert np.array_equal(sk, mk) and np.array_equal(sc, mc), j


def test_device_offsets_must_match_job_table(eng):
    parts = job_inputs(11)[:2]
    bases, off, jc = concat(parts)
    d_b = eng.dev_upload(bases)
    d_o = eng.dev_upload(off)
    try:
        jb = off[jc.astype(np.int64)].copy()
        eng.count_many_device(d_b, len(bases), d_o, len(off) - 1, jc, jb, 21, 15)
        jb2 = jb.copy()
        jb2[-1] -= 7  # the last job claims fewer bases than its chunks cover
        with pytest.raises(T.TunaGpuError, match="disagree"):
            eng.count_many_device(d_b, len(bases), d_o, len(off) - 1, jc, jb2, 21, 15)
    finally:
        eng.dev_free(d_b)
        eng.dev_free(d_o)


def test_batch_config_validation(eng):
    with pytest.raises(T.TunaGpuError):
        eng.set_batch_config(kmers_per_bucket=1000)
    with pytest.raises(T.TunaGpuError):
        eng.set_batch_config(load_factor=0.0)
    c = eng.batch_config()
    assert c.kmers_per_bucket >= 256 and c.use_graph in (0, 1)


def test_autotune_installs_a_config_with_identical_results(eng):
    parts = job_inputs(12)
    bases, off, jc = concat(parts)
    jb = off[jc.astype(np.int64)].copy()
    d_b = eng.dev_upload(bases)
    d_o = eng.dev_upload(off)
    old = eng.batch_config()
    try:
        say("autotune")
        best, ms = eng.autotune_many(d_b, len(bases), d_o, len(off) - 1, jc, jb, 21, 15)
        assert ms > 0 and best["kmers_per_bucket"] in (4096, 8192, 16384, 32768, 65536)
        assert eng.batch_config().as_dict() == best
        check_jobs(eng, parts, 21, 15, label="tuned")
    finally:
        eng.dev_free(d_b)
        eng.dev_free(d_o)
        eng.set_batch_config(**{k_: getattr(old, k_) for k_, _ in old._fields_})


@pytest.mark.gpu
def test_single_result_survives_many_run(eng):
    # a single job's result stays where the batch gathered it; a many-job run
    # that follows must move it first, so reading it afterwards gives the same
    parts = job_inputs(7)
    b, o = parts[1]
    eng.count(b, o, 31, 21, 16, 1, fetch=False)
    k1, c1 = eng.result(31)
    k1, c1 = _sorted(np.array(k1).reshape(len(c1), -1), np.array(c1))
    bases, off, jc = concat(parts)
    eng.count_many(bases, off, jc, 63, 21)
    k2, c2 = eng.result(31)
    k2, c2 = _sorted(np.array(k2).reshape(len(c2), -1), np.array(c2))
    say(f"  single job: {len(c1)} k-mers before and {len(c2)} after a many-job run")
    assert len(c1) > 0
    assert np.array_equal(k1, k2) and np.array_equal(c1, c2)


@pytest.mark.gpu
def test_single_result_survives_tile_change(eng):
    # changing the routing tiles frees the batch buffers: a single job's result
    # kept there must move first, so the copy and the device TSV/KFF of the
    # result are the same bytes before and after
    parts = job_inputs(8)
    b, o = parts[1]
    old = eng.batch_config()
    try:
        eng.count(b, o, 31, 21, 16, 1, fetch=False)
        k1, c1 = eng.result(31)
        tsv1, kff1 = eng.result_tsv(31), eng.result_kff(31)
        eng.set_batch_config(max_tiles=256 if old.max_tiles != 256 else 512)
        k2, c2 = eng.result(31)
        assert len(c1) > 0 and np.array_equal(k1, k2) and np.array_equal(c1, c2)
        assert eng.result_tsv(31) == tsv1 and eng.result_kff(31) == kff1
        # the next job on the reconfigured engine is counted as usual
        check = O.count(b, o, 31, 1)
        gk, gc = eng.count(b, o, 31, 21, 16, 1)
        gk, gc = _sorted(gk, gc)
        ok, oc = _sorted(check[0], check[1])
        assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
        say(f"  single job of {len(c1)} k-mers: same bytes after a tile change")
    finally:
        eng.set_batch_config(**{k_: getattr(old, k_) for k_, _ in old._fields_})


def test_result_store_across_key_widths(eng):
    """Counts that alternate between one-word (k <= 32) and two-word keys
    (k > 32) keep their result store sized by their results: a width change
    does not grow it."""
    b, off = T.parse_buffer(T.synth("kind=reads_fq genome=3000 seed=3 reads=40 rl=100 sub=0.01 nrate=0.0"))
    want = {}
    for k in (21, 33):
        ok, oc, _, _ = O.count(b, off, k)
        want[k] = len(oc)
    for i in range(60):
        k = (21, 33)[i % 2]
        gk, gc = eng.count(b, off, k, 11, 16)
        assert len(gc) == want[k], (i, k)
