"""Full-size workloads.  The 2 Gbp repetitive-genome job is compared with the
oracle's multithreaded two-phase count; all of them are also checked through
size-independent properties:
  * the counts add up to the number of k-mers, and every distinct k-mer is
    reported when ci = 1;
  * the multiset of (k-mer, count) pairs, reduced to an order-free checksum,
    is the same whatever the partitioning (bucket size) and whichever engine
    mode (batched launch sequence, plain host-driven sequence) counts it."""
import sys

import numpy as np
import pytest

import tuna_amd as T

pytestmark = pytest.mark.gpu

C2 = "kind=reads_fq genome=5000000 seed=2 gc=0.5 reads=1666667 rl=150 sub=0.005 nrate=0.0001"
C3 = ("kind=repeat_fq genome=100000000 seed=3 gc=0.45 cov=20 rl=150 sub=0.003 "
      "fam=50 flen=300 div=0.05 rep=0.35 tan=0.10")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def checksum(keys, cnt):
    """Sum over pairs of a mixed 64-bit hash of (key words, count): independent
    of the order of the pairs, sensitive to any changed, lost or split pair."""
    h = np.zeros(len(cnt), dtype=np.uint64)
    with np.errstate(over="ignore"):
        for w in range(keys.shape[1]):
            h ^= keys[:, w] * np.uint64(0x9E3779B97F4A7C15 + 2 * w)
            h = (h ^ (h >> np.uint64(31))) * np.uint64(0xBF58476D1CE4E5B9)
        h ^= cnt.astype(np.uint64) * np.uint64(0x94D049BB133111EB)
        h = (h ^ (h >> np.uint64(29))) * np.uint64(0xD6E8FEB86659FD93)
        return int(h.sum(dtype=np.uint64))


def summary(eng, k):
    keys, cnt = eng.result(k)
    return len(cnt), int(cnt.sum(dtype=np.uint64)), checksum(keys, cnt)


@pytest.mark.parametrize("name,spec,k,m,ci,n", [
    ("short reads 250 Mbp", C2, 31, 21, 2, 256),
    ("repetitive genome reads 2 Gbp", C3, 31, 23, 1, 4096),
    ("repetitive genome reads 2 Gbp, n=256", C3, 31, 23, 1, 256),
])
def test_full_size_properties(name, spec, k, m, ci, n):
    bases, off = T.synth_parsed(spec)
    say(f"{name}: {len(bases)} bases, {len(off) - 1} chunks")
    eng = T.Engine(0)
    d_b = eng.dev_upload(bases)
    d_o = eng.dev_upload(off)
    N, nch = len(bases), len(off) - 1
    del bases
    try:
        st = eng.count_device(d_b, N, d_o, nch, k, m, n, ci)
        ref = summary(eng, k)
        say(f"  batched: {st.n_kmers} k-mers, {st.n_unique} distinct, {ref[0]} reported")
        assert ref[0] == st.n_out
        if ci == 1:
            assert ref[1] == st.n_kmers and ref[0] == st.n_unique
        else:
            assert ref[1] <= st.n_kmers and ref[0] <= st.n_unique
        # another partitioning of the same input
        eng.set_batch_config(kmers_per_bucket=65536)
        st2 = eng.count_device(d_b, N, d_o, nch, k, m, n, ci)
        assert (st2.n_kmers, st2.n_unique, st2.n_superkmers) == (st.n_kmers, st.n_unique, st.n_superkmers)
        assert summary(eng, k) == ref
        say("  other bucket size: same multiset")
        # the plain host-driven sequence
        eng.set_plain_sequence(True)
        st3 = eng.count_device(d_b, N, d_o, nch, k, m, n, ci)
        assert (st3.n_kmers, st3.n_unique, st3.n_superkmers) == (st.n_kmers, st.n_unique, st.n_superkmers)
        assert summary(eng, k) == ref
        say("  plain sequence: same multiset")
    finally:
        eng.dev_free(d_b)
        eng.dev_free(d_o)
        eng.close()


C3_25X = C3.replace("cov=20", "cov=25")


def test_job_above_2_31_bases():
    """2.5 Gbp of repetitive-genome reads (the 2 Gbp workload's generator at
    25x): more than 2^31 bases, so the job is counted in slices.  With ci = 1
    the counts add up to the k-mers and every distinct k-mer is reported; the
    multiset is the same with another slice size, in the plain sequence, and
    counted from host memory (each slice uploaded alone)."""
    bases, off = T.synth_parsed(C3_25X)
    N, nch = len(bases), len(off) - 1
    say(f"2.5 Gbp reads: {N} bases, {nch} chunks")
    assert N > (1 << 31)
    eng = T.Engine(0)
    d_b = eng.dev_upload(bases)
    d_o = eng.dev_upload(off)
    try:
        st = eng.count_device(d_b, N, d_o, nch, 31, 23, 4096, 1)
        ref = summary(eng, 31)
        say(f"  default slices: {st.n_kmers} k-mers, {st.n_unique} distinct, {st.n_superkmers} superkmers")
        assert ref[0] == st.n_out == st.n_unique and ref[1] == st.n_kmers
        kmers = int(np.sum(np.maximum(np.diff(off.astype(np.int64)) - 30, 0)))
        assert st.n_kmers == kmers
        eng.set_batch_config(slice_bases=1 << 30)
        st2 = eng.count_device(d_b, N, d_o, nch, 31, 23, 4096, 1)
        assert (st2.n_kmers, st2.n_unique, st2.n_superkmers) == (st.n_kmers, st.n_unique, st.n_superkmers)
        assert summary(eng, 31) == ref
        say("  1 Gbp slices: same multiset")
        eng.set_plain_sequence(True)
        st3 = eng.count_device(d_b, N, d_o, nch, 31, 23, 4096, 1)
        assert (st3.n_kmers, st3.n_unique, st3.n_superkmers) == (st.n_kmers, st.n_unique, st.n_superkmers)
        assert summary(eng, 31) == ref
        say("  plain sequence: same multiset")
        eng.set_plain_sequence(False)
        eng.set_batch_config(slice_bases=0)
        eng.dev_free(d_b)
        d_b = None
        st4 = eng.count(bases, off, 31, 23, 4096, 1, fetch=False)
        assert (st4.n_kmers, st4.n_unique, st4.n_superkmers) == (st.n_kmers, st.n_unique, st.n_superkmers)
        assert summary(eng, 31) == ref
        say("  host-resident input, slices uploaded one at a time: same multiset")
    finally:
        if d_b is not None:
            eng.dev_free(d_b)
        eng.dev_free(d_o)
        eng.close()


def test_full_size_repeat_reads_against_oracle():
    """The 2 Gbp repetitive-genome job at its own settings (k=31 m=23 ci=1,
    n=4096) against the oracle's two-phase count on the host's threads: the
    same totals (k-mers, superkmers, distinct), exactly the same set of keys,
    and the same (key, count) multiset by an order-free checksum of the pairs."""
    import os
    import oracle as O
    try:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        threads = 1
    bases, off = T.synth_parsed(C3)
    say(f"repeat reads vs oracle: {len(bases)} bases, {len(off) - 1} chunks, {threads} host threads")
    eng = T.Engine(0)
    try:
        gk, gc = eng.count(bases, off, 31, 23, 4096, 1)
        st = eng.last_stats
    finally:
        eng.close()
    say(f"  device: {st.n_kmers} k-mers, {st.n_superkmers} superkmers, {st.n_unique} distinct")
    ok, oc, info = O.count_parallel(bases, off, 31, 23, 4096, 1, 2**64 - 1, threads)
    del bases, off
    say(f"  oracle: {info['kmers']} k-mers, {info['superkmers']} superkmers, {info['unique']} distinct "
        f"({info['phase1_s'] + info['phase2_s']:.1f} s)")
    assert (st.n_kmers, st.n_superkmers, st.n_unique, st.n_out) == \
        (info["kmers"], info["superkmers"], info["unique"], info["out"])
    assert len(gc) == len(oc) and int(gc.sum(dtype=np.uint64)) == int(oc.sum(dtype=np.uint64))
    assert checksum(gk, gc) == checksum(ok, oc)
    assert np.array_equal(np.sort(gk[:, 0]), np.sort(ok[:, 0]))
    say("  same keys, same (key, count) multiset")


def test_wide_job_above_2_32_bases():
    """k = 127 on a 2.2 Gbp read set A, then on A followed by a copy of itself
    (4.4 Gbp: the copy's k-mers sit past base 2^32, so the table's positions
    and the packed stream's offsets need more than 32 bits).  Every key of
    A + A is a key of A with twice its count: the same distinct keys, twice
    the k-mers and superkmers, and the (key, count) checksum of A + A equal
    to that of A with doubled counts."""
    spec = "kind=reads_fq genome=20000000 seed=9 gc=0.5 reads=14700000 rl=150 sub=0.001 nrate=0"
    bases, off = T.synth_parsed(spec)
    N, nch = len(bases), len(off) - 1
    say(f"wide above 2^32: A = {N} bases, {nch} chunks")
    assert 2 * N > 2**32
    eng = T.Engine(0)
    try:
        ka, ca = eng.count(bases, off, 127, 21, 16, 1)
        sa = eng.last_stats
        say(f"  A: {sa.n_kmers} k-mers, {sa.n_superkmers} superkmers, {sa.n_unique} distinct")
        ref = (len(ca), checksum(ka, ca.astype(np.uint64) * np.uint64(2)))
        del ka, ca
        bb = np.concatenate([bases, bases])
        oo = np.concatenate([off, off[1:] + np.uint64(N)])
        del bases, off
        kb, cb = eng.count(bb, oo, 127, 21, 16, 1)
        sb = eng.last_stats
        say(f"  A + A: {sb.n_kmers} k-mers, {sb.n_superkmers} superkmers, {sb.n_unique} distinct")
    finally:
        eng.close()
    assert (sb.n_kmers, sb.n_superkmers, sb.n_unique, sb.n_out) == \
        (2 * sa.n_kmers, 2 * sa.n_superkmers, sa.n_unique, sa.n_out)
    assert int(cb.sum(dtype=np.uint64)) == sb.n_kmers and int(cb.min()) >= 2
    assert (len(cb), checksum(kb, cb.astype(np.uint64))) == ref
    say("  every key twice as often")
