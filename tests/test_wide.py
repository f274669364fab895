"""k > 64: keys of 3 to 8 words, counted through the HBM table path, against
the oracle (whose multi-word counts the CPU tests pin to a string brute
force).  Bit-exact: sorted (key, count) lists, the k-mer, distinct, reported
and superkmer totals."""
import random
import sys

import numpy as np
import pytest

import oracle as O
import tuna_amd as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rs(rng, n, alph="ACGT"):
    return "".join(rng.choice(alph) for _ in range(n))


@pytest.fixture(scope="module")
def eng():
    e = T.Engine(0)
    yield e
    e.close()


def _sorted(keys, cnt):
    order = np.lexsort(tuple(keys[:, i] for i in range(keys.shape[1])))
    return keys[order], cnt[order]


def check(eng, strings, k, m, ci=1, cx=2**64 - 1, label=""):
    bases, off = O.chunks_of(strings)
    say(f"wide {label} k={k} m={m} ci={ci} N={len(bases)} chunks={len(off) - 1}")
    gk, gc = eng.count(bases, off, k, m, 16, ci, cx)
    st = eng.last_stats
    ok, oc, tot, nu = O.count(bases, off, k, ci, cx)
    sk = O.superkmers(bases, off, k, m)
    assert st.n_kmers == tot
    assert st.n_unique == nu
    assert st.n_out == len(oc) == len(gc)
    assert st.n_superkmers == len(sk[0])
    assert gk.shape[1] == (k + 31) // 32
    gk, gc = _sorted(gk, gc)
    ok, oc = _sorted(ok, oc)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gc, oc)
    return st


def mixed_inputs(rng, k):
    # short chunks (some below k), chunks of several strips, homopolymers and
    # short tandem repeats longer than one superkmer cap, and overlapping reads
    g = rs(rng, 3000)
    reads = []
    for _ in range(150):
        p = rng.randint(0, len(g) - 200)
        reads.append(g[p:p + 200])
    return ([rs(rng, rng.randint(1, 600)) for _ in range(120)] + [rs(rng, 9000), "A" * 3000, "CAG" * 1200,
            rs(rng, 500) + "T" * 700 + rs(rng, 300), "C" * (k - 1), "G" * k] + reads)


@pytest.mark.parametrize("k,m", [(65, 21), (77, 31), (96, 21), (127, 21), (127, 31), (128, 9), (129, 21),
                                 (160, 32), (200, 31), (255, 27), (256, 1)])
def test_wide_counts_match_oracle(eng, k, m):
    rng = random.Random(7000 + k + m)
    check(eng, mixed_inputs(rng, k), k, m, label="mixed")


@pytest.mark.parametrize("k", [65, 127, 200])
def test_wide_filters_and_repeated_calls(eng, k):
    rng = random.Random(k)
    strings = mixed_inputs(rng, k)
    check(eng, strings, k, 21, ci=2, label="ci2")
    check(eng, strings, k, 21, ci=3, cx=10, label="ci3cx10")
    # a k <= 64 count between two wide ones leaves both right
    check(eng, strings, 31, 21, label="narrow")
    check(eng, strings, k, 21, label="again")


def test_wide_edge_inputs(eng):
    # every chunk shorter than k; one chunk of exactly k; an empty input
    check(eng, ["ACGT" * 20, "A"], 127, 21, label="short")
    check(eng, ["ACGT" * 32], 128, 21, label="exact")
    bases, off = np.zeros(0, np.uint8), np.zeros(1, np.uint64)
    keys, cnt = eng.count(bases, off, 127, 21, 16, 1)
    assert len(cnt) == 0 and eng.last_stats.n_kmers == 0


def test_wide_reads_at_scale(eng):
    # 150 bp reads are one strip each; 250 bp reads at 20x of a 200 kbp genome
    text = T.synth("kind=reads_fq genome=200000 seed=31 cov=20 rl=250 sub=0.004 nrate=0.0001")
    bases, off = T.parse_buffer(text)
    say(f"wide reads: {len(bases)} bases")
    gk, gc = eng.count(bases, off, 127, 21, 16, 2)
    st = eng.last_stats
    ok, oc, tot, nu = O.count(bases, off, 127, 2)
    assert st.n_kmers == tot and st.n_unique == nu and len(gc) == len(oc)
    gk, gc = _sorted(gk, gc)
    ok, oc = _sorted(ok, oc)
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)


def test_batched_engine_refuses_k_above_64(eng):
    bases, off = O.chunks_of(["ACGT" * 50, "GGCA" * 40])
    with pytest.raises(T.TunaGpuError, match="batched engine counts k <= 64"):
        eng.count_many(bases, off, np.array([0, 1, 2], np.uint64), 127, 21)
    with pytest.raises(T.TunaGpuError, match="k <= 64"):
        eng.superkmers(bases, off, 127, 21, 16)


@pytest.mark.parametrize("cap", [1, 2, 3, 5])
def test_wide_deque_overflow_recounted(eng, cap):
    # a deque of `cap` entries overflows on most strips (an ascending pair of
    # m-mer hashes suffices for cap 1): those strips are counted again by the
    # pass that keeps the whole window, so totals and counts are unchanged
    rng = random.Random(900 + cap)
    eng.set_wide_deque(cap)
    try:
        check(eng, mixed_inputs(rng, 127), 127, 21, label=f"deque{cap}")
        check(eng, mixed_inputs(rng, 200), 200, 31, label=f"deque{cap}")
    finally:
        eng.set_wide_deque(16)


def test_wide_deque_setting_range(eng):
    for bad in (0, 17):
        with pytest.raises(T.TunaGpuError, match="deque entries"):
            eng.set_wide_deque(bad)


def test_wide_with_kernel_timing(eng):
    # with per-kernel timing the superkmer pass runs on the context stream
    # (one stream), otherwise beside the table insert: the same results
    rng = random.Random(77)
    strings = mixed_inputs(rng, 127)
    eng.set_profiling(True)
    try:
        check(eng, strings, 127, 21, ci=2, label="timed")
        names = set(eng.kernel_times())
        assert {"superkmers", "count", "drain", "job_gather"} <= names
    finally:
        eng.set_profiling(False)
        eng.reset_kernel_times()
    check(eng, strings, 127, 21, ci=2, label="untimed")
