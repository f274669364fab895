"""A single job larger than one pass is counted in slices of whole chunks,
each slice counted without the count filter and merged on the device
(tg_batch_config.slice_bases); an input in host memory (tg_count_host, or
files above the device input budget) is uploaded one slice at a time.  The slice size is forced small here so that
small inputs span many slices; every result is compared with the oracle,
bit-exact, in both single-job modes (batched engine, plain sequence)."""
import random
import sys

import numpy as np
import pytest

import oracle as O
import tuna_amd as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _sorted(keys, cnt):
    if keys.shape[1] == 1:
        order = np.lexsort((keys[:, 0],))
    else:
        order = np.lexsort((keys[:, 0], keys[:, 1]))
    return keys[order], cnt[order]


@pytest.fixture(scope="module", params=["batched", "plain"])
def seng(request):
    e = T.Engine(0)
    e.set_plain_sequence(request.param == "plain")
    old = e.batch_config()
    yield e
    e.set_batch_config(**{k_: getattr(old, k_) for k_, _ in old._fields_})
    e.close()


def inputs():
    rng = random.Random(12)
    reads = T.synth_parsed("kind=reads_fq genome=60000 seed=21 cov=12 rl=150 sub=0.005 nrate=0.001")
    genome = ">g\n" + "".join(rng.choice("ACGT") for _ in range(30000)) + "N" + \
             "".join(rng.choice("ACGT") for _ in range(9000)) + "\n"
    return [("reads", reads), ("genome", O.parse(genome.encode()))]


@pytest.mark.parametrize("k,m,ci,cx", [(31, 21, 1, 2**64 - 1), (31, 21, 2, 2**64 - 1), (21, 11, 2, 5),
                                       (63, 21, 2, 2**64 - 1), (17, 9, 1, 2**64 - 1)])
def test_sliced_counts_match_oracle(seng, k, m, ci, cx):
    for name, (b, o) in inputs():
        ok, oc, tot, nu = O.count(b, o, k, ci, cx)
        ok, oc = _sorted(ok, oc)
        seng.set_batch_config(slice_bases=0)
        one = seng.count(b, o, k, m, 16, ci, cx, fetch=False)
        for sl in (1024, 20000, 400000):
            seng.set_batch_config(slice_bases=sl)
            gk, gc = seng.count(b, o, k, m, 16, ci, cx)
            st = seng.last_stats
            say(f"  {name} k={k} ci={ci} slice {sl}: {len(gc)} reported of {st.n_unique} distinct")
            assert st.n_kmers == tot == one.n_kmers
            assert st.n_unique == nu == one.n_unique
            assert st.n_out == len(oc) == len(gc)
            # superkmers: chunks are never cut, so the totals add up exactly
            assert st.n_superkmers == one.n_superkmers
            gk, gc = _sorted(gk, gc)
            assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
        # device-formatted output of a sliced result
        assert seng.result_tsv(k) == T.format_tsv(k, *seng.result(k))


def test_sliced_device_resident_input(seng):
    b, o = inputs()[0][1]
    d_b = seng.dev_upload(b)
    d_o = seng.dev_upload(o)
    try:
        seng.set_batch_config(slice_bases=5000)
        st = seng.count_device(d_b, len(b), d_o, len(o) - 1, 31, 21, 16, 2)
        gk, gc = _sorted(*seng.result(31))
        ok, oc, tot, nu = O.count(b, o, 31, 2)
        ok, oc = _sorted(ok, oc)
        assert (st.n_kmers, st.n_unique, st.n_out) == (tot, nu, len(oc))
        assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
    finally:
        seng.dev_free(d_b)
        seng.dev_free(d_o)


def test_files_larger_than_the_input_budget(seng, tmp_path):
    # input files above the device input budget are gathered in host memory
    # and counted slice by slice (each slice uploaded alone); also gzip
    import gzip
    text = bytes(T.synth("kind=reads_fq genome=60000 seed=23 cov=12 rl=150 sub=0.005 nrate=0.001"))
    rng = random.Random(5)
    fa = (">a\n" + "".join(rng.choice("ACGT") for _ in range(20000)) + "\n>b\n" +
          "".join(rng.choice("ACGTN") for _ in range(15000)) + "\n").encode()
    files = [tmp_path / "r.fq", tmp_path / "g.fa.gz"]
    files[0].write_bytes(text)
    files[1].write_bytes(gzip.compress(fa))
    b1, o1 = O.parse(text)
    b2, o2 = O.parse(fa)
    b = np.concatenate([b1, b2])
    o = np.concatenate([o1, o2[1:] + o1[-1]]).astype(np.uint64)
    try:
        for k, m, ci in ((31, 21, 2), (63, 21, 1)):
            ok, oc, tot, nu = O.count(b, o, k, ci)
            ok, oc = _sorted(ok, oc)
            for budget, sl in ((0, 0), (4096, 0), (4096, 20000)):
                seng.set_input_budget(budget)
                seng.set_batch_config(slice_bases=sl)
                gk, gc = seng.count_files([str(f) for f in files], k, m, 16, ci, threads=4)
                st = seng.last_stats
                say(f"  files k={k} budget {budget} slice {sl}: {len(gc)} reported")
                assert (st.n_kmers, st.n_unique, st.n_out) == (tot, nu, len(oc))
                gk, gc = _sorted(gk, gc)
                assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
    finally:
        seng.set_input_budget(0)
        seng.set_batch_config(slice_bases=0)


def test_slice_config_validation(seng):
    with pytest.raises(T.TunaGpuError, match="slice_bases"):
        seng.set_batch_config(slice_bases=100)
    with pytest.raises(T.TunaGpuError, match="slice_bases"):
        seng.set_batch_config(slice_bases=1 << 31)


def test_sliced_phase1_exports(seng):
    # the superkmer and partition exports of an input larger than the slice
    # limit: the same superkmers (global starts), sub-partitions and reference
    # partition bytes as in one pass
    b, o = inputs()[1][1]
    r2, r2o = inputs()[0][1]
    bases = np.concatenate([b, r2])
    off = np.concatenate([o, r2o[1:] + o[-1]]).astype(np.uint64)
    seng.set_batch_config(slice_bases=0)
    one = seng.superkmers(bases, off, 31, 21, 64)
    parts_one = seng.partitions(bases, off, 31, 21, 64)
    seng.set_batch_config(slice_bases=3000)
    sl = seng.superkmers(bases, off, 31, 21, 64)
    parts_sl = seng.partitions(bases, off, 31, 21, 64)
    assert sl[4] == one[4]
    a = sorted(zip(*(x.tolist() for x in one[:4])))
    c = sorted(zip(*(x.tolist() for x in sl[:4])))
    assert a == c and len(a) == len(O.superkmers(bases, off, 31, 21)[0])
    assert [sorted(p) for p in parts_one] == [sorted(p) for p in parts_sl]
    seng.set_batch_config(slice_bases=0)
    cov_one = seng.debug_min_coverage(bases, off, 31, 21)
    seng.set_batch_config(slice_bases=3000)
    cov_sl = seng.debug_min_coverage(bases, off, 31, 21)
    assert all(np.array_equal(x, y) for x, y in zip(cov_one, cov_sl))
