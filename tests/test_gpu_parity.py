"""GPU parity tests: the HIP path through the C ABI against the oracle and the
reference's own golden vectors.  Bit-exact throughout (integer work)."""
import json
import os
import random
import sys

import numpy as np
import pytest

import oracle as O
import tuna_amd as T

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def eng():
    e = T.Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module", params=["batched", "plain"])
def ceng(request):
    """Counting engine in both single-job modes: the batched engine (one
    captured launch sequence) and the plain host-driven sequence."""
    e = T.Engine(0)
    e.set_plain_sequence(request.param == "plain")
    yield e
    e.close()


def rs(rng, n, alph="ACGT"):
    return "".join(rng.choice(alph) for _ in range(n))


def _ilog2(x):
    return x.bit_length() - 1


def _sorted_keys(keys, cnt):
    if keys.shape[1] == 1:
        order = np.lexsort((keys[:, 0],))
    else:
        order = np.lexsort((keys[:, 0], keys[:, 1]))
    return keys[order], cnt[order]


def check_counts(eng, bases, off, k, m, n=16, ci=1, cx=2**64 - 1, label=""):
    say(f"count {label} k={k} m={m} n={n} ci={ci} N={len(bases)} chunks={len(off) - 1}")
    gk, gc = eng.count(bases, off, k, m, n, ci, cx)
    st = eng.last_stats
    ok, oc, tot, nu = O.count(bases, off, k, ci, cx)
    assert st.n_kmers == tot
    assert st.n_unique == nu
    assert st.n_out == len(oc) == len(gc)
    gk, gc = _sorted_keys(gk, gc)
    ok, oc = _sorted_keys(ok, oc)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gc, oc)
    return st


def test_pack_matches_kache_codes(eng):
    rng = random.Random(1)
    s = rs(rng, 1000, "ACGTacgt").encode()
    say("pack 1000")
    w = eng.pack(s)
    codes = [O.lib().orc_kache_code(c) for c in s]
    exp = []
    for i in range(0, len(codes), 16):
        v = 0
        for j in range(16):
            v = (v << 2) | (codes[i + j] if i + j < len(codes) else 0)
        exp.append(v)
    assert list(w) == exp


@pytest.mark.parametrize("plain", [False, True])
def test_counts_from_misaligned_device_inputs(eng, plain):
    # an input that starts anywhere inside a device buffer (a batch or a job
    # inside a concatenation): packing reads it with aligned 16-byte loads
    # realigned in registers, and the last words base by base
    text = T.synth("kind=reads_fq genome=20000 seed=41 cov=4 rl=150 sub=0.01 nrate=0.002")
    b, o = T.parse_buffer(text)
    pad = np.frombuffer(b"ACGTTGCAACGTTGCAAC", dtype=np.uint8)
    eng.set_plain_sequence(plain)
    try:
        for shift in range(18):
            n = len(b) - (shift % 7)  # also inputs that end inside a 16-byte word
            oo = np.minimum(o, np.uint64(n))
            d_b = eng.dev_upload(np.concatenate([pad[:shift], b, pad[:5]]))
            d_o = eng.dev_upload(oo)
            try:
                say(f"misaligned input: shift {shift}, {n} bases, plain={plain}")
                eng.count_device(T.ctypes.c_void_p(d_b.value + shift), n, d_o, len(oo) - 1, 31, 21, 16, 1)
                gk, gc = eng.result(31)
                ok, oc, _, _ = O.count(b[:n], oo, 31, 1)
                gi = np.argsort(gk[:, 0], kind="stable")
                oi = np.argsort(ok[:, 0], kind="stable")
                assert np.array_equal(gk[gi], ok[oi]) and np.array_equal(gc[gi], oc[oi]), shift
            finally:
                eng.dev_free(d_b)
                eng.dev_free(d_o)
    finally:
        eng.set_plain_sequence(False)


@pytest.mark.parametrize("name", sorted(GOLD["inputs"]))
def test_partition_records_match_reference(eng, name):
    text = GOLD["inputs"][name].encode()
    bases, off = T.parse_buffer(text)
    for key in sorted(GOLD["records"]):
        nm, k, m, n = key.split("|")
        if nm != name:
            continue
        k, m, n = int(k), int(m), int(n)
        g = GOLD["records"][key]
        say(f"records {name} k={k} m={m} n={n}")
        parts = eng.partitions(bases, off, k, m, n)
        for p in range(n):
            blob = bytes.fromhex(g["parts"][p])
            ref, q = [], 0
            while q < len(blob):
                L = blob[q]
                ref.append(blob[q:q + 2 + (L + 3) // 4])
                q += 2 + (L + 3) // 4
            assert sorted(parts[p]) == sorted(ref), (key, p)


@pytest.mark.parametrize("k,m", [(31, 21), (31, 23), (21, 15), (63, 21), (11, 5), (5, 3), (9, 8),
                                 (64, 31), (32, 16), (17, 2), (3, 2), (45, 9), (33, 15), (7, 1),
                                 (57, 6), (64, 1)])
def test_superkmers_match_oracle(eng, k, m):
    rng = random.Random(k * 100 + m)
    strings = [rs(rng, rng.randint(1, 400)) for _ in range(300)]
    strings += [rs(rng, 9000), "A" * 3000 + rs(rng, 50), "CAG" * 1500, rs(rng, 5000, "AC")]
    bases, off = O.chunks_of(strings)
    n = 16
    say(f"superkmers k={k} m={m}")
    gs, gl, gm, gp, nsp = eng.superkmers(bases, off, k, m, n)
    os_, ol, om, oh, km = O.superkmers(bases, off, k, m)
    log2n, log2s = _ilog2(n), _ilog2(nsp) - _ilog2(n)
    sb = 18 - _ilog2(nsp)
    oh = oh.astype(np.uint64)
    # sub-partition = reference partition bits, then hash bits above the split bits
    opid = ((oh & np.uint64(n - 1)) << np.uint64(log2s)) | ((oh >> np.uint64(log2n + sb)) & np.uint64((1 << log2s) - 1))
    a = sorted(zip(gs.tolist(), gl.tolist(), gm.tolist(), gp.tolist()))
    b = sorted(zip(os_.tolist(), ol.tolist(), om.tolist(), opid.astype(np.uint32).tolist()))
    assert len(a) == len(b)
    assert a == b


@pytest.mark.parametrize("k,m", [(31, 21), (21, 15), (63, 21), (15, 7), (5, 3)])
def test_head_strips_caps_and_pieces(eng, ceng, k, m):
    # every chunk fits one strip (at most 2048 k-mers), so every block takes
    # the head-strip decision: superkmers capped at 256 - k k-mers
    # (homopolymers and short tandem repeats longer than 255 bases), fat
    # pieces, chunks of exactly k bases and of fewer
    rng = random.Random(4000 + k)
    strings = ["A" * 700, "AC" * 400, rs(rng, 300) + "T" * 500 + rs(rng, 40), "CAG" * 300,
               rs(rng, 2000), "G" * k, "C" * (k - 1), rs(rng, k)]
    strings += [rs(rng, rng.randint(1, 260)) for _ in range(200)]
    rng.shuffle(strings)
    bases, off = O.chunks_of(strings)
    say(f"head strips k={k} m={m}")
    gs, gl, gm, gp, nsp = eng.superkmers(bases, off, k, m, 16)
    os_, ol, om, oh, km = O.superkmers(bases, off, k, m)
    assert sorted(zip(gs.tolist(), gl.tolist(), gm.tolist())) == sorted(zip(os_.tolist(), ol.tolist(), om.tolist()))
    st = check_counts(ceng, bases, off, k, m, 16, label="head")
    assert st.n_superkmers == len(os_)


@pytest.mark.parametrize("k,m", [(31, 21), (21, 15), (63, 21), (2, 1), (5, 3), (15, 7), (32, 16),
                                 (33, 15), (45, 9), (64, 31), (25, 11), (57, 6), (64, 1)])
def test_counts_match_oracle_random_chunks(ceng, k, m):
    rng = random.Random(1000 + k)
    strings = [rs(rng, rng.randint(1, 300)) for _ in range(500)] + [rs(rng, 20000)]
    bases, off = O.chunks_of(strings)
    check_counts(ceng, bases, off, k, m, 16, label="random")


def test_counts_low_complexity_and_caps(ceng):
    rng = random.Random(7)
    strings = ["A" * 7000, "AC" * 3000, rs(rng, 3000) + "T" * 2600 + rs(rng, 3000),
               "CAGT" * 2000, ("ACGGT" * 40 + rs(rng, 5)) * 30]
    bases, off = O.chunks_of(strings)
    for (k, m) in [(31, 21), (63, 21), (21, 15)]:
        check_counts(ceng, bases, off, k, m, 4, label="lowcomplexity")


def test_counts_high_multiplicity_and_filters(ceng):
    rng = random.Random(8)
    g = rs(rng, 5000)
    reads = []
    for _ in range(3000):
        p = rng.randint(0, len(g) - 150)
        reads.append(g[p:p + 150])
    bases, off = O.chunks_of(reads)
    check_counts(ceng, bases, off, 31, 21, 16, label="coverage")
    check_counts(ceng, bases, off, 31, 21, 16, ci=2, label="ci2")
    check_counts(ceng, bases, off, 31, 21, 16, ci=50, cx=100, label="ci50cx100")
    check_counts(ceng, bases, off, 63, 21, 16, ci=3, label="k63ci3")


def test_counts_edge_inputs(ceng):
    # all chunks shorter than k, single chunk of exactly k, empty input
    for strings, k in [(["ACGT", "ACG", "A"], 5), (["ACGTACGTAC"], 10), ([], 31), (["ACGTACGTACGTA"], 13)]:
        bases, off = O.chunks_of(strings) if strings else (np.zeros(0, np.uint8), np.zeros(1, np.uint64))
        check_counts(ceng, bases, off, k, min(3, k - 1), 16, label="edge")


def test_counts_fastq_synthetic_matches_oracle(ceng):
    text = T.synth("kind=reads_fq genome=200000 seed=2 cov=20 rl=150 sub=0.005 nrate=0.0001")
    bases, off = T.parse_buffer(text)
    ob, oo = O.parse(text)
    assert np.array_equal(bases, ob) and np.array_equal(off, oo)
    check_counts(ceng, bases, off, 31, 21, 256, ci=2, label="reads_fq ci2")
    check_counts(ceng, bases, off, 31, 21, 256, ci=1, label="reads_fq")


def test_counts_genome_c1_shape(ceng):
    text = T.synth("kind=genome_fa len=4600000 seed=1")
    bases, off = T.parse_buffer(text)
    st = check_counts(ceng, bases, off, 31, 21, 16, label="C1 genome")
    _, _, _, _, km = O.superkmers(bases, off, 31, 21)
    assert st.n_superkmers == len(O.superkmers(bases, off, 31, 21)[0])


def test_repeat_genome_reads(ceng):
    text = T.synth("kind=repeat_fq genome=300000 seed=3 gc=0.45 cov=5 rl=150 sub=0.003")
    bases, off = T.parse_buffer(text)
    check_counts(ceng, bases, off, 31, 23, 64, label="repeat reads")


def test_repeat_reads_c3_settings(ceng):
    # the 2 Gbp repetitive workload's generator and settings (k=31 m=23,
    # n=4096) on a 2.5 Mbp genome at 20x (50 Mbp of reads), oracle-checked
    bases, off = T.synth_parsed("kind=repeat_fq genome=2500000 seed=3 gc=0.45 cov=20 rl=150 sub=0.003 "
                                "fam=50 flen=300 div=0.05 rep=0.35 tan=0.10")
    st = check_counts(ceng, bases, off, 31, 23, 4096, label="repeat reads, C3 settings")
    assert st.n_parts == 4096


def test_split_buckets_in_record_ranges():
    # 10 Mbp of repetitive reads from a 1 Mbp genome with 2^20 bases per
    # bucket wanted: 16 sub-partitions of about 10^5 distinct keys each, far
    # more than one LDS table, so every sub-partition splits into many work
    # items, every coarse bucket is split-ordered, and the fine scatter moves
    # each in four record ranges
    e = T.Engine(0)
    try:
        e.set_batch_config(kmers_per_bucket=1 << 20)
        bases, off = T.synth_parsed("kind=repeat_fq genome=1000000 seed=3 gc=0.45 cov=10 rl=150 sub=0.003 "
                                    "fam=50 flen=300 div=0.05 rep=0.35 tan=0.10")
        st = check_counts(e, bases, off, 31, 23, 4096, label="split buckets in record ranges")
        say(f"  sub-partitions {st.n_subparts}, work items {st.n_lds_groups}")
        assert st.n_lds_groups > 4 * st.n_subparts
    finally:
        e.close()


def test_repeated_calls_reuse_buffers(ceng):
    rng = random.Random(9)
    for trial in range(3):
        strings = [rs(rng, rng.randint(50, 2000)) for _ in range(50 - 20 * trial)]
        bases, off = O.chunks_of(strings)
        check_counts(ceng, bases, off, 31, 21, 16, label=f"reuse{trial}")


@pytest.mark.parametrize("k,m,n", [(31, 21, 64), (63, 21, 16), (21, 15, 256)])
def test_debug_partition_stats(k, m, n):
    # per-partition statistics (the reference's -dbg table report): insertions
    # and distinct keys per partition add up to the job's totals, the counts
    # equal the oracle's with the statistics on, and none exist before
    e = T.Engine(0)
    try:
        e.set_plain_sequence(True)
        with pytest.raises(T.TunaGpuError, match="no debug statistics"):
            e.debug_partitions(n)
        e.set_debug_stats(True)
        text = T.synth("kind=reads_fq genome=60000 seed=7 cov=12 rl=150 sub=0.005 nrate=0.001")
        b, o = T.parse_buffer(text)
        say(f"debug stats k={k} m={m} n={n}")
        keys, cnt = e.count(b, o, k, m, n)
        st = e.last_stats
        d = e.debug_partitions(n)
        assert len(d["inserted"]) == n
        assert int(d["inserted"].sum()) == st.n_kmers
        assert int(d["lds_unique"].sum()) <= st.n_unique
        if int(d["spilled"].sum()) == 0:
            assert int(d["lds_unique"].sum()) == st.n_unique
        assert np.all(d["table_cap"] % 1024 == 0) and np.all(d["lds_unique"] <= d["table_cap"])
        ok, oc, tot, nu = O.count(b, o, k, 1, 2**64 - 1)
        assert st.n_unique == nu and st.n_kmers == tot
        with pytest.raises(T.TunaGpuError):
            e.debug_partitions(n // 2)
    finally:
        e.close()


@pytest.mark.parametrize("k,m", [(31, 21), (63, 21), (21, 15), (17, 2)])
def test_debug_min_coverage_matches_oracle(eng, k, m):
    # k-mers per distinct minimizer (the reference's -dbg coverage histogram),
    # grouped here by the oracle's minimizer hash of each superkmer
    text = T.synth("kind=reads_fq genome=50000 seed=8 cov=10 rl=150 sub=0.005 nrate=0.001")
    b, o = T.parse_buffer(text)
    b2 = b"".join([b">lc\n", b"A" * 2000, b"CA" * 800, b"\n"])
    lb, lo = T.parse_buffer(b2)
    for bases, off in ((b, o), (lb, lo)):
        say(f"min coverage k={k} m={m} bases={len(bases)}")
        cov, cnt = eng.debug_min_coverage(bases, off, k, m)
        st, ln, mp, hs, km = O.superkmers(bases, off, k, m)
        per = {}
        for h, l in zip(hs.tolist(), ln.tolist()):
            per[h] = per.get(h, 0) + l - k + 1
        hist = {}
        for v in per.values():
            hist[v] = hist.get(v, 0) + 1
        exp = sorted(hist.items())
        assert list(zip(cov.tolist(), cnt.tolist())) == exp
        assert int((cov.astype(np.int64) * cnt.astype(np.int64)).sum()) == km


@pytest.mark.parametrize("n", [32768, 1 << 18])
def test_partition_export_beyond_sub_partitions(eng, n):
    # more reference partitions than device sub-partitions (2^14): records are
    # grouped by H & (n-1) on the host; every partition's record bytes equal
    # the oracle's records of that partition (orc_pack_record, pinned to the
    # reference's writer by the golden vectors)
    rng = random.Random(n)
    strings = [rs(rng, rng.randint(40, 3000)) for _ in range(60)] + ["ACG" * 900]
    bases, off = O.chunks_of(strings)
    k, m = 31, 21
    say(f"partitions n={n}")
    parts = eng.partitions(bases, off, k, m, n)
    st, ln, mp, hs, _ = O.superkmers(bases, off, k, m)
    want = {}
    for s0, l0, m0, h0 in zip(st.tolist(), ln.tolist(), mp.tolist(), hs.tolist()):
        rec = O.pack_record(bytes(bases[s0:s0 + l0]), m0, k, m)
        want.setdefault(h0 & (n - 1), []).append(rec)
    got = {P: sorted(r) for P, r in enumerate(parts) if r}
    assert got == {P: sorted(r) for P, r in want.items()}


@pytest.mark.parametrize("k,m", [(31, 21), (31, 23), (21, 15), (17, 9), (25, 5), (32, 16)])
def test_tandem_repeats_and_homopolymers(ceng, k, m):
    # records with a short period (tandem units of 1-12 bases, homopolymers),
    # alone and as reads with substitutions, between random flanks: their
    # k-mers recur every period and are counted with weights
    rng = random.Random(7000 + k)
    strings = []
    for unit_len in range(1, 13):
        unit = rs(rng, unit_len)
        arr = unit * (600 // unit_len + 3)
        strings.append(rs(rng, 40) + arr + rs(rng, 40))
        for _ in range(30):
            a = rng.randint(0, len(arr) - 150)
            r = list(arr[a:a + 150])
            if rng.random() < 0.5:
                r[rng.randint(0, 149)] = rng.choice("ACGT")
            strings.append("".join(r))
    strings += ["A" * 500, "AC" * 300, "ACGTTGCA" * 80, rs(rng, 3000)]
    bases, off = O.chunks_of(strings)
    check_counts(ceng, bases, off, k, m, 16, label="tandem repeats")
    check_counts(ceng, bases, off, k, m, 16, ci=3, cx=40, label="tandem repeats, ci=3 cx=40")


REF_COUNTS = json.load(open(os.path.join(HERE, "golden", "reference_counts.json")))


def test_counts_match_reference_table(ceng):
    """Every golden input and (k, m, ci, cx) of the reference-table fixture:
    the device result, formatted as TSV on the device, sorted, has the number
    of lines, the count sum and the SHA-256 that the reference's own
    Streaming_Kmer_Hash_Table gave (tests/golden/make_golden.py)."""
    import hashlib
    for key in sorted(REF_COUNTS):
        name, k, m, ci, cx = key.split("|")
        k, m, ci, cx = int(k), int(m), int(ci), int(cx)
        b, off = T.parse_buffer(GOLD["inputs"][name].encode())
        st = ceng.count(b, off, k, m, 16, ci, cx, fetch=False)
        lines = sorted(ceng.result_tsv(k).splitlines(keepends=True))
        want = REF_COUNTS[key]
        got = (len(lines), sum(int(l.split(b"\t")[1]) for l in lines), hashlib.sha256(b"".join(lines)).hexdigest())
        assert st.n_out == want["n"] and got == (want["n"], want["sum"], want["sha256"]), key
