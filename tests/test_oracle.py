"""CPU tests that pin the oracle (oracle/oracle.c).

Each oracle function is checked against (a) vectors produced by the reference
code itself (tests/golden/reference_vectors.json, from oracle/_ref) and (b)
independent mathematical facts: closed forms, strand symmetry, brute force.
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))

COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def rc(s):
    return "".join(COMP[c] for c in reversed(s.upper()))


def rs(rng, n, alph="ACGT"):
    return "".join(rng.choice(alph) for _ in range(n))


# ---------------------------------------------------------------- A1 -------
@pytest.mark.parametrize("name", sorted(GOLD["parse"]))
def test_parse_matches_reference_parser(name):
    b, off = O.parse(GOLD["inputs"][name])
    got = [bytes(b[int(off[i]):int(off[i + 1])]).decode() for i in range(len(off) - 1)]
    assert got == GOLD["parse"][name]


def test_parse_empty_raises():
    with pytest.raises(ValueError):
        O.parse(b"")


def test_parse_fasta_rules_by_hand():
    b, off = O.parse(">h1\nACGT\nacgt\nNNAC\n>h2 x\nGG>in\nTT\n")
    got = [bytes(b[int(off[i]):int(off[i + 1])]).decode() for i in range(len(off) - 1)]
    assert got == ["ACGTacgt", "AC", "GG", "TT"]
    # no header at all: nothing is read
    b, off = O.parse("ACGTACGT\n")
    assert len(off) == 1


def test_parse_fastq_rules_by_hand():
    text = "@r1\nACGTNACG\n+\nIIIIIIII\n@r2\nacgt\n+\nIIII\n"
    b, off = O.parse(text)
    got = [bytes(b[int(off[i]):int(off[i + 1])]).decode() for i in range(len(off) - 1)]
    assert got == ["ACGT", "ACG", "acgt"]


# ---------------------------------------------------------------- A3 -------
@pytest.mark.parametrize("key", sorted(GOLD["hash"])[:40])
def test_nthash_matches_reference_roller(key):
    m, s = key.split("|")
    m = int(m)
    assert [int(x) for x in O.nthash_roll(s, m)] == GOLD["hash"][key]


def test_nthash_roll_equals_closed_form_and_is_strand_symmetric():
    rng = random.Random(1)
    for m in (1, 2, 5, 11, 21, 31, 32):
        s = rs(rng, 150)
        roll = [int(x) for x in O.nthash_roll(s, m)]
        direct = [O.nthash_direct(s[i:i + m], m) for i in range(len(s) - m + 1)]
        assert roll == direct
        for i in range(0, len(s) - m + 1, 7):
            assert O.nthash_direct(s[i:i + m], m) == O.nthash_direct(rc(s[i:i + m]), m)


def test_nthash_single_base_closed_form():
    # m = 1: fwd = seed[b], rev = seed[b^2]; canonical of A equals that of T
    seeds = [0x3c8bfbb395c60474, 0x3193c18562a02b4c, 0x295549f54be24456, 0x20323ed082572324]
    nt = {"A": 0, "C": 1, "T": 2, "G": 3}
    for c in "ACGT":
        assert O.nthash_direct(c, 1) == seeds[nt[c]] ^ seeds[nt[c] ^ 2]


# ---------------------------------------------------------------- A4 -------
@pytest.mark.parametrize("key", sorted(GOLD["window"])[::3])
def test_window_matches_reference_minimizer_window(key):
    k, m, s = key.split("|")
    k, m = int(k), int(m)
    ref = np.array(GOLD["window"][key], dtype=np.uint64).reshape(-1, 2)
    H, P = O.window(O.nthash_roll(s, m), k - m + 1)
    assert np.array_equal(H, ref[:, 0]) and np.array_equal(P.astype(np.uint64), ref[:, 1])


def test_window_stream_equals_brute_rule_on_ties():
    rng = random.Random(3)
    for W in (1, 2, 3, 7, 11, 16, 43):
        for trial in range(20):
            M = rng.randint(W, 200)
            h = np.array([rng.choice([5, 7, 9, 2**64 - 2]) for _ in range(M)], dtype=np.uint64)
            assert all(np.array_equal(x, y) for x, y in zip(O.window(h, W), O.window(h, W, True)))


def test_window_brute_is_a_true_minimum():
    rng = random.Random(4)
    h = np.array([rng.randrange(0, 50) for _ in range(300)], dtype=np.uint64)
    for W in (3, 11):
        H, P = O.window(h, W, True)
        for t in range(len(H)):
            assert H[t] == h[t:t + W].min() and h[P[t]] == H[t] and t <= P[t] < t + W


def test_planted_tie_mistake_is_caught():
    """A window that prefers the older block on ties disagrees with the reference."""
    key = next(k for k in sorted(GOLD["window"]) if k.split("|")[2].startswith("CAG"))
    k, m, s = key.split("|")
    k, m = int(k), int(m)
    W = k - m + 1
    h = O.nthash_roll(s, m)
    ref = np.array(GOLD["window"][key], dtype=np.uint64).reshape(-1, 2)
    wrong = []
    for t in range(len(h) - W + 1):
        seg = h[t:t + W]
        wrong.append(t + int(np.argmin(seg)))  # leftmost everywhere
    assert not np.array_equal(np.array(wrong, np.uint64), ref[:, 1])


# -------------------------------------------------------------- A5/A6 ------
def _oracle_parts(text, k, m, n):
    b, off = O.parse(text)
    st, ln, mp, hs, km = O.superkmers(b, off, k, m)
    parts = [[] for _ in range(n)]
    for i in range(len(st)):
        s = bytes(b[int(st[i]):int(st[i]) + int(ln[i])])
        parts[int(hs[i]) & (n - 1)].append(O.pack_record(s, int(mp[i]), k, m))
    return parts, km, len(st)


def _split_records(blob, k, m):
    recs, q = [], 0
    wide = 2 * k - m > 255
    while q < len(blob):
        L = int.from_bytes(blob[q:q + 2], "little") if wide else blob[q]
        h = 4 if wide else 2
        recs.append(blob[q:q + h + (L + 3) // 4])
        q += h + (L + 3) // 4
    return recs


@pytest.mark.parametrize("key", sorted(GOLD["records"]))
def test_superkmer_records_match_reference(key):
    name, k, m, n = key.split("|")
    k, m, n = int(k), int(m), int(n)
    g = GOLD["records"][key]
    parts, km, nsk = _oracle_parts(GOLD["inputs"][name], k, m, n)
    assert km == g["kmers"] and nsk == g["superkmers"]
    for p in range(n):
        assert sorted(parts[p]) == sorted(_split_records(bytes.fromhex(g["parts"][p]), k, m))


def test_superkmer_invariants():
    rng = random.Random(5)
    strings = [rs(rng, rng.randint(1, 400)) for _ in range(40)] + ["A" * 700, "AC" * 400]
    b, off = O.chunks_of(strings)
    for (k, m) in [(31, 21), (21, 15), (63, 21), (5, 3)]:
        st, ln, mp, hs, km = O.superkmers(b, off, k, m)
        K = sum(max(len(s) - k + 1, 0) for s in strings)
        assert km == K == int((ln.astype(np.int64) - k + 1).sum())
        assert (ln <= 255).all() and (ln >= k).all() and (mp <= k - m).all()
        # every superkmer's minimizer is the window minimum of its first k-mer
        for i in range(0, len(st), 11):
            s = bytes(b[int(st[i]):int(st[i]) + k]).decode()
            hh = [O.nthash_direct(s[j:j + m], m) for j in range(k - m + 1)]
            assert min(hh) == int(hs[i]) and hh[int(mp[i])] == int(hs[i])


# ------------------------------------------------------------ A8/A9/A10 ----
def _brute_counts(strings, k):
    d = {}
    for s in strings:
        s = s.upper()
        for i in range(len(s) - k + 1):
            x = s[i:i + k]
            c = min(x, rc(x))
            d[c] = d.get(c, 0) + 1
    return d


def _key_to_label(row, k):
    key = sum(int(row[i]) << (64 * i) for i in range(len(row)))
    return "".join("ACGT"[(key >> (2 * (k - 1 - i))) & 3] for i in range(k))


@pytest.mark.parametrize("k", [2, 5, 15, 31, 32, 33, 45, 63, 64, 65, 96, 97, 127, 128, 200, 256])
def test_count_equals_python_brute_force(k):
    rng = random.Random(k)
    strings = [rs(rng, rng.randint(1, 300)) for _ in range(30)] + ["acgt" * 30, "T" * 100]
    b, off = O.chunks_of(strings)
    keys, cnt, tot, nu = O.count(b, off, k)
    brute = _brute_counts(strings, k)
    assert tot == sum(brute.values()) and nu == len(brute)
    got = {_key_to_label(keys[i], k): int(cnt[i]) for i in range(len(cnt))}
    assert got == brute


def test_count_filters():
    strings = ["ACGTACGTAAAAAAAAAAAA", "ACGTACGTCC"]
    b, off = O.chunks_of(strings)
    k = 4
    brute = _brute_counts(strings, k)
    keys, cnt, _, _ = O.count(b, off, k, ci=2, cx=5)
    got = {_key_to_label(keys[i], k): int(cnt[i]) for i in range(len(cnt))}
    assert got == {s: c for s, c in brute.items() if 2 <= c <= 5}


def test_canonical_invariants():
    rng = random.Random(6)
    for k in (1, 7, 31, 32, 33, 63, 64, 65, 127, 128, 129, 256):
        for _ in range(20):
            s = rs(rng, k)
            assert O.canonical(s, k) == O.canonical(rc(s), k)
            lab = O.tsv_line(O.canonical(s, k), k, 3).split("\t")[0]
            assert lab == min(s, rc(s))


def test_tsv_line_format():
    assert O.tsv_line(0b00011011, 4, 12) == "ACGT\t12\n"


def test_round_pow2():
    assert [O.round_pow2(x) for x in (0, 1, 2, 3, 5, 16, 17, 1000)] == [1, 1, 2, 4, 8, 16, 32, 1024]


@pytest.mark.parametrize("k,m,n,ci,cx,threads", [(31, 21, 16, 1, 2**64 - 1, 1), (31, 21, 256, 2, 2**64 - 1, 4),
                                                 (21, 15, 64, 1, 3, 3), (63, 21, 32, 2, 2**64 - 1, 4),
                                                 (33, 9, 8, 1, 2**64 - 1, 2), (5, 3, 4, 1, 2**64 - 1, 2)])
def test_parallel_count_equals_brute_force(k, m, n, ci, cx, threads):
    # the multithreaded two-phase CPU count (the CPU baseline's fallback) has
    # the same multiset as the sort-and-count, and the superkmer totals of
    # orc_superkmers, whatever the partitions and threads
    import tuna_amd as T
    b, o = T.synth_parsed("kind=reads_fq genome=30000 seed=9 cov=6 rl=150 sub=0.01 nrate=0.002")
    rng = random.Random(k)
    g = ">g\n" + "".join(rng.choice("ACGT") for _ in range(5000)) + "A" * 400 + "\n"
    gb, go = O.parse(g.encode())
    b = np.concatenate([b, gb])
    o = np.concatenate([o, go[1:] + o[-1]]).astype(np.uint64)
    pk, pc, info = O.count_parallel(b, o, k, m, n, ci, cx, threads)
    ok, oc, tot, nu = O.count(b, o, k, ci, cx)
    assert info["kmers"] == tot and info["unique"] == nu and info["out"] == len(oc)
    assert info["superkmers"] == len(O.superkmers(b, o, k, m)[0])
    sk = np.lexsort(tuple(pk[:, i] for i in range(pk.shape[1])))
    so = np.lexsort(tuple(ok[:, i] for i in range(ok.shape[1])))
    assert np.array_equal(pk[sk], ok[so]) and np.array_equal(pc[sk], oc[so])


REF_COUNTS = json.load(open(os.path.join(HERE, "golden", "reference_counts.json")))


@pytest.mark.parametrize("key", sorted(REF_COUNTS))
def test_count_matches_reference_table(key):
    """orc_count against the reference's own counting table: the fixture holds,
    for each input and (k, m, ci, cx), the number of reported k-mers, the sum of
    their counts and the SHA-256 of the sorted TSV lines that the reference's
    Streaming_Kmer_Hash_Table produced over the reference's partition buffers
    (tests/golden/make_golden.py)."""
    name, k, m, ci, cx = key.split("|")
    k, ci, cx = int(k), int(ci), int(cx)
    b, off = O.parse(GOLD["inputs"][name].encode())
    keys, cnt, _, _ = O.count(b, off, k, ci, cx)
    lines = sorted(O.tsv_line(int(keys[i, 0]) | (int(keys[i, 1]) << 64 if keys.shape[1] > 1 else 0), k,
                              int(cnt[i])).encode() for i in range(len(cnt)))
    want = REF_COUNTS[key]
    assert len(lines) == want["n"] and int(cnt.sum()) == want["sum"]
    assert hashlib.sha256(b"".join(lines)).hexdigest() == want["sha256"]
