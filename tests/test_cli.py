"""The `tuna` CLI and the C++ library header (include/tuna/tuna.hpp).

CPU tests: argument validation (exit codes and messages of src/CLI.cpp),
output formatters, gzip input.  GPU tests: end-to-end counts against the
oracle, KFF output, @list input, -tp, and -kt partition files against the
reference's own partition bytes (tests/golden)."""
import gzip
import json
import os
import random
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
import tuna_amd as T

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLI = os.path.join(ROOT, "tuna_amd", "tuna")
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))


def run_cli(*args, cwd=None):
    return subprocess.run([CLI, *args], capture_output=True, text=True, cwd=cwd, timeout=300)


def rs(rng, n, alph="ACGT"):
    return "".join(rng.choice(alph) for _ in range(n))


# ------------------------------------------------------------------ CPU ----
@pytest.mark.parametrize("args,msg", [
    (["-k", "300", "a", "b"], "k-mer length (-k 300) must be in [2, 256]"),
    (["-k", "31", "-m", "31", "a", "b"], "must be strictly less than k-mer length"),
    (["-m", "40", "a", "b"], "minimizer length (-m 40) must be in [1, 32]"),
    (["-t", "0", "a", "b"], "number of threads (-t) must be > 0"),
    (["-x", "a", "b"], "unknown option: -x"),
    (["only_one"], "need at least one input file and an output file"),
    (["-k", "abc", "a", "b"], "invalid integer for -k: abc"),
    (["-k"], "-k requires an argument"),
    (["@/nonexistent/list", "out"], "cannot open input list"),
])
def test_cli_rejects_bad_arguments(args, msg):
    r = run_cli(*args)
    assert r.returncode == 1
    assert "tuna: error: " in r.stderr and msg in r.stderr


def test_cli_help_and_version():
    r = run_cli("--help")
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = run_cli("-v")
    assert r.returncode == 0 and r.stderr.startswith("tuna 1.4.0")


def test_cli_k_above_64_without_partition_files(tmp_path):
    # k > 64 is counted without superkmer records: -tp, -kt and -dbg are refused
    # before any device work; k above 256 is refused by the parser
    f = tmp_path / "x.fa"
    f.write_text(">a\nACGT\n")
    for flag in ("-tp", "-kt", "-dbg"):
        r = run_cli("-k", "127", "-m", "21", flag, str(f), str(tmp_path / "o.tsv"))
        assert r.returncode == 1 and "unsupported combination k=127 with -tp, -kt or -dbg" in r.stderr, flag
    r = run_cli("-k", "257", "-m", "21", str(f), str(tmp_path / "o.tsv"))
    assert r.returncode == 1 and "tuna: error:" in r.stderr


def test_tsv_formatter_matches_oracle():
    rng = random.Random(3)
    for k in (5, 31, 32, 33, 63, 64, 65, 96, 127, 128, 200, 256):
        nw = (k + 31) // 32
        keys, counts, lines = [], [], []
        for _ in range(50):
            key = rng.getrandbits(2 * k)
            c = rng.randint(1, 10**9)
            keys.append([(key >> (64 * i)) & (2**64 - 1) for i in range(nw)])
            counts.append(c)
            lines.append(O.tsv_line(key, k, c))
        ka = np.array(keys, dtype=np.uint64).reshape(-1)
        ca = np.array(counts, dtype=np.uint32)
        buf = np.zeros(50 * (k + 12), dtype=np.uint8)
        n = T.lib().tg_format_tsv(k, ka.ctypes.data_as(T._u64p), ca.ctypes.data_as(T._u32p), 50,
                                  buf.ctypes.data_as(T.ctypes.c_char_p), len(buf))
        assert bytes(buf[:n]).decode() == "".join(lines)


def parse_kff(path):
    """Minimal KFF reader for files written by tg_write_kff."""
    d = open(path, "rb").read()
    assert d[:3] == b"KFF" and d[-3:] == b"KFF"
    assert d[5] == 0x1B and d[6] == 1 and d[7] == 1
    p = 8
    msz = struct.unpack(">I", d[p:p + 4])[0]
    p += 4 + msz
    assert d[p:p + 1] == b"v"
    nv = struct.unpack(">Q", d[p + 1:p + 9])[0]
    p += 9
    var = {}
    for _ in range(nv):
        e = d.index(b"\0", p)
        var[d[p:e].decode()] = struct.unpack(">Q", d[e + 1:e + 9])[0]
        p = e + 9
    assert d[p:p + 1] == b"r"
    nb = struct.unpack(">Q", d[p + 1:p + 9])[0]
    p += 9
    k = var["k"]
    kb = (k + 3) // 4
    out = {}
    for _ in range(nb):
        seq = d[p:p + kb]
        cnt = struct.unpack(">I", d[p + kb:p + kb + 4])[0]
        p += kb + 4
        lab = "".join("ACGT"[(seq[i // 4] >> (6 - 2 * (i % 4))) & 3] for i in range(k))
        out[lab] = cnt
    assert var["max"] == 1 and var["data_size"] == 4
    return k, out


def test_kff_writer_round_trip(tmp_path):
    rng = random.Random(4)
    for k in (7, 31, 33, 63):
        nw = 2 if k > 32 else 1
        exp, keys, counts = {}, [], []
        for _ in range(40):
            key = rng.getrandbits(2 * k)
            c = rng.randint(1, 2**32 - 1)
            keys.append([key & (2**64 - 1), key >> 64][:nw])
            counts.append(c)
            exp[O.tsv_line(key, k, 1).split("\t")[0]] = c
        ka = np.array(keys, dtype=np.uint64).reshape(-1)
        ca = np.array(counts, dtype=np.uint32)
        path = str(tmp_path / f"o{k}.kff")
        assert T.lib().tg_write_kff(path.encode(), k, ka.ctypes.data_as(T._u64p),
                                    ca.ctypes.data_as(T._u32p), 40) == 0
        kk, got = parse_kff(path)
        assert kk == k and got == exp


def test_gzip_input_reads_like_plain(tmp_path):
    text = GOLD["inputs"]["fastq_reads"].encode()
    p = tmp_path / "r.fq.gz"
    with gzip.open(p, "wb") as f:
        f.write(text)
    assert T.read_file(p) == text
    assert np.array_equal(T.parse_file(p)[1], T.parse_buffer(text)[1])


def test_empty_and_missing_inputs_report_reference_messages(tmp_path):
    e = tmp_path / "empty.fa"
    e.write_bytes(b"")
    with pytest.raises(T.TunaGpuError, match=f"Empty file: {e}"):
        T.read_file(e)
    m = tmp_path / "missing.fa"
    with pytest.raises(T.TunaGpuError, match=f"Cannot open file: {m}"):
        T.read_file(m)


def test_mapped_input_equals_read_input(tmp_path):
    data = T.synth("kind=reads_fq genome=20000 seed=4 reads=300 rl=150 sub=0.01 nrate=0.001").tobytes()
    p = tmp_path / "r.fq"
    p.write_bytes(data)
    L = T.lib()
    d, n = T._u8p(), T.ctypes.c_uint64()
    assert L.tg_map_file(str(p).encode(), T.ctypes.byref(d), T.ctypes.byref(n)) == 0
    assert n.value == len(data) and T.ctypes.string_at(d, n.value) == data
    L.tg_unmap_file(d, n)
    for name, content, msg in (("missing.fq", None, "Cannot open file"), ("empty.fq", b"", "Empty file")):
        q = tmp_path / name
        if content is not None:
            q.write_bytes(content)
        assert L.tg_map_file(str(q).encode(), T.ctypes.byref(d), T.ctypes.byref(n)) == -5
        assert L.tg_last_error().decode() == f"{msg}: {q}"


def test_auto_partitions_bounds(tmp_path):
    f = tmp_path / "g.fa"
    f.write_text(">a\n" + "ACGT" * 1000000 + "\n")
    arr = (T.ctypes.c_char_p * 1)(str(f).encode())
    T.lib().tg_auto_partitions.restype = T.ctypes.c_uint32
    n = T.lib().tg_auto_partitions(arr, 1, 31, 1)
    assert 16 <= n <= 32768 and n & (n - 1) == 0


# ------------------------------------------------------------------ GPU ----
def oracle_tsv(files_text, k, ci=1, cx=2**64 - 1):
    bs, offs, base = [], [], 0
    for t in files_text:
        b, o = O.parse(t)
        bs.append(b)
        offs.append(o[1:] + base)
        base += len(b)
    bases = np.concatenate(bs)
    off = np.concatenate([np.zeros(1, np.uint64)] + offs)
    keys, cnt, _, _ = O.count(bases, off, k, ci, cx)
    out = {}
    for i in range(len(cnt)):
        key = sum(int(keys[i, j]) << (64 * j) for j in range(keys.shape[1]))
        out[O.tsv_line(key, k, 1).split("\t")[0]] = int(cnt[i])
    return out


def read_tsv(path):
    out = {}
    for line in open(path):
        a, b = line.rstrip("\n").split("\t")
        assert a not in out
        out[a] = int(b)
    return out


@pytest.mark.gpu
def test_cli_counts_match_oracle(tmp_path):
    text = T.synth("kind=reads_fq genome=40000 seed=2 cov=15 rl=150 sub=0.005 nrate=0.0002")
    fq = tmp_path / "r.fq"
    fq.write_bytes(bytes(text))
    out = tmp_path / "o.tsv"
    r = run_cli("-k", "31", "-m", "21", "-n", "64", "-ci", "2", str(fq), str(out))
    assert r.returncode == 0, r.stderr
    exp = oracle_tsv([bytes(text)], 31, 2)
    assert read_tsv(out) == exp
    stats = dict(l.split(": ") for l in r.stderr.splitlines() if ": " in l and not l.startswith("tuna"))
    assert int(stats["unique_kmers"]) == len(exp) and int(stats["n_parts"]) == 64
    b, o = O.parse(bytes(text))
    assert int(stats["superkmers"]) == len(O.superkmers(b, o, 31, 21)[0])


@pytest.mark.gpu
@pytest.mark.parametrize("k,m", [(127, 21), (200, 31)])
def test_cli_counts_large_k_match_oracle(tmp_path, k, m):
    # k > 64 (the reference's default dispatch includes k = 127): TSV and KFF
    text = T.synth("kind=reads_fq genome=30000 seed=4 cov=12 rl=250 sub=0.004 nrate=0.0002")
    fq = tmp_path / "r.fq"
    fq.write_bytes(bytes(text))
    out = tmp_path / "o.tsv"
    r = run_cli("-k", str(k), "-m", str(m), "-ci", "2", str(fq), str(out))
    assert r.returncode == 0, r.stderr
    exp = oracle_tsv([bytes(text)], k, 2)
    assert read_tsv(out) == exp
    stats = dict(l.split(": ") for l in r.stderr.splitlines() if ": " in l and not l.startswith("tuna"))
    b, o = O.parse(bytes(text))
    assert int(stats["unique_kmers"]) == len(exp)
    assert int(stats["superkmers"]) == len(O.superkmers(b, o, k, m)[0])
    kff = tmp_path / "o.kff"
    r = run_cli("-k", str(k), "-m", str(m), "-ci", "2", str(fq), str(kff))
    assert r.returncode == 0, r.stderr
    kk, recs = parse_kff(kff)
    assert kk == k and recs == exp


@pytest.mark.gpu
def test_cli_debug_table_stats(tmp_path):
    # -dbg: the reference's per-partition table report (src/count.hpp:478-565),
    # with per-partition insertions and distinct keys pinned to the oracle's
    # superkmers (a k-mer's partition is its minimizer's), counts unchanged
    text = bytes(T.synth("kind=reads_fq genome=40000 seed=3 cov=15 rl=150 sub=0.005 nrate=0"))
    fq = tmp_path / "r.fq"
    fq.write_bytes(text)
    out, work = tmp_path / "o.tsv", tmp_path / "work"
    n, k, m = 16, 31, 21
    r = run_cli("-k", str(k), "-m", str(m), "-n", str(n), "-dbg", "-w", str(work), str(fq), str(out))
    assert r.returncode == 0, r.stderr
    assert read_tsv(out) == oracle_tsv([text], k)
    err = r.stderr
    for line in ("[debug] per-partition table stats (first 20):", "[debug] aggregate table stats (16 partitions):",
                 "dbg_load_mean: ", "dbg_unique_mean: ", "dbg_n_resizes: 0", "[debug] resize summary:",
                 "[debug] minimizer coverage (k-mers sharing one minimizer):"):
        assert line in err, line
    cov_rows = open(work / "debug_min_coverage.csv").read().splitlines()
    assert cov_rows[0] == "coverage,n_minimizers,total_kmers"
    assert sum(int(r.split(",")[2]) for r in cov_rows[1:]) == int(
        next(l for l in err.splitlines() if l.startswith("  total k-mers covered")).split(":")[1])
    rows = [l.split(",") for l in open(work / "debug_table_stats.csv").read().splitlines()]
    assert rows[0] == ["partition_id", "init_sz", "table_cap", "n_inserted", "n_unique", "load_factor",
                       "n_resizes", "resize_s", "n_overflow"]
    rows = rows[1:]
    assert [int(x[0]) for x in rows] == list(range(n))
    b, o = O.parse(text)
    st, ln, mp, hs, km = O.superkmers(b, o, k, m)
    part = (hs.astype(np.uint64) & np.uint64(n - 1)).astype(np.int64)
    for p_ in range(n):
        sel = part == p_
        ins = int((ln[sel].astype(np.int64) - k + 1).sum())
        # the partition's distinct k-mers: count its superkmers as chunks
        starts = st[sel].astype(np.uint64)
        sb = np.concatenate([b[int(s0):int(s0) + int(l0)] for s0, l0 in zip(starts, ln[sel])]) if len(starts) else b[:0]
        offs = np.concatenate([[0], np.cumsum(ln[sel].astype(np.uint64))]).astype(np.uint64)
        _, _, tot, nu = O.count(sb, offs, k, 1, 2**64 - 1)
        assert int(rows[p_][3]) == ins == tot, p_
        assert int(rows[p_][8]) == 0
        assert int(rows[p_][4]) == nu, p_
        assert int(rows[p_][1]) == int(rows[p_][2]) and int(rows[p_][2]) % 4096 == 0


@pytest.mark.gpu
def test_cli_kff_list_gzip_and_multi_file(tmp_path):
    rng = random.Random(6)
    texts = [">a\n" + rs(rng, 3000) + "\n", ">b\n" + rs(rng, 2000) + "N" + rs(rng, 500) + "\n"]
    paths = []
    for i, t in enumerate(texts):
        p = tmp_path / f"f{i}.fa.gz"
        with gzip.open(p, "wb") as f:
            f.write(t.encode())
        paths.append(str(p))
    lst = tmp_path / "list.txt"
    lst.write_text("\n".join(paths) + "\n")
    out = tmp_path / "o.kff"
    r = run_cli("-k", "25", "-m", "11", "-hp", "@" + str(lst), str(out))
    assert r.returncode == 0, r.stderr
    k, got = parse_kff(str(out))
    assert k == 25 and got == oracle_tsv([t.encode() for t in texts], 25)


@pytest.mark.gpu
def test_cli_partition_only_and_kept_partition_files(tmp_path):
    name = "homopolymer_cap"
    f = tmp_path / "h.fa"
    f.write_text(GOLD["inputs"][name])
    wd = tmp_path / "wd"
    r = run_cli("-k", "31", "-m", "21", "-n", "16", "-tp", "-kt", "-w", str(wd), str(f), str(tmp_path / "o.tsv"))
    assert r.returncode == 0, r.stderr
    g = GOLD["records"][f"{name}|31|21|16"]
    assert f"superkmers: {g['superkmers']}" in r.stderr
    for p in range(16):
        blob = (wd / f"hash_{p}.superkmers").read_bytes()
        mine, q = [], 0
        while q < len(blob):
            L = blob[q]
            mine.append(blob[q:q + 2 + (L + 3) // 4])
            q += 2 + (L + 3) // 4
        ref, q = [], 0
        rb = bytes.fromhex(g["parts"][p])
        while q < len(rb):
            L = rb[q]
            ref.append(rb[q:q + 2 + (L + 3) // 4])
            q += 2 + (L + 3) // 4
        assert sorted(mine) == sorted(ref)


@pytest.mark.gpu
def test_cli_empty_input_error(tmp_path):
    e = tmp_path / "e.fa"
    e.write_bytes(b"")
    r = run_cli(str(e), str(tmp_path / "o.tsv"))
    assert r.returncode == 1 and f"tuna: error: Empty file: {e}" in r.stderr


@pytest.mark.gpu
def test_cli_unwritable_output_reports_reference_message(tmp_path):
    f = tmp_path / "g.fa"
    f.write_text(">g\n" + "ACGTTGCA" * 50 + "\n")
    for out in (tmp_path / "no_dir" / "o.tsv", tmp_path / "no_dir" / "o.kff"):
        r = run_cli("-k", "21", "-m", "11", str(f), str(out))
        # src/pipeline.hpp:151
        assert r.returncode == 1 and f"tuna: error: cannot open output file: {out}" in r.stderr


CPP_TEST = r'''
#include "tuna/tuna.hpp"
#include <cstdio>
#include <map>
#include <thread>
#include <unordered_map>
int main(int argc, char** argv) {
    std::vector<std::string> files{argv[1]}, other{argv[2]};
    tuna::Options o; o.min_count = 2; o.partitions = 32;
    auto m = tuna::count_to<31, 21>(files, o);
    auto raw = tuna::count_to_raw<31, 21>(files, o);
    std::map<std::string, uint32_t> sorted(m.begin(), m.end());
    if (raw.size() != m.size()) { fprintf(stderr, "size mismatch\n"); return 2; }
    // the packed k-mers: labels, canonical form and a hashed container
    std::unordered_map<tuna::Kmer<31>, uint32_t, tuna::Kmer_Hasher<31>> by_kmer;
    for (auto& kv : raw) {
        std::string s; kv.first.get_label(s);
        if (m.at(s) != kv.second) return 3;
        if (kv.first.canonical() != kv.first || kv.first > kv.first.reverse_complement()) return 5;
        if (tuna::Kmer<31>(s) != kv.first || kv.first.as_int() != tuna::Kmer<31>(s).as_int()) return 6;
        by_kmer[kv.first] = kv.second;
    }
    if (by_kmer.size() != raw.size()) return 7;
    for (auto& kv : m) {
        tuna::Kmer<31> x(kv.first);
        if (by_kmer.at(x.reverse_complement().canonical()) != kv.second) return 8;
    }
    for (auto& kv : sorted) printf("%s\t%u\n", kv.first.c_str(), kv.second);
    // two threads counting different inputs at once: each result equals the
    // same count made alone
    auto alone_a = m;
    auto alone_b = tuna::count_to<31, 21>(other, o);
    std::unordered_map<std::string, uint32_t> got_a, got_b;
    for (int rep = 0; rep < 3; ++rep) {
        std::thread ta([&] { got_a = tuna::count_to<31, 21>(files, o); });
        std::thread tb([&] { got_b = tuna::count_to<31, 21>(other, o); });
        ta.join(); tb.join();
        if (got_a != alone_a || got_b != alone_b) { fprintf(stderr, "threaded mismatch\n"); return 9; }
    }
    // k > 64: keys of four words
    auto m127 = tuna::count_to<127, 21>(files, o);
    auto raw127 = tuna::count_to_raw<127, 21>(files, o);
    if (m127.size() != raw127.size() || m127.empty()) return 10;
    std::unordered_map<tuna::Kmer<127>, uint32_t, tuna::Kmer_Hasher<127>> by127;
    for (auto& kv : raw127) {
        std::string s; kv.first.get_label(s);
        if (s.size() != 127 || m127.at(s) != kv.second) return 11;
        if (kv.first.canonical() != kv.first || tuna::Kmer<127>(s) != kv.first) return 12;
        by127[kv.first] = kv.second;
    }
    if (by127.size() != raw127.size()) return 13;
    fprintf(stderr, "k127 %zu\n", m127.size());
    try { tuna::count<31>({}, [](std::string_view, uint32_t) {}); return 4; }
    catch (const std::invalid_argument& e) { fprintf(stderr, "%s\n", e.what()); }
    return 0;
}
'''


@pytest.mark.gpu
def test_cpp_api_count_to_and_raw(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(CPP_TEST)
    exe = tmp_path / "t"
    lib = os.path.join(ROOT, "tuna_amd")
    c = subprocess.run(["g++", "-std=c++17", "-O1", str(src), "-I", os.path.join(ROOT, "include"),
                        "-L", lib, "-ltuna_gpu", "-Wl,-rpath," + lib, "-pthread", "-o", str(exe)],
                       capture_output=True, text=True)
    assert c.returncode == 0, c.stderr
    text = T.synth("kind=reads_fq genome=20000 seed=7 cov=10 rl=150 sub=0.005")
    fq = tmp_path / "r.fq"
    fq.write_bytes(bytes(text))
    fq2 = tmp_path / "r2.fq"
    fq2.write_bytes(bytes(T.synth("kind=reads_fq genome=30000 seed=8 cov=12 rl=150 sub=0.005")))
    r = subprocess.run([str(exe), str(fq), str(fq2)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "tuna::count: no input files" in r.stderr
    got = {}
    for line in r.stdout.splitlines():
        a, b = line.split("\t")
        got[a] = int(b)
    assert got == oracle_tsv([bytes(text)], 31, 2)
    assert f"k127 {len(oracle_tsv([bytes(text)], 127, 2))}" in r.stderr


@pytest.mark.gpu
def test_cli_partition_count_rounding_and_auto(tmp_path):
    # -n is rounded up to a power of two (src/main.cpp:47-52); without -n the
    # auto-tuned count of this host is used; counts are the same either way
    text = T.synth("kind=reads_fq genome=20000 seed=9 cov=8 rl=150 sub=0.005")
    fq = tmp_path / "r.fq"
    fq.write_bytes(bytes(text))
    exp = oracle_tsv([bytes(text)], 31, 1)
    r = run_cli("-n", "100", str(fq), str(tmp_path / "a.tsv"))
    assert r.returncode == 0, r.stderr
    stats = dict(l.split(": ") for l in r.stderr.splitlines() if ": " in l and not l.startswith("tuna"))
    assert int(stats["n_parts"]) == 128 and "n=128" in r.stderr
    assert read_tsv(tmp_path / "a.tsv") == exp
    arr = (T.ctypes.c_char_p * 1)(str(fq).encode())
    T.lib().tg_auto_partitions.restype = T.ctypes.c_uint32
    auto = O.round_pow2(T.lib().tg_auto_partitions(arr, 1, 31, 1))
    r = run_cli(str(fq), str(tmp_path / "b.tsv"))
    assert r.returncode == 0, r.stderr
    stats = dict(l.split(": ") for l in r.stderr.splitlines() if ": " in l and not l.startswith("tuna"))
    assert int(stats["n_parts"]) == auto
    assert read_tsv(tmp_path / "b.tsv") == exp
