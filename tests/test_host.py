"""CPU tests of the product library's host side: the C ABI loads and exports
every declared symbol, the host chunker matches the reference parser, and the
synthetic generators are deterministic.  No GPU calls."""
import json
import os
import random
import re

import numpy as np
import pytest

import oracle as O
import tuna_amd as T

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "tuna_gpu.h")).read()
    declared = sorted(set(re.findall(r"\b(tg_[a-z_0-9]+)\s*\(", hdr)))
    assert declared, "no declarations found"
    lib = T.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(T.EXPORTS) == declared


@pytest.mark.parametrize("name", sorted(GOLD["parse"]))
def test_host_parser_matches_reference(name):
    b, off = T.parse_buffer(GOLD["inputs"][name].encode())
    got = [bytes(b[int(off[i]):int(off[i + 1])]).decode() for i in range(len(off) - 1)]
    assert got == GOLD["parse"][name]


def _random_text(rng):
    alph = "ACGTacgtNRY \r\n>"
    if rng.random() < 0.5:
        recs = []
        for i in range(rng.randint(1, 6)):
            seq = "".join(rng.choice(alph) if rng.random() < 0.08 else rng.choice("ACGT")
                          for _ in range(rng.randint(0, 300)))
            recs.append(">r%d\n%s\n" % (i, "\n".join(seq[j:j + 50] for j in range(0, len(seq), 50))))
        return "".join(recs)
    recs = []
    for i in range(rng.randint(1, 8)):
        seq = "".join(rng.choice("ACGTN") for _ in range(rng.randint(0, 200)))
        recs.append("@r%d\n%s\n+\n%s\n" % (i, seq, "?" * len(seq)))
    t = "".join(recs)
    if rng.random() < 0.3:
        t = t.rstrip("\n")
    return t


def test_host_parser_matches_oracle_on_random_text():
    rng = random.Random(11)
    for _ in range(300):
        t = _random_text(rng).encode()
        if not t:
            continue
        b1, o1 = T.parse_buffer(t)
        b2, o2 = O.parse(t)
        assert np.array_equal(o1, o2) and np.array_equal(b1, b2)


def _quirk_fastq(rng):
    """FASTQ where some records start with an empty header line whose '\n'
    sits at block offset 63 (the reference lexer then loses that line feed and
    the record spans five lines), and whose DNA may end at a multiple of 64."""
    out = ""
    for i in range(rng.randint(2, 12)):
        if rng.random() < 0.5:
            # a filler record that brings the next record start to offset 63
            pad = (63 - len(out)) % 64
            sl = rng.randint(0, 20)
            h = (pad - (2 * sl + 6)) % 64
            out += "@%s\n%s\n+\n%s\n" % ("x" * h, "A" * sl, "?" * sl)
            assert len(out) % 64 == 63
            hdr = ""
        else:
            hdr = "@r%d" % i
        seq = "".join(rng.choice("ACGTN") for _ in range(rng.randint(0, 90)))
        out += "%s\n%s\n+\n%s\n" % (hdr, seq, "?" * len(seq))
    if rng.random() < 0.5:
        out = out.rstrip("\n")
    if rng.random() < 0.3:
        out += "@z\n" + "A" * ((-len(out) - 3) % 64)  # DNA that ends the text at a multiple of 64
    return out


def test_parallel_parser_matches_sequential():
    """Any number of threads gives the sequential parse: cuts inside records,
    inside headers, inside chunks that continue over many lines, around the
    64-byte-block behaviours and at the text's ends."""
    rng = random.Random(12)
    texts = [GOLD["inputs"][n].encode() for n in sorted(GOLD["parse"])]
    texts += [_random_text(rng).encode() for _ in range(300)]
    texts += [_quirk_fastq(rng).encode() for _ in range(200)]
    # long multi-line and single-line FASTA records, headers across cuts
    for _ in range(20):
        g = "".join(rng.choice("ACGTN" if rng.random() < 0.01 else "ACGT") for _ in range(rng.randint(500, 5000)))
        w = rng.choice([0, 60, 80])
        body = "\n".join(g[j:j + w] for j in range(0, len(g), w)) if w else g
        texts.append(("junk\n>h1 " + "x" * rng.randint(0, 300) + "\n" + body + "\n>h2\n" + body[:777]).encode())
    for t in texts:
        if not t:
            continue
        b1, o1 = T.parse_buffer(t)
        for th in (2, 3, 5, 8, 16, 37, 64):
            b2, o2 = T.parse_buffer(t, threads=th)
            assert np.array_equal(o1, o2) and np.array_equal(b1, b2), (th, t[:80])


def test_host_parser_matches_oracle_on_block_boundary_fastq():
    rng = random.Random(3)
    for _ in range(200):
        t = _quirk_fastq(rng).encode()
        b1, o1 = T.parse_buffer(t)
        b2, o2 = O.parse(t)
        assert np.array_equal(o1, o2) and np.array_equal(b1, b2)


def test_parallel_parser_on_a_large_fastq():
    t = T.synth("kind=reads_fq genome=200000 seed=5 reads=20000 rl=150 sub=0.005 nrate=0.001")
    b1, o1 = T.parse_buffer(t)
    for th in (4, 16):
        b2, o2 = T.parse_buffer(t, threads=th)
        assert np.array_equal(o1, o2) and np.array_equal(b1, b2)


def test_host_parser_empty_raises():
    with pytest.raises(T.TunaGpuError):
        T.parse_buffer(b"")


def test_synth_is_deterministic_and_parses():
    a = T.synth("kind=reads_fq genome=20000 seed=2 reads=200 rl=150 sub=0.005 nrate=0.001")
    b = T.synth("kind=reads_fq genome=20000 seed=2 reads=200 rl=150 sub=0.005 nrate=0.001")
    c = T.synth("kind=reads_fq genome=20000 seed=3 reads=200 rl=150 sub=0.005 nrate=0.001")
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    bases, off = T.parse_buffer(a)
    assert len(off) >= 201 and int(off[-1]) <= 200 * 150
    g = T.synth("kind=genome_fa len=10000 seed=1")
    bases, off = T.parse_buffer(g)
    assert len(off) == 2 and int(off[-1]) == 10000


def auto_n(sizes, gz, k, threads, l2, l3, fd):
    L = T.lib()
    L.tg_auto_partitions_with.restype = T.ctypes.c_uint32
    L.tg_auto_partitions_with.argtypes = [T._u64p, T._u8p, T.ctypes.c_uint32, T.ctypes.c_uint32,
                                          T.ctypes.c_uint32, T.ctypes.c_uint64, T.ctypes.c_uint64,
                                          T.ctypes.c_uint64]
    s = np.array(sizes, dtype=np.uint64)
    g = np.array(gz, dtype=np.uint8)
    return L.tg_auto_partitions_with(s.ctypes.data_as(T._u64p), g.ctypes.data_as(T._u8p), len(sizes), k,
                                     threads, l2, l3, fd)


# Host shapes with the reference formula (src/setup.hpp:82-196) worked by hand:
#   max_parts = fd - 32;  writer_cap = largest 4096 * 2^i <= l2 / 32, at most 32768
#   est = 76% of the bytes (the largest file's when several files average < 50 MB;
#         gzip counts 6x);  per_kmer = (8 ceil(k/32) + 6) * 5 / 4 rounded up
#   n_l3 = pow2 >= ceil(t * est * per_kmer / (3/4 l3)), at least 16
#   n_tp = pow2 >= max(1, total >> 21), at least 16;  n = min(max(n_l3, n_tp), caps), >= 16
@pytest.mark.parametrize("case,expect", [
    # 4,660,000 B FASTA, k=31, t=1, L2 1 MiB, L3 32 MiB, fd 1024:
    #   est = 3,541,600; per_kmer = (14*5+3)/4 = 18; 3,541,600*18 = 63,748,800
    #   / 25,165,824 -> 3 -> n_l3 16; 4,660,000 >> 21 = 2 -> n_tp 16; n = 16
    ((([4660000], [0]), 31, 1, 1 << 20, 32 << 20, 1024), 16),
    # 4,188,888,785 B FASTQ, k=31, t=16, L2 512 KiB, L3 256 MiB, fd 1024:
    #   est = 3,183,555,476; 16*est*18 = 916,863,977,088 / 201,326,592 = 4554.1 -> 4555
    #   -> n_l3 8192; total >> 21 = 1997 -> n_tp 2048; n = 8192; caps: fd 992,
    #   writer 16384 (524288/32 = 16384) -> 992
    ((([4188888785], [0]), 31, 16, 512 << 10, 256 << 20, 1024), 992),
    # the same with fd 65536: caps 65504 and 16384 -> 8192
    ((([4188888785], [0]), 31, 16, 512 << 10, 256 << 20, 65536), 8192),
    # the same with L2 2 MiB (writer cap 32768 at most) and t = 64:
    #   64*est*18 = 3,667,455,908,352 / 201,326,592 = 18216.4 -> 32768; caps 32768 -> 32768
    ((([4188888785], [0]), 31, 64, 2 << 20, 256 << 20, 1 << 20), 32768),
    # 200 files of 5,000,000 B (average < 50 MB: the largest counts), k=31, t=8,
    #   L2 1 MiB, L3 32 MiB, fd 4096: est = 3,800,000; 8*3,800,000*18 = 547,200,000
    #   / 25,165,824 = 21.7 -> 22 -> n_l3 32; total 1e9 >> 21 = 476 -> n_tp 512; n = 512
    ((([5000000] * 200, [0] * 200), 31, 8, 1 << 20, 32 << 20, 4096), 512),
    # one 100,000,000 B gzip file (600,000,000 B counted), k=63, t=4, L2 256 KiB,
    #   L3 64 MiB, fd 2048: per_kmer = (22*5+3)/4 = 28; est = 456,000,000;
    #   4*456,000,000*28 = 51,072,000,000 / 50,331,648 = 1014.7 -> 1015 -> n_l3 1024;
    #   600,000,000 >> 21 = 286 -> n_tp 512; n = 1024; caps 2016, writer 8192 -> 1024
    ((([100000000], [1]), 63, 4, 256 << 10, 64 << 20, 2048), 1024),
    # tiny input, fd 40 (max_parts 8) -> min(16, 8) = 8, floor 16 -> 16
    ((([1000], [0]), 21, 1, 1 << 20, 32 << 20, 40), 16),
])
def test_auto_partitions_formula_pinned(case, expect):
    (sizes, gz), k, t, l2, l3, fd = case
    assert auto_n(sizes, gz, k, t, l2, l3, fd) == expect


def test_server_refuses_k_above_64_before_any_device_call():
    # parameter checks run on the host: the message names the entry points
    # that count such jobs
    with pytest.raises(T.TunaGpuError, match="the server counts k <= 64"):
        T.Server(127, 21)
    with pytest.raises(T.TunaGpuError, match="invalid k/m"):
        T.Server(300, 21)
