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
