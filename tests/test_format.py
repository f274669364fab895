"""Output formats built on the device (TSV lines, KFF raw-section records, whole
files) against the host formatters, which the CPU tests pin to the oracle's
line format and a KFF reader.  Byte-exact, in the result's own order."""
import os
import sys

import numpy as np
import pytest

import tuna_amd as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def eng():
    e = T.Engine(0)
    yield e
    e.close()


def host_tsv(k, keys, cnt):
    n = len(cnt)
    buf = np.zeros(max(n, 1) * (k + 12), dtype=np.uint8)
    ka = np.ascontiguousarray(keys.reshape(-1))
    got = T.lib().tg_format_tsv(k, ka.ctypes.data_as(T._u64p), cnt.ctypes.data_as(T._u32p), n,
                                buf.ctypes.data_as(T.ctypes.c_char_p), len(buf))
    assert got >= 0
    return bytes(buf[:got])


def host_kff(path, k, keys, cnt):
    ka = np.ascontiguousarray(keys.reshape(-1))
    assert T.lib().tg_write_kff(os.fsencode(path), k, ka.ctypes.data_as(T._u64p),
                                cnt.ctypes.data_as(T._u32p), len(cnt)) == 0
    return open(path, "rb").read()


def inputs():
    low = b">lc\n" + b"A" * 300000 + b"CG" * 70000 + b"\n"       # counts with 5-6 digits
    reads = T.synth("kind=reads_fq genome=60000 seed=21 cov=12 rl=150 sub=0.01 nrate=0.001")
    genome = T.synth("kind=genome_fa len=120000 seed=22")
    return [T.parse_buffer(low), T.parse_buffer(reads), T.parse_buffer(genome)]


@pytest.mark.parametrize("k,m", [(2, 1), (15, 7), (31, 21), (32, 15), (33, 17), (63, 21), (64, 31), (65, 21),
                                 (127, 21), (200, 31)])
def test_device_tsv_and_kff_records_match_host(eng, k, m, tmp_path):
    for i, (b, o) in enumerate(inputs()):
        say(f"format k={k} input {i}: {len(b)} bases")
        keys, cnt = eng.count(b, o, k, m, 16, 1)
        tsv = eng.result_tsv(k)
        assert tsv == host_tsv(k, keys, cnt)
        # ranges of the result format to the same bytes, in pieces
        n = len(cnt)
        cuts = sorted({0, n, n // 3, (2 * n) // 3, min(n, 511), min(n, 513)})
        parts = [eng.result_tsv(k, a, c - a) for a, c in zip(cuts, cuts[1:])]
        assert b"".join(parts) == tsv
        # KFF records are the file's bytes between header and footer
        rec = eng.result_kff(k)
        kb = (k + 3) // 4 + 4
        assert len(rec) == n * kb
        whole = host_kff(tmp_path / "h.kff", k, keys, cnt)
        if n == 0:  # every read shorter than k
            assert rec == b"" and len(eng.result_tsv(k)) == 0
            continue
        pos = whole.find(rec)
        assert pos > 0 and whole[pos + len(rec):].endswith(b"KFF")
        if n > 7:
            assert eng.result_kff(k, 5, 2) == rec[5 * kb:7 * kb]


def test_whole_file_writers_across_chunks(eng, tmp_path):
    # 4.6 M distinct k-mers: three device passes of the file writer
    b, o = T.parse_buffer(T.synth("kind=genome_fa len=4600000 seed=1"))
    keys, cnt = eng.count(b, o, 31, 21, 16, 1)
    say(f"whole-file writers: {len(cnt)} entries")
    assert len(cnt) > 2 * (1 << 21)
    t = tmp_path / "d.tsv"
    eng.result_write(31, t)
    assert open(t, "rb").read() == host_tsv(31, keys, cnt)
    kd = tmp_path / "d.kff"
    eng.result_write(31, kd, kff=True)
    assert open(kd, "rb").read() == host_kff(tmp_path / "h.kff", 31, keys, cnt)


def test_filters_empty_results_and_errors(eng, tmp_path):
    b, o = T.parse_buffer(b">a\nACGTACGTAC\n")
    keys, cnt = eng.count(b, o, 5, 3, 16, 2)    # ci = 2 keeps the repeated 5-mers only
    assert eng.result_tsv(5) == host_tsv(5, keys, cnt)
    keys, cnt = eng.count(b, o, 5, 3, 16, 100)  # nothing passes
    assert len(cnt) == 0 and eng.result_tsv(5) == b"" and eng.result_kff(5) == b""
    eng.result_write(5, tmp_path / "e.tsv")
    assert open(tmp_path / "e.tsv", "rb").read() == b""
    eng.result_write(5, tmp_path / "e.kff", kff=True)
    assert open(tmp_path / "e.kff", "rb").read() == host_kff(tmp_path / "h.kff", 5, keys, cnt)
    keys, cnt = eng.count(b, o, 5, 3, 16, 1)
    with pytest.raises(T.TunaGpuError, match="key width"):
        eng.result_tsv(40)
    with pytest.raises(T.TunaGpuError, match="out of range"):
        eng.result_tsv(5, len(cnt), 1)
    small = T.ctypes.create_string_buffer(4)
    assert T.lib().tg_result_tsv(eng._h, 5, 0, len(cnt), small, 4) == -4  # TG_ECAP
    assert small.raw == b"\0" * 4
    with pytest.raises(T.TunaGpuError, match="cannot open output file"):
        eng.result_write(5, tmp_path / "missing_dir" / "x.tsv")
