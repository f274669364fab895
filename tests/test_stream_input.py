"""CPU tests of the file reader behind tg_count_files: a gzip file parsed in
sections while it inflates, BGZF members inflated on many threads and plain
files parsed on many threads must all give the bases and chunk ends that
parsing the file's whole text gives (tg_parse_buffer on the text zlib's
gzread returns, which is how the reference reads a file,
src/seq_source.hpp:48-205).  Section boundaries fall inside headers, inside
records, inside chunks that run over many lines and around the lexer's
64-byte block behaviours.  No GPU calls."""
import gzip
import os
import random
import struct
import zlib

import numpy as np
import pytest

import tuna_amd as T
from test_host import _quirk_fastq, _random_text


def _bgzf(data, block=65280, level=6):
    """BGZF (SAM/BAM specification 4.1): gzip members of at most 64 KiB, each
    with a 'BC' extra subfield holding its size minus one, then the empty
    end-of-file member."""
    out = bytearray()
    for i in range(0, len(data), block):
        blk = data[i:i + block]
        c = zlib.compressobj(level, zlib.DEFLATED, -15)
        body = c.compress(blk) + c.flush()
        bsize = 18 + len(body) + 8
        out += b"\x1f\x8b\x08\x04\x00\x00\x00\x00\x00\xff" + struct.pack("<H", 6) + b"BC" + \
            struct.pack("<HH", 2, bsize - 1)
        out += body + struct.pack("<II", zlib.crc32(blk) & 0xFFFFFFFF, len(blk))
    out += bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
    return bytes(out)


def _write(tmp_path, name, payload):
    p = tmp_path / name
    p.write_bytes(payload)
    return str(p)


def _whole(text):
    if not text:
        return np.zeros(0, np.uint8), np.zeros(1, np.uint64)
    return T.parse_buffer(text)


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def _texts():
    rng = random.Random(21)
    out = []
    # many small records back to back: a few kB each, so small sections cut them everywhere
    for _ in range(6):
        out.append("".join(_random_text(rng) for _ in range(40)).encode())
    for _ in range(6):
        out.append(_quirk_fastq(rng).encode())
    q = ""
    while len(q) < 60000:
        q += _quirk_fastq(rng).rstrip("\n") + "\n"
    out.append(q.encode())
    # long multi-line FASTA records
    g = "".join(rng.choice("ACGTN" if rng.random() < 0.01 else "ACGT") for _ in range(40000))
    out.append(("junk\n>h1 " + "x" * 300 + "\n" + "\n".join(g[j:j + 60] for j in range(0, len(g), 60)) +
                "\n>h2\n" + g[:5000]).encode())
    out.append(T.synth("kind=reads_fq genome=20000 seed=7 reads=800 rl=150 sub=0.01 nrate=0.002").tobytes())
    return out


@pytest.mark.parametrize("section", [0, 64, 700, 4096])
def test_streamed_gzip_parse_equals_whole_text(tmp_path, section):
    for i, t in enumerate(_texts()):
        want = _whole(t)
        path = _write(tmp_path, f"t{i}.gz", gzip.compress(t, 6))
        got = T.parse_file(path, threads=3, section_bytes=section)
        assert _same(got, want), (i, section, t[:60])


@pytest.mark.parametrize("threads", [1, 4, 16])
def test_bgzf_and_plain_parse_equal_whole_text(tmp_path, threads):
    for i, t in enumerate(_texts()):
        want = _whole(t)
        for name, payload in ((f"b{i}.fq.gz", _bgzf(t, block=rnd_block(i))), (f"p{i}.txt", t)):
            got = T.parse_file(_write(tmp_path, name, payload), threads=threads, section_bytes=512)
            assert _same(got, want), (name, threads)


def rnd_block(i):
    return [65280, 1000, 333, 4096][i % 4]


def test_reader_equals_gzread_text(tmp_path):
    """Every route reads the text zlib's gzread reads: concatenated members,
    bytes after the last member, a member cut short, a '.gz' name on a plain
    file."""
    t = T.synth("kind=reads_fq genome=30000 seed=3 reads=600 rl=120 sub=0.01 nrate=0.001").tobytes()
    half = len(t) // 2
    cases = {
        "members.fq.gz": gzip.compress(t[:half]) + gzip.compress(t[half:]),
        "trailing.fq.gz": gzip.compress(t) + b"not a member\n",
        "cut.fq.gz": gzip.compress(t)[: len(gzip.compress(t)) * 2 // 3],
        "plain_named.fq.gz": t,
        "bgzf_cut.fq.gz": _bgzf(t, block=2000)[:-28 - 700],
    }
    for name, payload in cases.items():
        path = _write(tmp_path, name, payload)
        want = T.parse_buffer(T.read_file(path))
        for section in (0, 128):
            got = T.parse_file(path, threads=4, section_bytes=section)
            assert _same(got, want), name


def test_reader_errors_carry_reference_messages(tmp_path):
    cases = [
        (str(tmp_path / "missing.fa"), "Cannot open file: "),
        (_write(tmp_path, "empty.fa", b""), "Empty file: "),
        (str(tmp_path / "missing.fa.gz"), "Cannot open: "),
        (_write(tmp_path, "empty.fa.gz", b""), "Cannot read: "),
        (_write(tmp_path, "nothing.fa.gz", gzip.compress(b"")), "Cannot read: "),
    ]
    for path, msg in cases:
        with pytest.raises(T.TunaGpuError) as e:
            T.parse_file(path, threads=2)
        assert str(e.value) == msg + path


@pytest.fixture(scope="module")
def eng():
    e = T.Engine(0)
    yield e
    e.close()


@pytest.mark.gpu
def test_count_files_equals_count_of_parsed_text(eng, tmp_path):
    """tg_count_files (files read straight to the device in sections) gives
    the result tg_count_host gives on the same files' chunks, for plain,
    gzip, concatenated-gzip and BGZF files, one file or several."""
    rng = random.Random(5)
    t1 = T.synth("kind=reads_fq genome=60000 seed=8 cov=10 rl=150 sub=0.005 nrate=0.001").tobytes()
    t2 = T.synth("kind=genome_fa len=300000 seed=9").tobytes()
    t3 = "".join(_random_text(rng) for _ in range(200)).encode()
    files = [
        _write(tmp_path, "a.fq", t1),
        _write(tmp_path, "a.fq.gz", gzip.compress(t1)),
        _write(tmp_path, "b.fa.gz", gzip.compress(t2[:100000]) + gzip.compress(t2[100000:])),
        _write(tmp_path, "c.fq.gz", _bgzf(t1, block=7000)),
        _write(tmp_path, "d.fa", t3),
    ]
    sets = [[f] for f in files] + [files, files[::-1]]
    for k, m, ci in ((31, 21, 1), (21, 11, 2), (55, 19, 1)):
        for fs in sets:
            parts = [T.parse_buffer(T.read_file(f)) for f in fs]
            bases = np.concatenate([b for b, _ in parts])
            off, base = [np.zeros(1, np.uint64)], 0
            for b, o in parts:
                off.append(o[1:] + np.uint64(base))
                base += len(b)
            off = np.concatenate(off)
            wk, wc = eng.count(bases, off, k, m, 16, ci)
            ws = eng.last_stats
            gk, gc = eng.count_files(fs, k, m, 16, ci, threads=4)
            gs = eng.last_stats
            assert (gs.n_bases, gs.n_chunks, gs.n_kmers, gs.n_unique, gs.n_out) == \
                (ws.n_bases, ws.n_chunks, ws.n_kmers, ws.n_unique, ws.n_out), fs
            o1, o2 = np.lexsort(wk.T[::-1]), np.lexsort(gk.T[::-1])
            assert np.array_equal(wk[o1], gk[o2]) and np.array_equal(wc[o1], gc[o2]), fs


@pytest.mark.gpu
def test_count_files_reports_input_errors(eng, tmp_path):
    ok = _write(tmp_path, "ok.fa", b">a\nACGTACGTACGTACGTACGTACGTACGTACGTAA\n")
    for bad, msg in ((str(tmp_path / "none.fa"), "Cannot open file: "),
                     (_write(tmp_path, "e.fa.gz", b""), "Cannot read: ")):
        with pytest.raises(T.TunaGpuError) as e:
            eng.count_files([ok, bad], 21, 11)
        assert str(e.value).endswith(": " + msg + bad)  # the message after the status code
    gk, gc = eng.count_files([ok], 21, 11)
    assert len(gc) > 0
