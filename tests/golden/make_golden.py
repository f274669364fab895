"""Writes the golden vectors in this folder from the reference itself.

Runs oracle/_ref/ref_tuna — our driver compiled against the reference's own
headers (nt_hash.hpp, minimizer_window.hpp, partition_hash.hpp,
superkmer_io.hpp, seq_source.hpp/helicase) — on small hand-made inputs and
stores inputs and outputs as JSON.  Needs /root/reference (this container);
the committed JSON files are what the tests read.
"""
import hashlib
import json
import os
import random
import struct
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_tuna")

KM = [(31, 21), (31, 23), (21, 15), (63, 21), (11, 5), (15, 7), (5, 3), (9, 8), (33, 15),
      (25, 11), (17, 2), (64, 31), (32, 16), (3, 2), (45, 9)]


# (k, m) pairs of the count fixtures: the workloads' shapes and small ones
COUNT_KM = [(31, 21), (31, 23), (21, 15), (63, 21), (11, 5), (5, 3), (33, 15), (64, 31)]


def rs(rng, n, alph="ACGT"):
    return "".join(rng.choice(alph) for _ in range(n))


def fasta(rng, name, s, width=60):
    return ">" + name + "\n" + "\n".join(s[i:i + width] for i in range(0, len(s), width)) + "\n"


def fastq(reads):
    return "".join("@r%d\n%s\n+\n%s\n" % (i, s, "?" * len(s)) for i, s in enumerate(reads))


def inputs():
    rng = random.Random(20240601)
    cases = {}
    cases["fasta_multi"] = fasta(rng, "a", rs(rng, 2500)) + fasta(rng, "b desc", rs(rng, 700) + "N" + rs(rng, 300), 70)
    cases["homopolymer_cap"] = ">h\n" + "A" * 1200 + rs(rng, 120) + "CA" * 700 + rs(rng, 60) + "ACGTTGCA" * 90 + "\n"
    cases["fastq_reads"] = fastq([rs(rng, 150) for _ in range(60)] + [rs(rng, 70) + "N" + rs(rng, 79)])
    cases["mixed_case_cr"] = ">x\nacgtnACGTacgtnn\n>y desc\n" + rs(rng, 90).lower() + "\r\n" + rs(rng, 90) + "\n"
    cases["preamble_gt_midline"] = "junk ACGT before header\n>r1\n" + rs(rng, 80) + ">inline hdr\n" + rs(rng, 80) + "\n\n" + rs(rng, 40) + "\n"
    cases["iupac_spaces"] = ">r\n" + rs(rng, 60) + "R" + rs(rng, 60) + " " + rs(rng, 60) + "\t" + rs(rng, 60) + "\n"
    cases["repeat_families"] = fasta(rng, "rep", "".join(rs(rng, 40) + fam for fam in [rs(rng, 150)] * 6))
    cases["palindromes"] = ">p\n" + ("ACGT" * 10 + "AATT" * 10 + "GGCC" * 10) * 4 + "\n"
    # FASTQ whose last sequence run ends exactly at a 64-byte boundary at EOF
    body = "@r\n" + rs(rng, 61)
    cases["fastq_eof64"] = body
    # empty header line landing at block offset 63
    rd = rs(rng, 28)  # 2*28 + 7 = 63 bytes, so the empty line's '\n' is byte 63
    pre = "@a\n" + rd + "\n+\n" + "?" * len(rd) + "\n"
    assert len(pre) == 63
    cases["fastq_lf63"] = pre + "\n" + rs(rng, 50) + "\n+\n" + rs(rng, 50) + "\n@c\n" + rs(rng, 50) + "\n+\n" + "?" * 50 + "\n"
    return cases


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/ref_tuna not built (make -C oracle ref)")
    tmp = tempfile.mkdtemp()
    out = {"inputs": {}, "parse": {}, "records": {}, "hash": {}, "window": {}}
    rng = random.Random(99)
    for name, text in inputs().items():
        path = os.path.join(tmp, name)
        with open(path, "w") as f:
            f.write(text)
        out["inputs"][name] = text
        chunks = subprocess.run([REF, "parse", path], capture_output=True, check=True).stdout.split(b"\n")[:-1]
        out["parse"][name] = [c.decode() for c in chunks]
        for (k, m) in KM:
            for n in (1, 16):
                recf = path + ".rec"
                p = subprocess.run([REF, "records", str(k), str(m), str(n), path, recf],
                                   capture_output=True, text=True, check=True)
                data = open(recf, "rb").read()
                nn = struct.unpack_from("<I", data, 0)[0]
                pos, parts = 4, []
                for _ in range(nn):
                    sz = struct.unpack_from("<Q", data, pos)[0]
                    pos += 8
                    parts.append(data[pos:pos + sz].hex())
                    pos += sz
                st = dict(l.split(": ") for l in p.stderr.strip().split("\n"))
                out["records"]["%s|%d|%d|%d" % (name, k, m, n)] = {
                    "parts": parts, "kmers": int(st["kmers"]), "superkmers": int(st["superkmers"])}
    for m in sorted({m for _, m in KM}):
        for s in [rs(rng, 120), "A" * 50, "ACGT" * 15, rs(rng, 90, "AC")]:
            v = subprocess.run([REF, "hash", str(m), s], capture_output=True, text=True, check=True).stdout.split()
            out["hash"]["%d|%s" % (m, s)] = [int(x) for x in v]
    for (k, m) in KM:
        for s in [rs(rng, 260), "A" * 150, "ACGT" * 50, rs(rng, 240, "AC"), "CAG" * 60 + rs(rng, 30)]:
            v = subprocess.run([REF, "window", str(k), str(m), s], capture_output=True, text=True, check=True).stdout.split()
            out["window"]["%d|%d|%s" % (k, m, s)] = [int(x) for x in v]
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"), sort_keys=True)
    # counts from the reference's own table (the count mode of the driver
    # runs Streaming_Kmer_Hash_Table over the reference's partition buffers):
    # per case the number of reported k-mers, the sum of their counts and the
    # SHA-256 of the sorted "<label>\t<count>\n" lines
    counts = {}
    for name, text in inputs().items():
        path = os.path.join(tmp, name)
        for (k, m) in COUNT_KM:
            for ci, cx in ((1, 2**64 - 1), (2, 2**64 - 1), (1, 1), (2, 3)):
                o = path + ".tsv"
                subprocess.run([REF, "count", str(k), str(m), "16", "2", str(ci), str(cx), o, path],
                               capture_output=True, text=True, check=True)
                lines = sorted(open(o, "rb").read().splitlines(keepends=True))
                body = b"".join(lines)
                counts["%s|%d|%d|%d|%d" % (name, k, m, ci, cx)] = {
                    "n": len(lines), "sum": sum(int(l.split(b"\t")[1]) for l in lines),
                    "sha256": hashlib.sha256(body).hexdigest()}
    with open(os.path.join(HERE, "reference_counts.json"), "w") as f:
        json.dump(counts, f, separators=(",", ":"), sort_keys=True)


if __name__ == "__main__":
    main()
