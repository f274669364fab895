"""CPU oracle for the k-mer counting path — TEST INFRASTRUCTURE ONLY.

ctypes bindings to oracle/liboracle.so (a plain-C restatement of the
reference's algorithm, see oracle.c).  Only tests/, __graft_entry__.smoke()
and the cpu_baseline leg of bench.py may import this module, and only as the
checker; the product (tuna_amd) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
REF_BIN = os.path.join(_HERE, "_ref", "ref_tuna")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_parse.restype = ctypes.c_int64
        L.orc_parse.argtypes = [_u8p, ctypes.c_uint64, _u8p, _u64p, ctypes.c_uint64]
        L.orc_nthash_direct.restype = ctypes.c_uint64
        L.orc_nthash_direct.argtypes = [_u8p, ctypes.c_int]
        L.orc_nthash_roll.restype = None
        L.orc_nthash_roll.argtypes = [_u8p, ctypes.c_uint64, ctypes.c_int, _u64p]
        for fn in (L.orc_window_stream, L.orc_window_brute):
            fn.restype = None
            fn.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_int, _u64p, _u32p]
        L.orc_superkmers.restype = ctypes.c_int64
        L.orc_superkmers.argtypes = [_u8p, _u64p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                     _u64p, _u32p, _u32p, _u64p, ctypes.c_uint64, _u64p]
        L.orc_pack_record.restype = ctypes.c_uint64
        L.orc_pack_record.argtypes = [_u8p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                      ctypes.c_int, _u8p]
        L.orc_canonical.restype = None
        L.orc_canonical.argtypes = [_u8p, ctypes.c_int, _u64p]
        L.orc_count.restype = ctypes.c_int64
        L.orc_count.argtypes = [_u8p, _u64p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32,
                                ctypes.c_uint64, _u64p, _u32p, ctypes.c_uint64, _u64p, _u64p]
        L.orc_tsv_line.restype = ctypes.c_uint64
        L.orc_tsv_line.argtypes = [_u64p, ctypes.c_int, ctypes.c_uint32, ctypes.c_char_p]
        L.orc_count_parallel.restype = ctypes.c_int64
        L.orc_count_parallel.argtypes = [_u8p, _u64p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                         _u64p, _u32p, ctypes.c_uint64, _u64p,
                                         ctypes.POINTER(ctypes.c_double)]
        L.orc_round_pow2.restype = ctypes.c_uint32
        L.orc_round_pow2.argtypes = [ctypes.c_uint32]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def as_u8(x):
    if isinstance(x, (bytes, bytearray)):
        return np.frombuffer(bytes(x), dtype=np.uint8)
    if isinstance(x, str):
        return np.frombuffer(x.encode(), dtype=np.uint8)
    return np.ascontiguousarray(x, dtype=np.uint8)


def parse(data):
    """FASTA/FASTQ bytes -> (bases uint8, offsets uint64)."""
    d = as_u8(data)
    out = np.zeros(max(len(d), 1), dtype=np.uint8)
    cap = len(d) // 2 + 2
    off = np.zeros(cap, dtype=np.uint64)
    n = lib().orc_parse(_p(d, _u8p), len(d), _p(out, _u8p), _p(off, _u64p), cap)
    if n == -2:
        raise ValueError("Empty file")
    if n < 0:
        raise RuntimeError("orc_parse capacity")
    off = off[: n + 1].copy()
    return out[: int(off[-1])].copy(), off


def chunks_of(strings):
    """List of ACGT strings -> (bases, offsets)."""
    b = "".join(strings).encode()
    off = np.zeros(len(strings) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(s) for s in strings])
    return np.frombuffer(b, dtype=np.uint8).copy() if b else np.zeros(0, np.uint8), off


def nthash_direct(s, m):
    a = as_u8(s)
    return lib().orc_nthash_direct(_p(a, _u8p), m)


def nthash_roll(s, m):
    a = as_u8(s)
    n = len(a) - m + 1
    out = np.zeros(max(n, 1), dtype=np.uint64)
    if n > 0:
        lib().orc_nthash_roll(_p(a, _u8p), len(a), m, _p(out, _u64p))
    return out[: max(n, 0)]


def window(h, W, brute=False):
    h = np.ascontiguousarray(h, dtype=np.uint64)
    K = len(h) - W + 1
    H = np.zeros(max(K, 1), dtype=np.uint64)
    P = np.zeros(max(K, 1), dtype=np.uint32)
    if K > 0:
        fn = lib().orc_window_brute if brute else lib().orc_window_stream
        fn(_p(h, _u64p), len(h), W, _p(H, _u64p), _p(P, _u32p))
    return H[: max(K, 0)], P[: max(K, 0)]


def superkmers(bases, off, k, m):
    """-> (start, len, minpos, hash) arrays, total k-mers."""
    bases = as_u8(bases)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nch = len(off) - 1
    cap = max(int(off[-1]) if nch else 0, 1)
    st = np.zeros(cap, np.uint64)
    ln = np.zeros(cap, np.uint32)
    mp = np.zeros(cap, np.uint32)
    hs = np.zeros(cap, np.uint64)
    km = ctypes.c_uint64(0)
    n = lib().orc_superkmers(_p(bases, _u8p), _p(off, _u64p), nch, k, m, _p(st, _u64p),
                             _p(ln, _u32p), _p(mp, _u32p), _p(hs, _u64p), cap, ctypes.byref(km))
    if n < 0:
        raise RuntimeError("orc_superkmers capacity")
    return st[:n].copy(), ln[:n].copy(), mp[:n].copy(), hs[:n].copy(), km.value


def pack_record(ascii_bytes, min_pos, k, m):
    a = as_u8(ascii_bytes)
    out = np.zeros(len(a) // 4 + 8, dtype=np.uint8)
    n = lib().orc_pack_record(_p(a, _u8p), len(a), min_pos, k, m, _p(out, _u8p))
    return bytes(out[:n])


def key_words(k):
    """64-bit words of a k-mer key: ceil(k/32)."""
    return (k + 31) // 32


def canonical(s, k):
    a = as_u8(s)
    w = np.zeros(max(2, key_words(k)), dtype=np.uint64)
    lib().orc_canonical(_p(a, _u8p), k, _p(w, _u64p))
    return sum(int(w[i]) << (64 * i) for i in range(key_words(k)))


def count(bases, off, k, ci=1, cx=2**64 - 1):
    """-> (keys uint64 [n, nw], counts uint32 [n], total k-mers, distinct)."""
    bases = as_u8(bases)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nch = len(off) - 1
    nw = key_words(k)
    K = 0
    if nch:
        L = np.diff(off.astype(np.int64))
        K = int(np.maximum(L - k + 1, 0).sum())
    cap = max(K, 1)
    keys = np.zeros(cap * nw, np.uint64)
    cnt = np.zeros(cap, np.uint32)
    tot = ctypes.c_uint64(0)
    nu = ctypes.c_uint64(0)
    n = lib().orc_count(_p(bases, _u8p), _p(off, _u64p), nch, k, ci, cx, _p(keys, _u64p),
                        _p(cnt, _u32p), cap, ctypes.byref(tot), ctypes.byref(nu))
    if n < 0:
        raise RuntimeError("orc_count capacity")
    return keys[: n * nw].reshape(n, nw).copy(), cnt[:n].copy(), tot.value, nu.value


def count_parallel(bases, off, k, m, n_parts=256, ci=1, cx=2**64 - 1, threads=1, keep=True):
    """The two-phase multithreaded CPU count (cpu_count.c).  -> (keys [n, nw]
    or None, counts or None, {kmers, superkmers, unique, out, phase1_s,
    phase2_s}); keep=False only counts (the timed baseline)."""
    bases = as_u8(bases)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nch = len(off) - 1
    nw = 2 if k > 32 else 1
    cap = 0
    keys = cnt = None
    if keep:
        L = np.diff(off.astype(np.int64)) if nch else np.zeros(0, np.int64)
        cap = max(int(np.maximum(L - k + 1, 0).sum()), 1)
        keys = np.zeros(cap * nw, np.uint64)
        cnt = np.zeros(cap, np.uint32)
    st = np.zeros(3, np.uint64)
    ph = (ctypes.c_double * 2)()
    n = lib().orc_count_parallel(_p(bases, _u8p), _p(off, _u64p), nch, k, m, n_parts, ci, cx, threads,
                                 _p(keys, _u64p) if keep else None, _p(cnt, _u32p) if keep else None, cap,
                                 _p(st, _u64p), ph)
    if n < 0:
        raise RuntimeError("orc_count_parallel failed (%d)" % n)
    info = {"kmers": int(st[0]), "superkmers": int(st[1]), "unique": int(st[2]), "out": int(n),
            "phase1_s": ph[0], "phase2_s": ph[1]}
    if keep:
        return keys[: n * nw].reshape(n, nw).copy(), cnt[:n].copy(), info
    return None, None, info


def tsv_line(key, k, count):
    w = np.array([(key >> (64 * i)) & (2**64 - 1) for i in range(max(2, key_words(k)))], dtype=np.uint64)
    buf = ctypes.create_string_buffer(k + 16)
    n = lib().orc_tsv_line(_p(w, _u64p), k, count, buf)
    return buf.raw[:n].decode()


def round_pow2(n):
    return lib().orc_round_pow2(n)


def ref_available():
    return os.path.exists(REF_BIN)
