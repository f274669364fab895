"""tuna_amd — MI355X (gfx950) k-mer counting engine, Python bindings.

Thin ctypes layer over the C ABI in include/tuna_gpu.h (libtuna_gpu.so, built
in-tree by ``make -C tuna_amd/csrc``).  There is no CPU fallback: if the
library is missing, or no GPU is present, calls raise.

The Python surface mirrors the reference library API (include/tuna/tuna.hpp
of vicLeva/tuna: ``count``, ``count_to``, ``count_to_raw``, ``Options``)
for the counting path.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtuna_gpu.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p

# every symbol declared in include/tuna_gpu.h
EXPORTS = [
    "tg_ctx_create", "tg_ctx_destroy", "tg_last_error", "tg_set_profiling", "tg_kernel_times",
    "tg_reset_kernel_times", "tg_dev_alloc", "tg_dev_free", "tg_h2d", "tg_d2h", "tg_sync",
    "tg_count_host", "tg_count_device", "tg_result_size", "tg_result_copy", "tg_pack_host",
    "tg_superkmers_host", "tg_partitions_host", "tg_parse_buffer", "tg_read_file", "tg_free",
    "tg_auto_partitions", "tg_auto_partitions_with", "tg_format_tsv", "tg_write_kff", "tg_synth", "tg_synth_copy",
    "tg_synth_free", "tg_get_batch_config", "tg_set_batch_config", "tg_count_many",
    "tg_count_many_device", "tg_many_result_size", "tg_many_result_copy", "tg_many_result_device",
    "tg_synth_alloc", "tg_server_create", "tg_server_destroy", "tg_server_submit",
    "tg_server_poll", "tg_server_wait", "tg_server_result", "tg_server_release", "tg_server_stats",
    "tg_host_alloc", "tg_host_free", "tg_autotune_many", "tg_set_plain_sequence",
    "tg_set_debug_stats", "tg_set_wide_deque", "tg_set_input_budget", "tg_debug_partitions", "tg_debug_min_coverage",
    "tg_result_tsv", "tg_result_tsv_noalloc", "tg_result_kff", "tg_result_write", "tg_parse_buffer_mt", "tg_map_file",
    "tg_unmap_file", "tg_parse_file", "tg_count_files",
]


class TgParams(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("m", ctypes.c_uint32), ("n_partitions", ctypes.c_uint32),
                ("ci", ctypes.c_uint32), ("cx", ctypes.c_uint64)]


class TgStats(ctypes.Structure):
    _fields_ = [("n_bases", ctypes.c_uint64), ("n_chunks", ctypes.c_uint64),
                ("n_kmers", ctypes.c_uint64), ("n_superkmers", ctypes.c_uint64),
                ("n_unique", ctypes.c_uint64), ("n_out", ctypes.c_uint64),
                ("n_parts", ctypes.c_uint32), ("n_subparts", ctypes.c_uint32),
                ("n_lds_groups", ctypes.c_uint32), ("n_hbm_parts", ctypes.c_uint32),
                ("ms_phase1", ctypes.c_double), ("ms_phase2", ctypes.c_double)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class TgJobStats(ctypes.Structure):
    _fields_ = [("n_bases", ctypes.c_uint64), ("n_kmers", ctypes.c_uint64),
                ("n_superkmers", ctypes.c_uint64), ("n_unique", ctypes.c_uint64),
                ("n_out", ctypes.c_uint64)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class TgManyStats(ctypes.Structure):
    _fields_ = [("n_jobs", ctypes.c_uint32), ("n_batches", ctypes.c_uint32),
                ("n_graph_launches", ctypes.c_uint32), ("n_captures", ctypes.c_uint32),
                ("n_retries", ctypes.c_uint32), ("n_kmers", ctypes.c_uint64),
                ("n_unique", ctypes.c_uint64), ("n_out", ctypes.c_uint64),
                ("ms_wall", ctypes.c_double)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class TgBatchConfig(ctypes.Structure):
    _fields_ = [("kmers_per_bucket", ctypes.c_uint32), ("max_tiles", ctypes.c_uint32),
                ("batch_bases", ctypes.c_uint64), ("load_factor", ctypes.c_float),
                ("use_graph", ctypes.c_uint32), ("rho", ctypes.c_float), ("slice_bases", ctypes.c_uint64)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class TgServerConfig(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_uint32), ("max_batch_jobs", ctypes.c_uint32),
                ("max_batch_bases", ctypes.c_uint64), ("hbm_budget", ctypes.c_uint64)]


class TgRequestTimes(ctypes.Structure):
    _fields_ = [("submit_to_start_s", ctypes.c_double), ("submit_to_done_s", ctypes.c_double),
                ("lane", ctypes.c_uint32), ("batch_jobs", ctypes.c_uint32)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class TgServerInfo(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_uint32), ("lane_bases", ctypes.c_uint64),
                ("hbm_budget", ctypes.c_uint64), ("hbm_free", ctypes.c_uint64),
                ("queued", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("failed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("pinned_pool_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


TG_ETIMEOUT = -6


class TunaGpuError(RuntimeError):
    pass


def key_words(k):
    """64-bit words per key of a result: ceil(k/32), the reference Kmer<k> layout."""
    return (k + 31) // 32


def lib():
    """Load libtuna_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TunaGpuError(f"{LIB_PATH} not built: run __graft_entry__.build() or make -C tuna_amd/csrc")
    L = ctypes.CDLL(LIB_PATH)
    L.tg_last_error.restype = ctypes.c_char_p
    L.tg_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.tg_ctx_destroy.argtypes = [_vp]
    L.tg_ctx_destroy.restype = None
    L.tg_set_profiling.argtypes = [_vp, ctypes.c_int]
    L.tg_kernel_times.argtypes = [_vp, ctypes.c_char_p, ctypes.c_uint64,
                                  ctypes.POINTER(ctypes.c_double), _u64p, ctypes.c_int]
    L.tg_reset_kernel_times.argtypes = [_vp]
    L.tg_reset_kernel_times.restype = None
    L.tg_dev_alloc.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)]
    L.tg_dev_free.argtypes = [_vp, _vp]
    L.tg_h2d.argtypes = [_vp, _vp, _vp, ctypes.c_uint64]
    L.tg_d2h.argtypes = [_vp, _vp, _vp, ctypes.c_uint64]
    L.tg_sync.argtypes = [_vp]
    L.tg_count_host.argtypes = [_vp, ctypes.POINTER(TgParams), _u8p, ctypes.c_uint64, _u64p,
                                ctypes.c_uint64, ctypes.POINTER(TgStats)]
    L.tg_count_device.argtypes = [_vp, ctypes.POINTER(TgParams), _vp, ctypes.c_uint64, _vp,
                                  ctypes.c_uint64, ctypes.POINTER(TgStats)]
    L.tg_result_size.argtypes = [_vp]
    L.tg_result_size.restype = ctypes.c_uint64
    L.tg_result_copy.argtypes = [_vp, _u64p, _u32p, ctypes.c_uint64]
    L.tg_pack_host.argtypes = [_vp, _u8p, ctypes.c_uint64, _u32p]
    L.tg_superkmers_host.argtypes = [_vp, ctypes.POINTER(TgParams), _u8p, ctypes.c_uint64, _u64p,
                                     ctypes.c_uint64, _u64p, _u32p, _u32p, _u32p, ctypes.c_uint64,
                                     _u64p, _u32p]
    L.tg_partitions_host.argtypes = [_vp, ctypes.POINTER(TgParams), _u8p, ctypes.c_uint64, _u64p,
                                     ctypes.c_uint64, _u64p, _u8p, ctypes.c_uint64, _u64p,
                                     ctypes.c_uint64, _u64p, _u64p]
    L.tg_parse_buffer.restype = ctypes.c_int64
    L.tg_parse_buffer.argtypes = [_u8p, ctypes.c_uint64, _u8p, _u64p, ctypes.c_uint64]
    L.tg_parse_buffer_mt.restype = ctypes.c_int64
    L.tg_parse_buffer_mt.argtypes = [_u8p, ctypes.c_uint64, _u8p, _u64p, ctypes.c_uint64, ctypes.c_uint32]
    L.tg_read_file.argtypes = [ctypes.c_char_p, ctypes.POINTER(_u8p), _u64p]
    L.tg_parse_file.restype = ctypes.c_int64
    L.tg_parse_file.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(_u8p), _u64p,
                                ctypes.POINTER(_u64p)]
    L.tg_count_files.argtypes = [_vp, ctypes.POINTER(TgParams), ctypes.POINTER(ctypes.c_char_p), ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_void_p]
    L.tg_map_file.argtypes = [ctypes.c_char_p, ctypes.POINTER(_u8p), _u64p]
    L.tg_unmap_file.argtypes = [_u8p, ctypes.c_uint64]
    L.tg_unmap_file.restype = None
    L.tg_free.argtypes = [_vp]
    L.tg_free.restype = None
    L.tg_auto_partitions.restype = ctypes.c_uint32
    L.tg_auto_partitions.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_uint32]
    L.tg_format_tsv.restype = ctypes.c_int64
    L.tg_format_tsv.argtypes = [ctypes.c_uint32, _u64p, _u32p, ctypes.c_uint64, ctypes.c_char_p,
                                ctypes.c_uint64]
    L.tg_write_kff.argtypes = [ctypes.c_char_p, ctypes.c_uint32, _u64p, _u32p, ctypes.c_uint64]
    L.tg_result_tsv_noalloc.restype = ctypes.c_int64
    L.tg_result_tsv_noalloc.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_char_p,
                                        ctypes.c_uint64]
    L.tg_result_tsv.restype = ctypes.c_int64
    L.tg_result_tsv.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_char_p,
                                ctypes.c_uint64]
    L.tg_result_kff.restype = ctypes.c_int64
    L.tg_result_kff.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, _u8p,
                                ctypes.c_uint64]
    L.tg_result_write.argtypes = [_vp, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_int]
    L.tg_synth.restype = ctypes.c_int64
    L.tg_synth.argtypes = [ctypes.c_char_p]
    L.tg_synth_copy.argtypes = [_u8p, ctypes.c_uint64]
    L.tg_synth_free.restype = None
    L.tg_synth_alloc.restype = ctypes.c_int64
    L.tg_synth_alloc.argtypes = [ctypes.c_char_p, ctypes.POINTER(_u8p)]
    L.tg_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(TgParams),
                                   ctypes.POINTER(TgServerConfig), ctypes.POINTER(_vp)]
    L.tg_server_destroy.argtypes = [_vp]
    L.tg_server_destroy.restype = None
    L.tg_server_submit.argtypes = [_vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64, _u64p]
    L.tg_server_poll.argtypes = [_vp, ctypes.c_uint64]
    L.tg_server_wait.argtypes = [_vp, ctypes.c_uint64, ctypes.c_double]
    L.tg_server_result.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(_u64p), ctypes.POINTER(_u32p),
                                   _u64p, ctypes.POINTER(TgJobStats), ctypes.POINTER(TgRequestTimes)]
    L.tg_server_release.argtypes = [_vp, ctypes.c_uint64]
    L.tg_server_stats.argtypes = [_vp, ctypes.POINTER(TgServerInfo)]
    L.tg_host_alloc.argtypes = [ctypes.c_uint64, ctypes.POINTER(_vp)]
    L.tg_host_free.argtypes = [_vp]
    L.tg_set_plain_sequence.argtypes = [_vp, ctypes.c_int]
    L.tg_set_debug_stats.argtypes = [_vp, ctypes.c_int]
    L.tg_set_wide_deque.argtypes = [_vp, ctypes.c_uint32]
    L.tg_set_input_budget.argtypes = [_vp, ctypes.c_uint64]
    L.tg_debug_partitions.argtypes = [_vp, _u64p, _u64p, _u64p, _u64p, ctypes.c_uint64]
    L.tg_debug_partitions.restype = ctypes.c_int64
    L.tg_debug_min_coverage.argtypes = [_vp, ctypes.POINTER(TgParams), _u8p, ctypes.c_uint64, _u64p,
                                        ctypes.c_uint64, _u32p, _u64p, ctypes.c_uint64]
    L.tg_debug_min_coverage.restype = ctypes.c_int64
    L.tg_autotune_many.argtypes = [_vp, ctypes.POINTER(TgParams), _vp, ctypes.c_uint64, _vp,
                                   ctypes.c_uint64, _u64p, _u64p, ctypes.c_uint32,
                                   ctypes.POINTER(TgBatchConfig), ctypes.POINTER(ctypes.c_double)]
    L.tg_get_batch_config.argtypes = [_vp, ctypes.POINTER(TgBatchConfig)]
    L.tg_set_batch_config.argtypes = [_vp, ctypes.POINTER(TgBatchConfig)]
    L.tg_count_many.argtypes = [_vp, ctypes.POINTER(TgParams), _u8p, ctypes.c_uint64, _u64p,
                                ctypes.c_uint64, _u64p, ctypes.c_uint32, ctypes.POINTER(TgJobStats),
                                ctypes.POINTER(TgManyStats)]
    L.tg_count_many_device.argtypes = [_vp, ctypes.POINTER(TgParams), _vp, ctypes.c_uint64, _vp,
                                       ctypes.c_uint64, _u64p, _u64p, ctypes.c_uint32,
                                       ctypes.POINTER(TgJobStats), ctypes.POINTER(TgManyStats)]
    L.tg_many_result_size.argtypes = [_vp, ctypes.c_uint32]
    L.tg_many_result_size.restype = ctypes.c_uint64
    L.tg_many_result_copy.argtypes = [_vp, ctypes.c_uint32, _u64p, _u32p, ctypes.c_uint64]
    L.tg_many_result_device.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                        _u64p]
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def _check(rc):
    if rc != 0:
        raise TunaGpuError(f"tuna_gpu error {rc}: {lib().tg_last_error().decode()}")


def _check_size(v):
    """Entry points that return a byte count: negative values are errors."""
    if v < 0:
        _check(v)
    return v


def _u8(x):
    if isinstance(x, (bytes, bytearray)):
        return np.frombuffer(bytes(x), dtype=np.uint8)
    if isinstance(x, str):
        return np.frombuffer(x.encode(), dtype=np.uint8)
    return np.ascontiguousarray(x, dtype=np.uint8)


# ------------------------------------------------------------------ host ----
def parse_buffer(data, threads=1):
    """FASTA/FASTQ bytes -> (bases uint8, chunk offsets uint64) on the host,
    on `threads` host threads (the output does not depend on it)."""
    d = _u8(data)
    out = np.zeros(max(len(d), 1), dtype=np.uint8)
    cap = len(d) // 2 + 2
    off = np.zeros(cap, dtype=np.uint64)
    if threads > 1:
        n = lib().tg_parse_buffer_mt(_p(d, _u8p), len(d), _p(out, _u8p), _p(off, _u64p), cap, threads)
    else:
        n = lib().tg_parse_buffer(_p(d, _u8p), len(d), _p(out, _u8p), _p(off, _u64p), cap)
    if n < 0:
        raise TunaGpuError("Empty file" if n == -5 else f"parse failed ({n})")
    off = off[: n + 1].copy()
    return out[: int(off[-1])].copy(), off


def read_file(path):
    """Whole input file (gzip-decompressed for *.gz) as bytes."""
    L = lib()
    data = _u8p()
    n = ctypes.c_uint64(0)
    rc = L.tg_read_file(str(path).encode(), ctypes.byref(data), ctypes.byref(n))
    if rc != 0:
        raise TunaGpuError(L.tg_last_error().decode())
    try:
        return ctypes.string_at(data, n.value)
    finally:
        L.tg_free(data)


def parse_file(path, threads=1, section_bytes=0):
    """One input file -> (bases, chunk offsets), read as the engine reads it:
    gzip text parsed in sections of `section_bytes` while it inflates, BGZF
    members inflated on `threads` threads, plain files parsed on `threads`
    threads.  Equal to parse_buffer(read_file(path)) for every file."""
    L = lib()
    b = _u8p()
    n = ctypes.c_uint64(0)
    o = _u64p()
    nch = L.tg_parse_file(str(path).encode(), threads, section_bytes, ctypes.byref(b), ctypes.byref(n),
                          ctypes.byref(o))
    if nch < 0:
        raise TunaGpuError(L.tg_last_error().decode())
    try:
        bases = np.frombuffer(ctypes.string_at(b, n.value), dtype=np.uint8).copy()
        off = np.ctypeslib.as_array(o, shape=(nch + 1,)).copy()
    finally:
        L.tg_free(b)
        L.tg_free(o)
    return bases, off


def synth(spec):
    """Seeded synthetic FASTA/FASTQ text (see tg_synth in include/tuna_gpu.h)."""
    L = lib()
    n = L.tg_synth(spec.encode())
    if n < 0:
        raise TunaGpuError(f"bad synth spec {spec!r}")
    buf = np.empty(n, dtype=np.uint8)
    _check(L.tg_synth_copy(_p(buf, _u8p), n))
    L.tg_synth_free()
    return buf


def synth_parsed(spec):
    """Generate and parse one synthetic input -> (bases, off); thread-safe, so
    many samples can be generated in parallel threads."""
    L = lib()
    ptr = _u8p()
    n = L.tg_synth_alloc(spec.encode(), ctypes.byref(ptr))
    if n < 0:
        raise TunaGpuError(f"bad synth spec {spec!r}")
    try:
        text = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n]
        return parse_buffer(text)
    finally:
        L.tg_free(ptr)


class _Pinned:
    def __init__(self, nbytes):
        self.p = _vp()
        _check(lib().tg_host_alloc(max(nbytes, 1), ctypes.byref(self.p)))

    def __del__(self):
        if self.p:
            lib().tg_host_free(self.p)
            self.p = None


def format_tsv(k, keys, cnt):
    """TSV lines of (key, count) pairs formatted on the host (tg_format_tsv)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1)
    cnt = np.ascontiguousarray(cnt, dtype=np.uint32)
    n = len(cnt)
    buf = np.zeros(max(n, 1) * (k + 12), dtype=np.uint8)
    got = _check_size(lib().tg_format_tsv(k, _p(keys, _u64p), _p(cnt, _u32p), n,
                                          buf.ctypes.data_as(ctypes.c_char_p), len(buf)))
    return buf[:got].tobytes()


def pinned_copy(arr):
    """A copy of arr in pinned host memory (faster uploads); the memory lives as
    long as the returned array."""
    arr = np.ascontiguousarray(arr)
    owner = _Pinned(arr.nbytes)
    buf = (ctypes.c_uint8 * max(arr.nbytes, 1)).from_address(owner.p.value)
    buf.owner = owner  # the array keeps buf alive, buf keeps the pinned block
    out = np.frombuffer(buf, dtype=np.uint8, count=arr.nbytes).view(arr.dtype).reshape(arr.shape)
    out[...] = arr
    return out


class Server:
    """Request loop with continuous batching (tg_server_* in include/tuna_gpu.h)."""

    def __init__(self, k, m=21, ci=1, cx=2**64 - 1, lanes=2, max_batch_jobs=64, max_batch_bases=0,
                 hbm_budget=0, device=0):
        self.k = k
        self.nw = key_words(k)
        p = _params(k, m, 16, ci, cx)
        cfg = TgServerConfig(lanes, max_batch_jobs, max_batch_bases, hbm_budget)
        self._h = _vp()
        _check(lib().tg_server_create(device, ctypes.byref(p), ctypes.byref(cfg), ctypes.byref(self._h)))
        self._keep = {}

    def close(self):
        if self._h:
            lib().tg_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, bases, off):
        bases = bases if isinstance(bases, np.ndarray) and bases.dtype == np.uint8 else _u8(bases)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        rid = ctypes.c_uint64(0)
        _check(lib().tg_server_submit(self._h, bases.ctypes.data_as(_vp), len(bases),
                                      off.ctypes.data_as(_vp), len(off) - 1, ctypes.byref(rid)))
        self._keep[rid.value] = (bases, off)
        return rid.value

    def poll(self, rid):
        rc = lib().tg_server_poll(self._h, rid)
        if rc < 0:
            _check(rc)
        return rc == 1

    def wait(self, rid, timeout=-1.0):
        rc = lib().tg_server_wait(self._h, rid, timeout)
        if rc == TG_ETIMEOUT:
            return False
        _check(rc)
        return True

    def result(self, rid, copy=True):
        kp, cp = _u64p(), _u32p()
        n = ctypes.c_uint64(0)
        st, t = TgJobStats(), TgRequestTimes()
        _check(lib().tg_server_result(self._h, rid, ctypes.byref(kp), ctypes.byref(cp), ctypes.byref(n),
                                      ctypes.byref(st), ctypes.byref(t)))
        n = n.value
        if not copy:
            return None, None, st.as_dict(), t.as_dict()
        keys = np.ctypeslib.as_array(kp, shape=(max(n, 1) * self.nw,))[: n * self.nw].copy()
        cnt = np.ctypeslib.as_array(cp, shape=(max(n, 1),))[:n].copy()
        return keys.reshape(n, self.nw), cnt, st.as_dict(), t.as_dict()

    def release(self, rid):
        _check(lib().tg_server_release(self._h, rid))
        self._keep.pop(rid, None)

    def stats(self):
        i = TgServerInfo()
        _check(lib().tg_server_stats(self._h, ctypes.byref(i)))
        return i.as_dict()


# --------------------------------------------------------------- options ----
@dataclass
class Options:
    """Mirror of tuna::Options (include/tuna/tuna.hpp:59-66)."""
    threads: int = 1
    min_count: int = 1
    max_count: int = 2**32 - 1
    partitions: int = 0
    ram_gb: int = 0
    work_dir: str = ""


def _params(k, m, partitions=16, ci=1, cx=2**64 - 1):
    return TgParams(k=k, m=m, n_partitions=partitions or 16, ci=ci, cx=min(cx, 2**64 - 1))


# ---------------------------------------------------------------- engine ----
class Engine:
    """One device context: stream, resident buffers, last result."""

    def __init__(self, device=0):
        L = lib()
        h = _vp()
        _check(L.tg_ctx_create(device, ctypes.byref(h)))
        self._h = h
        self.last_stats = None

    def close(self):
        if self._h:
            lib().tg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- profiling
    def set_profiling(self, on=True):
        _check(lib().tg_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self):
        names = ctypes.create_string_buffer(8192)
        ms = (ctypes.c_double * 64)()
        ln = (ctypes.c_uint64 * 64)()
        n = lib().tg_kernel_times(self._h, names, 8192, ms, ln, 64)
        keys = names.raw.split(b"\0")[:n]
        return {keys[i].decode(): (ms[i], ln[i]) for i in range(min(n, 64))}

    def set_plain_sequence(self, on=True):
        """Run single jobs with the plain host-driven launch sequence (True)
        or through the batched engine as a batch of one (False, default)."""
        _check(lib().tg_set_plain_sequence(self._h, 1 if on else 0))

    def set_wide_deque(self, entries=16):
        """Deque entries per lane of the k > 64 superkmer pass (1..16); a strip
        that needs more is counted by the pass that keeps the whole window."""
        _check(lib().tg_set_wide_deque(self._h, int(entries)))

    def set_input_budget(self, nbytes=0):
        """Bytes of input count_files keeps on the device (0: half the free
        device memory); larger inputs are counted from host memory in slices."""
        _check(lib().tg_set_input_budget(self._h, int(nbytes)))

    def set_debug_stats(self, on=True):
        """Record table statistics per reference partition for jobs counted
        with the plain sequence (the reference's -dbg table report)."""
        _check(lib().tg_set_debug_stats(self._h, 1 if on else 0))

    def debug_partitions(self, n):
        """Per-partition statistics of the last job: k-mers inserted,
        distinct keys in the LDS tables, k-mers spilled, table slots."""
        a = [np.zeros(n, np.uint64) for _ in range(4)]
        got = lib().tg_debug_partitions(self._h, *[x.ctypes.data_as(_u64p) for x in a], n)
        _check_size(got)
        return dict(zip(("inserted", "lds_unique", "spilled", "table_cap"), [x[:got] for x in a]))

    def debug_min_coverage(self, bases, off, k, m, partitions=16, cap=1 << 16):
        """Minimizer coverage histogram: (coverage values ascending, number of
        minimizers with each)."""
        bases = _u8(bases)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        p = _params(k, m, partitions)
        cov = np.zeros(cap, np.uint32)
        cnt = np.zeros(cap, np.uint64)
        got = lib().tg_debug_min_coverage(self._h, ctypes.byref(p), _p(bases, _u8p), len(bases),
                                          _p(off, _u64p), len(off) - 1, _p(cov, _u32p), _p(cnt, _u64p), cap)
        _check_size(got)
        return cov[:got], cnt[:got]

    def reset_kernel_times(self):
        lib().tg_reset_kernel_times(self._h)

    # -- device memory
    def dev_upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = _vp()
        _check(lib().tg_dev_alloc(self._h, max(arr.nbytes, 1), ctypes.byref(p)))
        if arr.nbytes:
            _check(lib().tg_h2d(self._h, p, arr.ctypes.data_as(_vp), arr.nbytes))
        return p

    def dev_free(self, p):
        _check(lib().tg_dev_free(self._h, p))

    def sync(self):
        _check(lib().tg_sync(self._h))

    # -- counting
    def count(self, bases, off, k, m=21, partitions=16, ci=1, cx=2**64 - 1, fetch=True):
        bases = _u8(bases)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        p = _params(k, m, partitions, ci, cx)
        st = TgStats()
        _check(lib().tg_count_host(self._h, ctypes.byref(p), _p(bases, _u8p), len(bases),
                                   _p(off, _u64p), len(off) - 1, ctypes.byref(st)))
        self.last_stats = st
        return self.result(k) if fetch else st

    def count_files(self, files, k, m=21, partitions=16, ci=1, cx=2**64 - 1, threads=1, fetch=True):
        """Count FASTA/FASTQ(.gz) files read straight to the device."""
        p = _params(k, m, partitions, ci, cx)
        st = TgStats()
        names = (ctypes.c_char_p * max(len(files), 1))(*[str(f).encode() for f in files])
        _check(lib().tg_count_files(self._h, ctypes.byref(p), names, len(files), threads, ctypes.byref(st)))
        self.last_stats = st
        return self.result(k) if fetch else st

    def count_device(self, d_bases, n_bases, d_off, n_chunks, k, m=21, partitions=16, ci=1,
                     cx=2**64 - 1):
        p = _params(k, m, partitions, ci, cx)
        st = TgStats()
        _check(lib().tg_count_device(self._h, ctypes.byref(p), d_bases, n_bases, d_off,
                                     n_chunks, ctypes.byref(st)))
        self.last_stats = st
        return st

    def result(self, k):
        nw = key_words(k)
        n = lib().tg_result_size(self._h)
        keys = np.zeros(max(n, 1) * nw, dtype=np.uint64)
        cnt = np.zeros(max(n, 1), dtype=np.uint32)
        _check(lib().tg_result_copy(self._h, _p(keys, _u64p), _p(cnt, _u32p), n))
        return keys[: n * nw].reshape(n, nw), cnt[:n]

    def result_tsv(self, k, first=0, n=None):
        """TSV bytes of result entries [first, first+n), formatted on the device."""
        if n is None:
            n = lib().tg_result_size(self._h) - first
        size = _check_size(lib().tg_result_tsv(self._h, k, first, n, None, 0))
        buf = ctypes.create_string_buffer(max(size, 1))
        got = _check_size(lib().tg_result_tsv(self._h, k, first, n, buf, size))
        return buf.raw[:got]

    def result_kff(self, k, first=0, n=None):
        """KFF raw-section records of result entries [first, first+n)."""
        if n is None:
            n = lib().tg_result_size(self._h) - first
        size = _check_size(lib().tg_result_kff(self._h, k, first, n, None, 0))
        out = np.zeros(max(size, 1), dtype=np.uint8)
        got = _check_size(lib().tg_result_kff(self._h, k, first, n, _p(out, _u8p), size))
        return out[:got].tobytes()

    def result_write(self, k, path, kff=False):
        """Write the whole result as a TSV or KFF file, formatted on the device."""
        _check(lib().tg_result_write(self._h, k, os.fsencode(path), 1 if kff else 0))

    # -- many jobs per launch sequence
    def batch_config(self):
        c = TgBatchConfig()
        _check(lib().tg_get_batch_config(self._h, ctypes.byref(c)))
        return c

    def set_batch_config(self, **kw):
        c = self.batch_config()
        for key, v in kw.items():
            setattr(c, key, v)
        _check(lib().tg_set_batch_config(self._h, ctypes.byref(c)))
        return c

    def count_many(self, bases, off, job_chunk, k, m=21, ci=1, cx=2**64 - 1):
        """Count every job (chunks [job_chunk[j], job_chunk[j+1])) separately in
        shared launches; returns per-job stats (results via many_result)."""
        bases = _u8(bases)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        jc = np.ascontiguousarray(job_chunk, dtype=np.uint64)
        nj = len(jc) - 1
        p = _params(k, m, 16, ci, cx)
        st = (TgJobStats * max(nj, 1))()
        ms = TgManyStats()
        _check(lib().tg_count_many(self._h, ctypes.byref(p), _p(bases, _u8p), len(bases),
                                   _p(off, _u64p), len(off) - 1, _p(jc, _u64p), nj, st,
                                   ctypes.byref(ms)))
        self.last_many = ms
        return [st[j].as_dict() for j in range(nj)]

    def count_many_device(self, d_bases, n_bases, d_off, n_chunks, job_chunk, job_base, k, m=21,
                          ci=1, cx=2**64 - 1, stats=True):
        jc = np.ascontiguousarray(job_chunk, dtype=np.uint64)
        jb = np.ascontiguousarray(job_base, dtype=np.uint64)
        nj = len(jc) - 1
        p = _params(k, m, 16, ci, cx)
        st = (TgJobStats * max(nj, 1))()
        ms = TgManyStats()
        _check(lib().tg_count_many_device(self._h, ctypes.byref(p), d_bases, n_bases, d_off,
                                          n_chunks, _p(jc, _u64p), _p(jb, _u64p), nj, st,
                                          ctypes.byref(ms)))
        self.last_many = ms
        return [st[j].as_dict() for j in range(nj)] if stats else None

    def autotune_many(self, d_bases, n_bases, d_off, n_chunks, job_chunk, job_base, k, m=21, ci=1,
                      cx=2**64 - 1):
        """Pick and install the fastest batch configuration for these jobs."""
        jc = np.ascontiguousarray(job_chunk, dtype=np.uint64)
        jb = np.ascontiguousarray(job_base, dtype=np.uint64)
        p = _params(k, m, 16, ci, cx)
        best = TgBatchConfig()
        ms = ctypes.c_double(0)
        _check(lib().tg_autotune_many(self._h, ctypes.byref(p), d_bases, n_bases, d_off, n_chunks,
                                      _p(jc, _u64p), _p(jb, _u64p), len(jc) - 1, ctypes.byref(best),
                                      ctypes.byref(ms)))
        return best.as_dict(), ms.value

    def many_result(self, job, k):
        nw = key_words(k)
        n = lib().tg_many_result_size(self._h, job)
        keys = np.zeros(max(n, 1) * nw, dtype=np.uint64)
        cnt = np.zeros(max(n, 1), dtype=np.uint32)
        _check(lib().tg_many_result_copy(self._h, job, _p(keys, _u64p), _p(cnt, _u32p), n))
        return keys[: n * nw].reshape(n, nw), cnt[:n]

    def pack(self, bases):
        bases = _u8(bases)
        out = np.zeros((len(bases) + 15) // 16 + 1, dtype=np.uint32)
        _check(lib().tg_pack_host(self._h, _p(bases, _u8p), len(bases), _p(out, _u32p)))
        return out[: (len(bases) + 15) // 16]

    def superkmers(self, bases, off, k, m, partitions=16):
        bases = _u8(bases)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        p = _params(k, m, partitions)
        cap = max(int(off[-1]) if len(off) > 1 else 0, 1)
        st = np.zeros(cap, np.uint64)
        ln = np.zeros(cap, np.uint32)
        mp = np.zeros(cap, np.uint32)
        pid = np.zeros(cap, np.uint32)
        nsk = ctypes.c_uint64(0)
        nsp = ctypes.c_uint32(0)
        _check(lib().tg_superkmers_host(self._h, ctypes.byref(p), _p(bases, _u8p), len(bases),
                                        _p(off, _u64p), len(off) - 1, _p(st, _u64p), _p(ln, _u32p),
                                        _p(mp, _u32p), _p(pid, _u32p), cap, ctypes.byref(nsk),
                                        ctypes.byref(nsp)))
        n = nsk.value
        return st[:n], ln[:n], mp[:n], pid[:n], nsp.value

    def partitions(self, bases, off, k, m, partitions=16):
        """Records per reference partition, serialised in the reference format."""
        bases = _u8(bases)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        p = _params(k, m, partitions)
        n = partitions
        K = int(np.maximum(np.diff(off.astype(np.int64)) - k + 1, 0).sum()) if len(off) > 1 else 0
        bcap = 4 * len(bases) + 16 * K + 64
        rcap = K + 2
        part_off = np.zeros(n + 1, np.uint64)
        rb = np.zeros(bcap, np.uint8)
        ro = np.zeros(rcap, np.uint64)
        nsk = ctypes.c_uint64(0)
        nby = ctypes.c_uint64(0)
        _check(lib().tg_partitions_host(self._h, ctypes.byref(p), _p(bases, _u8p), len(bases),
                                        _p(off, _u64p), len(off) - 1, _p(part_off, _u64p),
                                        _p(rb, _u8p), bcap, _p(ro, _u64p), rcap,
                                        ctypes.byref(nsk), ctypes.byref(nby)))
        parts = []
        for P in range(n):
            recs = []
            for i in range(int(part_off[P]), int(part_off[P + 1])):
                recs.append(bytes(rb[int(ro[i]):int(ro[i + 1])]))
            parts.append(recs)
        return parts


_default = None


def default_engine():
    global _default
    if _default is None:
        _default = Engine(0)
    return _default


def kmer_label(key, k):
    """Uppercase ACGT label of a key (first base most significant)."""
    return "".join("ACGT"[(key >> (2 * (k - 1 - i))) & 3] for i in range(k))


def _read_inputs(files):
    if not files:
        raise ValueError("tuna::count: no input files")
    bs, offs, base = [], [], 0
    for f in files:
        b, o = parse_file(f)
        bs.append(b)
        offs.append(o[1:] + base if base else o[1:])
        base += len(b)
    bases = np.concatenate(bs) if bs else np.zeros(0, np.uint8)
    off = np.concatenate([np.zeros(1, np.uint64)] + offs)
    return bases, off


def count(k, m, files, cb, opts=None):
    """tuna::count<k,m>(files, cb, opts): cb(label, count) per counted k-mer."""
    opts = opts or Options()
    bases, off = _read_inputs(files)
    keys, cnt = default_engine().count(bases, off, k, m, opts.partitions or 16,
                                       opts.min_count, opts.max_count)
    for i in range(len(cnt)):
        key = int(keys[i, 0]) | (int(keys[i, 1]) << 64 if keys.shape[1] > 1 else 0)
        cb(kmer_label(key, k), int(cnt[i]))


def count_to(k, m, files, opts=None):
    """tuna::count_to<k,m>(files, opts) -> {label: count}."""
    out = {}
    count(k, m, files, lambda s, c: out.__setitem__(s, c), opts)
    return out


def count_to_raw(k, m, files, opts=None):
    """tuna::count_to_raw<k,m>(files, opts) -> (keys [n, ceil(k/32)] uint64, counts)."""
    opts = opts or Options()
    bases, off = _read_inputs(files)
    return default_engine().count(bases, off, k, m, opts.partitions or 16, opts.min_count,
                                  opts.max_count)
