"""tg_result_tsv_noalloc: the TSV entry point that allocates nothing (its
device scratch is a fixed array per GPU).  Called through the C ABI and
compared byte for byte with tg_result_tsv on the same entries, two calls in a
row with the second smaller, then again after another job; sizes beyond its
limit are refused."""
import ctypes
import sys

import numpy as np
import pytest

import tuna_amd as T

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 18


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def noalloc(eng, k, first, n):
    L = T.lib()
    L.tg_result_tsv_noalloc.restype = ctypes.c_int64
    L.tg_result_tsv_noalloc.argtypes = [T._vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                        ctypes.c_char_p, ctypes.c_uint64]
    size = L.tg_result_tsv_noalloc(eng._h, k, first, n, None, 0)
    if size < 0:
        raise T.TunaGpuError(L.tg_last_error().decode())
    buf = ctypes.create_string_buffer(max(size, 1))
    got = L.tg_result_tsv_noalloc(eng._h, k, first, n, buf, size)
    assert got == size
    return buf.raw[:got]


@pytest.mark.parametrize("k,m,ci", [(31, 21, 1), (63, 21, 1)])
def test_noalloc_tsv_equals_allocating_entry_point(k, m, ci):
    eng = T.Engine(0)
    try:
        b, o = T.synth_parsed("kind=reads_fq genome=400000 seed=31 cov=4 rl=150 sub=0.01 nrate=0.001")
        st = eng.count(b, o, k, m, 16, ci, fetch=False)
        n = st.n_out
        say(f"k={k}: result of {n} k-mers")
        assert n > NOALLOC_MAX
        first = 1000
        a = noalloc(eng, k, first, NOALLOC_MAX)
        assert a == eng.result_tsv(k, first, NOALLOC_MAX)
        assert a.count(b"\n") == NOALLOC_MAX
        # the second call is smaller and reuses the same scratch
        b2 = noalloc(eng, k, 17, 1000)
        assert b2 == eng.result_tsv(k, 17, 1000)
        assert noalloc(eng, k, n - 5, 5) == eng.result_tsv(k, n - 5, 5)
        assert noalloc(eng, k, 0, 0) == b""
        with pytest.raises(T.TunaGpuError, match="at most 262144 lines"):
            noalloc(eng, k, 0, NOALLOC_MAX + 1)
        # another job on the same context: the scratch keeps nothing from before
        b3, o3 = T.synth_parsed("kind=genome_fa len=50000 seed=32")
        st3 = eng.count(b3, o3, k, m, 16, 1, fetch=False)
        assert noalloc(eng, k, 0, st3.n_out) == eng.result_tsv(k, 0, st3.n_out)
    finally:
        eng.close()
