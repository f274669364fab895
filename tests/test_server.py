"""Serving loop: every request's result equals the oracle and is the same
whether the request runs alone or mixed with concurrent requests (bit-exact:
the tolerance for this integer path is zero)."""
import sys
import threading

import numpy as np
import pytest

import oracle as O
import tuna_amd as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _sorted(keys, cnt):
    if keys.shape[1] == 1:
        order = np.lexsort((keys[:, 0],))
    else:
        order = np.lexsort((keys[:, 0], keys[:, 1]))
    return keys[order], cnt[order]


def requests(n, seed=5, median=60000, hi=400000):
    # the serving workload's shape (own genome per request, log-normal length,
    # 150 bp reads at 15x, no errors), scaled down so the oracle stays quick
    out = []
    for i in range(n):
        spec = (f"kind=samples_fq seed={seed} index={i} median={median} sigma=1.0 lo=5000 hi={hi} "
                f"cov=15 rl=150 sub=0")
        out.append(T.synth_parsed(spec))
    return out


def check(res, b, o, k, ci):
    keys, cnt, st, _ = res
    ok, oc, tot, nu = O.count(b, o, k, ci)
    assert st["n_kmers"] == tot and st["n_unique"] == nu and st["n_out"] == len(oc)
    gk, gc = _sorted(keys, cnt)
    ok, oc = _sorted(ok, oc)
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc)
    return gk, gc


@pytest.mark.parametrize("k,m,ci", [(63, 21, 2), (31, 21, 1)])
def test_server_requests_match_oracle(k, m, ci):
    reqs = requests(10)
    say(f"server k={k} m={m} ci={ci}: {len(reqs)} requests, {sum(len(b) for b, _ in reqs)} bases")
    srv = T.Server(k, m, ci, lanes=2, max_batch_jobs=4)
    try:
        ids = [srv.submit(b, o) for b, o in reqs]
        for rid, (b, o) in zip(ids, reqs):
            assert srv.wait(rid, 120.0)
            check(srv.result(rid), b, o, k, ci)
            srv.release(rid)
        info = srv.stats()
        assert info["completed"] == len(reqs) and info["failed"] == 0 and info["queued"] == 0
    finally:
        srv.close()


def test_result_same_alone_and_in_mix():
    k, m, ci = 63, 21, 2
    reqs = [(T.pinned_copy(b), T.pinned_copy(o)) for b, o in requests(12, seed=6)]
    target = 3
    srv = T.Server(k, m, ci, lanes=2, max_batch_jobs=8)
    try:
        say("alone")
        rid = srv.submit(*reqs[target])
        assert srv.wait(rid, 120.0)
        alone = check(srv.result(rid), *reqs[target], k, ci)
        t_alone = srv.result(rid, copy=False)[3]
        srv.release(rid)
        assert t_alone["batch_jobs"] == 1
        say("in the mix: 4 threads submit 12 requests at once, the target among them")
        ids = {}
        lock = threading.Lock()

        def submit(js):
            for j in js:
                r = srv.submit(*reqs[j])
                with lock:
                    ids[j] = r

        th = [threading.Thread(target=submit, args=(list(range(t, 12, 4)),)) for t in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        mixed_batch = 0
        for j, r in sorted(ids.items()):
            assert srv.wait(r, 120.0)
            res = srv.result(r)
            if j == target:
                gk, gc = _sorted(res[0], res[1])
                assert np.array_equal(gk, alone[0]) and np.array_equal(gc, alone[1])
                mixed_batch = res[3]["batch_jobs"]
            else:
                check(res, *reqs[j], k, ci)
            srv.release(r)
        say(f"target ran in a batch of {mixed_batch} requests")
    finally:
        srv.close()


def test_admission_rejects_oversized_request():
    # a budget that admits ~1 M bases per lane batch
    srv = T.Server(31, 21, 1, lanes=1, hbm_budget=64 * (1 << 20))
    try:
        info = srv.stats()
        assert 0 < info["lane_bases"] <= (1 << 20)
        big = T.synth_parsed("kind=genome_fa len=2000000 seed=3")
        with pytest.raises(T.TunaGpuError, match="admission"):
            srv.submit(*big)
        assert srv.stats()["rejected"] == 1
        small = T.synth_parsed("kind=genome_fa len=200000 seed=4")
        rid = srv.submit(*small)
        assert srv.wait(rid, 60.0)
        check(srv.result(rid), *small, 31, 1)
        srv.release(rid)
    finally:
        srv.close()


def test_wait_poll_release_semantics():
    srv = T.Server(21, 15, 1, lanes=1)
    try:
        b, o = T.synth_parsed("kind=genome_fa len=300000 seed=8")
        rid = srv.submit(b, o)
        assert srv.wait(rid, 60.0)
        assert srv.poll(rid)
        srv.release(rid)
        with pytest.raises(T.TunaGpuError):
            srv.poll(rid)
        with pytest.raises(T.TunaGpuError):
            srv.submit(b, o[:-1])  # offsets do not end at n_bases
    finally:
        srv.close()


def test_requests_at_serving_workload_size():
    # requests of the serving workload at its stated size (own genome per
    # request, log-normal median 1 Mbp, 150 bp reads at 15x; k=63 m=21 ci=2):
    # the three smallest of the first 12 (each ~10-20 Mbases of reads) against
    # the oracle, and every one against the closed-form k-mer total
    k, m, ci = 63, 21, 2
    reqs = [T.synth_parsed(f"kind=samples_fq seed=5 index={i} median=1000000 sigma=1.0 lo=10000 hi=20000000 "
                           f"cov=15 rl=150 sub=0") for i in range(12)]
    say(f"{len(reqs)} requests, {sum(len(b) for b, _ in reqs)} bases")
    srv = T.Server(k, m, ci, lanes=2, max_batch_jobs=8)
    try:
        ids = [srv.submit(b, o) for b, o in reqs]
        small = set(np.argsort([len(b) for b, _ in reqs])[:3].tolist())
        for i, (rid, (b, o)) in enumerate(zip(ids, reqs)):
            assert srv.wait(rid, 300.0)
            res = srv.result(rid)
            L = np.diff(o.astype(np.int64))
            assert res[2]["n_kmers"] == int(np.maximum(L - k + 1, 0).sum())
            assert res[2]["n_out"] == len(res[1])
            if i in small:
                say(f"  request {i}: {len(b)} bases against the oracle")
                check(res, b, o, k, ci)
            srv.release(rid)
    finally:
        srv.close()
