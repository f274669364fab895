"""The tuna::Kmer<k> surface of include/tuna/tuna.hpp (the reference's
kache_hash::Kmer<k>, include/kache-hash/Kmer.hpp:29-230), checked on the CPU
against string arithmetic on the labels: construction from labels, reverse
complement, canonical form, front/back/base_at, comparisons, rolling in both
directions, prefix/suffix of a (k+1)-mer, lmer_at, as_int, packed bytes, and
the Kmer_Hasher functor inside an unordered_map."""
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

PROG = r'''
#include "tuna/tuna.hpp"
#include <cstdio>
#include <iostream>
#include <string>
#include <unordered_map>

template <uint16_t k>
int run(const std::string& a, const std::string& b, char nb, char pb) {
    using K = tuna::Kmer<k>;
    K x(a), y(b), xr = x.reverse_complement();
    std::cout << "rc " << xr << "\n";
    std::cout << "canon " << x.canonical() << " " << x.canonical(xr) << " " << *K::canonical(x, xr) << "\n";
    std::cout << "fb " << int(x.front()) << " " << int(x.back()) << " " << int(x.base_at(0)) << " "
              << int(x.base_at(k - 1)) << "\n";
    std::cout << "cmp " << (x < y) << (x > y) << (x == y) << (x != y) << (x == K(a)) << (x != K(a))
              << x.in_forward(K(a)) << "\n";
    K f = x, fr = xr;
    f.roll_to_next_kmer(nb, fr);
    std::cout << "next " << f << " " << fr << "\n";
    K g = x, gr = xr;
    g.roll_to_prev_kmer(tuna::DNA::Base((((pb >> 2) ^ (pb >> 1)) & 3)), gr);
    std::cout << "prev " << g << " " << gr << "\n";
    K h = x;
    h.roll_forward(tuna::DNA::Extended_Base::G);
    K q = x;
    q.roll_backward(tuna::DNA::Extended_Base::C);
    std::cout << "edge " << h << " " << q << "\n";
    tuna::Kmer<k + 1> e(a + std::string(1, nb));
    K pre, suf;
    pre.from_prefix(e);
    suf.from_suffix(e);
    std::cout << "ps " << pre << " " << suf << "\n";
    uint8_t bytes[(k + 3) / 4];
    x.write_packed_2bit_msb(bytes);
    std::cout << "bytes";
    for (unsigned i = 0; i < (k + 3) / 4; ++i) std::printf(" %02x", bytes[i]);
    std::fflush(stdout);
    std::cout << "\n";
    std::unordered_map<K, int, tuna::Kmer_Hasher<k>> m;
    m[x] += 1; m[K(a)] += 1; m[y] += 1; m[xr.canonical()] += 0;
    std::cout << "map " << m.size() << " " << m[x] << " " << (tuna::Kmer_Hasher<k>()(x) == x.to_u64())
              << " " << (x.to_u64(1) != x.to_u64(2)) << " " << K::num_words() << "\n";
    if constexpr (k <= 32) {
        std::cout << "int " << x.as_int() << " " << (K(x.as_int()) == x) << " " << x.template lmer_at<3>(1) << "\n";
    }
    return 0;
}

int main(int argc, char** argv) {
    const int k = std::atoi(argv[1]);
    const std::string a = argv[2], b = argv[3];
    const char nb = argv[4][0], pb = argv[5][0];
    switch (k) {
    case 5: return run<5>(a, b, nb, pb);
    case 31: return run<31>(a, b, nb, pb);
    case 32: return run<32>(a, b, nb, pb);
    case 33: return run<33>(a, b, nb, pb);
    case 63: return run<63>(a, b, nb, pb);
    case 64: return run<64>(a, b, nb, pb);
    case 127: return run<127>(a, b, nb, pb);
    }
    return 1;
}
'''

COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}
CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def rc(s):
    return "".join(COMP[c] for c in reversed(s))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("kmer")
    src = d / "k.cpp"
    src.write_text(PROG)
    out = d / "k"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", str(src), "-I", os.path.join(ROOT, "include"),
                        "-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(out)


@pytest.mark.parametrize("k", [5, 31, 32, 33, 63, 64, 127])
def test_kmer_surface_matches_label_arithmetic(exe, k):
    rng = random.Random(k)
    for trial in range(12):
        a = "".join(rng.choice("ACGT") for _ in range(k))
        b = a if trial == 0 else "".join(rng.choice("ACGT") for _ in range(k))
        if trial == 1:
            a = a[: k // 2] + rc(a[: k - k // 2])  # a prefix followed by its reverse complement
        nb, pb = rng.choice("ACGTacgt"), rng.choice("ACGT")
        low = a.lower() if trial == 2 else a
        r = subprocess.run([exe, str(k), low, b, nb, pb], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        out = {l.split(" ", 1)[0]: l.split(" ", 1)[1] for l in r.stdout.splitlines()}
        canon = min(a, rc(a))
        assert out["rc"] == rc(a)
        assert out["canon"].split() == [canon] * 3
        assert out["fb"].split() == [str(CODE[a[0]]), str(CODE[a[-1]]), str(CODE[a[-1]]), str(CODE[a[0]])]
        lt, gt, eq = a < b, a > b, a == b
        assert out["cmp"] == f"{int(lt)}{int(gt)}{int(eq)}{int(not eq)}101"
        n = nb.upper()
        assert out["next"].split() == [a[1:] + n, rc(a[1:] + n)]
        assert out["prev"].split() == [pb + a[:-1], rc(pb + a[:-1])]
        assert out["edge"].split() == [a[1:] + "G", "C" + a[:-1]]
        e = a + n
        assert out["ps"].split() == [e[:k], e[1:]]
        packed = []
        for i in range(0, k, 4):
            v = 0
            for j in range(4):
                v = (v << 2) | (CODE[a[i + j]] if i + j < k else 0)
            packed.append(f"{v:02x}")
        assert out["bytes"].split() == packed
        m = out["map"].split()
        distinct = len({a, b, canon})
        assert int(m[0]) == distinct and m[1] == ("3" if b == a else "2") and m[2] == "1" and m[3] == "1"
        assert int(m[4]) == (k + 31) // 32
        if k <= 32:
            val = 0
            for c in a:
                val = (val << 2) | CODE[c]
            iv = out["int"].split()
            lm = 0
            for c in a[1:4]:
                lm = (lm << 2) | CODE[c]
            assert iv == [str(val), "1", str(lm)]
