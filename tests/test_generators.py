"""The synthetic workload generators (zero-compression_amd/csrc/gen.cpp) pinned two ways, on the CPU: the SHA-256 of
every stream's first 1 MiB against tests/golden/generators.json, and a restatement here in Python of the published
algorithms they are built from (SplitMix64 seeding, xoshiro256**, Vigna 2018) and of the workload rules, which must
give the same bytes on a prefix of each stream."""
import hashlib
import json
import math
import os

import numpy as np

import pyzero
from golden.make_generators import digests

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "generators.json")))
M64 = (1 << 64) - 1


class Xoshiro:
    """xoshiro256** (next = rotl(s1 * 5, 7) * 9), its state from four SplitMix64 outputs of the seed"""

    def __init__(self, seed):
        self.s, x = [], seed
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & M64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
            self.s.append(z ^ (z >> 31))

    @staticmethod
    def rotl(v, k):
        return ((v << k) | (v >> (64 - k))) & M64

    def next(self):
        s = self.s
        r = (self.rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self.rotl(s[3], 45)
        return r

    def u01(self):  # (0, 1]
        return ((self.next() >> 11) + 1) * (1.0 / 9007199254740992.0)

    def geometric(self, mean):
        return 1 + math.floor(math.log(self.u01()) / math.log(1.0 - 1.0 / mean))

    def normal(self):
        a, b = self.u01(), self.u01()
        return math.sqrt(-2.0 * math.log(a)) * math.cos(6.283185307179586 * b)


def zipf_cdf(s=1.1):
    w = [math.pow(r, -s) for r in range(1, 257)]
    acc, cdf = 0.0, []
    for x in w:
        acc += x
        cdf.append(acc)
    return [c / acc for c in cdf[:-1]] + [2.0]


def lz_prefix(seed, n):
    """the LZ-friendly stream: with p = 0.5 (a literal run when nothing is written yet) a literal run of
    geometric(16) Zipf(1.1) bytes, or a back-copy of min(4 + geometric(20), 258) bytes from min(1 + geometric(4096),
    2^18, written) back"""
    g, cdf, out = Xoshiro(seed), zipf_cdf(), bytearray()
    while len(out) < n:
        if not out or g.u01() <= 0.5:
            ln = min(g.geometric(16.0), n - len(out))
            for _ in range(ln):
                u = g.u01() - 1e-300
                out.append(next(r for r, c in enumerate(cdf) if c > u))
        else:
            ln = min(4 + g.geometric(20.0), 258)
            off = min(1 + g.geometric(4096.0), 1 << 18, len(out))
            ln = min(ln, n - len(out))
            for _ in range(ln):
                out.append(out[len(out) - off])
    return bytes(out)


def exec_prefix(seed, n):
    """4 KiB pages: 30 % zeros, 40 % little-endian u32 arrays with round(N(0, 64)) steps from a uniform start, 30 %
    copies of 16..512 earlier bytes from page-aligned offsets of earlier pages (random bytes on page 0)"""
    g, out, P = Xoshiro(seed), bytearray(), 4096
    page = 0
    while len(out) < n:
        ln = min(P, n - len(out))
        kind = g.u01()
        if kind <= 0.3:
            out += bytes(ln)
        elif kind <= 0.7:
            v = g.next() >> 32
            pg = bytearray()
            for _ in range(0, ln, 4):
                x = 64.0 * g.normal()
                step = int(math.copysign(math.floor(abs(x) + 0.5), x))  # (lround: halves away from zero)
                v = (v + step) & 0xFFFFFFFF
                pg += v.to_bytes(4, "little")
            out += pg[:ln]
        elif page == 0:
            out += bytes(g.next() & 0xFF for _ in range(ln))
        else:
            pg = bytearray()
            while len(pg) < ln:
                c = 16 + g.next() % 497
                src = (g.next() % page) * P
                c = min(c, ln - len(pg))
                base = len(out)
                for k in range(c):
                    pg.append(out[src + k] if src + k < base else pg[src + k - base])
            out += pg
        page += 1
    return bytes(out)


def test_generator_digests_are_pinned():
    assert digests() == GOLD


def test_prng_restatement():
    for seed in (0, 2, 12345):
        g = Xoshiro(seed)
        want = b"".join(g.next().to_bytes(8, "little") for _ in range(512))
        assert pyzero.gen_random(seed, 4096).tobytes() == want
    assert pyzero.gen_random(2, 13).tobytes() == pyzero.gen_random(2, 16).tobytes()[:13]


def test_lz_and_exec_restatements():
    for seed in (0, 1):
        assert pyzero.gen_lz(seed, 1 << 16).tobytes() == lz_prefix(seed, 1 << 16)
    assert pyzero.gen_exec(3, 20 * 4096 + 100).tobytes() == exec_prefix(3, 20 * 4096 + 100)


def test_lognormal_sizes_rule():
    g = Xoshiro(4)
    want = [min(max(int(math.floor(math.exp(math.log(8192.0) + 2.0 * g.normal()) + 0.5)), 64), 1 << 26)
            for _ in range(32)]
    assert [int(x) for x in pyzero.gen_lognormal_sizes(4, 8192.0, 2.0, 64, 1 << 26, 32)] == want
    s = pyzero.gen_lognormal_sizes(7, 8192.0, 2.0, 64, 1 << 26, 20000)
    assert 7000 < float(np.median(s)) < 9500  # (the median of the distribution, before clipping: 8 KiB)
