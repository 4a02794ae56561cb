"""Awkward inputs for the block codec, deterministic from a fixed seed: empty and tiny inputs, long runs, short
periods, random bytes, far copies, mixed runs/noise/repeats, inputs exactly one block long and one byte either side,
a long literal run, matches of every length (the kinds of data the reference's tests/stress.py exercises), plus the
synthetic workloads' generators. Used by make_golden.py (with the compiled reference) and by the tests."""
import random

KiB = 1 << 10
MiB = 1 << 20


def corpus(small=False):
    rng = random.Random(20261203)
    scale = 4 if small else 1
    d = {}
    d["empty"] = b""
    d["one byte"] = b"x"
    d["three bytes"] = b"abc"
    d["four bytes"] = b"abcd"
    d["five bytes"] = b"abcde"
    d["zeros"] = bytes(MiB // scale + 5)
    d["one value"] = b"\xAB" * (300 * KiB // scale)
    d["period 2"] = b"ab" * (150 * KiB // scale)
    d["period 3"] = b"abc" * (100 * KiB // scale)
    d["period 7"] = b"abcdefg" * (40 * KiB // scale)
    d["period 300"] = rng.randbytes(300) * (1000 // scale)
    d["random"] = rng.randbytes(MiB // scale)
    x = bytearray(rng.randbytes(512 * KiB // scale))
    q = len(x) // 4
    x[2 * q: 3 * q] = x[:q]
    d["far copy"] = bytes(x)
    parts = []
    total = MiB // scale
    while sum(map(len, parts)) < total:
        k = rng.randrange(5)
        n = rng.randrange(1, 20000)
        if k == 0:
            parts.append(bytes([rng.randrange(256)]) * n)
        elif k == 1:
            parts.append(rng.randbytes(n))
        elif k == 2:
            parts.append(rng.randbytes(rng.randrange(1, 9)) * (n // 8 + 1))
        elif k == 3 and parts:
            parts.append(parts[rng.randrange(len(parts))][:n])
        else:
            parts.append(b"pack my box with five dozen liquor jugs " * (n // 40 + 1))
    d["mixed"] = b"".join(parts)
    bls = (16,) if small else (16, 18, 20)
    for bl in bls:
        for delta in (-1, 0, 1):
            d["block %d %+d" % (bl, delta)] = (b"hello world " * ((1 << bl) // 12 + 1))[: (1 << bl) + delta]
    d["long literals"] = rng.randbytes(150 * KiB // scale) + b"0123456789" * 200 + rng.randbytes(150 * KiB // scale)
    b = bytearray()
    for n in list(range(4, 40)) + [100, 274, 275, 1000, 4095, 4096, 5000, 70000]:
        s = rng.randbytes(8)
        b += s + s[:4] * (n // 4)
    d["every length"] = bytes(b)
    return d
