"""Writes tests/golden/golden.json from the reference's own CPU codec (oracle/_ref/libzeroref.so, built by
oracle/Makefile from /root/reference/src). Run in the build container, where the reference is mounted:

    make -C oracle && python tests/golden/make_golden.py

Vectors: CRC-32 of fixed inputs; normalize() and encode_stream() on histograms that exercise the excess and deficit
rules; and, for every corpus input, block size 2^16/2^18/2^20 and chain depth 8/32/256 (levels 1/5/9), the SHA-256
of the reference CpuCodec's output (per block: method byte, u32 length, payload).
"""
import hashlib
import json
import os
import random
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)
import numpy as np  # noqa: E402

from corpus import corpus  # noqa: E402
from oracle_lib import Ref  # noqa: E402


def block_digest(ref, data, block_log, depth):
    B = 1 << block_log
    h = hashlib.sha256()
    total = 0
    stored = 0
    for off in range(0, len(data), B):
        blk = data[off: off + B]
        payload, st = ref.compress_block(blk, depth)
        if st or len(payload) >= len(blk):
            h.update(struct.pack("<BI", 0, len(blk)))
            total += len(blk)
            stored += 1
        else:
            h.update(struct.pack("<BI", 1, len(payload)) + payload)
            total += len(payload)
    return h.hexdigest(), total, stored


def stream_cases():
    rng = random.Random(7)
    cases = [b"a", b"ab", b"a" * 1000, bytes(range(256)), bytes(range(256)) * 3 + b"\0" * 5000]
    cases.append(bytes(rng.choice([0, 1, 2, 3, 200]) for _ in range(3000)))
    cases.append(bytes(min(255, int(rng.expovariate(0.05))) for _ in range(20000)))  # skewed: excess rule
    cases.append(bytes([0]) * 4000 + bytes(range(1, 256)))  # many 1-count symbols: excess taken from the max
    cases.append(bytes([5]) * 3 + bytes([9]) * 3 + bytes([7]) * 3)  # deficit to the first max (a tie)
    cases.append(bytes(rng.randrange(256) for _ in range(70000)))  # near uniform: raw is smaller
    cases.append(bytes(rng.randrange(40) for _ in range(64)))
    cases.append(bytes(rng.randrange(40) for _ in range(65)))
    cases.append(bytes(rng.randrange(3) for _ in range(100000)))
    return cases


def main():
    ref = Ref()
    g = {"crc": [], "normalize": [], "streams": [], "blocks": []}
    rng = random.Random(11)
    for n in (0, 1, 3, 8, 9, 100, 4096, 70001):
        data = rng.randbytes(n)
        g["crc"].append({"hex": data.hex() if n <= 100 else None, "seed_n": [11, n], "crc": ref.crc32(data)})
    g["crc"].append({"hex": b"123456789".hex(), "seed_n": None, "crc": ref.crc32(b"123456789")})
    for s in stream_cases():
        counts = np.bincount(np.frombuffer(s, dtype=np.uint8), minlength=256).astype(np.uint32)
        g["normalize"].append({"counts": counts.tolist(), "freq": ref.normalize(counts).tolist()})
        sec = ref.encode_stream(s)
        g["streams"].append({"sym": s.hex() if len(s) <= 5000 else None,
                             "sym_sha256": hashlib.sha256(s).hexdigest(), "n": len(s),
                             "section_sha256": hashlib.sha256(sec).hexdigest(), "section_len": len(sec),
                             "section": sec.hex() if len(sec) <= 3000 else None})
    # the large stream cases are regenerated by the tests from the same function
    data = corpus()
    for name, d in data.items():
        for bl in (16, 18, 20):
            for depth in (8, 32, 256):
                dig, total, stored = block_digest(ref, d, bl, depth)
                g["blocks"].append({"set": name, "block_log": bl, "depth": depth, "n": len(d),
                                    "input_sha256": hashlib.sha256(d).hexdigest(), "sha256": dig,
                                    "comp": total, "stored": stored})
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(g, f, indent=0)
    print("wrote", os.path.join(HERE, "golden.json"), len(g["blocks"]), "block vectors")


if __name__ == "__main__":
    main()
