"""Writes tests/golden/generators.json: SHA-256 of the first 1 MiB of every synthetic workload stream and the first
lognormal sizes, so that a change to the generators (zero-compression_amd/csrc/gen.cpp), which would change every
workload and every measured number, cannot go unnoticed. Run after a deliberate change only."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "zero-compression_amd"))
import pyzero  # noqa: E402

STREAMS = [("lz", 0), ("lz", 1), ("random", 2), ("exec", 3)]


def digests():
    out = {}
    gen = {"lz": pyzero.gen_lz, "random": pyzero.gen_random, "exec": pyzero.gen_exec}
    for kind, seed in STREAMS:
        out[f"{kind}_{seed}"] = hashlib.sha256(gen[kind](seed, 1 << 20).tobytes()).hexdigest()
    out["lognormal_4"] = [int(x) for x in pyzero.gen_lognormal_sizes(4, 8192.0, 2.0, 64, 1 << 26, 32)]
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "generators.json"), "w") as f:
        json.dump(digests(), f, indent=1)
