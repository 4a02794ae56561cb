"""The oracle (oracle/zoracle.c) pinned against the reference's golden vectors, known answers, zlib and brute force.
CPU only."""
import hashlib
import json
import os
import random
import struct
import zlib

import numpy as np
import pytest

from golden.corpus import corpus
from golden.make_golden import stream_cases
from oracle_lib import Oracle, Ref, ref_available

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))
O = Oracle()
DEPTH_OF_LEVEL = {1: 8, 5: 32, 9: 256}  # src/main.cpp:45,54


def test_crc_known_answers():
    assert O.crc32(b"123456789") == 0xCBF43926
    assert O.crc32(b"") == 0
    rng = random.Random(3)
    for n in (1, 7, 8, 9, 1000, 65537):
        d = rng.randbytes(n)
        assert O.crc32(d) == zlib.crc32(d)


def test_crc_golden():
    for v in GOLD["crc"]:
        if v["seed_n"]:
            seed, n = v["seed_n"]
        data = bytes.fromhex(v["hex"]) if v["hex"] is not None else None
        if data is None:
            rng = random.Random(11)
            for m in (0, 1, 3, 8, 9, 100, 4096, 70001):
                d = rng.randbytes(m)
                if m == n:
                    data = d
                    break
        assert O.crc32(data) == v["crc"]


def test_crc_combine():
    rng = random.Random(4)
    for _ in range(40):
        a, b = rng.randbytes(rng.randrange(0, 5000)), rng.randbytes(rng.randrange(0, 5000))
        assert O.crc32_combine(zlib.crc32(a), zlib.crc32(b), len(b)) == zlib.crc32(a + b)
    # a long B: the exponent's high bits
    a, b = b"x", bytes(3 << 20)
    assert O.crc32_combine(zlib.crc32(a), zlib.crc32(b), len(b)) == zlib.crc32(a + b)


def test_normalize_golden_and_invariants():
    for v in GOLD["normalize"]:
        counts = np.array(v["counts"], dtype=np.uint32)
        f = O.normalize(counts)
        assert f.tolist() == v["freq"]
        assert int(f.sum()) == 4096
        assert all((f[counts > 0] >= 1)) and all(f[counts == 0] == 0)


def test_encode_stream_golden():
    cases = stream_cases()
    assert len(cases) == len(GOLD["streams"])
    for s, v in zip(cases, GOLD["streams"]):
        assert hashlib.sha256(s).hexdigest() == v["sym_sha256"]
        sec = O.encode_stream(s)
        assert len(sec) == v["section_len"]
        assert hashlib.sha256(sec).hexdigest() == v["section_sha256"]
        r, back = O.decode_stream(sec, len(s))
        assert r == len(sec) and back == s


def test_encode_stream_empty_and_damage():
    assert O.encode_stream(b"") == b"\x00\x00\x00\x00\x00"
    sec = O.encode_stream(bytes(random.Random(5).randrange(8) for _ in range(5000)))
    assert sec[0] == 1
    assert O.decode_stream(sec[:-1], 5000)[0] < 0  # body cut
    bad = bytearray(sec)
    bad[5 + 32] ^= 0x40  # a frequency: the sum is no longer 4096
    assert O.decode_stream(bytes(bad), 5000)[0] < 0


@pytest.mark.parametrize("depth", [8, 32, 256])
def test_cpu_blocks_match_reference(depth):
    data = corpus()
    for v in GOLD["blocks"]:
        if v["depth"] != depth:
            continue
        d = data[v["set"]]
        assert hashlib.sha256(d).hexdigest() == v["input_sha256"], v["set"]
        B = 1 << v["block_log"]
        h = hashlib.sha256()
        total = 0
        for off in range(0, len(d), B):
            blk = d[off: off + B]
            payload, st = O.cpu_compress_block(blk, depth)
            if st:
                h.update(struct.pack("<BI", 0, len(blk)))
                total += len(blk)
            else:
                h.update(struct.pack("<BI", 1, len(payload)) + payload)
                total += len(payload)
                if v["block_log"] == 18:
                    rc, back = O.decompress_block(payload, len(blk))
                    assert rc == 0 and back == blk
        assert total == v["comp"], (v["set"], v["block_log"])
        assert h.hexdigest() == v["sha256"], (v["set"], v["block_log"])


def test_decompress_rejects_damage():
    d = corpus(small=True)["mixed"][: 1 << 16]
    payload, st = O.cpu_compress_block(d, 32)
    assert not st
    rng = random.Random(9)
    rejected = 0
    for _ in range(200):
        b = bytearray(payload)
        for _ in range(rng.randrange(1, 6)):
            b[rng.randrange(len(b))] = rng.randrange(256)
        rc, back = O.decompress_block(bytes(b), len(d))
        rejected += rc != 0
        assert rc != 0 or len(back) == len(d)
    assert rejected > 20  # (damage inside raw literal bytes decodes to other bytes: the block CRC catches it)
    assert O.decompress_block(payload[:10], len(d))[0] != 0


# ---- the GPU encoder's model against a restatement of kernels/zero.cl's k_hash_count / k_bucket_scan / k_scatter /
# k_sort / k_match in Python, with every bucket sorted completely (the determinism rule) and the byte past the block
# read as 0
def _hash5(d, p):
    v = d[p] | d[p + 1] << 8 | d[p + 2] << 16 | d[p + 3] << 24
    w = d[p + 4] if p + 4 < len(d) else 0
    return ((v * 2654435761) ^ (w * 0x9E3779B1 + (w >> 3))) & 0xFFFFFFFF


def _zero_cl_srt(d, block_log):
    """srt of one block as zero.cl lays it out: entries p | tag << 20 grouped by bucket (k_scatter,
    kernels/zero.cl:92-103), each bucket sorted by entry value (k_sort's order, :108-144, over the whole bucket)."""
    hbits = min(block_log - 3, 15)
    n = len(d)
    buckets = [[] for _ in range(1 << hbits)]
    for p in range(n - 3 if n >= 4 else 0):  # positions with 4 bytes left (:64, :98)
        hv = _hash5(d, p)
        buckets[hv >> (32 - hbits)].append(p | (((hv >> (32 - hbits - 12)) & 0xFFF) << 20))
    return [e for bk in buckets for e in sorted(bk)]


def _brute_match(d, block_log, level):
    """k_match (kernels/zero.cl:172-236) over _zero_cl_srt, statement by statement."""
    ncandl, lencap = (2, 16) if level <= 2 else (4, 16) if level <= 6 else (8, 32)
    n = len(d)
    out = [0] * n
    sb = _zero_cl_srt(d, block_log)
    count = len(sb)

    def load4u(q):
        return d[q] | d[q + 1] << 8 | d[q + 2] << 16 | d[q + 3] << 24

    def match_len(a, b, maxlen):
        k = 0
        while k < maxlen and d[a + k] == d[b + k]:
            k += 1
        return k

    for i in range(count):
        e = sb[i]
        p = e & 0xFFFFF
        jc, any_ = [], False
        for k in range(1, ncandl + 1):
            c = sb[i - k] if i >= k else ~e & 0xFFFFFFFF
            j = c & 0xFFFFF
            ok = ((c ^ e) >> 20) == 0 and j < p
            jc.append(j if ok else p)
            any_ |= ok
        best, bestd = 0, 0
        if any_:
            pw = load4u(p)
            maxlen = min(lencap, n - p)
            if p > 0 and load4u(p - 1) == pw:
                best, bestd = 4 + match_len(p + 3, p + 4, maxlen - 4), 1
            for j in jc:
                if best >= maxlen:
                    break
                if j == p or load4u(j) != pw:
                    continue
                if best >= 4 and d[j + best] != d[p + best]:
                    continue
                ln = 4 + match_len(j + 4, p + 4, maxlen - 4)
                if ln > best or (ln == best and p - j < bestd):
                    best, bestd = ln, p - j
        out[p] = (best << 20) | (bestd - 1) if best >= 4 else 0
    return out


def _brute_tokens(d, match, level):
    lencap = 16 if level <= 6 else 32
    n = len(d)

    def bl(x):
        return x.bit_length() - 1

    lazy = []
    for p in range(n):
        v, v2 = match[p], match[p + 1] if p + 1 < n else 0
        l, l2 = v >> 20, v2 >> 20
        r = v
        if l and l2 and 8 * l2 - bl((v2 & 0xFFFFF) + 1) - 8 > 8 * l - bl((v & 0xFFFFF) + 1):
            r = 0
        lazy.append(r)
    toks, p = [], 0
    while p < n:
        v = lazy[p]
        l, dist = v >> 20, (v & 0xFFFFF) + 1
        if l == lencap:
            lim = min(4095, n - p)
            while l < lim and d[p + l] == d[p - dist + l]:
                l += 1
        if l:
            toks.append((p, l, dist))
            p += l
        else:
            p += 1
    return toks


@pytest.mark.parametrize("level", [1, 5, 9])
def test_gpu_model_brute_force(level):
    rng = random.Random(level)
    inputs = [b"", b"abc", b"abcd", b"aaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaa" * 3,
              bytes(rng.randrange(4) for _ in range(1500)),
              (b"the cat sat on the mat. " * 20 + bytes(rng.randrange(256) for _ in range(300))) * 3,
              bytes(rng.choice(b"ab") for _ in range(900)) + bytes(200) + b"xyz" * 100]
    for d in inputs:
        for block_log in (16, 18):
            m = O.gpu_match(d, block_log, level)
            assert m.tolist() == _brute_match(d, block_log, level)
            assert O.gpu_tokens(d, block_log, level) == _brute_tokens(d, m.tolist(), level)


def test_gpu_model_run_check_needs_a_candidate():
    """zero.cl runs the distance-1 check only inside `if (any)` (kernels/zero.cl:209-222). At the end of a run of
    one value (bytes p-1 .. p+3 equal, p+4 another), the run check alone gives a match of 4 at distance 1: it must
    give nothing when no earlier slot holds the position's tag, and decide when the only candidates are tag
    collisions whose bytes differ."""
    rng = random.Random(21)
    parts = []
    for _ in range(7000):  # runs of exactly 5 followed by a random byte, among random bytes
        q = rng.randrange(256)
        parts.append(bytes([q]) * 5 + bytes([(q + 1 + rng.randrange(255)) % 256]))
        parts.append(bytes(rng.randrange(256) for _ in range(rng.randrange(4))))
    d = b"".join(parts)[: 1 << 16]
    n = len(d)
    sb = _zero_cl_srt(d, 16)
    slot = {e & 0xFFFFF: i for i, e in enumerate(sb)}
    seen = {"none": 0, "collision": 0}
    for level in (1, 5, 9):
        ncand = 2 if level <= 2 else 4 if level <= 6 else 8
        m = O.gpu_match(d, 16, level)
        for p in range(1, n - 4):
            if not (d[p - 1: p + 3] == d[p: p + 4] and d[p + 4] != d[p + 3]):
                continue
            i = slot[p]
            cands = [sb[i - k] & 0xFFFFF for k in range(1, ncand + 1)
                     if i >= k and ((sb[i - k] ^ sb[i]) >> 20) == 0 and (sb[i - k] & 0xFFFFF) < p]
            if not cands:
                assert m[p] == 0, (level, p)
                seen["none"] += 1
            elif all(d[j: j + 4] != d[p: p + 4] for j in cands):
                assert m[p] == 4 << 20, (level, p)
                seen["collision"] += 1
    assert seen["none"] > 100 and seen["collision"] > 0, seen


def test_pack_back_half_matches_reference_encoder():
    # the CPU parse's tokens through the shared back half equal the reference's whole compress_block
    d = corpus(small=True)["mixed"][: 1 << 17]
    payload, _ = O.cpu_compress_block(d, 32)
    assert O.pack_tokens(d, O.cpu_tokens(d, 32)) == payload


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (the reference is not mounted)")
def test_oracle_equals_reference_on_random_inputs():
    ref = Ref()
    rng = random.Random(12)
    for _ in range(30):
        kind = rng.randrange(3)
        n = rng.randrange(0, 200000)
        if kind == 0:
            d = bytes(rng.randrange(1 + rng.randrange(255)) for _ in range(n // 8))
        elif kind == 1:
            d = bytes(rng.randbytes(rng.randrange(1, 50))) * (n // 50 + 1)
        else:
            d = bytes(np.frombuffer(rng.randbytes(n), dtype=np.uint8) & 0x0F)
        for depth in (8, 32):
            assert O.cpu_compress_block(d, depth) == ref.compress_block(d, depth)
        payload, _ = O.gpu_compress_block(d, 18, 5)
        rc, back = ref.decompress_block(payload, len(d))
        assert rc == 0 and back == d
