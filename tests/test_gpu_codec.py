"""Parity of the HIP codec with the oracle, through the C ABI (libzero_hip.so). GPU only."""
import random
import zlib

import numpy as np
import pytest

import pyzero
from golden.corpus import corpus
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
O = Oracle()
DEPTH = {1: 8, 5: 32, 9: 256}


def split(d, bl):
    B = 1 << bl
    return [d[o: o + B] for o in range(0, len(d), B)]


def oracle_blocks(d, bl, depth):
    """[(method, raw_len, payload)] as the reference's CpuCodec writes them (the oracle is byte-identical to it)."""
    out = []
    for blk in split(d, bl):
        payload, st = O.cpu_compress_block(blk, depth)
        out.append((0, len(blk), blk) if st else (1, len(blk), payload))
    return out


@pytest.mark.parametrize("bl", [16, 18, 20])
def test_decode_reference_archives(gpu, bl):
    """every block the reference CPU codec writes decodes on the GPU to the original bytes"""
    for name, d in corpus(small=bl != 18).items():
        for depth in ((8, 32, 256) if bl == 18 else (32,)):
            blocks = oracle_blocks(d, bl, depth)
            if not blocks:
                continue
            assert gpu.decompress(blocks, bl) == d, (name, bl, depth)


@pytest.mark.parametrize("bl,level", [(16, 1), (16, 5), (18, 1), (18, 5), (18, 9), (20, 5), (20, 9)])
def test_encode_equals_model(gpu, bl, level):
    """the GPU encoder's payloads are byte-identical to the oracle's model of it, and decode (GPU and CPU)"""
    data = corpus(small=True)
    data["lz"] = pyzero.gen_lz(0, 3 << 18).tobytes()
    data["exec"] = pyzero.gen_exec(3, 3 << 18).tobytes()
    for name, d in data.items():
        if not d:
            assert gpu.compress(d, bl, level) == []
            continue
        got = gpu.compress(d, bl, level)
        blocks = split(d, bl)
        assert len(got) == len(blocks)
        for k, (blk, (method, payload)) in enumerate(zip(blocks, got)):
            want, st = O.gpu_compress_block(blk, bl, level)
            if st:
                assert method == 0, (name, k)
            else:
                assert method == 1 and payload == want, (name, k, len(payload), len(want))
                rc, back = O.decompress_block(payload, len(blk))
                assert rc == 0 and back == blk
        dec = gpu.decompress([(m, len(b), p if m else b) for b, (m, p) in zip(blocks, got)], bl)
        assert dec == d, name


def test_crc32(gpu):
    rng = random.Random(1)
    for bl in (16, 18, 20):
        for n in (0, 1, 7, (1 << bl) - 1, 1 << bl, (3 << bl) + 12345):
            d = rng.randbytes(n)
            crc = gpu.crc32_blocks_host(d, bl)
            assert crc.tolist() == [zlib.crc32(b) for b in split(d, bl)], (bl, n)


def test_crc32_ranges(gpu):
    """CRCs of byte ranges (files' pieces) of the last compress input and decompress output, against zlib"""
    rng = random.Random(3)
    d = pyzero.gen_lz(4, (5 << 18) + 777).tobytes()
    n = len(d)
    cases = [(0, 0), (0, n), (n, 0), (0, 1), (n - 1, 1), (1 << 18, 1 << 18), (1000, (1 << 18) + 3), (5, (3 << 18) + 9)]
    cases += [(o, rng.randrange(0, n - o + 1)) for o in (rng.randrange(n) for _ in range(40))]
    got = gpu.compress(d, 18, 5)
    want = [zlib.crc32(d[o: o + k]) for o, k in cases]
    assert gpu.crc32_ranges(0, cases) == want
    assert gpu.crc32_ranges(0, []) == []
    with pytest.raises(pyzero.ZeroError):
        gpu.crc32_ranges(0, [(n, 1)])
    with pytest.raises(pyzero.ZeroError):
        gpu.crc32_ranges(1, [(0, 1)])  # (no decompress output since the compress)
    back = gpu.decompress([(m, len(b), p if m else b) for b, (m, p) in zip(split(d, 18), got)], 18)
    assert back == d
    assert gpu.crc32_ranges(1, cases) == want
    with pytest.raises(pyzero.ZeroError):
        gpu.crc32_ranges(0, [(0, 1)])  # (the decompress replaced the compress input)


def lz_from_tokens(rng, n, short_groups):
    """bytes and their (match position, length, offset) tokens, built by running random tokens: offsets are mostly
    taken again from the last three tokens (repeats, chains of them), and either short (groups close after a few
    tokens) or long (groups run to 255 tokens, across many chunks of 64 tokens)"""
    out, toks, last, recent = bytearray(), [], [1, 4, 8], [1, 4, 8]
    gstart, gt = 0, 0  # (the open group, tracked only to steer the offsets towards repeats)
    while True:
        L = rng.choice([0, 0, 1, 3, 7, 15, 16, 40]) if rng.random() < 0.9 else rng.randrange(300)
        L = max(L, 1 - len(out))
        M = 4 + (rng.randrange(12) if rng.random() < 0.8 else rng.randrange(300))
        i = len(out) + L
        if i + M > n:
            break
        out += bytes(rng.randrange(256) for _ in range(L))
        if rng.random() < 0.6:
            d = rng.choice(last + recent)
        elif short_groups:
            d = rng.randrange(1, 24)
        else:
            d = rng.randrange(min(i, 2000), i + 1)
        d = max(1, min(d, i))
        if gt == 255 or (gt and i - d < len(out) - L and i - d + min(M, d) > gstart):
            gstart, gt, recent = len(out) - L, 0, last
        for k in range(M):
            out.append(out[i - d + k])
        toks.append((i, M, d))
        last = [d] + last[:2]
        gt += 1
    out += bytes(rng.randrange(256) for _ in range(n - len(out)))
    return bytes(out), toks


@pytest.mark.parametrize("bl", [16, 18])
def test_decode_repeat_chains(gpu, bl):
    """blocks full of repeated offsets (chains of repeats inside a chunk of tokens, repeats of a group that began
    chunks earlier) decode like the sequential decoder"""
    rng = random.Random(bl)
    n = 1 << bl
    blocks, want = [], b""
    for short in (True, False, True, False):
        d, toks = lz_from_tokens(rng, n, short)
        payload = O.pack_tokens(d, toks)
        rc, back = O.decompress_block(payload, n)
        assert rc == 0 and back == d
        blocks.append((1, n, payload))
        want += d
    assert gpu.decompress(blocks, bl) == want


def test_damaged_payloads(gpu):
    """damage is refused or decodes like the sequential decoder; it never faults"""
    d = corpus(small=True)["mixed"][: 1 << 16]
    payload, st = O.cpu_compress_block(d, 32)
    assert not st
    rng = random.Random(5)
    cases = []
    for _ in range(300):
        b = bytearray(payload)
        for _ in range(rng.randrange(1, 8)):
            b[rng.randrange(len(b))] = rng.randrange(256)
        cases.append(bytes(b))
    cases += [payload[:k] for k in (0, 5, 19, 20, 25, len(payload) // 2, len(payload) - 1)]
    blocks = [(1, len(d), c) for c in cases]
    with pytest.raises(pyzero.DamagedError):
        gpu.decompress(blocks, 16)
    err = gpu.last_err
    for k, c in enumerate(cases):
        rc, back = O.decompress_block(c, len(d))
        if rc != 0:
            assert err[k] != 0, k  # the sequential decoder refuses it: so does the GPU
        elif err[k] == 0:
            assert gpu.decompress([(1, len(d), c)], 16) == back, k
        else:
            assert err[k] == pyzero.ZGPU_ERR_DAMAGED or err[k] & 0x40, (k, err[k])  # only the group rule adds refusals


def test_stored_and_method_errors(gpu):
    d = bytes(range(256)) * 256
    assert gpu.decompress([(0, len(d), d)], 16) == d
    with pytest.raises(pyzero.DamagedError):  # a block larger than the block size
        gpu.decompress([(0, len(d) + 1, d + b"x")], 16)
    with pytest.raises(pyzero.DamagedError):
        gpu.decompress([(0, len(d), d[:-1])], 16)
    with pytest.raises(pyzero.DamagedError):
        gpu.decompress([(2, 10, b"0123456789")], 16)


def test_large_roundtrip_device_resident(gpu):
    """64 MiB of the LZ workload through device-resident encode and decode; a sample decoded by the oracle"""
    bl, level = 18, 5
    n = 64 << 20
    d = pyzero.gen_lz(1, n)
    B = 1 << bl
    nb = n // B
    din = gpu.alloc(n)
    din.upload(d)
    tab = pyzero.block_table([B] * nb, src=np.arange(nb, dtype=np.uint64) * B, dst=np.arange(nb, dtype=np.uint64) * B)
    dtab = gpu.alloc(tab.nbytes)
    dtab.upload(tab)
    dout = gpu.alloc(nb * B)
    dlen = gpu.alloc(nb * 4)
    gpu.encode_blocks(din.ptr, dtab.ptr, nb, bl, level, dout.ptr, dlen.ptr)
    gpu.sync()
    lens = dlen.download(dtype=np.uint32)
    slabs = dout.download()
    assert (lens != pyzero.STORED).all()
    comp = [slabs[b * B: b * B + int(lens[b])].tobytes() for b in range(nb)]
    ratio = sum(map(len, comp)) / n
    assert 0.2 < ratio < 0.6, ratio
    for b in (0, 1, nb // 2, nb - 1):
        rc, back = O.decompress_block(comp[b], B)
        assert rc == 0 and back == d[b * B:(b + 1) * B].tobytes()
        want, _ = O.gpu_compress_block(d[b * B:(b + 1) * B], bl, level)
        assert comp[b] == want
    # decode on the device from a packed blob
    blob = np.frombuffer(b"".join(comp), dtype=np.uint8)
    dt = pyzero.block_table([B] * nb, comp_lens=[len(c) for c in comp], methods=[1] * nb)
    dcomp = gpu.alloc(blob.size)
    dcomp.upload(blob)
    dtab.upload(dt)
    draw = gpu.alloc(n)
    derr = gpu.alloc(nb * 4)
    gpu.decode_blocks(dcomp.ptr, dtab.ptr, nb, bl, draw.ptr, derr.ptr)
    gpu.sync()
    assert (derr.download(dtype=np.uint32) == 0).all()
    assert np.array_equal(draw.download(), d)
    crc = gpu.alloc(nb * 4)
    gpu.crc32_blocks(draw.ptr, dtab.ptr, nb, True, crc.ptr)
    gpu.sync()
    got = crc.download(dtype=np.uint32)
    assert got[0] == zlib.crc32(d[:B].tobytes()) and got[-1] == zlib.crc32(d[-B:].tobytes())
    for buf in (din, dtab, dout, dlen, dcomp, draw, derr, crc):
        buf.free()
