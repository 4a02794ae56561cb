"""Parity of the HIP codec with the oracle, through the C ABI (libzero_hip.so). GPU only."""
import random
import zlib

import numpy as np
import pytest

import hashlib
import json
import os
import struct

import pyzero
from golden.corpus import corpus
from golden.make_golden import stream_cases
from oracle_lib import Oracle, Ref, ref_available

pytestmark = pytest.mark.gpu
O = Oracle()
R = Ref() if ref_available() else None
DEPTH = {1: 8, 5: 32, 9: 256}
GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")))


def sections(payload):
    """the five stream sections of a ZB2 payload with their symbol counts (src/codec2.h:140-151), and the offset bits"""
    nt, nl, ne, nob, ng = struct.unpack_from("<5I", payload)
    out, pos = [], 20
    for cnt in (nl, nt, ne, nt, ng):
        ln = struct.unpack_from("<I", payload, pos + 1)[0]
        out.append((payload[pos: pos + 5 + ln], cnt))
        pos += 5 + ln
    assert len(payload) == pos + nob
    return out


def check_payload_against_reference(payload, blk):
    """the reference CPU decoder gives blk back, and the reference's encode_stream writes every one of the payload's
    five sections byte for byte from the symbols they hold"""
    if R is None:
        return
    rc, back = R.decompress_block(payload, len(blk))
    assert rc == 0 and back == blk
    for sec, cnt in sections(payload):
        r, syms = O.decode_stream(sec, cnt)
        assert r == len(sec)
        assert R.encode_stream(syms) == sec


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


@pytest.mark.parametrize("bl,level", [(16, 1), (16, 5), (16, 9), (18, 1), (18, 5), (18, 9), (20, 1), (20, 5), (20, 9)])
def test_encode_equals_model(gpu, bl, level):
    """the GPU encoder's payloads are byte-identical to the oracle's model of it, decode under the oracle and the
    reference's own decoder, and hold exactly the sections the reference's encode_stream writes for their symbols"""
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
                check_payload_against_reference(payload, blk)
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


# the GPU's error bits for each way the sequential decoder (oracle, = src/codec2.h:155-203) refuses a payload: its
# header (-1), its counts (-2), a stream section or the offset bits (-3, -4), the tokens (-5 .. -10). The GPU also
# checks the counts against the block's caps (E_COUNTS) before the streams, and the stream decoder runs beside the
# replay's checks, so a refusal may carry those bits too.
STREAM_BITS = pyzero.E_STREAM | pyzero.E_TABLE | pyzero.E_LANE
ALLOWED = {-1: pyzero.E_HDR, -2: pyzero.E_COUNTS, -3: STREAM_BITS | pyzero.E_COUNTS, -4: STREAM_BITS | pyzero.E_COUNTS}
for _rc in range(-10, -4):
    ALLOWED[_rc] = pyzero.E_TOKEN | pyzero.E_GROUP | pyzero.E_COUNTS


@pytest.mark.parametrize("bl", [16, 18, 20])
def test_damaged_payloads(gpu, bl):
    """damage at 2^16 / 2^18 / 2^20 in payloads of the GPU encoder at levels 1, 5, 9 and of the CPU encoder: every
    payload the sequential decoder refuses, the GPU refuses with the bits of that refusal; one it accepts decodes to
    the same bytes, or is refused by the group rule alone; nothing faults"""
    d = pyzero.gen_lz(7, 1 << bl).tobytes()
    rng = random.Random(bl)
    originals = [O.gpu_compress_block(d, bl, lv)[0] for lv in (1, 5, 9)] + [O.cpu_compress_block(d, 32)[0]]
    cases = []
    for payload in originals:
        for _ in range(60):
            b = bytearray(payload)
            for _ in range(rng.randrange(1, 8)):
                if rng.random() < 0.3:  # the header and section headers, where the counts and lengths are
                    b[rng.randrange(min(len(b), 64))] = rng.randrange(256)
                else:
                    b[rng.randrange(len(b))] = rng.randrange(256)
            cases.append(bytes(b))
        cases += [payload[:k] for k in (0, 5, 19, 20, 25, len(payload) // 2, len(payload) - 1)]
    blocks = [(1, len(d), c) for c in cases]
    with pytest.raises(pyzero.DamagedError):
        gpu.decompress(blocks, bl)
    err = [int(e) for e in gpu.last_err]
    refused_by_both = 0
    for k, c in enumerate(cases):
        rc, back = O.decompress_block(c, len(d))
        if rc != 0:
            assert err[k] != 0 and err[k] & ~ALLOWED[rc] == 0, (k, rc, hex(err[k]))
            refused_by_both += 1
        elif err[k] == 0:
            assert gpu.decompress([(1, len(d), c)], bl) == back, k
        else:
            assert err[k] == pyzero.E_GROUP, (k, hex(err[k]))  # only the group rule adds refusals
    assert refused_by_both > len(cases) // 4


def test_stored_and_method_errors(gpu):
    d = bytes(range(256)) * 256
    assert gpu.decompress([(0, len(d), d)], 16) == d
    with pytest.raises(pyzero.DamagedError):  # a block larger than the block size
        gpu.decompress([(0, len(d) + 1, d + b"x")], 16)
    with pytest.raises(pyzero.DamagedError):
        gpu.decompress([(0, len(d), d[:-1])], 16)
    with pytest.raises(pyzero.DamagedError):
        gpu.decompress([(2, 10, b"0123456789")], 16)


def _hash5(v, w):
    """the 5-byte hash of kernels/zero.cl:49-58 (v: the first 4 bytes little endian, w: the fifth)"""
    return ((v * 2654435761) ^ (w * 0x9E3779B1 + (w >> 3))) & 0xFFFFFFFF


@pytest.mark.parametrize("bl", [16, 18])
def test_encode_large_buckets(gpu, bl):
    """buckets of more than 4096 positions: a zero run with a few other keys in its bucket (sorted through a bitmap of
    positions) and with many (a 2-byte pattern whose windows 'ababa' fall in the zero run's bucket under another
    tag: the radix sort); payloads equal the oracle's model at levels 1, 5 and 9"""
    hb = bl - 3
    hz = _hash5(0, 0)
    hp = _hash5(16 | 246 << 8 | 16 << 16 | 246 << 24, 16)
    assert hz >> (32 - hb) == hp >> (32 - hb)  # one bucket
    assert (hz >> (20 - hb)) & 0xFFF != (hp >> (20 - hb)) & 0xFFF  # two tags
    B = 1 << bl
    rng = np.random.default_rng(bl)
    pat = np.tile(np.array([16, 246], np.uint8), 6000)
    blocks = []
    for few in (True, False):
        d = rng.integers(0, 256, B, dtype=np.uint8)
        d[1000:1000 + 24000] = 0
        d[30000:30000 + (600 if few else pat.size)] = pat[: 600 if few else pat.size]
        blocks.append(d.tobytes())
    data = b"".join(blocks)
    for level in (1, 5, 9):
        got = gpu.compress(data, bl, level)
        assert len(got) == len(blocks)
        for k, (blk, (method, payload)) in enumerate(zip(blocks, got)):
            want, st = O.gpu_compress_block(blk, bl, level)
            assert st == 0 and method == 1 and payload == want, (level, k)
        assert gpu.decompress([(m, len(b), p) for b, (m, p) in zip(blocks, got)], bl) == data


@pytest.mark.parametrize("bl", [18, 20])
def test_encode_buckets_past_16_bits(gpu, bl):
    """the hash kernel's bucket counters are 16-bit halves of LDS words, kept exact past 65535 by a log of the
    atomics that carried: two runs of one byte value, each far longer than 65536 positions, one counted in the low half
    of its word and one in the high half (so both a low half's carry and a whole word's wrap occur, several times at
    2^20); payloads equal the oracle's model and decode back"""
    hb = bl - 3 if bl - 3 < 15 else 15
    hz = _hash5(0, 0) >> (32 - hb)
    v = next(v for v in range(1, 256) if (_hash5(v * 0x01010101, v) >> (32 - hb)) & 1 != hz & 1)
    B = 1 << bl
    run = 100_000 if bl == 18 else 450_000
    rng = np.random.default_rng(7 + bl)
    d = rng.integers(0, 256, B, dtype=np.uint8)
    d[100:100 + run] = 0
    d[200 + run:200 + 2 * run] = v
    blk = d.tobytes()
    for level in (1, 5):
        (method, payload), = gpu.compress(blk, bl, level)
        want, st = O.gpu_compress_block(blk, bl, level)
        assert st == 0 and method == 1 and payload == want, level
        assert gpu.decompress([(method, B, payload)], bl) == blk


def test_encode_past_one_launch(gpu):
    """more than 2^31 bytes in one zgpu_encode_blocks call (8193 blocks of 2^18) runs as several launches and gives
    every block the payload it gets alone: the blocks are 128 copies of 64 blocks, so each must match its copy's
    payload from a small call"""
    bl, level = 18, 5
    B = 1 << bl
    base = pyzero.gen_lz(5, 64 * B)
    nb = 128 * 64 + 1
    din = gpu.alloc(nb * B)
    for k in range(128):
        din.upload(base, offset=k * 64 * B)
    din.upload(base[:B], offset=128 * 64 * B)
    src = np.arange(nb, dtype=np.uint64) * B
    tab = pyzero.block_table([B] * nb, src=src, dst=src)
    dtab = gpu.alloc(tab.nbytes)
    dtab.upload(tab)
    dout = gpu.alloc(nb * B)
    dlen = gpu.alloc(nb * 4)
    gpu.encode_blocks(din.ptr, dtab.ptr, nb, bl, level, dout.ptr, dlen.ptr)
    gpu.sync()
    lens = dlen.download(dtype=np.uint32)
    small = gpu.compress(base.tobytes(), bl, level)
    want = [len(p) if m else pyzero.STORED for m, p in small]
    assert [int(x) for x in lens[:64]] == want
    assert all((lens[k * 64: (k + 1) * 64] == lens[:64]).all() for k in range(128))
    assert int(lens[-1]) == want[0]
    for b in (0, 63, 8191, 8192):  # (either side of the launch boundary, and the last block)
        assert dout.download(int(lens[b]), offset=b * B).tobytes() == small[b % 64][1]
    for buf in (din, dtab, dout, dlen):
        buf.free()


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


def test_encode_stream_reference_bytes(gpu):
    """the GPU's normaliser and 64-lane rANS encoder write the reference's stream sections byte for byte: on the
    golden histograms that exercise normalize's excess and deficit tie rules (src/codec.h:41-66), on empty and
    one-symbol streams, and on random histograms of every skew"""
    cases = stream_cases()
    assert len(cases) == len(GOLD["streams"])
    for sym, v in zip(cases, GOLD["streams"]):
        sec = gpu.encode_stream(sym)
        assert len(sec) == v["section_len"] and hashlib.sha256(sec).hexdigest() == v["section_sha256"]
    rng = np.random.default_rng(8)
    extra = [b"", b"\x07", b"\x00" * 100000, bytes(range(256)) * 3]
    for n in (1, 63, 64, 65, 1000, 4095, 4096, 4097, 100000, 1 << 19):
        for a in (0.3, 1.0, 3.0):
            extra.append(np.minimum(rng.zipf(1 + a, n) - 1, 255).astype(np.uint8).tobytes())
        extra.append(rng.integers(0, 256, n, dtype=np.uint8).tobytes())
    for sym in extra:
        sec = gpu.encode_stream(sym)
        assert sec == O.encode_stream(sym), len(sym)
        if R is not None:
            assert sec == R.encode_stream(sym), len(sym)


@pytest.mark.parametrize("bl", [18, 20])
def test_decode_format_maxima(gpu, bl):
    """payloads at the format's limits decode like the sequential decoder: matches of 19 + 65535 bytes (the u16
    length escape; the CPU encoder stops at 4096), literal runs longer than 2^16 (the u24 escape), overlapping
    matches at periods 1..17 and far ones"""
    n = 1 << bl
    rng = random.Random(bl + 1)
    out, toks = bytearray(rng.randbytes(40)), []
    while True:
        M = rng.choice([65554, 65554, 65535, 40000, 19, 20, 4097, 5000])
        d = rng.choice([1, 2, 3, 4, 7, 8, 15, 16, 17, 1000, len(out)])
        d = max(1, min(d, len(out)))
        L = rng.choice([0, 1, 14, 15, 16, 300, 70000 if bl == 20 else 20000])
        if len(out) + L + M > n:
            break
        out += rng.randbytes(L)
        i = len(out)
        for k in range(M):
            out.append(out[i - d + k])
        toks.append((i, M, d))
    out += rng.randbytes(n - len(out))
    out = bytes(out)
    payload = O.pack_tokens(out, toks)
    rc, back = O.decompress_block(payload, n)
    assert rc == 0 and back == out
    if R is not None:
        rc, back = R.decompress_block(payload, n)
        assert rc == 0 and back == out
    assert max(m for _, m, _ in toks) == 65554
    assert gpu.decompress([(1, n, payload)] * 3, bl) == out * 3


def test_block_table_checks(gpu):
    """device block tables are checked by the kernels: an entry with raw_len > 2^block_log is refused (E_COUNTS on
    decode, ZGPU_BAD_BLOCK on encode) before any byte of it is touched, and the entries around it are unharmed.
    The oversized entry comes first, so that its slab is followed by another block's and the buffers hold 2^bl
    bytes more than the table names."""
    bl = 16
    B = 1 << bl
    d = pyzero.gen_lz(3, 3 * B)
    din = gpu.alloc(5 * B)
    din.upload(d)
    tab = pyzero.block_table([B + 1, B, B], src=np.array([0, B, 2 * B], dtype=np.uint64),
                             dst=np.array([0, B, 2 * B], dtype=np.uint64))
    dtab = gpu.alloc(tab.nbytes)
    dtab.upload(tab)
    dout = gpu.alloc(4 * B)
    dlen = gpu.alloc(16)
    gpu.encode_blocks(din.ptr, dtab.ptr, 3, bl, 5, dout.ptr, dlen.ptr)
    gpu.sync()
    lens = dlen.download(12, dtype=np.uint32)
    slabs = dout.download()
    assert lens[0] == pyzero.BAD_BLOCK
    for b in (1, 2):
        want, st = O.gpu_compress_block(d[b * B:(b + 1) * B].tobytes(), bl, 5)
        assert not st and slabs[b * B: b * B + int(lens[b])].tobytes() == want
    # decode: the first entry's raw_len is too large for both methods
    good = [slabs[b * B: b * B + int(lens[b])].tobytes() for b in (1, 2)]
    for method in (0, 1):
        first = d[:B].tobytes() + b"x" if method == 0 else good[0]
        blob = np.frombuffer(first + good[0] + good[1] + bytes(2 * B), dtype=np.uint8)
        cl = [len(first), len(good[0]), len(good[1])]
        t = pyzero.block_table([B + 1, B, B], comp_lens=cl, methods=[method, 1, 1],
                               dst=np.array([0, B, 2 * B], dtype=np.uint64))
        dcomp = gpu.alloc(blob.size)
        dcomp.upload(blob)
        dtab.upload(t)
        draw = gpu.alloc(5 * B)
        derr = gpu.alloc(16)
        gpu.decode_blocks(dcomp.ptr, dtab.ptr, 3, bl, draw.ptr, derr.ptr)
        gpu.sync()
        err = derr.download(12, dtype=np.uint32)
        assert err.tolist() == [pyzero.E_COUNTS, 0, 0], (method, err)
        assert draw.download(2 * B, offset=B).tobytes() == d[B:].tobytes()
        for buf in (dcomp, draw, derr):
            buf.free()
    for buf in (din, dtab, dout, dlen):
        buf.free()


def test_host_buffer_forms_in_chunks(gpu):
    """zgpu_compress / zgpu_decompress on 64 MiB at 2^16 (1024 blocks: four chunks over the two streams): payloads
    packed back to back equal the oracle's model, block CRCs equal zlib's, the round trip is exact, and the range CRCs
    cover the whole batch on the device"""
    bl, B = 16, 1 << 16
    d = pyzero.gen_lz(9, (1024 << 16) - 1234)
    d[5 * B: 9 * B] = pyzero.gen_random(10, 4 * B)  # some stored blocks among the compressed ones
    got = gpu.compress(d, bl, 5)
    assert len(got) == 1024
    assert [m for m, _ in got[5:9]] == [0, 0, 0, 0]
    rng = random.Random(4)
    for k in [0, 1, 255, 256, 511, 512, 767, 768, 1023] + rng.sample(range(1024), 12):
        blk = d[k * B:(k + 1) * B].tobytes()
        want, st = O.gpu_compress_block(blk, bl, 5)
        assert got[k] == ((0, b"") if st else (1, want)), k
    blocks = [(m, min(B, d.size - k * B), p if m else d[k * B:(k + 1) * B].tobytes()) for k, (m, p) in enumerate(got)]
    assert gpu.decompress(blocks, bl) == d.tobytes()
    assert gpu.last_crc.tolist() == [zlib.crc32(d[k * B:(k + 1) * B].tobytes()) for k in range(1024)]
    spans = [(0, d.size), (B - 7, 3 * B + 11), (d.size - 5, 5), (600 * B + 3, 100 * B)]
    assert gpu.crc32_ranges(1, spans) == [zlib.crc32(d[o: o + n].tobytes()) for o, n in spans]
