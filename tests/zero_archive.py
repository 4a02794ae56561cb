"""A .zero container writer and reader for tests (format: reference src/archive.h:1-11), with block payloads from
the oracle (byte-identical to the reference CPU codec). Test infrastructure: it checks the product's CLI in both
directions."""
import os
import struct
import zlib

from oracle_lib import Oracle

O = Oracle()


def write_archive(path, entries, block_log=18, depth=32, file_crc=None):
    """entries: [(name, bytes) for a file, (name, None) for a folder], in stream order; file_crc: {name: crc} to
    store instead of the right CRC (a table that disagrees with intact blocks)"""
    B = 1 << block_log
    stream = b"".join(d for _, d in entries if d is not None)
    nblocks = (len(stream) + B - 1) // B
    out = bytearray(b"ZERO" + bytes([2, block_log, 0, 0]) + struct.pack("<QQ", len(stream), nblocks))
    comp_total = 0
    for off in range(0, len(stream), B):
        blk = stream[off: off + B]
        payload, stored = O.cpu_compress_block(blk, depth)
        if stored:
            out += struct.pack("<IIB3xI", len(blk), len(blk), 0, zlib.crc32(blk)) + blk
            comp_total += len(blk)
        else:
            out += struct.pack("<IIB3xI", len(blk), len(payload), 1, zlib.crc32(blk)) + payload
            comp_total += len(payload)
    tbl = bytearray(struct.pack("<I", len(entries)))
    for name, d in entries:
        nb = name.encode()
        tbl += struct.pack("<B3xI", 1 if d is None else 0, len(nb)) + nb
        crc = 0 if d is None else (file_crc or {}).get(name, zlib.crc32(d))
        tbl += struct.pack("<QqII", 0 if d is None else len(d), 1700000000, 0, crc)
    table_off = 24 + comp_total + 16 * nblocks
    assert table_off == len(out)
    out += tbl + struct.pack("<QQI", table_off, len(tbl), zlib.crc32(bytes(tbl))) + b"ZEND"
    with open(path, "wb") as f:
        f.write(out)


def read_archive(path):
    """-> (block_log, [(name, bytes or None)]) decoded with the oracle; raises on any damage"""
    data = open(path, "rb").read()
    assert data[:4] == b"ZERO" and data[4] == 2
    block_log = data[5]
    total, nblocks = struct.unpack_from("<QQ", data, 8)
    table_off, table_len, table_crc = struct.unpack_from("<QQI", data, len(data) - 24)
    assert data[-4:] == b"ZEND" and zlib.crc32(data[table_off: table_off + table_len]) == table_crc
    pos = 24
    stream = bytearray()
    for _ in range(nblocks):
        raw_len, comp_len, method, crc = struct.unpack_from("<IIB3xI", data, pos)
        pos += 16
        body = data[pos: pos + comp_len]
        pos += comp_len
        if method == 0:
            assert comp_len == raw_len
            blk = body
        else:
            rc, blk = O.decompress_block(body, raw_len)
            assert rc == 0, "oracle decoder refused a block"
        assert zlib.crc32(blk) == crc
        stream += blk
    assert pos == table_off and len(stream) == total
    t = data[table_off: table_off + table_len]
    (n,) = struct.unpack_from("<I", t, 0)
    p = 4
    entries = []
    s = 0
    for _ in range(n):
        typ, nl = struct.unpack_from("<B3xI", t, p)
        p += 8
        name = t[p: p + nl].decode()
        p += nl
        size, mtime, attr, crc = struct.unpack_from("<QqII", t, p)
        p += 24
        if typ == 1:
            entries.append((name, None))
        else:
            d = bytes(stream[s: s + size])
            s += size
            assert zlib.crc32(d) == crc
            entries.append((name, d))
    return block_log, entries
