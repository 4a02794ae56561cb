"""zgpu_decode_blocks_noalloc (the decoder with its scratch in the library's fixed array) against
zgpu_decode_blocks, through the C ABI: two calls in a row, the second with a smaller input, and the size limit.
GPU only."""
import numpy as np
import pytest

import pyzero
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
O = Oracle()


def case(seed, nb, bl, damage):
    """nb blocks of 2^bl (the last shorter): LZ blocks as the oracle's model of the GPU encoder writes them, a
    stored random block, and `damage` damaged payloads; returns the raw bytes, the blob and the table"""
    B = 1 << bl
    rng = np.random.default_rng(seed)
    raw = pyzero.gen_lz(seed, nb * B - 777)
    raw[B: 2 * B] = pyzero.gen_random(seed + 1, B)
    bodies, meth = [], []
    for b in range(nb):
        blk = raw[b * B:(b + 1) * B].tobytes()
        payload, st = O.gpu_compress_block(blk, bl, 5)
        bodies.append(bytearray(blk if st else payload))
        meth.append(0 if st else 1)
    hit = rng.choice([k for k in range(nb) if meth[k]], damage, replace=False)
    for b in hit:
        for _ in range(4):
            bodies[b][int(rng.integers(len(bodies[b])))] = int(rng.integers(256))
    rl = [min(B, raw.size - b * B) for b in range(nb)]
    tab = pyzero.block_table(rl, comp_lens=[len(x) for x in bodies], methods=meth)
    return raw, np.frombuffer(b"".join(bytes(x) for x in bodies), dtype=np.uint8), tab, set(int(b) for b in hit)


def decode_both(gpu, raw, blob, tab, bl):
    nb = tab.size
    d_comp = gpu.alloc(blob.size)
    d_comp.upload(blob)
    d_tab = gpu.alloc(tab.nbytes)
    d_tab.upload(tab)
    outs = []
    for fn in (gpu.decode_blocks, gpu.decode_blocks_noalloc):
        d_out = gpu.alloc(raw.size)
        d_err = gpu.alloc(nb * 4)
        fn(d_comp.ptr, d_tab.ptr, nb, bl, d_out.ptr, d_err.ptr)
        gpu.sync()
        outs.append((d_out.download(), d_err.download(dtype=np.uint32)))
        d_out.free()
        d_err.free()
    d_comp.free()
    d_tab.free()
    return outs


def test_noalloc_decode_equals_decode(gpu):
    bl = 18
    limit = pyzero.NOALLOC_SCRATCH_BYTES
    assert gpu.lib.zgpu_decode_scratch_bytes(128, bl) <= limit < gpu.lib.zgpu_decode_scratch_bytes(140, bl)
    for seed, nb, damage in ((3, 128, 3), (4, 9, 1)):  # the second call smaller than the first
        raw, blob, tab, hit = case(seed, nb, bl, damage)
        (a, ea), (b, eb) = decode_both(gpu, raw, blob, tab, bl)
        assert np.array_equal(ea, eb), (seed, ea, eb)  # the same error bits, damaged blocks included
        assert all(ea[k] == 0 for k in range(nb) if k not in hit)
        ok = np.repeat(ea == 0, 1 << bl)[: raw.size]  # every block that decoded: the same bytes
        assert np.array_equal(a[ok], b[ok]), seed
        clean = np.repeat([k not in hit for k in range(nb)], 1 << bl)[: raw.size]
        assert np.array_equal(b[clean], raw[clean]), seed


def test_noalloc_refuses_more_than_its_scratch(gpu):
    bl = 18
    nb = 140
    tab = pyzero.block_table([1 << bl] * nb, comp_lens=[1 << bl] * nb, methods=[0] * nb)
    d_tab = gpu.alloc(tab.nbytes)
    d_tab.upload(tab)
    d_comp = gpu.alloc(nb << bl)
    d_out = gpu.alloc(nb << bl)
    d_err = gpu.alloc(nb * 4)
    with pytest.raises(pyzero.ZeroError) as e:
        gpu.decode_blocks_noalloc(d_comp.ptr, d_tab.ptr, nb, bl, d_out.ptr, d_err.ptr)
    assert e.value.code == pyzero.ZGPU_ERR_ARG and "scratch" in str(e.value)
    for buf in (d_tab, d_comp, d_out, d_err):
        buf.free()
