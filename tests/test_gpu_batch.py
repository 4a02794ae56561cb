"""The batched form (zgpu_plan_*): blocks of many uneven files planned by size class, through the C ABI. GPU only.

Every payload must be byte-identical to the plain path's for the same block (and so to the oracle's model of the
reference GPU encoder), whatever the class, the tuned shape or the graph replay; decode must give the files back."""
import numpy as np
import pytest

import pyzero
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
O = Oracle()


def files(seed):
    rng = np.random.default_rng(seed)
    sizes = [0, 1, 3, 4, 17, 100, 1000, 1023, 1024, 1025, 5000, 8192, 8193, 70000, 1 << 18, (1 << 18) + 5, 600000]
    sizes += [int(x) for x in np.clip(np.exp(np.log(8192) + 2.0 * rng.standard_normal(40)), 64, 1 << 20)]
    lz = pyzero.gen_lz(seed, sum(sizes) + 1)
    rnd = pyzero.gen_random(seed + 1, 1 << 16)
    out, off = [], 0
    for k, n in enumerate(sizes):
        if k % 7 == 3:  # some incompressible files, some zeros
            out.append(np.resize(rnd, n))
        elif k % 7 == 5:
            out.append(np.zeros(n, dtype=np.uint8))
        else:
            out.append(lz[off: off + n])
        off += n
    return out


def run_plan(gpu, data, sizes, bl, level, streams=4, tune=False, graphs=0, pass_bytes=None, match_wg=0):
    src, rl, fi = pyzero.file_blocks(sizes, bl)
    with pyzero.Plan(gpu, src, rl, bl, streams) as plan:
        if pass_bytes:
            plan.set_pass_bytes(pass_bytes)
        if match_wg:
            assert plan.match_workgroup(match_wg) == match_wg
        d_in = gpu.alloc(max(1, data.size))
        d_in.upload(data)
        d_slabs = gpu.alloc(max(1, plan.slab_bytes))
        d_len = gpu.alloc(max(1, plan.nb) * 4)
        if tune:
            plan.tune(d_in.ptr, level, d_slabs.ptr, d_len.ptr)
        plan.use_graphs(graphs > 0)
        results = []
        for _ in range(max(1, graphs)):
            plan.encode(d_in.ptr, level, d_slabs.ptr, d_len.ptr)
            gpu.sync()
            lens = d_len.download(dtype=np.uint32)[: plan.nb].copy()
            slabs = d_slabs.download()
            off = plan.slab_offsets()
            results.append([(0, b"") if lens[b] == pyzero.STORED else
                            (1, slabs[int(off[b]): int(off[b]) + int(lens[b])].tobytes()) for b in range(plan.nb)])
            d_slabs.upload(np.zeros(max(1, plan.slab_bytes), dtype=np.uint8))  # (a replay must write it all again)
        classes = plan.classes()
        run_plan.passes = plan.passes
        run_plan.slab_bytes = plan.slab_bytes
        run_plan.occupancy = plan.stage_occupancy()
        run_plan.match_wg = plan.match_workgroup()
        # decode through the plan: payloads back to back
        blocks = results[-1]
        comp = [p if m else data[int(s): int(s) + int(r)].tobytes() for (m, p), s, r in zip(blocks, src, rl)]
        coff = np.cumsum([0] + [len(c) for c in comp])[:-1].astype(np.uint64)
        plan.set_stored(coff, [len(c) for c in comp], [m for m, _ in blocks])
        d_comp = gpu.alloc(max(1, sum(len(c) for c in comp)))
        d_comp.upload(np.frombuffer(b"".join(comp) or b"\0", dtype=np.uint8))
        d_out = gpu.alloc(max(1, data.size))
        d_err = gpu.alloc(max(1, plan.nb) * 4)
        decoded = []
        for _ in range(max(1, graphs)):
            d_out.upload(np.zeros(max(1, data.size), dtype=np.uint8))
            plan.decode(d_comp.ptr, d_out.ptr, d_err.ptr)
            gpu.sync()
            assert not d_err.download(dtype=np.uint32)[: plan.nb].any()
            decoded.append(d_out.download()[: data.size].copy())
    return results, decoded, classes, (src, rl)


@pytest.mark.parametrize("bl,level", [(18, 5), (16, 1), (18, 9)])
def test_plan_equals_plain_path(gpu, bl, level):
    fs = files(bl * 10 + level)
    sizes = [f.size for f in fs]
    data = np.concatenate(fs)
    results, decoded, classes, (src, rl) = run_plan(gpu, data, sizes, bl, level)
    assert len(classes) > 3
    assert np.array_equal(decoded[0], data)
    B = 1 << bl
    k = 0
    for f in fs:
        plain = gpu.compress(f, bl, level) if f.size else []
        for j, want in enumerate(plain):
            assert results[0][k] == want, (k, f.size, j)
            if want[0] == 1 and k % 5 == 0:  # and the oracle's model of the reference GPU encoder, on a sample
                assert want[1] == O.gpu_compress_block(f[j * B:(j + 1) * B].tobytes(), bl, level)[0]
            k += 1
    assert k == len(results[0])


def test_plan_tuned_and_graph_replays(gpu):
    """the persistent stages' output does not depend on the tuned parameters (bucket bits, parse walkers, workgroups
    per compute unit, the match stage's workgroup size), on graph replays, or on how the blocks are cut into
    passes"""
    fs = files(7)
    sizes = [f.size for f in fs]
    data = np.concatenate(fs)
    base, _, _, _ = run_plan(gpu, data, sizes, 18, 5, streams=1)
    assert run_plan.passes == 1
    tuned, dec, classes, _ = run_plan(gpu, data, sizes, 18, 5, streams=3, tune=True, graphs=3)
    assert len(tuned) == 3 and len(dec) == 3
    assert set(run_plan.occupancy) >= {"hash", "match", "parse", "dlz"}
    assert run_plan.match_wg in (64, 128, 256, 512, 1024)
    for r in tuned:
        assert r == base[0]
    for d in dec:
        assert np.array_equal(d, data)
    for w in (64, 1024):
        other, dec, _, _ = run_plan(gpu, data, sizes, 18, 5, match_wg=w)
        assert run_plan.match_wg == w and other[0] == base[0], w
        assert np.array_equal(dec[0], data)
    # many passes: a class split over passes, passes of several classes, graph replays of them
    many, dec, _, _ = run_plan(gpu, data, sizes, 18, 5, graphs=2, pass_bytes=1 << 20)
    assert run_plan.passes >= max(2, run_plan.slab_bytes >> 20), (run_plan.passes, run_plan.slab_bytes)
    assert many[0] == base[0] and many[1] == base[0]
    for d in dec:
        assert np.array_equal(d, data)


def test_serving_loop_round_trips_and_mix_independence(gpu):
    """a short open-loop run: every request completes, every response round-trips, and requests served again one per
    batch give the same bytes as in the mix"""
    lz = pyzero.gen_lz(1, 64 << 20)
    rnd = pyzero.gen_random(2, 8 << 20)
    rep = gpu.serve_bench(lz, rnd, offered_GBps=4.0, seconds=0.25, size_median=64 * 1024, size_sigma=1.5,
                          size_max=8 << 20, hbm_budget=8 << 30, slots=2, seed=11, alone_checks=24)
    assert rep["requests"] > 100
    assert rep["compress_requests"] > 0 and rep["decompress_requests"] > 0
    assert rep["mismatches"] == 0
    assert rep["alone_checked"] == 24 and rep["alone_mismatches"] == 0
    assert rep["batches"] >= 2 and rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]


def reference_pool(data, bl):
    """every 2^bl block of data as the reference's CpuCodec writes it at level 5 (hash-chain depth 32,
    src/main.cpp:54): the compiled reference when it is built, else the oracle (byte-identical to it)"""
    from oracle_lib import Ref, ref_available
    if ref_available():
        return Ref().cpu_compress(data, bl, 32, 16)
    B = 1 << bl
    blob, clen, meth = [], [], []
    for o in range(0, data.size, B):
        blk = data[o: o + B].tobytes()
        payload, st = O.cpu_compress_block(blk, 32)
        blob.append(blk if st else payload)
        clen.append(len(blob[-1]))
        meth.append(0 if st else 1)
    return np.frombuffer(b"".join(blob), dtype=np.uint8), np.array(clen, dtype=np.uint32), np.array(meth, dtype=np.uint8)


def test_serving_at_c5_shape_against_the_oracle(gpu):
    """The serving loop at the C5 shape: Poisson arrivals at 20 GB/s offered, 70 % compress, lognormal sizes (median
    256 KiB, sigma 1.5, at most 256 MiB), LZ : random 9 : 1, levels 1 / 5 / 9, 2^18 blocks, a 64 GiB HBM budget, six
    batch slots; a 0.5 s arrival window. The decompress requests decode blocks the reference's CPU codec wrote. Every
    response is round-tripped by the server itself; a sample of them is checked here block by block: compress
    responses equal the oracle's model of the GPU encoder at the request's level and decode under the reference's
    decoder, decompress responses equal the pool's bytes."""
    from oracle_lib import Ref, ref_available
    bl, B = 18, 1 << 18
    lz = pyzero.gen_lz(1, 256 << 20)
    rnd = pyzero.gen_random(2, 64 << 20)
    pools = [reference_pool(lz, bl), reference_pool(rnd, bl)]
    assert (pools[1][2] == 0).all()  # (random bytes: every block stored)
    rep = gpu.serve_bench(lz, rnd, offered_GBps=20.0, seconds=0.5, size_median=256 * 1024, size_sigma=1.5,
                          size_max=256 << 20, hbm_budget=64 << 30, block_log=bl, slots=6, seed=23, alone_checks=32,
                          dec_pools=pools, sample_every=53, sample_cap=256, sample_bytes=192 << 20)
    print(f"C5 test: {rep['requests']} requests, sustained {rep['sustained_GBps']:.2f} GB/s, p50 {rep['p50_ms']:.1f} ms,"
          f" p99 {rep['p99_ms']:.1f} ms, {rep['batches']} batches, {rep['samples']} samples", flush=True)
    assert rep["requests"] > 5000
    assert rep["mismatches"] == 0
    assert rep["alone_checked"] == 32 and rep["alone_mismatches"] == 0
    assert rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]
    assert rep["sustained_GBps"] > 10.0  # (the server keeps up with most of the offered load even in a short window)
    ref = Ref() if ref_available() else None
    src = [lz, rnd]
    seen = {0: 0, 1: 0}
    levels = set()
    for f, body, lens in rep["sample"]:
        raw = src[f["pool"]][f["src_off"]: f["src_off"] + f["raw_len"]]
        assert raw.size == f["raw_len"]
        seen[f["op"]] += 1
        if f["op"] == 1:
            assert f["err"] == 0 and body == raw.tobytes()
            continue
        levels.add(f["level"])
        offs = np.concatenate([[0], np.cumsum([0 if x == pyzero.STORED else int(x) for x in lens])])
        assert offs[-1] == len(body)
        nb = f["nblk"]
        for k in sorted({0, nb - 1, nb // 3, (2 * nb) // 3}):  # (at most four blocks of a request: the model is slow)
            blk = raw[k * B:(k + 1) * B].tobytes()
            want, st = O.gpu_compress_block(blk, bl, f["level"])
            if st:
                assert lens[k] == pyzero.STORED
                continue
            got = body[offs[k]: offs[k + 1]]
            assert got == want, (f, k)
            if ref is not None:
                rc, back = ref.decompress_block(got, len(blk))
                assert rc == 0 and back == blk
    assert seen[0] > 20 and seen[1] > 5 and levels == {1, 5, 9}, (seen, levels)
