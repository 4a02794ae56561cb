"""Benchmark of the MI355X zero block codec on its headline workload.

Workload (one GPU): 4 GiB of LZ-friendly synthetic bytes (seed 1) and 1 GiB of uniform random bytes (seed 2), level 5,
blocks of 2^18 bytes, every buffer resident in HBM when a timed region starts. A step is one pass of the codec over the
whole 4 GiB input: ``--steps`` encode passes and ``--steps`` decode passes are timed separately (HIP events on the
codec's stream, bracketed by device synchronisation). ``value`` is the decode throughput in raw GB/s (10^9 B/s,
the reference's convention: raw bytes / time); the encode throughput, the ratio and the random-data numbers sit beside
it. ``roofline`` is the dominant decode kernel's algorithmic bytes (compressed + raw) per launch over its average
launch duration against the 8 TB/s HBM peak. ``cpu_baseline`` is the reference's own CPU codec (oracle/_ref, compiled
from the reference's headers) on a bounded sample of the same input, on this machine's cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--gib G]

``--workload uneven`` measures the batched form instead: 200,000 files with lognormal sizes (median 8 KiB, sigma 2,
clipped to 64 B .. 64 MiB) cut from the seed-1 LZ stream, level 5, 2^18 blocks, through one size-class plan
(autotuned, replayed as a hipGraph), against the plain path called once per file on a sample of the same files.

``--workload serve`` runs the serving loop under an open-loop load: Poisson arrivals at 20 GB/s of raw bytes for one
second, 70 % compress / 30 % decompress, sizes lognormal (median 256 KiB, sigma 1.5, at most 256 MiB), seed-1 LZ and
random bytes 9 : 1, levels uniform over 1 / 5 / 9, a 64 GiB HBM budget; latency from arrival to the result in host
memory. The same load at 8 GB/s is reported beside it.

``--workload sweep`` crosses block sizes 2^16 / 2^18 / 2^20 with levels 1 / 5 / 9 over 2 GiB of executable-like
pages (seed 3): ratio, encode and decode GB/s per point, every round trip checked.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "zero-compression_amd"))
HBM_PEAK = 8.0e12


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(data, bl, depth, budget_s=12.0):
    """The reference CpuCodec (oracle/_ref/libzeroref.so) on a sample of the input, all cores of this process."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Ref, ref_available
    if not ref_available():
        return None
    ref = Ref()
    threads = len(os.sched_getaffinity(0))
    sample = data[: 256 << 20]
    t0 = time.perf_counter()
    blob, clen, meth = ref.cpu_compress(sample, bl, depth, threads)
    tc = time.perf_counter() - t0
    raw_len = np.array([min(1 << bl, sample.size - o) for o in range(0, sample.size, 1 << bl)], dtype=np.uint32)
    reps, td = 0, 0.0
    while reps < 3 or (td < budget_s / 3 and reps < 20):
        t0 = time.perf_counter()
        rc, back = ref.cpu_decompress(blob, clen, meth, raw_len, threads)
        td += time.perf_counter() - t0
        reps += 1
        assert rc == 0
    assert np.array_equal(back, sample)
    return {
        "value": round(sample.size * reps / td / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "reference",
        "sample": f"first {sample.size >> 20} MiB of the seed-1 LZ input, level 5 (chain depth {depth}), 2^{bl} blocks: "
                  f"decode x{reps} (value), encode once",
        "compress_GBps": round(sample.size / tc / 1e9, 4), "ratio": round(float(clen.sum()) / sample.size, 4),
    }


def traffic_of(kernel, n, level, bl):
    """HBM bytes per launch of `kernel` on this workload, from the committed counter summary (profiles/traffic.json:
    rocprofv3 FETCH_SIZE and WRITE_SIZE passes, each scaled by what the counter reads for k_copy of known size), or
    None when the summary is for another workload or kernel."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        t = json.load(open(path))
    except (OSError, ValueError):
        return None
    w = t.get("workload", {})
    if (w.get("input_bytes"), w.get("level"), w.get("block_log")) != (n, level, bl):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k["hbm_bytes_per_launch"]


def bench_uneven(a, gpu):
    """The batched uneven-file form against per-file calls of the plain path (one JSON line)."""
    import pyzero

    bl, level = a.block_log, a.level
    B = 1 << bl
    sizes = pyzero.gen_lognormal_sizes(4, 8192.0, 2.0, 64, 1 << 26, a.files).astype(np.int64)
    total = int(sizes.sum())
    t0 = time.time()
    data = pyzero.gen_lz(1, total)
    log(f"uneven: {a.files} files, {total / 2**30:.2f} GiB in total (generated in {time.time() - t0:.1f} s)")
    src, rl, fi = pyzero.file_blocks(sizes, bl)
    nb = int(rl.size)
    d_in = gpu.alloc(total)
    d_in.upload(data)
    ev0, ev1 = gpu.event(), gpu.event()

    def timed(fn, steps):
        gpu.sync()
        gpu.record(ev0)
        for _ in range(steps):
            fn()
        gpu.record(ev1)
        gpu.sync()
        return gpu.elapsed_ms(ev0, ev1) / steps

    plan = pyzero.Plan(gpu, src, rl, bl, a.streams)
    d_slabs = gpu.alloc(plan.slab_bytes)
    d_len = gpu.alloc(nb * 4)
    t0 = time.time()
    plan.tune(d_in.ptr, level, d_slabs.ptr, d_len.ptr)
    log(f"uneven: tuned in {time.time() - t0:.1f} s; classes {plan.classes()}")
    plan.use_graphs(True)
    for _ in range(max(1, a.warmup)):
        plan.encode(d_in.ptr, level, d_slabs.ptr, d_len.ptr)
    enc_ms = timed(lambda: plan.encode(d_in.ptr, level, d_slabs.ptr, d_len.ptr), a.steps)
    lens = d_len.download(dtype=np.uint32).copy()
    stored = lens == pyzero.STORED
    clen = np.where(stored, rl, lens).astype(np.uint32)
    slabs = d_slabs.download()
    off = plan.slab_offsets()
    comp = np.concatenate([data[int(s): int(s) + int(r)] if st else slabs[int(o): int(o) + int(c)]
                           for s, r, st, o, c in zip(src, rl, stored, off, clen)])
    del slabs
    coff = np.concatenate([[0], np.cumsum(clen.astype(np.uint64))[:-1]]).astype(np.uint64)
    plan.set_stored(coff, clen, np.where(stored, 0, 1))
    d_comp = gpu.alloc(comp.size)
    d_comp.upload(comp)
    d_out = gpu.alloc(total)
    d_err = gpu.alloc(nb * 4)
    for _ in range(max(1, a.warmup)):
        plan.decode(d_comp.ptr, d_out.ptr, d_err.ptr)
    dec_ms = timed(lambda: plan.decode(d_comp.ptr, d_out.ptr, d_err.ptr), a.steps)
    ok = not d_err.download(dtype=np.uint32).any() and np.array_equal(d_out.download(), data)
    log(f"uneven batched: encode {total / enc_ms / 1e6:.1f} GB/s, decode {total / dec_ms / 1e6:.1f} GB/s, ratio "
        f"{comp.size / total:.4f}, round trip {'identical' if ok else 'MISMATCH'}")

    # the plain path, one call per file, on the first files (device-resident, same buffers)
    nfs = min(a.files, a.sample_files)
    nbs = int(np.searchsorted(fi, nfs))
    sample_bytes = int(rl[:nbs].sum())
    tab = pyzero.block_table(rl[:nbs], src=src[:nbs], dst=off[:nbs])
    d_tab = gpu.alloc(max(1, tab.nbytes))
    d_tab.upload(tab)
    dtab = pyzero.block_table(rl[:nbs], comp_lens=clen[:nbs], methods=np.where(stored[:nbs], 0, 1),
                              src=coff[:nbs], dst=src[:nbs])
    d_dtab = gpu.alloc(max(1, dtab.nbytes))
    d_dtab.upload(dtab)
    firsts = np.searchsorted(fi[:nbs], np.arange(nfs + 1))
    d_pslab = gpu.alloc(max(1, nbs) * B)
    d_plen = gpu.alloc(max(1, nbs) * 4)

    def perfile_encode():
        for k in range(nfs):
            b0, b1 = int(firsts[k]), int(firsts[k + 1])
            if b1 > b0:
                gpu.encode_blocks(d_in.ptr, d_tab.addr(b0 * 32), b1 - b0, bl, level, d_pslab.addr(b0 * B),
                                  d_plen.addr(b0 * 4))

    def perfile_decode():
        for k in range(nfs):
            b0, b1 = int(firsts[k]), int(firsts[k + 1])
            if b1 > b0:
                gpu.decode_blocks(d_comp.ptr, d_dtab.addr(b0 * 32), b1 - b0, bl, d_out.ptr, d_err.addr(b0 * 4))

    perfile_encode()
    pf_enc_ms = timed(perfile_encode, 1)
    plen = d_plen.download(nbs * 4, dtype=np.uint32)
    same = bool(np.array_equal(plen, lens[:nbs]))
    perfile_decode()
    pf_dec_ms = timed(perfile_decode, 1)
    # the batched form on the same sample
    with pyzero.Plan(gpu, src[:nbs], rl[:nbs], bl, a.streams) as sp:
        d_s2 = gpu.alloc(max(1, sp.slab_bytes))
        sp.tune(d_in.ptr, level, d_s2.ptr, d_len.ptr)
        sp.use_graphs(True)
        sp.encode(d_in.ptr, level, d_s2.ptr, d_len.ptr)
        sb_enc_ms = timed(lambda: sp.encode(d_in.ptr, level, d_s2.ptr, d_len.ptr), a.steps)
        sp.set_stored(coff[:nbs], clen[:nbs], np.where(stored[:nbs], 0, 1))
        sp.decode(d_comp.ptr, d_out.ptr, d_err.ptr)
        sb_dec_ms = timed(lambda: sp.decode(d_comp.ptr, d_out.ptr, d_err.ptr), a.steps)
    log(f"sample of {nfs} files ({sample_bytes / 2**20:.1f} MiB): per-file encode {sample_bytes / pf_enc_ms / 1e6:.2f} "
        f"GB/s, decode {sample_bytes / pf_dec_ms / 1e6:.2f} GB/s; batched encode {sample_bytes / sb_enc_ms / 1e6:.2f} "
        f"GB/s, decode {sample_bytes / sb_dec_ms / 1e6:.2f} GB/s; per-file payloads identical: {same}")
    line = {
        "metric": "batched uneven-file compress/decompress GB/s vs the plain per-file form",
        "value": round(total / (dec_ms * 1e-3) / 1e9, 3), "unit": "GB/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(dec_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": f"{a.files} files, lognormal sizes (median 8 KiB, sigma 2, 64 B..64 MiB), seed-1 LZ "
                               f"bytes, level {level}, 2^{bl} blocks, one plan; step = decode of every file",
                   "files": a.files, "input_bytes": total, "blocks": nb, "streams": a.streams},
        "batched": {"compress_GBps": round(total / (enc_ms * 1e-3) / 1e9, 3),
                    "decompress_GBps": round(total / (dec_ms * 1e-3) / 1e9, 3), "ratio": round(comp.size / total, 5),
                    "roundtrip_identical": ok, "classes": plan.classes()},
        "per_file_sample": {"files": nfs, "bytes": sample_bytes,
                            "per_file_compress_GBps": round(sample_bytes / (pf_enc_ms * 1e-3) / 1e9, 3),
                            "per_file_decompress_GBps": round(sample_bytes / (pf_dec_ms * 1e-3) / 1e9, 3),
                            "batched_compress_GBps": round(sample_bytes / (sb_enc_ms * 1e-3) / 1e9, 3),
                            "batched_decompress_GBps": round(sample_bytes / (sb_dec_ms * 1e-3) / 1e9, 3),
                            "payloads_identical": same},
    }
    plan.close()
    print(json.dumps(line), flush=True)


def bench_serve(a, gpu):
    """The serving loop at the configured offered load and at a lighter one (one JSON line)."""
    import pyzero

    t0 = time.time()
    lz = pyzero.gen_lz(1, 2 << 30)
    rnd = pyzero.gen_random(2, 256 << 20)
    log(f"serve: pools generated in {time.time() - t0:.1f} s")
    runs = {}
    for offered in (a.offered, 8.0):
        t0 = time.time()
        r = gpu.serve_bench(lz, rnd, offered_GBps=offered, seconds=a.seconds, block_log=a.block_log,
                            hbm_budget=64 << 30, slots=a.slots, seed=5, alone_checks=64)
        log(f"serve at {offered} GB/s offered: {r['requests']} requests, sustained {r['sustained_GBps']:.2f} GB/s, "
            f"p50 {r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms, max {r['max_ms']:.1f} ms, {r['batches']} batches "
            f"(mean {r['mean_batch_requests']:.1f} requests), mismatches {r['mismatches']}, alone vs mix "
            f"{r['alone_mismatches']}/{r['alone_checked']} ({time.time() - t0:.1f} s)")
        runs[offered] = r
    main_run = runs[a.offered]
    line = {
        "metric": "serving p50/p99 latency (arrival to result in host memory) and sustained GB/s",
        "value": round(main_run["sustained_GBps"], 3), "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(main_run["wall_s"] * 1e3, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": f"open-loop Poisson arrivals at {a.offered} GB/s for {a.seconds} s, 70 % compress, "
                               "lognormal sizes (median 256 KiB, sigma 1.5, <= 256 MiB), LZ : random 9 : 1, levels "
                               f"1/5/9, 2^{a.block_log} blocks, 64 GiB HBM budget, {a.slots} batch slots"},
        "runs": {str(k): {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                 for k, v in runs.items()},
    }
    print(json.dumps(line), flush=True)


def bench_sweep(a, gpu):
    """Block size x level over the executable-like workload (one JSON line with every point)."""
    import pyzero

    n = int(a.gib * (1 << 30)) if a.gib != 4.0 else 2 << 30
    t0 = time.time()
    data = pyzero.gen_exec(3, n)
    log(f"sweep: {n >> 20} MiB of executable-like pages generated in {time.time() - t0:.1f} s")
    d_in = gpu.alloc(n)
    d_in.upload(data)
    ev0, ev1 = gpu.event(), gpu.event()
    points = []
    for bl in (16, 18, 20):
        B = 1 << bl
        nb = n // B
        tab = pyzero.block_table([B] * nb, src=np.arange(nb, dtype=np.uint64) * B,
                                 dst=np.arange(nb, dtype=np.uint64) * B)
        d_tab = gpu.alloc(tab.nbytes)
        d_tab.upload(tab)
        batch = max(1, (1 << 30) >> bl)
        slabs = gpu.alloc(n)
        lens = gpu.alloc(nb * 4)
        for level in (1, 5, 9):
            def enc():
                for b0 in range(0, nb, batch):
                    k = min(batch, nb - b0)
                    gpu.encode_blocks(d_in.ptr, d_tab.addr(b0 * 32), k, bl, level, slabs.addr(b0 << bl),
                                      lens.addr(b0 * 4))
            enc()
            gpu.sync()
            gpu.record(ev0)
            for _ in range(a.steps):
                enc()
            gpu.record(ev1)
            gpu.sync()
            enc_ms = gpu.elapsed_ms(ev0, ev1) / a.steps
            ln = lens.download(dtype=np.uint32).copy()
            st = ln == pyzero.STORED
            clen = np.where(st, B, ln).astype(np.uint64)
            sl = slabs.download()
            blob = np.concatenate([data[b * B:(b + 1) * B] if st[b] else sl[b * B: b * B + int(ln[b])]
                                   for b in range(nb)])
            del sl
            dt = pyzero.block_table([B] * nb, comp_lens=clen, methods=np.where(st, 0, 1))
            d_dt = gpu.alloc(dt.nbytes)
            d_dt.upload(dt)
            d_comp = gpu.alloc(blob.size)
            d_comp.upload(blob)
            d_out = gpu.alloc(n)
            d_err = gpu.alloc(nb * 4)
            gpu.decode_blocks(d_comp.ptr, d_dt.ptr, nb, bl, d_out.ptr, d_err.ptr)
            gpu.sync()
            gpu.record(ev0)
            for _ in range(a.steps):
                gpu.decode_blocks(d_comp.ptr, d_dt.ptr, nb, bl, d_out.ptr, d_err.ptr)
            gpu.record(ev1)
            gpu.sync()
            dec_ms = gpu.elapsed_ms(ev0, ev1) / a.steps
            ok = not d_err.download(dtype=np.uint32).any() and np.array_equal(d_out.download(), data)
            p = {"block_log": bl, "level": level, "ratio": round(blob.size / n, 5),
                 "compress_GBps": round(n / (enc_ms * 1e-3) / 1e9, 2), "decompress_GBps": round(n / (dec_ms * 1e-3) / 1e9, 2),
                 "stored_blocks": int(st.sum()), "roundtrip_identical": bool(ok)}
            log(f"sweep {p}")
            points.append(p)
            for buf in (d_dt, d_comp, d_out, d_err):
                buf.free()
        for buf in (d_tab, slabs, lens):
            buf.free()
    l5 = [p for p in points if p["block_log"] == 18 and p["level"] == 5][0]
    line = {
        "metric": "block/level sweep: ratio and compress/decompress GB/s per point", "value": l5["decompress_GBps"],
        "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": 1, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": f"{n >> 20} MiB executable-like pages (seed 3), blocks 2^16/2^18/2^20 x levels 1/5/9; "
                               "value = decode GB/s at 2^18, level 5"},
        "points": points,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--gib", type=float, default=4.0, help="size of the LZ input in GiB")
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--block-log", type=int, default=18)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["plain", "uneven", "serve", "sweep"], default="plain")
    ap.add_argument("--offered", type=float, default=20.0, help="serve workload: offered load in GB/s")
    ap.add_argument("--seconds", type=float, default=1.0, help="serve workload: arrival window")
    ap.add_argument("--slots", type=int, default=6, help="serve workload: batches in flight")
    ap.add_argument("--files", type=int, default=200000, help="uneven workload: number of files")
    ap.add_argument("--sample-files", type=int, default=2000, help="uneven workload: files timed per-file")
    ap.add_argument("--streams", type=int, default=4, help="uneven workload: streams of the plan")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (blocks would shard across GPUs with one "
              "all-gather of block lengths). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import pyzero

    bl, level = a.block_log, a.level
    B = 1 << bl
    n = int(a.gib * (1 << 30)) // B * B
    nb = n // B
    gpu = pyzero.Gpu(0)
    props = gpu.props()
    log(f"device {props['name']} {props['arch']} {props['cu_count']} CUs")
    if a.workload != "plain":
        {"uneven": bench_uneven, "serve": bench_serve, "sweep": bench_sweep}[a.workload](a, gpu)
        gpu.close()
        return
    t0 = time.time()
    data = pyzero.gen_lz(1, n)
    rnd = pyzero.gen_random(2, 1 << 30)
    log(f"generated {n >> 20} MiB LZ + 1024 MiB random in {time.time() - t0:.1f} s")

    # ---- device buffers
    d_in = gpu.alloc(n)
    d_in.upload(data)
    tab = pyzero.block_table([B] * nb, src=np.arange(nb, dtype=np.uint64) * B, dst=np.arange(nb, dtype=np.uint64) * B)
    d_tab = gpu.alloc(tab.nbytes)
    d_tab.upload(tab)
    slabs = gpu.alloc(n)
    lens = gpu.alloc(nb * 4)
    batch = 4096  # blocks per encoder launch (1 GiB at 2^18)

    def encode_pass(din, dtab, k_blocks, out_slabs, out_lens):
        for b0 in range(0, k_blocks, batch):
            k = min(batch, k_blocks - b0)
            gpu.encode_blocks(din.ptr, dtab.addr(b0 * 32), k, bl, level, out_slabs.addr(b0 << bl),
                              out_lens.addr(b0 * 4))

    ev0, ev1 = gpu.event(), gpu.event()
    # ---- encode
    for _ in range(a.warmup):
        encode_pass(d_in, d_tab, nb, slabs, lens)
    gpu.sync()
    gpu.record(ev0)
    for _ in range(a.steps):
        encode_pass(d_in, d_tab, nb, slabs, lens)
    gpu.record(ev1)
    gpu.sync()
    enc_ms = gpu.elapsed_ms(ev0, ev1) / a.steps
    log(f"encode {enc_ms:.2f} ms/step = {n / enc_ms / 1e6:.1f} GB/s")
    gpu.set_profile(True)
    encode_pass(d_in, d_tab, min(nb, batch), slabs, lens)
    enc_stages = gpu.stage_times()
    gpu.set_profile(False)
    log("encode stages (ms, one 1 GiB launch): " + ", ".join(f"{k} {v:.2f}" for k, v in enc_stages))

    # ---- the archive: payloads packed back to back (host side, untimed)
    ln = lens.download(dtype=np.uint32).copy()
    sl = slabs.download()
    stored = ln == pyzero.STORED
    clen = np.where(stored, B, ln).astype(np.uint64)
    meth = np.where(stored, 0, 1).astype(np.uint32)
    parts = [sl[b * B: b * B + int(ln[b])] if not stored[b] else data[b * B:(b + 1) * B] for b in range(nb)]
    blob = np.concatenate(parts)
    del sl, parts
    comp_total = int(clen.sum())
    ratio = comp_total / n
    log(f"ratio {ratio:.4f} ({comp_total} / {n}), {int(stored.sum())} stored blocks")
    slabs.free()
    d_comp = gpu.alloc(blob.size)
    d_comp.upload(blob)
    dtab = pyzero.block_table([B] * nb, comp_lens=clen, methods=meth)
    d_dtab = gpu.alloc(dtab.nbytes)
    d_dtab.upload(dtab)
    d_out = gpu.alloc(n)
    d_err = gpu.alloc(nb * 4)

    # ---- decode
    for _ in range(a.warmup):
        gpu.decode_blocks(d_comp.ptr, d_dtab.ptr, nb, bl, d_out.ptr, d_err.ptr)
    gpu.sync()
    gpu.set_profile(True)
    stage_sum = {}
    gpu.record(ev0)
    for _ in range(a.steps):
        gpu.decode_blocks(d_comp.ptr, d_dtab.ptr, nb, bl, d_out.ptr, d_err.ptr)
        for k, v in gpu.stage_times():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    gpu.record(ev1)
    gpu.sync()
    gpu.set_profile(False)
    dec_ms = gpu.elapsed_ms(ev0, ev1) / a.steps
    stage_avg = {k: v / a.steps for k, v in stage_sum.items()}
    log(f"decode {dec_ms:.3f} ms/step = {n / dec_ms / 1e6:.1f} GB/s; stages {stage_avg}")
    # timed again without the per-stage events or syncs, for the headline
    gpu.record(ev0)
    for _ in range(a.steps):
        gpu.decode_blocks(d_comp.ptr, d_dtab.ptr, nb, bl, d_out.ptr, d_err.ptr)
    gpu.record(ev1)
    gpu.sync()
    dec_ms = gpu.elapsed_ms(ev0, ev1) / a.steps
    err = d_err.download(dtype=np.uint32)
    ok = bool((err == 0).all()) and np.array_equal(d_out.download(), data)
    log(f"decode {dec_ms:.3f} ms/step = {n / dec_ms / 1e6:.1f} GB/s, round trip {'identical' if ok else 'MISMATCH'}")
    # a plain device copy of the same size (k_copy): the bandwidth this GPU gives a streaming kernel
    gpu.copy(d_out.ptr, d_in.ptr, n)
    gpu.record(ev0)
    for _ in range(a.steps):
        gpu.copy(d_out.ptr, d_in.ptr, n)
    gpu.record(ev1)
    gpu.sync()
    copy_ms = gpu.elapsed_ms(ev0, ev1) / a.steps
    log(f"copy {n >> 20} MiB: {copy_ms:.3f} ms = {2 * n / copy_ms / 1e6:.1f} GB/s (read + write)")
    dom = max(stage_avg, key=stage_avg.get)
    alg = comp_total + n
    roof = {"bound": "hbm", "kernel": f"k_dec_{dom}", "achieved": round(alg / (stage_avg[dom] * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK / 1e9, "unit": "GB/s"}
    roof["frac"] = round(roof["achieved"] * 1e9 / HBM_PEAK, 4)
    roof["traffic"] = traffic_of(f"k_dec_{dom}", n, level, bl)
    roof["whole_decode_frac"] = round(alg / (dec_ms * 1e-3) / HBM_PEAK, 4)

    # ---- random 1 GiB: every block stored
    for buf in (d_comp, d_out):
        buf.free()
    nr = rnd.size // B
    r_in = gpu.alloc(rnd.size)
    r_in.upload(rnd)
    r_slabs = gpu.alloc(rnd.size)
    r_lens = gpu.alloc(nr * 4)
    rtab = pyzero.block_table([B] * nr, src=np.arange(nr, dtype=np.uint64) * B, dst=np.arange(nr, dtype=np.uint64) * B)
    r_tab = gpu.alloc(rtab.nbytes)
    r_tab.upload(rtab)
    encode_pass(r_in, r_tab, nr, r_slabs, r_lens)
    gpu.sync()
    gpu.record(ev0)
    for _ in range(a.steps):
        encode_pass(r_in, r_tab, nr, r_slabs, r_lens)
    gpu.record(ev1)
    gpu.sync()
    r_enc_ms = gpu.elapsed_ms(ev0, ev1) / a.steps
    rl = r_lens.download(dtype=np.uint32)
    r_stored = int((rl == pyzero.STORED).sum())
    rdt = pyzero.block_table([B] * nr, comp_lens=[B] * nr, methods=[0] * nr)
    r_tab.upload(rdt)
    r_out = gpu.alloc(rnd.size)
    r_err = gpu.alloc(nr * 4)
    gpu.decode_blocks(r_in.ptr, r_tab.ptr, nr, bl, r_out.ptr, r_err.ptr)
    gpu.sync()
    gpu.record(ev0)
    for _ in range(a.steps):
        gpu.decode_blocks(r_in.ptr, r_tab.ptr, nr, bl, r_out.ptr, r_err.ptr)
    gpu.record(ev1)
    gpu.sync()
    r_dec_ms = gpu.elapsed_ms(ev0, ev1) / a.steps
    r_ok = np.array_equal(r_out.download(), rnd)
    log(f"random: encode {rnd.size / r_enc_ms / 1e6:.1f} GB/s ({r_stored}/{nr} stored), "
        f"decode {rnd.size / r_dec_ms / 1e6:.1f} GB/s, {'identical' if r_ok else 'MISMATCH'}")

    cpu = None if a.no_cpu else cpu_baseline(data, bl, {1: 8, 9: 256}.get(level, 32))
    if cpu:
        log(f"cpu baseline: {cpu}")
    line = {
        "metric": "compress/decompress GB/s + ratio at L5/256KiB; p99 latency at fixed load",
        "value": round(n / (dec_ms * 1e-3) / 1e9, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dec_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": f"plain GPU drop-in: {n >> 30} GiB LZ-friendly seed 1 + 1 GiB random seed 2, "
                               f"level {level}, 2^{bl} blocks, device-resident; step = decode of the whole LZ input",
                   "level": level, "block_log": bl, "input_bytes": n, "blocks": nb},
        "decompress_GBps": round(n / (dec_ms * 1e-3) / 1e9, 3),
        "compress_GBps": round(n / (enc_ms * 1e-3) / 1e9, 3),
        "compress_ms_per_step": round(enc_ms, 3),
        "ratio": round(ratio, 5),
        "roundtrip_identical": bool(ok and r_ok),
        "random": {"compress_GBps": round(rnd.size / (r_enc_ms * 1e-3) / 1e9, 3),
                   "decompress_GBps": round(rnd.size / (r_dec_ms * 1e-3) / 1e9, 3),
                   "stored_blocks": r_stored, "blocks": nr},
        "decode_stage_ms": {k: round(v, 4) for k, v in stage_avg.items()},
        "copy_GBps": round(2 * n / (copy_ms * 1e-3) / 1e9, 1),
        "encode_stage_ms_per_GiB": {k: round(v, 3) for k, v in enc_stages},
        "roofline": roof,
        "cpu_baseline": cpu,
        "device": props["name"] or f"{props['arch']}, {props['cu_count']} CUs",  # (the name is empty on some boxes)
    }
    print(json.dumps(line), flush=True)
    gpu.close()


if __name__ == "__main__":
    main()
