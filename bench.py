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
(persistent stage kernels, autotuned, replayed as a hipGraph), against the plain path called once per file on every
file: GB/s of both, and p50/p99 latency of a backlog of all files submitted at once.

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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def time_ref(ref, data, bl, depth, threads, budget_s):
    """the reference CpuCodec's compress and decompress of data, each repeated to about budget_s / 2 of wall time;
    the round trip is checked"""
    reps_c, tc = 0, 0.0
    while reps_c < 1 or (tc < budget_s / 2 and reps_c < 50):
        t0 = time.perf_counter()
        blob, clen, meth = ref.cpu_compress(data, bl, depth, threads)
        tc += time.perf_counter() - t0
        reps_c += 1
    raw_len = np.array([min(1 << bl, data.size - o) for o in range(0, data.size, 1 << bl)], dtype=np.uint32)
    reps_d, td = 0, 0.0
    while reps_d < 3 or (td < budget_s / 2 and reps_d < 200):
        t0 = time.perf_counter()
        rc, back = ref.cpu_decompress(blob, clen, meth, raw_len, threads)
        td += time.perf_counter() - t0
        reps_d += 1
        assert rc == 0
    assert np.array_equal(back, data)
    return (data.size * reps_c / tc / 1e9, data.size * reps_d / td / 1e9, float(clen.sum()) / data.size, reps_c,
            reps_d)


def cpu_baseline(data, bl, depth, budget_s=8.0):
    """The reference's own CpuCodec (oracle/_ref/libzeroref.so, compiled from its headers) on this machine's cores:
    the C1 configuration as stated (64 MiB of the seed-0 LZ stream, level 5 = hash-chain depth 32, 2^18 blocks,
    `zero bench --cpu`'s work), and beside it a 256 MiB sample of this run's own input."""
    import pyzero
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Ref, ref_available
    if not ref_available():
        return None
    ref = Ref()
    threads = len(os.sched_getaffinity(0))
    c1 = pyzero.gen_lz(0, 64 << 20)
    c_gbps, d_gbps, ratio, rc_, rd_ = time_ref(ref, c1, 18, 32, threads, budget_s)
    s_c, s_d, s_ratio, _, _ = time_ref(ref, data[: 256 << 20], bl, depth, threads, budget_s / 2)
    return {
        "value": round(d_gbps, 3), "unit": "GB/s", "cores": threads, "kind": "reference",
        "cpu": cpu_model(),
        "sample": f"C1: 64 MiB of the seed-0 LZ stream, level 5 (chain depth 32), 2^18 blocks, all {threads} threads: "
                  f"decode x{rd_} (value), encode x{rc_}; round trip identical",
        "compress_GBps": round(c_gbps, 4), "ratio": round(ratio, 5),
        "same_workload_sample": {"input": f"first 256 MiB of this run's seed-1 input, depth {depth}, 2^{bl} blocks",
                                 "decompress_GBps": round(s_d, 3), "compress_GBps": round(s_c, 4),
                                 "ratio": round(s_ratio, 5)},
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

    plan = pyzero.Plan(gpu, src, rl, bl)
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

    # the plain path, one call per file, over every file: the 200k files are a backlog submitted at t = 0, and a
    # file's latency is the time until its results are on the device; an event after every 64th file's call times
    # the files of each group of 64 (the last file of a group exactly, the others at most 63 files late)
    nfs = a.files
    firsts = np.searchsorted(fi, np.arange(nfs + 1))
    maxb = int(np.max(np.diff(firsts))) if nfs else 1
    tab = pyzero.block_table(rl, src=src, dst=np.zeros(nb, dtype=np.uint64))
    d_tab = gpu.alloc(max(1, tab.nbytes))
    d_tab.upload(tab)
    dtab = pyzero.block_table(rl, comp_lens=clen, methods=np.where(stored, 0, 1), src=coff, dst=src)
    d_dtab = gpu.alloc(max(1, dtab.nbytes))
    d_dtab.upload(dtab)
    d_pslab = gpu.alloc(max(1, maxb) * B)  # (one file's slabs at a time)
    d_plen = gpu.alloc(max(1, nb) * 4)
    G = 64
    evs = [gpu.event() for _ in range((nfs + G - 1) // G + 1)]

    def perfile(kind):
        gpu.sync()
        gpu.record(evs[0])
        t0 = time.time()
        for k in range(nfs):
            b0, b1 = int(firsts[k]), int(firsts[k + 1])
            if b1 > b0:
                if kind == "encode":
                    # (the plain path writes block b's payload at out + (b << block_log): a file's blocks from its slab 0)
                    gpu.encode_blocks(d_in.ptr, d_tab.addr(b0 * 32), b1 - b0, bl, level, d_pslab.ptr, d_plen.addr(b0 * 4))
                else:
                    gpu.decode_blocks(d_comp.ptr, d_dtab.addr(b0 * 32), b1 - b0, bl, d_out.ptr, d_err.addr(b0 * 4))
            if k % G == G - 1 or k == nfs - 1:
                gpu.record(evs[k // G + 1])
        gpu.sync()
        host_s = time.time() - t0
        groups = np.array([gpu.elapsed_ms(evs[0], evs[i + 1]) for i in range((nfs + G - 1) // G)], dtype=np.float64)
        done = groups[np.arange(nfs) // G]
        return float(done.max()), done, host_s

    pf_enc_ms, pf_enc_done, pf_enc_host = perfile("encode")
    plen = d_plen.download(nb * 4, dtype=np.uint32)
    same = bool(np.array_equal(plen, lens))
    gpu._check(gpu.lib.zgpu_fill(gpu.h, d_out.ptr, 0, total, None))  # (the per-file decode must write it all again)
    pf_dec_ms, pf_dec_done, pf_dec_host = perfile("decode")
    pf_ok = not d_err.download(dtype=np.uint32).any() and np.array_equal(d_out.download(), data)
    # the batched form: one plan call over the backlog, every file done when it ends
    def pct(v, q):
        return round(float(np.percentile(v, q)), 3)
    lat = {"definition": "a backlog of every file submitted at t = 0; latency = time until the file's results are on "
                         "the device (per-file form: event after every 64th file's call; batched form: the plan "
                         "call's end, for every file)",
           "per_file_compress_ms": {"p50": pct(pf_enc_done, 50), "p99": pct(pf_enc_done, 99)},
           "per_file_decompress_ms": {"p50": pct(pf_dec_done, 50), "p99": pct(pf_dec_done, 99)},
           "batched_compress_ms": {"p50": round(enc_ms, 3), "p99": round(enc_ms, 3)},
           "batched_decompress_ms": {"p50": round(dec_ms, 3), "p99": round(dec_ms, 3)}}
    log(f"per-file form over all {nfs} files: encode {total / pf_enc_ms / 1e6:.2f} GB/s (host {pf_enc_host:.1f} s), "
        f"decode {total / pf_dec_ms / 1e6:.2f} GB/s, round trip {'identical' if pf_ok else 'MISMATCH'}, block sizes "
        f"identical to the batched form: {same}; latency {lat}")
    line = {
        "metric": "batched uneven-file compress/decompress GB/s and p50/p99 latency vs the plain per-file form",
        "value": round(total / (dec_ms * 1e-3) / 1e9, 3), "unit": "GB/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(dec_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": f"{a.files} files, lognormal sizes (median 8 KiB, sigma 2, 64 B..64 MiB), seed-1 LZ "
                               f"bytes, level {level}, 2^{bl} blocks, one plan; step = decode of every file",
                   "files": a.files, "input_bytes": total, "blocks": nb},
        "batched": {"compress_GBps": round(total / (enc_ms * 1e-3) / 1e9, 3),
                    "decompress_GBps": round(total / (dec_ms * 1e-3) / 1e9, 3), "ratio": round(comp.size / total, 5),
                    "roundtrip_identical": ok, "classes": plan.classes(), "passes": plan.passes,
                    "stage_workgroups_per_cu": plan.stage_occupancy()},
        "per_file": {"files": nfs, "bytes": total,
                     "compress_GBps": round(total / (pf_enc_ms * 1e-3) / 1e9, 3),
                     "decompress_GBps": round(total / (pf_dec_ms * 1e-3) / 1e9, 3),
                     "roundtrip_identical": pf_ok, "block_sizes_identical_to_batched": same},
        "latency": lat,
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
    batch = (1 << 30) >> bl  # blocks of the profiled encoder launch (1 GiB): its stages and its counters

    def encode_pass(din, dtab, k_blocks, out_slabs, out_lens):
        # one call over every block (the library launches at most 2 GiB of blocks at a time)
        gpu.encode_blocks(din.ptr, dtab.ptr, k_blocks, bl, level, out_slabs.ptr, out_lens.ptr)

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
    # the encoder: the whole pass over the same algorithmic bytes (raw read, payloads written), and its dominant
    # kernel, the match finder, whose own algorithmic bytes per position are the sorted entry read (4), the byte
    # (1) and the match written (4); stage times from one 1 GiB launch
    enc = dict(enc_stages)
    gib_bytes = min(nb, batch) * B
    enc_roof = {"bound": "hbm", "achieved": round(alg / (enc_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s"}
    enc_roof["frac"] = round(enc_roof["achieved"] * 1e9 / HBM_PEAK, 5)
    if "match" in enc:
        enc_roof["k_enc_match"] = {"algorithmic_bytes": 9 * gib_bytes, "ms": round(enc["match"], 3),
                                   "achieved": round(9 * gib_bytes / (enc["match"] * 1e-3) / 1e9, 1)}
        enc_roof["k_enc_match"]["frac"] = round(enc_roof["k_enc_match"]["achieved"] * 1e9 / HBM_PEAK, 5)
        enc_roof["k_enc_match"]["traffic"] = traffic_of("k_enc_match", n, level, bl)

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
        "encode_roofline": enc_roof,
        "cpu_baseline": cpu,
        "device": props["name"] or f"{props['arch']}, {props['cu_count']} CUs",  # (the name is empty on some boxes)
    }
    print(json.dumps(line), flush=True)
    gpu.close()


if __name__ == "__main__":
    main()
