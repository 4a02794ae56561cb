"""Python binding of the MI355X zero block codec (libzero_hip.so, C ABI in include/zero_hip.h) and of the
synthetic workload generators (libzerogen.so).

The codec object mirrors zero::Codec (reference src/archive.h:56-69): ``compress`` turns whole blocks of a buffer
into (method, payload) pairs, ``decompress`` turns a list of stored blocks back into their raw bytes. Device-resident
entry points (``encode_blocks`` / ``decode_blocks`` / ``crc32_blocks``) take device pointers, as the benchmark does.

There is no CPU fallback: without the built library on a GPU machine every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzero_hip.so")
GEN_PATH = os.path.join(HERE, "libzerogen.so")

ZGPU_OK = 0
ZGPU_ERR_DEVICE = -1
ZGPU_ERR_ARG = -2
ZGPU_ERR_DAMAGED = -3
ZGPU_ERR_NOMEM = -4
STORED = 0xFFFFFFFF
NOALLOC_SCRATCH_BYTES = 64 << 20
BAD_BLOCK = 0xFFFFFFFE
# per-block decoder error bits (err[b])
E_HDR, E_COUNTS, E_STREAM, E_TABLE, E_LANE, E_TOKEN, E_GROUP, E_METHOD = (1 << k for k in range(8))

# every symbol include/zero_hip.h declares (tests check the library exports them all)
EXPORTS = (
    "zgpu_open", "zgpu_close", "zgpu_props", "zgpu_last_error", "zgpu_sync",
    "zgpu_alloc", "zgpu_free", "zgpu_alloc_host", "zgpu_free_host", "zgpu_upload", "zgpu_download", "zgpu_fill", "zgpu_copy",
    "zgpu_stream_create", "zgpu_stream_destroy", "zgpu_stream_sync",
    "zgpu_event_create", "zgpu_event_destroy", "zgpu_event_record", "zgpu_event_elapsed_ms",
    "zgpu_encode_blocks", "zgpu_decode_blocks", "zgpu_decode_blocks_noalloc", "zgpu_crc32_blocks",
    "zgpu_encode_scratch_bytes", "zgpu_decode_scratch_bytes",
    "zgpu_compress", "zgpu_decompress", "zgpu_encode_stream", "zgpu_crc32", "zgpu_crc32_ranges",
    "zgpu_set_profile", "zgpu_stage_times",
    "zgpu_plan_create", "zgpu_plan_destroy", "zgpu_plan_slab_bytes", "zgpu_plan_slab_offsets", "zgpu_plan_classes",
    "zgpu_plan_encode", "zgpu_plan_set_stored", "zgpu_plan_decode", "zgpu_plan_tune", "zgpu_plan_use_graphs",
    "zgpu_plan_stage_occupancy", "zgpu_plan_set_stage_occupancy", "zgpu_plan_passes", "zgpu_plan_pass_bytes",
    "zgpu_plan_match_workgroup",
    "zgpu_plan_profile", "zgpu_plan_stage_ms",
    "zgpu_serve_bench",
)

BLOCK_DTYPE = np.dtype([("src", "<u8"), ("dst", "<u8"), ("raw_len", "<u4"), ("comp_len", "<u4"),
                        ("method", "<u4"), ("level", "<u4")])


class ZeroError(RuntimeError):
    """A codec failure; ``code`` is the ZGPU_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class DamagedError(ZeroError):
    """At least one block is damaged (the reference's zb::Error, src/codec.h:29-31)."""


class DeviceError(ZeroError):
    """The GPU failed (the reference's gpu::Error, src/gpu.h:17-27)."""


class Props(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128), ("arch", ctypes.c_char * 64), ("cu_count", ctypes.c_int),
                ("clock_khz", ctypes.c_int), ("hbm_bytes", ctypes.c_uint64), ("lds_per_block", ctypes.c_uint32),
                ("max_group", ctypes.c_uint32), ("integrated", ctypes.c_int)]


_lib = None
_gen = None
vp = ctypes.c_void_p
u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
i32 = ctypes.c_int


def load_library(path=LIB_PATH):
    """Loads libzero_hip.so; raises when it is not built (there is no other code path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ZeroError(ZGPU_ERR_DEVICE, f"{path} is not built: run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    sig = {
        "zgpu_open": (i32, [i32, ctypes.POINTER(vp)]), "zgpu_close": (None, [vp]),
        "zgpu_props": (i32, [vp, ctypes.POINTER(Props)]), "zgpu_last_error": (ctypes.c_char_p, [vp]),
        "zgpu_sync": (i32, [vp]),
        "zgpu_alloc": (i32, [vp, ctypes.c_size_t, ctypes.POINTER(vp)]), "zgpu_free": (i32, [vp, vp]),
        "zgpu_alloc_host": (i32, [vp, ctypes.c_size_t, ctypes.POINTER(vp)]), "zgpu_free_host": (i32, [vp, vp]),
        "zgpu_upload": (i32, [vp, vp, vp, ctypes.c_size_t]), "zgpu_download": (i32, [vp, vp, vp, ctypes.c_size_t]),
        "zgpu_fill": (i32, [vp, vp, ctypes.c_uint8, ctypes.c_size_t, vp]),
        "zgpu_copy": (i32, [vp, vp, vp, ctypes.c_size_t, vp]),
        "zgpu_plan_create": (i32, [vp, u32, vp, vp, u32, i32, ctypes.POINTER(vp)]),
        "zgpu_plan_destroy": (None, [vp]), "zgpu_plan_slab_bytes": (ctypes.c_uint64, [vp]),
        "zgpu_plan_slab_offsets": (i32, [vp, vp]), "zgpu_plan_classes": (i32, [vp, vp, vp, vp, vp, i32]),
        "zgpu_plan_encode": (i32, [vp, vp, i32, vp, vp]), "zgpu_plan_set_stored": (i32, [vp, vp, vp, vp]),
        "zgpu_plan_decode": (i32, [vp, vp, vp, vp]), "zgpu_plan_tune": (i32, [vp, vp, i32, vp, vp]),
        "zgpu_plan_use_graphs": (i32, [vp, i32]),
        "zgpu_plan_stage_occupancy": (i32, [vp, vp, i32]), "zgpu_plan_passes": (i32, [vp]),
        "zgpu_plan_pass_bytes": (i32, [vp, u64]), "zgpu_plan_set_stage_occupancy": (i32, [vp, vp, i32]),
        "zgpu_plan_match_workgroup": (i32, [vp, u32]),
        "zgpu_plan_profile": (i32, [vp, i32]), "zgpu_plan_stage_ms": (i32, [vp, vp, i32]),
        "zgpu_serve_bench": (i32, [vp, ctypes.POINTER(ServeCfg), vp, ctypes.c_uint64, vp, ctypes.c_uint64,
                                   ctypes.POINTER(ServeReport)]),
        "zgpu_stream_create": (i32, [vp, ctypes.POINTER(vp)]), "zgpu_stream_destroy": (i32, [vp, vp]),
        "zgpu_stream_sync": (i32, [vp, vp]),
        "zgpu_event_create": (i32, [vp, ctypes.POINTER(vp)]), "zgpu_event_destroy": (i32, [vp, vp]),
        "zgpu_event_record": (i32, [vp, vp, vp]),
        "zgpu_event_elapsed_ms": (i32, [vp, vp, vp, ctypes.POINTER(ctypes.c_float)]),
        "zgpu_encode_blocks": (i32, [vp, vp, vp, u32, u32, i32, vp, vp, vp]),
        "zgpu_decode_blocks": (i32, [vp, vp, vp, u32, u32, vp, vp, vp]),
        "zgpu_decode_blocks_noalloc": (i32, [vp, vp, vp, u32, u32, vp, vp, vp]),
        "zgpu_crc32_blocks": (i32, [vp, vp, vp, u32, i32, vp, vp]),
        "zgpu_encode_scratch_bytes": (ctypes.c_size_t, [u32, u32]),
        "zgpu_decode_scratch_bytes": (ctypes.c_size_t, [u32, u32]),
        "zgpu_compress": (i32, [vp, vp, u64, u32, i32, vp, u64, vp, vp]),
        "zgpu_decompress": (i32, [vp, vp, vp, vp, vp, u32, u32, vp, vp, vp]),
        "zgpu_encode_stream": (i32, [vp, vp, u32, vp, ctypes.POINTER(u32)]),
        "zgpu_crc32": (i32, [vp, vp, u64, u32, vp]),
        "zgpu_crc32_ranges": (i32, [vp, i32, u32, vp, vp, vp]),
        "zgpu_set_profile": (i32, [vp, i32]),
        "zgpu_stage_times": (i32, [vp, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_float), i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    """Address of a numpy array's data (contiguous)."""
    return ctypes.c_void_p(a.ctypes.data)


class ServeSample(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint32), ("level", ctypes.c_uint32), ("pool", ctypes.c_uint32),
                ("nblk", ctypes.c_uint32), ("src_off", ctypes.c_uint64), ("raw_len", ctypes.c_uint64),
                ("bytes_off", ctypes.c_uint64), ("bytes_len", ctypes.c_uint64), ("blk_len_off", ctypes.c_uint32),
                ("err", ctypes.c_uint32), ("latency_ms", ctypes.c_double)]


class ServeCfg(ctypes.Structure):
    _fields_ = [("offered_GBps", ctypes.c_double), ("seconds", ctypes.c_double), ("compress_share", ctypes.c_double),
                ("random_share", ctypes.c_double), ("size_median", ctypes.c_double), ("size_sigma", ctypes.c_double),
                ("size_max", ctypes.c_uint64), ("hbm_budget", ctypes.c_uint64), ("block_log", ctypes.c_uint32),
                ("slots", ctypes.c_int), ("seed", ctypes.c_uint64), ("alone_checks", ctypes.c_uint32),
                ("dec_comp", vp * 2), ("dec_comp_len", vp * 2), ("dec_method", vp * 2),
                ("sample_every", ctypes.c_uint32), ("sample_cap", ctypes.c_uint32),
                ("samples", ctypes.POINTER(ServeSample)), ("sample_bytes", vp), ("sample_bytes_cap", ctypes.c_uint64),
                ("sample_blk_len", vp), ("sample_blk_cap", ctypes.c_uint32)]


class ServeReport(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("requests", "compress_requests", "decompress_requests", "bytes")] + \
               [(n, ctypes.c_double) for n in ("p50_ms", "p99_ms", "max_ms", "mean_ms", "wall_s", "sustained_GBps")] + \
               [("batches", ctypes.c_uint64), ("mean_batch_requests", ctypes.c_double)] + \
               [(n, ctypes.c_uint64) for n in ("hbm_bytes", "slot_input_bytes", "mismatches", "alone_checked",
                                               "alone_mismatches", "samples")]


class DeviceBuffer:
    def __init__(self, gpu, nbytes):
        self.gpu = gpu
        self.nbytes = int(nbytes)
        p = vp()
        gpu._check(gpu.lib.zgpu_alloc(gpu.h, max(1, self.nbytes), ctypes.byref(p)))
        self.ptr = p

    def upload(self, arr, offset=0):
        a = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
        assert offset + a.size <= self.nbytes
        self.gpu._check(self.gpu.lib.zgpu_upload(self.gpu.h, vp(self.ptr.value + offset), _ptr(a), a.size))

    def download(self, nbytes=None, offset=0, dtype=np.uint8):
        n = self.nbytes - offset if nbytes is None else int(nbytes)
        out = np.empty(n, dtype=np.uint8)
        self.gpu._check(self.gpu.lib.zgpu_download(self.gpu.h, _ptr(out), vp(self.ptr.value + offset), n))
        return out.view(dtype)

    def addr(self, offset=0):
        return vp(self.ptr.value + offset)

    def free(self):
        if self.ptr is not None and self.ptr.value:
            self.gpu.lib.zgpu_free(self.gpu.h, self.ptr)
        self.ptr = None


class Gpu:
    """A device context (one host thread at a time)."""

    def __init__(self, ordinal=0):
        self.lib = load_library()
        h = vp()
        rc = self.lib.zgpu_open(ordinal, ctypes.byref(h))
        if rc != ZGPU_OK:
            raise DeviceError(rc, "cannot open the GPU")
        self.h = h

    def close(self):
        if self.h:
            self.lib.zgpu_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc == ZGPU_OK:
            return
        msg = self.lib.zgpu_last_error(self.h).decode(errors="replace")
        if rc == ZGPU_ERR_DAMAGED:
            raise DamagedError(rc, msg)
        if rc == ZGPU_ERR_DEVICE:
            raise DeviceError(rc, msg)
        raise ZeroError(rc, msg)

    def props(self):
        p = Props()
        self._check(self.lib.zgpu_props(self.h, ctypes.byref(p)))
        return {"name": p.name.decode(), "arch": p.arch.decode(), "cu_count": p.cu_count,
                "clock_khz": p.clock_khz, "hbm_bytes": p.hbm_bytes, "lds_per_block": p.lds_per_block}

    # ---- host-buffer forms (zero::Codec)
    def compress(self, data, block_log=18, level=5):
        """[(method, payload bytes)] per block: method 1 with the ZB2 payload, or 0 (stored, payload empty)."""
        a = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
        n = a.size
        B = 1 << block_log
        nb = (n + B - 1) // B
        if nb == 0:
            return []
        out = np.empty(max(1, n), dtype=np.uint8)
        lens = np.empty(nb, dtype=np.uint32)
        self._check(self.lib.zgpu_compress(self.h, _ptr(a), n, block_log, level, _ptr(out), out.size, _ptr(lens),
                                           None))
        res, o = [], 0
        for b in range(nb):
            if lens[b] == STORED:
                res.append((0, b""))
            else:
                res.append((1, out[o: o + int(lens[b])].tobytes()))
                o += int(lens[b])
        return res

    def decompress(self, blocks, block_log):
        """blocks: [(method, raw_len, comp bytes)] -> the raw bytes back to back. Raises DamagedError."""
        nb = len(blocks)
        if nb == 0:
            return b""
        comp = np.frombuffer(b"".join(bytes(c) for _, _, c in blocks) or b"\0", dtype=np.uint8)
        clen = np.array([len(c) for _, _, c in blocks], dtype=np.uint32)
        meth = np.array([m for m, _, _ in blocks], dtype=np.uint8)
        rlen = np.array([r for _, r, _ in blocks], dtype=np.uint32)
        out = np.empty(max(1, int(rlen.sum(dtype=np.uint64))), dtype=np.uint8)
        err = np.zeros(nb, dtype=np.uint32)
        crc = np.zeros(nb, dtype=np.uint32)
        rc = self.lib.zgpu_decompress(self.h, _ptr(comp), _ptr(clen), _ptr(meth), _ptr(rlen), nb, block_log,
                                      _ptr(out), _ptr(err), _ptr(crc))
        self.last_crc = crc
        self.last_err = err
        self._check(rc)
        return out[: int(rlen.sum(dtype=np.uint64))].tobytes()

    def encode_stream(self, sym):
        """one stream section (kind, len, body) of the symbols, by the GPU's rANS encoder (zb::encode_stream)"""
        a = np.frombuffer(bytes(sym), dtype=np.uint8) if not isinstance(sym, np.ndarray) else sym
        out = np.empty(a.size + 512, dtype=np.uint8)
        n = u32()
        self._check(self.lib.zgpu_encode_stream(self.h, _ptr(a) if a.size else None, a.size, _ptr(out),
                                                ctypes.byref(n)))
        return out[: n.value].tobytes()

    def crc32_blocks_host(self, data, block_log):
        a = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
        B = 1 << block_log
        nb = (a.size + B - 1) // B
        crc = np.zeros(max(1, nb), dtype=np.uint32)
        if nb:
            self._check(self.lib.zgpu_crc32(self.h, _ptr(a), a.size, block_log, _ptr(crc)))
        return crc[:nb]

    def crc32_ranges(self, which, ranges):
        """CRC-32 of (offset, length) ranges of the last compress input (which 0) or decompress output (which 1)."""
        off = np.array([r[0] for r in ranges], dtype=np.uint64)
        ln = np.array([r[1] for r in ranges], dtype=np.uint64)
        crc = np.zeros(max(1, len(ranges)), dtype=np.uint32)
        self._check(self.lib.zgpu_crc32_ranges(self.h, int(which), len(ranges), _ptr(off), _ptr(ln), _ptr(crc)))
        return [int(c) for c in crc[: len(ranges)]]

    # ---- device-resident forms
    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def encode_blocks(self, d_in, d_blocks, nb, block_log, level, d_out, d_out_len, stream=None):
        self._check(self.lib.zgpu_encode_blocks(self.h, d_in, d_blocks, nb, block_log, level, d_out, d_out_len,
                                                stream))

    def decode_blocks(self, d_comp, d_blocks, nb, block_log, d_out, d_err, stream=None):
        self._check(self.lib.zgpu_decode_blocks(self.h, d_comp, d_blocks, nb, block_log, d_out, d_err, stream))

    def decode_blocks_noalloc(self, d_comp, d_blocks, nb, block_log, d_out, d_err, stream=None):
        """decode_blocks with the library's fixed scratch array: allocates nothing (ZGPU_NOALLOC_SCRATCH_BYTES)"""
        self._check(self.lib.zgpu_decode_blocks_noalloc(self.h, d_comp, d_blocks, nb, block_log, d_out, d_err,
                                                        stream))

    def crc32_blocks(self, d_data, d_blocks, nb, use_dst, d_crc, stream=None):
        self._check(self.lib.zgpu_crc32_blocks(self.h, d_data, d_blocks, nb, int(use_dst), d_crc, stream))

    def sync(self):
        self._check(self.lib.zgpu_sync(self.h))

    def serve_bench(self, lz_pool, random_pool, offered_GBps=20.0, seconds=1.0, compress_share=0.7,
                    random_share=0.1, size_median=256 * 1024, size_sigma=1.5, size_max=256 << 20,
                    hbm_budget=64 << 30, block_log=18, slots=3, seed=5, alone_checks=64, dec_pools=None,
                    sample_every=0, sample_cap=256, sample_bytes=256 << 20):
        """The serving loop under an open-loop synthetic load (zgpu_serve_bench); returns its report as a dict.
        dec_pools: [(comp blob, comp_len[], method[])] for the LZ and the random pool, every 2^block_log block as the
        caller compressed it (the decompress requests' inputs). sample_every > 0: the report's "sample" holds
        (fields dict, response bytes, per-block payload sizes) of every sample_every-th request."""
        cfg = ServeCfg(offered_GBps, seconds, compress_share, random_share, size_median, size_sigma, size_max,
                       hbm_budget, block_log, slots, seed, alone_checks)
        keep = []
        if dec_pools is not None:
            for i, (blob, clen, meth) in enumerate(dec_pools):
                arrs = [np.ascontiguousarray(blob, dtype=np.uint8), np.ascontiguousarray(clen, dtype=np.uint32),
                        np.ascontiguousarray(meth, dtype=np.uint8)]
                keep += arrs
                cfg.dec_comp[i], cfg.dec_comp_len[i], cfg.dec_method[i] = (a.ctypes.data for a in arrs)
        if sample_every:
            samples = (ServeSample * sample_cap)()
            sbytes = np.empty(sample_bytes, dtype=np.uint8)
            slens = np.empty(sample_bytes // 1024 + 1024, dtype=np.uint32)
            cfg.sample_every, cfg.sample_cap, cfg.samples = sample_every, sample_cap, samples
            cfg.sample_bytes, cfg.sample_bytes_cap = sbytes.ctypes.data, sbytes.size
            cfg.sample_blk_len, cfg.sample_blk_cap = slens.ctypes.data, slens.size
        rep = ServeReport()
        lz = np.ascontiguousarray(lz_pool, dtype=np.uint8)
        rn = np.ascontiguousarray(random_pool, dtype=np.uint8)
        self._check(self.lib.zgpu_serve_bench(self.h, ctypes.byref(cfg), _ptr(lz), lz.size, _ptr(rn), rn.size,
                                              ctypes.byref(rep)))
        out = {n: getattr(rep, n) for n, _ in ServeReport._fields_}
        if sample_every:
            out["sample"] = []
            for k in range(rep.samples):
                sm = samples[k]
                f = {n: getattr(sm, n) for n, _ in ServeSample._fields_}
                body = sbytes[f["bytes_off"]: f["bytes_off"] + f["bytes_len"]].tobytes()
                lens = slens[f["blk_len_off"]: f["blk_len_off"] + f["nblk"]].copy() if f["op"] == 0 else None
                out["sample"].append((f, body, lens))
        return out

    def copy(self, dst, src, nbytes, stream=None):
        """device-to-device copy on the codec's stream (a plain kernel when both addresses are 16-byte aligned)"""
        self._check(self.lib.zgpu_copy(self.h, dst, src, nbytes, stream))

    def event(self):
        e = vp()
        self._check(self.lib.zgpu_event_create(self.h, ctypes.byref(e)))
        return e

    def record(self, ev, stream=None):
        self._check(self.lib.zgpu_event_record(self.h, ev, stream))

    def elapsed_ms(self, a, b):
        ms = ctypes.c_float()
        self._check(self.lib.zgpu_event_elapsed_ms(self.h, a, b, ctypes.byref(ms)))
        return ms.value

    def set_profile(self, on):
        self._check(self.lib.zgpu_set_profile(self.h, int(on)))

    def stage_times(self):
        names = (ctypes.c_char_p * 32)()
        ms = (ctypes.c_float * 32)()
        k = self.lib.zgpu_stage_times(self.h, names, ms, 32)
        if k < 0:
            self._check(k)
        return [(names[i].decode(), ms[i]) for i in range(k)]


class Plan:
    """The batched form (zgpu_plan_*): the blocks of many files of uneven lengths, planned together. src/raw_len
    give every block's place in the input buffer; block_log is the archive's block size."""

    def __init__(self, gpu, src, raw_len, block_log, streams=4):
        self.gpu, self.lib = gpu, gpu.lib
        self.src = np.ascontiguousarray(src, dtype=np.uint64)
        self.raw_len = np.ascontiguousarray(raw_len, dtype=np.uint32)
        self.nb = int(self.raw_len.size)
        self.block_log = block_log
        h = vp()
        gpu._check(self.lib.zgpu_plan_create(gpu.h, self.nb, _ptr(self.src), _ptr(self.raw_len), block_log, streams,
                                             ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h is not None and self.h.value:
            self.lib.zgpu_plan_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def slab_bytes(self):
        return int(self.lib.zgpu_plan_slab_bytes(self.h))

    def slab_offsets(self):
        off = np.empty(max(1, self.nb), dtype=np.uint64)
        self.gpu._check(self.lib.zgpu_plan_slab_offsets(self.h, _ptr(off)))
        return off[: self.nb]

    def classes(self):
        cap = 16
        a = [np.zeros(cap, dtype=np.uint32) for _ in range(4)]
        n = self.lib.zgpu_plan_classes(self.h, *[_ptr(x) for x in a], cap)
        return [dict(block_log=int(a[0][i]), blocks=int(a[1][i]), hash_bits=int(a[2][i]), walkers=int(a[3][i]))
                for i in range(min(n, cap))]

    def encode(self, d_in, level, d_slabs, d_out_len):
        self.gpu._check(self.lib.zgpu_plan_encode(self.h, d_in, level, d_slabs, d_out_len))

    def set_stored(self, comp_off, comp_len, method):
        self._st = [np.ascontiguousarray(comp_off, dtype=np.uint64), np.ascontiguousarray(comp_len, dtype=np.uint32),
                    np.ascontiguousarray(method, dtype=np.uint8)]
        self.gpu._check(self.lib.zgpu_plan_set_stored(self.h, *[_ptr(x) for x in self._st]))

    def decode(self, d_comp, d_out, d_err):
        self.gpu._check(self.lib.zgpu_plan_decode(self.h, d_comp, d_out, d_err))

    def tune(self, d_in, level, d_slabs, d_out_len):
        self.gpu._check(self.lib.zgpu_plan_tune(self.h, d_in, level, d_slabs, d_out_len))

    def use_graphs(self, on=True):
        self.gpu._check(self.lib.zgpu_plan_use_graphs(self.h, 1 if on else 0))

    STAGES = ("hash", "place", "sort_s", "sort_m", "sort_l", "match", "parse", "tokens", "rans", "pack",
              "dstreams", "dlz")

    def stage_occupancy(self):
        """workgroups per compute unit of each persistent stage"""
        occ = np.zeros(16, dtype=np.uint32)
        n = self.lib.zgpu_plan_stage_occupancy(self.h, _ptr(occ), 16)
        return dict(zip(self.STAGES, [int(x) for x in occ[:n]]))

    def profile(self, on=True):
        self.gpu._check(self.lib.zgpu_plan_profile(self.h, 1 if on else 0))

    def stage_ms(self):
        """each stage's time in the last profiled call, summed over the passes"""
        ms = np.zeros(16, dtype=np.float32)
        n = self.lib.zgpu_plan_stage_ms(self.h, _ptr(ms), 16)
        if n < 0:
            self.gpu._check(n)
        return {k: round(float(v), 3) for k, v in zip(self.STAGES, ms[:n]) if v > 0}

    def set_stage_occupancy(self, **occ):
        """workgroups per compute unit of the named stages"""
        cur = self.stage_occupancy()
        cur.update(occ)
        a = np.array([cur[k] for k in self.STAGES], dtype=np.uint32)
        self.gpu._check(self.lib.zgpu_plan_set_stage_occupancy(self.h, _ptr(a), a.size))

    @property
    def passes(self):
        return int(self.lib.zgpu_plan_passes(self.h))

    def set_pass_bytes(self, n):
        """at most n slab bytes per pass (the encoder's scratch is about 19 n)"""
        self.gpu._check(self.lib.zgpu_plan_pass_bytes(self.h, int(n)))

    def match_workgroup(self, threads=0):
        """the match stage's threads per workgroup; sets it first when threads is 64..1024"""
        r = int(self.lib.zgpu_plan_match_workgroup(self.h, int(threads)))
        if r < 0:
            self.gpu._check(r)
        return r


def file_blocks(sizes, block_log):
    """each file cut into blocks of 2^block_log (the last one shorter), files back to back: (src, raw_len, file)"""
    B = 1 << block_log
    src, rl, fi = [], [], []
    off = 0
    for k, n in enumerate(sizes):
        for o in range(0, int(n), B):
            src.append(off + o)
            rl.append(min(B, int(n) - o))
            fi.append(k)
        off += int(n)
    return (np.array(src, dtype=np.uint64), np.array(rl, dtype=np.uint32), np.array(fi, dtype=np.uint32))


def block_table(raw_lens, comp_lens=None, methods=None, src=None, dst=None):
    """A zgpu_block_t array: src/dst default to back-to-back placement of comp (or raw) and raw bytes."""
    raw = np.asarray(raw_lens, dtype=np.uint64)
    t = np.zeros(raw.size, dtype=BLOCK_DTYPE)
    t["raw_len"] = raw
    comp = raw if comp_lens is None else np.asarray(comp_lens, dtype=np.uint64)
    t["comp_len"] = comp if comp_lens is not None else 0
    t["method"] = 0 if methods is None else np.asarray(methods)
    t["src"] = np.concatenate([[0], np.cumsum(comp)[:-1]]) if src is None else src
    t["dst"] = np.concatenate([[0], np.cumsum(raw)[:-1]]) if dst is None else dst
    return t


# ---- synthetic workloads (csrc/gen.cpp)
def _gen_lib():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise ZeroError(ZGPU_ERR_ARG, f"{GEN_PATH} is not built")
        g = ctypes.CDLL(GEN_PATH)
        for name in ("zgen_lz", "zgen_random", "zgen_exec"):
            getattr(g, name).argtypes = [u64, vp, u64]
            getattr(g, name).restype = None
        g.zgen_lognormal_sizes.argtypes = [u64, ctypes.c_double, ctypes.c_double, u64, u64, vp, u64]
        g.zgen_lognormal_sizes.restype = None
        _gen = g
    return _gen


def gen_lz(seed, n):
    """LZ-friendly bytes: literal runs (geometric(16), Zipf(1.1) bytes) and back-copies (4+geometric(20) <= 258 at
    1+geometric(4096) <= 2^18), each with p = 0.5."""
    out = np.empty(n, dtype=np.uint8)
    _gen_lib().zgen_lz(seed, _ptr(out), n)
    return out


def gen_random(seed, n):
    out = np.empty(n, dtype=np.uint8)
    _gen_lib().zgen_random(seed, _ptr(out), n)
    return out


def gen_exec(seed, n):
    """4 KiB pages: 30 % zeros, 40 % uint32 arrays with N(0, 64) deltas, 30 % copies of earlier spans."""
    out = np.empty(n, dtype=np.uint8)
    _gen_lib().zgen_exec(seed, _ptr(out), n)
    return out


def gen_lognormal_sizes(seed, median, sigma, lo, hi, count):
    out = np.empty(count, dtype=np.uint64)
    _gen_lib().zgen_lognormal_sizes(seed, median, sigma, lo, hi, _ptr(out), count)
    return out
