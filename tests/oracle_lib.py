"""ctypes access to the oracle (oracle/liboracle.so, our C restatement) and to the compiled reference
(oracle/_ref/libzeroref.so). Test infrastructure: the checker, never the thing measured."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libzeroref.so")

vp, u8p, u32, u64, i32, sz = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t


class Token(ctypes.Structure):
    _fields_ = [("i", u32), ("m", u32), ("d", u32)]


def _buf(data):
    a = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else np.ascontiguousarray(data)
    return a, ctypes.c_void_p(a.ctypes.data if a.size else 0)


class Oracle:
    def __init__(self, path=ORACLE):
        L = ctypes.CDLL(path)
        L.zo_crc32_update.restype = u32
        L.zo_crc32_update.argtypes = [u32, vp, sz]
        L.zo_crc32_combine.restype = u32
        L.zo_crc32_combine.argtypes = [u32, u32, u64]
        L.zo_normalize.argtypes = [vp, u32, vp]
        L.zo_encode_stream.restype = sz
        L.zo_encode_stream.argtypes = [vp, u32, vp]
        L.zo_decode_stream.restype = ctypes.c_long
        L.zo_decode_stream.argtypes = [vp, sz, u32, vp]
        L.zo_cpu_compress_block.restype = sz
        L.zo_cpu_compress_block.argtypes = [vp, u32, i32, vp, ctypes.POINTER(i32)]
        L.zo_decompress_block.restype = i32
        L.zo_decompress_block.argtypes = [vp, sz, vp, u32]
        L.zo_gpu_match.argtypes = [vp, u32, u32, i32, vp]
        L.zo_gpu_lazy.argtypes = [vp, u32, vp]
        L.zo_gpu_parse.restype = sz
        L.zo_gpu_parse.argtypes = [vp, u32, i32, vp, vp]
        L.zo_gpu_compress_block.restype = sz
        L.zo_gpu_compress_block.argtypes = [vp, u32, u32, i32, vp, ctypes.POINTER(i32)]
        L.zo_pack_tokens.restype = sz
        L.zo_pack_tokens.argtypes = [vp, u32, vp, sz, vp]
        L.zo_cpu_parse.restype = sz
        L.zo_cpu_parse.argtypes = [vp, u32, i32, vp]
        self.L = L

    def crc32(self, data, crc=0):
        a, p = _buf(data)
        return self.L.zo_crc32_update(crc, p, a.size)

    def crc32_combine(self, a, b, len_b):
        return self.L.zo_crc32_combine(a, b, len_b)

    def normalize(self, counts):
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        f = np.zeros(256, dtype=np.uint16)
        self.L.zo_normalize(ctypes.c_void_p(c.ctypes.data), int(c.sum()), ctypes.c_void_p(f.ctypes.data))
        return f

    def encode_stream(self, sym):
        a, p = _buf(sym)
        out = np.zeros(a.size + 1024, dtype=np.uint8)
        n = self.L.zo_encode_stream(p, a.size, ctypes.c_void_p(out.ctypes.data))
        return out[:n].tobytes()

    def decode_stream(self, sec, n):
        a, p = _buf(sec)
        out = np.zeros(max(1, n), dtype=np.uint8)
        r = self.L.zo_decode_stream(p, a.size, n, ctypes.c_void_p(out.ctypes.data))
        return r, out[:n].tobytes()

    def cpu_compress_block(self, data, depth):
        a, p = _buf(data)
        out = np.zeros(4 * a.size + 4096, dtype=np.uint8)
        st = i32()
        n = self.L.zo_cpu_compress_block(p, a.size, depth, ctypes.c_void_p(out.ctypes.data), ctypes.byref(st))
        return out[:n].tobytes(), bool(st.value)

    def gpu_compress_block(self, data, block_log, level):
        a, p = _buf(data)
        out = np.zeros(4 * a.size + 4096, dtype=np.uint8)
        st = i32()
        n = self.L.zo_gpu_compress_block(p, a.size, block_log, level, ctypes.c_void_p(out.ctypes.data), ctypes.byref(st))
        return out[:n].tobytes(), bool(st.value)

    def gpu_match(self, data, block_log, level):
        a, p = _buf(data)
        m = np.zeros(max(1, a.size), dtype=np.uint32)
        self.L.zo_gpu_match(p, a.size, block_log, level, ctypes.c_void_p(m.ctypes.data))
        return m[: a.size]

    def gpu_tokens(self, data, block_log, level):
        a, p = _buf(data)
        m = np.zeros(max(1, a.size), dtype=np.uint32)
        lz = np.zeros(max(1, a.size), dtype=np.uint32)
        self.L.zo_gpu_match(p, a.size, block_log, level, ctypes.c_void_p(m.ctypes.data))
        self.L.zo_gpu_lazy(ctypes.c_void_p(m.ctypes.data), a.size, ctypes.c_void_p(lz.ctypes.data))
        tk = (Token * (a.size // 4 + 1))()
        k = self.L.zo_gpu_parse(p, a.size, level, ctypes.c_void_p(lz.ctypes.data), tk)
        return [(tk[t].i, tk[t].m, tk[t].d) for t in range(k)]

    def cpu_tokens(self, data, depth):
        a, p = _buf(data)
        tk = (Token * (a.size // 4 + 1))()
        k = self.L.zo_cpu_parse(p, a.size, depth, tk)
        return [(tk[t].i, tk[t].m, tk[t].d) for t in range(k)]

    def pack_tokens(self, data, tokens):
        a, p = _buf(data)
        tk = (Token * max(1, len(tokens)))()
        for t, (i, m, d) in enumerate(tokens):
            tk[t].i, tk[t].m, tk[t].d = i, m, d
        out = np.zeros(4 * a.size + 4096, dtype=np.uint8)
        n = self.L.zo_pack_tokens(p, a.size, tk, len(tokens), ctypes.c_void_p(out.ctypes.data))
        return out[:n].tobytes()

    def decompress_block(self, payload, n):
        a, p = _buf(payload)
        out = np.zeros(max(1, n), dtype=np.uint8)
        rc = self.L.zo_decompress_block(p, a.size, ctypes.c_void_p(out.ctypes.data), n)
        return rc, out[:n].tobytes()


def ref_available():
    return os.path.exists(REF)


class Ref:
    """The reference's own CPU codec (src/codec.h, codec2.h, crc32.h) behind oracle/ref_driver.cpp."""

    def __init__(self, path=REF):
        L = ctypes.CDLL(path)
        L.ref_crc32.restype = u32
        L.ref_crc32.argtypes = [vp, sz]
        L.ref_crc32_combine.restype = u32
        L.ref_crc32_combine.argtypes = [u32, u32, u64]
        L.ref_normalize.argtypes = [vp, u32, vp]
        L.ref_encode_stream.restype = sz
        L.ref_encode_stream.argtypes = [vp, u32, vp]
        L.ref_decode_stream.restype = ctypes.c_long
        L.ref_decode_stream.argtypes = [vp, sz, u32, vp]
        L.ref_compress_block.restype = sz
        L.ref_compress_block.argtypes = [vp, u32, i32, vp, sz, ctypes.POINTER(i32)]
        L.ref_decompress_block.restype = i32
        L.ref_decompress_block.argtypes = [vp, sz, vp, u32]
        L.ref_cpu_compress.restype = ctypes.c_long
        L.ref_cpu_compress.argtypes = [vp, sz, u32, i32, i32, vp, sz, vp, vp]
        L.ref_cpu_decompress.restype = i32
        L.ref_cpu_decompress.argtypes = [vp, vp, vp, vp, sz, i32, vp]
        self.L = L

    def crc32(self, data):
        a, p = _buf(data)
        return self.L.ref_crc32(p, a.size)

    def normalize(self, counts):
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        f = np.zeros(256, dtype=np.uint16)
        self.L.ref_normalize(ctypes.c_void_p(c.ctypes.data), int(c.sum()), ctypes.c_void_p(f.ctypes.data))
        return f

    def encode_stream(self, sym):
        a, p = _buf(sym)
        out = np.zeros(a.size + 1024, dtype=np.uint8)
        n = self.L.ref_encode_stream(p, a.size, ctypes.c_void_p(out.ctypes.data))
        return out[:n].tobytes()

    def compress_block(self, data, depth):
        a, p = _buf(data)
        cap = 4 * a.size + 4096
        out = np.zeros(cap, dtype=np.uint8)
        st = i32()
        n = self.L.ref_compress_block(p, a.size, depth, ctypes.c_void_p(out.ctypes.data), cap, ctypes.byref(st))
        return out[:n].tobytes(), bool(st.value)

    def decompress_block(self, payload, n):
        a, p = _buf(payload)
        out = np.zeros(max(1, n), dtype=np.uint8)
        rc = self.L.ref_decompress_block(p, a.size, ctypes.c_void_p(out.ctypes.data), n)
        return rc, out[:n].tobytes()

    def cpu_compress(self, data, block_log, depth, threads):
        """CpuCodec::compress: (blob, comp_len[], method[])"""
        a, p = _buf(data)
        B = 1 << block_log
        nb = (a.size + B - 1) // B
        out = np.zeros(a.size + 64, dtype=np.uint8)
        cl = np.zeros(max(1, nb), dtype=np.uint32)
        me = np.zeros(max(1, nb), dtype=np.uint8)
        n = self.L.ref_cpu_compress(p, a.size, block_log, depth, threads, ctypes.c_void_p(out.ctypes.data), out.size,
                                    ctypes.c_void_p(cl.ctypes.data), ctypes.c_void_p(me.ctypes.data))
        assert n >= 0
        return out[:n], cl[:nb], me[:nb]

    def cpu_decompress(self, blob, comp_len, method, raw_len, threads):
        raw_len = np.ascontiguousarray(raw_len, dtype=np.uint32)
        out = np.zeros(max(1, int(raw_len.sum())), dtype=np.uint8)
        blob = np.ascontiguousarray(blob)
        comp_len = np.ascontiguousarray(comp_len, dtype=np.uint32)
        method = np.ascontiguousarray(method, dtype=np.uint8)
        rc = self.L.ref_cpu_decompress(ctypes.c_void_p(blob.ctypes.data), ctypes.c_void_p(comp_len.ctypes.data),
                                       ctypes.c_void_p(method.ctypes.data), ctypes.c_void_p(raw_len.ctypes.data),
                                       comp_len.size, threads, ctypes.c_void_p(out.ctypes.data))
        return rc, out[: int(raw_len.sum())]
