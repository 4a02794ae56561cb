"""The C-ABI library loads and exports every symbol include/zero_hip.h declares (no GPU calls)."""
import ctypes
import os
import re

import pyzero

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "zero_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"\b(zgpu_\w+)\s*\(", src))


def test_header_and_binding_agree():
    assert declared() == set(pyzero.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pyzero.LIB_PATH)
    missing = [s for s in sorted(declared()) if not hasattr(lib, s)]
    assert not missing, missing


def test_scratch_sizes():
    lib = pyzero.load_library()
    # about 19 bytes of encoder scratch per input byte, under 2 for the decoder
    for bl in (16, 18, 20):
        nb = 64
        e = lib.zgpu_encode_scratch_bytes(nb, bl)
        d = lib.zgpu_decode_scratch_bytes(nb, bl)
        assert 10 * (nb << bl) < e < 24 * (nb << bl)
        assert (nb << bl) < d < 3 * (nb << bl)
        # (the encoder runs at most 2^31 bytes of blocks per launch, over scratch for one launch)
        per = (1 << 31) >> bl
        assert lib.zgpu_encode_scratch_bytes(4 * per, bl) == lib.zgpu_encode_scratch_bytes(per, bl)
        assert lib.zgpu_encode_scratch_bytes(per, bl) > lib.zgpu_encode_scratch_bytes(per // 2, bl)
