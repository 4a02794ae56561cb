"""The container's file checksums from per-batch range CRCs (host/archive.h FileCrcs), on the CPU: a stand-in codec
(tests/host/container_check.cpp) stores blocks and answers range_crcs() from the host copy of each batch. Archives it
writes are read back by the test reader, which checks every file's CRC with zlib."""
import os
import random
import subprocess

import pytest

from zero_archive import read_archive, write_archive

HERE = os.path.dirname(os.path.abspath(__file__))
HOST = os.path.join(HERE, "..", "zero-compression_amd", "host")


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bin") / "container_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-pthread", "-I", HOST, "-I", os.path.join(HERE, "..", "include"),
                    "-o", exe, os.path.join(HERE, "host", "container_check.cpp")], check=True)
    return exe


def tree(root):
    rng = random.Random(9)
    files = {
        "t/a.bin": rng.randbytes(200000),   # spans batches of 3 x 64 KiB
        "t/empty": b"",
        "t/s/b.txt": b"abc" * 50001,
        "t/s/c.bin": rng.randbytes(65536),  # exactly one block
        "t/z": b"z",
    }
    for name, d in files.items():
        os.makedirs(os.path.dirname(os.path.join(root, name)), exist_ok=True)
        with open(os.path.join(root, name), "wb") as f:
            f.write(d)
    return files


@pytest.mark.parametrize("host_sums", [False, True])
def test_file_crcs_from_ranges_round_trip(tool, tmp_path, host_sums):
    """files' CRCs from the codec's range CRCs, or (a codec without them) from the host; batches of 3 blocks, so
    files and their pieces cross batch boundaries and the pipeline's slots are all reused"""
    files = tree(str(tmp_path))
    env = {k: v for k, v in os.environ.items() if k != "CHECK_HOST_SUMS"}
    if host_sums:
        env["CHECK_HOST_SUMS"] = "1"
    arc = str(tmp_path / "a.zero")
    r = subprocess.run([tool, "c", arc, "16", "t"], cwd=str(tmp_path), capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr
    assert "range_crc_calls=" in r.stdout and r.stdout.strip().endswith("=0") == host_sums
    _, entries = read_archive(arc)  # (checks every block's and every file's CRC with zlib)
    assert {n: d for n, d in entries if d is not None} == files
    r = subprocess.run([tool, "t", arc], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr
    out = str(tmp_path / "out")
    r = subprocess.run([tool, "x", arc, out], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr
    for name, d in files.items():
        assert open(os.path.join(out, name), "rb").read() == d


def test_file_crc_mismatch_caught(tool, tmp_path):
    """a table CRC that disagrees with intact blocks is reported by t and x (random files: every block is stored,
    which the stand-in codec reads)"""
    rng = random.Random(4)
    files = [("d", None), ("d/x", rng.randbytes(150000)), ("d/e", b""), ("d/y", rng.randbytes(70000))]
    good, bad = str(tmp_path / "good.zero"), str(tmp_path / "bad.zero")
    write_archive(good, files, 16, 8)
    write_archive(bad, files, 16, 8, file_crc={"d/y": 1})
    r = subprocess.run([tool, "t", good], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for args in (["t", bad], ["x", bad, str(tmp_path / "out")]):
        r = subprocess.run([tool] + args, capture_output=True, text=True)
        assert r.returncode == 1 and "d/y does not match its checksum" in r.stderr, r.stderr
