"""The zero command line and the .zero container: CPU-only checks here, both-way archive parity on the GPU."""
import os
import random
import subprocess

import pytest

from zero_archive import read_archive, write_archive

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ZERO = os.path.join(ROOT, "zero-compression_amd", "zero")


def run(*args, cwd=None):
    return subprocess.run([ZERO, *args], capture_output=True, text=True, cwd=cwd, timeout=120)


def sample_tree(root):
    rng = random.Random(3)
    os.makedirs(os.path.join(root, "tree", "sub", "deeper"))
    files = {
        "tree/a.txt": b"hello world " * 5000,
        "tree/empty.bin": b"",
        "tree/sub/random.bin": rng.randbytes(300000),
        "tree/sub/deeper/runs.bin": bytes(100000) + b"\x01" * 70000 + b"xyz" * 30000,
        "tree/z_last.txt": b"the end\n",
    }
    for name, d in files.items():
        with open(os.path.join(root, name), "wb") as f:
            f.write(d)
    return files


def test_list_reference_archive(tmp_path):
    entries = [("d", None), ("d/x.bin", b"abc" * 100000), ("d/y.bin", bytes(range(256)) * 2000)]
    arc = str(tmp_path / "a.zero")
    write_archive(arc, entries, 16, 32)
    r = run("l", arc)
    assert r.returncode == 0, r.stderr
    assert "3 entries" in r.stdout and "d/x.bin" in r.stdout and "300000" in r.stdout


def test_damaged_archive_refused_without_gpu(tmp_path):
    arc = str(tmp_path / "a.zero")
    write_archive(arc, [("f", b"data" * 1000)], 16, 32)
    raw = bytearray(open(arc, "rb").read())
    raw[-30] ^= 0xFF  # the table
    open(arc, "wb").write(bytes(raw))
    r = run("l", arc)
    assert r.returncode == 1 and "archive damaged" in r.stderr


def test_cpu_option_refused():
    r = run("t", "x.zero", "--cpu")
    assert r.returncode == 1 and "--cpu" in r.stderr


@pytest.mark.gpu
def test_gpu_archives_both_ways(tmp_path):
    files = sample_tree(str(tmp_path))
    for bl, level in ((16, 1), (18, 5), (20, 9)):
        arc = str(tmp_path / f"ours_{bl}.zero")
        r = run("c", "tree", "-o", arc, "--block", str(bl), "--level", str(level), "-q", cwd=str(tmp_path))
        assert r.returncode == 0, r.stderr
        # our archive, read back by the reference CPU codec (the oracle is byte-identical to it)
        got_bl, entries = read_archive(arc)
        assert got_bl == bl
        got = {n: d for n, d in entries if d is not None}
        assert got == files
        r = run("t", arc, "-q")
        assert r.returncode == 0 and "OK" in r.stdout, r.stderr
        out = str(tmp_path / f"x_{bl}")
        r = run("x", arc, "-o", out, "-q")
        assert r.returncode == 0, r.stderr
        for name, d in files.items():
            assert open(os.path.join(out, name), "rb").read() == d
    # the reference CPU codec's archive, extracted on the GPU
    ref_arc = str(tmp_path / "ref.zero")
    write_archive(ref_arc, [("tree", None)] + [(n, d) for n, d in sorted(files.items())], 18, 32)
    out = str(tmp_path / "from_ref")
    r = run("x", ref_arc, "-o", out, "-q")
    assert r.returncode == 0, r.stderr
    for name, d in files.items():
        assert open(os.path.join(out, name), "rb").read() == d


@pytest.mark.gpu
def test_gpu_file_checksums(tmp_path):
    """files' CRCs come from the GPU: a table entry that disagrees with intact blocks is caught, in t and x"""
    files = [("a", b"x" * 100000), ("b", bytes(range(256)) * 700), ("c", b""), ("d", b"tail")]
    good = str(tmp_path / "good.zero")
    write_archive(good, files, 16, 8)
    assert run("t", good, "-q").returncode == 0
    bad = str(tmp_path / "bad.zero")
    write_archive(bad, files, 16, 8, file_crc={"b": 12345})
    for args in (("t", bad, "-q"), ("x", bad, "-o", str(tmp_path / "out"), "-q")):
        r = run(*args)
        assert r.returncode == 1 and "b does not match its checksum" in r.stderr, (args, r.stderr)


@pytest.mark.gpu
def test_gpu_damaged_archive_reports(tmp_path):
    arc = str(tmp_path / "a.zero")
    write_archive(arc, [("f", (b"abcdefgh" * 30000) + bytes(range(256)) * 100)], 16, 32)
    raw = bytearray(open(arc, "rb").read())
    rng = random.Random(2)
    for k in range(20):
        b = bytearray(raw)
        for _ in range(rng.randrange(1, 10)):
            b[rng.randrange(24, len(b) - 60)] = rng.randrange(256)
        p = str(tmp_path / f"d{k}.zero")
        open(p, "wb").write(bytes(b))
        r = run("t", p, "-q")
        assert r.returncode in (0, 1)
        assert "GPU" not in r.stderr or "damaged" in r.stderr or "checksum" in r.stderr
        if r.returncode == 1:
            assert "damaged" in r.stderr or "checksum" in r.stderr or "archive" in r.stderr
