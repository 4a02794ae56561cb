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


def env_with(**kv):
    e = {k: v for k, v in os.environ.items() if k not in ("ZERO_FAIL_GPU", "ZERO_BATCH_BYTES", "ZERO_PROFILE")}
    e.update({k: str(v) for k, v in kv.items()})
    return e


@pytest.mark.gpu
def test_gpu_many_batches_and_profile_lines(tmp_path):
    """batches of 3 blocks: files and their pieces cross batches, every pipeline slot and both codec buffers are
    reused; the archive is read back by the oracle's reader; ZERO_PROFILE prints the reference's [profile] lines"""
    files = sample_tree(str(tmp_path))
    arc = str(tmp_path / "m.zero")
    r = subprocess.run([ZERO, "c", "tree", "-o", arc, "--block", "16", "-q"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120, env=env_with(ZERO_BATCH_BYTES=3 << 16, ZERO_PROFILE=1))
    assert r.returncode == 0, r.stderr
    assert "[profile] compress" in r.stderr and "[profile] main, up to teardown" in r.stderr, r.stderr
    _, entries = read_archive(arc)
    assert {n: d for n, d in entries if d is not None} == files
    out = str(tmp_path / "mx")
    r = subprocess.run([ZERO, "x", arc, "-o", out, "-q"], capture_output=True, text=True, timeout=120,
                       env=env_with(ZERO_BATCH_BYTES=2 << 16, ZERO_PROFILE=1))
    assert r.returncode == 0, r.stderr
    assert "[profile] decompress" in r.stderr and "[profile] extract (s):" in r.stderr
    for name, d in files.items():
        assert open(os.path.join(out, name), "rb").read() == d


@pytest.mark.gpu
def test_gpu_failure_stops_the_job_and_removes_the_archive(tmp_path):
    """ZERO_FAIL_GPU=n fails the GPU's batch n + 1 as a lost device does (src/safe_codec.h:4,17,39). There is no CPU
    codec to finish on: the job exits with the device error, and the unfinished archive is removed
    (src/archive.h:347-352), on the first batch and after batches were already written"""
    sample_tree(str(tmp_path))
    for fail_after in (0, 2):
        arc = str(tmp_path / f"f{fail_after}.zero")
        r = subprocess.run([ZERO, "c", "tree", "-o", arc, "--block", "16", "-q"], cwd=str(tmp_path),
                           capture_output=True, text=True, timeout=120,
                           env=env_with(ZERO_FAIL_GPU=fail_after, ZERO_BATCH_BYTES=1 << 16))
        assert r.returncode == 1 and "the GPU was reset" in r.stderr, (fail_after, r.stderr)
        assert not os.path.exists(arc), fail_after
    arc = str(tmp_path / "good.zero")
    assert run("c", "tree", "-o", arc, "--block", "16", "-q", cwd=str(tmp_path)).returncode == 0
    r = subprocess.run([ZERO, "t", arc, "-q"], capture_output=True, text=True, timeout=120,
                       env=env_with(ZERO_FAIL_GPU=1, ZERO_BATCH_BYTES=1 << 16))
    assert r.returncode == 1 and "the GPU was reset" in r.stderr, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("bl", [16, 18, 20])
def test_gpu_damaged_archives_across_sizes_and_levels(tmp_path, bl):
    """archives our CLI writes at 2^bl and levels 1 / 5 / 9, damaged at random (tests/stress.py:119-138 does this to
    the reference): t never crashes or hangs and never reports a GPU fault; it says OK only for an archive that
    the oracle's reader also accepts, with the same files, and otherwise names the damage"""
    d = os.path.join(str(tmp_path), "in")
    os.makedirs(d)
    import pyzero
    rng = random.Random(bl)
    open(os.path.join(d, "lz.bin"), "wb").write(pyzero.gen_lz(bl, 3 << bl).tobytes())
    open(os.path.join(d, "exec.bin"), "wb").write(pyzero.gen_exec(bl + 1, (2 << bl) + 777).tobytes())
    for level in (1, 5, 9):
        arc = str(tmp_path / f"a{level}.zero")
        r = run("c", "in", "-o", arc, "--block", str(bl), "--level", str(level), "-q", cwd=str(tmp_path))
        assert r.returncode == 0, r.stderr
        raw = open(arc, "rb").read()
        for k in range(12):
            b = bytearray(raw)
            for _ in range(rng.randrange(1, 6)):
                b[rng.randrange(len(b))] = rng.randrange(256)
            p = str(tmp_path / f"d{level}_{k}.zero")
            open(p, "wb").write(bytes(b))
            r = run("t", p, "-q")
            print(f"2^{bl} L{level} #{k}: rc {r.returncode} {r.stderr.strip()[-90:]}", flush=True)
            assert r.returncode in (0, 1), (level, k, r.returncode, r.stderr)
            assert "the GPU was reset" not in r.stderr, (level, k, r.stderr)
            if r.returncode == 0:
                _, entries = read_archive(p)  # (raises if the oracle's reader sees damage)
                assert {n: x for n, x in entries if x is not None} == \
                    {"in/lz.bin": open(os.path.join(d, "lz.bin"), "rb").read(),
                     "in/exec.bin": open(os.path.join(d, "exec.bin"), "rb").read()}
            else:
                assert "damaged" in r.stderr or "checksum" in r.stderr or "archive" in r.stderr, r.stderr
