"""The BGEO reader parses files it does not control: built with
AddressSanitizer and UBSan (host code, g++), it reads the fixtures, every
truncation of a fixture and of a synthetic file with all attribute kinds,
seeded byte flips, and a header that claims 2^31 - 1 points with no data,
without a memory error and without allocating for points that are not there.
No GPU."""
import pathlib
import shutil
import struct
import subprocess

import numpy as np
import pytest

from test_bgeo import FIX, write_bgeo

ROOT = pathlib.Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("asan")
    exe = d / "bgeo_harness"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=all", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "bgeo_harness.cpp"),
           str(ROOT / "bachelor-thesis_amd" / "csrc" / "pm_bgeo.cpp"), "-lz", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def _run(exe, files):
    r = subprocess.run([str(exe)] + [str(f) for f in files], capture_output=True, text=True, timeout=300,
                       env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    return [tuple(int(v) for v in line.split()) for line in r.stdout.splitlines()]


def test_fixtures_and_truncations(harness, tmp_path):
    files = sorted(FIX.glob("*.bgeo")) + [FIX / "test.geo"]
    syn = tmp_path / "syn.bgeo"
    write_bgeo(syn, np.arange(30, dtype=np.float32).reshape(10, 3),
               [("v", 3, 5, None), ("id", 1, 1, None), ("s", 1, 4, ["x", "yy"])])
    for src in (FIX / "base.bgeo", syn):
        raw = src.read_bytes()
        for k in range(len(raw)):
            f = tmp_path / f"{src.stem}_cut{k}.bgeo"
            f.write_bytes(raw[:k])
            files.append(f)
    out = _run(harness, files)
    assert len(out) == len(files)
    ok = {f.name: o for f, o in zip(files, out)}
    assert ok["base.bgeo"] == (0, 0, 5, 5)
    # a cut inside the header or the point records is refused; the synthetic file ends with its points,
    # base.bgeo has primitive and detail sections after them, which positions do not need (Partio does
    # not check its reads of those either): cuts there still give every point
    syn_cuts = [ok[f"syn_cut{k}.bgeo"] for k in range(len(syn.read_bytes()))]
    assert all(o[0] != 0 and o[1] != 0 for o in syn_cuts)
    base_cuts = [ok[f"base_cut{k}.bgeo"] for k in range(len((FIX / "base.bgeo").read_bytes()))]
    first = next(k for k, o in enumerate(base_cuts) if o[1] == 0)
    assert first >= 41 + 5 * 16                                         # header + at least P of 5 points
    assert all(o[1] != 0 for o in base_cuts[:first])
    assert all(o == (0, 0, 5, 5) for o in base_cuts[first:])


def test_byte_flips_and_huge_header(harness, tmp_path):
    rng = np.random.default_rng(17)
    raw = bytearray((FIX / "scatter.bgeo").read_bytes())
    files = []
    for k in range(300):
        b = bytearray(raw)
        for _ in range(1 + k % 4):
            b[rng.integers(0, len(b))] = rng.integers(0, 256)
        f = tmp_path / f"flip{k}.bgeo"
        f.write_bytes(bytes(b))
        files.append(f)
    huge = tmp_path / "huge.bgeo"
    huge.write_bytes(b"Bgeo" + b"V" + struct.pack(">9i", 5, 2 ** 31 - 1, 0, 0, 0, 0, 0, 0, 0))
    files.append(huge)
    out = _run(harness, files)
    assert out[-1][1] == -7                                             # no data behind the claimed points
