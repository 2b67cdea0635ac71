"""The BGEO reader of the dataset IO (Dataset::Dataset / ReadFile,
Dataset.cpp:169-227, 292-306; Partio::readBGEO, vendor/partio/src/io/
BGEO.cpp:187-300) through the C ABI's host-side pm_bgeo_read: the reference's
own Partio fixtures and golden values, gzip, synthetic files with every
attribute kind Partio accepts, and the refusals.  No GPU."""
import gzip
import pathlib
import struct

import numpy as np
import pytest

import pmarch
from pmarch import _lib, scenes

FIX = pathlib.Path(__file__).resolve().parent / "golden" / "bgeo"


def write_bgeo(path, xyz, attrs=(), gz=False):
    """A classic BGEO V5 file: header, point attributes (name, size, type,
    defaults or indexed strings), points (P = x, y, z, 1 and the attribute
    words), no primitives, no detail attributes."""
    xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
    n = xyz.shape[0]
    out = bytearray(b"Bgeo" + b"V" + struct.pack(">9i", 5, n, 0, 0, 0, len(attrs), 0, 0, 0))
    words = []
    for name, size, typ, extra in attrs:
        out += struct.pack(">H", len(name)) + name.encode() + struct.pack(">Hi", size, typ)
        if typ in (0, 1, 5):
            out += struct.pack(f">{size}i", *([0] * size))
        elif typ == 4:
            out += struct.pack(">i", len(extra))
            for st in extra:
                out += struct.pack(">H", len(st)) + st.encode()
        elif typ == 2:
            pass
        words.append(size)
    for i in range(n):
        out += struct.pack(">4f", *xyz[i], 1.0)
        for k, size in enumerate(words):
            out += struct.pack(f">{size}i", *([i + k] * size))
    data = gzip.compress(bytes(out)) if gz else bytes(out)
    pathlib.Path(path).write_bytes(data)


def test_partio_fixtures_golden_values():
    base = _lib.bgeo_read(FIX / "base.bgeo")
    delta = _lib.bgeo_read(FIX / "delta.bgeo")
    assert base.shape == (5, 3) and delta.shape == (3, 3)
    assert np.array_equal(base[:, 0], np.array([0.0, 0.1, 0.2, 0.3, 0.4], np.float32))
    assert np.array_equal(delta[:, 0], np.array([0.1, 0.3, 0.5], np.float32))
    assert _lib.bgeo_read(FIX / "scatter.bgeo").shape[0] > 0
    assert _lib.bgeo_read(FIX / "baseidsfaceid.bgeo").shape[0] > 0


def test_gzip_and_plain_agree(tmp_path):
    raw = (FIX / "base.bgeo").read_bytes()
    (tmp_path / "base.bgeo.gz").write_bytes(gzip.compress(raw))
    assert np.array_equal(_lib.bgeo_read(tmp_path / "base.bgeo.gz"), _lib.bgeo_read(FIX / "base.bgeo"))


@pytest.mark.parametrize("gz", [False, True])
def test_synthetic_roundtrip_with_attributes(tmp_path, gz):
    sc = scenes.lattice_cube(0)
    attrs = [("v", 3, 5, None), ("id", 1, 1, None), ("life", 1, 0, None), ("state", 1, 4, ["a", "bb", "ccc"])]
    f = tmp_path / ("lattice.bgeo.gz" if gz else "lattice.bgeo")
    write_bgeo(f, sc.xyz, attrs, gz=gz)
    assert np.array_equal(_lib.bgeo_read(f), sc.xyz)


def test_refusals(tmp_path):
    L = pmarch.load()
    import ctypes as C
    n = C.c_uint64(0)
    assert L.pm_bgeo_read(str(FIX / "test.geo").encode(), None, 0, C.byref(n)) == -7        # ASCII GEO
    assert L.pm_bgeo_read(str(tmp_path / "missing.bgeo").encode(), None, 0, C.byref(n)) == -7
    write_bgeo(tmp_path / "str.bgeo", np.zeros((4, 3)), [("name", 1, 2, None)])
    assert L.pm_bgeo_read(str(tmp_path / "str.bgeo").encode(), None, 0, C.byref(n)) == -7   # string attribute
    raw = (FIX / "base.bgeo").read_bytes()
    (tmp_path / "cut.bgeo").write_bytes(raw[:100])
    xyz = np.zeros((16, 3), np.float32)
    assert L.pm_bgeo_read(str(tmp_path / "cut.bgeo").encode(), xyz.ctypes.data, 16, C.byref(n)) == -7
    assert L.pm_bgeo_read(str(FIX / "base.bgeo").encode(), xyz.ctypes.data, 4, C.byref(n)) == -4   # cap < n
    assert L.pm_bgeo_read(None, None, 0, C.byref(n)) == -1


def test_dataset_frame_files(tmp_path):
    """Dataset::Dataset enumerates prefix + i + suffix for i = 1, 2, ... while
    the file exists (Dataset.cpp:181-200)."""
    for i in (1, 2, 3, 5):
        write_bgeo(tmp_path / f"frame_{i}.bgeo", np.full((2, 3), i, np.float32))
    files = pmarch.dataset_files(str(tmp_path / "frame_"), ".bgeo")
    assert [f.name for f in files] == ["frame_1.bgeo", "frame_2.bgeo", "frame_3.bgeo"]
    assert len(pmarch.dataset_files(str(tmp_path / "frame_"), ".bgeo", count=2)) == 2
    frames = pmarch.read_dataset(str(tmp_path / "frame_"), ".bgeo")
    assert [float(x[0, 0]) for x in frames] == [1.0, 2.0, 3.0]
