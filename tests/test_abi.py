"""The C ABI library loads, exports every function include/pmarch.h declares, and
fails loudly without a GPU -- no GPU compute is called here."""
import ctypes
import re

import numpy as np
import pytest

import pmarch
from pmarch import _lib


def test_header_functions_exported():
    L = pmarch.load()
    names = pmarch.header_functions()
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_settings_layout_matches_header():
    txt = _lib.HEADER.read_text()
    body = re.search(r"typedef struct pm_settings \{(.*?)\} pm_settings;", txt, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"(int32_t|float)\s+(\w+)(\[\d+\])?;", body)
    assert [f[1] for f in fields] == [f[0] for f in _lib.PmSettings._fields_]
    assert ctypes.sizeof(_lib.PmSettings) == 96


def test_oracle_settings_fields_match():
    """The oracle's settings are pm_settings' first ten fields, then its
    anisotropic fields, with the same names and types."""
    import oracle as O
    ours = dict((f[0], f[1]) for f in _lib.PmSettings._fields_)
    theirs = O.Settings._fields_
    assert [f[0] for f in _lib.PmSettings._fields_][:10] == [f[0] for f in theirs[:10]]
    for name, typ in theirs:
        assert name in ours and ours[name] is typ, name


def test_no_gpu_fails_loudly():
    L = pmarch.load()
    p = ctypes.c_void_p()
    rc = L.pm_create(0, ctypes.byref(p))
    if rc == 0:
        L.pm_destroy(p)
        pytest.skip("a GPU is present on this machine")
    assert rc == -5                                  # PM_ERR_NODEV
    with pytest.raises(pmarch.PmError):
        pmarch.Marcher(0)


def test_null_context_is_rejected():
    L = pmarch.load()
    assert L.pm_set_frame(None, 0, None, 0, ctypes.c_float(0.1)) == -1
    assert L.pm_query(None) == -1
    assert L.pm_wait(None) == -1
    assert L.pm_last_error(None) == b"null context"
    xyz = np.zeros((4, 3), np.float32)
    assert L.pm_set_frame(None, 0, xyz.ctypes.data, 4, ctypes.c_float(0.1)) == -1
