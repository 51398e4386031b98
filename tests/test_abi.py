"""The C ABI: libom355.so loads (no GPU needed to load it) and exports every function that
include/om355.h declares, and the ctypes table covers exactly those names."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "om355.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(om_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for need in ("om_sweep", "om_sweep_hypothesis", "om_pm_refine", "om_box_zero_mean", "om_sample_view"):
        assert need in names


def test_library_exports_every_declared_symbol():
    from py_ortho_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libom355.so not built")
    L = _lib.lib()  # resolves every SIGNATURES entry
    names = _declared()
    assert sorted(_lib.SIGNATURES) == names
    for n in names:
        assert getattr(L, n) is not None
    assert L.om_version() >= 10000


def test_errors_map_to_reference_exceptions():
    from py_ortho_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libom355.so not built")
    with pytest.raises(ValueError):
        _lib.check(_lib.OM_ENSRC, "x")
    with pytest.raises(MemoryError):
        _lib.check(_lib.OM_ENOMEM, "x")
    with pytest.raises(RuntimeError):
        _lib.check(_lib.OM_EHIP, "x")
