"""CPU-side checks of the drop-in boundary: libflowhip.so loads and exports
every entry point include/flowhip.h declares, and the host binding maps the
status codes to the reference's error messages.  No compute call is made
(there is no GPU here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "flowhip.h")).read()
    return sorted(set(re.findall(r"\b(fh_[a-z_0-9]+)\s*\(", text)))


def test_header_declarations_match_binding():
    from flowlab_amd import hip
    bound = sorted(n for n, _, _ in hip.SIGNATURES)
    assert declared() == bound


def test_library_exports_every_symbol():
    from flowlab_amd import hip
    lib = C.CDLL(hip.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_binding_loads_and_reports_no_device_cleanly():
    from flowlab_amd import hip
    lib = hip.lib()
    assert lib.fh_device_count() >= 0
    assert isinstance(hip.last_error(), str)


def test_status_mapping():
    from flowlab_amd import hip
    with pytest.raises(hip.HipError, match="GPU device lost"):
        hip.check(hip.FH_ERR_DEVICE_LOST)
    hip.check(hip.FH_OK)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from flowlab_amd import hip
    monkeypatch.setattr(hip, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError, match="no CPU fallback"):
        hip.load()
