"""The C-ABI libraries load and export every function their headers declare
(no GPU needed: nothing is called)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(nomad_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", [("nomad_hip.h", "libnomad_hip.so"), ("nomad_tools.h", "libnomad_tools.so")])
def test_exports(header, lib):
    names = _declared(ROOT / "include" / header)
    assert len(names) >= 8
    so = ctypes.CDLL(str(ROOT / "nomad_amd" / "lib" / lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from nomad_amd import lib

    assert set(_declared(ROOT / "include" / "nomad_hip.h")) == set(lib.HIP_SYMBOLS)
    assert set(_declared(ROOT / "include" / "nomad_tools.h")) == set(lib.TOOLS_SYMBOLS)


def test_abi_version_and_error_string():
    from nomad_amd import lib

    L = lib.hip()
    assert L.nomad_abi_version() == 1
    assert isinstance(L.nomad_last_error(), bytes)


def test_argument_errors_before_any_device_call():
    """Argument checks of the batch and autotune entry points return
    NOMAD_EINVAL with a message, before touching a device."""
    from nomad_amd import lib

    L = lib.hip()
    assert L.nomad_train_batch(None, 1, 1, 0, None) == -1
    arr = (lib._P * 1)(None)
    assert L.nomad_train_batch(arr, 0, 1, 0, None) == -1
    assert L.nomad_train_batch(arr, 257, 1, 0, None) == -1
    assert b"nomad_train_batch" in L.nomad_last_error()
    best = lib.TuneTrial()
    assert L.nomad_autotune(None, None, None, None, 0, 0.1, ctypes.byref(best), None, 0, None) == -1
    assert b"nomad_autotune" in L.nomad_last_error()
    assert L.nomad_eval_noalloc(None, None) == -1
    assert b"nomad_eval_noalloc" in L.nomad_last_error()
    assert L.nomad_get_route_state(None, None, None, None, None, None, None, None, None) == -1
    assert L.nomad_set_visit_counts(None, None, 0) == -1
