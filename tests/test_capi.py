"""C ABI: the library loads without a GPU, exports every symbol the public
header declares, and reports errors as codes + messages (no exceptions)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from openmoc_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "openmoc_amd.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(moc_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_exported_and_bound():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTED)


def test_version_string():
    assert b"gfx950" in _lib.lib().moc_version()


def test_null_arguments_return_codes():
    lib = _lib.lib()
    assert lib.moc_tracks_generate(None, 0, None) == _lib.MOC_OK - 1
    assert b"null" in lib.moc_last_error()
    assert lib.moc_solver_set_tracks(None, None) != 0
    assert lib.moc_solver_get_keff(None, None) != 0


def test_malformed_problem_is_rejected():
    lib = _lib.lib()
    h = C.c_void_p()
    rc = lib.moc_tracks_generate(b"moc-problem 1\nbogus 1\n", 1, C.byref(h))
    assert rc == -1 and not h.value
    assert b"unknown key" in lib.moc_last_error()
    rc = lib.moc_tracks_generate(b"not a problem\n", 1, C.byref(h))
    assert rc == -1


def test_python_layer_raises_runtime_error():
    import openmoc_amd as om
    tg = om.TrackGenerator("moc-problem 1\ngroups 1\nend\n")
    with pytest.raises(RuntimeError):
        tg.generateTracks()
