"""The C ABI: the library loads and exports every entry point the public
header declares, and the Python binding covers them all. No GPU calls."""
import os
import re
import subprocess

import pytest

import robtop_amd
from robtop_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "robtop_amd.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "rt_create_from_voxels" in names and "rt_solve" in names and len(names) > 40


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "run python -m robtop_amd.build"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib.SIGNATURES) == declared()


def test_library_loads_and_reports_version():
    L = _lib.lib()
    assert b"gfx950" in L.rt_version()


def test_errors_are_reported_not_silent():
    L = _lib.lib()
    assert L.rt_num_levels(None) == _lib.RT_EARG
    assert b"null handle" in L.rt_last_error()
