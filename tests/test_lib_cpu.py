"""The C-ABI library without a GPU: it is built, loads, exports every symbol
include/opf.h declares, and its host-side mesh precompute matches the oracle
(no device compute is called here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import REPO

import opf_oracle as O

LIB = os.path.join(REPO, "softx-d-17-00015_amd", "opf", "libopf.so")
HEADER = os.path.join(REPO, "include", "opf.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(opf_[a-z0-9_]+)\s*\(", text)))


def test_library_built_and_exports_header_symbols():
    assert os.path.exists(LIB), "libopf.so not built (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from opf import _lib
    assert set(declared_symbols()) <= set(_lib.PROTOTYPES), \
        set(declared_symbols()) - set(_lib.PROTOTYPES)


@pytest.mark.parametrize("N", [16, 32, 48, 64, 96])
def test_host_mesh_matches_oracle(N):
    from opf import host_mesh
    me = host_mesh(N)
    ref = O.build_mesh(N)
    assert np.array_equal(me["r"], ref.r)
    tol = 1e-13
    assert np.max(np.abs(me["intrdr"] - ref.intrdr)) <= tol * np.abs(ref.intrdr).max()
    assert np.max(np.abs(me["dr0"] - ref.dr0)) <= tol * np.abs(ref.dr0).max()
    assert np.max(np.abs(me["dr1"] - ref.dr1)) <= tol * np.abs(ref.dr1).max()
    assert np.max(np.abs(me["dr1band"] - ref.dr[1])) <= tol * np.abs(ref.dr[1]).max()
    assert np.max(np.abs(me["radlap"] - ref.radLap)) <= tol * np.abs(ref.radLap).max()


def test_host_mesh_rejects_bad_sizes():
    from opf import host_mesh
    from opf._lib import OpfError
    with pytest.raises(OpfError):
        host_mesh(4)
