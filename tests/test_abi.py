"""The C ABI library and the CLI, without a GPU: symbols and host-side error paths."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR, ROOT


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "aknn.h")).read()
    return sorted(set(re.findall(r"\b(aknn_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(aknn):
    L = aknn.load()
    declared = _header_symbols()
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(aknn.EXPORTS) == declared


def test_library_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG_DIR, "libaknn.so")],
                         capture_output=True, text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for name in _header_symbols():
        assert name in names


def test_product_does_not_link_the_oracle():
    out = subprocess.run(["ldd", os.path.join(PKG_DIR, "libaknn.so")], capture_output=True, text=True).stdout
    assert "oracle" not in out
    out = subprocess.run(["nm", "-D", os.path.join(PKG_DIR, "libaknn.so")], capture_output=True, text=True).stdout
    assert "orc_" not in out


def test_handle_errors_without_device(aknn):
    L = aknn.load()
    h = L.aknn_create(0)
    try:
        assert L.aknn_calculate(h) != 0
        assert b"not initialized" in L.aknn_last_error(h)
        assert L.aknn_set_option(h, b"no_such_option", 1) != 0
        assert L.aknn_set_option(h, b"rng_order", 3) != 0
        assert L.aknn_set_option(h, b"seed", 7) == 0
        assert L.aknn_set_option(h, b"split", 2) != 0
        assert L.aknn_set_option(h, b"split", 1) == 0
        assert L.aknn_set_option(h, b"router_trees", -1) != 0
        assert L.aknn_set_bins(h, np.zeros(4, np.int32), 4, 2) != 0  # no data yet
        assert L.aknn_set_router(h, None) != 0
        assert L.aknn_save_result(h) != 0
    finally:
        L.aknn_destroy(h)


def test_router_rejects_bad_shapes_and_missing_device(aknn):
    L = aknn.load()
    assert not L.aknn_router_create(0, 16, 30, 8, 1)   # hidden not a multiple of 32
    assert not L.aknn_router_create(0, 16, 64, 4, 1)   # too few bins for the matrix-core tiles
    assert not L.aknn_router_create(97, 16, 64, 8, 1)  # no such device
    assert L.aknn_router_stats(None, np.zeros(7), 7) != 0


def _cli(*args):
    return subprocess.run([os.path.join(PKG_DIR, "aknn"), *args], capture_output=True, text=True, timeout=60)


def test_cli_wrong_argument_count():
    r = _cli("a", "b")
    assert r.returncode == 1
    assert "Wrong number of arguments!" in r.stdout


def test_cli_wrong_distance():
    r = _cli("in.csv", "out", "2", "10", "10", "10", "MANHATTAN", "0", "T")
    assert r.returncode == 2
    assert "Wrong distance algorithm!" in r.stdout


def test_cli_missing_device():
    r = _cli("in.csv", "out", "2", "10", "10", "10", "EUCLIDEAN", "97", "T")
    assert r.returncode == 3
    assert "Device with id 97 is not located in this computer" in r.stdout
