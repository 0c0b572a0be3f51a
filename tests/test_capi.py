"""The C-ABI library loads and exports every entry point include/hermite_hip.h
declares; size helpers agree with the reference record layouts; without a
gfx950 device hh_create fails with a status code and a message (no fallback).
No compute calls are made here."""
import ctypes as C
import os
import re
import subprocess

import pytest

import hermite_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hermite_hip.h")
DRIVER_HEADER = os.path.join(ROOT, "include", "hermite_driver.h")


def declared(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hh[rd]?_[a-z_0-9]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib,minimum", [(HEADER, "hip", 30), (DRIVER_HEADER, "host", 18)])
def test_every_declared_symbol_is_exported(header, lib, minimum):
    path = H.HIP_LIB if lib == "hip" else H.HOST_LIB
    names = declared(header)
    assert len(names) >= minimum
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (hh[rd]?_[a-z_0-9]+)\b", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = H.hip_lib() if lib == "hip" else H.host_lib()
    for n in names:
        assert getattr(L, n) is not None


def test_record_sizes_match_reference_layouts():
    L = H.hip_lib()
    for order, (p, f, jp, pr) in {4: (24, 6, 14, 7), 6: (28, 9, 20, 10), 8: (36, 12, 26, 13)}.items():
        assert L.hh_particle_words(order) == p == H.PSTRIDE[order]
        assert L.hh_force_words(order) == f == H.FSTRIDE[order]
        assert L.hh_jp_words(order) == jp == H.JPSTRIDE[order]
        assert L.hh_pred_words(order) == pr
    assert L.hh_particle_words(5) < 0


def test_argument_errors_without_device():
    L = H.hip_lib()
    h = C.c_void_p()
    assert L.hh_create(5, 100, 0, 0, C.byref(h)) == -1
    assert L.hh_create(6, 0, 0, 0, C.byref(h)) == -1
    assert L.hh_create(6, 100, 7, 0, C.byref(h)) == -1
    assert b"order" in L.hh_last_error(None) or b"variant" in L.hh_last_error(None)


def test_create_fails_loudly_without_gfx950():
    L = H.hip_lib()
    h = C.c_void_p()
    rc = L.hh_create(6, 128, 0, 0, C.byref(h))
    if rc == 0:  # running on a GPU host
        L.hh_destroy(h)
        pytest.skip("a gfx950 device is present")
    assert rc == -3
    assert L.hh_last_error(None)
    with pytest.raises(H.HHError):
        H.Gravity(6, 128, "fp64")
