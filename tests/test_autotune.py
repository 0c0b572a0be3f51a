"""Tuning-table handling on the host: nearest-N lookup, the text form read
by hh_load_tuning_file, and rejection of malformed tables (no GPU calls)."""
import ctypes as C
import json

import hermite_amd as H
from hermite_amd import autotune as A


def test_table_for_picks_nearest_tuned_size(tmp_path):
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"6/fp64/1024": [[64, 2, 1, 0], [2**31 - 1, 1, 1, 0]],
                                "6/fp64/65536": [[2**31 - 1, 1, 2, 32]],
                                "6/split/65536": [[2**31 - 1, 1, 1, 8]]}))
    assert A.table_for(6, "fp64", 60000, str(path)) == [(2**31 - 1, 1, 2, 32)]
    assert A.table_for(6, "fp64", 2000, str(path))[0] == (64, 2, 1, 0)
    assert A.table_for(6, "split", 65536, str(path)) == [(2**31 - 1, 1, 1, 8)]
    assert A.table_for(8, "fp64", 65536, str(path)) is None
    assert A.table_for(6, "fp64", 1 << 20, str(path)) is None  # beyond a factor of 4
    assert A.table_for(6, "fp64", 1024, "off") is None


def test_text_form_round_trip(tmp_path):
    rows = [(3, 2, 1, 16), (700, 1, 2, 0), (2**31 - 1, 1, 1, 4)]
    p = tmp_path / "t.txt"
    A.write_text(rows, str(p))
    lines = [ln.split() for ln in p.read_text().splitlines() if not ln.startswith("#")]
    assert [tuple(map(int, ln)) for ln in lines] == rows


def test_candidates_are_valid_geometries():
    for variant in ("fp64", "split", "exact"):
        for nact in (1, 64, 4096, 65536):
            c = A.candidates(6, variant, 65536, nact)
            assert c and all(k in (1, 2) and ni in (1, 2) and s >= 0 for k, ni, s in c)
            if variant == "exact":
                assert all(k == 1 and ni == 1 for k, ni, _ in c)
            if nact > 256:
                assert all(k == 1 for k, _, _ in c)


def test_table_entry_points_reject_bad_arguments():
    L = H.hip_lib()
    one = (C.c_int * 1)(1)
    assert L.hh_set_tuning_table(None, 1, one, one, one, one) == -1
    assert L.hh_load_tuning_file(None, b"/nonexistent") == -1
