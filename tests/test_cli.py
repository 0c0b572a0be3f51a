"""Command-line driver and file formats against the reference's own driver.

The fixtures (tests/golden/cli_fixtures.json.gz, written by
tests/golden/make_cli_golden.py) hold the parameter echo of the reference's
parameter parser (SRC/parameter.cpp, its TEST_MAIN) and the stdout, snapshots
and dumps of the reference's main (SRC/compile.cpp) on its generic FP64 back
end, for 4th, 6th and 8th order, started from a snapshot and restarted from
a dump.  CPU tests: parameter parsing and the dump / snapshot writers and
readers (byte-identical re-writes).  GPU tests: hermite_run with the EXACT
kernel reproduces the reference's output lines (all but wall-clock fields)
and writes byte-identical snapshot and dump files.
"""
import gzip
import json
import os
import re
import subprocess

import numpy as np
import pytest

import hermite_amd as H

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "cli_fixtures.json.gz")
RUNNER = os.path.join(os.path.dirname(HERE), "hermite_amd", "hermite_run")


@pytest.fixture(scope="module")
def fix():
    with gzip.open(FIX, "rt") as f:
        return json.load(f)


def run_cli(inp, cwd, args=()):
    return subprocess.run([RUNNER, *args], input=inp, capture_output=True, text=True, cwd=cwd, timeout=120)


def normalise(stdout):
    """Output lines without their wall-clock fields: print_stat rows keep
    t, block/particle steps, totals, mean nact and both energy errors; the
    profile's "##" row keeps nbody and nact."""
    out = []
    for ln in stdout.splitlines():
        f = ln.split()
        if ln.startswith("## Init N square"):
            out.append("## Init N square")
        elif ln.startswith("prediction bandwidth"):
            out.append("prediction bandwidth")
        elif ln.startswith("## ") and len(f) == 6 and f[1].isdigit():
            out.append(f"## {f[1]} {f[5]}")
        elif len(f) == 10 and re.fullmatch(r"\d+\.\d+", f[0]):
            out.append(" ".join(f[:8]))
        else:
            out.append(ln)
    return out


@pytest.mark.parametrize("case", ["plain", "inverse_and_comments", "strings"])
def test_parameter_echo_matches_reference(fix, case, tmp_path):
    ref = fix["params"][case]
    r = run_cli(ref["inp"], tmp_path, ["--params"])
    assert r.returncode == 0, r.stderr
    # the reference's self-test prints an empty line before the echo
    assert r.stdout == ref["stdout"].lstrip("\n")
    for ln in ref["stderr"].splitlines():
        assert ln in r.stderr.splitlines()


def test_parameter_check_rejects(tmp_path):
    r = run_cli("order 6\neta 0.5\neta_s 0.1\nsnapin x.dat\ntend 1\ndtmax 0.1\neps 0.01\n", tmp_path)
    assert r.returncode != 0 and "power of two" in r.stderr
    r = run_cli("order 5\neta 0.5\neta_s 0.1\nsnapin x.dat\ntend 1\ndtmax 0.0625\neps 0.01\n", tmp_path)
    assert r.returncode != 0 and "order" in r.stderr
    r = run_cli("order 6\neta 0.5\neta_s 0.1\ntend 1\ndtmax 0.0625\neps 0.01\n", tmp_path)
    assert r.returncode != 0 and "snapin or dumpin" in r.stderr


@pytest.mark.parametrize("order", [4, 6, 8])
def test_dump_and_snapshot_rewrite_identically(fix, order, tmp_path):
    case = fix["cases"][str(order)]
    for name, text in case["files"].items():
        src = tmp_path / ("in_" + name)
        src.write_text(text)
        dst = tmp_path / ("out_" + name)
        if "dump" in name:
            d = H.read_dump(str(src), order)
            assert len(d["P"]) == fix["n"]
            assert np.all(d["P"][:, 2] == d["tsys"])  # tlast = tsys on restore
            H.write_dump(str(dst), order, d["dumpid"], d["tsys"], d["e0"], d["eprev"], d["nstep_tot"],
                         d["nbstep_tot"], d["P"])
        else:
            sid, t, P = H.read_snapshot(str(src), order)
            pot = np.array([float.fromhex(ln.split()[-1]) for ln in text.splitlines()[3:]])
            H.write_snapshot(str(dst), order, sid, t, P, pot)
        assert dst.read_text() == text, name


def test_reader_errors(tmp_path):
    bad = tmp_path / "bad.dat"
    bad.write_text("1\n2\n0.0\n1 0X1P+0\n")
    with pytest.raises(H.HHError):
        H.read_snapshot(str(bad), 6)
    with pytest.raises(H.HHError):
        H.read_snapshot(str(tmp_path / "missing.dat"), 6)
    dump = tmp_path / "d.dat"
    dump.write_text("6 1\n2\n0.0\n0X1P+0 0X1P+0\n1 1\n")
    with pytest.raises(H.HHError):
        H.read_dump(str(dump), 8)  # order mismatch
    with pytest.raises(H.HHError):
        H.read_dump(str(dump), 6)  # records missing


def _check_run(ref_stdout, ref_files, r, cwd):
    assert r.returncode == 0, r.stderr[-2000:]
    assert normalise(r.stdout) == normalise(ref_stdout)
    for name, text in ref_files.items():
        assert (cwd / name).read_text() == text, name


@pytest.mark.gpu
@pytest.mark.parametrize("order", [4, 6, 8])
def test_hermite_run_reproduces_reference(fix, order, tmp_path):
    case = fix["cases"][str(order)]
    (tmp_path / "snap0.dat").write_text(case["snap0"])
    print(f"hermite_run order {order} from snapshot", flush=True)
    r = run_cli(case["inp"] + "variant exact\n", tmp_path)
    _check_run(case["stdout"], case["files"], r, tmp_path)


@pytest.mark.gpu
@pytest.mark.parametrize("order", [4, 6, 8])
def test_hermite_run_restart_reproduces_reference(fix, order, tmp_path):
    case = fix["cases"][str(order)]
    (tmp_path / "dump_in.dat").write_text(case["files"]["out_dump001.dat"])
    print(f"hermite_run order {order} from dump", flush=True)
    r = run_cli(case["inp_restart"] + "variant exact\n", tmp_path)
    _check_run(case["stdout_restart"], case["files_restart"], r, tmp_path)


@pytest.mark.gpu
def test_hermite_run_fp64_close_to_reference(fix, tmp_path):
    """The default FP64 kernel: same block-step counts are not guaranteed
    (timestep quantisation), energies within a loose bound."""
    case = fix["cases"]["6"]
    (tmp_path / "snap0.dat").write_text(case["snap0"])
    r = run_cli(case["inp"], tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    rows = [ln.split() for ln in r.stdout.splitlines() if re.match(r"^\d+\.\d+ ", ln)]
    ref = [ln.split() for ln in case["stdout"].splitlines() if re.match(r"^\d+\.\d+ ", ln)]
    assert [x[0] for x in rows] == [x[0] for x in ref]
    for a, b in zip(rows, ref):
        assert abs(float(a[7]) - float(b[7])) < 1e-7


@pytest.mark.parametrize("order,eta,eta_s", [(4, "0.1", "0.01"), (6, "0.5", "0.1"), (8, "0.75", "0.1")])
def test_runbench_parameters_are_the_reference_run_files(order, eta, eta_s, tmp_path):
    """hermite_amd.runbench feeds hermite_run the parameters of
    RUN/inp.{4th,6th,8th} (eta, eta_s; tend 2, dt_max 1/16, kt_for_eps 1,
    log / snapshot / dump intervals 0.25 / 2 / 1, snapin inp.dat)."""
    from hermite_amd import runbench

    r = run_cli(runbench.params(order, "fp64"), tmp_path, ["--params"])
    assert r.returncode == 0, r.stderr
    echo = dict(ln.split(" : ", 1) for ln in (r.stdout + r.stderr).splitlines() if " : " in ln)
    want = {"order": str(order), "eta": eta, "eta_s": eta_s, "tend": "2", "dtmax": "0.0625", "kt_for_eps": "1",
            "log_interval": "0.25", "snap_interval": "2", "dump_interval": "1", "snapin": "inp.dat",
            "variant": "fp64"}
    for k, v in want.items():
        assert echo.get(k) == v, (k, echo.get(k))


def _stat_rows(stdout):
    return [ln.split() for ln in stdout.splitlines() if re.match(r"^\d+\.\d+ ", ln)]


@pytest.mark.gpu
@pytest.mark.parametrize("order", [4, 6, 8])
def test_hermite_run_resident_mode(fix, order, tmp_path):
    """resident 1: block steps GPU-resident between output times.  The
    particles are not permuted between block steps, so the j summation order,
    and with it the last bits, differ from the reference's run: the same
    output times and files, the reference's energy errors to 1e-7, every
    particle in every snapshot and dump (the dumps' records all at the
    output time); a restart
    from the resident run's dump runs resident to the end."""
    case = fix["cases"][str(order)]
    (tmp_path / "snap0.dat").write_text(case["snap0"])
    print(f"hermite_run order {order} resident", flush=True)
    r = run_cli(case["inp"] + "variant exact\nresident 1\n", tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    rows, ref = _stat_rows(r.stdout), _stat_rows(case["stdout"])
    assert [x[0] for x in rows] == [x[0] for x in ref]
    for a, b in zip(rows, ref):
        assert abs(float(a[7]) - float(b[7])) < 1e-7
        assert abs(float(a[6]) - float(b[6])) < 1e-7
    _, _, P0 = H.read_snapshot(str(tmp_path / "snap0.dat"), order)
    for name in case["files"]:
        path = str(tmp_path / name)
        if "snap" in name:
            _, t, P = H.read_snapshot(path, order)
        else:
            d = H.read_dump(path, order)
            t, P = d["tsys"], d["P"]
            assert np.all(P[:, 2] == t), name  # every tlast at the output time
        assert sorted(P[:, 0]) == sorted(P0[:, 0]), name
        assert np.all(np.isfinite(P)), name
    (tmp_path / "dump_in.dat").write_text((tmp_path / "out_dump001.dat").read_text())
    r2 = run_cli(case["inp_restart"] + "variant exact\nresident 1\n", tmp_path)
    assert r2.returncode == 0, r2.stderr[-2000:]
    rows2, ref2 = _stat_rows(r2.stdout), _stat_rows(case["stdout_restart"])
    assert [x[0] for x in rows2] == [x[0] for x in ref2]
