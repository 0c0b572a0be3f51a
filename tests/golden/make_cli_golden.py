"""Golden outputs of the reference's own command-line driver.

Runs oracle/_ref/ref_main{4,6,8} (SRC/compile.cpp + SRC/parameter.cpp on the
reference's generic FP64 back end, built from the reference sources by
oracle/Makefile) on small Plummer snapshots written in the reference's masaki
format, once from the snapshot and once restarted from the dump it wrote,
and oracle/_ref/ref_param (parameter.cpp's self-test) on a few parameter
files.  Stores parameter files, input snapshots, stdout and every file the
runs wrote in tests/golden/cli_fixtures.json.gz.

    python tests/golden/make_cli_golden.py        (needs oracle/_ref built)
"""
import gzip
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import hermite_amd as H  # noqa: E402
from oracle import REF_DIR  # noqa: E402

N = 48
SEED = 11
# eta / eta_s of RUN/inp.{4th,6th,8th}; the 6th-order case takes its
# softening from kt_for_eps (set_thermal_eps), the others from epsinv
ORDERS = {4: ("0.1", "0.01", "epsinv 64"), 6: ("0.5", "0.1", "kt_for_eps 1.0"), 8: ("0.75", "0.1", "epsinv 64")}

PARAM_FILES = {
    "plain": "order 6\neta 0.5\neta_s 0.1\nsnapin inp.dat\ntend 2.0\ndtmaxinv 16\nkt_for_eps 1.0\n"
             "log_interval 0.25\nsnap_interval 2.0\ndump_interval 1.0\n",
    "inverse_and_comments": "# a comment\norder 8\n  eta 0.75\neta_s 0.1\n1bad 3\nbogus_tag 5\ndumpin d.dat\n"
                            "dumpid 3\nsnapid 2\ntend 1e-3\ndtmax 0.03125\nepsinv 1024\n",
    "strings": "order 4\nsnapin a.dat\nsnapout_base snap%04d.dat\ndumpout_base dump%03d.dat\ntend 1\n",
}


def params(order, start, extra=""):
    eta, eta_s, soft = ORDERS[order]
    return (f"order {order}\neta {eta}\neta_s {eta_s}\n{start}\nsnapout_base out_snap%03d.dat\n"
            f"dumpout_base out_dump%03d.dat\ntend 0.25\ndtmaxinv 16\n{soft}\nlog_interval 0.0625\n"
            f"snap_interval 0.125\ndump_interval 0.125\n{extra}")


def run(exe, inp, cwd):
    r = subprocess.run([exe], input=inp, capture_output=True, text=True, cwd=cwd, check=True)
    return r.stdout


def main():
    out = {"n": N, "seed": SEED, "cases": {}, "params": {}}
    for order in ORDERS:
        m, x, v = H.plummer(N, SEED)
        P = H.make_particles(order, m, x, v)
        with tempfile.TemporaryDirectory() as td:
            H.write_snapshot(os.path.join(td, "snap0.dat"), order, 0, 0.0, P, np.zeros(N))
            snap0 = open(os.path.join(td, "snap0.dat")).read()
            inp = params(order, "snapin snap0.dat")
            stdout = run(os.path.join(REF_DIR, f"ref_main{order}"), inp, td)
            files = {f: open(os.path.join(td, f)).read() for f in sorted(os.listdir(td)) if f.startswith("out_")}
        with tempfile.TemporaryDirectory() as td:
            with open(os.path.join(td, "dump_in.dat"), "w") as f:
                f.write(files["out_dump001.dat"])
            inp_r = params(order, "dumpin dump_in.dat")
            stdout_r = run(os.path.join(REF_DIR, f"ref_main{order}"), inp_r, td)
            files_r = {f: open(os.path.join(td, f)).read() for f in sorted(os.listdir(td)) if f.startswith("out_")}
        out["cases"][str(order)] = {"inp": inp, "snap0": snap0, "stdout": stdout, "files": files,
                                    "inp_restart": inp_r, "stdout_restart": stdout_r, "files_restart": files_r}
    for name, text in PARAM_FILES.items():
        r = subprocess.run([os.path.join(REF_DIR, "ref_param")], input=text, capture_output=True, text=True,
                           check=True)
        out["params"][name] = {"inp": text, "stdout": r.stdout, "stderr": r.stderr}
    path = os.path.join(HERE, "cli_fixtures.json.gz")
    with gzip.open(path, "wt") as f:
        json.dump(out, f)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
