"""The reference's own benchmark run (RUN/run.sh: the command line on the
``pl{N/1024}k`` Plummer inputs with RUN/inp.{4th,6th,8th}) through
``hermite_run`` on this back end, printing its ``##`` lines: the "Init N
square" rate and, per N-body time unit, wall time, Gflops (num_step * N *
60/97/144 over the wall time, counting everything, SRC/nbodysystem.h:147),
microseconds per block step and the mean active count.

    python -m hermite_amd.runbench --order 6 --n 65536 [--variant fp64] [--resident]

--resident runs the block steps GPU-resident between output times (the
"resident 1" tag); the particles then keep their order, so the j summation
order and the last bits of the run differ from the reference's.

The input is generated as RUN/gen-plum.c does (seed 19640916, radii below
100), and the parameters are those of RUN/inp.{4th,6th,8th}; with the host
driver the mean active counts therefore match the reference's logs
(RUN/gpu-*.log) line for line.
"""
import argparse
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

from . import make_particles, plummer, write_snapshot

# RUN/inp.4th, inp.6th, inp.8th: eta, eta_s (the other tags are the same)
_ETA = {4: (0.1, 0.01), 6: (0.5, 0.1), 8: (0.75, 0.1)}


def params(order: int, variant: str, resident: bool = False) -> str:
    eta, eta_s = _ETA[order]
    return (f"order {order}\neta   {eta}\neta_s {eta_s}\n\nsnapin     inp.dat\n\n"
            "tend       2.0\ndtmaxinv   16\nkt_for_eps 1.0\n\n"
            "log_interval  0.25\nsnap_interval 2.0\ndump_interval 1.0\n"
            f"variant {variant}\n" + ("resident 1\n" if resident else ""))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--order", type=int, default=6, choices=[4, 6, 8])
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--variant", default="fp64", choices=["fp64", "split", "exact"])
    ap.add_argument("--resident", action="store_true", help="GPU-resident block steps (tag resident 1)")
    args = ap.parse_args(argv)
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hermite_run")
    if not os.path.exists(exe):
        raise SystemExit("hermite_run is not built (run __graft_entry__.build())")
    m, x, v = plummer(args.n, 19640916, 100.0)
    P = make_particles(args.order, m, x, v)
    with tempfile.TemporaryDirectory() as wd:
        write_snapshot(os.path.join(wd, "inp.dat"), args.order, 0, 0.0, P, np.zeros(args.n))
        t0 = time.time()
        out = subprocess.run([exe], input=params(args.order, args.variant, args.resident), text=True, capture_output=True, cwd=wd)
    if out.returncode:
        sys.stderr.write(out.stderr)
        return out.returncode
    for ln in out.stdout.splitlines():
        if ln.startswith("##"):
            print(ln)
    print(f"# {time.time() - t0:.1f} s wall for t = 0 -> 2", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
