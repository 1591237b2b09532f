"""Writes the golden fixtures under tests/golden/ from the reference checkout.

    python tests/golden/make_golden.py      (needs /root/reference and
                                             oracle/_ref/probe: make -C oracle ref)

Fixtures are data only:
  * cert-{n}.json      the reference's certificates (n, stored F, non-zero count,
                       weights) and ladder.csv's row for that n;
  * probe_verify.json  what the reference's own CPU verifier (tools/probe.cpp
                       --verify, compiled from source into oracle/_ref/) prints
                       for each certificate, written in the saved-vector format;
  * probe_opt.json     best F and vector of tools/probe.cpp --opt runs (the
                       reference's fp64 CPU twin of the local-search kernel) for
                       small n, with the seeds/configs that rebuild their starts.
"""
from __future__ import annotations

import csv
import json
import os
import re
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
REF = Path("/root/reference")
PROBE = REPO / "oracle" / "_ref" / "probe"


def write_vec(path: Path, a, F) -> None:
    """Saved-vector text format (src/main.cu:329-338): comments, 'n N', 'F x', heights."""
    with open(path, "w") as f:
        f.write("# AlphaEvolve problem 2 -- autocorrelation / Sidon\n")
        f.write(f"n {len(a)}\n")
        f.write(f"F {F:.17e}\n")
        for x in a:
            f.write(f"{x!r}\n")


def parse_verify(text: str) -> dict:
    def grab(pat):
        m = re.search(pat, text)
        if not m:
            raise RuntimeError(f"probe output missing {pat!r}:\n{text}")
        return m.groups()
    n, = grab(r"n\s+=\s+(\d+)")
    S, = grab(r"sum a_i\s+=\s+(\S+)")
    M, k, kmax = grab(r"max_k b_k\s+=\s+(\S+)\s+at k = (\d+) of (\d+)")
    na, = grab(r"active peaks\s+=\s+(\d+)")
    F, = grab(r"\n  F\s+=\s+(\S+)")
    rb, Fb = grab(r"round-off bnd\s+=\s+(\S+) relative\s+->\s+F <= (\S+)")
    return dict(n=int(n), S=S, M=M, kmax=int(k), nActive=int(na), F=F, relBound=rb,
                F_upper=Fb)


def main() -> int:
    if not REF.exists() or not PROBE.exists():
        print("needs /root/reference and oracle/_ref/probe (make -C oracle ref)")
        return 1
    ladder = {}
    with open(REF / "ladder.csv") as f:
        for row in csv.reader(line for line in f if not line.startswith("#")):
            ladder[int(row[0])] = dict(F=row[1], non_zero_cells=int(row[2]))

    verify_out = {}
    env = dict(os.environ, OMP_NUM_THREADS="4")
    for cert in sorted(REF.glob("certificates/*.json")):
        c = json.loads(cert.read_text())
        n = c["n"]
        out = dict(n=n, F=c["F"], non_zero_cells=c["non_zero_cells"],
                   ladder=ladder.get(n), weights=c["weights"])
        (HERE / f"cert-{n}.json").write_text(json.dumps(out))
        with tempfile.TemporaryDirectory() as td:
            vp = Path(td) / f"sidon-{n}.txt"
            write_vec(vp, c["weights"], c["F"])
            r = subprocess.run([str(PROBE), "--verify", str(vp)], capture_output=True,
                               text=True, env=env)
            rec = parse_verify(r.stdout)
            rec["exit"] = r.returncode
            verify_out[str(n)] = rec
    (HERE / "probe_verify.json").write_text(json.dumps(verify_out, indent=1))

    # --opt: restarts = 2 -> starts of kind 0 (Fourier profile) and kind 1 (iid
    # uniform); the binary reports the better one.
    runs = []
    for (n, iters, seed, mode, pA, pB, lrA, lrB) in [
            (64, 300, 7, 0, 6, 4000, 0.08, 0.002),
            (96, 200, 11, 0, 8, 8, 1e-3, 1e-3),
            (48, 250, 3, 1, 6, 400, 0.5, 0.05),
            (128, 120, 5, 0, 4, 1e4, 0.085, 0.0015)]:
        with tempfile.TemporaryDirectory() as td:
            op = Path(td) / "opt.txt"
            cmd = [str(PROBE), "--opt", "--n", str(n), "--restarts", "2", "--iters", str(iters),
                   "--seed", str(seed), "--mode", str(mode), "--pa", repr(pA), "--pb", repr(pB),
                   "--lra", repr(lrA), "--lrb", repr(lrB), "--out", str(op)]
            r = subprocess.run(cmd, capture_output=True, text=True, env=env, check=True)
            lines = [ln for ln in op.read_text().splitlines() if ln.strip()]
            runs.append(dict(n=n, iters=iters, seed=seed, mode=mode, pA=pA, pB=pB, lrA=lrA,
                             lrB=lrB, restarts=2, F=lines[1].split()[1],
                             a=[float(x) for x in lines[2:]],
                             stdout=[ln for ln in r.stdout.splitlines()
                                     if ln.strip().startswith(("best", "median", "worst"))]))
    (HERE / "probe_opt.json").write_text(json.dumps(runs))
    print("wrote", sorted(p.name for p in HERE.glob("*.json")))
    return 0


if __name__ == "__main__":
    sys.exit(main())
