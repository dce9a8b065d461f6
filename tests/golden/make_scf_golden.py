"""Write SCF-level golden fixtures tests/golden/scf_<case>.json from the reference.

Runs the reference program (oracle/_ref/x2dhf_omp, built from the reference's
own sources by oracle/ref.mk) on the decks of tests/golden/decks -- all but
one relax orbitals and potentials with MCSOR in the col-wise order, the
algorithm of the pipelined GPU sweep; altsweeps_fh_91x121 uses the
reference's default SOR in the middle ordering with its altsweeps option
(backward sweeps for odd orbitals), and ne2_set03_181x151 and gaf_set01 are
the reference's own Ne2 and GaF decks (test-sets/ne2/set-03, gaf/set-01) --
and keeps, from its listing:
  * the orbital energy of every orbital after every SCF iteration,
  * the final total energy and its components, orbital energies and the
    number of (MC)SOR sweeps.
A listing produced earlier can be given instead of running the deck
(the configuration-size decks take minutes to tens of minutes); with
arguments only the named cases are written:

    python tests/golden/make_scf_golden.py [case=path/to/listing ...]

The configuration-3 deck runs 3000 iterations (about 15 minutes on 6
threads here) with the saturation test off, so that its stopping point does
not depend on rounding noise.
"""
import json
import os
import re
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "x2dhf_omp")

CASES = {
    "config1_h2_91x121": "config1_h2_91x121.data",
    "n2_91x121": "n2_91x121.data",
    "fh_91x121": "fh_91x121.data",
    "config2_n2_331x601": "config2_n2_331x601.data",
    "config3_fh_481x901": "config3_fh_481x901.data",
    "altsweeps_fh_91x121": "altsweeps_fh_91x121.data",
    "ne2_set03_181x151": "ref_ne2_set-03.data",
    "gaf_set01": "ref_gaf_set-01.data",
}

NUM = r"(-?\d\.\d+E[+-]\d+)"
ITER = re.compile(r"^\s*(\d+)\s+(\d+)\s+(sigma|pi|delta)\s*([gu]?)\s+" + NUM)
FINAL_ORB = re.compile(r"^\s+(\d+)\s+(sigma|pi|delta)\s+([gu]?)\s*" + NUM + r"\s+(-?\d\.\d+E[+-]\d+)\s*$")


def parse(text):
    iters = {}
    for line in text.splitlines():
        m = ITER.match(line)
        if m:
            it = int(m.group(1))
            iters.setdefault(it, []).append(float(m.group(5)))
    out = {"iterations": [iters[k] for k in sorted(iters)]}
    m = re.findall(r"total energy:\s+" + NUM, text)
    out["etot"] = float(m[-1])
    for key, pat in (("enuc", "nuclear attraction energy"), ("ekin", "kinetic energy"),
                     ("ecoul", "Coulomb energy"), ("eexch", "exchange energy")):
        mm = re.search(r"^\s+" + pat + r":\s+(-?\d+\.\d+)\s*$", text, re.M)
        out[key] = float(mm.group(1))
    out["sweeps"] = int(re.search(r"\(MC\)SOR iterations:\s+(\d+)", text).group(1))
    tail = text[text.rindex("orbital                 energy"):]
    eps = []
    for line in tail.splitlines()[1:]:
        m = FINAL_ORB.match(line)
        if not m:
            break
        eps.append(float(m.group(4)))
    out["eps"] = eps
    for reason in ("orbital energy threshold reached", "orbital normalization threshold reached",
                   "solution cannot be further improved", "scf iteration limit reached"):
        if reason in text:
            out["stop"] = reason
    return out


def run_deck(deck):
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(HERE, "decks", deck)) as fin:
            t0 = time.time()
            r = subprocess.run([REF], stdin=fin, cwd=d, env=dict(os.environ, OMP_NUM_THREADS="4"),
                               stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, check=True)
            wall = time.time() - t0
    return r.stdout, wall


if __name__ == "__main__":
    given = dict(a.split("=", 1) for a in sys.argv[1:])
    for name, deck in CASES.items():
        if given and name not in given:
            continue
        if name in given:
            text, wall = open(given[name]).read(), None
        else:
            text, wall = run_deck(deck)
        out = parse(text)
        out["deck"] = deck
        out["wall_s_here"] = wall
        with open(os.path.join(HERE, f"scf_{name}.json"), "w") as f:
            json.dump(out, f, indent=0)
        print(name, out["etot"], len(out["iterations"]), "iterations", out["sweeps"], "sweeps", out.get("stop"))
