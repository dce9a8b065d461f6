"""Golden vectors for the deck front end's grid rules: for every test set of
the reference (test-sets/*/set-*), the grid and nuclei cards of its
input.data and the grid its reference.lst reports (nu and mu point counts).
Reads the reference's files as text; writes tests/golden/deck_grids.json.

usage: python tests/golden/make_deck_grids.py [/root/reference]"""
import glob
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
out = []
for d in sorted(glob.glob(os.path.join(ref, "test-sets", "*", "set-*"))):
    inp, lst = os.path.join(d, "input.data"), os.path.join(d, "reference.lst")
    if not (os.path.exists(inp) and os.path.exists(lst)):
        continue
    grid = nuclei = None
    for line in open(inp, errors="replace"):
        t = line.split("!")[0].split()
        if not t:
            continue
        if t[0].lower() == "grid" and grid is None:
            grid = t[1:]
        if t[0].lower() == "nuclei" and nuclei is None:
            nuclei = t[1:]
    text = open(lst, errors="replace").read()
    m_nu = re.search(r"nu \(h_nu\)\s*=\s*(\d+)", text)
    m_mu = re.search(r"mu \(h_mu\)\s*=\s*(\d+)", text)
    if not (grid and nuclei and m_nu and m_mu):
        continue
    out.append({"set": os.path.relpath(d, os.path.join(ref, "test-sets")), "grid": grid, "nuclei": nuclei,
                "nu": int(m_nu.group(1)), "mu": int(m_mu.group(1))})
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "deck_grids.json")
with open(path, "w") as f:
    json.dump(out, f, indent=0)
print(f"{len(out)} test sets -> {path}")
