"""Write the golden fixtures tests/golden/<case>.npz from the reference.

Runs oracle/_ref/ref_harness (the reference x2dhf routines compiled from
their own sources by oracle/ref.mk, driven by oracle/ref_harness.f90) on the
decks in tests/golden/decks/ and keeps the slices of its dumps that the tests
check.  Needs the reference sources, so it only runs where they are mounted;
the fixtures it writes are plain arrays (numpy .npz, no pickles) and travel
with the repository.

    make -f oracle/ref.mk harness
    python tests/golden/make_golden.py [case ...]     (h2_61x91_orderings: every other ordering)
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
MAXORB = 36  # MAXORB of the reference params.h (ee is dimensioned maxorb x maxorb)

# case name -> (deck, traced orbital or None for the reference's first one)
CASES = {
    "h2_61x91": ("h2_61x91.data", None),
    "n2_61x91": ("n2_61x91.data", 1),
}
# the same decks in the reference's default (middle) ordering: only the
# relaxation vectors (inputs and the sor / mcsor outputs) are kept
LIGHT = {
    "h2_61x91_middle": ("h2_61x91_middle.data", None),
    "n2_61x91_middle": ("n2_61x91_middle.data", 1),
}
LIGHT_KEYS = ("sizes", "scal", "coef", "suppl", "vxi", "veta")
# the H2 61x91 deck in each of the reference's other mesh orderings (its
# `order` card; one MCSOR thread): the relaxation vectors of each go into
# one file, h2_61x91_orderings.npz, under "<ordering>/<key>"
ORDER_BASE = "h2_61x91_middle.data"
ORDER_NAMES = ("middlemt", "row-wise", "rrow-wis", "test1", "test2", "test3", "test4", "middle2")
ORDER_KEYS = ("k1_win", "k1_lhs", "k1_rhs", "k1_wout_mcsor1", "k1_wout_sor", "k1_wout_sor_bwd", "k1_isym",
              "k1_scal")


def run_case(name, deck, iorb, light=False, deck_text=None):
    with tempfile.TemporaryDirectory() as d:
        env = dict(os.environ, X2_DUMP=d, OMP_NUM_THREADS="2")
        if iorb is not None:
            env["X2_IORB"] = str(iorb)
        if deck_text is None:
            with open(os.path.join(HERE, "decks", deck)) as f:
                deck_text = f.read()
        subprocess.run([HARNESS], input=deck_text, text=True, cwd=d, env=env, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)

        def f64(n):
            return np.fromfile(os.path.join(d, n + ".f64"), dtype="<f8")

        def i32(n):
            return np.fromfile(os.path.join(d, n + ".i32"), dtype="<i4")

        nni, nmu, norb, _ = (int(v) for v in i32("sizes"))
        n = nni * nmu
        npad = (nni + 8) * (nmu + 8)
        io = int(i32("nexchpots")[0])
        pl = i32("potlist").reshape(-1, 5)[:io]          # in1, in2, dm, isym, ibexcp
        traced = int(pl[0, 0]) if iorb is None else iorb
        if iorb is None:
            traced = norb
        out = {
            "sizes": np.array([nni, nmu, norb, traced], dtype=np.int32),
            "scal": f64("scal"),
            "coef": np.concatenate([f64("dmu1"), f64("dmu2"), f64("dni1"), f64("dni2"), f64("exeven")]),
            "vmu": f64("vmu"), "vxi": f64("vxi"), "vni": f64("vni"), "veta": f64("veta"),
            "suppl": f64("suppl").reshape(14, n),
            "potlist": pl.astype(np.int32),
        }
        for pfx in ("k1_", "k2_"):
            if not os.path.exists(os.path.join(d, pfx + "win.f64")):
                continue
            for k in ("win", "wout_mcsor", "wout_mcsor1", "wout_sor", "wout_sor_bwd"):
                out[pfx + k] = f64(pfx + k)
            for k in ("lhs", "rhs"):
                out[pfx + k] = f64(pfx + k)
            out[pfx + "isym"] = i32(pfx + "isym")
            out[pfx + "scal"] = f64(pfx + "scal")

        if light == "vectors":
            return {k: out[k] for k in ORDER_KEYS}
        if light:
            keep = {k: v for k, v in out.items() if k in LIGHT_KEYS or k.startswith("k")}
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **keep)
            print(name, "->", os.path.getsize(os.path.join(HERE, name + ".npz")), "bytes")
            return

        def slices(arr, offs):
            return np.stack([arr[o - 1:o - 1 + n] for o in offs])

        orb = lambda a: a[: norb * n].reshape(norb, n)
        out["orb_init"] = orb(f64("orb_init"))
        out["orb_prep"] = orb(f64("orb_prep"))
        exch_prep = f64("exch_prep")
        # Coulomb potentials (first norb slots) + the potentials relaxed for the traced orbital
        out["coul_prep"] = exch_prep[: norb * n].reshape(norb, n)
        offs = [int(o) for o in pl[:, 4]]
        out["pots_prep"] = slices(exch_prep, offs)
        out["pots_asympt"] = slices(f64("exch_asympt"), offs)
        out["pots_relaxed"] = slices(f64("exch_relaxed"), offs)
        out["ee_prep"] = f64("ee_prep").reshape(MAXORB, MAXORB).T[:norb, :norb].copy()
        out["etot_prep"] = f64("etot_prep")
        out["norm_prep"] = f64("norm_prep")[:norb]
        out["fock1"] = f64("fock1")
        out["fock2"] = f64("fock2")
        out["coulombptr"] = orb(f64("coulombptr"))[traced - 1]
        out["exchangeptr"] = orb(f64("exchangeptr"))[traced - 1]
        out["orb_relaxed"] = orb(f64("orb_relaxed"))[traced - 1]
        out["orb_norm"] = orb(f64("orb_norm"))[traced - 1]
        out["orb_ortho"] = orb(f64("orb_ortho"))[traced - 1]
        out["norm_1"] = f64("norm_1")[:norb]
        out["ee_1"] = f64("ee_1").reshape(MAXORB, MAXORB).T[:norb, :norb].copy()
        out["oneel_1"] = f64("oneel_1")[:1]
        out["etot_1"] = f64("etot_1")
        # cmulti(iorb + (k-1)*norb), k = 1..mpole (scfUtils.F90:531-589)
        out["cmulti_prep"] = f64("cmulti_prep")[: 4 * norb].reshape(4, norb)
        out["cmulti_1"] = f64("cmulti_1")[: 4 * norb].reshape(4, norb)
        for k in ("excdi", "excqu", "excoc", "exche"):
            out[k + "_1"] = f64(k + "_1")[: norb * (norb + 1)]
            out[k + "_prep"] = f64(k + "_prep")[: norb * (norb + 1)]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, "->", os.path.getsize(os.path.join(HERE, name + ".npz")), "bytes")


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit("build the harness first: make -f oracle/ref.mk harness")
    only = set(sys.argv[1:])
    for name, (deck, iorb) in CASES.items():
        if not only or name in only:
            run_case(name, deck, iorb)
    if not only or "h2_61x91_orderings" in only:
        with open(os.path.join(HERE, "decks", ORDER_BASE)) as f:
            base = f.read()
        assert "order middle" in base
        vec = {}
        for o in ORDER_NAMES:
            got = run_case(o, None, None, light="vectors", deck_text=base.replace("order middle", f"order {o}"))
            vec.update({f"{o}/{k}": v for k, v in got.items()})
        np.savez_compressed(os.path.join(HERE, "h2_61x91_orderings.npz"), **vec)
        print("h2_61x91_orderings ->", os.path.getsize(os.path.join(HERE, "h2_61x91_orderings.npz")), "bytes")
    for name, (deck, iorb) in LIGHT.items():
        if not only or name in only:
            run_case(name, deck, iorb, light=True)
