"""Input-deck front end (x2dhf_amd.deck) against the reference's own decks and
listings: the grid rules reproduce the grid every reference test set reports
(tests/golden/deck_grids.json, from tests/golden/make_deck_grids.py), the
configuration decks give the job specifications the benchmark and the SCF
parity tests use, the reference's LiH and BH decks parse into their
molecules, and cards outside the path are refused."""
import json
import os

import pytest

from conftest import GOLDEN

DECKS = os.path.join(GOLDEN, "decks")


def _fields(spec):
    out = {}
    for name, _ in type(spec)._fields_:
        v = getattr(spec, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


def test_grid_rules_reproduce_reference_listings():
    from x2dhf_amd import deck
    vecs = json.load(open(os.path.join(GOLDEN, "deck_grids.json")))
    assert len(vecs) > 150
    for e in vecs:
        txt = (f"nuclei {' '.join(e['nuclei'])}\nconfig 0\n 1 sigma end\ngrid {' '.join(e['grid'])}\n"
               "orbpot hydrogen\nlcao\n 1.0 1 0 1.0 1.0 1 0 1.0\nstop\n")
        d = deck.parse(txt)
        assert (d.nni, d.nmu) == (e["nu"], e["mu"]), e["set"]


def test_mesh_rounding():
    from x2dhf_amd import deck
    # the reference-admissible sizes of the benchmark grids
    assert [deck.nnu_calc(n) for n in (91, 121, 169, 337, 505, 601, 901)] == [91, 121, 151, 331, 481, 601, 901]
    with pytest.raises(deck.DeckError):
        deck.nnu_calc(30)


@pytest.mark.parametrize("name,builder,kw", [
    ("config1_h2_91x121", "h2_spec", dict(nni=91, nmu=121, rinf=40.0, omega=(1.95, 1.95), maxscf=3000)),
    ("config2_n2_331x601", "n2_spec", dict(nni=331, nmu=601, rinf=40.0, maxscf=8000)),
    # configuration 3 runs a fixed 3000 iterations with the saturation test off
    # and unreachable thresholds, so that both programs stop at the same place
    ("config3_fh_481x901", "fh_spec", dict(nni=481, nmu=901, rinf=50.0, maxscf=3000, ienterm=16, inoterm=16,
                                           conv=(8000, 0, 0))),
])
def test_configuration_decks_give_the_benchmark_specs(name, builder, kw):
    from x2dhf_amd import deck, scf
    d = deck.load(os.path.join(DECKS, name + ".data"))
    kw = dict(dict(ienterm=12, inoterm=12), **kw)
    want = getattr(scf, builder)(**kw)
    assert _fields(d.spec()) == _fields(want)
    # these decks ask for MCSOR col-wise for orbitals and potentials
    assert (d.orb_method, d.pot_method, d.ordering) == (0, 0, 0)
    assert d.notes[-1] == "orbitals by MCSOR, potentials by MCSOR, col-wise ordering"


def test_reference_lih_deck():
    from x2dhf_amd import deck, scf
    d = deck.load(os.path.join(DECKS, "ref_lih_set-01.data"))
    # test-sets/lih/set-01/reference.lst: nu = 241, mu = 331, R_infty = 60
    assert (d.nni, d.nmu, d.rinf) == (241, 331, 60.0)
    assert (d.maxscf, d.ienterm, d.inoterm) == (10000, 12, 14)
    assert d.conv == (2000, 0, 0) and d.omega == (1.85, 1.85)
    # no mcsor / order cards: the reference's default, SOR in the middle ordering
    want = scf.lih_spec(nni=241, nmu=331, rinf=60.0, maxscf=10000, ienterm=12, inoterm=14, omega=(1.85, 1.85),
                        conv=(2000, 0, 0), orb_method=1, pot_method=1, ordering=1)
    assert _fields(d.spec()) == _fields(want)
    assert d.notes[-1] == "orbitals by SOR, potentials by SOR, middle ordering"


def test_reference_bh_deck():
    from x2dhf_amd import deck, scf
    d = deck.load(os.path.join(DECKS, "ref_bh-elprop_set-01.data"))
    # test-sets/bh-elprop/set-01/reference.lst: nu = 331, mu = 601, R_infty = 200
    assert (d.nni, d.nmu, d.rinf) == (331, 601, 200.0)
    assert (d.maxscf, d.ienterm, d.inoterm, d.ienterm4init) == (4000, 14, 16, 0)
    want = scf.bh_spec(nni=331, nmu=601, rinf=200.0, maxscf=4000, ienterm=14, inoterm=16, ienterm4init=0,
                       conv=(4000, 0, 0), orb_method=d.orb_method, pot_method=d.pot_method, ordering=d.ordering)
    assert _fields(d.spec()) == _fields(want)


def test_units_and_lcao_screening():
    from x2dhf_amd import deck
    d = deck.parse("method hf\nnuclei 3.0 1.0 1.5954 angstroms\nconfig 0\n 2 sigma end\ngrid 151 40.0\n"
                   "orbpot hydrogen\nlcao 2\n 1.0 2 0 1.0 1.0 1 0 0.0\n 1.0 1 0 0.0 0.0 1 0 0.0\n"
                   "scf 100 10 9 8\nomega 1.9\nmultipol 2.0 6\nsor4pot 12 2\nstop\n")
    assert abs(d.r - 1.5954 / deck.BOHR2ANG) < 1e-14
    assert d.lcao[0] == (1.0, 2, 0, 2.0, 1.0, 1, 0, 1.0)  # exponents Z - value
    assert (d.maxscf, d.ienterm, d.inoterm) == (100, 9, 8)
    assert d.omega == (1.9, -1.0) and (d.recalc_mm, d.mpole) == (2.0, 6)
    assert d.maxsorpot == (2, 12) and d.maxsororb == (1, 10)


@pytest.mark.parametrize("bad,why", [
    ("method dft\n", "only HF"),
    ("orbpot hf\n", "hydrogenic"),
    ("fefield 1e-4\n", "outside the relaxation path"),
    ("config 0\n 1 pi + - . end\n", "open-shell"),
    ("config 0\n 1 sigmaxx end\n", "wrong symmetry"),
])
def test_cards_outside_the_path_are_refused(bad, why):
    from x2dhf_amd import deck
    base = "nuclei 1.0 1.0 1.4\nconfig 0\n 1 sigma g end\ngrid 91 121 40.0\norbpot hydrogen\nlcao\n" \
           " 1.0 1 0 1.0 1.0 1 0 1.0\n"
    with pytest.raises(deck.DeckError, match=why):
        deck.parse(bad + base if not bad.startswith("config") else bad + base.replace(
            "config 0\n 1 sigma g end\n", "").replace("lcao\n 1.0 1 0 1.0 1.0 1 0 1.0\n", "lcao\n 1 1 0 1 1 1 0 1\n"))


def test_missing_cards_are_refused():
    from x2dhf_amd import deck
    with pytest.raises(deck.DeckError, match="lcao"):
        deck.parse("nuclei 1.0 1.0 1.4\nconfig 0\n 1 sigma g end\ngrid 91 121 40.0\norbpot hydrogen\nstop\n")


@pytest.mark.gpu
def test_h2_deck_runs_to_the_reference_energy():
    """The configuration-1 deck, parsed and run on the GPU, reaches the
    reference program's energy for the same deck."""
    from x2dhf_amd import deck, scf
    ref = json.load(open(os.path.join(GOLDEN, "scf_config1_h2_91x121.json")))
    d = deck.load(os.path.join(DECKS, "config1_h2_91x121.data"))
    print("H2 config-1 deck through the front end", flush=True)
    job = scf.Job(d.spec())
    try:
        res = job.run()
    finally:
        job.close()
    print(f"E = {res['etot']:.16f} (ref {ref['etot']:.16f})")
    assert abs(res["etot"] - ref["etot"]) < 1e-10


def test_relaxation_scheme_cards():
    """Without solver cards the reference relaxes orbitals and potentials by
    SOR in the middle ordering (initVariables.F90:206, 253); mcsor / mcsor-o
    switch the orbitals to MCSOR, mcsor-ce the potentials
    (inputInterface.F90:1915-2028), order picks the mesh ordering."""
    from x2dhf_amd import deck
    base = "method hf\nnuclei 1.0 1.0 1.4\nconfig 0\n 1 sigma g end\ngrid 61 91 20.0\norbpot hydrogen\n" \
           "lcao\n 1.0 1 0 1.0 1.0 1 0 1.0\n"
    for extra, want in (("", (1, 1, 1)), ("mcsor\n", (0, 1, 1)), ("mcsor-o 1 10 1\n", (0, 1, 1)),
                        ("mcsor\nmcsor-ce\n", (0, 0, 1)), ("order col-wise\n", (1, 1, 0)),
                        ("mcsor 1\nmcsor-ce 1\norder middle\n", (0, 0, 1))):
        d = deck.parse(base + extra + "stop\n")
        assert (d.orb_method, d.pot_method, d.ordering) == want, extra
        s = d.spec()
        assert (s.orb_method, s.pot_method, s.ordering) == want
    # every ordering of the reference's mesh, read as 8 characters
    for name, code in (("row-wise", 3), ("rrow-wise", 4), ("middlemt", 2), ("test4", 8), ("middle2", 9)):
        d = deck.parse(base + f"order {name}\nstop\n")
        assert d.ordering == code and d.spec().ordering == code
        assert deck.parse(deck.to_deck(d.spec())).ordering == code
    with pytest.raises(deck.DeckError, match="undefined type of ordering"):
        deck.parse(base + "order diagonal\nstop\n")
    # altsweeps (read as its first 8 characters, inputInterface.F90:736)
    d = deck.parse(base + "altsweeps\nstop\n")
    assert d.alt_sweeps == 1 and d.spec().alt_sweeps == 1
    assert d.notes[-1].endswith("forward/backward SOR sweeps")
    assert deck.parse(deck.to_deck(d.spec())).alt_sweeps == 1
    assert deck.parse(base + "stop\n").spec().alt_sweeps == 0


def test_spec_to_deck_round_trip():
    """to_deck writes the reference deck of a job specification: parsing it
    gives the same specification (served-stream molecules, both schemes)."""
    from x2dhf_amd import deck, serve
    for _, name, spec in serve.poisson_stream(12, 10.0, seed=4, prec=0):
        spec.schedule = 0
        txt = deck.to_deck(spec, title=name)
        back = deck.parse(txt).spec()
        # the reference rounds the grid down to 1 mod 30 (nnucalc / nmucalc)
        assert (back.nni, back.nmu) == (deck.nnu_calc(spec.nni), deck.round_to_mesh(spec.nmu))
        back.nni, back.nmu = spec.nni, spec.nmu
        assert _fields(back) == _fields(spec), name
    d = deck.load(os.path.join(DECKS, "config3_fh_481x901.data"))
    s = d.spec()
    assert _fields(deck.parse(deck.to_deck(s)).spec()) == _fields(s)
    assert "mcsor 8\nmcsor-ce" in deck.to_deck(s, mcsor_threads=8)
