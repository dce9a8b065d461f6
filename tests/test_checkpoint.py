"""The reference's checkpoint files (x2dhf_amd.checkpoint) against the
reference's own Be fixture (test-sets/be/set-01/be-hf.*, copied to
tests/golden/be-hf/): its layout is read back exactly and rewritten byte for
byte; loaded into a GPU job, the set-up pass gives the total energy the
reference's listing prints for the same files before any relaxation; and a
checkpoint written by this build resumes a job at the energy it was written
with."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

BE = os.path.join(GOLDEN, "be-hf", "be-hf")
DECKS = os.path.join(GOLDEN, "decks")
# test-sets/be/set-01/reference.lst:93, the energy after prepSCF on the files
E_BE_LISTING = -1.4573023352634815e+01


def test_be_fixture_layout():
    from x2dhf_amd import checkpoint as ckm
    ck = ckm.read(BE, ms=[0, 0])
    assert (ck.nni, ck.nmu, ck.norb, ck.nel, ck.maxorb) == (151, 181, 2, 4, 36)
    assert abs(ck.rinf - 35.0) < 1e-13 and (ck.z1, ck.z2) == (4.0, 0.0) and abs(ck.r - 2.386) < 1e-14
    assert ck.orbitals.shape == (2, 151 * 181) and ck.coulomb.shape == (2, 151 * 181)
    assert len(ck.exchange) == 1                                    # the 1s-2s pair
    # orbital energies of the converged Be functions (2 sigma, 1 sigma)
    assert abs(ck.eps[0] - (-0.3092702)) < 1e-6 and abs(ck.eps[1] - (-4.7326726)) < 1e-6
    # both orbitals normalised to 1e-6 (the reference stores the norms it had)
    assert np.all(np.abs(ck.norms[:2] - 1.0) < 1e-5)


def test_be_fixture_rewritten_byte_for_byte(tmp_path):
    from x2dhf_amd import checkpoint as ckm
    ck = ckm.read(BE)
    out = str(tmp_path / "be")
    ckm.write(out, ck)
    for ext in ("orb", "coul", "exch"):
        assert open(out + "." + ext, "rb").read() == open(BE + "." + ext, "rb").read(), ext
    h, h0 = ckm.read_header(out + ".dat"), ckm.read_header(BE + ".dat")
    for k in ("title", "nni", "nmu", "norb", "nel", "nexch", "rinf", "z1", "z2", "r"):
        assert h[k] == h0[k], k


def test_fortran_e_format():
    from x2dhf_amd import checkpoint as ckm
    assert ckm._fortran_e(34.99999999999999) == "   0.3499999999999999E+02"
    assert ckm._fortran_e(0.0) == "   0.0000000000000000E+00"
    assert ckm._fortran_e(-2.386000000000001) == "  -0.2386000000000001E+01"


def test_exchange_records_and_slots():
    from x2dhf_amd import checkpoint as ckm
    ms = [1, 0, 0, 0, 0, 0]                     # N2: 1 pi, 5 sigma (config 2)
    recs = ckm.exchange_records(ms)
    slots = ckm.exchange_slots(ms)
    # 15 pairs and the pi self-exchange: 16 exchange potentials (22 with Coulomb)
    assert len(recs) == 16 and sorted(slots.values()) == list(range(1, 17))
    assert set(recs) == set(slots)
    assert recs[0] == (1, 1, False)             # file order: iorb1 outer
    ms = [1, 1, 0]                              # two pi orbitals: a pi-pi pair has two records
    recs = ckm.exchange_records(ms)
    assert recs[:3] == [(1, 1, False), (1, 2, False), (1, 2, True)]
    assert len(recs) == len(ckm.exchange_slots(ms)) == 6


def test_synthetic_round_trip(tmp_path):
    from x2dhf_amd import checkpoint as ckm
    rng = np.random.default_rng(0)
    n = 31 * 61
    ck = ckm.Checkpoint(title="round trip", nni=31, nmu=61, rinf=25.0, z1=3.0, z2=1.0, r=3.015, norb=2, nel=4,
                        nexch=1, orbitals=rng.standard_normal((2, n)), coulomb=rng.standard_normal((2, n)),
                        exchange=[rng.standard_normal(n)], ee=np.diag(np.r_[-2.4, -0.3, np.zeros(34)]))
    out = str(tmp_path / "x")
    ckm.write(out, ck)
    back = ckm.read(out, ms=[0, 0])
    assert np.array_equal(back.orbitals, ck.orbitals) and np.array_equal(back.coulomb, ck.coulomb)
    assert np.array_equal(back.exchange[0], ck.exchange[0]) and np.array_equal(back.ee, ck.ee)
    assert (back.nni, back.nmu, back.norb, back.maxorb) == (31, 61, 2, 36)


def test_truncated_file_is_refused(tmp_path):
    from x2dhf_amd import checkpoint as ckm
    for ext in ("dat", "coul", "exch"):
        (tmp_path / f"b.{ext}").write_bytes(open(BE + "." + ext, "rb").read())
    (tmp_path / "b.orb").write_bytes(open(BE + ".orb", "rb").read()[:20000])
    with pytest.raises(ValueError, match="Fortran record"):
        ckm.read(str(tmp_path / "b"))


@pytest.mark.gpu
def test_be_checkpoint_energy_matches_reference_listing():
    """The Be deck (orbpot old) and its checkpoint: the set-up pass on the
    loaded functions gives the energy of reference.lst:93."""
    from x2dhf_amd import checkpoint as ckm, deck
    d = deck.load(os.path.join(DECKS, "ref_be_set-01.data"))
    ck = ckm.read(BE, ms=[m for m, _ in d.orbitals])
    print("Be checkpoint resume", flush=True)
    job = ckm.resume(d.spec(), ck)
    try:
        r = job.result()
    finally:
        job.close()
    print(f"E = {r['etot']:.16f} (listing {E_BE_LISTING:.16f}), diff {r['etot'] - E_BE_LISTING:.2e}")
    assert abs(r["etot"] - E_BE_LISTING) < 1e-10


@pytest.mark.gpu
def test_checkpoint_resumes_a_converged_job(tmp_path):
    """H2 (configuration 1 deck) to convergence, written, read back and
    resumed: the resumed job starts at the written energy and stops within a
    few iterations."""
    from x2dhf_amd import checkpoint as ckm, deck, scf
    d = deck.load(os.path.join(DECKS, "config1_h2_91x121.data"))
    print("H2 run for the checkpoint", flush=True)
    job = scf.Job(d.spec())
    try:
        res = job.run()
        ck = ckm.from_job(job, title="H2 91x121")
    finally:
        job.close()
    out = str(tmp_path / "h2")
    ckm.write(out, ck)
    back = ckm.read(out, ms=[0])
    print("resume", flush=True)
    job = ckm.resume(d.spec(), back)
    try:
        e0 = job.result()["etot"]
        for it in range(40):
            if job.iterate(1):
                break
        job.finish()
        e1 = job.result()["etot"]
    finally:
        job.close()
    print(f"written {res['etot']:.16f} resumed {e0:.16f} after {it + 1} iterations {e1:.16f}")
    assert abs(e0 - res["etot"]) < 1e-12
    assert abs(e1 - res["etot"]) < 1e-10
    assert it < 39
    ref = json.load(open(os.path.join(GOLDEN, "scf_config1_h2_91x121.json")))
    assert abs(e1 - ref["etot"]) < 1e-10
