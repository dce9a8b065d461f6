"""The SCF loop's host bookkeeping (the Fortran driver's energy_change and
convergence_control, doSCF.F90:546-744) run on the CPU with given orbital
energies and norms: the reference's stop rules, and the failure guards that
retire a job whose energies turn NaN or diverge instead of iterating to
maxscf.  No device call."""
import numpy as np

from x2dhf_amd import scf

EE0 = np.array([-15.7, -1.5, -0.6])


def converging(nit, rate=0.7, amp=1e-2):
    """Orbital energies approaching EE0 geometrically, norms approaching 1."""
    k = np.arange(1, nit + 1)[:, None]
    ee = EE0[None, :] + amp * rate ** k * np.array([1.0, -0.5, 0.25])[None, :]
    norms = 1.0 + 1e-3 * rate ** k * np.ones((1, 3))
    return ee, norms


def test_energy_threshold_stop():
    ee, nm = converging(200)
    why, it = scf.control_probe(ee, nm, EE0 + 0.05, maxscf=1000, ienterm=10, inoterm=20)
    # |d eps| < 1e-10 first at iteration `first`; then the count of further
    # such iterations must reach nscfextra = 3 before the next one stops
    # (doSCF.F90:648-668)
    assert why == "orbital energy threshold", why
    first = int(np.argmax(np.abs(np.diff(ee, axis=0)).max(axis=1) < 1e-10)) + 2
    assert it == first + 4, (it, first)


def test_iteration_limit_stop():
    ee, nm = converging(50, rate=0.99)
    why, it = scf.control_probe(ee, nm, EE0, maxscf=40, ienterm=12, inoterm=12)
    assert (why, it) == ("iteration limit", 40)


def test_nan_in_an_orbital_energy_retires_the_job():
    ee, nm = converging(100, rate=0.95)
    ee[17, 1] = np.nan
    why, it = scf.control_probe(ee, nm, EE0, maxscf=8000, ienterm=12, inoterm=12)
    assert (why, it) == ("not a number", 18)


def test_nan_in_a_norm_retires_the_job():
    ee, nm = converging(100, rate=0.95)
    nm[30, 2] = np.nan
    why, it = scf.control_probe(ee, nm, EE0, maxscf=8000, ienterm=12, inoterm=12)
    assert (why, it) == ("not a number", 31)


def test_divergence_retires_the_job():
    """An orbital-energy change above orbEnergyIncThld = 1e4
    (initVariables.F90:187-191) stops the SCF at that iteration."""
    ee, nm = converging(100, rate=0.95)
    ee[25:, 0] -= 2.0e4
    why, it = scf.control_probe(ee, nm, EE0, maxscf=8000, ienterm=12, inoterm=12)
    assert (why, it) == ("diverged", 26)
    # a large but admissible change does not
    ee, nm = converging(60, rate=0.95)
    ee[25:, 0] -= 5.0e3
    why, it = scf.control_probe(ee, nm, EE0, maxscf=60, ienterm=12, inoterm=12)
    assert (why, it) == ("iteration limit", 60)


def test_saturation_stop():
    """No decrease of the largest energy change for nenlast = 20 iterations
    after nscf2skip (doSCF.F90:586-603)."""
    ee, nm = converging(400, rate=0.9)
    rng = np.random.default_rng(1)
    ee[100:] += 1e-9 * rng.standard_normal(ee[100:].shape)   # stalls at 1e-9
    nm[100:] = 1.0 + 1e-9 * np.abs(rng.standard_normal(nm[100:].shape))
    why, it = scf.control_probe(ee, nm, EE0, maxscf=400, ienterm=12, inoterm=12, nscf2skip=50)
    assert why == "saturation", why
    assert 100 < it < 400
