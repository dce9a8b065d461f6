"""The SCF loop's end-of-iteration control on the device (x2g_scf_control,
one thread per job) against the host rules (the Fortran driver's
energy_change + convergence_control, doSCF.F90:546-744, run by
scf.control_probe on the CPU): every scenario of tests/test_scf_control.py
and random orbital-energy / norm sequences stop at the same iteration with
the same stop code, many jobs per launch.  Then whole batched SCF runs with
the control on the device and on the host give the same iterations and
bit-identical energies."""
import ctypes as C

import numpy as np
import pytest

from test_scf_control import EE0, converging

pytestmark = pytest.mark.gpu

MAXO = 16


def _scenarios():
    """(name, ee[nit, norb], norms[nit, norb], ee0, maxscf, ienterm, inoterm, nscf2skip)"""
    out = []
    ee, nm = converging(200)
    out.append(("energy threshold", ee, nm, EE0 + 0.05, 1000, 10, 20, 0))
    ee, nm = converging(50, rate=0.99)
    out.append(("iteration limit", ee, nm, EE0, 40, 12, 12, 0))
    ee, nm = converging(100, rate=0.95)
    ee[17, 1] = np.nan
    out.append(("nan energy", ee, nm, EE0, 8000, 12, 12, 0))
    ee, nm = converging(100, rate=0.95)
    nm[30, 2] = np.nan
    out.append(("nan norm", ee, nm, EE0, 8000, 12, 12, 0))
    ee, nm = converging(100, rate=0.95)
    ee[25:, 0] -= 2.0e4
    out.append(("diverged", ee, nm, EE0, 8000, 12, 12, 0))
    ee, nm = converging(400, rate=0.9)
    rng = np.random.default_rng(1)
    ee[100:] += 1e-9 * rng.standard_normal(ee[100:].shape)
    nm[100:] = 1.0 + 1e-9 * np.abs(rng.standard_normal(nm[100:].shape))
    out.append(("saturation", ee, nm, EE0, 400, 12, 12, 50))
    # norm threshold first: energies keep moving at 1e-9, norms settle
    ee, nm = converging(300, rate=0.9)
    ee[60:] += 1e-9 * np.sin(np.arange(240))[:, None]
    out.append(("norm threshold", ee, nm, EE0, 300, 12, 8, 0))
    # random sequences: geometric convergence with noise floors and bumps
    rng = np.random.default_rng(7)
    for k in range(24):
        norb = int(rng.integers(1, 9))
        nit = int(rng.integers(40, 400))
        e0 = -np.sort(rng.uniform(0.2, 30.0, norb))[::-1]
        rate = rng.uniform(0.7, 0.98)
        it = np.arange(1, nit + 1)[:, None]
        ee = e0[None, :] + rng.uniform(1e-3, 1e-1) * rate ** it * rng.standard_normal(norb)[None, :]
        ee += 10.0 ** rng.uniform(-13, -8) * rng.standard_normal((nit, norb))
        if rng.random() < 0.3:
            ee[int(rng.integers(5, nit)):, int(rng.integers(norb))] += rng.uniform(1e-6, 1e-3)
        nm = 1.0 + rng.uniform(1e-4, 1e-2) * rate ** it * np.abs(rng.standard_normal(norb))[None, :]
        nm += 10.0 ** rng.uniform(-13, -9) * np.abs(rng.standard_normal((nit, norb)))
        out.append((f"random {k}", ee, nm, e0 + rng.uniform(0.01, 0.1), int(rng.integers(nit // 2, nit + 50)),
                    int(rng.integers(7, 13)), int(rng.integers(6, 12)), int(rng.integers(0, 80))))
    return out


def _record(norb, ee0, maxscf, ienterm, inoterm, nscf2skip):
    """The control record of a job in the state the host rules start from
    (scf_opts defaults, the scf_job type's initial values)."""
    from x2dhf_amd import gpu
    c = gpu.ScfCtrl()
    c.ethld, c.nthld, c.thren4init = 10.0 ** -ienterm, 10.0 ** -inoterm, 10.0 ** -1
    c.recalc_mm, c.inc_thld = 1.15, 1.0e4
    for i in range(MAXO):
        c.ee[i] = ee0[i] if i < norb else 0.0
        c.eeold[i] = c.ee[i]
        c.orbnorm[i] = 1.0
    c.eeconv, c.eeconvprev, c.dnmaxprev, c.deltaeemax, c.deltanormmax = 10.0, 1e35, 1e35, 100.0, 100.0
    c.norb, c.maxscf = norb, maxscf
    c.nscf2skip = nscf2skip if nscf2skip > 0 else 600
    c.nenlast, c.nnolast, c.nscfextra, c.maxresets, c.polish = 20, 20, 3, 5, 0
    return c


def test_device_control_stops_where_the_host_rules_stop():
    from x2dhf_amd import gpu, scf
    sc = _scenarios()
    want = []
    for name, ee, nm, e0, maxscf, ient, inot, skip in sc:
        want.append(scf.control_probe(ee, nm, e0, maxscf=maxscf, ienterm=ient, inoterm=inot, nscf2skip=skip))
    ctx = gpu.Context(0)
    try:
        nbytes = C.sizeof(gpu.ScfCtrl)
        ctrl = [ctx.alloc(nbytes // 8) for _ in sc]
        rec = [ctx.alloc(2 * MAXO) for _ in sc]
        for k, (name, ee, nm, e0, maxscf, ient, inot, skip) in enumerate(sc):
            c = _record(ee.shape[1], e0, maxscf, ient, inot, skip)
            gpu.check(ctx.lib.x2g_h2d(ctx.h, ctrl[k].ptr, C.byref(c), nbytes))
        got = [None] * len(sc)
        running = list(range(len(sc)))
        it = 0
        while running:
            it += 1
            for k in running:
                ee, nm = sc[k][1], sc[k][2]
                r = np.zeros(2 * MAXO)
                if it <= ee.shape[0]:
                    r[:ee.shape[1]] = nm[it - 1]
                    r[MAXO:MAXO + ee.shape[1]] = ee[it - 1]
                rec[k].set(r)
            cp = (C.c_void_p * len(running))(*[ctrl[k].ptr for k in running])
            rp = (C.c_void_p * len(running))(*[rec[k].ptr for k in running])
            gpu.check(ctx.lib.x2g_scf_control(ctx.h, cp, rp, len(running)))
            ctx.sync()
            still = []
            for k in running:
                c = gpu.ScfCtrl()
                gpu.check(ctx.lib.x2g_d2h(ctx.h, C.byref(c), ctrl[k].ptr, nbytes))
                assert c.iscf == it
                if c.stop:
                    got[k] = (scf.STOPCODES[c.stopcode], c.iscf)
                elif it >= sc[k][1].shape[0]:
                    got[k] = ("running", sc[k][1].shape[0])
                else:
                    still.append(k)
            running = still
        for k, s in enumerate(sc):
            print(f"{s[0]}: host {want[k]} device {got[k]}")
        for k, s in enumerate(sc):
            assert got[k] == want[k], (s[0], got[k], want[k])
        # the scenarios cover every stop the rules have, except resets
        codes = {w[0] for w in want}
        assert {"orbital energy threshold", "orbital norm threshold", "iteration limit", "saturation",
                "diverged", "not a number"} <= codes, codes
    finally:
        ctx.close()


def _batched(specs):
    from x2dhf_amd import scf
    jobs = [scf.Job(specs[0])]
    jobs += [scf.Job(s, share=jobs[0]) for s in specs[1:]]
    try:
        for j in jobs:
            j.prep()
        stopped = [False] * len(jobs)
        while not all(stopped):
            stopped = scf.iterate_batched(jobs, stopped)
        for j in jobs:
            j.finish()
        return [j.result() for j in jobs]
    finally:
        for j in reversed(jobs):
            j.close()


@pytest.mark.parametrize("case", ["n2_fp64", "fh_mixed_polish"])
def test_batched_scf_device_control_equals_host_control(case):
    """Jobs stepped by the batched schedule make the same decisions with the
    control on the device as with the host rules: the same number of
    iterations, stop, and bit-identical energies.  N2 (FP64 sweeps, a
    hydrogenic start whose frozen orbitals are released by the control) and
    FH with FP32 correction sweeps and an FP64 finish of 20 iterations (the
    control switches the precision)."""
    from x2dhf_amd import scf
    if case == "n2_fp64":
        mk = lambda hc: [scf.n2_spec(91, 121, 40.0, ienterm=12, inoterm=12, maxscf=3000, schedule=1, prec=0,
                                     host_control=hc)]
    else:
        mk = lambda hc: [scf.fh_spec(91, 121, 40.0, ienterm=12, inoterm=12, maxscf=3000, schedule=1, prec=1,
                                     polish=20, host_control=hc),
                         scf.h2_spec(61, 91, 40.0, ienterm=11, inoterm=11, maxscf=2000, schedule=1, prec=1,
                                     polish=20, host_control=hc)]
    print(f"{case}: device control", flush=True)
    dev = _batched(mk(0))
    print(f"{case}: host control", flush=True)
    host = _batched(mk(1))
    for a, b in zip(dev, host):
        print(f"device: {a['iterations']} iterations, stop '{a['stop']}', E = {a['etot']:.15f}; "
              f"host: {b['iterations']}, '{b['stop']}', E = {b['etot']:.15f}")
        assert a["iterations"] == b["iterations"]
        assert a["stop"] == b["stop"]
        assert a["etot"] == b["etot"]
        assert np.array_equal(a["eps"], b["eps"])
