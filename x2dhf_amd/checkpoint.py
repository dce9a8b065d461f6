"""The reference's checkpoint files: <prefix>.dat / .orb / .coul / .exch.

x2dhf writes them at the end of a run (writeToDisk, diskInterface.F90:1143-1212,
and wtdisknat, dumpDataToDisk.F90:147-330) and reads them for `orbpot old`
(readFromDisk / readFuncs, diskInterface.F90:1099-1137, 247-779).  Native
format, the reference's default:

  .dat   text: the title (75 characters); date and version; "nnu nmu Rinf";
         "Z1 Z2 R"; "norb nel nexch"
  .orb   Fortran unformatted sequential records (4-byte length before and
         after each): the 15 address arrays i1b i2b i3b i1e ... i3mu (int32;
         maxorb, maxorb and maxorb(maxorb+1)/2 values of each triple); one
         record of nnu*nmu values per orbital; then orbNorm (maxorb), the
         Lagrange-multiplier matrix ee (maxorb x maxorb, column-major),
         cmulti (maxmpole*maxorb, moment m of orbital i at i + (m-1) norb),
         and eight exchange-moment arrays excdi,
         excqu, excoc, exche, exc5..exc8 (maxorb(maxorb+1) each)
  .coul  one record per orbital: its Coulomb potential
  .exch  the exchange potentials: for iorb1 = 1..norb, iorb2 = iorb1..norb,
         skipping iorb1 = iorb2 for sigma orbitals, one record, and a second
         one when iorb1 /= iorb2 and neither orbital is a sigma orbital

maxorb (36 in the reference's build, src/params.h) is read back from the
length of the address record.  The arrays keep the reference's layout
(column-major, nu fastest), the layout of this build's device arrays.  Files
are parsed with numpy from their bytes; nothing in them is executed.

Loading into a job maps the exchange records onto the job's exchange slots
(pairs k = i + j(j-1)/2 numbered with j outer, i inner; a second slot for
pi-pi pairs: x2dhf_amd/fortran/x2f_case.f90).  Writing from a job stores the
orbitals, potentials, norms, orbital energies (the diagonal of ee) and
multipole moments; the off-diagonal multipliers and exchange moments, which
the reference recomputes when it resumes, are written as zeros.
"""
import os
import time
from dataclasses import dataclass, field

import numpy as np

MAXORB = 36
MAXMPOLE = 8


@dataclass
class Checkpoint:
    title: str
    nni: int
    nmu: int
    rinf: float
    z1: float
    z2: float
    r: float
    norb: int
    nel: int
    nexch: int
    orbitals: np.ndarray                     # (norb, nni*nmu)
    coulomb: np.ndarray                      # (norb, nni*nmu)
    exchange: list                           # arrays in file order
    norms: np.ndarray = None                 # (maxorb,)
    ee: np.ndarray = None                    # (maxorb, maxorb)
    cmulti: np.ndarray = None                # (maxmpole*maxorb,)
    exc: np.ndarray = None                   # (8, maxorb*(maxorb+1))
    address: np.ndarray = None               # int32 address record as read
    stamp: str = ""
    maxorb: int = MAXORB

    @property
    def eps(self):
        """Orbital energies (diagonal of ee)."""
        return np.array([self.ee[i, i] for i in range(self.norb)])


# --------------------------------------------------------------------------
# Fortran sequential unformatted records

def _records(path):
    data = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(data):
        if pos + 4 > len(data):
            raise ValueError(f"{path}: truncated record marker at byte {pos}")
        n = int.from_bytes(data[pos:pos + 4], "little")
        end = pos + 4 + n
        if end + 4 > len(data) or int.from_bytes(data[end:end + 4], "little") != n:
            raise ValueError(f"{path}: record at byte {pos} is not a 4-byte-marker Fortran record")
        out.append(data[pos + 4:end])
        pos = end + 4
    return out


def _write_records(path, arrays):
    with open(path, "wb") as f:
        for a in arrays:
            b = np.ascontiguousarray(a).tobytes()
            n = len(b).to_bytes(4, "little")
            f.write(n)
            f.write(b)
            f.write(n)


def _f64(rec, count, what):
    a = np.frombuffer(rec, dtype="<f8")
    if count is not None and a.size != count:
        raise ValueError(f"{what}: {a.size} values, {count} expected")
    return a.copy()


# --------------------------------------------------------------------------
# the text header

def _fortran_e(v, width=25, digits=16):
    """Fortran E<width>.<digits> (0.dddE+xx)."""
    if v == 0.0:
        s = "0." + "0" * digits + "E+00"
    else:
        m, e = f"{abs(v):.{digits - 1}e}".split("e")
        s = ("-" if v < 0 else "") + "0." + m.replace(".", "") + f"E{int(e) + 1:+03d}"
    return s.rjust(width)


def read_header(path):
    lines = open(path).read().split("\n")
    if len(lines) < 5:
        raise ValueError(f"{path}: header of 5 lines expected")
    nni, nmu, rinf = lines[2].split()[:3]
    z1, z2, r = lines[3].split()[:3]
    norb, nel, nexch = lines[4].split()[:3]
    return dict(title=lines[0].rstrip(), stamp=lines[1].rstrip(), nni=int(nni), nmu=int(nmu),
                rinf=float(rinf.replace("D", "E")), z1=float(z1.replace("D", "E")), z2=float(z2.replace("D", "E")),
                r=float(r.replace("D", "E")), norb=int(norb), nel=int(nel), nexch=int(nexch))


def write_header(path, ck, version="x2dhf-amd"):
    stamp = time.strftime("%Y/%m/%d  %H:%M:%S.00  ")
    with open(path, "w") as f:
        f.write(f"{ck.title[:75]:<75}\n")
        f.write(f"{stamp:<25}{version:<15}\n")
        f.write(f"{ck.nni:5d}{ck.nmu:5d}{_fortran_e(ck.rinf)}   / nnu nmu Rinf\n")
        f.write(f"{_fortran_e(ck.z1)}{_fortran_e(ck.z2)}{_fortran_e(ck.r)}   / Z1 Z2 R\n")
        f.write(f"{ck.norb:5d}{ck.nel:5d}{ck.nexch:5d}    / norb nel nexch\n")


# --------------------------------------------------------------------------
# exchange records

def exchange_records(ms):
    """The exchange records of a configuration with magnetic quantum numbers
    ms (orbital order), in file order: [(iorb1, iorb2, second)] (1-based)."""
    out = []
    n = len(ms)
    for a in range(1, n + 1):
        for b in range(a, n + 1):
            if a == b and ms[a - 1] == 0:
                continue
            out.append((a, b, False))
            if a != b and ms[a - 1] != 0 and ms[b - 1] != 0:
                out.append((a, b, True))
    return out


def exchange_slots(ms):
    """{(i, j, second): slot} of the job's exchange slots (1-based), numbered
    as x2f_case.f90 numbers them."""
    slots, n = {}, 0
    for j in range(1, len(ms) + 1):
        for i in range(1, j + 1):
            if i == j and ms[i - 1] == 0:
                continue
            n += 1
            slots[(i, j, False)] = n
            if i != j and ms[i - 1] != 0 and ms[j - 1] != 0:
                n += 1
                slots[(i, j, True)] = n
    return slots


# --------------------------------------------------------------------------
# files

def read(prefix, ms=None):
    """Read <prefix>.dat/.orb/.coul/.exch.  ms (the orbitals' magnetic
    quantum numbers) is only used to check the number of exchange records."""
    h = read_header(prefix + ".dat")
    n = h["nni"] * h["nmu"]
    norb = h["norb"]
    orb = _records(prefix + ".orb")
    addr = np.frombuffer(orb[0], dtype="<i4").copy()
    # 5 x (2 maxorb + maxorb(maxorb+1)/2) address values
    per = addr.size // 5
    mo = int(round((-5 + np.sqrt(25 + 8 * per)) / 2))
    if 5 * (2 * mo + mo * (mo + 1) // 2) != addr.size:
        raise ValueError(f"{prefix}.orb: address record of {addr.size} values fits no maxorb")
    if len(orb) < 1 + norb:
        raise ValueError(f"{prefix}.orb: {len(orb) - 1} orbital records, {norb} expected")
    orbitals = np.stack([_f64(orb[1 + i], n, f"orbital {i + 1}") for i in range(norb)])
    rest = orb[1 + norb:]
    norms = ee = cmulti = exc = None
    if rest:
        norms = _f64(rest[0], mo, "orbNorm")
        ee = _f64(rest[1], None, "ee")
        ee = ee.reshape(mo, mo, order="F") if ee.size == mo * mo else np.diag(np.resize(ee, mo))
        cmulti = _f64(rest[2], None, "cmulti")
        exc = np.stack([_f64(r, None, "exchange moments") for r in rest[3:11]]) if len(rest) >= 11 else None
    coul = _records(prefix + ".coul")
    if len(coul) < norb:
        raise ValueError(f"{prefix}.coul: {len(coul)} records, {norb} expected")
    coulomb = np.stack([_f64(coul[i], n, f"Coulomb potential {i + 1}") for i in range(norb)])
    exchange = [_f64(r, n, "exchange potential") for r in _records(prefix + ".exch")] \
        if os.path.exists(prefix + ".exch") else []
    if ms is not None and len(exchange) != len(exchange_records(ms)):
        raise ValueError(f"{prefix}.exch: {len(exchange)} records, {len(exchange_records(ms))} expected")
    return Checkpoint(title=h["title"], nni=h["nni"], nmu=h["nmu"], rinf=h["rinf"], z1=h["z1"], z2=h["z2"],
                      r=h["r"], norb=norb, nel=h["nel"], nexch=h["nexch"], orbitals=orbitals, coulomb=coulomb,
                      exchange=exchange, norms=norms, ee=ee, cmulti=cmulti, exc=exc, address=addr,
                      stamp=h["stamp"], maxorb=mo)


def write(prefix, ck, version="x2dhf-amd"):
    """Write <prefix>.dat/.orb/.coul/.exch in the reference's native layout."""
    mo = ck.maxorb
    n = ck.nni * ck.nmu
    write_header(prefix + ".dat", ck, version)
    addr = ck.address
    if addr is None or addr.size != 5 * (2 * mo + mo * (mo + 1) // 2):
        addr = _address(ck)
    norms = np.zeros(mo) if ck.norms is None else np.resize(np.asarray(ck.norms, float), mo)
    ee = np.zeros((mo, mo)) if ck.ee is None else np.asarray(ck.ee, float)
    cm = np.zeros(MAXMPOLE * mo) if ck.cmulti is None else np.resize(np.asarray(ck.cmulti, float), MAXMPOLE * mo)
    exc = np.zeros((8, mo * (mo + 1))) if ck.exc is None else np.asarray(ck.exc, float)
    recs = [addr.astype("<i4")] + [np.asarray(o, "<f8").reshape(n) for o in ck.orbitals]
    recs += [norms.astype("<f8"), ee.astype("<f8").ravel(order="F"), cm.astype("<f8")]
    recs += [e.astype("<f8") for e in exc]
    _write_records(prefix + ".orb", recs)
    _write_records(prefix + ".coul", [np.asarray(c, "<f8").reshape(n) for c in ck.coulomb])
    _write_records(prefix + ".exch", [np.asarray(x, "<f8").reshape(n) for x in ck.exchange])


def _address(ck):
    """Address arrays of a fresh layout (orbitals, Coulomb and exchange
    potentials stored one after another, each nnu*nmu long)."""
    mo, n = ck.maxorb, ck.nni * ck.nmu
    mp = mo * (mo + 1) // 2
    a1 = np.zeros(mo, np.int32)
    a1[:ck.norb] = 1 + n * np.arange(ck.norb)
    e1 = np.zeros(mo, np.int32)
    e1[:ck.norb] = n
    a3 = np.zeros(mp, np.int32)
    k = len(ck.exchange)
    a3[:k] = 1 + n * np.arange(k)
    e3 = np.zeros(mp, np.int32)
    e3[:k] = n
    ng1, ng3 = (e1 > 0).astype(np.int32), (e3 > 0).astype(np.int32)
    mu1, mu3 = ng1 * ck.nmu, ng3 * ck.nmu
    return np.concatenate([a1, a1, a3, e1, e1, e3, e1, e1, e3, ng1, ng1, ng3, mu1, mu1, mu3]).astype(np.int32)


# --------------------------------------------------------------------------
# jobs

def load_into(job, ck):
    """Put a checkpoint's orbitals and potentials into a created job (before
    job.prep()); the job's grid, nuclei and orbitals must match it."""
    s = job.spec
    if (s.nni, s.nmu, s.norb) != (ck.nni, ck.nmu, ck.norb):
        raise ValueError(f"checkpoint {ck.nni}x{ck.nmu} with {ck.norb} orbitals does not match the job "
                         f"{s.nni}x{s.nmu} with {s.norb}")
    if max(abs(s.z1 - ck.z1), abs(s.z2 - ck.z2), abs(s.r - ck.r), abs(s.rinf - ck.rinf)) > 1e-10:
        raise ValueError("checkpoint nuclei / Rinf do not match the job")
    ms = [s.m[i] for i in range(s.norb)]
    recs = exchange_records(ms)
    if len(ck.exchange) != len(recs):
        raise ValueError(f"checkpoint has {len(ck.exchange)} exchange potentials, the job {len(recs)}")
    slots = exchange_slots(ms)
    for i in range(ck.norb):
        job.put(0, i + 1, ck.orbitals[i])
        job.put(1, i + 1, ck.coulomb[i])
    for rec, x in zip(recs, ck.exchange):
        job.put(2, slots[rec], x)


def from_job(job, title=""):
    """A checkpoint of a job's current orbitals, potentials, norms, orbital
    energies and multipole moments."""
    s = job.spec
    r = job.result()
    ms = [s.m[i] for i in range(s.norb)]
    slots = exchange_slots(ms)
    recs = exchange_records(ms)
    mo = MAXORB
    ee = np.zeros((mo, mo))
    norms = np.zeros(mo)
    cm = np.zeros(MAXMPOLE * mo)
    for i in range(s.norb):
        ee[i, i] = r["eps"][i]
        norms[i] = r["norm"][i]
        for k in range(min(MAXMPOLE, r["cmulti"].shape[1])):
            cm[i + k * s.norb] = r["cmulti"][i][k]  # cmulti(iorb + (m-1) norb), scfUtils.F90:1063
    nel = int(round(sum(s.occ[i] for i in range(s.norb))))
    return Checkpoint(title=title, nni=s.nni, nmu=s.nmu, rinf=s.rinf, z1=s.z1, z2=s.z2, r=s.r, norb=s.norb,
                      nel=nel, nexch=len(recs), orbitals=np.stack([job.get(0, i + 1) for i in range(s.norb)]),
                      coulomb=np.stack([job.get(1, i + 1) for i in range(s.norb)]),
                      exchange=[job.get(2, slots[rec]) for rec in recs], norms=norms, ee=ee, cmulti=cm)


def resume(spec, ck, ctx=None):
    """A job of spec started from checkpoint ck (orbpot old): the functions
    are loaded and, when the checkpoint carries the Lagrange multipliers,
    the set-up pass takes them and the orbitals as they are, as the
    reference's prepSCF does after a complete read (initAddData false);
    otherwise it normalises, orthogonalises and forms the multipliers."""
    from . import scf
    job = scf.Job(spec, ctx=ctx)
    try:
        load_into(job, ck)
        if ck.ee is not None:
            job.prep_restart(ck.ee)
        else:
            job.prep()
    except Exception:
        job.close()
        raise
    return job
