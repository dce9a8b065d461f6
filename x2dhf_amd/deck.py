"""Input decks in the reference's card grammar -> job specifications.

x2dhf reads a deck of cards (inputInterface.F90:681-818): one label per card,
then its items, blank-separated; '!' starts a comment (a whole line starting with '#' too); labels and
keywords count by their first 8 characters, as the reference reads them.
This module reads the cards that define a closed-shell Hartree-Fock job on the
relaxation path and builds the scf.JobSpec the GPU driver runs:

  title ...                          ignored
  method hf                          only HF (read_method, 2030-2073)
  nuclei Z1 Z2 R [angstrom]          R in bohr unless 'angstrom' (read_nuclei, 2155-2184)
  config q [Z1 Z2]                   then one card per orbital group:
     n  sigma|pi|delta|phi [g|u] [end]   n orbitals of that symmetry, closed shells only
                                         (read_config, 862-1077)
  grid nnu [nmu] Rinf                nnu and nmu rounded down to 1 mod 30; without nmu
                                     it follows from Rinf (read_grid 1491-1556,
                                     nmucalc 2961-3037, nnucalc 3043-3095)
  orbpot hydrogen [ienterm4init]     hydrogenic LCAO start, or
  orbpot old                         a start from a checkpoint (read_orbpot, 2257-2360)
  lcao [1|2]                         then one card per orbital:
     c1 n1 l1 Z1' c2 n2 l2 Z2' [..]  with '2' the exponents are Z - value (read_lcao, 1749-1881)
  scf maxscf [save ienterm inoterm verbose]   (read_scf, 2735-2783)
  conv nscf2skip [nenlast [nnolast]]          saturation test (read_conv, 1079-1102)
  omega w_orb [w_pot]                          (read_omega, 2186-2222)
  multipol recalc [mpole]                      (read_multipol, 2124-2153)
  sor [orb [pot]], sor4orb n [m], sor4pot n [m]   sweep counts (read_sor, 2806-2880)
  mcsor [n ..], mcsor-o [..]         orbitals by MCSOR (read_mcsor, read_mcsor_o, 1915-1995)
  mcsor-ce [..]                      potentials by MCSOR (read_mcsor_ce, 1997-2028); without
                                     these cards both are relaxed by SOR, the reference's default
  order name                         mesh ordering (read_order, 8 characters: col-wise,
                                     middle (the default, initVariables.F90:206), middlemt,
                                     row-wise, rrow-wis(e), test1..test4, middle2)
  altsweeps                          SOR sweeps forward for even and backward for odd
                                     orbitals (inputInterface.F90:736, doSCF.F90:403-411)
  print, prtevery, mmoments, plot, out4pair, debug, inout, stop   ignored

Any other card (DFT, external fields, fixed orbitals, other start orbitals,
...) is outside this path and raises DeckError, as does an open-shell
configuration.
"""
import math
from dataclasses import dataclass, field

from . import scf

BOHR2ANG = 0.529177210671212  # commons.F90:21
_SYM_M = {"sigma": 0, "pi": 1, "delta": 2, "phi": 3}
_IGNORED = {"title", "print", "prtevery", "mmoments", "plot", "out4pair", "debug", "inout", "stop"}
# the reference's mesh orderings by their 8-character `order` names
# (initCBAllocArrays.F90:2458-2810), as the X2G_ORDER_* codes
_ORDERINGS = {"col-wise": 0, "middle": 1, "middlemt": 2, "row-wise": 3, "rrow-wis": 4, "test1": 5, "test2": 6,
              "test3": 7, "test4": 8, "middle2": 9}


class DeckError(ValueError):
    pass


def round_to_mesh(n):
    """The largest n' <= n of the form 6k+1 and 5k+6 (both loops of
    nnucalc / nmucalc), i.e. n' = 1 (mod 30)."""
    if n < 31:
        raise DeckError(f"cannot find a commensurate grid size for {n} points")
    return n - (n - 1) % 30


def nnu_calc(n):
    """nnucalc (inputInterface.F90:3043-3095)."""
    return round_to_mesh(n)


def nmu_calc(n, nni, r, rinf):
    """nmucalc (inputInterface.F90:2961-3037): n = 0 takes the mu step equal
    to the nu step, pi/(nni-1), up to xi_max = 2 Rinf / R."""
    if n == 0:
        hnu = math.pi / (nni - 1)
        x = 2.0 * rinf / r
        n = int(math.floor(math.log(x + math.sqrt(x * x - 1.0)) / hnu + 1.0 + 0.5))
    return round_to_mesh(n)


@dataclass
class Deck:
    title: str = ""
    z1: float = 0.0
    z2: float = 0.0
    r: float = 0.0
    charge: int = 0
    orbitals: list = field(default_factory=list)   # (m, gu) in deck order
    nnu_card: int = 0
    nmu_card: int = 0
    rinf: float = 0.0
    nni: int = 0
    nmu: int = 0
    ienterm4init: int = 1
    lcao: list = field(default_factory=list)       # (c1, n1, l1, za, c2, n2, l2, zb)
    maxscf: int = 1000
    ienterm: int = 10
    inoterm: int = 10
    omega: tuple = (-1.0, -1.0)
    conv: tuple = (0, 0, 0)
    mpole: int = 4
    recalc_mm: float = 1.15
    maxsororb: tuple = (1, 10)
    maxsorpot: tuple = (1, 10)
    start: str = "hydrogen"                        # orbpot: hydrogen | old
    orb_method: int = 1                            # 0 MCSOR, 1 SOR (the reference's default)
    pot_method: int = 1
    ordering: int = 1                              # 0 col-wise, 1 middle (the default)
    alt_sweeps: int = 0                            # altsweeps card
    notes: list = field(default_factory=list)

    def spec(self, prec=0, schedule=0, device=0, polish=0):
        """The scf.JobSpec of this deck (polish: FP64 finish iterations of a
        prec = 1 job, see scf.make_spec).  For orbpot old the LCAO guess is a
        placeholder (one hydrogenic function per orbital), replaced by the
        checkpoint's functions (x2dhf_amd.checkpoint.resume)."""
        if self.start == "old" and not self.lcao:
            zc = self.z1 if self.z1 > 0 else self.z2
            self.lcao = [(1.0, m + 1, m, zc, 0.0, m + 1, m, zc) for m, _ in self.orbitals]
        if len(self.lcao) != len(self.orbitals):
            raise DeckError("the deck needs one lcao card per orbital")
        orbs = [scf.Orbital(m, gu, *lc) for (m, gu), lc in zip(self.orbitals, self.lcao)]
        return scf.make_spec(self.z1, self.z2, self.r, self.nni, self.nmu, self.rinf, orbs, maxscf=self.maxscf,
                             ienterm=self.ienterm, inoterm=self.inoterm, omega=self.omega, prec=prec,
                             schedule=schedule, device=device, maxsor=self.maxsororb, maxsorpot=self.maxsorpot,
                             mpole=self.mpole, ienterm4init=self.ienterm4init, recalc_mm=self.recalc_mm,
                             conv=self.conv, polish=polish, orb_method=self.orb_method,
                             pot_method=self.pot_method, ordering=self.ordering, alt_sweeps=self.alt_sweeps)


def _cards(text):
    for raw in text.splitlines():
        line = raw.split("!")[0].strip()  # '!' starts a comment, also after items
        if not line or line[0] == "#":
            continue
        yield line.split()


def _word(tok):
    """A label or keyword as the reference reads it: into a character*8
    variable, so only its first 8 characters count ('angstroms' is
    'angstrom')."""
    return tok.lower()[:8]


def _num(tok, kind, card):
    try:
        return kind(tok)
    except ValueError:
        raise DeckError(f"{card}: '{tok}' is not a number") from None


def _ints(items, card):
    out = []
    for t in items:
        try:
            out.append(int(t))
        except ValueError:
            break
    return out


def parse(text):
    """Parse a deck (the text of an x2dhf input file) into a Deck."""
    d = Deck()
    cards = list(_cards(text))
    seen = set()
    i = 0
    while i < len(cards):
        tok = cards[i]
        label, items = _word(tok[0]), tok[1:]
        i += 1
        seen.add(label)
        if label == "stop":
            break
        if label == "title":
            d.title = " ".join(items)
        elif label in _IGNORED:
            pass
        elif label == "method":
            if not items or _word(items[0]) != "hf":
                raise DeckError(f"method {' '.join(items)}: only HF runs on this path")
        elif label == "nuclei":
            if len(items) < 3:
                raise DeckError("nuclei: Z1 Z2 R expected")
            d.z1, d.z2, d.r = (_num(t, float, label) for t in items[:3])
            if len(items) > 3 and _word(items[3]) == "angstrom":
                d.r /= BOHR2ANG
            if abs(d.z1) + abs(d.z2) == 0.0 or d.r == 0.0:
                raise DeckError("nuclei: incomplete input")
        elif label == "config":
            if items:
                d.charge = _num(items[0], int, label)
            if len(items) >= 3:
                z1, z2 = _num(items[1], float, label), _num(items[2], float, label)
                d.z1, d.z2 = (z1 if z1 > 0 else d.z1), (z2 if z2 > 0 else d.z2)
            while True:
                if i >= len(cards):
                    raise DeckError("config: missing 'end'")
                t = [_word(x) for x in cards[i]]
                i += 1
                n = _num(t[0], int, "config")
                if len(t) < 2 or t[1] not in _SYM_M:
                    raise DeckError(f"config: wrong symmetry or symmetry higher than phi: {' '.join(t)}")
                rest = t[2:]
                gu = 0
                if rest and rest[0] in ("g", "u"):
                    gu = 1 if rest[0] == "g" else -1
                    rest = rest[1:]
                if rest and rest[0] != "end":
                    raise DeckError(f"config: open-shell entries '{' '.join(rest)}' (closed shells only)")
                d.orbitals += [(_SYM_M[t[1]], gu)] * n
                if rest and rest[0] == "end":
                    break
        elif label == "grid":
            nums = [_num(t, float, label) for t in items]
            if len(nums) >= 3 and nums[2] != 0.0:
                d.nnu_card, d.nmu_card, d.rinf = int(nums[0]), int(round(nums[1])), nums[2]
            elif len(nums) >= 2:
                d.nnu_card, d.nmu_card, d.rinf = int(nums[0]), 0, nums[1]
            else:
                raise DeckError("grid: nnu [nmu] Rinf expected")
        elif label == "orbpot":
            kind = _word(items[0]) if items else ""
            if kind == "old":
                # functions from a checkpoint (x2dhf_amd.checkpoint); no
                # frozen-orbital start (read_orbpot leaves lfixorb4init off)
                d.start, d.ienterm4init = "old", 0
            elif kind == "hydrogen":
                d.start = "hydrogen"
                n = _ints(items[1:], label)
                if n:
                    d.ienterm4init = n[0]
            else:
                raise DeckError(f"orbpot {' '.join(items)}: only the hydrogenic LCAO start and checkpoints "
                                "(orbpot old) run on this path")
        elif label == "lcao":
            incl = _ints(items, label)
            mode = incl[0] if incl else 1
            if mode not in (1, 2):
                raise DeckError(f"lcao {mode}: 1 or 2 expected")
            norb = len(d.orbitals)
            if norb == 0:
                raise DeckError("lcao before config")
            d.lcao = []
            for _ in range(norb):
                if i >= len(cards):
                    raise DeckError("lcao: fewer cards than orbitals")
                t = cards[i]
                i += 1
                if len(t) < 8:
                    raise DeckError(f"lcao: 8 items expected: {' '.join(t)}")
                c1, n1, l1, za = (_num(t[0], float, label), _num(t[1], int, label), _num(t[2], int, label),
                                  _num(t[3], float, label))
                c2, n2, l2, zb = (_num(t[4], float, label), _num(t[5], int, label), _num(t[6], int, label),
                                  _num(t[7], float, label))
                if mode == 2:
                    za, zb = d.z1 - za, d.z2 - zb
                d.lcao.append((c1, n1, l1, za, c2, n2, l2, zb))
        elif label == "scf":
            n = _ints(items, label)
            if n:
                d.maxscf = n[0]
            if len(n) > 2:
                d.ienterm = n[2]
            if len(n) > 3:
                d.inoterm = n[3]
        elif label == "conv":
            n = _ints(items, label)
            d.conv = tuple((n + [0, 0, 0])[:3])
        elif label == "omega":
            w = [_num(t, float, label) for t in items[:2]]
            if not w or all(x == 0.0 for x in w):
                raise DeckError("omega: values expected on the card")
            wo = w[0] if w[0] != 0.0 else d.omega[0]
            wp = w[1] if len(w) > 1 and w[1] != 0.0 else d.omega[1]
            d.omega = (wo, wp)
        elif label == "multipol":
            if items:
                f = _num(items[0], float, label)
                d.recalc_mm = f if f > 0 else 1e15
            n = _ints(items[1:], label)
            if n:
                if not 2 <= n[0] <= 8:
                    raise DeckError("multipol: mpole must be 2..8")
                d.mpole = n[0]
        elif label == "sor":
            n = _ints(items, label)
            if n:
                d.maxsororb = (d.maxsororb[0], n[0])
            if len(n) > 1:
                d.maxsorpot = (d.maxsorpot[0], n[1])
        elif label in ("sor4orb", "sor4pot"):
            n = _ints(items, label)
            cur = d.maxsororb if label == "sor4orb" else d.maxsorpot
            if n:
                cur = (n[1] if len(n) > 1 else cur[0], n[0])
            if label == "sor4orb":
                d.maxsororb = cur
            else:
                d.maxsorpot = cur
        elif label in ("mcsor", "mcsor-o"):
            d.orb_method = 0
        elif label == "mcsor-ce":
            d.pot_method = 0
        elif label == "order":
            name = (_word(items[0]) if items else "")[:8]  # read as character*8
            if name not in _ORDERINGS:
                raise DeckError(f"order '{name}': undefined type of ordering")
            d.ordering = _ORDERINGS[name]
        elif label == "altsweep":
            d.alt_sweeps = 1
        else:
            raise DeckError(f"card '{label}' is outside the relaxation path of this build")
    for need in ("nuclei", "config", "grid", "orbpot") + (("lcao",) if d.start == "hydrogen" else ()):
        if need not in seen:
            raise DeckError(f"the deck has no '{need}' card")
    d.nni = nnu_calc(d.nnu_card)
    d.nmu = nmu_calc(d.nmu_card, d.nni, d.r, d.rinf)
    scheme = {0: "MCSOR", 1: "SOR"}
    order = {v: k for k, v in _ORDERINGS.items()}[d.ordering]
    d.notes.append(f"orbitals by {scheme[d.orb_method]}, potentials by {scheme[d.pot_method]}, {order} ordering"
                   + (", forward/backward SOR sweeps" if d.alt_sweeps else ""))
    return d


def to_deck(spec, title="job", mcsor_threads=None):
    """The reference deck of a job specification (the reverse of parse, for
    the cards this path reads): the reference program runs it to the same
    SCF.  mcsor_threads: the thread count of the reference's mcsor card
    (orbitals by MCSOR with that many threads); None keeps the spec's
    schemes as cards.  Only the hydrogenic LCAO start is written."""
    sym = {0: "sigma", 1: "pi", 2: "delta", 3: "phi"}
    lines = [f"title {title}", "method hf", f"nuclei {spec.z1!r} {spec.z2!r} {spec.r!r}", "config 0"]
    for k in range(spec.norb):
        gu = {1: " g", -1: " u"}.get(spec.gu[k], "")
        lines.append(f"  1 {sym[spec.m[k]]}{gu}" + (" end" if k == spec.norb - 1 else ""))
    lines.append(f"grid {spec.nni} {spec.nmu} {spec.rinf!r}")
    lines.append("orbpot hydrogen" + (f" {spec.ienterm4init}" if spec.ienterm4init != 1 else ""))
    lines.append("lcao")
    for k in range(spec.norb):
        lines.append(f"  {spec.c1[k]!r} {spec.n1[k]} {spec.l1[k]} {spec.za[k]!r}   "
                     f"{spec.c2[k]!r} {spec.n2[k]} {spec.l2[k]} {spec.zb[k]!r}")
    if spec.omega_orb > 0:
        lines.append(f"omega {spec.omega_orb!r}" + (f" {spec.omega_pot!r}" if spec.omega_pot > 0 else ""))
    lines.append(f"scf {spec.maxscf} 1000 {spec.ienterm} {spec.inoterm} 3")
    if spec.nscf2skip > 0:
        lines.append(f"conv {spec.nscf2skip}" + (f" {spec.nenlast}" if spec.nenlast > 0 else "") +
                     (f" {spec.nnolast}" if spec.nnolast > 0 else ""))
    if spec.maxsororb[1] != 10 or spec.maxsorpot[1] != 10:
        lines.append(f"sor {spec.maxsororb[1]} {spec.maxsorpot[1]}")
    lines.append("order " + {v: k for k, v in _ORDERINGS.items()}[spec.ordering])
    if mcsor_threads is not None:
        lines += [f"mcsor {int(mcsor_threads)}", "mcsor-ce"]
    else:
        if spec.orb_method == 0:
            lines.append("mcsor")
        if spec.pot_method == 0:
            lines.append("mcsor-ce")
    if spec.alt_sweeps:
        lines.append("altsweeps")
    lines.append("stop")
    return "\n".join(lines) + "\n"


def load(path):
    """Parse the deck file at path."""
    with open(path) as f:
        return parse(f.read())
