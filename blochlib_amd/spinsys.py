"""Tensor-described spin systems in BlochLib's ``spins{}`` syntax.

This is the host-side mirror of ``SolidSys::read`` (src/QMspins/sys.cc:120-204) and of the
per-interaction readers it calls:

* ``numspin <n>`` / ``spins <n>``                              sys.cc:138-140
* ``T <isotope> <spin>``                                         sys.cc:141-156
* ``C <iso> <del> <eta> <spin> [alpha beta gamma]`` (Hz, deg)    csa.cc:400-440
* ``D <del> <spin1> <spin2> [alpha beta gamma]``                 dip.cc:285-323
* ``J <iso> <spin1> <spin2> [strong]``                           j.cc:171-202

Lines are split on white space like ``parse_param`` (src/utils/utils.cc:85-127).  Lines whose
first token starts with ``#`` are comments.  Errors follow BlochLib's ``BLEXCEPTION`` cases
(eta outside [0,1] csa.cc:57-64, too few fields csa.cc:423-426 / dip.cc:309-312 / j.cc:193-196,
spin index out of range) and raise :class:`BlochLibError`.

Only spin-1/2 nuclei are supported; the isotopes listed in ``SPIN_HALF`` are accepted.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SPIN_HALF = {"1H", "13C", "15N", "19F", "29Si", "31P", "3He", "89Y", "109Ag", "113Cd",
             "119Sn", "125Te", "129Xe", "183W", "195Pt", "199Hg", "205Tl", "207Pb", "e"}


class BlochLibError(RuntimeError):
    """Raised where BlochLib's BLEXCEPTION would fire (blassert.h:64-79)."""


@dataclass
class SolidSys:
    """Spin system: isotopes plus CSA / dipolar / J tables (sys.h:59-62)."""

    isotopes: list = field(default_factory=lambda: ["1H"])
    csa: list = field(default_factory=list)   # (iso, del, eta, spin, alpha, beta, gamma)
    dip: list = field(default_factory=list)   # (del, on1, on2, alpha, beta, gamma)
    jc: list = field(default_factory=list)    # (iso, on1, on2, strong)

    @property
    def spins(self) -> int:
        return len(self.isotopes)

    @property
    def N(self) -> int:
        return 1 << self.spins

    # ---------------------------------------------------------------- parsing
    @classmethod
    def from_text(cls, text: str) -> "SolidSys":
        return cls.read(text.splitlines())

    @classmethod
    def read(cls, lines) -> "SolidSys":
        numspin = 1
        names = {}
        toks = [ln.split() for ln in lines]
        toks = [t for t in toks if t and not t[0].startswith("#")]
        for t in toks:
            if t[0] in ("numspin", "spins") and len(t) >= 2:
                numspin = int(float(t[1]))
            elif t[0] in ("T", "type", "Type"):
                if len(t) < 3:
                    raise BlochLibError("'T' specifier with the wrong number of inputs: "
                                        "'Type 1H 2' ({Type flag} {nucleus} {on which spin})")
                names[int(t[2])] = t[1]
        if numspin < 1:
            raise BlochLibError("numspin must be >= 1")
        iso = ["1H"] * numspin
        for k, v in names.items():
            if not 0 <= k < numspin:
                raise BlochLibError(f"spin {k} out of range for numspin {numspin}")
            iso[k] = v
        sy = cls(isotopes=iso)
        for t in toks:
            head = t[0]
            if head in ("C", "CSA", "csa"):
                if len(t) <= 4:
                    raise BlochLibError("A CSA must have at least 4 parameters.. "
                                        "C <iso> <del> <eta> <spin> <alpha> <beta> <gamma>")
                vals = [float(x) for x in t[1:4]] + [int(float(t[4]))]
                ang = [float(x) for x in t[5:8]] + [0.0] * (3 - len(t[5:8]))
                sy.add_csa(*vals, *ang)
            elif head in ("D", "Dip", "dip", "DIP"):
                if len(t) < 4:
                    raise BlochLibError("A dipole must have at least 4 parameters.. "
                                        "D <coupling> <spin 1> <spin 2> <alpha> <beta> <gamma>")
                ang = [float(x) for x in t[4:7]] + [0.0] * (3 - len(t[4:7]))
                sy.add_dip(float(t[1]), int(float(t[2])), int(float(t[3])), *ang)
            elif head in ("J", "j", "jcop", "Jcop"):
                if len(t) <= 3:
                    raise BlochLibError("A j-coupling must have at least 3 parameters.. "
                                        "J <coupling> <spin 1> <spin 2>")
                strong = int(float(t[4])) if len(t) >= 5 else 0
                sy.add_j(float(t[1]), int(float(t[2])), int(float(t[3])), strong)
        sy.validate()
        return sy

    def add_csa(self, iso, delta, eta, spin, alpha=0.0, beta=0.0, gamma=0.0):
        if eta > 1.0 or eta < 0.0:
            raise BlochLibError(" 'eta' set to a negative or value >1...")
        self.csa.append((float(iso), float(delta), float(eta), int(spin),
                         float(alpha), float(beta), float(gamma)))

    def add_dip(self, delta, on1, on2, alpha=0.0, beta=0.0, gamma=0.0):
        self.dip.append((float(delta), int(on1), int(on2), float(alpha), float(beta), float(gamma)))

    def add_j(self, iso, on1, on2, strong=0):
        self.jc.append((float(iso), int(on1), int(on2), int(strong)))

    def validate(self):
        S = self.spins
        for name in self.isotopes:
            if name not in SPIN_HALF:
                raise BlochLibError(f"isotope '{name}' is not a spin-1/2 nucleus; only spin-1/2 "
                                    "systems are supported")
        for c in self.csa:
            if not 0 <= c[3] < S:
                raise BlochLibError(f"CSA on spin {c[3]} out of range")
        for d in self.dip:
            if not (0 <= d[1] < S and 0 <= d[2] < S) or d[1] == d[2]:
                raise BlochLibError(f"dipole ({d[1]},{d[2]}) invalid")
        for j in self.jc:
            if not (0 <= j[1] < S and 0 <= j[2] < S) or j[1] == j[2]:
                raise BlochLibError(f"J ({j[1]},{j[2]}) invalid")

    def copy(self) -> "SolidSys":
        return SolidSys(list(self.isotopes), list(self.csa), list(self.dip), list(self.jc))

    def set_spin_param(self, what: str, value: float):
        """alterSys(what, value): SolidSys::setSpinParam (src/QMspins/sys.cc:433-560) with its
        availability test (sys.cc:217-280).  'C<i>' and 'C<i>iso' set the isotropic shift of
        the CSA on spin i, 'C<i>del' / 'C<i>eta' / 'C<i>alpha' / 'C<i>beta' / 'C<i>gamma' (the
        third character decides) the other tensor values (csa.cc:46-116); 'D<i><j>' the
        dipolar coupling's del, 'J<i><j>' the J coupling's iso (dip.cc:48-57, j.cc:42-49);
        spin indices are single digits.  As the reference, a D or J change raises when the
        last coupling of its list is not the one named."""
        def avail():
            if len(what) < 2 or not what[1].isdigit():
                return False
            if what[0] == "C":
                return any(c[3] == int(what[1]) for c in self.csa)
            if what[0] in "DJ":
                if len(what) < 3 or not what[2].isdigit():
                    return False
                a, b = int(what[1]), int(what[2])
                lst = self.dip if what[0] == "D" else self.jc
                return any((x[1] == a and x[2] == b) or (x[1] == b and x[2] == a) for x in lst)
            return False

        if not avail():
            raise BlochLibError(f" Your spin parameter {what} Ain't no good")
        v = float(value)
        if what[0] == "C":
            on = int(what[1])
            field_of = {"e": 2, "d": 1, "i": 0, "a": 4, "b": 5, "g": 6}
            if len(what) >= 3:
                if what[2] not in field_of:
                    if len(what) >= 6 and what[2:5] == "sig" and what[5] in "123":
                        raise BlochLibError(f"alterSys({what}) (principal values) is not "
                                            "supported by this engine")
                    return  # the reference sets nothing
                k = field_of[what[2]]
            else:
                k = 0
            if k == 2 and (v > 1.0 or v < 0.0):
                raise BlochLibError(" 'eta' set to a negative or value >1...")
            for i, c in enumerate(self.csa):
                if c[3] == on:
                    c = list(c)
                    c[k] = v
                    self.csa[i] = tuple(c)
                    return
            return
        a, b = int(what[1]), int(what[2])
        lst = self.dip if what[0] == "D" else self.jc
        for i, x in enumerate(lst):
            if (x[1] == a and x[2] == b) or (x[1] == b and x[2] == a):
                x = list(x)
                x[0] = v
                lst[i] = tuple(x)
            elif i + 1 == len(lst):
                kind = "Dipole" if what[0] == "D" else "J"
                raise BlochLibError(f"no {kind} coupling exsits between spin {a} and {b}")

    def to_text(self) -> str:
        out = [f"numspin {self.spins}"]
        out += [f"T {n} {i}" for i, n in enumerate(self.isotopes)]
        out += ["C %.17g %.17g %.17g %d %.17g %.17g %.17g" % c for c in self.csa]
        out += ["D %.17g %d %d %.17g %.17g %.17g" % d for d in self.dip]
        out += ["J %.17g %d %d %d" % j for j in self.jc]
        return "\n".join(out) + "\n"

    # ---------------------------------------------------------------- flat tables
    def isotope_ids(self) -> np.ndarray:
        order = {}
        for n in self.isotopes:
            order.setdefault(n, len(order))
        return np.array([order[n] for n in self.isotopes], dtype=np.int32)

    def tables(self):
        """(isotope ids int32[S], csa f64[ncsa*7], dip f64[ndip*6], jc f64[nj*4])."""
        csa = np.array(self.csa, dtype=np.float64).reshape(-1, 7)
        dip = np.array(self.dip, dtype=np.float64).reshape(-1, 6)
        jc = np.array(self.jc, dtype=np.float64).reshape(-1, 4)
        return (self.isotope_ids(), np.ascontiguousarray(csa.ravel()),
                np.ascontiguousarray(dip.ravel()), np.ascontiguousarray(jc.ravel()))

    # ---------------------------------------------------------------- operators
    def op(self, expr: str) -> np.ndarray:
        """Dense N x N operator for a BlochLib operator string (``Ix``, ``Ip_0``, ``Iz_1``,
        ``Ix+Iy``, ``2*Iz_0``, ``Ix_0+Ix_1`` ...), the spin-operator subset of HamiltonianGen
        (src/QMspins/driver/ham_gen.cc:269-360).  Spin 0 is the leftmost Kronecker factor."""
        return spin_operator_expr(self.spins, expr)


_SINGLE = {
    "x": np.array([[0, 0.5], [0.5, 0]], dtype=np.complex128),
    "y": np.array([[0, -0.5j], [0.5j, 0]], dtype=np.complex128),
    "z": np.array([[0.5, 0], [0, -0.5]], dtype=np.complex128),
    "p": np.array([[0, 1], [0, 0]], dtype=np.complex128),
    "m": np.array([[0, 0], [1, 0]], dtype=np.complex128),
    "e": np.eye(2, dtype=np.complex128),
}


def single_spin_op(S: int, cart: str, spin: int) -> np.ndarray:
    """mps_expand(sps_I?, spin) (multispinop.cc:51-70)."""
    out = np.ones((1, 1), dtype=np.complex128)
    for i in range(S):
        out = np.kron(out, _SINGLE[cart] if i == spin else np.eye(2))
    return out


def total_spin_op(S: int, cart: str) -> np.ndarray:
    if cart == "e":
        return np.eye(1 << S, dtype=np.complex128)
    return sum(single_spin_op(S, cart, i) for i in range(S))


def _atom(S: int, tok: str) -> np.ndarray:
    tok = tok.strip()
    if len(tok) >= 2 and tok[0] == "I" and tok[1] in "xyzpme":
        if len(tok) == 2:
            return total_spin_op(S, tok[1])
        if len(tok) >= 4 and tok[2] == "_" and tok[3:].isdigit():
            k = int(tok[3:])
            if not 0 <= k < S:
                raise BlochLibError(f"spin operator '{tok}' names spin {k} of {S}")
            return single_spin_op(S, tok[1], k)
    raise BlochLibError(f"cannot evaluate operator string '{tok}'")


def spin_operator_expr(S: int, expr: str) -> np.ndarray:
    """Evaluate sums/differences of [coef*]op[*op...] terms."""
    e = expr.replace(" ", "")
    if not e:
        raise BlochLibError("empty operator string")
    terms, cur, depth = [], "", 0
    for ch in e:
        if ch in "+-" and cur and cur[-1] not in "eE*":
            terms.append(cur)
            cur = ch
        else:
            cur += ch
    terms.append(cur)
    N = 1 << S
    total = np.zeros((N, N), dtype=np.complex128)
    for t in terms:
        sign = 1.0
        if t[0] in "+-":
            sign = -1.0 if t[0] == "-" else 1.0
            t = t[1:]
        mat = np.eye(N, dtype=np.complex128) * sign
        for f in t.split("*"):
            try:
                mat = mat * float(f)
            except ValueError:
                mat = mat @ _atom(S, f)
        total += mat
    return total
