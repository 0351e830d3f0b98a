"""solid-2.0 input files (``.sim``) for the GPU engine: the reader and a runner of the pulse
language over the whole powder at once.

Reference behaviour restated (examples/solid-2.0/src):
  * sections ``spins{}``, ``parameters{}``, ``pulses{}`` (solid.cc:115-193); ``spins{}`` is
    SolidSys::read (src/QMspins/sys.cc:120-204), its subsections ``spin1{}``, ``spin2{}`` ...
    named systems of the global lines plus their own (MultiSystem, multisystem.cc:104-181);
  * ``parameters{}``: ``name=expr`` assignments over the ParamSet defaults (paramset.cc:44-57)
    and ``powder{}`` / ``powder1{}`` ... sections with ``aveType`` (a file or a built-in
    generator), ``thetaStep``, ``phiStep``, ``gammaStep`` (multipowder.cc:56-114); the first
    powder is the default one (multipowder.cc:140-154);
  * expressions: the Parser's grammar, functions and constants (src/utils/parser.cc);
  * ``pulses{}`` statements in SequenceRun::decide's order: assignments, ``detect(op)``,
    ``ro(op)``, ``amplitude(a)``, ``offset(o)``, ``wr=``, ``rotor=``, ``spinsys(name)``,
    ``powder(name)``, ``alterSys(what, value)``, ``loop``/``if``/``elseif``/``else`` with
    ScriptParse's resume rule, ``fid()`` and indexed ``fid(i)`` in 1-D, 2-D (``2D()``) and ptop
    (``ptop()``) modes, every ``savefid*`` form (sequencerun_funcs.cc:553-894), sub-sequence
    sections used by ``use(name[, repeat[, hold]])`` / ``reuse(...)`` (sequencerun_funcs.cc:
    465-551; SequenceParse's unrolling and readPulses' evaluation, pulsedata.cc:287-378), and
    pulse lines ``ISO:pulse(t[, phase_deg[, amp[, offset]]]) | ISO:delay(t)``
    (pulsedata.cc:166-280);
  * propagators split each pulse of length tau into ceil(tau/maxtstep) steps unless
    tau <= mintstep or wr == 0 (propogation.cc:257-300); each pulse propagates with the spin
    system named before it (pulsedata.cc:313-317, 362);
  * every orientation's rho is a device state (blfid_state_*) that the used propagators advance
    and every fid() reads; a plain fid() is the gamma-COMPUTE (or static) FID from it.
A main-section ``cycler(op)`` is a propagator-list entry that projects every orientation's rho
onto op where the list is applied (sequencerun.cc:164-178); inside a sub-sequence it has no
effect, as the reference's propagator of a used sequence never applies it.
Not supported (BlochLibError names the statement): ``reset()`` (inside the reference's
orientation loop it discards the powder sum and disables every later save), quadrupolar
interactions, a ``powder()`` switch after the program has pulsed or acquired, a powder switch
inside a sub-sequence, and a held sub-sequence whose pulses use several spin systems.  The
engine has no per-crystallite CPU path to fall back on.
"""
from __future__ import annotations

import ast
import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import engine, fidio
from .powder import Powder
from .spinsys import BlochLibError, SolidSys

DEG2RAD = math.pi / 180.0
RAD2DEG = 180.0 / math.pi
MAGIC_RAD = math.acos(1.0 / math.sqrt(3.0))

PARAM_DEFAULTS = {"wr": 0.0, "npts1D": 256.0, "npts2D": 1.0, "sw": 20000.0, "sw2": 0.0,
                  "maxtstep": 1e-6, "mintstep": 1e-7, "Bfield": 400e6, "rotor": 0.0,
                  "gammaSteps": 1.0, "roeq": "Iz", "detect": "Ip", "filesave": "soliddata",
                  "multiInt": "no"}
STRING_PARAMS = ("roeq", "detect", "filesave", "multiInt")


# ------------------------------------------------------------------------------ expressions
# The Parser of the reference (src/utils/parser.cc), restated: a tokenizer (GetNext, :301-367),
# a recursive-descent grammar that emits a postfix stack (:370-555: one level for + - comparisons
# % && ||, one for * /, one for ^, then unary signs, parentheses, operands and functions) and a
# stack machine (evaluate/Compute, :571-790).  Its quirks are kept because scripts are
# evaluated by them: + - comparisons % && || share one precedence level, left to right; * and
# / bind tighter; ^ is right-associative and binds tighter still; a unary - or ! applies to
# the operand before a following ^ (-2^2 = 4); two adjacent operator characters form one token
# whose first character decides (2*-3 = 6, 2+-3 = 5); functions take one argument; x, y, z
# and t are the Parser's internal variables and always read 0; an evaluation error (0/0, ln or
# log of a non-positive value, ...) stops the machine and the value is the bottom of its stack.
# An expression the grammar rejects raises BlochLibError (the reference evaluates it anyway).
_SYMBOL = set("+-*/%^=()<>&|!")
_MATHSYM = set("+-*/%^=<>&|!")
_MATFUNCS = {"abs", "acos", "asin", "atan", "cos", "exp", "floor", "frac", "int", "ln", "log",
             "pow", "round", "ceil", "sin", "sqrt", "tan", "trunc", "sinh", "cosh", "tanh",
             "CHSSGN", "FAC"}
_CONSTS = {"e": math.e, "pi": math.pi, "Pi": math.pi, "PI": math.pi, "deg2rad": DEG2RAD,
           "rad2deg": RAD2DEG}
_INTERNAL = ("x", "y", "z", "t")
_LOGLIM, _LNLIM = 1.0e-10, 1.0e-36
_singular = [0]  # Compute's function-static SINGULAR flag (parser.cc:574), process-wide


def _cint(v: float) -> int:
    """C's (int) conversion of a double: truncation toward zero."""
    if v != v or v in (math.inf, -math.inf):
        return 0
    return int(v)


def _ctrunc_mod(a: int, b: int) -> int:
    if b == 0:
        raise BlochLibError("integer modulo by zero in expression")
    return int(math.fmod(a, b))


class _Parser:
    def __init__(self, expr: str, known):
        self.s = expr
        self.n = 0
        self.known = known
        self.chunk = ""
        self.group = ""
        self.stack = []   # postfix: ("N", value) | ("V", name) | ("F", name)
        self.error = False

    def c0(self):
        return self.chunk[0] if self.chunk else "\0"

    def next(self):
        s, n = self.s, self.n
        self.chunk = ""
        if n >= len(s):
            return
        while n < len(s) and s[n].isspace():
            n += 1
        if n >= len(s):
            self.n = n
            return
        ch = s[n]
        if ch in _SYMBOL:
            self.group = "SYMBOL"
            self.chunk = ch
            n += 1
            if n < len(s) and s[n - 1] in _MATHSYM and s[n] in _MATHSYM:
                self.chunk += s[n]
                n += 1
        elif ch.isalpha():
            while n < len(s) and s[n] not in _SYMBOL and (n == 0 or s[n - 1] != "]"):
                if not s[n].isspace():
                    self.chunk += s[n]
                n += 1
            if self.chunk in _MATFUNCS:
                self.group = "MATFUNC"
            elif self.chunk in _INTERNAL or self.known(self.chunk):
                self.group = "VARIABLE"
            else:
                self.error = True
                raise BlochLibError(f"Parser::Prase Error...thing \"{self.chunk}\" for "
                                    f"\"{s}\" not found")
        elif ch.isdigit() or ch == ".":
            while n < len(s):
                if s[n].isdigit() or s[n] == ".":
                    self.chunk += s[n]
                    self.group = "NUMBER"
                    n += 1
                elif s[n] == "e" and n + 1 < len(s) and (s[n + 1] in "+-" or s[n + 1].isdigit()):
                    self.chunk += s[n:n + 2]
                    self.group = "NUMBER"
                    n += 2
                else:
                    break
        self.n = n

    def parse(self):
        self.next()
        if self.chunk:
            self.sum_level()
        else:
            self.error = True
        return self

    def sum_level(self):
        self.product_level()
        sign = self.c0()
        sec = self.chunk[1] if len(self.chunk) > 1 else "\0"
        while sign in "+-<>%&|!=" and sign != "\0":
            self.next()
            self.product_level()
            op = {("+",): "PLUS", ("-",): "MINUS"}.get((sign,))
            if op is None:
                op = {(">", "\0"): "GREATER", ("<", "\0"): "LESS", ("=", "="): "EQEQ",
                      ("!", "="): "NEEQ", (">", "="): "GREATEREQ", ("<", "="): "LESSEQ",
                      ("&", "&"): "ANDAND", ("|", "|"): "OROR"}.get((sign, sec))
            if op is None and sign == "%":
                op = "MOD"
            self.stack.append(("F", op or ""))
            sign = self.c0()
            sec = self.chunk[1] if len(self.chunk) > 1 else "\0"

    def product_level(self):
        self.power_level()
        sign = self.c0()
        while sign in "*/" and sign != "\0":
            self.next()
            self.power_level()
            self.stack.append(("F", "MULT" if sign == "*" else "DIV"))
            sign = self.c0()

    def power_level(self):
        self.unary()
        if self.c0() == "^":
            self.next()
            self.power_level()
            self.stack.append(("F", "pow"))

    def unary(self):
        sign = " "
        if (self.group == "SYMBOL" and self.c0() == "+") or self.c0() in "-!" and self.chunk:
            sign = self.c0()
            self.next()
        self.parenthesised()
        if sign == "-":
            self.stack.append(("F", "CHSSGN"))
        if sign == "!":
            self.stack.append(("F", "NOT"))

    def parenthesised(self):
        if self.c0() == "(" and self.group == "SYMBOL":
            self.next()
            self.sum_level()
            if self.c0() != ")":
                self.error = True
            self.next()
        else:
            self.operand()

    def operand(self):
        if self.group == "NUMBER" and self.chunk:
            self.stack.append(("N", _atof(self.chunk)))
            self.next()
        elif self.group == "VARIABLE" and self.chunk:
            self.stack.append(("V", self.chunk))
            self.next()
        else:
            self.function()

    def function(self):
        if self.group == "MATFUNC" and self.chunk:
            name = self.chunk
            self.next()
            if self.c0() == "(":
                self.next()
                self.sum_level()
                if self.c0() != ")":
                    self.error = True
                self.stack.append(("F", name))
                self.next()
        else:
            self.error = True


def _atof(txt: str) -> float:
    """C atof: the longest numeric prefix."""
    m = re.match(r"[0-9]*\.?[0-9]*(?:[eE][+-]?[0-9]+)?", txt)
    try:
        return float(m.group(0)) if m and m.group(0) not in ("", ".") else 0.0
    except ValueError:
        return 0.0


def _compute(mf: str, st: list, tos: list) -> bool:
    """One Compute step on the stack machine; returns True on an evaluation error."""
    def pop():
        tos[0] -= 1
        return st[tos[0]] if tos[0] >= 0 else 0.0

    def push(v):
        if tos[0] >= len(st):
            st.extend([0.0] * len(st))
        st[tos[0]] = float(v)
        tos[0] += 1

    binary = {"PLUS": lambda a, b: a + b, "MINUS": lambda a, b: a - b,
              "GREATER": lambda a, b: float(a > b), "LESS": lambda a, b: float(a < b),
              "EQEQ": lambda a, b: float(a == b), "NEEQ": lambda a, b: float(a != b),
              "GREATEREQ": lambda a, b: float(a >= b), "LESSEQ": lambda a, b: float(a <= b),
              "ANDAND": lambda a, b: float(bool(_cint(a)) and bool(_cint(b))),
              "OROR": lambda a, b: float(bool(_cint(a)) or bool(_cint(b))),
              "MOD": lambda a, b: float(_ctrunc_mod(_cint(a), _cint(b)))}
    if mf in binary:
        t1 = pop()
        t2 = pop()
        push(binary[mf](t2, t1))
        return False
    if mf == "MULT":
        t1 = pop()
        t2 = pop()
        if abs(t1) > 10e10 or abs(t2) > 10e10:
            push((math.sqrt(t2) * math.sqrt(t1)) ** 2 if t1 >= 0 and t2 >= 0 else math.nan)
            if _singular[0]:
                return True
            _singular[0] = 0
        else:
            push(t2 * t1)
        return False
    if mf == "DIV":
        t1 = pop()
        t2 = pop()
        if abs(t1) < _LOGLIM and abs(t2) < _LOGLIM:
            push(0.0)
            _singular[0] = 1
            return True
        if abs(t1) < _LNLIM:
            push(0.0)
            return True
        push(t2 / t1)
        return False
    if mf == "pow":
        t1 = pop()
        t2 = pop()
        try:
            v = math.pow(t2, t1)
        except (ValueError, OverflowError, ZeroDivisionError):
            v = math.nan if t2 < 0 else math.inf
        push(v)
        return False
    if mf in ("CHSSGN", "NOT", "abs", "acos", "asin", "atan", "cos", "exp", "frac", "int",
              "ln", "log", "round", "ceil", "sin", "sqrt", "tan", "trunc", "floor", "sinh",
              "cosh", "tanh"):
        t1 = pop()
        if mf == "CHSSGN":
            push(-t1)
        elif mf == "NOT":
            push(float(t1 == 0.0))
        elif mf == "abs":
            push(abs(t1))
        elif mf in ("acos", "asin"):
            if abs(t1) > 1:
                push(0.0)
                return True
            push(math.acos(t1) if mf == "acos" else math.asin(t1))
        elif mf == "ln":
            if t1 > _LNLIM:
                push(math.log(t1))
            else:
                return True
        elif mf == "log":
            if t1 > _LOGLIM:
                push(math.log10(t1))
            else:
                return True
        elif mf == "sqrt":
            if t1 < 0.0:
                push(0.0)
                return True
            push(math.sqrt(t1))
        elif mf == "tan":
            if abs(math.cos(t1)) < _LNLIM:
                push(0.0)
                return True
            push(math.tan(t1))
        elif mf == "frac":
            push(t1 - _cint(t1))
        elif mf in ("int", "trunc"):
            push(float(_cint(t1)))
        elif mf == "round":
            push(float(_cint(t1 + 0.5)) if t1 >= 0.0 else float(_cint(t1 - 0.5)))
        else:
            fn = {"atan": math.atan, "cos": math.cos, "exp": math.exp, "ceil": math.ceil,
                  "sin": math.sin, "floor": math.floor, "sinh": math.sinh, "cosh": math.cosh,
                  "tanh": math.tanh}[mf]
            try:
                push(fn(t1))
            except OverflowError:
                push(math.inf)
        return False
    return False   # an unknown operator (FAC, or none) changes nothing


def evaluate(expr: str, vars: dict) -> float:
    """Value of a Parser expression; names resolve to numeric entries of `vars`, then to the
    Parser's constants (e, pi, Pi, PI, deg2rad, rad2deg)."""
    def known(name):
        v = vars.get(name)
        return (v is not None and not isinstance(v, str)) or name in _CONSTS

    p = _Parser(expr, known).parse()
    if p.error:
        # the reference's Parser flags the error and its callers read whatever the machine
        # holds; an unparsable expression is an input error here
        raise BlochLibError(f"cannot evaluate expression '{expr}'")
    st = [0.0] * 50
    tos = [0]
    for kind, val in p.stack:
        if kind == "N":
            _compute_push(st, tos, val)
        elif kind == "V":
            if val in _INTERNAL:
                v = 0.0
            else:
                v = vars.get(val)
                v = float(v) if v is not None and not isinstance(v, str) else _CONSTS.get(val, 0.0)
            _compute_push(st, tos, v)
        elif _compute(val, st, tos):
            break
    return float(st[0])


def _compute_push(st, tos, v):
    if tos[0] >= len(st):
        st.extend([0.0] * len(st))
    st[max(tos[0], 0)] = float(v)
    tos[0] = max(tos[0], 0) + 1


# ------------------------------------------------------------------------------ sections
@dataclass
class Section:
    name: str
    items: list = field(default_factory=list)  # str lines and nested Sections

    def sub(self, name):
        for it in self.items:
            if isinstance(it, Section) and it.name == name:
                return it
        return None

    def lines(self):
        return [it for it in self.items if isinstance(it, str)]


def parse_sections(text: str) -> Section:
    root = Section("")
    stack = [root]
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        while line:
            if "{" in line and ("}" not in line or line.index("{") < line.index("}")):
                head, line = line.split("{", 1)
                if head.strip():
                    pre, _, name = head.strip().rpartition(" ")
                    if pre.strip():
                        stack[-1].items.append(pre.strip())
                    s = Section(name.strip())
                else:
                    s = Section("")
                stack[-1].items.append(s)
                stack.append(s)
                line = line.strip()
            elif "}" in line:
                head, line = line.split("}", 1)
                if head.strip():
                    stack[-1].items.append(head.strip())
                if len(stack) == 1:
                    raise BlochLibError("unbalanced '}' in input file")
                stack.pop()
                line = line.strip()
            else:
                stack[-1].items.append(line)
                line = ""
    if len(stack) != 1:
        raise BlochLibError("unbalanced '{' in input file")
    return root


def _inside(s: str) -> str:
    a, b = s.find("("), s.rfind(")")
    return s[a + 1:b] if a >= 0 and b > a else ""


def _split_args(s: str):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
            continue
        depth += ch == "("
        depth -= ch == ")"
        cur += ch
    if cur.strip():
        out.append(cur)
    return [x.strip() for x in out]


_ASSIGN = re.compile(r"^([A-Za-z_][A-Za-z0-9_]*)\s*=\s*(.+)$")


@dataclass
class Segment:
    """One pulse line: its time span and per-channel (isotope, amp Hz, phase rad, offset Hz),
    and the spin system it propagates with (a spinsys() name, 'default' the first)."""
    t1: float
    t2: float
    rotor_rad: float
    channels: list
    sys: str = "default"


# ------------------------------------------------------------------------------ program
class SimFile:
    """A parsed solid-2.0 input file."""

    def __init__(self, text: str, base_dir: str = "."):
        self.base_dir = base_dir
        root = parse_sections(text)
        spins = [s for s in root.items if isinstance(s, Section) and s.name.startswith("spins")]
        if not spins:
            raise BlochLibError("no spins{} section")
        if len(spins) > 1:
            raise BlochLibError("one spins{} section (several systems are its spin1{}, spin2{} "
                                "... subsections)")
        self.systems = self._systems(spins[0])
        self.system = self.systems[min(self.systems)]
        self.params = dict(PARAM_DEFAULTS)
        self.powders = {}
        self.powder = None
        par = root.sub("parameters")
        if par is not None:
            for ln in par.lines():
                self._param_line(ln)
            self.powders = self._powders(par)
            if self.powders:
                self.powder = self.powders[min(self.powders)]
        self.pulses = root.sub("pulses")
        if self.pulses is None:
            raise BlochLibError("no pulses{} section")

    @classmethod
    def read(cls, path: str) -> "SimFile":
        with open(path) as f:
            return cls(f.read(), os.path.dirname(os.path.abspath(path)))

    def _param_line(self, ln: str):
        m = _ASSIGN.match(ln)
        if m:
            name, val = m.group(1), m.group(2).strip()
        else:
            tok = ln.split(None, 1)
            if len(tok) != 2:
                return
            name, val = tok[0], tok[1].strip()
        self.set_param(name, val)

    def set_param(self, name: str, val: str):
        if name in STRING_PARAMS:
            self.params[name] = val
            return
        try:
            self.params[name] = evaluate(val, self.params)
        except BlochLibError:
            self.params[name] = val

    @staticmethod
    def _numbered(sec: Section, base: str, limit_key: str) -> list:
        """The subsections base1, base2, ... present in order (counted while the next one
        exists, up to the section's own `limit_key` value, 100 by default), as MultiSystem::
        parse and MultiPowder::parse count them (multisystem.cc:104-132,
        multipowder.cc:56-87)."""
        limit = 100
        for ln in sec.lines():
            tok = ln.replace("=", " ").split()
            if len(tok) >= 2 and tok[0] == limit_key:
                try:
                    limit = int(float(tok[1]))
                except ValueError:
                    pass
        out = []
        while sec.sub(f"{base}{len(out) + 1}") is not None and len(out) <= limit:
            out.append(sec.sub(f"{base}{len(out) + 1}"))
        return out

    def _systems(self, spins: Section) -> dict:
        """MultiSystem (multisystem.cc:104-152): the spins{} lines outside subsections are the
        global part; each subsection spin1{}, spin2{} ... is a system read from the global
        lines followed by its own; without subsections the global lines are the one system
        'spin1'.  Every system must keep the Hilbert dimension.  'default' is the first name
        in order (the std::map's first), spin1."""
        glob = spins.lines()
        subs = self._numbered(spins, "spin", "numsys")
        out = {}
        if not subs:
            out["spin1"] = SolidSys.from_text("\n".join(glob))
        for i, sec in enumerate(subs):
            out[f"spin{i + 1}"] = SolidSys.from_text("\n".join(glob + _flatten(sec)))
        dims = {sy.N for sy in out.values()}
        if len(dims) > 1:
            raise BlochLibError("the spin systems of spins{} differ in Hilbert dimension")
        return out

    def _powders(self, par: Section) -> dict:
        """MultiPowder::parse (multipowder.cc:56-87): subsections powder1{}, powder2{} ... or
        one powder{}."""
        subs = self._numbered(par, "powder", "numpows")
        if subs:
            return {f"powder{i + 1}": self._powder(sec) for i, sec in enumerate(subs)}
        one = par.sub("powder")
        return {"powder": self._powder(one)} if one is not None else {}

    def _powder(self, sec: Section) -> Powder:
        kv = {}
        for ln in sec.lines():
            tok = ln.replace("=", " ").split(None, 1)
            if len(tok) == 2:
                kv[tok[0]] = tok[1].strip()
        ave = kv.get("aveType")
        if ave is None:
            raise BlochLibError("powder section without aveType")
        th = int(evaluate(kv.get("thetaStep", "0"), self.params))
        ph = int(evaluate(kv.get("phiStep", "0"), self.params))
        ga = int(evaluate(kv.get("gammaStep", "0"), self.params))
        if ave in ("zcw", "zcw2d", "zcw2D", "alderman", "sphere", "rect", "sophie", "liquid"):
            return Powder.generate(ave, th, ph, ga)
        path = ave if os.path.isabs(ave) else os.path.join(self.base_dir, ave)
        return Powder.read(path)

    # -------------------------------------------------------------------------- running
    def run(self, outdir: str = ".", precision: str = "fp64", verbose: bool = False) -> dict:
        """Execute the pulses{} program; returns {"fid": complex FID, "sw": sw, "files": [...]}"""
        return _Runner(self, outdir, precision, verbose).run()


def _flatten(sec: Section):
    out = []
    for it in sec.items:
        out.extend(_flatten(it) if isinstance(it, Section) else [it])
    return out


def _nows(s: str) -> str:
    return re.sub(r"\s", "", s)


def _is_assignment(s: str) -> bool:
    """The reference's assignment test (sequencerun.cc:302-304): an '=' that does not start an
    '==' and is not preceded by '!'."""
    i = s.find("=")
    return i >= 0 and s.find("==") != i and s.find("!=") != i - 1


# ------------------------------------------------------------------------------ control flow
def run_script(lines, decide, vars: dict):
    """ScriptParse::parse (src/utils/scriptparse.cc:412-435): loop(v=a:b) / loop(v=a:step:b)
    ... end and if(c) ... [elseif(c) ...] [else ...] end around `decide(line)` for every other
    line.  Loop bounds are evaluated once; the loop variable is read back after every pass
    (scriptparse.cc:195-201).  The if structure is restated with the reference's resume index
    (scriptparse.cc:392-396): parseIfs returns its start plus the bodies' line count, marker
    lines not counted, so with elseif/else branches the caller resumes that many lines early
    and passes the tail of the structure to `decide` again."""
    i = 0
    while i < len(lines):
        w = _nows(lines[i])
        if w.startswith("loop("):
            i = _parse_loop(lines, i, decide, vars)
        elif w.startswith("if("):
            i = _parse_if(lines, i, decide, vars)
        else:
            decide(lines[i])
        i += 1


def _parse_loop(p, i, decide, vars):
    """ScriptParse::parseLoops (scriptparse.cc:111-209); returns the index of the loop's end."""
    arg = _inside(_nows(p[i]))
    parts = arg.split(":")
    if len(parts) not in (2, 3) or "=" not in parts[0]:
        raise BlochLibError(f" Bad 'loop' usage for \"{p[i]}\" should be \"loop(var=begin:end)\" "
                            "or \"loop(var=begin:step:end)\", ")
    lhs, rhs = parts[0].split("=", 1)
    if len(parts) == 2:
        step, end = 1.0, evaluate(parts[1], vars)
    else:
        step, end = evaluate(parts[1], vars), evaluate(parts[2], vars)
    vars[lhs] = evaluate(rhs, vars)
    body, numl, stop = [], 0, None
    for k in range(i + 1, len(p)):
        w = _nows(p[k])
        if w.startswith("loop(") or w.startswith("if("):
            numl += 1
        if w == "end":
            if numl <= 0:
                stop = k
                break
            numl -= 1
        body.append(p[k])
    if stop is None:
        raise BlochLibError(" bad syntax (section ended before loop 'end' was found)")
    while float(vars[lhs]) <= end:
        run_script(body, decide, vars)
        vars[lhs] = float(vars[lhs]) + step
    return stop


def _parse_if(p, start, decide, vars):
    """ScriptParse::parseIfs (scriptparse.cc:213-409), including its resume index."""
    ifin, elin, elifs, eltests = [], [], [], []
    iftest, gotif = 0, False
    i = start
    arg = ""
    while i < len(p):
        arg = _nows(p[i])
        gotelif = False
        numl = 0
        if arg.startswith("if") and not arg.startswith("if("):
            raise BlochLibError(f" Bad 'if' Syntax requires an argument\n at line \"{p[i]}\"")
        if arg.startswith("if("):
            cond = _inside(arg)
            if not cond:
                raise BlochLibError(" 'if' requires an argument")
            iftest = int(evaluate(cond, vars))
            gotif = True
            numl = 0
            for k in range(i + 1, len(p)):
                arg = _nows(p[k])
                if arg.startswith("loop(") or arg.startswith("if("):
                    numl += 1
                if arg == "end":
                    numl -= 1
                if numl <= 0 and arg.startswith("else"):
                    i = k - 1
                    break
                if numl == -1 and arg.startswith("end"):
                    i = k
                    break
                ifin.append(p[k])
            if arg.startswith("end") and numl <= 0:
                break
        if i < len(p):
            arg = _nows(p[i])
        if arg.startswith("elseif"):
            if not arg.startswith("elseif("):
                raise BlochLibError(f" Bad 'else if' Syntax requires an argument\n at line "
                                    f"\"{p[i]}\"")
            cond = _inside(arg)
            if not cond:
                raise BlochLibError(" 'else if' requires an argument")
            eltests.append(int(evaluate(cond, vars)))
            gotelif = True
            numl = 0
            tml = []
            for k in range(i + 1, len(p)):
                arg = _nows(p[k])
                if arg.startswith("loop(") or arg.startswith("if("):
                    numl += 1
                if arg == "end":
                    numl -= 1
                if numl <= 0 and arg.startswith("else"):
                    elifs.append(tml)
                    i = k - 1
                    break
                if numl == -1 and arg.startswith("end"):
                    elifs.append(tml)
                    i = k
                    break
                tml.append(p[k])
            if arg.startswith("end") and numl <= 0:
                break
        if i < len(p):
            arg = _nows(p[i])
        if arg.startswith("else") and not gotelif:
            numl = 0
            for k in range(i + 1, len(p)):
                arg = _nows(p[k])
                if arg.startswith("loop(") or arg.startswith("if("):
                    numl += 1
                if arg == "end":
                    numl -= 1
                if numl == -1 and arg.startswith("end"):
                    i = k
                    break
                elin.append(p[k])
            break
        i += 1
    retend = len(ifin) + start + sum(len(x) for x in elifs) + len(elin)
    did = False
    if iftest != 0:
        run_script(ifin, decide, vars)
        did = True
    if not did:
        for body, t in zip(elifs, eltests):
            if t != 0:
                run_script(body, decide, vars)
                did = True
                break
    if not did and elin:
        run_script(elin, decide, vars)
    return retend if gotif else start


# ------------------------------------------------------------------------------ runner
@dataclass
class _Prop:
    """One used propagator (a PropList entry, sequencerun.h:102-150): its maxtstep-split steps
    (t1, dt, rotor_rad, channels), the spin system and spinning speed it was computed with,
    and how a fid() applies it: `repeat` times, and with `hold` only at the first point."""
    steps: list
    sys: SolidSys
    wr: float
    repeat: int = 1
    hold: bool = False
    # a cycler(op) entry instead: no steps, rho <- Tr(rho op^H) op / Tr(op op^H)
    proj: np.ndarray | None = field(default=None, compare=False)


def _rf_tables(st):
    """rf channel tables (isotope, amp[n], phase_rad[n], offset[n]) of a step list."""
    n = len(st)
    chans = {}
    for k, (_, _, _, cl) in enumerate(st):
        seen = {}
        for iso, a, ph, o in cl:
            j = seen.get(iso, 0)
            seen[iso] = j + 1
            key = (iso, j)
            if key not in chans:
                chans[key] = (np.zeros(n), np.zeros(n), np.zeros(n))
            chans[key][0][k] = a
            chans[key][1][k] = ph
            chans[key][2][k] = o
    return [(iso, a, ph, o) for (iso, _), (a, ph, o) in chans.items()]


class _Runner:
    """SequenceRun (examples/solid-2.0/src/sequencerun.cc, sequencerun_funcs.cc) over the whole
    powder at once.  The reference runs the pulses{} program once per powder orientation
    (sequencerun.cc:512-531); its control flow does not depend on the orientation, so the
    runner runs it once, and every fid() computes all orientations together on the GPU.  The
    density operator of every orientation (the reference's `ro`) is a device State: ro() resets
    it, and the propagators a fid() applies are queued and applied on the device when a later
    computation needs the propagated rho."""

    def __init__(self, sim: SimFile, outdir: str, precision: str, verbose: bool):
        self.sim = sim
        self.p = sim.params
        self.outdir = outdir
        self.precision = precision
        self.verbose = verbose
        self.subs = {it.name: it for it in sim.pulses.items if isinstance(it, Section)}
        # the named spin systems (alterSys changes the current one in place) and the current
        # system and powder names (sequencerun.cc:229 'default', spinsys() / powder())
        self.systems = {k: v.copy() for k, v in sim.systems.items()}
        self.usesys = "default"
        self.usepow = "default"
        self.sub_cache = {}      # sub-sequence -> [segments of its last parse, _Prop or None]
        self.ptop = False
        self.is2D = False
        self.pt2D = -1
        self.to_run = True       # show() turns the run into a listing (sequencerun_funcs.cc:306)
        self.proplist = []
        self.current_time = 0.0
        self.amp = 0.0
        self.offset = 0.0
        npts = int(self.p["npts1D"])
        self.fid1 = np.zeros(npts, dtype=np.complex128)
        self.fid2 = np.zeros((max(int(self.p["npts2D"]), 0), npts), dtype=np.complex128)
        self.roeq = self.sys.op(str(self.p["roeq"]))
        # rho of every orientation = rho_op propagated by ops[] in order (ops[:synced] are
        # already applied to the device state when has_base)
        self.rho_op = self.roeq
        self.ops = []
        self.state = None
        self.has_base = False
        self.synced = 0
        self.files = []
        self.not_saved = True

    def log(self, *a):
        if self.verbose:
            print("[sim]", *a, flush=True)

    def system(self, use: str) -> SolidSys:
        """MultiSystem::getSystem (multisystem.cc:165-181): 'default' is the first system."""
        if use == "default":
            return self.systems[min(self.systems)]
        if use not in self.systems:
            raise BlochLibError(f"SoildSys '{use}' does not exsist")
        return self.systems[use]

    @property
    def sys(self) -> SolidSys:
        return self.system(self.usesys)

    def run(self):
        lines = [it for it in self.sim.pulses.items if isinstance(it, str)]
        try:
            run_script(lines, self.statement, self.p)
            if self.not_saved:
                self.save("savefid()")   # sequencerun.cc:533-536
        finally:
            if self.state is not None:
                self.state.close()
                self.state = None
        return {"fid": self.fid1.copy(), "fid2d": self.fid2.copy() if self.is2D else None,
                "sw": float(self.p["sw"]), "files": self.files}

    # ---------------------------------------------------------------- statements
    def statement(self, s: str):
        """SequenceRun::decide (sequencerun.cc:272-336), in its order."""
        w = _nows(s)
        if not w:
            return
        if "pulse(" in s or "delay(" in s:
            self.pulse_line(w)
        elif "amplitude(" in s:
            inner = _inside(w)
            if not inner:
                raise BlochLibError(f" Bad 'amplitude' usage for \"{s}\"")
            self.amp = evaluate(_split_args(inner)[0], self.p)
        elif "offset(" in s:
            inner = _inside(w)
            if not inner:
                raise BlochLibError(f" Bad 'offset' usage for \"{s}\"")
            self.offset = evaluate(_split_args(inner)[0], self.p)
        elif "spinsys(" in s:
            # setSystem (sequencerun_funcs.cc:179-193): later pulses and fid() use it
            if not _inside(w):
                raise BlochLibError(f" Bad 'spinsys' usage for \"{s}\"\n should be "
                                    "\"spinsys(sys)\" ")
            self.system(_inside(w))
            self.usesys = _inside(w)
        elif "powder(" in s:
            self.set_powder(w, s)
        elif "cycler(" in s:
            # doCycler (sequencerun_funcs.cc:232-256): a propagator-list entry that projects
            # ro onto the operator where the list is applied (sequencerun.cc:164-178)
            inner = _inside(w)
            if not inner:
                raise BlochLibError(f" Bad 'cycler' usage for \"{s}\"\n should be "
                                    "\"cycler(SpinOp)\" ")
            if self.to_run and inner not in ("Fe", "Ie"):
                self.proplist.append(_Prop([], self.sys.copy(), float(self.p["wr"]),
                                           proj=self.sys.op(inner)))
        elif "dumpState()" in s:
            return
        elif _is_assignment(s):
            m = _ASSIGN.match(s.strip())
            if not m:
                raise BlochLibError(f" Bad 'variable' usage for \"{s}\"")
            self.sim.set_param(m.group(1), m.group(2))
        elif "ptop()" in s:
            self.ptop = True
        elif "show()" in s:
            self.to_run = False
        elif "2D()" in s:
            self.is2D = True
            if self.pt2D == -1:
                self.pt2D = 0
        elif "ro(" in s:
            inner = _inside(w)
            if inner:
                if inner == "roeq":
                    inner = str(self.p["roeq"])
                self.set_rho(self.sys.op(inner))
            else:
                self.set_rho(self.roeq)
        elif "detect(" in s:
            if _inside(w):
                self.p["detect"] = _inside(w)
        elif "U(" in s:
            return
        elif "use(" in s:
            self.use(w)
        elif "reset(" in s:
            # ResetPars (sequencerun_funcs.cc:349-362) runs inside the orientation loop: it
            # zeroes the accumulated powder FID and clears canSave, so the reference keeps one
            # orientation's data and saves nothing afterwards
            raise BlochLibError("reset() is not supported by this engine (the reference's "
                                "reset() discards the powder sum and disables later saves)")
        elif "alterSys(" in s:
            args = _split_args(_inside(w))
            if not _inside(w) or len(args) < 2:
                raise BlochLibError(f" Bad 'use' usage for \"{s}\"\n should be "
                                    "\"alterSys(what, number)\" ")
            if self.to_run:
                self.sys.set_spin_param(args[0], evaluate(args[1], self.p))
        elif any(c in s for c in ("savefid(", "savefidmatlab(", "savefidtext(",
                                  "savefidbinary(")):
            self.save(w)
        elif "fid(" in s:
            self.do_fid(w)
        # anything else is SequenceParse::decideSeP's business (sequenceparse.cc:431-467): in
        # the main section it changes nothing

    def segment(self, w: str, t0: float, rotor_rad: float, amp: float, off: float) -> Segment:
        """PulseData::read (pulsedata.cc:165-278): channels split by '|', 'iso:pulse(t[, phase
        deg[, amp[, offset]]])' or 'iso:delay(t)'; the last channel sets the length."""
        chans = []
        dur = 0.0
        for part in w.split("|"):
            if part.count(":") != 1:
                raise BlochLibError(f"bad pulse line '{w}'")
            iso, act = part.split(":", 1)
            args = _split_args(_inside(act))
            if "pulse(" in act:
                if not args:
                    raise BlochLibError(f"bad pulse line '{w}'")
                dt = evaluate(args[0], self.p)
                ph = evaluate(args[1], self.p) * DEG2RAD if len(args) >= 2 else 0.0
                a = evaluate(args[2], self.p) if len(args) >= 3 else amp
                o = evaluate(args[3], self.p) if len(args) >= 4 else off
                chans.append((iso, a, ph, o))
            elif "delay(" in act:
                if not args:
                    raise BlochLibError(f"bad delay line '{w}'")
                dt = evaluate(args[0], self.p)
                chans.append((iso, 0.0, 0.0, 0.0))
            else:
                raise BlochLibError(f"bad pulse line '{w}'")
            dur = dt
        return Segment(t0, t0 + dur, rotor_rad, chans)

    def split_steps(self, segs) -> list:
        """Propogation::propogator's step split (propogation.cc:257-300): a pulse of length tau
        becomes ceil(tau / maxtstep) equal steps unless tau <= mintstep or wr == 0."""
        wr = float(self.p["wr"])
        maxt, mint = float(self.p["maxtstep"]), float(self.p["mintstep"])
        out = []
        for sg in segs:
            tau = sg.t2 - sg.t1
            if tau <= 0:
                continue
            m = int(math.ceil(tau / maxt)) if maxt > 0 else 1
            if tau <= mint or m < 1 or wr == 0:
                m = 1
            d = tau / m
            for j in range(m):
                out.append((sg.t1 + j * d, d, sg.rotor_rad, sg.channels))
        return out

    def pulse_line(self, w: str):
        """doPulseData (sequencerun_funcs.cc:101-134): a main-section pulse, timed from the
        current time, its propagator used once."""
        seg = self.segment(w, self.current_time, float(self.p["rotor"]) * DEG2RAD, self.amp,
                           self.offset)
        if self.to_run:
            self.proplist.append(_Prop(self.split_steps([seg]), self.sys.copy(),
                                       float(self.p["wr"])))
        self.current_time = seg.t2

    def expand_sub(self, sec: Section) -> list:
        """A sub-sequence's pulse list: SequenceParse unrolls its loops and ifs and records the
        remaining lines (sequenceparse.cc:386-430, addLine :116-128 turns 'rotor=x' into
        'rotor(x)'), then readPulses evaluates the recorded lines in order (pulsedata.cc:
        287-378): rotor() in degrees (default: the magic angle), amplitude(), offset(),
        assignments, and pulse/delay lines timed from 0.  Variables are read when a line is
        evaluated, so a loop variable used inside a loop's pulses reads its final value, as
        in the reference."""
        if any(isinstance(it, Section) for it in sec.items):
            raise BlochLibError("nested sections inside a sub-sequence")
        rec = []

        def record(line):
            w = _nows(line)
            if any(k in line for k in ("pulse(", "delay(", "amplitude(", "offset(", "rotor=",
                                       "wr=", "rotor(", "spinsys(", "powder(", "U(",
                                       "cycler(")) or (
                    "=" in line and line.find("==") != line.find("=")
                    and line.find("!=") != line.find("=") + 1):
                rec.append("rotor(" + w.split("=", 1)[1] + ")" if w.startswith("rotor=") else w)

        run_script(sec.lines(), record, self.p)
        segs = []
        t, rot, amp, off, use = 0.0, MAGIC_RAD, 0.0, 0.0, "default"
        for w in rec:
            if "rotor(" in w:
                rot = evaluate(_inside(w), self.p) * DEG2RAD
            elif "spinsys(" in w:
                use = _inside(w) or "default"     # readPulses (pulsedata.cc:313-317)
                self.system(use)
            elif "powder(" in w:
                if (_inside(w) or "default") not in ("default", self.usepow):
                    raise BlochLibError("a powder() switch inside a sub-sequence (orientations "
                                        "of another powder per pulse) is not supported")
            elif "amplitude(" in w:
                amp = evaluate(_inside(w), self.p)
            elif "offset(" in w:
                off = evaluate(_inside(w), self.p)
            elif "cycler(" in w:
                # recorded on the next pulse (pulsedata.cc:337-340, 364), but the propagator
                # of a used sequence (Propogation::propogator) never applies it: no effect
                continue
            elif "=" in w and w.find("==") != w.find("=") and w.find("!=") != w.find("=") + 1:
                lhs, rhs = w.split("=", 1)
                self.p[lhs] = evaluate(rhs, self.p)
            elif "pulse(" in w or "delay(" in w:
                sg = self.segment(w, t, rot, amp, off)
                sg.sys = use
                segs.append(sg)
                t = sg.t2
        return segs

    def seq_props(self, segs) -> list:
        """A sub-sequence's propagators: one per run of consecutive pulses on the same spin
        system, each with that system (alterSys changes applied so far included)."""
        out = []
        for sg in segs:
            if out and out[-1][0] == sg.sys:
                out[-1][1].append(sg)
            else:
                out.append((sg.sys, [sg]))
        return [_Prop(self.split_steps(g), self.system(u).copy(), float(self.p["wr"]))
                for u, g in out]

    def use(self, w: str):
        """showUsed (sequencerun_funcs.cc:465-551): use(seq[, repeat[, hold]]) parses the
        sub-sequence again, reuse() keeps the pulse list and propagator of its last use."""
        rehash = not w.startswith("reuse(")
        inner = _inside(w)
        args = _split_args(inner)
        if not inner or not args:
            raise BlochLibError(f" Bad 'use' usage for \"{w}\"\n should be \"use(seq)\", "
                                "\"use(seq, repeat)\", \"use(seq, repeat, hold)\" ")
        repeat = int(evaluate(args[1], self.p)) if len(args) >= 2 else 1
        if repeat < 0:
            raise BlochLibError(f" Bad 'repeat' usage for \"{w}\"\n repeat should be >=0 "
                                "(use none or more...) ")
        held = False
        if len(args) >= 3:
            held = args[2] == "hold" or bool(evaluate(args[2], self.p))
        sec = self.subs.get(args[0])
        if sec is None:
            raise BlochLibError(f" Sub Pulse section does NOT exsist use... \"{w}\"")
        ent = self.sub_cache.get(args[0])
        if rehash or ent is None:
            ent = [self.expand_sub(sec), None]
            self.sub_cache[args[0]] = ent
        if repeat > 0 and self.to_run:
            if ent[1] is None:
                ent[1] = self.seq_props(ent[0])
            props = ent[1]
            if len(props) == 1:
                pr = props[0]
                self.proplist.append(_Prop(pr.steps, pr.sys, pr.wr, repeat, held))
            elif held:
                raise BlochLibError("a held sub-sequence whose pulses use several spin systems "
                                    "is not supported")
            else:
                # the whole sequence `repeat` times, its parts in order
                for _ in range(repeat):
                    self.proplist.extend(_Prop(pr.steps, pr.sys, pr.wr) for pr in props)
        if not held and ent[0]:
            self.current_time = ent[0][-1].t2 * repeat

    # ---------------------------------------------------------------- rho of every orientation
    def set_rho(self, op):
        """setRo (sequencerun_funcs.cc:315-331)."""
        self.rho_op = op
        self.ops = []
        self.has_base = False
        self.synced = 0

    def clean(self) -> bool:
        return not self.ops

    def queue(self, props, times: int = 1):
        """Apply the propagators (and cycler projections) in order, the whole list `times`
        times (lazily)."""
        props = [p for p in props if (p.steps or p.proj is not None) and p.repeat > 0]
        if not props or times <= 0:
            return
        if len(props) == 1:
            self.ops.append((props[0], props[0].repeat * times))
        elif times == 1:
            self.ops.extend((p, p.repeat) for p in props)
        else:
            per = self.period(props)
            if per is not None:
                self.ops.append((per, times))
            else:
                for _ in range(times):
                    self.ops.extend((p, p.repeat) for p in props)

    def period(self, props):
        """The propagators applied once each, in order, as one _Prop (their repeats expanded),
        or None when they differ in spin system, spinning speed or rotor angle or include a
        cycler projection."""
        if any(p.proj is not None for p in props):
            return None
        props = [p for p in props if p.steps and p.repeat > 0]
        if not props:
            return None
        p0 = props[0]
        rots = {x[2] for p in props for x in p.steps}
        if len(rots) > 1 or any(p.wr != p0.wr or p.sys != p0.sys for p in props):
            return None
        steps = []
        for p in props:
            steps.extend(p.steps * p.repeat)
        if len(steps) > (1 << 20):
            raise BlochLibError("a pulse period of more than 2^20 steps")
        return _Prop(steps, p0.sys, p0.wr)

    def materialize(self) -> "engine.State":
        pw = self.powder()
        if self.state is None:
            self.state = engine.State(self.sys.spins, pw)
        if not self.has_base:
            self.state.set(self.rho_op)
            self.has_base = True
            self.synced = 0
        for pr, reps in self.ops[self.synced:]:
            if pr.proj is not None:
                self.state.project(pr.proj)      # idempotent: once for any repeat count
                continue
            groups = []
            for x in pr.steps:
                if groups and groups[-1][0] == x[2]:
                    groups[-1][1].append(x)
                else:
                    groups.append((x[2], [x]))
            for _ in range(1 if len(groups) == 1 else reps):
                for rot, st in groups:
                    self.state.apply(pr.sys, pr.wr, rot * RAD2DEG, [x[1] for x in st],
                                     rf=_rf_tables(st), step_t1=[x[0] for x in st],
                                     repeat=reps if len(groups) == 1 else 1)
        self.synced = len(self.ops)
        return self.state

    def powder(self) -> Powder:
        if not self.sim.powders:
            raise BlochLibError("no powder{} section in parameters")
        if self.usepow == "default":
            return self.sim.powders[min(self.sim.powders)]
        if self.usepow not in self.sim.powders:
            raise BlochLibError(f"Powder '{self.usepow}' does not exsist")
        return self.sim.powders[self.usepow]

    def set_powder(self, w: str, s: str):
        """setPowder (sequencerun_funcs.cc:212-228): a different powder restarts the FID
        accumulation.  The orientation loop of the reference restarts with it; here every
        orientation of the program is a device state, so a switch is accepted only before any
        pulse, propagator or fid() of the program."""
        inner = _inside(w)
        if not inner:
            raise BlochLibError(f" Bad 'powder' usage for \"{s}\"\n should be \"powder(pow)\" ")
        if inner == self.usepow:
            return
        if self.proplist or self.ops or self.state is not None or np.any(self.fid1):
            raise BlochLibError("powder() after pulses or fid() (a second powder within one "
                                "program) is not supported")
        old, self.usepow = self.usepow, inner
        try:
            self.powder()
        except BlochLibError:
            self.usepow = old
            raise
        self.fid1[:] = 0

    def det(self):
        return self.sys.op(str(self.p["detect"]))

    def trace(self) -> complex:
        """sum_k w_k Tr(rho_k det) / Tr(det det^H) over the orientations."""
        return self.materialize().trace(self.det(), normalize=True)

    def spin_fid(self, npts: int) -> np.ndarray:
        """hFID.FID(ro, detect, ...) / Tr(det det^H) summed over the powder
        (sequencerun_funcs.cc:690-735): gamma-COMPUTE when spinning, the static FID otherwise,
        from the current time; sets sw to the adjusted sweep width."""
        wr, rotor, sw = float(self.p["wr"]), float(self.p["rotor"]), float(self.p["sw"])
        if self.clean():
            f, sw2 = engine.powder_fid(self.sys, wr, rotor, sw, npts, self.powder(), self.rho_op,
                                       self.det(), t0=self.current_time, normalize=True,
                                       precision=self.precision)
        else:
            f, sw2 = engine.state_fid(self.sys, wr, rotor, sw, npts, self.powder(),
                                      self.materialize(), self.det(), t0=self.current_time,
                                      normalize=True, precision=self.precision)
        self.p["sw"] = sw2
        return f

    def ptop_points(self, props, npts: int) -> np.ndarray:
        """Points m = 0..npts-1 of Tr(P^m rho P^-m det) / Tr(det det^H) summed over the powder,
        P the propagators applied once each, from the current rho; rho is left after npts-1
        applications."""
        out = np.zeros(npts, dtype=np.complex128)
        if npts <= 0:
            return out
        live = [p for p in props if (p.steps or p.proj is not None) and p.repeat > 0]
        per = self.period(live)
        if not live:
            out[:] = self.trace()
            return out
        if per is None:
            for m in range(npts):
                out[m] = self.trace()
                if m + 1 < npts:
                    self.queue(live, 1)
            return out
        rot = per.steps[0][2] * RAD2DEG
        dt = [x[1] for x in per.steps]
        t1 = [x[0] for x in per.steps]
        rf = _rf_tables(per.steps)
        if self.clean():
            out = engine.pulse_fid(per.sys, per.wr, rot, dt, npts, self.powder(), self.rho_op,
                                   self.det(), rf=rf, normalize=True, precision=self.precision,
                                   step_t1=t1)
        else:
            out = engine.state_pulse_fid(per.sys, per.wr, rot, dt, npts, self.powder(),
                                         self.materialize(), self.det(), rf=rf, step_t1=t1)
        self.queue([per], npts - 1)
        return out

    # ---------------------------------------------------------------- fid() and savefid()
    def do_fid(self, w: str):
        """SequenceRun::doFID (sequencerun_funcs.cc:553-746)."""
        inner = _inside(w)
        idx = -1
        npts = int(self.p["npts1D"])
        if len(self.fid1) != npts:     # resizeAndPreserve(npts1D)
            f = np.zeros(npts, dtype=np.complex128)
            f[:min(npts, len(self.fid1))] = self.fid1[:npts]
            self.fid1 = f
        if inner.strip():
            idx = int(evaluate(inner, self.p))
            if idx < 0:
                raise BlochLibError(f"fid({inner}): negative point index {idx}")
            if not self.is2D:
                if idx >= len(self.fid1):
                    self.fid1 = np.concatenate([self.fid1, np.zeros(idx + 1 - len(self.fid1),
                                                                    dtype=np.complex128)])
                    self.p["npts1D"] = float(idx + 1)
            elif idx >= self.fid2.shape[0]:
                grow = np.zeros((idx + 1 - self.fid2.shape[0], npts), dtype=np.complex128)
                if self.fid2.shape[1] != npts:
                    raise BlochLibError("npts1D changed after the 2-D data was sized")
                self.fid2 = np.concatenate([self.fid2, grow])
                self.p["npts2D"] = float(idx + 1)
        if not self.to_run:
            self.log("collect FID" + (f" at point {idx}" if idx != -1 else ""))
            return
        if self.ptop:
            self.fid_ptop(idx, npts)
        else:
            self.fid_plain(idx, npts)

    def fid_plain(self, idx: int, npts: int):
        if not self.is2D:
            # the used propagators stay in the list (numProp is not reset on this branch)
            self.queue(self.proplist, 1)
            if idx != -1:
                self.fid1[idx] += self.trace()
            else:
                self.fid1 += self.spin_fid(npts)
                self.log(f"fid: {len(self.powder())} crystallites, {npts} points")
            return
        if self.pt2D == 0:
            self.queue(self.proplist, 1)
        elif any(p.steps and p.repeat > 0 and not p.hold for p in self.proplist):
            # subSequence::propogate with fidpt = pt2D > 0 applies holdU, which is only set
            # at fidpt == 0 (sequencerun.cc:181-194)
            raise BlochLibError("fid() without an index after the first 2-D row with "
                                "propagators before it is not supported (the reference "
                                "applies a propagator it has not computed)")
        else:
            self.queue([p for p in self.proplist if p.proj is not None], 1)
        row = idx if idx != -1 else self.pt2D
        if row >= self.fid2.shape[0]:
            raise BlochLibError(f"2-D row {row} beyond npts2D = {self.fid2.shape[0]}")
        if self.fid2.shape[1] != npts:
            raise BlochLibError("npts1D changed after the 2-D data was sized")
        self.fid2[row] += self.spin_fid(npts)
        self.log(f"2-D row {row}: {len(self.powder())} crystallites, {npts} points")
        if idx == -1:
            self.pt2D += 1
        self.proplist = []

    def fid_ptop(self, idx: int, npts: int):
        allp = list(self.proplist)
        free = [p for p in allp if not p.hold]
        if not self.is2D:
            if idx != -1:
                self.queue(allp, 1)
                self.fid1[idx] += self.trace()
                if idx == 0 and self.current_time != 0:
                    self.p["sw"] = 1.0 / self.current_time
            elif npts > 0:
                if self.clean() and len(free) == len(allp):
                    f = self.ptop_points(allp, npts)
                else:
                    f = np.zeros(npts, dtype=np.complex128)
                    f[0] = self.trace()
                    if npts > 1:
                        self.queue(allp, 1)
                        f[1:] = self.ptop_points(free, npts - 1)
                self.fid1 += f
                self.log(f"ptop fid: {len(self.powder())} crystallites, {npts} points")
            self.proplist = []
            return
        if idx == -1:
            return      # the reference does nothing for fid() in 2-D ptop mode
        if self.fid2.shape[1] != npts:
            raise BlochLibError("npts1D changed after the 2-D data was sized")
        if self.clean() and len(free) == len(allp) and self.period(allp) is not None:
            row = self.ptop_points(allp, npts + 1)[1:]
        else:
            self.queue(allp, 1)
            row = self.ptop_points(free, npts)
        self.fid2[idx] += row
        self.log(f"ptop 2-D row {idx}: {len(self.powder())} crystallites, {npts} points")
        self.proplist = []

    def save(self, w: str):
        """SequenceRun::saveFID (sequencerun_funcs.cc:752-894)."""
        if w.startswith("savefidmatlab("):
            what = "matlab"
        elif w.startswith("savefidbinary("):
            what = "binary"
        elif w.startswith("savefid(") or w.startswith("savefidtext("):
            what = "text"
        else:
            return
        args = _split_args(_inside(w))
        fname = args[0] if args else str(self.p["filesave"])
        path = os.path.join(self.outdir, fname)
        if not self.to_run:
            return
        if len(args) <= 1:
            if self.is2D:
                if what == "matlab":
                    self.files.append(fidio.save_matlab(path, {"vdat": self.fid2}))
                elif what == "binary":
                    self.files.append(fidio.save_fid2d_binary(self.fid2, path))
                else:
                    self.files.append(fidio.save_fid2d_text(self.fid2, path))
            else:
                self.files.append(fidio.save_fid(
                    {"matlab": "savefidmatlab", "binary": "savefidbinary",
                     "text": "savefid"}[what], self.fid1, path, float(self.p["sw"])))
            self.not_saved = False
            self.log("saved", self.files[-1])
            return
        idx = int(evaluate(args[1], self.p))
        rows = self.fid2.shape[0]
        if idx < 0 or idx > rows or (self.is2D and idx == rows):
            raise BlochLibError(f" Bad 'savefid' usage for \"{w}\"\n the index counter should "
                                "be >=0 and <npts2D")
        path = path + str(idx)
        cur = self.fid2[idx] if self.is2D else self.fid1
        dt = float(self.current_time) / 2 if self.ptop else 1.0 / float(self.p["sw"]) / 2
        # the text file is always written first; the MAT-file and binary forms then replace
        # it and hold the 1-D data (fid1D), not the row (sequencerun_funcs.cc:866-889)
        written = [fidio.plotter_fid(cur, path, dt)]
        if what == "matlab":
            written.append(fidio.save_matlab(path, {"vdat": self.fid1,
                                                    "sw1": float(self.p["sw"])}))
        elif what == "binary":
            written.append(fidio.save_fid_binary(self.fid1, path,
                                                 sw1="%.4f" % float(self.p["sw"])))
        else:
            written.append(fidio.plotter_fid(cur, path, dt))
        for f in written:
            if f not in self.files:
                self.files.append(f)
        self.not_saved = False
