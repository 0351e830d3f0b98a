"""solid-2.0 input files (``.sim``) for the GPU engine: the reader and a runner of the subset of
the pulse language that the engine computes.

Reference behaviour restated (examples/solid-2.0/src):
  * sections ``spins{}``, ``parameters{}``, ``pulses{}`` (solid.cc:115-193); ``spins{}`` is
    SolidSys::read (src/QMspins/sys.cc:120-204); several spin systems are not supported;
  * ``parameters{}``: ``name=expr`` assignments over the ParamSet defaults (paramset.cc:44-57)
    and ``powder{}`` / ``powder1{}`` ... sections with ``aveType`` (a file or a built-in
    generator), ``thetaStep``, ``phiStep``, ``gammaStep`` (multipowder.cc:56-114); the first
    powder is the default one (multipowder.cc:140-154);
  * expressions: + - * / ^ with the Parser's functions and constants e, pi, Pi, PI, deg2rad,
    rad2deg (src/utils/parser.cc:77-102, 219-246, 640-720);
  * ``pulses{}`` statements: assignments, ``detect(op)``, ``ro(op)``, ``amplitude(a)``,
    ``offset(o)``, ``wr=``, ``rotor=``, ``fid()``, ``savefid/savefidtext/savefidbinary/
    savefidmatlab(name)`` (sequencerun_funcs.cc:752-894), ``ptop()``, sub-sequence sections
    ``name{...}`` used by ``use(name[, repeat])`` / ``reuse(name[, repeat])``
    (sequencerun_funcs.cc:465-551), and pulse lines ``ISO:pulse(t[, phase_deg[, amp[,
    offset]]]) | ISO:delay(t)`` (pulsedata.cc:166-280) inside sub-sequences or, in ptop mode,
    in the main section (doPulseData, sequencerun_funcs.cc:101-135); ``loop(v=a:b) ... end``
    inside sub-sequences;
  * a plain ``fid()`` (no pulses before it) is oneFID's powder FID divided by Tr(det det^H)
    (sequencerun_funcs.cc:685-705, the whole powder in one engine call) and sets sw to the
    adjusted value; a ptop ``fid()`` records Tr(rho det)/Tr(det det^H) and applies every used
    propagator between points (sequencerun_funcs.cc:641-655);
  * propagators split each pulse of length tau into ceil(tau/maxtstep) steps unless
    tau <= mintstep or wr == 0 (propogation.cc:257-300); pulses read inside a sub-sequence take
    their rotor angle from the sub-sequence's ``rotor()`` with the magic angle as default
    (pulsedata.cc:287-300), pulses of the main section the ``rotor`` parameter; a
    sub-sequence's times start at 0 each time it is used (sequencerun.cc:122-162).
  * pulses before a plain ``fid()``: every crystallite's rho is propagated through the used
    propagators, then its gamma-COMPUTE FID runs from the current time (sequencerun_funcs.cc:
    693-703) -- the engine's preparation stage (blfid_plan_set_prep).
Everything else (2-D acquisition, indexed ``fid(i)``, main-section ``loop``/``if``, ``hold``,
several spin systems or powders in one program) raises BlochLibError: the engine has no
per-crystallite CPU path to fall back on.
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
def _frac(x):
    return x - int(x)


def _round(x):
    return float(int(x + 0.5)) if x >= 0.0 else float(int(x - 0.5))


_FUNCS = {
    "abs": abs, "acos": lambda x: math.acos(x) if abs(x) <= 1 else 0.0,
    "asin": lambda x: math.asin(x) if abs(x) <= 1 else 0.0, "atan": math.atan,
    "cos": math.cos, "exp": math.exp, "floor": math.floor, "frac": _frac,
    "int": lambda x: float(int(x)), "ln": lambda x: math.log(x) if x > 1e-36 else 0.0,
    "log": lambda x: math.log10(x) if x > 1e-10 else 0.0, "pow": math.pow, "round": _round,
    "ceil": math.ceil, "sin": math.sin, "sqrt": lambda x: math.sqrt(x) if x >= 0 else 0.0,
    "tan": math.tan, "trunc": lambda x: float(int(x)), "sinh": math.sinh, "cosh": math.cosh,
    "tanh": math.tanh,
}
_CONSTS = {"e": math.e, "pi": math.pi, "Pi": math.pi, "PI": math.pi, "deg2rad": DEG2RAD,
           "rad2deg": RAD2DEG}


def evaluate(expr: str, vars: dict) -> float:
    """Numeric value of a Parser expression; names resolve to numeric parameters."""
    src = expr.strip().replace("^", "**")
    try:
        tree = ast.parse(src, mode="eval")
    except SyntaxError as e:
        raise BlochLibError(f"cannot evaluate expression '{expr}'") from e

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, ast.BinOp):
            a, b = ev(n.left), ev(n.right)
            if isinstance(n.op, ast.Add):
                return a + b
            if isinstance(n.op, ast.Sub):
                return a - b
            if isinstance(n.op, ast.Mult):
                return a * b
            if isinstance(n.op, ast.Div):
                return a / b
            if isinstance(n.op, ast.Pow):
                return a ** b
            if isinstance(n.op, ast.Mod):
                return math.fmod(a, b)
        if isinstance(n, ast.Name):
            if n.id in vars and not isinstance(vars[n.id], str):
                return float(vars[n.id])
            if n.id in _CONSTS:
                return _CONSTS[n.id]
            raise BlochLibError(f"unknown variable '{n.id}' in '{expr}'")
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS:
            return float(_FUNCS[n.func.id](*[ev(a) for a in n.args]))
        raise BlochLibError(f"cannot evaluate expression '{expr}'")

    return ev(tree)


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
    """One pulse line: its time span and per-channel (isotope, amp Hz, phase rad, offset Hz)."""
    t1: float
    t2: float
    rotor_rad: float
    channels: list


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
            raise BlochLibError("several spin systems (spins1{}, spins2{} ...) are not supported")
        self.system = SolidSys.from_text("\n".join(_flatten(spins[0])))
        self.params = dict(PARAM_DEFAULTS)
        self.powder = None
        par = root.sub("parameters")
        if par is not None:
            pows = [s for s in par.items if isinstance(s, Section) and s.name.startswith("powder")]
            pows.sort(key=lambda s: s.name)
            for ln in par.lines():
                self._param_line(ln)
            if pows:
                self.powder = self._powder(pows[0])
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


class _Runner:
    def __init__(self, sim: SimFile, outdir: str, precision: str, verbose: bool):
        self.sim = sim
        self.p = sim.params
        self.outdir = outdir
        self.precision = precision
        self.verbose = verbose
        self.subs = {it.name: it for it in sim.pulses.items if isinstance(it, Section)}
        self.ptop = False
        self.props = []          # used propagators: list of (segments, repeat)
        self.current_time = 0.0
        self.amp = 0.0
        self.offset = 0.0
        self.ro = None
        self.fid = None
        self.files = []

    def log(self, *a):
        if self.verbose:
            print("[sim]", *a, flush=True)

    def run(self):
        for it in self.sim.pulses.items:
            if isinstance(it, Section):
                continue            # sub-sequence definitions
            self.statement(it)
        return {"fid": self.fid, "sw": float(self.p["sw"]), "files": self.files}

    # statements of the main section
    def statement(self, s: str):
        w = s.replace(" ", "")
        low = w.lower()
        if not w:
            return
        if w.startswith("loop(") or w.startswith("if(") or w == "end":
            raise BlochLibError(f"'{s}': loops and conditions in the main pulse section are not "
                                "supported by this engine")
        if w.startswith("2D(") or w.startswith("2d("):
            raise BlochLibError("2-D acquisition is not supported by this engine")
        if w.startswith("show("):
            return
        if w.startswith("ptop("):
            self.ptop = True
            return
        if w.startswith("savefid"):
            cmd = w[:w.index("(")]
            args = _split_args(_inside(w))
            if len(args) > 1:
                raise BlochLibError("indexed savefid (2-D) is not supported by this engine")
            name = args[0] if args else str(self.p["filesave"])
            if self.fid is None:
                return
            path = os.path.join(self.outdir, name)
            self.files.append(fidio.save_fid(cmd, self.fid, path, float(self.p["sw"])))
            self.log("saved", self.files[-1])
            return
        if w.startswith("fid("):
            if _inside(w).strip():
                raise BlochLibError("indexed fid(i) is not supported by this engine")
            self.do_fid()
            return
        if w.startswith("detect("):
            self.p["detect"] = _inside(w)
            return
        if w.startswith("ro("):
            self.ro = _inside(w)
            return
        if w.startswith("amplitude("):
            self.amp = evaluate(_inside(w), self.p)
            return
        if w.startswith("offset("):
            self.offset = evaluate(_inside(w), self.p)
            return
        if w.startswith("use(") or w.startswith("reuse("):
            self.use(w)
            return
        if "pulse(" in low or "delay(" in low:
            seg = self.segment(w, self.current_time, self.p["rotor"] * DEG2RAD, self.amp,
                               self.offset)
            self.props.append(([seg], 1))
            self.current_time = seg.t2
            return
        m = _ASSIGN.match(s)
        if m:
            self.sim.set_param(m.group(1), m.group(2))
            return
        raise BlochLibError(f"statement '{s}' is not supported by this engine")

    def segment(self, w: str, t0: float, rotor_rad: float, amp: float, off: float) -> Segment:
        chans = []
        dur = 0.0
        for part in w.split("|"):
            if ":" not in part:
                raise BlochLibError(f"bad pulse line '{w}'")
            iso, act = part.split(":", 1)
            args = _split_args(_inside(act))
            if not args:
                raise BlochLibError(f"bad pulse line '{w}'")
            dt = evaluate(args[0], self.p)
            if act.startswith("pulse("):
                ph = evaluate(args[1], self.p) * DEG2RAD if len(args) >= 2 else 0.0
                a = evaluate(args[2], self.p) if len(args) >= 3 else amp
                o = evaluate(args[3], self.p) if len(args) >= 4 else off
                chans.append((iso, a, ph, o))
            elif act.startswith("delay("):
                chans.append((iso, 0.0, 0.0, 0.0))
            else:
                raise BlochLibError(f"bad pulse line '{w}'")
            dur = dt  # the last channel sets the length (PulseData::read)
        return Segment(t0, t0 + dur, rotor_rad, chans)

    def expand_sub(self, sec: Section):
        """readPulses over a sub-sequence: segments timed from 0."""
        segs = []
        state = {"t": 0.0, "rot": MAGIC_RAD, "amp": 0.0, "off": 0.0}

        def run(items):
            i = 0
            while i < len(items):
                it = items[i]
                if isinstance(it, Section):
                    raise BlochLibError("nested sections inside a sub-sequence")
                w = it.replace(" ", "")
                if w.startswith("loop("):
                    depth, j = 1, i + 1
                    while j < len(items):
                        wj = items[j].replace(" ", "") if isinstance(items[j], str) else ""
                        if wj.startswith("loop("):
                            depth += 1
                        elif wj == "end":
                            depth -= 1
                            if depth == 0:
                                break
                        j += 1
                    if j >= len(items):
                        raise BlochLibError("loop without end")
                    var, rng = _inside(w).split("=", 1)
                    a, b = rng.split(":", 1)
                    lo, hi = int(evaluate(a, self.p)), int(evaluate(b, self.p))
                    for v in range(lo, hi + 1):
                        self.p[var.strip()] = float(v)
                        run(items[i + 1:j])
                    i = j + 1
                    continue
                if w.startswith("rotor("):
                    state["rot"] = evaluate(_inside(w), self.p) * DEG2RAD
                elif w.startswith("amplitude("):
                    state["amp"] = evaluate(_inside(w), self.p)
                elif w.startswith("offset("):
                    state["off"] = evaluate(_inside(w), self.p)
                elif w.startswith("spinsys(") or w.startswith("powder("):
                    if _inside(w) not in ("", "default"):
                        raise BlochLibError("several spin systems / powders are not supported")
                elif "pulse(" in w or "delay(" in w:
                    seg = self.segment(w, state["t"], state["rot"], state["amp"], state["off"])
                    segs.append(seg)
                    state["t"] = seg.t2
                elif w.startswith("if(") or w == "end":
                    raise BlochLibError("conditions inside sub-sequences are not supported")
                elif _ASSIGN.match(it):
                    m = _ASSIGN.match(it)
                    self.sim.set_param(m.group(1), m.group(2))
                elif w:
                    raise BlochLibError(f"statement '{it}' inside a sub-sequence is not supported")
                i += 1

        run(sec.items)
        return segs

    def use(self, w: str):
        args = _split_args(_inside(w))
        if not args:
            raise BlochLibError(f"Bad 'use' usage for \"{w}\"")
        sec = self.subs.get(args[0])
        if sec is None:
            raise BlochLibError(f"Sub Pulse section does NOT exsist use... \"{w}\"")
        repeat = int(evaluate(args[1], self.p)) if len(args) >= 2 else 1
        if repeat < 0:
            raise BlochLibError("repeat should be >=0")
        if len(args) >= 3:
            raise BlochLibError("'hold' (2-D) is not supported by this engine")
        segs = self.expand_sub(sec)
        if segs:
            self.current_time = segs[-1].t2 * repeat
        if repeat > 0:
            self.props.append((segs, repeat))

    def steps(self):
        """Flatten the used propagators into engine steps (t1, dt, rotor, channels)."""
        wr = float(self.p["wr"])
        maxt, mint = float(self.p["maxtstep"]), float(self.p["mintstep"])
        out = []
        for segs, repeat in self.props:
            one = []
            for sg in segs:
                tau = sg.t2 - sg.t1
                if tau <= 0:
                    continue
                m = int(math.ceil(tau / maxt)) if maxt > 0 else 1
                if tau <= mint or m < 1 or wr == 0:
                    m = 1
                d = tau / m
                for j in range(m):
                    one.append((sg.t1 + j * d, d, sg.rotor_rad, sg.channels))
            out.extend(one * repeat)
        return out

    def do_fid(self):
        sy = self.sim.system
        pw = self.sim.powder
        if pw is None:
            raise BlochLibError("no powder{} section in parameters")
        npts = int(self.p["npts1D"])
        det = sy.op(str(self.p["detect"]))
        ro = sy.op(self.ro if self.ro is not None else str(self.p["roeq"]))
        wr, rotor = float(self.p["wr"]), float(self.p["rotor"])
        if not self.ptop:
            st = self.steps() if self.props else []
            if st:
                # pulses before a plain fid(): every crystallite's rho is propagated through the
                # used sub-sequences, then its gamma-COMPUTE FID runs from the current time
                # (sequencerun_funcs.cc:693-703, oneFID::setStartTime(currentTime))
                rots = {x[2] for x in st}
                if len(rots) > 1:
                    raise BlochLibError("pulses with different rotor angles in one propagator "
                                        "are not supported by this engine")
                t1 = np.array([x[0] for x in st])
                dt = np.array([x[1] for x in st])
                f, sw = engine.prep_fid(sy, wr, rotor, float(self.p["sw"]), npts, pw, ro, det,
                                        dt, rf=self.rf_tables(st), step_t1=t1,
                                        t0=self.current_time,
                                        prep_rotor_deg=next(iter(rots)) * RAD2DEG,
                                        normalize=True, precision=self.precision)
                self.log(f"fid after {len(st)} pulse steps: {len(pw)} crystallites, {npts} "
                         f"points, sw {sw}")
            else:
                f, sw = engine.powder_fid(sy, wr, rotor, float(self.p["sw"]), npts, pw, ro, det,
                                          normalize=True, precision=self.precision)
                self.log(f"fid: {len(pw)} crystallites, {npts} points, sw {sw}")
            self.p["sw"] = sw
            self.props = []
        else:
            st = self.steps()
            if not st:
                f = np.full(npts, np.sum(pw.weight) * np.trace(ro @ det)
                            / np.trace(det @ det.conj().T), dtype=np.complex128)
            else:
                rots = {s[2] for s in st}
                if len(rots) > 1:
                    raise BlochLibError("pulses with different rotor angles in one propagator "
                                        "are not supported by this engine")
                t1 = np.array([s[0] for s in st])
                dt = np.array([s[1] for s in st])
                rf = self.rf_tables(st)
                f = engine.pulse_fid(sy, wr, next(iter(rots)) * RAD2DEG, dt, npts, pw, ro, det,
                                     rf=rf, normalize=True, precision=self.precision,
                                     step_t1=t1)
                self.log(f"ptop fid: {len(pw)} crystallites, {len(st)} steps, {npts} points")
            self.props = []
        self.fid = f if self.fid is None else self.fid + f

    @staticmethod
    def rf_tables(st):
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
