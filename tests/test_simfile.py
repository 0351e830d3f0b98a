"""The .sim reader: sections, parameters, expressions, powders, sub-sequence expansion into
engine steps, and the explicit errors for what the engine does not compute."""
import math
import os

import numpy as np
import pytest

from blochlib_amd.powder import Powder
from blochlib_amd.simfile import MAGIC_RAD, SimFile, _Runner, evaluate, parse_sections
from blochlib_amd.spinsys import BlochLibError

STATIC = """
# two CSA-only protons, static
spins{
    numspin 2
    T 1H 0
    T 1H 1
    C 5000  2134 0 0     # iso del eta spin
    C -5000 2789 0.5 1
}
parameters{
    powder{
        aveType ../crystals/rep2000
    }
    maxtstep=5e-6
    npts1D=512
    sw=40000
    roeq= Iz
    detect=Ip
}
pulses{
    wr=0
    rotor=0
    detect(Ip)
    ro(Ix)
    fid()
    savefidtext(statfid)
}
"""

C7 = """
spins{
    numspin 2
    T 13C 0
    T 13C 1
    D -2000 0 1
}
parameters{
    powder{
        aveType zcw
        thetaStep 233
        phiStep 144
    }
    maxtstep=1e-6
    npts1D=64
    roeq= Iz_0+Iz_1
    detect=Iz_0
}
pulses{
    wr=5000
    rotor=rad2deg*acos(1/sqrt(3))
    amp=7*wr
    stph=0
    phst=360/7
    t360=1/amp
    sub1{
        amplitude(amp)
        loop(oo=0:6)
            13C:pulse(2*t360, stph)
            stph = stph +phst
        end
    }
    ptop()
    reuse(sub1)
    fid()
    savefidtext(c7fid)
}
"""


def test_expressions():
    v = {"wr": 5000.0, "amp": 35000.0}
    assert math.isclose(evaluate("rad2deg*acos(1/sqrt(3))", v), math.degrees(MAGIC_RAD))
    assert evaluate("7*wr", v) == 35000.0
    assert evaluate("2^10", v) == 1024.0
    assert evaluate("-sqrt(-1)", v) == 0.0          # the Parser pushes 0 for a bad sqrt
    assert evaluate("log(1000)+ln(e)", v) == 4.0
    with pytest.raises(BlochLibError):
        evaluate("nosuch+1", v)
    with pytest.raises(BlochLibError):
        evaluate("__import__('os')", v)


def test_sections_nest_and_strip_comments():
    root = parse_sections("a{\n x=1 # c\n b{\n y\n }\n}\nz{ }\n")
    a = root.sub("a")
    assert a.lines() == ["x=1"] and a.sub("b").lines() == ["y"] and root.sub("z") is not None
    with pytest.raises(BlochLibError):
        parse_sections("a{\n")


def test_static_file_reads(tmp_path):
    sim = SimFile(STATIC, base_dir=str(tmp_path / "x"))
    assert sim.system.spins == 2 and sim.system.isotopes == ["1H", "1H"]
    assert sim.params["npts1D"] == 512 and sim.params["sw"] == 40000 and sim.params["roeq"] == "Iz"
    assert sim.params["maxtstep"] == 5e-6 and sim.params["mintstep"] == 1e-7   # default kept
    assert len(sim.powder) == len(Powder.read("rep2000"))


def test_c7_subsequence_expands_to_steps():
    sim = SimFile(C7)
    assert len(sim.powder) == 232 and math.isclose(sim.powder.weight.sum(), 1.0)
    r = _Runner(sim, ".", "fp64", False)
    for s in ["wr=5000", "rotor=rad2deg*acos(1/sqrt(3))", "amp=7*wr", "stph=0", "phst=360/7",
              "t360=1/amp", "ptop()", "reuse(sub1)"]:
        r.statement(s)
    (segs, rep), = r.props
    assert rep == 1 and len(segs) == 7
    t360 = 1 / 35000.0
    for j, sg in enumerate(segs):
        assert math.isclose(sg.t1, 2 * t360 * j, abs_tol=1e-15)
        assert math.isclose(sg.t2 - sg.t1, 2 * t360)
        assert sg.rotor_rad == MAGIC_RAD
        (iso, a, ph, o), = sg.channels
        assert iso == "13C" and a == 35000.0 and o == 0.0
        assert math.isclose(ph, math.radians(j * 360 / 7))
    assert math.isclose(r.current_time, 14 * t360)
    st = r.steps()
    # each 57.1 us pulse becomes ceil(57.1/1) = 58 steps of maxtstep or less
    per = math.ceil(2 * t360 / 1e-6)
    assert len(st) == 7 * per and all(math.isclose(s[1], 2 * t360 / per) for s in st)
    assert math.isclose(sum(s[1] for s in st), 2 / 5000.0)
    rf = r.rf_tables(st)
    assert len(rf) == 1 and rf[0][0] == "13C" and np.all(rf[0][1] == 35000.0)


def test_repeat_reuses_the_same_times():
    sim = SimFile(C7.replace("reuse(sub1)", "reuse(sub1, 3)"))
    r = _Runner(sim, ".", "fp64", False)
    for s in ["wr=5000", "amp=7*wr", "stph=0", "phst=360/7", "t360=1/amp", "ptop()",
              "reuse(sub1, 3)"]:
        r.statement(s)
    st = r.steps()
    n = len(st) // 3
    assert [s[0] for s in st[:n]] == [s[0] for s in st[n:2 * n]]


def test_main_section_pulses_in_ptop_mode_are_timed_by_current_time():
    sim = SimFile(C7)
    r = _Runner(sim, ".", "fp64", False)
    for s in ["wr=0", "ptop()", "1H:pulse(1e-6, 90, 50000) | 13C:delay(1e-6)",
              "13C:delay(3e-6)"]:
        r.statement(s)
    assert [(p[0][0].t1, p[0][0].t2) for p in r.props] == [(0.0, 1e-6), (1e-6, 4e-6)]
    (iso, a, ph, o), (iso2, a2, _, _) = r.props[0][0][0].channels
    assert (iso, a, iso2, a2) == ("1H", 50000.0, "13C", 0.0) and math.isclose(ph, math.pi / 2)


@pytest.mark.parametrize("stmt", ["loop(i=0:3)", "2D()", "fid(3)", "savefid(x, 2)",
                                  "if(i>2)", "reuse(sub1, 1, hold)", "frobnicate(1)"])
def test_unsupported_statements_raise(stmt):
    sim = SimFile(C7)
    r = _Runner(sim, ".", "fp64", False)
    with pytest.raises(BlochLibError):
        r.statement(stmt)


def test_plain_fid_after_pulses_runs_the_preparation(monkeypatch, tmp_path):
    """Pulses before a plain fid() (sequencerun_funcs.cc:693-703): the runner hands the pulse
    steps (maxtstep-split, main-section times, rf tables), the start time (the pulses' end) and
    the pulses' rotor angle to the engine's preparation; the engine call is intercepted here."""
    from blochlib_amd import engine
    seen = {}

    def fake(sy, wr, rotor, sw, npts, pw, rho, det, dt, rf=(), step_t1=None, t0=None,
             prep_rotor_deg=None, normalize=True, precision="fp64"):
        seen.update(wr=wr, rotor=rotor, sw=sw, npts=npts, dt=np.asarray(dt), rf=rf,
                    t1=np.asarray(step_t1), t0=t0, prot=prep_rotor_deg, rho=rho)
        return np.zeros(npts, complex), 40000.0

    monkeypatch.setattr(engine, "prep_fid", fake)
    mas = STATIC.replace("wr=0", "wr=10000").replace("rotor=0", "rotor=54.7356")
    sim = SimFile(mas.replace("fid()", "1H:pulse(12e-6, 90, 25000)\n    fid()"))
    sim.base_dir = os.path.dirname(__file__)
    out = sim.run(outdir=str(tmp_path))
    assert out["sw"] == 40000.0
    assert seen["wr"] == 10000.0 and math.isclose(seen["rotor"], 54.7356)
    # 12 us in steps of at most maxtstep = 5 us: three equal steps, from t = 0
    assert np.allclose(seen["dt"], 4e-6) and np.allclose(seen["t1"], [0.0, 4e-6, 8e-6])
    assert math.isclose(seen["t0"], 12e-6) and math.isclose(seen["prot"], 54.7356)
    (iso, amp, ph, off), = seen["rf"]
    assert iso == "1H" and np.allclose(amp, 25000.0) and np.allclose(ph, math.pi / 2)


def test_cli_reports_errors(tmp_path, capsys):
    from blochlib_amd import solid
    p = tmp_path / "bad.sim"
    p.write_text(C7.replace("reuse(sub1)", "reuse(nosub)"))
    assert solid.main([str(p)]) == 1
    assert "does NOT exsist" in capsys.readouterr().err
