"""The .sim reader: sections, parameters, expressions, powders, sub-sequence expansion into
engine steps, and the explicit errors for what the engine does not compute."""
import math
import os

import numpy as np
import pytest

from blochlib_amd.powder import Powder
from blochlib_amd import fidio
from blochlib_amd.simfile import (MAGIC_RAD, SimFile, _Runner, _rf_tables, evaluate,
                                   parse_sections, run_script)
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
    (pr,) = r.proplist
    assert pr.repeat == 1 and not pr.hold and pr.wr == 5000.0
    segs = r.sub_cache["sub1"][0]
    assert len(segs) == 7
    t360 = 1 / 35000.0
    for j, sg in enumerate(segs):
        assert math.isclose(sg.t1, 2 * t360 * j, abs_tol=1e-15)
        assert math.isclose(sg.t2 - sg.t1, 2 * t360)
        assert sg.rotor_rad == MAGIC_RAD
        (iso, a, ph, o), = sg.channels
        assert iso == "13C" and a == 35000.0 and o == 0.0
        assert math.isclose(ph, math.radians(j * 360 / 7))
    assert math.isclose(r.current_time, 14 * t360)
    st = pr.steps
    # each 57.1 us pulse becomes ceil(57.1/1) = 58 steps of maxtstep or less
    per = math.ceil(2 * t360 / 1e-6)
    assert len(st) == 7 * per and all(math.isclose(s[1], 2 * t360 / per) for s in st)
    assert math.isclose(sum(s[1] for s in st), 2 / 5000.0)
    rf = _rf_tables(st)
    assert len(rf) == 1 and rf[0][0] == "13C" and np.all(rf[0][1] == 35000.0)


def test_repeat_reuses_the_same_times():
    sim = SimFile(C7.replace("reuse(sub1)", "reuse(sub1, 3)"))
    r = _Runner(sim, ".", "fp64", False)
    for s in ["wr=5000", "amp=7*wr", "stph=0", "phst=360/7", "t360=1/amp", "ptop()",
              "reuse(sub1, 3)"]:
        r.statement(s)
    (pr,) = r.proplist
    assert pr.repeat == 3
    st = r.period([pr]).steps
    n = len(st) // 3
    assert [s[0] for s in st[:n]] == [s[0] for s in st[n:2 * n]] == [s[0] for s in pr.steps]
    # currentTime = end of the pulse list times the repeat count (sequencerun_funcs.cc:537-539)
    assert math.isclose(r.current_time, 3 * 2 / 5000.0)


def test_main_section_pulses_in_ptop_mode_are_timed_by_current_time():
    sim = SimFile(C7)
    r = _Runner(sim, ".", "fp64", False)
    for s in ["wr=0", "ptop()", "1H:pulse(1e-6, 90, 50000) | 13C:delay(1e-6)",
              "13C:delay(3e-6)"]:
        r.statement(s)
    assert [(p.steps[0][0], p.steps[0][0] + p.steps[0][1]) for p in r.proplist] == \
        [(0.0, 1e-6), (1e-6, 4e-6)]
    (iso, a, ph, o), (iso2, a2, _, _) = r.proplist[0].steps[0][3]
    assert (iso, a, iso2, a2) == ("1H", 50000.0, "13C", 0.0) and math.isclose(ph, math.pi / 2)


@pytest.mark.parametrize("stmt", ["reset()", "cycler()", "spinsys(other)", "powder(p2)",
                                  "alterSys(Q1del, 3)", "alterSys(C7iso, 3)", "use(nosub)",
                                  "use(sub1, -1)", "1H:pulse(1e-6):x"])
def test_unsupported_or_bad_statements_raise(stmt):
    sim = SimFile(C7)
    r = _Runner(sim, ".", "fp64", False)
    with pytest.raises(BlochLibError):
        r.statement(stmt)


def test_unknown_statements_change_nothing():
    """Lines that no branch of SequenceRun::decide takes go to decideSeP, which records them
    for nothing in the main section (sequenceparse.cc:431-467)."""
    sim = SimFile(C7)
    r = _Runner(sim, ".", "fp64", False)
    for s in ["frobnicate(1)", "rotor(30)", "end", "else", "2D(16)"]:
        r.statement(s)
    assert not r.proplist and not r.is2D and r.p["rotor"] == 0.0


def test_plain_fid_after_pulses_applies_them_to_the_state(monkeypatch, tmp_path):
    """Pulses before a plain fid() (sequencerun_funcs.cc:686-703): the runner applies the pulse
    steps (maxtstep-split, main-section times, rf tables, the pulses' rotor angle) to every
    orientation's rho and runs the FID from the state at the current time."""
    calls = fake_engine(monkeypatch)
    mas = STATIC.replace("wr=0", "wr=10000").replace("rotor=0", "rotor=54.7356")
    sim = SimFile(mas.replace("fid()", "1H:pulse(12e-6, 90, 25000)\n    fid()"))
    sim.base_dir = os.path.dirname(__file__)
    out = sim.run(outdir=str(tmp_path))
    assert out["sw"] == 40000.0
    kinds = [c[0] for c in calls]
    assert kinds == ["create", "set", "apply", "state_fid"]
    _, sy, wr, rot, dt, rf, t1, rep = calls[2]
    assert wr == 10000.0 and math.isclose(rot, 54.7356) and rep == 1
    # 12 us in steps of at most maxtstep = 5 us: three equal steps, from t = 0
    assert np.allclose(dt, 4e-6) and np.allclose(t1, [0.0, 4e-6, 8e-6])
    (iso, amp, ph, off), = rf
    assert iso == "1H" and np.allclose(amp, 25000.0) and np.allclose(ph, math.pi / 2)
    assert math.isclose(calls[3][1]["t0"], 12e-6) and calls[3][1]["wr"] == 10000.0


# ------------------------------------------------------------------ control flow and quirks
def _record(lines, vars=None):
    seen = []
    run_script(lines, lambda s: seen.append((s, dict(vars or {}).get("i"))), vars or {})
    return [s for s, _ in seen]


def test_loops_with_step_and_nesting():
    v = {}
    seen = []
    run_script(["loop(i=0:2:6)", "loop(j=1:2)", "x", "end", "end"],
               lambda s: seen.append((s, v["i"], v["j"])), v)
    assert seen == [("x", i, j) for i in (0, 2, 4, 6) for j in (1, 2)]
    assert v["i"] == 8.0        # the variable keeps the value that ended the loop


def test_if_elseif_else_and_the_resume_quirk():
    """parseIfs returns start + the bodies' line count, its marker lines not counted
    (scriptparse.cc:392-396): the caller resumes as many lines early as the structure has
    elseif/else markers, here at C1, and passes the rest of the structure to decide again."""
    for i, want in [(1, ["A", "C1", "C2", "end"]), (2, ["B", "C1", "C2", "end"]),
                    (3, ["C1", "C2", "C1", "C2", "end"])]:
        v = {"i": float(i)}
        seen = []
        run_script(["if(i==1)", "A", "elseif(i==2)", "B", "else", "C1", "C2", "end"],
                   seen.append, v)
        assert seen == want, (i, seen)
    v = {"i": 0.0}
    seen = []
    run_script(["if(i==1)", "A", "else", "B", "end", "Z"], seen.append, v)
    assert seen == ["B", "B", "end", "Z"]
    v = {"i": 1.0}
    seen = []
    run_script(["if(i==1)", "A", "end", "Z"], seen.append, v)
    assert seen == ["A", "end", "Z"]


def test_expression_quirks_of_the_reference_parser():
    v = {"i": 3.0, "n": 8.0, "t": 5.0}
    assert evaluate("-2^2", v) == 4.0          # unary minus before ^
    assert evaluate("2*-3", v) == 6.0          # '*-' is one token
    assert evaluate("i%2==1", v) == 1.0        # one precedence level, left to right
    assert evaluate("1+2>2", v) == 1.0
    assert evaluate("i!=(n-1)", v) == 1.0
    assert evaluate("t", v) == 0.0             # x, y, z, t are the Parser's own zeros
    assert evaluate("5+1/0", v) == 5.0         # an evaluation error stops the machine
    assert evaluate("2^3^2", v) == 512.0       # ^ is right-associative
    assert evaluate("-7%3", v) == -1.0         # C remainder on truncated integers


def test_sub_sequence_reads_loop_variables_at_their_final_value():
    """SequenceParse unrolls the loop, readPulses evaluates the recorded lines afterwards: a
    loop variable inside the pulses reads its last value, an assignment chain is applied in
    order (sequenceparse.cc:386-430, pulsedata.cc:287-378)."""
    text = C7.replace("""        loop(oo=0:6)
            13C:pulse(2*t360, stph)
            stph = stph +phst
        end""", """        loop(oo=0:2)
            13C:pulse(t360, oo*10)
            13C:pulse(t360, stph)
            stph = stph +phst
        end""")
    sim = SimFile(text)
    r = _Runner(sim, ".", "fp64", False)
    for s in ["wr=5000", "amp=7*wr", "stph=0", "phst=360/7", "t360=1/amp"]:
        r.statement(s)
    segs = r.expand_sub(r.subs["sub1"])
    ph = [round(math.degrees(sg.channels[0][2]), 9) for sg in segs]
    assert ph[0::2] == [30.0, 30.0, 30.0]
    assert ph[1::2] == [round(k * 360 / 7, 9) for k in range(3)]


def test_alter_sys_follows_set_spin_param():
    from blochlib_amd.spinsys import SolidSys
    sy = SolidSys.from_text("numspin 3\nT 1H 0\nT 1H 1\nT 1H 2\nC 100 2000 0.5 1\n"
                            "D 500 0 1\nD 700 1 2\nJ 30 0 2 0\n")
    sy.set_spin_param("C1iso", 250)
    sy.set_spin_param("C1eta", 0.25)
    sy.set_spin_param("C1", 300)
    sy.set_spin_param("C1gamma", 45)
    assert sy.csa[0] == (300.0, 2000.0, 0.25, 1, 0.0, 0.0, 45.0)
    sy.set_spin_param("D21", 900)          # either spin order; the last dipole matches
    assert sy.dip[1][0] == 900.0
    with pytest.raises(BlochLibError):      # the reference raises: its last dipole is D12
        sy.set_spin_param("D01", 600)
    sy.set_spin_param("J20", 40)
    assert sy.jc[0][0] == 40.0
    for bad in ("C0iso", "D02", "Q1del", "C1eta9"):
        with pytest.raises(BlochLibError):
            sy.set_spin_param(bad, 1.0 if bad != "C1eta9" else 2.0)


# ------------------------------------------------------------------ runner with a fake engine
def fake_engine(monkeypatch):
    """Replace the engine's device entry points by recorders (CPU-only runner tests)."""
    from blochlib_amd import engine
    calls = []

    class FakeState:
        def __init__(self, nspins, pw):
            self.N = 1 << nspins
            calls.append(("create", nspins, len(pw)))

        def set(self, rho):
            calls.append(("set", np.array(rho)))

        def apply(self, sy, wr, rotor_deg, step_dt, rf=(), step_t1=None, repeat=1):
            calls.append(("apply", sy, wr, rotor_deg, np.asarray(step_dt), rf,
                          np.asarray(step_t1), repeat))

        def trace(self, det, normalize=True):
            calls.append(("trace",))
            return 1.0 + 0j

        def project(self, op):
            calls.append(("project", np.array(op)))

        def close(self):
            pass

    def fid_stub(kind):
        def f(sy, wr, rotor, *a, **kw):
            npts = a[1] if kind in ("powder_fid", "state_fid") else a[1]
            calls.append((kind, dict(wr=wr, rotor=rotor, t0=kw.get("t0"), args=a, sys=sy)))
            n = a[1] if kind in ("powder_fid", "state_fid") else a[1]
            out = np.full(n, len(calls), dtype=complex)
            return (out, 40000.0) if kind in ("powder_fid", "state_fid") else out
        return f

    monkeypatch.setattr(engine, "State", FakeState)
    for k in ("powder_fid", "state_fid", "pulse_fid", "state_pulse_fid"):
        monkeypatch.setattr(engine, k, fid_stub(k))
    return calls


MAS2D = STATIC.replace("wr=0", "wr=10000").replace("rotor=0", "rotor=54.7356").replace(
    """    detect(Ip)
    ro(Ix)
    fid()
    savefidtext(statfid)""", """    fidpts=4
    2D()
    detect(Ip)
    ro(Ix)
    dwell2D=0.00002
    loop(i=0:fidpts-1)
        fid(i)
        if(i!=(fidpts-1))
            1H:delay(dwell2D)
        end
    end
    savefidmatlab(2dmas)""")


def test_2d_rows_carry_rho_through_the_delays(monkeypatch, tmp_path):
    """examples/solid-2.0/examples/mas/mas-2d.sim: row i is the FID after i delays, from
    the current time; each fid(i) applies only the delay used since the previous one."""
    calls = fake_engine(monkeypatch)
    sim = SimFile(MAS2D)
    sim.base_dir = os.path.dirname(__file__)
    res = sim.run(outdir=str(tmp_path))
    kinds = [c[0] for c in calls]
    assert kinds == ["powder_fid", "create", "set", "apply", "state_fid", "apply", "state_fid",
                     "apply", "state_fid"]
    t0 = [c[1]["t0"] for c in calls if c[0] in ("powder_fid", "state_fid")]
    assert np.allclose(t0, [0.0, 2e-5, 4e-5, 6e-5])
    for c in calls:
        if c[0] == "apply":
            assert np.isclose(c[4].sum(), 2e-5) and c[7] == 1
    assert res["fid2d"].shape == (4, 512)
    assert [res["fid2d"][i, 0] for i in range(4)] == [1, 5, 7, 9]
    m = fidio.read_fid_matlab(res["files"][0])
    assert np.array_equal(m["vdat"], res["fid2d"])


POSTC7_2D = C7.replace("""    ptop()
    reuse(sub1)
    fid()
    savefidtext(c7fid)""", """    fidpt=3
    2D()
    loop(m=0:fidpt-1)
        reuse(sub1, m)
        13C:pulse(1e-6, 270, 250000)
        fid(m)
        ro(Iz_0+Iz_1)
    end
    savefidbinary(2dc7)""")


def test_2d_reuse_repeat_then_pulse_per_row(monkeypatch, tmp_path):
    """examples/solid-2.0/examples/c7s/postc7-2d.sim: row m applies the sub-sequence m times
    (one propagator, repeated) and the pulse, then ro() resets every orientation's rho."""
    calls = fake_engine(monkeypatch)
    sim = SimFile(POSTC7_2D)
    res = sim.run(outdir=str(tmp_path))
    kinds = [c[0] for c in calls]
    assert kinds == ["create", "set", "apply", "state_fid",
                     "set", "apply", "apply", "state_fid",
                     "set", "apply", "apply", "state_fid"]
    applies = [c for c in calls if c[0] == "apply"]
    assert [a[7] for a in applies] == [1, 1, 1, 2, 1]
    # the pulse starts where the repeated sub-sequence ended
    assert np.isclose(applies[2][6][0], 1 * 2 / 5000.0) and np.isclose(applies[4][6][0], 2 * 2 / 5000.0)
    t0 = [c[1]["t0"] for c in calls if c[0] == "state_fid"]
    assert np.allclose(t0, [1e-6, 2 / 5000.0 + 1e-6, 4 / 5000.0 + 1e-6])
    got = fidio.read_fid2d_binary(res["files"][0])
    assert np.array_equal(got, res["fid2d"])


def test_ptop_indexed_fid_and_held_propagators(monkeypatch, tmp_path):
    """ptop fid(i) applies every used propagator once (held ones too, fidpt = 0) and records
    one point; sw = 1/currentTime at i = 0 (sequencerun_funcs.cc:626-637)."""
    calls = fake_engine(monkeypatch)
    text = C7.replace("""    ptop()
    reuse(sub1)
    fid()
    savefidtext(c7fid)""", """    ptop()
    ro(Iz_0)
    loop(i=0:2)
        reuse(sub1, 2, hold)
        13C:pulse(1e-6, 0, 250000)
        fid(i)
    end
    savefidtext(tdr)""")
    sim = SimFile(text)
    res = sim.run(outdir=str(tmp_path))
    kinds = [c[0] for c in calls]
    assert kinds == ["create", "set"] + ["apply", "apply", "trace"] * 3
    reps = [c[7] for c in calls if c[0] == "apply"]
    assert reps == [2, 1] * 3
    assert np.allclose(res["fid"][:3], 1.0) and res["sw"] == 1.0 / 1e-6


def test_ptop_2d_row_is_points_after_each_application(monkeypatch, tmp_path):
    """2-D ptop fid(m): point i is recorded after the i+1-th application (held propagators
    only at the first), sequencerun_funcs.cc:656-676."""
    calls = fake_engine(monkeypatch)
    text = C7.replace("""    reuse(sub1)
    fid()
    savefidtext(c7fid)""", """    2D()
    loop(m=0:1)
        use(sub1)
        fid(m)
        ro(Iz_0+Iz_1)
    end
    savefid(c7rows)""")
    sim = SimFile(text)
    res = sim.run(outdir=str(tmp_path))
    kinds = [c[0] for c in calls]
    assert kinds == ["pulse_fid", "pulse_fid"]
    for c in calls:
        assert c[1]["args"][1] == 65          # npts1D + 1 points, the first dropped
    got = fidio.read_fid2d_text(res["files"][0])
    assert np.allclose(got, res["fid2d"], rtol=1e-5)


def test_show_turns_the_run_into_a_listing(monkeypatch, tmp_path):
    calls = fake_engine(monkeypatch)
    sim = SimFile(C7.replace("    ptop()", "    show()\n    ptop()"))
    res = sim.run(outdir=str(tmp_path))
    assert calls == [] and res["files"] == []


def test_indexed_savefid_writes_text_then_the_1d_forms(monkeypatch, tmp_path):
    fake_engine(monkeypatch)
    sim = SimFile(POSTC7_2D.replace("savefidbinary(2dc7)",
                                    "savefidmatlab(r, 1)\n    savefidbinary(b, 2)"))
    res = sim.run(outdir=str(tmp_path))
    names = [os.path.basename(f) for f in res["files"]]
    assert names == ["r1", "b2"]
    m = fidio.read_fid_matlab(res["files"][0])
    assert m["sw1"] == 40000.0 and np.array_equal(m["vdat"], res["fid"])
    with open(res["files"][1], "rb") as fh:
        assert fh.readline() == b"sw1 =40000.0000\n"
    with pytest.raises(BlochLibError):
        SimFile(POSTC7_2D.replace("savefidbinary(2dc7)", "savefid(x, 3)")).run(
            outdir=str(tmp_path))


def test_cli_reports_errors(tmp_path, capsys):
    from blochlib_amd import solid
    p = tmp_path / "bad.sim"
    p.write_text(C7.replace("reuse(sub1)", "reuse(nosub)"))
    assert solid.main([str(p)]) == 1
    assert "does NOT exsist" in capsys.readouterr().err


MULTI = """
spins{
    numspin 2
    T 1H 0
    T 1H 1
    C 1000 2000 0 0
    spin1{
        D -3000 0 1
    }
    spin2{
        D -6000 0 1
        J 50 0 1
    }
}
parameters{
    powder1{
        aveType ../crystals/rep2000
    }
    powder2{
        aveType zcw
        thetaStep 233
        phiStep 144
    }
    npts1D=64
    sw=40000
    roeq= Iz
}
pulses{
    sub1{
        amplitude(100000)
        1H:pulse(2.5e-6, 0)
        spinsys(spin2)
        1H:delay(10e-6)
    }
    wr=0
    rotor=0
    PROGRAM
}
"""


def multi(program):
    return SimFile(MULTI.replace("PROGRAM", program))


def test_spin_systems_and_powders_are_named_like_multisystem():
    """MultiSystem / MultiPowder: spin1{}, spin2{} extend the global lines; powder1{},
    powder2{} are named powders; 'default' is the first of each; without subsections the
    global lines are 'spin1'."""
    sim = multi("fid()")
    assert sorted(sim.systems) == ["spin1", "spin2"] and sorted(sim.powders) == ["powder1", "powder2"]
    s1, s2 = sim.systems["spin1"], sim.systems["spin2"]
    assert len(s1.csa) == len(s2.csa) == 1 and len(s1.dip) == len(s2.dip) == 1
    assert s1.dip[0][0] == -3000 and s2.dip[0][0] == -6000 and len(s2.jc) == 1 and not s1.jc
    assert sim.system is s1 and sim.powder is sim.powders["powder1"]
    one = SimFile(STATIC)
    assert list(one.systems) == ["spin1"] and list(one.powders) == ["powder"]
    bad = MULTI.replace("        J 50 0 1\n", "        T 1H 2\n        J 50 0 2\n")
    with pytest.raises(BlochLibError):
        SimFile(bad.replace("PROGRAM", "fid()"))


def test_spinsys_switches_the_system_of_later_pulses_and_fids(monkeypatch, tmp_path):
    """Main section: a pulse before spinsys(spin2) propagates with spin1, the fid() after it
    uses spin2 (sequencerun_funcs.cc:101-134, 179-193, 594-595)."""
    calls = fake_engine(monkeypatch)
    sim = multi("""detect(Ip)
    ro(Ix)
    amplitude(50000)
    1H:pulse(5e-6)
    spinsys(spin2)
    fid()""")
    sim.run(outdir=str(tmp_path))
    applies = [c for c in calls if c[0] == "apply"]
    fids = [c for c in calls if c[0] in ("powder_fid", "state_fid")]
    assert len(applies) == 1 and applies[0][1] == sim.systems["spin1"]
    assert len(fids) == 1 and fids[0][1]["sys"] == sim.systems["spin2"]
    with pytest.raises(BlochLibError):
        multi("spinsys(spin3)\n    fid()").run(outdir=str(tmp_path))


def test_spinsys_inside_a_subsequence_splits_its_propagator(monkeypatch, tmp_path):
    """readPulses gives each pulse the system named before it in the sequence (pulsedata.cc:
    313-317, 362): use(sub1, 2) applies [pulse on spin1, delay on spin2] twice, in order."""
    calls = fake_engine(monkeypatch)
    sim = multi("""detect(Ip)
    ro(Ix)
    use(sub1, 2)
    fid()""")
    sim.run(outdir=str(tmp_path))
    applies = [c for c in calls if c[0] == "apply"]
    systems = [c[1] for c in applies]
    s1, s2 = sim.systems["spin1"], sim.systems["spin2"]
    assert systems == [s1, s2, s1, s2]
    assert all(c[7] == 1 for c in applies)
    with pytest.raises(BlochLibError):          # held: one propagator per sequence
        multi("use(sub1, 2, hold)\n    fid()").run(outdir=str(tmp_path))


def test_powder_selects_a_named_powder_before_the_program_runs(monkeypatch, tmp_path):
    """powder(powder2) before any pulse or fid() makes the FID use that powder; a switch after
    a fid() (a second powder within one program) is refused."""
    calls = fake_engine(monkeypatch)
    sim = multi("""powder(powder2)
    detect(Ip)
    ro(Ix)
    fid()""")
    sim.run(outdir=str(tmp_path))
    fids = [c for c in calls if c[0] in ("powder_fid", "state_fid")]
    assert len(fids[0][1]["args"][2]) == len(sim.powders["powder2"]) != len(sim.powders["powder1"])
    with pytest.raises(BlochLibError):
        multi("fid()\n    powder(powder2)\n    fid()").run(outdir=str(tmp_path))
    with pytest.raises(BlochLibError):
        multi("powder(powder7)\n    fid()").run(outdir=str(tmp_path))


def test_cycler_projects_rho_where_the_list_is_applied(monkeypatch, tmp_path):
    """Main-section cycler(op) is a propagator-list entry (doCycler, sequencerun_funcs.cc:
    232-256): the fid() after it applies the pulse, then the projection onto op; cycler(Ie)
    adds nothing; a cycler line inside a sub-sequence has no effect."""
    calls = fake_engine(monkeypatch)
    sim = multi("""detect(Ip)
    ro(Ix)
    use(sub1)
    cycler(Iz)
    cycler(Ie)
    fid()""".replace("use(sub1)", "amplitude(50000)\n    1H:pulse(5e-6)"))
    sim.run(outdir=str(tmp_path))
    kinds = [c[0] for c in calls]
    assert kinds == ["create", "set", "apply", "project", "state_fid"], kinds
    proj = [c[1] for c in calls if c[0] == "project"][0]
    assert np.array_equal(proj, sim.system.op("Iz"))
    calls.clear()
    sub = MULTI.replace("        spinsys(spin2)\n", "        cycler(Iz)\n")
    sim = SimFile(sub.replace("PROGRAM", "detect(Ip)\n    ro(Ix)\n    use(sub1)\n    fid()"))
    sim.run(outdir=str(tmp_path))
    assert "project" not in [c[0] for c in calls]
