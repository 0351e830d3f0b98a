"""CPU checks of the drop-in boundary: the shared library loads, exports every entry point
declared in include/blfid.h, and fails cleanly (no crash, typed status) without a device."""
import ctypes
import os
import re

import numpy as np
import pytest

from blochlib_amd import engine
from blochlib_amd.spinsys import SolidSys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "blfid.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(blfid_[a-z_]+)\s*\(", txt)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ["blfid_plan_create", "blfid_plan_destroy", "blfid_plan_run", "blfid_plan_fetch",
                 "blfid_powder_fid", "blfid_plan_set_operators", "blfid_plan_set_powder",
                 "blfid_last_error", "blfid_plan_sw"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    L = engine.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_version_string():
    assert b"gfx950" in engine.lib().blfid_version()


def test_create_rejects_bad_input_before_touching_device():
    sy = SolidSys.from_text("numspin 2\nC 0 100 0.5 0\n")
    with pytest.raises(engine.EngineError) as e:
        engine.Plan(sy, 10000.0, 54.7, 640000.0, 0)  # npts = 0 (oneFID.h:78-83)
    assert e.value.status == 1
    bad = SolidSys(isotopes=["1H"] * 2, csa=[(0.0, 100.0, 1.5, 0, 0.0, 0.0, 0.0)])
    with pytest.raises(engine.EngineError) as e:
        engine.Plan(bad, 10000.0, 54.7, 640000.0, 16)  # eta > 1 (csa.cc:57-64)
    assert e.value.status == 1


def test_no_device_is_a_typed_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("device present")
    sy = SolidSys.from_text("numspin 2\nC 0 100 0.5 0\n")
    with pytest.raises(engine.EngineError) as e:
        engine.Plan(sy, 10000.0, 54.7, 640000.0, 16)
    assert e.value.status == 3
