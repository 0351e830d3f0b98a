"""Measurement plumbing on the CPU: the rocpd summariser, the counter-summary parser that turns
FETCH_SIZE / WRITE_SIZE into calibrated bytes per launch, and the k_corr flop count."""
import sqlite3

import numpy as np

import bench
from blochlib_amd import profsum


def make_db(path):
    con = sqlite3.connect(path)
    con.execute("create table kernels (name text, start integer, end integer)")
    con.executemany("insert into kernels values (?, ?, ?)", [
        ("void k_exp<4>(DevPlan, int)", 0, 2_000_000),
        ("void k_exp<4>(DevPlan, int)", 5_000_000, 9_000_000),
        ("k_corr<float, true>(DevPlan, int)", 0, 1_000_000),
    ])
    con.execute("create table counters_collection (kernel_name text, dispatch_id integer, "
                "counter_name text, value real)")
    # two instances per dispatch (e.g. per XCD) are summed, then averaged over dispatches
    con.executemany("insert into counters_collection values (?, ?, ?, ?)", [
        ("k_corr<float, true>(DevPlan, int)", 1, "FETCH_SIZE", 100.0),
        ("k_corr<float, true>(DevPlan, int)", 1, "FETCH_SIZE", 50.0),
        ("k_corr<float, true>(DevPlan, int)", 2, "FETCH_SIZE", 250.0),
    ])
    con.commit()
    con.close()


def test_profsum_kernel_and_counter_tables(tmp_path):
    db = tmp_path / "run_results.db"
    make_db(db)
    kt = profsum.kernel_table(sqlite3.connect(db))
    assert kt[0][0] == "k_exp<4>" and kt[0][1] == 2
    assert abs(kt[0][2] - 6.0) < 1e-12 and abs(kt[0][3] - 3.0) < 1e-12
    ct = profsum.counter_table(sqlite3.connect(db))
    assert ct == [("k_corr<float, true>", "FETCH_SIZE", 2, 200.0, 250.0)]
    text = profsum.summarize(str(tmp_path), "python3 bench.py")
    assert "k_corr<float, true>" in text and "# command" in text


def test_profiled_traffic_calibration(tmp_path, monkeypatch):
    f = tmp_path / "summary.txt"
    f.write_text(
        "# calibration copy bytes: 1073741824\n"
        "__amd_rocclr_copyBuffer            FETCH_SIZE   10   2000.0   2097152.0\n"
        "__amd_rocclr_copyBuffer            WRITE_SIZE   10   2000.0   1048576.0\n"
        "k_corr<float, true>                FETCH_SIZE   74   1000.0   3000.0\n"
        "k_corr<float, true>                WRITE_SIZE   74    100.0    300.0\n"
        "k_corr<float, false>               FETCH_SIZE   26   2000.0   3000.0\n"
        "k_corr<float, false>               WRITE_SIZE   26    200.0    300.0\n")
    monkeypatch.setattr(bench, "COUNTER_SUMMARY", str(f))
    t = bench.profiled_traffic("k_corr", "c64")
    # the copy read 2 GiB for 1 GiB moved: fetch counts are halved, writes taken as they are
    assert abs(t["fetch_factor"] - 0.5) < 1e-12 and abs(t["write_factor"] - 1.0) < 1e-12
    fetch = (74 * 1000.0 + 26 * 2000.0) / 100 * 1024 * 0.5
    write = (74 * 100.0 + 26 * 200.0) / 100 * 1024
    assert abs(t["bytes"] - (fetch + write)) < 1e-6
    assert bench.profiled_traffic("k_exp", "c64") is None


def test_corr_flop_count():
    n, P = 128, 1024
    # passes 8, 8, 2: per transform 2 x 16 radix-8 butterflies (56 + 7 x 6 flops) and 64
    # radix-2 ones (4 flops), three transforms, the scaled product 8 flops per point
    fft = 3 * (2 * 16 * (56 + 42) + 64 * 4) + 8 * n
    assert bench.corr_flops_per_elem(n, P) == fft + 18 * n + 14 * P
    # n = 64: passes 8, 8 (the last of radix 8 without twiddles)
    assert bench.corr_flops_per_elem(64, P) == 3 * (8 * 98 + 8 * 56) + 8 * 64 + 18 * 64 + 14 * P
    # below 16 points the direct sum
    assert bench.corr_flops_per_elem(8, P) == 8 * 64 + 18 * 8 + 14 * P
    # not a power of two: the direct cyclic sum
    assert bench.corr_flops_per_elem(96, P) == 8 * 96 * 96 + 18 * 96 + 14 * P
    assert np.isclose(bench.corr_flops_per_elem(1024, P), 8 * 1024.0 ** 2 + 18 * 1024 + 14 * P)
    # autocorrelation (R = conj(Q)): two transforms, |q|^2 / n, Q staging alone
    auto = 2 * (2 * 16 * (56 + 42) + 64 * 4) + 4 * n
    assert bench.corr_flops_per_elem(n, P, True) == auto + 9 * n + 14 * P
    # the direct sum is the same either way
    assert bench.corr_flops_per_elem(96, P, True) == bench.corr_flops_per_elem(96, P)


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, bench.__file__, "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_family_roofline_bound_and_units():
    """The bound follows what a family issues (matrix cores for the complex64 chains and Q
    products, vector ALUs for the propagator, the correlation and every FP64 kernel), and a
    family whose flops per algorithmic byte sit below the ridge point is rated against HBM."""
    assert bench.family("k_exp<5>") == "k_exp" and bench.family("k_corr") == "k_corr"
    d = {"ms": 1000.0, "launches": 10, "flops": 50e12, "bytes": 1e9}   # 50 TFLOP/s, 50 kflop/B
    r = bench.family_roofline("k_dyson", d, "c64")
    assert r["bound"] == "mfma" and r["unit"] == "TFLOP/s" and abs(r["achieved"] - 50.0) < 1e-9
    assert abs(r["frac"] - 50.0 / bench.FP32_PEAK_TFLOPS) < 1e-12
    assert abs(r["alg_bytes_per_launch"] - 1e8) < 1e-3
    assert bench.family_roofline("k_exp", d, "c64")["bound"] == "valu"
    assert bench.family_roofline("k_corr", d, "c64")["bound"] == "valu"
    r64 = bench.family_roofline("k_dyson", d, "fp64")
    assert r64["bound"] == "valu" and r64["peak"] == bench.FP64_PEAK_TFLOPS
    # 1 flop per byte, far below the ridge (~20 flop/B): HBM bound, achieved in GB/s
    low = {"ms": 1000.0, "launches": 1, "flops": 4e12, "bytes": 4e12}
    h = bench.family_roofline("k_corr", low, "c64")
    assert h["bound"] == "hbm" and h["unit"] == "GB/s" and abs(h["achieved"] - 4000.0) < 1e-6
    assert abs(h["frac"] - 0.5) < 1e-12
    # no flops counted (eigen-solver): bytes against HBM
    e = bench.family_roofline("k_eig", {"ms": 1.0, "launches": 1, "flops": 0.0, "bytes": 8e9}, "c64")
    assert e["bound"] == "hbm" and abs(e["achieved"] - 8000.0) < 1e-6


def test_cpu_info_caps_at_one_gpu_share():
    cores, model = bench.cpu_info()
    assert 1 <= cores <= 16 and isinstance(model, str)


def test_tuning_table_round_trip_and_saved_file():
    """The 16-int launch table <-> dict used by bench.py and the saved autotuner table: every
    entry survives the round trip (JSON keys as strings too), and the committed table parses."""
    import json
    import os
    from blochlib_amd import engine
    v = np.zeros(engine.TUNE_LEN, np.int32)
    v[[1, 3]] = 1                      # two-kernel propagator for tile classes 2 and 4
    v[5] = 2                           # correlation variant
    v[6 + engine.TUNE_FAMILIES.index("qr")] = 2048
    v[15] = 0b0101                     # other Q-product form for classes 1 and 3
    d = engine.tuning_to_dict(v)
    assert d["method"][2] == "two" and d["method"][1] == "fused"
    assert d["grid"]["qr"] == 2048 and d["qr_flip"] == {1: True, 2: False, 3: True, 4: False}
    assert np.array_equal(engine.tuning_to_array(d), v)
    assert np.array_equal(engine.tuning_to_array(json.loads(json.dumps(d))), v)
    with open(bench.TUNING_FILE) as f:
        saved = json.load(f)
    t = engine.tuning_to_array(saved["tuning"])
    assert t.shape == (engine.TUNE_LEN,) and t[15] in range(16) and os.path.basename(
        bench.TUNING_FILE) == "mi355x.json"
