"""Benchmark: powder-averaged MAS FID throughput on one MI355X.

Workload (the batched configuration the headline metric is quoted on): 64 seeded homonuclear
1H clusters of 4-8 spins (Hilbert dimension 16-256), MAS 15 kHz at the magic angle, 128 steps
per rotor period (sw = 1.92 MHz), ZCW 3722 orientations, 1024-point FID per cluster,
rho = Ix, detect = I+ (gamma-COMPUTE path).

One step = one pass of the engine over the whole workload: 64 powder FIDs, i.e.
64 * 3722 * 128 = 30.5 M crystallite-steps.  Inputs (spin-system tables, operators, powder
sets) are resident in HBM before the timed region; the timed region is launch-to-completion
of every plan's run (plans dealt round robin to --streams HIP streams, default 1).

The default precision is c64 (complex64 propagators and products on matrix cores, FP64
eigen-solver, phases, FID points and powder sum: solid-2.0's own form); the plain FP64 form is
then timed on the same workload too, and every c64 FID is compared with its FP64 counterpart.

Prints one JSON line: metric, value (crystallite-steps/s), FIDs/s, per-FID latency p50/p99
inside a step, accuracy against the FP64 oracle, the roofline of the dominant kernel (HIP
events around its launches during the timed steps; HBM traffic from the committed counter
summary of this command) and a CPU baseline (the oracle, i.e. the reference algorithm
restated in C, on a bounded sample, one thread per sample).

--workload ptop measures the pulse-program path instead: 13C + 6 1H (N=128) under MAS
12.5 kHz with 100 kHz phase-alternating 1H rf, 256 steps per period, ZCW 1154, 2048
rotor-synchronised points, one FID per CSA parameter set (--sets of them per step).
--workload serve runs the serving mix (Poisson arrivals admitted by device-memory footprint).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from blochlib_amd import engine, workloads  # noqa: E402

# launch choices of the autotuner saved for this device (blfid_autotune on the batched workload)
TUNING_FILE = os.path.join(ROOT, "blochlib_amd", "tuning", "mi355x.json")
TUNING_SOURCE = ["defaults (no saved table)"]

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
DTYPE = {"fp64": "f64", "c64": "c64 (f32 products, f64 eig/phases/sums)",
         "bf16x3": "c64 (bf16x3 split products, f32 exp, f64 eig/phases/sums)"}
METRIC = "crystallite-steps/s + FIDs/s at ≤1e-5 rel L2 vs FP64 oracle; p50/p99 latency"
# counter summaries of this benchmark's default command (FETCH_SIZE and WRITE_SIZE passes)
COUNTER_SUMMARY = os.path.join(ROOT, "profiles", "config2_batched_{prec}_hbm_counters.txt")


def corr_flops_per_elem(n: int, P: int, autoc: bool = False) -> float:
    """Executed flops of k_corr per correlation element and crystallite (as the engine counts
    them): for n a power of two from 16 to 512 three mixed-radix FFTs (radix-8 passes of a
    56-flop DFT and 7 twiddle products per butterfly, the last pass of radix 2, 4 or 8 without
    twiddles) and the scaled product, else the direct sum; plus twiddle staging and the FID
    sums.  autoc: the autocorrelation variant (R = conj(Q)) with two transforms, a |q|^2 / n
    product (4 flops per point) and the staging of Q alone (9 n)."""
    c = 8.0 * n * n
    stage = 18.0 * n
    if 16 <= n <= 512 and (n & (n - 1)) == 0:
        ln = int(n).bit_length() - 1
        npass = (ln + 2) // 3
        rl = 1 << (ln - 3 * (npass - 1))
        dft = {2: 4.0, 4: 16.0, 8: 56.0}[rl]
        t = 2.0 if autoc else 3.0
        c = t * (npass - 1) * (n / 8.0) * 98.0 + t * (n / rl) * dft + (4.0 if autoc else 8.0) * n
        if autoc:
            stage = 9.0 * n
    return c + stage + 14.0 * P


def corr_is_autocorrelation(wls) -> bool:
    """R = conj(Q) for every workload: rho == (det + det^H) / 2 exactly and det has no entry
    pair (a, b), (b, a) (so no correlation element has a reverse element)."""
    for w in wls:
        rho, det = w.operators()
        if not np.array_equal(rho, 0.5 * (det + det.conj().T)):
            return False
        if np.any((det != 0) & (det.T != 0)):
            return False
    return True


# What each kernel family issues on MI355X, per precision: "mfma" where its products run on the
# matrix cores (complex64 Dyson chain, B = A W and Q products), "valu" where they run on the
# vector ALUs (the propagator's real Taylor GEMMs, the FFT correlation, every FP64 product).
ISSUES = {"k_exp": "valu", "k_dyson": "mfma", "k_expdyson": "mfma", "k_bmul": "mfma", "k_qr": "mfma",
          "k_corr": "valu", "k_eig": "valu", "k_reduce": "valu", "k_static_fid": "valu"}


def family(kernel: str) -> str:
    return kernel.split("<")[0]


def cpu_info():
    """Host cores this process may use (capped at the 16-core share of one GPU on the box) and
    the CPU model name."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return min(avail, 16), model


def profiled_traffic(kernel: str, precision: str):
    """HBM bytes per launch of `kernel` from the committed counter summary of this workload:
    FETCH_SIZE + WRITE_SIZE (KiB per dispatch, averaged over the kernel's dispatches; all
    template variants whose name starts with `kernel`), each scaled by the known size of the
    calibration copy over what the same counter read for it (the largest runtime copy
    dispatch).  None when the summary is absent."""
    path = COUNTER_SUMMARY.format(prec=precision)
    if not os.path.exists(path):
        return None
    copy_bytes = None
    vals = {}  # counter -> [dispatches, sum of means * dispatches]
    cal = {}
    for line in open(path):
        if line.startswith("# calibration copy bytes:"):
            copy_bytes = float(line.split(":")[1])
            continue
        f = line.split()
        if len(f) < 5 or f[-4] not in ("FETCH_SIZE", "WRITE_SIZE"):
            continue
        name, cname = " ".join(f[:-4]), f[-4]
        n, mean, mx = int(f[-3]), float(f[-2]), float(f[-1])
        if name.startswith("__amd_rocclr_copyBuffer"):
            cal[cname] = max(cal.get(cname, 0.0), mx)
        elif name == kernel or name.startswith(kernel + "<"):
            v = vals.setdefault(cname, [0, 0.0])
            v[0] += n
            v[1] += mean * n
    if copy_bytes is None or set(vals) != {"FETCH_SIZE", "WRITE_SIZE"} or set(cal) != set(vals):
        return None
    factor = {c: copy_bytes / (cal[c] * 1024.0) for c in cal}
    fetch = vals["FETCH_SIZE"][1] / vals["FETCH_SIZE"][0] * 1024.0 * factor["FETCH_SIZE"]
    write = vals["WRITE_SIZE"][1] / vals["WRITE_SIZE"][0] * 1024.0 * factor["WRITE_SIZE"]
    return {"bytes": fetch + write, "fetch_bytes": fetch, "write_bytes": write,
            "fetch_factor": factor["FETCH_SIZE"], "write_factor": factor["WRITE_SIZE"],
            "source": os.path.relpath(path, ROOT)}


def family_roofline(fam: str, d: dict, precision: str) -> dict:
    """Roofline of one kernel family over the timed steps (all its tile variants): executed
    algorithmic flops and algorithmic HBM bytes (as the engine counts them per launch) over the
    family's summed launch time.  The bound follows what the family issues (ISSUES): matrix-
    core ("mfma") or vector ("valu") FP32 for complex64, vector FP64 for the FP64 form; when its
    flops per algorithmic byte fall below MI355X's ridge point it is "hbm" with GB/s."""
    fp64 = precision == "fp64"
    bound = "valu" if fp64 else ISSUES.get(fam, "valu")
    peak = FP64_PEAK_TFLOPS if fp64 else FP32_PEAK_TFLOPS
    sec = max(d["ms"], 1e-9) * 1e-3
    launches = max(d["launches"], 1)
    tflops = d["flops"] / sec / 1e12
    gbs = d["bytes"] / sec / 1e9
    out = {"kernel": fam, "bound": bound, "achieved": tflops, "peak": peak, "unit": "TFLOP/s",
           "frac": tflops / peak, "traffic": None, "avg_launch_ms": d["ms"] / launches,
           "launches": d["launches"], "ms_total": d["ms"],
           "flops_per_launch": d["flops"] / launches,
           "alg_bytes_per_launch": d["bytes"] / launches, "alg_gbs": gbs}
    if d["flops"] > 0 and d["bytes"] > 0 and \
            d["flops"] / d["bytes"] < peak * 1e12 / (HBM_PEAK_GBS * 1e9):
        out.update({"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": gbs / HBM_PEAK_GBS, "achieved_tflops": tflops,
                    "flop_peak_tflops": peak})
    elif d["flops"] == 0 and d["bytes"] > 0:
        # no flops counted (the iterative eigen-solver): its bytes against HBM
        out.update({"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": gbs / HBM_PEAK_GBS})
    out["note"] = ("achieved = executed algorithmic flops (block-diagonal algorithm, unpadded "
                   "sizes) or algorithmic HBM bytes of the family per launch / its average "
                   "launch time (HIP events on the launch stream); peak = MI355X published "
                   + ("FP64" if fp64 else "FP32 (matrix for mfma, vector for valu; the vector "
                      "figure assumes packed FP32, which this build does not emit, so its "
                      "practical vector ceiling is half)")
                   + "; bound from what the family issues")
    return out


def hip():
    L = ctypes.CDLL("libamdhip64.so")
    L.hipDeviceSynchronize.restype = ctypes.c_int
    L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                      ctypes.c_void_p]
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    L.hipEventDestroy.argtypes = [ctypes.c_void_p]
    return L


def copy_calibration(H, mib: int):
    """One device-to-device copy of `mib` MiB (the runtime's copy kernel): under a counter
    profile its FETCH_SIZE / WRITE_SIZE per dispatch calibrate the kernels' byte counts."""
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    n = mib << 20
    H.hipMalloc(ctypes.byref(a), n)
    H.hipMalloc(ctypes.byref(b), n)
    H.hipMemset(a, 1, n)
    H.hipDeviceSynchronize()
    H.hipMemcpy(b, a, n, 3)  # hipMemcpyDeviceToDevice
    H.hipDeviceSynchronize()
    H.hipFree(a)
    H.hipFree(b)
    print(f"[bench] calibration copy of {mib} MiB done", flush=True)


def cpu_baseline(wls, sample_mode: str):
    """The oracle (the reference gamma-COMPUTE algorithm restated in C, FP64, dense N x N, its
    GEMMs on a register-blocked 4x4 complex AVX2 kernel) threaded over crystallites on all the
    host cores of this process (capped at 16, one GPU's share of the box).  For each spin count
    of the cluster mix, C crystallites of one cluster (one per thread) are timed as a whole;
    median of 3 repetitions (one for 8 spins, whose single pass already takes ~30-60 s).  The
    crystallite-steps/s of the mix follow from those per-spin-count rates and the mix's
    composition.  sample_mode "quick" skips 8 spins (scaled from 7 by N^3, and said so)."""
    from oracle import oracle as O
    cores, model = cpu_info()
    reps = {}
    counts = {}
    for w in wls:
        reps.setdefault(w.meta["spins"], w)
        counts[w.meta["spins"]] = counts.get(w.meta["spins"], 0) + 1
    spins = sorted(reps)
    measure = spins if sample_mode == "full" else [x for x in spins if x <= 7]
    sec = {}  # wall seconds per crystallite with `cores` threads busy
    detail = {}
    t_all = time.perf_counter()
    for S in measure:
        w = reps[S]
        rho, det = w.operators()
        pw = w.powder().subset(np.arange(cores) * 37 % w.powder_size)
        runs = []
        for _ in range(3 if S <= 7 else 1):
            t = time.perf_counter()
            O.powder_fid(w.system, w.wr, w.rotor, w.sw, w.npts, pw, rho, det, nthreads=cores)
            runs.append(time.perf_counter() - t)
        sec[S] = float(np.median(runs)) / cores
        detail[S] = [round(x, 2) for x in runs]
        print(f"[bench] cpu baseline: {S} spins, {cores} crystallites on {cores} threads: "
              f"{detail[S]} s", flush=True)
    scaled = [S for S in spins if S not in sec]
    for S in scaled:
        sec[S] = sec[max(measure)] * 8.0 ** (S - max(measure))
    n = wls[0].steps
    K = wls[0].powder_size
    cs_total = sum(counts[S] * K * n for S in spins)
    sec_total = sum(counts[S] * K * sec[S] for S in spins)
    return {
        "value": cs_total / sec_total,
        "unit": "crystallite-steps/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": model,
        "sample": ("oracle (reference gamma-COMPUTE algorithm in C, FP64, dense N x N, GEMMs on a "
                   "register-blocked 4x4 complex AVX2/FMA kernel, the Fave sum with a complex exp "
                   f"per term as the reference) on {cores} threads, {cores} crystallites of one "
                   f"cluster per spin count {measure} timed as a whole (wall s per repetition "
                   f"{detail}; median of 3, one pass for 8 spins), weighted by the 64-cluster mix"
                   + (f"; spin counts {scaled} scaled by N^3" if scaled else "")
                   + f"; {time.perf_counter() - t_all:.0f} s of sampling on {model}"),
        "seconds_per_crystallite_all_cores": {str(k): v for k, v in sec.items()},
        "fids_per_s": len(wls) / sec_total,
    }


def cpu_baseline_ptop(w):
    """Oracle (explicit Dyson product with rf, then rho <- U rho U^H per point, FP64) threaded
    over crystallites on all host cores of this process (capped at 16): C crystallites of one
    parameter set at 16 and at 48 points (median of 3 each) separate the period propagator from
    the per-point cost, combined for the full point count."""
    from oracle import oracle as O
    cores, model = cpu_info()
    pw = w.powder().subset(np.arange(cores) * 37 % w.powder_size)
    rho, det = w.operators()
    t = []
    for npts in (16, 48):
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            O.ptop_fid(w.system, w.wr, w.rotor, w.steps, npts, pw, rho, det, rf=w.rf,
                       nthreads=cores)
            runs.append(time.perf_counter() - t0)
        t.append(float(np.median(runs)) / cores)
    per_point = max((t[1] - t[0]) / 32.0, 0.0)
    fixed = max(t[0] - 16 * per_point, 0.0)
    sec = fixed + w.npts * per_point
    return {
        "value": w.steps / sec,
        "unit": "crystallite-steps/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": model,
        "sample": ("oracle (explicit rf Dyson product + point-by-point U rho U^H, FP64, dense "
                   f"N=128, blocked AVX2 GEMMs) on {cores} threads, {cores} crystallites of "
                   f"parameter set 0 at 16 and 48 points (median of 3): {fixed:.3f} s for the "
                   f"period propagator + {per_point * 1e3:.2f} ms per point per crystallite with "
                   f"all threads busy, combined for {w.npts} points = {sec:.3f} s per "
                   f"crystallite; {model}"),
    }


def accuracy_ptop(wls, precision, ncryst=2, npts=64):
    from oracle import oracle as O
    w = wls[0]
    pw = w.powder().subset(np.arange(ncryst) * 577 % w.powder_size)
    rho, det = w.operators()
    g = engine.ptop_fid(w.system, w.wr, w.rotor, w.steps, npts, pw, rho, det, rf=w.rf,
                        precision=precision)
    c = O.ptop_fid(w.system, w.wr, w.rotor, w.steps, npts, pw, rho, det, rf=w.rf)
    return float(np.linalg.norm(g - c) / np.linalg.norm(c))


def accuracy(wls, precision, limit_spins=6, ncryst=2):
    """rel-L2 of the engine vs the FP64 oracle on a few crystallites of the small clusters."""
    from oracle import oracle as O
    worst = 0.0
    seen = set()
    for w in wls:
        S = w.meta["spins"]
        if S > limit_spins or S in seen:
            continue
        seen.add(S)
        pw = w.powder().subset(np.arange(ncryst) * 997 % w.powder_size)
        rho, det = w.operators()
        g, _ = engine.powder_fid(w.system, w.wr, w.rotor, w.sw, w.npts, pw, rho, det,
                                 precision=precision)
        c, _ = O.powder_fid(w.system, w.wr, w.rotor, w.sw, w.npts, pw, rho, det)
        worst = max(worst, float(np.linalg.norm(g - c) / np.linalg.norm(c)))
    return worst


def bench_serve(args):
    """Serving mix: Poisson arrivals at 200 requests/s of seeded 1H clusters (2..10 spins,
    log-uniform), MAS 5-25 kHz, ZCW 1154 or 3722, 32-128 steps per period, 256-2048 points;
    the first --requests requests of the seeded stream, admitted by device-memory footprint
    onto 4 streams.  Reports FIDs/s over the makespan and per-request latency percentiles."""
    from blochlib_amd import serve
    from oracle import oracle as O
    reqs = serve.config5_requests(args.requests, max_spins=args.max_spins)
    cs = sum(len(r.powder) * int(round(r.sw / r.wr)) for r in reqs)
    print(f"[bench] serving {len(reqs)} requests ({cs / 1e6:.1f} M crystallite-steps)", flush=True)
    srv = serve.FIDServer(streams=4, precision=args.precision, policy=args.serve_policy,
                          slice_flops=args.slice_flops if args.slice_flops > 0 else float("inf"))
    warm = serve.config5_requests(2, seed=999, max_spins=4)
    srv.run(warm)
    t = time.perf_counter()
    done = srv.run(reqs, on_done=lambda r: print(f"[bench] request {r.rid} S={r.system.spins} "
                                                   f"arrived {r.arrival:.2f}s admitted "
                                                   f"{r.t_admit:.2f}s done at {r.t_done:.2f}s",
                                                   flush=True))
    span = time.perf_counter() - t
    switches = srv.switches
    srv.close()
    lat = np.array([r.latency for r in done]) * 1e3
    # accuracy: the smallest requests against the oracle
    worst = 0.0
    for r in sorted(done, key=lambda r: r.system.N * len(r.powder))[:3]:
        ref, _ = O.powder_fid(r.system, r.wr, r.rotor, r.sw, r.npts, r.powder, r.rho, r.det)
        worst = max(worst, float(np.linalg.norm(r.fid - ref) / np.linalg.norm(ref)))
    out = {
        "metric": METRIC, "value": cs / span, "unit": "crystallite-steps/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": span * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": ("serving mix: Poisson 200 req/s, seeded 1H clusters of 2-"
                                f"{args.max_spins} spins (log-uniform), MAS 5-25 kHz, ZCW 1154/3722, "
                                "32-128 steps/period, 256-2048 points; first "
                                f"{len(reqs)} requests of the seeded stream"),
                   "requests": len(reqs), "streams": 4},
        "fids_per_s": len(done) / span,
        "latency_ms_per_fid": {"p50": float(np.percentile(lat, 50)),
                               "p99": float(np.percentile(lat, 99)),
                               "note": "arrival to result on the host, queueing included"},
        "latency_ms_by_spins": {
            f"{lo}-{hi}": {"requests": int(m.sum()), "p50": float(np.percentile(lat[m], 50)),
                           "max": float(lat[m].max())}
            for lo, hi in ((2, 6), (7, 8), (9, 10))
            for m in [np.array([lo <= r.system.spins <= hi for r in done])] if m.any()},
        "slice_flops": args.slice_flops, "policy": args.serve_policy,
        "stream_switches": switches,
        "rel_l2_vs_oracle": worst,
    }
    print(json.dumps(out), flush=True)


def measure(H, wls, ptop, precision, steps, warmup, args, calibrate=False):
    """Plans for every workload item resident on the device, `warmup` untimed passes, then
    `steps` timed passes (launch to completion of every plan, HIP events per plan on its
    stream, per-kernel HIP-event times from the engine).  Plans are dealt round robin to
    `--streams` non-blocking HIP streams (default priority), each with its own workspace."""
    nst = max(1, args.streams)
    streams = []
    for _ in range(nst):
        st = ctypes.c_void_p()
        H.hipStreamCreateWithFlags(ctypes.byref(st), 1)
        streams.append(st)
    wss = [engine.Workspace(args.workspace_gb * 1e9 / nst) for _ in range(nst)]
    plans = []
    t0 = time.perf_counter()
    for w in wls:
        if ptop:
            p = engine.Plan.ptop(w.system, w.wr, w.rotor, w.steps, w.npts, w.rf,
                                 precision=precision)
        else:
            p = engine.Plan(w.system, w.wr, w.rotor, w.sw, w.npts, precision=precision)
        rho, det = w.operators()
        p.set_operators(rho, det)
        p.set_powder(w.powder())
        p.set_workspace(wss[len(plans) % nst])
        plans.append(p)
    print(f"[bench] {precision}: {len(plans)} plans resident in "
          f"{time.perf_counter() - t0:.1f}s", flush=True)
    if not ptop and args.tune_crystallites > 0:
        # launch choices timed on the leading crystallites of every plan (outside the timed
        # region, like plan creation); the same FIDs, bit for bit or within the c64 tolerance
        t1 = time.perf_counter()
        tuning = engine.autotune(plans, crystallites=args.tune_crystallites)
        tuning["seconds"] = time.perf_counter() - t1
        tuning["source"] = f"blfid_autotune in this run ({precision} plans)"
        print(f"[bench] {precision}: autotuned in {tuning['seconds']:.1f}s: {tuning}", flush=True)
        if args.save_tuning and precision == args.precision:
            with open(args.save_tuning, "w") as f:
                json.dump({"tuning": {k: tuning[k] for k in ("method", "corr", "grid", "qr_flip")},
                           "made_by": "python bench.py --tune-crystallites "
                                      f"{args.tune_crystallites} --save-tuning <file>",
                           "workload": ("batched: 64 seeded 1H clusters of 4-8 spins, MAS 15 kHz, "
                                        "128 steps/period, ZCW 3722, 1024 points; " + precision),
                           "pass_ms": tuning["pass_ms"]}, f, indent=1)
    else:
        tuning = dict(engine.tuning(), source=TUNING_SOURCE[0])

    def step(events=None):
        for i, p in enumerate(plans):
            st = streams[i % nst]
            if events is not None:
                H.hipEventRecord(events[2 * i], st)
            p.run(True, stream=st.value)
            if events is not None:
                H.hipEventRecord(events[2 * i + 1], st)

    if calibrate and args.copy_calibration:
        copy_calibration(H, args.copy_calibration)
    for i in range(warmup):
        step()
        H.hipDeviceSynchronize()
        print(f"[bench] {precision}: warmup {i + 1}/{warmup} done", flush=True)
    for p in plans:
        p.set_profiling(True)
    evs = []
    for _ in range(2 * len(plans)):
        e = ctypes.c_void_p()
        H.hipEventCreate(ctypes.byref(e))
        evs.append(e)
    lat = []
    H.hipDeviceSynchronize()
    t_start = time.perf_counter()
    for s in range(steps):
        step(evs)
        H.hipDeviceSynchronize()
        for i in range(len(plans)):
            ms = ctypes.c_float()
            H.hipEventElapsedTime(ctypes.byref(ms), evs[2 * i], evs[2 * i + 1])
            lat.append(ms.value)
        print(f"[bench] {precision}: step {s + 1}/{steps} done", flush=True)
    H.hipDeviceSynchronize()
    elapsed = time.perf_counter() - t_start
    # per-kernel times over the timed steps
    agg = {}
    for p in plans:
        for k, v in p.phase_times().items():
            a = agg.setdefault(k, {"ms": 0.0, "launches": 0, "flops": 0.0, "bytes": 0.0})
            for q in a:
                a[q] += v[q]
    fids = [p.fetch() for p in plans]
    for p in plans:
        p.close()
    for ws in wss:
        ws.close()
    for e in evs:
        H.hipEventDestroy(e)
    for st in streams:
        H.hipStreamDestroy(st)
    return {"elapsed": elapsed, "lat": lat, "agg": agg, "fids": fids, "streams": nst,
            "tuning": tuning}


def dense_config3(H, precision: str, K: int = 64) -> dict:
    """The dense 10-spin configuration (N = 1024, blocks up to 252 states, MAS 20 kHz, 64 steps,
    512 points) on K evenly spaced ZCW 1154 crystallites: one warm-up, one timed run of the
    requested form and of the FP64 form, and the requested form's FID against the FP64 one."""
    w = workloads.config3()
    rho, det = w.operators()
    pw = w.powder().subset(np.arange(0, w.powder_size, w.powder_size // K)[:K])
    res = {}
    for prec in dict.fromkeys((precision, "fp64")):
        p = engine.Plan(w.system, w.wr, w.rotor, w.sw, w.npts, precision=prec)
        try:
            p.set_operators(rho, det)
            p.set_powder(pw)
            p.run()
            p.fetch()
            H.hipDeviceSynchronize()
            t = time.perf_counter()
            p.run()
            f = p.fetch()
            res[prec] = (time.perf_counter() - t, f)
        finally:
            p.close()
    el, f = res[precision]
    el64, f64 = res["fp64"]
    print(f"[bench] dense config 3: K={K} {precision} {el:.3f}s, fp64 {el64:.3f}s", flush=True)
    return {"workload": "dense 10-spin 1H cluster (N=1024, blocks up to 252 states), MAS 20 kHz, "
                        "64 steps/period, 512 points, rho=Ix, detect=I+",
            "crystallites": K, "precision": precision, "unit": "crystallite-steps/s",
            "value": K * w.steps / el, "ms": el * 1e3,
            "fp64_form_value": K * w.steps / el64,
            "rel_l2_vs_fp64_gpu": float(np.linalg.norm(f - f64) / np.linalg.norm(f64))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--clusters", type=int, default=64)
    ap.add_argument("--cpu-sample", choices=["full", "quick", "none"], default="full")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--workspace-gb", type=float, default=120.0,
                    help="device working memory shared out over the streams (GB)")
    ap.add_argument("--streams", type=int, default=1,
                    help="HIP streams the plans are dealt to, round robin")
    ap.add_argument("--workload", choices=["batched", "ptop", "serve"], default="batched")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="one device copy of MIB MiB before the warm-up (counter calibration)")
    ap.add_argument("--requests", type=int, default=48, help="requests of the serving mix")
    ap.add_argument("--max-spins", type=int, default=10, help="largest cluster of the serving mix")
    ap.add_argument("--slice-flops", type=float, default=4e12,
                    help="serving: work per powder slice of a request (0: whole requests)")
    ap.add_argument("--serve-policy", choices=["srpt", "fifo"], default="srpt",
                    help="serving: streams go to the least remaining work (switched at slice "
                         "boundaries) or to requests in arrival order")
    ap.add_argument("--sets", type=int, default=8, help="parameter sets per step (ptop)")
    ap.add_argument("--precision", choices=["fp64", "c64", "bf16x3"], default="c64",
                    help="complex64 propagators/products with FP64 eigen-solver, phases and "
                         "reductions (solid-2.0's form), complex128 throughout, or the complex64 "
                         "form with its complex products as three bf16 products each")
    ap.add_argument("--no-fp64-form", action="store_true",
                    help="skip timing the FP64 form beside the c64 headline")
    ap.add_argument("--tune-crystallites", type=int, default=0,
                    help="run the launch autotuner on (at least) this many leading crystallites "
                         "of every plan before the warm-up (0: use the saved table, "
                         "blochlib_amd/tuning/mi355x.json, or the defaults without it)")
    ap.add_argument("--save-tuning", default="", metavar="FILE",
                    help="write the table the autotuner chose (with --tune-crystallites) as JSON")
    ap.add_argument("--tuning", default="",
                    help="launch choices as JSON (the engine.tuning() dict), set instead of the "
                         "saved table")
    ap.add_argument("--dense-k", type=int, default=64,
                    help="crystallites of the dense 10-spin configuration timed beside the "
                         "headline (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not implemented "
              "(crystallites would shard across GPUs with one P-point all-reduce).",
              file=sys.stderr)
        sys.exit(2)

    if args.tuning:
        engine.set_tuning(json.loads(args.tuning))
        TUNING_SOURCE[0] = "--tuning"
    elif args.tune_crystallites <= 0 and os.path.exists(TUNING_FILE):
        with open(TUNING_FILE) as f:
            engine.set_tuning(json.load(f)["tuning"])
        TUNING_SOURCE[0] = os.path.relpath(TUNING_FILE, ROOT) + " (saved blfid_autotune table)"
    if args.workload == "serve":
        return bench_serve(args)
    H = hip()
    ptop = args.workload == "ptop"
    wls = workloads.config4(args.sets) if ptop else workloads.config2(args.clusters)
    cs_per_step = sum(w.powder_size * w.steps for w in wls)
    m = measure(H, wls, ptop, args.precision, args.steps, args.warmup, args, calibrate=True)
    elapsed, lat, agg = m["elapsed"], m["lat"], m["agg"]
    nst = m["streams"]
    fams = {}
    for k, v in agg.items():
        f = fams.setdefault(family(k), {"ms": 0.0, "launches": 0, "flops": 0.0, "bytes": 0.0})
        for q in f:
            f[q] += v[q]
    # the dominant kernel family by summed kernel time over the timed steps
    dom = max(fams, key=lambda k: fams[k]["ms"])
    roof = family_roofline(dom, fams[dom], args.precision)
    traffic = None if ptop else profiled_traffic(dom, args.precision)
    value = cs_per_step * args.steps / elapsed
    peak = FP64_PEAK_TFLOPS if args.precision == "fp64" else FP32_PEAK_TFLOPS
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "crystallite-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": ({
            "workload": ("heteronuclear rf: 13C + 6 1H (N=128), MAS 12.5 kHz, magic angle, "
                         "256 steps/period, 100 kHz 0/180 phase-alternating 1H rf (8 cycles per "
                         "period), ZCW 1154, 2048 rotor-synchronised points, rho=Ix_0, "
                         "detect=Ip_0, one FID per seeded CSA set"),
            "precision": args.precision,
            "sets": len(wls),
            "streams": nst,
            "crystallites": wls[0].powder_size,
            "steps_per_period": wls[0].steps,
            "npts": wls[0].npts,
            "crystallite_steps_per_step": cs_per_step,
        } if ptop else {
            "workload": ("batched uneven sizes: 64 seeded 1H clusters of 4-8 spins (N=16-256), "
                         "MAS 15 kHz, magic angle, 128 steps/period, ZCW 3722, 1024-point FID, "
                         "rho=Ix, detect=I+"),
            "precision": args.precision,
            "clusters": len(wls),
            "streams": nst,
            "crystallites": wls[0].powder_size,
            "steps_per_period": wls[0].steps,
            "npts": wls[0].npts,
            "crystallite_steps_per_step": cs_per_step,
        }),
        "fids_per_s": len(wls) * args.steps / elapsed,
        "latency_ms_per_fid": {"p50": float(np.percentile(lat, 50)),
                               "p99": float(np.percentile(lat, 99)),
                               "note": ("per-plan (cluster or parameter set) powder FID time on its stream "
                                        "inside a step (plans on other streams run alongside)")},
        "kernels": {k: {"ms_total": v["ms"], "launches": v["launches"],
                        "tflops": v["flops"] / max(v["ms"], 1e-9) / 1e9}
                    for k, v in sorted(agg.items(), key=lambda kv: -kv[1]["ms"])},
        "roofline": roof,
        "tuning": m.get("tuning"),
        "families": {k: family_roofline(k, v, args.precision)
                     for k, v in sorted(fams.items(), key=lambda kv: -kv[1]["ms"])},
    }
    if traffic:
        # counter-measured HBM bytes per launch of the family (committed profile of this
        # command) against its algorithmic bytes per launch: re-reads and spills show above 1
        roof["traffic"] = traffic["bytes"]
        roof["traffic_ratio"] = traffic["bytes"] / max(roof["alg_bytes_per_launch"], 1.0)
        roof["traffic_detail"] = traffic
    if dom == "k_corr":
        n, P = wls[0].steps, wls[0].npts
        d = fams[dom]
        ratio = (14.0 * n * n + 14.0 * P) / corr_flops_per_elem(n, P, corr_is_autocorrelation(wls))
        roof["canonical_flops_per_launch"] = ratio * d["flops"] / d["launches"]
        roof["canonical_tflops"] = ratio * d["flops"] / (d["ms"] * 1e-3) / 1e12
        roof["canonical_note"] = (
            "reference algorithm's count for this kernel (14 n^2 + 14 P flops per correlation "
            "element and crystallite: direct cyclic sum with a complex exp-multiply per term and "
            "the FID recurrence); the kernel executes it with FFTs")
    if args.precision != "fp64" and not args.no_fp64_form:
        # the plain FP64 form on the same workload (one timed step) and the c64 FIDs against it
        m64 = measure(H, wls, ptop, "fp64", 1, 1, args)
        errs = [float(np.linalg.norm(a - b) / np.linalg.norm(b))
                for a, b in zip(m["fids"], m64["fids"])]
        out["fp64_form"] = {
            "value": cs_per_step / m64["elapsed"],
            "ms_per_step": m64["elapsed"] * 1e3,
            "rel_l2_c64_vs_fp64_gpu": {"max": max(errs), "median": float(np.median(errs)),
                                       "fids": len(errs),
                                       "note": "every FID of the step at full size"},
        }
    if not ptop and args.dense_k > 0:
        out["dense_config3"] = dense_config3(H, args.precision, args.dense_k)
    if not args.no_accuracy:
        out["rel_l2_vs_oracle"] = (accuracy_ptop(wls, args.precision) if ptop
                                   else accuracy(wls, args.precision))
    if args.cpu_sample != "none":
        cb = cpu_baseline_ptop(wls[0]) if ptop else cpu_baseline(wls, args.cpu_sample)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample",
                                                  "cpu_model")}
        out["speedup_vs_cpu_core"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
