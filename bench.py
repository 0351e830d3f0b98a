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

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
DTYPE = {"fp64": "f64", "c64": "c64 (f32 products, f64 eig/phases/sums)",
         "bf16x3": "c64 (bf16x3 split products, f32 exp, f64 eig/phases/sums)"}
METRIC = "crystallite-steps/s + FIDs/s at ≤1e-5 rel L2 vs FP64 oracle; p50/p99 latency"
# counter summaries of this benchmark's default command (FETCH_SIZE and WRITE_SIZE passes)
COUNTER_SUMMARY = os.path.join(ROOT, "profiles", "config2_batched_{prec}_hbm_counters.txt")


def corr_flops_per_elem(n: int, P: int) -> float:
    """Executed flops of k_corr per correlation element and crystallite (as the engine counts
    them): for n a power of two from 16 to 512 three mixed-radix FFTs (radix-8 passes of a
    56-flop DFT and 7 twiddle products per butterfly, the last pass of radix 2, 4 or 8 without
    twiddles) and the scaled product, else the direct sum; plus twiddle staging and the FID
    sums."""
    c = 8.0 * n * n
    if 16 <= n <= 512 and (n & (n - 1)) == 0:
        ln = int(n).bit_length() - 1
        npass = (ln + 2) // 3
        rl = 1 << (ln - 3 * (npass - 1))
        dft = {2: 4.0, 4: 16.0, 8: 56.0}[rl]
        c = 3.0 * (npass - 1) * (n / 8.0) * 98.0 + 3.0 * (n / rl) * dft + 8.0 * n
    return c + 18.0 * n + 14.0 * P


def corr_bytes_per_elem_step(wls, precision: str) -> float:
    """Bytes k_corr must read per correlation element and step: Q in the storage precision,
    plus the same again when R is not conj(Q) alone -- rho is not (det + det^H) / 2, or det has
    entries in both (a, b) and (b, a) so that some elements have a reverse element."""
    cs = 16.0 if precision == "fp64" else 8.0
    for w in wls:
        rho, det = w.operators()
        scale = max(np.abs(det).max(), 1.0)
        rosym = np.abs(rho - 0.5 * (det + det.conj().T)).max() <= 1e-12 * scale
        if not rosym or np.any((det != 0) & (det.T != 0)):
            return 2.0 * cs
    return cs


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
    """Time the oracle on one crystallite of a representative cluster per spin count,
    each sample on its own thread (single-threaded oracle), and extrapolate the config's
    crystallite-steps/s per core from the cluster mix."""
    from oracle import oracle as O
    reps = {}
    for w in wls:
        reps.setdefault(w.meta["spins"], w)
    counts = {}
    for w in wls:
        counts[w.meta["spins"]] = counts.get(w.meta["spins"], 0) + 1
    spins = sorted(reps)
    measure = spins if sample_mode == "full" else [s for s in spins if s <= 7]

    def one(S):
        w = reps[S]
        pw = w.powder().subset([0])
        rho, det = w.operators()
        t = time.perf_counter()
        O.powder_fid(w.system, w.wr, w.rotor, w.sw, w.npts, pw, rho, det, nthreads=1)
        return S, time.perf_counter() - t

    with cf.ThreadPoolExecutor(max_workers=len(measure)) as ex:
        t_s = dict(ex.map(one, measure))
    if sample_mode != "full":  # extrapolate the largest spin counts by N^3
        for S in spins:
            if S not in t_s:
                t_s[S] = t_s[max(t_s)] * 8.0 ** (S - max(t_s))
    n = wls[0].steps
    K = wls[0].powder_size
    cs_total = sum(counts[S] * K * n for S in spins)
    sec_total = sum(counts[S] * K * t_s[S] for S in spins)
    return {
        "value": cs_total / sec_total,
        "unit": "crystallite-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": ("oracle (reference gamma-COMPUTE algorithm in C, FP64, dense N x N) on 1 "
                   f"crystallite of the first cluster with S spins for S in {measure}, each on "
                   "its own thread; rate extrapolated over the 64-cluster mix by spin count "
                   f"(seconds per crystallite: { {s: round(v, 3) for s, v in t_s.items()} })"),
        "seconds_per_crystallite": {str(k): v for k, v in t_s.items()},
        "fids_per_s": len(wls) / sec_total,
    }


def cpu_baseline_ptop(w):
    """Oracle (explicit Dyson product with rf, then rho <- U rho U^H per point, FP64, one
    thread) on one crystallite of one parameter set: the period propagator and the per-point
    cost are timed separately (16 and 48 points) and combined for the full point count."""
    from oracle import oracle as O
    pw = w.powder().subset([0])
    rho, det = w.operators()
    t = []
    for npts in (16, 48):
        t0 = time.perf_counter()
        O.ptop_fid(w.system, w.wr, w.rotor, w.steps, npts, pw, rho, det, rf=w.rf)
        t.append(time.perf_counter() - t0)
    per_point = max((t[1] - t[0]) / 32.0, 0.0)
    fixed = max(t[0] - 16 * per_point, 0.0)
    sec = fixed + w.npts * per_point
    return {
        "value": w.steps / sec,
        "unit": "crystallite-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": ("oracle (explicit rf Dyson product + point-by-point U rho U^H, FP64, dense "
                   f"N=128) on 1 crystallite of parameter set 0, timed at 16 and 48 points: "
                   f"{fixed:.3f} s for the period propagator + {per_point * 1e3:.2f} ms per "
                   f"point, combined for {w.npts} points = {sec:.2f} s per crystallite"),
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
    srv = serve.FIDServer(streams=4, precision=args.precision)
    warm = serve.config5_requests(2, seed=999, max_spins=4)
    srv.run(warm)
    t = time.perf_counter()
    done = srv.run(reqs, on_done=lambda r: print(f"[bench] request {r.rid} S={r.system.spins} "
                                                   f"done at {r.t_done:.2f}s", flush=True))
    span = time.perf_counter() - t
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
            a = agg.setdefault(k, {"ms": 0.0, "launches": 0, "flops": 0.0})
            a["ms"] += v["ms"]
            a["launches"] += v["launches"]
            a["flops"] += v["flops"]
    fids = [p.fetch() for p in plans]
    for p in plans:
        p.close()
    for ws in wss:
        ws.close()
    for e in evs:
        H.hipEventDestroy(e)
    for st in streams:
        H.hipStreamDestroy(st)
    return {"elapsed": elapsed, "lat": lat, "agg": agg, "fids": fids, "streams": nst}


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
    ap.add_argument("--sets", type=int, default=8, help="parameter sets per step (ptop)")
    ap.add_argument("--precision", choices=["fp64", "c64", "bf16x3"], default="c64",
                    help="complex64 propagators/products with FP64 eigen-solver, phases and "
                         "reductions (solid-2.0's form), complex128 throughout, or the complex64 "
                         "form with its complex products as three bf16 products each")
    ap.add_argument("--no-fp64-form", action="store_true",
                    help="skip timing the FP64 form beside the c64 headline")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not implemented "
              "(crystallites would shard across GPUs with one P-point all-reduce).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "serve":
        return bench_serve(args)
    H = hip()
    ptop = args.workload == "ptop"
    wls = workloads.config4(args.sets) if ptop else workloads.config2(args.clusters)
    cs_per_step = sum(w.powder_size * w.steps for w in wls)
    m = measure(H, wls, ptop, args.precision, args.steps, args.warmup, args, calibrate=True)
    elapsed, lat, agg = m["elapsed"], m["lat"], m["agg"]
    nst = m["streams"]
    dom = max(agg, key=lambda k: agg[k]["ms"])
    traffic = None if ptop else profiled_traffic(dom, args.precision)
    d = agg[dom]
    avg_ms = d["ms"] / d["launches"]
    achieved = d["flops"] / d["launches"] / (avg_ms * 1e-3) / 1e12
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
        "roofline": {
            "kernel": dom,
            "bound": "mfma",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic["bytes"] if traffic else None,
            "avg_launch_ms": avg_ms,
            "flops_per_launch": d["flops"] / d["launches"],
            "note": ("achieved = executed algorithmic flops of the kernel (block-diagonal "
                     "algorithm, unpadded sizes) per launch / its average launch time (HIP events "
                     "on the launch stream); peak = MI355X "
                     + ("FP64" if args.precision == "fp64" else "FP32") + " dense"),
        },
    }
    if traffic:
        out["roofline"]["traffic_detail"] = traffic
        out["roofline"]["achieved_hbm_gbs"] = traffic["bytes"] / (avg_ms * 1e-3) / 1e9
    if dom == "k_corr":
        n, P = wls[0].steps, wls[0].npts
        # algorithmic HBM bytes: every step of every correlation element is read once as Q, and
        # a second time (the reverse element of Q, or the stored R) unless R = conj(Q) there
        elems = d["flops"] / d["launches"] / corr_flops_per_elem(n, P)
        per_step = corr_bytes_per_elem_step(wls, args.precision)
        alg_bytes = elems * n * per_step
        if d["flops"] / d["launches"] / alg_bytes < peak * 1e12 / (HBM_PEAK_GBS * 1e9):
            # below the ridge point: HBM bound
            out["roofline"].update({
                "bound": "hbm", "achieved": alg_bytes / (avg_ms * 1e-3) / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": alg_bytes / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "bytes_per_launch": alg_bytes, "achieved_tflops": achieved,
                "flop_peak_tflops": peak,
                "note": ("achieved = algorithmic HBM bytes per launch (%g B per correlation "
                         "element-step) / the average launch time (HIP events on the launch "
                         "stream); the kernel's executed flops per algorithmic byte are below "
                         "MI355X's ridge point, so HBM (8 TB/s) is its roof; achieved_tflops = "
                         "executed flops rate" % per_step),
            })
        else:
            out["roofline"]["bytes_per_launch"] = alg_bytes
            out["roofline"]["note"] += (
                "; algorithmic HBM bytes %g per correlation element-step (%.3g GB per launch): "
                "executed flops per byte above the ridge point" % (per_step, alg_bytes / 1e9))
        ratio = (14.0 * n * n + 14.0 * P) / corr_flops_per_elem(n, P)
        out["roofline"]["canonical_flops_per_launch"] = ratio * d["flops"] / d["launches"]
        out["roofline"]["canonical_tflops"] = ratio * achieved
        out["roofline"]["canonical_note"] = (
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
    if not args.no_accuracy:
        out["rel_l2_vs_oracle"] = (accuracy_ptop(wls, args.precision) if ptop
                                   else accuracy(wls, args.precision))
    if args.cpu_sample != "none":
        cb = cpu_baseline_ptop(wls[0]) if ptop else cpu_baseline(wls, args.cpu_sample)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["speedup_vs_cpu_core"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
