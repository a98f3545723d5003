#!/usr/bin/env python3
"""Benchmark of the split-step Crank-Nicolson dipolar GPE solver on one MI355X.

Workload (default): plain imaginary-time FP64 solve at 256^3, dx = dy = dz =
0.1 (box [-12.8, 12.8)^3), dt = 5e-4, gamma/nu/lambda = 0.5/1/1.5, G0 = 40,
GDD0 = 30 (OPTION = 2: g = 80, gd = 60), spherical cut-off R = 12.8, Gaussian
initial wave function.  A step is one full split step (dipolar convolution,
calcnu, x/y/z Crank-Nicolson sweeps, Simpson norm and renormalisation) over
the whole grid with psi resident in HBM.

Prints one JSON line: grid-pt-steps/s = Nx*Ny*Nz*K / t over K timed steps,
the roofline of the dominant kernel (algorithmic bytes / mean launch time,
measured with HIP events on the solver's stream) and the CPU baseline (the
FP64 oracle on the host cores, bounded sample of the same workload).  With
the default workload it adds the same grid in mixed precision ("mixed") and
the serving stream ("serving": the first --serve-requests requests replayed
through the request loop, all at once for capacity and at their Poisson
arrival times for latency, and the same requests as back-to-back plain FP64
solves).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "dbec-gp-mpi-cuda_amd")
sys.path.insert(0, os.path.join(PKG, "python"))

METRIC = "grid-pt-steps/s at 256³ (FP64 vs mixed) + solves/s, p50/p99 at |Δμ/μ|≤1e-8"
HBM_PEAK_GBS = 8000.0
STREAM_FIXTURE = os.path.join(REPO, "tests", "golden", "serve_stream.json")
WORKLOADS = {
    # name: (n, h, dt, cutoff, mode, prec)
    "imag256_fp64": ((256, 256, 256), 0.1, 0.0005, 12.8, "imag", "fp64"),
    "imag256_mixed": ((256, 256, 256), 0.1, 0.0005, 12.8, "imag", "mixed"),
    "real256x384_fp64": ((256, 256, 384), 0.1, 0.0002, 12.8, "real", "fp64"),
    # serving-stream sizes (dx = 16/N, cutoff 8), for tuning runs
    "real192_mixed": ((192, 192, 192), 16 / 192, 0.001, 8.0, "real", "mixed"),
    "real192_fp64": ((192, 192, 192), 16 / 192, 0.001, 8.0, "real", "fp64"),
    "imag192_mixed": ((192, 192, 192), 16 / 192, 0.01, 8.0, "imag", "mixed"),
    "imag192_fp64": ((192, 192, 192), 16 / 192, 0.01, 8.0, "imag", "fp64"),
    "real160_mixed": ((160, 160, 160), 16 / 160, 0.001, 8.0, "real", "mixed"),
    "imag160_mixed": ((160, 160, 160), 16 / 160, 0.01, 8.0, "imag", "mixed"),
    "imag160_fp64": ((160, 160, 160), 16 / 160, 0.01, 8.0, "imag", "fp64"),
    "real96_mixed": ((96, 96, 96), 16 / 96, 0.001, 8.0, "real", "mixed"),
    "imag96_mixed": ((96, 96, 96), 16 / 96, 0.01, 8.0, "imag", "mixed"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def b_alg_per_point(mode, prec, nz):
    """Algorithmic HBM bytes per grid point per split step of the reference
    operator sequence: each operator one streaming pass, one half-spectrum
    read+write per FFT axis with the x forward * potdd * inverse fused, Thomas
    intermediates on chip.  imag: 15e + 8ch + eh; real: 21e + 8ch + eh, with
    e the real element size, c the complex one, h = (Nz/2+1)/Nz."""
    e = 8 if prec == "fp64" else 4
    c = 2 * e
    h = (nz // 2 + 1) / nz
    base = 15 * e if mode == "imag" else 21 * e
    return base + 8 * c * h + e * h


def host_description(threads):
    """The host the CPU baseline runs on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable, "omp_num_threads": threads,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "")}


def cpu_baseline(n, h, dt, cutoff, mode, steps, threads, coupling=(80.0, 60.0)):
    """The FP64 oracle (OpenMP) on the host: `steps` split steps of the
    workload (Gaussian psi0), timed around the step loop only."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    import dbecpy

    lib = dbecpy.load(os.path.join(REPO, "oracle", "liboracle.so"))
    p = dbecpy.make_params(n[0], n[1], n[2], h, h, h, dt, 0.5, 1.0, 1.5, cutoff)
    m = dbecpy.IMAG if mode == "imag" else dbecpy.REAL
    with dbecpy.Solver(lib, p, m) as s:
        s.set_psi(dbecpy.gaussian_psi0(p))
        s.calcnorm(rescale=(m == dbecpy.IMAG))
        s.set_coupling(*coupling)
        s.step(1)  # first touch of every array
        t0 = time.perf_counter()
        s.step(steps)
        t = time.perf_counter() - t0
    return n[0] * n[1] * n[2] * steps / t, t


def copy_calibration(nbytes):
    """One device-to-device copy of `nbytes` (the runtime's blit kernel), the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def percentiles(xs):
    import numpy as np
    a = np.asarray(xs, dtype=float)
    return {"p50": float(np.percentile(a, 50)), "p95": float(np.percentile(a, 95)),
            "p99": float(np.percentile(a, 99))}


def serve_bench(lib, count, baseline_count, rate, max_batch, check_every, timeout, capacity_only=False,
                polish_tol=None):
    """The serving stream (first `count` requests, Poisson arrivals at `rate`
    per second, mixed precision) replayed in real time against the request
    loop, and its first `baseline_count` requests as back-to-back plain FP64
    solves (one at a time through dbec.h, arrivals queued FIFO)."""
    import numpy as np
    import dbecpy
    import dbecserve as ds

    stream = ds.load_stream(STREAM_FIXTURE, count, prec=dbecpy.MIXED)
    if rate != 20.0:  # the fixture's arrivals are at 20 requests/s
        stream = [(t * 20.0 / rate, r) for t, r in stream]
    # capacity: the same requests all submitted at once
    if polish_tol is None:
        polish_tol = ds.DEFAULT_POLISH_TOL
    kw = {"polish_tol": polish_tol}
    with ds.Server(lib, device=0, max_batch=max_batch, check_every=check_every, **kw) as srv:
        t0 = time.perf_counter()
        srv.run_all([r for _, r in stream], timeout=timeout)
        cap_wall = time.perf_counter() - t0
    log(f"serving capacity: {count} requests at once in {cap_wall:.2f} s = {count / cap_wall:.2f} solves/s")
    if capacity_only:
        return {"requests": count, "precision": "mixed", "max_batch": max_batch, "check_every": check_every,
                "polish_tol": polish_tol, "capacity_solves_per_s": count / cap_wall, "capacity_wall_s": cap_wall}
    with ds.Server(lib, device=0, max_batch=max_batch, check_every=check_every, **kw) as srv:
        t0 = time.perf_counter()
        res = ds.serve_stream(srv, stream, timeout=timeout)
        wall = time.perf_counter() - t0
    lat = [r["latency"] for r in res]
    status = {}
    for r in res:
        status[r["status"]] = status.get(r["status"], 0) + 1
    out = {
        "requests": count, "rate_per_s": rate, "precision": "mixed", "max_batch": max_batch,
        "check_every": check_every, "polish_tol": polish_tol, "capacity_solves_per_s": count / cap_wall,
        "wall_s": wall, "solves_per_s": count / wall,
        "latency_s": percentiles(lat), "mean_steps": float(np.mean([r["steps"] for r in res])),
        "status_counts": {["converged", "steps_done", "not_converged", "diverged"][k]: v for k, v in status.items()},
    }
    log(f"serving: {count} requests in {wall:.2f} s = {count / wall:.2f} solves/s, "
        f"p50/p95/p99 {out['latency_s']['p50']:.2f}/{out['latency_s']['p95']:.2f}/{out['latency_s']['p99']:.2f} s")
    if baseline_count:
        sub = stream[:baseline_count]
        fp64 = [ds.make_request(r.id, r.nx, r.mode, dbecpy.FP64, r.g / r.par, r.gd / r.par,
                                (r.vgamma, r.vnu, r.vlambda), r.dt, r.max_steps, r.mutol) for _, r in sub]
        service = []
        for i, q in enumerate(fp64):
            t1 = time.perf_counter()
            ds.plain_solve(lib, q, check_every=check_every)
            service.append(time.perf_counter() - t1)
            if (i + 1) % 200 == 0 and i + 1 < len(fp64):
                log(f"back-to-back FP64: {i + 1} of {len(fp64)} requests, {sum(service):.1f} s")
        # FIFO single server: start = max(arrival, previous finish)
        fin, blat = 0.0, []
        for (arr, _), sv in zip(sub, service):
            fin = max(arr, fin) + sv
            blat.append(fin - arr)
        out["baseline_back_to_back_fp64"] = {
            "requests": baseline_count, "solves_per_s": baseline_count / sum(service),
            "latency_s": percentiles(blat), "mean_service_s": float(np.mean(service)),
        }
        out["capacity_speedup_vs_back_to_back"] = (out["capacity_solves_per_s"]
                                                   / out["baseline_back_to_back_fp64"]["solves_per_s"])
        log(f"back-to-back FP64: {baseline_count} requests, {baseline_count / sum(service):.2f} solves/s")
    return out


# Couplings of the reference's example input (AS = 100, ADD = 132.7,
# NATOMS = 5000, AHO = 1e-6: G0 = 4 pi AS N a0 / AHO, GDD0 = 3 ADD N a0 / AHO),
# whose ground state is stable (eps_dd = 1.33); the 256^3 workload's own
# couplings (G0 = 40, GDD0 = 30, eps_dd = 3.1) collapse to a grid-scale spike.
STABLE_G0, STABLE_GDD0 = 332.4918, 105.33272


def converge(s, check, tol, cap, trace=None):
    """dbec_step in chunks of `check` steps with calcmuen after each, until
    |mu - mu_prev| <= tol |mu| (or `cap` steps).  Returns (steps, mu, en)."""
    steps, prev = 0, None
    while steps < cap:
        s.step(check, want_norm=False)
        steps += check
        mu, en = s.calcmuen()
        if trace is not None:
            trace.append((steps, mu, abs(mu - prev) / abs(mu) if prev else None))
        if prev is not None and abs(mu - prev) <= tol * abs(mu):
            return steps, mu, en
        prev = mu
    return steps, mu, en


def tts_bench(lib, n, h, dt, cutoff, check, mutol, polish_tol, cap):
    """Time to solution of a ground state at the bench grid on stable
    couplings: plain FP64 from the Gaussian to |dmu/mu| <= mutol between
    checks, against the mixed form to |dmu/mu| <= polish_tol followed by an
    FP64 polish (psi handed over through the host) to the same mutol.  Both
    include their mu checks; errors of the polished result against FP64."""
    import numpy as np
    import dbecpy

    p = dbecpy.make_params(n[0], n[1], n[2], h, h, h, dt, 0.5, 1.0, 1.5, cutoff, device=0)
    g, gd = 2 * STABLE_G0, 2 * STABLE_GDD0
    psi0 = dbecpy.gaussian_psi0(p)
    out = {"grid": list(n), "dt": dt, "G0": STABLE_G0, "GDD0": STABLE_GDD0, "check_every": check,
           "mutol": mutol, "polish_tol": polish_tol}
    with dbecpy.Solver(lib, p, dbecpy.IMAG, dbecpy.FP64) as d:
        d.set_psi(psi0)
        d.calcnorm(rescale=True)
        d.set_coupling(g, gd)
        d.synchronize()
        t0 = time.perf_counter()
        k64, mu64, en64 = converge(d, check, mutol, cap)
        t64 = time.perf_counter() - t0
        rms64, psi64 = d.calcrms(), d.get_psi()
    log(f"tts: FP64 {k64} steps in {t64:.2f} s, mu/par {mu64 / 2:.12f}")
    trace = []
    with dbecpy.Solver(lib, p, dbecpy.IMAG, dbecpy.MIXED) as m:
        m.set_psi(psi0)
        m.calcnorm(rescale=True)
        m.set_coupling(g, gd)
        m.synchronize()
        t0 = time.perf_counter()
        km, mum, _ = converge(m, check, polish_tol, cap, trace)
        psim = m.get_psi()
        tm = time.perf_counter() - t0
    with dbecpy.Solver(lib, p, dbecpy.IMAG, dbecpy.FP64) as d:
        t0 = time.perf_counter()
        d.set_psi(psim)
        d.calcnorm(rescale=True)
        d.set_coupling(g, gd)
        kp, mup, enp = converge(d, check, mutol, cap)
        tp = time.perf_counter() - t0
        rmsp, psip = d.calcrms(), d.get_psi()
    log(f"tts: mixed {km} steps in {tm:.2f} s (mu/par {mum / 2:.12f}), FP64 polish {kp} steps in {tp:.2f} s, "
        f"mu/par {mup / 2:.12f}; speedup {t64 / (tm + tp):.2f}x")
    out.update({
        "fp64": {"steps": k64, "seconds": t64, "mu_over_par": mu64 / 2, "E_over_par": en64 / 2},
        "mixed": {"steps": km, "seconds": tm, "mu_over_par": mum / 2,
                  "dmu_trace": [[k, r] for k, _, r in trace if r is not None][-8:]},
        "polish": {"steps": kp, "seconds": tp, "mu_over_par": mup / 2, "E_over_par": enp / 2},
        "tts_speedup": t64 / (tm + tp),
        "error_vs_fp64": {
            "mu": abs(mup - mu64) / abs(mu64), "E": abs(enp - en64) / abs(en64),
            "rms": float(np.max(np.abs(np.asarray(rmsp) - np.asarray(rms64)) / np.abs(np.asarray(rms64)))),
            "max_dpsi_over_max_psi": float(np.max(np.abs(psip - psi64)) / np.max(np.abs(psi64))),
            "mixed_phase_mu": abs(mum - mu64) / abs(mu64),
        },
    })
    return out


def traffic_for(workload, kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (profiles/traffic_<workload>.json), or None."""
    path = os.path.join(REPO, "profiles", f"traffic_{workload}.json")
    try:
        d = json.load(open(path))
        k = d["kernels"][kernel]
        return k["fetch_bytes"] + k["write_bytes"]
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="imag256_fp64", choices=sorted(WORKLOADS))
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu-cfg0", type=int, default=1, help="also time the 64^3 x 2000-step CPU configuration")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed-steps", type=int, default=1000,
                    help="also time the same grid in mixed precision for this many steps (imag256 only)")
    ap.add_argument("--tts", type=int, default=1, help="time to solution, FP64 vs mixed + FP64 polish (imag256)")
    ap.add_argument("--tts-check", type=int, default=100)
    ap.add_argument("--tts-polish-tol", type=float, default=1e-7)
    ap.add_argument("--tts-cap", type=int, default=60000)
    ap.add_argument("--serve-requests", type=int, default=200, help="requests of the serving stream to replay")
    ap.add_argument("--serve-baseline", type=int, default=200, help="of those, run back to back in plain FP64")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-batch", type=int, default=16)
    ap.add_argument("--serve-check", type=int, default=50)
    ap.add_argument("--serve-polish-tol", type=float, default=None,
                    help="request loop polish_tol (default: the binding's)")
    ap.add_argument("--serve-capacity-only", action="store_true",
                    help="run only the serving capacity measurement (the first --serve-requests requests at once)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    rank = int(os.environ.get("RANK", "0"))
    if args.gpus != 1 or world != 1:
        if rank == 0:
            print("bench.py: this build runs on one GPU; a single solve does not shard across GPUs here "
                  "(see DESIGN.md, 'How the path would shard'). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import dbecpy

    if args.serve_capacity_only:
        lib = dbecpy.load(dbecpy.HIP_LIB)
        out = serve_bench(lib, args.serve_requests, 0, args.serve_rate, args.serve_batch, args.serve_check,
                          timeout=600.0, capacity_only=True, polish_tol=args.serve_polish_tol)
        print(json.dumps({"serving_capacity": out}), flush=True)
        return
    if args.copy_calibration_mb:
        copy_calibration(args.copy_calibration_mb << 20)
        log(f"copy calibration: {args.copy_calibration_mb} MiB device-to-device")
    n, h, dt, cutoff, mode, prec = WORKLOADS[args.workload]
    lib = dbecpy.load(dbecpy.HIP_LIB)
    p = dbecpy.make_params(n[0], n[1], n[2], h, h, h, dt, 0.5, 1.0, 1.5, cutoff, device=0)
    m = dbecpy.IMAG if mode == "imag" else dbecpy.REAL
    pr = dbecpy.FP64 if prec == "fp64" else dbecpy.MIXED
    npts = n[0] * n[1] * n[2]
    log(f"workload {args.workload}: grid {n}, {mode} time, {prec}")
    s = dbecpy.Solver(lib, p, m, pr)
    tuning = s.tuning()
    log(f"launch choices: {tuning}")
    s.set_psi(dbecpy.gaussian_psi0(p))
    s.calcnorm(rescale=(m == dbecpy.IMAG))
    s.set_coupling(80.0, 60.0)
    s.step(args.warmup, want_norm=False)
    s.synchronize()
    log(f"warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    s.step(args.steps, want_norm=False)
    s.synchronize()
    t = time.perf_counter() - t0
    log(f"timed {args.steps} steps: {t:.3f} s")
    value = npts * args.steps / t
    ms_step = 1e3 * t / args.steps
    norm = s.calcnorm()
    mu, en = s.calcmuen()
    # per-kernel HIP-event timing right after the timed region (the device at
    # its working clocks): every launch of a few steps repeated
    # --profile-reps times back to back on the solver's stream, so each
    # figure is the pipelined time of one launch.  The repeats alter psi, so
    # this comes after the observables.
    stats = s.profile(args.profile_reps)
    s.close()
    log("per-kernel: " + ", ".join(f"{k['name']} {k['avg_ms'] * 1e3:.1f}us" for k in stats))
    dom = max(stats, key=lambda k: k["avg_ms"] * k["launches"])
    achieved = dom["alg_bytes"] / (dom["avg_ms"] * 1e-3) / 1e9
    step_kernel_ms = sum(k["avg_ms"] * k["launches"] for k in stats)
    # kernels that run once per dbec_step call (the first forward z pass, the
    # last z sweep, the final norm and rescale; launches 0) count once per
    # timed call
    step_bytes = sum(k["alg_bytes"] * (k["launches"] if k["launches"] >= 1 else 1.0 / args.steps) for k in stats)

    balg = b_alg_per_point(mode, prec, n[2])
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{mode}-time dipolar GPE split step, {n[0]}x{n[1]}x{n[2]}, {prec}",
            "grid": list(n), "dx": h, "dt": dt, "trap": [0.5, 1.0, 1.5], "G0": 40.0, "GDD0": 30.0,
            "cutoff": cutoff, "option": 2, "psi0": "Gaussian",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom["name"],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_for(args.workload, dom["name"]),
            "alg_bytes_per_launch": dom["alg_bytes"],
            "avg_launch_ms": dom["avg_ms"],
        },
        # the whole step: the bytes its kernels move (algorithmic bytes of
        # this design's passes, per step) over the timed step time; and the
        # reference operator sequence's B_alg over the same time, an effective
        # rate only (it counts passes this design does not make)
        "step_roofline": {
            "bytes_per_step": step_bytes,
            "bytes_per_point": step_bytes / npts,
            "achieved_gbs": step_bytes / (ms_step * 1e-3) / 1e9,
            "frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "kernel_sum_ms": step_kernel_ms,
            "b_alg_bytes_per_point": balg,
            "b_alg_effective_gbs": balg * npts / (ms_step * 1e-3) / 1e9,
        },
        "kernels": {k["name"]: {"avg_us": k["avg_ms"] * 1e3, "gbs": k["alg_bytes"] / (k["avg_ms"] * 1e-3) / 1e9,
                                "launches_per_step": k["launches"]} for k in stats},
        "observables": {"norm": norm, "mu_over_par": mu / 2.0, "E_over_par": en / 2.0},
        "tuning": tuning,
        "cpu_baseline": None,
    }
    if args.mixed_steps and mode == "imag" and prec == "fp64":
        sm = dbecpy.Solver(lib, p, m, dbecpy.MIXED)
        sm.set_psi(dbecpy.gaussian_psi0(p))
        sm.calcnorm(rescale=True)
        sm.set_coupling(80.0, 60.0)
        sm.step(args.warmup, want_norm=False)
        sm.synchronize()
        t0 = time.perf_counter()
        sm.step(args.mixed_steps, want_norm=False)
        sm.synchronize()
        tm = time.perf_counter() - t0
        sm.close()
        line["mixed"] = {"value": npts * args.mixed_steps / tm, "unit": "grid-pt-steps/s", "steps": args.mixed_steps,
                         "ms_per_step": 1e3 * tm / args.mixed_steps, "speedup_vs_fp64": (npts * args.mixed_steps / tm) / value,
                         "dtype": "f32+f64"}
        log(f"mixed: {1e3 * tm / args.mixed_steps:.3f} ms/step, {line['mixed']['speedup_vs_fp64']:.2f}x FP64")
    if args.tts and mode == "imag" and prec == "fp64":
        line["mixed_tts"] = tts_bench(lib, n, h, dt, cutoff, args.tts_check, 1e-8, args.tts_polish_tol,
                                      args.tts_cap)
    if args.serve_requests and mode == "imag" and prec == "fp64":
        try:
            line["serving"] = serve_bench(lib, args.serve_requests, args.serve_baseline, args.serve_rate,
                                          args.serve_batch, args.serve_check, timeout=600.0,
                                          polish_tol=args.serve_polish_tol)
        except Exception as e:  # report, keep the throughput line
            log(f"serving benchmark failed: {e!r}")
            line["serving"] = {"error": repr(e)}
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        # before the oracle runs: its OpenMP runtime binds the calling thread
        # to one CPU, which would then be all that the affinity query sees
        os.environ.setdefault("OMP_PROC_BIND", "close")
        host = host_description(threads)
        log(f"cpu baseline: oracle, {args.cpu_steps} steps at {n}, {threads} threads")
        v, tc = cpu_baseline(n, h, dt, cutoff, mode, args.cpu_steps, threads)
        line["cpu_baseline"] = {
            "value": v, "unit": "grid-pt-steps/s", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_steps} FP64 split steps of the same {n[0]}x{n[1]}x{n[2]} {mode}-time workload "
                      f"({tc:.1f} s), OpenMP oracle",
            "host": host,
        }
        if args.cpu_cfg0:
            # the 64^3 CPU configuration in full: dx = 0.25, dt = 0.002, 2000
            # steps, G0 = 40, GDD0 = 30, R = 8, Gaussian psi0
            v0, t0 = cpu_baseline((64, 64, 64), 0.25, 0.002, 8.0, "imag", 2000, threads)
            line["cpu_baseline"]["cfg0_full"] = {
                "value": v0, "unit": "grid-pt-steps/s", "seconds": t0,
                "workload": "imag time 64x64x64, dx 0.25, dt 0.002, 2000 steps, G0 40, GDD0 30, R 8"}
            log(f"cpu baseline 64^3 x 2000: {t0:.1f} s, {v0:.3e} grid-pt-steps/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
