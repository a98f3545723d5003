#!/usr/bin/env python3
"""Named configuration runs whose results are kept as records (Markdown with
the command, inputs and numbers).

    campaign.py real-quench --variant workload|stable --work DIR --out FILE
        The real-time configuration through the product programs, as a user
        of imag3d-mpicuda / real3d-mpicuda would run it: imag3d at
        256x256x384 (dx = 0.1, dt = 5e-4, R = 12.8, OPTION = 2) writes its
        ground state's density (NRUNOUT = imag3d-den), then real3d
        (NSTP = 0: psi = sqrt(density), NPAS = 0, NRUN = 20000, dt = 2e-4,
        GPAR = 1, GDDPAR = 1.5: gd quenched 30 -> 45 for the workload
        couplings) tracks rms(t) every 100 steps (DYNAOUT) and norm, mu, the
        reference's E (of |psi|) and the phase-gradient energy every 1000
        steps (OBSSTP / OBSOUT).  Variant `workload`: G0 = 40, GDD0 = 30,
        imaginary time for 5000 steps (the 256^3 workload's count); variant
        `stable`: the reference example's G0 = 332.4918, GDD0 = 105.33272,
        imaginary time to |dmu/mu| <= 1e-8 (MUTOL, checks every 500 steps).

    campaign.py mixed-512 --out FILE
        512^3 (dx = 0.05, dt = 1.25e-4, R = 12.8) from the Gaussian: plain
        FP64 for K + 200 steps against the mixed form for K steps followed by
        a 200-step FP64 polish; step times and the differences in mu, E, rms
        and max|d psi|.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import dbecpy  # noqa: E402

STABLE = (332.4918, 105.33272)
WORKLOAD = (40.0, 30.0)


def write_input(path, d):
    with open(path, "w") as f:
        for k, v in d.items():
            f.write(f"   {k} = {v}\n")


def run_prog(prog, work, inp):
    t0 = time.time()
    r = subprocess.run([os.path.join(dbecpy.BIN_DIR, prog), "-i", inp], cwd=work, capture_output=True, text=True)
    if r.returncode != 0:
        raise SystemExit(f"{prog} failed ({r.returncode}): {r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    print(f"[campaign] {prog} done in {time.time() - t0:.1f} s", flush=True)
    return time.time() - t0


def table(rows, header):
    out = ["| " + " | ".join(header) + " |", "|" + "---:|" * len(header)]
    for r in rows:
        out.append("| " + " | ".join(r) + " |")
    return "\n".join(out)


def real_quench(args):
    g0, gd0 = STABLE if args.variant == "stable" else WORKLOAD
    os.makedirs(args.work, exist_ok=True)
    grid = dict(OPTION="2", NATOMS="5000", NX="256", NY="256", NZ="384", DX="0.1", DY="0.1", DZ="0.1",
                GAMMA="0.5", NU="1.0", LAMBDA="1.5", CUTOFF="12.8", G0=repr(g0), GDD0=repr(gd0))
    imag = dict(grid, DT="0.0005", NSTP="0", NRUN="0", OUTPUT="imag3d-out", RMSOUT="imag3d-rms",
                NRUNOUT="imag3d-den", OUTSTPX="1", OUTSTPY="1", OUTSTPZ="1", OBSSTP="500", OBSOUT="imag3d-obs")
    if args.variant == "stable":
        imag.update(NPAS="40000", MUTOL="1e-8")
    else:
        imag.update(NPAS="5000")
    real = dict(grid, DT="0.0002", NSTP="0", NPAS="0", NRUN=str(args.real_steps), GPAR="1.", GDDPAR="1.5",
                OUTPUT="real3d-out", RMSOUT="real3d-rms", DYNAOUT="real3d-dyna", OUTSTPT="100", OBSSTP="1000",
                OBSOUT="real3d-obs")
    write_input(os.path.join(args.work, "imag3d-input"), imag)
    write_input(os.path.join(args.work, "real3d-input"), real)
    t_imag = run_prog("imag3d", args.work, "imag3d-input")
    t_real = run_prog("real3d", args.work, "real3d-input")
    rd = lambda f: open(os.path.join(args.work, f)).read()
    dyna = np.loadtxt(os.path.join(args.work, "real3d-dyna.txt"))
    obs = np.array([list(map(float, l.split())) for l in rd("real3d-obs.txt").splitlines() if not l.startswith("#")])
    iobs = [l for l in rd("imag3d-obs.txt").splitlines() if not l.startswith("#")]
    norm, e_ref, e_ph = obs[:, 2], obs[:, 4], obs[:, 9]
    summary = {
        "variant": args.variant, "G0": g0, "GDD0": gd0, "real_steps": args.real_steps,
        "imag_wall_s": t_imag, "real_wall_s": t_real,
        "norm_drift_max": float(np.max(np.abs(norm - norm[0]))),
        "E_ref_range": [float(e_ref.min()), float(e_ref.max())],
        "E_phase_range": [float(e_ph.min()), float(e_ph.max())],
        "E_phase_drift_rel": float(np.ptp(e_ph) / abs(e_ph[0])),
        "rms_first": dyna[0, 1:].tolist(), "rms_last": dyna[-1, 1:].tolist(),
    }
    lines = [f"# Real-time quench, 256x256x384, {args.variant} couplings", "",
             f"Command: `python dbec-gp-mpi-cuda_amd/python/campaign.py real-quench --variant {args.variant} "
             f"--work <dir> --out {args.out}` (product programs `imag3d`, `real3d` on one MI355X).", "",
             "## Inputs", "", "imag3d:", "", "```", rd("imag3d-input").rstrip(), "```", "", "real3d:", "", "```",
             rd("real3d-input").rstrip(), "```", "",
             "## imag3d summary (imag3d-out.txt)", "", "```", rd("imag3d-out.txt").rstrip(), "```", "",
             "Cadence (imag3d-obs.txt: step t norm mu/par E/par <r> <x> <y> <z>), last rows:", "", "```"]
    lines += iobs[-4:] + ["```", "", "## real3d summary (real3d-out.txt)", "", "```", rd("real3d-out.txt").rstrip(),
                          "```", "",
                          "## Cadence every 1000 steps (real3d-obs.txt)", "",
                          "E/par is the reference's energy (kinetic term of |psi|); Ephase/par has the phase "
                          "gradient in the kinetic term (the energy the dynamics conserves).", ""]
    lines.append(table([[f"{int(r[0])}", f"{r[1]:.4f}", f"{r[2]:.12f}", f"{r[3]:.8f}", f"{r[4]:.8f}",
                         f"{r[9]:.8f}", f"{r[6]:.6f}", f"{r[7]:.6f}", f"{r[8]:.6f}"] for r in obs],
                       ["step", "t", "norm", "mu/par", "E/par", "Ephase/par", "<x>", "<y>", "<z>"]))
    lines += ["", "## rms(t) every 1000 steps (real3d-dyna.txt, every 100 steps in the file)", ""]
    lines.append(table([[f"{r[0]:.4f}", f"{r[1]:.6f}", f"{r[2]:.6f}", f"{r[3]:.6f}", f"{r[4]:.6f}"]
                        for r in dyna[::10]], ["t", "<r>", "<x>", "<y>", "<z>"]))
    lines += ["", "## Summary", "", "```json", json.dumps(summary, indent=1), "```", ""]
    open(args.out, "w").write("\n".join(lines))
    print(json.dumps(summary), flush=True)


def mixed_512(args):
    n = 512
    p = dbecpy.make_params(n, n, n, 0.05, 0.05, 0.05, 0.000125, 0.5, 1.0, 1.5, 12.8, device=0)
    g, gd = 2 * WORKLOAD[0], 2 * WORKLOAD[1]
    lib = dbecpy.load(dbecpy.HIP_LIB)
    psi0 = dbecpy.gaussian_psi0(p)
    K, P = args.steps, args.polish
    out = {"grid": [n, n, n], "dx": 0.05, "dt": 0.000125, "cutoff": 12.8, "G0": WORKLOAD[0], "GDD0": WORKLOAD[1],
           "mixed_steps": K, "polish_steps": P}
    with dbecpy.Solver(lib, p, dbecpy.IMAG, dbecpy.FP64) as d:
        d.set_psi(psi0)
        d.calcnorm(rescale=True)
        d.set_coupling(g, gd)
        d.step(5)
        d.synchronize()
        t0 = time.perf_counter()
        d.step(K + P - 5, want_norm=False)
        d.synchronize()
        t64 = time.perf_counter() - t0
        mu64, en64 = d.calcmuen()
        rms64, psi64 = d.calcrms(), d.get_psi()
    print(f"[campaign] FP64 {K + P} steps: {1e3 * t64 / (K + P - 5):.3f} ms/step", flush=True)
    with dbecpy.Solver(lib, p, dbecpy.IMAG, dbecpy.MIXED) as m:
        m.set_psi(psi0)
        m.calcnorm(rescale=True)
        m.set_coupling(g, gd)
        m.step(5)
        m.synchronize()
        t0 = time.perf_counter()
        m.step(K - 5, want_norm=False)
        m.synchronize()
        tm = time.perf_counter() - t0
        mum, enm = m.calcmuen()
        psim = m.get_psi()
    with dbecpy.Solver(lib, p, dbecpy.IMAG, dbecpy.FP64) as d:
        d.set_psi(psim)
        d.calcnorm(rescale=True)
        d.set_coupling(g, gd)
        d.step(P)
        mup, enp = d.calcmuen()
        rmsp, psip = d.calcrms(), d.get_psi()
    err = lambda a, b: abs(a - b) / abs(b)
    out.update({
        "fp64_ms_per_step": 1e3 * t64 / (K + P - 5), "mixed_ms_per_step": 1e3 * tm / (K - 5),
        "step_speedup": (t64 / (K + P - 5)) / (tm / (K - 5)),
        "mu_over_par": {"fp64": mu64 / 2, "mixed": mum / 2, "mixed+polish": mup / 2},
        "E_over_par": {"fp64": en64 / 2, "mixed": enm / 2, "mixed+polish": enp / 2},
        "error_vs_fp64": {"mu": err(mup, mu64), "E": err(enp, en64),
                          "rms": float(np.max(np.abs(np.asarray(rmsp) - rms64) / np.abs(rms64))),
                          "max_dpsi_over_max_psi": float(np.max(np.abs(psip - psi64)) / np.max(np.abs(psi64))),
                          "mu_before_polish": err(mum, mu64)},
    })
    lines = ["# 512^3 mixed precision + FP64 polish against plain FP64", "",
             f"Command: `python dbec-gp-mpi-cuda_amd/python/campaign.py mixed-512 --steps {K} --polish {P} "
             f"--out {args.out}` (C ABI through dbecpy, one MI355X).", "",
             "Grid 512^3, dx = 0.05 (box [-12.8, 12.8)^3), dt = 1.25e-4, R = 12.8, trap 0.5/1/1.5, "
             f"G0 = {WORKLOAD[0]}, GDD0 = {WORKLOAD[1]} (OPTION 2), Gaussian psi0.  FP64: {K + P} steps.  "
             f"Mixed: {K} steps, then psi (FP32 widened to FP64 on the host) and {P} FP64 steps.", "",
             "```json", json.dumps(out, indent=1), "```", ""]
    open(args.out, "w").write("\n".join(lines))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    a = sub.add_parser("real-quench")
    a.add_argument("--variant", choices=["workload", "stable"], required=True)
    a.add_argument("--work", required=True)
    a.add_argument("--out", required=True)
    a.add_argument("--real-steps", type=int, default=20000)
    b = sub.add_parser("mixed-512")
    b.add_argument("--steps", type=int, default=1000)
    b.add_argument("--polish", type=int, default=200)
    b.add_argument("--out", required=True)
    args = ap.parse_args()
    {"real-quench": real_quench, "mixed-512": mixed_512}[args.what](args)


if __name__ == "__main__":
    main()
