#!/usr/bin/env python3
"""A whole 256^3 FP64 run against the oracle (not collected by pytest: it
takes about 12 minutes of oracle time on 16 host threads).  Runs the HIP
library and the FP64 oracle side by side from the Gaussian psi0 and compares
them every --chunk steps in norm, mu, E, rms and max|d psi|; writes a
Markdown record.

    python tests/long_parity_256.py --variant workload --steps 5000 --out records/parity_256_workload_5000.md
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "dbec-gp-mpi-cuda_amd", "python"))
import dbecpy  # noqa: E402

STABLE = (332.4918, 105.33272)
WORKLOAD = (40.0, 30.0)


def table(rows, header):
    out = ["| " + " | ".join(header) + " |", "|" + "---:|" * len(header)]
    return "\n".join(out + ["| " + " | ".join(r) + " |" for r in rows])


def oracle_256(args):
    """The 256^3 workload against the FP64 oracle over a whole run."""
    oracle_lib = os.path.join(REPO, "oracle", "liboracle.so")
    g0, gd0 = STABLE if args.variant == "stable" else WORKLOAD
    p = dbecpy.make_params(256, 256, 256, 0.1, 0.1, 0.1, 0.0005, 0.5, 1.0, 1.5, 12.8)
    psi0 = dbecpy.gaussian_psi0(p)
    hip, ora = dbecpy.load(dbecpy.HIP_LIB), dbecpy.load(oracle_lib)
    rows, t_gpu, t_cpu = [], 0.0, 0.0
    with dbecpy.Solver(hip, p) as g, dbecpy.Solver(ora, p) as o:
        for s in (g, o):
            s.set_psi(psi0)
            s.calcnorm(rescale=True)
            s.set_coupling(2 * g0, 2 * gd0)
        done = 0
        while done < args.steps:
            k = min(args.chunk, args.steps - done)
            t0 = time.perf_counter()
            ng = g.step(k)
            t_gpu += time.perf_counter() - t0
            t0 = time.perf_counter()
            no = o.step(k)
            t_cpu += time.perf_counter() - t0
            done += k
            pg, po = g.get_psi(), o.get_psi()
            d = float(np.max(np.abs(pg - po)) / np.max(np.abs(po)))
            del pg, po
            mg, mo = g.calcmuen(), o.calcmuen()
            rg, ro = g.calcrms(), o.calcrms()
            row = {"step": done, "norm_rel": abs(ng - no) / abs(no), "mu_rel": abs(mg[0] - mo[0]) / abs(mo[0]),
                   "E_rel": abs(mg[1] - mo[1]) / abs(mo[1]),
                   "rms_rel": float(np.max(np.abs(np.asarray(rg) - ro) / np.abs(ro))), "max_dpsi_rel": d,
                   "mu_over_par": mo[0] / 2, "E_over_par": mo[1] / 2, "r": float(ro[0])}
            rows.append(row)
            print(f"[campaign] {done} steps: " + ", ".join(f"{k} {v:.3e}" for k, v in row.items() if k != "step"),
                  flush=True)
    worst = {k: max(r[k] for r in rows) for k in ("norm_rel", "mu_rel", "E_rel", "rms_rel", "max_dpsi_rel")}
    out = {"variant": args.variant, "G0": g0, "GDD0": gd0, "steps": args.steps, "gpu_s": t_gpu,
           "oracle_s": t_cpu, "oracle_threads": os.environ.get("OMP_NUM_THREADS", "default"), "worst": worst}
    lines = [f"# 256^3 FP64 imaginary time against the oracle, {args.steps} steps ({args.variant} couplings)", "",
             f"Command: `python tests/long_parity_256.py --variant {args.variant} "
             f"--steps {args.steps} --chunk {args.chunk} --out {args.out}` (HIP library on one MI355X, the "
             "OpenMP FP64 oracle on the same host).", "",
             f"Grid 256^3, dx = 0.1, dt = 5e-4, R = 12.8, trap 0.5/1/1.5, G0 = {g0}, GDD0 = {gd0} "
             "(OPTION 2), Gaussian psi0 normalised.  Relative differences GPU vs oracle after each chunk; "
             "max_dpsi_rel = max|psi_gpu - psi_oracle| / max|psi_oracle|.", ""]
    lines.append(table([[str(r["step"]), f"{r['norm_rel']:.1e}", f"{r['mu_rel']:.1e}", f"{r['E_rel']:.1e}",
                         f"{r['rms_rel']:.1e}", f"{r['max_dpsi_rel']:.1e}", f"{r['mu_over_par']:.10g}",
                         f"{r['E_over_par']:.10g}", f"{r['r']:.8f}"] for r in rows],
                       ["step", "norm", "mu", "E", "rms", "max dpsi", "mu/par", "E/par", "<r>"]))
    lines += ["", "```json", json.dumps(out, indent=1), "```", ""]
    open(args.out, "w").write("\n".join(lines))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--variant", choices=["workload", "stable"], required=True)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--chunk", type=int, default=500)
    ap.add_argument("--out", required=True)
    oracle_256(ap.parse_args())


if __name__ == "__main__":
    main()
