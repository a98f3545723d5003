#!/usr/bin/env python3
"""Writes serve_stream.json: the serving stream of 2000 requests, drawn once.

Poisson arrivals at 20 requests/s; per request, in this order of draws from
numpy's PCG64 generator with seed 0: the inter-arrival gap (exponential,
mean 1/20 s), the grid index into N in {48, 64, 96, 128, 160, 192}, G0 ~
U[5, 60], GDD0 ~ U[0, 40], gamma, nu, lambda ~ U[0.3, 3.0], and a uniform
draw u with u < 0.15 marking a real-time run.  The settings the stream leaves
open are fixed here and stored with it: dx = 16/N (cubic grid, box
[-8, 8)^3), cutoff R = 8 (half the box), OPTION = 2 (g = 2 G0, gd = 2 GDD0),
imaginary time dt = 0.01 to |dmu/mu| <= 1e-8 between checks (at most 20000
steps), real time dt = 0.001 for 5000 steps from the Gaussian psi0.

    python tests/golden/make_serve_stream.py            # rewrites the fixture
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SIZES = (48, 64, 96, 128, 160, 192)
SETTINGS = {"count": 2000, "rate_per_s": 20.0, "seed": 0, "generator": "numpy PCG64",
            "box": 16.0, "cutoff": 8.0, "par": 2.0, "dt_imag": 0.01, "dt_real": 0.001,
            "real_steps": 5000, "imag_cap": 20000, "mutol": 1e-8,
            "fields": ["t_arrival", "n", "G0", "GDD0", "gamma", "nu", "lambda", "real"]}


def draw(count=2000, rate=20.0, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    t, rows = 0.0, []
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        n = SIZES[int(rng.integers(0, len(SIZES)))]
        g0 = rng.uniform(5.0, 60.0)
        gd0 = rng.uniform(0.0, 40.0)
        trap = (rng.uniform(0.3, 3.0), rng.uniform(0.3, 3.0), rng.uniform(0.3, 3.0))
        real = int(rng.uniform() < 0.15)
        rows.append([t, n, g0, gd0, trap[0], trap[1], trap[2], real])
    return rows


def main():
    out = dict(SETTINGS, requests=draw(SETTINGS["count"], SETTINGS["rate_per_s"], SETTINGS["seed"]))
    with open(os.path.join(HERE, "serve_stream.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
