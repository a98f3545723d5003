"""The launch choices that the autotuner picks (FFT tile multipliers, x/y
sweep row width and boundary peeling, z sweep staged or direct, the LDS
layout of the FP64 fused inverse pass) change no result: the same solves run
in child processes with the built-in rules, with every choice forced the
other way, and with the tuner on, and psi must agree bit for bit."""
import json
import os
import subprocess
import sys

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu

# Each case: grid, mode, precision.  Non-cubic grids so that x and y differ.
SCRIPT = r"""
import hashlib, json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import dbecpy
lib = dbecpy.load(dbecpy.HIP_LIB)
out = {}
for (nx, ny, nz), mode, prec in (((96, 64, 128), dbecpy.IMAG, dbecpy.FP64), ((96, 64, 128), dbecpy.IMAG, dbecpy.MIXED),
                                 ((64, 96, 48), dbecpy.REAL, dbecpy.FP64), ((64, 96, 48), dbecpy.REAL, dbecpy.MIXED)):
    p = dbecpy.make_params(nx, ny, nz, 0.25, 0.2, 0.3, 0.002, 0.5, 1.0, 1.5, 6.0)
    with dbecpy.Solver(lib, p, mode, prec) as s:
        psi0 = dbecpy.gaussian_psi0(p)
        s.set_psi(psi0 if mode == dbecpy.IMAG else psi0.astype(np.complex128))
        if mode == dbecpy.IMAG:
            s.calcnorm(rescale=True)
        s.set_coupling(2 * 20.0, 2 * 10.0)
        s.step(12)
        out[f"{nx}x{ny}x{nz}-{mode}-{prec}"] = hashlib.sha256(s.get_psi().tobytes()).hexdigest()
print(json.dumps(out))
"""

FORCED = {
    "one": {**{f"DBEC_FFT_LM_{p}{b}": "1" for p in ("ZFWD", "YPASS", "XCONV", "FUSED") for b in (64, 32)},
            "DBEC_SWEEP_XY_ROWB": "128", "DBEC_SWEEP_XY_ROWB32": "128", "DBEC_SWEEP_XY_PEEL": "0",
            "DBEC_SWEEP_Z_DIRECT": "1", "DBEC_FFT_ZINV_SWZ": "1"},
    "two": {**{f"DBEC_FFT_LM_{p}{b}": "2" for p in ("ZFWD", "YPASS", "XCONV", "FUSED") for b in (64, 32)},
            "DBEC_SWEEP_XY_ROWB": "256", "DBEC_SWEEP_XY_ROWB32": "256", "DBEC_SWEEP_XY_PEEL": "1",
            "DBEC_SWEEP_Z_DIRECT": "0", "DBEC_FFT_ZINV_SWZ": "0"},
}


def run(extra):
    env = {k: v for k, v in os.environ.items()
           if not k.startswith(("DBEC_FFT_LM_", "DBEC_FFT_ZINV_SWZ", "DBEC_SWEEP_", "DBEC_AUTOTUNE"))}
    env.update(extra)
    print(f"[gpu-case] autotune child {sorted(extra.items())[:2]}...", flush=True)
    r = subprocess.run([sys.executable, "-c", SCRIPT, os.path.join(PKG, "python")], env=env, capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_tuned_choices_leave_results_bit_identical():
    rules = run({"DBEC_AUTOTUNE": "0"})
    assert len(rules) == 4
    assert run({"DBEC_AUTOTUNE": "0", **FORCED["one"]}) == rules
    assert run({"DBEC_AUTOTUNE": "0", **FORCED["two"]}) == rules
    assert run({}) == rules


# The FP64 fused inverse z pass has two LDS layouts (padded and XOR-swizzled;
# with complex psi the swizzled one also reads the twiddles from the global
# table); each z length has its own radix plan, and the plans' first radices
# (4, 8, 16) set the swizzle.
LAYOUT_SCRIPT = r"""
import hashlib, json, sys
sys.path.insert(0, sys.argv[1])
import dbecpy
lib = dbecpy.load(dbecpy.HIP_LIB)
out = {}
for mode, name in ((dbecpy.IMAG, "imag"), (dbecpy.REAL, "real")):
    for nz in (48, 80, 96, 192, 256, 384, 512):
        p = dbecpy.make_params(32, 48, nz, 0.25, 0.25, 0.1, 0.002, 0.5, 1.0, 1.5, 6.0)
        with dbecpy.Solver(lib, p, mode, dbecpy.FP64) as s:
            s.set_psi(dbecpy.gaussian_psi0(p))
            s.calcnorm(rescale=True)
            s.set_coupling(2 * 20.0, 2 * 10.0)
            s.step(4)
            out[name + str(nz)] = hashlib.sha256(s.get_psi().tobytes()).hexdigest()
print(json.dumps(out))
"""


def test_fused_inverse_layouts_bit_identical():
    def layout(v):
        env = {k: x for k, x in os.environ.items()
               if not k.startswith(("DBEC_FFT_LM_", "DBEC_FFT_ZINV_SWZ", "DBEC_SWEEP_", "DBEC_AUTOTUNE"))}
        env.update({"DBEC_AUTOTUNE": "0", "DBEC_FFT_ZINV_SWZ": v})
        print(f"[gpu-case] fused inverse layout {v}", flush=True)
        r = subprocess.run([sys.executable, "-c", LAYOUT_SCRIPT, os.path.join(PKG, "python")], env=env,
                           capture_output=True, text=True, timeout=180)
        assert r.returncode == 0, r.stderr[-2000:]
        return json.loads(r.stdout.strip().splitlines()[-1])
    padded = layout("0")
    assert len(padded) == 14
    assert layout("1") == padded
