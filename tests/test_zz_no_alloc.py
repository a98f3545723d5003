"""The allocation-free density entry point (dbec_density_noalloc) against
dbec_density through the C ABI: every extract kind, several output steps, two
calls in a row on contexts of different grids (the second smaller, so the
shared device scratch still holds the first call's larger contents), FP64 and
mixed, imaginary and real time.  Bit-identical values are required: both
entry points run the same reduction kernels and the same |psi|^2."""
import numpy as np
import pytest

import dbecpy
from conftest import params, smooth_field

pytestmark = pytest.mark.gpu

KINDS = range(10)
STEPS = [(1, 1, 1), (2, 3, 2), (3, 1, 4)]


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def check_all_kinds(s):
    for kind in KINDS:
        for o in STEPS:
            a = s.density(kind, *o)
            b = s.density_noalloc(kind, *o)
            assert a.shape == b.shape, (kind, o)
            assert np.array_equal(a, b), (kind, o, float(np.max(np.abs(a - b))))


@pytest.mark.parametrize("mode,prec", [(dbecpy.IMAG, dbecpy.FP64), (dbecpy.REAL, dbecpy.FP64),
                                       (dbecpy.IMAG, dbecpy.MIXED), (dbecpy.REAL, dbecpy.MIXED)],
                         ids=["imag-fp64", "real-fp64", "imag-mixed", "real-mixed"])
def test_noalloc_density_equals_density_twice(hip_lib, mode, prec):
    for n in ((64, 80, 96), (32, 48, 16)):  # the second call's grid is smaller
        say(f"density_noalloc vs density: mode {mode} prec {prec} grid {n}")
        p = params(n=n, h=(0.25, 0.2, 0.3), dt=0.002, cutoff=5.0)
        with dbecpy.Solver(hip_lib, p, mode, prec) as s:
            s.set_psi(smooth_field(p, seed=5, complex_=(mode == dbecpy.REAL)))
            s.set_coupling(40.0, 30.0)
            s.step(2)  # leaves a pending renormalisation in imaginary time
            check_all_kinds(s)


def test_noalloc_density_rejects_oversize_extract(hip_lib):
    """The full 3-D extract of a 128^3 grid needs 16 MiB of scratch: rejected
    with DBEC_EINVAL and a message, before any launch; a sampled one fits."""
    say("density_noalloc: oversize extract rejected")
    p = params(n=(128, 128, 128), h=(0.2, 0.2, 0.2), dt=0.002, cutoff=6.0)
    with dbecpy.Solver(hip_lib, p) as s:
        s.set_psi(dbecpy.gaussian_psi0(p))
        with pytest.raises(dbecpy.DbecError) as e:
            s.density_noalloc(dbecpy.DEN_XYZ, 1, 1, 1)
        assert e.value.code == dbecpy.DBEC_EINVAL
        assert "scratch" in str(e.value)
        a, b = s.density(dbecpy.DEN_XYZ, 2, 2, 2), s.density_noalloc(dbecpy.DEN_XYZ, 2, 2, 2)
        assert np.array_equal(a, b)
