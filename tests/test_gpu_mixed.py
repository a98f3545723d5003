"""The mixed-precision form (FP32 / complex64 bulk, FP64 reductions,
renormalisation and Thomas coefficients) against the plain FP64 path, and
the FP64 polish that follows it."""
import sys

import numpy as np
import pytest

import dbecpy
from conftest import rel

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


# the reference's stable physics: G0 = 332.4918, GDD0 = 105.33272 (OPTION = 2)
G, GD = 2 * 332.4918, 2 * 105.33272


def make(lib, prec, n=64, h=0.3, dt=0.004):
    p = dbecpy.make_params(n, n, n, h, h, h, dt, 0.5, 1.0, 1.5, 10.0)
    s = dbecpy.Solver(lib, p, dbecpy.IMAG, prec)
    s.set_psi(dbecpy.gaussian_psi0(p))
    s.calcnorm(rescale=True)
    s.set_coupling(G, GD)
    return s


def test_mixed_tracks_fp64(hip_lib):
    say("mixed vs fp64 64^3 x 1000")
    with make(hip_lib, dbecpy.FP64) as d, make(hip_lib, dbecpy.MIXED) as m:
        for _ in range(4):
            d.step(250)
            m.step(250)
        mud, mum = d.calcmuen()[0], m.calcmuen()[0]
        say(f"  mu fp64 {mud:.12f} mixed {mum:.12f} rel {abs(mum - mud) / abs(mud):.2e}")
        assert abs(mum - mud) / abs(mud) < 1e-5
        assert rel(m.get_psi(), d.get_psi()) < 1e-4
        assert m.calcnorm() == pytest.approx(1.0, abs=1e-6)


def test_mixed_then_fp64_polish_converges(hip_lib):
    """Converge in mixed precision, then continue in FP64 from the mixed state:
    the polished mu agrees with the FP64 ground state to 1e-8."""
    say("mixed + polish vs fp64 ground state 48^3")
    steps = 3000
    with make(hip_lib, dbecpy.FP64, n=48, h=0.4, dt=0.01) as d, make(hip_lib, dbecpy.MIXED, n=48, h=0.4, dt=0.01) as m:
        d.step(steps)
        m.step(steps)
        p = d.p
        with dbecpy.Solver(hip_lib, p, dbecpy.IMAG, dbecpy.FP64) as pol:
            pol.set_psi(m.get_psi())
            pol.set_coupling(G, GD)
            pol.step(400)
            d.step(400)
            mud, mup = d.calcmuen()[0], pol.calcmuen()[0]
            say(f"  mu fp64 {mud:.14f} polished {mup:.14f} rel {abs(mup - mud) / abs(mud):.2e}")
            assert abs(mup - mud) / abs(mud) < 1e-8


def test_fp64_accumulators_leave_nothing_to_compensate(hip_lib):
    """The mixed form's reductions accumulate in FP64.  Each term of the
    Simpson norm is an FP32 value squared (exact in FP64) times an FP64
    weight, so the device sum differs from a compensated sum of the same
    terms (math.fsum, exact up to the rounding of each product) by the FP64
    accumulation error alone.  Measured on a mixed ground state at 128^3:
    below 1e-13 relative, five orders under the 1e-8 bar and far under the
    FP32 state error, which is why the sums carry no Kahan compensation."""
    import math

    say("mixed: FP64 accumulation against a compensated host sum, 128^3")
    n, h = 128, 0.2
    p = dbecpy.make_params(n, n, n, h, h, h, 0.005, 0.5, 1.0, 1.5, 10.0)
    with dbecpy.Solver(hip_lib, p, dbecpy.IMAG, dbecpy.MIXED) as m:
        m.set_psi(dbecpy.gaussian_psi0(p))
        m.calcnorm(rescale=True)
        m.set_coupling(G, GD)
        m.step(300)
        psi = m.get_psi()  # FP32 values, held exactly in float64
        norm_dev = m.calcnorm()
        rms_dev = m.calcrms()

    def simpson_w(n, h):
        w = np.zeros(n)
        for i in range(1, n - 1, 2):
            w[i - 1] += 1.0
            w[i] += 4.0
            w[i + 1] += 1.0
        if n % 2 == 0:
            w[n - 1] += 5.0 / 4.0
            w[n - 2] += 8.0 / 4.0
            w[n - 3] -= 1.0 / 4.0
        return w * h / 3.0

    w = simpson_w(n, h)
    p2 = psi * psi
    x2 = ((np.arange(n) - n // 2) * h) ** 2
    rows = np.array([[math.fsum(w * p2[i, j]) for j in range(n)] for i in range(n)])  # z
    cols = np.array([math.fsum(w * rows[i]) for i in range(n)])  # y
    norm_host = math.sqrt(math.fsum(w * cols))  # x
    rx_host = math.sqrt(math.fsum(w * x2 * cols))
    d_norm = abs(norm_dev - norm_host) / norm_host
    d_rx = abs(rms_dev[1] - rx_host) / rx_host
    say(f"  norm device {norm_dev!r} compensated {norm_host!r}: {d_norm:.1e}; <x> {d_rx:.1e}")
    assert d_norm < 1e-13 and d_rx < 1e-13
