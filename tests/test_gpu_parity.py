"""Parity of the HIP library with the FP64 oracle, through the C ABI.

Per operator on small grids (every transform length family), multi-step
trajectories (imaginary and real time), and the reference's example chain via
the product programs.  FP64 tolerance: 1e-10 relative in mu, E, rms and
max|d psi| (the bar of the plain path); single operators are checked at 1e-12.
"""
import os
import sys

import numpy as np
import pytest

import dbecpy
from conftest import params, rel, smooth_field
from golden_cli import check_chain, run_chain

pytestmark = pytest.mark.gpu

TOL_OP = 1e-12
TOL = 1e-10


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


GRIDS = [(32, 48, 64), (64, 64, 64), (48, 80, 96), (128, 96, 80), (160, 192, 16)]


@pytest.mark.parametrize("mode", [dbecpy.IMAG, dbecpy.REAL], ids=["imag", "real"])
@pytest.mark.parametrize("n", GRIDS, ids=lambda n: "x".join(map(str, n)))
def test_operator_parity(hip_lib, oracle_lib, mode, n):
    say(f"operator parity mode={mode} grid={n}")
    p = params(n=n, h=(0.25, 0.2, 0.3), dt=0.004, cutoff=5.0)
    f = smooth_field(p, seed=7, complex_=(mode == dbecpy.REAL))
    with dbecpy.Solver(hip_lib, p, mode) as g, dbecpy.Solver(oracle_lib, p, mode) as o:
        for s in (g, o):
            s.set_psi(f)
            s.set_coupling(80.0, 60.0)
        for s in (g, o):
            s.calcpsidd2()
        assert rel(g.get_psidd2(), o.get_psidd2()) < TOL_OP, "calcpsidd2"
        for name in ("calcnu", "calclux", "calcluy", "calcluz"):
            for s in (g, o):
                getattr(s, name)()
            assert rel(g.get_psi(), o.get_psi()) < TOL_OP, name
        ng, no = g.calcnorm(rescale=(mode == dbecpy.IMAG)), o.calcnorm(rescale=(mode == dbecpy.IMAG))
        assert ng == pytest.approx(no, rel=TOL_OP)
        assert rel(g.get_psi(), o.get_psi()) < TOL_OP, "calcnorm rescale"
        mg, mo = g.calcmuen(), o.calcmuen()
        assert mg == pytest.approx(mo, rel=TOL_OP)
        assert g.calcmuen_phase() == pytest.approx(o.calcmuen_phase(), rel=TOL_OP)
        assert g.calcrms() == pytest.approx(o.calcrms(), rel=TOL_OP)
        assert g.psi_center() == pytest.approx(o.psi_center(), rel=TOL_OP, abs=1e-300)
        for kind in range(10):
            for o3 in ((1, 1, 1), (2, 3, 2)):
                a, b = g.density(kind, *o3), o.density(kind, *o3)
                assert a.shape == b.shape
                assert rel(a, b) < TOL_OP, (kind, o3)


@pytest.mark.parametrize("mode", [dbecpy.IMAG, dbecpy.REAL], ids=["imag", "real"])
def test_tables_match(hip_lib, oracle_lib, mode):
    say(f"tables mode={mode}")
    p = params(n=(32, 48, 64))
    with dbecpy.Solver(hip_lib, p, mode) as g, dbecpy.Solver(oracle_lib, p, mode) as o:
        assert np.array_equal(g.table(dbecpy.TABLE_POT), o.table(dbecpy.TABLE_POT))
        assert np.array_equal(g.table(dbecpy.TABLE_POTDD), o.table(dbecpy.TABLE_POTDD))
        for ax in range(3):
            for t in (dbecpy.TABLE_CALPHA, dbecpy.TABLE_CGAMMA):
                assert rel(g.table(t, ax), o.table(t, ax)) < 1e-15


@pytest.mark.parametrize("mode", [dbecpy.IMAG, dbecpy.REAL], ids=["imag", "real"])
@pytest.mark.parametrize("n", [(32, 48, 64), (48, 32, 80)], ids=lambda n: "x".join(map(str, n)))
def test_step_with_nonzero_faces(hip_lib, oracle_lib, mode, n):
    """A field that is O(1) on every face, so the periodic face copies of the
    dipolar potential (kernel2a/2b) change calcnu on the face rows, which then
    enter the sweeps' boundary terms.  One and two composite steps."""
    say(f"nonzero faces mode={mode} grid={n}")
    p = params(n=n, h=(0.3, 0.25, 0.2), dt=0.004, cutoff=4.0)
    x, y, z = dbecpy.grid(p)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    f = 1.0 + 0.4 * np.cos(0.7 * X + 0.3 * Y) * np.sin(0.5 * Z + 0.2) + 0.1 * X / (1 + Y * Y)
    if mode == dbecpy.REAL:
        f = f * np.exp(1j * (0.3 * X - 0.2 * Z))
    with dbecpy.Solver(hip_lib, p, mode) as g, dbecpy.Solver(oracle_lib, p, mode) as o:
        for s in (g, o):
            s.set_psi(f)
            s.set_coupling(60.0, 200.0)
        for _ in range(2):
            a, b = g.step(1), o.step(1)
            if mode == dbecpy.IMAG:
                assert a == pytest.approx(b, rel=TOL_OP)
            assert rel(g.get_psi(), o.get_psi()) < TOL_OP


def test_imag_trajectory_cfg0(hip_lib, oracle_lib):
    """64^3, dx=0.25, dt=0.002, 2000 steps, gamma/nu/lambda = 0.5/1/1.5, G0=40,
    GDD0=30, R=8, Gaussian psi0: mu, E, rms every 500 steps and max|d psi|."""
    say("imag trajectory 64^3 x 2000")
    p = dbecpy.make_params(64, 64, 64, 0.25, 0.25, 0.25, 0.002, 0.5, 1.0, 1.5, 8.0)
    psi0 = dbecpy.gaussian_psi0(p)
    with dbecpy.Solver(hip_lib, p) as g, dbecpy.Solver(oracle_lib, p) as o:
        for s in (g, o):
            s.set_psi(psi0)
            s.calcnorm(rescale=True)
            s.set_coupling(2 * 40.0, 2 * 30.0)
        for chunk in range(4):
            ng, no = g.step(500), o.step(500)
            assert ng == pytest.approx(no, rel=TOL)
            assert g.calcmuen() == pytest.approx(o.calcmuen(), rel=TOL)
            assert g.calcrms() == pytest.approx(o.calcrms(), rel=TOL)
            d = rel(g.get_psi(), o.get_psi())
            say(f"  after {500 * (chunk + 1)} steps: max|dpsi|/max|psi| = {d:.3e}")
            assert d < TOL


def test_imag_ramp_matches(hip_lib, oracle_lib):
    """NSTP ramp: g += g_stp before every step, accumulated."""
    say("imag ramp 48x64x32")
    p = params(n=(48, 64, 32), h=(0.3, 0.25, 0.3), dt=0.003, cutoff=6.0)
    with dbecpy.Solver(hip_lib, p) as g, dbecpy.Solver(oracle_lib, p) as o:
        for s in (g, o):
            s.set_psi(dbecpy.gaussian_psi0(p))
            s.calcnorm(rescale=True)
            s.set_coupling(0.0, 0.0)
            s.step(37, 2 * 40.0 / 37, 2 * 30.0 / 37)
        assert g.get_coupling() == o.get_coupling()
        assert rel(g.get_psi(), o.get_psi()) < TOL


def test_real_trajectory(hip_lib, oracle_lib):
    say("real trajectory 48x64x80 x 300")
    p = params(n=(48, 64, 80), h=(0.3, 0.25, 0.2), dt=0.001, cutoff=6.0)
    f = dbecpy.gaussian_psi0(p).astype(np.complex128) * np.exp(1j * 0.2 * dbecpy.grid(p)[0][:, None, None])
    with dbecpy.Solver(hip_lib, p, dbecpy.REAL) as g, dbecpy.Solver(oracle_lib, p, dbecpy.REAL) as o:
        for s in (g, o):
            s.set_psi(f)
            s.set_coupling(2 * 40.0, 2 * 45.0)
        for _ in range(3):
            g.step(100, want_norm=False)
            o.step(100, want_norm=False)
            assert rel(g.get_psi(), o.get_psi()) < TOL
            assert g.calcrms() == pytest.approx(o.calcrms(), rel=TOL)
            assert g.calcnorm() == pytest.approx(o.calcnorm(), rel=TOL)


# Couplings of the reference's example input (G0 = 4 pi AS N a0 / AHO,
# GDD0 = 3 ADD N a0 / AHO with AS = 100, ADD = 132.7, N = 5000, AHO = 1e-6):
# a stable ground state (eps_dd = 1.33), unlike the 256^3 workload's G0 = 40,
# GDD0 = 30 (eps_dd = 3.1), which drive psi towards collapse.
STABLE = (2 * 332.4918, 2 * 105.33272)


def check_imag_256(hip_lib, oracle_lib, coupling, steps, chunk):
    p = dbecpy.make_params(256, 256, 256, 0.1, 0.1, 0.1, 0.0005, 0.5, 1.0, 1.5, 12.8)
    psi0 = dbecpy.gaussian_psi0(p)
    with dbecpy.Solver(hip_lib, p) as g, dbecpy.Solver(oracle_lib, p) as o:
        for s in (g, o):
            s.set_psi(psi0)
            s.calcnorm(rescale=True)
            s.set_coupling(*coupling)
        for k in range(steps // chunk):
            assert g.step(chunk) == pytest.approx(o.step(chunk), rel=TOL)
            d = rel(g.get_psi(), o.get_psi())
            mg, mo = g.calcmuen(), o.calcmuen()
            say(f"  after {chunk * (k + 1)} steps: max|dpsi|/max|psi| {d:.2e}, "
                f"mu/par {mg[0] / 2:.10f} (|dmu/mu| {abs(mg[0] - mo[0]) / abs(mo[0]):.1e})")
            assert d < TOL
            assert mg == pytest.approx(mo, rel=TOL)
            assert g.calcrms() == pytest.approx(o.calcrms(), rel=TOL)


def test_imag_256_workload_500_steps(hip_lib, oracle_lib):
    """The 256^3 workload (dx = 0.1, dt = 5e-4, R = 12.8, G0 = 40, GDD0 = 30,
    Gaussian psi0), 500 FP64 steps: mu, E, rms and max|d psi| against the
    oracle every 100 steps at 1e-10."""
    say("imag 256^3 workload couplings x 500")
    check_imag_256(hip_lib, oracle_lib, (80.0, 60.0), 500, 100)


def test_imag_256_stable_500_steps(hip_lib, oracle_lib):
    """The same grid on stable couplings (the reference example's G0, GDD0),
    500 FP64 steps against the oracle at 1e-10."""
    say("imag 256^3 stable couplings x 500")
    check_imag_256(hip_lib, oracle_lib, STABLE, 500, 100)


def test_real_256x256x384_from_imag_state(hip_lib, oracle_lib):
    """The real-time workload's grid (256x256x384, dx = 0.1, dt = 2e-4, R =
    12.8): an imaginary-time FP64 state (200 steps of dt = 5e-4 on stable
    couplings, on the GPU) propagated for 100 real-time steps with gd
    quenched by 1.5, GPU against oracle: psi, norm, rms, mu and E (both
    kinetic forms) at 1e-10."""
    say("real 256x256x384 x 100 from an imaginary-time state")
    pi = dbecpy.make_params(256, 256, 384, 0.1, 0.1, 0.1, 0.0005, 0.5, 1.0, 1.5, 12.8)
    with dbecpy.Solver(hip_lib, pi) as s:
        s.set_psi(dbecpy.gaussian_psi0(pi))
        s.calcnorm(rescale=True)
        s.set_coupling(*STABLE)
        s.step(200)
        psi = s.get_psi().astype(np.complex128)
    p = dbecpy.make_params(256, 256, 384, 0.1, 0.1, 0.1, 0.0002, 0.5, 1.0, 1.5, 12.8)
    with dbecpy.Solver(hip_lib, p, dbecpy.REAL) as g, dbecpy.Solver(oracle_lib, p, dbecpy.REAL) as o:
        for s in (g, o):
            s.set_psi(psi)
            s.set_coupling(STABLE[0], 1.5 * STABLE[1])
        del psi
        for k in range(2):
            g.step(50, want_norm=False)
            o.step(50, want_norm=False)
            d = rel(g.get_psi(), o.get_psi())
            say(f"  after {50 * (k + 1)} steps: max|dpsi|/max|psi| {d:.2e}")
            assert d < TOL
            assert g.calcnorm() == pytest.approx(o.calcnorm(), rel=TOL)
            assert g.calcrms() == pytest.approx(o.calcrms(), rel=TOL)
        assert g.calcmuen() == pytest.approx(o.calcmuen(), rel=TOL)
        assert g.calcmuen_phase() == pytest.approx(o.calcmuen_phase(), rel=TOL)


def test_imag_512_fp64_and_mixed(hip_lib, oracle_lib):
    """512^3 (dx = 0.05, dt = 1.25e-4, R = 12.8): two FP64 steps against the
    oracle, and 100 steps of the mixed form against FP64 (mu to 1e-6)."""
    say("imag 512^3: FP64 x 2 vs oracle, mixed x 100 vs FP64")
    p = dbecpy.make_params(512, 512, 512, 0.05, 0.05, 0.05, 0.000125, 0.5, 1.0, 1.5, 12.8)
    psi0 = dbecpy.gaussian_psi0(p)
    with dbecpy.Solver(hip_lib, p) as g, dbecpy.Solver(oracle_lib, p) as o:
        for s in (g, o):
            s.set_psi(psi0)
            s.calcnorm(rescale=True)
            s.set_coupling(80.0, 60.0)
        assert g.step(2) == pytest.approx(o.step(2), rel=TOL)
        assert rel(g.get_psi(), o.get_psi()) < TOL
    del psi0
    with dbecpy.Solver(hip_lib, p) as d, dbecpy.Solver(hip_lib, p, dbecpy.IMAG, dbecpy.MIXED) as m:
        psi0 = dbecpy.gaussian_psi0(p)
        for s in (d, m):
            s.set_psi(psi0)
            s.calcnorm(rescale=True)
            s.set_coupling(80.0, 60.0)
            s.step(100)
        mud, mum = d.calcmuen()[0], m.calcmuen()[0]
        say(f"  mu fp64 {mud:.12g} mixed {mum:.12g} rel {abs(mum - mud) / abs(mud):.2e}")
        assert abs(mum - mud) / abs(mud) < 1e-6


# Every supported transform length on every axis: grid i takes LENGTHS[i],
# LENGTHS[i + 5] and LENGTHS[i + 9] (mod 13) for x, y, z, so each length's
# radix plan runs in the z passes (zfwd, zinv, fused inverse), the y passes
# and the x convolution, and its sweeps along each axis.
LENGTHS = (16, 32, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512)
ROTATED = [(LENGTHS[i], LENGTHS[(i + 5) % 13], LENGTHS[(i + 9) % 13]) for i in range(13)]


@pytest.mark.parametrize("n", ROTATED, ids=lambda n: "x".join(map(str, n)))
def test_every_length_on_every_axis(hip_lib, oracle_lib, n):
    """calcpsidd2 and one composite step against the oracle (FP64, imaginary
    and real time, 1e-12), and one mixed-precision step against FP64."""
    say(f"every length on every axis: grid {n}")
    p = params(n=n, h=tuple(16.0 / k for k in n), dt=0.002, cutoff=6.0)
    for mode in (dbecpy.IMAG, dbecpy.REAL):
        f = smooth_field(p, seed=3, complex_=(mode == dbecpy.REAL))
        with dbecpy.Solver(hip_lib, p, mode) as g, dbecpy.Solver(oracle_lib, p, mode) as o:
            for s in (g, o):
                s.set_psi(f)
                s.set_coupling(40.0, 30.0)
                s.calcpsidd2()
            assert rel(g.get_psidd2(), o.get_psidd2()) < TOL_OP, ("calcpsidd2", mode)
            a, b = g.step(1), o.step(1)
            if mode == dbecpy.IMAG:
                assert a == pytest.approx(b, rel=TOL_OP)
            assert rel(g.get_psi(), o.get_psi()) < TOL_OP, ("step", mode)
            with dbecpy.Solver(hip_lib, p, mode, dbecpy.MIXED) as m:
                m.set_psi(f)
                m.set_coupling(40.0, 30.0)
                m.step(1)
                d = rel(m.get_psi(), g.get_psi())
                say(f"  mode {mode}: mixed step vs FP64 {d:.2e}")
                assert d < 1e-5, ("mixed step", mode)


def test_reference_chain_product_programs(tmp_path):
    """The product's imag3d -> real3d chain on the reference's example inputs
    against the reference's shipped outputs."""
    say("reference chain via product programs")
    work = run_chain(dbecpy.BIN_DIR, str(tmp_path), timeout=300)
    rep = check_chain(work, den_tol=1e-11)
    say(f"  last-digit flips: { {k: v for k, v in rep.items() if k.endswith('.txt')} }")
