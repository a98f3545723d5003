"""Pins of the oracle's functions by mathematics that does not restate them:
library routines (numpy FFT, dense linear solves), closed-form integrals,
symmetries and invariants.  Grids are small so the file runs in seconds."""
import numpy as np
import pytest

import dbecpy
from conftest import params, rel, smooth_field


def face_copies(phi):
    """Periodic face copies of calcpsidd2_kernel2a/2b (imag3d.cu:1591-1625) at
    one rank, applied to an array, written as the three copy loops."""
    nx, ny, nz = phi.shape
    out = phi.copy()
    out[nx - 1, : ny - 1, : nz - 1] = phi[0, : ny - 1, : nz - 1]
    out[: nx - 1, ny - 1, : nz - 1] = phi[: nx - 1, 0, : nz - 1]
    out[: nx - 1, : ny - 1, nz - 1] = phi[: nx - 1, : ny - 1, 0]
    return out


@pytest.mark.parametrize("mode", [dbecpy.IMAG, dbecpy.REAL])
def test_convolution_equals_numpy_rfft(oracle_lib, mode):
    p = params(n=(16, 24, 20), cutoff=3.0)
    f = smooth_field(p, seed=1, complex_=(mode == dbecpy.REAL))
    with dbecpy.Solver(oracle_lib, p, mode) as s:
        s.set_psi(f)
        s.calcpsidd2()
        phi = s.get_psidd2()
        half = s.table(dbecpy.TABLE_POTDD).reshape(16, 24, 11)
    rho = np.abs(f) ** 2
    ref = np.fft.irfftn(np.fft.rfftn(rho) * half, s=rho.shape, axes=(0, 1, 2)) * rho.size
    assert rel(phi, face_copies(ref)) < 1e-13


def test_potdd_symmetries(oracle_lib):
    # dx = dz and Nx = Nz so that kx index m and kz index m have the same |k|
    p = params(n=(32, 16, 32), h=(0.3, 0.25, 0.3), cutoff=4.0)
    with dbecpy.Solver(oracle_lib, p) as s:
        half = s.table(dbecpy.TABLE_POTDD).reshape(32, 16, 17)
    assert half[0, 0, 0] == 0.0
    # even in kx and ky: index i and N - i hold the same number
    assert np.array_equal(half[1:], half[1:][::-1])
    assert np.array_equal(half[:, 1:], half[:, 1:][:, ::-1])
    # angular factor (3 kz^2/k^2 - 1): +2 on the kz axis, -1 on the kx axis
    for m in range(1, 16):
        assert half[0, 0, m] / half[m, 0, 0] == pytest.approx(-2.0, rel=1e-13)


@pytest.mark.parametrize("mode", [dbecpy.IMAG, dbecpy.REAL])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sweep_equals_dense_crank_nicolson(oracle_lib, mode, axis):
    """calclux/y/z solve (1 - dt/2 K) psi' = (1 + dt/2 K) psi on the interior,
    K = k D2 / h^2 with k = 1/(3-par) (imaginary time) or i/(3-par) (real
    time), boundary values psi'[0] = 0 and psi'[N-1] = psi[N-1] during the
    solve, then both ends set to 0."""
    p = params(n=(12, 14, 10), h=(0.3, 0.2, 0.25), dt=0.01)
    f = smooth_field(p, seed=2, complex_=(mode == dbecpy.REAL))
    with dbecpy.Solver(oracle_lib, p, mode) as s:
        s.set_psi(f)
        [s.calclux, s.calcluy, s.calcluz][axis]()
        out = s.get_psi()
    h = (p.dx, p.dy, p.dz)[axis]
    kin = 1.0 / (3.0 - p.par) * (1j if mode == dbecpy.REAL else 1.0)
    r = p.dt / h / h * kin / 2.0
    fa = np.moveaxis(f, axis, -1)
    oa = np.moveaxis(out, axis, -1)
    n = fa.shape[-1]
    m = n - 2
    M = np.diag(np.full(m, 1 + 2 * r)) + np.diag(np.full(m - 1, -r), 1) + np.diag(np.full(m - 1, -r), -1)
    for idx in np.ndindex(fa.shape[:-1]):
        v = fa[idx]
        rhs = (1 - 2 * r) * v[1:-1] + r * (v[2:] + v[:-2])
        rhs[-1] += r * v[-1]  # psi'[N-1] = psi[N-1] while solving; psi'[0] = 0
        sol = np.linalg.solve(M, rhs)
        assert np.max(np.abs(oa[idx][1:-1] - sol)) < 1e-13 * max(1.0, np.max(np.abs(sol)))
        assert oa[idx][0] == 0 and oa[idx][-1] == 0


def test_simpson_norm_exact_for_quadratics(oracle_lib):
    """calcnorm integrates |psi|^2 = (a + x^2)(b + y^2)(c + z^2) exactly
    (Simpson is exact to degree 3, the even-N end correction to degree 2)."""
    p = params(n=(16, 20, 24), h=(0.3, 0.25, 0.2))
    x, y, z = dbecpy.grid(p)
    a, b, c = 1.3, 0.7, 2.1
    rho = (a + x * x)[:, None, None] * (b + y * y)[None, :, None] * (c + z * z)[None, None, :]

    def I(t, k):
        return k * (t[-1] - t[0]) + (t[-1] ** 3 - t[0] ** 3) / 3.0

    exact = I(x, a) * I(y, b) * I(z, c)
    with dbecpy.Solver(oracle_lib, p) as s:
        s.set_psi(np.sqrt(rho))
        n = s.calcnorm(rescale=True)
        assert n * n == pytest.approx(exact, rel=1e-13)
        # rescale multiplies by 1/norm: the new norm is 1
        assert s.calcnorm() == pytest.approx(1.0, rel=1e-14)


def test_rms_and_mu_closed_forms_uniform_field(oracle_lib):
    """psi = 1: <x^2> = int x^2 dV, and with g = gd = 0 the chemical
    potential is int pot dV (all integrands of degree 2, integrated exactly)."""
    p = params(n=(16, 20, 24), h=(0.3, 0.25, 0.2), trap=(0.5, 1.0, 1.5))
    x, y, z = dbecpy.grid(p)
    L = [t[-1] - t[0] for t in (x, y, z)]
    M2 = [(t[-1] ** 3 - t[0] ** 3) / 3.0 for t in (x, y, z)]
    with dbecpy.Solver(oracle_lib, p) as s:
        s.set_psi(np.ones((16, 20, 24)))
        r = s.calcrms()
        assert r[1] ** 2 == pytest.approx(M2[0] * L[1] * L[2], rel=1e-13)
        assert r[2] ** 2 == pytest.approx(L[0] * M2[1] * L[2], rel=1e-13)
        assert r[3] ** 2 == pytest.approx(L[0] * L[1] * M2[2], rel=1e-13)
        assert r[0] ** 2 == pytest.approx(r[1] ** 2 + r[2] ** 2 + r[3] ** 2, rel=1e-14)
        s.set_coupling(0.0, 0.0)
        mu, en = s.calcmuen()
        g2 = np.array([p.vgamma, p.vnu, p.vlambda]) ** 2
        pot_int = 0.5 * p.par * (g2[0] * M2[0] * L[1] * L[2] + g2[1] * L[0] * M2[1] * L[2]
                                 + g2[2] * L[0] * L[1] * M2[2])
        assert mu == pytest.approx(pot_int, rel=1e-13)
        assert en == pytest.approx(pot_int, rel=1e-13)


def test_gaussian_initial_observables(oracle_lib):
    """The normalised Gaussian of initpsi: rms -> (1/(2 gamma))^(1/2) etc.,
    |psi(0)|^2 -> pi^(-3/2) (gamma nu lambda)^(1/2), mu at g = 0 -> (gamma + nu +
    lambda)/2 * par (printed values: imag3d-out.txt:17, imag3d-rms.txt:6)."""
    p = params(n=(128, 96, 80), h=(0.2, 0.2, 0.2), dt=0.005, cutoff=10.0)
    with dbecpy.Solver(oracle_lib, p) as s:
        s.set_psi(dbecpy.gaussian_psi0(p))
        assert s.calcnorm(rescale=True) == pytest.approx(1.0, abs=5e-5)
        r = s.calcrms()
        assert r[1:] == pytest.approx([1.0, np.sqrt(0.5), np.sqrt(1 / 3.0)], abs=5e-6)
        assert abs(s.psi_center()) ** 2 == pytest.approx(np.pi ** -1.5 * np.sqrt(0.75), rel=1e-4)
        mu, en = s.calcmuen()
        assert mu / p.par == pytest.approx(1.5, abs=1e-3)
        assert f"{mu / p.par:.5f}" == "1.49956"


def test_real_time_conserves_norm(oracle_lib):
    """Real-time split step: calcnu has unit modulus and each CN sweep is a
    Cayley transform (unitary in the discrete l2 norm), so for a field that
    vanishes at the Dirichlet walls sum |psi|^2 is conserved to round-off (the
    Simpson-weighted norm is not an invariant of the scheme)."""
    p = params(n=(24, 24, 24), h=(0.5, 0.5, 0.5), dt=0.002, trap=(1.0, 1.0, 1.5), cutoff=4.0)
    f = dbecpy.gaussian_psi0(p) * np.exp(1j * 0.3 * dbecpy.grid(p)[0][:, None, None])
    with dbecpy.Solver(oracle_lib, p, dbecpy.REAL) as s:
        s.set_psi(f)
        s.set_coupling(40.0, 30.0)
        n0 = s.calcnorm()
        l0 = np.sum(np.abs(s.get_psi()) ** 2)
        s.step(20, want_norm=False)
        assert np.sum(np.abs(s.get_psi()) ** 2) == pytest.approx(l0, rel=1e-12)
        assert s.calcnorm() == pytest.approx(n0, rel=1e-2)


def test_imag_time_energy_decreases(oracle_lib):
    p = params(n=(24, 24, 24), h=(0.35, 0.35, 0.35), dt=0.005, cutoff=4.0)
    with dbecpy.Solver(oracle_lib, p) as s:
        s.set_psi(dbecpy.gaussian_psi0(p))
        s.calcnorm(rescale=True)
        s.set_coupling(30.0, 20.0)
        es = []
        for _ in range(6):
            s.step(10)
            es.append(s.calcmuen()[1])
        assert all(b < a for a, b in zip(es, es[1:]))


def test_convolution_equals_direct_dft_on_tiny_grid(oracle_lib):
    """On an 8^3 grid the convolution equals a brute-force O(N^2) DFT of
    |psi|^2, multiplied by potdd (mirrored to all kz) and transformed back."""
    p = params(n=(8, 8, 8), h=(0.5, 0.5, 0.5), cutoff=2.0)
    f1 = smooth_field(p, seed=3)
    with dbecpy.Solver(oracle_lib, p) as s:
        half = s.table(dbecpy.TABLE_POTDD).reshape(8, 8, 5)
        s.set_psi(f1)
        s.calcpsidd2()
        a = s.get_psidd2()
    rho = f1 ** 2
    k = np.arange(8)
    W = np.exp(-2j * np.pi * np.outer(k, k) / 8)
    F = np.einsum("ia,jb,kc,abc->ijk", W, W, W, rho)
    full = np.empty((8, 8, 8))
    full[:, :, :5] = half
    full[:, :, 5:] = half[:, :, 1:4][:, :, ::-1]
    Wi = np.conj(W)
    phi = np.einsum("ia,jb,kc,abc->ijk", Wi, Wi, Wi, F * full).real
    assert rel(a, face_copies(phi)) < 1e-13


def test_phase_energy_pins(oracle_lib):
    """calcmuen_phase: the kinetic density of the complex psi.
    * imaginary time: identical to calcmuen;
    * a positive psi times a constant phase e^{i theta}: the squares of the
      real and imaginary parts' derivatives add up to those of |psi|, so it
      equals calcmuen (which differentiates |psi|) to round-off;
    * psi = phi e^{i k x} with phi > 0: the flow adds k^2/(3 - par) per unit
      of |psi|^2, so E and mu grow by k^2 norm^2 / (3 - par), while calcmuen
      does not see the flow.  The five-point stencil's product-rule error is
      O(h^4) of phi's fifth derivative (here about 1e-3 of the flow term); it
      scales as h^4, which a grid twice as fine shows."""
    # boxes wide enough that phi is negligible where the end rules of the
    # stencil (zero, then three-point) apply
    p = params(n=(48, 40, 40), h=(0.3, 0.25, 0.2), cutoff=4.0)
    phi = np.abs(dbecpy.gaussian_psi0(p))
    with dbecpy.Solver(oracle_lib, p, dbecpy.IMAG) as s:
        s.set_psi(phi)
        s.set_coupling(30.0, 20.0)
        assert s.calcmuen_phase() == s.calcmuen()
    x = dbecpy.grid(p)[0][:, None, None]
    with dbecpy.Solver(oracle_lib, p, dbecpy.REAL) as s:
        s.set_coupling(30.0, 20.0)
        s.set_psi(phi * np.exp(0.7j))
        ref_mu, ref_en = s.calcmuen()
        mu, en = s.calcmuen_phase()
        assert mu == pytest.approx(ref_mu, rel=1e-13) and en == pytest.approx(ref_en, rel=1e-13)
        k = 0.5
        s.set_psi(phi * np.exp(1j * k * x))
        assert s.calcmuen() == pytest.approx((ref_mu, ref_en), rel=1e-13)
        norm2 = s.calcnorm() ** 2
        mu, en = s.calcmuen_phase()
        flow = k * k * norm2 / (3.0 - p.par)
        err = abs((en - ref_en) - flow) / flow
        assert err < 2e-3
        assert mu - ref_mu == pytest.approx(en - ref_en, rel=1e-12)
    # half the spacing: the discretisation error falls by about 2^4
    p2 = params(n=(96, 80, 80), h=(0.15, 0.125, 0.1), cutoff=4.0)
    phi2 = np.abs(dbecpy.gaussian_psi0(p2))
    x2 = dbecpy.grid(p2)[0][:, None, None]
    with dbecpy.Solver(oracle_lib, p2, dbecpy.REAL) as s:
        s.set_coupling(30.0, 20.0)
        s.set_psi(phi2)
        e0 = s.calcmuen_phase()[1]
        s.set_psi(phi2 * np.exp(1j * k * x2))
        flow2 = k * k * s.calcnorm() ** 2 / (3.0 - p2.par)
        err2 = abs((s.calcmuen_phase()[1] - e0) - flow2) / flow2
    assert 8.0 < err / err2 < 32.0, (err, err2)
