"""The product programs' extension keys on the GPU (OBSSTP, OBSOUT, MUTOL,
PRECISION, POLISH), against the oracle-linked programs and against the FP64
product run.  The cadence observables, the MUTOL stop point, the summary
lines and the density files of FP64 runs agree with the oracle; a mixed run
with an FP64 polish ends at the FP64 ground state's printed mu and E."""
import sys

import numpy as np
import pytest

import dbecpy
from test_cli import BASE, phase_lines, run, write

pytestmark = pytest.mark.gpu

CASE = dict(BASE, NX="32", NY="32", NZ="32", DX="0.5", DY="0.5", DZ="0.5", DT="0.01", G0="20.", GDD0="10.",
            CUTOFF="8.", NPAS="4000", NRUN="100", OUTPUT="out", RMSOUT="rms", NRUNOUT="den", OUTSTPX="1",
            OUTSTPY="1", OUTSTPZ="1", OBSSTP="50", OBSOUT="obs")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def obs_rows(path):
    return np.array([list(map(float, l.split())) for l in open(path) if not l.startswith("#")])


def summary_numbers(path):
    """The numbers of the phase-end summary lines ("After ... iter.:")."""
    out = []
    for l in phase_lines(path):
        if l.startswith("After"):
            out.append([float(t) for t in l.split(":")[1].split()])
    return np.array(out)


def run_case(bindir, d, tmp):
    tmp.mkdir()
    write(tmp / "in", d)
    r = run(bindir, "imag3d", ["-i", "in"], tmp)
    assert r.returncode == 0, r.stderr
    return tmp


def test_cadence_and_mutol_match_oracle(oracle_bin, tmp_path):
    """FP64 with OBSSTP = 50, MUTOL = 1e-8: the product stops NPAS at the
    same step as the oracle, and every cadence row, the summary and the
    1-D densities agree to 1e-10."""
    say("imag3d OBSSTP/OBSOUT/MUTOL: product vs oracle")
    d = dict(CASE, MUTOL="1e-8")
    a = run_case(dbecpy.BIN_DIR, d, tmp_path / "product")
    b = run_case(oracle_bin, d, tmp_path / "oracle")
    stop_a = [l for l in phase_lines(a / "out.txt") if "converged" in l]
    stop_b = [l for l in phase_lines(b / "out.txt") if "converged" in l]
    say(f"  product: {stop_a}; oracle: {stop_b}")
    # NPAS and NRUN both stop on MUTOL; the stop steps must be the oracle's
    assert len(stop_a) == len(stop_b) == 2
    assert [l.split(":")[0] for l in stop_a] == [l.split(":")[0] for l in stop_b]
    ra, rb = obs_rows(a / "obs.txt"), obs_rows(b / "obs.txt")
    assert ra.shape == rb.shape and np.array_equal(ra[:, 0], rb[:, 0])
    assert np.allclose(ra[:, 2:], rb[:, 2:], rtol=1e-10, atol=0)
    sa, sb = summary_numbers(a / "out.txt"), summary_numbers(b / "out.txt")
    assert np.allclose(sa, sb, rtol=0, atol=1.01e-5)  # printed to 4-5 decimals
    for ax in "xyz":
        da = np.fromfile(a / f"den1d_{ax}.bin")
        db = np.fromfile(b / f"den1d_{ax}.bin")
        assert np.allclose(da, db, rtol=1e-10, atol=1e-14 * np.abs(db).max())


def test_mixed_with_polish_reaches_the_fp64_ground_state(tmp_path):
    """PRECISION = MIXED with MUTOL = 1e-6 (the FP32 floor) and POLISH = 400
    FP64 steps, against the FP64 product run converged to MUTOL = 1e-8: the
    printed NRUN mu, E and rms agree to the printed digits (1e-5)."""
    say("imag3d PRECISION = MIXED + POLISH vs FP64")
    f = run_case(dbecpy.BIN_DIR, dict(CASE, MUTOL="1e-8"), tmp_path / "fp64")
    m = run_case(dbecpy.BIN_DIR, dict(CASE, MUTOL="1e-6", PRECISION="MIXED", POLISH="400"), tmp_path / "mixed")
    assert any("POLISH: 400 FP64 steps" in l for l in phase_lines(m / "out.txt"))
    sf, sm = summary_numbers(f / "out.txt"), summary_numbers(m / "out.txt")
    say(f"  NRUN fp64 {sf[-1]}; mixed + polish {sm[-1]}")
    assert np.allclose(sm[-1], sf[-1], rtol=0, atol=2.01e-5)
    rows = obs_rows(m / "obs.txt")
    assert rows.shape[1] == 9  # imaginary time: no phase-energy column


def test_real_time_cadence_matches_oracle(oracle_bin, tmp_path):
    """real3d with OBSSTP from an imaginary-time ground state (imag3d-den.bin
    of the product's imag3d), gd quenched by GDDPAR = 1.2: the cadence rows
    of the product and the oracle programs agree to 1e-10, each row carries
    the reference's E (of |psi|) and the energy with the phase gradient, and
    the flow makes the latter the larger."""
    import shutil

    say("real3d OBSSTP: product vs oracle, phase energy column")
    im = run_case(dbecpy.BIN_DIR, dict(CASE, GDD0="4.", MUTOL="1e-8", NRUN="0", NRUNOUT="imag3d-den"),
                  tmp_path / "imag")
    r = dict(BASE, NX="32", NY="32", NZ="32", DX="0.5", DY="0.5", DZ="0.5", DT="0.001", G0="20.", GDD0="4.",
             CUTOFF="8.", NSTP="0", NPAS="0", NRUN="2000", GPAR="1.", GDDPAR="1.2", OUTPUT="out", OBSSTP="200",
             OBSOUT="obs")
    rows = []
    for name, bindir in (("product", dbecpy.BIN_DIR), ("oracle", oracle_bin)):
        d = tmp_path / name
        d.mkdir()
        shutil.copy(im / "imag3d-den.bin", d / "imag3d-den.bin")
        write(d / "in", r)
        res = run(bindir, "real3d", ["-i", "in"], d)
        assert res.returncode == 0, res.stdout + res.stderr
        rows.append(obs_rows(d / "obs.txt"))
    a, b = rows
    assert a.shape == b.shape == (10, 10)
    assert np.allclose(a, b, rtol=1e-10, atol=0)
    norm, e_ref, e_phase = a[:, 2], a[:, 4], a[:, 9]
    say(f"  norm {norm.min():.6f}..{norm.max():.6f}; E ref {e_ref[0]:.6f}..{e_ref[-1]:.6f}; "
        f"E phase {e_phase[0]:.6f}..{e_phase[-1]:.6f}")
    assert np.all(np.abs(norm - 1.0) < 1e-3)
    assert np.all(e_phase > e_ref)
