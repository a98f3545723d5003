"""The oracle against the reference's own shipped outputs.

The reference ships the summary, RMS, real-time rms series and 1-D density
files of its example inputs (output/imag3d-mpicuda, output/real3d-mpicuda; a
16-rank GPU run with cuFFT).  The oracle, driven by the same command-line
program as the product, reruns the imag3d -> imag3d-den.bin -> real3d chain
and must agree with every printed digit and with the full-precision density
vectors.
"""
from conftest import GOLDEN  # noqa: F401
from golden_cli import check_chain, run_chain


def test_oracle_chain_matches_reference_outputs(oracle_bin, tmp_path):
    work = run_chain(oracle_bin, str(tmp_path))
    rep = check_chain(work, den_tol=1e-13)
    # the oracle reproduces every printed digit (no last-digit flips)
    flips = {k: v for k, v in rep.items() if k.endswith(".txt") and v}
    assert not flips, flips
