"""Run the imag3d -> imag3d-den.bin -> real3d chain of a pair of programs on the
reference's example inputs and compare with the reference's shipped outputs
(tests/golden/ref_*)."""
import os
import re
import shutil
import subprocess

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run_chain(bin_dir, work, env=None, timeout=900):
    os.makedirs(work, exist_ok=True)
    for f in ("imag3d-input", "real3d-input"):
        shutil.copy(os.path.join(GOLDEN, f), work)
    for prog in ("imag3d", "real3d"):
        r = subprocess.run([os.path.join(bin_dir, prog), "-i", f"{prog}-input"], cwd=work, env=env,
                           capture_output=True, text=True, timeout=timeout)
        assert r.returncode == 0, f"{prog} failed: {r.stdout}\n{r.stderr}"
    return work


_num = re.compile(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?")


def _last_place(tok):
    """Size of one unit in the last printed digit of a %f / %e token."""
    m = re.match(r"[-+]?\d+\.(\d+)(?:[eE]([-+]?\d+))?$", tok)
    digits, exp = len(m.group(1)), int(m.group(2) or 0)
    return 10.0 ** (exp - digits)


def compare_text(ours_path, ref_path):
    """Line-by-line comparison; numbers may differ by at most one unit in the
    last printed digit.  Timing lines and the MPI rank count are skipped.
    Returns (mismatches, last-digit flips)."""
    ours = open(ours_path).read().splitlines()
    ref = open(ref_path).read().splitlines()
    bad, flips = [], 0
    assert len(ours) == len(ref), f"{ours_path}: {len(ours)} lines vs {len(ref)}"
    for a, b in zip(ours, ref):
        if "wall-clock time" in b:
            assert a.split(":")[0] == b.split(":")[0]
            continue
        if "MPI_NUM_PROCS" in b:
            a, b = a.rsplit("=", 1)[0], b.rsplit("=", 1)[0]
        ta, tb = _num.findall(a), _num.findall(b)
        if _num.sub("#", a) != _num.sub("#", b) or len(ta) != len(tb):
            bad.append((a, b))
            continue
        for x, y in zip(ta, tb):
            if x == y:
                continue
            if abs(float(x) - float(y)) <= 1.0001 * _last_place(y):
                flips += 1
            else:
                bad.append((a, b))
                break
    return bad, flips


def den1d_rel(ours_path, ref_path):
    a = np.fromfile(ours_path, "<f8").reshape(-1, 2)
    b = np.fromfile(ref_path, "<f8").reshape(-1, 2)
    assert a.shape == b.shape
    assert np.array_equal(a[:, 0], b[:, 0])
    return float(np.max(np.abs(a[:, 1] - b[:, 1])) / np.max(np.abs(b[:, 1])))


def check_chain(work, den_tol):
    """Assert the chain's outputs agree with the reference's; returns a report."""
    rep = {}
    for name, sub in (("imag3d-out.txt", "ref_imag3d"), ("imag3d-rms.txt", "ref_imag3d"),
                      ("real3d-out.txt", "ref_real3d"), ("real3d-rms.txt", "ref_real3d"),
                      ("real3d-dyna.txt", "ref_real3d")):
        bad, flips = compare_text(os.path.join(work, name), os.path.join(GOLDEN, sub, name))
        assert not bad, f"{name}: {bad[:3]}"
        rep[name] = flips
    for pfx, sub in (("imag3d-den1d", "ref_imag3d"), ("real3d-den1d", "ref_real3d")):
        for ax in "xyz":
            f = f"{pfx}_{ax}.bin"
            r = den1d_rel(os.path.join(work, f), os.path.join(GOLDEN, sub, f))
            assert r < den_tol, f"{f}: rel {r}"
            rep[f] = r
    return rep
