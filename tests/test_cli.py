"""Command-line contract of imag3d / real3d (imag3d.cu:71-79, 650-829;
real3d.cu:675-883, 925-931): usage text, parameter-file errors, required and
conditional keys, exit codes.  Every case fails before any device work, so it
runs on the CPU for both the product programs and the oracle-linked ones."""
import os
import subprocess

import pytest

import dbecpy
from conftest import ORACLE_BIN

BASE = {
    "OPTION": "2", "NATOMS": "5000", "G0": "40.", "GDD0": "30.", "CUTOFF": "8.", "NX": "16", "NY": "16",
    "NZ": "16", "DX": "0.5", "DY": "0.5", "DZ": "0.5", "DT": "0.002", "GAMMA": "0.5", "NU": "1.",
    "LAMBDA": "1.5", "NSTP": "0", "NPAS": "2", "NRUN": "0",
}


def bins():
    out = []
    if os.path.exists(os.path.join(dbecpy.BIN_DIR, "imag3d")):
        out.append(dbecpy.BIN_DIR)
    out.append(ORACLE_BIN)
    return out


def write(path, d, extra=""):
    with open(path, "w") as f:
        f.write("# test parameters\n")
        for k, v in d.items():
            f.write(f"   {k} = {v}\n")
        f.write(extra)


def run(bindir, prog, args, cwd):
    return subprocess.run([os.path.join(bindir, prog)] + args, cwd=cwd, capture_output=True, text=True,
                          timeout=60)


@pytest.fixture(params=bins(), ids=lambda b: "product" if "amd" in b else "oracle")
def bindir(request, oracle_bin):
    return request.param


@pytest.mark.parametrize("prog", ["imag3d", "real3d"])
@pytest.mark.parametrize("args", [[], ["-x", "f"], ["-i"], ["-i", "a", "b"]])
def test_usage(bindir, prog, args, tmp_path):
    r = run(bindir, prog, args, tmp_path)
    assert r.returncode == 1
    assert r.stderr == f"Usage: {os.path.join(bindir, prog)} -i <input parameter file> \n"


@pytest.mark.parametrize("prog", ["imag3d", "real3d"])
def test_wrong_parameter_file(bindir, prog, tmp_path):
    r = run(bindir, prog, ["-i", "missing-file"], tmp_path)
    assert r.returncode == 1 and r.stderr == "Wrong input parameter file.\n"
    # a file without any "KEY = VALUE" line counts as wrong too (cfg_init returns 0)
    (tmp_path / "empty").write_text("# nothing\nKEY=VALUE\n")
    r = run(bindir, prog, ["-i", "empty"], tmp_path)
    assert r.returncode == 1 and r.stderr == "Wrong input parameter file.\n"


MISSING = [
    ("OPTION", "OPTION is not defined in the configuration file\n"),
    ("NATOMS", "NATOMS is not defined in the configuration file.\n"),
    ("NX", "NX is not defined in the configuration file.\n"),
    ("NZ", "Nz is not defined in the configuration file.\n"),
    ("DT", "DT is not defined in the configuration file.\n"),
    ("LAMBDA", "LAMBDA is not defined in the configuration file.\n"),
    ("NRUN", "NRUN is not defined in the configuration file.\n"),
    ("CUTOFF", "CUTOFF is not defined in the configuration file.\n"),
]


@pytest.mark.parametrize("key,msg", MISSING)
def test_missing_required_key(bindir, key, msg, tmp_path):
    d = dict(BASE)
    del d[key]
    write(tmp_path / "in", d)
    r = run(bindir, "imag3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1 and r.stderr == msg


def test_g0_fallback_needs_aho_and_as(bindir, tmp_path):
    d = dict(BASE)
    del d["G0"]
    write(tmp_path / "in", d)
    r = run(bindir, "imag3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1 and r.stderr == "AHO is not defined in the configuration file.\n"
    d["AHO"] = "1.e-6"
    write(tmp_path / "in", d)
    r = run(bindir, "imag3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1 and r.stderr == "AS is not defined in the configuration file.\n"


def test_real3d_needs_gpar_gddpar_outstpt(bindir, tmp_path):
    d = dict(BASE)
    write(tmp_path / "in", d)
    r = run(bindir, "real3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1 and r.stderr == "GPAR is not defined in the configuration file.\n"
    d["GPAR"] = "1.5"
    write(tmp_path / "in", d)
    r = run(bindir, "real3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1 and r.stderr == "GDDPAR is not defined in the configuration file.\n"
    d["GDDPAR"] = "1."
    d["DYNAOUT"] = "dyn"
    write(tmp_path / "in", d)
    r = run(bindir, "real3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1 and r.stderr == "OUTSTPT is not defined in the configuration file.\n"


def test_outstp_required_with_density_output(bindir, tmp_path):
    d = dict(BASE, NRUNOUT="den")
    write(tmp_path / "in", d)
    r = run(bindir, "imag3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1 and r.stderr == "OUTSTPX is not defined in the configuration file.\n"


@pytest.mark.parametrize("val", ["3", "-1"])
def test_potmem_range(bindir, val, tmp_path):
    write(tmp_path / "in", dict(BASE, POTMEM=val))
    r = run(bindir, "imag3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1 and r.stderr == "POTMEM can be 0, 1 or 2.\n"


def test_first_occurrence_wins_and_spacing(oracle_bin, tmp_path):
    """cfg_read returns the first match; "KEY=VALUE" without spaces is not a
    key line (sscanf "%s = %s", src/utils/cfg.c:43)."""
    d = dict(BASE, NPAS="1")
    write(tmp_path / "in", d, extra="   NPAS = 9999\nNRUN=5\n   OUTPUT = o\n")
    r = run(oracle_bin, "imag3d", ["-i", "in"], tmp_path)
    assert r.returncode == 0, r.stderr
    out = (tmp_path / "o.txt").read_text()
    assert "NSTP = 0, NPAS = 1, NRUN = 0" in out


def test_real3d_without_imag3d_den(bindir, tmp_path):
    write(tmp_path / "in", dict(BASE, GPAR="1.5", GDDPAR="1."))
    r = run(bindir, "real3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1
    assert r.stdout.endswith("Run the program using the input file to read. e.g.: imag3d-den.bin\n")


# ---- extension keys (absent from the reference; defaults leave its behaviour)

@pytest.mark.parametrize("extra,msg", [
    ({"MUTOL": "1e-8"}, "OBSSTP is not defined in the configuration file.\n"),
    ({"OBSOUT": "obs"}, "OBSSTP is not defined in the configuration file.\n"),
    ({"OBSOUT": "obs", "OBSSTP": "0"}, "OBSSTP must be positive.\n"),
    ({"MUTOL": "-1", "OBSSTP": "5"}, "MUTOL must be positive.\n"),
    ({"PRECISION": "FP16"}, "PRECISION can be FP64 or MIXED.\n"),
    ({"POLISH": "10"}, "POLISH applies to an imaginary-time run with PRECISION = MIXED.\n"),
])
def test_extension_key_errors(bindir, extra, msg, tmp_path):
    write(tmp_path / "in", dict(BASE, **extra))
    r = run(bindir, "imag3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1 and r.stderr == msg


def test_mutol_real_time_rejected(bindir, tmp_path):
    write(tmp_path / "in", dict(BASE, GPAR="1.", GDDPAR="1.", MUTOL="1e-8", OBSSTP="5"))
    r = run(bindir, "real3d", ["-i", "in"], tmp_path)
    assert r.returncode == 1 and r.stderr == "MUTOL applies to imaginary time only.\n"


CONV = dict(BASE, G0="20.", GDD0="10.", CUTOFF="4.", DX="0.6", DY="0.6", DZ="0.6", DT="0.01", NPAS="3000",
            NRUN="0", OUTPUT="out")


def phase_lines(path):
    return [l for l in open(path).read().splitlines() if l.startswith(("After", " *"))]


def test_mutol_stops_npas_where_a_fixed_count_ends(oracle_bin, tmp_path):
    """MUTOL ends NPAS at the first OBSSTP point with |dmu/mu| <= MUTOL; the
    state is the one a run of exactly that many NPAS steps reaches (the
    observables at the cadence leave psi unchanged)."""
    a = tmp_path / "a"
    a.mkdir()
    write(a / "in", dict(CONV, OBSSTP="50", OBSOUT="obs", MUTOL="1e-7"))
    r = run(oracle_bin, "imag3d", ["-i", "in"], a)
    assert r.returncode == 0, r.stderr
    lines = phase_lines(a / "out.txt")
    stop = [l for l in lines if "NPAS converged after" in l]
    assert len(stop) == 1, lines
    n = int(stop[0].split("after")[1].split()[0])
    assert 0 < n < 3000 and n % 50 == 0
    rows = [list(map(float, l.split())) for l in open(a / "obs.txt") if not l.startswith("#")]
    assert int(rows[-1][0]) == n
    mu1, mu0 = rows[-1][3], rows[-2][3]
    assert abs(mu1 - mu0) <= 1e-7 * abs(mu1)
    assert all(abs(r1[3] - r0[3]) > 1e-7 * abs(r1[3]) for r0, r1 in zip(rows[:-2], rows[1:-1]))
    b = tmp_path / "b"
    b.mkdir()
    write(b / "in", dict(CONV, NPAS=str(n)))
    r = run(oracle_bin, "imag3d", ["-i", "in"], b)
    assert r.returncode == 0, r.stderr
    assert [l for l in lines if l.startswith("After")] == phase_lines(b / "out.txt")


def test_obs_cadence_leaves_results_unchanged(oracle_bin, tmp_path):
    out = []
    for name, extra in (("plain", {}), ("obs", {"OBSSTP": "7", "OBSOUT": "obs"})):
        d = tmp_path / name
        d.mkdir()
        write(d / "in", dict(CONV, NPAS="40", NRUN="9", NRUNOUT="den", OUTSTPX="1", OUTSTPY="1", OUTSTPZ="1",
                             **extra))
        r = run(oracle_bin, "imag3d", ["-i", "in"], d)
        assert r.returncode == 0, r.stderr
        out.append((phase_lines(d / "out.txt"), (d / "den1d_x.bin").read_bytes()))
    assert out[0] == out[1]
    rows = [l for l in open(tmp_path / "obs" / "obs.txt") if not l.startswith("#")]
    assert [int(l.split()[0]) for l in rows] == [7, 14, 21, 28, 35, 47]
