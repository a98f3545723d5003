"""The C ABI library: loads, exports every entry point declared in
include/dbec.h, and fails cleanly (status code + message, no exit) on bad
arguments.  No compute calls: runs on the CPU."""
import ctypes
import os
import re

import pytest

import dbecpy
from conftest import REPO, params


def declared():
    hdr = open(os.path.join(REPO, "include", "dbec.h")).read()
    return sorted(set(re.findall(r"\b(dbec_[a-z0-9_]+)\s*\(", hdr)))


def test_header_matches_binding():
    assert declared() == sorted(dbecpy.SYMBOLS)


def declared_serve():
    hdr = open(os.path.join(REPO, "include", "dbec_serve.h")).read()
    return sorted(set(re.findall(r"\b(dbec_server_[a-z0-9_]+)\s*\(", hdr)))


def test_serve_header_matches_binding_and_library():
    import dbecserve
    assert declared_serve() == sorted(dbecserve.SERVE_SYMBOLS)
    if not os.path.exists(dbecpy.HIP_LIB):
        pytest.fail("libdbec_hip.so is not built")
    lib = ctypes.CDLL(dbecpy.HIP_LIB, mode=ctypes.RTLD_LOCAL)
    for name in declared_serve():
        assert hasattr(lib, name), name


@pytest.mark.parametrize("which", ["product", "oracle"])
def test_library_exports_every_symbol(which, oracle_lib):
    path = dbecpy.HIP_LIB if which == "product" else os.path.join(REPO, "oracle", "liboracle.so")
    if which == "product" and not os.path.exists(path):
        pytest.fail("libdbec_hip.so is not built")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    for name in declared():
        assert hasattr(lib, name), name


def test_product_rejects_bad_parameters():
    lib = dbecpy.load()
    p = params(n=(30, 32, 32))  # 30 is not a supported transform length
    with pytest.raises(dbecpy.DbecError) as e:
        dbecpy.Solver(lib, p)
    assert e.value.code in (dbecpy.DBEC_EUNSUPPORTED, dbecpy.DBEC_EDEVICE)
    p = params(n=(32, 32, 32), dt=-1.0)
    with pytest.raises(dbecpy.DbecError) as e:
        dbecpy.Solver(lib, p)
    assert e.value.code in (dbecpy.DBEC_EINVAL, dbecpy.DBEC_EDEVICE)


def test_product_without_device_reports_error():
    """With no HIP device (this container) creation fails with a status and a
    message instead of exiting; with a device it succeeds."""
    lib = dbecpy.load()
    p = params(n=(32, 32, 32))
    try:
        s = dbecpy.Solver(lib, p)
    except dbecpy.DbecError as e:
        assert e.code == dbecpy.DBEC_EDEVICE
        assert "device" in str(e).lower()
    else:
        s.close()
