"""The serving stream's fixture (tests/golden/serve_stream.json): it is what
its generator draws, and its requests have the stated distribution.  CPU only."""
import importlib.util
import json
import os

import numpy as np

import dbecpy
import dbecserve as ds
from conftest import GOLDEN

PATH = os.path.join(GOLDEN, "serve_stream.json")


def generator():
    spec = importlib.util.spec_from_file_location("make_serve_stream", os.path.join(GOLDEN, "make_serve_stream.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_fixture_is_the_generators_draw():
    g = generator()
    d = json.load(open(PATH))
    assert d["requests"] == g.draw(d["count"], d["rate_per_s"], d["seed"])
    assert {k: d[k] for k in g.SETTINGS} == g.SETTINGS


def test_stream_distribution():
    d = json.load(open(PATH))
    rows = np.array(d["requests"])
    assert len(rows) == 2000
    t = rows[:, 0]
    assert np.all(np.diff(t) > 0)
    assert abs(len(t) / t[-1] - 20.0) < 1.0  # Poisson rate 20/s
    assert set(rows[:, 1].astype(int)) == {48, 64, 96, 128, 160, 192}
    assert rows[:, 2].min() >= 5 and rows[:, 2].max() <= 60
    assert rows[:, 3].min() >= 0 and rows[:, 3].max() <= 40
    assert rows[:, 4:7].min() >= 0.3 and rows[:, 4:7].max() <= 3.0
    assert abs(rows[:, 7].mean() - 0.15) < 0.03


def test_load_stream_requests():
    st = ds.load_stream(PATH, 50, prec=dbecpy.MIXED)
    assert len(st) == 50
    for i, (t, r) in enumerate(st):
        assert r.id == i and r.prec == dbecpy.MIXED and r.nx == r.ny == r.nz
        assert r.dx == 16.0 / r.nx and r.cutoff == 8.0 and r.par == 2.0
        if r.mode == dbecpy.REAL:
            assert r.dt == 0.001 and r.max_steps == 5000
        else:
            assert r.dt == 0.01 and r.max_steps == 20000 and r.mutol == 1e-8
