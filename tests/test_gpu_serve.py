"""The request loop (include/dbec_serve.h) on the GPU.

* a served FP64 ground state equals the same solve run alone through dbec.h
  (host Gaussian, dbec_step, calcmuen every check) to 1e-10 in mu, E, rms;
* a solve's result is the same alone and in a mix of other sizes, kinds and
  couplings, bit for bit (no kernel mixes data of two solves);
* a memory budget that holds one bucket at a time still serves a stream of
  two sizes, evicting idle buckets, never holding more than the budget;
* mixed precision ground states agree with FP64 to 1e-6 in mu (the error of
  the FP32 bulk arithmetic, stated per result); a collapsing, strongly
  dipolar request ends (converged to the grid-scale state or flagged
  diverged) instead of running to its step cap.
"""
import os
import sys

import numpy as np
import pytest

import dbecpy
import dbecserve as ds

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def stable(rid, n, prec=dbecpy.FP64, mode=dbecpy.IMAG, g0=20.0, gd0=8.0, trap=(0.8, 1.0, 1.6), steps=4000,
           dt=None):
    if dt is None:
        dt = 0.01 if mode == dbecpy.IMAG else 0.001
    return ds.make_request(rid, n, mode, prec, g0, gd0, trap, dt, steps if mode == dbecpy.IMAG else 300,
                           mutol=1e-8)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_served_equals_plain_fp64(hip_lib):
    say("served FP64 ground states vs plain dbec_step solves")
    reqs = [stable(0, 32), stable(1, 48, g0=30.0, gd0=5.0, trap=(1.2, 0.7, 1.5)),
            stable(2, 32, mode=dbecpy.REAL, g0=10.0, gd0=4.0)]
    with ds.Server(hip_lib, max_batch=4, check_every=50) as srv:
        got = srv.run_all(reqs, timeout=120)
    for r in reqs:
        a, b = got[r.id], ds.plain_solve(hip_lib, r, check_every=50)
        say(f"  id {r.id}: status {a['status']} steps {a['steps']} mu {a['mu']:.12g} (plain {b['steps']} {b['mu']:.12g})")
        assert a["status"] == b["status"] and a["steps"] == b["steps"]
        assert rel(a["mu"], b["mu"]) < 1e-10 and rel(a["en"], b["en"]) < 1e-10
        assert np.allclose(a["rms"], b["rms"], rtol=1e-10, atol=0)
        if r.mode == dbecpy.IMAG:
            assert a["status"] == ds.CONVERGED


def test_alone_equals_in_the_mix(hip_lib):
    say("alone vs in the mix")
    probe = [stable(100, 48, prec=dbecpy.MIXED), stable(101, 48, prec=dbecpy.MIXED, mode=dbecpy.REAL)]
    others = [stable(200, 48, prec=dbecpy.MIXED, g0=40.0, gd0=20.0, trap=(2.0, 0.5, 1.0)),
              stable(201, 64, prec=dbecpy.MIXED, g0=15.0),
              stable(202, 32, prec=dbecpy.MIXED, mode=dbecpy.REAL),
              stable(203, 48, prec=dbecpy.MIXED, mode=dbecpy.REAL, g0=25.0, gd0=12.0),
              stable(204, 48, prec=dbecpy.FP64)]
    with ds.Server(hip_lib, max_batch=8, check_every=40) as srv:
        alone = srv.run_all(probe, timeout=120)
    with ds.Server(hip_lib, max_batch=8, check_every=40) as srv:
        mix = srv.run_all(others[:2] + probe[:1] + others[2:] + probe[1:], timeout=120)
    for r in probe:
        a, b = alone[r.id], mix[r.id]
        say(f"  id {r.id}: steps {a['steps']} mu {a['mu']!r} / {b['mu']!r}")
        for k in ("status", "steps", "mu", "en", "norm"):
            assert a[k] == b[k], k
        assert a["rms"] == b["rms"]


def test_memory_budget_evicts_idle_buckets(hip_lib):
    say("memory budget, two sizes")
    reqs = [stable(300 + i, 32 if i % 2 == 0 else 48, prec=dbecpy.MIXED, steps=400) for i in range(6)]
    with ds.Server(hip_lib, max_batch=2, check_every=50) as probe:
        probe.run_all(reqs[1:2], timeout=60)
        big, _ = probe.usage()
    with ds.Server(hip_lib, max_batch=2, check_every=50, mem_budget=big * 1.05) as srv:
        for r in reqs:
            srv.submit(r)
        got, peak = {}, 0.0
        while srv.pending() > 0:
            for x in srv.progress(True):
                got[x["id"]] = x
            peak = max(peak, srv.usage()[0])
        assert peak <= big * 1.05
    assert sorted(got) == [r.id for r in reqs]
    assert all(x["status"] in (ds.CONVERGED, ds.NOT_CONVERGED) for x in got.values())


def test_mixed_close_to_fp64_and_collapse_stays_finite(hip_lib):
    say("mixed (+ FP64 polish) vs FP64 ground state; collapsing request")
    r64, r32 = stable(400, 48), stable(401, 48, prec=dbecpy.MIXED)
    hard = ds.make_request(402, 48, dbecpy.IMAG, dbecpy.MIXED, 5.0, 40.0, (0.5, 0.5, 2.5), 0.01, 3000, 1e-8)
    with ds.Server(hip_lib, max_batch=4, check_every=50, polish_tol=0.0) as srv:
        pure = srv.run_all([stable(403, 48, prec=dbecpy.MIXED)], timeout=120)[403]
    with ds.Server(hip_lib, max_batch=4, check_every=50, polish_tol=1e-6) as srv:
        got = srv.run_all([r64, r32, hard], timeout=120)
    a, b = got[400], got[401]
    say(f"  FP64 mu {a['mu']:.12g}; mixed {pure['mu']:.12g} rel {rel(pure['mu'], a['mu']):.2e}; "
        f"mixed + polish {b['mu']:.12g} rel {rel(b['mu'], a['mu']):.2e} ({b['steps_fp64']} of {b['steps']} steps FP64)")
    assert pure["steps_fp64"] == 0 and rel(pure["mu"], a["mu"]) < 1e-6 and rel(pure["en"], a["en"]) < 1e-6
    assert b["status"] == ds.CONVERGED and 0 < b["steps_fp64"] < b["steps"]
    assert rel(b["mu"], a["mu"]) < 1e-8 and rel(b["en"], a["en"]) < 1e-8
    h = got[402]
    say(f"  collapsing request: status {h['status']} steps {h['steps']} mu {h['mu']:.6g}")
    assert h["status"] in (ds.CONVERGED, ds.DIVERGED)


def test_polish_migration_then_slot_reuse(hip_lib):
    """Mixed slots of one per bucket: when a ground state moves to its FP64
    polish, the freed mixed slot takes the next queued request in the same
    progress call.  The widen that reads the slot is ordered before the new
    request's initial state is written, so both results equal the same
    requests served alone, bit for bit."""
    say("polish migration, then the freed slot is reused")
    a = stable(500, 48, prec=dbecpy.MIXED)
    b = stable(501, 48, prec=dbecpy.MIXED, g0=25.0, gd0=6.0, trap=(0.9, 1.1, 1.4))
    alone = {}
    for r in (a, b):
        with ds.Server(hip_lib, max_batch=1, check_every=50, polish_tol=1e-6) as srv:
            alone.update(srv.run_all([r], timeout=120))
    with ds.Server(hip_lib, max_batch=1, check_every=50, polish_tol=1e-6) as srv:
        both = srv.run_all([a, b], timeout=120)
    for r in (a, b):
        x, y = alone[r.id], both[r.id]
        say(f"  id {r.id}: {x['steps']} steps ({x['steps_fp64']} FP64) mu {x['mu']!r} / {y['mu']!r}")
        assert x["steps_fp64"] > 0
        for k in ("status", "steps", "steps_fp64", "mu", "en", "norm"):
            assert x[k] == y[k], k
        assert x["rms"] == y["rms"]


STREAM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "serve_stream.json")


def stream_slice():
    """From the serving stream: the first three ground states and the first
    two real-time runs with N in {96, 128, 160, 192}, in stream order."""
    imag, real = [], []
    for _, r in ds.load_stream(STREAM, prec=dbecpy.MIXED):
        if r.nx < 96:
            continue
        if r.mode == dbecpy.IMAG and len(imag) < 3:
            imag.append(r)
        elif r.mode == dbecpy.REAL and len(real) < 2:
            real.append(r)
        if len(imag) == 3 and len(real) == 2:
            return imag + real
    raise AssertionError("stream slice not found")


def test_stream_slice_alone_equals_in_the_mix(hip_lib):
    """Five requests of the serving stream at its larger grids (96-192), both
    modes, mixed precision with FP64 polish: served together and each alone,
    every result is the same bit for bit."""
    reqs = stream_slice()
    say("stream slice: " + ", ".join(f"id {r.id} N {r.nx} {'real' if r.mode else 'imag'}" for r in reqs))
    with ds.Server(hip_lib, max_batch=8, check_every=50) as srv:
        mix = srv.run_all(reqs, timeout=300)
    for r in reqs:
        with ds.Server(hip_lib, max_batch=8, check_every=50) as srv:
            a = srv.run_all([r], timeout=300)[r.id]
        b = mix[r.id]
        say(f"  id {r.id}: status {a['status']} steps {a['steps']} ({a['steps_fp64']} FP64) mu {a['mu']!r}")
        for k in ("status", "steps", "steps_fp64"):
            assert a[k] == b[k], (r.id, k)
        # bit for bit; a diverged (collapsing) solve reports NaN in both runs
        for k in ("mu", "en", "norm"):
            assert np.array_equal(np.float64(a[k]), np.float64(b[k]), equal_nan=True), (r.id, k)
        assert np.array_equal(np.array(a["rms"]), np.array(b["rms"]), equal_nan=True)


def test_stream_slice_fp64_against_oracle(hip_lib, oracle_lib):
    """A ground state and a real-time run of the stream slice in FP64, 100
    steps each, served against the same requests solved by the oracle: mu,
    E and rms at 1e-10."""
    reqs = stream_slice()
    picks = [r for r in reqs if r.mode == dbecpy.IMAG][:1] + [r for r in reqs if r.mode == dbecpy.REAL][:1]
    fp64 = [ds.make_request(r.id, r.nx, r.mode, dbecpy.FP64, r.g / r.par, r.gd / r.par,
                            (r.vgamma, r.vnu, r.vlambda), r.dt, 100, mutol=1e-300, cutoff=r.cutoff)
            for r in picks]
    with ds.Server(hip_lib, max_batch=4, check_every=50) as srv:
        got = srv.run_all(fp64, timeout=300)
    for r in fp64:
        a, b = got[r.id], ds.plain_solve(oracle_lib, r, check_every=50)
        say(f"  id {r.id} N {r.nx}: mu {a['mu']!r} oracle {b['mu']!r}")
        assert a["steps"] == b["steps"] == 100
        assert rel(a["mu"], b["mu"]) < 1e-10 and rel(a["en"], b["en"]) < 1e-10
        assert np.allclose(a["rms"], b["rms"], rtol=1e-10, atol=0)
