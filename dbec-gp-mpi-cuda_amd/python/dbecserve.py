"""ctypes binding of the request loop (include/dbec_serve.h) and the
mixed-size request stream used to measure it.

The server batches solves of equal grid/dx/dt/cutoff/mode/precision (a size
bucket) into shared launches and admits requests into free slots between
chunks of steps.  `load_stream` reads the serving stream from its JSON
fixture (tests/golden/serve_stream.json, written once by
tests/golden/make_serve_stream.py): Poisson arrivals at 20/s, cubic grids N in
{48, 64, 96, 128, 160, 192} with dx = 16/N, G0 ~ U[5, 60], GDD0 ~ U[0, 40],
gamma, nu, lambda ~ U[0.3, 3], 85 % ground states to |dmu/mu| < 1e-8 and
15 % real-time runs, together with the settings the stream fixes (dt, cutoff,
step counts).
"""
import ctypes
import os
import time

import numpy as np

import dbecpy

SERVE_SYMBOLS = [
    "dbec_server_create", "dbec_server_destroy", "dbec_server_time", "dbec_server_submit",
    "dbec_server_progress", "dbec_server_pending", "dbec_server_usage", "dbec_server_last_error",
]

CONVERGED, STEPS_DONE, NOT_CONVERGED, DIVERGED = range(4)


class Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("mem_budget", ctypes.c_double), ("max_batch", ctypes.c_int),
                ("check_every", ctypes.c_long), ("polish_tol", ctypes.c_double)]


class Request(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_long), ("mode", ctypes.c_int), ("prec", ctypes.c_int),
        ("nx", ctypes.c_long), ("ny", ctypes.c_long), ("nz", ctypes.c_long),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
        ("vgamma", ctypes.c_double), ("vnu", ctypes.c_double), ("vlambda", ctypes.c_double),
        ("cutoff", ctypes.c_double), ("par", ctypes.c_double),
        ("g", ctypes.c_double), ("gd", ctypes.c_double),
        ("max_steps", ctypes.c_long), ("mutol", ctypes.c_double),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_long), ("status", ctypes.c_int), ("steps", ctypes.c_long), ("steps_fp64", ctypes.c_long),
        ("norm", ctypes.c_double), ("mu", ctypes.c_double), ("en", ctypes.c_double),
        ("rms", ctypes.c_double * 4),
        ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
    ]

    def as_dict(self):
        return {"id": self.id, "status": self.status, "steps": self.steps, "steps_fp64": self.steps_fp64,
                "norm": self.norm, "mu": self.mu,
                "en": self.en, "rms": list(self.rms), "t_submit": self.t_submit, "t_start": self.t_start,
                "t_done": self.t_done}


def declare(lib):
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    sig = {
        "dbec_server_create": ([ctypes.POINTER(Config), ctypes.POINTER(P)], I),
        "dbec_server_destroy": ([P], I),
        "dbec_server_time": ([P], ctypes.c_double),
        "dbec_server_submit": ([P, ctypes.POINTER(Request)], I),
        "dbec_server_progress": ([P, I, ctypes.POINTER(Result), I, ctypes.POINTER(I)], I),
        "dbec_server_pending": ([P], ctypes.c_long),
        "dbec_server_usage": ([P, D, ctypes.POINTER(I)], I),
        "dbec_server_last_error": ([P], ctypes.c_char_p),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


def make_request(rid, n, mode, prec, g0, gd0, trap, dt, max_steps, mutol=1e-8, cutoff=None, box=16.0, par=2.0):
    """A cubic-grid request: N^3 points, dx = box / N, couplings g = par * G0."""
    h = box / n
    return Request(id=rid, mode=mode, prec=prec, nx=n, ny=n, nz=n, dx=h, dy=h, dz=h, dt=dt,
                   vgamma=trap[0], vnu=trap[1], vlambda=trap[2],
                   cutoff=box / 2 if cutoff is None else cutoff, par=par,
                   g=par * g0, gd=par * gd0, max_steps=max_steps, mutol=mutol)


# A mixed ground state moves to FP64 once |dmu/mu| <= this between checks (or
# once it stops falling below 1e-6, FP32's rounding floor).
DEFAULT_POLISH_TOL = 1e-6


class Server:
    def __init__(self, lib, device=0, mem_budget=0.0, max_batch=16, check_every=50, polish_tol=DEFAULT_POLISH_TOL):
        self.lib = declare(lib)
        self.h = ctypes.c_void_p()
        cfg = Config(device=device, mem_budget=mem_budget, max_batch=max_batch, check_every=check_every,
                     polish_tol=polish_tol)
        rc = self.lib.dbec_server_create(ctypes.byref(cfg), ctypes.byref(self.h))
        if rc != 0:
            raise dbecpy.DbecError(rc, "server creation failed")
        self._buf = (Result * 64)()

    def _check(self, rc):
        if rc != 0:
            raise dbecpy.DbecError(rc, self.lib.dbec_server_last_error(self.h).decode())

    def time(self):
        return self.lib.dbec_server_time(self.h)

    def submit(self, req):
        self._check(self.lib.dbec_server_submit(self.h, ctypes.byref(req)))

    def progress(self, wait=True):
        n = ctypes.c_int(0)
        self._check(self.lib.dbec_server_progress(self.h, 1 if wait else 0, self._buf, 64, ctypes.byref(n)))
        return [self._buf[i].as_dict() for i in range(n.value)]

    def pending(self):
        return self.lib.dbec_server_pending(self.h)

    def usage(self):
        b, k = ctypes.c_double(0), ctypes.c_int(0)
        self._check(self.lib.dbec_server_usage(self.h, ctypes.byref(b), ctypes.byref(k)))
        return b.value, k.value

    def run_all(self, reqs, timeout=600.0):
        """Submit every request now and serve them to completion."""
        for r in reqs:
            self.submit(r)
        out, t0 = [], time.time()
        while self.pending() > 0:
            out += self.progress(True)
            if time.time() - t0 > timeout:
                raise TimeoutError(f"{self.pending()} requests still pending")
        return {r["id"]: r for r in out}

    def close(self):
        if self.h:
            self.lib.dbec_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def load_stream(path, count=None, prec=dbecpy.MIXED):
    """The serving stream of the JSON fixture at `path`: (arrival time,
    Request) pairs in arrival order, the first `count` of them (all when
    None), in precision `prec`.  Request ids are the stream positions."""
    import json

    d = json.load(open(path))
    rows = d["requests"][: (count if count is not None else len(d["requests"]))]
    out = []
    for i, (t, n, g0, gd0, ga, nu, la, real) in enumerate(rows):
        n = int(n)
        if real:
            req = make_request(i, n, dbecpy.REAL, prec, g0, gd0, (ga, nu, la), d["dt_real"], d["real_steps"],
                               cutoff=d["cutoff"], box=d["box"], par=d["par"])
        else:
            req = make_request(i, n, dbecpy.IMAG, prec, g0, gd0, (ga, nu, la), d["dt_imag"], d["imag_cap"],
                               mutol=d["mutol"], cutoff=d["cutoff"], box=d["box"], par=d["par"])
        out.append((float(t), req))
    return out


def serve_stream(server, stream, timeout=600.0):
    """Replays arrivals in real time against the server; returns results with
    latency = completion - arrival (seconds)."""
    base = server.time()
    i, out, t_end = 0, [], time.time() + timeout
    arrival = {}
    while i < len(stream) or server.pending() > 0:
        now = server.time() - base
        while i < len(stream) and stream[i][0] <= now:
            arrival[stream[i][1].id] = base + stream[i][0]
            server.submit(stream[i][1])
            i += 1
        for r in server.progress(True):
            r["latency"] = r["t_done"] - arrival[r["id"]]
            out.append(r)
        if time.time() > t_end:
            raise TimeoutError("serving stream did not finish")
    return out


def plain_solve(lib, req, check_every=50):
    """One request as a plain FP64 solve through dbec.h: host Gaussian,
    dbec_step in chunks, calcmuen at every check (the back-to-back baseline)."""
    p = dbecpy.make_params(req.nx, req.ny, req.nz, req.dx, req.dy, req.dz, req.dt, req.vgamma, req.vnu,
                           req.vlambda, req.cutoff, par=req.par)
    with dbecpy.Solver(lib, p, req.mode, dbecpy.FP64) as s:
        s.set_psi(dbecpy.gaussian_psi0(p) if req.mode == dbecpy.IMAG
                  else dbecpy.gaussian_psi0(p).astype(np.complex128))
        if req.mode == dbecpy.IMAG:
            s.calcnorm(rescale=True)
        s.set_coupling(req.g, req.gd)
        steps, mu_prev, status = 0, None, None
        while status is None:
            k = min(check_every, req.max_steps - steps) if req.mode == dbecpy.REAL else check_every
            s.step(k, want_norm=False)
            steps += k
            if req.mode == dbecpy.IMAG:
                mu, en = s.calcmuen()
                if not np.isfinite(mu):
                    status = DIVERGED
                elif mu_prev is not None and abs(mu - mu_prev) <= req.mutol * abs(mu):
                    status = CONVERGED
                elif steps >= req.max_steps:
                    status = NOT_CONVERGED
                mu_prev = mu
            elif steps >= req.max_steps:
                mu, en = s.calcmuen()
                status = STEPS_DONE
        return {"id": req.id, "status": status, "steps": steps, "mu": mu, "en": en, "rms": s.calcrms()}
