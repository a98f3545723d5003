"""ctypes binding of the solver's C ABI (include/dbec.h).

The same binding drives any library that implements the ABI; the product is
libdbec_hip.so next to this package (HIP, gfx950).  Operators keep the names
of the reference's per-step functions (calcpsidd2, calcnu, calclux, calcluy,
calcluz, calcnorm, calcmuen, calcrms; src/imag3d-mpicuda/imag3d-mpicuda.cu
and src/real3d-mpicuda/real3d-mpicuda.cu).
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_LIB = os.path.join(PKG_DIR, "libdbec_hip.so")
BIN_DIR = os.path.join(PKG_DIR, "bin")

PI_REF = 3.14159265358979  # the literal of imag3d.cu:69

DBEC_OK, DBEC_EINVAL, DBEC_ENOMEM, DBEC_EDEVICE, DBEC_EUNSUPPORTED, DBEC_ESTATE = range(6)
IMAG, REAL = 0, 1
FP64, MIXED = 0, 1
DEN_X, DEN_Y, DEN_Z, DEN_XY, DEN_XZ, DEN_YZ, CUT_XY0, CUT_X0Z, CUT_0YZ, DEN_XYZ = range(10)
TABLE_POT, TABLE_POTDD, TABLE_CALPHA, TABLE_CGAMMA = range(4)

SYMBOLS = [
    "dbec_create", "dbec_destroy", "dbec_last_error", "dbec_set_psi", "dbec_get_psi",
    "dbec_set_coupling", "dbec_get_coupling", "dbec_calcpsidd2", "dbec_get_psidd2",
    "dbec_calcnu", "dbec_calclux", "dbec_calcluy", "dbec_calcluz", "dbec_calcnorm",
    "dbec_calcmuen", "dbec_calcmuen_phase", "dbec_calcrms", "dbec_psi_center", "dbec_step", "dbec_density_count",
    "dbec_density", "dbec_density_noalloc", "dbec_get_table", "dbec_synchronize", "dbec_profile_step", "dbec_tuning",
]


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_long), ("ny", ctypes.c_long), ("nz", ctypes.c_long),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("vgamma", ctypes.c_double), ("vnu", ctypes.c_double), ("vlambda", ctypes.c_double),
        ("cutoff", ctypes.c_double), ("par", ctypes.c_double), ("pi", ctypes.c_double),
        ("device", ctypes.c_int),
    ]


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("avg_ms", ctypes.c_double),
                ("alg_bytes", ctypes.c_double), ("launches", ctypes.c_double)]


class DbecError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dbec error {code}: {msg}")
        self.code = code


def load(path=HIP_LIB):
    """Load a library implementing dbec.h and declare the signatures."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run make in {PKG_DIR})")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    P, D, L, I = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_long, ctypes.c_int
    sig = {
        "dbec_create": ([ctypes.POINTER(Params), I, I, P, ctypes.POINTER(P)], I),
        "dbec_destroy": ([P], I),
        "dbec_last_error": ([P], ctypes.c_char_p),
        "dbec_set_psi": ([P, D], I),
        "dbec_get_psi": ([P, D], I),
        "dbec_set_coupling": ([P, ctypes.c_double, ctypes.c_double], I),
        "dbec_get_coupling": ([P, D, D], I),
        "dbec_calcpsidd2": ([P], I),
        "dbec_get_psidd2": ([P, D], I),
        "dbec_calcnu": ([P], I),
        "dbec_calclux": ([P], I),
        "dbec_calcluy": ([P], I),
        "dbec_calcluz": ([P], I),
        "dbec_calcnorm": ([P, D, I], I),
        "dbec_calcmuen": ([P, D, D], I),
        "dbec_calcmuen_phase": ([P, D, D], I),
        "dbec_calcrms": ([P, D], I),
        "dbec_psi_center": ([P, D], I),
        "dbec_step": ([P, L, ctypes.c_double, ctypes.c_double, D], I),
        "dbec_density_count": ([P, I, L, L, L], L),
        "dbec_density": ([P, I, L, L, L, D], I),
        "dbec_density_noalloc": ([P, I, L, L, L, D], I),
        "dbec_get_table": ([P, I, I, D], I),
        "dbec_synchronize": ([P], I),
        "dbec_profile_step": ([P, I, ctypes.POINTER(KernelStat), I, ctypes.POINTER(I)], I),
        "dbec_tuning": ([P, ctypes.c_char_p, L], I),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


def make_params(nx, ny, nz, dx, dy, dz, dt, gamma, nu, lam, cutoff, par=2.0, pi=PI_REF, device=-1):
    return Params(nx, ny, nz, dx, dy, dz, dt, gamma, nu, lam, cutoff, par, pi, device)


def grid(p):
    """x[i] = (i - N/2) h (initpsi, imag3d.cu:842-855)."""
    x = (np.arange(p.nx) - p.nx // 2) * p.dx
    y = (np.arange(p.ny) - p.ny // 2) * p.dy
    z = (np.arange(p.nz) - p.nz // 2) * p.dz
    return x, y, z


def gaussian_psi0(p, pi=PI_REF):
    """Anisotropic Gaussian of initpsi (imag3d.cu:835-867), [x][y][z]."""
    x, y, z = grid(p)
    cpsi = np.sqrt(pi * np.sqrt(pi / (p.vgamma * p.vnu * p.vlambda)))
    e = (p.vgamma * (x * x)[:, None, None] + p.vnu * (y * y)[None, :, None]
         + p.vlambda * (z * z)[None, None, :])
    return np.exp(-0.5 * e) / cpsi


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Solver:
    """One solve (context) of a library implementing dbec.h."""

    def __init__(self, lib, params, mode=IMAG, prec=FP64, stream=None):
        self.lib = lib
        self.p = params
        self.mode = mode
        self.prec = prec
        self.shape = (params.nx, params.ny, params.nz)
        h = ctypes.c_void_p()
        rc = lib.dbec_create(ctypes.byref(params), mode, prec, stream, ctypes.byref(h))
        if rc != DBEC_OK:
            raise DbecError(rc, lib.dbec_last_error(None).decode())
        self.h = h

    def _chk(self, rc):
        if rc != DBEC_OK:
            raise DbecError(rc, self.lib.dbec_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.dbec_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------
    def set_psi(self, psi):
        if self.mode == REAL:
            a = np.ascontiguousarray(np.asarray(psi, dtype=np.complex128)).view(np.float64)
        else:
            a = np.ascontiguousarray(np.asarray(psi, dtype=np.float64))
        assert a.size == np.prod(self.shape) * (2 if self.mode == REAL else 1)
        self._chk(self.lib.dbec_set_psi(self.h, _dp(a)))

    def get_psi(self):
        if self.mode == REAL:
            a = np.empty(self.shape, dtype=np.complex128)
            self._chk(self.lib.dbec_get_psi(self.h, _dp(a.view(np.float64))))
        else:
            a = np.empty(self.shape, dtype=np.float64)
            self._chk(self.lib.dbec_get_psi(self.h, _dp(a)))
        return a

    def set_coupling(self, g, gd):
        self._chk(self.lib.dbec_set_coupling(self.h, g, gd))

    def get_coupling(self):
        g, gd = ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.dbec_get_coupling(self.h, ctypes.byref(g), ctypes.byref(gd)))
        return g.value, gd.value

    # -- operators -----------------------------------------------------------
    def calcpsidd2(self):
        self._chk(self.lib.dbec_calcpsidd2(self.h))

    def get_psidd2(self):
        a = np.empty(self.shape, dtype=np.float64)
        self._chk(self.lib.dbec_get_psidd2(self.h, _dp(a)))
        return a

    def calcnu(self):
        self._chk(self.lib.dbec_calcnu(self.h))

    def calclux(self):
        self._chk(self.lib.dbec_calclux(self.h))

    def calcluy(self):
        self._chk(self.lib.dbec_calcluy(self.h))

    def calcluz(self):
        self._chk(self.lib.dbec_calcluz(self.h))

    def calcnorm(self, rescale=False):
        v = ctypes.c_double()
        self._chk(self.lib.dbec_calcnorm(self.h, ctypes.byref(v), 1 if rescale else 0))
        return v.value

    def calcmuen(self):
        mu, en = ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.dbec_calcmuen(self.h, ctypes.byref(mu), ctypes.byref(en)))
        return mu.value, en.value

    def calcmuen_phase(self):
        """mu and E with the phase gradient in the kinetic term (dbec_calcmuen_phase)."""
        mu, en = ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.dbec_calcmuen_phase(self.h, ctypes.byref(mu), ctypes.byref(en)))
        return mu.value, en.value

    def calcrms(self):
        r = np.zeros(4)
        self._chk(self.lib.dbec_calcrms(self.h, _dp(r)))
        return r

    def psi_center(self):
        r = np.zeros(2)
        self._chk(self.lib.dbec_psi_center(self.h, _dp(r)))
        return complex(r[0], r[1])

    def step(self, n, g_stp=0.0, gd_stp=0.0, want_norm=True):
        v = ctypes.c_double(0.0)
        self._chk(self.lib.dbec_step(self.h, n, g_stp, gd_stp, ctypes.byref(v) if want_norm else None))
        return v.value

    def synchronize(self):
        self._chk(self.lib.dbec_synchronize(self.h))

    def density(self, kind, ox=1, oy=1, oz=1):
        n = self.lib.dbec_density_count(self.h, kind, ox, oy, oz)
        if n < 0:
            raise DbecError(DBEC_EINVAL, "bad density request")
        a = np.empty(n, dtype=np.float64)
        self._chk(self.lib.dbec_density(self.h, kind, ox, oy, oz, _dp(a)))
        return a

    def density_noalloc(self, kind, ox=1, oy=1, oz=1):
        """dbec_density_noalloc: the same extract through the library's fixed
        device scratch (no device allocation)."""
        n = self.lib.dbec_density_count(self.h, kind, ox, oy, oz)
        if n < 0:
            raise DbecError(DBEC_EINVAL, "bad density request")
        a = np.empty(n, dtype=np.float64)
        self._chk(self.lib.dbec_density_noalloc(self.h, kind, ox, oy, oz, _dp(a)))
        return a

    def table(self, which, axis=0):
        nx, ny, nz = self.shape
        c = 2 if self.mode == REAL else 1
        n = {TABLE_POT: nx * ny * nz, TABLE_POTDD: nx * ny * (nz // 2 + 1)}.get(
            which, c * ((nx, ny, nz)[axis] - 1))
        a = np.empty(n, dtype=np.float64)
        self._chk(self.lib.dbec_get_table(self.h, which, axis, _dp(a)))
        return a

    def tuning(self):
        """The context's launch choices as text (dbec_tuning)."""
        buf = ctypes.create_string_buffer(512)
        self._chk(self.lib.dbec_tuning(self.h, buf, len(buf)))
        return buf.value.decode()

    def profile(self, reps=10):
        stats = (KernelStat * 64)()
        n = ctypes.c_int()
        self._chk(self.lib.dbec_profile_step(self.h, reps, stats, 64, ctypes.byref(n)))
        return [dict(name=s.name.decode(), avg_ms=s.avg_ms, alg_bytes=s.alg_bytes, launches=s.launches)
                for s in stats[: n.value]]
