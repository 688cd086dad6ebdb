"""ctypes wrapper of the oracle (TEST INFRASTRUCTURE, NOT PRODUCT CODE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker / timed CPU baseline.  Arrays are numpy float64
in the reference layout: A has shape (8, nz, ny, nx), padded fields
(nz+2ng, ny+2ng, nx+2ng), all C-contiguous (= Fortran order of the
reference's (x, y, z) indices).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libppsoracle.so")

OK, ERR_NAN_OR_TINY_DIAG, ERR_VOLUME_SOURCE, ERR_NOT_IN_KERNEL, ERR_BAD_SIZE, \
    ERR_DIVERGED, ERR_SOLID, ERR_UNSUPPORTED, ERR_NOMEM = range(9)


class OrcGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("ng", ctypes.c_int), ("bc", ctypes.c_int * 6),
                ("xlength", ctypes.c_double), ("ylength", ctypes.c_double),
                ("zlength", ctypes.c_double)]

    @classmethod
    def make(cls, nx, ny=None, nz=None, ng=2, bc=(0, 0, 0, 0, 0, 0), lengths=(1.0, 1.0, 1.0)):
        g = cls()
        g.nx = nx
        g.ny = nx if ny is None else ny
        g.nz = nx if nz is None else nz
        g.ng = ng
        for q in range(6):
            g.bc[q] = int(bc[q])
        g.xlength, g.ylength, g.zlength = (float(v) for v in lengths)
        return g

    @property
    def shape_p(self):
        return (self.nz + 2 * self.ng, self.ny + 2 * self.ng, self.nx + 2 * self.ng)

    @property
    def shape_A(self):
        return (8, self.nz, self.ny, self.nx)

    def interior(self, p):
        g = self.ng
        return p[..., g:g + self.nz, g:g + self.ny, g:g + self.nx]


_L = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def lib():
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing at {LIB_PATH}: run oracle/build_oracle.py")
    L = ctypes.CDLL(LIB_PATH)
    G = ctypes.POINTER(OrcGrid)
    IP = ctypes.POINTER(ctypes.c_int)
    DP = ctypes.POINTER(ctypes.c_double)
    sig = {
        "orc_coords": (None, [_I, _I, _D, _P, _P, _P, _P]),
        "orc_pressure_masks": (None, [G, _P, _P, _P]),
        "orc_setup_poisson": (_I, [G, _P, _P, _P, _P, _P, _P, _P, _D, _D, _P]),
        "orc_setup_poisson_bp": (_I, [G, _P, _P, _P, _P, _P, _P, _P, _D, _D, _P, _P]),
        "orc_check_poisson_setup": (_I, [G, _P, _P, _P, _P, _P]),
        "orc_residual_field": (None, [G, _P, _P, _P]),
        "orc_compute_residual": (_D, [G, _P, _P, _I]),
        "orc_calc_residuals": (None, [G, _P, _P, DP]),
        "orc_apply_bc": (None, [G, _P]),
        "orc_redblack_relax": (None, [G, _P, _P, _D]),
        "orc_fill_coefficients": (None, [IP, _P, _P]),
        "orc_coarse_from_fine": (None, [IP, _P, _P]),
        "orc_coarse_fine_interp": (None, [IP, _I, _I, _P, _P]),
        "orc_newsolver_mg": (_I, [G, _P, _P, _D, _D, _I, _I, _I, _I, IP, DP, _P]),
        "orc_newsolver_std": (_I, [G, _P, _P, _D, _D, _I, _I, IP, DP, _P]),
        "orc_free_solver": (_I, [G, _P, _P, _P, _P, _I, _D, _D, _I, IP, DP, _P]),
        "orc_setup_testmg": (None, [G, _P, _P]),
        "orc_get_mgtest_err": (_D, [G, _P]),
        "orc_project_velocity": (None, [G, _P, _P, _P, _P, _P, _P, _P, _P, _D]),
        "orc_pcg_solve": (_I, [G, _P, _P, _D, _I, DP]),
        "orc_linfunc": (_I, [G, _P, _D, _D, _I, _I, _P]),
        "orc_set_threads": (None, [_I]),
        "orc_max_threads": (_I, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _L = L
    return L


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous expected"
    return ctypes.c_void_p(a.ctypes.data)


def coords(N, ng, length):
    out = [np.empty(N + 2 * ng) for _ in range(4)]
    lib().orc_coords(N, ng, float(length), *[_p(a) for a in out])
    return tuple(out)


def pressure_masks(g):
    um, vm, wm = (np.empty(g.shape_p) for _ in range(3))
    lib().orc_pressure_masks(ctypes.byref(g), _p(um), _p(vm), _p(wm))
    return um, vm, wm


def setup_poisson(g, u, v, w, um, vm, wm, rho, dt=1.0, volume_source=0.0, bp=None):
    """SetupPoisson + Poisson_BCs + check_and_debug_Poisson; bp is the
    reference's BoundaryPressure(1:6) (None: zeros).  Returns (rc, A)."""
    A = np.empty(g.shape_A)
    bpa = np.zeros(6) if bp is None else np.ascontiguousarray(bp, dtype=np.float64)
    assert bpa.shape == (6,)
    rc = lib().orc_setup_poisson_bp(ctypes.byref(g), _p(u), _p(v), _p(w), _p(um), _p(vm), _p(wm),
                                    _p(rho), dt, volume_source, _p(bpa), _p(A))
    return rc, A


def check_poisson_setup(g, A, um, vm, wm):
    row = np.empty((g.nz, g.ny, g.nx))
    rc = lib().orc_check_poisson_setup(ctypes.byref(g), _p(A), _p(um), _p(vm), _p(wm), _p(row))
    return rc, row


def residual_field(g, A, p):
    r = np.empty((g.nz, g.ny, g.nx))
    lib().orc_residual_field(ctypes.byref(g), _p(A), _p(p), _p(r))
    return r


def compute_residual(g, A, p, norm=2):
    return lib().orc_compute_residual(ctypes.byref(g), _p(A), _p(p), norm)


def calc_residuals(g, A, p):
    out = (ctypes.c_double * 3)()
    lib().orc_calc_residuals(ctypes.byref(g), _p(A), _p(p), out)
    return tuple(out)


def apply_bc(g, p):
    lib().orc_apply_bc(ctypes.byref(g), _p(p))


def redblack_relax(g, A, p, beta):
    lib().orc_redblack_relax(ctypes.byref(g), _p(A), _p(p), beta)


def fill_coefficients(nf, Af):
    nc = [v // 2 for v in nf]
    Ac = np.zeros((8, nc[2], nc[1], nc[0]))
    arr = (ctypes.c_int * 3)(*nf)
    lib().orc_fill_coefficients(arr, _p(Af), _p(Ac))
    return Ac


def coarse_from_fine(nf, Af, Ac):
    arr = (ctypes.c_int * 3)(*nf)
    lib().orc_coarse_from_fine(arr, _p(Af), _p(Ac))
    return Ac


def coarse_fine_interp(nf, ng, sym, ef, ec):
    arr = (ctypes.c_int * 3)(*nf)
    lib().orc_coarse_fine_interp(arr, ng, sym, _p(ef), _p(ec))
    return ef


def newsolver_mg(g, A, p, maxError, beta=1.2, maxit=50, norm=2, nrelax=3, sym_mg=0):
    """NewSolverMG on p (in place).  Returns (rc, it, tres2, history)."""
    it, t2 = ctypes.c_int(), ctypes.c_double()
    hist = np.zeros(max(maxit, 1))
    rc = lib().orc_newsolver_mg(ctypes.byref(g), _p(A), _p(p), maxError, beta, maxit, norm,
                                nrelax, sym_mg, ctypes.byref(it), ctypes.byref(t2), _p(hist))
    return rc, it.value, t2.value, hist[:min(it.value, maxit)]


def newsolver_std(g, A, p, maxError, beta=1.2, maxit=50, norm=2):
    it, t2 = ctypes.c_int(), ctypes.c_double()
    hist = np.zeros(max(maxit, 1))
    rc = lib().orc_newsolver_std(ctypes.byref(g), _p(A), _p(p), maxError, beta, maxit, norm,
                                 ctypes.byref(it), ctypes.byref(t2), _p(hist))
    return rc, it.value, t2.value, hist[:min(it.value, maxit)]


def free_solver(g, A, p, pcmask, pgas, solver_flag, maxError, beta=1.2, maxit=50):
    """FreeSolver on p (in place); pcmask int32 and pgas float64, both padded
    like p.  Returns (rc, it, tres2, history)."""
    assert pcmask.dtype == np.int32 and pcmask.flags["C_CONTIGUOUS"]
    it, t2 = ctypes.c_int(), ctypes.c_double()
    hist = np.zeros(max(maxit, 1))
    rc = lib().orc_free_solver(ctypes.byref(g), _p(A), _p(p), ctypes.c_void_p(pcmask.ctypes.data),
                               _p(pgas), solver_flag, maxError, beta, maxit, ctypes.byref(it),
                               ctypes.byref(t2), _p(hist))
    return rc, it.value, t2.value, hist[:min(it.value, maxit)]


def setup_testmg(g, rho):
    A = np.empty(g.shape_A)
    lib().orc_setup_testmg(ctypes.byref(g), _p(rho), _p(A))
    return A


def get_mgtest_err(g, p):
    return lib().orc_get_mgtest_err(ctypes.byref(g), _p(p))


def project_velocity(g, u, v, w, um, vm, wm, rho, p, dt=1.0):
    lib().orc_project_velocity(ctypes.byref(g), _p(u), _p(v), _p(w), _p(um), _p(vm), _p(wm),
                               _p(rho), _p(p), dt)


def linfunc(g, cvof, a1, a2, mean_flag=101, nfilter=0):
    """vofmodules.f90:171-218 then do_all_ghost; returns the padded field."""
    out = np.empty(g.shape_p)
    rc = lib().orc_linfunc(ctypes.byref(g), _p(cvof), a1, a2, mean_flag, nfilter, _p(out))
    if rc != 0:
        raise ValueError("linfunc: mean_flag must be 101 or 102 and nfilter 0 or 1")
    return out


def set_threads(n):
    """Threads of the NewSolverMG restatement (1 = serial)."""
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def pcg_solve(g, A, p, rtol=1e-14, maxit=100000):
    rel = ctypes.c_double()
    its = lib().orc_pcg_solve(ctypes.byref(g), _p(A), _p(p), rtol, maxit, ctypes.byref(rel))
    return its, rel.value
