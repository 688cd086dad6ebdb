"""paris-git_amd -- MI355X pressure-Poisson solver for PARIS.

Python mirror of the reference's linear-solver interface over the C ABI of
``libpps.so`` (include/pps.h):

* :meth:`Solver.new_solver` / :meth:`Solver.new_solver_mg` follow
  ``NewSolver(A,p,maxError,beta,maxit,it,ierr,norm)`` (newsolver.f90:34-57)
  and ``NewSolverMG(...,tres2)`` (MGsolver.f90:346-405) on host arrays;
* :meth:`Solver.solve` is the device-resident form on torch tensors;
* the producer / consumer of the system (SetupPoisson, SetPressureBC,
  check_poisson_setup, calcResiduals, project_velocity) are free functions;
* :class:`PressureSolve` is a ``torch.autograd.Function`` whose backward is
  the adjoint solve plus the fused face kernel.

Array layout is the reference's Fortran order: ``A`` holds 8 planes of
``nx*ny*nz`` doubles (torch shape ``(8, nz, ny, nx)``), ``p`` carries ``ng``
ghost layers (torch shape ``(nz+2ng, ny+2ng, nx+2ng)``).

Nothing here computes on the CPU: every operation calls libpps, and a
missing library raises instead of falling back.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpps.so")

# enum pps_status
OK = 0
ERR_NAN_OR_TINY_DIAG = 1
ERR_VOLUME_SOURCE = 2
ERR_NOT_IN_KERNEL = 3
ERR_BAD_SIZE = 4
ERR_DIVERGED = 5
ERR_NOT_CONVERGED = 6
ERR_NONSYMMETRIC = 7
ERR_BAD_ARG = 8
ERR_HIP = 9
ERR_NO_DEVICE = 10
ERR_UNSUPPORTED = 11
ERR_SCRATCH = 12

# enum pps_method / pps_outer
MG_REFERENCE = 0
MG_FP64 = 1
MG_MIXED = 2
SOR = 3
OUTER_CG = 0
OUTER_REFINE = 1
SMOOTH_FUSED = 0
SMOOTH_COLOUR_PASSES = 1
TRANSFER_CONSTANT = 0
TRANSFER_TRILINEAR = 1
TRANSFERS = {"constant": TRANSFER_CONSTANT, "trilinear": TRANSFER_TRILINEAR}

METHODS = {"reference": MG_REFERENCE, "fp64": MG_FP64, "mixed": MG_MIXED, "sor": SOR}

# boundary codes (modules.f90:870-877)
WALL, PERIODIC, SHEAR, VEL_IN, VEL_OUT = 0, 1, 2, 3, 4


class PPSError(RuntimeError):
    """A libpps call returned a non-zero status."""

    def __init__(self, code, what=""):
        self.code = code
        msg = _lib().pps_status_string(code).decode() if _LIBH is not None else str(code)
        super().__init__(f"{what}: {msg} (status {code})" if what else f"{msg} (status {code})")


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("ng", ctypes.c_int), ("bdry_cond", ctypes.c_int * 6),
                ("xlength", ctypes.c_double), ("ylength", ctypes.c_double),
                ("zlength", ctypes.c_double)]

    @classmethod
    def make(cls, nx, ny=None, nz=None, ng=2, bdry_cond=(0, 0, 0, 0, 0, 0),
             lengths=(1.0, 1.0, 1.0)):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        g = cls()
        g.nx, g.ny, g.nz, g.ng = nx, ny, nz, ng
        for q in range(6):
            g.bdry_cond[q] = int(bdry_cond[q])
        g.xlength, g.ylength, g.zlength = (float(v) for v in lengths)
        return g

    @property
    def shape_A(self):
        return (8, self.nz, self.ny, self.nx)

    @property
    def shape_p(self):
        return (self.nz + 2 * self.ng, self.ny + 2 * self.ng, self.nx + 2 * self.ng)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    def interior(self, p):
        """View of the interior cells of a padded array (numpy or torch)."""
        g = self.ng
        return p[g:g + self.nz, g:g + self.ny, g:g + self.nx]

    def periodic(self):
        return tuple(int(self.bdry_cond[d] == 1) for d in range(3))


class Options(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("outer", ctypes.c_int), ("nrelax", ctypes.c_int),
                ("sym_mg", ctypes.c_int), ("nu_pre", ctypes.c_int), ("nu_post", ctypes.c_int),
                ("check_setup", ctypes.c_int), ("multigrid", ctypes.c_int),
                ("smoother", ctypes.c_int), ("sweep_rows", ctypes.c_int),
                ("sweep_planes", ctypes.c_int),
                ("tail_cells", ctypes.c_int), ("transfer", ctypes.c_int)]


_LIBH = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)
_GP = ctypes.POINTER(Grid)


def _lib():
    """Load libpps.so; raise if it has not been built."""
    global _LIBH
    if _LIBH is not None:
        return _LIBH
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libpps.so is missing at {LIB_PATH}: build it first "
                           "(python -c 'import __graft_entry__ as g; g.build()')")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "pps_default_options": (_I, [ctypes.POINTER(Options)]),
        "pps_init": (_I, [_GP, ctypes.POINTER(Options), ctypes.POINTER(_P)]),
        "pps_finalize": (_I, [_P]),
        "pps_set_options": (_I, [_P, ctypes.POINTER(Options)]),
        "pps_set_cell_widths": (_I, [_P, _P, _P, _P]),
        "pps_new_solver": (_I, [_P, _P, _P, _D, _D, _I, _IP, _IP, _I]),
        "pps_new_solver_mg": (_I, [_P, _P, _P, _D, _D, _I, _IP, _IP, _I, _DP]),
        "pps_solve_device": (_I, [_P, _P, _P, _D, _D, _I, _I, _IP, _DP, _P]),
        "pps_solve_device_noalloc": (_I, [_P, _P, _P, _D, _D, _I, _I, _IP, _DP, _P]),
        "pps_solve_fields_device": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _D, _D, _P, _D, _I, _IP,
                                         _DP, _P, _P, _P, _P]),
        "pps_solve_project_device": (_I, [_P, _P, _P, _D, _D, _I, _I, _IP, _DP, _P, _P, _P, _P, _P,
                                          _P, _P, _D, _P, _P, _P, _P]),
        "pps_last_history": (_I, [_P, _DP, _I]),
        "pps_last_timing": (_I, [_P, _DP, _DP]),
        "pps_adjoint_device": (_I, [_P, _P, _P, _P, _D, _I, _IP, _P]),
        "pps_solve_faces_device": (_I, [_P, _P, _P, _P, _P, _P, _P, _D, _I, _IP, _DP, _P]),
        "pps_adjoint_faces_device": (_I, [_P, _P, _P, _P, _P, _P, _P, _D, _I, _IP, _P]),
        "pps_face_gradient_device": (_I, [_GP, _P, _P, _P, _P, _P, _P, _P]),
        "pps_pressure_masks_device": (_I, [_GP, _P, _P, _P, _P]),
        "pps_linfunc_device": (_I, [_GP, _P, _D, _D, _I, _I, _P, _P]),
        "pps_setup_poisson_device": (_I, [_GP, _P, _P, _P, _P, _P, _P, _P, _D, _D, _P, _P]),
        "pps_setup_poisson_bp_device": (_I, [_GP, _P, _P, _P, _P, _P, _P, _P, _D, _D, _P, _P, _P]),
        "pps_setup_poisson_coords_device": (_I, [_GP, _P, _P, _P, _P, _P, _P, _P, _D, _D, _P, _P, _P,
                                                 _P, _P, _P, _P, _P]),
        "pps_check_setup_device": (_I, [_GP, _P, _P]),
        "pps_check_poisson_setup_device": (_I, [_GP, _P, _P, _P, _P, _P]),
        "pps_remove_rhs_mean_device": (_I, [_GP, _P, _P]),
        "pps_calc_residuals_device": (_I, [_GP, _P, _P, _DP, _P]),
        "pps_project_velocity_device": (_I, [_GP, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P]),
        "pps_project_velocity_coords_device": (_I, [_GP, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P,
                                                    _P, _P, _P, _P]),
        "pps_synth_fields_device": (_I, [_GP, _P, _I, _D, _D, _D, ctypes.c_ulonglong,
                                         _P, _P, _P, _P, _P]),
        "pps_synth_drops": (_I, [_GP, _I, _D, _D, ctypes.c_ulonglong, _P]),
        "pps_coords": (_I, [_I, _I, _D, _P, _P, _P, _P]),
        "pps_status_string": (ctypes.c_char_p, [_I]),
        "pps_faces_to_A_device": (_I, [_GP, _P, _P, _P, _P, _P, _P, _P]),
        "pps_profile_enable": (_I, [_P, _I]),
        "pps_profile_read": (_I, [_P, _DP, ctypes.POINTER(ctypes.c_longlong), _DP]),
        "pps_free_solver": (_I, [_P, _P, _P, _P, _P, _I, _D, _D, _I, _IP, _IP, _DP]),
        "pps_free_solver_device": (_I, [_P, _P, _P, _P, _P, _I, _D, _D, _I, _IP, _DP, _P]),
        "pps_poi_solve": (_I, [_P, _P, _P, _D, _I, _IP, _I]),
        "pps_poi_solve_device": (_I, [_P, _P, _P, _D, _I, _IP, _I, _P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIBH = L
    return L


def exported_symbols():
    """Names of the C-ABI functions declared in include/pps.h."""
    return [
        "pps_default_options", "pps_init", "pps_finalize", "pps_set_options", "pps_set_cell_widths",
        "pps_new_solver", "pps_new_solver_mg", "pps_solve_device", "pps_solve_device_noalloc",
        "pps_solve_project_device", "pps_solve_fields_device",
        "pps_last_history",
        "pps_last_timing", "pps_adjoint_device", "pps_face_gradient_device",
        "pps_solve_faces_device", "pps_adjoint_faces_device",
        "pps_pressure_masks_device", "pps_linfunc_device", "pps_setup_poisson_device",
        "pps_setup_poisson_bp_device", "pps_setup_poisson_coords_device",
        "pps_check_setup_device", "pps_check_poisson_setup_device",
        "pps_remove_rhs_mean_device", "pps_calc_residuals_device",
        "pps_project_velocity_device", "pps_project_velocity_coords_device",
        "pps_synth_fields_device", "pps_synth_drops",
        "pps_coords", "pps_status_string", "pps_faces_to_A_device",
        "pps_profile_enable", "pps_profile_read", "pps_free_solver", "pps_free_solver_device",
        "pps_poi_solve", "pps_poi_solve_device",
    ]


def _check(rc, what, allow=(OK,)):
    if rc not in allow:
        raise PPSError(rc, what)
    return rc


def _ptr(t, n=None, what="array"):
    """Device (or host) address of a contiguous float64 torch tensor / numpy
    array; with n given, the array must hold exactly n elements (the
    kernels index the extents the grid implies, so a wrong size would read
    or write past the allocation)."""
    if isinstance(t, np.ndarray):
        if not (t.flags["C_CONTIGUOUS"] or t.flags["F_CONTIGUOUS"]):
            raise ValueError(f"{what} must be contiguous")
        if t.dtype != np.float64:
            raise TypeError(f"{what} must be float64, got {t.dtype}")
        if n is not None and t.size != n:
            raise ValueError(f"{what} must hold {n} elements, got {t.size}")
        return ctypes.c_void_p(t.ctypes.data)
    import torch
    if t.dtype != torch.float64:
        raise TypeError(f"{what} must be a float64 tensor, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    if n is not None and t.numel() != n:
        raise ValueError(f"{what} must hold {n} elements, got {t.numel()}")
    return ctypes.c_void_p(t.data_ptr())


def _iptr(t, n=None, what="mask"):
    """Address of a contiguous int32 numpy array / torch tensor (masks) of
    n elements when n is given."""
    if isinstance(t, np.ndarray):
        if t.dtype != np.int32 or not t.flags["C_CONTIGUOUS"]:
            raise TypeError(f"{what} must be a C-contiguous int32 array")
        size = t.size
    else:
        import torch
        if t.dtype != torch.int32 or not t.is_contiguous():
            raise TypeError(f"{what} must be a contiguous int32 tensor")
        size = t.numel()
    if n is not None and size != n:
        raise ValueError(f"{what} must hold {n} elements, got {size}")
    return ctypes.c_void_p(t.data_ptr() if not isinstance(t, np.ndarray) else t.ctypes.data)


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _need_gpu(*ts):
    import torch
    for t in ts:
        if isinstance(t, np.ndarray) or not t.is_cuda or t.dtype != torch.float64:
            raise TypeError("expected a CUDA (HIP) float64 tensor")


def _nA(grid):
    return 8 * grid.cells


def _np(grid):
    z, y, x = grid.shape_p
    return z * y * x


def _padded_ptrs(grid, **arrays):
    """Pointers of padded float64 arrays (shape of p), in keyword order."""
    n = _np(grid)
    return [_ptr(a, n, name) for name, a in arrays.items()]


def _spacing_arrays(grid, dx, dxh, dy, dyh, dz, dzh):
    """The reference's spacing arrays with ghosts (n + 2 ng values per
    axis) as contiguous float64 numpy arrays, sizes checked."""
    out = []
    for a, m, name in ((dx, grid.nx, "dx"), (dxh, grid.nx, "dxh"), (dy, grid.ny, "dy"),
                       (dyh, grid.ny, "dyh"), (dz, grid.nz, "dz"), (dzh, grid.nz, "dzh")):
        arr = np.ascontiguousarray(a, dtype=np.float64)
        _ptr(arr, m + 2 * grid.ng, name)
        out.append(arr)
    return out


def default_options():
    o = Options()
    _check(_lib().pps_default_options(ctypes.byref(o)), "pps_default_options")
    return o


class Solver:
    """A solver handle (init_MG / finalize_MG, MGsolver.f90:34-82)."""

    def __init__(self, grid, method="fp64", outer=OUTER_CG, nrelax=3, sym_mg=0, nu_pre=2,
                 nu_post=2, check_setup=0, multigrid=1, smoother=SMOOTH_FUSED, sweep_rows=0,
                 sweep_planes=0, tail_cells=0, transfer=TRANSFER_CONSTANT):
        self.grid = grid
        o = default_options()
        o.method = METHODS[method] if isinstance(method, str) else int(method)
        o.outer, o.nrelax, o.sym_mg = int(outer), int(nrelax), int(sym_mg)
        o.nu_pre, o.nu_post = int(nu_pre), int(nu_post)
        o.check_setup, o.multigrid = int(check_setup), int(multigrid)
        o.smoother = int(smoother)
        o.sweep_rows, o.sweep_planes = int(sweep_rows), int(sweep_planes)
        o.tail_cells = int(tail_cells)
        o.transfer = TRANSFERS[transfer] if isinstance(transfer, str) else int(transfer)
        self.options = o
        self.has_cell_widths = False
        h = ctypes.c_void_p()
        _check(_lib().pps_init(ctypes.byref(grid), ctypes.byref(o), ctypes.byref(h)), "pps_init")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib().pps_finalize(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_method(self, method):
        self.options.method = METHODS[method] if isinstance(method, str) else int(method)
        _check(_lib().pps_set_options(self._h, ctypes.byref(self.options)), "pps_set_options")

    def set_cell_widths(self, dx=None, dy=None, dz=None):
        """Cell widths of a stretched grid (numpy, nx / ny / nz values), or
        None to clear: the fast methods then weight each row by its cell
        volume, which symmetrises SetupPoisson's system."""
        if dx is None and dy is None and dz is None:
            _check(_lib().pps_set_cell_widths(self._h, None, None, None), "pps_set_cell_widths")
            self.has_cell_widths = False
            return
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (dx, dy, dz)]
        g = self.grid
        ptrs = [_ptr(a, m, nm) for a, m, nm in zip(arrs, (g.nx, g.ny, g.nz), ("dx", "dy", "dz"))]
        _check(_lib().pps_set_cell_widths(self._h, *ptrs), "pps_set_cell_widths")
        self.has_cell_widths = True

    # -- reference signatures on host arrays ---------------------------
    def new_solver(self, A, p, maxError, beta=1.2, maxit=50, norm=2):
        """NewSolver: p (host, float64, padded) is updated in place.
        Returns (it, ierr); it == maxit+1 when not converged."""
        it, ierr = ctypes.c_int(), ctypes.c_int()
        g = self.grid
        rc = _lib().pps_new_solver(self._h, _ptr(A, _nA(g), "A"), _ptr(p, _np(g), "p"), maxError,
                                   beta, maxit,
                                   ctypes.byref(it), ctypes.byref(ierr), norm)
        _check(rc, "NewSolver", allow=(OK, ERR_NOT_CONVERGED))
        return it.value, ierr.value

    def new_solver_mg(self, A, p, maxError, beta=1.2, maxit=50, norm=2):
        """NewSolverMG: returns (it, ierr, tres2)."""
        it, ierr, t2 = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        g = self.grid
        rc = _lib().pps_new_solver_mg(self._h, _ptr(A, _nA(g), "A"), _ptr(p, _np(g), "p"), maxError,
                                      beta, maxit,
                                      ctypes.byref(it), ctypes.byref(ierr), norm,
                                      ctypes.byref(t2))
        _check(rc, "NewSolverMG", allow=(OK, ERR_NOT_CONVERGED))
        return it.value, ierr.value, t2.value

    def free_solver(self, A, p, pcmask, P_gas, solver_flag, maxError, beta=1.2, maxit=50):
        """FreeSolver (freesurface.f90:700-808) on host arrays: p (float64)
        is updated in place; pcmask (int32) and P_gas (float64, or None for
        solver_flag 2) are padded like p.  Returns (it, ierr, tres2)."""
        it, ierr, t2 = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        g = self.grid
        rc = _lib().pps_free_solver(self._h, _ptr(A, _nA(g), "A"), _ptr(p, _np(g), "p"),
                                    _iptr(pcmask, _np(g), "pcmask"),
                                    _ptr(P_gas, _np(g), "P_gas") if P_gas is not None else None,
                                    solver_flag,
                                    maxError, beta, maxit, ctypes.byref(it), ctypes.byref(ierr),
                                    ctypes.byref(t2))
        _check(rc, "FreeSolver", allow=(OK, ERR_NOT_CONVERGED))
        return it.value, ierr.value, t2.value

    def free_solve(self, A, p, pcmask, P_gas, solver_flag, maxError, beta=1.2, maxit=50,
                   stream=None):
        """FreeSolver in place on device tensors (pcmask int32, P_gas float64
        or None for solver_flag 2, both padded like p); returns (it, tres2)."""
        _need_gpu(A, p)
        if not pcmask.is_cuda:
            raise TypeError("pcmask must be a CUDA (HIP) int32 tensor")
        it, t2 = ctypes.c_int(), ctypes.c_double()
        g = self.grid
        rc = _lib().pps_free_solver_device(self._h, _ptr(A, _nA(g), "A"), _ptr(p, _np(g), "p"),
                                           _iptr(pcmask, _np(g), "pcmask"),
                                           _ptr(P_gas, _np(g), "P_gas") if P_gas is not None else None,
                                           solver_flag, maxError, beta, maxit, ctypes.byref(it),
                                           ctypes.byref(t2), _stream(stream))
        _check(rc, "pps_free_solver_device", allow=(OK, ERR_NOT_CONVERGED))
        return it.value, t2.value

    def poi_solve(self, A, p, maxError, maxit=100, solver_type=2):
        """poi_solve (modules.f90:2411) on host arrays: relative tolerance
        ||r||/||A8|| < maxError; returns num_iterations."""
        n = ctypes.c_int()
        g = self.grid
        rc = _lib().pps_poi_solve(self._h, _ptr(A, _nA(g), "A"), _ptr(p, _np(g), "p"), maxError,
                                  maxit, ctypes.byref(n),
                                  solver_type)
        _check(rc, "poi_solve", allow=(OK, ERR_NOT_CONVERGED))
        return n.value

    def poi_solve_device(self, A, p, maxError, maxit=100, solver_type=2, stream=None):
        """poi_solve on device tensors; returns num_iterations."""
        _need_gpu(A, p)
        n = ctypes.c_int()
        g = self.grid
        rc = _lib().pps_poi_solve_device(self._h, _ptr(A, _nA(g), "A"), _ptr(p, _np(g), "p"),
                                         maxError, maxit,
                                         ctypes.byref(n), solver_type, _stream(stream))
        _check(rc, "pps_poi_solve_device", allow=(OK, ERR_NOT_CONVERGED))
        return n.value

    # -- device-resident form ---------------------------------------------
    def solve(self, A, p, maxError, beta=1.2, maxit=50, norm=2, stream=None, allow_maxit=True):
        """Solve in place on device tensors; returns (it, tres2)."""
        _need_gpu(A, p)
        it, t2 = ctypes.c_int(), ctypes.c_double()
        g = self.grid
        rc = _lib().pps_solve_device(self._h, _ptr(A, _nA(g), "A"), _ptr(p, _np(g), "p"), maxError,
                                     beta, maxit, norm,
                                     ctypes.byref(it), ctypes.byref(t2), _stream(stream))
        _check(rc, "pps_solve_device",
               allow=(OK, ERR_NOT_CONVERGED) if allow_maxit else (OK,))
        return it.value, t2.value

    def solve_noalloc(self, A, p, maxError, beta=1.2, maxit=50, norm=2, stream=None,
                      allow_maxit=True):
        """solve() through pps_solve_device_noalloc: no device allocation
        (fast methods, norm 2, grids up to about 1.1 million cells: 104^3 mixed, 108^3 FP64);
        returns (it, tres2)."""
        _need_gpu(A, p)
        it, t2 = ctypes.c_int(), ctypes.c_double()
        g = self.grid
        rc = _lib().pps_solve_device_noalloc(self._h, _ptr(A, _nA(g), "A"), _ptr(p, _np(g), "p"),
                                             maxError, beta, maxit, norm,
                                             ctypes.byref(it), ctypes.byref(t2), _stream(stream))
        _check(rc, "pps_solve_device_noalloc",
               allow=(OK, ERR_NOT_CONVERGED) if allow_maxit else (OK,))
        return it.value, t2.value

    def solve_fields(self, fields, p, maxError, dt=1.0, volume_source=0.0, maxit=50,
                     project=False, stream=None, allow_maxit=True):
        """Assemble the system from the fields inside the solve (A is never
        formed) and solve in place on p; fields = dict(u, v, w, um, vm, wm,
        rho) of padded device tensors.  With project=True also returns the
        corrected velocities.  Returns (it, tres2[, u_out, v_out, w_out])."""
        import torch
        f = fields
        _need_gpu(p, f["u"], f["v"], f["w"], f["um"], f["vm"], f["wm"], f["rho"])
        g = self.grid
        npad = _np(g)
        outs = [torch.empty_like(f[c]) for c in ("u", "v", "w")] if project else [None] * 3
        it, t2 = ctypes.c_int(), ctypes.c_double()
        rc = _lib().pps_solve_fields_device(
            self._h, *[_ptr(f[c], npad, c) for c in ("u", "v", "w", "um", "vm", "wm", "rho")],
            dt, volume_source, _ptr(p, npad, "p"), maxError, maxit, ctypes.byref(it), ctypes.byref(t2),
            *[_ptr(o) if o is not None else None for o in outs], _stream(stream))
        _check(rc, "pps_solve_fields_device",
               allow=(OK, ERR_NOT_CONVERGED) if allow_maxit else (OK,))
        if project:
            return (it.value, t2.value) + tuple(outs)
        return it.value, t2.value

    def solve_project(self, A, p, u, v, w, umask, vmask, wmask, rho, dt, maxError, beta=1.2,
                      maxit=50, norm=2, stream=None, allow_maxit=True):
        """Solve in place on p, then project_velocity (paris.f90:898-908) into
        new tensors; returns (it, tres2, u_out, v_out, w_out).  The fast
        methods apply the correction inside their last residual pass."""
        import torch
        _need_gpu(A, p, u, v, w, umask, vmask, wmask, rho)
        uo, vo, wo = torch.empty_like(u), torch.empty_like(v), torch.empty_like(w)
        it, t2 = ctypes.c_int(), ctypes.c_double()
        g = self.grid
        npad = _np(g)
        rc = _lib().pps_solve_project_device(
            self._h, _ptr(A, _nA(g), "A"), _ptr(p, npad, "p"), maxError, beta, maxit, norm,
            ctypes.byref(it), ctypes.byref(t2), *[_ptr(a, npad, nm) for a, nm in (
                (u, "u"), (v, "v"), (w, "w"), (umask, "umask"), (vmask, "vmask"), (wmask, "wmask"),
                (rho, "rho"))],
            dt, _ptr(uo), _ptr(vo), _ptr(wo), _stream(stream))
        _check(rc, "pps_solve_project_device",
               allow=(OK, ERR_NOT_CONVERGED) if allow_maxit else (OK,))
        return it.value, t2.value, uo, vo, wo

    def adjoint(self, A, g, lam, rtol=1e-12, maxit=100, stream=None):
        """Solve M^T lam = g (g, lam padded device tensors; M^T = M unless cell
        widths are set); returns iterations."""
        _need_gpu(A, g, lam)
        it = ctypes.c_int()
        gr = self.grid
        rc = _lib().pps_adjoint_device(self._h, _ptr(A, _nA(gr), "A"), _ptr(g, _np(gr), "g"),
                                       _ptr(lam, _np(gr), "lam"), rtol, maxit,
                                       ctypes.byref(it), _stream(stream))
        _check(rc, "pps_adjoint_device")
        return it.value

    def solve_faces(self, cface, weights, b, p, maxError, maxit=50, stream=None, allow_maxit=True):
        """faces_to_A(grid, cface, weights, b) then solve(), without forming
        A: the rows come from the face values inside the setup pass (fast
        methods).  Returns (it, tres2)."""
        _need_gpu(cface, b, p)
        it, t2 = ctypes.c_int(), ctypes.c_double()
        g = self.grid
        rc = _lib().pps_solve_faces_device(self._h, _ptr(cface, 3 * g.cells, "cface"),
                                           *_face_weight_ptrs(g, weights), _ptr(b, g.cells, "b"),
                                           _ptr(p, _np(g), "p"), maxError, maxit, ctypes.byref(it),
                                           ctypes.byref(t2), _stream(stream))
        _check(rc, "pps_solve_faces_device",
               allow=(OK, ERR_NOT_CONVERGED) if allow_maxit else (OK,))
        return it.value, t2.value

    def adjoint_faces(self, cface, weights, g, lam, rtol=1e-12, maxit=100, stream=None):
        """adjoint() for the system of face values cface: reuses the
        hierarchy of the last solve_faces on the same cface tensor."""
        _need_gpu(cface, g, lam)
        it = ctypes.c_int()
        gr = self.grid
        rc = _lib().pps_adjoint_faces_device(self._h, _ptr(cface, 3 * gr.cells, "cface"),
                                             *_face_weight_ptrs(gr, weights), _ptr(g, _np(gr), "g"),
                                             _ptr(lam, _np(gr), "lam"), rtol, maxit, ctypes.byref(it),
                                             _stream(stream))
        _check(rc, "pps_adjoint_faces_device")
        return it.value

    def history(self, cap=100000):
        buf = (ctypes.c_double * cap)()
        n = _lib().pps_last_history(self._h, buf, cap)
        return list(buf[:max(n, 0)])

    def timing(self):
        """(solve_ms, setup_ms) of the last solve, HIP events."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _lib().pps_last_timing(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def profile(self, on=True, every=1):
        """Time finest-level smoother launches with HIP events (the kernel's
        own start and end); every > 1 times one launch in `every`."""
        _check(_lib().pps_profile_enable(self._h, (max(int(every), 1) if on else 0)),
               "pps_profile_enable")

    def profile_read(self):
        """(kernel_ms, launches, algorithmic_bytes) since the last read."""
        ms, n, b = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
        _check(_lib().pps_profile_read(self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)),
               "pps_profile_read")
        return ms.value, n.value, b.value


# ---------------------------------------------------------------------------
# producer / consumer of the system, synthetic inputs
# ---------------------------------------------------------------------------
def coords(N, ng, length):
    """x, xh, dx, dxh as paris.f90:2959-3102 builds them (numpy)."""
    arrs = [np.empty(N + 2 * ng) for _ in range(4)]
    _check(_lib().pps_coords(N, ng, float(length), *[_ptr(a) for a in arrs]), "pps_coords")
    return tuple(arrs)


def synth_drops(grid, ndrop, rmin, rmax, seed):
    out = np.zeros((max(ndrop, 1), 4))
    _check(_lib().pps_synth_drops(ctypes.byref(grid), ndrop, rmin, rmax, seed, _ptr(out)),
           "pps_synth_drops")
    return out[:ndrop]


def synth_fields(grid, drops, rho1, rho2, width=1.5, seed=1, device="cuda"):
    """rho, u, v, w (padded device tensors) of a droplet field."""
    import torch
    drops = np.ascontiguousarray(np.asarray(drops, dtype=np.float64).reshape(-1, 4))
    rho, u, v, w = (torch.empty(grid.shape_p, dtype=torch.float64, device=device) for _ in range(4))
    dp = _ptr(drops) if len(drops) else ctypes.c_void_p(0)
    _check(_lib().pps_synth_fields_device(ctypes.byref(grid), dp, len(drops), rho1, rho2, width,
                                          seed, _ptr(rho), _ptr(u), _ptr(v), _ptr(w), _stream()),
           "pps_synth_fields_device")
    return rho, u, v, w


ARITH_MEAN, HARM_MEAN = 101, 102


def linfunc(grid, cvof, a1, a2, mean_flag=ARITH_MEAN, nfilter=0):
    """linfunc + do_all_ghost (vofmodules.f90:171-218, modules.f90:2062-2072):
    the padded property field (a1 where cvof = 0, a2 where 1) from the
    padded volume fraction; a new device tensor."""
    import torch
    _need_gpu(cvof)
    out = torch.empty(grid.shape_p, dtype=torch.float64, device=cvof.device)
    _check(_lib().pps_linfunc_device(ctypes.byref(grid), _ptr(cvof, _np(grid), "cvof"), float(a1),
                                     float(a2),
                                     int(mean_flag), int(nfilter), _ptr(out), _stream()),
           "pps_linfunc_device")
    return out


def pressure_masks(grid, device="cuda"):
    import torch
    um, vm, wm = (torch.empty(grid.shape_p, dtype=torch.float64, device=device) for _ in range(3))
    _check(_lib().pps_pressure_masks_device(ctypes.byref(grid), _ptr(um), _ptr(vm), _ptr(wm),
                                            _stream()), "pps_pressure_masks_device")
    return um, vm, wm


def setup_poisson(grid, u, v, w, um, vm, wm, rho, dt=1.0, volume_source=0.0,
                  boundary_pressure=None):
    """SetupPoisson + Poisson_BCs: the 8-slot system (device tensor).
    boundary_pressure: the reference's BoundaryPressure(1:6) (-x, +x, -y,
    +y, -z, +z) read by the pressure codes 5 / 6; None means zeros."""
    import torch
    _need_gpu(u, v, w, um, vm, wm, rho)
    A = torch.empty(grid.shape_A, dtype=torch.float64, device=rho.device)
    fields = _padded_ptrs(grid, u=u, v=v, w=w, umask=um, vmask=vm, wmask=wm, rho=rho)
    bp = np.zeros(6) if boundary_pressure is None else np.ascontiguousarray(
        boundary_pressure, dtype=np.float64)
    _check(_lib().pps_setup_poisson_bp_device(ctypes.byref(grid), *fields, dt,
                                              volume_source, _ptr(bp, 6, "boundary_pressure"),
                                              _ptr(A), _stream()),
           "pps_setup_poisson_device")
    return A


def setup_poisson_coords(grid, u, v, w, um, vm, wm, rho, dx, dxh, dy, dyh, dz, dzh, dt=1.0,
                         volume_source=0.0):
    """SetupPoisson on a stretched grid: dx ... dzh are the reference's
    spacing arrays with ghosts (numpy, n + 2 ng values each)."""
    import torch
    _need_gpu(u, v, w, um, vm, wm, rho)
    A = torch.empty(grid.shape_A, dtype=torch.float64, device=rho.device)
    fields = _padded_ptrs(grid, u=u, v=v, w=w, umask=um, vmask=vm, wmask=wm, rho=rho)
    sp = _spacing_arrays(grid, dx, dxh, dy, dyh, dz, dzh)
    _check(_lib().pps_setup_poisson_coords_device(ctypes.byref(grid), *fields, dt,
                                                  volume_source, *[_ptr(a) for a in sp],
                                                  _ptr(A), _stream()),
           "pps_setup_poisson_coords_device")
    return A


def remove_rhs_mean(grid, A):
    _need_gpu(A)
    _check(_lib().pps_remove_rhs_mean_device(ctypes.byref(grid), _ptr(A, _nA(grid), "A"), _stream()),
           "pps_remove_rhs_mean_device")
    return A


def check_setup(grid, A, umask=None, vmask=None, wmask=None):
    """check_poisson_setup: returns the status code (0 = well posed).  With
    the face masks, solid cells are read from them as the reference does;
    without, they are the cells with six zero couplings."""
    _need_gpu(A)
    if umask is None and vmask is None and wmask is None:
        return _lib().pps_check_setup_device(ctypes.byref(grid), _ptr(A, _nA(grid), "A"), _stream())
    _need_gpu(umask, vmask, wmask)
    return _lib().pps_check_poisson_setup_device(
        ctypes.byref(grid), _ptr(A, _nA(grid), "A"),
        *_padded_ptrs(grid, umask=umask, vmask=vmask, wmask=wmask), _stream())


def calc_residuals(grid, A, p):
    """calcResiduals: (L1, L2, Linf)."""
    _need_gpu(A, p)
    out = (ctypes.c_double * 3)()
    _check(_lib().pps_calc_residuals_device(ctypes.byref(grid), _ptr(A, _nA(grid), "A"),
                                            _ptr(p, _np(grid), "p"), out, _stream()),
           "pps_calc_residuals_device")
    return tuple(out)


def project_velocity(grid, u, v, w, um, vm, wm, rho, p, dt=1.0):
    _need_gpu(u, v, w, um, vm, wm, rho, p)
    fields = _padded_ptrs(grid, u=u, v=v, w=w, umask=um, vmask=vm, wmask=wm, rho=rho, p=p)
    _check(_lib().pps_project_velocity_device(ctypes.byref(grid), *fields, dt, _stream()),
           "pps_project_velocity_device")


def project_velocity_coords(grid, u, v, w, um, vm, wm, rho, p, dx, dxh, dy, dyh, dz, dzh, dt=1.0):
    """project_velocity on a stretched grid (in place on u, v, w); dx ... dzh
    are the reference's spacing arrays with ghosts (numpy)."""
    _need_gpu(u, v, w, um, vm, wm, rho, p)
    fields = _padded_ptrs(grid, u=u, v=v, w=w, umask=um, vmask=vm, wmask=wm, rho=rho, p=p)
    sp = _spacing_arrays(grid, dx, dxh, dy, dyh, dz, dzh)
    _check(_lib().pps_project_velocity_coords_device(ctypes.byref(grid), *fields, dt,
                                                     *[_ptr(a) for a in sp], _stream()),
           "pps_project_velocity_coords_device")


def face_weights(grid, dt=1.0, device="cuda"):
    """Geometric weights dt/(dx*dxh) of the +x/+y/+z faces (per plane)."""
    import torch
    out = []
    for N, L in ((grid.nx, grid.xlength), (grid.ny, grid.ylength), (grid.nz, grid.zlength)):
        x, xh, dx, dxh = coords(N, grid.ng, L)
        sl = slice(grid.ng, grid.ng + N)
        out.append(torch.tensor(dt / (dx[sl] * dxh[sl]), dtype=torch.float64, device=device))
    return tuple(out)


def _face_weight_ptrs(grid, weights):
    wx, wy, wz = weights
    _need_gpu(wx, wy, wz)
    return [_ptr(wx, grid.nx, "wx"), _ptr(wy, grid.ny, "wy"), _ptr(wz, grid.nz, "wz")]


def face_gradient(grid, p, lam, weights):
    """dJ/dc_f of the +x, +y, +z face of every cell for A = w c (the map of
    faces_to_A): (3, nz, ny, nx); wall faces get 0."""
    import torch
    _need_gpu(p, lam)
    g = torch.empty((3, grid.nz, grid.ny, grid.nx), dtype=torch.float64, device=p.device)
    _check(_lib().pps_face_gradient_device(ctypes.byref(grid), _ptr(p, _np(grid), "p"),
                                           _ptr(lam, _np(grid), "lam"),
                                           *_face_weight_ptrs(grid, weights), _ptr(g), _stream()),
           "pps_face_gradient_device")
    return g


def faces_to_A(grid, cface, weights, b):
    """8-slot system from face values c (3, nz, ny, nx) and an interior RHS."""
    import torch
    _need_gpu(cface, b)
    A = torch.empty(grid.shape_A, dtype=torch.float64, device=cface.device)
    _check(_lib().pps_faces_to_A_device(ctypes.byref(grid), _ptr(cface, 3 * grid.cells, "cface"),
                                        *_face_weight_ptrs(grid, weights),
                                        _ptr(b, grid.cells, "b"), _ptr(A), _stream()),
           "pps_faces_to_A_device")
    return A


# ---------------------------------------------------------------------------
# benchmark / test configurations (synthetic droplet fields)
# ---------------------------------------------------------------------------
CONFIGS = {
    # 64^3 walls, one sphere R=0.25, density ratio 10
    "sphere64": dict(n=64, bc=(0, 0, 0, 0, 0, 0), drops=[(0.5, 0.5, 0.5, 0.25)], rho2=10.0),
    # 128^3 walls, one sphere, density ratio 1000
    "sphere128": dict(n=128, bc=(0, 0, 0, 0, 0, 0), drops=[(0.5, 0.5, 0.5, 0.25)], rho2=1000.0),
    # 256^3 periodic-x / walls-yz, 64 droplets R in [0.03, 0.08], ratio 1000
    "drops256": dict(n=256, bc=(1, 0, 0, 1, 0, 0), ndrop=64, rmin=0.03, rmax=0.08, rho2=1000.0),
    # 512^3 walls, 512 droplets R in [0.01, 0.04], ratio 1000
    "drops512": dict(n=512, bc=(0, 0, 0, 0, 0, 0), ndrop=512, rmin=0.01, rmax=0.04, rho2=1000.0),
}


def make_problem(name=None, n=None, bc=(0, 0, 0, 0, 0, 0), drops=None, ndrop=0, rmin=0.0,
                 rmax=0.0, rho1=1.0, rho2=1000.0, seed=1, width=1.5, ng=2, dt=1.0):
    """Build a droplet-field pressure system on the GPU.

    Returns (grid, A, fields) with A the 8-slot system (right-hand side
    mean removed) and fields = dict(rho, u, v, w, um, vm, wm).
    """
    if name is not None:
        cfg = dict(CONFIGS[name])
        n = cfg.pop("n")
        bc = cfg.pop("bc")
        drops = cfg.pop("drops", None)
        ndrop = cfg.pop("ndrop", 0)
        rmin = cfg.pop("rmin", 0.0)
        rmax = cfg.pop("rmax", 0.0)
        rho2 = cfg.pop("rho2", rho2)
    grid = Grid.make(n, ng=ng, bdry_cond=bc)
    if drops is None:
        drops = synth_drops(grid, ndrop, rmin, rmax, seed)
    rho, u, v, w = synth_fields(grid, drops, rho1, rho2, width, seed)
    um, vm, wm = pressure_masks(grid)
    A = setup_poisson(grid, u, v, w, um, vm, wm, rho, dt)
    remove_rhs_mean(grid, A)
    return grid, A, dict(rho=rho, u=u, v=v, w=w, um=um, vm=vm, wm=wm)


# ---------------------------------------------------------------------------
# autograd
# ---------------------------------------------------------------------------
def _autograd():
    import torch

    class PressureSolve(torch.autograd.Function):
        """p = M(c)^-1 b with M assembled from face values c.

        forward(cface (3,nz,ny,nx), b (nz,ny,nx), solver, weights, tol, maxit)
        returns the interior pressure (nz,ny,nx).  With the fast methods the
        8-slot system is never formed: the solve takes its rows from the
        face values (without cell widths on the solver).  backward solves the adjoint M lam = grad_p on the same
        hierarchy (M is symmetric) and forms grad_c = -w (lam_L - lam_R)
        (p_L - p_R) with the fused face kernel; grad_b = lam.  The forward
        pressure is kept, not recomputed.
        """

        @staticmethod
        def forward(ctx, cface, b, solver, weights, tol=1e-12, maxit=200):
            grid = solver.grid
            cf = cface.detach().contiguous()
            bc = b.detach().contiguous()
            p = torch.zeros(grid.shape_p, dtype=torch.float64, device=b.device)
            # rms(b) in one pass over b (no squared temporary)
            bn = float(torch.linalg.vector_norm(bc)) / math.sqrt(bc.numel())
            tol_abs = tol * bn if bn > 0 else tol
            if solver.options.method in (METHODS["fp64"], METHODS["mixed"]) and not solver.has_cell_widths:
                solver.solve_faces(cf, weights, bc, p, tol_abs, maxit, allow_maxit=False)
                A = None
            else:
                A = faces_to_A(grid, cf, weights, bc)
                solver.solve(A, p, tol_abs, 1.0, maxit, 2, allow_maxit=False)
            ctx.solver, ctx.weights, ctx.tol, ctx.maxit = solver, weights, tol, maxit
            ctx.A = A
            ctx.save_for_backward(cf, p)
            return grid.interior(p).clone()

        @staticmethod
        def backward(ctx, grad_p):
            cf, p = ctx.saved_tensors
            solver = ctx.solver
            grid = solver.grid
            g = torch.zeros(grid.shape_p, dtype=torch.float64, device=p.device)
            grid.interior(g).copy_(grad_p)
            lam = torch.empty_like(g)
            if ctx.A is None:
                solver.adjoint_faces(cf, ctx.weights, g, lam, rtol=ctx.tol, maxit=ctx.maxit)
            else:
                solver.adjoint(ctx.A, g, lam, rtol=ctx.tol, maxit=ctx.maxit)
            ctx.A = None
            gc = face_gradient(grid, p, lam, ctx.weights)
            return gc, grid.interior(lam).clone(), None, None, None, None

    return PressureSolve


def __getattr__(name):
    if name == "PressureSolve":
        cls = _autograd()
        globals()["PressureSolve"] = cls
        return cls
    raise AttributeError(name)
