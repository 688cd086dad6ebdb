"""The C ABI: libpps.so loads and exports every function include/pps.h
declares; host-only helpers agree with the oracle.  No GPU compute here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "pps.h")).read()
    # function declarations: a return type at the start of a line, then the name
    pat = r"^(?:int|const char \*)\s*(pps_[A-Za-z0-9_]+)\s*\("
    return sorted(set(re.findall(pat, txt, flags=re.M)))


def test_library_exports_every_declared_symbol(pps):
    L = pps._lib()
    decl = _declared()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(L, name), name
    # the Python mirror binds exactly the declared set
    assert sorted(pps.exported_symbols()) == decl


def test_no_device_is_an_error_not_a_fallback(pps):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    g = pps.Grid.make(16)
    with pytest.raises(pps.PPSError) as e:
        pps.Solver(g)
    assert e.value.code == pps.ERR_NO_DEVICE


def test_coords_match_oracle(pps, orc):
    for N, L in ((16, 1.0), (24, 2.0), (128, 1.0)):
        a = pps.coords(N, 2, L)
        b = orc.coords(N, 2, L)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_synth_drops_deterministic(pps):
    g = pps.Grid.make(64)
    d1 = pps.synth_drops(g, 64, 0.03, 0.08, 7)
    d2 = pps.synth_drops(g, 64, 0.03, 0.08, 7)
    assert np.array_equal(d1, d2)
    assert d1.shape == (64, 4)
    assert np.all((d1[:, :3] >= 0) & (d1[:, :3] < 1))
    assert np.all((d1[:, 3] >= 0.03) & (d1[:, 3] <= 0.08))
    d3 = pps.synth_drops(g, 64, 0.03, 0.08, 8)
    assert not np.array_equal(d1, d3)


def test_status_strings_and_bad_args(pps):
    L = pps._lib()
    assert L.pps_status_string(pps.ERR_NOT_IN_KERNEL) == b"constant p not in kernel"
    assert L.pps_status_string(999) == b"unknown status"
    g = pps.Grid.make(16)
    g.bdry_cond[0] = 1  # periodic x- without periodic x+ (paris.f90:2980-2986)
    h = ctypes.c_void_p()
    assert L.pps_init(ctypes.byref(g), None, ctypes.byref(h)) == pps.ERR_BAD_ARG
    o = pps.default_options()
    assert (o.method, o.nrelax, o.nu_pre, o.nu_post, o.multigrid) == (pps.MG_FP64, 3, 2, 2, 1)


def test_fortran_module_compiles():
    import shutil
    import subprocess
    flang = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(flang):
        pytest.skip("no amdflang")
    src = os.path.join(ROOT, "paris-git_amd", "fortran", "pps_paris.f90")
    out = os.path.join(ROOT, "paris-git_amd", "_build", "fortran_test")
    os.makedirs(out, exist_ok=True)
    r = subprocess.run([flang, "-c", src, "-J", out, "-o", os.path.join(out, "m.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_wrong_dtype_or_size_is_refused_before_the_call(pps):
    """The Python mirror checks dtype and element count of every array it
    passes: a float32 array, or one whose size does not match the grid's
    extents, raises instead of reaching a kernel that would index past it."""
    import torch
    g = pps.Grid.make(8)
    with pytest.raises(TypeError):
        pps._ptr(np.zeros(10, dtype=np.float32))
    with pytest.raises(TypeError):
        pps._ptr(torch.zeros(10, dtype=torch.float32))
    with pytest.raises(ValueError):
        pps._ptr(np.zeros(g.shape_A[1:]), pps._nA(g), "A")
    with pytest.raises(ValueError):
        pps._ptr(torch.zeros((g.nz, g.ny, g.nx), dtype=torch.float64), pps._np(g), "p")
    with pytest.raises(ValueError):
        pps._ptr(torch.zeros((4, 4), dtype=torch.float64).t(), 16)
    with pytest.raises(TypeError):
        pps._iptr(np.zeros(g.shape_p, dtype=np.int64), pps._np(g))
    assert pps._ptr(np.zeros(g.shape_A), pps._nA(g), "A").value
    # device entry points refuse host or non-float64 tensors
    with pytest.raises(TypeError):
        pps._need_gpu(torch.zeros(4, dtype=torch.float64))
    with pytest.raises(TypeError):
        pps.check_setup(g, torch.zeros(g.shape_A, dtype=torch.float32))
    # spacing arrays must have n + 2 ng entries
    with pytest.raises(ValueError):
        pps._spacing_arrays(g, *[np.ones(g.nx) for _ in range(6)])
