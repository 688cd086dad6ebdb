"""pps_solve_device_noalloc against pps_solve_device, through the C ABI.

The no-allocation entry point runs the same kernels on buffers carved from
the library's fixed scratch array, so its pressure, iteration count and
tres2 must equal pps_solve_device's bit for bit.  Two calls in a row per
method, the second on a smaller grid (the array keeps the first call's
contents), and device memory must not shrink across a call.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (n, boundary codes, droplets): a 64^3 walled sphere (density ratio 1000),
# then a 32^3 periodic-x box of 6 droplets
CASES = [
    (64, (0, 0, 0, 0, 0, 0), [(0.5, 0.5, 0.5, 0.25)]),
    (32, (1, 0, 0, 1, 0, 0), None),
]


def _free_bytes():
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


@pytest.mark.parametrize("method", ["mixed", "fp64"])
def test_noalloc_equals_solve_device(pps, method):
    for n, bc, drops in CASES:
        print(f"noalloc vs solve_device: {method} {n}^3 bc={bc}", flush=True)
        if drops is None:
            grid, A, _ = pps.make_problem(n=n, bc=bc, ndrop=6, rmin=0.08, rmax=0.16, seed=5)
        else:
            grid, A, _ = pps.make_problem(n=n, bc=bc, drops=drops, seed=5)
        brms = float(torch.sqrt((A[7] ** 2).mean()))
        tol = 1e-10 * brms
        with pps.Solver(grid, method=method) as s:
            p_ref = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
            it_ref, t2_ref = s.solve(A, p_ref, tol, 1.2, 100, allow_maxit=False)
        with pps.Solver(grid, method=method) as s:
            p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
            free0 = _free_bytes()
            it, t2 = s.solve_noalloc(A, p, tol, 1.2, 100, allow_maxit=False)
            free1 = _free_bytes()
        print(f"   it {it} / {it_ref}, tres2/rms(b) {t2 / brms:.3e} / {t2_ref / brms:.3e}, "
              f"device free bytes {free0} -> {free1}", flush=True)
        assert it == it_ref and t2 == t2_ref
        assert torch.equal(p, p_ref)
        assert t2 < tol
        # nothing allocated on the device (the plain solve takes 5-35 MB here)
        assert free0 - free1 < (1 << 20)


def test_noalloc_limits(pps):
    """Refusals come before any work: a grid beyond the scratch array, a
    method without the fast path, another norm."""
    g = pps.Grid.make(128)
    A = torch.zeros(g.shape_A, dtype=torch.float64, device="cuda")
    p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
    with pps.Solver(g, method="mixed") as s:
        with pytest.raises(pps.PPSError) as e:
            s.solve_noalloc(A, p, 1e-10, 1.2, 10)
        assert e.value.code == pps.ERR_SCRATCH
    g = pps.Grid.make(16)
    A = torch.zeros(g.shape_A, dtype=torch.float64, device="cuda")
    p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
    with pps.Solver(g, method="reference") as s:
        with pytest.raises(pps.PPSError) as e:
            s.solve_noalloc(A, p, 1e-10, 1.2, 10)
        assert e.value.code == pps.ERR_UNSUPPORTED
    with pps.Solver(g, method="fp64") as s:
        with pytest.raises(pps.PPSError) as e:
            s.solve_noalloc(A, p, 1e-10, 1.2, 10, norm=1)
        assert e.value.code == pps.ERR_UNSUPPORTED
