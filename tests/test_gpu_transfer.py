"""The fast V-cycle's grid transfers: the default pair (8-child average,
piecewise-constant prolongation) and the trilinear pair (trilinear
prolongation with its transpose, full weighting over 4 x 4 x 4 fine cells).

* One V-cycle is a symmetric operator for both pairs (R = P^T / 8 and a
  symmetric smoothing order): <V u, v> = <u, V v> for random mean-free u, v.
  A restriction that is not the exact transpose of the prolongation (a
  wrong weight, a wrong wall or periodic wrap) breaks this.
* With trilinear transfers both forms reach the oracle's converged solution
  (Jacobi-PCG to a 1e-15 relative residual) to 1e-10 in the max norm.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _host(t):
    return np.ascontiguousarray(t.detach().cpu().numpy())


def _vcycle(pps, grid, A, rhs, method, transfer):
    """One V-cycle applied to rhs from a zero start (one step of the plain
    refinement outer iteration): returns V rhs (interior)."""
    B = A.clone()
    B[7] = rhs
    with pps.Solver(grid, method=method, outer=pps.OUTER_REFINE, transfer=transfer) as s:
        p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
        s.solve(B, p, 1e-300, 1.2, 1)
    return grid.interior(p).clone()


@pytest.mark.parametrize("transfer", ["constant", "trilinear"])
@pytest.mark.parametrize("method,tol", [("fp64", 1e-12), ("mixed", 1e-5)])
@pytest.mark.parametrize("n,bc", [(32, (0,) * 6), (32, (1, 0, 1, 1, 0, 1)), (64, (1, 0, 0, 1, 0, 0))])
def test_vcycle_is_symmetric(pps, transfer, method, tol, n, bc):
    _say("V-cycle symmetry", transfer, method, n, bc)
    grid, A, _ = pps.make_problem(n=n, bc=bc, drops=[(0.3, 0.5, 0.6, 0.2), (0.8, 0.3, 0.4, 0.15)],
                                  rho2=1000.0, seed=6)
    gen = torch.Generator(device="cuda").manual_seed(n)
    u = torch.randn((n, n, n), generator=gen, device="cuda", dtype=torch.float64)
    v = torch.randn((n, n, n), generator=gen, device="cuda", dtype=torch.float64)
    u -= u.mean()
    v -= v.mean()
    Vu = _vcycle(pps, grid, A, u, method, transfer)
    Vv = _vcycle(pps, grid, A, v, method, transfer)
    a, b = float((Vu * v).sum()), float((u * Vv).sum())
    scale = float(Vu.norm() * v.norm())
    _say(f"   <Vu,v>={a:.12e} <u,Vv>={b:.12e} rel {abs(a - b) / scale:.2e}")
    assert abs(a - b) <= tol * scale
    # and V is not trivially symmetric: it is not a multiple of the identity
    assert float((Vu - (Vu * u).sum() / (u * u).sum() * u).norm()) > 1e-3 * float(Vu.norm())


@pytest.mark.parametrize("method", ["fp64", "mixed"])
@pytest.mark.parametrize("n,bc,rho2", [(32, (0,) * 6, 1000.0), (32, (1, 0, 0, 1, 0, 0), 1000.0),
                                       (32, (1,) * 6, 100.0), (64, (0, 1, 0, 0, 1, 0), 10.0)])
def test_trilinear_matches_oracle(pps, orc, method, n, bc, rho2):
    _say("trilinear vs oracle", method, n, bc, rho2)
    grid, A, _ = pps.make_problem(n=n, bc=bc, drops=[(0.3, 0.5, 0.6, 0.2), (0.8, 0.3, 0.4, 0.15)],
                                  rho2=rho2, seed=7)
    og = orc.OrcGrid.make(n, bc=bc)
    Ah = _host(A)
    ph = np.zeros(og.shape_p)
    its, rel = orc.pcg_solve(og, Ah, ph, rtol=1e-15, maxit=200000)
    assert rel <= 2e-15
    ref = og.interior(ph) - og.interior(ph).mean()
    brms = float(np.sqrt((Ah[7] ** 2).mean()))
    with pps.Solver(grid, method=method, transfer="trilinear") as s:
        p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
        it, t2 = s.solve(A, p, 1e-14 * brms, 1.2, 200)
    x = og.interior(_host(p))
    x = x - x.mean()
    err = np.abs(x - ref).max() / np.abs(ref).max()
    _say(f"   it={it} rel.res={t2 / brms:.2e} err={err:.2e}")
    assert it <= 60 and err <= 1e-10


@pytest.mark.parametrize("name", ["sphere128", "drops256"])
def test_transfers_at_size(pps, orc, name):
    """Both transfer pairs on the 128^3 and 256^3 workloads to the metric's
    1e-10: the oracle's residual of each result meets it, iterations and
    times are reported, and an adjoint solve (a smooth right-hand side)
    converges with each pair."""
    _say("transfers at size", name)
    from helpers import sine_target
    grid, A, _ = pps.make_problem(name)
    og = orc.OrcGrid.make(grid.nx, bc=tuple(grid.bdry_cond))
    Ah = _host(A)
    brms = float(np.sqrt((Ah[7] ** 2).mean()))
    sols = {}
    for method in ("mixed", "fp64"):
        for transfer in ("constant", "trilinear"):
            with pps.Solver(grid, method=method, transfer=transfer) as s:
                p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
                s.solve(A, p, 1e-10 * brms, 1.2, 100)   # warm-up
                p.zero_()
                it, t2 = s.solve(A, p, 1e-10 * brms, 1.2, 100, allow_maxit=False)
                ms = s.timing()[0]
                pt = sine_target(grid, seed=21)
                g = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
                grid.interior(g).copy_(pt - pt.mean())
                lam = torch.empty_like(g)
                ita = s.adjoint(A, g, lam, rtol=1e-10, maxit=100)
            r = orc.residual_field(og, Ah, _host(p))
            rel = float(np.sqrt((r ** 2).mean()) / brms)
            _say(f"   {method} {transfer}: it={it} {ms:.2f} ms oracle rel.res {rel:.2e} adjoint it={ita}")
            assert rel < 1e-10 and it <= 30 and ita <= 40
            x = og.interior(_host(p))
            sols[(method, transfer)] = x - x.mean()
    ref = sols[("fp64", "constant")]
    for k, x in sols.items():
        d = np.abs(x - ref).max() / np.abs(ref).max()
        _say(f"   {k} vs fp64/constant {d:.2e}")
        assert d <= 1e-7
