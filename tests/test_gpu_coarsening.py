"""General coarsening of the fast V-cycle (sizes and cell shapes the
reference's multigrid refuses, MGsolver.f90:44-46):

* odd sizes on wall axes: the last coarse cell of an odd axis has a single
  child (50^3, 49 x 64 x 40, 37 x 30 x 45);
* semi-coarsening: axes whose cells are more than 1.5 times as long as the
  shortest are not halved until the cells are nearly cubic (96 x 48 x 80 in
  a unit box, 64 x 64 x 32 in a 1 x 1 x 4 box);
* a periodic axis of odd size is never halved, and a level that has one is
  the coarsest (more than 512 cells there is a bad size).

Each solve is pinned by the oracle: its converged solution (Jacobi-PCG to
its rounding floor) to 1e-10 in the max norm with the null-space mean
removed, and the oracle's residual of the returned pressure.  One V-cycle
on these irregular hierarchies is a symmetric operator (CG needs it).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _host(t):
    return np.ascontiguousarray(t.detach().cpu().numpy())


def _mf(x):
    return x - x.mean()


def _system(pps, shape, bc=(0,) * 6, lengths=(1.0, 1.0, 1.0), rho2=1000.0, seed=3):
    g = pps.Grid.make(*shape, bdry_cond=bc, lengths=lengths)
    drops = [(0.3 * lengths[0], 0.5 * lengths[1], 0.45 * lengths[2], 0.2 * min(lengths)),
             (0.7 * lengths[0], 0.35 * lengths[1], 0.6 * lengths[2], 0.12 * min(lengths))]
    rho, u, v, w = pps.synth_fields(g, drops, 1.0, rho2, 1.5, seed=seed)
    um, vm, wm = pps.pressure_masks(g)
    A = pps.setup_poisson(g, u, v, w, um, vm, wm, rho, 1.0)
    pps.remove_rhs_mean(g, A)
    return g, A


CASES = [
    ((50, 50, 50), (0,) * 6, (1.0, 1.0, 1.0)),
    ((49, 64, 40), (0,) * 6, (1.0, 1.0, 1.0)),
    ((37, 30, 45), (0,) * 6, (1.0, 1.0, 1.0)),
    ((64, 49, 40), (1, 0, 0, 1, 0, 0), (1.0, 1.0, 1.0)),
    ((96, 48, 80), (0,) * 6, (1.0, 1.0, 1.0)),
    ((64, 64, 32), (0,) * 6, (1.0, 1.0, 4.0)),
]


@pytest.mark.parametrize("shape,bc,lengths", CASES)
def test_irregular_hierarchy_matches_oracle(pps, orc, shape, bc, lengths):
    _say("irregular hierarchy", shape, bc, lengths)
    g, A = _system(pps, shape, bc, lengths)
    og = orc.OrcGrid.make(*shape, bc=bc, lengths=lengths)
    Ah = _host(A)
    brms = float(np.sqrt((Ah[7] ** 2).mean()))
    ph = np.zeros(og.shape_p)
    its, rel = orc.pcg_solve(og, Ah, ph, rtol=0.0, maxit=4000)
    _say(f"   oracle PCG {its} iterations, rel. residual {rel:.2e}")
    assert rel <= 1e-14
    ref = _mf(og.interior(ph))
    for method in ("fp64", "mixed"):
        with pps.Solver(g, method=method) as s:
            p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
            it, t2 = s.solve(A, p, 1e-10 * brms, 1.2, 60, allow_maxit=False)
            r = orc.residual_field(og, Ah, _host(p))
            rel10 = float(np.sqrt((r ** 2).mean()) / brms)
            p.zero_()
            it2, t22 = s.solve(A, p, 1e-14 * brms, 1.2, 80, allow_maxit=False)
        err = float(np.abs(_mf(og.interior(_host(p))) - ref).max() / np.abs(ref).max())
        _say(f"   {method}: 1e-10 in {it} it (oracle rel. residual {rel10:.2e}); "
             f"to stagnation {it2} it, err_inf {err:.2e}")
        assert rel10 < 1e-10 and it <= 30
        assert err <= 1e-10


def _vcycle(pps, grid, A, rhs, method):
    B = A.clone()
    B[7] = rhs
    with pps.Solver(grid, method=method, outer=pps.OUTER_REFINE) as s:
        p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
        s.solve(B, p, 1e-300, 1.2, 1)
    return grid.interior(p).clone()


@pytest.mark.parametrize("shape,bc,lengths", [CASES[0], CASES[3], CASES[5]])
@pytest.mark.parametrize("method,tol", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_irregular_vcycle_is_symmetric(pps, shape, bc, lengths, method, tol):
    _say("irregular V-cycle symmetry", shape, bc, lengths, method)
    g, A = _system(pps, shape, bc, lengths)
    nz, ny, nx = shape[2], shape[1], shape[0]
    gen = torch.Generator(device="cuda").manual_seed(11)
    u = torch.randn((nz, ny, nx), generator=gen, device="cuda", dtype=torch.float64)
    v = torch.randn((nz, ny, nx), generator=gen, device="cuda", dtype=torch.float64)
    u -= u.mean()
    v -= v.mean()
    Vu = _vcycle(pps, g, A, u, method)
    Vv = _vcycle(pps, g, A, v, method)
    a, b = float((Vu * v).sum()), float((u * Vv).sum())
    scale = float(Vu.norm() * v.norm())
    _say(f"   <Vu,v>={a:.12e} <u,Vv>={b:.12e} rel {abs(a - b) / scale:.2e}")
    assert abs(a - b) <= tol * scale


def test_periodic_odd_sizes(pps):
    """A periodic axis of odd size is not halved and its level is the
    coarsest: 49^3 fully periodic has 117,649 cells there (a bad size), while
    a 7 x 8 x 8 periodic box (448 cells) is solved exactly at once."""
    _say("periodic odd sizes")
    g = pps.Grid.make(49, bdry_cond=(1,) * 6)
    A = torch.zeros((8, 49, 49, 49), dtype=torch.float64, device="cuda")
    A[0:6] = 1.0
    A[6] = 6.0
    p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
    with pps.Solver(g, method="mixed") as s:
        with pytest.raises(pps.PPSError) as e:
            s.solve(A, p, 1e-8, 1.2, 10)
    assert e.value.code == pps.ERR_BAD_SIZE
    g, A = _system(pps, (7, 8, 8), bc=(1,) * 6, rho2=10.0)
    brms = float(torch.sqrt((A[7] ** 2).mean()))
    with pps.Solver(g, method="fp64") as s:
        p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it, t2 = s.solve(A, p, 1e-12 * brms, 1.2, 20, allow_maxit=False)
    _say(f"   7x8x8 periodic: it={it} rel.res {t2 / brms:.2e}")
    assert it <= 3
