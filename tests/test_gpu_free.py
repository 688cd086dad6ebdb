"""FreeSolver (freesurface.f90:700-855) through the C ABI against the
oracle: the masked RB-SOR sweeps keep the reference's arithmetic, so p and
the iteration count match; tres2 differs only by the order of its sum."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _host(t):
    return np.ascontiguousarray(t.detach().cpu().numpy())


def _case(pps, orc, n, bc, flag, seed):
    g = pps.Grid.make(n, bdry_cond=bc)
    og = orc.OrcGrid.make(n, bc=bc)
    _, At, _ = pps.make_problem(n=n, bc=bc, drops=[(0.4, 0.5, 0.6, 0.25)], rho2=100.0, seed=seed)
    rng = np.random.default_rng(seed)
    m = np.zeros(og.shape_p, dtype=np.int32)
    vals, prob = ([0, 1, 2, 3], [0.7, 0.1, 0.1, 0.1]) if flag == 1 else ([0, 1, 2, 3], None)
    og.interior(m)[...] = rng.choice(vals, size=(n, n, n), p=prob).astype(np.int32)
    pgas = rng.standard_normal(og.shape_p)
    p0 = np.zeros(og.shape_p)
    og.interior(p0)[...] = rng.standard_normal((n, n, n))
    return g, og, At, m, pgas, p0


@pytest.mark.parametrize("flag,bc", [(1, (0,) * 6), (1, (1, 0, 0, 1, 0, 0)), (2, (0,) * 6),
                                     (2, (1,) * 6)])
def test_free_solver_matches_oracle(pps, orc, flag, bc):
    _say("free solver", flag, bc)
    n = 16
    g, og, At, m, pgas, p0 = _case(pps, orc, n, bc, flag, seed=5 + flag)
    A = _host(At)
    tol = 1e-10 * np.abs(A[6]).max()
    pr = p0.copy()
    rc, it, t2, hist = orc.free_solver(og, A, pr, m, pgas, flag, tol, 1.3, 5000)
    assert rc == 0 and it < 5000
    with pps.Solver(g, method="sor") as s:
        p = torch.from_numpy(p0).cuda()
        git, gt2 = s.free_solve(At, p, torch.from_numpy(m).cuda(),
                                torch.from_numpy(pgas).cuda() if flag == 1 else None, flag, tol,
                                1.3, 5000)
        ghist = np.array(s.history())
    # interiors (the GPU keeps face ghosts only; edges and corners of p are
    # never read by the 7-point stencil)
    pg, po = og.interior(_host(p)), og.interior(pr)
    d = np.abs(pg - po).max() / np.abs(po).max()
    _say(f"   it={git} (oracle {it}) max rel diff {d:.1e} bit-exact {np.array_equal(pg, po)}")
    assert git == it
    assert d <= 1e-13
    assert gt2 == pytest.approx(t2, rel=1e-12)
    assert np.allclose(ghist, hist, rtol=1e-12, atol=0)


def test_free_solver_host_entry_point_and_errors(pps, orc):
    _say("free solver host / errors")
    n = 16
    g, og, At, m, pgas, p0 = _case(pps, orc, n, (0,) * 6, 1, seed=11)
    A = _host(At)
    tol = 1e-10 * np.abs(A[6]).max()
    with pps.Solver(g, method="sor") as s:
        ph = p0.copy()
        it_h, ierr, t2_h = s.free_solver(A, ph, m, pgas, 1, tol, 1.3, 5000)
        pd = torch.from_numpy(p0).cuda()
        it_d, t2_d = s.free_solve(At, pd, torch.from_numpy(m).cuda(), torch.from_numpy(pgas).cuda(),
                                  1, tol, 1.3, 5000)
        assert ierr == 0 and it_h == it_d and t2_h == t2_d
        assert np.array_equal(ph, _host(pd))
        # not converged: it = maxit + 1
        pn = p0.copy()
        it_n, _, _ = s.free_solver(A, pn, m, pgas, 1, tol, 1.3, 3)
        assert it_n == 4
        # solver_flag outside {1, 2}; NaN in A
        with pytest.raises(pps.PPSError):
            s.free_solver(A, p0.copy(), m, pgas, 0, tol)
        # NaN in an active cell's coefficients (inactive cells are not checked,
        # as in catch_divergence_fs)
        B = A.copy()
        B[2, 3, 4, 5] = np.nan
        m2 = m.copy()
        og.interior(m2)[3, 4, 5] = 0
        with pytest.raises(pps.PPSError):
            s.free_solver(B, p0.copy(), m2, pgas, 1, tol)
