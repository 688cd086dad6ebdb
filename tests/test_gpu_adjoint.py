"""Adjoint: the backward of PressureSolve (adjoint solve on the same
hierarchy + fused face kernel) against central finite differences of the
forward solve, and the adjoint solve against the oracle's residual."""
import numpy as np
import pytest

from helpers import faces_from_rho, objective, objective_diff, random_faces, sine_target

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n,bc,rho2", [(32, (0,) * 6, 1000.0), (16, (1, 0, 0, 1, 0, 0), 100.0)])
def test_adjoint_gradient_vs_central_fd(pps, n, bc, rho2):
    _say("adjoint FD", n, bc, rho2)
    grid, A, f = pps.make_problem(n=n, bc=bc, drops=[(0.5, 0.5, 0.5, 0.25)], rho2=rho2, seed=5)
    cf = faces_from_rho(grid, f["rho"])
    b = A[7].clone()
    w = pps.face_weights(grid)
    pt = sine_target(grid)
    PS = pps.PressureSolve
    with pps.Solver(grid, method="fp64") as s:
        c = cf.clone().requires_grad_(True)
        p = PS.apply(c, b, s, w, 1e-14, 300)
        J = objective(p, pt)
        J.backward()
        g = c.grad.detach()
        # the assembled operator equals SetupPoisson's (face values from rho)
        A2 = pps.faces_to_A(grid, cf, w, b)
        assert torch.allclose(A2[:7], A[:7], rtol=1e-14, atol=0)

        eps = 1e-6
        fd, ad = [], []
        with torch.no_grad():
            for (d, k, j, i) in random_faces(grid, 40, 7):
                pm = []
                for sgn in (1.0, -1.0):
                    cc = cf.clone()
                    cc[d, k, j, i] += sgn * eps
                    pm.append(PS.apply(cc, b, s, w, 1e-14, 300))
                fd.append(objective_diff(pm[0], pm[1], pt) / (2 * eps))
                ad.append(float(g[d, k, j, i]))
        fd, ad = np.array(fd), np.array(ad)
        rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
        _say(f"   |g|={np.linalg.norm(fd):.3e} rel.err={rel:.2e}")
        assert rel <= 1e-6


def test_adjoint_gradient_128_200_faces(pps):
    """The adjoint workload at 128^3: the single-sphere field, density ratio
    1000, J = 0.5||p - p_target||^2 with 8 random sine modes; the gradient on
    200 random faces against central differences (eps = 1e-6, FP64 solves
    to a 1e-14 relative residual)."""
    _say("adjoint 128^3, 200 faces")
    grid, A, f = pps.make_problem("sphere128")
    cf = faces_from_rho(grid, f["rho"])
    b = A[7].clone()
    w = pps.face_weights(grid)
    pt = sine_target(grid, seed=11)
    PS = pps.PressureSolve
    tol = 1e-14
    with pps.Solver(grid, method="fp64") as s:
        c = cf.clone().requires_grad_(True)
        p = PS.apply(c, b, s, w, tol, 300)
        objective(p, pt).backward()
        g = c.grad.detach()
        eps = 1e-6
        fd, ad = [], []
        with torch.no_grad():
            for (d, k, j, i) in random_faces(grid, 200, 13):
                pm = []
                for sgn in (1.0, -1.0):
                    cc = cf.clone()
                    cc[d, k, j, i] += sgn * eps
                    pm.append(PS.apply(cc, b, s, w, tol, 300))
                fd.append(objective_diff(pm[0], pm[1], pt) / (2 * eps))
                ad.append(float(g[d, k, j, i]))
    fd, ad = np.array(fd), np.array(ad)
    rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    worst = np.max(np.abs(ad - fd)) / np.max(np.abs(fd))
    _say(f"   |g|={np.linalg.norm(fd):.3e} rel.err(2-norm)={rel:.2e} max-err/max={worst:.2e}")
    assert rel <= 1e-6


def test_adjoint_gradient_vs_oracle_solver_fd(pps, orc):
    """The adjoint gradient against central differences of the FP64 CPU
    solver (the oracle's CG on the same 8-slot systems, assembled from the
    perturbed faces)."""
    _say("adjoint vs oracle-solver FD")
    n = 16
    grid, A, f = pps.make_problem(n=n, drops=[(0.5, 0.5, 0.5, 0.25)], rho2=1000.0, seed=4)
    og = orc.OrcGrid.make(n)
    cf = faces_from_rho(grid, f["rho"])
    b = A[7].clone()
    w = pps.face_weights(grid)
    pt = sine_target(grid, seed=2)
    PS = pps.PressureSolve
    with pps.Solver(grid, method="fp64") as s:
        c = cf.clone().requires_grad_(True)
        p = PS.apply(c, b, s, w, 1e-14, 300)
        objective(p, pt).backward()
        g = c.grad.detach()

    def oracle_p(cc):
        Ah = np.ascontiguousarray(pps.faces_to_A(grid, cc, w, b).cpu().numpy())
        ph = np.zeros(og.shape_p)
        its, rel = orc.pcg_solve(og, Ah, ph, rtol=1e-15, maxit=100000)
        assert rel <= 1e-14
        return torch.from_numpy(og.interior(ph).copy()).cuda()

    eps = 1e-6
    fd, ad = [], []
    for (d, k, j, i) in random_faces(grid, 30, 17):
        pm = []
        for sgn in (1.0, -1.0):
            cc = cf.clone()
            cc[d, k, j, i] += sgn * eps
            pm.append(oracle_p(cc))
        fd.append(objective_diff(pm[0], pm[1], pt) / (2 * eps))
        ad.append(float(g[d, k, j, i]))
    fd, ad = np.array(fd), np.array(ad)
    rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    _say(f"   |g|={np.linalg.norm(fd):.3e} rel.err={rel:.2e}")
    assert rel <= 1e-6


def test_adjoint_repeated_calls_on_caller_stream(pps):
    """The adjoint reads its right-hand side after the work the caller queued
    on the same (default) stream: repeated calls with the same g on one
    handle give the same lam bit for bit, each satisfying M lam = g."""
    _say("adjoint repeated calls")
    grid, A, f = pps.make_problem("drops256")
    pt = sine_target(grid, seed=21)
    gi = pt - pt.mean()
    gn = float(gi.pow(2).mean().sqrt())
    lams = []
    with pps.Solver(grid, method="mixed") as s:
        for _ in range(3):
            g = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
            grid.interior(g).copy_(gi)
            lam = torch.empty_like(g)
            s.adjoint(A, g, lam, rtol=1e-10, maxit=60)
            B = A.clone()
            B[7] = gi
            res = pps.calc_residuals(grid, B, lam)[1] / gn
            _say(f"   true rel. residual {res:.2e}")
            assert res <= 1e-10
            lams.append(lam)
    assert torch.equal(lams[0], lams[1]) and torch.equal(lams[1], lams[2])


def test_adjoint_solve_residual(pps, orc):
    _say("adjoint residual")
    grid, A, f = pps.make_problem(n=32, drops=[(0.5, 0.5, 0.5, 0.25)], rho2=1000.0, seed=8)
    og = orc.OrcGrid.make(32)
    g = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
    gi = torch.randn((32, 32, 32), dtype=torch.float64, device="cuda")
    grid.interior(g).copy_(gi - gi.mean())
    lam = torch.empty_like(g)
    with pps.Solver(grid, method="mixed") as s:
        it = s.adjoint(A, g, lam, rtol=1e-12, maxit=100)
    assert it < 30
    B = A.clone()
    B[7] = grid.interior(g)
    r = orc.residual_field(og, np.ascontiguousarray(B.cpu().numpy()), np.ascontiguousarray(lam.cpu().numpy()))
    assert np.sqrt((r ** 2).mean()) < 1e-12 * float(torch.sqrt((gi - gi.mean()).pow(2).mean()))
