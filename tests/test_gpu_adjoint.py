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


def test_adjoint_zero_right_hand_side(pps):
    """The backward of p.sum() on an all-walls box: the adjoint right-hand
    side is constant, hence zero after the mean-free projection, and the
    gradients are exactly zero (no non-convergence error)."""
    _say("adjoint zero right-hand side")
    grid, A, f = pps.make_problem(n=16, drops=[(0.5, 0.5, 0.5, 0.25)], rho2=100.0, seed=2)
    cf = faces_from_rho(grid, f["rho"])
    b = A[7].clone().requires_grad_(True)
    w = pps.face_weights(grid)
    for method in ("fp64", "mixed"):
        with pps.Solver(grid, method=method) as s:
            c = cf.clone().requires_grad_(True)
            p = pps.PressureSolve.apply(c, b, s, w, 1e-12, 100)
            p.sum().backward()
            assert torch.count_nonzero(c.grad) == 0
            assert torch.count_nonzero(b.grad) == 0
            b.grad = None


def test_adjoint_after_method_switch(pps, orc):
    """FP64 solve of one system, mixed solve of another, back to FP64, then
    the adjoint of the second system: the adjoint uses the second system's
    operator in the current precision (its residual is checked with the
    oracle)."""
    _say("adjoint after a method switch")
    g1, A1, _ = pps.make_problem(n=32, drops=[(0.3, 0.5, 0.5, 0.2)], rho2=1000.0, seed=3)
    g2, A2, _ = pps.make_problem(n=32, drops=[(0.7, 0.4, 0.6, 0.15)], rho2=10.0, seed=4)
    og = orc.OrcGrid.make(32)
    gi = torch.randn((32, 32, 32), dtype=torch.float64, device="cuda")
    gi -= gi.mean()
    g = torch.zeros(g1.shape_p, dtype=torch.float64, device="cuda")
    g1.interior(g).copy_(gi)
    lam = torch.empty_like(g)
    with pps.Solver(g1, method="fp64") as s:
        p = torch.zeros(g1.shape_p, dtype=torch.float64, device="cuda")
        s.solve(A1, p, 1e-10, 1.2, 50)
        s.set_method("mixed")
        p2 = torch.zeros(g1.shape_p, dtype=torch.float64, device="cuda")
        s.solve(A2, p2, 1e-10, 1.2, 50)
        s.set_method("fp64")
        it = s.adjoint(A2, g, lam, rtol=1e-12, maxit=100)
    B = A2.clone()
    B[7] = gi
    r = orc.residual_field(og, np.ascontiguousarray(B.cpu().numpy()),
                           np.ascontiguousarray(lam.cpu().numpy()))
    rel = np.sqrt((r ** 2).mean()) / float(gi.pow(2).mean().sqrt())
    _say(f"   it={it} oracle rel. residual {rel:.2e}")
    assert rel <= 1e-12


def test_adjoint_gradient_on_zero_faces(pps):
    """Faces with c_f = 0 (cut couplings inside the box) still carry the
    gradient of the map A = w c: central differences at c_f = 0 agree."""
    _say("adjoint on zero faces")
    grid, A, f = pps.make_problem(n=16, drops=[(0.5, 0.5, 0.5, 0.25)], rho2=100.0, seed=6)
    cf = faces_from_rho(grid, f["rho"])
    faces = random_faces(grid, 12, 23)
    for (d, k, j, i) in faces:
        cf[d, k, j, i] = 0.0
    b = A[7].clone()
    w = pps.face_weights(grid)
    pt = sine_target(grid, seed=4)
    PS = pps.PressureSolve
    with pps.Solver(grid, method="fp64") as s:
        c = cf.clone().requires_grad_(True)
        p = PS.apply(c, b, s, w, 1e-14, 300)
        objective(p, pt).backward()
        gr = c.grad.detach()
        eps = 1e-6
        fd, ad = [], []
        with torch.no_grad():
            for (d, k, j, i) in faces:
                pm = []
                for sgn in (1.0, -1.0):
                    cc = cf.clone()
                    cc[d, k, j, i] += sgn * eps
                    pm.append(PS.apply(cc, b, s, w, 1e-14, 300))
                fd.append(objective_diff(pm[0], pm[1], pt) / (2 * eps))
                ad.append(float(gr[d, k, j, i]))
    fd, ad = np.array(fd), np.array(ad)
    rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    _say(f"   |g|={np.linalg.norm(fd):.3e} rel.err={rel:.2e}")
    assert np.count_nonzero(ad) == len(ad)
    assert rel <= 1e-6


def test_adjoint_128_directional_vs_oracle_solver(pps, orc):
    """The 128^3 adjoint workload against the FP64 CPU solver: a central
    difference of J along a random relative perturbation v of every open
    face, each J from the oracle's CG (Jacobi-PCG, a fixed 2000 iterations,
    which reach its rounding floor of about 1e-15 relative residual; the two
    solves then run the same arithmetic sequence, so their rounding errors
    largely cancel in the difference) on the 8-slot system of the perturbed
    faces, against <dJ/dc, v> from the adjoint solve and the face kernel."""
    _say("adjoint 128^3 directional derivative vs oracle solver")
    grid, A, f = pps.make_problem("sphere128")
    og = orc.OrcGrid.make(128)
    cf = faces_from_rho(grid, f["rho"])
    b = A[7].clone()
    w = pps.face_weights(grid)
    pt = sine_target(grid, seed=11)
    PS = pps.PressureSolve
    with pps.Solver(grid, method="fp64") as s:
        c = cf.clone().requires_grad_(True)
        p = PS.apply(c, b, s, w, 1e-14, 300)
        objective(p, pt).backward()
        g = c.grad.detach()
    gen = torch.Generator(device="cuda").manual_seed(9)
    v = (torch.rand(cf.shape, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1) * cf
    eps = 1e-6

    def oracle_p(cc):
        Ah = np.ascontiguousarray(pps.faces_to_A(grid, cc, w, b).cpu().numpy())
        ph = np.zeros(og.shape_p)
        its, rel = orc.pcg_solve(og, Ah, ph, rtol=0.0, maxit=2000)
        _say(f"   oracle PCG {its} iterations, rel. residual {rel:.2e}")
        assert rel <= 1e-14
        return torch.from_numpy(og.interior(ph).copy()).cuda()

    pm = [oracle_p(cf + sgn * eps * v) for sgn in (1.0, -1.0)]
    fd = objective_diff(pm[0], pm[1], pt) / (2 * eps)
    ad = float((g * v).sum())
    rel = abs(ad - fd) / abs(fd)
    _say(f"   adjoint {ad:.8e} oracle central differences {fd:.8e} rel.err {rel:.2e}")
    assert rel <= 1e-6
