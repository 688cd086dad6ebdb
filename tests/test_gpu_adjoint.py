"""Adjoint: the backward of PressureSolve (adjoint solve on the same
hierarchy + fused face kernel) against central finite differences of the
forward solve, and the adjoint solve against the oracle's residual."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _faces_from_rho(grid, rho):
    """c_f = 2/(rho_L+rho_R) on every +x/+y/+z face; closed (0) on walls."""
    ng = grid.ng
    R = rho[ng - 1:ng + grid.nz + 1, ng - 1:ng + grid.ny + 1, ng - 1:ng + grid.nx + 1]
    c = R[1:-1, 1:-1, 1:-1]
    cx = 2.0 / (c + R[1:-1, 1:-1, 2:])
    cy = 2.0 / (c + R[1:-1, 2:, 1:-1])
    cz = 2.0 / (c + R[2:, 1:-1, 1:-1])
    per = grid.periodic()
    if not per[0]:
        cx[:, :, -1] = 0.0
    if not per[1]:
        cy[:, -1, :] = 0.0
    if not per[2]:
        cz[-1, :, :] = 0.0
    return torch.stack([cx, cy, cz]).contiguous()


def _target(grid, seed=3):
    """p_target: 8 random-amplitude sine modes, wavenumbers 1..4."""
    rng = np.random.default_rng(seed)
    x = (np.arange(grid.nx) + 0.5) / grid.nx
    y = (np.arange(grid.ny) + 0.5) / grid.ny
    z = (np.arange(grid.nz) + 0.5) / grid.nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    t = np.zeros_like(X)
    for _ in range(8):
        k = rng.integers(1, 5, size=3)
        a = rng.uniform(-1, 1)
        ph = rng.uniform(0, 2 * np.pi)
        t += a * np.sin(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z) + ph)
    return torch.from_numpy(t).cuda()


def _J(p, pt):
    p0 = p - p.mean()
    return 0.5 * ((p0 - pt) ** 2).sum()


@pytest.mark.parametrize("n,bc,rho2", [(32, (0,) * 6, 1000.0), (16, (1, 0, 0, 1, 0, 0), 100.0)])
def test_adjoint_gradient_vs_central_fd(pps, n, bc, rho2):
    _say("adjoint FD", n, bc, rho2)
    grid, A, f = pps.make_problem(n=n, bc=bc, drops=[(0.5, 0.5, 0.5, 0.25)], rho2=rho2, seed=5)
    cf = _faces_from_rho(grid, f["rho"])
    b = A[7].clone()
    w = pps.face_weights(grid)
    pt = _target(grid)
    PS = pps.PressureSolve
    with pps.Solver(grid, method="fp64") as s:
        c = cf.clone().requires_grad_(True)
        p = PS.apply(c, b, s, w, 1e-13, 300)
        J = _J(p, pt)
        J.backward()
        g = c.grad.detach()
        # the assembled operator equals SetupPoisson's (face values from rho)
        A2 = pps.faces_to_A(grid, cf, w, b)
        assert torch.allclose(A2[:7], A[:7], rtol=1e-14, atol=0)

        rng = np.random.default_rng(7)
        per = grid.periodic()
        faces = []
        while len(faces) < 40:
            d = int(rng.integers(0, 3))
            k, j, i = (int(rng.integers(0, m)) for m in (grid.nz, grid.ny, grid.nx))
            last = (i == grid.nx - 1, j == grid.ny - 1, k == grid.nz - 1)[d]
            if last and not per[d]:
                continue
            faces.append((d, k, j, i))
        eps = 1e-6
        fd, ad = [], []
        with torch.no_grad():
            for (d, k, j, i) in faces:
                Jpm = []
                for sgn in (1.0, -1.0):
                    cc = cf.clone()
                    cc[d, k, j, i] += sgn * eps
                    Jpm.append(float(_J(PS.apply(cc, b, s, w, 1e-13, 300), pt)))
                fd.append((Jpm[0] - Jpm[1]) / (2 * eps))
                ad.append(float(g[d, k, j, i]))
        fd, ad = np.array(fd), np.array(ad)
        rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
        _say(f"   |g|={np.linalg.norm(fd):.3e} rel.err={rel:.2e}")
        assert rel < 1e-5


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
