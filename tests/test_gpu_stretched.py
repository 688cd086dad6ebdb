"""Stretched grids (xform, paris.f90:3056-3102).  SetupPoisson's system
(modules.f90:2737-2746) is not symmetric there: a face's two couplings use
the widths of the two cells, dx(i) A2(i) = dx(i+1) A1(i+1).  Given the cell
widths, the fast methods weight each row by its cell volume.  The system is
assembled on the host by a numpy restatement of SetupPoisson on the
reference's stretched coordinates; the solution is pinned by the oracle's
residual of the caller's (unweighted) system and by the two fast forms
agreeing."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _host(t):
    return np.ascontiguousarray(t.detach().cpu().numpy())


def _axis(N, ng, L, form):
    """x, dx, dxh of paris.f90:3056-3102 (Fortran index q = 1..N+2ng at
    array index q-1)."""
    nt = N + 2 * ng
    s = (np.arange(1, nt + 1) - ng) / N
    xh = L * (form * s * (0.5 - s) * (1.0 - s) + s)
    x = np.empty(nt)
    x[1:] = 0.5 * (xh[1:] + xh[:-1])
    x[0] = 2.0 * xh[0] - x[1]
    dxh = np.empty(nt)
    dxh[:-1] = x[1:] - x[:-1]
    dxh[-1] = dxh[-2]
    dx = np.empty(nt)
    dx[1:] = xh[1:] - xh[:-1]
    dx[0] = dx[1]
    return dx, dxh


def _setup_poisson(g, form, rho, u, v, w, um, vm, wm, dt, spacings=None):
    """modules.f90:2737-2746 on padded (z, y, x) arrays; walls only (codes
    0, Poisson_BCs changes nothing)."""
    ng, nx, ny, nz = g.ng, g.nx, g.ny, g.nz
    dx, dxh = _axis(nx, ng, 1.0, form[0])
    dy, dyh = _axis(ny, ng, 1.0, form[1])
    dz, dzh = _axis(nz, ng, 1.0, form[2])
    if spacings is not None:
        spacings.extend([dx, dxh, dy, dyh, dz, dzh])
    I, J, K = slice(ng, ng + nx), slice(ng, ng + ny), slice(ng, ng + nz)
    Im, Jm, Km = slice(ng - 1, ng + nx - 1), slice(ng - 1, ng + ny - 1), slice(ng - 1, ng + nz - 1)
    Ip, Jp, Kp = slice(ng + 1, ng + nx + 1), slice(ng + 1, ng + ny + 1), slice(ng + 1, ng + nz + 1)
    r = rho
    c = r[K, J, I]
    A = np.empty((8, nz, ny, nx))
    A[0] = 2 * dt * um[K, J, Im] / (dx[I] * dxh[Im] * (r[K, J, Im] + c))
    A[1] = 2 * dt * um[K, J, I] / (dx[I] * dxh[I] * (r[K, J, Ip] + c))
    A[2] = 2 * dt * vm[K, Jm, I] / (dy[J][:, None] * dyh[Jm][:, None] * (r[K, Jm, I] + c))
    A[3] = 2 * dt * vm[K, J, I] / (dy[J][:, None] * dyh[J][:, None] * (r[K, Jp, I] + c))
    A[4] = 2 * dt * wm[Km, J, I] / (dz[K][:, None, None] * dzh[Km][:, None, None] * (r[Km, J, I] + c))
    A[5] = 2 * dt * wm[K, J, I] / (dz[K][:, None, None] * dzh[K][:, None, None] * (r[Kp, J, I] + c))
    A[6] = ((((A[0] + A[1]) + A[2]) + A[3]) + A[4]) + A[5]
    A[7] = -(((u[K, J, I] - u[K, J, Im]) / dx[I]
              + (v[K, J, I] - v[K, Jm, I]) / dy[J][:, None])
             + (w[K, J, I] - w[Km, J, I]) / dz[K][:, None, None])
    return A, dx[I], dy[J], dz[K]


def _problem(pps, n, form):
    g = pps.Grid.make(*n)
    drops = [(0.3, 0.5, 0.5, 0.2), (0.7, 0.4, 0.6, 0.15)]
    rho, u, v, w = pps.synth_fields(g, drops, 1.0, 1000.0, 1.5, seed=6)
    um, vm, wm = pps.pressure_masks(g)
    A, dx, dy, dz = _setup_poisson(g, form, *[_host(t) for t in (rho, u, v, w, um, vm, wm)], 1.0)
    # compatibility on the stretched grid: the volume-weighted mean of A8
    vol = dz[:, None, None] * dy[None, :, None] * dx[None, None, :]
    A[7] -= (vol * A[7]).sum() / vol.sum()
    return g, np.ascontiguousarray(A), dx, dy, dz


@pytest.mark.parametrize("n", [(32, 32, 32), (64, 32, 48)])
def test_fast_solves_stretched_grid_with_cell_widths(pps, orc, n):
    _say("stretched", n)
    g, A, dx, dy, dz = _problem(pps, n, (0.8, 0.5, 0.6))
    og = orc.OrcGrid.make(*n)
    At = torch.from_numpy(A).cuda()
    brms = np.sqrt((A[7] ** 2).mean())
    sols = {}
    for m in ("fp64", "mixed"):
        with pps.Solver(g, method=m) as s:
            p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
            with pytest.raises(pps.PPSError) as e:
                s.solve(At, p, 1e-10 * brms, 1.2, 100)
            assert e.value.code == pps.ERR_NONSYMMETRIC
            s.set_cell_widths(dx, dy, dz)
            p.zero_()
            it, t2 = s.solve(At, p, 1e-10 * brms, 1.2, 100, allow_maxit=False)
            _say(f"   {m}: it={it} rel.res={t2 / brms:.2e}")
            assert it <= 40
            r = orc.residual_field(og, A, _host(p))
            # the reported residual is the caller's unweighted one
            assert np.sqrt((r ** 2).mean()) < 2e-10 * brms
            assert abs(np.sqrt((r ** 2).mean()) - t2) <= 1e-3 * t2 + 1e-14 * brms
            x = _host(g.interior(p))
            sols[m] = x - x.mean()
    d = np.abs(sols["mixed"] - sols["fp64"]).max() / np.abs(sols["fp64"]).max()
    assert d < 1e-8


def _m_apply(A, L, transpose):
    """M L or M^T L for M = diag(A7) - offdiag(A1..A6) on (z, y, x) arrays,
    walls (the couplings through them are zero)."""
    out = A[6] * L
    if transpose:
        # row i's +x coupling A2(i) multiplies the unknown at i+1, etc.
        out[:, :, 1:] -= A[1][:, :, :-1] * L[:, :, :-1]
        out[:, :, :-1] -= A[0][:, :, 1:] * L[:, :, 1:]
        out[:, 1:] -= A[3][:, :-1] * L[:, :-1]
        out[:, :-1] -= A[2][:, 1:] * L[:, 1:]
        out[1:] -= A[5][:-1] * L[:-1]
        out[:-1] -= A[4][1:] * L[1:]
    else:
        out[:, :, 1:] -= A[0][:, :, 1:] * L[:, :, :-1]
        out[:, :, :-1] -= A[1][:, :, :-1] * L[:, :, 1:]
        out[:, 1:] -= A[2][:, 1:] * L[:, :-1]
        out[:, :-1] -= A[3][:, :-1] * L[:, 1:]
        out[1:] -= A[4][1:] * L[:-1]
        out[:-1] -= A[5][:-1] * L[1:]
    return out


@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_adjoint_on_stretched_grid(pps, method):
    """With cell widths the caller's M is not symmetric and the adjoint
    solves M^T lam = g: a numpy M^T applied to the returned lam gives g back,
    while M lam does not."""
    _say("stretched adjoint", method)
    g, A, dx, dy, dz = _problem(pps, (32, 24, 40), (0.8, 0.5, 0.6))
    At = torch.from_numpy(A).cuda()
    # the numpy operator agrees with the oracle's residual on a random field
    rng = np.random.default_rng(3)
    gi = rng.standard_normal((g.nz, g.ny, g.nx))
    gi -= gi.mean()
    G = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
    g.interior(G).copy_(torch.from_numpy(gi))
    lam = torch.zeros_like(G)
    with pps.Solver(g, method=method) as s:
        s.set_cell_widths(dx, dy, dz)
        it = s.adjoint(At, G, lam, rtol=1e-11, maxit=100)
    L = _host(g.interior(lam))
    grms = np.sqrt((gi ** 2).mean())
    rel_t = np.sqrt(((_m_apply(A, L, True) - gi) ** 2).mean()) / grms
    rel_n = np.sqrt(((_m_apply(A, L, False) - gi) ** 2).mean()) / grms
    _say(f"   it={it} |M^T lam - g|={rel_t:.2e} |M lam - g|={rel_n:.2e}")
    assert it <= 40
    assert rel_t < 1e-9
    assert rel_n > 1e-3


def test_m_apply_matches_oracle_residual(orc):
    """The numpy M of the adjoint test is the oracle's: b - M p equals the
    oracle's residual field for a random p (walls)."""
    import types
    g = types.SimpleNamespace(nx=12, ny=10, nz=8, ng=2)
    rng = np.random.default_rng(4)
    pad = (g.nz + 4, g.ny + 4, g.nx + 4)
    rho = 1.0 + 9.0 * rng.random(pad)
    u, v, w = (rng.standard_normal(pad) for _ in range(3))
    um, vm, wm = (np.ones(pad) for _ in range(3))
    um[:, :, 1] = um[:, :, g.nx + 1] = 0.0
    vm[:, 1, :] = vm[:, g.ny + 1, :] = 0.0
    wm[1] = wm[g.nz + 1] = 0.0
    A, *_ = _setup_poisson(g, (0.8, 0.5, 0.6), rho, u, v, w, um, vm, wm, 1.0)
    og = orc.OrcGrid.make(g.nx, g.ny, g.nz)
    P = np.zeros(og.shape_p)
    P[2:-2, 2:-2, 2:-2] = rng.standard_normal((g.nz, g.ny, g.nx))
    r = orc.residual_field(og, np.ascontiguousarray(A), P)
    assert np.allclose(A[7] - _m_apply(A, P[2:-2, 2:-2, 2:-2], False), r, rtol=0, atol=1e-9 * np.abs(r).max())


def test_stretched_assembly_on_device(pps):
    """pps_setup_poisson_coords_device with the reference's stretched
    spacings equals the numpy restatement of SetupPoisson bit for bit."""
    _say("stretched assembly")
    g = pps.Grid.make(24, 16, 20)
    drops = [(0.3, 0.5, 0.5, 0.2)]
    rho, u, v, w = pps.synth_fields(g, drops, 1.0, 1000.0, 1.5, seed=7)
    um, vm, wm = pps.pressure_masks(g)
    sp = []
    A, *_ = _setup_poisson(g, (0.7, 0.4, 0.9), *[_host(t) for t in (rho, u, v, w, um, vm, wm)], 0.5,
                           spacings=sp)
    Ad = pps.setup_poisson_coords(g, u, v, w, um, vm, wm, rho, *sp, dt=0.5)
    assert np.array_equal(_host(Ad), A)


def test_stretched_projection(pps):
    """project_velocity on stretched spacings (paris.f90:898-908 with the
    reference's dxh, dyh, dzh) equals the numpy restatement bit for bit; the
    fused solve + projection refuses a handle with cell widths (it uses the
    uniform spacings)."""
    _say("stretched projection")
    g = pps.Grid.make(24, 16, 20)
    ng = g.ng
    rho, u, v, w = pps.synth_fields(g, [(0.4, 0.5, 0.5, 0.2)], 1.0, 1000.0, 1.5, seed=8)
    um, vm, wm = pps.pressure_masks(g)
    p = torch.randn(g.shape_p, dtype=torch.float64, device="cuda")
    sp = []
    _setup_poisson(g, (0.7, 0.4, 0.9), *[_host(t) for t in (rho, u, v, w, um, vm, wm)], 0.5,
                   spacings=sp)
    dx, dxh, dy, dyh, dz, dzh = sp
    dt = 0.5
    hu, hv, hw, hp, hr = (_host(t) for t in (u, v, w, p, rho))
    hum, hvm, hwm = (_host(t) for t in (um, vm, wm))
    K, J, I = slice(ng, ng + g.nz), slice(ng, ng + g.ny), slice(ng, ng + g.nx)
    Iu, Jv, Kw = slice(ng, ng + g.nx - 1), slice(ng, ng + g.ny - 1), slice(ng, ng + g.nz - 1)
    Iu1, Jv1, Kw1 = slice(ng + 1, ng + g.nx), slice(ng + 1, ng + g.ny), slice(ng + 1, ng + g.nz)
    eu, ev, ew = hu.copy(), hv.copy(), hw.copy()
    eu[K, J, Iu] = hu[K, J, Iu] - dt * (2.0 * hum[K, J, Iu] / dxh[Iu]) * (hp[K, J, Iu1] - hp[K, J, Iu]) / (
        hr[K, J, Iu1] + hr[K, J, Iu])
    ev[K, Jv, I] = hv[K, Jv, I] - dt * (2.0 * hvm[K, Jv, I] / dyh[Jv][:, None]) * (
        hp[K, Jv1, I] - hp[K, Jv, I]) / (hr[K, Jv1, I] + hr[K, Jv, I])
    ew[Kw, J, I] = hw[Kw, J, I] - dt * (2.0 * hwm[Kw, J, I] / dzh[Kw][:, None, None]) * (
        hp[Kw1, J, I] - hp[Kw, J, I]) / (hr[Kw1, J, I] + hr[Kw, J, I])
    pps.project_velocity_coords(g, u, v, w, um, vm, wm, rho, p, *sp, dt=dt)
    for a, b in ((u, eu), (v, ev), (w, ew)):
        assert np.array_equal(_host(a), b)
    with pps.Solver(g, method="mixed") as s:
        s.set_cell_widths(dx[I], dy[J], dz[K])
        A = torch.zeros((8,) + (g.nz, g.ny, g.nx), dtype=torch.float64, device="cuda")
        q = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        with pytest.raises(pps.PPSError) as e:
            s.solve_project(A, q, u, v, w, um, vm, wm, rho, dt, 1e-8, 1.2, 5)
        assert e.value.code == pps.ERR_UNSUPPORTED


def test_uniform_widths_change_nothing(pps):
    """Equal widths normalise to weights of exactly 1: the solve is the
    unweighted one bit for bit."""
    _say("uniform widths")
    g, A, f = pps.make_problem(n=32, drops=[(0.5, 0.5, 0.5, 0.25)], rho2=100.0)
    brms = float(torch.sqrt((A[7] ** 2).mean()))
    out = []
    for widths in (None, np.full(32, 1.0 / 32)):
        with pps.Solver(g, method="mixed") as s:
            if widths is not None:
                s.set_cell_widths(widths, widths, widths)
            p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
            it, t2 = s.solve(A, p, 1e-10 * brms, 1.2, 50)
            out.append((it, t2, p.clone()))
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert torch.equal(out[0][2], out[1][2])


def test_cell_widths_errors(pps):
    _say("cell widths errors")
    g, A, f = pps.make_problem(n=16, drops=[(0.5, 0.5, 0.5, 0.25)], rho2=10.0)
    with pps.Solver(g, method="fp64") as s:
        bad = np.full(16, 1.0 / 16)
        bad[3] = 0.0
        with pytest.raises(pps.PPSError) as e:
            s.set_cell_widths(bad, bad, bad)
        assert e.value.code == pps.ERR_BAD_ARG
        s.set_cell_widths(np.full(16, 0.1), np.full(16, 0.1), np.full(16, 0.1))
        s.set_cell_widths()   # cleared: the unweighted solve again
        brms = float(torch.sqrt((A[7] ** 2).mean()))
        p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it, t2 = s.solve(A, p, 1e-10 * brms, 1.2, 50, allow_maxit=False)
        assert t2 < 1e-10 * brms
