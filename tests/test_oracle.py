"""Pins the oracle (CPU restatement of the reference path) against what the
mathematics fixes: closed forms, invariants, an independent direct solver,
and the reference's own known-answer test (Tests/PoissonMG)."""
import numpy as np
import pytest

from helpers import droplet_fields, exact_solution, sparse_matrix, system, wrap_ghosts


def _mf(x):
    return x - x.mean()


def test_coords_uniform_closed_form(orc):
    # paris.f90:3046-3102 with xform=0 and N a power of two: every spacing is exact
    for N, L in ((16, 1.0), (32, 2.0), (64, 1.0)):
        x, xh, dx, dxh = orc.coords(N, 2, L)
        h = L / N
        q = np.arange(N + 4)
        assert np.array_equal(xh, L * ((q + 1 - 2) / N))
        assert np.array_equal(dx, np.full(N + 4, h))
        assert np.array_equal(dxh, np.full(N + 4, h))
        assert np.allclose(x[1:], (q[1:] + 1 - 2.5) * h, rtol=0, atol=1e-15)


def test_assembly_closed_form_uniform_density(orc):
    # rho = 1, dt = 1, h = 1/16: every open coupling is exactly 1/h^2, walls 0,
    # and a linear velocity u = x gives div = 1 -> A8 = -1
    g = orc.OrcGrid.make(16)
    rho = np.ones(g.shape_p)
    x = orc.coords(16, 2, 1.0)[1]  # face positions xh
    u = np.broadcast_to(x[None, None, :], g.shape_p).copy()
    v = np.zeros(g.shape_p)
    w = np.zeros(g.shape_p)
    um, vm, wm = orc.pressure_masks(g)
    rc, A = orc.setup_poisson(g, u, v, w, um, vm, wm, rho, 1.0, 0.0)
    assert rc == 0
    h2 = 1.0 / 16 ** 2
    assert np.all(A[0][:, :, 1:] == 1.0 / h2) and np.all(A[0][:, :, 0] == 0.0)
    assert np.all(A[1][:, :, :-1] == 1.0 / h2) and np.all(A[1][:, :, -1] == 0.0)
    assert np.all(A[2][:, 0, :] == 0.0) and np.all(A[3][:, -1, :] == 0.0)
    assert np.all(A[4][0] == 0.0) and np.all(A[5][-1] == 0.0)
    assert np.array_equal(A[6], A[0] + A[1] + A[2] + A[3] + A[4] + A[5])
    # u is the face coordinate, so (u(i)-u(i-1))/dx = 1 except at the walls,
    # where the mask does not touch u (the walls' u is the coordinate too)
    assert np.allclose(A[7], -1.0, rtol=0, atol=1e-12)


def test_assembly_invariants_droplets(orc):
    g = orc.OrcGrid.make(16, bc=(1, 0, 0, 1, 0, 0))
    A, f = system(orc, g, [(0.3, 0.5, 0.5, 0.2), (0.95, 0.4, 0.6, 0.15)], rho2=1000.0, remove_mean=False)
    # symmetry A2(i) = A1(i+1) (periodic wrap in x), zero row sums
    assert np.array_equal(A[1], np.roll(A[0], -1, axis=2))
    assert np.array_equal(A[3][:, :-1], A[2][:, 1:])
    assert np.array_equal(A[5][:-1], A[4][1:])
    assert np.array_equal(A[6], ((((A[0] + A[1]) + A[2]) + A[3]) + A[4]) + A[5])
    # closed box in y/z, periodic in x: the divergence sums to zero
    assert abs(A[7].sum()) < 1e-9 * np.abs(A[7]).sum()
    # the face coupling is dt*c_f/h^2 with c_f = 2/(rho_L+rho_R)
    rho = g.interior(f["rho"])
    c = 2.0 / (rho[:, :, :-1] + rho[:, :, 1:])
    assert np.allclose(A[1][:, :, :-1], c * 16 ** 2, rtol=1e-15)
    rc, _ = orc.check_poisson_setup(g, np.ascontiguousarray(A), f["um"], f["vm"], f["wm"])
    assert rc == 0


def test_check_poisson_setup_errors(orc):
    g = orc.OrcGrid.make(8)
    A, f = system(orc, g, [(0.5, 0.5, 0.5, 0.25)])
    masks = (f["um"], f["vm"], f["wm"])
    B = A.copy(); B[3, 2, 2, 2] = np.nan
    assert orc.check_poisson_setup(g, B, *masks)[0] == orc.ERR_NAN_OR_TINY_DIAG
    B = A.copy(); B[6, 1, 1, 1] = 0.0
    assert orc.check_poisson_setup(g, B, *masks)[0] == orc.ERR_NAN_OR_TINY_DIAG
    B = A.copy(); B[7] += 1e-8
    assert orc.check_poisson_setup(g, B, *masks)[0] == orc.ERR_VOLUME_SOURCE
    B = A.copy(); B[7] -= 1e-8  # the reference's test is one-sided
    assert orc.check_poisson_setup(g, B, *masks)[0] == orc.OK
    B = A.copy(); B[6, 3, 3, 3] *= 1.001
    assert orc.check_poisson_setup(g, B, *masks)[0] == orc.ERR_NOT_IN_KERNEL


def test_sparse_matrix_is_the_residual_operator(orc):
    # the scipy matrix and the oracle's residual agree on random vectors
    g = orc.OrcGrid.make(8, bc=(1, 0, 1, 1, 0, 1))
    A, _ = system(orc, g, [(0.5, 0.5, 0.5, 0.3)], rho2=50.0)
    rng = np.random.default_rng(5)
    p = np.zeros(g.shape_p)
    g.interior(p)[...] = rng.standard_normal((8, 8, 8))
    wrap_ghosts(g, p)
    r = orc.residual_field(g, A, p)
    M = sparse_matrix(g, A)
    r2 = A[7].ravel() - M @ g.interior(p).ravel()
    assert np.allclose(r.ravel(), r2, rtol=0, atol=1e-12 * np.abs(r2).max())
    # and M is symmetric with the constant vector in its kernel
    assert abs(M - M.T).max() == 0.0
    assert np.abs(M @ np.ones(M.shape[0])).max() < 1e-10 * abs(M).max()


def test_rb_sor_colour_order(orc):
    # with beta = 1 one RB sweep leaves every black (odd) cell's equation
    # satisfied exactly; the red cells in general not
    g = orc.OrcGrid.make(8)
    A, _ = system(orc, g, [(0.5, 0.5, 0.5, 0.25)], rho2=10.0)
    rng = np.random.default_rng(1)
    p = np.zeros(g.shape_p)
    g.interior(p)[...] = rng.standard_normal((8, 8, 8))
    orc.redblack_relax(g, A, p, 1.0)
    r = orc.residual_field(g, A, p)
    k, j, i = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    black = (i + j + k) % 2 == 1
    scale = np.abs(A[6] * g.interior(p)).max()
    assert np.abs(r[black]).max() < 1e-12 * scale
    assert np.abs(r[~black]).max() > 1e-6 * scale


def test_rb_sor_fixed_point_is_solution(orc):
    g = orc.OrcGrid.make(8)
    A, _ = system(orc, g, [(0.5, 0.5, 0.5, 0.25)], rho2=10.0)
    x = exact_solution(g, A)
    p = np.zeros(g.shape_p)
    g.interior(p)[...] = x
    q = p.copy()
    orc.redblack_relax(g, A, q, 1.2)
    assert np.abs(q - p).max() < 1e-11 * np.abs(x).max()


def test_transfer_operators_on_constants_and_sums(orc):
    nf = [8, 4, 6]
    rng = np.random.default_rng(2)
    Af = rng.standard_normal((8, nf[2], nf[1], nf[0]))
    Ac = orc.fill_coefficients(nf, Af)
    # the 8-child average preserves the total / 8 and maps constants to constants
    assert np.allclose(Ac.sum(axis=(1, 2, 3)), Af.sum(axis=(1, 2, 3)) / 8, rtol=1e-13)
    Ac2 = orc.fill_coefficients(nf, np.full_like(Af, 3.25))
    assert np.all(Ac2 == 3.25)
    Acr = np.zeros_like(Ac)
    orc.coarse_from_fine(nf, Af, Acr)
    assert np.allclose(Acr[7], Ac[7], rtol=1e-14)
    # piecewise-constant interpolation: every child equals its parent
    ng = 2
    ec = np.zeros((nf[2] // 2 + 4, nf[1] // 2 + 4, nf[0] // 2 + 4))
    ec[2:-2, 2:-2, 2:-2] = rng.standard_normal((nf[2] // 2, nf[1] // 2, nf[0] // 2))
    ef = np.zeros((nf[2] + 4, nf[1] + 4, nf[0] + 4))
    orc.coarse_fine_interp(nf, ng, 0, ef, ec)
    inner = ef[2:-2, 2:-2, 2:-2]
    par = np.repeat(np.repeat(np.repeat(ec[2:-2, 2:-2, 2:-2], 2, 0), 2, 1), 2, 2)
    assert np.array_equal(inner, par)
    # sym_MG with a zeroed fine array (as the reference calls it) gives 0
    ef2 = np.zeros_like(ef)
    orc.coarse_fine_interp(nf, ng, 1, ef2, ec)
    assert np.all(ef2 == 0.0)


@pytest.mark.parametrize("bc,rho2", [((0,) * 6, 10.0), ((1, 0, 0, 1, 0, 0), 1000.0),
                                     ((1,) * 6, 100.0)])
def test_newsolver_mg_converges_to_direct_solution(orc, bc, rho2):
    g = orc.OrcGrid.make(16, bc=bc)
    A, _ = system(orc, g, [(0.5, 0.5, 0.5, 0.25), (0.2, 0.8, 0.3, 0.12)], rho2=rho2, seed=4)
    x = exact_solution(g, A)
    p = np.zeros(g.shape_p)
    # the reference V-cycle contracts slowly at high density ratio (its
    # child-averaged coarse operators are 4x too stiff), so give it room
    rc, it, t2, hist = orc.newsolver_mg(g, A, p, 1e-300, 1.2, 600)
    assert rc == 0 and it == 601
    got = _mf(g.interior(p))
    assert np.abs(got - _mf(x)).max() <= 1e-10 * np.abs(x).max()
    # the residual history is that of a convergent iteration, down to the
    # FP64 floor
    assert hist[-1] < 1e-11 * hist[0]


def test_newsolver_mg_iteration_semantics(orc):
    g = orc.OrcGrid.make(16)
    A, _ = system(orc, g, [(0.5, 0.5, 0.5, 0.25)], rho2=10.0)
    p = np.zeros(g.shape_p)
    rc, it, t2, hist = orc.newsolver_mg(g, A, p, 1e-6, 1.2, 100)
    # it is the first iteration whose tres2 < maxError; tres2 is the RMS
    assert rc == 0 and t2 < 1e-6 and it == len(hist) and np.all(hist[:-1] >= 1e-6)
    r = orc.residual_field(g, A, p)
    assert np.isclose(t2, np.sqrt((r ** 2).mean()), rtol=1e-12)
    # not converged: it = maxit + 1
    p = np.zeros(g.shape_p)
    rc, it, t2, hist = orc.newsolver_mg(g, A, p, 1e-300, 1.2, 3)
    assert it == 4 and len(hist) == 3
    # bad size
    g2 = orc.OrcGrid.make(12)
    A2 = np.ones((8, 12, 12, 12))
    assert orc.newsolver_mg(g2, A2, np.zeros(g2.shape_p), 1e-3)[0] == orc.ERR_BAD_SIZE


def test_newsolver_std_converges_and_catches_nan(orc):
    g = orc.OrcGrid.make(8)
    A, _ = system(orc, g, [(0.5, 0.5, 0.5, 0.25)], rho2=10.0)
    x = exact_solution(g, A)
    p = np.zeros(g.shape_p)
    rc, it, t2, hist = orc.newsolver_std(g, A, p, 1e-11 * np.sqrt((A[7] ** 2).mean()), 1.7, 20000)
    assert rc == 0 and it <= 20000
    assert np.abs(_mf(g.interior(p)) - _mf(x)).max() < 1e-8 * np.abs(x).max()
    B = A.copy(); B[2, 1, 1, 1] = np.nan
    p = np.zeros(g.shape_p)
    assert orc.newsolver_std(g, B, p, 1e-6, 1.2, 10)[0] == orc.ERR_DIVERGED


def test_discrete_eigenfunction_closed_form(orc):
    # constant density, periodic box: sin(2 pi x) sin(2 pi y) sin(2 pi z) at
    # cell centres is an eigenvector of the 7-point operator with eigenvalue
    # sum_d 4 sin^2(pi h_d) / h_d^2, so the exact discrete solution is known
    g = orc.OrcGrid.make(16, bc=(1,) * 6)
    rho = np.ones(g.shape_p)
    A = orc.setup_testmg(g, rho)
    x = orc.coords(16, 2, 1.0)[0][2:18]
    S = np.sin(2 * np.pi * x)[None, None, :] * np.sin(2 * np.pi * x)[None, :, None] * \
        np.sin(2 * np.pi * x)[:, None, None]
    A[7] = S
    h = 1.0 / 16
    lam = 3 * 4 * np.sin(np.pi * h) ** 2 / h ** 2
    p = np.zeros(g.shape_p)
    orc.newsolver_mg(g, A, p, 1e-300, 1.0, 60)
    assert np.abs(_mf(g.interior(p)) - S / lam).max() < 1e-13 / lam * 1e2


def test_reference_poisson_mg_known_answer(orc):
    # Tests/PoissonMG: 32x16x16 periodic, XLENGTH=2, BETA=1, MAXIT=20,
    # MAXERROR=1d-300 -> last residual < 1e-8 and L1 error vs the sine < 0.1
    # (run.sh:15).  One rank here (the reference runs two).
    g = orc.OrcGrid.make(32, 16, 16, bc=(1,) * 6, lengths=(2.0, 1.0, 1.0))
    rho = np.ones(g.shape_p)
    A = orc.setup_testmg(g, rho)
    p = np.zeros(g.shape_p)
    rc, it, t2, hist = orc.newsolver_mg(g, A, p, 1e-300, 1.0, 20)
    assert rc == 0 and it == 21
    assert hist[-1] < 1e-8
    err = orc.get_mgtest_err(g, p)
    assert err < 0.1
    # and it is far inside the threshold: the discrete solution of the
    # continuous-sine right-hand side differs from the sine by O(h^2)
    assert err < 0.02


def test_project_velocity_makes_divergence_free(orc):
    g = orc.OrcGrid.make(16, bc=(1, 0, 0, 1, 0, 0))
    A, f = system(orc, g, [(0.5, 0.5, 0.5, 0.25)], rho2=100.0, remove_mean=False)
    x = exact_solution(g, A)
    p = np.zeros(g.shape_p)
    g.interior(p)[...] = x
    wrap_ghosts(g, p)
    u, v, w = f["u"].copy(), f["v"].copy(), f["w"].copy()
    orc.project_velocity(g, u, v, w, f["um"], f["vm"], f["wm"], f["rho"], p, 1.0)
    wrap_ghosts(g, u)
    ng = 2
    div = ((u[ng:-ng, ng:-ng, ng:-ng] - u[ng:-ng, ng:-ng, ng - 1:-ng - 1]) +
           (v[ng:-ng, ng:-ng, ng:-ng] - v[ng:-ng, ng - 1:-ng - 1, ng:-ng]) +
           (w[ng:-ng, ng:-ng, ng:-ng] - w[ng - 1:-ng - 1, ng:-ng, ng:-ng])) * 16
    div0 = np.abs(A[7]).max()
    assert np.abs(div).max() < 1e-9 * div0


def test_calc_residuals_closed_forms(orc):
    g = orc.OrcGrid.make(8)
    A, _ = system(orc, g, [(0.5, 0.5, 0.5, 0.25)])
    p = np.zeros(g.shape_p)
    L1, L2, Li = orc.calc_residuals(g, A, p)
    assert np.isclose(L1, np.abs(A[7]).mean(), rtol=1e-13)
    assert np.isclose(L2, np.sqrt((A[7] ** 2).mean()), rtol=1e-13)
    assert Li == np.abs(A[7]).max()


def test_pcg_matches_direct(orc):
    g = orc.OrcGrid.make(16, bc=(1, 0, 0, 1, 0, 0))
    A, _ = system(orc, g, [(0.4, 0.5, 0.5, 0.2)], rho2=1000.0)
    x = exact_solution(g, A)
    p = np.zeros(g.shape_p)
    its, rel = orc.pcg_solve(g, A, p, 1e-14, 100000)
    assert rel < 1e-13
    assert np.abs(_mf(g.interior(p)) - _mf(x)).max() < 1e-10 * np.abs(x).max()


def test_newsolver_mg_threads_do_not_change_the_algorithm(orc):
    """The OpenMP loops of the restatement (used to time the reference
    algorithm on all host cores) update independent cells only: the iterate
    is bit-identical to the serial run, the residual history equal to
    summation-order rounding."""
    g = orc.OrcGrid.make(32, bc=(1, 0, 0, 1, 0, 0))
    A, _ = system(orc, g, [(0.3, 0.5, 0.5, 0.2)], rho2=1000.0, seed=4)
    out = []
    for nt in (1, max(2, orc.max_threads())):
        orc.set_threads(nt)
        p = np.zeros(g.shape_p)
        rc, it, t2, hist = orc.newsolver_mg(g, A, p, 1e-300, 1.2, 6)
        out.append((p, hist))
    orc.set_threads(1)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.allclose(out[0][1], out[1][1], rtol=1e-12, atol=0)


# --- FreeSolver (freesurface.f90:700-855) -----------------------------------

def _fs_mask(g, rng, values):
    m = np.zeros(g.shape_p, dtype=np.int32)
    g.interior(m)[...] = rng.choice(values, size=(g.nz, g.ny, g.nx)).astype(np.int32)
    return m


def test_free_solver_with_every_cell_active_is_newsolver_std(orc):
    """solver_flag 1 and pcmask == 0 everywhere: the masked sweep is
    RedBlackRelax itself, so from the same (ghost-wrapped) start a fixed
    number of sweeps (maxError 0 never stops either solver) gives
    NewSolver_std's p bit for bit."""
    g = orc.OrcGrid.make(12, bc=(1, 0, 0, 1, 0, 0))
    A, _ = system(orc, g, [(0.4, 0.5, 0.6, 0.25)], rho2=100.0)
    rng = np.random.default_rng(3)
    p0 = np.zeros(g.shape_p)
    g.interior(p0)[...] = rng.standard_normal((g.nz, g.ny, g.nx))
    orc.apply_bc(g, p0)   # FreeSolver wraps the ghosts first (freesurface.f90:727)
    p1, p2 = p0.copy(), p0.copy()
    m = np.zeros(g.shape_p, dtype=np.int32)
    rc1, it1, _, _ = orc.free_solver(g, A, p1, m, np.zeros(g.shape_p), 1, 0.0, 1.3, 7)
    rc2, it2, _, _ = orc.newsolver_std(g, A, p2, 0.0, 1.3, 7)
    assert rc1 == rc2 == 0 and it1 == it2 == 8
    assert np.array_equal(p1, p2)


def test_free_solver_dirichlet_planes_closed_form(orc):
    """Uniform density, walls, no source; the planes i = 0 and i = nx-1 are
    gas cells (pcmask 1) holding P_gas = 0 and 1.  The discrete solution on
    the liquid cells is exactly linear in i, p = i / (nx - 1); gas cells keep
    P_gas, and the stop is on the max-norm residual."""
    g = orc.OrcGrid.make(12, 6, 6)
    A, _ = system(orc, g, [], rho2=1.0, remove_mean=False)
    A[7] = 0.0
    m = np.zeros(g.shape_p, dtype=np.int32)
    ng = g.ng
    m[ng:ng + g.nz, ng:ng + g.ny, ng] = 1
    m[ng:ng + g.nz, ng:ng + g.ny, ng + g.nx - 1] = 1
    pgas = np.zeros(g.shape_p)
    pgas[..., ng + g.nx - 1] = 1.0
    p = np.zeros(g.shape_p)
    tol = 1e-12 * A[6].max()
    rc, it, t2, hist = orc.free_solver(g, A, p, m, pgas, 1, tol, 1.7, 20000)
    assert rc == 0 and it < 20000
    x = g.interior(p)
    lin = np.arange(g.nx) / (g.nx - 1)
    assert np.abs(x - lin[None, None, :]).max() < 1e-9
    assert np.all(x[..., 0] == 0.0) and np.all(x[..., -1] == 1.0)
    r = orc.residual_field(g, A, p)[..., 1:-1]
    assert np.abs(r).max() < tol
    assert t2 == pytest.approx(np.sqrt((r ** 2).sum()) / r.size, rel=1e-8)


def test_free_solver_flag2_mask_rules(orc):
    """solver_flag 2 updates pcmask 1 and 2 cells, zeroes pcmask 3 cells and
    leaves pcmask 0 cells (and P_gas) alone; converged active cells satisfy
    their equations."""
    g = orc.OrcGrid.make(10, bc=(1, 0, 0, 1, 0, 0))
    A, _ = system(orc, g, [(0.5, 0.4, 0.5, 0.3)], rho2=50.0)
    rng = np.random.default_rng(9)
    m = _fs_mask(g, rng, [0, 1, 2, 3])
    p = np.zeros(g.shape_p)
    g.interior(p)[...] = rng.standard_normal((g.nz, g.ny, g.nx))
    p_in = p.copy()
    scale = np.abs(A[6]).max()
    rc, it, t2, hist = orc.free_solver(g, A, p, m, np.full(g.shape_p, 7.0), 2, 1e-10 * scale, 1.2, 20000)
    assert rc == 0 and it < 20000
    mi, x, x0 = g.interior(m), g.interior(p), g.interior(p_in)
    assert np.all(x[mi == 3] == 0.0)
    assert np.array_equal(x[mi == 0], x0[mi == 0])
    act = (mi == 1) | (mi == 2)
    r = orc.residual_field(g, A, p)
    assert np.abs(r[act]).max() < 1e-10 * scale
    assert t2 == pytest.approx(np.sqrt((r[act] ** 2).sum()) / act.sum(), rel=1e-8)
    assert len(hist) == it and hist[-1] == t2


def test_free_solver_rejects_flag_and_nan(orc):
    g = orc.OrcGrid.make(8)
    A, _ = system(orc, g, [(0.5, 0.5, 0.5, 0.25)], rho2=10.0)
    m = np.zeros(g.shape_p, dtype=np.int32)
    z = np.zeros(g.shape_p)
    assert orc.free_solver(g, A, z.copy(), m, z, 0, 1e-8)[0] == orc.ERR_UNSUPPORTED
    B = A.copy(); B[3, 2, 2, 2] = np.nan
    assert orc.free_solver(g, B, z.copy(), m, z, 1, 1e-8)[0] == orc.ERR_DIVERGED
