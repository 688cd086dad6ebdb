"""Further pins of the oracle by closed forms (CPU only).

* The sequencing of MG_iteration / NewSolverMG (MGsolver.f90:253-405): on a
  diagonal system (every coupling zero, A7 = d > 0) each red-black SOR sweep
  acts cell by cell, so k sweeps from zero of d e = r give
  e = a r / d with a = 1 - (1 - beta)^k and leave the residual c r with
  c = 1 - a.  The whole V-cycle then has a closed form in a, c, the 8-child
  mean R and the injection P, which fixes the sweep counts, the one extra
  bottom sweep on level 2 against the post-smoothing residual, the
  residual-correction order of the up leg and the stale level-2 residual the
  next cycle starts from when the hierarchy has only two levels (beta = 1
  would solve the diagonal system in one sweep and hide all of this, so
  beta is 0.8 or 1.2).
* Poisson_BCs codes 3 / 4 (modules.f90:2755-2892): on a flux face the outward
  coupling is 0, A7 is the sum of the remaining couplings, and every other
  entry equals the wall assembly of the same fields.
"""
import numpy as np
import pytest

from helpers import droplet_fields


def _R(x):
    """8-child mean of a (nz, ny, nx) field."""
    z, y, w = x.shape
    return x.reshape(z // 2, 2, y // 2, 2, w // 2, 2).mean(axis=(1, 3, 5))


def _P(x):
    """Piecewise-constant injection to the 8 children."""
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def _diag_vcycle(r, d, lvl, beta, k):
    """Correction of one MG_iteration on the diagonal system at level `lvl`
    (lvl = log2 n) for right-hand side r; returns (correction, residual left
    in slot 8)."""
    c = (1.0 - beta) ** k
    a = 1.0 - c
    if lvl == 2:
        # nrelax sweeps from zero, residual c r, then one sweep of d e = c r
        # starting from e = a r / d
        q = (1.0 - beta) * a + beta * c
        return q * r / d, c * r
    s = c * r
    ec, _ = _diag_vcycle(_R(s), _R(d), lvl - 1, beta, k)
    E = _P(ec)
    t = s - d * E
    corr = a * r / d + E + a * t / d
    return corr, c * t


@pytest.mark.parametrize("n", [4, 8, 16])
@pytest.mark.parametrize("nrelax", [1, 2, 3])
@pytest.mark.parametrize("beta", [0.8, 1.2])
def test_newsolver_mg_sequencing_on_a_diagonal_system(orc, n, nrelax, beta):
    lvl = int(np.log2(n))
    g = orc.OrcGrid.make(n)
    rng = np.random.default_rng(n + 10 * nrelax)
    A = np.zeros(g.shape_A)
    d = rng.uniform(1.0, 3.0, (n, n, n))
    A[6] = d
    A[7] = rng.standard_normal((n, n, n))
    p0 = np.zeros(g.shape_p)
    g.interior(p0)[...] = rng.standard_normal((n, n, n))
    # closed form over three cycles
    p = g.interior(p0).copy()
    slot8 = A[7] - d * p
    want, hist = [], []
    for _ in range(3):
        corr, slot8 = _diag_vcycle(slot8, d, lvl, beta, nrelax)
        p = p + corr
        if lvl > 2:
            # the up leg leaves the true residual of the new iterate
            assert np.allclose(slot8, A[7] - d * p, rtol=0, atol=1e-13 * np.abs(A[7]).max())
        want.append(p.copy())
        hist.append(np.sqrt(((A[7] - d * p) ** 2).mean()))
    for m in (1, 2, 3):
        ph = p0.copy()
        rc, it, t2, h = orc.newsolver_mg(g, A, ph, 1e-300, beta, m, nrelax=nrelax)
        assert rc == 0 and it == m + 1
        got = g.interior(ph)
        assert np.allclose(got, want[m - 1], rtol=0, atol=1e-13 * np.abs(want[m - 1]).max())
        r0 = np.sqrt(((A[7] - d * g.interior(p0)) ** 2).mean())
        assert np.allclose(h, hist[:m], rtol=1e-10, atol=1e-14 * r0)
    if lvl == 2:
        # two levels: the next cycle restarts from the post-smoothing
        # residual, not the residual of the corrected iterate
        c = (1.0 - beta) ** nrelax
        q = (1.0 - beta) * (1.0 - c) + beta * c
        r0 = A[7] - d * g.interior(p0)
        assert np.allclose(want[1], g.interior(p0) + q * r0 / d * (1.0 + c), rtol=0,
                           atol=1e-13 * np.abs(want[1]).max())


def _assemble(orc, g, seed=3):
    rho, u, v, w = droplet_fields(g, [(0.4, 0.5, 0.6, 0.25)], rho2=100.0, seed=seed, orc=orc)
    um, vm, wm = orc.pressure_masks(g)
    rc, A = orc.setup_poisson(g, u, v, w, um, vm, wm, rho, 1.0, 0.0)
    assert rc == 0
    return A


# flux faces: x- takes code 3, x+ codes 3 and 4, the other faces code 3
# (modules.f90:2768-2872); code 4 elsewhere is not a flux condition
FLUX = {0: (3,), 3: (3, 4), 1: (3,), 4: (3,), 2: (3,), 5: (3,)}
# (slot of the outward coupling, axis of the numpy array, index of the face cells)
FACE = {0: (0, 2, 0), 3: (1, 2, -1), 1: (2, 1, 0), 4: (3, 1, -1), 2: (4, 0, 0), 5: (5, 0, -1)}


@pytest.mark.parametrize("bc", [(3, 0, 0, 4, 0, 0), (3, 0, 0, 3, 0, 0), (0, 3, 3, 0, 3, 3),
                                (3, 3, 3, 4, 3, 3), (4, 0, 0, 0, 4, 4)])
def test_poisson_bcs_flux_faces(orc, bc):
    n = 8
    g = orc.OrcGrid.make(n, bc=bc)
    g0 = orc.OrcGrid.make(n, bc=(0,) * 6)
    A = _assemble(orc, g)
    W = _assemble(orc, g0)
    other = np.ones((n, n, n), dtype=bool)
    for f in range(6):
        slot, ax, idx = FACE[f]
        sl = [slice(None)] * 3
        sl[ax] = idx
        sl = tuple(sl)
        if bc[f] in FLUX[f]:
            # the outward coupling dropped, A7 the sum of the rest
            assert np.all(A[slot][sl] == 0.0)
            rest = sum(A[s][sl] for s in range(6))
            assert np.allclose(A[6][sl], rest, rtol=4e-16 * 6, atol=0)
            # every other entry of the face rows as on a wall
            for s in list(range(6)) + [7]:
                if s != slot:
                    assert np.array_equal(A[s][sl], W[s][sl]), (f, s)
            assert np.allclose(A[6][sl], W[6][sl], rtol=4e-16 * 6, atol=0)
            other[sl] = False
        elif bc[f] == 4:
            # not a flux face (and not a wall): the boundary coupling stays
            assert np.all(A[slot][sl] > 0.0)
            other[sl] = False
    # rows off the flux faces: bit for bit the wall assembly
    for s in range(8):
        assert np.array_equal(A[s][other], W[s][other]), s
    # interior faces stay symmetric
    assert np.array_equal(A[1][:, :, :-1], A[0][:, :, 1:])
    assert np.array_equal(A[3][:, :-1], A[2][:, 1:])
    assert np.array_equal(A[5][:-1], A[4][1:])


# ---------------------------------------------------------------------------
# Poisson_BCs codes 5 / 6 (pressure conditions): closed forms on uniform
# density with zero velocities, where the pressure solution is known
# ---------------------------------------------------------------------------
def _uniform_system(orc, n, bc, bp, u_of_x=None):
    g = orc.OrcGrid.make(n, bc=bc)
    rho = np.ones(g.shape_p)
    u = np.zeros(g.shape_p)
    if u_of_x is not None:
        xh = orc.coords(n, 2, 1.0)[1]
        u[...] = u_of_x(xh)[None, None, :]
    v = np.zeros(g.shape_p)
    w = np.zeros(g.shape_p)
    um, vm, wm = orc.pressure_masks(g)
    rc, A = orc.setup_poisson(g, u, v, w, um, vm, wm, rho, 1.0, 0.0, bp=bp)
    assert rc == 0
    return g, A


def _direct(g, A):
    import scipy.sparse.linalg as spla
    from helpers import sparse_matrix
    M = sparse_matrix(g, A).tocsc()
    return spla.spsolve(M, A[7].ravel()).reshape(g.nz, g.ny, g.nx)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_pressure_code5_gives_the_linear_profile(orc, axis):
    """Code 5 on both faces of one axis, walls elsewhere: the solution is
    the straight line through the two boundary pressures at the faces (each
    face row is P_0 = P_inner / 3 + 2/3 P_b, the linear extrapolation)."""
    n = 8
    bc = [0] * 6
    bc[axis] = bc[axis + 3] = 5
    bp = np.zeros(6)
    bp[2 * axis], bp[2 * axis + 1] = 1.5, -0.5
    g, A = _uniform_system(orc, n, bc, bp)
    p = _direct(g, A)
    xc = (np.arange(n) + 0.5) / n
    line = bp[2 * axis] + (bp[2 * axis + 1] - bp[2 * axis]) * xc
    shape = [1, 1, 1]
    shape[2 - axis] = n
    want = np.broadcast_to(line.reshape(shape), (n, n, n))
    assert np.abs(p - want).max() < 1e-12


@pytest.mark.parametrize("bc", [(5, 5, 5, 0, 0, 0), (0, 0, 0, 5, 5, 0), (5, 5, 0, 5, 5, 0)])
def test_pressure_code5_corners_keep_a_constant(orc, bc):
    """Corner cells of several code-5 faces add the conditions of each face:
    with one boundary pressure P on every such face, p = P solves the
    system exactly."""
    P = 0.75
    g, A = _uniform_system(orc, 8, bc, np.full(6, P))
    p = _direct(g, A)
    assert np.abs(p - P).max() < 1e-12
    # the corner rows themselves
    r = A[7] - A[6] * P + sum(A[s] for s in range(6)) * P
    assert np.abs(r).max() < 1e-13


def test_pressure_code5_zplus_writes_its_corner_couplings_into_the_first_plane(orc):
    """The z+ block of code 5 assigns the y-couplings of its four corner cells
    at the ks plane (modules.f90:2884-2887): with walls on y those entries
    become 0 there, and nowhere else does the first plane differ from the
    wall assembly."""
    n = 8
    g, A = _uniform_system(orc, n, (0, 0, 0, 0, 0, 5), np.full(6, 0.3))
    _, W = _uniform_system(orc, n, (0,) * 6, None)
    first = np.ones((n, n), dtype=bool)
    for (j, i, s) in ((0, 0, 3), (0, n - 1, 3), (n - 1, 0, 2), (n - 1, n - 1, 2)):
        assert A[s][0, j, i] == 0.0 and W[s][0, j, i] > 0.0
        first[j, i] = False
    for s in range(8):
        assert np.array_equal(A[s][0][first], W[s][0][first])
    # the last plane is the code-5 row: A7 = 1, A6 (inward z) = 1/3
    assert np.all(A[6][-1] == 1.0) and np.all(A[4][-1] == 1.0 / 3.0)
    assert np.all(A[7][-1] == 2.0 / 3.0 * 0.3)


def test_pressure_code6_rows(orc):
    """Code 6 on x+ (p and du/dn at the face): the face rows are satisfied
    by P_b + alpha (x - 1)^2 for any alpha (value P_b and zero slope at the
    face), the system with zero velocity is solved by p = P_b, and the face
    rows' source does not see du/dx."""
    n = 8
    P = -1.25
    bp = np.zeros(6)
    bp[1] = P
    g, A = _uniform_system(orc, n, (0, 0, 0, 6, 0, 0), bp)
    p = _direct(g, A)
    assert np.abs(p - P).max() < 1e-12
    xc = (np.arange(n) + 0.5) / n
    for alpha in (1.0, -3.0):
        q = np.broadcast_to(P + alpha * (xc - 1.0) ** 2, (n, n, n))
        # residual of the face rows only (x- neighbour, the face itself)
        r = A[7][..., -1] - A[6][..., -1] * q[..., -1] + A[0][..., -1] * q[..., -2] + \
            (A[2][..., -1] + A[3][..., -1] + A[4][..., -1] + A[5][..., -1]) * q[..., -1]
        # y / z walls couple nothing outside, interior y / z neighbours equal q
        assert np.abs(r).max() < 1e-10 * np.abs(A[6]).max()
    # u = x: du/dx = 1 everywhere, removed from the face rows only
    _, B = _uniform_system(orc, n, (0, 0, 0, 6, 0, 0), bp, u_of_x=lambda xh: xh)
    assert np.allclose(B[7][..., -1], A[7][..., -1], rtol=0, atol=1e-12)
    assert np.allclose(B[7][..., :-1], A[7][..., :-1] - 1.0, rtol=0, atol=1e-12)
