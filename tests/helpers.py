"""CPU-side problem builders for the oracle tests (numpy only).

These make inputs, not answers: the droplet density and random face
velocities are fed to the oracle's own SetupPoisson restatement.
"""
import numpy as np


def droplet_fields(g, drops, rho1=1.0, rho2=10.0, width=1.5, seed=0, orc=None):
    """Padded rho, u, v, w for an OrcGrid (walls or periodic per g.bc)."""
    X, Y, Z = g.shape_p[2], g.shape_p[1], g.shape_p[0]
    ng = g.ng
    x = orc.coords(g.nx, ng, g.xlength)[0]
    y = orc.coords(g.ny, ng, g.ylength)[0]
    z = orc.coords(g.nz, ng, g.zlength)[0]
    h = min(g.xlength / g.nx, g.ylength / g.ny, g.zlength / g.nz)
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    c = np.zeros((Z, Y, X))
    per = [g.bc[0] == 1, g.bc[1] == 1, g.bc[2] == 1]
    L = [g.xlength, g.ylength, g.zlength]
    for (cx, cy, cz, R) in drops:
        d = []
        for q, (co, cc) in enumerate(((xx, cx), (yy, cy), (zz, cz))):
            dd = co - cc
            if per[q]:
                dd = dd - L[q] * np.rint(dd / L[q])
            d.append(dd)
        s = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2) - R
        c += 0.5 * (1.0 - np.tanh(s / (width * h)))
    c = np.minimum(c, 1.0)
    rho = c * (rho2 - rho1) + rho1
    rng = np.random.default_rng(seed)
    u, v, w = (rng.uniform(-1.0, 1.0, (Z, Y, X)) for _ in range(3))
    # periodic fields: make the ghosts exact copies of the wrapped interior
    for f in (rho, u, v, w):
        wrap_ghosts(g, f)
    # zero normal velocity on wall faces (face index -1 and n-1)
    sl = slice(None)
    if not per[0]:
        u[:, :, ng - 1] = 0.0
        u[:, :, ng + g.nx - 1] = 0.0
    if not per[1]:
        v[:, ng - 1, :] = 0.0
        v[:, ng + g.ny - 1, :] = 0.0
    if not per[2]:
        w[ng - 1, :, :] = 0.0
        w[ng + g.nz - 1, :, :] = 0.0
    del sl
    return rho, u, v, w


def wrap_ghosts(g, f):
    """All ghost layers of a padded array from the wrapped interior in
    periodic directions (x, then y, then z)."""
    ng = g.ng
    n = [g.nx, g.ny, g.nz]
    for d in range(3):
        if g.bc[d] != 1:
            continue
        ax = 2 - d
        f2 = np.moveaxis(f, ax, 0)
        for q in range(ng):
            f2[q] = f2[n[d] + q]
            f2[n[d] + ng + q] = f2[ng + q]
    return f


def system(orc, g, drops, rho2=10.0, seed=0, remove_mean=True):
    """8-slot A from the oracle's SetupPoisson on a droplet field."""
    rho, u, v, w = droplet_fields(g, drops, rho2=rho2, seed=seed, orc=orc)
    um, vm, wm = orc.pressure_masks(g)
    rc, A = orc.setup_poisson(g, u, v, w, um, vm, wm, rho, 1.0, 0.0)
    assert rc == 0, rc
    if remove_mean:
        A[7] -= A[7].mean()
    return A, dict(rho=rho, u=u, v=v, w=w, um=um, vm=vm, wm=wm)


def sparse_matrix(g, A):
    """The 8-slot system as a scipy CSR matrix M (M p = b) on the interior,
    periodic couplings wrapped, wall couplings (to fixed ghosts) dropped."""
    import scipy.sparse as sp
    nx, ny, nz = g.nx, g.ny, g.nz
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [A[6].ravel()]
    shifts = [(0, -1, 2, 0), (1, +1, 2, 0), (2, -1, 1, 1), (3, +1, 1, 1), (4, -1, 0, 2), (5, +1, 0, 2)]
    for slot, off, ax, d in shifts:
        nb = np.roll(idx, -off, axis=ax)
        a = A[slot].copy()
        if g.bc[d] != 1:
            # couplings through the wall reach a ghost: drop them
            edge = [slice(None)] * 3
            edge[ax] = 0 if off < 0 else -1
            a[tuple(edge)] = 0.0
        rows.append(idx.ravel())
        cols.append(nb.ravel())
        vals.append(-a.ravel())
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return M


def exact_solution(g, A):
    """Mean-free solution of the singular Neumann/periodic system by a
    bordered direct solve (scipy)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    M = sparse_matrix(g, A)
    n = M.shape[0]
    one = sp.csr_matrix(np.ones((n, 1)))
    K = sp.bmat([[M, one], [one.T, None]], format="csc")
    rhs = np.concatenate([A[7].ravel(), [0.0]])
    x = spla.spsolve(K, rhs)
    return x[:n].reshape(g.nz, g.ny, g.nx)


# ---------------------------------------------------------------------------
# adjoint helpers (torch tensors on the solver's device)
# ---------------------------------------------------------------------------
def faces_from_rho(grid, rho):
    """c_f = 2/(rho_L+rho_R) on every +x/+y/+z face; closed (0) on walls."""
    import torch
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


def sine_target(grid, seed=3, device="cuda"):
    """p_target: 8 random-amplitude sine modes, wavenumbers 1..4."""
    import torch
    rng = np.random.default_rng(seed)
    ax = [(torch.arange(m, dtype=torch.float64, device=device) + 0.5) / m
          for m in (grid.nz, grid.ny, grid.nx)]
    z = ax[0].view(-1, 1, 1)
    y = ax[1].view(1, -1, 1)
    x = ax[2].view(1, 1, -1)
    t = torch.zeros((grid.nz, grid.ny, grid.nx), dtype=torch.float64, device=device)
    for _ in range(8):
        k = rng.integers(1, 5, size=3)
        a = rng.uniform(-1, 1)
        ph = rng.uniform(0, 2 * np.pi)
        t += a * torch.sin(2 * np.pi * (int(k[0]) * x + int(k[1]) * y + int(k[2]) * z) + ph)
    return t


def objective(p, pt):
    """J = 0.5 ||(p - mean p) - p_target||^2."""
    p0 = p - p.mean()
    return 0.5 * ((p0 - pt) ** 2).sum()


def objective_diff(pp, pm, pt):
    """J(p+) - J(p-) evaluated difference-first, so its rounding is relative
    to the difference and not to J."""
    d = pp - pm
    d0 = d - d.mean()
    sm = 0.5 * (pp + pm)
    s0 = sm - sm.mean()
    return float((d0 * (s0 - pt)).sum())


def random_faces(grid, count, seed):
    """Random (direction, k, j, i) faces, skipping closed wall faces."""
    rng = np.random.default_rng(seed)
    per = grid.periodic()
    faces = []
    while len(faces) < count:
        d = int(rng.integers(0, 3))
        k, j, i = (int(rng.integers(0, m)) for m in (grid.nz, grid.ny, grid.nx))
        last = (i == grid.nx - 1, j == grid.ny - 1, k == grid.nz - 1)[d]
        if last and not per[d]:
            continue
        faces.append((d, k, j, i))
    return faces
