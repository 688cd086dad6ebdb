"""The solve from face values (pps_solve_faces_device): the rows of the
system that pps_faces_to_A_device forms, taken from the face values inside
the fast solve's setup pass, so that A is never written.  Against
faces_to_A followed by pps_solve_device: p with its ghosts, the iteration
count and tres2 bit for bit; the adjoint from face values
(pps_adjoint_faces_device) against pps_adjoint_device on that A, on the
hierarchy of the solve and set up afresh."""
import numpy as np
import pytest

from helpers import faces_from_rho

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


CASES = [
    ((32, 32, 32), (0,) * 6),
    ((32, 24, 40), (1, 0, 0, 1, 0, 0)),
    ((16, 16, 16), (1,) * 6),
]


def _system(pps, shape, bc, seed):
    grid, A, f = pps.make_problem(n=shape[0], bc=bc, drops=[(0.45, 0.5, 0.55, 0.22)], rho2=1000.0,
                                  seed=seed) if shape[0] == shape[1] == shape[2] else (None, None, None)
    if grid is None:
        grid = pps.Grid.make(*shape, bdry_cond=bc)
        rho, u, v, w = pps.synth_fields(grid, [(0.45, 0.5, 0.55, 0.22)], 1.0, 1000.0, 1.5, seed=seed)
        um, vm, wm = pps.pressure_masks(grid)
        A = pps.setup_poisson(grid, u, v, w, um, vm, wm, rho, 1.0)
        pps.remove_rhs_mean(grid, A)
        f = dict(rho=rho)
    cf = faces_from_rho(grid, f["rho"])
    w = pps.face_weights(grid)
    # weights that differ plane by plane, so that a transposed or shifted
    # weight shows
    gen = torch.Generator(device="cuda").manual_seed(seed)
    w = tuple(wi * (1.0 + 0.25 * torch.rand(wi.shape, generator=gen, device="cuda", dtype=torch.float64))
              for wi in w)
    b = A[7].clone()
    return grid, cf, w, b


@pytest.mark.parametrize("shape,bc", CASES)
@pytest.mark.parametrize("method", ["mixed", "fp64"])
def test_solve_faces_equals_faces_to_A_then_solve(pps, shape, bc, method):
    _say("solve_faces", method, shape, bc)
    grid, cf, w, b = _system(pps, shape, bc, seed=4)
    A = pps.faces_to_A(grid, cf, w, b)
    brms = float(torch.sqrt((b ** 2).mean()))
    tol = 1e-10 * brms
    with pps.Solver(grid, method=method) as s:
        p_ref = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
        it_ref, t2_ref = s.solve(A, p_ref, tol, 1.2, 60, allow_maxit=False)
        p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
        it, t2 = s.solve_faces(cf, w, b, p, tol, 60, allow_maxit=False)
    _say(f"   it {it} / {it_ref}, tres2/rms(b) {t2 / brms:.2e}")
    assert it == it_ref and t2 == t2_ref and t2 < tol
    assert torch.equal(p, p_ref)


@pytest.mark.parametrize("shape,bc", [CASES[0], CASES[1]])
@pytest.mark.parametrize("method", ["mixed", "fp64"])
def test_adjoint_faces_equals_adjoint_on_A(pps, shape, bc, method):
    _say("adjoint_faces", method, shape, bc)
    grid, cf, w, b = _system(pps, shape, bc, seed=6)
    A = pps.faces_to_A(grid, cf, w, b)
    gen = torch.Generator(device="cuda").manual_seed(9)
    g = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
    grid.interior(g).copy_(torch.randn(grid.shape_A[1:], generator=gen, device="cuda", dtype=torch.float64))
    brms = float(torch.sqrt((b ** 2).mean()))
    with pps.Solver(grid, method=method) as s:
        lam_ref = torch.empty_like(g)
        it_ref = s.adjoint(A, g, lam_ref, rtol=1e-11, maxit=60)
        # on the hierarchy of a solve from the same face values
        p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
        s.solve_faces(cf, w, b, p, 1e-10 * brms, 60, allow_maxit=False)
        lam = torch.empty_like(g)
        it = s.adjoint_faces(cf, w, g, lam, rtol=1e-11, maxit=60)
        assert it == it_ref and torch.equal(lam, lam_ref)
        # set up afresh from the face values (the last operator was A)
        s.solve(A, p, 1e-10 * brms, 1.2, 60, allow_maxit=False)
        lam2 = torch.empty_like(g)
        it2 = s.adjoint_faces(cf, w, g, lam2, rtol=1e-11, maxit=60)
    _say(f"   it {it} / {it_ref} / {it2}")
    assert it2 == it_ref and torch.equal(lam2, lam_ref)


def test_adjoint_faces_with_other_weights_sets_up_afresh(pps):
    """The hierarchy of a face-value solve is known by the addresses of the
    face values and of the weights: the same values with other weights is
    another system."""
    _say("adjoint_faces, other weights")
    grid, cf, w, b = _system(pps, (32, 32, 32), (0,) * 6, seed=8)
    w2 = tuple(wi * 1.5 for wi in w)
    A2 = pps.faces_to_A(grid, cf, w2, b)
    gen = torch.Generator(device="cuda").manual_seed(3)
    g = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
    grid.interior(g).copy_(torch.randn(grid.shape_A[1:], generator=gen, device="cuda", dtype=torch.float64))
    brms = float(torch.sqrt((b ** 2).mean()))
    with pps.Solver(grid, method="mixed") as s:
        lam_ref = torch.empty_like(g)
        it_ref = s.adjoint(A2, g, lam_ref, rtol=1e-11, maxit=60)
    with pps.Solver(grid, method="mixed") as s:
        p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
        s.solve_faces(cf, w, b, p, 1e-10 * brms, 60, allow_maxit=False)
        lam = torch.empty_like(g)
        it = s.adjoint_faces(cf, w2, g, lam, rtol=1e-11, maxit=60)
    _say(f"   it {it} / {it_ref}")
    assert it == it_ref and torch.equal(lam, lam_ref)


def test_solve_faces_refusals(pps):
    grid, cf, w, b = _system(pps, (16, 16, 16), (0,) * 6, seed=2)
    p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
    with pps.Solver(grid, method="reference") as s:
        with pytest.raises(pps.PPSError) as e:
            s.solve_faces(cf, w, b, p, 1e-8)
        assert e.value.code == pps.ERR_UNSUPPORTED
    # cell widths weight SetupPoisson's rows; the face-value system is
    # symmetric already and takes none
    with pps.Solver(grid, method="mixed") as s:
        s.set_cell_widths(np.full(16, 1 / 16), np.full(16, 1 / 16), np.full(16, 1 / 16))
        for call in (lambda: s.solve_faces(cf, w, b, p, 1e-8),
                     lambda: s.adjoint_faces(cf, w, p, torch.empty_like(p))):
            with pytest.raises(pps.PPSError) as e:
                call()
            assert e.value.code == pps.ERR_UNSUPPORTED
