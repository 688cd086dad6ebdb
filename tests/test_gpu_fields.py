"""The projection step from the fields (pps_solve_fields_device): SetupPoisson
evaluated inside the fast solve's setup pass, A never formed, against
pps_setup_poisson_device followed by pps_solve_device (and by
pps_solve_project_device with the velocity outputs), bit for bit: p with its
ghosts, the iteration count, tres2 and the corrected velocities."""
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _fields(pps, g, seed, rho2=1000.0):
    rho, u, v, w = pps.synth_fields(g, [(0.4, 0.55, 0.5, 0.22), (0.8, 0.2, 0.3, 0.12)], 1.0, rho2, 1.5,
                                    seed=seed)
    um, vm, wm = pps.pressure_masks(g)
    return dict(u=u, v=v, w=w, um=um, vm=vm, wm=wm, rho=rho)


# walls; periodic x; fully periodic; inflow (3) / outflow (4) codes;
# a non-cubic box
CASES = [
    ((32, 32, 32), (0, 0, 0, 0, 0, 0)),
    ((32, 32, 32), (1, 0, 0, 1, 0, 0)),
    ((16, 16, 16), (1, 1, 1, 1, 1, 1)),
    ((32, 16, 16), (3, 0, 1, 4, 0, 1)),
    ((64, 32, 48), (0, 0, 0, 0, 0, 0)),
]


@pytest.mark.parametrize("shape,bc", CASES)
@pytest.mark.parametrize("method", ["mixed", "fp64"])
def test_solve_fields_equals_assemble_then_solve(pps, shape, bc, method):
    _say("solve_fields", method, shape, bc)
    g = pps.Grid.make(*shape, bdry_cond=bc)
    f = _fields(pps, g, seed=11)
    dt = 0.7
    A = pps.setup_poisson(g, f["u"], f["v"], f["w"], f["um"], f["vm"], f["wm"], f["rho"], dt)
    brms = float(torch.sqrt((A[7] ** 2).mean()))
    tol = 1e-10 * brms
    with pps.Solver(g, method=method) as s:
        p_ref = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it_ref, t2_ref = s.solve(A, p_ref, tol, 1.2, 40)
        p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it, t2 = s.solve_fields(f, p, tol, dt=dt, maxit=40)
    _say(f"   it {it} / {it_ref}  tres2/rms(b) {t2 / brms:.3e}")
    assert it == it_ref and t2 == t2_ref
    assert torch.equal(p, p_ref)
    # with the velocity correction
    with pps.Solver(g, method=method) as s:
        p_ref = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it_ref, t2_ref, uo_r, vo_r, wo_r = s.solve_project(
            A, p_ref, f["u"], f["v"], f["w"], f["um"], f["vm"], f["wm"], f["rho"], dt, tol, 1.2, 40)
        p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it, t2, uo, vo, wo = s.solve_fields(f, p, tol, dt=dt, maxit=40, project=True)
    assert it == it_ref and t2 == t2_ref
    assert torch.equal(p, p_ref)
    assert torch.equal(uo, uo_r) and torch.equal(vo, vo_r) and torch.equal(wo, wo_r)


def test_solve_fields_converges_on_drops256(pps):
    """The 256^3 benchmark field from its fields: the same solution as the
    A-based solve, bit for bit (mixed form)."""
    _say("solve_fields drops256")
    g, A, f = pps.make_problem("drops256")
    # make_problem removed the mean of A8; assemble again without it
    A = pps.setup_poisson(g, f["u"], f["v"], f["w"], f["um"], f["vm"], f["wm"], f["rho"], 1.0)
    brms = float(torch.sqrt((A[7] ** 2).mean()))
    tol = 1e-10 * brms
    with pps.Solver(g, method="mixed") as s:
        p_ref = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it_ref, t2_ref = s.solve(A, p_ref, tol, 1.2, 50, allow_maxit=False)
        del A
        p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it, t2 = s.solve_fields(f, p, tol, maxit=50, allow_maxit=False)
    _say(f"   it {it} / {it_ref}  tres2/rms(b) {t2 / brms:.3e}")
    assert it == it_ref and t2 == t2_ref and t2 < tol
    assert torch.equal(p, p_ref)


def test_solve_fields_refusals(pps):
    g = pps.Grid.make(16, bdry_cond=(5, 0, 0, 5, 0, 0))
    f = _fields(pps, g, seed=2)
    p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
    with pps.Solver(g, method="mixed") as s:
        with pytest.raises(pps.PPSError) as e:
            s.solve_fields(f, p, 1e-10)
        assert e.value.code == pps.ERR_UNSUPPORTED
    g = pps.Grid.make(16)
    f = _fields(pps, g, seed=2)
    with pps.Solver(g, method="reference") as s:
        with pytest.raises(pps.PPSError) as e:
            s.solve_fields(f, p, 1e-10)
        assert e.value.code == pps.ERR_UNSUPPORTED
