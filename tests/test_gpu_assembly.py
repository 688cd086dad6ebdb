"""Assembly and its checks on the GPU against the oracle: the pressure
boundary codes 5 / 6 of Poisson_BCs with non-zero BoundaryPressure, solid
cells (all six face masks closed) through SetupPoisson's solid rule, and
check_poisson_setup with the masks, every error code included."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _host(t):
    return np.ascontiguousarray(t.detach().cpu().numpy())


BP = np.array([0.7, -0.4, 1.3, 0.2, -0.9, 0.5])


@pytest.mark.parametrize("bc", [(5, 0, 0, 5, 0, 0), (5, 5, 5, 5, 5, 5), (5, 0, 1, 6, 0, 1),
                                (0, 5, 5, 4, 5, 5), (3, 3, 5, 6, 3, 5), (1, 5, 0, 1, 5, 5)])
@pytest.mark.parametrize("shape", [(16, 12, 8), (8, 8, 8)])
def test_pressure_codes_bit_exact(pps, orc, bc, shape):
    _say("pressure codes", bc, shape)
    nx, ny, nz = shape
    g = pps.Grid.make(nx, ny, nz, bdry_cond=bc)
    og = orc.OrcGrid.make(nx, ny, nz, bc=bc)
    rho, u, v, w = pps.synth_fields(g, [(0.3, 0.5, 0.5, 0.3)], 1.0, 100.0, 1.5, seed=4)
    um, vm, wm = pps.pressure_masks(g)
    A = pps.setup_poisson(g, u, v, w, um, vm, wm, rho, 0.5, boundary_pressure=BP)
    rc, oA = orc.setup_poisson(og, _host(u), _host(v), _host(w), _host(um), _host(vm), _host(wm),
                               _host(rho), 0.5, 0.0, bp=BP)
    assert rc == 0
    assert np.array_equal(_host(A), oA)
    # the reference and SOR methods take the non-symmetric rows: a few
    # cycles match the oracle's NewSolverMG / NewSolver_std
    if all(s % 2 == 0 for s in shape) and nx == ny == nz:
        p0 = np.zeros(og.shape_p)
        rc, it, t2, hist = orc.newsolver_std(og, oA, p0, 1e-300, 1.2, 5)
        with pps.Solver(g, method="sor") as s:
            p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
            s.solve(A, p, 1e-300, 1.2, 5)
            ghist = np.array(s.history())
        assert np.allclose(ghist, hist, rtol=1e-9, atol=0)
        assert np.abs(og.interior(_host(p)) - og.interior(p0)).max() <= \
            1e-12 * np.abs(og.interior(p0)).max()


def _solid_fields(pps, g, cells):
    """Droplet fields with the six face masks of `cells` closed and the
    velocities on those faces zero (a solid obstacle)."""
    rho, u, v, w = pps.synth_fields(g, [(0.5, 0.5, 0.5, 0.25)], 1.0, 10.0, 1.5, seed=8)
    um, vm, wm = pps.pressure_masks(g)
    G = g.ng
    for (i, j, k) in cells:
        I, J, K = i + G, j + G, k + G
        um[K, J, I - 1] = um[K, J, I] = 0.0
        vm[K, J - 1, I] = vm[K, J, I] = 0.0
        wm[K - 1, J, I] = wm[K, J, I] = 0.0
        u[K, J, I - 1] = u[K, J, I] = 0.0
        v[K, J - 1, I] = v[K, J, I] = 0.0
        w[K - 1, J, I] = w[K, J, I] = 0.0
    return rho, u, v, w, um, vm, wm


SOLID = [(3, 4, 5), (4, 4, 5), (10, 2, 7), (0, 0, 0)]


def test_solid_cells_assembly_and_check(pps, orc):
    _say("solid cells")
    bc = (0, 0, 1, 0, 0, 1)
    g = pps.Grid.make(12, 10, 8, bdry_cond=bc)
    og = orc.OrcGrid.make(12, 10, 8, bc=bc)
    rho, u, v, w, um, vm, wm = _solid_fields(pps, g, SOLID)
    H = {k: _host(t) for k, t in dict(rho=rho, u=u, v=v, w=w, um=um, vm=vm, wm=wm).items()}
    A = pps.setup_poisson(g, u, v, w, um, vm, wm, rho, 1.0)
    rc, oA = orc.setup_poisson(og, H["u"], H["v"], H["w"], H["um"], H["vm"], H["wm"], H["rho"])
    assert rc == 0
    assert np.array_equal(_host(A), oA)
    for (i, j, k) in SOLID:
        assert oA[6, k, j, i] == 1.0 and np.all(oA[:6, k, j, i] == 0.0)
    # check_poisson_setup: well posed, then every error code, device == oracle
    masks = (H["um"], H["vm"], H["wm"])
    cases = {"ok": lambda B: None}

    def nan(B):
        B[2, 3, 3, 3] = np.nan

    def tiny(B):
        B[6, 1, 2, 3] = 1e-60

    def source(B):
        B[7] += 1e-8

    def kernel(B):
        B[6, 5, 5, 5] *= 1.001

    def fluid_with_no_couplings(B):
        # a fluid cell (open masks) with zero couplings and A7 = 1: not in
        # the kernel by the masks, although its couplings look solid
        B[:6, 2, 6, 8] = 0.0
        B[6, 2, 6, 8] = 1.0

    cases.update(nan=nan, tiny=tiny, source=source, kernel=kernel, fluid=fluid_with_no_couplings)
    for name, f in cases.items():
        B = oA.copy()
        f(B)
        want, _ = orc.check_poisson_setup(og, B, *masks)
        got = pps.check_setup(g, torch.from_numpy(B).cuda(), um, vm, wm)
        _say(f"   {name}: oracle {want} device {got}")
        assert got == want
        if name == "fluid":
            assert want == orc.ERR_NOT_IN_KERNEL
            # without masks the couplings make it look solid
            assert pps.check_setup(g, torch.from_numpy(B).cuda()) == pps.OK


def test_solid_rule_errors_match_oracle(pps, orc):
    """check_and_debug_Poisson: a solid cell with a non-zero source, and a
    tiny diagonal outside a solid (dt = 0), are errors on both sides."""
    _say("solid rule errors")
    g = pps.Grid.make(8)
    og = orc.OrcGrid.make(8)
    rho, u, v, w, um, vm, wm = _solid_fields(pps, g, [(3, 3, 3)])
    G = g.ng
    u2 = u.clone()
    u2[3 + G, 3 + G, 3 + G - 1] = 0.5   # flux into the solid cell: A8 > 0
    for name, uu, dt in (("source in solid", u2, 1.0), ("dt = 0", u, 0.0)):
        rc, _ = orc.setup_poisson(og, _host(uu), _host(v), _host(w), _host(um), _host(vm),
                                  _host(wm), _host(rho), dt, 0.0)
        with pytest.raises(pps.PPSError) as e:
            pps.setup_poisson(g, uu, v, w, um, vm, wm, rho, dt)
        _say(f"   {name}: oracle {rc} device {e.value.code}")
        assert rc == orc.ERR_SOLID and e.value.code == pps.ERR_NAN_OR_TINY_DIAG
