"""poi_solve (modules.f90:2411-2593), the HYPRE entry on the same 8-slot
system, through the C ABI: every solver type runs the fast MG-PCG to the
relative residual HYPRE's SetTol gives; the result is checked with the
oracle's residual and against the oracle's solution run to stagnation."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _host(t):
    return np.ascontiguousarray(t.detach().cpu().numpy())


def _mf(x):
    return x - x.mean()


@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_poi_solve_matches_oracle(pps, orc, method):
    _say("poi_solve", method)
    n, bc = 32, (1, 0, 0, 1, 0, 0)
    og = orc.OrcGrid.make(n, bc=bc)
    g, At, _ = pps.make_problem(n=n, bc=bc, drops=[(0.3, 0.5, 0.5, 0.2), (0.7, 0.4, 0.6, 0.15)],
                                rho2=1000.0, seed=3)
    A = _host(At)
    pref = np.zeros(og.shape_p)
    orc.newsolver_mg(og, A, pref, 1e-300, 1.2, 3000)
    ref = _mf(og.interior(pref))
    bn = np.linalg.norm(A[7])
    results = []
    with pps.Solver(g, method=method) as s:
        for st in (1, 2, 3, 4):
            p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
            nit = s.poi_solve_device(At, p, 1e-13, 100, st)
            ph = _host(p)
            rel = np.linalg.norm(orc.residual_field(og, A, ph)) / bn
            err = np.abs(_mf(og.interior(ph)) - ref).max() / np.abs(ref).max()
            _say(f"   type {st}: it={nit} ||r||/||b|| {rel:.2e} err {err:.2e}")
            assert nit <= 100 and rel < 1e-13 and err <= 1e-10
            results.append(ph)
        # every solver type is the same algorithm
        for ph in results[1:]:
            assert np.array_equal(ph, results[0])
        # host entry point
        phh = np.zeros(og.shape_p)
        nit_h = s.poi_solve(A, phh, 1e-13, 100, 2)
        assert np.array_equal(og.interior(phh), og.interior(results[0]))
        # maxit reached: num_iterations = maxit
        p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        assert s.poi_solve_device(At, p, 1e-13, 2, 2) == 2
        with pytest.raises(pps.PPSError):
            s.poi_solve_device(At, p, 1e-13, 100, 5)
