"""Every benchmark configuration at its own size, through the C ABI.

* 64^3 (one sphere, density ratio 10, walls) and 128^3 (one sphere, density
  ratio 1000, walls): the oracle's Jacobi-PCG, run to a 1e-15 relative
  residual, gives the reference solution; each GPU form is compared with it
  in the max norm (null-space mean removed) at the metric's 1e-10 stop and
  solved on to stagnation, where the bar is 1e-10.
* 256^3 (64 droplets, periodic x, walls y/z, density ratio 1000): the
  oracle's solver is too slow at this size inside a test, so each form's
  1e-10 result is pinned by the oracle's residual of the returned pressure,
  and the forms are compared with each other at the stop and at stagnation.
"""
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


def _err(x, ref):
    return float(np.abs(_mf(x) - ref).max() / np.abs(ref).max())


def _solve(pps, grid, A, method, rtol, brms, maxit=200):
    with pps.Solver(grid, method=method) as s:
        p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
        it, t2 = s.solve(A, p, rtol * brms, 1.2, maxit)
    return p, it, t2


def _oracle_reference(orc, og, Ah):
    """The oracle's converged solution (Jacobi-PCG to a 1e-15 relative
    residual), mean removed."""
    ph = np.zeros(og.shape_p)
    its, rel = orc.pcg_solve(og, Ah, ph, rtol=1e-15, maxit=200000)
    _say(f"   oracle PCG: {its} iterations, relative residual {rel:.2e}")
    assert rel <= 2e-15
    return _mf(og.interior(ph))


# measured maxima of the max-norm error at the 1e-10 stop (one MI355X),
# asserted with a factor-10 margin: the stop bounds the residual, not the
# error, which grows with the condition number (about 1e7 at 128^3 with
# density ratio 1000)
STOP_ERR = {"sphere64": 1e-8, "sphere128": 1e-7}


@pytest.mark.parametrize("name", ["sphere64", "sphere128"])
def test_config_vs_oracle_solution(pps, orc, name):
    _say("config vs oracle solution", name)
    grid, A, _ = pps.make_problem(name)
    og = orc.OrcGrid.make(grid.nx)
    Ah = _host(A)
    brms = float(np.sqrt((Ah[7] ** 2).mean()))
    ref = _oracle_reference(orc, og, Ah)
    for method in ("fp64", "mixed"):
        p, it, t2 = _solve(pps, grid, A, method, 1e-10, brms)
        ph = _host(p)
        r = orc.residual_field(og, Ah, ph)
        rel = float(np.sqrt((r ** 2).mean()) / brms)
        e_stop = _err(og.interior(ph), ref)
        p2, it2, t22 = _solve(pps, grid, A, method, 1e-14, brms)
        e_stag = _err(og.interior(_host(p2)), ref)
        _say(f"   {method}: 1e-10 stop: it={it} oracle rel. residual {rel:.2e} "
             f"err_inf {e_stop:.2e}; to stagnation: it={it2} rel.res {t22 / brms:.2e} "
             f"err_inf {e_stag:.2e}")
        assert it <= 20 and rel < 1e-10
        assert e_stop <= STOP_ERR[name]
        assert e_stag <= 1e-10


def test_drops256_forms(pps, orc):
    """The 256^3 benchmark field: both forms to the metric's 1e-10, pinned by
    the oracle's residual; mixed against FP64 at the stop and at
    stagnation."""
    _say("drops256 forward, both forms")
    grid, A, _ = pps.make_problem("drops256")
    og = orc.OrcGrid.make(256, bc=(1, 0, 0, 1, 0, 0))
    Ah = _host(A)
    brms = float(np.sqrt((Ah[7] ** 2).mean()))
    stop, stag = {}, {}
    for method in ("fp64", "mixed"):
        p, it, t2 = _solve(pps, grid, A, method, 1e-10, brms)
        ph = _host(p)
        r = orc.residual_field(og, Ah, ph)
        rel = float(np.sqrt((r ** 2).mean()) / brms)
        _say(f"   {method}: it={it} reported {t2 / brms:.2e} oracle rel. residual {rel:.2e}")
        assert it <= 20 and rel < 1e-10
        assert np.isclose(rel, t2 / brms, rtol=1e-6)
        stop[method] = _mf(og.interior(ph))
        del r, ph
        p2, it2, t22 = _solve(pps, grid, A, method, 1e-14, brms)
        ph2 = _host(p2)
        r2 = orc.residual_field(og, Ah, ph2)
        rel2 = float(np.sqrt((r2 ** 2).mean()) / brms)
        _say(f"   {method} to stagnation: it={it2} oracle rel. residual {rel2:.2e}")
        assert rel2 < 1e-13
        stag[method] = _mf(og.interior(ph2))
        del p, p2, r2, ph2
    ref = stag["fp64"]
    d_stag = _err(stag["mixed"], ref)
    e64 = _err(stop["fp64"], ref)
    e32 = _err(stop["mixed"], ref)
    d_stop = _err(stop["mixed"], stop["fp64"])
    _say(f"   at stagnation mixed vs fp64 {d_stag:.2e}; at the 1e-10 stop vs the stagnated "
         f"fp64 solution: fp64 {e64:.2e}, mixed {e32:.2e}; mixed vs fp64 {d_stop:.2e}")
    assert d_stag <= 1e-10
    assert max(e64, e32, d_stop) <= 1e-7
