"""The fast solve's restart (fast_solve in pps_fast.hip): the outer CG stops
on its recursive residual, the final check forms the true residual b - M x
of x + alpha d (the last iteration's update, applied inside the check) and,
when that misses the tolerance while the recursion met it, iterates on from
x with the update applied first.  A tolerance below the rounding floor of
the true residual forces that path: the recursion falls below it, the true
residual does not, and each restart repeats that (four attempts).  The solve then returns not converged (it = maxit + 1)
with p at the rounding floor (the reported tres2 and the oracle's residual
of the returned p both there) and agreeing with the oracle's converged
solution."""
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


@pytest.mark.parametrize("method", ["mixed", "fp64"])
def test_restart_below_the_rounding_floor(pps, orc, method):
    _say("restart below the rounding floor", method)
    shape, bc = (32, 32, 32), (1, 0, 0, 1, 0, 0)
    g = pps.Grid.make(*shape, bdry_cond=bc)
    rho, u, v, w = pps.synth_fields(g, [(0.4, 0.5, 0.5, 0.2), (0.75, 0.3, 0.6, 0.12)], 1.0, 1000.0, 1.5,
                                    seed=5)
    um, vm, wm = pps.pressure_masks(g)
    A = pps.setup_poisson(g, u, v, w, um, vm, wm, rho, 1.0)
    pps.remove_rhs_mean(g, A)
    Ah = _host(A)
    og = orc.OrcGrid.make(*shape, bc=bc)
    brms = float(np.sqrt((Ah[7] ** 2).mean()))
    ph = np.zeros(og.shape_p)
    its, rel = orc.pcg_solve(og, Ah, ph, rtol=0.0, maxit=3000)
    assert rel <= 1e-14
    ref = _mf(og.interior(ph))
    # the true residual stops near 4e-16 rms(b), the recursion goes on to
    # about 5e-17 (then stalls near 3e-18): measured on an MI355X
    tol = 1e-16 * brms
    maxit = 60
    with pps.Solver(g, method=method) as s:
        p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it, t2 = s.solve(A, p, tol, 1.2, maxit, allow_maxit=True)
        hist = s.history()
    first = next((k for k, h in enumerate(hist) if h < tol), None)
    _say(f"   it {it}, {len(hist)} iterations in all, recursion below tol after {first}, "
         f"tres2/rms(b) {t2 / brms:.2e}")
    assert it == maxit + 1 and t2 >= tol
    # the recursion met the tolerance and the solve went on: a restart
    assert first is not None and len(hist) > first + 1
    # at the rounding floor two evaluations of b - M x differ by their own
    # rounding, so the reported tres2 and the oracle's are bounded, not equal
    r = orc.residual_field(og, Ah, _host(p))
    true = float(np.sqrt((r ** 2).mean()))
    assert true / brms < 1e-13 and t2 / brms < 1e-13
    err = float(np.abs(_mf(og.interior(_host(p))) - ref).max() / np.abs(ref).max())
    _say(f"   oracle residual {true / brms:.2e}, err_inf {err:.2e}")
    assert err <= 1e-10
