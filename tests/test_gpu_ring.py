"""Deferred x updates of the fast solve (outer_iterate in pps_fast.hip): the
CG directions stay in a ring of 9 slots and k_flush applies them to x
together, when the ring is full and after the last iteration, each as the
same fma in iteration order.  The no-allocation solve keeps the former
form (k_dq_tiled applies the previous direction every iteration), so the
two must give the same p, iteration count and tres2 bit for bit, also over
runs long enough to fill the ring more than once and through restarts."""
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("method", ["mixed", "fp64"])
@pytest.mark.parametrize("rel", [1e-14, 1e-16])
def test_deferred_updates_equal_per_iteration_updates(pps, method, rel):
    _say("deferred vs per-iteration x updates", method, rel)
    shape, bc = (32, 32, 32), (1, 0, 0, 1, 0, 0)
    g = pps.Grid.make(*shape, bdry_cond=bc)
    rho, u, v, w = pps.synth_fields(g, [(0.4, 0.5, 0.5, 0.2), (0.75, 0.3, 0.6, 0.12)], 1.0, 1000.0, 1.5,
                                    seed=5)
    um, vm, wm = pps.pressure_masks(g)
    A = pps.setup_poisson(g, u, v, w, um, vm, wm, rho, 1.0)
    pps.remove_rhs_mean(g, A)
    brms = float(torch.sqrt((A[7] ** 2).mean()))
    # 1e-16 is below the true residual's rounding floor: the solve restarts
    # until maxit (the recursion meets it after about 12 iterations)
    tol, maxit = rel * brms, 60
    with pps.Solver(g, method=method) as s:
        p_ref = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it_ref, t2_ref = s.solve(A, p_ref, tol, 1.2, maxit, allow_maxit=True)
        n_hist = len(s.history())
    with pps.Solver(g, method=method) as s:
        p = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
        it, t2 = s.solve_noalloc(A, p, tol, 1.2, maxit, allow_maxit=True)
    _say(f"   it {it} / {it_ref}, {n_hist} iterations recorded, tres2/rms(b) {t2 / brms:.3e} / {t2_ref / brms:.3e}")
    # the ring (9 slots) filled at least once
    assert n_hist >= 10
    assert it == it_ref and t2 == t2_ref
    assert torch.equal(p, p_ref)
