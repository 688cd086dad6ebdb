"""The 512^3 workload: 512 droplets (radii uniform in [0.01, 0.04]), density
ratio 1000, walls; the forward FP32-V-cycle / FP64-refined solve and the
adjoint backward in one autograd call.  Checked through size-independent
properties, since the oracle's solver cannot reach this size in a test:
the oracle's own residual of the returned pressure and of the adjoint
solution, agreement with the FP64 method, and a central-difference
directional derivative of the gradient."""
import time

import numpy as np
import pytest

from helpers import faces_from_rho, objective, objective_diff, sine_target

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _padded(grid, interior):
    p = torch.zeros(grid.shape_p, dtype=torch.float64, device=interior.device)
    grid.interior(p).copy_(interior)
    return p


def _rel_rms_residual(orc, og, A_host, p, rhs):
    r = orc.residual_field(og, A_host, np.ascontiguousarray(p.cpu().numpy()))
    return float(np.sqrt((r ** 2).mean()) / np.sqrt((rhs ** 2).mean()))


def test_drops512_forward_backward(pps, orc):
    _say("512^3 forward + backward")
    grid, A, f = pps.make_problem("drops512")
    og = orc.OrcGrid.make(grid.nx)
    cf = faces_from_rho(grid, f["rho"])
    b = A[7].clone().requires_grad_(True)
    w = pps.face_weights(grid)
    pt = sine_target(grid, seed=21)
    PS = pps.PressureSolve
    with pps.Solver(grid, method="mixed") as s:
        c = cf.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p = PS.apply(c, b, s, w, 1e-10, 200)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J = objective(p, pt)
        try:
            J.backward()
        except Exception:
            _say("   adjoint history", ["%.2e" % v for v in s.history()[:20]])
            raise
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        _say(f"   forward {1e3 * (t1 - t0):.1f} ms, backward {1e3 * (t2 - t1):.1f} ms")
        g = c.grad.detach()
        lam = b.grad.detach()
        p = p.detach()

        # forward: the oracle's residual of p, relative to the right-hand side
        A_host = np.ascontiguousarray(pps.faces_to_A(grid, cf, w, b.detach()).cpu().numpy())
        rf = _rel_rms_residual(orc, og, A_host, _padded(grid, p), A_host[7])
        _say(f"   forward rel. residual (oracle) {rf:.2e}")
        assert rf <= 1e-10

        # backward: M lam = P0(p0 - p_target), checked with the oracle's residual
        p0 = p - p.mean()
        rhs = (p0 - pt) - (p0 - pt).mean()
        B = A_host.copy()
        B[7] = rhs.cpu().numpy()
        rb = _rel_rms_residual(orc, og, B, _padded(grid, lam), B[7])
        _say(f"   adjoint rel. residual (oracle) {rb:.2e}")
        assert rb <= 1e-10
        del B, A_host

        # agreement with the FP64 method (pinned against the oracle at
        # smaller sizes), null-space mean removed
        with pps.Solver(grid, method="fp64") as s64:
            p64 = PS.apply(cf, b.detach(), s64, w, 1e-10, 200)
        d = (p - p.mean()) - (p64 - p64.mean())
        err = float(d.abs().max() / (p64 - p64.mean()).abs().max())
        _say(f"   mixed vs fp64 max-norm rel. difference {err:.2e}")
        # measured 5.3e-15 on an MI355X
        assert err <= 1e-13

        # a directional derivative: central differences of J along a random
        # relative perturbation of every open face, against <dJ/dc, v>
        # (single faces carry too little of J at this size for a difference
        # quotient to resolve them at the solver's tolerance)
        gen = torch.Generator(device="cuda").manual_seed(5)
        v = (torch.rand(cf.shape, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1) * cf
        eps = 1e-6
        with torch.no_grad():
            pm = [PS.apply(cf + sgn * eps * v, b.detach(), s, w, 1e-13, 300) for sgn in (1.0, -1.0)]
        fd = objective_diff(pm[0], pm[1], pt) / (2 * eps)
        ad = float((g * v).sum())
        rel = abs(ad - fd) / abs(fd)
        _say(f"   directional derivative: adjoint {ad:.6e} central differences {fd:.6e} rel.err {rel:.2e}")
        assert rel <= 1e-6
