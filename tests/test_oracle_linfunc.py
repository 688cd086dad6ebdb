"""The oracle's linfunc (vofmodules.f90:171-218) + do_all_ghost
(modules.f90:2062-2072) pinned by what the mathematics fixes: the end
values and the harmonic mean of the two maps, the filter as the tensor
product of [1, 2, 1] / 4 (an independent numpy correlation), constants and
linear fields kept by the filter, and the ghost wrap of periodic axes
composed axis by axis."""
import numpy as np
import pytest


def _grid(orc, n, bc):
    return orc.OrcGrid.make(*n, bc=bc)


def test_linfunc_maps(orc):
    g = _grid(orc, (6, 5, 4), (0,) * 6)
    a1, a2 = 1.0, 1000.0
    for c, arith, harm in ((0.0, a1, a1), (1.0, a2, a2), (0.5, 0.5 * (a1 + a2), 2 * a1 * a2 / (a1 + a2))):
        f = np.full(g.shape_p, c)
        assert np.allclose(orc.linfunc(g, f, a1, a2, 101, 0), arith, rtol=1e-15, atol=0)
        assert np.allclose(orc.linfunc(g, f, a1, a2, 102, 0), harm, rtol=1e-14, atol=0)
    with pytest.raises(ValueError):
        orc.linfunc(g, np.zeros(g.shape_p), a1, a2, 103, 0)
    with pytest.raises(ValueError):
        orc.linfunc(g, np.zeros(g.shape_p), a1, a2, 101, 2)


def test_linfunc_filter_is_the_121_tensor_product(orc):
    g = _grid(orc, (7, 6, 5), (0,) * 6)
    rng = np.random.default_rng(1)
    c = rng.random(g.shape_p)
    f = orc.linfunc(g, c, 0.0, 1.0, 101, 1)     # a1 = 0, a2 = 1: the filtered cvof itself
    w1 = np.array([0.25, 0.5, 0.25])
    w = w1[:, None, None] * w1[None, :, None] * w1[None, None, :]
    inner = np.zeros_like(c[1:-1, 1:-1, 1:-1])
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                inner += w[dz, dy, dx] * c[dz:dz + inner.shape[0], dy:dy + inner.shape[1], dx:dx + inner.shape[2]]
    assert np.allclose(f[1:-1, 1:-1, 1:-1], inner, rtol=1e-14, atol=1e-15)
    # the outermost ghost layer is not filtered
    edge = np.ones(g.shape_p, bool)
    edge[1:-1, 1:-1, 1:-1] = False
    assert np.array_equal(f[edge], c[edge])
    # constants exactly, linear fields to rounding
    assert np.array_equal(orc.linfunc(g, np.full(g.shape_p, 0.375), 0.0, 1.0, 101, 1), np.full(g.shape_p, 0.375))
    lin = np.add.outer(np.add.outer(0.01 * np.arange(g.shape_p[0]), 0.02 * np.arange(g.shape_p[1])),
                       0.03 * np.arange(g.shape_p[2]))
    fl = orc.linfunc(g, lin, 0.0, 1.0, 101, 1)
    assert np.allclose(fl[1:-1, 1:-1, 1:-1], lin[1:-1, 1:-1, 1:-1], rtol=0, atol=1e-14)


def test_linfunc_ghost_wrap(orc):
    """x periodic, y and z walls: x ghosts take the wrapped values over the
    full y, z extent (corners included); y and z ghosts keep their own."""
    n, ng = (6, 5, 4), 2
    rng = np.random.default_rng(2)
    gw = _grid(orc, n, (0,) * 6)
    gp = _grid(orc, n, (1, 0, 0, 1, 0, 0))
    c = rng.random(gw.shape_p)
    fw = orc.linfunc(gw, c, 1.0, 10.0, 102, 1)
    fp = orc.linfunc(gp, c, 1.0, 10.0, 102, 1)
    nx = n[0]
    # (z, y, x) arrays: x ghosts are the first and last ng columns
    assert np.array_equal(fp[:, :, :ng], fw[:, :, nx:nx + ng])
    assert np.array_equal(fp[:, :, ng + nx:], fw[:, :, ng:2 * ng])
    assert np.array_equal(fp[:, :, ng:ng + nx], fw[:, :, ng:ng + nx])
    # all periodic: a corner ghost is the fully wrapped interior value
    ga = _grid(orc, n, (1,) * 6)
    fa = orc.linfunc(ga, c, 1.0, 10.0, 102, 1)
    assert fa[0, 0, 0] == fa[n[2], n[1], n[0]]
    assert fa[-1, -1, -1] == fa[ng + 1, ng + 1, ng + 1]
