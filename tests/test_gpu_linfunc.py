"""pps_linfunc_device against the oracle (vofmodules.f90:171-218 +
modules.f90:2062-2072): bit for bit on the padded field, both means, both
filter settings, walls and periodic axes, a non-cubic grid."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("bc", [(0,) * 6, (1, 0, 0, 1, 0, 0), (0, 1, 1, 0, 1, 1), (1,) * 6])
@pytest.mark.parametrize("mean_flag", [101, 102])
@pytest.mark.parametrize("nfilter", [0, 1])
def test_linfunc_bit_exact(pps, orc, bc, mean_flag, nfilter):
    print("linfunc", bc, mean_flag, nfilter, flush=True)
    n = (12, 10, 8)
    g = pps.Grid.make(*n, bdry_cond=bc)
    og = orc.OrcGrid.make(*n, bc=bc)
    rng = np.random.default_rng(5)
    c = rng.random(og.shape_p)
    c[c < 0.3] = 0.0
    c[c > 0.8] = 1.0
    ref = orc.linfunc(og, c, 1.2, 1000.0, mean_flag, nfilter)
    f = pps.linfunc(g, torch.from_numpy(c).cuda(), 1.2, 1000.0, mean_flag, nfilter)
    assert np.array_equal(f.cpu().numpy(), ref)


def test_linfunc_errors(pps):
    g = pps.Grid.make(8)
    c = torch.zeros(g.shape_p, dtype=torch.float64, device="cuda")
    for mf, nf in ((7, 0), (101, 2)):
        with pytest.raises(pps.PPSError) as e:
            pps.linfunc(g, c, 1.0, 2.0, mf, nf)
        assert e.value.code == pps.ERR_BAD_ARG
