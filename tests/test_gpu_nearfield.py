"""Near-field set-up on the GPU (mnp_near_terms, mnp_colsum_far) against the host set-up
(greenret_refined.py:287-516, compgreen_stat.py:219-464)."""
import numpy as np
import pytest

from conftest import banner

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import mnpbem_amd as M                                   # noqa: E402
from mnpbem_amd.nearfield import NearField               # noqa: E402


def _cases():
    drude = M.EpsDrude(10, 9.065, 0.0708)
    yield "sphere400", M.ComParticle([M.EpsConst(1.0), drude], [M.trisphere(400, 20)], [[2, 1]], 1)
    yield "rod-quads", M.ComParticle([M.EpsConst(1.0), drude], [M.trirod(10, 40, [15, 15, 15])], [[2, 1]], 1)
    p1 = M.trisphere(256, 20).shift([0, 0, -11.0])
    p2 = M.trisphere(256, 20).shift([0, 0, 11.0])
    yield "dimer", M.ComParticle([M.EpsConst(1.0), drude], [p1, p2], [[2, 1], [2, 1]], 1, 2)


@pytest.mark.parametrize("name,cp", list(_cases()))
def test_device_nearfield_matches_host(name, cp):
    banner("nearfield device vs host: {} n={}".format(name, cp.nfaces))
    host = NearField(cp)
    dev = NearField(cp, device=torch.device("cuda", 0))
    assert np.array_equal(host.rows, dev.rows) and np.array_equal(host.cols, dev.cols)
    for a, b in ((host.g, dev.g), (host.f, dev.f)):
        scale = np.abs(a).max(axis=0, keepdims=True) + 1e-300
        assert np.max(np.abs(a - b) / scale) < 1e-12
    assert np.max(np.abs(host.closed_correction - dev.closed_correction)) < 1e-10 * np.abs(host.closed_correction).max()
