"""The HBM streaming probes compute what they state (so the bandwidth they
report is for that work): copy, write, the RK4-stage streams and the stage
tile pattern, on small buffers against numpy."""
import numpy as np
import pytest

from phix_amd import capi

from helpers import stream

pytestmark = pytest.mark.gpu


def test_probe_kernels_compute_what_they_state(dev):
    import torch
    rng = np.random.default_rng(7)
    n = 1 << 16
    a, b, c = (torch.from_numpy(rng.standard_normal(n).astype(np.float32)).to(dev) for _ in range(3))
    d, e = torch.zeros_like(a), torch.zeros_like(a)
    P = capi.ptr
    capi.call("phx_hbm_probe", 2, P(a), P(d), None, None, None, n, 0, 0, 0, stream())
    torch.cuda.synchronize()
    assert torch.equal(d, a)
    capi.call("phx_hbm_probe", 3, P(a), P(b), P(c), P(d), P(e), n, 0, 0, 64, stream())
    torch.cuda.synchronize()
    assert torch.allclose(d, b + 0.5 * a, rtol=0, atol=1e-6) and torch.allclose(e, c + 0.25 * a, rtol=0, atol=1e-6)
    capi.call("phx_hbm_probe", 1, P(e), None, None, None, None, n, 0, 0, 64, stream())
    capi.call("phx_hbm_probe", 0, P(a), P(b), None, None, None, n, 0, 0, 64, stream())
    torch.cuda.synchronize()
    assert torch.all(e == 1.0)


@pytest.mark.parametrize("tile,side", [(0, 64), (1, 128), (2, 256)])
def test_probe_tile_pattern_cells_only(dev, tile, side):
    """Kind 5 writes d = b + u/2 and c = c + u/4 on every interior cell (u = a)."""
    import torch
    S = side + 2
    rng = np.random.default_rng(tile)
    u, cc, ac = (rng.standard_normal((S, S, S)).astype(np.float32) for _ in range(3))
    tu, tc, ta = (torch.from_numpy(x).to(dev) for x in (u, cc, ac))
    to = torch.zeros_like(tu)
    P = capi.ptr
    capi.call("phx_hbm_probe", 5, P(tu), P(tc), P(ta), P(to), None, S ** 3, side, tile, 4, stream())
    torch.cuda.synchronize()
    inner = (slice(1, S - 1),) * 3
    np.testing.assert_allclose(to.cpu().numpy()[inner], cc[inner] + 0.5 * u[inner], rtol=0, atol=1e-6)
    np.testing.assert_allclose(ta.cpu().numpy()[inner], ac[inner] + 0.25 * u[inner], rtol=0, atol=1e-6)
