"""zone_gather (zonal Manning n -> per-cell n for the autograd Function):
the forward equals the dense one-hot product Z @ n bit for bit, and the
gradient is Z^T g, summed over the cells as a column sum of Z * g."""
import numpy as np
import torch

from cuteflow_amd.adjoint import zone_gather, zone_map


def test_zone_gather_forward_and_gradient():
    rng = np.random.default_rng(3)
    zones = rng.integers(0, 8, size=5000)
    Z = zone_map(zones, 8)
    zidx = torch.as_tensor(zones, dtype=torch.int64)
    n = torch.tensor(rng.uniform(0.02, 0.08, 8), dtype=torch.float64, requires_grad=True)
    cells = zone_gather(n, zidx, Z)
    assert torch.equal(cells, (Z @ n).detach())
    g = torch.tensor(rng.standard_normal(5000), dtype=torch.float64)
    (cells * g).sum().backward()
    assert torch.equal(n.grad, (Z * g.unsqueeze(1)).sum(dim=0))
    assert torch.allclose(n.grad, Z.t() @ g, rtol=1e-13, atol=0.0)
