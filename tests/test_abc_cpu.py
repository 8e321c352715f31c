"""Learned absorbing boundary, host side: band ownership, network layout and homogeneity."""
import numpy as np
import pytest
import torch

from elastic_isotropic_operators_amd import abc_learn as A


@pytest.mark.parametrize("nz,nx,S", [(40, 56, 10), (128, 228, 10), (37, 91, 7), (30, 30, 11)])
def test_bands_cover_every_edge_cell_once(nz, nx, S):
    own = np.zeros((nx, nz), np.int32)
    for (x0, x1, z0, z1) in A.band_regions(nz, nx, S):
        own[x0:x1, z0:z1] += 1
    F = A.FAT
    want = np.zeros_like(own)
    want[F:nx - F, F:nz - F] = 1
    want[F + S:nx - F - S, F + S:nz - F - S] = 0
    assert np.array_equal(own, want)


def test_features_crop_and_zero_outside():
    nz, nx, S = 40, 56, 10
    newer = torch.randn(5, nx, nz, dtype=torch.float64)
    older = torch.randn(5, nx, nz, dtype=torch.float64)
    scale = torch.tensor([1.0, 2.0, 4.0, 8.0, 16.0])
    reg = A.band_regions(nz, nx, S)
    f = A.features(newer, older, scale, reg)
    for (x0, x1, z0, z1), x in zip(reg, f):
        assert x.shape == (10, x1 - x0 + 12, z1 - z0 + 12)
    # x-min band: crop column c is array column x0 - 6 + c; columns < 0 are zero
    x = f[0]
    assert torch.all(x[:, :2] == 0)                      # x = -2, -1
    c, r = 5, 9                                          # x = 3, z = -2 + 9 = 7
    assert torch.allclose(x[3, c, r], newer[3, 3, 7] / 8.0)
    assert torch.allclose(x[5 + 1, c, r], older[1, 3, 7] / 2.0)


def test_network_layout_and_homogeneity():
    net = A.AbcNet("cpu", seed=3)
    flat = net.flat()
    assert flat.size == 4 * (16 * 10 + 16 * 16 + 5 * 16) * 25
    # flat order: edge-major, then layer, then torch layout
    assert np.array_equal(flat[:4000], net.W[0][0].detach().numpy().reshape(-1))
    off = 4 * 0 + 12400 * 2 + 4000 + 6400
    assert np.array_equal(flat[off:off + 2000], net.W[2][2].detach().numpy().reshape(-1))
    x = torch.randn(2, 10, 22, 30)
    with torch.no_grad():
        y = net.edge_forward(1, x)
        assert y.shape == (2, 5, 10, 18)
        # no biases + leaky ReLU: positively homogeneous of degree 1
        assert torch.allclose(net.edge_forward(1, 3.5 * x), 3.5 * y, rtol=1e-5, atol=1e-6)
        yb = net.edge_forward(1, x, bf16=True)
        assert float((yb - y).norm() / y.norm()) < 3e-2


def test_save_load_roundtrip(tmp_path):
    net = A.AbcNet("cpu", seed=5)
    A.save(tmp_path / "w.npz", net, np.arange(1, 6, dtype=np.float32), 10, info={"val": [0.1, 0.2]})
    net2, scale, strip = A.load(tmp_path / "w.npz", device="cpu")
    assert strip == 10 and np.array_equal(scale, np.arange(1, 6, dtype=np.float32))
    assert np.array_equal(net2.flat(), net.flat())
