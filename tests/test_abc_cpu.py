"""Learned absorbing boundary, host side: frame ownership, network inputs, layout and homogeneity."""
import numpy as np
import pytest
import torch

from elastic_isotropic_operators_amd import abc_learn as A


@pytest.mark.parametrize("nz,nx", [(40, 56), (128, 228), (37, 91)])
def test_frames_cover_the_stencil_read_cells_once(nz, nx):
    own = np.zeros((nx, nz), np.int32)
    for (x0, x1, z0, z1) in A.frame_regions(nz, nx):
        own[x0:x1, z0:z1] += 1
    F = A.FAT
    want = np.ones_like(own)
    want[F:nx - F, F:nz - F] = 0          # update cells: not written
    want[:F, :F] = want[:F, nz - F:] = want[nx - F:, :F] = want[nx - F:, nz - F:] = 0  # corners: never read
    assert np.array_equal(own, want)


def test_features_read_update_cells_only():
    nz, nx = 40, 56
    newer = torch.randn(5, nx, nz, dtype=torch.float64)
    older = torch.randn(5, nx, nz, dtype=torch.float64)
    scale = torch.tensor([1.0, 2.0, 4.0, 8.0, 16.0, 0.5, 0.25, 0.125, 1.0, 2.0])
    reg = A.frame_regions(nz, nx)
    f = A.features(newer, older, scale, reg)
    for (x0, x1, z0, z1), x in zip(reg, f):
        assert x.shape == (10, x1 - x0 + 12, z1 - z0 + 12)
    # x-min frame: crop column c is array column c - 6; only columns >= 4 (update cells) are nonzero
    x = f[0]
    assert torch.all(x[:, :10] == 0)
    c, r = 11, 9                                         # x = 5, z = 4 - 6 + 9 = 7
    assert torch.allclose(x[3, c, r], newer[3, 5, 7] / 8.0)
    assert torch.allclose(x[5 + 1, c, r], (newer[1, 5, 7] - older[1, 5, 7]) / 0.25)
    assert torch.all(x[:, :, :6] == 0)                   # z < 4


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
        assert float((net.edge_forward(1, 3.5 * x) - 3.5 * y).abs().max()) < 1e-5 * float(y.abs().max()) * 3.5 + 1e-6
        yb = net.edge_forward(1, x, bf16=True)
        assert float((yb - y).norm() / y.norm()) < 3e-2


def test_row_weights_follow_the_stencil():
    for e, axis, near in ((0, 2, -1), (1, 2, 0), (2, 3, -1), (3, 3, 0)):
        w = A.row_weights(e, device="cpu", floor=0.0)
        assert w.shape[axis] == 4 and float(w.reshape(-1)[near]) == 1.0
        second = float(w.reshape(-1)[1 if near == 0 else 2])
        assert abs(second - (A.STENCIL[1] / A.STENCIL[0]) ** 2) < 1e-9
        assert float(A.row_weights(e, device="cpu", floor=0.05).min()) == pytest.approx(0.05)


def test_save_load_roundtrip(tmp_path):
    net = A.AbcNet("cpu", seed=5)
    A.save(tmp_path / "w.npz", net, np.arange(1, 11, dtype=np.float32), 10, info={"val": [0.1, 0.2]})
    net2, scale, strip, on = A.load(tmp_path / "w.npz", device="cpu")
    assert strip == 10 and on == 1 and np.array_equal(scale, np.arange(1, 11, dtype=np.float32))
    assert np.array_equal(net2.flat(), net.flat())
