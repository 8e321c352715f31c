"""CPU tests of the product's host side: C-ABI library exports, C++ host layer (parameter
validation, media, spatial interpolation) against the oracle, numpy host prep, and the
error behaviour of the drop-in operators that fires before any GPU work."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from elastic_isotropic_operators_amd import _capi, geometry as G, workloads as W
from elastic_isotropic_operators_amd.sep import genericIO

from helpers import operator_case, oracle_grid, reg_positions, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_c_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "eio.h")).read()
    names = set(re.findall(r"^\s*(?:int|const char\*|void)\s+(eio_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 17
    lib = ctypes.CDLL(_capi.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(_capi.EXPORTED) <= names


def test_grid_check_rejects_grids_over_the_32bit_byte_limit():
    """eio_grid_check (no device): 5*nz*nx*precision must stay below 2^31 bytes; the grid just
    under the limit passes, the one just over it is EIO_ERR_ARG, in FP32 and FP64."""
    for prec in (4, 8):
        nz = 16384
        nx_max = ((1 << 31) - 1) // (5 * nz * prec)
        assert _capi.grid_check(nz, nx_max, 100, 2, 10.0, 10.0, 0.001, 20, prec) == 0
        with pytest.raises(_capi.EioError, match=r"error 1: grid too large"):
            _capi.grid_check(nz, nx_max + 1, 100, 2, 10.0, 10.0, 0.001, 20, prec)
    with pytest.raises(_capi.EioError, match="Subsampling parameter too high"):
        _capi.grid_check(100, 100, 100, 101, 10.0, 10.0, 0.001, 20, 4)
    with pytest.raises(_capi.EioError, match="Padding value is too high"):
        _capi.grid_check(100, 100, 100, 2, 10.0, 10.0, 0.001, 500, 4)


def test_fwi_group_and_segment_choice():
    """misfit_and_gradient's group size and store segment: full coarse storage (one segment of
    nts-1 intervals) only for a group that fits; a caller's group that does not fit, or no room
    for even one shot, takes checkpoints (interval `segment`, 0 = ceil(sqrt(nts-1)));
    checkpoint=L forces checkpoints (all shots in one group unless a group is given)."""
    from elastic_isotropic_operators_amd.fwi import group_and_segment as gs
    nts = 3000
    assert gs(None, 7, 32, nts, 0) == (7, nts - 1)          # default: as many as fit
    assert gs(4, 7, 32, nts, 0) == (4, nts - 1)             # smaller group than fits
    assert gs(10, 7, 32, nts, 0) == (10, 0)                 # caller's group too big: checkpoints
    assert gs(10, 7, 32, nts, 55) == (10, 55)
    assert gs(None, 0, 32, nts, 0) == (1, 0)                # not even one shot fits
    assert gs(None, 40, 32, nts, 0) == (32, nts - 1)        # clipped to the shot count
    assert gs(None, 7, 32, nts, 0, checkpoint=0) == (32, 0)
    assert gs(8, 7, 32, nts, 0, checkpoint=55) == (8, 55)


def test_fwi_cache_group():
    """The default FWI shot group keeps the gradient passes' working set (40 values per cell)
    within the 192 MiB shot-group budget: one shot of the 600 x 1592 C3 grid (153 MB), more of
    small grids, never zero."""
    from elastic_isotropic_operators_amd.fwi import cache_group, CACHE_GROUP_BYTES
    assert CACHE_GROUP_BYTES == 192 << 20
    assert cache_group(600, 1592, 4) == 1
    assert cache_group(280, 472, 4) == 9          # 21.1 MB per shot
    assert cache_group(5000, 5000, 8) == 1


def _conv_elastic(args):
    par = genericIO.io(params=args)
    el = genericIO.defaultIO.getVector(par.getString("elasticParam"), ndims=3)
    from elastic_isotropic_operators_amd import elasticParamConvertModule as C
    c = el.clone()
    C.ElasticConv(el, 1).forward(False, el, c)
    return par, c


def test_fdparam_media_and_checks(tmp_path):
    from elastic_isotropic_operators_amd import pyElastic_iso_float_nl as nl
    w = W.small(nz=40, nx=56, nts=100, sub=2)
    args, _ = operator_case(tmp_path, w)
    par, c = _conv_elastic(args)
    fd = nl.fdParamElastic(c, par)
    assert (fd.nts, fd.sub, fd.ntw) == (100, 4, 397)
    med = O.media(oracle_grid(w), w["elastic"])
    for k in range(5):
        assert rel_l2(fd.getMedia()[k], med[k]) < 1e-6
    # Courant > 0.45 (fdParamElastic.cpp:256-267)
    bad = dict(par.param, sub="1", dts="0.004")
    with pytest.raises(RuntimeError, match="Courant"):
        nl.fdParamElastic(c, bad)
    # dispersion < 3 (:269-281)
    with pytest.raises(RuntimeError, match="Dispersion"):
        nl.fdParamElastic(c, dict(par.param, fMax="200"))
    # block size (:283-294, NLF:204-208)
    with pytest.raises(RuntimeError, match="(?i)block size"):
        nl.fdParamElastic(c, dict(par.param, blockSize="8"))
    # 2*sub > SUB_MAX (:29-32)
    with pytest.raises(RuntimeError, match="SUB_MAX"):
        nl.fdParamElastic(c, dict(par.param, sub="51"))
    # parfile / model consistency (:296-324)
    with pytest.raises(RuntimeError, match="nz not consistent"):
        nl.fdParamElastic(c, dict(par.param, nz=str(w["nz"] + 16)))


def test_fdparam_time_axes_known_answer(tmp_path):
    """sub=10 -> fine sub-sampling 20, ntw = 39981 for nts = 2000 (NB/Nonlinear.ipynb cell 23)."""
    from elastic_isotropic_operators_amd import pyElastic_iso_float_nl as nl
    w = W.small(nz=40, nx=56, nts=2000, sub=10)
    w["dts"] = 0.004
    args, _ = operator_case(tmp_path, w, extra=dict(dts=0.004, fMax=5))
    par, c = _conv_elastic(args)
    fd = nl.fdParamElastic(c, par)
    assert (fd.sub, fd.ntw) == (20, 39981) and abs(fd.dtw - 0.0002) < 1e-9


def test_space_interp_matches_oracle(tmp_path):
    from elastic_isotropic_operators_amd import Elastic_iso_float_prop as PF
    w = W.small(nz=40, nx=56, nts=100, sub=2)
    args, geo = operator_case(tmp_path, w, nExp=3)
    par, c = _conv_elastic(args)
    sc, sx, sz, sxz, _ = PF.buildSourceGeometry(par, c)
    rc, rx, rz, rxz, _ = PF.buildReceiversGeometry(par, c)
    for e in range(3):
        ref = reg_positions(w, geo["sz"][e], geo["sx"][e])
        for g, lst in enumerate((sc, sx, sz, sxz)):
            assert list(lst[e].getRegPosUnique()) == list(ref[g])
    ref = reg_positions(w, geo["rz"], geo["rx"])
    for g, lst in enumerate((rc, rx, rz, rxz)):
        assert list(lst[0].getRegPosUnique()) == list(ref[g])
        wts = np.array(lst[0].getWeights()).reshape(-1, 4)
        assert np.allclose(wts.sum(axis=1), 1.0)


def test_space_interp_out_of_bounds():
    from elastic_isotropic_operators_amd import pyElastic_iso_float_nl as nl
    from elastic_isotropic_operators_amd.sep import Hypercube
    h = Hypercube.hypercube(axes=[Hypercube.axis(n=50, d=10.0), Hypercube.axis(n=60, d=10.0)])
    with pytest.raises(RuntimeError, match="out of bounds in the z"):
        nl.spaceInterpGpu(np.array([5.0], np.float32), np.array([300.0], np.float32), h, 10, "linear", 1, 0, 0, 0)
    # sinc / gauss keep nFilt cells from the edges (spaceInterpGpu.cpp:378-403)
    with pytest.raises(RuntimeError, match="out of bounds in the x"):
        nl.spaceInterpGpu(np.array([200.0], np.float32), np.array([35.0], np.float32), h, 10, "sinc", 4, 0, 0, 0)
    nl.spaceInterpGpu(np.array([200.0], np.float32), np.array([35.0], np.float32), h, 10, "linear", 4, 0, 0, 0)
    with pytest.raises(RuntimeError, match="dipole acquisition not implemented"):
        nl.spaceInterpGpu(np.array([200.0], np.float32), np.array([300.0], np.float32), h, 10, "gauss", 2, 1, 0, 0)
    with pytest.raises(RuntimeError, match="method not defined"):
        nl.spaceInterpGpu(np.array([200.0], np.float32), np.array([300.0], np.float32), h, 10, "cubic", 2, 0, 0, 0)


@pytest.mark.parametrize("method,nfilt", [("sinc", 1), ("sinc", 4), ("gauss", 1), ("gauss", 3), ("linear", 1)])
def test_sinc_gauss_interp_matches_oracle(method, nfilt):
    """The C++ interpolator (pybind spaceInterpGpu) and the numpy one (torch FWI engine) against
    the oracle's weights: same stencil, same unique regular list, weights to float rounding;
    the C++ and numpy versions compute in the same float/double order and agree to an ulp."""
    from elastic_isotropic_operators_amd import pyElastic_iso_float_nl as nl
    from elastic_isotropic_operators_amd.sep import Hypercube
    rng = np.random.default_rng(17)
    nz, oz, dz, ox, dx = 70, -40.0, 10.0, -45.0, 12.5
    z = (100.0 + 400.0 * rng.random(30)).astype(np.float32)
    x = (100.0 + 500.0 * rng.random(30)).astype(np.float32)
    z[0] = oz + 20 * dz  # on a node
    h = Hypercube.hypercube(axes=[Hypercube.axis(n=nz, o=oz, d=dz), Hypercube.axis(n=80, o=ox, d=dx)])
    cpp = nl.spaceInterpGpu(z, x, h, 10, method, nfilt, 0, 0, 0)
    npy = G.Interp(nz, oz, dz, ox, dx, z, x, method, nfilt)
    idx, w, uniq = O.interp(nz, oz, dz, ox, dx, z.astype(np.float64), x.astype(np.float64), method, nfilt)
    nf = idx.shape[1]
    assert list(cpp.getRegPosUnique()) == list(uniq) == list(npy.uniq)
    assert np.array_equal(np.array(cpp.getRegPos()).reshape(-1, nf), idx)
    assert np.array_equal(npy.idx, idx)
    wc = np.array(cpp.getWeights(), np.float64).reshape(-1, nf)
    # float positions (as the reference computes them) against the oracle's double ones
    assert np.abs(wc - w).max() < 1e-5 * np.abs(w).max()
    assert np.abs(wc - npy.w).max() <= 2 * np.spacing(np.float32(np.abs(w).max()))
    # forward / adjoint of the C++ interpolator are transposes of each other on these weights
    reg = rng.standard_normal((uniq.size, 10)).astype(np.float32)
    irr = rng.standard_normal((30, 10)).astype(np.float32)
    a = float(np.sum(npy.forward(reg).astype(np.float64) * irr))
    b = float(np.sum(reg.astype(np.float64) * npy.adjoint(irr)))
    assert abs(a - b) < 1e-5 * abs(a)


@pytest.mark.parametrize("zsh,xsh", [(1.0, 0.0), (-0.5, 1.5)])
def test_dipole_interp_matches_oracle(zsh, xsh):
    """The C++ interpolator with dipole=1 (pybind spaceInterpGpu, spaceInterpGpu.cpp:108-137)
    against the oracle's dipole stencil: same 8 points per device, unique list, weights to
    float rounding."""
    from elastic_isotropic_operators_amd import pyElastic_iso_float_nl as nl
    from elastic_isotropic_operators_amd.sep import Hypercube
    rng = np.random.default_rng(19)
    nz, oz, dz, ox, dx = 70, -40.0, 10.0, -45.0, 12.5
    z = (150.0 + 300.0 * rng.random(20)).astype(np.float32)
    x = (150.0 + 400.0 * rng.random(20)).astype(np.float32)
    h = Hypercube.hypercube(axes=[Hypercube.axis(n=nz, o=oz, d=dz), Hypercube.axis(n=80, o=ox, d=dx)])
    cpp = nl.spaceInterpGpu(z, x, h, 10, "linear", 1, 1, zsh, xsh)
    idx, w, uniq = O.interp(nz, oz, dz, ox, dx, z.astype(np.float64), x.astype(np.float64), "linear", 1, 1,
                            zsh, xsh)
    assert list(cpp.getRegPosUnique()) == list(uniq)
    assert np.array_equal(np.array(cpp.getRegPos()).reshape(-1, 8), idx)
    wc = np.array(cpp.getWeights(), np.float64).reshape(-1, 8)
    assert np.abs(wc - w).max() < 1e-5


def test_duplicate_gpu_ids_rejected_before_device_use(tmp_path):
    from elastic_isotropic_operators_amd import Elastic_iso_float_prop as PF
    w = W.small(nz=40, nx=56, nts=100, sub=2)
    args, _ = operator_case(tmp_path, w, extra=dict(iGpu="0,0"))
    (model, data, elastic, par, sc, sx, sz, sxz, rc, rx, rz, rxz, _) = PF.nonlinearOpInitFloat(args)
    with pytest.raises(RuntimeError, match="duplicates"):
        PF.nonlinearPropElasticShotsGpu(model, data, elastic, par, sc, sx, sz, sxz, rc, rx, rz, rxz)


def test_numpy_host_prep_matches_oracle():
    for w in (W.small(), W.gaussian_c0()):
        g = oracle_grid(w)
        med64 = O.media(g, w["elastic"])
        med32 = G.scaled_media(w["elastic"], G.dtw_of(w["dts"], w["sub"]))
        for k in range(5):
            assert rel_l2(med32[k], med64[k]) < 1e-6
        for zc, xc in ((w["src_z"][0], w["src_x"][0]), (w["rec_z"], w["rec_x"])):
            a = reg_positions(w, zc, xc)
            b = [ip.uniq for ip in G.staggered_interps(w["nz"], w["oz"], w["dz"], w["ox"], w["dx"], zc, xc)]
            assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_workload_sizes():
    c0 = W.gaussian_c0()
    assert (c0["nz"], c0["nx"], W.min_pad(c0)) == (280, 472, 30)
    assert W.cell_updates(c0) == 272 * 464 * 1998
    shots = W.layered_batch_c2(64)
    tot = sum(W.cell_updates(w) for w in shots)
    assert 4e11 < tot < 6e11


def test_autotune_signature_and_candidates():
    from elastic_isotropic_operators_amd import autotune
    a = autotune.signature([(300, 600, 7996, 4, 600), (400, 800, 5000, 4, 800)])
    b = autotune.signature([(400, 800, 5000, 4, 800), (300, 600, 7996, 4, 600)])
    assert a == b and len(a) == 16
    assert a != autotune.signature([(300, 600, 7996, 4, 600)])
    # the shot-group budget changes which form is fastest: it is part of the key
    assert a != autotune.signature([(300, 600, 7996, 4, 600), (400, 800, 5000, 4, 800)], group_bytes=0)
    assert a != autotune.signature([(300, 600, 7996, 4, 600), (400, 800, 5000, 4, 800)], group_streams=1)
    c = autotune.candidates()
    assert (1, 0) in c and (2, 0) in c and len(set(c)) == len(c)
