"""The drop-in operator layer (parfile -> *OpInit* -> compiled pybind operators -> C ABI ->
HIP) against the CPU oracle's shot-level pipeline (FP64), and its adjoint relations.

Mirrors how the reference scripts use the operators (nonlinearPythonElasticFloatMain.py,
BornPythonElasticFloatMain.py): forward, adjoint, add semantics, setBackground, dot tests.
"""
import numpy as np
import pytest

import oracle as O
from elastic_isotropic_operators_amd import Elastic_iso_double_prop as PD
from elastic_isotropic_operators_amd import Elastic_iso_float_prop as PF
from elastic_isotropic_operators_amd import workloads as W

from helpers import operator_case, oracle_grid, rel_l2

pytestmark = pytest.mark.gpu


def _small(nts=150, seed=0):
    return W.small(nz=40, nx=56, nts=nts, sub=2, seed=seed)


def _nl_op(P, args, init):
    (model, data, elastic, par, sc, sx, sz, sxz, rc, rx, rz, rxz, _) = init(args)
    op = P.nonlinearPropElasticShotsGpu(model, data, elastic, par, sc, sx, sz, sxz, rc, rx, rz, rxz)
    return op, model, data, elastic


def test_nl_forward_and_adjoint_vs_oracle(tmp_path):
    w = _small()
    args, geo = operator_case(tmp_path, w, nExp=3)
    op, model, data, elastic = _nl_op(PF, args, PF.nonlinearOpInitFloat)
    model.getNdArray()[0] = w["model"].astype(np.float32)
    op.forward(False, model, data)
    g = oracle_grid(w)
    d = data.getNdArray()
    for e in range(3):
        ref = O.nl_fwd_shot(g, w["elastic"], geo["sz"][e], geo["sx"][e], geo["rz"], geo["rx"], w["model"])
        assert np.linalg.norm(ref) > 0
        assert rel_l2(d[e], ref) < 1e-5, (e, rel_l2(d[e], ref))
    # adjoint of random data, stacked over shots into the single source signal
    rng = np.random.default_rng(3)
    dd = data.clone()
    dd.getNdArray()[...] = rng.standard_normal(dd.getNdArray().shape)
    mm = model.clone()
    op.adjoint(False, mm, dd)
    ref = sum(O.nl_adj_shot(g, w["elastic"], geo["sz"][e], geo["sx"][e], geo["rz"], geo["rx"],
                            dd.getNdArray()[e].astype(np.float64)) for e in range(3))
    assert rel_l2(mm.getNdArray()[0], ref) < 1e-5


def test_nl_add_semantics(tmp_path):
    w = _small(nts=80)
    args, geo = operator_case(tmp_path, w, nExp=2)
    op, model, data, elastic = _nl_op(PF, args, PF.nonlinearOpInitFloat)
    model.getNdArray()[0] = w["model"]
    op.forward(False, model, data)
    d1 = data.getNdArray().copy()
    op.forward(True, model, data)
    assert rel_l2(data.getNdArray(), 2 * d1) < 1e-6


def test_nl_dot_tests_fp32_and_fp64(tmp_path):
    w = _small(nts=100, seed=1)
    args, _ = operator_case(tmp_path, w, nExp=2)
    op32, *_ = _nl_op(PF, args, PF.nonlinearOpInitFloat)
    assert op32.dotTestCpp(False, 1e-5)
    op64, *_ = _nl_op(PD, args, PD.nonlinearOpInitDouble)
    assert op64.dotTestCpp(False, 1e-12)
    e1, e2 = op64.dotTest(False, 1e-12)
    assert e1 < 1e-12 and e2 < 1e-12


def test_set_background_matches_new_operator(tmp_path):
    w = _small(nts=80, seed=2)
    args, geo = operator_case(tmp_path, w, nExp=1)
    op, model, data, elastic = _nl_op(PF, args, PF.nonlinearOpInitFloat)
    model.getNdArray()[0] = w["model"]
    el2 = elastic.clone()
    el2.getNdArray()[...] *= np.float32(1.02)
    op.setBackground(el2)
    op.forward(False, model, data)
    g = oracle_grid(w)
    ref = O.nl_fwd_shot(g, w["elastic"] * 1.02, geo["sz"][0], geo["sx"][0], geo["rz"], geo["rx"], w["model"])
    assert rel_l2(data.getNdArray()[0], ref) < 1e-5


def _born(P, args, init):
    (model, data, elastic, par, sig, sc, sx, sz, sxz, rc, rx, rz, rxz, _) = init(args)
    op = P.BornElasticShotsGpu(model, data, elastic, par, sig, sc, sx, sz, sxz, rc, rx, rz, rxz)
    return op, model, data, elastic, sig


def test_born_forward_adjoint_vs_oracle(tmp_path):
    w = _small(nts=120, seed=4)
    args, geo = operator_case(tmp_path, w, nExp=2)
    op, model, data, elastic, sig = _born(PF, args, PF.BornOpInitFloat)
    dm = W.born_perturbation(w, seed=5)
    model.getNdArray()[...] = dm
    op.forward(False, model, data)
    g = oracle_grid(w)
    for e in range(2):
        ref = O.born_fwd_shot(g, w["elastic"], geo["sz"][e], geo["sx"][e], w["model"], geo["rz"], geo["rx"], dm)
        assert np.linalg.norm(ref) > 0
        assert rel_l2(data.getNdArray()[e], ref) < 1e-5, (e, rel_l2(data.getNdArray()[e], ref))
    rng = np.random.default_rng(6)
    dd = data.clone()
    dd.getNdArray()[...] = rng.standard_normal(dd.getNdArray().shape)
    mm = model.clone()
    op.adjoint(False, mm, dd)
    ref = O.born_adj_shots(g, w["elastic"], geo["sz"], geo["sx"], w["model"], geo["rz"], geo["rx"],
                           dd.getNdArray().astype(np.float64))
    assert rel_l2(mm.getNdArray(), ref) < 1e-5, rel_l2(mm.getNdArray(), ref)


def test_c1_born_forward_and_adjoint_vs_fp64_oracle(tmp_path):
    """Workload C1 at full size (the C0 grid, 8 shots at 50-cell spacing, the seeded 5 %
    perturbation) through the float Born operator: the Born data of the perturbation against
    the FP64 oracle (1e-5) and the Born adjoint of that data (the imaging condition) against
    the oracle's (2e-5)."""
    w = W.gaussian_c1()
    dm = w["dm"]
    args, geo = operator_case(tmp_path, w, nExp=8, zSource=3, xSource=26, spacing=50, nRec=200,
                              depthRec=1, oRec=1, dRec=2)
    op, model, data, elastic, sig = _born(PF, args, PF.BornOpInitFloat)
    model.getNdArray()[...] = dm
    op.forward(False, model, data)
    assert op.pyOp.lastPropagationMs() > 0   # device time of the call, as bench.py reads it
    g = oracle_grid(w)
    ref = np.stack([O.born_fwd_shot(g, w["elastic"], geo["sz"][e], geo["sx"][e], w["model"], geo["rz"],
                                    geo["rx"], dm) for e in range(8)])
    assert np.linalg.norm(ref) > 0
    assert rel_l2(data.getNdArray(), ref) < 1e-5, rel_l2(data.getNdArray(), ref)
    img = model.clone()
    op.adjoint(False, img, data)
    ref_img = O.born_adj_shots(g, w["elastic"], geo["sz"], geo["sx"], w["model"], geo["rz"], geo["rx"],
                               data.getNdArray().astype(np.float64))
    assert rel_l2(img.getNdArray(), ref_img) < 2e-5, rel_l2(img.getNdArray(), ref_img)


def test_born_dot_test_fp64(tmp_path):
    w = _small(nts=100, seed=7)
    args, _ = operator_case(tmp_path, w, nExp=2)
    op, *_ = _born(PD, args, PD.BornOpInitDouble)
    e1, e2 = op.dotTest(False, 1e-12)
    assert e1 < 1e-12 and e2 < 1e-12


def test_fwi_operator_equals_nl_forward(tmp_path):
    w = _small(nts=80, seed=8)
    args, _ = operator_case(tmp_path, w, nExp=2)
    (elastic, elasticConv, data, sources, par, sc, sx, sz, sxz, rc, rx, rz, rxz, _) = PF.nonlinearFwiOpInitFloat(args)
    fwi = PF.nonlinearFwiPropElasticShotsGpu(elasticConv, data, sources, par, sc, sx, sz, sxz, rc, rx, rz, rxz)
    d_fwi = data.clone()
    fwi.forward(False, elasticConv, d_fwi)
    op, model, data2, el2 = _nl_op(PF, args, PF.nonlinearOpInitFloat)
    model.getNdArray()[...] = sources.getNdArray()
    op.forward(False, model, data2)
    assert rel_l2(d_fwi.getNdArray(), data2.getNdArray()) < 1e-6


def test_forward_and_adjoint_wavefield(tmp_path):
    """forwardWavefield / adjointWavefield / getWavefield with saveWavefield=1 and
    wavefieldShotNumber=1: the wavefield (axes z, x, component, time) interpolated to the
    receivers with the operator's own geometry reproduces the recorded data of that shot, and
    the scaled adjoint wavefield interpolated to the source reproduces the adjoint model."""
    from elastic_isotropic_operators_amd import geometry as G
    w = _small(nts=90)
    args, geo = operator_case(tmp_path, w, nExp=3, extra=dict(saveWavefield=1, wavefieldShotNumber=1))
    op, model, data, elastic = _nl_op(PF, args, PF.nonlinearOpInitFloat)
    model.getNdArray()[0] = w["model"].astype(np.float32)
    op.forwardWavefield(False, model, data)
    wv = op.getWavefield()
    ax = wv.getHyper().axes
    assert [a.n for a in ax] == [w["nz"], w["nx"], 5, w["nts"]]
    wf = wv.getNdArray().astype(np.float64)          # (nts, 5, nx, nz)
    d = data.getNdArray()[1]
    d2 = data.clone()
    op.forward(False, model, d2)
    # the plain forward takes two fine steps per launch, the wavefield-saving one one step: the
    # same arithmetic per cell, equal to FP32 rounding (the compiler may contract the two
    # kernels' multiply-adds differently)
    assert rel_l2(d2.getNdArray(), data.getNdArray()) < 1e-6
    ri = G.staggered_interps(w["nz"], w["oz"], w["dz"], w["ox"], w["dx"], geo["rz"], geo["rx"])
    for k, gi in enumerate(G.GRID_OF_COMP):
        reg = wf[:, k].reshape(w["nts"], -1)[:, ri[gi].uniq].T
        assert rel_l2(ri[gi].forward(reg), d[k]) < 1e-5
    # adjoint
    rng = np.random.default_rng(4)
    dd = data.clone()
    dd.getNdArray()[...] = rng.standard_normal(dd.getNdArray().shape)
    mm = model.clone()
    op.adjointWavefield(False, mm, dd)
    wa = op.getWavefield().getNdArray().astype(np.float64)
    si = G.staggered_interps(w["nz"], w["oz"], w["dz"], w["ox"], w["dx"], geo["sz"][1], geo["sx"][1])
    mm1 = model.clone()
    op.adjoint(False, mm1, dd)
    # the plain adjoint takes two reverse fine steps per launch, the wavefield-saving one one
    # step: the same operator within FP32 rounding
    assert rel_l2(mm1.getNdArray(), mm.getNdArray()) < 1e-6
    # shot 1's own contribution to the (shot-stacked) adjoint model
    dd1 = dd.clone()
    dd1.getNdArray()[...] = 0
    dd1.getNdArray()[1] = dd.getNdArray()[1]
    m1 = model.clone()
    op.adjoint(False, m1, dd1)
    for k, gi in enumerate(G.GRID_OF_COMP):
        reg = wa[:, k].reshape(w["nts"], -1)[:, si[gi].uniq].T
        assert rel_l2(si[gi].forward(reg), m1.getNdArray()[0, k]) < 1e-5


def test_free_surface_operator_vs_oracle(tmp_path):
    """surfaceCondition=1 through the parfile: the padding rule puts the free surface on the
    first interior row, the NL operator matches the oracle's shot-level free-surface forward,
    and the FP64 operator passes the dot test; the Born operator refuses it."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(5)
    nz, nx, nts = 40, 56, 120
    vp = gaussian_filter(2200 + 400 * rng.random((nx, nz)), 3)
    rho = 1800 + 300 * gaussian_filter(rng.random((nx, nz)), 3)
    w = W._grid_dict(np.stack([vp, vp / 1.8, rho]).astype(np.float32), 10.0, 10.0, 10, 10, nts, 0.002, 2, 20.0,
                     surfaceCondition=1)
    w["model"] = W.explosive_model(nts, w["dts"], 20.0, 0.05)
    args, geo = operator_case(tmp_path, w, nExp=2, zSource=1, depthRec=1, extra=dict(surfaceCondition=1))
    op, model, data, elastic = _nl_op(PF, args, PF.nonlinearOpInitFloat)
    model.getNdArray()[0] = w["model"]
    op.forward(False, model, data)
    g = O.grid(w["nz"], w["nx"], w["nts"], 2 * w["sub"], w["dz"], w["dx"], w["dts"], W.min_pad(w), w["alphaCos"],
               w["oz"], w["ox"], surface=1)
    for e in range(2):
        ref = O.nl_fwd_shot(g, w["elastic"], geo["sz"][e], geo["sx"][e], geo["rz"], geo["rx"], w["model"])
        assert rel_l2(data.getNdArray()[e], ref) < 1e-5, (e, rel_l2(data.getNdArray()[e], ref))
    op64, *_ = _nl_op(PD, args, PD.nonlinearOpInitDouble)
    assert op64.dotTestCpp(False, 1e-12)
    (m, d, el, par, sig, sc, sx, sz, sxz, rc, rx, rz, rxz, _) = PF.BornOpInitFloat(args)
    with pytest.raises(RuntimeError, match="nonlinear operator only"):
        PF.BornElasticShotsGpu(m, d, el, par, sig, sc, sx, sz, sxz, rc, rx, rz, rxz)


@pytest.mark.parametrize("smeth,rmeth", [(("sinc", 3), ("gauss", 2)), (("gauss", 1), ("sinc", 1)),
                                         (("sinc", 4), ("linear", 1))])
def test_sinc_gauss_acquisition_vs_oracle(tmp_path, smeth, rmeth):
    """sourceInterpMethod / recInterpMethod = sinc or gauss (spaceInterpGpu.cpp:139-207)
    through the drop-in NL operator against the oracle with the same stencil."""
    from helpers import oracle_nl_irregular
    w = _small(nts=120, seed=5)
    extra = dict(sourceInterpMethod=smeth[0], sourceInterpNumFilters=smeth[1],
                 recInterpMethod=rmeth[0], recInterpNumFilters=rmeth[1])
    args, geo = operator_case(tmp_path, w, nExp=2, zSource=5, depthRec=3, extra=extra)
    op, model, data, elastic = _nl_op(PF, args, PF.nonlinearOpInitFloat)
    model.getNdArray()[0] = w["model"].astype(np.float32)
    op.forward(False, model, data)
    g = oracle_grid(w)
    for e in range(2):
        ref = oracle_nl_irregular(w, g, geo["sz"][e], geo["sx"][e], geo["rz"], geo["rx"], w["model"],
                                  smeth, rmeth)
        assert np.linalg.norm(ref) > 0
        assert rel_l2(data.getNdArray()[e], ref) < 1e-5, (e, rel_l2(data.getNdArray()[e], ref))
    op64, *_ = _nl_op(PD, args, PD.nonlinearOpInitDouble)
    assert op64.dotTestCpp(False, 1e-12)


@pytest.mark.parametrize("zsh,xsh", [(1.0, 0.0), (-0.5, 1.5)])
def test_dipole_acquisition_vs_oracle(tmp_path, zsh, xsh):
    """dipole=1 with zDipoleShift / xDipoleShift (spaceInterpGpu.cpp:108-137: sources and
    receivers read and inject f(p) - f(p + shift)) through the drop-in NL operator against the
    oracle's pipeline with the same dipole stencils; FP64 dot test of the operator."""
    from helpers import oracle_nl_irregular
    w = _small(nts=120, seed=6)
    extra = dict(dipole=1, zDipoleShift=zsh, xDipoleShift=xsh)
    args, geo = operator_case(tmp_path, w, nExp=2, zSource=5, depthRec=3, extra=extra)
    op, model, data, elastic = _nl_op(PF, args, PF.nonlinearOpInitFloat)
    model.getNdArray()[0] = w["model"].astype(np.float32)
    op.forward(False, model, data)
    g = oracle_grid(w)
    dip = ("linear", 1, 1, zsh, xsh)
    for e in range(2):
        ref = oracle_nl_irregular(w, g, geo["sz"][e], geo["sx"][e], geo["rz"], geo["rx"], w["model"], dip, dip)
        assert np.linalg.norm(ref) > 0
        assert rel_l2(data.getNdArray()[e], ref) < 1e-5, (e, rel_l2(data.getNdArray()[e], ref))
    op64, *_ = _nl_op(PD, args, PD.nonlinearOpInitDouble)
    assert op64.dotTestCpp(False, 1e-12)
