"""GPU tests of the FWI problem layer on the HIP engine.

The masked, component-selected, vp/vs/rho-parameterised gradient of FwiProblem against the
same chain built from the FP64 oracle (NL forward, full-storage Born adjoint) and the
reference's conversion Jacobian in numpy; and a few projected L-BFGS iterations on a small
inversion that must lower the misfit, move the model towards the true one and stay in the
bounds."""
import numpy as np
import pytest

import oracle as O
from elastic_isotropic_operators_amd import seis_utils as su

from helpers import fwi_case, oracle_grid, rel_l2

pytestmark = pytest.mark.gpu


def _engine(w, g, dtype):
    from elastic_isotropic_operators_amd.fwi import ElasticFwi
    grid = dict(nz=w["nz"], nx=w["nx"], dz=w["dz"], dx=w["dx"], oz=w["oz"], ox=w["ox"], nts=w["nts"],
                dts=w["dts"], sub=w["sub"], minPad=g.minPad, alphaCos=w["alphaCos"])
    return ElasticFwi(grid, w["src_z"], w["src_x"], w["model"], w["rec_z"], w["rec_x"], dtype=dtype)


@pytest.mark.parametrize("tdtype", ["float32", "float64"])
def test_problem_gradient_vs_fp64_oracle_chain(tdtype):
    import torch
    from elastic_isotropic_operators_amd import fwi_problem as P
    dtype = getattr(torch, tdtype)
    w, g, m0, dtrue = fwi_case(nts=400, dts=0.001, nExp=2, nz=48, nx=72)
    sz, sx, rz, rx, src = w["src_z"], w["src_x"], w["rec_z"], w["rec_x"], w["model"]
    comps = "vx,vz,p"
    d_full = np.stack([O.nl_fwd_shot(g, m0 + dtrue, sz[e], sx[e], rz, rx, src) for e in range(2)])
    d_obs = su.datcomp_fwd(comps, d_full)
    rng = np.random.default_rng(7)
    data_mask = rng.uniform(0.5, 1.0, d_obs.shape)
    grad_mask = np.ones(m0.shape)
    grad_mask[:, :, :12] = 0.0  # mute the shallow part, as a water-layer mask would
    vvr = su.vpvsrho_to_rholamemu(m0, -1)
    # oracle chain
    F = np.stack([O.nl_fwd_shot(g, m0, sz[e], sx[e], rz, rx, src) for e in range(2)])
    r = data_mask * su.datcomp_fwd(comps, F) - data_mask * d_obs
    gd = su.datcomp_adj(comps, data_mask * r, F.shape)
    g_lame = O.born_adj_shots(g, m0, sz, sx, src, rz, rx, gd)
    ref = grad_mask * su.conv_jacobian_adj(vvr, g_lame)
    phi_ref = 0.5 * float(np.sum(r * r))

    eng = _engine(w, g, dtype)
    prob = P.FwiProblem(eng, d_obs, components=comps, mod_par=1, grad_mask=grad_mask, data_mask=data_mask)
    phi, grad, _ = prob.evaluate(torch.tensor(vvr, dtype=dtype, device="cuda"))
    tol = 1e-4 if dtype == torch.float32 else 1e-10
    assert abs(phi - phi_ref) / phi_ref < tol
    assert rel_l2(grad.double().cpu().numpy(), ref) < tol, rel_l2(grad.double().cpu().numpy(), ref)
    assert prob.n_forward == 1 and prob.n_gradient == 1


def test_lbfgs_small_inversion_lowers_misfit_within_bounds():
    import torch
    from elastic_isotropic_operators_amd import fwi_problem as P
    w, g, m0, dtrue = fwi_case(nts=400, dts=0.001, nExp=3, nz=48, nx=72)
    sz, sx, rz, rx, src = w["src_z"], w["src_x"], w["rec_z"], w["rec_x"], w["model"]
    m_true = m0 + dtrue
    d_obs = np.stack([O.nl_fwd_shot(g, m_true, sz[e], sx[e], rz, rx, src) for e in range(3)])
    vvr0 = su.vpvsrho_to_rholamemu(m0, -1)
    vvr_true = su.vpvsrho_to_rholamemu(m_true, -1)
    lo = np.stack([np.full(m0.shape[1:], v) for v in (0.9 * vvr0[0].min(), 0.9 * vvr0[1].min(),
                                                          0.95 * vvr0[2].min())])
    hi = np.stack([np.full(m0.shape[1:], v) for v in (1.1 * vvr0[0].max(), 1.1 * vvr0[1].max(),
                                                          1.05 * vvr0[2].max())])
    eng = _engine(w, g, torch.float32)
    prob = P.FwiProblem(eng, d_obs, mod_par=1, min_bound=lo, max_bound=hi, dtw=eng.dtw,
                        dz=w["dz"], dx=w["dx"])
    m = torch.tensor(vvr0, dtype=torch.float32, device="cuda")
    res = P.lbfgs(prob, m, 4, step_frac=0.005)
    print("objective", res.objective, res.stop_reason, "forwards", res.n_forward, "gradients", res.n_gradient)
    assert res.iterations == 4
    assert res.objective[-1] < 0.5 * res.objective[0]
    assert all(b < a for a, b in zip(res.objective, res.objective[1:]))
    mf = res.model.double().cpu().numpy()
    assert (mf >= lo - 1e-3).all() and (mf <= hi + 1e-3).all()
    # the vp update correlates with the true vp perturbation
    dv = (mf - vvr0)[0].ravel()
    dt = (vvr_true - vvr0)[0].ravel()
    assert float(dv @ dt) > 0.0
    assert res.n_gradient == res.iterations + 1


def test_efwi_driver_runs_and_writes_outputs(tmp_path):
    """The efwiMain-style driver: parfile geometry, SEP files in and out, two L-BFGS
    iterations; its starting objective equals the FP64 oracle's misfit."""
    from elastic_isotropic_operators_amd import efwi
    from elastic_isotropic_operators_amd.sep import Hypercube, SepVector, genericIO
    from helpers import operator_case
    w, g, m0, dtrue = fwi_case(nts=300, dts=0.001, nExp=2, nz=48, nx=72)
    args, c = operator_case(tmp_path, w, nExp=2, zSource=3, depthRec=1)
    comps = "vx,vz,p"
    d_full = np.stack([O.nl_fwd_shot(g, m0 + dtrue, c["sz"][e], c["sx"][e], c["rz"], c["rx"], w["model"])
                       for e in range(2)])
    d_obs = su.datcomp_fwd(comps, d_full)
    hyp = Hypercube.hypercube(axes=[Hypercube.axis(n=w["nts"], d=w["dts"]), Hypercube.axis(n=d_obs.shape[2]),
                                    Hypercube.axis(n=3), Hypercube.axis(n=2)])
    dv = SepVector.getSepVector(hyp)
    dv.getNdArray()[...] = d_obs
    genericIO.defaultIO.writeVector(str(tmp_path / "data.H"), dv)
    F0 = np.stack([O.nl_fwd_shot(g, m0, c["sz"][e], c["sx"][e], c["rz"], c["rx"], w["model"]) for e in range(2)])
    r0 = su.datcomp_fwd(comps, F0) - d_obs
    phi_ref = 0.5 * float(np.sum(r0 * r0))
    folder = tmp_path / "inv"
    res = efwi.main(args + ["data=%s" % (tmp_path / "data.H"), "comp=%s" % comps, "nIter=2",
                            "folder=%s" % folder, "solver=lbfgs", "iterSampling=1", "pyinfo=0",
                            "minBound_par2=500.0"])
    obj = genericIO.defaultIO.getVector(str(folder / "inv_obj.H")).getNdArray()
    assert obj.shape == (3,) and obj[2] < obj[1] < obj[0]
    assert abs(obj[0] - phi_ref) / phi_ref < 1e-4
    mod = genericIO.defaultIO.getVector(str(folder / "inv_model.H"), ndims=4).getNdArray()
    assert mod.shape == (3, 3, w["nx"], w["nz"])
    np.testing.assert_allclose(mod[-1], res.model.cpu().numpy(), rtol=1e-6)
    grad = genericIO.defaultIO.getVector(str(folder / "inv_grad.H"), ndims=4).getNdArray()
    assert grad.shape == mod.shape and np.isfinite(grad).all()
    resid = genericIO.defaultIO.getVector(str(folder / "inv_res.H"), ndims=4).getNdArray()
    assert resid.shape == d_obs.shape
    assert (folder / "inv_logFile").read_text().count("objective") == 3


def test_engine_with_sinc_receivers_matches_oracle():
    """The torch FWI engine with sinc sources and gaussian receivers records the oracle's data."""
    import torch
    from helpers import oracle_nl_irregular
    w, g, m0, _ = fwi_case(nts=200, dts=0.001, nExp=2, nz=48, nx=72)
    from elastic_isotropic_operators_amd.fwi import ElasticFwi
    grid = dict(nz=w["nz"], nx=w["nx"], dz=w["dz"], dx=w["dx"], oz=w["oz"], ox=w["ox"], nts=w["nts"],
                dts=w["dts"], sub=w["sub"], minPad=g.minPad, alphaCos=w["alphaCos"])
    eng = ElasticFwi(grid, w["src_z"], w["src_x"], w["model"], w["rec_z"], w["rec_x"], dtype=torch.float64,
                     src_interp=("sinc", 2), rec_interp=("gauss", 2))
    d = eng.data(torch.tensor(m0, dtype=torch.float64, device="cuda")).cpu().numpy()
    for e in range(2):
        ref = oracle_nl_irregular(w, g, w["src_z"][e], w["src_x"][e], w["rec_z"], w["rec_x"], w["model"],
                                  ("sinc", 2), ("gauss", 2))
        assert rel_l2(d[e], ref) < 1e-6, rel_l2(d[e], ref)
