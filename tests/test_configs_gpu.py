"""Each configuration's workload at its own size on the HIP path, against the FP64 oracle or
the FP64 GPU path (itself pinned to the oracle to 1e-11 on the small cases of the other test
files).

C0  NL forward, 280x472 padded Gaussian-anomaly model, nts 1000, through the float operator
    (parfile -> nonlinearOpInitFloat -> pybind -> C ABI -> two-step kernel), against the FP64
    oracle's shot pipeline: data within 1e-5 relative L2.
C1  Born forward and adjoint, 8 shots on the C0 grid, through the double operator against the
    FP64 oracle (1e-11), and the FP32 / FP64 Born dot tests at that size.
C2  the 64-shot uneven layered batch in one batched propagation (two fine steps per launch,
    tile width chosen per call): the shortest and the longest shot against the FP64 oracle at
    full nts (1e-5), and every shot against its FP64-GPU twin (1e-5).
C3  in tests/test_fwi_gpu.py (full-size checkpointed gradient, sqrt(nts) checkpoints).
C4  the learned-boundary kernel on the 1028x3028 grid against the BF16 torch network, and the
    plain 10- and 60-cell sponges' reflection ordering on the 1000x3000 model, nts 4000.
"""
import sys

import numpy as np
import pytest

import oracle as O
from elastic_isotropic_operators_amd import _capi, geometry as G, workloads as W

from helpers import GRID_OF_COMP, operator_case, oracle_grid, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ C0
def test_c0_nl_forward_vs_fp64_oracle(tmp_path):
    from elastic_isotropic_operators_amd import Elastic_iso_float_prop as PF
    O.set_ftz(True)  # FP64 denormals flushed on the oracle threads (values below 2.2e-308)
    w = W.gaussian_c0()
    args, geo = operator_case(tmp_path, w, nExp=1, zSource=3, xSource=201, nRec=200, depthRec=1, oRec=1, dRec=2)
    (model, data, elastic, par, sc, sx, sz, sxz, rc, rx, rz, rxz, _) = PF.nonlinearOpInitFloat(args)
    op = PF.nonlinearPropElasticShotsGpu(model, data, elastic, par, sc, sx, sz, sxz, rc, rx, rz, rxz)
    model.getNdArray()[0] = w["model"]
    op.forward(False, model, data)
    ref = O.nl_fwd_shot(oracle_grid(w), w["elastic"], geo["sz"][0], geo["sx"][0], geo["rz"], geo["rx"], w["model"])
    got = data.getNdArray()[0]
    err = rel_l2(got, ref)
    per = [rel_l2(got[k], ref[k]) for k in range(5)]
    _say("C0 NL forward FP32 vs FP64 oracle: rel L2 %.3e (per component %s)" % (err, ["%.1e" % e for e in per]))
    assert np.linalg.norm(ref) > 0
    assert err < 1e-5, (err, per)


# ------------------------------------------------------------------ C1
def _c1_born(tmp_path, P, init):
    w = W.gaussian_c1()
    args, geo = operator_case(tmp_path, w, nExp=8, zSource=3, xSource=26, spacing=50, nRec=200,
                              depthRec=1, oRec=1, dRec=2)
    (model, data, elastic, par, sig, sc, sx, sz, sxz, rc, rx, rz, rxz, _) = init(args)
    op = P.BornElasticShotsGpu(model, data, elastic, par, sig, sc, sx, sz, sxz, rc, rx, rz, rxz)
    return w, geo, op, model, data


def test_c1_born_fp64_vs_fp64_oracle(tmp_path):
    """Workload C1 through the double Born operator: data of the perturbation and the image of
    that data against the FP64 oracle at 1e-11."""
    from elastic_isotropic_operators_amd import Elastic_iso_double_prop as PD
    O.set_ftz(False)
    w, geo, op, model, data = _c1_born(tmp_path, PD, PD.BornOpInitDouble)
    dm = w["dm"]
    model.getNdArray()[...] = dm
    op.forward(False, model, data)
    g = oracle_grid(w)
    ref = np.stack([O.born_fwd_shot(g, w["elastic"], geo["sz"][e], geo["sx"][e], w["model"], geo["rz"], geo["rx"], dm)
                    for e in range(8)])
    err = rel_l2(data.getNdArray(), ref)
    _say("C1 Born forward FP64 vs FP64 oracle: rel L2 %.3e" % err)
    assert np.linalg.norm(ref) > 0
    assert err < 1e-11, err
    img = model.clone()
    op.adjoint(False, img, data)
    ref_img = O.born_adj_shots(g, w["elastic"], geo["sz"], geo["sx"], w["model"], geo["rz"], geo["rx"],
                               data.getNdArray().astype(np.float64))
    err = rel_l2(img.getNdArray(), ref_img)
    _say("C1 Born adjoint FP64 vs FP64 oracle: rel L2 %.3e" % err)
    assert err < 1e-11, err


@pytest.mark.parametrize("prec", ["float", "double"])
def test_c1_born_dot_test_at_size(tmp_path, prec):
    """Operator::dotTest of the Born operator on workload C1 (8 shots, 280x472, nts 1000):
    1e-5 in FP32, 1e-12 in FP64, with and without add."""
    if prec == "float":
        from elastic_isotropic_operators_amd import Elastic_iso_float_prop as P
        init, tol = P.BornOpInitFloat, 1e-5
    else:
        from elastic_isotropic_operators_amd import Elastic_iso_double_prop as P
        init, tol = P.BornOpInitDouble, 1e-12
    w, geo, op, model, data = _c1_born(tmp_path, P, init)
    e1, e2 = op.dotTest(False, tol)
    _say("C1 Born dot test %s: %.2e %.2e" % (prec, e1, e2))
    assert e1 < tol and e2 < tol, (e1, e2)


# ------------------------------------------------------------------ C2
_C2 = {}


def _c2_shots(ctx, shots_w, dtype):
    """Upload the batch as bench.py does (FP32 or FP64 media and scaled regular source traces on
    the device, receiver buffers); returns (shots, keep, per-shot (receiver interps, device
    receiver buffers))."""
    shots, keep, outs = [], [], []
    isz = np.dtype(dtype).itemsize
    for w in shots_w:
        sub_fd = 2 * w["sub"]
        dtw = G.dtw_of(w["dts"], w["sub"])
        media = G.scaled_media(w["elastic"], dtw, dtype)
        mid = ctx.model_add(w["nz"], w["nx"], w["nts"], sub_fd, w["dz"], w["dx"], w["dts"], W.min_pad(w), media,
                            w["alphaCos"])
        si = G.staggered_interps(w["nz"], w["oz"], w["dz"], w["ox"], w["dx"], w["src_z"][0], w["src_x"][0])
        ri = G.staggered_interps(w["nz"], w["oz"], w["dz"], w["ox"], w["dx"], w["rec_z"], w["rec_x"])
        src = G.nl_source_traces(si, media, w["model"], dtw, w["dx"], w["dz"], dtype)
        spos = [ip.uniq for ip in si]
        rpos = [ip.uniq for ip in ri]
        dsrc, drec = [], []
        for k in range(5):
            p = ctx.alloc(src[k].nbytes)
            ctx.h2d(p, src[k])
            dsrc.append(p)
            drec.append(ctx.alloc(rpos[GRID_OF_COMP[k]].size * w["nts"] * isz))
        shot, kp = ctx.make_shot(mid, spos, rpos, dsrc, drec)
        shots.append(shot)
        keep.append((kp, spos, rpos, dsrc, drec))
        outs.append((ri, rpos, drec))
    return shots, keep, outs


def _c2_free(ctx, keep):
    for kp, spos, rpos, dsrc, drec in keep:
        for p in dsrc + drec:
            ctx.free(p)


def _c2_read(ctx, w, out, dtype):
    ri, rpos, drec = out
    reg = []
    for k in range(5):
        a = np.empty((rpos[GRID_OF_COMP[k]].size, w["nts"]), dtype)
        ctx.d2h(a, drec[k])
        reg.append(a)
    return reg


def _c2_fp32():
    """The FP32 batch, propagated once per test session: regular traces of every shot."""
    if "f32" not in _C2:
        shots_w = W.layered_batch_c2()
        ctx = _capi.Context(0, np.float32)
        ctx.set_time_block(2, _capi.VARIANT_AUTO)
        shots, keep, outs = _c2_shots(ctx, shots_w, np.float32)
        _say("C2: 64 shots uploaded (FP32)")
        ctx.nl_forward(shots, device_ptrs=True)
        _say("C2: FP32 batch propagated in %.1f ms of device time" % ctx.last_timing()[0])
        regs = [_c2_read(ctx, w, o, np.float32) for w, o in zip(shots_w, outs)]
        _c2_free(ctx, keep)
        ctx.close()
        _C2["w"], _C2["f32"] = shots_w, regs
    return _C2["w"], _C2["f32"]


@pytest.mark.parametrize("which", ["shortest", "longest"])
def test_c2_batch_shot_vs_fp64_oracle(which):
    """One shot of the 64-shot batch, propagated with the other 63 in one batched time loop,
    against the FP64 oracle's shot pipeline (interpolation, scaling, propagation, receiver
    interpolation) at its full trace length."""
    shots_w, regs = _c2_fp32()
    nts = [w["nts"] for w in shots_w]
    i = int(np.argmin(nts)) if which == "shortest" else int(np.argmax(nts))
    w = shots_w[i]
    ri = G.staggered_interps(w["nz"], w["oz"], w["dz"], w["ox"], w["dx"], w["rec_z"], w["rec_x"])
    got = np.stack([ri[GRID_OF_COMP[k]].forward(regs[i][k].astype(np.float64)) for k in range(5)])
    O.set_ftz(True)  # FP64 denormals flushed on the oracle threads (values below 2.2e-308)
    _say("C2 %s shot %d: %dx%d, nts %d: oracle running" % (which, i, w["nz"], w["nx"], w["nts"]))
    ref = O.nl_fwd_shot(oracle_grid(w), w["elastic"], w["src_z"][0], w["src_x"][0], w["rec_z"], w["rec_x"],
                        w["model"])
    err = rel_l2(got, ref)
    per = [rel_l2(got[k], ref[k]) for k in range(5)]
    _say("C2 %s shot %d vs FP64 oracle: rel L2 %.3e (per component %s)" % (which, i, err, ["%.1e" % e for e in per]))
    assert np.linalg.norm(ref) > 0
    assert err < 1e-5, (err, per)


def test_c2_every_shot_vs_fp64_gpu_twin():
    """All 64 shots of the FP32 batch against the same batch propagated in FP64 on the GPU."""
    shots_w, regs = _c2_fp32()
    ctx = _capi.Context(0, np.float64)
    shots, keep, outs = _c2_shots(ctx, shots_w, np.float64)
    ctx.nl_forward(shots, device_ptrs=True)
    _say("C2: FP64 batch propagated in %.1f ms of device time" % ctx.last_timing()[0])
    worst = 0.0
    for i, (w, o) in enumerate(zip(shots_w, outs)):
        ref = _c2_read(ctx, w, o, np.float64)
        err = rel_l2(np.concatenate([r.ravel() for r in regs[i]]), np.concatenate([r.ravel() for r in ref]))
        worst = max(worst, err)
        assert err < 1e-5, (i, err)
    _c2_free(ctx, keep)
    ctx.close()
    _say("C2: worst shot FP32 vs FP64 GPU rel L2 %.3e" % worst)


# ------------------------------------------------------------------ C4
def test_c4_learned_boundary_kernel_at_size():
    """The learned-boundary kernel with the committed network on the 1028x3028 grid (the
    1000x3000 interior, a 10-cell strip and the frame): one step of two random states; update
    cells equal the plain step's, frame cells equal the BF16 torch network on them."""
    import os
    import torch
    from elastic_isotropic_operators_amd import abc_learn as A
    path = os.path.join(os.path.dirname(A.__file__), "data", "abc_strip10.npz")
    net, scale, strip, on = A.load(path)
    vvr, nts = W.learned_abc_c4()
    w = A.padded(vvr, strip)
    nz, nx = w["nz"], w["nx"]
    assert (nz, nx) == (1028, 3028)
    ctx = _capi.Context(0, np.float32)
    dtw = G.dtw_of(0.001, 2)
    media = G.scaled_media(w["elastic"], dtw, np.float32)
    mid = ctx.model_add(nz, nx, nts, 4, w["dz"], w["dx"], 0.001, w["minPad"], media, 0.99)
    g = torch.Generator(device="cuda").manual_seed(5)
    st = torch.randn((2, 2, 5, nx, nz), device="cuda", generator=g)
    st *= torch.tensor(scale[:5], device="cuda").view(1, 1, 5, 1, 1)  # network inputs of order 1
    plain = st.clone()
    ctx.step_states(mid, 2, plain.data_ptr())
    ctx.set_abc(mid, on, net.flat(), scale)
    learned = st.clone()
    ctx.step_states(mid, 2, learned.data_ptr())
    torch.cuda.synchronize()
    F = A.FAT
    assert torch.equal(learned[:, 1], plain[:, 1])
    assert torch.equal(learned[:, 0, :, F:nx - F, F:nz - F], plain[:, 0, :, F:nx - F, F:nz - F])
    reg = A.frame_regions(nz, nx)
    sc = torch.tensor(scale, device="cuda")
    depth = on // 10 if on // 10 else F
    worst = 0.0
    for i in range(2):
        feats = A.features(plain[i, 0], plain[i, 1], sc, reg, diff=(on % 10 == 1))
        with torch.no_grad():
            for e in range(4):
                ref = (net.edge_forward(e, feats[e][None], bf16=True)[0] * sc[:5].view(5, 1, 1)).double()
                got = A.region(learned[i, 0], reg)[e].double()
                err = float((got - ref).norm() / ref.norm())
                worst = max(worst, err)
                assert err < 2e-3, (i, e, err)
    _say("C4 learned-boundary kernel vs BF16 torch network: worst frame rel L2 %.2e (depth %d)" % (worst, depth))
    ctx.close()


def test_c4_plain_sponge_reflection_ordering():
    """Full C4 run (1000x3000 model, nts 4000): the wide 120-cell sponge is the reference; the
    plain 10-cell sponge reflects more than the plain 60-cell one, and both reflect."""
    from elastic_isotropic_operators_amd import abc_learn as A
    vvr, nts = W.learned_abc_c4()
    res = A.reflection_test(vvr, None, None, strip=10, mid=60, wide=120, nts=nts, rec_every=4,
                            variants=("plain", "mid"), log=_say)
    assert res["wide120"]["refl"] == 0.0
    assert res["plain10"]["refl"] > res["plain60"]["refl"] > 0, res
