"""GPU tests of the checkpointed FWI gradient.

C-ABI level: eio_fwi_forward records the same data as the NL forward, and eio_fwi_gradient
(Born adjoint with segment recomputation from checkpoints into a ring of coarse slices) equals
the oracle's full-storage Born adjoint, for segment lengths that exercise every ring case.
torch level: ElasticFwi's autograd gradient against the FP64 oracle gradient (<= 1e-4 relative
L2 in FP32) and against finite differences of the objective on 10 directions.
"""
import numpy as np
import pytest

import oracle as O
from elastic_isotropic_operators_amd import _capi

from helpers import (GRID_OF_COMP, fwi_case, fwi_directions, oracle_grid, random_traces, reg_positions, rel_l2,
                     scattered_positions)

pytestmark = pytest.mark.gpu

TOL = {np.float32: 1e-5, np.float64: 1e-11}


def _reg_case(nz=40, nx=56, nts=120, sub=2, seed=40, nshot=1):
    from elastic_isotropic_operators_amd import workloads as W
    w = W.small(nz=nz, nx=nx, nts=nts, sub=sub, seed=seed, nExp=nshot)
    g = oracle_grid(w)
    med = O.media(g, w["elastic"], np.float64)
    spos = [reg_positions(w, w["src_z"][e], w["src_x"][e]) for e in range(nshot)]
    rpos = reg_positions(w, w["rec_z"], w["rec_x"])
    return w, g, med, spos, rpos


def _zeros_like_traces(pos, nts, dtype):
    return [np.zeros((pos[GRID_OF_COMP[k]].size, nts), dtype) for k in range(5)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("segment", [0, 1, 2, 5, 118, 119, 500])
def test_checkpointed_gradient_matches_full_storage_oracle(dtype, segment):
    """nts = 120 (119 coarse intervals): segment 0 -> 11 (10 checkpoints, a short last
    segment), 1 and 2 (ring of 4/5 slices), 118 (one interval in the last segment), 119 and
    500 (one segment, no checkpoint)."""
    w, g, med, spos, rpos = _reg_case()
    spos = spos[0]
    rng = np.random.default_rng(41)
    src = random_traces(rng, spos, g.nts, np.float64)
    res = random_traces(rng, rpos, g.nts, np.float64)
    ref_d = O.nl_fwd_reg(g, med, spos, src, rpos, np.float64)
    src_fine = [O.time_interp(s, g.dts, g.sub) for s in src]
    ref_g = O.born_adj_reg(g, med, spos, src_fine, rpos, res, np.float64)

    ctx = _capi.Context(0, dtype)
    mid = ctx.model_add(g.nz, g.nx, g.nts, g.sub, g.dz, g.dx, g.dts, g.minPad, med, g.alphaCos)
    srcc = [np.ascontiguousarray(s, dtype) for s in src]
    rec = _zeros_like_traces(rpos, g.nts, dtype)
    shot, keep = ctx.make_shot(mid, spos, rpos, srcc, rec)
    h = ctx.fwi_forward([shot], segment)
    L, nck, nbytes = ctx.fwi_info(h)
    Lexp = min(segment if segment > 0 else 11, g.nts - 1)
    assert L == Lexp and nck == -(-(g.nts - 1) // Lexp) - 1 and nbytes > 0
    for k in range(5):
        assert rel_l2(rec[k], ref_d[k]) < TOL[dtype], (k, rel_l2(rec[k], ref_d[k]))

    resc = [np.ascontiguousarray(r, dtype) for r in res]
    shot2, keep2 = ctx.make_shot(mid, spos, rpos, srcc, resc)
    out = np.zeros((5, g.nx, g.nz), dtype)
    ctx.fwi_gradient(h, [shot2], out)
    assert rel_l2(out, ref_g) < 10 * TOL[dtype], rel_l2(out, ref_g)
    # the store can be used again and gives the same gradient
    out2 = np.zeros((5, g.nx, g.nz), dtype)
    ctx.fwi_gradient(h, [shot2], out2)
    assert np.array_equal(out, out2)
    ctx.fwi_release(h)
    with pytest.raises(_capi.EioError, match="unknown FWI store"):
        ctx.fwi_gradient(h, [shot2], out2)


@pytest.mark.parametrize("variant2", [4, 10, 11])
@pytest.mark.parametrize("segment,sub", [(1, 2), (2, 1), (5, 2), (118, 1), (0, 2), (119, 2)])
def test_two_step_checkpointed_forward_and_recompute_match_oracle(segment, sub, variant2):
    """FP32 with two fine steps per launch everywhere: the checkpointed forward (checkpoints
    taken from the level pair that holds the state at each segment start), the two-step
    segment recompute (no receiver recording) and the two-step receiver pass, against the FP64
    oracle's data and full-storage Born adjoint; every ring case (segment 1: four slices) and
    sub_fd 2 and 4; the 56 x 32 row-lane forward (variant 4) and the 56 x 24 quad-lane one
    (variant 11, whose last segment's decimating forward runs the row-lane kernel of its
    width)."""
    w, g, med, spos, rpos = _reg_case(sub=sub)
    spos = spos[0]
    rng = np.random.default_rng(141)
    src = random_traces(rng, spos, g.nts, np.float64)
    res = random_traces(rng, rpos, g.nts, np.float64)
    ref_d = O.nl_fwd_reg(g, med, spos, src, rpos, np.float64)
    src_fine = [O.time_interp(s, g.dts, g.sub) for s in src]
    ref_g = O.born_adj_reg(g, med, spos, src_fine, rpos, res, np.float64)
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, variant2)
    ctx.set_adjoint_time_block(2, 7)
    mid = ctx.model_add(g.nz, g.nx, g.nts, g.sub, g.dz, g.dx, g.dts, g.minPad, med, g.alphaCos)
    srcc = [np.ascontiguousarray(x, np.float32) for x in src]
    rec = _zeros_like_traces(rpos, g.nts, np.float32)
    shot, keep = ctx.make_shot(mid, spos, rpos, srcc, rec)
    h = ctx.fwi_forward([shot], segment)
    for k in range(5):
        assert rel_l2(rec[k], ref_d[k]) < 1e-5, (k, rel_l2(rec[k], ref_d[k]))
    shot2, keep2 = ctx.make_shot(mid, spos, rpos, srcc, [np.ascontiguousarray(r, np.float32) for r in res])
    out = np.zeros((5, g.nx, g.nz), np.float32)
    ctx.fwi_gradient(h, [shot2], out)
    assert rel_l2(out, ref_g) < 1e-4, rel_l2(out, ref_g)
    out2 = np.zeros((5, g.nx, g.nz), np.float32)  # the store used again: its ring is recomputed
    ctx.fwi_gradient(h, [shot2], out2)
    assert np.array_equal(out, out2)
    ctx.fwi_release(h)


@pytest.mark.parametrize("variant2", [4, 0, 9, 10, 11, 12])
def test_two_step_forward_with_whole_record_matches_oracle(variant2):
    """FP32, one segment (the whole coarse record): the forward runs the two-step kernel with
    the decimation fused; data and gradient against the FP64 oracle's full-storage Born
    adjoint, at sub_fd 4 (two launches per coarse interval)."""
    w, g, med, spos, rpos = _reg_case(sub=2, nshot=2)
    rng = np.random.default_rng(47)
    srcs = [random_traces(rng, spos[e], g.nts, np.float64) for e in range(2)]
    res = [random_traces(rng, rpos, g.nts, np.float64) for _ in range(2)]
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, variant2)
    mid = ctx.model_add(g.nz, g.nx, g.nts, g.sub, g.dz, g.dx, g.dts, g.minPad, med, g.alphaCos)
    shots, keep, recs = [], [], []
    for e in range(2):
        rec = _zeros_like_traces(rpos, g.nts, np.float32)
        s, kp = ctx.make_shot(mid, spos[e], rpos, [np.ascontiguousarray(x, np.float32) for x in srcs[e]], rec)
        shots.append(s)
        keep.append(kp)
        recs.append(rec)
    h = ctx.fwi_forward(shots, g.nts - 1)
    assert ctx.fwi_info(h)[1] == 0
    ref_g = np.zeros((5, g.nx, g.nz))
    for e in range(2):
        ref_d = O.nl_fwd_reg(g, med, spos[e], srcs[e], rpos, np.float64)
        for k in range(5):
            assert rel_l2(recs[e][k], ref_d[k]) < 1e-5
        src_fine = [O.time_interp(x, g.dts, g.sub) for x in srcs[e]]
        ref_g += np.asarray(O.born_adj_reg(g, med, spos[e], src_fine, rpos, res[e], np.float64))
    shots2 = []
    for e in range(2):
        s, kp = ctx.make_shot(mid, spos[e], rpos, [np.ascontiguousarray(x, np.float32) for x in srcs[e]],
                              [np.ascontiguousarray(x, np.float32) for x in res[e]])
        shots2.append(s)
        keep.append(kp)
    dm = np.zeros((5, g.nx, g.nz), np.float32)
    ctx.fwi_gradient(h, shots2, dm)
    ctx.fwi_release(h)
    assert rel_l2(dm, ref_g) < 1e-4, rel_l2(dm, ref_g)


@pytest.mark.parametrize("va,sub,segment", [(0, 1, 0), (1, 2, 59), (2, 1, 59), (3, 2, 0), (1, 1, 0), (4, 1, 0),
                                             (4, 2, 59), (5, 2, 0), (6, 1, 59), (7, 2, 59),
                                             (8, 1, 0), (9, 2, 59), (10, 2, 59), (11, 1, 0), (10, 1, 0)])
def test_two_step_adjoint_gradient_matches_oracle(va, sub, segment):
    """FP32 FWI gradient with two reverse fine steps per launch of the receiver pass, over the
    checkpointed store (segment 0 -> 8: recomputed segments, one-step tiles) and the whole
    record (59): two shots, residuals at 200 receivers per grid scattered over several tiles,
    against the FP64 oracle's full-storage Born adjoint and against one step per launch.
    sub = 1 -> sub_fd 2 (one pair per coarse interval), 2 -> sub_fd 4 (two pairs)."""
    from elastic_isotropic_operators_amd import workloads as W
    w = W.small(nz=130, nx=120, nts=60, sub=sub, seed=90)
    g = oracle_grid(w)
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(91 + va)
    spos = [scattered_positions(rng, g.nz, g.nx, 3) for _ in range(2)]
    rpos = scattered_positions(rng, g.nz, g.nx, 200)
    srcs = [random_traces(rng, spos[e], g.nts, np.float64) for e in range(2)]
    res = [random_traces(rng, rpos, g.nts, np.float64) for _ in range(2)]
    ref_g = np.zeros((5, g.nx, g.nz))
    for e in range(2):
        src_fine = [O.time_interp(x, g.dts, g.sub) for x in srcs[e]]
        ref_g += np.asarray(O.born_adj_reg(g, med, spos[e], src_fine, rpos, res[e], np.float64))
    ctx = _capi.Context(0, np.float32)
    mid = ctx.model_add(g.nz, g.nx, g.nts, g.sub, g.dz, g.dx, g.dts, g.minPad, med, g.alphaCos)
    keep, shots, shots2 = [], [], []
    for e in range(2):
        s, kp = ctx.make_shot(mid, spos[e], rpos, [np.ascontiguousarray(x, np.float32) for x in srcs[e]],
                              _zeros_like_traces(rpos, g.nts, np.float32))
        shots.append(s)
        keep.append(kp)
        s, kp = ctx.make_shot(mid, spos[e], rpos, [np.ascontiguousarray(x, np.float32) for x in srcs[e]],
                              [np.ascontiguousarray(x, np.float32) for x in res[e]])
        shots2.append(s)
        keep.append(kp)
    h = ctx.fwi_forward(shots, segment)
    one = np.zeros((5, g.nx, g.nz), np.float32)
    ctx.fwi_gradient(h, shots2, one)
    ctx.set_adjoint_time_block(2, va)
    two = np.zeros((5, g.nx, g.nz), np.float32)
    ctx.fwi_gradient(h, shots2, two)
    ctx.fwi_release(h)
    assert rel_l2(two, ref_g) < 1e-4, rel_l2(two, ref_g)
    assert rel_l2(two, one) < 1e-5, rel_l2(two, one)


def test_checkpointed_gradient_several_shots():
    """Three shots in one batch: stacked gradient equals the sum of per-shot oracle adjoints."""
    dtype = np.float32
    w, g, med, spos, rpos = _reg_case(nz=48, nx=72, nts=90, sub=1, seed=42, nshot=3)
    rng = np.random.default_rng(43)
    ctx = _capi.Context(0, dtype)
    mid = ctx.model_add(g.nz, g.nx, g.nts, g.sub, g.dz, g.dx, g.dts, g.minPad, med, g.alphaCos)
    shots, keep, resid, ref = [], [], [], np.zeros((5, g.nx, g.nz))
    srcs = []
    for e in range(3):
        src = random_traces(rng, spos[e], g.nts, np.float64)
        res = random_traces(rng, rpos, g.nts, np.float64)
        ref += O.born_adj_reg(g, med, spos[e], [O.time_interp(s, g.dts, g.sub) for s in src], rpos, res)
        srcc = [np.ascontiguousarray(s, dtype) for s in src]
        rec = _zeros_like_traces(rpos, g.nts, dtype)
        s, kp = ctx.make_shot(mid, spos[e], rpos, srcc, rec)
        shots.append(s)
        keep.append((kp, srcc, rec))
        resid.append([np.ascontiguousarray(r, dtype) for r in res])
        srcs.append(srcc)
    h = ctx.fwi_forward(shots, 4)
    shots2 = []
    for e in range(3):
        s, kp = ctx.make_shot(mid, spos[e], rpos, srcs[e], resid[e])
        shots2.append(s)
        keep.append(kp)
    out = np.zeros((5, g.nx, g.nz), dtype)
    ctx.fwi_gradient(h, shots2, out)
    assert rel_l2(out, ref) < 1e-4, rel_l2(out, ref)
    # a gradient call with other shots than the store's is an argument error
    with pytest.raises(_capi.EioError, match="differ"):
        ctx.fwi_gradient(h, shots2[:2], out)
    ctx.fwi_release(h)


# ------------------------------------------------------------------ torch layer
def _engine(w, g, dtype, segment=0):
    import torch
    from elastic_isotropic_operators_amd.fwi import ElasticFwi
    grid = dict(nz=w["nz"], nx=w["nx"], dz=w["dz"], dx=w["dx"], oz=w["oz"], ox=w["ox"], nts=w["nts"],
                dts=w["dts"], sub=w["sub"], minPad=g.minPad, alphaCos=w["alphaCos"])
    return ElasticFwi(grid, w["src_z"], w["src_x"], w["model"], w["rec_z"], w["rec_x"], dtype=dtype,
                      segment=segment)


def _oracle_gradient(w, g, m0, dobs):
    sz, sx, rz, rx, src = w["src_z"], w["src_x"], w["rec_z"], w["rec_x"], w["model"]
    res = np.stack([O.nl_fwd_shot(g, m0, sz[e], sx[e], rz, rx, src) for e in range(sz.shape[0])]) - dobs
    return O.born_adj_shots(g, m0, sz, sx, src, rz, rx, res), res


@pytest.mark.parametrize("tdtype", ["float32", "float64"])
def test_autograd_gradient_vs_fp64_oracle(tdtype):
    import torch
    dtype = getattr(torch, tdtype)
    w, g, m0, dtrue = fwi_case(nts=400, dts=0.001, nExp=2, nz=48, nx=72)
    sz, sx, rz, rx, src = w["src_z"], w["src_x"], w["rec_z"], w["rec_x"], w["model"]
    dobs = np.stack([O.nl_fwd_shot(g, m0 + dtrue, sz[e], sx[e], rz, rx, src) for e in range(2)])
    ref, res_ref = _oracle_gradient(w, g, m0, dobs)
    eng = _engine(w, g, dtype)
    m = torch.tensor(m0, dtype=dtype, device="cuda", requires_grad=True)
    d_obs = torch.tensor(dobs, dtype=dtype, device="cuda")
    loss = eng.misfit(m, d_obs)
    loss.backward()
    grad = m.grad.double().cpu().numpy()
    tol = 1e-4 if dtype == torch.float32 else 1e-10
    assert rel_l2(grad, ref) < tol, rel_l2(grad, ref)
    assert abs(loss.item() - 0.5 * np.sum(res_ref ** 2)) / (0.5 * np.sum(res_ref ** 2)) < tol * 10


def test_autograd_gradient_finite_differences():
    """FP64 GPU: central finite differences of the objective on 10 smooth random directions
    against <gradient, direction>; the bar is the O(dts^2) coarse-time gap of the Born
    operator that the oracle test pins (< 2 % at dts = 0.5 ms)."""
    import torch
    w, g, m0, dtrue = fwi_case(nts=800, dts=0.0005)
    eng = _engine(w, g, torch.float64)
    mt = torch.tensor(m0 + dtrue, device="cuda")
    with torch.no_grad():
        d_obs, _, _ = eng.forward_data(mt)
    m = torch.tensor(m0, device="cuda", requires_grad=True)
    eng.misfit(m, d_obs).backward()
    grad = m.grad.cpu().numpy()

    def J(x):
        with torch.no_grad():
            return float(eng.misfit(torch.tensor(x, device="cuda"), d_obs))

    fd, ad = [], []
    for v in fwi_directions(m0, 10):
        eps = 1e-2
        fd.append((J(m0 + eps * v) - J(m0 - eps * v)) / (2 * eps))
        ad.append(np.sum(grad * v))
    fd, ad = np.array(fd), np.array(ad)
    gap = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
    corr = fd @ ad / np.linalg.norm(fd) / np.linalg.norm(ad)
    assert gap < 0.02 and corr > 0.9999, (gap, corr)


def test_autograd_chain_rule_through_conversion_and_datcomp():
    """Gradient with respect to (vp, vs, rho) through the torch conversion and the vx,vz,p
    component selection equals ElasticConvJab.adjoint of the (rho, lambda, mu) gradient of the
    same objective (elasticParamConvertModule.py:91-108, dataCompModule.py:79-107)."""
    import torch
    from elastic_isotropic_operators_amd import seis_utils as su
    from elastic_isotropic_operators_amd.fwi import datcomp, vpvsrho_to_rholamemu
    w, g, m0, dtrue = fwi_case(nts=300, dts=0.001)
    eng = _engine(w, g, torch.float64, segment=7)
    vvr = torch.tensor(w["vvr"].astype(np.float64), device="cuda", requires_grad=True)
    with torch.no_grad():
        d_obs = datcomp(eng.forward_data(vpvsrho_to_rholamemu(vvr) * 1.01)[0], "vx,vz,p")
    eng.misfit(vpvsrho_to_rholamemu(vvr), d_obs, components="vx,vz,p").backward()
    m = vpvsrho_to_rholamemu(vvr.detach()).clone().requires_grad_()
    eng.misfit(m, d_obs, components="vx,vz,p").backward()
    ref = su.conv_jacobian_adj(w["vvr"].astype(np.float64), m.grad.cpu().numpy())
    assert rel_l2(vvr.grad.cpu().numpy(), ref) < 1e-10


def test_full_size_gradient_fp32_within_1e4_of_fp64():
    """Two shots (the first and the last) of the 32-shot FWI workload at full size (600x1592
    padded, nts=3000, sub=2) through the autograd path with the forward wavefield checkpointed
    every ceil(sqrt(nts-1)) = 55 coarse intervals (54 checkpoints, each segment recomputed in the
    backward): the FP32 gradient is within 1e-4 relative L2 of the FP64 one (the FP64 GPU path
    matches the FP64 oracle to 1e-10 on the small cases above)."""
    import torch
    from elastic_isotropic_operators_amd import workloads as W
    from elastic_isotropic_operators_amd.fwi import ElasticFwi, datcomp
    wt, w0 = W.fwi_c3(nshots=32)
    e = [0, 31]
    seg = int(np.ceil(np.sqrt(w0["nts"] - 1)))
    assert seg == 55
    grid = dict(nz=w0["nz"], nx=w0["nx"], dz=w0["dz"], dx=w0["dx"], oz=w0["oz"], ox=w0["ox"], nts=w0["nts"],
                dts=w0["dts"], sub=w0["sub"], minPad=W.min_pad(w0), alphaCos=w0["alphaCos"])
    grads = {}
    d_obs = None
    for dt in (torch.float64, torch.float32):
        eng = ElasticFwi(grid, w0["src_z"][e], w0["src_x"][e], w0["model"], w0["rec_z"], w0["rec_x"], dtype=dt,
                         segment=seg)
        if d_obs is None:
            with torch.no_grad():
                d_obs = datcomp(eng.forward_data(torch.tensor(wt["elastic"], dtype=dt, device="cuda"))[0])
        m = torch.tensor(w0["elastic"], dtype=dt, device="cuda", requires_grad=True)
        eng.misfit(m, d_obs.to(dt), components="vx,vz,p").backward()
        assert eng.last_store_info[:2] == (seg, 54), eng.last_store_info
        grads[dt] = m.grad.double().cpu().numpy()
        del eng, m
        torch.cuda.empty_cache()
    assert np.linalg.norm(grads[torch.float64]) > 0
    err = rel_l2(grads[torch.float32], grads[torch.float64])
    print("C3 checkpointed gradient FP32 vs FP64 (shots 0, 31): rel L2 %.3e" % err, flush=True)
    assert err < 1e-4, err


def test_autograd_gradient_is_deterministic():
    """Three gradients of the same objective are bitwise equal (guards the ordering between
    torch's stream, which produces media and residuals, and the propagator's stream)."""
    import torch
    w, g, m0, dtrue = fwi_case(nts=400, dts=0.001, nExp=2, nz=48, nx=72)
    eng = _engine(w, g, torch.float32, segment=9)
    with torch.no_grad():
        d_obs = eng.forward_data(torch.tensor(m0 + dtrue, dtype=torch.float32, device="cuda"))[0]
    grads = []
    for _ in range(3):
        m = torch.tensor(m0, dtype=torch.float32, device="cuda", requires_grad=True)
        eng.misfit(m, d_obs).backward()
        grads.append(m.grad.cpu().numpy())
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


@pytest.mark.parametrize("tdtype", ["float64", "float32"])
def test_grouped_full_storage_gradient_equals_checkpointed(tdtype):
    """misfit_and_gradient (shot groups, full coarse storage, no recompute) gives the misfit and
    gradient of the checkpointed autograd path."""
    import torch
    dtype = getattr(torch, tdtype)
    w, g, m0, dtrue = fwi_case(nts=300, dts=0.001, nExp=3, nz=48, nx=72)
    sz, sx, rz, rx, src = w["src_z"], w["src_x"], w["rec_z"], w["rec_x"], w["model"]
    dobs = np.stack([O.nl_fwd_shot(g, m0 + dtrue, sz[e], sx[e], rz, rx, src) for e in range(3)])
    eng = _engine(w, g, dtype, segment=7)
    m = torch.tensor(m0, dtype=dtype, device="cuda", requires_grad=True)
    d_obs = torch.tensor(dobs, dtype=dtype, device="cuda")
    from elastic_isotropic_operators_amd.fwi import datcomp
    comps = "vx,vz,p"
    loss = eng.misfit(m, datcomp(d_obs, comps), comps)
    loss.backward()
    ref = m.grad.double()
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    for group in (1, 2, None):
        phi, grad = eng.misfit_and_gradient(m.detach(), datcomp(d_obs, comps), comps, group=group)
        assert abs(phi - float(loss.detach())) / float(loss.detach()) < tol
        assert float(torch.linalg.norm(grad.double() - ref) / torch.linalg.norm(ref)) < tol, group
    assert eng.last_store_info[1] == 0          # one segment: no checkpoints
    # the stores borrow the context's workspace; freeing it between calls changes nothing
    eng.ctx.release_workspace()
    phi2, grad2 = eng.misfit_and_gradient(m.detach(), datcomp(d_obs, comps), comps, group=2)
    assert float(torch.linalg.norm(grad2.double() - ref) / torch.linalg.norm(ref)) < tol
