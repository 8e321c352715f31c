"""GPU parity of the C ABI (libeio_hip.so) against the CPU oracle at the regular-grid level.

The C ABI takes scaled media, regular device positions and regular traces, exactly the
level of the reference's GPU seam (propShotsElastic{Fwd,Adj}Gpu, BornShots{Fwd,Adj}Gpu).
Bars: FP32 within 1e-5 relative L2 of the FP64 oracle; FP64 within 1e-11; dot-product
tests 1e-5 (FP32) and 1e-12 (FP64).
"""
import numpy as np
import pytest

import oracle as O
from elastic_isotropic_operators_amd import _capi
from elastic_isotropic_operators_amd import workloads as W

from helpers import GRID_OF_COMP, oracle_grid, reg_positions, rel_l2, random_traces, scattered_positions

pytestmark = pytest.mark.gpu

TOL = {np.float32: 1e-5, np.float64: 1e-11}
DOT_TOL = {np.float32: 1e-5, np.float64: 1e-12}
# Born adjoint (image) bars: the imaging sum accumulates over every coarse sample; FP32 reaches
# 1e-6 .. 1e-5 on these cases
BADJ_TOL = {np.float32: 2e-5, np.float64: 1e-10}


def _case(nz=40, nx=56, nts=120, sub=1, seed=0):
    w = W.small(nz=nz, nx=nx, nts=nts, sub=sub, seed=seed)
    g = oracle_grid(w)
    med = O.media(g, w["elastic"], np.float64)
    spos = reg_positions(w, w["src_z"][0], w["src_x"][0])
    rpos = reg_positions(w, w["rec_z"], w["rec_x"])
    return w, g, med, spos, rpos


VARIANT = [0]


def _ctx(dtype):
    c = _capi.Context(0, dtype)
    c.set_variant(VARIANT[0])
    return c


def _model(ctx, w, g, med):
    return ctx.model_add(g.nz, g.nx, g.nts, g.sub, g.dz, g.dx, g.dts, g.minPad, med, g.alphaCos)


def _gpu_nl_fwd(ctx, mid, spos, rpos, src, nts, dtype):
    rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, nts), dtype) for k in range(5)]
    shot, keep = ctx.make_shot(mid, spos, rpos, [np.ascontiguousarray(s, dtype) for s in src], rec)
    ctx.nl_forward([shot])
    return rec


def _gpu_nl_adj(ctx, mid, spos, rpos, dat, nts, dtype):
    src = [np.zeros((spos[GRID_OF_COMP[k]].size, nts), dtype) for k in range(5)]
    shot, keep = ctx.make_shot(mid, spos, rpos, src, [np.ascontiguousarray(d, dtype) for d in dat])
    ctx.nl_adjoint([shot])
    return src


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("sub", [1, 2])
def test_nl_forward_parity(dtype, sub):
    w, g, med, spos, rpos = _case(sub=sub)
    rng = np.random.default_rng(1)
    src = random_traces(rng, spos, g.nts, np.float64)
    ref = O.nl_fwd_reg(g, med, spos, src, rpos, np.float64)
    ctx = _ctx(dtype)
    mid = _model(ctx, w, g, med)
    out = _gpu_nl_fwd(ctx, mid, spos, rpos, src, g.nts, dtype)
    for k in range(5):
        assert np.linalg.norm(ref[k]) > 0
        assert rel_l2(out[k], ref[k]) < TOL[dtype], (k, rel_l2(out[k], ref[k]))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nl_adjoint_parity(dtype):
    w, g, med, spos, rpos = _case(sub=1)
    rng = np.random.default_rng(2)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    ref = O.nl_adj_reg(g, med, spos, rpos, dat, np.float64)
    ctx = _ctx(dtype)
    mid = _model(ctx, w, g, med)
    out = _gpu_nl_adj(ctx, mid, spos, rpos, dat, g.nts, dtype)
    for k in range(5):
        assert rel_l2(out[k], ref[k]) < TOL[dtype], (k, rel_l2(out[k], ref[k]))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nl_dot_product(dtype):
    w, g, med, spos, rpos = _case(sub=1, seed=3)
    rng = np.random.default_rng(11)
    m = [rng.standard_normal((spos[GRID_OF_COMP[k]].size, g.nts)).astype(dtype) for k in range(5)]
    d = [rng.standard_normal((rpos[GRID_OF_COMP[k]].size, g.nts)).astype(dtype) for k in range(5)]
    ctx = _ctx(dtype)
    mid = _model(ctx, w, g, med)
    Fm = _gpu_nl_fwd(ctx, mid, spos, rpos, m, g.nts, dtype)
    Ftd = _gpu_nl_adj(ctx, mid, spos, rpos, d, g.nts, dtype)
    a = sum(np.sum(Fm[k].astype(np.float64) * d[k]) for k in range(5))
    b = sum(np.sum(m[k].astype(np.float64) * Ftd[k]) for k in range(5))
    assert abs(a - b) / abs(b) < DOT_TOL[dtype], (a, b)


def _born_inputs(w, g, spos, seed):
    rng = np.random.default_rng(seed)
    ntw = (g.nts - 1) * g.sub + 1
    src_fine = random_traces(rng, spos, ntw, np.float64)
    from scipy.ndimage import gaussian_filter
    dm5 = np.stack([gaussian_filter(rng.standard_normal((g.nx, g.nz)), 2.0) for _ in range(5)])
    dm5[:, :4, :] = 0
    dm5[:, -4:, :] = 0
    dm5[:, :, :4] = 0
    dm5[:, :, -4:] = 0
    return src_fine, dm5


def _gpu_born_fwd(ctx, mid, spos, rpos, src_fine, dm5, nts, dtype):
    rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, nts), dtype) for k in range(5)]
    shot, keep = ctx.make_shot(mid, spos, rpos, [np.ascontiguousarray(s, dtype) for s in src_fine], rec)
    ctx.born_forward([shot], dm5)
    return rec


def _gpu_born_adj(ctx, mid, spos, rpos, src_fine, dat, nx, nz, dtype):
    out = np.zeros((5, nx, nz), dtype)
    shot, keep = ctx.make_shot(mid, spos, rpos, [np.ascontiguousarray(s, dtype) for s in src_fine],
                               [np.ascontiguousarray(d, dtype) for d in dat])
    ctx.born_adjoint([shot], out)
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_born_forward_parity(dtype):
    w, g, med, spos, rpos = _case(sub=1, seed=4)
    src_fine, dm5 = _born_inputs(w, g, spos, 5)
    ref = O.born_fwd_reg(g, med, spos, src_fine, dm5, rpos, np.float64)
    ctx = _ctx(dtype)
    mid = _model(ctx, w, g, med)
    out = _gpu_born_fwd(ctx, mid, spos, rpos, src_fine, dm5, g.nts, dtype)
    for k in range(5):
        assert np.linalg.norm(ref[k]) > 0
        assert rel_l2(out[k], ref[k]) < TOL[dtype], (k, rel_l2(out[k], ref[k]))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_born_adjoint_parity(dtype):
    w, g, med, spos, rpos = _case(sub=2, seed=6)
    src_fine, _ = _born_inputs(w, g, spos, 7)
    rng = np.random.default_rng(8)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    ref = O.born_adj_reg(g, med, spos, src_fine, rpos, dat, np.float64)
    ctx = _ctx(dtype)
    mid = _model(ctx, w, g, med)
    out = _gpu_born_adj(ctx, mid, spos, rpos, src_fine, dat, g.nx, g.nz, dtype)
    assert rel_l2(out, ref) < BADJ_TOL[dtype], rel_l2(out, ref)
    for k in range(5):
        assert rel_l2(out[k], ref[k]) < BADJ_TOL[dtype], (k, rel_l2(out[k], ref[k]))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_born_dot_product(dtype):
    w, g, med, spos, rpos = _case(sub=1, seed=9)
    src_fine, dm5 = _born_inputs(w, g, spos, 10)
    rng = np.random.default_rng(12)
    d = [rng.standard_normal((rpos[GRID_OF_COMP[k]].size, g.nts)) for k in range(5)]
    ctx = _ctx(dtype)
    mid = _model(ctx, w, g, med)
    Bm = _gpu_born_fwd(ctx, mid, spos, rpos, src_fine, dm5, g.nts, dtype)
    Btd = _gpu_born_adj(ctx, mid, spos, rpos, src_fine, [x.astype(dtype) for x in d], g.nx, g.nz, dtype)
    dm5c = dm5.astype(dtype).astype(np.float64)
    a = sum(np.sum(Bm[k].astype(np.float64) * d[k].astype(dtype)) for k in range(5))
    b = np.sum(dm5c * Btd.astype(np.float64))
    assert abs(a - b) / abs(b) < DOT_TOL[dtype], (a, b)


def test_batched_uneven_shots_match_single_shots():
    """Several models of different size and trace length in one batched call equal
    per-shot oracle runs (the persistent batch only changes scheduling, not arithmetic)."""
    dtype = np.float32
    ctx = _ctx(dtype)
    cases = [_case(nz=40, nx=56, nts=120, sub=1, seed=20), _case(nz=56, nx=40, nts=90, sub=2, seed=21),
             _case(nz=33, nx=70, nts=150, sub=1, seed=22)]
    shots, recs, refs, keep = [], [], [], []
    rng = np.random.default_rng(23)
    for (w, g, med, spos, rpos) in cases:
        mid = _model(ctx, w, g, med)
        src = random_traces(rng, spos, g.nts, np.float64)
        refs.append(O.nl_fwd_reg(g, med, spos, src, rpos, np.float64))
        rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, g.nts), dtype) for k in range(5)]
        srcc = [np.ascontiguousarray(s, dtype) for s in src]
        shot, kp = ctx.make_shot(mid, spos, rpos, srcc, rec)
        keep.append((kp, srcc))
        shots.append(shot)
        recs.append(rec)
    ctx.nl_forward(shots)
    for rec, ref in zip(recs, refs):
        for k in range(5):
            assert rel_l2(rec[k], ref[k]) < 1e-5, (k, rel_l2(rec[k], ref[k]))


def test_invalid_positions_rejected():
    """A device in the FAT halo is an argument error, not a fault."""
    w, g, med, spos, rpos = _case()
    ctx = _ctx(np.float32)
    mid = _model(ctx, w, g, med)
    bad = [np.array([1], np.int32)] * 4  # iz = 1 is inside the FAT halo
    rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, g.nts), np.float32) for k in range(5)]
    src = [np.zeros((1, g.nts), np.float32) for _ in range(5)]
    shot, keep = ctx.make_shot(mid, bad, rpos, src, rec)
    with pytest.raises(_capi.EioError):
        ctx.nl_forward([shot])


def test_empty_batch_rejected():
    """A call with no shots is an argument error."""
    ctx = _ctx(np.float32)
    with pytest.raises(_capi.EioError):
        ctx.nl_forward([])
    with pytest.raises(_capi.EioError):
        ctx.born_forward([], np.zeros((5, 4, 4), np.float32))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_shots_without_sources_or_receivers(dtype):
    """A batch of a normal shot, a shot with no receivers and a shot with no sources (empty
    position lists and traces): the normal shot equals the oracle, the source-less shot records
    exactly zero, for the NL forward (FP32: two steps per launch, automatic tile width), the NL
    adjoint and the Born forward."""
    w = W.small(nz=90, nx=110, nts=40, sub=1, seed=180)
    g = oracle_grid(w)
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(181)
    spos = scattered_positions(rng, g.nz, g.nx, 6)
    rpos = scattered_positions(rng, g.nz, g.nx, 20)
    none = [np.zeros(0, np.int32)] * 4
    ctx = _capi.Context(0, dtype)
    if dtype == np.float32:
        ctx.set_time_block(2, _capi.VARIANT_AUTO)
        ctx.set_adjoint_time_block(2, _capi.VARIANT_AUTO)
    mid = _model(ctx, w, g, med)

    def traces(pos, nt):
        return [np.zeros((pos[GRID_OF_COMP[k]].size, nt), dtype) for k in range(5)]

    # NL forward
    src = random_traces(rng, spos, g.nts, dtype)
    ref = O.nl_fwd_reg(g, med, spos, [s.astype(np.float64) for s in src], rpos, np.float64)
    rec_a, rec_c = traces(rpos, g.nts), traces(rpos, g.nts)
    for k in range(5):
        rec_c[k][...] = 1.0  # must be overwritten with zeros
    shots, keep = [], []
    for sp, rp, sr, rc in ((spos, rpos, src, rec_a), (spos, none, src, traces(none, g.nts)),
                           (none, rpos, traces(none, g.nts), rec_c)):
        sh, kp = ctx.make_shot(mid, sp, rp, sr, rc)
        shots.append(sh)
        keep.append(kp)
    ctx.nl_forward(shots)
    for k in range(5):
        assert rel_l2(rec_a[k], ref[k]) < TOL[dtype], (k, rel_l2(rec_a[k], ref[k]))
        assert not np.any(rec_c[k])
    # NL adjoint: the receiver-less shot extracts exactly zero
    dat = random_traces(rng, rpos, g.nts, dtype)
    ref_a = O.nl_adj_reg(g, med, spos, rpos, [d.astype(np.float64) for d in dat], np.float64)
    out_a, out_b = traces(spos, g.nts), traces(spos, g.nts)
    for k in range(5):
        out_b[k][...] = 1.0
    shots, keep = [], []
    for sp, rp, sr, rc in ((spos, rpos, out_a, dat), (spos, none, out_b, traces(none, g.nts)),
                           (none, rpos, traces(none, g.nts), dat)):
        sh, kp = ctx.make_shot(mid, sp, rp, sr, rc)
        shots.append(sh)
        keep.append(kp)
    ctx.nl_adjoint(shots)
    for k in range(5):
        assert rel_l2(out_a[k], ref_a[k]) < TOL[dtype], (k, rel_l2(out_a[k], ref_a[k]))
        assert not np.any(out_b[k])
    # Born forward
    src_fine, dm5 = _born_inputs(w, g, spos, 182)
    ref_b = O.born_fwd_reg(g, med, spos, src_fine, dm5, rpos, np.float64)
    ntw = (g.nts - 1) * g.sub + 1
    rec_a, rec_c = traces(rpos, g.nts), traces(rpos, g.nts)
    for k in range(5):
        rec_c[k][...] = 1.0
    shots, keep = [], []
    sf = [np.ascontiguousarray(s, dtype) for s in src_fine]
    for sp, rp, sr, rc in ((spos, rpos, sf, rec_a), (spos, none, sf, traces(none, g.nts)),
                           (none, rpos, traces(none, ntw), rec_c)):
        sh, kp = ctx.make_shot(mid, sp, rp, sr, rc)
        shots.append(sh)
        keep.append(kp)
    ctx.born_forward(shots, dm5)
    for k in range(5):
        assert rel_l2(rec_a[k], ref_b[k]) < TOL[dtype], (k, rel_l2(rec_a[k], ref_b[k]))
        assert not np.any(rec_c[k])


@pytest.mark.parametrize("v", range(6))
def test_all_fp32_variants_forward_adjoint(v):
    """Every FP32 step-kernel variant (tile shape / occupancy) gives the same NL forward and
    adjoint to FP32 accuracy."""
    if v >= len(_capi.Context.variants()):
        pytest.skip("variant not compiled")
    w, g, med, spos, rpos = _case(nz=70, nx=90, nts=80, sub=1, seed=30)
    rng = np.random.default_rng(31)
    src = random_traces(rng, spos, g.nts, np.float64)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    ref_f = O.nl_fwd_reg(g, med, spos, src, rpos, np.float64)
    ref_a = O.nl_adj_reg(g, med, spos, rpos, dat, np.float64)
    ctx = _capi.Context(0, np.float32)
    ctx.set_variant(v)
    mid = _model(ctx, w, g, med)
    out_f = _gpu_nl_fwd(ctx, mid, spos, rpos, src, g.nts, np.float32)
    out_a = _gpu_nl_adj(ctx, mid, spos, rpos, dat, g.nts, np.float32)
    for k in range(5):
        assert rel_l2(out_f[k], ref_f[k]) < 1e-5, (v, k, rel_l2(out_f[k], ref_f[k]))
        assert rel_l2(out_a[k], ref_a[k]) < 1e-5, (v, k, rel_l2(out_a[k], ref_a[k]))


def test_born_adjoint_with_reused_storage_of_another_shape():
    """The coarse Vx/Vz storage is a context workspace that is never cleared: a Born call on a
    larger grid leaves nonzero values where the next, smaller grid has its FAT halo.  The
    imaging condition must not read them (regression test)."""
    dtype = np.float32
    ctx = _ctx(dtype)
    w1, g1, med1, sp1, rp1 = _case(nz=64, nx=80, nts=100, sub=1, seed=50)
    src1, _ = _born_inputs(w1, g1, sp1, 51)
    d1 = random_traces(np.random.default_rng(52), rp1, g1.nts, np.float64)
    m1 = _model(ctx, w1, g1, med1)
    _gpu_born_adj(ctx, m1, sp1, rp1, src1, d1, g1.nx, g1.nz, dtype)
    w, g, med, spos, rpos = _case(nz=40, nx=56, nts=100, sub=1, seed=53)
    src_fine, _ = _born_inputs(w, g, spos, 54)
    dat = random_traces(np.random.default_rng(55), rpos, g.nts, np.float64)
    ref = O.born_adj_reg(g, med, spos, src_fine, rpos, dat, np.float64)
    mid = _model(ctx, w, g, med)
    out = _gpu_born_adj(ctx, mid, spos, rpos, src_fine, dat, g.nx, g.nz, dtype)
    assert rel_l2(out, ref) < 2e-5, rel_l2(out, ref)


@pytest.mark.parametrize("v", range(13))
@pytest.mark.parametrize("sub_fd", [2, 3])
def test_two_step_nl_forward_matches_oracle(v, sub_fd):
    """The temporally blocked NL forward (two fine steps per launch, step 1 recomputed on the
    tile halo) equals the oracle for every variant, with sources and receivers scattered over
    the grid (many on tile edges, so halo injection is exercised) and with an odd number of
    fine steps (sub_fd = 3, nts = 60: 177 steps, the last launch does one step)."""
    if v >= len(_capi.Context.time_block_variants()):
        pytest.skip("variant not compiled")
    w = W.small(nz=150, nx=130, nts=60, sub=1, seed=60)
    g = O.grid(w["nz"], w["nx"], w["nts"], sub_fd, w["dz"], w["dx"], w["dts"], W.min_pad(w),
               w["alphaCos"], w["oz"], w["ox"])
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(61 + v)
    spos = scattered_positions(rng, g.nz, g.nx, 40)
    rpos = scattered_positions(rng, g.nz, g.nx, 60)
    src = random_traces(rng, spos, g.nts, np.float64)
    ref = O.nl_fwd_reg(g, med, spos, src, rpos, np.float64)
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, v)
    mid = _model(ctx, w, g, med)
    out = _gpu_nl_fwd(ctx, mid, spos, rpos, src, g.nts, np.float32)
    for k in range(5):
        assert np.linalg.norm(ref[k]) > 0
        assert rel_l2(out[k], ref[k]) < 1e-5, (v, k, rel_l2(out[k], ref[k]))


def test_auto_time_block_variant_matches_explicit():
    """EIO_VARIANT_AUTO chooses the two-step tile width per call: a one-shot 150 x 130 batch
    (6 tiles of 56 x 32, under the threshold of 1024) equals the oracle and the explicit 56 x 8
    variants (NL forward: variant 12, streamed stage; NL adjoint: variant 10); a batch of 4 shots of 600 x 1600 (2200 tiles) equals
    the explicit 56 x 32 variants (Born forward, Born adjoint) and, for its NL forward, which
    runs in two shot groups and so takes 56 x 8 with the streamed stage, agrees with them to
    FP32 rounding."""
    w = W.small(nz=150, nx=130, nts=30, sub=1, seed=170)
    g = oracle_grid(w)
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(171)
    spos = scattered_positions(rng, g.nz, g.nx, 20)
    rpos = scattered_positions(rng, g.nz, g.nx, 40)
    src = random_traces(rng, spos, g.nts, np.float64)
    ref = O.nl_fwd_reg(g, med, spos, src, rpos, np.float64)
    ctx = _capi.Context(0, np.float32)
    mid = _model(ctx, w, g, med)
    out = {}
    for v in (12, _capi.VARIANT_AUTO):
        ctx.set_time_block(2, v)
        out[v] = _gpu_nl_fwd(ctx, mid, spos, rpos, src, g.nts, np.float32)
    for k in range(5):
        assert rel_l2(out[_capi.VARIANT_AUTO][k], ref[k]) < 1e-5, (k, rel_l2(out[_capi.VARIANT_AUTO][k], ref[k]))
        assert np.array_equal(out[_capi.VARIANT_AUTO][k], out[12][k]), k
    # the adjoint's automatic choice on the same small batch: 56 x 8 (variant 10)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    ref_a = O.nl_adj_reg(g, med, spos, rpos, dat, np.float64)
    adj = {}
    for va in (10, _capi.VARIANT_AUTO):
        ctx.set_adjoint_time_block(2, va)
        adj[va] = _gpu_nl_adj(ctx, mid, spos, rpos, dat, g.nts, np.float32)
    for k in range(5):
        assert rel_l2(adj[_capi.VARIANT_AUTO][k], ref_a[k]) < 1e-5, (k, rel_l2(adj[_capi.VARIANT_AUTO][k], ref_a[k]))
        assert rel_l2(adj[_capi.VARIANT_AUTO][k], adj[10][k]) < 1e-6
    # above the threshold: 4 shots of 600 x 1600, a few coarse samples
    wb = W.small(nz=600, nx=1600, nts=6, sub=1, seed=172)
    gb = oracle_grid(wb)
    medb = O.media(gb, wb["elastic"], np.float64)
    midb = _model(ctx, wb, gb, medb)
    ntw = (gb.nts - 1) * gb.sub + 1
    shots, keep, recs = {}, [], {}
    dm5 = rng.standard_normal((5, gb.nx, gb.nz)) * 1e-3
    for v in (4, _capi.VARIANT_AUTO):
        ctx.set_time_block(2, v)
        for kind in ("nl", "born"):
            rng_v = np.random.default_rng(173)
            batch, rec_all = [], []
            for e in range(4):
                sp = scattered_positions(rng_v, gb.nz, gb.nx, 10)
                rp = scattered_positions(rng_v, gb.nz, gb.nx, 30)
                tr = random_traces(rng_v, sp, ntw if kind == "born" else gb.nts, np.float32)
                rec = [np.zeros((rp[GRID_OF_COMP[k]].size, gb.nts), np.float32) for k in range(5)]
                shot, kp = ctx.make_shot(midb, sp, rp, tr, rec)
                batch.append(shot)
                keep.append(kp)
                rec_all.append(rec)
            if kind == "nl":
                ctx.nl_forward(batch)
            else:
                ctx.born_forward(batch, dm5)
            recs[(v, kind)] = rec_all
    for kind in ("nl", "born"):
        for e in range(4):
            for k in range(5):
                a, b = recs[(_capi.VARIANT_AUTO, kind)][e][k], recs[(4, kind)][e][k]
                assert np.linalg.norm(b) > 0
                assert rel_l2(a, b) < 1e-6, (kind, e, k, rel_l2(a, b))
    # the Born adjoint of the large batch: 56 x 32 (variant 7)
    img = {}
    for va in (7, _capi.VARIANT_AUTO):
        ctx.set_adjoint_time_block(2, va)
        rng_v = np.random.default_rng(174)
        batch = []
        for e in range(4):
            sp = scattered_positions(rng_v, gb.nz, gb.nx, 10)
            rp = scattered_positions(rng_v, gb.nz, gb.nx, 30)
            tr = random_traces(rng_v, sp, ntw, np.float32)
            dd = random_traces(rng_v, rp, gb.nts, np.float32)
            shot, kp = ctx.make_shot(midb, sp, rp, tr, dd)
            batch.append(shot)
            keep.append(kp)
        img[va] = np.zeros((5, gb.nx, gb.nz), np.float32)
        ctx.born_adjoint(batch, img[va])
    assert np.linalg.norm(img[7]) > 0
    assert rel_l2(img[_capi.VARIANT_AUTO], img[7]) < 1e-6, rel_l2(img[_capi.VARIANT_AUTO], img[7])


@pytest.mark.parametrize("va", range(12))
@pytest.mark.parametrize("sub_fd", [2, 4, 3])
def test_two_step_born_adjoint_matches_oracle(va, sub_fd):
    """The Born adjoint with two reverse fine steps per launch of the receiver pass (four state
    levels, step 1 recomputed on the tile halo, data injected on the halo in step 1 and on
    owned cells in step 2, the pair's adjoint fields summed in registers) against the FP64
    oracle and against one step per launch, with 300 receivers per grid scattered over a
    grid of several tiles each way.  sub_fd 2: one pair per coarse interval, 4: two (slot its
    read-modified-written by the second), 3: odd, one step per launch."""
    if va >= len(_capi.Context.adjoint_time_block_variants()):
        pytest.skip("variant not compiled")
    w = W.small(nz=150, nx=130, nts=50, sub=1, seed=80)
    g = O.grid(w["nz"], w["nx"], w["nts"], sub_fd, w["dz"], w["dx"], w["dts"], W.min_pad(w),
               w["alphaCos"], w["oz"], w["ox"])
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(81 + va)
    spos = scattered_positions(rng, g.nz, g.nx, 20)
    rpos = scattered_positions(rng, g.nz, g.nx, 300)
    src_fine, _ = _born_inputs(w, g, spos, 82)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    ref = O.born_adj_reg(g, med, spos, src_fine, rpos, dat, np.float64)
    ctx = _capi.Context(0, np.float32)
    mid = _model(ctx, w, g, med)
    one = _gpu_born_adj(ctx, mid, spos, rpos, src_fine, dat, g.nx, g.nz, np.float32)
    ctx.set_adjoint_time_block(2, va)
    two = _gpu_born_adj(ctx, mid, spos, rpos, src_fine, dat, g.nx, g.nz, np.float32)
    for k in range(5):
        assert np.linalg.norm(ref[k]) > 0
        assert rel_l2(two[k], ref[k]) < 2e-5, (va, k, rel_l2(two[k], ref[k]))
        assert rel_l2(two[k], one[k]) < 1e-5, (va, k, rel_l2(two[k], one[k]))
    # a second call reuses the four levels and the slots: same result
    again = _gpu_born_adj(ctx, mid, spos, rpos, src_fine, dat, g.nx, g.nz, np.float32)
    assert np.array_equal(again, two)


@pytest.mark.parametrize("va", range(12))
def test_two_step_nl_adjoint_matches_oracle(va):
    """The NL adjoint with two reverse fine steps per launch (data injected on the step-1 halo,
    source traces extracted from u1 and u2 at owned source cells) against the FP64 oracle and
    against one step per launch: 120 sources and 200 receivers per grid scattered over a grid
    of several tiles each way, nts 50, sub_fd 4."""
    if va >= len(_capi.Context.adjoint_time_block_variants()):
        pytest.skip("variant not compiled")
    w = W.small(nz=150, nx=130, nts=50, sub=2, seed=95)
    g = oracle_grid(w)
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(96 + va)
    spos = scattered_positions(rng, g.nz, g.nx, 120)
    rpos = scattered_positions(rng, g.nz, g.nx, 200)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    ref = O.nl_adj_reg(g, med, spos, rpos, dat, np.float64)
    ctx = _capi.Context(0, np.float32)
    mid = _model(ctx, w, g, med)
    one = _gpu_nl_adj(ctx, mid, spos, rpos, dat, g.nts, np.float32)
    ctx.set_adjoint_time_block(2, va)
    two = _gpu_nl_adj(ctx, mid, spos, rpos, dat, g.nts, np.float32)
    for k in range(5):
        assert np.linalg.norm(ref[k]) > 0
        assert rel_l2(two[k], ref[k]) < 1e-5, (va, k, rel_l2(two[k], ref[k]))
        assert rel_l2(two[k], one[k]) < 1e-5, (va, k, rel_l2(two[k], one[k]))


def test_two_step_nl_adjoint_pairs_across_coarse_intervals():
    """sub_fd 3 and an even step count (nts 51): pairs of reverse steps straddle coarse
    intervals, which the NL adjoint allows (each step injects and extracts at its own time)."""
    w = W.small(nz=60, nx=70, nts=51, sub=1, seed=160)
    g = O.grid(w["nz"], w["nx"], w["nts"], 3, w["dz"], w["dx"], w["dts"], W.min_pad(w),
               w["alphaCos"], w["oz"], w["ox"])
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(161)
    spos = scattered_positions(rng, g.nz, g.nx, 20)
    rpos = scattered_positions(rng, g.nz, g.nx, 30)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    ref = O.nl_adj_reg(g, med, spos, rpos, dat, np.float64)
    ctx = _capi.Context(0, np.float32)
    ctx.set_adjoint_time_block(2, 7)
    mid = _model(ctx, w, g, med)
    out = _gpu_nl_adj(ctx, mid, spos, rpos, dat, g.nts, np.float32)
    for k in range(5):
        assert rel_l2(out[k], ref[k]) < 1e-5, (k, rel_l2(out[k], ref[k]))


def test_two_step_nl_dot_product():
    """<F m, d> = <m, F^T d> with two fine steps per launch both ways (FP32 bar 1e-5)."""
    w, g, med, spos, rpos = _case(nz=120, nx=100, nts=80, sub=1, seed=97)
    rng = np.random.default_rng(98)
    m = [rng.standard_normal((spos[GRID_OF_COMP[k]].size, g.nts)).astype(np.float32) for k in range(5)]
    d = [rng.standard_normal((rpos[GRID_OF_COMP[k]].size, g.nts)).astype(np.float32) for k in range(5)]
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, 4)
    ctx.set_adjoint_time_block(2, 7)
    mid = _model(ctx, w, g, med)
    Fm = _gpu_nl_fwd(ctx, mid, spos, rpos, m, g.nts, np.float32)
    Ftd = _gpu_nl_adj(ctx, mid, spos, rpos, d, g.nts, np.float32)
    a = sum(np.sum(Fm[k].astype(np.float64) * d[k]) for k in range(5))
    b = sum(np.sum(m[k].astype(np.float64) * Ftd[k]) for k in range(5))
    assert abs(a - b) / abs(b) < DOT_TOL[np.float32], (a, b)


@pytest.mark.parametrize("v", [0, 1, 4, 7, 8, 9, 10, 11, 12])
def test_two_step_born_source_pass_matches_oracle(v):
    """Born forward and adjoint whose source and scattered passes take two fine steps per launch
    (fine-sampled sources injected on the step-1 halo and Vx/Vz decimated into the whole
    record; secondary sources added at every cell of both steps, receivers recorded): the FP64
    oracle and one step per launch, with 30 sources and 200 receivers per grid scattered over
    several tiles each way; sub_fd 4."""
    w = W.small(nz=150, nx=130, nts=50, sub=2, seed=100)
    g = oracle_grid(w)
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(101 + v)
    spos = scattered_positions(rng, g.nz, g.nx, 30)
    rpos = scattered_positions(rng, g.nz, g.nx, 200)
    src_fine, dm5 = _born_inputs(w, g, spos, 102)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    ref_f = O.born_fwd_reg(g, med, spos, src_fine, dm5, rpos, np.float64)
    ref_a = O.born_adj_reg(g, med, spos, src_fine, rpos, dat, np.float64)
    ctx = _capi.Context(0, np.float32)
    mid = _model(ctx, w, g, med)
    one_f = _gpu_born_fwd(ctx, mid, spos, rpos, src_fine, dm5, g.nts, np.float32)
    one_a = _gpu_born_adj(ctx, mid, spos, rpos, src_fine, dat, g.nx, g.nz, np.float32)
    ctx.set_time_block(2, v)
    two_f = _gpu_born_fwd(ctx, mid, spos, rpos, src_fine, dm5, g.nts, np.float32)
    two_a = _gpu_born_adj(ctx, mid, spos, rpos, src_fine, dat, g.nx, g.nz, np.float32)
    for k in range(5):
        assert np.linalg.norm(ref_f[k]) > 0
        assert rel_l2(two_f[k], ref_f[k]) < 1e-5, (v, k, rel_l2(two_f[k], ref_f[k]))
        assert rel_l2(two_f[k], one_f[k]) < 1e-5, (v, k, rel_l2(two_f[k], one_f[k]))
    assert rel_l2(two_a, ref_a) < 2e-5, rel_l2(two_a, ref_a)
    assert rel_l2(two_a, one_a) < 1e-5, rel_l2(two_a, one_a)


def test_born_with_odd_fine_subsampling_and_time_block_two():
    """A model with an odd fine sub-sampling (sub_fd 3) and an even step count: the Born passes
    whose pairs would straddle coarse intervals take one step per launch, so forward and
    adjoint still equal the oracle with the two-step setting on."""
    w = W.small(nz=60, nx=70, nts=51, sub=1, seed=150)
    g = O.grid(w["nz"], w["nx"], w["nts"], 3, w["dz"], w["dx"], w["dts"], W.min_pad(w),
               w["alphaCos"], w["oz"], w["ox"])
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(151)
    spos = scattered_positions(rng, g.nz, g.nx, 8)
    rpos = scattered_positions(rng, g.nz, g.nx, 30)
    src_fine, dm5 = _born_inputs(w, g, spos, 152)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    ref_f = O.born_fwd_reg(g, med, spos, src_fine, dm5, rpos, np.float64)
    ref_a = O.born_adj_reg(g, med, spos, src_fine, rpos, dat, np.float64)
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, 4)
    ctx.set_adjoint_time_block(2, 7)
    mid = _model(ctx, w, g, med)
    out_f = _gpu_born_fwd(ctx, mid, spos, rpos, src_fine, dm5, g.nts, np.float32)
    out_a = _gpu_born_adj(ctx, mid, spos, rpos, src_fine, dat, g.nx, g.nz, np.float32)
    for k in range(5):
        assert rel_l2(out_f[k], ref_f[k]) < 1e-5, (k, rel_l2(out_f[k], ref_f[k]))
    assert rel_l2(out_a, ref_a) < 2e-5, rel_l2(out_a, ref_a)


def test_two_step_born_adjoint_dot_product():
    """<B m, d> = <m, B^T d> with the two-step receiver pass (FP32 bar 1e-5)."""
    w, g, med, spos, rpos = _case(nz=120, nx=100, nts=80, sub=1, seed=85)
    src_fine, dm5 = _born_inputs(w, g, spos, 86)
    rng = np.random.default_rng(87)
    d = [rng.standard_normal((rpos[GRID_OF_COMP[k]].size, g.nts)) for k in range(5)]
    ctx = _capi.Context(0, np.float32)
    ctx.set_adjoint_time_block(2, 0)
    mid = _model(ctx, w, g, med)
    Bm = _gpu_born_fwd(ctx, mid, spos, rpos, src_fine, dm5, g.nts, np.float32)
    Btd = _gpu_born_adj(ctx, mid, spos, rpos, src_fine, [x.astype(np.float32) for x in d], g.nx, g.nz,
                        np.float32)
    dm5c = dm5.astype(np.float32).astype(np.float64)
    a = sum(np.sum(Bm[k].astype(np.float64) * d[k].astype(np.float32)) for k in range(5))
    b = np.sum(dm5c * Btd.astype(np.float64))
    assert abs(a - b) / abs(b) < DOT_TOL[np.float32], (a, b)


def test_two_step_uneven_batch_matches_single_step():
    """An uneven batch (different grids, trace lengths and fine sub-samplings, even and odd
    step counts) gives the same data with two steps per launch as with one."""
    rng = np.random.default_rng(70)
    cases = [_case(nz=70, nx=96, nts=90, sub=1, seed=71), _case(nz=96, nx=70, nts=61, sub=2, seed=72),
             _case(nz=45, nx=120, nts=120, sub=1, seed=73)]
    outs = {}
    for tb in (1, 2):
        ctx = _capi.Context(0, np.float32)
        ctx.set_time_block(tb, 0)
        shots, keep, recs = [], [], []
        for i, (w, g, med, spos, rpos) in enumerate(cases):
            if i == 1:  # odd fine sub-sampling through the C ABI: odd step count
                g = O.grid(g.nz, g.nx, g.nts, 3, g.dz, g.dx, g.dts, g.minPad, g.alphaCos)
            mid = _model(ctx, w, g, med)
            src = random_traces(np.random.default_rng(80 + i), spos, g.nts, np.float32)
            rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, g.nts), np.float32) for k in range(5)]
            s, kp = ctx.make_shot(mid, spos, rpos, src, rec)
            shots.append(s)
            keep.append((kp, src))
            recs.append(rec)
        ctx.nl_forward(shots)
        outs[tb] = recs
    for a, b in zip(outs[1], outs[2]):
        for k in range(5):
            assert rel_l2(b[k], a[k]) < 2e-6, (k, rel_l2(b[k], a[k]))


def test_two_step_nl_adjoint_batch_of_different_grids():
    """One NL adjoint call over shots on three grids of different shape (same time axis, as the
    adjoint requires): two reverse steps per launch give each shot's one-step result, and a
    shot's result does not depend on the batch it runs in."""
    cases = [_case(nz=70, nx=96, nts=60, sub=1, seed=121), _case(nz=96, nx=70, nts=60, sub=1, seed=122),
             _case(nz=45, nx=120, nts=60, sub=1, seed=123)]
    outs = {}
    for tb in (1, 2, "alone"):
        ctx = _capi.Context(0, np.float32)
        ctx.set_adjoint_time_block(1 if tb == 1 else 2, 7)
        srcs = []
        batches = [[0, 1, 2]] if tb != "alone" else [[0], [1], [2]]
        per = {}
        for batch in batches:
            shots, keep = [], []
            for i in batch:
                w, g, med, spos, rpos = cases[i]
                mid = _model(ctx, w, g, med)
                dat = random_traces(np.random.default_rng(130 + i), rpos, g.nts, np.float32)
                src = [np.zeros((spos[GRID_OF_COMP[k]].size, g.nts), np.float32) for k in range(5)]
                s, kp = ctx.make_shot(mid, spos, rpos, src, dat)
                shots.append(s)
                keep.append((kp, dat))
                per[i] = src
            ctx.nl_adjoint(shots)
        outs[tb] = [per[i] for i in range(3)]
    for i in range(3):
        for k in range(5):
            assert np.linalg.norm(outs[1][i][k]) > 0
            assert rel_l2(outs[2][i][k], outs[1][i][k]) < 2e-6, (i, k)
            assert np.array_equal(outs[2][i][k], outs["alone"][i][k]), (i, k)


def test_autotuner_choice_keeps_parity():
    """The autotuner times every candidate kernel on the caller's batch with a step limit,
    leaves the fastest set on the context and lifts the limit: the next full call is exact."""
    from elastic_isotropic_operators_amd import autotune
    w, g, med, spos, rpos = _case(nz=70, nx=90, nts=80, sub=1, seed=90)
    rng = np.random.default_rng(91)
    src = random_traces(rng, spos, g.nts, np.float32)
    ref = O.nl_fwd_reg(g, med, spos, src, rpos, np.float64)
    ctx = _capi.Context(0, np.float32)
    mid = _model(ctx, w, g, med)
    rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, g.nts), np.float32) for k in range(5)]
    shot, keep = ctx.make_shot(mid, spos, rpos, src, rec)
    grids = [(g.nz, g.nx, (g.nts - 1) * g.sub, sum(p.size for p in spos), sum(p.size for p in rpos))]
    res = autotune.tune_nl_forward(ctx, [shot], grids, steps=32, device_ptrs=False, use_cache=False)
    assert len(res["table"]) == len(autotune.candidates()) and not res["cached"]
    assert all(v > 0 for v in res["table"].values())
    ctx.nl_forward([shot])
    for k in range(5):
        assert rel_l2(rec[k], ref[k]) < 1e-5, (res["choice"], k, rel_l2(rec[k], ref[k]))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_saved_wavefields_match_oracle(dtype):
    """eio_nl_{forward,adjoint}_wavefield: the saved coarse wavefield of one shot of a batch
    equals the oracle's, and the data of the call equal a plain call's."""
    w, g, med, spos, rpos = _case(nz=40, nx=56, nts=70, sub=2, seed=95)
    rng = np.random.default_rng(96)
    src = random_traces(rng, spos, g.nts, np.float64)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    ref_d, ref_w = O.nl_fwd_reg(g, med, spos, src, rpos, np.float64, wavefield=True)
    ref_m, ref_wa = O.nl_adj_reg(g, med, spos, rpos, dat, np.float64, wavefield=True)
    ctx = _ctx(dtype)
    mid = _model(ctx, w, g, med)
    tol = TOL[dtype]
    # forward: the saved shot is the second of a 2-shot batch (the first one is a decoy)
    recs = [[np.zeros((rpos[GRID_OF_COMP[k]].size, g.nts), dtype) for k in range(5)] for _ in range(2)]
    srcs = [[np.ascontiguousarray(s * (1 + e), dtype) for s in src] for e in range(2)]
    shots, keep = [], []
    for e in range(2):
        s, kp = ctx.make_shot(mid, spos, rpos, srcs[1 - e], recs[e])
        shots.append(s)
        keep.append(kp)
    wf = np.zeros((g.nts, 5, g.nx, g.nz), dtype)
    ctx.nl_forward_wavefield(shots, 1, wf)
    assert rel_l2(wf, ref_w) < tol, rel_l2(wf, ref_w)
    for k in range(5):
        assert rel_l2(recs[1][k], ref_d[k]) < tol
    # adjoint
    mods = [np.zeros((spos[GRID_OF_COMP[k]].size, g.nts), dtype) for k in range(5)]
    s, kp = ctx.make_shot(mid, spos, rpos, mods, [np.ascontiguousarray(d, dtype) for d in dat])
    wa = np.zeros((g.nts, 5, g.nx, g.nz), dtype)
    ctx.nl_adjoint_wavefield([s], 0, wa)
    assert rel_l2(wa, ref_wa) < tol, rel_l2(wa, ref_wa)
    for k in range(5):
        assert rel_l2(mods[k], ref_m[k]) < tol
    with pytest.raises(_capi.EioError, match="out of range"):
        ctx.nl_adjoint_wavefield([s], 1, wa)


def _fs_case(nts=80, seed=0):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    nz, nx = 70, 90
    vp = gaussian_filter(2200 + 400 * rng.random((nx, nz)), 3)
    rho = 1800 + 300 * gaussian_filter(rng.random((nx, nz)), 3)
    w = W._grid_dict(np.stack([vp, vp / 1.8, rho]).astype(np.float32), 10.0, 10.0, 10, 10, nts, 0.002, 1, 20.0,
                     surfaceCondition=1)
    g = O.grid(w["nz"], w["nx"], w["nts"], 2, w["dz"], w["dx"], w["dts"], W.min_pad(w), w["alphaCos"], w["oz"],
               w["ox"], surface=1)
    return w, g, O.media(g, w["elastic"], np.float64)


@pytest.mark.parametrize("dtype,variant", [(np.float64, 0)] + [(np.float32, v) for v in range(6)])
def test_free_surface_nl_forward_adjoint_match_oracle(dtype, variant):
    """Free surface (surface=1): forward and adjoint equal the oracle for every FP32 variant,
    with sources and receivers scattered over the grid including the rows above the surface;
    the GPU adjoint passes the dot test."""
    w, g, med = _fs_case()
    rng = np.random.default_rng(100 + variant)
    sp, rp = scattered_positions(rng, g.nz, g.nx, 20), scattered_positions(rng, g.nz, g.nx, 30)
    src = random_traces(rng, sp, g.nts, np.float64)
    dat = random_traces(rng, rp, g.nts, np.float64)
    ref_d = O.nl_fwd_reg(g, med, sp, src, rp, np.float64)
    ref_m = O.nl_adj_reg(g, med, sp, rp, dat, np.float64)
    ctx = _capi.Context(0, dtype)
    if dtype == np.float32:
        ctx.set_variant(variant)
        ctx.set_time_block(2, 0)  # free-surface shots fall back to the one-step kernel
    mid = ctx.model_add(g.nz, g.nx, g.nts, g.sub, g.dz, g.dx, g.dts, g.minPad, med, g.alphaCos, surface=1)
    out_d = _gpu_nl_fwd(ctx, mid, sp, rp, src, g.nts, dtype)
    out_m = _gpu_nl_adj(ctx, mid, sp, rp, dat, g.nts, dtype)
    for k in range(5):
        assert rel_l2(out_d[k], ref_d[k]) < TOL[dtype], (k, rel_l2(out_d[k], ref_d[k]))
        assert rel_l2(out_m[k], ref_m[k]) < TOL[dtype], (k, rel_l2(out_m[k], ref_m[k]))
    a = sum(np.sum(out_d[k].astype(np.float64) * dat[k].astype(dtype)) for k in range(5))
    mm = [np.ascontiguousarray(s, dtype).astype(np.float64) for s in src]
    b = sum(np.sum(mm[k] * out_m[k].astype(np.float64)) for k in range(5))
    assert abs(a - b) / abs(b) < DOT_TOL[dtype], (a, b)


def test_free_surface_rejected_by_born():
    w, g, med = _fs_case(nts=20)
    ctx = _capi.Context(0, np.float32)
    mid = ctx.model_add(g.nz, g.nx, g.nts, g.sub, g.dz, g.dx, g.dts, g.minPad, med, g.alphaCos, surface=1)
    spos = [np.array([20 * g.nz + 20], np.int32)] * 4
    rpos = [np.array([30 * g.nz + 20], np.int32)] * 4
    ntw = (g.nts - 1) * g.sub + 1
    src = [np.zeros((1, ntw), np.float32) for _ in range(5)]
    rec = [np.zeros((1, g.nts), np.float32) for _ in range(5)]
    shot, keep = ctx.make_shot(mid, spos, rpos, src, rec)
    with pytest.raises(_capi.EioError, match="nonlinear operator only"):
        ctx.born_forward([shot], np.zeros((5, g.nx, g.nz), np.float32))


def test_graph_mode_matches_eager_on_repeated_calls():
    """eio_ctx_set_graph(1): each call captures its time loop (state cleared by fill kernels,
    never hipMemsetAsync) into a hipGraph and replays it.  Two graph calls of each operator,
    with eager calls of another operator and host synchronisations between them, give the
    eager results bit for bit (same kernels in the same order); FP32, two steps per launch."""
    w = W.small(nz=100, nx=90, nts=40, sub=1, seed=110)
    g = oracle_grid(w)
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(111)
    spos = scattered_positions(rng, g.nz, g.nx, 10)
    rpos = scattered_positions(rng, g.nz, g.nx, 40)
    src = random_traces(rng, spos, g.nts, np.float64)
    dat = random_traces(rng, rpos, g.nts, np.float64)
    src_fine, dm5 = _born_inputs(w, g, spos, 112)
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, 4)
    ctx.set_adjoint_time_block(2, 7)
    mid = _model(ctx, w, g, med)
    ops = {
        "nl_fwd": lambda: _gpu_nl_fwd(ctx, mid, spos, rpos, src, g.nts, np.float32),
        "nl_adj": lambda: _gpu_nl_adj(ctx, mid, spos, rpos, dat, g.nts, np.float32),
        "born_fwd": lambda: _gpu_born_fwd(ctx, mid, spos, rpos, src_fine, dm5, g.nts, np.float32),
        "born_adj": lambda: _gpu_born_adj(ctx, mid, spos, rpos, src_fine, dat, g.nx, g.nz, np.float32),
    }
    eager = {k: f() for k, f in ops.items()}
    names = list(ops)
    for i, k in enumerate(names):
        for rep in range(3):
            ctx.set_graph(1)
            out = ops[k]()                       # each call synchronises its stream on return
            ctx.set_graph(0)
            ops[names[(i + 1) % len(names)]]()   # eager work between the graph calls
            assert all(np.array_equal(a, b) for a, b in zip(out, eager[k])), (k, rep)
    # one graph built per operator, replayed by its later calls
    builds, launches, updates = ctx.graph_info()
    assert (builds, launches) == (4, 12), (builds, launches, updates)


def test_graph_replay_of_a_batch_with_changing_inputs():
    """Graph mode on a two-shot batch: the first call builds the graph, later calls with new
    source traces (same device buffers) replay it unchanged and give the eager results."""
    cases = [_case(nz=60, nx=80, nts=60, sub=2, seed=120), _case(nz=40, nx=100, nts=50, sub=2, seed=121)]
    rng = np.random.default_rng(122)
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, _capi.VARIANT_AUTO)
    shots, srcs, recs, keep = [], [], [], []
    for (w, g, med, spos, rpos) in cases:
        mid = _model(ctx, w, g, med)
        src = [np.zeros((spos[GRID_OF_COMP[k]].size, g.nts), np.float32) for k in range(5)]
        rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, g.nts), np.float32) for k in range(5)]
        shot, kp = ctx.make_shot(mid, spos, rpos, src, rec)
        shots.append(shot)
        srcs.append(src)
        recs.append(rec)
        keep.append(kp)
    for rep in range(4):
        for (w, g, med, spos, rpos), src in zip(cases, srcs):
            new = random_traces(rng, spos, g.nts, np.float32)
            for k in range(5):
                src[k][...] = new[k]
        ctx.set_graph(0)
        ctx.nl_forward(shots)
        eager = [[r.copy() for r in rec] for rec in recs]
        ctx.set_graph(1)
        ctx.nl_forward(shots)
        for rec, ref in zip(recs, eager):
            assert all(np.array_equal(a, b) for a, b in zip(rec, ref)), rep
    builds, launches, updates = ctx.graph_info()
    assert (builds, launches) == (1, 4), (builds, launches, updates)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_grid_over_the_32bit_byte_limit_rejected(dtype):
    """A grid whose 5-field buffers reach 2^31 bytes is an argument error at model registration,
    before any allocation or launch (the kernels use 32-bit buffer offsets)."""
    isz = np.dtype(dtype).itemsize
    nz = 16384
    nx = (1 << 31) // (5 * nz * isz) + 1        # just over the limit
    media = np.empty((5, nx, nz), dtype)           # never touched: virtual memory only
    ctx = _capi.Context(0, dtype)
    with pytest.raises(_capi.EioError, match="2\\^31") as ei:
        ctx.model_add(nz, nx, 100, 2, 10.0, 10.0, 0.001, 20, media)
    assert "error 1:" in str(ei.value)          # EIO_ERR_ARG
    ctx.close()


def test_learned_boundary_rejected_outside_nl_forward():
    """With a learned boundary attached, the NL adjoint, Born and FWI calls refuse to run (their
    operator would not be the forward's adjoint); the NL forward runs; detached, all run."""
    from elastic_isotropic_operators_amd import abc_learn as A
    w, g, med, spos, rpos = _case(nts=40, sub=2, seed=130)
    ctx = _capi.Context(0, np.float32)
    mid = _model(ctx, w, g, med)
    rng = np.random.default_rng(131)
    src = [np.ascontiguousarray(x, np.float32) for x in random_traces(rng, spos, g.nts, np.float64)]
    rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, g.nts), np.float32) for k in range(5)]
    shot, keep = ctx.make_shot(mid, spos, rpos, src, rec)
    net = A.AbcNet("cuda", seed=3)
    ctx.set_abc(mid, 1, net.flat(), np.ones(10, np.float32))
    ctx.nl_forward([shot])
    dm = np.zeros((5, g.nx, g.nz), np.float32)
    for call in (lambda: ctx.nl_adjoint([shot]), lambda: ctx.born_adjoint([shot], dm),
                 lambda: ctx.born_forward([shot], dm), lambda: ctx.fwi_forward([shot], 0)):
        with pytest.raises(_capi.EioError, match="learned boundary"):
            call()
    ctx.set_abc(mid, 0)
    ctx.nl_adjoint([shot])
    h = ctx.fwi_forward([shot], 0)
    ctx.fwi_release(h)
    ctx.close()


def _persistent_batch(ctx, cases, seed):
    rng = np.random.default_rng(seed)
    shots, recs, keep = [], [], []
    for (w, g, med, spos, rpos) in cases:
        mid = _model(ctx, w, g, med)
        src = [np.ascontiguousarray(x, np.float32) for x in random_traces(rng, spos, g.nts, np.float64)]
        rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, g.nts), np.float32) for k in range(5)]
        shot, kp = ctx.make_shot(mid, spos, rpos, src, rec)
        shots.append(shot)
        recs.append(rec)
        keep.append((kp, src))
    return shots, recs, keep


@pytest.mark.parametrize("v2", [_capi.VARIANT_AUTO, 0, 4, 8, 10, 11, 12])
def test_persistent_launch_equals_launched_pairs(v2):
    """The whole two-step NL forward in one persistent launch (tiles wait for their neighbours'
    previous pair through flags) gives the launched form's traces bit for bit, on an uneven
    batch (grids, trace lengths, sub-samplings, odd step counts) with scattered sources and
    receivers, and on a batch with more tiles than co-resident workgroups (several tiles per
    workgroup)."""
    rng = np.random.default_rng(140)
    cases = []
    for i, (nz, nx, nts, sub) in enumerate([(90, 130, 70, 1), (150, 90, 50, 2), (60, 200, 61, 1)]):
        w = W.small(nz=nz, nx=nx, nts=nts, sub=sub, seed=141 + i)
        g = oracle_grid(w)
        med = O.media(g, w["elastic"], np.float64)
        cases.append((w, g, med, scattered_positions(rng, g.nz, g.nx, 12), scattered_positions(rng, g.nz, g.nx, 30)))
    big = W.small(nz=1400, nx=2400, nts=12, sub=1, seed=145)   # ~1000 tiles of 56 x 32
    gb = oracle_grid(big)
    cases.append((big, gb, O.media(gb, big["elastic"], np.float64), scattered_positions(rng, gb.nz, gb.nx, 40),
                  scattered_positions(rng, gb.nz, gb.nx, 80)))
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, v2)
    shots, recs, keep = _persistent_batch(ctx, cases, 146)
    ctx.nl_forward(shots)
    ref = [[r.copy() for r in rec] for rec in recs]
    ctx.set_persistent(1)
    for rep in range(2):
        for rec in recs:
            for r in rec:
                r[...] = np.nan
        ctx.nl_forward(shots)
        total_ms, nlaunch, _ = ctx.last_timing()
        assert nlaunch == 1
        for i, (rec, rr) in enumerate(zip(recs, ref)):
            assert all(np.array_equal(a, b) for a, b in zip(rec, rr)), (v2, rep, i)
    ctx.close()


def test_persistent_launch_c0_vs_fp64_oracle():
    """Workload C0 (one 280x472 shot, nts 1000, sub_fd 2) in one persistent launch against the
    FP64 oracle: receivers within 1e-5."""
    w = W.gaussian_c0()
    g = oracle_grid(w)
    med = O.media(g, w["elastic"], np.float64)
    spos = reg_positions(w, w["src_z"][0], w["src_x"][0])
    rpos = reg_positions(w, w["rec_z"], w["rec_x"])
    src = random_traces(np.random.default_rng(150), spos, g.nts, np.float64)
    O.set_ftz(True)
    ref = O.nl_fwd_reg(g, med, spos, src, rpos, np.float64)
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, _capi.VARIANT_AUTO)
    ctx.set_persistent(2)
    mid = _model(ctx, w, g, med)
    out = _gpu_nl_fwd(ctx, mid, spos, rpos, src, g.nts, np.float32)
    assert ctx.last_timing()[1] == 1           # one launch: the batch fits co-resident
    O.set_ftz(False)
    for k in range(5):
        assert rel_l2(out[k], ref[k]) < 1e-5, (k, rel_l2(out[k], ref[k]))
    ctx.close()


def test_persistent_launch_abort_drains_and_reports():
    """A spin limit of one poll forces the persistent launch's bounded neighbour waits to give
    up: every workgroup returns, the call reports EIO_ERR_RUNTIME, and the context runs the
    next call normally."""
    w, g, med, spos, rpos = _case(nz=300, nx=400, nts=200, sub=1, seed=160)
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, _capi.VARIANT_AUTO)
    mid = _model(ctx, w, g, med)
    rng = np.random.default_rng(161)
    src = random_traces(rng, spos, g.nts, np.float64)
    ref = _gpu_nl_fwd(ctx, mid, spos, rpos, src, g.nts, np.float32)
    ctx.set_persistent(1, 1)
    with pytest.raises(_capi.EioError, match="spin limit"):
        _gpu_nl_fwd(ctx, mid, spos, rpos, src, g.nts, np.float32)
    ctx.set_persistent(1, 1 << 20)
    out = _gpu_nl_fwd(ctx, mid, spos, rpos, src, g.nts, np.float32)
    assert all(np.array_equal(a, b) for a, b in zip(out, ref))
    ctx.close()


@pytest.mark.parametrize("v3", [0, 1, 2])
@pytest.mark.parametrize("nts,sub_fd", [(60, 2), (62, 2), (60, 3)])
def test_three_step_nl_forward_matches_oracle(v3, nts, sub_fd):
    """Three fine steps per launch (u1 on an 8-cell and u2 on a 4-cell halo, recomputed per
    tile) equal the oracle for the three tile widths (24, 32 and 8 columns), with sources and receivers scattered over the
    grid (many on tile edges: sources injected into every tile's copies of u1 and u2), and step
    counts 118, 122 and 177 (the last launch takes 1, 2 and 3 steps)."""
    w = W.small(nz=150, nx=130, nts=nts, sub=1, seed=62)
    g = O.grid(w["nz"], w["nx"], w["nts"], sub_fd, w["dz"], w["dx"], w["dts"], W.min_pad(w),
               w["alphaCos"], w["oz"], w["ox"])
    med = O.media(g, w["elastic"], np.float64)
    rng = np.random.default_rng(63 + v3)
    spos = scattered_positions(rng, g.nz, g.nx, 40)
    rpos = scattered_positions(rng, g.nz, g.nx, 60)
    src = random_traces(rng, spos, g.nts, np.float64)
    ref = O.nl_fwd_reg(g, med, spos, src, rpos, np.float64)
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(3, v3)
    mid = _model(ctx, w, g, med)
    out = _gpu_nl_fwd(ctx, mid, spos, rpos, src, g.nts, np.float32)
    nlaunch = ctx.last_timing()[1]
    assert nlaunch == -(-((g.nts - 1) * g.sub) // 3), nlaunch   # ceil(fine steps / 3) launches
    for k in range(5):
        assert np.linalg.norm(ref[k]) > 0
        assert rel_l2(out[k], ref[k]) < 1e-5, (v3, k, rel_l2(out[k], ref[k]))


def test_three_step_uneven_batch_matches_oracle():
    """An uneven batch (grids, trace lengths and sub-samplings differ, so shots finish in
    different launches, each with its own 1-, 2- or 3-step last launch) through the three-step
    forward equals the oracle per shot."""
    rng = np.random.default_rng(64)
    cases = []
    for i, (nz, nx, nts, sub) in enumerate([(90, 130, 70, 1), (150, 90, 50, 2), (60, 200, 61, 1)]):
        w = W.small(nz=nz, nx=nx, nts=nts, sub=sub, seed=65 + i)
        g = oracle_grid(w)
        med = O.media(g, w["elastic"], np.float64)
        cases.append((w, g, med, scattered_positions(rng, g.nz, g.nx, 12), scattered_positions(rng, g.nz, g.nx, 30)))
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(3, _capi.VARIANT_AUTO)
    shots, recs, refs, keep = [], [], [], []
    for (w, g, med, spos, rpos) in cases:
        mid = _model(ctx, w, g, med)
        src = random_traces(rng, spos, g.nts, np.float64)
        refs.append(O.nl_fwd_reg(g, med, spos, src, rpos, np.float64))
        rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, g.nts), np.float32) for k in range(5)]
        srcc = [np.ascontiguousarray(x, np.float32) for x in src]
        shot, kp = ctx.make_shot(mid, spos, rpos, srcc, rec)
        shots.append(shot)
        recs.append(rec)
        keep.append((kp, srcc))
    ctx.nl_forward(shots)
    for rec, ref in zip(recs, refs):
        for k in range(5):
            assert rel_l2(rec[k], ref[k]) < 1e-5, (k, rel_l2(rec[k], ref[k]))


@pytest.mark.parametrize("tb,tiny,streams", [(1, False, 2), (2, False, 2), (3, False, 2), (2, True, 2),
                                             (2, False, 1)])
def test_shot_groups_equal_one_group_and_the_oracle(tb, tiny, streams):
    """Shot groups (each group runs all its steps before the next starts) give the receiver
    traces of the one-group call bit for bit, and the oracle's, with one, two and three fine
    steps per launch, on an uneven batch that a small working-set budget splits into several
    groups (shots of different lengths inside one group finish in different launches); `tiny`:
    a budget below every shot's working set, so each shot is a group of its own; `streams`: the
    groups run on one stream or on two at once (each with half the budget)."""
    rng = np.random.default_rng(71)
    sizes = [(90, 130, 70, 1), (150, 90, 50, 2), (60, 200, 61, 1), (120, 100, 40, 2), (70, 70, 66, 1)]
    cases = []
    for i, (nz, nx, nts, sub) in enumerate(sizes):
        w = W.small(nz=nz, nx=nx, nts=nts, sub=sub, seed=72 + i)
        g = oracle_grid(w)
        med = O.media(g, w["elastic"], np.float64)
        cases.append((w, g, med, scattered_positions(rng, g.nz, g.nx, 8), scattered_positions(rng, g.nz, g.nx, 20)))
    ctx = _capi.Context(0, np.float32)
    # one kernel for both calls (the automatic two-step choice differs between a grouped and an
    # ungrouped call): variant 12, which the automatic choice takes for a grouped call
    ctx.set_time_block(tb, 12 if tb == 2 else _capi.VARIANT_AUTO)
    shots, recs, refs, keep = [], [], [], []
    for (w, g, med, spos, rpos) in cases:
        mid = _model(ctx, w, g, med)
        src = random_traces(rng, spos, g.nts, np.float64)
        refs.append(O.nl_fwd_reg(g, med, spos, src, rpos, np.float64))
        rec = [np.zeros((rpos[GRID_OF_COMP[k]].size, g.nts), np.float32) for k in range(5)]
        srcc = [np.ascontiguousarray(x, np.float32) for x in src]
        shot, kp = ctx.make_shot(mid, spos, rpos, srcc, rec)
        shots.append(shot)
        recs.append(rec)
        keep.append((kp, srcc))
    ctx.set_shot_group_streams(streams)
    ctx.set_shot_group_bytes(0)
    ctx.nl_forward(shots)
    assert ctx.last_shot_groups() == 1
    one = [[r.copy() for r in rec] for rec in recs]
    per_cell = (15 if tb == 1 else 25) * 4
    ctx.set_shot_group_bytes(1 if tiny else 2 * max(c[1].nz * c[1].nx for c in cases) * per_cell)
    for r in recs:
        for x in r:
            x[...] = 0
    ctx.nl_forward(shots)
    assert ctx.last_shot_groups() == len(cases) if tiny else ctx.last_shot_groups() >= 2, ctx.last_shot_groups()
    for rec, r1, ref in zip(recs, one, refs):
        for k in range(5):
            assert np.array_equal(rec[k], r1[k]), k
            assert rel_l2(rec[k], ref[k]) < 1e-5, (k, rel_l2(rec[k], ref[k]))
    if tb == 2:
        # the automatic variant of a grouped call is variant 12: the same traces again
        grouped = [[r.copy() for r in rec] for rec in recs]
        ctx.set_time_block(2, _capi.VARIANT_AUTO)
        ctx.nl_forward(shots)
        assert ctx.last_shot_groups() >= 2
        for rec, r1 in zip(recs, grouped):
            for k in range(5):
                assert np.array_equal(rec[k], r1[k]), k
    ctx.set_shot_group_bytes(_capi.GROUP_AUTO)
    ctx.close()


def test_pinned_host_buffers_round_trip():
    """eio_host_alloc: page-locked host memory (the operator modules' per-call trace buffers).
    A host -> device -> host copy through two such buffers returns the same bytes; 0 bytes
    gives NULL and freeing NULL is allowed."""
    import ctypes
    L = _capi.lib()
    p = ctypes.c_void_p(1)
    _capi.check(L.eio_host_alloc(0, ctypes.byref(p)))
    assert p.value is None
    _capi.check(L.eio_host_free(None))
    ctx = _capi.Context(0, np.float32)
    n = 1 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    _capi.check(L.eio_host_alloc(4 * n, ctypes.byref(a)))
    _capi.check(L.eio_host_alloc(4 * n, ctypes.byref(b)))
    d = ctx.alloc(4 * n)
    try:
        x = np.ctypeslib.as_array((ctypes.c_float * n).from_address(a.value))
        y = np.ctypeslib.as_array((ctypes.c_float * n).from_address(b.value))
        x[:] = np.random.default_rng(5).standard_normal(n).astype(np.float32)
        y[:] = 0
        _capi.check(L.eio_memcpy(ctx.h, ctypes.c_void_p(d), a, 4 * n, 1))
        _capi.check(L.eio_memcpy(ctx.h, b, ctypes.c_void_p(d), 4 * n, 2))
        _capi.check(L.eio_synchronize(ctx.h))
        assert np.array_equal(x, y)
    finally:
        ctx.free(d)
        _capi.check(L.eio_host_free(a))
        _capi.check(L.eio_host_free(b))
        ctx.close()
