"""The allocation-free NL forward (eio_nl_forward_noalloc): its workspace comes from the
library's per-GPU scratch array instead of hipMalloc.  Its traces equal eio_nl_forward's on
the same inputs (same kernels, so bit for bit), on two calls in a row, the second with a
smaller batch; a batch whose workspace exceeds the array is rejected before any launch."""
import numpy as np
import pytest

import oracle as O
from elastic_isotropic_operators_amd import _capi
from elastic_isotropic_operators_amd import workloads as W

from helpers import GRID_OF_COMP, oracle_grid, random_traces, reg_positions, scattered_positions

pytestmark = pytest.mark.gpu


def _upload(ctx, w, dtype, rng, scattered=0):
    """One shot of workload w with device traces; returns (shot, keep, receiver buffers)."""
    g = oracle_grid(w)
    med = O.media(g, w["elastic"], np.float64)
    mid = ctx.model_add(g.nz, g.nx, g.nts, g.sub, g.dz, g.dx, g.dts, g.minPad, med, g.alphaCos)
    if scattered:
        spos = scattered_positions(rng, g.nz, g.nx, scattered)
        rpos = scattered_positions(rng, g.nz, g.nx, 2 * scattered)
    else:
        spos = reg_positions(w, w["src_z"][0], w["src_x"][0])
        rpos = reg_positions(w, w["rec_z"], w["rec_x"])
    src = [np.ascontiguousarray(s, dtype) for s in random_traces(rng, spos, g.nts, np.float64)]
    dsrc, drec, shapes = [], [], []
    for k in range(5):
        p = ctx.alloc(max(src[k].nbytes, 1))
        ctx.h2d(p, src[k])
        dsrc.append(p)
        shape = (rpos[GRID_OF_COMP[k]].size, g.nts)
        shapes.append(shape)
        drec.append(ctx.alloc(max(int(np.prod(shape)) * np.dtype(dtype).itemsize, 1)))
    shot, keep = ctx.make_shot(mid, spos, rpos, dsrc, drec)
    return shot, (keep, spos, rpos, dsrc), (drec, shapes)


def _poison(ctx, recs, dtype):
    for drec, shapes in recs:
        for p, s in zip(drec, shapes):
            ctx.h2d(p, np.full(s, np.nan, dtype))


def _read(ctx, recs, dtype):
    out = []
    for drec, shapes in recs:
        for p, s in zip(drec, shapes):
            a = np.empty(s, dtype)
            ctx.d2h(a, p)
            out.append(a)
    return out


CASES = [("f32-one-step", np.float32, 1, 0), ("f32-two-step", np.float32, 2, 0),
         ("f32-persistent", np.float32, 2, 2), ("f64", np.float64, 1, 0)]


@pytest.mark.parametrize("name,dtype,tb,persistent", CASES, ids=[c[0] for c in CASES])
def test_noalloc_forward_equals_nl_forward(name, dtype, tb, persistent):
    ctx = _capi.Context(0, dtype)
    if dtype == np.float32:
        ctx.set_time_block(tb, _capi.VARIANT_AUTO)
        ctx.set_persistent(persistent)
    rng = np.random.default_rng(700)
    batches = [
        [W.small(nz=120, nx=260, nts=90, sub=2, seed=701), W.small(nz=70, nx=90, nts=71, sub=1, seed=702)],
        [W.small(nz=40, nx=56, nts=60, sub=1, seed=703)],   # second call: a smaller batch
    ]
    for bi, ws in enumerate(batches):
        shots, keep, recs = [], [], []
        for i, w in enumerate(ws):
            shot, kp, rec = _upload(ctx, w, dtype, rng, scattered=12 if i == 0 else 0)
            shots.append(shot)
            keep.append(kp)
            recs.append(rec)
        _poison(ctx, recs, dtype)
        ctx.nl_forward(shots, device_ptrs=True)
        ref = _read(ctx, recs, dtype)
        _poison(ctx, recs, dtype)
        ctx.nl_forward_noalloc(shots)
        got = _read(ctx, recs, dtype)
        assert sum(float(np.abs(r).sum()) for r in ref) > 0
        for j, (a, b) in enumerate(zip(got, ref)):
            assert np.array_equal(a, b), (name, bi, j)
    ctx.close()


def test_noalloc_rejects_batches_beyond_the_scratch():
    """A 1100 x 2100 shot needs 4 levels x 5 fields x 2.31 M cells x 4 B = 185 MB of state,
    more than the 64 MiB array: EIO_ERR_ARG naming the scratch, receivers untouched."""
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, _capi.VARIANT_AUTO)
    rng = np.random.default_rng(710)
    w = W.small(nz=1092, nx=2092, nts=20, sub=1, seed=711)
    shot, keep, rec = _upload(ctx, w, np.float32, rng)
    _poison(ctx, [rec], np.float32)
    with pytest.raises(_capi.EioError, match="scratch"):
        ctx.nl_forward_noalloc([shot])
    assert all(np.isnan(a).all() for a in _read(ctx, [rec], np.float32))
    ctx.close()
