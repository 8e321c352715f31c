"""Autotuner of the FP32 NL step kernels: picks the tile shape and the time blocking.

Candidates are every single-step variant (tile 64 x WX*R, R x-cells per lane, requested
waves per SIMD) and every two-step variant (owned tile 56 x WX*R2, two fine steps per
launch).  Each candidate runs the caller's own batch for a bounded number of fine steps
(``eio_ctx_set_step_limit``) and is timed with the HIP events of the context
(``eio_ctx_last_timing``); the fastest per fine step wins and is set on the context.
Results are cached per batch signature (grid sizes, step counts, device counts, the variant
tables and a hash of the built library) in
``$XDG_CACHE_HOME/eio/autotune.json`` (or ``~/.cache``), so later runs of the same batch
skip the sweep.  The reference has no tuning (fixed 16 x 16 CUDA blocks, varDeclare.h:5).
"""
import hashlib
import json
import os

from . import _capi


def _cache_path():
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "eio", "autotune.json")


_LIB_HASH = []


def _library_hash():
    """Hash of the built libeio_hip.so: a rebuilt library (its variant tables may differ, or
    its kernels run at other speeds) never reuses an older entry."""
    if not _LIB_HASH:
        h = hashlib.sha1()
        with open(_capi.LIB_PATH, "rb") as f:
            for chunk in iter(lambda: f.read(1 << 20), b""):
                h.update(chunk)
        _LIB_HASH.append(h.hexdigest()[:16])
    return _LIB_HASH[0]


def signature(grids, group_bytes=_capi.GROUP_AUTO, group_streams=2):
    """Batch key: sorted (nz, nx, ntw-1, n_src, n_rec) of the shots, the shot-group budget and
    streams, the single-, two- and three-step variant tables and the library build."""
    items = sorted(tuple(int(x) for x in g) for g in grids)
    key = {"shots": items, "group_bytes": int(group_bytes), "group_streams": int(group_streams),
           "variants": _capi.Context.variants(),
           "time_block_variants": _capi.Context.time_block_variants(),
           "time_block3_variants": _capi.Context.time_block3_variants(), "lib": _library_hash()}
    return hashlib.sha1(json.dumps(key).encode()).hexdigest()[:16]


def candidates():
    out = [(1, v) for v in range(len(_capi.Context.variants()))]
    out += [(2, v) for v in range(len(_capi.Context.time_block_variants()))]
    out += [(3, v) for v in range(len(_capi.Context.time_block3_variants()))]
    return out


def _apply(ctx, cand):
    tb, v = cand
    if tb == 1:
        ctx.set_time_block(1, 0)
        ctx.set_variant(v)
    else:
        ctx.set_time_block(tb, v)


def tune_nl_forward(ctx, shots, grids, steps=192, device_ptrs=True, use_cache=True, log=None,
                    group_bytes=_capi.GROUP_AUTO, group_streams=2):
    """Time every candidate on `shots` (eio_shot list of the caller's batch; outputs are
    overwritten with partial data) for `steps` fine steps and set the fastest on `ctx`.
    grids: per shot (nz, nx, ntw-1, n_src, n_rec) for the cache key; group_bytes,
    group_streams: the shot-group settings of `ctx` (part of the key: they change which form is
    fastest).
    Returns {"choice": (time_block, variant), "table": {cand: ms per fine step}, "cached": bool}."""
    key = signature(grids, group_bytes, group_streams)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            cache = json.load(open(path))
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        choice = tuple(cache[key]["choice"])
        _apply(ctx, choice)
        return {"choice": choice, "table": cache[key]["table"], "cached": True}
    table = {}
    nfine = max(1, min(int(steps), max(int(g[2]) for g in grids)))
    try:
        ctx.set_step_limit(steps)
        for cand in candidates():
            _apply(ctx, cand)
            ctx.nl_forward(shots, device_ptrs=device_ptrs)   # plan, workspace, first-touch
            ctx.nl_forward(shots, device_ptrs=device_ptrs)
            total_ms, nlaunch, step_ms = ctx.last_timing()
            per_step = step_ms * nlaunch / float(nfine)
            table["%d:%d" % cand] = per_step
            if log:
                log("autotune %s: %.4f ms per fine step" % (cand, per_step))
    finally:
        ctx.set_step_limit(0)
    best = min(table, key=table.get)
    choice = tuple(int(x) for x in best.split(":"))
    _apply(ctx, choice)
    if use_cache:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            cache[key] = {"choice": list(choice), "table": table}
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(cache, f)
            os.replace(tmp, path)
        except OSError:
            pass
    return {"choice": choice, "table": table, "cached": False}
