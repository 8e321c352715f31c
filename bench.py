#!/usr/bin/env python3
"""Benchmark of the elastic propagator on one MI355X.

Headline (value, GCell/s): workload "c2", 64 layered shots of uneven size (interior nz
300-600, nx 600-1800, nts 2000-4000 at 1 ms, sub=2, Ricker 8 Hz, 40-cell sponge), FP32, all
shots propagated by the nonlinear forward operator in one batched time loop.  A "step" is
one full batched propagation of the 64 shots.  Inputs (media, source traces) are resident
in HBM before the timed region; outputs stay in HBM.

FWI leg (fwi_gradient, s/shot): the 32-shot elastic FWI gradient (600x1592 padded grid,
nts=3000, sub=2) through the torch FWI engine (ElasticFwi.misfit_and_gradient: shot groups
sized to HBM keep their whole coarse record, two fine steps per launch in the forward and the
receiver pass), plus the relative L2 distance of the FP32 gradient from the FP64 gradient on
two of the shots.  It runs in a child process because torch must be imported before the
propagator library is loaded (both then share torch's HIP runtime), while the headline leg and
its profiles use the system runtime without torch.

Born leg (born): workload C1 (8 shots on the 280x472 grid) through the drop-in float and
double Born operators, forward and adjoint wall times and FP32 vs FP64.  Learned-boundary
leg (learned_boundary): reflection energy and device time of the plain and learned absorbing
boundaries on the 1000x3000 model (needs the committed weights).

Prints ONE JSON line with GCell/s (cell-update = one update-region cell for one fine step),
the HBM roofline of the fused step kernel (80 algorithmic bytes per FP32 cell-update,
timed with HIP events on the kernel's stream), the legs and a CPU baseline: the FP32
OpenMP oracle on a bounded sample of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from elastic_isotropic_operators_amd import _capi, autotune, geometry as G, workloads as W  # noqa: E402

BYTES_PER_CELL = {"nl_fwd": 80}
HBM_PEAK_GBS = 8000.0
METRIC = "GCell-updates/s + FWI-gradient s/shot at \u22641e-4 rel-L2 vs FP64, 1 MI355X"
# measured bytes per cell-update of the dominant kernel per (time block, variant, grouping)
TRAFFIC_FILES = [os.path.join(ROOT, "profiles", f) for f in
                 ("k_step2_fwd_layered_batch_c2_traffic.json", "k_step2_fwd_fs_layered_batch_c2_grouped_traffic.json",
                  "k_step2_fwd_fs56x8_layered_batch_c2_grouped_traffic.json",
                  "k_step2_fwd_fs56x8_layered_batch_c2_grouped_2streams_traffic.json")]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def prepare_shots(ctx, shots_w, dtype=np.float32):
    """Upload media and source traces; allocate receiver traces.
    Returns (shots, keep, cells, grids) with grids = per shot (nz, nx, ntw-1, n_src, n_rec)."""
    shots, keep, cells, grids = [], [], 0, []
    for w in shots_w:
        sub_fd = 2 * w["sub"]
        dtw = G.dtw_of(w["dts"], w["sub"])
        media = G.scaled_media(w["elastic"], dtw, dtype)
        mid = ctx.model_add(w["nz"], w["nx"], w["nts"], sub_fd, w["dz"], w["dx"], w["dts"],
                            W.min_pad(w), media, w["alphaCos"])
        si = G.staggered_interps(w["nz"], w["oz"], w["dz"], w["ox"], w["dx"], w["src_z"][0],
                                 w["src_x"][0])
        ri = G.staggered_interps(w["nz"], w["oz"], w["dz"], w["ox"], w["dx"], w["rec_z"], w["rec_x"])
        src = G.nl_source_traces(si, media, w["model"], dtw, w["dx"], w["dz"], dtype)
        spos = [ip.uniq for ip in si]
        rpos = [ip.uniq for ip in ri]
        dsrc, drec = [], []
        for k in range(5):
            p = ctx.alloc(src[k].nbytes)
            ctx.h2d(p, src[k])
            dsrc.append(p)
            drec.append(ctx.alloc(rpos[G.GRID_OF_COMP[k]].size * w["nts"] * np.dtype(dtype).itemsize))
        shot, kp = ctx.make_shot(mid, spos, rpos, dsrc, drec)
        keep.append((kp, spos, rpos, dsrc, drec))
        shots.append(shot)
        cells += W.cell_updates(w)
        grids.append((w["nz"], w["nx"], (w["nts"] - 1) * sub_fd, sum(p.size for p in spos),
                      sum(p.size for p in rpos)))
    return shots, keep, cells, grids


def cpu_baseline(shots_w, target_s=12.0):
    """FP32 OpenMP oracle on the workload's first shots, in order, each with its trace length
    cut to what the remaining time budget allows, until about `target_s` seconds of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    O.set_ftz(True)

    def per_coarse(w):
        return (w["nz"] - 8) * (w["nx"] - 8) * 2 * w["sub"]

    def run(w, nts):
        g = O.grid(w["nz"], w["nx"], nts, 2 * w["sub"], w["dz"], w["dx"], w["dts"], W.min_pad(w),
                   w["alphaCos"], w["oz"], w["ox"])
        t0 = time.perf_counter()
        O.nl_fwd_shot(g, w["elastic"], w["src_z"][0], w["src_x"][0], w["rec_z"], w["rec_x"],
                      w["model"][:, :, :nts], np.float32)
        return time.perf_counter() - t0, per_coarse(w) * (nts - 1)

    el, cells = run(shots_w[0], 20)
    rate = cells / max(el, 1e-9)
    total_el, total_cells, parts = 0.0, 0, []
    for i, w in enumerate(shots_w[:8]):
        left = target_s - total_el
        if left < 0.5:
            break
        nts = int(min(w["nts"], max(20, left * rate / per_coarse(w))))
        el, cells = run(w, nts)
        total_el += el
        total_cells += cells
        rate = total_cells / total_el
        parts.append("shot %d (%dx%d, %d of %d coarse samples)" % (i, w["nz"], w["nx"], nts, w["nts"]))
    model = ""
    try:  # the host's CPU model string (as lscpu prints it)
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=total_cells / total_el / 1e9, unit="GCell/s", cores=O.num_threads(), kind="port",
                sample="%s of the workload, FP32 oracle with flush-to-zero, %.1f s" % (", ".join(parts), total_el),
                host_cpu=model, host_nproc=os.cpu_count())


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    O.set_ftz(True)
    return O


def _oracle_grid(O, w, nts=None):
    return O.grid(w["nz"], w["nx"], w["nts"] if nts is None else nts, 2 * w["sub"], w["dz"], w["dx"], w["dts"],
                  W.min_pad(w), w["alphaCos"], w["oz"], w["ox"])


def c0_leg(args):
    """Unbatched leg (workload C0: one 280x472 shot, nts 1000, sub_fd 2): the two-step NL
    forward with one launch per pair of fine steps, the same loop replayed from the context's
    cached hipGraph, the whole loop in one persistent launch, and the three-step forward (one
    launch per three fine steps); device time per call (HIP events) after warm-up calls, best
    of `args.c0_reps`; plus the FP32 oracle on the host cores (the whole shot)."""
    w = W.gaussian_c0()
    cells = W.cell_updates(w)
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, _capi.VARIANT_AUTO)
    shots, keep, _, _ = prepare_shots(ctx, [w])
    out = {"workload": w["name"], "cell_updates": cells, "dtype": "f32"}
    for mode in ("launched", "graph", "persistent", "three_step"):
        ctx.set_graph(mode == "graph")
        ctx.set_persistent(1 if mode == "persistent" else 0)
        ctx.set_time_block(3 if mode == "three_step" else 2, _capi.VARIANT_AUTO)
        for _ in range(2):
            ctx.nl_forward(shots, device_ptrs=True)
        ms = []
        for _ in range(args.c0_reps):
            ctx.nl_forward(shots, device_ptrs=True)
            tot, nl, _ = ctx.last_timing()
            ms.append(tot)
        best = float(min(ms))
        out[mode] = {"ms": round(best, 3), "gcell_per_s": round(cells / best * 1e-6, 2), "launches": int(nl)}
        log("c0 %s: %.3f ms per propagation (%d launches), %.2f GCell/s" % (mode, best, nl, cells / best * 1e-6))
    b, l, u = ctx.graph_info()
    out["graph"]["builds"], out["graph"]["replays"] = b, l - b
    ctx.set_graph(0)
    ctx.set_persistent(0)
    ctx.close()
    if not args.no_cpu:
        O = _oracle()
        t0 = time.perf_counter()
        O.nl_fwd_shot(_oracle_grid(O, w), w["elastic"], w["src_z"][0], w["src_x"][0], w["rec_z"], w["rec_x"],
                      w["model"], np.float32)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cells / el / 1e9, 3), "unit": "GCell/s", "cores": O.num_threads(),
                               "kind": "port", "sample": "the whole shot, FP32 oracle with flush-to-zero, %.1f s" % el}
        log("c0 cpu: %.1f s, %.3f GCell/s" % (el, cells / el / 1e9))
    return out


def fwi_cpu_baseline(w0, target_s=10.0):
    """FP32 oracle FWI gradient of shot 0 on the host cores (NL forward of the predicted data +
    Born adjoint of the residual: the reference's three propagation passes), trace length cut to
    about `target_s` seconds and extrapolated linearly in nts to the full length."""
    O = _oracle()

    def run(nts):
        g = _oracle_grid(O, w0, nts)
        m = np.ascontiguousarray(w0["model"][:, :, :nts])
        t0 = time.perf_counter()
        d = O.nl_fwd_shot(g, w0["elastic"], w0["src_z"][0], w0["src_x"][0], w0["rec_z"], w0["rec_x"], m, np.float32)
        O.born_adj_shots(g, w0["elastic"], w0["src_z"][:1], w0["src_x"][:1], m, w0["rec_z"], w0["rec_x"],
                         d[None].astype(np.float32), np.float32)
        return time.perf_counter() - t0

    el = run(21)
    n = int(min(w0["nts"], max(21, 1 + 20 * target_s / max(el, 1e-3))))
    el = run(n)
    s_shot = el * (w0["nts"] - 1) / (n - 1)
    return {"value": round(s_shot, 2), "unit": "s/shot", "cores": O.num_threads(), "kind": "port",
            "sample": "shot 0, %d of %d coarse samples (%.1f s), extrapolated linearly in nts; FP32 oracle, "
                      "NL forward + Born adjoint" % (n, w0["nts"], el)}


def fwi_child(args):
    """FWI gradient leg (child process): torch first, then the propagator library."""
    import torch
    from elastic_isotropic_operators_amd.fwi import ElasticFwi, datcomp

    t0 = time.perf_counter()
    wt, w0 = W.fwi_c3(nshots=args.fwi_shots, scale=args.fwi_scale)
    grid = dict(nz=w0["nz"], nx=w0["nx"], dz=w0["dz"], dx=w0["dx"], oz=w0["oz"], ox=w0["ox"],
                nts=w0["nts"], dts=w0["dts"], sub=w0["sub"], minPad=W.min_pad(w0), alphaCos=w0["alphaCos"])
    comps = "vx,vz,p"

    adj = tuple(int(x) for x in args.fwi_adj.split(":"))

    def engine(dtype, sel):
        return ElasticFwi(grid, w0["src_z"][sel], w0["src_x"][sel], w0["model"], w0["rec_z"], w0["rec_x"],
                          dtype=dtype, adj_time_block=adj)

    eng = engine(torch.float32, slice(None))
    with torch.no_grad():
        d_obs = datcomp(eng.forward_data(torch.tensor(wt["elastic"], device="cuda"))[0], comps)
    torch.cuda.synchronize()
    log("fwi: workload %s, observed data in %.1f s" % (w0["name"], time.perf_counter() - t0))
    m = torch.tensor(w0["elastic"], device="cuda")
    group = eng.default_group()
    log("fwi: %d shots per full-storage group (at most %d fit in memory)" % (group, eng.full_storage_group()))
    g32 = None

    def gradient():
        nonlocal g32
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, g32 = eng.misfit_and_gradient(m, d_obs, comps, group=group)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    for i in range(args.fwi_warmup):
        log("fwi warmup %d: %.2f s" % (i, gradient()))
    times = []
    for i in range(args.fwi_steps):
        times.append(gradient())
        log("fwi gradient %d: %.2f s (%.3f s/shot)" % (i, times[-1], times[-1] / eng.nshot))
    L, nck, nbytes = eng.last_store_info
    # per shot: forward (nts-1 coarse intervals) + recompute of every segment but the last
    # (L intervals each; none with full storage) + adjoint, at cell_updates / (nts-1)
    # cell-updates per interval
    per_interval = W.cell_updates(w0) / (w0["nts"] - 1)
    grad_cells = eng.nshot * per_interval * (2 * (w0["nts"] - 1) + L * nck)
    g32 = g32.double()
    # FP32 vs FP64 on two of the shots at full size
    sel = [0, eng.nshot // 2]
    gs = {}
    for dt in (torch.float64, torch.float32):
        e2 = engine(dt, sel)
        m2 = torch.tensor(w0["elastic"], dtype=dt, device="cuda")
        gs[dt] = e2.misfit_and_gradient(m2, d_obs[sel].to(dt), comps)[1].double()
        del e2, m2
    rel = float(torch.linalg.norm(gs[torch.float32] - gs[torch.float64]) / torch.linalg.norm(gs[torch.float64]))
    log("fwi: FP32 vs FP64 gradient on shots %s: rel L2 %.2e" % (sel, rel))
    t = float(np.mean(times))
    # checkpointed form (the configuration's sqrt(nts) snapshots): the autograd path over all
    # shots at once, torch.autograd.Function -> eio_fwi_forward / eio_fwi_gradient
    ck = None
    if args.fwi_segment >= 0:
        eng.ctx.release_workspace()
        seg = args.fwi_segment or int(np.ceil(np.sqrt(w0["nts"] - 1)))
        engc = ElasticFwi(grid, w0["src_z"], w0["src_x"], w0["model"], w0["rec_z"], w0["rec_x"],
                          dtype=torch.float32, segment=seg, adj_time_block=adj)

        def grad_ck():
            mm = torch.tensor(w0["elastic"], device="cuda", requires_grad=True)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            engc.misfit(mm, d_obs, components=comps).backward()
            torch.cuda.synchronize()
            return time.perf_counter() - t1, mm.grad.double()

        tw, _ = grad_ck()
        log("fwi checkpointed warmup: %.2f s" % tw)
        tc, gck = grad_ck()
        Lc, nckc, nbc = engc.last_store_info
        rel_ck = float(torch.linalg.norm(gck - g32) / torch.linalg.norm(g32))
        log("fwi checkpointed (interval %d, %d checkpoints, %.1f GB): %.2f s (%.3f s/shot), rel L2 vs full "
            "storage %.2e" % (Lc, nckc, nbc / 1e9, tc, tc / engc.nshot, rel_ck))
        ck = {"s_per_shot": round(tc / engc.nshot, 4), "s_per_gradient": round(tc, 3), "checkpoint_interval": Lc,
              "checkpoints": nckc, "store_bytes": nbc, "rel_l2_vs_full_storage_fp32": rel_ck,
              "gcell_per_s": round(eng.nshot * per_interval * (2 * (w0["nts"] - 1) + Lc * nckc) / tc / 1e9, 2),
              "method": "torch.autograd.Function over all shots, forward wavefield checkpointed"}
        engc.ctx.release_workspace()
        del engc
    out = {"s_per_shot": round(t / eng.nshot, 4), "s_per_gradient": round(t, 3), "shots": eng.nshot,
           "workload": w0["name"], "dtype": "f32", "steps": args.fwi_steps, "warmup": args.fwi_warmup,
           "gcell_per_s": round(grad_cells / t / 1e9, 2), "rel_l2_vs_fp64": rel, "fp64_check_shots": sel,
           "checkpoint_interval": L, "checkpoints": nck, "store_bytes": nbytes, "shots_per_group": group,
           "method": "shot groups sized to the Infinity Cache (records within HBM), full coarse Vx/Vz storage, gradient per group",
           "adjoint_time_block": list(adj),
           "full_storage_bytes": 2 * 4 * w0["nz"] * w0["nx"] * w0["nts"] * eng.nshot,
           "grad_norm": float(torch.linalg.norm(g32)), "checkpointed": ck}
    if not args.no_cpu:
        t0 = time.perf_counter()
        out["cpu_baseline"] = fwi_cpu_baseline(w0)
        log("fwi cpu baseline: %.1f s/shot (%.1f s)" % (out["cpu_baseline"]["value"], time.perf_counter() - t0))
    print("FWI_JSON " + json.dumps(out), flush=True)


def run_fwi_leg(args):
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--fwi-child", "--fwi-shots", str(args.fwi_shots),
           "--fwi-steps", str(args.fwi_steps), "--fwi-warmup", str(args.fwi_warmup),
           "--fwi-scale", str(args.fwi_scale), "--fwi-adj", args.fwi_adj, "--fwi-segment", str(args.fwi_segment)]
    if args.no_cpu:
        cmd.append("--no-cpu")
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
    for line in r.stdout.splitlines():
        if line.startswith("FWI_JSON "):
            return json.loads(line[len("FWI_JSON "):])
    raise RuntimeError("FWI leg failed (exit %d)" % r.returncode)


def born_child(args):
    """Born leg (child process): workload C1 (the C0 grid and background, 8 shots at 50-cell
    spacing, the seeded 5 % perturbation) through the drop-in operators (parfile ->
    BornOpInit{Float,Double} -> BornElasticShotsGpu), as a reference script runs them: Born
    forward of the perturbation, then the Born adjoint of that data.  Wall time of each call
    (each returns synchronised), after one warm-up call; FP32 against the double operator."""
    import tempfile
    from elastic_isotropic_operators_amd import Elastic_iso_double_prop as PD
    from elastic_isotropic_operators_amd import Elastic_iso_float_prop as PF
    from elastic_isotropic_operators_amd.sep import Hypercube, SepVector, genericIO
    w = W.gaussian_c1()
    dm = w["dm"]
    res = {}
    with tempfile.TemporaryDirectory() as td:
        hyp = Hypercube.hypercube(axes=[Hypercube.axis(n=w["nz"], o=w["oz"], d=w["dz"]),
                                        Hypercube.axis(n=w["nx"], o=w["ox"], d=w["dx"]), Hypercube.axis(n=3)])
        mod = SepVector.getSepVector(hyp)
        mod.getNdArray()[...] = w["vvr"]
        mfile = os.path.join(td, "model.H")
        genericIO.defaultIO.writeVector(mfile, mod)
        src = SepVector.getSepVector(Hypercube.hypercube(
            axes=[Hypercube.axis(n=w["nts"], d=w["dts"]), Hypercube.axis(n=1), Hypercube.axis(n=5)]))
        src.getNdArray()[...] = w["model"]
        sfile = os.path.join(td, "sources.H")
        genericIO.defaultIO.writeVector(sfile, src)
        par = dict(nts=w["nts"], dts=w["dts"], sub=w["sub"], nz=w["nz"], nx=w["nx"], dz=w["dz"], dx=w["dx"],
                   fMax=w["fMax"], mod_par=1, zPadMinus=w["zPadMinus"], zPadPlus=w["zPadPlus"],
                   xPadMinus=w["xPadMinus"], xPadPlus=w["xPadPlus"], nExp=8, zSource=3, xSource=26,
                   spacingShots=50, nReceiver=200, depthReceiver=1, oReceiver=1, dReceiver=2, iGpu=0,
                   blockSize=16, elasticParam=mfile, sources=sfile)
        cargs = ["%s=%s" % kv for kv in par.items()]
        for P, init, name in ((PF, PF.BornOpInitFloat, "f32"), (PD, PD.BornOpInitDouble, "f64")):
            (model, data, elastic, parobj, sig, sc, sx, sz, sxz, rc, rx, rz, rxz, _) = init(cargs)
            op = P.BornElasticShotsGpu(model, data, elastic, parobj, sig, sc, sx, sz, sxz, rc, rx, rz, rxz)
            model.getNdArray()[...] = dm
            op.forward(False, model, data)
            t0 = time.perf_counter()
            op.forward(False, model, data)
            tf = time.perf_counter() - t0
            df = op.pyOp.lastPropagationMs() / 1e3   # device time of the compiled operator's last call
            img = model.clone()
            op.adjoint(False, img, data)
            t0 = time.perf_counter()
            op.adjoint(False, img, data)
            ta = time.perf_counter() - t0
            da = op.pyOp.lastPropagationMs() / 1e3
            res[name] = (tf, ta, data.getNdArray().astype(np.float64), img.getNdArray().astype(np.float64), df, da)
            log("born %s: forward %.3f s (device %.3f s), adjoint %.3f s (device %.3f s)" % (name, tf, df, ta, da))
    if not args.no_cpu:
        # FP32 oracle Born forward + adjoint on the host cores, shot by shot until ~10 s,
        # extrapolated to the 8 shots
        O = _oracle()
        g = _oracle_grid(O, w)
        sz, sx = W.source_coords(w, 3, 26, nExp=8, spacingShots=50)
        rz, rx = W.receiver_coords(w, 200, 1, 1, 2)
        dsh = res["f32"][2]
        el, n = 0.0, 0
        while n < 8 and el < 10.0:
            t0 = time.perf_counter()
            O.born_fwd_shot(g, w["elastic"], sz[n], sx[n], w["model"], rz, rx, dm, np.float32)
            O.born_adj_shots(g, w["elastic"], sz[n:n + 1], sx[n:n + 1], w["model"], rz, rx,
                             dsh[n:n + 1].astype(np.float32), np.float32)
            el += time.perf_counter() - t0
            n += 1
        cpu = {"value": round(el / n * 8, 2), "unit": "s per Born forward + adjoint of the 8 shots",
               "cores": O.num_threads(), "kind": "port",
               "sample": "%d of 8 shots (%.1f s), extrapolated; FP32 oracle" % (n, el)}
        log("born cpu: %.2f s per forward + adjoint of 8 shots (%d shots timed)" % (cpu["value"], n))
    ntw = (w["nts"] - 1) * 2 * w["sub"] + 1
    cells = 8 * (w["nz"] - 8) * (w["nx"] - 8) * (ntw - 1)   # one pass of the 8 shots
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    tf, ta, d32, i32, df, da = res["f32"]
    out = {"workload": w["name"], "shots": 8, "dtype": "f32", "cell_updates_per_pass": cells,
           "forward_s": round(tf, 4), "adjoint_s": round(ta, 4),
           "forward_gcell_per_s": round(2 * cells / tf / 1e9, 2),   # source + scattered pass
           "adjoint_gcell_per_s": round(2 * cells / ta / 1e9, 2),   # source + receiver pass
           "device_forward_s": round(df, 4), "device_adjoint_s": round(da, 4),
           "device_forward_gcell_per_s": round(2 * cells / df / 1e9, 2),
           "device_adjoint_gcell_per_s": round(2 * cells / da / 1e9, 2),
           "f64_forward_s": round(res["f64"][0], 4), "f64_adjoint_s": round(res["f64"][1], 4),
           "rel_l2_data_f32_vs_f64": rel(d32, res["f64"][2]), "rel_l2_image_f32_vs_f64": rel(i32, res["f64"][3]),
           "timing": "forward_s/adjoint_s: wall time of the operator call (host interpolation and transfers "
                     "included); device_*: the call's device time on the context stream (HIP events)"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu
    print("BORN_JSON " + json.dumps(out), flush=True)


def run_born_leg(args):
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--born-child"] + (["--no-cpu"] if args.no_cpu else [])
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=600)
    for line in r.stdout.splitlines():
        if line.startswith("BORN_JSON "):
            return json.loads(line[len("BORN_JSON "):])
    raise RuntimeError("Born leg failed (exit %d)" % r.returncode)


ABC_WEIGHTS = os.path.join(ROOT, "elastic_isotropic_operators_amd", "data", "abc_strip10.npz")


def abc_child(args):
    """Learned-boundary leg (child process): one shot on the 1000 x 3000 layered model with
    the wide (120-cell) sponge as reference, the plain 10- and 60-cell sponges and the learned
    10-cell boundary; reflection energy at receivers inside the model and device time."""
    import torch  # noqa: F401  (before the propagator library)
    from elastic_isotropic_operators_amd import abc_learn as A
    net, scale, strip, on = A.load(ABC_WEIGHTS)
    vvr, nts = W.learned_abc_c4(scale=args.abc_scale)
    res = A.reflection_test(vvr, net, scale, strip=strip, mid=60, wide=120, nts=nts, rec_every=4, on=on,
                            time_block=(1, 2), log=log)
    out = {"workload": "learned_abc_c4_%dx%d_nts%d" % (vvr.shape[2], vvr.shape[1], nts), "strip": strip,
           "runs": res, "weights": os.path.relpath(ABC_WEIGHTS, ROOT)}
    print("ABC_JSON " + json.dumps(out), flush=True)


def run_abc_leg(args):
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--abc-child", "--abc-scale", str(args.abc_scale)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
    for line in r.stdout.splitlines():
        if line.startswith("ABC_JSON "):
            return json.loads(line[len("ABC_JSON "):])
    raise RuntimeError("learned-boundary leg failed (exit %d)" % r.returncode)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c2", "c0"])
    ap.add_argument("--scale", type=float, default=1.0, help="shrink every size (quick runs)")
    ap.add_argument("--nshots", type=int, default=64)
    ap.add_argument("--graph", type=int, default=0)
    ap.add_argument("--variant", type=int, default=-1, help="FP32 step-kernel variant (-1: default)")
    ap.add_argument("--time-block", type=int, default=1, help="fine steps per launch (1, 2 or 3)")
    ap.add_argument("--group-bytes", type=int, default=-1,
                    help="working-set budget of the batch's shot groups (-1: automatic, 192 MiB; 0: one group)")
    ap.add_argument("--group-streams", type=int, default=2, help="streams the shot groups run on (1 or 2)")
    ap.add_argument("--variant2", type=int, default=0, help="two-step kernel variant")
    ap.add_argument("--variant3", type=int, default=1, help="three-step kernel variant")
    ap.add_argument("--autotune", type=int, default=1,
                    help="pick time block and variant by timing every candidate (overrides the flags)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nts-cap", type=int, default=0, help="cap every trace length (profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=1, help="measure a plain copy kernel")
    ap.add_argument("--fwi", type=int, default=1, help="run the FWI gradient leg")
    ap.add_argument("--fwi-shots", type=int, default=32)
    ap.add_argument("--fwi-steps", type=int, default=2)
    ap.add_argument("--fwi-warmup", type=int, default=1)
    ap.add_argument("--fwi-scale", type=float, default=1.0)
    ap.add_argument("--fwi-adj", default="2:-1",
                    help="reverse fine steps per launch of the gradient's receiver pass and the two-step "
                         "adjoint variant, as steps:variant")
    ap.add_argument("--fwi-segment", type=int, default=0,
                    help="checkpoint interval of the checkpointed FWI gradient (0: ceil(sqrt(nts-1)); -1: skip it)")
    ap.add_argument("--c0", type=int, default=1, help="run the unbatched C0 leg (launched, graph, persistent, three-step)")
    ap.add_argument("--c0-reps", type=int, default=5)
    ap.add_argument("--fwi-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--abc", type=int, default=1, help="run the learned-boundary leg (needs its weights)")
    ap.add_argument("--born", type=int, default=1, help="run the Born operator leg (workload C1)")
    ap.add_argument("--born-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--abc-scale", type=float, default=1.0)
    ap.add_argument("--abc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.fwi_child:
        fwi_child(args)
        return
    if args.born_child:
        born_child(args)
        return
    if args.abc_child:
        abc_child(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: it runs on a single MI355X; over several GPUs the "
              "shots would be partitioned (one batch per GPU) with one all-reduce of the FWI "
              "gradient and misfit.", file=sys.stderr)
        sys.exit(2)

    t_build = time.perf_counter()
    if args.workload == "c2":
        shots_w = W.layered_batch_c2(args.nshots, seed=2, scale=args.scale)
        wl_name = "layered_batch_c2_%dshots" % args.nshots
        if args.nts_cap:
            for w in shots_w:
                w["nts"] = min(w["nts"], args.nts_cap)
                w["model"] = np.ascontiguousarray(w["model"][:, :, :w["nts"]])
            wl_name += "_nts%d" % args.nts_cap
    else:
        shots_w = [W.gaussian_c0()]
        wl_name = shots_w[0]["name"]
    log("workload %s built in %.1f s" % (wl_name, time.perf_counter() - t_build))

    ctx = _capi.Context(0, np.float32)
    ctx.set_graph(args.graph)
    if args.variant >= 0:
        ctx.set_variant(args.variant)
    ctx.set_time_block(args.time_block, args.variant3 if args.time_block == 3 else args.variant2)
    ctx.set_shot_group_bytes(args.group_bytes)
    ctx.set_shot_group_streams(args.group_streams)
    shots, keep, cells, grids = prepare_shots(ctx, shots_w)
    log("uploaded %d shots, %.3e cell-updates per step" % (len(shots), cells))
    tuned = None
    if args.autotune:
        tuned = autotune.tune_nl_forward(ctx, shots, grids, log=log, group_bytes=args.group_bytes,
                                         group_streams=args.group_streams)
        args.time_block, v = tuned["choice"]
        if args.time_block == 3:
            args.variant3 = v
        elif args.time_block == 2:
            args.variant2 = v
        else:
            args.variant = v
        log("autotune choice: time block %d, variant %d%s" % (args.time_block, v,
                                                             " (cached)" if tuned["cached"] else ""))

    for i in range(args.warmup):
        t0 = time.perf_counter()
        ctx.nl_forward(shots, device_ptrs=True)
        log("warmup %d: %.3f s" % (i, time.perf_counter() - t0))
    _capi.check(_capi.lib().eio_synchronize(ctx.h))
    times, step_ms, nlaunch = [], [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        ctx.nl_forward(shots, device_ptrs=True)
        _capi.check(_capi.lib().eio_synchronize(ctx.h))
        dt = time.perf_counter() - t0
        tot, nst, sms = ctx.last_timing()
        times.append(dt)
        step_ms.append(sms)
        nlaunch.append(nst)
        log("step %d: %.3f s, %.2f GCell/s, loop %.1f ms over %d launches" %
            (i, dt, cells / dt / 1e9, sms * nst, nst))
    elapsed = sum(times)
    ms_per_step = elapsed / args.steps * 1e3
    value = cells * args.steps / elapsed / 1e9
    loop_s = float(np.mean([s * n for s, n in zip(step_ms, nlaunch)])) / 1e3
    achieved = BYTES_PER_CELL["nl_fwd"] * cells / loop_s / 1e9
    if (args.time_block == 3 and args.variant3 < 0) or (args.time_block == 2 and args.variant2 < 0):
        kname = "k_step%d_fwd (the library's automatic variant per call, %d fine steps per launch)" % (
            args.time_block, args.time_block)
    elif args.time_block == 3:
        R3, WX3, OCC3 = _capi.Context.time_block3_variants()[args.variant3]
        kname = "k_step3_fwd<float,%d,%d,%d> (3 fine steps per launch)" % (R3, WX3, OCC3)
    elif args.time_block == 2:
        R2, WX2, OCC2 = _capi.Context.time_block_variants()[args.variant2]
        layout = _capi.Context.time_block_layout(args.variant2)
        if layout == 2:
            kname = "k_step2q_fwd<float,%d> (56x24 tile, 4-row lane quads, 2 fine steps per launch)" % OCC2
        else:
            kname = "k_step2_fwd<float,%d,%d,%d%s> (2 fine steps per launch)" % (
                R2, WX2, OCC2, ",FS" if layout == 1 else "")
    else:
        R, WX, OCC = _capi.Context.variants()[args.variant if args.variant >= 0 else 1]
        kname = "k_step_fwd<float,%d,%d,%d,M_NL>" % (R, WX, OCC)
    groups = ctx.last_shot_groups()
    if groups > 1:
        kname += " in %d shot groups" % groups
    res = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GCell/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl_name, "shots": len(shots), "cell_updates_per_step": cells,
                   "scale": args.scale, "graph": args.graph, "variant": args.variant,
                   "time_block": args.time_block, "variant2": args.variant2, "variant3": args.variant3,
                   "group_bytes": args.group_bytes, "group_streams": args.group_streams, "shot_groups": groups,
                   "autotune": None if tuned is None else {"choice": list(tuned["choice"]),
                                                           "ms_per_fine_step": tuned["table"],
                                                           "cached": tuned["cached"]}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": kname,
                     "mean_launch_ms": round(float(np.mean(step_ms)), 5),
                     "launches_per_step": int(nlaunch[-1]),
                     "bytes_per_cell_update": BYTES_PER_CELL["nl_fwd"],
                     "concurrent_streams": 2 if groups > 1 and args.group_streams == 2 and not args.graph else 1,
                     "note": ("achieved counts the reference algorithm's 80 B per cell-update over the timed "
                              "loop; in shot groups the kernel's data is served mostly by the 256 MB Infinity "
                              "Cache, and with two streams two launches run at once, so mean_launch_ms (loop "
                              "time per launch) is half a kernel's own duration")
                     if groups > 1 else "achieved counts the reference algorithm's 80 B per cell-update"},
    }
    if args.copy_probe:
        nb = 2 << 30
        cms = ctx.copy_probe(nb, 10, kind=1)
        res["roofline"]["copy_probe_GBs"] = round(2 * nb / (cms / 1e3) / 1e9, 1)
        log("copy probe (one float4 per thread, nontemporal stores): %.3f ms for 2 x %d bytes -> %.0f GB/s" %
            (cms, nb, 2 * nb / cms / 1e6))
        cms4 = ctx.copy_probe(nb, 10, kind=2)
        res["roofline"]["copy_probe_word_GBs"] = round(2 * nb / (cms4 / 1e3) / 1e9, 1)
        log("copy probe (one 4-byte word per thread): %.3f ms -> %.0f GB/s" % (cms4, 2 * nb / cms4 / 1e6))
        cms0 = ctx.copy_probe(nb, 10, kind=0)
        res["roofline"]["copy_probe_gridstride_GBs"] = round(2 * nb / (cms0 / 1e3) / 1e9, 1)
        log("copy probe (grid-stride float4): %.3f ms -> %.0f GB/s" % (cms0, 2 * nb / cms0 / 1e6))
    for tf in TRAFFIC_FILES:
        # measured bytes per cell-update of this kernel (rocprofv3 FETCH_SIZE/WRITE_SIZE,
        # copy-kernel calibrated) times the cell-updates of this run's average launch
        if not os.path.exists(tf) or args.workload != "c2":
            continue
        tr = json.load(open(tf))
        v = args.variant2 if args.time_block == 2 else args.variant if args.time_block == 1 else args.variant3
        if (tr["time_block"] == args.time_block and tr.get("variant", tr.get("variant2")) == v
                and (tr.get("shot_groups", 1) > 1) == (groups > 1)
                and tr.get("group_streams", 1) == res["roofline"]["concurrent_streams"]):
            upd_per_launch = cells / float(nlaunch[-1])
            res["roofline"]["traffic"] = round(tr["bytes_per_cell_update"] * upd_per_launch)
            res["roofline"]["traffic_bytes_per_cell_update"] = round(tr["bytes_per_cell_update"], 2)
            res["roofline"]["traffic_source"] = os.path.relpath(tf, ROOT)
            if "traffic_level" in tr:
                res["roofline"]["traffic_level"] = tr["traffic_level"]
    # the secondary legs run in child processes; one that fails is reported in its own field
    # and does not cost the headline line
    legs = []
    if args.c0:
        legs.append(("unbatched_c0", "unbatched C0", c0_leg))
    if args.fwi:
        legs.append(("fwi_gradient", "fwi", run_fwi_leg))
    if args.born:
        legs.append(("born", "born", run_born_leg))
    if args.abc and os.path.exists(ABC_WEIGHTS):
        legs.append(("learned_boundary", "learned-boundary", run_abc_leg))
    for key, name, run in legs:
        t0 = time.perf_counter()
        try:
            res[key] = run(args)
        except Exception as e:  # noqa: BLE001 (reported in the JSON line)
            res[key] = {"error": "%s: %s" % (type(e).__name__, e)}
        log("%s leg %.1f s" % (name, time.perf_counter() - t0))
    if not args.no_cpu:
        t0 = time.perf_counter()
        res["cpu_baseline"] = cpu_baseline(shots_w)
        log("cpu baseline %.1f s" % (time.perf_counter() - t0))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
