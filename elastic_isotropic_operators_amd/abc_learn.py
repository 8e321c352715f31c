"""Learned absorbing boundary: the network, its training data, its training and its test.

A model with a narrow plain sponge of S cells (the reference's cosine taper,
kernelsGpuElastic.cu:1192-1221, with minPad = S) gets a per-edge 3-layer CNN that corrects
the S-cell band of the new time level after every fine step (kernels in csrc/eio_abc.h,
C ABI eio_model_set_abc).  The network is trained to make the narrow model behave like a
wide one:

  data      for seeded layered models, the propagator runs one shot on the model padded by
            a WIDE sponge (120 cells) and snapshots the raw state (u_k, u_k+1) and
            (u_k+1, u_k+2) at pairs of consecutive fine steps (eio_nl_forward_snapshots);
  input     the snapshot (u_k, u_k+1) cropped to the narrow grid (frame outside it zeroed)
            takes ONE plain narrow step (eio_nl_step_states): v = the narrow new level;
            features = (v, u_k+1) on the band plus 6 cells around it;
  target    u_k+2 of the wide run minus v on the band: what the narrow boundary gets wrong
            in one step (its reflection and its damping).

Network (per edge x-min, x-max, z-min, z-max): three "valid" 5x5 convolutions 10 -> 16 -> 16
-> 5, leaky ReLU (slope 0.1) between them, no biases, inputs and outputs in units of a
per-field scale.  Without biases the network is positively homogeneous, so its corrections
scale with the wavefield.  Training: Adam on the relative squared error of the correction,
per-sample normalised.  The FP32 torch network here is the trainer; the GPU kernel runs the
same network with BF16 weights and activations and FP32 accumulation, which
`edge_forward(..., bf16=True)` emulates.

The test of a trained network is the boundary reflection energy: receivers inside the
model record a shot with the wide sponge (the reference), the plain narrow sponge, a plain
intermediate sponge and the learned narrow boundary; the energy of each record's difference
from the reference, relative to the reference's energy, measures what the boundary sends
back.
"""
import time

import numpy as np
import torch
import torch.nn.functional as F

from . import _capi
from . import geometry as G
from . import seis_utils as su
from . import workloads as W

FAT = 4
LEAK = 0.1
MARGIN = 6                       # receptive-field half width of three 5x5 layers
CIN = (10, 16, 16)
COUT = (16, 16, 5)
EDGES = ("xmin", "xmax", "zmin", "zmax")


# ------------------------------------------------------------------ geometry
def band_regions(nz, nx, S):
    """Output rectangles (x0, x1, z0, z1) of the four edge bands: the x-min and x-max bands
    own every update row, the z-min and z-max bands the columns between them."""
    return [(FAT, FAT + S, FAT, nz - FAT), (nx - FAT - S, nx - FAT, FAT, nz - FAT),
            (FAT + S, nx - FAT - S, FAT, FAT + S), (FAT + S, nx - FAT - S, nz - FAT - S, nz - FAT)]


def features(newer, older, scale, regions):
    """Per edge, the network input (10, X + 12, Z + 12): the two levels divided by the field
    scale (multiplied by its FP32 inverse, as the kernel does), zero outside the array."""
    inv = (1.0 / scale.float()).to(newer.dtype).view(5, 1, 1)
    u = torch.cat([newer * inv, older * inv])
    up = F.pad(u, (MARGIN, MARGIN, MARGIN, MARGIN))
    return [up[:, x0:x1 + 2 * MARGIN, z0:z1 + 2 * MARGIN] for (x0, x1, z0, z1) in regions]


def band(u, regions):
    return [u[:, x0:x1, z0:z1] for (x0, x1, z0, z1) in regions]


# ------------------------------------------------------------------ network
def _bf16(x):
    return x.to(torch.bfloat16).to(torch.float32)


class AbcNet:
    """Weights W[e][l], torch Conv2d layout [out][in][5][5] (first kernel axis along x)."""

    def __init__(self, device="cuda", seed=0, weights=None):
        g = torch.Generator().manual_seed(seed)
        self.W = []
        for e in range(4):
            layers = []
            for l in range(3):
                if weights is not None:
                    w = torch.as_tensor(weights[e][l], dtype=torch.float32)
                else:
                    std = (2.0 / (CIN[l] * 25)) ** 0.5 * (0.1 if l == 2 else 1.0)
                    w = torch.randn(COUT[l], CIN[l], 5, 5, generator=g) * std
                layers.append(w.to(device).requires_grad_(True))
            self.W.append(layers)

    def parameters(self):
        return [w for layers in self.W for w in layers]

    def edge_forward(self, e, x, bf16=False):
        """x (B, 10, X + 12, Z + 12) -> correction (B, 5, X, Z) in scale units."""
        h = _bf16(x) if bf16 else x
        for l in range(3):
            w = _bf16(self.W[e][l]) if bf16 else self.W[e][l]
            h = F.conv2d(h, w)
            if l < 2:
                h = F.leaky_relu(h, LEAK)
                if bf16:
                    h = _bf16(h)
        return h

    def flat(self):
        """The C ABI weight array (eio_model_set_abc): edges, layers, [out][in][5][5]."""
        return np.concatenate([self.W[e][l].detach().float().cpu().numpy().reshape(-1)
                               for e in range(4) for l in range(3)]).astype(np.float32)

    def state(self):
        return {"%s_%d" % (EDGES[e], l): self.W[e][l].detach().cpu().numpy() for e in range(4) for l in range(3)}

    @classmethod
    def from_state(cls, st, device="cuda"):
        return cls(device, weights=[[st["%s_%d" % (EDGES[e], l)] for l in range(3)] for e in range(4)])


def save(path, net, scale, strip, info=None):
    st = net.state()
    st["scale"] = np.asarray(scale, np.float32)
    st["strip"] = np.int32(strip)
    for k, v in (info or {}).items():
        st["info_" + k] = np.asarray(v)
    np.savez(path, **st)


def load(path, device="cuda"):
    z = np.load(path, allow_pickle=False)
    st = {k: z[k] for k in z.files}
    return AbcNet.from_state(st, device), st["scale"].astype(np.float32), int(st["strip"])


# ------------------------------------------------------------------ propagation helpers
def padded(vvr, pad, dz=10.0, dx=10.0):
    """Interior (3, nx, nz) [vp, vs, rho] padded by `pad` sponge cells and the FAT frame by
    edge replication; interior origin at (0, 0).  Returns a dict with the grid and (rho,
    lambda, mu)."""
    p = pad + FAT
    v = np.pad(np.asarray(vvr, np.float64), ((0, 0), (p, p), (p, p)), mode="edge")
    return dict(nz=v.shape[2], nx=v.shape[1], dz=float(dz), dx=float(dx), oz=-p * dz, ox=-p * dx,
                elastic=su.vpvsrho_to_rholamemu(v, 1), minPad=int(pad), pad=int(pad))


class Shot:
    """One explosive shot of a padded model on a context: media, regular source traces and
    receiver buffers on the device (torch tensors)."""

    def __init__(self, ctx, w, nts, dts, sub, src_z, src_x, rec_z, rec_x, fdom=8.0, delay=0.15, alpha=0.99):
        self.ctx, self.w, self.nts = ctx, w, nts
        dtw = G.dtw_of(dts, sub)
        self.media = G.scaled_media(w["elastic"], dtw, np.float32)
        self.mid = ctx.model_add(w["nz"], w["nx"], nts, 2 * sub, w["dz"], w["dx"], dts, w["minPad"], self.media, alpha)
        si = G.staggered_interps(w["nz"], w["oz"], w["dz"], w["ox"], w["dx"], [src_z], [src_x])
        self.ri = G.staggered_interps(w["nz"], w["oz"], w["dz"], w["ox"], w["dx"], rec_z, rec_x)
        model = W.explosive_model(nts, dts, fdom, delay)
        src = G.nl_source_traces(si, self.media, model, dtw, w["dx"], w["dz"], np.float32)
        self.src = [torch.tensor(s, device="cuda") for s in src]
        self.rec = [torch.zeros((self.ri[G.GRID_OF_COMP[k]].uniq.size, nts), device="cuda") for k in range(5)]
        self.spos = [ip.uniq for ip in si]
        self.rpos = [ip.uniq for ip in self.ri]
        self.ntw = (nts - 1) * 2 * sub + 1
        self.cells = (w["nz"] - 2 * FAT) * (w["nx"] - 2 * FAT) * (self.ntw - 1)

    def shot(self):
        s, keep = self.ctx.make_shot(self.mid, self.spos, self.rpos, [t.data_ptr() for t in self.src],
                                     [t.data_ptr() for t in self.rec])
        self._keep = keep
        return s

    def forward(self):
        """Receiver data (5, nRec, nts) on the host and the device time of the propagation."""
        torch.cuda.synchronize()
        self.ctx.nl_forward([self.shot()], device_ptrs=True)
        total_ms = self.ctx.last_timing()[0]
        data = np.stack([self.ri[G.GRID_OF_COMP[k]].forward(self.rec[k].cpu().numpy().astype(np.float64))
                         for k in range(5)])
        return data, total_ms

    def snapshots(self, steps):
        """Raw (older, newer) states after the given fine steps: (n, 2, 5, nx, nz) on the device."""
        out = torch.empty((len(steps), 2, 5, self.w["nx"], self.w["nz"]), device="cuda")
        torch.cuda.synchronize()
        self.ctx.nl_forward_snapshots([self.shot()], 0, steps, out.data_ptr())
        return out


def crop_states(states, off, nx, nz):
    """Wide-grid states (n, 2, 5, NX, NZ) -> narrow grid (n, 2, 5, nx, nz) at offset `off`
    cells, with the narrow FAT frame zeroed (a propagation never writes it)."""
    s = states[..., off:off + nx, off:off + nz].contiguous()
    s[..., :FAT, :] = 0
    s[..., nx - FAT:, :] = 0
    s[..., :, :FAT] = 0
    s[..., :, nz - FAT:] = 0
    return s


# ------------------------------------------------------------------ training data
def generate(n_models=200, seed=4, nz_i=100, nx_i=200, strip=10, wide=120, nts=1500, dts=0.001, sub=2,
             pairs_per_model=12, first_step=600, device=0, log=print):
    """Training samples from wide-sponge runs of seeded layered models (the ranges of the
    batched-shot workload).  Returns per edge (X, Y): X (N, 10, ., .) features and Y (N, 5, ., .)
    corrections in field units, plus the raw per-field band RMS of the features."""
    ctx = _capi.Context(device, np.float32)
    rng = np.random.default_rng(seed)
    Xs = [[] for _ in range(4)]
    Ys = [[] for _ in range(4)]
    t0 = time.time()
    for m in range(n_models):
        vvr = W.layered_vvr(rng, nz_i, nx_i)
        ww = padded(vvr, wide)
        wn = padded(vvr, strip)
        sz = float(rng.uniform(20, nz_i - 20)) * 10.0
        sx = float(rng.uniform(20, nx_i - 20)) * 10.0
        wshot = Shot(ctx, ww, nts, dts, sub, sz, sx, [50.0], [50.0])
        nshot = Shot(ctx, wn, nts, dts, sub, sz, sx, [50.0], [50.0])
        ks = np.sort(rng.choice(np.arange(first_step, wshot.ntw - 4, 2), size=pairs_per_model, replace=False))
        steps = np.stack([ks, ks + 1], axis=1).reshape(-1)
        snaps = wshot.snapshots(steps)
        off = wide - strip
        cur = crop_states(snaps[0::2], off, wn["nx"], wn["nz"])      # (u_k, u_k+1)
        nxt = crop_states(snaps[1::2], off, wn["nx"], wn["nz"])      # (u_k+1, u_k+2)
        del snaps
        torch.cuda.synchronize()
        ctx.step_states(nshot.mid, cur.shape[0], cur.data_ptr())     # cur -> (v, u_k+1)
        torch.cuda.synchronize()
        reg = band_regions(wn["nz"], wn["nx"], strip)
        one = torch.ones(5, device="cuda")
        for i in range(cur.shape[0]):
            feats = features(cur[i, 0], cur[i, 1], one, reg)
            tgt = band(nxt[i, 1] - cur[i, 0], reg)
            for e in range(4):
                Xs[e].append(feats[e])
                Ys[e].append(tgt[e])
        ctx.model_remove(wshot.mid)
        ctx.model_remove(nshot.mid)
        if log and (m + 1) % 20 == 0:
            log("abc data: %d/%d models, %.1f s" % (m + 1, n_models, time.time() - t0))
    data = [(torch.stack(Xs[e]), torch.stack(Ys[e])) for e in range(4)]
    # per-field scale: RMS over all bands of the new level
    sq = sum((d[0][:, :5] ** 2).sum(dim=(0, 2, 3)) for d in data)
    cnt = sum(d[0][:, 0].numel() for d in data)
    scale = torch.sqrt(sq / cnt).clamp_min(1e-30)
    ctx.close()
    return data, scale


# ------------------------------------------------------------------ training
def train(data, scale, iters=4000, lr=2e-3, batch=32, seed=0, val_frac=0.1, log=print):
    """Adam on the per-sample relative squared error of the corrections; returns the network
    and the held-out relative L2 error per edge."""
    torch.manual_seed(seed)
    inv = (1.0 / scale).view(1, 5, 1, 1)
    sets = []
    for X, Y in data:
        Xn = X.clone()
        Xn[:, :5] *= inv
        Xn[:, 5:] *= inv
        Yn = Y * inv
        a = torch.sqrt((Xn ** 2).mean(dim=(1, 2, 3), keepdim=True))
        a = torch.maximum(a, 1e-3 * a.max())
        Xn, Yn = Xn / a, Yn / a
        n = Xn.shape[0]
        nv = max(1, int(val_frac * n))
        sets.append((Xn[nv:], Yn[nv:], Xn[:nv], Yn[:nv]))
    net = AbcNet("cuda", seed)
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, iters)
    g = torch.Generator(device="cuda").manual_seed(seed)
    t0 = time.time()
    for it in range(iters):
        opt.zero_grad()
        loss = 0.0
        for e in range(4):
            Xt, Yt = sets[e][0], sets[e][1]
            idx = torch.randint(0, Xt.shape[0], (batch,), device="cuda", generator=g)
            pred = net.edge_forward(e, Xt[idx])
            y = Yt[idx]
            loss = loss + ((pred - y) ** 2).sum() / (y ** 2).sum().clamp_min(1e-30)
        loss.backward()
        opt.step()
        sched.step()
        if log and (it % 500 == 0 or it == iters - 1):
            log("abc train: iter %d loss %.4f (%.1f s)" % (it, float(loss) / 4, time.time() - t0))
    val = []
    with torch.no_grad():
        for e in range(4):
            Xv, Yv = sets[e][2], sets[e][3]
            err = sum(float(((net.edge_forward(e, Xv[i:i + 64]) - Yv[i:i + 64]) ** 2).sum())
                      for i in range(0, Xv.shape[0], 64))
            val.append((err / float((Yv ** 2).sum())) ** 0.5)
    return net, val


# ------------------------------------------------------------------ reflection test
def reflection_test(vvr, net=None, scale=None, strip=10, mid=60, wide=120, nts=1500, dts=0.001, sub=2,
                    src=None, rec_every=2, inset=3, device=0, time_block=2, log=print):
    """Shot records of the same interior model under the wide sponge (reference), the plain
    `strip` sponge, the plain `mid` sponge and (with a network) the learned `strip` boundary.
    Receivers: every `rec_every` cells along the four interior edges, `inset` cells in.
    Returns {name: dict(refl=relative energy of the difference from the reference,
    ms=device time, gcells=GCell/s of the propagated grid, interior_gcells=interior cells per
    second)}."""
    ctx = _capi.Context(device, np.float32)
    nx_i, nz_i = vvr.shape[1], vvr.shape[2]
    d = 10.0
    zs = np.arange(inset, nz_i - inset, rec_every) * d
    xs = np.arange(inset, nx_i - inset, rec_every) * d
    rz = np.concatenate([np.full(xs.size, inset * d), np.full(xs.size, (nz_i - 1 - inset) * d), zs, zs])
    rx = np.concatenate([xs, xs, np.full(zs.size, inset * d), np.full(zs.size, (nx_i - 1 - inset) * d)])
    if src is None:
        src = (0.3 * nz_i * d, 0.5 * nx_i * d)
    ntw = (nts - 1) * 2 * sub + 1
    interior = nz_i * nx_i * (ntw - 1)
    out = {}
    runs = [("wide%d" % wide, wide, False), ("plain%d" % strip, strip, False), ("plain%d" % mid, mid, False)]
    if net is not None:
        runs.append(("learned%d" % strip, strip, True))
    ref = None
    for name, pad, learned in runs:
        w = padded(vvr, pad)
        shot = Shot(ctx, w, nts, dts, sub, src[0], src[1], rz, rx)
        if learned:
            ctx.set_abc(shot.mid, strip, net.flat(), scale)
        ctx.set_time_block(1 if learned else time_block, 0 if learned else 6)
        shot.forward()  # warm-up
        data, ms = shot.forward()
        if ref is None:
            ref = data
        refl = float(((data - ref) ** 2).sum() / (ref ** 2).sum())
        out[name] = dict(refl=refl, ms=ms, gcells=shot.cells / ms * 1e-6, interior_gcells=interior / ms * 1e-6,
                         grid=[w["nz"], w["nx"]])
        if log:
            log("abc %s: grid %dx%d, %.1f ms, %.2f GCell/s, reflection energy %.3e" %
                (name, w["nz"], w["nx"], ms, out[name]["gcells"], refl))
        ctx.model_remove(shot.mid)
    ctx.close()
    return out


def main(argv=None):
    """Generate, train, save and test: python -m elastic_isotropic_operators_amd.abc_learn
    --out weights.npz [--models 200 --iters 4000 --test-nz 1000 --test-nx 3000 --test-nts 4000]"""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--models", type=int, default=200)
    ap.add_argument("--pairs", type=int, default=12)
    ap.add_argument("--iters", type=int, default=4000)
    ap.add_argument("--strip", type=int, default=10)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--test-nz", type=int, default=200)
    ap.add_argument("--test-nx", type=int, default=400)
    ap.add_argument("--test-nts", type=int, default=1500)
    args = ap.parse_args(argv)

    def log(m):
        print("[abc] " + m, flush=True)

    data, scale = generate(n_models=args.models, seed=args.seed, strip=args.strip, pairs_per_model=args.pairs,
                           log=log)
    log("scale " + " ".join("%.3e" % v for v in scale.tolist()))
    net, val = train(data, scale, iters=args.iters, log=log)
    log("held-out relative error per edge " + " ".join("%.3f" % v for v in val))
    del data
    save(args.out, net, scale.cpu().numpy(), args.strip, info={"val": val, "models": args.models,
                                                               "iters": args.iters, "seed": args.seed})
    vvr = W.layered_vvr(np.random.default_rng(1234), args.test_nz, args.test_nx)
    res = reflection_test(vvr, net, scale.cpu().numpy(), strip=args.strip, nts=args.test_nts, log=log)
    print(json.dumps({"val": val, "test": res}), flush=True)


if __name__ == "__main__":
    main()
