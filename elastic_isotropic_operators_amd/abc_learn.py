"""Learned absorbing boundary: the network, its training data, its training and its test.

Without help, the stencil of the outermost update cells reads zeros in the 4-cell FAT frame
and the grid edge reflects; the reference hides that behind a cosine sponge
(kernelsGpuElastic.cu:1192-1221) that must be wide to work.  Here a model keeps a narrow
sponge of S cells and gets, per edge, a 3-layer CNN that writes the new level's FAT frame
after every fine step (kernels in csrc/eio_abc.h, C ABI eio_model_set_abc): the values a
much wider grid would hold there, predicted from the first update cells of the new and the
previous level, so the stencil continues the outgoing wavefield instead of reflecting it.

  data      for seeded layered models, the propagator runs one shot on the model padded by
            a WIDE sponge (120 cells) and snapshots the raw state (u_k, u_k+1) at random fine
            steps (eio_nl_forward_snapshots);
  input     the snapshot cropped to the narrow grid (interior + S + frame): the new level and
            the difference new - previous (formed before rounding: one fine step changes a
            field by about a percent), on the update cells only (the frame and the outside
            read as zero), each per-field scaled;
  target    u_k+1 of the wide run on the narrow grid's frame cells.

Network (per edge x-min, x-max, z-min, z-max): three "valid" 5x5 convolutions 10 -> 16 -> 16
-> 5, leaky ReLU (slope 0.1) between them, no biases, inputs and outputs in units of a
per-field scale.  Without biases the network is positively homogeneous, so its output scales
with the wavefield.  Training: Adam on the per-sample relative squared error, each frame row
weighted by the square of the stencil coefficient that reads it from the first update cell
(the outermost frame rows barely matter to the stencil and are the hardest to predict).  The
network's forward and backward during training run on the library's training kernels
(eio_abc_conv_forward / eio_abc_conv_backward: BF16 MFMA, FP32 accumulation, the inference
kernel's rounding points; torch supplies the loss, the leaky-ReLU derivative and Adam);
`edge_forward(..., bf16=True)` is the torch restatement the kernels are checked against.

The test of a trained network is the boundary reflection energy: receivers inside the model
record a shot with the wide sponge (the reference), the plain narrow sponge, a plain
intermediate sponge and the learned narrow boundary; the energy of each record's difference
from the reference, relative to the reference's energy, measures what the boundary sends
back.
"""
import ctypes
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
def frame_regions(nz, nx):
    """Output rectangles (x0, x1, z0, z1) of the four edge frames: x-min and x-max own the
    update rows, z-min and z-max the update columns (the frame corners are never read)."""
    return [(0, FAT, FAT, nz - FAT), (nx - FAT, nx, FAT, nz - FAT),
            (FAT, nx - FAT, 0, FAT), (FAT, nx - FAT, nz - FAT, nz)]


def features(newer, older, scale, regions, diff=True):
    """Per edge, the network input (10, X + 12, Z + 12): the new level and the difference
    new - previous (diff) or the previous level on the update cells (zero on the frame and
    outside the array), multiplied by the FP32 inverses of scale[0..4] and scale[5..9] as the
    kernel does."""
    nx, nz = newer.shape[1], newer.shape[2]
    inv = (1.0 / scale.float()).to(newer.dtype).view(10, 1, 1)
    u = torch.zeros((10, nx, nz), dtype=newer.dtype, device=newer.device)
    a = newer[:, FAT:nx - FAT, FAT:nz - FAT]
    b = older[:, FAT:nx - FAT, FAT:nz - FAT]
    u[:5, FAT:nx - FAT, FAT:nz - FAT] = a * inv[:5]
    u[5:, FAT:nx - FAT, FAT:nz - FAT] = ((a - b) if diff else b) * inv[5:]
    up = F.pad(u, (MARGIN, MARGIN, MARGIN, MARGIN))
    return [up[:, x0:x1 + 2 * MARGIN, z0:z1 + 2 * MARGIN] for (x0, x1, z0, z1) in regions]


def region(u, regions):
    return [u[:, x0:x1, z0:z1] for (x0, x1, z0, z1) in regions]


# stencil coefficients (nonlinearPropElasticGpuFunctions.cu:159-167, times the cell size)
# that read frame rows 1, 2, 3, 4 cells out from the first update cell
STENCIL = (1.196289062541883, 0.079752604188901, 0.009570312506634, 6.975446437140719e-4)


def row_weights(edge, device="cuda", floor=0.05, depth=4):
    """Loss weight of each frame row of an edge (shape broadcastable to (B, 5, X, Z)): the
    squared stencil coefficient, relative to the first, but at least `floor` (a row the loss
    ignores can take values large enough to destabilise the propagation)."""
    w = torch.tensor([max(c * c / STENCIL[0] ** 2, floor) if r < depth else 0.0 for r, c in enumerate(STENCIL)],
                     device=device)   # distance 1..4; rows beyond `depth` are not written
    if edge == 0:      # x-min frame: x = 3 is next to the update region
        return w.flip(0).view(1, 1, 4, 1)
    if edge == 1:
        return w.view(1, 1, 4, 1)
    if edge == 2:
        return w.flip(0).view(1, 1, 1, 4)
    return w.view(1, 1, 1, 4)


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


# ------------------------------------------------------------------ training kernels (HIP)
# One 5x5 layer of the network on the library's MFMA kernels (eio_abc_conv_forward /
# eio_abc_conv_backward): NHWC, 16 channels, BF16 activations and weights at the inference
# kernel's rounding points, FP32 accumulation; the backward gives the input gradient and the
# weight gradient (summed in a fixed order).  Straight-through rounding, as torch's casts.
_TRAIN = {}


def _train_ctx():
    if "ctx" not in _TRAIN:
        L = _capi.lib()
        vp, ci = ctypes.c_void_p, ctypes.c_int
        L.eio_abc_conv_forward.argtypes = [vp, vp, ci, ci, ci, vp, ci, vp, vp]
        L.eio_abc_conv_forward.restype = ci
        L.eio_abc_conv_backward.argtypes = [vp, vp, vp, ci, ci, ci, vp, vp, vp]
        L.eio_abc_conv_backward.restype = ci
        _TRAIN["lib"] = L
        _TRAIN["ctx"] = _capi.Context(torch.cuda.current_device(), np.float32)
    return _TRAIN["lib"], _TRAIN["ctx"]


class Conv5Layer(torch.autograd.Function):
    """x (B, H, W, 16) FP32 holding BF16 values (W >= 20), w (16, 16, 5, 5) -> BF16(leaky(z))
    (leaky) or z (the last layer) with z the valid 5x5 convolution, (B, H-4, W-4, 16)."""

    @staticmethod
    def forward(fctx, x, w, leaky):
        L, ctx = _train_ctx()
        a = x.detach().to(torch.bfloat16).contiguous()
        wc = w.detach().float().contiguous()
        B, H, Wd, C = a.shape
        assert C == 16 and wc.shape == (16, 16, 5, 5)
        z = torch.empty((B, H - 4, Wd - 4, 16), dtype=torch.float32, device=x.device)
        h = torch.empty((B, H - 4, Wd - 4, 16), dtype=torch.bfloat16, device=x.device)
        torch.cuda.synchronize()
        _capi.check(L.eio_abc_conv_forward(ctx.h, a.data_ptr(), B, H, Wd, wc.data_ptr(), int(bool(leaky)),
                                           z.data_ptr(), h.data_ptr()))
        fctx.save_for_backward(a, wc, z)
        fctx.leaky = bool(leaky)
        return h.float() if leaky else z.clone()

    @staticmethod
    def backward(fctx, dh):
        L, ctx = _train_ctx()
        a, wc, z = fctx.saved_tensors
        dz = dh.float()
        if fctx.leaky:
            dz = torch.where(z > 0, dz, LEAK * dz)
        dz = dz.contiguous()
        B, H, Wd, _ = a.shape
        da = torch.empty(a.shape, dtype=torch.float32, device=dz.device) if fctx.needs_input_grad[0] else None
        dw = torch.empty((16, 16, 5, 5), dtype=torch.float32, device=dz.device)
        torch.cuda.synchronize()
        _capi.check(L.eio_abc_conv_backward(ctx.h, a.data_ptr(), dz.data_ptr(), B, H, Wd, wc.data_ptr(),
                                            None if da is None else da.data_ptr(), dw.data_ptr()))
        return da, dw, None


def edge_forward_hip(net, e, x):
    """net.edge_forward(e, x, bf16=True) on the training kernels: x (B, 10, X + 12, Z + 12) ->
    (B, 5, X, Z), differentiable in x and in the network's weights.  The long axis of a frame
    is the kernels' W axis (z frames are transposed, and their taps with them)."""
    h = F.pad(x, (0, 0, 0, 0, 0, 16 - x.shape[1])).permute(0, 2, 3, 1)   # (B, X+12, Z+12, 16)
    tr = e >= 2
    if tr:
        h = h.transpose(1, 2)
    for l in range(3):
        w = net.W[e][l]
        w = F.pad(w, (0, 0, 0, 0, 0, 16 - w.shape[1], 0, 16 - w.shape[0]))
        if tr:
            w = w.transpose(2, 3)
        h = Conv5Layer.apply(h.contiguous(), w.contiguous(), l < 2)
    if tr:
        h = h.transpose(1, 2)
    return h[..., :5].permute(0, 3, 1, 2)


def save(path, net, scale, strip, on=1, info=None):
    """on: the eio_model_set_abc input mode (1: new level and difference, 2: both levels)."""
    st = net.state()
    st["scale"] = np.asarray(scale, np.float32)
    st["strip"] = np.int32(strip)
    st["on"] = np.int32(on)
    for k, v in (info or {}).items():
        st["info_" + k] = np.asarray(v)
    np.savez(path, **st)


def load(path, device="cuda"):
    """(network, scale[10], strip, on)."""
    z = np.load(path, allow_pickle=False)
    st = {k: z[k] for k in z.files}
    return AbcNet.from_state(st, device), st["scale"].astype(np.float32), int(st["strip"]), int(st.get("on", 1))


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


def crop(states, off, nx, nz):
    """Wide-grid states (..., NX, NZ) -> the narrow grid (..., nx, nz) at offset `off` cells."""
    return states[..., off:off + nx, off:off + nz]


# ------------------------------------------------------------------ training data
def generate(n_models=200, seed=4, nz_i=100, nx_i=200, strip=10, wide=120, nts=1500, dts=0.001, sub=2,
             snaps_per_model=24, first_step=600, device=0, log=print):
    """Training samples from wide-sponge runs of seeded layered models (the ranges of the
    batched-shot workload): per edge (X, Y) with X (N, 10, ., .) the new and the previous level
    on the update cells and Y (N, 5, ., .) the frame, in field units."""
    ctx = _capi.Context(device, np.float32)
    rng = np.random.default_rng(seed)
    Xs = [[] for _ in range(4)]
    Ys = [[] for _ in range(4)]
    t0 = time.time()
    one = torch.ones(10, device="cuda")
    for m in range(n_models):
        vvr = W.layered_vvr(rng, nz_i, nx_i)
        ww = padded(vvr, wide)
        wn = padded(vvr, strip)
        sz = float(rng.uniform(20, nz_i - 20)) * 10.0
        sx = float(rng.uniform(20, nx_i - 20)) * 10.0
        wshot = Shot(ctx, ww, nts, dts, sub, sz, sx, [50.0], [50.0])
        steps = np.sort(rng.choice(np.arange(first_step, wshot.ntw - 2), size=snaps_per_model, replace=False))
        snaps = crop(wshot.snapshots(steps), wide - strip, wn["nx"], wn["nz"])
        reg = frame_regions(wn["nz"], wn["nx"])
        for i in range(snaps.shape[0]):
            feats = features(snaps[i, 1], snaps[i, 0], one, reg, diff=False)
            tgt = region(snaps[i, 1], reg)
            for e in range(4):
                Xs[e].append(feats[e].clone())
                Ys[e].append(tgt[e].clone())
        del snaps
        ctx.model_remove(wshot.mid)
        if log and (m + 1) % 20 == 0:
            log("abc data: %d/%d models, %.1f s" % (m + 1, n_models, time.time() - t0))
    data = [(torch.stack(Xs[e]), torch.stack(Ys[e])) for e in range(4)]
    ctx.close()
    return data


def inputs(X, diff):
    """Raw (new, previous) samples -> the network's channels (new, new - previous or previous)."""
    return torch.cat([X[:, :5], X[:, :5] - X[:, 5:]], dim=1) if diff else X


def input_scale(data, diff):
    """Per-channel RMS over the update cells of all samples: the 10 input scales."""
    cnt = sum(float((d[0][:, 0] != 0).sum()) for d in data)
    sq = sum((inputs(d[0], diff) ** 2).sum(dim=(0, 2, 3)) for d in data)
    return torch.sqrt(sq / max(cnt, 1.0)).clamp_min(1e-30)


# ------------------------------------------------------------------ differentiable narrow step
FD_C = (1.196289062541883, -0.079752604188901, 0.009570312506634, -6.975446437140719e-4)


class TorchStep:
    """One fine step of the NL forward without sources, in torch (differentiable), on states
    (B, 5, nx, nz): new = d (d old + A cur) on the update cells (the fused form of the
    reference's step, kernelsGpuElastic.cu:298-376, and sponge, :1192-1221), zero elsewhere.
    Used to train the network through rollouts; checked against the GPU step."""

    def __init__(self, media, nz, nx, dz, dx, minPad, alpha=0.99, device="cuda"):
        F_ = FAT
        self.nz, self.nx = nz, nx
        self.med = torch.as_tensor(np.asarray(media, np.float32), device=device)[:, F_:nx - F_, F_:nz - F_]
        self.cx = [c / dx for c in FD_C]
        self.cz = [c / dz for c in FD_C]
        iz = np.arange(nz)[None, :]
        ix = np.arange(nx)[:, None]
        j = np.minimum(np.minimum(iz - F_, ix - F_), np.minimum(nz - iz - 1 - F_, nx - ix - 1 - F_))
        tab = np.array([alpha + (1.0 - alpha) * np.cos(np.pi / minPad * (minPad - jj)) for jj in range(max(minPad, 1))],
                       np.float32) if minPad > 0 else np.ones(1, np.float32)
        d = np.where(j < minPad, tab[np.clip(j, 0, max(minPad, 1) - 1)], np.float32(1.0)).astype(np.float32)
        self.d = torch.as_tensor(d[F_:nx - F_, F_:nz - F_], device=device)

    def _dx(self, f, plus):
        nx, F_ = self.nx, FAT
        out = 0.0
        for k in range(4):
            if plus:   # D+ f[i] = sum c_k (f[i+1+k] - f[i-k])
                out = out + self.cx[k] * (f[..., F_ + 1 + k:nx - F_ + 1 + k, F_:self.nz - F_]
                                          - f[..., F_ - k:nx - F_ - k, F_:self.nz - F_])
            else:      # D- f[i] = sum c_k (f[i+k] - f[i-1-k])
                out = out + self.cx[k] * (f[..., F_ + k:nx - F_ + k, F_:self.nz - F_]
                                          - f[..., F_ - 1 - k:nx - F_ - 1 - k, F_:self.nz - F_])
        return out

    def _dz(self, f, plus):
        nz, F_ = self.nz, FAT
        out = 0.0
        for k in range(4):
            if plus:
                out = out + self.cz[k] * (f[..., F_:self.nx - F_, F_ + 1 + k:nz - F_ + 1 + k]
                                          - f[..., F_:self.nx - F_, F_ - k:nz - F_ - k])
            else:
                out = out + self.cz[k] * (f[..., F_:self.nx - F_, F_ + k:nz - F_ + k]
                                          - f[..., F_:self.nx - F_, F_ - 1 - k:nz - F_ - 1 - k])
        return out

    def __call__(self, old, cur):
        F_ = FAT
        m = self.med
        vx, vz, sxx, szz, sxz = (cur[:, i] for i in range(5))
        dvx_p = self._dx(vx, True)
        dvz_p = self._dz(vz, True)
        a = torch.stack([m[0] * (self._dx(sxx, False) + self._dz(sxz, True)),
                         m[1] * (self._dz(szz, False) + self._dx(sxz, True)),
                         m[2] * dvx_p + m[3] * dvz_p,
                         m[3] * dvx_p + m[2] * dvz_p,
                         m[4] * (self._dz(vx, False) + self._dx(vz, False))], dim=1)
        d = self.d
        new = d * (d * old[..., F_:self.nx - F_, F_:self.nz - F_] + a)
        return F.pad(new, (F_, F_, F_, F_))


def place_frames(outs, shape, regions):
    """Frame outputs per edge (B, 5, X, Z) placed into a zero tensor of `shape`."""
    fr = torch.zeros(shape, dtype=outs[0].dtype, device=outs[0].device)
    for (x0, x1, z0, z1), o in zip(regions, outs):
        fr[..., x0:x1, z0:z1] = o
    return fr


def rollout(net, step, older, newer, scale, steps, diff=True):
    """`steps` fine steps of the narrow model with the network writing the frames: returns the
    new levels (steps, B, 5, nx, nz).  older/newer (B, 5, nx, nz); newer carries its frame."""
    nx, nz = newer.shape[-2], newer.shape[-1]
    regs = frame_regions(nz, nx)
    inv = (1.0 / scale).view(1, 10, 1, 1)
    outs = []
    for _ in range(steps):
        new = step(older, newer)
        # network input as the kernel forms it: update cells of (new, new - newer), scaled
        prev = F.pad(newer[..., FAT:nx - FAT, FAT:nz - FAT], (FAT,) * 4)
        u = torch.cat([new, (new - prev) if diff else prev], dim=1) * inv
        up = F.pad(u, (MARGIN,) * 4)
        fr = [net.edge_forward(e, up[..., x0:x1 + 2 * MARGIN, z0:z1 + 2 * MARGIN]) * scale[:5].view(1, 5, 1, 1)
              for e, (x0, x1, z0, z1) in enumerate(regs)]
        new = new + place_frames(fr, new.shape, regs)
        outs.append(new)
        older, newer = newer, new
    return torch.stack(outs)


def generate_sequences(n_models=100, seed=5, nz_i=100, nx_i=200, strip=10, wide=120, nts=1500, dts=0.001, sub=2,
                       seq_per_model=2, length=33, first_step=700, device=0, log=print):
    """Consecutive wide-run states on the narrow grid: per sequence (length, 2, 5, nx, nz),
    plus each model's scaled media for the narrow step."""
    ctx = _capi.Context(device, np.float32)
    rng = np.random.default_rng(seed)
    seqs, meds = [], []
    t0 = time.time()
    for m in range(n_models):
        vvr = W.layered_vvr(rng, nz_i, nx_i)
        ww = padded(vvr, wide)
        wn = padded(vvr, strip)
        sz = float(rng.uniform(20, nz_i - 20)) * 10.0
        sx = float(rng.uniform(20, nx_i - 20)) * 10.0
        wshot = Shot(ctx, ww, nts, dts, sub, sz, sx, [50.0], [50.0])
        starts = np.sort(rng.choice(np.arange(first_step, wshot.ntw - length - 2, length + 1), size=seq_per_model,
                                    replace=False))
        steps = np.concatenate([np.arange(s0, s0 + length) for s0 in starts])
        snaps = crop(wshot.snapshots(steps), wide - strip, wn["nx"], wn["nz"]).contiguous()
        for q in range(seq_per_model):
            seqs.append(snaps[q * length:(q + 1) * length].clone())
            meds.append(G.scaled_media(wn["elastic"], G.dtw_of(dts, sub), np.float32))
        del snaps
        ctx.model_remove(wshot.mid)
        if log and (m + 1) % 20 == 0:
            log("abc sequences: %d/%d models, %.1f s" % (m + 1, n_models, time.time() - t0))
    ctx.close()
    return seqs, meds, wn


def train_rollout(net, seqs, meds, wn, scale, iters=1500, lr=2e-4, horizons=(4, 8, 16, 32), batch=4, seed=1,
                  diff=True, log=print):
    """Fine-tune the network through rollouts of the narrow model: from a wide-run state, roll
    out H steps with the network writing the frames and penalise the relative squared error of
    every new level against the wide run (update cells and frame), per-field scaled.  H grows
    over the run (curriculum)."""
    torch.manual_seed(seed)
    g = torch.Generator().manual_seed(seed)
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    steps = [TorchStep(md, wn["nz"], wn["nx"], wn["dz"], wn["dx"], wn["minPad"]) for md in meds]
    inv5 = (1.0 / scale[:5]).view(1, 1, 5, 1, 1)
    t0 = time.time()
    skipped = 0
    L = seqs[0].shape[0]
    for it in range(iters):
        H = horizons[min(len(horizons) - 1, it * len(horizons) // iters)]
        opt.zero_grad()
        loss = 0.0
        for _ in range(batch):
            i = int(torch.randint(0, len(seqs), (1,), generator=g))
            s0 = int(torch.randint(0, L - H, (1,), generator=g))
            sq = seqs[i]
            older, newer = sq[s0, 0][None], sq[s0, 1][None]
            pred = rollout(net, steps[i], older, newer, scale, H, diff)    # (H, 1, 5, nx, nz)
            tgt = sq[s0 + 1:s0 + 1 + H, 1][:, None]
            err = ((pred - tgt) * inv5) ** 2
            ref = (tgt * inv5) ** 2
            loss = loss + err.sum() / ref.sum().clamp_min(1e-30)
        if not bool(torch.isfinite(loss)):  # a diverged rollout: skip, keep the weights
            skipped += 1
            continue
        (loss / batch).backward()
        torch.nn.utils.clip_grad_norm_(net.parameters(), 1.0)
        opt.step()
        if log and (it % 100 == 0 or it == iters - 1):
            log("abc rollout: iter %d horizon %d loss %.3e skipped %d (%.1f s)" % (
                it, H, float(loss.detach()) / batch, skipped, time.time() - t0))
    return net


# ------------------------------------------------------------------ training
def train(data, scale, iters=4000, lr=2e-3, batch=32, seed=0, val_frac=0.1, row_floor=0.05, diff=True, depth=4,
          log=print, engine="hip"):
    """Adam on the per-sample relative squared error of the corrections; returns the network
    and the held-out relative L2 error per edge.  engine "hip": the network's forward and
    backward run on the library's training kernels (BF16 MFMA, the inference kernel's
    rounding); "torch": FP32 torch convolutions."""
    torch.manual_seed(seed)
    inv = (1.0 / scale).view(1, 10, 1, 1)
    sets = []
    for X, Y in data:
        Xn = inputs(X, diff) * inv
        Yn = Y * inv[:, :5]
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
            pred = edge_forward_hip(net, e, Xt[idx]) if engine == "hip" else net.edge_forward(e, Xt[idx])
            y = Yt[idx]
            rw = row_weights(e, floor=row_floor, depth=depth)
            loss = loss + (rw * (pred - y) ** 2).sum() / (rw * y ** 2).sum().clamp_min(1e-30)
        loss.backward()
        opt.step()
        sched.step()
        if log and (it % 500 == 0 or it == iters - 1):
            log("abc train: iter %d loss %.4f (%.1f s)" % (it, float(loss.detach()) / 4, time.time() - t0))
    val = []
    with torch.no_grad():
        for e in range(4):
            Xv, Yv = sets[e][2], sets[e][3]
            rw = row_weights(e, floor=row_floor, depth=depth)
            err = sum(float((rw * (net.edge_forward(e, Xv[i:i + 64], bf16=True) - Yv[i:i + 64]) ** 2).sum())
                      for i in range(0, Xv.shape[0], 64))
            val.append((err / float((rw * Yv ** 2).sum())) ** 0.5)
    return net, val


# ------------------------------------------------------------------ reflection test
def _edge_receivers(nz_i, nx_i, rec_every, inset, d=10.0):
    """Receivers every `rec_every` cells along the four interior edges, `inset` cells in."""
    zs = np.arange(inset, nz_i - inset, rec_every) * d
    xs = np.arange(inset, nx_i - inset, rec_every) * d
    rz = np.concatenate([np.full(xs.size, inset * d), np.full(xs.size, (nz_i - 1 - inset) * d), zs, zs])
    rx = np.concatenate([xs, xs, np.full(zs.size, inset * d), np.full(zs.size, (nx_i - 1 - inset) * d)])
    return rz, rx


def reflection_test(vvr, net=None, scale=None, strip=10, mid=60, wide=120, nts=1500, dts=0.001, sub=2,
                    src=None, rec_every=2, inset=3, device=0, time_block=(1, 2), on=1, log=print,
                    variants=("plain", "mid", "learned", "nosponge"), damping=None):
    """Shot records of the same interior model under the wide sponge (reference), the plain
    `strip` sponge, the plain `mid` sponge and (with a network) the learned `strip` boundary.
    Receivers: every `rec_every` cells along the four interior edges, `inset` cells in.
    Runs without a network are timed with each forward form in `time_block` (fine steps per
    launch, the library's automatic tile choice for each) and report the fastest; the learned
    runs take one fine step per launch (the network runs after every step).
    Returns {name: dict(refl=relative energy of the difference from the reference,
    ms=device time of the fastest form, ms_by_time_block, gcells=GCell/s of the propagated
    grid, interior_gcells=interior cells per second)}."""
    forms_plain = tuple(time_block) if isinstance(time_block, (tuple, list)) else (int(time_block),)
    ctx = _capi.Context(device, np.float32)
    nx_i, nz_i = vvr.shape[1], vvr.shape[2]
    d = 10.0
    rz, rx = _edge_receivers(nz_i, nx_i, rec_every, inset, d)
    if src is None:
        src = (0.3 * nz_i * d, 0.5 * nx_i * d)
    ntw = (nts - 1) * 2 * sub + 1
    interior = nz_i * nx_i * (ntw - 1)
    out = {}
    runs = [("wide%d" % wide, wide, False)]
    if "plain" in variants:
        runs.append(("plain%d" % strip, strip, False))
    if "mid" in variants:
        runs.append(("plain%d" % mid, mid, False))
    if net is not None and "learned" in variants:
        runs.append(("learned%d" % strip, strip, True))
    if net is not None and "nosponge" in variants:
        runs.append(("learned%d_nosponge" % strip, strip, "nosponge"))
    if damping is not None:
        runs.append(("fitted%d" % strip, strip, "fitted"))
    ref = None
    for name, pad, learned in runs:
        w = padded(vvr, pad)
        if learned == "nosponge":
            w["minPad"] = 0
        shot = Shot(ctx, w, nts, dts, sub, src[0], src[1], rz, rx)
        if learned == "fitted":
            ctx.set_damping(shot.mid, np.asarray(damping, np.float32))
            learned = False
        if learned:
            ctx.set_abc(shot.mid, on, net.flat(), scale)
        times, data = {}, None
        for tb in ((1,) if learned else forms_plain):
            ctx.set_time_block(tb, _capi.VARIANT_AUTO)
            shot.forward()  # warm-up
            d_tb, times[tb] = shot.forward()
            if data is None:
                data = d_tb
        ms = min(times.values())
        if ref is None:
            ref = data
        refl = float(((data - ref) ** 2).sum() / (ref ** 2).sum())
        if not np.isfinite(refl):
            refl = float("inf")
        out[name] = dict(refl=refl, ms=ms, ms_by_time_block={str(k): v for k, v in times.items()},
                         gcells=shot.cells / ms * 1e-6, interior_gcells=interior / ms * 1e-6,
                         grid=[w["nz"], w["nx"]])
        if log:
            log("abc %s: grid %dx%d, %.1f ms (%s), %.2f GCell/s, reflection energy %.3e" %
                (name, w["nz"], w["nx"], ms, ", ".join("%s step%s %.1f" % (k, "" if k == 1 else "s", v)
                                                          for k, v in times.items()),
                 out[name]["gcells"], refl))
        ctx.model_remove(shot.mid)
    ctx.close()
    return out


# ------------------------------------------------------------------ fitted damping table
def cosine_table(strip, alpha=0.99):
    """The reference's sponge factors (nonlinearPropElasticGpuFunctions.cu:195-202)."""
    j = np.arange(strip)
    return alpha + (1.0 - alpha) * np.cos(np.pi / strip * (strip - j))


def fit_damping(strip=10, wide=120, n_models=3, seed=55, nz_i=200, nx_i=400, nts=1500, dts=0.001, sub=2,
                evals=400, rec_every=2, inset=3, log=print):
    """A boundary that is stable by construction: the narrow strip's damping table itself,
    `strip` factors in (0, 1] (d = exp(-exp(theta)), so no table can add energy), fitted to
    wide-sponge runs that the propagator generates.  Minimises the mean reflection energy
    (relative to the wide run's record) over n_models seeded layered models with Nelder-Mead
    from the cosine table; returns (table, {"start": .., "best": .., "evals": ..})."""
    from scipy.optimize import minimize
    ctx = _capi.Context(0, np.float32)
    ctx.set_time_block(2, _capi.VARIANT_AUTO)
    rng = np.random.default_rng(seed)
    rz, rx = _edge_receivers(nz_i, nx_i, rec_every, inset)
    cases = []
    for i in range(n_models):
        vvr = W.layered_vvr(rng, nz_i, nx_i)
        src = (float(rng.uniform(0.2, 0.8)) * nz_i * 10.0, float(rng.uniform(0.2, 0.8)) * nx_i * 10.0)
        ref_shot = Shot(ctx, padded(vvr, wide), nts, dts, sub, src[0], src[1], rz, rx)
        ref, _ = ref_shot.forward()
        ctx.model_remove(ref_shot.mid)
        cases.append((Shot(ctx, padded(vvr, strip), nts, dts, sub, src[0], src[1], rz, rx), ref))

    def loss(theta):
        table = np.exp(-np.exp(np.clip(theta, -30.0, 3.0))).astype(np.float32)
        table = np.maximum(table, np.float32(1e-6))
        r = 0.0
        for shot, ref in cases:
            ctx.set_damping(shot.mid, table)
            data, _ = shot.forward()
            e = float(((data - ref) ** 2).sum() / (ref ** 2).sum())
            r += e if np.isfinite(e) else 1e30
        return r / len(cases)

    theta0 = np.log(-np.log(cosine_table(strip)))
    start = loss(theta0)
    res = minimize(loss, theta0, method="Nelder-Mead",
                   options={"maxfev": evals, "xatol": 1e-3, "fatol": 1e-9 * start, "adaptive": True})
    table = np.exp(-np.exp(np.clip(res.x, -30.0, 3.0)))
    if log:
        log("fitted damping: mean reflection %.3e (cosine table %.3e) after %d evaluations; table %s" %
            (res.fun, start, res.nfev, " ".join("%.4f" % v for v in table)))
    for shot, _ in cases:
        ctx.model_remove(shot.mid)
    ctx.close()
    return table, {"start": start, "best": float(res.fun), "evals": int(res.nfev)}


def with_gain(net, gamma):
    """The network with its output scaled by gamma (the last layer's weights times gamma: the
    network is positively homogeneous in its output layer)."""
    st = {"%s_%d" % (EDGES[e], l): net.W[e][l].detach().cpu().numpy() * (gamma if l == 2 else 1.0)
          for e in range(4) for l in range(3)}
    return AbcNet.from_state(st, net.W[0][0].device)


def calibrate_gain(net, scale, on, strip=10, gammas=(1.0, 0.7, 0.5, 0.35, 0.25, 0.15, 0.1), n_models=3, seed=77,
                   nz_i=200, nx_i=400, nts=1500, log=print):
    """Output gain of the network chosen on held-out layered models: the gain (and sponge or
    none) with the lowest mean reflection relative to the plain narrow sponge among those whose
    full-length rollouts stay bounded (within twice the plain reflection) on every model; gain
    0 (the plain sponge) unless one beats it on average.  Returns (best dict, table)."""
    rng = np.random.default_rng(seed)
    vvrs = [W.layered_vvr(rng, nz_i, nx_i) for _ in range(n_models)]
    plain = []
    for v in vvrs:
        r = reflection_test(v, None, scale, strip=strip, nts=nts, on=on, log=None, variants=("plain",))
        plain.append(r["plain%d" % strip]["refl"])
    table = []
    best = {"gamma": 0.0, "sponge": True, "ratio": 1.0}
    for gm in gammas:
        ng = with_gain(net, gm)
        rs = [reflection_test(v, ng, scale, strip=strip, nts=nts, on=on, log=None, variants=("learned", "nosponge"))
              for v in vvrs]
        for key, sponge in (("learned%d" % strip, True), ("learned%d_nosponge" % strip, False)):
            ratios = [r[key]["refl"] / p for r, p in zip(rs, plain)]
            ratio = float(np.mean(ratios)) if all(np.isfinite(ratios)) and max(ratios) < 2.0 else float("inf")
            table.append({"gamma": gm, "sponge": sponge, "ratio_to_plain": ratio,
                          "per_model": [float(x) for x in ratios]})
            if log:
                log("abc gain %.2f sponge %d: reflection / plain = %s" % (gm, sponge, " ".join("%.3g" % x for x in ratios)))
            if ratio < best["ratio"]:
                best = {"gamma": gm, "sponge": sponge, "ratio": ratio}
    return best, table


def main(argv=None):
    """Generate, train, save and test: python -m elastic_isotropic_operators_amd.abc_learn
    --out weights.npz [--models 200 --iters 4000 --test-nz 1000 --test-nx 3000 --test-nts 4000]"""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--models", type=int, default=200)
    ap.add_argument("--snaps", type=int, default=24)
    ap.add_argument("--iters", type=int, default=4000)
    ap.add_argument("--strip", type=int, default=10)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--test-nz", type=int, default=200)
    ap.add_argument("--test-nx", type=int, default=400)
    ap.add_argument("--test-nts", type=int, default=1500)
    ap.add_argument("--row-floor", type=float, nargs="+", default=[0.05])
    ap.add_argument("--rollout", type=int, default=0, help="rollout fine-tuning iterations (0: none)")
    ap.add_argument("--seq-models", type=int, default=100)
    ap.add_argument("--modes", nargs="+", default=["level"], choices=["diff", "level"])
    ap.add_argument("--depths", type=int, nargs="+", default=[4])
    ap.add_argument("--engine", default="hip", choices=["hip", "torch"],
                    help="training forward/backward: the library's MFMA kernels or torch FP32")
    args = ap.parse_args(argv)

    def log(m):
        print("[abc] " + m, flush=True)

    data = generate(n_models=args.models, seed=args.seed, strip=args.strip, snaps_per_model=args.snaps, log=log)
    vvr = W.layered_vvr(np.random.default_rng(1234), args.test_nz, args.test_nx)
    seqs, meds, wn = (generate_sequences(n_models=args.seq_models, strip=args.strip, log=log) if args.rollout
                      else (None, None, None))
    out = []
    for mode in args.modes:
        diff = mode == "diff"
        on = 1 if diff else 2
        scale = input_scale(data, diff)
        log("mode %s scale %s" % (mode, " ".join("%.3e" % v for v in scale.tolist())))
        for fl in args.row_floor:
          for depth in args.depths:
            net, val = train(data, scale, iters=args.iters, row_floor=fl, diff=diff, depth=depth, log=log,
                             engine=args.engine)
            log("mode %s row floor %g depth %d: held-out relative error per edge %s" % (
                mode, fl, depth, " ".join("%.3f" % v for v in val)))
            if args.rollout:
                net = train_rollout(net, seqs, meds, wn, scale, iters=args.rollout, diff=diff, log=log)
            on_d = on + 10 * depth
            best, table = calibrate_gain(net, scale.cpu().numpy(), on_d, strip=args.strip, log=log)
            log("mode %s row floor %g depth %d: chosen gain %.2f sponge %d, reflection / plain %.3f" % (
                mode, fl, depth, best["gamma"], best["sponge"], best["ratio"]))
            netg = with_gain(net, best["gamma"])
            tag = "" if len(args.modes) * len(args.row_floor) * len(args.depths) == 1 else "_%s_floor%g_depth%d" % (
                mode, fl, depth)
            save(args.out.replace(".npz", tag + ".npz"), netg, scale.cpu().numpy(), args.strip, on_d,
                 info={"val": val, "models": args.models, "iters": args.iters, "seed": args.seed, "row_floor": fl,
                       "rollout": args.rollout, "gain": best["gamma"], "sponge": int(best["sponge"]),
                       "calibration_ratio": best["ratio"], "depth": depth})
            res = reflection_test(vvr, netg, scale.cpu().numpy(), strip=args.strip, nts=args.test_nts, on=on_d,
                                  log=log)
            out.append({"mode": mode, "row_floor": fl, "depth": depth, "val": val, "gain": best,
                        "calibration": table, "test": res})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
