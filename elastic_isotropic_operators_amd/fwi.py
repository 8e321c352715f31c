"""Elastic FWI on the GPU: the checkpointed gradient behind ``torch.autograd.Function``.

The reference's FWI script (efwiMain.py:229-476) minimises

    f(m) = 1/2 || C F(m) - d_obs ||^2

with F the nonlinear shot operator evaluated at the elastic model m = (rho, lambda, mu)
(nonlinearFwiPropElasticShotsGpu, Elastic_iso_float_prop.py:1107-1151), C the data-component
selection of ElasticDatComp (dataCompModule.py:9-107), and takes as gradient the Born adjoint
at the current model applied to C^T (C F(m) - d_obs) (BornElasticShotsGpu.cpp:228-378).

Here the forward and the gradient are one hand-written HIP pair behind the C ABI
(eio_fwi_forward / eio_fwi_gradient): the forward saves the batch state every
~sqrt(nts) coarse samples instead of the reference's full nz*nx*nts storage of Vx and Vz,
and the gradient recomputes one segment at a time.  Everything stays resident in HBM:
media, traces and the gradient are torch tensors whose device addresses go straight to the
C ABI (torch's stream is synchronized before each call; the call synchronizes its own).
torch autograd supplies C^T and the chain rule through any parameter conversion written in
torch (``vpvsrho_to_rholamemu``).

The gradient ignores the model dependence of the vx/vz source scaling (rhoxDtw, rhozDtw at
the source positions), as the reference's Born-based gradient does.

Import order: torch's ROCm wheel carries its own HIP runtime with the system runtime's soname;
torch must be imported before libeio_hip.so is loaded, so the library binds to that runtime
and torch's device pointers are valid for it.
"""
import weakref

import numpy as np
import torch

from . import _capi
from . import geometry as G

GRID_OF_COMP = G.GRID_OF_COMP
_COMP = {"vx": 0, "vz": 1, "sxx": 2, "szz": 3, "sxz": 4}


# ------------------------------------------------------------------ torch host math
def _sx_adj(f):
    """staggerX::adjoint (stagger.cpp:38-57) on (..., nx, nz)."""
    out = torch.empty_like(f)
    out[..., 1:-1, :] = 0.5 * (f[..., 1:-1, :] + f[..., :-2, :])
    out[..., 0, :] = 0.5 * f[..., 0, :]
    out[..., -1, :] = f[..., -1, :] + 0.5 * f[..., -2, :]
    return out


def _sz_adj(f):
    """staggerZ::adjoint (stagger.cpp:95-114)."""
    out = torch.empty_like(f)
    out[..., 1:-1] = 0.5 * (f[..., 1:-1] + f[..., :-2])
    out[..., 0] = 0.5 * f[..., 0]
    out[..., -1] = f[..., -1] + 0.5 * f[..., -2]
    return out


def _sx_fwd(f):
    """staggerX::forward (stagger.cpp:18-35)."""
    out = torch.empty_like(f)
    out[..., :-1, :] = 0.5 * (f[..., :-1, :] + f[..., 1:, :])
    out[..., -1, :] = f[..., -1, :]
    return out


def _sz_fwd(f):
    """staggerZ::forward (stagger.cpp:75-93)."""
    out = torch.empty_like(f)
    out[..., :-1] = 0.5 * (f[..., :-1] + f[..., 1:])
    out[..., -1] = f[..., -1]
    return out


def scaled_media(model, dtw):
    """(3, nx, nz) [rho, lambda, mu] -> (5, nx, nz) rhoxDtw, rhozDtw, lamb2MuDtw, lambDtw,
    muxzDtw (fdParamElastic.cpp:118-148)."""
    rho, lam, mu = model[0], model[1], model[2]
    two = 2.0 * dtw
    return torch.stack([two / _sx_adj(rho), two / _sz_adj(rho), two * (lam + 2.0 * mu), two * lam,
                        two * _sz_adj(_sx_adj(mu))]).contiguous()


def slices_to_gradient(acc, media, dtw):
    """Stacked imaging condition (5, nx, nz) -> (drho, dlambda, dmu) (3, nx, nz):
    scale by the media and 2*dtw, then the stagger transposes (BornElasticShotsGpu.cpp:325-371)."""
    two = 2.0 * dtw
    a0 = acc[0] * media[0]
    a1 = acc[1] * media[1]
    drho = _sx_fwd(a0) + _sz_fwd(a1)
    dlam = two * acc[2]
    dmu = two * acc[3] + _sx_fwd(_sz_fwd(two * acc[4]))
    return torch.stack([drho, dlam, dmu])


def vpvsrho_to_rholamemu(m):
    """ElasticConv.forward with mod_par=1 (elasticParamConvertModule.py:23-47), in torch so
    autograd applies the conversion Jacobian (ElasticConvJab.adjoint) in the backward."""
    vp, vs, rho = m[0], m[1], m[2]
    return torch.stack([rho, rho * (vp * vp - 2.0 * vs * vs), rho * vs * vs])


def _datcomp_adjoint(sampled, components):
    """ElasticDatComp.adjoint (dataCompModule.py:79-107) on (nShot, nComp, nRec, nts)."""
    shape = (sampled.shape[0], 5) + tuple(sampled.shape[2:])
    out = torch.zeros(shape, dtype=sampled.dtype, device=sampled.device)
    for i, c in enumerate(components.lower().split(",")):
        if c == "p":
            out[:, 2] += 0.5 * sampled[:, i]
            out[:, 3] += 0.5 * sampled[:, i]
        else:
            out[:, _COMP[c]] += sampled[:, i]
    return out


def datcomp(data, components="vx,vz,p"):
    """ElasticDatComp.forward (dataCompModule.py:45-77) on (nShot, 5, nRec, nts)."""
    out = []
    for c in components.lower().split(","):
        out.append(0.5 * (data[:, 2] + data[:, 3]) if c == "p" else data[:, _COMP[c]])
    return torch.stack(out, dim=1)


# ------------------------------------------------------------------ engine
class _Store:
    """A checkpoint store held by the context until released (or garbage collected)."""

    def __init__(self, engine, handle, srcs):
        self._ctx = weakref.ref(engine.ctx)
        self.handle = handle
        self.srcs = srcs  # the scaled source tensors the store's forward used

    def release(self):
        ctx = self._ctx()
        if self.handle is not None and ctx is not None and ctx.h:
            ctx.fwi_release(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# Working-set budget of a shot group, as the library's automatic shot groups of the NL forward
# (EIO_GROUP_AUTO: 192 MiB of the MI355X's 256 MB Infinity Cache).
CACHE_GROUP_BYTES = 192 << 20
# Values per grid cell that one step of the gradient's passes touches: the receiver pass's four
# adjoint state levels and the media (25), its adjoint-field slots (10) and the record's Vx/Vz
# slices (4), rounded up.
GRADIENT_VALUES_PER_CELL = 40


def cache_group(nz, nx, itemsize, budget=CACHE_GROUP_BYTES):
    """Shots per group whose gradient passes keep their working set within `budget` (at least 1)."""
    return max(1, int(budget // (GRADIENT_VALUES_PER_CELL * int(nz) * int(nx) * int(itemsize))))


def group_and_segment(group, fit, nshot, nts, segment, checkpoint=None):
    """Shots per group and the store's segment length (coarse intervals) of one
    misfit_and_gradient call.  fit: shots whose whole coarse Vx/Vz record fits in memory;
    group: the caller's group size (None: `fit`); checkpoint: None chooses (full storage when
    the group fits, else checkpoints), else a checkpoint interval that is always used (0:
    ceil(sqrt(nts-1))).  A group that does not fit takes checkpoints (interval `segment`,
    0: ceil(sqrt(nts-1))) instead of failing for memory."""
    g = fit if group is None else int(group)
    if checkpoint is not None:
        return max(1, min(nshot if group is None else g, nshot)), int(checkpoint)
    if 1 <= g <= fit:
        return min(g, nshot), nts - 1
    return max(1, min(g, nshot)), int(segment)


class ElasticFwi:
    """Shots of one model grid, resident on one GPU.

    grid:    dict with nz, nx, dz, dx, oz, ox (padded grid incl. the FAT halo), nts, dts,
             sub (user sub-sampling; the fine step is dts / (2 sub)), minPad, alphaCos
    src_z, src_x: (nShot, nSrc) source coordinates
    sources: (5, nSrc, nts) shared by all shots, or (nShot, 5, nSrc, nts)
    rec_z, rec_x: (nRec,) receiver coordinates shared by all shots
    segment: checkpoint interval in coarse samples; 0: the whole record (no checkpoint, nothing
             recomputed) when the batch's full coarse Vx/Vz storage fits in 60 % of the free
             device memory, else ceil(sqrt(nts-1))
    src_interp, rec_interp: (method, nfilt) of the spatial interpolation of sources and
             receivers: ("linear", 1), ("sinc", n) or ("gauss", n) (spaceInterpGpu)
    time_block: (fine steps per launch, two-step kernel variant) of the FP32 forwards;
             variant _capi.VARIANT_AUTO picks the tile width per call from the batch size
    adj_time_block: (reverse fine steps per launch, two-step adjoint variant) of the FP32
             gradient's receiver pass (VARIANT_AUTO as for time_block)
    """

    def __init__(self, grid, src_z, src_x, sources, rec_z, rec_x, device=0, dtype=torch.float32,
                 segment=0, variant=None, src_interp=("linear", 1), rec_interp=("linear", 1),
                 time_block=(2, _capi.VARIANT_AUTO),
                 adj_time_block=(2, _capi.VARIANT_AUTO)):
        if not torch.cuda.is_available():
            if _capi._lib is not None:
                raise RuntimeError("torch sees no GPU because libeio_hip.so was loaded before torch: "
                                   "import torch first so both share torch's HIP runtime")
            raise RuntimeError("ElasticFwi needs a GPU (the propagator is HIP-only)")
        self.g = dict(grid)
        self.nz, self.nx, self.nts = int(grid["nz"]), int(grid["nx"]), int(grid["nts"])
        self.sub_fd = 2 * int(grid["sub"])
        self.dtw = float(grid["dts"]) / self.sub_fd
        self.dtype = dtype
        self.device = torch.device("cuda", device)
        self.segment = int(segment)
        self.ctx = _capi.Context(device, np.float32 if dtype == torch.float32 else np.float64)
        if variant is not None:
            self.ctx.set_variant(variant)
        # a forward that keeps its whole coarse record runs two fine steps per launch
        # (time_block = (steps, two-step variant); (1, 0): one step per launch)
        self.ctx.set_time_block(*time_block)
        self.ctx.set_adjoint_time_block(*adj_time_block)
        self.mid = None
        self.last_store_info = None  # (interval, checkpoints, device bytes) of the last store
        src_z = np.atleast_2d(np.asarray(src_z, np.float64))
        src_x = np.atleast_2d(np.asarray(src_x, np.float64))
        self.nshot = src_z.shape[0]
        sig = np.asarray(sources, np.float64)
        if sig.ndim == 3:
            sig = np.broadcast_to(sig, (self.nshot,) + sig.shape)
        if sig.shape[0] != self.nshot or sig.shape[1] != 5 or sig.shape[3] != self.nts:
            raise ValueError("sources must be (5, nSrc, nts) or (nShot, 5, nSrc, nts)")
        oz, ox, dz, dx = grid["oz"], grid["ox"], grid["dz"], grid["dx"]
        self.area = 1.0 / (float(dx) * float(dz))
        # sources: unscaled regular traces per shot and component (irregular -> regular,
        # spaceInterpGpu.cpp:313-376), plus the positions the media scaling is read at
        self.src_pos, self.src_reg = [], []
        for e in range(self.nshot):
            ips = G.staggered_interps(self.nz, oz, dz, ox, dx, src_z[e], src_x[e], *src_interp)
            self.src_pos.append([ip.uniq for ip in ips])
            self.src_reg.append([torch.as_tensor(ips[GRID_OF_COMP[k]].adjoint(sig[e, k]), dtype=dtype,
                                                 device=self.device) for k in range(5)])
        self.src_idx = [[torch.as_tensor(self.src_pos[e][GRID_OF_COMP[k]].astype(np.int64), device=self.device)
                         for k in range(2)] for e in range(self.nshot)]
        rips = G.staggered_interps(self.nz, oz, dz, ox, dx, np.asarray(rec_z, np.float64),
                                   np.asarray(rec_x, np.float64), *rec_interp)
        self.rec_pos = [ip.uniq for ip in rips]
        self.nrec = rips[0].n_irreg
        self.rec_map = [torch.as_tensor(ip.map, device=self.device) for ip in rips]
        self.rec_w = [torch.as_tensor(ip.w, dtype=dtype, device=self.device) for ip in rips]

    # -------------------------------------------------------------- C-ABI plumbing
    def _torch_ready(self):
        """The propagator runs on its context's own stream; torch's work on the inputs must be
        complete first.  (torch's default stream is the legacy null stream, handle 0, which
        eio_ctx_set_stream would read as 'own stream' and which does not order work on a
        non-blocking stream.)  The C-ABI calls synchronize their stream before returning, so
        their outputs are ready for torch."""
        torch.cuda.current_stream(self.device).synchronize()

    def _set_media(self, media):
        n = self.nz * self.nx
        ptrs = [media.data_ptr() + k * n * media.element_size() for k in range(5)]
        if self.mid is None:
            g = self.g
            self.mid = self.ctx.model_add(self.nz, self.nx, self.nts, self.sub_fd, g["dz"], g["dx"], g["dts"],
                                          g["minPad"], None, g.get("alphaCos", 0.99), device_ptrs=ptrs)
        else:
            self.ctx.model_set_media(self.mid, device_ptrs=ptrs)

    def _sel(self, sel):
        return list(range(self.nshot)) if sel is None else list(sel)

    def _scaled_sources(self, media, sel=None):
        """nonlinearPropElasticGpu.cpp:109-144: vx, vz by rhoxDtw, rhozDtw at the device
        positions, stresses by 2*dtw, all by 1/(dx dz); for the shots `sel` (default all)."""
        out = []
        for e in self._sel(sel):
            per = []
            for k in range(5):
                r = self.src_reg[e][k]
                if k < 2:
                    f = media[k].reshape(-1)[self.src_idx[e][k]].unsqueeze(1)
                    per.append((r * f * self.area).contiguous())
                else:
                    per.append((r * (2.0 * self.dtw) * self.area).contiguous())
            out.append(per)
        return out

    def _shots(self, srcs, recs, sel=None):
        """C-ABI shots of the shots `sel`; srcs and recs are indexed by position in `sel`."""
        shots, keep = [], []
        for i, e in enumerate(self._sel(sel)):
            rec_ptr = [recs[k][i].data_ptr() if recs[k].shape[1] else 0 for k in range(5)]
            src_ptr = [srcs[i][k].data_ptr() if srcs[i][k].numel() else 0 for k in range(5)]
            s, kp = self.ctx.make_shot(self.mid, self.src_pos[e], self.rec_pos, src_ptr, rec_ptr)
            shots.append(s)
            keep.append(kp)
        return shots, keep

    def _reg_to_irreg(self, recs):
        """spaceInterpGpu forward (regular -> irregular) on (nShot, nReg, nts) per component."""
        out = []
        for k in range(5):
            gi = GRID_OF_COMP[k]
            reg = recs[k]
            mp, w = self.rec_map[gi], self.rec_w[gi]
            out.append((reg[:, mp, :] * w[None, :, :, None]).sum(dim=2))
        return torch.stack(out, dim=1)

    def _irreg_to_reg(self, data):
        """spaceInterpGpu adjoint on (nShot, 5, nRec, nts)."""
        recs = []
        ns = data.shape[0]
        for k in range(5):
            gi = GRID_OF_COMP[k]
            mp, w = self.rec_map[gi], self.rec_w[gi]
            reg = torch.zeros((ns, self.rec_pos[gi].size, self.nts), dtype=self.dtype, device=self.device)
            contrib = (data[:, k, :, None, :] * w[None, :, :, None]).reshape(ns, -1, self.nts)
            reg.index_add_(1, mp.reshape(-1), contrib)
            recs.append(reg.contiguous())
        return recs

    def _empty_recs(self, ns=None):
        ns = self.nshot if ns is None else ns
        return [torch.empty((ns, self.rec_pos[GRID_OF_COMP[k]].size, self.nts), dtype=self.dtype,
                            device=self.device) for k in range(5)]

    # -------------------------------------------------------------- operations
    def forward_data(self, model, checkpoint=False):
        """F(model) for all shots -> (nShot, 5, nRec, nts); with checkpoint=True also the
        store the gradient needs.  Returns (data, store or None, media)."""
        model = model.detach().to(device=self.device, dtype=self.dtype)
        media = scaled_media(model, self.dtw)
        srcs = self._scaled_sources(media)
        recs = self._empty_recs()
        self._torch_ready()
        self._set_media(media)
        shots, keep = self._shots(srcs, recs)
        store = None
        if checkpoint:
            seg = self.segment
            if seg == 0 and self.full_storage_group() >= self.nshot:
                seg = self.nts - 1
            h = self.ctx.fwi_forward(shots, seg, device_ptrs=True)
            store = _Store(self, h, srcs)
            self.last_store_info = self.ctx.fwi_info(h)
        else:
            self.ctx.nl_forward(shots, device_ptrs=True)
        return self._reg_to_irreg(recs), store, media

    def gradient(self, store, media, grad_data):
        """Born adjoint of grad_data (nShot, 5, nRec, nts) at the model of `store`, as
        (drho, dlambda, dmu) (3, nx, nz)."""
        if store is None or store.handle is None:
            raise RuntimeError("no checkpoint store: run the forward with gradients enabled")
        recs = self._irreg_to_reg(grad_data.to(self.dtype).contiguous())
        acc = torch.zeros((5, self.nx, self.nz), dtype=self.dtype, device=self.device)
        self._torch_ready()
        self._set_media(media)
        shots, keep = self._shots(store.srcs, recs)
        n = self.nz * self.nx
        ptrs = [acc.data_ptr() + k * n * acc.element_size() for k in range(5)]
        self.ctx.fwi_gradient(store.handle, shots, ptrs, device_ptrs=True)
        return slices_to_gradient(acc, media, self.dtw)

    def store_info(self, store):
        return self.ctx.fwi_info(store.handle)

    def last_timing(self):
        return self.ctx.last_timing()

    def data(self, model):
        """Differentiable F(model): autograd calls the checkpointed gradient.  Without grad
        mode (or for a model that needs no gradient) no checkpoints are kept."""
        if torch.is_grad_enabled() and model.requires_grad:
            return ElasticFwiFunction.apply(model, self)
        return self.forward_data(model)[0]

    def full_storage_group(self, margin=0.6):
        """Shots per group whose full coarse Vx/Vz storage (2 nz nx nts values per shot) fits in
        `margin` of the free device memory (0 when not even one shot fits)."""
        free, _ = torch.cuda.mem_get_info(self.device)
        # the record (2 nts fields) + state levels (4 x 5), recompute state (4 x 5), adjoint-field
        # slots (2 x 5), accumulators (5) and traces, rounded up
        per_shot = (2 * self.nts + 60) * self.nz * self.nx * torch.tensor([], dtype=self.dtype).element_size()
        return int(margin * free // per_shot)

    def default_group(self):
        """Shots per group of misfit_and_gradient when the caller gives none: as many as keep
        the passes' working set in the Infinity Cache (cache_group), at most as many as keep
        their whole coarse record in memory; 0 when not even one record fits."""
        if getattr(self, "_group", None) is None:
            self._group = self.full_storage_group()
        if self._group < 1:
            return 0
        return min(self._group, cache_group(self.nz, self.nx, torch.tensor([], dtype=self.dtype).element_size()))

    def misfit_and_gradient(self, model, d_obs, components=None, group=None, checkpoint=None):
        """1/2 || C F(model) - d_obs ||^2 and its gradient with respect to `model` (rho, lambda,
        mu), shot group by shot group: each group's forward keeps the whole coarse Vx/Vz
        record (one segment, no checkpoints, the reference's full storage,
        BornElasticGpuFunctions.cu:1128-1191, sized to the device) and its gradient follows at
        once, so no segment is recomputed.  The misfit is a sum over shots, so the gradient of
        each group needs only its own residual.  group: shots per group (default:
        default_group(), the shots whose passes stay in the Infinity Cache and whose records fit
        in memory); a group that does not fit, or checkpoint = an interval (0: ceil(sqrt(
        nts-1)), all shots in one group unless `group` is given), takes checkpoints instead."""
        model = model.detach().to(device=self.device, dtype=self.dtype)
        d_obs = torch.as_tensor(d_obs, device=self.device, dtype=self.dtype)
        media = scaled_media(model, self.dtw)
        # sized once: later calls reuse the context's workspace, which free memory no longer shows
        if getattr(self, "_group", None) is None:
            self._group = self.full_storage_group()
        if group is None and checkpoint is None and self._group >= 1:
            group = self.default_group()
        g, segment = group_and_segment(group, self._group, self.nshot, self.nts, self.segment, checkpoint)
        grad = torch.zeros((3, self.nx, self.nz), dtype=self.dtype, device=self.device)
        phi = 0.0
        for s0 in range(0, self.nshot, g):
            sel = list(range(s0, min(self.nshot, s0 + g)))
            srcs = self._scaled_sources(media, sel)
            recs = self._empty_recs(len(sel))
            self._torch_ready()
            self._set_media(media)
            shots, keep = self._shots(srcs, recs, sel)
            h = self.ctx.fwi_forward(shots, segment, device_ptrs=True)
            try:
                data = self._reg_to_irreg(recs)
                d = datcomp(data, components) if components else data
                r = d - d_obs[sel]
                phi += 0.5 * float(torch.sum(r.double() * r.double()))
                gd = _datcomp_adjoint(r, components) if components else r
                grecs = self._irreg_to_reg(gd.contiguous())
                acc = torch.zeros((5, self.nx, self.nz), dtype=self.dtype, device=self.device)
                self._torch_ready()
                shots, keep = self._shots(srcs, grecs, sel)
                n = self.nz * self.nx
                ptrs = [acc.data_ptr() + k * n * acc.element_size() for k in range(5)]
                self.ctx.fwi_gradient(h, shots, ptrs, device_ptrs=True)
                self.last_store_info = self.ctx.fwi_info(h)
            finally:
                self.ctx.fwi_release(h)
            grad += slices_to_gradient(acc, media, self.dtw)
        return phi, grad

    def misfit(self, model, d_obs, components=None):
        """1/2 || C F(model) - d_obs ||^2 (C = ElasticDatComp when components is given)."""
        d = self.data(model)
        if components:
            d = datcomp(d, components)
        r = d - d_obs
        return 0.5 * torch.sum(r * r)


class ElasticFwiFunction(torch.autograd.Function):
    """model (3, nx, nz) [rho, lambda, mu] -> data (nShot, 5, nRec, nts).  The backward is the
    hand-written Born adjoint with checkpoint recomputation (one use per forward)."""

    @staticmethod
    def forward(ctx, model, engine):
        data, store, media = engine.forward_data(model, checkpoint=True)
        ctx.engine, ctx.store, ctx.media = engine, store, media
        return data

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad):
        if ctx.store is None:
            raise RuntimeError("the checkpoint store of this forward was already used")
        g = ctx.engine.gradient(ctx.store, ctx.media, grad.contiguous())
        ctx.store.release()
        ctx.store = None
        return g.to(grad.dtype), None
