"""Shared helpers for the parity tests: workload -> regular-level inputs."""
import numpy as np

import oracle as O
from elastic_isotropic_operators_amd import workloads as W

GRID_OF_COMP = (1, 2, 0, 0, 3)


def oracle_grid(w):
    return O.grid(w["nz"], w["nx"], w["nts"], 2 * w["sub"], w["dz"], w["dx"], w["dts"],
                  W.min_pad(w), w["alphaCos"], w["oz"], w["ox"])


def reg_positions(w, zc, xc):
    """Unique regular positions on the 4 staggered grids (centre, X, Z, XZ)."""
    oz, ox, dz, dx = w["oz"], w["ox"], w["dz"], w["dx"]
    origins = [(oz, ox), (oz, ox - 0.5 * dx), (oz - 0.5 * dz, ox), (oz - 0.5 * dz, ox - 0.5 * dx)]
    return [O.interp(w["nz"], o[0], dz, o[1], dx, np.atleast_1d(zc), np.atleast_1d(xc))[2]
            for o in origins]


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def random_traces(rng, spos, nt, dtype, smooth=True):
    out = []
    for k in range(5):
        t = rng.standard_normal((spos[GRID_OF_COMP[k]].size, nt))
        if smooth:
            t = np.cumsum(t, axis=1)
            t -= t.mean(axis=1, keepdims=True)
        out.append(np.ascontiguousarray(t, dtype=dtype))
    return out


def operator_case(tmp_path, w, nExp=1, zSource=4, xSource=None, spacing=None, nRec=None, depthRec=2,
                  oRec=2, dRec=2, extra=None):
    """Write the padded VpVsRho model and a Ricker source file and return (args, coords) for
    the parfile-driven init functions, plus the oracle-side coordinates of the same geometry."""
    from elastic_isotropic_operators_amd.sep import Hypercube, SepVector, genericIO
    nx_int = w["nx"] - 8 - w["xPadMinus"] - w["xPadPlus"]
    xSource = xSource if xSource is not None else nx_int // 3
    spacing = spacing if spacing is not None else max(1, nx_int // (2 * nExp))
    nRec = nRec if nRec is not None else nx_int // 2 - 2
    hyp = Hypercube.hypercube(axes=[Hypercube.axis(n=w["nz"], o=w["oz"], d=w["dz"]),
                                    Hypercube.axis(n=w["nx"], o=w["ox"], d=w["dx"]),
                                    Hypercube.axis(n=3)])
    mod = SepVector.getSepVector(hyp)
    mod.getNdArray()[...] = w["vvr"]
    mfile = str(tmp_path / "model.H")
    genericIO.defaultIO.writeVector(mfile, mod)
    src = SepVector.getSepVector(Hypercube.hypercube(
        axes=[Hypercube.axis(n=w["nts"], d=w["dts"]), Hypercube.axis(n=1), Hypercube.axis(n=5)]))
    src.getNdArray()[...] = w["model"][:, :, :]
    sfile = str(tmp_path / "sources.H")
    genericIO.defaultIO.writeVector(sfile, src)
    par = dict(nts=w["nts"], dts=w["dts"], sub=w["sub"], nz=w["nz"], nx=w["nx"], dz=w["dz"], dx=w["dx"],
               fMax=w["fMax"], mod_par=1, zPadMinus=w["zPadMinus"], zPadPlus=w["zPadPlus"],
               xPadMinus=w["xPadMinus"], xPadPlus=w["xPadPlus"], nExp=nExp, zSource=zSource,
               xSource=xSource, spacingShots=spacing, nReceiver=nRec, depthReceiver=depthRec,
               oReceiver=oRec, dReceiver=dRec, iGpu=0, blockSize=16, elasticParam=mfile, sources=sfile)
    if extra:
        par.update(extra)
    args = ["%s=%s" % (k, v) for k, v in par.items()]
    sz, sx = W.source_coords(w, zSource, xSource, nExp=nExp, spacingShots=spacing)
    rz, rx = W.receiver_coords(w, nRec, depthRec, oRec, dRec)
    return args, dict(sz=sz, sx=sx, rz=rz, rx=rx)


def scattered_positions(rng, nz, nx, count):
    """`count` distinct update-region positions (ix*nz+iz) on each of the 4 grids, spread over
    the whole grid so that many fall on or next to tile edges."""
    out = []
    for _ in range(4):
        iz = rng.integers(4, nz - 4, size=4 * count)
        ix = rng.integers(4, nx - 4, size=4 * count)
        p = np.unique(ix * nz + iz)[:count]
        out.append(rng.permutation(p).astype(np.int32))
    return out


def smooth_field(rng, shape, sigma=4.0):
    from scipy.ndimage import gaussian_filter
    f = gaussian_filter(rng.standard_normal(shape), sigma)
    return f / np.sqrt(np.mean(f ** 2))


def fwi_case(nts=800, dts=0.0005, fdom=12.0, seed=0, nz=48, nx=64, nExp=1):
    """Small FWI problem: smooth random background m0 = (rho, lambda, mu), a 3 % smooth
    perturbation dtrue for the observed data, explosive Ricker sources."""
    w = W.small(nz=nz, nx=nx, nts=nts, sub=1, fDom=fdom, nExp=nExp, seed=seed)
    w["dts"] = dts
    w["model"] = W.explosive_model(nts, dts, fdom, 0.05)
    m0 = w["elastic"].astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    dtrue = np.stack([0.03 * m0[k].mean() * smooth_field(rng, m0.shape[1:]) for k in range(3)])
    return w, oracle_grid(w), m0, dtrue


def fwi_directions(m0, n, seed=33, frac=0.01):
    rng = np.random.default_rng(seed)
    return [np.stack([frac * m0[k].mean() * smooth_field(rng, m0.shape[1:]) for k in range(3)])
            for _ in range(n)]


def oracle_nl_irregular(w, g, src_z, src_x, rec_z, rec_x, src_model, src_interp=("linear", 1),
                        rec_interp=("linear", 1)):
    """One NL shot through the oracle with any spatial interpolation: irregular sources ->
    regular (interpolation adjoint, scaled by rhoxDtw / rhozDtw at the regular points or
    2 dtw, and 1/(dx dz): nonlinearPropElasticGpu.cpp:109-144) -> oracle regular-level NL
    forward -> irregular receivers.  Returns (5, nRec, nts) in FP64."""
    med = O.media(g, w["elastic"])
    sips = O.staggered_interp(g, src_z, src_x, *src_interp)   # (method, nfilt[, dipole, zshift, xshift])
    rips = O.staggered_interp(g, rec_z, rec_x, *rec_interp)
    dtw = g.dts / g.sub
    area = 1.0 / (g.dx * g.dz)
    src = []
    for k in range(5):
        ip = sips[GRID_OF_COMP[k]]
        reg = O.irreg_to_reg(ip, np.asarray(src_model[k], np.float64))
        scale = med[k].reshape(-1)[ip[2]][:, None] if k < 2 else 2.0 * dtw
        src.append(reg * scale * area)
    dreg = O.nl_fwd_reg(g, med, [ip[2] for ip in sips], src, [ip[2] for ip in rips])
    return np.stack([O.reg_to_irreg(rips[GRID_OF_COMP[k]], dreg[k]) for k in range(5)])
