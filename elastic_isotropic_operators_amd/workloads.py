"""Synthetic workloads (models, geometry, wavelets) for the propagator.

Each builder returns a plain dict that the operator layer, the tests and the
benchmark consume.  Geometry follows the reference parfile conventions of
Elastic_iso_float_prop.py:205-362 and :525-658: 1-based source/receiver sample
indices shifted by the top/left padding and the FAT halo, converted to physical
coordinates with the padded model origin.

Workloads (sizes are interior sizes, before padding):
  nl_small / born_small   a few-second case for parity tests
  gaussian_c0             200 x 400 Gaussian-anomaly model, nts=1000, 1 shot (NL forward)
  gaussian_c1             same grid, 8 shots, smoothed random perturbations (Born)
  layered_batch_c2        64 layered shots of uneven grid and trace length (batched NL)
  fwi_c3                  32-shot elastic FWI gradient on a faulted layered model with salt
  learned_abc_c4          1000 x 3000 layered model for the learned-boundary comparison
"""
import numpy as np

from . import seis_utils as su

FAT = 4


def _grid_dict(model_vvr, dz, dx, zPad, xPad, nts, dts, sub, fMax, alphaCos=0.99,
               surfaceCondition=0, mod_par=1):
    """Pad a (3, nx, nz) [vp, vs, rho] model and convert it to (rho, lambda, mu)."""
    padded, s = su.pad_model(model_vvr, zPad, xPad, FAT, 16, surfaceCondition)
    elastic = su.vpvsrho_to_rholamemu(padded.astype(np.float64), mod_par)
    oz = 0.0 - (FAT + s["zPadMinus"]) * dz
    ox = 0.0 - (FAT + s["xPadMinus"]) * dx
    return dict(vvr=padded, elastic=elastic, nz=s["nz"], nx=s["nx"], dz=float(dz), dx=float(dx),
                oz=oz, ox=ox, zPadMinus=s["zPadMinus"], zPadPlus=s["zPadPlus"],
                xPadMinus=s["xPadMinus"], xPadPlus=s["xPadPlus"], nts=int(nts), dts=float(dts),
                sub=int(sub), fMax=float(fMax), alphaCos=float(alphaCos),
                surfaceCondition=surfaceCondition, blockSize=16, fat=FAT)


def source_coords(w, zSource, xSource, nExp=1, spacingShots=0):
    """buildSourceGeometry (Elastic_iso_float_prop.py:311-347): one point source per shot."""
    ozS = zSource - 1 + w["zPadMinus"] + FAT
    oxS = xSource - 1 + w["xPadMinus"] + FAT
    z = np.full((nExp, 1), w["oz"] + ozS * w["dz"])
    x = np.array([[w["ox"] + (oxS + i * spacingShots) * w["dx"]] for i in range(nExp)])
    return z, x


def receiver_coords(w, nReceiver, depthReceiver=1, oReceiver=1, dReceiver=1):
    """buildReceiversGeometry (Elastic_iso_float_prop.py:615-641)."""
    ozR = depthReceiver - 1 + w["zPadMinus"] + FAT
    oxR = oReceiver - 1 + w["xPadMinus"] + FAT
    irec = np.arange(nReceiver)
    z = w["oz"] + ozR * w["dz"] + 0.0 * irec
    x = w["ox"] + oxR * w["dx"] + dReceiver * w["dx"] * irec
    return z, x


def explosive_model(nts, dts, fDom, timeDelay, nSrc=1):
    """Model vector (5, nSrc, nts): the same Ricker on sxx and szz."""
    r = su.ricker(nts, dts, fDom, timeDelay)
    m = np.zeros((5, nSrc, nts), np.float32)
    m[2, :, :] = r
    m[3, :, :] = r
    return m


def _gaussian_model(nz, nx, dv=300.0, sigma=20.0, vp0=2500.0):
    iz = np.arange(nz)[None, :]
    ix = np.arange(nx)[:, None]
    r2 = (iz - nz // 2) ** 2 + (ix - nx // 2) ** 2
    vp = vp0 + dv * np.exp(-r2 / (2.0 * sigma * sigma))
    vs = vp / 1.73
    rho = np.full_like(vp, 2000.0)
    return np.stack([vp, vs, rho]).astype(np.float32)


def gaussian_c0():
    """Plain NL forward workload: 200x400 at 10 m, 30-cell sponge, nts=1000, dt=1 ms."""
    w = _grid_dict(_gaussian_model(200, 400), 10.0, 10.0, 30, 30, 1000, 0.001, 1, 37.5)
    w["src_z"], w["src_x"] = source_coords(w, 3, 201)
    w["rec_z"], w["rec_x"] = receiver_coords(w, 200, 1, 1, 2)
    w["model"] = explosive_model(w["nts"], w["dts"], 15.0, 0.1)
    w["name"] = "gaussian_c0_nl_fwd_280x472_nts1000"
    return w


def born_perturbation(w, seed=1, rms_frac=0.05, sigma=5.0):
    """Seeded Gaussian noise smoothed with a 5-cell Gaussian at 5 % RMS, mapped to
    (drho, dlambda, dmu) with the conversion Jacobian at the background, zero in the padding."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    vvr = w["vvr"].astype(np.float64)
    d = np.zeros_like(vvr)
    x0, z0 = FAT + w["xPadMinus"], FAT + w["zPadMinus"]
    x1, z1 = w["nx"] - FAT - w["xPadPlus"], w["nz"] - FAT - w["zPadPlus"]
    for k in range(3):
        noise = gaussian_filter(rng.standard_normal((x1 - x0, z1 - z0)), sigma, mode="nearest")
        bg = vvr[k, x0:x1, z0:z1]
        noise *= rms_frac * np.sqrt(np.mean(bg ** 2)) / np.sqrt(np.mean(noise ** 2))
        d[k, x0:x1, z0:z1] = noise
    return su.conv_jacobian_fwd(vvr, d)


def gaussian_c1():
    """Born workload: C0 grid and background, 8 shots at 50-cell spacing."""
    w = gaussian_c0()
    w["src_z"], w["src_x"] = source_coords(w, 3, 26, nExp=8, spacingShots=50)
    w["dm"] = born_perturbation(w)
    w["name"] = "gaussian_c1_born_280x472_nts1000_8shots"
    return w


def small(nz=40, nx=56, nts=120, sub=1, zPad=10, xPad=10, nExp=1, seed=0, fDom=25.0):
    """A few-second case for parity tests (random smooth model, explosive source)."""
    rng = np.random.default_rng(seed)
    vp = 2200.0 + 400.0 * rng.random((nx, nz))
    from scipy.ndimage import gaussian_filter
    vp = gaussian_filter(vp, 3.0)
    vs = vp / 1.8
    rho = 1800.0 + 300.0 * gaussian_filter(rng.random((nx, nz)), 3.0)
    w = _grid_dict(np.stack([vp, vs, rho]).astype(np.float32), 10.0, 10.0, zPad, xPad, nts,
                   0.002, sub, 20.0)
    w["src_z"], w["src_x"] = source_coords(w, 4, nx // 3, nExp=nExp, spacingShots=max(1, nx // (2 * nExp)))
    w["rec_z"], w["rec_x"] = receiver_coords(w, nx // 2 - 2, 2, 2, 2)
    w["model"] = explosive_model(nts, w["dts"], fDom, 0.05)
    w["name"] = "small_%dx%d_nts%d" % (w["nz"], w["nx"], nts)
    return w


def layered_vvr(rng, nz, nx, nlay_range=(5, 15), vp_range=(1500.0, 4500.0),
                vpvs_range=(1.6, 2.2)):
    """Seeded random layered model; density from Gardner (rho = 310 vp^0.25)."""
    nl = int(rng.integers(nlay_range[0], nlay_range[1] + 1))
    depths = np.sort(rng.integers(1, nz, size=nl - 1))
    vp_l = rng.uniform(*vp_range, size=nl)
    r_l = rng.uniform(*vpvs_range, size=nl)
    layer = np.searchsorted(depths, np.arange(nz), side="right")
    vp = np.broadcast_to(vp_l[layer][None, :], (nx, nz)).astype(np.float64)
    vs = vp / r_l[layer][None, :]
    rho = 310.0 * vp ** 0.25
    return np.stack([vp, vs, rho]).astype(np.float32)


def layered_batch_c2(nshots=64, seed=2, scale=1.0):
    """64 layered shots of uneven grid and trace length (nz 300-600, nx 600-1800,
    nts 2000-4000 at 1 ms, sub=2, Ricker 8 Hz, 40-cell sponge).  ``scale`` < 1
    shrinks every size for quick runs."""
    rng = np.random.default_rng(seed)
    shots = []
    for i in range(nshots):
        nz = int(rng.integers(300, 601) * scale)
        nx = int(rng.integers(600, 1801) * scale)
        nts = int(rng.integers(2000, 4001) * scale)
        vvr = layered_vvr(rng, nz, nx)
        w = _grid_dict(vvr, 10.0, 10.0, 40, 40, nts, 0.001, 2, 20.0)
        w["src_z"], w["src_x"] = source_coords(w, 3, nx // 2 + 1)
        w["rec_z"], w["rec_x"] = receiver_coords(w, nx // 2, 1, 1, 2)
        w["model"] = explosive_model(nts, w["dts"], 8.0, 0.15)
        w["name"] = "c2_shot%02d" % i
        shots.append(w)
    return shots


def _faulted_salt_vvr(rng, nz, nx, nfault=3):
    """Layered model (layered_vvr ranges) cut by `nfault` dipping normal faults and one
    salt-like body (vp 4500, vs 2600, Gardner density)."""
    vvr = layered_vvr(rng, nz, nx).astype(np.float64)
    iz = np.arange(nz)[None, :].astype(np.float64)
    ix = np.arange(nx)[:, None].astype(np.float64)
    for _ in range(nfault):
        x0 = rng.uniform(0.15, 0.85) * nx
        dip = np.deg2rad(rng.uniform(45.0, 75.0)) * (1 if rng.random() < 0.5 else -1)
        throw = int(rng.integers(max(2, nz // 60), max(3, nz // 20) + 1))
        hanging = ix > x0 + iz / np.tan(dip)      # cells beyond the fault plane
        shifted = np.concatenate([np.repeat(vvr[:, :, :1], throw, axis=2), vvr[:, :, :-throw]], axis=2)
        vvr = np.where(hanging[None], shifted, vvr)
    zc, xc = rng.uniform(0.35, 0.6) * nz, rng.uniform(0.3, 0.7) * nx
    a, b = rng.uniform(0.08, 0.15) * nx, rng.uniform(0.12, 0.2) * nz
    ang = np.arctan2(iz - zc, ix - xc)
    r = 1.0 + 0.15 * np.sin(3.0 * ang + rng.uniform(0, 2 * np.pi))
    salt = ((ix - xc) / a) ** 2 + ((iz - zc) / b) ** 2 < r ** 2
    vvr[0][salt] = 4500.0
    vvr[1][salt] = 2600.0
    vvr[2][salt] = 310.0 * 4500.0 ** 0.25
    return vvr.astype(np.float32)


def fwi_c3(nshots=32, seed=3, scale=1.0, smooth_cells=25.0):
    """Elastic FWI gradient workload: 500x1500 interior at 10 m (40-cell sponge -> 600x1592
    padded), nts=3000 at 1 ms, sub=2, Ricker 8 Hz, `nshots` evenly spaced surface shots,
    receivers every 2 cells at the surface.  True model: seeded layered model with 3 dipping
    faults and a salt body; starting model: the true model smoothed with a 25-cell Gaussian.
    Returns (true workload, starting workload); both share geometry and sources."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    nz, nx, nts = int(500 * scale), int(1500 * scale), int(3000 * scale)
    vvr = _faulted_salt_vvr(rng, nz, nx)
    vvr0 = np.stack([gaussian_filter(vvr[k].astype(np.float64), smooth_cells * scale, mode="nearest")
                     for k in range(3)]).astype(np.float32)
    out = []
    for v, tag in ((vvr, "true"), (vvr0, "start")):
        w = _grid_dict(v, 10.0, 10.0, 40, 40, nts, 0.001, 2, 20.0)
        sp = max(1, nx // nshots)
        w["src_z"], w["src_x"] = source_coords(w, 3, sp // 2 + 1, nExp=nshots, spacingShots=sp)
        w["rec_z"], w["rec_x"] = receiver_coords(w, nx // 2, 1, 1, 2)
        w["model"] = explosive_model(nts, w["dts"], 8.0, 0.15)
        w["name"] = "fwi_c3_%s_%dx%d_nts%d_%dshots" % (tag, w["nz"], w["nx"], nts, nshots)
        out.append(w)
    return out[0], out[1]


def learned_abc_c4(seed=44, scale=1.0):
    """Learned-boundary test model: a seeded layered model (the batched-shot ranges) on a
    1000 x 3000 interior at 10 m, nts=4000 at 1 ms, sub=2; returns (vvr interior (3, nx, nz),
    nts).  The networks are trained on other seeds."""
    rng = np.random.default_rng(seed)
    nz, nx, nts = int(1000 * scale), int(3000 * scale), int(4000 * scale)
    return layered_vvr(rng, nz, nx), nts


def cell_updates(w):
    """(nz-8)(nx-8)(ntw-1) with ntw = (nts-1)*2*sub + 1."""
    ntw = (w["nts"] - 1) * 2 * w["sub"] + 1
    return (w["nz"] - 2 * FAT) * (w["nx"] - 2 * FAT) * (ntw - 1)


def min_pad(w):
    zPad = min(w["zPadMinus"], w["zPadPlus"]) if w["surfaceCondition"] == 0 else w["zPadPlus"]
    return min(zPad, min(w["xPadMinus"], w["xPadPlus"]))
