"""numpy host preparation for the device-level (C ABI) path.

Used by the torch FWI layer and the benchmark, which keep traces and media resident in
HBM and call the C ABI directly.  The drop-in operator classes use the C++ host layer
(csrc/host) that implements the same rules.

  scaled_media        fdParamElastic.cpp:118-148 (stagger rho and mu, scale by 2*dtw)
  Interp              spaceInterpGpu.cpp:78-207, 288-376 (bilinear/sinc/gauss weights, unique list)
  staggered grids     Elastic_iso_float_prop.py:224-249 (centre, X, Z, XZ origins)
  source scaling      nonlinearPropElasticGpu.cpp:109-144
"""
import numpy as np

GRID_OF_COMP = (1, 2, 0, 0, 3)


def stagger_x_adj(f):
    """staggerX::adjoint (stagger.cpp:38-57) on (..., nx, nz)."""
    out = np.empty_like(f)
    out[..., 1:-1, :] = 0.5 * (f[..., 1:-1, :] + f[..., :-2, :])
    out[..., 0, :] = 0.5 * f[..., 0, :]
    out[..., -1, :] = f[..., -1, :] + 0.5 * f[..., -2, :]
    return out


def stagger_z_adj(f):
    """staggerZ::adjoint (stagger.cpp:95-114)."""
    out = np.empty_like(f)
    out[..., 1:-1] = 0.5 * (f[..., 1:-1] + f[..., :-2])
    out[..., 0] = 0.5 * f[..., 0]
    out[..., -1] = f[..., -1] + 0.5 * f[..., -2]
    return out


def stagger_x_fwd(f):
    """staggerX::forward (stagger.cpp:18-35)."""
    out = np.empty_like(f)
    out[..., :-1, :] = 0.5 * (f[..., :-1, :] + f[..., 1:, :])
    out[..., -1, :] = f[..., -1, :]
    return out


def stagger_z_fwd(f):
    """staggerZ::forward (stagger.cpp:75-93)."""
    out = np.empty_like(f)
    out[..., :-1] = 0.5 * (f[..., :-1] + f[..., 1:])
    out[..., -1] = f[..., -1]
    return out


def dtw_of(dts, sub_user):
    return np.float32(dts) / np.float32(2 * sub_user)


def scaled_media(elastic, dtw, dtype=np.float32):
    """(3, nx, nz) [rho, lambda, mu] -> (5, nx, nz) rhoxDtw, rhozDtw, lamb2MuDtw, lambDtw, muxzDtw."""
    el = np.asarray(elastic, dtype=dtype)
    rho, lam, mu = el[0], el[1], el[2]
    two = dtype(2.0) * dtype(dtw)
    out = np.empty((5,) + rho.shape, dtype=dtype)
    out[0] = two / stagger_x_adj(rho)
    out[1] = two / stagger_z_adj(rho)
    out[2] = two * (lam + dtype(2) * mu)
    out[3] = two * lam
    out[4] = two * stagger_z_adj(stagger_x_adj(mu))
    return out


def _sinc_pi(x):
    """sin(x)/x in double with the Taylor series near 0."""
    x = np.asarray(x, np.float64)
    small = np.abs(x) < 1.2207e-4
    safe = np.where(small, 1.0, x)
    x2 = x * x
    return np.where(small, 1.0 - x2 / 6.0 + x2 * x2 / 120.0, np.sin(safe) / safe)


def _normal_pdf(x, mean, sd):
    return np.exp(-((x - mean) ** 2) / (2.0 * sd * sd)) / (sd * 2.5066282746310002)


class Interp:
    """Spatial interpolation of devices on one staggered grid (spaceInterpGpu.cpp:11-207):
    linear (bilinear, 4 points), or sinc / gauss on the (2 nfilt)^2 points from
    int(position) - nfilt + 1 to int(position) + nfilt, x-major, with separable weights
    normalised to sum 1 (sinc(pi u) sinc(pi v), or normal densities of standard deviation
    2 cells).  Coordinates in float32, kernels in double rounded to float32, sums in float32,
    as there.  The whole stencil is kept for any nfilt (the reference keeps 4 entries per
    device, consistent only for nfilt = 1: spaceInterpGpu.cpp:38-41)."""

    def __init__(self, nz, oz, dz, ox, dx, zc, xc, method="linear", nfilt=1):
        zc = np.atleast_1d(np.asarray(zc, np.float32))
        xc = np.atleast_1d(np.asarray(xc, np.float32))
        oz32, dz32, ox32, dx32 = np.float32(oz), np.float32(dz), np.float32(ox), np.float32(dx)
        wz = (zc - oz32) / dz32
        wx = (xc - ox32) / dx32
        zr = wz.astype(np.int64)
        xr = wx.astype(np.int64)
        if method == "linear":
            wz = np.float32(1.0) - (wz - zr.astype(np.float32))
            wx = np.float32(1.0) - (wx - xr.astype(np.float32))
            b = xr * nz + zr
            self.idx = np.stack([b, b + 1, b + nz, b + nz + 1], axis=1)
            self.w = np.stack([wz * wx, (1 - wz) * wx, wz * (1 - wx), (1 - wz) * (1 - wx)],
                              axis=1).astype(np.float64)
        elif method in ("sinc", "gauss"):
            n2 = 2 * int(nfilt)
            idx, w = [], []
            for ix in range(n2):
                for iz in range(n2):
                    px = xr + ix - nfilt + 1
                    pz = zr + iz - nfilt + 1
                    xcur = px.astype(np.float32) * dx32 + ox32
                    zcur = pz.astype(np.float32) * dz32 + oz32
                    if method == "sinc":
                        xin = (xc - xcur) / dx32
                        zin = (zc - zcur) / dz32
                        v = _sinc_pi(np.pi * zin.astype(np.float64)) * _sinc_pi(np.pi * xin.astype(np.float64))
                    else:
                        v = (_normal_pdf(xcur.astype(np.float64), xc.astype(np.float64), float(np.float32(2) * dx32))
                             * _normal_pdf(zcur.astype(np.float64), zc.astype(np.float64),
                                           float(np.float32(2) * dz32)))
                    idx.append(px * nz + pz)
                    w.append(v.astype(np.float32))
            total = np.zeros(zc.shape, np.float32)
            for v in w:  # sequential float sum, as the reference's loop
                total = (total + v).astype(np.float32)
            self.idx = np.stack(idx, axis=1)
            self.w = np.stack([v / total for v in w], axis=1).astype(np.float64)
        else:
            raise ValueError("**** ERROR: Space interp method not defined ****")
        flat = self.idx.reshape(-1)
        uniq, first = np.unique(flat, return_index=True)
        order = np.argsort(first)               # first-seen order (spaceInterpGpu.cpp:288-311)
        self.uniq = uniq[order].astype(np.int32)
        lookup = {int(p): i for i, p in enumerate(self.uniq)}
        self.map = np.array([lookup[int(p)] for p in flat], np.int64).reshape(self.idx.shape)
        self.n_irreg = zc.size

    def adjoint(self, irr):
        """irregular (nIrreg, nt) -> regular (nUniq, nt)."""
        reg = np.zeros((self.uniq.size, irr.shape[1]), dtype=irr.dtype)
        for f in range(self.idx.shape[1]):
            np.add.at(reg, self.map[:, f], (self.w[:, f:f + 1] * irr).astype(irr.dtype))
        return reg

    def forward(self, reg):
        """regular (nUniq, nt) -> irregular (nIrreg, nt)."""
        out = np.zeros((self.n_irreg, reg.shape[1]), dtype=reg.dtype)
        for f in range(self.idx.shape[1]):
            out += (self.w[:, f:f + 1] * reg[self.map[:, f]]).astype(reg.dtype)
        return out


def staggered_interps(nz, oz, dz, ox, dx, zc, xc, method="linear", nfilt=1):
    """The four grids: centre, X (x-dx/2), Z (z-dz/2), XZ."""
    origins = [(oz, ox), (oz, ox - 0.5 * dx), (oz - 0.5 * dz, ox), (oz - 0.5 * dz, ox - 0.5 * dx)]
    return [Interp(nz, o[0], dz, o[1], dx, zc, xc, method, nfilt) for o in origins]


def parfile_source_coords(par, oz, dz, ox, dx, read_vector=None):
    """Per-shot source coordinates of buildSourceGeometry (Elastic_iso_float_prop.py:261-347):
    from sourceGeomFile ([nExp][nSimSou][x, y, z] in SEP axis order), else one device per
    shot at zSource, xSource + iExp*spacingShots (1-based samples shifted by zPadMinus,
    xPadMinus and fat).  Returns ([(zc, xc) per shot], (n, o, d) of the source axis)."""
    nExp = par.getInt("nExp")
    geom_file = par.getString("sourceGeomFile", "None")
    if geom_file != "None":
        geom = read_vector(geom_file, ndims=3).getNdArray()
        if nExp != geom.shape[2]:
            raise ValueError("ERROR! nExp (%d) not consistent with number of shots provided within "
                             "sourceGeomFile (%d)" % (nExp, geom.shape[2]))
        coords = [(np.array(geom[2, :, e], np.float64), np.array(geom[0, :, e], np.float64))
                  for e in range(nExp)]
        return coords, (nExp, 1.0, 1.0)
    fat = par.getInt("fat", 4)
    ozS = par.getInt("zSource") - 1 + par.getInt("zPadMinus") + fat
    oxS = par.getInt("xSource") - 1 + par.getInt("xPadMinus") + fat
    sp = par.getInt("spacingShots")
    coords = [(np.array([oz + ozS * dz]), np.array([ox + (oxS + e * sp) * dx])) for e in range(nExp)]
    return coords, (nExp, ox + oxS * dx, sp * dx)


def parfile_receiver_coords(par, oz, dz, ox, dx, read_vector=None):
    """Receiver coordinates of buildReceiversGeometry (Elastic_iso_float_prop.py:584-641):
    from recGeomFile, else nReceiver devices at depthReceiver from oReceiver every dReceiver
    samples.  Returns (zc, xc, (n, o, d) of the receiver axis)."""
    geom_file = par.getString("recGeomFile", "none")
    if geom_file != "none":
        geom = read_vector(geom_file, ndims=3).getNdArray()
        xc, zc = np.array(geom[0, :, 0], np.float64), np.array(geom[2, :, 0], np.float64)
        n = par.getInt("nReceiver")
        if n != len(xc):
            raise ValueError("ERROR [buildReceiversGeometry]: Number of receiver coordinates (%s) not "
                             "consistent with parameter file (nReceiver=%d)" % (len(xc), n))
        return zc, xc, (n, 1.0, 1.0)
    fat = par.getInt("fat", 4)
    ozR = par.getInt("depthReceiver") - 1 + par.getInt("zPadMinus") + fat
    n = par.getInt("nReceiver")
    oxR = par.getInt("oReceiver") - 1 + par.getInt("xPadMinus") + fat
    dR = par.getInt("dReceiver")
    irec = np.arange(n)
    return oz + ozR * dz + 0 * irec, ox + oxR * dx + dR * dx * irec, (n, ox + oxR * dx, dR * dx)


def nl_source_traces(interps, media, model, dtw, dx, dz, dtype=np.float32):
    """Irregular model (5, nSrc, nts) -> 5 regular scaled source traces."""
    two = dtype(2.0) * dtype(dtw)
    area = dtype(1.0) / (dtype(dx) * dtype(dz))
    out = []
    for k in range(5):
        ip = interps[GRID_OF_COMP[k]]
        reg = ip.adjoint(np.asarray(model[k], dtype=dtype))
        if k == 0:
            scale = media[0].reshape(-1)[ip.uniq][:, None]
        elif k == 1:
            scale = media[1].reshape(-1)[ip.uniq][:, None]
        else:
            scale = two
        out.append(np.ascontiguousarray((reg * scale) * area, dtype=dtype))
    return out
