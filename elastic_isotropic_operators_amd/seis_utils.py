"""Host-side seismic utilities on the caller path of the propagator.

numpy restatements of the small reference utilities that surround the operators:
model padding (padElasticFileGpuMain.cpp), the Ricker wavelet (waveletFloatMain.py),
elastic parameter conversion and its Jacobian (elasticParamConvertModule.py) and the
data-component sampler (dataCompModule.py).  Array layouts follow the reference
numpy views (slowest axis first): elastic models are (3, nx, nz), data are
(nShot, 5, nRec, nts).
"""
import math

import numpy as np


def pad_sizes(nz, nx, zPad, xPad, fat=4, blockSize=16, surfaceCondition=0):
    """Padding rule of seis_utils_float/src/padElasticFileGpuMain.cpp:31-79.

    Returns dict(zPadMinus, zPadPlus, xPadMinus, xPadPlus, nz, nx) where nz/nx
    include padding and the FAT halo (so nz-2*fat is a multiple of blockSize).
    """
    if surfaceCondition == 0:
        nzTotal = zPad * 2 + nz
        nbBlockz = int(math.ceil(np.float32(nzTotal) / np.float32(blockSize)))
        zPadPlus = nbBlockz * blockSize - nz - zPad
        zPadMinus = zPad
    elif surfaceCondition == 1:
        nzTotal = zPad + nz + fat
        nbBlockz = int(math.ceil(np.float32(nzTotal) / np.float32(blockSize)))
        zPadMinus = fat
        zPadPlus = nbBlockz * blockSize - nz - zPadMinus
    else:
        raise ValueError("ERROR UNKNOWN SURFACE CONDITION PARAMETER")
    nxTotal = xPad * 2 + nx
    nbBlockx = int(math.ceil(np.float32(nxTotal) / np.float32(blockSize)))
    xPadPlus = nbBlockx * blockSize - nx - xPad
    return dict(zPadMinus=zPadMinus, zPadPlus=zPadPlus, xPadMinus=xPad, xPadPlus=xPadPlus,
                nz=zPadMinus + zPadPlus + nz + 2 * fat, nx=xPad + xPadPlus + nx + 2 * fat)


def pad_model(model, zPad, xPad, fat=4, blockSize=16, surfaceCondition=0):
    """Edge-replication padding (padElasticFileGpuMain.cpp:96-125).

    model: (nPar, nx, nz) -> (padded (nPar, nxNew, nzNew), sizes dict).
    """
    model = np.asarray(model)
    npar, nx, nz = model.shape
    s = pad_sizes(nz, nx, zPad, xPad, fat, blockSize, surfaceCondition)
    zm, zp, xm, xp = s["zPadMinus"], s["zPadPlus"], s["xPadMinus"], s["xPadPlus"]
    out = np.zeros((npar, s["nx"], s["nz"]), dtype=model.dtype)
    x0, z0 = fat + xm, fat + zm
    out[:, x0:x0 + nx, z0:z0 + nz] = model
    out[:, x0:x0 + nx, :z0] = model[:, :, :1]
    out[:, x0:x0 + nx, z0 + nz:] = model[:, :, nz - 1:nz]
    out[:, :x0, :] = out[:, x0:x0 + 1, :]
    out[:, x0 + nx:, :] = out[:, x0 + nx - 1:x0 + nx, :]
    return out, s


def ricker(nts, dts, fDom, timeDelay=0.0, ots=0.0):
    """Ricker wavelet, waveletFloatMain.py:88-107: (1-2 a t^2) exp(-a t^2), a=(pi f)^2."""
    alpha = (np.pi * fDom) * (np.pi * fDom)
    t = ots + np.arange(nts) * dts - timeDelay
    return ((1.0 - 2.0 * alpha * t * t) * np.exp(-1.0 * alpha * t * t)).astype(np.float32)


def vpvsrho_to_rholamemu(m, conv_type=1):
    """ElasticConv.forward (elasticParamConvertModule.py:23-47); m is (3, nx, nz) [vp, vs, rho]."""
    m = np.asarray(m)
    out = np.zeros_like(m)
    if conv_type == 1:
        out[0] = m[2]
        out[1] = m[2] * (m[0] * m[0] - 2.0 * m[1] * m[1])
        out[2] = m[2] * m[1] * m[1]
    elif conv_type == 2:
        out[0] = m[2] * 1e3
        out[1] = m[2] * (m[0] * m[0] - 2.0 * m[1] * m[1]) * 1e9
        out[2] = m[2] * m[1] * m[1] * 1e9
    elif conv_type == -1:
        out[0] = np.sqrt((m[1] + 2 * m[2]) / m[0])
        out[1] = np.sqrt(m[2] / m[0])
        out[2] = m[0]
    else:
        raise ValueError("ERROR! Unsupported conv_type (current value: %s). Please redefine it!"
                         % conv_type)
    return out


def conv_jacobian_fwd(bg, dm):
    """ElasticConvJab.forward (elasticParamConvertModule.py:72-89) at background bg=[vp,vs,rho]."""
    vp, vs, rho = bg
    out = np.zeros_like(dm)
    out[0] = dm[2]
    out[1] = (vp * vp - 2.0 * vs * vs) * dm[2] - 4.0 * vs * rho * dm[1] + 2.0 * vp * rho * dm[0]
    out[2] = vs * vs * dm[2] + 2.0 * vs * rho * dm[1]
    return out


def conv_jacobian_adj(bg, dd):
    """ElasticConvJab.adjoint (elasticParamConvertModule.py:91-108)."""
    vp, vs, rho = bg
    out = np.zeros_like(dd)
    out[0] = 2.0 * vp * rho * dd[1]
    out[1] = 2.0 * vs * rho * dd[2] - 4.0 * vs * rho * dd[1]
    out[2] = dd[0] + (vp * vp - 2.0 * vs * vs) * dd[1] + vs * vs * dd[2]
    return out


_COMP = {"vx": 0, "vz": 1, "sxx": 2, "szz": 3, "sxz": 4}


def datcomp_fwd(components, data):
    """ElasticDatComp.forward (dataCompModule.py:45-77): data (nShot,5,nRec,nts) -> sampled."""
    comps = [c.lower() for c in components.split(",")]
    out = np.zeros((data.shape[0], len(comps)) + data.shape[2:], dtype=data.dtype)
    for i, c in enumerate(comps):
        if c == "p":
            out[:, i] = 0.5 * (data[:, 2] + data[:, 3])
        else:
            out[:, i] = data[:, _COMP[c]]
    return out


def datcomp_adj(components, sampled, nshape):
    """ElasticDatComp.adjoint (dataCompModule.py:79-107)."""
    comps = [c.lower() for c in components.split(",")]
    out = np.zeros(nshape, dtype=sampled.dtype)
    for i, c in enumerate(comps):
        if c == "p":
            out[:, 2] += 0.5 * sampled[:, i]
            out[:, 3] += 0.5 * sampled[:, i]
        else:
            out[:, _COMP[c]] += sampled[:, i]
    return out
