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
