"""ctypes front-end of the CPU oracle (``libeio_oracle.so``, built from eio_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, never by the product package.

Every function mirrors one piece of the reference arithmetic; see the header of
eio_oracle.cpp for the reference file:line each one follows.  Arrays use the
reference layouts in numpy (C order, slowest axis first):
  elastic param   (3, nx, nz)        rho, lambda, mu
  media           (5, nx, nz)        rhoxDtw, rhozDtw, lamb2MuDtw, lambDtw, muxzDtw
  shot model/data (5, nTrace, nts)   vx, vz, sxx, szz, sxz
  Born model      (3, nx, nz)        drho, dlambda, dmu
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libeio_oracle.so")


class OracleGrid(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int), ("nx", ctypes.c_int), ("nts", ctypes.c_int),
                ("sub", ctypes.c_int), ("minPad", ctypes.c_int),
                ("dz", ctypes.c_double), ("dx", ctypes.c_double), ("dts", ctypes.c_double),
                ("alphaCos", ctypes.c_double), ("oz", ctypes.c_double), ("ox", ctypes.c_double),
                ("surface", ctypes.c_int)]


def build():
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
    return _lib


def grid(nz, nx, nts, sub_fd, dz, dx, dts, minPad, alphaCos=0.99, oz=0.0, ox=0.0, surface=0):
    """Grid description; ``sub_fd`` is the fine sub-sampling (twice the user ``sub``);
    surface=1: free surface at row 2*FAT."""
    return OracleGrid(int(nz), int(nx), int(nts), int(sub_fd), int(minPad), float(dz), float(dx),
                      float(dts), float(alphaCos), float(oz), float(ox), int(surface))


def _suffix(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def media(g, elastic, dtype=np.float64):
    el = _c(elastic, dtype)
    out = np.zeros((5, g.nx, g.nz), dtype=dtype)
    getattr(lib(), "oracle_media_" + _suffix(dtype))(ctypes.byref(g), _ptr(el), _ptr(out))
    return out


def consts(g, dtype=np.float64):
    """FD coefficients (z then x), time filter F (2*(sub+1)), damping table (minPad)."""
    coef = np.zeros(8, dtype)
    filt = np.zeros(2 * (g.sub + 1), dtype)
    damp = np.zeros(max(g.minPad, 1), dtype)
    getattr(lib(), "oracle_consts_" + _suffix(dtype))(ctypes.byref(g), _ptr(coef), _ptr(filt),
                                                      _ptr(damp))
    return coef, filt, damp[:g.minPad]


def stagger(which, field, dtype=np.float64):
    """which: 'x_adj', 'z_adj', 'x_fwd', 'z_fwd' on a (nx, nz) field."""
    code = {"x_adj": 0, "z_adj": 1, "x_fwd": 2, "z_fwd": 3}[which]
    f = _c(field, dtype)
    out = np.zeros_like(f)
    nx, nz = f.shape
    getattr(lib(), "oracle_stagger_" + _suffix(dtype))(code, nz, nx, _ptr(f), _ptr(out))
    return out


def time_interp(coarse, dts, sub_fd, dtype=np.float64):
    c = _c(coarse, dtype)
    ndev, nts = c.shape
    ntw = (nts - 1) * sub_fd + 1
    out = np.zeros((ndev, ntw), dtype)
    getattr(lib(), "oracle_time_interp_" + _suffix(dtype))(nts, ctypes.c_double(dts), sub_fd,
                                                           ndev, _ptr(c), _ptr(out))
    return out


_METHODS = {"linear": 0, "sinc": 1, "gauss": 2}


def interp(nz, oz, dz, ox, dx, zc, xc, method="linear", nfilt=1, dipole=0, zshift=0.0, xshift=0.0):
    """Spatial interpolation weights of one staggered grid: (idx[nDev,nf], w[nDev,nf], unique
    positions), nf = 4 (linear), 8 (linear with dipole=1: the stencil of the position shifted
    by (zshift dz, xshift dx) with negated weights as entries 4..7) or (2 nfilt)^2 (sinc,
    gauss)."""
    zc = _c(zc, np.float64)
    xc = _c(xc, np.float64)
    nd = zc.size
    if dipole == 1:
        if method != "linear":
            raise ValueError("dipole acquisition is implemented for linear interpolation only")
        idx = np.zeros((nd, 8), np.int32)
        w = np.zeros((nd, 8), np.float64)
        uniq = np.zeros(8 * nd, np.int32)
        f = lib().oracle_interp_dipole
        f.restype = ctypes.c_int
        nu = f(nz, ctypes.c_double(oz), ctypes.c_double(dz), ctypes.c_double(ox), ctypes.c_double(dx), nd,
               _ptr(zc), _ptr(xc), ctypes.c_double(zshift), ctypes.c_double(xshift), _ptr(idx), _ptr(w),
               _ptr(uniq))
        return idx, w, uniq[:nu].copy()
    m = _METHODS[method]
    nf = 4 if m == 0 else (2 * nfilt) ** 2
    idx = np.zeros((nd, nf), np.int32)
    w = np.zeros((nd, nf), np.float64)
    uniq = np.zeros(nf * nd, np.int32)
    f = lib().oracle_interp_m
    f.restype = ctypes.c_int
    nu = f(nz, ctypes.c_double(oz), ctypes.c_double(dz), ctypes.c_double(ox), ctypes.c_double(dx),
           nd, _ptr(zc), _ptr(xc), int(m), int(nfilt), _ptr(idx), _ptr(w), _ptr(uniq))
    return idx, w, uniq[:nu].copy()


def staggered_interp(g, zc, xc, method="linear", nfilt=1, dipole=0, zshift=0.0, xshift=0.0):
    """interp on the four staggered grids of g (centre, X, Z, XZ)."""
    origins = [(g.oz, g.ox), (g.oz, g.ox - 0.5 * g.dx), (g.oz - 0.5 * g.dz, g.ox),
               (g.oz - 0.5 * g.dz, g.ox - 0.5 * g.dx)]
    return [interp(g.nz, o[0], g.dz, o[1], g.dx, zc, xc, method, nfilt, dipole, zshift, xshift) for o in origins]


def irreg_to_reg(ip, irr):
    """Adjoint of the interpolation: irregular (nDev, nt) -> regular (nUniq, nt)."""
    idx, w, uniq = ip
    pos = {int(p): i for i, p in enumerate(uniq)}
    reg = np.zeros((uniq.size, irr.shape[1]), np.float64)
    for d in range(idx.shape[0]):
        for f in range(idx.shape[1]):
            reg[pos[int(idx[d, f])]] += w[d, f] * irr[d]
    return reg


def reg_to_irreg(ip, reg):
    """Forward of the interpolation: regular (nUniq, nt) -> irregular (nDev, nt)."""
    idx, w, uniq = ip
    pos = {int(p): i for i, p in enumerate(uniq)}
    out = np.zeros((idx.shape[0], reg.shape[1]), np.float64)
    for d in range(idx.shape[0]):
        for f in range(idx.shape[1]):
            out[d] += w[d, f] * reg[pos[int(idx[d, f])]]
    return out


# ------------------------------------------------------------------ regular level
def _pos_arrays(pos4):
    arrs = [np.ascontiguousarray(p, dtype=np.int32) for p in pos4]
    ptrs = (ctypes.c_void_p * 4)(*[a.ctypes.data for a in arrs])
    ns = (ctypes.c_int * 4)(*[a.size for a in arrs])
    return arrs, ptrs, ns


GRID_OF_COMP = (1, 2, 0, 0, 3)  # vx->X, vz->Z, sxx/szz->centre, sxz->XZ


def nl_fwd_reg(g, med, spos, src_traces, rpos, dtype=np.float64, wavefield=False):
    """Regular-level NL forward.  spos/rpos: 4 position arrays (centre, X, Z, XZ);
    src_traces: 5 arrays (nReg, nts).  Returns 5 data arrays (nRegRec, nts), and with
    wavefield=True also the coarse-sampled wavefield (nts, 5, nx, nz)."""
    med = _c(med, dtype)
    sa, sp, sn = _pos_arrays(spos)
    ra, rp, rn = _pos_arrays(rpos)
    mods = [_c(t, dtype) for t in src_traces]
    dats = [np.zeros((ra[GRID_OF_COMP[k]].size, g.nts), dtype) for k in range(5)]
    mp = (ctypes.c_void_p * 5)(*[a.ctypes.data for a in mods])
    dp = (ctypes.c_void_p * 5)(*[a.ctypes.data for a in dats])
    if wavefield:
        w = np.zeros((g.nts, 5, g.nx, g.nz), dtype)
        getattr(lib(), "oracle_nl_fwd_reg_wfld_" + _suffix(dtype))(ctypes.byref(g), _ptr(med), sp, sn,
                                                                   mp, rp, rn, dp, _ptr(w))
        return dats, w
    getattr(lib(), "oracle_nl_fwd_reg_" + _suffix(dtype))(ctypes.byref(g), _ptr(med), sp, sn, mp,
                                                          rp, rn, dp)
    return dats


def nl_adj_reg(g, med, spos, rpos, dat_traces, dtype=np.float64, wavefield=False):
    """Regular-level NL adjoint; with wavefield=True also the adjoint wavefield before any
    scaling (nts, 5, nx, nz)."""
    med = _c(med, dtype)
    sa, sp, sn = _pos_arrays(spos)
    ra, rp, rn = _pos_arrays(rpos)
    dats = [_c(t, dtype) for t in dat_traces]
    mods = [np.zeros((sa[GRID_OF_COMP[k]].size, g.nts), dtype) for k in range(5)]
    mp = (ctypes.c_void_p * 5)(*[a.ctypes.data for a in mods])
    dp = (ctypes.c_void_p * 5)(*[a.ctypes.data for a in dats])
    if wavefield:
        w = np.zeros((g.nts, 5, g.nx, g.nz), dtype)
        getattr(lib(), "oracle_nl_adj_reg_wfld_" + _suffix(dtype))(ctypes.byref(g), _ptr(med), sp, sn,
                                                                   mp, rp, rn, dp, _ptr(w))
        return mods, w
    getattr(lib(), "oracle_nl_adj_reg_" + _suffix(dtype))(ctypes.byref(g), _ptr(med), sp, sn, mp,
                                                          rp, rn, dp)
    return mods


def born_fwd_reg(g, med, spos, src_fine, dm5, rpos, dtype=np.float64):
    med = _c(med, dtype)
    dm5 = _c(dm5, dtype)
    sa, sp, sn = _pos_arrays(spos)
    ra, rp, rn = _pos_arrays(rpos)
    srcs = [_c(t, dtype) for t in src_fine]
    dats = [np.zeros((ra[GRID_OF_COMP[k]].size, g.nts), dtype) for k in range(5)]
    fp = (ctypes.c_void_p * 5)(*[a.ctypes.data for a in srcs])
    dp = (ctypes.c_void_p * 5)(*[a.ctypes.data for a in dats])
    getattr(lib(), "oracle_born_fwd_reg_" + _suffix(dtype))(ctypes.byref(g), _ptr(med), sp, sn, fp,
                                                            _ptr(dm5), rp, rn, dp)
    return dats


def born_adj_reg(g, med, spos, src_fine, rpos, dat_traces, dtype=np.float64):
    med = _c(med, dtype)
    sa, sp, sn = _pos_arrays(spos)
    ra, rp, rn = _pos_arrays(rpos)
    srcs = [_c(t, dtype) for t in src_fine]
    dats = [_c(t, dtype) for t in dat_traces]
    dm5 = np.zeros((5, g.nx, g.nz), dtype)
    fp = (ctypes.c_void_p * 5)(*[a.ctypes.data for a in srcs])
    dp = (ctypes.c_void_p * 5)(*[a.ctypes.data for a in dats])
    getattr(lib(), "oracle_born_adj_reg_" + _suffix(dtype))(ctypes.byref(g), _ptr(med), sp, sn, fp,
                                                            _ptr(dm5), rp, rn, dp)
    return dm5


# ------------------------------------------------------------------ shot level
def _xy(a):
    return _c(np.atleast_1d(a), np.float64)


def nl_fwd_shot(g, elastic, src_z, src_x, rec_z, rec_x, model, dtype=np.float64):
    """One shot: model (5, nSrc, nts) -> data (5, nRec, nts)."""
    el = _c(elastic, dtype)
    sz, sx, rz, rx = _xy(src_z), _xy(src_x), _xy(rec_z), _xy(rec_x)
    mo = _c(model, dtype)
    data = np.zeros((5, rz.size, g.nts), dtype)
    getattr(lib(), "oracle_nl_fwd_shot_" + _suffix(dtype))(
        ctypes.byref(g), _ptr(el), sz.size, _ptr(sz), _ptr(sx), rz.size, _ptr(rz), _ptr(rx),
        _ptr(mo), _ptr(data))
    return data


def nl_adj_shot(g, elastic, src_z, src_x, rec_z, rec_x, data, dtype=np.float64):
    el = _c(elastic, dtype)
    sz, sx, rz, rx = _xy(src_z), _xy(src_x), _xy(rec_z), _xy(rec_x)
    da = _c(data, dtype)
    model = np.zeros((5, sz.size, g.nts), dtype)
    getattr(lib(), "oracle_nl_adj_shot_" + _suffix(dtype))(
        ctypes.byref(g), _ptr(el), sz.size, _ptr(sz), _ptr(sx), rz.size, _ptr(rz), _ptr(rx),
        _ptr(model), _ptr(da))
    return model


def born_fwd_shot(g, elastic, src_z, src_x, src_sig, rec_z, rec_x, dm, dtype=np.float64):
    el = _c(elastic, dtype)
    sz, sx, rz, rx = _xy(src_z), _xy(src_x), _xy(rec_z), _xy(rec_x)
    sig = _c(src_sig, dtype)
    dm = _c(dm, dtype)
    data = np.zeros((5, rz.size, g.nts), dtype)
    getattr(lib(), "oracle_born_fwd_shot_" + _suffix(dtype))(
        ctypes.byref(g), _ptr(el), sz.size, _ptr(sz), _ptr(sx), _ptr(sig), rz.size, _ptr(rz),
        _ptr(rx), _ptr(dm), _ptr(data))
    return data


def born_adj_shots(g, elastic, src_z, src_x, src_sig, rec_z, rec_x, data, dtype=np.float64):
    """Born adjoint stacked over shots.  src_z/src_x: (nShot, nSrc); data (nShot, 5, nRec, nts)."""
    el = _c(elastic, dtype)
    sz = _c(np.atleast_2d(src_z), np.float64)
    sx = _c(np.atleast_2d(src_x), np.float64)
    rz, rx = _xy(rec_z), _xy(rec_x)
    sig = _c(src_sig, dtype)
    da = _c(data, dtype)
    nshot, nsrc = sz.shape
    dm = np.zeros((3, g.nx, g.nz), dtype)
    getattr(lib(), "oracle_born_adj_shots_" + _suffix(dtype))(
        ctypes.byref(g), _ptr(el), nshot, nsrc, _ptr(sz), _ptr(sx), _ptr(sig), rz.size, _ptr(rz),
        _ptr(rx), _ptr(dm), _ptr(da))
    return dm


def free_steps(g, med, older, current, nsteps, dtype=np.float64):
    """`nsteps` fine steps of the NL loop body (step, damping of both levels, swap) without
    sources or receivers from the levels (older, current), each (5, nx, nz); returns the new
    (older, current)."""
    st = np.ascontiguousarray(np.stack([older, current]), dtype=dtype)
    m = _c(med, dtype)
    getattr(lib(), "oracle_free_steps_" + _suffix(dtype))(ctypes.byref(g), _ptr(m), _ptr(st), int(nsteps))
    return st[0], st[1]


def damp_field(g, dtype=np.float64):
    """(nx, nz): the factor the sponge damping applies to each cell (1 outside the sponge and in
    the FAT halo)."""
    out = np.zeros((g.nx, g.nz), dtype)
    getattr(lib(), "oracle_damp_field_" + _suffix(dtype))(ctypes.byref(g), _ptr(out))
    return out


def set_ftz(on=True):
    """Flush denormals to zero on the OpenMP threads (FP32 timing twin)."""
    lib().oracle_set_ftz(int(bool(on)))


def num_threads():
    f = lib().oracle_num_threads
    f.restype = ctypes.c_int
    return int(f())
