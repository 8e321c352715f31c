"""Elastic FWI driver with the parameters and files of the reference's efwiMain.py.

    python -m elastic_isotropic_operators_amd.efwi elasticParam=init.H par=par.p data=data.H \\
        sources=wavelet.H nIter=20 folder=fwi solver=lbfgs comp=vx,vz,p mod_par=1

Reads what efwiMain.py reads (efwiMain.py:1-88 documents the parameters): the padded initial
model (SEP [z, x, 3]), the source wavelet ([t, 5] or [t, 1, 5]), the observed data
([t, rec, comp, shot] as dataCompMain writes it), the acquisition parameters of the NL
operator (zSource/xSource/spacingShots or sourceGeomFile, nReceiver/depthReceiver/
oReceiver/dReceiver or recGeomFile), maskGradient, dataMaskFile, minBound_par1..3 /
maxBound_par1..3 or minBoundVector / maxBoundVector, solver (lbfgs | nlcg | sd), stepper,
initStep, illumination (L-BFGS initial inverse Hessian), sourceInterpMethod/recInterpMethod
(linear, sinc, gauss) with their NumFilters, nIter, folder, prefix, saveObj,
saveRes, saveGrad, saveModel, iterSampling.  Every shot runs batched on one GPU with the
checkpointed gradient (fwi.ElasticFwi); the problem and solvers are fwi_problem's.

Written files (SEP .H + .H@ xdr_float), under folder/prefix (prefix defaults to the folder
name, inversionUtils.py:21-26):
  _obj.H    the objective at the initial model and after each iteration
  _model.H  [z, x, 3, k]: the model every iterSampling iterations and the last one
  _grad.H   [z, x, 3, k]: the gradient at the same iterations
  _res.H    [t, rec, comp, shot]: the residual of the last saved iteration
  _logFile  one line per iteration

Not available: Dask clients (one GPU runs every shot), spline=1 (the B-spline model
operator, interpBSplineModule, is not part of the reference repository), the free surface
in the gradient (the Born operator refuses surfaceCondition=1).
"""
import os
import sys
import time

import numpy as np


def _read(path, ndims):
    from .sep import genericIO
    return genericIO.defaultIO.getVector(path, ndims=ndims)


def _write(path, arr, axes):
    """arr in numpy order (last SEP axis first); axes [(n, o, d)] in SEP order."""
    from .sep import Hypercube, SepVector, genericIO
    hyper = Hypercube.hypercube(axes=[Hypercube.axis(n=n, o=o, d=d) for n, o, d in axes])
    vec = SepVector.getSepVector(hyper)
    vec.getNdArray()[...] = np.asarray(arr, np.float32).reshape(vec.getNdArray().shape)
    genericIO.defaultIO.writeVector(path, vec)


def _bounds(par, shape, which):
    """createBoundVectors (efwiMain.py:118-174): a bound vector file, or per-parameter
    constants, or None."""
    f = par.getString(which + "BoundVector", "no%sBoundVectorFile" % which.capitalize())
    if f != "no%sBoundVectorFile" % which.capitalize():
        return _read(f, 3).getNdArray().astype(np.float64)
    inf = -np.inf if which == "min" else np.inf
    vals = [par.getFloat("%sBound_par%d" % (which, k), inf) for k in (1, 2, 3)]
    if all(v == inf for v in vals):
        return None
    out = np.empty(shape)
    for k in range(3):
        out[k] = vals[k]
    return out


def build_problem(par, device=0):
    """The engine, the problem and the starting model from an efwiMain-style parameter set."""
    import torch
    from . import geometry as G
    from .fwi import ElasticFwi
    from .fwi_problem import FwiProblem

    if par.getInt("spline", 0) == 1:
        raise NotImplementedError("spline=1 needs interpBSplineModule, which is not part of the reference "
                                  "repository and is not available in this build")
    if par.getInt("surfaceCondition", 0) != 0:
        raise RuntimeError("surfaceCondition=1 (free surface) is available for the nonlinear operator only")
    ep = _read(par.getString("elasticParam"), 3)
    ax = ep.getHyper().axes
    model0 = ep.getNdArray().astype(np.float64)                     # (3, nx, nz)
    nz, oz, dz = ax[0].n, ax[0].o, ax[0].d
    nx, ox, dx = ax[1].n, ax[1].o, ax[1].d
    nts, dts, sub = par.getInt("nts"), par.getFloat("dts"), par.getInt("sub")
    zPad = min(par.getInt("zPadMinus"), par.getInt("zPadPlus"))
    xPad = min(par.getInt("xPadMinus"), par.getInt("xPadPlus"))
    grid = dict(nz=nz, nx=nx, dz=dz, dx=dx, oz=oz, ox=ox, nts=nts, dts=dts, sub=sub,
                minPad=min(zPad, xPad), alphaCos=par.getFloat("alphaCos", 0.99))
    coords, _ = G.parfile_source_coords(par, oz, dz, ox, dx, _read)
    nsrc = {len(c[0]) for c in coords}
    if len(nsrc) != 1:
        raise ValueError("every shot must have the same number of sources")
    src_z = np.stack([c[0] for c in coords])
    src_x = np.stack([c[1] for c in coords])
    rec_z, rec_x, _ = G.parfile_receiver_coords(par, oz, dz, ox, dx, _read)
    sig = _read(par.getString("sources"), 3).getNdArray().astype(np.float64)
    if sig.size == 5 * nts:
        sig = np.broadcast_to(sig.reshape(5, 1, nts), (5, src_z.shape[1], nts))
    elif sig.shape != (5, src_z.shape[1], nts):
        raise ValueError("sources must hold [nts, 5] or [nts, nSrc, 5] samples")
    data_file = par.getString("data", "noDataFile")
    if data_file == "noDataFile":
        raise IOError("**** ERROR: User did not provide data file (data) ****\n")
    d_obs = _read(data_file, 4).getNdArray().astype(np.float64)     # (nShot, nComp, nRec, nts)
    comp = par.getString("comp")
    components = None if comp == "vx,vz,sxx,szz,sxz" else comp
    ncomp = 5 if components is None else len(comp.split(","))
    if d_obs.shape != (src_z.shape[0], ncomp, rec_z.size, nts):
        raise ValueError("ERROR! The input data have different size of the expected inversion data! Check "
                         "your arguments and paramater file")
    gm = par.getString("maskGradient", "NoMask")
    grad_mask = None if gm == "NoMask" else _read(gm, 3).getNdArray().astype(np.float64)
    dm = par.getString("dataMaskFile", "noDataMask")
    data_mask = None if dm == "noDataMask" else _read(dm, 4).getNdArray().astype(np.float64)
    dtype = torch.float64 if par.getString("precision", "float") == "double" else torch.float32
    if par.getInt("dipole", 0) != 0:
        raise NotImplementedError("dipole=1 is available in the operator classes, not in the FWI driver")
    src_interp = (par.getString("sourceInterpMethod", "linear"), par.getInt("sourceInterpNumFilters", 4))
    rec_interp = (par.getString("recInterpMethod", "linear"), par.getInt("recInterpNumFilters", 4))
    eng = ElasticFwi(grid, src_z, src_x, sig, rec_z, rec_x, device=device, dtype=dtype,
                     segment=par.getInt("checkpointInterval", 0),
                     src_interp=(src_interp[0], 1 if src_interp[0] == "linear" else src_interp[1]),
                     rec_interp=(rec_interp[0], 1 if rec_interp[0] == "linear" else rec_interp[1]))
    prob = FwiProblem(eng, d_obs, components=components, mod_par=par.getInt("mod_par", 0),
                      grad_mask=grad_mask, data_mask=data_mask,
                      min_bound=_bounds(par, model0.shape, "min"), max_bound=_bounds(par, model0.shape, "max"),
                      dtw=eng.dtw, dz=dz, dx=dx)
    m0 = torch.tensor(model0, dtype=dtype, device=eng.device)
    return eng, prob, m0, ep.getHyper().axes


def main(argv=None):
    import torch  # before the propagator library: both share torch's HIP runtime
    from .sep import genericIO
    from .fwi_problem import solve

    par = genericIO.io(params=argv if argv is not None else sys.argv)
    n_iter = par.getInt("nIter")
    folder = par.getString("folder")
    os.makedirs(folder, exist_ok=True)
    prefix = par.getString("prefix", "None")
    inv = os.path.join(folder, os.path.basename(folder.rstrip("/")) if prefix == "None" else prefix)
    every = max(1, par.getInt("iterSampling", 10))
    save = {k: par.getInt(k, 1) for k in ("saveObj", "saveRes", "saveGrad", "saveModel")}
    pyinfo = par.getInt("pyinfo", 1)
    solver = par.getString("solver")
    stepper = par.getString("stepper", "default")

    eng, prob, m0, axes = build_problem(par)
    model_axes = [(a.n, a.o, a.d) for a in axes[:2]] + [(3, 0.0, 1.0)]
    kw = {}
    init_step = par.getFloat("initStep", -1.0)
    if init_step > 0:
        kw["init_step"] = init_step
    if solver == "lbfgs":
        kw["stepper"] = "backtracking" if stepper in ("default", "linear") else stepper
        illum = par.getString("illumination", "noIllum")
        if illum != "noIllum":
            kw["h0"] = torch.tensor(_read(illum, 3).getNdArray(), dtype=m0.dtype, device=m0.device)
    else:
        kw["stepper"] = "parabolic" if stepper == "default" else stepper
    models, grads, objs, saved = [], [], [], []
    log = open(inv + "_logFile", "w")
    t0 = time.time()

    def save_iterate(it, m, g, r):
        saved.append(it)
        models.append(m.double().cpu().numpy())
        grads.append(g.double().cpu().numpy())
        k = len(models)
        if save["saveModel"]:
            _write(inv + "_model.H", np.stack(models), model_axes + [(k, 0.0, 1.0)])
        if save["saveGrad"]:
            _write(inv + "_grad.H", np.stack(grads), model_axes + [(k, 0.0, 1.0)])
        if save["saveRes"] and r is not None:
            rr = r.double().cpu().numpy()
            _write(inv + "_res.H", rr, [(rr.shape[3], par.getFloat("ots", 0.0), par.getFloat("dts")),
                                        (rr.shape[2], 0.0, 1.0), (rr.shape[1], 0.0, 1.0),
                                        (rr.shape[0], 0.0, 1.0)])

    def record(it, m, phi, g, r):
        objs.append(phi)
        line = "iter %d objective %.9e time %.1f s" % (it, phi, time.time() - t0)
        log.write(line + "\n")
        log.flush()
        if pyinfo:
            print(line, flush=True)
        if it % every == 0 or it == n_iter:
            save_iterate(it, m, g, r)
        if save["saveObj"]:
            _write(inv + "_obj.H", np.array(objs), [(len(objs), 0.0, 1.0)])

    res = solve(prob, m0, n_iter, solver=solver, callback=record, **kw)
    if res.iterations not in saved:  # stopped early between two saved iterations
        save_iterate(res.iterations, res.model, res.gradient, None)
    msg = "stopped: %s after %d iterations; %d forwards, %d gradients, %.1f s" % (
        res.stop_reason, res.iterations, res.n_forward, res.n_gradient, res.seconds)
    log.write(msg + "\n")
    log.close()
    if pyinfo:
        print(msg)
    return res


if __name__ == "__main__":
    main()
