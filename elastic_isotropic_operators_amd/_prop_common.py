"""Precision-generic body of Elastic_iso_{float,double}_prop.

Same functions, classes, argument meaning and errors as the reference's python operator
layer (python_float/Elastic_iso_float_prop.py): parfile-driven geometry builders
(:205-362, :525-658), operator init functions (:842-932, :1012-1104, :1155-1252) and
operator wrappers (:935-1009, :1107-1151, :1255-1311).  The compiled modules are the
package's pyElastic_iso_{float,double}_{nl,born}; Dask clients are not supported (one
process drives the single GPU and batches all shots on it).
"""
import sys

import numpy as np

from .sep import Hypercube, SepVector, genericIO, pyOperator as Op
from . import elasticParamConvertModule as ElaConv
from . import geometry as G


def _storage(real):
    return "dataFloat" if real == "float" else "dataDouble"


def make(real):
    if real == "float":
        from . import pyElastic_iso_float_nl as pynl
        from . import pyElastic_iso_float_born as pyborn
    else:
        from . import pyElastic_iso_double_nl as pynl
        from . import pyElastic_iso_double_born as pyborn
    storage = _storage(real)
    spaceInterpGpu = pynl.spaceInterpGpu
    ns = {}

    def _params(args):
        return genericIO.io(params=args if args is not None else sys.argv)

    def _grid_hypers(elasticParam):
        ax = elasticParam.getHyper().axes
        zAxis = Hypercube.axis(n=ax[0].n, o=ax[0].o, d=ax[0].d)
        zAxisShifted = Hypercube.axis(n=ax[0].n, o=ax[0].o - 0.5 * ax[0].d, d=ax[0].d)
        xAxis = Hypercube.axis(n=ax[1].n, o=ax[1].o, d=ax[1].d)
        xAxisShifted = Hypercube.axis(n=ax[1].n, o=ax[1].o - 0.5 * ax[1].d, d=ax[1].d)
        paramAxis = Hypercube.axis(n=ax[2].n, o=ax[2].o, d=ax[2].d) if len(ax) > 2 else Hypercube.axis(n=3)
        return [Hypercube.hypercube(axes=[zAxis, xAxis, paramAxis]),
                Hypercube.hypercube(axes=[zAxis, xAxisShifted, paramAxis]),
                Hypercube.hypercube(axes=[zAxisShifted, xAxis, paramAxis]),
                Hypercube.hypercube(axes=[zAxisShifted, xAxisShifted, paramAxis])]

    def _interp4(zc, xc, hypers, nts, method, nfilt, dipole, zsh, xsh):
        return [spaceInterpGpu(zc, xc, h, nts, method, nfilt, dipole, zsh, xsh) for h in hypers]

    def buildSourceGeometry(parObject, elasticParam):
        """Elastic_iso_float_prop.py:205-362."""
        dipole = parObject.getInt("dipole", 0)
        zDipoleShift = parObject.getFloat("zDipoleShift", 0.0)
        xDipoleShift = parObject.getFloat("xDipoleShift", 0.0)
        nts = parObject.getInt("nts")
        nExp = parObject.getInt("nExp")
        ax = elasticParam.getHyper().axes
        oz, dz, ox, dx = ax[0].o, ax[0].d, ax[1].o, ax[1].d
        hypers = _grid_hypers(elasticParam)
        method = parObject.getString("sourceInterpMethod", "linear")
        nfilt = parObject.getInt("sourceInterpNumFilters", 4)
        lists = [[], [], [], []]
        coords, (n, o, d) = G.parfile_source_coords(parObject, oz, dz, ox, dx, genericIO.defaultIO.getVector)
        sourceAxis = Hypercube.axis(n=n, o=o, d=d)
        for zarr, xarr in coords:
            zc = SepVector.getSepVector(ns=[len(zarr)], storage=storage)
            xc = SepVector.getSepVector(ns=[len(xarr)], storage=storage)
            zc.getNdArray()[:] = zarr
            xc.getNdArray()[:] = xarr
            for l, s in zip(lists, _interp4(zc, xc, hypers, nts, method, nfilt, dipole,
                                           zDipoleShift, xDipoleShift)):
                l.append(s)
        return lists[0], lists[1], lists[2], lists[3], sourceAxis

    def buildReceiversGeometry(parObject, elasticParam):
        """Elastic_iso_float_prop.py:525-658."""
        dipole = parObject.getInt("dipole", 0)
        zDipoleShift = parObject.getFloat("zDipoleShift", 0.0)
        xDipoleShift = parObject.getFloat("xDipoleShift", 0.0)
        nts = parObject.getInt("nts")
        ax = elasticParam.getHyper().axes
        oz, dz, ox, dx = ax[0].o, ax[0].d, ax[1].o, ax[1].d
        hypers = _grid_hypers(elasticParam)
        method = parObject.getString("recInterpMethod", "linear")
        nfilt = parObject.getInt("recInterpNumFilters", 4)
        zarr, xarr, (n, o, d) = G.parfile_receiver_coords(parObject, oz, dz, ox, dx,
                                                          genericIO.defaultIO.getVector)
        receiverAxis = Hypercube.axis(n=n, o=o, d=d)
        zc = SepVector.getSepVector(ns=[len(zarr)], storage=storage)
        xc = SepVector.getSepVector(ns=[len(xarr)], storage=storage)
        zc.getNdArray()[:] = zarr
        xc.getNdArray()[:] = xarr
        grids = _interp4(zc, xc, hypers, nts, method, nfilt, dipole, zDipoleShift, xDipoleShift)
        return [grids[0]], [grids[1]], [grids[2]], [grids[3]], receiverAxis

    def _read_elastic(parObject, convert):
        elasticParam = parObject.getString("elasticParam", "noElasticParamFile")
        if elasticParam == "noElasticParamFile":
            print("**** ERROR: User did not provide elastic parameter file ****\n")
            sys.exit()
        ep = genericIO.defaultIO.getVector(elasticParam, ndims=3, storage=storage)
        mod_par = parObject.getInt("mod_par", 0)
        conv = ep
        if mod_par != 0:
            convOp = ElaConv.ElasticConv(ep, mod_par)
            conv = ep.clone()
            convOp.forward(False, ep, conv)
        return ep, conv

    def nonlinearOpInitFloat(args=None, client=None):
        """Elastic_iso_float_prop.py:842-932."""
        if client:
            raise NotImplementedError("Dask clients are not supported: all shots run batched on one GPU")
        parObject = _params(args)
        nts = parObject.getInt("nts", -1)
        timeAxis = Hypercube.axis(n=nts, o=parObject.getFloat("ots", 0.0), d=parObject.getFloat("dts", -1.0))
        dummyAxis = Hypercube.axis(n=1)
        wavefieldAxis = Hypercube.axis(n=5)
        sourceGeomFile = parObject.getString("sourceGeomFile", "None")
        if sourceGeomFile != "None":
            g = genericIO.defaultIO.getVector(sourceGeomFile, ndims=3)
            modelHyper = Hypercube.hypercube(axes=[timeAxis, g.getHyper().getAxis(2), wavefieldAxis, dummyAxis])
        else:
            modelHyper = Hypercube.hypercube(axes=[timeAxis, dummyAxis, wavefieldAxis, dummyAxis])
        modelFloat = SepVector.getSepVector(modelHyper, storage=storage)
        _, elasticParamFloat = _read_elastic(parObject, True)
        sc, sx, sz, sxz, sourceAxis = buildSourceGeometry(parObject, elasticParamFloat)
        rc, rx, rz, rxz, receiverAxis = buildReceiversGeometry(parObject, elasticParamFloat)
        dataFloat = SepVector.getSepVector(
            Hypercube.hypercube(axes=[timeAxis, receiverAxis, wavefieldAxis, sourceAxis]), storage=storage)
        return (modelFloat, dataFloat, elasticParamFloat, parObject, sc, sx, sz, sxz, rc, rx, rz, rxz,
                modelFloat)

    class nonlinearPropElasticShotsGpu(Op.Operator):
        """Wrapper encapsulating the compiled non-linear propagator (Elastic_iso_float_prop.py:935-1009)."""

        def __init__(self, domain, range, elasticParam, paramP, sourcesVectorCenterGrid, sourcesVectorXGrid,
                     sourcesVectorZGrid, sourcesVectorXZGrid, receiversVectorCenterGrid, receiversVectorXGrid,
                     receiversVectorZGrid, receiversVectorXZGrid):
            self.setDomainRange(domain, range)
            if "getCpp" in dir(elasticParam):
                elasticParam = elasticParam.getCpp()
            if "getCpp" in dir(paramP):
                paramP = paramP.getCpp()
            ax = elasticParam.getHyper().axes
            self._grid_axes = (Hypercube.axis(n=ax[0].n, o=ax[0].o, d=ax[0].d),
                               Hypercube.axis(n=ax[1].n, o=ax[1].o, d=ax[1].d))
            self._dts = paramP.getFloat("dts")
            self._ots = paramP.getFloat("ots", 0.0)
            self.pyOp = pynl.nonlinearPropElasticShotsGpu(
                elasticParam, paramP.param, sourcesVectorCenterGrid, sourcesVectorXGrid, sourcesVectorZGrid,
                sourcesVectorXZGrid, receiversVectorCenterGrid, receiversVectorXGrid, receiversVectorZGrid,
                receiversVectorXZGrid)

        def __str__(self):
            return " NLOper "

        def forward(self, add, model, data):
            if "getCpp" in dir(model):
                model = model.getCpp()
            if "getCpp" in dir(data):
                data = data.getCpp()
            with pynl.ostream_redirect():
                self.pyOp.forward(add, model, data)

        def adjoint(self, add, model, data):
            if "getCpp" in dir(model):
                model = model.getCpp()
            if "getCpp" in dir(data):
                data = data.getCpp()
            with pynl.ostream_redirect():
                self.pyOp.adjoint(add, model, data)

        def setBackground(self, elasticParam):
            if "getCpp" in dir(elasticParam):
                elasticParam = elasticParam.getCpp()
            with pynl.ostream_redirect():
                self.pyOp.setBackground(elasticParam)

        def forwardWavefield(self, add, model, data):
            """Forward, keeping the wavefield of shot wavefieldShotNumber when saveWavefield=1
            (Elastic_iso_float_prop.py:971-979)."""
            if "getCpp" in dir(model):
                model = model.getCpp()
            if "getCpp" in dir(data):
                data = data.getCpp()
            with pynl.ostream_redirect():
                self.pyOp.forwardWavefield(add, model, data)

        def adjointWavefield(self, add, model, data):
            """Adjoint, keeping the scaled adjoint wavefield (Elastic_iso_float_prop.py:991-999)."""
            if "getCpp" in dir(model):
                model = model.getCpp()
            if "getCpp" in dir(data):
                data = data.getCpp()
            with pynl.ostream_redirect():
                self.pyOp.adjointWavefield(add, model, data)

        def getWavefield(self):
            """The saved wavefield as a float4D vector with axes z, x, component (5), time
            (Elastic_iso_float_prop.py:1001-1003; seismicElasticOperator2D.cpp:174-191)."""
            w = self.pyOp.getWavefield()
            hz, hx = self._grid_axes
            hyper = Hypercube.hypercube(axes=[hz, hx, Hypercube.axis(n=5, o=0.0, d=1.0),
                                              Hypercube.axis(n=w.shape[0], o=self._ots, d=self._dts)])
            vec = SepVector.getSepVector(hyper, storage=_storage(real))
            vec.getNdArray()[...] = w
            return vec

        def dotTestCpp(self, verb=False, maxError=.00001):
            with pynl.ostream_redirect():
                return self.pyOp.dotTest(verb, maxError)

    def nonlinearFwiOpInitFloat(args=None, client=None):
        """Elastic_iso_float_prop.py:1012-1104."""
        if client:
            raise NotImplementedError("Dask clients are not supported: all shots run batched on one GPU")
        parObject = _params(args)
        elasticParamFloat, elasticParamFloatConv = _read_elastic(parObject, True)
        nts = parObject.getInt("nts", -1)
        timeAxis = Hypercube.axis(n=nts, o=parObject.getFloat("ots", 0.0), d=parObject.getFloat("dts", -1.0))
        dummyAxis = Hypercube.axis(n=1)
        wavefieldAxis = Hypercube.axis(n=5)
        sourceFloat = SepVector.getSepVector(
            Hypercube.hypercube(axes=[timeAxis, dummyAxis, wavefieldAxis, dummyAxis]), storage=storage)
        sourcesFile = parObject.getString("sources", "noSourcesFile")
        if sourcesFile == "noSourcesFile":
            raise IOError("**** ERROR: User did not provide seismic sources file ****")
        sourcesTemp = genericIO.defaultIO.getVector(sourcesFile, ndims=3)
        src = sourcesTemp.getNdArray()
        # the reference's FWI inputs hold [nts, 5]; Born-style [nts, 1, 5] files have the same samples
        sourceFloat.getNdArray()[0, :, 0, :] = src.reshape(5, nts) if src.size == 5 * nts else src
        sc, sx, sz, sxz, sourceAxis = buildSourceGeometry(parObject, elasticParamFloat)
        rc, rx, rz, rxz, receiverAxis = buildReceiversGeometry(parObject, elasticParamFloat)
        dataFloat = SepVector.getSepVector(
            Hypercube.hypercube(axes=[timeAxis, receiverAxis, wavefieldAxis, sourceAxis]), storage=storage)
        return (elasticParamFloat, elasticParamFloatConv, dataFloat, sourceFloat, parObject, sc, sx, sz, sxz,
                rc, rx, rz, rxz, elasticParamFloat)

    class nonlinearFwiPropElasticShotsGpu(Op.Operator):
        """Non-linear operator of the elastic parameters (Elastic_iso_float_prop.py:1107-1151)."""

        def __init__(self, domain, range, sources, paramP, sourcesVectorCenterGrid, sourcesVectorXGrid,
                     sourcesVectorZGrid, sourcesVectorXZGrid, receiversVectorCenterGrid, receiversVectorXGrid,
                     receiversVectorZGrid, receiversVectorXZGrid):
            self.setDomainRange(domain, range)
            if "getCpp" in dir(domain):
                domain = domain.getCpp()
            if "getCpp" in dir(paramP):
                paramP = paramP.getCpp()
            if "getCpp" in dir(sources):
                sources = sources.getCpp()
            self.sources = sources.clone()
            self.pyOp = pynl.nonlinearPropElasticShotsGpu(
                domain, paramP.param, sourcesVectorCenterGrid, sourcesVectorXGrid, sourcesVectorZGrid,
                sourcesVectorXZGrid, receiversVectorCenterGrid, receiversVectorXGrid, receiversVectorZGrid,
                receiversVectorXZGrid)

        def __str__(self):
            return " NLOper "

        def forward(self, add, model, data):
            self.setBackground(model)
            if "getCpp" in dir(data):
                data = data.getCpp()
            with pynl.ostream_redirect():
                self.pyOp.forward(add, self.sources, data)

        def setBackground(self, elasticParam):
            if "getCpp" in dir(elasticParam):
                elasticParam = elasticParam.getCpp()
            with pynl.ostream_redirect():
                self.pyOp.setBackground(elasticParam)

    def BornOpInitFloat(args=None, client=None):
        """Elastic_iso_float_prop.py:1155-1252."""
        if client:
            raise NotImplementedError("Dask clients are not supported: all shots run batched on one GPU")
        parObject = _params(args)
        _, elasticParamFloat = _read_elastic(parObject, True)
        nts = parObject.getInt("nts", -1)
        timeAxis = Hypercube.axis(n=nts, o=parObject.getFloat("ots", 0.0), d=parObject.getFloat("dts", -1.0))
        wavefieldAxis = Hypercube.axis(n=5)
        sourcesFile = parObject.getString("sources", "noSourcesFile")
        if sourcesFile == "noSourcesFile":
            raise IOError("**** ERROR: User did not provide seismic sources file ****")
        sourcesSignalsFloat = genericIO.defaultIO.getVector(sourcesFile, ndims=3, storage=storage)
        sourcesSignalsVector = [sourcesSignalsFloat]
        modelFloat = SepVector.getSepVector(elasticParamFloat.getHyper(), storage=storage)
        sc, sx, sz, sxz, sourceAxis = buildSourceGeometry(parObject, elasticParamFloat)
        rc, rx, rz, rxz, receiverAxis = buildReceiversGeometry(parObject, elasticParamFloat)
        dataFloat = SepVector.getSepVector(
            Hypercube.hypercube(axes=[timeAxis, receiverAxis, wavefieldAxis, sourceAxis]), storage=storage)
        return (modelFloat, dataFloat, elasticParamFloat, parObject, sourcesSignalsVector, sc, sx, sz, sxz,
                rc, rx, rz, rxz, modelFloat)

    class BornElasticShotsGpu(Op.Operator):
        """Wrapper encapsulating the compiled elastic Born propagator (Elastic_iso_float_prop.py:1255-1311)."""

        def __init__(self, domain, range, elasticParam, paramP, sourcesSignalsVector, sourcesVectorCenterGrid,
                     sourcesVectorXGrid, sourcesVectorZGrid, sourcesVectorXZGrid, receiversVectorCenterGrid,
                     receiversVectorXGrid, receiversVectorZGrid, receiversVectorXZGrid):
            self.setDomainRange(domain, range)
            if "getCpp" in dir(elasticParam):
                elasticParam = elasticParam.getCpp()
            if "getCpp" in dir(paramP):
                paramP = paramP.getCpp()
            sourcesSignalsVector = [s.getCpp() if "getCpp" in dir(s) else s for s in sourcesSignalsVector]
            self.pyOp = pyborn.BornElasticShotsGpu(
                elasticParam, paramP.param, sourcesSignalsVector, sourcesVectorCenterGrid, sourcesVectorXGrid,
                sourcesVectorZGrid, sourcesVectorXZGrid, receiversVectorCenterGrid, receiversVectorXGrid,
                receiversVectorZGrid, receiversVectorXZGrid)

        def __str__(self):
            return " BornOp "

        def forward(self, add, model, data):
            if "getCpp" in dir(model):
                model = model.getCpp()
            if "getCpp" in dir(data):
                data = data.getCpp()
            with pynl.ostream_redirect():
                self.pyOp.forward(add, model, data)

        def adjoint(self, add, model, data):
            if "getCpp" in dir(model):
                model = model.getCpp()
            if "getCpp" in dir(data):
                data = data.getCpp()
            with pynl.ostream_redirect():
                self.pyOp.adjoint(add, model, data)

        def setBackground(self, elasticParam):
            if "getCpp" in dir(elasticParam):
                elasticParam = elasticParam.getCpp()
            with pynl.ostream_redirect():
                self.pyOp.setBackground(elasticParam)

    suffix = "Float" if real == "float" else "Double"
    ns.update(dict(
        spaceInterpGpu=spaceInterpGpu,
        buildSourceGeometry=buildSourceGeometry,
        buildReceiversGeometry=buildReceiversGeometry,
        nonlinearPropElasticShotsGpu=nonlinearPropElasticShotsGpu,
        nonlinearFwiPropElasticShotsGpu=nonlinearFwiPropElasticShotsGpu,
        BornElasticShotsGpu=BornElasticShotsGpu,
    ))
    ns["nonlinearOpInit" + suffix] = nonlinearOpInitFloat
    ns["nonlinearFwiOpInit" + suffix] = nonlinearFwiOpInitFloat
    ns["BornOpInit" + suffix] = BornOpInitFloat
    return ns
