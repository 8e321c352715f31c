"""Elastic data component sampler (seis_utils/python/dataCompModule.py): selects vx, vz, sxx,
szz, sxz and the pressure p = 0.5 (sxx + szz) from [time, receiver, 5, shot] data."""
from collections import Counter

from .sep import Hypercube, SepVector, pyOperator
from . import seis_utils as su


class ElasticDatComp(pyOperator.Operator):
    def __init__(self, components, domain):
        hyper = domain.getHyper()
        timeAxis, recAxis, shotAxis = hyper.getAxis(1), hyper.getAxis(2), hyper.getAxis(4)
        self.comp_list = components.split(",")
        if any(v > 1 for v in Counter(self.comp_list).values()):
            raise ValueError("ERROR! A component was provided multiple times! Check your input arguments")
        compAxis = Hypercube.axis(n=len(self.comp_list))
        rng = SepVector.getSepVector(Hypercube.hypercube(axes=[timeAxis, recAxis, compAxis, shotAxis]),
                                     storage="dataFloat" if domain.getNdArray().dtype.itemsize == 4 else "dataDouble")
        self.setDomainRange(domain, rng)
        self.comp_list = [c.lower() for c in self.comp_list]
        if not any(c in self.comp_list for c in ("vx", "vz", "sxx", "szz", "sxz", "p")):
            raise ValueError("ERROR! Provided unknown data components: %s" % components)
        self.components = ",".join(self.comp_list)

    def forward(self, add, model, data):
        self.checkDomainRange(model, data)
        if not add:
            data.zero()
        data.getNdArray()[...] += su.datcomp_fwd(self.components, model.getNdArray())

    def adjoint(self, add, model, data):
        self.checkDomainRange(model, data)
        if not add:
            model.zero()
        m = model.getNdArray()
        m[...] += su.datcomp_adj(self.components, data.getNdArray(), m.shape)
