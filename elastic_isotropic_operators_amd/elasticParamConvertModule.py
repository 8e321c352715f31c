"""Elastic parameter conversion operators (seis_utils/python/elasticParamConvertModule.py).

ElasticConv: VpVsRho <-> RhoLameMu (conv_type 1, 2, -1); ElasticConvJab: its Jacobian at a
background model with forward/adjoint.  Vectors are SEP-style (axis order reversed in numpy).
"""
from .sep import pyOperator
from . import seis_utils as su


class ElasticConv(pyOperator.Operator):
    def __init__(self, domain, conv_type):
        self.setDomainRange(domain, domain)
        self.conv_type = conv_type

    def forward(self, add, model, data):
        self.checkDomainRange(model, data)
        if not add:
            data.zero()
        out = su.vpvsrho_to_rholamemu(model.getNdArray(), self.conv_type)
        data.getNdArray()[...] += out.astype(data.getNdArray().dtype)


class ElasticConvJab(pyOperator.Operator):
    def __init__(self, domain, background, conv_type):
        self.setDomainRange(domain, domain)
        self.background = domain.clone()
        self.conv_type = conv_type
        if background is not None:
            self.setBackground(background)

    def forward(self, add, model, data):
        self.checkDomainRange(model, data)
        if self.conv_type not in (1, 2):
            raise ValueError("ERROR! Unsupported conv_type (current value: %s). Please redefine it!" % self.conv_type)
        if not add:
            data.zero()
        data.getNdArray()[...] += su.conv_jacobian_fwd(self.background.getNdArray(), model.getNdArray())

    def adjoint(self, add, model, data):
        self.checkDomainRange(model, data)
        if self.conv_type not in (1, 2):
            raise ValueError("ERROR! Unsupported conv_type (current value: %s). Please redefine it!" % self.conv_type)
        if not add:
            model.zero()
        model.getNdArray()[...] += su.conv_jacobian_adj(self.background.getNdArray(), data.getNdArray())

    def setBackground(self, background):
        self.background.copy(background)
        if self.conv_type == 2:
            self.background.getNdArray()[:] *= 1e3
