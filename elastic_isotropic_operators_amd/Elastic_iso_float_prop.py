"""FP32 operator layer with the reference's names (python_float/Elastic_iso_float_prop.py).

nonlinearOpInitFloat, nonlinearFwiOpInitFloat, BornOpInitFloat, buildSourceGeometry,
buildReceiversGeometry and the operator classes nonlinearPropElasticShotsGpu,
nonlinearFwiPropElasticShotsGpu, BornElasticShotsGpu.
"""
from ._prop_common import make as _make

globals().update(_make("float"))
