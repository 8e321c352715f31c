"""Import name used by the reference scripts."""
from elastic_isotropic_operators_amd import pyElastic_iso_float_nl as _m
globals().update({k: v for k, v in vars(_m).items() if not k.startswith("__")})
