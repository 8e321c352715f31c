"""Import name used by the reference scripts."""
from elastic_isotropic_operators_amd.sep.genericIO import *  # noqa: F401,F403
from elastic_isotropic_operators_amd.sep import genericIO as _m
globals().update({k: v for k, v in vars(_m).items() if not k.startswith("__")})
