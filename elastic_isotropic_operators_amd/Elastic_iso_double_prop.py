"""FP64 operator layer (python_double/Elastic_iso_double_prop.py): same API as the FP32 one
with ...Double init function names, used for dot-product tests at 1e-12."""
from ._prop_common import make as _make

globals().update(_make("double"))
