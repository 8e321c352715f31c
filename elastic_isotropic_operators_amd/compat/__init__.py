"""Top-level import names of the reference (``import Elastic_iso_float_prop``, ``import genericIO``,
``import pyElastic_iso_float_nl`` ...): put this directory on PYTHONPATH to run reference-style
scripts unchanged against this package."""
