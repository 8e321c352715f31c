"""Minimal stand-ins for the SEP python API subset used by the operator scripts.

The reference builds on the external SEP ``genericIO`` / ``SepVector`` / ``Hypercube`` /
``pyOperator`` packages (un-vendored submodules).  These modules provide the calls the
operator layer and its scripts make (axes, hypercubes, numpy-backed vectors with the SEP
axis order, key=value parameters and parfiles, SEP .H/.H@ reading and writing, the
Operator base with its dot-product test), so the drop-in scripts run unchanged.
"""
