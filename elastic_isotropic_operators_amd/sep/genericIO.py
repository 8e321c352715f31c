"""genericIO subset: key=value parameters (command line and ``par=`` files) and SEP files.

``io(params=sys.argv)`` collects key=value arguments; ``par=file`` arguments are expanded
in place (later keys win, as in SEP).  ``defaultIO.getVector`` / ``writeVector`` read and
write SEP ``.H`` headers with their binary ``.H@`` files (xdr_float = big-endian float32,
native_float = little-endian float32).
"""
import os
import re
import sys

import numpy as np

from . import Hypercube, SepVector

_TOKEN = re.compile(r'(?<![#\w])([A-Za-z_]\w*)=("[^"]*"|\'[^\']*\'|[^\s#]+)')


def _strip(v):
    if len(v) >= 2 and v[0] == v[-1] and v[0] in "\"'":
        return v[1:-1]
    return v


def parse_parfile(path):
    """key=value pairs of a SEP parfile (``#`` starts a comment)."""
    out = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0]
            for k, v in _TOKEN.findall(line):
                out[k] = _strip(v)
    return out


class io:
    def __init__(self, params=None, **kw):
        self.param = {}
        if params is None:
            params = sys.argv
        if isinstance(params, dict):
            items = ["%s=%s" % (k, v) for k, v in params.items()]
        else:
            items = list(params)
        for a in items:
            if "=" not in str(a):
                continue
            k, v = str(a).split("=", 1)
            if k == "par":
                self.param.update(parse_parfile(_strip(v)))
            else:
                self.param[k] = _strip(v)

    def _get(self, key, default, conv):
        if key in self.param:
            return conv(self.param[key])
        if default is None:
            raise ValueError("ERROR! Parameter %s not provided" % key)
        return default

    def getInt(self, key, default=None):
        return self._get(key, default, lambda s: int(float(s)))

    def getFloat(self, key, default=None):
        return self._get(key, default, float)

    def getString(self, key, default=None):
        return self._get(key, default, str)

    def getBool(self, key, default=None):
        return self._get(key, default, lambda s: s.lower() in ("1", "y", "yes", "true"))

    def getInts(self, key, default=None):
        if key in self.param:
            return [int(x) for x in self.param[key].split(",") if x != ""]
        if default is None:
            raise ValueError("ERROR! Parameter %s not provided" % key)
        return list(default) if isinstance(default, (list, tuple)) else [default]

    def getFloats(self, key, default=None):
        if key in self.param:
            return [float(x) for x in self.param[key].split(",") if x != ""]
        if default is None:
            raise ValueError("ERROR! Parameter %s not provided" % key)
        return list(default) if isinstance(default, (list, tuple)) else [default]

    def getCpp(self):
        return self


def read_header(path):
    keys = {}
    with open(path) as f:
        for line in f:
            for k, v in _TOKEN.findall(line):
                keys[k] = _strip(v)
    return keys


class _DefaultIO:
    def getVector(self, path, ndims=None, storage="dataFloat"):
        h = read_header(path)
        axes = []
        for i in range(1, 10):
            if "n%d" % i not in h:
                break
            axes.append(Hypercube.axis(n=int(h["n%d" % i]), o=float(h.get("o%d" % i, 0.0)),
                                       d=float(h.get("d%d" % i, 1.0)), label=h.get("label%d" % i, "")))
        while len(axes) > 1 and axes[-1].n == 1 and (ndims is None or len(axes) > ndims):
            axes.pop()
        if ndims is not None:
            while len(axes) < ndims:
                axes.append(Hypercube.axis(n=1))
        binpath = h.get("in", path + "@")
        if not os.path.isabs(binpath):
            cand = os.path.join(os.path.dirname(os.path.abspath(path)), binpath)
            binpath = cand if os.path.exists(cand) else binpath
        fmt = h.get("data_format", "xdr_float")
        dt = ">f4" if fmt.startswith("xdr") else "<f4"
        n = int(np.prod([a.n for a in axes]))
        data = np.fromfile(binpath, dtype=dt, count=n).astype(np.float32)
        vec = SepVector.getSepVector(Hypercube.hypercube(axes=axes), storage=storage)
        vec.getNdArray()[...] = data.reshape(vec.getNdArray().shape)
        return vec

    def writeVector(self, path, vec):
        hyper = vec.getHyper()
        binpath = path + "@"
        with open(path, "w") as f:
            for i, a in enumerate(hyper.axes, start=1):
                f.write("n%d=%d o%d=%g d%d=%g\n" % (i, a.n, i, a.o, i, a.d))
            f.write('esize=4\ndata_format="xdr_float"\nin="%s"\n' % os.path.abspath(binpath))
        vec.getNdArray().astype(">f4").tofile(binpath)


defaultIO = _DefaultIO()
