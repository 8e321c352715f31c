"""SEP hypercube subset: axes (n, o, d, label, unit); first axis fastest."""


class axis:
    def __init__(self, n=1, o=0.0, d=1.0, label="", unit="", axisInfo=None):
        if axisInfo is not None:
            n, o, d = axisInfo.n, axisInfo.o, axisInfo.d
        self.n = int(n)
        self.o = float(o)
        self.d = float(d)
        self.label = label
        self.unit = unit

    def __repr__(self):
        return "axis(n=%d, o=%g, d=%g)" % (self.n, self.o, self.d)


class hypercube:
    def __init__(self, axes=None, ns=None, os=None, ds=None):
        if axes is None:
            ns = list(ns)
            os = list(os) if os is not None else [0.0] * len(ns)
            ds = list(ds) if ds is not None else [1.0] * len(ns)
            axes = [axis(n=n, o=o, d=d) for n, o, d in zip(ns, os, ds)]
        self.axes = [axis(n=a.n, o=a.o, d=a.d, label=a.label, unit=a.unit) for a in axes]

    def getNdim(self):
        return len(self.axes)

    def getAxis(self, i):
        """1-based, like the SEP API."""
        return self.axes[i - 1]

    def getNs(self):
        return [a.n for a in self.axes]

    def getN123(self):
        n = 1
        for a in self.axes:
            n *= a.n
        return n

    def getCpp(self):
        return self

    def checkSame(self, other):
        return [a.n for a in self.axes] == [b.n for b in other.axes]

    def __repr__(self):
        return "hypercube(%s)" % self.axes
