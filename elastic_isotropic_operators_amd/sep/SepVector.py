"""numpy-backed SEP vectors.

getNdArray() returns the storage with the SEP axis order reversed (slowest axis first),
as the reference scripts index it (e.g. ``model.getNdArray()[2, :, :]`` is the third
elastic parameter).  getCpp() returns the vector itself: the compiled operator modules
read the numpy storage in place.
"""
import numpy as np

from . import Hypercube


class vector:
    def __init__(self, hyper, arr=None, dtype=np.float32):
        self.hyper = hyper if isinstance(hyper, Hypercube.hypercube) else Hypercube.hypercube(axes=hyper)
        shape = tuple(a.n for a in reversed(self.hyper.axes))
        if arr is None:
            arr = np.zeros(shape, dtype=dtype)
        else:
            arr = np.asarray(arr, dtype=dtype)
            if arr.shape != shape:
                arr = arr.reshape(shape)
        self.arr = np.ascontiguousarray(arr)

    # -- SEP API subset
    def getNdArray(self):
        return self.arr

    def getHyper(self):
        return self.hyper

    def getCpp(self):
        return self

    def clone(self):
        return self.__class__(self.hyper, self.arr.copy(), self.arr.dtype)

    def cloneSpace(self):
        return self.__class__(self.hyper, None, self.arr.dtype)

    def checkSame(self, other):
        return self.arr.shape == other.getNdArray().shape

    def zero(self):
        self.arr[...] = 0
        return self

    def set(self, val):
        self.arr[...] = val
        return self

    def scale(self, sc):
        self.arr *= sc
        return self

    def rand(self, seed=None):
        rng = np.random.default_rng(seed)
        self.arr[...] = rng.uniform(-1.0, 1.0, self.arr.shape)
        return self

    random = rand

    def copy(self, other):
        self.arr[...] = other.getNdArray()
        return self

    def scaleAdd(self, other, sc1=1.0, sc2=1.0):
        self.arr[...] = sc1 * self.arr + sc2 * other.getNdArray()
        return self

    def add(self, other):
        self.arr += other.getNdArray()
        return self

    def multiply(self, other):
        self.arr *= other.getNdArray()
        return self

    def dot(self, other):
        return float(np.sum(self.arr.astype(np.float64) * other.getNdArray().astype(np.float64)))

    def norm(self, N=2):
        return float(np.linalg.norm(self.arr.astype(np.float64).ravel(), N))

    def max(self):
        return float(self.arr.max())

    def min(self):
        return float(self.arr.min())


class floatVector(vector):
    def __init__(self, hyper=None, arr=None, dtype=np.float32, fromCpp=None):
        if fromCpp is not None:
            hyper, arr = fromCpp.getHyper(), fromCpp.getNdArray()
        super().__init__(hyper, arr, np.float32)


class doubleVector(vector):
    def __init__(self, hyper=None, arr=None, dtype=np.float64, fromCpp=None):
        if fromCpp is not None:
            hyper, arr = fromCpp.getHyper(), fromCpp.getNdArray()
        super().__init__(hyper, arr, np.float64)


def getSepVector(hyper=None, ns=None, storage="dataFloat", axes=None, **kw):
    if hyper is None:
        if axes is not None:
            hyper = Hypercube.hypercube(axes=axes)
        else:
            hyper = Hypercube.hypercube(ns=ns)
    if storage == "dataDouble":
        return doubleVector(hyper)
    if storage != "dataFloat":
        raise ValueError("unsupported storage %s" % storage)
    return floatVector(hyper)
