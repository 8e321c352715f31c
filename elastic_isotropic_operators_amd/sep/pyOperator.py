"""Operator base class (the pySolver pyOperator subset used by the operator wrappers)."""
import time

import numpy as np


class Operator:
    def setDomainRange(self, domain, range):
        self.domain = domain.cloneSpace()
        self.range = range.cloneSpace()

    def getDomain(self):
        return self.domain

    def getRange(self):
        return self.range

    def checkDomainRange(self, x, y):
        if not self.domain.checkSame(x):
            raise ValueError("Provided x vector does not match operator domain")
        if not self.range.checkSame(y):
            raise ValueError("Provided y vector does not match operator range")

    def forward(self, add, model, data):
        raise NotImplementedError

    def adjoint(self, add, model, data):
        raise NotImplementedError

    def dotTest(self, verb=False, maxError=1e-4):
        """<F m, d> = <m, F' d> with and without add (pySolver pyOperator.dotTest)."""
        d1 = self.domain.clone().rand()
        r1 = self.range.clone().rand()
        d2 = self.domain.clone()
        r2 = self.range.clone()
        t0 = time.time()
        self.forward(False, d1, r2)
        t1 = time.time()
        self.adjoint(False, d2, r1)
        t2 = time.time()
        dt1, dt2 = r1.dot(r2), d1.dot(d2)
        err1 = abs((dt1 - dt2) / dt2)
        self.forward(True, d1, r2)
        self.adjoint(True, d2, r1)
        dt3, dt4 = r1.dot(r2), d1.dot(d2)
        err2 = abs((dt3 - dt4) / dt4)
        if verb:
            print("Dot-product test of forward and adjoint operators")
            print("Applying forward operator add=False\n Runs in: %s seconds" % (t1 - t0))
            print("Applying adjoint operator add=False\n Runs in: %s seconds" % (t2 - t1))
            print("Dot products add=False: domain=%s range=%s " % (dt2, dt1))
            print("Absolute error: %s" % abs(dt1 - dt2))
            print("Relative error: %s \n" % err1)
            print("Dot products add=True: domain=%s range=%s " % (dt4, dt3))
            print("Relative error: %s \n" % err2)
        if err1 > maxError or err2 > maxError:
            raise Warning("Dot products failure: relative errors %s %s > %s" % (err1, err2, maxError))
        return err1, err2
