"""FWI problem layer and solvers on the CPU.

The problem layer (objective, masks, bounds, component selection, parameter conversion chain)
is checked with an analytic torch engine in place of the GPU propagator: its gradient must
equal autograd's gradient of the same objective, so a dropped mask, a transposed component
map or a wrong Jacobian term fails.  The solvers are checked on problems whose minimiser is
known in closed form.  The GPU engine itself is covered by the -m gpu tests."""
import math

import numpy as np
import pytest
import torch

from elastic_isotropic_operators_amd import fwi_problem as P


class _Store:
    def __init__(self, lame):
        self.lame = lame
        self.released = False

    def release(self):
        self.released = True


class AnalyticEngine:
    """data[e, c, r, t] = tanh(sum_k W[e, c, r, t, k] * feature_k(lame)), a smooth nonlinear
    stand-in for F; its gradient is autograd's."""

    def __init__(self, nx=6, nz=5, nshot=2, nrec=3, nts=4, seed=0):
        g = torch.Generator().manual_seed(seed)
        self.device = torch.device("cpu")
        self.dtype = torch.float64
        self.shape = (3, nx, nz)
        self.W = torch.randn(nshot, 5, nrec, nts, 3 * nx * nz, generator=g, dtype=torch.float64) * 0.3
        self.scale = torch.tensor([2000.0, 8e9, 4e9], dtype=torch.float64)[:, None, None]
        self.n_store = 0

    def _f(self, lame):
        x = torch.log(lame / self.scale).reshape(-1)
        return torch.tanh(self.W @ x)

    def forward_data(self, lame, checkpoint=False):
        store = None
        if checkpoint:
            store = _Store(lame.detach().clone())
            self.n_store += 1
        return self._f(lame.detach()), store, None

    def gradient(self, store, media, grad_data):
        lame = store.lame.clone().requires_grad_(True)
        d = self._f(lame)
        (g,) = torch.autograd.grad(d, lame, grad_outputs=grad_data)
        return g


def _vvr_model(nx=6, nz=5, seed=1):
    rng = np.random.default_rng(seed)
    vp = 2500.0 + 200.0 * rng.standard_normal((nx, nz))
    vs = vp / 1.8
    rho = 2100.0 + 50.0 * rng.standard_normal((nx, nz))
    return torch.tensor(np.stack([vp, vs, rho]), dtype=torch.float64)


def _autograd_gradient(eng, m, mod_par, d_obs, components, data_mask):
    """Gradient of 1/2||W_d C F(g(m)) - W_d d_obs||^2 by autograd (exact derivative)."""
    m = m.clone().requires_grad_(True)
    d = eng._f(P.elastic_conv(m, mod_par))
    if components:
        d = P.datcomp_forward(d, components)
    if data_mask is not None:
        d = d * data_mask
        d_obs = d_obs * data_mask
    r = d - d_obs
    phi = 0.5 * torch.sum(r * r)
    (g,) = torch.autograd.grad(phi, m)
    return float(phi.detach()), g


@pytest.mark.parametrize("mod_par,components,masked", [(0, None, False), (1, "vx,vz,p", True),
                                                        (1, "sxz,p,vz", False), (0, "p", True)])
def test_problem_gradient_equals_autograd(mod_par, components, masked):
    eng = AnalyticEngine()
    m_true = _vvr_model(seed=2)
    m = _vvr_model(seed=3)
    if mod_par == 0:
        m_true, m = P.elastic_conv(m_true, 1), P.elastic_conv(m, 1)
    d = eng._f(P.elastic_conv(m_true, mod_par))
    d_obs = P.datcomp_forward(d, components) if components else d
    rng = np.random.default_rng(5)
    data_mask = torch.tensor(rng.uniform(0, 1, d_obs.shape)) if masked else None
    grad_mask = torch.tensor(rng.uniform(0, 1, m.shape)) if masked else None
    prob = P.FwiProblem(eng, d_obs, components=components, mod_par=mod_par, grad_mask=grad_mask,
                        data_mask=data_mask)
    phi, g, r = prob.evaluate(m)
    phi_ref, g_ref = _autograd_gradient(eng, m, mod_par, d_obs, components, data_mask)
    if grad_mask is not None:
        g_ref = g_ref * grad_mask
    assert phi == pytest.approx(phi_ref, rel=1e-12)
    assert float(torch.linalg.norm(g - g_ref) / torch.linalg.norm(g_ref)) < 1e-10
    assert prob.n_forward == 1 and prob.n_gradient == 1  # the gradient reused the kept forward


def test_mod_par2_gradient_is_reference_jacobian():
    """mod_par 2 (km/s, g/cm^3): ElasticConvJab at setBackground(m) = 1e-3 x the exact
    derivative, so the gradient is 1e-3 x autograd's."""
    eng = AnalyticEngine(seed=4)
    m = _vvr_model(seed=6) * 1e-3
    d_obs = eng._f(P.elastic_conv(_vvr_model(seed=7) * 1e-3, 2))
    prob = P.FwiProblem(eng, d_obs, mod_par=2)
    _, g, _ = prob.evaluate(m)
    _, g_ref = _autograd_gradient(eng, m, 2, d_obs, None, None)
    assert float(torch.linalg.norm(g - 1e-3 * g_ref) / torch.linalg.norm(1e-3 * g_ref)) < 1e-10


def test_conv_jacobian_adjoint_is_vjp():
    m = _vvr_model(seed=8)
    g = torch.randn(m.shape, dtype=torch.float64)
    mm = m.clone().requires_grad_(True)
    (ref,) = torch.autograd.grad(P.elastic_conv(mm, 1), mm, grad_outputs=g)
    got = P.elastic_conv_jacobian_adjoint(m, g, 1)
    assert float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref)) < 1e-13


def test_datcomp_adjoint_dot():
    rng = np.random.default_rng(9)
    x = torch.tensor(rng.standard_normal((2, 5, 3, 7)))
    for comps in ("vx,vz,p", "p", "sxx,szz,sxz,vx,vz", "p,sxx"):
        y = torch.tensor(rng.standard_normal((2, len(comps.split(",")), 3, 7)))
        lhs = float(torch.sum(P.datcomp_forward(x, comps) * y))
        rhs = float(torch.sum(x * P.datcomp_adjoint(y, comps)))
        assert lhs == pytest.approx(rhs, rel=1e-13)
    with pytest.raises(ValueError):
        P.datcomp_forward(x, "vx,vx")
    with pytest.raises(ValueError):
        P.datcomp_forward(x, "pressure")


def test_bounds_projection_and_courant_guard():
    eng = AnalyticEngine()
    m = _vvr_model()
    d_obs = eng._f(P.elastic_conv(m, 1))
    lo = torch.tensor([2400.0, 1300.0, 2050.0], dtype=torch.float64)[:, None, None].expand(m.shape)
    prob = P.FwiProblem(eng, d_obs, mod_par=1, min_bound=lo, max_bound=None)
    pm = prob.project(m)
    assert bool((pm >= lo).all()) and bool(((pm == m) | (pm == lo)).all())
    # vmax 2 dtw / d: at dtw=1e-3, d=10, vmax ~ 3000 gives 0.6 > 0.45
    prob = P.FwiProblem(eng, d_obs, mod_par=1, dtw=1e-3, dz=10.0, dx=10.0)
    with pytest.raises(P.StabilityError):
        prob.objective(m)
    prob = P.FwiProblem(eng, d_obs, mod_par=1, dtw=5e-4, dz=10.0, dx=10.0)
    assert prob.objective(m) == pytest.approx(0.0, abs=1e-20)


class FunctionProblem:
    """An analytic objective behind the problem interface the solvers use."""

    def __init__(self, f, lo=None, hi=None, blowup=None):
        self.f, self.lo, self.hi, self.blowup = f, lo, hi, blowup
        self.n_forward = self.n_gradient = 0
        self.last_residual = None

    def project(self, m):
        if self.lo is not None:
            m = torch.maximum(m, self.lo)
        if self.hi is not None:
            m = torch.minimum(m, self.hi)
        return m

    def objective(self, m, keep=False):
        m = self.project(m)
        if self.blowup is not None and float(m.abs().max()) > self.blowup:
            raise P.StabilityError("blow-up")
        self.n_forward += 1
        return float(self.f(m))

    def gradient(self, m):
        m = self.project(m).clone().requires_grad_(True)
        self.n_gradient += 1
        (g,) = torch.autograd.grad(self.f(m), m)
        return g

    def evaluate(self, m):
        return self.objective(m), self.gradient(m), None

    def close(self):
        pass


def _rosenbrock(m):
    x = m.reshape(-1)
    return torch.sum(100.0 * (x[1:] - x[:-1] ** 2) ** 2 + (1.0 - x[:-1]) ** 2)


@pytest.mark.parametrize("memory", [None, 5])
def test_lbfgs_rosenbrock(memory):
    prob = FunctionProblem(_rosenbrock)
    m0 = torch.tensor([-1.2, 1.0, -1.2, 1.0, 0.5, -0.3], dtype=torch.float64)
    res = P.lbfgs(prob, m0, 300, memory=memory, init_step=1e-3, tol_grad=1e-12)
    assert float((res.model - 1.0).abs().max()) < 1e-6
    assert all(b <= a + 1e-12 for a, b in zip(res.objective, res.objective[1:]))  # monotone


def _quadratic(n=20, cond=100.0, seed=0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = q @ np.diag(np.geomspace(1.0, cond, n)) @ q.T
    b = rng.standard_normal(n)
    A_t, b_t = torch.tensor(A), torch.tensor(b)
    return (lambda m: 0.5 * m @ (A_t @ m) - b_t @ m), np.linalg.solve(A, b)


@pytest.mark.parametrize("solver,kw", [("lbfgs", {}), ("nlcg", {"beta_type": "FR"}),
                                       ("nlcg", {"beta_type": "PRP"}), ("nlcg", {"beta_type": "HS"}),
                                       ("nlcg", {"beta_type": "DY"}),
                                       ("nlcg", {"beta_type": "PRP", "stepper": "backtracking"})])
def test_solvers_quadratic(solver, kw):
    f, xs = _quadratic()
    prob = FunctionProblem(f)
    res = P.solve(prob, torch.zeros(20, dtype=torch.float64), 400, solver=solver, tol_grad=1e-10, **kw)
    assert np.linalg.norm(res.model.numpy() - xs) / np.linalg.norm(xs) < 1e-5, res.stop_reason


def test_steepest_descent_decreases():
    f, xs = _quadratic(n=8, cond=10.0, seed=3)
    prob = FunctionProblem(f)
    res = P.solve(prob, torch.zeros(8, dtype=torch.float64), 200, solver="sd", tol_grad=1e-8)
    assert np.linalg.norm(res.model.numpy() - xs) / np.linalg.norm(xs) < 1e-4
    assert all(b < a for a, b in zip(res.objective, res.objective[1:]))


def test_lbfgs_box_constrained_separable():
    """min sum d_i (x_i - c_i)^2 over a box: the minimiser is clip(c, lo, hi)."""
    rng = np.random.default_rng(4)
    d = torch.tensor(rng.uniform(1.0, 10.0, 12))
    c = torch.tensor(rng.uniform(-2.0, 2.0, 12))
    lo, hi = torch.full((12,), -1.0, dtype=torch.float64), torch.full((12,), 1.0, dtype=torch.float64)
    prob = FunctionProblem(lambda m: torch.sum(d * (m - c) ** 2), lo=lo, hi=hi)
    res = P.lbfgs(prob, torch.zeros(12, dtype=torch.float64), 100, init_step=0.05)
    assert float((res.model - torch.clamp(c, -1.0, 1.0)).abs().max()) < 1e-6


def test_line_search_backs_off_unstable_models():
    """A trial the propagator would refuse (StabilityError) counts as phi = inf."""
    f, xs = _quadratic(n=6, cond=5.0, seed=5)
    prob = FunctionProblem(f, blowup=50.0)
    res = P.lbfgs(prob, torch.zeros(6, dtype=torch.float64), 50, init_step=1e4, tol_grad=1e-10)
    assert np.linalg.norm(res.model.numpy() - xs) / np.linalg.norm(xs) < 1e-6


def test_lbfgs_on_problem_layer_decreases_and_respects_bounds():
    eng = AnalyticEngine(seed=11)
    m_true = _vvr_model(seed=12)
    m0 = _vvr_model(seed=13)
    d_obs = P.datcomp_forward(eng._f(P.elastic_conv(m_true, 1)), "vx,vz,p")
    lo = torch.tensor([2300.0, 1200.0, 1950.0], dtype=torch.float64)[:, None, None].expand(m0.shape).clone()
    hi = torch.tensor([2900.0, 1700.0, 2250.0], dtype=torch.float64)[:, None, None].expand(m0.shape).clone()
    prob = P.FwiProblem(eng, d_obs, components="vx,vz,p", mod_par=1, min_bound=lo, max_bound=hi)
    seen = []
    res = P.lbfgs(prob, m0, 8, callback=lambda it, m, phi, g, r: seen.append((it, phi)))
    assert res.iterations == 8 and [s[0] for s in seen] == list(range(0, 9))
    assert [s[1] for s in seen] == res.objective
    assert res.objective[-1] < 0.2 * res.objective[0]
    assert bool((res.model >= lo).all()) and bool((res.model <= hi).all())
    assert eng.n_store == prob.n_forward  # every forward kept its checkpoints for the gradient


def test_parfile_coordinates_match_workload_geometry(tmp_path):
    """The parfile geometry rules (Elastic_iso_float_prop.py:311-347, 615-641) and the source
    and receiver geometry files."""
    from elastic_isotropic_operators_amd import geometry as G, workloads as W
    from elastic_isotropic_operators_amd.sep import Hypercube, SepVector, genericIO
    w = W.small(nz=40, nx=56, nts=50, nExp=3)
    par = genericIO.io(params=["nExp=3", "zSource=3", "xSource=5", "spacingShots=7", "zPadMinus=%d" % w["zPadMinus"],
                               "xPadMinus=%d" % w["xPadMinus"], "nReceiver=9", "depthReceiver=2",
                               "oReceiver=4", "dReceiver=3"])
    coords, (n, o, d) = G.parfile_source_coords(par, w["oz"], w["dz"], w["ox"], w["dx"])
    sz, sx = W.source_coords(w, 3, 5, nExp=3, spacingShots=7)
    np.testing.assert_allclose(np.stack([c[0] for c in coords]), sz)
    np.testing.assert_allclose(np.stack([c[1] for c in coords]), sx)
    assert (n, d) == (3, 7 * w["dx"]) and o == pytest.approx(sx[0, 0])
    rz, rx, (nr, orr, dr) = G.parfile_receiver_coords(par, w["oz"], w["dz"], w["ox"], w["dx"])
    rz2, rx2 = W.receiver_coords(w, 9, 2, 4, 3)
    np.testing.assert_allclose(rz, rz2)
    np.testing.assert_allclose(rx, rx2)
    assert nr == 9 and dr == 3 * w["dx"]
    # geometry files: [x, y, z] along the slowest axis, devices, shots along the fastest
    geom = SepVector.getSepVector(Hypercube.hypercube(axes=[Hypercube.axis(n=3), Hypercube.axis(n=2),
                                                            Hypercube.axis(n=3)]))
    a = geom.getNdArray()               # (3 coords, 2 devices, 3 shots)
    a[0] = np.arange(6).reshape(2, 3) * 10.0 + 100.0
    a[2] = np.arange(6).reshape(2, 3) * 5.0 + 30.0
    genericIO.defaultIO.writeVector(str(tmp_path / "sg.H"), geom)
    par = genericIO.io(params=["nExp=3", "sourceGeomFile=%s" % (tmp_path / "sg.H"), "nReceiver=2",
                               "recGeomFile=%s" % (tmp_path / "sg.H")])
    coords, (n, o, d) = G.parfile_source_coords(par, 0.0, 1.0, 0.0, 1.0, genericIO.defaultIO.getVector)
    assert (n, o, d) == (3, 1.0, 1.0)
    for e in range(3):
        np.testing.assert_allclose(coords[e][0], a[2, :, e])
        np.testing.assert_allclose(coords[e][1], a[0, :, e])
    rz, rx, _ = G.parfile_receiver_coords(par, 0.0, 1.0, 0.0, 1.0, genericIO.defaultIO.getVector)
    np.testing.assert_allclose(rz, a[2, :, 0])
    np.testing.assert_allclose(rx, a[0, :, 0])
    par = genericIO.io(params=["nExp=2", "sourceGeomFile=%s" % (tmp_path / "sg.H")])
    with pytest.raises(ValueError):
        G.parfile_source_coords(par, 0.0, 1.0, 0.0, 1.0, genericIO.defaultIO.getVector)


def test_efwi_rejects_unavailable_options():
    from elastic_isotropic_operators_amd import efwi
    from elastic_isotropic_operators_amd.sep import genericIO
    with pytest.raises(NotImplementedError):
        efwi.build_problem(genericIO.io(params=["spline=1"]))
    with pytest.raises(RuntimeError):
        efwi.build_problem(genericIO.io(params=["surfaceCondition=1"]))
