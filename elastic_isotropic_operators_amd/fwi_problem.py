"""Elastic FWI problem layer and solvers in torch, around the GPU-resident gradient.

The reference's FWI script (efwiMain.py:229-476) assembles and minimises

    phi(m) = 1/2 || W_d C F(g(m)) - W_d d_obs ||^2      subject to  lo <= m <= hi

  m    model in the inversion parameterisation, (3, nx, nz): [rho, lambda, mu] (mod_par 0),
       [vp, vs, rho] in m/s and kg/m^3 (mod_par 1) or km/s and g/cm^3 (mod_par 2)
  g    ElasticConv (elasticParamConvertModule.py:23-47), chained by CombNonlinearOp
       (efwiMain.py:331-340)
  F    the nonlinear shot operator (nonlinearFwiPropElasticShotsGpu)
  C    ElasticDatComp, the data components (efwiMain.py:345-368)
  W_d  the data mask (dataMaskFile, efwiMain.py:371-389): applied to the modelled data and,
       once, to the observed data
  gradient  W_m J_g^T B^T C^T W_d^T r, with B the Born operator at g(m) and W_m the gradient
       mask (maskGradient, efwiMain.py:258-265,342-343)
  bounds   minBound_par1..3 / maxBound_par1..3 or bound vectors (efwiMain.py:118-174): the
       model is clipped to [lo, hi] before every evaluation.

The Jacobian of g is the reference's ElasticConvJab (elasticParamConvertModule.py:51-117) at
the background the solver sets through setBackground.  For mod_par 1 that is the exact
derivative.  For mod_par 2 setBackground multiplies the background by 1e3
(elasticParamConvertModule.py:110-117) while the formula stays the m/s one, which equals
1e-3 times the exact derivative of the km/s conversion: the gradient carries that factor,
its direction is exact.

Solvers.  The reference hands the problem to pySolver (ProblemL2NonLinear, LBFGSsolver,
NLCGsolver and its steppers), a third-party library that is not vendored and has no pinned
version.  The solvers below restate the textbook methods, not pySolver's code (Nocedal &
Wright, Numerical Optimization, 2nd ed.: L-BFGS two-loop recursion, Alg. 7.4; nonlinear
conjugate gradients, section 5.2; backtracking with quadratic/cubic interpolation, section
3.5), so their iterates are not expected to equal pySolver's.  Every objective evaluation
is one batched NL forward on the GPU; every gradient one checkpointed Born adjoint that
reuses the checkpoints of the accepted trial's forward, so a line search that accepts its
first trial costs one forward and one gradient per iteration.
"""
import math
import time

import torch

_COMP = {"vx": 0, "vz": 1, "sxx": 2, "szz": 3, "sxz": 4}


class StabilityError(ValueError):
    """The model breaks the propagator's Courant limit (fdParamElastic.cpp:256-267)."""


# ------------------------------------------------------------------ parameterisation
def elastic_conv(m, mod_par):
    """ElasticConv.forward (elasticParamConvertModule.py:23-47) on (3, nx, nz); mod_par 0 is
    the identity ([rho, lambda, mu] inverted directly)."""
    if mod_par == 0:
        return m
    if mod_par not in (1, 2):
        raise ValueError("ERROR! Unsupported conv_type (current value: %s). Please redefine it!" % mod_par)
    a, b = (1e3, 1e9) if mod_par == 2 else (1.0, 1.0)
    vp, vs, rho = m[0], m[1], m[2]
    return torch.stack([a * rho, b * rho * (vp * vp - 2.0 * vs * vs), b * rho * vs * vs])


def elastic_conv_jacobian_adjoint(m, g, mod_par):
    """ElasticConvJab.adjoint (elasticParamConvertModule.py:91-108) of g (3, nx, nz) at
    setBackground(m) (:110-117)."""
    if mod_par == 0:
        return g
    if mod_par not in (1, 2):
        raise ValueError("ERROR! Unsupported conv_type (current value: %s). Please redefine it!" % mod_par)
    bg = m * 1e3 if mod_par == 2 else m
    vp, vs, rho = bg[0], bg[1], bg[2]
    return torch.stack([2.0 * vp * rho * g[1],
                        2.0 * vs * rho * g[2] - 4.0 * vs * rho * g[1],
                        g[0] + (vp * vp - 2.0 * vs * vs) * g[1] + vs * vs * g[2]])


# ------------------------------------------------------------------ data components
def _comp_list(components):
    comps = [c.strip().lower() for c in components.split(",")]
    if len(set(comps)) != len(comps):
        raise ValueError("ERROR! A component was provided multiple times! Check your input arguments")
    for c in comps:
        if c != "p" and c not in _COMP:
            raise ValueError("ERROR! Provided unknown data components: %s" % components)
    return comps


def datcomp_forward(data, components):
    """ElasticDatComp.forward (dataCompModule.py:45-77): (nShot, 5, nRec, nts) -> (nShot, nComp,
    nRec, nts), p = (sxx + szz) / 2."""
    out = [0.5 * (data[:, 2] + data[:, 3]) if c == "p" else data[:, _COMP[c]] for c in _comp_list(components)]
    return torch.stack(out, dim=1)


def datcomp_adjoint(sampled, components):
    """ElasticDatComp.adjoint (dataCompModule.py:79-107)."""
    shape = (sampled.shape[0], 5) + tuple(sampled.shape[2:])
    out = torch.zeros(shape, dtype=sampled.dtype, device=sampled.device)
    for i, c in enumerate(_comp_list(components)):
        if c == "p":
            out[:, 2] += 0.5 * sampled[:, i]
            out[:, 3] += 0.5 * sampled[:, i]
        else:
            out[:, _COMP[c]] += sampled[:, i]
    return out


def _dot(a, b):
    return float(torch.sum(a.double() * b.double()))


# ------------------------------------------------------------------ problem
class FwiProblem:
    """phi(m), its gradient and the bound projection for one set of shots.

    engine      an ElasticFwi (all shots resident on one GPU), or any object with .device,
                .dtype, .forward_data(lame, checkpoint) -> (data, store, media) and
                .gradient(store, media, grad_data) -> (3, nx, nz)
    d_obs       observed data, (nShot, 5, nRec, nts) or (nShot, nComp, nRec, nts) with
                `components`
    components  ElasticDatComp list ("vx,vz,p", ...); None keeps all five
    mod_par     0, 1 or 2 as above
    grad_mask   (3, nx, nz) gradient mask, or None
    data_mask   same shape as d_obs, or None
    min_bound, max_bound  (3, nx, nz) tensors, scalars or None
    courant_limit, dtw, dz, dx  when dtw is given, a model whose vmax*2*dtw/min(dz, dx)
                exceeds the limit raises StabilityError (the reference's setBackground
                refuses it: fdParamElastic.cpp:256-267)
    """

    def __init__(self, engine, d_obs, components=None, mod_par=0, grad_mask=None, data_mask=None,
                 min_bound=None, max_bound=None, courant_limit=0.45, dtw=None, dz=None, dx=None):
        self.engine = engine
        dev, dt = engine.device, engine.dtype

        def to(x):
            return None if x is None else torch.as_tensor(x, dtype=dt, device=dev)

        self.components = components
        if components:
            _comp_list(components)
        self.mod_par = int(mod_par)
        if self.mod_par not in (0, 1, 2):
            raise ValueError("ERROR! Unsupported conv_type (current value: %s). Please redefine it!" % mod_par)
        self.data_mask = to(data_mask)
        d = to(d_obs)
        self.d_obs = d * self.data_mask if self.data_mask is not None else d
        self.grad_mask = to(grad_mask)
        self.lo, self.hi = to(min_bound), to(max_bound)
        self.courant_limit = courant_limit
        self.dtw, self.dz, self.dx = dtw, dz, dx
        self.n_forward = 0
        self.n_gradient = 0
        self.last_residual = None
        self._kept = None

    # -------------------------------------------------------------- pieces
    def project(self, m):
        """Clip to the bounds (the solver's bound projection)."""
        if self.lo is not None:
            m = torch.maximum(m, self.lo)
        if self.hi is not None:
            m = torch.minimum(m, self.hi)
        return m

    def courant(self, lame):
        """vmax * 2 dtw / min(dz, dx) of a [rho, lambda, mu] model (fdParamElastic.cpp:81-98,256-267)."""
        vp = torch.sqrt(torch.clamp((lame[1] + 2.0 * lame[2]) / lame[0], min=0.0))
        return float(vp.max()) * 2.0 * self.dtw / min(self.dz, self.dx)

    def residual(self, data):
        d = datcomp_forward(data, self.components) if self.components else data
        if self.data_mask is not None:
            d = d * self.data_mask
        return d - self.d_obs

    def _release(self):
        if self._kept is not None:
            store = self._kept[1]
            if store is not None and hasattr(store, "release"):
                store.release()
            self._kept = None

    # -------------------------------------------------------------- evaluations
    def objective(self, m, keep=False):
        """phi at project(m).  keep=True keeps the forward's checkpoints so that gradient() at
        the same model needs no second forward."""
        m = self.project(m)
        lame = elastic_conv(m, self.mod_par)
        if not bool(torch.isfinite(lame).all()) or bool((lame[0] <= 0).any()):
            raise StabilityError("model has non-finite or non-positive density")
        if self.dtw is not None:
            c = self.courant(lame)
            if not c <= self.courant_limit:
                raise StabilityError("Courant number %.3f exceeds %.2f" % (c, self.courant_limit))
        self._release()
        data, store, media = self.engine.forward_data(lame, checkpoint=keep)
        self.n_forward += 1
        r = self.residual(data)
        phi = 0.5 * _dot(r, r)
        self.last_residual = r
        if keep:
            self._kept = (m, store, media, r)
        return phi

    def gradient(self, m):
        """W_m J_g^T B^T C^T W_d^T r at project(m)."""
        m = self.project(m)
        if self._kept is None or not torch.equal(self._kept[0], m):
            self.objective(m, keep=True)
        m, store, media, r = self._kept
        gd = r * self.data_mask if self.data_mask is not None else r
        if self.components:
            gd = datcomp_adjoint(gd, self.components)
        g = self.engine.gradient(store, media, gd)
        self.n_gradient += 1
        self._release()
        g = elastic_conv_jacobian_adjoint(m, g, self.mod_par)
        if self.grad_mask is not None:
            g = g * self.grad_mask
        return g

    def evaluate(self, m):
        """(phi, gradient, residual) at project(m)."""
        phi = self.objective(m, keep=True)
        r = self.last_residual
        return phi, self.gradient(m), r

    def close(self):
        self._release()


# ------------------------------------------------------------------ solvers
class SolverResult:
    def __init__(self):
        self.model = None
        self.objective = []      # phi at the starting model and after every iteration
        self.gradient = None
        self.iterations = 0
        self.stop_reason = ""
        self.n_forward = 0
        self.n_gradient = 0
        self.seconds = 0.0


def _trial(problem, m, alpha, p):
    """phi at project(m + alpha p), keeping its checkpoints; inf for an unstable model."""
    mt = problem.project(m + alpha * p)
    try:
        phi = problem.objective(mt, keep=True)
    except StabilityError:
        return mt, math.inf
    return mt, phi


def _free_direction(problem, m, p):
    """p with the entries that push against an active bound set to zero (those the bound
    projection would not let move)."""
    pm = float(p.abs().max())
    if pm == 0.0:
        return p
    mm = float(m.abs().max())
    moved = problem.project(m + (1e-6 * (mm if mm > 0 else 1.0) / pm) * p) - m
    return torch.where(moved != 0, p, torch.zeros_like(p))


def backtracking(problem, m, phi0, g, p, alpha, c1=1e-4, max_trials=12):
    """Armijo backtracking with quadratic then cubic interpolation of phi(alpha) (Nocedal &
    Wright section 3.5), each new step kept in [0.1, 0.5] of the last one.  The sufficient
    decrease is measured on the projected displacement, g'(P(m + alpha p) - m).  Returns
    (model, phi, alpha) of the accepted trial, whose checkpoints the problem keeps, or None."""
    dphi0 = _dot(g, p)
    prev = None
    for _ in range(max_trials):
        mt, phi = _trial(problem, m, alpha, p)
        dec = _dot(g, mt - m)
        if math.isfinite(phi) and dec < 0.0 and phi <= phi0 + c1 * dec:
            return mt, phi, alpha
        if not math.isfinite(phi):
            new = 0.1 * alpha
        elif prev is None:
            den = 2.0 * (phi - phi0 - dphi0 * alpha)
            new = -dphi0 * alpha * alpha / den if den > 0 else 0.5 * alpha
        else:
            a0, f0 = prev
            a1, f1 = alpha, phi
            d1 = f1 - phi0 - dphi0 * a1
            d0 = f0 - phi0 - dphi0 * a0
            den = a0 * a0 * a1 * a1 * (a1 - a0)
            ca = (a0 * a0 * d1 - a1 * a1 * d0) / den
            cb = (-a0 * a0 * a0 * d1 + a1 * a1 * a1 * d0) / den
            if abs(ca) < 1e-300:
                new = -dphi0 / (2.0 * cb) if cb > 0 else 0.5 * alpha
            else:
                disc = cb * cb - 3.0 * ca * dphi0
                new = (-cb + math.sqrt(disc)) / (3.0 * ca) if disc >= 0 else 0.5 * alpha
        if not math.isfinite(phi):
            prev = None
        else:
            prev = (alpha, phi)
        alpha = min(max(new, 0.1 * alpha), 0.5 * alpha)
    return None


def parabolic(problem, m, phi0, g, p, alpha, max_trials=12):
    """Step from a parabola through phi(0), phi(alpha), phi(2 alpha): the best of the two
    trials and the parabola's minimum is accepted if it lowers phi, else alpha is halved.
    Returns as backtracking."""
    for _ in range(max_trials):
        m1, f1 = _trial(problem, m, alpha, p)
        if not math.isfinite(f1):
            alpha *= 0.1
            continue
        m2, f2 = _trial(problem, m, 2.0 * alpha, p)
        curv = f2 - 2.0 * f1 + phi0
        cands = [(f1, alpha, m1), (f2, 2.0 * alpha, m2)]
        if curv > 0:
            a = alpha * (3.0 * phi0 - 4.0 * f1 + f2) / (2.0 * curv)
            if a > 0:
                mp, fp = _trial(problem, m, a, p)
                cands.append((fp, a, mp))
        best = min(cands, key=lambda t: t[0])
        if best[0] < phi0:
            # gradient() re-runs the forward unless the best trial was the last one evaluated
            fb, ab, mb = best
            return mb, fb, ab
        alpha *= 0.5
    return None


def _line_search(kind):
    if kind in ("backtracking", "linear", "default"):
        return backtracking
    if kind in ("parabolic", "parabolicNew"):
        return parabolic
    raise ValueError("ERROR! Provided unknonw stepper type: %s" % kind)


def _first_step(m, p, init_step, step_frac):
    if init_step and init_step > 0:
        return float(init_step)
    pm, mm = float(p.abs().max()), float(m.abs().max())
    return step_frac * (mm if mm > 0 else 1.0) / pm if pm > 0 else 1.0


def _stop(res, it, n_iter, phi, phi0, gnorm, g0norm, tol_obj, tol_grad):
    if it >= n_iter:
        res.stop_reason = "maximum number of iterations"
    elif tol_obj and phi <= tol_obj * phi0:
        res.stop_reason = "objective below %g of its initial value" % tol_obj
    elif tol_grad and gnorm <= tol_grad * g0norm:
        res.stop_reason = "gradient norm below %g of its initial value" % tol_grad
    else:
        return False
    return True


def lbfgs(problem, m0, n_iter, memory=None, h0=None, init_step=None, step_frac=0.01,
          stepper="backtracking", tol_obj=0.0, tol_grad=0.0, callback=None):
    """Projected L-BFGS.  Directions from the two-loop recursion (Nocedal & Wright Alg. 7.4)
    over the last `memory` pairs (all pairs when None), initial inverse Hessian h0 (a diagonal
    tensor: the reference's illumination option, efwiMain.py:404-410) or the scaling
    s'y / y'y; pairs with s'y <= 0 are skipped.  The first step length is `init_step`
    (initStep) or the one moving the largest model entry by step_frac of the model's maximum;
    later iterations try the unit step first.  callback(iteration, model, phi, gradient,
    residual) runs at the starting model (iteration 0) and after every iteration."""
    search = _line_search(stepper)
    t0 = time.time()
    res = SolverResult()
    m = problem.project(m0.detach().clone())
    phi, g, _ = problem.evaluate(m)
    res.objective.append(phi)
    phi0, g0norm = phi, math.sqrt(_dot(g, g))
    if callback is not None:
        callback(0, m, phi, g, problem.last_residual)
    pairs = []
    it = 0
    while not _stop(res, it, n_iter, phi, phi0, math.sqrt(_dot(g, g)), g0norm, tol_obj, tol_grad):
        q = g.clone()
        alphas = []
        for s, y, rho in reversed(pairs):
            a = rho * _dot(s, q)
            q -= a * y
            alphas.append(a)
        if h0 is not None:
            r = h0 * q
        elif pairs:
            s, y, _ = pairs[-1]
            r = (_dot(s, y) / _dot(y, y)) * q
        else:
            r = q
        for (s, y, rho), a in zip(pairs, reversed(alphas)):
            b = rho * _dot(y, r)
            r += (a - b) * s
        p = _free_direction(problem, m, -r)
        if _dot(p, g) >= 0.0:  # not a descent direction: restart from the gradient
            pairs.clear()
            p = _free_direction(problem, m, -(h0 * g if h0 is not None else g))
        alpha = 1.0 if pairs else _first_step(m, p, init_step, step_frac)
        step = search(problem, m, phi, g, p, alpha)
        if step is None:
            res.stop_reason = "line search failed"
            break
        m_new, phi_new, _ = step
        g_new = problem.gradient(m_new)
        s, y = m_new - m, g_new - g
        sy = _dot(s, y)
        if sy > 1e-12 * math.sqrt(_dot(s, s) * _dot(y, y)):
            pairs.append((s, y, 1.0 / sy))
            if memory and len(pairs) > memory:
                pairs.pop(0)
        m, phi, g = m_new, phi_new, g_new
        it += 1
        res.objective.append(phi)
        if callback is not None:
            callback(it, m, phi, g, problem.last_residual)
    res.model, res.gradient, res.iterations = m, g, it
    res.n_forward, res.n_gradient = problem.n_forward, problem.n_gradient
    res.seconds = time.time() - t0
    problem.close()
    return res


_BETAS = ("FR", "PRP", "HS", "DY", "SD")


def nlcg(problem, m0, n_iter, beta_type="FR", init_step=None, step_frac=0.01, stepper="parabolic",
         tol_obj=0.0, tol_grad=0.0, callback=None):
    """Projected nonlinear conjugate gradients (Nocedal & Wright section 5.2) with beta
    Fletcher-Reeves (FR), Polak-Ribiere+ (PRP), Hestenes-Stiefel (HS), Dai-Yuan (DY) or 0
    (SD: steepest descent).  A direction that is not a descent direction restarts from the
    negative gradient.  Initial step of later iterations: alpha_prev g_prev'p_prev / g'p
    (Nocedal & Wright eq. 3.60)."""
    beta_type = beta_type.upper()
    if beta_type not in _BETAS:
        raise ValueError("unknown beta_type %s (one of %s)" % (beta_type, ", ".join(_BETAS)))
    search = _line_search(stepper)
    t0 = time.time()
    res = SolverResult()
    m = problem.project(m0.detach().clone())
    phi, g, _ = problem.evaluate(m)
    res.objective.append(phi)
    phi0, g0norm = phi, math.sqrt(_dot(g, g))
    if callback is not None:
        callback(0, m, phi, g, problem.last_residual)
    p = _free_direction(problem, m, -g)
    alpha_prev, gp_prev = None, None
    it = 0
    while not _stop(res, it, n_iter, phi, phi0, math.sqrt(_dot(g, g)), g0norm, tol_obj, tol_grad):
        gp = _dot(g, p)
        if alpha_prev is None:
            alpha = _first_step(m, p, init_step, step_frac)
        else:
            alpha = alpha_prev * gp_prev / gp
        step = search(problem, m, phi, g, p, alpha)
        if step is None:
            res.stop_reason = "line search failed"
            break
        m_new, phi_new, alpha_prev = step
        gp_prev = gp
        g_new = problem.gradient(m_new)
        y = g_new - g
        gg = _dot(g, g)
        if beta_type == "FR":
            beta = _dot(g_new, g_new) / gg
        elif beta_type == "PRP":
            beta = max(0.0, _dot(g_new, y) / gg)
        elif beta_type == "HS":
            beta = _dot(g_new, y) / _dot(p, y)
        elif beta_type == "DY":
            beta = _dot(g_new, g_new) / _dot(p, y)
        else:
            beta = 0.0
        p = _free_direction(problem, m_new, -g_new + beta * p)
        if _dot(p, g_new) >= 0.0:
            p = _free_direction(problem, m_new, -g_new)
        m, phi, g = m_new, phi_new, g_new
        it += 1
        res.objective.append(phi)
        if callback is not None:
            callback(it, m, phi, g, problem.last_residual)
    res.model, res.gradient, res.iterations = m, g, it
    res.n_forward, res.n_gradient = problem.n_forward, problem.n_gradient
    res.seconds = time.time() - t0
    problem.close()
    return res


def solve(problem, m0, n_iter, solver="lbfgs", **kw):
    """efwiMain.py:392-420: solver=lbfgs | nlcg | sd."""
    if solver == "lbfgs":
        return lbfgs(problem, m0, n_iter, **kw)
    if solver == "nlcg":
        return nlcg(problem, m0, n_iter, **kw)
    if solver == "sd":
        kw["beta_type"] = "SD"
        return nlcg(problem, m0, n_iter, **kw)
    raise ValueError("ERROR! Provided unknonw solver type: %s" % solver)
