"""Misfit, adjoint source and model update — the steps either side of the solver.

- ``L2DistanceLoss``: f = 1/2 ||modelled - observed||^2, adjoint source = +residual for the
  modelled traces (stride/optimisation/loss/l2_distance.py:36-66).
- ``apply_prec``: illumination preconditioning of a gradient by its ``prec``
  (stride/problem/data.py:409-446): smooth, scale to mean ``prec_scale``, add 1, divide.
- ``GradientDescent``: ``variable -= step * direction`` in the variable's transformed space
  (slowness for vp; stride/optimisation/optimisers/optimiser.py:242-330).
- ``LBFGS``: limited-memory BFGS two-loop recursion with a backtracking line search (no
  reference counterpart: Stride ships only gradient descent).

Arrays may be numpy arrays or torch tensors (on the GPU the misfit and residual stay in HBM).
"""

import numpy as np
import scipy.ndimage

__all__ = ['L2DistanceLoss', 'apply_prec', 'GradientDescent', 'LBFGS']


def _xp(a):
    try:
        import torch
        if isinstance(a, torch.Tensor):
            return torch
    except ImportError:
        pass
    return np


def _unwrap(a):
    """The array of a ``Traces``-like container (anything with ``.data`` that is not
    itself an array or tensor)."""
    if isinstance(a, np.ndarray) or _xp(a) is not np:
        return a
    return getattr(a, 'data', a)


class L2DistanceLoss:
    def __init__(self, **kwargs):
        self.residual = None

    def forward(self, modelled, observed):
        modelled = _unwrap(modelled)
        observed = _unwrap(observed)
        xp = _xp(modelled)
        residual = modelled - observed
        self.residual = residual
        if xp is np:
            fun = 0.5 * float(np.sum(residual.astype(np.float64) ** 2))
        else:
            fun = 0.5 * float((residual.double() ** 2).sum())
        return fun, residual

    def adjoint(self, d_fun=1.0, modelled_needs_grad=True, observed_needs_grad=False):
        grad_modelled = d_fun * self.residual if modelled_needs_grad else None
        grad_observed = -d_fun * self.residual if observed_needs_grad else None
        self.residual = None
        return grad_modelled, grad_observed


def apply_prec(grad, prec, prec_scale=4.0, prec_smooth=1.0, prec_op=None):
    grad = np.array(grad, dtype=np.float32, copy=True)
    if prec is None:
        return grad
    prec = np.array(prec, dtype=np.float64, copy=True)
    if prec_smooth is not None:
        prec = scipy.ndimage.gaussian_filter(prec, prec_smooth)
    factor = np.sum(np.abs(prec))
    if factor > 1e-31:
        factor = prec_scale * prec.size / factor
        prec = prec * factor + 1
        if prec_op is not None:
            prec = prec_op(prec)
        nz = np.abs(prec) > 0
        grad[nz] *= 1 / prec[nz]
    return grad


class GradientDescent:
    """``x <- x - step * direction`` applied to ``transform(x)`` (1/vp for vp), then mapped
    back, with optional bounds on the physical variable."""

    def __init__(self, step_size=1.0, transform=None, inverse=None, bounds=None):
        self.step_size = step_size
        self.transform = transform
        self.inverse = inverse
        self.bounds = bounds

    def step(self, variable, direction, step_size=None):
        step = self.step_size if step_size is None else step_size
        x = variable if self.transform is None else self.transform(variable)
        x = x - step * direction
        out = x if self.inverse is None else self.inverse(x)
        if self.bounds is not None:
            out = np.clip(out, *self.bounds)
        return out.astype(np.float32)


class LBFGS:
    """Limited-memory BFGS on a flat float64 vector.

    ``minimize(x0, fun_and_grad, iterations)`` with ``fun_and_grad(x) -> (f, g)``; every call
    counts as one solver call (forward + adjoint). The first step is scaled so that the
    largest update is ``first_step`` (in the units of x)."""

    def __init__(self, memory=8, first_step=1.0, c1=1e-4, max_backtracks=6, bounds=None):
        self.memory = memory
        self.first_step = first_step
        self.c1 = c1
        self.max_backtracks = max_backtracks
        self.bounds = bounds
        self.calls = 0
        self.history = []
        self.proposals_kept = 0  # first steps from first_direction that were kept

    def _clip(self, x):
        return x if self.bounds is None else np.clip(x, *self.bounds)

    def minimize(self, x0, fun_and_grad, iterations, callback=None, first_direction=None, gated=False):
        """``first_direction(x, g)``, when given, proposes the first step of the run (the step
        is -r for its return value r, None for the scaled gradient); it is tried at full length
        once and, when it fails the sufficient-decrease test, replaced by the scaled gradient.
        ``gated``: ``fun_and_grad(x, threshold)`` may return ``(f, None)`` without the gradient
        when ``f`` exceeds ``threshold`` (a rejected trial then costs the forward pass only)."""
        x = np.asarray(x0, dtype=np.float64).copy()
        f, g = fun_and_grad(x, None) if gated else fun_and_grad(x)
        self.calls += 1
        self.history.append((self.calls, f))
        S, Y = [], []
        for it in range(iterations):
            q = g.copy()
            alphas = []
            for s, y in reversed(list(zip(S, Y))):
                rho = 1.0 / np.dot(y, s)
                a = rho * np.dot(s, q)
                alphas.append((a, rho, s, y))
                q -= a * y
            if S:
                gamma = np.dot(S[-1], Y[-1]) / np.dot(Y[-1], Y[-1])
                r = gamma * q
            else:
                r = q * (self.first_step / max(np.max(np.abs(q)), 1e-30))
            for a, rho, s, y in reversed(alphas):
                b = rho * np.dot(y, r)
                r += s * (a - b)
            d = -r
            slope = np.dot(g, d)
            if slope >= 0:  # not a descent direction: restart from steepest descent
                S, Y = [], []
                d = -g * (self.first_step / max(np.max(np.abs(g)), 1e-30))
                slope = np.dot(g, d)
            trials = []
            if it == 0 and first_direction is not None:
                rp = first_direction(x, g)
                if rp is not None and np.dot(g, -rp) < 0:
                    trials.append((-np.asarray(rp, np.float64), 1))  # tried once at full length
            trials.append((d, self.max_backtracks))
            accepted = False
            for k, (d, tries) in enumerate(trials):
                slope = np.dot(g, d)
                t = 1.0
                for _ in range(tries):
                    xn = self._clip(x + t * d)
                    threshold = f + self.c1 * t * slope
                    fn, gn = fun_and_grad(xn, threshold) if gated else fun_and_grad(xn)
                    self.calls += 1
                    if gn is not None and fn <= threshold:
                        accepted = True
                        break
                    t *= 0.5
                if accepted:
                    self.proposals_kept += int(k + 1 < len(trials))
                    break
            if not accepted:
                break
            s_vec, y_vec = xn - x, gn - g
            if np.dot(s_vec, y_vec) > 1e-12 * np.linalg.norm(s_vec) * np.linalg.norm(y_vec):
                S.append(s_vec)
                Y.append(y_vec)
                if len(S) > self.memory:
                    S.pop(0)
                    Y.pop(0)
            x, f, g = xn, fn, gn
            self.history.append((self.calls, f))
            if callback is not None:
                callback(x, f)
        return x, f
