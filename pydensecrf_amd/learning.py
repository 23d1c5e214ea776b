"""Parameter learning over the GPU gradients.

The reference's learning loop, host side: an energy over a parameter vector whose value
and gradient come from ``DenseCRF::gradient`` (densecrf.cpp:238-297) with one of its
objectives (objective.h:34-75), the parameter get/set of densecrf.cpp:298-360, and the
L-BFGS driver and checks of optimization.h:30-40 / optimization.cpp:68-130
(``minimizeLBFGS``, ``numericGradient``, ``gradient``, ``gradCheck``,
``computeFunction``).

The model is a ``pydensecrf_amd.capi.Model``; every energy evaluation is one training
forward and one adjoint on the GPU.  The parameter vector concatenates, in this order
and for the groups switched on: the unary parameters (a logistic unary's L, packed
column-major), the label-compatibility parameters (packed like
``labelCompatibilityParameters``) and the kernel parameters (``kernelParameters``).

The minimiser is liblbfgs 1.10's L-BFGS as the reference links it
(pydensecrf/densecrf/external/liblbfgs/lib/lbfgs.c, built without SSE; parameters of
optimization.cpp:80-84: 6 correction pairs, epsilon 1e-6 on ||g|| / max(1, ||x||), at
most 50 iterations, the More-Thuente line search with ftol 1e-4, gtol 0.9, xtol 1e-16,
steps in [1e-20, 1e20], at most 40 evaluations per search), restated in ``lbfgs``:
x is kept in double and rounded to float32 for every evaluation, as the reference's
callback does, and every vector reduction is a sequential double sum, so on the same
energy values the iterates are the reference's.
"""
import numpy as np

from . import capi

LOG_LIKELIHOOD = capi.OBJECTIVE_LOG_LIKELIHOOD
HAMMING = capi.OBJECTIVE_HAMMING
IOU = capi.OBJECTIVE_IOU


class CRFEnergy:
    """E(x) = -objective(x) + l2 / 2 * ||x||^2 over the chosen parameter groups.

    model: capi.Model with its unary and pairwise terms; gt: int16 labels (N,), values
    outside [0, L) ignored; n_iter: mean-field iterations; objective: LOG_LIKELIHOOD
    (param = robust), HAMMING (class_weight, or param = class_weight_pow) or IOU.
    unary / pairwise / kernel: learn the logistic unary's parameters, the
    compatibilities, the kernel parameters.
    """

    def __init__(self, model, gt, n_iter, objective=LOG_LIKELIHOOD, param=0.0, class_weight=None,
                 unary=False, pairwise=True, kernel=False, l2=0.0):
        self.model = model
        self.gt = np.ascontiguousarray(gt, dtype=np.int16)
        if self.gt.shape != (model.N,):
            raise ValueError("gt must have one label per variable (%d), got %s" % (model.N, self.gt.shape))
        self.n_iter = int(n_iter)
        self.objective, self.param = int(objective), float(param)
        self.class_weight = None if class_weight is None else np.asarray(class_weight, np.float32)
        self.unary, self.pairwise, self.kernel = bool(unary), bool(pairwise), bool(kernel)
        self.l2 = float(l2)
        if self.unary and model.unary_params().size == 0:
            raise ValueError("the model's unary has no parameters (set a logistic unary to learn it)")
        self.sizes = [model.unary_params().size if self.unary else 0,
                      model.compat_params().size if self.pairwise else 0,
                      model.kernel_params().size if self.kernel else 0]
        self.evaluations = 0

    def initial_value(self):
        parts = []
        if self.unary:
            parts.append(self.model.unary_params())
        if self.pairwise:
            parts.append(self.model.compat_params())
        if self.kernel:
            parts.append(self.model.kernel_params())
        return np.concatenate(parts).astype(np.float64) if parts else np.zeros(0)

    def set(self, x):
        """Write the parameter vector into the model."""
        x = np.asarray(x, dtype=np.float32)
        if x.size != sum(self.sizes):
            raise ValueError("expected %d parameters, got %d" % (sum(self.sizes), x.size))
        a, b = self.sizes[0], self.sizes[0] + self.sizes[1]
        if self.unary:
            self.model.set_unary_params(x[:a])
        if self.pairwise:
            self.model.set_compat_params(x[a:b])
        if self.kernel:
            kp = x[b:]
            if not np.array_equal(kp, self.model.kernel_params()):  # rebuilds the lattices
                self.model.set_kernel_params(kp)

    def gradient(self, x):
        """(E(x), dE/dx), both float64."""
        self.set(x)
        r, _, cg, _, kg = self.model.gradient_objective(self.n_iter, self.gt, self.objective, self.param,
                                                        self.class_weight, kernel=self.kernel, unary_grad=False)
        self.evaluations += 1
        parts = []
        if self.unary:
            parts.append(self.model.unary_param_gradient())
        if self.pairwise:
            parts.append(cg)
        if self.kernel:
            parts.append(kg)
        x = np.asarray(x, dtype=np.float64)
        g = -np.concatenate(parts).astype(np.float64) if parts else np.zeros(0)
        return -r + 0.5 * self.l2 * float(x @ x), g + self.l2 * x


# liblbfgs status codes (lbfgs.h)
LBFGS_SUCCESS = 0
LBFGS_ALREADY_MINIMIZED = 2
LBFGSERR_INVALID_N = -1020
LBFGSERR_OUTOFINTERVAL = -1003
LBFGSERR_INCORRECT_TMINMAX = -1002
LBFGSERR_ROUNDING_ERROR = -1001
LBFGSERR_MINIMUMSTEP = -1000
LBFGSERR_MAXIMUMSTEP = -999
LBFGSERR_MAXIMUMLINESEARCH = -998
LBFGSERR_MAXIMUMITERATION = -997
LBFGSERR_WIDTHTOOSMALL = -996
LBFGSERR_INVALIDPARAMETERS = -995
LBFGSERR_INCREASEGRADIENT = -994

# the reference's settings (optimization.cpp:80-84 over liblbfgs's defaults)
LBFGS_PARAMS = dict(m=6, epsilon=1e-6, max_iterations=50, ftol=1e-4, gtol=0.9, xtol=1e-16, min_step=1e-20,
                    max_step=1e20, max_linesearch=40)

_F64 = np.float64


def _dot(a, b):
    """a . b summed left to right in double (a plain loop's order)."""
    return _F64(np.cumsum(a * b)[-1]) if a.size else _F64(0.0)


def _norm(a):
    return np.sqrt(_dot(a, a))


def _larger(a, b):
    return a if a >= b else b


def _smaller(a, b):
    return a if a <= b else b


def _cubic_min(u, fu, du, v, fv, dv):
    """Minimiser of the cubic through (u, fu, du) and (v, fv, dv)."""
    span = v - u
    theta = (fu - fv) * 3 / span + du + dv
    sc = _larger(_larger(abs(theta), abs(du)), abs(dv))
    a = theta / sc
    gamma = sc * np.sqrt(a * a - (du / sc) * (dv / sc))
    if v < u:
        gamma = -gamma
    ratio = (gamma - du + theta) / (gamma - du + gamma + dv)
    return u + ratio * span


def _cubic_min_safe(u, fu, du, v, fv, dv, lo, hi):
    """As _cubic_min, seen from v; an unbounded cubic gives the bound it heads for."""
    span = v - u
    theta = (fu - fv) * 3 / span + du + dv
    sc = _larger(_larger(abs(theta), abs(du)), abs(dv))
    a = theta / sc
    gamma = sc * np.sqrt(_larger(_F64(0.0), a * a - (du / sc) * (dv / sc)))
    if u < v:
        gamma = -gamma
    ratio = (gamma - dv + theta) / (gamma - dv + gamma + du)
    if ratio < 0.0 and gamma != 0.0:
        return v - ratio * span
    return hi if a < 0 else lo


def _quad_min(u, fu, du, v, fv):
    """Minimiser of the quadratic through (u, fu, du) and (v, fv)."""
    span = v - u
    return u + du / ((fu - fv) / span + du) / 2 * span


def _secant_min(u, du, v, dv):
    """Zero of the derivative's secant through (u, du) and (v, dv)."""
    return v + dv / (dv - du) * (u - v)


def _trial_step(best, other, trial, lo, hi, bracketed):
    """One update of the Moré-Thuente interval of uncertainty.  best / other / trial are
    (step, value, derivative) triples of the best step so far, the interval's other end
    and the step just evaluated.  Returns (status, best, other, next step, bracketed);
    a nonzero status leaves everything as it was."""
    x, fx, dx = best
    y, fy, dy = other
    t, ft, dt = trial
    opposite = dt * (dx / abs(dx)) < 0.0
    if bracketed:
        if t <= _smaller(x, y) or _larger(x, y) <= t:
            return LBFGSERR_OUTOFINTERVAL, best, other, t, bracketed
        if 0.0 <= dx * (t - x):
            return LBFGSERR_INCREASEGRADIENT, best, other, t, bracketed
        if hi < lo:
            return LBFGSERR_INCORRECT_TMINMAX, best, other, t, bracketed
    if fx < ft:
        # higher value: the minimum is bracketed; the cubic step when it is the closer
        # to x, else halfway between the cubic and the quadratic step
        bracketed, near_upper = True, True
        mc = _cubic_min(x, fx, dx, t, ft, dt)
        mq = _quad_min(x, fx, dx, t, ft)
        nxt = mc if abs(mc - x) < abs(mq - x) else mc + 0.5 * (mq - mc)
    elif opposite:
        # lower value, derivatives of opposite sign: bracketed; the cubic step when it
        # is the farther from t, else the secant step
        bracketed, near_upper = True, False
        mc = _cubic_min(x, fx, dx, t, ft, dt)
        mq = _secant_min(x, dx, t, dt)
        nxt = mc if abs(mc - t) > abs(mq - t) else mq
    elif abs(dt) < abs(dx):
        # lower value, same sign, shrinking derivative
        near_upper = True
        mc = _cubic_min_safe(x, fx, dx, t, ft, dt, lo, hi)
        mq = _secant_min(x, dx, t, dt)
        if bracketed:
            nxt = mc if abs(t - mc) < abs(t - mq) else mq
        else:
            nxt = mc if abs(t - mc) > abs(t - mq) else mq
    else:
        # lower value, same sign, derivative not shrinking
        near_upper = False
        if bracketed:
            nxt = _cubic_min(t, ft, dt, y, fy, dy)
        else:
            nxt = hi if x < t else lo
    if fx < ft:
        other = trial
    else:
        if opposite:
            other = best
        best = trial
    if hi < nxt:
        nxt = hi
    if nxt < lo:
        nxt = lo
    if bracketed and near_upper:
        cap = best[0] + 0.66 * (other[0] - best[0])
        if best[0] < other[0]:
            if cap < nxt:
                nxt = cap
        elif nxt < cap:
            nxt = cap
    return 0, best, other, nxt, bracketed


def _line_search(evaluate, xp, f0, g0, d, stp, p):
    """Moré-Thuente search along d from xp for a step with sufficient decrease
    (ftol) and curvature (gtol).  Returns (evaluations or a negative status, x, f, g,
    step); f and g are those of the last evaluation."""
    if stp <= 0.0:
        return LBFGSERR_INVALIDPARAMETERS, xp, f0, g0, stp
    dg0 = _dot(g0, d)
    if 0.0 < dg0:
        return LBFGSERR_INCREASEGRADIENT, xp, f0, g0, stp
    bracketed, stage1, uinfo, count = False, True, 0, 0
    dgtest = p["ftol"] * dg0
    width = _F64(p["max_step"] - p["min_step"])
    prev_width = 2.0 * width
    best = (_F64(0.0), f0, dg0)
    other = (_F64(0.0), f0, dg0)
    x, f, g = xp, f0, g0
    while True:
        if bracketed:
            lo, hi = _smaller(best[0], other[0]), _larger(best[0], other[0])
        else:
            lo, hi = best[0], stp + 4.0 * (stp - best[0])
        if stp < p["min_step"]:
            stp = _F64(p["min_step"])
        if p["max_step"] < stp:
            stp = _F64(p["max_step"])
        if bracketed and ((stp <= lo or hi <= stp) or p["max_linesearch"] <= count + 1 or uinfo != 0 or
                          hi - lo <= p["xtol"] * hi):
            stp = best[0]
        x = xp + stp * d
        f, g = evaluate(x)
        dg = _dot(g, d)
        ftest = f0 + stp * dgtest
        count += 1
        if bracketed and ((stp <= lo or hi <= stp) or uinfo != 0):
            return LBFGSERR_ROUNDING_ERROR, x, f, g, stp
        if stp == p["max_step"] and f <= ftest and dg <= dgtest:
            return LBFGSERR_MAXIMUMSTEP, x, f, g, stp
        if stp == p["min_step"] and (ftest < f or dgtest <= dg):
            return LBFGSERR_MINIMUMSTEP, x, f, g, stp
        if bracketed and hi - lo <= p["xtol"] * hi:
            return LBFGSERR_WIDTHTOOSMALL, x, f, g, stp
        if p["max_linesearch"] <= count:
            return LBFGSERR_MAXIMUMLINESEARCH, x, f, g, stp
        if f <= ftest and abs(dg) <= p["gtol"] * (-dg0):
            return count, x, f, g, stp
        if stage1 and f <= ftest and _smaller(p["ftol"], p["gtol"]) * dg0 <= dg:
            stage1 = False
        if stage1 and ftest < f and f <= best[1]:
            # the modified function psi(s) = f(s) - s dgtest until psi has a
            # nonpositive value and a nonnegative derivative
            def shift(tr):
                return (tr[0], tr[1] - tr[0] * dgtest, tr[2] - dgtest)

            def unshift(tr):
                return (tr[0], tr[1] + tr[0] * dgtest, tr[2] + dgtest)

            uinfo, best, other, stp, bracketed = _trial_step(shift(best), shift(other), shift((stp, f, dg)),
                                                             lo, hi, bracketed)
            best, other = unshift(best), unshift(other)
        else:
            uinfo, best, other, stp, bracketed = _trial_step(best, other, (stp, f, dg), lo, hi, bracketed)
        if bracketed:
            if 0.66 * prev_width <= abs(other[0] - best[0]):
                stp = best[0] + 0.5 * (other[0] - best[0])
            prev_width = width
            width = abs(other[0] - best[0])


def lbfgs(fun, x0, **params):
    """Minimise fun from x0 with limited-memory BFGS (Nocedal 1980, two-loop recursion)
    and the Moré-Thuente line search, as liblbfgs 1.10's lbfgs() with the reference's
    settings (LBFGS_PARAMS, overridable).  fun(x float32 (n,)) -> (value, gradient).
    Returns (x float64, f, status): f is the last evaluation's value (after a failed
    line search x is the previous iterate, as in liblbfgs)."""
    p = dict(LBFGS_PARAMS, **params)
    x = np.array(x0, dtype=np.float64).ravel()
    n = x.size
    if n <= 0:
        return x, _F64(np.nan), LBFGSERR_INVALID_N

    def evaluate(xv):
        val, grad = fun(xv.astype(np.float32))
        return _F64(val), np.asarray(grad, np.float32).astype(np.float64).ravel()

    with np.errstate(all="ignore"):
        f, g = evaluate(x)
        d = -g
        if _norm(g) / _larger(_norm(x), _F64(1.0)) <= p["epsilon"]:
            return x, f, LBFGS_ALREADY_MINIMIZED
        step = 1.0 / _norm(d)
        pairs = []  # (s, y, y.s), oldest first, at most m
        k = 1
        while True:
            xp, gp = x, g
            ls, x, f, g, step = _line_search(evaluate, xp, f, gp, d, step, p)
            if ls < 0:
                return xp, f, ls
            if _norm(g) / _larger(_norm(x), _F64(1.0)) <= p["epsilon"]:
                return x, f, LBFGS_SUCCESS
            if p["max_iterations"] != 0 and p["max_iterations"] < k + 1:
                return x, f, LBFGSERR_MAXIMUMITERATION
            s, y = x - xp, g - gp
            ys, yy = _dot(y, s), _dot(y, y)
            pairs.append((s, y, ys))
            if len(pairs) > p["m"]:
                pairs.pop(0)
            k += 1
            d = -g
            alpha = []
            for sj, yj, ysj in reversed(pairs):
                a = _dot(sj, d) / ysj
                alpha.append(a)
                d = d + (-a) * yj
            d = d * (ys / yy)
            for (sj, yj, ysj), a in zip(pairs, reversed(alpha)):
                beta = _dot(yj, d) / ysj
                d = d + (a - beta) * sj
            step = _F64(1.0)


def minimize_lbfgs(energy, restart=0, verbose=False, **params):
    """minimizeLBFGS (optimization.cpp:68-103): lbfgs from energy.initial_value(), run
    again from where it stopped, up to `restart` more times, while the energy still
    falls; the model is left at the returned float32 parameters."""
    x = np.asarray(energy.initial_value(), np.float64)
    last = 1e100
    status = None
    for i in range(restart + 1):
        x, fx, status = lbfgs(lambda v: energy.gradient(v), x, **params)
        if verbose:
            print("L-BFGS run %d: f = %.6g, status %d" % (i, fx, status))
        if last > fx:
            last = fx
        else:
            break
    out = x.astype(np.float32)
    energy.set(out)
    return out


def numeric_gradient(energy, x, eps=1e-3):
    """numericGradient (optimization.cpp:104-115): central differences of E."""
    x = np.asarray(x, dtype=np.float64)
    g = np.empty_like(x)
    for i in range(x.size):
        xx = x.copy()
        xx[i] = x[i] + eps
        v1 = energy.gradient(xx)[0]
        xx[i] = x[i] - eps
        v0 = energy.gradient(xx)[0]
        g[i] = (v1 - v0) / (2 * eps)
    energy.set(x)
    return g


def gradient(energy, x):
    """gradient (optimization.cpp:116-120): dE/dx."""
    return energy.gradient(x)[1]


def grad_check(energy, x, eps=1e-3):
    """gradCheck (optimization.cpp:121-126): ||numeric - analytic gradient||."""
    ng = numeric_gradient(energy, x, eps)
    return float(np.linalg.norm(ng - energy.gradient(x)[1]))


def compute_function(energy, x, dx, n_samples=100):
    """computeFunction (optimization.cpp:128-134): E along x + i dx, i = 0 .. n_samples-1."""
    x, dx = np.asarray(x, np.float64), np.asarray(dx, np.float64)
    return np.array([energy.gradient(x + i * dx)[0] for i in range(n_samples)])
