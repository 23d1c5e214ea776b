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

The reference minimises with liblbfgs 1.10 (not part of the reference sources here);
this module uses scipy's L-BFGS-B with the same iteration cap (50) and restart rule
(run again while the energy still falls).  scipy's stopping test is on the largest
projected gradient entry (``gtol``), liblbfgs's on ||g|| / max(1, ||x||) (``epsilon``,
1e-6 in the reference): trajectories differ, the optimum they approach does not.
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


def minimize_lbfgs(energy, restart=0, verbose=False, max_iterations=50, gtol=1e-6):
    """minimizeLBFGS (optimization.cpp:68-103): L-BFGS from energy.initial_value(), run
    again up to `restart` times while the energy still falls; the model is left at the
    returned parameters."""
    from scipy.optimize import minimize

    x = energy.initial_value()
    best = np.inf
    for i in range(restart + 1):
        res = minimize(energy.gradient, x, jac=True, method="L-BFGS-B",
                       options={"maxiter": max_iterations, "gtol": gtol})
        if verbose:
            print("L-BFGS run %d: f = %.6g, %d iterations, %s" % (i, res.fun, res.nit, res.message))
        if res.fun < best:
            best, x = res.fun, res.x
        else:
            break
    energy.set(x)
    return x.astype(np.float32)


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
