"""The learning driver (pydensecrf_amd.learning) on a known energy, without a GPU:
minimize_lbfgs finds the minimum of a convex quadratic and stops its restarts once the
energy no longer falls; numeric_gradient, gradient, grad_check and compute_function
follow optimization.cpp:104-134 on a closed-form gradient."""
import numpy as np

from pydensecrf_amd import learning


class Quadratic:
    """E(x) = 1/2 (x - c)^T A (x - c), A symmetric positive definite; set() records x."""

    def __init__(self, n=6, seed=0):
        rng = np.random.default_rng(seed)
        B = rng.standard_normal((n, n))
        self.A = B @ B.T + n * np.eye(n)
        self.c = rng.standard_normal(n)
        self.x0 = np.zeros(n)
        self.last = None
        self.evaluations = 0

    def initial_value(self):
        return self.x0.copy()

    def set(self, x):
        self.last = np.asarray(x, np.float64)

    def gradient(self, x):
        self.evaluations += 1
        d = np.asarray(x, np.float64) - self.c
        return 0.5 * d @ self.A @ d, self.A @ d


def test_minimize_lbfgs_quadratic():
    e = Quadratic()
    x = learning.minimize_lbfgs(e, restart=3)
    np.testing.assert_allclose(x, e.c, atol=1e-5)
    np.testing.assert_allclose(e.last, x, atol=1e-6)  # the model is left at the result


def test_restart_stops_when_energy_stalls():
    e = Quadratic(seed=1)
    learning.minimize_lbfgs(e, restart=0)
    one = e.evaluations
    e2 = Quadratic(seed=1)
    learning.minimize_lbfgs(e2, restart=5)
    # the second run starts at the optimum and does not improve: no third run
    assert e2.evaluations <= 3 * one


def test_numeric_gradient_and_checks():
    e = Quadratic(seed=2)
    x = np.linspace(-1, 1, 6)
    np.testing.assert_allclose(learning.numeric_gradient(e, x, 1e-4), e.A @ (x - e.c), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(learning.gradient(e, x), e.A @ (x - e.c))
    assert learning.grad_check(e, x, 1e-4) < 1e-5
    dx = np.full(6, 0.1)
    f = learning.compute_function(e, x, dx, 5)
    want = [e.gradient(x + i * dx)[0] for i in range(5)]
    np.testing.assert_allclose(f, want)
