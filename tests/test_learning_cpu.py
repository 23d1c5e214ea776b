"""The learning driver (pydensecrf_amd.learning) without a GPU.

* lbfgs / minimize_lbfgs against the reference's minimizeLBFGS (liblbfgs 1.10, built
  from the reference sources): tests/golden/lbfgs_small.npz holds every point the
  reference evaluated and its result, on Rosenbrock and on a CRF learning energy whose
  values the oracle reproduces bit for bit; the restatement must evaluate the same
  points and return the same result.
* minimize_lbfgs finds the minimum of a convex quadratic and stops its restarts once
  the energy no longer falls; numeric_gradient, gradient, grad_check and
  compute_function follow optimization.cpp:104-134 on a closed-form gradient.
"""
import os

import numpy as np
import pytest

from pydensecrf_amd import learning

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _traced(fun, trace):
    def f(x):
        trace.append(x.copy())
        return fun(x)
    return f


@pytest.mark.parametrize("i", [0, 1])
def test_lbfgs_rosenbrock_matches_reference(i):
    from tests.helpers import rosenbrock
    z = np.load(os.path.join(GOLDEN, "lbfgs_small.npz"), allow_pickle=False)
    tr = []
    x, _, status = learning.lbfgs(_traced(rosenbrock, tr), z["r%d_x0" % i])
    ref = z["r%d_trace" % i]
    assert len(tr) == len(ref)
    assert np.array_equal(np.array(tr), ref)
    assert np.array_equal(x.astype(np.float32), z["r%d_x" % i])
    assert status in (learning.LBFGS_SUCCESS, learning.LBFGSERR_MAXIMUMITERATION)


@pytest.mark.parametrize("name,kernel", [("crf", True), ("crfc", False)])
def test_minimize_lbfgs_crf_energy_matches_reference(name, kernel):
    """minimizeLBFGS with one restart on -LogLikelihood + l2 of a small CRF (Diagonal
    compatibilities, with or without the DIAG kernel parameters), the energy evaluated by
    the oracle (whose DenseCRF::gradient is the reference's, bit for bit, for these
    terms)."""
    from oracle import pyoracle as po
    from tests.helpers import crf_learning_energy, crf_learning_model
    z = np.load(os.path.join(GOLDEN, "lbfgs_small.npz"), allow_pickle=False)
    crf, gt = crf_learning_model(po.OracleCRF)
    x0 = np.concatenate([crf.compat_params()] + ([crf.kernel_params()] if kernel else [])).astype(np.float32)
    assert np.array_equal(x0, z[name + "_x0"])
    fun = crf_learning_energy(crf, gt, kernel=kernel)

    class Energy:
        def __init__(self):
            self.trace, self.values, self.last = [], [], None

        def initial_value(self):
            return x0

        def gradient(self, x):
            self.trace.append(np.asarray(x, np.float32).copy())
            f, g = fun(np.asarray(x, np.float32))
            self.values.append(f)
            return f, g

        def set(self, x):
            self.last = x

    e = Energy()
    x = learning.minimize_lbfgs(e, restart=1)
    assert len(e.trace) == len(z[name + "_trace"])
    assert np.array_equal(np.array(e.trace), z[name + "_trace"])
    np.testing.assert_array_equal(np.array(e.values), z[name + "_values"])
    assert np.array_equal(x, z[name + "_x"]) and np.array_equal(e.last, x)


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
