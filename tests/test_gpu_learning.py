"""The learning loop on the GPU gradients: CRFEnergy's gradient against central
differences of its own value (numericGradient / gradCheck, optimization.cpp:104-126),
and minimize_lbfgs lowering the energy for each objective."""
import numpy as np
import pytest

from pydensecrf_amd import capi, learning

pytestmark = pytest.mark.gpu


def _model(seed, N=240, L=3, F=4, logistic=True):
    rng = np.random.default_rng(seed)
    g = capi.Model.create(N, L)
    if logistic:
        g.set_logistic_unary((rng.standard_normal((L, F)) * 0.4).astype(np.float32),
                             rng.standard_normal((F, N)).astype(np.float32))
    else:
        g.set_unary(rng.standard_normal((L, N)).astype(np.float32))
    # 1-D features: both blur axes of the lattice are the same operator, the filter is
    # symmetric and the reference's kernel gradient is the exact gradient
    g.add_pairwise((rng.standard_normal((1, N)) * 3).astype(np.float32), 1.5, capi.DIAG_KERNEL, 3)
    g.add_pairwise((rng.standard_normal((1, N)) * 2).astype(np.float32),
                   (rng.standard_normal((L, L)) * 0.5).astype(np.float32), capi.FULL_KERNEL, 2)
    gt = rng.integers(0, L, N).astype(np.int16)
    return g, gt


@pytest.mark.parametrize("objective", [learning.LOG_LIKELIHOOD, learning.HAMMING, learning.IOU])
def test_energy_gradient_central_differences(objective):
    g, gt = _model(7)
    e = learning.CRFEnergy(g, gt, 3, objective=objective, unary=True, pairwise=True, kernel=True, l2=0.01)
    x = e.initial_value()
    assert x.size == 3 * 4 + 1 + 6 + 1 + 1
    _, an = e.gradient(x)
    ng = learning.numeric_gradient(e, x, 1e-2)
    assert np.linalg.norm(ng - an) <= 2e-2 * np.linalg.norm(an), (ng, an)
    assert learning.grad_check(e, x, 1e-2) <= 2e-2 * np.linalg.norm(an)


@pytest.mark.parametrize("objective,param", [(learning.LOG_LIKELIHOOD, 0.0), (learning.HAMMING, 0.5),
                                             (learning.IOU, 0.0)])
def test_minimize_lbfgs_lowers_energy(objective, param):
    g, gt = _model(11, logistic=False)
    e = learning.CRFEnergy(g, gt, 3, objective=objective, param=param, pairwise=True)
    x0 = e.initial_value()
    f0 = e.gradient(x0)[0]
    x = learning.minimize_lbfgs(e, restart=1, max_iterations=15)
    f1 = e.gradient(x)[0]
    assert f1 < f0
    np.testing.assert_array_equal(g.compat_params(), x)  # the model is left at the result


def test_minimize_lbfgs_reaches_reference_optimum():
    """The compatibility-learning problem of tests/golden/lbfgs_small.npz ("crfc": the
    reference's minimizeLBFGS with one restart on -LogLikelihood + l2 over Diagonal
    compatibilities, an energy smooth in its parameters) run with the GPU's energy.  The
    GPU's values differ from the reference's in the last float bits, so the iterates
    differ a little; the optimum reached agrees to 1e-4 of the parameters' scale and its
    energy to 1e-6.  (With the bilateral term's kernel parameters the energy is
    discontinuous, d = 5, and runs that differ in rounding end at different points.)"""
    import os

    from tests.helpers import crf_learning_problem
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lbfgs_small.npz"),
                allow_pickle=False)
    img, U, gt = crf_learning_problem()
    g = capi.Model.create_2d(16, 12, 3)
    g.set_unary(U)
    g.add_gaussian(3.0, 2.0, np.array([1.0, 2.0, 1.5], np.float32))
    g.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, img, np.array([3.0, 4.0, 5.0], np.float32))
    e = learning.CRFEnergy(g, gt, 3, pairwise=True, kernel=False, l2=1e-2)
    assert np.array_equal(e.initial_value().astype(np.float32), z["crfc_x0"])
    x = learning.minimize_lbfgs(e, restart=1)
    ref = z["crfc_x"]
    err = float(np.abs(x - ref).max()) / float(np.abs(ref).max())
    f_gpu = e.gradient(x)[0]
    f_ref = float(z["crfc_values"].min())
    print("GPU L-BFGS: %d evaluations (reference %d), max |x - x_ref| / max |x_ref| = %.3g, E %.9f (reference "
          "best %.9f)" % (e.evaluations, len(z["crfc_trace"]), err, f_gpu, f_ref), flush=True)
    assert err <= 1e-4
    assert abs(f_gpu - f_ref) <= 1e-6 * abs(f_ref)
