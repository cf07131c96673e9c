"""zlh_mlp_gradient_noalloc (the gradient entry point that allocates no device memory) against
zlh_mlp_gradient on the same inputs, on two calls in a row, the second with fewer samples, in
both training forms; and the refusal of more samples than its static array holds."""
import numpy as np
import pytest

from zllbm_amd import closure, hip
from tests import scenarios as sc

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 9)).astype(np.float32)
    w = rng.normal(size=9)
    T = np.logaddexp(0.0, X @ w * 0.5).astype(np.float32)
    return X, T


@pytest.mark.parametrize("form", [hip.MLP_FP32, hip.MLP_BF16_MFMA])
def test_noalloc_gradient_matches_allocating_entry_point(form):
    grid, kw = sc.periodic_tgv(3, 4, nu=0.02)
    s = hip.Solver(grid, **kw)
    p = closure.init_params(5)
    # the full array first, then a smaller call: the second must not see the first's samples
    for n, seed in ((hip.MLP_NOALLOC_MAX_SAMPLES, 1), (3001, 2)):
        X, T = _data(n, seed)
        loss_a, ga = s.mlp_gradient(p, X, T, form=form)
        loss_b, gb = s.mlp_gradient_noalloc(p, X, T, form=form)
        print("form %d, n %d: loss %.6e / %.6e, max |dg| %.3e of max |g| %.3e"
              % (form, n, loss_a, loss_b, np.max(np.abs(ga - gb)), np.max(np.abs(ga))))
        # same kernel and inputs; only the order of the floating-point atomic sums may differ
        assert abs(loss_a - loss_b) <= 1e-5 * abs(loss_a)
        assert np.max(np.abs(ga - gb)) <= 1e-5 * np.max(np.abs(ga))


def test_noalloc_gradient_refuses_more_samples_than_it_holds():
    grid, kw = sc.periodic_tgv(3, 4, nu=0.02)
    s = hip.Solver(grid, **kw)
    X, T = _data(hip.MLP_NOALLOC_MAX_SAMPLES + 1, 3)
    with pytest.raises(RuntimeError, match="at most 65536 samples"):
        s.mlp_gradient_noalloc(closure.init_params(5), X, T)
