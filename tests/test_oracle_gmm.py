"""The colour-GMM oracle (oracle/gmm_oracle.py) pinned on the CPU: closed
forms of the score, scipy's multivariate normal, a planted mistake, the
quantisation's rounding, and EM against scikit-learn's GaussianMixture from
the same start."""
import warnings

import numpy as np
import pytest

import gmm_oracle as M

LOG2PI = np.log(2 * np.pi)


def test_single_component_closed_form():
    # N(0, I): score(x) = 1.5 log(2 pi) + |x|^2 / 2
    x = np.array([[0, 0, 0], [1, 2, 2], [3, 0, 4]], np.uint8)
    s = M.score([1.0], [[0, 0, 0]], [[1, 1, 1]], x)
    assert np.allclose(s, 1.5 * LOG2PI + np.array([0.0, 4.5, 12.5]), rtol=0, atol=1e-12)
    # mean (10, 20, 30), variances (4, 9, 16), at the mean and one sigma off in each channel
    x = np.array([[10, 20, 30], [12, 23, 34]], np.uint8)
    s = M.score([1.0], [[10, 20, 30]], [[4, 9, 16]], x)
    base = 0.5 * (3 * LOG2PI + np.log(4) + np.log(9) + np.log(16))
    assert np.allclose(s, [base, base + 1.5], rtol=0, atol=1e-12)


def test_mixture_identities():
    rng = np.random.default_rng(1)
    x = rng.integers(0, 256, (500, 3)).astype(np.uint8)
    one = M.score([1.0], [[100, 50, 20]], [[300, 200, 400]], x)
    # two identical halves are the same component
    two = M.score([0.5, 0.5], [[100, 50, 20]] * 2, [[300, 200, 400]] * 2, x)
    assert np.allclose(one, two, rtol=1e-13)
    # a component of weight 0 changes nothing
    zero = M.score([1.0, 0.0], [[100, 50, 20], [0, 0, 0]], [[300, 200, 400], [1, 1, 1]], x)
    assert np.allclose(one, zero, rtol=1e-13)


def test_scores_match_scipy():
    from scipy.special import logsumexp
    from scipy.stats import multivariate_normal
    rng = np.random.default_rng(2)
    x = rng.integers(0, 256, (2000, 4)).astype(np.uint8)
    for k in (1, 3, 8):
        w = rng.dirichlet(np.ones(k))
        mu = rng.uniform(0, 255, (k, 3))
        v = rng.uniform(5, 2000, (k, 3))
        ref = -logsumexp(np.stack([np.log(w[i]) + multivariate_normal(mu[i], np.diag(v[i])).logpdf(
            x[:, :3].astype(np.float64)) for i in range(k)], axis=1), axis=1)
        got = M.score(w, mu, v, x)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-10), k
        # planted mistake in the oracle: h without its factor 0.5
        c, m, h = M.constants(w, mu, v)
        bad = -M.lse(M.terms(c, m, 2 * h, M.colours(x)))
        assert not np.allclose(bad, ref, rtol=1e-6)


def test_quantise_rounding_and_bounds():
    s = np.array([-3.0, 0.0, 0.25, 0.35, 2.5, 3.5, 1e9, np.inf])
    q = M.quantise(s, 10.0, 100.0)
    # negative -> 0; 2.5 -> 25 exactly; ties to even at scale 1; capped at 100
    assert list(q) == [0, 0, 2, 4, 25, 35, 1000, 1000]
    assert list(M.quantise(np.array([0.5, 1.5, 2.5]), 1.0, 10.0)) == [0, 2, 2]


def test_em_matches_sklearn():
    from sklearn.exceptions import ConvergenceWarning
    from sklearn.mixture import GaussianMixture
    rng = np.random.default_rng(3)
    centres = np.array([[40, 60, 90], [200, 150, 80], [120, 120, 120]])
    x = np.clip(np.concatenate([rng.normal(c, 15, (700, 3)) for c in centres]), 0, 255)
    x = np.rint(x).astype(np.uint8)
    rng.shuffle(x)
    k, iters, min_covar = 3, 12, 0.001
    w, mu, v, ll, it = M.fit(x, k, iters, init_covar=30.0, min_covar=min_covar)
    assert it == iters
    n = len(x)
    idx = [i * (n - 1) // (k - 1) for i in range(k)]
    xf = x.astype(np.float64)
    g = GaussianMixture(k, covariance_type="diag", max_iter=iters, tol=0.0, reg_covar=min_covar,
                        weights_init=np.full(k, 1.0 / k), means_init=xf[idx],
                        precisions_init=np.full((k, 3), 1.0 / 30.0), init_params="random")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", ConvergenceWarning)
        g.fit(xf)
    # sklearn adds 10 * eps to each component's responsibility sum
    assert np.allclose(w, g.weights_, rtol=1e-9)
    assert np.allclose(mu, g.means_, rtol=1e-9)
    assert np.allclose(v, g.covariances_, rtol=1e-9)
    # the fitted mixture separates the three colour clusters
    assert np.allclose(np.sort(mu[:, 0]), np.sort(centres[:, 0]), atol=3)


def test_em_tolerance_stops_early():
    rng = np.random.default_rng(4)
    x = rng.integers(0, 256, (800, 3)).astype(np.uint8)
    _, _, _, _, it = M.fit(x, 2, 200, tol=1e-3)
    assert 1 < it < 200
