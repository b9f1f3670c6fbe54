"""CPU oracle of the colour GMMs (float64 numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the shipped
library never imports it.

Restates the reference's GMM (GMM.pyx:18-420, gmm.cl:33-432) in float64:
  constants  per component c_i = log w_i - 0.5 * ((log(2 pi v_i0) +
             log(2 pi v_i1)) + log(2 pi v_i2)), h_ij = 0.5 / v_ij (the
             reference folds the same terms into its matrix A, initA and em2
             gmm.cl:33-72, 241-306)
  score      per pixel -(m + log(sum_i exp(e_i - m))), e_i = c_i - ((d0^2 h_i0
             + d1^2 h_i1) + d2^2 h_i2), m = max e_i (eval / score_img2d,
             gmm.cl:314-432), every operation in the order the device uses
  unaries    rint(scale * min(max(score, 0), cap)), half to even
  fit        EM from initA's start (means at samples i (n-1)/(k-1), weights
             1/k, variances init_covar), responsibilities (em1
             gmm.cl:75-141), weight = sum r / n, mean = sum r x / sum r,
             var = sum r x^2 / sum r - mean^2 + min_covar (em2
             gmm.cl:183-310); a component with no responsibility keeps its
             mean and variance.  The reference's initA reads the green
             channel through the mask 0xFF0 (gmm.cl:53); the restatement
             takes the whole channel.
"""
import numpy as np

TWO_PI = 2.0 * np.pi


def constants(weights, means, var):
    w = np.asarray(weights, np.float64)
    mu = np.asarray(means, np.float64).reshape(-1, 3)
    v = np.asarray(var, np.float64).reshape(-1, 3)
    with np.errstate(divide="ignore"):
        c = np.log(w) - 0.5 * ((np.log(TWO_PI * v[:, 0]) + np.log(TWO_PI * v[:, 1]))
                               + np.log(TWO_PI * v[:, 2]))
    return c, mu, 0.5 / v


def colours(rgb):
    """float64 [n][3] colours of a uint8 [..][3 or 4] image or sample set."""
    a = np.asarray(rgb, np.uint8)
    return a.reshape(-1, a.shape[-1])[:, :3].astype(np.float64)


def terms(c, mu, h, x):
    d = x[:, None, :] - mu[None, :, :]
    t = (d * d) * h[None, :, :]
    return c[None, :] - ((t[..., 0] + t[..., 1]) + t[..., 2])


def lse(e):
    m = e.max(axis=1)
    s = np.zeros(e.shape[0])
    for i in range(e.shape[1]):
        s = s + np.exp(e[:, i] - m)
    return m + np.log(s)


def score(weights, means, var, rgb):
    """Negative log-likelihood of every colour (float64)."""
    c, mu, h = constants(weights, means, var)
    return -lse(terms(c, mu, h, colours(rgb)))


def quantise(s, scale, cap):
    return np.rint(scale * np.minimum(np.maximum(s, 0.0), cap)).astype(np.int32)


def unaries(fg, bg, rgb, scale, cap):
    """(source, sink) int32: the quantised scores under bg and fg; fg / bg
    are (weights, means, var)."""
    return quantise(score(*bg, rgb), scale, cap), quantise(score(*fg, rgb), scale, cap)


def fit(samples, k, iters, tol=0.0, init_covar=30.0, min_covar=0.001, init=None):
    """EM; returns (weights, means, var, mean log-likelihood of the last E
    step, iterations run)."""
    x = colours(samples)
    n = x.shape[0]
    if init is None:
        idx = [i * (n - 1) // (k - 1) if k > 1 else 0 for i in range(k)]
        w = np.full(k, 1.0 / k)
        mu = x[idx].copy()
        v = np.full((k, 3), float(init_covar))
    else:
        w, mu, v = (np.array(a, np.float64).reshape(s) for a, s in zip(init, ((k,), (k, 3), (k, 3))))
    ll, prev, it = np.nan, np.nan, 0
    while it < iters:
        c, m_, h = constants(w, mu, v)
        e = terms(c, m_, h, x)
        l = lse(e)
        ll = l.sum() / n
        r = np.exp(e - l[:, None])
        N = r.sum(axis=0)
        S = r.T @ x
        Q = r.T @ (x * x)
        w = N / n
        for i in range(k):
            if N[i] > 0:
                mu[i] = S[i] / N[i]
                v[i] = Q[i] / N[i] - mu[i] * mu[i] + min_covar
                v[i][~(v[i] > 0)] = min_covar
        it += 1
        if tol > 0:
            if it > 1 and abs(ll - prev) < tol:
                break
            prev = ll
    return w, mu, v, ll, it
