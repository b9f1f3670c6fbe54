"""GMM: a colour Gaussian mixture (diagonal covariances) fitted by EM and
scored per pixel on the device.

Drop-in for the reference's GMM (GMM.pyx:18-420): the same constructor,
fit(d_samples, nSamples, retParams) and score(d_population, dScore), with
the OpenCL context accepted and ignored.  Samples and populations are
packed RGBA uint32 device buffers, as in the reference (or uint8 [n][3|4]
host arrays, uploaded).  libgcmi works in float64 and keeps the model as
weights, means and variances (the reference packs them into the matrix A of
gmm.cl).  The scores are the negative log-likelihoods the reference writes;
GraphCut.upload_gmm turns a background and a foreground model's scores into
the cut's unaries.
"""
import ctypes

import numpy as np

from gcmi_lib import DeviceBuffer, Gmm, check, gmm, lib, ptr

INIT_COVAR = 30.0   # GMM.pyx:7
MIN_COVAR = 0.001   # GMM.pyx:8
TOL = 0.01          # change of log-likelihood taken as converged (GMM.pyx:328)


def _samples(d_samples, n=None):
    """(device pointer, channels, count, keep-alive) of samples."""
    if isinstance(d_samples, np.ndarray):
        a = np.ascontiguousarray(d_samples)
        if a.dtype == np.uint32:
            a = a.reshape(-1).view(np.uint8).reshape(-1, 4)
        a = np.ascontiguousarray(a.reshape(-1, a.shape[-1]), np.uint8)
        if a.shape[1] not in (3, 4):
            raise ValueError("samples: uint8 [n][3|4] or packed uint32 RGBA")
        buf = DeviceBuffer(max(a.nbytes, 4)).upload(a)
        return buf.ptr, a.shape[1], a.shape[0] if n is None else int(n), buf
    count = d_samples.size // 4 if n is None else int(n)
    return ptr(d_samples), 4, count, None


class GMM:
    def __init__(self, context=None, n_iterations=65, n_components=4, cap_samples=None,
                 min_covar=None, init_covar=None):
        self.n_iterations = int(n_iterations)
        self.n_components = int(n_components)
        self.cap_samples = cap_samples
        self.init_covar = float(init_covar) if init_covar else INIT_COVAR
        self.min_covar = float(min_covar) if min_covar else MIN_COVAR
        self.has_converged = False
        self.has_preset_wmc = False
        self.model = None          # gcmi_lib.Gmm of the current parameters
        self.logLiklihoods = []
        self.iterations = 0

    def set_params(self, weights, means, covars):
        """Preset weights, means and (diagonal) covariances; the next fit
        starts from them (the reference's has_preset_wmc)."""
        self.model = gmm(weights, means, covars)
        self.has_preset_wmc = True

    def params(self):
        """(weights, means, covars) of the current model."""
        return self.model.params()

    def fit(self, d_samples, nSamples=None, retParams=None):
        p, ch, n, keep = _samples(d_samples, nSamples)
        if n < self.n_components:
            raise ValueError("nSamples < nComp: {0}, {1}".format(n, self.n_components))
        warm = self.model is not None and (self.has_converged or self.has_preset_wmc)
        m = self.model if warm else Gmm()
        ll = ctypes.c_double()
        its = ctypes.c_int()
        check(lib().gcmi_gmm_fit(p, ch, n, self.n_components, self.n_iterations, TOL,
                                 self.init_covar, self.min_covar, 1 if warm else 0,
                                 ctypes.byref(m), ctypes.byref(ll), ctypes.byref(its)))
        del keep
        self.model = m
        self.iterations = its.value
        self.logLiklihoods = [ll.value]
        self.has_converged = its.value < self.n_iterations
        self.has_preset_wmc = False
        if retParams:
            return self.params()

    def score(self, d_population, dScore):
        """dScore[i] (float32) = negative log-likelihood of pixel i."""
        if self.model is None:
            raise ValueError("fit() or set_params() before score()")
        p, ch, n, keep = _samples(d_population)
        check(lib().gcmi_gmm_score(ctypes.byref(self.model), p, ch, n, ptr(dScore), None))
        check(lib().gcmi_stream_sync(None))
        del keep
