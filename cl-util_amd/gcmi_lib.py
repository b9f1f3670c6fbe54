"""ctypes binding of libgcmi.so and device-buffer plumbing.

The library is the only compute path: there is no CPU fallback.  If the
shared object is missing, importing the drop-in classes raises
GcmiUnavailable; if no GPU is visible, every compute call raises
GcmiError(ENODEVICE).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgcmi.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "gcmi.h")

OK = 0
EINVAL = -1
ENOMEM = -2
ENODEVICE = -3
ENOTCONVERGED = -4
EHIP = -5
EBUDGET = -6
EBUSY = -7

_NAMES = {EINVAL: "EINVAL", ENOMEM: "ENOMEM", ENODEVICE: "ENODEVICE",
          ENOTCONVERGED: "ENOTCONVERGED", EHIP: "EHIP", EBUDGET: "EBUDGET",
          EBUSY: "EBUSY"}


class GcmiUnavailable(ImportError):
    """libgcmi.so is not built."""


class GcmiError(RuntimeError):
    def __init__(self, code, message):
        self.code = code
        super().__init__("%s (%d): %s" % (_NAMES.get(code, "?"), code, message))


class NotConverged(GcmiError):
    pass


_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p
_int = ctypes.c_int
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t

# name -> argtypes; every function returns int unless listed in _RESTYPE
SIGNATURES = {
    "gcmi_last_error": [],
    "gcmi_version": [],
    "gcmi_device_count": [ctypes.POINTER(_int)],
    "gcmi_malloc": [ctypes.POINTER(_vp), _sz],
    "gcmi_free": [_vp],
    "gcmi_host_alloc": [ctypes.POINTER(_vp), _sz],
    "gcmi_host_free": [_vp],
    "gcmi_memcpy_h2d": [_vp, _vp, _sz],
    "gcmi_memcpy_d2h": [_vp, _vp, _sz],
    "gcmi_memcpy_d2d": [_vp, _vp, _sz],
    "gcmi_memset_i32": [_vp, _i32, _sz, _vp],
    "gcmi_stream_sync": [_vp],
    "gcmi_device_sync": [],
    "gcmi_prefix_sum": [_vp, _vp, _int, _vp],
    "gcmi_prefix_sum_timed": [_vp, _vp, _int, _vp, _i64p],
    "gcmi_stream_compact": [_vp, _vp, _vp, _int, _vp],
    "gcmi_tiles_fill": [_vp, _int, _i32, _vp],
    "gcmi_tiles_flag": [_vp, _vp, _int, _int, _i32, _vp],
    "gcmi_tiles_flag_logical": [_vp, _vp, _vp, _int, _int, _int, _i32, _i32, _int, _vp],
    "gcmi_tiles_increment_logical": [_vp, _vp, _int, _int, _int, _i32, _i32, _int, _i32, _vp],
    "gcmi_tiles_build": [_vp, _int, _int, _i32, _vp, _vp, _vp],
    "gcmi_tiles_build_noalloc": [_vp, _int, _int, _i32, _vp, _vp, _vp],
    "gcmi_tiles_build_logical": [_vp, _vp, _int, _int, _int, _i32, _i32, _int, _vp, _vp, _vp],
    "gcmi_cut_create": [_int, _int, _int, ctypes.POINTER(_vp)],
    "gcmi_cut_destroy": [_vp],
    "gcmi_cut_upload": [_vp, _vp, _vp, _vp, _vp],
    "gcmi_cut_upload_device": [_vp, _vp, _vp, _vp, _vp],
    "gcmi_cut_run": [_vp, _int, _int, _i64, _i64p],
    "gcmi_cut_labels": [_vp, _vp],
    "gcmi_cut_push": [_vp, _int],
    "gcmi_cut_relabel": [_vp, _i32p],
    "gcmi_cut_start_bfs": [_vp, _i32p],
    "gcmi_cut_is_completed": [_vp, ctypes.POINTER(_int), _i64p],
    "gcmi_cut_tile_flags": [_vp, _int, _vp],
    "gcmi_cut_labels_device": [_vp, _vp],
    "gcmi_cut_stats": [_vp, _i64p, _int],
    "gcmi_cut_heights": [_vp, _vp],
    "gcmi_cut_geometry": [_vp, ctypes.POINTER(_int)],
    "gcmi_cut_field": [_vp, _int, _vp],
    "gcmi_cut_profile": [_vp, _int],
    "gcmi_cut_kernel_times": [_vp, ctypes.POINTER(ctypes.c_double), _i64p, _int],
    "gcmi_cut_edit": [_vp, _int, _vp, _vp, _vp],
    "gcmi_cut_paint": [_vp, _int, _vp, ctypes.c_double, _i32, _i32],
    "gcmi_cut_labels_packed": [_vp, _vp],
    "gcmi_cut_forget": [_vp],
    "gcmi_cut_set_option": [_vp, _int, _int],
    "gcmi_batch_set_option": [_vp, _int, _int],
    "gcmi_server_set_option": [_vp, _int, _int],
    "gcmi_batch_paint": [_vp, _int, _int, _vp, ctypes.c_double, _i32, _i32],
    "gcmi_server_submit_polyline": [_vp, _int, _int, _vp, ctypes.c_double, _i32, _i32, _i64p],
    "gcmi_server_replay": [_vp, _int, _vp, _i64p],
    "gcmi_batch_create": [_int, _vp, _vp, _int, ctypes.POINTER(_vp)],
    "gcmi_batch_destroy": [_vp],
    "gcmi_batch_upload": [_vp, _int, _vp, _vp, _vp, _vp],
    "gcmi_batch_upload_device": [_vp, _int, _vp, _vp, _vp, _vp],
    "gcmi_batch_edit": [_vp, _int, _int, _vp, _vp, _vp],
    "gcmi_batch_run": [_vp, _int, _int, _i64, _i64p],
    "gcmi_batch_run_subset": [_vp, _int, _vp, _int, _int, _i64, _i64p],
    "gcmi_batch_labels": [_vp, _int, _vp],
    "gcmi_batch_stats": [_vp, _i64p, _int],
    "gcmi_batch_profile": [_vp, _int],
    "gcmi_batch_kernel_times": [_vp, ctypes.POINTER(ctypes.c_double), _i64p, _int],
    "gcmi_server_create": [_int, _sz, _int, _int, _int, ctypes.POINTER(_vp)],
    "gcmi_server_create_lanes": [_int, _sz, _int, _int, _int, _int, ctypes.POINTER(_vp)],
    "gcmi_server_destroy": [_vp],
    "gcmi_server_set_params": [_vp, _int, _int],
    "gcmi_server_open": [_vp, _int, _int, ctypes.POINTER(_int)],
    "gcmi_server_close": [_vp, _int],
    "gcmi_server_submit_image": [_vp, _int, _vp, _vp, _vp, _vp, _i64p],
    "gcmi_server_submit_stroke": [_vp, _int, _int, _vp, _vp, _vp, _i64p],
    "gcmi_server_wait": [_vp, _i64, _int, _i64p],
    "gcmi_server_mask": [_vp, _int, _vp, _i64p],
    "gcmi_server_stats": [_vp, _i64p, _int],
    "gcmi_server_forget": [_vp, _i64],
    "gcmi_rgb_model_default": [_vp],
    "gcmi_cut_upload_rgb": [_vp, _vp, _int, _vp],
    "gcmi_cut_capacities": [_vp, _vp, _vp, _vp, _vp],
    "gcmi_batch_upload_rgb": [_vp, _int, _vp, _int, _vp],
    "gcmi_server_set_rgb_model": [_vp, _vp],
    "gcmi_server_submit_rgb": [_vp, _int, _vp, _int, _i64p],
    "gcmi_tiles_filter": [_vp, _int, _int, _int, _int, _int, _int, _i32, _vp, _vp],
    "gcmi_growcut_create": [_int, _int, _vp, _int, _int, ctypes.POINTER(_vp)],
    "gcmi_growcut_destroy": [_vp],
    "gcmi_growcut_label": [_vp, _int, _vp, _int],
    "gcmi_growcut_evolve": [_vp, _int, ctypes.POINTER(_int), ctypes.POINTER(_int)],
    "gcmi_growcut_read": [_vp, _vp, _vp, _vp, ctypes.POINTER(_i32)],
    "gcmi_gmm_score": [_vp, _vp, _int, ctypes.c_longlong, _vp, _vp],
    "gcmi_gmm_fit": [_vp, _int, ctypes.c_longlong, _int, _int, ctypes.c_double, ctypes.c_double,
                     ctypes.c_double, _int, _vp, ctypes.POINTER(ctypes.c_double),
                     ctypes.POINTER(_int)],
    "gcmi_cut_upload_rgb_gmm": [_vp, _vp, _int, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double],
    "gcmi_batch_upload_rgb_gmm": [_vp, _int, _vp, _int, _vp, _vp, _vp, ctypes.c_double,
                                  ctypes.c_double],
}

REQ_IMAGE, REQ_STROKE, REQ_POLYLINE, REQ_RGB = 0, 1, 2, 3
PAIR_EXP, PAIR_INV = 0, 1
OPT_LOCAL_RELABEL = 1
OPT_DEVICE_ROUNDS = 2
OPT_GRAPH = 3
TILES_ACTIVE, TILES_BORDER, TILES_CHANGED, TILES_EXCESS = 0, 1, 2, 3
FIELDS = {"height": 0, "excess": 1, "up": 2, "down": 3, "left": 4, "right": 5}


class RgbModel(ctypes.Structure):
    """struct gcmi_rgb_model of gcmi.h: capacities from an 8-bit colour image."""
    _fields_ = [("mu_bg", ctypes.c_double * 3), ("mu_fg", ctypes.c_double * 3),
                ("unary_scale", ctypes.c_double), ("unary_div", ctypes.c_double),
                ("unary_cap", ctypes.c_double), ("pair_kind", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("pair_scale", ctypes.c_double),
                ("inv_beta", ctypes.c_double), ("inv_eps", ctypes.c_double)]


def rgb_model(**fields):
    """The library's default colour model with the given fields replaced
    (mu_bg / mu_fg as 3-sequences, pair_kind PAIR_EXP or PAIR_INV)."""
    m = RgbModel()
    lib().gcmi_rgb_model_default(ctypes.byref(m))
    for k, v in fields.items():
        if k in ("mu_bg", "mu_fg"):
            getattr(m, k)[:] = [float(x) for x in v]
        elif not hasattr(m, k):
            raise TypeError("unknown colour-model field %r" % k)
        else:
            setattr(m, k, v)
    return m


def rgb_image(image):
    """A C-contiguous uint8 [H][W][3 or 4] array and its channel count."""
    a = np.ascontiguousarray(image, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] not in (3, 4):
        raise ValueError("expected a [H][W][3] or [H][W][4] uint8 image, got shape %s"
                         % (a.shape,))
    return a, a.shape[2]


GMM_MAX_COMPONENTS = 8


class Gmm(ctypes.Structure):
    """struct gcmi_gmm of gcmi.h: a colour mixture with diagonal covariances."""
    _fields_ = [("k", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("weight", ctypes.c_double * GMM_MAX_COMPONENTS),
                ("mean", (ctypes.c_double * 3) * GMM_MAX_COMPONENTS),
                ("var", (ctypes.c_double * 3) * GMM_MAX_COMPONENTS)]

    def params(self):
        """(weights [k], means [k][3], variances [k][3]) as float64 arrays."""
        k = self.k
        return (np.array(self.weight[:k]), np.array([list(m) for m in self.mean[:k]]),
                np.array([list(v) for v in self.var[:k]]))


def gmm(weights, means, var):
    """A Gmm from k weights and [k][3] means and diagonal variances."""
    w = np.asarray(weights, np.float64).reshape(-1)
    mu = np.asarray(means, np.float64).reshape(-1, 3)
    v = np.asarray(var, np.float64).reshape(-1, 3)
    if not (1 <= w.size <= GMM_MAX_COMPONENTS) or mu.shape[0] != w.size or v.shape[0] != w.size:
        raise ValueError("a GMM has 1..%d components with a mean and variance each"
                         % GMM_MAX_COMPONENTS)
    g = Gmm()
    g.k = w.size
    for i in range(w.size):
        g.weight[i] = w[i]
        g.mean[i][:] = [float(x) for x in mu[i]]
        g.var[i][:] = [float(x) for x in v[i]]
    return g


class Request(ctypes.Structure):
    """struct gcmi_request of gcmi.h (one request of a replayed trace)."""
    _fields_ = [("t_ns", ctypes.c_int64), ("session", ctypes.c_int32),
                ("kind", ctypes.c_int32), ("n", ctypes.c_int32),
                ("source", ctypes.c_int32), ("sink", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("width", ctypes.c_double),
                ("p0", ctypes.c_void_p), ("p1", ctypes.c_void_p),
                ("p2", ctypes.c_void_p), ("p3", ctypes.c_void_p)]


KERNEL_CLASSES = ("push_first", "push", "fold", "bfs_init", "bfs_round", "flood",
                  "labels", "compact")
_RESTYPE = {"gcmi_last_error": ctypes.c_char_p, "gcmi_version": ctypes.c_char_p,
            "gcmi_rgb_model_default": None}

_lib = None


def lib():
    """The loaded library (raises GcmiUnavailable if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GcmiUnavailable(
                "libgcmi.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def last_error():
    msg = lib().gcmi_last_error()
    return msg.decode() if msg else ""


def check(rc):
    if rc != OK:
        cls = NotConverged if rc == ENOTCONVERGED else GcmiError
        raise cls(rc, last_error())
    return rc


def device_count():
    n = ctypes.c_int()
    check(lib().gcmi_device_count(ctypes.byref(n)))
    return n.value


def ptr(obj):
    """Device address of a DeviceBuffer, a torch tensor or an int."""
    if obj is None:
        return None
    if isinstance(obj, int):
        return obj
    if hasattr(obj, "ptr"):
        return obj.ptr
    if hasattr(obj, "data_ptr"):
        return obj.data_ptr()
    raise TypeError("not a device buffer: %r" % type(obj))


def host_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _PinnedBlock:
    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().gcmi_host_alloc(ctypes.byref(p), max(1, int(nbytes))))
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().gcmi_host_free(ctypes.c_void_p(self.ptr))
            self.ptr = None


def pinned_empty(shape, dtype):
    """A numpy array in page-locked host memory (freed with the array)."""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    blk = _PinnedBlock(n)
    buf = (ctypes.c_uint8 * max(1, n)).from_address(blk.ptr)
    buf._gcmi_block = blk  # the memory lives as long as the buffer
    return np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)


def pinned_copy(a):
    """A C-contiguous copy of `a` in page-locked host memory."""
    a = np.asarray(a)
    out = pinned_empty(a.shape, a.dtype)
    out[...] = a
    return out


class DeviceBuffer:
    """Owned device allocation (the role cl.Buffer plays in the reference)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().gcmi_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value

    @property
    def size(self):
        return self.nbytes

    def upload(self, array):
        a = np.ascontiguousarray(array)
        if a.nbytes > self.nbytes:
            raise ValueError("host array larger than buffer")
        check(lib().gcmi_memcpy_h2d(self.ptr, host_ptr(a), a.nbytes))
        return self

    def download(self, dtype, count=None):
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("readback larger than buffer")
        check(lib().gcmi_memcpy_d2h(host_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            if _lib is not None:
                _lib.gcmi_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
