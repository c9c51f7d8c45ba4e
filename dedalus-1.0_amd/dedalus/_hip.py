"""ctypes binding to libddx.so, the MI355X pseudospectral step.

The library is built in-tree (``build_ddx.py`` next to this package) and
loaded from this directory.  There is no CPU fallback: every entry point
raises ``DDXError`` when the library or a HIP device is missing.

Device buffers are owned by ``DevBuf`` objects (freed when collected); the
library never frees caller memory.  ``Plan`` wraps one grid (shape, box
lengths) and owns its stream and workspace.
"""

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libddx.so")

# enums of include/ddx.h
F64, F32 = 0, 1
# dealiasing rules (ddx_plan_create_ex)
DEALIAS_23, DEALIAS_SPHERICAL, DEALIAS_NONE = 0, 1, 2
HYDRO, BOUSSINESQ, MHD = 0, 1, 2
RK2MID, RK2TRAP = 0, 1
TUNE_WIDE_TILES, TUNE_XCD_ORDER, TUNE_FUSED_YEPI, TUNE_RADIX = 0, 1, 2, 3
EULER, ETD1, ETD2RK1, ETD2RK2 = 0, 1, 2, 3

ERR_NAMES = {-1: "DDX_ERR_ARG", -2: "DDX_ERR_HIP", -3: "DDX_ERR_UNSUPPORTED",
             -4: "DDX_ERR_NOMEM", -5: "DDX_ERR_NODEV"}


class DDXError(RuntimeError):
    """Failure reported by libddx (or the library is missing)."""

    def __init__(self, code, msg):
        self.code = code
        RuntimeError.__init__(self, "%s: %s" % (ERR_NAMES.get(code, code), msg))


class Params(ctypes.Structure):
    """Mirror of ddx_params."""
    _fields_ = [("physics", ctypes.c_int),
                ("viscosity_order", ctypes.c_int),
                ("nu", ctypes.c_double),
                ("kappa", ctypes.c_double),
                ("eta", ctypes.c_double),
                ("g", ctypes.c_double),
                ("alpha_t", ctypes.c_double),
                ("beta", ctypes.c_double),
                ("bdir", ctypes.c_int),
                ("rho0", ctypes.c_double)]


# name -> (restype, argtypes); every exported symbol of include/ddx.h and
# include/ddx_closure.h
_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SZ = ctypes.c_size_t
_PP = ctypes.POINTER(ctypes.c_void_p)
_I64P = ctypes.POINTER(ctypes.c_int64)
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

SIGNATURES = {
    "ddx_init": (_I, [_I]),
    "ddx_last_error": (ctypes.c_char_p, []),
    "ddx_version": (_I, []),
    "ddx_plan_create": (_I, [_I, _I64P, _DP, _I, _PP]),
    "ddx_plan_create_ex": (_I, [_I, _I64P, _DP, _I, _I, _PP]),
    "ddx_dealias": (_I, [_VP, _VP]),
    "ddx_bwd_combine": (_I, [_VP, _I, _PP, _I, _IP, _IP, _IP, _DP, _PP]),
    "ddx_plan_create_batched": (_I, [_I, ctypes.c_int64, ctypes.c_int64, _DP, _I, _PP]),
    "ddx_plan_destroy": (_I, [_VP]),
    "ddx_plan_kshape": (_I, [_VP, _I64P, _I64P, _I64P, _I64P]),
    "ddx_plan_ktable": (_I, [_VP, _I, _DP, ctypes.c_int64]),
    "ddx_plan_set_stream": (_I, [_VP, _VP]),
    "ddx_plan_set_graph": (_I, [_VP, _I]),
    "ddx_plan_set_tuning": (_I, [_VP, _I, _I]),
    "ddx_alloc": (_I, [_SZ, _PP]),
    "ddx_free": (_I, [_VP]),
    "ddx_h2d": (_I, [_VP, _VP, _VP, _SZ]),
    "ddx_d2h": (_I, [_VP, _VP, _VP, _SZ]),
    "ddx_zero": (_I, [_VP, _VP, _SZ]),
    "ddx_copy": (_I, [_VP, _VP, _VP, _SZ]),
    "ddx_sync": (_I, [_VP]),
    "ddx_pack": (_I, [_VP, _VP, _VP]),
    "ddx_unpack": (_I, [_VP, _VP, _VP]),
    "ddx_fwd": (_I, [_VP, _VP, _VP]),
    "ddx_bwd": (_I, [_VP, _VP, _VP]),
    "ddx_fwd_many": (_I, [_VP, _I, _PP, _PP]),
    "ddx_bwd_many": (_I, [_VP, _I, _PP, _PP]),
    "ddx_deriv": (_I, [_VP, _VP, _VP, _I]),
    "ddx_project": (_I, [_VP, _PP]),
    "ddx_update": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _D, _I, _D]),
    "ddx_update_if": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP, _D]),
    "ddx_rhs": (_I, [_VP, ctypes.POINTER(Params), _PP, _PP, _PP]),
    "ddx_step": (_I, [_VP, _I, ctypes.POINTER(Params), _PP, _PP, _PP, _PP, _PP, _I, _D]),
    "ddx_step_noalloc": (_I, [_VP, _I, ctypes.POINTER(Params), _PP, _PP, _PP, _PP, _PP, _I, _D]),
    "ddx_step_xforced": (_I, [_VP, _I, ctypes.POINTER(Params), _PP, _PP, _PP, _PP, _PP, _PP, _I, _D]),
    "ddx_step_xin": (_I, [_VP, _I, ctypes.POINTER(Params), _PP, _PP, _PP, _PP, _PP, _PP, _PP, _D]),
    "ddx_kcombine": (_I, [_VP, _VP, _I, _PP, _DP, ctypes.POINTER(_I), _I]),
    "ddx_xcombine": (_I, [_VP, _VP, _I, _PP, _PP, _DP, _I]),
    "ddx_xreduce": (_I, [_VP, _VP, _I, _DP]),
    "ddx_kreduce": (_I, [_VP, _VP, _VP, _I, _DP]),
    "ddx_kreduce_members": (_I, [_VP, _VP, _VP, _I, _VP]),
    "ddx_prof_enable": (_I, [_VP, _I]),
    "ddx_prof_read": (_I, [_VP, _I, ctypes.POINTER(_I), ctypes.c_char_p, _I64P, _DP, _DP]),
    # include/ddx_closure.h
    "ddx_conv3x3": (_I, [_VP, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _I, _VP, _VP]),
    "ddx_conv3x3_wgrad": (_I, [_VP, _I, _I, _I, _I, _I, _VP, _VP, _VP, _I, _VP]),
    "ddx_conv_colsum": (_I, [_VP, _VP, ctypes.c_int64, _I, _VP, _I, _VP]),
    "ddx_conv3x3_grads": (_I, [_VP, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _I, _VP, _VP]),
    "ddx_conv_wprep": (_I, [_VP, _VP, _I, _I, _I, _I, _VP, _VP]),
    "ddx_conv_wunpack": (_I, [_VP, _VP, _I, _I, _I, _I, _VP]),
    "ddx_conv_pack": (_I, [_VP, ctypes.c_int64, _I, _I, _PP, ctypes.c_int64, _I, _VP, _VP, _VP]),
    "ddx_conv_unpack": (_I, [_VP, ctypes.c_int64, _I, _I, _VP, _PP, ctypes.c_int64, _I, _VP, _PP]),
    "ddx_box_filter": (_I, [_VP, _I, _I, _I, _I, _VP, _VP]),
    "ddx_conv_net": (_I, [_VP, _I, _I, _I, _I, ctypes.POINTER(_I), _PP, _PP, _VP, _VP, _PP, _VP,
                          _PP, _PP, _VP]),
    "ddx_mfma_probe": (_I, [_VP, _I, _VP, _VP, _VP]),
}

# enums of include/ddx_closure.h
CONV_LINEAR, CONV_GELU, CONV_GELU_GRAD = 0, 1, 2
CONV_MAXCH = 16

_lib = None
_lock = threading.Lock()
_device = None


def load_library(path=LIB_PATH):
    """Load libddx.so and declare its signatures (no device needed).

    torch, when installed, is imported first: its wheel carries its own HIP
    runtime under the same soname (libamdhip64.so.7) as the system one that
    libddx.so links, and the first one loaded serves the whole process.
    Loading the system runtime first would leave torch a second HIP/HSA
    runtime that finds no device."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise DDXError(-5, "libddx.so is not built (%s); run build_ddx.py" % path)
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            lib = ctypes.CDLL(path)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise DDXError(rc, _lib.ddx_last_error().decode())


def init(device=0):
    """Select the HIP device once per process."""
    global _device
    lib = load_library()
    if _device is None:
        _check(lib.ddx_init(int(device)))
        _device = int(device)
    return _device


def ptr_array(ptrs):
    """ctypes void* array from ints / None."""
    arr = (ctypes.c_void_p * max(len(ptrs), 1))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


def compact_real(ref, kref_shape, kc_shape, cut):
    """Reference-layout real k array -> compact float64 array of
    Ky * Kz * Kxp entries (the modes ddx_pack keeps, in its order; zero on
    the padding columns kx index > cut[0]).  Reference layouts: (kx, ky) in
    2D, (ky, kz, kx) in 3D (representations.py:204-233)."""
    ref = np.asarray(ref, dtype=np.float64)
    if ref.shape != tuple(kref_shape):
        raise ValueError("k array shape %s, expected %s" % (ref.shape, tuple(kref_shape)))
    Ky, Kz, Kxp = kc_shape
    cx, cy, cz = cut
    ndim = len(kref_shape)
    ny = kref_shape[1] if ndim == 2 else kref_shape[0]
    slab = np.arange(Ky)
    fy = np.where(slab <= cy, slab, slab + (ny - Ky))
    col = np.arange(min(Kxp, cx + 1))
    out = np.zeros((Ky, Kz, Kxp))
    if ndim == 2:
        out[:, 0, :col.size] = ref[np.ix_(col, fy)].T
    else:
        nz = kref_shape[1]
        row = np.arange(Kz)
        fz = np.where(row <= cz, row, row + (nz - Kz))
        out[:, :, :col.size] = ref[np.ix_(fy, fz, col)]
    return out.reshape(-1)


class DevBuf:
    """Device allocation of nbytes (zero-filled on request)."""

    def __init__(self, nbytes):
        init()
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(_lib.ddx_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.ddx_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __int__(self):
        return self.ptr


class Plan:
    """One grid: xshape in reference order, box lengths, precision.

    With ``batch=B`` (2D only) the plan advances B independent members of
    the grid at once: x arrays are (B, ny, nx), reference k arrays
    (B, nx/2+1, ny), compact k buffers hold all members."""

    def __init__(self, shape, length, precision=F64, batch=None, dealias=DEALIAS_23):
        init()
        self.shape = tuple(int(n) for n in shape)
        self.length = tuple(float(x) for x in length)
        self.ndim = len(self.shape)
        self.precision = precision
        self.batch = int(batch) if batch else None
        self.dealias_rule = int(dealias)
        if self.batch and self.dealias_rule != DEALIAS_23:
            raise ValueError("batched plans use the 2/3 rule")
        ln = (ctypes.c_double * self.ndim)(*self.length)
        p = ctypes.c_void_p()
        if self.batch:
            if self.ndim != 2:
                raise ValueError("batched plans are 2D")
            _check(_lib.ddx_plan_create_batched(self.batch, self.shape[0], self.shape[1], ln,
                                                precision, ctypes.byref(p)))
        else:
            sh = (ctypes.c_int64 * self.ndim)(*self.shape)
            _check(_lib.ddx_plan_create_ex(self.ndim, sh, ln, precision, self.dealias_rule,
                                           ctypes.byref(p)))
        self.handle = p
        kref = (ctypes.c_int64 * 3)()
        kc = (ctypes.c_int64 * 3)()
        cut = (ctypes.c_int64 * 3)()
        nkc = ctypes.c_int64()
        _check(_lib.ddx_plan_kshape(p, kref, kc, cut, ctypes.byref(nkc)))
        if self.batch:
            self.kref_shape = (kref[2], kref[0], kref[1])
            self.xshape = (self.batch,) + self.shape
        else:
            self.kref_shape = tuple(kref[i] for i in range(2 if self.ndim == 2 else 3))
            self.xshape = self.shape
        self.kc_shape = tuple(kc[i] for i in range(3))
        self.cut = tuple(cut[i] for i in range(3))
        self.nkc = nkc.value
        self.npts = int(np.prod(self.xshape))
        self.nref = int(np.prod(self.kref_shape))
        self.cbytes = 16 if precision == F64 else 8
        self.rbytes = 8 if precision == F64 else 4
        self.ktables = []
        for ax, n in ((0, self.cut[0] + 1), (1, self.kc_shape[0]), (2, self.kc_shape[1])):
            buf = np.zeros(n, dtype=np.float64)
            _check(_lib.ddx_plan_ktable(p, ax, buf.ctypes.data_as(_DP), n))
            self.ktables.append(buf)
        self._stage = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.ddx_plan_destroy(h)
            self.handle = None

    # --- buffers --------------------------------------------------------
    def alloc_k(self, zero=True):
        b = DevBuf(self.nkc * self.cbytes)
        if zero:
            self.zero(b)
        return b

    def alloc_x(self, zero=True):
        b = DevBuf(self.npts * self.rbytes)
        if zero:
            self.zero(b)
        return b

    def staging(self):
        """Device buffer for one reference-layout k array (layout conversion)."""
        if self._stage is None:
            self._stage = DevBuf(self.nref * 16)
        return self._stage

    def zero(self, buf):
        _check(_lib.ddx_zero(self.handle, int(buf), buf.nbytes))

    def sync(self):
        _check(_lib.ddx_sync(self.handle))

    def set_graph(self, on):
        _check(_lib.ddx_plan_set_graph(self.handle, 1 if on else 0))

    def set_tuning(self, knob, value):
        """A kernel choice (TUNE_*): bit-identical results, except FP64 tile
        widths and the radix plan (TUNE_RADIX), equal to rounding; -1
        restores the built-in rule."""
        _check(_lib.ddx_plan_set_tuning(self.handle, int(knob), int(value)))

    def set_stream(self, stream):
        """Launch on an existing hipStream_t (an int handle); None or 0 restores
        the plan's own stream."""
        _check(_lib.ddx_plan_set_stream(self.handle, ctypes.c_void_p(int(stream) if stream else None)))

    def copy(self, dst, src, nbytes=None):
        n = min(dst.nbytes, src.nbytes) if nbytes is None else int(nbytes)
        _check(_lib.ddx_copy(self.handle, int(dst), int(src), n))

    # --- instrumentation --------------------------------------------------
    def prof_enable(self, on=True):
        _check(_lib.ddx_prof_enable(self.handle, 1 if on else 0))

    def prof_read(self):
        """{kernel class: (launches, total ms, total algorithmic bytes)}."""
        m = 32
        n = ctypes.c_int()
        names = ctypes.create_string_buffer(32 * m)
        launches = (ctypes.c_int64 * m)()
        ms = (ctypes.c_double * m)()
        by = (ctypes.c_double * m)()
        _check(_lib.ddx_prof_read(self.handle, m, ctypes.byref(n), names, launches, ms, by))
        out = {}
        for i in range(n.value):
            if launches[i]:
                nm = names.raw[32 * i:32 * i + 32].split(b"\0", 1)[0].decode()
                out[nm] = (int(launches[i]), float(ms[i]), float(by[i]))
        return out

    def h2d(self, buf, arr):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > buf.nbytes:
            raise ValueError("host array larger than device buffer")
        _check(_lib.ddx_h2d(self.handle, int(buf), arr.ctypes.data, arr.nbytes))

    def d2h(self, arr, buf):
        if not arr.flags.c_contiguous or arr.nbytes > buf.nbytes:
            raise ValueError("host array must be contiguous and fit the buffer")
        _check(_lib.ddx_d2h(self.handle, arr.ctypes.data, int(buf), arr.nbytes))

    # --- layout -----------------------------------------------------------
    def upload_k(self, kc, kref):
        """Reference-layout complex array (host) -> compact device buffer."""
        kref = np.ascontiguousarray(kref, dtype=np.complex128)
        if kref.shape != self.kref_shape:
            raise ValueError("k array shape %s, expected %s" % (kref.shape, self.kref_shape))
        st = self.staging()
        self.h2d(st, kref)
        _check(_lib.ddx_pack(self.handle, int(st), int(kc)))

    def download_k(self, kc, out=None):
        """Compact device buffer -> reference-layout complex array (host)."""
        if out is None:
            out = np.empty(self.kref_shape, dtype=np.complex128)
        st = self.staging()
        _check(_lib.ddx_unpack(self.handle, int(kc), int(st)))
        self.d2h(out, st)
        return out

    def upload_x(self, xd, x):
        """Real x-space array (host) -> device, in the plan's precision."""
        x = np.ascontiguousarray(x, dtype=np.float64 if self.precision == F64 else np.float32)
        if x.shape != self.xshape:
            raise ValueError("x array shape %s, expected %s" % (x.shape, self.xshape))
        self.h2d(xd, x)

    def download_x(self, xd, out=None):
        """Device x-space array -> float64 host array."""
        if out is None:
            out = np.empty(self.xshape, dtype=np.float64)
        if self.precision == F64:
            self.d2h(out, xd)
        else:
            tmp = np.empty(self.xshape, dtype=np.float32)
            self.d2h(tmp, xd)
            out[...] = tmp
        return out

    # --- numerics ---------------------------------------------------------
    def fwd(self, x, kc):
        _check(_lib.ddx_fwd(self.handle, int(x), int(kc)))

    def bwd(self, kc, x):
        _check(_lib.ddx_bwd(self.handle, int(kc), int(x)))

    def fwd_many(self, xs, kcs):
        """fwd of several fields in one call (one launch per pass)."""
        n = len(xs)
        a = (ctypes.c_void_p * n)(*[int(b) for b in xs])
        b = (ctypes.c_void_p * n)(*[int(b) for b in kcs])
        _check(_lib.ddx_fwd_many(self.handle, n, a, b))

    def bwd_many(self, kcs, xs):
        """bwd of several fields in one call (one launch per pass)."""
        n = len(kcs)
        a = (ctypes.c_void_p * n)(*[int(b) for b in kcs])
        b = (ctypes.c_void_p * n)(*[int(b) for b in xs])
        _check(_lib.ddx_bwd_many(self.handle, n, a, b))

    def deriv(self, kc, out, dim):
        _check(_lib.ddx_deriv(self.handle, int(kc), int(out), int(dim)))

    def project(self, comps):
        _check(_lib.ddx_project(self.handle, ptr_array([int(c) for c in comps])))

    def update(self, kind, y, out, F1, F2, coeff, order, dt):
        _check(_lib.ddx_update(self.handle, kind, int(y), int(out), int(F1),
                               int(F2) if F2 is not None else None,
                               float(coeff), int(order), float(dt)))

    def bwd_combine(self, kcs, recipes, xs):
        """x-space outputs of linear combinations of compact k buffers:
        recipes[q] = [(src index, dim (-1: none, 0/1/2: i k_x/y/z), coef), ...]
        with one or two terms; xs[q] receives the backward transform of the
        q-th combination."""
        n = len(recipes)
        nterm = (ctypes.c_int * n)(*[len(r) for r in recipes])
        src = (ctypes.c_int * (2 * n))()
        dim = (ctypes.c_int * (2 * n))()
        coef = (ctypes.c_double * (2 * n))()
        for q, r in enumerate(recipes):
            for t, (s_, d_, c_) in enumerate(r):
                src[2 * q + t], dim[2 * q + t], coef[2 * q + t] = int(s_), int(d_), float(c_)
        _check(_lib.ddx_bwd_combine(self.handle, len(kcs), ptr_array([int(k) for k in kcs]), n,
                                    nterm, src, dim, coef, ptr_array([int(x) for x in xs])))

    def dealias(self, kc):
        """The plan's dealiasing rule in place on a compact k buffer."""
        _check(_lib.ddx_dealias(self.handle, int(kc)))

    def update_if(self, kind, y, out, F1, F2, intfactor, dt):
        """The sweep of ``update`` with a per-mode integrating factor: a
        device buffer of nkc float64 values in the compact layout
        (``compact_real`` of the reference-layout array)."""
        _check(_lib.ddx_update_if(self.handle, kind, int(y), int(out), int(F1),
                                  int(F2) if F2 is not None else None, int(intfactor), float(dt)))

    def upload_real_k(self, ref, buf=None):
        """Reference-layout real k array (host) -> compact float64 device
        buffer (a new one unless ``buf`` is given)."""
        if self.batch:
            raise ValueError("per-mode k arrays are gathered for plain plans only")
        host = compact_real(ref, self.kref_shape, self.kc_shape, self.cut)
        if buf is None:
            buf = DevBuf(host.nbytes)
        self.h2d(buf, host)
        return buf

    def kcombine(self, out, terms, accumulate=False):
        """out (+)= sum coef * M(op) * src over (src, coef, op) terms (op -1, 0..2, 3)."""
        n = len(terms)
        src = ptr_array([int(t[0]) for t in terms])
        coef = (ctypes.c_double * max(n, 1))(*[float(t[1]) for t in terms])
        op = (ctypes.c_int * max(n, 1))(*[int(t[2]) for t in terms])
        _check(_lib.ddx_kcombine(self.handle, int(out), n, src, coef, op, 1 if accumulate else 0))

    def xcombine(self, out, terms, accumulate=False):
        """out (+)= sum coef * a * b over (a, b or None, coef) terms (x space)."""
        n = len(terms)
        a = ptr_array([int(t[0]) for t in terms])
        b = ptr_array([int(t[1]) if t[1] is not None else None for t in terms])
        coef = (ctypes.c_double * max(n, 1))(*[float(t[2]) for t in terms])
        _check(_lib.ddx_xcombine(self.handle, int(out), n, a, b, coef, 1 if accumulate else 0))

    def xreduce(self, x, op):
        r = ctypes.c_double()
        _check(_lib.ddx_xreduce(self.handle, int(x), int(op), ctypes.byref(r)))
        return r.value

    def kreduce(self, k0, op, k1=None):
        r = ctypes.c_double()
        _check(_lib.ddx_kreduce(self.handle, int(k0), int(k1) if k1 is not None else None,
                                int(op), ctypes.byref(r)))
        return r.value

    def kreduce_members(self, k0, op, k1=None, out=None):
        """Per-member reductions of a 2D plan into a device buffer of
        nbatch doubles (asynchronous); returns the buffer."""
        nb = self.batch or 1
        if out is None:
            out = DevBuf(8 * nb)
        _check(_lib.ddx_kreduce_members(self.handle, int(k0), int(k1) if k1 is not None else None,
                                        int(op), ctypes.c_void_p(int(out))))
        return out

    def rhs(self, params, state, deriv, forcing=None):
        fz = None
        if forcing is not None:
            fz = ptr_array([int(f) if f is not None else None for f in forcing])
        _check(_lib.ddx_rhs(self.handle, ctypes.byref(params),
                            ptr_array([int(s) for s in state]),
                            ptr_array([int(d) for d in deriv]), fz))

    def step(self, integrator, params, state, stage, deriv1, deriv2, nsteps, dt, forcing=None,
             xforcing=None, xin=None):
        """ddx_step: deriv2=None leaves the second right-hand side unstored.
        xforcing (x-space arrays, one per state component or None) selects
        ddx_step_xforced; xin (the x-space inputs of the first right-hand
        side, 2D, one step) selects ddx_step_xin."""
        st = ptr_array([int(s) for s in stage]) if stage is not None else None
        d2 = ptr_array([int(d) for d in deriv2]) if deriv2 is not None else None
        fz = None
        if forcing is not None:
            fz = ptr_array([int(f) if f is not None else None for f in forcing])
        if xin is not None:
            if int(nsteps) != 1:
                raise DDXError(-1, "x-space inputs apply to a single step")
            xz = ptr_array([int(f) if f is not None else None for f in xforcing]) if xforcing is not None else None
            _check(_lib.ddx_step_xin(self.handle, integrator, ctypes.byref(params),
                                     ptr_array([int(s) for s in state]), st,
                                     ptr_array([int(d) for d in deriv1]), d2,
                                     fz, xz, ptr_array([int(x) for x in xin]), float(dt)))
            return
        if xforcing is not None:
            xz = ptr_array([int(f) if f is not None else None for f in xforcing])
            _check(_lib.ddx_step_xforced(self.handle, integrator, ctypes.byref(params),
                                         ptr_array([int(s) for s in state]), st,
                                         ptr_array([int(d) for d in deriv1]), d2,
                                         fz, xz, int(nsteps), float(dt)))
            return
        _check(_lib.ddx_step(self.handle, integrator, ctypes.byref(params),
                             ptr_array([int(s) for s in state]), st,
                             ptr_array([int(d) for d in deriv1]), d2,
                             fz, int(nsteps), float(dt)))


    def step_noalloc(self, integrator, params, state, stage, deriv1, deriv2, nsteps, dt,
                     forcing=None):
        """ddx_step_noalloc: ddx_step on the library's static device scratch."""
        st = ptr_array([int(s) for s in stage]) if stage is not None else None
        d2 = ptr_array([int(d) for d in deriv2]) if deriv2 is not None else None
        fz = None
        if forcing is not None:
            fz = ptr_array([int(f) if f is not None else None for f in forcing])
        _check(_lib.ddx_step_noalloc(self.handle, integrator, ctypes.byref(params),
                                     ptr_array([int(s) for s in state]), st,
                                     ptr_array([int(d) for d in deriv1]), d2,
                                     fz, int(nsteps), float(dt)))


def make_params(physics, nu=0., viscosity_order=1, kappa=0., eta=0., g=1., alpha_t=1.,
                beta=1., bdir=2, rho0=1.):
    return Params(physics, int(viscosity_order), float(nu), float(kappa), float(eta),
                  float(g), float(alpha_t), float(beta), int(bdir), float(rho0))


_plans = {}


def get_plan(shape, length, precision=F64, dealias=DEALIAS_23):
    """Shared plan per (shape, length, precision, dealiasing rule)."""
    key = (tuple(int(n) for n in shape), tuple(float(x) for x in length), precision, int(dealias))
    p = _plans.get(key)
    if p is None:
        p = Plan(shape, length, precision, dealias=dealias)
        _plans[key] = p
    return p
