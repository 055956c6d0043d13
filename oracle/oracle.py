"""CPU oracle for the Kirchhoff + Convolve1D + CGLS/LSQR hot path.

TEST INFRASTRUCTURE ONLY.  This module restates the reference PyLops
algorithms on the CPU so that the HIP product can be checked against them.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; nothing under ``pylops_amd`` does.

Contents (reference file:line each part follows):

* geometry / analytic traveltime tables / gradients
  -- pylops/waveeqprocessing/kirchhoff.py (K) 457-500, 502-607, 642-657
* dynamic amplitude and angle tables, six/rix, apertures
  -- K 283-301, 316-330, 343-393, 411-439
* ``KirchhoffOracle`` -- the operator: kernels in ``liboracle.so`` (plain C,
  same loop nests as K 683-1001), orchestration K 1048-1143
* ``conv_same`` / ``conv_same_rows`` -- Convolve1D short-filter path
  (one filter, or one per trace), convolve1d.py 57-115
* ``conv_long_forward`` / ``conv_long_adjoint`` -- the filter longer than
  the signal, convolve1d.py 118-199
* ``cgls`` / ``lsqr`` -- pylops/optimization/cls_basic.py 418-708, 835-1351
* ``eikonal_fsm`` -- traveltime tables of the eikonal mode, K 609-638
  (the reference calls scikit-fmm, absent here: restated as fast sweeping
  from skfmm's zero contour, pinned by closed forms)
* ``ista`` / ``fista`` and their thresholds -- pylops/optimization/
  cls_sparsity.py 36-200, 1313-2160 (synthesis form, no SOp)
* ``dottest`` -- pylops/utils/dottest.py 10-132
* ``ricker`` / ``aperture_taper`` -- pylops/utils/wavelets.py 167-206,
  pylops/utils/tapers.py 18-56, 115-151
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        _lib = ctypes.CDLL(_LIBPATH)
        _declare(_lib)
    return _lib


_p = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double


def _declare(L):
    for t in ("f64", "f32", "x64y32"):
        for k in ("kin_fwd", "kin_adj"):
            f = getattr(L, f"or_{k}_{t}")
            f.argtypes = [_p, _p, _i, _i, _i, _i64, _d, _p, _p]
            f.restype = None
        for k in ("trav_fwd", "trav_adj"):
            f = getattr(L, f"or_{k}_{t}")
            f.argtypes = [_p, _p, _i, _i, _i64, _p, _p]
            f.restype = None
        for k in ("dyn_fwd", "dyn_adj"):
            f = getattr(L, f"or_{k}_{t}")
            f.argtypes = [_p, _p, _i, _i, _i, _i64, _d, _p, _p, _p, _p, _p,
                          _d, _d, _p, _i, _p, _p, _d, _d, _p, _p]
            f.restype = None
    L.or_conv_same_f64.argtypes = [_p, _p, _i64, _i, _p, _i]
    L.or_conv_same_f64.restype = None
    L.or_num_threads.argtypes = []
    L.or_num_threads.restype = _i
    L.or_set_num_threads.argtypes = [_i]
    L.or_set_num_threads.restype = None


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------------------
# wavelets / tapers
# ---------------------------------------------------------------------------
def ricker(t, f0=10):
    """Ricker wavelet on the symmetric axis (wavelets.py:167-206)."""
    t = np.asarray(t)
    if len(t) % 2 == 0:
        t = t[:-1]
    tt = np.concatenate((np.flipud(-t[1:]), t), axis=0)
    arg = (np.pi * f0 * tt) ** 2
    w = (1 - 2 * arg) * np.exp(-arg)
    return w, tt, int(np.argmax(np.abs(w)))


def aperture_taper():
    """taper(41, 20, "hanning")[20:] (K:412, tapers.py:18-56)."""
    han = np.hanning(2 * 20 - 1)
    first = han[:20]
    full = np.concatenate([first, np.ones(41 - 40), first[::-1]])
    return full[20:]


# ---------------------------------------------------------------------------
# geometry and tables
# ---------------------------------------------------------------------------
def identify_geometry(z, x, srcs, recs, y=None):
    """K:457-500."""
    ns, nr = srcs.shape[1], recs.shape[1]
    nz, nx = len(z), len(x)
    dz, dx = abs(z[1] - z[0]), abs(x[1] - x[0])
    if y is None:
        return dict(ndims=2, shiftdim=0, dims=(nx, nz), ny=1, nx=nx, nz=nz,
                    ns=ns, nr=nr, dy=None, dx=dx, dz=dz,
                    dsamp=(dx, dz), origin=(x[0], z[0]))
    ny = len(y)
    dy = abs(y[1] - y[0])
    return dict(ndims=3, shiftdim=1, dims=(ny, nx, nz), ny=ny, nx=nx, nz=nz,
                ns=ns, nr=nr, dy=dy, dx=dx, dz=dz,
                dsamp=(dy, dx, dz), origin=(y[0], x[0], z[0]))


def analytic_tables(z, x, srcs, recs, vel, y=None):
    """Analytic traveltime/distance tables and gradients (K:570-607, 642-648).

    The squared distance is accumulated per axis in the reference's order
    (x term, z term, then y term in 3D) so that float64 results agree to the
    last bit."""
    g = identify_geometry(z, x, srcs, recs, y)
    sd = g["shiftdim"]
    if g["ndims"] == 2:
        X = np.repeat(np.asarray(x, dtype=np.float64), len(z))
        Z = np.tile(np.asarray(z, dtype=np.float64), len(x))
        Y = None
    else:
        X = np.tile(np.repeat(np.asarray(x, dtype=np.float64), len(z)), len(y))
        Z = np.tile(np.asarray(z, dtype=np.float64), len(y) * len(x))
        Y = np.repeat(np.asarray(y, dtype=np.float64), len(x) * len(z))

    def dist2(pts):
        out = np.empty((X.size, pts.shape[1]))
        for i in range(pts.shape[1]):
            p = pts[:, i]
            d = (X - p[0 + sd]) ** 2 + (Z - p[1 + sd]) ** 2
            if Y is not None:
                d = d + (Y - p[0]) ** 2
            out[:, i] = d
        return out

    ts = np.sqrt(dist2(srcs)) / vel
    tr = np.sqrt(dist2(recs)) / vel
    ds = ts * vel
    dr = tr * vel
    gs = gradient_tables(ts, g["dims"], g["dsamp"])
    gr = gradient_tables(tr, g["dims"], g["dsamp"])
    return ts, tr, ds, dr, gs, gr


def _gradient_1axis(f, h, axis):
    """Second-order central differences inside, first-order one-sided at the
    two ends, with uniform spacing h -- the rule np.gradient(edge_order=1)
    applies (used by K:643-648 and K:351-360)."""
    f = np.moveaxis(f, axis, 0)
    out = np.empty_like(f)
    out[1:-1] = (f[2:] - f[:-2]) / (2.0 * h)
    out[0] = (f[1] - f[0]) / h
    out[-1] = (f[-1] - f[-2]) / h
    return np.moveaxis(out, 0, axis)


def gradient_tables(tab, dims, dsamp=None):
    """Gradients of a (ni, n) table reshaped to (*dims, n) along every spatial
    axis; dsamp=None means unit spacing (the byot path, K:351-360)."""
    n = tab.shape[1]
    f = tab.reshape(*dims, n)
    return [_gradient_1axis(f, 1.0 if dsamp is None else float(dsamp[a]), a)
            for a in range(len(dims))]


def dynamic_tables(ndims, dims, ds, dr, gs, gr):
    """Amplitude (K:316-330) and angle (K:364-393) tables.  ds/dr may be None
    when the caller provides amplitudes (byot)."""
    ni = int(np.prod(dims))
    amps = ampr = None
    if ds is not None:
        maxdist = 1e-2 * (np.max(ds) + np.max(dr))
        if ndims == 2:
            amps = 1.0 / np.sqrt(ds + maxdist)
            ampr = 1.0 / np.sqrt(dr + maxdist)
        else:
            amps = 1.0 / (ds + maxdist)
            ampr = 1.0 / (dr + maxdist)
    ns, nr = gs[0].shape[-1], gr[0].shape[-1]
    if ndims == 2:
        angs = np.arctan2(gs[0], gs[1]).reshape(ni, ns)
        angr = np.arctan2(gr[0], gr[1]).reshape(ni, nr)
    else:
        gs3 = np.stack(gs)
        gr3 = np.stack(gr)
        sgn = np.sign(gs3[1])
        angs = (sgn * np.arccos(gs3[-1] / np.sqrt(np.sum(gs3 ** 2, axis=0)))).reshape(ni, ns)
        # the receiver angle takes the source-side sign (K:387-393); with
        # ns != nr the reference broadcast fails, so only ns == nr or a
        # broadcastable sign is representable
        angr = (sgn * np.arccos(gr3[-1] / np.sqrt(np.sum(gr3 ** 2, axis=0)))).reshape(ni, nr)
    return amps, ampr, angs, angr


def six_rix(ndims, x, srcs, recs):
    """Source/receiver x-indices for the offset aperture (K:289-301)."""
    ns, nr = srcs.shape[1], recs.shape[1]
    if ndims == 2:
        dx = x[1] - x[0]
        six = np.repeat(((srcs[0] - x[0]) // dx).astype(int), nr)
        rix = np.tile(((recs[0] - x[0]) // dx).astype(int), ns)
        return six.astype(np.float64), rix.astype(np.float64)
    return np.zeros(ns * nr), np.zeros(ns * nr)


def apertures(ndims, nx, aperture, angleaperture):
    """Aperture and angle-aperture limits (K:414-439)."""
    if aperture is not None and ndims == 3:
        aperture = None
    if aperture is None:
        ap = np.array([float(nx + 1)])
    else:
        ap = np.atleast_1d(np.asarray(aperture, dtype=np.float64))
    if ap.size == 1:
        ap = np.array([0.8 * ap[0], ap[0]])
    if angleaperture is None:
        angleaperture = [0.0, 1000.0]
    ang = np.deg2rad(np.atleast_1d(np.asarray(angleaperture, dtype=np.float64)))
    if ang.size == 1:
        ang = np.array([0.8 * ang[0], ang[0]])
    return ap, ang


# ---------------------------------------------------------------------------
# Convolve1D (short filter)
# ---------------------------------------------------------------------------
def padded_filter(h, offset):
    """Offset padding of the filter (convolve1d.py:72-86); the result always
    has odd length."""
    h = np.asarray(h, dtype=np.float64)
    nh = h.size
    if offset > nh - 1:
        raise ValueError("offset must be smaller than h.shape[axis] - 1")
    off = 2 * (nh // 2 - int(offset))
    if nh % 2 == 0:
        off -= 1
    if off > 0:
        h = np.concatenate([np.zeros(off), h])
    elif off < 0:
        h = np.concatenate([h, np.zeros(-off)])
    return h


def conv_same(x2d, h_odd):
    """y[n] = sum_k h[k] x[n + (nh-1)/2 - k] along the last axis, float64."""
    x2d = _c64(x2d)
    h = _c64(h_odd)
    y = np.empty_like(x2d)
    lib().or_conv_same_f64(_ptr(x2d), _ptr(y), x2d.shape[0], x2d.shape[1],
                           _ptr(h), h.size)
    return y


def conv_same_rows(x2d, h2d_odd):
    """conv_same with one odd filter per row (an N-D filter of
    _Convolve1Dshort, C1D:57-115, broadcast to the rows)."""
    x2d = np.asarray(x2d, dtype=np.float64)
    h2d = np.broadcast_to(np.asarray(h2d_odd, dtype=np.float64),
                          (x2d.shape[0], np.shape(h2d_odd)[-1]))
    return np.vstack([conv_same(x2d[r:r + 1], h2d[r]) for r in range(x2d.shape[0])])


def _long_shift(n, offset):
    """C1D:139-141: 2 (n//2 - offset), minus one for even n."""
    sh = 2 * (n // 2 - int(offset))
    return sh - 1 if n % 2 == 0 else sh


def conv_long_forward(x, h, offset):
    """_Convolve1Dlong._matvec (C1D:165-174) for one 1-D model x (n) and
    filter h (nh > n): the model padded by the offset rule, then
    fftconvolve(h, xp, mode="same"), i.e. the centre nh samples of the full
    convolution."""
    x = np.asarray(x, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    sh = _long_shift(x.size, offset)
    xp = np.pad(x, (max(sh, 0), -min(sh, 0)))
    full = np.convolve(h, xp)
    s0 = (full.size - h.size) // 2
    return full[s0:s0 + h.size]


def conv_long_adjoint(d, h, offset, n):
    """_Convolve1Dlong._rmatvec (C1D:176-199) for one 1-D data row d (nh):
    the data padded the other way, fully convolved with the flipped filter,
    and the n samples from len//2 - n//2 kept."""
    d = np.asarray(d, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    sh = _long_shift(n, offset)
    dp = np.pad(d, (-min(sh, 0), max(sh, 0)))
    full = np.convolve(h[::-1], dp)
    s0 = full.size // 2 - n // 2
    return full[s0:s0 + n]


# ---------------------------------------------------------------------------
# Kirchhoff operator
# ---------------------------------------------------------------------------
class KirchhoffOracle:
    """CPU restatement of pylops.waveeqprocessing.Kirchhoff (K:234-455,
    1048-1143) for mode "analytic" and "byot".  ``dtype`` selects the
    output element type (float32 reproduces per-`+=` float32 rounding)."""

    def __init__(self, z, x, t, srcs, recs, vel, wav, wavcenter, y=None,
                 mode="analytic", wavfilter=False, dynamic=False, trav=None,
                 amp=None, aperture=None, angleaperture=90.0,
                 dtype="float64"):
        g = identify_geometry(z, x, srcs, recs, y)
        self.g = g
        self.ndims = g["ndims"]
        self.nx, self.nz, self.ny = g["nx"], g["nz"], g["ny"]
        self.ns, self.nr = g["ns"], g["nr"]
        self.dt = t[1] - t[0]
        self.nt = len(t)
        self.dtype = np.dtype(dtype)
        self.dynamic = dynamic
        dims = g["dims"]
        self.ni = int(np.prod(dims))
        self.travsrcrec = True
        if dynamic:
            self.six, self.rix = six_rix(self.ndims, x, srcs, recs)
        if mode == "analytic":
            ts, tr, ds, dr, gs, gr = analytic_tables(z, x, srcs, recs, vel, y)
            self.trav_srcs, self.trav_recs = ts, tr
            if dynamic:
                self.amp_srcs, self.amp_recs, self.angle_srcs, self.angle_recs = \
                    dynamic_tables(self.ndims, dims, ds, dr, gs, gr)
        elif mode == "byot":
            if isinstance(trav, tuple):
                self.trav_srcs, self.trav_recs = (_c64(trav[0]), _c64(trav[1]))
            else:
                self.travsrcrec = False
                self.trav = _c64(trav)
            if dynamic:
                self.amp_srcs, self.amp_recs = _c64(amp[0]), _c64(amp[1])
                gs = gradient_tables(self.trav_srcs, dims, None)
                gr = gradient_tables(self.trav_recs, dims, None)
                _, _, self.angle_srcs, self.angle_recs = dynamic_tables(
                    self.ndims, dims, None, None, gs, gr)
        else:
            raise NotImplementedError("oracle covers analytic and byot")
        if not self.travsrcrec:
            q = self.trav / self.dt
            self.itrav = np.ascontiguousarray(q.astype("int32"))
            self.travd = np.ascontiguousarray(q - self.itrav)
        self.wav_f64 = np.asarray(wav).dtype == np.float64
        if wavfilter:
            wav = wavelet_filter(wav, self.dt, srcs.shape[0], recs.shape[0], self.ndims)
        self.h = padded_filter(wav, wavcenter)
        self.aptap = aperture_taper()
        self.aperture, self.angleaperture = apertures(self.ndims, self.nx, aperture,
                                                      angleaperture)
        if dynamic:
            if np.isscalar(vel):
                self.vel = vel * np.ones(self.ni)
            else:
                self.vel = _c64(np.asarray(vel).ravel())
        self.dims = (dims if self.ndims == 2 else (dims[0] * dims[1], dims[2]))
        self.dimsd = (self.ns, self.nr, self.nt)
        self.shape = (self.ns * self.nr * self.nt, self.ni)

    def _sfx(self):
        return "f32" if self.dtype == np.float32 else "f64"

    def _kernel_args(self):
        if self.dynamic:
            return (self.ns, self.nr, self.nt, self.ni, float(self.dt),
                    _ptr(self.vel), _ptr(_c64(self.trav_srcs)), _ptr(_c64(self.trav_recs)),
                    _ptr(_c64(self.amp_srcs)), _ptr(_c64(self.amp_recs)),
                    float(self.aperture[0]), float(self.aperture[1]), _ptr(self.aptap),
                    self.nz, _ptr(self.six), _ptr(self.rix),
                    float(self.angleaperture[0]), float(self.angleaperture[1]),
                    _ptr(_c64(self.angle_srcs)), _ptr(_c64(self.angle_recs)))
        return (self.ns, self.nr, self.nt, self.ni, float(self.dt),
                _ptr(_c64(self.trav_srcs)), _ptr(_c64(self.trav_recs)))

    def spread(self, m):
        """The kernel alone (K:727-889): y = zeros, accumulate."""
        m = np.ascontiguousarray(np.asarray(m).ravel(), dtype=self.dtype)
        y = np.zeros(self.ns * self.nr * self.nt, dtype=self.dtype)
        L = lib()
        if not self.travsrcrec:
            getattr(L, "or_trav_fwd_" + self._sfx())(
                _ptr(m), _ptr(y), self.ns * self.nr, self.nt, self.ni,
                _ptr(self.itrav), _ptr(self.travd))
        elif self.dynamic:
            keep = self._kernel_args()
            getattr(L, "or_dyn_fwd_" + self._sfx())(_ptr(m), _ptr(y), *keep)
        else:
            keep = self._kernel_args()
            getattr(L, "or_kin_fwd_" + self._sfx())(_ptr(m), _ptr(y), *keep)
        return y

    def gather(self, d, in_f64=False):
        """The kernel alone (K:754-1001).  ``in_f64`` feeds a float64 input
        to a float32 operator (the reference's adjoint after a float64
        wavelet correlation)."""
        sfx = self._sfx()
        if in_f64 and sfx == "f32":
            sfx = "x64y32"
            d = np.ascontiguousarray(np.asarray(d).ravel(), dtype=np.float64)
        else:
            d = np.ascontiguousarray(np.asarray(d).ravel(), dtype=self.dtype)
        y = np.zeros(self.ni, dtype=self.dtype)
        L = lib()
        if not self.travsrcrec:
            getattr(L, "or_trav_adj_" + sfx)(
                _ptr(d), _ptr(y), self.ns * self.nr, self.nt, self.ni,
                _ptr(self.itrav), _ptr(self.travd))
        elif self.dynamic:
            keep = self._kernel_args()
            getattr(L, "or_dyn_adj_" + sfx)(_ptr(d), _ptr(y), *keep)
        else:
            keep = self._kernel_args()
            getattr(L, "or_kin_adj_" + sfx)(_ptr(d), _ptr(y), *keep)
        return y

    def matvec(self, m):
        """K:1048-1094: kernel, then the wavelet convolution."""
        y = self.spread(m).reshape(self.ns * self.nr, self.nt)
        return conv_same(y, self.h).ravel()

    def rmatvec(self, d):
        """K:1096-1143: correlation with the wavelet, then the kernel."""
        d = np.asarray(d, dtype=np.float64).reshape(self.ns * self.nr, self.nt)
        dc = conv_same(d, self.h[::-1])
        return self.gather(dc, in_f64=self.wav_f64).astype(np.float64)


def wavelet_filter(wav, dt, dimsrc, dimrec, dimv):
    """K:659-681 (2D sqrt(i w) / 3D -i w spectral shaping)."""
    f = np.fft.rfftfreq(len(wav), dt)
    W = np.fft.rfft(wav, n=len(wav))
    if dimsrc == 2 and dimv == 2:
        Wf = W * np.sqrt(1j * 2 * np.pi * f)
    elif dimsrc == 3 and dimrec == 3 and dimv == 3:
        Wf = W * (-1j * 2 * np.pi * f)
    else:
        raise NotImplementedError("2.D wavelet currently not available")
    return np.fft.irfft(Wf, n=len(wav))


# ---------------------------------------------------------------------------
# dense operator wrapper (for brute-force checks and solver tests)
# ---------------------------------------------------------------------------
class DenseOp:
    def __init__(self, A):
        self.A = np.asarray(A, dtype=np.float64)
        self.shape = self.A.shape

    def matvec(self, x):
        return self.A @ x

    def rmatvec(self, y):
        return self.A.T @ y


# ---------------------------------------------------------------------------
# solvers (float64)
# ---------------------------------------------------------------------------
def cgls(Op, y, x0=None, niter=10, damp=0.0, tol=1e-4):
    """CGLS as pylops/optimization/cls_basic.py:418-630 (non-preallocated
    branch).  Returns (x, istop, iiter, r1norm, r2norm, cost)."""
    y = np.asarray(y, dtype=np.float64)
    dampsq = damp ** 2
    if x0 is None:
        x = np.zeros(Op.shape[1])
        s = y.copy()
        c = Op.rmatvec(s)
    else:
        x = np.asarray(x0, dtype=np.float64).copy()
        s = y - Op.matvec(x)
        c = Op.rmatvec(s) - damp * x           # raw damp here (CB:477)
    q = Op.matvec(c)
    kold = abs(c.dot(c))
    cost = [float(np.linalg.norm(s))]
    cost1 = [float(np.sqrt(cost[0] ** 2 + damp * abs(x.dot(x))))]
    iiter = 0
    while iiter < niter and kold > tol:
        a = kold / (q.dot(q) + dampsq * c.dot(c))
        x = x + a * c
        s = s - a * q
        r = Op.rmatvec(s) - dampsq * x
        k = abs(r.dot(r))
        b = k / kold
        c = r + b * c
        q = Op.matvec(c)
        kold = k
        iiter += 1
        cost.append(float(np.linalg.norm(s)))
        cost1.append(float(np.sqrt(cost[iiter] ** 2 + dampsq * abs(x.dot(x)))))
    istop = 1 if kold < tol else 2
    return x, istop, iiter, kold, cost1[iiter], np.array(cost)


def lsqr(Op, y, x0=None, damp=0.0, atol=1e-8, btol=1e-8, conlim=1e8,
         niter=10, calc_var=True):
    """LSQR as pylops/optimization/cls_basic.py:835-1186 (non-preallocated
    branch), including the scalar `var` update of CB:1066-1069.  Returns the
    11-tuple of CB:1339-1351."""
    y = np.asarray(y, dtype=np.float64)
    m, n = Op.shape
    var = np.zeros(n) if calc_var else None
    iiter = 0
    istop = 0
    ctol = 1.0 / conlim if conlim > 0 else 0
    anorm = 0.0
    acond = 0.0
    dampsq = damp ** 2
    ddnorm = 0.0
    res2 = 0.0
    xnorm = 0.0
    xxnorm = 0.0
    z = 0.0
    cs2 = -1.0
    sn2 = 0.0
    if x0 is None:
        x = np.zeros(n)
        u = y.copy()
    else:
        x = np.asarray(x0, dtype=np.float64).copy()
        u = y - Op.matvec(x)
    alfa = 0.0
    beta = np.linalg.norm(u)
    if beta > 0:
        u = u / beta
        v = Op.rmatvec(u)
        alfa = np.linalg.norm(v)
        if alfa > 0:
            v = v / alfa
    else:
        v = x.copy()
        alfa = 0
    w = v.copy()
    arnorm = alfa * beta
    arnorm0 = arnorm
    rhobar = alfa
    phibar = beta
    bnorm = beta
    rnorm = beta
    r1norm = rnorm
    r2norm = rnorm
    cost = [float(rnorm)]
    test1 = test2 = test3 = 0.0
    while iiter < niter and istop == 0:
        u = Op.matvec(v) - alfa * u
        beta = np.linalg.norm(u)
        if beta > 0:
            u = u / beta
            anorm = np.linalg.norm([anorm, alfa, beta, damp])
            v = Op.rmatvec(u) - beta * v
            alfa = np.linalg.norm(v)
            if alfa > 0:
                v = v / alfa
        rhobar1 = np.linalg.norm([rhobar, damp])
        cs1 = rhobar / rhobar1
        sn1 = damp / rhobar1
        psi = sn1 * phibar
        phibar = cs1 * phibar
        rho = np.linalg.norm([rhobar1, beta])
        cs = rhobar1 / rho
        sn = beta / rho
        theta = sn * alfa
        rhobar = -cs * alfa
        phi = cs * phibar
        phibar = sn * phibar
        tau = sn * phi
        t1 = phi / rho
        t2 = -theta / rho
        dk = w / rho
        x = x + t1 * w
        w = v + t2 * w
        ddnorm = ddnorm + np.linalg.norm(dk) ** 2
        if calc_var:
            var = var + np.dot(dk, dk)
        delta = sn2 * rho
        gambar = -cs2 * rho
        rhs = phi - delta * z
        zbar = rhs / gambar
        xnorm = np.sqrt(xxnorm + zbar ** 2)
        gamma = np.linalg.norm([gambar, theta])
        cs2 = gambar / gamma
        sn2 = theta / gamma
        z = rhs / gamma
        xxnorm = xxnorm + z ** 2.0
        acond = anorm * np.sqrt(ddnorm)
        res1 = phibar ** 2
        res2 = res2 + psi ** 2
        rnorm = np.sqrt(res1 + res2)
        arnorm = alfa * abs(tau)
        r1sq = rnorm ** 2 - dampsq * xxnorm
        r1norm = np.sqrt(abs(r1sq))
        cost.append(float(r1norm))
        if r1sq < 0:
            r1norm = -r1norm
        r2norm = rnorm
        test1 = rnorm / bnorm
        test2 = arnorm / arnorm0
        test3 = 1.0 / acond
        tt1 = test1 / (1.0 + anorm * xnorm / bnorm)
        rtol = btol + atol * anorm * xnorm / bnorm
        if iiter >= niter:
            istop = 7
        if 1 + test3 <= 1:
            istop = 6
        if 1 + test2 <= 1:
            istop = 5
        if 1 + tt1 <= 1:
            istop = 4
        if test3 <= ctol:
            istop = 3
        if test2 <= atol:
            istop = 2
        if test1 <= rtol:
            istop = 1
        iiter += 1
    return (x, istop, iiter, r1norm, r2norm, anorm, acond, arnorm, xnorm, var,
            np.array(cost))


_EIK_BIG = 1e30
_EIK_R = 2


def _solve_upwind(a, b, g, sl, three):
    """Largest t with sum_k ((t - b_k)/g_k)^2 = s^2 over the axes whose
    upwind value a_k lies below t, taken in increasing a_k: the Godunov
    update with one-sided differences (t - b_k)/g_k per axis (first order:
    b = a, g = h; second order: b = (4a - a2)/3, g = 2h/3).  Vectorised,
    same operation order as the device kernel.  Returns (t, ok) with ok
    False where a quadratic had no real root."""
    a0, a1, a2 = a
    b0, b1, b2 = b
    g0, g1, g2 = g

    def swap(u, v):
        m = v[0] < u[0]
        return (tuple(np.where(m, y, x) for x, y in zip(u, v)),
                tuple(np.where(m, x, y) for x, y in zip(u, v)))

    (a0, b0, g0), (a1, b1, g1) = swap((a0, b0, g0), (a1, b1, g1))
    if three:
        (a1, b1, g1), (a2, b2, g2) = swap((a1, b1, g1), (a2, b2, g2))
        (a0, b0, g0), (a1, b1, g1) = swap((a0, b0, g0), (a1, b1, g1))
    t1 = b0 + sl * g0
    w0 = 1.0 / (g0 * g0)
    w1 = 1.0 / (g1 * g1)
    A = w0 + w1
    B = b0 * w0 + b1 * w1
    C = b0 * b0 * w0 + b1 * b1 * w1 - sl * sl
    D = B * B - A * C
    t2 = (B + np.sqrt(np.maximum(D, 0.0))) / A
    m2 = t1 > a1
    t = np.where(m2, t2, t1)
    ok = ~m2 | (D >= 0.0)
    if three:
        w2 = 1.0 / (g2 * g2)
        A3 = A + w2
        B3 = B + b2 * w2
        C3 = C + b2 * b2 * w2
        D3 = B3 * B3 - A3 * C3
        t3 = (B3 + np.sqrt(np.maximum(D3, 0.0))) / A3
        m3 = m2 & (t2 > a2)
        t = np.where(m3, t3, t)
        ok = ok & (~m3 | (D3 >= 0.0))
    return t, ok


def _axis_terms(l1, r1, l2, r2, h, order):
    """Upwind value a, difference base b and step g of one axis from the
    neighbours one (l1, r1) and two (l2, r2) cells away on each side
    (_EIK_BIG where absent): the smaller side is upwind; second order uses
    (3t - 4a + a2)/(2h) when the second cell on that side is known and not
    above the first."""
    left = l1 <= r1
    a = np.where(left, l1, r1)
    h = np.broadcast_to(np.asarray(h, dtype=np.float64), a.shape)
    if order == 1:
        return a, a, h
    a2 = np.where(left, l2, r2)
    two = (a < _EIK_BIG) & (a2 < _EIK_BIG) & (a2 <= a)
    return a, np.where(two, (4.0 * a - a2) / 3.0, a), np.where(two, (2.0 / 3.0) * h, h)


def eikonal_fsm(slow, src, h, max_iter=20, order=2):
    """First-arrival times T[n][ny, nx, nz] from the zero contour of phi = -1
    at each source cell src[s] = (ix, iy, iz) and +1 elsewhere, for the
    slowness slow[ny, nx, nz] (2D: ny = 1, hy unused), by fast sweeping
    (Zhao 2005) with Godunov upwind updates.  Cells within two cells (per
    axis) of the source hold the straight-ray time from the contour (a
    diamond through the half-way points to the neighbouring cell centres)
    and stay fixed.  First-order sweeps run to convergence; with order 2,
    second-order sweeps (one-sided differences over two cells where the
    upwind values are monotone, first order where the quadratic has no root)
    then lower the values further until they converge too.  Returns
    (T, rounds), rounds counting both phases."""
    slow = np.asarray(slow, dtype=np.float64)
    ny, nx, nz = slow.shape
    hx, hy, hz = h
    three = ny > 1
    if not three:
        hy = 1.0
    src = np.asarray(src, dtype=np.int64).reshape(-1, 3)
    n = len(src)
    IY, IX, IZ = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    T = np.full((n, ny, nx, nz), _EIK_BIG)
    frozen = np.zeros((n, ny, nx, nz), dtype=bool)
    for k, (sx, sy, sz) in enumerate(src):
        box = ((np.abs(IX - sx) <= _EIK_R) & (np.abs(IY - sy) <= _EIK_R)
               & (np.abs(IZ - sz) <= _EIK_R))
        dx = (IX - sx) * hx
        dy = (IY - sy) * hy
        dz = (IZ - sz) * hz
        rho = np.sqrt(dx * dx + dy * dy + dz * dz)
        S = 2.0 * (np.abs(dx) / hx + np.abs(dz) / hz + (np.abs(dy) / hy if three else 0.0))
        sl = 0.5 * (slow + slow[sy, sx, sz])
        with np.errstate(divide="ignore", invalid="ignore"):
            v = np.where(S > 0, sl * rho * np.maximum(0.0, 1.0 - 1.0 / S), 0.0)
        T[k][box] = v[box]
        frozen[k] = box
    ndiag = (nx - 1) + (ny - 1) + (nz - 1) + 1
    ndir = 8 if three else 4
    BIG = _EIK_BIG

    def nb(cy, cx, cz, axis, off):
        """T at the neighbour off cells along axis, BIG outside the grid."""
        if axis == 0:
            j, nn = cx + off, nx
            v = T[:, cy, np.clip(j, 0, nx - 1), cz]
        elif axis == 1:
            j, nn = cy + off, ny
            v = T[:, np.clip(j, 0, ny - 1), cx, cz]
        else:
            j, nn = cz + off, nz
            v = T[:, cy, cx, np.clip(j, 0, nz - 1)]
        return np.where((j >= 0) & (j < nn), v, BIG)

    rounds = 0
    for phase in ((1, 2) if order == 2 else (1,)):
        for _ in range(max_iter):
            rounds += 1
            changed = False
            for q in range(ndir):
                fx = (q & 1) != 0
                fz = (q & 2) != 0
                fy = (q & 4) != 0
                PX = np.where(fx, nx - 1 - IX, IX)
                PY = np.where(fy, ny - 1 - IY, IY)
                PZ = np.where(fz, nz - 1 - IZ, IZ)
                diag = PX + PY + PZ
                for d in range(ndiag):
                    cy, cx, cz = np.nonzero(diag == d)
                    if len(cx) == 0:
                        continue
                    Tc = T[:, cy, cx, cz]
                    sl = np.broadcast_to(slow[cy, cx, cz], Tc.shape)
                    terms = []
                    for axis, hh in ((0, hx), (2, hz), (1, hy)):
                        if axis == 1 and not three:
                            big = np.full_like(Tc, BIG)
                            terms.append((big, big, np.broadcast_to(hh, Tc.shape)))
                            continue
                        l2 = nb(cy, cx, cz, axis, -2) if phase == 2 else None
                        r2 = nb(cy, cx, cz, axis, 2) if phase == 2 else None
                        terms.append(_axis_terms(nb(cy, cx, cz, axis, -1),
                                                 nb(cy, cx, cz, axis, 1), l2, r2, hh, phase))
                    a = tuple(tt[0] for tt in terms)
                    b = tuple(tt[1] for tt in terms)
                    g = tuple(tt[2] for tt in terms)
                    t, ok = _solve_upwind(a, b, g, sl, three)
                    if phase == 2:
                        hs = tuple(np.broadcast_to(hh, Tc.shape) for hh in (hx, hz, hy))
                        t1, _ = _solve_upwind(a, a, hs, sl, three)
                        t = np.where(ok, t, t1)
                    upd = (t < Tc) & ~frozen[:, cy, cx, cz]
                    if np.any(upd & (Tc - t > 1e-13 * t)):
                        changed = True
                    T[:, cy, cx, cz] = np.where(upd, t, Tc)
            if not changed:
                break
    return T, rounds


def soft_threshold(x, thresh):
    """cls_sparsity.py:64-93 (real case)."""
    return np.maximum(np.abs(x) - thresh, 0.0) * np.sign(x)


def hard_threshold(x, thresh):
    """cls_sparsity.py:36-61."""
    x1 = x.copy()
    x1[np.abs(x) <= np.sqrt(2 * thresh)] = 0
    return x1


def half_threshold(x, thresh):
    """cls_sparsity.py:96-133 (the closed-form proximal map of
    thresh * |u|^(1/2), Chen, Shen and Suter 2016)."""
    arg = np.ones_like(x)
    nz = x != 0
    arg[nz] = (thresh / 8.0) * (np.abs(x[nz]) / 3.0) ** (-1.5)
    arg = np.clip(arg, -1, 1)
    phi = 2.0 / 3.0 * np.arccos(arg)
    x1 = 2.0 / 3.0 * x * (1 + np.cos(2.0 * np.pi / 3.0 - phi))
    x1[np.abs(x) <= (54 ** (1.0 / 3.0) / 4.0) * thresh ** (2.0 / 3.0)] = 0
    return x1


_THRESH = {"soft": soft_threshold, "hard": hard_threshold, "half": half_threshold}


def shrinkage(Op, y, alpha, niter=10, eps=0.1, tol=1e-10, threshkind="soft", x0=None,
              fast=True, decay=None, S=None):
    """ISTA (fast=False, cls_sparsity.py:1700-1800) or FISTA (fast=True,
    cls_sparsity.py:2010-2120) with a given step ``alpha``; ``S`` (a dense
    matrix) is the analysis operator SOp: the threshold acts on S^H xu and
    S maps back (cls_sparsity.py:1734-1758).  Returns ``(x, niter, cost)``."""
    thr = _THRESH[threshkind]
    thresh = eps * alpha * 0.5
    x = np.zeros(Op.shape[1]) if x0 is None else np.asarray(x0, dtype=np.float64).copy()
    z = x.copy()
    t = 1.0
    cost = []
    it = 0
    xupdate = np.inf
    while it < niter and xupdate > tol:
        xold = x.copy()
        base = z if fast else x
        res = y - Op.matvec(base)
        xu = base + alpha * Op.rmatvec(res)
        d = 1.0 if decay is None else decay[it]
        if S is None:
            x = thr(xu, d * thresh)
        else:
            x = S @ thr(S.conj().T @ xu, d * thresh)
        if fast:
            told = t
            t = (1.0 + np.sqrt(1.0 + 4.0 * t ** 2)) / 2.0
            z = x + ((told - 1.0) / t) * (x - xold)
            costdata = 0.5 * np.linalg.norm(y - Op.matvec(x)) ** 2
        else:
            costdata = 0.5 * np.linalg.norm(res) ** 2
        xupdate = np.linalg.norm(x - xold)
        cost.append(costdata + eps * np.linalg.norm(x, ord=1))
        it += 1
    return x, it, np.array(cost)


def dottest_values(Op, seed=0, dtype=np.float64):
    """(yy, xx) = (v.(Op u), u.(Op^H v)) on N(0,1) vectors (dottest.py:97-112)."""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(Op.shape[1]).astype(dtype)
    v = rng.standard_normal(Op.shape[0]).astype(dtype)
    yy = np.vdot(Op.matvec(u), v)
    xx = np.vdot(u, Op.rmatvec(v))
    return yy, xx
