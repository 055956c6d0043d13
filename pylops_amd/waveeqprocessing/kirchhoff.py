"""Kirchhoff demigration / migration operator on MI355X.

Drop-in for ``pylops.waveeqprocessing.Kirchhoff``
(pylops/waveeqprocessing/kirchhoff.py, K:234-1143): same constructor
signature, attributes, ``dims``/``dimsd`` and exceptions.  The engine
kernels (K:683-1001) and the wavelet Convolve1D (K:407-409) run as HIP
kernels of libkirchx.so on the device; tables are built in float64 on the
device and stored as per-tile integer sample bases plus fp32 / fp16 / fp64
sample offsets.
"""
import ctypes
import logging
import warnings

import numpy as np
import torch

from pylops_amd import _lib
from pylops_amd.linearoperator import LinearOperator
from pylops_amd.signalprocessing.convolve1d import Convolve1D
from pylops_amd.utils._internal import _value_or_sized_to_array
from pylops_amd.utils.backend import is_torch, like_input, to_device
from pylops_amd.utils.decorators import reshaped
from pylops_amd.utils.tapers import taper
from pylops_amd.waveeqprocessing import _tables as T

__all__ = ["Kirchhoff"]

logger = logging.getLogger(__name__)

_ENGINES = ("hip", "cuda", "numba", "numpy")
# small attributes whose reassignment invalidates the device tables
_PLAIN_ATTRS = ("vel", "aperture", "angleaperture", "six", "rix", "tile", "tabledtype")
# launch-shape attributes: copied into the plan, the tables stay
_LAUNCH_ATTRS = ("rblock", "sorder", "gsplit", "sside", "gtiles")


class _Table:
    """A table attribute: stored as given (numpy or torch), read back as a
    read-only numpy array (the reference's attributes are numpy arrays), and
    any reassignment marks the device copy stale (tests reassign e.g.
    ``Op.trav_srcs = ...``).  The device tables are built from the stored
    table, so an edit in place could not reach them: it raises instead
    (``ValueError: assignment destination is read-only``); assign a new
    array to change a table."""

    def __set_name__(self, owner, name):
        self.name = name

    def __get__(self, obj, objtype=None):
        if obj is None:
            return self
        if self.name not in obj._tab:
            raise AttributeError(self.name)
        host = obj._tab_host.get(self.name)
        if host is None:
            host = _host(obj._tab[self.name]).view()
            host.flags.writeable = False
            obj._tab_host[self.name] = host
        return host

    def __set__(self, obj, value):
        obj._tab[self.name] = value
        obj._tab_host.pop(self.name, None)
        object.__setattr__(obj, "_dirty", True)


class Kirchhoff(LinearOperator):
    r"""Kirchhoff demigration operator (see the reference docstring,
    K:40-232).

    Parameters are those of the reference: ``z, x, t, srcs, recs, vel, wav,
    wavcenter, y=None, mode="eikonal", wavfilter=False, dynamic=False,
    trav=None, amp=None, aperture=None, angleaperture=90.0, snell=None,
    engine="hip", dtype="float64", name="K"``.  ``engine`` accepts the
    reference names ("numpy", "numba", "cuda"); every engine runs the HIP
    kernels.  Extra keyword-only parameters:

    tabledtype : ``"float16"`` stores the receiver traveltime offsets and,
        when dynamic, the receiver weights (scaled to at most 1, the source
        weights by the same factor) in half precision (float32 operators
        only; the source tables stay float32); default is the operator
        dtype.
    tile : ``(tz, tx, ty)`` image tile of 64 points.
    device : torch device of the tables (default: current HIP device).
    rblock : receivers (or sources, see ``sside``) per forward workgroup
        (0 = automatic).
    sorder : forward lane order, ``KX_ORDER_NATURAL`` (1) or
        ``KX_ORDER_ROTATED`` (2); 0 = automatic.
    gsplit : adjoint waves per tile group, 1, 2 or 4; 0 = automatic.
    gtiles : neighbouring image tiles per adjoint wave, sharing its staged
        trace windows, 1, 2 or 4; 0 = automatic.
    eikonal_iter : sweep rounds per phase of the eikonal solver at most
        (``mode="eikonal"``); tables that do not converge within them raise
        ``RuntimeError``.
    sside : what a forward workgroup blocks, ``KX_SIDE_RECEIVERS`` (1: one
        source, a block of receivers) or ``KX_SIDE_SOURCES`` (2: one
        receiver, a block of sources; float32 operators); 0 = automatic.
        Every choice gives bitwise the same data.

    ``waveeqprocessing.autotune`` measures the fastest ``tile``, ``rblock``,
    ``sorder``, ``sside`` and ``gsplit`` for a geometry.
    """

    trav_srcs = _Table()
    trav_recs = _Table()
    trav = _Table()
    amp_srcs = _Table()
    amp_recs = _Table()
    angle_srcs = _Table()
    angle_recs = _Table()

    def __init__(self, z, x, t, srcs, recs, vel, wav, wavcenter, y=None,
                 mode="eikonal", wavfilter=False, dynamic=False, trav=None, amp=None,
                 aperture=None, angleaperture=90.0, snell=None, engine="hip",
                 dtype="float64", name="K", *, tabledtype=None, tile=None,
                 device=None, rblock=0, sorder=0, gsplit=0, sside=0, gtiles=0,
                 eikonal_iter=20):
        object.__setattr__(self, "_dirty", True)
        object.__setattr__(self, "_tab", {})
        object.__setattr__(self, "_tab_host", {})
        warnings.warn(
            "Kirchhoff tables are best passed as (sources, receivers) tuples; "
            "the single combined table is kept for back-compatibility only.",
            FutureWarning)
        if engine not in _ENGINES:
            raise KeyError("engine must be hip, numpy, numba or cuda")
        _lib.require_device()
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        dev = self.device

        (self.ndims, _, dims, self.ny, self.nx, self.nz, ns, nr, _, _, _, dsamp,
         origin) = Kirchhoff._identify_geometry(z, x, srcs, recs, y=y)
        self.dt = t[1] - t[0]
        self.nt = len(t)
        self.dynamic = dynamic
        if self.dynamic:
            self.six, self.rix = T.source_receiver_indices(self.ndims, x, srcs, recs)

        self.travsrcrec = True
        ni = int(np.prod(dims))
        if mode in ("analytic", "eikonal"):
            if mode == "eikonal":
                # the reference solves with scikit-fmm (K:609-638); here the
                # device fast-sweeping solver (kx_eikonal)
                ts, tr, ds, dr, gs, gr = T.eikonal_tables(srcs, recs, vel, origin, dsamp, dims,
                                                          dev, self.dynamic,
                                                          max_iter=eikonal_iter)
            else:
                if not isinstance(vel, (float, int)):
                    raise ValueError("vel must be scalar for mode=analytical")
                ts, tr, ds, dr, gs, gr = T.analytic(z, x, srcs, recs, vel, y, dims, dsamp, dev)
            self._set("trav_srcs", ts)
            self._set("trav_recs", tr)
            if self.dynamic:
                a_s, a_r, self.maxdist = T.amplitudes(self.ndims, ds, dr)
                self._set("amp_srcs", a_s)
                self._set("amp_recs", a_r)
            del ds, dr
        elif mode == "byot":
            if isinstance(trav, tuple):
                self._set("trav_srcs", trav[0])
                self._set("trav_recs", trav[1])
            else:
                self.travsrcrec = False
                self._set("trav", trav)
            if self.dynamic and not self.travsrcrec:
                raise NotImplementedError("separate traveltime tables must be provided "
                                          "when selecting mode=dynamic")
            if self.dynamic:
                if isinstance(amp, tuple):
                    self._set("amp_srcs", amp[0])
                    self._set("amp_recs", amp[1])
                else:
                    raise NotImplementedError("separate amplitude tables must be provided ")
                gs = T.gradient(T.as_f64(trav[0], dev), dims, None)
                gr = T.gradient(T.as_f64(trav[1], dev), dims, None)
        else:
            raise NotImplementedError("method must be analytic, eikonal or byot")

        if self.dynamic:
            angs, angr = T.angles(self.ndims, gs, gr, ni)
            self._set("angle_srcs", angs)
            self._set("angle_recs", angr)
            del gs, gr

        if wavfilter:
            self.wav = self._wavelet_reshaping(np.asarray(wav), self.dt, srcs.shape[0],
                                               recs.shape[0], self.ndims)
        else:
            self.wav = wav
        self.cop = Convolve1D((ns * nr, self.nt), h=np.asarray(self.wav), offset=wavcenter,
                              axis=1, dtype=dtype)
        # the reference's forward convolves its traces with scipy, which
        # returns the result type of traces and wavelet: a float32 operator
        # with a float64 wavelet (or wavfilter=True, whose reshaped wavelet
        # is float64) returns float64 data (C1D:99-106)
        object.__setattr__(self, "_fwd_dtype",
                           np.result_type(np.dtype(dtype), np.asarray(self.wav).dtype))
        self.aperturetap = taper(41, 20, "hanning")[20:]

        if aperture is not None and self.ndims == 3:
            aperture = None
            warnings.warn("Aperture is forced to None as currently not implemented in 3D")
        if aperture is not None:
            warnings.warn("Aperture is currently defined as ratio of offset over depth, "
                          "and may be not ideal for highly heterogeneous media")
        ap = (self.nx + 1,) if aperture is None else _value_or_sized_to_array(aperture)
        if len(ap) == 1:
            ap = np.array([0.8 * ap[0], ap[0]])
        self.aperture = ap
        angleaperture = [0.0, 1000.0] if angleaperture is None else angleaperture
        angap = np.deg2rad(_value_or_sized_to_array(angleaperture))
        if len(angap) == 1:
            angap = np.array([0.8 * angap[0], angap[0]])
        self.angleaperture = angap

        self.ns, self.nr = ns, nr
        self.nsnr = ns * nr
        self.ni = ni
        dims = tuple(dims) if self.ndims == 2 else (dims[0] * dims[1], dims[2])
        super().__init__(dtype=np.dtype(dtype), dims=dims, dimsd=(ns, nr, self.nt), name=name)
        if self.dynamic:
            self.vel = (vel.flatten() if not isinstance(vel, (float, int))
                        else vel * np.ones(ni))

        if self.dtype not in (np.float32, np.float64):
            raise NotImplementedError("dtype must be float32 or float64")
        if tabledtype is None:
            self.tabledtype = self.dtype
        else:
            self.tabledtype = np.dtype(tabledtype)
            if self.tabledtype == np.float16 and self.dtype != np.float32:
                raise NotImplementedError("float16 tables need a float32 operator")
            if self.tabledtype not in (np.float16, self.dtype):
                raise NotImplementedError(f"tabledtype {self.tabledtype} with dtype {self.dtype}")
        self.tile = tuple(tile) if tile is not None else T.default_tile(self.ndims, self.nz)
        self.rblock = int(rblock)
        self.sorder = int(sorder)
        self.sside = int(sside)
        self.gsplit = int(gsplit)
        self.gtiles = int(gtiles)
        self.engine = engine
        if engine == "cuda" and not self.travsrcrec:
            raise NotImplementedError("engine='cuda' not implemented for traveltimes "
                                      "provided in one table")
        if engine in ("numpy", "numba"):
            logger.info("engine=%s runs the HIP kernels of pylops_amd", engine)
        self._build_device_tables()

    # ------------------------------------------------------------------
    # table attributes: reassignment marks the device copy stale
    # ------------------------------------------------------------------
    def _set(self, name, value):
        setattr(self, name, value)

    def __setattr__(self, name, value):
        object.__setattr__(self, name, value)
        if name in _PLAIN_ATTRS:
            object.__setattr__(self, "_dirty", True)
        elif name in _LAUNCH_ATTRS and "_plan" in self.__dict__:
            setattr(self._plan, name, int(value))

    @property
    def itrav(self):
        """Integer sample of the combined table (K:395-398), on request."""
        if self.travsrcrec:
            raise AttributeError("itrav exists only for a single combined table")
        return (np.asarray(_host(self.trav)) / self.dt).astype("int32")

    @property
    def travd(self):
        if self.travsrcrec:
            raise AttributeError("travd exists only for a single combined table")
        q = np.asarray(_host(self.trav)) / self.dt
        return q - q.astype("int32")

    # ------------------------------------------------------------------
    # reference static helpers
    # ------------------------------------------------------------------
    @staticmethod
    def _identify_geometry(z, x, srcs, recs, y=None):
        """K:457-500."""
        ns, nr = srcs.shape[1], recs.shape[1]
        nz, nx = len(z), len(x)
        dz = np.abs(z[1] - z[0])
        dx = np.abs(x[1] - x[0])
        if y is None:
            return (2, 0, np.array([nx, nz]), 1, nx, nz, ns, nr, None, dx, dz,
                    np.array([dx, dz]), np.array([x[0], z[0]]))
        ny = len(y)
        dy = np.abs(y[1] - y[0])
        return (3, 1, np.array([ny, nx, nz]), ny, nx, nz, ns, nr, dy, dx, dz,
                np.array([dy, dx, dz]), np.array([y[0], x[0], z[0]]))

    @staticmethod
    def _traveltime_table(z, x, srcs, recs, vel, y=None, mode="eikonal"):
        """K:502-657 (numpy outputs, float64).  mode="analytic" is computed
        with torch on the current HIP device when one is visible and on the
        host otherwise; mode="eikonal" runs the device eikonal solver."""
        (ndims, _, dims, _, _, _, _, _, _, _, _, dsamp, origin) = \
            Kirchhoff._identify_geometry(z, x, srcs, recs, y=y)
        if mode == "analytic":
            if not isinstance(vel, (float, int)):
                raise ValueError("vel must be scalar for mode=analytical")
            dev = torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")
            out = T.analytic(z, x, srcs, recs, vel, y, dims, dsamp, dev)
        elif mode == "eikonal":
            _lib.require_device()
            dev = torch.device("cuda", torch.cuda.current_device())
            out = T.eikonal_tables(srcs, recs, vel, origin, dsamp, dims, dev, True)
        else:
            raise NotImplementedError("method must be analytic or eikonal")
        ts, tr, ds, dr = (o.cpu().numpy() for o in out[:4])
        gs = [g.cpu().numpy() for g in out[4]]
        gr = [g.cpu().numpy() for g in out[5]]
        return ts, tr, ds, dr, gs, gr

    def _wavelet_reshaping(self, wav, dt, dimsrc, dimrec, dimv):
        """K:659-681: sqrt(i w) shaping in 2D, -i w in 3D."""
        f = np.fft.rfftfreq(len(wav), dt)
        W = np.fft.rfft(wav, n=len(wav))
        if dimsrc == 2 and dimv == 2:
            Wfilt = W * np.sqrt(1j * 2 * np.pi * f)
        elif (dimsrc == 2 or dimrec == 2) and dimv == 3:
            raise NotImplementedError("2.D wavelet currently not available")
        elif dimsrc == 3 and dimrec == 3 and dimv == 3:
            Wfilt = W * (-1j * 2 * np.pi * f)
        return np.fft.irfft(Wfilt, n=len(wav))

    # ------------------------------------------------------------------
    # device tables and plan
    # ------------------------------------------------------------------
    def _fmt(self):
        if self.tabledtype == np.float16:
            return _lib.KX_F16
        return _lib.KX_F64 if self.dtype == np.float64 else _lib.KX_F32

    def _build_device_tables(self):
        dev = self.device
        ny = self.ny if self.ndims == 3 else 1
        plan = T.make_geom_plan(self.nx, ny, self.nz, self.tile)
        fmt = self._fmt()
        inv_dt = 1.0 / float(self.dt)
        keep = {}
        if self.travsrcrec:
            ts = T.as_f64(self._tab["trav_srcs"], dev)
            tr = T.as_f64(self._tab["trav_recs"], dev)
            ns, nr = self.ns, self.nr
        else:
            # one combined table: the kernels' receiver side (trace
            # isrc * nr + irec = the table's column), a zero source table
            ts = torch.zeros((self.ni, 1), dtype=torch.float64, device=dev)
            tr = T.as_f64(self._tab["trav"], dev)
            ns, nr = 1, self.nsnr
        if ts.shape != (self.ni, ns) or tr.shape != (self.ni, nr):
            raise ValueError(f"traveltime tables of shape {tuple(ts.shape)} / "
                             f"{tuple(tr.shape)} do not match (ni, n) = ({self.ni}, {ns}) / "
                             f"({self.ni}, {nr})")
        neg = bool((ts.min() + tr.min()) < 0)
        # Entries that no trace sample can take, whatever the other side's
        # time: a sum at or past sample nt - 1, or at or before -1 (the
        # reference's bounds tests, K:743-751), kept 2 samples clear of
        # either edge.  They are encoded apart (T.encode), so that huge
        # times standing for unreachable points cost their tile's other
        # points no precision.
        s_min, s_max = float(ts.min()) * inv_dt, float(ts.max()) * inv_dt
        r_min, r_max = float(tr.min()) * inv_dt, float(tr.max()) * inv_dt
        nt = self.nt
        lo_s, hi_s = -3.0 - r_max, nt + 1.0 - r_min
        lo_r, hi_r = -3.0 - s_max, nt + 1.0 - s_min
        if fmt == _lib.KX_F16:
            # a float16 offset reaches only 65504 samples past its base
            out = (tr < lo_r * self.dt) | (tr >= hi_r * self.dt)
            if bool(out.any()):
                keep_r = tr[~out]
                reached = float(keep_r.min()) * inv_dt if keep_r.numel() else 0.0
                if nt + 4.0 > 65504.0 + min(0.0, s_min) + min(0.0, reached):
                    raise NotImplementedError(
                        f"float16 tables cannot mark the unreachable receiver times of "
                        f"nt={nt} traces; use float32 tables")
                del keep_r
            del out
        # float16 offsets on the receiver side only: the source tables (a
        # few rows per tile) stay float32, so that a traveltime sum carries
        # one float16 rounding
        keep["off_s"], keep["base_s"] = T.encode(plan, ts, inv_dt,
                                                 _lib.KX_F32 if fmt == _lib.KX_F16 else fmt,
                                                 lo_s, hi_s)
        keep["off_r"], keep["base_r"] = T.encode(plan, tr, inv_dt, fmt, lo_r, hi_r)
        del ts, tr
        wfmt = _lib.KX_F64 if self.dtype == np.float64 else _lib.KX_F32
        if self.dynamic:
            vel = T.as_f64(self.vel, dev).reshape(-1)
            amin, amax = float(self.angleaperture[0]), float(self.angleaperture[1])
            tab = self._tab
            ps, qs = T.side_weights(T.as_f64(tab["amp_srcs"], dev),
                                    T.as_f64(tab["angle_srcs"], dev),
                                    amin, amax, self.aperturetap, scale=2.0 / vel)
            pr, qr = T.side_weights(T.as_f64(tab["amp_recs"], dev),
                                    T.as_f64(tab["angle_recs"], dev),
                                    amin, amax, self.aperturetap)
            rfmt = wfmt
            if fmt == _lib.KX_F16:
                # float16 receiver weights, scaled to at most 1 (any units
                # fit float16), the source weights by the same factor
                sc = float(torch.sqrt(pr * pr + qr * qr).max())
                if sc > 0.0:
                    ps, qs, pr, qr = ps * sc, qs * sc, pr / sc, qr / sc
                rfmt = _lib.KX_F16
            keep["wp_s"] = T.tile(plan, ps, wfmt)
            keep["wq_s"] = T.tile(plan, qs, wfmt)
            keep["wp_r"] = T.tile(plan, pr, rfmt)
            keep["wq_r"] = T.tile(plan, qr, rfmt)
            # |ps pr + qs qr| <= |(ps, qs)| |(pr, qr)|: bounds every weight
            plan.wamp = T.weight_bound(keep["wp_s"], keep["wq_s"]) * \
                T.weight_bound(keep["wp_r"], keep["wq_r"])
            # a typical weight: far below wamp, the forward accumulates in
            # float64 (the fixed-point rounding follows wamp)
            plan.wtyp = T.weight_typical(keep["wp_s"], keep["wq_s"]) * \
                T.weight_typical(keep["wp_r"], keep["wq_r"])
            izmin = T.aperture_izmin(self.six, self.rix, self.nz, float(self.aperture[1]))
            keep["izmin"] = torch.as_tensor(izmin, device=dev)
        else:
            plan.wamp = plan.wtyp = 1.0
        # forward scratch (accumulation scale), written by every forward
        keep["aux"] = torch.zeros(_lib.KX_AUX_BYTES // 4, dtype=torch.int32, device=dev)
        tdt = torch.float64 if self.dtype == np.float64 else torch.float32
        keep["h"] = torch.as_tensor(np.asarray(self.cop.h, dtype=np.float64), dtype=tdt,
                                    device=dev)
        keep["hstar"] = torch.as_tensor(np.asarray(self.cop.hstar, dtype=np.float64),
                                        dtype=tdt, device=dev)
        if self._fwd_dtype != self.dtype:
            # float64 wavelet of a float32 operator: float64 convolutions
            keep["h64"] = torch.as_tensor(np.asarray(self.cop.h, dtype=np.float64),
                                          device=dev)
            keep["hstar64"] = torch.as_tensor(np.asarray(self.cop.hstar, dtype=np.float64),
                                              device=dev)
        plan.dtype = wfmt
        plan.tab_fmt = fmt
        plan.dynamic = int(self.dynamic)
        if self.dynamic and not bool(keep["izmin"].any()):
            # no trace has an aperture gate: the adjoint skips the test
            plan.dynamic = _lib.KX_DYNAMIC_NOGATE
        plan.ns, plan.nr, plan.nt = ns, nr, self.nt
        plan.neg_trav = int(neg)
        plan.rblock = self.rblock
        plan.sorder = self.sorder
        plan.sside = self.sside
        plan.gsplit = self.gsplit
        plan.gtiles = self.gtiles
        plan.nh = keep["h"].numel()
        for k, v in keep.items():
            setattr(plan, k, v.data_ptr())
        object.__setattr__(self, "_dev_tables", keep)
        object.__setattr__(self, "_plan", plan)
        object.__setattr__(self, "_dirty", False)

    @property
    def plan(self):
        """The kx_plan (ctypes) the kernels are launched with."""
        if self._dirty:
            self._build_device_tables()
        return self._plan

    def _work(self, device, tdtype):
        return torch.empty(self.nsnr * self.nt, dtype=tdtype, device=device)

    def _into(self, x, out, adjoint):
        """Device fast path for the solvers: apply the operator (adjoint if
        ``adjoint``) to the contiguous device tensor ``x`` and write the
        result into ``out``, reusing one cached workspace, so that a captured
        iteration always touches the same buffers."""
        plan = self.plan
        ws = getattr(self, "_ws", None)
        if ws is None or ws.dtype != x.dtype or ws.device != x.device:
            ws = torch.empty(self.nsnr * self.nt, dtype=x.dtype, device=x.device)
            object.__setattr__(self, "_ws", ws)
        L = _lib.load()
        fn = L.kx_adjoint if adjoint else L.kx_forward
        rc = fn(ctypes.byref(plan), _lib.ptr(x), _lib.ptr(out), _lib.ptr(ws),
                _lib.stream_ptr(x.device))
        _lib.check(rc, "kx_adjoint" if adjoint else "kx_forward")
        if adjoint:
            self.rmatvec_count += 1
        else:
            self.matvec_count += 1
        return out

    def _launch(self, fn, x, nout, what):
        plan = self.plan
        xd = to_device(x.reshape(-1), self.dtype, self.device)
        out = torch.empty(nout, dtype=xd.dtype, device=self.device)
        work = self._work(self.device, xd.dtype)
        rc = fn(ctypes.byref(plan), _lib.ptr(xd), _lib.ptr(out), _lib.ptr(work),
                _lib.stream_ptr(self.device))
        _lib.check(rc, what)
        return like_input(out, x)

    # ------------------------------------------------------------------
    # products (K:1048-1143)
    # ------------------------------------------------------------------
    def _conv64(self, x, h):
        """float64 "same" convolution of every trace of x with h (device)."""
        out = torch.empty(self.nsnr * self.nt, dtype=torch.float64, device=self.device)
        rc = _lib.load().kx_conv(_lib.ptr(x), _lib.ptr(out), self.nsnr, self.nt, _lib.ptr(h),
                                 h.numel(), _lib.KX_F64, _lib.stream_ptr(self.device))
        _lib.check(rc, "kx_conv")
        return out

    @reshaped
    def _matvec(self, x):
        if self._fwd_dtype != self.dtype:
            # float32 traces, convolved in float64 with the float64 wavelet
            # (the reference's scipy upcast, C1D:99-106): float64 data
            traces = self.spread(to_device(x.reshape(-1), self.dtype, self.device))
            return like_input(self._conv64(traces.double(), self._dev_tables["h64"]), x)
        return self._launch(_lib.load().kx_forward, x, self.nsnr * self.nt, "kx_forward")

    @reshaped
    def _rmatvec(self, x):
        if self._fwd_dtype != self.dtype:
            # correlation in float64 (scipy's upcast), then the float32 gather
            self.plan  # noqa: B018  (rebuilds stale device tables)
            xd = to_device(x.reshape(-1), np.float64, self.device)
            dc = self._conv64(xd, self._dev_tables["hstar64"])
            return like_input(self.gather(dc.to(torch.float32)), x)
        return self._launch(_lib.load().kx_adjoint, x, self.ni, "kx_adjoint")

    # several columns at once: the reference's matmat applies matvec column
    # by column (LO:436-460), and every column here is bitwise its matvec.
    # A float32 forward spreads up to 8 models per workgroup
    # (kx_spread_models: the tables of a tile are read and the sample index
    # computed once for all of them), then convolves all their traces in
    # one launch; a float32 adjoint correlates all traces in one launch and
    # gathers 4 or 2 models per wave (kx_gather_models).  float64 products
    # run the columns one after another through the single-problem kernels,
    # on one stream, into one output buffer.  (A
    # persistent batch of k copies of this operator measured 25% slower on
    # the 601x300 dynamic forward than k launches: one problem's launch
    # already fills the GPU, and its receiver rows stay in L2 longer.)
    _MAX_MODELS = 64

    def _multi(self, X, adjoint):
        k = X.shape[1]
        nout = self.shape[1] if adjoint else self.shape[0]
        xd = to_device(X, self.dtype, self.device).T.contiguous()
        out = torch.empty((k, nout), dtype=xd.dtype, device=self.device)
        if xd.dtype == torch.float32:
            plan = self.plan
            L = _lib.load()
            st = _lib.stream_ptr(self.device)
            ntrace = self.nsnr * self.nt
            work = torch.empty((min(k, self._MAX_MODELS), ntrace), dtype=xd.dtype,
                               device=self.device)
            for j0 in range(0, k, self._MAX_MODELS):
                nm = min(self._MAX_MODELS, k - j0)
                if adjoint:
                    # correlate every column's traces with the wavelet, then
                    # gather all models (kx_adjoint's two steps, K at once)
                    rc = L.kx_conv(_lib.ptr(xd[j0:j0 + nm]), _lib.ptr(work), nm * self.nsnr,
                                   self.nt, plan.hstar, plan.nh, _lib.KX_F32, st)
                    _lib.check(rc, "kx_conv")
                    rc = L.kx_gather_models(ctypes.byref(plan), _lib.ptr(work),
                                            _lib.ptr(out[j0:j0 + nm]), nm, st)
                    _lib.check(rc, "kx_gather_models")
                else:
                    rc = L.kx_spread_models(ctypes.byref(plan), _lib.ptr(xd[j0:j0 + nm]),
                                            _lib.ptr(work), nm, st)
                    _lib.check(rc, "kx_spread_models")
                    rc = L.kx_conv(_lib.ptr(work), _lib.ptr(out[j0:j0 + nm]), nm * self.nsnr,
                                   self.nt, plan.h, plan.nh, _lib.KX_F32, st)
                    _lib.check(rc, "kx_conv")
            # one product per column, as the reference's matmat counts them
            if adjoint:
                self.rmatvec_count += k
            else:
                self.matvec_count += k
            return like_input(out.T, X)
        for j in range(k):  # each counts one matvec, as the reference's columns do
            self._into(xd[j], out[j], adjoint)
        return like_input(out.T, X)

    def _matmat(self, X):
        if X.shape[1] == 1 or self._fwd_dtype != self.dtype:
            return super()._matmat(X)
        return self._multi(X, False)

    def _rmatmat(self, X):
        if X.shape[1] == 1 or self._fwd_dtype != self.dtype:
            return super()._rmatmat(X)
        return self._multi(X, True)

    # kernels alone, without the wavelet (the engine seam, K:1003-1046)
    def spread(self, m):
        """The forward kernel alone (no wavelet): traces (ns*nr, nt)."""
        plan = self.plan
        md = to_device(m.reshape(-1), self.dtype, self.device)
        out = torch.empty(self.nsnr * self.nt, dtype=md.dtype, device=self.device)
        rc = _lib.load().kx_spread(ctypes.byref(plan), _lib.ptr(md), _lib.ptr(out),
                                   _lib.stream_ptr(self.device))
        _lib.check(rc, "kx_spread")
        return like_input(out, m)

    def gather(self, d):
        """The adjoint kernel alone (no wavelet correlation): image (ni,)."""
        plan = self.plan
        dd = to_device(d.reshape(-1), self.dtype, self.device)
        out = torch.empty(self.ni, dtype=dd.dtype, device=self.device)
        rc = _lib.load().kx_gather(ctypes.byref(plan), _lib.ptr(dd), _lib.ptr(out),
                                   _lib.stream_ptr(self.device))
        _lib.check(rc, "kx_gather")
        return like_input(out, d)


def _host(a):
    return a.detach().cpu().numpy() if is_torch(a) else np.asarray(a)
